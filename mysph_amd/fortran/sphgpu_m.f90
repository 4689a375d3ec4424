!! Fortran binding of libsphgpu (include/sphgpu.h) over iso_c_binding.
!!
!! The derived types are bind(C) mirrors of sphgpu_material / sphgpu_config;
!! every entry point is declared with the C name it binds. Status codes are
!! the negative values of sphgpu_status; sphgpu_check turns a failure into
!! `error stop` with the library's message, which is how the reference
!! reports fatal conditions (src/NNPS/asymmetricSearch.f90:38-39).
!!
!! Presets: sphgpu_preset fills the per-case constants the reference selects
!! by `nick` (eos.f90:35-40, viscosity.f90:14-33, artificialViscosity.f90:
!! 30-79, externalForce.f90:62-101, artificialTerm.f90:27-32), including the
!! single-precision literals the reference promotes to real(8).
module sphgpu_m
    use, intrinsic :: iso_c_binding
    implicit none
    private

    integer, parameter, public :: SPHGPU_MAX_MATERIALS = 8

    !! sphgpu_status
    integer(c_int), parameter, public :: SPHGPU_OK = 0, SPHGPU_E_OUT_OF_RANGE = -1, &
        SPHGPU_E_NONFINITE = -2, SPHGPU_E_PAIR_CAPACITY = -3, SPHGPU_E_BAD_CONFIG = -4, &
        SPHGPU_E_HIP = -5, SPHGPU_E_ARG = -6, SPHGPU_E_CAPACITY = -7
    !! sphgpu_precision
    integer(c_int32_t), parameter, public :: SPHGPU_FP64 = 0, SPHGPU_MIXED = 1
    !! sphgpu_eos
    integer(c_int32_t), parameter, public :: SPHGPU_EOS_NONE = 0, SPHGPU_EOS_TAIT = 1, &
        SPHGPU_EOS_LINEAR = 2, SPHGPU_EOS_ISOTHERMAL = 3
    !! sphgpu_gravity_mode
    integer(c_int32_t), parameter, public :: SPHGPU_GRAVITY_ALL = 0, SPHGPU_GRAVITY_SKIP_SOLID = 1
    !! sphgpu_field
    integer(c_int32_t), parameter, public :: SPHGPU_F_X = 0, SPHGPU_F_V = 1, SPHGPU_F_RHO = 2, &
        SPHGPU_F_P = 3, SPHGPU_F_E = 4, SPHGPU_F_MASS = 5, SPHGPU_F_H = 6, SPHGPU_F_C = 7, &
        SPHGPU_F_DISP = 8, SPHGPU_F_STRESS = 9, SPHGPU_F_SHEAR = 10, SPHGPU_F_DVDT = 11, &
        SPHGPU_F_DRHODT = 12, SPHGPU_F_DEDT = 13, SPHGPU_F_DSDT = 14, SPHGPU_F_DIVX = 15, &
        SPHGPU_F_DIVV = 16, SPHGPU_F_ETA = 17

    type, bind(C), public :: sphgpu_material
        integer(c_int32_t) :: type
        integer(c_int32_t) :: eos
        real(c_double) :: rho0
        real(c_double) :: K
        real(c_double) :: gamma
        real(c_double) :: eta
        real(c_double) :: sigma_y
    end type sphgpu_material

    type, bind(C), public :: sphgpu_config
        integer(c_int32_t) :: dim
        integer(c_int32_t) :: skf
        integer(c_int32_t) :: pa_sph
        integer(c_int32_t) :: summation_density
        integer(c_int32_t) :: normalize_density
        integer(c_int32_t) :: viscosity
        integer(c_int32_t) :: gravity
        integer(c_int32_t) :: gravity_mode
        integer(c_int32_t) :: external_force
        integer(c_int32_t) :: artificial_viscosity
        integer(c_int32_t) :: artificial_stress
        integer(c_int32_t) :: kernel_correction
        integer(c_int32_t) :: precision
        integer(c_int32_t) :: n_materials
        real(c_double) :: gravity_g
        real(c_double) :: av_alpha, av_beta, av_psi
        real(c_double) :: lj_factor, lj_r0, lj_p1, lj_p2
        real(c_double) :: pen_factor, pen_n1, pen_n2
        real(c_double) :: couple_delta
        real(c_double) :: as_lambda, as_delta, as_exponent
        real(c_double) :: dt
        type(sphgpu_material) :: mat(SPHGPU_MAX_MATERIALS)
        integer(c_int32_t) :: fluid_visc_skip_solid
        integer(c_int32_t) :: sle
    end type sphgpu_config

    public :: sphgpu_abi_version, sphgpu_create, sphgpu_upload, sphgpu_derivatives, &
              sphgpu_advance, sphgpu_download, sphgpu_upload_field, sphgpu_export_pairs, &
              sphgpu_neighbour_stats, &
              sphgpu_grid_info, sphgpu_timing, sphgpu_step, sphgpu_last_error, sphgpu_destroy, &
              sphgpu_checkpoint_size, sphgpu_checkpoint_save, sphgpu_checkpoint_restore, &
              sphgpu_context_info
    public :: sphgpu_check, sphgpu_error_message, sphgpu_base_config, sphgpu_add_material, &
              sphgpu_preset, sphgpu_tait_b, sphgpu_sound_speed, f32

    interface
        integer(c_int) function sphgpu_abi_version() bind(C, name="sphgpu_abi_version")
            import :: c_int
        end function sphgpu_abi_version

        integer(c_int) function sphgpu_create(cfg, capacity, ctx) bind(C, name="sphgpu_create")
            import :: c_int, c_int64_t, c_ptr, sphgpu_config
            type(sphgpu_config), intent(in) :: cfg
            integer(c_int64_t), value :: capacity
            type(c_ptr), intent(out) :: ctx
        end function sphgpu_create

        !! Pointers are c_loc of contiguous arrays, or c_null_ptr where the
        !! header allows NULL.
        integer(c_int) function sphgpu_upload(ctx, n_real, n_dummy, x, v, mass, rho, p, e, c, &
                                              h, shear, type, state, boundary) &
                bind(C, name="sphgpu_upload")
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: n_real, n_dummy
            type(c_ptr), value :: x, v, mass, rho, p, e, c, h, shear, type, state, boundary
        end function sphgpu_upload

        integer(c_int) function sphgpu_derivatives(ctx) bind(C, name="sphgpu_derivatives")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function sphgpu_derivatives

        integer(c_int) function sphgpu_advance(ctx, nsteps) bind(C, name="sphgpu_advance")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nsteps
        end function sphgpu_advance

        integer(c_int) function sphgpu_download(ctx, field, dst) bind(C, name="sphgpu_download")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: field
            type(c_ptr), value :: dst
        end function sphgpu_download

        integer(c_int) function sphgpu_upload_field(ctx, field, src) bind(C, name="sphgpu_upload_field")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: field
            type(c_ptr), value :: src
        end function sphgpu_upload_field

        integer(c_int) function sphgpu_export_pairs(ctx, offsets, nbr, cap, needed) &
                bind(C, name="sphgpu_export_pairs")
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: ctx, offsets, nbr
            integer(c_int64_t), value :: cap
            integer(c_int64_t), intent(out) :: needed
        end function sphgpu_export_pairs

        integer(c_int) function sphgpu_neighbour_stats(ctx, out4) bind(C, name="sphgpu_neighbour_stats")
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int64_t), intent(out) :: out4(4)
        end function sphgpu_neighbour_stats

        integer(c_int) function sphgpu_grid_info(ctx, out9) bind(C, name="sphgpu_grid_info")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: out9(9)
        end function sphgpu_grid_info

        integer(c_int) function sphgpu_timing(ctx, advance_ms, force_kernel_ms, force_launches) &
                bind(C, name="sphgpu_timing")
            import :: c_int, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: advance_ms, force_kernel_ms
            integer(c_int64_t), intent(out) :: force_launches
        end function sphgpu_timing

        integer(c_int64_t) function sphgpu_step(ctx) bind(C, name="sphgpu_step")
            import :: c_int64_t, c_ptr
            type(c_ptr), value :: ctx
        end function sphgpu_step

        type(c_ptr) function sphgpu_last_error(ctx) bind(C, name="sphgpu_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function sphgpu_last_error

        subroutine sphgpu_destroy(ctx) bind(C, name="sphgpu_destroy")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end subroutine sphgpu_destroy

        !! Exact checkpoint: every device array and the step state
        integer(c_int) function sphgpu_checkpoint_size(ctx, nbytes) bind(C, name="sphgpu_checkpoint_size")
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), intent(out) :: nbytes
        end function sphgpu_checkpoint_size

        integer(c_int) function sphgpu_checkpoint_save(ctx, buf, nbytes) bind(C, name="sphgpu_checkpoint_save")
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: ctx, buf
            integer(c_int64_t), value :: nbytes
        end function sphgpu_checkpoint_save

        integer(c_int) function sphgpu_checkpoint_restore(buf, nbytes, ctx) &
                bind(C, name="sphgpu_checkpoint_restore")
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: buf
            integer(c_int64_t), value :: nbytes
            type(c_ptr), intent(out) :: ctx
        end function sphgpu_checkpoint_restore

        integer(c_int) function sphgpu_context_info(ctx, cfg, counts) bind(C, name="sphgpu_context_info")
            import :: c_int, c_ptr, c_int64_t, sphgpu_config
            type(c_ptr), value :: ctx
            type(sphgpu_config), intent(out) :: cfg
            integer(c_int64_t), intent(out) :: counts(3)
        end function sphgpu_context_info

        integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function c_strlen
    end interface

contains

    !! Value of a default-real (single-precision) literal promoted to real(8).
    elemental real(8) function f32(x)
        real(8), intent(in) :: x
        f32 = real(real(x, 4), 8)
    end function f32

    function sphgpu_error_message(ctx) result(msg)
        type(c_ptr), intent(in) :: ctx
        character(len=:), allocatable :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: chars(:)
        integer :: n, i
        p = sphgpu_last_error(ctx)
        if (.not. c_associated(p)) then
            msg = ""
            return
        end if
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: msg)
        do i = 1, n
            msg(i:i) = chars(i)
        end do
    end function sphgpu_error_message

    !! Stop on a failed call, with the call name, status and the library's
    !! message.
    subroutine sphgpu_check(status, ctx, what)
        integer(c_int), intent(in) :: status
        type(c_ptr), intent(in) :: ctx
        character(len=*), intent(in) :: what
        character(len=:), allocatable :: msg
        if (status == SPHGPU_OK) return
        msg = sphgpu_error_message(ctx)
        write(*, "(A, A, A, I0, A, A)") " sphgpu error in ", what, " (status ", status, "): ", msg
        error stop "libsphgpu call failed"
    end subroutine sphgpu_check

    !! The reference's defaults for a nick with no special case
    !! (getValueFromToml.f90:99-150, externalForce.f90:62-72,
    !! artificialViscosity.f90:75-78, artificialTerm.f90:17-32).
    function sphgpu_base_config(dim) result(c)
        integer, intent(in) :: dim
        type(sphgpu_config) :: c
        integer :: k
        c%dim = dim
        c%skf = 1
        c%pa_sph = 2
        c%summation_density = 1
        c%normalize_density = 0
        c%viscosity = 0
        c%gravity = 0
        c%gravity_mode = SPHGPU_GRAVITY_ALL
        c%external_force = 0
        c%artificial_viscosity = 0
        c%artificial_stress = 0
        c%kernel_correction = 0
        c%precision = SPHGPU_FP64
        c%n_materials = 0
        c%gravity_g = -9.81_8
        c%av_alpha = 1; c%av_beta = 1; c%av_psi = f32(0.1_8)
        c%lj_factor = f32(1e-2_8); c%lj_r0 = f32(1.25e-5_8); c%lj_p1 = 12; c%lj_p2 = 4
        c%pen_factor = 1e5_8; c%pen_n1 = 6; c%pen_n2 = 4
        c%couple_delta = f32(0.001_8)
        c%as_lambda = f32(0.2_8); c%as_delta = 0; c%as_exponent = 4
        c%dt = 5e-4_8
        c%fluid_visc_skip_solid = 0
        c%sle = 0
        do k = 1, SPHGPU_MAX_MATERIALS
            c%mat(k) = sphgpu_material(0, 0, 0.0_8, 0.0_8, 1.0_8, 0.0_8, 0.0_8)
        end do
    end function sphgpu_base_config

    subroutine sphgpu_add_material(c, type, eos, rho0, K, gamma, eta, sigma_y)
        type(sphgpu_config), intent(inout) :: c
        integer, intent(in) :: type, eos
        real(8), intent(in) :: rho0, K, gamma, eta
        real(8), intent(in), optional :: sigma_y
        integer :: n
        n = c%n_materials + 1
        if (n > SPHGPU_MAX_MATERIALS) error stop "sphgpu_add_material: too many materials"
        c%mat(n)%type = abs(type)
        c%mat(n)%eos = eos
        c%mat(n)%rho0 = rho0
        c%mat(n)%K = K
        c%mat(n)%gamma = gamma
        c%mat(n)%eta = eta
        c%mat(n)%sigma_y = 0
        if (present(sigma_y)) c%mat(n)%sigma_y = sigma_y
        c%n_materials = n
    end subroutine sphgpu_add_material

    !! B = c^2 rho0 / gamma, in the order eos.f90:41 evaluates it.
    real(8) function sphgpu_tait_b(c0, rho0, gamma)
        real(8), intent(in) :: c0, rho0, gamma
        sphgpu_tait_b = c0 * c0 * rho0 / gamma
    end function sphgpu_tait_b

    real(8) function sphgpu_sound_speed(c, type)
        type(sphgpu_config), intent(in) :: c
        integer, intent(in) :: type
        integer :: k
        sphgpu_sound_speed = 0
        do k = 1, c%n_materials
            if (c%mat(k)%type /= abs(type)) cycle
            select case (c%mat(k)%eos)
            case (SPHGPU_EOS_TAIT)
                sphgpu_sound_speed = sqrt(c%mat(k)%K * c%mat(k)%gamma / c%mat(k)%rho0)
            case (SPHGPU_EOS_LINEAR)
                sphgpu_sound_speed = sqrt(c%mat(k)%K)
            end select
        end do
    end function sphgpu_sound_speed

    !! Per-nick constants of the reference for the cases this library runs.
    !! `r0` is the wall-force range the reference takes on first entry of
    !! external_force for "dam_break" (externalForce.f90:84-87). Returns
    !! .false. for a nick whose constants are not defined here (the
    !! reference leaves several of them uninitialised, eos.f90:35-40).
    logical function sphgpu_preset(nick, dim, r0, c) result(ok)
        character(len=*), intent(in) :: nick
        integer, intent(in) :: dim
        real(8), intent(in) :: r0
        type(sphgpu_config), intent(inout) :: c
        ok = .true.
        c%n_materials = 0
        select case (nick)
        case ("dam_break")
            c%lj_factor = 10; c%lj_r0 = r0
            c%couple_delta = f32(0.2_8)
            !! water: Tait with c = 50 (eos.f90:37), eta = 1e-3 (viscosity.f90:18)
            call sphgpu_add_material(c, 2, SPHGPU_EOS_TAIT, 1000.0_8, &
                                     sphgpu_tait_b(50.0_8, 1000.0_8, 7.0_8), 7.0_8, f32(1e-3_8))
        case ("can_beam", "can_beam_3d")
            c%av_alpha = 0.5_8; c%av_beta = 0.5_8; c%av_psi = f32(0.1_8)
            if (dim == 2) c%as_delta = f32(1e-3_8)
            !! type 102: K = 2e7, rho0 = 1100 (eos.f90:219-226), G = 4.27e6 (viscosity.f90:24)
            call sphgpu_add_material(c, 102, SPHGPU_EOS_LINEAR, 1100.0_8, f32(2e7_8) / 1100.0_8, &
                                     1.0_8, f32(4.27e6_8))
        case default
            ok = .false.
        end select
    end function sphgpu_preset

end module sphgpu_m
