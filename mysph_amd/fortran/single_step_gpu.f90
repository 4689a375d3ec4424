!! Drop-in replacement of the reference's external subroutine `single_step`
!! (src/TimeIntegration/singleStep.f90:1-211) that evaluates the step's
!! derivatives on the GPU through libsphgpu.
!!
!! It has the reference definition's dummy arguments (the explicit interface
!! leap_frog calls through, timeIntegration.f90:7-19), so it links in place
!! of singleStep.f90's object with no change to the caller. Per call:
!!   1. dummy particles are regenerated by the reference's own
!!      gen_dummy_particle, as singleStep.f90:57 does;
!!   2. Particles(1:ntotal+ndummy) are packed into structure-of-arrays
!!      buffers and uploaded (the derived type with allocatable components is
!!      not C-interoperable);
!!   3. sphgpu_derivatives runs BGGS, density, EOS, internal/external and
!!      artificial forces and the fixed boundary (singleStep.f90:83-142);
!!   4. Delta, dSdt and the fields the reference updates in place (Pressure,
!!      Density under summation density, Stress, Viscosity, the divergences,
!!      SmoothingLength when it evolves) are written back.
!! Neighbour lists stay on the device, so Particles%neighborNum/List/w/dwdx
!! remain zero, as the reference leaves them on return (:199-204).
!!
!! Switches the GPU path does not implement (open boundary, artificial heat,
!! shifting, averaged velocity, the DSPH density scheme SPH=4)
!! stop with a message instead of running something else.
module single_step_gpu_m
    use, intrinsic :: iso_c_binding
    use sphgpu_m
    implicit none
    private
    type(c_ptr), save, public :: gpu_ctx = c_null_ptr
    integer(c_int64_t), save :: gpu_capacity = 0
    real(c_double), allocatable, target, save, public :: sx(:), sv(:), smass(:), srho(:), sp(:), &
        se(:), sc(:), sh(:), sshear(:), sout(:)
    integer(c_int32_t), allocatable, target, save, public :: stype(:), sstate(:), sbnd(:)
    public :: single_step_gpu_open, single_step_gpu_close
contains

    subroutine single_step_gpu_open(cfg, capacity, dim)
        type(sphgpu_config), intent(in) :: cfg
        integer(c_int64_t), intent(in) :: capacity
        integer, intent(in) :: dim
        if (c_associated(gpu_ctx)) return
        call sphgpu_check(sphgpu_create(cfg, capacity, gpu_ctx), gpu_ctx, "sphgpu_create")
        gpu_capacity = capacity
        allocate(sx(dim * capacity), sv(dim * capacity), smass(capacity), srho(capacity), &
                 sp(capacity), se(capacity), sc(capacity), sh(capacity), &
                 sshear(dim * dim * capacity), sout(dim * dim * capacity))
        allocate(stype(capacity), sstate(capacity), sbnd(capacity))
    end subroutine single_step_gpu_open

    subroutine single_step_gpu_close()
        if (.not. c_associated(gpu_ctx)) return
        call sphgpu_destroy(gpu_ctx)
        gpu_ctx = c_null_ptr
        deallocate(sx, sv, smass, srho, sp, se, sc, sh, sshear, sout, stype, sstate, sbnd)
    end subroutine single_step_gpu_close

end module single_step_gpu_m


subroutine single_step(ntotal, ndummy, nbuffer, Particles, Delta, aver_v, Shear, dSdt)
    use, intrinsic :: iso_c_binding
    use ctrl_dict,    only: Config, Field, Project
    use SPH,          only: Particle, Update
    use dummy_part_m, only: gen_dummy_particle
    use sphgpu_m
    use single_step_gpu_m
    implicit none
    integer,        intent(inout) :: ntotal
    integer,        intent(inout) :: ndummy, nbuffer
    type(Particle), intent(inout) :: Particles(:)
    type(Update),   intent(inout) :: Delta(:)
    real(8),        intent(inout) :: aver_v(:, :)
    real(8), intent(in),    optional :: Shear(:, :, :)
    real(8), intent(inout), optional :: dSdt(:, :, :)

    type(sphgpu_config) :: cfg
    integer :: N, dim, i, a, b
    integer(c_int64_t) :: st(4)
    type(c_ptr) :: pshear

    dim = Field%Dim
    if (Config%dummy_parti_w) call gen_dummy_particle(ntotal, ndummy, Particles)
    if (Config%open_boundary_w) error stop "single_step (GPU): open boundaries are not supported"
    nbuffer = 0
    N = ntotal + ndummy

    if (.not. c_associated(gpu_ctx)) then
        if (Config%pa_sph == 4) error stop "single_step (GPU): SPH = 4 (DSPH) is not supported"
        if (Config%arti_heat_w .or. Config%shifting_w .or. Config%aver_velocity_w) &
            error stop "single_step (GPU): artificial heat, shifting and averaged velocity are not supported"
        cfg = sphgpu_base_config(dim)
        !! externalForce.f90:84-87 takes the wall range from particle 2 on first entry
        if (.not. sphgpu_preset(Project%nick, dim, abs(Particles(2)%x(1) - Particles(2)%x(2)), cfg)) &
            error stop "single_step (GPU): no constants for this case nick"
        cfg%skf = Config%skf
        cfg%pa_sph = Config%pa_sph
        cfg%summation_density = merge(1, 0, Config%sum_density_w)
        cfg%normalize_density = merge(1, 0, Config%norm_dens_w)
        cfg%viscosity = merge(1, 0, Config%viscosity_w)
        cfg%gravity = merge(1, 0, Config%gravity_w)
        cfg%external_force = merge(1, 0, Config%ex_force_w)
        cfg%artificial_viscosity = merge(1, 0, Config%arti_visc_w)
        cfg%artificial_stress = merge(1, 0, Config%arti_stress_w)
        cfg%kernel_correction = merge(1, 0, Config%kernel_correciton_w)
        cfg%dt = Config%delta_t
        cfg%sle = Config%sle
        cfg%precision = SPHGPU_FP64
        call single_step_gpu_open(cfg, int(size(Particles), c_int64_t), dim)
    end if

    !! pack: vectors particle-major, tensors [i][a][b] = Stress(a, b)
    do i = 1, N
        do a = 1, dim
            sx((i-1)*dim + a) = Particles(i)%x(a)
            sv((i-1)*dim + a) = Particles(i)%v(a)
        end do
        smass(i) = Particles(i)%Mass
        srho(i)  = Particles(i)%Density
        sp(i)    = Particles(i)%Pressure
        se(i)    = Particles(i)%InternalEnergy
        sc(i)    = Particles(i)%SoundSpeed
        sh(i)    = Particles(i)%SmoothingLength
        stype(i)  = Particles(i)%Type
        sstate(i) = Particles(i)%State
        sbnd(i)   = Particles(i)%Boundary
        if (present(Shear)) then
            do a = 1, dim
                do b = 1, dim
                    sshear((i-1)*dim*dim + (a-1)*dim + b) = Shear(a, b, i)
                end do
            end do
        end if
    end do
    !! without Shear the library's deviatoric stress starts at zero
    pshear = c_null_ptr
    if (present(Shear)) pshear = c_loc(sshear)
    call sphgpu_check(sphgpu_upload(gpu_ctx, int(ntotal, c_int64_t), int(ndummy, c_int64_t), &
                      c_loc(sx), c_loc(sv), c_loc(smass), c_loc(srho), c_loc(sp), c_loc(se), &
                      c_loc(sc), c_loc(sh), pshear, c_loc(stype), c_loc(sstate), c_loc(sbnd)), &
                      gpu_ctx, "sphgpu_upload")
    !! Shepard normalisation reads the previous step's divergencePosition
    !! (density.f90:25-37, :67-76); an upload starts it at zero
    if (Config%sum_density_w .and. Config%norm_dens_w) then
        do i = 1, N
            sout(i) = Particles(i)%divergencePosition
        end do
        call sphgpu_check(sphgpu_upload_field(gpu_ctx, SPHGPU_F_DIVX, c_loc(sout)), gpu_ctx, &
                          "sphgpu_upload_field divx")
    end if
    call sphgpu_check(sphgpu_derivatives(gpu_ctx), gpu_ctx, "sphgpu_derivatives")

    !! Delta = in + av + ex + as (singleStep.f90:139-142), after the fixed boundary
    call sphgpu_check(sphgpu_download(gpu_ctx, SPHGPU_F_DVDT, c_loc(sout)), gpu_ctx, "download dvdt")
    do i = 1, ntotal
        do a = 1, dim
            Delta(i)%Velocity(a) = sout((i-1)*dim + a)
        end do
    end do
    call sphgpu_check(sphgpu_download(gpu_ctx, SPHGPU_F_DEDT, c_loc(sout)), gpu_ctx, "download dedt")
    do i = 1, ntotal
        Delta(i)%Energy = sout(i)
    end do
    if (Config%sum_density_w) then
        call sphgpu_check(sphgpu_download(gpu_ctx, SPHGPU_F_RHO, c_loc(sout)), gpu_ctx, "download rho")
        do i = 1, ntotal
            Particles(i)%Density = sout(i)
        end do
    else
        call sphgpu_check(sphgpu_download(gpu_ctx, SPHGPU_F_DRHODT, c_loc(sout)), gpu_ctx, "download drhodt")
        do i = 1, ntotal
            Delta(i)%Density = sout(i)
        end do
    end if
    if (present(dSdt)) then
        call sphgpu_check(sphgpu_download(gpu_ctx, SPHGPU_F_DSDT, c_loc(sout)), gpu_ctx, "download dsdt")
        do i = 1, ntotal
            do a = 1, dim
                do b = 1, dim
                    dSdt(a, b, i) = sout((i-1)*dim*dim + (a-1)*dim + b)
                end do
            end do
        end do
    end if
    !! in-place updates of internal_force (EOS over every slot) and divergence
    call sphgpu_check(sphgpu_download(gpu_ctx, SPHGPU_F_P, c_loc(sout)), gpu_ctx, "download p")
    do i = 1, N
        Particles(i)%Pressure = sout(i)
    end do
    if (Config%viscosity_w) then      !! internalForce.f90:45
        call sphgpu_check(sphgpu_download(gpu_ctx, SPHGPU_F_ETA, c_loc(sout)), gpu_ctx, "download eta")
        do i = 1, N
            Particles(i)%Viscosity = sout(i)
        end do
    end if
    !! Stress = Shear - p I for solids only (internalForce.f90:175-180)
    if (present(Shear)) then
        call sphgpu_check(sphgpu_download(gpu_ctx, SPHGPU_F_STRESS, c_loc(sout)), gpu_ctx, "download stress")
        do i = 1, N
            if (abs(Particles(i)%Type) <= 100) cycle
            do a = 1, dim
                do b = 1, dim
                    Particles(i)%Stress(a, b) = sout((i-1)*dim*dim + (a-1)*dim + b)
                end do
            end do
        end do
    end if
    call sphgpu_check(sphgpu_download(gpu_ctx, SPHGPU_F_DIVX, c_loc(sout)), gpu_ctx, "download divx")
    do i = 1, ntotal
        Particles(i)%divergencePosition = sout(i)
    end do
    call sphgpu_check(sphgpu_download(gpu_ctx, SPHGPU_F_DIVV, c_loc(sout)), gpu_ctx, "download divv")
    do i = 1, ntotal
        Particles(i)%divergenceVelocity = sout(i)
    end do
    !! h_upgrade (singleStep.f90:136, hsml.f90:8-38) ran on the GPU
    if (Config%sle /= 0) then
        call sphgpu_check(sphgpu_download(gpu_ctx, SPHGPU_F_H, c_loc(sout)), gpu_ctx, "download h")
        do i = 1, ntotal
            Particles(i)%SmoothingLength = sout(i)
        end do
    end if

    if (mod(Config%i_time_step, Config%print_interval) == 0 .and. Config%print_statistics_w) then
        call sphgpu_check(sphgpu_neighbour_stats(gpu_ctx, st), gpu_ctx, "sphgpu_neighbour_stats")
        write(*, "(A)") " >> Statistics: number of neighbors per particle (GPU search):"
        write(*, "(A, I0)") "    maximal neighbors: ", st(2)
        write(*, "(A, I0)") "    minimal neighbors: ", st(3)
        write(*, "(A, G0)") "    Average number of neighbors: ", real(st(1)) / max(ntotal, 1)
        write(*, "(A, I0)") "    Number of particles without neighbors: ", st(4)
    end if
end subroutine single_step
