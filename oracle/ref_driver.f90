!! Test-infrastructure driver for the compiled reference (oracle/_ref).
!!
!! This program is the only new code linked into oracle/_ref/ref_driver; every
!! physics routine it calls is compiled unchanged from the reference sources
!! (src/**/*.f90 minus getValueFromToml.f90 and main.f90, which need toml-f).
!! It replaces main.f90 (src/main.f90:1-29) and the toml reader
!! (src/Initialize/getValueFromToml.f90:80-189): the Config/Field/Project
!! globals are filled from the values the case's sph.toml would give.
!!
!! Usage:  ref_driver <case> <mode> <nsteps> <out.bin> [in.bin]
!!   case : dam_break | can_beam | can_beam_3d | custom
!!   mode : init   case set-up plus generated dummy particles, state dump
!!          pairs  BGGS neighbour lists of the initial state
!!          deriv  one call of single_step (step 1), derivatives + state dump
!!          run    leap_frog for nsteps steps, state dump
!!          output the reference's writer on the set-up state (files under
!!                 ./output and ./vtk; out.bin is not written)
!!          bggs   (case custom) BGGS on particles read from in.bin
!!          extforce (case custom) BGGS, then external_force alone, on the
!!                 full particle state read from in.bin (custom state format)
!!   in.bin (optional, named cases, modes deriv/run): initial velocity override
!!   case custom with modes deriv / run / extforce reads a whole particle
!!   state and its switches from in.bin (see read_custom_state); deriv keeps
!!   the slots past ntotal as dummy sources, run (leap_frog, which counts no
!!   dummies unless it generates them) integrates real particles only.
!!
!! Output: a stream of records  name(char*8) kind(int32: 8=real64, 4=int32)
!!         count(int64) data(count values).
program ref_driver
    use ctrl_dict, only: Project, Config, Field
    use sph,       only: Particle, Update, allocateParticleList
    use input_m,   only: damBreak, cantileverBeam, cantileverBeam3D
    use dummy_part_m, only: gen_dummy_particle
    use APS_M,     only: BGGS
    use time_integration_m, only: time_integration, single_step
    use output_m,  only: output
    implicit none

    character(len=64)  :: case_name, mode, arg
    character(len=512) :: out_path, in_path
    integer :: nsteps, ntotal, ndummy, nbuffer, n, i, u, nargs
    type(Particle), allocatable :: P(:), Q(:)
    type(Update),   allocatable :: Delta(:)
    real(8), allocatable :: aver_v(:, :), Shear(:, :, :), dSdt(:, :, :)
    logical :: has_in

    nargs = command_argument_count()
    if (nargs < 4) error stop "usage: ref_driver case mode nsteps out.bin [in.bin]"
    call get_command_argument(1, case_name)
    call get_command_argument(2, mode)
    call get_command_argument(3, arg)
    read(arg, *) nsteps
    call get_command_argument(4, out_path)
    has_in = nargs >= 5
    if (has_in) call get_command_argument(5, in_path)

    call set_defaults()

    if (trim(case_name) == "custom") then
        if (trim(mode) == "bggs") then
            call run_custom_bggs()
        else
            call run_custom_state()
        end if
        stop
    end if

    call configure_case(trim(case_name))
    call allocateParticleList(P, Field%Maxn, Field%Dim, Field%pairNum)
    ntotal = 0
    select case (trim(case_name))
    case ("dam_break")
        call damBreak(ntotal, P)
    case ("can_beam")
        call cantileverBeam(ntotal, P)
    case ("can_beam_3d")
        call cantileverBeam3D(ntotal, P)
    case default
        error stop "unknown case"
    end select
    if (has_in) call read_velocity_override()
    call apply_dummy_sound_speed(ntotal)

    ndummy = 0
    nbuffer = 0
    select case (trim(mode))
    case ("init")
        if (Config%dummy_parti_w) call gen_dummy_particle(ntotal, ndummy, P)
        call dump_state(ntotal, ndummy)
    case ("output")
        !! the reference's own writer (output.f90:7-83) on the set-up state:
        !! output/Type_<t>_0.dat and vtk/Type_<t>_0.vtk under the working directory
        if (Config%dummy_parti_w) call gen_dummy_particle(ntotal, ndummy, P)
        Project%out_path = "output"
        Project%vtk_path = "vtk"
        call execute_command_line("mkdir -p output vtk")
        call output(0, P(1:ntotal+ndummy))
    case ("pairs")
        if (Config%dummy_parti_w) call gen_dummy_particle(ntotal, ndummy, P)
        call BGGS(P(1:ntotal), P(1:ntotal+ndummy), skipItsSelf=.true.)
        call dump_pairs(ntotal)
    case ("deriv")
        allocate(Delta(Field%Maxn))
        do i = 1, Field%Maxn
            Delta(i)%Density = 0
            Delta(i)%Energy  = 0
            allocate(Delta(i)%Velocity(Field%Dim), source=0._8)
        end do
        allocate(aver_v(Field%Dim, Field%Maxn), source=0._8)
        allocate(Shear(Field%Dim, Field%Dim, Field%Maxn), source=0._8)
        allocate(dSdt(Field%Dim, Field%Dim, Field%Maxn), source=0._8)
        Config%i_time_step = 1
        call single_step(ntotal, ndummy, nbuffer, P, Delta, aver_v, Shear, dSdt)
        call dump_state(ntotal, ndummy)
        call dump_delta(ntotal)
    case ("run")
        Config%max_time_step = nsteps
        call time_integration(ntotal, P)
        !! leap_frog keeps ndummy local; count the dummies on a scratch copy
        !! so that the dumped dummy pressures stay those of the last step
        if (Config%dummy_parti_w) then
            Q = P
            call gen_dummy_particle(ntotal, ndummy, Q)
        end if
        call dump_state(ntotal, ndummy)
    case default
        error stop "unknown mode"
    end select

contains

    subroutine set_defaults()
        Project%in_path = "."
        Project%out_path = "."
        Project%vtk_path = "."
        Project%project_name = "oracle"
        Project%nick = "unknown"
        Field%Dim = 2
        Field%Maxn = 5000
        Field%ntotal = 0
        Field%pairNum = 20
        Config%delta_t = 5e-4_8
        Config%max_time_step = 1
        Config%i_time_step = 0
        Config%print_interval = huge(0)
        Config%save_interval = huge(0)
        Config%monitor_particle = 1
        Config%pa_sph = 2
        Config%nnps = 2
        Config%sle = 0
        Config%skf = 1
        Config%sum_density_w = .true.
        Config%aver_velocity_w = .false.
        Config%read_ini_file_w = .false.
        Config%dummy_parti_w = .false.
        Config%dp_input = .false.
        Config%viscosity_w = .false.
        Config%gravity_w = .false.
        Config%ex_force_w = .false.
        Config%arti_visc_w = .false.
        Config%arti_heat_w = .false.
        Config%norm_dens_w = .false.
        Config%write_dp_w = .false.
        Config%write_vtk_w = .false.
        Config%write_dp_vtk_w = .false.
        Config%print_statistics_w = .false.
        Config%open_boundary_w = .false.
        Config%kernel_correciton_w = .false.
        Config%shifting_w = .false.
        Config%arti_stress_w = .false.
#ifdef _OPENMP
        Config%nthreads = 1
        Config%chunkSize = 1
#endif
    end subroutine set_defaults

    !! Values from data/<case>/sph.toml, with the reader's defaults
    !! (getValueFromToml.f90:99-155) for keys the file omits.
    subroutine configure_case(name)
        character(len=*), intent(in) :: name
        select case (name)
        case ("dam_break")         !! data/damBreak/sph.toml:1-37
            Project%nick = "dam_break"
            Field%Dim = 2; Field%Maxn = 25000; Field%pairNum = 40
            Config%pa_sph = 2; Config%skf = 1; Config%sle = 0
            Config%delta_t = 1e-3_8
            Config%sum_density_w = .false.
            Config%dummy_parti_w = .true.
            Config%viscosity_w = .true.
            Config%gravity_w = .true.
            Config%arti_visc_w = .true.
        case ("can_beam")          !! data/cantileverBeam/sph.toml:1-58
            Project%nick = "can_beam"
            Field%Dim = 2; Field%Maxn = 505; Field%pairNum = 60
            Config%pa_sph = 2; Config%skf = 1; Config%sle = 0
            Config%delta_t = 2e-6_8
            Config%sum_density_w = .false.
            Config%dummy_parti_w = .true.
            Config%viscosity_w = .true.
            Config%gravity_w = .true.
            Config%arti_visc_w = .true.
            Config%kernel_correciton_w = .true.
            Config%arti_stress_w = .false.
        case ("can_beam_3d")       !! data/cantileverBeam3D/sph.toml
            Project%nick = "can_beam_3d"
            Field%Dim = 3; Field%Maxn = 9045; Field%pairNum = 400
            Config%pa_sph = 2; Config%skf = 1; Config%sle = 0
            Config%delta_t = 2e-6_8
            Config%sum_density_w = .false.
            Config%dummy_parti_w = .true.
            Config%viscosity_w = .true.
            Config%gravity_w = .true.
            Config%arti_visc_w = .true.
            Config%kernel_correciton_w = .true.
        end select
        !! getValueFromToml.f90:152-154
        if (Config%dummy_parti_w .or. Config%gravity_w) Config%ex_force_w = .true.
        call apply_env_overrides()
    end subroutine configure_case

    !! Optional switches for variants of a shipped case (e.g. artificial
    !! stress on for the beam), read from the environment.
    subroutine apply_env_overrides()
        character(len=64) :: buf
        integer :: stat
        call get_environment_variable("REF_ARTI_STRESS", buf, status=stat)
        if (stat == 0 .and. trim(buf) == "1") Config%arti_stress_w = .true.
        call get_environment_variable("REF_SUM_DENSITY", buf, status=stat)
        if (stat == 0 .and. trim(buf) == "1") Config%sum_density_w = .true.
        call get_environment_variable("REF_NORM_DENSITY", buf, status=stat)
        if (stat == 0 .and. trim(buf) == "1") Config%norm_dens_w = .true.
        call get_environment_variable("REF_DT", buf, status=stat)
        if (stat == 0) read(buf, *) Config%delta_t
        call get_environment_variable("REF_SLE", buf, status=stat)
        if (stat == 0) read(buf, *) Config%sle
        call get_environment_variable("REF_PAIRNUM", buf, status=stat)
        if (stat == 0) read(buf, *) Field%pairNum
        call get_environment_variable("REF_PA_SPH", buf, status=stat)
        if (stat == 0) read(buf, *) Config%pa_sph
        call get_environment_variable("REF_SKF", buf, status=stat)
        if (stat == 0) read(buf, *) Config%skf
#ifdef _OPENMP
        call get_environment_variable("REF_THREADS", buf, status=stat)
        if (stat == 0) read(buf, *) Config%nthreads
#endif
    end subroutine apply_env_overrides

    !! REF_DUMMY_C: SoundSpeed of every slot past the real particles. The dummy
    !! generators never set it (dummyParticles.f90:544-659), so the value
    !! survives their per-step regeneration; the Riemann schemes (SPH = 3)
    !! divide by rho*c summed over a pair and need it nonzero.
    subroutine apply_dummy_sound_speed(nreal)
        integer, intent(in) :: nreal
        character(len=64) :: buf
        integer :: stat, i
        real(8) :: c
        call get_environment_variable("REF_DUMMY_C", buf, status=stat)
        if (stat /= 0) return
        read(buf, *) c
        do i = nreal + 1, size(P)
            P(i)%SoundSpeed = c
        end do
    end subroutine apply_dummy_sound_speed

    subroutine read_velocity_override()
        integer :: iu, d
        real(8), allocatable :: vv(:)
        allocate(vv(Field%Dim * ntotal))
        open(newunit=iu, file=trim(in_path), access="stream", form="unformatted", status="old")
        read(iu) vv
        close(iu)
        do i = 1, ntotal
            do d = 1, Field%Dim
                P(i)%v(d) = vv((i-1)*Field%Dim + d)
            end do
        end do
    end subroutine read_velocity_override

    !! Custom BGGS input: int32 dim, skf, ntotal, n, kpair; then real64
    !! x(dim*n) (particle-major), h(n); int32 state(n).
    subroutine run_custom_bggs()
        integer :: iu, d, hdr(5)
        real(8), allocatable :: xx(:), hh(:)
        integer, allocatable :: st(:)
        open(newunit=iu, file=trim(in_path), access="stream", form="unformatted", status="old")
        read(iu) hdr
        Field%Dim = hdr(1); Config%skf = hdr(2); ntotal = hdr(3); n = hdr(4)
        Field%pairNum = hdr(5); Field%Maxn = n
        allocate(xx(Field%Dim*n), hh(n), st(n))
        read(iu) xx, hh, st
        close(iu)
        call allocateParticleList(P, n, Field%Dim, Field%pairNum)
        do i = 1, n
            do d = 1, Field%Dim
                P(i)%x(d) = xx((i-1)*Field%Dim + d)
            end do
            P(i)%SmoothingLength = hh(i)
            P(i)%State = st(i)
        end do
        call BGGS(P(1:ntotal), P(1:n), skipItsSelf=.true.)
        call dump_pairs(ntotal)
    end subroutine run_custom_bggs

    !! Custom state input: int32 dim, skf, ntotal, n, kpair, flags, pa_sph,
    !! sle; char(16) nick; real64 dt; then real64 x(dim*n), v(dim*n) (both
    !! particle-major), mass(n), rho(n), p(n), e(n), c(n), h(n); int32
    !! type(n), state(n), boundary(n). flags: 1 viscosity, 2 gravity,
    !! 4 ex_force, 8 arti_visc, 16 kernel correction, 32 sum_density,
    !! 64 norm_dens, 128 arti_stress.
    subroutine read_custom_state()
        integer :: iu, d, hdr(8), flags
        character(len=16) :: nk
        real(8) :: dt
        real(8), allocatable :: xx(:), vv(:), ms(:), rh(:), pp(:), ee(:), cc(:), hh(:)
        integer, allocatable :: ty(:), st(:), bd(:)
        open(newunit=iu, file=trim(in_path), access="stream", form="unformatted", status="old")
        read(iu) hdr, nk, dt
        Field%Dim = hdr(1); Config%skf = hdr(2); ntotal = hdr(3); n = hdr(4)
        Field%pairNum = hdr(5); flags = hdr(6); Config%pa_sph = hdr(7); Config%sle = hdr(8)
        Field%Maxn = n
        Project%nick = trim(nk)
        Config%delta_t = dt
        Config%viscosity_w = iand(flags, 1) /= 0
        Config%gravity_w = iand(flags, 2) /= 0
        Config%ex_force_w = iand(flags, 4) /= 0
        Config%arti_visc_w = iand(flags, 8) /= 0
        Config%kernel_correciton_w = iand(flags, 16) /= 0
        Config%sum_density_w = iand(flags, 32) /= 0
        Config%norm_dens_w = iand(flags, 64) /= 0
        Config%arti_stress_w = iand(flags, 128) /= 0
        allocate(xx(Field%Dim*n), vv(Field%Dim*n), ms(n), rh(n), pp(n), ee(n), cc(n), hh(n))
        allocate(ty(n), st(n), bd(n))
        read(iu) xx, vv, ms, rh, pp, ee, cc, hh, ty, st, bd
        close(iu)
        call allocateParticleList(P, n, Field%Dim, Field%pairNum)
        do i = 1, n
            do d = 1, Field%Dim
                P(i)%x(d) = xx((i-1)*Field%Dim + d)
                P(i)%v(d) = vv((i-1)*Field%Dim + d)
            end do
            P(i)%Mass = ms(i); P(i)%Density = rh(i); P(i)%Pressure = pp(i)
            P(i)%InternalEnergy = ee(i); P(i)%SoundSpeed = cc(i); P(i)%SmoothingLength = hh(i)
            P(i)%Type = ty(i); P(i)%State = st(i); P(i)%Boundary = bd(i)
        end do
    end subroutine read_custom_state

    subroutine run_custom_state()
        use ex_force_m, only: external_force
        real(8), allocatable :: exdvdt(:, :)
        call read_custom_state()
        select case (trim(mode))
        case ("extforce")
            !! singleStep.f90:84 then :116: the pair list of this state, then
            !! the external force alone (it reads x, h, SoundSpeed and Type)
            call BGGS(P(1:ntotal), P(1:n), skipItsSelf=.true.)
            allocate(exdvdt(Field%Dim, n), source=0._8)
            call external_force(ntotal, P, exdvdt)
            open(newunit=u, file=trim(out_path), access="stream", form="unformatted", status="replace")
            call put_i4(u, "ntotal", [ntotal])
            call put_i4(u, "dim", [Field%Dim])
            call put_i4(u, "count", P(1:ntotal)%neighborNum)
            call put_r8(u, "exdvdt", reshape(exdvdt(:, 1:ntotal), [Field%Dim*ntotal]))
            close(u)
        case ("deriv")
            allocate(Delta(Field%Maxn))
            do i = 1, Field%Maxn
                Delta(i)%Density = 0
                Delta(i)%Energy  = 0
                allocate(Delta(i)%Velocity(Field%Dim), source=0._8)
            end do
            allocate(aver_v(Field%Dim, Field%Maxn), source=0._8)
            allocate(Shear(Field%Dim, Field%Dim, Field%Maxn), source=0._8)
            allocate(dSdt(Field%Dim, Field%Dim, Field%Maxn), source=0._8)
            Config%i_time_step = 1
            ndummy = n - ntotal
            nbuffer = 0
            call single_step(ntotal, ndummy, nbuffer, P, Delta, aver_v, Shear, dSdt)
            call dump_state(ntotal, n - ntotal)
            call dump_delta(ntotal)
        case ("run")
            if (n /= ntotal) error stop "custom run: leap_frog integrates real particles only (n == ntotal)"
            Config%max_time_step = nsteps
            call time_integration(ntotal, P)
            call dump_state(ntotal, 0)
        case default
            error stop "unknown mode for case custom"
        end select
    end subroutine run_custom_state

    subroutine put_r8(iu, name, a)
        integer, intent(in) :: iu
        character(len=*), intent(in) :: name
        real(8), intent(in) :: a(:)
        character(len=8) :: nm
        nm = name
        write(iu) nm, 8_4, int(size(a), 8), a
    end subroutine put_r8

    subroutine put_i4(iu, name, a)
        integer, intent(in) :: iu
        character(len=*), intent(in) :: name
        integer, intent(in) :: a(:)
        character(len=8) :: nm
        nm = name
        write(iu) nm, 4_4, int(size(a), 8), a
    end subroutine put_i4

    subroutine dump_state(nreal, ndum)
        integer, intent(in) :: nreal, ndum
        integer :: m, d, dd, k
        real(8), allocatable :: vec(:), ten(:)
        m = nreal + ndum
        open(newunit=u, file=trim(out_path), access="stream", form="unformatted", status="replace")
        call put_i4(u, "ntotal", [nreal])
        call put_i4(u, "ndummy", [ndum])
        call put_i4(u, "dim", [Field%Dim])
        allocate(vec(Field%Dim*m), ten(Field%Dim*Field%Dim*m))
        do i = 1, m
            do d = 1, Field%Dim
                vec((i-1)*Field%Dim + d) = P(i)%x(d)
            end do
        end do
        call put_r8(u, "x", vec)
        do i = 1, m
            do d = 1, Field%Dim
                vec((i-1)*Field%Dim + d) = P(i)%v(d)
            end do
        end do
        call put_r8(u, "v", vec)
        do i = 1, m
            do d = 1, Field%Dim
                vec((i-1)*Field%Dim + d) = P(i)%Displacement(d)
            end do
        end do
        call put_r8(u, "disp", vec)
        k = 0
        do i = 1, m
            do d = 1, Field%Dim
                do dd = 1, Field%Dim
                    k = k + 1
                    ten(k) = P(i)%Stress(d, dd)
                end do
            end do
        end do
        call put_r8(u, "stress", ten)
        call put_r8(u, "mass", P(1:m)%Mass)
        call put_r8(u, "rho", P(1:m)%Density)
        call put_r8(u, "p", P(1:m)%Pressure)
        call put_r8(u, "e", P(1:m)%InternalEnergy)
        call put_r8(u, "c", P(1:m)%SoundSpeed)
        call put_r8(u, "h", P(1:m)%SmoothingLength)
        call put_r8(u, "eta", P(1:m)%Viscosity)
        call put_r8(u, "divx", P(1:m)%divergencePosition)
        call put_r8(u, "divv", P(1:m)%divergenceVelocity)
        call put_i4(u, "type", P(1:m)%Type)
        call put_i4(u, "state", P(1:m)%State)
        call put_i4(u, "boundary", P(1:m)%Boundary)
        close(u)
    end subroutine dump_state

    subroutine dump_delta(nreal)
        integer, intent(in) :: nreal
        integer :: d, dd, k
        real(8), allocatable :: vec(:), ten(:)
        open(newunit=u, file=trim(out_path), access="stream", form="unformatted", &
             status="old", position="append")
        allocate(vec(Field%Dim*nreal), ten(Field%Dim*Field%Dim*nreal))
        do i = 1, nreal
            do d = 1, Field%Dim
                vec((i-1)*Field%Dim + d) = Delta(i)%Velocity(d)
            end do
        end do
        call put_r8(u, "dvdt", vec)
        call put_r8(u, "drhodt", Delta(1:nreal)%Density)
        call put_r8(u, "dedt", Delta(1:nreal)%Energy)
        k = 0
        do i = 1, nreal
            do d = 1, Field%Dim
                do dd = 1, Field%Dim
                    k = k + 1
                    ten(k) = dSdt(d, dd, i)
                end do
            end do
        end do
        call put_r8(u, "dsdt", ten)
        close(u)
    end subroutine dump_delta

    subroutine dump_pairs(nreal)
        integer, intent(in) :: nreal
        integer :: k, d, total
        integer, allocatable :: cnt(:), nbr(:)
        real(8), allocatable :: ww(:), dw(:)
        allocate(cnt(nreal))
        total = 0
        do i = 1, nreal
            cnt(i) = P(i)%neighborNum
            total = total + cnt(i)
        end do
        allocate(nbr(total), ww(total), dw(total*Field%Dim))
        total = 0
        do i = 1, nreal
            do k = 1, cnt(i)
                total = total + 1
                nbr(total) = P(i)%neighborList(k)
                ww(total) = P(i)%w(k)
                do d = 1, Field%Dim
                    dw((total-1)*Field%Dim + d) = P(i)%dwdx(d, k)
                end do
            end do
        end do
        open(newunit=u, file=trim(out_path), access="stream", form="unformatted", status="replace")
        call put_i4(u, "ntotal", [nreal])
        call put_i4(u, "dim", [Field%Dim])
        call put_i4(u, "count", cnt)
        call put_i4(u, "nbr", nbr)
        call put_r8(u, "w", ww)
        call put_r8(u, "dwdx", dw)
        close(u)
    end subroutine dump_pairs

end program ref_driver
