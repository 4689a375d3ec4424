!! mysph_host: Fortran host program of the GPU SPH solver.
!!
!! Usage:  mysph_host <project_dir> [--setup-only <file>] [--final <file>]
!!                    [--no-output]
!!
!! Like the reference program (src/main.f90:1-29 with
!! getValueFromToml.f90:17-78), it reads <project_dir>/sph.toml, builds the
!! case, runs maxTimeStep leap-frog steps and writes Type_<type>_<step> .dat
!! files to <project_dir>/output and .vtk files to <project_dir>/vtk every
!! saveInterval steps. The particle state stays on the GPU between outputs;
!! the host only chooses the cadence (sphgpu_advance) and downloads fields to
!! write them.
!!
!! sph.toml keys read (the reference's, with its defaults):
!!   name, nick
!!   [Parameter] DIM, SKF, deltaT, maxTimeStep (1000), saveInterval (500),
!!     printInterval (100), printStatistics (true), SPH (2), SLE (0), summationDensity,
!!     normalizeDensity, viscosity, gravity, externalForce,
!!     artificialViscosity, artificialStress, kernelCorrection
!!   [postProcess] writeVTKfile (true), writeDummyParticle (true)
!! and the case description, which the reference keeps in code:
!!   no [Case] table: the reference's own case for the nick (dam_break,
!!     can_beam, can_beam_3d), with every [Parameter] switch read as the
!!     reference reads it (its defaults where the file is silent); another
!!     nick stops the program
!!   [Case] kind = "reference" | "dam_break" | "file", precision = "fp64" | "mixed"
!!     dam_break: dx, column = [..], tank = [..], layers (3), h_over_dx (1.2),
!!                jitter (0.01), alpha (0.1), seed (20261203)
!!     file: file = "<particle file>" (records as write_dump writes them);
!!           the physics constants then come from the nick preset.
!! Switches in [Parameter] override the case's defaults when present.
!!
!! --setup-only <file>: build the case, write it to <file> (and the step-0
!! output files unless --no-output) and stop without touching the GPU.
!! --final <file>: write the final state to <file>.
!! --checkpoint <file>: write an exact checkpoint of the run at the end
!! (sphgpu_checkpoint_save).
!! --resume <file>: continue from such a checkpoint (the case is still built
!! from sph.toml for its particle types and counts, which must match) up to
!! maxTimeStep; the result equals an uninterrupted run bit for bit.
!! --abi: print the library's ABI version and the size of sphgpu_config as
!! this binding lays it out, then stop.
program mysph_host
    use, intrinsic :: iso_c_binding
    use, intrinsic :: iso_fortran_env, only: int32, int64
    use sphgpu_m
    use toml_lite_m
    use cases_m
    use output_m
    implicit none

    interface
        integer(c_int) function c_mkdir(path, mode) bind(C, name="mkdir")
            import :: c_int, c_char
            character(kind=c_char), intent(in) :: path(*)
            integer(c_int), value :: mode
        end function c_mkdir
    end interface

    character(len=1024) :: arg
    character(len=:), allocatable :: project_dir, setup_only, final_path, msg, name, nick, kind, &
                                     prec, out_dir, vtk_dir, file, ckpt_out, ckpt_in
    logical :: ok, no_output, write_vtk, write_dummy, print_stats, report
    integer :: i, nargs, dim, layers, nc, nt, max_step, save_every, print_every, step, chunk, n
    integer(int64) :: seed, t0, t1, rate, st4(4)
    real(8) :: dx, h_over_dx, jitter, alpha, column(3), tank(3), secs, adv_ms, force_ms
    integer(c_int64_t) :: launches, nbytes, counts(3)
    integer(c_int8_t), allocatable, target :: blob(:)
    type(sphgpu_config) :: cfg_ck
    type(particle_set), target :: ps
    type(sphgpu_config) :: cfg
    type(c_ptr) :: ctx
    type(frame_t) :: frame

    project_dir = "."
    setup_only = ""
    final_path = ""
    ckpt_out = ""
    ckpt_in = ""
    no_output = .false.
    nargs = command_argument_count()
    i = 1
    do while (i <= nargs)
        call get_command_argument(i, arg)
        select case (trim(arg))
        case ("--setup-only")
            i = i + 1
            call get_command_argument(i, arg)
            setup_only = trim(arg)
        case ("--final")
            i = i + 1
            call get_command_argument(i, arg)
            final_path = trim(arg)
        case ("--checkpoint")
            i = i + 1
            call get_command_argument(i, arg)
            ckpt_out = trim(arg)
        case ("--resume")
            i = i + 1
            call get_command_argument(i, arg)
            ckpt_in = trim(arg)
        case ("--no-output")
            no_output = .true.
        case ("--abi")
            write(*, "(A, I0, A, I0)") " abi ", sphgpu_abi_version(), " config_bytes ", c_sizeof(cfg)
            stop
        case default
            project_dir = trim(arg)
        end select
        i = i + 1
    end do

    call toml_load(project_dir // "/sph.toml", ok, msg)
    if (.not. ok) then
        write(*, "(A)") " mysph_host: " // msg
        error stop "cannot read sph.toml"
    end if
    name = toml_string("name", "untitled")
    nick = toml_string("nick", "unknown")
    dim = toml_int("Parameter.DIM", 2)
    if (dim /= 2 .and. dim /= 3) error stop "mysph_host: DIM must be 2 or 3"
    !! a reference sph.toml has no [Case] table: the case is the one the
    !! reference builds in code for its nick (input.f90:1031-1405); a nick
    !! this host cannot build is refused rather than replaced
    if (toml_has("Case.kind")) then
        kind = toml_string("Case.kind", "")
    else if (is_reference_case(nick)) then
        kind = "reference"
    else
        write(*, "(A)") " mysph_host: sph.toml has no [Case] table and nick '" // nick // &
            "' is not a case this host builds (dam_break, can_beam, can_beam_3d);" // &
            " describe it with [Case] kind = ""dam_break"" or ""file"""
        error stop "mysph_host: unknown case"
    end if
    prec = toml_string("Case.precision", "fp64")
    write(*, "(A)") " Project: " // name // " (" // nick // "), case " // kind // ", " // prec

    select case (kind)
    case ("reference")
        if (.not. is_reference_case(nick)) error stop "mysph_host: kind = reference needs a reference nick"
        call case_reference(nick, ps, cfg)
        if (ps%dim /= dim) error stop "mysph_host: DIM differs from the reference case's dimension"
        cfg%precision = precision_code(prec)
    case ("dam_break")
        dx = toml_real("Case.dx", 0.02_8)
        if (dim == 2) then
            column(1:2) = [1.0_8, 2.0_8]
            tank(1:2) = [4.0_8, 3.0_8]
        else
            column = [0.6_8, 0.6_8, 1.2_8]
            tank = [1.6_8, 0.6_8, 1.3_8]
        end if
        call toml_reals("Case.column", column, nc)
        if (nc /= 0 .and. nc /= dim) error stop "mysph_host: Case.column needs DIM values"
        call toml_reals("Case.tank", tank, nt)
        if (nt /= 0 .and. nt /= dim) error stop "mysph_host: Case.tank needs DIM values"
        layers = toml_int("Case.layers", 3)
        h_over_dx = toml_real("Case.h_over_dx", 1.2_8)
        jitter = toml_real("Case.jitter", 0.01_8)
        alpha = toml_real("Case.alpha", 0.1_8)
        seed = int(toml_int("Case.seed", int(SEED_JITTER)), int64)
        call case_dam_break(dim, dx, column(1:dim), tank(1:dim), layers, h_over_dx, jitter, alpha, &
                            seed, precision_code(prec), ps, cfg)
    case ("file")
        file = toml_string("Case.file", "")
        if (file(1:1) /= "/") file = project_dir // "/" // file
        call read_set(file, ps, ok, msg)
        if (.not. ok) then
            write(*, "(A)") " mysph_host: " // msg
            error stop "cannot read the particle file"
        end if
        if (ps%dim /= dim) error stop "mysph_host: particle file dimension differs from DIM"
        cfg = sphgpu_base_config(dim)
        !! externalForce.f90:84-87 takes the wall range from particle 2
        if (.not. sphgpu_preset(nick, dim, abs(ps%x(1, 2) - ps%x(2, 2)), cfg)) &
            error stop "mysph_host: no constants for this nick"
        cfg%precision = precision_code(prec)
    case default
        error stop "mysph_host: Case.kind must be reference, dam_break or file"
    end select
    call refuse_unsupported()
    if (kind == "reference") then
        call reference_switches(cfg)
    else
        call apply_switches(cfg)
    end if

    max_step = toml_int("Parameter.maxTimeStep", 1000)
    save_every = toml_int("Parameter.saveInterval", 500)
    print_every = toml_int("Parameter.printInterval", 100)
    print_stats = toml_bool("Parameter.printStatistics", .true.)
    write_vtk = toml_bool("postProcess.writeVTKfile", .true.)
    write_dummy = toml_bool("postProcess.writeDummyParticle", .true.)
    n = ps%n_real + ps%n_dummy
    write(*, "(A, I0, A, I0, A, ES12.5)") " Particles: ", ps%n_real, " real + ", ps%n_dummy, &
        " fixed boundary; dt = ", cfg%dt

    out_dir = project_dir // "/output"
    vtk_dir = project_dir // "/vtk"
    if (.not. no_output) then
        call make_dir(out_dir)
        if (write_vtk) call make_dir(vtk_dir)
    end if

    if (len(setup_only) > 0) then
        call write_dump(setup_only, ps)
        if (.not. no_output) then
            call frame_from_set(ps, frame)
            call write_output(out_dir, vtk_dir, 0, frame, write_vtk, write_dummy)
        end if
        write(*, "(A)") " Set-up written to " // setup_only
        stop
    end if

    step = 0
    if (len(ckpt_in) > 0) then
        call read_blob(ckpt_in, blob)
        call sphgpu_check(sphgpu_checkpoint_restore(c_loc(blob), int(size(blob), c_int64_t), ctx), &
                          c_null_ptr, "sphgpu_checkpoint_restore")
        deallocate(blob)
        call sphgpu_check(sphgpu_context_info(ctx, cfg_ck, counts), ctx, "sphgpu_context_info")
        if (counts(1) /= ps%n_real .or. counts(2) /= ps%n_dummy) &
            error stop "mysph_host: the checkpoint holds another particle set than sph.toml builds"
        step = int(sphgpu_step(ctx)) - 1
        write(*, "(A, I0)") " Resumed from " // ckpt_in // " after step ", step
    else
        call sphgpu_check(sphgpu_create(cfg, int(n, c_int64_t), ctx), ctx, "sphgpu_create")
        call sphgpu_check(sphgpu_upload(ctx, int(ps%n_real, c_int64_t), int(ps%n_dummy, c_int64_t), &
                          c_loc(ps%x), c_loc(ps%v), c_loc(ps%mass), c_loc(ps%rho), c_loc(ps%p), &
                          c_loc(ps%e), c_loc(ps%c), c_loc(ps%h), c_null_ptr, c_loc(ps%type), &
                          c_loc(ps%state), c_loc(ps%boundary)), ctx, "sphgpu_upload")
    end if
    call frame_allocate(frame, dim, n)
    if (.not. no_output .and. save_every > 0 .and. step == 0) then
        call download(ctx, frame)
        call write_output(out_dir, vtk_dir, 0, frame, write_vtk, write_dummy)
    end if

    call system_clock(t0, rate)
    do while (step < max_step)
        chunk = max_step - step
        if (print_every > 0) chunk = min(chunk, print_every - mod(step, print_every))
        if (save_every > 0) chunk = min(chunk, save_every - mod(step, save_every))
        call sphgpu_check(sphgpu_advance(ctx, int(chunk, c_int32_t)), ctx, "sphgpu_advance")
        step = step + chunk
        !! Fortran does not short-circuit .and.: mod by zero would trap
        report = .false.
        if (print_every > 0) report = mod(step, print_every) == 0
        if (report) then
            call sphgpu_check(sphgpu_timing(ctx, adv_ms, force_ms, launches), ctx, "sphgpu_timing")
            write(*, "(A, I0, A, I0, A, F9.3, A, F8.3, A)") " step ", step, " / ", max_step, &
                ": ", adv_ms / chunk, " ms/step (force kernel ", force_ms, " ms)"
            if (print_stats) then
                call sphgpu_check(sphgpu_neighbour_stats(ctx, st4), ctx, "sphgpu_neighbour_stats")
                write(*, "(A, I0, A, I0, A, G0, A, I0)") "   neighbours: max ", st4(2), ", min ", &
                    st4(3), ", average ", real(st4(1)) / max(ps%n_real, 1), ", without ", st4(4)
            end if
        end if
        if (.not. no_output .and. save_every > 0) then
            if (mod(step, save_every) == 0) then
                call download(ctx, frame)
                call write_output(out_dir, vtk_dir, step, frame, write_vtk, write_dummy)
            end if
        end if
    end do
    call system_clock(t1)
    secs = real(t1 - t0, 8) / real(rate, 8)
    write(*, "(A, I0, A, F10.3, A, ES12.5, A)") " Done: ", max_step, " steps in ", secs, " s, ", &
        real(ps%n_real, 8) * max_step / max(secs, 1e-9_8), " particle-steps/s (incl. output)"

    if (len(final_path) > 0) then
        call download(ctx, frame)
        ps%x = frame%x
        ps%v = frame%v
        ps%rho = frame%rho
        ps%p = frame%p
        ps%e = frame%e
        call write_dump(final_path, ps)
    end if
    if (len(ckpt_out) > 0) then
        call sphgpu_check(sphgpu_checkpoint_size(ctx, nbytes), ctx, "sphgpu_checkpoint_size")
        allocate(blob(nbytes))
        call sphgpu_check(sphgpu_checkpoint_save(ctx, c_loc(blob), nbytes), ctx, "sphgpu_checkpoint_save")
        call write_blob(ckpt_out, blob)
        write(*, "(A, I0, A)") " Checkpoint written to " // ckpt_out // " (", nbytes, " bytes)"
    end if
    call sphgpu_destroy(ctx)

contains

    subroutine write_blob(path, b)
        character(len=*), intent(in) :: path
        integer(c_int8_t), intent(in) :: b(:)
        integer :: u
        open(newunit=u, file=path, access="stream", form="unformatted", status="replace", action="write")
        write(u) b
        close(u)
    end subroutine write_blob

    subroutine read_blob(path, b)
        character(len=*), intent(in) :: path
        integer(c_int8_t), allocatable, intent(out) :: b(:)
        integer :: u, ios
        integer(int64) :: sz
        open(newunit=u, file=path, access="stream", form="unformatted", status="old", action="read", &
             iostat=ios)
        if (ios /= 0) error stop "mysph_host: cannot open the checkpoint file"
        inquire(unit=u, size=sz)
        allocate(b(sz))
        read(u) b
        close(u)
    end subroutine read_blob

    integer function precision_code(p)
        character(len=*), intent(in) :: p
        select case (p)
        case ("fp64")
            precision_code = SPHGPU_FP64
        case ("mixed")
            precision_code = SPHGPU_MIXED
        case default
            error stop "mysph_host: Case.precision must be fp64 or mixed"
        end select
    end function precision_code

    !! The reference's own reading of [Parameter] (getValueFromToml.f90:
    !! 117-154): every switch takes the file's value or the reference's
    !! default, whatever the case would otherwise set.
    subroutine reference_switches(c)
        type(sphgpu_config), intent(inout) :: c
        logical :: dummy
        c%pa_sph = toml_int("Parameter.SPH", 2)
        c%skf = toml_int("Parameter.SKF", 1)
        c%sle = toml_int("Parameter.SLE", 0)
        c%dt = toml_real("Parameter.deltaT", f32(5e-4_8))      !! dble(5e-4)
        c%summation_density = merge(1, 0, toml_bool("Parameter.summationDensity", .true.))
        c%normalize_density = merge(1, 0, toml_bool("Parameter.normalizeDensity", .false.))
        c%viscosity = merge(1, 0, toml_bool("Parameter.viscosity", .false.))
        c%gravity = merge(1, 0, toml_bool("Parameter.gravity", .false.))
        c%artificial_viscosity = merge(1, 0, toml_bool("Parameter.artificialViscosity", .false.))
        c%artificial_stress = merge(1, 0, toml_bool("Parameter.artificialStress", .false.))
        c%kernel_correction = merge(1, 0, toml_bool("Parameter.kernelCorrection", .false.))
        dummy = toml_bool("Parameter.dummyParticle", .false.)
        c%external_force = merge(1, 0, toml_bool("Parameter.externalForce", .false.) .or. dummy &
                                       .or. c%gravity /= 0)
    end subroutine reference_switches

    !! Switches the reference reads that this solver does not implement stop
    !! the run instead of being ignored.
    subroutine refuse_unsupported()
        character(len=32), parameter :: keys(6) = [character(len=32) :: "averageVelocity", &
            "artificialHeat", "openBoundary", "shifting", "readInitialFile", "dpInput"]
        integer :: k
        do k = 1, size(keys)
            if (toml_bool("Parameter." // trim(keys(k)), .false.)) then
                write(*, "(A)") " mysph_host: [Parameter] " // trim(keys(k)) // &
                    " = true is not supported by the GPU solver"
                error stop "mysph_host: unsupported switch"
            end if
        end do
    end subroutine refuse_unsupported

    !! [Parameter] switches of the reference (getValueFromToml.f90:117-154)
    !! over the case's own settings, where the file gives them
    subroutine apply_switches(c)
        type(sphgpu_config), intent(inout) :: c
        if (toml_has("Parameter.SKF")) c%skf = toml_int("Parameter.SKF", 1)
        if (toml_has("Parameter.SPH")) c%pa_sph = toml_int("Parameter.SPH", 2)
        if (toml_has("Parameter.SLE")) c%sle = toml_int("Parameter.SLE", 0)
        if (toml_has("Parameter.deltaT")) c%dt = toml_real("Parameter.deltaT", c%dt)
        call flag("Parameter.summationDensity", c%summation_density)
        call flag("Parameter.normalizeDensity", c%normalize_density)
        call flag("Parameter.viscosity", c%viscosity)
        call flag("Parameter.gravity", c%gravity)
        call flag("Parameter.externalForce", c%external_force)
        call flag("Parameter.artificialViscosity", c%artificial_viscosity)
        call flag("Parameter.artificialStress", c%artificial_stress)
        call flag("Parameter.kernelCorrection", c%kernel_correction)
        !! getValueFromToml.f90:152-154: gravity (or dummy particles) turn on external forces
        if (c%gravity /= 0) c%external_force = 1
        if (toml_bool("Parameter.dummyParticle", .false.)) c%external_force = 1
    end subroutine apply_switches

    subroutine flag(key, v)
        character(len=*), intent(in) :: key
        integer(c_int32_t), intent(inout) :: v
        if (toml_has(key)) v = merge(1, 0, toml_bool(key, .false.))
    end subroutine flag

    subroutine make_dir(path)
        character(len=*), intent(in) :: path
        integer(c_int) :: rc
        logical :: there
        inquire(file=path // "/.", exist=there)
        if (there) return
        rc = c_mkdir(path // c_null_char, int(o'755', c_int))
        if (rc /= 0) then
            inquire(file=path // "/.", exist=there)
            if (.not. there) error stop "mysph_host: cannot create an output directory"
        end if
    end subroutine make_dir

    subroutine download(c, f)
        type(c_ptr), intent(in) :: c
        type(frame_t), intent(inout), target :: f
        call get(SPHGPU_F_X, f%x)
        call get(SPHGPU_F_V, f%v)
        call get(SPHGPU_F_DISP, f%disp)
        call get(SPHGPU_F_STRESS, f%stress)
        call get1(SPHGPU_F_MASS, f%mass)
        call get1(SPHGPU_F_RHO, f%rho)
        call get1(SPHGPU_F_P, f%p)
        call get1(SPHGPU_F_E, f%e)
        call get1(SPHGPU_F_C, f%c)
        call get1(SPHGPU_F_H, f%h)
        call get1(SPHGPU_F_ETA, f%eta)
        call get1(SPHGPU_F_DIVX, f%divx)
        f%type = ps%type
        f%state = ps%state
        f%boundary = ps%boundary
    contains
        subroutine get(field, a)
            integer(c_int32_t), intent(in) :: field
            real(8), intent(inout), target, contiguous :: a(:, :)
            call sphgpu_check(sphgpu_download(c, field, c_loc(a)), c, "sphgpu_download")
        end subroutine get
        subroutine get1(field, a)
            integer(c_int32_t), intent(in) :: field
            real(8), intent(inout), target, contiguous :: a(:)
            call sphgpu_check(sphgpu_download(c, field, c_loc(a)), c, "sphgpu_download")
        end subroutine get1
    end subroutine download

    subroutine frame_from_set(s, f)
        type(particle_set), intent(in) :: s
        type(frame_t), intent(inout) :: f
        integer :: m
        m = s%n_real + s%n_dummy
        call frame_allocate(f, s%dim, m)
        f%x = s%x
        f%v = s%v
        f%mass = s%mass
        f%rho = s%rho
        f%p = s%p
        f%e = s%e
        f%c = s%c
        f%h = s%h
        f%type = s%type
        f%state = s%state
        f%boundary = s%boundary
    end subroutine frame_from_set

end program mysph_host
