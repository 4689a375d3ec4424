!! Particle sets of the Fortran host: the dam-break family (2D and 3D, fixed
!! boundary lattice, seeded jitter) and particle files, in structure-of-arrays
!! form ready for sphgpu_upload.
!!
!! The dam break follows the same recipe and the same floating-point
!! operation order as mysph_amd/cases.py (dam_break), so both hosts build
!! identical inputs bit for bit:
!!   - fluid lattice (i + 1/2) dx per axis, last axis fastest;
!!   - jitter +-jitter*dx from SplitMix64 of key = seed XOR (4 id + axis),
!!     value 2 (top 53 bits / 2^53) - 1;
!!   - hydrostatic p = 1000 g max(H - z, 0), rho from the inverse Tait EOS
!!     (eos.f90:46-47 form) with libm pow;
!!   - `layers` of fixed boundary particles outside the open-top tank.
!! Physics constants: Tait c0 = 10 sqrt(2 g H), Monaghan AV, gravity,
!! continuity density, dt = cfl h / (c0 (1 + 1.2 alpha)) (courantNumber.f90:
!! 27-29 at rest).
module cases_m
    use, intrinsic :: iso_fortran_env, only: int32, int64
    use sphgpu_m
    implicit none
    private

    type, public :: particle_set
        integer :: dim = 2
        integer :: n_real = 0
        integer :: n_dummy = 0
        real(8), allocatable :: x(:, :), v(:, :)      !! (dim, n)
        real(8), allocatable :: mass(:), rho(:), p(:), e(:), c(:), h(:)
        integer(int32), allocatable :: type(:), state(:), boundary(:)
    end type particle_set

    integer(int64), parameter, public :: SEED_JITTER = 20261203_int64

    public :: case_dam_break, dam_break_config, allocate_set, splitmix64, jitter_unit, &
              write_dump, read_set, put_r8, put_i4, case_reference, is_reference_case

contains

    subroutine allocate_set(ps, dim, n)
        type(particle_set), intent(inout) :: ps
        integer, intent(in) :: dim, n
        ps%dim = dim
        allocate(ps%x(dim, n), ps%v(dim, n), source=0.0_8)
        allocate(ps%mass(n), ps%rho(n), ps%p(n), ps%e(n), ps%c(n), ps%h(n), source=0.0_8)
        allocate(ps%type(n), ps%state(n), ps%boundary(n), source=0_int32)
    end subroutine allocate_set

    !! a + b mod 2^64 on int64 bit patterns, without signed overflow.
    elemental integer(int64) function add64(a, b)
        integer(int64), intent(in) :: a, b
        integer(int64), parameter :: M32 = 4294967295_int64
        integer(int64) :: lo, hi
        lo = iand(a, M32) + iand(b, M32)
        hi = shiftr(a, 32) + shiftr(b, 32) + shiftr(lo, 32)
        add64 = ior(iand(lo, M32), shiftl(iand(hi, M32), 32))
    end function add64

    !! a * b mod 2^64 on int64 bit patterns, in 16-bit limbs (every partial
    !! sum stays below 2^35).
    elemental integer(int64) function mul64(a, b)
        integer(int64), intent(in) :: a, b
        integer(int64), parameter :: M16 = 65535_int64
        integer(int64) :: a0, a1, a2, a3, b0, b1, b2, b3, c0, c1, c2, c3
        a0 = iand(a, M16); a1 = iand(shiftr(a, 16), M16)
        a2 = iand(shiftr(a, 32), M16); a3 = shiftr(a, 48)
        b0 = iand(b, M16); b1 = iand(shiftr(b, 16), M16)
        b2 = iand(shiftr(b, 32), M16); b3 = shiftr(b, 48)
        c0 = a0 * b0
        c1 = a0 * b1 + a1 * b0 + shiftr(c0, 16)
        c2 = a0 * b2 + a1 * b1 + a2 * b0 + shiftr(c1, 16)
        c3 = a0 * b3 + a1 * b2 + a2 * b1 + a3 * b0 + shiftr(c2, 16)
        mul64 = ior(ior(iand(c0, M16), shiftl(iand(c1, M16), 16)), &
                    ior(shiftl(iand(c2, M16), 32), shiftl(iand(c3, M16), 48)))
    end function mul64

    !! SplitMix64 output function of key + golden gamma.
    elemental integer(int64) function splitmix64(key)
        integer(int64), intent(in) :: key
        integer(int64), parameter :: GAMMA = -7046029254386353131_int64   ! 0x9E3779B97F4A7C15
        integer(int64), parameter :: M1 = -4658895280553007687_int64      ! 0xBF58476D1CE4E5B9
        integer(int64), parameter :: M2 = -7723592293110705685_int64      ! 0x94D049BB133111EB
        integer(int64) :: z
        z = add64(key, GAMMA)
        z = mul64(ieor(z, shiftr(z, 30)), M1)
        z = mul64(ieor(z, shiftr(z, 27)), M2)
        splitmix64 = ieor(z, shiftr(z, 31))
    end function splitmix64

    !! Uniform value in [-1, 1) for (particle id, axis), axis = 0-based.
    elemental real(8) function jitter_unit(id, axis, seed)
        integer(int64), intent(in) :: id, seed
        integer, intent(in) :: axis
        integer(int64) :: z
        z = splitmix64(ieor(seed, ior(shiftl(id, 2), int(axis, int64))))
        jitter_unit = 2.0_8 * (real(shiftr(z, 11), 8) * (1.0_8 / 9007199254740992.0_8)) - 1.0_8
    end function jitter_unit

    !! Physics of the dam-break benchmark (alpha, beta = 0, psi = 0.1).
    function dam_break_config(dim, column_height, h, alpha, cfl, precision) result(c)
        integer, intent(in) :: dim, precision
        real(8), intent(in) :: column_height, h, alpha, cfl
        type(sphgpu_config) :: c
        real(8) :: c0
        c = sphgpu_base_config(dim)
        c%summation_density = 0
        c%viscosity = 0
        c%gravity = 1
        c%external_force = 1
        c%artificial_viscosity = 1
        c%av_alpha = alpha; c%av_beta = 0.0_8; c%av_psi = 0.1_8
        c%lj_r0 = 0.0_8     !! the fixed particles carry the wall
        c0 = 10.0_8 * sqrt(2.0_8 * 9.81_8 * column_height)
        c%dt = cfl * h / (c0 * (1.0_8 + 1.2_8 * alpha))
        c%precision = precision
        call sphgpu_add_material(c, 2, SPHGPU_EOS_TAIT, 1000.0_8, sphgpu_tait_b(c0, 1000.0_8, 7.0_8), &
                                 7.0_8, f32(1e-3_8))
    end function dam_break_config

    !! Dam break: `column` = (width, height) in 2D or (x, y, z) in 3D; `tank` =
    !! (length, wall height) or (length, width, wall height).
    subroutine case_dam_break(dim, dx, column, tank, layers, h_over_dx, jitter, alpha, seed, &
                              precision, ps, cfg)
        integer, intent(in) :: dim, layers, precision
        real(8), intent(in) :: dx, column(:), tank(:), h_over_dx, jitter, alpha
        integer(int64), intent(in) :: seed
        type(particle_set), intent(out) :: ps
        type(sphgpu_config), intent(out) :: cfg
        integer :: ncol(3), ntank(3), lo(3), hi(3), i0, i1, i2, d, nf, nb, k, idx(3)
        real(8) :: top, h, B, c0, g, m, jd

        top = column(dim)
        h = h_over_dx * dx
        cfg = dam_break_config(dim, top, h, alpha, 0.2_8, precision)
        B = cfg%mat(1)%K
        c0 = sphgpu_sound_speed(cfg, 2)     !! sqrt(B gamma / rho0)
        g = 9.81_8
        m = 1000.0_8 * dx ** real(dim, 8)
        jd = jitter * dx
        ncol = 1
        ntank = 1
        do d = 1, dim
            ncol(d) = nint(column(d) / dx)
            ntank(d) = nint(tank(d) / dx)
        end do
        nf = product(ncol(1:dim))
        !! boundary cells: [-layers, ntank + layers) on the horizontal axes,
        !! [-layers, ntank) vertically, minus the open-top interior
        lo = 0
        hi = 1
        do d = 1, dim - 1
            lo(d) = -layers
            hi(d) = ntank(d) + layers
        end do
        lo(dim) = -layers
        hi(dim) = ntank(dim)
        nb = 0
        do i0 = lo(1), hi(1) - 1
            do i1 = lo(2), hi(2) - 1
                do i2 = lo(3), hi(3) - 1
                    idx = [i0, i1, i2]
                    if (.not. interior(idx)) nb = nb + 1
                end do
            end do
        end do
        call allocate_set(ps, dim, nf + nb)
        ps%n_real = nf
        ps%n_dummy = nb
        !! fluid, last axis fastest
        k = 0
        do i0 = 0, ncol(1) - 1
            do i1 = 0, ncol(2) - 1
                do i2 = 0, ncol(3) - 1
                    idx = [i0, i1, i2]
                    k = k + 1
                    do d = 1, dim
                        ps%x(d, k) = (real(idx(d), 8) + 0.5_8) * dx
                        if (jitter /= 0.0_8) ps%x(d, k) = ps%x(d, k) + &
                            jd * jitter_unit(int(k - 1, int64), d - 1, seed)
                    end do
                    call fill(k, 2)
                end do
            end do
        end do
        do i0 = lo(1), hi(1) - 1
            do i1 = lo(2), hi(2) - 1
                do i2 = lo(3), hi(3) - 1
                    idx = [i0, i1, i2]
                    if (interior(idx)) cycle
                    k = k + 1
                    do d = 1, dim
                        ps%x(d, k) = (real(idx(d), 8) + 0.5_8) * dx
                    end do
                    call fill(k, -2)
                end do
            end do
        end do

    contains

        logical function interior(ix)
            integer, intent(in) :: ix(3)
            integer :: dd
            interior = ix(dim) >= 0
            do dd = 1, dim - 1
                interior = interior .and. ix(dd) >= 0 .and. ix(dd) < ntank(dd)
            end do
        end function interior

        subroutine fill(kk, t)
            integer, intent(in) :: kk, t
            ps%mass(kk) = m
            ps%p(kk) = 1000.0_8 * g * max(top - ps%x(dim, kk), 0.0_8)
            ps%rho(kk) = (ps%p(kk) / B + 1.0_8) ** (1.0_8 / 7.0_8) * 1000.0_8
            ps%c(kk) = c0
            ps%h(kk) = h
            ps%type(kk) = t
        end subroutine fill

    end subroutine case_dam_break

    !! The cases the reference builds in code for these nicks
    !! (input.f90:1031-1088 dam break with dummyParticles.f90:544-659;
    !! input.f90:1336-1405 cantilever beams), in the same floating-point
    !! operation order as mysph_amd/cases.py (reference_dam_break,
    !! reference_can_beam, reference_can_beam_3d), which equals the
    !! reference's bit for bit. The reference regenerates the dam break's wall
    !! particles every step; they never move, so they are built once.
    logical function is_reference_case(nick)
        character(len=*), intent(in) :: nick
        select case (nick)
        case ("dam_break", "can_beam", "can_beam_3d")
            is_reference_case = .true.
        case default
            is_reference_case = .false.
        end select
    end function is_reference_case

    subroutine case_reference(nick, ps, cfg)
        character(len=*), intent(in) :: nick
        type(particle_set), intent(out) :: ps
        type(sphgpu_config), intent(out) :: cfg
        select case (nick)
        case ("dam_break")
            call reference_dam_break(ps, cfg)
        case ("can_beam")
            call reference_beam(2, ps, cfg)
        case ("can_beam_3d")
            call reference_beam(3, ps, cfg)
        case default
            error stop "case_reference: not a reference case"
        end select
    end subroutine case_reference

    subroutine reference_dam_break(ps, cfg)
        type(particle_set), intent(out) :: ps
        type(sphgpu_config), intent(out) :: cfg
        real(8) :: dx, g981, B, ddx, lf, xx, yy, pk, wall(4)
        integer :: nx, ny, nf, nd, k, i, j, l, nyw, nxw, nbaf
        real(8), allocatable :: wx(:, :), wm(:)
        integer, allocatable :: hyd(:)     !! 1 hydrostatic, 0 lid (rho0, p 0)
        dx = f32(0.2_8)
        g981 = f32(9.81_8)
        B = sphgpu_tait_b(50.0_8, 1000.0_8, 7.0_8)
        nx = int((25.0_8 - 0.0_8) / dx)
        ny = int((25.0_8 - 0.0_8) / dx)
        nf = nx * ny
        !! walls, dx re-derived from particle 2 (dummyParticles.f90:561-564):
        !! particle 2 is (i, j) = (1, 2)
        ddx = abs(real(real(1, 4) - 0.5_4, 8) * dx - real(real(2, 4) - 0.5_4, 8) * dx)
        wall = [0.0_8, 100.0_8, 0.0_8, 40.0_8]
        nyw = floor((wall(4) - wall(3)) / ddx) + 4
        nxw = floor((wall(2) - wall(1)) / ddx) + 1
        nbaf = floor(3 / (ddx / 2))
        nd = 4 * (2 * nyw + nxw)
        do l = 1, 4
            nd = nd + max(nbaf - l + 1, 0)
        end do
        nd = nd + 4 * nxw
        allocate(wx(2, nd), wm(nd), hyd(nd))
        k = 0
        do l = 1, 4
            lf = real(real(l, 4) - 0.5_4, 8)
            do i = 1, nyw
                yy = wall(3) + real(real(i - 4, 4) - 0.5_4, 8) * ddx
                call add(wall(1) - lf * ddx, yy, 1000.0_8 * ddx * ddx, 1)
            end do
            do i = 1, nyw
                yy = wall(3) + real(real(i - 4, 4) - 0.5_4, 8) * ddx
                call add(wall(2) + lf * ddx, yy, 1000.0_8 * ddx * ddx, 1)
            end do
            do i = 1, nxw
                xx = wall(1) + real(real(i, 4) - 0.5_4, 8) * ddx
                call add(xx, wall(3) - lf * ddx, 1000.0_8 * ddx * ddx, -1)
            end do
        end do
        do l = 1, 4
            do i = 1, nbaf - l + 1
                xx = (wall(1) + wall(2)) / 2 + real(real(i + l, 4) - 1.5_4, 8) * ddx / 2
                yy = wall(3) + real(real(i, 4) - 0.5_4, 8) * ddx / 2
                call add(xx, yy, 1000.0_8 * ddx * ddx / 4, 1)
            end do
        end do
        do l = 1, 4
            lf = real(real(l, 4) + 0.5_4, 8)
            do i = 1, nxw
                xx = wall(1) + real(real(i, 4) - 0.5_4, 8) * ddx
                call add(xx, wall(4) - lf * ddx, 1000.0_8 * ddx * ddx, 0)
            end do
        end do
        if (k /= nd) error stop "reference_dam_break: wall count"
        call allocate_set(ps, 2, nf + nd)
        ps%n_real = nf
        ps%n_dummy = nd
        k = 0
        do i = 1, nx
            do j = 1, ny
                k = k + 1
                ps%x(1, k) = 0.0_8 + real(real(i, 4) - 0.5_4, 8) * dx
                ps%x(2, k) = 0.0_8 + real(real(j, 4) - 0.5_4, 8) * dx
                ps%mass(k) = 1000.0_8 * dx * dx
                ps%p(k) = 1000.0_8 * g981 * (25.0_8 - ps%x(2, k))
                ps%rho(k) = (ps%p(k) / B + 1.0_8) ** (1.0_8 / 7.0_8) * 1000.0_8
                ps%c(k) = 50.0_8
                ps%type(k) = 2
                ps%h(k) = dx
            end do
        end do
        do i = 1, nd
            k = nf + i
            ps%x(:, k) = wx(:, i)
            ps%mass(k) = wm(i)
            ps%rho(k) = 1000.0_8
            ps%type(k) = -2
            ps%h(k) = ddx
            if (hyd(i) == 1) then
                pk = 1000.0_8 * g981 * max(25.0_8 - wx(2, i), 0.0_8)
                ps%p(k) = pk
                ps%rho(k) = (pk / B + 1.0_8) ** (1.0_8 / 7.0_8) * 1000.0_8
            end if
        end do
        cfg = sphgpu_base_config(2)
        if (.not. sphgpu_preset("dam_break", 2, ddx, cfg)) error stop "reference_dam_break: preset"
        cfg%summation_density = 0
        cfg%viscosity = 1
        cfg%gravity = 1
        cfg%external_force = 1
        cfg%artificial_viscosity = 1
        cfg%dt = 1e-3_8
    contains
        subroutine add(ax, ay, am, h)
            real(8), intent(in) :: ax, ay, am
            integer, intent(in) :: h
            k = k + 1
            wx(1, k) = ax
            wx(2, k) = ay
            wm(k) = am
            hyd(k) = h
        end subroutine add
    end subroutine reference_dam_break

    !! data/cantileverBeam (101 x 5, dx = 1e-3) and data/cantileverBeam3D
    !! (201 x 5 x 9, dx = 5e-4): type 102, h = 2 dx, first column clamped.
    subroutine reference_beam(dim, ps, cfg)
        integer, intent(in) :: dim
        type(particle_set), intent(out) :: ps
        type(sphgpu_config), intent(out) :: cfg
        real(8) :: dx, bulk, rho0, c
        integer :: nx, ny, nz, i, j, kk, k
        if (dim == 2) then
            dx = f32(1e-3_8)
            nx = 101; ny = 5; nz = 1
        else
            dx = f32(5e-4_8)
            nx = 201; ny = 5; nz = 9
        end if
        bulk = f32(2e7_8)
        rho0 = 1100.0_8
        c = sqrt(bulk / rho0)
        call allocate_set(ps, dim, nx * ny * nz)
        ps%n_real = nx * ny * nz
        ps%n_dummy = 0
        k = 0
        do i = 1, nx
            do j = 1, ny
                do kk = 1, nz
                    k = k + 1
                    ps%x(1, k) = (i - 1) * dx
                    ps%x(2, k) = (j - 1 - ny / 2) * dx
                    if (dim == 3) ps%x(3, k) = (kk - 1 - nz / 2) * dx
                    ps%rho(k) = rho0
                    if (dim == 2) then
                        ps%mass(k) = rho0 * (dx * dx)
                    else
                        ps%mass(k) = rho0 * (dx * dx * dx)
                    end if
                    ps%c(k) = c
                    ps%type(k) = 102
                    ps%h(k) = dx * 2
                    if (i == 1) ps%boundary(k) = 1
                end do
            end do
        end do
        cfg = sphgpu_base_config(dim)
        if (dim == 2) then
            if (.not. sphgpu_preset("can_beam", 2, 0.0_8, cfg)) error stop "reference_beam: preset"
        else
            if (.not. sphgpu_preset("can_beam_3d", 3, 0.0_8, cfg)) error stop "reference_beam: preset"
        end if
        cfg%summation_density = 0
        cfg%viscosity = 1
        cfg%gravity = 1
        cfg%external_force = 1
        cfg%artificial_viscosity = 1
        cfg%kernel_correction = 1
        cfg%dt = 2e-6_8
    end subroutine reference_beam

    !! Records of the particle file format: name (8 chars), kind (int32: 8 =
    !! real64, 4 = int32), count (int64), data.
    subroutine put_r8(u, name, a)
        integer, intent(in) :: u
        character(len=*), intent(in) :: name
        real(8), intent(in) :: a(:)
        character(len=8) :: nm
        nm = name
        write(u) nm, 8_int32, int(size(a), int64), a
    end subroutine put_r8

    subroutine put_i4(u, name, a)
        integer, intent(in) :: u
        character(len=*), intent(in) :: name
        integer(int32), intent(in) :: a(:)
        character(len=8) :: nm
        nm = name
        write(u) nm, 4_int32, int(size(a), int64), a
    end subroutine put_i4

    subroutine write_dump(path, ps)
        character(len=*), intent(in) :: path
        type(particle_set), intent(in) :: ps
        integer :: u, n
        n = ps%n_real + ps%n_dummy
        open(newunit=u, file=path, access="stream", form="unformatted", status="replace")
        call put_i4(u, "ntotal", [int(ps%n_real, int32)])
        call put_i4(u, "ndummy", [int(ps%n_dummy, int32)])
        call put_i4(u, "dim", [int(ps%dim, int32)])
        call put_r8(u, "x", reshape(ps%x(:, 1:n), [ps%dim * n]))
        call put_r8(u, "v", reshape(ps%v(:, 1:n), [ps%dim * n]))
        call put_r8(u, "mass", ps%mass(1:n))
        call put_r8(u, "rho", ps%rho(1:n))
        call put_r8(u, "p", ps%p(1:n))
        call put_r8(u, "e", ps%e(1:n))
        call put_r8(u, "c", ps%c(1:n))
        call put_r8(u, "h", ps%h(1:n))
        call put_i4(u, "type", ps%type(1:n))
        call put_i4(u, "state", ps%state(1:n))
        call put_i4(u, "boundary", ps%boundary(1:n))
        close(u)
    end subroutine write_dump

    !! Read a particle file (the format write_dump writes). Required records:
    !! ntotal, ndummy, dim, x, mass, rho, h, type; others default to zero.
    subroutine read_set(path, ps, ok, msg)
        character(len=*), intent(in) :: path
        type(particle_set), intent(out) :: ps
        logical, intent(out) :: ok
        character(len=:), allocatable, intent(out) :: msg
        integer :: u, ios, n, dim, nreal, ndum
        character(len=8) :: nm
        integer(int32) :: kind
        integer(int64) :: cnt
        real(8), allocatable :: r(:)
        integer(int32), allocatable :: iv(:)
        logical :: have(5)

        ok = .false.
        msg = ""
        nreal = -1; ndum = -1; dim = -1
        have = .false.
        open(newunit=u, file=path, access="stream", form="unformatted", status="old", iostat=ios)
        if (ios /= 0) then
            msg = "cannot open " // path
            return
        end if
        !! first pass: sizes
        do
            read(u, iostat=ios) nm, kind, cnt
            if (ios /= 0) exit
            if (kind == 4) then
                allocate(iv(cnt))
                read(u) iv
                select case (trim(nm))
                case ("ntotal"); nreal = iv(1)
                case ("ndummy"); ndum = iv(1)
                case ("dim"); dim = iv(1)
                end select
                deallocate(iv)
            else
                allocate(r(cnt))
                read(u) r
                deallocate(r)
            end if
        end do
        if (nreal < 0 .or. ndum < 0 .or. (dim /= 2 .and. dim /= 3)) then
            msg = "particle file lacks ntotal/ndummy/dim"
            close(u)
            return
        end if
        n = nreal + ndum
        call allocate_set(ps, dim, n)
        ps%n_real = nreal
        ps%n_dummy = ndum
        rewind(u)
        do
            read(u, iostat=ios) nm, kind, cnt
            if (ios /= 0) exit
            if (kind == 4) then
                allocate(iv(cnt))
                read(u) iv
                if (any(trim(nm) == [character(len=8) :: "type", "state", "boundary"]) .and. cnt /= n) then
                    msg = "record " // trim(nm) // " has the wrong length"
                    close(u)
                    return
                end if
                select case (trim(nm))
                case ("type"); ps%type = iv; have(5) = .true.
                case ("state"); ps%state = iv
                case ("boundary"); ps%boundary = iv
                end select
                deallocate(iv)
            else
                allocate(r(cnt))
                read(u) r
                select case (trim(nm))
                case ("x", "v")
                    if (cnt /= int(dim, int64) * n) then
                        msg = "record " // trim(nm) // " has the wrong length"
                        close(u)
                        return
                    end if
                    if (trim(nm) == "x") then
                        ps%x = reshape(r, [dim, n]); have(1) = .true.
                    else
                        ps%v = reshape(r, [dim, n])
                    end if
                case ("mass", "rho", "p", "e", "c", "h")
                    if (cnt /= n) then
                        msg = "record " // trim(nm) // " has the wrong length"
                        close(u)
                        return
                    end if
                    select case (trim(nm))
                    case ("mass"); ps%mass = r; have(2) = .true.
                    case ("rho"); ps%rho = r; have(3) = .true.
                    case ("p"); ps%p = r
                    case ("e"); ps%e = r
                    case ("c"); ps%c = r
                    case ("h"); ps%h = r; have(4) = .true.
                    end select
                end select
                deallocate(r)
            end if
        end do
        close(u)
        if (.not. all(have)) then
            msg = "particle file lacks one of x, mass, rho, h, type"
            return
        end if
        ok = .true.
    end subroutine read_set

end module cases_m
