!! Output of the Fortran host in the reference's file layouts, so its results
!! diff against reference runs and feed the same plotting scripts:
!!   - one text file per particle type, Type_<type>_<index>.dat, with the
!!     header and the per-particle record of output.f90:87-244 and
!!     SPH.f90:56-93 (index, type, state, then x, v, mass, density,
!!     pressure, internal energy, sound speed, smoothing length, viscosity,
!!     position divergence, displacement, stress row by row, ES15.7; solids
!!     add the von Mises stress, ES24.16);
!!   - one legacy ASCII VTK file per type, Type_<type>_<index>.vtk, with the
!!     fields and order of output.f90:246-352.
!! Type 0 particles are filed with type 4 if present, else type 5, as the
!! reference does (output.f90:36-48).
module output_m
    use, intrinsic :: iso_fortran_env, only: int32
    implicit none
    private

    !! One snapshot of the particle state on the host (downloaded fields).
    type, public :: frame_t
        integer :: dim = 2
        integer :: n = 0
        real(8), allocatable :: x(:, :), v(:, :), disp(:, :)   !! (dim, n)
        real(8), allocatable :: stress(:, :)                    !! (dim*dim, n): Stress(a,b) at (a-1)*dim+b
        real(8), allocatable :: mass(:), rho(:), p(:), e(:), c(:), h(:), eta(:), divx(:)
        integer(int32), allocatable :: type(:), state(:), boundary(:)
    end type frame_t

    public :: frame_allocate, write_output, von_mises, timestamp

contains

    subroutine frame_allocate(f, dim, n)
        type(frame_t), intent(inout) :: f
        integer, intent(in) :: dim, n
        f%dim = dim
        f%n = n
        allocate(f%x(dim, n), f%v(dim, n), f%disp(dim, n), f%stress(dim * dim, n), source=0.0_8)
        allocate(f%mass(n), f%rho(n), f%p(n), f%e(n), f%c(n), f%h(n), f%eta(n), f%divx(n), source=0.0_8)
        allocate(f%type(n), f%state(n), f%boundary(n), source=0_int32)
    end subroutine frame_allocate

    !! von Mises stress of output.f90:354-377 from Stress(a,b) = s((a-1)*dim+b).
    pure real(8) function von_mises(s, dim) result(m)
        real(8), intent(in) :: s(:)
        integer, intent(in) :: dim
        select case (dim)
        case (1)
            m = abs(s(1))
        case (2)
            m = sqrt((s(1) + s(4))**2 - 3 * (s(1) * s(4)) + 3 * (s(2)**2))
        case default
            m = sqrt((s(1) + s(5) + s(9))**2 &
                     - 3 * (s(5) * s(9) + s(9) * s(1) + s(1) * s(5)) &
                     + 3 * (s(2)**2 + s(3)**2 + s(6)**2))
        end select
    end function von_mises

    function timestamp() result(t)
        character(len=19) :: t
        character(len=8) :: date
        character(len=10) :: time
        call date_and_time(date, time)
        t = date(1:4) // "-" // date(5:6) // "-" // date(7:8) // " " // &
            time(1:2) // ":" // time(3:4) // ":" // time(5:6)
    end function timestamp

    !! Write the .dat (to out_dir) and .vtk (to vtk_dir, if vtk) files of
    !! every particle type present in f; fixed boundary particles (negative
    !! type) only if with_dummy.
    subroutine write_output(out_dir, vtk_dir, index, f, vtk, with_dummy)
        character(len=*), intent(in) :: out_dir, vtk_dir
        integer, intent(in) :: index
        type(frame_t), intent(in) :: f
        logical, intent(in) :: vtk, with_dummy
        integer :: num(-200:200), t, i, k, tt
        integer, allocatable :: ids(:)
        character(len=:), allocatable :: base

        num = 0
        do i = 1, f%n
            t = max(-200, min(200, int(f%type(i))))
            num(t) = num(t) + 1
        end do
        allocate(ids(f%n))
        do t = -200, 200
            if (t == 0) cycle
            tt = t
            k = 0
            do i = 1, f%n
                if (filed_type(int(f%type(i))) == tt) then
                    k = k + 1
                    ids(k) = i
                end if
            end do
            if (k == 0) cycle
            if (tt < 0 .and. .not. with_dummy) cycle
            base = "Type_" // itoa(tt) // "_" // itoa(index)
            call write_dat(out_dir // "/" // base // ".dat", f, ids(1:k))
            if (vtk) call write_vtk(vtk_dir // "/" // base // ".vtk", f, ids(1:k))
        end do

    contains

        integer function filed_type(ty)
            integer, intent(in) :: ty
            filed_type = ty
            if (ty == 0) then
                if (num(4) /= 0) then
                    filed_type = 4
                else
                    filed_type = 5
                end if
            end if
        end function filed_type

    end subroutine write_output

    subroutine write_dat(path, f, ids)
        character(len=*), intent(in) :: path
        type(frame_t), intent(in) :: f
        integer, intent(in) :: ids(:)
        integer :: u, j, i, dim, nval
        logical :: solid
        character(len=128) :: fmt

        dim = f%dim
        solid = f%type(ids(1)) > 100
        nval = 8 + 3 * dim + dim * dim
        if (solid) then
            write(fmt, "(A, I0, A)") "(2(I8), 2X, ", nval, "(ES15.7e0, 2X), ES24.16e3)"
        else
            write(fmt, "(A, I0, A)") "(2(I8), 2X, ", nval, "(ES15.7e0, 2X))"
        end if
        open(newunit=u, file=path, status="replace", action="write")
        select case (dim)
        case (2)
            if (solid) then
                write(u, 1001) "Index", "Type", "State", "X", "Y", "U", "V", "Mass", "Density", &
                    "Pressure", "InternalEnergy", "SoundSpeed", "SmoothingLength", "Viscosity", &
                    "DivDistance", "Displacement_X", "Displacement_Y", "StressXX", "StressXY", &
                    "StressYX", "StressYY", "vonMises"
            else
                write(u, 1001) "Index", "Type", "State", "X", "Y", "U", "V", "Mass", "Density", &
                    "Pressure", "InternalEnergy", "SoundSpeed", "SmoothingLength", "Viscosity", &
                    "DivDistance", "Displacement_X", "Displacement_Y", "StressXX", "StressXY", &
                    "StressYX", "StressYY"
            end if
        case default
            if (solid) then
                write(u, 1001) "Index", "Type", "State", "X", "Y", "Z", "U", "V", "W", "Mass", &
                    "Density", "Pressure", "InternalEnergy", "SoundSpeed", "SmoothingLength", &
                    "Viscosity", "DivDistance", "Displacement_X", "Displacement_Y", "Displacement_Z", &
                    "StressXX", "StressXY", "StressXZ", "StressYX", "StressYY", "StressYZ", &
                    "StressZX", "StressZY", "StressZZ", "vonMises"
            else
                write(u, 1001) "Index", "Type", "State", "X", "Y", "Z", "U", "V", "W", "Mass", &
                    "Density", "Pressure", "InternalEnergy", "SoundSpeed", "SmoothingLength", &
                    "Viscosity", "DivDistance", "Displacement_X", "Displacement_Y", "Displacement_Z", &
                    "StressXX", "StressXY", "StressXZ", "StressYX", "StressYY", "StressYZ", &
                    "StressZX", "StressZY", "StressZZ"
            end if
        end select
        do j = 1, size(ids)
            i = ids(j)
            write(u, "(I8)", advance="no") i
            if (solid) then
                write(u, fmt) f%type(i), f%state(i), f%x(:, i), f%v(:, i), f%mass(i), f%rho(i), &
                    f%p(i), f%e(i), f%c(i), f%h(i), f%eta(i), f%divx(i), f%disp(:, i), &
                    f%stress(:, i), von_mises(f%stress(:, i), dim)
            else
                write(u, fmt) f%type(i), f%state(i), f%x(:, i), f%v(:, i), f%mass(i), f%rho(i), &
                    f%p(i), f%e(i), f%c(i), f%h(i), f%eta(i), f%divx(i), f%disp(:, i), &
                    f%stress(:, i)
            end if
        end do
        close(u)
1001    format(3(A8), *(A17))
    end subroutine write_dat

    subroutine write_vtk(path, f, ids)
        character(len=*), intent(in) :: path
        type(frame_t), intent(in) :: f
        integer, intent(in) :: ids(:)
        integer :: u, j, i, d, dd, dim, n

        dim = f%dim
        n = size(ids)
        open(newunit=u, file=path, status="replace", action="write")
        write(u, "(A)") "# vtk DataFile Version 3.0, " // timestamp()
        write(u, "(A)") "paraview_vtk_output"
        write(u, "(A)") "ASCII"
        write(u, "(A)") "DATASET UNSTRUCTURED_GRID"
        write(u, "(A, I0, A)") "POINTS ", n, " float"
        do j = 1, n
            write(u, 1001) real(f%x(:, ids(j))), (0.0, d = 1, 3 - dim)
        end do
        write(u, "(A, I0)") "POINT_DATA ", n
        call scalar("Mass", f%mass)
        call scalar("Density", f%rho)
        call scalar("Pressure", f%p)
        call scalar("InternalEnergy", f%e)
        call iscalar("Type", f%type)
        call iscalar("State", f%state)
        call iscalar("Boundary", f%boundary)
        call scalar("SmoothingLength", f%h)
        call scalar("DivDistance", f%divx)
        write(u, "(A)") "VECTORS U float"
        do j = 1, n
            write(u, 1001) real(f%v(:, ids(j))), (0.0, d = 1, 3 - dim)
        end do
        write(u, "(A)") "VECTORS Displacement float"
        do j = 1, n
            write(u, 1001) real(f%disp(:, ids(j))), (0.0, d = 1, 3 - dim)
        end do
        !! column-major Stress(d, dd), then zeros up to nine values
        write(u, "(A)") "TENSORS Stress float"
        do j = 1, n
            i = ids(j)
            write(u, 1001) ((real(f%stress((d - 1) * dim + dd, i)), d = 1, dim), dd = 1, dim), &
                           (0.0, d = 1, 9 - dim**2)
        end do
        if (f%type(ids(1)) > 100) then
            write(u, "(A)") "SCALARS vonMises float 1"
            write(u, "(A)") "LOOKUP_TABLE DEFAULT"
            do j = 1, n
                write(u, 1001) real(von_mises(f%stress(:, ids(j)), dim))
            end do
        end if
        close(u)
1001    format(*(ES24.16e3, 3X))

    contains

        subroutine scalar(name, a)
            character(len=*), intent(in) :: name
            real(8), intent(in) :: a(:)
            integer :: jj
            write(u, "(A)") "SCALARS " // name // " float 1"
            write(u, "(A)") "LOOKUP_TABLE DEFAULT"
            do jj = 1, n
                write(u, 1001) real(a(ids(jj)))
            end do
1001        format(*(ES24.16e3, 3X))
        end subroutine scalar

        subroutine iscalar(name, a)
            character(len=*), intent(in) :: name
            integer(int32), intent(in) :: a(:)
            integer :: jj
            write(u, "(A)") "SCALARS " // name // " int 1"
            write(u, "(A)") "LOOKUP_TABLE DEFAULT"
            do jj = 1, n
                write(u, "(I0)") a(ids(jj))
            end do
        end subroutine iscalar

    end subroutine write_vtk

    function itoa(i) result(s)
        integer, intent(in) :: i
        character(len=:), allocatable :: s
        character(len=16) :: b
        write(b, "(I0)") i
        s = trim(b)
    end function itoa

end module output_m
