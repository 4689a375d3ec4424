!! Reader for the subset of TOML that case files use: [section] headers,
!! `key = value` lines with strings, integers, reals, booleans and flat
!! arrays of numbers, and `#` comments. Keys are looked up as
!! "section.key" ("key" before the first header), case-sensitively, with a
!! default for a missing key, as the reference's reader does
!! (src/Initialize/getValueFromToml.f90:95-185, which uses toml-f).
module toml_lite_m
    implicit none
    private

    integer, parameter :: KEYLEN = 96, VALLEN = 512
    type :: entry_t
        character(len=KEYLEN) :: key = ""
        character(len=VALLEN) :: val = ""
    end type entry_t

    type(entry_t), allocatable, save :: table(:)
    integer, save :: nent = 0

    public :: toml_load, toml_has, toml_int, toml_real, toml_bool, toml_string, toml_reals, &
              toml_count

contains

    !! Parse `path`; ok = .false. with msg set on a syntax error or a
    !! missing file.
    subroutine toml_load(path, ok, msg)
        character(len=*), intent(in) :: path
        logical, intent(out) :: ok
        character(len=:), allocatable, intent(out) :: msg
        character(len=1024) :: line
        character(len=KEYLEN) :: section
        integer :: u, ios, lineno, eq, k
        character(len=:), allocatable :: s, key, val

        ok = .false.
        msg = ""
        nent = 0
        if (allocated(table)) deallocate(table)
        allocate(table(64))
        section = ""
        open(newunit=u, file=path, status="old", action="read", iostat=ios)
        if (ios /= 0) then
            msg = "cannot open " // trim(path)
            return
        end if
        lineno = 0
        do
            read(u, "(A)", iostat=ios) line
            if (ios /= 0) exit
            lineno = lineno + 1
            s = trim(adjustl(strip_comment(line)))
            if (len(s) == 0) cycle
            if (s(1:1) == "[") then
                k = index(s, "]")
                if (k < 3) then
                    msg = "bad section header at line " // itoa(lineno)
                    close(u)
                    return
                end if
                section = trim(adjustl(s(2:k-1)))
                cycle
            end if
            eq = index(s, "=")
            if (eq < 2) then
                msg = "expected key = value at line " // itoa(lineno)
                close(u)
                return
            end if
            key = trim(adjustl(s(1:eq-1)))
            val = trim(adjustl(s(eq+1:)))
            if (len(val) == 0) then
                msg = "missing value at line " // itoa(lineno)
                close(u)
                return
            end if
            if (len_trim(section) > 0) key = trim(section) // "." // key
            call put(key, val)
        end do
        close(u)
        ok = .true.
    end subroutine toml_load

    !! Text before a '#' that is not inside a quoted string.
    function strip_comment(line) result(s)
        character(len=*), intent(in) :: line
        character(len=:), allocatable :: s
        integer :: i
        character(len=1) :: q
        q = " "
        do i = 1, len(line)
            if (q == " ") then
                if (line(i:i) == '"' .or. line(i:i) == "'") then
                    q = line(i:i)
                else if (line(i:i) == "#") then
                    s = line(1:i-1)
                    return
                end if
            else if (line(i:i) == q) then
                q = " "
            end if
        end do
        s = line
    end function strip_comment

    subroutine put(key, val)
        character(len=*), intent(in) :: key, val
        type(entry_t), allocatable :: tmp(:)
        integer :: i
        do i = 1, nent
            if (table(i)%key == key) then
                table(i)%val = val
                return
            end if
        end do
        if (nent == size(table)) then
            allocate(tmp(2 * size(table)))
            tmp(1:nent) = table(1:nent)
            call move_alloc(tmp, table)
        end if
        nent = nent + 1
        table(nent)%key = key
        table(nent)%val = val
    end subroutine put

    integer function find(key)
        character(len=*), intent(in) :: key
        integer :: i
        find = 0
        do i = 1, nent
            if (table(i)%key == key) then
                find = i
                return
            end if
        end do
    end function find

    integer function toml_count()
        toml_count = nent
    end function toml_count

    logical function toml_has(key)
        character(len=*), intent(in) :: key
        toml_has = find(key) > 0
    end function toml_has

    integer function toml_int(key, default)
        character(len=*), intent(in) :: key
        integer, intent(in) :: default
        integer :: k, ios
        toml_int = default
        k = find(key)
        if (k == 0) return
        read(table(k)%val, *, iostat=ios) toml_int
        if (ios /= 0) call bad(key, "an integer")
    end function toml_int

    real(8) function toml_real(key, default)
        character(len=*), intent(in) :: key
        real(8), intent(in) :: default
        integer :: k, ios
        toml_real = default
        k = find(key)
        if (k == 0) return
        read(table(k)%val, *, iostat=ios) toml_real
        if (ios /= 0) call bad(key, "a number")
    end function toml_real

    logical function toml_bool(key, default)
        character(len=*), intent(in) :: key
        logical, intent(in) :: default
        integer :: k
        toml_bool = default
        k = find(key)
        if (k == 0) return
        select case (trim(table(k)%val))
        case ("true")
            toml_bool = .true.
        case ("false")
            toml_bool = .false.
        case default
            call bad(key, "true or false")
        end select
    end function toml_bool

    function toml_string(key, default) result(s)
        character(len=*), intent(in) :: key, default
        character(len=:), allocatable :: s
        integer :: k, n
        k = find(key)
        if (k == 0) then
            s = default
            return
        end if
        s = trim(table(k)%val)
        n = len(s)
        if (n >= 2 .and. (s(1:1) == '"' .or. s(1:1) == "'") .and. s(n:n) == s(1:1)) then
            s = s(2:n-1)
        else
            call bad(key, "a quoted string")
        end if
    end function toml_string

    !! Flat array of numbers `[a, b, ...]`; n = number read (0 if missing).
    subroutine toml_reals(key, a, n)
        character(len=*), intent(in) :: key
        real(8), intent(inout) :: a(:)
        integer, intent(out) :: n
        integer :: k, i, j, ios
        character(len=VALLEN) :: v
        n = 0
        k = find(key)
        if (k == 0) return
        v = table(k)%val
        i = index(v, "[")
        j = index(v, "]", back=.true.)
        if (i == 0 .or. j <= i) call bad(key, "an array of numbers")
        v = v(i+1:j-1)
        do i = 1, len_trim(v)
            if (v(i:i) == ",") n = n + 1
        end do
        n = n + 1
        if (len_trim(v) == 0) n = 0
        if (n > size(a)) call bad(key, "an array of at most " // itoa(size(a)) // " numbers")
        if (n > 0) then
            read(v, *, iostat=ios) a(1:n)
            if (ios /= 0) call bad(key, "an array of numbers")
        end if
    end subroutine toml_reals

    subroutine bad(key, what)
        character(len=*), intent(in) :: key, what
        write(*, "(A)") " toml: value of " // trim(key) // " is not " // what
        error stop "invalid case file"
    end subroutine bad

    function itoa(i) result(s)
        integer, intent(in) :: i
        character(len=:), allocatable :: s
        character(len=16) :: b
        write(b, "(I0)") i
        s = trim(b)
    end function itoa

end module toml_lite_m
