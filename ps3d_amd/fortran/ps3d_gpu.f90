! ps3d_gpu.f90 -- iso_c_binding interface of the MI355X step library
! (include/ps3d_gpu.h, include/ps3d_nc.h): what a ps3d host program uses in place
! of its inversion, stepper and NetCDF modules.
module ps3d_gpu
    use iso_c_binding
    implicit none

    ! enum ps3d_field
    integer(c_int), parameter :: PS3D_SVOR = 0, PS3D_VOR = 1, PS3D_VEL = 2, PS3D_SVEL = 3, &
                                 PS3D_SVORTS = 4, PS3D_PRES = 5, PS3D_DELTA = 6
    ! enum ps3d_filtering, ps3d_stepper, ps3d_pretype, ps3d_length_scale
    integer(c_int), parameter :: PS3D_FILTER_HOU_LI = 0, PS3D_FILTER_23RULE = 1
    integer(c_int), parameter :: PS3D_STEPPER_IMPL_RK4 = 0, PS3D_STEPPER_CN2 = 1
    integer(c_int), parameter :: PS3D_PRE_CONSTANT = 0, PS3D_PRE_VORCH = 1, PS3D_PRE_BFMAX = 2, &
                                 PS3D_PRE_ROLL_MEAN_MAX_STRAIN = 3, PS3D_PRE_MAX_STRAIN = 4, &
                                 PS3D_PRE_US_MAX_STRAIN = 5
    integer(c_int), parameter :: PS3D_LS_KOLMOGOROV = 0, PS3D_LS_GEOPHYSICAL = 1
    ! length of a diagnostics record (enum ps3d_diag), of ps3d's field_stats record
    ! (enum ps3d_stat) and of one field_stats file record (t, the statistics,
    ! helicity, dt, prefactor: PS3D_NC_STATS_ROW of ps3d_nc.h)
    integer, parameter :: PS3D_NDIAG = 19
    integer, parameter :: PS3D_NSTAT = 40
    integer, parameter :: PS3D_NC_STATS_LEN = PS3D_NSTAT + 4
    ! ps3d_gpu_moments: ke, helicity, cross-helicity mean and largest |u x w|
    integer, parameter :: PS3D_NMOMENT = 4

    type, bind(C) :: ps3d_params
        integer(c_int) :: nx, ny, nz
        real(c_double) :: lower(3), extent(3)
        integer(c_int) :: filtering, stepper, nnu
        real(c_double) :: prediss
        integer(c_int) :: pretype, roll_mean_win_size, length_scale
        real(c_double) :: alpha, limit, dt_fixed, f_cor(3)
    end type ps3d_params

    interface
        subroutine ps3d_params_default(p) bind(C)
            import :: ps3d_params
            type(ps3d_params), intent(out) :: p
        end subroutine
        integer(c_int) function ps3d_gpu_create(p, h) bind(C)
            import :: c_int, c_ptr, ps3d_params
            type(ps3d_params), intent(in) :: p
            type(c_ptr), intent(out) :: h
        end function
        integer(c_int) function ps3d_gpu_destroy(h) bind(C)
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int) function ps3d_gpu_set_vorticity(h, vor) bind(C)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: vor(*)
        end function
        integer(c_int) function ps3d_gpu_vor2vel(h) bind(C)
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int) function ps3d_gpu_derived(h) bind(C)
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int) function ps3d_gpu_adapt(h, t, dt, diag) bind(C)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), value :: t
            real(c_double), intent(out) :: dt
            real(c_double), intent(out) :: diag(*)
        end function
        integer(c_int) function ps3d_gpu_advance_n(h, t, n, diag) bind(C)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(inout) :: t
            integer(c_int), value :: n
            real(c_double), intent(out) :: diag(*)
        end function
        integer(c_int) function ps3d_gpu_advance_until(h, t, t_stop, nmax, diag, nsteps) bind(C)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(inout) :: t
            real(c_double), value :: t_stop
            integer(c_int), value :: nmax
            real(c_double), intent(out) :: diag(*)
            integer(c_int), intent(out) :: nsteps
        end function
        integer(c_int) function ps3d_gpu_field_stats(h, diag, stats) bind(C)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            type(c_ptr), value :: diag
            real(c_double), intent(out) :: stats(*)
        end function
        integer(c_int) function ps3d_gpu_moments(h, out) bind(C)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: out(*)
        end function
        ! genspec's spectrum: vel = c_null_ptr for the state's velocity; spec = c_null_ptr
        ! (with cap 0) only sets nshell and dk
        integer(c_int) function ps3d_gpu_spectrum(h, vel, cap, nshell, dk, spec, num) bind(C)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            type(c_ptr), value :: vel
            integer(c_int), value :: cap
            integer(c_int), intent(out) :: nshell
            real(c_double), intent(out) :: dk
            type(c_ptr), value :: spec
            type(c_ptr), value :: num
        end function
        integer(c_int) function ps3d_gpu_get(h, which, f) bind(C)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int), value :: which
            real(c_double), intent(out) :: f(*)
        end function
        type(c_ptr) function ps3d_gpu_last_error() bind(C)
            import :: c_ptr
        end function

        integer(c_int) function ps3d_nc_grid(path, ncells, extent, origin) bind(C)
            import :: c_int, c_char, c_double
            character(kind=c_char), intent(in) :: path(*)
            integer(c_int), intent(out) :: ncells(3)
            real(c_double), intent(out) :: extent(3), origin(3)
        end function
        integer(c_int) function ps3d_nc_records(path, nrecords, t_last) bind(C)
            import :: c_int, c_char, c_double
            character(kind=c_char), intent(in) :: path(*)
            integer(c_int), intent(out) :: nrecords
            real(c_double), intent(out) :: t_last
        end function
        integer(c_int) function ps3d_nc_read_fields(path, step, t, nx, ny, nz, vel, vor) bind(C)
            import :: c_int, c_char, c_double, c_ptr
            character(kind=c_char), intent(in) :: path(*)
            integer(c_int), value :: step, nx, ny, nz
            real(c_double), intent(out) :: t
            type(c_ptr), value :: vel
            real(c_double), intent(out) :: vor(*)
        end function
        integer(c_int) function ps3d_nc_create_fields(path, nx, ny, nz, lower, extent, flags) bind(C)
            import :: c_int, c_char, c_double
            character(kind=c_char), intent(in) :: path(*)
            integer(c_int), value :: nx, ny, nz, flags
            real(c_double), intent(in) :: lower(3), extent(3)
        end function
        integer(c_int) function ps3d_nc_append_fields(path, t, nx, ny, nz, vel, vor, pres, delta) bind(C)
            import :: c_int, c_char, c_double
            character(kind=c_char), intent(in) :: path(*)
            real(c_double), value :: t
            integer(c_int), value :: nx, ny, nz
            real(c_double), intent(in) :: vel(*), vor(*), pres(*), delta(*)
        end function
        integer(c_int) function ps3d_nc_create_stats(path, nx, ny, nz, lower, extent) bind(C)
            import :: c_int, c_char, c_double
            character(kind=c_char), intent(in) :: path(*)
            integer(c_int), value :: nx, ny, nz
            real(c_double), intent(in) :: lower(3), extent(3)
        end function
        integer(c_int) function ps3d_nc_stats_row(diag, stats, row) bind(C)
            import :: c_int, c_double
            real(c_double), intent(in) :: diag(*), stats(*)
            real(c_double), intent(out) :: row(*)
        end function
        integer(c_int) function ps3d_nc_append_stats(path, rows, nrows) bind(C)
            import :: c_int, c_char, c_double
            character(kind=c_char), intent(in) :: path(*)
            real(c_double), intent(in) :: rows(*)
            integer(c_int), value :: nrows
        end function
        type(c_ptr) function ps3d_nc_last_error() bind(C)
            import :: c_ptr
        end function
        integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function
    end interface

contains

    ! The message behind a C string pointer.
    function c_message(p) result(msg)
        type(c_ptr), intent(in) :: p
        character(len=:), allocatable :: msg
        character(kind=c_char), pointer :: chars(:)
        integer :: n, i
        if (.not. c_associated(p)) then
            msg = ''
            return
        endif
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: msg)
        do i = 1, n
            msg(i:i) = chars(i)
        enddo
    end function

    ! Stop with the library's message on a nonzero status (the reference's mpi_stop,
    ! src/mpi/mpi_utils.f90:15-42).
    subroutine gpu_check(st, what)
        integer(c_int), intent(in) :: st
        character(*), intent(in) :: what
        if (st /= 0) then
            write(*, '(a)') 'Error in ' // what // ': ' // c_message(ps3d_gpu_last_error())
            error stop 1
        endif
    end subroutine

    subroutine nc_check(st, what)
        integer(c_int), intent(in) :: st
        character(*), intent(in) :: what
        if (st /= 0) then
            write(*, '(a)') 'Error in ' // what // ': ' // c_message(ps3d_nc_last_error())
            error stop 1
        endif
    end subroutine

    ! NUL-terminated copy of a Fortran string for the C entry points.
    function cstr(s) result(c)
        character(*), intent(in) :: s
        character(kind=c_char, len=len_trim(s) + 1) :: c
        c = trim(s) // c_null_char
    end function

end module ps3d_gpu
