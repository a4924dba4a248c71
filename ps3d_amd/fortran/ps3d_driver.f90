! ps3d_driver.f90 -- the ps3d program (src/ps3d.f90) on the MI355X step library.
!
!   ps3d_driver --config <file>
!
! Reads the &PS3D namelist of src/utils/options.f90:97-141 (same groups and names:
! field_file, field_step, stepper, vor_visc, filtering, output, time), the grid box
! and initial vorticity from the field file (setup_domain_and_parameters and
! read_netcdf_fields, src/utils/utils.f90:91-134, src/fields/field_netcdf.f90:276-433),
! runs the time loop of src/ps3d.f90:114-126 with the output rule of write_step
! (src/utils/utils.f90:53-89) and writes <basename>_fields.nc and
! <basename>_field_stats.nc (ps3d's 40 statistics per record).  Files are NetCDF
! classic: 64-bit offset, or 64-bit data (CDF-5) where a field record exceeds 4 GiB.
! Extension: time%dt_fixed > 0 fixes the step (ps3d always adapts it).
program ps3d_driver
    use iso_c_binding
    use ps3d_gpu
    implicit none

    ! --- &PS3D namelist (src/utils/options.f90:13-91) ---
    type info
        double precision                  :: field_freq         = 1.0d0
        logical                           :: write_fields       = .true.
        character(len=32), dimension(128) :: field_list         = ''
        double precision                  :: field_stats_freq   = 1.0d0
        logical                           :: write_field_stats  = .true.
        logical                           :: overwrite          = .false.
        character(len=512)                :: basename           = ''
        logical                           :: l_balanced         = .false.
    end type info
    type visc_type
        integer            :: nnu = 3
        double precision   :: prediss = 10.0d0
        character(len=20)  :: pretype = 'roll-mean-max-strain'
        integer            :: roll_mean_win_size = 1000
        character(len=11)  :: length_scale = 'Kolmogorov'
    end type visc_type
    type time_info_type
        double precision :: initial      = 0.0d0
        double precision :: limit        = 0.0d0
        double precision :: alpha        = 0.1d0
        logical          :: precise_stop = .false.
        double precision :: dt_fixed     = 0.0d0
    end type time_info_type

    character(len=512)   :: field_file = ''
    integer              :: field_step = -1
    character(len=16)    :: stepper = 'impl-diff-rk4'
    character(len=8)     :: filtering = 'Hou & Li'
    type(info)           :: output
    type(visc_type)      :: vor_visc
    type(time_info_type) :: time
    namelist /PS3D/ field_file, field_step, stepper, vor_visc, filtering, output, time

    character(len=512) :: cfg, arg
    character(len=:), allocatable :: ffile, sfile
    type(ps3d_params) :: p
    type(c_ptr) :: h
    integer(c_int) :: ncells(3), nrec, st
    real(c_double) :: extent(3), origin(3), t, t0, tlast, dt, tstop
    real(c_double) :: row(PS3D_NDIAG), stats(PS3D_NSTAT), srow(PS3D_NC_STATS_LEN)
    logical :: wf, ws
    integer(c_int) :: nmax
    ! steps queued per call of ps3d_gpu_advance_until (replays of one captured step)
    integer(c_int), parameter :: nchunk = 32
    real(c_double) :: rows(PS3D_NDIAG, nchunk)
    integer(c_int) :: nd
    integer :: k
    real(c_double), allocatable, target :: vor(:), vel(:), pres(:), delta(:)
    integer :: ios, fn, i, nfw, nsfw
    integer(c_size_t) :: nv
    logical :: exists
    double precision, parameter :: eps = epsilon(1.0d0)

    ! --- command line (src/ps3d.f90:155-183) ---
    cfg = ''
    i = 1
    do while (i <= command_argument_count())
        call get_command_argument(i, arg)
        if (trim(arg) == '--config') then
            i = i + 1
            call get_command_argument(i, cfg)
        endif
        i = i + 1
    enddo
    if (len_trim(cfg) == 0) then
        write(*, '(a)') 'Usage: ps3d_driver --config <file>'
        error stop 1
    endif

    ! --- read_config_file ---
    inquire(file=trim(cfg), exist=exists)
    if (.not. exists) then
        write(*, '(a)') 'Error: input file "' // trim(cfg) // '" does not exist.'
        error stop 1
    endif
    open(action='read', file=trim(cfg), iostat=ios, newunit=fn)
    read(nml=PS3D, iostat=ios, unit=fn)
    if (ios /= 0) then
        write(*, '(a)') 'Error: invalid Namelist format.'
        error stop 1
    endif
    close(fn)

    ! --- setup_domain_and_parameters: grid, box and start time from the field file ---
    call nc_check(ps3d_nc_grid(cstr(field_file), ncells, extent, origin), 'reading the field file grid')
    call nc_check(ps3d_nc_records(cstr(field_file), nrec, tlast), 'reading the field file records')

    call ps3d_params_default(p)
    p%nx = ncells(1)
    p%ny = ncells(2)
    p%nz = ncells(3)
    p%lower = origin
    p%extent = extent
    select case (trim(filtering))
        case ('Hou & Li')
            p%filtering = PS3D_FILTER_HOU_LI
        case ('2/3-rule')
            p%filtering = PS3D_FILTER_23RULE
        case default
            write(*, '(a)') "Error: no filter called '" // trim(filtering) // "'."
            error stop 1
    end select
    select case (trim(stepper))
        case ('impl-diff-rk4')
            p%stepper = PS3D_STEPPER_IMPL_RK4
        case ('cn2')
            p%stepper = PS3D_STEPPER_CN2
        case default
            write(*, '(a)') "No stepper called '" // trim(stepper) // "' available."
            error stop 1
    end select
    p%nnu = vor_visc%nnu
    p%prediss = vor_visc%prediss
    select case (trim(vor_visc%pretype))
        case ('constant')
            p%pretype = PS3D_PRE_CONSTANT
        case ('vorch')
            p%pretype = PS3D_PRE_VORCH
        case ('bfmax')
            p%pretype = PS3D_PRE_BFMAX
        case ('roll-mean-max-strain')
            p%pretype = PS3D_PRE_ROLL_MEAN_MAX_STRAIN
        case ('max-strain')
            p%pretype = PS3D_PRE_MAX_STRAIN
        case ('us-max-strain')
            p%pretype = PS3D_PRE_US_MAX_STRAIN
        case default
            write(*, '(a)') "Error: no prefactor type '" // trim(vor_visc%pretype) // "'."
            error stop 1
    end select
    p%roll_mean_win_size = vor_visc%roll_mean_win_size
    if (trim(vor_visc%length_scale) == 'geophysical') then
        p%length_scale = PS3D_LS_GEOPHYSICAL
    else
        p%length_scale = PS3D_LS_KOLMOGOROV
    endif
    p%alpha = time%alpha
    ! adapt clips every step to time%limit - t (src/stepper/advance.f90:325-333);
    ! time%precise_stop is read but not used by the step (options.f90:167)
    p%limit = time%limit
    p%dt_fixed = time%dt_fixed

    nv = int(p%nx, c_size_t) * int(p%ny, c_size_t) * int(p%nz + 1, c_size_t)
    allocate(vor(3 * nv), vel(3 * nv), pres(nv), delta(nv))

    ! --- setup_fields: the vorticity of record field_step (src/utils/utils.f90:136-184) ---
    call nc_check(ps3d_nc_read_fields(cstr(field_file), int(merge(field_step, -1, field_step >= 1), c_int), &
                                      t0, p%nx, p%ny, p%nz, c_null_ptr, vor), 'reading the initial vorticity')
    call gpu_check(ps3d_gpu_create(p, h), 'ps3d_gpu_create')
    call gpu_check(ps3d_gpu_set_vorticity(h, vor), 'ps3d_gpu_set_vorticity')

    ! --- setup_output_files (resume as create_netcdf_field_file does) ---
    if (len_trim(output%basename) == 0) then
        write(*, '(a)') 'Error: output%basename is empty.'
        error stop 1
    endif
    ffile = trim(output%basename) // '_fields.nc'
    sfile = trim(output%basename) // '_field_stats.nc'
    t = t0
    nfw = 0
    nsfw = 0
    if (t0 > 0.0d0) nfw = int(t0 / output%field_freq)
    call open_or_resume(ffile, .true.)
    call open_or_resume(sfile, .false.)

    ! --- run (src/ps3d.f90:114-126): advance = vor2vel, adapt, write_step, source, step ---
    ! The loop test t < limit runs on the device: each call queues up to nchunk steps
    ! and stops before the first one that would start at or after the next output
    ! time of write_step (t + eps >= n * freq, src/utils/utils.f90:62-89).  At an
    ! output time the physical fields (vor2vel, pressure and divergence) are formed
    ! first; the fields are written, and the field_stats record takes the adapt
    ! values of the step that starts there (its diagnostics row).
    do while (t < time%limit)
        wf = output%write_fields .and. t >= dble(nfw) * output%field_freq - eps
        ws = output%write_field_stats .and. t >= dble(nsfw) * output%field_stats_freq - eps
        if (wf .or. ws) then
            call gpu_check(ps3d_gpu_vor2vel(h), 'ps3d_gpu_vor2vel')
            call gpu_check(ps3d_gpu_derived(h), 'ps3d_gpu_derived')
        endif
        if (wf) then
            call write_fields(t)
            nfw = nfw + 1
        endif
        if (ws) then
            call gpu_check(ps3d_gpu_field_stats(h, c_null_ptr, stats), 'ps3d_gpu_field_stats')
            nsfw = nsfw + 1
        endif
        tstop = time%limit
        if (output%write_fields) tstop = min(tstop, dble(nfw) * output%field_freq - eps)
        if (output%write_field_stats) tstop = min(tstop, dble(nsfw) * output%field_stats_freq - eps)
        nmax = nchunk
        if (tstop <= t) then
            ! an output interval shorter than the step: one step, then write again
            tstop = time%limit
            nmax = 1
        endif
        call gpu_check(ps3d_gpu_advance_until(h, t, tstop, nmax, rows, nd), 'ps3d_gpu_advance_until')
        if (nd == 0 .and. t < tstop) then
            write(*, '(a)') 'Error: no step taken before the stop time.'
            error stop 1
        endif
        if (ws .and. nd > 0) then
            call nc_check(ps3d_nc_stats_row(rows(:, 1), stats, srow), 'write_netcdf_field_stats')
            call nc_check(ps3d_nc_append_stats(cstr(sfile), srow, 1_c_int), 'write_netcdf_field_stats')
        endif
        do k = 1, nd
            write(*, '(a, es22.14, a, es22.14)') ' At time', rows(1, k), ' and time step', rows(2, k)
        enddo
    enddo

    ! write_last_step (src/utils/utils.f90:53-60): forced write at the final time
    if (t > t0) then
        call gpu_check(ps3d_gpu_vor2vel(h), 'ps3d_gpu_vor2vel')
        call gpu_check(ps3d_gpu_derived(h), 'ps3d_gpu_derived')
        if (output%write_fields) call write_fields(t)
        if (output%write_field_stats) then
            call gpu_check(ps3d_gpu_field_stats(h, c_null_ptr, stats), 'ps3d_gpu_field_stats')
            call gpu_check(ps3d_gpu_adapt(h, t, dt, row), 'ps3d_gpu_adapt')
            call nc_check(ps3d_nc_stats_row(row, stats, srow), 'write_netcdf_field_stats')
            call nc_check(ps3d_nc_append_stats(cstr(sfile), srow, 1_c_int), 'write_netcdf_field_stats')
        endif
    endif

    call gpu_check(ps3d_gpu_destroy(h), 'ps3d_gpu_destroy')

contains

    ! write_netcdf_fields of the state of the last vor2vel and derived (pressure and
    ! horizontal divergence)
    subroutine write_fields(tt)
        real(c_double), intent(in) :: tt
        call gpu_check(ps3d_gpu_get(h, PS3D_VEL, vel), 'ps3d_gpu_get')
        call gpu_check(ps3d_gpu_get(h, PS3D_VOR, vor), 'ps3d_gpu_get')
        call gpu_check(ps3d_gpu_get(h, PS3D_PRES, pres), 'ps3d_gpu_get')
        call gpu_check(ps3d_gpu_get(h, PS3D_DELTA, delta), 'ps3d_gpu_get')
        call nc_check(ps3d_nc_append_fields(cstr(ffile), tt, p%nx, p%ny, p%nz, vel, vor, pres, delta), &
                      'write_netcdf_fields')
    end subroutine

    ! An existing file with records is continued, whatever output%overwrite says, as
    ! create_netcdf_field_file does (field_netcdf.f90:75-95); otherwise it is
    ! created.  Records at or before the last time in the file are not written
    ! again: the counters start past it.
    subroutine open_or_resume(path, fields)
        character(*), intent(in) :: path
        logical, intent(in) :: fields
        integer(c_int) :: n
        real(c_double) :: tl
        logical :: ex
        inquire(file=path, exist=ex)
        if (ex) then
            if (ps3d_nc_records(cstr(path), n, tl) == 0) then
                if (n > 0) then
                    if (fields) then
                        nfw = max(nfw, int(tl / output%field_freq) + 1)
                    else
                        nsfw = max(nsfw, int(tl / output%field_stats_freq) + 1)
                    endif
                    return
                endif
            endif
        endif
        if (fields) then
            call nc_check(ps3d_nc_create_fields(cstr(path), p%nx, p%ny, p%nz, p%lower, p%extent, 1_c_int), &
                          'create_netcdf_field_file')
        else
            call nc_check(ps3d_nc_create_stats(cstr(path), p%nx, p%ny, p%nz, p%lower, p%extent), &
                          'create_netcdf_field_stats_file')
        endif
    end subroutine

end program ps3d_driver
