! =============================================================================
! Fortran binding of libepic_gpu.so (ISO_C_BINDING).
!
! A Fortran driver replaces the reference's per-timestep core with these
! procedures; the names and argument meanings follow the reference:
!   par2grid(l_reuse)       parcel_interpl.f90:146
!   grid2par(add)           parcel_interpl.f90:405
!   vor2vel()               inversion.f90:25
!   vorticity_tendency()    inversion.f90:321
!   ls_rk_step(t)           ls_rk.f90:88   (fixed, or adaptive after
!                           epic_gpu_set_adaptive_dt)
!   get_time_step(t)        rk_utils.f90:73
!   parcel_split()          parcel_split.f90:37
!   parcel_merge()          parcel_merge.f90:48
!   init_parcel_correction  parcel_correction.f90:42
!   apply_vortcor           parcel_correction.f90:89
!   apply_laplace(l_reuse)  parcel_correction.f90:142
!   apply_gradient(prefactor, max_compression, l_reuse)
!                           parcel_correction.f90:183
! Parcel and field arrays keep the reference's layouts: position(3, n),
! vorticity(3, n), B(5, n), volume(n), ... and f(-1:nz+1, -1:ny+1, -1:nx+1).
! A non-zero status from the library stops the run with its message, as the
! reference's mpi_exit_on_error does (mpi_utils.f90:15-33).
! =============================================================================
module epic_gpu_mod
    use iso_c_binding
    implicit none

    private

    type, bind(C) :: epic_grid_t
        integer(c_int) :: nx, ny, nz
        real(c_double) :: lower(3), extent(3)
    end type epic_grid_t

    type, bind(C) :: epic_phys_t
        real(c_double) :: glat, lambda_c, q0
        real(c_double) :: f_cor(3)
    end type epic_phys_t

    type, bind(C) :: epic_flags_t
        integer(c_int) :: zeta_bndry_zero(2)
        integer(c_int) :: rk_order
        real(c_double) :: vmin
    end type epic_flags_t

    ! epic_time_t: time options (options.f90:99-107)
    type, bind(C) :: epic_time_t
        integer(c_int) :: fixed
        real(c_double) :: fixed_dt, alpha, limit
        integer(c_int) :: precise_stop
    end type epic_time_t

    integer(c_int), parameter, public :: EPIC_FIELD_VOLG = 0, EPIC_FIELD_VORTG = 1, EPIC_FIELD_TBUOYG = 2, &
                                         EPIC_FIELD_DBUOYG = 3, EPIC_FIELD_HUMG = 4, EPIC_FIELD_VELOG = 5, &
                                         EPIC_FIELD_VELGRADG = 6, EPIC_FIELD_VTEND = 7

    interface
        integer(c_int) function c_epic_create(grid, phys, flags, max_parcels, h) bind(C, name="epic_create")
            import :: c_int, c_int64_t, c_ptr, epic_grid_t, epic_phys_t, epic_flags_t
            type(epic_grid_t), intent(in)  :: grid
            type(epic_phys_t), intent(in)  :: phys
            type(epic_flags_t), intent(in) :: flags
            integer(c_int64_t), value      :: max_parcels
            type(c_ptr), intent(out)       :: h
        end function
        integer(c_int) function c_epic_destroy(h) bind(C, name="epic_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function
        type(c_ptr) function c_epic_last_error(h) bind(C, name="epic_last_error")
            import :: c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int) function c_epic_set_parcels(h, n, pos, vor, B, vol, buoy, hum) bind(C, name="epic_set_parcels")
            import :: c_int, c_int64_t, c_ptr, c_double
            type(c_ptr), value        :: h
            integer(c_int64_t), value :: n
            real(c_double), intent(in) :: pos(*), vor(*), B(*), vol(*), buoy(*), hum(*)
        end function
        integer(c_int) function c_epic_get_parcels(h, n, pos, vor, B, vol, buoy, hum) bind(C, name="epic_get_parcels")
            import :: c_int, c_int64_t, c_ptr, c_double
            type(c_ptr), value        :: h
            integer(c_int64_t), value :: n
            real(c_double), intent(out) :: pos(*), vor(*), B(*), vol(*), buoy(*), hum(*)
        end function
        integer(c_int) function c_epic_get_field(h, which, out) bind(C, name="epic_get_field")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: h
            integer(c_int), value :: which
            real(c_double), intent(out) :: out(*)
        end function
        integer(c_int) function c_epic_par2grid(h, l_reuse) bind(C, name="epic_par2grid")
            import :: c_int, c_ptr
            type(c_ptr), value    :: h
            integer(c_int), value :: l_reuse
        end function
        integer(c_int) function c_epic_grid2par(h, add) bind(C, name="epic_grid2par")
            import :: c_int, c_ptr
            type(c_ptr), value    :: h
            integer(c_int), value :: add
        end function
        integer(c_int) function c_epic_vor2vel(h) bind(C, name="epic_vor2vel")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int) function c_epic_vorticity_tendency(h) bind(C, name="epic_vorticity_tendency")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int) function c_epic_ls_rk_step(h, t, dt, t_initial) bind(C, name="epic_ls_rk_step")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value          :: h
            real(c_double), intent(inout) :: t
            real(c_double), value       :: dt, t_initial
        end function
        integer(c_int) function c_epic_ls_rk_step_adaptive(h, t, opts, t_initial, dt) &
                bind(C, name="epic_ls_rk_step_adaptive")
            import :: c_int, c_ptr, c_double, epic_time_t
            type(c_ptr), value            :: h
            real(c_double), intent(inout) :: t
            type(epic_time_t), intent(in) :: opts
            real(c_double), value         :: t_initial
            real(c_double), intent(out)   :: dt
        end function
        integer(c_int) function c_epic_get_time_step(h, t, opts, dt, gmax, bmax) bind(C, name="epic_get_time_step")
            import :: c_int, c_ptr, c_double, epic_time_t
            type(c_ptr), value            :: h
            real(c_double), value         :: t
            type(epic_time_t), intent(in) :: opts
            real(c_double), intent(out)   :: dt, gmax, bmax
        end function
        integer(c_int) function c_epic_parcel_split(h, lambda_max, amax, n_split) bind(C, name="epic_parcel_split")
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value              :: h
            real(c_double), value           :: lambda_max, amax
            integer(c_int64_t), intent(out) :: n_split
        end function
        integer(c_int) function c_epic_parcel_merge(h, n_merge) bind(C, name="epic_parcel_merge")
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value              :: h
            integer(c_int64_t), intent(out) :: n_merge
        end function
        integer(c_int) function c_epic_apply_laplace(h, l_reuse) bind(C, name="epic_apply_laplace")
            import :: c_int, c_ptr
            type(c_ptr), value    :: h
            integer(c_int), value :: l_reuse
        end function
        integer(c_int) function c_epic_apply_gradient(h, prefactor, max_compression, l_reuse) &
                bind(C, name="epic_apply_gradient")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: h
            real(c_double), value :: prefactor, max_compression
            integer(c_int), value :: l_reuse
        end function
        integer(c_int) function c_epic_init_parcel_correction(h, vor_bar) bind(C, name="epic_init_parcel_correction")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value          :: h
            real(c_double), intent(out) :: vor_bar(3)
        end function
        integer(c_int) function c_epic_apply_vortcor(h) bind(C, name="epic_apply_vortcor")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int64_t) function c_epic_num_parcels(h) bind(C, name="epic_num_parcels")
            import :: c_int64_t, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function
    end interface

    ! the member this run works on (the reference's module globals)
    type(c_ptr) :: handle = c_null_ptr
    integer(c_int64_t) :: num_parcels = 0
    double precision :: fixed_dt = 0.0d0, time_initial = 0.0d0
    logical :: adaptive_dt = .false.
    type(epic_time_t) :: time_opts

    public :: epic_gpu_init, epic_gpu_finalise, epic_gpu_upload_parcels, epic_gpu_download_parcels, &
              epic_gpu_get_field, epic_gpu_num_parcels, par2grid, grid2par, vor2vel, vorticity_tendency, &
              ls_rk_step, get_time_step, epic_gpu_set_adaptive_dt, parcel_split, parcel_merge, init_parcel_correction, apply_vortcor, apply_laplace, &
              apply_gradient

contains

    subroutine check(rc, where)
        integer(c_int), intent(in) :: rc
        character(*), intent(in)   :: where
        character(kind=c_char), pointer :: msg(:)
        type(c_ptr) :: p
        integer :: n
        if (rc == 0) return
        p = c_epic_last_error(handle)
        n = int(c_strlen(p))
        call c_f_pointer(p, msg, [n])
        write(*, *) "epic_gpu: ", where, " failed (", rc, "): ", msg(1:n)
        error stop 1
    end subroutine check

    ! epic_create: field_alloc + parcels%allocate + init_inversion + ls_rk_setup
    subroutine epic_gpu_init(nx, ny, nz, lower, extent, glat, lambda_c, q0, f_cor, zeta_zero, rk_order, &
                             max_parcels, dt, t_initial)
        integer, intent(in)          :: nx, ny, nz, rk_order
        double precision, intent(in) :: lower(3), extent(3), glat, lambda_c, q0, f_cor(3), dt, t_initial
        logical, intent(in)          :: zeta_zero(2)
        integer(c_int64_t), intent(in) :: max_parcels
        type(epic_grid_t)  :: g
        type(epic_phys_t)  :: p
        type(epic_flags_t) :: f
        g%nx = nx
        g%ny = ny
        g%nz = nz
        g%lower = lower
        g%extent = extent
        p%glat = glat
        p%lambda_c = lambda_c
        p%q0 = q0
        p%f_cor = f_cor
        f%zeta_bndry_zero = merge(1, 0, zeta_zero)
        f%rk_order = rk_order
        f%vmin = 0.0d0
        call check(c_epic_create(g, p, f, max_parcels, handle), "epic_create")
        fixed_dt = dt
        time_initial = t_initial
    end subroutine epic_gpu_init

    subroutine epic_gpu_finalise
        integer(c_int) :: rc
        rc = c_epic_destroy(handle)
        handle = c_null_ptr
    end subroutine epic_gpu_finalise

    subroutine epic_gpu_upload_parcels(n, position, vorticity, B, volume, buoyancy, humidity)
        integer, intent(in)          :: n
        double precision, intent(in) :: position(3, n), vorticity(3, n), B(5, n), volume(n), buoyancy(n), humidity(n)
        num_parcels = int(n, c_int64_t)
        call check(c_epic_set_parcels(handle, num_parcels, position, vorticity, B, volume, buoyancy, humidity), &
                   "epic_set_parcels")
    end subroutine epic_gpu_upload_parcels

    subroutine epic_gpu_download_parcels(n, position, vorticity, B, volume, buoyancy, humidity)
        integer, intent(in)           :: n
        double precision, intent(out) :: position(3, n), vorticity(3, n), B(5, n), volume(n), buoyancy(n), humidity(n)
        call check(c_epic_get_parcels(handle, int(n, c_int64_t), position, vorticity, B, volume, buoyancy, humidity), &
                   "epic_get_parcels")
    end subroutine epic_gpu_download_parcels

    ! f(-1:nz+1, -1:ny+1, -1:nx+1[, ncomp])
    subroutine epic_gpu_get_field(which, f)
        integer(c_int), intent(in)    :: which
        double precision, intent(out) :: f(*)
        call check(c_epic_get_field(handle, which, f), "epic_get_field")
    end subroutine epic_gpu_get_field

    subroutine par2grid(l_reuse)
        logical, optional :: l_reuse
        integer(c_int) :: flag
        flag = -1
        if (present(l_reuse)) flag = merge(1, 0, l_reuse)
        call check(c_epic_par2grid(handle, flag), "par2grid")
    end subroutine par2grid

    subroutine grid2par(add)
        logical, optional, intent(in) :: add
        integer(c_int) :: flag
        flag = -1
        if (present(add)) flag = merge(1, 0, add)
        call check(c_epic_grid2par(handle, flag), "grid2par")
    end subroutine grid2par

    subroutine vor2vel
        call check(c_epic_vor2vel(handle), "vor2vel")
    end subroutine vor2vel

    subroutine vorticity_tendency
        call check(c_epic_vorticity_tendency(handle), "vorticity_tendency")
    end subroutine vorticity_tendency

    ! time%l_use_fixed_dt = .false.: the adaptive step of get_time_step with
    ! time%alpha, time%limit and time%precise_stop
    subroutine epic_gpu_set_adaptive_dt(alpha, limit, precise_stop)
        double precision, intent(in) :: alpha, limit
        logical, intent(in)          :: precise_stop
        adaptive_dt = .true.
        time_opts%fixed = 0
        time_opts%fixed_dt = fixed_dt
        time_opts%alpha = alpha
        time_opts%limit = limit
        time_opts%precise_stop = merge(1, 0, precise_stop)
    end subroutine epic_gpu_set_adaptive_dt

    subroutine ls_rk_step(t)
        double precision, intent(inout) :: t
        real(c_double) :: tt, dt
        tt = t
        if (adaptive_dt) then
            call check(c_epic_ls_rk_step_adaptive(handle, tt, time_opts, time_initial, dt), "ls_rk_step")
        else
            call check(c_epic_ls_rk_step(handle, tt, fixed_dt, time_initial), "ls_rk_step")
        endif
        t = tt
    end subroutine ls_rk_step

    ! get_time_step from the current fields (after par2grid and vor2vel)
    function get_time_step(t) result(dt)
        double precision, intent(in) :: t
        double precision :: dt
        type(epic_time_t) :: o
        real(c_double) :: h, gmax, bmax
        o = time_opts
        if (.not. adaptive_dt) then
            o%fixed = 1
            o%fixed_dt = fixed_dt
        endif
        call check(c_epic_get_time_step(handle, t, o, h, gmax, bmax), "get_time_step")
        dt = h
    end function get_time_step

    ! parcel_split with the reference's defaults (parcel%lambda_max = 4,
    ! amax from the grid); the parcel count grows by the number of splits
    subroutine parcel_split(lambda_max)
        double precision, optional, intent(in) :: lambda_max
        integer(c_int64_t) :: n_split
        real(c_double) :: lam
        lam = 4.0d0
        if (present(lambda_max)) lam = lambda_max
        call check(c_epic_parcel_split(handle, lam, 0.0d0, n_split), "parcel_split")
        num_parcels = c_epic_num_parcels(handle)
    end subroutine parcel_split

    subroutine parcel_merge
        integer(c_int64_t) :: n_merge
        call check(c_epic_parcel_merge(handle, n_merge), "parcel_merge")
        num_parcels = c_epic_num_parcels(handle)
    end subroutine parcel_merge

    subroutine init_parcel_correction
        real(c_double) :: vor_bar(3)
        call check(c_epic_init_parcel_correction(handle, vor_bar), "init_parcel_correction")
    end subroutine init_parcel_correction

    subroutine apply_vortcor
        call check(c_epic_apply_vortcor(handle), "apply_vortcor")
    end subroutine apply_vortcor

    subroutine apply_laplace(l_reuse)
        logical, optional, intent(in) :: l_reuse
        integer(c_int) :: flag
        flag = -1
        if (present(l_reuse)) flag = merge(1, 0, l_reuse)
        call check(c_epic_apply_laplace(handle, flag), "apply_laplace")
    end subroutine apply_laplace

    subroutine apply_gradient(prefactor, max_compression, l_reuse)
        double precision, intent(in)  :: prefactor, max_compression
        logical, optional, intent(in) :: l_reuse
        integer(c_int) :: flag
        flag = -1
        if (present(l_reuse)) flag = merge(1, 0, l_reuse)
        call check(c_epic_apply_gradient(handle, prefactor, max_compression, flag), "apply_gradient")
    end subroutine apply_gradient

    integer function epic_gpu_num_parcels()
        epic_gpu_num_parcels = int(c_epic_num_parcels(handle))
    end function epic_gpu_num_parcels

end module epic_gpu_mod
