! Drives the GPU core from Fortran through epic_gpu_mod, the way EPIC's run
! loop (epic3d.f90:128-161) would: a 16^3 Beltrami member on
! [-pi/2, pi/2]^3 with a 2x2x2 lattice of spherical parcels
! (parcel_init.f90:31-152), vorticity omega = alpha u sampled at the parcel
! centres (beltrami_3d.f90:103-122), then two iterations of the run loop:
! apply_vortcor, a fixed-dt RK step, parcel_merge, parcel_split and two
! rounds of apply_laplace / apply_gradient.  Prints the final time, the
! parcel count and checksums of the final parcel state and of the gridded
! volume, which a caller compares with the C-ABI run of the same member.
program drive_beltrami
    use iso_c_binding, only : c_int64_t
    use epic_gpu_mod
    implicit none
    integer, parameter :: nx = 16, ny = 16, nz = 16, npd = 2
    integer, parameter :: n = nx * ny * nz * npd ** 3
    double precision, parameter :: pi = acos(-1.0d0)
    double precision :: lower(3), extent(3), dx(3), vcell, abc, l23, x, y, z, alpha, fk2l2, t
    double precision, allocatable :: pos(:, :), vor(:, :), B(:, :), vol(:), buoy(:), hum(:)
    double precision :: volg(-1:nz+1, -1:ny+1, -1:nx+1)
    integer :: ix, iy, iz, i, j, k, l, step, cor_iter, m

    lower = -0.5d0 * pi
    extent = pi
    dx = extent / dble((/nx, ny, nz/))
    vcell = product(dx)
    allocate(pos(3, n), vor(3, n), B(5, n), vol(n), buoy(n), hum(n))
    alpha = sqrt(3.0d0)
    fk2l2 = alpha / 2.0d0
    l = 0
    do iz = 0, nz - 1
        do iy = 0, ny - 1
            do ix = 0, nx - 1
                do k = 1, npd
                    do j = 1, npd
                        do i = 1, npd
                            l = l + 1
                            pos(1, l) = lower(1) + ix * dx(1) + dx(1) * ((dble(i) - 0.5d0) / npd)
                            pos(2, l) = lower(2) + iy * dx(2) + dx(2) * ((dble(j) - 0.5d0) / npd)
                            pos(3, l) = lower(3) + iz * dx(3) + dx(3) * ((dble(k) - 0.5d0) / npd)
                        enddo
                    enddo
                enddo
            enddo
        enddo
    enddo
    vol = vcell / dble(npd ** 3)
    abc = 0.75d0 * vol(1) / pi
    l23 = abc ** (2.0d0 / 3.0d0)
    B = 0.0d0
    B(1, :) = l23
    B(4, :) = l23
    do l = 1, n
        x = pos(1, l)
        y = pos(2, l)
        z = pos(3, l)
        vor(1, l) = fk2l2 * (sin(z) - alpha * cos(z)) * sin(x + y)
        vor(2, l) = fk2l2 * (sin(z) + alpha * cos(z)) * sin(x + y)
        vor(3, l) = alpha * cos(z) * cos(x + y)
    enddo
    buoy = 0.0d0
    hum = 0.0d0

    call epic_gpu_init(nx, ny, nz, lower, extent, 0.0d0, 0.0d0, 0.0d0, (/0.0d0, 0.0d0, 0.0d0/), &
                       (/.true., .true./), 4, int(2 * n, c_int64_t), 0.01d0, 0.0d0)
    call epic_gpu_upload_parcels(n, pos, vor, B, vol, buoy, hum)
    call init_parcel_correction
    t = 0.0d0
    do step = 1, 2
        call apply_vortcor
        call ls_rk_step(t)
        call parcel_merge
        call parcel_split
        do cor_iter = 1, 2
            call apply_laplace((cor_iter > 1))
            call apply_gradient(1.8d0, 0.5d0, .true.)
        enddo
    enddo
    m = epic_gpu_num_parcels()
    if (m /= n) then
        write(*, *) "parcel count changed: ", m
        error stop 1
    endif
    call epic_gpu_download_parcels(n, pos, vor, B, vol, buoy, hum)
    call par2grid
    call epic_gpu_get_field(EPIC_FIELD_VOLG, volg)
    write(*, '(a, f8.4, a, i8)') "t=", t, " n=", n
    write(*, '(a, 6es24.16)') "checksums", sum(abs(pos)), sum(abs(vor)), sum(B(1, :)), sum(B(4, :)), &
        sum(vol), sum(volg(0:nz, 0:ny-1, 0:nx-1))
    call epic_gpu_finalise
end program drive_beltrami
