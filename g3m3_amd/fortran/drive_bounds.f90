! A Fortran driver fragment in the reference's style for the nested-grid
! boundary calls: bndry_flanks into the wrk arrays (code/3d_main.f90:
! 2351-2387), bndry_flanks with the state as both target and source, the
! aliased call of :2945-2976, and bndry_corer (:3045-3060), each through the
! library's full-signature entry point by an implicit interface, on host
! arrays registered once.  The 34 fields of each set are the slices of one
! 5-D array, passed in the reference's argument order.
!
!   drive_bounds IN OUT
!
! IN  (stream access): int32 nx,ny,nz,n_grids,box; float64 t_old(n_grids),
!     t_new(n_grids), t1, t2; float32 gx0,gx1,gy0,gy1,gz0,gz1 (each
!     n_grids); then the arrays cur, old, wrk, each (nx,ny,nz,n_grids,34) in
!     the field order rho,px,py,pz,presx,presy,presz,presxy,presxz,presyz of
!     q, h and o, then epres,bx,by,bz
! OUT (stream access): wrk, then cur
program drive_bounds
    implicit none
    integer nx, ny, nz, n_grids, box
    real(8), allocatable :: t_old(:), t_new(:)
    real(8) t1, t2
    real, allocatable :: gx0(:), gx1(:), gy0(:), gy1(:), gz0(:), gz1(:)
    real, allocatable, dimension(:,:,:,:,:) :: cur, old, wrk
    real, allocatable :: work(:,:,:)
    integer g3m3_init, g3m3_register, g3m3_to_device, g3m3_to_host, g3m3_last_error
    external g3m3_init, g3m3_register, g3m3_to_device, g3m3_to_host, g3m3_last_error
    integer ierr
    integer(8) nb
    character(len=512) fin, fout

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(10, file=trim(fin), access='stream', form='unformatted', status='old')
    read(10) nx, ny, nz, n_grids, box
    allocate(t_old(n_grids), t_new(n_grids), gx0(n_grids), gx1(n_grids), gy0(n_grids), &
        gy1(n_grids), gz0(n_grids), gz1(n_grids))
    read(10) t_old, t_new, t1, t2
    read(10) gx0, gx1, gy0, gy1, gz0, gz1
    allocate(cur(nx,ny,nz,n_grids,34), old(nx,ny,nz,n_grids,34), wrk(nx,ny,nz,n_grids,34), &
        work(nx,ny,nz))
    read(10) cur, old, wrk
    close(10)

    ierr = g3m3_init()
    if (ierr /= 0) stop 'g3m3_init failed'
    nb = 4_8 * nx * ny * nz * n_grids * 34
    if (g3m3_register(cur, nb) /= 0 .or. g3m3_to_device(cur) /= 0) stop 'register cur'
    if (g3m3_register(old, nb) /= 0 .or. g3m3_to_device(old) /= 0) stop 'register old'
    if (g3m3_register(wrk, nb) /= 0 .or. g3m3_to_device(wrk) /= 0) stop 'register wrk'
    if (g3m3_register(work, 4_8 * nx * ny * nz) /= 0) stop 'register work'

    call g3m3_bndry_flanks(wrk(:,:,:,:,1), wrk(:,:,:,:,2), wrk(:,:,:,:,3), wrk(:,:,:,:,4), &
        wrk(:,:,:,:,5), wrk(:,:,:,:,6), wrk(:,:,:,:,7), wrk(:,:,:,:,8), wrk(:,:,:,:,9), &
        wrk(:,:,:,:,10), wrk(:,:,:,:,11), wrk(:,:,:,:,12), wrk(:,:,:,:,13), &
        wrk(:,:,:,:,14), wrk(:,:,:,:,15), wrk(:,:,:,:,16), wrk(:,:,:,:,17), &
        wrk(:,:,:,:,18), wrk(:,:,:,:,19), wrk(:,:,:,:,20), wrk(:,:,:,:,21), &
        wrk(:,:,:,:,22), wrk(:,:,:,:,23), wrk(:,:,:,:,24), wrk(:,:,:,:,25), &
        wrk(:,:,:,:,26), wrk(:,:,:,:,27), wrk(:,:,:,:,28), wrk(:,:,:,:,29), &
        wrk(:,:,:,:,30), wrk(:,:,:,:,31), wrk(:,:,:,:,32), wrk(:,:,:,:,33), &
        wrk(:,:,:,:,34), cur(:,:,:,:,1), cur(:,:,:,:,2), cur(:,:,:,:,3), cur(:,:,:,:,4), &
        cur(:,:,:,:,5), cur(:,:,:,:,6), cur(:,:,:,:,7), cur(:,:,:,:,8), cur(:,:,:,:,9), &
        cur(:,:,:,:,10), cur(:,:,:,:,11), cur(:,:,:,:,12), cur(:,:,:,:,13), &
        cur(:,:,:,:,14), cur(:,:,:,:,15), cur(:,:,:,:,16), cur(:,:,:,:,17), &
        cur(:,:,:,:,18), cur(:,:,:,:,19), cur(:,:,:,:,20), cur(:,:,:,:,21), &
        cur(:,:,:,:,22), cur(:,:,:,:,23), cur(:,:,:,:,24), cur(:,:,:,:,25), &
        cur(:,:,:,:,26), cur(:,:,:,:,27), cur(:,:,:,:,28), cur(:,:,:,:,29), &
        cur(:,:,:,:,30), cur(:,:,:,:,31), cur(:,:,:,:,32), cur(:,:,:,:,33), &
        cur(:,:,:,:,34), old(:,:,:,:,1), old(:,:,:,:,2), old(:,:,:,:,3), old(:,:,:,:,4), &
        old(:,:,:,:,5), old(:,:,:,:,6), old(:,:,:,:,7), old(:,:,:,:,8), old(:,:,:,:,9), &
        old(:,:,:,:,10), old(:,:,:,:,11), old(:,:,:,:,12), old(:,:,:,:,13), &
        old(:,:,:,:,14), old(:,:,:,:,15), old(:,:,:,:,16), old(:,:,:,:,17), &
        old(:,:,:,:,18), old(:,:,:,:,19), old(:,:,:,:,20), old(:,:,:,:,21), &
        old(:,:,:,:,22), old(:,:,:,:,23), old(:,:,:,:,24), old(:,:,:,:,25), &
        old(:,:,:,:,26), old(:,:,:,:,27), old(:,:,:,:,28), old(:,:,:,:,29), &
        old(:,:,:,:,30), old(:,:,:,:,31), old(:,:,:,:,32), old(:,:,:,:,33), &
        old(:,:,:,:,34), work, nx, ny, nz, n_grids, box, t_old, t_new, t1, gx0, gx1, gy0, &
        gy1, gz0, gz1)
    call g3m3_bndry_flanks(cur(:,:,:,:,1), cur(:,:,:,:,2), cur(:,:,:,:,3), cur(:,:,:,:,4), &
        cur(:,:,:,:,5), cur(:,:,:,:,6), cur(:,:,:,:,7), cur(:,:,:,:,8), cur(:,:,:,:,9), &
        cur(:,:,:,:,10), cur(:,:,:,:,11), cur(:,:,:,:,12), cur(:,:,:,:,13), &
        cur(:,:,:,:,14), cur(:,:,:,:,15), cur(:,:,:,:,16), cur(:,:,:,:,17), &
        cur(:,:,:,:,18), cur(:,:,:,:,19), cur(:,:,:,:,20), cur(:,:,:,:,21), &
        cur(:,:,:,:,22), cur(:,:,:,:,23), cur(:,:,:,:,24), cur(:,:,:,:,25), &
        cur(:,:,:,:,26), cur(:,:,:,:,27), cur(:,:,:,:,28), cur(:,:,:,:,29), &
        cur(:,:,:,:,30), cur(:,:,:,:,31), cur(:,:,:,:,32), cur(:,:,:,:,33), &
        cur(:,:,:,:,34), cur(:,:,:,:,1), cur(:,:,:,:,2), cur(:,:,:,:,3), cur(:,:,:,:,4), &
        cur(:,:,:,:,5), cur(:,:,:,:,6), cur(:,:,:,:,7), cur(:,:,:,:,8), cur(:,:,:,:,9), &
        cur(:,:,:,:,10), cur(:,:,:,:,11), cur(:,:,:,:,12), cur(:,:,:,:,13), &
        cur(:,:,:,:,14), cur(:,:,:,:,15), cur(:,:,:,:,16), cur(:,:,:,:,17), &
        cur(:,:,:,:,18), cur(:,:,:,:,19), cur(:,:,:,:,20), cur(:,:,:,:,21), &
        cur(:,:,:,:,22), cur(:,:,:,:,23), cur(:,:,:,:,24), cur(:,:,:,:,25), &
        cur(:,:,:,:,26), cur(:,:,:,:,27), cur(:,:,:,:,28), cur(:,:,:,:,29), &
        cur(:,:,:,:,30), cur(:,:,:,:,31), cur(:,:,:,:,32), cur(:,:,:,:,33), &
        cur(:,:,:,:,34), old(:,:,:,:,1), old(:,:,:,:,2), old(:,:,:,:,3), old(:,:,:,:,4), &
        old(:,:,:,:,5), old(:,:,:,:,6), old(:,:,:,:,7), old(:,:,:,:,8), old(:,:,:,:,9), &
        old(:,:,:,:,10), old(:,:,:,:,11), old(:,:,:,:,12), old(:,:,:,:,13), &
        old(:,:,:,:,14), old(:,:,:,:,15), old(:,:,:,:,16), old(:,:,:,:,17), &
        old(:,:,:,:,18), old(:,:,:,:,19), old(:,:,:,:,20), old(:,:,:,:,21), &
        old(:,:,:,:,22), old(:,:,:,:,23), old(:,:,:,:,24), old(:,:,:,:,25), &
        old(:,:,:,:,26), old(:,:,:,:,27), old(:,:,:,:,28), old(:,:,:,:,29), &
        old(:,:,:,:,30), old(:,:,:,:,31), old(:,:,:,:,32), old(:,:,:,:,33), &
        old(:,:,:,:,34), work, nx, ny, nz, n_grids, box, t_old, t_new, t2, gx0, gx1, gy0, &
        gy1, gz0, gz1)
    call g3m3_bndry_corer(cur(:,:,:,:,1), cur(:,:,:,:,5), cur(:,:,:,:,6), cur(:,:,:,:,7), &
        cur(:,:,:,:,2), cur(:,:,:,:,3), cur(:,:,:,:,4), cur(:,:,:,:,11), cur(:,:,:,:,15), &
        cur(:,:,:,:,16), cur(:,:,:,:,17), cur(:,:,:,:,12), cur(:,:,:,:,13), &
        cur(:,:,:,:,14), cur(:,:,:,:,21), cur(:,:,:,:,25), cur(:,:,:,:,26), &
        cur(:,:,:,:,27), cur(:,:,:,:,22), cur(:,:,:,:,23), cur(:,:,:,:,24), &
        cur(:,:,:,:,31), cur(:,:,:,:,32), cur(:,:,:,:,33), cur(:,:,:,:,34), cur(:,:,:,:,8), &
        cur(:,:,:,:,9), cur(:,:,:,:,10), cur(:,:,:,:,18), cur(:,:,:,:,19), cur(:,:,:,:,20), &
        cur(:,:,:,:,28), cur(:,:,:,:,29), cur(:,:,:,:,30), nx, ny, nz, n_grids, box, gx0, &
        gx1, gy0, gy1, gz0, gz1)
    ierr = g3m3_last_error()
    if (ierr /= 0) stop 'g3m3 call failed'

    if (g3m3_to_host(wrk) /= 0 .or. g3m3_to_host(cur) /= 0) stop 'g3m3_to_host failed'
    open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(11) wrk, cur
    close(11)
end program
