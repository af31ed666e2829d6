! The driver fragment of drive_fields.f90 through explicit interfaces: the
! same calls with `use g3m3_iface` (g3m3_iface.f90, bind(c) interfaces whose
! dummies are arrays or scalars as in the reference subroutines), so the
! compiler checks every argument against the reference's declarations.  The
! list counts nummid and numzero are arrays (code/sub_BandE.f90:22-23).
!
!   drive_iface IN OUT   (the files of drive_fields)
program drive_iface
    use g3m3_iface
    implicit none
    integer nx, ny, nz, n_grids, box, mbndry, nummid(1), numzero(1), mmid, mzero
    real rx, delt, rmassq, rmassh, rmasso, reynolds, resist, gamma, gamma1, smallbit
    real, allocatable, dimension(:,:,:,:) :: bx, by, bz, bx0, by0, bz0, oldbx, oldby, oldbz, &
        qpx, qpy, qpz, qrho, hpx, hpy, hpz, hrho, opx, opy, opz, orho, &
        epres, wrkepres, oldepres, rst
    real, allocatable, dimension(:,:,:) :: curx, cury, curz, bsx, bsy, bsz, btot, &
        evx, evy, evz, tvx, tvy, tvz, efldx, efldy, efldz
    integer ijmid(1,3,1), ijzero(1,3,1)
    integer ierr
    integer(8) n4, n3
    character(len=512) fin, fout

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(10, file=trim(fin), access='stream', form='unformatted', status='old')
    read(10) nx, ny, nz, n_grids, box
    read(10) rx, delt, rmassq, rmassh, rmasso, reynolds, resist, gamma, gamma1, smallbit
    allocate(bx(nx,ny,nz,n_grids), by(nx,ny,nz,n_grids), bz(nx,ny,nz,n_grids), &
        bx0(nx,ny,nz,n_grids), by0(nx,ny,nz,n_grids), bz0(nx,ny,nz,n_grids), &
        oldbx(nx,ny,nz,n_grids), oldby(nx,ny,nz,n_grids), oldbz(nx,ny,nz,n_grids), &
        qpx(nx,ny,nz,n_grids), qpy(nx,ny,nz,n_grids), qpz(nx,ny,nz,n_grids), &
        qrho(nx,ny,nz,n_grids), hpx(nx,ny,nz,n_grids), hpy(nx,ny,nz,n_grids), &
        hpz(nx,ny,nz,n_grids), hrho(nx,ny,nz,n_grids), opx(nx,ny,nz,n_grids), &
        opy(nx,ny,nz,n_grids), opz(nx,ny,nz,n_grids), orho(nx,ny,nz,n_grids), &
        epres(nx,ny,nz,n_grids), wrkepres(nx,ny,nz,n_grids), oldepres(nx,ny,nz,n_grids), &
        rst(nx,ny,nz,1))
    allocate(curx(nx,ny,nz), cury(nx,ny,nz), curz(nx,ny,nz), bsx(nx,ny,nz), &
        bsy(nx,ny,nz), bsz(nx,ny,nz), btot(nx,ny,nz), evx(nx,ny,nz), evy(nx,ny,nz), &
        evz(nx,ny,nz), tvx(nx,ny,nz), tvy(nx,ny,nz), tvz(nx,ny,nz), efldx(nx,ny,nz), &
        efldy(nx,ny,nz), efldz(nx,ny,nz))
    read(10) bx, by, bz, bx0, by0, bz0, oldbx, oldby, oldbz
    read(10) qpx, qpy, qpz, qrho, hpx, hpy, hpz, hrho, opx, opy, opz, orho
    read(10) epres, wrkepres, oldepres
    close(10)
    rst = 0.
    ijmid = 1
    ijzero = 1
    mbndry = 1
    nummid = 0
    numzero = 0
    mmid = 1
    mzero = 1

    ! start-up: device mirrors of every array the calls touch
    ierr = g3m3_init()
    if (ierr /= 0) stop 'g3m3_init failed'
    n4 = 4_8 * nx * ny * nz * n_grids
    n3 = 4_8 * nx * ny * nz
    call reg4(bx); call reg4(by); call reg4(bz)
    call reg4(bx0); call reg4(by0); call reg4(bz0)
    call reg4(oldbx); call reg4(oldby); call reg4(oldbz)
    call reg4(qpx); call reg4(qpy); call reg4(qpz); call reg4(qrho)
    call reg4(hpx); call reg4(hpy); call reg4(hpz); call reg4(hrho)
    call reg4(opx); call reg4(opy); call reg4(opz); call reg4(orho)
    call reg4(epres); call reg4(wrkepres); call reg4(oldepres)
    ierr = g3m3_register(rst, 4_8 * nx * ny * nz)
    ierr = g3m3_to_device(rst)
    call reg3(curx); call reg3(cury); call reg3(curz)
    call reg3(bsx); call reg3(bsy); call reg3(bsz); call reg3(btot)
    call reg3(evx); call reg3(evy); call reg3(evz)
    call reg3(tvx); call reg3(tvy); call reg3(tvz)
    call reg3(efldx); call reg3(efldy); call reg3(efldz)

    ! the step-2 field sequence of a box-step, reference argument lists
    call g3m3_calcur(bx, by, bz, nx, ny, nz, n_grids, box, curx, cury, curz, rx, rx, rx)
    call g3m3_bfield_total(bx, by, bz, bx0, by0, bz0, bsx, bsy, bsz, btot, &
        nx, ny, nz, n_grids, box, smallbit)
    call g3m3_fnd_evel(qpx, qpy, qpz, qrho, hpx, hpy, hpz, hrho, opx, opy, opz, orho, &
        curx, cury, curz, evx, evy, evz, tvx, tvy, tvz, &
        nx, ny, nz, n_grids, box, rmassq, rmassh, rmasso, reynolds)
    call g3m3_bande(efldx, efldy, efldz, bsx, bsy, bsz, curx, cury, curz, evx, evy, evz, &
        btot, epres, qrho, hrho, orho, rst, resist, reynolds, &
        nx, ny, nz, n_grids, box, rmassq, rmassh, rmasso, &
        ijmid, nummid, ijzero, mbndry, numzero, mmid, mzero, rx, rx, rx)
    call g3m3_push_elec(wrkepres, oldepres, epres, evx, evy, evz, gamma, gamma1, &
        nx, ny, nz, n_grids, box, delt, rx, rx, rx)
    call g3m3_push_bfld(bx, by, bz, oldbx, oldby, oldbz, efldx, efldy, efldz, &
        nx, ny, nz, n_grids, box, delt, rx, rx, rx)
    ierr = g3m3_last_error()
    if (ierr /= 0) stop 'g3m3 call failed'

    ! output point: bring results back to the host arrays
    ierr = g3m3_to_host(curx) + g3m3_to_host(cury) + g3m3_to_host(curz)
    ierr = ierr + g3m3_to_host(bsx) + g3m3_to_host(bsy) + g3m3_to_host(bsz) + g3m3_to_host(btot)
    ierr = ierr + g3m3_to_host(evx) + g3m3_to_host(evy) + g3m3_to_host(evz)
    ierr = ierr + g3m3_to_host(tvx) + g3m3_to_host(tvy) + g3m3_to_host(tvz)
    ierr = ierr + g3m3_to_host(efldx) + g3m3_to_host(efldy) + g3m3_to_host(efldz)
    ierr = ierr + g3m3_to_host(epres) + g3m3_to_host(bx) + g3m3_to_host(by) + g3m3_to_host(bz)
    if (ierr /= 0) stop 'g3m3_to_host failed'
    open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(11) curx, cury, curz, bsx, bsy, bsz, btot, evx, evy, evz, tvx, tvy, tvz
    write(11) efldx, efldy, efldz, epres, bx, by, bz
    close(11)

contains

    subroutine reg4(a)
        real a(nx,ny,nz,n_grids)
        if (g3m3_register(a, n4) /= 0) stop 'g3m3_register failed'
        if (g3m3_to_device(a) /= 0) stop 'g3m3_to_device failed'
    end subroutine

    subroutine reg3(a)
        real a(nx,ny,nz)
        if (g3m3_register(a, n3) /= 0) stop 'g3m3_register failed'
    end subroutine

end program
