! Generated by g3m3_amd/fortran/make_iface.py from include/g3m3.h; do not edit.
! bind(c) interfaces of the g3m3 C-ABI's subroutine entry points.  Each
! dummy is an array or a scalar as in the reference subroutine of the same
! name, so `use g3m3_iface` then `call g3m3_<name>(<the reference's
! arguments>)` is checked by the compiler.  The reference's LOGICAL dummies
! are logical(c_int) (4 bytes, as the library reads them: nonzero = .true.).
! Arrays are host arrays registered with g3m3_register or device arrays.
module g3m3_iface
    use, intrinsic :: iso_c_binding
    implicit none
    interface
        integer(c_int) function g3m3_init() bind(c, name='g3m3_init')
            import :: c_int
        end function
        integer(c_int) function g3m3_last_error() bind(c, name='g3m3_last_error')
            import :: c_int
        end function
        subroutine g3m3_sync() bind(c, name='g3m3_sync')
        end subroutine
        integer(c_int) function g3m3_register(host, bytes) bind(c, name='g3m3_register')
            import :: c_int, c_long_long
            type(*), dimension(*) :: host
            integer(c_long_long), value :: bytes
        end function
        integer(c_int) function g3m3_unregister(host) bind(c, name='g3m3_unregister')
            import :: c_int
            type(*), dimension(*) :: host
        end function
        integer(c_int) function g3m3_to_device(host) bind(c, name='g3m3_to_device')
            import :: c_int
            type(*), dimension(*) :: host
        end function
        integer(c_int) function g3m3_to_host(host) bind(c, name='g3m3_to_host')
            import :: c_int
            type(*), dimension(*) :: host
        end function
        subroutine g3m3_store_array(bx, bx0, nx, ny, nz, n_grids, box) bind(c, name='g3m3_store_array')
            import :: c_float, c_int
            real(c_float), intent(inout) :: bx(*)
            real(c_float), intent(in) :: bx0(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
        end subroutine
        subroutine g3m3_calcur(bx, by, bz, nx, ny, nz, n_grids, box, curx, cury, curz, &
            rx, ry, rz) bind(c, name='g3m3_calcur')
            import :: c_float, c_int
            real(c_float), intent(in) :: bx(*)
            real(c_float), intent(in) :: by(*)
            real(c_float), intent(in) :: bz(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(inout) :: curx(*)
            real(c_float), intent(inout) :: cury(*)
            real(c_float), intent(inout) :: curz(*)
            real(c_float), intent(in) :: rx
            real(c_float), intent(in) :: ry
            real(c_float), intent(in) :: rz
        end subroutine
        subroutine g3m3_fnd_evel(qpx, qpy, qpz, qrho, hpx, hpy, hpz, hrho, opx, opy, &
            opz, orho, curx, cury, curz, evx, evy, evz, tvx, tvy, tvz, nx, ny, nz, &
            n_grids, box, rmassq, rmassh, rmasso, reynolds) bind(c, name='g3m3_fnd_evel')
            import :: c_float, c_int
            real(c_float), intent(in) :: qpx(*)
            real(c_float), intent(in) :: qpy(*)
            real(c_float), intent(in) :: qpz(*)
            real(c_float), intent(in) :: qrho(*)
            real(c_float), intent(in) :: hpx(*)
            real(c_float), intent(in) :: hpy(*)
            real(c_float), intent(in) :: hpz(*)
            real(c_float), intent(in) :: hrho(*)
            real(c_float), intent(in) :: opx(*)
            real(c_float), intent(in) :: opy(*)
            real(c_float), intent(in) :: opz(*)
            real(c_float), intent(in) :: orho(*)
            real(c_float), intent(in) :: curx(*)
            real(c_float), intent(in) :: cury(*)
            real(c_float), intent(in) :: curz(*)
            real(c_float), intent(inout) :: evx(*)
            real(c_float), intent(inout) :: evy(*)
            real(c_float), intent(inout) :: evz(*)
            real(c_float), intent(inout) :: tvx(*)
            real(c_float), intent(inout) :: tvy(*)
            real(c_float), intent(inout) :: tvz(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: rmassq
            real(c_float), intent(in) :: rmassh
            real(c_float), intent(in) :: rmasso
            real(c_float), intent(in) :: reynolds
        end subroutine
        subroutine g3m3_bande(efldx, efldy, efldz, bsx, bsy, bsz, curx, cury, curz, &
            evx, evy, evz, btot, epres, qrho, hrho, orho, rst, resist, reynolds, nx, &
            ny, nz, n_grids, box, rmassq, rmassh, rmasso, ijmid, nummid, ijzero, &
            mbndry, numzero, mmid, mzero, rx, ry, rz) bind(c, name='g3m3_bande')
            import :: c_float, c_int
            real(c_float), intent(inout) :: efldx(*)
            real(c_float), intent(inout) :: efldy(*)
            real(c_float), intent(inout) :: efldz(*)
            real(c_float), intent(in) :: bsx(*)
            real(c_float), intent(in) :: bsy(*)
            real(c_float), intent(in) :: bsz(*)
            real(c_float), intent(in) :: curx(*)
            real(c_float), intent(in) :: cury(*)
            real(c_float), intent(in) :: curz(*)
            real(c_float), intent(in) :: evx(*)
            real(c_float), intent(in) :: evy(*)
            real(c_float), intent(in) :: evz(*)
            real(c_float), intent(in) :: btot(*)
            real(c_float), intent(in) :: epres(*)
            real(c_float), intent(in) :: qrho(*)
            real(c_float), intent(in) :: hrho(*)
            real(c_float), intent(in) :: orho(*)
            real(c_float), intent(in) :: rst(*)
            real(c_float), intent(in) :: resist
            real(c_float), intent(in) :: reynolds
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: rmassq
            real(c_float), intent(in) :: rmassh
            real(c_float), intent(in) :: rmasso
            integer(c_int), intent(in) :: ijmid(*)
            integer(c_int), intent(in) :: nummid(*)
            integer(c_int), intent(in) :: ijzero(*)
            integer(c_int), intent(in) :: mbndry
            integer(c_int), intent(in) :: numzero(*)
            integer(c_int), intent(in) :: mmid
            integer(c_int), intent(in) :: mzero
            real(c_float), intent(in) :: rx
            real(c_float), intent(in) :: ry
            real(c_float), intent(in) :: rz
        end subroutine
        subroutine g3m3_bande_eta(efldx, efldy, efldz, bsx, bsy, bsz, curx, cury, curz, &
            evx, evy, evz, btot, epres, qrho, hrho, orho, rst, resist, reynolds, nx, &
            ny, nz, n_grids, box, rmassq, rmassh, rmasso, ijmid, nummid, ijzero, &
            mbndry, numzero, mmid, mzero, rx, ry, rz, eta) bind(c, name='g3m3_bande_eta')
            import :: c_float, c_int
            real(c_float), intent(inout) :: efldx(*)
            real(c_float), intent(inout) :: efldy(*)
            real(c_float), intent(inout) :: efldz(*)
            real(c_float), intent(in) :: bsx
            real(c_float), intent(in) :: bsy
            real(c_float), intent(in) :: bsz
            real(c_float), intent(in) :: curx(*)
            real(c_float), intent(in) :: cury(*)
            real(c_float), intent(in) :: curz(*)
            real(c_float), intent(in) :: evx
            real(c_float), intent(in) :: evy
            real(c_float), intent(in) :: evz
            real(c_float), intent(in) :: btot
            real(c_float), intent(in) :: epres
            real(c_float), intent(in) :: qrho(*)
            real(c_float), intent(in) :: hrho(*)
            real(c_float), intent(in) :: orho(*)
            real(c_float), intent(in) :: rst
            real(c_float), intent(in) :: resist
            real(c_float), intent(in) :: reynolds
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: rmassq
            real(c_float), intent(in) :: rmassh
            real(c_float), intent(in) :: rmasso
            integer(c_int), intent(in) :: ijmid
            integer(c_int), intent(in) :: nummid
            integer(c_int), intent(in) :: ijzero
            integer(c_int), intent(in) :: mbndry
            integer(c_int), intent(in) :: numzero
            integer(c_int), intent(in) :: mmid
            integer(c_int), intent(in) :: mzero
            real(c_float), intent(in) :: rx
            real(c_float), intent(in) :: ry
            real(c_float), intent(in) :: rz
            real(c_float), intent(in) :: eta
        end subroutine
        subroutine g3m3_bfield_total(bx, by, bz, bx0, by0, bz0, bsx, bsy, bsz, btot, &
            nx, ny, nz, n_grids, box, smallbit) bind(c, name='g3m3_bfield_total')
            import :: c_float, c_int
            real(c_float), intent(in) :: bx(*)
            real(c_float), intent(in) :: by(*)
            real(c_float), intent(in) :: bz(*)
            real(c_float), intent(in) :: bx0(*)
            real(c_float), intent(in) :: by0(*)
            real(c_float), intent(in) :: bz0(*)
            real(c_float), intent(inout) :: bsx(*)
            real(c_float), intent(inout) :: bsy(*)
            real(c_float), intent(inout) :: bsz(*)
            real(c_float), intent(inout) :: btot(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: smallbit
        end subroutine
        subroutine g3m3_push_elec(wrkepres, oldepres, epres, evx, evy, evz, gamma, &
            gamma1, nx, ny, nz, n_grids, box, delt, rx, ry, rz) bind(c, name='g3m3_push_elec')
            import :: c_float, c_int
            real(c_float), intent(in) :: wrkepres(*)
            real(c_float), intent(in) :: oldepres(*)
            real(c_float), intent(inout) :: epres(*)
            real(c_float), intent(in) :: evx(*)
            real(c_float), intent(in) :: evy(*)
            real(c_float), intent(in) :: evz(*)
            real(c_float), intent(in) :: gamma
            real(c_float), intent(in) :: gamma1
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: delt
            real(c_float), intent(in) :: rx
            real(c_float), intent(in) :: ry
            real(c_float), intent(in) :: rz
        end subroutine
        subroutine g3m3_push_ion(qrho, qpresx, qpresy, qpresz, qpx, qpy, qpz, oldqrho, &
            oldqpresx, oldqpresy, oldqpresz, oldqpx, oldqpy, oldqpz, wrkqrho, &
            wrkqpresx, wrkqpresy, wrkqpresz, wrkqpx, wrkqpy, wrkqpz, qpresxy, qpresxz, &
            qpresyz, oldqpresxy, oldqpresxz, oldqpresyz, wrkqpresxy, wrkqpresxz, &
            wrkqpresyz, bsx, bsy, bsz, btot, efldx, efldy, efldz, rmassq, qvx, qvy, &
            qvz, tvx, tvy, tvz, gamma, gamma1, nx, ny, nz, n_grids, box, delt, grav, &
            re_equiv, reynolds, grd_xmin, grd_xmax, grd_ymin, grd_ymax, grd_zmin, &
            grd_zmax, isotropic) bind(c, name='g3m3_push_ion')
            import :: c_float, c_int
            real(c_float), intent(inout) :: qrho(*)
            real(c_float), intent(inout) :: qpresx(*)
            real(c_float), intent(inout) :: qpresy(*)
            real(c_float), intent(inout) :: qpresz(*)
            real(c_float), intent(inout) :: qpx(*)
            real(c_float), intent(inout) :: qpy(*)
            real(c_float), intent(inout) :: qpz(*)
            real(c_float), intent(in) :: oldqrho(*)
            real(c_float), intent(in) :: oldqpresx(*)
            real(c_float), intent(in) :: oldqpresy(*)
            real(c_float), intent(in) :: oldqpresz(*)
            real(c_float), intent(in) :: oldqpx(*)
            real(c_float), intent(in) :: oldqpy(*)
            real(c_float), intent(in) :: oldqpz(*)
            real(c_float), intent(in) :: wrkqrho(*)
            real(c_float), intent(in) :: wrkqpresx(*)
            real(c_float), intent(in) :: wrkqpresy(*)
            real(c_float), intent(in) :: wrkqpresz(*)
            real(c_float), intent(in) :: wrkqpx(*)
            real(c_float), intent(in) :: wrkqpy(*)
            real(c_float), intent(in) :: wrkqpz(*)
            real(c_float), intent(inout) :: qpresxy(*)
            real(c_float), intent(inout) :: qpresxz(*)
            real(c_float), intent(inout) :: qpresyz(*)
            real(c_float), intent(in) :: oldqpresxy(*)
            real(c_float), intent(in) :: oldqpresxz(*)
            real(c_float), intent(in) :: oldqpresyz(*)
            real(c_float), intent(in) :: wrkqpresxy(*)
            real(c_float), intent(in) :: wrkqpresxz(*)
            real(c_float), intent(in) :: wrkqpresyz(*)
            real(c_float), intent(in) :: bsx(*)
            real(c_float), intent(in) :: bsy(*)
            real(c_float), intent(in) :: bsz(*)
            real(c_float), intent(in) :: btot(*)
            real(c_float), intent(in) :: efldx(*)
            real(c_float), intent(in) :: efldy(*)
            real(c_float), intent(in) :: efldz(*)
            real(c_float), intent(in) :: rmassq
            real(c_float), intent(inout) :: qvx(*)
            real(c_float), intent(inout) :: qvy(*)
            real(c_float), intent(inout) :: qvz(*)
            real(c_float), intent(in) :: tvx(*)
            real(c_float), intent(in) :: tvy(*)
            real(c_float), intent(in) :: tvz(*)
            real(c_float), intent(in) :: gamma
            real(c_float), intent(in) :: gamma1
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: delt
            real(c_float), intent(in) :: grav
            real(c_float), intent(in) :: re_equiv
            real(c_float), intent(in) :: reynolds
            real(c_float), intent(in) :: grd_xmin(*)
            real(c_float), intent(in) :: grd_xmax(*)
            real(c_float), intent(in) :: grd_ymin(*)
            real(c_float), intent(in) :: grd_ymax(*)
            real(c_float), intent(in) :: grd_zmin(*)
            real(c_float), intent(in) :: grd_zmax(*)
            logical(c_int), intent(in) :: isotropic
        end subroutine
        subroutine g3m3_push_bfld(bx, by, bz, oldbx, oldby, oldbz, efldx, efldy, efldz, &
            nx, ny, nz, n_grids, box, delt, rx, ry, rz) bind(c, name='g3m3_push_bfld')
            import :: c_float, c_int
            real(c_float), intent(inout) :: bx(*)
            real(c_float), intent(inout) :: by(*)
            real(c_float), intent(inout) :: bz(*)
            real(c_float), intent(in) :: oldbx(*)
            real(c_float), intent(in) :: oldby(*)
            real(c_float), intent(in) :: oldbz(*)
            real(c_float), intent(in) :: efldx(*)
            real(c_float), intent(in) :: efldy(*)
            real(c_float), intent(in) :: efldz(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: delt
            real(c_float), intent(in) :: rx
            real(c_float), intent(in) :: ry
            real(c_float), intent(in) :: rz
        end subroutine
        subroutine g3m3_set_rho(qrho, qpresx, qpresy, qpresz, qpresxy, qpresxz, &
            qpresyz, rmassq, hrho, hpresx, hpresy, hpresz, hpresxy, hpresxz, hpresyz, &
            rmassh, orho, opresx, opresy, opresz, opresxy, opresxz, opresyz, rmasso, &
            epres, nx, ny, nz, n_grids, box, o_conc) bind(c, name='g3m3_set_rho')
            import :: c_float, c_int
            real(c_float), intent(inout) :: qrho(*)
            real(c_float), intent(inout) :: qpresx(*)
            real(c_float), intent(inout) :: qpresy(*)
            real(c_float), intent(inout) :: qpresz(*)
            real(c_float), intent(inout) :: qpresxy(*)
            real(c_float), intent(inout) :: qpresxz(*)
            real(c_float), intent(inout) :: qpresyz(*)
            real(c_float), intent(in) :: rmassq
            real(c_float), intent(inout) :: hrho(*)
            real(c_float), intent(inout) :: hpresx(*)
            real(c_float), intent(inout) :: hpresy(*)
            real(c_float), intent(inout) :: hpresz(*)
            real(c_float), intent(inout) :: hpresxy(*)
            real(c_float), intent(inout) :: hpresxz(*)
            real(c_float), intent(inout) :: hpresyz(*)
            real(c_float), intent(in) :: rmassh
            real(c_float), intent(inout) :: orho(*)
            real(c_float), intent(inout) :: opresx(*)
            real(c_float), intent(inout) :: opresy(*)
            real(c_float), intent(inout) :: opresz(*)
            real(c_float), intent(inout) :: opresxy(*)
            real(c_float), intent(inout) :: opresxz(*)
            real(c_float), intent(inout) :: opresyz(*)
            real(c_float), intent(in) :: rmasso
            real(c_float), intent(inout) :: epres(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: o_conc
        end subroutine
        subroutine g3m3_set_speed_agrd(nx, ny, nz, n_grids, box, rmassq, rmassh, &
            rmasso, isotropic, smallbit, aniso_limit, vlim, qvlim, hvlim, ovlim, qclim, &
            hclim, oclim, cslim, bx, by, bz, bx0, by0, bz0, qrho, qpx, qpy, qpz, &
            qpresx, qpresy, qpresz, qpresxy, qpresxz, qpresyz, hrho, hpx, hpy, hpz, &
            hpresx, hpresy, hpresz, hpresxy, hpresxz, hpresyz, orho, opx, opy, opz, &
            opresx, opresy, opresz, opresxy, opresxz, opresyz, epres, btot, pxmax, &
            pymax, pzmax, csmax, alfmax, fastest) bind(c, name='g3m3_set_speed_agrd')
            import :: c_float, c_int
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: rmassq
            real(c_float), intent(in) :: rmassh
            real(c_float), intent(in) :: rmasso
            logical(c_int), intent(in) :: isotropic
            real(c_float), intent(in) :: smallbit
            real(c_float), intent(in) :: aniso_limit
            real(c_float), intent(in) :: vlim
            real(c_float), intent(in) :: qvlim
            real(c_float), intent(in) :: hvlim
            real(c_float), intent(in) :: ovlim
            real(c_float), intent(in) :: qclim
            real(c_float), intent(in) :: hclim
            real(c_float), intent(in) :: oclim
            real(c_float), intent(in) :: cslim
            real(c_float), intent(in) :: bx(*)
            real(c_float), intent(in) :: by(*)
            real(c_float), intent(in) :: bz(*)
            real(c_float), intent(in) :: bx0(*)
            real(c_float), intent(in) :: by0(*)
            real(c_float), intent(in) :: bz0(*)
            real(c_float), intent(inout) :: qrho(*)
            real(c_float), intent(inout) :: qpx(*)
            real(c_float), intent(inout) :: qpy(*)
            real(c_float), intent(inout) :: qpz(*)
            real(c_float), intent(inout) :: qpresx(*)
            real(c_float), intent(inout) :: qpresy(*)
            real(c_float), intent(inout) :: qpresz(*)
            real(c_float), intent(inout) :: qpresxy(*)
            real(c_float), intent(inout) :: qpresxz(*)
            real(c_float), intent(inout) :: qpresyz(*)
            real(c_float), intent(inout) :: hrho(*)
            real(c_float), intent(inout) :: hpx(*)
            real(c_float), intent(inout) :: hpy(*)
            real(c_float), intent(inout) :: hpz(*)
            real(c_float), intent(inout) :: hpresx(*)
            real(c_float), intent(inout) :: hpresy(*)
            real(c_float), intent(inout) :: hpresz(*)
            real(c_float), intent(inout) :: hpresxy(*)
            real(c_float), intent(inout) :: hpresxz(*)
            real(c_float), intent(inout) :: hpresyz(*)
            real(c_float), intent(inout) :: orho(*)
            real(c_float), intent(inout) :: opx(*)
            real(c_float), intent(inout) :: opy(*)
            real(c_float), intent(inout) :: opz(*)
            real(c_float), intent(inout) :: opresx(*)
            real(c_float), intent(inout) :: opresy(*)
            real(c_float), intent(inout) :: opresz(*)
            real(c_float), intent(inout) :: opresxy(*)
            real(c_float), intent(inout) :: opresxz(*)
            real(c_float), intent(inout) :: opresyz(*)
            real(c_float), intent(inout) :: epres(*)
            real(c_float), intent(inout) :: btot(*)
            real(c_float), intent(inout) :: pxmax
            real(c_float), intent(inout) :: pymax
            real(c_float), intent(inout) :: pzmax
            real(c_float), intent(inout) :: csmax
            real(c_float), intent(inout) :: alfmax
            real(c_float), intent(inout) :: fastest
        end subroutine
        subroutine g3m3_set_speed_agrd_noalloc(nx, ny, nz, n_grids, box, rmassq, &
            rmassh, rmasso, isotropic, smallbit, aniso_limit, vlim, qvlim, hvlim, &
            ovlim, qclim, hclim, oclim, cslim, bx, by, bz, bx0, by0, bz0, qrho, qpx, &
            qpy, qpz, qpresx, qpresy, qpresz, qpresxy, qpresxz, qpresyz, hrho, hpx, &
            hpy, hpz, hpresx, hpresy, hpresz, hpresxy, hpresxz, hpresyz, orho, opx, &
            opy, opz, opresx, opresy, opresz, opresxy, opresxz, opresyz, epres, btot, &
            pxmax, pymax, pzmax, csmax, alfmax, fastest) bind(c, name='g3m3_set_speed_agrd_noalloc')
            import :: c_float, c_int
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: rmassq
            real(c_float), intent(in) :: rmassh
            real(c_float), intent(in) :: rmasso
            logical(c_int), intent(in) :: isotropic
            real(c_float), intent(in) :: smallbit
            real(c_float), intent(in) :: aniso_limit
            real(c_float), intent(in) :: vlim
            real(c_float), intent(in) :: qvlim
            real(c_float), intent(in) :: hvlim
            real(c_float), intent(in) :: ovlim
            real(c_float), intent(in) :: qclim
            real(c_float), intent(in) :: hclim
            real(c_float), intent(in) :: oclim
            real(c_float), intent(in) :: cslim
            real(c_float), intent(in) :: bx(*)
            real(c_float), intent(in) :: by(*)
            real(c_float), intent(in) :: bz(*)
            real(c_float), intent(in) :: bx0(*)
            real(c_float), intent(in) :: by0(*)
            real(c_float), intent(in) :: bz0(*)
            real(c_float), intent(inout) :: qrho(*)
            real(c_float), intent(inout) :: qpx(*)
            real(c_float), intent(inout) :: qpy(*)
            real(c_float), intent(inout) :: qpz(*)
            real(c_float), intent(inout) :: qpresx(*)
            real(c_float), intent(inout) :: qpresy(*)
            real(c_float), intent(inout) :: qpresz(*)
            real(c_float), intent(inout) :: qpresxy(*)
            real(c_float), intent(inout) :: qpresxz(*)
            real(c_float), intent(inout) :: qpresyz(*)
            real(c_float), intent(inout) :: hrho(*)
            real(c_float), intent(inout) :: hpx(*)
            real(c_float), intent(inout) :: hpy(*)
            real(c_float), intent(inout) :: hpz(*)
            real(c_float), intent(inout) :: hpresx(*)
            real(c_float), intent(inout) :: hpresy(*)
            real(c_float), intent(inout) :: hpresz(*)
            real(c_float), intent(inout) :: hpresxy(*)
            real(c_float), intent(inout) :: hpresxz(*)
            real(c_float), intent(inout) :: hpresyz(*)
            real(c_float), intent(inout) :: orho(*)
            real(c_float), intent(inout) :: opx(*)
            real(c_float), intent(inout) :: opy(*)
            real(c_float), intent(inout) :: opz(*)
            real(c_float), intent(inout) :: opresx(*)
            real(c_float), intent(inout) :: opresy(*)
            real(c_float), intent(inout) :: opresz(*)
            real(c_float), intent(inout) :: opresxy(*)
            real(c_float), intent(inout) :: opresxz(*)
            real(c_float), intent(inout) :: opresyz(*)
            real(c_float), intent(inout) :: epres(*)
            real(c_float), intent(inout) :: btot(*)
            real(c_float), intent(inout) :: pxmax
            real(c_float), intent(inout) :: pymax
            real(c_float), intent(inout) :: pzmax
            real(c_float), intent(inout) :: csmax
            real(c_float), intent(inout) :: alfmax
            real(c_float), intent(inout) :: fastest
        end subroutine
        subroutine g3m3_lap_plasma(rho, px, py, pz, presx, presy, presz, presxy, &
            presxz, presyz, wrkrho, wrkpx, wrkpy, wrkpz, wrkpresx, wrkpresy, wrkpresz, &
            wrkpresxy, wrkpresxz, wrkpresyz, vx, vy, vz, nx, ny, nz, n_grids, box, &
            chirho, chipxyz, chierg, delt, isotropic) bind(c, name='g3m3_lap_plasma')
            import :: c_float, c_int
            real(c_float), intent(in) :: rho(*)
            real(c_float), intent(in) :: px(*)
            real(c_float), intent(in) :: py(*)
            real(c_float), intent(in) :: pz(*)
            real(c_float), intent(in) :: presx(*)
            real(c_float), intent(in) :: presy(*)
            real(c_float), intent(in) :: presz(*)
            real(c_float), intent(in) :: presxy(*)
            real(c_float), intent(in) :: presxz(*)
            real(c_float), intent(in) :: presyz(*)
            real(c_float), intent(inout) :: wrkrho(*)
            real(c_float), intent(inout) :: wrkpx(*)
            real(c_float), intent(inout) :: wrkpy(*)
            real(c_float), intent(inout) :: wrkpz(*)
            real(c_float), intent(inout) :: wrkpresx(*)
            real(c_float), intent(inout) :: wrkpresy(*)
            real(c_float), intent(inout) :: wrkpresz(*)
            real(c_float), intent(inout) :: wrkpresxy(*)
            real(c_float), intent(inout) :: wrkpresxz(*)
            real(c_float), intent(inout) :: wrkpresyz(*)
            real(c_float), intent(in) :: vx(*)
            real(c_float), intent(in) :: vy(*)
            real(c_float), intent(in) :: vz(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: chirho
            real(c_float), intent(in) :: chipxyz
            real(c_float), intent(in) :: chierg
            real(c_float), intent(in) :: delt
            logical(c_int), intent(in) :: isotropic
        end subroutine
        subroutine g3m3_lap_elec(ppres, wrkppres, vx, vy, vz, nx, ny, nz, n_grids, box, &
            chirho, chipxyz, chierg, delt) bind(c, name='g3m3_lap_elec')
            import :: c_float, c_int
            real(c_float), intent(in) :: ppres(*)
            real(c_float), intent(inout) :: wrkppres(*)
            real(c_float), intent(in) :: vx(*)
            real(c_float), intent(in) :: vy(*)
            real(c_float), intent(in) :: vz(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: chirho
            real(c_float), intent(in) :: chipxyz
            real(c_float), intent(in) :: chierg
            real(c_float), intent(in) :: delt
        end subroutine
        subroutine g3m3_lap_bfld(bx, by, bz, wrkbx, wrkby, wrkbz, vx, vy, vz, nx, ny, &
            nz, n_grids, box, chirho, chipxyz, chierg, delt) bind(c, name='g3m3_lap_bfld')
            import :: c_float, c_int
            real(c_float), intent(in) :: bx(*)
            real(c_float), intent(in) :: by(*)
            real(c_float), intent(in) :: bz(*)
            real(c_float), intent(inout) :: wrkbx(*)
            real(c_float), intent(inout) :: wrkby(*)
            real(c_float), intent(inout) :: wrkbz(*)
            real(c_float), intent(in) :: vx(*)
            real(c_float), intent(in) :: vy(*)
            real(c_float), intent(in) :: vz(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: chirho
            real(c_float), intent(in) :: chipxyz
            real(c_float), intent(in) :: chierg
            real(c_float), intent(in) :: delt
        end subroutine
        subroutine g3m3_fnd_vtot(qpx, qpy, qpz, qrho, hpx, hpy, hpz, hrho, opx, opy, &
            opz, orho, vx, vy, vz, nx, ny, nz, n_grids, box, rmassq, rmassh, rmasso) bind(c, name='g3m3_fnd_vtot')
            import :: c_float, c_int
            real(c_float), intent(in) :: qpx(*)
            real(c_float), intent(in) :: qpy(*)
            real(c_float), intent(in) :: qpz(*)
            real(c_float), intent(in) :: qrho(*)
            real(c_float), intent(in) :: hpx(*)
            real(c_float), intent(in) :: hpy(*)
            real(c_float), intent(in) :: hpz(*)
            real(c_float), intent(in) :: hrho(*)
            real(c_float), intent(in) :: opx(*)
            real(c_float), intent(in) :: opy(*)
            real(c_float), intent(in) :: opz(*)
            real(c_float), intent(in) :: orho(*)
            real(c_float), intent(inout) :: vx(*)
            real(c_float), intent(inout) :: vy(*)
            real(c_float), intent(inout) :: vz(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: rmassq
            real(c_float), intent(in) :: rmassh
            real(c_float), intent(in) :: rmasso
        end subroutine
        subroutine g3m3_species_velocity(rho, px, py, pz, vx, vy, vz, nx, ny, nz, &
            n_grids, box, smallbit) bind(c, name='g3m3_species_velocity')
            import :: c_float, c_int
            real(c_float), intent(in) :: rho(*)
            real(c_float), intent(in) :: px(*)
            real(c_float), intent(in) :: py(*)
            real(c_float), intent(in) :: pz(*)
            real(c_float), intent(inout) :: vx(*)
            real(c_float), intent(inout) :: vy(*)
            real(c_float), intent(inout) :: vz(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: smallbit
        end subroutine
        subroutine g3m3_fcsmooth(px, oldpx, wrkpx, nx, ny, nz, n_grids, box, chipx, &
            tempx, tempy, tempz) bind(c, name='g3m3_fcsmooth')
            import :: c_float, c_int
            real(c_float), intent(inout) :: px(*)
            real(c_float), intent(in) :: oldpx(*)
            real(c_float), intent(in) :: wrkpx(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: chipx
            real(c_float), intent(inout) :: tempx(*)
            real(c_float), intent(inout) :: tempy(*)
            real(c_float), intent(inout) :: tempz(*)
        end subroutine
        subroutine g3m3_flux_correct(qrho, qpresx, qpresy, qpresz, qpresxy, qpresxz, &
            qpresyz, qpx, qpy, qpz, wrkqrho, wrkqpresx, wrkqpresy, wrkqpresz, &
            wrkqpresxy, wrkqpresxz, wrkqpresyz, wrkqpx, wrkqpy, wrkqpz, oldqrho, &
            oldqpresx, oldqpresy, oldqpresz, oldqpresxy, oldqpresxz, oldqpresyz, &
            oldqpx, oldqpy, oldqpz, hrho, hpresx, hpresy, hpresz, hpresxy, hpresxz, &
            hpresyz, hpx, hpy, hpz, wrkhrho, wrkhpresx, wrkhpresy, wrkhpresz, &
            wrkhpresxy, wrkhpresxz, wrkhpresyz, wrkhpx, wrkhpy, wrkhpz, oldhrho, &
            oldhpresx, oldhpresy, oldhpresz, oldhpresxy, oldhpresxz, oldhpresyz, &
            oldhpx, oldhpy, oldhpz, orho, opresx, opresy, opresz, opresxy, opresxz, &
            opresyz, opx, opy, opz, wrkorho, wrkopresx, wrkopresy, wrkopresz, &
            wrkopresxy, wrkopresxz, wrkopresyz, wrkopx, wrkopy, wrkopz, oldorho, &
            oldopresx, oldopresy, oldopresz, oldopresxy, oldopresxz, oldopresyz, &
            oldopx, oldopy, oldopz, epres, wrkepres, oldepres, bx, by, bz, wrkbx, &
            wrkby, wrkbz, oldbx, oldby, oldbz, vvx, vvy, vvz, nx, ny, nz, n_grids, box, &
            difrho, diferg, xspac, isotropic) bind(c, name='g3m3_flux_correct')
            import :: c_float, c_int
            real(c_float), intent(inout) :: qrho(*)
            real(c_float), intent(inout) :: qpresx(*)
            real(c_float), intent(inout) :: qpresy(*)
            real(c_float), intent(inout) :: qpresz(*)
            real(c_float), intent(inout) :: qpresxy(*)
            real(c_float), intent(inout) :: qpresxz(*)
            real(c_float), intent(inout) :: qpresyz(*)
            real(c_float), intent(inout) :: qpx(*)
            real(c_float), intent(inout) :: qpy(*)
            real(c_float), intent(inout) :: qpz(*)
            real(c_float), intent(in) :: wrkqrho(*)
            real(c_float), intent(in) :: wrkqpresx(*)
            real(c_float), intent(in) :: wrkqpresy(*)
            real(c_float), intent(in) :: wrkqpresz(*)
            real(c_float), intent(in) :: wrkqpresxy(*)
            real(c_float), intent(in) :: wrkqpresxz(*)
            real(c_float), intent(in) :: wrkqpresyz(*)
            real(c_float), intent(in) :: wrkqpx(*)
            real(c_float), intent(in) :: wrkqpy(*)
            real(c_float), intent(in) :: wrkqpz(*)
            real(c_float), intent(in) :: oldqrho(*)
            real(c_float), intent(in) :: oldqpresx(*)
            real(c_float), intent(in) :: oldqpresy(*)
            real(c_float), intent(in) :: oldqpresz(*)
            real(c_float), intent(in) :: oldqpresxy(*)
            real(c_float), intent(in) :: oldqpresxz(*)
            real(c_float), intent(in) :: oldqpresyz(*)
            real(c_float), intent(in) :: oldqpx(*)
            real(c_float), intent(in) :: oldqpy(*)
            real(c_float), intent(in) :: oldqpz(*)
            real(c_float), intent(inout) :: hrho(*)
            real(c_float), intent(inout) :: hpresx(*)
            real(c_float), intent(inout) :: hpresy(*)
            real(c_float), intent(inout) :: hpresz(*)
            real(c_float), intent(inout) :: hpresxy(*)
            real(c_float), intent(inout) :: hpresxz(*)
            real(c_float), intent(inout) :: hpresyz(*)
            real(c_float), intent(inout) :: hpx(*)
            real(c_float), intent(inout) :: hpy(*)
            real(c_float), intent(inout) :: hpz(*)
            real(c_float), intent(in) :: wrkhrho(*)
            real(c_float), intent(in) :: wrkhpresx(*)
            real(c_float), intent(in) :: wrkhpresy(*)
            real(c_float), intent(in) :: wrkhpresz(*)
            real(c_float), intent(in) :: wrkhpresxy(*)
            real(c_float), intent(in) :: wrkhpresxz(*)
            real(c_float), intent(in) :: wrkhpresyz(*)
            real(c_float), intent(in) :: wrkhpx(*)
            real(c_float), intent(in) :: wrkhpy(*)
            real(c_float), intent(in) :: wrkhpz(*)
            real(c_float), intent(in) :: oldhrho(*)
            real(c_float), intent(in) :: oldhpresx(*)
            real(c_float), intent(in) :: oldhpresy(*)
            real(c_float), intent(in) :: oldhpresz(*)
            real(c_float), intent(in) :: oldhpresxy(*)
            real(c_float), intent(in) :: oldhpresxz(*)
            real(c_float), intent(in) :: oldhpresyz(*)
            real(c_float), intent(in) :: oldhpx(*)
            real(c_float), intent(in) :: oldhpy(*)
            real(c_float), intent(in) :: oldhpz(*)
            real(c_float), intent(inout) :: orho(*)
            real(c_float), intent(inout) :: opresx(*)
            real(c_float), intent(inout) :: opresy(*)
            real(c_float), intent(inout) :: opresz(*)
            real(c_float), intent(inout) :: opresxy(*)
            real(c_float), intent(inout) :: opresxz(*)
            real(c_float), intent(inout) :: opresyz(*)
            real(c_float), intent(inout) :: opx(*)
            real(c_float), intent(inout) :: opy(*)
            real(c_float), intent(inout) :: opz(*)
            real(c_float), intent(in) :: wrkorho(*)
            real(c_float), intent(in) :: wrkopresx(*)
            real(c_float), intent(in) :: wrkopresy(*)
            real(c_float), intent(in) :: wrkopresz(*)
            real(c_float), intent(in) :: wrkopresxy(*)
            real(c_float), intent(in) :: wrkopresxz(*)
            real(c_float), intent(in) :: wrkopresyz(*)
            real(c_float), intent(in) :: wrkopx(*)
            real(c_float), intent(in) :: wrkopy(*)
            real(c_float), intent(in) :: wrkopz(*)
            real(c_float), intent(in) :: oldorho(*)
            real(c_float), intent(in) :: oldopresx(*)
            real(c_float), intent(in) :: oldopresy(*)
            real(c_float), intent(in) :: oldopresz(*)
            real(c_float), intent(in) :: oldopresxy(*)
            real(c_float), intent(in) :: oldopresxz(*)
            real(c_float), intent(in) :: oldopresyz(*)
            real(c_float), intent(in) :: oldopx(*)
            real(c_float), intent(in) :: oldopy(*)
            real(c_float), intent(in) :: oldopz(*)
            real(c_float), intent(inout) :: epres(*)
            real(c_float), intent(in) :: wrkepres(*)
            real(c_float), intent(in) :: oldepres(*)
            real(c_float), intent(inout) :: bx(*)
            real(c_float), intent(inout) :: by(*)
            real(c_float), intent(inout) :: bz(*)
            real(c_float), intent(in) :: wrkbx(*)
            real(c_float), intent(in) :: wrkby(*)
            real(c_float), intent(in) :: wrkbz(*)
            real(c_float), intent(in) :: oldbx(*)
            real(c_float), intent(in) :: oldby(*)
            real(c_float), intent(in) :: oldbz(*)
            real(c_float), intent(inout) :: vvx(*)
            real(c_float), intent(inout) :: vvy(*)
            real(c_float), intent(inout) :: vvz(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: difrho
            real(c_float), intent(in) :: diferg
            real(c_float), intent(in) :: xspac(*)
            logical(c_int), intent(in) :: isotropic
        end subroutine
        subroutine g3m3_bndry_outer(qrho, qpresx, qpresy, qpresz, qpx, qpy, qpz, hrho, &
            hpresx, hpresy, hpresz, hpx, hpy, hpz, orho, opresx, opresy, opresz, opx, &
            opy, opz, epres, qpresxy, qpresxz, qpresyz, hpresxy, hpresxz, hpresyz, &
            opresxy, opresxz, opresyz, rmassq, rmassh, rmasso, bx, by, bz, bx0, by0, &
            bz0, nx, ny, nz, n_grids, srho, rho_frac, o_conc, spress, spx, spy, spz, &
            sbx_wind, sby_wind, sbz_wind, ti_te, isotropic) bind(c, name='g3m3_bndry_outer')
            import :: c_float, c_int
            real(c_float), intent(inout) :: qrho(*)
            real(c_float), intent(inout) :: qpresx(*)
            real(c_float), intent(inout) :: qpresy(*)
            real(c_float), intent(inout) :: qpresz(*)
            real(c_float), intent(inout) :: qpx(*)
            real(c_float), intent(inout) :: qpy(*)
            real(c_float), intent(inout) :: qpz(*)
            real(c_float), intent(inout) :: hrho(*)
            real(c_float), intent(inout) :: hpresx(*)
            real(c_float), intent(inout) :: hpresy(*)
            real(c_float), intent(inout) :: hpresz(*)
            real(c_float), intent(inout) :: hpx(*)
            real(c_float), intent(inout) :: hpy(*)
            real(c_float), intent(inout) :: hpz(*)
            real(c_float), intent(inout) :: orho(*)
            real(c_float), intent(inout) :: opresx(*)
            real(c_float), intent(inout) :: opresy(*)
            real(c_float), intent(inout) :: opresz(*)
            real(c_float), intent(inout) :: opx(*)
            real(c_float), intent(inout) :: opy(*)
            real(c_float), intent(inout) :: opz(*)
            real(c_float), intent(inout) :: epres(*)
            real(c_float), intent(inout) :: qpresxy(*)
            real(c_float), intent(inout) :: qpresxz(*)
            real(c_float), intent(inout) :: qpresyz(*)
            real(c_float), intent(inout) :: hpresxy(*)
            real(c_float), intent(inout) :: hpresxz(*)
            real(c_float), intent(inout) :: hpresyz(*)
            real(c_float), intent(inout) :: opresxy(*)
            real(c_float), intent(inout) :: opresxz(*)
            real(c_float), intent(inout) :: opresyz(*)
            real(c_float), intent(in) :: rmassq
            real(c_float), intent(in) :: rmassh
            real(c_float), intent(in) :: rmasso
            real(c_float), intent(inout) :: bx(*)
            real(c_float), intent(inout) :: by(*)
            real(c_float), intent(inout) :: bz(*)
            real(c_float), intent(in) :: bx0(*)
            real(c_float), intent(in) :: by0(*)
            real(c_float), intent(in) :: bz0(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            real(c_float), intent(in) :: srho
            real(c_float), intent(in) :: rho_frac
            real(c_float), intent(in) :: o_conc
            real(c_float), intent(in) :: spress
            real(c_float), intent(in) :: spx
            real(c_float), intent(in) :: spy
            real(c_float), intent(in) :: spz
            real(c_float), intent(in) :: sbx_wind
            real(c_float), intent(in) :: sby_wind
            real(c_float), intent(in) :: sbz_wind
            real(c_float), intent(in) :: ti_te
            logical(c_int), intent(in) :: isotropic
        end subroutine
        subroutine g3m3_bndry_inner(qrho, qpresx, qpresy, qpresz, qpx, qpy, qpz, &
            rmassq, hrho, hpresx, hpresy, hpresz, hpx, hpy, hpz, rmassh, orho, opresx, &
            opresy, opresz, opx, opy, opz, rmasso, epres, bx, by, bz, qpresxy, qpresxz, &
            qpresyz, hpresxy, hpresxz, hpresyz, opresxy, opresxz, opresyz, nx, ny, nz, &
            n_grids, parm_srf, parm_mid, parm_zero, ijsrf, numsrf, ijmid, nummid, &
            ijzero, numzero, mbndry, msrf, mmid, mzero, erho, epress, alpha_e, ti_te, &
            o_conc, re_equiv, r_inner, sbx_wind, spress, grd_xmin, grd_xmax, grd_ymin, &
            grd_ymax, grd_zmin, grd_zmax, v_rot, r_rot) bind(c, name='g3m3_bndry_inner')
            import :: c_float, c_int
            real(c_float), intent(inout) :: qrho(*)
            real(c_float), intent(inout) :: qpresx(*)
            real(c_float), intent(inout) :: qpresy(*)
            real(c_float), intent(inout) :: qpresz(*)
            real(c_float), intent(inout) :: qpx(*)
            real(c_float), intent(inout) :: qpy(*)
            real(c_float), intent(inout) :: qpz(*)
            real(c_float), intent(in) :: rmassq
            real(c_float), intent(inout) :: hrho(*)
            real(c_float), intent(inout) :: hpresx(*)
            real(c_float), intent(inout) :: hpresy(*)
            real(c_float), intent(inout) :: hpresz(*)
            real(c_float), intent(inout) :: hpx(*)
            real(c_float), intent(inout) :: hpy(*)
            real(c_float), intent(inout) :: hpz(*)
            real(c_float), intent(in) :: rmassh
            real(c_float), intent(inout) :: orho(*)
            real(c_float), intent(inout) :: opresx(*)
            real(c_float), intent(inout) :: opresy(*)
            real(c_float), intent(inout) :: opresz(*)
            real(c_float), intent(inout) :: opx(*)
            real(c_float), intent(inout) :: opy(*)
            real(c_float), intent(inout) :: opz(*)
            real(c_float), intent(in) :: rmasso
            real(c_float), intent(inout) :: epres(*)
            real(c_float), intent(inout) :: bx(*)
            real(c_float), intent(inout) :: by(*)
            real(c_float), intent(inout) :: bz(*)
            real(c_float), intent(inout) :: qpresxy(*)
            real(c_float), intent(inout) :: qpresxz(*)
            real(c_float), intent(inout) :: qpresyz(*)
            real(c_float), intent(inout) :: hpresxy(*)
            real(c_float), intent(inout) :: hpresxz(*)
            real(c_float), intent(inout) :: hpresyz(*)
            real(c_float), intent(inout) :: opresxy(*)
            real(c_float), intent(inout) :: opresxz(*)
            real(c_float), intent(inout) :: opresyz(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            real(c_float), intent(in) :: parm_srf(*)
            real(c_float), intent(in) :: parm_mid(*)
            real(c_float), intent(in) :: parm_zero(*)
            integer(c_int), intent(in) :: ijsrf(*)
            integer(c_int), intent(in) :: numsrf(*)
            integer(c_int), intent(in) :: ijmid(*)
            integer(c_int), intent(in) :: nummid(*)
            integer(c_int), intent(in) :: ijzero(*)
            integer(c_int), intent(in) :: numzero(*)
            integer(c_int), intent(in) :: mbndry
            integer(c_int), intent(in) :: msrf
            integer(c_int), intent(in) :: mmid
            integer(c_int), intent(in) :: mzero
            real(c_float), intent(in) :: erho
            real(c_float), intent(in) :: epress
            real(c_float), intent(in) :: alpha_e
            real(c_float), intent(in) :: ti_te
            real(c_float), intent(in) :: o_conc
            real(c_float), intent(in) :: re_equiv
            real(c_float), intent(in) :: r_inner
            real(c_float), intent(in) :: sbx_wind
            real(c_float), intent(in) :: spress
            real(c_float), intent(in) :: grd_xmin(*)
            real(c_float), intent(in) :: grd_xmax(*)
            real(c_float), intent(in) :: grd_ymin(*)
            real(c_float), intent(in) :: grd_ymax(*)
            real(c_float), intent(in) :: grd_zmin(*)
            real(c_float), intent(in) :: grd_zmax(*)
            real(c_float), intent(in) :: v_rot
            real(c_float), intent(in) :: r_rot
        end subroutine
        subroutine g3m3_flanks(wrkrho, rho, oldrho, nx, ny, nz, n_grids, box, t_new, &
            t_old, t, work, grd_xmin, grd_xmax, grd_ymin, grd_ymax, grd_zmin, grd_zmax) bind(c, name='g3m3_flanks')
            import :: c_double, c_float, c_int
            real(c_float), intent(inout) :: wrkrho(*)
            real(c_float), intent(in) :: rho(*)
            real(c_float), intent(in) :: oldrho(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_double), intent(in) :: t_new(*)
            real(c_double), intent(in) :: t_old(*)
            real(c_double), intent(in) :: t
            real(c_float), intent(inout) :: work(*)
            real(c_float), intent(in) :: grd_xmin(*)
            real(c_float), intent(in) :: grd_xmax(*)
            real(c_float), intent(in) :: grd_ymin(*)
            real(c_float), intent(in) :: grd_ymax(*)
            real(c_float), intent(in) :: grd_zmin(*)
            real(c_float), intent(in) :: grd_zmax(*)
        end subroutine
        subroutine g3m3_corer(rho, nx, ny, nz, n_grids, box, grd_xmin, grd_xmax, &
            grd_ymin, grd_ymax, grd_zmin, grd_zmax) bind(c, name='g3m3_corer')
            import :: c_float, c_int
            real(c_float), intent(inout) :: rho(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: grd_xmin(*)
            real(c_float), intent(in) :: grd_xmax(*)
            real(c_float), intent(in) :: grd_ymin(*)
            real(c_float), intent(in) :: grd_ymax(*)
            real(c_float), intent(in) :: grd_zmin(*)
            real(c_float), intent(in) :: grd_zmax(*)
        end subroutine
        subroutine g3m3_bndry_flanks(wrkqrho, wrkqpx, wrkqpy, wrkqpz, wrkqpresx, &
            wrkqpresy, wrkqpresz, wrkqpresxy, wrkqpresxz, wrkqpresyz, wrkhrho, wrkhpx, &
            wrkhpy, wrkhpz, wrkhpresx, wrkhpresy, wrkhpresz, wrkhpresxy, wrkhpresxz, &
            wrkhpresyz, wrkorho, wrkopx, wrkopy, wrkopz, wrkopresx, wrkopresy, &
            wrkopresz, wrkopresxy, wrkopresxz, wrkopresyz, wrkepres, wrkbx, wrkby, &
            wrkbz, qrho, qpx, qpy, qpz, qpresx, qpresy, qpresz, qpresxy, qpresxz, &
            qpresyz, hrho, hpx, hpy, hpz, hpresx, hpresy, hpresz, hpresxy, hpresxz, &
            hpresyz, orho, opx, opy, opz, opresx, opresy, opresz, opresxy, opresxz, &
            opresyz, epres, bx, by, bz, oldqrho, oldqpx, oldqpy, oldqpz, oldqpresx, &
            oldqpresy, oldqpresz, oldqpresxy, oldqpresxz, oldqpresyz, oldhrho, oldhpx, &
            oldhpy, oldhpz, oldhpresx, oldhpresy, oldhpresz, oldhpresxy, oldhpresxz, &
            oldhpresyz, oldorho, oldopx, oldopy, oldopz, oldopresx, oldopresy, &
            oldopresz, oldopresxy, oldopresxz, oldopresyz, oldepres, oldbx, oldby, &
            oldbz, work, nx, ny, nz, n_grids, box, t_old, t_new, t, grd_xmin, grd_xmax, &
            grd_ymin, grd_ymax, grd_zmin, grd_zmax) bind(c, name='g3m3_bndry_flanks')
            import :: c_double, c_float, c_int
            real(c_float), intent(inout) :: wrkqrho(*)
            real(c_float), intent(inout) :: wrkqpx(*)
            real(c_float), intent(inout) :: wrkqpy(*)
            real(c_float), intent(inout) :: wrkqpz(*)
            real(c_float), intent(inout) :: wrkqpresx(*)
            real(c_float), intent(inout) :: wrkqpresy(*)
            real(c_float), intent(inout) :: wrkqpresz(*)
            real(c_float), intent(inout) :: wrkqpresxy(*)
            real(c_float), intent(inout) :: wrkqpresxz(*)
            real(c_float), intent(inout) :: wrkqpresyz(*)
            real(c_float), intent(inout) :: wrkhrho(*)
            real(c_float), intent(inout) :: wrkhpx(*)
            real(c_float), intent(inout) :: wrkhpy(*)
            real(c_float), intent(inout) :: wrkhpz(*)
            real(c_float), intent(inout) :: wrkhpresx(*)
            real(c_float), intent(inout) :: wrkhpresy(*)
            real(c_float), intent(inout) :: wrkhpresz(*)
            real(c_float), intent(inout) :: wrkhpresxy(*)
            real(c_float), intent(inout) :: wrkhpresxz(*)
            real(c_float), intent(inout) :: wrkhpresyz(*)
            real(c_float), intent(inout) :: wrkorho(*)
            real(c_float), intent(inout) :: wrkopx(*)
            real(c_float), intent(inout) :: wrkopy(*)
            real(c_float), intent(inout) :: wrkopz(*)
            real(c_float), intent(inout) :: wrkopresx(*)
            real(c_float), intent(inout) :: wrkopresy(*)
            real(c_float), intent(inout) :: wrkopresz(*)
            real(c_float), intent(inout) :: wrkopresxy(*)
            real(c_float), intent(inout) :: wrkopresxz(*)
            real(c_float), intent(inout) :: wrkopresyz(*)
            real(c_float), intent(inout) :: wrkepres(*)
            real(c_float), intent(inout) :: wrkbx(*)
            real(c_float), intent(inout) :: wrkby(*)
            real(c_float), intent(inout) :: wrkbz(*)
            real(c_float), intent(in) :: qrho(*)
            real(c_float), intent(in) :: qpx(*)
            real(c_float), intent(in) :: qpy(*)
            real(c_float), intent(in) :: qpz(*)
            real(c_float), intent(in) :: qpresx(*)
            real(c_float), intent(in) :: qpresy(*)
            real(c_float), intent(in) :: qpresz(*)
            real(c_float), intent(in) :: qpresxy(*)
            real(c_float), intent(in) :: qpresxz(*)
            real(c_float), intent(in) :: qpresyz(*)
            real(c_float), intent(in) :: hrho(*)
            real(c_float), intent(in) :: hpx(*)
            real(c_float), intent(in) :: hpy(*)
            real(c_float), intent(in) :: hpz(*)
            real(c_float), intent(in) :: hpresx(*)
            real(c_float), intent(in) :: hpresy(*)
            real(c_float), intent(in) :: hpresz(*)
            real(c_float), intent(in) :: hpresxy(*)
            real(c_float), intent(in) :: hpresxz(*)
            real(c_float), intent(in) :: hpresyz(*)
            real(c_float), intent(in) :: orho(*)
            real(c_float), intent(in) :: opx(*)
            real(c_float), intent(in) :: opy(*)
            real(c_float), intent(in) :: opz(*)
            real(c_float), intent(in) :: opresx(*)
            real(c_float), intent(in) :: opresy(*)
            real(c_float), intent(in) :: opresz(*)
            real(c_float), intent(in) :: opresxy(*)
            real(c_float), intent(in) :: opresxz(*)
            real(c_float), intent(in) :: opresyz(*)
            real(c_float), intent(in) :: epres(*)
            real(c_float), intent(in) :: bx(*)
            real(c_float), intent(in) :: by(*)
            real(c_float), intent(in) :: bz(*)
            real(c_float), intent(in) :: oldqrho(*)
            real(c_float), intent(in) :: oldqpx(*)
            real(c_float), intent(in) :: oldqpy(*)
            real(c_float), intent(in) :: oldqpz(*)
            real(c_float), intent(in) :: oldqpresx(*)
            real(c_float), intent(in) :: oldqpresy(*)
            real(c_float), intent(in) :: oldqpresz(*)
            real(c_float), intent(in) :: oldqpresxy(*)
            real(c_float), intent(in) :: oldqpresxz(*)
            real(c_float), intent(in) :: oldqpresyz(*)
            real(c_float), intent(in) :: oldhrho(*)
            real(c_float), intent(in) :: oldhpx(*)
            real(c_float), intent(in) :: oldhpy(*)
            real(c_float), intent(in) :: oldhpz(*)
            real(c_float), intent(in) :: oldhpresx(*)
            real(c_float), intent(in) :: oldhpresy(*)
            real(c_float), intent(in) :: oldhpresz(*)
            real(c_float), intent(in) :: oldhpresxy(*)
            real(c_float), intent(in) :: oldhpresxz(*)
            real(c_float), intent(in) :: oldhpresyz(*)
            real(c_float), intent(in) :: oldorho(*)
            real(c_float), intent(in) :: oldopx(*)
            real(c_float), intent(in) :: oldopy(*)
            real(c_float), intent(in) :: oldopz(*)
            real(c_float), intent(in) :: oldopresx(*)
            real(c_float), intent(in) :: oldopresy(*)
            real(c_float), intent(in) :: oldopresz(*)
            real(c_float), intent(in) :: oldopresxy(*)
            real(c_float), intent(in) :: oldopresxz(*)
            real(c_float), intent(in) :: oldopresyz(*)
            real(c_float), intent(in) :: oldepres(*)
            real(c_float), intent(in) :: oldbx(*)
            real(c_float), intent(in) :: oldby(*)
            real(c_float), intent(in) :: oldbz(*)
            real(c_float), intent(inout) :: work(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_double), intent(in) :: t_old(*)
            real(c_double), intent(in) :: t_new(*)
            real(c_double), intent(in) :: t
            real(c_float), intent(in) :: grd_xmin(*)
            real(c_float), intent(in) :: grd_xmax(*)
            real(c_float), intent(in) :: grd_ymin(*)
            real(c_float), intent(in) :: grd_ymax(*)
            real(c_float), intent(in) :: grd_zmin(*)
            real(c_float), intent(in) :: grd_zmax(*)
        end subroutine
        subroutine g3m3_bndry_corer(qrho, qpresx, qpresy, qpresz, qpx, qpy, qpz, hrho, &
            hpresx, hpresy, hpresz, hpx, hpy, hpz, orho, opresx, opresy, opresz, opx, &
            opy, opz, epres, bx, by, bz, qpresxy, qpresxz, qpresyz, hpresxy, hpresxz, &
            hpresyz, opresxy, opresxz, opresyz, nx, ny, nz, n_grids, box, grd_xmin, &
            grd_xmax, grd_ymin, grd_ymax, grd_zmin, grd_zmax) bind(c, name='g3m3_bndry_corer')
            import :: c_float, c_int
            real(c_float), intent(inout) :: qrho(*)
            real(c_float), intent(inout) :: qpresx(*)
            real(c_float), intent(inout) :: qpresy(*)
            real(c_float), intent(inout) :: qpresz(*)
            real(c_float), intent(inout) :: qpx(*)
            real(c_float), intent(inout) :: qpy(*)
            real(c_float), intent(inout) :: qpz(*)
            real(c_float), intent(inout) :: hrho(*)
            real(c_float), intent(inout) :: hpresx(*)
            real(c_float), intent(inout) :: hpresy(*)
            real(c_float), intent(inout) :: hpresz(*)
            real(c_float), intent(inout) :: hpx(*)
            real(c_float), intent(inout) :: hpy(*)
            real(c_float), intent(inout) :: hpz(*)
            real(c_float), intent(inout) :: orho(*)
            real(c_float), intent(inout) :: opresx(*)
            real(c_float), intent(inout) :: opresy(*)
            real(c_float), intent(inout) :: opresz(*)
            real(c_float), intent(inout) :: opx(*)
            real(c_float), intent(inout) :: opy(*)
            real(c_float), intent(inout) :: opz(*)
            real(c_float), intent(inout) :: epres(*)
            real(c_float), intent(inout) :: bx(*)
            real(c_float), intent(inout) :: by(*)
            real(c_float), intent(inout) :: bz(*)
            real(c_float), intent(inout) :: qpresxy(*)
            real(c_float), intent(inout) :: qpresxz(*)
            real(c_float), intent(inout) :: qpresyz(*)
            real(c_float), intent(inout) :: hpresxy(*)
            real(c_float), intent(inout) :: hpresxz(*)
            real(c_float), intent(inout) :: hpresyz(*)
            real(c_float), intent(inout) :: opresxy(*)
            real(c_float), intent(inout) :: opresxz(*)
            real(c_float), intent(inout) :: opresyz(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: grd_xmin(*)
            real(c_float), intent(in) :: grd_xmax(*)
            real(c_float), intent(in) :: grd_ymin(*)
            real(c_float), intent(in) :: grd_ymax(*)
            real(c_float), intent(in) :: grd_zmin(*)
            real(c_float), intent(in) :: grd_zmax(*)
        end subroutine
        subroutine g3m3_flanks_synced(rho, nx, ny, nz, n_grids, box, grd_xmin, &
            grd_xmax, grd_ymin, grd_ymax, grd_zmin, grd_zmax) bind(c, name='g3m3_flanks_synced')
            import :: c_float, c_int
            real(c_float), intent(inout) :: rho(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: grd_xmin(*)
            real(c_float), intent(in) :: grd_xmax(*)
            real(c_float), intent(in) :: grd_ymin(*)
            real(c_float), intent(in) :: grd_ymax(*)
            real(c_float), intent(in) :: grd_zmin(*)
            real(c_float), intent(in) :: grd_zmax(*)
        end subroutine
        subroutine g3m3_banded_tpose(nbands, npts, divB_band_pos, banded_matrix, x_vec, &
            tp_result) bind(c, name='g3m3_banded_tpose')
            import :: c_float, c_int
            integer(c_int), intent(in) :: nbands
            integer(c_int), intent(in) :: npts
            integer(c_int), intent(in) :: divB_band_pos(*)
            real(c_float), intent(in) :: banded_matrix(*)
            real(c_float), intent(in) :: x_vec(*)
            real(c_float), intent(inout) :: tp_result(*)
        end subroutine
        subroutine g3m3_divb_correct(nx, ny, nz, n_grids, box, nbands, divB_band_pos, &
            xspac, divB_errlim, bx, by, bz) bind(c, name='g3m3_divb_correct')
            import :: c_float, c_int
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            integer(c_int), intent(in) :: nbands
            integer(c_int), intent(in) :: divB_band_pos(*)
            real(c_float), intent(in) :: xspac(*)
            real(c_float), intent(in) :: divB_errlim
            real(c_float), intent(inout) :: bx(*)
            real(c_float), intent(inout) :: by(*)
            real(c_float), intent(inout) :: bz(*)
        end subroutine
        integer(c_int) function g3m3_divb_correct_ex(nx, ny, nz, n_grids, box, xspac, &
            divB_errlim, bx, by, bz, mode, max_iters, iters, residual2) bind(c, name='g3m3_divb_correct_ex')
            import :: c_double, c_float, c_int, c_long_long
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: xspac(*)
            real(c_float), intent(in) :: divB_errlim
            real(c_float), intent(inout) :: bx(*)
            real(c_float), intent(inout) :: by(*)
            real(c_float), intent(inout) :: bz(*)
            integer(c_int), intent(in) :: mode
            integer(c_long_long), intent(in) :: max_iters
            integer(c_long_long), intent(inout) :: iters
            real(c_double), intent(inout) :: residual2
        end function
        subroutine g3m3_trilin_interp(sxyz, gridpts, qty, scdata) bind(c, name='g3m3_trilin_interp')
            import :: c_float
            real(c_float), intent(in) :: sxyz(*)
            real(c_float), intent(in) :: gridpts(*)
            real(c_float), intent(in) :: qty(*)
            real(c_float), intent(inout) :: scdata
        end subroutine
        subroutine g3m3_sparse_solver(nbands, npts, divB_band_pos, divB_errlim, &
            recurse, redo_lim, divB, soln_vec, residual2, banded_matrix) bind(c, name='g3m3_sparse_solver')
            import :: c_float, c_int
            integer(c_int), intent(in) :: nbands
            integer(c_int), intent(in) :: npts
            integer(c_int), intent(in) :: divB_band_pos(*)
            real(c_float), intent(in) :: divB_errlim
            logical(c_int), intent(in) :: recurse
            integer(c_int), intent(in) :: redo_lim
            real(c_float), intent(in) :: divB(*)
            real(c_float), intent(inout) :: soln_vec(*)
            real(c_float), intent(inout) :: residual2
            real(c_float), intent(in) :: banded_matrix(*)
        end subroutine
        subroutine g3m3_fnd_pres(px, py, pz, pxy, pxz, pyz, p_para, p_perp, p_cross, &
            vvx, vvy, vvz, bxt, byt, bzt, nx, ny, nz, n_grids, box) bind(c, name='g3m3_fnd_pres')
            import :: c_float, c_int
            real(c_float), intent(in) :: px(*)
            real(c_float), intent(in) :: py(*)
            real(c_float), intent(in) :: pz(*)
            real(c_float), intent(in) :: pxy(*)
            real(c_float), intent(in) :: pxz(*)
            real(c_float), intent(in) :: pyz(*)
            real(c_float), intent(inout) :: p_para(*)
            real(c_float), intent(inout) :: p_perp(*)
            real(c_float), intent(inout) :: p_cross(*)
            real(c_float), intent(in) :: vvx(*)
            real(c_float), intent(in) :: vvy(*)
            real(c_float), intent(in) :: vvz(*)
            real(c_float), intent(in) :: bxt(*)
            real(c_float), intent(in) :: byt(*)
            real(c_float), intent(in) :: bzt(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
        end subroutine
        subroutine g3m3_findgrid(xyz, n_grids, grid_minvals, grid_maxvals, xspac, &
            gridpt, re_equiv, cname, cname_len) bind(c, name='g3m3_findgrid')
            import :: c_char, c_float, c_int, c_size_t
            real(c_float), intent(in) :: xyz(*)
            integer(c_int), intent(in) :: n_grids
            real(c_float), intent(in) :: grid_minvals(*)
            real(c_float), intent(in) :: grid_maxvals(*)
            real(c_float), intent(in) :: xspac(*)
            integer(c_int), intent(inout) :: gridpt(*)
            real(c_float), intent(in) :: re_equiv
            character(kind=c_char), intent(in) :: cname(*)
            integer(c_size_t), value :: cname_len
        end subroutine
        subroutine g3m3_sample_fields(npts, sxyz, n_grids, grid_minvals, grid_maxvals, &
            xspac, re_equiv, nx, ny, nz, nq, fields, scdata, gridpt) bind(c, name='g3m3_sample_fields')
            import :: c_float, c_int, c_ptr
            integer(c_int), intent(in) :: npts
            real(c_float), intent(in) :: sxyz(*)
            integer(c_int), intent(in) :: n_grids
            real(c_float), intent(in) :: grid_minvals(*)
            real(c_float), intent(in) :: grid_maxvals(*)
            real(c_float), intent(in) :: xspac(*)
            real(c_float), intent(in) :: re_equiv
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: nq
            type(c_ptr), intent(in) :: fields(*)
            real(c_float), intent(inout) :: scdata(*)
            integer(c_int), intent(inout) :: gridpt(*)
        end subroutine
        subroutine g3m3_find_aux_cube(cname, i_craft, scout, n_grids, grid_minvals, &
            grid_maxvals, xspac, re_equiv, xcraftn, zcraftn, ut, ntimesn, recordingn, &
            gridpt, sxyz, cube_vertices, cname_len) bind(c, name='g3m3_find_aux_cube')
            import :: c_char, c_float, c_int, c_size_t
            character(kind=c_char), intent(in) :: cname(*)
            integer(c_int), intent(in) :: i_craft
            integer(c_int), intent(in) :: scout
            integer(c_int), intent(in) :: n_grids
            real(c_float), intent(in) :: grid_minvals(*)
            real(c_float), intent(in) :: grid_maxvals(*)
            real(c_float), intent(in) :: xspac(*)
            real(c_float), intent(in) :: re_equiv
            real(c_float), intent(in) :: xcraftn(*)
            real(c_float), intent(in) :: zcraftn(*)
            real(c_float), intent(in) :: ut
            integer(c_int), intent(inout) :: ntimesn(*)
            logical(c_int), intent(inout) :: recordingn
            integer(c_int), intent(inout) :: gridpt(*)
            real(c_float), intent(inout) :: sxyz(*)
            real(c_float), intent(inout) :: cube_vertices(*)
            integer(c_size_t), value :: cname_len
        end subroutine
        subroutine g3m3_sample_craft(ncraft, xcraft, zcraft, ut, n_grids, grid_minvals, &
            grid_maxvals, xspac, re_equiv, nx, ny, nz, nq, fields, sxyz, scdata, gridpt) bind(c, name='g3m3_sample_craft')
            import :: c_float, c_int, c_ptr
            integer(c_int), intent(in) :: ncraft
            real(c_float), intent(in) :: xcraft(*)
            real(c_float), intent(in) :: zcraft(*)
            real(c_float), intent(in) :: ut
            integer(c_int), intent(in) :: n_grids
            real(c_float), intent(in) :: grid_minvals(*)
            real(c_float), intent(in) :: grid_maxvals(*)
            real(c_float), intent(in) :: xspac(*)
            real(c_float), intent(in) :: re_equiv
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: nq
            type(c_ptr), intent(in) :: fields(*)
            real(c_float), intent(inout) :: sxyz(*)
            real(c_float), intent(inout) :: scdata(*)
            integer(c_int), intent(inout) :: gridpt(*)
        end subroutine
        subroutine g3m3_flux_counter(qpx, qpy, qpz, hpx, hpy, hpz, opx, opy, opz, nx, &
            ny, nz, n_grids, box, rmassq, rmassh, rmasso, grid_spacing, qflux_in, &
            qflux_out, hflux_in, hflux_out, oflux_in, oflux_out) bind(c, name='g3m3_flux_counter')
            import :: c_float, c_int
            real(c_float), intent(in) :: qpx(*)
            real(c_float), intent(in) :: qpy(*)
            real(c_float), intent(in) :: qpz(*)
            real(c_float), intent(in) :: hpx(*)
            real(c_float), intent(in) :: hpy(*)
            real(c_float), intent(in) :: hpz(*)
            real(c_float), intent(in) :: opx(*)
            real(c_float), intent(in) :: opy(*)
            real(c_float), intent(in) :: opz(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: box
            real(c_float), intent(in) :: rmassq
            real(c_float), intent(in) :: rmassh
            real(c_float), intent(in) :: rmasso
            real(c_float), intent(in) :: grid_spacing
            real(c_float), intent(inout) :: qflux_in
            real(c_float), intent(inout) :: qflux_out
            real(c_float), intent(inout) :: hflux_in
            real(c_float), intent(inout) :: hflux_out
            real(c_float), intent(inout) :: oflux_in
            real(c_float), intent(inout) :: oflux_out
        end subroutine
        subroutine g3m3_dipole(abx, aby, abz, ax, ay, az, rot_angle, xdip, ydip, zdip, &
            sin_tilt, cos_tilt, b0) bind(c, name='g3m3_dipole')
            import :: c_float
            real(c_float), intent(inout) :: abx
            real(c_float), intent(inout) :: aby
            real(c_float), intent(inout) :: abz
            real(c_float), intent(in) :: ax
            real(c_float), intent(in) :: ay
            real(c_float), intent(in) :: az
            real(c_float), intent(in) :: rot_angle
            real(c_float), intent(in) :: xdip
            real(c_float), intent(in) :: ydip
            real(c_float), intent(in) :: zdip
            real(c_float), intent(in) :: sin_tilt
            real(c_float), intent(in) :: cos_tilt
            real(c_float), intent(in) :: b0
        end subroutine
        subroutine g3m3_mak_dip_grd(bx0, by0, bz0, nx, ny, nz, n_grids, mbndry, box, &
            ijzero, numzero, mzero, r_inner, grd_xmin, grd_xmax, grd_ymin, grd_ymax, &
            grd_zmin, grd_zmax, rot_angle, xdip, ydip, zdip, sin_tilt, cos_tilt, b0) bind(c, name='g3m3_mak_dip_grd')
            import :: c_float, c_int
            real(c_float), intent(inout) :: bx0(*)
            real(c_float), intent(inout) :: by0(*)
            real(c_float), intent(inout) :: bz0(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: n_grids
            integer(c_int), intent(in) :: mbndry
            integer(c_int), intent(in) :: box
            integer(c_int), intent(in) :: ijzero(*)
            integer(c_int), intent(in) :: numzero(*)
            integer(c_int), intent(in) :: mzero
            real(c_float), intent(in) :: r_inner
            real(c_float), intent(in) :: grd_xmin(*)
            real(c_float), intent(in) :: grd_xmax(*)
            real(c_float), intent(in) :: grd_ymin(*)
            real(c_float), intent(in) :: grd_ymax(*)
            real(c_float), intent(in) :: grd_zmin(*)
            real(c_float), intent(in) :: grd_zmax(*)
            real(c_float), intent(in) :: rot_angle
            real(c_float), intent(in) :: xdip
            real(c_float), intent(in) :: ydip
            real(c_float), intent(in) :: zdip
            real(c_float), intent(in) :: sin_tilt
            real(c_float), intent(in) :: cos_tilt
            real(c_float), intent(in) :: b0
        end subroutine
        subroutine g3m3_set_resist(rst, nx, ny, nz, mbndry, resist, ijzero, numzero, &
            ijmid, nummid, ijsrf, numsrf, msrf, mmid, mzero, b0) bind(c, name='g3m3_set_resist')
            import :: c_float, c_int
            real(c_float), intent(inout) :: rst(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: mbndry
            real(c_float), intent(in) :: resist
            integer(c_int), intent(in) :: ijzero(*)
            integer(c_int), intent(in) :: numzero(*)
            integer(c_int), intent(in) :: ijmid(*)
            integer(c_int), intent(in) :: nummid(*)
            integer(c_int), intent(in) :: ijsrf(*)
            integer(c_int), intent(in) :: numsrf(*)
            integer(c_int), intent(in) :: msrf
            integer(c_int), intent(in) :: mmid
            integer(c_int), intent(in) :: mzero
            real(c_float), intent(in) :: b0
        end subroutine
        subroutine g3m3_sample_fields_adjoint(npts, sxyz, n_grids, grid_minvals, &
            grid_maxvals, xspac, re_equiv, nx, ny, nz, nq, fields_bar, scdata_bar) bind(c, name='g3m3_sample_fields_adjoint')
            import :: c_float, c_int, c_ptr
            integer(c_int), intent(in) :: npts
            real(c_float), intent(in) :: sxyz(*)
            integer(c_int), intent(in) :: n_grids
            real(c_float), intent(in) :: grid_minvals(*)
            real(c_float), intent(in) :: grid_maxvals(*)
            real(c_float), intent(in) :: xspac(*)
            real(c_float), intent(in) :: re_equiv
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            integer(c_int), intent(in) :: nq
            type(c_ptr), intent(in) :: fields_bar(*)
            real(c_float), intent(in) :: scdata_bar(*)
        end subroutine
        subroutine g3m3_push_ion_adjoint(src, bsx, bsy, bsz, btot, efldx, efldy, efldz, &
            tvx, tvy, tvz, out_bar, src_bar, old_bar, bsx_bar, bsy_bar, bsz_bar, &
            btot_bar, efldx_bar, efldy_bar, efldz_bar, tvx_bar, tvy_bar, tvz_bar, nx, &
            ny, nz, rmass, delt, grav, re_equiv, reynolds, grd_xmin, grd_xmax, &
            grd_ymin, grd_ymax, grd_zmin, grd_zmax, work) bind(c, name='g3m3_push_ion_adjoint')
            import :: c_float, c_int, c_ptr
            type(c_ptr), intent(in) :: src(*)
            real(c_float), intent(in) :: bsx(*)
            real(c_float), intent(in) :: bsy(*)
            real(c_float), intent(in) :: bsz(*)
            real(c_float), intent(in) :: btot(*)
            real(c_float), intent(in) :: efldx(*)
            real(c_float), intent(in) :: efldy(*)
            real(c_float), intent(in) :: efldz(*)
            real(c_float), intent(in) :: tvx(*)
            real(c_float), intent(in) :: tvy(*)
            real(c_float), intent(in) :: tvz(*)
            type(c_ptr), intent(in) :: out_bar(*)
            type(c_ptr), intent(in) :: src_bar(*)
            type(c_ptr), intent(in) :: old_bar(*)
            real(c_float), intent(inout) :: bsx_bar(*)
            real(c_float), intent(inout) :: bsy_bar(*)
            real(c_float), intent(inout) :: bsz_bar(*)
            real(c_float), intent(inout) :: btot_bar(*)
            real(c_float), intent(inout) :: efldx_bar(*)
            real(c_float), intent(inout) :: efldy_bar(*)
            real(c_float), intent(inout) :: efldz_bar(*)
            real(c_float), intent(inout) :: tvx_bar(*)
            real(c_float), intent(inout) :: tvy_bar(*)
            real(c_float), intent(inout) :: tvz_bar(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            real(c_float), intent(in) :: rmass
            real(c_float), intent(in) :: delt
            real(c_float), intent(in) :: grav
            real(c_float), intent(in) :: re_equiv
            real(c_float), intent(in) :: reynolds
            real(c_float), intent(in) :: grd_xmin
            real(c_float), intent(in) :: grd_xmax
            real(c_float), intent(in) :: grd_ymin
            real(c_float), intent(in) :: grd_ymax
            real(c_float), intent(in) :: grd_zmin
            real(c_float), intent(in) :: grd_zmax
            real(c_float), intent(inout) :: work(*)
        end subroutine
        subroutine g3m3_lap_plasma_adjoint(q, vx, vy, vz, w_bar, q_bar, vx_bar, vy_bar, &
            vz_bar, nx, ny, nz, chirho, chipxyz, delt, work) bind(c, name='g3m3_lap_plasma_adjoint')
            import :: c_float, c_int, c_ptr
            type(c_ptr), intent(in) :: q(*)
            real(c_float), intent(in) :: vx(*)
            real(c_float), intent(in) :: vy(*)
            real(c_float), intent(in) :: vz(*)
            type(c_ptr), intent(in) :: w_bar(*)
            type(c_ptr), intent(in) :: q_bar(*)
            real(c_float), intent(inout) :: vx_bar(*)
            real(c_float), intent(inout) :: vy_bar(*)
            real(c_float), intent(inout) :: vz_bar(*)
            integer(c_int), intent(in) :: nx
            integer(c_int), intent(in) :: ny
            integer(c_int), intent(in) :: nz
            real(c_float), intent(in) :: chirho
            real(c_float), intent(in) :: chipxyz
            real(c_float), intent(in) :: delt
            real(c_float), intent(inout) :: work(*)
        end subroutine
        subroutine g3m3_species_velocity_adjoint(rho, px, py, pz, vx_bar, vy_bar, &
            vz_bar, rho_bar, px_bar, py_bar, pz_bar, n, smallbit) bind(c, name='g3m3_species_velocity_adjoint')
            import :: c_float, c_int
            real(c_float), intent(in) :: rho(*)
            real(c_float), intent(in) :: px(*)
            real(c_float), intent(in) :: py(*)
            real(c_float), intent(in) :: pz(*)
            real(c_float), intent(in) :: vx_bar(*)
            real(c_float), intent(in) :: vy_bar(*)
            real(c_float), intent(in) :: vz_bar(*)
            real(c_float), intent(inout) :: rho_bar(*)
            real(c_float), intent(inout) :: px_bar(*)
            real(c_float), intent(inout) :: py_bar(*)
            real(c_float), intent(inout) :: pz_bar(*)
            integer(c_int), intent(in) :: n
            real(c_float), intent(in) :: smallbit
        end subroutine
    end interface
end module g3m3_iface
