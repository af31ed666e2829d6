! Test helper: Fortran unformatted sequential I/O with the fluid-file record
! sequence of the reference's restart files (one t record, 37 field
! records, one planet-list record).  It converts between that format and a
! raw blob (t, the 37 arrays, then the list arrays, no record markers) so a
! test can check the library's reader and writer against the Fortran
! runtime's own record framing.
!
!   fluid_rw w SPEC BLOB FLUID    blob -> fluid file
!   fluid_rw r SPEC FLUID BLOB    fluid file -> blob
! SPEC: nx ny nz n_grids mbndry msrf mmid mzero (list-directed)
program fluid_rw
    implicit none
    character(len=512) :: mode, spec, fin, fout
    integer :: nx, ny, nz, ng, mb, msrf, mmid, mzero, a
    double precision :: t
    real, allocatable :: f(:,:,:,:,:)
    real, allocatable :: ps(:,:,:), pm(:,:,:), pz(:,:,:)
    integer, allocatable :: is(:,:,:), im(:,:,:), iz(:,:,:), ns(:), nm(:), nzr(:)

    call get_command_argument(1, mode)
    call get_command_argument(2, spec)
    call get_command_argument(3, fin)
    call get_command_argument(4, fout)
    open(9, file=trim(spec), status='old')
    read(9, *) nx, ny, nz, ng, mb, msrf, mmid, mzero
    close(9)
    allocate(f(nx, ny, nz, ng, 37))
    allocate(ps(mb, 7, msrf), pm(mb, 7, mmid), pz(mb, 7, mzero))
    allocate(is(mb, 3, msrf), im(mb, 3, mmid), iz(mb, 3, mzero), ns(mb), nm(mb), nzr(mb))

    if (mode(1:1) == 'w') then
        open(10, file=trim(fin), access='stream', form='unformatted', status='old')
        read(10) t, f, ps, pm, pz, iz, nzr, im, nm, is, ns
        close(10)
        open(11, file=trim(fout), form='unformatted', status='replace')
        write(11) t
        do a = 1, 37
            write(11) f(:, :, :, :, a)
        end do
        write(11) ps, pm, pz, iz, nzr, im, nm, is, ns
        close(11)
    else
        open(10, file=trim(fin), form='unformatted', status='old')
        read(10) t
        do a = 1, 37
            read(10) f(:, :, :, :, a)
        end do
        read(10) ps, pm, pz, iz, nzr, im, nm, is, ns
        close(10)
        open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
        write(11) t, f, ps, pm, pz, iz, nzr, im, nm, is, ns
        close(11)
    end if
end program
