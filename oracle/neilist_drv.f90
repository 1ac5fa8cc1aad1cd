! Checker driver for the reference neighbour list (neilist/neilist.F90:148-187).
! Test infrastructure only: links against the reference's own neilist object.
!
! usage: neilist_ref <in.bin> <out.bin>
!   in  (stream): int32 natom, int32 periodic, float64 rcut, float64 cell(3,3) (column k =
!                 lattice vector k), float64 posin(3,natom), float64 magin(3,natom)
!   out (stream): int32 natom, natombox, nei; int32 nrangestart(natom), nrangestop(natom),
!                 nlist(nei), origin(natombox), nshift(3,natombox);
!                 float64 dist(nei), eist(3,nei), dist_mom(nei), eist_mom(3,nei),
!                 dist_momi(natom), eist_momi(3,natom), pos(3,natombox)
program neilist_ref
    use neilistglobal
    implicit none
    character(len=4096) :: fin, fout
    integer :: iper, nei, u
    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u) natom, iper, rcut
    read(u) cell
    allocate(posin(3, natom), magin(3, natom))
    read(u) posin, magin
    close(u)
    periodic = (iper /= 0)
    nresize = 0
    call neilist(.false., .false.)
    nei = 0
    if (natom > 0) nei = nrangestop(natom)
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) natom, natombox, nei
    write(u) nrangestart(1:natom), nrangestop(1:natom), nlist(1:nei), origin(1:natombox), &
             nshift(1:3, 1:natombox)
    write(u) dist(1:nei), eist(1:3, 1:nei), dist_mom(1:nei), eist_mom(1:3, 1:nei), &
             dist_momi(1:natom), eist_momi(1:3, 1:natom), pos(1:3, 1:natombox)
    close(u)
end program neilist_ref
