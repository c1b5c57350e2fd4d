! Independent Fortran check of the restart record layout used by
! rsdft_develop_amd/restart.py (wf.dat1 and vrho.dat1 of RSDFT, one grid
! rank, real bands).  Mode "read": reads the two files in the record
! sequence of the reference's writers and prints a summary.  Mode "write":
! writes both files for a 4x3x2 grid with known values.
program restart_io
  implicit none
  character(256) :: mode, fwf, fvr
  integer :: ML, ML1, ML2, ML3, MB, MB1, MB2, MK, MS, MMBZ, IOC, OC, TW
  integer :: npar(0:7), two(2), nel, nat, i, n
  integer, allocatable :: LL(:,:), ki(:), zn(:)
  real(8), allocatable :: occ(:), u(:), rho(:), vl(:), vh(:), vx(:), pos(:,:)
  real(8) :: aa(3,3), bb(3,3), kbb(3,1), ef
  character(64) :: ver, cmt

  call get_command_argument(1, mode)
  call get_command_argument(2, fwf)
  call get_command_argument(3, fvr)

  if (trim(mode) == "read") then
    open(10, file=trim(fwf), form="unformatted", status="old")
    read(10) i
    read(10) ML, ML1, ML2, ML3
    read(10) MB, MB1, MB2
    read(10) MK, MS, MMBZ
    read(10) IOC, OC, TW
    allocate(LL(3, ML), occ(MB), u(ML))
    read(10) LL
    read(10) occ
    read(10) aa, bb, kbb
    read(10) npar
    read(10) two
    read(10) two
    read(10) two
    read(10) two
    print '(a,i0,1x,i0,1x,i0,1x,i0,1x,i0,1x,i0,1x,i0)', "wf ", i, ML, ML1, ML2, ML3, MB, MB2 - MB1 + 1
    print '(a,3i6)', "LL last ", LL(:, ML)
    print '(a,es24.16)', "occ sum ", sum(occ)
    print '(a,es24.16)', "bb11 ", bb(1,1)
    do n = MB1, MB2
      read(10) u
      print '(a,i0,1x,es24.16)', "band ", n, sum(u * (/ (dble(i), i = 1, ML) /))
    end do
    close(10)
    open(11, file=trim(fvr), form="unformatted", status="old")
    read(11) ver
    read(11) ML, ML1, ML2, ML3
    deallocate(LL)
    allocate(LL(3, ML), rho(ML), vl(ML), vh(ML), vx(ML))
    read(11) LL
    read(11) cmt
    read(11) aa
    read(11) cmt
    read(11) nel, nat
    allocate(pos(3, nat), ki(nat), zn(nel))
    read(11) pos
    read(11) ki
    read(11) zn
    read(11) cmt
    read(11) ef
    read(11) rho
    read(11) vl
    read(11) vh
    read(11) vx
    close(11)
    print '(a,a)', "version ", trim(ver)
    print '(a,i0,1x,i0,1x,i0,1x,i0)', "atoms ", nel, nat, ki(nat), zn(1)
    print '(a,es24.16)', "efermi ", ef
    print '(a,4es24.16)', "fields ", sum(rho), sum(vl), sum(vh), sum(vx)
  else
    ML1 = 4; ML2 = 3; ML3 = 2; ML = ML1 * ML2 * ML3; MB = 3
    allocate(LL(3, ML), occ(MB), u(ML))
    i = 0
    do n = 0, ML3 - 1
      do MK = 0, ML2 - 1
        do MS = 0, ML1 - 1
          i = i + 1
          LL(:, i) = (/ MS, MK, n /)
        end do
      end do
    end do
    occ = (/ 2.d0, 1.5d0, 0.25d0 /)
    aa = 0.d0; aa(1,1) = 4.d0; aa(2,2) = 3.d0; aa(3,3) = 2.d0
    bb = 0.d0; kbb = 0.d0
    npar = 1
    open(10, file=trim(fwf), form="unformatted", status="replace")
    write(10) -1
    write(10) ML, ML1, ML2, ML3
    write(10) MB, 1, MB
    write(10) 1, 1, 1
    write(10) 0, 3, 1
    write(10) LL
    write(10) occ
    write(10) aa, bb, kbb
    write(10) npar
    write(10) 0, ML
    write(10) 0, MB
    write(10) 0, 1
    write(10) 0, 1
    do n = 1, MB
      u = (/ (dble(i) + 100.d0 * n, i = 1, ML) /)
      write(10) u
    end do
    close(10)
  end if
end program restart_io
