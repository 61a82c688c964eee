! Reads a parameter file or an acquisition file with Fortran list-directed
! input, one READ per record group as the reference's readers do
! (subreadpar.f90:10-20; main_redatum_DAW.f90:121, then subacqui.f90:29-57 on
! the reopened file),
! and prints what it read, so that the library's C++ reader can be checked
! against a Fortran runtime.  Output: one value per line; reals as ES17.9
! (enough digits to recover a REAL exactly); "ERROR <iostat>" on a failed READ.
!
!   list_input par FILE | list_input acqui FILE
program list_input
  implicit none
  character(len=16) :: what
  character(len=512) :: path
  character(len=160) :: a, b
  integer :: ios, i1, i2, i3, i4, i, nsrc, nrec
  real :: r1, r2, r3, r4
  call get_command_argument(1, what)
  call get_command_argument(2, path)
  open(10, file=path, status='old')
  if (what == 'par') then
     read(10, *, iostat=ios) i1
     call chk(ios); print '(a,i0)', 'mode ', i1
     a = ''; b = ''
     read(10, *, iostat=ios) a, b
     call chk(ios); print '(a,a)', 'name_true ', trim(a); print '(a,a)', 'name_init ', trim(b)
     a = ''
     read(10, *, iostat=ios) a
     call chk(ios); print '(a,a)', 'name_acqui ', trim(a)
     a = ''
     read(10, *, iostat=ios) a
     call chk(ios); print '(a,a)', 'name_src ', trim(a)
     r1 = -7; r2 = -7; r3 = -7; r4 = -7
     read(10, *, iostat=ios) r1, r2, r3, r4
     call chk(ios)
     print '(a,es17.9)', 'flow ', r1; print '(a,es17.9)', 'fhig ', r2
     print '(a,es17.9)', 'llow ', r3; print '(a,es17.9)', 'lhig ', r4
     i1 = -7; i2 = -7; i3 = -7; i4 = -7; r1 = -7; r2 = -7; r3 = -7
     read(10, *, iostat=ios) i1, i2, i3, r1, i4, r2, r3
     call chk(ios)
     print '(a,i0)', 'n1 ', i1; print '(a,i0)', 'n2 ', i2; print '(a,i0)', 'npml ', i3
     print '(a,es17.9)', 'h ', r1; print '(a,i0)', 'nt ', i4
     print '(a,es17.9)', 'dt ', r2; print '(a,es17.9)', 'dto ', r3
     i1 = -7; i2 = -7
     read(10, *, iostat=ios) i1, i2
     call chk(ios); print '(a,i0)', 'type_src ', i1; print '(a,i0)', 'type_rec ', i2
     i1 = -7
     read(10, *, iostat=ios) i1
     call chk(ios); print '(a,i0)', 'ofs ', i1
     r1 = -7
     read(10, *, iostat=ios) r1
     call chk(ios); print '(a,es17.9)', 'mu ', r1
     i1 = -7
     read(10, *, iostat=ios) i1
     call chk(ios); print '(a,i0)', 'oio ', i1
     r1 = -7; r2 = -7
     read(10, *, iostat=ios) r1, r2
     call chk(ios); print '(a,es17.9)', 'error ', r1; print '(a,es17.9)', 'maxiter ', r2
  else
     read(10, *, iostat=ios) nsrc, nrec
     call chk(ios)
     print '(a,i0,1x,i0)', 'counts ', nsrc, nrec
     ! the points are read after reopening the file and skipping one record
     close(10)
     open(10, file=path, status='old')
     read(10, *)
     do i = 1, nsrc + nrec
        read(10, *, iostat=ios) r1, r2
        call chk(ios)
        print '(a,es17.9,1x,es17.9)', 'zx ', r1, r2
     end do
  end if
  close(10)
contains
  subroutine chk(ios)
    integer, intent(in) :: ios
    if (ios /= 0) then
       print '(a,i0)', 'ERROR ', ios
       stop
    end if
  end subroutine chk
end program list_input
