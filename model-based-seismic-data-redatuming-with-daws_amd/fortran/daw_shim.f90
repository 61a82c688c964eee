! Drop-in Fortran layer over libdaw.so (include/daw.h) for the reference's
! driver (main_redatum_DAW.f90, sub_redatum_DAW.f90).
!
! It provides, with the reference's names and argument lists:
!   * module kwsinc        (kwsinc.f90): irad, nalpha, allocatekws2d,
!                          kwsinc2d, deallocatekws2d, subsourcekws
!   * module minresModule  (minresModule.f90): filmartin2d
!   * external subroutines submodeling, submodeling_adjoint_esource,
!     submodeling_adjoint_esource_data (submodeling2.f90),
!     subwavefieldresample (subwavefieldresample.f90)
! sublinearminres, whose arguments are the driver's grid_type/inv_type, lives
! in daw_shim_minres.f90 and is compiled against the driver's own module
! `parameters`.  Every libdaw error becomes `error stop` with its message,
! the reference's way of failing.
module daw_capi
  use iso_c_binding
  implicit none
  interface
     function daw_last_error() bind(C, name='daw_last_error') result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function daw_last_error
     integer(c_int) function daw_kwsinc2d(ofs) bind(C, name='daw_kwsinc2d')
       import :: c_int
       integer(c_int), value :: ofs
     end function daw_kwsinc2d
     integer(c_int) function daw_deallocatekws2d() bind(C, name='daw_deallocatekws2d')
       import :: c_int
     end function daw_deallocatekws2d
     integer(c_int) function daw_subsourcekws(z, x, h, ofs, ix, ialpha, nfs) bind(C, name='daw_subsourcekws')
       import :: c_int, c_float
       real(c_float), value :: z, x, h
       integer(c_int), value :: ofs
       integer(c_int) :: ix(2), ialpha(2), nfs
     end function daw_subsourcekws
     integer(c_int) function daw_submodeling(vp, n1, n2, npml, h, ofs, nt, dt, s, type_src, type_rec, is5, ir, nr, &
          sis, wavefield) bind(C, name='daw_submodeling')
       import :: c_int, c_float, c_ptr
       real(c_float) :: vp(*), s(*), sis(*)
       integer(c_int), value :: n1, n2, npml, ofs, nt, type_src, type_rec, nr
       real(c_float), value :: h, dt
       integer(c_int) :: is5(5), ir(*)
       type(c_ptr), value :: wavefield
     end function daw_submodeling
     integer(c_int) function daw_submodeling_adjoint_esource(vp, n1, n2, npml, h, ofs, nt, dt, res, type_rec, nr, &
          ir, sis, wavefieldp) bind(C, name='daw_submodeling_adjoint_esource')
       import :: c_int, c_float, c_ptr
       real(c_float) :: vp(*), res(*), sis(*)
       integer(c_int), value :: n1, n2, npml, ofs, nt, type_rec, nr
       real(c_float), value :: h, dt
       integer(c_int) :: ir(*)
       type(c_ptr), value :: wavefieldp
     end function daw_submodeling_adjoint_esource
     integer(c_int) function daw_submodeling_adjoint_esource_data(vp, n1, n2, npml, h, ofs, nt, dt, res, type_rec, &
          nr, ir, sis) bind(C, name='daw_submodeling_adjoint_esource_data')
       import :: c_int, c_float
       real(c_float) :: vp(*), res(*), sis(*)
       integer(c_int), value :: n1, n2, npml, ofs, nt, type_rec, nr
       real(c_float), value :: h, dt
       integer(c_int) :: ir(*)
     end function daw_submodeling_adjoint_esource_data
     integer(c_int) function daw_subwavefieldresample(n1, n2, npml, nt, dt, ir, nr, sis, wavefield) &
          bind(C, name='daw_subwavefieldresample')
       import :: c_int, c_float
       integer(c_int), value :: n1, n2, npml, nt, nr
       real(c_float), value :: dt
       integer(c_int) :: ir(*)
       real(c_float) :: sis(*), wavefield(*)
     end function daw_subwavefieldresample
     integer(c_int) function daw_sublinearminres(c0, n1, n2, npml, h, ofs, nt, dt, type_rec, nrec, ir, mu, maxiter, &
          error, deltadr, deltade, hist, info) bind(C, name='daw_sublinearminres')
       import :: c_int, c_float, c_ptr
       real(c_float) :: c0(*), deltadr(*), deltade(*)
       integer(c_int), value :: n1, n2, npml, ofs, nt, type_rec, nrec
       real(c_float), value :: h, dt, mu, maxiter, error
       integer(c_int) :: ir(*)
       type(c_ptr), value :: hist, info
     end function daw_sublinearminres
     integer(c_int) function daw_sublinearminres_log(c0, n1, n2, npml, h, ofs, nt, dt, type_rec, nrec, ir, mu, &
          maxiter, error, deltadr, deltade, log_path) bind(C, name='daw_sublinearminres_log')
       import :: c_int, c_float, c_char
       real(c_float) :: c0(*), deltadr(*), deltade(*)
       integer(c_int), value :: n1, n2, npml, ofs, nt, type_rec, nrec
       real(c_float), value :: h, dt, mu, maxiter, error
       integer(c_int) :: ir(*)
       character(kind=c_char) :: log_path(*)
     end function daw_sublinearminres_log
     integer(c_int) function daw_subresample(xin, xout, n1in, n2, n1out, d1in, d1out) bind(C, name='daw_subresample')
       import :: c_int, c_float
       real(c_float) :: xin(*), xout(*)
       integer(c_int), value :: n1in, n2, n1out
       real(c_float), value :: d1in, d1out
     end function daw_subresample
     integer(c_int) function daw_filmartin2d(din, nr, nt, dt, flow, fhig, llow, lhig, dout) bind(C, name='daw_filmartin2d')
       import :: c_int, c_float
       real(c_float) :: din(*), dout(*)
       integer(c_int), value :: nr, nt
       real(c_float), value :: dt, flow, fhig, llow, lhig
     end function daw_filmartin2d
  end interface
contains
  subroutine daw_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (rc == 0) return
    text = ' '
    call c_f_pointer(daw_last_error(), msg, [512])
    do i = 1, 512
       if (msg(i) == c_null_char) exit
       text(i:i) = msg(i)
    end do
    write(*, '(a,a,a,i0,a,a)') 'libdaw: ', where, ' failed (', rc, '): ', trim(text)
    error stop 1
  end subroutine daw_check
end module daw_capi

! kwsinc.f90:22-637 -- the table itself stays inside libdaw (per-point tiles)
module kwsinc
  use daw_capi
  implicit none
  integer, parameter :: irad(1:3) = 10
  integer, parameter :: nalpha(1:3) = 100
contains
  subroutine allocatekws2d
  end subroutine allocatekws2d

  subroutine kwsinc2d(ifreesurf)
    integer :: ifreesurf
    call daw_check(daw_kwsinc2d(int(ifreesurf, c_int)), 'kwsinc2d')
  end subroutine kwsinc2d

  subroutine deallocatekws2d
    call daw_check(daw_deallocatekws2d(), 'deallocatekws2d')
  end subroutine deallocatekws2d

  subroutine subsourcekws(ndim, xs, dz, ifs, ix, ialpha, nalpha_frees)
    integer, intent(in) :: ndim, ifs
    real, intent(in) :: xs(ndim), dz
    integer, intent(out) :: ix(ndim), ialpha(ndim), nalpha_frees
    integer(c_int) :: cix(2), cia(2), cnf
    if (ndim /= 2) then
       write(*, *) 'libdaw: subsourcekws supports ndim = 2 (the 2-D path)'
       error stop 1
    end if
    call daw_check(daw_subsourcekws(real(xs(1), c_float), real(xs(2), c_float), real(dz, c_float), &
         int(ifs, c_int), cix, cia, cnf), 'subsourcekws')
    ix = cix
    ialpha = cia
    nalpha_frees = cnf
  end subroutine subsourcekws
end module kwsinc

! minresModule.f90:486-614 (filmartin2d); fhig/lhig are updated in place when
! fhig exceeds the Nyquist frequency, as the reference does (:523-527)
module minresModule
  use daw_capi
  implicit none
contains
  subroutine filmartin2d(deltadeo, nr, nt, dt, flow, fhig, llow, lhig, deltade)
    integer :: nr, nt
    real :: deltadeo(nt, nr), deltade(nt, nr)
    real :: dt, flow, fhig, llow, lhig
    real :: fmax
    fmax = (1. / dt) / 2
    if (fhig > fmax) then
       fhig = fmax
       lhig = 0.
    end if
    call daw_check(daw_filmartin2d(deltadeo, int(nr, c_int), int(nt, c_int), real(dt, c_float), real(flow, c_float), &
         real(fhig, c_float), real(llow, c_float), real(lhig, c_float), deltade), 'filmartin2d')
  end subroutine filmartin2d
end module minresModule

! submodeling2.f90:11 -- is0 is unused by the reference as well
subroutine submodeling(is0, vp, n1, n2, npml, h, ofs, nt, dt, s, type_src, type_rec, is, ir, nr, sis, wavefield)
  use daw_capi
  implicit none
  integer :: is0, n1, n2, npml, ofs, nt, type_src, type_rec, nr
  real :: h, dt
  real, target :: vp(n1, n2), s(nt), sis(nt, nr), wavefield(n1, n2, nt)
  integer :: is(5), ir(nr, 5)
  call daw_check(daw_submodeling(vp, n1, n2, npml, h, ofs, nt, dt, s, type_src, type_rec, is, ir, nr, sis, &
       c_loc(wavefield)), 'submodeling')
end subroutine submodeling

! submodeling2.f90:523
subroutine submodeling_adjoint_esource(is, vp, n1, n2, npml, h, ofs, nt, dt, res, type_rec, nr, ir, sis, wavefieldp)
  use daw_capi
  implicit none
  integer :: is, n1, n2, npml, ofs, nt, type_rec, nr
  real :: h, dt
  real, target :: vp(n1, n2), res(nt, nr), sis(nt, nr), wavefieldp(n1, n2, nt)
  integer :: ir(nr, 5)
  call daw_check(daw_submodeling_adjoint_esource(vp, n1, n2, npml, h, ofs, nt, dt, res, type_rec, nr, ir, sis, &
       c_loc(wavefieldp)), 'submodeling_adjoint_esource')
end subroutine submodeling_adjoint_esource

! submodeling2.f90:880
subroutine submodeling_adjoint_esource_data(is, vp, n1, n2, npml, h, ofs, nt, dt, res, type_rec, nr, ir, sis)
  use daw_capi
  implicit none
  integer :: is, n1, n2, npml, ofs, nt, type_rec, nr
  real :: h, dt
  real :: vp(n1, n2), res(nt, nr), sis(nt, nr)
  integer :: ir(nr, 5)
  call daw_check(daw_submodeling_adjoint_esource_data(vp, n1, n2, npml, h, ofs, nt, dt, res, type_rec, nr, ir, sis), &
       'submodeling_adjoint_esource_data')
end subroutine submodeling_adjoint_esource_data

! subwavefieldresample.f90:1
subroutine subwavefieldresample(n1, n2, npml, nt, dt, ir, nr, sis, wavefield)
  use daw_capi
  implicit none
  integer :: n1, n2, npml, nt, nr
  real :: dt
  integer :: ir(nr, 5)
  real :: wavefield(n1, n2, nt), sis(nt, nr)
  call daw_check(daw_subwavefieldresample(n1, n2, npml, nt, dt, ir, nr, sis, wavefield), 'subwavefieldresample')
end subroutine subwavefieldresample

! subresample_hicks.f90:25 (the time resampler the reference ships without
! linking; a driver that adds it gets this one)
subroutine subresample(xin, xout, n1in, n2, n1out, d1in, d1out)
  use daw_capi
  implicit none
  integer :: n1in, n2, n1out
  real :: d1in, d1out
  real :: xin(n1in, n2), xout(n1out, n2)
  call daw_check(daw_subresample(xin, xout, n1in, n2, n1out, d1in, d1out), 'subresample')
end subroutine subresample
