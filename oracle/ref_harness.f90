! Test harness over the reference's own Fortran routines.
!
! TEST INFRASTRUCTURE ONLY.  This file is ours; it is compiled together with
! the reference sources (read in place from the reference tree by
! oracle/build_ref.sh) into oracle/_ref/libdawref.so.  It exposes C entry
! points (bind(C)) for the module procedures and derived-type interfaces that
! ctypes cannot reach directly:
!   kwsinc module (kwsinc.f90:94-179, 212-378, 582-606),
!   sublinearminres / MINRES (sublinearminres.f90:7-86, minresModule.f90:10-360),
!   filmartin2d (minresModule.f90:486-614), REALFT2S (libfourier.f90:199-276),
!   subacqui (subacqui.f90:12-62).
! The external subroutines submodeling_, submodeling_adjoint_esource_,
! submodeling_adjoint_esource_data_, subwavefieldresample_ and subresiduals_
! are called straight from C (all arguments by reference).
module dawref_harness
  use iso_c_binding
  use kwsinc
  use parameters
  use minresModule
  implicit none

contains

  subroutine dawref_kwsinc_init(ofs) bind(C, name='dawref_kwsinc_init')
    integer(c_int), value :: ofs
    if (allocated(coef_hicks2D_monopole)) call deallocatekws2d
    call allocatekws2d
    call kwsinc2d(int(ofs))
  end subroutine dawref_kwsinc_init

  subroutine dawref_kwsinc_free() bind(C, name='dawref_kwsinc_free')
    if (allocated(coef_hicks2D_monopole)) call deallocatekws2d
  end subroutine dawref_kwsinc_free

  ! 1-D monopole table, out(n+11, ialpha), n = -10..10, ialpha = 1..100
  subroutine dawref_hicks1d(out) bind(C, name='dawref_hicks1d')
    real(c_float), intent(out) :: out(21, 100)
    call allocatekws1d
    call kwsinc1d
    out = coef_hicks1D_monopole
    call deallocatekws1d
  end subroutine dawref_hicks1d

  ! one 21x21 tile (i1 fastest) of coef_hicks2D_monopole(:,:,ia1,ia2,ip)
  subroutine dawref_hicks2d(ia1, ia2, ip, out) bind(C, name='dawref_hicks2d')
    integer(c_int), value :: ia1, ia2, ip
    real(c_float), intent(out) :: out(21, 21)
    out = coef_hicks2D_monopole(:, :, ia1, ia2, ip)
  end subroutine dawref_hicks2d

  subroutine dawref_bessi0(x, y) bind(C, name='dawref_bessi0')
    real(c_float), value :: x
    real(c_float), intent(out) :: y
    y = BESSI0(x)
  end subroutine dawref_bessi0

  subroutine dawref_subsourcekws(z, x, h, ofs, ix, ialpha, nfs) bind(C, name='dawref_subsourcekws')
    real(c_float), value :: z, x, h
    integer(c_int), value :: ofs
    integer(c_int), intent(out) :: ix(2), ialpha(2), nfs
    real :: xs(2)
    integer :: ix_(2), ia_(2), nf_
    xs(1) = z
    xs(2) = x
    call subsourcekws(2, xs, h, int(ofs), ix_, ia_, nf_)
    ix = ix_
    ialpha = ia_
    nfs = nf_
  end subroutine dawref_subsourcekws

  ! subacqui on the acquisition file 'fname' (<= 26 characters, cwd-relative)
  subroutine dawref_subacqui(fname, flen, h, npml, ofs, nsrc, nrec, src_tab, rec_tab) &
       bind(C, name='dawref_subacqui')
    character(kind=c_char), intent(in) :: fname(*)
    integer(c_int), value :: flen, npml, ofs, nsrc, nrec
    real(c_float), value :: h
    integer(c_int), intent(out) :: src_tab(nsrc, 5), rec_tab(nrec, 5)
    type(grid_type) :: grid
    character(len=26) :: name_acqui
    integer :: i
    integer, allocatable :: st(:, :), rt(:, :)
    name_acqui = ' '
    do i = 1, min(int(flen), 26)
       name_acqui(i:i) = fname(i)
    end do
    grid%h = h
    grid%npml = npml
    grid%ofs = ofs
    grid%nsrc = nsrc
    grid%nrec = nrec
    allocate(st(nsrc, 5), rt(nrec, 5))
    call subacqui(name_acqui, grid, st, rt)
    src_tab = st
    rec_tab = rt
  end subroutine dawref_subacqui

  ! sublinearminres with its grid/inv carriers filled from plain arguments
  subroutine dawref_linearminres(n1, n2, npml, h, nt, dt, ofs, type_rec, nrec, mu, maxiter, &
       c0, rec_tab, deltadr, deltade) bind(C, name='dawref_linearminres')
    integer(c_int), value :: n1, n2, npml, nt, ofs, type_rec, nrec, maxiter
    real(c_float), value :: h, dt, mu
    real(c_float), intent(in) :: c0(n1, n2), deltadr(nt, nrec)
    integer(c_int), intent(in) :: rec_tab(nrec, 5)
    real(c_float), intent(out) :: deltade(nt, nrec)
    type(grid_type) :: grid
    type(inv_type) :: inv
    integer, allocatable :: rt(:, :)
    grid%n1 = n1
    grid%n2 = n2
    grid%npml = npml
    grid%h = h
    grid%nt = nt
    grid%dt = dt
    grid%ofs = ofs
    grid%type_rec = type_rec
    grid%nrec = nrec
    grid%mu = mu
    grid%maxiter = real(maxiter)
    grid%error = 0.
    grid%oio = 0
    allocate(grid%c0(n1, n2))
    grid%c0 = c0
    allocate(inv%deltadr(nt, nrec), inv%deltade(nt, nrec))
    inv%deltadr = deltadr
    inv%deltade = 0.
    allocate(rt(nrec, 5))
    rt = rec_tab
    call sublinearminres(1, grid, inv, rt)
    deltade = inv%deltade
  end subroutine dawref_linearminres

  ! sublinearminres with grid%oio = 1: as dawref_linearminres, and the
  ! reference's MINRES writes its log MINRES_00001.txt into the working
  ! directory (sublinearminres.f90:51-61)
  subroutine dawref_linearminres_log(n1, n2, npml, h, nt, dt, ofs, type_rec, nrec, mu, maxiter, &
       c0, rec_tab, deltadr, deltade) bind(C, name='dawref_linearminres_log')
    integer(c_int), value :: n1, n2, npml, nt, ofs, type_rec, nrec, maxiter
    real(c_float), value :: h, dt, mu
    real(c_float), intent(in) :: c0(n1, n2), deltadr(nt, nrec)
    integer(c_int), intent(in) :: rec_tab(nrec, 5)
    real(c_float), intent(out) :: deltade(nt, nrec)
    type(grid_type) :: grid
    type(inv_type) :: inv
    integer, allocatable :: rt(:, :)
    grid%n1 = n1
    grid%n2 = n2
    grid%npml = npml
    grid%h = h
    grid%nt = nt
    grid%dt = dt
    grid%ofs = ofs
    grid%type_rec = type_rec
    grid%nrec = nrec
    grid%mu = mu
    grid%maxiter = real(maxiter)
    grid%error = 0.
    grid%oio = 1
    allocate(grid%c0(n1, n2))
    grid%c0 = c0
    allocate(inv%deltadr(nt, nrec), inv%deltade(nt, nrec))
    inv%deltadr = deltadr
    inv%deltade = 0.
    allocate(rt(nrec, 5))
    rt = rec_tab
    call sublinearminres(1, grid, inv, rt)
    deltade = inv%deltade
  end subroutine dawref_linearminres_log

  ! MINRES stopped after itnlim iterations, returning its scalar outputs; the
  ! shift is formed the way sublinearminres.f90:64-72 forms it.
  subroutine dawref_minres(n1, n2, npml, h, nt, dt, ofs, type_rec, nrec, mu, itnlim, &
       c0, rec_tab, deltadr, x, fac, istop, itn, scal) bind(C, name='dawref_minres')
    integer(c_int), value :: n1, n2, npml, nt, ofs, type_rec, nrec, itnlim
    real(c_float), value :: h, dt, mu
    real(c_float), intent(in) :: c0(n1, n2), deltadr(nt, nrec)
    integer(c_int), intent(in) :: rec_tab(nrec, 5)
    real(c_float), intent(out) :: x(nt * nrec), fac
    integer(c_int), intent(out) :: istop, itn
    real(c_float), intent(out) :: scal(5)   ! Anorm, Acond, rnorm, Arnorm, ynorm
    type(grid_type) :: grid
    type(inv_type) :: inv
    integer, allocatable :: rt(:, :)
    real, allocatable :: b(:), x0(:), xx(:)
    real :: nr_dr, nr_de, shift, Anorm, Acond, rnorm, Arnorm, ynorm
    integer :: nn, is_, it_
    grid%n1 = n1
    grid%n2 = n2
    grid%npml = npml
    grid%h = h
    grid%nt = nt
    grid%dt = dt
    grid%ofs = ofs
    grid%type_rec = type_rec
    grid%nrec = nrec
    grid%mu = mu
    grid%maxiter = real(itnlim)
    grid%error = 0.
    grid%oio = 0
    allocate(grid%c0(n1, n2))
    grid%c0 = c0
    allocate(inv%deltadr(nt, nrec), inv%deltade(nt, nrec))
    inv%deltadr = deltadr
    allocate(rt(nrec, 5))
    rt = rec_tab
    nn = nt * nrec
    allocate(b(nn), x0(nn), xx(nn))
    call sub2d21d(inv%deltadr, nrec, nt, nn, b)
    x0 = 0.
    call submodeling_adjoint_esource_data(1, grid%c0, n1, n2, npml, h, ofs, nt, dt, &
         inv%deltadr, type_rec, nrec, rt, inv%deltade)
    call datanorm(inv%deltadr, nt, nrec, nr_dr)
    call datanorm(inv%deltade, nt, nrec, nr_de)
    fac = nr_dr / nr_de * mu
    shift = -fac
    call MINRES(nn, b, shift, .false., .false., x0, xx, int(itnlim), -111, 0., &
         is_, it_, Anorm, Acond, rnorm, Arnorm, ynorm, 1, grid, inv, rt)
    x = xx
    istop = is_
    itn = it_
    scal(1) = Anorm
    scal(2) = Acond
    scal(3) = rnorm
    scal(4) = Arnorm
    scal(5) = ynorm
  end subroutine dawref_minres

  subroutine dawref_filmartin2d(din, nr, nt, dt, flow, fhig, llow, lhig, dout) &
       bind(C, name='dawref_filmartin2d')
    integer(c_int), value :: nr, nt
    real(c_float), value :: dt, flow, fhig, llow, lhig
    real(c_float), intent(in) :: din(nt, nr)
    real(c_float), intent(out) :: dout(nt, nr)
    real :: a, b, c, d
    real, allocatable :: tmp(:, :)
    a = flow
    b = fhig
    c = llow
    d = lhig
    allocate(tmp(nt, nr))
    tmp = din
    call filmartin2d(tmp, int(nr), int(nt), dt, a, b, c, d, dout)
  end subroutine dawref_filmartin2d

  subroutine dawref_realft2s(n, isign, dre, dim_) bind(C, name='dawref_realft2s')
    integer(c_int), value :: n, isign
    real(c_double), intent(inout) :: dre(n), dim_(n)
    real(8), allocatable :: buf(:)
    allocate(buf(n))
    call REALFT2S(int(n), int(isign), dre, dim_, buf)
  end subroutine dawref_realft2s

  ! subhicks1d (subresample_hicks.f90:87) on a zeroed table: it writes the
  ! entries inside the window's support and leaves the rest as they were
  subroutine dawref_subhicks1d(out) bind(C, name='dawref_subhicks1d')
    real(c_float), intent(out) :: out(-4:4, 10)
    external :: subhicks1d
    out = 0.
    call subhicks1d(out, 4, 10)
  end subroutine dawref_subhicks1d

  ! subresample (subresample_hicks.f90:25)
  subroutine dawref_subresample(xin, xout, n1in, n2, n1out, d1in, d1out) bind(C, name='dawref_subresample')
    integer(c_int), value :: n1in, n2, n1out
    real(c_float), value :: d1in, d1out
    real(c_float), intent(in) :: xin(n1in, n2)
    real(c_float), intent(out) :: xout(n1out, n2)
    external :: subresample
    real :: a, b
    a = d1in
    b = d1out
    call subresample(xin, xout, int(n1in), int(n2), int(n1out), a, b)
  end subroutine dawref_subresample

end module dawref_harness
