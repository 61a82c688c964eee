! Calls the drop-in shim exactly as the reference's driver calls the
! reference routines (sub_redatum_DAW.f90:59-60, :82-83, :115, :131-134,
! :238-240; main_redatum_DAW.f90:92-93; subacqui.f90:32), reading its inputs
! from in.bin and writing every output to out.bin (stream access).
program shim_driver
  use parameters
  use kwsinc
  use minresModule
  implicit none
  integer :: hdr(7), n1, n2, npml, nt, nr, ndat, ofs, u, ix(2), ia(2), nf
  real :: fh(3), h, dt, mu, xs(2)
  real, allocatable :: vp(:, :), vp0(:, :), s(:), res(:, :), sis(:, :), wf(:, :, :), sis2(:, :), sis3(:, :)
  real, allocatable :: wfp(:, :, :), sisd(:, :), filt(:, :), rsmp(:, :)
  integer :: ntout
  integer, allocatable :: src(:), rec(:, :), dat(:, :)
  type(grid_type) :: grid
  type(inv_type) :: inv
  real :: fl, fhg, ll, lh

  open(newunit=u, file='in.bin', access='stream', form='unformatted', status='old')
  read(u) hdr
  read(u) fh
  n1 = hdr(1); n2 = hdr(2); npml = hdr(3); nt = hdr(4); nr = hdr(5); ndat = hdr(6); ofs = hdr(7)
  h = fh(1); dt = fh(2); mu = fh(3)
  allocate(vp(n1, n2), vp0(n1, n2), s(nt), src(5), rec(nr, 5), dat(ndat, 5), res(nt, nr))
  read(u) vp, vp0, s, src, rec, dat, res
  close(u)
  allocate(sis(nt, nr), wf(n1, n2, nt), sis2(nt, nr), sis3(nt, nr), wfp(n1, n2, nt), sisd(nt, ndat), filt(nt, nr))

  call allocatekws2d
  call kwsinc2d(ofs)
  xs = (/ 15.0, 123.4 /)
  call subsourcekws(2, xs, h, ofs, ix, ia, nf)
  call submodeling(1, vp, n1, n2, npml, h, ofs, nt, dt, s, 0, 0, src, rec, nr, sis, wf)
  call submodeling_adjoint_esource_data(1, vp0, n1, n2, npml, h, ofs, nt, dt, res, 0, nr, rec, sis2)
  call submodeling_adjoint_esource(1, vp0, n1, n2, npml, h, ofs, nt, dt, res, 0, nr, rec, sis3, wfp)
  call subwavefieldresample(n1, n2, npml, nt, dt, dat, ndat, sisd, wf)
  fl = 3.0; fhg = 30.0; ll = 2.0; lh = 10.0
  call filmartin2d(sis, nr, nt, dt, fl, fhg, ll, lh, filt)
  ! dt -> 2.5 dt, nto as sub_redatum_DAW.f90:24 sizes it
  ntout = nint(float(nt - 1) * dt / (2.5 * dt)) + 1
  allocate(rsmp(ntout, nr))
  call subresample(sis, rsmp, nt, nr, ntout, dt, 2.5 * dt)

  grid%n1 = n1; grid%n2 = n2; grid%npml = npml; grid%h = h; grid%nt = nt; grid%dt = dt
  grid%ofs = ofs; grid%type_rec = 0; grid%nrec = nr; grid%mu = mu; grid%maxiter = 4.; grid%error = 0.
  grid%oio = 1
  allocate(grid%c0(n1, n2), inv%deltadr(nt, nr), inv%deltade(nt, nr))
  grid%c0 = vp0
  inv%deltadr = res
  call sublinearminres(1, grid, inv, rec)
  call deallocatekws2d

  open(newunit=u, file='out.bin', access='stream', form='unformatted', status='replace')
  write(u) ix, ia, nf
  write(u) sis, wf, sis2, sis3, wfp, sisd, filt, inv%deltade, rsmp
  close(u)
end program shim_driver

! The driver's shot-name routine (subname.f90:2-20: five digits, zero padded),
! which the shim's sublinearminres calls for the MINRES log name.
subroutine subname(i, name)
  implicit none
  integer :: i
  character(len=160) :: name
  name = ' '
  write(name, '(i5.5)') i
end subroutine subname
