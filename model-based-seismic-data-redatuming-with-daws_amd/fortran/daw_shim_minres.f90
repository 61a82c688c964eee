! sublinearminres (sublinearminres.f90:7-86) over libdaw: the MINRES solve of
! (H + fac I) deltade = deltadr runs on the GPU with FP64 Lanczos scalars.
! Compiled against the driver's module `parameters` (main_redatum_DAW.f90:5-37),
! whose grid_type / inv_type fields it reads: n1, n2, npml, h, ofs, nt, dt,
! type_rec, nrec, mu, maxiter, error, c0 and inv%deltadr; it writes
! inv%deltade as the reference does (:82).  With grid%oio = 1 MINRES writes
! its log to MINRES_<istot>.txt as the reference's does (:51-61, 78-80); the
! name comes from the driver's subname (subname.f90).
subroutine sublinearminres(istot, grid, inv, ir)
  use parameters
  use daw_capi
  implicit none
  type(grid_type) :: grid
  type(inv_type) :: inv
  integer :: istot
  integer :: ir(grid%nrec, 5)
  character(len=160) :: namei
  if (grid%oio == 1) then
     call subname(istot, namei)
     call daw_check(daw_sublinearminres_log(grid%c0, grid%n1, grid%n2, grid%npml, grid%h, grid%ofs, grid%nt, &
          grid%dt, grid%type_rec, grid%nrec, ir, grid%mu, grid%maxiter, grid%error, inv%deltadr, inv%deltade, &
          'MINRES_'//trim(namei)//'.txt'//c_null_char), 'sublinearminres')
  else
     call daw_check(daw_sublinearminres(grid%c0, grid%n1, grid%n2, grid%npml, grid%h, grid%ofs, grid%nt, grid%dt, &
          grid%type_rec, grid%nrec, ir, grid%mu, grid%maxiter, grid%error, inv%deltadr, inv%deltade, c_null_ptr, &
          c_null_ptr), 'sublinearminres')
  end if
end subroutine sublinearminres
