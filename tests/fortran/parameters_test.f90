! Test-side module `parameters` carrying the grid_type / inv_type fields the
! reference's driver defines (main_redatum_DAW.f90:10-35), so the shim's
! sublinearminres can be linked and exercised without the driver.
module parameters
  type grid_type
     integer :: mode
     integer :: n1, n2, npml, nt, nto
     real :: h, dt, t0, dto, fhig, flow, llow, lhig
     real :: peakfreq, maxfreq, mu, maxiter, error
     real, allocatable :: ctrue(:, :), c0(:, :), s(:)
     integer :: type_src, type_rec
     integer :: ofs
     integer :: oio
     integer :: nsrc, nrec
     real, allocatable :: xs(:, :), xr(:, :)
     integer, allocatable :: is(:, :), ir(:, :)
  end type grid_type
  type inv_type
     real, allocatable :: dataobs(:, :), datar(:, :), deltadr(:, :)
     real, allocatable :: deltade(:, :), datae(:, :), dataesc(:, :)
     real, allocatable :: dataescf(:, :), dataobsf(:, :)
  end type inv_type
end module parameters
