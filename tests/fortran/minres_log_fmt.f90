! Writes a MINRES log with Fortran's own formatted output, from values read
! on standard input, so that the library's C++ writer of the same log can be
! checked against a Fortran runtime.  The record layouts are those MINRES
! writes (minresModule.f90: heading :66-67, iteration lines :322-324, final
! status :337-342, formats :346-358); every value is a default REAL as there.
!
! stdin: n itnlim rtol shift / Anorm Acond rnorm ynorm Arnorm / istop itn /
!        then per iteration: itn x1 qrnorm_over_beta1 Anorm Acond printed
program minres_log_fmt
  implicit none
  integer :: n, itnlim, istop, itn, k, kk, printed
  real :: rtol, shift, anorm, acond, rnorm, ynorm, arnorm, x1, qr, an, ac
  character(len=256) :: path
  character(len=51) :: stopmsg(-1:8)
  stopmsg(-1) = 'beta2 = 0.  If M = I, b and x are eigenvectors of A'
  stopmsg(0) = 'beta1 = 0.  The exact solution is  x = 0'
  stopmsg(1) = 'Requested accuracy achieved, as determined by rtol'
  stopmsg(2) = 'Reasonable accuracy achieved, given eps'
  stopmsg(3) = 'x has converged to an eigenvector'
  stopmsg(4) = 'Acond has exceeded 0.1/eps'
  stopmsg(5) = 'The iteration limit was reached'
  stopmsg(6) = 'Aprod  does not define a symmetric matrix'
  stopmsg(7) = 'Msolve does not define a symmetric matrix'
  stopmsg(8) = 'Msolve does not define a pos-def preconditioner'
  call get_command_argument(1, path)
  read(*,*) n, itnlim, rtol, shift
  read(*,*) anorm, acond, rnorm, ynorm, arnorm
  read(*,*) istop, itn
  open(20, file=path, status='replace')
  write(20, 10) ' Enter MINRES.  ', n, .false., .false., itnlim, rtol, shift
  do k = 1, itn
     read(*,*) kk, x1, qr, an, ac, printed
     if (printed /= 0) then
        if (kk == 1) write(20, 20)
        write(20, 30) kk, x1, qr, an, ac
        if (mod(kk, 10) == 0) write(20, 40)
     end if
  end do
  write(20, 50) ' Exit  MINRES.  ', istop, itn, ' Exit  MINRES.  ', anorm, acond, &
       ' Exit  MINRES.  ', rnorm, ynorm, arnorm
  write(20, 60) ' Exit  MINRES.  ', stopmsg(istop)
  close(20)
10 format(// 1p, a, 5x, 'Solution of symmetric   Ax = b' / ' n      =', i7, 5x, 'checkA =', l4, 12x, &
          'precon =', l4 / ' itnlim =', i7, 5x, 'rtol   =', e11.2, 5x, 'shift  =', e23.14)
20 format(// 5x, 'itn', 8x, 'x(1)', 10x, 'norm(r)', 3x, 'norm(A)', 3x, 'cond(A)')
30 format(1p, i8, e19.10, 3e10.2)
40 format(1x)
50 format(/ 1p, a, 5x, 'istop =', i3, 14x, 'itn   =', i8 / a, 5x, 'Anorm =', e12.4, 5x, 'Acond =', e12.4 &
          / a, 5x, 'rnorm =', e12.4, 5x, 'ynorm =', e12.4, 5x, 'Arnorml=', e12.4)
60 format(a, 5x, a)
end program minres_log_fmt
