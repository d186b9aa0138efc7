! mmh_legacy.f90 -- drop-in replacements for the reference's solver entry points.
!
! Each subroutine keeps the name, argument list and in/out behaviour of the reference routine it
! replaces, so an unmodified reference-style driver links against libmmh_fortran.so instead of the
! reference objects; the work runs on the GPU through the C ABI (module mmh_gpu).  Default REAL
! arguments as in the reference; the _f64 siblings take real(8).
!
!   multigrid        mmh/multigrid.f90:16      relax           mmh/gauss1.f90:64
!   calcdefp         mmh/limiter.f90:17        relaxing        mmh/relaxing.f90:21
!   nsmooth          mmh/limiter.f90:677       scalar2gradient mmh/gradient.f90:1
!   defp2delta       mmh/reconstruct.f90:1
!
! The reference takes NG from a compile-time header; here calcdefp, relaxing and nsmooth use the
! size set by mmh_gpu_init(NG), and multigrid uses it as the NG of its sweep schedule
! (nu = 4*NG/N per level) when it is larger than the call's grid.

subroutine multigrid(u, rhs, defpot, ubig, dx, nx, ny, nz, nu, nred, iopt)
  use mmh_gpu
  implicit none
  integer nx, ny, nz, iopt, nred, nu
  real u(nx, ny, nz), defpot(nx, ny, nz), rhs(nx, ny, nz), dx, ubig(nu, nx, ny, nz)
  if (nx /= ny .or. ny /= nz) then
    write(*,*) 'multigrid: only perfect cubes supported'
    error stop 1
  end if
  if (nred /= 1) then
    write(*,*) 'multigrid: only nred=1 supported, nred=', nred
    error stop 1
  end if
  call mmh_check(mmh_multigrid_host_f(mmh_ctx(), u, rhs, defpot, ubig, dx, nx, nu, iopt, max(mmh_ng, nx)), &
                 'multigrid')
end subroutine multigrid

subroutine multigrid_f64(u, rhs, defpot, ubig, dx, nx, ny, nz, nu, nred, iopt)
  use mmh_gpu
  implicit none
  integer nx, ny, nz, iopt, nred, nu
  real(8) u(nx, ny, nz), defpot(nx, ny, nz), rhs(nx, ny, nz), dx, ubig(nu, nx, ny, nz)
  if (nx /= ny .or. ny /= nz .or. nred /= 1) then
    write(*,*) 'multigrid_f64: only perfect cubes and nred=1 supported'
    error stop 1
  end if
  call mmh_check(mmh_multigrid_host_d(mmh_ctx(), u, rhs, defpot, ubig, dx, nx, nu, iopt, max(mmh_ng, nx)), &
                 'multigrid_f64')
end subroutine multigrid_f64

subroutine relax(arr, resid, rhs, deformpot, u, rinf, dx, na1, na2, na3, nu, nrelax, iopt)
  use mmh_gpu
  implicit none
  integer iopt, na1, na2, na3, nrelax, nu
  real arr(na1, na2, na3), resid((na1 + 1) / 2, (na2 + 1) / 2, (na3 + 1) / 2), rhs(na1, na2, na3), rinf, &
       u(nu, na1, na2, na3), deformpot(na1, na2, na3), dx
  if (na1 /= na2 .or. na2 /= na3) then
    write(*,*) 'relax: only perfect cubes supported'
    error stop 1
  end if
  call mmh_check(mmh_relax_host_f(mmh_ctx(), arr, resid, rhs, deformpot, u, nu, rinf, dx, na1, nrelax, iopt), 'relax')
end subroutine relax

subroutine relax_f64(arr, resid, rhs, deformpot, u, rinf, dx, na1, na2, na3, nu, nrelax, iopt)
  use mmh_gpu
  implicit none
  integer iopt, na1, na2, na3, nrelax, nu
  real(8) arr(na1, na2, na3), resid((na1 + 1) / 2, (na2 + 1) / 2, (na3 + 1) / 2), rhs(na1, na2, na3), rinf, &
          u(nu, na1, na2, na3), deformpot(na1, na2, na3), dx
  call mmh_check(mmh_relax_host_d(mmh_ctx(), arr, resid, rhs, deformpot, u, nu, rinf, dx, na1, nrelax, iopt), &
                 'relax_f64')
end subroutine relax_f64

subroutine calcdefp(defp, tmp, tmp2, def, u, dtold1, dtaumesh, nfluid)
  use mmh_gpu
  implicit none
  integer nfluid
  real defp(*), tmp(*), tmp2(*), def(*), u(*), dtold1, dtaumesh
  call mmh_check(mmh_calcdefp_host_f(mmh_ctx(), defp, tmp, tmp2, def, u, dtold1, dtaumesh, nfluid, &
                 mmh_need_ng('calcdefp'), real(mmh_compressmax)), 'calcdefp')
end subroutine calcdefp

subroutine calcdefp_f64(defp, tmp, tmp2, def, u, dtold1, dtaumesh, nfluid)
  use mmh_gpu
  implicit none
  integer nfluid
  real(8) defp(*), tmp(*), tmp2(*), def(*), u(*), dtold1, dtaumesh
  call mmh_check(mmh_calcdefp_host_d(mmh_ctx(), defp, tmp, tmp2, def, u, dtold1, dtaumesh, nfluid, &
                 mmh_need_ng('calcdefp_f64'), mmh_compressmax), 'calcdefp_f64')
end subroutine calcdefp_f64

subroutine relaxing(u, def, defp, cmax, dt, nfluidcmp)
  use mmh_gpu
  implicit none
  integer nfluidcmp
  real u(*), def(*), defp(*), cmax, dt
  call mmh_check(mmh_relaxing_host_f(mmh_ctx(), u, def, defp, dt, nfluidcmp, mmh_need_ng('relaxing')), 'relaxing')
end subroutine relaxing

subroutine relaxing_f64(u, def, defp, cmax, dt, nfluidcmp)
  use mmh_gpu
  implicit none
  integer nfluidcmp
  real(8) u(*), def(*), defp(*), cmax, dt
  call mmh_check(mmh_relaxing_host_d(mmh_ctx(), u, def, defp, dt, nfluidcmp, mmh_need_ng('relaxing_f64')), &
                 'relaxing_f64')
end subroutine relaxing_f64

subroutine nsmooth(a)
  use mmh_gpu
  implicit none
  real a(*)
  call mmh_check(mmh_nsmooth_host_f(mmh_ctx(), a, mmh_need_ng('nsmooth')), 'nsmooth')
end subroutine nsmooth

subroutine scalar2gradient(indata, outdata, nn)
  use mmh_gpu
  implicit none
  integer(4) :: nn(3)
  real(4) :: indata(nn(1), nn(2), nn(3)), outdata(nn(1), nn(2), nn(3), 3)
  if (nn(1) /= nn(2) .or. nn(2) /= nn(3)) then
    write(*,*) 'scalar2gradient: only perfect cubes supported'
    error stop 1
  end if
  call mmh_check(mmh_gradient_host_f(mmh_ctx(), indata, outdata, nn(1)), 'scalar2gradient')
end subroutine scalar2gradient

subroutine defp2delta(indata, outdata, nn)
  use mmh_gpu
  implicit none
  integer(4) :: nn(3)
  real(4) :: indata(nn(1), nn(2), nn(3)), outdata(nn(1), nn(2), nn(3))
  if (nn(1) /= nn(2) .or. nn(2) /= nn(3)) then
    write(*,*) 'defp2delta: only perfect cubes supported'
    error stop 1
  end if
  call mmh_check(mmh_delta_host_f(mmh_ctx(), indata, outdata, nn(1)), 'defp2delta')
end subroutine defp2delta

! C-callable setter for hosts that cannot use the module (e.g. a ctypes harness).
subroutine mmh_set_grid(ng, compressmax) bind(C, name='mmh_set_grid')
  use mmh_gpu
  implicit none
  integer(c_int), value :: ng
  real(c_double), value :: compressmax
  call mmh_gpu_init(int(ng), compressmax)
end subroutine mmh_set_grid
