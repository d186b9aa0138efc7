! dropin_driver.F90 -- a caller written the way the reference's own driver calls its solver: the
! external procedures multigrid, relax, calcdefp, relaxing, nsmooth, scalar2gradient and defp2delta
! with implicit interfaces, default REAL arrays, every argument by reference (mmh/main.f90:95-164,
! mmh/limiter.f90:408).  TEST INFRASTRUCTURE.  The same source is linked twice:
!   - against the reference's objects (oracle/build_ref.sh, -DNG=<n>): oracle/_ref/dropin_ref_ng<n>
!   - against libmmh_fortran.so, the GPU drop-ins (tests/dropin/build.sh, -DMMH_GPU):
!     tests/dropin/_bin/dropin_gpu_ng<n>
! The only difference between the two is the drop-ins' one extra call, mmh_gpu_init(NG), which
! gives the library the grid size the reference fixes at compile time.
!
!   dropin_<ref|gpu>_ng<n> <rho.bin> <def.bin> <out.bin>
! rho, def: raw FP32 stream files of NG^3 values.  out.bin receives, in this order:
!   relax        phi (NG^3), restricted residual ((NG/2)^3), rinf (1)      relax(..., 8, 19)
!   multigrid    phi (NG^3), projected rhs (NG^3)                         one V-cycle
!   outer loop   def, defp, u(1,:,:,:) (NG^3 each) after 2 iterations of calcdefp / def += defp /
!                relaxing from def = 0
!   spectral     scalar2gradient (3 NG^3) and defp2delta (NG^3) of that def
program dropin_driver
#ifdef MMH_GPU
  use mmh_gpu, only: mmh_gpu_init
#endif
  implicit none
  integer, parameter :: ng = NG, nfluid = 5
  real, allocatable :: u(:, :, :, :), def(:, :, :), defp(:, :, :), tmp(:, :, :), tmp2(:, :, :)
  real, allocatable :: rho(:, :, :), dmesh(:, :, :), phi(:, :, :), rhs(:, :, :), resid(:, :, :)
  real, allocatable :: grad(:, :, :, :), delta(:, :, :)
  real :: rinf, gpad(3), compressmax
  integer :: nn(3), iter, fu
  character(512) :: frho, fdef, fout
  ! COMMON /globalpa/ of the reference (mmh/globalpa.f90): compressmax is its fourth word
  common /globalpa/ gpad, compressmax

  call get_command_argument(1, frho)
  call get_command_argument(2, fdef)
  call get_command_argument(3, fout)
  allocate(u(nfluid, ng, ng, ng), def(ng, ng, ng), defp(ng, ng, ng), tmp(ng, ng, ng), tmp2(ng + 2, ng, ng))
  allocate(rho(ng, ng, ng), dmesh(ng, ng, ng), phi(ng, ng, ng), rhs(ng, ng, ng), resid(ng / 2, ng / 2, ng / 2))
  allocate(grad(ng, ng, ng, 3), delta(ng, ng, ng))
  open(newunit=fu, file=frho, access='stream', form='unformatted', status='old')
  read(fu) rho
  close(fu)
  open(newunit=fu, file=fdef, access='stream', form='unformatted', status='old')
  read(fu) dmesh
  close(fu)
  compressmax = 10
#ifdef MMH_GPU
  call mmh_gpu_init(ng, 10d0)
#endif
  u = 0
  u(1, :, :, :) = rho

  open(newunit=fu, file=fout, access='stream', form='unformatted', status='replace')
  ! relax on a deformed mesh: 8 sweeps, then the restricted residual and its max-norm
  rhs = (1 - rho) / 30
  call nsmooth(rhs)
  phi = 0
  resid = 0
  call relax(phi, resid, rhs, dmesh, u, rinf, 1., ng, ng, ng, nfluid, 8, 19)
  write(fu) phi, resid, rinf
  ! one multigrid V-cycle on the same system (rhs projected in place)
  phi = 0
  call multigrid(phi, rhs, dmesh, u, 1., ng, ng, ng, nfluid, 1, 2)
  write(fu) phi, rhs
  ! the reconstruction iteration (mmh/main.f90:89-110)
  def = 0
  do iter = 1, 2
    call calcdefp(defp, tmp, tmp2, def, u, 1., 1., 5)
    def = def + defp
    call relaxing(u, def, defp, 1., 1., 5)
  end do
  write(fu) def, defp, u(1, :, :, :)
  nn = ng
  call scalar2gradient(def, grad, nn)
  call defp2delta(def, delta, nn)
  write(fu) grad, delta
  close(fu)
end program dropin_driver
