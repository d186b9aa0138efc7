! mmh_gpu.f90 -- iso_c_binding interface of the MI355X moving-mesh reconstruction library.
!
! Fortran host code reaches the HIP kernels through the C ABI declared in include/mmh.h.  This
! module binds those entry points and holds the settings the reference keeps at compile time or in
! COMMON blocks: the grid size NG (mmh/dimen.f90:1) and compressmax (mmh/globalpa.f90, set at
! mmh/main.f90:56).  The reference-named drop-ins live in mmh_legacy.f90 as external procedures so
! that their link names (multigrid_, relax_, calcdefp_, relaxing_, ...) are the reference's.
module mmh_gpu
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: MMH_OK = 0, MMH_EINVAL = 1, MMH_EDEVICE = 2, MMH_ENOMEM = 3, MMH_ENOCONV = 4

  integer :: mmh_ng = 0                 ! grid size used where the reference uses NG
  real(c_double) :: mmh_compressmax = 10.0d0

  interface
    function mmh_default_ctx() bind(C, name='mmh_default_ctx') result(c)
      import :: c_ptr
      type(c_ptr) :: c
    end function

    function mmh_error_string(ctx) bind(C, name='mmh_error_string') result(p)
      import :: c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr) :: p
    end function

    function mmh_relax_host_f(ctx, arr, resid, rhs, deformpot, u, nu, rinf, dx, n, nrelax, iopt) &
        bind(C, name='mmh_relax_host_f') result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: ctx
      real(c_float) :: arr(*), resid(*), rhs(*), deformpot(*), u(*), rinf
      integer(c_int), value :: nu, n, nrelax, iopt
      real(c_float), value :: dx
      integer(c_int) :: rc
    end function

    function mmh_relax_host_d(ctx, arr, resid, rhs, deformpot, u, nu, rinf, dx, n, nrelax, iopt) &
        bind(C, name='mmh_relax_host_d') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: arr(*), resid(*), rhs(*), deformpot(*), u(*), rinf
      integer(c_int), value :: nu, n, nrelax, iopt
      real(c_double), value :: dx
      integer(c_int) :: rc
    end function

    function mmh_multigrid_host_f(ctx, u, rhs, defpot, ubig, dx, n, nu, iopt, ng) &
        bind(C, name='mmh_multigrid_host_f') result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: ctx
      real(c_float) :: u(*), rhs(*), defpot(*), ubig(*)
      real(c_float), value :: dx
      integer(c_int), value :: n, nu, iopt, ng
      integer(c_int) :: rc
    end function

    function mmh_multigrid_host_d(ctx, u, rhs, defpot, ubig, dx, n, nu, iopt, ng) &
        bind(C, name='mmh_multigrid_host_d') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: u(*), rhs(*), defpot(*), ubig(*)
      real(c_double), value :: dx
      integer(c_int), value :: n, nu, iopt, ng
      integer(c_int) :: rc
    end function

    function mmh_calcdefp_host_f(ctx, defp, tmp, tmp2, def, u, dtold, dtaumesh, nfluid, n, compressmax) &
        bind(C, name='mmh_calcdefp_host_f') result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: ctx
      real(c_float) :: defp(*), tmp(*), tmp2(*), def(*), u(*)
      real(c_float), value :: dtold, dtaumesh, compressmax
      integer(c_int), value :: nfluid, n
      integer(c_int) :: rc
    end function

    function mmh_calcdefp_host_d(ctx, defp, tmp, tmp2, def, u, dtold, dtaumesh, nfluid, n, compressmax) &
        bind(C, name='mmh_calcdefp_host_d') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: defp(*), tmp(*), tmp2(*), def(*), u(*)
      real(c_double), value :: dtold, dtaumesh, compressmax
      integer(c_int), value :: nfluid, n
      integer(c_int) :: rc
    end function

    function mmh_relaxing_host_f(ctx, u, def, defp, dt, nfluid, n) bind(C, name='mmh_relaxing_host_f') result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: ctx
      real(c_float) :: u(*), def(*), defp(*)
      real(c_float), value :: dt
      integer(c_int), value :: nfluid, n
      integer(c_int) :: rc
    end function

    function mmh_relaxing_host_d(ctx, u, def, defp, dt, nfluid, n) bind(C, name='mmh_relaxing_host_d') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: u(*), def(*), defp(*)
      real(c_double), value :: dt
      integer(c_int), value :: nfluid, n
      integer(c_int) :: rc
    end function

    function mmh_gradient_host_f(ctx, phi, psi, n) bind(C, name='mmh_gradient_host_f') result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: ctx
      real(c_float) :: phi(*), psi(*)
      integer(c_int), value :: n
      integer(c_int) :: rc
    end function

    function mmh_delta_host_f(ctx, phi, delta, n) bind(C, name='mmh_delta_host_f') result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: ctx
      real(c_float) :: phi(*), delta(*)
      integer(c_int), value :: n
      integer(c_int) :: rc
    end function

    function mmh_nsmooth_host_f(ctx, a, n) bind(C, name='mmh_nsmooth_host_f') result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: ctx
      real(c_float) :: a(*)
      integer(c_int), value :: n
      integer(c_int) :: rc
    end function

    function mmh_lognormal_host_d(ctx, rho, n, sigma, slope, ks_frac, seed) &
        bind(C, name='mmh_lognormal_host_d') result(rc)
      import :: c_ptr, c_int, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double) :: rho(*)
      integer(c_int), value :: n
      real(c_double), value :: sigma, slope, ks_frac
      integer(c_int64_t), value :: seed
      integer(c_int) :: rc
    end function

    function mmh_reconstruct_host_f(ctx, u, def, defp, n, nfluid, niter, dtold, dtaumesh, dt, compressmax, diag) &
        bind(C, name='mmh_reconstruct_host_f') result(rc)
      import :: c_ptr, c_int, c_float, c_double
      type(c_ptr), value :: ctx
      real(c_float) :: u(*), def(*), defp(*)
      integer(c_int), value :: n, nfluid, niter
      real(c_float), value :: dtold, dtaumesh, dt, compressmax
      real(c_double) :: diag(*)
      integer(c_int) :: rc
    end function

    function c_strlen(s) bind(C, name='strlen') result(l)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: l
    end function
  end interface

contains

  ! Sets the grid size the reference fixes at compile time (NG) for the drop-ins whose argument
  ! lists carry no size (calcdefp, relaxing, nsmooth) and for multigrid's sweep schedule.
  subroutine mmh_gpu_init(ng, compressmax)
    integer, intent(in) :: ng
    real(c_double), intent(in), optional :: compressmax
    mmh_ng = ng
    if (present(compressmax)) mmh_compressmax = compressmax
  end subroutine

  function mmh_ctx() result(c)
    type(c_ptr) :: c
    c = mmh_default_ctx()
    if (.not. c_associated(c)) then
      write(*,*) 'mmh_gpu: no HIP device available'
      error stop 2
    end if
  end function

  ! The reference stops on invalid input (mmh/multigrid.f90:28-35); the drop-ins do the same with
  ! the library's message.
  subroutine mmh_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: l
    if (rc == MMH_OK) return
    p = mmh_error_string(mmh_default_ctx())
    l = int(c_strlen(p))
    call c_f_pointer(p, msg, [l])
    write(*,*) trim(where), ': ', msg(1:l)
    error stop 1
  end subroutine

  integer function mmh_need_ng(where) result(n)
    character(*), intent(in) :: where
    if (mmh_ng <= 0) then
      write(*,*) trim(where), ': call mmh_gpu_init(NG) first'
      error stop 1
    end if
    n = mmh_ng
  end function

end module mmh_gpu
