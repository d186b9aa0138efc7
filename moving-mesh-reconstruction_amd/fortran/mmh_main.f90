! mmh_main.f90 -- reconstruction driver: density file in; deformation potential, displacement and
! reconstructed density files out, with the solve on the GPU.
!
! Reads the run description from a namelist file (first command-line argument):
!
!   &mmhrun workdir='./', ng=64, niter=300, idstring='_2' /
!
! and keeps the reference's on-disk conventions (mmh/main.f90:60-65, 120-179, 184-218): raw
! headerless stream files of default REAL in Fortran order, N^3 values each, named
!   input     <workdir><zstring><inf><NG:i4.4><idstring><prefix>
!   outputs   <workdir><dir><zstring><stem><idstring><NG:i4.4><prefix>
!             stem def (dir fndef), psix/psiy/psiz (fndisp), recon (fndelta)
!   restart   <workdir><dir><zstring>restartu<step:i4.4>-<idstring><NG:i4.4><prefix>  (fnu, density)
!             <workdir><dir><zstring>restartdef<step:i4.4>-<idstring><NG:i4.4><prefix> (fndef)
! With restart = .true. the iteration resumes after restartstep from those files; with
! flag_check = .true. they are written for step niter at the end.  dry_run = .true. only prints the
! files a run would read and write.  The grid size is a run-time parameter here.
program mmh_main
  use iso_c_binding
  use mmh_gpu
  implicit none
  character(512) :: workdir = './'
  character(16) :: zstring = '0.000'
  character(64) :: inf = 'den', idstring = '_2', prefix = '.bin'
  character(64) :: fnu = '', fndef = '', fndelta = '', fndisp = ''
  integer :: ng = 64, niter = 300, restartstep = 150
  logical :: restart = .false., flag_def = .true., flag_gradient = .true., flag_density = .true.
  logical :: flag_check = .true., dry_run = .false.
  real :: compressmax = 10.
  namelist /mmhrun/ workdir, zstring, inf, idstring, prefix, fnu, fndef, fndelta, fndisp, ng, niter, &
    restartstep, restart, flag_def, flag_gradient, flag_density, flag_check, dry_run, compressmax

  character(512) :: arg, fn
  character(4) :: gstring, sstring
  character(16) :: ngdigits
  integer :: u_nml, ios, begin, nrun, it, i
  real, allocatable :: u(:, :, :, :), def(:, :, :), defp(:, :, :), work(:, :, :), grad(:, :, :, :)
  real(8), allocatable :: diag(:, :)
  real(8) :: umean, t0, t1

  if (command_argument_count() < 1) then
    write(*,*) 'usage: mmh_main <namelist file with &mmhrun ... />'
    error stop 1
  end if
  call get_command_argument(1, arg)
  open(newunit=u_nml, file=trim(arg), status='old', action='read', iostat=ios)
  if (ios /= 0) then
    write(*,*) 'mmh_main: cannot open ', trim(arg)
    error stop 1
  end if
  read(u_nml, nml=mmhrun, iostat=ios)
  close(u_nml)
  if (ios /= 0) then
    write(*,*) 'mmh_main: bad namelist &mmhrun in ', trim(arg)
    error stop 1
  end if
  if (ng < 16 .or. iand(ng, ng - 1) /= 0) then
    write(*,*) 'mmh_main: ng must be a power of two >= 16, ng=', ng
    error stop 1
  end if
  write(gstring, '(i4.4)') ng
  write(ngdigits, '(i0)') ng
  if (len_trim(fnu) == 0) fnu = 'u' // trim(ngdigits) // '/'
  if (len_trim(fndef) == 0) fndef = 'def' // trim(ngdigits) // '/'
  if (len_trim(fndelta) == 0) fndelta = 'delta' // trim(ngdigits) // '/'
  if (len_trim(fndisp) == 0) fndisp = 'disp' // trim(ngdigits) // '/'
  write(*,*) 'grid:', ng
  begin = 1
  if (restart) begin = restartstep + 1
  nrun = max(0, niter - begin + 1)

  if (dry_run) then
    if (restart) then
      call show('reading', restartname(fnu, 'restartu', restartstep))
      call show('reading', restartname(fndef, 'restartdef', restartstep))
    else
      call show('reading', inputname())
    end if
    if (flag_check) then
      call show('writing', restartname(fnu, 'restartu', niter))
      call show('writing', restartname(fndef, 'restartdef', niter))
    end if
    if (flag_def) call show('writing', outname(fndef, 'def'))
    if (flag_gradient) then
      call show('writing', outname(fndisp, 'psix'))
      call show('writing', outname(fndisp, 'psiy'))
      call show('writing', outname(fndisp, 'psiz'))
    end if
    if (flag_density) call show('writing', outname(fndelta, 'recon'))
    write(*,*) 'iterations:', begin, niter
    stop
  end if

  allocate(u(5, ng, ng, ng), def(ng, ng, ng), defp(ng, ng, ng), work(ng, ng, ng))
  u = 0.
  def = 0.
  if (restart) then
    write(*,*) 'RESTART from step', restartstep
    call readfield(restartname(fnu, 'restartu', restartstep), work)
    u(1, :, :, :) = work
    call readfield(restartname(fndef, 'restartdef', restartstep), def)
  else
    call readfield(inputname(), work)
    u(1, :, :, :) = work
  end if
  umean = sum(real(u(1, :, :, :), 8)) / real(ng, 8)**3
  write(*,*) 'u mean=', real(umean)
  if (.not. restart .and. abs(umean - 1.d0) > 1.d-5) then
    ! the reference requires a mean density of 1 (mmh/main.f90:72-77)
    write(*,*) 'mmh_main: mean density must be 1 to 1e-5, got', umean
    error stop 1
  end if

  allocate(diag(7, max(nrun, 1)))
  call cpu_time(t0)
  call mmh_gpu_init(ng, real(compressmax, c_double))
  call mmh_check(mmh_reconstruct_host_f(mmh_ctx(), u, def, defp, ng, 5, nrun, 1., 1., 1., compressmax, diag), &
                 'reconstruct')
  call cpu_time(t1)
  do it = 1, nrun
    write(*, '(a,i6,a,es14.6,a,es14.6,a,es14.6,a,2es14.6,a,i0,a,i0)') ' iter=', begin + it - 1, &
      ' defp sigma=', diag(1, it), ' def sigma=', diag(2, it), ' u sigma=', diag(3, it), &
      ' u min max=', diag(4, it), diag(5, it), ' u<0: ', nint(diag(6, it)), ' ncompress: ', nint(diag(7, it))
  end do
  write(*,*) 'iterations done:', nrun, ' seconds:', t1 - t0

  if (flag_check) then
    work = u(1, :, :, :)
    call writefield(restartname(fnu, 'restartu', niter), work)
    call writefield(restartname(fndef, 'restartdef', niter), def)
  end if
  if (flag_def) call writefield(outname(fndef, 'def'), def)
  if (flag_gradient) then
    allocate(grad(ng, ng, ng, 3))
    call mmh_check(mmh_gradient_host_f(mmh_ctx(), def, grad, ng), 'scalar2gradient')
    call writefield(outname(fndisp, 'psix'), grad(:, :, :, 1))
    call writefield(outname(fndisp, 'psiy'), grad(:, :, :, 2))
    call writefield(outname(fndisp, 'psiz'), grad(:, :, :, 3))
  end if
  if (flag_density) then
    call mmh_check(mmh_delta_host_f(mmh_ctx(), def, work, ng), 'defp2delta')
    call writefield(outname(fndelta, 'recon'), work)
  end if
  write(*,*) 'END'

contains

  function inputname() result(f)
    character(1024) :: f
    f = trim(workdir) // trim(zstring) // trim(inf) // gstring // trim(idstring) // trim(prefix)
  end function

  function outname(dir, stem) result(f)
    character(*), intent(in) :: dir, stem
    character(1024) :: f
    f = trim(workdir) // trim(dir) // trim(zstring) // stem // trim(idstring) // gstring // trim(prefix)
  end function

  function restartname(dir, stem, step) result(f)
    character(*), intent(in) :: dir, stem
    integer, intent(in) :: step
    character(1024) :: f
    write(sstring, '(i4.4)') step
    f = trim(workdir) // trim(dir) // trim(zstring) // stem // sstring // '-' // trim(idstring) // gstring // &
        trim(prefix)
  end function

  subroutine show(what, f)
    character(*), intent(in) :: what, f
    write(*, '(a,": ",a)') what, trim(f)
  end subroutine

  subroutine readfield(f, a)
    character(*), intent(in) :: f
    real, intent(out) :: a(:, :, :)
    integer :: un
    call show('reading', f)
    open(newunit=un, file=trim(f), status='old', access='stream', form='unformatted', iostat=ios)
    if (ios /= 0) then
      write(*,*) 'mmh_main: cannot open ', trim(f)
      error stop 1
    end if
    read(un, iostat=ios) a
    close(un)
    if (ios /= 0) then
      write(*,*) 'mmh_main: short file ', trim(f)
      error stop 1
    end if
  end subroutine

  subroutine writefield(f, a)
    character(*), intent(in) :: f
    real, intent(in) :: a(:, :, :)
    integer :: un
    call show('writing', f)
    open(newunit=un, file=trim(f), status='replace', access='stream', form='unformatted', iostat=ios)
    if (ios /= 0) then
      write(*,*) 'mmh_main: cannot create ', trim(f)
      error stop 1
    end if
    write(un) a
    close(un)
  end subroutine

end program mmh_main
