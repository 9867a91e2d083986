! he4_tddft_gpu.f90 -- Fortran host loop of the He-4 TDDFT + Bohmian impurity
! time integration on the GPU library.
!
! Follows the reference driver: first potential evaluation and energy
! (DFT4He3d.f90:630-653), zeroed particle histories (:841-845), then per
! iteration steprk for the first three iterations and steppc after, a
! potential evaluation, the particle means of units 81-83 and the energy
! every pener iterations (:849-885).
!
! Usage:  he4_tddft_gpu [namelist-file]      (default: namelist on stdin)
!
! Namelist /input/ uses the reference's names where they exist:
!   nx ny nz xmax ymax zmax N_par mimpur ironx core4 l lambdah tzmean tzsurf
!   nq deltat deltatps niter pener r_cutoff_gs umax_gs
! plus the cored Morse parameters d_morse re_morse alpha_morse, the precision
! (0 FP64, 1 mixed), tune_steps (> 0: run the launch-shape autotuner with that
! many PC iterations per candidate after the start; results are unchanged)
! and the files:
!   filein   stream file: psi(nx,ny,nz) complex(8), rimp(N_par,3), vimp(N_par,3)
!   fileout  stream file: psi, rimp, vimp, aimp after the last iteration
!            (with dry_run: fvlj4, wcgk, kalfs, vq from hetddft_make_kernels)
!   filelog  text: "means iter time errHe errimp errvimp <r> <v> <a>" per
!            iteration, "energy iter <12 terms>" at start and every pener,
!            and "diag iter <23 values of hetddft_diagnostics>" every pener
!            (He count and centre-of-mass moments, chemical potential, total
!            current, impurity mean-square radii, particle min/max: the
!            driver's output of DFT4He3d.f90:879-930)
program he4_tddft_gpu
  use, intrinsic :: iso_c_binding
  use hetddft_mod
  implicit none

  integer(c_int) :: nx, ny, nz, n_par, ironx, nq, precision
  real(c_double) :: xmax, ymax, zmax, mimpur, l, lambdah, tzmean, tzsurf
  real(c_double) :: deltat, deltatps, r_cutoff_gs, umax_gs, d_morse, re_morse, alpha_morse
  real(c_double) :: time0
  character(len=3) :: core4
  integer :: niter, pener, tune_steps
  logical :: dry_run
  character(len=512) :: filein, fileout, filelog

  namelist /input/ nx, ny, nz, xmax, ymax, zmax, n_par, mimpur, ironx, core4, l, &
                   lambdah, tzmean, tzsurf, nq, deltat, deltatps, niter, pener, &
                   r_cutoff_gs, umax_gs, d_morse, re_morse, alpha_morse, precision, &
                   time0, filein, fileout, filelog, dry_run, tune_steps

  type(hetddft_params) :: p
  type(hetddft_morse) :: morse
  type(hetddft_tuning) :: tuning
  type(hetddft_droplet) :: droplet
  type(c_ptr) :: ctx
  complex(c_double_complex), allocatable :: psi(:,:,:)
  real(c_double), allocatable :: rimp(:,:), vimp(:,:), aimp(:,:)
  real(c_double), allocatable :: fvlj4(:,:,:), wcgk(:,:,:), kalfs(:,:,:), vq(:,:,:)
  real(c_double) :: errHe, errimp, errvimp, means(9), e(12), diag(23), time
  integer :: iter, narg, u, ulog, ios
  character(len=512) :: nmlfile

  ! defaults: the library's Orsay-Trento constants, a small box
  ctx = c_null_ptr
  call hetddft_check(hetddft_default_params(p), ctx, 'hetddft_default_params')
  nx = p%nx; ny = p%ny; nz = p%nz
  xmax = p%xmax; ymax = p%ymax; zmax = p%zmax
  n_par = p%n_par; mimpur = p%mimpur; ironx = p%ironx
  core4 = 'OTC'; if (p%core_otc == 0) core4 = 'OT '
  l = p%l; lambdah = p%lambdah; tzmean = p%tzmean; tzsurf = p%tzsurf
  precision = p%precision
  nq = 1500; deltat = 0.d0; deltatps = 5.d-4; niter = 10; pener = 5; time0 = 0.d0
  d_morse = 20.d0; re_morse = 5.d0; alpha_morse = 0.8d0; r_cutoff_gs = 2.d0; umax_gs = 0.d0
  filein = 'state.in'; fileout = 'state.out'; filelog = 'steps.log'; dry_run = .false.
  tune_steps = 0

  narg = command_argument_count()
  if (narg >= 1) then
    call get_command_argument(1, nmlfile)
    open(newunit=u, file=trim(nmlfile), status='old', action='read')
    read(u, nml=input)
    close(u)
  else
    read(*, nml=input)
  end if

  ! time step in K^-1 from picoseconds (DFT4He3d.f90:230-245)
  if (deltat == 0.d0) then
    if (deltatps == 0.d0) then
      write(0, *) 'You must specify either Deltat (in kelvin) or Deltatps (in picosecond)'
      error stop 1
    end if
    deltat = deltatps / 7.63822291d0
  else
    deltatps = deltat * 7.63822291d0
  end if

  p%nx = nx; p%ny = ny; p%nz = nz
  p%xmax = xmax; p%ymax = ymax; p%zmax = zmax
  p%n_par = n_par; p%mimpur = mimpur; p%ironx = ironx
  p%core_otc = 0; if (core4 == 'OTC') p%core_otc = 1
  p%l = l; p%lambdah = lambdah; p%tzmean = tzmean; p%tzsurf = tzsurf
  p%precision = precision
  morse%d = d_morse; morse%re = re_morse; morse%alpha = alpha_morse
  morse%rc = r_cutoff_gs; morse%umax = umax_gs; morse%nq = nq

  ! k-space kernels: fforma, initcg, kalfs, FT_V_spline (DFT4He3d.f90:517-570)
  allocate(fvlj4(nx/2+1, ny, nz), wcgk(nx/2+1, ny, nz), kalfs(nx/2+1, ny, nz), vq(nx/2+1, ny, nz))
  call hetddft_check(hetddft_make_kernels(p, morse, fvlj4, wcgk, kalfs, vq), ctx, &
                     'hetddft_make_kernels')

  if (dry_run) then
    write(*, '("hetddft_params bytes ",i0,", hetddft_morse bytes ",i0,", hetddft_tuning bytes ",i0, &
              & ", hetddft_droplet bytes ",i0)') c_sizeof(p), c_sizeof(morse), c_sizeof(tuning), c_sizeof(droplet)
    open(newunit=u, file=trim(fileout), access='stream', form='unformatted', status='replace')
    write(u) fvlj4, wcgk, kalfs, vq
    close(u)
    stop
  end if

  allocate(psi(nx, ny, nz), rimp(n_par, 3), vimp(n_par, 3), aimp(n_par, 3))
  open(newunit=u, file=trim(filein), access='stream', form='unformatted', status='old', &
       action='read', iostat=ios)
  if (ios /= 0) then
    write(0, '("cannot open ",a)') trim(filein)
    error stop 1
  end if
  read(u) psi, rimp, vimp
  close(u)

  call hetddft_check(hetddft_create(p, ctx), c_null_ptr, 'hetddft_create')
  call hetddft_check(hetddft_set_kernels(ctx, fvlj4, wcgk, kalfs, vq), ctx, 'hetddft_set_kernels')
  call hetddft_check(hetddft_set_state(ctx, psi, rimp, vimp), ctx, 'hetddft_set_state')

  open(newunit=ulog, file=trim(filelog), status='replace', action='write')

  ! den = conjg(psi)*psi; potenimp; poten; forceimp; zeroed histories
  call hetddft_check(hetddft_start(ctx), ctx, 'hetddft_start')
  if (tune_steps > 0) call hetddft_check(hetddft_autotune(ctx, deltat, int(tune_steps, c_int), tuning), ctx, &
                                         'hetddft_autotune')
  call hetddft_check(hetddft_particle_means(ctx, means), ctx, 'hetddft_particle_means')
  write(ulog, '("means ",i8,13es26.17e3)') 0, time0, 0.d0, 0.d0, 0.d0, means
  call gpu_energy(ctx, e)
  write(ulog, '("energy ",i8,12es26.17e3)') 0, e

  do iter = 1, niter
    if (iter <= 3) then
      call gpu_steprk(ctx, deltat)
      errHe = 0.d0; errimp = 0.d0; errvimp = 0.d0
    else
      call gpu_steppc(ctx, deltat, errHe, errimp, errvimp)
    end if
    time = time0 + iter * deltatps
    call gpu_veval(ctx)
    call hetddft_check(hetddft_particle_means(ctx, means), ctx, 'hetddft_particle_means')
    write(ulog, '("means ",i8,13es26.17e3)') iter, time, errHe, errimp, errvimp, means
    if (mod(iter, pener) == 0) then
      call gpu_energy(ctx, e)
      write(ulog, '("energy ",i8,12es26.17e3)') iter, e
      call hetddft_check(hetddft_diagnostics(ctx, diag), ctx, 'hetddft_diagnostics')
      write(ulog, '("diag ",i8,23es26.17e3)') iter, diag
    end if
  end do
  close(ulog)

  call hetddft_check(hetddft_get_state(ctx, psi, rimp, vimp, aimp), ctx, 'hetddft_get_state')
  open(newunit=u, file=trim(fileout), access='stream', form='unformatted', status='replace')
  write(u) psi, rimp, vimp, aimp
  close(u)
  call hetddft_check(hetddft_destroy(ctx), c_null_ptr, 'hetddft_destroy')
end program he4_tddft_gpu
