! hetddft_mod.f90 -- ISO_C_BINDING interface of libhetddft (include/hetddft.h)
! for Fortran hosts of the He-4 Orsay-Trento TDDFT solver.
!
! The reference driver calls steprk(deltat), steppc(deltat,errHe,errimp,errvimp)
! (steprk.f90:1, steppc.f90:1) and potenimp(); poten(); forceimp()
! (potenimp.f90:1,24, poten.f90:5) on module globals.  With this module a host
! keeps its own arrays, hands them to a device context once, and replaces each
! of those calls by the gpu_* routine of the same meaning below.  Arrays are
! passed unchanged: psi(nx,ny,nz) complex(8), real fields (nx,ny,nz),
! k-space kernels (nx/2+1,ny,nz), particles (N_par,3).
!
! Every hetddft_* function returns 0 or a negative error code; the gpu_*
! wrappers stop the program with the library's message on failure, as the
! reference stops on its own errors.
module hetddft_mod
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: HETDDFT_OK = 0
  integer(c_int), parameter, public :: HETDDFT_E_ARG = -1
  integer(c_int), parameter, public :: HETDDFT_E_HIP = -2
  integer(c_int), parameter, public :: HETDDFT_E_NOMEM = -3
  integer(c_int), parameter, public :: HETDDFT_E_STATE = -4
  integer(c_int), parameter, public :: HETDDFT_E_NONFINITE = -5
  integer(c_int), parameter, public :: HETDDFT_FP64 = 0
  integer(c_int), parameter, public :: HETDDFT_MIXED = 1

  ! field ids of hetddft_get_field
  integer(c_int), parameter, public :: HETDDFT_F_DEN = 0, HETDDFT_F_POT4 = 1, &
       HETDDFT_F_UIMP = 2, HETDDFT_F_DENCG = 3, HETDDFT_F_DELJ4 = 4, HETDDFT_F_FALFS = 5, &
       HETDDFT_F_DXDEN = 6, HETDDFT_F_DYDEN = 7, HETDDFT_F_DZDEN = 8, &
       HETDDFT_F_INTXALF = 9, HETDDFT_F_INTYALF = 10, HETDDFT_F_INTZALF = 11, &
       HETDDFT_F_DENX = 12, HETDDFT_F_PSIX = 13, HETDDFT_F_HPSIX = 14, &
       HETDDFT_F_FXCONV = 15, HETDDFT_F_FYCONV = 16, HETDDFT_F_FZCONV = 17, &
       HETDDFT_F_CORREL = 18, HETDDFT_F_UALPHAS_CONV = 19

  ! Mirrors struct hetddft_params; names follow namelist /input/ and modules.f90.
  type, bind(C), public :: hetddft_params
    integer(c_int) :: nx, ny, nz
    real(c_double) :: xmax, ymax, zmax
    integer(c_int) :: n_par
    real(c_double) :: mimpur
    integer(c_int) :: ironx
    integer(c_int) :: core_otc
    real(c_double) :: lambdah, tzmean, tzsurf
    real(c_double) :: cp4, cpp4, alphas, l, den0s, h2o2m4, c, beta, den_m, h4, eps4, sigma4
    integer(c_int) :: precision
  end type hetddft_params

  ! Mirrors struct hetddft_morse (cored Morse He-impurity potential).
  type, bind(C), public :: hetddft_morse
    real(c_double) :: d, re, alpha, rc, umax
    integer(c_int) :: nq
  end type hetddft_morse

  ! Mirrors struct hetddft_tuning (launch shapes; index 1..5 = HETDDFT_TUNE_GRAD..RK + 1).
  type, bind(C), public :: hetddft_tuning
    integer(c_int) :: tile(5), zc(5), prefetch(5)
    integer(c_int) :: fft_z_tile, fft_z_wgs
    real(c_double) :: ms_per_step
  end type hetddft_tuning

  ! Mirrors struct hetddft_droplet (fresh start built on the device; the
  ! seed is a 64-bit unsigned value in C, passed through c_long_long).
  type, bind(C), public :: hetddft_droplet
    real(c_double) :: n_he, rho0, r0, a, cavity, cavity_w, ripple
    real(c_double) :: clm(24), centre(3), lamda(3)
    real(c_double) :: phase_noise, sigma, v0(3)
    integer(c_long_long) :: seed
  end type hetddft_droplet

  public :: hetddft_default_params, hetddft_create, hetddft_destroy, hetddft_last_error, &
            hetddft_make_kernels, hetddft_set_kernels, hetddft_set_state, hetddft_start, &
            hetddft_potential, hetddft_step_rk, hetddft_step_pc, hetddft_last_errors, &
            hetddft_iterate, hetddft_particle_means, hetddft_energy, hetddft_get_state, &
            hetddft_get_field, hetddft_convolve, hetddft_profile, hetddft_profile_read, &
            hetddft_sync, hetddft_device_bytes, hetddft_graph_info, hetddft_diagnostics, &
            hetddft_current, hetddft_checkpoint_size, hetddft_checkpoint_save, hetddft_checkpoint_load, &
            hetddft_sample_particles, hetddft_init_droplet, hetddft_init_kernels, hetddft_get_kernels, &
            hetddft_autotune, hetddft_get_tuning, hetddft_set_tuning, &
            hetddft_batch_create, hetddft_batch_iterate, hetddft_batch_last_error, hetddft_batch_destroy
  public :: hetddft_message, hetddft_check, gpu_steprk, gpu_steppc, gpu_veval, gpu_energy

  interface
    integer(c_int) function hetddft_default_params(p) bind(C, name="hetddft_default_params")
      import :: c_int, hetddft_params
      type(hetddft_params), intent(out) :: p
    end function

    integer(c_int) function hetddft_create(p, ctx) bind(C, name="hetddft_create")
      import :: c_int, c_ptr, hetddft_params
      type(hetddft_params), intent(in) :: p
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function hetddft_destroy(ctx) bind(C, name="hetddft_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function hetddft_last_error(ctx) bind(C, name="hetddft_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function hetddft_make_kernels(p, morse, fvlj4, wcgk, kalfs, vq) &
        bind(C, name="hetddft_make_kernels")
      import :: c_int, c_double, hetddft_params, hetddft_morse
      type(hetddft_params), intent(in) :: p
      type(hetddft_morse), intent(in) :: morse
      real(c_double), intent(out) :: fvlj4(*), wcgk(*), kalfs(*), vq(*)
    end function

    integer(c_int) function hetddft_set_kernels(ctx, fvlj4, wcgk, kalfs, vq) &
        bind(C, name="hetddft_set_kernels")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: fvlj4(*), wcgk(*), kalfs(*), vq(*)
    end function

    integer(c_int) function hetddft_set_state(ctx, psi, rimp, vimp) bind(C, name="hetddft_set_state")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: psi(*)
      real(c_double), intent(in) :: rimp(*), vimp(*)
    end function

    integer(c_int) function hetddft_start(ctx) bind(C, name="hetddft_start")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function hetddft_potential(ctx) bind(C, name="hetddft_potential")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function hetddft_step_rk(ctx, deltat) bind(C, name="hetddft_step_rk")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat
    end function

    integer(c_int) function hetddft_step_pc(ctx, deltat, err) bind(C, name="hetddft_step_pc")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat
      real(c_double), intent(out) :: err(3)
    end function

    integer(c_int) function hetddft_last_errors(ctx, err) bind(C, name="hetddft_last_errors")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: err(3)
    end function

    integer(c_int) function hetddft_iterate(ctx, deltat, err) bind(C, name="hetddft_iterate")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat
      real(c_double), intent(out) :: err(3)
    end function

    integer(c_int) function hetddft_particle_means(ctx, out) bind(C, name="hetddft_particle_means")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(9)
    end function

    integer(c_int) function hetddft_energy(ctx, out) bind(C, name="hetddft_energy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(12)
    end function

    integer(c_int) function hetddft_get_state(ctx, psi, rimp, vimp, aimp) bind(C, name="hetddft_get_state")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(out) :: psi(*)
      real(c_double), intent(out) :: rimp(*), vimp(*), aimp(*)
    end function

    integer(c_int) function hetddft_get_field(ctx, which, out) bind(C, name="hetddft_get_field")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: which
      real(c_double), intent(out) :: out(*)
    end function

    ! in(i), out(i) are c_loc() of the i-th real field; kern is a half-grid kernel.
    integer(c_int) function hetddft_convolve(ctx, nf, in, kern, out) bind(C, name="hetddft_convolve")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nf
      type(c_ptr), intent(in) :: in(*)
      real(c_double), intent(in) :: kern(*)
      type(c_ptr), intent(in) :: out(*)
    end function

    integer(c_int) function hetddft_profile(ctx, enable) bind(C, name="hetddft_profile")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
    end function

    integer(c_int) function hetddft_profile_read(ctx, cls, ms_total, launches, bytes_total) &
        bind(C, name="hetddft_profile_read")
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: ctx
      integer(c_int), value :: cls
      real(c_double), intent(out) :: ms_total
      integer(c_long_long), intent(out) :: launches
      real(c_double), intent(out) :: bytes_total
    end function

    integer(c_int) function hetddft_sync(ctx) bind(C, name="hetddft_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function hetddft_diagnostics(ctx, out) bind(C, name="hetddft_diagnostics")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(23)
    end function

    integer(c_int) function hetddft_init_kernels(ctx, morse) bind(C, name="hetddft_init_kernels")
      import :: c_int, c_ptr, hetddft_morse
      type(c_ptr), value :: ctx
      type(hetddft_morse), intent(in) :: morse
    end function

    integer(c_int) function hetddft_get_kernels(ctx, fvlj4, wcgk, kalfs, vq) bind(C, name="hetddft_get_kernels")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: fvlj4(*), wcgk(*), kalfs(*), vq(*)
    end function

    integer(c_int) function hetddft_sample_particles(ctx, denx, lo, hi, seed, max_attempts, rimp) &
        bind(C, name="hetddft_sample_particles")
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: denx(*), lo(3), hi(3)
      integer(c_long_long), value :: seed, max_attempts
      real(c_double), intent(out) :: rimp(*)
    end function

    integer(c_int) function hetddft_init_droplet(ctx, d) bind(C, name="hetddft_init_droplet")
      import :: c_int, c_ptr, hetddft_droplet
      type(c_ptr), value :: ctx
      type(hetddft_droplet), intent(in) :: d
    end function

    integer(c_int) function hetddft_checkpoint_size(ctx, bytes) bind(C, name="hetddft_checkpoint_size")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx
      integer(c_long_long), intent(out) :: bytes
    end function

    integer(c_int) function hetddft_checkpoint_save(ctx, buf) bind(C, name="hetddft_checkpoint_save")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, buf
    end function

    integer(c_int) function hetddft_checkpoint_load(ctx, buf) bind(C, name="hetddft_checkpoint_load")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, buf
    end function

    integer(c_int) function hetddft_current(ctx, jx, jy, jz) bind(C, name="hetddft_current")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: jx(*), jy(*), jz(*)
    end function

    integer(c_int) function hetddft_graph_info(ctx, count, enabled) bind(C, name="hetddft_graph_info")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: count, enabled
    end function

    integer(c_long_long) function hetddft_device_bytes(ctx) bind(C, name="hetddft_device_bytes")
      import :: c_long_long, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function hetddft_autotune(ctx, deltat, steps, out) bind(C, name="hetddft_autotune")
      import :: c_int, c_ptr, c_double, hetddft_tuning
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat
      integer(c_int), value :: steps
      type(hetddft_tuning), intent(out) :: out
    end function

    integer(c_int) function hetddft_get_tuning(ctx, out) bind(C, name="hetddft_get_tuning")
      import :: c_int, c_ptr, hetddft_tuning
      type(c_ptr), value :: ctx
      type(hetddft_tuning), intent(out) :: out
    end function

    integer(c_int) function hetddft_set_tuning(ctx, t) bind(C, name="hetddft_set_tuning")
      import :: c_int, c_ptr, hetddft_tuning
      type(c_ptr), value :: ctx
      type(hetddft_tuning), intent(in) :: t
    end function

    ! ctxs(i) are the instances' contexts; deltat(i) their time steps.
    integer(c_int) function hetddft_batch_create(ctxs, n, nthreads, batch) bind(C, name="hetddft_batch_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(in) :: ctxs(*)
      integer(c_int), value :: n, nthreads
      type(c_ptr), intent(out) :: batch
    end function

    integer(c_int) function hetddft_batch_iterate(batch, deltat, steps, err, failed) &
        bind(C, name="hetddft_batch_iterate")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: batch
      real(c_double), intent(in) :: deltat(*)
      integer(c_int), value :: steps
      real(c_double), intent(out) :: err(3, *)
      integer(c_int), intent(out) :: failed
    end function

    type(c_ptr) function hetddft_batch_last_error(batch) bind(C, name="hetddft_batch_last_error")
      import :: c_ptr
      type(c_ptr), value :: batch
    end function

    integer(c_int) function hetddft_batch_destroy(batch) bind(C, name="hetddft_batch_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function

    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! Message of the last failure on ctx (c_null_ptr: of the last hetddft_create).
  function hetddft_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: s
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    s = hetddft_last_error(ctx)
    if (.not. c_associated(s)) then
      msg = ''
      return
    end if
    n = int(c_strlen(s))
    call c_f_pointer(s, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function hetddft_message

  ! Stops with the library's message when rc is an error code.
  subroutine hetddft_check(rc, ctx, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: what
    if (rc /= HETDDFT_OK) then
      write(0, '("hetddft: ",a," failed (",i0,"): ",a)') what, rc, hetddft_message(ctx)
      error stop 2
    end if
  end subroutine hetddft_check

  ! call steprk(deltat)
  subroutine gpu_steprk(ctx, deltat)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(in) :: deltat
    call hetddft_check(hetddft_step_rk(ctx, deltat), ctx, 'steprk')
  end subroutine gpu_steprk

  ! call steppc(deltat, errHe, errimp, errvimp)
  subroutine gpu_steppc(ctx, deltat, errHe, errimp, errvimp)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(in) :: deltat
    real(c_double), intent(out) :: errHe, errimp, errvimp
    real(c_double) :: err(3)
    call hetddft_check(hetddft_step_pc(ctx, deltat, err), ctx, 'steppc')
    errHe = err(1); errimp = err(2); errvimp = err(3)
  end subroutine gpu_steppc

  ! call potenimp(); call poten(); call forceimp()
  subroutine gpu_veval(ctx)
    type(c_ptr), intent(in) :: ctx
    call hetddft_check(hetddft_potential(ctx), ctx, 'potenimp/poten/forceimp')
  end subroutine gpu_veval

  ! call energy(); e = (etot, etot4, ekin4, elj4, ealphas, esolid, ecor4,
  !                     eimpu, ekinx, ekinqx, eHeX, norm)
  subroutine gpu_energy(ctx, e)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out) :: e(12)
    call hetddft_check(hetddft_energy(ctx, e), ctx, 'energy')
  end subroutine gpu_energy

end module hetddft_mod
