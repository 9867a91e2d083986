/*
 * hetddft.h -- C-ABI of the MI355X (gfx950) time-step library for the
 * Orsay-Trento He-4 TDDFT solver with Bohmian impurity test particles.
 *
 * Drop-in boundary for the per-step subroutines of the reference Fortran code
 * (bcntls2016/4hetddft-isotropic-bohmian).  The reference talks through module
 * globals; here the same state lives on the GPU inside an opaque context and
 * crosses the boundary only in hetddft_set_* / hetddft_get_* calls.
 *
 * Array conventions (identical to the Fortran arrays, so a Fortran host can
 * pass its arrays through ISO_C_BINDING unchanged):
 *   real field      f(nx,ny,nz)        double[nz][ny][nx]         (x fastest)
 *   complex field   psi(nx,ny,nz)      double[nz][ny][nx][2]      (re,im)
 *   half spectrum   k(nx/2+1,ny,nz)    double[nz][ny][nx/2+1]     (real kernels)
 *   particles       rimp(N_par,3)      double[3][N_par]           (x comps first)
 *
 * Every function returns 0 on success or a negative HETDDFT_E_* code; the
 * message of the last failure is available from hetddft_last_error().
 * Calls on one context are synchronous with respect to their host outputs
 * and must not overlap.  One context per host thread.
 */
#ifndef HETDDFT_H
#define HETDDFT_H

#ifdef __cplusplus
extern "C" {
#endif

#define HETDDFT_OK 0
#define HETDDFT_E_ARG -1        /* bad argument or unsupported grid size   */
#define HETDDFT_E_HIP -2        /* HIP runtime failure                     */
#define HETDDFT_E_NOMEM -3      /* device allocation failed                */
#define HETDDFT_E_STATE -4      /* call out of order (e.g. no kernels set) */
#define HETDDFT_E_NONFINITE -5  /* non-finite error estimate in a PC step  */

#define HETDDFT_FP64 0          /* every field FP64                         */
#define HETDDFT_MIXED 1         /* potential fields and FFTs FP32; psi,
                                   reductions and particles FP64            */

typedef struct hetddft_ctx hetddft_ctx;

/* Parameters of one droplet + impurity instance.  Field names follow the
 * reference namelist /input/ (DFT4He3d.f90:147-165) and modules.f90. */
typedef struct hetddft_params {
  int nx, ny, nz;            /* grid; each in the tstgrid list (tstgrid.f90:8-12), nx even */
  double xmax, ymax, zmax;   /* box is [-xmax, xmax) etc., hx = 2 xmax/nx */
  int n_par;                 /* number of Bohmian test particles, 0 .. 2^27 */
  double mimpur;             /* impurity mass in u */
  int ironx;                 /* smoothing passes of the particle histogram */
  int core_otc;              /* 1: core4='OTC' (alpha_s term on), 0: 'OT ' */
  double lambdah, tzmean, tzsurf;  /* time constant timec (DFT4He3d.f90:773-782) */
  /* Orsay-Trento functional (modules.f90:122-141, 201-213) */
  double cp4, cpp4, alphas, l, den0s, h2o2m4, C, beta, den_m, h4, eps4, sigma4;
  int precision;             /* HETDDFT_FP64 or HETDDFT_MIXED */
} hetddft_params;

/* Cored Morse He-impurity pair potential V(r)=D[e^{-2a(r-re)}-2e^{-a(r-re)}],
 * V = umax for r < rc (Select_Pot core convention, V_impur.f90:179-189). */
typedef struct hetddft_morse {
  double D, re, alpha, rc, umax;
  int nq;                    /* momenta of the spline transform (Nq) */
} hetddft_morse;

/* Fills the Orsay-Trento defaults and a 64^3 box of +-16 A. */
int hetddft_default_params(hetddft_params* p);

/* Allocates device state.  Replaces dimen() + fftini() (dimen.f90:8-146,
 * fft.f90:5-76). */
int hetddft_create(const hetddft_params* p, hetddft_ctx** out);

/* Frees the context; NULL is a no-op (HETDDFT_OK). */
int hetddft_destroy(hetddft_ctx* ctx);

/* Message of the last failure on ctx; with ctx == NULL, the message of the
 * last failed hetddft_create or hetddft_particle_means_dev on this thread. */
const char* hetddft_last_error(const hetddft_ctx* ctx);

/* Host-side generator of the k-space kernels on the half grid:
 * fvlj4 (fforma.f90:11-61), wcgk (initcg.f90:5-51), kalfs
 * (DFT4He3d.f90:566-570) and the FP32-rounded Vq of the cored Morse potential
 * (FT_V_spline.f90:14-276 via DFT4He3d.f90:527-536).  Any pointer may be NULL. */
int hetddft_make_kernels(const hetddft_params* p, const hetddft_morse* morse, double* fvlj4,
                         double* wcgk, double* kalfs, double* vq);

/* Uploads the k-space kernels (the arrays fforma/initcg/FT_V_spline leave in
 * fvlj4, wcgk, kalfs and real(vq)). */
int hetddft_set_kernels(hetddft_ctx* ctx, const double* fvlj4, const double* wcgk,
                        const double* kalfs, const double* vq);

/* Uploads psi (complex, interleaved), rimp and vimp (Fortran (N_par,3)). */
int hetddft_set_state(hetddft_ctx* ctx, const double* psi, const double* rimp, const double* vimp);

/* den = conjg(psi)*psi, first potential evaluation, zeroed histories and
 * accumulators (DFT4He3d.f90:630-634, 841-845; dimen.f90:78-79, 131-134):
 * a context that has run before restarts exactly like a new one. */
int hetddft_start(hetddft_ctx* ctx);

/* potenimp(); poten(); forceimp()  (potenimp.f90:1-73, poten.f90:5-129). */
int hetddft_potential(hetddft_ctx* ctx);

/* steprk(deltat)  (steprk.f90:1-128); deltat in K^-1. */
int hetddft_step_rk(hetddft_ctx* ctx, double deltat);

/* steppc(deltat, errHe, errimp, errvimp)  (steppc.f90:1-155).  err may be
 * NULL: the step then runs without a host synchronisation and the error
 * estimates stay on the device (hetddft_last_errors fetches them, and
 * hetddft_sync checks them).  A non-finite estimate is HETDDFT_E_NONFINITE
 * (the reference prints errHe every step, steppc.f90:109-112). */
int hetddft_step_pc(hetddft_ctx* ctx, double deltat, double err[3]);

int hetddft_last_errors(hetddft_ctx* ctx, double err[3]);

/* One iteration of the reference time loop (DFT4He3d.f90:849-863):
 * steprk for the first three calls after hetddft_start, steppc after, then
 * potenimp/poten/forceimp.  err (may be NULL) gets the PC errors, or zeros
 * after an RK step. */
int hetddft_iterate(hetddft_ctx* ctx, double deltat, double err[3]);

/* Means of rimp (A), vimp (A/ps) and aimp (A/ps^2) (DFT4He3d.f90:864-878). */
int hetddft_particle_means(hetddft_ctx* ctx, double out[9]);

/* The same means with no context and no allocation: rimp, vimp and aimp are
 * DEVICE arrays (N_par,3) of the current GPU, out is a DEVICE array of 9
 * doubles written as hetddft_particle_means writes its host array (same sums,
 * bit for bit), and the work is queued on `stream` (a hipStream_t; NULL is
 * the null stream) without a host synchronisation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * partial sums live in a fixed 8 KB device array of the library (one per GPU,
 * held until the process ends), so n_par is limited to
 * HETDDFT_MEANS_DEV_MAX_PAR; calls that overlap on one GPU share that array
 * and race.  Errors: hetddft_last_error(NULL). */
#define HETDDFT_MEANS_DEV_MAX_PAR (1 << 30)
int hetddft_particle_means_dev(const double* rimp, const double* vimp, const double* aimp, int n_par,
                               double* out, void* stream);

/* energy() (energy.f90:12-129) plus the norm sum(den)*dxyz:
 * out = {etot, etot4, ekin4, elj4, ealphas, esolid, ecor4, eimpu, ekinx,
 *        ekinqx, eHeX, norm}. */
int hetddft_energy(hetddft_ctx* ctx, double out[12]);

/* Downloads psi, rimp, vimp, aimp (any pointer may be NULL). */
int hetddft_get_state(hetddft_ctx* ctx, double* psi, double* rimp, double* vimp, double* aimp);

/* Downloads one real field (as double) by id.  The force convolutions
 * FXCONV..FZCONV (-V*d_c den, potenimp.f90:39-50) are read only by the
 * trilinear gather at the particles' cell corners, so with n_par > 0 a
 * potential evaluation writes only their z planes that hold such corners
 * (their values there are those of the full transform); the other planes
 * keep what they held before. */
#define HETDDFT_F_DEN 0
#define HETDDFT_F_POT4 1
#define HETDDFT_F_UIMP 2
#define HETDDFT_F_DENCG 3
#define HETDDFT_F_DELJ4 4
#define HETDDFT_F_FALFS 5       /* 1 - denalf/den0s (term_alfa.f90:28-30) */
#define HETDDFT_F_DXDEN 6
#define HETDDFT_F_DYDEN 7
#define HETDDFT_F_DZDEN 8
#define HETDDFT_F_INTXALF 9
#define HETDDFT_F_INTYALF 10
#define HETDDFT_F_INTZALF 11
#define HETDDFT_F_DENX 12
#define HETDDFT_F_PSIX 13
#define HETDDFT_F_HPSIX 14
#define HETDDFT_F_FXCONV 15
#define HETDDFT_F_FYCONV 16
#define HETDDFT_F_FZCONV 17
#define HETDDFT_F_CORREL 18
#define HETDDFT_F_UALPHAS_CONV 19
int hetddft_get_field(hetddft_ctx* ctx, int which, double* out);

/* Standalone batched 3D r2c + c2r convolution of nf real fields with one
 * real k-space kernel: out_i = c2r(r2c(in_i) * kern) * renor.  Exposes the
 * transform layer (fft.f90) for testing and for hosts that only offload the
 * FFTs.  Arrays are host pointers in the layouts above. */
int hetddft_convolve(hetddft_ctx* ctx, int nf, const double* const* in, const double* kern,
                     double* const* out);

/* Per-kernel-class timing.  hetddft_profile(ctx, 1) resets the counters and
 * brackets every later launch of the classes below with HIP events on the
 * context's stream; hetddft_profile_read returns the summed event time (ms),
 * the number of launches and the algorithmic bytes those launches move
 * (inputs read once + outputs written once).  Profiling adds events between
 * kernels; leave it off for throughput runs. */
#define HETDDFT_PROF_XYF 0   /* r2c along x and y of whole planes          */
#define HETDDFT_PROF_XF 1    /* r2c along x (planes too large for LDS)      */
#define HETDDFT_PROF_YF 2    /* forward FFT along y                         */
#define HETDDFT_PROF_Z 3     /* z FFT, k-space kernel product, inverse z    */
#define HETDDFT_PROF_YI 4    /* inverse FFT along y                         */
#define HETDDFT_PROF_XI 5    /* c2r along x                                 */
#define HETDDFT_PROF_XYI 6   /* c2r along y and x of whole planes           */
#define HETDDFT_PROF_POT 7   /* alpha_s divergence + pot4 assembly          */
#define HETDDFT_PROF_PCP 8   /* Hamming predictor/modifier of psi           */
#define HETDDFT_PROF_PCC 9   /* Hamming corrector of psi + errHe            */
#define HETDDFT_PROF_GRAD 10 /* grad(den)                                    */
#define HETDDFT_PROF_N 11
int hetddft_profile(hetddft_ctx* ctx, int enable);
int hetddft_profile_read(hetddft_ctx* ctx, int cls, double* ms_total, long long* launches,
                         double* bytes_total);

/* Per-step diagnostics on the device, in one pass over the grid (the
 * driver's per-step output, DFT4He3d.f90:879-930, and the chemical potential
 * and current of the dynamics):
 *   out[HETDDFT_DIAG_NORM]         sum(den) dxyz                (DFT4He3d.f90:884)
 *   out[HETDDFT_DIAG_CM + c]       sum(den x_c) dxyz; the reference's centre of
 *                                  mass is this / n4        (r_cm.f90:23-58)
 *   out[HETDDFT_DIAG_MU]           <psi|H|psi>/<psi|psi>, H = -h2o2m4 lap + pot4
 *   out[HETDDFT_DIAG_JTOT + c]     integral of j_c, j = 2 h2o2m4 Im(conj(psi) grad psi)
 *   out[HETDDFT_DIAG_MSX + c]      sqrt(sum(denx x_c^2) dxyz)  (DFT4He3d.f90:918-928)
 *   out[HETDDFT_DIAG_PMINMAX + 2k] min, + 2k + 1 max of vimp(:,k) (k = 0..2) and
 *                                  aimp(:,k-3) (k = 3..5)   (DFT4He3d.f90:900-903)
 * Derivatives are the 13-point stencils of the step (zero outside the box). */
#define HETDDFT_DIAG_NORM 0
#define HETDDFT_DIAG_CM 1
#define HETDDFT_DIAG_MU 4
#define HETDDFT_DIAG_JTOT 5
#define HETDDFT_DIAG_MSX 8
#define HETDDFT_DIAG_PMINMAX 11
#define HETDDFT_DIAG_N 23
int hetddft_diagnostics(hetddft_ctx* ctx, double out[HETDDFT_DIAG_N]);

/* The four k-space kernels of hetddft_make_kernels computed on the device and
 * stored in the context as hetddft_set_kernels stores them (replaces the pair
 * make_kernels + set_kernels, with no host arrays).  The spline of the
 * potential and the spline over momenta stay on the host (sequential and
 * small); the momentum sums of FT_V_spline (FT_V_spline.f90:151-225, ~10^5
 * terms per momentum) and every half-grid point of fvlj4 (fforma.f90:44-56),
 * wcgk (initcg.f90:37-42), kalfs (DFT4He3d.f90:566-570) and Vq (rounded to
 * FP32 as DFT4He3d.f90:533) run on the device.  The sums are added in another
 * order than on the host, so values differ from hetddft_make_kernels in the
 * last bits (and a Vq entry by one FP32 ulp where that moves a rounding). */
int hetddft_init_kernels(hetddft_ctx* ctx, const hetddft_morse* morse);

/* The context's kernels back in the layout of hetddft_make_kernels (FP32
 * values widened in mixed precision); any pointer may be NULL. */
int hetddft_get_kernels(hetddft_ctx* ctx, double* fvlj4, double* wcgk, double* kalfs, double* vq);

/* Test-particle positions drawn by rejection from a density, as sorteo
 * (sorteo.f90:1-37, called at DFT4He3d.f90:598 for starts from density files):
 * a uniform point in the box [lo[c], hi[c]] and a uniform value in
 * [0, max(denx)) are drawn until the value lies below the trilinear
 * interpolation of denx at the point (interpola3denx, sorteo.f90:40-74).
 * denx is nx*ny*nz doubles (x fastest); rimp receives (N_par, 3).  The random
 * numbers are Philox4x64-10 keyed by (seed, 0), block (particle, attempt, 0, 0),
 * 53-bit uniforms: the reference's random_number stream is compiler-specific,
 * so the positions follow the reference's distribution, not its values.
 * HETDDFT_E_STATE if a particle is not accepted within max_attempts. */
int hetddft_sample_particles(hetddft_ctx* ctx, const double* denx, const double lo[3], const double hi[3],
                             unsigned long long seed, long long max_attempts, double* rimp);

/* A fresh start built on the device: psi, rimp and vimp of a Woods-Saxon
 * droplet and a Gaussian test-particle cloud, as the reference's mode-0 start
 * (DFT4He3d.f90:585-605) builds them from a density file: psi = sqrt(den)
 * exp(i (x lamdax + y lamday + z lamdaz)), every particle with velocity
 * v0 * Ktops (A/ps to A*K, 7.63822291).  Here den is the closed form
 *   rho0 / (1 + exp((r - R)/a)),  r = |x - centre|,
 *   R = r0 n_he^(1/3) + ripple * s / max|s| over the grid,
 *   s = sum over l = 1..4, m = -l..l of clm[k] P_l^|m|(cos theta) T_m(phi)
 *       (k runs over (l, m) in that order; P with the Condon-Shortley
 *       phase; T_m = cos(m phi) for m >= 0, sin(|m| phi) for m < 0),
 * divided by 1 + exp(-(r - cavity)/cavity_w) when cavity > 0, and rescaled
 * so that sum(den) dxyz = n_he, on the grid x(ix) = -xmax + hx (ix - 1).
 * phase_noise > 0 adds sigma_phi * g to the phase at every point and
 * sigma > 0 places particle p at centre + sigma (g0, g1, g2); g are standard
 * normals from Philox4x64-10 (the generator of hetddft_sample_particles) by
 * Box-Muller: for point i, key (seed, 1) and counter (i, 0, 0, 0); for
 * particle p, key (seed, 2) and counter (p, 0, 0, 0), words 0-1 giving g0 and
 * g1, words 2-3 g2; g = sqrt(-2 ln u) cos(2 pi v) (sin for g1) with
 * u = ((w0 >> 11) + 1) / 2^53 and v = (w1 >> 11) / 2^53.  sigma <= 0 leaves
 * rimp as it is (for hetddft_sample_particles); vimp is set either way.
 * Replaces the state like hetddft_set_state (hetddft_start must follow).
 * No host-side field is formed or copied, so an ensemble's instances are set
 * up without host work per grid point.  HETDDFT_E_ARG for n_he, rho0, r0 or
 * a not positive, negative ripple, phase_noise or cavity, or cavity > 0 with
 * cavity_w not positive. */
typedef struct hetddft_droplet {
  double n_he;           /* He atoms */
  double rho0;           /* 0.0218 A^-3 */
  double r0;             /* 2.22 A */
  double a;              /* surface width, 0.7 A */
  double cavity;         /* cavity radius (A); 0: none */
  double cavity_w;       /* cavity edge width (A) */
  double ripple;         /* surface ripple amplitude (A); 0: none */
  double clm[24];        /* ripple coefficients */
  double centre[3];      /* droplet and cloud centre (A) */
  double lamda[3];       /* plane-wave boost (xlamdax, xlamday, xlamdaz; A^-1) */
  double phase_noise;    /* sigma_phi (rad) */
  double sigma;          /* particle cloud width (A) */
  double v0[3];          /* particle velocity (A/ps) */
  unsigned long long seed;
} hetddft_droplet;
int hetddft_init_droplet(hetddft_ctx* ctx, const hetddft_droplet* d);

/* Full-history checkpoint for bit-exact resume.  No counterpart in the
 * reference, whose restarts read psi and the particles back (readenc.f90:96-114,
 * mode 2) and begin again with three RK steps; this one keeps everything the
 * next steps read: psi with its PC slots (psi history x3, H psi history x2,
 * pc) and the RK accumulator, the particles with their slots and qr, qv, pcr,
 * pcv, the slot order and the step counter.  _size gives the bytes of the
 * image; _save writes it to a host buffer of that size; _load restores it into
 * a context of the same grid, N_par and precision (after
 * hetddft_set_kernels) and re-runs the potential evaluation, after which
 * hetddft_iterate continues exactly as the saved context would have. */
int hetddft_checkpoint_size(hetddft_ctx* ctx, long long* bytes);
int hetddft_checkpoint_save(hetddft_ctx* ctx, void* buf);
int hetddft_checkpoint_load(hetddft_ctx* ctx, const void* buf);

/* Current density j = 2 h2o2m4 Im(conj(psi) grad psi) (units of psi^2 A/K^-1)
 * into host arrays of nx*ny*nz doubles (x fastest); any pointer may be NULL.
 * The reference writes no current (pcurr is a dump period only); this is the
 * field behind HETDDFT_DIAG_JTOT. */
int hetddft_current(hetddft_ctx* ctx, double* jx, double* jy, double* jz);

/* Waits for all work queued on the context's stream.  When the last call was
 * a PC step (or an iteration ending in one) whose error estimates were not
 * fetched, it reduces them first and returns HETDDFT_E_NONFINITE if one is
 * not finite: a run driven with err = NULL learns of a diverged state here. */
int hetddft_sync(hetddft_ctx* ctx);

/* With HETDDFT_GRAPH=1 (both streams) or 2 (particle chain on the main
 * stream) in the environment at create time, hetddft_iterate captures its PC
 * steps as HIP graphs, one per state of the rotating history slots (two in
 * steady state), and replays them; by default the steps are launched
 * directly.  *count receives the number of captured graphs of the context;
 * *enabled is 0 when graphs are off (by default, or after a failed capture).
 * No counterpart in the reference (a diagnostic of this library's launch
 * path). */
int hetddft_graph_info(hetddft_ctx* ctx, int* count, int* enabled);

/* Device bytes allocated by the context. */
long long hetddft_device_bytes(const hetddft_ctx* ctx);

/* Launch shapes and their autotuner.  No counterpart in the reference (its
 * loops have no launch shapes).  Each stencil kernel class below runs on a
 * z-marching engine with a tile of TX x TY columns per workgroup
 * (tile 0 = 64x4, 1 = 32x8, 2 = 32x16, 3 = 64x8) and zc z planes per
 * workgroup, with (prefetch = 1) or without (0) the next plane's loads in
 * flight while a plane is computed (without: fewer registers, more resident
 * waves); the register-resident z transform of 64-, 128- and 256-point
 * z lines takes fft_z_tile consecutive ky per work item (8, 16 or 32; 0 when
 * the grid has no such choice) and runs fft_z_wgs persistent workgroups per
 * CU (1..8) over the work items.  Every choice computes each output value with
 * the same operations in the same order, so the state after any step is the
 * same bit for bit whatever the tuning (the errHe estimate sums per-workgroup
 * partials, so its last bits follow the corrector's workgroup count).
 *
 * hetddft_autotune times each candidate on the context's own state: it saves
 * a checkpoint (hetddft_checkpoint_save), runs `steps` profiled PC iterations
 * (hetddft_iterate with deltat) per candidate -- every tile shape with z chunks giving
 * about 512, 768, 1024 and 2048 workgroups, with and without prefetch, then
 * each z-transform tile width with 1..4 workgroups per CU -- keeps the fastest shape per class, restores the checkpoint
 * (hetddft_checkpoint_load) and applies the choice.  It needs
 * hetddft_start first.  A later call in the same process for the same grid,
 * precision, N_par and core reuses the result without timing.  out (may be
 * NULL) receives the choice and ms_per_step, the mean wall time of one PC
 * iteration of the chosen shapes over `steps` iterations.  steps <= 0
 * selects 4. */
#define HETDDFT_TUNE_GRAD 0 /* grad(den)                                   */
#define HETDDFT_TUNE_POT 1  /* alpha_s divergence + pot4 assembly          */
#define HETDDFT_TUNE_PCP 2  /* Hamming predictor/modifier of psi           */
#define HETDDFT_TUNE_PCC 3  /* Hamming corrector of psi + errHe            */
#define HETDDFT_TUNE_RK 4   /* RK-Gill stage (takes the predictor's shape) */
#define HETDDFT_TUNE_NSTENCIL 5
typedef struct hetddft_tuning {
  int tile[HETDDFT_TUNE_NSTENCIL];
  int zc[HETDDFT_TUNE_NSTENCIL];
  int prefetch[HETDDFT_TUNE_NSTENCIL];
  int fft_z_tile;
  int fft_z_wgs;
  double ms_per_step;
} hetddft_tuning;
int hetddft_autotune(hetddft_ctx* ctx, double deltat, int steps, hetddft_tuning* out);
int hetddft_get_tuning(hetddft_ctx* ctx, hetddft_tuning* out);
/* Applies a tuning (from hetddft_get_tuning or a saved autotune result);
 * tile outside 0..3, zc < 1, prefetch other than 0 or 1, fft_z_wgs outside
 * 0..8 (0 keeps the current one), or an fft_z_tile this grid does not offer
 * (other than 0, which keeps the current one) is HETDDFT_E_ARG. */
int hetddft_set_tuning(hetddft_ctx* ctx, const hetddft_tuning* t);

/* Batched variant: an ensemble of independent instances (contexts of any
 * grids, N_par and precisions, each created, loaded and started as above)
 * advanced together -- the reference's time loop (DFT4He3d.f90:848-863) run
 * for every instance.  No counterpart in the reference, which runs one
 * droplet per process.  hetddft_batch_create borrows the n context pointers
 * (they must outlive the batch and must not be used by other threads while
 * hetddft_batch_iterate runs) and spreads them over nthreads host threads
 * (nthreads <= 0: min(n, 8)), largest grids first, so each thread holds about
 * the same number of grid points.  hetddft_batch_iterate runs `steps`
 * hetddft_iterate calls on every instance with deltat[i] for instance i: each
 * thread submits one iteration of each of its instances in turn, with no
 * synchronisation between iterations, and the call returns when every
 * instance has finished.  Instances with equal hetddft_params, launch
 * shapes (hetddft_set_tuning) and stream options form groups of up to 8 at
 * hetddft_batch_create; a group whose members are at the same iteration
 * and get the same deltat advances in shared launches (one launch per
 * kernel carries every member, on the first member's streams), otherwise
 * its members advance one by one.  Either way every instance's results are
 * bit-identical to those of hetddft_iterate on that instance alone.
 * err (may be NULL) receives 3 doubles per instance,
 * the PC error estimates of its last iteration; they are reduced and checked
 * for every instance in any case, so a diverged instance fails the call with
 * HETDDFT_E_NONFINITE.  On a failure the other
 * threads stop after their current iteration, and the return code and the
 * message (hetddft_batch_last_error) are those of the first failing instance,
 * whose index goes to *failed (may be NULL; -1 when all succeed). */
typedef struct hetddft_batch hetddft_batch;
int hetddft_batch_create(hetddft_ctx* const* ctxs, int n, int nthreads, hetddft_batch** out);
int hetddft_batch_iterate(hetddft_batch* b, const double* deltat, int steps, double* err, int* failed);
const char* hetddft_batch_last_error(const hetddft_batch* b);
/* Frees the batch (not its contexts); NULL is a no-op (HETDDFT_OK). */
int hetddft_batch_destroy(hetddft_batch* b);

#ifdef __cplusplus
}
#endif
#endif
