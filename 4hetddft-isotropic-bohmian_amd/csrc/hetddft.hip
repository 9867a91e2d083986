// MI355X (gfx950) engine behind the hetddft C-ABI (include/hetddft.h).
//
// One context = one droplet+impurity instance resident in HBM.  The per-step
// work of the reference (steprk.f90, steppc.f90, potenimp.f90, poten.f90,
// term_alfa.f90, energy.f90) is issued as a fixed sequence of hand-written
// kernels on one HIP stream; history slots rotate by pointer, never by copy.
#include <hip/hip_ext.h>

#include "../../include/hetddft.h"
#include "common.h"
#include "fft_kernels.h"
#include "field_kernels.h"
#include "init_kernels.h"
#include "kinit.h"
#include "kinit_device.h"
#include "stencil25.h"

#include <algorithm>
#include <atomic>
#include <new>
#include <thread>
#include <chrono>
#include <mutex>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

using namespace hetd;

namespace {

const int kValid[] = {1,  2,  3,  4,  5,  6,  8,  9,  10, 12, 15, 16, 18, 20, 24, 25, 27, 30,
                      32, 36, 40, 45, 48, 50, 54, 60, 64, 72, 75, 80, 81, 90, 96, 100, 108, 120,
                      125, 128, 135, 144, 150, 160, 162, 180, 192, 200, 216, 225, 240, 243, 250, 256};

bool tstgrid_ok(int n) {
  for (int v : kValid)
    if (v == n) return true;
  return false;
}

FftPlan make_plan(int n) {
  FftPlan p{};
  p.n = n;
  int m = n;
  const int order[] = {8, 4, 2, 5, 3};
  while (m > 1) {
    int r = 0;
    for (int c : order)
      if (m % c == 0) {
        r = c;
        break;
      }
    if (r == 0 || p.nstages >= kMaxStages) {
      p.nstages = -1;
      return p;
    }
    p.radix[p.nstages++] = r;
    m /= r;
  }
  return p;
}

Stencil13 make_stencil(int order, double h) {
  // Centre rows of the 13-point table (mates.f:827-832 for K=1, :874-879 for
  // K=2), offsets -6..6, times 1/h^order as Deriv3D_p does.
  static const long long d1n[6] = {6, -15, 5, -1, 1, -1}, d1d[6] = {7, 56, 63, 56, 385, 5544};
  static const long long d2n[6] = {12, -15, 10, -1, 2, -1}, d2d[6] = {7, 56, 189, 112, 1925, 16632};
  Stencil13 s{};
  const double sto = 1.0 / (order == 1 ? h : h * h);
  for (int j = 1; j <= 6; ++j) {
    double c;
    if (order == 1) {
      c = (double)d1n[j - 1] / (double)d1d[j - 1];
      s.w[6 + j] = c * sto;
      s.w[6 - j] = (-c) * sto;
    } else {
      c = (double)d2n[j - 1] / (double)d2d[j - 1];
      s.w[6 + j] = c * sto;
      s.w[6 - j] = c * sto;
    }
  }
  s.w[6] = order == 1 ? 0.0 * sto : ((double)(-5369) / 1800.0) * sto;
  return s;
}

}  // namespace

// FFT kernel tables (filled in by select_fft below)
template <class T>
struct FftKernelsT {
  void (*xf)(XfArgs<T>, Grid, FftPlan, const cplx<T>*, const cplx<T>*) = nullptr;
  void (*xi)(XiArgs<T, T>, Grid, FftPlan, const cplx<T>*, const cplx<T>*) = nullptr;
  void (*yf)(YArgs<T>, Grid, FftPlan, const cplx<T>*) = nullptr;
  void (*yi)(YArgs<T>, Grid, FftPlan, const cplx<T>*) = nullptr;
  void (*zc)(ZArgs<T>, Grid, FftPlan, const cplx<T>*) = nullptr;
  int tyz = 16;        // z-pass tile width (consecutive ky)
  size_t lds_x = 0, lds_y = 0, lds_z = 0;
  // persistent z pass (compile-time lengths)
  void (*zp)(ZArgs<T>, int, Grid, const cplx<T>*) = nullptr;
  size_t lds_zp = 0;
  int typ = 0;
  // register-resident z pass at ky-tile widths 8, 16, 32 (same arithmetic
  // per column, so the same results); zp/lds_zp/typ hold the selected one
  int nzv = 0;
  void (*zpv[3])(ZArgs<T>, int, Grid, const cplx<T>*) = {};
  size_t lds_zpv[3] = {};
  int typv[3] = {};
  // whole-plane x+y kernels (when a plane fits in LDS)
  void (*xyf)(XfArgs<T>, Grid, const cplx<T>*, const cplx<T>*, const cplx<T>*) = nullptr;
  void (*xyi)(XiArgs<T, T>, Grid, const cplx<T>*, const cplx<T>*, const cplx<T>*) = nullptr;
  size_t lds_xy = 0;
  // the same generic plane kernels in 512-thread workgroups, for launches of
  // at least two planes per CU (null for the 128 x 128 kernels)
  void (*xyf_h)(XfArgs<T>, Grid, const cplx<T>*, const cplx<T>*, const cplx<T>*) = nullptr;
  void (*xyi_h)(XiArgs<T, T>, Grid, const cplx<T>*, const cplx<T>*, const cplx<T>*) = nullptr;
  // persistent form of the 128 x 128 FP64 inverse plane kernel (one
  // workgroup per CU), used when a launch holds at least two planes per CU
  void (*xyip)(XiArgs<T, T>, Grid, int, const cplx<T>*, const cplx<T>*) = nullptr;
};
using FftKernels64 = FftKernelsT<double>;
using FftKernels32 = FftKernelsT<float>;

// Device resources of one transform precision: twiddle tables, spectrum
// work buffers (internal order [kz][kx][ky]) and the k-space kernels.
constexpr int kSpectra = 8;
struct FftSet {
  void* tw_xh = nullptr;  // exp(-2 pi i m/(nx/2))
  void* twh = nullptr;    // exp(-2 pi i k/nx), k = 0..nx/2
  void* tw_y = nullptr;
  void* tw_z = nullptr;
  void* tw_p128 = nullptr;  // regrouped twiddles of the 128 x 128 plane kernels (TQ, TX, TR)
  void* spec[kSpectra] = {};
  void* kfvlj4 = nullptr;
  void* kwcgk = nullptr;
  void* kkalfs = nullptr;
  void* kvq = nullptr;
  size_t lds_x = 0, lds_y = 0, lds_z = 0;
  bool ready = false;
};

struct hetddft_ctx {
  hetddft_params p{};
  Grid g{};
  double hx = 0, hy = 0, hz = 0, dxyz = 0, renor = 0, m_imp = 0, h2o2mx = 0;
  hipStream_t st = nullptr;
  // Side stream for the particle work (histogram, ironing, force gather,
  // particle predictor/corrector), which is latency-bound and runs beside the
  // field kernels on st.  Every entry point forks st2 from st on entry and
  // joins it back before returning, so callers see one ordered stream.
  hipStream_t st2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_chain = nullptr, ev_b3 = nullptr, ev_join = nullptr, ev_force = nullptr;
  std::string err;
  long long bytes = 0;
  bool kernels_set = false, started = false;
  int iter_in_run = 0;
  bool mixed = false;
  size_t tsz = 8;  // bytes of the FFT/field real type

  Stencil13 d1x{}, d1y{}, d1z{}, d2x{}, d2y{}, d2z{};
  FftPlan pl_xh{}, pl_y{}, pl_z{};
  int zc = 16;  // default z planes per workgroup of the stencil engine
  int xcd = 1;  // XCD-grouped tile order in the stencil engine
  int tile = 0; // default tile shape of the stencil engine (with_tile)
  // launch shape of each stencil kernel class (HETDDFT_TUNE_*): tile shape
  // and z planes per workgroup; set from tile/zc at create, changed by
  // hetddft_set_tuning / hetddft_autotune.  Results do not depend on them.
  int tile_k[HETDDFT_TUNE_NSTENCIL] = {};
  int zc_k[HETDDFT_TUNE_NSTENCIL] = {};
  int pf_k[HETDDFT_TUNE_NSTENCIL] = {1, 1, 1, 1, 1};
  int cus = 256;      // compute units of the device
  int zp_per_cu = 4;  // resident workgroups per CU of the persistent z pass

  // device memory
  std::vector<void*> allocs;
  FftSet fs;       // transforms of the time step (FP64 or FP32)
  FftSet fs64e;    // FP64 transforms for energy() in mixed mode
  void* e64[5] = {};  // FP64 den, denx, dencg, delj4, uimp for energy() in mixed mode
  c64* timec = nullptr;
  // psi side (always FP64 complex)
  c64* psi = nullptr;
  c64* spare = nullptr;
  c64* old[3] = {nullptr, nullptr, nullptr};
  c64* hold[2] = {nullptr, nullptr};
  c64* pc = nullptr;
  c64* q = nullptr;
  int ioldp[3] = {0, 1, 2}, ioldh[2] = {0, 1};
  // potential side (R = double or float)
  enum {
    F_DEN, F_POT4, F_UIMP, F_DENCG, F_DELJ4, F_FALFS, F_DX, F_DY, F_DZ, F_IX, F_IY, F_IZ,
    F_DENXA, F_PSIX, F_HPSIX, F_FCX, F_FCY, F_FCZ, F_STO1, F_UACONV, F_DENXB, F_COUNT
  };
  void* fld[F_COUNT] = {};
  void* denx = nullptr;  // whichever of DENXA/DENXB holds the ironed histogram
  unsigned int* counts = nullptr;
  int* reg = nullptr;           // occupied-region bookkeeping (field_kernels.h)
  double* fgather = nullptr;    // quantum part of the force at the gather corners (24 N_par doubles)
  int* hpart = nullptr;         // k_hist's partial boxes (chains run in order on one stream)
  bool iron_fused = true;       // ironing passes in one launch (HETDDFT_IRON_FUSED=0: one per pass)
  // The PC error partials of the last step (part[0..err_nb) for psi,
  // part[err_nb..+2 err_pb) for the particles) are reduced into res only when
  // asked for, before anything else reuses part, or dropped by the next step.
  bool err_pending = false;
  int err_nb = 0, err_pb = 0;
  // The particle density chain (histogram -> ironing -> psix -> hpsix) has
  // two buffer sets, each with its own counts and live/work regions.  A chain
  // queued on st2 writes the set that is not current; the potential
  // evaluation that consumes it makes that set current (fld[] entries,
  // counts, reg, denx).  Fields read between steps always belong to the
  // positions of the last potential evaluation, as in the reference.
  struct ChainSet {
    void* denxa = nullptr;
    void* denxb = nullptr;
    void* psix = nullptr;
    void* hpsix = nullptr;
    unsigned int* counts = nullptr;
    int* reg = nullptr;
    void* denx = nullptr;
  };
  ChainSet cset[2];
  int cs = 0;
  // PC steps keep the particle integrators on st2 beside the field work
  // (false with HIP graphs, whose captures fork and join st2 within a step)
  bool lean = true;
  bool chain_pending = false;  // a chain for the next potential evaluation is queued on st2
  bool side_pending = false;   // st2 holds work that st has not waited for
  // particles
  double* rimp = nullptr;
  double* vimp = nullptr;
  double* aimp = nullptr;
  double* rold[3] = {};
  double* vold[3] = {};
  double* aold[2] = {};
  double* qr = nullptr;
  double* qv = nullptr;
  double* pcr = nullptr;
  double* pcv = nullptr;
  int ioldr[3] = {0, 1, 2}, ioldv[3] = {0, 1, 2}, iolda[2] = {0, 1};
  // reductions
  double* part = nullptr;
  long long part_len = 0;
  double* res = nullptr;       // device results
  double* jcur = nullptr;      // current density (3 fields, FP64), allocated by the first hetddft_current
  double* hres = nullptr;      // pinned host copy
  double last_err[3] = {0, 0, 0};
  // Particle means reduced together with the PC error estimates, so that a
  // host that asks for both every step (DFT4He3d.f90:855-878) waits for the
  // GPU once per step: means_wanted records that the last fetch was
  // followed by a particle_means call, means_ready that means_cache holds
  // the means of the current particle state (any step or state change
  // clears it).
  bool means_wanted = false, means_ready = false;
  double means_cache[9] = {};
  // PC steps replayed from hipGraphs (hetddft_iterate): one graph per host
  // state of the rotating slots (two in steady state), keyed by that state.
  // In a dry pass the step's host logic runs without launching anything,
  // which advances the slots exactly as the captured step does.
  bool graph_on = false;
  bool graph_one_stream = false;  // capture with the particle chain on st
  bool capturing = false;         // st is being captured into a graph
  bool one_stream = false;        // set while such a capture runs
  bool dry = false;
  std::map<std::string, hipGraphExec_t> graphs;
  std::vector<hipGraph_t> graph_src;
  // per-kernel-class timing with HIP events on st (hetddft_profile)
  struct ProfRec {
    hipEvent_t a, b;
    int cls;
    double bytes;
  };
  bool prof_on = false;
  std::vector<ProfRec> prof_pending;
  std::vector<hipEvent_t> prof_pool;
  double prof_ms[HETDDFT_PROF_N] = {};
  long long prof_n[HETDDFT_PROF_N] = {};
  double prof_bytes[HETDDFT_PROF_N] = {};
  // bytes per class that kernels count on the device (transform planes of
  // jobs whose extent only the device knows)
  unsigned long long* prof_dev = nullptr;
  FftKernels64 fft64;
  FftKernels32 fft32;
};

namespace {

thread_local std::string g_create_error;

hipEvent_t prof_event(hetddft_ctx* c) {
  if (!c->prof_pool.empty()) {
    hipEvent_t e = c->prof_pool.back();
    c->prof_pool.pop_back();
    return e;
  }
  // timing only: no system-scope fence (no cache writeback inside the
  // measured interval)
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
  return e;
}

// Times launches on st when profiling.  Every kernel launch on st in the
// step path goes through PLAUNCH (profiled class cls >= 0) or SLAUNCH
// (cls = -1), which hand an end event to hipExtLaunchKernel: the event
// carries the dispatch's completion timestamp.  Kernels on st run back to
// back, so a profiled launch lasts from the end of the previous launch on
// st to its own end -- the convention of a profiler's kernel trace, where
// each kernel of a stream starts when the previous one ends.  No start event
// is attached: a separate start packet would add its own gap to the
// measurement.  With profiling off the event is null and PLAUNCH is a plain
// launch.
struct ProfScope {
  hetddft_ctx* c;
  int cls;
  double bytes;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(hetddft_ctx* ctx, int k, double by) : c(ctx), cls(k), bytes(by) {
    if (!c->prof_on) return;
    b = prof_event(c);
  }
  ~ProfScope() {
    if (b) c->prof_pending.push_back({a, b, cls, bytes});
  }
};

// Marks the current point of st in the end-event chain, so that a wait for
// the other stream is not counted into the next profiled launch.
int prof_mark(hetddft_ctx* c) {
  if (!c->prof_on) return HETDDFT_OK;
  hipEvent_t e = prof_event(c);
  if (!e || hipEventRecord(e, c->st) != hipSuccess) {
    c->err = "prof_mark: hipEventRecord failed";
    return HETDDFT_E_HIP;
  }
  c->prof_pending.push_back({nullptr, e, -1, 0.0});
  return HETDDFT_OK;
}

#define PLAUNCH(ps, kern, grid, block, lds, ...)                                               \
  do {                                                                                         \
    if (c->dry) break;                                                                         \
    if ((ps).b)                                                                                \
      hipExtLaunchKernelGGL(kern, grid, block, lds, c->st, (ps).a, (ps).b, 0u, __VA_ARGS__);   \
    else                                                                                       \
      hipLaunchKernelGGL(kern, grid, block, lds, c->st, __VA_ARGS__);                          \
  } while (0)

// A launch on st after which `done` (an event, or null) is complete.  The
// launch signals it itself, as its stop event, so no marker packet sits on st
// between this kernel and the next; with profiling on (the launch carries the
// profiling event) or under graph capture, done is recorded after the launch.
#define PLAUNCH_DONE(ps, done, kern, grid, block, lds, ...)                                    \
  do {                                                                                         \
    if (c->dry) break;                                                                         \
    if ((done) && !(ps).b && !c->capturing) {                                                  \
      hipExtLaunchKernelGGL(kern, grid, block, lds, c->st, nullptr, (done), 0u, __VA_ARGS__);  \
    } else {                                                                                   \
      PLAUNCH(ps, kern, grid, block, lds, __VA_ARGS__);                                        \
      if (done) HIPCHK(c, hipEventRecord((done), c->st));                                      \
    }                                                                                          \
  } while (0)

// A launch on the side stream.
#define S2LAUNCH(kern, grid, block, lds, ...)                                                    \
  do {                                                                                          \
    if (!c->dry) hipLaunchKernelGGL(kern, grid, block, lds, c->one_stream ? c->st : c->st2, __VA_ARGS__); \
  } while (0)

// A launch on st outside the profiled classes (keeps the end-event chain).
#define SLAUNCH(kern, grid, block, lds, ...)                  \
  do {                                                        \
    ProfScope ps_(c, -1, 0.0);                                \
    PLAUNCH(ps_, kern, grid, block, lds, __VA_ARGS__);        \
  } while (0)

int fail(hetddft_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(ctx, expr)                                                                       \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess) return fail(ctx, HETDDFT_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define LAUNCHCHK(ctx)                                                                            \
  do {                                                                                            \
    hipError_t e_ = hipGetLastError();                                                            \
    if (e_ != hipSuccess) return fail(ctx, HETDDFT_E_HIP, "kernel launch: %s (%s:%d)",            \
                                      hipGetErrorString(e_), __FILE__, __LINE__);                 \
  } while (0)

#define RET(expr)             \
  do {                        \
    int r_ = (expr);          \
    if (r_ != HETDDFT_OK) return r_; \
  } while (0)

// Work queued on st so far precedes work queued on st2 from now on.
int fork_side(hetddft_ctx* c) {
  if (c->dry || c->one_stream) return HETDDFT_OK;
  HIPCHK(c, hipEventRecord(c->ev_fork, c->st));
  HIPCHK(c, hipStreamWaitEvent(c->st2, c->ev_fork, 0));
  return HETDDFT_OK;
}

// Work queued on st2 so far precedes work queued on st from now on.
int join_side(hetddft_ctx* c, hipEvent_t e) {
  if (c->dry || c->one_stream) return HETDDFT_OK;
  HIPCHK(c, hipEventRecord(e, c->st2));
  HIPCHK(c, hipStreamWaitEvent(c->st, e, 0));
  return HETDDFT_OK;
}

int dalloc(hetddft_ctx* c, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(c, HETDDFT_E_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  c->allocs.push_back(*p);
  c->bytes += (long long)bytes;
  hipError_t e2 = hipMemset(*p, 0, bytes);
  if (e2 != hipSuccess) return fail(c, HETDDFT_E_HIP, "hipMemset: %s", hipGetErrorString(e2));
  return HETDDFT_OK;
}

template <class T>
int upload_twiddles(hetddft_ctx* c, void** dst, int n, int count, int denom) {
  // table[m] = exp(-2 pi i m/denom), m = 0..count-1, built in long double
  std::vector<cplx<T>> h(count);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int m = 0; m < count; ++m) {
    const long double a = -2.0L * pi * (long double)m / (long double)denom;
    h[m].x = (T)cosl(a);
    h[m].y = (T)sinl(a);
  }
  (void)n;
  RET(dalloc(c, dst, sizeof(cplx<T>) * count));
  HIPCHK(c, hipMemcpy(*dst, h.data(), sizeof(cplx<T>) * count, hipMemcpyHostToDevice));
  return HETDDFT_OK;
}

// Regrouped twiddles of the 128 x 128 plane kernels (fft_kernels.h):
// TQ[16m+q] = W^(qm), TX[8j+k1] = W^(2jk1), TR[2r+i0] = W^(8ri0), W^k =
// exp(-2 pi i k/128) formed as in upload_twiddles (the same values as its
// table).
template <class T>
int upload_twiddles_p128(hetddft_ctx* c, void** dst) {
  std::vector<cplx<T>> h(kP128Regrouped);
  const long double pi = 3.141592653589793238462643383279502884L;
  auto w = [&](int k) {
    const long double a = -2.0L * pi * (long double)k / 128.0L;
    return mk((T)cosl(a), (T)sinl(a));
  };
  for (int q = 0; q < 16; ++q)
    for (int m = 0; m < 8; ++m) h[kP128GQ + 16 * m + q] = w(q * m);
  for (int k1 = 0; k1 < 8; ++k1)
    for (int j = 0; j < 8; ++j) h[kP128GX + 8 * j + k1] = w(2 * j * k1);
  for (int i0 = 0; i0 < 2; ++i0)
    for (int r = 0; r < 8; ++r) h[kP128GR + 2 * r + i0] = w(8 * r * i0);
  RET(dalloc(c, dst, sizeof(cplx<T>) * kP128Regrouped));
  HIPCHK(c, hipMemcpy(*dst, h.data(), sizeof(cplx<T>) * kP128Regrouped, hipMemcpyHostToDevice));
  return HETDDFT_OK;
}

inline unsigned blocks_for(long long n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

// ---- FFT kernel selection: compile-time lengths for the common grid sizes,
// runtime radix plans for the rest of the tstgrid list.
template <class T>
using FftKernels = FftKernelsT<T>;

template <class T, int MH>
void pick_x(FftKernels<T>& k, int m) {
  if constexpr (MH == 64 || MH == 128) {
    // half-length transform in registers (one LDS exchange)
    constexpr int F1 = 8, F2 = MH / 8;             // forward: stage A full, 16-wide HBM rows
    constexpr int I1 = MH / 8, I2 = 8;             // inverse: stage B stores runs of MH/8 pairs
    k.xf = k_fft_xreg<T, T, MH, F1, F2>;
    k.xi = k_fft_xireg<T, T, MH, I1, I2>;
    const size_t bf = xreg_bytes<T, MH, F1, F2>(), bi = xireg_bytes<T, MH, I1, I2>();
    k.lds_x = bf > bi ? bf : bi;
  } else {
    k.xf = k_fft_xf<T, T, MH>;
    k.xi = k_fft_xi<T, T, MH>;
    k.lds_x = (MH > 0 ? 1 : 2) * tile_bytes<T>(kRowsX, m);
  }
}
template <class T, int N>
void pick_y(FftKernels<T>& k, int n) {
  if constexpr (N == 64 || N == 128 || N == 256) {
    // register-resident y pass (one LDS exchange)
    constexpr int R1 = N == 64 ? 8 : 16;
    constexpr int R2 = N / R1;
    k.yf = k_fft_yreg<T, false, N, R1, R2>;
    k.yi = k_fft_yreg<T, true, N, R1, R2>;
    k.lds_y = yreg_bytes<T, N, R1, R2>();
  } else {
    k.yf = k_fft_y<T, false, N>;
    k.yi = k_fft_y<T, true, N>;
    k.lds_y = (N > 0 ? 1 : 2) * tile_bytes<T>(kLinesY, n);
  }
}
template <class T, int N>
void pick_z(FftKernels<T>& k, int n) {
  constexpr int TY = tile_z<T, N>();
  k.zc = k_fft_zconv<T, N, TY>;
  k.tyz = TY;
  k.lds_z = (N > 0 ? 1 : 2) * tile_bytes<T>(k.tyz, n);
  if constexpr (N == 48 || N == 64 || N == 80 || N == 96 || N == 128 || N == 256) {
    // register-resident z pass (two LDS exchanges per transform pair), at
    // ky-tile widths 8, 16 and 32 for the autotuner; 32 by default
    constexpr int R1 = N == 256 ? 16 : 8;
    constexpr int R2 = N / R1;
    k.nzv = 3;
    k.zpv[0] = k_fft_zreg<T, N, 8, R1, R2>;
    k.lds_zpv[0] = zreg_bytes<T, N, 8, R1, R2>();
    k.typv[0] = 8;
    k.zpv[1] = k_fft_zreg<T, N, 16, R1, R2>;
    k.lds_zpv[1] = zreg_bytes<T, N, 16, R1, R2>();
    k.typv[1] = 16;
    k.zpv[2] = k_fft_zreg<T, N, 32, R1, R2>;
    k.lds_zpv[2] = zreg_bytes<T, N, 32, R1, R2>();
    k.typv[2] = 32;
    k.zp = k.zpv[2];
    k.lds_zp = k.lds_zpv[2];
    k.typ = k.typv[2];
  } else if constexpr (N > 0) {
    // persistent z pass with the next item's first-stage rows prefetched
    k.zp = k_fft_zpipe<T, N, TY>;
    k.lds_zp = zpipe_bytes<T, N, TY>();
    k.typ = TY;
  }
}

// Generic plane kernels run in 1024-thread workgroups while a launch gives a
// CU fewer than two planes (shorter per-plane latency), and in 512-thread
// workgroups from two planes per CU on, so that one CU holds more planes at
// different phases of their load, transform and store.  Both forms compute
// every value with the same operations in the same order.
constexpr int kPlaneThreadsHalf = 512;
inline bool plane_half(int items, int cus) { return items >= 2 * cus; }

template <class T, int MH, int NY>
void pick_xy(FftKernels<T>& k) {
  constexpr size_t b = plane_bytes<T>(MH, NY);
  if constexpr (MH == 64 && NY == 128) {
    k.xyf = k_fft_xyf128<T, T>;
    k.xyi = k_fft_xyi128<T, T>;
    k.lds_xy = kPlane128Lds * sizeof(cplx<T>);
    if constexpr (sizeof(T) == 8) k.xyip = k_fft_xyi128p<T, T>;
  } else if constexpr (b <= 160 * 1024) {
    k.xyf = k_fft_xyf<T, T, MH, NY, kPlaneThreads>;
    k.xyi = k_fft_xyi<T, T, MH, NY, kPlaneThreads>;
    k.xyf_h = k_fft_xyf<T, T, MH, NY, kPlaneThreadsHalf>;
    k.xyi_h = k_fft_xyi<T, T, MH, NY, kPlaneThreadsHalf>;
    k.lds_xy = b;
  }
}

template <class T>
FftKernels<T> select_fft(int m, int ny, int nz) {
  FftKernels<T> k;
  if (m * 2 == ny) {
    switch (m) {
      case 12: pick_xy<T, 12, 24>(k); break;
      case 16: pick_xy<T, 16, 32>(k); break;
      case 24: pick_xy<T, 24, 48>(k); break;
      case 32: pick_xy<T, 32, 64>(k); break;
      case 40: pick_xy<T, 40, 80>(k); break;
      case 48: pick_xy<T, 48, 96>(k); break;
      case 64: pick_xy<T, 64, 128>(k); break;
      default: break;
    }
  }
#define HETD_X(M) \
  case M: pick_x<T, M>(k, m); break;
  switch (m) {
    HETD_X(12) HETD_X(16) HETD_X(20) HETD_X(24) HETD_X(32) HETD_X(40) HETD_X(48) HETD_X(64) HETD_X(128)
    default: pick_x<T, 0>(k, m); break;
  }
#undef HETD_X
#define HETD_YZ(N) \
  case N: pick_y<T, N>(k, ny); break;
  switch (ny) {
    HETD_YZ(24) HETD_YZ(32) HETD_YZ(48) HETD_YZ(64) HETD_YZ(80) HETD_YZ(96) HETD_YZ(128) HETD_YZ(256)
    default: pick_y<T, 0>(k, ny); break;
  }
#undef HETD_YZ
#define HETD_YZ(N) \
  case N: pick_z<T, N>(k, nz); break;
  switch (nz) {
    HETD_YZ(24) HETD_YZ(32) HETD_YZ(48) HETD_YZ(64) HETD_YZ(80) HETD_YZ(96) HETD_YZ(128) HETD_YZ(256)
    default: pick_z<T, 0>(k, nz); break;
  }
#undef HETD_YZ
  return k;
}

template <class T>
FftKernels<T>& fft_of(hetddft_ctx* c);
template <>
FftKernels<double>& fft_of<double>(hetddft_ctx* c) {
  return c->fft64;
}
template <>
FftKernels<float>& fft_of<float>(hetddft_ctx* c) {
  return c->fft32;
}

template <class T>
int fft_alloc_tables(hetddft_ctx* c, FftSet& fs) {
  const int m = c->g.nx / 2;
  RET(upload_twiddles<T>(c, &fs.tw_xh, m, m, m));
  RET(upload_twiddles<T>(c, &fs.twh, c->g.nx, m + 1, c->g.nx));
  RET(upload_twiddles<T>(c, &fs.tw_y, c->g.ny, c->g.ny, c->g.ny));
  RET(upload_twiddles<T>(c, &fs.tw_z, c->g.nz, c->g.nz, c->g.nz));
  if (m == 64 && c->g.ny == 128) RET(upload_twiddles_p128<T>(c, &fs.tw_p128));
  FftKernels<T>& k = fft_of<T>(c);
  k = select_fft<T>(m, c->g.ny, c->g.nz);
  fs.lds_x = k.lds_x;
  fs.lds_y = k.lds_y;
  fs.lds_z = k.lds_z;
  const int lim = 160 * 1024;
  if (fs.lds_x > (size_t)lim || fs.lds_y > (size_t)lim || fs.lds_z > (size_t)lim)
    return fail(c, HETDDFT_E_ARG, "FFT tiles exceed LDS");
  HIPCHK(c, hipFuncSetAttribute((const void*)k.xf, hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  HIPCHK(c, hipFuncSetAttribute((const void*)k.xi, hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  HIPCHK(c, hipFuncSetAttribute((const void*)k.yf, hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  HIPCHK(c, hipFuncSetAttribute((const void*)k.yi, hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  HIPCHK(c, hipFuncSetAttribute((const void*)k.zc, hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  if (k.xyf) {
    HIPCHK(c, hipFuncSetAttribute((const void*)k.xyf, hipFuncAttributeMaxDynamicSharedMemorySize, lim));
    HIPCHK(c, hipFuncSetAttribute((const void*)k.xyi, hipFuncAttributeMaxDynamicSharedMemorySize, lim));
    if (k.xyf_h) {
      HIPCHK(c, hipFuncSetAttribute((const void*)k.xyf_h, hipFuncAttributeMaxDynamicSharedMemorySize, lim));
      HIPCHK(c, hipFuncSetAttribute((const void*)k.xyi_h, hipFuncAttributeMaxDynamicSharedMemorySize, lim));
    }
  }
  if (k.xyip) HIPCHK(c, hipFuncSetAttribute((const void*)k.xyip, hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  if (k.zp) HIPCHK(c, hipFuncSetAttribute((const void*)k.zp, hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  for (int v = 0; v < k.nzv; ++v)
    HIPCHK(c, hipFuncSetAttribute((const void*)k.zpv[v], hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  return HETDDFT_OK;
}

// ---- one batched convolution: out_j = c2r(r2c(in_i) * sign*kernel) ------
template <class T>
struct ConvJob {
  int input;
  const void* kern;
  double sign;
  void* out;        // real field T*
  int epi = XI_STORE;
  double ec = 0.0;
  const int* zbox = nullptr;  // device box whose z planes the output needs (z_range), null: all
};

// Five passes over HBM (see fft_kernels.h): inputs i < nin are formed by
// the x-pass producers in xf_in; every job reads one input spectrum.  When
// each input feeds exactly one job, the z pass works in place (a work item
// reads its whole column tile before writing it back, and distinct items
// cover disjoint columns), so a batch needs only nin spectrum buffers.
// Spectrum buffers of one instance's batch: input i in spec[i]; job k in
// place in its input's buffer when every input feeds exactly one job, else in
// spec[nin + k].  False when the batch does not fit the instance's buffers.
template <class T>
bool batch_buffers(const FftSet& fs, int nin, const std::vector<ConvJob<T>>& jobs, std::vector<void*>& sin,
                   std::vector<void*>& sout) {
  const int nout = (int)jobs.size();
  int uses[kMaxBatch] = {};
  for (const ConvJob<T>& j : jobs) {
    if (j.input < 0 || j.input >= nin || nin > kMaxBatch) return false;
    uses[j.input]++;
  }
  bool inplace = nout == nin;
  for (int i = 0; i < nin && inplace; ++i) inplace = uses[i] == 1;
  if (nin < 1 || nout < 1 || (!inplace && nin + nout > kSpectra)) return false;
  for (int i = 0; i < nin; ++i) sin.push_back(fs.spec[i]);
  for (int k = 0; k < nout; ++k) sout.push_back(inplace ? fs.spec[jobs[k].input] : fs.spec[nin + k]);
  return true;
}

template <class T>
int conv_batch_in(hetddft_ctx* c, FftSet& fs, const XfArgs<T>& xf_in, int nin, const std::vector<ConvJob<T>>& jobs,
                  void* const* spec_in, void* const* spec_out, hipEvent_t done = nullptr);

template <class T>
int conv_batch(hetddft_ctx* c, FftSet& fs, const XfArgs<T>& xf_in, int nin,
               const std::vector<ConvJob<T>>& jobs) {
  std::vector<void*> sin, sout;
  if (!batch_buffers<T>(fs, nin, jobs, sin, sout))
    return fail(c, HETDDFT_E_ARG, "conv_batch: %d in / %d out", nin, (int)jobs.size());
  return conv_batch_in<T>(c, fs, xf_in, nin, jobs, sin.data(), sout.data());
}

// One launch per pass for all inputs and jobs given: input i's spectrum in
// spec_in[i], job k's in spec_out[k] (its input's buffer when in place).
// fs supplies the twiddle tables.  The jobs of several instances of one grid
// go into one batch this way.  done (or null): an event that completes with
// the batch's last launch.
template <class T>
int conv_batch_in(hetddft_ctx* c, FftSet& fs, const XfArgs<T>& xf_in, int nin, const std::vector<ConvJob<T>>& jobs,
                  void* const* spec_in, void* const* spec_out, hipEvent_t done) {
  const Grid& g = c->g;
  const int nout = (int)jobs.size();
  if (nin < 1 || nout < 1 || nin > kMaxBatch || nout > kMaxBatch || nout > kMaxJobs)
    return fail(c, HETDDFT_E_ARG, "conv_batch: %d in / %d out", nin, nout);
  for (const ConvJob<T>& j : jobs)
    if (j.input < 0 || j.input >= nin) return fail(c, HETDDFT_E_ARG, "conv_batch: job input %d", j.input);
  XfArgs<T> xa = xf_in;
  for (int i = 0; i < nin; ++i) xa.out[i] = spec_in[i];
  FftKernels<T>& K = fft_of<T>(c);
  const dim3 gx((unsigned)((g.ny + kRowsX - 1) / kRowsX), (unsigned)g.nz, (unsigned)nin);
  const unsigned ylines = (unsigned)(((long long)g.mx * g.nz + kLinesY - 1) / kLinesY);
  const double Rb = (double)g.n * sizeof(T), Sb = 2.0 * (double)g.h * sizeof(T), Kb = (double)g.h * sizeof(T);
  // algorithmic bytes of the x(+y) pass: each distinct input field read once
  // (a field that several producers use, e.g. falfs or the density gradient,
  // counts once), each spectrum written once
  std::vector<const void*> fields;
  for (int i = 0; i < nin; ++i) {
    const int nsrc = xa.op[i] == XF_GRAD_DOT ? 6 : (xa.op[i] == XF_CORREL || xa.op[i] == XF_MUL) ? 2 : 1;
    for (int k = 0; k < nsrc; ++k)
      if (std::find(fields.begin(), fields.end(), (const void*)xa.in[i][k]) == fields.end())
        fields.push_back(xa.in[i][k]);
  }
  const double in_bytes = (double)fields.size() * Rb;
  if (K.xyf) {
    ProfScope ps(c, HETDDFT_PROF_XYF, in_bytes + nin * Sb);
    // the 128 x 128 plane kernels take the regrouped twiddles in the first slot
    const bool half = K.xyf_h && plane_half(g.nz * nin, c->cus);
    PLAUNCH(ps, half ? K.xyf_h : K.xyf, dim3(g.nz, nin), dim3(half ? kPlaneThreadsHalf : kPlaneThreads), K.lds_xy,
        xa, g, (const cplx<T>*)(fs.tw_p128 ? fs.tw_p128 : fs.tw_xh), (const cplx<T>*)fs.twh,
        (const cplx<T>*)fs.tw_y);
    LAUNCHCHK(c);
  } else {
    {
      ProfScope ps(c, HETDDFT_PROF_XF, in_bytes + nin * Sb);
      PLAUNCH(ps, K.xf, gx, dim3(kThreads), fs.lds_x, xa, g, c->pl_xh, (const cplx<T>*)fs.tw_xh,
                                                    (const cplx<T>*)fs.twh);
    }
    LAUNCHCHK(c);
    YArgs<T> ya{};
    for (int i = 0; i < nin; ++i) ya.buf[i] = spec_in[i];
    ProfScope ps(c, HETDDFT_PROF_YF, nin * 2.0 * Sb);
    PLAUNCH(ps, K.yf, dim3(ylines, 1, nin), dim3(kThreads), fs.lds_y, ya, g, c->pl_y, (const cplx<T>*)fs.tw_y);
    LAUNCHCHK(c);
  }
  ZArgs<T> za{};
  XiArgs<T, T> xi{};
  YArgs<T> yb{};
  xi.renor = c->renor;
  // jobs with a z box: the device counts what the passes move for them
  // (profiling only); the host counts the rest
  int nfull = 0;
  auto pbytes = [&](int cls) { return c->prof_on ? c->prof_dev + cls : (unsigned long long*)nullptr; };
  za.pbytes = pbytes(HETDDFT_PROF_Z);
  xi.pbytes = pbytes(K.xyi ? HETDDFT_PROF_XYI : HETDDFT_PROF_XI);
  yb.pbytes = pbytes(HETDDFT_PROF_YI);
  for (int k = 0; k < nout; ++k) {
    const ConvJob<T>& j = jobs[k];
    za.in[k] = spec_in[j.input];
    za.kern[k] = (const T*)j.kern;
    za.ksign[k] = (T)j.sign;
    void* os = spec_out[k];
    za.out[k] = os;
    za.zbox[k] = j.zbox;
    yb.buf[k] = os;
    yb.zbox[k] = j.zbox;
    xi.in[k] = os;
    xi.out[k] = (T*)j.out;
    xi.epi[k] = j.epi;
    xi.ec[k] = j.ec;
    xi.zbox[k] = j.zbox;
    if (!j.zbox) ++nfull;
  }
  // z pass: each distinct input spectrum and k-space kernel read once, each
  // job's output spectrum written once
  std::vector<const void*> kerns;
  std::vector<int> zin;
  for (const ConvJob<T>& j : jobs) {
    if (std::find(kerns.begin(), kerns.end(), j.kern) == kerns.end()) kerns.push_back(j.kern);
    if (std::find(zin.begin(), zin.end(), j.input) == zin.end()) zin.push_back(j.input);
  }
  const double z_bytes = ((double)zin.size() + nfull) * Sb + (double)kerns.size() * Kb;
  if (K.zp) {
    // persistent grid: a few workgroups per CU, each walking the work items
    const long long items = (long long)((g.ny + K.typ - 1) / K.typ) * g.mx * nout;
    const unsigned grid = (unsigned)std::min<long long>(items, (long long)c->cus * c->zp_per_cu);
    ProfScope ps(c, HETDDFT_PROF_Z, z_bytes);
    PLAUNCH(ps, K.zp, dim3(grid), dim3(kThreads), K.lds_zp, za, nout, g, (const cplx<T>*)fs.tw_z);
  } else {
    ProfScope ps(c, HETDDFT_PROF_Z, z_bytes);
    PLAUNCH(ps, K.zc, dim3((g.ny + K.tyz - 1) / K.tyz, g.mx, nout), dim3(kThreads), fs.lds_z,
        za, g, c->pl_z, (const cplx<T>*)fs.tw_z);
  }
  LAUNCHCHK(c);
  if (K.xyi) {
    ProfScope ps(c, HETDDFT_PROF_XYI, nfull * (Sb + Rb));
    const int items = g.nz * nout;
    if (K.xyip && items >= 2 * c->cus)
      PLAUNCH_DONE(ps, done, K.xyip, dim3(c->cus), dim3(kPlaneThreads), K.lds_xy, xi, g, nout,
                   (const cplx<T>*)fs.tw_y, (const cplx<T>*)fs.tw_p128);
    else if (K.xyi_h && plane_half(items, c->cus))
      PLAUNCH_DONE(ps, done, K.xyi_h, dim3(g.nz, nout), dim3(kPlaneThreadsHalf), K.lds_xy, xi, g,
                   (const cplx<T>*)fs.tw_xh, (const cplx<T>*)fs.twh, (const cplx<T>*)fs.tw_y);
    else
      PLAUNCH_DONE(ps, done, K.xyi, dim3(g.nz, nout), dim3(kPlaneThreads), K.lds_xy, xi, g,
                   (const cplx<T>*)(fs.tw_p128 ? fs.tw_p128 : fs.tw_xh), (const cplx<T>*)fs.twh,
                   (const cplx<T>*)fs.tw_y);
    LAUNCHCHK(c);
  } else {
    {
      ProfScope ps(c, HETDDFT_PROF_YI, nfull * 2.0 * Sb);
      PLAUNCH(ps, K.yi, dim3(ylines, 1, nout), dim3(kThreads), fs.lds_y, yb, g, c->pl_y, (const cplx<T>*)fs.tw_y);
    }
    LAUNCHCHK(c);
    ProfScope ps(c, HETDDFT_PROF_XI, nfull * (Sb + Rb));
    const dim3 gxi((unsigned)((g.ny + kRowsX - 1) / kRowsX), (unsigned)g.nz, (unsigned)nout);
    PLAUNCH_DONE(ps, done, K.xi, gxi, dim3(kThreads), fs.lds_x, xi, g, c->pl_xh, (const cplx<T>*)fs.tw_xh,
                 (const cplx<T>*)fs.twh);
    LAUNCHCHK(c);
  }
  return HETDDFT_OK;
}

// Tile shapes of the z-marching stencil engine (c->tile selects one).
using TileA = Tile25<64, 4>;
using TileB = Tile25<32, 8>;
using TileC = Tile25<32, 16>;
using TileD = Tile25<64, 8>;

template <class Fn>
auto with_tile(int tile, Fn&& fn) -> decltype(fn(TileA{})) {
  switch (tile) {
    case 1: return fn(TileB{});
    case 2: return fn(TileC{});
    case 3: return fn(TileD{});
    default: return fn(TileA{});
  }
}
// Tile shape and software pipelining (pf) of a stencil launch:
// fn(tile type, std::integral_constant<bool, pf>).
template <class Fn>
void with_shape(int tile, int pf, Fn&& fn) {
  with_tile(tile, [&](auto sh) {
    if (pf)
      fn(sh, std::integral_constant<bool, true>{});
    else
      fn(sh, std::integral_constant<bool, false>{});
    return 0;
  });
}

// Workgroups of the engine: TX x TY columns, zc planes each.
template <class SH>
dim3 grid25z(const Grid& g, int zc) {
  return dim3((unsigned)((g.nx + SH::TX - 1) / SH::TX), (unsigned)((g.ny + SH::TY - 1) / SH::TY),
              (unsigned)((g.nz + zc - 1) / zc));
}
// ... for stencil class k (HETDDFT_TUNE_*)
template <class SH>
dim3 grid25(const hetddft_ctx* c, int k) {
  return grid25z<SH>(c->g, c->zc_k[k]);
}
// ... for nm members (member-major in z)
template <class SH>
dim3 grid25m(const hetddft_ctx* c, int k, int nm) {
  dim3 d = grid25<SH>(c, k);
  d.z *= (unsigned)nm;
  return d;
}
// Workgroup count of class k's launch.
long long blocks25(const hetddft_ctx* c, int k) {
  const dim3 d = with_tile(c->tile_k[k], [&](auto sh) { return grid25<decltype(sh)>(c, k); });
  return (long long)d.x * d.y * d.z;
}

BoxArgs box_args(const hetddft_ctx* c) {
  BoxArgs b{};
  b.g = c->g;
  b.xmin = -c->p.xmax;
  b.ymin = -c->p.ymax;
  b.zmin = -c->p.zmax;
  b.hx = c->hx;
  b.hy = c->hy;
  b.hz = c->hz;
  b.npar = c->p.n_par;
  return b;
}

GatherArgs gather_args(const hetddft_ctx* c) {
  GatherArgs ga{};
  ga.b = box_args(c);
  ga.d1[0] = c->d1x;
  ga.d1[1] = c->d1y;
  ga.d1[2] = c->d1z;
  ga.m_imp = c->m_imp;
  return ga;
}

// Grid-stride launches over a device-side region.  The chain runs beside
// the field kernels, so it takes few CUs.
unsigned region_blocks() { return 256u; }

// Instances advanced together: every launch goes to the streams of the
// first (lead) instance and carries the data of all.  The members share
// grid, parameters, launch shapes, step counter and slot state, so each step's
// host logic is the same for all of them.  A single instance is a group of
// one.
struct Group {
  hetddft_ctx* m[kMaxMembers] = {};
  int n = 0;
};
inline Group solo(hetddft_ctx* c) {
  Group g;
  g.m[0] = c;
  g.n = 1;
  return g;
}

// Particle density chain of potenimp (denpart.f90, ironing.f90) of every
// member, launched on the lead c's side stream: histogram, ironing and psi_X
// (one launch each for all members), then H psi_X.
template <class T>
int issue_chain(Group& G) {
  hetddft_ctx* c = G.m[0];
  const Grid& g = c->g;
  // ---- potenimp: denpart (denpart.f90:5-49) on the occupied region only,
  // queued on the side stream into the set that is not current
  const BoxArgs b = box_args(c);
  if (c->p.n_par > 0) {
    HistTab t{};
    for (int i = 0; i < G.n; ++i) {
      hetddft_ctx* m = G.m[i];
      t.r[i] = m->rimp;
      t.cnt[i] = m->cset[m->cs ^ 1].counts;
      t.part[i] = m->hpart;
    }
    S2LAUNCH(k_hist, dim3(blocks_for(c->p.n_par), G.n), dim3(kBlock), 0, b, t);
    LAUNCHCHK(c);
  }
  const int nb = c->p.n_par > 0 ? (int)blocks_for(c->p.n_par) : 0;
  const int margin = c->p.ironx + 6;
  const bool fused_iron = c->iron_fused && c->p.ironx >= 1 && c->p.ironx <= 3;
  const double hxyz = c->hx * c->hy * c->hz;
  const double npar = (double)c->p.n_par;
  if (fused_iron) {
    // all passes (and the region update) in one launch
    IronTab<T> t{};
    for (int i = 0; i < G.n; ++i) {
      hetddft_ctx* m = G.m[i];
      hetddft_ctx::ChainSet& S = m->cset[m->cs ^ 1];
      t.reg[i] = S.reg;
      t.part[i] = m->hpart;
      t.cnt[i] = S.counts;
      t.dst[i] = (T*)S.denxa;
      t.psix[i] = (T*)S.psix;
      S.denx = S.denxa;
    }
    auto fused = [&](auto np) {
      S2LAUNCH((k_iron_fused<T, decltype(np)::value>), dim3(region_blocks(), G.n), dim3(kBlock), 0, g, t, nb,
               margin, npar, hxyz);
    };
    if (c->p.ironx == 1) fused(std::integral_constant<int, 1>{});
    else if (c->p.ironx == 2) fused(std::integral_constant<int, 2>{});
    else fused(std::integral_constant<int, 3>{});
    LAUNCHCHK(c);
  } else {
    const int npass = c->p.ironx > 0 ? c->p.ironx : 1;
    for (int i = 0; i < G.n; ++i) {
      hetddft_ctx* m = G.m[i];
      hetddft_ctx::ChainSet& S = m->cset[m->cs ^ 1];
      S2LAUNCH(k_reg_union, dim3(1), dim3(64), 0, g, S.reg, margin, (const int*)m->hpart, nb);
      LAUNCHCHK(c);
      T* cur = (T*)S.denxa;
      T* oth = (T*)S.denxb;
      T* psix = (T*)S.psix;
      S2LAUNCH((k_iron<T>), dim3(region_blocks()), dim3(kBlock), 0, g, (const int*)S.reg, S.counts,
               (const T*)nullptr, cur, npass == 1 ? psix : (T*)nullptr, npar, hxyz, c->p.ironx > 0 ? 1 : 0);
      LAUNCHCHK(c);
      for (int it = 1; it < npass; ++it) {
        S2LAUNCH((k_iron<T>), dim3(region_blocks()), dim3(kBlock), 0, g, (const int*)S.reg,
                 (const unsigned int*)nullptr, (const T*)cur, oth, it == npass - 1 ? psix : (T*)nullptr, npar,
                 hxyz, 1);
        LAUNCHCHK(c);
        std::swap(cur, oth);
      }
      S.denx = cur;
    }
  }
  // the field side needs denx (batch 1); H psi_X (denpart.f90:42-46) enters
  // only the force gather after batch 3, so it stays off the point st waits for
  if (!c->dry && !c->one_stream) HIPCHK(c, hipEventRecord(c->ev_chain, c->st2));
  {
    HpsixTab<T> t{};
    for (int i = 0; i < G.n; ++i) {
      hetddft_ctx* m = G.m[i];
      hetddft_ctx::ChainSet& S = m->cset[m->cs ^ 1];
      t.reg[i] = S.reg;
      t.psix[i] = (const T*)S.psix;
      t.hpsix[i] = (T*)S.hpsix;
      t.cnt[i] = S.counts;
    }
    S2LAUNCH((k_hpsix<T>), dim3(region_blocks(), G.n), dim3(kBlock), 0, g, t, compact<double, true>(c->d2x),
             compact<double, true>(c->d2y), compact<double, true>(c->d2z), c->h2o2mx, margin);
    LAUNCHCHK(c);
  }
  for (int i = 0; i < G.n; ++i) {
    G.m[i]->chain_pending = true;
    G.m[i]->side_pending = true;
  }
  return HETDDFT_OK;
}

// Makes chain set cs current.
void use_chain_set(hetddft_ctx* c) {
  const hetddft_ctx::ChainSet& S = c->cset[c->cs];
  c->fld[hetddft_ctx::F_DENXA] = S.denxa;
  c->fld[hetddft_ctx::F_DENXB] = S.denxb;
  c->fld[hetddft_ctx::F_PSIX] = S.psix;
  c->fld[hetddft_ctx::F_HPSIX] = S.hpsix;
  c->counts = S.counts;
  c->reg = S.reg;
  c->denx = S.denx;
}

// st waits for everything queued on st2.
int settle(hetddft_ctx* c) {
  if (!c->side_pending) return HETDDFT_OK;
  RET(join_side(c, c->ev_join));
  c->side_pending = false;
  return HETDDFT_OK;
}

// potenimp(); poten(); forceimp().  The particle density chain runs on the
// side stream (issued here, or by the caller when chain_issued) beside the
// density-side work on st, which waits for it only where denx enters.
// With leave_side (the PC step's layout), the force gather stays on st2
// without a join: st does not read particle state, and the next potential
// evaluation's wait for the particle chain (queued on st2 after the gather)
// orders it after the gather.
template <class T>
int veval(Group& G, bool leave_side = false) {
  hetddft_ctx* c = G.m[0];
  const Grid& g = c->g;
  for (int i = 0; i < G.n; ++i) G.m[i]->means_ready = false;
  auto F = [&](hetddft_ctx* m, int id) { return (T*)m->fld[id]; };
  const bool otc = c->p.core_otc != 0;
  if (!c->chain_pending) {
    RET(fork_side(c));
    RET(issue_chain<T>(G));
  }
  // ---- poten: derden (derden.f90:26-28), every member in one launch
  {
    ProfScope psg(c, HETDDFT_PROF_GRAD, 4.0 * (double)g.n * sizeof(T));
    with_shape(c->tile_k[HETDDFT_TUNE_GRAD], c->pf_k[HETDDFT_TUNE_GRAD], [&](auto sh, auto pf) {
      using SH = decltype(sh);
      constexpr bool PF = decltype(pf)::value;
      GradTab<T> t{};
      t.nzb = (int)grid25<SH>(c, HETDDFT_TUNE_GRAD).z;
      for (int i = 0; i < G.n; ++i) {
        hetddft_ctx* m = G.m[i];
        t.f[i] = F(m, hetddft_ctx::F_DEN);
        t.fx[i] = F(m, hetddft_ctx::F_DX);
        t.fy[i] = F(m, hetddft_ctx::F_DY);
        t.fz[i] = F(m, hetddft_ctx::F_DZ);
      }
      PLAUNCH(psg, (k_grad3_25<T, SH, PF>), grid25m<SH>(c, HETDDFT_TUNE_GRAD, G.n), dim3(SH::TX, SH::TY), 0, g,
              c->zc_k[HETDDFT_TUNE_GRAD], c->xcd, t, compact<T, false>(c->d1x), compact<T, false>(c->d1y),
              compact<T, false>(c->d1z));
      return 0;
    });
    LAUNCHCHK(c);
  }
  // the particle density chain enters here: its set becomes current
  if (!c->dry && !c->one_stream) HIPCHK(c, hipStreamWaitEvent(c->st, c->ev_chain, 0));
  RET(prof_mark(c));
  for (int i = 0; i < G.n; ++i) {
    hetddft_ctx* m = G.m[i];
    m->cs ^= 1;
    use_chain_set(m);
    m->chain_pending = false;
    m->side_pending = false;
  }
  // Each batch: the members' inputs and jobs in one launch per pass, as
  // many members per launch as the argument arrays hold (kMaxBatch inputs and
  // jobs: larger kernel arguments cost every launch several microseconds).
  auto run_batch = [&](auto&& build, hipEvent_t done = nullptr) -> int {
    XfArgs<T> xa{};
    int nin = 0;
    std::vector<ConvJob<T>> jobs;
    std::vector<void*> sin, sout;
    for (int i = 0; i < G.n; ++i) {
      hetddft_ctx* m = G.m[i];
      XfArgs<T> xm{};
      std::vector<ConvJob<T>> jm;
      const int nm = build(m, xm, jm);
      if (nin > 0 && (nin + nm > kMaxBatch || (int)(jobs.size() + jm.size()) > kMaxBatch)) {
        RET(conv_batch_in<T>(c, c->fs, xa, nin, jobs, sin.data(), sout.data()));
        xa = XfArgs<T>{};
        nin = 0;
        jobs.clear();
        sin.clear();
        sout.clear();
      }
      if (nm > kMaxBatch || (int)jm.size() > kMaxBatch || !batch_buffers<T>(m->fs, nm, jm, sin, sout))
        return fail(c, HETDDFT_E_ARG, "potential batch does not fit");
      for (int k = 0; k < nm; ++k) {
        xa.op[nin + k] = xm.op[k];
        for (int q = 0; q < 6; ++q) xa.in[nin + k][q] = xm.in[k][q];
        xa.c[nin + k][0] = xm.c[k][0];
        xa.c[nin + k][1] = xm.c[k][1];
      }
      for (ConvJob<T> j : jm) {
        j.input += nin;
        jobs.push_back(j);
      }
      nin += nm;
    }
    return conv_batch_in<T>(c, c->fs, xa, nin, jobs, sin.data(), sout.data(), done);
  };
  // batch 1: denx*Vq -> uimp (potenimp.f90:15-17); den*{wcgk, fvlj4, kalfs}
  //          -> dencg, delj4, falfs = 1 - denalf/den0s (poten.f90:27-50,94-99;
  //          term_alfa.f90:28-30)
  RET(run_batch([&](hetddft_ctx* m, XfArgs<T>& xa, std::vector<ConvJob<T>>& jobs) {
    xa.op[0] = XF_COPY;
    xa.in[0][0] = (T*)m->denx;
    xa.op[1] = XF_COPY;
    xa.in[1][0] = F(m, hetddft_ctx::F_DEN);
    jobs.push_back({0, m->fs.kvq, 1.0, F(m, hetddft_ctx::F_UIMP)});
    jobs.push_back({1, m->fs.kwcgk, 1.0, F(m, hetddft_ctx::F_DENCG)});
    jobs.push_back({1, m->fs.kfvlj4, 1.0, F(m, hetddft_ctx::F_DELJ4)});
    if (otc) jobs.push_back({1, m->fs.kkalfs, 1.0, F(m, hetddft_ctx::F_FALFS), XI_ONE_MINUS_DIV, m->p.den0s});
    return 2;
  }));
  // batch 2: correlation term (poten.f90:62-74); alpha_s integrals
  //          int_c = G*(falfs d_c den) (term_alfa.f90:32-52)
  RET(run_batch([&](hetddft_ctx* m, XfArgs<T>& xa, std::vector<ConvJob<T>>& jobs) {
    xa.op[0] = XF_CORREL;
    xa.in[0][0] = F(m, hetddft_ctx::F_DEN);
    xa.in[0][1] = F(m, hetddft_ctx::F_DENCG);
    xa.c[0][0] = m->p.cp4;
    xa.c[0][1] = m->p.cpp4;
    jobs.push_back({0, m->fs.kwcgk, 1.0, F(m, hetddft_ctx::F_STO1)});
    int nin = 1;
    if (otc) {
      const int d[3] = {hetddft_ctx::F_DX, hetddft_ctx::F_DY, hetddft_ctx::F_DZ};
      const int o[3] = {hetddft_ctx::F_IX, hetddft_ctx::F_IY, hetddft_ctx::F_IZ};
      for (int k = 0; k < 3; ++k) {
        xa.op[1 + k] = XF_MUL;
        xa.in[1 + k][0] = F(m, hetddft_ctx::F_FALFS);
        xa.in[1 + k][1] = F(m, d[k]);
        jobs.push_back({1 + k, m->fs.kkalfs, 1.0, F(m, o[k])});
      }
      nin = 4;
    }
    return nin;
  }));
  // batch 3: grad(den).int -> ualphas convolution (term_alfa.f90:65-80);
  //          -Vq*FFT(d_c den) -> force convolutions (potenimp.f90:39-50).
  // Its last launch completes ev_fork, which the force gather waits for.
  const bool gather = c->p.n_par > 0;
  const bool fork_on_launch = gather && !c->dry && !c->one_stream;
  RET(run_batch([&](hetddft_ctx* m, XfArgs<T>& xa, std::vector<ConvJob<T>>& jobs) {
    int nin = 0;
    if (otc) {
      xa.op[0] = XF_GRAD_DOT;
      xa.in[0][0] = F(m, hetddft_ctx::F_DX);
      xa.in[0][1] = F(m, hetddft_ctx::F_DY);
      xa.in[0][2] = F(m, hetddft_ctx::F_DZ);
      xa.in[0][3] = F(m, hetddft_ctx::F_IX);
      xa.in[0][4] = F(m, hetddft_ctx::F_IY);
      xa.in[0][5] = F(m, hetddft_ctx::F_IZ);
      jobs.push_back({0, m->fs.kkalfs, 1.0, F(m, hetddft_ctx::F_UACONV)});
      nin = 1;
    }
    // The force fields enter only the trilinear gather (interpola3,
    // potenimp.f90:71), which reads them at the particles' cell corners: the
    // inverse passes produce the z planes of the gather box (reg, written by
    // this evaluation's particle chain) and leave the others as they were.
    const int d[3] = {hetddft_ctx::F_DX, hetddft_ctx::F_DY, hetddft_ctx::F_DZ};
    const int o[3] = {hetddft_ctx::F_FCX, hetddft_ctx::F_FCY, hetddft_ctx::F_FCZ};
    for (int k = 0; k < 3; ++k) {
      xa.op[nin] = XF_COPY;
      xa.in[nin][0] = F(m, d[k]);
      ConvJob<T> j{nin, m->fs.kvq, -1.0, F(m, o[k])};
      if (m->p.n_par > 0) j.zbox = m->reg + kRegGather;
      jobs.push_back(j);
      ++nin;
    }
    return nin;
  }, fork_on_launch ? c->ev_fork : nullptr));
  // ---- forceimp gather (potenimp.f90:52-71), on the side stream beside
  // the pot4 assembly below: the two read disjoint fields and write aimp and
  // pot4 respectively
  if (gather) {
    if (fork_on_launch) HIPCHK(c, hipStreamWaitEvent(c->st2, c->ev_fork, 0));
    ForceQTab<T> tq{};
    BlendTab<T> tb{};
    for (int i = 0; i < G.n; ++i) {
      hetddft_ctx* m = G.m[i];
      tq.reg[i] = m->reg;
      tq.rp[i] = m->rimp;
      tq.denx[i] = (const T*)m->denx;
      tq.psix[i] = F(m, hetddft_ctx::F_PSIX);
      tq.hpsix[i] = F(m, hetddft_ctx::F_HPSIX);
      tq.out[i] = m->fgather;
      tb.reg[i] = m->reg;
      tb.rp[i] = m->rimp;
      tb.fcx[i] = F(m, hetddft_ctx::F_FCX);
      tb.fcy[i] = F(m, hetddft_ctx::F_FCY);
      tb.fcz[i] = F(m, hetddft_ctx::F_FCZ);
      tb.fq[i] = m->fgather;
      tb.acc[i] = m->aimp;
    }
    S2LAUNCH((k_force_q<T>), dim3(region_blocks(), G.n), dim3(kBlock), 0, gather_args(c), tq);
    LAUNCHCHK(c);
    S2LAUNCH((k_force_blend<T>), dim3(blocks_for(3LL * c->p.n_par), G.n), dim3(kBlock), 0, gather_args(c), tb);
    LAUNCHCHK(c);
  }
  // pot4 assembly with the alpha_s divergence (term_alfa.f90:85-113,
  // poten.f90:80-124), every member in one launch (the members share the
  // functional's constants)
  {
    Pot25T<T> p5{};
    p5.g = g;
    p5.d1x = compact<T, false>(c->d1x);
    p5.d1y = compact<T, false>(c->d1y);
    p5.d1z = compact<T, false>(c->d1z);
    p5.a0 = (T)(c->p.cp4 / 2.0);
    p5.a1 = (T)(c->p.cpp4 / 3.0);
    p5.den0s = (T)c->p.den0s;
    p5.h2o2m4 = (T)c->p.h2o2m4;
    p5.alphas = (T)c->p.alphas;
    p5.C = (T)c->p.C;
    p5.beta = (T)c->p.beta;
    p5.den_m = (T)c->p.den_m;
    p5.otc = otc ? 1 : 0;
    p5.zc = c->zc_k[HETDDFT_TUNE_POT];
    p5.xcd = c->xcd;
    ProfScope ps(c, HETDDFT_PROF_POT, (otc ? 11.0 : 6.0) * (double)g.n * sizeof(T));
    with_shape(c->tile_k[HETDDFT_TUNE_POT], c->pf_k[HETDDFT_TUNE_POT], [&](auto sh, auto pf) {
      using SH = decltype(sh);
      constexpr bool PF = decltype(pf)::value;
      PotTab<T> t{};
      t.nzb = (int)grid25<SH>(c, HETDDFT_TUNE_POT).z;
      for (int i = 0; i < G.n; ++i) {
        hetddft_ctx* m = G.m[i];
        t.den[i] = F(m, hetddft_ctx::F_DEN);
        t.dencg[i] = F(m, hetddft_ctx::F_DENCG);
        t.delj4[i] = F(m, hetddft_ctx::F_DELJ4);
        t.sto1c[i] = F(m, hetddft_ctx::F_STO1);
        t.uimp[i] = F(m, hetddft_ctx::F_UIMP);
        t.falfs[i] = F(m, hetddft_ctx::F_FALFS);
        t.intx[i] = F(m, hetddft_ctx::F_IX);
        t.inty[i] = F(m, hetddft_ctx::F_IY);
        t.intz[i] = F(m, hetddft_ctx::F_IZ);
        t.uaconv[i] = F(m, hetddft_ctx::F_UACONV);
        t.pot4[i] = F(m, hetddft_ctx::F_POT4);
      }
      PLAUNCH(ps, (k_pot_final25<T, SH, PF>), grid25m<SH>(c, HETDDFT_TUNE_POT, G.n), dim3(SH::TX, SH::TY), 0, p5,
              t);
      return 0;
    });
    LAUNCHCHK(c);
  }
  if (leave_side) {
    for (int i = 0; i < G.n; ++i) G.m[i]->side_pending = true;
  } else {
    // H psi_X (and the gather, with particles) ran on the side stream
    RET(join_side(c, c->ev_force));
    RET(prof_mark(c));
  }
  return HETDDFT_OK;
}

template <class T>
int veval(hetddft_ctx* c, bool leave_side = false) {
  Group G = solo(c);
  return veval<T>(G, leave_side);
}

int run_veval(hetddft_ctx* c) { return c->mixed ? veval<float>(c) : veval<double>(c); }

Ham25 ham_args(const hetddft_ctx* c) {
  Ham25 a{};
  a.g = c->g;
  a.d2x = compact<double, true>(c->d2x);
  a.d2y = compact<double, true>(c->d2y);
  a.d2z = compact<double, true>(c->d2z);
  a.h2o2m4 = c->p.h2o2m4;
  a.timec_z = c->timec;
  a.zc = c->zc;
  a.xcd = c->xcd;
  return a;
}
// ... with the z chunk of stencil class k
Ham25 ham_args(const hetddft_ctx* c, int k) {
  Ham25 a = ham_args(c);
  a.zc = c->zc_k[k];
  return a;
}

template <class T>
int step_rk_t(Group& G, double dt) {
  hetddft_ctx* c = G.m[0];
  for (int i = 0; i < G.n; ++i) G.m[i]->means_ready = false;
  const double s2 = std::sqrt(2.0);
  const double arun[4] = {0.5, 1.0 - 1.0 / s2, 1.0 + 1.0 / s2, 1.0 / 6.0};
  const double brun[4] = {2.0, 1.0, 1.0, 2.0};
  const double crun[4] = {0.5, 1.0 - 1.0 / s2, 1.0 + 1.0 / s2, 0.5};
  const int n3 = 3 * c->p.n_par;
  const Ham25 ha = ham_args(c, HETDDFT_TUNE_RK);
  for (int j = 0; j < 4; ++j) {
    c64* outs[kMaxMembers] = {};
    c64* h1os[kMaxMembers] = {};
    double* r1os[kMaxMembers] = {};
    double* v1os[kMaxMembers] = {};
    double* a1os[kMaxMembers] = {};
    for (int i = 0; i < G.n; ++i) {
      hetddft_ctx* m = G.m[i];
      if (j == 0) {
        // psiold(3)=psiold(2); psiold(2)=psiold(1); psiold(1)=psi; hpsiold(2)=hpsiold(1)
        c64* freed = m->old[2];
        m->old[2] = m->old[1];
        m->old[1] = m->old[0];
        m->old[0] = m->psi;  // the current psi becomes psiold(1); new psi goes to `freed`
        outs[i] = freed;
        c64* hf = m->hold[1];
        m->hold[1] = m->hold[0];
        m->hold[0] = hf;
        h1os[i] = hf;
        double* rf = m->rold[2];
        m->rold[2] = m->rold[1];
        m->rold[1] = m->rold[0];
        m->rold[0] = rf;
        r1os[i] = rf;
        double* vf = m->vold[2];
        m->vold[2] = m->vold[1];
        m->vold[1] = m->vold[0];
        m->vold[0] = vf;
        v1os[i] = vf;
        double* af = m->aold[1];
        m->aold[1] = m->aold[0];
        m->aold[0] = af;
        a1os[i] = af;
      } else {
        outs[i] = m->spare;
      }
    }
    with_shape(c->tile_k[HETDDFT_TUNE_RK], c->pf_k[HETDDFT_TUNE_RK], [&](auto sh, auto pf) {
      using SH = decltype(sh);
      constexpr bool PF = decltype(pf)::value;
      RkTab<T> t{};
      t.nzb = (int)grid25<SH>(c, HETDDFT_TUNE_RK).z;
      for (int i = 0; i < G.n; ++i) {
        hetddft_ctx* m = G.m[i];
        t.timec[i] = m->timec;
        t.psi[i] = m->psi;
        t.pot4[i] = (const T*)m->fld[hetddft_ctx::F_POT4];
        t.q[i] = m->q;
        t.out[i] = outs[i];
        t.den[i] = (T*)m->fld[hetddft_ctx::F_DEN];
        t.h1[i] = h1os[i];
      }
      SLAUNCH((k_rk_stage25<T, SH, PF>), grid25m<SH>(c, HETDDFT_TUNE_RK, G.n), dim3(SH::TX, SH::TY), 0, ha, t,
              arun[j], brun[j], crun[j], dt);
      return 0;
    });
    LAUNCHCHK(c);
    PartRkTab pt{};
    for (int i = 0; i < G.n; ++i) {
      hetddft_ctx* m = G.m[i];
      if (j == 0) {
        m->psi = outs[i];
      } else {
        m->spare = m->psi;
        m->psi = outs[i];
      }
      pt.r[i] = m->rimp;
      pt.v[i] = m->vimp;
      pt.acc[i] = m->aimp;
      pt.qr[i] = m->qr;
      pt.qv[i] = m->qv;
      pt.r1[i] = r1os[i];
      pt.v1[i] = v1os[i];
      pt.a1[i] = a1os[i];
    }
    if (n3 > 0) {
      SLAUNCH(k_part_rk, dim3(blocks_for(n3), G.n), dim3(kBlock), 0, n3, pt, arun[j], brun[j], crun[j], dt);
      LAUNCHCHK(c);
    }
    if (j <= 2) RET(veval<T>(G));
  }
  for (int i = 0; i < G.n; ++i) {
    hetddft_ctx* m = G.m[i];
    for (int k = 0; k < 3; ++k) m->ioldp[k] = m->ioldr[k] = m->ioldv[k] = k;
    m->ioldh[0] = m->iolda[0] = 0;
    m->ioldh[1] = m->iolda[1] = 1;
  }
  return HETDDFT_OK;
}

template <class T>
int step_rk_t(hetddft_ctx* c, double dt) {
  Group G = solo(c);
  return step_rk_t<T>(G, dt);
}

template <class T>
int step_pc_t(Group& G, double dt) {
  hetddft_ctx* c = G.m[0];
  for (int i = 0; i < G.n; ++i) G.m[i]->means_ready = false;
  const double c112 = 112.0 / 121.0, c9 = 9.0 / 121.0, c4o3 = 4.0 / 3.0;
  const Grid& g = c->g;
  const int n3 = 3 * c->p.n_par;
  // errHe partial sums: one per workgroup of the corrector
  const unsigned nb = (unsigned)blocks25(c, HETDDFT_TUNE_PCC);
  // In the lean layout the particle predictor, the chain, the force gather
  // and the particle corrector all run on st2, in that order, and st2 waits
  // for st only before each gather (the force fields of batch 3).
  const bool lean = c->lean && !c->one_stream;
  // Particle predictor (steppc.f90:64-87), then the density chain of the
  // predicted positions on the side stream, beside the psi predictor.
  auto particle_side = [&]() -> int {
    {
      PartPredTab pt{};
      for (int i = 0; i < G.n; ++i) {
        hetddft_ctx* m = G.m[i];
        m->err_pending = false;  // the previous step's estimates were not asked for
        const int* R = m->ioldr;
        const int* V = m->ioldv;
        int* A = m->iolda;
        pt.r[i] = m->rimp;
        pt.v[i] = m->vimp;
        pt.acc[i] = m->aimp;
        pt.r3[i] = m->rold[R[2]];
        pt.v1[i] = m->vold[V[0]];
        pt.v2[i] = m->vold[V[1]];
        pt.v3[i] = m->vold[V[2]];
        pt.a1[i] = m->aold[A[0]];
        pt.a2[i] = m->aold[A[1]];
        pt.pcr[i] = m->pcr;
        pt.pcv[i] = m->pcv;
        std::swap(A[0], A[1]);
      }
      if (n3 > 0) {
        if (lean)
          S2LAUNCH(k_part_predict, dim3(blocks_for(n3), G.n), dim3(kBlock), 0, n3, pt, c4o3 * dt, c112);
        else
          SLAUNCH(k_part_predict, dim3(blocks_for(n3), G.n), dim3(kBlock), 0, n3, pt, c4o3 * dt, c112);
        LAUNCHCHK(c);
      }
    }
    if (!lean) RET(fork_side(c));
    RET(issue_chain<T>(G));
    return HETDDFT_OK;
  };
  auto psi_predictor = [&]() -> int {
    // psi predictor + modifier (steppc.f90:36-62), every member in one launch
    {
      const Ham25 hp = ham_args(c, HETDDFT_TUNE_PCP);
      // reads psi, psi_old(3), two H psi slots, pc and pot4; writes psi (into
      // slot 3), one H psi slot and den
      ProfScope pspred(c, HETDDFT_PROF_PCP, (double)g.n * (16.0 * 7 + 2.0 * sizeof(T)));
      with_shape(c->tile_k[HETDDFT_TUNE_PCP], c->pf_k[HETDDFT_TUNE_PCP], [&](auto sh, auto pf) {
        using SH = decltype(sh);
        constexpr bool PF = decltype(pf)::value;
        PcpTab<T> t{};
        t.nzb = (int)grid25<SH>(c, HETDDFT_TUNE_PCP).z;
        for (int i = 0; i < G.n; ++i) {
          hetddft_ctx* m = G.m[i];
          const int* P = m->ioldp;
          const int* H = m->ioldh;
          t.timec[i] = m->timec;
          t.psi[i] = m->psi;
          t.pot4[i] = (const T*)m->fld[hetddft_ctx::F_POT4];
          t.old3[i] = m->old[P[2]];
          t.h1[i] = m->hold[H[0]];
          t.h2[i] = m->hold[H[1]];
          t.pc[i] = m->pc;
          t.den[i] = (T*)m->fld[hetddft_ctx::F_DEN];
        }
        PLAUNCH(pspred, (k_pc_predict25<T, SH, PF>), grid25m<SH>(c, HETDDFT_TUNE_PCP, G.n), dim3(SH::TX, SH::TY),
                0, hp, t, c4o3 * dt, c112);
        return 0;
      });
    }
    LAUNCHCHK(c);
    return HETDDFT_OK;
  };
  // In the lean layout the streams do not wait for each other here, so the
  // psi predictor goes first: after a host synchronisation (a host that
  // fetches the error estimates every step) the compute stream then has its
  // longest kernel queued while the side stream's launches are still being
  // issued.  The fork of the other layout must follow the particle predictor.
  if (lean) {
    RET(psi_predictor());
    RET(particle_side());
  } else {
    RET(particle_side());
    RET(psi_predictor());
  }
  for (int i = 0; i < G.n; ++i) {
    hetddft_ctx* m = G.m[i];
    int* P = m->ioldp;
    int* H = m->ioldh;
    c64* x = m->psi;
    m->psi = m->old[P[2]];
    m->old[P[2]] = x;
    std::swap(H[0], H[1]);
  }
  RET(veval<T>(G, lean));
  // Particle corrector (steppc.f90:122-151), then the density chain of the
  // corrected positions on the side stream beside the psi corrector; the
  // driver's potential evaluation (DFT4He3d.f90:861) consumes that chain.
  {
    PartCorrTab pt{};
    for (int i = 0; i < G.n; ++i) {
      hetddft_ctx* m = G.m[i];
      int* R = m->ioldr;
      int* V = m->ioldv;
      const int* A = m->iolda;
      pt.r[i] = m->rimp;
      pt.v[i] = m->vimp;
      pt.acc[i] = m->aimp;
      pt.r3[i] = m->rold[R[2]];
      pt.r2[i] = m->rold[R[1]];
      pt.v1[i] = m->vold[V[0]];
      pt.v2[i] = m->vold[V[1]];
      pt.v3[i] = m->vold[V[2]];
      pt.a1[i] = m->aold[A[0]];
      pt.a2[i] = m->aold[A[1]];
      pt.pcr[i] = m->pcr;
      pt.pcv[i] = m->pcv;
      pt.part[i] = m->part + nb;  // after the psi partials (disjoint from the psi corrector's)
      int t = R[2];
      R[2] = R[1];
      R[1] = R[0];
      R[0] = t;
      t = V[2];
      V[2] = V[1];
      V[1] = V[0];
      V[0] = t;
    }
    if (n3 > 0) {
      const unsigned pb = blocks_for(n3);
      if (lean)
        S2LAUNCH(k_part_correct, dim3(pb, G.n), dim3(kBlock), 0, n3, pt, 3.0 * dt, c9);
      else
        SLAUNCH(k_part_correct, dim3(pb, G.n), dim3(kBlock), 0, n3, pt, 3.0 * dt, c9);
      LAUNCHCHK(c);
    }
  }
  if (!lean) RET(fork_side(c));
  RET(issue_chain<T>(G));
  // psi corrector (steppc.f90:97-120), every member in one launch; the
  // error sums stay on the device until fetch_errors asks for them
  {
    const Ham25 hc = ham_args(c, HETDDFT_TUNE_PCC);
    // reads psi, two psi_old slots, two H psi slots, pc and pot4; writes pc,
    // psi and den
    ProfScope pscorr(c, HETDDFT_PROF_PCC, (double)g.n * (16.0 * 8 + 2.0 * sizeof(T)));
    with_shape(c->tile_k[HETDDFT_TUNE_PCC], c->pf_k[HETDDFT_TUNE_PCC], [&](auto sh, auto pf) {
      using SH = decltype(sh);
      constexpr bool PF = decltype(pf)::value;
      PccTab<T> t{};
      t.nzb = (int)grid25<SH>(c, HETDDFT_TUNE_PCC).z;
      for (int i = 0; i < G.n; ++i) {
        hetddft_ctx* m = G.m[i];
        const int* P = m->ioldp;
        const int* H = m->ioldh;
        t.timec[i] = m->timec;
        t.psi[i] = m->psi;
        t.pot4[i] = (const T*)m->fld[hetddft_ctx::F_POT4];
        t.o3[i] = m->old[P[2]];
        t.o2[i] = m->old[P[1]];
        t.h1[i] = m->hold[H[0]];
        t.h2[i] = m->hold[H[1]];
        t.pc[i] = m->pc;
        t.out[i] = m->spare;
        t.den[i] = (T*)m->fld[hetddft_ctx::F_DEN];
        t.part[i] = m->part;
      }
      PLAUNCH(pscorr, (k_pc_correct25<T, SH, PF>), grid25m<SH>(c, HETDDFT_TUNE_PCC, G.n), dim3(SH::TX, SH::TY),
              0, hc, t, 3.0 * dt, c9, c112);
      return 0;
    });
  }
  LAUNCHCHK(c);
  for (int i = 0; i < G.n; ++i) {
    hetddft_ctx* m = G.m[i];
    int* P = m->ioldp;
    std::swap(m->psi, m->spare);
    m->err_pending = true;
    m->err_nb = (int)nb;
    m->err_pb = n3 > 0 ? (int)blocks_for(n3) : 0;
    const int t = P[2];
    P[2] = P[1];
    P[1] = P[0];
    P[0] = t;
  }
  return HETDDFT_OK;
}

template <class T>
int step_pc_t(hetddft_ctx* c, double dt) {
  Group G = solo(c);
  return step_pc_t<T>(G, dt);
}

// Reduces the last PC step's error partials into res (see err_pending).
int flush_errors(hetddft_ctx* c) {
  if (!c->err_pending) return HETDDFT_OK;
  c->err_pending = false;
  SLAUNCH(k_finish_sums, dim3(1), dim3(kBlock), 0, (const double*)c->part, c->err_nb, 1, c->res);
  LAUNCHCHK(c);
  if (c->err_pb > 0)
    SLAUNCH(k_finish_sums, dim3(2), dim3(kBlock), 0, (const double*)(c->part + c->err_nb), c->err_pb, 2, c->res + 1);
  else
    SLAUNCH(k_fill_zero, dim3(1), dim3(64), 0, c->res + 1, 2);
  LAUNCHCHK(c);
  return HETDDFT_OK;
}

// The nine particle means (r, v, a per axis) into res[16..24], on st.
int launch_means(hetddft_ctx* c) {
  const int np = c->p.n_par;
  const unsigned nb = std::min(blocks_for(np), 64u);
  hipLaunchKernelGGL(k_part_sums, dim3(nb, 9), dim3(kBlock), 0, c->st, np, (const double*)c->rimp,
                     (const double*)c->vimp, (const double*)c->aimp, c->part);
  LAUNCHCHK(c);
  hipLaunchKernelGGL(k_finish_sums, dim3(9), dim3(kBlock), 0, c->st, (const double*)c->part, (int)nb, 9,
                     c->res + 16);
  LAUNCHCHK(c);
  return HETDDFT_OK;
}

// The means written by the reference's driver (DFT4He3d.f90:864-878) from
// the sums in hres[16..24]: positions in A, velocities in A/ps,
// accelerations in A/ps^2.
void means_from_host(const hetddft_ctx* c, double out[9]) {
  const int np = c->p.n_par;
  const double pstok = 1.0 / 7.63822291;
  for (int k = 0; k < 9; ++k) {
    double v = c->hres[16 + k] / np;
    if (k >= 3 && k < 6) v = v * pstok;
    if (k >= 6) v = v * pstok * pstok;
    out[k] = v;
  }
}

int fetch_errors(hetddft_ctx* c, double err[3]) {
  RET(settle(c));
  RET(flush_errors(c));
  // the means too when the last fetch was followed by a particle_means call
  // (one wait for the GPU instead of two); the error partials are reduced
  // before the means reuse the partials buffer (same stream)
  const bool means = c->means_wanted && c->p.n_par > 0;
  c->means_wanted = false;
  if (means) RET(launch_means(c));
  HIPCHK(c, hipMemcpyAsync(c->hres, c->res, (means ? 25 : 3) * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (means) {
    means_from_host(c, c->means_cache);
    c->means_ready = true;
  }
  const double c9 = 9.0 / 121.0;
  const double e0 = c9 * c->hres[0] / (double)c->g.n;
  const double e1 = c->hres[1] * 0.3333333333 / (double)c->p.n_par;
  const double e2 = c->hres[2] * 0.3333333333 / (double)c->p.n_par;
  c->last_err[0] = e0;
  c->last_err[1] = c->p.n_par > 0 ? e1 : 0.0;
  c->last_err[2] = c->p.n_par > 0 ? e2 : 0.0;
  if (err) std::memcpy(err, c->last_err, sizeof c->last_err);
  if (!std::isfinite(e0) || !std::isfinite(c->last_err[1]) || !std::isfinite(c->last_err[2]))
    return fail(c, HETDDFT_E_NONFINITE, "non-finite PC error estimate (errHe=%g)", e0);
  return HETDDFT_OK;
}

template <class T>
int field_to_host(hetddft_ctx* c, const void* dsrc, double* out) {
  const size_t n = (size_t)c->g.n;
  if (sizeof(T) == sizeof(double)) {
    HIPCHK(c, hipMemcpyAsync(out, dsrc, n * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
  } else {
    std::vector<float> tmp(n);
    HIPCHK(c, hipMemcpyAsync(tmp.data(), dsrc, n * sizeof(float), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (size_t i = 0; i < n; ++i) out[i] = tmp[i];
  }
  return HETDDFT_OK;
}

int upload_real(hetddft_ctx* c, void* dst, const double* src, size_t n, bool as_float) {
  if (!as_float) {
    HIPCHK(c, hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<float> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = (float)src[i];
    HIPCHK(c, hipMemcpy(dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice));
  }
  return HETDDFT_OK;
}

// k-space kernel from the Fortran (mx,ny,nz) layout into the internal
// spectrum order [kz][kx][ky]
int upload_kernel(hetddft_ctx* c, void* dst, const double* src, bool as_float) {
  const Grid& g = c->g;
  std::vector<double> t((size_t)g.h);
  for (int iz = 0; iz < g.nz; ++iz)
    for (int iy = 0; iy < g.ny; ++iy)
      for (int ix = 0; ix < g.mx; ++ix)
        t[((size_t)iz * g.mx + ix) * g.ny + iy] = src[((size_t)iz * g.ny + iy) * g.mx + ix];
  return upload_real(c, dst, t.data(), t.size(), as_float);
}

template <class T>
__global__ void k_to_double(const T* __restrict__ a, double* __restrict__ b, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}

template <class T>
int energy_t(hetddft_ctx* c, double out[12]) {
  const Grid& g = c->g;
  auto F = [&](int id) { return (const T*)c->fld[id]; };
  // The density-side terms (kinetic aside) are large and cancel; they are
  // evaluated on FP64 fields: in mixed mode den is rebuilt from psi and the
  // convolutions behind uimp, dencg and delj4 are redone in FP64.
  const double *den64, *dencg64, *delj464, *uimp64;
  if (c->mixed) {
    double* d64 = (double*)c->e64[0];
    double* dx64 = (double*)c->e64[1];
    hipLaunchKernelGGL((k_density<double>), dim3(blocks_for(g.n)), dim3(kBlock), 0, c->st, (const c64*)c->psi, d64,
                       g.n);
    LAUNCHCHK(c);
    hipLaunchKernelGGL((k_to_double<T>), dim3(blocks_for(g.n)), dim3(kBlock), 0, c->st, (const T*)c->denx, dx64, g.n);
    LAUNCHCHK(c);
    XfArgs<double> xa{};
    xa.op[0] = XF_COPY;
    xa.in[0][0] = dx64;
    xa.op[1] = XF_COPY;
    xa.in[1][0] = d64;
    std::vector<ConvJob<double>> jobs;
    jobs.push_back({0, c->fs64e.kvq, 1.0, c->e64[4]});
    jobs.push_back({1, c->fs64e.kwcgk, 1.0, c->e64[2]});
    jobs.push_back({1, c->fs64e.kfvlj4, 1.0, c->e64[3]});
    RET(conv_batch<double>(c, c->fs64e, xa, 2, jobs));
    den64 = d64;
    dencg64 = (const double*)c->e64[2];
    delj464 = (const double*)c->e64[3];
    uimp64 = (const double*)c->e64[4];
  } else {
    den64 = (const double*)c->fld[hetddft_ctx::F_DEN];
    dencg64 = (const double*)c->fld[hetddft_ctx::F_DENCG];
    delj464 = (const double*)c->fld[hetddft_ctx::F_DELJ4];
    uimp64 = (const double*)c->fld[hetddft_ctx::F_UIMP];
  }
  EnergyArgs ea{};
  ea.g = g;
  ea.d1x = c->d1x;
  ea.d1y = c->d1y;
  ea.d1z = c->d1z;
  ea.aux1 = 0.5 * c->p.cp4;
  ea.aux2 = (1.0 / 3.0) * c->p.cpp4;
  ea.den0s = c->p.den0s;
  ea.beta = c->p.beta;
  ea.den_m = c->p.den_m;
  ea.otc = c->p.core_otc;
  const unsigned nb = blocks_for(g.n);
  hipLaunchKernelGGL((k_energy<T>), dim3(nb), dim3(kBlock), 0, c->st, ea, (const c64*)c->psi, den64, dencg64,
                     delj464, F(hetddft_ctx::F_FALFS), F(hetddft_ctx::F_DX), F(hetddft_ctx::F_DY),
                     F(hetddft_ctx::F_DZ), F(hetddft_ctx::F_IX), F(hetddft_ctx::F_IY), F(hetddft_ctx::F_IZ),
                     F(hetddft_ctx::F_PSIX), F(hetddft_ctx::F_HPSIX), uimp64, c->part);
  LAUNCHCHK(c);
  hipLaunchKernelGGL(k_finish_sums, dim3(8), dim3(kBlock), 0, c->st, (const double*)c->part, (int)nb, 8, c->res + 4);
  LAUNCHCHK(c);
  const int n3 = 3 * c->p.n_par;
  if (n3 > 0) {
    const unsigned pb = std::min(blocks_for(n3), 1024u);
    hipLaunchKernelGGL(k_sq_sum, dim3(pb), dim3(kBlock), 0, c->st, n3, (const double*)c->vimp, c->part + 8 * nb);
    LAUNCHCHK(c);
    hipLaunchKernelGGL(k_finish_sums, dim3(1), dim3(kBlock), 0, c->st, (const double*)(c->part + 8 * nb), (int)pb, 1,
                       c->res + 12);
    LAUNCHCHK(c);
  } else {
    HIPCHK(c, hipMemsetAsync(c->res + 12, 0, sizeof(double), c->st));
  }
  HIPCHK(c, hipMemcpyAsync(c->hres + 4, c->res + 4, 9 * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  const double* s = c->hres + 4;
  const hetddft_params& p = c->p;
  const double dxyz = c->dxyz;
  const double ekin4 = p.h2o2m4 * s[0] * dxyz;
  const double elj4 = 0.5 * s[1] * dxyz;
  const double ecor4 = s[2] * dxyz;
  const double ealphas = p.core_otc ? -p.h2o2m4 * 0.5 * p.alphas * s[3] * dxyz : 0.0;
  const double esolid = p.C * s[4] * dxyz;
  double etot4 = ekin4 + elj4 + ecor4 + esolid;
  if (p.core_otc) etot4 = etot4 + ealphas;
  const double ekinx = p.n_par > 0 ? 0.5 * c->m_imp * s[8] / p.n_par : 0.0;
  const double ekinqx = s[5] * dxyz;
  const double ehex = s[6] * dxyz;
  const double eimpu = ekinx + ehex + ekinqx;
  const double vals[12] = {etot4 + eimpu, etot4, ekin4, elj4, ealphas, esolid, ecor4, eimpu, ekinx, ekinqx, ehex,
                           s[7] * dxyz};
  std::memcpy(out, vals, sizeof vals);
  return HETDDFT_OK;
}

}  // namespace

// ---- allocation-free particle means -------------------------------------
// Device scratch of hetddft_particle_means_dev, the entry point that
// allocates nothing: kNoAllocParts partial sums for each of the 9 components,
// then the 9 sums.  Every GPU holds its own copy; host code reaches it only
// through the address hipGetSymbolAddress returns on that GPU.
constexpr int kNoAllocParts = 64;
constexpr int kScratchDoubles = 1024;
static_assert(9 * kNoAllocParts + 9 <= kScratchDoubles, "particle-means scratch too small");
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// Address of this GPU's scratch, looked up once per device (current device).
int scratch_of_current_device(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    g_create_error = std::string("hipGetDevice: ") + hipGetErrorString(e);
    return HETDDFT_E_HIP;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    g_create_error = "device ordinal beyond the scratch table";
    return HETDDFT_E_ARG;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      g_create_error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return HETDDFT_E_HIP;
    }
    g_scratch_dev[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_dev[dev];
  return HETDDFT_OK;
}

}  // namespace

// ============================ C-ABI ======================================
// Host state that a PC step and the following potential evaluation advance:
// the rotating psi / H psi / particle slots, the chain set, the flags, and
// the step size the kernels are given.
struct StepState {
  c64* psi;
  c64* spare;
  c64* old[3];
  c64* hold[2];
  int ioldp[3], ioldh[2], ioldr[3], ioldv[3], iolda[2];
  void* fld[hetddft_ctx::F_COUNT];
  void* denx;
  unsigned int* counts;
  int* reg;
  void* csdenx[2];
  int cs;
  bool chain_pending, side_pending;
  double dt;
};

StepState save_step_state(const hetddft_ctx* c, double dt) {
  StepState s;
  std::memset(&s, 0, sizeof s);  // the key below compares bytes, padding included
  s.psi = c->psi;
  s.spare = c->spare;
  std::copy(c->old, c->old + 3, s.old);
  std::copy(c->hold, c->hold + 2, s.hold);
  std::copy(c->ioldp, c->ioldp + 3, s.ioldp);
  std::copy(c->ioldh, c->ioldh + 2, s.ioldh);
  std::copy(c->ioldr, c->ioldr + 3, s.ioldr);
  std::copy(c->ioldv, c->ioldv + 3, s.ioldv);
  std::copy(c->iolda, c->iolda + 2, s.iolda);
  std::copy(c->fld, c->fld + hetddft_ctx::F_COUNT, s.fld);
  s.denx = c->denx;
  s.counts = c->counts;
  s.reg = c->reg;
  s.csdenx[0] = c->cset[0].denx;
  s.csdenx[1] = c->cset[1].denx;
  s.cs = c->cs;
  s.chain_pending = c->chain_pending;
  s.side_pending = c->side_pending;
  s.dt = dt;
  return s;
}

void load_step_state(hetddft_ctx* c, const StepState& s) {
  c->psi = s.psi;
  c->spare = s.spare;
  std::copy(s.old, s.old + 3, c->old);
  std::copy(s.hold, s.hold + 2, c->hold);
  std::copy(s.ioldp, s.ioldp + 3, c->ioldp);
  std::copy(s.ioldh, s.ioldh + 2, c->ioldh);
  std::copy(s.ioldr, s.ioldr + 3, c->ioldr);
  std::copy(s.ioldv, s.ioldv + 3, c->ioldv);
  std::copy(s.iolda, s.iolda + 2, c->iolda);
  std::copy(s.fld, s.fld + hetddft_ctx::F_COUNT, c->fld);
  c->denx = s.denx;
  c->counts = s.counts;
  c->reg = s.reg;
  c->cset[0].denx = s.csdenx[0];
  c->cset[1].denx = s.csdenx[1];
  c->cs = s.cs;
  c->chain_pending = s.chain_pending;
  c->side_pending = s.side_pending;
}

void clear_graphs(hetddft_ctx* c) {
  for (auto& kv : c->graphs) hipGraphExecDestroy(kv.second);
  for (hipGraph_t g : c->graph_src) hipGraphDestroy(g);
  c->graphs.clear();
  c->graph_src.clear();
}

template <class T>
int pc_iteration(Group& G, double dt) {
  hetddft_ctx* c = G.m[0];
  RET(step_pc_t<T>(G, dt));
  return veval<T>(G, c->lean && !c->one_stream);
}
template <class T>
int pc_iteration(hetddft_ctx* c, double dt) {
  Group G = solo(c);
  return pc_iteration<T>(G, dt);
}

// One PC step and the driver's potential evaluation.  With graphs on
// (HETDDFT_GRAPH=1 or 2 at create), the launches of a given slot state are
// captured once and replayed afterwards; the host logic then runs dry to
// advance the slots.  Mode 1 captures both streams (the side stream joins
// through the fork event and rejoins where the potential evaluation waits for
// the particle chain), mode 2 puts the particle chain on st.  A capture that
// fails leaves graphs off for this instance and runs the step directly.
// Graphs are off by default: on ROCm 7.2 / MI355X their replay was slower
// than direct launches (128^3 mixed: 0.39x with both streams, 0.92x on one;
// the 32-instance ensemble: 0.44x).
int iterate_pc(hetddft_ctx* c, double dt) {
  auto direct = [&]() { return c->mixed ? pc_iteration<float>(c, dt) : pc_iteration<double>(c, dt); };
  if (!c->graph_on || c->prof_on || c->chain_pending || c->side_pending) return direct();
  const StepState s0 = save_step_state(c, dt);
  const std::string key(reinterpret_cast<const char*>(&s0), sizeof s0);
  auto it = c->graphs.find(key);
  if (it != c->graphs.end()) {
    c->dry = true;
    const int rc = direct();
    c->dry = false;
    if (rc != HETDDFT_OK) {
      load_step_state(c, s0);
      return rc;
    }
    HIPCHK(c, hipGraphLaunch(it->second, c->st));
    return HETDDFT_OK;
  }
  if (c->graphs.size() >= 12) clear_graphs(c);  // a new step size, or set_state
  hipGraph_t gr = nullptr;
  hipGraphExec_t ex = nullptr;
  bool ok = hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal) == hipSuccess;
  if (ok) {
    c->one_stream = c->graph_one_stream;
    c->capturing = true;
    const int rc = direct();
    c->capturing = false;
    c->one_stream = false;
    const hipError_t e = hipStreamEndCapture(c->st, &gr);
    ok = rc == HETDDFT_OK && e == hipSuccess && gr != nullptr &&
         hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0) == hipSuccess;
  }
  if (!ok) {
    if (ex) hipGraphExecDestroy(ex);
    if (gr) hipGraphDestroy(gr);
    (void)hipGetLastError();
    load_step_state(c, s0);
    c->graph_on = false;
    return direct();
  }
  c->graphs[key] = ex;
  c->graph_src.push_back(gr);
  HIPCHK(c, hipGraphLaunch(ex, c->st));
  return HETDDFT_OK;
}

// ---- full-history checkpoint ------------------------------------------------
// Header, then the device arrays in a fixed order: psi; its PC slots (psi
// history x3, H psi history x2, pc) and the RK accumulator q, by buffer
// index; the particles r, v, a; their slots r x3, v x3, a x2; qr, qv, pcr,
// pcv.  The slot order (ioldp, ioldh, ioldr, ioldv, iolda) and the step
// counter are in the header.
struct CkptHeader {
  char magic[8];  // "HETDCKP1"
  int nx, ny, nz, n_par, precision, iter_in_run;
  int ioldp[3], ioldh[2], ioldr[3], ioldv[3], iolda[2];
  int reserved[8];
};

struct CkptPart {
  void* dev;
  size_t bytes;
};

std::vector<CkptPart> ckpt_parts(hetddft_ctx* c) {
  const size_t n = (size_t)c->g.n * 16, np3 = (size_t)3 * c->p.n_par * 8;
  std::vector<CkptPart> v;
  v.push_back({c->psi, n});
  for (int k = 0; k < 3; ++k) v.push_back({c->old[k], n});
  for (int k = 0; k < 2; ++k) v.push_back({c->hold[k], n});
  v.push_back({c->pc, n});
  v.push_back({c->q, n});
  if (np3) {
    for (double* a : {c->rimp, c->vimp, c->aimp}) v.push_back({a, np3});
    for (int k = 0; k < 3; ++k) v.push_back({c->rold[k], np3});
    for (int k = 0; k < 3; ++k) v.push_back({c->vold[k], np3});
    for (int k = 0; k < 2; ++k) v.push_back({c->aold[k], np3});
    for (double* a : {c->qr, c->qv, c->pcr, c->pcv}) v.push_back({a, np3});
  }
  return v;
}

long long ckpt_bytes(hetddft_ctx* c) {
  long long b = sizeof(CkptHeader);
  for (const CkptPart& q : ckpt_parts(c)) b += (long long)q.bytes;
  return b;
}

DiagArgs diag_args(const hetddft_ctx* c) {
  DiagArgs a{};
  a.g = c->g;
  a.d1x = c->d1x;
  a.d1y = c->d1y;
  a.d1z = c->d1z;
  a.d2x = c->d2x;
  a.d2y = c->d2y;
  a.d2z = c->d2z;
  a.xmin = -c->p.xmax;
  a.ymin = -c->p.ymax;
  a.zmin = -c->p.zmax;
  a.hx = c->hx;
  a.hy = c->hy;
  a.hz = c->hz;
  a.h2o2m4 = c->p.h2o2m4;
  return a;
}

// ---- launch-shape tuning ----------------------------------------------
template <class T>
int fft_z_tile_of(hetddft_ctx* c) {
  const FftKernels<T>& k = fft_of<T>(c);
  return k.nzv > 0 ? k.typ : 0;
}
int fft_z_tile(hetddft_ctx* c) { return c->mixed ? fft_z_tile_of<float>(c) : fft_z_tile_of<double>(c); }

// Selects the register-resident z transform with ky tile width w; false if
// this grid offers no such kernel.
template <class T>
bool select_fft_z_tile(hetddft_ctx* c, int w) {
  FftKernels<T>& k = fft_of<T>(c);
  for (int v = 0; v < k.nzv; ++v)
    if (k.typv[v] == w) {
      k.zp = k.zpv[v];
      k.lds_zp = k.lds_zpv[v];
      k.typ = k.typv[v];
      return true;
    }
  return false;
}

void tuning_of(hetddft_ctx* c, hetddft_tuning* t) {
  for (int k = 0; k < HETDDFT_TUNE_NSTENCIL; ++k) {
    t->tile[k] = c->tile_k[k];
    t->zc[k] = c->zc_k[k];
    t->prefetch[k] = c->pf_k[k];
  }
  t->fft_z_tile = fft_z_tile(c);
  t->fft_z_wgs = c->zp_per_cu;
}

int apply_tuning(hetddft_ctx* c, const hetddft_tuning* t) {
  for (int k = 0; k < HETDDFT_TUNE_NSTENCIL; ++k)
    if (t->tile[k] < 0 || t->tile[k] > 3 || t->zc[k] < 1 || (t->prefetch[k] != 0 && t->prefetch[k] != 1))
      return fail(c, HETDDFT_E_ARG, "tuning: class %d has tile %d, zc %d, prefetch %d", k, t->tile[k], t->zc[k],
                  t->prefetch[k]);
  {
    // the corrector's errHe partials (one per workgroup) must fit the
    // partial-sum buffer
    const int k = HETDDFT_TUNE_PCC;
    const dim3 d = with_tile(t->tile[k], [&](auto sh) {
      return grid25z<decltype(sh)>(c->g, std::min(t->zc[k], c->g.nz));
    });
    if ((long long)d.x * d.y * d.z > (long long)kDiagSlots * blocks_for(c->g.n))
      return fail(c, HETDDFT_E_ARG, "tuning: corrector grid of %u workgroups is too large", d.x * d.y * d.z);
  }
  if (t->fft_z_wgs < 0 || t->fft_z_wgs > 8)
    return fail(c, HETDDFT_E_ARG, "tuning: fft_z_wgs %d outside 0..8", t->fft_z_wgs);
  if (t->fft_z_tile != 0) {
    const bool ok = c->mixed ? select_fft_z_tile<float>(c, t->fft_z_tile) : select_fft_z_tile<double>(c, t->fft_z_tile);
    if (!ok) return fail(c, HETDDFT_E_ARG, "tuning: no z transform with ky tile %d on this grid", t->fft_z_tile);
  }
  for (int k = 0; k < HETDDFT_TUNE_NSTENCIL; ++k) {
    c->tile_k[k] = t->tile[k];
    c->zc_k[k] = std::min(t->zc[k], c->g.nz);
    c->pf_k[k] = t->prefetch[k];
  }
  if (t->fft_z_wgs > 0) c->zp_per_cu = t->fft_z_wgs;
  clear_graphs(c);  // captured steps hold the old launch shapes
  return HETDDFT_OK;
}

// z chunks giving about 512, 768, 1024 and 2048 workgroups for a tile shape
// (2..32 planes each), without repeats.
std::vector<int> zc_candidates(const hetddft_ctx* c, int tile) {
  const long long cols = with_tile(tile, [&](auto sh) {
    using SH = decltype(sh);
    return (long long)((c->g.nx + SH::TX - 1) / SH::TX) * ((c->g.ny + SH::TY - 1) / SH::TY);
  });
  std::vector<int> out;
  for (long long target : {512LL, 768LL, 1024LL, 2048LL}) {
    long long zc = (cols * c->g.nz + target - 1) / target;
    zc = std::max(2LL, std::min(32LL, std::min<long long>(zc, c->g.nz)));
    if (std::find(out.begin(), out.end(), (int)zc) == out.end()) out.push_back((int)zc);
  }
  return out;
}

// Mean ms per launch of each profiled class over `steps` PC iterations
// (one untimed iteration first).
static int time_classes(hetddft_ctx* c, double dt, int steps, double ms[HETDDFT_PROF_N]);
// Wall ms per PC iteration over `steps` iterations (one untimed first).
static int time_steps(hetddft_ctx* c, double dt, int steps, double* ms);

static std::mutex g_tune_mu;
static std::map<std::string, hetddft_tuning> g_tune_cache;

static std::string tune_key(const hetddft_ctx* c) {
  char b[128];
  std::snprintf(b, sizeof b, "%d/%d/%d/%d/%d/%d", c->p.nx, c->p.ny, c->p.nz, c->p.precision, c->p.n_par,
                c->p.core_otc);
  return b;
}

// ---- ensemble executor -------------------------------------------------
struct hetddft_batch {
  std::vector<hetddft_ctx*> ctx;
  // instances that can advance in shared launches (same grid, parameters and
  // launch shapes), at most kMaxMembers each
  std::vector<std::vector<int>> groups;
  std::vector<std::vector<int>> part;  // groups of each host thread
  std::string err;
};

namespace {

// Contexts with equal keys hold the same grid, parameters and launch shapes,
// so one step's launches can carry all of them.
std::string group_key(const hetddft_ctx* c) {
  std::string k(reinterpret_cast<const char*>(&c->p), sizeof c->p);
  int v[4 * HETDDFT_TUNE_NSTENCIL + 8];
  int n = 0;
  for (int i = 0; i < HETDDFT_TUNE_NSTENCIL; ++i) {
    v[n++] = c->tile_k[i];
    v[n++] = c->zc_k[i];
    v[n++] = c->pf_k[i];
  }
  v[n++] = c->zp_per_cu;
  v[n++] = c->mixed ? fft_z_tile_of<float>(const_cast<hetddft_ctx*>(c)) : fft_z_tile_of<double>(const_cast<hetddft_ctx*>(c));
  v[n++] = c->xcd;
  v[n++] = c->iron_fused ? 1 : 0;
  v[n++] = c->lean ? 1 : 0;
  v[n++] = c->graph_on ? 1 : 0;
  k.append(reinterpret_cast<const char*>(v), n * sizeof(int));
  return k;
}

// Members at the same step (step counter, whether a chain is queued) and
// with the same step size advance as one group; each keeps its own chain
// set and slot order.
bool same_state(const hetddft_ctx* a, const hetddft_ctx* b) {
  return a->started && b->started && a->iter_in_run == b->iter_in_run && a->chain_pending == b->chain_pending &&
         !a->prof_on && !b->prof_on;
}

// One iteration of the reference loop (DFT4He3d.f90:849-863) for every
// member, with shared launches; as hetddft_iterate with err = NULL.
int iterate_group(Group& G, double dt) {
  hetddft_ctx* c = G.m[0];
  if (c->iter_in_run < 3) {
    RET(settle(c));
    for (int i = 0; i < G.n; ++i) G.m[i]->chain_pending = false;
    RET(c->mixed ? step_rk_t<float>(G, dt) : step_rk_t<double>(G, dt));
    RET(c->mixed ? veval<float>(G) : veval<double>(G));
  } else {
    RET(c->mixed ? pc_iteration<float>(G, dt) : pc_iteration<double>(G, dt));
  }
  for (int i = 0; i < G.n; ++i) G.m[i]->iter_in_run++;
  return HETDDFT_OK;
}

}  // namespace

extern "C" {

int hetddft_default_params(hetddft_params* p) {
  if (!p) return HETDDFT_E_ARG;
  std::memset(p, 0, sizeof *p);
  p->nx = p->ny = p->nz = 64;
  p->xmax = p->ymax = p->zmax = 16.0;
  p->n_par = 1000;
  p->mimpur = 40.0;
  p->ironx = 3;
  p->core_otc = 1;
  p->lambdah = 0.0;
  p->tzmean = 0.0;
  p->tzsurf = 1.0;
  p->cp4 = -2.41186e4;
  p->cpp4 = 1.85850e6;
  p->alphas = 54.31;
  p->l = 1.0;
  p->den0s = 0.04;
  p->h2o2m4 = 6.05969638298;
  p->C = 3.1577504e4;
  p->beta = 40.0;
  p->den_m = 0.37;
  p->h4 = 2.190323;
  p->eps4 = 10.22;
  p->sigma4 = 2.556;
  p->precision = HETDDFT_FP64;
  return HETDDFT_OK;
}

const char* hetddft_last_error(const hetddft_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

long long hetddft_device_bytes(const hetddft_ctx* ctx) { return ctx ? ctx->bytes : 0; }

int hetddft_destroy(hetddft_ctx* c) {
  if (!c) return HETDDFT_OK;
  if (c->st) hipStreamSynchronize(c->st);
  if (c->st2) hipStreamSynchronize(c->st2);
  clear_graphs(c);
  for (hipEvent_t e : {c->ev_fork, c->ev_chain, c->ev_b3, c->ev_join, c->ev_force})
    if (e) hipEventDestroy(e);

  if (c->st2) hipStreamDestroy(c->st2);
  for (auto& r : c->prof_pending) {
    if (r.a) hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  for (hipEvent_t e : c->prof_pool) hipEventDestroy(e);
  for (void* p : c->allocs) hipFree(p);
  if (c->hres) hipHostFree(c->hres);
  if (c->st) hipStreamDestroy(c->st);
  delete c;
  return HETDDFT_OK;
}

int hetddft_create(const hetddft_params* p, hetddft_ctx** out) {
  if (!p || !out) return HETDDFT_E_ARG;
  *out = nullptr;
  hetddft_ctx* c = new hetddft_ctx();
  c->p = *p;
  auto bad = [&](int code) {
    g_create_error = c->err;  // readable through hetddft_last_error(NULL)
    hetddft_destroy(c);
    return code;
  };
  if (!tstgrid_ok(p->nx) || !tstgrid_ok(p->ny) || !tstgrid_ok(p->nz) || (p->nx % 2) != 0 || p->nx < 4 ||
      p->ny < 2 || p->nz < 2) {
    fail(c, HETDDFT_E_ARG, "grid %dx%dx%d not supported (tstgrid list, nx even)", p->nx, p->ny, p->nz);
    return bad(HETDDFT_E_ARG);
  }
  // particle kernels address 3 N_par doubles with 32-bit byte offsets
  if (p->n_par < 0 || p->n_par > kMaxNpar || p->ironx < 0 || p->mimpur <= 0 || !(p->xmax > 0) || !(p->ymax > 0) || !(p->zmax > 0) ||
      (p->precision != HETDDFT_FP64 && p->precision != HETDDFT_MIXED) || p->tzsurf == 0.0) {
    fail(c, HETDDFT_E_ARG, "invalid parameters");
    return bad(HETDDFT_E_ARG);
  }
  c->mixed = p->precision == HETDDFT_MIXED;
  c->tsz = c->mixed ? 4 : 8;
  Grid& g = c->g;
  g.nx = p->nx;
  g.ny = p->ny;
  g.nz = p->nz;
  g.mx = p->nx / 2 + 1;
  g.n = (long long)p->nx * p->ny * p->nz;
  g.mp = g.mx;
  g.h = (long long)g.mx * p->ny * p->nz;
  c->hx = 2.0 * std::fabs(p->xmax) / p->nx;
  c->hy = 2.0 * std::fabs(p->ymax) / p->ny;
  c->hz = 2.0 * std::fabs(p->zmax) / p->nz;
  c->dxyz = c->hx * c->hy * c->hz;
  c->renor = 1.0 / ((double)p->nx * p->ny * p->nz);
  c->m_imp = p->mimpur * 0.0206149591055;
  c->h2o2mx = p->h2o2m4 * (4.002602 / p->mimpur);
  c->d1x = make_stencil(1, c->hx);
  c->d1y = make_stencil(1, c->hy);
  c->d1z = make_stencil(1, c->hz);
  c->d2x = make_stencil(2, c->hx);
  c->d2y = make_stencil(2, c->hy);
  c->d2z = make_stencil(2, c->hz);
  {
    // tile shape, and a z chunk giving about HETDDFT_ZBLOCKS workgroups
    // (4..32 planes each)
    const char* t = std::getenv("HETDDFT_TILE");
    if (t) c->tile = std::max(0, std::min(3, std::atoi(t)));
    const long long target = 1024;
    const long long cols = with_tile(c->tile, [&](auto sh) {
      using SH = decltype(sh);
      return (int)(((p->nx + SH::TX - 1) / SH::TX) * ((p->ny + SH::TY - 1) / SH::TY));
    });
    long long zc = cols * p->nz / target;
    c->zc = (int)std::max(4LL, std::min(32LL, zc));
    const char* z = std::getenv("HETDDFT_ZC");
    if (z && std::atoi(z) > 0) c->zc = std::atoi(z);
    const char* x = std::getenv("HETDDFT_XCD");
    if (x) c->xcd = std::atoi(x) != 0;
    const char* fi = std::getenv("HETDDFT_IRON_FUSED");
    if (fi) c->iron_fused = std::atoi(fi) != 0;
    const char* gph = std::getenv("HETDDFT_GRAPH");
    c->graph_on = gph && std::atoi(gph) != 0;
    c->graph_one_stream = gph && std::atoi(gph) == 2;
    c->lean = !c->graph_on;
    for (int k = 0; k < HETDDFT_TUNE_NSTENCIL; ++k) {
      c->tile_k[k] = c->tile;
      c->zc_k[k] = c->zc;
    }
  }
  c->pl_xh = make_plan(p->nx / 2);
  c->pl_y = make_plan(p->ny);
  c->pl_z = make_plan(p->nz);
  if (c->pl_xh.nstages < 0 || c->pl_y.nstages < 0 || c->pl_z.nstages < 0) {
    fail(c, HETDDFT_E_ARG, "no radix plan for grid %dx%dx%d", p->nx, p->ny, p->nz);
    return bad(HETDDFT_E_ARG);
  }
  if (hipStreamCreate(&c->st) != hipSuccess) {
    fail(c, HETDDFT_E_HIP, "hipStreamCreate failed");
    return bad(HETDDFT_E_HIP);
  }
  if (hipStreamCreate(&c->st2) != hipSuccess) {
    fail(c, HETDDFT_E_HIP, "hipStreamCreate failed");
    return bad(HETDDFT_E_HIP);
  }
  // Cross-stream ordering between the two streams of this device: a
  // device-scope release is enough (no system-scope cache writeback per
  // record).
  const unsigned evflags = hipEventDisableTiming | hipEventDisableSystemFence;
  for (hipEvent_t* e : {&c->ev_fork, &c->ev_chain, &c->ev_b3, &c->ev_join, &c->ev_force})
    if (hipEventCreateWithFlags(e, evflags) != hipSuccess) {
      fail(c, HETDDFT_E_HIP, "hipEventCreate failed");
      return bad(HETDDFT_E_HIP);
    }
  {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
      c->cus = cus;
  }
  int rc = c->mixed ? fft_alloc_tables<float>(c, c->fs) : fft_alloc_tables<double>(c, c->fs);
  if (rc) return bad(rc);
  if (c->mixed) rc = fft_alloc_tables<double>(c, c->fs64e);
  if (rc) return bad(rc);
  const size_t n = (size_t)g.n, h = (size_t)g.h;
  auto A = [&](void** ptr, size_t bytes) { return dalloc(c, ptr, bytes); };
  std::vector<std::pair<void**, size_t>> want;
  want.push_back({(void**)&c->psi, n * 16});
  want.push_back({(void**)&c->spare, n * 16});
  for (int k = 0; k < 3; ++k) want.push_back({(void**)&c->old[k], n * 16});
  for (int k = 0; k < 2; ++k) want.push_back({(void**)&c->hold[k], n * 16});
  want.push_back({(void**)&c->pc, n * 16});
  want.push_back({(void**)&c->q, n * 16});
  for (int k = 0; k < hetddft_ctx::F_COUNT; ++k) want.push_back({&c->fld[k], n * c->tsz});
  want.push_back({(void**)&c->counts, n * 4});
  want.push_back({(void**)&c->reg, kRegLen * sizeof(int)});
  want.push_back({(void**)&c->fgather, (size_t)24 * std::max(p->n_par, 1) * 8});
  want.push_back({(void**)&c->hpart, (size_t)kBoxPart * blocks_for(std::max(p->n_par, 1)) * sizeof(int)});
  for (int k = 0; k < kSpectra; ++k) want.push_back({&c->fs.spec[k], h * 2 * c->tsz});
  want.push_back({&c->fs.kfvlj4, h * c->tsz});
  want.push_back({&c->fs.kwcgk, h * c->tsz});
  want.push_back({&c->fs.kkalfs, h * c->tsz});
  want.push_back({&c->fs.kvq, h * c->tsz});
  if (c->mixed) {
    // FP64 re-evaluation of the energy functional: 2 inputs + 3 outputs
    for (int k = 0; k < 5; ++k) want.push_back({&c->fs64e.spec[k], h * 16});
    want.push_back({&c->fs64e.kfvlj4, h * 8});
    want.push_back({&c->fs64e.kwcgk, h * 8});
    want.push_back({&c->fs64e.kvq, h * 8});
    for (int k = 0; k < 5; ++k) want.push_back({&c->e64[k], n * 8});
  }
  const size_t np3 = (size_t)3 * std::max(p->n_par, 1);
  for (double** pp : {&c->rimp, &c->vimp, &c->aimp, &c->qr, &c->qv, &c->pcr, &c->pcv})
    want.push_back({(void**)pp, np3 * 8});
  for (int k = 0; k < 3; ++k) want.push_back({(void**)&c->rold[k], np3 * 8});
  for (int k = 0; k < 3; ++k) want.push_back({(void**)&c->vold[k], np3 * 8});
  for (int k = 0; k < 2; ++k) want.push_back({(void**)&c->aold[k], np3 * 8});
  c->part_len = (long long)kDiagSlots * blocks_for(g.n) + 2LL * blocks_for((long long)np3) + 2048;
  want.push_back({(void**)&c->part, (size_t)c->part_len * 8});
  want.push_back({(void**)&c->res, 64 * 8});
  want.push_back({(void**)&c->prof_dev, HETDDFT_PROF_N * sizeof(unsigned long long)});
  want.push_back({(void**)&c->timec, (size_t)g.nz * 16});
  for (auto& w : want) {
    rc = A(w.first, w.second);
    if (rc) return bad(rc);
  }
  if (hipHostMalloc((void**)&c->hres, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    fail(c, HETDDFT_E_NOMEM, "hipHostMalloc failed");
    return bad(HETDDFT_E_NOMEM);
  }
  std::memset(c->hres, 0, 64 * sizeof(double));
  {
    // two particle-chain buffer sets; set 0 uses the fields allocated above
    hetddft_ctx::ChainSet& s0 = c->cset[0];
    s0.denxa = c->fld[hetddft_ctx::F_DENXA];
    s0.denxb = c->fld[hetddft_ctx::F_DENXB];
    s0.psix = c->fld[hetddft_ctx::F_PSIX];
    s0.hpsix = c->fld[hetddft_ctx::F_HPSIX];
    s0.counts = c->counts;
    s0.reg = c->reg;
    s0.denx = s0.denxa;
    hetddft_ctx::ChainSet& s1 = c->cset[1];
    for (void** f : {&s1.denxa, &s1.denxb, &s1.psix, &s1.hpsix}) {
      rc = A(f, n * c->tsz);
      if (rc) return bad(rc);
    }
    rc = A((void**)&s1.counts, n * 4);
    if (rc) return bad(rc);
    rc = A((void**)&s1.reg, kRegLen * sizeof(int));
    if (rc) return bad(rc);
    s1.denx = s1.denxa;
    // occupied region: nothing live yet
    int r0[kRegLen];
    for (int k = 0; k < kRegLen; ++k) r0[k] = k % 6 < 3 ? 0x7fffffff : -1;
    for (int* rg : {s0.reg, s1.reg})
      if (hipMemcpy(rg, r0, sizeof r0, hipMemcpyHostToDevice) != hipSuccess) {
        fail(c, HETDDFT_E_HIP, "region init failed");
        return bad(HETDDFT_E_HIP);
      }
    c->cs = 0;
    use_chain_set(c);
  }
  // timec = cmplx(Lambdah*(1+tanh((z-tzmean)/tzsurf)), 1) with the real part
  // rounded to REAL(4) by the KIND-less cmplx (DFT4He3d.f90:773-782)
  std::vector<c64> tc(g.nz);
  for (int iz = 0; iz < g.nz; ++iz) {
    const double z = -p->zmax + c->hz * iz;
    const double re = p->lambdah * (1.0 + std::tanh((z - p->tzmean) / p->tzsurf));
    tc[iz].x = (double)(float)re;
    tc[iz].y = 1.0;
  }
  if (hipMemcpy(c->timec, tc.data(), g.nz * sizeof(c64), hipMemcpyHostToDevice) != hipSuccess) {
    fail(c, HETDDFT_E_HIP, "timec upload failed");
    return bad(HETDDFT_E_HIP);
  }
  *out = c;
  return HETDDFT_OK;
}

int hetddft_make_kernels(const hetddft_params* p, const hetddft_morse* mo, double* fvlj4, double* wcgk,
                         double* kalfs, double* vq) {
  if (!p || p->nx <= 0 || p->ny <= 0 || p->nz <= 0) return HETDDFT_E_ARG;
  using namespace kinit;
  const int nx = p->nx, ny = p->ny, nz = p->nz, mx = nx / 2 + 1;
  const double hx = 2.0 * std::fabs(p->xmax) / nx, hy = 2.0 * std::fabs(p->ymax) / ny,
               hz = 2.0 * std::fabs(p->zmax) / nz;
  const double pi = 4.0 * std::atan(1.0), twopi = 2.0 * pi;
  std::vector<double> px(mx), py(ny), pz(nz);
  for (int i = 0; i < mx; ++i) px[i] = pfreq(i, nx, hx);
  for (int i = 0; i < ny; ++i) py[i] = pfreq(i, ny, hy);
  for (int i = 0; i < nz; ++i) pz[i] = pfreq(i, nz, hz);
  const size_t H = (size_t)mx * ny * nz;
  std::vector<double> pm(H);
  for (int iz = 0; iz < nz; ++iz)
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < mx; ++ix)
        pm[ix + (size_t)mx * (iy + (size_t)ny * iz)] = std::sqrt((px[ix] * px[ix] + py[iy] * py[iy]) + pz[iz] * pz[iz]);
  const double h4 = p->h4;
  if (wcgk)
    for (size_t i = 0; i < H; ++i) {
      const double pp = pm[i];
      if (pp > 0.0) {
        const double arg = twopi * pp * h4;
        const double aux = std::sin(arg) - arg * std::cos(arg);
        wcgk[i] = std::fabs(aux) < 1e-10 ? 1.0 : 3.0 * aux / (arg * arg * arg);
      } else {
        wcgk[i] = 1.0;
      }
    }
  if (kalfs)
    for (size_t i = 0; i < H; ++i) {
      const double t = pi * p->l * pm[i];
      kalfs[i] = std::exp(-(t * t));
    }
  if (fvlj4) {
    const double eps = p->eps4, sigma = p->sigma4;
    const double sigma6 = std::pow(sigma, 6), aux1 = 8.0 * eps;
    const double pi4 = 4.0 * pi, eps4 = 4.0 * eps, s12 = sigma6 * sigma6;
    const double b0 = pi4 * eps4 * s12 / (9.0 * std::pow(h4, 9)) + (-pi4 * eps4 * sigma6 / (3.0 * std::pow(h4, 3)));
    for (size_t i = 0; i < H; ++i) {
      const double pp = pm[i];
      if (std::fabs(pp) < 1e-10) {
        fvlj4[i] = b0;
      } else {
        const double a2 = twopi * pp;
        fvlj4[i] = (aux1 * sigma6 * (sigma6 * sinx11(a2, h4) - sinx5(a2, h4))) / pp;
      }
    }
  }
  if (vq) {
    MorsePot pot{};
    pot.D = mo ? mo->D : 20.0;
    pot.re = mo ? mo->re : 5.0;
    pot.alpha = mo ? mo->alpha : 0.8;
    pot.rc = mo ? mo->rc : 2.0;
    pot.umax = 0.0;
    {
      const double e = std::exp(-pot.alpha * (pot.rc - pot.re));
      pot.umax = (mo && mo->umax != 0.0) ? mo->umax : pot.D * (e * e - 2.0 * e);
    }
    const int nq = (mo && mo->nq > 1) ? mo->nq : 1500;
    const double pmaxx = 1.0 / (2.0 * hx), pmaxy = 1.0 / (2.0 * hy), pmaxz = 1.0 / (2.0 * hz);
    const double qmax = std::sqrt(pmaxx * pmaxx + pmaxy * pmaxy + pmaxz * pmaxz);
    const double hq = qmax / (nq - 1);
    FtvSpline ft(pot, qmax, hq);
    for (size_t i = 0; i < H; ++i) vq[i] = (double)(float)ft(pm[i]);
  }
  return HETDDFT_OK;
}

int hetddft_init_kernels(hetddft_ctx* c, const hetddft_morse* mo) {
  if (!c) return HETDDFT_E_ARG;
  using namespace kinit;
  const hetddft_params* p = &c->p;
  const double hx = 2.0 * std::fabs(p->xmax) / p->nx, hy = 2.0 * std::fabs(p->ymax) / p->ny,
               hz = 2.0 * std::fabs(p->zmax) / p->nz;
  MorsePot pot{};
  pot.D = mo ? mo->D : 20.0;
  pot.re = mo ? mo->re : 5.0;
  pot.alpha = mo ? mo->alpha : 0.8;
  pot.rc = mo ? mo->rc : 2.0;
  {
    const double e = std::exp(-pot.alpha * (pot.rc - pot.re));
    pot.umax = (mo && mo->umax != 0.0) ? mo->umax : pot.D * (e * e - 2.0 * e);
  }
  const int nq = (mo && mo->nq > 1) ? mo->nq : 1500;
  const double pmaxx = 1.0 / (2.0 * hx), pmaxy = 1.0 / (2.0 * hy), pmaxz = 1.0 / (2.0 * hz);
  const double qmax = std::sqrt(pmaxx * pmaxx + pmaxy * pmaxy + pmaxz * pmaxz);
  const double hq = qmax / (nq - 1);
  const FtvSetup su(pot, qmax, hq);
  const int N = FtvSetup::N, np = (int)su.P.size();
  RET(settle(c));
  std::vector<void*> tmp;
  auto talloc = [&](size_t bytes) -> double* {
    void* q = nullptr;
    if (hipMalloc(&q, bytes) != hipSuccess) return nullptr;
    tmp.push_back(q);
    return (double*)q;
  };
  auto cleanup = [&]() {
    for (void* q : tmp) hipFree(q);
  };
  double *dx = talloc(N * 8), *da = talloc(N * 8), *db = talloc(N * 8), *dc = talloc(N * 8), *dd = talloc(N * 8);
  double *dP = talloc(np * 8), *dv = talloc(np * 8), *sA = talloc(np * 8), *sB = talloc(np * 8), *sC = talloc(np * 8),
         *sD = talloc(np * 8);
  for (double* q : {dx, da, db, dc, dd, dP, dv, sA, sB, sC, sD})
    if (!q) {
      cleanup();
      return fail(c, HETDDFT_E_NOMEM, "init_kernels: hipMalloc failed");
    }
  hipError_t e = hipSuccess;
  auto up = [&](double* dst, const std::vector<double>& v) {
    if (e == hipSuccess) e = hipMemcpy(dst, v.data(), v.size() * 8, hipMemcpyHostToDevice);
  };
  up(dx, su.x);
  up(da, su.a);
  up(db, su.b);
  up(dc, su.c);
  up(dd, su.d);
  up(dP, su.P);
  std::vector<double> vpp(np);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_ftv_sums, dim3(np), dim3(256), 0, c->st, (const double*)dx, (const double*)da,
                       (const double*)db, (const double*)dc, (const double*)dd, N, (const double*)dP, su.rc, su.fc, dv);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  if (e == hipSuccess) e = hipMemcpy(vpp.data(), dv, np * 8, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    cleanup();
    return fail(c, HETDDFT_E_HIP, "init_kernels: %s", hipGetErrorString(e));
  }
  const FtvSpline ft(su.P, vpp);
  up(dP, ft.P);
  up(sA, ft.A);
  up(sB, ft.B);
  up(sC, ft.C);
  up(sD, ft.D);
  KinitArgs ka{};
  ka.nx = p->nx;
  ka.ny = p->ny;
  ka.nz = p->nz;
  ka.mx = p->nx / 2 + 1;
  ka.hx = hx;
  ka.hy = hy;
  ka.hz = hz;
  ka.h4 = p->h4;
  ka.l = p->l;
  ka.eps = p->eps4;
  ka.sigma = p->sigma4;
  ka.P = dP;
  ka.A = sA;
  ka.B = sB;
  ka.C = sC;
  ka.D = sD;
  ka.np = np;
  const unsigned nb = blocks_for(c->g.h, 256);
  if (e == hipSuccess) {
    if (c->mixed)
      hipLaunchKernelGGL((k_make_kernels<float>), dim3(nb), dim3(256), 0, c->st, ka, (double*)c->fs64e.kfvlj4,
                         (double*)c->fs64e.kwcgk, (double*)nullptr, (double*)c->fs64e.kvq, (float*)c->fs.kfvlj4,
                         (float*)c->fs.kwcgk, (float*)c->fs.kkalfs, (float*)c->fs.kvq);
    else
      hipLaunchKernelGGL((k_make_kernels<double>), dim3(nb), dim3(256), 0, c->st, ka, (double*)c->fs.kfvlj4,
                         (double*)c->fs.kwcgk, (double*)c->fs.kkalfs, (double*)c->fs.kvq, (double*)nullptr,
                         (double*)nullptr, (double*)nullptr, (double*)nullptr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  cleanup();
  if (e != hipSuccess) return fail(c, HETDDFT_E_HIP, "init_kernels: %s", hipGetErrorString(e));
  c->kernels_set = true;
  return HETDDFT_OK;
}

int hetddft_get_kernels(hetddft_ctx* c, double* fvlj4, double* wcgk, double* kalfs, double* vq) {
  if (!c) return HETDDFT_E_ARG;
  if (!c->kernels_set) return fail(c, HETDDFT_E_STATE, "kernels not set");
  RET(settle(c));
  HIPCHK(c, hipStreamSynchronize(c->st));
  const Grid& g = c->g;
  std::vector<double> t((size_t)g.h);
  std::vector<float> tf(c->mixed ? (size_t)g.h : 0);
  const void* src[4] = {c->fs.kfvlj4, c->fs.kwcgk, c->fs.kkalfs, c->fs.kvq};
  double* dst[4] = {fvlj4, wcgk, kalfs, vq};
  for (int k = 0; k < 4; ++k) {
    if (!dst[k]) continue;
    if (c->mixed) {
      HIPCHK(c, hipMemcpy(tf.data(), src[k], (size_t)g.h * 4, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < (size_t)g.h; ++i) t[i] = (double)tf[i];
    } else {
      HIPCHK(c, hipMemcpy(t.data(), src[k], (size_t)g.h * 8, hipMemcpyDeviceToHost));
    }
    for (int iz = 0; iz < g.nz; ++iz)
      for (int iy = 0; iy < g.ny; ++iy)
        for (int ix = 0; ix < g.mx; ++ix)
          dst[k][((size_t)iz * g.ny + iy) * g.mx + ix] = t[((size_t)iz * g.mx + ix) * g.ny + iy];
  }
  return HETDDFT_OK;
}

int hetddft_set_kernels(hetddft_ctx* c, const double* fvlj4, const double* wcgk, const double* kalfs,
                        const double* vq) {
  if (!c || !fvlj4 || !wcgk || !kalfs || !vq) return fail(c, HETDDFT_E_ARG, "null kernel array");
  const size_t h = (size_t)c->g.h;
  (void)h;
  RET(upload_kernel(c, c->fs.kfvlj4, fvlj4, c->mixed));
  RET(upload_kernel(c, c->fs.kwcgk, wcgk, c->mixed));
  RET(upload_kernel(c, c->fs.kkalfs, kalfs, c->mixed));
  RET(upload_kernel(c, c->fs.kvq, vq, c->mixed));
  if (c->mixed) {
    RET(upload_kernel(c, c->fs64e.kfvlj4, fvlj4, false));
    RET(upload_kernel(c, c->fs64e.kwcgk, wcgk, false));
    RET(upload_kernel(c, c->fs64e.kvq, vq, false));
  }
  c->kernels_set = true;
  return HETDDFT_OK;
}

int hetddft_set_state(hetddft_ctx* c, const double* psi, const double* rimp, const double* vimp) {
  if (!c) return HETDDFT_E_ARG;
  c->means_ready = false;
  RET(settle(c));
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->chain_pending = false;
  if (psi) HIPCHK(c, hipMemcpy(c->psi, psi, (size_t)c->g.n * 16, hipMemcpyHostToDevice));
  const size_t np3 = (size_t)3 * c->p.n_par;
  if (rimp && np3) HIPCHK(c, hipMemcpy(c->rimp, rimp, np3 * 8, hipMemcpyHostToDevice));
  if (vimp && np3) HIPCHK(c, hipMemcpy(c->vimp, vimp, np3 * 8, hipMemcpyHostToDevice));
  c->started = false;
  return HETDDFT_OK;
}

int hetddft_start(hetddft_ctx* c) {
  if (!c) return HETDDFT_E_ARG;
  if (!c->kernels_set) return fail(c, HETDDFT_E_STATE, "hetddft_set_kernels not called");
  RET(settle(c));
  c->chain_pending = false;
  const Grid& g = c->g;
  if (c->mixed)
    hipLaunchKernelGGL((k_density<float>), dim3(blocks_for(g.n)), dim3(kBlock), 0, c->st, (const c64*)c->psi,
                       (float*)c->fld[hetddft_ctx::F_DEN], g.n);
  else
    hipLaunchKernelGGL((k_density<double>), dim3(blocks_for(g.n)), dim3(kBlock), 0, c->st, (const c64*)c->psi,
                       (double*)c->fld[hetddft_ctx::F_DEN], g.n);
  LAUNCHCHK(c);
  RET(run_veval(c));
  // every accumulator starts at zero, as after allocation in the reference
  // (dimen.f90:78-79,131-134; DFT4He3d.f90:841-845), also on a context that
  // has run before: pc and q of psi, pcr and pcv of the particles, the
  // particle history slots and the RK accumulators qr, qv
  const size_t np3 = (size_t)3 * std::max(c->p.n_par, 1);
  for (double* pp : {c->qr, c->qv, c->pcr, c->pcv, c->rold[0], c->rold[1], c->rold[2], c->vold[0], c->vold[1],
                     c->vold[2], c->aold[0], c->aold[1]})
    HIPCHK(c, hipMemsetAsync(pp, 0, np3 * 8, c->st));
  HIPCHK(c, hipMemsetAsync(c->pc, 0, (size_t)g.n * 16, c->st));
  HIPCHK(c, hipMemsetAsync(c->q, 0, (size_t)g.n * 16, c->st));
  HIPCHK(c, hipMemsetAsync(c->res, 0, 3 * sizeof(double), c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->err_pending = false;
  std::memset(c->last_err, 0, sizeof c->last_err);
  for (int k = 0; k < 3; ++k) c->ioldp[k] = c->ioldr[k] = c->ioldv[k] = k;
  c->ioldh[0] = c->iolda[0] = 0;
  c->ioldh[1] = c->iolda[1] = 1;
  clear_graphs(c);
  c->started = true;
  c->iter_in_run = 0;
  return HETDDFT_OK;
}

int hetddft_potential(hetddft_ctx* c) {
  if (!c) return HETDDFT_E_ARG;
  if (!c->kernels_set) return fail(c, HETDDFT_E_STATE, "hetddft_set_kernels not called");
  RET(run_veval(c));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return HETDDFT_OK;
}

int hetddft_step_rk(hetddft_ctx* c, double dt) {
  if (!c) return HETDDFT_E_ARG;
  if (!c->started) return fail(c, HETDDFT_E_STATE, "hetddft_start not called");
  // the particles move on st in the RK stages: a chain queued for the old
  // positions is not used
  RET(settle(c));
  c->chain_pending = false;
  RET(c->mixed ? step_rk_t<float>(c, dt) : step_rk_t<double>(c, dt));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return HETDDFT_OK;
}

int hetddft_step_pc(hetddft_ctx* c, double dt, double err[3]) {
  if (!c) return HETDDFT_E_ARG;
  if (!c->started) return fail(c, HETDDFT_E_STATE, "hetddft_start not called");
  RET(c->mixed ? step_pc_t<float>(c, dt) : step_pc_t<double>(c, dt));
  if (err) return fetch_errors(c, err);
  return HETDDFT_OK;
}

int hetddft_last_errors(hetddft_ctx* c, double err[3]) {
  if (!c || !err) return HETDDFT_E_ARG;
  return fetch_errors(c, err);
}

int hetddft_iterate(hetddft_ctx* c, double dt, double err[3]) {
  if (!c) return HETDDFT_E_ARG;
  if (!c->started) return fail(c, HETDDFT_E_STATE, "hetddft_start not called");
  if (c->iter_in_run < 3) {
    RET(settle(c));
    c->chain_pending = false;
    RET(c->mixed ? step_rk_t<float>(c, dt) : step_rk_t<double>(c, dt));
    RET(run_veval(c));
    if (err) {
      err[0] = err[1] = err[2] = 0.0;
      HIPCHK(c, hipStreamSynchronize(c->st));
    }
  } else {
    RET(iterate_pc(c, dt));
    if (err) RET(fetch_errors(c, err));
  }
  c->iter_in_run++;
  return HETDDFT_OK;
}

int hetddft_particle_means(hetddft_ctx* c, double out[9]) {
  if (!c || !out) return HETDDFT_E_ARG;
  c->means_wanted = true;
  if (c->means_ready) {
    std::memcpy(out, c->means_cache, sizeof c->means_cache);
    return HETDDFT_OK;
  }
  RET(settle(c));
  RET(flush_errors(c));
  const int np = c->p.n_par;
  if (np == 0) {
    std::memset(out, 0, 9 * sizeof(double));
    return HETDDFT_OK;
  }
  RET(launch_means(c));
  HIPCHK(c, hipMemcpyAsync(c->hres + 16, c->res + 16, 9 * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  means_from_host(c, c->means_cache);
  c->means_ready = true;
  std::memcpy(out, c->means_cache, sizeof c->means_cache);
  return HETDDFT_OK;
}

int hetddft_particle_means_dev(const double* rimp, const double* vimp, const double* aimp, int n_par,
                               double* out, void* stream) {
  auto bad = [](int code, const char* msg) {
    g_create_error = msg;
    return code;
  };
  if (n_par < 0 || n_par > HETDDFT_MEANS_DEV_MAX_PAR)
    return bad(HETDDFT_E_ARG, "hetddft_particle_means_dev: n_par outside 0..2^30");
  if (!out || (n_par > 0 && (!rimp || !vimp || !aimp)))
    return bad(HETDDFT_E_ARG, "hetddft_particle_means_dev: null array");
  double* scr = nullptr;
  const int rc = scratch_of_current_device(&scr);
  if (rc != HETDDFT_OK) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (n_par == 0) {
    hipLaunchKernelGGL(k_fill_zero, dim3(1), dim3(64), 0, st, out, 9);
  } else {
    // the kernels and block counts of hetddft_particle_means, so the sums
    // are the same bit for bit
    const unsigned nb = std::min(blocks_for(n_par), (unsigned)kNoAllocParts);
    hipLaunchKernelGGL(k_part_sums, dim3(nb, 9), dim3(kBlock), 0, st, n_par, rimp, vimp, aimp, scr);
    hipLaunchKernelGGL(k_finish_sums, dim3(9), dim3(kBlock), 0, st, (const double*)scr, (int)nb, 9,
                       scr + 9 * kNoAllocParts);
    hipLaunchKernelGGL(k_part_means_scale, dim3(1), dim3(64), 0, st, (const double*)(scr + 9 * kNoAllocParts),
                       n_par, 1.0 / 7.63822291, out);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    g_create_error = std::string("hetddft_particle_means_dev: ") + hipGetErrorString(e);
    return HETDDFT_E_HIP;
  }
  return HETDDFT_OK;
}

int hetddft_diagnostics(hetddft_ctx* c, double out[HETDDFT_DIAG_N]) {
  if (!c || !out) return HETDDFT_E_ARG;
  if (!c->started) return fail(c, HETDDFT_E_STATE, "hetddft_start not called");
  RET(settle(c));
  RET(flush_errors(c));
  const Grid& g = c->g;
  const DiagArgs da = diag_args(c);
  const unsigned nb = blocks_for(g.n);
  if (c->mixed)
    hipLaunchKernelGGL((k_diag<float>), dim3(nb), dim3(kBlock), 0, c->st, da, (const c64*)c->psi,
                       (const float*)c->fld[hetddft_ctx::F_POT4], (const float*)c->denx, c->part);
  else
    hipLaunchKernelGGL((k_diag<double>), dim3(nb), dim3(kBlock), 0, c->st, da, (const c64*)c->psi,
                       (const double*)c->fld[hetddft_ctx::F_POT4], (const double*)c->denx, c->part);
  LAUNCHCHK(c);
  hipLaunchKernelGGL(k_finish_sums, dim3(kDiagSlots), dim3(kBlock), 0, c->st, (const double*)c->part, (int)nb,
                     kDiagSlots, c->res + 32);
  LAUNCHCHK(c);
  const int np = c->p.n_par;
  if (np > 0) {
    hipLaunchKernelGGL(k_part_minmax, dim3(6), dim3(kBlock), 0, c->st, np, (const double*)c->vimp,
                       (const double*)c->aimp, c->res + 32 + kDiagSlots);
    LAUNCHCHK(c);
  }
  HIPCHK(c, hipMemcpyAsync(c->hres + 32, c->res + 32, (kDiagSlots + 12) * sizeof(double), hipMemcpyDeviceToHost,
                           c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  const double* r = c->hres + 32;
  const double dv = c->dxyz;
  out[HETDDFT_DIAG_NORM] = r[0] * dv;
  for (int k = 0; k < 3; ++k) {
    out[HETDDFT_DIAG_CM + k] = r[1 + k] * dv;
    out[HETDDFT_DIAG_JTOT + k] = 2.0 * c->p.h2o2m4 * r[8 + k] * dv;
    out[HETDDFT_DIAG_MSX + k] = std::sqrt(r[4 + k] * dv);
  }
  out[HETDDFT_DIAG_MU] = r[7] / r[0];
  for (int k = 0; k < 12; ++k) out[HETDDFT_DIAG_PMINMAX + k] = np > 0 ? r[kDiagSlots + k] : 0.0;
  return HETDDFT_OK;
}

int hetddft_checkpoint_size(hetddft_ctx* c, long long* bytes) {
  if (!c || !bytes) return HETDDFT_E_ARG;
  *bytes = ckpt_bytes(c);
  return HETDDFT_OK;
}

// Header to hdr (host memory), arrays to data: host memory (kind D2H) or
// device memory (kind D2D, the autotuner's snapshot).
static int checkpoint_save_to(hetddft_ctx* c, void* hdr, void* data, hipMemcpyKind kind) {
  if (!c->started) return fail(c, HETDDFT_E_STATE, "hetddft_start not called");
  RET(settle(c));
  HIPCHK(c, hipStreamSynchronize(c->st));
  CkptHeader h;
  std::memset(&h, 0, sizeof h);
  std::memcpy(h.magic, "HETDCKP1", 8);
  h.nx = c->p.nx;
  h.ny = c->p.ny;
  h.nz = c->p.nz;
  h.n_par = c->p.n_par;
  h.precision = c->p.precision;
  h.iter_in_run = c->iter_in_run;
  std::copy(c->ioldp, c->ioldp + 3, h.ioldp);
  std::copy(c->ioldh, c->ioldh + 2, h.ioldh);
  std::copy(c->ioldr, c->ioldr + 3, h.ioldr);
  std::copy(c->ioldv, c->ioldv + 3, h.ioldv);
  std::copy(c->iolda, c->iolda + 2, h.iolda);
  std::memcpy(hdr, &h, sizeof h);
  unsigned char* o = static_cast<unsigned char*>(data);
  for (const CkptPart& q : ckpt_parts(c)) {
    HIPCHK(c, hipMemcpyAsync(o, q.dev, q.bytes, kind, c->st));
    o += q.bytes;
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  return HETDDFT_OK;
}

int hetddft_checkpoint_save(hetddft_ctx* c, void* buf) {
  if (!c || !buf) return HETDDFT_E_ARG;
  return checkpoint_save_to(c, buf, static_cast<unsigned char*>(buf) + sizeof(CkptHeader), hipMemcpyDeviceToHost);
}

// Header from hdr (host memory), arrays from data (host: kind H2D; device:
// kind D2D).
static int checkpoint_load_from(hetddft_ctx* c, const void* hdr, const void* data, hipMemcpyKind kind) {
  c->means_ready = false;
  if (!c->kernels_set) return fail(c, HETDDFT_E_STATE, "hetddft_set_kernels not called");
  CkptHeader h;
  std::memcpy(&h, hdr, sizeof h);
  if (std::memcmp(h.magic, "HETDCKP1", 8) != 0) return fail(c, HETDDFT_E_ARG, "checkpoint: bad magic");
  if (h.nx != c->p.nx || h.ny != c->p.ny || h.nz != c->p.nz || h.n_par != c->p.n_par ||
      h.precision != c->p.precision)
    return fail(c, HETDDFT_E_ARG, "checkpoint of a %dx%dx%d grid, N_par %d, precision %d does not fit this context",
                h.nx, h.ny, h.nz, h.n_par, h.precision);
  auto perm_ok = [](const int* a, int n) {
    int seen = 0;
    for (int k = 0; k < n; ++k) {
      if (a[k] < 0 || a[k] >= n || (seen >> a[k]) & 1) return false;
      seen |= 1 << a[k];
    }
    return true;
  };
  if (!perm_ok(h.ioldp, 3) || !perm_ok(h.ioldh, 2) || !perm_ok(h.ioldr, 3) || !perm_ok(h.ioldv, 3) ||
      !perm_ok(h.iolda, 2) || h.iter_in_run < 0)
    return fail(c, HETDDFT_E_ARG, "checkpoint: corrupt slot order");
  RET(settle(c));
  HIPCHK(c, hipStreamSynchronize(c->st));
  const unsigned char* o = static_cast<const unsigned char*>(data);
  for (const CkptPart& q : ckpt_parts(c)) {
    HIPCHK(c, hipMemcpyAsync(q.dev, o, q.bytes, kind, c->st));
    o += q.bytes;
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  std::copy(h.ioldp, h.ioldp + 3, c->ioldp);
  std::copy(h.ioldh, h.ioldh + 2, c->ioldh);
  std::copy(h.ioldr, h.ioldr + 3, c->ioldr);
  std::copy(h.ioldv, h.ioldv + 3, c->ioldv);
  std::copy(h.iolda, h.iolda + 2, c->iolda);
  c->iter_in_run = h.iter_in_run;
  c->chain_pending = false;
  // the error estimates of steps run before the load belong to another state
  c->err_pending = false;
  std::memset(c->last_err, 0, sizeof c->last_err);
  HIPCHK(c, hipMemsetAsync(c->res, 0, 3 * sizeof(double), c->st));
  clear_graphs(c);
  // den and the potential fields of the restored state (the fields the saved
  // context held after its last potential evaluation)
  if (c->mixed)
    hipLaunchKernelGGL((k_density<float>), dim3(blocks_for(c->g.n)), dim3(kBlock), 0, c->st, (const c64*)c->psi,
                       (float*)c->fld[hetddft_ctx::F_DEN], c->g.n);
  else
    hipLaunchKernelGGL((k_density<double>), dim3(blocks_for(c->g.n)), dim3(kBlock), 0, c->st, (const c64*)c->psi,
                       (double*)c->fld[hetddft_ctx::F_DEN], c->g.n);
  LAUNCHCHK(c);
  RET(run_veval(c));
  RET(settle(c));
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->started = true;
  return HETDDFT_OK;
}

int hetddft_checkpoint_load(hetddft_ctx* c, const void* buf) {
  if (!c || !buf) return HETDDFT_E_ARG;
  return checkpoint_load_from(c, buf, static_cast<const unsigned char*>(buf) + sizeof(CkptHeader), hipMemcpyHostToDevice);
}

int hetddft_sample_particles(hetddft_ctx* c, const double* denx, const double lo[3], const double hi[3],
                             unsigned long long seed, long long max_attempts, double* rimp) {
  if (!c || !denx || !lo || !hi || !rimp || max_attempts <= 0) return HETDDFT_E_ARG;
  const int np = c->p.n_par;
  if (np == 0) return HETDDFT_OK;
  const Grid& g = c->g;
  SortArgs a{};
  a.g = g;
  // x(1) and x(2) - x(1) from the grid arrays x(ix) = -xmax + hx (ix-1)
  a.x0 = -c->p.xmax;
  a.y0 = -c->p.ymax;
  a.z0 = -c->p.zmax;
  a.dx = (-c->p.xmax + c->hx) - a.x0;
  a.dy = (-c->p.ymax + c->hy) - a.y0;
  a.dz = (-c->p.zmax + c->hz) - a.z0;
  double mx = denx[0];
  for (long long i = 1; i < g.n; ++i) mx = std::max(mx, denx[i]);
  if (!(mx > 0.0)) return fail(c, HETDDFT_E_ARG, "sample_particles: denx has no positive value");
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = lo[k];
    a.hi[k] = hi[k];
  }
  a.denmax = mx;
  a.seed = seed;
  a.max_attempts = max_attempts;
  RET(settle(c));
  double* dd = nullptr;
  double* dr = nullptr;
  int* flag = nullptr;
  auto cleanup = [&]() {
    if (dd) hipFree(dd);
    if (dr) hipFree(dr);
    if (flag) hipFree(flag);
  };
  if (hipMalloc((void**)&dd, (size_t)g.n * 8) != hipSuccess || hipMalloc((void**)&dr, (size_t)3 * np * 8) != hipSuccess ||
      hipMalloc((void**)&flag, sizeof(int)) != hipSuccess) {
    cleanup();
    return fail(c, HETDDFT_E_NOMEM, "sample_particles: hipMalloc failed");
  }
  hipError_t e = hipMemcpy(dd, denx, (size_t)g.n * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(flag, 0, sizeof(int));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_sorteo, dim3(blocks_for(np)), dim3(kBlock), 0, c->st, a, np, (const double*)dd, dr, flag);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  int failed = 0;
  if (e == hipSuccess) e = hipMemcpy(&failed, flag, sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess && !failed) e = hipMemcpy(rimp, dr, (size_t)3 * np * 8, hipMemcpyDeviceToHost);
  cleanup();
  if (e != hipSuccess) return fail(c, HETDDFT_E_HIP, "sample_particles: %s", hipGetErrorString(e));
  if (failed) return fail(c, HETDDFT_E_STATE, "sample_particles: a particle was not accepted in %lld attempts",
                          max_attempts);
  return HETDDFT_OK;
}

int hetddft_init_droplet(hetddft_ctx* c, const hetddft_droplet* d) {
  if (!c || !d) return HETDDFT_E_ARG;
  c->means_ready = false;
  if (!(d->n_he > 0.0) || !(d->rho0 > 0.0) || !(d->r0 > 0.0) || !(d->a > 0.0) || !(d->ripple >= 0.0) ||
      !(d->phase_noise >= 0.0) || !(d->cavity >= 0.0) || (d->cavity > 0.0 && !(d->cavity_w > 0.0)))
    return fail(c, HETDDFT_E_ARG, "init_droplet: invalid droplet parameters");
  const Grid& g = c->g;
  DropArgs a{};
  a.g = g;
  // x(ix) = -xmax + hx (ix - 1), as the grid arrays of the reference
  a.x0 = -c->p.xmax;
  a.y0 = -c->p.ymax;
  a.z0 = -c->p.zmax;
  a.hx = c->hx;
  a.hy = c->hy;
  a.hz = c->hz;
  a.cx = d->centre[0];
  a.cy = d->centre[1];
  a.cz = d->centre[2];
  a.rho0 = d->rho0;
  a.R = d->r0 * std::cbrt(d->n_he);
  a.a = d->a;
  a.rb = d->cavity;
  a.w = d->cavity_w;
  a.ripple = d->ripple;
  for (int k = 0; k < 24; ++k) a.clm[k] = d->clm[k];
  a.kx = d->lamda[0];
  a.ky = d->lamda[1];
  a.kz = d->lamda[2];
  a.sphi = d->phase_noise;
  a.n_he = d->n_he;
  a.dxyz = c->dxyz;
  a.seed = d->seed;
  RET(settle(c));
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->chain_pending = false;
  double* tmp = nullptr;  // kDropBlocks partials, then max and sum
  if (hipMalloc((void**)&tmp, (kDropBlocks + 2) * sizeof(double)) != hipSuccess)
    return fail(c, HETDDFT_E_NOMEM, "init_droplet: hipMalloc failed");
  double* smax = tmp + kDropBlocks;
  double* sum = smax + 1;
  if (a.ripple > 0.0) {
    hipLaunchKernelGGL(k_drop_surfmax, dim3(kDropBlocks), dim3(kBlock), 0, c->st, a, tmp);
    hipLaunchKernelGGL(k_drop_maxfin, dim3(1), dim3(kBlock), 0, c->st, (const double*)tmp, kDropBlocks, smax);
  }
  hipLaunchKernelGGL(k_drop_sum, dim3(kDropBlocks), dim3(kBlock), 0, c->st, a, (const double*)smax, tmp);
  hipLaunchKernelGGL(k_finish_sums, dim3(1), dim3(kBlock), 0, c->st, (const double*)tmp, kDropBlocks, 1, sum);
  hipLaunchKernelGGL(k_drop_psi, dim3(blocks_for(g.n)), dim3(kBlock), 0, c->st, a, (const double*)smax,
                     (const double*)sum, (c64*)c->psi);
  const int np = c->p.n_par;
  if (np > 0) {
    const double ktops = 7.63822291;  // A/ps -> A*K (DFT4He3d.f90:600-602)
    const double sg = d->sigma > 0.0 ? d->sigma : 0.0;
    if (d->sigma > 0.0)
      hipLaunchKernelGGL(k_drop_cloud, dim3(blocks_for(np)), dim3(kBlock), 0, c->st, np, a.cx, a.cy, a.cz, sg,
                         d->v0[0] * ktops, d->v0[1] * ktops, d->v0[2] * ktops, a.seed, c->rimp, c->vimp);
    else
      hipLaunchKernelGGL(k_drop_velocity, dim3(blocks_for(np)), dim3(kBlock), 0, c->st, np, d->v0[0] * ktops,
                         d->v0[1] * ktops, d->v0[2] * ktops, c->vimp);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  hipFree(tmp);
  if (e != hipSuccess) return fail(c, HETDDFT_E_HIP, "init_droplet: %s", hipGetErrorString(e));
  c->started = false;
  return HETDDFT_OK;
}

int hetddft_current(hetddft_ctx* c, double* jx, double* jy, double* jz) {
  if (!c) return HETDDFT_E_ARG;
  if (!c->started) return fail(c, HETDDFT_E_STATE, "hetddft_start not called");
  RET(settle(c));
  const Grid& g = c->g;
  if (!c->jcur) RET(dalloc(c, (void**)&c->jcur, (size_t)3 * g.n * sizeof(double)));
  double* d[3] = {c->jcur, c->jcur + g.n, c->jcur + 2 * g.n};
  hipLaunchKernelGGL(k_current, dim3(blocks_for(g.n)), dim3(kBlock), 0, c->st, diag_args(c), (const c64*)c->psi, d[0],
                     d[1], d[2]);
  LAUNCHCHK(c);
  HIPCHK(c, hipStreamSynchronize(c->st));
  double* h[3] = {jx, jy, jz};
  for (int k = 0; k < 3; ++k)
    if (h[k]) HIPCHK(c, hipMemcpy(h[k], d[k], (size_t)g.n * sizeof(double), hipMemcpyDeviceToHost));
  return HETDDFT_OK;
}

int hetddft_energy(hetddft_ctx* c, double out[12]) {
  if (!c || !out) return HETDDFT_E_ARG;
  RET(settle(c));
  if (!c->started) return fail(c, HETDDFT_E_STATE, "hetddft_start not called");
  RET(flush_errors(c));
  return c->mixed ? energy_t<float>(c, out) : energy_t<double>(c, out);
}

int hetddft_get_state(hetddft_ctx* c, double* psi, double* rimp, double* vimp, double* aimp) {
  if (!c) return HETDDFT_E_ARG;
  RET(settle(c));
  HIPCHK(c, hipStreamSynchronize(c->st));
  const size_t np3 = (size_t)3 * c->p.n_par;
  if (psi) HIPCHK(c, hipMemcpy(psi, c->psi, (size_t)c->g.n * 16, hipMemcpyDeviceToHost));
  if (rimp && np3) HIPCHK(c, hipMemcpy(rimp, c->rimp, np3 * 8, hipMemcpyDeviceToHost));
  if (vimp && np3) HIPCHK(c, hipMemcpy(vimp, c->vimp, np3 * 8, hipMemcpyDeviceToHost));
  if (aimp && np3) HIPCHK(c, hipMemcpy(aimp, c->aimp, np3 * 8, hipMemcpyDeviceToHost));
  return HETDDFT_OK;
}

int hetddft_get_field(hetddft_ctx* c, int which, double* out) {
  if (!c || !out) return HETDDFT_E_ARG;
  RET(settle(c));
  static const int map[20] = {hetddft_ctx::F_DEN,    hetddft_ctx::F_POT4,  hetddft_ctx::F_UIMP,
                              hetddft_ctx::F_DENCG,  hetddft_ctx::F_DELJ4, hetddft_ctx::F_FALFS,
                              hetddft_ctx::F_DX,     hetddft_ctx::F_DY,    hetddft_ctx::F_DZ,
                              hetddft_ctx::F_IX,     hetddft_ctx::F_IY,    hetddft_ctx::F_IZ,
                              -1,                    hetddft_ctx::F_PSIX,  hetddft_ctx::F_HPSIX,
                              hetddft_ctx::F_FCX,    hetddft_ctx::F_FCY,   hetddft_ctx::F_FCZ,
                              hetddft_ctx::F_STO1,   hetddft_ctx::F_UACONV};
  if (which < 0 || which >= 20) return fail(c, HETDDFT_E_ARG, "unknown field id %d", which);
  const void* src = which == HETDDFT_F_DENX ? c->denx : c->fld[map[which]];
  return c->mixed ? field_to_host<float>(c, src, out) : field_to_host<double>(c, src, out);
}

int hetddft_convolve(hetddft_ctx* c, int nf, const double* const* in, const double* kern, double* const* out) {
  if (!c || nf < 1 || nf > 5 || !in || !kern || !out) return fail(c, HETDDFT_E_ARG, "convolve: bad arguments");
  RET(settle(c));
  const size_t n = (size_t)c->g.n;
  // stage inputs in the force-convolution fields, kernel in the alpha_s slot
  const int stage[5] = {hetddft_ctx::F_FCX, hetddft_ctx::F_FCY, hetddft_ctx::F_FCZ, hetddft_ctx::F_STO1,
                        hetddft_ctx::F_UACONV};
  const int dst[5] = {hetddft_ctx::F_IX, hetddft_ctx::F_IY, hetddft_ctx::F_IZ, hetddft_ctx::F_DENCG,
                      hetddft_ctx::F_DELJ4};
  for (int i = 0; i < nf; ++i) RET(upload_real(c, c->fld[stage[i]], in[i], n, c->mixed));
  RET(upload_kernel(c, c->fs.kkalfs, kern, c->mixed));
  c->kernels_set = false;
  auto run = [&](auto tag) -> int {
    using T = decltype(tag);
    XfArgs<T> xa{};
    std::vector<ConvJob<T>> jobs;
    for (int i = 0; i < nf; ++i) {
      xa.op[i] = XF_COPY;
      xa.in[i][0] = (const T*)c->fld[stage[i]];
      jobs.push_back({i, c->fs.kkalfs, 1.0, c->fld[dst[i]]});
    }
    RET(conv_batch<T>(c, c->fs, xa, nf, jobs));
    for (int i = 0; i < nf; ++i) RET(field_to_host<T>(c, c->fld[dst[i]], out[i]));
    return HETDDFT_OK;
  };
  return c->mixed ? run(float{}) : run(double{});
}

int hetddft_profile(hetddft_ctx* c, int enable) {
  if (!c) return HETDDFT_E_ARG;
  RET(settle(c));
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (auto& r : c->prof_pending) {
    if (r.a) c->prof_pool.push_back(r.a);
    c->prof_pool.push_back(r.b);
  }
  c->prof_pending.clear();
  for (int k = 0; k < HETDDFT_PROF_N; ++k) {
    c->prof_ms[k] = 0.0;
    c->prof_n[k] = 0;
    c->prof_bytes[k] = 0.0;
  }
  HIPCHK(c, hipMemset(c->prof_dev, 0, HETDDFT_PROF_N * sizeof(unsigned long long)));
  c->prof_on = enable != 0;
  return HETDDFT_OK;
}

int hetddft_profile_read(hetddft_ctx* c, int cls, double* ms_total, long long* launches, double* bytes_total) {
  if (!c || cls < 0 || cls >= HETDDFT_PROF_N) return HETDDFT_E_ARG;
  RET(settle(c));
  HIPCHK(c, hipStreamSynchronize(c->st));
  hipEvent_t prev = nullptr;  // end event of the previous launch on st
  for (auto& r : c->prof_pending) {
    if (r.cls >= 0 && prev) {
      float ms = 0.0f;
      HIPCHK(c, hipEventElapsedTime(&ms, prev, r.b));
      c->prof_ms[r.cls] += ms;
      c->prof_n[r.cls] += 1;
      c->prof_bytes[r.cls] += r.bytes;
    }
    prev = r.b;
  }
  for (auto& r : c->prof_pending) {
    if (r.a) c->prof_pool.push_back(r.a);
    c->prof_pool.push_back(r.b);
  }
  c->prof_pending.clear();
  {
    unsigned long long dev[HETDDFT_PROF_N];
    HIPCHK(c, hipMemcpy(dev, c->prof_dev, sizeof dev, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemset(c->prof_dev, 0, sizeof dev));
    for (int k = 0; k < HETDDFT_PROF_N; ++k) c->prof_bytes[k] += (double)dev[k];
  }
  if (ms_total) *ms_total = c->prof_ms[cls];
  if (launches) *launches = c->prof_n[cls];
  if (bytes_total) *bytes_total = c->prof_bytes[cls];
  return HETDDFT_OK;
}

int hetddft_graph_info(hetddft_ctx* c, int* count, int* enabled) {
  if (!c) return HETDDFT_E_ARG;
  if (count) *count = (int)c->graphs.size();
  if (enabled) *enabled = c->graph_on ? 1 : 0;
  return HETDDFT_OK;
}

int hetddft_sync(hetddft_ctx* c) {
  if (!c) return HETDDFT_E_ARG;
  // a PC step whose error estimates nobody fetched: reduce and check them
  // here, so a run driven with err = NULL still reports a diverged state
  if (c->err_pending) return fetch_errors(c, nullptr);
  RET(settle(c));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return HETDDFT_OK;
}

int hetddft_get_tuning(hetddft_ctx* c, hetddft_tuning* out) {
  if (!c || !out) return HETDDFT_E_ARG;
  tuning_of(c, out);
  out->ms_per_step = 0.0;
  return HETDDFT_OK;
}

int hetddft_set_tuning(hetddft_ctx* c, const hetddft_tuning* t) {
  if (!c || !t) return HETDDFT_E_ARG;
  RET(settle(c));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return apply_tuning(c, t);
}

int hetddft_autotune(hetddft_ctx* c, double dt, int steps, hetddft_tuning* out) {
  if (!c) return HETDDFT_E_ARG;
  if (!c->started) return fail(c, HETDDFT_E_STATE, "hetddft_start not called");
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(c, HETDDFT_E_ARG, "autotune: deltat %g", dt);
  if (steps <= 0) steps = 4;
  const std::string key = tune_key(c);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      RET(hetddft_set_tuning(c, &it->second));
      if (out) *out = it->second;
      return HETDDFT_OK;
    }
  }
  // Snapshot of the state, restored at the end.  In device memory when it
  // fits, so that the restore neither waits on a host copy nor leaves the
  // compute units idle for it; in host memory otherwise.
  long long sz = 0;
  RET(hetddft_checkpoint_size(c, &sz));
  CkptHeader snap_hdr;
  std::vector<unsigned char> img;
  void* dimg = nullptr;
  const size_t data_bytes = (size_t)sz - sizeof(CkptHeader);
  if (hipMalloc(&dimg, data_bytes) != hipSuccess) {
    (void)hipGetLastError();
    dimg = nullptr;
    img.resize(data_bytes);
  }
  {
    const int rs = checkpoint_save_to(c, &snap_hdr, dimg ? dimg : img.data(),
                                      dimg ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost);
    if (rs != HETDDFT_OK) {
      if (dimg) (void)hipFree(dimg);
      return rs;
    }
  }
  const bool graphs = c->graph_on;
  c->graph_on = false;
  hetddft_tuning best;
  tuning_of(c, &best);
  int rc = HETDDFT_OK;
  auto run = [&]() -> int {
    // the PC classes need the RK start-up behind them
    while (c->iter_in_run < 3) RET(hetddft_iterate(c, dt, nullptr));
    // stencil classes: every tile shape x z chunk, fastest per class
    const int cls_prof[4] = {HETDDFT_PROF_GRAD, HETDDFT_PROF_POT, HETDDFT_PROF_PCP, HETDDFT_PROF_PCC};
    const int cls_tune[4] = {HETDDFT_TUNE_GRAD, HETDDFT_TUNE_POT, HETDDFT_TUNE_PCP, HETDDFT_TUNE_PCC};
    struct Cand {
      double ms;
      int tile, zc, pf;
    };
    std::vector<Cand> cand[4];
    auto set_cls = [&](hetddft_tuning& t, int j, const Cand& k) {
      t.tile[cls_tune[j]] = k.tile;
      t.zc[cls_tune[j]] = k.zc;
      t.prefetch[cls_tune[j]] = k.pf;
      if (cls_tune[j] == HETDDFT_TUNE_PCP) {
        t.tile[HETDDFT_TUNE_RK] = k.tile;
        t.zc[HETDDFT_TUNE_RK] = k.zc;
        t.prefetch[HETDDFT_TUNE_RK] = k.pf;
      }
    };
    {
      for (int tile = 0; tile < 4; ++tile)
        for (int zc : zc_candidates(c, tile))
          for (int pf = 1; pf >= 0; --pf) {
          hetddft_tuning t = best;
          for (int k = 0; k < HETDDFT_TUNE_NSTENCIL; ++k) {
            t.tile[k] = tile;
            t.zc[k] = zc;
            t.prefetch[k] = pf;
          }
          RET(apply_tuning(c, &t));
          double ms[HETDDFT_PROF_N];
          RET(time_classes(c, dt, steps, ms));
          for (int j = 0; j < 4; ++j)
            if (ms[cls_prof[j]] > 0.0) cand[j].push_back({ms[cls_prof[j]], tile, zc, pf});
          }
      for (int j = 0; j < 4; ++j) {
        std::sort(cand[j].begin(), cand[j].end(), [](const Cand& a, const Cand& b) { return a.ms < b.ms; });
        if (!cand[j].empty()) set_cls(best, j, cand[j][0]);
      }
    }
    RET(apply_tuning(c, &best));
    // A class's launch time, taken in isolation, misses what a shape does to
    // the launches around it (the cache contents it leaves, the side-stream
    // kernels it shares the GPU with): among the three fastest shapes of each
    // class, keep the one with the shortest whole PC step.
    {
      double cur = 0.0;
      RET(time_steps(c, dt, 2 * steps, &cur));
      for (int j = 0; j < 4; ++j)
        for (size_t r = 1; r < cand[j].size() && r < 3; ++r) {
          hetddft_tuning t = best;
          set_cls(t, j, cand[j][r]);
          RET(apply_tuning(c, &t));
          double w = 0.0;
          RET(time_steps(c, dt, 2 * steps, &w));
          if (w < cur) {
            cur = w;
            best = t;
          }
        }
      RET(apply_tuning(c, &best));
    }
    // z transform: tile width x persistent workgroups per CU; the three
    // fastest by launch time, then by whole PC step as above
    if (best.fft_z_tile != 0) {
      struct ZCand {
        double ms;
        int w, o;
      };
      std::vector<ZCand> zc;
      for (int w : {8, 16, 32})
        for (int o = 1; o <= 4; ++o) {
          hetddft_tuning t = best;
          t.fft_z_tile = w;
          t.fft_z_wgs = o;
          if (apply_tuning(c, &t) != HETDDFT_OK) continue;
          double ms[HETDDFT_PROF_N];
          RET(time_classes(c, dt, steps, ms));
          if (ms[HETDDFT_PROF_Z] > 0.0) zc.push_back({ms[HETDDFT_PROF_Z], w, o});
        }
      std::sort(zc.begin(), zc.end(), [](const ZCand& a, const ZCand& b) { return a.ms < b.ms; });
      double cur = 1e300;
      hetddft_tuning keep = best;
      for (size_t r = 0; r < zc.size() && r < 3; ++r) {
        hetddft_tuning t = best;
        t.fft_z_tile = zc[r].w;
        t.fft_z_wgs = zc[r].o;
        RET(apply_tuning(c, &t));
        double w = 0.0;
        RET(time_steps(c, dt, 2 * steps, &w));
        if (w < cur) {
          cur = w;
          keep = t;
        }
      }
      best = keep;
      RET(apply_tuning(c, &best));
    }
    // wall time of the chosen shapes
    return time_steps(c, dt, steps, &best.ms_per_step);
  };
  rc = run();
  c->prof_on = false;
  c->graph_on = graphs;
  const int rc2 = checkpoint_load_from(c, &snap_hdr, dimg ? dimg : img.data(),
                                       dimg ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice);
  if (dimg) (void)hipFree(dimg);
  if (rc != HETDDFT_OK) return rc;
  if (rc2 != HETDDFT_OK) return rc2;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = best;
  }
  if (out) *out = best;
  return HETDDFT_OK;
}

int hetddft_batch_create(hetddft_ctx* const* ctxs, int n, int nthreads, hetddft_batch** out) {
  if (!out) return HETDDFT_E_ARG;
  *out = nullptr;
  if (!ctxs || n <= 0) return HETDDFT_E_ARG;
  for (int i = 0; i < n; ++i)
    if (!ctxs[i]) return HETDDFT_E_ARG;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j)
      if (ctxs[i] == ctxs[j]) return HETDDFT_E_ARG;
  auto* b = new (std::nothrow) hetddft_batch();
  if (!b) return HETDDFT_E_NOMEM;
  b->ctx.assign(ctxs, ctxs + n);
  // groups of equal keys, at most kMaxMembers each, in instance order
  std::vector<std::string> keys(n);
  for (int i = 0; i < n; ++i) keys[i] = group_key(ctxs[i]);
  std::vector<bool> used(n, false);
  for (int i = 0; i < n; ++i) {
    if (used[i]) continue;
    std::vector<int> grp{i};
    used[i] = true;
    for (int j = i + 1; j < n && (int)grp.size() < kMaxMembers; ++j)
      if (!used[j] && keys[j] == keys[i]) {
        grp.push_back(j);
        used[j] = true;
      }
    b->groups.push_back(grp);
  }
  const int ng = (int)b->groups.size();
  const int nt = std::max(1, std::min(ng, nthreads > 0 ? nthreads : 8));
  // largest groups first, each to the thread with the fewest grid points so far
  std::vector<int> order(ng);
  for (int i = 0; i < ng; ++i) order[i] = i;
  auto points = [&](int gi) { return (long long)b->groups[gi].size() * ctxs[b->groups[gi][0]]->g.n; };
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return points(a) > points(c); });
  b->part.assign(nt, {});
  std::vector<long long> load(nt, 0);
  for (int gi : order) {
    const int t = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    b->part[t].push_back(gi);
    load[t] += points(gi);
  }
  for (auto& v : b->part) std::sort(v.begin(), v.end());
  *out = b;
  return HETDDFT_OK;
}

int hetddft_batch_iterate(hetddft_batch* b, const double* dt, int steps, double* err, int* failed) {
  if (failed) *failed = -1;
  if (!b || !dt || steps < 0) return HETDDFT_E_ARG;
  const int n = (int)b->ctx.size();
  std::atomic<bool> stop{false};
  std::mutex mu;
  int first_rc = HETDDFT_OK, first_i = -1;
  auto record = [&](int i, int rc) {
    std::lock_guard<std::mutex> lk(mu);
    if (first_i < 0) {
      first_i = i;
      first_rc = rc;
    }
    stop = true;
  };
  // One group of the batch: shared launches on the lead's streams when the
  // members are in the same state with the same step, else one instance at
  // a time.
  auto drive_group = [&](const std::vector<int>& idx) -> bool {
    bool shared = idx.size() > 1;
    for (size_t k = 1; k < idx.size() && shared; ++k)
      shared = same_state(b->ctx[idx[0]], b->ctx[idx[k]]) && dt[idx[k]] == dt[idx[0]];
    if (!shared) return false;
    Group G;
    for (int i : idx) G.m[G.n++] = b->ctx[i];
    hetddft_ctx* lead = G.m[0];
    // the members' own streams are idle before the lead's carry their data
    for (int k = 0; k < G.n; ++k) {
      hetddft_ctx* m = G.m[k];
      int rc = settle(m);
      if (rc == HETDDFT_OK && (hipStreamSynchronize(m->st) != hipSuccess || hipStreamSynchronize(m->st2) != hipSuccess))
        rc = fail(m, HETDDFT_E_HIP, "hipStreamSynchronize failed");
      if (rc != HETDDFT_OK) {
        record(idx[k], rc);
        return true;
      }
    }
    // flags of the members other than the lead follow the lead's streams
    for (int s = 0; s < steps && !stop; ++s) {
      const int rc = iterate_group(G, dt[idx[0]]);
      if (rc != HETDDFT_OK) {
        record(idx[0], rc);
        return true;
      }
    }
    int rc = settle(lead);
    if (rc == HETDDFT_OK && hipStreamSynchronize(lead->st) != hipSuccess)
      rc = fail(lead, HETDDFT_E_HIP, "hipStreamSynchronize failed");
    if (rc != HETDDFT_OK) {
      record(idx[0], rc);
      return true;
    }
    for (int k = 0; k < G.n; ++k) G.m[k]->side_pending = false;
    return true;
  };
  auto drive = [&](const std::vector<int>& mine) {
    std::vector<int> singles;
    for (int gi : mine)
      if (!drive_group(b->groups[gi]))
        for (int i : b->groups[gi]) singles.push_back(i);
    for (int s = 0; s < steps && !stop; ++s)
      for (int i : singles) {
        const int rc = hetddft_iterate(b->ctx[i], dt[i], nullptr);
        if (rc != HETDDFT_OK) {
          record(i, rc);
          return;
        }
      }
    // every instance's last estimates are reduced and checked here
    for (int gi : mine)
      for (int i : b->groups[gi]) {
        const int rc = hetddft_sync(b->ctx[i]);
        if (rc != HETDDFT_OK) {
          record(i, rc);
          return;
        }
      }
  };
  std::vector<std::thread> th;
  th.reserve(b->part.size());
  for (size_t t = 1; t < b->part.size(); ++t) th.emplace_back(drive, std::cref(b->part[t]));
  drive(b->part[0]);
  for (auto& x : th) x.join();
  if (first_i >= 0) {
    b->err = std::string("instance ") + std::to_string(first_i) + ": " + hetddft_last_error(b->ctx[first_i]);
    if (failed) *failed = first_i;
    return first_rc;
  }
  if (err)
    for (int i = 0; i < n; ++i) {
      if (b->ctx[i]->iter_in_run <= 3) {  // the last iteration was an RK step
        err[3 * i] = err[3 * i + 1] = err[3 * i + 2] = 0.0;
        continue;
      }
      const int rc = hetddft_last_errors(b->ctx[i], err + 3 * i);
      if (rc != HETDDFT_OK) {
        b->err = std::string("instance ") + std::to_string(i) + ": " + hetddft_last_error(b->ctx[i]);
        if (failed) *failed = i;
        return rc;
      }
    }
  b->err.clear();
  return HETDDFT_OK;
}

const char* hetddft_batch_last_error(const hetddft_batch* b) { return b ? b->err.c_str() : ""; }

int hetddft_batch_destroy(hetddft_batch* b) {
  delete b;
  return HETDDFT_OK;
}

}  // extern "C"

static int time_steps(hetddft_ctx* c, double dt, int steps, double* ms) {
  RET(hetddft_iterate(c, dt, nullptr));
  RET(settle(c));
  HIPCHK(c, hipStreamSynchronize(c->st));
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < steps; ++i) RET(hetddft_iterate(c, dt, nullptr));
  RET(settle(c));
  HIPCHK(c, hipStreamSynchronize(c->st));
  const auto t1 = std::chrono::steady_clock::now();
  *ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / steps;
  return HETDDFT_OK;
}

static int time_classes(hetddft_ctx* c, double dt, int steps, double ms[HETDDFT_PROF_N]) {
  RET(hetddft_iterate(c, dt, nullptr));
  RET(hetddft_profile(c, 1));
  for (int i = 0; i < steps; ++i) RET(hetddft_iterate(c, dt, nullptr));
  for (int k = 0; k < HETDDFT_PROF_N; ++k) {
    double t = 0.0;
    long long n = 0;
    RET(hetddft_profile_read(c, k, &t, &n, nullptr));
    ms[k] = n > 0 ? t / (double)n : 0.0;
  }
  return hetddft_profile(c, 0);
}
