// Batched 3D real-to-complex / complex-to-real transform passes.
//
// The reference runs 22 FFTW r2c/c2r transforms per potential evaluation
// (poten.f90, term_alfa.f90, potenimp.f90; plans in fft.f90:43-65).  Here a
// 3D convolution is five batched passes over HBM:
//   XF  real rows -> half-length complex FFT + Hermitian split (r2c along x),
//       with the pointwise producer of the transformed field fused on load,
//       written transposed so that ky is the fastest index;
//   YF  complex FFT along y on contiguous lines;
//   Z   forward FFT along z, multiplication by one or more real k-space
//       kernels, inverse FFT along z, one output spectrum per kernel;
//   YI  inverse complex FFT along y on contiguous lines;
//   XI  transposed read, Hermitian merge + half-length inverse FFT (c2r
//       along x), times renor = 1/(nx ny nz) (fft.f90:125-207).
// Every launch carries several fields so one launch fills the chip.
//
// Internal half spectra are stored S[kz][kx][ky] (ky fastest, kx < nx/2+1);
// k-space kernels use the same order.  Every global access moves rows of at
// least 16 consecutive complex values.
//
// Kernels come in a compile-time-length form (N > 0: radix plan, tiles and
// index arithmetic fold to constants; radix stages run in place through
// registers, one LDS buffer) and a runtime-plan form (N == 0, two LDS
// buffers) for every other size of the tstgrid list.
#pragma once
#include "fft_device.h"

namespace hetd {

constexpr int kMaxBatch = 8;
constexpr int kMaxJobs = 12;
constexpr int kThreads = 256;
constexpr int kRowsX = 16;   // rows (consecutive y) per x-pass workgroup
constexpr int kLinesY = 16;  // contiguous lines per y-pass workgroup

// Output planes of a convolution job.  A job whose output is read only inside
// a box (the force convolutions, which the trilinear gather reads at the
// particles' cell corners) carries a device pointer to that box, (x0, y0, z0,
// x1, y1, z1) inclusive and 0-based, empty when z1 < z0; the inverse passes
// then produce its z planes z0..z1 only.  After the z pass the spectra are
// indexed by real-space z and the x+y transforms of a plane use that plane
// alone, so the planes produced hold the bits of the full transform.  A null
// box: every plane.  The box is read on the device, where the particle chain
// wrote it, so the host needs no copy of it.
struct ZRange {
  int lo, hi;
};
__device__ inline ZRange z_range(const int* box, int nz) {
  if (!box) return {0, nz - 1};
  return {box[2], box[5]};
}
// The bytes a pass moves for the planes of its boxed jobs are counted on the
// device (the host does not know the boxes), into *pbytes when profiling is
// on (non-null): the first thread of the launch adds plane_bytes per plane
// of every boxed job, once.
__device__ inline bool first_thread() {
  return threadIdx.x == 0 && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0;
}
template <class A>
__device__ inline void count_boxed_planes(const A& a, int njobs, int nz, long long plane_bytes) {
  if (!a.pbytes || !first_thread()) return;
  long long b = 0;
  for (int j = 0; j < njobs; ++j)
    if (a.zbox[j]) {
      const ZRange r = z_range(a.zbox[j], nz);
      b += (long long)max(r.hi - r.lo + 1, 0) * plane_bytes;
    }
  if (b > 0) atomicAdd(a.pbytes, (unsigned long long)b);
}

// Consecutive ky per z-pass workgroup: 256-byte rows up to nz = 128, half
// that for longer (or runtime-length) lines to bound the LDS tile.
template <class T, int N>
__host__ __device__ constexpr int tile_z() {
  return (N > 0 && N <= 128) ? (sizeof(T) == 8 ? 8 : 16) : (sizeof(T) == 8 ? 8 : 16);
}

// ---- compile-time radix plan ---------------------------------------------
__host__ __device__ constexpr int pick_radix(int m) {
  return (m % 8 == 0) ? 8 : (m % 4 == 0) ? 4 : (m % 2 == 0) ? 2 : (m % 5 == 0) ? 5 : (m % 3 == 0) ? 3 : m;
}

// Radix for the next stage of a length-n transform over L lines on THR
// threads: 16 first for n >= 128 (one LDS pass fewer than 8 x 8 x 2 for
// 128) when the stage still gives at least half the threads a butterfly.
__host__ __device__ constexpr int pick_radix_for(int m, int n, int L, int thr) {
  return (m % 16 == 0 && n >= 128 && 2 * L * (n / 16) >= thr) ? 16 : pick_radix(m);
}

// One in-place Stockham stage over L lines: every thread loads its
// butterflies into registers, the block synchronises, then stores.
template <class T, bool INV, int R, int N, int NS, int L>
__device__ inline void stage_ip(cplx<T>* buf, int ls, const cplx<T>* __restrict__ tw) {
  constexpr int NB = N / R;
  constexpr int TWS = N / (NS * R);
  constexpr int TOT = L * NB;
  constexpr int PER = (TOT + kThreads - 1) / kThreads;
  cplx<T> v[PER][R];
  int dst[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = threadIdx.x + q * kThreads;
    dst[q] = -1;
    if (TOT % kThreads == 0 || b < TOT) {
      const int l = b / NB;
      const int j = b - l * NB;
      const cplx<T>* s = buf + l * ls;
      const int k = j % NS;
#pragma unroll
      for (int t = 0; t < R; ++t) v[q][t] = s[j + t * NB];
      if constexpr (NS > 1) {
#pragma unroll
        for (int t = 1; t < R; ++t) {
          const cplx<T> w = tw[t * k * TWS];
          v[q][t] = INV ? tmulc(v[q][t], w) : tmul(v[q][t], w);
        }
      }
      if constexpr (R == 2) bfly2<T, INV>(v[q]);
      if constexpr (R == 3) bfly3<T, INV>(v[q]);
      if constexpr (R == 4) bfly4<T, INV>(v[q]);
      if constexpr (R == 5) bfly5<T, INV>(v[q]);
      if constexpr (R == 8) bfly8<T, INV>(v[q]);
      if constexpr (R == 16) bfly16<T, INV>(v[q]);
      dst[q] = l * ls + (j / NS) * NS * R + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    if (TOT % kThreads == 0 || dst[q] >= 0) {
#pragma unroll
      for (int t = 0; t < R; ++t) buf[dst[q] + t * NS] = v[q][t];
    }
  }
  __syncthreads();
}

template <class T, bool INV, int N, int NS, int REM, int L>
__device__ inline void stages_ip(cplx<T>* buf, int ls, const cplx<T>* __restrict__ tw) {
  if constexpr (REM > 1) {
    constexpr int R = pick_radix_for(REM, N, L, kThreads);
    stage_ip<T, INV, R, N, NS, L>(buf, ls, tw);
    stages_ip<T, INV, N, NS * R, REM / R, L>(buf, ls, tw);
  }
}

// Transform L lines held in b0.  Compile-time N runs in place and returns
// b0; N == 0 ping-pongs through b1 and returns whichever holds the result.
// Callers must synchronise before the first call (data in LDS).
template <class T, bool INV, int N, int L>
__device__ inline cplx<T>* fft_tile(cplx<T>* b0, cplx<T>* b1, int ls, const FftPlan& pl,
                                    const cplx<T>* __restrict__ tw) {
  if constexpr (N > 0) {
    stages_ip<T, INV, N, 1, N, L>(b0, ls, tw);
    return b0;
  } else {
    const int r = fft_lines<T, INV>(b0, b1, L, ls, pl, tw);
    return r ? b1 : b0;
  }
}

// LDS bytes of a tile of L lines of length n (line stride n+1).
template <class T>
__host__ __device__ constexpr size_t tile_bytes(int L, int n) {
  return (size_t)L * (n + 1) * sizeof(cplx<T>);
}

// Pointwise producers fused into the forward x pass (the array the reference
// hands to dfftw_execute is formed on the fly from its inputs).
enum XfOp : int {
  XF_COPY = 0,      // f = a
  XF_CORREL = 1,    // f = den*dencg*(cp4+cpp4*dencg)           (poten.f90:62-64)
  XF_MUL = 2,       // f = a*b, e.g. falfs*d                    (term_alfa.f90:32-36)
  XF_GRAD_DOT = 3,  // f = dx*ix + dy*iy + dz*iz                (term_alfa.f90:65-69)
};

template <class R>
struct XfArgs {
  const R* in[kMaxBatch][6];
  int op[kMaxBatch];
  void* out[kMaxBatch];     // cplx<T>* half spectra
  double c[kMaxBatch][2];   // per-field op constants: (cp4, cpp4)
};

template <class R>
__device__ inline double xf_value(const XfArgs<R>& a, int f, long long i) {
  switch (a.op[f]) {
    case XF_CORREL: {
      double den = (double)a.in[f][0][i], dcg = (double)a.in[f][1][i];
      return den * dcg * (a.c[f][0] + a.c[f][1] * dcg);
    }
    case XF_MUL:
      return (double)a.in[f][0][i] * (double)a.in[f][1][i];
    case XF_GRAD_DOT: {
      double s = (double)a.in[f][0][i] * (double)a.in[f][3][i] +
                 (double)a.in[f][1][i] * (double)a.in[f][4][i];
      return s + (double)a.in[f][2][i] * (double)a.in[f][5][i];
    }
    default:
      return (double)a.in[f][0][i];
  }
}

// r2c along x for kRowsX consecutive y rows of one z plane.
// grid = (ceil(ny/kRowsX), nz, nfields).  MH > 0: compile-time nx/2.
template <class T, class R, int MH>
__global__ void __launch_bounds__(kThreads) k_fft_xf(XfArgs<R> a, Grid g, FftPlan pl,
                                                     const cplx<T>* __restrict__ tw,
                                                     const cplx<T>* __restrict__ twh) {
  extern __shared__ unsigned char smem_raw[];
  constexpr int ROWS = kRowsX;
  const int m = MH > 0 ? MH : g.nx / 2;
  const int nx = 2 * m;
  const int ls = m + 1;
  cplx<T>* b0 = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* b1 = b0 + ROWS * ls;
  const int f = blockIdx.z;
  const int y0 = blockIdx.x * ROWS;
  const int iz = blockIdx.y;
  const int nrow = (g.ny - y0) < ROWS ? (g.ny - y0) : ROWS;
  T* b0r = reinterpret_cast<T*>(b0);
  const long long src0 = ((long long)iz * g.ny + y0) * nx;
  for (int e = threadIdx.x; e < ROWS * nx; e += kThreads) {
    const int r = e / nx, c = e - r * nx;
    T v = T(0);
    if (r < nrow) v = (T)xf_value(a, f, src0 + e);
    b0r[2 * (r * ls) + c] = v;
  }
  __syncthreads();
  cplx<T>* Z = fft_tile<T, false, MH, ROWS>(b0, b1, ls, pl, tw);
  // Hermitian split in place, one thread per (k, m-k) pair:
  // E = (Z[k] + conj Z[m-k])/2, O = (Z[k] - conj Z[m-k])/(2i), X[k] = E + W^k O
  const int half = m / 2 + 1;
  for (int e = threadIdx.x; e < ROWS * half; e += kThreads) {
    const int r = e / half, k = e - r * half;
    cplx<T>* z = Z + r * ls;
    const int k2 = m - k;  // == m for k == 0
    const cplx<T> za = z[k];
    const cplx<T> zb = z[k == 0 ? 0 : k2];
    {
      const cplx<T> zc = conj(zb);
      const cplx<T> E = scale(za + zc, T(0.5));
      const cplx<T> D = za - zc;
      const cplx<T> O = mk(T(0.5) * D.y, T(-0.5) * D.x);
      z[k] = E + tmul(twh[k], O);
    }
    if (k2 != k) {
      const cplx<T> zc = conj(za);
      const cplx<T> E = scale(zb + zc, T(0.5));
      const cplx<T> D = zb - zc;
      const cplx<T> O = mk(T(0.5) * D.y, T(-0.5) * D.x);
      z[k2] = E + tmul(twh[k2], O);
    }
  }
  __syncthreads();
  // transposed store: S[iz][kx][y0 + r]
  cplx<T>* out = reinterpret_cast<cplx<T>*>(a.out[f]) + (long long)iz * g.mx * g.ny + y0;
  const int M = m + 1;
  for (int e = threadIdx.x; e < ROWS * M; e += kThreads) {
    const int kx = e / ROWS, r = e - kx * ROWS;
    if (r < nrow) out[(long long)kx * g.ny + r] = Z[r * ls + kx];
  }
}

// c2r along x: transposed read of kRowsX consecutive ky columns, real rows
// out, times renor.  grid = (ceil(ny/kRowsX), nz, nfields).
// Epilogues of the inverse x pass: the c2r result v = z*renor is stored as
// is, or as 1 - v/c (falfs = 1 - denalf/den0s, term_alfa.f90:28-30).
enum XiEpi : int { XI_STORE = 0, XI_ONE_MINUS_DIV = 1 };

template <class T, class R>
struct XiArgs {
  const void* in[kMaxBatch];   // cplx<T>* half spectra
  R* out[kMaxBatch];
  int epi[kMaxBatch];
  double ec[kMaxBatch];
  const int* zbox[kMaxBatch];  // output planes (z_range), null: all
  unsigned long long* pbytes;  // bytes of the boxed jobs (profiling), or null
  double renor;
};

// Whether job f needs plane iz.
template <class T, class R>
__device__ inline bool xi_plane_needed(const XiArgs<T, R>& a, int f, int iz, int nz) {
  if (!a.zbox[f]) return true;
  const ZRange r = z_range(a.zbox[f], nz);
  return iz >= r.lo && iz <= r.hi;
}

template <class T, class R, int MH>
__global__ void __launch_bounds__(kThreads) k_fft_xi(XiArgs<T, R> a, Grid g, FftPlan pl,
                                                     const cplx<T>* __restrict__ tw,
                                                     const cplx<T>* __restrict__ twh) {
  extern __shared__ unsigned char smem_raw[];
  constexpr int ROWS = kRowsX;
  const int m = MH > 0 ? MH : g.nx / 2;
  const int nx = 2 * m;
  const int ls = m + 1;
  const int M = m + 1;
  cplx<T>* b0 = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* b1 = b0 + ROWS * ls;
  const int f = blockIdx.z;
  const int y0 = blockIdx.x * ROWS;
  const int iz = blockIdx.y;
  const int nrow = (g.ny - y0) < ROWS ? (g.ny - y0) : ROWS;
  count_boxed_planes(a, gridDim.z, g.nz, (long long)g.ny * (M * sizeof(cplx<T>) + nx * sizeof(R)));
  if (!xi_plane_needed(a, f, iz, g.nz)) return;
  const cplx<T>* in = reinterpret_cast<const cplx<T>*>(a.in[f]) + (long long)iz * g.mx * g.ny + y0;
  for (int e = threadIdx.x; e < ROWS * M; e += kThreads) {
    const int kx = e / ROWS, r = e - kx * ROWS;
    b0[r * ls + kx] = (r < nrow) ? in[(long long)kx * g.ny + r] : mk(T(0), T(0));
  }
  __syncthreads();
  // Hermitian merge in place:
  // Y[k] = (X[k] + conj X[m-k]) + i (X[k] - conj X[m-k]) W^{-k}
  const int half = m / 2 + 1;
  for (int e = threadIdx.x; e < ROWS * half; e += kThreads) {
    const int r = e / half, k = e - r * half;
    cplx<T>* x = b0 + r * ls;
    const int k2 = m - k;
    const cplx<T> xa = x[k], xb = x[k2];
    {
      const cplx<T> xc = conj(xb);
      const cplx<T> s = xa + xc;
      const cplx<T> d = tmulc(xa - xc, twh[k]);
      x[k] = mk(s.x - d.y, s.y + d.x);
    }
    if (k != 0 && k2 != k) {
      const cplx<T> xc = conj(xa);
      const cplx<T> s = xb + xc;
      const cplx<T> d = tmulc(xb - xc, twh[k2]);
      x[k2] = mk(s.x - d.y, s.y + d.x);
    }
  }
  __syncthreads();
  const cplx<T>* Zc = fft_tile<T, true, MH, ROWS>(b0, b1, ls, pl, tw);
  const T* zr = reinterpret_cast<const T*>(Zc);
  R* out = a.out[f] + ((long long)iz * g.ny + y0) * nx;
  const T renor = (T)a.renor;
  const int epi = a.epi[f];
  const double ec = a.ec[f];
  for (int e = threadIdx.x; e < nrow * nx; e += kThreads) {
    const int r = e / nx, c = e - r * nx;
    const T v = zr[2 * (r * ls) + c] * renor;
    out[e] = epi == XI_ONE_MINUS_DIV ? (R)(1.0 - (double)(R)v / ec) : (R)v;
  }
}

// Complex FFT along y on contiguous lines (line q = kz*mx + kx, length ny).
// grid = (ceil(mx*nz/kLinesY), 1, nfields).  In place.
template <class T>
struct YArgs {
  void* buf[kMaxBatch];
  const int* zbox[kMaxBatch];  // inverse pass: output planes (z_range), null: all
  unsigned long long* pbytes;  // bytes of the boxed jobs (profiling), or null
};

// Inverse y pass: whether a workgroup's lines q0 .. q0+nl-1 (line q lies in
// plane q / mx) reach the planes its job needs.
template <class T>
__device__ inline bool y_lines_needed(const YArgs<T>& a, int f, const Grid& g, int q0, int nl, int n) {
  count_boxed_planes(a, gridDim.z, g.nz, 2LL * g.mx * n * (long long)sizeof(cplx<T>));
  if (!a.zbox[f]) return true;
  const ZRange r = z_range(a.zbox[f], g.nz);
  return q0 / g.mx <= r.hi && (q0 + nl - 1) / g.mx >= r.lo;
}

template <class T, bool INV, int N>
__global__ void __launch_bounds__(kThreads) k_fft_y(YArgs<T> a, Grid g, FftPlan pl, const cplx<T>* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  constexpr int L = kLinesY;
  const int n = N > 0 ? N : g.ny;
  const int ls = n + 1;
  cplx<T>* b0 = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* b1 = b0 + L * ls;
  const int nlines = g.mx * g.nz;
  const int q0 = blockIdx.x * L;
  const int nl = (nlines - q0) < L ? (nlines - q0) : L;
  if (INV && !y_lines_needed(a, blockIdx.z, g, q0, nl, n)) return;
  cplx<T>* base = reinterpret_cast<cplx<T>*>(a.buf[blockIdx.z]) + (long long)q0 * n;
  for (int e = threadIdx.x; e < L * n; e += kThreads) {
    const int l = e / n, j = e - l * n;
    b0[l * ls + j] = (l < nl) ? base[e] : mk(T(0), T(0));
  }
  __syncthreads();
  const cplx<T>* Y = fft_tile<T, INV, N, L>(b0, b1, ls, pl, tw);
  for (int e = threadIdx.x; e < nl * n; e += kThreads) {
    const int l = e / n, j = e - l * n;
    base[e] = Y[l * ls + j];
  }
}

// z pass of a convolution, one job per workgroup: forward FFT along z of
// the job's input, multiplication by sign*kernel(kx,ky,kz), inverse FFT along
// z, store to the job's output spectrum.  Tile = TY consecutive ky at fixed
// kx.  grid = (ceil(ny/TY), mx, njobs).
template <class T>
struct ZArgs {
  const void* in[kMaxJobs];
  const T* kern[kMaxJobs];
  T ksign[kMaxJobs];
  void* out[kMaxJobs];
  const int* zbox[kMaxJobs];   // output planes (z_range), null: all
  unsigned long long* pbytes;  // bytes written for the boxed jobs (profiling), or null
};

template <class T, int N, int TY>
__global__ void __launch_bounds__(kThreads) k_fft_zconv(ZArgs<T> a, Grid g, FftPlan pl,
                                                        const cplx<T>* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  const int n = N > 0 ? N : g.nz;
  const int ls = n + 1;
  cplx<T>* b0 = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* b1 = b0 + TY * ls;                             // runtime-plan scratch
  const long long pstride = (long long)g.mx * g.ny;      // stride between kz planes
  const int ky0 = blockIdx.x * TY;
  const int nty = (g.ny - ky0) < TY ? (g.ny - ky0) : TY;
  const int kx = blockIdx.y;
  const int jb = blockIdx.z;
  const long long off = (long long)kx * g.ny + ky0;
  const cplx<T>* src = reinterpret_cast<const cplx<T>*>(a.in[jb]) + off;
  for (int e = threadIdx.x; e < TY * n; e += kThreads) {
    const int j = e / TY, l = e - j * TY;
    b0[l * ls + j] = (l < nty) ? src[(long long)j * pstride + l] : mk(T(0), T(0));
  }
  __syncthreads();
  cplx<T>* F = fft_tile<T, false, N, TY>(b0, b1, ls, pl, tw);
  const T* kern = a.kern[jb] + off;
  const T sg = a.ksign[jb];
  for (int e = threadIdx.x; e < TY * n; e += kThreads) {
    const int j = e / TY, l = e - j * TY;
    if (l < nty) {
      const cplx<T> v = F[l * ls + j];
      const T kv = kern[(long long)j * pstride + l];
      const cplx<T> w = mk(v.x * kv, v.y * kv);
      F[l * ls + j] = sg < T(0) ? mk(-w.x, -w.y) : w;
    }
  }
  __syncthreads();
  cplx<T>* other = (F == b0) ? b1 : b0;
  const cplx<T>* Y = fft_tile<T, true, N, TY>(F, other, ls, pl, tw);
  cplx<T>* dst = reinterpret_cast<cplx<T>*>(a.out[jb]) + off;
  // a boxed job: its planes only
  const ZRange zr = z_range(a.zbox[jb], n);
  const unsigned span = (unsigned)(zr.hi - zr.lo);
  count_boxed_planes(a, gridDim.z, n, pstride * (long long)sizeof(cplx<T>));
  for (int e = threadIdx.x; e < TY * n; e += kThreads) {
    const int j = e / TY, l = e - j * TY;
    if (l < nty && (unsigned)(j - zr.lo) <= span) dst[(long long)j * pstride + l] = Y[l * ls + j];
  }
}

// Z pass, persistent form for compile-time lengths.  gridDim.x workgroups walk
// the work items (ky tile, kx, job), ky tile fastest, with stride gridDim.x.
// While one item is transformed in LDS, the next item's column tile and
// kernel values are already on their way into registers, so HBM traffic
// overlaps the LDS work instead of every workgroup loading, transforming and
// storing in lockstep.  Twiddles are staged in LDS once: no other global load
// is outstanding behind the prefetch during the transform.  Arithmetic is
// that of k_fft_zconv (same stages, twiddles and kernel product).
template <class T, int N, int TY>
__host__ __device__ constexpr size_t zpipe_bytes() {
  return (size_t)(TY * (N + 1) + N) * sizeof(cplx<T>);
}

template <class T, int N, int TY>
__global__ void __launch_bounds__(kThreads) k_fft_zpipe(ZArgs<T> a, int njobs, Grid g,
                                                        const cplx<T>* __restrict__ twg) {
  static_assert(N > 0 && kThreads % TY == 0, "tile width must divide the workgroup");
  extern __shared__ unsigned char smem_raw[];
  constexpr int ls = N + 1;
  constexpr int RSTEP = kThreads / TY;            // rows between a thread's values
  constexpr int PER = (N + RSTEP - 1) / RSTEP;    // values per thread
  constexpr bool FULL = N % RSTEP == 0;           // every thread's last row exists
  cplx<T>* tile = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* tw = tile + TY * ls;
  const int l = threadIdx.x % TY;         // ky within the tile
  const int j0 = threadIdx.x / TY;        // first kz row of this thread
  const int ntile = (g.ny + TY - 1) / TY;
  const int per_job = ntile * g.mx;
  const int total = per_job * njobs;
  const long long pstride = (long long)g.mx * g.ny;
  for (int i = threadIdx.x; i < N; i += kThreads) tw[i] = twg[i];

  cplx<T> v[PER];
  T kv[PER];
  auto fetch = [&](int item) {
    const int jb = item / per_job;
    const int r = item - jb * per_job;
    const int kx = r / ntile;
    const int ky = (r - kx * ntile) * TY + l;
    const bool ok = ky < g.ny;
    const long long off = (long long)kx * g.ny + (ok ? ky : 0);
    const cplx<T>* src = reinterpret_cast<const cplx<T>*>(a.in[jb]) + off;
    const T* kp = a.kern[jb] + off;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const bool in = ok && (FULL || j0 + q * RSTEP < N);
      const long long o = in ? (long long)(j0 + q * RSTEP) * pstride : 0;
      v[q] = in ? src[o] : mk(T(0), T(0));
      kv[q] = in ? kp[o] : T(0);
    }
  };

  count_boxed_planes(a, njobs, N, pstride * (long long)sizeof(cplx<T>));
  int w = blockIdx.x;
  if (w < total) fetch(w);
  for (; w < total; w += gridDim.x) {
    const int jb = w / per_job;
    const int r = w - jb * per_job;
    const int kx = r / ntile;
    const int ky = (r - kx * ntile) * TY + l;
    T kc[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      if (FULL || j0 + q * RSTEP < N) tile[l * ls + j0 + q * RSTEP] = v[q];
      kc[q] = kv[q];
    }
    __syncthreads();
    if (w + (int)gridDim.x < total) fetch(w + gridDim.x);
    stages_ip<T, false, N, 1, N, TY>(tile, ls, tw);
    const T sg = a.ksign[jb];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      if (FULL || j0 + q * RSTEP < N) {
        cplx<T>& x = tile[l * ls + j0 + q * RSTEP];
        const cplx<T> p = mk(x.x * kc[q], x.y * kc[q]);
        x = sg < T(0) ? mk(-p.x, -p.y) : p;
      }
    }
    __syncthreads();
    stages_ip<T, true, N, 1, N, TY>(tile, ls, tw);
    const int* zb = a.zbox[jb];
    if (ky < g.ny) {
      cplx<T>* dst = reinterpret_cast<cplx<T>*>(a.out[jb]) + (long long)kx * g.ny + ky;
      if (!zb) {
#pragma unroll
        for (int q = 0; q < PER; ++q)
          if (FULL || j0 + q * RSTEP < N) dst[(long long)(j0 + q * RSTEP) * pstride] = tile[l * ls + j0 + q * RSTEP];
      } else {
        // a boxed job: its planes only
        const ZRange zr = z_range(zb, N);
        const unsigned span = (unsigned)(zr.hi - zr.lo);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          const int kz = j0 + q * RSTEP;
          if ((FULL || kz < N) && (unsigned)(kz - zr.lo) <= span) dst[(long long)kz * pstride] = tile[l * ls + kz];
        }
      }
    }

    __syncthreads();
  }
}

// Radix-R butterfly on v[0..R) (R = 6, 8, 10, 12 or 16).
template <class T, bool INV, int R>
__device__ inline void bfly(cplx<T>* v) {
  static_assert(R == 6 || R == 8 || R == 10 || R == 12 || R == 16, "register radix");
  if constexpr (R == 6) bfly_ab<T, INV, 2, 3>(v);
  if constexpr (R == 8) bfly8<T, INV>(v);
  if constexpr (R == 10) bfly_ab<T, INV, 2, 5>(v);
  if constexpr (R == 12) bfly_ab<T, INV, 4, 3>(v);
  if constexpr (R == 16) bfly16<T, INV>(v);
}

// Z pass with the transforms held in registers between two LDS exchanges,
// for lengths N = R1 * R2 (R1 in {8, 16}; R2 in {8, 16}, or 6, 10, 12 for
// N = 48, 80, 96).  Per work item (TY consecutive
// ky of one kx, one job) the four Stockham stages run as:
//   A  forward radix-R1 stage straight from HBM (rows kz = j + R2 t), written
//      to LDS in Stockham order (j R1 + t);
//   B  forward radix-R2 stage from LDS (rows j + R1 t) with twiddles W_N^(t j):
//      the thread then holds X[j + R1 t], multiplies it by the kernel values
//      read from HBM,
//   C  and runs the first inverse stage (radix R2, no twiddles) on the same
//      registers, written to LDS (j R2 + t);
//   D  inverse radix-R1 stage from LDS (rows j + R2 t) with twiddles
//      conj W_N^(t j), stored straight to HBM (rows kz = j + R2 t).
// Four LDS passes per item instead of twelve.  Threads take consecutive ky
// first, so every HBM access moves rows of TY consecutive complex values and
// LDS accesses of a wavefront walk columns (stride N+1: no bank conflicts).
// Persistent over items like k_fft_zpipe; the next item's stage-A rows are
// loaded while the current item runs B, C and D.
// Workgroups are dispatched to the 8 XCDs in turn (workgroup p on XCD p % 8),
// and each XCD has its own L2.  xcd_order(p, nb) is p's place in an order
// that gives every XCD one contiguous range of [0, nb).
__device__ inline int xcd_order(int p, int nb) {
  constexpr int NXCD = 8;
  const int q = nb / NXCD, r = nb % NXCD, k = p % NXCD;
  return k * q + min(k, r) + p / NXCD;
}

template <class T, int N, int TY, int R1, int R2>
__host__ __device__ constexpr size_t zreg_bytes() {
  return (size_t)(TY * (N + 1) + N) * sizeof(cplx<T>);
}

template <class T, int N, int TY, int R1, int R2>
__global__ void __launch_bounds__(kThreads) k_fft_zreg(ZArgs<T> a, int njobs, Grid g,
                                                       const cplx<T>* __restrict__ twg) {
  static_assert(R1 * R2 == N && (R1 == 8 || R1 == 16) && (R2 == 6 || R2 == 8 || R2 == 10 || R2 == 12 || R2 == 16),
                "N = R1 * R2");
  extern __shared__ unsigned char smem_raw[];
  constexpr int ls = N + 1;
  constexpr int NA = TY * R2;  // stage A / D butterflies
  constexpr int NB = TY * R1;  // stage B / C butterflies
  constexpr int PA = (NA + kThreads - 1) / kThreads;
  constexpr int PB = (NB + kThreads - 1) / kThreads;
  cplx<T>* tile = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* tw = tile + TY * ls;
  const int tid = threadIdx.x;
  const int ntile = (g.ny + TY - 1) / TY;
  const int ncol = ntile * g.mx;  // column tiles (kx, ky tile)
  const int total = ncol * njobs;
  const long long pstride = (long long)g.mx * g.ny;
  for (int i = tid; i < N; i += kThreads) tw[i] = twg[i];

  // Items (column tile, job) in xcd_order, the jobs of a column tile next to
  // each other: each XCD walks its own range of column tiles, and the jobs
  // that share an input spectrum or a k-space kernel read it under one L2
  // at about the same time (+1.1% FP64 and +0.6% mixed PC steps/s at 128^3;
  // the same order without the XCD grouping, or the XCD grouping of the x+y
  // plane kernels, measured slower).
  struct Item {
    int jb, kx, ky0;
  };
  auto decode = [&](int w) {
    const int l = xcd_order(w, total);
    Item it;
    const int col = l / njobs;
    it.jb = l - col * njobs;
    it.kx = col / ntile;
    it.ky0 = (col - it.kx * ntile) * TY;
    return it;
  };
  // stage-A rows of an item: butterfly b = l + TY j, elements kz = j + R2 t
  cplx<T> va[PA][R1];
  auto load_a = [&](const Item& it) {
    const cplx<T>* src = reinterpret_cast<const cplx<T>*>(a.in[it.jb]) + (long long)it.kx * g.ny + it.ky0;
#pragma unroll
    for (int q = 0; q < PA; ++q) {
      const int b = tid + q * kThreads;
      const int l = b % TY, j = b / TY;
      const bool ok = (NA % kThreads == 0 || b < NA) && it.ky0 + l < g.ny;
#pragma unroll
      for (int t = 0; t < R1; ++t) {
        const long long o = ok ? (long long)(j + R2 * t) * pstride + l : 0;
        va[q][t] = ok ? src[o] : mk(T(0), T(0));
      }
    }
  };

  count_boxed_planes(a, njobs, N, pstride * (long long)sizeof(cplx<T>));
  int w = blockIdx.x;
  if (w < total) load_a(decode(w));
  __syncthreads();  // twiddles
  for (; w < total; w += gridDim.x) {
    const Item it = decode(w);
    // A: forward radix R1 (first stage, no twiddles) -> LDS
#pragma unroll
    for (int q = 0; q < PA; ++q) {
      const int b = tid + q * kThreads;
      if (NA % kThreads == 0 || b < NA) {
        const int l = b % TY, j = b / TY;
        bfly<T, false, R1>(va[q]);
#pragma unroll
        for (int t = 0; t < R1; ++t) tile[l * ls + j * R1 + t] = va[q][t];
      }
    }
    __syncthreads();
    if (w + (int)gridDim.x < total) load_a(decode(w + gridDim.x));
    // B reads (kernel values from HBM first, so their latency overlaps)
    const T* kern = a.kern[it.jb] + (long long)it.kx * g.ny + it.ky0;
    const T sg = a.ksign[it.jb];
    cplx<T> u[PB][R2];
    T kv[PB][R2];
#pragma unroll
    for (int q = 0; q < PB; ++q) {
      const int b = tid + q * kThreads;
      const int l = b % TY, j = b / TY;
      const bool ok = (NB % kThreads == 0 || b < NB) && it.ky0 + l < g.ny;
#pragma unroll
      for (int t = 0; t < R2; ++t) {
        const long long o = ok ? (long long)(j + R1 * t) * pstride + l : 0;
        kv[q][t] = ok ? kern[o] : T(0);
      }
      if (NB % kThreads == 0 || b < NB) {
#pragma unroll
        for (int t = 0; t < R2; ++t) u[q][t] = tile[l * ls + j + R1 * t];
      }
    }
    // every stage-B read of a column precedes the stage-C writes into it
    __syncthreads();
    // B: twiddles + forward radix R2; kernel product; C: inverse radix R2
#pragma unroll
    for (int q = 0; q < PB; ++q) {
      const int b = tid + q * kThreads;
      if (NB % kThreads == 0 || b < NB) {
        const int l = b % TY, j = b / TY;
#pragma unroll
        for (int t = 1; t < R2; ++t) u[q][t] = tmul(u[q][t], tw[t * j]);
        bfly<T, false, R2>(u[q]);
#pragma unroll
        for (int t = 0; t < R2; ++t) {
          const cplx<T> pr = mk(u[q][t].x * kv[q][t], u[q][t].y * kv[q][t]);
          u[q][t] = sg < T(0) ? mk(-pr.x, -pr.y) : pr;
        }
        bfly<T, true, R2>(u[q]);
#pragma unroll
        for (int t = 0; t < R2; ++t) tile[l * ls + j * R2 + t] = u[q][t];
      }
    }
    __syncthreads();
    // D: twiddles + inverse radix R1 -> HBM (a boxed job: its planes only)
    cplx<T>* dst = reinterpret_cast<cplx<T>*>(a.out[it.jb]) + (long long)it.kx * g.ny + it.ky0;
    const int* zb = a.zbox[it.jb];
#pragma unroll
    for (int q = 0; q < PA; ++q) {
      const int b = tid + q * kThreads;
      if (NA % kThreads == 0 || b < NA) {
        const int l = b % TY, j = b / TY;
        cplx<T> v[R1];
#pragma unroll
        for (int t = 0; t < R1; ++t) v[t] = tile[l * ls + j + R2 * t];
#pragma unroll
        for (int t = 1; t < R1; ++t) v[t] = tmulc(v[t], tw[t * j]);
        bfly<T, true, R1>(v);
        if (it.ky0 + l < g.ny) {
          if (!zb) {
#pragma unroll
            for (int t = 0; t < R1; ++t) dst[(long long)(j + R2 * t) * pstride + l] = v[t];
          } else {
            const ZRange zr = z_range(zb, N);
            const unsigned span = (unsigned)(zr.hi - zr.lo);
            const int jl = j - zr.lo;
#pragma unroll
            for (int t = 0; t < R1; ++t)
              if ((unsigned)(jl + R2 * t) <= span) dst[(long long)(j + R2 * t) * pstride + l] = v[t];
          }
        }
      }
    }

    __syncthreads();  // stage-D reads done before the next item's stage-A writes
  }
}

// Y pass with the transform held in registers around one LDS exchange, for
// lengths N = R1 * R2 (R1, R2 in {8, 16}).  A workgroup takes kLinesY
// contiguous lines (line q = kz*mx + kx, length ny, elements contiguous):
//   A  radix-R1 stage straight from HBM (elements j + R2 t), written to LDS
//      at j (R1+1) + t (the +1 keeps consecutive j in different banks);
//   B  radix-R2 stage from LDS (elements j + R1 t) with twiddles W_N^(t j)
//      (conjugate for the inverse); the thread then holds X[j + R1 t] and
//      stores it straight to HBM.
// Threads take consecutive j first, so every HBM access moves runs of R2 (A)
// or R1 (B) consecutive complex values of a line.  In place: a workgroup
// reads all of its lines before it writes any.  grid = (ceil(mx*nz/kLinesY),
// 1, nfields).
template <class T, int N, int R1, int R2>
__host__ __device__ constexpr size_t yreg_bytes() {
  return (size_t)(kLinesY * R2 * (R1 + 1) + N) * sizeof(cplx<T>);
}

template <class T, bool INV, int N, int R1, int R2>
__global__ void __launch_bounds__(kThreads) k_fft_yreg(YArgs<T> a, Grid g, FftPlan pl,
                                                       const cplx<T>* __restrict__ twg) {
  static_assert(R1 * R2 == N && (R1 == 8 || R1 == 16) && (R2 == 8 || R2 == 16), "N = R1 * R2");
  constexpr int TL = kLinesY;
  constexpr int LS = R2 * (R1 + 1);  // LDS words of a line, in complex values
  constexpr int NA = TL * R2;        // stage-A butterflies
  constexpr int NB = TL * R1;        // stage-B butterflies
  constexpr int PA = (NA + kThreads - 1) / kThreads;
  constexpr int PB = (NB + kThreads - 1) / kThreads;
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* tile = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* tw = tile + TL * LS;
  const int tid = threadIdx.x;
  (void)pl;
  const int nlines = g.mx * g.nz;
  const int q0 = blockIdx.x * TL;
  if (INV && !y_lines_needed(a, blockIdx.z, g, q0, min(TL, nlines - q0), N)) return;
  for (int i = tid; i < N; i += kThreads) tw[i] = twg[i];
  cplx<T>* base = reinterpret_cast<cplx<T>*>(a.buf[blockIdx.z]) + (long long)q0 * N;
  // A: loads of every line of the workgroup, radix R1, to LDS
  cplx<T> v[PA][R1];
#pragma unroll
  for (int q = 0; q < PA; ++q) {
    const int b = tid + q * kThreads;
    const int j = b % R2, l = b / R2;
    const bool ok = (NA % kThreads == 0 || b < NA) && q0 + l < nlines;
#pragma unroll
    for (int t = 0; t < R1; ++t) v[q][t] = ok ? base[(long long)l * N + j + R2 * t] : mk(T(0), T(0));
  }
#pragma unroll
  for (int q = 0; q < PA; ++q) {
    const int b = tid + q * kThreads;
    if (NA % kThreads == 0 || b < NA) {
      const int j = b % R2, l = b / R2;
      bfly<T, INV, R1>(v[q]);
#pragma unroll
      for (int t = 0; t < R1; ++t) tile[l * LS + j * (R1 + 1) + t] = v[q][t];
    }
  }
  __syncthreads();
  // B: twiddles, radix R2, to HBM
#pragma unroll
  for (int q = 0; q < PB; ++q) {
    const int b = tid + q * kThreads;
    if (NB % kThreads == 0 || b < NB) {
      const int j = b % R1, l = b / R1;
      cplx<T> u[R2];
#pragma unroll
      for (int t = 0; t < R2; ++t) u[t] = tile[l * LS + t * (R1 + 1) + j];
#pragma unroll
      for (int t = 1; t < R2; ++t) u[t] = INV ? tmulc(u[t], tw[t * j]) : tmul(u[t], tw[t * j]);
      bfly<T, INV, R2>(u);
      if (q0 + l < nlines) {
#pragma unroll
        for (int t = 0; t < R2; ++t) base[(long long)l * N + j + R1 * t] = u[t];
      }
    }
  }
}

// Producer pairs z[t] = (f[2n], f[2n+1]), n = j + STRIDE t (t < CNT), of one
// row of the x pass: the XF_* producer of field f, formed as in xf_value.
template <class T, class R, int CNT, int STRIDE>
__device__ inline void xrow_pairs(const XfArgs<R>& a, int f, long long row0, int j, cplx<T>* z) {
  auto ld = [&](int k, int t) { return reinterpret_cast<const cplx<R>*>(a.in[f][k] + row0)[j + STRIDE * t]; };
  switch (a.op[f]) {
    case XF_CORREL: {
      const double c0 = a.c[f][0], c1 = a.c[f][1];
      auto fn = [&](R den, R dcg) {
        const double d = (double)den, g = (double)dcg;
        return (T)(d * g * (c0 + c1 * g));
      };
#pragma unroll
      for (int t = 0; t < CNT; ++t) {
        const cplx<R> u = ld(0, t), w = ld(1, t);
        z[t] = mk(fn(u.x, w.x), fn(u.y, w.y));
      }
      break;
    }
    case XF_MUL: {
#pragma unroll
      for (int t = 0; t < CNT; ++t) {
        const cplx<R> u = ld(0, t), w = ld(1, t);
        z[t] = mk((T)u.x * (T)w.x, (T)u.y * (T)w.y);  // the correctly rounded product
      }
      break;
    }
    case XF_GRAD_DOT: {
#pragma unroll
      for (int t = 0; t < CNT; ++t) {
        cplx<R> v[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) v[k] = ld(k, t);
        const double sx = (double)v[0].x * (double)v[3].x + (double)v[1].x * (double)v[4].x;
        const double sy = (double)v[0].y * (double)v[3].y + (double)v[1].y * (double)v[4].y;
        z[t] = mk((T)(sx + (double)v[2].x * (double)v[5].x), (T)(sy + (double)v[2].y * (double)v[5].y));
      }
      break;
    }
    default: {
#pragma unroll
      for (int t = 0; t < CNT; ++t) {
        const cplx<R> u = ld(0, t);
        z[t] = mk((T)u.x, (T)u.y);
      }
      break;
    }
  }
}

// X pass (r2c along x, nx = 2 MH, MH = R1 * R2) with the half-length
// transform in registers, for kRowsX consecutive y rows of one z plane:
//   A  radix R1 on the producer pairs z[n], n = j + R2 t, straight from HBM,
//      to LDS at j (R1+1) + t;
//   B  radix R2 from LDS (elements j + R1 t) with twiddles W_MH^(t j): the
//      thread holds Z[j + R1 t] and writes it to LDS in natural order;
// then the Hermitian split and the transposed store of k_fft_xf.  grid =
// (ceil(ny/kRowsX), nz, nfields); tw = exp(-2 pi i k/MH), twh = exp(-2 pi i k/nx).
template <class T, int MH, int R1, int R2>
__host__ __device__ constexpr size_t xreg_bytes() {
  constexpr int a = kRowsX * R2 * (R1 + 1), b = kRowsX * (MH + 1);
  return (size_t)(a > b ? a : b) * sizeof(cplx<T>);
}

template <class T, class R, int MH, int R1, int R2>
__global__ void __launch_bounds__(kThreads) k_fft_xreg(XfArgs<R> a, Grid g, FftPlan pl,
                                                       const cplx<T>* __restrict__ tw,
                                                       const cplx<T>* __restrict__ twh) {
  static_assert(R1 * R2 == MH && (R1 == 8 || R1 == 16) && (R2 == 8 || R2 == 16), "MH = R1 * R2");
  constexpr int ROWS = kRowsX, NX = 2 * MH, M = MH + 1, ls = MH + 1;
  constexpr int LS = R2 * (R1 + 1);
  constexpr int NA = ROWS * R2, NB = ROWS * R1;
  constexpr int PA = (NA + kThreads - 1) / kThreads, PB = (NB + kThreads - 1) / kThreads;
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* tile = reinterpret_cast<cplx<T>*>(smem_raw);
  (void)pl;
  const int tid = threadIdx.x;
  const int f = blockIdx.z;
  const int y0 = blockIdx.x * ROWS;
  const int iz = blockIdx.y;
  const int nrow = (g.ny - y0) < ROWS ? (g.ny - y0) : ROWS;
  // A
  cplx<T> v[PA][R1];
#pragma unroll
  for (int q = 0; q < PA; ++q) {
    const int b = tid + q * kThreads;
    const int j = b % R2, r = b / R2;
    if ((NA % kThreads == 0 || b < NA) && r < nrow) {
      xrow_pairs<T, R, R1, R2>(a, f, ((long long)iz * g.ny + y0 + r) * NX, j, v[q]);
    } else {
#pragma unroll
      for (int t = 0; t < R1; ++t) v[q][t] = mk(T(0), T(0));
    }
  }
#pragma unroll
  for (int q = 0; q < PA; ++q) {
    const int b = tid + q * kThreads;
    if (NA % kThreads == 0 || b < NA) {
      const int j = b % R2, r = b / R2;
      bfly<T, false, R1>(v[q]);
#pragma unroll
      for (int t = 0; t < R1; ++t) tile[r * LS + j * (R1 + 1) + t] = v[q][t];
    }
  }
  __syncthreads();
  // B
  cplx<T> u[PB][R2];
#pragma unroll
  for (int q = 0; q < PB; ++q) {
    const int b = tid + q * kThreads;
    if (NB % kThreads == 0 || b < NB) {
      const int j = b % R1, r = b / R1;
#pragma unroll
      for (int t = 0; t < R2; ++t) u[q][t] = tile[r * LS + t * (R1 + 1) + j];
#pragma unroll
      for (int t = 1; t < R2; ++t) u[q][t] = tmul(u[q][t], tw[t * j]);
      bfly<T, false, R2>(u[q]);
    }
  }
  __syncthreads();  // every stage-B read precedes the natural-order writes below
  cplx<T>* Z = tile;
#pragma unroll
  for (int q = 0; q < PB; ++q) {
    const int b = tid + q * kThreads;
    if (NB % kThreads == 0 || b < NB) {
      const int j = b % R1, r = b / R1;
#pragma unroll
      for (int t = 0; t < R2; ++t) Z[r * ls + j + R1 * t] = u[q][t];
    }
  }
  __syncthreads();
  // Hermitian split in place, one thread per (k, MH-k) pair (as k_fft_xf)
  constexpr int half = MH / 2 + 1;
  for (int e = tid; e < ROWS * half; e += kThreads) {
    const int r = e / half, k = e - r * half;
    cplx<T>* z = Z + r * ls;
    const int k2 = MH - k;  // == MH for k == 0
    const cplx<T> za = z[k];
    const cplx<T> zb = z[k == 0 ? 0 : k2];
    {
      const cplx<T> zc = conj(zb);
      const cplx<T> E = scale(za + zc, T(0.5));
      const cplx<T> D = za - zc;
      const cplx<T> O = mk(T(0.5) * D.y, T(-0.5) * D.x);
      z[k] = E + tmul(twh[k], O);
    }
    if (k2 != k) {
      const cplx<T> zc = conj(za);
      const cplx<T> E = scale(zb + zc, T(0.5));
      const cplx<T> D = zb - zc;
      const cplx<T> O = mk(T(0.5) * D.y, T(-0.5) * D.x);
      z[k2] = E + tmul(twh[k2], O);
    }
  }
  __syncthreads();
  // transposed store: S[iz][kx][y0 + r]
  cplx<T>* out = reinterpret_cast<cplx<T>*>(a.out[f]) + (long long)iz * g.mx * g.ny + y0;
  for (int e = tid; e < ROWS * M; e += kThreads) {
    const int kx = e / ROWS, r = e - kx * ROWS;
    if (r < nrow) out[(long long)kx * g.ny + r] = Z[r * ls + kx];
  }
}

// Inverse x pass (c2r, nx = 2 MH) with the half-length transform in
// registers: the transposed read and Hermitian merge of k_fft_xi, then
//   A  inverse radix R1 from LDS (elements j + R2 t), written back over the
//      merged rows at j (R1+1) + t (one LDS buffer: more workgroups per CU);
//   B  conjugate twiddles W_MH^(-t j), inverse radix R2: the thread holds
//      z[j + R1 t] = (x[2n], x[2n+1]) and stores it straight to HBM, times
//      renor, with the XI epilogue.
template <class T, int MH, int R1, int R2>
__host__ __device__ constexpr size_t xireg_bytes() {
  constexpr int a = kRowsX * (MH + 1), b = kRowsX * R2 * (R1 + 1);
  return (size_t)(a > b ? a : b) * sizeof(cplx<T>);
}

template <class T, class R, int MH, int R1, int R2>
__global__ void __launch_bounds__(kThreads) k_fft_xireg(XiArgs<T, R> a, Grid g, FftPlan pl,
                                                        const cplx<T>* __restrict__ tw,
                                                        const cplx<T>* __restrict__ twh) {
  static_assert(R1 * R2 == MH && (R1 == 8 || R1 == 16) && (R2 == 8 || R2 == 16), "MH = R1 * R2");
  constexpr int ROWS = kRowsX, NX = 2 * MH, M = MH + 1, ls = MH + 1;
  constexpr int LS = R2 * (R1 + 1);
  constexpr int NA = ROWS * R2, NB = ROWS * R1;
  constexpr int PA = (NA + kThreads - 1) / kThreads, PB = (NB + kThreads - 1) / kThreads;
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* b0 = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* tile = b0;
  (void)pl;
  const int tid = threadIdx.x;
  const int f = blockIdx.z;
  const int y0 = blockIdx.x * ROWS;
  const int iz = blockIdx.y;
  const int nrow = (g.ny - y0) < ROWS ? (g.ny - y0) : ROWS;
  count_boxed_planes(a, gridDim.z, g.nz, (long long)g.ny * (M * sizeof(cplx<T>) + NX * sizeof(R)));
  if (!xi_plane_needed(a, f, iz, g.nz)) return;
  const cplx<T>* in = reinterpret_cast<const cplx<T>*>(a.in[f]) + (long long)iz * g.mx * g.ny + y0;
  {
    // transposed read: every load of the thread is issued before the first
    // LDS store, so the reads of a row block are in flight together
    constexpr int NL = (ROWS * M + kThreads - 1) / kThreads;
    cplx<T> ld[NL];
#pragma unroll
    for (int q = 0; q < NL; ++q) {
      const int e = tid + q * kThreads;
      const int kx = e / ROWS, r = e - kx * ROWS;
      ld[q] = (e < ROWS * M && r < nrow) ? in[(long long)kx * g.ny + r] : mk(T(0), T(0));
    }
#pragma unroll
    for (int q = 0; q < NL; ++q) {
      const int e = tid + q * kThreads;
      const int kx = e / ROWS, r = e - kx * ROWS;
      if (e < ROWS * M) b0[r * ls + kx] = ld[q];
    }
  }
  __syncthreads();
  // Hermitian merge in place (as k_fft_xi)
  constexpr int half = MH / 2 + 1;
  for (int e = tid; e < ROWS * half; e += kThreads) {
    const int r = e / half, k = e - r * half;
    cplx<T>* x = b0 + r * ls;
    const int k2 = MH - k;
    const cplx<T> xa = x[k], xb = x[k2];
    {
      const cplx<T> xc = conj(xb);
      const cplx<T> s = xa + xc;
      const cplx<T> d = tmulc(xa - xc, twh[k]);
      x[k] = mk(s.x - d.y, s.y + d.x);
    }
    if (k != 0 && k2 != k) {
      const cplx<T> xc = conj(xa);
      const cplx<T> s = xb + xc;
      const cplx<T> d = tmulc(xb - xc, twh[k2]);
      x[k2] = mk(s.x - d.y, s.y + d.x);
    }
  }
  __syncthreads();
  // A
  cplx<T> v[PA][R1];
#pragma unroll
  for (int q = 0; q < PA; ++q) {
    const int b = tid + q * kThreads;
    if (NA % kThreads == 0 || b < NA) {
      const int j = b % R2, r = b / R2;
#pragma unroll
      for (int t = 0; t < R1; ++t) v[q][t] = b0[r * ls + j + R2 * t];
      bfly<T, true, R1>(v[q]);
    }
  }
  __syncthreads();  // every merged value is read before the tile overwrites the rows
#pragma unroll
  for (int q = 0; q < PA; ++q) {
    const int b = tid + q * kThreads;
    if (NA % kThreads == 0 || b < NA) {
      const int j = b % R2, r = b / R2;
#pragma unroll
      for (int t = 0; t < R1; ++t) tile[r * LS + j * (R1 + 1) + t] = v[q][t];
    }
  }
  __syncthreads();
  // B -> HBM
  const T renor = (T)a.renor;
  const bool omd = a.epi[f] == XI_ONE_MINUS_DIV;
  const double ec = a.ec[f];
#pragma unroll
  for (int q = 0; q < PB; ++q) {
    const int b = tid + q * kThreads;
    if (NB % kThreads == 0 || b < NB) {
      const int j = b % R1, r = b / R1;
      cplx<T> u[R2];
#pragma unroll
      for (int t = 0; t < R2; ++t) u[t] = tile[r * LS + t * (R1 + 1) + j];
#pragma unroll
      for (int t = 1; t < R2; ++t) u[t] = tmulc(u[t], tw[t * j]);
      bfly<T, true, R2>(u);
      if (r < nrow) {
        cplx<R>* dst = reinterpret_cast<cplx<R>*>(a.out[f] + ((long long)iz * g.ny + y0 + r) * NX);
#pragma unroll
        for (int t = 0; t < R2; ++t) {
          const T vx = u[t].x * renor, vy = u[t].y * renor;
          cplx<R> o;
          o.x = omd ? (R)(1.0 - (double)(R)vx / ec) : (R)vx;
          o.y = omd ? (R)(1.0 - (double)(R)vy / ec) : (R)vy;
          dst[j + R1 * t] = o;
        }
      }
    }
  }
}

}  // namespace hetd

namespace hetd {

// ---- whole-plane kernels ---------------------------------------------------
// When one z plane of the half spectrum (nx/2+1 by ny complex) fits in LDS,
// the x and y transforms run in one pass: XYF reads the real plane once and
// writes the plane's (kx, ky) spectrum as one contiguous block; XYI does the
// reverse.  This drops the separate y passes (two full spectrum sweeps per
// convolution).

constexpr int kPlaneThreads = 1024;

// In-place Stockham stage with compile-time line stride LS and element
// stride ES over L lines, THR threads.
template <class T, bool INV, int R, int N, int NS, int L, int THR, int LS, int ES>
__device__ inline void stage_g(cplx<T>* buf, const cplx<T>* __restrict__ tw) {
  constexpr int NB = N / R;
  constexpr int TWS = N / (NS * R);
  constexpr int TOT = L * NB;
  constexpr int PER = (TOT + THR - 1) / THR;
  cplx<T> v[PER][R];
  int dst[PER];
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = tid + q * THR;
    dst[q] = -1;
    if (TOT % THR == 0 || b < TOT) {
      const int l = b / NB;
      const int j = b - l * NB;
      const cplx<T>* s = buf + l * LS;
      const int k = j % NS;
#pragma unroll
      for (int t = 0; t < R; ++t) v[q][t] = s[(j + t * NB) * ES];
      if constexpr (NS > 1) {
#pragma unroll
        for (int t = 1; t < R; ++t) {
          const cplx<T> w = tw[t * k * TWS];
          v[q][t] = INV ? tmulc(v[q][t], w) : tmul(v[q][t], w);
        }
      }
      if constexpr (R == 2) bfly2<T, INV>(v[q]);
      if constexpr (R == 3) bfly3<T, INV>(v[q]);
      if constexpr (R == 4) bfly4<T, INV>(v[q]);
      if constexpr (R == 5) bfly5<T, INV>(v[q]);
      if constexpr (R == 6 || R == 8 || R == 10 || R == 12 || R == 16) bfly<T, INV, R>(v[q]);
      dst[q] = l * LS + ((j / NS) * NS * R + k) * ES;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    if (TOT % THR == 0 || dst[q] >= 0) {
#pragma unroll
      for (int t = 0; t < R; ++t) buf[dst[q] + t * NS * ES] = v[q][t];
    }
  }
  __syncthreads();
}

// Radix of the next plane-kernel stage: a remaining length of 6, 10 or 12
// is one register butterfly (one LDS pass instead of two).
__host__ __device__ constexpr int pick_radix_g(int m, int n, int L, int thr) {
  return (m == 6 || m == 10 || m == 12) ? m : pick_radix_for(m, n, L, thr);
}

template <class T, bool INV, int N, int NS, int REM, int L, int THR, int LS, int ES>
__device__ inline void stages_g(cplx<T>* buf, const cplx<T>* __restrict__ tw) {
  if constexpr (REM > 1) {
    constexpr int R = pick_radix_g(REM, N, L, THR);
    static_assert(R == 2 || R == 3 || R == 4 || R == 5 || R == 6 || R == 8 || R == 10 || R == 12 || R == 16,
                  "plane-kernel radix");
    stage_g<T, INV, R, N, NS, L, THR, LS, ES>(buf, tw);
    stages_g<T, INV, N, NS * R, REM / R, L, THR, LS, ES>(buf, tw);
  }
}

template <class T>
__host__ __device__ constexpr size_t plane_bytes(int mh, int ny) {
  return (size_t)(mh + 1) * ny * sizeof(cplx<T>);
}

// Producer values of a real z plane (NY rows of NX) into LDS rows of stride
// 2M reals.  Each thread's loads are all issued before the first value is
// used (in chunks for the six-input producer), so a workgroup keeps its whole
// plane in flight; the arithmetic is that of xf_value.
template <class T, class R, int NX, int NY, int M, int THR>
__device__ inline void load_plane_real(const XfArgs<R>& a, int f, long long src0, T* Pr) {
  constexpr int NE = NX * NY;
  constexpr int CNT = (NE + THR - 1) / THR;
  constexpr bool FULL = NE % THR == 0;
  const int tid = threadIdx.x;
  // element q of this thread exists (the last round may be partial)
  auto has = [&](int q) { return FULL || tid + q * THR < NE; };
  auto put = [&](int q, double v) {
    if (!has(q)) return;
    const int e = tid + q * THR;
    const int y = e / NX, c = e - y * NX;
    Pr[2 * M * y + c] = (T)v;
  };
  // index of element q, clamped to a valid element for the partial round
  auto at = [&](int q) { return src0 + (has(q) ? tid + q * THR : 0); };
  switch (a.op[f]) {
    case XF_CORREL: {
      R u[CNT], w[CNT];
#pragma unroll
      for (int q = 0; q < CNT; ++q) {
        u[q] = a.in[f][0][at(q)];
        w[q] = a.in[f][1][at(q)];
      }
      const double c0 = a.c[f][0], c1 = a.c[f][1];
#pragma unroll
      for (int q = 0; q < CNT; ++q) {
        const double den = (double)u[q], dcg = (double)w[q];
        put(q, den * dcg * (c0 + c1 * dcg));
      }
      break;
    }
    case XF_MUL: {
      R u[CNT], w[CNT];
#pragma unroll
      for (int q = 0; q < CNT; ++q) {
        u[q] = a.in[f][0][at(q)];
        w[q] = a.in[f][1][at(q)];
      }
#pragma unroll
      for (int q = 0; q < CNT; ++q) put(q, (double)u[q] * (double)w[q]);
      break;
    }
    case XF_GRAD_DOT: {
      constexpr int CH = CNT >= 4 ? 4 : CNT;
#pragma unroll
      for (int q0 = 0; q0 < CNT; q0 += CH) {
        R v[6][CH];
#pragma unroll
        for (int k = 0; k < 6; ++k)
#pragma unroll
          for (int q = 0; q < CH; ++q)
            if (q0 + q < CNT) v[k][q] = a.in[f][k][at(q0 + q)];
#pragma unroll
        for (int q = 0; q < CH; ++q) {
          if (q0 + q >= CNT) continue;
          const double s = (double)v[0][q] * (double)v[3][q] + (double)v[1][q] * (double)v[4][q];
          put(q0 + q, s + (double)v[2][q] * (double)v[5][q]);
        }
      }
      break;
    }
    default: {
      R u[CNT];
#pragma unroll
      for (int q = 0; q < CNT; ++q) u[q] = a.in[f][0][at(q)];
#pragma unroll
      for (int q = 0; q < CNT; ++q) put(q, (double)u[q]);
      break;
    }
  }
}

// r2c in x and y of one z plane.  grid = (nz, nfields), THR threads (at
// least NY).  LDS P[ky][kx], row length M = MH+1 (real rows hold 2*MH values
// first).
template <class T, class R, int MH, int NY, int THR>
__global__ void __launch_bounds__(THR) k_fft_xyf(XfArgs<R> a, Grid g, const cplx<T>* __restrict__ twx,
                                                 const cplx<T>* __restrict__ twh,
                                                 const cplx<T>* __restrict__ twy) {
  static_assert(THR >= NY, "the column-0 unpack takes one thread per ky");
  extern __shared__ unsigned char smem_raw[];
  constexpr int M = MH + 1;
  constexpr int NX = 2 * MH;
  cplx<T>* P = reinterpret_cast<cplx<T>*>(smem_raw);
  T* Pr = reinterpret_cast<T*>(smem_raw);
  const int iz = blockIdx.x;
  const int f = blockIdx.y;
  const long long src0 = (long long)iz * NY * NX;
  load_plane_real<T, R, NX, NY, M, THR>(a, f, src0, Pr);
  __syncthreads();
  stages_g<T, false, MH, 1, MH, NY, THR, M, 1>(P, twx);
  // Hermitian split per row, one item per pair (k, MH-k), k = 1..MH/2.  The
  // k = MH/2 item also forms the DC and Nyquist values (both real) and
  // stores them packed as X0 + i X_MH in column 0, so that the y transform
  // runs over MH columns: one full round of butterflies per stage.
  constexpr int QH = MH / 2;
  for (int e = threadIdx.x; e < NY * QH; e += THR) {
    const int y = e / QH, k = 1 + (e - y * QH);
    cplx<T>* z = P + y * M;
    const int k2 = MH - k;
    const cplx<T> za = z[k];
    const cplx<T> zb = z[k2];
    {
      const cplx<T> zc = conj(zb);
      const cplx<T> E = scale(za + zc, T(0.5));
      const cplx<T> D = za - zc;
      const cplx<T> O = mk(T(0.5) * D.y, T(-0.5) * D.x);
      z[k] = E + tmul(twh[k], O);
    }
    if (k2 != k) {
      const cplx<T> zc = conj(za);
      const cplx<T> E = scale(zb + zc, T(0.5));
      const cplx<T> D = zb - zc;
      const cplx<T> O = mk(T(0.5) * D.y, T(-0.5) * D.x);
      z[k2] = E + tmul(twh[k2], O);
    }
    if (k == QH) {
      const cplx<T> z0 = z[0];
      z[0] = mk(z0.x + z0.y, z0.x - z0.y);
    }
  }
  __syncthreads();
  stages_g<T, false, NY, 1, NY, MH, THR, 1, M>(P, twy);
  // unpack column 0: W = FFT(X0) + i FFT(X_MH) with both inputs real, so
  // FFT(X0)[k] = (W[k] + conj W[-k])/2 and FFT(X_MH)[k] = -i (W[k] - conj W[-k])/2
  for (int ky = threadIdx.x; ky <= NY / 2; ky += THR) {
    const int km = (NY - ky) % NY;
    const cplx<T> wa = P[ky * M], wb = P[km * M];
    const cplx<T> sa = wa + conj(wb), da = wa - conj(wb);
    const cplx<T> sb = wb + conj(wa), db = wb - conj(wa);
    P[ky * M] = scale(sa, T(0.5));
    P[ky * M + MH] = mk(T(0.5) * da.y, T(-0.5) * da.x);
    if (km != ky) {
      P[km * M] = scale(sb, T(0.5));
      P[km * M + MH] = mk(T(0.5) * db.y, T(-0.5) * db.x);
    }
  }
  __syncthreads();
  cplx<T>* out = reinterpret_cast<cplx<T>*>(a.out[f]) + (long long)iz * M * NY;
  for (int e = threadIdx.x; e < M * NY; e += THR) {
    const int kx = e / NY, ky = e - kx * NY;
    out[e] = P[ky * M + kx];
  }
}

// c2r in y and x of one z plane, times renor, with the XI epilogues.  THR
// threads, at least NY (the threads of kx = 0 also read column MH).
template <class T, class R, int MH, int NY, int THR>
__global__ void __launch_bounds__(THR) k_fft_xyi(XiArgs<T, R> a, Grid g, const cplx<T>* __restrict__ twx,
                                                 const cplx<T>* __restrict__ twh,
                                                 const cplx<T>* __restrict__ twy) {
  static_assert(THR >= NY, "the first NY elements must be the threads' first elements");
  extern __shared__ unsigned char smem_raw[];
  constexpr int M = MH + 1;
  constexpr int NX = 2 * MH;
  cplx<T>* P = reinterpret_cast<cplx<T>*>(smem_raw);
  const T* Pr = reinterpret_cast<const T*>(smem_raw);
  const int iz = blockIdx.x;
  const int f = blockIdx.y;
  count_boxed_planes(a, gridDim.y, g.nz, (long long)M * NY * sizeof(cplx<T>) + (long long)NY * NX * sizeof(R));
  if (!xi_plane_needed(a, f, iz, g.nz)) return;
  const cplx<T>* in = reinterpret_cast<const cplx<T>*>(a.in[f]) + (long long)iz * M * NY;
  // Columns kx = 0 and MH have Hermitian symmetry in ky (their inverse y
  // transforms are real): they travel packed as col0 + i col_MH in column 0,
  // so the y transform runs over MH columns.
  {
    constexpr int NE = MH * NY;
    constexpr int CNT = (NE + THR - 1) / THR;
    constexpr bool FULL = NE % THR == 0;
    cplx<T> v[CNT];
#pragma unroll
    for (int q = 0; q < CNT; ++q) {
      const int e = threadIdx.x + q * THR;
      v[q] = (FULL || e < NE) ? in[e] : mk(T(0), T(0));
    }
    // threads holding kx = 0 (the first NY elements) also read column MH
    cplx<T> u = mk(T(0), T(0));
    if (threadIdx.x < NY) u = in[MH * NY + threadIdx.x];
#pragma unroll
    for (int q = 0; q < CNT; ++q) {
      const int e = threadIdx.x + q * THR;
      if (!(FULL || e < NE)) continue;
      const int kx = e / NY, ky = e - kx * NY;
      cplx<T> w = v[q];
      if (kx == 0) w = mk(w.x - u.y, w.y + u.x);
      P[ky * M + kx] = w;
    }
  }
  __syncthreads();
  stages_g<T, true, NY, 1, NY, MH, THR, 1, M>(P, twy);
  // Hermitian merge per row, one item per pair (k, MH-k), k = 1..MH/2; the
  // k = MH/2 item also merges the real DC and Nyquist values held in column 0.
  constexpr int QH = MH / 2;
  for (int e = threadIdx.x; e < NY * QH; e += THR) {
    const int y = e / QH, k = 1 + (e - y * QH);
    cplx<T>* x = P + y * M;
    const int k2 = MH - k;
    const cplx<T> xa = x[k], xb = x[k2];
    {
      const cplx<T> xc = conj(xb);
      const cplx<T> s = xa + xc;
      const cplx<T> d = tmulc(xa - xc, twh[k]);
      x[k] = mk(s.x - d.y, s.y + d.x);
    }
    if (k2 != k) {
      const cplx<T> xc = conj(xa);
      const cplx<T> s = xb + xc;
      const cplx<T> d = tmulc(xb - xc, twh[k2]);
      x[k2] = mk(s.x - d.y, s.y + d.x);
    }
    if (k == QH) {
      const cplx<T> w = x[0];
      const cplx<T> x0 = mk(w.x, T(0)), xm = mk(w.y, T(0));
      const cplx<T> s = x0 + xm;
      const cplx<T> d = tmulc(x0 - xm, twh[0]);
      x[0] = mk(s.x - d.y, s.y + d.x);
    }
  }
  __syncthreads();
  stages_g<T, true, MH, 1, MH, NY, THR, M, 1>(P, twx);
  R* out = a.out[f] + (long long)iz * NY * NX;
  const T renor = (T)a.renor;
  const int epi = a.epi[f];
  const double ec = a.ec[f];
  for (int e = threadIdx.x; e < NY * NX; e += THR) {
    const int y = e / NX, c = e - y * NX;
    const T v = Pr[2 * M * y + c] * renor;
    out[e] = epi == XI_ONE_MINUS_DIV ? (R)(1.0 - (double)(R)v / ec) : (R)v;
  }
}

// ---- register-resident kernels for 128 x 128 planes ------------------------
// The same r2c / c2r of one z plane as k_fft_xyf / k_fft_xyi, with every
// radix stage in registers.  Per plane (1024 threads, 16 wavefronts):
//   x   each wavefront owns 8 rows; a row's 64-point half-length transform is
//       8 x 8: radix 8 straight from HBM, one exchange inside the wavefront,
//       radix 8.  The Hermitian split needs X[k] and X[64-k]: the stage-B
//       lanes of a row hold k1 in the order 1 7 2 6 3 5 0 4, so partner
//       columns sit in neighbouring lanes and one DPP lane swap brings them
//       together (k1 = 0 and 4 pair with themselves).
//   xy  one block-wide exchange through the plane (rows of 64 complex values,
//       column index xor-swizzled per row, no bank conflicts for the row
//       writes and the column reads);
//   y   each wavefront owns 4 columns; the 128-point transform is 8 x 16:
//       radix 8, one exchange inside the wavefront, radix 8, and a radix-2
//       step across neighbouring lanes (DPP swap), stored straight to HBM.
// XYI runs the same steps in reverse with conjugate twiddles.  Column 0
// carries DC + i Nyquist as in k_fft_xyf; its unpack is a fourth, small
// exchange inside wavefront 0.  tw[m] = exp(-2 pi i m/128), m < 128.
// Twiddles of the 128 x 128 plane kernels: tw[m] = W^m = exp(-2 pi i m/128),
// m < 128, and the products the radix stages take, regrouped so that a
// lane reads its twiddles of one stage at a lane base plus immediate
// offsets (no index arithmetic), neighbouring lanes reading neighbouring
// entries (no LDS bank conflicts):
//   TQ[16 m + q] = W^(q m)    (y transform, first stage; q < 16, m < 8)
//   TX[8 j + k1] = W^(2 j k1) (x transform, second stage; k1, j < 8)
//   TR[2 r + i0] = W^(8 r i0) (y transform, radix-2 step; i0 < 2, r < 8)
// The device table of the regrouped ones holds TQ, TX, TR in that order;
// the workgroup stages tw and the regrouped table together in LDS.
constexpr int kP128Regrouped = 128 + 64 + 16;
constexpr int kP128GQ = 0, kP128GX = 128, kP128GR = 192;               // in the device table
constexpr int kP128TQ = 128, kP128TX = 128 + 64 + 64, kP128TR = 128 + 192;  // in the LDS copy
constexpr int kPlane128Lds = (128 * 64 + 128 + kP128Regrouped);  // plane + staged twiddles, in complex values
// FP32 planes take 66.5 KB of LDS: two workgroups per CU when each thread
// keeps to 64 VGPRs (8 wavefronts per SIMD)
template <class T>
constexpr int kPlane128Waves = sizeof(T) == 4 ? 8 : 4;

// Wavefront-local LDS ordering: lanes of one wavefront run in lockstep and
// LDS executes one wavefront's instructions in order, so an exchange inside a
// wavefront needs only that the compiler keep its stores before its loads.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  __builtin_amdgcn_sched_barrier(0);
}

// value held by lane ^ 1 (DPP quad_perm [1,0,3,2])
__device__ inline float lane_swap1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ inline double lane_swap1(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, 0xB1, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xB1, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}
template <class T>
__device__ inline cplx<T> lane_swap1(cplx<T> v) {
  return mk(lane_swap1(v.x), lane_swap1(v.y));
}

// X[64 - k] for k = k1 + 8 k2: own value for k1 = 0 and 4, the partner
// lane's otherwise.  Selects values (a conditional over array elements would
// select addresses and keep the arrays out of registers).
template <class T>
__device__ inline cplx<T> pick3(int k1, cplx<T> a0, cplx<T> a4, cplx<T> ap) {
  const cplx<T> r = k1 == 4 ? a4 : ap;
  return k1 == 0 ? a0 : r;
}

// plane element (y, x), x < 64
__device__ inline int p128_at(int y, int x) {
  return y * 64 + (x ^ ((y & 15) ^ ((y & 1) << 3) ^ ((y >> 6) * 12)));
}
// x exchange of a wavefront: row yl < 8, first-stage butterfly j, output k1
__device__ inline int p128_xe(int yl, int j, int k1) { return yl * 64 + 8 * (j ^ yl) + (k1 ^ j); }
// y exchange of a wavefront: column c < 4, first-stage butterfly j < 16, output m1
__device__ inline int p128_ye(int c, int j, int m1) {
  const int rot = ((j & 1) << 2) | ((j >> 1) & 3);
  return c * 128 + 8 * (j ^ (2 * c)) + (m1 ^ rot ^ ((j >> 3) << 2));
}
// stage-B lane s of a row holds k1 = 1 7 2 6 3 5 0 4 (partner k1 of 8-k1 at s ^ 1)
__device__ inline int p128_k1(int s) { return (0x40536271 >> (4 * s)) & 7; }

// The same maps as a lane's base XOR a compile-time constant.  Each map
// packs fields of 3 bits that do not overlap, so XOR with a constant that
// touches only the fields indexed by the loop variable moves exactly that
// index:
//   p128_xe(yl, s, o)          = (yl*64 + 8 (s^yl) + s)   ^ o
//   p128_xe(yl, j, k1)         = (72 yl + k1)              ^ 9 j
//   p128_at(y, k1 + 8 k2)      = (64 y + (k1 ^ sw(y)))     ^ 8 k2
//   p128_ye(c, q, m)           = (128 c + 8 (q^2c) + (rot(q) ^ 4 (q>>3))) ^ m
//   p128_ye(c, i0 + 2 i1, m1)  = (128 c + 8 (i0^2c) + (m1 ^ 4 i0))        ^ 17 i1
// with sw(y) the row swizzle of p128_at.  Byte offsets (element index times
// the element size, a power of two) keep the form, and the wavefront's
// exchange area (512 elements from w*512) adds bits above the XOR, so
// every such access costs one XOR besides the LDS instruction.  Read rows
// q + 16 t of a column (q < 16) and write rows m1 + 8 r + 64 i0 (m1 < 8) of
// a column take two bases and immediate offsets.
__device__ inline int p128_sw(int y) { return (y & 15) ^ ((y & 1) << 3) ^ ((y >> 6) * 12); }
template <class T>
__device__ __forceinline__ cplx<T>& lds_at(cplx<T>* base, unsigned byte_off) {
  return *reinterpret_cast<cplx<T>*>(reinterpret_cast<char*>(base) + byte_off);
}

// Producer values z[t] = (f[2n], f[2n+1]), n = s + 8t, of one row of 128.
template <class T, class R>
__device__ inline void p128_load_row(const XfArgs<R>& a, int f, long long row0, int s, cplx<T>* z) {
  auto ld = [&](int k, int t) { return reinterpret_cast<const cplx<R>*>(a.in[f][k] + row0)[s + 8 * t]; };
  switch (a.op[f]) {
    case XF_CORREL: {
      cplx<R> u[8], w[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) u[t] = ld(0, t), w[t] = ld(1, t);
      const double c0 = a.c[f][0], c1 = a.c[f][1];
      auto fn = [&](R den, R dcg) {
        const double d = (double)den, g = (double)dcg;
        return (T)(d * g * (c0 + c1 * g));
      };
#pragma unroll
      for (int t = 0; t < 8; ++t) z[t] = mk(fn(u[t].x, w[t].x), fn(u[t].y, w[t].y));
      break;
    }
    case XF_MUL: {
      cplx<R> u[8], w[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) u[t] = ld(0, t), w[t] = ld(1, t);
#pragma unroll
      for (int t = 0; t < 8; ++t)
        z[t] = mk((T)u[t].x * (T)w[t].x, (T)u[t].y * (T)w[t].y);  // the correctly rounded product
      break;
    }
    case XF_GRAD_DOT: {
      // values per chunk: FP32 planes keep to 64 VGPRs; FP64 planes have room for 4
      constexpr int CH = sizeof(T) == 8 ? 4 : 2;
#pragma unroll
      for (int t0 = 0; t0 < 8; t0 += CH) {
        cplx<R> v[6][CH];
#pragma unroll
        for (int k = 0; k < 6; ++k)
#pragma unroll
          for (int q = 0; q < CH; ++q) v[k][q] = ld(k, t0 + q);
#pragma unroll
        for (int q = 0; q < CH; ++q) {
          const double sx = (double)v[0][q].x * (double)v[3][q].x + (double)v[1][q].x * (double)v[4][q].x;
          const double sy = (double)v[0][q].y * (double)v[3][q].y + (double)v[1][q].y * (double)v[4][q].y;
          z[t0 + q] = mk((T)(sx + (double)v[2][q].x * (double)v[5][q].x),
                         (T)(sy + (double)v[2][q].y * (double)v[5][q].y));
        }
        // keep the next chunk's loads below this one (bounds live registers)
        __builtin_amdgcn_sched_barrier(0);
      }
      break;
    }
    default: {
      cplx<R> u[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) u[t] = ld(0, t);
#pragma unroll
      for (int t = 0; t < 8; ++t) z[t] = mk((T)u[t].x, (T)u[t].y);
      break;
    }
  }
}

// The value itself, opaque to the optimiser (it cannot hoist what derives
// from it out of a loop).
__device__ __forceinline__ int opaque_int(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

// r2c in x and y of one 128 x 128 z plane from the producer values v of row
// y = 8w + lane/8 (complex s + 8t, s = lane%8), written to out (the plane's
// half spectrum).  twx: tw for the Hermitian split and txr: TX for the x
// pass (global or staged); tws: the staged copies in LDS (tw, then TQ, TX,
// TR at kP128TQ, kP128TX, kP128TR).  mid() runs after the barrier that
// completes the plane in LDS.
template <class T, class Mid>
__device__ __forceinline__ void xyf128_plane(cplx<T>* P, const cplx<T>* twx, const cplx<T>* txr,
                                             const cplx<T>* tws, int tid, cplx<T>* v, cplx<T>* out, Mid mid) {
  constexpr int NY = 128, MH = 64;
  const int w = tid >> 6, lane = tid & 63;
  // ---- x: thread (row y, lane s of the row)
  const int yl = lane >> 3, s = lane & 7, y = 8 * w + yl;
  const int k1 = p128_k1(s);
  constexpr unsigned E = sizeof(cplx<T>);
  bfly<T, false, 8>(v);  // v[o] = Y[j = s][o]
  {
    // p128_xe(yl, s, o) and p128_xe(yl, j, k1) in the wavefront's area
    const unsigned bs = (unsigned)(w * 512 + yl * 64 + 8 * (s ^ yl) + s) * E;
    const unsigned br = (unsigned)(w * 512 + 72 * yl + k1) * E;
#pragma unroll
    for (int o = 0; o < 8; ++o) lds_at(P, bs ^ (o * E)) = v[o];
    wave_lds_sync();
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = lds_at(P, br ^ (9 * j * E));
  }
#pragma unroll
  for (int j = 1; j < 8; ++j) v[j] = tmul(v[j], txr[8 * j + k1]);
  bfly<T, false, 8>(v);  // v[k2] = X[k1 + 8 k2]
  // Hermitian split Z[k] = E + W_128^k O with X[64 - k] from the partner lane
  {
    cplx<T> pv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) pv[q] = lane_swap1(v[q]);
    cplx<T> zk[8];
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) {
      const cplx<T> za = v[k2];
      const cplx<T> zb = pick3(k1, v[(8 - k2) & 7], v[7 - k2], pv[7 - k2]);
      const cplx<T> zc = conj(zb);
      const cplx<T> E = scale(za + zc, T(0.5));
      const cplx<T> D = za - zc;
      const cplx<T> O = mk(T(0.5) * D.y, T(-0.5) * D.x);
      zk[k2] = E + tmul(twx[k1 + 8 * k2], O);
      // k = 0: DC + i Nyquist, both real
      if (k2 == 0 && k1 == 0) zk[0] = mk(za.x + za.y, za.x - za.y);
    }
    // p128_at(y, k1 + 8 k2)
    const unsigned bp = (unsigned)(y * 64 + (k1 ^ p128_sw(y))) * E;
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) lds_at(P, bp ^ (8 * k2 * E)) = zk[k2];
  }
  __syncthreads();
  mid();
  // ---- y: thread (column kx = 4w + c, lane q of the column)
  const int c = lane >> 4, q = lane & 15, kx = 4 * w + c;
  {
    // p128_at(q + 16 t, kx): rows t >= 4 (y >> 6 = 1) flip the swizzle by 12
    const unsigned b0 = (unsigned)(q * 64 + (kx ^ p128_sw(q))) * E;
    const unsigned b1 = (unsigned)(q * 64 + (kx ^ p128_sw(q) ^ 12)) * E;
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = lds_at(P, (t < 4 ? b0 : b1) + 1024 * t * E);
  }
  __syncthreads();  // every plane read precedes the column exchanges below
  bfly<T, false, 8>(v);  // v[m1]
#pragma unroll
  for (int m = 1; m < 8; ++m) v[m] = tmul(v[m], tws[kP128TQ + 16 * m + q]);
  cplx<T>* ye = P + w * 512;
  const int m1 = q >> 1, i0 = q & 1;
  {
    // p128_ye(c, q, m) and p128_ye(c, i0 + 2 i1, m1) in the wavefront's area
    const int rot = ((q & 1) << 2) | ((q >> 1) & 3);
    const unsigned bs = (unsigned)(w * 512 + 128 * c + 8 * (q ^ (2 * c)) + (rot ^ ((q >> 3) << 2))) * E;
    const unsigned br = (unsigned)(w * 512 + 128 * c + 8 * (i0 ^ (2 * c)) + (m1 ^ (i0 << 2))) * E;
#pragma unroll
    for (int m = 0; m < 8; ++m) lds_at(P, bs ^ (m * E)) = v[m];
    wave_lds_sync();
#pragma unroll
    for (int i1 = 0; i1 < 8; ++i1) v[i1] = lds_at(P, br ^ (17 * i1 * E));
  }
  bfly<T, false, 8>(v);  // B_i0[r]
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = tmul(v[r], tws[kP128TR + 2 * r + i0]);
  // ky = m1 + 8 r + 64 i0: v + o (i0 = 0) or o - v (i0 = 1), as one fma
  // with the exact factor +-1
  {
    const T sg = i0 ? T(-1) : T(1);
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const cplx<T> o = lane_swap1(v[r]);
      v[r] = mk(fmad(v[r].x, sg, o.x), fmad(v[r].y, sg, o.y));
    }
  }
  if (w == 0) {
    // column 0 holds FFT(X0) + i FFT(X_MH) of two real columns:
    // FFT(X0)[k] = (W[k] + conj W[-k])/2, FFT(X_MH)[k] = -i (W[k] - conj W[-k])/2
    if (c == 0) {
#pragma unroll
      for (int r = 0; r < 8; ++r) ye[m1 + 8 * r + 64 * i0] = v[r];
    }
    wave_lds_sync();
    if (c == 0) {
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int ky = m1 + 8 * r + 64 * i0;
        const cplx<T> wa = v[r], wb = ye[(NY - ky) & (NY - 1)];
        const cplx<T> sa = wa + conj(wb), da = wa - conj(wb);
        out[ky] = scale(sa, T(0.5));
        out[MH * NY + ky] = mk(T(0.5) * da.y, T(-0.5) * da.x);
      }
    }
  }
  if (kx != 0) {
#pragma unroll
    for (int r = 0; r < 8; ++r) out[kx * NY + m1 + 8 * r + 64 * i0] = v[r];
  }
}

// Stages tw (128 values) and the regrouped table twp (TQ, TX, TR) in LDS.
template <class T>
__device__ __forceinline__ void p128_stage_twiddles(cplx<T>* tws, const cplx<T>* __restrict__ tw,
                                                    const cplx<T>* __restrict__ twp, int tid) {
  if (tid < 128) tws[tid] = tw[tid];
  else if (tid < 128 + kP128Regrouped) tws[tid] = twp[tid - 128];
}

// r2c in x and y of one 128 x 128 z plane.  grid = (nz, nfields), 1024 threads.
// twp: the regrouped twiddles (TQ, TX, TR), tw: W^m, m < 128.  FP32 reads
// every twiddle from the LDS copy (13-15% faster than global reads at
// 128^3); FP64 reads the x-pass twiddles from global memory, which spares it
// a barrier.
template <class T, class R>
__global__ void __launch_bounds__(1024, kPlane128Waves<T>) k_fft_xyf128(XfArgs<R> a, Grid g, const cplx<T>* __restrict__ twp,
                                                     const cplx<T>* __restrict__ twh,
                                                     const cplx<T>* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  constexpr int NX = 128, NY = 128, M = 65;
  constexpr bool kStaged = sizeof(T) == 4;
  cplx<T>* P = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* tws = P + NY * 64;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int iz = blockIdx.x, f = blockIdx.y;
  const int y = 8 * w + (lane >> 3), s = lane & 7;
  p128_stage_twiddles<T>(tws, tw, twp, tid);
  cplx<T> v[8];
  p128_load_row<T, R>(a, f, ((long long)iz * NY + y) * NX, s, v);
  if constexpr (kStaged) __syncthreads();  // twiddles staged
  cplx<T>* out = reinterpret_cast<cplx<T>*>(a.out[f]) + (long long)iz * M * NY;
  xyf128_plane<T>(P, kStaged ? tws : tw, kStaged ? tws + kP128TX : twp + kP128GX, tws, tid, v, out, [] {});
}

// c2r in y and x of one 128 x 128 z plane from the spectrum columns v
// (column kx = 4w + lane/16, rows ky = q + 16t, q = lane%16; column 0 holds
// col0 + i col_MH), times renor, with the XI epilogue of field f, to the
// rows of dst (the plane's first row).  mid() runs after the barrier that
// completes the plane in LDS.
template <class T, class R, class Mid>
__device__ __forceinline__ void xyi128_plane(const XiArgs<T, R>& a, int f, cplx<T>* P, const cplx<T>* tyq,
                                             const cplx<T>* tyr, const cplx<T>* tws, int tid, cplx<T>* v, R* dst0,
                                             Mid mid) {
  constexpr int NX = 128;
  const int w = tid >> 6, lane = tid & 63;
  const int c = lane >> 4, q = lane & 15, kx = 4 * w + c;
  constexpr unsigned E = sizeof(cplx<T>);
  bfly<T, true, 8>(v);
#pragma unroll
  for (int m = 1; m < 8; ++m) v[m] = tmulc(v[m], tyq[16 * m + q]);
  {
    const int m1 = q >> 1, i0 = q & 1;
    {
      // p128_ye(c, q, m) and p128_ye(c, i0 + 2 i1, m1) in the wavefront's area
      const int rot = ((q & 1) << 2) | ((q >> 1) & 3);
      const unsigned bs = (unsigned)(w * 512 + 128 * c + 8 * (q ^ (2 * c)) + (rot ^ ((q >> 3) << 2))) * E;
      const unsigned br = (unsigned)(w * 512 + 128 * c + 8 * (i0 ^ (2 * c)) + (m1 ^ (i0 << 2))) * E;
#pragma unroll
      for (int m = 0; m < 8; ++m) lds_at(P, bs ^ (m * E)) = v[m];
      wave_lds_sync();
#pragma unroll
      for (int i1 = 0; i1 < 8; ++i1) v[i1] = lds_at(P, br ^ (17 * i1 * E));
    }
    bfly<T, true, 8>(v);
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = tmulc(v[r], tyr[2 * r + i0]);
    // v + o (i0 = 0) or o - v (i0 = 1), as one fma with the exact factor +-1
    const T sg = i0 ? T(-1) : T(1);
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const cplx<T> o = lane_swap1(v[r]);
      v[r] = mk(fmad(v[r].x, sg, o.x), fmad(v[r].y, sg, o.y));
    }
    __syncthreads();  // every wavefront's column exchange is read before the plane is written
    {
      // p128_at(m1 + 8 r + 64 i0, kx): odd r flip the swizzle by 8
      const int swl = m1 ^ ((m1 & 1) << 3) ^ (12 * i0);
      const unsigned be = (unsigned)(m1 * 64 + 4096 * i0 + (kx ^ swl)) * E;
      const unsigned bo = (unsigned)(m1 * 64 + 4096 * i0 + (kx ^ swl ^ 8)) * E;
#pragma unroll
      for (int r = 0; r < 8; ++r) lds_at(P, ((r & 1) ? bo : be) + 512 * r * E) = v[r];
    }
  }
  __syncthreads();
  mid();
  // ---- x: thread (row y, lane s), stage-A lanes hold k1 = p128_k1(s)
  const int yl = lane >> 3, s = lane & 7, y = 8 * w + yl;
  const int k1 = p128_k1(s);
  {
    // p128_at(y, k1 + 8 k2)
    const unsigned bp = (unsigned)(y * 64 + (k1 ^ p128_sw(y))) * E;
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) v[k2] = lds_at(P, bp ^ (8 * k2 * E));
  }
  {
    // Hermitian merge X[k] = s + i d, s = Z[k] + conj Z[64-k],
    // d = (Z[k] - conj Z[64-k]) conj(W_128^k); k = 0 merges DC and Nyquist
    cplx<T> pv[8];
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) pv[k2] = lane_swap1(v[k2]);
    cplx<T> xk[8];
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) {
      const cplx<T> xa = v[k2];
      cplx<T> xc = conj(pick3(k1, v[(8 - k2) & 7], v[7 - k2], pv[7 - k2]));
      cplx<T> xd = xa;
      if (k2 == 0 && k1 == 0) {
        xd = mk(xa.x, T(0));
        xc = mk(xa.y, T(0));
      }
      const cplx<T> sm = xd + xc;
      const cplx<T> d = tmulc(xd - xc, tws[k1 + 8 * k2]);
      xk[k2] = mk(sm.x - d.y, sm.y + d.x);
    }
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) v[k2] = xk[k2];
  }
  bfly<T, true, 8>(v);  // v[j] = G[k1][j]
#pragma unroll
  for (int j = 1; j < 8; ++j) v[j] = tmulc(v[j], tws[kP128TX + 8 * j + k1]);
  {
    // p128_xe(yl, j, k1) and p128_xe(yl, s, o) in the wavefront's area
    const unsigned bs = (unsigned)(w * 512 + 72 * yl + k1) * E;
    const unsigned br = (unsigned)(w * 512 + yl * 64 + 8 * (s ^ yl) + s) * E;
#pragma unroll
    for (int j = 0; j < 8; ++j) lds_at(P, bs ^ (9 * j * E)) = v[j];
    wave_lds_sync();
#pragma unroll
    for (int o = 0; o < 8; ++o) v[o] = lds_at(P, br ^ (o * E));
  }
  bfly<T, true, 8>(v);  // v[t] = z[s + 8t] = (x[2n], x[2n+1])
  cplx<R>* dst = reinterpret_cast<cplx<R>*>(dst0 + (long long)y * NX);
  const T renor = (T)a.renor;
  const bool omd = a.epi[f] == XI_ONE_MINUS_DIV;
  const double ec = a.ec[f];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const T vx = v[t].x * renor, vy = v[t].y * renor;
    cplx<R> o;
    o.x = omd ? (R)(1.0 - (double)(R)vx / ec) : (R)vx;
    o.y = omd ? (R)(1.0 - (double)(R)vy / ec) : (R)vy;
    dst[s + 8 * t] = o;
  }
}

// Spectrum columns of thread tid in plane iz of spectrum `in`: kx = 4w +
// lane/16, rows q + 16t; column 0 packed as col0 + i col_MH.
template <class T>
__device__ __forceinline__ void p128_columns(const cplx<T>* in, int tid, cplx<T>* v) {
  constexpr int NY = 128, MH = 64;
  const int lane = tid & 63, kx = 4 * (tid >> 6) + (lane >> 4), q = lane & 15;
#pragma unroll
  for (int t = 0; t < 8; ++t) v[t] = ld_stream(in + kx * NY + q + 16 * t);
}
template <class T>
__device__ __forceinline__ void p128_pack_col0(const cplx<T>* in, int tid, cplx<T>* v) {
  constexpr int NY = 128, MH = 64;
  const int lane = tid & 63, kx = 4 * (tid >> 6) + (lane >> 4), q = lane & 15;
  if (kx == 0) {
    // columns 0 and MH (real inverse transforms) travel packed as col0 + i col_MH
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const cplx<T> u = ld_stream(in + MH * NY + q + 16 * t);
      v[t] = mk(v[t].x - u.y, v[t].y + u.x);
    }
  }
}

// c2r in y and x of one 128 x 128 z plane, times renor, with the XI epilogues.
// twp: the regrouped twiddles (TQ, TX, TR), tw: W^m, m < 128; the y pass
// reads TQ and TR from global memory (it runs before any barrier).
template <class T, class R>
__global__ void __launch_bounds__(1024, kPlane128Waves<T>) k_fft_xyi128(XiArgs<T, R> a, Grid g, const cplx<T>* __restrict__ twp,
                                                     const cplx<T>* __restrict__ twh,
                                                     const cplx<T>* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  constexpr int NX = 128, NY = 128, M = 65;
  cplx<T>* P = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* tws = P + NY * 64;
  const int tid = threadIdx.x;
  const int iz = blockIdx.x, f = blockIdx.y;
  count_boxed_planes(a, gridDim.y, g.nz, (long long)M * NY * sizeof(cplx<T>) + (long long)NY * NX * sizeof(R));
  if (!xi_plane_needed(a, f, iz, g.nz)) return;
  const cplx<T>* in = reinterpret_cast<const cplx<T>*>(a.in[f]) + (long long)iz * M * NY;
  cplx<T> v[8];
  p128_columns<T>(in, tid, v);
  p128_pack_col0<T>(in, tid, v);
  p128_stage_twiddles<T>(tws, tw, twp, tid);
  xyi128_plane<T, R>(a, f, P, twp + kP128GQ, twp + kP128GR, tws, tid, v, a.out[f] + (long long)iz * NY * NX,
                     [] {});
}

// Persistent form of k_fft_xyi128 for a plane that fills the LDS (FP64: 133
// KB, one workgroup per CU, so a workgroup's loads, transforms and stores
// would follow one another).  The items are the planes each job needs
// (z_range), in job order; workgroups walk them, item = blockIdx.x + k
// gridDim.x, and load the next item's spectrum columns once the current plane
// is in LDS, so the loads are in flight during its x pass and stores (28
// instead of 30 us per 4-field launch at 128^3).  Same arithmetic as
// k_fft_xyi128.  grid = (a few per CU), 1024 threads.
template <class T, class R>
__global__ void __launch_bounds__(1024, 4) k_fft_xyi128p(XiArgs<T, R> a, Grid g, int njobs,
                                                         const cplx<T>* __restrict__ tw,
                                                         const cplx<T>* __restrict__ twp) {
  extern __shared__ unsigned char smem_raw[];
  constexpr int NX = 128, NY = 128, M = 65;
  cplx<T>* P = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* tws = P + NY * 64;
  const int tid0 = threadIdx.x;
  p128_stage_twiddles<T>(tws, tw, twp, tid0);
  int lo[kMaxBatch], cnt[kMaxBatch];
  int nitems = 0;
#pragma unroll
  for (int j = 0; j < kMaxBatch; ++j) {
    const ZRange r = j < njobs ? z_range(a.zbox[j], g.nz) : ZRange{0, -1};
    lo[j] = r.lo;
    cnt[j] = max(r.hi - r.lo + 1, 0);
    nitems += cnt[j];
  }
  // item -> (job f, plane iz)
  auto decode = [&](int item, int& f, int& iz) {
    f = 0;
    iz = 0;
    bool found = false;
#pragma unroll
    for (int j = 0; j < kMaxBatch; ++j) {
      if (!found && item < cnt[j]) {
        f = j;
        iz = lo[j] + item;
        found = true;
      } else if (!found) {
        item -= cnt[j];
      }
    }
  };
  auto spec_of = [&](int item) {
    int f, iz;
    decode(item, f, iz);
    return reinterpret_cast<const cplx<T>*>(a.in[f]) + (long long)iz * M * NY;
  };
  count_boxed_planes(a, njobs, g.nz, (long long)M * NY * sizeof(cplx<T>) + (long long)NY * NX * sizeof(R));
  cplx<T> nxt[8];
  int it = blockIdx.x;
  if (it < nitems) p128_columns<T>(spec_of(it), tid0, nxt);
  __syncthreads();  // twiddles staged
  for (; it < nitems; it += gridDim.x) {
    const int tid = opaque_int(tid0);
    int f, iz;
    decode(it, f, iz);
    cplx<T> v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = nxt[t];
    p128_pack_col0<T>(reinterpret_cast<const cplx<T>*>(a.in[f]) + (long long)iz * M * NY, tid, v);
    const int nit = it + gridDim.x;
    xyi128_plane<T, R>(a, f, P, tws + kP128TQ, tws + kP128TR, tws, tid, v, a.out[f] + (long long)iz * NY * NX, [&] {
      if (nit < nitems) p128_columns<T>(spec_of(nit), tid, nxt);
    });
    __syncthreads();  // the exchange areas are read before the next item writes them
  }
}

}  // namespace hetd
