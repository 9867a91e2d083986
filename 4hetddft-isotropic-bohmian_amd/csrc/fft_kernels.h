// Batched 3D real-to-complex / complex-to-real transform passes.
//
// The reference runs 22 FFTW r2c/c2r transforms per potential evaluation
// (poten.f90, term_alfa.f90, potenimp.f90; plans in fft.f90:43-65).  Here a
// 3D convolution is five batched passes over HBM:
//   XF  real rows -> half-length complex FFT + Hermitian split (r2c along x),
//       with the pointwise producer of the transformed field fused on load;
//   YF  complex FFT along y on tiles of consecutive kx;
//   Z   forward FFT along z, multiplication by one or more real k-space
//       kernels, inverse FFT along z, one output spectrum per kernel;
//   YI  inverse complex FFT along y;
//   XI  Hermitian merge + half-length inverse FFT (c2r along x), times
//       renor = 1/(nx ny nz) (fft.f90:125-207).
// Every launch carries several fields so one launch fills the chip.
#pragma once
#include "fft_device.h"

namespace hetd {

constexpr int kMaxBatch = 8;
constexpr int kMaxJobs = 12;

// Pointwise producers fused into the forward x pass (the array the reference
// hands to dfftw_execute is formed on the fly from its inputs).
enum XfOp : int {
  XF_COPY = 0,      // f = a
  XF_CORREL = 1,    // f = den*dencg*(cp4+cpp4*dencg)           (poten.f90:62-64)
  XF_FALFS_D = 2,   // f = (1 - denalf/den0s)*d                 (term_alfa.f90:28-36)
  XF_GRAD_DOT = 3,  // f = dx*ix + dy*iy + dz*iz                (term_alfa.f90:65-69)
};

template <class R>
struct XfArgs {
  const R* in[kMaxBatch][6];
  int op[kMaxBatch];
  void* out[kMaxBatch];     // cplx<T>* half spectra
  double c[kMaxBatch][2];   // per-field op constants: (cp4, cpp4) or (den0s, -)
};

template <class R>
__device__ inline double xf_value(const XfArgs<R>& a, int f, long long i) {
  switch (a.op[f]) {
    case XF_CORREL: {
      double den = (double)a.in[f][0][i], dcg = (double)a.in[f][1][i];
      return den * dcg * (a.c[f][0] + a.c[f][1] * dcg);
    }
    case XF_FALFS_D: {
      double fal = 1.0 - (double)a.in[f][0][i] / a.c[f][0];
      return fal * (double)a.in[f][1][i];
    }
    case XF_GRAD_DOT: {
      double s = (double)a.in[f][0][i] * (double)a.in[f][3][i] +
                 (double)a.in[f][1][i] * (double)a.in[f][4][i];
      return s + (double)a.in[f][2][i] * (double)a.in[f][5][i];
    }
    default:
      return (double)a.in[f][0][i];
  }
}

// r2c along x.  grid = (ceil(ny*nz/rows), nfields); dynamic LDS holds two
// buffers of rows x (m+1) complex, m = nx/2.
template <class T, class R>
__global__ void __launch_bounds__(256) k_fft_xf(XfArgs<R> a, Grid g, FftPlan pl, int rows,
                                                const cplx<T>* __restrict__ tw,
                                                const cplx<T>* __restrict__ twh) {
  extern __shared__ unsigned char smem_raw[];
  const int m = g.nx / 2;
  const int ls = m + 1;
  cplx<T>* b0 = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* b1 = b0 + rows * ls;
  const int f = blockIdx.y;
  const long long nrows = (long long)g.ny * g.nz;
  const long long row0 = (long long)blockIdx.x * rows;
  const int nrow = (int)((nrows - row0) < rows ? (nrows - row0) : rows);
  // load: x[2j] -> re, x[2j+1] -> im of complex j
  T* b0r = reinterpret_cast<T*>(b0);
  for (int e = threadIdx.x; e < rows * g.nx; e += blockDim.x) {
    const int r = e / g.nx, c = e - r * g.nx;
    T v = T(0);
    if (r < nrow) v = (T)xf_value(a, f, (row0 + r) * g.nx + c);
    b0r[2 * (r * ls) + c] = v;
  }
  __syncthreads();
  const int res = fft_lines<T, false>(b0, b1, rows, ls, pl, tw);
  const cplx<T>* Z = res ? b1 : b0;
  cplx<T>* out = reinterpret_cast<cplx<T>*>(a.out[f]);
  const int M = m + 1;
  for (int e = threadIdx.x; e < nrow * M; e += blockDim.x) {
    const int r = e / M, k = e - r * M;
    const cplx<T>* z = Z + r * ls;
    const cplx<T> zk = z[k == m ? 0 : k];
    const cplx<T> zc = conj(z[k == 0 ? 0 : m - k]);
    // E = (Z[k] + conj Z[m-k])/2, O = (Z[k] - conj Z[m-k])/(2i), X = E + W^k O
    const cplx<T> E = scale(zk + zc, T(0.5));
    const cplx<T> D = zk - zc;
    const cplx<T> O = mk(T(0.5) * D.y, T(-0.5) * D.x);
    const cplx<T> X = E + cmul(twh[k], O);
    out[(row0 + r) * M + k] = X;
  }
}

// c2r along x: half spectrum rows -> real rows, times renor.
template <class T, class R>
struct XiArgs {
  const void* in[kMaxBatch];   // cplx<T>* half spectra
  R* out[kMaxBatch];
  double renor;
};

template <class T, class R>
__global__ void __launch_bounds__(256) k_fft_xi(XiArgs<T, R> a, Grid g, FftPlan pl, int rows,
                                                const cplx<T>* __restrict__ tw,
                                                const cplx<T>* __restrict__ twh) {
  extern __shared__ unsigned char smem_raw[];
  const int m = g.nx / 2;
  const int ls = m + 1;
  const int M = m + 1;
  cplx<T>* b0 = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* b1 = b0 + rows * ls;
  const int f = blockIdx.y;
  const long long nrows = (long long)g.ny * g.nz;
  const long long row0 = (long long)blockIdx.x * rows;
  const int nrow = (int)((nrows - row0) < rows ? (nrows - row0) : rows);
  const cplx<T>* in = reinterpret_cast<const cplx<T>*>(a.in[f]);
  for (int e = threadIdx.x; e < rows * m; e += blockDim.x) {
    const int r = e / m, k = e - r * m;
    cplx<T> y = mk(T(0), T(0));
    if (r < nrow) {
      const cplx<T>* x = in + (row0 + r) * M;
      const cplx<T> xk = x[k];
      const cplx<T> xc = conj(x[m - k]);
      // Y[k] = (X[k] + conj X[m-k]) + i (X[k] - conj X[m-k]) W^{-k}
      const cplx<T> s = xk + xc;
      const cplx<T> d = cmulc(xk - xc, twh[k]);
      y = mk(s.x - d.y, s.y + d.x);
    }
    b0[r * ls + k] = y;
  }
  __syncthreads();
  const int res = fft_lines<T, true>(b0, b1, rows, ls, pl, tw);
  const T* zr = reinterpret_cast<const T*>(res ? b1 : b0);
  R* out = a.out[f];
  const T renor = (T)a.renor;
  for (int e = threadIdx.x; e < nrow * g.nx; e += blockDim.x) {
    const int r = e / g.nx, c = e - r * g.nx;
    out[(row0 + r) * g.nx + c] = (R)(zr[2 * (r * ls) + c] * renor);
  }
}

// Complex FFT along y (stride M) for tiles of tx consecutive kx at fixed kz.
// grid = (ntiles, nz, nfields).  In place.
template <class T>
struct YArgs {
  void* buf[kMaxBatch];
};

template <class T, bool INV>
__global__ void __launch_bounds__(256) k_fft_y(YArgs<T> a, Grid g, FftPlan pl, int tx,
                                               const cplx<T>* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  const int n = g.ny;
  const int ls = n + 1;
  cplx<T>* b0 = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* b1 = b0 + tx * ls;
  const int M = g.mx;
  const int kx0 = blockIdx.x * tx;
  const int ntx = (M - kx0) < tx ? (M - kx0) : tx;
  cplx<T>* base = reinterpret_cast<cplx<T>*>(a.buf[blockIdx.z]) + (long long)blockIdx.y * n * M + kx0;
  for (int e = threadIdx.x; e < tx * n; e += blockDim.x) {
    const int j = e / tx, l = e - j * tx;
    b0[l * ls + j] = (l < ntx) ? base[(long long)j * M + l] : mk(T(0), T(0));
  }
  __syncthreads();
  const int res = fft_lines<T, INV>(b0, b1, tx, ls, pl, tw);
  const cplx<T>* Y = res ? b1 : b0;
  for (int e = threadIdx.x; e < tx * n; e += blockDim.x) {
    const int j = e / tx, l = e - j * tx;
    if (l < ntx) base[(long long)j * M + l] = Y[l * ls + j];
  }
}

// z pass of a convolution: forward FFT along z of input i, then for each
// job of that input: multiply by sign*kernel(kx,ky,kz), inverse FFT along z,
// store to the job's output spectrum.  grid = (ntiles, ny, ninputs).
template <class T>
struct ZArgs {
  const void* in[kMaxBatch];
  int job0[kMaxBatch + 1];
  const T* kern[kMaxJobs];
  T ksign[kMaxJobs];
  void* out[kMaxJobs];
};

template <class T>
__global__ void __launch_bounds__(256) k_fft_zconv(ZArgs<T> a, Grid g, FftPlan pl, int tx,
                                                   const cplx<T>* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  const int n = g.nz;
  const int ls = n + 1;
  cplx<T>* bA = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* bB = bA + tx * ls;
  cplx<T>* bC = bB + tx * ls;
  const int M = g.mx;
  const long long pstride = (long long)M * g.ny;  // stride between kz planes
  const int kx0 = blockIdx.x * tx;
  const int ntx = (M - kx0) < tx ? (M - kx0) : tx;
  const int ky = blockIdx.y;
  const int ii = blockIdx.z;
  const long long off = (long long)ky * M + kx0;
  const cplx<T>* src = reinterpret_cast<const cplx<T>*>(a.in[ii]) + off;
  for (int e = threadIdx.x; e < tx * n; e += blockDim.x) {
    const int j = e / tx, l = e - j * tx;
    bA[l * ls + j] = (l < ntx) ? src[(long long)j * pstride + l] : mk(T(0), T(0));
  }
  __syncthreads();
  int res = fft_lines<T, false>(bA, bB, tx, ls, pl, tw);
  cplx<T>* F = res ? bB : bA;     // forward spectrum, kept
  cplx<T>* W0 = res ? bA : bB;    // free buffers for the inverse
  cplx<T>* W1 = bC;
  for (int jb = a.job0[ii]; jb < a.job0[ii + 1]; ++jb) {
    const T* kern = a.kern[jb] + off;
    const T sg = a.ksign[jb];
    for (int e = threadIdx.x; e < tx * n; e += blockDim.x) {
      const int j = e / tx, l = e - j * tx;
      cplx<T> v = F[l * ls + j];
      if (l < ntx) {
        const T kv = kern[(long long)j * pstride + l];
        v = mk(v.x * kv, v.y * kv);
        if (sg < T(0)) v = mk(-v.x, -v.y);
      }
      W0[l * ls + j] = v;
    }
    __syncthreads();
    const int r2 = fft_lines<T, true>(W0, W1, tx, ls, pl, tw);
    const cplx<T>* Y = r2 ? W1 : W0;
    cplx<T>* dst = reinterpret_cast<cplx<T>*>(a.out[jb]) + off;
    for (int e = threadIdx.x; e < tx * n; e += blockDim.x) {
      const int j = e / tx, l = e - j * tx;
      if (l < ntx) dst[(long long)j * pstride + l] = Y[l * ls + j];
    }
    __syncthreads();
  }
}

}  // namespace hetd
