// Mixed-radix (2,3,4,5,8,16) Stockham FFT stages on LDS-resident lines.
//
// The 3D r2c/c2r transforms of the reference (FFTW plans, fft.f90:43-65) are
// built from these 1D stages: each workgroup stages a tile of lines in LDS,
// runs every radix stage between two LDS buffers, and writes the result
// back with coalesced HBM accesses.  Forward uses exp(-2 pi i jk/n) (FFTW
// FORWARD); inverse is unnormalised, like FFTW BACKWARD.
#pragma once
#include "common.h"

namespace hetd {

template <class T>
struct RadixConst {
  static constexpr T s3 = T(0.86602540378443864676372317075294);   // sin(pi/3)
  static constexpr T c51 = T(0.30901699437494742410229341718282);  // cos(2pi/5)
  static constexpr T c52 = T(-0.80901699437494742410229341718282); // cos(4pi/5)
  static constexpr T s51 = T(0.95105651629515357211643933337938);  // sin(2pi/5)
  static constexpr T s52 = T(0.58778525229247312916870595463907);  // sin(4pi/5)
  static constexpr T r2 = T(0.70710678118654752440084436210485);   // 1/sqrt(2)
};

// Twiddle products of the transforms with fused multiply-adds.  The FFT
// layer has no reference operation order to keep (the reference calls
// FFTW/MKL, fft.f90:43-65, whose results are pinned by the DFT definition),
// and one rounding per component is more accurate than two.
__device__ inline float fmad(float a, float b, float c) { return fmaf(a, b, c); }
__device__ inline double fmad(double a, double b, double c) { return fma(a, b, c); }
template <class T>
__device__ inline cplx<T> tmul(cplx<T> a, cplx<T> b) {
  return mk(fmad(a.x, b.x, -(a.y * b.y)), fmad(a.x, b.y, a.y * b.x));
}
template <class T>
__device__ inline cplx<T> tmulc(cplx<T> a, cplx<T> b) {  // a * conj(b)
  return mk(fmad(a.x, b.x, a.y * b.y), fmad(a.y, b.x, -(a.x * b.y)));
}

// multiply by -i (forward) or +i (inverse)
template <class T, bool INV>
__device__ inline cplx<T> rot90(cplx<T> a) {
  return INV ? mk(-a.y, a.x) : mk(a.y, -a.x);
}

template <class T, bool INV>
__device__ inline void bfly2(cplx<T>* v) {
  cplx<T> a = v[0], b = v[1];
  v[0] = a + b;
  v[1] = a - b;
}

template <class T, bool INV>
__device__ inline void bfly4(cplx<T>* v) {
  cplx<T> a = v[0] + v[2], b = v[0] - v[2];
  cplx<T> c = v[1] + v[3], d = rot90<T, INV>(v[1] - v[3]);
  v[0] = a + c;
  v[2] = a - c;
  v[1] = b + d;
  v[3] = b - d;
}

template <class T, bool INV>
__device__ inline void bfly3(cplx<T>* v) {
  const T s3 = RadixConst<T>::s3;
  cplx<T> t1 = v[1] + v[2];
  cplx<T> t2 = v[0] - scale(t1, T(0.5));
  cplx<T> t3 = rot90<T, INV>(scale(v[1] - v[2], s3));
  v[0] = v[0] + t1;
  v[1] = t2 + t3;
  v[2] = t2 - t3;
}

template <class T, bool INV>
__device__ inline void bfly5(cplx<T>* v) {
  const T c1 = RadixConst<T>::c51, c2 = RadixConst<T>::c52;
  const T s1 = RadixConst<T>::s51, s2 = RadixConst<T>::s52;
  cplx<T> b1 = v[1] + v[4], b2 = v[2] + v[3];
  cplx<T> d1 = v[1] - v[4], d2 = v[2] - v[3];
  cplx<T> a0 = v[0];
  cplx<T> e1 = a0 + scale(b1, c1) + scale(b2, c2);
  cplx<T> e2 = a0 + scale(b2, c1) + scale(b1, c2);
  cplx<T> f1 = rot90<T, INV>(scale(d1, s1) + scale(d2, s2));
  cplx<T> f2 = rot90<T, INV>(scale(d1, s2) - scale(d2, s1));
  v[0] = a0 + b1 + b2;
  v[1] = e1 + f1;
  v[4] = e1 - f1;
  v[2] = e2 + f2;
  v[3] = e2 - f2;
}

template <class T, bool INV>
__device__ inline void bfly8(cplx<T>* v) {
  const T r = RadixConst<T>::r2;
  cplx<T> e[4] = {v[0], v[2], v[4], v[6]};
  cplx<T> o[4] = {v[1], v[3], v[5], v[7]};
  bfly4<T, INV>(e);
  bfly4<T, INV>(o);
  // o[k] *= W8^k (W8 = exp(-+2 pi i/8))
  cplx<T> o1 = INV ? mk(r * (o[1].x - o[1].y), r * (o[1].x + o[1].y))
                   : mk(r * (o[1].x + o[1].y), r * (o[1].y - o[1].x));
  cplx<T> o2 = rot90<T, INV>(o[2]);
  cplx<T> o3 = INV ? mk(-r * (o[3].x + o[3].y), r * (o[3].x - o[3].y))
                   : mk(r * (o[3].y - o[3].x), -r * (o[3].x + o[3].y));
  v[0] = e[0] + o[0];
  v[4] = e[0] - o[0];
  v[1] = e[1] + o1;
  v[5] = e[1] - o1;
  v[2] = e[2] + o2;
  v[6] = e[2] - o2;
  v[3] = e[3] + o3;
  v[7] = e[3] - o3;
}

// Radix 16 as 4 x 4: inner length-4 DFTs over n1 (stride 4), twiddles
// W16^(n2 k1), outer length-4 DFTs over n2; X[k1 + 4 k2] in natural order.
template <class T, bool INV>
__device__ inline void bfly16(cplx<T>* v) {
  const T c1 = T(0.92387953251128675612818318939679);  // cos(pi/8)
  const T s1 = T(0.38268343236508977172845998403040);  // sin(pi/8)
  const T r = RadixConst<T>::r2;
  const T sg = INV ? T(1) : T(-1);
  cplx<T> y[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) y[n2][n1] = v[4 * n1 + n2];
    bfly4<T, INV>(y[n2]);
  }
  const cplx<T> w1 = mk(c1, sg * s1);    // W16^1
  const cplx<T> w2 = mk(r, sg * r);      // W16^2
  const cplx<T> w3 = mk(s1, sg * c1);    // W16^3
  const cplx<T> w6 = mk(-r, sg * r);     // W16^6
  const cplx<T> w9 = mk(-c1, -sg * s1);  // W16^9
  y[1][1] = tmul(y[1][1], w1);
  y[1][2] = tmul(y[1][2], w2);
  y[1][3] = tmul(y[1][3], w3);
  y[2][1] = tmul(y[2][1], w2);
  y[2][2] = rot90<T, INV>(y[2][2]);
  y[2][3] = tmul(y[2][3], w6);
  y[3][1] = tmul(y[3][1], w3);
  y[3][2] = tmul(y[3][2], w6);
  y[3][3] = tmul(y[3][3], w9);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    cplx<T> z[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
    bfly4<T, INV>(z);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = z[k2];
  }
}

// cos and sin of 2 pi m / r for the composite radices r = 6, 10, 12.
__host__ __device__ constexpr double root_cos(int r, int m) {
  constexpr double c30 = 0.86602540378443864676372317075294;  // cos(pi/6)
  constexpr double c36 = 0.80901699437494742410229341718282;  // cos(pi/5)
  constexpr double c72 = 0.30901699437494742410229341718282;  // cos(2pi/5)
  constexpr double t12[12] = {1, c30, 0.5, 0, -0.5, -c30, -1, -c30, -0.5, 0, 0.5, c30};
  constexpr double t10[10] = {1, c36, c72, -c72, -c36, -1, -c36, -c72, c72, c36};
  return r == 6 ? t12[(2 * m) % 12] : r == 12 ? t12[m % 12] : t10[m % 10];
}
__host__ __device__ constexpr double root_sin(int r, int m) {
  constexpr double c30 = 0.86602540378443864676372317075294;  // sin(pi/3)
  constexpr double s36 = 0.58778525229247312916870595463907;  // sin(pi/5)
  constexpr double s72 = 0.95105651629515357211643933337938;  // sin(2pi/5)
  constexpr double t12[12] = {0, 0.5, c30, 1, c30, 0.5, 0, -0.5, -c30, -1, -c30, -0.5};
  constexpr double t10[10] = {0, s36, s72, s72, s36, 0, -s36, -s72, -s72, -s36};
  return r == 6 ? t12[(2 * m) % 12] : r == 12 ? t12[m % 12] : t10[m % 10];
}
static_assert(root_cos(10, 3) < 0 && root_sin(12, 4) > 0.86 && root_cos(6, 2) == -0.5, "root tables");

// DFT of length A in registers (A = 2, 3, 4, 5, 8), natural order.
template <class T, bool INV, int A>
__device__ inline void small_dft(cplx<T>* v) {
  static_assert(A == 2 || A == 3 || A == 4 || A == 5 || A == 8, "small DFT length");
  if constexpr (A == 2) bfly2<T, INV>(v);
  if constexpr (A == 3) bfly3<T, INV>(v);
  if constexpr (A == 4) bfly4<T, INV>(v);
  if constexpr (A == 5) bfly5<T, INV>(v);
  if constexpr (A == 8) bfly8<T, INV>(v);
}

// Radix A*B in registers, as bfly16 is 4 x 4: length-A DFTs over n1
// (elements B n1 + n2), twiddles W_{AB}^(n2 k1), length-B DFTs over n2;
// X[k1 + A k2] in natural order.  W = exp(-2 pi i/AB), conjugate for INV.
template <class T, bool INV, int A, int B>
__device__ inline void bfly_ab(cplx<T>* v) {
  constexpr int R = A * B;
  cplx<T> y[B][A];
#pragma unroll
  for (int n2 = 0; n2 < B; ++n2) {
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) y[n2][n1] = v[B * n1 + n2];
    small_dft<T, INV, A>(y[n2]);
  }
#pragma unroll
  for (int n2 = 1; n2 < B; ++n2)
#pragma unroll
    for (int k1 = 1; k1 < A; ++k1) {
      const T wc = T(root_cos(R, n2 * k1)), ws = T(root_sin(R, n2 * k1));
      y[n2][k1] = tmul(y[n2][k1], mk(wc, INV ? ws : -ws));
    }
#pragma unroll
  for (int k1 = 0; k1 < A; ++k1) {
    cplx<T> z[B];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) z[n2] = y[n2][k1];
    small_dft<T, INV, B>(z);
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = z[k2];
  }
}

// One Stockham stage of radix R over L lines of length n (line stride ls in
// LDS).  tw is the length-n table exp(-2 pi i m/n).
template <class T, bool INV, int R>
__device__ inline void stage_r(const cplx<T>* __restrict__ src, cplx<T>* __restrict__ dst, int L,
                               int n, int ls, int Ns, const cplx<T>* __restrict__ tw) {
  const int nb = n / R;
  const int twstep = n / (Ns * R);
  const int total = L * nb;
  for (int b = threadIdx.x; b < total; b += blockDim.x) {
    const int l = b / nb;
    const int j = b - l * nb;
    const cplx<T>* s = src + l * ls;
    cplx<T>* d = dst + l * ls;
    const int k = j % Ns;
    cplx<T> v[R];
#pragma unroll
    for (int t = 0; t < R; ++t) v[t] = s[j + t * nb];
    if (Ns > 1) {
#pragma unroll
      for (int t = 1; t < R; ++t) {
        cplx<T> w = tw[t * k * twstep];
        v[t] = INV ? tmulc(v[t], w) : tmul(v[t], w);
      }
    }
    if constexpr (R == 2) bfly2<T, INV>(v);
    if constexpr (R == 3) bfly3<T, INV>(v);
    if constexpr (R == 4) bfly4<T, INV>(v);
    if constexpr (R == 5) bfly5<T, INV>(v);
    if constexpr (R == 8) bfly8<T, INV>(v);
    const int base = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int t = 0; t < R; ++t) d[base + t * Ns] = v[t];
  }
}

// Full 1D transform of L lines; data starts in buf[0], buf[1] is scratch.
// Returns the index (0 or 1) of the buffer holding the result.  Every thread
// of the block must call it; it synchronises between stages.
template <class T, bool INV>
__device__ inline int fft_lines(cplx<T>* buf0, cplx<T>* buf1, int L, int ls, const FftPlan& pl,
                                const cplx<T>* __restrict__ tw) {
  cplx<T>* bufs[2] = {buf0, buf1};
  int cur = 0;
  int Ns = 1;
  for (int s = 0; s < pl.nstages; ++s) {
    const int r = pl.radix[s];
    const cplx<T>* src = bufs[cur];
    cplx<T>* dst = bufs[cur ^ 1];
    switch (r) {
      case 2: stage_r<T, INV, 2>(src, dst, L, pl.n, ls, Ns, tw); break;
      case 3: stage_r<T, INV, 3>(src, dst, L, pl.n, ls, Ns, tw); break;
      case 4: stage_r<T, INV, 4>(src, dst, L, pl.n, ls, Ns, tw); break;
      case 5: stage_r<T, INV, 5>(src, dst, L, pl.n, ls, Ns, tw); break;
      default: stage_r<T, INV, 8>(src, dst, L, pl.n, ls, Ns, tw); break;
    }
    __syncthreads();
    Ns *= r;
    cur ^= 1;
  }
  return cur;
}

}  // namespace hetd
