// Device-side k-space kernel construction (hetddft_init_kernels): the same
// kernels as the host generators of kinit.h -- fvlj4 (fforma.f90:44-56),
// wcgk (initcg.f90:37-42), kalfs (DFT4He3d.f90:566-570) and Vq by the spline
// Fourier transform of FT_V_spline.f90 -- evaluated on the device: the
// momentum sums of the transform (one workgroup per momentum, about 10^5
// spline pieces each) and every half-grid point of the four kernels, written
// straight into the transform's internal [kz][kx][ky] order.
#pragma once
#include "common.h"

namespace hetd {

constexpr double kPiD = 3.141592653589793238462643383279502884;

__device__ inline double d_pfreq(int i, int n, double h) {
  const double hp = 1.0 / (n * h);
  const double pmax = 1.0 / (2.0 * h);
  return i <= n / 2 ? hp * i : -pmax + hp * (i - n / 2);
}

// Si(x), as kinit::sine_integral: power series for x <= 4, else the modified
// Lentz continued fraction of E1(ix) with Si = pi/2 + Im(E1(ix)).
__device__ inline double d_sine_integral(double x) {
  const double sg = x < 0 ? -1.0 : 1.0;
  x = fabs(x);
  if (x <= 4.0) {
    double term = x, sum = x;
    const double x2 = x * x;
    for (int k = 1; k < 60; ++k) {
      term *= -x2 / ((2.0 * k) * (2.0 * k + 1.0));
      const double add = term / (2.0 * k + 1.0);
      sum += add;
      if (fabs(add) < 1e-18 * fabs(sum)) break;
    }
    return sg * sum;
  }
  const double tiny = 1e-300;
  cplx<double> b = mk(1.0, x);
  cplx<double> c = mk(1.0 / tiny, 0.0);
  auto inv = [](cplx<double> z) {
    const double n = z.x * z.x + z.y * z.y;
    return mk(z.x / n, -z.y / n);
  };
  cplx<double> d = inv(b);
  cplx<double> h = d;
  for (int i = 2; i < 100000; ++i) {
    const double a = -double(i - 1) * double(i - 1);
    b.x += 2.0;
    d = inv(mk(a * d.x + b.x, a * d.y + b.y));
    const cplx<double> ac = inv(c);
    c = mk(b.x + a * ac.x, b.y + a * ac.y);
    const cplx<double> del = cmul(c, d);
    h = cmul(h, del);
    if (fabs(del.x - 1.0) + fabs(del.y) < 1e-17) break;
  }
  h = cmul(h, mk(cos(x), -sin(x)));
  return sg * (kPiD * 0.5 + h.y);
}

__device__ inline double d_sinint(double a, double x) { return kPiD * 0.5 - d_sine_integral(a * x); }

__device__ inline double d_sinx11(double a, double x) {
  const double y = a * x, y2 = y * y;
  const double db = 40320.0 - 720.0 * y2 + 24.0 * (y2 * y2) - 2.0 * pow(y2, 3) + 1.0 * pow(y2, 4);
  const double dc = 362880.0 - 5040.0 * y2 + 120.0 * (y2 * y2) - 6.0 * pow(y2, 3) + 1.0 * pow(y2, 4);
  const double x9 = pow(x, 9);
  const double res1 = (a * db * cos(y) + dc * sin(y) / x) / (3628800.0 * x9);
  const double res2 = -(pow(a, 10) * d_sinint(a, x)) / 3628800.0;
  return res1 + res2;
}

__device__ inline double d_sinx5(double a, double x) {
  const double y = a * x, y2 = y * y;
  const double db = -2.0 + 1.0 * y2, dc = -6.0 + 1.0 * y2;
  const double res1 = -(a * db * cos(y) + dc * sin(y) / x) / (24.0 * pow(x, 3));
  const double res2 = (pow(a, 4) * d_sinint(a, x)) / 24.0;
  return res1 + res2;
}

// V(p) for the momenta of the grid: workgroup i sums the analytic Fourier
// integrals of the n-1 spline pieces at p = P[i] (FT_V_spline.f90:151-225).
__global__ void __launch_bounds__(256) k_ftv_sums(const double* __restrict__ x, const double* __restrict__ a,
                                                   const double* __restrict__ b, const double* __restrict__ c,
                                                   const double* __restrict__ d, int n, const double* __restrict__ P,
                                                   double rc, double fc, double* __restrict__ vpp) {
  __shared__ double sh[4];
  const int i = blockIdx.x;
  const double p = P[i];
  const double q = 2.0 * kPiD * p;
  const double Qpi = 4.0 * kPiD;
  double Fi = 0.0;
  if (p == 0.0) {
    for (int j = threadIdx.x; j < n - 1; j += blockDim.x) {
      const double u = x[j], w = x[j + 1];
      Fi += a[j] / 3.0 * (pow(w, 3) - pow(u, 3)) + b[j] / 4.0 * (pow(w, 4) - pow(u, 4)) +
            c[j] / 5.0 * (pow(w, 5) - pow(u, 5)) + d[j] / 6.0 * (pow(w, 6) - pow(u, 6));
    }
  } else {
    const double c1 = 1.0 / q, c2 = c1 * c1, c3 = c2 * c1, c4 = c2 * c2, c5 = c4 * c1;
    for (int j = threadIdx.x; j < n - 1; j += blockDim.x) {
      const double rj = x[j], rj1 = x[j + 1];
      double sj, cj, sj1, cj1;
      sincos(q * rj, &sj, &cj);
      sincos(q * rj1, &sj1, &cj1);
      const double A = a[j], B = b[j], C = c[j], D = d[j];
      Fi += ((A + (B + (C + D * rj) * rj) * rj) * rj * c1 * cj - (A + (B + (C + D * rj1) * rj1) * rj1) * rj1 * c1 * cj1 +
             (A + (2.0 * B + (3.0 * C + 4.0 * D * rj1) * rj1) * rj1) * c2 * sj1 -
             (A + (2.0 * B + (3.0 * C + 4.0 * D * rj) * rj) * rj) * c2 * sj +
             (2.0 * B + (6.0 * C + 12.0 * D * rj1) * rj1) * c3 * cj1 - (2.0 * B + (6.0 * C + 12.0 * D * rj) * rj) * c3 * cj +
             (6.0 * C + 24.0 * D * rj) * c4 * sj - (6.0 * C + 24.0 * D * rj1) * c4 * sj1 + 24.0 * D * c5 * (cj - cj1));
    }
  }
  for (int off = 32; off > 0; off >>= 1) Fi += __shfl_down(Fi, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = Fi;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += sh[k];
    if (p == 0.0) {
      vpp[i] = (rc * rc * rc * fc / 3.0 + s) * Qpi;
    } else {
      const double c1 = 1.0 / q;
      const double vp = fc * c1 * (sin(q * rc) * c1 - rc * cos(q * rc));
      vpp[i] = (vp + s) * c1 * Qpi;
    }
  }
}

struct KinitArgs {
  int nx, ny, nz, mx;
  double hx, hy, hz;
  double h4, l;          // hard-core radius of wcgk/fvlj4, alpha_s kernel width
  double eps, sigma;     // Lennard-Jones
  const double* P;       // momentum grid of the Vq spline, np points
  const double* A;
  const double* B;
  const double* C;
  const double* D;
  int np;
};

// Interval of pc in the sorted grid P[0..n) by bisection (mates.f:4-34),
// 0-based.
__device__ inline int d_findi(const double* xs, int n, double xc) {
  if (xc > xs[n - 1]) return n - 2;
  if (xc < xs[0]) return 0;
  int i = 1, i1 = n;
  for (;;) {
    const int it = (i + i1) / 2;
    if (xc > xs[it - 1]) i = it;
    if (xc <= xs[it - 1]) i1 = it;
    if (i1 - i <= 1) break;
  }
  return i - 1;
}

// The four kernels at every half-grid point.  out64[k] (FP64, may be null)
// and outT[k] (T, may be null) are in internal order [kz][kx][ky];
// k = 0 fvlj4, 1 wcgk, 2 kalfs, 3 vq.
template <class T>
__global__ void __launch_bounds__(256) k_make_kernels(KinitArgs a, double* __restrict__ o64_0, double* __restrict__ o64_1,
                                                      double* __restrict__ o64_2, double* __restrict__ o64_3,
                                                      T* __restrict__ oT_0, T* __restrict__ oT_1, T* __restrict__ oT_2,
                                                      T* __restrict__ oT_3) {
  const long long h = (long long)a.mx * a.ny * a.nz;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= h) return;
  // internal index e = (iz * mx + ix) * ny + iy
  const int iy = (int)(e % a.ny);
  const long long r = e / a.ny;
  const int ix = (int)(r % a.mx), iz = (int)(r / a.mx);
  const double px = d_pfreq(ix, a.nx, a.hx), py = d_pfreq(iy, a.ny, a.hy), pz = d_pfreq(iz, a.nz, a.hz);
  const double pp = sqrt((px * px + py * py) + pz * pz);
  const double pi = kPiD, twopi = 2.0 * kPiD;
  double wcgk = 1.0;
  if (pp > 0.0) {
    const double arg = twopi * pp * a.h4;
    const double aux = sin(arg) - arg * cos(arg);
    wcgk = fabs(aux) < 1e-10 ? 1.0 : 3.0 * aux / (arg * arg * arg);
  }
  const double t = pi * a.l * pp;
  const double kalfs = exp(-(t * t));
  double fvlj4;
  {
    const double sigma6 = pow(a.sigma, 6), aux1 = 8.0 * a.eps;
    if (fabs(pp) < 1e-10) {
      const double pi4 = 4.0 * pi, eps4 = 4.0 * a.eps, s12 = sigma6 * sigma6;
      fvlj4 = pi4 * eps4 * s12 / (9.0 * pow(a.h4, 9)) + (-pi4 * eps4 * sigma6 / (3.0 * pow(a.h4, 3)));
    } else {
      const double a2 = twopi * pp;
      fvlj4 = (aux1 * sigma6 * (sigma6 * d_sinx11(a2, a.h4) - d_sinx5(a2, a.h4))) / pp;
    }
  }
  const int k = d_findi(a.P, a.np, pp);
  const double vq = (double)(float)(a.A[k] + pp * (a.B[k] + pp * (a.C[k] + pp * a.D[k])));
  if (o64_0) o64_0[e] = fvlj4;
  if (o64_1) o64_1[e] = wcgk;
  if (o64_2) o64_2[e] = kalfs;
  if (o64_3) o64_3[e] = vq;
  if (oT_0) oT_0[e] = (T)fvlj4;
  if (oT_1) oT_1[e] = (T)wcgk;
  if (oT_2) oT_2[e] = (T)kalfs;
  if (oT_3) oT_3[e] = (T)vq;
}

}  // namespace hetd
