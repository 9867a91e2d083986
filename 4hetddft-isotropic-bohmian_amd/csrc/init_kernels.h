// Initial state of a droplet built on the device: He density, psi and the
// test-particle cloud.
//
// The reference's fresh start (mode 0, DFT4He3d.f90:585-605) forms
// psi = sqrt(den) exp(i (x lamdax + y lamday + z lamdaz)) from a density read
// from file, draws the particles with sorteo and gives every particle the
// velocity v0 * Ktops.  Here the density is a closed form, so an instance
// needs no host-side field:
//   rho(r) = rho0 / (1 + exp((r - R(theta, phi))/a)),  r = |x - centre|,
//   R = r0 N^(1/3) + ripple * s(theta, phi) / max|s|,
//   s = sum_{l=1..4} sum_{m=-l..l} c_lm P_l^|m|(cos theta) T_m(phi),
//       T_m = cos(m phi) (m >= 0), sin(|m| phi) (m < 0),
//       P_l^m with the Condon-Shortley phase;
//   times 1/(1 + exp(-(r - r_b)/w)) for a cavity of radius r_b;
//   rescaled so that sum(rho) dxyz = N;
//   psi = sqrt(rho) (cos th, sin th),  th = k.x + sigma_phi * g(point);
//   particle p at centre + sigma * (g0, g1, g2)(p), velocity v0 * Ktops.
// g are standard normals from Philox4x64-10 (the generator of sorteo) by
// Box-Muller: g = sqrt(-2 ln u) cos(2 pi v) (and sin for the second of a
// pair), u = (w0 >> 11) + 1 over 2^53 in (0, 1], v = (w1 >> 11) / 2^53; the
// phase of point i uses key (seed, 1) and counter (i, 0, 0, 0), particle p
// key (seed, 2) and counter (p, 0, 0, 0), words (0, 1) for g0 and g1 and
// (2, 3) for g2.  Sums and the maximum go through fixed-order block
// reductions, so the state does not depend on the launch.
#pragma once
#include "field_kernels.h"

namespace hetd {

struct DropArgs {
  Grid g;
  double x0, y0, z0, hx, hy, hz;  // x(ix) = x0 + hx ix, 0-based
  double cx, cy, cz;              // droplet centre
  double rho0, R, a;              // Woods-Saxon
  double rb, w;                   // cavity radius and width (rb = 0: none)
  double ripple;                  // amplitude (0: none)
  double clm[24];                 // l = 1..4, m = -l..l in that order
  double kx, ky, kz;              // plane-wave boost (A^-1)
  double sphi;                    // phase-noise sigma (rad)
  double n_he, dxyz;
  unsigned long long seed;
};

constexpr int kDropBlocks = 1024;

__device__ inline double drop_u(unsigned long long w) { return (double)((w >> 11) + 1ull) * (1.0 / 9007199254740992.0); }

// The angular ripple s(theta, phi) at offset (dx, dy, dz) from the centre;
// cos theta = dz/r (0 at r = 0), phi = atan2(dy, dx).
__device__ inline double drop_surf(const DropArgs& a, double dx, double dy, double dz, double r) {
  const double x = r > 0.0 ? dz / r : 0.0;
  const double phi = atan2(dy, dx);
  const double s2 = 1.0 - x * x;
  const double s = sqrt(s2 > 0.0 ? s2 : 0.0);
  const double x2 = x * x;
  // P_l^m(x), m = 0..l, Condon-Shortley phase (closed forms for l <= 4)
  double P[5][5];
  P[1][0] = x;
  P[1][1] = -s;
  P[2][0] = 0.5 * (3.0 * x2 - 1.0);
  P[2][1] = -3.0 * x * s;
  P[2][2] = 3.0 * s * s;
  P[3][0] = 0.5 * (5.0 * x2 - 3.0) * x;
  P[3][1] = -1.5 * (5.0 * x2 - 1.0) * s;
  P[3][2] = 15.0 * x * s * s;
  P[3][3] = -15.0 * s * s * s;
  P[4][0] = 0.125 * ((35.0 * x2 - 30.0) * x2 + 3.0);
  P[4][1] = -2.5 * (7.0 * x2 - 3.0) * x * s;
  P[4][2] = 7.5 * (7.0 * x2 - 1.0) * s * s;
  P[4][3] = -105.0 * x * s * s * s;
  P[4][4] = 105.0 * s * s * s * s;
  double acc = 0.0;
  int k = 0;
#pragma unroll
  for (int l = 1; l <= 4; ++l) {
#pragma unroll
    for (int m = -l; m <= l; ++m) {
      const int am = m < 0 ? -m : m;
      const double t = m >= 0 ? cos(m * phi) : sin(am * phi);
      acc += a.clm[k++] * (P[l][am] * t);
    }
  }
  return acc;
}

struct DropPoint {
  double x, y, z, dx, dy, dz, r;
};
__device__ inline DropPoint drop_point(const DropArgs& a, long long i) {
  const Grid& g = a.g;
  const int ix = (int)(i % g.nx);
  const long long q = i / g.nx;
  const int iy = (int)(q % g.ny), iz = (int)(q / g.ny);
  DropPoint p;
  p.x = a.x0 + a.hx * ix;
  p.y = a.y0 + a.hy * iy;
  p.z = a.z0 + a.hz * iz;
  p.dx = p.x - a.cx;
  p.dy = p.y - a.cy;
  p.dz = p.z - a.cz;
  p.r = sqrt((p.dx * p.dx + p.dy * p.dy) + p.dz * p.dz);
  return p;
}

// Unnormalised density at point i; smax = max|s| over the grid.
__device__ inline double drop_rho(const DropArgs& a, const DropPoint& p, double smax) {
  double R = a.R;
  if (a.ripple > 0.0) {
    const double sm = smax > 1e-300 ? smax : 1e-300;
    R = R + a.ripple * (drop_surf(a, p.dx, p.dy, p.dz, p.r) / sm);
  }
  double rho = a.rho0 / (1.0 + exp((p.r - R) / a.a));
  if (a.rb > 0.0) rho = rho / (1.0 + exp(-(p.r - a.rb) / a.w));
  return rho;
}

// part[b] = max over block b's points of |s|; grid kDropBlocks x kBlock.
__global__ void __launch_bounds__(kBlock) k_drop_surfmax(DropArgs a, double* __restrict__ part) {
  __shared__ double sh[kBlock / 64];
  double m = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < a.g.n; i += (long long)gridDim.x * blockDim.x) {
    const DropPoint p = drop_point(a, i);
    m = fmax(m, fabs(drop_surf(a, p.dx, p.dy, p.dz, p.r)));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) r = fmax(r, sh[k]);
    part[blockIdx.x] = r;
  }
}

// out = max of part[0..n); one block.
__global__ void __launch_bounds__(kBlock) k_drop_maxfin(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double sh[kBlock / 64];
  double m = 0.0;
  for (int b = threadIdx.x; b < n; b += blockDim.x) m = fmax(m, part[b]);
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) r = fmax(r, sh[k]);
    *out = r;
  }
}

// part[b] = block b's sum of rho (grid-stride order, then the block tree).
__global__ void __launch_bounds__(kBlock) k_drop_sum(DropArgs a, const double* __restrict__ smax,
                                                     double* __restrict__ part) {
  __shared__ double sh[kBlock / 64];
  const double sm = a.ripple > 0.0 ? *smax : 0.0;
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < a.g.n; i += (long long)gridDim.x * blockDim.x)
    s += drop_rho(a, drop_point(a, i), sm);
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// psi = sqrt(rho * N/(S dxyz)) (cos th, sin th), S = *sum.
__global__ void __launch_bounds__(kBlock) k_drop_psi(DropArgs a, const double* __restrict__ smax,
                                                     const double* __restrict__ sum, c64* __restrict__ psi) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.g.n) return;
  const double sm = a.ripple > 0.0 ? *smax : 0.0;
  const double scale = a.n_he / (*sum * a.dxyz);
  const DropPoint p = drop_point(a, i);
  const double amp = sqrt(drop_rho(a, p, sm) * scale);
  double th = (p.x * a.kx + p.y * a.ky) + p.z * a.kz;
  if (a.sphi > 0.0) {
    U64x4 c;
    c.v[0] = (unsigned long long)i;
    c.v[1] = c.v[2] = c.v[3] = 0ull;
    const U64x4 w = philox4x64(c, a.seed, 1ull);
    const double gn = sqrt(-2.0 * log(drop_u(w.v[0]))) * cos(6.283185307179586 * u53(w.v[1]));
    th = th + a.sphi * gn;
  }
  psi[i] = mk(amp * cos(th), amp * sin(th));
}

// Particle p: r = centre + sigma (g0, g1, g2), v = vel; rimp and vimp are
// (N_par, 3), component-major.
__global__ void __launch_bounds__(kBlock) k_drop_cloud(int np, double cx, double cy, double cz, double sigma,
                                                       double vx, double vy, double vz, unsigned long long seed,
                                                       double* __restrict__ rimp, double* __restrict__ vimp) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  U64x4 c;
  c.v[0] = (unsigned long long)p;
  c.v[1] = c.v[2] = c.v[3] = 0ull;
  const U64x4 w = philox4x64(c, seed, 2ull);
  const double r01 = sqrt(-2.0 * log(drop_u(w.v[0])));
  const double t01 = 6.283185307179586 * u53(w.v[1]);
  const double r23 = sqrt(-2.0 * log(drop_u(w.v[2])));
  const double t23 = 6.283185307179586 * u53(w.v[3]);
  rimp[p] = cx + sigma * (r01 * cos(t01));
  rimp[np + p] = cy + sigma * (r01 * sin(t01));
  rimp[2 * np + p] = cz + sigma * (r23 * cos(t23));
  vimp[p] = vx;
  vimp[np + p] = vy;
  vimp[2 * np + p] = vz;
}

// vimp = vel for every particle (positions untouched).
__global__ void __launch_bounds__(kBlock) k_drop_velocity(int np, double vx, double vy, double vz,
                                                          double* __restrict__ vimp) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  vimp[p] = vx;
  vimp[np + p] = vy;
  vimp[2 * np + p] = vz;
}

}  // namespace hetd
