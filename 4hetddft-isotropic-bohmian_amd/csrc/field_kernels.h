// Grid-point kernels: 13-point stencils, Hamiltonian application fused with
// the RK-Gill / Hamming updates of psi, potential assembly, NGP histogram,
// ironing, trilinear force gather, particle updates and reductions.
//
// Floating-point expressions keep the reference's operand order so that the
// FP64 path reproduces the CPU restatement of the reference to rounding.
#pragma once
#include "common.h"

namespace hetd {

constexpr int kBlock = 256;

struct Idx3 {
  int ix, iy, iz;
};

__device__ inline Idx3 idx3(long long i, const Grid& g) {
  // 32-bit unsigned index arithmetic: grids hold at most 256^3 points
  Idx3 r;
  const unsigned u = (unsigned)i;
  const unsigned pl = (unsigned)g.nx * (unsigned)g.ny;
  const unsigned iz = u / pl;
  const unsigned rem = u - iz * pl;
  const unsigned iy = rem / (unsigned)g.nx;
  r.iz = (int)iz;
  r.iy = (int)iy;
  r.ix = (int)(rem - iy * (unsigned)g.nx);
  return r;
}

// 1D 13-point derivative along one axis with f == 0 outside the box
// (derivnD with icon=13: init_deriv_parallel.f90:428-535, redef.f:121-170).
// The sum runs over offsets -6..6 in order, adding w*0 for out-of-box
// points.  Every load is issued unconditionally from an in-box address (the
// centre for out-of-box offsets) so all 13 can be in flight together.
template <class R, class W = Stencil13>
__device__ inline double d13(const R* __restrict__ f, long long i, int pos, int n, long long stride,
                             const W& w) {
  double v[13];
#pragma unroll
  for (int m = -6; m <= 6; ++m) {
    const bool inb = (unsigned)(pos + m) < (unsigned)n;
    const double t = (double)f[i + (inb ? m * stride : 0)];
    v[m + 6] = inb ? t : 0.0;
  }
  double acc = 0.0;
#pragma unroll
  for (int m = 0; m < 13; ++m) acc = acc + wat(w, m) * v[m];
  return acc;
}

__device__ inline c64 d13c(const c64* __restrict__ f, long long i, int pos, int n, long long stride,
                           const Stencil13& w) {
  c64 v[13];
#pragma unroll
  for (int m = -6; m <= 6; ++m) {
    const bool inb = (unsigned)(pos + m) < (unsigned)n;
    const c64 t = f[i + (inb ? m * stride : 0)];
    v[m + 6] = inb ? t : mk(0.0, 0.0);
  }
  double ax = 0.0, ay = 0.0;
#pragma unroll
  for (int m = 0; m < 13; ++m) {
    ax = ax + w.w[m] * v[m].x;
    ay = ay + w.w[m] * v[m].y;
  }
  return mk(ax, ay);
}

// ---- block reduction (sum of doubles), fixed tree order ------------------
__device__ inline double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int k = 0; k < nw; ++k) s += sh[k];
  }
  return s;  // valid in thread 0
}

// p[0..n) = 0 (a kernel rather than a memset, so that it can sit in a
// captured graph).
__global__ void k_fill_zero(double* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// Sum of nblocks partial sums per slot; one block per slot, fixed order.
__global__ void k_finish_sums(const double* __restrict__ part, int nblocks, int nslots,
                              double* __restrict__ out) {
  __shared__ double sh[16];
  const int s = blockIdx.x;
  double v = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) v += part[(long long)s * nblocks + b];
  v = block_sum(v, sh);
  if (threadIdx.x == 0) out[s] = v;
}

// den = |psi|^2 (DFT4He3d.f90:630 with conjg(psi)*psi)
template <class R>
__global__ void k_density(const c64* __restrict__ psi, R* __restrict__ den, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const c64 v = psi[i];
  den[i] = (R)(v.x * v.x + v.y * v.y);
}

// ---- impurity side ------------------------------------------------------
// NGP histogram (denpart.f90:24-32): ix = int(1.5 + (r - x1)/h); counts are
// exact integers, so the atomic order does not matter.  The histogram also
// records the bounding box of the occupied cells in reg[0..5].
struct BoxArgs {
  Grid g;
  double xmin, ymin, zmin, hx, hy, hz;
  int npar;
};

// Cell of a particle for the trilinear gather (interpola3,
// potenimp.f90:91-106): ix = int((r - x1)/h + 1); valid when 1 <= ix < n.
struct GatherCell {
  int ix, iy, iz;
  double r, s, t;
  bool ok;
};

// from the scaled offsets a = (r - x1)/h + 1 of the three components
__device__ inline GatherCell gather_cell_at(const BoxArgs& b, double ax, double ay, double az) {
  GatherCell q;
  // 1 <= int(a) <= n-1  <=>  1 <= a < n; the comparisons are false for a
  // non-finite position, which is then outside (and never converted)
  q.ok = ax >= 1.0 && ax < (double)b.g.nx && ay >= 1.0 && ay < (double)b.g.ny && az >= 1.0 &&
         az < (double)b.g.nz;
  q.ix = q.ok ? (int)ax : 0;
  q.iy = q.ok ? (int)ay : 0;
  q.iz = q.ok ? (int)az : 0;
  q.r = ax - q.ix;
  q.s = ay - q.iy;
  q.t = az - q.iz;
  return q;
}

__device__ inline GatherCell gather_cell(const BoxArgs& b, const double* __restrict__ rp, int i) {
  const int np = b.npar;
  return gather_cell_at(b, (rp[i] - b.xmin) / b.hx + 1, (rp[np + i] - b.ymin) / b.hy + 1,
                        (rp[2 * np + i] - b.zmin) / b.hz + 1);
}

// reg layout: [0..5] bbox of occupied cells (imin,jmin,kmin,imax,jmax,kmax);
// [6..11] region that may hold non-zero denx/psix/hpsix; [12..17] region
// recomputed by this evaluation; [18..23] bbox of the trilinear-gather cell
// corners (interpola3) of the particles inside the box.  Regions are
// inclusive 0-based boxes.
constexpr int kRegBox = 0, kRegLive = 6, kRegWork = 12, kRegGather = 18, kRegLen = 24;

// Partial boxes of k_hist: per workgroup, the occupied-cell box then the
// gather-corner box, (min x,y,z, max x,y,z) each; empty: min 0x7fffffff, max -1.
constexpr int kBoxPart = 12;

__device__ inline void box_reduce_wave(int (&v)[kBoxPart]) {
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int d = 0; d < kBoxPart; ++d) {
      const int o = __shfl_down(v[d], off, 64);
      v[d] = d % 6 < 3 ? min(v[d], o) : max(v[d], o);
    }
}

// Occupied-cell box v[0..5] and gather-corner box v[6..11] -> reg; then
// work = live U expand(bbox, margin); live := expand(bbox, margin).
// live := expand(box, margin), clamped to the grid (empty box: empty).
__device__ inline void box_expand(Grid g, const int* box, int margin, int (&e)[6]) {
  const int n[3] = {g.nx, g.ny, g.nz};
  const bool empty = box[3] < 0;
  for (int d = 0; d < 3; ++d) {
    e[d] = empty ? 0x7fffffff : max(box[d] - margin, 0);
    e[3 + d] = empty ? -1 : min(box[3 + d] + margin, n[d] - 1);
  }
}

// work = live U e, from the live region before this evaluation.
__device__ inline void work_region(const int* __restrict__ reg, const int (&e)[6], int (&w)[6]) {
  const bool lempty = reg[kRegLive + 3] < 0;
  for (int d = 0; d < 3; ++d) {
    w[d] = lempty ? e[d] : min(reg[kRegLive + d], e[d]);
    w[3 + d] = lempty ? e[3 + d] : max(reg[kRegLive + 3 + d], e[3 + d]);
  }
}

// Combines nb partial boxes of k_hist over one wavefront.
__device__ inline void box_reduce_parts(const int* __restrict__ part, int nb, int (&v)[kBoxPart]) {
#pragma unroll
  for (int d = 0; d < kBoxPart; ++d) v[d] = d % 6 < 3 ? 0x7fffffff : -1;
  for (int k = (int)(threadIdx.x & 63); k < nb; k += 64)
#pragma unroll
    for (int d = 0; d < kBoxPart; ++d) {
      const int o = part[k * kBoxPart + d];
      v[d] = d % 6 < 3 ? min(v[d], o) : max(v[d], o);
    }
  box_reduce_wave(v);
}

// Occupied-cell box v[0..5] and gather-corner box v[6..11] -> reg; then
// work = live U expand(bbox, margin); live := expand(bbox, margin).
__device__ inline void reg_union_apply(Grid g, int* __restrict__ reg, int margin, const int (&v)[kBoxPart]) {
  for (int d = 0; d < 6; ++d) {
    reg[kRegBox + d] = v[d];
    reg[kRegGather + d] = v[6 + d];
  }
  int e[6], w[6];
  box_expand(g, v, margin, e);
  work_region(reg, e, w);
  for (int d = 0; d < 6; ++d) {
    reg[kRegWork + d] = w[d];
    reg[kRegLive + d] = e[d];
  }
}

// NGP histogram (denpart.f90:20-31).  Each workgroup reduces its particles'
// boxes in registers and LDS and writes one partial, which the next kernel of
// the chain (k_iron_fused, or k_reg_union) combines, so no atomics contend on
// the boxes.  (Combining them in the
// last workgroup to finish, behind a device-scope fence in every workgroup,
// measured 3-4% slower per step at 128^3.)
__device__ __forceinline__ void hist_body(const BoxArgs& b, const double* __restrict__ r,
                                                 unsigned int* __restrict__ cnt, int* __restrict__ part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int ix = 0, iy = 0, iz = 0;
  bool in = false;
  // (r - x1)/h once per component: the NGP index takes int(1.5 + .), the
  // gather cell int(. + 1) (the same values as two separate divisions, and
  // few enough registers for the kernel to start beside the psi kernels)
  double ux = 0.0, uy = 0.0, uz = 0.0;
  if (i < b.npar) {
    ux = (r[i] - b.xmin) / b.hx;
    uy = (r[b.npar + i] - b.ymin) / b.hy;
    uz = (r[2 * b.npar + i] - b.zmin) / b.hz;
    const double fx = 1.5 + ux;
    const double fy = 1.5 + uy;
    const double fz = 1.5 + uz;
    // 1 <= int(f) <= n  <=>  1 <= f < n+1; false for a non-finite position,
    // which is dropped like one outside the box (and never converted)
    in = fx >= 1.0 && fx < b.g.nx + 1.0 && fy >= 1.0 && fy < b.g.ny + 1.0 && fz >= 1.0 && fz < b.g.nz + 1.0;
    if (in) {
      ix = (int)fx;
      iy = (int)fy;
      iz = (int)fz;
    }
  }
  if (in) {
    const long long j = (ix - 1) + (long long)b.g.nx * ((iy - 1) + (long long)b.g.ny * (iz - 1));
    atomicAdd(&cnt[j], 1u);
  }
  // the gather cell's corners are (ix-1 .. ix) 0-based in each direction
  GatherCell q{};
  q.ok = false;
  if (i < b.npar) q = gather_cell_at(b, ux + 1, uy + 1, uz + 1);
  int v[kBoxPart] = {in ? ix - 1 : 0x7fffffff, in ? iy - 1 : 0x7fffffff, in ? iz - 1 : 0x7fffffff,
                     in ? ix - 1 : -1, in ? iy - 1 : -1, in ? iz - 1 : -1,
                     q.ok ? q.ix - 1 : 0x7fffffff, q.ok ? q.iy - 1 : 0x7fffffff, q.ok ? q.iz - 1 : 0x7fffffff,
                     q.ok ? q.ix : -1, q.ok ? q.iy : -1, q.ok ? q.iz : -1};
  box_reduce_wave(v);
  constexpr int W = kBlock / 64;
  __shared__ int sw[W][kBoxPart];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int d = 0; d < kBoxPart; ++d) sw[w][d] = v[d];
  __syncthreads();
  if (threadIdx.x < kBoxPart) {
    const int d = threadIdx.x;
    int m = sw[0][d];
    for (int k = 1; k < W; ++k) m = d % 6 < 3 ? min(m, sw[k][d]) : max(m, sw[k][d]);
    part[blockIdx.x * kBoxPart + d] = m;
  }
}

// Combines the nb partial boxes of k_hist (none without particles) and
// updates the regions.  One wavefront.
__global__ void k_reg_union(Grid g, int* __restrict__ reg, int margin, const int* __restrict__ part, int nb) {
  int v[kBoxPart];
  box_reduce_parts(part, nb, v);
  if (threadIdx.x == 0) reg_union_apply(g, reg, margin, v);
}

// Grid-stride iteration over an inclusive region box.
struct RegionIter {
  int x0, y0, z0, nxr, nyr, vol;
};

__device__ inline RegionIter region_iter(const int* reg) {
  RegionIter it;
  it.x0 = reg[0];
  it.y0 = reg[1];
  it.z0 = reg[2];
  it.nxr = reg[3] - reg[0] + 1;
  it.nyr = reg[4] - reg[1] + 1;
  const int nzr = reg[5] - reg[2] + 1;
  it.vol = (reg[3] < 0 || it.nxr <= 0 || it.nyr <= 0 || nzr <= 0) ? 0 : it.nxr * it.nyr * nzr;
  return it;
}

__device__ inline Idx3 region_point(const RegionIter& it, int v) {
  Idx3 p;
  const int a = v / it.nxr;
  p.ix = it.x0 + (v - a * it.nxr);
  const int b = a / it.nyr;
  p.iy = it.y0 + (a - b * it.nyr);
  p.iz = it.z0 + b;
  return p;
}

// denx = count/N_par/hxyz (denpart.f90:33); optionally the first ironing
// pass (ironing.f90:16-28) is applied on the fly.
template <class R>
__device__ inline double denx_of(const unsigned int* cnt, const R* src, long long j, double npar,
                                 double hxyz) {
  return cnt ? ((double)cnt[j] / npar) / hxyz : (double)src[j];
}

// When psix_out is given, the pass also writes psix = sqrt(denx)
// (denpart.f90:40) for the value it produces.
template <class R>
__global__ void __launch_bounds__(kBlock) k_iron(Grid g, const int* __restrict__ reg,
                                                  const unsigned int* __restrict__ cnt,
                                                  const R* __restrict__ src, R* __restrict__ dst,
                                                  R* __restrict__ psix_out, double npar, double hxyz,
                                                  int smooth) {
  const RegionIter it = region_iter(reg + kRegWork);
  for (int vv = blockIdx.x * blockDim.x + threadIdx.x; vv < it.vol; vv += gridDim.x * blockDim.x) {
  const Idx3 p = region_point(it, vv);
  const long long i = p.ix + (long long)g.nx * (p.iy + (long long)g.ny * p.iz);
  const double c = denx_of(cnt, src, i, npar, hxyz);
  double v = c;
  if (smooth) {
    if (!(p.ix == 0 || p.iy == 0 || p.iz == 0 || p.ix == g.nx - 1 || p.iy == g.ny - 1 || p.iz == g.nz - 1)) {
      const long long sy = g.nx, sz = (long long)g.nx * g.ny;
      double s = denx_of(cnt, src, i - 1, npar, hxyz) + denx_of(cnt, src, i + 1, npar, hxyz);
      s = s + denx_of(cnt, src, i - sy, npar, hxyz);
      s = s + denx_of(cnt, src, i + sy, npar, hxyz);
      s = s + denx_of(cnt, src, i - sz, npar, hxyz);
      s = s + denx_of(cnt, src, i + sz, npar, hxyz);
      v = 8.333333333333333e-2 * s + 0.5 * c;
    }
  }
  dst[i] = (R)v;
  if (psix_out) psix_out[i] = (R)sqrt((double)(R)v);
  }
}

// denx and psix of NP ironing passes (1 <= NP <= 3) in one launch, with the
// region update of k_reg_union folded in: every workgroup combines the
// histogram's partial boxes and forms the work region itself (the live
// region it reads is not written in this launch); workgroup 0 stores the
// boxes and the work region, and k_hpsix, which runs after every workgroup
// here has read the old live region, stores the new one.  Each
// workgroup takes 8^3-point tiles of the work region and keeps the levels in
// LDS -- the histogram densities on the tile grown by NP, pass k on the tile
// grown by NP - k -- so the halo is recomputed instead of exchanged through
// HBM between launches.  Every value is formed exactly as in k_iron (same
// operand order, level >= 1 rounded to R); points outside the grid are
// never read by points inside it, and points outside the work region hold
// zero at every level (the region covers the histogram grown by more than
// NP), so the results equal NP launches of k_iron bit for bit.
template <class R, int NP>
__device__ __forceinline__ void iron_fused_body(const Grid& g, int* __restrict__ reg,
                                                        const int* __restrict__ part, int nb, int margin,
                                                        const unsigned int* __restrict__ cnt, R* __restrict__ dst,
                                                        R* __restrict__ psix, double npar, double hxyz) {
  constexpr int TB = 8;
  constexpr int L0 = TB + 2 * NP;
  constexpr int L1 = NP >= 2 ? TB + 2 * (NP - 1) : 1;
  constexpr int L2 = NP >= 3 ? TB + 2 * (NP - 2) : 1;
  __shared__ double lv0[L0 * L0 * L0];
  __shared__ R lv1[L1 * L1 * L1];
  __shared__ R lv2[L2 * L2 * L2];
  __shared__ int w[6];
  if (threadIdx.x < 64) {
    int v[kBoxPart];
    box_reduce_parts(part, nb, v);
    if (threadIdx.x == 0) {
      int e[6], wr[6];
      box_expand(g, v, margin, e);
      work_region(reg, e, wr);
      for (int d = 0; d < 6; ++d) w[d] = wr[d];
      if (blockIdx.x == 0)
        for (int d = 0; d < 6; ++d) {
          reg[kRegBox + d] = v[d];
          reg[kRegGather + d] = v[6 + d];
          reg[kRegWork + d] = wr[d];
        }
    }
  }
  __syncthreads();
  const RegionIter it = region_iter(w);
  if (it.vol == 0) return;
  const int nzr = w[5] - w[2] + 1;
  const int ntx = (it.nxr + TB - 1) / TB, nty = (it.nyr + TB - 1) / TB, ntz = (nzr + TB - 1) / TB;
  const int ntiles = ntx * nty * ntz;
  const int tid = threadIdx.x;
  auto interior = [&](int x, int y, int z) {
    return !(x == 0 || y == 0 || z == 0 || x == g.nx - 1 || y == g.ny - 1 || z == g.nz - 1);
  };
  auto ingrid = [&](int x, int y, int z) {
    return x >= 0 && y >= 0 && z >= 0 && x < g.nx && y < g.ny && z < g.nz;
  };
  // one smoothing pass from a level of side S (values as double via ld) to
  // the level of side S - 2 whose origin is one point further in
  auto smooth = [&](auto ld, int S, int e, int ox, int oy, int oz) -> double {
    const int D = S - 2;
    const int dz = e / (D * D), rem = e - dz * D * D, dy = rem / D, dx = rem - dy * D;
    const int x = ox + dx, y = oy + dy, z = oz + dz;
    const int sc = (dx + 1) + S * ((dy + 1) + S * (dz + 1));
    const double c = ld(sc);
    if (!ingrid(x, y, z)) return 0.0;
    if (!interior(x, y, z)) return c;
    double s = ld(sc - 1) + ld(sc + 1);
    s = s + ld(sc - S);
    s = s + ld(sc + S);
    s = s + ld(sc - S * S);
    s = s + ld(sc + S * S);
    return 8.333333333333333e-2 * s + 0.5 * c;
  };
  auto ld0 = [&](int k) { return lv0[k]; };
  auto ld1 = [&](int k) { return (double)lv1[k]; };
  auto ld2 = [&](int k) { return (double)lv2[k]; };
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int tz = t / (ntx * nty), tr = t - tz * ntx * nty, ty = tr / ntx, tx = tr - ty * ntx;
    const int ox = it.x0 + tx * TB, oy = it.y0 + ty * TB, oz = it.z0 + tz * TB;
    {
      // every count load of the thread is issued before the first LDS store
      constexpr int NL = (L0 * L0 * L0 + kBlock - 1) / kBlock;
      unsigned int cv[NL];
#pragma unroll
      for (int q = 0; q < NL; ++q) {
        const int e = tid + q * kBlock;
        const int lz = e / (L0 * L0), rem = e - lz * L0 * L0, ly = rem / L0, lx = rem - ly * L0;
        const int x = ox - NP + lx, y = oy - NP + ly, z = oz - NP + lz;
        cv[q] = (e < L0 * L0 * L0 && ingrid(x, y, z)) ? cnt[x + (long long)g.nx * (y + (long long)g.ny * z)] : 0u;
      }
#pragma unroll
      for (int q = 0; q < NL; ++q) {
        const int e = tid + q * kBlock;
        if (e < L0 * L0 * L0) lv0[e] = ((double)cv[q] / npar) / hxyz;
      }
    }
    __syncthreads();
    if constexpr (NP >= 2) {
      for (int e = tid; e < L1 * L1 * L1; e += blockDim.x)
        lv1[e] = (R)smooth(ld0, L0, e, ox - (NP - 1), oy - (NP - 1), oz - (NP - 1));
      __syncthreads();
    }
    if constexpr (NP >= 3) {
      for (int e = tid; e < L2 * L2 * L2; e += blockDim.x)
        lv2[e] = (R)smooth(ld1, L1, e, ox - (NP - 2), oy - (NP - 2), oz - (NP - 2));
      __syncthreads();
    }
    for (int e = tid; e < TB * TB * TB; e += blockDim.x) {
      double v;
      if constexpr (NP == 1) v = smooth(ld0, L0, e, ox, oy, oz);
      else if constexpr (NP == 2) v = smooth(ld1, L1, e, ox, oy, oz);
      else v = smooth(ld2, L2, e, ox, oy, oz);
      const int dz = e / (TB * TB), rem = e - dz * TB * TB, dy = rem / TB, dx = rem - dy * TB;
      const int x = ox + dx, y = oy + dy, z = oz + dz;
      if (x <= w[3] && y <= w[4] && z <= w[5]) {
        const long long i = x + (long long)g.nx * (y + (long long)g.ny * z);
        dst[i] = (R)v;
        psix[i] = (R)sqrt((double)(R)v);
      }
    }
    __syncthreads();
  }
}

// hpsix = -h2o2mx*(d2x+d2y+d2z)psix (denpart.f90:42-46).  It also clears
// the NGP counts over the work region (which holds every occupied cell), so
// the counts are zero everywhere between chains and the next histogram needs
// no clearing pass, and stores live := expand(box, margin) (the value it
// already has when k_reg_union ran).
template <class R>
__device__ __forceinline__ void hpsix_body(const Grid& g, int* __restrict__ reg, const R* __restrict__ psix,
                                                   const Stencil7<double, true>& wx, const Stencil7<double, true>& wy,
                                                   const Stencil7<double, true>& wz, double h2o2mx,
                                                   R* __restrict__ hpsix, unsigned int* __restrict__ cnt,
                                                   int margin) {
  const RegionIter it = region_iter(reg + kRegWork);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    int e[6];
    box_expand(g, reg + kRegBox, margin, e);
    for (int d = 0; d < 6; ++d) reg[kRegLive + d] = e[d];
  }
  for (int vv = blockIdx.x * blockDim.x + threadIdx.x; vv < it.vol; vv += gridDim.x * blockDim.x) {
  const Idx3 p = region_point(it, vv);
  const long long i = p.ix + (long long)g.nx * (p.iy + (long long)g.ny * p.iz);
  cnt[i] = 0u;  // the first ironing pass has read the counts
  const double sx = d13(psix, i, p.ix, g.nx, 1, wx);
  const double sy = d13(psix, i, p.iy, g.ny, g.nx, wy);
  const double sz = d13(psix, i, p.iz, g.nz, (long long)g.nx * g.ny, wz);
  hpsix[i] = (R)(-h2o2mx * ((sx + sy) + sz));
  }
}

struct GatherArgs {
  BoxArgs b;
  Stencil13 d1[3];
  double m_imp;
};

// Quantum part of F_c at grid point (ix,iy,iz), 0-based (potenimp.f90:52-69):
// sto7*(d_c psix*hpsix - psix*d_c hpsix), sto7 = 1/denx if > 0.  The
// component is a template argument so that every read of the kernel's
// argument block has a fixed offset; a run-time choice among its fields
// makes the compiler copy the whole block into private memory.
template <int C, class R>
__device__ __forceinline__ double force_q_c(const GatherArgs& a, const R* __restrict__ denx,
                                            const R* __restrict__ psix, const R* __restrict__ hpsix, int ix,
                                            int iy, int iz) {
  const Grid& g = a.b.g;
  const long long j = ix + (long long)g.nx * (iy + (long long)g.ny * iz);
  const double dx = (double)denx[j];
  const double sto7 = dx > 0.0 ? 1.0 / dx : 0.0;
  const int pos = C == 0 ? ix : (C == 1 ? iy : iz);
  const int n = C == 0 ? g.nx : (C == 1 ? g.ny : g.nz);
  const long long st = C == 0 ? 1 : (C == 1 ? (long long)g.nx : (long long)g.nx * g.ny);
  const double s1 = d13(psix, j, pos, n, st, a.d1[C]);
  const double s4 = d13(hpsix, j, pos, n, st, a.d1[C]);
  return sto7 * (s1 * (double)hpsix[j] - (double)psix[j] * s4);
}

template <class R>
__device__ __forceinline__ double force_q(const GatherArgs& a, int c, const R* __restrict__ denx,
                                          const R* __restrict__ psix, const R* __restrict__ hpsix, int ix, int iy,
                                          int iz) {
  if (c == 0) return force_q_c<0>(a, denx, psix, hpsix, ix, iy, iz);
  if (c == 1) return force_q_c<1>(a, denx, psix, hpsix, ix, iy, iz);
  return force_q_c<2>(a, denx, psix, hpsix, ix, iy, iz);
}

// F_c = (fconv + quantum part)/m_imp at grid point j.
template <class R>
__device__ inline double force_of(const GatherArgs& a, const R* __restrict__ fconv, long long j, double q) {
  return ((double)fconv[j] + q) / a.m_imp;
}

// The quantum part of F_c wherever the trilinear gather reads it.  When the
// gather box (the union of the particles' cell corners) holds at most 8
// points per particle, once per (component, box point): out[c*vol + local],
// local x fastest; evaluated once per grid point instead of once per
// (particle, corner), it shares the stencils of the corners that several
// particles' cells have in common.  Otherwise (a sparse cloud) once per
// (component, corner, particle): out[(c*8 + corner)*N_par + i], corner bits
// (dx,dy,dz) = (bit2,bit1,bit0) in the order x000..x111 of blend_103.  The
// buffer holds 24 N_par doubles either way.
template <class R>
__device__ __forceinline__ void force_q_body(const GatherArgs& a, const int* __restrict__ reg,
                                                    const double* __restrict__ rp, const R* __restrict__ denx,
                                                    const R* __restrict__ psix, const R* __restrict__ hpsix,
                                                    double* __restrict__ out) {
  const RegionIter it = region_iter(reg + kRegGather);
  const int np = a.b.npar;
  if (it.vol <= 8 * np) {
    const int n = 3 * it.vol;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
      const int c = v / it.vol;
      const Idx3 p = region_point(it, v - c * it.vol);
      out[v] = force_q(a, c, denx, psix, hpsix, p.ix, p.iy, p.iz);
    }
  } else {
    const int n = 24 * np;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
      const int slot = v / np;
      const int i = v - slot * np;
      const int c = slot >> 3, k = slot & 7;
      const GatherCell q = gather_cell(a.b, rp, i);
      out[v] = q.ok ? force_q(a, c, denx, psix, hpsix, q.ix - 1 + ((k >> 2) & 1), q.iy - 1 + ((k >> 1) & 1),
                              q.iz - 1 + (k & 1))
                    : 0.0;
    }
  }
}

// Trilinear blend of F_c = (fconv + quantum part)/m_imp at the particles
// (blend_103, potenimp.f90:278-286) from the 8 corners x000..x111 of each
// particle's cell; zero outside [1, n-1] (potenimp.f90:104-105).  One
// thread per (particle, component); the quantum parts in either layout of
// k_force_q.
template <class R>
__device__ __forceinline__ void force_blend_body(const GatherArgs& a, const int* __restrict__ reg,
                                                        const double* __restrict__ rp,
                                                        const R* __restrict__ fcx, const R* __restrict__ fcy,
                                                        const R* __restrict__ fcz, const double* __restrict__ fq,
                                                        double* __restrict__ acc) {
  const int np = a.b.npar;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * np) return;
  const int c = t / np;
  const int i = t - c * np;
  const GatherCell q = gather_cell(a.b, rp, i);
  if (!q.ok) {
    acc[t] = 0.0;
    return;
  }
  const RegionIter it = region_iter(reg + kRegGather);
  const R* fc = c == 0 ? fcx : (c == 1 ? fcy : fcz);
  const Grid& g = a.b.g;
  const long long j0 = (q.ix - 1) + (long long)g.nx * ((q.iy - 1) + (long long)g.ny * (q.iz - 1));
  const long long gy = g.nx, gz = (long long)g.nx * g.ny;
  const bool box = it.vol <= 8 * np;
  const double* f = box ? fq + (long long)c * it.vol + (q.ix - 1 - it.x0) +
                              it.nxr * ((q.iy - 1 - it.y0) + it.nyr * (q.iz - 1 - it.z0))
                        : fq + (long long)c * 8 * np + i;
  const int sx = box ? 1 : 4 * np, sy = box ? it.nxr : 2 * np, sz = box ? it.nxr * it.nyr : np;
  double x[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int dx = (k >> 2) & 1, dy = (k >> 1) & 1, dz = k & 1;
    x[k] = force_of(a, fc, j0 + dx + gy * dy + gz * dz, f[sx * dx + sy * dy + sz * dz]);
  }
  const double x000 = x[0], x001 = x[1], x010 = x[2], x011 = x[3];
  const double x100 = x[4], x101 = x[5], x110 = x[6], x111 = x[7];
  const double rr = q.r, ss = q.s, tt = q.t;
  acc[t] = 1.0 * (+x000) + rr * (-x000 + x100) + ss * (-x000 + x010) + tt * (-x000 + x001) +
           rr * ss * (+x000 - x100 - x010 + x110) + rr * tt * (+x000 - x100 - x001 + x101) +
           ss * tt * (+x000 - x010 - x001 + x011) +
           rr * ss * tt * (-x000 + x100 + x010 + x001 - x011 - x101 - x110 + x111);
}

// ---- particle integrators (steprk.f90:78-105, steppc.f90:64-87,122-151) --
// All particle arrays are (N_par,3) column-major: entry c*N_par + i.
// Element i of every array is addressed as the array's (uniform) base plus
// one 32-bit byte offset 8i, so the loads and stores take the base from
// scalar registers and one VGPR of address serves all of them (3 N_par
// doubles stay below 4 GiB: hetddft_create rejects N_par > kMaxNpar).  The kernels then
// keep to 32 VGPRs and start beside the psi kernels (see kSideVgprs).
constexpr int kMaxNpar = 1 << 27;  // 3 * 8 * kMaxNpar bytes < 4 GiB

__device__ __forceinline__ void part_rk_body(int n3, double* __restrict__ r, double* __restrict__ v,
                          const double* __restrict__ acc, double* __restrict__ qr, double* __restrict__ qv,
                          double* __restrict__ r1_out, double* __restrict__ v1_out,
                          double* __restrict__ a1_out, double ca, double cb, double cc, double dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  const unsigned o = 8u * (unsigned)i;
  const double vi = el(v, o), ai = el(acc, o);
  double st = ca * (vi - cb * el(qr, o));
  el(qr, o) = (el(qr, o) + 3.0 * st) - cc * vi;
  if (r1_out) el(r1_out, o) = el(r, o);
  el(r, o) = el(r, o) + dt * st;
  st = ca * (ai - cb * el(qv, o));
  el(qv, o) = (el(qv, o) + 3.0 * st) - cc * ai;
  if (a1_out) el(a1_out, o) = ai;
  if (v1_out) el(v1_out, o) = vi;
  el(v, o) = vi + dt * st;
}

__device__ __forceinline__ void part_predict_body(int n3, double* __restrict__ r, const double* __restrict__ v,
                               const double* __restrict__ acc, double* r3, const double* __restrict__ v1,
                               const double* __restrict__ v2, double* v3, const double* __restrict__ a1,
                               double* a2, double* __restrict__ pcr, double* __restrict__ pcv,
                               double c4o3dt, double c112) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  const unsigned o = 8u * (unsigned)i;
  double st = el(r3, o) + c4o3dt * ((2.0 * el(v, o) - el(v1, o)) + 2.0 * el(v2, o));
  el(r3, o) = el(r, o);
  el(r, o) = st - c112 * el(pcr, o);
  el(pcr, o) = st;
  st = el(v3, o) + c4o3dt * ((2.0 * el(acc, o) - el(a1, o)) + 2.0 * el(a2, o));
  el(v3, o) = st - c112 * el(pcv, o);
  el(pcv, o) = st;
  el(a2, o) = el(acc, o);
}

__device__ __forceinline__ void part_correct_body(int n3, double* __restrict__ r, double* __restrict__ v,
                               const double* __restrict__ acc, const double* __restrict__ r3,
                               const double* __restrict__ r2, const double* __restrict__ v1,
                               const double* __restrict__ v2, double* v3, const double* __restrict__ a1,
                               const double* __restrict__ a2, double* __restrict__ pcr,
                               double* __restrict__ pcv, double* __restrict__ part, double dt3,
                               double c9) {
  __shared__ double sh[16];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double er = 0.0, ev = 0.0;
  if (i < n3) {
    const unsigned o = 8u * (unsigned)i;
    double st = 0.125 * ((9.0 * el(r3, o) - el(r2, o)) + dt3 * ((el(v3, o) + 2.0 * el(v, o)) - el(v1, o)));
    double p = el(pcr, o) - st;
    el(pcr, o) = p;
    el(r, o) = st + c9 * p;
    er = fabs(c9 * p);
    st = 0.125 * ((9.0 * el(v, o) - el(v2, o)) + dt3 * ((el(acc, o) + 2.0 * el(a1, o)) - el(a2, o)));
    p = el(pcv, o) - st;
    el(pcv, o) = p;
    el(v3, o) = el(v, o);
    el(v, o) = st + c9 * p;
    ev = fabs(c9 * p);
  }
  double s = block_sum(er, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  __syncthreads();
  s = block_sum(ev, sh);
  if (threadIdx.x == 0) part[gridDim.x + blockIdx.x] = s;
}

// Per-component sums of r, v and a for the per-step means (DFT4He3d.f90:864-878).
__global__ void k_part_sums(int np, const double* __restrict__ r, const double* __restrict__ v,
                            const double* __restrict__ acc, double* __restrict__ part) {
  __shared__ double sh[16];
  const int comp = blockIdx.y;  // 0..8
  const double* src = comp < 3 ? r : (comp < 6 ? v : acc);
  const int c = comp % 3;
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x)
    s += src[(long long)c * np + i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[comp * gridDim.x + blockIdx.x] = s;
}

// Means from the 9 particle sums, scaled as hetddft_particle_means scales
// them on the host: r in A, v in A/ps, a in A/ps^2 (DFT4He3d.f90:864-878).
__global__ void k_part_means_scale(const double* __restrict__ sums, int np, double pstok, double* __restrict__ out) {
  const int k = threadIdx.x;
  if (k >= 9) return;
  double v = sums[k] / np;
  if (k >= 3 && k < 6) v = v * pstok;
  if (k >= 6) v = v * pstok * pstok;
  out[k] = v;
}

// ---- energy (energy.f90:32-124) ----------------------------------------
struct EnergyArgs {
  Grid g;
  Stencil13 d1x, d1y, d1z;
  double aux1, aux2;   // cp4/2, cpp4/3
  double den0s, beta, den_m;
  int otc;
};

// Slots: 0 |grad psi|^2, 1 delj4*den, 2 correlation, 3 alpha_s, 4 solid,
// 5 psix*hpsix, 6 uimp*den, 7 den (norm).
template <class R>
__global__ void __launch_bounds__(kBlock) k_energy(EnergyArgs a, const c64* __restrict__ psi,
                                                    const double* __restrict__ den,
                                                    const double* __restrict__ dencg,
                                                    const double* __restrict__ delj4, const R* __restrict__ falfs,
                                                    const R* __restrict__ dxd, const R* __restrict__ dyd,
                                                    const R* __restrict__ dzd, const R* __restrict__ ix,
                                                    const R* __restrict__ iy, const R* __restrict__ iz,
                                                    const R* __restrict__ psix, const R* __restrict__ hpsix,
                                                    const double* __restrict__ uimp, double* __restrict__ part) {
  __shared__ double sh[16];
  const Grid& g = a.g;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (i < g.n) {
    const Idx3 p = idx3(i, g);
    const c64 gx = d13c(psi, i, p.ix, g.nx, 1, a.d1x);
    const c64 gy = d13c(psi, i, p.iy, g.ny, g.nx, a.d1y);
    const c64 gz = d13c(psi, i, p.iz, g.nz, (long long)g.nx * g.ny, a.d1z);
    const double ax = hypot(gx.x, gx.y), ay = hypot(gy.x, gy.y), az = hypot(gz.x, gz.y);
    v[0] = ax * ax + ay * ay + az * az;
    const double dn = (double)den[i], dcg = (double)dencg[i];
    v[1] = (double)delj4[i] * dn;
    v[2] = dn * (dcg * dcg) * (a.aux1 + a.aux2 * dcg);
    if (a.otc) {
      const double fal = (double)falfs[i];
      v[3] = fal * ((double)dxd[i] * (double)ix[i] + (double)dyd[i] * (double)iy[i] +
                    (double)dzd[i] * (double)iz[i]);
    }
    v[4] = dn * (1.0 + tanh(a.beta * (dn - a.den_m)));
    v[5] = (double)psix[i] * (double)hpsix[i];
    v[6] = (double)uimp[i] * dn;
    v[7] = dn;
  }
  for (int s = 0; s < 8; ++s) {
    const double t = block_sum(v[s], sh);
    if (threadIdx.x == 0) part[(long long)s * gridDim.x + blockIdx.x] = t;
    __syncthreads();
  }
}

__global__ void k_sq_sum(int n, const double* __restrict__ v, double* __restrict__ part) {
  __shared__ double sh[16];
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += v[i] * v[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// ---- per-step diagnostics ---------------------------------------------------
// The driver's per-step output (DFT4He3d.f90:879-930): the He count
// sum(den) dxyz, the centre-of-mass moments of r_cm (r_cm.f90:23-58), the
// impurity mean-square radii sqrt(sum(denx x^2) dxyz) (:918-928); and the
// chemical potential <psi|H|psi>/<psi|psi> with H = -h2o2m4 lap + pot4 and the
// current j = 2 h2o2m4 Im(conj(psi) grad psi), with the 13-point stencils of
// the step.  x(ix) = -xmax + hx (ix-1) (DFT4He3d.f90:302-310).
struct DiagArgs {
  Grid g;
  Stencil13 d1x, d1y, d1z, d2x, d2y, d2z;
  double xmin, ymin, zmin, hx, hy, hz;
  double h2o2m4;
};

constexpr int kDiagSlots = 11;

// Slots: 0 |psi|^2, 1-3 |psi|^2 (x, y, z), 4-6 denx (x^2, y^2, z^2),
// 7 Re conj(psi) (-h2o2m4 lap psi + pot4 psi), 8-10 Im conj(psi) d_c psi.
template <class R>
__global__ void __launch_bounds__(kBlock) k_diag(DiagArgs a, const c64* __restrict__ psi,
                                                  const R* __restrict__ pot4, const R* __restrict__ denx,
                                                  double* __restrict__ part) {
  __shared__ double sh[16];
  const Grid& g = a.g;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v[kDiagSlots] = {};
  if (i < g.n) {
    const Idx3 p = idx3(i, g);
    const double x = a.xmin + a.hx * p.ix, y = a.ymin + a.hy * p.iy, z = a.zmin + a.hz * p.iz;
    const c64 ps = psi[i];
    const double rho = ps.x * ps.x + ps.y * ps.y;
    v[0] = rho;
    v[1] = rho * x;
    v[2] = rho * y;
    v[3] = rho * z;
    const double dx = (double)denx[i];
    v[4] = dx * (x * x);
    v[5] = dx * (y * y);
    v[6] = dx * (z * z);
    const long long sxy = (long long)g.nx * g.ny;
    const c64 lap = (d13c(psi, i, p.ix, g.nx, 1, a.d2x) + d13c(psi, i, p.iy, g.ny, g.nx, a.d2y)) +
                    d13c(psi, i, p.iz, g.nz, sxy, a.d2z);
    const double pv = (double)pot4[i];
    const c64 hp = mk(pv * ps.x - a.h2o2m4 * lap.x, pv * ps.y - a.h2o2m4 * lap.y);
    v[7] = ps.x * hp.x + ps.y * hp.y;
    const c64 gx = d13c(psi, i, p.ix, g.nx, 1, a.d1x);
    const c64 gy = d13c(psi, i, p.iy, g.ny, g.nx, a.d1y);
    const c64 gz = d13c(psi, i, p.iz, g.nz, sxy, a.d1z);
    v[8] = ps.x * gx.y - ps.y * gx.x;
    v[9] = ps.x * gy.y - ps.y * gy.x;
    v[10] = ps.x * gz.y - ps.y * gz.x;
  }
  for (int s = 0; s < kDiagSlots; ++s) {
    const double t = block_sum(v[s], sh);
    if (threadIdx.x == 0) part[(long long)s * gridDim.x + blockIdx.x] = t;
    __syncthreads();
  }
}

// j = 2 h2o2m4 Im(conj(psi) grad psi) at every point.
__global__ void __launch_bounds__(kBlock) k_current(DiagArgs a, const c64* __restrict__ psi,
                                                     double* __restrict__ jx, double* __restrict__ jy,
                                                     double* __restrict__ jz) {
  const Grid& g = a.g;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  const Idx3 p = idx3(i, g);
  const c64 ps = psi[i];
  const long long sxy = (long long)g.nx * g.ny;
  const c64 gx = d13c(psi, i, p.ix, g.nx, 1, a.d1x);
  const c64 gy = d13c(psi, i, p.iy, g.ny, g.nx, a.d1y);
  const c64 gz = d13c(psi, i, p.iz, g.nz, sxy, a.d1z);
  const double c = 2.0 * a.h2o2m4;
  jx[i] = c * (ps.x * gx.y - ps.y * gx.x);
  jy[i] = c * (ps.x * gy.y - ps.y * gy.x);
  jz[i] = c * (ps.x * gz.y - ps.y * gz.x);
}

// min and max of column c of an (N_par, 3) particle array: one block per
// (array, component); out[2k] = min, out[2k+1] = max, k = 3*array + c.
__global__ void __launch_bounds__(kBlock) k_part_minmax(int np, const double* __restrict__ v,
                                                         const double* __restrict__ acc, double* __restrict__ out) {
  __shared__ double smin[kBlock], smax[kBlock];
  const int k = blockIdx.x;  // 0..5
  const double* col = (k < 3 ? v : acc) + (long long)(k % 3) * np;
  double lo = INFINITY, hi = -INFINITY;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    lo = fmin(lo, col[i]);
    hi = fmax(hi, col[i]);
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * k] = smin[0];
    out[2 * k + 1] = smax[0];
  }
}

// ---- particle sampling (sorteo.f90:1-37) --------------------------------------
// Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC11): a counter-based generator,
// so particle i's attempt a draws the block of counter (i, a, 0, 0) under the
// key (seed, 0) with no state shared between threads.
struct U64x4 {
  unsigned long long v[4];
};

__device__ inline U64x4 philox4x64(U64x4 c, unsigned long long k0, unsigned long long k1) {
  constexpr unsigned long long M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
  constexpr unsigned long long W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    const unsigned long long hi0 = __umul64hi(M0, c.v[0]), lo0 = M0 * c.v[0];
    const unsigned long long hi1 = __umul64hi(M1, c.v[2]), lo1 = M1 * c.v[2];
    U64x4 o;
    o.v[0] = hi1 ^ c.v[1] ^ k0;
    o.v[1] = lo1;
    o.v[2] = hi0 ^ c.v[3] ^ k1;
    o.v[3] = lo0;
    c = o;
  }
  return c;
}

__device__ inline double u53(unsigned long long v) { return (double)(v >> 11) * (1.0 / 9007199254740992.0); }

struct SortArgs {
  Grid g;
  double x0, y0, z0, dx, dy, dz;  // x(1) and x(2) - x(1) of the grid arrays
  double lo[3], hi[3];
  double denmax;
  unsigned long long seed;
  long long max_attempts;
};

// One thread per particle: uniform point in [lo, hi], uniform dum in
// [0, denmax); accepted when dum < trilinear denx at the point (interpola3denx,
// sorteo.f90:40-74).  rimp is (N_par, 3); a particle not accepted within
// max_attempts sets *failed.
__global__ void __launch_bounds__(kBlock) k_sorteo(SortArgs a, int np, const double* __restrict__ denx,
                                                    double* __restrict__ rimp, int* __restrict__ failed) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const Grid& g = a.g;
  for (long long t = 0; t < a.max_attempts; ++t) {
    U64x4 c;
    c.v[0] = (unsigned long long)i;
    c.v[1] = (unsigned long long)t;
    c.v[2] = 0;
    c.v[3] = 0;
    const U64x4 w = philox4x64(c, a.seed, 0ull);
    double xyz[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) xyz[k] = u53(w.v[k]) * (a.hi[k] - a.lo[k]) + a.lo[k];
    const double dum = u53(w.v[3]) * a.denmax;
    const double ax = (xyz[0] - a.x0) / a.dx + 1.0;
    const double ay = (xyz[1] - a.y0) / a.dy + 1.0;
    const double az = (xyz[2] - a.z0) / a.dz + 1.0;
    const int ix = (int)ax, iy = (int)ay, iz = (int)az;
    double f = 0.0;
    if (!(ix >= g.nx || iy >= g.ny || iz >= g.nz || ix <= 0 || iy <= 0 || iz <= 0)) {
      const double r = ax - ix, s = ay - iy, u = az - iz;
      auto d = [&](int p, int q, int k) { return denx[(p - 1) + (long long)g.nx * ((q - 1) + (long long)g.ny * (k - 1))]; };
      const double x000 = d(ix, iy, iz), x001 = d(ix, iy, iz + 1), x010 = d(ix, iy + 1, iz), x011 = d(ix, iy + 1, iz + 1);
      const double x100 = d(ix + 1, iy, iz), x101 = d(ix + 1, iy, iz + 1), x110 = d(ix + 1, iy + 1, iz),
                   x111 = d(ix + 1, iy + 1, iz + 1);
      f = 1.0 * (+x000) + r * (-x000 + x100) + s * (-x000 + x010) + u * (-x000 + x001) +
          r * s * (+x000 - x100 - x010 + x110) + r * u * (+x000 - x100 - x001 + x101) +
          s * u * (+x000 - x010 - x001 + x011) + r * s * u * (-x000 + x100 + x010 + x001 - x011 - x101 - x110 + x111);
    }
    if (dum < f) {
      rimp[i] = xyz[0];
      rimp[np + i] = xyz[1];
      rimp[2 * np + i] = xyz[2];
      return;
    }
  }
  atomicOr(failed, 1);
}

// ---- particle-side launches carrying several members -----------------------
// The histogram and the integrators run on the side stream beside the psi
// kernels, which leave as few as 32 VGPRs per SIMD lane free (4 waves of
// 120 in the widest tuned shapes); a side wave of at most 32 VGPRs starts at
// once instead of queueing until the psi launch drains.
// blockIdx.y is the member; each member's work uses blockIdx.x and gridDim.x
// exactly as a launch of its own.  All members share the box and step.
struct HistTab {
  const double* r[kMaxMembers];
  unsigned int* cnt[kMaxMembers];
  int* part[kMaxMembers];
};
__global__ void __launch_bounds__(kBlock) k_hist(BoxArgs b, HistTab t) {
  const int m = blockIdx.y;
  hist_body(b, t.r[m], t.cnt[m], t.part[m]);
}

template <class R>
struct IronTab {
  int* reg[kMaxMembers];
  const int* part[kMaxMembers];
  const unsigned int* cnt[kMaxMembers];
  R* dst[kMaxMembers];
  R* psix[kMaxMembers];
};
template <class R, int NP>
__global__ void __launch_bounds__(kBlock) k_iron_fused(Grid g, IronTab<R> t, int nb, int margin, double npar,
                                                        double hxyz) {
  const int m = blockIdx.y;
  iron_fused_body<R, NP>(g, t.reg[m], t.part[m], nb, margin, t.cnt[m], t.dst[m], t.psix[m], npar, hxyz);
}

template <class R>
struct HpsixTab {
  int* reg[kMaxMembers];
  const R* psix[kMaxMembers];
  R* hpsix[kMaxMembers];
  unsigned int* cnt[kMaxMembers];
};
template <class R>
__global__ void __launch_bounds__(kBlock) k_hpsix(Grid g, HpsixTab<R> t, Stencil7<double, true> wx,
                                                   Stencil7<double, true> wy, Stencil7<double, true> wz,
                                                   double h2o2mx, int margin) {
  const int m = blockIdx.y;
  hpsix_body<R>(g, t.reg[m], t.psix[m], wx, wy, wz, h2o2mx, t.hpsix[m], t.cnt[m], margin);
}

template <class R>
struct ForceQTab {
  const int* reg[kMaxMembers];
  const double* rp[kMaxMembers];
  const R* denx[kMaxMembers];
  const R* psix[kMaxMembers];
  const R* hpsix[kMaxMembers];
  double* out[kMaxMembers];
};
template <class R>
__global__ void __launch_bounds__(kBlock) k_force_q(GatherArgs a, ForceQTab<R> t) {
  const int m = blockIdx.y;
  force_q_body<R>(a, t.reg[m], t.rp[m], t.denx[m], t.psix[m], t.hpsix[m], t.out[m]);
}

template <class R>
struct BlendTab {
  const int* reg[kMaxMembers];
  const double* rp[kMaxMembers];
  const R* fcx[kMaxMembers];
  const R* fcy[kMaxMembers];
  const R* fcz[kMaxMembers];
  const double* fq[kMaxMembers];
  double* acc[kMaxMembers];
};
template <class R>
__global__ void __launch_bounds__(kBlock) k_force_blend(GatherArgs a, BlendTab<R> t) {
  const int m = blockIdx.y;
  force_blend_body<R>(a, t.reg[m], t.rp[m], t.fcx[m], t.fcy[m], t.fcz[m], t.fq[m], t.acc[m]);
}

struct PartRkTab {
  double* r[kMaxMembers];
  double* v[kMaxMembers];
  const double* acc[kMaxMembers];
  double* qr[kMaxMembers];
  double* qv[kMaxMembers];
  double* r1[kMaxMembers];
  double* v1[kMaxMembers];
  double* a1[kMaxMembers];
};
__global__ void __launch_bounds__(kBlock) k_part_rk(int n3, PartRkTab t, double ca, double cb, double cc, double dt) {
  const int m = blockIdx.y;
  part_rk_body(n3, t.r[m], t.v[m], t.acc[m], t.qr[m], t.qv[m], t.r1[m], t.v1[m], t.a1[m], ca, cb, cc, dt);
}

struct PartPredTab {
  double* r[kMaxMembers];
  const double* v[kMaxMembers];
  const double* acc[kMaxMembers];
  double* r3[kMaxMembers];
  const double* v1[kMaxMembers];
  const double* v2[kMaxMembers];
  double* v3[kMaxMembers];
  const double* a1[kMaxMembers];
  double* a2[kMaxMembers];
  double* pcr[kMaxMembers];
  double* pcv[kMaxMembers];
};
__global__ void __launch_bounds__(kBlock) k_part_predict(int n3, PartPredTab t, double c4o3dt, double c112) {
  const int m = blockIdx.y;
  part_predict_body(n3, t.r[m], t.v[m], t.acc[m], t.r3[m], t.v1[m], t.v2[m], t.v3[m], t.a1[m], t.a2[m], t.pcr[m],
                    t.pcv[m], c4o3dt, c112);
}

struct PartCorrTab {
  double* r[kMaxMembers];
  double* v[kMaxMembers];
  const double* acc[kMaxMembers];
  const double* r3[kMaxMembers];
  const double* r2[kMaxMembers];
  const double* v1[kMaxMembers];
  const double* v2[kMaxMembers];
  double* v3[kMaxMembers];
  const double* a1[kMaxMembers];
  const double* a2[kMaxMembers];
  double* pcr[kMaxMembers];
  double* pcv[kMaxMembers];
  double* part[kMaxMembers];
};
__global__ void __launch_bounds__(kBlock) k_part_correct(int n3, PartCorrTab t, double dt3, double c9) {
  const int m = blockIdx.y;
  part_correct_body(n3, t.r[m], t.v[m], t.acc[m], t.r3[m], t.r2[m], t.v1[m], t.v2[m], t.v3[m], t.a1[m], t.a2[m],
                    t.pcr[m], t.pcv[m], t.part[m], dt3, c9);
}

}  // namespace hetd
