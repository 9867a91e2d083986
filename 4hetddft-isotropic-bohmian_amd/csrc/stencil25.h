// z-marching ("2.5D") engine for the 13-point star stencils of derivnD
// (icon=13, zero outside the box: init_deriv_parallel.f90:428-535).
//
// A workgroup of TX x TY threads owns one (x, y) column tile and walks
// a chunk of z planes.  Each thread keeps the 13 z-neighbours of its column
// in a register queue; the x and y neighbours come from an LDS copy of the
// current plane with a 6-point halo.  Per output point this costs one fresh
// HBM read plus the halo reads, instead of 37 scattered loads.
//
// Sums run over the offsets -6..6 in order with zeros outside the box, the
// same sequence of operations as the direct form (d13/d13c).
#pragma once
#include <type_traits>
#include "common.h"

namespace hetd {

// Tile shape: TX x TY columns per workgroup (one thread each), LDS plane
// tile of LX x LY with the 6-point halo.  Per output point a tile reads
// about 1 + 12/TX + 12/TY + 12/zc values of the stencilled field, so wider
// tiles and longer z chunks cut re-reads at the cost of fewer workgroups.
template <int TXv, int TYv>
struct Tile25 {
  static constexpr int TX = TXv, TY = TYv, LX = TXv + 12, LY = TYv + 12, N = TXv * TYv;
};

__device__ inline double vzero(double) { return 0.0; }
__device__ inline float vzero(float) { return 0.0f; }
__device__ inline c64 vzero(c64) { return mk(0.0, 0.0); }
__device__ inline double vmadd(double acc, double w, double v) { return acc + w * v; }
__device__ inline float vmadd(float acc, float w, float v) { return acc + w * v; }
__device__ inline c64 vmadd(c64 acc, double w, c64 v) { return mk(acc.x + w * v.x, acc.y + w * v.y); }

// acc + w*v as written (FU = false: separate rounding of the product, the
// reference's order) or fused (FU = true: one rounding, half the FP64
// instructions; used by the mixed path, whose psi update has no bitwise
// reference to follow).
template <bool FU>
__device__ inline double madd(double acc, double w, double v) {
  if constexpr (FU) return fma(w, v, acc);
  else return vmadd(acc, w, v);
}
template <bool FU>
__device__ inline float madd(float acc, float w, float v) {
  if constexpr (FU) return fmaf(w, v, acc);
  else return vmadd(acc, w, v);
}
template <bool FU>
__device__ inline c64 madd(c64 acc, double w, c64 v) {
  if constexpr (FU) return mk(fma(w, v.x, acc.x), fma(w, v.y, acc.y));
  else return vmadd(acc, w, v);
}

// Store of a real field that the next kernels read.  An FP64 field streams
// (st_stream: it would not stay in the caches beside the halo rows of the
// stencil kernels anyway); an FP32 field of the mixed path is stored plainly,
// so that the next kernels find much of it in the caches (streaming the FP32
// gradient and pot4 cost 1% of the mixed PC step at 128^3).
template <class R>
__device__ __forceinline__ void st_field(R* p, R v) {
  if constexpr (sizeof(R) == 8) st_stream(p, v);
  else *p = v;
}

template <class V, class R>
__device__ inline V vload(const R* f, int i) {
  return (V)f[i];
}
template <>
__device__ inline c64 vload<c64, c64>(const c64* f, int i) {
  return f[i];
}

struct Col {
  int x, y, tx, ty, tid;
  int bx, by, bz;  // tile of this workgroup (x tile, y tile, z chunk)
  int m;           // member (instance) of a launch that carries several
  int blk;         // workgroup index within the member
  bool in;
};


// Tile of this workgroup.  Workgroups are dispatched round-robin over the 8
// XCDs, each with its own L2.  With xcd set, XCD k takes a contiguous range
// of tiles (x tile fastest, then y tile, then z chunk), so the tiles that
// share x/y halos and z-queue planes read them through the same L2 at about
// the same time.  The mapping is a bijection for any block count.
// A launch that carries several members has nzb z chunks per member in
// gridDim.z (member-major): logical workgroup l of member m = l / (gx gy nzb).
template <class SH>
__device__ inline Col col_of(const Grid& g, int xcd, int nzb) {
  Col c;
  c.tx = threadIdx.x;
  c.ty = threadIdx.y;
  c.tid = c.tx + SH::TX * c.ty;
  const int gx = gridDim.x, gy = gridDim.y;
  const int p = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  int l = p;
  if (xcd) {
    constexpr int NXCD = 8;
    const int nb = gx * gy * gridDim.z;
    const int q = nb / NXCD, r = nb % NXCD, k = p % NXCD;
    l = k * q + min(k, r) + p / NXCD;
  }
  const int per = gx * gy * nzb;
  c.m = l / per;
  l -= c.m * per;
  c.blk = l;
  c.bx = l % gx;
  c.by = (l / gx) % gy;
  c.bz = l / (gx * gy);
  c.x = c.bx * SH::TX + c.tx;
  c.y = c.by * SH::TY + c.ty;
  c.in = c.x < g.nx && c.y < g.ny;
  return c;
}

// Element offset of (x, y, z); grids are at most 256^3, so 32 bits hold it.
__device__ inline int gidx(const Grid& g, int x, int y, int z) {
  return x + g.nx * (y + g.ny * z);
}

// Queue of the 13 z-neighbours of column (x, y): q[k] = f(z - 6 + k).
template <class V, class R>
struct ZQueue {
  V q[13];
  __device__ inline void init(const R* f, const Grid& g, const Col& c, int z0) {
    const int xc = min(c.x, g.nx - 1), yc = min(c.y, g.ny - 1);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      const int z = z0 - 6 + k;
      const V t = vload<V, R>(f, gidx(g, xc, yc, min(max(z, 0), g.nz - 1)));
      q[k] = (c.in && z >= 0 && z < g.nz) ? t : vzero(V());
    }
  }
  // Value entering the queue at plane z: f(z + 6), zero outside the box.
  __device__ inline V next(const R* f, const Grid& g, const Col& c, int z) const {
    const int zz = z + 6;
    const int xc = min(c.x, g.nx - 1), yc = min(c.y, g.ny - 1);
    const V t = vload<V, R>(f, gidx(g, xc, yc, min(zz, g.nz - 1)));
    return (c.in && zz < g.nz) ? t : vzero(V());
  }
  __device__ inline void fetch(const R* f, const Grid& g, const Col& c, int z) { q[12] = next(f, g, c, z); }
  __device__ inline void shift() {
#pragma unroll
    for (int k = 0; k < 12; ++k) q[k] = q[k + 1];
  }
  __device__ inline V centre() const { return q[6]; }
  template <bool FU = false, class W>
  __device__ inline V dz(const W& w) const {
    V acc = vzero(V());
#pragma unroll
    for (int m = 0; m < 13; ++m) acc = madd<FU>(acc, wat(w, m), q[m]);
    return acc;
  }
};

// The same tile in two steps, so that the loads of the next plane can be in
// flight while the current plane is computed: plane_load issues the halo
// loads into registers, plane_store writes them (and the centre) to LDS.
// Only the star halo is loaded: 6 rows above and below the tile's TX
// columns and 6 columns left and right of its TY rows.  The interior comes
// from the z queue and the corners of the LX x LY tile are never read.
template <class SH>
struct Halo {
  static constexpr int NYH = 12 * SH::TX;    // y halo: rows 0..5 and TY+6..TY+11
  static constexpr int NH = NYH + 12 * SH::TY;
  static constexpr int NE = (NH + SH::N - 1) / SH::N;
};

// Tile position (lx, ly) of halo element h < NH.
template <class SH>
__device__ inline void halo_pos(int h, int& lx, int& ly) {
  if (h < Halo<SH>::NYH) {
    const int r = h / SH::TX;
    lx = 6 + (h - r * SH::TX);
    ly = r < 6 ? r : r + SH::TY;
  } else {
    const int hx = h - Halo<SH>::NYH;
    const int r = hx / 12, k = hx - r * 12;
    ly = 6 + r;
    lx = k < 6 ? k : k + SH::TX;
  }
}

template <class SH, class V>
struct PlaneRegs {
  static constexpr int NE = Halo<SH>::NE;
  V v[NE];
};

template <class SH, class V, class R>
__device__ inline void plane_load(PlaneRegs<SH, V>& pr, const R* f, const Grid& g, const Col& c, int z) {
  const int x0 = c.bx * SH::TX - 6, y0 = c.by * SH::TY - 6;
#pragma unroll
  for (int k = 0; k < PlaneRegs<SH, V>::NE; ++k) {
    int lx, ly;
    halo_pos<SH>(min(c.tid + k * SH::N, Halo<SH>::NH - 1), lx, ly);
    const int x = x0 + lx, y = y0 + ly;
    const bool inb = x >= 0 && x < g.nx && y >= 0 && y < g.ny;
    const int xc = min(max(x, 0), g.nx - 1), yc = min(max(y, 0), g.ny - 1);
    const V t = vload<V, R>(f, gidx(g, xc, yc, z));
    pr.v[k] = inb ? t : vzero(V());
  }
}

template <class SH, class V>
__device__ inline void plane_store(V* tile, const PlaneRegs<SH, V>& pr, const Col& c, V centre) {
  tile[(c.ty + 6) * SH::LX + c.tx + 6] = centre;
#pragma unroll
  for (int k = 0; k < PlaneRegs<SH, V>::NE; ++k) {
    const int h = c.tid + k * SH::N;
    if (h < Halo<SH>::NH) {
      int lx, ly;
      halo_pos<SH>(h, lx, ly);
      tile[ly * SH::LX + lx] = pr.v[k];
    }
  }
}

template <class SH, bool FU = false, class V, class W>
__device__ inline V tile_dx(const V* tile, const Col& c, const W& w) {
  const V* row = tile + (c.ty + 6) * SH::LX + c.tx;
  V acc = vzero(V());
#pragma unroll
  for (int m = 0; m < 13; ++m) acc = madd<FU>(acc, wat(w, m), row[m]);
  return acc;
}

template <class SH, bool FU = false, class V, class W>
__device__ inline V tile_dy(const V* tile, const Col& c, const W& w) {
  const V* colp = tile + c.ty * SH::LX + c.tx + 6;
  V acc = vzero(V());
#pragma unroll
  for (int m = 0; m < 13; ++m) acc = madd<FU>(acc, wat(w, m), colp[m * SH::LX]);
  return acc;
}

// Block sum for the 2D engine blocks (SH::N threads).
template <class SH>
__device__ inline double block_sum25(double v, double* sh, int tid) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if ((tid & 63) == 0) sh[tid >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (tid == 0)
    for (int k = 0; k < SH::N / 64; ++k) s += sh[k];
  return s;
}

// ---- psi-side kernels on the engine ------------------------------------
struct Ham25 {
  Grid g;
  Stencil7<double, true> d2x, d2y, d2z;
  double h2o2m4;
  const c64* timec_z;
  int zc;   // z planes per workgroup
  int xcd;  // XCD-grouped tile order (col_of)
};

// Hpsi at the column's current plane (steppc.f90:36-46).  FU: fused
// stencil sums (the mixed path).
template <class SH, bool FU = false>
__device__ inline c64 ham25(const Ham25& a, const c64* tile, const ZQueue<c64, c64>& q, const Col& c,
                            double pv, int z) {
  const c64 sx = tile_dx<SH, FU>(tile, c, a.d2x);
  const c64 sy = tile_dy<SH, FU>(tile, c, a.d2y);
  const c64 sz = q.template dz<FU>(a.d2z);
  const c64 lap = (sx + sy) + sz;
  const c64 ps = q.centre();
  const c64 t = mk(lap.x * a.h2o2m4 - pv * ps.x, lap.y * a.h2o2m4 - pv * ps.y);
  return cmul(a.timec_z[z], t);
}

// Hamming predictor + modifier (steppc.f90:46-56).  PF (software pipeline):
// the tile, queue entry and centre values of plane z+1 are loaded while plane
// z is computed.  Without PF each plane's loads are issued in its own
// iteration, which needs about 40 fewer VGPRs (one more resident wave per
// SIMD) and gives the same results.
// The reference stores the predicted value in pc (pc = Sto1c, :55); here pc
// keeps the previous corrector's value and the corrector forms Sto1c again
// from the modified psi it reads, Sto1c = psi + c112 pc (the modifier
// inverted, equal up to one rounding): one 16 B store per point fewer.
template <class R, class SH, bool PF>
__device__ __forceinline__ void pc_predict25(const Ham25& a, const Col& c, const c64* __restrict__ psi,
                                             const R* __restrict__ pot4, c64* old3,
                                             const c64* __restrict__ h1, c64* h2,
                                             const c64* __restrict__ pc, R* __restrict__ den,
                                             double c4o3dt, double c112) {
  __shared__ c64 tile[SH::LX * SH::LY];
  const Grid& g = a.g;
  const int z0 = c.bz * a.zc;
  const int z1 = min(z0 + a.zc, g.nz);
  const int xc = min(c.x, g.nx - 1), yc = min(c.y, g.ny - 1);
  ZQueue<c64, c64> q;
  q.init(psi, g, c, z0);
  PlaneRegs<SH, c64> pr;
  c64 qn, o3n, a1n, a2n, pcn;
  double pvn;
  auto prefetch = [&](int z) {
    qn = q.next(psi, g, c, z);
    plane_load<SH>(pr, psi, g, c, z);
    const int j = gidx(g, xc, yc, z);
    pvn = (double)ld_stream(pot4 + j);
    o3n = ld_stream(old3 + j);
    a1n = ld_stream(h1 + j);
    a2n = ld_stream(h2 + j);
    pcn = ld_stream(pc + j);
  };
  if constexpr (PF) {
    if (z0 < z1) prefetch(z0);
  }
  for (int z = z0; z < z1; ++z) {
    if constexpr (!PF) prefetch(z);
    q.q[12] = qn;
    plane_store<SH>(tile, pr, c, q.q[6]);
    const double pv = pvn;
    const c64 o3 = o3n, a1 = a1n, a2 = a2n, pci = pcn;
    __syncthreads();
    if constexpr (PF) {
      if (z + 1 < z1) prefetch(z + 1);
    }
    if (c.in) {
      const int i = gidx(g, c.x, c.y, z);
      const c64 s4 = ham25<SH, std::is_same<R, float>::value>(a, tile, q, c, pv, z);
      const c64 s1 = mk(o3.x + c4o3dt * ((2.0 * s4.x - a1.x) + 2.0 * a2.x),
                        o3.y + c4o3dt * ((2.0 * s4.y - a1.y) + 2.0 * a2.y));
      const c64 np = mk(s1.x - c112 * pci.x, s1.y - c112 * pci.y);
      st_stream(old3 + i, np);
      st_stream(h2 + i, s4);
      st_field(den + i, (R)(np.x * np.x + np.y * np.y));
    }
    __syncthreads();
    q.shift();
  }
}

// Hamming corrector + final value + errHe partial sums (steppc.f90:97-112),
// software-pipelined like the predictor.  pc holds the previous corrector's
// value; the predicted value is psi + c112 pc (see pc_predict25).
template <class R, class SH, bool PF>
__device__ __forceinline__ void pc_correct25(const Ham25& a, const Col& c, const c64* __restrict__ psi,
                                             const R* __restrict__ pot4,
                                             const c64* __restrict__ o3p,
                                             const c64* __restrict__ o2p,
                                             const c64* __restrict__ h1p,
                                             const c64* __restrict__ h2p,
                                             c64* __restrict__ pc, c64* __restrict__ psi_out,
                                             R* __restrict__ den, double* __restrict__ part,
                                             double dt3, double c9, double c112) {
  __shared__ c64 tile[SH::LX * SH::LY];
  __shared__ double sh[SH::N / 64];
  const Grid& g = a.g;
  const int z0 = c.bz * a.zc;
  const int z1 = min(z0 + a.zc, g.nz);
  const int xc = min(c.x, g.nx - 1), yc = min(c.y, g.ny - 1);
  ZQueue<c64, c64> q;
  q.init(psi, g, c, z0);
  PlaneRegs<SH, c64> pr;
  c64 qn, v3n, v2n, a1n, a2n, pcn;
  double pvn;
  auto prefetch = [&](int z) {
    qn = q.next(psi, g, c, z);
    plane_load<SH>(pr, psi, g, c, z);
    const int j = gidx(g, xc, yc, z);
    pvn = (double)ld_stream(pot4 + j);
    v3n = ld_stream(o3p + j);
    v2n = ld_stream(o2p + j);
    a1n = ld_stream(h1p + j);
    a2n = ld_stream(h2p + j);
    pcn = ld_stream(pc + j);
  };
  if constexpr (PF) {
    if (z0 < z1) prefetch(z0);
  }
  double e = 0.0;
  for (int z = z0; z < z1; ++z) {
    if constexpr (!PF) prefetch(z);
    q.q[12] = qn;
    plane_store<SH>(tile, pr, c, q.q[6]);
    const double pv = pvn;
    const c64 v3 = v3n, v2 = v2n, a1 = a1n, a2 = a2n, pci = pcn;
    __syncthreads();
    if constexpr (PF) {
      if (z + 1 < z1) prefetch(z + 1);
    }
    if (c.in) {
      const int i = gidx(g, c.x, c.y, z);
      const c64 s4 = ham25<SH, std::is_same<R, float>::value>(a, tile, q, c, pv, z);
      const c64 s5 = mk(0.125 * ((9.0 * v3.x - v2.x) + dt3 * ((s4.x + 2.0 * a1.x) - a2.x)),
                        0.125 * ((9.0 * v3.y - v2.y) + dt3 * ((s4.y + 2.0 * a1.y) - a2.y)));
      const c64 ps = q.centre();
      const c64 s1 = mk(ps.x + c112 * pci.x, ps.y + c112 * pci.y);
      const c64 npc = mk(s1.x - s5.x, s1.y - s5.y);
      st_stream(pc + i, npc);
      const c64 np = mk(s5.x + c9 * npc.x, s5.y + c9 * npc.y);
      st_stream(psi_out + i, np);
      st_field(den + i, (R)(np.x * np.x + np.y * np.y));
      e += hypot(npc.x, npc.y);
    }
    __syncthreads();
    q.shift();
  }
  const double s = block_sum25<SH>(e, sh, c.tid);
  if (c.tid == 0) part[c.blk] = s;
}

// RK-Gill stage (steprk.f90:56-74), software-pipelined like the predictor.
template <class R, class SH, bool PF>
__device__ __forceinline__ void rk_stage25(const Ham25& a, const Col& c, const c64* __restrict__ psi,
                                           const R* __restrict__ pot4, c64* __restrict__ qacc,
                                           c64* __restrict__ psi_out, R* __restrict__ den,
                                           c64* __restrict__ h1_out, double ca, double cb,
                                           double cc, double dt) {
  __shared__ c64 tile[SH::LX * SH::LY];
  const Grid& g = a.g;
  const int z0 = c.bz * a.zc;
  const int z1 = min(z0 + a.zc, g.nz);
  const int xc = min(c.x, g.nx - 1), yc = min(c.y, g.ny - 1);
  ZQueue<c64, c64> q;
  q.init(psi, g, c, z0);
  PlaneRegs<SH, c64> pr;
  c64 qn, qin;
  double pvn;
  auto prefetch = [&](int z) {
    qn = q.next(psi, g, c, z);
    plane_load<SH>(pr, psi, g, c, z);
    const int j = gidx(g, xc, yc, z);
    pvn = (double)ld_stream(pot4 + j);
    qin = ld_stream(qacc + j);
  };
  if constexpr (PF) {
    if (z0 < z1) prefetch(z0);
  }
  for (int z = z0; z < z1; ++z) {
    if constexpr (!PF) prefetch(z);
    q.q[12] = qn;
    plane_store<SH>(tile, pr, c, q.q[6]);
    const double pv = pvn;
    const c64 qi = qin;
    __syncthreads();
    if constexpr (PF) {
      if (z + 1 < z1) prefetch(z + 1);
    }
    if (c.in) {
      const int i = gidx(g, c.x, c.y, z);
      const c64 s4 = ham25<SH, std::is_same<R, float>::value>(a, tile, q, c, pv, z);
      const c64 s1 = mk(ca * (s4.x - cb * qi.x), ca * (s4.y - cb * qi.y));
      st_stream(qacc + i, mk((qi.x + 3.0 * s1.x) - cc * s4.x, (qi.y + 3.0 * s1.y) - cc * s4.y));
      if (h1_out) st_stream(h1_out + i, s4);
      const c64 ps = q.centre();
      const c64 np = mk(ps.x + dt * s1.x, ps.y + dt * s1.y);
      st_stream(psi_out + i, np);
      st_field(den + i, (R)(np.x * np.x + np.y * np.y));
    }
    __syncthreads();
    q.shift();
  }
}

// grad(den) with the first-derivative stencil (derden.f90:26-28),
// software-pipelined like the predictor.  The arithmetic type is the field
// type: FP64 in the FP64 path, FP32 (weights rounded to FP32) on the FP32
// potential side of the mixed path.
template <class R, class SH, bool PF>
__device__ __forceinline__ void grad3_25(const Grid& g, int zc, const Col& c, const R* __restrict__ f,
                                         const Stencil7<R, false>& wx, const Stencil7<R, false>& wy,
                                         const Stencil7<R, false>& wz,
                                         R* __restrict__ fx, R* __restrict__ fy,
                                         R* __restrict__ fz) {
  __shared__ R tile[SH::LX * SH::LY];
  const int z0 = c.bz * zc;
  const int z1 = min(z0 + zc, g.nz);
  ZQueue<R, R> q;
  q.init(f, g, c, z0);
  PlaneRegs<SH, R> pr;
  R qn = R(0);
  auto prefetch = [&](int z) {
    qn = q.next(f, g, c, z);
    plane_load<SH>(pr, f, g, c, z);
  };
  if constexpr (PF) {
    if (z0 < z1) prefetch(z0);
  }
  for (int z = z0; z < z1; ++z) {
    if constexpr (!PF) prefetch(z);
    q.q[12] = qn;
    plane_store<SH>(tile, pr, c, q.q[6]);
    __syncthreads();
    if constexpr (PF) {
      if (z + 1 < z1) prefetch(z + 1);
    }
    if (c.in) {
      const int i = gidx(g, c.x, c.y, z);
      constexpr bool FU = std::is_same<R, float>::value;
      st_field(fx + i, (R)tile_dx<SH, FU>(tile, c, wx));
      st_field(fy + i, (R)tile_dy<SH, FU>(tile, c, wy));
      st_field(fz + i, (R)q.template dz<FU>(wz));
    }
    __syncthreads();
    q.shift();
  }
}

}  // namespace hetd

namespace hetd {

// ---- alpha_s divergence + pot4 on the engine ---------------------------
// term_alfa.f90:85-113 and poten.f90:80-124: each product g_c = falfs*int_c
// is differentiated along its own axis only, so g_x needs an x-halo row
// tile, g_y a y-halo column tile and g_z the z queue.  A is the arithmetic
// type: FP64 in the FP64 path (the reference's operation order), FP32 on the
// FP32 potential side of the mixed path.
template <class A>
struct Pot25T {
  Grid g;
  Stencil7<A, false> d1x, d1y, d1z;
  A a0, a1;                 // cp4/2, cpp4/3
  A den0s, h2o2m4, alphas;
  A C, beta, den_m;
  int otc;
  int zc;
  int xcd;
};

// Solid penalty C (1 + t + beta rho (1 - t^2)), t = tanh(beta (rho - rho_m))
// (poten.f90:114-118).  FP64: the reference's form.  FP32: with
// s = 1 + t = 2 / (1 + exp(-2x)) and 1 - t^2 = s (2 - s), which keeps the
// relative accuracy of s where t is close to -1 (1 + t would lose it all to
// the rounding of t, and C is large).
__device__ inline double penalty(double dn, double C, double beta, double den_m) {
  const double t = tanh(beta * (dn - den_m));
  return C * (1.0 + t + beta * dn * (1.0 - t * t));
}
__device__ inline float penalty(float dn, float C, float beta, float den_m) {
  const float x = beta * (dn - den_m);
  const float sg = 2.0f / (1.0f + __expf(-2.0f * x));
  return C * (sg + beta * dn * (sg * (2.0f - sg)));
}

template <class R, class SH, bool PF>
__device__ __forceinline__ void pot_final25(const Pot25T<R>& a, const Col& c, const R* __restrict__ den,
                                            const R* __restrict__ dencg,
                                            const R* __restrict__ delj4,
                                            const R* __restrict__ sto1c,
                                            const R* __restrict__ uimp,
                                            const R* __restrict__ falfs,
                                            const R* __restrict__ intx,
                                            const R* __restrict__ inty,
                                            const R* __restrict__ intz,
                                            const R* __restrict__ uaconv, R* __restrict__ pot4) {
  using A = R;
  __shared__ A tx_tile[SH::TY * SH::LX];    // g_x, rows of this tile with x halo
  __shared__ A ty_tile[SH::LY * SH::TX];    // g_y, columns of this tile with y halo
  constexpr int NX = (SH::TY * SH::LX + SH::N - 1) / SH::N;
  constexpr int NY = (SH::LY * SH::TX + SH::N - 1) / SH::N;
  const Grid& g = a.g;
  const int z0 = c.bz * a.zc;
  const int z1 = min(z0 + a.zc, g.nz);
  const int xc = min(c.x, g.nx - 1), yc = min(c.y, g.ny - 1);
  A q[13];
  if (a.otc) {
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      const int z = z0 - 6 + k;
      const int j = gidx(g, xc, yc, min(max(z, 0), g.nz - 1));
      const A t = (A)falfs[j] * (A)intz[j];
      q[k] = (c.in && z >= 0 && z < g.nz) ? t : A(0);
    }
  }
  const int x0 = c.bx * SH::TX - 6, y0 = c.by * SH::TY - 6;
  // Software pipeline: the raw operands of plane z+1 (tile products, queue
  // entry, centre values) are loaded while plane z is computed; products
  // are formed when the plane is used, as (A)a * (A)b.
  R fxn[NX], ixn[NX], fyn[NY], iyn[NY], fzn = R(0), izn = R(0);
  bool okx[NX], oky[NY], okz = false;
  R dnn = R(0), dcgn = R(0), dljn = R(0), s1n = R(0), uin = R(0), uan = R(0);
  auto prefetch = [&](int z) {
    if (a.otc) {
      const int zz = z + 6;
      const int jz = gidx(g, xc, yc, min(zz, g.nz - 1));
      fzn = falfs[jz];
      izn = intz[jz];
      okz = c.in && zz < g.nz;
#pragma unroll
      for (int k = 0; k < NX; ++k) {
        const int e = c.tid + k * SH::N;
        const int ly = e / SH::LX, lx = e - ly * SH::LX;
        const int x = x0 + lx, y = c.by * SH::TY + ly;
        okx[k] = x >= 0 && x < g.nx && y < g.ny;
        const int j = gidx(g, min(max(x, 0), g.nx - 1), min(y, g.ny - 1), z);
        fxn[k] = falfs[j];
        ixn[k] = intx[j];
      }
#pragma unroll
      for (int k = 0; k < NY; ++k) {
        const int e = c.tid + k * SH::N;
        const int ly = e / SH::TX, lx = e - ly * SH::TX;
        const int x = c.bx * SH::TX + lx, y = y0 + ly;
        oky[k] = x < g.nx && y >= 0 && y < g.ny;
        const int j = gidx(g, min(x, g.nx - 1), min(max(y, 0), g.ny - 1), z);
        fyn[k] = falfs[j];
        iyn[k] = inty[j];
      }
      uan = ld_stream(uaconv + gidx(g, xc, yc, z));
    }
    const int i = gidx(g, xc, yc, z);
    dnn = ld_stream(den + i);
    dcgn = ld_stream(dencg + i);
    dljn = ld_stream(delj4 + i);
    s1n = ld_stream(sto1c + i);
    uin = ld_stream(uimp + i);
  };
  if constexpr (PF) {
    if (z0 < z1) prefetch(z0);
  }
  for (int z = z0; z < z1; ++z) {
    if constexpr (!PF) prefetch(z);
    if (a.otc) {
      q[12] = okz ? (A)fzn * (A)izn : A(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) {
        const int e = c.tid + k * SH::N;
        if (e < SH::TY * SH::LX) tx_tile[e] = okx[k] ? (A)fxn[k] * (A)ixn[k] : A(0);
      }
#pragma unroll
      for (int k = 0; k < NY; ++k) {
        const int e = c.tid + k * SH::N;
        if (e < SH::LY * SH::TX) ty_tile[e] = oky[k] ? (A)fyn[k] * (A)iyn[k] : A(0);
      }
    }
    const A dn = (A)dnn, dcg = (A)dcgn, dlj = (A)dljn, s1 = (A)s1n;
    const A ui = (A)uin, uac = (A)uan;
    __syncthreads();
    if constexpr (PF) {
      if (z + 1 < z1) prefetch(z + 1);
    }
    if (c.in) {
      const int i = gidx(g, c.x, c.y, z);
      A p4 = (dlj + dcg * dcg * (a.a0 + a.a1 * dcg)) + s1;
      if (a.otc) {
        A s4 = A(0), s5 = A(0), s6 = A(0);
        const A* rx = tx_tile + c.ty * SH::LX + c.tx;
        const A* cy = ty_tile + c.tx + c.ty * SH::TX;
        constexpr bool FU = std::is_same<R, float>::value;
#pragma unroll
        for (int m = 0; m < 13; ++m) s4 = madd<FU>(s4, a.d1x.at(m), rx[m]);
#pragma unroll
        for (int m = 0; m < 13; ++m) s5 = madd<FU>(s5, a.d1y.at(m), cy[m * SH::TX]);
#pragma unroll
        for (int m = 0; m < 13; ++m) s6 = madd<FU>(s6, a.d1z.at(m), q[m]);
        const A ua = uac / a.den0s;
        const A ual = (((ua + s4) + s5) + s6) * a.h2o2m4 * a.alphas;
        p4 = p4 + ual;
      }
      st_field(pot4 + i, (R)((p4 + ui) + penalty(dn, a.C, a.beta, a.den_m)));
    }
    __syncthreads();
    if (a.otc) {
#pragma unroll
      for (int k = 0; k < 12; ++k) q[k] = q[k + 1];
    }
  }
}

// ---- launches carrying one or several members -----------------------------
// Per-member pointers of each stencil kernel; nzb = z chunks per member.
// Every member computes exactly what a launch of its own computes.
template <class R>
struct PcpTab {
  int nzb;
  const c64* timec[kMaxMembers];
  const c64* psi[kMaxMembers];
  const R* pot4[kMaxMembers];
  c64* old3[kMaxMembers];
  const c64* h1[kMaxMembers];
  c64* h2[kMaxMembers];
  const c64* pc[kMaxMembers];
  R* den[kMaxMembers];
};
template <class R>
struct PccTab {
  int nzb;
  const c64* timec[kMaxMembers];
  const c64* psi[kMaxMembers];
  const R* pot4[kMaxMembers];
  const c64* o3[kMaxMembers];
  const c64* o2[kMaxMembers];
  const c64* h1[kMaxMembers];
  const c64* h2[kMaxMembers];
  c64* pc[kMaxMembers];
  c64* out[kMaxMembers];
  R* den[kMaxMembers];
  double* part[kMaxMembers];
};
template <class R>
struct RkTab {
  int nzb;
  const c64* timec[kMaxMembers];
  const c64* psi[kMaxMembers];
  const R* pot4[kMaxMembers];
  c64* q[kMaxMembers];
  c64* out[kMaxMembers];
  R* den[kMaxMembers];
  c64* h1[kMaxMembers];
};
template <class R>
struct GradTab {
  int nzb;
  const R* f[kMaxMembers];
  R* fx[kMaxMembers];
  R* fy[kMaxMembers];
  R* fz[kMaxMembers];
};
template <class R>
struct PotTab {
  int nzb;
  const R* den[kMaxMembers];
  const R* dencg[kMaxMembers];
  const R* delj4[kMaxMembers];
  const R* sto1c[kMaxMembers];
  const R* uimp[kMaxMembers];
  const R* falfs[kMaxMembers];
  const R* intx[kMaxMembers];
  const R* inty[kMaxMembers];
  const R* intz[kMaxMembers];
  const R* uaconv[kMaxMembers];
  R* pot4[kMaxMembers];
};

template <class R, class SH, bool PF = true>
__global__ void __launch_bounds__(SH::N) k_pc_predict25(Ham25 a, PcpTab<R> t, double c4o3dt, double c112) {
  const Col c = col_of<SH>(a.g, a.xcd, t.nzb);
  const int m = c.m;
  Ham25 am = a;
  am.timec_z = t.timec[m];
  pc_predict25<R, SH, PF>(am, c, t.psi[m], t.pot4[m], t.old3[m], t.h1[m], t.h2[m], t.pc[m], t.den[m], c4o3dt,
                          c112);
}

template <class R, class SH, bool PF = true>
__global__ void __launch_bounds__(SH::N) k_pc_correct25(Ham25 a, PccTab<R> t, double dt3, double c9,
                                                         double c112) {
  const Col c = col_of<SH>(a.g, a.xcd, t.nzb);
  const int m = c.m;
  Ham25 am = a;
  am.timec_z = t.timec[m];
  pc_correct25<R, SH, PF>(am, c, t.psi[m], t.pot4[m], t.o3[m], t.o2[m], t.h1[m], t.h2[m], t.pc[m], t.out[m],
                          t.den[m], t.part[m], dt3, c9, c112);
}

template <class R, class SH, bool PF = true>
__global__ void __launch_bounds__(SH::N) k_rk_stage25(Ham25 a, RkTab<R> t, double ca, double cb, double cc,
                                                           double dt) {
  const Col c = col_of<SH>(a.g, a.xcd, t.nzb);
  const int m = c.m;
  Ham25 am = a;
  am.timec_z = t.timec[m];
  rk_stage25<R, SH, PF>(am, c, t.psi[m], t.pot4[m], t.q[m], t.out[m], t.den[m], t.h1[m], ca, cb, cc, dt);
}

template <class R, class SH, bool PF = true>
__global__ void __launch_bounds__(SH::N) k_grad3_25(Grid g, int zc, int xcd, GradTab<R> t, Stencil7<R, false> wx,
                                                         Stencil7<R, false> wy, Stencil7<R, false> wz) {
  const Col c = col_of<SH>(g, xcd, t.nzb);
  const int m = c.m;
  grad3_25<R, SH, PF>(g, zc, c, t.f[m], wx, wy, wz, t.fx[m], t.fy[m], t.fz[m]);
}

template <class R, class SH, bool PF = true>
__global__ void __launch_bounds__(SH::N) k_pot_final25(Pot25T<R> a, PotTab<R> t) {
  const Col c = col_of<SH>(a.g, a.xcd, t.nzb);
  const int m = c.m;
  pot_final25<R, SH, PF>(a, c, t.den[m], t.dencg[m], t.delj4[m], t.sto1c[m], t.uimp[m], t.falfs[m], t.intx[m],
                         t.inty[m], t.intz[m], t.uaconv[m], t.pot4[m]);
}

}  // namespace hetd
