// Shared device/host definitions for the MI355X (gfx950) TDDFT time-step library.
//
// Field layout in HBM follows the reference's Fortran arrays byte for byte:
// a real field f(nx,ny,nz) is stored x-fastest, index ix + nx*(iy + ny*iz);
// a half spectrum (nx/2+1, ny, nz) likewise with M = nx/2+1 in place of nx
// (modules.f90, dimen.f90:29-143).  Complex values are interleaved (re, im).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <type_traits>

namespace hetd {

template <class T>
struct alignas(2 * sizeof(T)) cplx {
  T x, y;
};

using c64 = cplx<double>;
using c32 = cplx<float>;

template <class T>
__host__ __device__ inline cplx<T> mk(T a, T b) {
  cplx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <class T>
__host__ __device__ inline cplx<T> operator+(cplx<T> a, cplx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <class T>
__host__ __device__ inline cplx<T> operator-(cplx<T> a, cplx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <class T>
__host__ __device__ inline cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class T>
__host__ __device__ inline cplx<T> cmulc(cplx<T> a, cplx<T> b) {  // a * conj(b)
  return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <class T>
__host__ __device__ inline cplx<T> conj(cplx<T> a) { return mk(a.x, -a.y); }
template <class T>
__host__ __device__ inline cplx<T> scale(cplx<T> a, T s) { return mk(a.x * s, a.y * s); }

// Element at byte offset off from base.  With base the same for a whole
// wavefront (a kernel argument plus uniform terms, held in scalar
// registers) and off a 32-bit per-lane offset, the access compiles to one
// global load or store of the scalar-base + vector-offset form: no 64-bit
// address arithmetic per access, and one VGPR of address for every array
// indexed by the same offset.  Callers keep off below 4 GiB.
template <class P>
__device__ __forceinline__ P& el(P* base, unsigned off) {
  using B = typename std::conditional<std::is_const<P>::value, const char, char>::type;
  return *reinterpret_cast<P*>(reinterpret_cast<B*>(base) + off);
}

// Streaming access for values a kernel reads or writes once: the load or
// store carries the non-temporal hint, so the L2 keeps those lines at low
// priority and they do not push out lines that neighbouring workgroups read
// again (the halo rows of the stencil kernels).
__device__ __forceinline__ double ld_stream(const double* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float ld_stream(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ c64 ld_stream(const c64* p) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  const d2 v = __builtin_nontemporal_load(reinterpret_cast<const d2*>(p));
  return mk(v.x, v.y);
}
__device__ __forceinline__ c32 ld_stream(const c32* p) {
  typedef float f2 __attribute__((ext_vector_type(2)));
  const f2 v = __builtin_nontemporal_load(reinterpret_cast<const f2*>(p));
  return mk(v.x, v.y);
}
__device__ __forceinline__ void st_stream(double* p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_stream(float* p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_stream(c64* p, c64 v) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  d2 w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<d2*>(p));
}

// Maximum FFT plan depth (radix stages) and supported line length.
constexpr int kMaxStages = 8;
constexpr int kMaxN = 256;

// Radix plan of one 1D transform length, passed to kernels by value.
struct FftPlan {
  int n;                      // transform length
  int nstages;                // number of radix stages
  int radix[kMaxStages];      // radix of each stage (2,3,4,5,8)
};

// 13-point stencil weights for offsets -6..6, already scaled by 1/h^order.
struct Stencil13 {
  double w[13];
};
// The same weights rounded to FP32, for the stencils of the mixed-precision
// potential side (FP32 fields, FP32 arithmetic).
struct Stencil13f {
  float w[13];
};
template <class A>
struct StencilOf {
  using type = Stencil13;
};
template <>
struct StencilOf<float> {
  using type = Stencil13f;
};
template <class A>
using StencilA = typename StencilOf<A>::type;

// A 13-point stencil stored by its centre symmetry: u[0] is the centre
// weight, u[j] the weight of offset +j, and offset -j has u[j] (SYM, the
// second derivative) or -u[j] (antisymmetric, the first derivative).  The
// first- and second-derivative tables are exactly (anti)symmetric, and
// (-u)*f = -(u*f) exactly, so sums over at(m) equal sums over the full table
// bit for bit.  7 values instead of 13 keep three axes' weights in scalar
// registers.
template <class A, bool SYM>
struct Stencil7 {
  A u[7];
  __host__ __device__ constexpr A at(int m) const {  // offset m - 6, m = 0..12
    return m > 6 ? u[m - 6] : (m < 6 ? (SYM ? u[6 - m] : -u[6 - m]) : u[0]);
  }
};
template <class A, bool SYM>
inline Stencil7<A, SYM> compact(const Stencil13& s) {
  Stencil7<A, SYM> r;
  r.u[0] = (A)s.w[6];
  for (int j = 1; j <= 6; ++j) r.u[j] = (A)s.w[6 + j];
  return r;
}

// Weight m (offset m - 6) of any stencil form.
__host__ __device__ inline double wat(const Stencil13& s, int m) { return s.w[m]; }
template <class A, bool SYM>
__host__ __device__ inline A wat(const Stencil7<A, SYM>& s, int m) {
  return s.at(m);
}

inline Stencil13f to_f32(const Stencil13& s) {
  Stencil13f r;
  for (int m = 0; m < 13; ++m) r.w[m] = (float)s.w[m];
  return r;
}
template <class A>
inline StencilA<A> stencil_as(const Stencil13& s);
template <>
inline Stencil13 stencil_as<double>(const Stencil13& s) {
  return s;
}
template <>
inline Stencil13f stencil_as<float>(const Stencil13& s) {
  return to_f32(s);
}

struct Grid {
  int nx, ny, nz;
  int mx;           // nx/2 + 1
  int mp;           // padded spectrum row length (>= mx, 128-byte multiple)
  long long n;      // nx*ny*nz
  long long h;      // mp*ny*nz (padded half spectrum)
};

// Instances of one grid that a launch carries at most (the member tables of
// the kernels hold one pointer per instance).
constexpr int kMaxMembers = 8;

}  // namespace hetd
