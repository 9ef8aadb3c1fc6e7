// mcw_device.h -- device-side building blocks of the transport kernels.
//
// One segment of a history is the body of MaCaw's CollisionKernel::onSegment
// (src/raykernels/CollisionKernel.C:98-235) fused with the MOOSE TraceRay step
// that produced it: exit side of the current element, one collision-distance
// draw, track-length tally, then cross / reflect / leak or collide.
//
// Arithmetic contract (what makes GPU histories bit-identical to a
// history-based FP64 CPU run of the same algorithm): compiled with
// -ffp-contract=off, every expression is written with explicit association,
// elementary functions below use only IEEE +,-,*,/ and sqrt, and integer
// RNG arithmetic is exact.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mcw_types.h"

namespace mcw {

// ------------------------------------------------------------------ bits

// First plane slot of element e (internal id); NS is the mesh's side count
// (4, 6) or 0 for a mixed mesh.
template <int NS> __device__ __forceinline__ int64_t slot_of(const MeshView &M, int64_t e)
{
  if (NS > 0) return e * NS;
  return e < M.n_tet ? 4 * e : 4 * M.n_tet + 8 * (e - M.n_tet);
}

// Atomics on device memory through explicit global (address space 1)
// pointers.  A pointer read from memory is generic to the compiler, which
// then emits FLAT atomics; the tally pointers of the request server are.
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) unsigned long long gu64;
__device__ __forceinline__ void atomic_add_dev(double *p, double v)
{
  (void)__builtin_amdgcn_global_atomic_fadd_f64((gdouble *)p, v);
}
__device__ __forceinline__ void atomic_add_dev(unsigned long long *p, unsigned long long v)
{
  (void)__hip_atomic_fetch_add((gu64 *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Atomic add of v at p for the lanes with `on`.  Lanes that share an address
// add their sum in one atomic: the first pending lane's address is matched
// against the others, repeatedly while a match joins at least two lanes (at
// most 4 groups); the remaining lanes add one by one.  Lanes that start at one
// source point share their first elements, and same-address atomics
// serialise.  Call with the whole wave.
__device__ __forceinline__ double wave_sum_dbl(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ void atomic_add_agg(bool on, double *p, double v)
{
  const unsigned long long a = (unsigned long long)p;
  bool todo = on;
  for (int g = 0; g < 4; ++g) {
    const unsigned long long m = __ballot(todo);
    if (m == 0ull) return;
    const int leader = __builtin_ctzll(m);
    const unsigned long long al = (unsigned long long)__shfl((long long)a, leader, 64);
    const bool same = todo && a == al;
    if (__popcll(__ballot(same)) < 2) break;
    const double s = wave_sum_dbl(same ? v : 0.0);
    if (__lane_id() == leader) atomic_add_dev(p, s);
    if (same) todo = false;
  }
  if (todo) atomic_add_dev(p, v);
}

__device__ __forceinline__ uint64_t dbits(double x) { return (uint64_t)__double_as_longlong(x); }
__device__ __forceinline__ double dfrom(uint64_t u) { return __longlong_as_double((long long)u); }

// ------------------------------------------------------------------ elementary functions
//
// log: fdlibm 5.3 e_log.c method (Sun Microsystems, 1993): x = 2^k (1+f),
// s = f/(2+f), log(1+f) = f - (hfsq - s (hfsq + R(s^2))), R a degree-14
// minimax polynomial.  sin/cos(2 pi xi): exact quadrant reduction of 4 xi and
// fdlibm's __kernel_sin / __kernel_cos polynomials on [-pi/4, pi/4].

__device__ __forceinline__ double log_d(double x)
{
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  const double two54 = 1.80143985094819840000e+16;
  const double Lg1 = 6.666666666666735130e-01;
  const double Lg2 = 3.999999999940941908e-01;
  const double Lg3 = 2.857142874366239149e-01;
  const double Lg4 = 2.222219843214978396e-01;
  const double Lg5 = 1.818357216161805012e-01;
  const double Lg6 = 1.531383769920937332e-01;
  const double Lg7 = 1.479819860511658591e-01;

  uint64_t u = dbits(x);
  int32_t hx = (int32_t)(u >> 32);
  uint32_t lx = (uint32_t)u;
  int32_t k = 0;
  if (hx < 0x00100000) {
    if (((hx & 0x7fffffff) | lx) == 0) return -two54 / 0.0;
    if (hx < 0) return (x - x) / 0.0;
    k -= 54;
    x *= two54;
    u = dbits(x);
    hx = (int32_t)(u >> 32);
  }
  if (hx >= 0x7ff00000) return x + x;
  k += (hx >> 20) - 1023;
  hx &= 0x000fffff;
  int32_t i = (hx + 0x95f64) & 0x100000;
  u = ((uint64_t)(uint32_t)(hx | (i ^ 0x3ff00000)) << 32) | (u & 0xffffffffull);
  x = dfrom(u);
  k += (i >> 20);
  double f = x - 1.0;
  double dk = (double)k;
  if ((0x000fffff & (2 + hx)) < 3) {
    if (f == 0.0) {
      if (k == 0) return 0.0;
      return dk * ln2_hi + dk * ln2_lo;
    }
    double R = f * f * (0.5 - 0.33333333333333333 * f);
    if (k == 0) return f - R;
    return dk * ln2_hi - ((R - dk * ln2_lo) - f);
  }
  double s = f / (2.0 + f);
  double z = s * s;
  i = hx - 0x6147a;
  double w = z * z;
  int32_t j = 0x6b851 - hx;
  double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
  double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
  i |= j;
  double R = t2 + t1;
  if (i > 0) {
    double hfsq = 0.5 * f * f;
    if (k == 0) return f - (hfsq - s * (hfsq + R));
    return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
  }
  if (k == 0) return f - s * (f - R);
  return dk * ln2_hi - ((s * (f - R) - dk * ln2_lo) - f);
}

__device__ __forceinline__ double ksin_d(double x)
{
  const double S1 = -1.66666666666666324348e-01;
  const double S2 = 8.33333333332248946124e-03;
  const double S3 = -1.98412698298579493134e-04;
  const double S4 = 2.75573137070700676789e-06;
  const double S5 = -2.50507602534068634195e-08;
  const double S6 = 1.58969099521155010221e-10;
  double z = x * x;
  double v = z * x;
  double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
  return x + v * (S1 + z * r);
}

__device__ __forceinline__ double kcos_d(double x)
{
  const double C1 = 4.16666666666666019037e-02;
  const double C2 = -1.38888888888741095749e-03;
  const double C3 = 2.48015872894767294178e-05;
  const double C4 = -2.75573143513906633035e-07;
  const double C5 = 2.08757232129817482790e-09;
  const double C6 = -1.13596475577881948265e-11;
  int32_t ix = (int32_t)(dbits(x) >> 32) & 0x7fffffff;
  double z = x * x;
  double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
  if (ix < 0x3FD33333) return 1.0 - (0.5 * z - z * r);
  double qx;
  if (ix > 0x3fe90000) qx = 0.28125;
  else qx = dfrom((uint64_t)(uint32_t)(ix - 0x00200000) << 32);
  double hz = 0.5 * z - qx;
  double a = 1.0 - qx;
  return a - (hz - z * r);
}

__device__ __forceinline__ void sincos_2pi_d(double xi, double *s, double *c)
{
  const double pio2 = 1.57079632679489661923;
  double t = 4.0 * xi;
  double q = (double)(int64_t)(t + 0.5);
  double r = t - q;
  double a = r * pio2;
  double sa = ksin_d(a), ca = kcos_d(a);
  int iq = ((int)q) & 3;
  if (iq == 0) { *s = sa; *c = ca; }
  else if (iq == 1) { *s = ca; *c = -sa; }
  else if (iq == 2) { *s = -sa; *c = -ca; }
  else { *s = -ca; *c = sa; }
}

__device__ __forceinline__ double cos_halfpi_d(double xi)
{
  double s, c;
  sincos_2pi_d(0.25 * xi, &s, &c);
  return c;
}

// ------------------------------------------------------------------ OpenMC 0.13 RNG
// 64-bit LCG (M = 6364136223846793005, C = 1442695040888963407), PCG-RXS-M-XS
// output permutation (O'Neill 2014), unit interval by 2^-64; skip-ahead by
// Brown (1994).  Streams: 0 tracking, 1 source, 2 URR.

constexpr uint64_t PRN_MULT = 6364136223846793005ull;
constexpr uint64_t PRN_ADD = 1442695040888963407ull;
constexpr uint64_t PRN_STRIDE = 152917ull;

__device__ __forceinline__ double prn(uint64_t &s)
{
  s = PRN_MULT * s + PRN_ADD;
  uint64_t word = ((s >> ((s >> 59u) + 5u)) ^ s) * 12605985483714917081ull;
  uint64_t out = (word >> 43u) ^ word;
  return (double)out * 0x1p-64;
}

// n steps of the LCG are the affine map s -> g s + c (mod 2^64).  Brown's
// skip-ahead squares the one-step map once per bit of n and composes the
// squares where n has a one.  The squares do not depend on n, so they are
// fixed at compile time: LCG_POW2[k] is the map of 2^k steps, and a jump costs
// two 64-bit products per set bit of n instead of four per bit.
struct LcgMap {
  uint64_t g, c;
};
struct LcgPow2 {
  LcgMap m[64];
};
constexpr LcgPow2 lcg_pow2_table()
{
  LcgPow2 t{};
  uint64_t g = PRN_MULT, c = PRN_ADD;
  for (int k = 0; k < 64; ++k) {
    t.m[k] = LcgMap{g, c};
    c = (g + 1) * c;
    g *= g;
  }
  return t;
}
static __constant__ LcgPow2 LCG_POW2 = lcg_pow2_table();

// The map of n steps; the maps are powers of one map, so they commute.
__device__ __forceinline__ LcgMap lcg_jump(uint64_t n)
{
  uint64_t g = 1, c = 0;
  for (int k = 0; n != 0; ++k, n >>= 1) {
    if (n & 1u) {
      const LcgMap p = LCG_POW2.m[k];
      c = c * p.g + p.c;
      g *= p.g;
    }
  }
  return LcgMap{g, c};
}

__device__ __forceinline__ uint64_t future_seed(uint64_t n, uint64_t seed)
{
  const LcgMap j = lcg_jump(n);
  return j.g * seed + j.c;
}

__device__ __forceinline__ uint64_t init_seed(int64_t id, int stream, uint64_t master)
{
  return future_seed((uint64_t)id * PRN_STRIDE, master + (uint64_t)stream);
}

// The tracking (stream 0) and source (stream 1) seeds of history id: one jump
// of id * PRN_STRIDE steps applied to both stream seeds.
__device__ __forceinline__ void init_seeds01(int64_t id, uint64_t master, uint64_t &s0, uint64_t &s1)
{
  const LcgMap j = lcg_jump((uint64_t)id * PRN_STRIDE);
  s0 = j.g * master + j.c;
  s1 = j.g * (master + 1) + j.c;
}

// MaCaw stores seeds in Real ray aux data (CollisionKernel.C:132-134,157-159).
__device__ __forceinline__ uint64_t seed_roundtrip(uint64_t s)
{
  double d = (double)s;
  if (d >= 18446744073709551616.0) return 0;
  return (uint64_t)d;
}

__device__ __forceinline__ uint64_t digest_step(uint64_t h, int64_t elem, int ev)
{
  h ^= (uint64_t)(elem * 8 + ev);
  return h * 0x100000001b3ull;
}

// ------------------------------------------------------------------ cross sections
//
// Pointwise nuclides on a unionized grid U (all nuclides' points merged).
// union_index finds k = largest index with U[k] <= E: a log-spaced bucket
// table (staged in LDS by the kernels) narrows the range to a few entries,
// then a binary search in U finishes.  For nuclide n, uidx[k][n] is then the
// bracketing point of its own grid (no point of any grid lies strictly
// between U[k] and U[k+1]), so the interpolation equals OpenMC's per-nuclide
// lookup: largest j with E_j <= E, clamped to [0, n-2], lin-lin.

struct Micro {
  double t, a, s, f;
};

// -1 when E < U[0].
__device__ __forceinline__ int32_t union_index(const XSView &X, const int32_t *hash, double E)
{
  const int32_t n = X.n_union;
  const double *__restrict__ U = X.U;
  if (E < U[0]) return -1;
  if (E >= U[n - 1]) return n - 1;
  double pos = (log_d(E) - X.logU0) * X.inv_dl;
  int32_t b = pos > 0.0 ? (int32_t)pos : 0;
  if (b > X.n_hash - 1) b = X.n_hash - 1;
  const int32_t b0 = b > 0 ? b - 1 : 0;
  const int32_t b1 = b + 2 < X.n_hash ? b + 2 : X.n_hash;
  int32_t lo = hash[b0];
  int32_t hi = hash[b1] + 1;
  if (hi > n - 1) hi = n - 1;
  // U[lo] <= E < U[hi]
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (U[mid] <= E) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void micro_at(const XSView &X, int32_t nuc, int32_t k, double E, Micro &m)
{
  const int64_t base = X.grid_off[nuc];
  const int32_t np = X.n_points[nuc];
  int32_t j = k < 0 ? -1 : X.uidx[(int64_t)k * X.n_nuc + nuc];
  double f;
  if (j < 0) {
    j = 0;
    f = 0.0;
  } else if (j >= np - 1) {
    j = np - 2;
    f = 1.0;
  } else {
    f = -1.0; // set below from the bracketing energies
  }
  const double4 r0 = X.rec[base + j];
  const double4 r1 = X.rec[base + j + 1];
  if (f < 0.0) f = (E - r0.x) / (r1.x - r0.x);
  const double g = 1.0 - f;
  m.t = g * r0.y + f * r1.y;
  m.a = g * r0.z + f * r1.z;
  m.f = g * r0.w + f * r1.w;
  m.s = g * X.scatter[base + j] + f * X.scatter[base + j + 1];
}

__device__ __forceinline__ double macro_total(const XSView &X, int mat, int32_t k, double E)
{
  if (mat < 0) return 0.0;
  double t = 0.0;
  const int32_t b = X.mat_off[mat], e = X.mat_off[mat + 1];
  for (int32_t q = b; q < e; ++q) {
    Micro u;
    micro_at(X, X.mat_nuc[q], k, E, u);
    t = t + X.mat_dens[q] * u.t;
  }
  return t;
}

// Macroscopic total and nu-fission at (mat, E) in one pass over the material.
__device__ __forceinline__ double macro_total_nf(const XSView &X, int mat, int32_t k, double E, double &nusf)
{
  nusf = 0.0;
  if (mat < 0) return 0.0;
  double t = 0.0, nf = 0.0;
  const int32_t b = X.mat_off[mat], e = X.mat_off[mat + 1];
  for (int32_t q = b; q < e; ++q) {
    const int32_t n = X.mat_nuc[q];
    Micro u;
    micro_at(X, n, k, E, u);
    t = t + X.mat_dens[q] * u.t;
    if (X.fis) {
      const double4 fp = X.fis[n];
      nf = nf + X.mat_dens[q] * ((fp.x + fp.y * E) * u.f);
    }
  }
  nusf = nf;
  return t;
}

constexpr int XS_GATHER = 3; // materials of up to this many nuclides take the gathered lookups

// A gathered lookup for a material of at most XS_GATHER nuclides: its loads
// run in rounds for all nuclides together -- ids and densities, then grid
// offsets and union-grid indices -- and only the bracketing records are read
// nuclide by nuclide, so a lookup waits for two rounds plus one per nuclide
// instead of four per nuclide.  The arithmetic is micro_at's and the sums
// run in material order: values equal the per-nuclide loops' to the bit.
struct MatIdx {
  int32_t n;
  int32_t nuc[XS_GATHER], np[XS_GATHER], j[XS_GATHER];
  double dn[XS_GATHER];
  int64_t lo[XS_GATHER];
};

__device__ __forceinline__ void mat_index(const XSView &X, int32_t b, int32_t n, int32_t k, MatIdx &I)
{
  I.n = n;
#pragma unroll
  for (int i = 0; i < XS_GATHER; ++i) {
    const int32_t q = b + (i < n ? i : 0); // past n: a repeat of the first, unused
    I.nuc[i] = X.mat_nuc[q];
    I.dn[i] = X.mat_dens[q];
  }
#pragma unroll
  for (int i = 0; i < XS_GATHER; ++i) {
    I.lo[i] = X.grid_off[I.nuc[i]];
    I.np[i] = X.n_points[I.nuc[i]];
    I.j[i] = k < 0 ? -1 : X.uidx[(int64_t)k * X.n_nuc + I.nuc[i]];
  }
}

// Micro cross sections of slot i of a gathered lookup (scatter only with S).
template <bool S>
__device__ __forceinline__ void micro_slot(const XSView &X, const MatIdx &I, int i, double E, Micro &m)
{
  int32_t j = I.j[i];
  double f;
  if (j < 0) {
    j = 0;
    f = 0.0;
  } else if (j >= I.np[i] - 1) {
    j = I.np[i] - 2;
    f = 1.0;
  } else {
    f = -1.0; // set below from the bracketing energies
  }
  const int64_t base = I.lo[i];
  const double4 r0 = X.rec[base + j];
  const double4 r1 = X.rec[base + j + 1];
  if (f < 0.0) f = (E - r0.x) / (r1.x - r0.x);
  const double g = 1.0 - f;
  m.t = g * r0.y + f * r1.y;
  m.a = g * r0.z + f * r1.z;
  m.f = g * r0.w + f * r1.w;
  m.s = S ? g * X.scatter[base + j] + f * X.scatter[base + j + 1] : 0.0;
}

// macro_total (and with NF, macro_total_nf) with a gathered lookup when the
// material is small enough.
template <bool NF>
__device__ __forceinline__ double macro_total_g(const XSView &X, int mat, int32_t k, double E, double &nusf)
{
  nusf = 0.0;
  if (mat < 0) return 0.0;
  const int32_t b = X.mat_off[mat], n = X.mat_off[mat + 1] - b;
  if (n > XS_GATHER) return NF ? macro_total_nf(X, mat, k, E, nusf) : macro_total(X, mat, k, E);
  MatIdx I;
  mat_index(X, b, n, k, I);
  double t = 0.0, nf = 0.0;
#pragma unroll
  for (int i = 0; i < XS_GATHER; ++i) {
    if (i < n) {
      Micro u;
      micro_slot<false>(X, I, i, E, u);
      t = t + I.dn[i] * u.t;
      if (NF && X.fis) {
        const double4 fp = X.fis[I.nuc[i]];
        nf = nf + I.dn[i] * ((fp.x + fp.y * E) * u.f);
      }
    }
  }
  nusf = nf;
  return t;
}

struct Macro {
  double t, a, s, f, kf; // kf: kappa-fission, sum of N kappa sigma_f
};

__device__ __forceinline__ void macro_all(const XSView &X, int mat, int32_t k, double E, Macro &M)
{
  M.t = M.a = M.s = M.f = M.kf = 0.0;
  if (mat < 0) return;
  const int32_t b = X.mat_off[mat], e = X.mat_off[mat + 1];
  for (int32_t q = b; q < e; ++q) {
    Micro u;
    const int32_t n = X.mat_nuc[q];
    micro_at(X, n, k, E, u);
    double dn = X.mat_dens[q];
    M.t = M.t + dn * u.t;
    M.a = M.a + dn * u.a;
    M.s = M.s + dn * u.s;
    M.f = M.f + dn * u.f;
    M.kf = M.kf + dn * (X.kappa[n] * u.f);
  }
}

// macro_all with a gathered lookup when the material is small enough.
__device__ __forceinline__ void macro_all_g(const XSView &X, int mat, int32_t k, double E, Macro &M)
{
  M.t = M.a = M.s = M.f = M.kf = 0.0;
  if (mat < 0) return;
  const int32_t b = X.mat_off[mat], n = X.mat_off[mat + 1] - b;
  if (n > XS_GATHER) {
    macro_all(X, mat, k, E, M);
    return;
  }
  MatIdx I;
  mat_index(X, b, n, k, I);
#pragma unroll
  for (int i = 0; i < XS_GATHER; ++i) {
    if (i < n) {
      Micro u;
      micro_slot<true>(X, I, i, E, u);
      const double dn = I.dn[i];
      M.t = M.t + dn * u.t;
      M.a = M.a + dn * u.a;
      M.s = M.s + dn * u.s;
      M.f = M.f + dn * u.f;
      M.kf = M.kf + dn * (X.kappa[I.nuc[i]] * u.f);
    }
  }
}

// ------------------------------------------------------------------ geometry

// Exit side of element `e` for a ray at p along u: smallest non-negative
// distance over sides whose outward normal has n.u > 0, the incoming side
// excluded, ties to the lowest side index.  Returns -1 when no side qualifies.
template <int NS>
__device__ __forceinline__ int exit_side(const double4 *__restrict__ pl, double x, double y, double z,
                                         double u, double v, double w, int inface, double &L)
{
  L = __builtin_inf();
  int fx = -1;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double4 q = pl[s];
    double den = (q.x * u + q.y * v) + q.z * w;
    bool ok = (s != inface) && (den > 0.0);
    double num = q.w - ((q.x * x + q.y * y) + q.z * z);
    double t = num / den;
    if (!(t > 0.0)) t = 0.0;
    if (ok && t < L) {
      L = t;
      fx = s;
    }
  }
  return fx;
}

// FP32 exit side from NSD FP32 planes, the position given relative to the
// planes' origin (dx, dy, dz) with S = |dx| + |dy| + |dz|.  Each candidate
// distance t carries an error bound e (evaluation and rounding of FP32
// operands: |dt| <= (16u(|c| + S) + 16u|t|)/den + 4u|t|, u = 2^-24; the
// quotients use the hardware reciprocal, within 1 ulp, so t = num * rcp(den)
// is within 3u|t| of num / den, where an IEEE division was within u).  When
// the best candidate's interval overlaps another's, or a side is within 1e-5
// of grazing, `recheck` is set and the caller must redo the selection in
// FP64; otherwise the chosen side is the one FP64 would choose, and L is the
// FP32 distance.
template <int NSD>
__device__ __forceinline__ int exit_side_f32(const float4 *__restrict__ pl, float dx, float dy, float dz, float S,
                                             double u, double v, double w, int inface, double &L, bool &recheck)
{
  const float uu = 0x1p-24f;
  const float ux = (float)u, uy = (float)v, uz = (float)w;
  float t[NSD], e[NSD];
  bool ok[NSD];
  bool graze = false;
#pragma unroll
  for (int s = 0; s < NSD; ++s) {
    const float4 q = pl[s];
    const float den = (q.x * ux + q.y * uy) + q.z * uz;
    const float num = q.w - ((q.x * dx + q.y * dy) + q.z * dz);
    const float rd = __builtin_amdgcn_rcpf(den);
    float ts = num * rd;
    ts = ts > 0.0f ? ts : 0.0f;
    ok[s] = (s != inface) && (den > 0.0f);
    graze |= (s != inface) && (fabsf(den) <= 1e-5f);
    t[s] = ts;
    e[s] = (16.0f * uu * (fabsf(q.w) + S) + 16.0f * uu * ts) * rd + 4.0f * uu * ts;
  }
  int fx = -1;
  float best = __builtin_inff();
#pragma unroll
  for (int s = 0; s < NSD; ++s)
    if (ok[s] && t[s] < best) {
      best = t[s];
      fx = s;
    }
  bool overlap = false;
  if (fx >= 0) {
    const float hi = best + e[fx];
#pragma unroll
    for (int s = 0; s < NSD; ++s)
      if (ok[s] && s != fx && t[s] - e[s] <= hi) overlap = true;
  }
  recheck = graze || overlap || fx < 0;
  L = (double)best;
  return fx;
}

// FP32 exit side of a TET4 (planes relative to its vertex v0).
__device__ __forceinline__ int exit_side_p32(const MixedTet &R, double x, double y, double z, double u, double v,
                                             double w, int inface, double &L, bool &recheck)
{
  const float dx = (float)(x - R.v0[0]), dy = (float)(y - R.v0[1]), dz = (float)(z - R.v0[2]);
  const float S = fabsf(dx) + fabsf(dy) + fabsf(dz);
  return exit_side_f32<4>(R.pl, dx, dy, dz, S, u, v, w, inface, L, recheck);
}

// FP32 exit side of a HEX8 (absolute planes).
__device__ __forceinline__ int exit_side_p32(const MixedHex &R, double x, double y, double z, double u, double v,
                                             double w, int inface, double &L, bool &recheck)
{
  const float dx = (float)x, dy = (float)y, dz = (float)z;
  const float S = fabsf(dx) + fabsf(dy) + fabsf(dz);
  return exit_side_f32<6>(R.pl, dx, dy, dz, S, u, v, w, inface, L, recheck);
}

// OpenMC rotate_angle with phi = 2 pi xi.
__device__ __forceinline__ void rotate_angle(double &u, double &v, double &w, double mu, double xi)
{
  double sinphi, cosphi;
  sincos_2pi_d(xi, &sinphi, &cosphi);
  double u0 = u, v0 = v, w0 = w;
  double a = sqrt(fmax(0.0, 1.0 - mu * mu));
  double b = sqrt(fmax(0.0, 1.0 - w0 * w0));
  if (b > 1e-10) {
    u = mu * u0 + a * (u0 * w0 * cosphi - v0 * sinphi) / b;
    v = mu * v0 + a * (v0 * w0 * cosphi + u0 * sinphi) / b;
    w = mu * w0 - a * b * cosphi;
  } else {
    b = sqrt(1.0 - v0 * v0);
    u = mu * u0 + a * (u0 * v0 * cosphi + w0 * sinphi) / b;
    v = mu * v0 - a * b * cosphi;
    w = mu * w0 + a * (v0 * w0 * cosphi - u0 * sinphi) / b;
  }
}

// ------------------------------------------------------------------ tallies

__device__ __forceinline__ void tally_score(const TallyView &T, const XSView &X, int64_t elem, double E,
                                            int32_t k, double wfl, int mat)
{
  int64_t fi;
  if (T.elem_base) {
    fi = T.elem_base[elem];
    if (fi < 0) return;
  } else {
    fi = elem;
  }
  if (T.n_edges > 0) {
    const double *ed = T.edges;
    const int ne = T.n_edges;
    if (E < ed[0] || E > ed[ne - 1]) return;
    int lo = 0, hi = ne - 1; // ed[lo] <= E, bin = largest lo <= ne-2
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (ed[mid] <= E) lo = mid;
      else hi = mid;
    }
    fi += (int64_t)lo * T.energy_stride;
  }
  double *row = T.value + fi * T.n_score_bins;
  if (T.simple_flux) {
    atomic_add_dev(row, wfl);
    return;
  }
  const int nsc = T.n_scores;
  for (int ni = 0; ni < T.n_nuclides; ++ni) {
    const int nuc = T.nuclides[ni];
    Macro M = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (nuc < 0) {
      if (T.needs_xs) macro_all_g(X, mat, k, E, M);
    } else if (mat >= 0) {
      // a nuclide bin scores nothing, flux included, in a material without
      // that nuclide (OpenMC 0.13 score_tracklength_tally /
      // score_collision_tally skip the bin)
      int32_t q = X.mat_off[mat];
      const int32_t qe = X.mat_off[mat + 1];
      while (q < qe && X.nuc_of[X.mat_nuc[q]] != nuc) ++q;
      if (q == qe) continue;
      if (T.needs_xs) {
        Micro mu;
        micro_at(X, X.mat_nuc[q], k, E, mu);
        const double dn = X.mat_dens[q];
        M.t = dn * mu.t;
        M.a = dn * mu.a;
        M.s = dn * mu.s;
        M.f = dn * mu.f;
        M.kf = dn * (X.kappa[X.mat_nuc[q]] * mu.f);
      }
    }
    for (int si = 0; si < nsc; ++si) {
      double m;
      switch (T.scores[si]) {
        case 0: m = 1.0; break;
        case 1: m = M.t; break;
        case 2: m = M.a; break;
        case 3: m = M.s; break;
        case 4: m = M.f; break;
        default: m = M.kf; break;
      }
      // scores are never negative and VALUE starts at +0, so adding an
      // exact zero changes nothing: skip the atomic (a kappa-fission
      // score outside the fuel, for instance)
      const double v = wfl * m;
      if (v != 0.0) atomic_add_dev(row + ni * nsc + si, v);
    }
  }
}

} // namespace mcw
