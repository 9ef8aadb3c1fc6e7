// mcw_kernels.hip -- gfx950 transport kernels.
//
// Source sites for a chunk of histories are sampled first (k_source_bank:
// one thread per history, coalesced structure-of-arrays writes).  Two
// schedules then transport them with the same per-history arithmetic:
//
//  PLAIN (event-based, host-driven).  The live particles sit compacted in a
//  structure-of-arrays bank.  Each cycle k_segment advances every live
//  particle by one segment (exit side, distance draw, track-length tally,
//  cross / reflect / leak) and appends it to the next live bank or to the
//  collision bank; k_collide then collides the latter (sample nuclide,
//  absorb or scatter, collision tally, new union-grid index) and appends the
//  survivors to the next live bank.  Appends are wavefront-aggregated
//  (ballot + prefix: one atomic per wavefront), so reads and writes stay
//  coalesced.
//
//  BATCHED (persistent, fused).  k_persistent keeps particles in registers.
//  A wavefront steps segments for its lanes while at least `lanes_min` of
//  them are in flight; lanes that reach a collision or die wait, then the
//  wavefront runs the collisions together and refills idle lanes together
//  from the source bank.  Expensive events (collisions, refills) are thus
//  paid once per batch of lanes instead of once per segment.  The unionized
//  energy-grid bucket table is staged in LDS once per workgroup.
//
// Both reproduce, per history, the FP64 CPU reference to the bit.
#include <hip/hip_runtime.h>

#include "mcw_device.h"
#include "mcw_types.h"

namespace mcw {

enum { EV_CROSS = 0, EV_REFLECT = 1, EV_LEAK = 2, EV_SCATTER = 3, EV_ABSORB = 4, EV_CUTOFF = 5,
       EV_LOST = 6, EV_COLLIDE = 7 };
enum { FATE_ALIVE = 0, FATE_LEAKED = 1, FATE_ABSORBED = 2, FATE_CUTOFF = 3, FATE_LOST = 4 };

struct PState {
  double x, y, z, u, v, w, E, sigt;
  uint64_t seed;
  int64_t hid;
  uint64_t h;
  int32_t elem, inface, mat, k;
  int32_t nseg, ncoll;
  double *tv;   // request server: the request's flux tally
  int32_t req;  // request server: slice of the round
  double x0, y0, z0; // record mode: where the current element visit began
  int32_t visit;     // record mode: index of the current visit
  // eigenvalue batches: macroscopic nu-fission (cached with sigt), the
  // history's keff scores and its fission-site count
  double nusf, ktl, kcol, kabs;
  int32_t prog;
};

// Event counts of one wavefront.  Every field is wavefront-uniform (sums of
// ballots), so the compiler keeps them in scalar registers.
struct WaveCount {
  unsigned long long seg, coll, absn, leak, cut, lost, refl, hex;
  __device__ void zero() { seg = coll = absn = leak = cut = lost = refl = hex = 0; }
};

// ------------------------------------------------------------------ wave helpers

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += (unsigned long long)__shfl_xor((long long)v, off, 64);
  return v;
}

__device__ __forceinline__ double wave_sum_d(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ unsigned long long wave_pop(bool p) { return (unsigned long long)__popcll(__ballot(p)); }

// Count the outcome of one segment step (every lane of the wavefront calls it).
__device__ __forceinline__ void count_segment(WaveCount &w, bool stepped, int ev, bool hex = false)
{
  w.seg += wave_pop(stepped && ev != 6); // a history lost before its segment adds none
  w.hex += wave_pop(stepped && ev != 6 && hex);
  w.leak += wave_pop(stepped && ev == 2);
  w.refl += wave_pop(stepped && ev == 1);
  w.lost += wave_pop(stepped && ev == 6);
}

// Count the outcome of one collision step (every lane calls it).
__device__ __forceinline__ void count_collision(WaveCount &w, bool collided, int ev)
{
  w.coll += wave_pop(collided);
  w.absn += wave_pop(collided && ev == 4);
  w.cut += wave_pop(collided && ev == 5);
}

__device__ void flush_counts(Counters *c, const WaveCount &l, double track)
{
  const double tr = wave_sum_d(track);
  if (__lane_id() == 0) {
    if (l.seg) atomicAdd(&c->segments, l.seg);
    if (l.coll) atomicAdd(&c->collisions, l.coll);
    if (l.absn) atomicAdd(&c->absorptions, l.absn);
    if (l.leak) atomicAdd(&c->leaks, l.leak);
    if (l.cut) atomicAdd(&c->cutoffs, l.cut);
    if (l.lost) atomicAdd(&c->lost, l.lost);
    if (l.refl) atomicAdd(&c->reflections, l.refl);
    if (l.hex) atomicAdd(&c->hex_segments, l.hex);
    if (tr != 0.0) unsafeAtomicAdd(&c->track, tr);
  }
}

__device__ __forceinline__ int wave_count(bool p) { return __popcll(__ballot(p)); }

// A value every lane of the wavefront holds alike, kept in scalar registers.
__device__ __forceinline__ int64_t wave_uniform(int64_t v)
{
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Workgroup-aggregated append to two queues at once (256 threads, 4
// wavefronts; every thread must call it): one atomic per queue per
// workgroup instead of one per wavefront.  Appends from every wavefront of
// the event kernels land on the same two counters, and same-address atomics
// serialise.  Slots are consecutive in thread order within the workgroup.
__device__ __forceinline__ void block_append2(unsigned long long *count_a, bool want_a, int64_t &slot_a,
                                              unsigned long long *count_b, bool want_b, int64_t &slot_b)
{
  __shared__ unsigned s_cnt[2][4];
  __shared__ unsigned long long s_base[2];
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  const unsigned long long ma = __ballot(want_a), mb = __ballot(want_b);
  if (lane == 0) {
    s_cnt[0][wave] = (unsigned)__popcll(ma);
    s_cnt[1][wave] = (unsigned)__popcll(mb);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned ta = s_cnt[0][0] + s_cnt[0][1] + s_cnt[0][2] + s_cnt[0][3];
    const unsigned tb = s_cnt[1][0] + s_cnt[1][1] + s_cnt[1][2] + s_cnt[1][3];
    s_base[0] = ta && count_a ? atomicAdd(count_a, (unsigned long long)ta) : 0ull;
    s_base[1] = tb && count_b ? atomicAdd(count_b, (unsigned long long)tb) : 0ull;
  }
  __syncthreads();
  unsigned pa = 0, pb = 0;
  for (int w = 0; w < wave; ++w) {
    pa += s_cnt[0][w];
    pb += s_cnt[1][w];
  }
  const unsigned long long below = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
  slot_a = want_a ? (int64_t)(s_base[0] + pa + (unsigned)__popcll(ma & below)) : -1;
  slot_b = want_b ? (int64_t)(s_base[1] + pb + (unsigned)__popcll(mb & below)) : -1;
  __syncthreads(); // s_cnt and s_base are reused by the next call
}

// ------------------------------------------------------------------ source

__device__ void sample_point_in_tet(const double *p0, const double *p1, const double *p2, const double *p3,
                                    uint64_t &seed, double *out)
{
  // Rocchini & Cignoni, Journal of Graphics Tools 5(4), 2000.
  double s = prn(seed), t = prn(seed), u = prn(seed);
  if (s + t > 1.0) {
    s = 1.0 - s;
    t = 1.0 - t;
  }
  if (t + u > 1.0) {
    double tmp = u;
    u = 1.0 - s - t;
    t = 1.0 - tmp;
  } else if (s + t + u > 1.0) {
    double tmp = u;
    u = s + t + u - 1.0;
    s = 1.0 - t - tmp;
  }
  double a = 1.0 - s - t - u;
#pragma unroll
  for (int k = 0; k < 3; ++k) out[k] = ((a * p0[k] + s * p1[k]) + t * p2[k]) + u * p3[k];
}

__constant__ int8_t c_kuhn[6][4] = {{0, 1, 2, 6}, {0, 2, 3, 6}, {0, 3, 7, 6},
                                    {0, 7, 4, 6}, {0, 4, 5, 6}, {0, 5, 1, 6}};

__device__ double tet_volume6_d(const double *a, const double *b, const double *c, const double *d)
{
  double ax = b[0] - a[0], ay = b[1] - a[1], az = b[2] - a[2];
  double bx = c[0] - a[0], by = c[1] - a[1], bz = c[2] - a[2];
  double cx = d[0] - a[0], cy = d[1] - a[1], cz = d[2] - a[2];
  double det = (ax * (by * cz - bz * cy) - ay * (bx * cz - bz * cx)) + az * (bx * cy - by * cx);
  return fabs(det);
}

// Lowest caller-index element whose side planes contain p within tol, or -1
// (the rule of the host-side locate and of the CPU reference).
__device__ int32_t locate_point_d(const MeshView &M, const double *p, double tol)
{
  int32_t cell[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double f = (p[k] - M.grid_lo[k]) * M.grid_inv[k];
    f = f < (double)(M.grid_n[k] - 1) ? f : (double)(M.grid_n[k] - 1);
    cell[k] = f > 0.0 ? (int32_t)f : 0; // NaN and below the box: cell 0
  }
  const int64_t ci = (int64_t)cell[0] + (int64_t)M.grid_n[0] * ((int64_t)cell[1] + (int64_t)M.grid_n[1] * cell[2]);
  int32_t best = -1, best_orig = 0x7fffffff;
  for (int32_t q = M.cell_start[ci]; q < M.cell_start[ci + 1]; ++q) {
    const int32_t e = M.cell_elem[q];
    int64_t s0;
    int ns;
    if (M.stride > 0) {
      s0 = (int64_t)e * M.stride;
      ns = M.stride;
    } else {
      s0 = slot_of<0>(M, e);
      ns = M.erec[2 * (int64_t)e + 1].w;
    }
    bool in = true;
    for (int s = 0; s < ns && in; ++s) {
      const double4 pl = M.plane[s0 + s];
      const double v = ((pl.x * p[0] + pl.y * p[1]) + pl.z * p[2]) - pl.w;
      if (v > tol) in = false;
    }
    if (in) {
      const int32_t o = M.orig_of_int[e];
      if (o < best_orig) {
        best_orig = o;
        best = e;
      }
    }
  }
  return best;
}

// OpenMC fixed-source initialize_history (OpenMCStudy.C:317-363 drives it):
// seeds for particle id (batch-1)*n + i + 1, then space, angle, energy
// sampled on stream 1 (sample_external_source).
__device__ void sample_site(const Params &P, const SourceView &S, uint64_t master, int64_t batch, int64_t n_particles,
                            int64_t hid, double *p, double *dir, double &E, uint64_t &track_seed, int32_t &elem_out)
{
  const int64_t id = (batch - 1) * n_particles + hid + 1;
  uint64_t ss;
  init_seeds01(id, master, track_seed, ss);
  const MeshView &M = P.mesh;
  int64_t elem;
  if (S.space == 0) {
    p[0] = S.point[0];
    p[1] = S.point[1];
    p[2] = S.point[2];
    elem = S.point_elem;
  } else if (S.space == 1) {
    // uniform in the box, resampled while outside the mesh
    int32_t e = -1;
    for (int it = 0; it < 1000 && e < 0; ++it) {
#pragma unroll
      for (int k = 0; k < 3; ++k) p[k] = S.box_lo[k] + prn(ss) * (S.box_hi[k] - S.box_lo[k]);
      e = locate_point_d(M, p, 1e-12);
      // OpenMC only_fissionable: a site outside a fissionable material is rejected too
      if (e >= 0 && S.only_fissionable) {
        const int32_t mt = M.erec[2 * (int64_t)e + 1].z;
        if (mt < 0 || !P.xs.mat_fissionable[mt]) e = -1;
      }
    }
    elem = e; // internal id, -1 when the box misses the mesh
  } else {
    const int64_t ne = M.n_elem;
    const double target = prn(ss) * M.cdf[ne - 1];
    // first e with cdf[e] > target.  The guide table brackets it: bucket j
    // of G holds targets in [j, j+1) * total / G and guide[j] is the first e
    // with cdf[e] > j * total / G, so the answer lies in [guide[j-1],
    // guide[j+2]] even when rounding moves j by one.  A binary search
    // finishes; if the two neighbours of its result show that the bracket
    // missed, the search runs again over the whole CDF.
    auto first_above = [&](int64_t lo, int64_t hi) {
      while (lo < hi) {
        const int64_t mid = lo + ((hi - lo) >> 1);
        if (M.cdf[mid] > target) hi = mid;
        else lo = mid + 1;
      }
      return lo;
    };
    int64_t lo = 0, hi = ne;
    if (M.n_guide > 0) {
      const int64_t G = M.n_guide;
      int64_t j = (int64_t)(target * M.cdf_guide_scale);
      j = j < 0 ? 0 : (j > G ? G : j);
      lo = M.cdf_guide[j > 0 ? j - 1 : 0];
      hi = M.cdf_guide[j + 2 <= G ? j + 2 : G];
    }
    int64_t e = first_above(lo, hi);
    if (!((e == 0 || !(M.cdf[e - 1] > target)) && (e == ne || M.cdf[e] > target))) e = first_above(0, ne);
    elem = e < ne ? e : ne - 1;
    // all-TET4 meshes: connectivity at a fixed stride, no offset or type load
    const bool tet = M.stride == 4 || M.type[elem] == 4;
    const int32_t *c = M.stride == 4 ? M.conn + (M.conn_off0 + 4 * elem) : M.conn + M.conn_off[elem];
    const double *X = M.xyz;
    if (tet) {
      sample_point_in_tet(X + 3 * (int64_t)c[0], X + 3 * (int64_t)c[1], X + 3 * (int64_t)c[2],
                          X + 3 * (int64_t)c[3], ss, p);
    } else {
      double vols[6], tot = 0.0;
      for (int k = 0; k < 6; ++k) {
        tot = tot + tet_volume6_d(X + 3 * (int64_t)c[c_kuhn[k][0]], X + 3 * (int64_t)c[c_kuhn[k][1]],
                                  X + 3 * (int64_t)c[c_kuhn[k][2]], X + 3 * (int64_t)c[c_kuhn[k][3]]);
        vols[k] = tot;
      }
      double tt = prn(ss) * tot;
      int k = 0;
      while (k < 5 && !(vols[k] > tt)) ++k;
      sample_point_in_tet(X + 3 * (int64_t)c[c_kuhn[k][0]], X + 3 * (int64_t)c[c_kuhn[k][1]],
                          X + 3 * (int64_t)c[c_kuhn[k][2]], X + 3 * (int64_t)c[c_kuhn[k][3]], ss, p);
    }
  }
  if (S.angle == 0) {
    double xi_phi = prn(ss);
    double mu = 2.0 * prn(ss) - 1.0;
    double s, c;
    sincos_2pi_d(xi_phi, &s, &c);
    double r = sqrt(1.0 - mu * mu);
    dir[0] = mu;
    dir[1] = r * c;
    dir[2] = r * s;
  } else {
    // p(mu) ~ 1 + beta mu about the axis: mu = -2c / (1 + sqrt(1 - 2 beta c)),
    // c = (1 - beta/2) - 2 xi, the root of F(mu) = xi without cancellation;
    // azimuth 2 pi xi' about the axis
    const double cq = (1.0 - 0.5 * S.beta) - 2.0 * prn(ss);
    double mu = -2.0 * cq / (1.0 + sqrt(1.0 - 2.0 * S.beta * cq));
    if (mu > 1.0) mu = 1.0;
    if (mu < -1.0) mu = -1.0;
    dir[0] = S.axis[0];
    dir[1] = S.axis[1];
    dir[2] = S.axis[2];
    rotate_angle(dir[0], dir[1], dir[2], mu, prn(ss));
  }
  if (S.energy == 0) {
    E = S.E0;
  } else {
    const double a = S.watt_a, b = S.watt_b;
    for (int it = 0;; ++it) {
      double r1 = prn(ss), r2 = prn(ss), r3 = prn(ss);
      double c = cos_halfpi_d(r3);
      double wmax = -a * (log_d(r1) + log_d(r2) * c * c);
      double Ew = wmax + 0.25 * a * a * b + (-1.0 + 2.0 * prn(ss)) * sqrt(a * a * b * wmax);
      if (Ew < P.xs.e_max || it >= 1000) {
        E = Ew < P.xs.e_max ? Ew : P.xs.e_min;
        break;
      }
    }
  }
  if (P.roundtrip) track_seed = seed_roundtrip(track_seed);
  elem_out = S.space == 2 ? M.int_of_orig[elem] : (int32_t)elem;
}

__device__ __forceinline__ void start_history(PState &st, int64_t hid)
{
  st.inface = -1;
  st.mat = -2; // forces Sigma_t at the first segment
  st.sigt = 0.0;
  st.hid = hid;
  st.h = 0xcbf29ce484222325ull;
  st.nseg = 0;
  st.ncoll = 0;
  st.x0 = st.x;
  st.y0 = st.y;
  st.z0 = st.z;
  st.visit = 0;
  st.nusf = 0.0;
  st.ktl = 0.0;
  st.kcol = 0.0;
  st.kabs = 0.0;
  st.prog = 0;
}

// Record mode: close the current element visit of a traced history with
// event `ev` (the visit was in internal element `e`) and start the next one
// at the current position.
template <bool REC> __device__ __forceinline__ void trace_visit(const Params &P, PState &st, int32_t e, int ev)
{
  if (!REC || st.hid >= P.trace_n) return;
  const unsigned long long k = atomicAdd(&P.cnt->trace_count, 1ull);
  if ((int64_t)k < P.trace_cap) {
    TraceRec &R = P.trace[k];
    R.x0 = st.x0;
    R.y0 = st.y0;
    R.z0 = st.z0;
    R.x1 = st.x;
    R.y1 = st.y;
    R.z1 = st.z;
    R.E = st.E;
    R.seed = st.seed;
    R.hid = st.hid;
    R.elem = e < 0 ? -1 : P.mesh.orig_of_int[e];
    R.visit = st.visit;
    R.event = ev;
    R.pad = 0;
  }
  ++st.visit;
  st.x0 = st.x;
  st.y0 = st.y;
  st.z0 = st.z;
}

// ------------------------------------------------------------------ one segment

__device__ __forceinline__ int64_t orig_elem(const MeshView &M, int32_t e) { return e < 0 ? -1 : M.orig_of_int[e]; }

__device__ __forceinline__ int32_t pick_nbr(const int4 &r0, const int4 &r1, int fx)
{
  // masked OR over the six codes: keeps the record in registers (a select
  // chain is turned into an indexed stack array by the compiler)
  const int32_t m0 = -(int32_t)(fx == 0), m1 = -(int32_t)(fx == 1), m2 = -(int32_t)(fx == 2);
  const int32_t m3 = -(int32_t)(fx == 3), m4 = -(int32_t)(fx == 4), m5 = -(int32_t)(fx == 5);
  return (r0.x & m0) | (r0.y & m1) | (r0.z & m2) | (r0.w & m3) | (r1.x & m4) | (r1.y & m5);
}

// Geometry and distance draw of one segment.  Returns the event that ends it;
// EV_COLLIDE leaves the particle at the collision point for collide().
// P32 (TET4 only): exit side from the FP32 mixed record, FP64 recheck of
// near ties (counted in `rechecks`).
// With `tl_ptr`, a plain per-element flux (the request server's, or the one
// track-length tally when P.tl_fast) is not added here but returned as the
// address (nullptr: no bin) and value, for the caller's wave-aggregated add.
template <int NS, bool REC, bool P32, bool SERVE = false, bool EIG = false>
__device__ __forceinline__ int segment(const Params &P, PState &st, double &track_sum, unsigned &rechecks,
                                       double **tl_ptr = nullptr, double *tl_val = nullptr, bool *in_hex = nullptr)
{
  const MeshView &M = P.mesh;
  if (st.elem < 0) return EV_LOST; // source site outside the mesh
  if (st.nseg >= P.max_seg) {
    if (REC) st.h = digest_step(st.h, orig_elem(M, st.elem), EV_LOST);
    trace_visit<REC>(P, st, st.elem, EV_LOST);
    return EV_LOST;
  }
  int4 r0, r1;
  int32_t mat;
  double L;
  int fx;
  int64_t slot0;
  bool fp32 = false;
  if (P32) {
    // FP32 record of a TET4 (every element of a TET4 mesh; in a mixed mesh
    // the ids below n_tet) or of a HEX8 of a mixed mesh
    if (NS == 4 || st.elem < M.n_tet) {
      const MixedTet &R = M.mtet[st.elem];
      fp32 = true;
      const int4 code = *(const int4 *)R.code;
      r0 = code;
      r1 = make_int4(-1, -1, R.mat, 4);
      mat = R.mat;
      slot0 = (int64_t)st.elem * 4; // TET4 are first in a mixed mesh
      bool recheck;
      fx = exit_side_p32(R, st.x, st.y, st.z, st.u, st.v, st.w, st.inface, L, recheck);
      if (recheck) {
        fx = exit_side<4>(M.plane + slot0, st.x, st.y, st.z, st.u, st.v, st.w, st.inface, L);
        ++rechecks;
      }
    } else if (NS == 0) {
      const MixedHex &R = M.mhex[st.elem - M.n_tet];
      fp32 = true;
      if (in_hex) *in_hex = true;
      r0 = *(const int4 *)R.code;
      const int2 c45 = *(const int2 *)(R.code + 4);
      r1 = make_int4(c45.x, c45.y, R.mat, 6);
      mat = R.mat;
      slot0 = slot_of<0>(M, st.elem);
      bool recheck;
      fx = exit_side_p32(R, st.x, st.y, st.z, st.u, st.v, st.w, st.inface, L, recheck);
      if (recheck) {
        fx = exit_side<6>(M.plane + slot0, st.x, st.y, st.z, st.u, st.v, st.w, st.inface, L);
        ++rechecks;
      }
    }
  }
  if (!fp32) {
    if (NS == 0 && st.elem >= M.n_tet) {
      // HEX8 of a mixed mesh: its neighbour record sits in slot 6 of its
      // plane block (one cache line fewer per segment)
      slot0 = slot_of<0>(M, st.elem);
      const int4 *er = (const int4 *)(M.plane + slot0 + 6);
      r0 = er[0];
      r1 = er[1];
      if (in_hex) *in_hex = true;
      mat = r1.z;
      fx = exit_side<6>(M.plane + slot0, st.x, st.y, st.z, st.u, st.v, st.w, st.inface, L);
    } else {
      r0 = M.erec[2 * (int64_t)st.elem];
      r1 = M.erec[2 * (int64_t)st.elem + 1];
      mat = r1.z;
      slot0 = (int64_t)st.elem * (NS > 0 ? NS : 4); // TET4 are first in a mixed mesh
      if (NS == 6) fx = exit_side<6>(M.plane + slot0, st.x, st.y, st.z, st.u, st.v, st.w, st.inface, L);
      else fx = exit_side<4>(M.plane + slot0, st.x, st.y, st.z, st.u, st.v, st.w, st.inface, L);
    }
  }
  if (mat != st.mat) {
    if (EIG) st.sigt = macro_total_g<true>(P.xs, mat, st.k, st.E, st.nusf);
    else {
      double unused;
      st.sigt = macro_total_g<false>(P.xs, mat, st.k, st.E, unused);
    }
    st.mat = mat;
  }
  if (fx < 0) {
    if (REC) st.h = digest_step(st.h, orig_elem(M, st.elem), EV_LOST);
    trace_visit<REC>(P, st, st.elem, EV_LOST);
    return EV_LOST;
  }
  const double xi = prn(st.seed);
  const double d = st.sigt > 0.0 ? -log_d(xi) / st.sigt : __builtin_inf();
  const double track = (d < L) ? d : L;
  ++st.nseg;
  track_sum += track;
  // keff track-length estimator, weight 1 (CollisionKernel.C:152-154)
  if (EIG) st.ktl = st.ktl + track * st.nusf;
  if (SERVE || P.tl_fast) {
    double *bin;
    if (SERVE) {
      bin = st.tv + st.elem;
    } else {
      const int64_t fi = P.tally[0].elem_base ? P.tally[0].elem_base[st.elem] : st.elem;
      bin = fi >= 0 ? P.tally[0].value + fi : nullptr;
    }
    if (tl_ptr) {
      *tl_ptr = bin;
      *tl_val = track;
    } else if (bin) {
      atomic_add_dev(bin, track);
    }
  } else {
    for (int t = 0; t < P.n_tl; ++t) tally_score(P.tally[t], P.xs, st.elem, st.E, st.k, track, mat);
  }
  if (d > L) {
    const int32_t e_visit = st.elem;
    st.x = st.x + L * st.u;
    st.y = st.y + L * st.v;
    st.z = st.z + L * st.w;
    int32_t nb = pick_nbr(r0, r1, fx);
    if (NS == 0 && nb >= 0 && (nb & 7) == 7) {
      // split interface: the tet on the exit point's side of the diagonal
      const SplitRec &R = M.split[nb >> 3];
      const double sg = ((R.pl.x * st.x + R.pl.y * st.y) + R.pl.z * st.z) - R.pl.w;
      nb = sg <= 0.0 ? R.code[0] : R.code[1];
    }
    int ev;
    if (nb >= 0) {
      ev = EV_CROSS;
      if (REC) st.h = digest_step(st.h, orig_elem(M, st.elem), ev);
      st.elem = nb >> 3;
      st.inface = nb & 7;
    } else {
      const int32_t bid = -1 - nb;
      const int kind = bid < M.n_bc ? M.bc_kind[bid] : 0;
      if (kind == 1) {
        ev = EV_REFLECT;
        if (REC) st.h = digest_step(st.h, orig_elem(M, st.elem), ev);
        const double4 q = M.plane[slot0 + fx];
        double dot = (st.u * q.x + st.v * q.y) + st.w * q.z;
        st.u = st.u - 2.0 * dot * q.x;
        st.v = st.v - 2.0 * dot * q.y;
        st.w = st.w - 2.0 * dot * q.z;
        st.inface = fx;
      } else {
        ev = EV_LEAK;
        if (REC) st.h = digest_step(st.h, orig_elem(M, st.elem), ev);
      }
    }
    if (P.roundtrip) st.seed = seed_roundtrip(st.seed);
    trace_visit<REC>(P, st, e_visit, ev);
    return ev;
  }
  // collision: back-step from the exit point (CollisionKernel.C:172-175)
  const double back = L - d;
  st.x = (st.x + L * st.u) - back * st.u;
  st.y = (st.y + L * st.v) - back * st.v;
  st.z = (st.z + L * st.w) - back * st.w;
  return EV_COLLIDE;
}

// Collision at the current point, in two parts around the fission bank's
// slot reservation (which the whole wavefront makes together):
//   collide_pick: collision-estimator tallies (and the keff collision
//     estimator), sample the nuclide (OpenMC sample_nuclide) and, in an
//     eigenvalue batch with a fissionable nuclide, the number of fission
//     neutrons (create_fission_sites: floor(nu_t) plus one with probability
//     frac(nu_t), nu_t = w / keff * nu sigma_f / sigma_t);
//   collide_finish: bank those sites (isotropic: mu, then phi; a prompt /
//     delayed draw, always prompt; a Watt energy), then absorb (analog, with
//     the keff absorption estimator) or scatter elastically (isotropic in CM,
//     target at rest) with an energy cutoff at the grid minimum.  A scatter
//     looks up the new energy's union-grid index through `hash`.
// Dynamic LDS of the persistent kernel: small flux tallies summed per
// workgroup (P.lds_tally track-length bins, then P.lds_coll collision bins).
extern __shared__ double s_tally[];

struct CollPick {
  int32_t nuc;
  double t, a, nuf; // picked nuclide's total, absorption and nu-fission at E
};

template <bool REC, bool EIG>
__device__ __forceinline__ int64_t collide_pick(const Params &P, PState &st, CollPick &cp)
{
  const XSView &X = P.xs;
  const int32_t mat = st.mat;
  if (REC) ++st.ncoll;
  if (P.lds_coll) { // the persistent kernel's LDS copy of a small collision flux tally
    const TallyView &T = P.tally[P.n_tl];
    const int64_t fi = T.elem_base ? T.elem_base[st.elem] : st.elem;
    if (fi >= 0) atomicAdd(s_tally + P.lds_tally + fi, 1.0 / st.sigt);
  } else {
    for (int t = P.n_tl; t < P.n_tl + P.n_coll; ++t)
      tally_score(P.tally[t], X, st.elem, st.E, st.k, 1.0 / st.sigt, mat);
  }
  if (EIG) st.kcol = st.kcol + st.nusf / st.sigt;
  const int32_t b = X.mat_off[mat], e = X.mat_off[mat + 1];
  const double cutoff = prn(st.seed) * st.sigt;
  double prob = 0.0;
  int32_t pick = e - 1;
  Micro mu;
  if (e - b <= XS_GATHER) {
    // small material: a gathered lookup, the same sampling
    MatIdx I;
    mat_index(X, b, e - b, st.k, I);
    bool done = false;
#pragma unroll
    for (int i = 0; i < XS_GATHER; ++i) {
      if (!done && b + i < e) {
        micro_slot<false>(X, I, i, st.E, mu);
        prob = prob + I.dn[i] * mu.t;
        if (prob >= cutoff) {
          pick = b + i;
          done = true;
        }
      }
    }
  } else {
    for (int32_t q = b; q < e; ++q) {
      micro_at(X, X.mat_nuc[q], st.k, st.E, mu);
      prob = prob + X.mat_dens[q] * mu.t;
      if (prob >= cutoff) {
        pick = q;
        break;
      }
    }
  }
  cp.nuc = X.mat_nuc[pick];
  cp.t = mu.t;
  cp.a = mu.a;
  cp.nuf = 0.0;
  int64_t nu = 0;
  if (EIG) {
    if (X.fis) {
      const double4 fp = X.fis[cp.nuc];
      cp.nuf = (fp.x + fp.y * st.E) * mu.f;
    }
    if (X.fissionable && X.fissionable[cp.nuc] && mu.t > 0.0) {
      const double nu_t = ((1.0 / P.fb.keff) * cp.nuf) / mu.t;
      nu = nu_t < 9.0e15 ? (int64_t)nu_t : (int64_t)9.0e15;
      if (prn(st.seed) <= nu_t - (double)nu) ++nu;
      // a history's progeny index has PROGENY_BITS bits: every reserved slot
      // below the bank's capacity is then written
      const int64_t room = ((int64_t)1 << PROGENY_BITS) - st.prog;
      if (nu > room) nu = room;
    }
  }
  return nu;
}

// One fission neutron (OpenMC sample_fission_neutron) from tracking-stream
// state `s`: mu = 2 xi - 1, then phi = 2 pi xi (isotropic), a prompt /
// delayed draw (no delayed data: always prompt), then a Watt energy
// resampled while at or above the grid's top (after 1000 tries the grid
// minimum).  Advances `s` past the site's draws.
__device__ __forceinline__ void sample_fission_site(const XSView &X, int32_t nuc, uint64_t &s, double &u, double &v,
                                                    double &w, double &E)
{
  const double4 fp = X.fis[nuc];
  const double fmu = 2.0 * prn(s) - 1.0;
  double sphi, cphi;
  sincos_2pi_d(prn(s), &sphi, &cphi);
  const double sn = sqrt(1.0 - fmu * fmu);
  u = fmu;
  v = sn * cphi;
  w = sn * sphi;
  (void)prn(s);
  const double a = fp.z, bw = fp.w;
  for (int it = 0;; ++it) {
    const double r1 = prn(s), r2 = prn(s), r3 = prn(s);
    const double c = cos_halfpi_d(r3);
    const double wmax = -a * (log_d(r1) + log_d(r2) * c * c);
    const double Ew = wmax + 0.25 * a * a * bw + (-1.0 + 2.0 * prn(s)) * sqrt(a * a * bw * wmax);
    if (Ew < X.e_max || it >= 1000) {
      E = Ew < X.e_max ? Ew : X.e_min;
      return;
    }
  }
}

__device__ __forceinline__ void lcg_skip(uint64_t &s) { s = PRN_MULT * s + PRN_ADD; }

// Fission sites go to slots base, base + 1, ...; from the first slot past the
// bank's capacity on, the collision's sites are not sampled, as OpenMC stops
// at a full bank.  A site's directions and energy are not computed here: the
// slot keeps the stream state before its draws and the nuclide, and
// k_fission_sites evaluates them after the batch, for all sites at once.
// Here only the stream advances past the site's draws: three, then a Watt
// attempt of four.  When r1 r2 > watt_t the first attempt is certainly below
// the grid's top (Ew <= (sqrt(w) + a sqrt(b)/2)^2 with w <= -a ln(r1 r2)), so
// the site takes exactly seven draws; otherwise (about 1e-5 of sites) the site
// is sampled in full to find where its draws end.
template <bool REC, bool EIG>
__device__ __forceinline__ int collide_finish(const Params &P, PState &st, const CollPick &cp, int64_t nu,
                                              int64_t base, const int32_t *hash, int32_t *written = nullptr)
{
  const XSView &X = P.xs;
  if (EIG) {
    const FissionView &F = P.fb;
    for (int64_t q = 0; q < nu; ++q) {
      const int64_t slot = base + q;
      if (slot >= F.cap) break;
      ++*written;
      const uint64_t s0 = st.seed;
      uint64_t s = s0;
      lcg_skip(s); // mu
      lcg_skip(s); // phi
      lcg_skip(s); // prompt or delayed
      const double r1 = prn(s), r2 = prn(s);
      if (r1 * r2 > X.watt_t[cp.nuc]) {
        lcg_skip(s); // r3
        lcg_skip(s); // r4
        st.seed = s;
      } else {
        double u, v, w, E;
        st.seed = s0;
        sample_fission_site(X, cp.nuc, st.seed, u, v, w, E);
      }
      F.x[slot] = st.x;
      F.y[slot] = st.y;
      F.z[slot] = st.z;
      F.elem[slot] = st.elem;
      F.key[slot] = ((uint64_t)st.hid << PROGENY_BITS) | (uint64_t)st.prog;
      F.seed[slot] = s0;
      F.nuc[slot] = cp.nuc;
      ++st.prog;
    }
  }
  int ev;
  if (cp.a > prn(st.seed) * cp.t) {
    ev = EV_ABSORB;
    if (EIG) st.kabs = st.kabs + cp.nuf / cp.a;
  } else {
    const double A = X.awr[cp.nuc];
    const double mu_cm = 2.0 * prn(st.seed) - 1.0;
    const double q2 = (A * A + 2.0 * A * mu_cm) + 1.0;
    const double Ep = st.E * q2 / ((A + 1.0) * (A + 1.0));
    if (Ep < X.e_min) {
      ev = EV_CUTOFF;
      st.E = Ep;
    } else {
      const double mu_lab = (1.0 + A * mu_cm) / sqrt(q2);
      const double xi_phi = prn(st.seed);
      rotate_angle(st.u, st.v, st.w, mu_lab, xi_phi);
      st.E = Ep;
      st.k = union_index(X, hash, Ep);
      st.mat = -2; // energy changed: recompute Sigma_t
      ev = EV_SCATTER;
    }
  }
  if (REC) st.h = digest_step(st.h, orig_elem(P.mesh, st.elem), ev);
  st.inface = -1;
  if (P.roundtrip) st.seed = seed_roundtrip(st.seed);
  if (ev != EV_SCATTER) trace_visit<REC>(P, st, st.elem, ev);
  return ev;
}

// A whole collision of a fixed-source history.
template <bool REC>
__device__ __forceinline__ int collide(const Params &P, PState &st, const int32_t *hash)
{
  CollPick cp;
  collide_pick<REC, false>(P, st, cp);
  return collide_finish<REC, false>(P, st, cp, 0, 0, hash);
}

// Slots for n items per lane (n <= 2^24, so the wavefront's sum fits 32
// bits), consecutive in lane order, from the wavefront's current chunk of
// fission-bank slots [next, end); when the chunk is short, the wavefront
// takes a new one of max(need, FB_CHUNK) slots with one atomic (the rest of
// the old chunk stays empty).  A single counter taken once per collision
// phase by every wavefront serialises: same-address atomics cost about as
// much as a third of the batch.  Returns this lane's first slot.  Every lane
// must call it; next and end are wavefront-uniform.
constexpr int32_t FB_CHUNK = 256;
__device__ __forceinline__ int64_t wave_reserve_chunk(unsigned long long *count, int32_t n, int64_t &next,
                                                      int64_t &end)
{
  const int lane = __lane_id();
  int32_t incl = n;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int32_t y = __shfl_up(incl, off, 64);
    if (lane >= off) incl += y;
  }
  const int32_t total = __shfl(incl, 63, 64);
  if (total == 0) return 0;
  if (next + total > end) {
    const int64_t want = total > FB_CHUNK ? total : FB_CHUNK;
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(count, (unsigned long long)want);
    next = (int64_t)__shfl((long long)base, 0, 64);
    end = next + want;
  }
  const int64_t mine = next + (int64_t)(incl - n);
  next += total;
  return mine;
}

// Per-history keff scores of a history that ended.
__device__ __forceinline__ void store_keff(const Params &P, const PState &st)
{
  const int64_t n = P.n_particles;
  P.fb.k_hist[st.hid] = st.ktl;
  P.fb.k_hist[n + st.hid] = st.kcol;
  P.fb.k_hist[2 * n + st.hid] = st.kabs;
}

__device__ __forceinline__ int fate_of(int ev)
{
  switch (ev) {
    case EV_LEAK: return FATE_LEAKED;
    case EV_ABSORB: return FATE_ABSORBED;
    case EV_CUTOFF: return FATE_CUTOFF;
    case EV_LOST: return FATE_LOST;
    default: return FATE_ALIVE;
  }
}

__device__ __forceinline__ bool is_death(int ev)
{
  return ev == EV_LEAK || ev == EV_LOST || ev == EV_ABSORB || ev == EV_CUTOFF;
}

__device__ void write_record(const Params &P, const PState &st, int fate)
{
  const int64_t i = st.hid;
  const RecordView &R = P.rec;
  R.digest[i] = st.h;
  R.seed[i] = st.seed;
  R.E[i] = st.E;
  R.x[i] = st.x;
  R.y[i] = st.y;
  R.z[i] = st.z;
  R.nseg[i] = st.nseg;
  R.ncoll[i] = st.ncoll;
  R.elem[i] = (int32_t)orig_elem(P.mesh, st.elem);
  R.fate[i] = fate;
}

// ------------------------------------------------------------------ bank I/O

template <bool record>
__device__ __forceinline__ void load(const BankView &B, int64_t s, PState &st)
{
  st.x = B.x[s]; st.y = B.y[s]; st.z = B.z[s];
  st.u = B.u[s]; st.v = B.v[s]; st.w = B.w[s];
  st.E = B.E[s]; st.sigt = B.sigt[s];
  st.seed = B.seed[s];
  st.hid = B.hid[s];
  st.elem = B.elem[s]; st.inface = B.inface[s]; st.mat = B.mat[s]; st.k = B.k[s];
  st.nseg = B.nseg[s];
  if (record) {
    st.h = B.digest[s];
    st.ncoll = B.ncoll[s];
  } else {
    st.h = 0;
    st.ncoll = 0;
  }
}

template <bool record>
__device__ __forceinline__ void store(const BankView &B, int64_t s, const PState &st)
{
  B.x[s] = st.x; B.y[s] = st.y; B.z[s] = st.z;
  B.u[s] = st.u; B.v[s] = st.v; B.w[s] = st.w;
  B.E[s] = st.E; B.sigt[s] = st.sigt;
  B.seed[s] = st.seed;
  B.hid[s] = st.hid;
  B.elem[s] = st.elem; B.inface[s] = st.inface; B.mat[s] = st.mat; B.k[s] = st.k;
  B.nseg[s] = st.nseg;
  if (record) {
    B.digest[s] = st.h;
    B.ncoll[s] = st.ncoll;
  }
}

__device__ __forceinline__ void load_site(const SourceBank &S, int64_t i, PState &st)
{
  st.x = S.x[i]; st.y = S.y[i]; st.z = S.z[i];
  st.u = S.u[i]; st.v = S.v[i]; st.w = S.w[i];
  st.E = S.E[i];
  st.seed = S.seed[i];
  st.elem = S.elem[i];
}

__device__ __forceinline__ void stage_hash(const XSView &X, int32_t *s_hash)
{
  for (int i = threadIdx.x; i <= X.n_hash; i += blockDim.x) s_hash[i] = X.hash[i];
  __syncthreads();
}

} // namespace mcw

using namespace mcw;

// ------------------------------------------------------------------ source bank

// Sites of histories hist_begin .. hist_begin + n - 1 into slots 0 .. n-1.
extern "C" __global__ void __launch_bounds__(256) k_source_bank(Params P, SourceBank S, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double p[3], d[3], E;
  uint64_t seed;
  int32_t elem;
  sample_site(P, P.src, P.master_seed, P.batch, P.n_particles, P.hist_begin + i, p, d, E, seed, elem);
  S.x[i] = p[0]; S.y[i] = p[1]; S.z[i] = p[2];
  S.u[i] = d[0]; S.v[i] = d[1]; S.w[i] = d[2];
  S.E[i] = E;
  S.seed[i] = seed;
  S.elem[i] = elem;
}

// Sites of a request-server round: slot i belongs to the slice whose range
// holds it.
extern "C" __global__ void __launch_bounds__(256) k_source_req(Params P, SourceBank S, int32_t n_slices, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const ReqSlice *R = P.slices;
  int32_t lo = 0, hi = n_slices; // last slice with first <= i
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (R[mid].first <= i) lo = mid;
    else hi = mid;
  }
  const ReqSlice &r = R[lo];
  double p[3], d[3], E;
  uint64_t seed;
  int32_t elem;
  sample_site(P, r.src, r.master_seed, 1, r.n_particles, r.hist0 + (i - r.first), p, d, E, seed, elem);
  S.x[i] = p[0]; S.y[i] = p[1]; S.z[i] = p[2];
  S.u[i] = d[0]; S.v[i] = d[1]; S.w[i] = d[2];
  S.E[i] = E;
  S.seed[i] = seed;
  S.elem[i] = elem;
  S.req[i] = lo;
}

// ------------------------------------------------------------------ PLAIN kernels

// Live bank <- source bank (slot i = history hist_begin + i).
template <bool REC>
__global__ void __launch_bounds__(256) k_bank_init(Params P, SourceBank S, BankView B, int64_t n, int q)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) P.cnt->q_count[q] = (unsigned long long)n;
  if (i >= n) return;
  PState st;
  load_site(S, i, st);
  start_history(st, P.hist_begin + i);
  st.k = union_index(P.xs, P.xs.hash, st.E);
  store<REC>(B, i, st);
}

// One segment for every particle of the live bank `in` (q_count[qin] of them).
template <int NS, bool REC>
__global__ void __launch_bounds__(256) k_segment(Params P, BankView in, int qin, BankView out, int qout,
                                                 BankView coll, int qcoll)
{
  Counters *C = P.cnt;
  const int64_t n = (int64_t)C->q_count[qin];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  WaveCount wc;
  wc.zero();
  double track = 0.0;
  unsigned rechecks = 0;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    const int64_t i = base + threadIdx.x;
    const bool valid = i < n;
    int ev = -1;
    bool hx = NS == 6;
    PState st;
    if (valid) {
      load<REC>(in, i, st);
      ev = segment<NS, REC, false>(P, st, track, rechecks, nullptr, nullptr, &hx);
      if (REC && is_death(ev)) write_record(P, st, fate_of(ev));
    }
    count_segment(wc, valid, ev, hx);
    const bool to_coll = valid && ev == EV_COLLIDE;
    const bool to_live = valid && (ev == EV_CROSS || ev == EV_REFLECT);
    int64_t pc, pl;
    block_append2(&C->q_count[qcoll], to_coll, pc, &C->q_count[qout], to_live, pl);
    if (to_coll) store<REC>(coll, pc, st);
    if (to_live) store<REC>(out, pl, st);
  }
  flush_counts(C, wc, track);
}

template <bool REC>
__global__ void __launch_bounds__(256) k_collide(Params P, BankView coll, int qcoll, BankView out, int qout)
{
  Counters *C = P.cnt;
  const int64_t n = (int64_t)C->q_count[qcoll];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  WaveCount wc;
  wc.zero();
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    const int64_t i = base + threadIdx.x;
    const bool valid = i < n;
    int ev = -1;
    PState st;
    if (valid) {
      load<REC>(coll, i, st);
      ev = collide<REC>(P, st, P.xs.hash);
      if (REC && ev != EV_SCATTER) write_record(P, st, fate_of(ev));
    }
    count_collision(wc, valid, ev);
    const bool to_live = valid && ev == EV_SCATTER;
    int64_t pl, unused;
    block_append2(&C->q_count[qout], to_live, pl, nullptr, false, unused);
    if (to_live) store<REC>(out, pl, st);
  }
  flush_counts(C, wc, 0.0);
}

extern "C" __global__ void k_reset_work(Counters *C)
{
  if (threadIdx.x == 0) C->next_hist = 0ull;
}

extern "C" __global__ void k_reset_queue(Counters *C, int q0, int q1)
{
  if (threadIdx.x == 0) {
    if (q0 >= 0) C->q_count[q0] = 0;
    if (q1 >= 0) C->q_count[q1] = 0;
  }
}

// ------------------------------------------------------------------ BATCHED kernel

enum { PH_IDLE = 0, PH_TRANSPORT = 1, PH_COLLIDE = 2 };

// WPE: minimum wavefronts per SIMD requested from the register allocator
// (0 = compiler's choice); a tuning knob trading spills for occupancy.
// Per-request segment and collision counts of the request server.  Each lane
// keeps a running sum (segments << 32 | collisions) for the request of its
// last history; it is added to the request's counters when the lane moves on
// to another request, and at the end of the kernel.  Adding at every history
// end instead puts every history of a request on the same two addresses,
// and same-address atomics serialise.
struct ReqAcc {
  unsigned long long v = 0ull;
  int32_t req = -1;
};

// Add v of the lanes with `todo` to their requests' counters: one atomic pair
// per distinct request among those lanes.  Call with the whole wave.
__device__ __forceinline__ void req_flush(const Params &P, bool todo, int32_t req, unsigned long long v)
{
  for (;;) {
    const unsigned long long m = __ballot(todo);
    if (m == 0ull) break;
    const int leader = __builtin_ctzll(m);
    const int32_t r = __shfl(req, leader, 64);
    const bool in = todo && req == r;
    const unsigned long long hi = wave_sum(in ? (v >> 32) : 0ull);
    const unsigned long long lo = wave_sum(in ? (v & 0xffffffffull) : 0ull);
    if (__lane_id() == leader) {
      atomic_add_dev(P.slices[r].counts, hi);
      atomic_add_dev(P.slices[r].counts + 1, lo);
    }
    if (in) todo = false;
  }
}

// A history of request st.req ended on this lane when `dying`: fold its counts
// into the lane's sum (nseg, ncoll < 2^20 each; the sum is flushed before
// its segment half could pass 2^32).  Call with the whole wave.
__device__ __forceinline__ void req_account(const Params &P, bool dying, const PState &st, ReqAcc &a)
{
  req_flush(P, dying && a.req != st.req && a.v != 0ull, a.req, a.v);
  if (dying) {
    if (a.req != st.req) {
      a.v = 0ull;
      a.req = st.req;
    }
    a.v += ((unsigned long long)st.nseg << 32) | (unsigned long long)(uint32_t)st.ncoll;
  }
  const bool full = (a.v >> 32) >= (1ull << 31);
  req_flush(P, full, a.req, a.v);
  if (full) a.v = 0ull;
}

// ------------------------------------------------------------------ staged flux tally

// One wavefront's LDS stage: up to STG_S records per bucket; per bucket the
// pool position its next full bucket goes to (a multiple of STG_CH means the
// current chunk is full, or there is none yet); the next unused chunk of the
// wavefront's slice.
struct StageWave {
  double sc[STG_NB][STG_S];
  uint16_t el[STG_NB][STG_S];
  uint32_t cnt[STG_NB];
  uint32_t pos[STG_NB];
  uint32_t next;
  uint32_t pad;
};

static_assert(STG_NB == 64, "a wavefront clears its stage one bucket per lane");
static_assert(STG_CH % STG_S == 0 && STG_CH % 64 == 0 && STG_CH < 65536, "chunks hold whole stage writes");
static_assert(64 % STG_S == 0, "lane groups of STG_S");
static_assert(4 * sizeof(StageWave) + (STG_HASH + 1) * 4 <= 160 * 1024 / 4, "four workgroups of 256 per CU");

constexpr uint32_t STG_NONE = 0xffffffffu;

// Lane group g = lane / STG_S moves the first nrec records of bucket bb
// (act: the group has a bucket) from the stage to the pool, one record per
// lane.  The group's first lane takes the next chunk of the wavefront's slice
// (wave: the wavefront's index in the launch) when the current one is full;
// past the slice's end the records go to the tally as device atomics.  A
// chunk's record count is written when it fills, and by the wavefront's final
// call (final) when it stays partly filled.  Whole wavefront.
__device__ __forceinline__ void stage_write(const Params &P, StageWave &W, bool act, uint32_t bb, uint32_t nrec,
                                            int64_t wave, bool final)
{
  const StageView &V = P.stg;
  const int lane = __lane_id();
  const uint32_t r = (uint32_t)(lane & (STG_S - 1));
  uint32_t pos = act ? W.pos[bb] : 0u;
  const bool need = act && r == 0 && nrec > 0 && pos % STG_CH == 0;
  const unsigned long long nm = __ballot(need);
  if (nm) {
    const uint32_t first = W.next;
    if (need) {
      const uint32_t ci = first + (uint32_t)__popcll(nm & ((1ull << lane) - 1ull));
      if (ci < (uint32_t)V.wave_chunks) {
        const int64_t chunk = wave * V.wave_chunks + ci;
        V.owner[chunk] = (uint8_t)bb;
        pos = (uint32_t)(chunk * STG_CH);
      } else {
        pos = STG_NONE;
      }
    }
    if (lane == 0) W.next = first + (uint32_t)__popcll(nm);
  }
  pos = (uint32_t)__shfl((int)pos, lane & ~(STG_S - 1), 64);
  if (act && r < nrec) {
    const double v = W.sc[bb][r];
    const uint32_t l = W.el[bb][r];
    if (pos != STG_NONE) {
      // streaming stores: the records are read once, by the sums after the
      // launch, and should not displace element records from the caches
      __builtin_nontemporal_store(v, V.sc + ((int64_t)pos + r));
      __builtin_nontemporal_store((uint16_t)l, V.el + ((int64_t)pos + r));
    } else {
      atomic_add_dev(P.tally[0].value + (((int64_t)bb << STG_BITS) | l), v);
    }
  }
  if (act && r == 0) {
    if (pos != STG_NONE) {
      const uint32_t np = pos + nrec;
      W.pos[bb] = np;
      if (np % STG_CH == 0) {
        if (nrec > 0) V.fill[np / STG_CH - 1] = (uint16_t)STG_CH;
      } else if (final) {
        V.fill[np / STG_CH] = (uint16_t)(np % STG_CH);
      }
    }
    W.cnt[bb] = 0u;
  }
}

// Stages lane scores (on: the lane has one for bin e) and writes the buckets
// they fill, eight per pass; a lane that finds its bucket full tries again
// after the write.  Whole wavefront.
__device__ __forceinline__ void stage_add(const Params &P, StageWave &W, bool on, uint32_t e, double v, int64_t wave)
{
  const uint32_t b = e >> STG_BITS;
  const uint32_t loc = e & ((1u << STG_BITS) - 1u);
  const int g = __lane_id() / STG_S;
  bool pend = on;
  for (;;) {
    uint32_t p = STG_S;
    if (pend) {
      p = atomicAdd(&W.cnt[b], 1u);
      if (p < STG_S) {
        W.sc[b][p] = v;
        W.el[b][p] = (uint16_t)loc;
        pend = false;
      }
    }
    // the lane that took a bucket's last slot names it (one lane per bucket)
    unsigned long long full = __ballot(p == STG_S - 1);
    while (full) {
      unsigned long long m = full;
      for (int i = 0; i < g; ++i) m &= m - 1ull;
      const bool act = m != 0ull;
      const uint32_t bb = (uint32_t)__shfl((int)b, act ? __builtin_ctzll(m) : 0, 64);
      stage_write(P, W, act, bb, STG_S, wave, false);
      for (int i = 0; i < 64 / STG_S; ++i) full &= full - 1ull;
    }
    if (__ballot(pend) == 0ull) break;
  }
}

// The wavefront's last records (fewer than STG_S per bucket) and the counts
// of its partly filled chunks.  Whole wavefront.
__device__ __forceinline__ void stage_drain(const Params &P, StageWave &W, int64_t wave)
{
  const uint32_t g = (uint32_t)(__lane_id() / STG_S);
  for (uint32_t j = 0; j < (uint32_t)STG_NB; j += 64 / STG_S) {
    const uint32_t bb = j + g;
    const bool act = bb < (uint32_t)P.stg.n_buckets;
    stage_write(P, W, act, bb, act ? W.cnt[bb] : 0u, wave, true);
  }
}

// Chunks per bucket: bcount[b] += chunks whose owner is b.
extern "C" __global__ void __launch_bounds__(256) k_stage_count(StageView V)
{
  __shared__ unsigned h[STG_NB];
  if (threadIdx.x < STG_NB) h[threadIdx.x] = 0u;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < V.n_chunks; c += stride) {
    const unsigned o = V.owner[c];
    if (o < (unsigned)STG_NB) atomicAdd(&h[o], 1u);
  }
  __syncthreads();
  if (threadIdx.x < STG_NB && h[threadIdx.x]) atomicAdd(V.bcount + threadIdx.x, (unsigned long long)h[threadIdx.x]);
}

constexpr int STG_TILE = 4; // chunks per thread of k_stage_scatter

// Lists each taken chunk under its bucket: bucket b's chunks fill list[off(b)
// .. off(b) + bcount[b]), off the exclusive prefix sum of bcount, in any
// order, as chunk << 9 | records; bcount[STG_NB + b] is the bucket's cursor
// (zero before the call).
extern "C" __global__ void __launch_bounds__(256) k_stage_scatter(StageView V)
{
  __shared__ unsigned h[STG_NB];
  __shared__ unsigned long long base[STG_NB];
  if (threadIdx.x < STG_NB) h[threadIdx.x] = 0u;
  __syncthreads();
  const int64_t c0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * STG_TILE;
  unsigned o[STG_TILE], rank[STG_TILE];
#pragma unroll
  for (int q = 0; q < STG_TILE; ++q) {
    o[q] = c0 + q < V.n_chunks ? (unsigned)V.owner[c0 + q] : (unsigned)STG_FREE;
    rank[q] = o[q] < (unsigned)STG_NB ? atomicAdd(&h[o[q]], 1u) : 0u;
  }
  __syncthreads();
  if (threadIdx.x < STG_NB) {
    const int b = threadIdx.x;
    unsigned long long off = 0;
    for (int j = 0; j < b; ++j) off += V.bcount[j];
    base[b] = h[b] ? off + atomicAdd(V.bcount + STG_NB + b, (unsigned long long)h[b]) : 0ull;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < STG_TILE; ++q)
    if (o[q] < (unsigned)STG_NB) {
      const uint32_t n = min((uint32_t)V.fill[c0 + q], (uint32_t)STG_CH);
      V.list[base[o[q]] + rank[q]] = (uint32_t)(c0 + q) << 9 | n;
    }
}

constexpr int STG_UNROLL = 4; // chunks in flight per wavefront of k_stage_reduce

// Sums bucket blockIdx.x's records in LDS (workgroups blockIdx.y share its
// chunks, STG_UNROLL at a time per wavefront) and adds the sums to
// tally[bucket bins].
extern "C" __global__ void __launch_bounds__(1024) k_stage_reduce(StageView V, double *tally, int64_t n_bins)
{
  static_assert(STG_CH <= 256 && (1 << 23) * (int64_t)STG_CH >= ((int64_t)1 << 31), "list entries: chunk << 9 | count");
  __shared__ double acc[1 << STG_BITS];
  const int b = blockIdx.x;
  for (int i = threadIdx.x; i < (1 << STG_BITS); i += blockDim.x) acc[i] = 0.0;
  unsigned long long off = 0;
  for (int j = 0; j < b; ++j) off += V.bcount[j];
  const int64_t cnt = (int64_t)V.bcount[b];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.y * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t waves = (int64_t)gridDim.y * (blockDim.x >> 6);
  constexpr int Q = STG_CH / 64;
  constexpr uint32_t MASK = (1u << STG_BITS) - 1u;
  // every load below is unconditional (indices clamped into the chunk or the
  // list, results masked after): loads under a branch each waited for their
  // data before the next was issued
  for (int64_t i = wave; i < cnt; i += STG_UNROLL * waves) {
    uint32_t ent[STG_UNROLL];
#pragma unroll
    for (int u = 0; u < STG_UNROLL; ++u) {
      const int64_t j = i + u * waves;
      const uint32_t e = V.list[off + (j < cnt ? j : cnt - 1)];
      ent[u] = j < cnt ? e : 0u;
    }
    double v[STG_UNROLL][Q];
    uint32_t l[STG_UNROLL][Q];
#pragma unroll
    for (int u = 0; u < STG_UNROLL; ++u) {
      const int n = (int)(ent[u] & 511u);
      const int64_t at = (int64_t)(ent[u] >> 9) * STG_CH;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int k = q * 64 + lane;
        const int64_t r = at + (k < n ? k : 0);
        v[u][q] = V.sc[r];
        l[u][q] = V.el[r];
      }
    }
#pragma unroll
    for (int u = 0; u < STG_UNROLL; ++u) {
      const int n = (int)(ent[u] & 511u);
#pragma unroll
      for (int q = 0; q < Q; ++q)
        if (q * 64 + lane < n) atomicAdd(&acc[l[u][q] & MASK], v[u][q]);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < (1 << STG_BITS); i += blockDim.x) {
    const int64_t e = ((int64_t)b << STG_BITS) + i;
    if (e < n_bins && acc[i] != 0.0) atomic_add_dev(tally + e, acc[i]);
  }
}

// ------------------------------------------------------------------ request-server tally cache

// A request's flux tally is its own array over the mesh, so a round's bins
// are (request slice, element) pairs: too many to stage in LDS.  Point
// sources send every history's first segments to the same few elements,
// whose device atomics then serialise.  Each wavefront of the serving kernel
// keeps the sums of its last SC_N (slice, element) keys in LDS: an add to a
// cached key stays in LDS, and a key that needs the slot adds the slot's sum
// to the device tally first.  Every cached sum is added at the end of the
// launch.
constexpr int SC_N = 256;
constexpr uint32_t SC_EMPTY = 0xffffffffu;
constexpr int SC_ELEM_BITS = 26;

struct ServeCache {
  double val[SC_N];
  uint32_t key[SC_N]; // slice << SC_ELEM_BITS | element, SC_EMPTY when free
  uint8_t claim[SC_N];
};

static_assert(4 * sizeof(ServeCache) + (STG_HASH + 1) * 4 <= 160 * 1024 / 4, "four serving workgroups per CU");

// Volatile pointers into LDS (address space 3): a plain volatile pointer is a
// generic one, which turns the accesses into flat instructions.
template <class T> using lds_volatile = __attribute__((address_space(3))) volatile T *;

__device__ __forceinline__ double *serve_bin(const Params &P, uint32_t key)
{
  return P.slices[key >> SC_ELEM_BITS].tally + (key & ((1u << SC_ELEM_BITS) - 1u));
}

// Adds v to slice `slice`'s bin `elem` (bin: its address) through the
// wavefront's cache.  Lanes with the same key are summed first; then each
// round, one lane per slot (the one whose id the slot's claim byte keeps)
// updates it.  Keys that do not fit (slice >= 63 or elem >= 2^26) go to the
// device directly.  Whole wavefront.
__device__ __forceinline__ void serve_cache_add(const Params &P, ServeCache &K, bool on, int32_t slice,
                                                int32_t elem, double *bin, double v)
{
  const int lane = __lane_id();
  const bool fits = on && slice >= 0 && slice < 63 && (uint32_t)elem < (1u << SC_ELEM_BITS);
  if (on && !fits) atomic_add_dev(bin, v);
  const uint32_t key = ((uint32_t)slice << SC_ELEM_BITS) | (uint32_t)elem;
  bool cand = fits, pend = fits;
  double vv = v;
  for (int g = 0; g < 4; ++g) {
    const unsigned long long m = __ballot(cand);
    if (m == 0ull) break;
    const int leader = __builtin_ctzll(m);
    const uint32_t kl = (uint32_t)__shfl((int)key, leader, 64);
    const bool same = cand && key == kl;
    if (__popcll(__ballot(same)) < 2) break;
    const double sum = wave_sum_dbl(same ? vv : 0.0);
    if (same) {
      cand = false;
      if (lane == leader) vv = sum;
      else pend = false;
    }
  }
  const uint32_t slot = (key ^ (key >> 7) ^ (key >> SC_ELEM_BITS)) & (SC_N - 1);
  // volatile: a lane must read back the claim byte the wavefront's store left
  // (another lane's id when several claimed the slot), not its own store
  // forwarded by the compiler; and the key and sum other lanes wrote
  const lds_volatile<uint8_t> claim = (lds_volatile<uint8_t>)K.claim;
  const lds_volatile<uint32_t> keys = (lds_volatile<uint32_t>)K.key;
  const lds_volatile<double> vals = (lds_volatile<double>)K.val;
  while (__ballot(pend)) {
    if (pend) claim[slot] = (uint8_t)lane;
    if (pend && claim[slot] == (uint8_t)lane) {
      const uint32_t k = keys[slot];
      if (k == key) {
        vals[slot] = vals[slot] + vv;
      } else {
        if (k != SC_EMPTY) atomic_add_dev(serve_bin(P, k), vals[slot]);
        keys[slot] = key;
        vals[slot] = vv;
      }
      pend = false;
    }
  }
}

// The cached sums, at the end of the launch.  Whole wavefront.
__device__ __forceinline__ void serve_cache_flush(const Params &P, ServeCache &K)
{
  const lds_volatile<uint32_t> keys = (lds_volatile<uint32_t>)K.key;
  const lds_volatile<double> vals = (lds_volatile<double>)K.val;
  for (int i = __lane_id(); i < SC_N; i += 64) {
    const uint32_t k = keys[i];
    if (k != SC_EMPTY) atomic_add_dev(serve_bin(P, k), vals[i]);
  }
}

template <int NS, bool REC, int WPE, bool P32, bool SERVE = false, bool EIG = false, bool STG = false>
// Mixed-mesh variants (NS = 0) of the fixed-source and serving kernels are
// held to 128 VGPRs, four waves per SIMD, which they measured faster at than
// at three (C5: 150 against 167 ms per batch), and so are the staging
// variants without records and the TET4 serving variants; WPE > 0 asks for
// more.
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : ((NS == 0 && !EIG) || (STG && !REC) || (SERVE && NS == 4) ? 4 : 1), 8)))
k_persistent(Params P, SourceBank S)
{
  // the staging and serving variants keep every other energy bucket (a 16 KB
  // table) to leave LDS for four wavefront stages or tally caches per workgroup
  __shared__ int32_t s_hash[STG || SERVE ? STG_HASH + 1 : MAX_HASH + 1];
  // a small per-element flux tally (a few hundred elements) is summed per
  // workgroup in LDS and added to the device tally once at the end: every
  // segment's atomic on one of a few hundred lines serialised in the caches
  const int32_t n_lds = SERVE || STG ? 0 : P.lds_tally + P.lds_coll;
  for (int i = threadIdx.x; i < n_lds; i += blockDim.x) s_tally[i] = 0.0;
  StageWave &W = reinterpret_cast<StageWave *>(s_tally)[STG ? threadIdx.x >> 6 : 0];
  const int64_t wave_ix = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (STG) {
    W.cnt[__lane_id()] = 0u;
    W.pos[__lane_id()] = 0u;
    if (__lane_id() == 0) W.next = 0u;
  }
  ServeCache &K = reinterpret_cast<ServeCache *>(s_tally)[SERVE ? threadIdx.x >> 6 : 0];
  if (SERVE)
    for (int i = __lane_id(); i < SC_N; i += 64) K.key[i] = SC_EMPTY;
  stage_hash(P.xs, s_hash);
  Counters *C = P.cnt;
  WaveCount wc;
  wc.zero();
  double track = 0.0;
  unsigned rechecks = 0;
  PState st;
  st.elem = 0;
  ReqAcc racc;
  int phase = PH_IDLE;
  bool exhausted = false;
  const int64_t n_hist = P.hist_end - P.hist_begin;
  const int lanes_min = P.lanes_min < 1 ? 1 : (P.lanes_min > 64 ? 64 : P.lanes_min);
  int64_t fb_next = 0, fb_end = 0;     // eigenvalue: the wavefront's chunk of bank slots
  int64_t h_next = 0, h_end = 0;       // the wavefront's chunk of history slots
  // chunks of 64 save atomics on the shared counter when histories are short
  // (C2: 52.8 -> 52.5 ms); with exact_refill a wavefront takes exactly what
  // its idle lanes need, since a chunk's unstarted histories wait for their
  // own wavefront's lanes and stretch the end of a launch of long histories
  // (C5: FP64 167 -> 157 ms, mixed precision 148 -> 126 ms)
  const bool chunked = !P.exact_refill;
  int32_t fb_written = 0;              // eigenvalue: this lane's sites written
  int64_t fb_req = 0;                  // eigenvalue: and asked for
  for (;;) {
    // refill idle lanes together from the wavefront's chunk of history slots
    // (a new chunk of at least 64 is one atomic on the shared counter; every
    // wavefront taking its refills from it directly serialised there)
    const bool need = phase == PH_IDLE && !exhausted;
    int64_t slot = -1;
    {
      const unsigned long long mask = __ballot(need);
      if (mask) {
        const int cnt = __popcll(mask);
        const int lane = __lane_id();
        const unsigned long long below = lane == 0 ? 0ull : (mask & ((~0ull) >> (64 - lane)));
        const int rank = __popcll(below);
        const int64_t avail = h_end - h_next;
        int64_t nb = 0, ne = 0;
        if (avail < cnt) {
          const int64_t short_by = cnt - avail;
          const int64_t want = chunked && short_by < 64 ? 64 : short_by;
          unsigned long long b0 = 0;
          if (lane == 0) b0 = atomicAdd(&C->next_hist, (unsigned long long)want);
          nb = (int64_t)__shfl((long long)b0, 0, 64);
          ne = nb + want;
        }
        if (need) slot = rank < avail ? h_next + rank : nb + (rank - avail);
        if (avail < cnt) {
          h_next = nb + (cnt - avail);
          h_end = ne;
        } else {
          h_next += cnt;
        }
        h_next = wave_uniform(h_next);
        h_end = wave_uniform(h_end);
      }
    }
    if (need) {
      if (slot < n_hist) {
        load_site(S, slot, st);
        start_history(st, P.hist_begin + (S.slot ? (int64_t)S.slot[slot] : slot));
        if (SERVE) {
          st.req = S.req[slot];
          st.tv = P.slices[st.req].tally;
        }
        st.k = union_index(P.xs, s_hash, st.E);
        phase = PH_TRANSPORT;
      } else {
        exhausted = true;
      }
    }
    if (__ballot(phase != PH_IDLE) == 0) break;
    // segments while enough lanes are in flight
    do {
      const bool go = phase == PH_TRANSPORT;
      int ev = -1;
      double *tl_ptr = nullptr;
      double tl_val = 0.0;
      bool hx = NS == 6;
      if (go) {
        ev = segment<NS, REC, P32, SERVE, EIG>(P, st, track, rechecks, &tl_ptr, &tl_val, &hx);
        if (ev == EV_COLLIDE) {
          phase = PH_COLLIDE;
        } else if (is_death(ev)) {
          if (REC) write_record(P, st, fate_of(ev));
          if (EIG) store_keff(P, st);
          phase = PH_IDLE;
        }
      }
      // the track-length bins are in LDS only when P.lds_tally covers them;
      // P.lds_coll alone (a large track-length tally beside a small
      // collision one) keeps the track-length adds on the device
      if (STG) {
        stage_add(P, W, tl_ptr != nullptr, tl_ptr ? (uint32_t)(tl_ptr - P.tally[0].value) : 0u, tl_val, wave_ix);
      } else if (!SERVE && P.lds_tally) {
        if (tl_ptr) atomicAdd(s_tally + (tl_ptr - P.tally[0].value), tl_val);
      } else if (SERVE) {
        // the segment's element (st.elem has moved on past a crossing)
        serve_cache_add(P, K, tl_ptr != nullptr, st.req, tl_ptr ? (int32_t)(tl_ptr - st.tv) : 0, tl_ptr, tl_val);
      } else if (P.tl_fast) {
        atomic_add_agg(tl_ptr != nullptr, tl_ptr, tl_val);
      }
      if (SERVE) req_account(P, go && is_death(ev), st, racc);
      count_segment(wc, go, ev, hx);
    } while (wave_count(phase == PH_TRANSPORT) >= lanes_min);
    // collisions together (in an eigenvalue batch the wavefront reserves its
    // fission-bank slots between sampling the yields and the sites)
    const bool hit = phase == PH_COLLIDE;
    int ev = -1;
    if (EIG) {
      CollPick cp;
      int64_t nu = 0;
      if (hit) nu = collide_pick<REC, true>(P, st, cp);
      fb_req += nu;
      const int64_t base = wave_reserve_chunk(&C->fission_slots, (int32_t)nu, fb_next, fb_end);
      if (hit) ev = collide_finish<REC, true>(P, st, cp, nu, base, s_hash, &fb_written);
    } else if (hit) {
      ev = collide<REC>(P, st, s_hash);
    }
    if (hit) {
      if (ev == EV_SCATTER) {
        phase = PH_TRANSPORT;
      } else {
        if (REC) write_record(P, st, fate_of(ev));
        if (EIG) store_keff(P, st);
        phase = PH_IDLE;
      }
    }
    if (SERVE) {
      if (hit && !REC) ++st.ncoll;
      req_account(P, hit && ev != EV_SCATTER, st, racc);
    }
    count_collision(wc, hit, ev);
  }
  if (SERVE) req_flush(P, racc.v != 0ull, racc.req, racc.v);
  if (SERVE) serve_cache_flush(P, K);
  if (STG) stage_drain(P, W, wave_ix);
  flush_counts(C, wc, track);
  if (EIG) {
    const unsigned long long w = wave_sum((unsigned long long)fb_written);
    const unsigned long long r = wave_sum((unsigned long long)fb_req);
    if (__lane_id() == 0 && r) {
      atomicAdd(&C->fission_sites, w);
      atomicAdd(&C->fission_requested, r);
    }
  }
  if (P32) {
    const unsigned long long r = wave_sum(rechecks);
    if (__lane_id() == 0 && r) atomicAdd(&C->rechecks, r);
  }
  if (!SERVE && n_lds) { // every wave of the workgroup gets here once
    __syncthreads();
    for (int i = threadIdx.x; i < n_lds; i += blockDim.x)
      if (s_tally[i] != 0.0) {
        double *dst = i < P.lds_tally ? P.tally[0].value + i : P.tally[P.n_tl].value + (i - P.lds_tally);
        atomic_add_dev(dst, s_tally[i]);
      }
  }
}

// ------------------------------------------------------------------ utilities

extern "C" __global__ void k_fill_u64(unsigned long long *p, int64_t n, unsigned long long v)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

extern "C" __global__ void k_fill(double *p, int64_t n, double v)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// dst[e] = src[perm[e]] / norm for e < n (internal -> caller element order),
// then dst[n .. n + tail) = src[n .. n + tail) bit for bit (a request's
// segment and collision counters).
extern "C" __global__ void k_gather_scale(const double *src, const int32_t *perm, int64_t n, double norm, double *dst,
                                          int tail)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n + tail; i += stride)
    dst[i] = i < n ? src[perm[i]] / norm : src[i];
}

extern "C" __global__ void k_zero_counters(Counters *C)
{
  unsigned long long *p = (unsigned long long *)C;
  const int n = sizeof(Counters) / 8;
  if (threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// VALUE -> SUM += VALUE/norm, SUM_SQ += (VALUE/norm)^2
extern "C" __global__ void k_accumulate(double *value, double *sum, double *sum_sq, int64_t n, double norm)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double v = value[i] / norm;
    sum[i] += v;
    sum_sq[i] += v * v;
  }
}

// Direction and energy of the fission sites in slots 0 .. n-1 from their
// saved stream states (the draws collide_finish stepped over); empty slots
// (key ~0) are skipped.
extern "C" __global__ void k_fission_sites(XSView X, FissionView F, int64_t n)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (F.key[i] == ~0ull) continue;
    uint64_t s = F.seed[i];
    double u, v, w, E;
    sample_fission_site(X, F.nuc[i], s, u, v, w, E);
    F.u[i] = u;
    F.v[i] = v;
    F.w[i] = w;
    F.E[i] = E;
  }
}

// Elementary functions and RNG on the device, for bit-equality tests.
extern "C" __global__ void k_math_probe(const double *x, int64_t n, double *log_out, double *sin_out,
                                        double *cos_out, double *sqrt_out, const uint64_t *seeds,
                                        double *prn_out, uint64_t *seed_out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  log_out[i] = log_d(x[i]);
  double s, c;
  sincos_2pi_d(x[i] - floor(x[i]), &s, &c);
  sin_out[i] = s;
  cos_out[i] = c;
  sqrt_out[i] = sqrt(x[i]);
  uint64_t sd = seeds[i];
  prn_out[i] = prn(sd);
  seed_out[i] = future_seed((uint64_t)i * PRN_STRIDE + 7, seeds[i]);
}

// The device's FP64 atomic-add rate: every thread adds 1.0 `iters` times at
// pseudo-random positions of a[0 .. mask] (mask + 1 a power of two), with
// the same instruction the tallies use.  The tally kernels are measured
// against this rate: on MI355X every FP64 atomic is performed at the memory
// side, whatever the array's size.
extern "C" __global__ void __launch_bounds__(256) k_atomic_probe(double *a, uint64_t mask, int32_t iters)
{
  uint64_t s = 0x9E3779B97F4A7C15ull * ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x + 1);
  for (int32_t i = 0; i < iters; ++i) {
    s = PRN_MULT * s + PRN_ADD;
    atomic_add_dev(a + ((s >> 20) & mask), 1.0);
  }
}

// Random 128-byte line reads: each thread reads 16 bytes from `iters`
// pseudo-random lines of a (lines x 128 B) array, eight independent reads in
// flight at a time, and writes one sum so that no read is dropped.  The
// gather kernels' ceiling: a segment reads one or two such lines.
extern "C" __global__ void __launch_bounds__(256) k_gather_probe(const double4 *a, uint64_t mask, int32_t iters,
                                                                 double *out)
{
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t s = 0x9E3779B97F4A7C15ull * (tid + 1);
  double acc = 0.0;
  for (int32_t i = 0; i < iters; i += 8) {
    uint64_t idx[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      s = PRN_MULT * s + PRN_ADD;
      idx[k] = (s >> 20) & mask;
    }
    double4 v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = a[4 * idx[k]];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += v[k].x + v[k].w;
  }
  out[tid] = acc;
}

// Union-grid lookup on the device, for tests: k and (total, absorption,
// scatter, fission) of nuclide `nuc` at each energy.
extern "C" __global__ void k_xs_probe(XSView X, int32_t nuc, const double *E, int64_t n, int32_t *k_out,
                                      double *xs_out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t k = union_index(X, X.hash, E[i]);
  Micro m;
  micro_at(X, nuc, k, E[i], m);
  k_out[i] = k;
  xs_out[4 * i] = m.t;
  xs_out[4 * i + 1] = m.a;
  xs_out[4 * i + 2] = m.s;
  xs_out[4 * i + 3] = m.f;
}

namespace mcw {
// Host-side launch helpers (the templates live in this translation unit).
static inline unsigned grid_for(int64_t n, int block, int64_t cap)
{
  int64_t b = (n + block - 1) / block;
  if (cap > 0 && b > cap) b = cap;
  return (unsigned)(b < 1 ? 1 : b);
}

hipError_t launch_source_bank(hipStream_t s, const Params &P, const SourceBank &S, int64_t n)
{
  hipLaunchKernelGGL(k_source_bank, dim3(grid_for(n, 256, 0)), dim3(256), 0, s, P, S, n);
  return hipGetLastError();
}

// Kernels are instantiated per side count (4 TET4, 6 HEX8, 0 mixed) and
// record mode, so that neither costs registers where it is not used.
#define MCW_BY_NS_REC(ns, rec, F, ...)                                                            \
  do {                                                                                           \
    if (rec) {                                                                                   \
      if (ns == 4) F(4, true, __VA_ARGS__);                                                      \
      else if (ns == 6) F(6, true, __VA_ARGS__);                                                 \
      else F(0, true, __VA_ARGS__);                                                              \
    } else {                                                                                     \
      if (ns == 4) F(4, false, __VA_ARGS__);                                                     \
      else if (ns == 6) F(6, false, __VA_ARGS__);                                                \
      else F(0, false, __VA_ARGS__);                                                             \
    }                                                                                            \
  } while (0)

hipError_t launch_bank_init(hipStream_t s, const Params &P, const SourceBank &S, const BankView &B, int64_t n, int q)
{
  if (P.record) hipLaunchKernelGGL(k_bank_init<true>, dim3(grid_for(n, 256, 0)), dim3(256), 0, s, P, S, B, n, q);
  else hipLaunchKernelGGL(k_bank_init<false>, dim3(grid_for(n, 256, 0)), dim3(256), 0, s, P, S, B, n, q);
  return hipGetLastError();
}

#define MCW_SEG(NS, REC, ...) hipLaunchKernelGGL((k_segment<NS, REC>), __VA_ARGS__)
hipError_t launch_segment(int ns, unsigned blocks, hipStream_t s, const Params &P, const BankView &in, int qin,
                          const BankView &out, int qout, const BankView &coll, int qcoll)
{
  MCW_BY_NS_REC(ns, P.record, MCW_SEG, dim3(blocks), dim3(256), 0, s, P, in, qin, out, qout, coll, qcoll);
  return hipGetLastError();
}

hipError_t launch_collide(unsigned blocks, hipStream_t s, const Params &P, const BankView &coll, int qcoll,
                          const BankView &out, int qout)
{
  if (P.record) hipLaunchKernelGGL(k_collide<true>, dim3(blocks), dim3(256), 0, s, P, coll, qcoll, out, qout);
  else hipLaunchKernelGGL(k_collide<false>, dim3(blocks), dim3(256), 0, s, P, coll, qcoll, out, qout);
  return hipGetLastError();
}

template <int WPE> const void *persistent_kernel_w(int ns, bool rec, bool p32)
{
  if (p32 && ns == 4) return rec ? (const void *)k_persistent<4, true, WPE, true> : (const void *)k_persistent<4, false, WPE, true>;
  if (p32 && ns == 0) return rec ? (const void *)k_persistent<0, true, WPE, true> : (const void *)k_persistent<0, false, WPE, true>;
  if (rec) {
    if (ns == 4) return (const void *)k_persistent<4, true, WPE, false>;
    if (ns == 6) return (const void *)k_persistent<6, true, WPE, false>;
    return (const void *)k_persistent<0, true, WPE, false>;
  }
  if (ns == 4) return (const void *)k_persistent<4, false, WPE, false>;
  if (ns == 6) return (const void *)k_persistent<6, false, WPE, false>;
  return (const void *)k_persistent<0, false, WPE, false>;
}

// Eigenvalue-batch variants (compiler-chosen occupancy).
template <bool REC> const void *eigen_kernel_r(int ns, bool p32)
{
  if (ns == 4) return p32 ? (const void *)k_persistent<4, REC, 0, true, false, true>
                          : (const void *)k_persistent<4, REC, 0, false, false, true>;
  if (ns == 6) return (const void *)k_persistent<6, REC, 0, false, false, true>;
  return p32 ? (const void *)k_persistent<0, REC, 0, true, false, true>
             : (const void *)k_persistent<0, REC, 0, false, false, true>;
}

// Staged-flux-tally variants (TET4 meshes, fixed-source batches: an
// eigenvalue kernel with a stage needs 180 VGPRs and measured slower).
const void *staged_kernel(bool rec, bool p32)
{
  if (p32) return rec ? (const void *)k_persistent<4, true, 0, true, false, false, true>
                      : (const void *)k_persistent<4, false, 0, true, false, false, true>;
  return rec ? (const void *)k_persistent<4, true, 0, false, false, false, true>
             : (const void *)k_persistent<4, false, 0, false, false, false, true>;
}

// The kernel variant for side count, record mode, occupancy hint, precision,
// run mode and tally staging.
const void *persistent_kernel(int ns, int wpe, bool rec, bool p32, bool eig, bool stg)
{
  if (stg && ns == 4 && wpe == 0 && !eig) return staged_kernel(rec, p32);
  if (stg) return nullptr;
  if (eig) return rec ? eigen_kernel_r<true>(ns, p32) : eigen_kernel_r<false>(ns, p32);
  return wpe == 5 ? persistent_kernel_w<5>(ns, rec, p32) : persistent_kernel_w<0>(ns, rec, p32);
}

// Dynamic LDS of a persistent launch: the small LDS tallies, or with a staged
// tally one StageWave per wavefront.
size_t persistent_lds_bytes(const Params &P, int block)
{
  if (P.stg.n_buckets > 0) return (size_t)(block / 64) * sizeof(StageWave);
  return (size_t)(P.lds_tally + P.lds_coll) * sizeof(double);
}

hipError_t launch_persistent(int ns, int wpe, bool p32, unsigned blocks, int block, hipStream_t s, const Params &P,
                             const SourceBank &S)
{
  const void *k = persistent_kernel(ns, wpe, P.record != 0, p32, P.eigen != 0, P.stg.n_buckets > 0);
  if (!k) return hipErrorInvalidValue;
  void *args[] = {(void *)&P, (void *)&S};
  return hipLaunchKernel(k, dim3(blocks), dim3(block), args, persistent_lds_bytes(P, block), s);
}

// Sums a staged flux tally's pool into tally[n_bins] (n_bins <= STG_NB << STG_BITS):
// chunks per bucket, the chunk list, then the per-bucket sums.  bcount holds
// 2 STG_NB zeros before the call.
hipError_t launch_stage_reduce(hipStream_t s, const StageView &V, double *tally, int64_t n_bins, int split)
{
  const unsigned gc = (unsigned)std::min<int64_t>((V.n_chunks + 255) / 256, 4096);
  hipLaunchKernelGGL(k_stage_count, dim3(gc), dim3(256), 0, s, V);
  const unsigned gs = (unsigned)((V.n_chunks + 256 * STG_TILE - 1) / (256 * STG_TILE));
  hipLaunchKernelGGL(k_stage_scatter, dim3(gs), dim3(256), 0, s, V);
  hipLaunchKernelGGL(k_stage_reduce, dim3((unsigned)V.n_buckets, (unsigned)split), dim3(1024), 0, s, V, tally,
                     n_bins);
  return hipGetLastError();
}

hipError_t launch_source_req(hipStream_t s, const Params &P, const SourceBank &S, int32_t n_slices, int64_t n)
{
  hipLaunchKernelGGL(k_source_req, dim3(grid_for(n, 256, 0)), dim3(256), 0, s, P, S, n_slices, n);
  return hipGetLastError();
}

// Request-server transport kernel (no records, compiler-chosen occupancy).
const void *serve_kernel(int ns, bool p32)
{
  if (ns == 4) return p32 ? (const void *)k_persistent<4, false, 0, true, true> : (const void *)k_persistent<4, false, 0, false, true>;
  if (ns == 6) return (const void *)k_persistent<6, false, 0, false, true>;
  return p32 ? (const void *)k_persistent<0, false, 0, true, true> : (const void *)k_persistent<0, false, 0, false, true>;
}

// Dynamic LDS of a serving launch of 256-thread workgroups: a tally cache per wavefront.
size_t serve_lds_bytes() { return 4 * sizeof(ServeCache); }

hipError_t launch_serve(int ns, bool p32, unsigned blocks, hipStream_t s, const Params &P, const SourceBank &S)
{
  void *args[] = {(void *)&P, (void *)&S};
  return hipLaunchKernel(serve_kernel(ns, p32), dim3(blocks), dim3(256), args, serve_lds_bytes(), s);
}

hipError_t launch_reset_queue(hipStream_t s, Counters *C, int q0, int q1)
{
  hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, s, C, q0, q1);
  return hipGetLastError();
}

hipError_t launch_reset_work(hipStream_t s, Counters *C)
{
  hipLaunchKernelGGL(k_reset_work, dim3(1), dim3(64), 0, s, C);
  return hipGetLastError();
}

hipError_t launch_fill(hipStream_t s, double *p, int64_t n, double v)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_gather_scale(hipStream_t s, const double *src, const int32_t *perm, int64_t n, double norm,
                               double *dst, int tail)
{
  hipLaunchKernelGGL(k_gather_scale, dim3(grid_for(n + tail, 256, 4096)), dim3(256), 0, s, src, perm, n, norm, dst,
                     tail);
  return hipGetLastError();
}

hipError_t launch_zero_counters(hipStream_t s, Counters *C)
{
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, C);
  return hipGetLastError();
}

hipError_t launch_fill_u64(hipStream_t s, unsigned long long *p, int64_t n, unsigned long long v)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_fission_sites(hipStream_t s, const XSView &X, const FissionView &F, int64_t n)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fission_sites, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, X, F, n);
  return hipGetLastError();
}

hipError_t launch_accumulate(hipStream_t s, double *value, double *sum, double *sum_sq, int64_t n, double norm)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_accumulate, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, value, sum, sum_sq, n, norm);
  return hipGetLastError();
}

hipError_t launch_math_probe(hipStream_t s, const double *x, int64_t n, double *lo, double *so, double *co,
                             double *sq, const uint64_t *seeds, double *po, uint64_t *seo)
{
  hipLaunchKernelGGL(k_math_probe, dim3(grid_for(n, 256, 0)), dim3(256), 0, s, x, n, lo, so, co, sq, seeds, po, seo);
  return hipGetLastError();
}

hipError_t launch_atomic_probe(hipStream_t s, double *a, uint64_t mask, unsigned blocks, int32_t iters)
{
  hipLaunchKernelGGL(k_atomic_probe, dim3(blocks), dim3(256), 0, s, a, mask, iters);
  return hipGetLastError();
}

hipError_t launch_gather_probe(hipStream_t s, const double4 *a, uint64_t mask, unsigned blocks, int32_t iters,
                               double *out)
{
  hipLaunchKernelGGL(k_gather_probe, dim3(blocks), dim3(256), 0, s, a, mask, iters, out);
  return hipGetLastError();
}

hipError_t launch_xs_probe(hipStream_t s, const XSView &X, int32_t nuc, const double *E, int64_t n, int32_t *k,
                           double *xs)
{
  hipLaunchKernelGGL(k_xs_probe, dim3(grid_for(n, 256, 0)), dim3(256), 0, s, X, nuc, E, n, k, xs);
  return hipGetLastError();
}
} // namespace mcw
