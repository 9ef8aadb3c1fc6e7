/*
 * oracle.cpp -- history-based FP64 CPU restatement of MaCaw's transport path.
 *
 * TEST INFRASTRUCTURE.  This is the oracle the GPU library is checked
 * against; only the test suite, the smoke check and the CPU-baseline leg of
 * the benchmark load it.  The product library never links or calls it.
 *
 * What it restates, and from where (paths relative to the reference tree):
 *   - per-history source and seeds ............ src/userobjects/OpenMCStudy.C:297-364
 *     (OpenMC initialize_history / sample_external_source, stream 1)
 *   - the per-segment body ..................... src/raykernels/CollisionKernel.C:98-235
 *       xs at the particle energy (:143), one distance draw d = -ln(xi)/Sigma_t
 *       (:150), track length min(d, L) (:152-154), cross if d > L (:163),
 *       otherwise collide at x_exit - (L - d) u (:172-175), absorb -> stop (:226-230)
 *   - element walk, exit side, RayBCs ........... MOOSE ray_tracing TraceRay,
 *       ReflectRayBC, KillRayBC (test/tests/study/openmc_study.i:49-54)
 *   - collision physics ........................ OpenMC physics.cpp (sample_nuclide,
 *       absorption, elastic scatter off a target at rest, rotate_angle)
 *   - tallies and batch close ................... src/userobjects/OpenMCTally.C:103-225,
 *       src/auxkernels/OpenMCTallyAux.C:110-203, src/userobjects/OpenMCStudy.C:374-402
 *   - eigenvalue runs ........................... keff track-length estimator
 *       (CollisionKernel.C:152-154), OpenMC's collision/absorption estimators and
 *       create_fission_sites / sample_fission_neutron (called through event_collide,
 *       CollisionKernel.C:185), fission bank resampling OpenMCStudy::synchronizeBanks
 *       (src/userobjects/OpenMCStudy.C:454-607), generation and average keff
 *       (OpenMCStudy.C:404-452, OpenMC calculate_generation_keff /
 *       calculate_average_keff)
 *   - element temperatures ...................... CollisionKernel.C:109 (sqrt(kT) from the
 *       element's CONSTANT MONOMIAL temperature), OpenMC Nuclide::calculate_xs with the
 *       default "nearest" temperature method (the table whose kT is closest to
 *       sqrt(kT)^2, ties to the lower temperature), looked up per segment
 *
 * Synthetic nuclides are pointwise tables: lin-lin interpolation at the
 * bracketing grid index (largest j with E_j <= E).  Elastic scattering is
 * isotropic in the centre-of-mass frame: E' = E (A^2 + 2 A mu + 1)/(A+1)^2,
 * mu_lab = (1 + A mu)/sqrt(A^2 + 2 A mu + 1) (Lamarsh, Introduction to Nuclear
 * Reactor Theory, ch. 3).
 *
 * Floating point: compiled with -ffp-contract=off; every expression is
 * written in the same association as the GPU kernels so that per-history
 * results (crossed-element sequence, collision count, final state) agree to
 * the bit.
 */
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../include/macaw_amd.h"
#include "orc_math.h"
#include "orc_rng.h"

namespace {

const int TET_SIDES[4][3] = {{0, 2, 1}, {0, 1, 3}, {1, 2, 3}, {2, 0, 3}};
const int HEX_SIDES[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5},
                             {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 5, 6, 7}};
const int HEX_KUHN[6][4] = {{0, 1, 2, 6}, {0, 2, 3, 6}, {0, 3, 7, 6},
                            {0, 7, 4, 6}, {0, 4, 5, 6}, {0, 5, 1, 6}};

enum { EV_CROSS = 0, EV_REFLECT = 1, EV_LEAK = 2, EV_SCATTER = 3, EV_ABSORB = 4, EV_CUTOFF = 5,
       EV_LOST = 6 };

inline uint64_t digest_step(uint64_t h, int64_t elem, int ev)
{
  h ^= (uint64_t)(elem * 8 + ev);
  return h * 0x100000001b3ULL;
}

struct Nuclide {
  double awr;
  std::vector<double> E, t, a, s, f;
  bool fissionable = false;
  double nu0 = 0.0, nu1 = 0.0;      // nu(E) = nu0 + nu1 E
  double watt_a = 0.0, watt_b = 0.0; // fission spectrum
  double temperature = 293.6;        // K
  double kappa = 0.0;                // eV released per fission (kappa-fission scores)
};

// Boltzmann constant [eV/K] (OpenMC K_BOLTZMANN).
constexpr double K_BOLTZMANN = 8.617333262e-5;

struct Material {
  int id;
  std::vector<int> nuc;
  std::vector<double> dens;
};

struct Filter {
  int type;
  std::vector<int64_t> ids;
  std::vector<double> edges;
  std::vector<int64_t> elem_bin; // CELL / UNIVERSE: element -> bin or -1
  int64_t n_bins;
};

struct Tally {
  int id;
  int estimator;
  std::vector<Filter> filters;
  std::vector<int64_t> strides;
  std::vector<int> scores;
  std::vector<int> nuclides;
  int64_t n_filter_bins;
  int64_t n_score_bins;
  std::vector<double> value, sum, sum_sq;
};

struct Oracle {
  std::string err;
  mcw_options opt;
  // mesh
  int64_t n_nodes = 0, n_elem = 0;
  std::vector<double> xyz;
  std::vector<uint8_t> type;
  std::vector<int64_t> off;
  std::vector<int32_t> conn, sub;
  std::vector<int64_t> face_off;   // n_elem + 1
  std::vector<double> plane;       // 4 per face slot
  std::vector<int64_t> nbr;        // neighbour element, or -1 - boundary id
  std::vector<int32_t> nbr_side;   // neighbour's side index
  // split interfaces: a HEX8 quad side covered by two TET4 triangle sides
  // that cut it along a diagonal; split_of[slot] indexes `splits` (-1: none)
  struct Split {
    double m[4];        // plane through the diagonal, normal to the quad: m.x - c
    int64_t elem[2];    // tet on the side m.x - c <= 0, tet on the side > 0
    int32_t side[2];    // their side on the quad
  };
  std::vector<int64_t> split_of;
  std::vector<Split> splits;
  std::vector<double> cdf;         // element volume CDF
  std::map<int, int> bc;           // boundary id -> kind
  // physics
  std::vector<Nuclide> nucs;
  std::vector<std::vector<Nuclide>> nuc_temps;          // further tables per nuclide
  std::vector<std::vector<const Nuclide *>> tables;     // every table, ascending temperature
  std::vector<double> elem_T;                           // K per element; empty: first tables
  double T_tol = 10.0;
  std::vector<Material> mats;
  std::map<int, int> block_mat;      // block -> material id
  std::vector<int32_t> elem_mat;     // element -> material index (-1 void)
  bool have_source = false;
  mcw_source src;
  int64_t src_elem = -1;
  double energy_max = 0, energy_min = 0;
  std::vector<Tally> tallies;
  std::vector<mcw_history> hist;
  // eigenvalue runs
  double keff = 1.0, k_sum = 0.0, k_sum_sq = 0.0;
  int32_t n_active = 0;
  int64_t eig_batch = 0, eig_n = 0;
  std::vector<mcw_site> fbank;    // last batch's fission bank, (parent, progeny) order
  std::vector<mcw_site> next_src; // next batch's source
  mcw_keff_result kres{};
};

bool fail(Oracle *o, const std::string &m)
{
  o->err = m;
  return false;
}

// ---------------------------------------------------------------- mesh

int n_sides(uint8_t t) { return t == MCW_TET4 ? 4 : 6; }

void face_nodes(const Oracle &o, int64_t e, int s, int32_t *out, int *nn)
{
  const int32_t *c = &o.conn[o.off[e]];
  if (o.type[e] == MCW_TET4) {
    for (int k = 0; k < 3; ++k) out[k] = c[TET_SIDES[s][k]];
    *nn = 3;
  } else {
    for (int k = 0; k < 4; ++k) out[k] = c[HEX_SIDES[s][k]];
    *nn = 4;
  }
}

// Plane through the face's three lowest-numbered nodes, unit normal.
void canonical_plane(const Oracle &o, const int32_t *fn, int nn, double *p)
{
  int32_t s[4];
  for (int k = 0; k < nn; ++k) s[k] = fn[k];
  std::sort(s, s + nn);
  const double *p0 = &o.xyz[3 * (int64_t)s[0]];
  const double *p1 = &o.xyz[3 * (int64_t)s[1]];
  const double *p2 = &o.xyz[3 * (int64_t)s[2]];
  double ax = p1[0] - p0[0], ay = p1[1] - p0[1], az = p1[2] - p0[2];
  double bx = p2[0] - p0[0], by = p2[1] - p0[1], bz = p2[2] - p0[2];
  double nx = ay * bz - az * by;
  double ny = az * bx - ax * bz;
  double nz = ax * by - ay * bx;
  double len = std::sqrt((nx * nx + ny * ny) + nz * nz);
  nx = nx / len;
  ny = ny / len;
  nz = nz / len;
  p[0] = nx;
  p[1] = ny;
  p[2] = nz;
  p[3] = (nx * p0[0] + ny * p0[1]) + nz * p0[2];
}

double tet_volume6(const double *a, const double *b, const double *c, const double *d)
{
  double ax = b[0] - a[0], ay = b[1] - a[1], az = b[2] - a[2];
  double bx = c[0] - a[0], by = c[1] - a[1], bz = c[2] - a[2];
  double cx = d[0] - a[0], cy = d[1] - a[1], cz = d[2] - a[2];
  double det = (ax * (by * cz - bz * cy) - ay * (bx * cz - bz * cx)) + az * (bx * cy - by * cx);
  return std::fabs(det);
}

double elem_volume6(const Oracle &o, int64_t e)
{
  const int32_t *c = &o.conn[o.off[e]];
  if (o.type[e] == MCW_TET4)
    return tet_volume6(&o.xyz[3 * (int64_t)c[0]], &o.xyz[3 * (int64_t)c[1]],
                       &o.xyz[3 * (int64_t)c[2]], &o.xyz[3 * (int64_t)c[3]]);
  double v = 0.0;
  for (int k = 0; k < 6; ++k)
    v = v + tet_volume6(&o.xyz[3 * (int64_t)c[HEX_KUHN[k][0]]], &o.xyz[3 * (int64_t)c[HEX_KUHN[k][1]]],
                        &o.xyz[3 * (int64_t)c[HEX_KUHN[k][2]]], &o.xyz[3 * (int64_t)c[HEX_KUHN[k][3]]]);
  return v;
}

bool build_mesh(Oracle *o, const int32_t *face_boundary)
{
  const int64_t ne = o->n_elem;
  o->face_off.assign(ne + 1, 0);
  for (int64_t e = 0; e < ne; ++e) {
    if (o->type[e] != MCW_TET4 && o->type[e] != MCW_HEX8) return fail(o, "unsupported element type");
    int64_t nn = o->off[e + 1] - o->off[e];
    if (nn != o->type[e]) return fail(o, "connectivity size does not match element type");
    o->face_off[e + 1] = o->face_off[e] + n_sides(o->type[e]);
  }
  const int64_t nf = o->face_off[ne];
  o->plane.assign(4 * nf, 0.0);
  o->nbr.assign(nf, -1);
  o->nbr_side.assign(nf, -1);

  // match faces by their sorted node tuples
  struct Key {
    int32_t n[4];
    int64_t slot;
  };
  std::vector<Key> keys(nf);
  for (int64_t e = 0; e < ne; ++e)
    for (int s = 0; s < n_sides(o->type[e]); ++s) {
      int32_t fn[4];
      int nn;
      face_nodes(*o, e, s, fn, &nn);
      std::sort(fn, fn + nn);
      Key &k = keys[o->face_off[e] + s];
      for (int q = 0; q < 4; ++q) k.n[q] = q < nn ? fn[q] : -1;
      k.slot = o->face_off[e] + s;
    }
  std::vector<Key> sorted = keys;
  std::sort(sorted.begin(), sorted.end(), [](const Key &a, const Key &b) {
    for (int q = 0; q < 4; ++q)
      if (a.n[q] != b.n[q]) return a.n[q] < b.n[q];
    return a.slot < b.slot;
  });
  auto same = [](const Key &a, const Key &b) {
    return a.n[0] == b.n[0] && a.n[1] == b.n[1] && a.n[2] == b.n[2] && a.n[3] == b.n[3];
  };
  // slot -> element lookup
  std::vector<int64_t> slot_elem(nf);
  for (int64_t e = 0; e < ne; ++e)
    for (int64_t s = o->face_off[e]; s < o->face_off[e + 1]; ++s) slot_elem[s] = e;
  std::vector<int64_t> open; // sides with no side of equal node set
  for (size_t i = 0; i < sorted.size();) {
    size_t j = i + 1;
    while (j < sorted.size() && same(sorted[i], sorted[j])) ++j;
    if (j - i > 2) return fail(o, "non-manifold face shared by more than two elements");
    if (j - i == 2) {
      int64_t a = sorted[i].slot, b = sorted[i + 1].slot;
      int64_t ea = slot_elem[a], eb = slot_elem[b];
      o->nbr[a] = eb;
      o->nbr_side[a] = (int32_t)(b - o->face_off[eb]);
      o->nbr[b] = ea;
      o->nbr_side[b] = (int32_t)(a - o->face_off[ea]);
    } else {
      open.push_back(sorted[i].slot);
    }
    i = j;
  }
  // split interfaces: a quad side (q0 q1 q2 q3, libMesh order) whose two
  // halves along diagonal q0-q2 or q1-q3 are open triangle sides of tets
  o->split_of.assign(nf, -1);
  o->splits.clear();
  std::map<std::array<int32_t, 3>, int64_t> tri;
  auto key3 = [](int32_t a, int32_t b, int32_t c) {
    std::array<int32_t, 3> k = {a, b, c};
    std::sort(k.begin(), k.end());
    return k;
  };
  for (int64_t a : open)
    if (o->type[slot_elem[a]] == MCW_TET4) {
      int32_t fn[4];
      int nn;
      face_nodes(*o, slot_elem[a], (int)(a - o->face_off[slot_elem[a]]), fn, &nn);
      tri[key3(fn[0], fn[1], fn[2])] = a;
    }
  std::vector<std::array<int32_t, 2>> split_diag;
  std::vector<std::array<int64_t, 3>> split_slots; // quad, triangle 0, triangle 1
  for (int64_t a : open) {
    const int64_t e = slot_elem[a];
    if (o->type[e] != MCW_HEX8) continue;
    int32_t q[4];
    int nn;
    face_nodes(*o, e, (int)(a - o->face_off[e]), q, &nn);
    for (int d = 0; d < 2; ++d) {
      auto t0 = tri.find(key3(q[d], q[d + 1], q[d + 2]));
      auto t1 = tri.find(key3(q[d + 2], q[(d + 3) & 3], q[d]));
      if (t0 == tri.end() || t1 == tri.end()) continue;
      const int64_t s0 = t0->second, s1 = t1->second;
      tri.erase(t0);
      tri.erase(tri.find(key3(q[d + 2], q[(d + 3) & 3], q[d])));
      o->split_of[a] = (int64_t)split_slots.size();
      split_slots.push_back({a, s0, s1});
      split_diag.push_back({std::min(q[d], q[d + 2]), std::max(q[d], q[d + 2])});
      for (int64_t ts : {s0, s1}) {
        o->nbr[ts] = e;
        o->nbr_side[ts] = (int32_t)(a - o->face_off[e]);
      }
      break;
    }
  }
  for (int64_t a : open) {
    if (o->split_of[a] >= 0 || o->nbr[a] >= 0) continue;
    int32_t bid = face_boundary ? face_boundary[a] : 0;
    if (bid < 0) return fail(o, "external side without a boundary id");
    o->nbr[a] = -1 - (int64_t)bid;
  }
  // outward planes: canonical plane, flipped to point away from the centroid
  for (int64_t e = 0; e < ne; ++e) {
    const int32_t *c = &o->conn[o->off[e]];
    int nn = o->type[e];
    double cx = 0, cy = 0, cz = 0;
    for (int k = 0; k < nn; ++k) {
      cx = cx + o->xyz[3 * (int64_t)c[k]];
      cy = cy + o->xyz[3 * (int64_t)c[k] + 1];
      cz = cz + o->xyz[3 * (int64_t)c[k] + 2];
    }
    cx = cx / nn;
    cy = cy / nn;
    cz = cz / nn;
    for (int s = 0; s < n_sides(o->type[e]); ++s) {
      int32_t fn[4];
      int fnn;
      face_nodes(*o, e, s, fn, &fnn);
      double *p = &o->plane[4 * (o->face_off[e] + s)];
      canonical_plane(*o, fn, fnn, p);
      double side = ((p[0] * cx + p[1] * cy) + p[2] * cz) - p[3];
      if (side > 0.0) {
        p[0] = -p[0];
        p[1] = -p[1];
        p[2] = -p[2];
        p[3] = -p[3];
      }
    }
  }
  // split planes: through the diagonal (lower node id first), normal to the
  // quad's outward normal n: m = (n x d)/|n x d|, c = m.x_a; the tet whose
  // off-diagonal vertex has m.v - c <= 0 takes exit points with m.x - c <= 0
  for (size_t k = 0; k < split_slots.size(); ++k) {
    const double *n = &o->plane[4 * split_slots[k][0]];
    const double *xa = &o->xyz[3 * (int64_t)split_diag[k][0]];
    const double *xb = &o->xyz[3 * (int64_t)split_diag[k][1]];
    const double dx = xb[0] - xa[0], dy = xb[1] - xa[1], dz = xb[2] - xa[2];
    double mx = n[1] * dz - n[2] * dy;
    double my = n[2] * dx - n[0] * dz;
    double mz = n[0] * dy - n[1] * dx;
    const double len = std::sqrt((mx * mx + my * my) + mz * mz);
    mx = mx / len;
    my = my / len;
    mz = mz / len;
    Oracle::Split S;
    S.m[0] = mx;
    S.m[1] = my;
    S.m[2] = mz;
    S.m[3] = (mx * xa[0] + my * xa[1]) + mz * xa[2];
    const int64_t t0 = split_slots[k][1], t1 = split_slots[k][2];
    int32_t fn[4];
    int nn;
    face_nodes(*o, slot_elem[t0], (int)(t0 - o->face_off[slot_elem[t0]]), fn, &nn);
    int32_t v = fn[0];
    for (int q = 0; q < 3; ++q)
      if (fn[q] != split_diag[k][0] && fn[q] != split_diag[k][1]) v = fn[q];
    const double *xv = &o->xyz[3 * (int64_t)v];
    const double sv = ((mx * xv[0] + my * xv[1]) + mz * xv[2]) - S.m[3];
    const int64_t lo = sv <= 0.0 ? t0 : t1, hi = sv <= 0.0 ? t1 : t0;
    S.elem[0] = slot_elem[lo];
    S.side[0] = (int32_t)(lo - o->face_off[slot_elem[lo]]);
    S.elem[1] = slot_elem[hi];
    S.side[1] = (int32_t)(hi - o->face_off[slot_elem[hi]]);
    o->splits.push_back(S);
  }
  // volume CDF (sequential sum in element order)
  o->cdf.assign(ne, 0.0);
  double acc = 0.0;
  for (int64_t e = 0; e < ne; ++e) {
    const double v = elem_volume6(*o, e);
    if (!(v > 0.0)) return fail(o, "degenerate element (zero volume)");
    acc = acc + v;
    o->cdf[e] = acc;
  }
  return true;
}

// Lowest-index element whose face planes all contain p within tol.
int64_t locate(const Oracle &o, const double *p, double tol)
{
  for (int64_t e = 0; e < o.n_elem; ++e) {
    bool in = true;
    for (int64_t s = o.face_off[e]; s < o.face_off[e + 1] && in; ++s) {
      const double *q = &o.plane[4 * s];
      double v = ((q[0] * p[0] + q[1] * p[1]) + q[2] * p[2]) - q[3];
      if (v > tol) in = false;
    }
    if (in) return e;
  }
  return -1;
}

// ---------------------------------------------------------------- physics

struct Micro {
  double t, a, s, f;
};

// largest j with E_j <= E, clamped to [0, n-2]; lin-lin interpolation
void micro_xs(const Nuclide &n, double E, Micro *m)
{
  const int64_t np = (int64_t)n.E.size();
  int64_t j;
  double f;
  if (!(E > n.E[0])) {
    j = 0;
    f = 0.0;
  } else if (E >= n.E[np - 1]) {
    j = np - 2;
    f = 1.0;
  } else {
    int64_t lo = 0, hi = np - 1; // E_lo <= E < E_hi
    while (hi - lo > 1) {
      int64_t mid = lo + (hi - lo) / 2;
      if (n.E[mid] <= E) lo = mid;
      else hi = mid;
    }
    j = lo;
    f = (E - n.E[j]) / (n.E[j + 1] - n.E[j]);
  }
  double g = 1.0 - f;
  m->t = g * n.t[j] + f * n.t[j + 1];
  m->a = g * n.a[j] + f * n.a[j + 1];
  m->s = g * n.s[j] + f * n.s[j + 1];
  m->f = n.f.empty() ? 0.0 : g * n.f[j] + f * n.f[j + 1];
}

struct Macro {
  double t, a, s, f, kf; // kf: kappa-fission, sum of N kappa sigma_f
};

// sqrt(kT) of an element [sqrt(eV)]; -1 when no temperatures are set.
inline double elem_sqrtkT(const Oracle &o, int64_t elem)
{
  return o.elem_T.empty() ? -1.0 : std::sqrt(K_BOLTZMANN * o.elem_T[elem]);
}

// The table of nuclide q read at sqrt(kT) = skT: OpenMC's "nearest" method
// over the nuclide's temperatures in ascending order, strict improvement, so
// a tie keeps the lower temperature.  skT < 0: the first table given.
const Nuclide &nuc_at(const Oracle &o, int q, double skT)
{
  if (skT < 0.0) return o.nucs[q];
  const std::vector<const Nuclide *> &tb = o.tables[q];
  const double kT = skT * skT;
  size_t it = 0;
  double dmin = INFINITY;
  for (size_t t = 0; t < tb.size(); ++t) {
    const double d = std::fabs(K_BOLTZMANN * tb[t]->temperature - kT);
    if (d < dmin) {
      dmin = d;
      it = t;
    }
  }
  return *tb[it];
}

void macro_xs(const Oracle &o, int mat, double E, double skT, Macro *M)
{
  M->t = M->a = M->s = M->f = M->kf = 0.0;
  if (mat < 0) return;
  const Material &m = o.mats[mat];
  for (size_t i = 0; i < m.nuc.size(); ++i) {
    Micro u;
    const Nuclide &n = nuc_at(o, m.nuc[i], skT);
    micro_xs(n, E, &u);
    M->kf = M->kf + m.dens[i] * (n.kappa * u.f);
    M->t = M->t + m.dens[i] * u.t;
    M->a = M->a + m.dens[i] * u.a;
    M->s = M->s + m.dens[i] * u.s;
    M->f = M->f + m.dens[i] * u.f;
  }
}

double macro_total(const Oracle &o, int mat, double E, double skT)
{
  if (mat < 0) return 0.0;
  const Material &m = o.mats[mat];
  double t = 0.0;
  for (size_t i = 0; i < m.nuc.size(); ++i) {
    Micro u;
    micro_xs(nuc_at(o, m.nuc[i], skT), E, &u);
    t = t + m.dens[i] * u.t;
  }
  return t;
}

// OpenMC sample_nuclide (physics.cpp): cutoff = xi Sigma_t, then the first
// nuclide (material order) whose running sum of N_i sigma_t,i reaches it; the
// last one when rounding leaves the sum short.  Returns the position in the
// material's list.
int sample_nuclide(const Oracle &o, const Material &m, double E, double skT, double sigt, uint64_t *seed)
{
  const double cutoff = orc_prn(seed) * sigt;
  double prob = 0.0;
  for (size_t q = 0; q < m.nuc.size(); ++q) {
    Micro mu;
    micro_xs(nuc_at(o, m.nuc[q], skT), E, &mu);
    prob = prob + m.dens[q] * mu.t;
    if (prob >= cutoff) return (int)q;
  }
  return (int)m.nuc.size() - 1;
}

// OpenMC rotate_angle with phi = 2 pi xi
void rotate_angle(double *u, double *v, double *w, double mu, double xi)
{
  double sinphi, cosphi;
  orc_sincos_2pi(xi, &sinphi, &cosphi);
  double u0 = *u, v0 = *v, w0 = *w;
  double a = std::sqrt(std::max(0.0, 1.0 - mu * mu));
  double b = std::sqrt(std::max(0.0, 1.0 - w0 * w0));
  if (b > 1e-10) {
    *u = mu * u0 + a * (u0 * w0 * cosphi - v0 * sinphi) / b;
    *v = mu * v0 + a * (v0 * w0 * cosphi + u0 * sinphi) / b;
    *w = mu * w0 - a * b * cosphi;
  } else {
    b = std::sqrt(1.0 - v0 * v0);
    *u = mu * u0 + a * (u0 * v0 * cosphi + w0 * sinphi) / b;
    *v = mu * v0 - a * b * cosphi;
    *w = mu * w0 + a * (v0 * w0 * cosphi - u0 * sinphi) / b;
  }
}

// Elastic scattering off a target at rest, isotropic in the CM frame:
// E' = E (A^2 + 2 A mu + 1)/(A+1)^2, mu_lab = (1 + A mu)/sqrt(A^2 + 2 A mu + 1).
void elastic_kinematics(double E, double A, double mu_cm, double *Ep, double *mu_lab)
{
  double q2 = (A * A + 2.0 * A * mu_cm) + 1.0;
  *Ep = E * q2 / ((A + 1.0) * (A + 1.0));
  *mu_lab = (1.0 + A * mu_cm) / std::sqrt(q2);
}

// ---------------------------------------------------------------- source

struct Site {
  double x, y, z, u, v, w, E;
  int64_t elem;
};

void sample_point_in_tet(const double *p0, const double *p1, const double *p2, const double *p3,
                         uint64_t *seed, double *out)
{
  // Rocchini & Cignoni, "Generating random points in a tetrahedron",
  // Journal of Graphics Tools 5(4), 2000.
  double s = orc_prn(seed), t = orc_prn(seed), u = orc_prn(seed);
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
  for (int k = 0; k < 3; ++k) out[k] = ((a * p0[k] + s * p1[k]) + t * p2[k]) + u * p3[k];
}

// OpenMC watt_spectrum: a Maxwellian w = -a (ln r1 + ln r2 cos^2(pi r3 / 2))
// plus a^2 b / 4 + (2 r4 - 1) sqrt(a^2 b w), resampled while at or above the
// top of the energy grid (after 1000 tries the grid minimum is taken).
double sample_watt(const Oracle &o, double a, double b, uint64_t *seed)
{
  for (int it = 0;; ++it) {
    double r1 = orc_prn(seed), r2 = orc_prn(seed), r3 = orc_prn(seed);
    double c = orc_cos_halfpi(r3);
    double wmax = -a * (orc_log(r1) + orc_log(r2) * c * c);
    double E = wmax + 0.25 * a * a * b + (-1.0 + 2.0 * orc_prn(seed)) * std::sqrt(a * a * b * wmax);
    if (E < o.energy_max || it >= 1000) return E < o.energy_max ? E : o.energy_min;
  }
}

bool sample_source(const Oracle &o, uint64_t *seed, Site *st)
{
  const mcw_source &S = o.src;
  double p[3];
  if (S.space == MCW_SPACE_POINT) {
    p[0] = S.point[0];
    p[1] = S.point[1];
    p[2] = S.point[2];
    st->elem = o.src_elem;
  } else if (S.space == MCW_SPACE_MESH_VOLUME) {
    double target = orc_prn(seed) * o.cdf[o.n_elem - 1];
    int64_t e = std::upper_bound(o.cdf.begin(), o.cdf.end(), target) - o.cdf.begin();
    if (e >= o.n_elem) e = o.n_elem - 1;
    const int32_t *c = &o.conn[o.off[e]];
    if (o.type[e] == MCW_TET4) {
      sample_point_in_tet(&o.xyz[3 * (int64_t)c[0]], &o.xyz[3 * (int64_t)c[1]], &o.xyz[3 * (int64_t)c[2]],
                          &o.xyz[3 * (int64_t)c[3]], seed, p);
    } else {
      double vols[6], tot = 0.0;
      for (int k = 0; k < 6; ++k) {
        tot = tot + tet_volume6(&o.xyz[3 * (int64_t)c[HEX_KUHN[k][0]]], &o.xyz[3 * (int64_t)c[HEX_KUHN[k][1]]],
                                &o.xyz[3 * (int64_t)c[HEX_KUHN[k][2]]], &o.xyz[3 * (int64_t)c[HEX_KUHN[k][3]]]);
        vols[k] = tot;
      }
      double tt = orc_prn(seed) * tot;
      int k = 0;
      while (k < 5 && !(vols[k] > tt)) ++k;
      sample_point_in_tet(&o.xyz[3 * (int64_t)c[HEX_KUHN[k][0]]], &o.xyz[3 * (int64_t)c[HEX_KUHN[k][1]]],
                          &o.xyz[3 * (int64_t)c[HEX_KUHN[k][2]]], &o.xyz[3 * (int64_t)c[HEX_KUHN[k][3]]], seed,
                          p);
    }
    st->elem = e;
  } else if (S.space == MCW_SPACE_BOX) {
    // uniform in the box, resampled while outside the mesh (OpenMC rejects
    // source sites that are not in a cell); the element is the lowest-index
    // one containing the point
    int64_t e = -1;
    for (int it = 0; it < 1000 && e < 0; ++it) {
      for (int k = 0; k < 3; ++k) p[k] = S.box_lo[k] + orc_prn(seed) * (S.box_hi[k] - S.box_lo[k]);
      e = locate(o, p, 1e-12);
      // OpenMC IndependentSource::sample, only_fissionable: the site's
      // material must hold a fissionable nuclide
      if (e >= 0 && S.only_fissionable) {
        bool fis = false;
        if (o.elem_mat[e] >= 0)
          for (int q : o.mats[o.elem_mat[e]].nuc) fis |= o.nucs[q].fissionable;
        if (!fis) e = -1;
      }
    }
    st->elem = e;
  } else {
    return false;
  }
  st->x = p[0];
  st->y = p[1];
  st->z = p[2];

  if (S.angle == MCW_ANGLE_ISOTROPIC) {
    double xi_phi = orc_prn(seed);
    double mu = 2.0 * orc_prn(seed) - 1.0;
    double s, c;
    orc_sincos_2pi(xi_phi, &s, &c);
    double r = std::sqrt(1.0 - mu * mu);
    st->u = mu;
    st->v = r * c;
    st->w = r * s;
  } else if (S.angle == MCW_ANGLE_LINEAR_MU) {
    // p(mu) ~ 1 + beta mu about the axis: mu solves F(mu) = xi, i.e.
    // (beta/2) mu^2 + mu + c = 0 with c = (1 - beta/2) - 2 xi, taken as the
    // root -2c / (1 + sqrt(1 - 2 beta c)) (no cancellation, beta = 0 included);
    // the azimuth is 2 pi xi' about the axis (rotate_angle)
    const double cq = (1.0 - 0.5 * S.beta) - 2.0 * orc_prn(seed);
    double mu = -2.0 * cq / (1.0 + std::sqrt(1.0 - 2.0 * S.beta * cq));
    if (mu > 1.0) mu = 1.0;
    if (mu < -1.0) mu = -1.0;
    double u = S.axis[0], v = S.axis[1], w = S.axis[2];
    rotate_angle(&u, &v, &w, mu, orc_prn(seed));
    st->u = u;
    st->v = v;
    st->w = w;
  } else {
    return false;
  }

  st->E = S.energy == MCW_ENERGY_MONO ? S.E0 : sample_watt(o, S.watt_a, S.watt_b, seed);
  return true;
}

// ---------------------------------------------------------------- tallies

struct Scratch {
  std::vector<std::vector<double>> value; // per tally
};

void score(const Oracle &o, const Tally &t, std::vector<double> &val, int64_t elem, double E, double wfl,
           int mat, double skT)
{
  int64_t fi = 0;
  for (size_t k = 0; k < t.filters.size(); ++k) {
    const Filter &f = t.filters[k];
    int64_t bin;
    if (f.type == MCW_FILTER_CELL || f.type == MCW_FILTER_UNIVERSE) {
      bin = f.elem_bin[elem];
    } else if (f.type == MCW_FILTER_ENERGY) {
      const std::vector<double> &ed = f.edges;
      if (E < ed.front() || E > ed.back()) return;
      bin = std::upper_bound(ed.begin(), ed.end(), E) - ed.begin() - 1;
      if (bin > (int64_t)ed.size() - 2) bin = (int64_t)ed.size() - 2;
    } else {
      bin = 0;
    }
    if (bin < 0) return;
    fi += bin * t.strides[k];
  }
  double *row = &val[fi * t.n_score_bins];
  const int nsc = (int)t.scores.size();
  for (size_t ni = 0; ni < t.nuclides.size(); ++ni) {
    int nuc = t.nuclides[ni];
    Macro Ma = {0, 0, 0, 0, 0};
    bool need_xs = false;
    for (int sc : t.scores) need_xs |= sc != MCW_SCORE_FLUX;
    if (nuc < 0) {
      if (need_xs) macro_xs(o, mat, E, skT, &Ma);
    } else if (mat >= 0) {
      // OpenMC 0.13 score_tracklength_tally / score_collision_tally: a
      // nuclide the material does not hold scores nothing in its bin
      const Material &m = o.mats[mat];
      size_t q = 0;
      while (q < m.nuc.size() && m.nuc[q] != nuc) ++q;
      if (q == m.nuc.size()) continue;
      if (need_xs) {
        Micro mu = {0, 0, 0, 0};
        const double dens = m.dens[q];
        micro_xs(nuc_at(o, nuc, skT), E, &mu);
        Ma.t = dens * mu.t;
        Ma.a = dens * mu.a;
        Ma.s = dens * mu.s;
        Ma.f = dens * mu.f;
        Ma.kf = dens * (o.nucs[nuc].kappa * mu.f);
      }
    }
    for (int si = 0; si < nsc; ++si) {
      double m;
      switch (t.scores[si]) {
        case MCW_SCORE_FLUX: m = 1.0; break;
        case MCW_SCORE_TOTAL: m = Ma.t; break;
        case MCW_SCORE_ABSORPTION: m = Ma.a; break;
        case MCW_SCORE_SCATTER: m = Ma.s; break;
        case MCW_SCORE_FISSION: m = Ma.f; break;
        default: m = Ma.kf; break;
      }
      row[ni * nsc + si] += wfl * m;
    }
  }
}

// ---------------------------------------------------------------- transport

struct Counters {
  int64_t segments = 0, collisions = 0, absorptions = 0, leaks = 0, cutoffs = 0, lost = 0,
          reflections = 0, hex_segments = 0;
  double track = 0.0;
};

struct TraceSink {
  std::vector<int32_t> *elems = nullptr;
  std::vector<int32_t> *events = nullptr;
  std::vector<double> *lengths = nullptr;
};

// One history's part of an eigenvalue batch: the keff its fission yield is
// divided by, where its fission sites go, and its three keff estimator sums.
struct EigenHistory {
  double keff = 1.0;
  bool tally = true;              // inactive batches score no tallies
  const mcw_site *source = nullptr; // batch > 1: the history's source site
  std::vector<mcw_site> *sites = nullptr;
  double k_tl = 0.0, k_col = 0.0, k_abs = 0.0;
};

// Macroscopic total and nu-fission at (mat, E).
double macro_total_nf(const Oracle &o, int mat, double E, double skT, double *nusf)
{
  *nusf = 0.0;
  if (mat < 0) return 0.0;
  const Material &m = o.mats[mat];
  double t = 0.0, nf = 0.0;
  for (size_t i = 0; i < m.nuc.size(); ++i) {
    const Nuclide &n = nuc_at(o, m.nuc[i], skT);
    Micro u;
    micro_xs(n, E, &u);
    t = t + m.dens[i] * u.t;
    nf = nf + m.dens[i] * ((n.nu0 + n.nu1 * E) * u.f);
  }
  *nusf = nf;
  return t;
}

void transport_history(const Oracle &o, int64_t batch, int64_t n_particles, int64_t i, Scratch &sc,
                       Counters &cnt, mcw_history *rec, TraceSink *trace, EigenHistory *eg = nullptr)
{
  const int64_t id = (batch - 1) * n_particles + i + 1;
  const uint64_t master = o.opt.master_seed;
  const bool rt = o.opt.seed_roundtrip != 0;
  const int64_t max_seg = o.opt.max_segments > 0 ? o.opt.max_segments : (int64_t)1 << 20;
  uint64_t seed = orc_init_seed(id, 0, master);
  Site st;
  if (eg && eg->source) {
    // a fission site of the previous batch (OpenMC initialize_history takes
    // source_bank[i]; only the tracking stream is seeded here)
    const mcw_site &q = *eg->source;
    st.x = q.r[0]; st.y = q.r[1]; st.z = q.r[2];
    st.u = q.u[0]; st.v = q.u[1]; st.w = q.u[2];
    st.E = q.E;
    st.elem = q.element;
  } else {
    uint64_t sseed = orc_init_seed(id, 1, master);
    sample_source(o, &sseed, &st);
  }
  if (rt) seed = orc_seed_roundtrip(seed);
  const bool tally_on = !eg || eg->tally;
  double nusf = 0.0;
  int32_t progeny = 0;

  double x = st.x, y = st.y, z = st.z, u = st.u, v = st.v, w = st.w, E = st.E;
  const double wgt = 1.0;
  int64_t elem = st.elem;
  int inface = -1;
  int fate = MCW_FATE_ALIVE;
  int32_t nseg = 0, ncoll = 0;
  uint64_t h = 0xcbf29ce484222325ULL;
  int last_mat = -2;
  double last_E = -1.0, last_skT = -2.0, sigt = 0.0;
  const std::vector<Tally> &T = o.tallies;

  if (elem < 0) fate = MCW_FATE_LOST;
  while (fate == MCW_FATE_ALIVE) {
    if (nseg >= max_seg) {
      fate = MCW_FATE_LOST;
      h = digest_step(h, elem, EV_LOST);
      break;
    }
    const int64_t seg_elem = elem;
    const int mat = o.elem_mat[elem];
    // OpenMC recomputes the material xs when the material, E or sqrt(kT) changed
    const double skT = elem_sqrtkT(o, elem);
    if (mat != last_mat || E != last_E || skT != last_skT) {
      sigt = eg ? macro_total_nf(o, mat, E, skT, &nusf) : macro_total(o, mat, E, skT);
      last_mat = mat;
      last_E = E;
      last_skT = skT;
    }
    // exit side: smallest non-negative distance over outgoing planes,
    // the incoming side excluded, ties to the lowest side index
    const int64_t f0 = o.face_off[elem];
    const int ns = (int)(o.face_off[elem + 1] - f0);
    double L = INFINITY;
    int fx = -1;
    for (int s = 0; s < ns; ++s) {
      if (s == inface) continue;
      const double *q = &o.plane[4 * (f0 + s)];
      double den = (q[0] * u + q[1] * v) + q[2] * w;
      if (!(den > 0.0)) continue;
      double num = q[3] - ((q[0] * x + q[1] * y) + q[2] * z);
      double t = num / den;
      if (!(t > 0.0)) t = 0.0;
      if (t < L) {
        L = t;
        fx = s;
      }
    }
    if (fx < 0) {
      fate = MCW_FATE_LOST;
      h = digest_step(h, elem, EV_LOST);
      break;
    }
    // one collision-distance draw per segment (CollisionKernel.C:150)
    double xi = orc_prn(&seed);
    double d = sigt > 0.0 ? -orc_log(xi) / sigt : INFINITY;
    double track = (d < L) ? d : L;
    ++nseg;
    if (o.type[elem] == MCW_HEX8) ++cnt.hex_segments;
    cnt.track += wgt * track;
    // keff track-length estimator (CollisionKernel.C:152-154)
    if (eg) eg->k_tl = eg->k_tl + (wgt * track) * nusf;
    if (tally_on)
      for (size_t k = 0; k < T.size(); ++k)
        if (T[k].estimator == MCW_EST_TRACKLENGTH) score(o, T[k], sc.value[k], elem, E, wgt * track, mat, skT);

    int ev;
    if (d > L) {
      x = x + L * u;
      y = y + L * v;
      z = z + L * w;
      int64_t nb = o.nbr[f0 + fx];
      int nside = o.nbr_side[f0 + fx];
      const int64_t sp = o.split_of[f0 + fx];
      if (sp >= 0) { // split interface: the tet on the exit point's side of the diagonal
        const Oracle::Split &S = o.splits[sp];
        const double sg = ((S.m[0] * x + S.m[1] * y) + S.m[2] * z) - S.m[3];
        const int k = sg <= 0.0 ? 0 : 1;
        nb = S.elem[k];
        nside = S.side[k];
      }
      if (nb >= 0) {
        ev = EV_CROSS;
        h = digest_step(h, elem, ev);
        inface = nside;
        elem = nb;
      } else {
        int bid = (int)(-1 - nb);
        auto it = o.bc.find(bid);
        int kind = it == o.bc.end() ? MCW_BC_VACUUM : it->second;
        if (kind == MCW_BC_REFLECT) {
          ev = EV_REFLECT;
          h = digest_step(h, elem, ev);
          const double *q = &o.plane[4 * (f0 + fx)];
          double dot = (u * q[0] + v * q[1]) + w * q[2];
          u = u - 2.0 * dot * q[0];
          v = v - 2.0 * dot * q[1];
          w = w - 2.0 * dot * q[2];
          inface = fx;
          ++cnt.reflections;
        } else {
          ev = EV_LEAK;
          h = digest_step(h, elem, ev);
          fate = MCW_FATE_LEAKED;
          ++cnt.leaks;
        }
      }
    } else {
      // back-step from the exit point (CollisionKernel.C:172-175)
      double back = L - d;
      x = (x + L * u) - back * u;
      y = (y + L * v) - back * v;
      z = (z + L * w) - back * w;
      ++ncoll;
      if (tally_on)
        for (size_t k = 0; k < T.size(); ++k)
          if (T[k].estimator == MCW_EST_COLLISION) score(o, T[k], sc.value[k], elem, E, wgt / sigt, mat, skT);
      // keff collision estimator (OpenMC Particle::event_collide)
      if (eg) eg->k_col = eg->k_col + (wgt * nusf) / sigt;
      const Material &m = o.mats[mat];
      const int pick = sample_nuclide(o, m, E, skT, sigt, &seed);
      Micro mu;
      const Nuclide &nuc = nuc_at(o, m.nuc[pick], skT);
      micro_xs(nuc, E, &mu);
      const double nuf = (nuc.nu0 + nuc.nu1 * E) * mu.f;
      if (eg && nuc.fissionable && mu.t > 0.0) {
        // OpenMC create_fission_sites: nu_t = w / keff * nu sigma_f / sigma_t
        // neutrons expected, floor(nu_t) plus one with probability frac(nu_t)
        const double nu_t = (((wgt / eg->keff) * 1.0) * nuf) / mu.t;
        int64_t nu = nu_t < 9.0e15 ? (int64_t)nu_t : (int64_t)9.0e15;
        if (orc_prn(&seed) <= nu_t - (double)nu) ++nu;
        nu = std::min<int64_t>(nu, ((int64_t)1 << 24) - progeny); // progeny index: 24 bits
        // more than the bank holds (3 n_particles) cannot be kept
        for (int64_t q = 0; q < nu && q < 3 * n_particles; ++q) {
          // sample_fission_neutron: isotropic (mu, then phi), the prompt /
          // delayed choice (no delayed data: always prompt), then the energy
          mcw_site f;
          f.r[0] = x; f.r[1] = y; f.r[2] = z;
          const double fmu = 2.0 * orc_prn(&seed) - 1.0;
          double sphi, cphi;
          orc_sincos_2pi(orc_prn(&seed), &sphi, &cphi);
          const double sn = std::sqrt(1.0 - fmu * fmu);
          f.u[0] = fmu;
          f.u[1] = sn * cphi;
          f.u[2] = sn * sphi;
          (void)orc_prn(&seed);
          f.E = sample_watt(o, nuc.watt_a, nuc.watt_b, &seed);
          f.parent = i;
          f.progeny = progeny++;
          f.element = (int32_t)elem;
          eg->sites->push_back(f);
        }
      }
      if (mu.a > orc_prn(&seed) * mu.t) {
        ev = EV_ABSORB;
        fate = MCW_FATE_ABSORBED;
        ++cnt.absorptions;
        // keff absorption estimator (OpenMC absorption())
        if (eg) eg->k_abs = eg->k_abs + (wgt * nuf) / mu.a;
      } else {
        double mu_cm = 2.0 * orc_prn(&seed) - 1.0;
        double Ep, mu_lab;
        elastic_kinematics(E, nuc.awr, mu_cm, &Ep, &mu_lab);
        if (Ep < o.energy_min) {
          ev = EV_CUTOFF;
          fate = MCW_FATE_CUTOFF;
          E = Ep;
          ++cnt.cutoffs;
        } else {
          double xi_phi = orc_prn(&seed);
          rotate_angle(&u, &v, &w, mu_lab, xi_phi);
          E = Ep;
          ev = EV_SCATTER;
        }
      }
      h = digest_step(h, elem, ev);
      inface = -1;
    }
    if (trace) {
      trace->elems->push_back((int32_t)seg_elem);
      trace->events->push_back(ev);
      trace->lengths->push_back(track);
    }
    if (rt) seed = orc_seed_roundtrip(seed);
  }
  cnt.segments += nseg;
  cnt.collisions += ncoll;
  if (fate == MCW_FATE_LOST) ++cnt.lost;
  if (rec) {
    rec->id = id;
    rec->digest = h;
    rec->seed = seed;
    rec->E = E;
    rec->x = x;
    rec->y = y;
    rec->z = z;
    rec->n_segments = nseg;
    rec->n_collisions = ncoll;
    rec->elem = (int32_t)elem;
    rec->fate = fate;
  }
}

// Sum of v in a fixed order that a GPU reduction can follow: blocks of 1024
// consecutive values (missing ones 0), each folded by 256 lanes as
// ((v[t] + v[t+256]) + v[t+512]) + v[t+768], then halved pairwise
// a[t] += a[t+s] for s = 128, 64, ..., 1; repeated over the block sums until
// one value is left.
double tree_sum(std::vector<double> v)
{
  if (v.empty()) return 0.0;
  for (;;) {
    const int64_t n = (int64_t)v.size();
    const int64_t nb = (n + 1023) / 1024;
    std::vector<double> out((size_t)nb);
    for (int64_t b = 0; b < nb; ++b) {
      auto at = [&](int64_t i) { return i < n ? v[(size_t)i] : 0.0; };
      double a[256];
      for (int t = 0; t < 256; ++t) {
        const int64_t i = b * 1024 + t;
        a[t] = ((at(i) + at(i + 256)) + at(i + 512)) + at(i + 768);
      }
      for (int s2 = 128; s2 > 0; s2 >>= 1)
        for (int t = 0; t < s2; ++t) a[t] = a[t] + a[t + s2];
      out[(size_t)b] = a[0];
    }
    if (nb == 1) return out[0];
    v.swap(out);
  }
}

// OpenMCStudy::synchronizeBanks on one rank (OpenMCStudy.C:454-607): every
// banked site is copied floor(n / total) times when fewer than n were banked,
// and once more with probability p = (n mod total, or n) / total, one draw per
// site from the tracking stream of seed id `batch`; then the list is cut to n,
// or filled up with the last sites of the bank.
std::vector<mcw_site> synchronize_bank(const std::vector<mcw_site> &bank, int64_t batch, int64_t n,
                                       uint64_t master)
{
  const int64_t total = (int64_t)bank.size();
  uint64_t seed = orc_init_seed(batch, 0, master);
  const int64_t needed = total < n ? n % total : n;
  const double p_sample = (double)needed / (double)total;
  std::vector<mcw_site> out;
  out.reserve((size_t)n);
  for (int64_t i = 0; i < total; ++i) {
    if (total < n)
      for (int64_t j = 1; j <= n / total; ++j) out.push_back(bank[(size_t)i]);
    if (orc_prn(&seed) < p_sample) out.push_back(bank[(size_t)i]);
  }
  const int64_t fin = (int64_t)out.size();
  if (fin > n) {
    out.resize((size_t)n);
  } else if (fin < n) {
    const int64_t more = n - fin;
    for (int64_t i = 0; i < more; ++i) out.push_back(bank[(size_t)std::max<int64_t>(0, total - more + i)]);
  }
  return out;
}

bool check_ready(Oracle *o)
{
  if (o->n_elem == 0) return fail(o, "no mesh");
  if (o->nucs.empty() || o->mats.empty()) return fail(o, "no materials");
  if (!o->have_source) return fail(o, "no source");
  // element -> material through the block map.  Every mapped id must name a
  // material (CollisionKernel.C:85-90); the material used is the one with
  // that id, or with MCW_MATERIAL_BY_INDEX the list entry id - 1, as
  // `p->material() = _block_to_openmc_materials.at(subdomain) - 1`
  // (CollisionKernel.C:110) indexes OpenMC's materials vector
  o->elem_mat.assign(o->n_elem, -1);
  std::map<int, int> id_to_idx;
  for (size_t i = 0; i < o->mats.size(); ++i) id_to_idx[o->mats[i].id] = (int)i;
  const bool by_index = o->opt.material_index == MCW_MATERIAL_BY_INDEX;
  for (int64_t e = 0; e < o->n_elem; ++e) {
    auto b = o->block_mat.find(o->sub[e]);
    if (b == o->block_mat.end()) continue;
    auto m = id_to_idx.find(b->second);
    if (m == id_to_idx.end()) return fail(o, "Could not find material " + std::to_string(b->second));
    if (by_index) {
      if (b->second < 1 || b->second > (int)o->mats.size())
        return fail(o, "material index " + std::to_string(b->second) + " - 1 is outside the materials");
      o->elem_mat[e] = b->second - 1;
    } else {
      o->elem_mat[e] = m->second;
    }
  }
  // every nuclide's tables in ascending temperature (OpenMC kTs_ order)
  o->nuc_temps.resize(o->nucs.size());
  o->tables.assign(o->nucs.size(), {});
  for (size_t q = 0; q < o->nucs.size(); ++q) {
    o->tables[q].push_back(&o->nucs[q]);
    for (auto &t : o->nuc_temps[q]) o->tables[q].push_back(&t);
    std::stable_sort(o->tables[q].begin(), o->tables[q].end(),
                     [](const Nuclide *a, const Nuclide *b) { return a->temperature < b->temperature; });
  }
  if (!o->elem_T.empty()) {
    if ((int64_t)o->elem_T.size() != o->n_elem) return fail(o, "one temperature per element");
    for (int64_t e = 0; e < o->n_elem; ++e) {
      if (o->elem_mat[e] < 0) continue;
      for (int q : o->mats[o->elem_mat[e]].nuc)
        if (std::fabs(nuc_at(*o, q, elem_sqrtkT(*o, e)).temperature - o->elem_T[e]) > o->T_tol)
          return fail(o, "no cross sections for nuclide " + std::to_string(q) + " near element " +
                             std::to_string(e) + "'s temperature");
    }
  }
  o->energy_min = INFINITY;
  o->energy_max = -INFINITY;
  for (auto &tb : o->tables)
    for (const Nuclide *n : tb) {
      o->energy_min = std::min(o->energy_min, n->E.front());
      o->energy_max = std::max(o->energy_max, n->E.back());
    }
  // a shared grid range is assumed by the energy cutoff
  return true;
}

} // namespace

// ======================================================================== C API
extern "C" {

void *orc_create(const mcw_options *opt)
{
  Oracle *o = new Oracle();
  mcw_options d;
  memset(&d, 0, sizeof d);
  d.master_seed = 1;
  o->opt = opt ? *opt : d;
  if (o->opt.master_seed == 0) o->opt.master_seed = 1;
  return o;
}

void orc_destroy(void *h) { delete (Oracle *)h; }

const char *orc_last_error(void *h) { return ((Oracle *)h)->err.c_str(); }

int orc_mesh_set(void *h, int64_t n_nodes, const double *xyz, int64_t n_elem, const uint8_t *elem_type,
                 const int64_t *conn_offsets, const int32_t *conn, const int32_t *subdomain,
                 const int32_t *face_boundary)
{
  Oracle *o = (Oracle *)h;
  o->n_nodes = n_nodes;
  o->n_elem = n_elem;
  o->xyz.assign(xyz, xyz + 3 * n_nodes);
  o->type.assign(elem_type, elem_type + n_elem);
  o->off.assign(conn_offsets, conn_offsets + n_elem + 1);
  o->conn.assign(conn, conn + conn_offsets[n_elem]);
  o->sub.assign(subdomain, subdomain + n_elem);
  return build_mesh(o, face_boundary) ? 0 : 1;
}

int orc_bc_set(void *h, int32_t bid, int32_t kind)
{
  ((Oracle *)h)->bc[bid] = kind;
  return 0;
}

int orc_xs_set(void *h, int32_t n, const mcw_nuclide *nuc)
{
  Oracle *o = (Oracle *)h;
  o->nucs.clear();
  for (int i = 0; i < n; ++i) {
    Nuclide q;
    q.awr = nuc[i].awr;
    int np = nuc[i].n_points;
    if (np < 2) return fail(o, "nuclide grid needs at least 2 points") ? 0 : 1;
    q.E.assign(nuc[i].energy, nuc[i].energy + np);
    q.t.assign(nuc[i].total, nuc[i].total + np);
    q.a.assign(nuc[i].absorption, nuc[i].absorption + np);
    q.s.assign(nuc[i].scatter, nuc[i].scatter + np);
    if (nuc[i].fission) {
      q.f.assign(nuc[i].fission, nuc[i].fission + np);
      q.fissionable = true;
    }
    q.nu0 = nuc[i].nu[0];
    q.nu1 = nuc[i].nu[1];
    q.watt_a = nuc[i].watt[0];
    q.watt_b = nuc[i].watt[1];
    q.kappa = nuc[i].kappa;
    if (nuc[i].temperature > 0.0) q.temperature = nuc[i].temperature;
    o->nucs.push_back(std::move(q));
  }
  o->nuc_temps.assign(o->nucs.size(), {});
  return 0;
}

/* A further table of nuclide `q` at t->temperature; awr, nu and the Watt
 * spectrum stay the first table's. */
int orc_xs_add_temperature(void *h, int32_t q, const mcw_nuclide *t)
{
  Oracle *o = (Oracle *)h;
  if (q < 0 || q >= (int32_t)o->nucs.size()) return fail(o, "no such nuclide") ? 0 : 1;
  const int np = t->n_points;
  if (np < 2) return fail(o, "nuclide grid needs at least 2 points") ? 0 : 1;
  Nuclide n = o->nucs[q];
  n.temperature = t->temperature > 0.0 ? t->temperature : 293.6;
  n.E.assign(t->energy, t->energy + np);
  n.t.assign(t->total, t->total + np);
  n.a.assign(t->absorption, t->absorption + np);
  n.s.assign(t->scatter, t->scatter + np);
  n.f.clear();
  if (t->fission) n.f.assign(t->fission, t->fission + np);
  if (n.f.empty() == o->nucs[q].fissionable) return fail(o, "fission data on some temperatures only") ? 0 : 1;
  o->nuc_temps[q].push_back(std::move(n));
  return 0;
}

/* Element temperatures [K] (n = 0: none, every element reads first tables). */
int orc_temperature_set(void *h, int64_t n, const double *kelvin, double tolerance)
{
  Oracle *o = (Oracle *)h;
  o->elem_T.assign(kelvin, kelvin + n);
  o->T_tol = tolerance;
  return 0;
}

int orc_materials_set(void *h, int32_t n, const mcw_material *m)
{
  Oracle *o = (Oracle *)h;
  o->mats.clear();
  for (int i = 0; i < n; ++i) {
    Material q;
    q.id = m[i].id;
    q.nuc.assign(m[i].nuclide, m[i].nuclide + m[i].n_nuclides);
    q.dens.assign(m[i].density, m[i].density + m[i].n_nuclides);
    o->mats.push_back(std::move(q));
  }
  return 0;
}

int orc_block_material_map(void *h, int32_t n, const int32_t *blocks, const int32_t *materials)
{
  Oracle *o = (Oracle *)h;
  o->block_mat.clear();
  for (int i = 0; i < n; ++i) o->block_mat[blocks[i]] = materials[i];
  return 0;
}

int orc_source_set(void *h, const mcw_source *s)
{
  Oracle *o = (Oracle *)h;
  o->src = *s;
  o->have_source = true;
  o->src_elem = -1;
  if (s->space == MCW_SPACE_POINT) {
    o->src_elem = locate(*o, s->point, 1e-12);
    if (o->src_elem < 0) return fail(o, "point source outside the mesh") ? 0 : 1;
  }
  if (s->angle == MCW_ANGLE_LINEAR_MU) {
    const double n2 = (s->axis[0] * s->axis[0] + s->axis[1] * s->axis[1]) + s->axis[2] * s->axis[2];
    if (!(std::fabs(n2 - 1.0) <= 1e-12)) return fail(o, "source axis must be a unit vector") ? 0 : 1;
    if (!(std::fabs(s->beta) <= 1.0)) return fail(o, "source beta must be in [-1, 1]") ? 0 : 1;
  } else if (s->angle != MCW_ANGLE_ISOTROPIC) {
    return fail(o, "source angle type not supported") ? 0 : 1;
  }
  return 0;
}

int orc_tally_add(void *h, int32_t id, int32_t estimator, int32_t n_filters, const mcw_filter *filters,
                  int32_t n_scores, const int32_t *scores, int32_t n_nuclides, const int32_t *nuclides)
{
  Oracle *o = (Oracle *)h;
  Tally t;
  t.id = id;
  t.estimator = estimator;
  for (int k = 0; k < n_filters; ++k) {
    Filter f;
    f.type = filters[k].type;
    if (f.type == MCW_FILTER_CELL || f.type == MCW_FILTER_UNIVERSE) {
      if (filters[k].n_bins > 0) {
        f.ids.assign(filters[k].ids, filters[k].ids + filters[k].n_bins);
      } else if (f.type == MCW_FILTER_CELL) {
        for (int64_t e = 0; e < o->n_elem; ++e) f.ids.push_back(e);
      } else {
        std::vector<int64_t> b(o->sub.begin(), o->sub.end());
        std::sort(b.begin(), b.end());
        b.erase(std::unique(b.begin(), b.end()), b.end());
        f.ids = b;
      }
      f.n_bins = (int64_t)f.ids.size();
      std::map<int64_t, int64_t> pos;
      for (int64_t b = 0; b < f.n_bins; ++b) pos.emplace(f.ids[b], b);
      f.elem_bin.assign(o->n_elem, -1);
      for (int64_t e = 0; e < o->n_elem; ++e) {
        auto it = pos.find(f.type == MCW_FILTER_CELL ? e : (int64_t)o->sub[e]);
        if (it != pos.end()) f.elem_bin[e] = it->second;
      }
    } else if (f.type == MCW_FILTER_ENERGY) {
      f.edges.assign(filters[k].edges, filters[k].edges + filters[k].n_bins);
      f.n_bins = (int64_t)f.edges.size() - 1;
    } else {
      f.n_bins = 1;
    }
    t.filters.push_back(std::move(f));
  }
  // OpenMC Tally::set_strides: the last filter varies fastest
  t.strides.assign(t.filters.size(), 0);
  int64_t stride = 1;
  for (int k = (int)t.filters.size() - 1; k >= 0; --k) {
    t.strides[k] = stride;
    stride *= t.filters[k].n_bins;
  }
  t.n_filter_bins = stride;
  t.scores.assign(scores, scores + n_scores);
  if (n_nuclides > 0) t.nuclides.assign(nuclides, nuclides + n_nuclides);
  else t.nuclides.push_back(-1);
  t.n_score_bins = (int64_t)t.scores.size() * (int64_t)t.nuclides.size();
  t.value.assign(t.n_filter_bins * t.n_score_bins, 0.0);
  t.sum = t.value;
  t.sum_sq = t.value;
  o->tallies.push_back(std::move(t));
  return 0;
}

int orc_tally_get(void *h, int32_t id, int32_t kind, double *out, size_t n)
{
  Oracle *o = (Oracle *)h;
  for (auto &t : o->tallies)
    if (t.id == id) {
      const std::vector<double> &src = kind == MCW_RESULT_VALUE ? t.value : kind == MCW_RESULT_SUM ? t.sum : t.sum_sq;
      if (n < src.size()) return fail(o, "output too small") ? 0 : 1;
      std::copy(src.begin(), src.end(), out);
      return 0;
    }
  return fail(o, "no such tally") ? 0 : 1;
}

/* Run one batch with `threads` host threads (0 = hardware concurrency).
 * Histories are split into contiguous ranges; tallies are reduced in
 * thread order. */
int orc_run_batch(void *h, int64_t batch, int64_t n_particles, int32_t threads, int32_t record,
                  mcw_batch_stats *out)
{
  Oracle *o = (Oracle *)h;
  if (!check_ready(o)) return 1;
  const bool eig = o->opt.run_mode == MCW_RUN_EIGENVALUE;
  bool active = true;
  if (eig) {
    if (batch == 1) {
      o->keff = 1.0;
      o->k_sum = o->k_sum_sq = 0.0;
      o->n_active = 0;
      o->next_src.clear();
    } else if (batch != o->eig_batch + 1 || n_particles != o->eig_n ||
               (int64_t)o->next_src.size() != n_particles) {
      return fail(o, "eigenvalue batches run in order 1, 2, ... with one particle count") ? 0 : 1;
    }
    active = batch > o->opt.n_inactive;
  }
  int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  if (nt > n_particles) nt = (int)std::max<int64_t>(1, n_particles);
  if (record) o->hist.assign(n_particles, mcw_history());
  else o->hist.clear();
  std::vector<Scratch> scr(nt);
  std::vector<Counters> cnt(nt);
  for (auto &s : scr) {
    s.value.resize(o->tallies.size());
    for (size_t k = 0; k < o->tallies.size(); ++k) s.value[k].assign(o->tallies[k].value.size(), 0.0);
  }
  std::vector<std::vector<mcw_site>> tsites(eig ? nt : 0);
  std::vector<double> k_tl(eig ? n_particles : 0), k_col(eig ? n_particles : 0), k_abs(eig ? n_particles : 0);
  auto work = [&](int t) {
    int64_t lo = n_particles * t / nt, hi = n_particles * (t + 1) / nt;
    for (int64_t i = lo; i < hi; ++i) {
      if (!eig) {
        transport_history(*o, batch, n_particles, i, scr[t], cnt[t], record ? &o->hist[i] : nullptr, nullptr);
        continue;
      }
      EigenHistory eh;
      eh.keff = o->keff;
      eh.tally = active;
      eh.source = batch > 1 ? &o->next_src[(size_t)i] : nullptr;
      eh.sites = &tsites[t];
      transport_history(*o, batch, n_particles, i, scr[t], cnt[t], record ? &o->hist[i] : nullptr, nullptr, &eh);
      k_tl[(size_t)i] = eh.k_tl;
      k_col[(size_t)i] = eh.k_col;
      k_abs[(size_t)i] = eh.k_abs;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto &th : pool) th.join();

  Counters tot;
  for (int t = 0; t < nt; ++t) {
    tot.segments += cnt[t].segments;
    tot.collisions += cnt[t].collisions;
    tot.absorptions += cnt[t].absorptions;
    tot.leaks += cnt[t].leaks;
    tot.cutoffs += cnt[t].cutoffs;
    tot.lost += cnt[t].lost;
    tot.reflections += cnt[t].reflections;
    tot.hex_segments += cnt[t].hex_segments;
    tot.track += cnt[t].track;
  }
  // batch close: VALUE / total weight -> SUM, SUM_SQ (OpenMC accumulate);
  // an inactive batch scores nothing
  const double norm = (double)n_particles;
  for (size_t k = 0; k < o->tallies.size(); ++k) {
    Tally &T = o->tallies[k];
    for (size_t b = 0; b < T.value.size(); ++b) {
      double v = 0.0;
      for (int t = 0; t < nt; ++t) v += scr[t].value[k][b];
      T.value[b] = v;
      if (!active) continue;
      double val = v / norm;
      T.sum[b] += val;
      T.sum_sq[b] += val * val;
    }
  }
  if (eig) {
    o->fbank.clear();
    for (auto &v : tsites) o->fbank.insert(o->fbank.end(), v.begin(), v.end());
    mcw_keff_result &K = o->kres;
    memset(&K, 0, sizeof K);
    K.batch = batch;
    K.fission_sites = (int64_t)o->fbank.size();
    K.keff_used = o->keff;
    K.k_generation = tree_sum(k_tl) / norm;
    K.k_collision = tree_sum(k_col) / norm;
    K.k_absorption = tree_sum(k_abs) / norm;
    // calculate_average_keff: the generation value while inactive, else the
    // mean over active batches and the standard deviation of that mean
    if (active) {
      ++o->n_active;
      o->k_sum += K.k_generation;
      o->k_sum_sq += K.k_generation * K.k_generation;
      const double n = (double)o->n_active;
      o->keff = o->k_sum / n;
      K.keff_std = o->n_active > 1 ? std::sqrt((o->k_sum_sq / n - o->keff * o->keff) / (n - 1.0)) : 0.0;
    } else {
      o->keff = K.k_generation;
    }
    K.keff = o->keff;
    K.n_active = o->n_active;
    o->eig_batch = batch;
    o->eig_n = n_particles;
    if (o->fbank.empty()) {
      o->next_src.clear();
      return fail(o, "No fission sites sampled on any process, cannot sample a new batch") ? 0 : 1;
    }
    o->next_src = synchronize_bank(o->fbank, batch, n_particles, o->opt.master_seed);
  }
  if (out) {
    memset(out, 0, sizeof *out);
    out->histories = n_particles;
    out->segments = tot.segments;
    out->collisions = tot.collisions;
    out->absorptions = tot.absorptions;
    out->leaks = tot.leaks;
    out->cutoffs = tot.cutoffs;
    out->lost = tot.lost;
    out->reflections = tot.reflections;
    out->hex_segments = tot.hex_segments;
    out->track_length = tot.track;
  }
  return 0;
}

int orc_keff_get(void *h, mcw_keff_result *out)
{
  *out = ((Oracle *)h)->kres;
  return 0;
}

static int copy_sites(const std::vector<mcw_site> &v, mcw_site *out, size_t max, int64_t *n_out)
{
  if (n_out) *n_out = (int64_t)v.size();
  if (out) std::copy(v.begin(), v.begin() + std::min(max, v.size()), out);
  return 0;
}

int orc_fission_bank_get(void *h, mcw_site *out, size_t max, int64_t *n_out)
{
  return copy_sites(((Oracle *)h)->fbank, out, max, n_out);
}

int orc_source_bank_get(void *h, mcw_site *out, size_t max, int64_t *n_out)
{
  return copy_sites(((Oracle *)h)->next_src, out, max, n_out);
}

/* The fixed-order sum the GPU reduces per-history keff scores with. */
double orc_tree_sum(const double *v, int64_t n) { return tree_sum(std::vector<double>(v, v + n)); }

/* synchronize_bank on caller sites, for tests. */
int orc_synchronize_bank(const mcw_site *bank, int64_t total, int64_t batch, int64_t n, uint64_t master,
                         mcw_site *out)
{
  if (total < 1 || n < 1) return 1;
  std::vector<mcw_site> b(bank, bank + total);
  std::vector<mcw_site> r = synchronize_bank(b, batch, n, master);
  std::copy(r.begin(), r.end(), out);
  return 0;
}

int orc_history_get(void *h, mcw_history *out, size_t n)
{
  Oracle *o = (Oracle *)h;
  if (n < o->hist.size()) return fail(o, "output too small") ? 0 : 1;
  std::copy(o->hist.begin(), o->hist.end(), out);
  return 0;
}

/* Per-segment trace of one history: for each segment the element it lies
 * in, the event that ends it and its track length.  Returns the number of segments written. */
int64_t orc_trace_history(void *h, int64_t batch, int64_t n_particles, int64_t i, int32_t *elems,
                          int32_t *events, double *lengths, int64_t max_n)
{
  Oracle *o = (Oracle *)h;
  if (!check_ready(o)) return -1;
  std::vector<int32_t> el, ev;
  std::vector<double> le;
  TraceSink ts;
  ts.elems = &el;
  ts.events = &ev;
  ts.lengths = &le;
  Scratch sc;
  sc.value.resize(o->tallies.size());
  for (size_t k = 0; k < o->tallies.size(); ++k) sc.value[k].assign(o->tallies[k].value.size(), 0.0);
  Counters c;
  transport_history(*o, batch, n_particles, i, sc, c, nullptr, &ts);
  int64_t n = std::min<int64_t>((int64_t)ev.size(), max_n);
  for (int64_t k = 0; k < n; ++k) {
    elems[k] = el[k];
    events[k] = ev[k];
    lengths[k] = le[k];
  }
  return (int64_t)ev.size();
}

/* Derived mesh data for tests. */
int64_t orc_mesh_n_faces(void *h) { return ((Oracle *)h)->face_off.back(); }
void orc_mesh_planes(void *h, double *out)
{
  Oracle *o = (Oracle *)h;
  std::copy(o->plane.begin(), o->plane.end(), out);
}
void orc_mesh_neighbors(void *h, int64_t *nbr, int32_t *side)
{
  Oracle *o = (Oracle *)h;
  std::copy(o->nbr.begin(), o->nbr.end(), nbr);
  std::copy(o->nbr_side.begin(), o->nbr_side.end(), side);
}
/* Split interfaces: per side slot the split index (-1: none); per split the
 * plane (4 doubles) and the two (element, side) pairs. */
int64_t orc_mesh_n_splits(void *h) { return (int64_t)((Oracle *)h)->splits.size(); }
void orc_mesh_splits(void *h, int64_t *split_of, double *planes, int64_t *elems, int32_t *sides)
{
  Oracle *o = (Oracle *)h;
  std::copy(o->split_of.begin(), o->split_of.end(), split_of);
  for (size_t k = 0; k < o->splits.size(); ++k) {
    for (int q = 0; q < 4; ++q) planes[4 * k + q] = o->splits[k].m[q];
    for (int q = 0; q < 2; ++q) {
      elems[2 * k + q] = o->splits[k].elem[q];
      sides[2 * k + q] = o->splits[k].side[q];
    }
  }
}
void orc_mesh_cdf(void *h, double *out)
{
  Oracle *o = (Oracle *)h;
  std::copy(o->cdf.begin(), o->cdf.end(), out);
}
int64_t orc_locate(void *h, const double *p, double tol) { return locate(*(Oracle *)h, p, tol); }

/* Source sites for particles 1..n of a batch (x,y,z,u,v,w,E,elem as double). */
int orc_sample_sources(void *h, int64_t batch, int64_t n_particles, int64_t n, double *out)
{
  Oracle *o = (Oracle *)h;
  if (!check_ready(o)) return 1;
  for (int64_t i = 0; i < n; ++i) {
    uint64_t s = orc_init_seed((batch - 1) * n_particles + i + 1, 1, o->opt.master_seed);
    Site st;
    if (!sample_source(*o, &s, &st)) return fail(o, "unsupported source") ? 0 : 1;
    double *r = out + 8 * i;
    r[0] = st.x; r[1] = st.y; r[2] = st.z; r[3] = st.u; r[4] = st.v; r[5] = st.w; r[6] = st.E;
    r[7] = (double)st.elem;
  }
  return 0;
}

/* Microscopic lookup for tests: out = {total, absorption, scatter, fission}. */
int orc_micro_xs(void *h, int32_t nuc, double E, double *out)
{
  Oracle *o = (Oracle *)h;
  Micro m;
  micro_xs(o->nucs[nuc], E, &m);
  out[0] = m.t; out[1] = m.a; out[2] = m.s; out[3] = m.f;
  return 0;
}

double orc_macro_total(void *h, int32_t mat, double E) { return macro_total(*(Oracle *)h, mat, E, -1.0); }

/* One nuclide pick of the collision step in material `mat` (list position)
 * at energy E from *seed (advanced past its draw), for tests: the picked
 * nuclide's index in the nuclide list. */
int32_t orc_sample_nuclide_x(void *h, int32_t mat, double E, uint64_t *seed)
{
  const Oracle &o = *(Oracle *)h;
  const Material &m = o.mats[mat];
  return m.nuc[sample_nuclide(o, m, E, -1.0, macro_total(o, mat, E, -1.0), seed)];
}

/* Elementary functions and RNG, exported for tests. */
double orc_log_x(double x) { return orc_log(x); }
void orc_sincos_2pi_x(double xi, double *s, double *c) { orc_sincos_2pi(xi, s, c); }
double orc_prn_x(uint64_t *seed) { return orc_prn(seed); }
uint64_t orc_step_x(uint64_t s) { return orc_step(s); }
uint64_t orc_future_seed_x(uint64_t n, uint64_t s) { return orc_future_seed(n, s); }
uint64_t orc_init_seed_x(int64_t id, int32_t stream, uint64_t master) { return orc_init_seed(id, stream, master); }
uint64_t orc_seed_roundtrip_x(uint64_t s) { return orc_seed_roundtrip(s); }
void orc_rotate_angle_x(double *uvw, double mu, double xi) { rotate_angle(&uvw[0], &uvw[1], &uvw[2], mu, xi); }
void orc_elastic_x(double E, double A, double mu_cm, double *Ep, double *mu_lab)
{
  elastic_kinematics(E, A, mu_cm, Ep, mu_lab);
}

} // extern "C"
