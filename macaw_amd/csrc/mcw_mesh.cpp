// mcw_mesh.cpp -- host-side derivation of the device mesh layout.
//
// From libMesh-ordered connectivity (the mesh OpenMCStudy maps to OpenMC
// cells, src/userobjects/OpenMCStudy.C:152-187) this builds what the
// element walk needs on the device:
//   - one outward unit plane per element side.  A side shared by two
//     elements gets one canonical plane (through its three lowest-numbered
//     nodes) used with opposite signs on the two sides, so a point that
//     leaves one element through the plane is exactly on the neighbour's;
//   - face neighbours by matching sorted side-node tuples; in mixed meshes a
//     HEX8 side covered by two TET4 sides (split along a diagonal) becomes a
//     split interface resolved by the exit point's side of the diagonal;
//   - boundary ids of external sides;
//   - the element-volume CDF of volumetric sources (sequential sum).
// Planar sides are required (warn_non_planar in the reference's inputs,
// examples/beavrs_core/run_3d_core_cycle1.i:102-104): a HEX8 side whose
// fourth node lies off the plane beyond a relative 1e-9 is rejected.
#include "mcw_mesh.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <numeric>
#include <thread>

namespace mcw {

static const int TET_SIDE[4][3] = {{0, 2, 1}, {0, 1, 3}, {1, 2, 3}, {2, 0, 3}};
static const int HEX_SIDE[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5},
                                   {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 5, 6, 7}};
const int HEX_KUHN_TETS[6][4] = {{0, 1, 2, 6}, {0, 2, 3, 6}, {0, 3, 7, 6},
                                 {0, 7, 4, 6}, {0, 4, 5, 6}, {0, 5, 1, 6}};

static inline int sides_of(uint8_t t) { return t == 4 ? 4 : 6; }

static void side_nodes(const HostMesh &m, int64_t e, int s, int32_t *out, int &nn)
{
  const int32_t *c = m.conn.data() + m.conn_off[e];
  if (m.type[e] == 4) {
    nn = 3;
    for (int k = 0; k < 3; ++k) out[k] = c[TET_SIDE[s][k]];
  } else {
    nn = 4;
    for (int k = 0; k < 4; ++k) out[k] = c[HEX_SIDE[s][k]];
  }
}

double tet_volume6(const double *a, const double *b, const double *c, const double *d)
{
  double ax = b[0] - a[0], ay = b[1] - a[1], az = b[2] - a[2];
  double bx = c[0] - a[0], by = c[1] - a[1], bz = c[2] - a[2];
  double cx = d[0] - a[0], cy = d[1] - a[1], cz = d[2] - a[2];
  double det = (ax * (by * cz - bz * cy) - ay * (bx * cz - bz * cx)) + az * (bx * cy - by * cx);
  return std::fabs(det);
}

static double volume6(const HostMesh &m, int64_t e)
{
  const int32_t *c = m.conn.data() + m.conn_off[e];
  const double *X = m.xyz.data();
  if (m.type[e] == 4)
    return tet_volume6(X + 3 * (int64_t)c[0], X + 3 * (int64_t)c[1], X + 3 * (int64_t)c[2], X + 3 * (int64_t)c[3]);
  double v = 0.0;
  for (int k = 0; k < 6; ++k)
    v = v + tet_volume6(X + 3 * (int64_t)c[HEX_KUHN_TETS[k][0]], X + 3 * (int64_t)c[HEX_KUHN_TETS[k][1]],
                        X + 3 * (int64_t)c[HEX_KUHN_TETS[k][2]], X + 3 * (int64_t)c[HEX_KUHN_TETS[k][3]]);
  return v;
}

std::string derive_mesh(HostMesh &m, const int32_t *face_boundary)
{
  const int64_t ne = m.n_elem;
  m.face_off.assign(ne + 1, 0);
  m.uniform_sides = 0;
  bool all_tet = true, all_hex = true;
  for (int64_t e = 0; e < ne; ++e) {
    uint8_t t = m.type[e];
    if (t != 4 && t != 8) return "unsupported element type " + std::to_string((int)t) + " at element " + std::to_string(e);
    if (m.conn_off[e + 1] - m.conn_off[e] != t) return "connectivity size does not match element type at element " + std::to_string(e);
    all_tet &= t == 4;
    all_hex &= t == 8;
    m.face_off[e + 1] = m.face_off[e] + sides_of(t);
  }
  if (all_tet) m.uniform_sides = 4;
  if (all_hex) m.uniform_sides = 6;
  const int64_t nf = m.face_off[ne];
  if (nf >= ((int64_t)1 << 31)) return "mesh too large: more than 2^31 element sides";
  if (ne >= ((int64_t)1 << 28)) return "mesh too large: more than 2^28 elements";
  for (int64_t i = 0; i < m.conn_off[ne]; ++i)
    if (m.conn[i] < 0 || m.conn[i] >= m.n_nodes) return "connectivity refers to a node out of range";
  for (int64_t i = 0; i < 3 * m.n_nodes; ++i)
    if (!std::isfinite(m.xyz[i])) return "node coordinates must be finite (node " + std::to_string(i / 3) + ")";

  // -- neighbours: sort side keys
  struct Key {
    int32_t n[4];
    int32_t slot;
  };
  std::vector<Key> keys(nf);
  std::vector<int32_t> slot_elem(nf);
  for (int64_t e = 0; e < ne; ++e)
    for (int s = 0; s < sides_of(m.type[e]); ++s) {
      int32_t fn[4];
      int nn;
      side_nodes(m, e, s, fn, nn);
      std::sort(fn, fn + nn);
      Key &k = keys[m.face_off[e] + s];
      for (int q = 0; q < 4; ++q) k.n[q] = q < nn ? fn[q] : -1;
      k.slot = (int32_t)(m.face_off[e] + s);
      slot_elem[m.face_off[e] + s] = (int32_t)e;
    }
  std::sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
    if (a.n[0] != b.n[0]) return a.n[0] < b.n[0];
    if (a.n[1] != b.n[1]) return a.n[1] < b.n[1];
    if (a.n[2] != b.n[2]) return a.n[2] < b.n[2];
    if (a.n[3] != b.n[3]) return a.n[3] < b.n[3];
    return a.slot < b.slot;
  });
  m.nbr.assign(nf, -1);
  m.boundary_ids.clear();
  std::vector<int32_t> open; // sides with no side of equal node set
  for (size_t i = 0; i < keys.size();) {
    size_t j = i + 1;
    while (j < keys.size() && keys[j].n[0] == keys[i].n[0] && keys[j].n[1] == keys[i].n[1] &&
           keys[j].n[2] == keys[i].n[2] && keys[j].n[3] == keys[i].n[3])
      ++j;
    if (j - i > 2) return "non-manifold side shared by more than two elements";
    if (j - i == 2) {
      int32_t a = keys[i].slot, b = keys[i + 1].slot;
      int32_t ea = slot_elem[a], eb = slot_elem[b];
      m.nbr[a] = eb * 8 + (int32_t)(b - m.face_off[eb]);
      m.nbr[b] = ea * 8 + (int32_t)(a - m.face_off[ea]);
    } else {
      open.push_back(keys[i].slot);
    }
    i = j;
  }
  keys.clear();
  keys.shrink_to_fit();

  // -- split interfaces (mixed meshes): a HEX8 side q0 q1 q2 q3 whose two
  // halves along the diagonal q0-q2 or q1-q3 are open TET4 sides.  The hex
  // side's code is split*8 + 7; the two triangles' code is the hex side.
  m.splits.clear();
  std::vector<std::array<int32_t, 3>> split_slots; // hex side, triangle, triangle
  std::vector<std::array<int32_t, 2>> split_diag;  // diagonal nodes, ascending
  if (m.uniform_sides == 0) {
    std::map<std::array<int32_t, 3>, int32_t> tri;
    auto key3 = [](int32_t a, int32_t b, int32_t c) {
      std::array<int32_t, 3> k = {a, b, c};
      std::sort(k.begin(), k.end());
      return k;
    };
    for (int32_t a : open)
      if (m.type[slot_elem[a]] == 4) {
        int32_t fn[4];
        int nn;
        side_nodes(m, slot_elem[a], (int)(a - m.face_off[slot_elem[a]]), fn, nn);
        tri[key3(fn[0], fn[1], fn[2])] = a;
      }
    for (int32_t a : open) {
      const int32_t e = slot_elem[a];
      if (m.type[e] != 8) continue;
      int32_t q[4];
      int nn;
      side_nodes(m, e, (int)(a - m.face_off[e]), q, nn);
      for (int d = 0; d < 2; ++d) {
        auto t0 = tri.find(key3(q[d], q[d + 1], q[d + 2]));
        auto t1 = tri.find(key3(q[d + 2], q[(d + 3) & 3], q[d]));
        if (t0 == tri.end() || t1 == tri.end()) continue;
        const int32_t s0 = t0->second, s1 = t1->second;
        tri.erase(t0);
        tri.erase(tri.find(key3(q[d + 2], q[(d + 3) & 3], q[d])));
        if (split_slots.size() >= ((size_t)1 << 28)) return "mesh too large: more than 2^28 split sides";
        m.nbr[a] = (int32_t)split_slots.size() * 8 + 7;
        split_slots.push_back({a, s0, s1});
        split_diag.push_back({std::min(q[d], q[d + 2]), std::max(q[d], q[d + 2])});
        m.nbr[s0] = e * 8 + (int32_t)(a - m.face_off[e]);
        m.nbr[s1] = m.nbr[s0];
        break;
      }
    }
  }
  for (int32_t a : open) {
    if (m.nbr[a] >= 0) continue; // matched to a split interface
    int32_t bid = face_boundary ? face_boundary[a] : 0;
    if (bid < 0) return "external side without a boundary id (element " + std::to_string(slot_elem[a]) + ")";
    m.nbr[a] = -1 - bid;
    m.boundary_ids.insert(bid);
  }
  std::vector<int32_t>().swap(open);

  // -- outward planes
  m.plane.assign(4 * nf, 0.0);
  const double *X = m.xyz.data();
  std::string err;
  auto planes = [&](int64_t lo, int64_t hi, std::string *perr) {
    for (int64_t e = lo; e < hi; ++e) {
      const int32_t *c = m.conn.data() + m.conn_off[e];
      const int nn = m.type[e];
      double cx = 0.0, cy = 0.0, cz = 0.0;
      for (int k = 0; k < nn; ++k) {
        cx = cx + X[3 * (int64_t)c[k]];
        cy = cy + X[3 * (int64_t)c[k] + 1];
        cz = cz + X[3 * (int64_t)c[k] + 2];
      }
      cx = cx / nn;
      cy = cy / nn;
      cz = cz / nn;
      for (int s = 0; s < sides_of(m.type[e]); ++s) {
        int32_t fn[4];
        int fnn;
        side_nodes(m, e, s, fn, fnn);
        std::sort(fn, fn + fnn);
        const double *p0 = X + 3 * (int64_t)fn[0];
        const double *p1 = X + 3 * (int64_t)fn[1];
        const double *p2 = X + 3 * (int64_t)fn[2];
        double ax = p1[0] - p0[0], ay = p1[1] - p0[1], az = p1[2] - p0[2];
        double bx = p2[0] - p0[0], by = p2[1] - p0[1], bz = p2[2] - p0[2];
        double nx = ay * bz - az * by;
        double ny = az * bx - ax * bz;
        double nz = ax * by - ay * bx;
        double len = std::sqrt((nx * nx + ny * ny) + nz * nz);
        if (!(len > 0.0)) {
          *perr = "degenerate side " + std::to_string(s) + " of element " + std::to_string(e);
          return;
        }
        nx = nx / len;
        ny = ny / len;
        nz = nz / len;
        double cc = (nx * p0[0] + ny * p0[1]) + nz * p0[2];
        if (fnn == 4) {
          const double *p3 = X + 3 * (int64_t)fn[3];
          double off = ((nx * p3[0] + ny * p3[1]) + nz * p3[2]) - cc;
          double scale = std::sqrt((ax * ax + ay * ay) + az * az) + std::sqrt((bx * bx + by * by) + bz * bz);
          if (std::fabs(off) > 1e-9 * scale) {
            *perr = "non-planar side " + std::to_string(s) + " of element " + std::to_string(e);
            return;
          }
        }
        double side = ((nx * cx + ny * cy) + nz * cz) - cc;
        double *p = m.plane.data() + 4 * (m.face_off[e] + s);
        if (side > 0.0) {
          p[0] = -nx;
          p[1] = -ny;
          p[2] = -nz;
          p[3] = -cc;
        } else {
          p[0] = nx;
          p[1] = ny;
          p[2] = nz;
          p[3] = cc;
        }
      }
    }
  };
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (ne < 100000) nt = 1;
  std::vector<std::string> errs(nt);
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t)
    pool.emplace_back(planes, ne * t / nt, ne * (t + 1) / nt, &errs[t]);
  for (auto &th : pool) th.join();
  for (auto &e : errs)
    if (!e.empty()) return e;

  // -- split planes: through the diagonal (lower node id first), normal to
  // the hex side's outward normal n: m = (n x d)/|n x d|, c = m.x_a.  The tet
  // whose off-diagonal vertex has m.v - c <= 0 takes exit points with
  // m.x - c <= 0, the other the rest.
  for (size_t k = 0; k < split_slots.size(); ++k) {
    const double *n = m.plane.data() + 4 * (int64_t)split_slots[k][0];
    const double *xa = X + 3 * (int64_t)split_diag[k][0];
    const double *xb = X + 3 * (int64_t)split_diag[k][1];
    const double dx = xb[0] - xa[0], dy = xb[1] - xa[1], dz = xb[2] - xa[2];
    double mx = n[1] * dz - n[2] * dy;
    double my = n[2] * dx - n[0] * dz;
    double mz = n[0] * dy - n[1] * dx;
    const double len = std::sqrt((mx * mx + my * my) + mz * mz);
    if (!(len > 0.0)) return "degenerate split side of element " + std::to_string(slot_elem[split_slots[k][0]]);
    mx = mx / len;
    my = my / len;
    mz = mz / len;
    HostMesh::Split S;
    S.m[0] = mx;
    S.m[1] = my;
    S.m[2] = mz;
    S.m[3] = (mx * xa[0] + my * xa[1]) + mz * xa[2];
    const int32_t t0 = split_slots[k][1], t1 = split_slots[k][2];
    int32_t fn[4];
    int nn;
    side_nodes(m, slot_elem[t0], (int)(t0 - m.face_off[slot_elem[t0]]), fn, nn);
    int32_t v = fn[0];
    for (int q = 0; q < 3; ++q)
      if (fn[q] != split_diag[k][0] && fn[q] != split_diag[k][1]) v = fn[q];
    const double *xv = X + 3 * (int64_t)v;
    const double sv = ((mx * xv[0] + my * xv[1]) + mz * xv[2]) - S.m[3];
    const int32_t lo = sv <= 0.0 ? t0 : t1, hi = sv <= 0.0 ? t1 : t0;
    S.elem[0] = slot_elem[lo];
    S.side[0] = (int32_t)(lo - m.face_off[slot_elem[lo]]);
    S.elem[1] = slot_elem[hi];
    S.side[1] = (int32_t)(hi - m.face_off[slot_elem[hi]]);
    m.splits.push_back(S);
  }

  // -- volume CDF, sequential in element order
  m.cdf.assign(ne, 0.0);
  std::vector<double> vol(ne);
  {
    std::vector<std::thread> p2;
    for (unsigned t = 0; t < nt; ++t)
      p2.emplace_back([&, t] {
        for (int64_t e = ne * t / nt; e < ne * (t + 1) / nt; ++e) vol[e] = volume6(m, e);
      });
    for (auto &th : p2) th.join();
  }
  double acc = 0.0;
  for (int64_t e = 0; e < ne; ++e) {
    if (!(vol[e] > 0.0)) return "degenerate element " + std::to_string(e) + " (zero volume)";
    acc = acc + vol[e];
    m.cdf[e] = acc;
  }
  {
    const int64_t G = std::min<int64_t>(ne, (int64_t)1 << 22);
    m.cdf_guide.assign(G + 1, 0);
    int64_t e = 0;
    for (int64_t j = 0; j <= G; ++j) {
      const double t = acc * (double)j / (double)G;
      while (e < ne && !(m.cdf[e] > t)) ++e;
      m.cdf_guide[j] = (int32_t)e;
    }
  }

  // -- Morton order of element centroids (21 bits per axis), ties by index;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = 0; i < m.n_nodes; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], X[3 * i + k]);
      hi[k] = std::max(hi[k], X[3 * i + k]);
    }
  auto spread = [](uint64_t v) {
    v &= 0x1fffff;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
  };
  // TET4 first, then HEX8, each group along the curve, so that an element's
  // type follows from its internal id (id < n_tet)
  struct MKey {
    int32_t hex;
    uint64_t code;
    int32_t e;
    bool operator<(const MKey &b) const
    {
      if (hex != b.hex) return hex < b.hex;
      if (code != b.code) return code < b.code;
      return e < b.e;
    }
  };
  std::vector<MKey> key(ne);
  for (int64_t e = 0; e < ne; ++e) {
    const int32_t *c = m.conn.data() + m.conn_off[e];
    const int nn = m.type[e];
    uint64_t code = 0;
    for (int k = 0; k < 3; ++k) {
      double cc = 0.0;
      for (int q = 0; q < nn; ++q) cc += X[3 * (int64_t)c[q] + k];
      cc /= nn;
      double t = hi[k] > lo[k] ? (cc - lo[k]) / (hi[k] - lo[k]) : 0.0;
      uint64_t q = (uint64_t)std::min(2097151.0, std::max(0.0, t * 2097152.0));
      code |= spread(q) << k;
    }
    key[e] = {m.type[e] == 8 ? 1 : 0, code, (int32_t)e};
  }
  std::sort(key.begin(), key.end());
  m.orig_of_int.resize(ne);
  m.int_of_orig.resize(ne);
  m.n_tet = 0;
  for (int64_t i = 0; i < ne; ++i) {
    m.orig_of_int[i] = key[i].e;
    m.int_of_orig[key[i].e] = (int32_t)i;
    m.n_tet += key[i].hex == 0;
  }
  return "";
}

void build_locator(HostMesh &m)
{
  if (m.have_grid) return;
  const int64_t ne = m.n_elem;
  const double *X = m.xyz.data();
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = 0; i < m.n_nodes; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], X[3 * i + k]);
      hi[k] = std::max(hi[k], X[3 * i + k]);
    }
  double ext[3];
  for (int k = 0; k < 3; ++k) ext[k] = std::max(hi[k] - lo[k], 1e-300);
  // cells shaped like the elements: per axis, the mean extent of an element's
  // bounding box, scaled so that there are about as many cells as elements
  // (an extruded mesh's elements can be 1000 times taller than wide; cubic
  // cells would then list hundreds of candidates each)
  double mean_ext[3] = {0.0, 0.0, 0.0};
  for (int64_t o = 0; o < ne; ++o) {
    double blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t q = m.conn_off[o]; q < m.conn_off[o + 1]; ++q)
      for (int k = 0; k < 3; ++k) {
        blo[k] = std::min(blo[k], X[3 * (int64_t)m.conn[q] + k]);
        bhi[k] = std::max(bhi[k], X[3 * (int64_t)m.conn[q] + k]);
      }
    for (int k = 0; k < 3; ++k) mean_ext[k] += (bhi[k] - blo[k]) / (double)ne;
  }
  double ratio = 1.0; // prod over axes of ext / mean element extent
  for (int k = 0; k < 3; ++k) {
    mean_ext[k] = std::min(std::max(mean_ext[k], 1e-12 * ext[k]), ext[k]);
    ratio *= ext[k] / mean_ext[k];
  }
  const double alpha = std::cbrt(ratio / (double)std::max<int64_t>(1, ne)); // cell side / mean extent
  for (int k = 0; k < 3; ++k) {
    m.grid_n[k] = (int32_t)std::min(1024.0, std::max(1.0, std::ceil(ext[k] / (alpha * mean_ext[k]))));
    m.grid_lo[k] = lo[k];
    m.grid_inv[k] = m.grid_n[k] / ext[k];
  }
  const int64_t nc = (int64_t)m.grid_n[0] * m.grid_n[1] * m.grid_n[2];
  // element bounding boxes (internal order), widened by a tolerance
  auto cells_of = [&](int64_t i, int32_t *c0, int32_t *c1) {
    const int64_t o = m.orig_of_int[i];
    double blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t q = m.conn_off[o]; q < m.conn_off[o + 1]; ++q)
      for (int k = 0; k < 3; ++k) {
        blo[k] = std::min(blo[k], X[3 * (int64_t)m.conn[q] + k]);
        bhi[k] = std::max(bhi[k], X[3 * (int64_t)m.conn[q] + k]);
      }
    for (int k = 0; k < 3; ++k) {
      const double tol = 1e-9 * ext[k];
      c0[k] = (int32_t)std::max(0.0, std::floor((blo[k] - tol - lo[k]) * m.grid_inv[k]));
      c1[k] = (int32_t)std::min((double)m.grid_n[k] - 1, std::floor((bhi[k] + tol - lo[k]) * m.grid_inv[k]));
    }
  };
  m.cell_start.assign(nc + 1, 0);
  for (int64_t i = 0; i < ne; ++i) {
    int32_t c0[3], c1[3];
    cells_of(i, c0, c1);
    for (int32_t z = c0[2]; z <= c1[2]; ++z)
      for (int32_t y = c0[1]; y <= c1[1]; ++y)
        for (int32_t x = c0[0]; x <= c1[0]; ++x) ++m.cell_start[x + (int64_t)m.grid_n[0] * (y + (int64_t)m.grid_n[1] * z) + 1];
  }
  for (int64_t c = 0; c < nc; ++c) m.cell_start[c + 1] += m.cell_start[c];
  m.cell_elem.resize(m.cell_start[nc]);
  std::vector<int32_t> fill(m.cell_start.begin(), m.cell_start.end() - 1);
  for (int64_t i = 0; i < ne; ++i) {
    int32_t c0[3], c1[3];
    cells_of(i, c0, c1);
    for (int32_t z = c0[2]; z <= c1[2]; ++z)
      for (int32_t y = c0[1]; y <= c1[1]; ++y)
        for (int32_t x = c0[0]; x <= c1[0]; ++x)
          m.cell_elem[fill[x + (int64_t)m.grid_n[0] * (y + (int64_t)m.grid_n[1] * z)]++] = (int32_t)i;
  }
  m.have_grid = true;
}

static bool contains(const HostMesh &m, int64_t e, const double *p, double tol)
{
  for (int64_t s = m.face_off[e]; s < m.face_off[e + 1]; ++s) {
    const double *q = m.plane.data() + 4 * s;
    const double v = ((q[0] * p[0] + q[1] * p[1]) + q[2] * p[2]) - q[3];
    if (v > tol) return false;
  }
  return true;
}

int64_t locate_point(const HostMesh &m, const double *p, double tol)
{
  if (!m.have_grid) {
    for (int64_t e = 0; e < m.n_elem; ++e)
      if (contains(m, e, p, tol)) return e;
    return -1;
  }
  // grid cell of p (clamped); its candidates are every element whose widened
  // bounding box meets the cell, so the lowest caller index among those that
  // contain p is the brute-force answer
  int64_t cell[3];
  for (int k = 0; k < 3; ++k) {
    const double f = std::min((p[k] - m.grid_lo[k]) * m.grid_inv[k], (double)(m.grid_n[k] - 1));
    cell[k] = f > 0.0 ? (int64_t)f : 0; // NaN and below the box: cell 0
  }
  const int64_t ci = cell[0] + (int64_t)m.grid_n[0] * (cell[1] + (int64_t)m.grid_n[1] * cell[2]);
  int64_t best = -1;
  for (int32_t q = m.cell_start[ci]; q < m.cell_start[ci + 1]; ++q) {
    const int64_t e = m.orig_of_int[m.cell_elem[q]];
    if ((best < 0 || e < best) && contains(m, e, p, tol)) best = e;
  }
  return best;
}

} // namespace mcw
