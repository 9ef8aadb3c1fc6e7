// objects.cpp -- MOOSE-shaped host objects (see objects.h).
#include "objects.h"

#include "exodus.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <set>
#include <sstream>

namespace mcwhost {

// ------------------------------------------------------------------ parameters

std::string InputParameters::getString(const std::string &k) const { return node_->get(k); }
std::string InputParameters::getString(const std::string &k, const std::string &d) const
{
  return node_->has(k) ? node_->get(k) : d;
}
std::vector<std::string> InputParameters::getWords(const std::string &k) const
{
  return node_->has(k) ? split_words(node_->get(k)) : std::vector<std::string>{};
}

static long long to_int(const InputParameters &p, const std::string &k, const std::string &v)
{
  try {
    size_t pos = 0;
    long long x = std::stoll(v, &pos);
    if (pos != v.size()) throw std::invalid_argument(v);
    return x;
  } catch (const std::exception &) {
    paramError(p, k, "expected an integer, got '" + v + "'");
  }
}

static double to_real(const InputParameters &p, const std::string &k, const std::string &v)
{
  try {
    size_t pos = 0;
    double x = std::stod(v, &pos);
    if (pos != v.size()) throw std::invalid_argument(v);
    return x;
  } catch (const std::exception &) {
    paramError(p, k, "expected a number, got '" + v + "'");
  }
}

long long InputParameters::getInt(const std::string &k) const { return to_int(*this, k, node_->get(k)); }
long long InputParameters::getInt(const std::string &k, long long d) const { return node_->has(k) ? getInt(k) : d; }
double InputParameters::getReal(const std::string &k) const { return to_real(*this, k, node_->get(k)); }
double InputParameters::getReal(const std::string &k, double d) const { return node_->has(k) ? getReal(k) : d; }
bool InputParameters::getBool(const std::string &k, bool d) const
{
  if (!node_->has(k)) return d;
  std::string v = node_->get(k);
  std::transform(v.begin(), v.end(), v.begin(), ::tolower);
  if (v == "true" || v == "1" || v == "yes") return true;
  if (v == "false" || v == "0" || v == "no") return false;
  paramError(*this, k, "expected a boolean, got '" + v + "'");
}
std::vector<long long> InputParameters::getInts(const std::string &k) const
{
  std::vector<long long> out;
  for (auto &w : getWords(k)) out.push_back(to_int(*this, k, w));
  return out;
}
std::vector<double> InputParameters::getReals(const std::string &k) const
{
  std::vector<double> out;
  for (auto &w : getWords(k)) out.push_back(to_real(*this, k, w));
  return out;
}

void paramError(const InputParameters &p, const std::string &param, const std::string &msg)
{
  throw MooseError(p.path() + "/" + param + ": " + msg);
}

void mooseError(const std::string &msg) { throw MooseError(msg); }

static void check(mcw_ctx *ctx, int rc, const std::string &what)
{
  if (rc != 0) mooseError(what + ": " + (ctx ? mcw_last_error(ctx) : "no context"));
}

// ------------------------------------------------------------------ mesh

// GeneratedMeshGenerator boundary ids (MOOSE): back, bottom, right, top, left, front.
static const char *GEN_BOUNDARY[6] = {"back", "bottom", "right", "top", "left", "front"};
// Kuhn split of a hex (libMesh node order) into 6 tets around diagonal 0-6.
static const int KUHN[6][4] = {{0, 1, 2, 6}, {0, 2, 3, 6}, {0, 3, 7, 6}, {0, 7, 4, 6}, {0, 4, 5, 6}, {0, 5, 1, 6}};
static const int TET_SIDE[4][3] = {{0, 2, 1}, {0, 1, 3}, {1, 2, 3}, {2, 0, 3}};

static void generated_mesh(const InputParameters &p, MeshData &m)
{
  if (p.getInt("dim", 3) != 3) paramError(p, "dim", "only dim = 3 meshes are traced here");
  const long long nx = p.getInt("nx", 1), ny = p.getInt("ny", 1), nz = p.getInt("nz", 1);
  const double x0 = p.getReal("xmin", 0), x1 = p.getReal("xmax", 1), y0 = p.getReal("ymin", 0),
               y1 = p.getReal("ymax", 1), z0 = p.getReal("zmin", 0), z1 = p.getReal("zmax", 1);
  const std::string et = p.getString("elem_type", "HEX8");
  if (et != "HEX8" && et != "TET4") paramError(p, "elem_type", "supported element types: HEX8, TET4");
  m = MeshData();
  for (long long k = 0; k <= nz; ++k)
    for (long long j = 0; j <= ny; ++j)
      for (long long i = 0; i <= nx; ++i) {
        m.xyz.push_back(x0 + (x1 - x0) * i / nx);
        m.xyz.push_back(y0 + (y1 - y0) * j / ny);
        m.xyz.push_back(z0 + (z1 - z0) * k / nz);
      }
  auto nid = [&](long long i, long long j, long long k) { return (int32_t)(i + (nx + 1) * (j + (ny + 1) * k)); };
  for (int b = 0; b < 6; ++b) m.boundary_names[GEN_BOUNDARY[b]] = b;
  for (long long k = 0; k < nz; ++k)
    for (long long j = 0; j < ny; ++j)
      for (long long i = 0; i < nx; ++i) {
        int32_t h[8] = {nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                        nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)};
        if (et == "HEX8") {
          m.type.push_back(MCW_HEX8);
          m.conn.insert(m.conn.end(), h, h + 8);
          m.conn_off.push_back((int64_t)m.conn.size());
          m.subdomain.push_back(0);
          for (int s = 0; s < 6; ++s) m.face_boundary.push_back(s);
        } else {
          for (int t = 0; t < 6; ++t) {
            int32_t tet[4];
            for (int q = 0; q < 4; ++q) tet[q] = h[KUHN[t][q]];
            m.type.push_back(MCW_TET4);
            m.conn.insert(m.conn.end(), tet, tet + 4);
            m.conn_off.push_back((int64_t)m.conn.size());
            m.subdomain.push_back(0);
            // boundary id of an external side: the cube face it lies on
            for (int s = 0; s < 4; ++s) {
              double c[3] = {0, 0, 0};
              for (int q = 0; q < 3; ++q)
                for (int d = 0; d < 3; ++d) c[d] += m.xyz[3 * (int64_t)tet[TET_SIDE[s][q]] + d] / 3.0;
              int b = 0;
              const double tol = 1e-12 * (std::fabs(x1 - x0) + std::fabs(y1 - y0) + std::fabs(z1 - z0));
              if (std::fabs(c[2] - z0) < tol) b = 0;
              else if (std::fabs(c[1] - y0) < tol) b = 1;
              else if (std::fabs(c[0] - x1) < tol) b = 2;
              else if (std::fabs(c[1] - y1) < tol) b = 3;
              else if (std::fabs(c[0] - x0) < tol) b = 4;
              else if (std::fabs(c[2] - z1) < tol) b = 5;
              m.face_boundary.push_back(b);
            }
          }
        }
      }
}

static void subdomain_bbox(const InputParameters &p, MeshData &m)
{
  auto bl = p.getReals("bottom_left"), tr = p.getReals("top_right");
  if (bl.size() != 3 || tr.size() != 3) paramError(p, "bottom_left", "bottom_left and top_right need 3 coordinates");
  const int32_t id = (int32_t)p.getInt("block_id");
  const bool inside = p.getString("location", "INSIDE") == "INSIDE";
  for (int64_t e = 0; e < m.n_elem(); ++e) {
    double c[3] = {0, 0, 0};
    const int nn = (int)(m.conn_off[e + 1] - m.conn_off[e]);
    for (int64_t q = m.conn_off[e]; q < m.conn_off[e + 1]; ++q)
      for (int d = 0; d < 3; ++d) c[d] += m.xyz[3 * (int64_t)m.conn[q] + d] / nn;
    bool in = true;
    for (int d = 0; d < 3; ++d) in &= c[d] >= std::min(bl[d], tr[d]) && c[d] <= std::max(bl[d], tr[d]);
    if (in == inside) m.subdomain[e] = id;
  }
}

// Exodus side (1-based) -> libMesh side, per element type.  HEX8: Exodus
// faces 1-4 are libMesh sides 1-4, face 5 (nodes 1,4,3,2) is side 0 and face
// 6 (5,6,7,8) side 5.  TET4: faces 1-3 are sides 1-3, face 4 (1,3,2) is side
// 0.  QUAD4: edge k joins nodes k, k+1 (edge k-1 here).
static int exodus_side(uint8_t type, int64_t s)
{
  if (type == MCW_HEX8) {
    static const int H[7] = {-1, 1, 2, 3, 4, 0, 5};
    return s >= 1 && s <= 6 ? H[s] : -1;
  }
  if (type == MCW_TET4) return s >= 1 && s <= 4 ? (int)(s % 4) : -1;
  return s >= 1 && s <= 4 ? (int)(s - 1) : -1; // QUAD4
}

static int n_sides(uint8_t type) { return type == MCW_HEX8 ? 6 : 4; }

// MOOSE FileMeshGenerator: an Exodus II mesh of TET4 and HEX8 blocks, or a
// 2D mesh of QUAD4 blocks for an extruder.  Elements are numbered block by
// block as stored; a block's id is its subdomain.  A side in several side
// sets keeps the smallest id.
static void file_mesh(const InputParameters &p, MeshData &m, const std::string &input_dir)
{
  std::string file = p.getString("file");
  if (!file.empty() && file[0] != '/' && !input_dir.empty()) file = input_dir + "/" + file;
  ExodusMesh X;
  try {
    X = read_exodus(file);
  } catch (const std::exception &e) {
    paramError(p, "file", e.what());
  }
  m = MeshData();
  m.dim = X.dim;
  m.xyz = X.xyz;
  std::vector<int64_t> slot_off{0};
  for (auto &B : X.blocks) {
    std::string t = B.elem_type;
    for (auto &ch : t) ch = (char)toupper(ch);
    uint8_t type;
    int need;
    if (X.dim == 3 && (t == "TETRA" || t == "TETRA4" || t == "TET4" || t == "TET")) {
      type = MCW_TET4;
      need = 4;
    } else if (X.dim == 3 && (t == "HEX" || t == "HEX8")) {
      type = MCW_HEX8;
      need = 8;
    } else if (X.dim == 2 && (t == "QUAD" || t == "QUAD4" || t == "SHELL4")) {
      type = MESH_QUAD4;
      need = 4;
    } else {
      paramError(p, "file", "block " + std::to_string(B.id) + ": element type '" + B.elem_type + "' in a " +
                                std::to_string(X.dim) + "D mesh is not supported (TET4, HEX8; QUAD4 to extrude)");
    }
    if (B.nodes_per_elem != need)
      paramError(p, "file", "block " + std::to_string(B.id) + ": " + std::to_string(B.nodes_per_elem) +
                                " nodes per element for " + B.elem_type);
    const int64_t ne = (int64_t)B.conn.size() / need;
    for (int64_t e = 0; e < ne; ++e) {
      m.type.push_back(type);
      for (int q = 0; q < need; ++q) m.conn.push_back((int32_t)B.conn[(size_t)(e * need + q)]);
      m.conn_off.push_back((int64_t)m.conn.size());
      m.subdomain.push_back((int32_t)B.id);
      slot_off.push_back(slot_off.back() + n_sides(type));
    }
  }
  if (m.type.empty()) paramError(p, "file", "the mesh has no elements");
  m.face_boundary.assign((size_t)slot_off.back(), -1);
  for (auto &S : X.sidesets) {
    for (size_t k = 0; k < S.elem.size(); ++k) {
      const int64_t e = S.elem[k];
      if (e < 0 || e >= m.n_elem())
        paramError(p, "file", "side set " + std::to_string(S.id) + " names an element outside the mesh");
      const int side = exodus_side(m.type[e], S.side[k]);
      if (side < 0) paramError(p, "file", "side set " + std::to_string(S.id) + " names an unknown side");
      int32_t &b = m.face_boundary[(size_t)(slot_off[e] + side)];
      if (b < 0 || S.id < b) b = (int32_t)S.id;
    }
    if (!S.name.empty()) m.boundary_names[S.name] = (int32_t)S.id;
  }
}

// MOOSE TransformGenerator: TRANSLATE or SCALE by vector_value.
static void transform_mesh(const InputParameters &p, MeshData &m)
{
  const std::string t = p.getString("transform");
  const std::vector<double> v = p.getReals("vector_value");
  if (v.size() != 3) paramError(p, "vector_value", "vector_value needs 3 components");
  for (int64_t i = 0; i < m.n_nodes(); ++i)
    for (int d = 0; d < 3; ++d) {
      double &x = m.xyz[3 * i + d];
      if (t == "TRANSLATE") x = x + v[d];
      else if (t == "SCALE") x = x * v[d];
      else paramError(p, "transform", "transform must be TRANSLATE or SCALE");
    }
}

// MOOSE FancyExtruderGenerator (AdvancedExtruderGenerator): a QUAD4 mesh swept
// along `direction` through elevation zones of `heights`, zone i in
// num_layers[i] equal layers, into HEX8.  Layer j's nodes are the 2D nodes
// moved by the height below layer j; elements are numbered layer by layer.
// An edge's boundary ids carry to the side it sweeps; the first layer's
// bottoms get bottom_boundary and the last layer's tops top_boundary.  A quad
// wound clockwise about the direction is taken in reverse node order, so the
// hex has positive volume.
static void extrude_mesh(const InputParameters &p, MeshData &m)
{
  if (m.dim != 2) paramError(p, "input", "the extruder needs a 2D (QUAD4) mesh");
  const std::vector<double> h = p.getReals("heights");
  const std::vector<long long> nl = p.getInts("num_layers");
  if (h.empty() || h.size() != nl.size()) paramError(p, "num_layers", "one number of layers per height");
  for (size_t i = 0; i < h.size(); ++i)
    if (!(h[i] > 0.0) || nl[i] < 1) paramError(p, "heights", "heights must be positive, with at least one layer each");
  std::vector<double> dir = p.isParamValid("direction") ? p.getReals("direction") : std::vector<double>{0, 0, 1};
  if (dir.size() != 3) paramError(p, "direction", "direction needs 3 components");
  const double dn = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  if (!(dn > 0.0)) paramError(p, "direction", "direction must not be zero");
  for (double &d : dir) d /= dn;
  const int32_t bottom = (int32_t)p.getInt("bottom_boundary"), top = (int32_t)p.getInt("top_boundary");
  std::vector<double> off{0.0};
  for (size_t i = 0; i < h.size(); ++i)
    for (long long k = 1; k <= nl[i]; ++k) off.push_back(off.back() + h[i] / (double)nl[i]);
  const int64_t L = (int64_t)off.size() - 1, n2 = m.n_nodes(), e2 = m.n_elem();
  if ((L + 1) * n2 > INT32_MAX) paramError(p, "num_layers", "too many nodes after extrusion");
  MeshData o;
  o.dim = 3;
  o.boundary_names = m.boundary_names;
  for (int64_t j = 0; j <= L; ++j)
    for (int64_t i = 0; i < n2; ++i)
      for (int d = 0; d < 3; ++d) o.xyz.push_back(m.xyz[3 * i + d] + dir[d] * off[j]);
  for (int64_t j = 0; j < L; ++j)
    for (int64_t e = 0; e < e2; ++e) {
      const int32_t *c = &m.conn[m.conn_off[e]];
      const double *a = &m.xyz[3 * (int64_t)c[0]], *b = &m.xyz[3 * (int64_t)c[1]], *q = &m.xyz[3 * (int64_t)c[2]];
      // quad normal (b - a) x (q - a) against the direction
      const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, w[3] = {q[0] - a[0], q[1] - a[1], q[2] - a[2]};
      const double nz = (u[1] * w[2] - u[2] * w[1]) * dir[0] + (u[2] * w[0] - u[0] * w[2]) * dir[1] +
                        (u[0] * w[1] - u[1] * w[0]) * dir[2];
      const bool rev = nz < 0.0;
      const int32_t r[4] = {c[0], rev ? c[3] : c[1], c[2], rev ? c[1] : c[3]};
      o.type.push_back(MCW_HEX8);
      for (int k = 0; k < 4; ++k) o.conn.push_back((int32_t)(r[k] + j * n2));
      for (int k = 0; k < 4; ++k) o.conn.push_back((int32_t)(r[k] + (j + 1) * n2));
      o.conn_off.push_back((int64_t)o.conn.size());
      o.subdomain.push_back(m.subdomain[e]);
      const int32_t *fb = &m.face_boundary[(size_t)(4 * e)];
      o.face_boundary.push_back(j == 0 ? bottom : -1);
      for (int k = 0; k < 4; ++k) o.face_boundary.push_back(fb[rev ? 3 - k : k]);
      o.face_boundary.push_back(j == L - 1 ? top : -1);
    }
  m = std::move(o);
}

MeshData build_mesh(const HitNode &mb, const std::string &input_dir)
{
  MeshData m;
  bool any = false;
  for (auto &c : mb.children) {
    InputParameters p(c.get());
    const std::string type = p.getString("type");
    if (type == "GeneratedMeshGenerator") {
      generated_mesh(p, m);
      any = true;
    } else if (type == "FileMeshGenerator") {
      file_mesh(p, m, input_dir);
      any = true;
    } else if (type == "SubdomainBoundingBoxGenerator" || type == "TransformGenerator" ||
               type == "FancyExtruderGenerator" || type == "AdvancedExtruderGenerator") {
      if (!any) paramError(p, "input", "needs a mesh generated before it");
      if (type == "SubdomainBoundingBoxGenerator") subdomain_bbox(p, m);
      else if (type == "TransformGenerator") transform_mesh(p, m);
      else extrude_mesh(p, m);
    } else {
      paramError(p, "type", "unsupported mesh generator '" + type + "'");
    }
  }
  if (!any) mooseError("[Mesh] does not generate a mesh");
  if (m.dim != 3) mooseError("[Mesh] ends with a 2D mesh: rays are traced in 3D meshes (extrude it)");
  return m;
}

// ------------------------------------------------------------------ synthetic nuclides

static NuclideData resonance_table(const NuclideData &base, long long np, double e0, double e1, double sa0,
                                   double sa_flat, double sp, double ff, const std::vector<double> &Er,
                                   const std::vector<double> &G, const std::vector<double> &pk,
                                   const std::vector<double> &fg);

// Pointwise nuclide: 1/v absorption plus single-level Breit-Wigner resonances
// and constant potential scattering on a log grid; resonance parameters are
// drawn from the OpenMC generator seeded with init_seed(seed, 0).  With
// `temperatures`, one table per temperature T, every resonance broadened for
// the Doppler width G_D = sqrt(4 k T E_r / awr): width sqrt(G^2 + G_D^2),
// peak scaled to keep width x peak; without, one unbroadened table at 293.6 K.
static NuclideData synthetic_nuclide(const InputParameters &p)
{
  NuclideData n;
  n.name = p.path().substr(p.path().rfind('/') + 1);
  n.awr = p.getReal("awr");
  const long long np = p.getInt("n_points", 20000);
  const double e0 = p.getReal("e_min", 1e-5), e1 = p.getReal("e_max", 2e7);
  const double sa0 = p.getReal("sigma_a0", 1.0), sp = p.getReal("sigma_p", 4.7);
  const double sa_flat = p.getReal("sigma_a_flat", 0.0);
  const double ff = p.getReal("fission_fraction", 0.0);
  const long long nres = p.getInt("resonances", 50);
  if (np < 2) paramError(p, "n_points", "a grid needs at least 2 points");
  if (!(ff >= 0.0 && ff <= 1.0)) paramError(p, "fission_fraction", "the fission fraction must be in [0, 1]");
  if (ff > 0.0) {
    const std::vector<double> nu = p.isParamValid("nu") ? p.getReals("nu") : std::vector<double>{2.43};
    if (nu.empty() || nu.size() > 2) paramError(p, "nu", "nu takes one or two coefficients: nu(E) = nu0 + nu1 E");
    n.nu[0] = nu[0];
    n.nu[1] = nu.size() > 1 ? nu[1] : 0.0;
    n.watt[0] = p.getReal("watt_a", 0.988e6);
    n.watt[1] = p.getReal("watt_b", 2.249e-6);
  }
  n.kappa = p.getReal("kappa", 0.0); // eV per fission, read by kappa-fission scores
  if (!(n.kappa >= 0.0)) paramError(p, "kappa", "kappa (eV per fission) must be >= 0");
  uint64_t s = (uint64_t)p.getInt("seed", 1) * 152917ull;
  auto prn = [&]() {
    s = 6364136223846793005ull * s + 1442695040888963407ull;
    uint64_t w = ((s >> ((s >> 59u) + 5u)) ^ s) * 12605985483714917081ull;
    return (double)((w >> 43u) ^ w) * 0x1p-64;
  };
  std::vector<double> Er(nres), G0(nres), pk0(nres), fg(nres);
  for (long long r = 0; r < nres; ++r) {
    Er[r] = std::exp(std::log(1.0) + prn() * (std::log(1e5) - std::log(1.0)));
    G0[r] = std::exp(std::log(1e-2) + prn() * (std::log(10.0) - std::log(1e-2)));
    pk0[r] = std::exp(std::log(10.0) + prn() * (std::log(1e4) - std::log(10.0)));
    fg[r] = 0.2 + 0.6 * prn();
  }
  std::vector<double> temps{0.0};
  if (p.isParamValid("temperatures")) {
    temps = p.getReals("temperatures");
    if (temps.empty()) paramError(p, "temperatures", "give at least one temperature");
    for (size_t k = 0; k < temps.size(); ++k) {
      if (!(temps[k] > 0.0 && std::isfinite(temps[k])))
        paramError(p, "temperatures", "temperatures must be positive (K)");
      for (size_t j = 0; j < k; ++j)
        if (temps[j] == temps[k]) paramError(p, "temperatures", "a temperature is given twice");
    }
  }
  std::vector<NuclideData> tables;
  for (double T : temps) {
    std::vector<double> G = G0, pk = pk0;
    if (T > 0.0)
      for (long long r = 0; r < nres; ++r) {
        const double gd2 = 4.0 * MCW_K_BOLTZMANN * T * Er[r] / n.awr;
        const double ge = std::sqrt(G0[r] * G0[r] + gd2);
        pk[r] = pk0[r] * G0[r] / ge;
        G[r] = ge;
      }
    tables.push_back(resonance_table(n, np, e0, e1, sa0, sa_flat, sp, ff, Er, G, pk, fg));
    if (T > 0.0) tables.back().temperature = T;
  }
  NuclideData out = std::move(tables[0]);
  for (size_t k = 1; k < tables.size(); ++k) out.temps.push_back(std::move(tables[k]));
  return out;
}

// One table of a synthetic nuclide: base's name, awr and fission data, the
// given resonances.
static NuclideData resonance_table(const NuclideData &base, long long np, double e0, double e1, double sa0,
                                   double sa_flat, double sp, double ff, const std::vector<double> &Er,
                                   const std::vector<double> &G, const std::vector<double> &pk,
                                   const std::vector<double> &fg)
{
  NuclideData n = base;
  const long long nres = (long long)Er.size();
  n.E.resize(np);
  n.t.resize(np);
  n.a.resize(np);
  n.s.resize(np);
  for (long long j = 0; j < np; ++j) {
    double E = j == 0 ? e0 : (j == np - 1 ? e1 : e0 * std::pow(e1 / e0, (double)j / (np - 1)));
    double a = sa0 * std::sqrt(0.0253 / E) + sa_flat, sc = sp;
    for (long long r = 0; r < nres; ++r) {
      double x = 2.0 * (E - Er[r]) / G[r];
      double shape = pk[r] / (1.0 + x * x);
      a += shape * fg[r] * std::sqrt(Er[r] / E);
      sc += shape * (1.0 - fg[r]);
    }
    n.E[j] = E;
    n.a[j] = a;
    n.s[j] = sc;
    n.t[j] = a + sc;
  }
  if (ff > 0.0) {
    n.f.resize(np);
    for (long long j = 0; j < np; ++j) n.f[j] = ff * n.a[j];
  }
  return n;
}

// ------------------------------------------------------------------ CollisionKernel

CollisionKernel::CollisionKernel(const InputParameters &p, Problem &fe) : _params(p), _fe(fe)
{
  _verbose = p.getBool("verbose", false);
  const std::string T = p.getString("temperature");
  _temperature = T;
  // a coupled variable may be given as a constant (temperature = 300)
  char *end = nullptr;
  std::strtod(T.c_str(), &end);
  const bool constant = !T.empty() && end && *end == '\0';
  auto it = constant ? fe.aux_vars.end() : fe.aux_vars.find(T);
  if (constant) {
    // nothing to check
  } else if (it == fe.aux_vars.end()) {
    paramError(p, "temperature", "no variable named '" + T + "'");
  } else if (it->second != 0) {
    // Check that the temperature variable is a constant monomial (CollisionKernel.C:52-57)
    paramError(p, "temperature", "Only CONST MONOMIAL temperatures (current order: " + std::to_string(it->second) +
                                     ") are supported.");
  }
  if (p.isParamValid("z_coord")) p.getReal("z_coord"); // 2D runs only; a 3D mesh ignores it
  auto blocks = p.getInts("blocks");
  auto materials = p.getInts("materials");
  if (blocks.size() != materials.size())
    paramError(p, "blocks", "The blocks parameter must be the same size as the materials parameter.");
  for (size_t i = 0; i < blocks.size(); ++i) _block_to_openmc_materials.emplace((int)blocks[i], (int)materials[i]);
}

void CollisionKernel::initialSetup()
{
  // Check that all materials do exist, otherwise it will crash at XS calculation
  for (auto &m : _block_to_openmc_materials) {
    bool found = false;
    for (auto &mat : _fe.materials) found |= mat.id == m.second;
    if (!found) mooseError("Could not find material " + std::to_string(m.second) + " in OpenMC materials.");
  }
  std::set<int> present(_fe.mesh.subdomain.begin(), _fe.mesh.subdomain.end());
  for (auto &b : _block_to_openmc_materials)
    if (!present.count(b.first))
      std::cerr << "*** Warning *** Could not find specified block " << b.first << " for the collision kernel.\n";
}

// ------------------------------------------------------------------ RayBCs

RayBC::RayBC(const InputParameters &p, Problem &fe, int kind) : _kind(kind)
{
  for (auto &name : p.getWords("boundary")) {
    auto it = fe.mesh.boundary_names.find(name);
    if (it != fe.mesh.boundary_names.end()) {
      _ids.push_back(it->second);
      continue;
    }
    try {
      _ids.push_back((int32_t)std::stoi(name));
    } catch (const std::exception &) {
      paramError(p, "boundary", "the boundary '" + name + "' was not found in the mesh");
    }
  }
}

void RayBC::apply(mcw_ctx *ctx) const
{
  for (int32_t b : _ids) check(ctx, mcw_bc_set(ctx, b, _kind), "RayBC");
}

// ------------------------------------------------------------------ OpenMCTally

// OpenMCTally.h:71-88: universe < cell < energy < others
static int filter_rank(const std::string &f)
{
  if (f == "universe") return 0;
  if (f == "cell") return 1;
  if (f == "energy") return 2;
  return 3;
}

OpenMCTally::OpenMCTally(const InputParameters &p, Problem &fe) : _params(p), _fe(fe)
{
  _id = (int)p.getInt("id");
  const std::string particle = p.getString("particle_type", "neutron");
  const std::string est = p.getString("estimator", "COLLISION");
  _scores = p.getWords("scores");
  _filters = p.getWords("filters");
  if (!p.isParamValid("scores")) paramError(p, "scores", "missing required parameter");
  if (!p.isParamValid("filters")) paramError(p, "filters", "missing required parameter");
  _nuclides = p.isParamValid("nuclides") ? p.getWords("nuclides") : std::vector<std::string>{"total"};
  _energy_bins = p.getReals("energy_bins");
  _cell_bins = p.getInts("cell_bins");
  _block_bins = p.getInts("block_bins");
  // Parameter checks (OpenMCTally.C:87-101); the track-length estimator is
  // supported here as well as the collision estimator
  if (particle != "neutron") paramError(p, "particle_type", "Only neutrons are currently supported.");
  if (est == "COLLISION") _estimator = MCW_EST_COLLISION;
  else if (est == "TRACKLENGTH") _estimator = MCW_EST_TRACKLENGTH;
  else paramError(p, "estimator", "Only collision and tracklength estimators are currently supported");
  const std::string exec = p.getString("execute_on", "initial");
  std::string lower = exec;
  std::transform(lower.begin(), lower.end(), lower.begin(), ::tolower);
  if (lower != "initial")
    paramError(p, "execute_on",
               "execute_on must be INITIAL to ensure tallies are created once at the beginning of the simulation");
  if (std::find(_filters.begin(), _filters.end(), "energy") != _filters.end() && !p.isParamValid("energy_bins"))
    mooseError("The energy bins parameter should be supplied if an energy filter is requested.");
}

void OpenMCTally::initialize()
{
  // Sort the filters into a specific order
  std::stable_sort(_filters.begin(), _filters.end(),
                   [](const std::string &a, const std::string &b) { return filter_rank(a) < filter_rank(b); });
  std::vector<mcw_filter> f(_filters.size());
  std::vector<std::vector<int64_t>> ids(_filters.size());
  _filter_bins.clear();
  _universe_ids.clear();
  for (size_t i = 0; i < _filters.size(); ++i) {
    const std::string &name = _filters[i];
    memset(&f[i], 0, sizeof f[i]);
    if (name == "energy") {
      f[i].type = MCW_FILTER_ENERGY;
      f[i].n_bins = (int32_t)_energy_bins.size();
      f[i].edges = _energy_bins.data();
      _filter_bins.push_back((int64_t)_energy_bins.size() - 1);
    } else if (name == "particle") {
      f[i].type = MCW_FILTER_PARTICLE;
      _filter_bins.push_back(1);
    } else if (name == "universe") {
      f[i].type = MCW_FILTER_UNIVERSE;
      if (_params.isParamValid("block_bins")) {
        for (auto b : _block_bins) ids[i].push_back(b);
      } else {
        std::set<int64_t> s(_fe.mesh.subdomain.begin(), _fe.mesh.subdomain.end());
        ids[i].assign(s.begin(), s.end());
      }
      _universe_ids = ids[i];
      f[i].n_bins = (int32_t)ids[i].size();
      f[i].ids = ids[i].data();
      _filter_bins.push_back((int64_t)ids[i].size());
    } else if (name == "cell") {
      f[i].type = MCW_FILTER_CELL;
      if (_params.isParamValid("cell_bins"))
        for (auto c : _cell_bins) ids[i].push_back(c);
      else
        for (int64_t e = 0; e < _fe.mesh.n_elem(); ++e) ids[i].push_back(e);
      f[i].n_bins = (int32_t)ids[i].size();
      f[i].ids = ids[i].data();
      _filter_bins.push_back((int64_t)ids[i].size());
      _cell_pos.clear();
      for (int64_t b = 0; b < (int64_t)ids[i].size(); ++b) _cell_pos.emplace(ids[i][b], b);
    } else {
      mooseError("Unrecognized filter");
    }
  }
  _strides.assign(_filters.size(), 0);
  int64_t stride = 1;
  for (int k = (int)_filters.size() - 1; k >= 0; --k) {
    _strides[k] = stride;
    stride *= _filter_bins[k];
  }
  std::vector<int32_t> sc;
  for (auto &s : _scores) {
    if (s == "flux") sc.push_back(MCW_SCORE_FLUX);
    else if (s == "total") sc.push_back(MCW_SCORE_TOTAL);
    else if (s == "absorption") sc.push_back(MCW_SCORE_ABSORPTION);
    else if (s == "scatter" || s == "elastic") sc.push_back(MCW_SCORE_SCATTER);
    else if (s == "fission") sc.push_back(MCW_SCORE_FISSION);
    else if (s == "kappa-fission") sc.push_back(MCW_SCORE_KAPPA_FISSION);
    else paramError(_params, "scores", "unknown score '" + s + "'");
  }
  // nuclide bins: "all" = every nuclide + total, "total" alone, else the list + total
  _nuclide_idx.clear();
  auto index_of = [&](const std::string &n) {
    for (size_t q = 0; q < _fe.nuclides.size(); ++q)
      if (_fe.nuclides[q].name == n) return (int32_t)q;
    paramError(_params, "nuclides", "unknown nuclide '" + n + "'");
  };
  if (_nuclides[0] == "all") {
    for (size_t q = 0; q < _fe.nuclides.size(); ++q) _nuclide_idx.push_back((int32_t)q);
    _nuclide_idx.push_back(-1);
  } else if (_nuclides[0] != "total") {
    for (auto &n : _nuclides) _nuclide_idx.push_back(index_of(n));
    _nuclide_idx.push_back(-1);
  } else {
    _nuclide_idx.push_back(-1);
  }
  check(_fe.study->ctx(),
        mcw_tally_add(_fe.study->ctx(), _id, _estimator, (int32_t)f.size(), f.data(), (int32_t)sc.size(), sc.data(),
                      (int32_t)_nuclide_idx.size(), _nuclide_idx.data()),
        "OpenMCTally " + _params.path());
}

int64_t OpenMCTally::cellBin(int64_t elem) const
{
  auto it = _cell_pos.find(elem);
  return it == _cell_pos.end() ? -1 : it->second;
}

// ------------------------------------------------------------------ OpenMCTallyAux

OpenMCTallyAux::OpenMCTallyAux(const InputParameters &p, Problem &fe) : _params(p), _fe(fe)
{
  if (!p.isParamValid("tally_id")) mooseError("Use tally ids, for now");
  _tally_id = (int)p.getInt("tally_id");
  const std::string g = p.getString("granularity");
  if (g == "global") _granularity = 0;
  else if (g == "universe") _granularity = 1;
  else if (g == "cell") _granularity = 2;
  else paramError(p, "granularity", "granularity must be one of global, universe, cell");
  _score = p.getString("score");
  // required MooseEnums with defaults (OpenMCTallyAux.C:34-43): an input may leave them out
  p.getString("estimator", "COLLISION");
  p.getString("particle_type", "neutron");
  _nuclide = p.getString("nuclide", "all");
  _energy_bin = (int)p.getInt("energy_bin", -1);
  _variable = p.getString("variable");
  // where (score, nuclide) sits among a tally's score bins: OpenMC's
  // results_ layout (nuclide-major, the default), or the index MaCaw's aux
  // computes, score_bin * n_nuclides + nuc_bin (OpenMCTallyAux.C:116-119);
  // they differ when a tally has several scores and several nuclide bins
  const std::string layout = p.getString("score_index_layout", "openmc");
  if (layout != "openmc" && layout != "macaw")
    paramError(p, "score_index_layout", "score_index_layout must be openmc or macaw");
  _score_major = layout == "macaw";
}

double OpenMCTallyAux::computeValue(int64_t elem, const std::vector<double> &sum) const
{
  const OpenMCTally *t = _fe.tally(_tally_id);
  if (!t) mooseError("no tally with id " + std::to_string(_tally_id));
  // nuclide bin: the last bin is the total when individual nuclides are binned
  const std::vector<int32_t> &nb = t->nuclideIndices();
  int nuc_bin = 0;
  if (nb.size() > 1) {
    if (_nuclide == "all") {
      nuc_bin = (int)nb.size() - 1;
    } else {
      nuc_bin = -1;
      for (size_t q = 0; q + 1 < nb.size(); ++q)
        if (_fe.nuclides[nb[q]].name == _nuclide) nuc_bin = (int)q;
      if (nuc_bin < 0) paramError(_params, "nuclide", "nuclide '" + _nuclide + "' is not binned in the tally");
    }
  }
  auto it = std::find(t->scores().begin(), t->scores().end(), _score);
  if (it == t->scores().end()) paramError(_params, "score", "The specified score does not exist in given tally");
  const int score_bin = (int)(it - t->scores().begin());
  const int64_t score_index = _score_major ? (int64_t)score_bin * (int64_t)nb.size() + nuc_bin
                                           : (int64_t)nuc_bin * (int64_t)t->scores().size() + score_bin;
  const int64_t nsb = t->nScoreBins();

  int64_t univ_end = 1, univ_stride = 1, energy_start = 0, energy_end = 1, energy_stride = 1, cell_stride = 1;
  int64_t cell_bin = 0, univ_bin = -1;
  bool has_cell = false, has_univ = false;
  for (size_t i = 0; i < t->filters().size(); ++i) {
    const std::string &f = t->filters()[i];
    if (f == "universe") {
      univ_end = t->filterBins()[i];
      univ_stride = t->strides()[i];
      has_univ = true;
      const auto &ids = t->universeIds();
      auto u = std::find(ids.begin(), ids.end(), (int64_t)_fe.mesh.subdomain[elem]);
      univ_bin = u == ids.end() ? -1 : (int64_t)(u - ids.begin());
    } else if (f == "energy") {
      if (_energy_bin < 0) {
        energy_end = t->filterBins()[i];
      } else {
        energy_start = _energy_bin;
        energy_end = energy_start + 1;
      }
      energy_stride = t->strides()[i];
    } else if (f == "cell") {
      // If the local cell is not in the tally cell filter, tally value is 0
      has_cell = true;
      cell_stride = t->strides()[i];
      cell_bin = t->cellBin(elem);
      if (cell_bin < 0) return 0.0;
    }
  }
  double val = 0.0;
  switch (_granularity) {
    case 0:
      mooseError("Global Tally retrieving currently not supported");
    case 1:
      if (!has_univ) mooseError("Tally does not have a universe filter");
      if (has_cell)
        mooseError("Universe granularity does not currently support tallies"
                   "with both universe and cell filters");
      if (univ_bin < 0) return 0.0;
      for (int64_t j = energy_start; j < energy_end; ++j)
        val += sum[(univ_bin * univ_stride + j * energy_stride) * nsb + score_index];
      break;
    default:
      if (!has_cell) mooseError("Specified tally does not contain a cell filter");
      for (int64_t i = 0; i < univ_end; ++i)
        for (int64_t j = energy_start; j < energy_end; ++j)
          val += sum[(i * univ_stride + cell_bin * cell_stride + j * energy_stride) * nsb + score_index];
      break;
  }
  return val;
}

// ------------------------------------------------------------------ OpenMCStudy

OpenMCStudy::OpenMCStudy(const InputParameters &p, Problem &fe) : _params(p), _fe(fe)
{
  _verbose = p.getBool("verbose", true);
  const std::string exec = p.getString("execute_on", "TIMESTEP_END");
  if (exec != "TIMESTEP_END" && exec != "timestep_end")
    paramError(p, "execute_on", "the study runs one batch per time step: execute_on = TIMESTEP_END");
}

void OpenMCStudy::initialSetup()
{
  if (_verbose) std::cout << "Initializing OpenMC simulation" << std::endl;
  Problem &fe = _fe;
  int rc = mcw_create(&_ctx, &fe.options);
  check(_ctx, rc, "mcw_create");
  const MeshData &m = fe.mesh;
  check(_ctx,
        mcw_mesh_set(_ctx, m.n_nodes(), m.xyz.data(), m.n_elem(), m.type.data(), m.conn_off.data(), m.conn.data(),
                     m.subdomain.data(), m.face_boundary.data()),
        "mcw_mesh_set");
  std::vector<mcw_nuclide> nd;
  for (auto &n : fe.nuclides)
    nd.push_back(mcw_nuclide{(int32_t)n.E.size(), n.awr, n.E.data(), n.t.data(), n.a.data(), n.s.data(),
                             n.f.empty() ? nullptr : n.f.data(), {n.nu[0], n.nu[1]}, {n.watt[0], n.watt[1]},
                             n.temperature, n.kappa});
  check(_ctx, mcw_xs_set(_ctx, (int32_t)nd.size(), nd.data()), "mcw_xs_set");
  for (size_t q = 0; q < fe.nuclides.size(); ++q)
    for (auto &n : fe.nuclides[q].temps) {
      mcw_nuclide t{(int32_t)n.E.size(), n.awr, n.E.data(), n.t.data(), n.a.data(), n.s.data(),
                    n.f.empty() ? nullptr : n.f.data(), {n.nu[0], n.nu[1]}, {n.watt[0], n.watt[1]}, n.temperature,
                    n.kappa};
      check(_ctx, mcw_xs_add_temperature(_ctx, (int32_t)q, &t), "mcw_xs_add_temperature");
    }
  // CollisionKernel.C:109: sqrt(kT) from the element's temperature variable
  if (!fe.elem_temperature.empty())
    check(_ctx,
          mcw_temperature_set(_ctx, (int64_t)fe.elem_temperature.size(), fe.elem_temperature.data(),
                              fe.temperature_tolerance),
          "mcw_temperature_set");
  std::vector<mcw_material> md;
  for (auto &mt : fe.materials)
    md.push_back(mcw_material{mt.id, (int32_t)mt.nuclides.size(), mt.nuclides.data(), mt.densities.data()});
  check(_ctx, mcw_materials_set(_ctx, (int32_t)md.size(), md.data()), "mcw_materials_set");
  check(_ctx, mcw_source_set(_ctx, &fe.source), "mcw_source_set");
}

OpenMCStudy::~OpenMCStudy()
{
  if (!_ctx) return;
  if (_verbose) std::cout << "Finalizing OpenMC" << std::endl;
  mcw_destroy(_ctx);
}

void OpenMCStudy::generateRays()
{
  ++_t_step;
  if (_verbose) std::cout << "Generating new rays" << std::endl;
  if (_fe.particles < 1) mooseError("No Rays were moved to _rays in defineRays()");
}

void OpenMCStudy::execute()
{
  if (_verbose) std::cout << "Executing study" << std::endl;
  check(_ctx, mcw_run_batch(_ctx, _t_step, _fe.particles, &_stats), "mcw_run_batch");
}

void OpenMCStudy::postExecuteStudy()
{
  if (_verbose)
    std::cout << "Batch " << _t_step << ": " << _stats.histories << " histories, " << _stats.segments << " segments, "
              << _stats.collisions << " collisions, leakage " << (double)_stats.leaks / _stats.histories << ", "
              << _stats.time_ms << " ms" << std::endl;
  if (_fe.options.run_mode == MCW_RUN_EIGENVALUE) {
    check(_ctx, mcw_keff_get(_ctx, &_keff), "mcw_keff_get");
    // OpenMC's output is silenced, so keff is printed (OpenMCStudy.C:398-401)
    std::cout << "Keff " << _keff.keff << " (" << _keff.keff_std << ") Generation: " << _keff.k_generation
              << std::endl;
  }
}

// ------------------------------------------------------------------ Problem

static bool active(const HitNode &parent, const HitNode &child)
{
  if (!parent.has("active")) return !parent.has("inactive") || [&] {
    for (auto &w : split_words(parent.get("inactive")))
      if (w == child.name) return false;
    return true;
  }();
  for (auto &w : split_words(parent.get("active")))
    if (w == child.name) return true;
  return false;
}

Problem::Problem(std::unique_ptr<HitNode> root, const std::string &input) : _root(std::move(root)), _input(input)
{
  HitNode &R = *_root;
  HitNode *mb = R.child("Mesh");
  if (!mb) mooseError("the input has no [Mesh] block");
  const size_t slash = input.rfind('/');
  mesh = build_mesh(*mb, slash == std::string::npos ? std::string(".") : input.substr(0, slash));
  for (const char *vb : {"Variables", "AuxVariables"})
    if (HitNode *v = R.child(vb))
      for (auto &c : v->children) {
        InputParameters p(c.get());
        std::string order = p.getString("order", "FIRST");
        aux_vars[c->name] = order == "CONSTANT" ? 0 : (order == "FIRST" ? 1 : 2);
        aux_family[c->name] = p.getString("family", "LAGRANGE");
        aux_initial[c->name] = p.getReal("initial_condition", 0.0);
      }
  HitNode *mc = R.child("MonteCarlo");
  if (!mc) mooseError("the input has no [MonteCarlo] block (settings, source, nuclides, materials)");
  read_monte_carlo(*mc);

  if (HitNode *rk = R.child("RayKernels"))
    for (auto &c : rk->children) {
      if (!active(*rk, *c)) continue;
      InputParameters p(c.get());
      if (p.getString("type") != "CollisionKernel") paramError(p, "type", "unknown RayKernel type");
      raykernels.push_back(std::make_unique<CollisionKernel>(p, *this));
    }
  if (!raykernels.empty()) {
    bool several = false;
    for (auto &n : nuclides) several |= !n.temps.empty();
    // one table per nuclide reads the same data at every temperature
    if (several) elem_temperature = element_values(raykernels[0]->temperatureVariable());
  }
  if (HitNode *rb = R.child("RayBCs"))
    for (auto &c : rb->children) {
      if (!active(*rb, *c)) continue;
      InputParameters p(c.get());
      const std::string t = p.getString("type");
      if (t == "ReflectRayBC") raybcs.push_back(std::make_unique<RayBC>(p, *this, MCW_BC_REFLECT));
      else if (t == "KillRayBC") raybcs.push_back(std::make_unique<RayBC>(p, *this, MCW_BC_VACUUM));
      else paramError(p, "type", "unknown RayBC type");
    }
  HitNode *uo = R.child("UserObjects");
  if (!uo) mooseError("the input has no [UserObjects] block");
  for (auto &c : uo->children) {
    if (!active(*uo, *c)) continue;
    InputParameters p(c.get());
    const std::string t = p.getString("type");
    if (t == "OpenMCStudy") {
      if (study) paramError(p, "type", "only one OpenMCStudy per input");
      study = std::make_unique<OpenMCStudy>(p, *this);
    }
  }
  if (!study) mooseError("the input has no OpenMCStudy user object");
  for (auto &c : uo->children) {
    if (!active(*uo, *c)) continue;
    InputParameters p(c.get());
    const std::string t = p.getString("type");
    if (t == "OpenMCTally") tallies.push_back(std::make_unique<OpenMCTally>(p, *this));
    else if (t != "OpenMCStudy") paramError(p, "type", "unknown UserObject type '" + t + "'");
  }
  if (HitNode *ak = R.child("AuxKernels"))
    for (auto &c : ak->children) {
      if (!active(*ak, *c)) continue;
      InputParameters p(c.get());
      const std::string t = p.getString("type");
      if (t != "OpenMCTallyAux") continue; // other aux kernels are not on this path
      auxkernels.push_back(std::make_unique<OpenMCTallyAux>(p, *this));
      if (!aux_vars.count(auxkernels.back()->variable()))
        paramError(p, "variable", "no auxiliary variable '" + auxkernels.back()->variable() + "'");
    }
  if (HitNode *ex = R.child("Executioner")) {
    InputParameters p(ex);
    if (p.getString("type", "Transient") != "Transient")
      mooseError("Unable to set the number of batches for a non Transient Executioner");
  }
  file_base = input.substr(0, input.rfind('.')) + "_out";
  if (HitNode *o = R.child("Outputs")) {
    InputParameters p(o);
    csv = p.getBool("csv", true);
    if (p.isParamValid("file_base")) file_base = p.getString("file_base");
    for (auto &c : o->children) {
      if (!active(*o, *c)) continue;
      InputParameters q(c.get());
      const std::string type = q.getString("type", "");
      if (type != "RayTracingExodus")
        paramError(q, "type", "output type '" + type + "' is not supported (RayTracingExodus is)");
      if (q.isParamValid("study") && q.getString("study") != "study" && !R.find("UserObjects/" + q.getString("study")))
        paramError(q, "study", "no study '" + q.getString("study") + "'");
      for (auto &w : q.getWords("execute_on"))
        if (w != "final") paramError(q, "execute_on", "only execute_on = final is supported for ray output");
      ray_outputs.push_back(c->name);
    }
  }
}

Problem::~Problem() = default;

const OpenMCTally *Problem::tally(int id) const
{
  for (auto &t : tallies)
    if (t->id() == id) return t.get();
  return nullptr;
}

void Problem::read_monte_carlo(const HitNode &mcb)
{
  InputParameters p(&mcb);
  mcw_default_options(&options);
  particles = p.getInt("particles");
  n_batches = (int)p.getInt("batches");
  options.master_seed = (uint64_t)p.getInt("seed", 1);
  options.seed_roundtrip = p.getBool("seed_roundtrip", false) ? 1 : 0;
  const std::string mode = p.getString("mode", "batched");
  if (mode == "batched") options.mode = MCW_MODE_BATCHED;
  else if (mode == "plain") options.mode = MCW_MODE_PLAIN;
  else paramError(p, "mode", "mode must be batched or plain");
  options.lanes_min = (int32_t)p.getInt("lanes_min", 0);
  const std::string run_mode = p.getString("run_mode", "fixed source");
  if (run_mode == "eigenvalue") options.run_mode = MCW_RUN_EIGENVALUE;
  else if (run_mode == "fixed source" || run_mode == "fixed_source") options.run_mode = MCW_RUN_FIXED_SOURCE;
  else paramError(p, "run_mode", "run_mode must be eigenvalue or fixed source");
  options.n_inactive = (int32_t)p.getInt("inactive", 0);
  // how a CollisionKernel material id selects a material: by id, or as
  // MaCaw's `id - 1` index into the material list (CollisionKernel.C:110)
  const std::string mi = p.getString("material_index", "id");
  if (mi == "id") options.material_index = MCW_MATERIAL_BY_ID;
  else if (mi == "index") options.material_index = MCW_MATERIAL_BY_INDEX;
  else paramError(p, "material_index", "material_index must be id or index");
  // OpenMC settings temperature_tolerance; by default the nearest table is
  // read however far, as OpenMC does for the sqrt(kT) MaCaw sets per segment
  temperature_tolerance = p.getReal("temperature_tolerance", INFINITY);
  if (!(temperature_tolerance >= 0.0)) paramError(p, "temperature_tolerance", "the tolerance must be >= 0 K");
  if (options.n_inactive < 0) paramError(p, "inactive", "the number of inactive batches must not be negative");
  if (options.run_mode == MCW_RUN_EIGENVALUE && options.n_inactive >= n_batches)
    paramError(p, "inactive", "an eigenvalue run needs more batches than inactive batches");
  if (options.run_mode == MCW_RUN_EIGENVALUE && options.mode != MCW_MODE_BATCHED)
    paramError(p, "mode", "eigenvalue runs need mode = batched");
  HitNode *nb = mcb.child("Nuclides");
  HitNode *mb = mcb.child("Materials");
  HitNode *sb = mcb.child("source");
  if (!nb || !mb || !sb) mooseError("[MonteCarlo] needs [Nuclides], [Materials] and [source] sub-blocks");
  for (auto &c : nb->children) {
    InputParameters q(c.get());
    if (q.getString("type", "SyntheticNuclide") != "SyntheticNuclide") paramError(q, "type", "unknown nuclide type");
    nuclides.push_back(synthetic_nuclide(q));
  }
  for (auto &c : mb->children) {
    InputParameters q(c.get());
    MaterialData m;
    m.id = (int32_t)q.getInt("id");
    for (auto &n : q.getWords("nuclides")) {
      int32_t idx = -1;
      for (size_t k = 0; k < nuclides.size(); ++k)
        if (nuclides[k].name == n) idx = (int32_t)k;
      if (idx < 0) paramError(q, "nuclides", "unknown nuclide '" + n + "'");
      m.nuclides.push_back(idx);
    }
    m.densities = q.getReals("densities");
    if (m.densities.size() != m.nuclides.size()) paramError(q, "densities", "one density per nuclide");
    materials.push_back(m);
  }
  InputParameters s(sb);
  memset(&source, 0, sizeof source);
  const std::string space = s.getString("space", "point");
  if (space == "point") {
    source.space = MCW_SPACE_POINT;
    auto pt = s.getReals("point");
    if (pt.size() != 3) paramError(s, "point", "a point needs 3 coordinates");
    for (int k = 0; k < 3; ++k) source.point[k] = pt[k];
  } else if (space == "volume") {
    source.space = MCW_SPACE_MESH_VOLUME;
  } else if (space == "box" || space == "fission") {
    // settings.xml <space type="box"> or <space type="fission"> (a box whose
    // sites must fall in fissionable materials), <parameters>lower_left upper_right</parameters>
    source.space = MCW_SPACE_BOX;
    source.only_fissionable = space == "fission" ? 1 : 0;
    auto lo = s.getReals("lower_left"), hi = s.getReals("upper_right");
    if (lo.size() != 3) paramError(s, "lower_left", "a box corner needs 3 coordinates");
    if (hi.size() != 3) paramError(s, "upper_right", "a box corner needs 3 coordinates");
    for (int k = 0; k < 3; ++k) {
      if (!(hi[k] >= lo[k])) paramError(s, "upper_right", "the box's upper corner must not lie below its lower corner");
      source.box_lo[k] = lo[k];
      source.box_hi[k] = hi[k];
    }
  } else {
    paramError(s, "space", "space must be point, box, fission or volume");
  }
  source.angle = MCW_ANGLE_ISOTROPIC;
  if (s.isParamValid("watt_a")) {
    source.energy = MCW_ENERGY_WATT;
    source.watt_a = s.getReal("watt_a");
    source.watt_b = s.getReal("watt_b");
  } else {
    source.energy = MCW_ENERGY_MONO;
    source.E0 = s.getReal("energy", 2e6);
  }
}

// Per-element values of a CONSTANT MONOMIAL variable: its initial_condition
// (MOOSE's default 0), then each active [ICs] ConstantIC on it in input
// order, on its `block`s or everywhere.
std::vector<double> Problem::element_values(const std::string &var) const
{
  char *end = nullptr;
  const double c = std::strtod(var.c_str(), &end);
  if (!var.empty() && end && *end == '\0') return std::vector<double>(mesh.n_elem(), c); // a constant
  std::vector<double> v(mesh.n_elem(), aux_initial.count(var) ? aux_initial.at(var) : 0.0);
  if (HitNode *ic = _root->child("ICs"))
    for (auto &c : ic->children) {
      if (!active(*ic, *c)) continue;
      InputParameters p(c.get());
      if (p.getString("variable") != var) continue;
      if (p.getString("type") != "ConstantIC") paramError(p, "type", "only ConstantIC sets the temperature");
      const double x = p.getReal("value");
      std::set<int> blocks;
      if (p.isParamValid("block"))
        for (long long b : p.getInts("block")) blocks.insert((int)b);
      for (int64_t e = 0; e < mesh.n_elem(); ++e)
        if (blocks.empty() || blocks.count(mesh.subdomain[e])) v[e] = x;
    }
  return v;
}

std::map<int, int> Problem::blockMap() const
{
  std::map<int, int> bm;
  for (auto &k : raykernels)
    for (auto &e : k->blockToMaterial()) bm[e.first] = e.second;
  return bm;
}

void Problem::run()
{
  // initial setup (EXEC_INITIAL): study, kernels, BCs, tallies
  for (auto &k : raykernels) k->initialSetup();
  study->initialSetup();
  std::map<int, int> bm = blockMap();
  std::vector<int32_t> blocks, mats;
  for (auto &e : bm) {
    blocks.push_back(e.first);
    mats.push_back(e.second);
  }
  mcw_ctx *ctx = study->ctx();
  if (blocks.empty()) mooseError("no CollisionKernel maps a block to a material");
  check(ctx, mcw_block_material_map(ctx, (int32_t)blocks.size(), blocks.data(), mats.data()), "block map");
  for (auto &b : raybcs) b->apply(ctx);
  for (auto &t : tallies) {
    std::cout << "Creating new tally" << std::endl;
    t->initialize();
  }
  std::vector<std::vector<double>> rows;
  std::map<int, std::vector<double>> sums;
  for (int step = 1; step <= n_batches; ++step) {
    if (step == n_batches && !ray_outputs.empty()) {
      // RayTracingExodus at execute_on = final: trace every ray of the last batch
      mcw_options o = options;
      o.mode = MCW_MODE_BATCHED;
      o.record_histories = 1;
      o.trace_histories = (int32_t)std::min<int64_t>(particles, INT32_MAX);
      check(ctx, mcw_options_set(ctx, &o), "mcw_options_set");
    }
    study->generateRays();
    study->execute();
    study->postExecuteStudy();
    const mcw_batch_stats &st = study->lastStats();
    std::vector<double> row = {(double)step, (double)st.segments, (double)st.collisions,
                               (double)st.leaks / (double)st.histories, st.histories / (st.time_ms * 1e-3)};
    if (options.run_mode == MCW_RUN_EIGENVALUE) {
      const mcw_keff_result &k = study->lastKeff();
      row.insert(row.end(), {k.k_generation, k.keff, k.keff_std});
    }
    for (auto &a : auxkernels) {
      int64_t nf = 0, ns = 0;
      check(ctx, mcw_tally_shape(ctx, a->tallyId(), &nf, &ns), "mcw_tally_shape");
      std::vector<double> &sum = sums[a->tallyId()];
      sum.assign((size_t)(nf * ns), 0.0);
      check(ctx, mcw_tally_get(ctx, a->tallyId(), MCW_RESULT_SUM, sum.data(), sum.size()), "mcw_tally_get");
      double tot = 0.0;
      for (int64_t e = 0; e < mesh.n_elem(); ++e) tot += a->computeValue(e, sum);
      row.push_back(tot);
    }
    rows.push_back(row);
  }
  if (!ray_outputs.empty()) {
    int64_t n = 0;
    check(ctx, mcw_trace_get(ctx, nullptr, 0, &n), "mcw_trace_get");
    std::vector<mcw_trace_record> recs((size_t)n);
    if (n) check(ctx, mcw_trace_get(ctx, recs.data(), recs.size(), &n), "mcw_trace_get");
    const mcw_batch_stats &st = study->lastStats();
    if (st.trace_records > n)
      std::cerr << "*** Warning: ray output keeps " << n << " of " << st.trace_records << " visit records" << std::endl;
    for (auto &name : ray_outputs) {
      const std::string path = file_base + "_" + name + ".e";
      write_rays_exodus(path, recs, n_batches, particles, options.master_seed ? options.master_seed : 1,
                        path.substr(path.rfind('/') + 1));
    }
  }
  if (csv) {
    write_outputs(rows);
    for (auto &a : auxkernels) {
      std::ofstream f(file_base + "_" + a->variable() + ".csv");
      f << "elem," << a->variable() << "\n";
      f.precision(17);
      for (int64_t e = 0; e < mesh.n_elem(); ++e) f << e << "," << a->computeValue(e, sums[a->tallyId()]) << "\n";
    }
  }
}

void Problem::write_outputs(const std::vector<std::vector<double>> &rows)
{
  std::ofstream f(file_base + ".csv");
  f << "time,segments,collisions,leakage,histories_per_s";
  if (options.run_mode == MCW_RUN_EIGENVALUE) f << ",k_generation,keff,keff_std";
  for (auto &a : auxkernels) f << "," << a->variable() << "_total";
  f << "\n";
  f.precision(17);
  for (auto &r : rows) {
    for (size_t k = 0; k < r.size(); ++k) f << (k ? "," : "") << r[k];
    f << "\n";
  }
}

// Binary dump of the problem as the library received it, for replay on a
// CPU reference: little-endian, int64 counts then raw arrays.
void Problem::dump(const std::string &path) const
{
  std::ofstream f(path, std::ios::binary);
  auto w64 = [&](int64_t v) { f.write((const char *)&v, 8); };
  auto wd = [&](const void *p, size_t n) { f.write((const char *)p, n); };
  f.write("MCWD0005", 8);
  w64(mesh.n_nodes());
  wd(mesh.xyz.data(), mesh.xyz.size() * 8);
  w64(mesh.n_elem());
  wd(mesh.type.data(), mesh.type.size());
  wd(mesh.conn_off.data(), mesh.conn_off.size() * 8);
  w64((int64_t)mesh.conn.size());
  wd(mesh.conn.data(), mesh.conn.size() * 4);
  wd(mesh.subdomain.data(), mesh.subdomain.size() * 4);
  w64((int64_t)mesh.face_boundary.size());
  wd(mesh.face_boundary.data(), mesh.face_boundary.size() * 4);
  std::map<int, int> bc;
  for (auto &b : raybcs)
    for (int32_t id : b->ids()) bc[id] = b->kind();
  w64((int64_t)bc.size());
  for (auto &e : bc) {
    w64(e.first);
    w64(e.second);
  }
  std::map<int, int> bm = blockMap();
  w64((int64_t)bm.size());
  for (auto &e : bm) {
    w64(e.first);
    w64(e.second);
  }
  auto table = [&](const NuclideData &n) {
    wd(&n.awr, 8);
    w64((int64_t)n.E.size());
    wd(n.E.data(), n.E.size() * 8);
    wd(n.t.data(), n.t.size() * 8);
    wd(n.a.data(), n.a.size() * 8);
    wd(n.s.data(), n.s.size() * 8);
    w64(n.f.empty() ? 0 : 1);
    if (!n.f.empty()) wd(n.f.data(), n.f.size() * 8);
    wd(n.nu, 16);
    wd(n.watt, 16);
    wd(&n.temperature, 8);
    wd(&n.kappa, 8);
  };
  w64((int64_t)nuclides.size());
  for (auto &n : nuclides) {
    table(n);
    w64((int64_t)n.temps.size());
    for (auto &t : n.temps) table(t);
  }
  w64((int64_t)materials.size());
  for (auto &m : materials) {
    w64(m.id);
    w64((int64_t)m.nuclides.size());
    wd(m.nuclides.data(), m.nuclides.size() * 4);
    wd(m.densities.data(), m.densities.size() * 8);
  }
  wd(&source, sizeof source);
  w64(particles);
  w64(n_batches);
  w64(options.seed_roundtrip);
  w64((int64_t)options.master_seed);
  w64(options.run_mode);
  w64(options.n_inactive);
  w64((int64_t)elem_temperature.size());
  wd(elem_temperature.data(), elem_temperature.size() * 8);
  wd(&temperature_tolerance, 8);
  w64(options.material_index);
}

} // namespace mcwhost
