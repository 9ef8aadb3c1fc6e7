// san_oracle.cpp -- drives the CPU oracle and the library's host-side mesh
// derivation (mcw_mesh.cpp) through their C and C++ interfaces, for an
// AddressSanitizer + UndefinedBehaviorSanitizer build.
//
// Problems are built here in C++: a reflective HEX8 box, the same box with
// one corner cell split into six Kuhn TET4 (split interfaces between tets and
// hexes), fixed-source and eigenvalue batches, general tallies (filters,
// reaction scores, nuclide bins), element temperatures, and the setup errors
// the oracle must report instead of crashing.  Exit status 0 and no
// sanitizer report means every access stayed in bounds.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "macaw_amd.h"
#include "mcw_mesh.h"

extern "C" {
void *orc_create(const mcw_options *opt);
void orc_destroy(void *h);
const char *orc_last_error(void *h);
int orc_mesh_set(void *h, int64_t n_nodes, const double *xyz, int64_t n_elem, const uint8_t *elem_type,
                 const int64_t *conn_offsets, const int32_t *conn, const int32_t *subdomain,
                 const int32_t *face_boundary);
int orc_bc_set(void *h, int32_t bid, int32_t kind);
int orc_xs_set(void *h, int32_t n, const mcw_nuclide *nuc);
int orc_xs_add_temperature(void *h, int32_t q, const mcw_nuclide *t);
int orc_temperature_set(void *h, int64_t n, const double *kelvin, double tolerance);
int orc_materials_set(void *h, int32_t n, const mcw_material *m);
int orc_block_material_map(void *h, int32_t n, const int32_t *blocks, const int32_t *materials);
int orc_source_set(void *h, const mcw_source *s);
int orc_tally_add(void *h, int32_t id, int32_t estimator, int32_t n_filters, const mcw_filter *filters,
                  int32_t n_scores, const int32_t *scores, int32_t n_nuclides, const int32_t *nuclides);
int orc_tally_get(void *h, int32_t id, int32_t kind, double *out, size_t n);
int orc_run_batch(void *h, int64_t batch, int64_t n_particles, int32_t threads, int32_t record,
                  mcw_batch_stats *out);
int orc_history_get(void *h, mcw_history *out, size_t n);
int orc_keff_get(void *h, mcw_keff_result *out);
int orc_fission_bank_get(void *h, mcw_site *out, size_t max, int64_t *n_out);
int orc_source_bank_get(void *h, mcw_site *out, size_t max, int64_t *n_out);
int orc_sample_sources(void *h, int64_t batch, int64_t n_particles, int64_t n, double *out);
int64_t orc_locate(void *h, const double *p, double tol);
}

namespace {

int failures = 0;

void expect(bool ok, const std::string &what)
{
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what.c_str());
    ++failures;
  }
}

struct Mesh {
  std::vector<double> xyz;
  std::vector<uint8_t> type;
  std::vector<int64_t> off{0};
  std::vector<int32_t> conn, sub;
};

// n^3 HEX8 cells on [0, 1]^3 in libMesh node order; cells with i, j, k <
// m_tet are split into the six Kuhn TET4 along the 000 -> 111 diagonal
Mesh box(int n, int m_tet)
{
  Mesh M;
  const int m1 = n + 1;
  for (int i = 0; i <= n; ++i)
    for (int j = 0; j <= n; ++j)
      for (int k = 0; k <= n; ++k) {
        M.xyz.push_back((double)i / n);
        M.xyz.push_back((double)j / n);
        M.xyz.push_back((double)k / n);
      }
  auto node = [&](int i, int j, int k) { return (int32_t)((i * m1 + j) * m1 + k); };
  static const int KUHN[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < n; ++k) {
        const int blk = (i + j + k) % 2;
        if (i < m_tet && j < m_tet && k < m_tet) {
          for (auto &ax : KUHN) {
            int p[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 1}};
            p[1][ax[0]] = 1;
            p[2][ax[0]] = 1;
            p[2][ax[1]] = 1;
            int32_t t[4];
            for (int q = 0; q < 4; ++q) t[q] = node(i + p[q][0], j + p[q][1], k + p[q][2]);
            // orient positively
            const double *a = &M.xyz[3 * t[0]], *b = &M.xyz[3 * t[1]], *c = &M.xyz[3 * t[2]], *d = &M.xyz[3 * t[3]];
            const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            const double v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
            const double w[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
            const double vol = u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) +
                               u[2] * (v[0] * w[1] - v[1] * w[0]);
            if (vol < 0) std::swap(t[1], t[2]);
            M.conn.insert(M.conn.end(), t, t + 4);
            M.type.push_back(MCW_TET4);
            M.off.push_back((int64_t)M.conn.size());
            M.sub.push_back(blk);
          }
        } else {
          const int32_t h[8] = {node(i, j, k),         node(i + 1, j, k),     node(i + 1, j + 1, k),
                                node(i, j + 1, k),     node(i, j, k + 1),     node(i + 1, j, k + 1),
                                node(i + 1, j + 1, k + 1), node(i, j + 1, k + 1)};
          M.conn.insert(M.conn.end(), h, h + 8);
          M.type.push_back(MCW_HEX8);
          M.off.push_back((int64_t)M.conn.size());
          M.sub.push_back(blk);
        }
      }
  return M;
}

struct Table {
  std::vector<double> E, t, a, s, f;
  mcw_nuclide n;
};

// 1/v absorption plus one resonance and constant scattering on a log grid
Table nuclide(int np, double awr, double sa0, double sp, double fis_frac, double T)
{
  Table X;
  for (int i = 0; i < np; ++i) {
    const double E = 1e-5 * std::pow(2e12, (double)i / (np - 1));
    const double res = 40.0 / (1.0 + std::pow((E - 6.7) / 0.05, 2));
    const double a = sa0 * std::sqrt(0.0253 / E) + 0.5 * res;
    const double s = sp + 0.5 * res;
    X.E.push_back(E);
    X.a.push_back(a);
    X.s.push_back(s);
    X.t.push_back(a + s);
    X.f.push_back(fis_frac * a);
  }
  std::memset(&X.n, 0, sizeof X.n);
  X.n.n_points = np;
  X.n.awr = awr;
  X.n.energy = X.E.data();
  X.n.total = X.t.data();
  X.n.absorption = X.a.data();
  X.n.scatter = X.s.data();
  X.n.fission = fis_frac > 0 ? X.f.data() : nullptr;
  X.n.nu[0] = 2.43;
  X.n.watt[0] = 0.988e6;
  X.n.watt[1] = 2.249e-6;
  X.n.temperature = T;
  X.n.kappa = 193.4e6;
  return X;
}

void load(void *o, const Mesh &M, std::vector<Table> &nucs, int reflect)
{
  expect(orc_mesh_set(o, (int64_t)M.xyz.size() / 3, M.xyz.data(), (int64_t)M.type.size(), M.type.data(),
                      M.off.data(), M.conn.data(), M.sub.data(), nullptr) == 0,
         "mesh_set");
  expect(orc_bc_set(o, 0, reflect ? MCW_BC_REFLECT : MCW_BC_VACUUM) == 0, "bc_set");
  std::vector<mcw_nuclide> nd;
  for (auto &t : nucs) nd.push_back(t.n);
  expect(orc_xs_set(o, (int32_t)nd.size(), nd.data()) == 0, "xs_set");
  const int32_t n0[3] = {0, 1, 2}, n1[2] = {2, 0};
  const double d0[3] = {0.04, 0.02, 0.01}, d1[2] = {0.05, 0.03};
  const mcw_material mats[2] = {{7, 3, n0, d0}, {3, 2, n1, d1}};
  expect(orc_materials_set(o, 2, mats) == 0, "materials_set");
  const int32_t blocks[2] = {0, 1}, ids[2] = {3, 7};
  expect(orc_block_material_map(o, 2, blocks, ids) == 0, "block_material_map");
}

void add_tallies(void *o, int n_elem)
{
  const double edges[4] = {1e-5, 1.0, 1e4, 2e7};
  std::vector<int64_t> cells;
  for (int64_t e = n_elem - 1; e >= 0; e -= 3) cells.push_back(e);
  const int32_t flux[1] = {MCW_SCORE_FLUX};
  const int32_t many[6] = {MCW_SCORE_FLUX, MCW_SCORE_TOTAL, MCW_SCORE_ABSORPTION, MCW_SCORE_SCATTER,
                           MCW_SCORE_FISSION, MCW_SCORE_KAPPA_FISSION};
  const int32_t total[1] = {-1}, nucs[3] = {0, 2, -1};
  mcw_filter f1[1] = {{MCW_FILTER_CELL, 0, nullptr, nullptr}};
  mcw_filter f2[2] = {{MCW_FILTER_UNIVERSE, 0, nullptr, nullptr}, {MCW_FILTER_ENERGY, 4, nullptr, edges}};
  mcw_filter f3[3] = {{MCW_FILTER_PARTICLE, 0, nullptr, nullptr},
                      {MCW_FILTER_CELL, (int32_t)cells.size(), cells.data(), nullptr},
                      {MCW_FILTER_ENERGY, 4, nullptr, edges}};
  expect(orc_tally_add(o, 1, MCW_EST_TRACKLENGTH, 1, f1, 1, flux, 1, total) == 0, "tally 1");
  expect(orc_tally_add(o, 2, MCW_EST_COLLISION, 2, f2, 6, many, 3, nucs) == 0, "tally 2");
  expect(orc_tally_add(o, 3, MCW_EST_TRACKLENGTH, 3, f3, 6, many, 3, nucs) == 0, "tally 3");
}

void run_fixed(const Mesh &M, std::vector<Table> &nucs, int space, bool temps)
{
  mcw_options opt;
  std::memset(&opt, 0, sizeof opt);
  opt.master_seed = 1;
  void *o = orc_create(&opt);
  load(o, M, nucs, space == MCW_SPACE_POINT);
  if (temps) {
    std::vector<Table> more;
    more.push_back(nuclide(150, 1.0, 0.3, 20.0, 0.0, 900.0));
    expect(orc_xs_add_temperature(o, 0, &more[0].n) == 0, "xs_add_temperature");
    std::vector<double> T(M.type.size());
    for (size_t e = 0; e < T.size(); ++e) T[e] = 300.0 + 900.0 * (double)e / (double)T.size();
    expect(orc_temperature_set(o, (int64_t)T.size(), T.data(), INFINITY) == 0, "temperature_set");
  }
  mcw_source src;
  std::memset(&src, 0, sizeof src);
  src.space = space;
  src.angle = MCW_ANGLE_LINEAR_MU;
  src.axis[2] = 1.0;
  src.beta = 0.4;
  src.energy = MCW_ENERGY_WATT;
  src.watt_a = 0.988e6;
  src.watt_b = 2.249e-6;
  src.point[0] = src.point[1] = src.point[2] = 0.53;
  for (int k = 0; k < 3; ++k) {
    src.box_lo[k] = 0.3;
    src.box_hi[k] = 1.4; // partly outside: sites are resampled
  }
  expect(orc_source_set(o, &src) == 0, std::string("source_set: ") + orc_last_error(o));
  add_tallies(o, (int)M.type.size());
  const int64_t n = 300;
  for (int64_t b = 1; b <= 2; ++b) {
    mcw_batch_stats st;
    const int rc = orc_run_batch(o, b, n, 3, 1, &st);
    expect(rc == 0, std::string("run_batch: ") + orc_last_error(o));
    if (rc) break;
    expect(st.lost == 0 && st.segments > 0, "fixed-source counts");
  }
  std::vector<mcw_history> h(n);
  expect(orc_history_get(o, h.data(), h.size()) == 0, "history_get");
  std::vector<double> t(1 << 16);
  for (int id = 1; id <= 3; ++id)
    for (int kind = 0; kind < 3; ++kind) expect(orc_tally_get(o, id, kind, t.data(), t.size()) == 0, "tally_get");
  expect(orc_tally_get(o, 2, 0, t.data(), 1) != 0, "tally_get into a short buffer must fail");
  std::vector<double> sites(8 * 500);
  expect(orc_sample_sources(o, 1, 500, 500, sites.data()) == 0, "sample_sources");
  const double p[3] = {0.25, 0.75, 0.1};
  expect(orc_locate(o, p, 1e-12) >= 0, "locate");
  orc_destroy(o);
}

void run_eigen(const Mesh &M)
{
  std::vector<Table> nucs;
  nucs.push_back(nuclide(120, 235.0, 30.0, 10.0, 0.85, 0.0));
  nucs.push_back(nuclide(90, 1.0, 0.3, 20.0, 0.0, 0.0));
  nucs.push_back(nuclide(60, 16.0, 0.01, 3.8, 0.0, 0.0));
  mcw_options opt;
  std::memset(&opt, 0, sizeof opt);
  opt.master_seed = 1;
  opt.run_mode = MCW_RUN_EIGENVALUE;
  opt.n_inactive = 1;
  void *o = orc_create(&opt);
  load(o, M, nucs, 1);
  mcw_source src;
  std::memset(&src, 0, sizeof src);
  src.space = MCW_SPACE_BOX;
  src.energy = MCW_ENERGY_MONO;
  src.E0 = 2e6;
  src.only_fissionable = 1;
  for (int k = 0; k < 3; ++k) src.box_hi[k] = 1.0;
  expect(orc_source_set(o, &src) == 0, "eigen source_set");
  add_tallies(o, (int)M.type.size());
  for (int64_t b = 1; b <= 3; ++b) {
    mcw_batch_stats st;
    const int rc = orc_run_batch(o, b, 200, 2, 1, &st);
    expect(rc == 0, std::string("eigen run_batch: ") + orc_last_error(o));
    if (rc) break;
  }
  mcw_keff_result k;
  expect(orc_keff_get(o, &k) == 0 && k.n_active == 2, "keff_get");
  int64_t nb = 0;
  expect(orc_fission_bank_get(o, nullptr, 0, &nb) == 0, "fission_bank_get");
  std::vector<mcw_site> s((size_t)nb + 1);
  expect(orc_fission_bank_get(o, s.data(), s.size(), &nb) == 0, "fission_bank_get (copy)");
  expect(orc_source_bank_get(o, s.data(), s.size(), &nb) == 0 && nb == 200, "source_bank_get");
  orc_destroy(o);
}

void errors(const Mesh &M, std::vector<Table> &nucs)
{
  mcw_options opt;
  std::memset(&opt, 0, sizeof opt);
  opt.material_index = MCW_MATERIAL_BY_INDEX; // material id 7 is past the list of 2
  void *o = orc_create(&opt);
  load(o, M, nucs, 0);
  mcw_source src;
  std::memset(&src, 0, sizeof src);
  src.space = MCW_SPACE_MESH_VOLUME;
  src.energy = MCW_ENERGY_MONO;
  src.E0 = 1e6;
  expect(orc_source_set(o, &src) == 0, "error case source_set");
  mcw_batch_stats st;
  expect(orc_run_batch(o, 1, 10, 1, 0, &st) != 0, "material index out of range must fail");
  src.space = MCW_SPACE_POINT;
  src.point[0] = 5.0;
  expect(orc_source_set(o, &src) != 0, "point source outside the mesh must fail");
  src.point[0] = 0.5;
  src.angle = MCW_ANGLE_LINEAR_MU;
  src.axis[0] = 1.0;
  src.axis[1] = 1.0;
  expect(orc_source_set(o, &src) != 0, "non-unit axis must fail");
  orc_destroy(o);
}

void host_mesh(const Mesh &M)
{
  mcw::HostMesh h;
  h.n_nodes = (int64_t)M.xyz.size() / 3;
  h.n_elem = (int64_t)M.type.size();
  h.xyz = M.xyz;
  h.type = M.type;
  h.conn_off = M.off;
  h.conn = M.conn;
  h.subdomain = M.sub;
  const std::string err = mcw::derive_mesh(h, nullptr);
  expect(err.empty(), "derive_mesh: " + err);
  mcw::build_locator(h);
  for (int q = 0; q < 200; ++q) {
    const double p[3] = {(q % 7 + 0.5) / 7.0, (q % 11 + 0.5) / 11.0, (q % 13 + 0.5) / 13.0};
    expect(mcw::locate_point(h, p, 1e-12) >= 0, "locate_point");
  }
  const double out[3] = {2.0, 0.5, 0.5};
  expect(mcw::locate_point(h, out, 1e-12) < 0, "locate_point outside");
}

} // namespace

int main()
{
  std::vector<Table> nucs;
  nucs.push_back(nuclide(150, 1.0, 0.3, 20.0, 0.0, 0.0));
  nucs.push_back(nuclide(100, 12.0, 0.0035, 4.7, 0.0, 0.0));
  nucs.push_back(nuclide(80, 56.0, 2.56, 11.0, 0.0, 0.0));
  const Mesh hex = box(3, 0), mixed = box(3, 2);
  host_mesh(hex);
  host_mesh(mixed);
  run_fixed(hex, nucs, MCW_SPACE_POINT, false);
  run_fixed(mixed, nucs, MCW_SPACE_MESH_VOLUME, true);
  run_fixed(mixed, nucs, MCW_SPACE_BOX, false);
  run_eigen(hex);
  run_eigen(mixed);
  errors(hex, nucs);
  if (failures) return 1;
  std::printf("sanitized oracle and host mesh: ok\n");
  return 0;
}
