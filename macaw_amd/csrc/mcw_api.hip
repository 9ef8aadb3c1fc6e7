// mcw_api.hip -- the C ABI (include/macaw_amd.h): context, device memory,
// problem upload and the batch driver.
//
// A context plays the part of OpenMCStudy + OpenMC's global state for one
// GPU: mcw_create/mcw_destroy bracket openmc_init/openmc_finalize
// (src/userobjects/OpenMCStudy.C:127-147,213-215), mcw_run_batch is one
// Transient step (generateRays + execute + postExecuteStudy,
// OpenMCStudy.C:218-402) and mcw_tally_get is what OpenMCTallyAux reads
// (src/auxkernels/OpenMCTallyAux.C:180-203).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/macaw_amd.h"
#include "mcw_mesh.h"
#include "mcw_types.h"

namespace mcw {
hipError_t launch_source_bank(hipStream_t s, const Params &P, const SourceBank &S, int64_t n);
hipError_t launch_bank_init(hipStream_t s, const Params &P, const SourceBank &S, const BankView &B, int64_t n, int q);
hipError_t launch_segment(int ns, unsigned blocks, hipStream_t s, const Params &P, const BankView &in, int qin,
                          const BankView &out, int qout, const BankView &coll, int qcoll);
hipError_t launch_collide(unsigned blocks, hipStream_t s, const Params &P, const BankView &coll, int qcoll,
                          const BankView &out, int qout);
hipError_t launch_persistent(int ns, int wpe, bool p32, unsigned blocks, int block, hipStream_t s, const Params &P,
                             const SourceBank &S);
const void *persistent_kernel(int ns, int wpe, bool rec, bool p32, bool eig, bool stg);
size_t persistent_lds_bytes(const Params &P, int block);
hipError_t launch_stage_reduce(hipStream_t s, const StageView &V, double *tally, int64_t n_bins, int split);
hipError_t launch_tree_sum(hipStream_t s, const double *in, int64_t n, double *scratch, double *out);
hipError_t launch_fission_sites(hipStream_t s, const XSView &X, const FissionView &F, int64_t n);
hipError_t launch_fill_u64(hipStream_t s, unsigned long long *p, int64_t n, unsigned long long v);
hipError_t fission_sort_temp_bytes(int64_t n, size_t *bytes);
hipError_t launch_fission_sort(hipStream_t s, const FissionView &F, int64_t n, int end_bit, uint64_t *keys_tmp,
                               uint32_t *perm, uint32_t *idx, void *tmp, size_t tmp_bytes);
hipError_t launch_sync_count(hipStream_t s, int64_t total, int64_t base, double p_sample, uint64_t seed0,
                             int64_t *counts);
hipError_t scan_temp_bytes(int64_t n, size_t *bytes);
hipError_t launch_scan(hipStream_t s, const int64_t *in, int64_t *out, int64_t n, void *tmp, size_t tmp_bytes);
hipError_t launch_sync_scatter(hipStream_t s, const FissionView &F, const uint32_t *perm, const int64_t *counts,
                               const int64_t *offsets, int64_t total, int64_t n, const SourceBank &out,
                               uint64_t *out_key, int64_t next_batch, uint64_t master, int roundtrip);
hipError_t launch_sites_out(hipStream_t s, const double *x, const double *y, const double *z, const double *u,
                            const double *v, const double *w, const double *E, const int32_t *elem,
                            const uint64_t *key, const uint32_t *perm, int64_t n, const int32_t *orig_of_int,
                            mcw_site *out);
hipError_t launch_reset_queue(hipStream_t s, Counters *C, int q0, int q1);
hipError_t launch_reset_work(hipStream_t s, Counters *C);
hipError_t launch_fill(hipStream_t s, double *p, int64_t n, double v);
hipError_t launch_zero_counters(hipStream_t s, Counters *C);
hipError_t launch_gather_scale(hipStream_t s, const double *src, const int32_t *perm, int64_t n, double norm,
                               double *dst, int tail);
hipError_t launch_accumulate(hipStream_t s, double *value, double *sum, double *sum_sq, int64_t n, double norm);
hipError_t launch_math_probe(hipStream_t s, const double *x, int64_t n, double *lo, double *so, double *co,
                             double *sq, const uint64_t *seeds, double *po, uint64_t *seo);
hipError_t launch_xs_probe(hipStream_t s, const XSView &X, int32_t nuc, const double *E, int64_t n, int32_t *k,
                           double *xs);
hipError_t launch_atomic_probe(hipStream_t s, double *a, uint64_t mask, unsigned blocks, int32_t iters);
hipError_t launch_gather_probe(hipStream_t s, const double4 *a, uint64_t mask, unsigned blocks, int32_t iters,
                               double *out);
hipError_t launch_source_req(hipStream_t s, const Params &P, const SourceBank &S, int32_t n_slices, int64_t n);
const void *serve_kernel(int ns, bool p32);
hipError_t sort_temp_bytes(int64_t n, size_t *bytes);
hipError_t launch_sort_sites(hipStream_t s, const MeshView &M, const SourceBank &in, const SourceBank &out,
                             int64_t n, int shift, int bits, uint32_t mat_mask, uint32_t *k0, uint32_t *k1,
                             uint32_t *v0, uint32_t *v1, void *tmp, size_t tmp_bytes);
hipError_t launch_sort_bank(hipStream_t s, const MeshView &M, const BankView &in, const BankView &out, int64_t n,
                            bool rec, Counters *C, int qout, int shift, int bits, uint32_t mat_mask, uint32_t *k0,
                            uint32_t *k1, uint32_t *v0, uint32_t *v1, void *tmp, size_t tmp_bytes);
hipError_t launch_serve(int ns, bool p32, unsigned blocks, hipStream_t s, const Params &P, const SourceBank &S);
size_t serve_lds_bytes();
} // namespace mcw

using namespace mcw;

namespace {

struct McwError {
  std::string msg;
};

#define HIPCHK(expr)                                                                         \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess)                                                                    \
      throw McwError{std::string(#expr) + ": " + hipGetErrorString(_e)};                     \
  } while (0)

// Device allocation owned by a context.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release()
  {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t n)
  {
    if (n == 0) n = 8;
    if (p && bytes >= n) return;
    release();
    HIPCHK(hipMalloc(&p, n));
    bytes = n;
  }
  template <class T> void upload(const std::vector<T> &v, hipStream_t s)
  {
    alloc(v.size() * sizeof(T));
    if (!v.empty()) HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  }
  template <class T> T *as() const { return (T *)p; }
};

struct HostNuclide {
  double awr;
  std::vector<double> E;
  std::vector<double> t, a, s, f;
  double nu[2] = {0.0, 0.0}, watt[2] = {0.0, 0.0};
  double temperature = 293.6; // K
  double kappa = 0.0;         // eV per fission
};

constexpr double K_BOLTZMANN = MCW_K_BOLTZMANN;

struct HostMaterial {
  int32_t id;
  std::vector<int32_t> nuc;
  std::vector<double> dens;
};

struct HostTally {
  int32_t id, estimator;
  std::vector<int32_t> ftype;
  std::vector<std::vector<int64_t>> fids;
  std::vector<std::vector<double>> fedges;
  std::vector<int32_t> scores, nuclides;
  int64_t n_filter_bins = 1, n_score_bins = 1;
  std::vector<int64_t> strides, nbins;
  // device
  DevBuf value, sum, sum_sq, elem_base, edges, d_scores, d_nuclides;
  bool identity = false;
  int64_t max_base = -1; // largest spatial bin base an element scores into (-1: none)
  int32_t n_edges = 0;
  int64_t energy_stride = 0;
};

} // namespace

struct mcw_ctx {
  mcw_options opt;
  std::string err;
  hipStream_t stream = nullptr;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  int n_cu = 256;

  HostMesh mesh;
  bool have_mesh = false;
  std::map<int32_t, int32_t> bc;
  std::vector<HostNuclide> nucs;
  std::vector<std::vector<HostNuclide>> nuc_temps; // further tables per nuclide
  std::vector<double> elem_T;                      // element temperatures (caller order), empty: none
  double T_tol = 10.0;
  int32_t n_eff_mats = 0;                          // (material, table choice) pairs on the device
  std::vector<HostMaterial> mats;
  std::map<int32_t, int32_t> block_mat;
  bool have_block_map = false;
  mcw_source src;
  bool have_src = false;
  int64_t src_elem = -1;
  std::vector<std::unique_ptr<HostTally>> tallies;
  bool dirty = true;

  // device problem
  DevBuf plane, erec, mtet, mhex, split, bc_kind, cdf, cdf_guide, xyz, conn, conn_off, type, int_of_orig, orig_of_int;
  DevBuf cell_start, cell_elem;
  DevBuf grid_off, npts, rec, scatter, fis, fissionable, watt_t, nuc_of, kappa, mat_fis, awr, mat_off, mat_nuc, mat_dens, uni, uidx, uhash;
  DevBuf uhash2; // every other entry of uhash: the table of the staging kernels
  DevBuf counters;
  Counters *h_counters = nullptr; // pinned
  // staged flux tally: the record pool (scores, bins), per-chunk owner and
  // count, the chunk list by bucket and the per-bucket counts and cursors;
  // segments per history of the last staged batch
  DevBuf stg_sc, stg_el, stg_owner, stg_fill, stg_list, stg_count;
  int64_t stg_chunks = 0;
  double stg_seg_per_hist = 0.0;
  int64_t tl_bins = 0; // bins of the one plain track-length flux tally (tl_fast), else 0
  // source bank, live/collision banks of the event-based path, records
  DevBuf sb[9];
  int64_t sb_cap = 0;
  // sorting: (key, slot) pairs, radix-sort scratch, the sorted source bank
  DevBuf sort_k[2], sort_v[2], sort_tmp;
  int64_t sort_cap = 0;
  size_t sort_tmp_bytes = 0;
  DevBuf sb2[10];
  int64_t sb2_cap = 0;
  // visit records of traced histories (record mode)
  DevBuf trace;
  int64_t trace_cap = 0, trace_count = 0;
  int sort_shift = 0, sort_bits = 32;
  uint32_t sort_mat_mask = 0;
  DevBuf bank[3][17];
  int64_t bank_cap = 0;
  DevBuf r_digest, r_seed, r_E, r_x, r_y, r_z, r_nseg, r_ncoll, r_elem, r_fate;
  int64_t rec_n = 0;
  int64_t last_batch_n = 0;
  int64_t last_batch = 0;
  double e_min = 0, e_max = 0;
  int32_t n_bc = 0;

  // eigenvalue runs: the fission bank (11 arrays: x y z u v w E elem key seed
  // nuc), its
  // sort (keys, slot permutation), per-history keff scores, the resampling's
  // counts and offsets, and the next batch's source (9 arrays + key)
  DevBuf fb[11], fb_keys2, fb_perm, fb_perm2, fb_tmp, k_hist, k_red, sync_cnt, sync_off, scan_tmp, gen[10];
  int64_t fb_cap = 0, gen_cap = 0, khist_cap = 0, sync_cap = 0;
  size_t fb_tmp_bytes = 0, scan_tmp_bytes = 0;
  int64_t eig_batch = 0, eig_n = 0, eig_sites = 0;
  bool eig_src_ready = false;
  double keff = 1.0, k_sum = 0.0, k_sum_sq = 0.0;
  int32_t n_active = 0;
  mcw_keff_result kres{};

  Params base{};
  // this device's copy of the static scratch array (mcw_run_batch_noalloc)
  unsigned char *scratch = nullptr;
};

namespace {

int fail(mcw_ctx *c, const std::string &m)
{
  if (c) c->err = m;
  return 1;
}

template <class F> int guard(mcw_ctx *c, F &&f)
{
  if (!c) return 1;
  try {
    HIPCHK(hipSetDevice(c->opt.device));
    f();
    return 0;
  } catch (const McwError &e) {
    return fail(c, e.msg);
  } catch (const std::exception &e) {
    return fail(c, e.what());
  }
}

void require(bool ok, const std::string &m)
{
  if (!ok) throw McwError{m};
}

// A source's numbers: finite positions, positive finite energy parameters.
void check_source_values(const mcw_source &src)
{
  require(src.only_fissionable == 0 || (src.only_fissionable == 1 && src.space == MCW_SPACE_BOX),
          "only_fissionable is 0, or 1 for a box source");
  for (int k = 0; k < 3; ++k) {
    if (src.space == MCW_SPACE_POINT) require(std::isfinite(src.point[k]), "point source coordinates must be finite");
    if (src.space == MCW_SPACE_BOX)
      require(std::isfinite(src.box_lo[k]) && std::isfinite(src.box_hi[k]), "box source bounds must be finite");
  }
  if (src.energy == MCW_ENERGY_MONO)
    require(std::isfinite(src.E0) && src.E0 > 0.0, "source energy must be positive and finite");
  if (src.energy == MCW_ENERGY_WATT)
    require(std::isfinite(src.watt_a) && src.watt_a > 0.0 && std::isfinite(src.watt_b) && src.watt_b > 0.0,
            "Watt spectrum parameters a and b must be positive and finite");
}

// Isotropic, or linear in mu about a unit axis with |beta| <= 1.
void check_angle(const mcw_source &src)
{
  if (src.angle == MCW_ANGLE_LINEAR_MU) {
    const double n2 = (src.axis[0] * src.axis[0] + src.axis[1] * src.axis[1]) + src.axis[2] * src.axis[2];
    require(std::fabs(n2 - 1.0) <= 1e-12, "source axis must be a unit vector");
    require(std::fabs(src.beta) <= 1.0, "source beta must be in [-1, 1]");
  } else {
    require(src.angle == MCW_ANGLE_ISOTROPIC, "source angle type not supported");
  }
}

// ------------------------------------------------------------------ scratch
// Device memory of mcw_run_batch_noalloc: the batch's counters and its
// source bank live in one static array per device instead of context
// allocations.  Each GPU the library runs on holds its own copy for the life
// of the process.  Host code never passes `scratch` itself to a kernel or a
// copy: the name is a host-side stand-in, and a kernel handed it would fault.
// It uses the device address that hipGetSymbolAddress returns for the device
// that is current, looked up once per device and kept below.
constexpr size_t SCRATCH_BYTES = (size_t)64 << 20; // 64 MiB per device
constexpr size_t SCRATCH_COUNTERS = 256;            // Counters, padded to 256 B
// source bank: 8 arrays of 8 B (x y z u v w E seed) and 1 of 4 B (element)
// per history, each array starting on a 256-byte boundary
constexpr int64_t SCRATCH_CHUNK =
    (int64_t)((SCRATCH_BYTES - SCRATCH_COUNTERS - 9 * 256) / 68) / 256 * 256;
static_assert(sizeof(Counters) <= SCRATCH_COUNTERS, "counters fit their scratch slot");
static_assert(SCRATCH_CHUNK == MCW_NOALLOC_CHUNK, "the header states the scratch bank's size");
static_assert(SCRATCH_COUNTERS + 8 * 8 * SCRATCH_CHUNK + 4 * SCRATCH_CHUNK <= SCRATCH_BYTES, "the bank fits");
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

std::mutex scratch_mu;
std::map<int, unsigned char *> scratch_dev; // device ordinal -> device address of `scratch`

// The device address of this device's scratch array (device `dev` must be
// current); looked up once per device.
unsigned char *scratch_of(int dev)
{
  std::lock_guard<std::mutex> lk(scratch_mu);
  auto it = scratch_dev.find(dev);
  if (it != scratch_dev.end()) return it->second;
  unsigned char *ptr = nullptr;
  HIPCHK(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
  require(ptr != nullptr, "no device address for the scratch array");
  scratch_dev[dev] = ptr;
  return ptr;
}

// OpenMC init_seed on the host: id * 152917 steps of the LCG from master +
// stream (Brown's arbitrary-stride skip-ahead).
uint64_t mcwhost_init_seed(int64_t id, uint64_t stream, uint64_t master)
{
  const uint64_t M = 6364136223846793005ull, C = 1442695040888963407ull;
  uint64_t n = (uint64_t)id * 152917ull, g = M, cc = C, g_new = 1, c_new = 0;
  while (n) {
    if (n & 1) {
      g_new *= g;
      c_new = c_new * g + cc;
    }
    cc = (g + 1) * cc;
    g *= g;
    n >>= 1;
  }
  return g_new * (master + stream) + c_new;
}

// Upload everything the kernels read; validate the problem as
// CollisionKernel::initialSetup does (CollisionKernel.C:78-96).
void prepare(mcw_ctx *c, bool need_source = true)
{
  if (!c->dirty && (c->have_src || !need_source)) return;
  require(c->have_mesh, "no mesh: call mcw_mesh_set");
  require(!c->nucs.empty(), "no nuclear data: call mcw_xs_set");
  require(!c->mats.empty(), "no materials: call mcw_materials_set");
  require(c->have_block_map, "no block to material map: call mcw_block_material_map");
  require(c->have_src || !need_source, "no source: call mcw_source_set");
  if (!c->have_src) {
    memset(&c->src, 0, sizeof c->src);
    c->src.space = MCW_SPACE_MESH_VOLUME;
    c->src.E0 = 1e6;
  }
  hipStream_t s = c->stream;
  HostMesh &m = c->mesh;

  // boundary conditions: every external boundary needs a RayBC
  int32_t max_bid = -1;
  for (int32_t b : m.boundary_ids) max_bid = std::max(max_bid, b);
  std::vector<int8_t> kinds(max_bid + 1, 0);
  for (int32_t b : m.boundary_ids) {
    auto it = c->bc.find(b);
    require(it != c->bc.end(), "boundary " + std::to_string(b) + " has external sides but no RayBC");
    kinds[b] = (int8_t)it->second;
  }
  c->n_bc = max_bid + 1;

  // block -> OpenMC material id; every id must exist (CollisionKernel.C:85-90).
  // BY_ID uses the material with that id; BY_INDEX the (id-1)-th material of
  // the list, as MaCaw's `material() = id - 1` (CollisionKernel.C:110)
  std::map<int32_t, int32_t> id_to_idx;
  for (size_t i = 0; i < c->mats.size(); ++i) id_to_idx[c->mats[i].id] = (int32_t)i;
  const bool by_index = c->opt.material_index == MCW_MATERIAL_BY_INDEX;
  require(c->opt.material_index == MCW_MATERIAL_BY_ID || by_index, "unknown material_index option");
  std::map<int32_t, int32_t> used; // material id -> list position
  for (auto &bm : c->block_mat) {
    require(id_to_idx.count(bm.second), "Could not find material " + std::to_string(bm.second) + " in OpenMC materials.");
    if (by_index)
      require(bm.second >= 1 && bm.second <= (int32_t)c->mats.size(),
              "material index " + std::to_string(bm.second) + " - 1 is outside the " +
                  std::to_string(c->mats.size()) + " materials");
    used[bm.second] = by_index ? bm.second - 1 : id_to_idx[bm.second];
  }
  std::vector<int32_t> emat(m.n_elem, -1);
  for (int64_t e = 0; e < m.n_elem; ++e) {
    auto b = c->block_mat.find(m.subdomain[e]);
    if (b != c->block_mat.end()) emat[e] = used[b->second];
  }
  const int32_t n_nuc = (int32_t)c->nucs.size();
  for (auto &mt : c->mats)
    for (int32_t q : mt.nuc)
      require(q >= 0 && q < n_nuc, "material " + std::to_string(mt.id) + " refers to an unknown nuclide");

  // temperatures, OpenMC's "nearest" method: for each element and nuclide
  // the table whose kT is closest to the element's kT = (sqrt(k T))^2 (ties
  // to the lower temperature).  Elements with one material and one table
  // choice share a device material; every (nuclide, table) in use is a
  // device nuclide ("instance"), instances 0 .. n-1 being the first tables.
  auto table = [&](int32_t q, int32_t t) -> const HostNuclide & {
    return t == 0 ? c->nucs[q] : c->nuc_temps[q][t - 1];
  };
  std::vector<std::vector<int32_t>> t_order(n_nuc); // tables by ascending temperature
  for (int32_t q = 0; q < n_nuc; ++q) {
    const int32_t nt = 1 + (int32_t)c->nuc_temps[q].size();
    for (int32_t t = 0; t < nt; ++t) t_order[q].push_back(t);
    std::stable_sort(t_order[q].begin(), t_order[q].end(),
                     [&](int32_t a, int32_t b) { return table(q, a).temperature < table(q, b).temperature; });
  }
  std::vector<std::vector<double>> t_kT(n_nuc); // k T of each table, ascending
  for (int32_t q = 0; q < n_nuc; ++q)
    for (int32_t t : t_order[q]) t_kT[q].push_back(K_BOLTZMANN * table(q, t).temperature);
  std::vector<std::pair<int32_t, int32_t>> inst; // (nuclide, table)
  std::vector<std::vector<int32_t>> inst_of(n_nuc); // [nuclide][table] -> instance, -1 none yet
  for (int32_t q = 0; q < n_nuc; ++q) {
    inst_of[q].assign(t_order[q].size(), -1);
    inst_of[q][0] = q;
    inst.push_back({q, 0});
  }
  const bool temps = !c->elem_T.empty();
  require(!temps || (int64_t)c->elem_T.size() == m.n_elem,
          "mcw_temperature_set: one temperature per element of the mesh");
  std::vector<int32_t> eff(m.n_elem, -1);    // caller element -> device material
  std::vector<std::vector<int32_t>> eff_key; // material, then one instance per nuclide of it
  if (!temps) {
    // every element reads its nuclides' first tables: device material = material
    for (size_t mi = 0; mi < c->mats.size(); ++mi) {
      std::vector<int32_t> key{(int32_t)mi};
      key.insert(key.end(), c->mats[mi].nuc.begin(), c->mats[mi].nuc.end());
      eff_key.push_back(std::move(key));
    }
    eff = emat;
  } else {
    std::unordered_map<uint64_t, std::vector<int32_t>> by_hash; // key hash -> device materials
    std::vector<int32_t> key;
    for (int64_t e = 0; e < m.n_elem; ++e) {
      const int32_t mi = emat[e];
      if (mi < 0) continue;
      const double sq = std::sqrt(K_BOLTZMANN * c->elem_T[e]);
      const double kT = sq * sq;
      key.clear();
      key.push_back(mi);
      uint64_t h = 0xcbf29ce484222325ull ^ (uint64_t)mi;
      for (int32_t q : c->mats[mi].nuc) {
        const std::vector<double> &a = t_kT[q];
        size_t best = 0;
        double dmin = INFINITY;
        for (size_t k = 0; k < a.size(); ++k) {
          const double d = std::fabs(a[k] - kT);
          if (d < dmin) {
            dmin = d;
            best = k;
          }
        }
        const int32_t t = t_order[q][best];
        if (!(std::fabs(table(q, t).temperature - c->elem_T[e]) <= c->T_tol))
          throw McwError{"no cross sections for nuclide " + std::to_string(q) + " at or near " +
               std::to_string(c->elem_T[e]) + " K (element " + std::to_string(e) + ")"};
        int32_t &id = inst_of[q][t];
        if (id < 0) {
          id = (int32_t)inst.size();
          inst.push_back({q, t});
        }
        key.push_back(id);
        h = (h ^ (uint64_t)id) * 0x100000001b3ull;
      }
      std::vector<int32_t> &cand = by_hash[h];
      int32_t found = -1;
      for (int32_t k : cand)
        if (eff_key[k] == key) found = k;
      if (found < 0) {
        found = (int32_t)eff_key.size();
        eff_key.push_back(key);
        cand.push_back(found);
      }
      eff[e] = found;
    }
  }
  c->n_eff_mats = (int32_t)eff_key.size();

  // mesh in internal (Morton) order: planes per side slot, and per element
  // one 32-byte record of neighbour codes, material and side count
  // (a mixed mesh lays out its TET4 first in 4-slot, 128-byte blocks and its
  // HEX8 after in 8-slot, 256-byte blocks, so that an element's planes are
  // found from its id alone and start on a cache line)
  const int64_t ne = m.n_elem;
  auto slot_of = [&](int64_t i) -> int64_t {
    if (m.uniform_sides) return i * m.uniform_sides;
    return i < m.n_tet ? 4 * i : 4 * m.n_tet + 8 * (i - m.n_tet);
  };
  std::vector<double4> planes((size_t)slot_of(ne), make_double4(0.0, 0.0, 0.0, 0.0));
  std::vector<int4> erec(2 * ne);
  for (int64_t i = 0; i < ne; ++i) {
    const int64_t o = m.orig_of_int[i];
    const int64_t f0 = m.face_off[o];
    const int ns = (int)(m.face_off[o + 1] - f0);
    const int64_t slot = slot_of(i);
    int32_t code[6] = {-1, -1, -1, -1, -1, -1};
    for (int k = 0; k < ns; ++k) {
      const double *q = &m.plane[4 * (f0 + k)];
      planes[slot + k] = make_double4(q[0], q[1], q[2], q[3]);
      const int32_t nb = m.nbr[f0 + k];
      code[k] = nb < 0 || (nb & 7) == 7 ? nb : m.int_of_orig[nb >> 3] * 8 + (nb & 7);
    }
    erec[2 * i] = make_int4(code[0], code[1], code[2], code[3]);
    erec[2 * i + 1] = make_int4(code[4], code[5], eff[o], ns);
    if (!m.uniform_sides && i >= m.n_tet) memcpy(&planes[slot + 6], &erec[2 * i], 2 * sizeof(int4));
  }
  c->plane.upload(planes, s);
  c->erec.upload(erec, s);
  // split interfaces, with internal codes
  std::vector<SplitRec> sp(m.splits.size());
  for (size_t k = 0; k < m.splits.size(); ++k) {
    const HostMesh::Split &h = m.splits[k];
    sp[k].pl = make_double4(h.m[0], h.m[1], h.m[2], h.m[3]);
    for (int q = 0; q < 2; ++q) sp[k].code[q] = m.int_of_orig[h.elem[q]] * 8 + h.side[q];
    sp[k].pad[0] = sp[k].pad[1] = 0;
  }
  c->split.upload(sp, s);
  // mixed-precision records of the TET4 (internal ids 0 .. n_tet-1): FP32
  // planes relative to vertex 0
  if (m.n_tet > 0) {
    std::vector<MixedTet> mt(m.n_tet);
    for (int64_t i = 0; i < m.n_tet; ++i) {
      const int64_t o = m.orig_of_int[i];
      const double *v0 = &m.xyz[3 * (int64_t)m.conn[m.conn_off[o]]];
      MixedTet &R = mt[i];
      memset(&R, 0, sizeof R);
      const int64_t s0 = slot_of(i);
      for (int k = 0; k < 3; ++k) R.v0[k] = v0[k];
      for (int sd = 0; sd < 4; ++sd) {
        const double4 q = planes[s0 + sd];
        const double rel = q.w - ((q.x * v0[0] + q.y * v0[1]) + q.z * v0[2]);
        R.pl[sd] = make_float4((float)q.x, (float)q.y, (float)q.z, (float)rel);
      }
      R.code[0] = erec[2 * i].x;
      R.code[1] = erec[2 * i].y;
      R.code[2] = erec[2 * i].z;
      R.code[3] = erec[2 * i].w;
      R.mat = erec[2 * i + 1].z;
    }
    c->mtet.upload(mt, s);
  }
  // mixed meshes: mixed-precision records of the HEX8 (internal ids n_tet ..)
  if (!m.uniform_sides && ne > m.n_tet) {
    std::vector<MixedHex> mh(ne - m.n_tet);
    for (int64_t i = m.n_tet; i < ne; ++i) {
      MixedHex &R = mh[i - m.n_tet];
      memset(&R, 0, sizeof R);
      const int64_t s0 = slot_of(i);
      for (int sd = 0; sd < 6; ++sd) {
        const double4 q = planes[s0 + sd];
        R.pl[sd] = make_float4((float)q.x, (float)q.y, (float)q.z, (float)q.w);
      }
      const int4 a = erec[2 * i], b = erec[2 * i + 1];
      R.code[0] = a.x; R.code[1] = a.y; R.code[2] = a.z; R.code[3] = a.w;
      R.code[4] = b.x; R.code[5] = b.y;
      R.mat = b.z;
    }
    c->mhex.upload(mh, s);
  }
  c->bc_kind.upload(kinds, s);
  c->cdf.upload(m.cdf, s);
  c->cdf_guide.upload(m.cdf_guide, s);
  c->xyz.upload(m.xyz, s);
  c->conn.upload(m.conn, s);
  c->conn_off.upload(m.conn_off, s);
  c->type.upload(m.type, s);
  c->int_of_orig.upload(m.int_of_orig, s);
  c->orig_of_int.upload(m.orig_of_int, s);

  // nuclear data: per-point records (E, total, absorption, fission), scatter
  // apart (collisions and eigenvalue scores read fission; only scatter
  // tally scores read scatter)
  std::vector<int64_t> goff;
  std::vector<int32_t> np;
  std::vector<double> aw, sct;
  std::vector<double4> recs;
  c->e_min = INFINITY;
  c->e_max = -INFINITY;
  for (int32_t q = 0; q < n_nuc; ++q) // the grid range over every table loaded
    for (size_t t = 0; t < t_order[q].size(); ++t) {
      c->e_min = std::min(c->e_min, table(q, (int32_t)t).E.front());
      c->e_max = std::max(c->e_max, table(q, (int32_t)t).E.back());
    }
  std::vector<double> uni;
  std::vector<int32_t> nuc_of;
  for (auto &iq : inst) {
    const HostNuclide &n = table(iq.first, iq.second);
    nuc_of.push_back(iq.first);
    goff.push_back((int64_t)recs.size());
    np.push_back((int32_t)n.E.size());
    aw.push_back(n.awr);
    for (size_t j = 0; j < n.E.size(); ++j) {
      recs.push_back(make_double4(n.E[j], n.t[j], n.a[j], n.f.empty() ? 0.0 : n.f[j]));
      sct.push_back(n.s[j]);
    }
    uni.insert(uni.end(), n.E.begin(), n.E.end());
  }
  c->nuc_of.upload(nuc_of, s);
  c->grid_off.upload(goff, s);
  c->npts.upload(np, s);
  c->rec.upload(recs, s);
  c->scatter.upload(sct, s);
  c->awr.upload(aw, s);
  // fission yield and spectrum per nuclide, and which nuclides are fissionable
  std::vector<double4> fpar;
  std::vector<int32_t> fable;
  for (auto &iq : inst) {
    const HostNuclide &n = table(iq.first, iq.second);
    fpar.push_back(make_double4(n.nu[0], n.nu[1], n.watt[0], n.watt[1]));
    fable.push_back(n.f.empty() ? 0 : 1);
  }
  c->fis.upload(fpar, s);
  c->fissionable.upload(fable, s);
  std::vector<double> kap; // per instance: the nuclide's first table's kappa
  for (auto &iq : inst) kap.push_back(c->nucs[iq.first].kappa);
  c->kappa.upload(kap, s);
  // Watt first-attempt test: Ew <= (sqrt(w) + a sqrt(b) / 2)^2 and
  // w <= -a ln(r1 r2), so r1 r2 > exp(-w* / a) with w* = (sqrt(e_max) -
  // a sqrt(b) / 2)^2 keeps Ew below e_max; 1e-6 of w* is margin for rounding
  std::vector<double> wt;
  for (auto &iq : inst) {
    const HostNuclide &n = table(iq.first, iq.second);
    const double a = n.watt[0], b = n.watt[1];
    const double r = std::sqrt(c->e_max) - 0.5 * a * std::sqrt(b);
    wt.push_back(n.f.empty() || !(a > 0.0) || !(r > 0.0) ? 2.0 : std::exp(-(r * r) * (1.0 - 1e-6) / a));
  }
  c->watt_t.upload(wt, s);
  // unionized grid, per-nuclide index map and log-bucket table
  std::sort(uni.begin(), uni.end());
  uni.erase(std::unique(uni.begin(), uni.end()), uni.end());
  const int64_t nu = (int64_t)uni.size();
  const int32_t nn = (int32_t)inst.size();
  require(nu < ((int64_t)1 << 31) / std::max(1, nn), "unionized grid too large");
  std::vector<int32_t> uidx((size_t)nu * nn);
  for (int32_t q = 0; q < nn; ++q) {
    const std::vector<double> &E = table(inst[q].first, inst[q].second).E;
    int64_t j = -1;
    for (int64_t k = 0; k < nu; ++k) {
      while (j + 1 < (int64_t)E.size() && E[j + 1] <= uni[k]) ++j;
      uidx[(size_t)k * nn + q] = (int32_t)j;
    }
  }
  const int32_t nh = MAX_HASH;
  const double logU0 = std::log(uni.front());
  const double dl = (std::log(uni.back()) - logU0) / nh;
  std::vector<int32_t> hash(nh + 1);
  for (int32_t b = 0; b < nh; ++b) {
    const double Eb = std::exp(logU0 + b * dl);
    int64_t k = (int64_t)(std::upper_bound(uni.begin(), uni.end(), Eb) - uni.begin()) - 1;
    hash[b] = (int32_t)std::max<int64_t>(0, k);
  }
  hash[nh] = (int32_t)(nu - 1);
  c->uni.upload(uni, s);
  c->uidx.upload(uidx, s);
  c->uhash.upload(hash, s);
  static_assert(MAX_HASH == 2 * STG_HASH, "the staging table keeps every other bucket");
  std::vector<int32_t> hash2(STG_HASH + 1);
  for (int32_t b = 0; b <= STG_HASH; ++b) hash2[b] = hash[2 * b];
  c->uhash2.upload(hash2, s);
  std::vector<int32_t> moff{0}, mnuc;
  std::vector<double> mden;
  for (auto &key : eff_key) {
    const HostMaterial &mt = c->mats[key[0]];
    for (size_t q = 0; q < mt.nuc.size(); ++q) {
      mnuc.push_back(key[1 + q]);
      mden.push_back(mt.dens[q]);
    }
    moff.push_back((int32_t)mnuc.size());
  }
  c->mat_off.upload(moff, s);
  c->mat_nuc.upload(mnuc, s);
  c->mat_dens.upload(mden, s);
  std::vector<int32_t> mfis; // device material has a fissionable nuclide (OpenMC Material::fissionable_)
  for (size_t d = 0; d + 1 < moff.size(); ++d) {
    int32_t f = 0;
    for (int32_t q = moff[d]; q < moff[d + 1]; ++q) f |= fable[mnuc[q]];
    mfis.push_back(f);
  }
  c->mat_fis.upload(mfis, s);

  // point locator for point and box sources (and for the request server)
  build_locator(m);
  c->cell_start.upload(m.cell_start, s);
  c->cell_elem.upload(m.cell_elem, s);
  // source
  check_angle(c->src);
  check_source_values(c->src);
  c->src_elem = -1;
  if (c->src.space == MCW_SPACE_POINT) {
    c->src_elem = locate_point(m, c->src.point, 1e-12);
    require(c->src_elem >= 0, "point source outside the mesh");
  } else if (c->src.space == MCW_SPACE_BOX) {
    for (int k = 0; k < 3; ++k)
      require(c->src.box_hi[k] >= c->src.box_lo[k], "box source: box_hi must not be below box_lo");
  } else {
    require(c->src.space == MCW_SPACE_MESH_VOLUME, "unknown source space type");
  }

  require(c->src.energy == MCW_ENERGY_MONO || c->src.energy == MCW_ENERGY_WATT, "source energy type not supported");

  c->counters.alloc(sizeof(Counters));
  if (!c->h_counters) HIPCHK(hipHostMalloc((void **)&c->h_counters, sizeof(Counters), hipHostMallocDefault));

  // views
  Params &P = c->base;
  P = Params{};
  P.mesh.plane = c->plane.as<double4>();
  P.mesh.erec = c->erec.as<int4>();
  P.mesh.stride = m.uniform_sides;
  P.mesh.n_bc = c->n_bc;
  P.mesh.bc_kind = c->bc_kind.as<int8_t>();
  P.mesh.n_elem = m.n_elem;
  P.mesh.cdf = c->cdf.as<double>();
  P.mesh.cdf_guide = c->cdf_guide.as<int32_t>();
  P.mesh.n_guide = (int64_t)m.cdf_guide.size() - 1;
  P.mesh.cdf_guide_scale = m.n_elem > 0 ? (double)P.mesh.n_guide / m.cdf[m.n_elem - 1] : 0.0;
  P.mesh.xyz = c->xyz.as<double>();
  P.mesh.conn = c->conn.as<int32_t>();
  P.mesh.conn_off = c->conn_off.as<int64_t>();
  P.mesh.conn_off0 = m.conn_off.empty() ? 0 : m.conn_off[0];
  P.mesh.type = c->type.as<uint8_t>();
  P.mesh.int_of_orig = c->int_of_orig.as<int32_t>();
  P.mesh.orig_of_int = c->orig_of_int.as<int32_t>();
  P.mesh.mtet = m.n_tet > 0 ? c->mtet.as<MixedTet>() : nullptr;
  P.mesh.mhex = !m.uniform_sides && m.n_elem > m.n_tet ? c->mhex.as<MixedHex>() : nullptr;
  P.mesh.n_tet = m.n_tet;
  P.mesh.split = m.splits.empty() ? nullptr : c->split.as<SplitRec>();
  for (int k = 0; k < 3; ++k) {
    P.mesh.grid_lo[k] = m.grid_lo[k];
    P.mesh.grid_inv[k] = m.grid_inv[k];
    P.mesh.grid_n[k] = m.grid_n[k];
  }
  P.mesh.cell_start = c->cell_start.as<int32_t>();
  P.mesh.cell_elem = c->cell_elem.as<int32_t>();
  P.xs.grid_off = c->grid_off.as<int64_t>();
  P.xs.n_points = c->npts.as<int32_t>();
  P.xs.rec = c->rec.as<double4>();
  P.xs.scatter = c->scatter.as<double>();
  P.xs.fis = c->fis.as<double4>();
  P.xs.fissionable = c->fissionable.as<int32_t>();
  P.xs.watt_t = c->watt_t.as<double>();
  P.xs.awr = c->awr.as<double>();
  P.xs.U = c->uni.as<double>();
  P.xs.uidx = c->uidx.as<int32_t>();
  P.xs.hash = c->uhash.as<int32_t>();
  P.xs.n_union = (int32_t)nu;
  P.xs.n_hash = nh;
  P.xs.logU0 = logU0;
  P.xs.inv_dl = 1.0 / dl;
  P.xs.mat_off = c->mat_off.as<int32_t>();
  P.xs.mat_nuc = c->mat_nuc.as<int32_t>();
  P.xs.mat_dens = c->mat_dens.as<double>();
  P.xs.n_nuc = (int32_t)inst.size();
  P.xs.n_mat = c->n_eff_mats;
  P.xs.nuc_of = c->nuc_of.as<int32_t>();
  P.xs.kappa = c->kappa.as<double>();
  P.xs.mat_fissionable = c->mat_fis.as<int32_t>();
  P.xs.e_min = c->e_min;
  P.xs.e_max = c->e_max;
  P.src.space = c->src.space;
  P.src.angle = c->src.angle;
  P.src.energy = c->src.energy;
  P.src.only_fissionable = c->src.only_fissionable;
  for (int k = 0; k < 3; ++k) {
    P.src.point[k] = c->src.point[k];
    P.src.box_lo[k] = c->src.box_lo[k];
    P.src.box_hi[k] = c->src.box_hi[k];
    P.src.axis[k] = c->src.axis[k];
  }
  P.src.beta = c->src.beta;
  P.src.E0 = c->src.E0;
  P.src.watt_a = c->src.watt_a;
  P.src.watt_b = c->src.watt_b;
  P.src.point_elem = c->src_elem >= 0 ? m.int_of_orig[c->src_elem] : -1;

  // tallies: track-length ones first, then collision ones
  std::vector<HostTally *> order;
  for (auto &t : c->tallies)
    if (t->estimator == MCW_EST_TRACKLENGTH) order.push_back(t.get());
  int n_tl = (int)order.size();
  for (auto &t : c->tallies)
    if (t->estimator == MCW_EST_COLLISION) order.push_back(t.get());
  require((int)order.size() <= MAX_TALLIES, "too many tallies (max 8)");
  P.n_tallies = (int32_t)order.size();
  P.n_tl = n_tl;
  P.n_coll = (int32_t)order.size() - n_tl;
  for (size_t k = 0; k < order.size(); ++k) {
    HostTally &t = *order[k];
    TallyView &V = P.tally[k];
    V.estimator = t.estimator;
    V.n_scores = (int32_t)t.scores.size();
    V.n_nuclides = (int32_t)t.nuclides.size();
    V.n_score_bins = (int32_t)t.n_score_bins;
    V.simple_flux = (t.scores.size() == 1 && t.scores[0] == MCW_SCORE_FLUX && t.nuclides.size() == 1) ? 1 : 0;
    V.needs_xs = 0;
    for (int32_t sc : t.scores) V.needs_xs |= sc != MCW_SCORE_FLUX;
    V.n_edges = t.n_edges;
    V.energy_stride = t.energy_stride;
    V.elem_base = t.identity ? nullptr : t.elem_base.as<int64_t>();
    V.edges = t.n_edges ? t.edges.as<double>() : nullptr;
    V.scores = t.d_scores.as<int32_t>();
    V.nuclides = t.d_nuclides.as<int32_t>();
    V.value = t.value.as<double>();
  }
  P.tl_fast = (n_tl == 1 && P.tally[0].simple_flux && P.tally[0].n_edges == 0) ? 1 : 0;
  c->tl_bins = P.tl_fast ? order[0]->n_filter_bins * order[0]->n_score_bins : 0;
  // up to 1024 bins (8 KB) beside the 32 KB hash table keep four 256-thread
  // workgroups per CU
  P.lds_tally = 0;
  P.lds_coll = 0;
  if (P.tl_fast && order[0]->n_filter_bins * order[0]->n_score_bins <= 1024)
    P.lds_tally = (int32_t)(order[0]->n_filter_bins * order[0]->n_score_bins);
  if (P.n_coll == 1) {
    const TallyView &T = P.tally[n_tl];
    const int64_t bins = order[n_tl]->n_filter_bins * order[n_tl]->n_score_bins;
    if (T.simple_flux && T.n_edges == 0 && P.lds_tally + bins <= 1024) P.lds_coll = (int32_t)bins;
  }
  // every LDS index a kernel branch can write lies inside that branch's own
  // range: track-length bins [0, lds_tally), collision bins [lds_tally,
  // lds_tally + lds_coll)
  if (P.lds_tally) {
    const HostTally &t = *order[0];
    require(P.tl_fast && t.n_score_bins == 1 && t.max_base < P.lds_tally,
            "internal: track-length LDS tally range does not cover its bins");
  }
  if (P.lds_coll) {
    const HostTally &t = *order[n_tl];
    require(P.n_coll == 1 && P.tally[n_tl].simple_flux && t.n_score_bins == 1 && t.max_base < P.lds_coll &&
                P.lds_tally + P.lds_coll <= 1024,
            "internal: collision LDS tally range does not cover its bins");
  }
  P.trace = nullptr;
  P.trace_n = 0;
  P.trace_cap = 0;
  P.master_seed = c->opt.master_seed ? c->opt.master_seed : 1;
  P.roundtrip = c->opt.seed_roundtrip;
  P.lanes_min = c->opt.lanes_min > 0 ? std::min(c->opt.lanes_min, 64) : 40;
  P.exact_refill = m.uniform_sides == 0 ? 1 : 0; // mixed meshes: graded, long histories
  P.max_seg = c->opt.max_segments > 0 ? c->opt.max_segments : ((int64_t)1 << 20);
  P.cnt = c->counters.as<Counters>();
  HIPCHK(hipStreamSynchronize(s));
  c->dirty = false;
}

// Build a tally's device layout (OpenMC Tally::set_strides / init_results).
void build_tally(mcw_ctx *c, HostTally &t)
{
  const HostMesh &m = c->mesh;
  const size_t nf = t.ftype.size();
  t.nbins.assign(nf, 1);
  std::vector<std::vector<int64_t>> elem_bin(nf);
  for (size_t k = 0; k < nf; ++k) {
    int32_t ty = t.ftype[k];
    if (ty == MCW_FILTER_CELL || ty == MCW_FILTER_UNIVERSE) {
      std::vector<int64_t> ids = t.fids[k];
      if (ids.empty()) {
        if (ty == MCW_FILTER_CELL) {
          ids.resize(m.n_elem);
          for (int64_t e = 0; e < m.n_elem; ++e) ids[e] = e;
        } else {
          std::vector<int64_t> b(m.subdomain.begin(), m.subdomain.end());
          std::sort(b.begin(), b.end());
          b.erase(std::unique(b.begin(), b.end()), b.end());
          ids = b;
        }
      }
      t.nbins[k] = (int64_t)ids.size();
      std::map<int64_t, int64_t> pos;
      for (int64_t b = 0; b < (int64_t)ids.size(); ++b) pos.emplace(ids[b], b);
      elem_bin[k].assign(m.n_elem, -1);
      for (int64_t e = 0; e < m.n_elem; ++e) {
        auto it = pos.find(ty == MCW_FILTER_CELL ? e : (int64_t)m.subdomain[e]);
        if (it != pos.end()) elem_bin[k][e] = it->second;
      }
    } else if (ty == MCW_FILTER_ENERGY) {
      require(t.fedges[k].size() >= 2, "energy filter needs at least two edges");
      for (size_t q = 1; q < t.fedges[k].size(); ++q)
        require(t.fedges[k][q] > t.fedges[k][q - 1], "energy filter edges must increase");
      t.nbins[k] = (int64_t)t.fedges[k].size() - 1;
    } else if (ty == MCW_FILTER_PARTICLE) {
      t.nbins[k] = 1;
    } else {
      throw McwError{"Unrecognized filter"};
    }
  }
  t.strides.assign(nf, 0);
  int64_t stride = 1;
  for (int k = (int)nf - 1; k >= 0; --k) {
    t.strides[k] = stride;
    stride *= t.nbins[k];
  }
  t.n_filter_bins = stride;
  t.n_score_bins = (int64_t)t.scores.size() * (int64_t)t.nuclides.size();
  // spatial base per element
  std::vector<int64_t> base(m.n_elem, 0);
  bool any_spatial = false;
  t.n_edges = 0;
  t.energy_stride = 0;
  for (size_t k = 0; k < nf; ++k) {
    if (t.ftype[k] == MCW_FILTER_CELL || t.ftype[k] == MCW_FILTER_UNIVERSE) {
      any_spatial = true;
      for (int64_t e = 0; e < m.n_elem; ++e) {
        if (base[e] < 0) continue;
        base[e] = elem_bin[k][e] < 0 ? -1 : base[e] + elem_bin[k][e] * t.strides[k];
      }
    } else if (t.ftype[k] == MCW_FILTER_ENERGY) {
      require(t.n_edges == 0, "at most one energy filter per tally");
      t.n_edges = (int32_t)t.fedges[k].size();
      t.energy_stride = t.strides[k];
      t.edges.upload(t.fedges[k], c->stream);
    }
  }
  t.identity = any_spatial;
  if (any_spatial)
    for (int64_t e = 0; e < m.n_elem && t.identity; ++e) t.identity = base[e] == e;
  if (!any_spatial) {
    // no spatial filter: every element scores into spatial bin 0
    t.identity = false;
  }
  // device order: base of internal element i is that of caller element orig_of_int[i];
  // an identity cell tally is stored by internal id and permuted on read-out
  std::vector<int64_t> base_int(m.n_elem);
  t.max_base = -1;
  for (int64_t i = 0; i < m.n_elem; ++i) {
    base_int[i] = base[m.orig_of_int[i]];
    t.max_base = std::max(t.max_base, base_int[i]);
  }
  t.elem_base.upload(base_int, c->stream);
  t.d_scores.upload(t.scores, c->stream);
  t.d_nuclides.upload(t.nuclides, c->stream);
  const size_t n = (size_t)(t.n_filter_bins * t.n_score_bins);
  t.value.alloc(n * 8);
  t.sum.alloc(n * 8);
  t.sum_sq.alloc(n * 8);
  HIPCHK(launch_fill(c->stream, t.value.as<double>(), (int64_t)n, 0.0));
  HIPCHK(launch_fill(c->stream, t.sum.as<double>(), (int64_t)n, 0.0));
  HIPCHK(launch_fill(c->stream, t.sum_sq.as<double>(), (int64_t)n, 0.0));
}

// Three structure-of-arrays banks of the event-based path: live (in),
// live (out) and collision.  17 arrays each, see BankView.
void ensure_bank(mcw_ctx *c, int64_t n)
{
  if (c->bank_cap >= n) return;
  const size_t d = (size_t)n * 8, i = (size_t)n * 4;
  for (int b = 0; b < 3; ++b)
    for (int k = 0; k < 17; ++k) c->bank[b][k].alloc(k < 11 ? d : i);
  c->bank_cap = n;
}

BankView bank_view(mcw_ctx *c, int b)
{
  DevBuf *a = c->bank[b];
  BankView B;
  B.x = a[0].as<double>(); B.y = a[1].as<double>(); B.z = a[2].as<double>();
  B.u = a[3].as<double>(); B.v = a[4].as<double>(); B.w = a[5].as<double>();
  B.E = a[6].as<double>(); B.sigt = a[7].as<double>();
  B.seed = a[8].as<uint64_t>(); B.hid = a[9].as<int64_t>(); B.digest = a[10].as<uint64_t>();
  B.elem = a[11].as<int32_t>(); B.inface = a[12].as<int32_t>(); B.mat = a[13].as<int32_t>();
  B.k = a[14].as<int32_t>(); B.nseg = a[15].as<int32_t>(); B.ncoll = a[16].as<int32_t>();
  return B;
}

void ensure_records(mcw_ctx *c, int64_t n)
{
  if (c->rec_n >= n) return;
  size_t d = (size_t)n * 8, i = (size_t)n * 4;
  c->r_digest.alloc(d); c->r_seed.alloc(d); c->r_E.alloc(d); c->r_x.alloc(d); c->r_y.alloc(d); c->r_z.alloc(d);
  c->r_nseg.alloc(i); c->r_ncoll.alloc(i); c->r_elem.alloc(i); c->r_fate.alloc(i);
  c->rec_n = n;
}

SourceBank source_bank(mcw_ctx *c, int64_t n)
{
  if (c->sb_cap < n) {
    for (int k = 0; k < 9; ++k) c->sb[k].alloc((size_t)n * (k < 8 ? 8 : 4));
    c->sb_cap = n;
  }
  SourceBank S;
  S.x = c->sb[0].as<double>(); S.y = c->sb[1].as<double>(); S.z = c->sb[2].as<double>();
  S.u = c->sb[3].as<double>(); S.v = c->sb[4].as<double>(); S.w = c->sb[5].as<double>();
  S.E = c->sb[6].as<double>(); S.seed = c->sb[7].as<uint64_t>(); S.elem = c->sb[8].as<int32_t>();
  return S;
}

// Key layout of the (material, element) sort: element id in the low bits,
// material index above it when both fit in 32 bits (else the element alone).
void sort_layout(mcw_ctx *c)
{
  int eb = 1;
  while (eb < 31 && ((int64_t)1 << eb) <= c->mesh.n_elem) ++eb;
  int mb = 0;
  while (((int64_t)1 << mb) < (int64_t)c->n_eff_mats) ++mb;
  if (eb + mb <= 32) {
    c->sort_shift = eb;
    c->sort_bits = eb + mb;
    c->sort_mat_mask = mb > 0 ? (uint32_t)(((uint64_t)1 << mb) - 1) : 0u;
  } else {
    c->sort_shift = 0;
    c->sort_bits = eb;
    c->sort_mat_mask = 0u;
  }
}

void ensure_sort(mcw_ctx *c, int64_t n)
{
  require(n < ((int64_t)1 << 31), "sorted banks hold fewer than 2^31 particles");
  sort_layout(c);
  if (c->sort_cap >= n) return;
  for (int k = 0; k < 2; ++k) {
    c->sort_k[k].alloc((size_t)n * 4);
    c->sort_v[k].alloc((size_t)n * 4);
  }
  HIPCHK(sort_temp_bytes(n, &c->sort_tmp_bytes));
  c->sort_tmp.alloc(c->sort_tmp_bytes);
  c->sort_cap = n;
}

// Source sites of a chunk sorted by (material, element) into the second bank.
SourceBank sort_sites(mcw_ctx *c, const Params &P, const SourceBank &S, int64_t n)
{
  ensure_sort(c, n);
  if (c->sb2_cap < n) {
    for (int k = 0; k < 10; ++k) c->sb2[k].alloc((size_t)n * (k < 8 ? 8 : 4));
    c->sb2_cap = n;
  }
  SourceBank O;
  O.x = c->sb2[0].as<double>(); O.y = c->sb2[1].as<double>(); O.z = c->sb2[2].as<double>();
  O.u = c->sb2[3].as<double>(); O.v = c->sb2[4].as<double>(); O.w = c->sb2[5].as<double>();
  O.E = c->sb2[6].as<double>(); O.seed = c->sb2[7].as<uint64_t>(); O.elem = c->sb2[8].as<int32_t>();
  O.slot = c->sb2[9].as<int32_t>();
  HIPCHK(launch_sort_sites(c->stream, P.mesh, S, O, n, c->sort_shift, c->sort_bits, c->sort_mat_mask,
                           c->sort_k[0].as<uint32_t>(), c->sort_k[1].as<uint32_t>(), c->sort_v[0].as<uint32_t>(),
                           c->sort_v[1].as<uint32_t>(), c->sort_tmp.p, c->sort_tmp_bytes));
  return O;
}

// Resident workgroups of the persistent kernel over the whole device.
int persistent_blocks(mcw_ctx *c, int ns, int wpe, bool rec, bool p32, int block, bool eig = false,
                      size_t lds_bytes = 0, bool stg = false)
{
  int per_cu = 0;
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persistent_kernel(ns, wpe, rec, p32, eig, stg), block,
                                                      lds_bytes));
  if (per_cu < 1) per_cu = 1;
  return c->n_cu * per_cu;
}

// The staged tally's pool for a launch of n histories over `waves` resident
// wavefronts and nb buckets: a slice per wavefront with room for its share of
// the segments per history of the last staged batch (128 before one ran) plus
// a quarter, and a partly filled chunk per bucket; at most 2^31 records over
// all (tally_stage > 0: about that many, at least one chunk per wavefront).
// Records beyond a slice go to device atomics.  The pool only grows.
StageView stage_view(mcw_ctx *c, int64_t n, int64_t waves, int32_t nb)
{
  const double per_hist = std::max(128.0, 1.25 * c->stg_seg_per_hist);
  int64_t wc = (int64_t)std::ceil(per_hist * (double)n / (double)waves / STG_CH) + nb;
  if (c->opt.tally_stage > 0) wc = std::min<int64_t>(wc, c->opt.tally_stage / ((int64_t)STG_CH * waves));
  wc = std::max<int64_t>(1, std::min<int64_t>(wc, (((int64_t)1 << 31) / STG_CH) / waves));
  const int64_t chunks = wc * waves;
  if (chunks > c->stg_chunks) {
    c->stg_sc.alloc((size_t)chunks * STG_CH * sizeof(double));
    c->stg_el.alloc((size_t)chunks * STG_CH * sizeof(uint16_t));
    c->stg_owner.alloc((size_t)(chunks + 7) / 8 * 8);
    c->stg_fill.alloc((size_t)chunks * sizeof(uint16_t));
    c->stg_list.alloc((size_t)chunks * sizeof(uint32_t));
    c->stg_chunks = chunks;
  }
  c->stg_count.alloc(2 * STG_NB * sizeof(unsigned long long));
  StageView V{};
  V.sc = c->stg_sc.as<double>();
  V.el = c->stg_el.as<uint16_t>();
  V.owner = c->stg_owner.as<uint8_t>();
  V.fill = c->stg_fill.as<uint16_t>();
  V.list = c->stg_list.as<uint32_t>();
  V.bcount = c->stg_count.as<unsigned long long>();
  V.wave_chunks = wc;
  V.n_chunks = chunks;
  V.n_buckets = nb;
  return V;
}

// Slots of the fission bank for batches of n histories: the 3 n sites OpenMC
// keeps, plus two chunks (FB_CHUNK = 256 slots) per resident wavefront that
// partly used chunks may leave empty.
int64_t fission_slots(mcw_ctx *c, int64_t n) { return 3 * n + 2 * 256 * (int64_t)c->n_cu * 32; }

// Eigenvalue batches of n histories: the fission bank with its sort and
// resampling scratch, the per-history keff scores, and the next batch's
// source.
void ensure_fission(mcw_ctx *c, int64_t n)
{
  const int64_t cap = fission_slots(c, n);
  require(cap < ((int64_t)1 << 31), "eigenvalue batches hold at most about 714 million particles");
  if (c->fb_cap < cap) {
    for (int k = 0; k < 11; ++k) c->fb[k].alloc((size_t)cap * (k == 7 || k == 10 ? 4 : 8));
    c->fb_keys2.alloc((size_t)cap * 8);
    c->fb_perm.alloc((size_t)cap * 4);
    c->fb_perm2.alloc((size_t)cap * 4);
    HIPCHK(fission_sort_temp_bytes(cap, &c->fb_tmp_bytes));
    c->fb_tmp.alloc(c->fb_tmp_bytes);
    c->sync_cnt.alloc((size_t)cap * 8);
    c->sync_off.alloc((size_t)cap * 8);
    HIPCHK(scan_temp_bytes(cap, &c->scan_tmp_bytes));
    c->scan_tmp.alloc(c->scan_tmp_bytes);
    c->fb_cap = cap;
  }
  if (c->khist_cap < n) {
    c->k_hist.alloc((size_t)n * 3 * 8);
    c->k_red.alloc(((size_t)(n + 1023) / 1024 * 2 + 8) * 8);
    c->khist_cap = n;
  }
  if (c->gen_cap < n) {
    for (int k = 0; k < 10; ++k) c->gen[k].alloc((size_t)n * (k == 8 ? 4 : 8));
    c->gen_cap = n;
  }
}

// The fission bank of batches of n histories (3 n sites kept; the slots
// include room for partly used chunks).
FissionView fission_view(mcw_ctx *c, int64_t n)
{
  FissionView F;
  F.x = c->fb[0].as<double>(); F.y = c->fb[1].as<double>(); F.z = c->fb[2].as<double>();
  F.u = c->fb[3].as<double>(); F.v = c->fb[4].as<double>(); F.w = c->fb[5].as<double>();
  F.E = c->fb[6].as<double>();
  F.elem = c->fb[7].as<int32_t>();
  F.key = c->fb[8].as<uint64_t>();
  F.seed = c->fb[9].as<uint64_t>();
  F.nuc = c->fb[10].as<int32_t>();
  F.cap = std::min<int64_t>(fission_slots(c, n), c->fb_cap);
  F.k_hist = c->k_hist.as<double>();
  F.keff = c->keff;
  return F;
}

// The next batch's source (histories h0 ..) as a source bank.
SourceBank gen_bank(mcw_ctx *c, int64_t h0)
{
  SourceBank S;
  S.x = c->gen[0].as<double>() + h0; S.y = c->gen[1].as<double>() + h0; S.z = c->gen[2].as<double>() + h0;
  S.u = c->gen[3].as<double>() + h0; S.v = c->gen[4].as<double>() + h0; S.w = c->gen[5].as<double>() + h0;
  S.E = c->gen[6].as<double>() + h0;
  S.seed = c->gen[7].as<uint64_t>() + h0;
  S.elem = c->gen[8].as<int32_t>() + h0;
  return S;
}

// End of an eigenvalue batch (OpenMCStudy::postExecuteStudy ->
// finalize_generation, OpenMCStudy.C:374-452): reduce the keff scores, update
// keff as calculate_generation_keff / calculate_average_keff do, sort the
// fission bank and resample it into the next batch's source.
void eigen_close(mcw_ctx *c, const Params &P, int64_t batch, int64_t n, bool active)
{
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = c->stream;
  double *kr = c->k_red.as<double>();
  double *scratch = kr + 8;
  for (int k = 0; k < 3; ++k) HIPCHK(launch_tree_sum(s, P.fb.k_hist + k * n, n, scratch, kr + k));
  double ks[3];
  HIPCHK(hipMemcpyAsync(ks, kr, sizeof ks, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(c->h_counters, P.cnt, sizeof(Counters), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const int64_t requested = (int64_t)c->h_counters->fission_requested;
  const int64_t written = (int64_t)c->h_counters->fission_sites;
  const int64_t total = std::min<int64_t>(written, 3 * n);
  const FissionView F = fission_view(c, n);
  const int64_t slots = std::min<int64_t>((int64_t)c->h_counters->fission_slots, F.cap);
  mcw_keff_result &K = c->kres;
  memset(&K, 0, sizeof K);
  K.batch = batch;
  K.fission_sites = total;
  K.sites_dropped = requested - total;
  K.keff_used = c->keff;
  K.k_generation = ks[0] / (double)n;
  K.k_collision = ks[1] / (double)n;
  K.k_absorption = ks[2] / (double)n;
  if (active) {
    ++c->n_active;
    c->k_sum += K.k_generation;
    c->k_sum_sq += K.k_generation * K.k_generation;
    const double na = (double)c->n_active;
    c->keff = c->k_sum / na;
    K.keff_std = c->n_active > 1 ? std::sqrt((c->k_sum_sq / na - c->keff * c->keff) / (na - 1.0)) : 0.0;
  } else {
    c->keff = K.k_generation;
  }
  K.keff = c->keff;
  K.n_active = c->n_active;
  c->eig_batch = batch;
  c->eig_n = n;
  c->eig_sites = total;
  c->eig_src_ready = false;
  require(total > 0, "No fission sites sampled on any process, cannot sample a new batch");
  // the sites' directions and energies; OpenMC order (empty slots, key ~0,
  // sort after every site: one key bit above the largest history); past
  // 3 n sites the bank keeps the first in that order; then synchronizeBanks
  HIPCHK(launch_fission_sites(s, P.xs, F, slots));
  int hb = 0;
  while (((int64_t)1 << hb) < n) ++hb;
  HIPCHK(launch_fission_sort(s, F, slots, PROGENY_BITS + hb + 1, c->fb_keys2.as<uint64_t>(),
                             c->fb_perm.as<uint32_t>(), c->fb_perm2.as<uint32_t>(), c->fb_tmp.p, c->fb_tmp_bytes));
  const int64_t base = total < n ? n / total : 0;
  const int64_t needed = total < n ? n % total : n;
  const double p_sample = (double)needed / (double)total;
  uint64_t seed0 = mcwhost_init_seed(batch, 0, c->opt.master_seed);
  HIPCHK(launch_sync_count(s, total, base, p_sample, seed0, c->sync_cnt.as<int64_t>()));
  HIPCHK(launch_scan(s, c->sync_cnt.as<int64_t>(), c->sync_off.as<int64_t>(), total, c->scan_tmp.p,
                     c->scan_tmp_bytes));
  HIPCHK(launch_sync_scatter(s, F, c->fb_perm.as<uint32_t>(), c->sync_cnt.as<int64_t>(), c->sync_off.as<int64_t>(),
                             total, n, gen_bank(c, 0), c->gen[9].as<uint64_t>(), batch + 1, c->opt.master_seed,
                             c->opt.seed_roundtrip));
  HIPCHK(hipStreamSynchronize(s));
  c->eig_src_ready = true;
  K.close_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Event-based transport of histories hist_begin .. hist_end (sites already
// in the source bank): segment and collision kernels alternate until no
// particle is alive.  Cycles go in groups of 8, after which the live count is
// read back (kernels read the exact count from device memory and loop over
// it, so a grid only needs to be large enough to fill the device).  With
// use_graph, a group is one replay of a captured graph (fixed grids); with
// sort_cadence, the live bank is sorted by (material, element) between
// groups.
void run_plain(mcw_ctx *c, Params &P, const SourceBank &S, int64_t n)
{
  // the event kernels have no dynamic LDS: tallies go straight to the device
  P.lds_tally = 0;
  P.lds_coll = 0;
  hipStream_t s = c->stream;
  const int ns = c->mesh.uniform_sides;
  const int block = 256;
  ensure_bank(c, n);
  BankView B[3] = {bank_view(c, 0), bank_view(c, 1), bank_view(c, 2)};
  HIPCHK(launch_bank_init(s, P, S, B[0], n, 0));
  const int64_t max_blocks = (int64_t)c->n_cu * 8;
  const bool graph = c->opt.use_graph != 0;
  const int64_t cadence = c->opt.sort_cadence > 0 ? ((int64_t)c->opt.sort_cadence + 7) / 8 * 8 : 0;
  if (cadence) ensure_sort(c, n);
  hipGraphExec_t exec[2] = {nullptr, nullptr}; // 8 cycles starting from live bank 0 / 1
  struct ExecGuard {
    hipGraphExec_t *e;
    ~ExecGuard()
    {
      for (int k = 0; k < 2; ++k)
        if (e[k]) (void)hipGraphExecDestroy(e[k]);
    }
  } exec_guard{exec};
  auto group = [&](int first, int64_t blocks) {
    int cur = first;
    for (int k = 0; k < 8; ++k) {
      const int nxt = 1 - cur;
      HIPCHK(launch_reset_queue(s, P.cnt, nxt, 2));
      HIPCHK(launch_segment(ns, (unsigned)blocks, s, P, B[cur], cur, B[nxt], nxt, B[2], 2));
      HIPCHK(launch_collide((unsigned)blocks, s, P, B[2], 2, B[nxt], nxt));
      cur = nxt;
    }
  };
  int64_t bound = n;
  int cur = 0;
  for (int64_t it = 8;; it += 8) {
    if (graph) {
      if (!exec[cur]) {
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
          group(cur, max_blocks);
        } catch (...) {
          (void)hipStreamEndCapture(s, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        HIPCHK(hipStreamEndCapture(s, &g));
        const hipError_t e = hipGraphInstantiate(&exec[cur], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(e);
      }
      HIPCHK(hipGraphLaunch(exec[cur], s));
    } else {
      group(cur, std::max<int64_t>(1, std::min<int64_t>((bound + block - 1) / block, max_blocks)));
    }
    HIPCHK(hipMemcpyAsync(&c->h_counters->q_count[cur], &P.cnt->q_count[cur], 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    bound = (int64_t)c->h_counters->q_count[cur];
    if (bound == 0) break;
    if (cadence && it % cadence == 0) {
      const int nxt = 1 - cur;
      HIPCHK(launch_sort_bank(s, P.mesh, B[cur], B[nxt], bound, P.record != 0, P.cnt, nxt, c->sort_shift,
                              c->sort_bits, c->sort_mat_mask, c->sort_k[0].as<uint32_t>(),
                              c->sort_k[1].as<uint32_t>(), c->sort_v[0].as<uint32_t>(), c->sort_v[1].as<uint32_t>(),
                              c->sort_tmp.p, c->sort_tmp_bytes));
      cur = nxt;
    }
  }
}

} // namespace

// ======================================================================== C ABI
extern "C" {

void mcw_default_options(mcw_options *o)
{
  if (!o) return;
  memset(o, 0, sizeof *o);
  o->mode = MCW_MODE_BATCHED;
  o->precision = MCW_PREC_FP64;
  o->master_seed = 1;
}

int mcw_create(mcw_ctx **out, const mcw_options *opts)
{
  if (!out) return 1;
  *out = nullptr;
  mcw_ctx *c = new mcw_ctx();
  mcw_default_options(&c->opt);
  if (opts) c->opt = *opts;
  if (c->opt.master_seed == 0) c->opt.master_seed = 1;
  int rc = guard(c, [&] {
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    require(c->opt.device >= 0 && c->opt.device < ndev, "no such HIP device " + std::to_string(c->opt.device));
    require(c->opt.sort_cadence >= 0, "sort_cadence must not be negative");
    require(c->opt.use_graph == 0 || c->opt.use_graph == 1, "use_graph must be 0 or 1");
    require(c->opt.run_mode == MCW_RUN_FIXED_SOURCE || c->opt.run_mode == MCW_RUN_EIGENVALUE, "unknown run_mode");
    require(c->opt.n_inactive >= 0, "n_inactive must not be negative");
    require(c->opt.tally_stage >= -1, "tally_stage must be -1, 0 or a positive record count");
    HIPCHK(hipSetDevice(c->opt.device));
    c->scratch = scratch_of(c->opt.device);
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto &e : c->ev) HIPCHK(hipEventCreate(&e));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, c->opt.device));
    c->n_cu = prop.multiProcessorCount;
  });
  *out = c; // returned even on failure so mcw_last_error can be read
  return rc;
}

int mcw_destroy(mcw_ctx *c)
{
  if (!c) return 0;
  (void)hipSetDevice(c->opt.device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->tallies.clear();
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->h_counters) (void)hipHostFree(c->h_counters);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

const char *mcw_last_error(const mcw_ctx *c) { return c ? c->err.c_str() : "null context"; }

int mcw_options_set(mcw_ctx *c, const mcw_options *o)
{
  return guard(c, [&] {
    require(o != nullptr, "null options");
    require(o->device == c->opt.device, "the device of a context cannot change");
    require((o->master_seed ? o->master_seed : 1) == c->opt.master_seed, "master_seed of a context cannot change");
    require(o->seed_roundtrip == c->opt.seed_roundtrip, "seed_roundtrip of a context cannot change");
    require(o->mode == MCW_MODE_PLAIN || o->mode == MCW_MODE_BATCHED, "unknown mode");
    require(o->precision == MCW_PREC_FP64 || o->precision == MCW_PREC_MIXED, "unknown precision");
    require(o->lanes_min >= 0 && o->lanes_min <= 64, "lanes_min must be in 0..64");
    require(o->block_size == 0 || (o->block_size % 64 == 0 && o->block_size <= 256),
            "block_size must be 64, 128, 192 or 256");
    require(o->sort_cadence >= 0, "sort_cadence must not be negative");
    require(o->use_graph == 0 || o->use_graph == 1, "use_graph must be 0 or 1");
    require(o->trace_histories >= 0, "trace_histories must not be negative");
    require(o->run_mode == MCW_RUN_FIXED_SOURCE || o->run_mode == MCW_RUN_EIGENVALUE, "unknown run_mode");
    require(o->n_inactive >= 0, "n_inactive must not be negative");
    require(o->material_index == MCW_MATERIAL_BY_ID || o->material_index == MCW_MATERIAL_BY_INDEX,
            "unknown material_index option");
    require(o->tally_stage >= -1, "tally_stage must be -1, 0 or a positive record count");
    const bool reprep = o->max_segments != c->opt.max_segments || o->lanes_min != c->opt.lanes_min ||
                        o->material_index != c->opt.material_index;
    c->opt = *o;
    if (c->opt.master_seed == 0) c->opt.master_seed = 1;
    if (reprep) c->dirty = true;
  });
}

int mcw_mesh_set(mcw_ctx *c, int64_t n_nodes, const double *xyz, int64_t n_elem, const uint8_t *elem_type,
                 const int64_t *conn_offsets, const int32_t *conn, const int32_t *subdomain,
                 const int32_t *face_boundary)
{
  return guard(c, [&] {
    require(n_nodes > 0 && n_elem > 0 && xyz && elem_type && conn_offsets && conn && subdomain,
            "mcw_mesh_set: empty mesh or null array");
    HostMesh &m = c->mesh;
    m = HostMesh();
    m.n_nodes = n_nodes;
    m.n_elem = n_elem;
    m.xyz.assign(xyz, xyz + 3 * n_nodes);
    m.type.assign(elem_type, elem_type + n_elem);
    m.conn_off.assign(conn_offsets, conn_offsets + n_elem + 1);
    require(m.conn_off[0] == 0, "conn_offsets must start at 0");
    for (int64_t e = 0; e < n_elem; ++e) require(m.conn_off[e + 1] >= m.conn_off[e], "conn_offsets must not decrease");
    m.conn.assign(conn, conn + m.conn_off[n_elem]);
    m.subdomain.assign(subdomain, subdomain + n_elem);
    std::string e = derive_mesh(m, face_boundary);
    require(e.empty(), e);
    c->have_mesh = true;
    c->dirty = true;
    // tally layouts depend on the mesh
    for (auto &t : c->tallies) build_tally(c, *t);
  });
}

int mcw_bc_set(mcw_ctx *c, int32_t boundary_id, int32_t kind)
{
  return guard(c, [&] {
    require(boundary_id >= 0, "boundary ids must be non-negative");
    require(kind == MCW_BC_VACUUM || kind == MCW_BC_REFLECT, "unknown RayBC kind");
    c->bc[boundary_id] = kind;
    c->dirty = true;
  });
}

static HostNuclide read_nuclide(int i, const mcw_nuclide &q)
{
  require(q.n_points >= 2 && q.energy && q.total && q.absorption && q.scatter,
          "nuclide " + std::to_string(i) + ": need >= 2 points and total/absorption/scatter tables");
  require(q.awr > 0.0, "nuclide " + std::to_string(i) + ": awr must be positive");
  require(std::isfinite(q.temperature) && q.temperature >= 0.0,
          "nuclide " + std::to_string(i) + ": temperature must be finite and non-negative");
  require(std::isfinite(q.kappa) && q.kappa >= 0.0, "nuclide " + std::to_string(i) + ": kappa must be >= 0");
  HostNuclide h;
  h.awr = q.awr;
  h.kappa = q.kappa;
  if (q.temperature > 0.0) h.temperature = q.temperature;
  h.E.assign(q.energy, q.energy + q.n_points);
  require(h.E[0] > 0.0, "nuclide " + std::to_string(i) + ": energies must be positive");
  for (int j = 1; j < q.n_points; ++j)
    require(h.E[j] > h.E[j - 1], "nuclide " + std::to_string(i) + ": energy grid must increase");
  h.t.assign(q.total, q.total + q.n_points);
  h.a.assign(q.absorption, q.absorption + q.n_points);
  h.s.assign(q.scatter, q.scatter + q.n_points);
  if (q.fission) {
    h.f.assign(q.fission, q.fission + q.n_points);
    for (int k = 0; k < 2; ++k) {
      require(std::isfinite(q.nu[k]) && std::isfinite(q.watt[k]),
              "nuclide " + std::to_string(i) + ": nu and Watt parameters must be finite");
      h.nu[k] = q.nu[k];
      h.watt[k] = q.watt[k];
    }
    require(q.watt[0] > 0.0 && q.watt[1] >= 0.0,
            "nuclide " + std::to_string(i) + ": the fission spectrum needs Watt a > 0 and b >= 0");
  }
  return h;
}

int mcw_xs_set(mcw_ctx *c, int32_t n, const mcw_nuclide *nuc)
{
  return guard(c, [&] {
    require(n > 0 && nuc, "mcw_xs_set: no nuclides");
    std::vector<HostNuclide> v;
    for (int i = 0; i < n; ++i) v.push_back(read_nuclide(i, nuc[i]));
    c->nucs = std::move(v);
    c->nuc_temps.assign(c->nucs.size(), {});
    c->dirty = true;
  });
}

int mcw_xs_add_temperature(mcw_ctx *c, int32_t nuclide, const mcw_nuclide *table)
{
  return guard(c, [&] {
    require(table, "mcw_xs_add_temperature: no table");
    require(nuclide >= 0 && nuclide < (int32_t)c->nucs.size(), "mcw_xs_add_temperature: no such nuclide");
    HostNuclide h = read_nuclide(nuclide, *table);
    const HostNuclide &b = c->nucs[nuclide];
    require(h.f.empty() == b.f.empty(),
            "nuclide " + std::to_string(nuclide) + ": every temperature needs fission data, or none");
    require(h.temperature != b.temperature,
            "nuclide " + std::to_string(nuclide) + ": a table at this temperature is loaded");
    for (auto &o : c->nuc_temps[nuclide])
      require(h.temperature != o.temperature,
              "nuclide " + std::to_string(nuclide) + ": a table at this temperature is loaded");
    // one set of reaction products per nuclide
    h.awr = b.awr;
    h.kappa = b.kappa;
    for (int k = 0; k < 2; ++k) {
      h.nu[k] = b.nu[k];
      h.watt[k] = b.watt[k];
    }
    c->nuc_temps[nuclide].push_back(std::move(h));
    c->dirty = true;
  });
}

int mcw_temperature_set(mcw_ctx *c, int64_t n_elem, const double *kelvin, double tolerance)
{
  return guard(c, [&] {
    require(tolerance >= 0.0, "mcw_temperature_set: tolerance must be >= 0 (INFINITY: no limit)");
    if (n_elem == 0) {
      c->elem_T.clear(); // back to every nuclide's first table
    } else {
      require(n_elem > 0 && kelvin, "mcw_temperature_set: no temperatures");
      for (int64_t e = 0; e < n_elem; ++e)
        require(std::isfinite(kelvin[e]) && kelvin[e] >= 0.0,
                "element " + std::to_string(e) + ": temperature must be finite and non-negative");
      c->elem_T.assign(kelvin, kelvin + n_elem);
    }
    c->T_tol = tolerance;
    c->dirty = true;
  });
}

int mcw_materials_set(mcw_ctx *c, int32_t n, const mcw_material *m)
{
  return guard(c, [&] {
    require(n > 0 && m, "mcw_materials_set: no materials");
    c->mats.clear();
    for (int i = 0; i < n; ++i) {
      HostMaterial h;
      h.id = m[i].id;
      require(m[i].n_nuclides > 0, "material " + std::to_string(h.id) + " has no nuclides");
      h.nuc.assign(m[i].nuclide, m[i].nuclide + m[i].n_nuclides);
      h.dens.assign(m[i].density, m[i].density + m[i].n_nuclides);
      c->mats.push_back(std::move(h));
    }
    c->dirty = true;
  });
}

int mcw_block_material_map(mcw_ctx *c, int32_t n, const int32_t *blocks, const int32_t *materials)
{
  return guard(c, [&] {
    require(n > 0 && blocks && materials, "The blocks parameter must be the same size as the materials parameter.");
    c->block_mat.clear();
    for (int i = 0; i < n; ++i) c->block_mat[blocks[i]] = materials[i];
    c->have_block_map = true;
    c->dirty = true;
  });
}

int mcw_source_set(mcw_ctx *c, const mcw_source *s)
{
  return guard(c, [&] {
    require(s != nullptr, "null source");
    c->src = *s;
    c->have_src = true;
    c->dirty = true;
  });
}

int mcw_tally_add(mcw_ctx *c, int32_t id, int32_t estimator, int32_t n_filters, const mcw_filter *filters,
                  int32_t n_scores, const int32_t *scores, int32_t n_nuclides, const int32_t *nuclides)
{
  return guard(c, [&] {
    require(estimator == MCW_EST_TRACKLENGTH || estimator == MCW_EST_COLLISION,
            "only track-length and collision estimators are supported");
    require(n_scores > 0 && scores, "a tally needs at least one score");
    for (auto &t : c->tallies) require(t->id != id, "tally id " + std::to_string(id) + " already exists");
    auto t = std::make_unique<HostTally>();
    t->id = id;
    t->estimator = estimator;
    for (int k = 0; k < n_filters; ++k) {
      t->ftype.push_back(filters[k].type);
      std::vector<int64_t> ids;
      std::vector<double> edges;
      if (filters[k].type == MCW_FILTER_ENERGY) {
        require(filters[k].edges != nullptr, "energy filter without edges");
        edges.assign(filters[k].edges, filters[k].edges + filters[k].n_bins);
      } else if (filters[k].n_bins > 0) {
        require(filters[k].ids != nullptr, "filter bins without ids");
        ids.assign(filters[k].ids, filters[k].ids + filters[k].n_bins);
      }
      t->fids.push_back(ids);
      t->fedges.push_back(edges);
    }
    for (int k = 0; k < n_scores; ++k)
      require(scores[k] >= MCW_SCORE_FLUX && scores[k] <= MCW_SCORE_KAPPA_FISSION, "unknown score");
    t->scores.assign(scores, scores + n_scores);
    if (n_nuclides > 0) t->nuclides.assign(nuclides, nuclides + n_nuclides);
    else t->nuclides.push_back(-1);
    if (c->have_mesh) build_tally(c, *t);
    c->tallies.push_back(std::move(t));
    c->dirty = true;
  });
}

int mcw_tally_shape(mcw_ctx *c, int32_t id, int64_t *nf, int64_t *ns)
{
  return guard(c, [&] {
    for (auto &t : c->tallies)
      if (t->id == id) {
        require(c->have_mesh, "tally shape needs the mesh");
        if (nf) *nf = t->n_filter_bins;
        if (ns) *ns = t->n_score_bins;
        return;
      }
    throw McwError{"no tally with id " + std::to_string(id)};
  });
}

int mcw_tally_reset(mcw_ctx *c)
{
  return guard(c, [&] {
    for (auto &t : c->tallies) {
      const int64_t n = t->n_filter_bins * t->n_score_bins;
      HIPCHK(launch_fill(c->stream, t->value.as<double>(), n, 0.0));
      HIPCHK(launch_fill(c->stream, t->sum.as<double>(), n, 0.0));
      HIPCHK(launch_fill(c->stream, t->sum_sq.as<double>(), n, 0.0));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

// The source bank of a chunk of up to SCRATCH_CHUNK histories inside the
// device's scratch array, after its counters.
static SourceBank scratch_bank(mcw_ctx *c)
{
  unsigned char *p = c->scratch + SCRATCH_COUNTERS;
  auto take = [&](size_t elem_bytes) {
    unsigned char *q = p;
    p += ((size_t)SCRATCH_CHUNK * elem_bytes + 255) / 256 * 256;
    return q;
  };
  SourceBank S;
  S.x = (double *)take(8); S.y = (double *)take(8); S.z = (double *)take(8);
  S.u = (double *)take(8); S.v = (double *)take(8); S.w = (double *)take(8);
  S.E = (double *)take(8); S.seed = (uint64_t *)take(8); S.elem = (int32_t *)take(4);
  return S;
}

// One batch.  noalloc: the caller cannot allocate (mcw_run_batch_noalloc);
// the counters and the source bank come from the device's scratch array and
// every other buffer must exist already.
static void run_batch_impl(mcw_ctx *c, int64_t batch, int64_t n_particles, mcw_batch_stats *out, bool noalloc)
{
  {
    require(batch >= 1, "batch numbers start at 1");
    require(n_particles >= 1, "No Rays were moved to _rays in defineRays()");
    if (noalloc) {
      require(!c->dirty && c->have_src && c->h_counters,
              "mcw_run_batch_noalloc needs a context already set up: run mcw_run_batch once after "
              "changing the problem or the options (set-up allocates device memory)");
      require(c->opt.run_mode == MCW_RUN_FIXED_SOURCE, "mcw_run_batch_noalloc runs fixed-source batches only");
      require(c->opt.mode == MCW_MODE_BATCHED, "mcw_run_batch_noalloc needs mode BATCHED");
      require(c->opt.sort_cadence == 0, "mcw_run_batch_noalloc does not sort (sort_cadence must be 0)");
      require(c->opt.trace_histories == 0, "mcw_run_batch_noalloc keeps no visit records (trace_histories 0)");
      require(!c->opt.record_histories || c->rec_n >= n_particles,
              "mcw_run_batch_noalloc: history records for " + std::to_string(n_particles) +
                  " histories were never allocated; run mcw_run_batch with record_histories at this size first");
      require(c->opt.bank_capacity <= SCRATCH_CHUNK,
              "mcw_run_batch_noalloc: bank_capacity " + std::to_string(c->opt.bank_capacity) +
                  " exceeds the scratch bank of " + std::to_string(SCRATCH_CHUNK) + " histories");
    }
    prepare(c);
    const bool eig = c->opt.run_mode == MCW_RUN_EIGENVALUE;
    bool active = true;
    if (eig) {
      require(c->opt.mode == MCW_MODE_BATCHED, "eigenvalue runs need mode BATCHED");
      if (batch == 1) {
        c->keff = 1.0;
        c->k_sum = c->k_sum_sq = 0.0;
        c->n_active = 0;
        c->eig_src_ready = false;
      } else {
        require(batch == c->eig_batch + 1 && n_particles == c->eig_n && c->eig_src_ready,
                "eigenvalue batches run in order 1, 2, ... with one particle count");
      }
      require(c->keff > 0.0 && std::isfinite(c->keff), "keff must be positive and finite");
      active = batch > c->opt.n_inactive;
      ensure_fission(c, n_particles);
    }
    hipStream_t s = c->stream;
    Params P = c->base;
    P.batch = batch;
    P.n_particles = n_particles;
    if (noalloc) P.cnt = (Counters *)c->scratch; // zeroed below by a kernel, as the context's
    if (eig) {
      P.eigen = 1;
      P.fb = fission_view(c, n_particles);
      HIPCHK(launch_fill(s, P.fb.k_hist, 3 * n_particles, 0.0));
      HIPCHK(launch_fill_u64(s, (unsigned long long *)P.fb.key, P.fb.cap, ~0ull));
    if (!active) { // inactive batches score no tallies
        P.n_tl = 0;
        P.n_coll = 0;
        P.tl_fast = 0;
        P.lds_tally = 0;
        P.lds_coll = 0;
      }
    }
    P.record = c->opt.record_histories ? 1 : 0;
    if (P.record) {
      ensure_records(c, n_particles);
      P.rec.digest = c->r_digest.as<uint64_t>();
      P.rec.seed = c->r_seed.as<uint64_t>();
      P.rec.E = c->r_E.as<double>();
      P.rec.x = c->r_x.as<double>();
      P.rec.y = c->r_y.as<double>();
      P.rec.z = c->r_z.as<double>();
      P.rec.nseg = c->r_nseg.as<int32_t>();
      P.rec.ncoll = c->r_ncoll.as<int32_t>();
      P.rec.elem = c->r_elem.as<int32_t>();
      P.rec.fate = c->r_fate.as<int32_t>();
    }
    if (c->opt.trace_histories > 0) {
      require(c->opt.mode == MCW_MODE_BATCHED && P.record,
              "trace_histories needs mode BATCHED and record_histories");
      const int64_t tn = std::min<int64_t>(c->opt.trace_histories, n_particles);
      const int64_t cap = std::min<int64_t>(std::max<int64_t>(tn * 256, (int64_t)1 << 16), (int64_t)1 << 24);
      if (c->trace_cap < cap) {
        c->trace.alloc((size_t)cap * sizeof(TraceRec));
        c->trace_cap = cap;
      }
      P.trace = c->trace.as<TraceRec>();
      P.trace_n = tn;
      P.trace_cap = c->trace_cap;
    }
    for (auto &t : c->tallies) HIPCHK(launch_fill(s, t->value.as<double>(), t->n_filter_bins * t->n_score_bins, 0.0));
    HIPCHK(launch_zero_counters(s, P.cnt));
    const int64_t cap_default = noalloc ? SCRATCH_CHUNK : (int64_t)1 << 24;
    const int64_t cap = c->opt.bank_capacity > 0 ? c->opt.bank_capacity : cap_default;
    const int ns = c->mesh.uniform_sides;
    float kernel_ms = 0.0f, stage_ms = 0.0f;
    int64_t launches = 0;
    bool staged = false;
    HIPCHK(hipEventRecord(c->ev[0], s));
    for (int64_t h0 = 0; h0 < n_particles; h0 += cap) {
      const int64_t h1 = std::min(n_particles, h0 + cap);
      P.hist_begin = h0;
      P.hist_end = h1;
      SourceBank S;
      if (eig && batch > 1) {
        S = gen_bank(c, h0); // resampled fission sites of the previous batch
      } else {
        S = noalloc ? scratch_bank(c) : source_bank(c, h1 - h0);
        HIPCHK(launch_source_bank(s, P, S, h1 - h0));
      }
      if (c->opt.mode == MCW_MODE_PLAIN) {
        run_plain(c, P, S, h1 - h0);
      } else {
        if (c->opt.sort_cadence > 0) S = sort_sites(c, P, S, h1 - h0);
        const int block = c->opt.block_size > 0 ? c->opt.block_size : 256;
        const int wpe = c->opt.waves_per_eu == 5 ? 5 : 0;
        const bool p32 = c->opt.precision == MCW_PREC_MIXED && ns != 6;
        require(block % 64 == 0 && block <= 256, "block_size must be 64, 128, 192 or 256");
        // the one plain per-element flux tally staged in LDS (no device atomic per segment)
        const bool stg = c->opt.tally_stage >= 0 && !noalloc && !eig && ns == 4 && wpe == 0 && block == 256 &&
                         P.tl_fast && P.lds_tally == 0 && P.lds_coll == 0 && c->tl_bins > 1024 &&
                         c->tl_bins <= ((int64_t)STG_NB << STG_BITS);
        Params PK = P;
        if (stg) {
          PK.stg.n_buckets = (int32_t)((c->tl_bins + (1 << STG_BITS) - 1) >> STG_BITS);
          PK.xs.hash = c->uhash2.as<int32_t>();
          PK.xs.n_hash = STG_HASH;
          PK.xs.inv_dl = P.xs.inv_dl * 0.5;
        }
        int64_t blocks = std::min<int64_t>(persistent_blocks(c, ns, wpe, P.record != 0, p32, block, eig,
                                                             persistent_lds_bytes(PK, block), stg),
                                           (h1 - h0 + block - 1) / block);
        if (blocks < 1) blocks = 1;
        if (stg) {
          PK.stg = stage_view(c, h1 - h0, blocks * (block / 64), PK.stg.n_buckets);
          HIPCHK(launch_fill_u64(s, (unsigned long long *)PK.stg.owner, (PK.stg.n_chunks + 7) / 8, ~0ull));
          HIPCHK(launch_fill_u64(s, PK.stg.bcount, 2 * STG_NB, 0ull));
        }
        // the work counter restarts for every chunk
        HIPCHK(launch_reset_work(s, P.cnt));
        HIPCHK(hipEventRecord(c->ev[2], s));
        HIPCHK(launch_persistent(ns, wpe, p32, (unsigned)blocks, block, s, PK, S));
        if (stg) {
          HIPCHK(hipEventRecord(c->ev[4], s));
          // one reduce workgroup per CU (each holds a bucket's 128 KB FP64 array in LDS)
          const int split = std::max(1, c->n_cu / PK.stg.n_buckets);
          HIPCHK(launch_stage_reduce(s, PK.stg, P.tally[0].value, c->tl_bins, split));
        }
        HIPCHK(hipEventRecord(c->ev[3], s));
        HIPCHK(hipEventSynchronize(c->ev[3]));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
        kernel_ms += ms;
        if (stg) {
          HIPCHK(hipEventElapsedTime(&ms, c->ev[4], c->ev[3]));
          stage_ms += ms;
          staged = true;
        }
        ++launches;
      }
    }
    HIPCHK(hipEventRecord(c->ev[1], s));
    // batch close: VALUE / total source weight -> SUM, SUM_SQ (active batches)
    if (active)
      for (auto &t : c->tallies)
        HIPCHK(launch_accumulate(s, t->value.as<double>(), t->sum.as<double>(), t->sum_sq.as<double>(),
                                 t->n_filter_bins * t->n_score_bins, (double)n_particles));
    if (eig) eigen_close(c, P, batch, n_particles, active);
    HIPCHK(hipMemcpyAsync(c->h_counters, P.cnt, sizeof(Counters), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    float total_ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&total_ms, c->ev[0], c->ev[1]));
    c->last_batch_n = n_particles;
    c->last_batch = batch;
    if (staged) c->stg_seg_per_hist = (double)c->h_counters->segments / (double)n_particles;
    c->trace_count = P.trace_n > 0 ? std::min<int64_t>((int64_t)c->h_counters->trace_count, c->trace_cap) : 0;
    if (out) {
      const Counters &k = *c->h_counters;
      memset(out, 0, sizeof *out);
      out->histories = n_particles;
      out->segments = (int64_t)k.segments;
      out->collisions = (int64_t)k.collisions;
      out->absorptions = (int64_t)k.absorptions;
      out->leaks = (int64_t)k.leaks;
      out->cutoffs = (int64_t)k.cutoffs;
      out->lost = (int64_t)k.lost;
      out->reflections = (int64_t)k.reflections;
      out->track_length = k.track;
      out->time_ms = total_ms;
      out->kernel_ms = c->opt.mode == MCW_MODE_PLAIN ? total_ms : kernel_ms;
      out->kernel_launches = launches;
      out->fp64_rechecks = (int64_t)k.rechecks;
      out->hex_segments = (int64_t)k.hex_segments;
      out->trace_records = (int64_t)k.trace_count;
      out->stage_ms = stage_ms;
    }
  }
}

int mcw_run_batch(mcw_ctx *c, int64_t batch, int64_t n_particles, mcw_batch_stats *out)
{
  return guard(c, [&] { run_batch_impl(c, batch, n_particles, out, false); });
}

int mcw_run_batch_noalloc(mcw_ctx *c, int64_t batch, int64_t n_particles, mcw_batch_stats *out)
{
  return guard(c, [&] { run_batch_impl(c, batch, n_particles, out, true); });
}

int mcw_keff_get(mcw_ctx *c, mcw_keff_result *out)
{
  return guard(c, [&] {
    require(out != nullptr, "null output");
    require(c->eig_batch > 0, "no eigenvalue batch has run");
    *out = c->kres;
  });
}

static void copy_sites(mcw_ctx *c, const SourceBank &S, const uint64_t *key, const uint32_t *perm, int64_t n,
                       mcw_site *out, size_t max)
{
  const int64_t m = std::min<int64_t>(n, (int64_t)max);
  if (!out || m <= 0) return;
  DevBuf tmp;
  tmp.alloc((size_t)m * sizeof(mcw_site));
  HIPCHK(launch_sites_out(c->stream, S.x, S.y, S.z, S.u, S.v, S.w, S.E, S.elem, key, perm, m,
                          c->orig_of_int.as<int32_t>(), tmp.as<mcw_site>()));
  HIPCHK(hipMemcpyAsync(out, tmp.p, (size_t)m * sizeof(mcw_site), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
}

int mcw_fission_bank_get(mcw_ctx *c, mcw_site *out, size_t max, int64_t *n_out)
{
  return guard(c, [&] {
    require(c->eig_batch > 0, "no eigenvalue batch has run");
    if (n_out) *n_out = c->eig_sites;
    const FissionView F = fission_view(c, c->eig_n);
    SourceBank S;
    S.x = F.x; S.y = F.y; S.z = F.z; S.u = F.u; S.v = F.v; S.w = F.w; S.E = F.E; S.elem = F.elem;
    copy_sites(c, S, F.key, c->fb_perm.as<uint32_t>(), c->eig_sites, out, max);
  });
}

int mcw_source_bank_get(mcw_ctx *c, mcw_site *out, size_t max, int64_t *n_out)
{
  return guard(c, [&] {
    require(c->eig_src_ready, "no resampled source: run an eigenvalue batch first");
    if (n_out) *n_out = c->eig_n;
    copy_sites(c, gen_bank(c, 0), c->gen[9].as<uint64_t>(), nullptr, c->eig_n, out, max);
  });
}

int mcw_trace_get(mcw_ctx *c, mcw_trace_record *out, size_t max, int64_t *n_out)
{
  return guard(c, [&] {
    const int64_t n = c->trace_count;
    if (n_out) *n_out = n;
    if (!out || max == 0 || n == 0) return;
    std::vector<TraceRec> tmp((size_t)n);
    HIPCHK(hipMemcpyAsync(tmp.data(), c->trace.p, (size_t)n * sizeof(TraceRec), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::sort(tmp.begin(), tmp.end(), [](const TraceRec &a, const TraceRec &b) {
      return a.hid != b.hid ? a.hid < b.hid : a.visit < b.visit;
    });
    const size_t m = std::min<size_t>(max, (size_t)n);
    for (size_t i = 0; i < m; ++i) {
      const TraceRec &r = tmp[i];
      mcw_trace_record &o = out[i];
      o.history = r.hid;
      o.visit = r.visit;
      o.element = r.elem;
      o.start[0] = r.x0;
      o.start[1] = r.y0;
      o.start[2] = r.z0;
      o.end[0] = r.x1;
      o.end[1] = r.y1;
      o.end[2] = r.z1;
      o.energy = r.E;
      o.seed_tracking = r.seed;
      o.event = r.event;
      o.pad = 0;
    }
  });
}

int mcw_tally_get(mcw_ctx *c, int32_t id, int32_t kind, double *out, size_t n)
{
  return guard(c, [&] {
    for (auto &t : c->tallies)
      if (t->id == id) {
        const size_t need = (size_t)(t->n_filter_bins * t->n_score_bins);
        require(n >= need, "output buffer holds " + std::to_string(n) + " values, tally has " + std::to_string(need));
        require(kind >= MCW_RESULT_VALUE && kind <= MCW_RESULT_SUM_SQ, "unknown result kind");
        const DevBuf &b = kind == MCW_RESULT_VALUE ? t->value : kind == MCW_RESULT_SUM ? t->sum : t->sum_sq;
        if (!t->identity) {
          HIPCHK(hipMemcpyAsync(out, b.p, need * 8, hipMemcpyDeviceToHost, c->stream));
          HIPCHK(hipStreamSynchronize(c->stream));
        } else {
          std::vector<double> tmp(need);
          HIPCHK(hipMemcpyAsync(tmp.data(), b.p, need * 8, hipMemcpyDeviceToHost, c->stream));
          HIPCHK(hipStreamSynchronize(c->stream));
          const int64_t w = t->n_score_bins;
          const std::vector<int32_t> &io = c->mesh.int_of_orig;
          for (int64_t e = 0; e < (int64_t)io.size(); ++e)
            std::copy(&tmp[(size_t)io[e] * w], &tmp[(size_t)io[e] * w] + w, out + (size_t)e * w);
        }
        return;
      }
    throw McwError{"no tally with id " + std::to_string(id)};
  });
}

int mcw_history_get(mcw_ctx *c, mcw_history *out, size_t n)
{
  return guard(c, [&] {
    require(c->opt.record_histories, "histories are recorded only with options.record_histories = 1");
    const int64_t N = c->last_batch_n;
    require((int64_t)n >= N, "output buffer too small for the batch's histories");
    std::vector<uint64_t> dg(N), sd(N);
    std::vector<double> E(N), x(N), y(N), z(N);
    std::vector<int32_t> ns(N), nc(N), el(N), fa(N);
    hipStream_t s = c->stream;
    HIPCHK(hipMemcpyAsync(dg.data(), c->r_digest.p, N * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(sd.data(), c->r_seed.p, N * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(E.data(), c->r_E.p, N * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(x.data(), c->r_x.p, N * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(y.data(), c->r_y.p, N * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(z.data(), c->r_z.p, N * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(ns.data(), c->r_nseg.p, N * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(nc.data(), c->r_ncoll.p, N * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(el.data(), c->r_elem.p, N * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(fa.data(), c->r_fate.p, N * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int64_t i = 0; i < N; ++i) {
      mcw_history &h = out[i];
      h.id = (c->last_batch - 1) * N + i + 1;
      h.digest = dg[i];
      h.seed = sd[i];
      h.E = E[i];
      h.x = x[i];
      h.y = y[i];
      h.z = z[i];
      h.n_segments = ns[i];
      h.n_collisions = nc[i];
      h.elem = el[i];
      h.fate = fa[i];
    }
  });
}

/* Diagnostics: evaluate the device log, sin/cos(2 pi frac(x)), sqrt, one
 * random draw and a seed skip-ahead on n inputs (host arrays). */
int mcw_debug_math_probe(mcw_ctx *c, const double *x, int64_t n, double *log_out, double *sin_out,
                         double *cos_out, double *sqrt_out, const uint64_t *seeds, double *prn_out,
                         uint64_t *seed_out)
{
  return guard(c, [&] {
    hipStream_t s = c->stream;
    DevBuf dx, dl, ds, dc, dq, dse, dp, dso;
    size_t b = (size_t)n * 8;
    dx.alloc(b); dl.alloc(b); ds.alloc(b); dc.alloc(b); dq.alloc(b); dse.alloc(b); dp.alloc(b); dso.alloc(b);
    HIPCHK(hipMemcpyAsync(dx.p, x, b, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dse.p, seeds, b, hipMemcpyHostToDevice, s));
    HIPCHK(launch_math_probe(s, dx.as<double>(), n, dl.as<double>(), ds.as<double>(), dc.as<double>(),
                             dq.as<double>(), dse.as<uint64_t>(), dp.as<double>(), dso.as<uint64_t>()));
    HIPCHK(hipMemcpyAsync(log_out, dl.p, b, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(sin_out, ds.p, b, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(cos_out, dc.p, b, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(sqrt_out, dq.p, b, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(prn_out, dp.p, b, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(seed_out, dso.p, b, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  });
}

int mcw_debug_atomic_rate(mcw_ctx *c, int64_t n_bins, double *adds_per_s, double *sum_out)
{
  return guard(c, [&] {
    require(n_bins >= 1024 && n_bins <= ((int64_t)1 << 28) && (n_bins & (n_bins - 1)) == 0,
            "n_bins must be a power of two in 2^10 .. 2^28");
    require(adds_per_s != nullptr, "null output");
    hipStream_t s = c->stream;
    DevBuf a;
    a.alloc((size_t)n_bins * 8);
    HIPCHK(launch_fill(s, a.as<double>(), n_bins, 0.0));
    const unsigned blocks = (unsigned)c->n_cu * 16;
    const int32_t iters = 256;
    HIPCHK(launch_atomic_probe(s, a.as<double>(), (uint64_t)(n_bins - 1), blocks, 8)); // warm
    HIPCHK(hipEventRecord(c->ev[2], s));
    HIPCHK(launch_atomic_probe(s, a.as<double>(), (uint64_t)(n_bins - 1), blocks, iters));
    HIPCHK(hipEventRecord(c->ev[3], s));
    HIPCHK(hipEventSynchronize(c->ev[3]));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
    *adds_per_s = (double)blocks * 256.0 * iters / (ms * 1e-3);
    if (sum_out) {
      // every add landed: the array sums to the number of adds
      std::vector<double> h((size_t)n_bins);
      HIPCHK(hipMemcpyAsync(h.data(), a.p, (size_t)n_bins * 8, hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      double t = 0.0;
      for (double v : h) t += v;
      *sum_out = t;
    }
  });
}

int mcw_debug_gather_rate(mcw_ctx *c, int64_t n_lines, double *lines_per_s)
{
  return guard(c, [&] {
    require(n_lines >= 1024 && n_lines <= ((int64_t)1 << 28) && (n_lines & (n_lines - 1)) == 0,
            "n_lines must be a power of two in 2^10 .. 2^28");
    require(lines_per_s != nullptr, "null output");
    hipStream_t s = c->stream;
    DevBuf a, o;
    a.alloc((size_t)n_lines * 128);
    HIPCHK(launch_fill(s, a.as<double>(), n_lines * 16, 1.0));
    const unsigned blocks = (unsigned)c->n_cu * 16;
    o.alloc((size_t)blocks * 256 * 8);
    const int32_t iters = 256;
    HIPCHK(launch_gather_probe(s, a.as<double4>(), (uint64_t)(n_lines - 1), blocks, 8, o.as<double>())); // warm
    HIPCHK(hipEventRecord(c->ev[2], s));
    HIPCHK(launch_gather_probe(s, a.as<double4>(), (uint64_t)(n_lines - 1), blocks, iters, o.as<double>()));
    HIPCHK(hipEventRecord(c->ev[3], s));
    HIPCHK(hipEventSynchronize(c->ev[3]));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
    *lines_per_s = (double)blocks * 256.0 * iters / (ms * 1e-3);
  });
}

int mcw_debug_xs_probe(mcw_ctx *c, int32_t nuc, const double *E, int64_t n, int32_t *k_out, double *xs_out)
{
  return guard(c, [&] {
    prepare(c);
    require(nuc >= 0 && nuc < (int32_t)c->nucs.size(), "no such nuclide");
    hipStream_t s = c->stream;
    DevBuf dE, dk, dx;
    dE.alloc((size_t)n * 8);
    dk.alloc((size_t)n * 4);
    dx.alloc((size_t)n * 32);
    HIPCHK(hipMemcpyAsync(dE.p, E, (size_t)n * 8, hipMemcpyHostToDevice, s));
    HIPCHK(launch_xs_probe(s, c->base.xs, nuc, dE.as<double>(), n, dk.as<int32_t>(), dx.as<double>()));
    HIPCHK(hipMemcpyAsync(k_out, dk.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(xs_out, dx.p, (size_t)n * 32, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  });
}

} // extern "C"

// ======================================================================== request server
//
// Continuous batching of fixed-source requests over resident meshes.  A
// request is admitted when an in-flight slot is free and its memory (flux
// tally + its share of a round's source bank, min(N, round budget) x 76 B)
// fits the budget.  Each mesh runs rounds: one fused launch (k_source_req +
// the persistent kernel in its request-server form) over at most
// `round_budget` histories, filled from the mesh's active requests by
// shortest remaining work first (a request admitted more than `age_ms` ago
// goes first, so none starves).  Few requests per round keep few tallies
// hot: a round spread over many requests' tallies runs markedly slower.  A mesh has three round slots,
// each with its own stream, counters and source bank, so the next round's
// workgroups fill the compute units that the current round's tail (its last,
// longest histories) leaves idle; requests that arrive meanwhile join at the
// next round, and rounds of different meshes overlap as well.  Admission
// fills and result copies run on a third stream per mesh: a finished
// request's tally is gathered into caller element order and normalised on the
// device, then copied to pinned host memory; the request is done when that
// copy has landed.  Device tallies and pinned buffers are pooled per mesh, so
// the loop makes no allocation once the pools have grown to the in-flight
// peak.  A request's histories use seeds (batch 1, its particle count, master
// seed request_id + 1), so its result is the same alone or in a mix up to the
// summation order of the FP64 tally atomics.

namespace {

struct ServeReq {
  mcw_request r;
  double t_submit = 0, t_admit = -1, t_done = -1;
  int64_t scheduled = 0, done = 0, bytes = 0;
  int32_t point_elem = -1; // internal element of a point source
  int slot = -1;           // pooled tally / pinned buffer in its mesh
  hipEvent_t copied = nullptr;
  bool copying = false;
  unsigned long long segments = 0, collisions = 0;
};

// A request's device buffer: n_elem flux values (internal order), then its
// segment and collision counters.
constexpr int REQ_TAIL = 2;
constexpr int ROUND_SLOTS = 3;

struct RoundSlot {
  hipStream_t stream = nullptr;
  hipEvent_t done_ev = nullptr;
  bool busy = false;
  std::vector<std::pair<ServeReq *, int64_t>> round;
  DevBuf slices, counters;
  ReqSlice *h_slices = nullptr; // pinned, one per in-flight request
  DevBuf sb[10];
  int64_t sb_cap = 0;
  ~RoundSlot()
  {
    if (stream) (void)hipStreamSynchronize(stream);
    if (h_slices) (void)hipHostFree(h_slices);
    if (done_ev) (void)hipEventDestroy(done_ev);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct ServeMesh {
  mcw_ctx *ctx = nullptr;
  RoundSlot rs[ROUND_SLOTS];
  hipStream_t aux = nullptr; // admission fills, result gathers and copies
  hipEvent_t fill_ev = nullptr;
  bool fill_recorded = false;
  DevBuf stage;
  int blocks = 0;
  std::vector<std::unique_ptr<DevBuf>> tallies; // n_elem + REQ_TAIL doubles
  std::vector<double *> pinned;                 // n_elem + REQ_TAIL doubles
  std::vector<int> free_slots;
  ~ServeMesh()
  {
    if (aux) (void)hipStreamSynchronize(aux);
    for (double *p : pinned)
      if (p) (void)hipHostFree(p);
    if (fill_ev) (void)hipEventDestroy(fill_ev);
    if (aux) (void)hipStreamDestroy(aux);
  }
  int acquire()
  {
    if (!free_slots.empty()) {
      const int k = free_slots.back();
      free_slots.pop_back();
      return k;
    }
    const size_t n = (size_t)ctx->mesh.n_elem + REQ_TAIL;
    tallies.push_back(std::make_unique<DevBuf>());
    tallies.back()->alloc(n * 8);
    pinned.push_back(nullptr);
    HIPCHK(hipHostMalloc((void **)&pinned.back(), n * 8, hipHostMallocDefault));
    return (int)tallies.size() - 1;
  }
};

} // namespace

struct mcw_server {
  std::string err;
  std::vector<std::unique_ptr<ServeMesh>> meshes;
  std::deque<ServeReq *> pending, done;
  std::vector<ServeReq *> active;
  int32_t max_inflight = 8;
  int64_t bank_bytes = 0, used = 0;
  // rounds of 2^18 histories: a short request waits for at most a short
  // round, and more, smaller rounds overlap each other's tails (C4 mix:
  // p50 81 -> 27 ms, saturated 7.4e7 -> 7.9e7 histories/s against 2^21)
  int64_t round_budget = (int64_t)1 << 18;
  double age_ms = 250.0; // admitted longer ago than this: served first
  std::chrono::steady_clock::time_point t0;
  double now() const
  {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

int sfail(mcw_server *s, const std::string &m)
{
  if (s) s->err = m;
  return 1;
}

template <class F> int sguard(mcw_server *s, F &&f)
{
  if (!s) return 1;
  try {
    f();
    return 0;
  } catch (const McwError &e) {
    return sfail(s, e.msg);
  } catch (const std::exception &e) {
    return sfail(s, e.what());
  }
}

int64_t request_bytes(const mcw_server *s, const ServeReq *q)
{
  const mcw_ctx *c = s->meshes[q->r.mesh_slot]->ctx;
  return (c->mesh.n_elem + REQ_TAIL) * 8 + std::min(q->r.n_particles, s->round_budget) * 76;
}

void release_request(mcw_server *s, ServeReq *q)
{
  if (q->slot >= 0) s->meshes[q->r.mesh_slot]->free_slots.push_back(q->slot);
  q->slot = -1;
  if (q->copied) (void)hipEventDestroy(q->copied);
  q->copied = nullptr;
  delete q;
}

// Every round holding q's histories is complete: gather + normalise on the
// device, then copy to its pinned buffer; the copy's event marks it done.
void start_copy(ServeMesh &m, ServeReq *q)
{
  mcw_ctx *c = m.ctx;
  const int64_t ne = c->mesh.n_elem;
  HIPCHK(launch_gather_scale(m.aux, m.tallies[q->slot]->as<double>(), c->int_of_orig.as<int32_t>(), ne,
                             (double)q->r.n_particles, m.stage.as<double>(), REQ_TAIL));
  HIPCHK(hipMemcpyAsync(m.pinned[q->slot], m.stage.p, (size_t)(ne + REQ_TAIL) * 8, hipMemcpyDeviceToHost, m.aux));
  HIPCHK(hipEventRecord(q->copied, m.aux));
  q->copying = true;
}

// Launch one round in a free slot of mesh mi; false when there is no work.
bool launch_round(mcw_server *s, int mi, RoundSlot &R)
{
  ServeMesh &m = *s->meshes[mi];
  mcw_ctx *c = m.ctx;
  std::vector<ServeReq *> mine;
  for (ServeReq *q : s->active)
    if (q->r.mesh_slot == mi && q->scheduled < q->r.n_particles) mine.push_back(q);
  if (mine.empty()) return false;
  // shortest remaining work first; a request admitted more than age_ms ago
  // goes ahead of the others (oldest first), so none starves
  const double now = s->now();
  std::stable_sort(mine.begin(), mine.end(), [&](const ServeReq *a, const ServeReq *b) {
    const bool oa = now - a->t_admit > s->age_ms, ob = now - b->t_admit > s->age_ms;
    if (oa != ob) return oa;
    if (oa) return a->t_admit < b->t_admit;
    return a->r.n_particles - a->scheduled < b->r.n_particles - b->scheduled;
  });
  R.round.clear();
  int64_t total = 0;
  int32_t n_sl = 0;
  for (size_t j = 0; j < mine.size() && total < s->round_budget; ++j) {
    ServeReq *q = mine[j];
    const int64_t rem = q->r.n_particles - q->scheduled;
    // a round that holds short requests takes no long request's bulk: the
    // long one goes to the next round (the mesh's other slot), so the short
    // ones finish when their own small round does, not with the bulk
    if (total > 0 && rem > 4 * std::max(total, s->round_budget / 16)) break;
    const int64_t cnt = std::min(rem, s->round_budget - total);
    ReqSlice &r = R.h_slices[n_sl++];
    memset(&r, 0, sizeof r);
    r.first = total;
    r.count = cnt;
    r.hist0 = q->scheduled;
    r.n_particles = q->r.n_particles;
    r.master_seed = (uint64_t)q->r.request_id + 1;
    r.tally = m.tallies[q->slot]->as<double>();
    r.counts = (unsigned long long *)(r.tally + c->mesh.n_elem);
    const mcw_source &src = q->r.source;
    r.src.space = src.space;
    r.src.angle = src.angle;
    r.src.energy = src.energy;
    r.src.only_fissionable = src.only_fissionable;
    for (int k = 0; k < 3; ++k) {
      r.src.point[k] = src.point[k];
      r.src.box_lo[k] = src.box_lo[k];
      r.src.box_hi[k] = src.box_hi[k];
      r.src.axis[k] = src.axis[k];
    }
    r.src.beta = src.beta;
    r.src.E0 = src.E0;
    r.src.watt_a = src.watt_a;
    r.src.watt_b = src.watt_b;
    r.src.point_elem = q->point_elem;
    R.round.push_back({q, cnt});
    q->scheduled += cnt;
    total += cnt;
  }
  hipStream_t st = R.stream;
  if (R.sb_cap < total) {
    for (int k = 0; k < 10; ++k) R.sb[k].alloc((size_t)s->round_budget * (k < 8 ? 8 : 4) + 64);
    R.sb_cap = s->round_budget;
  }
  SourceBank S;
  S.x = R.sb[0].as<double>(); S.y = R.sb[1].as<double>(); S.z = R.sb[2].as<double>();
  S.u = R.sb[3].as<double>(); S.v = R.sb[4].as<double>(); S.w = R.sb[5].as<double>();
  S.E = R.sb[6].as<double>(); S.seed = R.sb[7].as<uint64_t>(); S.elem = R.sb[8].as<int32_t>();
  S.req = R.sb[9].as<int32_t>();
  // admission fills of the tallies this round adds to come first
  if (m.fill_recorded) HIPCHK(hipStreamWaitEvent(st, m.fill_ev, 0));
  // the slot's previous round (and so its slice upload) completed before this call
  HIPCHK(hipMemcpyAsync(R.slices.p, R.h_slices, (size_t)n_sl * sizeof(ReqSlice), hipMemcpyHostToDevice, st));
  Params P = c->base;
  P.n_tallies = P.n_tl = P.n_coll = 0;
  P.tl_fast = 0;
  P.lds_tally = 0;
  P.lds_coll = 0;
  P.record = 0;
  P.batch = 1;
  P.hist_begin = 0;
  P.hist_end = total;
  P.slices = R.slices.as<ReqSlice>();
  P.cnt = R.counters.as<Counters>();
  // the serving kernel keeps every other energy bucket in LDS (room for its tally caches)
  P.xs.hash = c->uhash2.as<int32_t>();
  P.xs.n_hash = STG_HASH;
  P.xs.inv_dl = c->base.xs.inv_dl * 0.5;
  const int ns = c->mesh.uniform_sides;
  const bool p32 = c->opt.precision == MCW_PREC_MIXED && ns != 6;
  HIPCHK(launch_reset_work(st, P.cnt));
  HIPCHK(launch_source_req(st, P, S, n_sl, total));
  const unsigned blocks = (unsigned)std::min<int64_t>(m.blocks, (total + 255) / 256);
  HIPCHK(launch_serve(ns, p32, blocks < 1 ? 1 : blocks, st, P, S));
  HIPCHK(hipEventRecord(R.done_ev, st));
  R.busy = true;
  return true;
}

} // namespace

extern "C" {

int mcw_server_create(mcw_server **out, mcw_ctx *const *meshes, int32_t n_meshes, int32_t max_inflight,
                      int64_t bank_bytes)
{
  if (!out) return 1;
  mcw_server *s = new mcw_server();
  *out = s;
  return sguard(s, [&] {
    require(n_meshes > 0 && meshes, "mcw_server_create: no meshes");
    require(max_inflight > 0, "max_inflight must be positive");
    s->max_inflight = max_inflight;
    s->bank_bytes = bank_bytes > 0 ? bank_bytes : ((int64_t)16 << 30);
    s->t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < n_meshes; ++i) s->meshes.push_back(std::make_unique<ServeMesh>());
    for (int i = 0; i < n_meshes; ++i) {
      mcw_ctx *c = meshes[i];
      require(c != nullptr, "null mesh context");
      for (int j = 0; j < i; ++j) require(meshes[j] != c, "a mesh context is listed twice");
      HIPCHK(hipSetDevice(c->opt.device));
      if (guard(c, [&] { prepare(c, false); }) != 0) throw McwError{"mesh " + std::to_string(i) + ": " + c->err};
      ServeMesh &m = *s->meshes[i];
      m.ctx = c;
      HIPCHK(hipStreamCreateWithFlags(&m.aux, hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&m.fill_ev, hipEventDisableTiming));
      m.stage.alloc((size_t)(c->mesh.n_elem + REQ_TAIL) * 8);
      for (RoundSlot &R : m.rs) {
        HIPCHK(hipStreamCreateWithFlags(&R.stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&R.done_ev, hipEventDisableTiming));
        HIPCHK(hipHostMalloc((void **)&R.h_slices, (size_t)max_inflight * sizeof(ReqSlice), hipHostMallocDefault));
        R.slices.alloc((size_t)max_inflight * sizeof(ReqSlice));
        R.counters.alloc(sizeof(Counters));
        HIPCHK(launch_zero_counters(R.stream, R.counters.as<Counters>()));
      }
      int per_cu = 0;
      const int ns = c->mesh.uniform_sides;
      HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, serve_kernel(ns, c->opt.precision == MCW_PREC_MIXED && ns != 6), 256, serve_lds_bytes()));
      m.blocks = c->n_cu * std::max(1, per_cu);
      // the pools of request tallies and pinned result buffers start at the
      // in-flight limit, so the loop allocates nothing, and an admission's
      // cost does not include a pinned allocation
      for (int k = 0; k < max_inflight; ++k) m.free_slots.push_back(m.acquire());
    }
  });
}

int mcw_server_destroy(mcw_server *s)
{
  if (!s) return 0;
  for (auto &mp : s->meshes) {
    for (RoundSlot &R : mp->rs)
      if (R.stream) (void)hipStreamSynchronize(R.stream);
    if (mp->aux) (void)hipStreamSynchronize(mp->aux);
  }
  for (auto *q : s->pending) release_request(s, q);
  for (auto *q : s->active) release_request(s, q);
  for (auto *q : s->done) release_request(s, q);
  delete s;
  return 0;
}

const char *mcw_server_last_error(const mcw_server *s) { return s ? s->err.c_str() : "null server"; }

int mcw_server_tune(mcw_server *s, int64_t round_histories, double age_ms)
{
  return sguard(s, [&] {
    require(round_histories >= 1024 && round_histories <= ((int64_t)1 << 24),
            "round_histories must be in 1024 .. 2^24");
    require(age_ms > 0.0 && std::isfinite(age_ms), "age_ms must be positive and finite");
    for (auto &mp : s->meshes)
      for (RoundSlot &R : mp->rs) require(!R.busy, "mcw_server_tune: rounds are running; tune before submitting");
    require(s->active.empty() && s->pending.empty(), "mcw_server_tune: requests are queued; tune before submitting");
    s->round_budget = round_histories;
    s->age_ms = age_ms;
  });
}

int mcw_submit(mcw_server *s, const mcw_request *req)
{
  return sguard(s, [&] {
    require(req != nullptr, "null request");
    require(req->mesh_slot >= 0 && req->mesh_slot < (int32_t)s->meshes.size(), "request names an unknown mesh slot");
    require(req->n_particles >= 1, "a request needs at least one particle");
    const int sp = req->source.space;
    require(sp == MCW_SPACE_POINT || sp == MCW_SPACE_BOX || sp == MCW_SPACE_MESH_VOLUME, "unknown source space type");
    check_angle(req->source);
    check_source_values(req->source);
    require(req->source.energy == MCW_ENERGY_MONO || req->source.energy == MCW_ENERGY_WATT,
            "source energy type not supported");
    auto *q = new ServeReq();
    q->r = *req;
    q->t_submit = s->now();
    if (sp == MCW_SPACE_POINT) {
      const mcw_ctx *c = s->meshes[req->mesh_slot]->ctx;
      const int64_t e = locate_point(c->mesh, req->source.point, 1e-12);
      q->point_elem = e >= 0 ? c->mesh.int_of_orig[e] : -1;
    }
    s->pending.push_back(q);
  });
}

int mcw_poll(mcw_server *s, int32_t *n_done)
{
  return sguard(s, [&] {
    // finished rounds: start the result copy of every request whose last
    // histories were in the round
    for (auto &mp : s->meshes) {
      ServeMesh &m = *mp;
      for (RoundSlot &R : m.rs) {
        if (!R.busy) continue;
        hipError_t e = hipEventQuery(R.done_ev);
        if (e == hipErrorNotReady) continue;
        HIPCHK(e);
        R.busy = false;
        for (auto &rc : R.round) {
          rc.first->done += rc.second;
          if (rc.first->done == rc.first->r.n_particles) start_copy(m, rc.first);
        }
        R.round.clear();
      }
    }
    // landed copies
    for (size_t i = 0; i < s->active.size();) {
      ServeReq *q = s->active[i];
      if (q->copying) {
        hipError_t e = hipEventQuery(q->copied);
        if (e != hipErrorNotReady) {
          HIPCHK(e);
          const int64_t ne = s->meshes[q->r.mesh_slot]->ctx->mesh.n_elem;
          const double *tail = s->meshes[q->r.mesh_slot]->pinned[q->slot] + ne;
          memcpy(&q->segments, tail, 8);
          memcpy(&q->collisions, tail + 1, 8);
          q->t_done = s->now();
          s->used -= q->bytes;
          s->active.erase(s->active.begin() + (ptrdiff_t)i);
          s->done.push_back(q);
          continue;
        }
      }
      ++i;
    }
    // admissions
    std::vector<bool> filled(s->meshes.size(), false);
    while (!s->pending.empty() && (int32_t)s->active.size() < s->max_inflight) {
      ServeReq *q = s->pending.front();
      const int64_t need = request_bytes(s, q);
      if (s->used + need > s->bank_bytes && !s->active.empty()) break;
      ServeMesh &m = *s->meshes[q->r.mesh_slot];
      mcw_ctx *c = m.ctx;
      HIPCHK(hipEventCreateWithFlags(&q->copied, hipEventDisableTiming));
      q->slot = m.acquire();
      HIPCHK(launch_fill(m.aux, m.tallies[q->slot]->as<double>(), c->mesh.n_elem + REQ_TAIL, 0.0));
      filled[q->r.mesh_slot] = true;
      q->bytes = need;
      s->used += need;
      q->t_admit = s->now();
      s->pending.pop_front();
      s->active.push_back(q);
    }
    for (size_t i = 0; i < s->meshes.size(); ++i)
      if (filled[i]) {
        HIPCHK(hipEventRecord(s->meshes[i]->fill_ev, s->meshes[i]->aux));
        s->meshes[i]->fill_recorded = true;
      }
    // next rounds into free slots
    for (int i = 0; i < (int)s->meshes.size(); ++i)
      for (RoundSlot &R : s->meshes[i]->rs)
        if (!R.busy && !launch_round(s, i, R)) break;
    if (n_done) *n_done = (int32_t)s->done.size();
  });
}

int mcw_result(mcw_server *s, mcw_request_status *status, double *flux, size_t n)
{
  return sguard(s, [&] {
    require(!s->done.empty(), "no finished request");
    ServeReq *q = s->done.front();
    const ServeMesh &m = *s->meshes[q->r.mesh_slot];
    const size_t ne = (size_t)m.ctx->mesh.n_elem;
    require(flux == nullptr || n >= ne, "flux buffer too small for the request's mesh");
    if (status) {
      status->request_id = q->r.request_id;
      status->histories = q->r.n_particles;
      status->segments = (int64_t)q->segments;
      status->collisions = (int64_t)q->collisions;
      status->t_submit_ms = q->t_submit;
      status->t_admit_ms = q->t_admit;
      status->t_done_ms = q->t_done;
    }
    if (flux) memcpy(flux, m.pinned[q->slot], ne * 8);
    s->done.pop_front();
    release_request(s, q);
  });
}

} // extern "C"
