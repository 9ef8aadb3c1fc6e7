// mcw_types.h -- device views shared by the host runtime and the kernels.
//
// HBM layout (all arrays device-resident, built once per problem):
//   mesh   plane[slot]   double4 (nx, ny, nz, c), outward unit normal; a
//                        TET4 element's 4 planes are one 128-byte record
//          nbr[slot]     int32: neighbour*8 + neighbour side, split*8 + 7 (mixed
//                        meshes: a HEX8 side covered by two TET4 sides), or
//                        -1 - boundary
//          elem_mat[e]   int32 material index (-1 void)
//   xs     energy[]      per-nuclide grids, packed (grid_off[n] .. +n_points[n])
//          rec[]         double4 (E, total, absorption, fission) per point; scatter apart
//   bank   SoA particle state (x,y,z,u,v,w,E,sigt f64; seed u64; elem, inface,
//          mat i32) sized for one chunk of histories
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcw {

constexpr int MAX_TALLIES = 8;

// Mixed-precision TET4 record: one 128-byte line per element.  Planes are
// FP32, with offsets relative to the element's first vertex v0 (kept in
// FP64), so FP32 distances carry errors of order 2^-24 of the element size
// rather than of the domain size.
struct __align__(128) MixedTet {
  float4 pl[4];      // nx, ny, nz, n.(p_side - v0)
  double v0[3];
  int32_t code[4];   // neighbour codes, as in erec
  int32_t mat;
  int32_t pad[5];
};

// Mixed-precision HEX8 record of a mixed mesh: one 128-byte line per element.
// The six planes are FP32 with absolute offsets (there is no room for a
// reference vertex), so FP32 distances carry errors of order 2^-24 of the
// domain size; the FP64 re-solve of near ties keeps the exit side exact.
struct __align__(128) MixedHex {
  float4 pl[6];      // nx, ny, nz, c
  int32_t code[6];   // neighbour codes, as in erec
  int32_t mat;
  int32_t pad;
};

// Split interface of a mixed mesh: a HEX8 side covered by two TET4 sides
// that cut it along a diagonal.  An exit point x through the hex side enters
// code[0] when pl.xyz . x - pl.w <= 0, else code[1] (internal id * 8 + side).
struct SplitRec {
  double4 pl;
  int32_t code[2];
  int32_t pad[2];
};

// Elements live on the device with the TET4 first and the HEX8 after, each
// group in Morton order of their centroids ("internal" ids); the caller's ids appear only in source sampling (the
// volume CDF is in caller order), records and tally read-out.
struct MeshView {
  const double4 *plane;      // per side slot (internal order): outward unit normal, offset;
                             // element e's sides start at slot_of(M, e): 4e (TET4 meshes),
                             // 6e (HEX8 meshes); mixed meshes 4e for the TET4 (e < n_tet),
                             // 4 n_tet + 8 (e - n_tet) for the HEX8 (cache-line aligned; the
                             // HEX8 block's slot 6 holds its erec record, so a segment in a
                             // hex reads two lines, not three)
  const int4 *erec;          // per element, 2 x int4: neighbour code of sides 0..5
                             // (internal id * 8 + side, or -1 - boundary id), material, side count
  int32_t stride;            // sides per element for single-type meshes (4 or 6), 0 = mixed
  int32_t n_bc;
  const int8_t *bc_kind;     // per boundary id
  int64_t n_elem;
  // caller order, for volumetric sources
  const double *cdf;
  const int32_t *cdf_guide;   // cdf_guide[j] = first e with cdf[e] > j * cdf[n_elem-1] / n_guide
  int64_t n_guide;
  double cdf_guide_scale;     // n_guide / cdf[n_elem-1]
  const double *xyz;
  const int32_t *conn;
  const int64_t *conn_off;
  int64_t conn_off0;          // conn_off[0]; all-TET4 meshes (stride 4) hold element e at conn_off0 + 4e
  const uint8_t *type;
  const int32_t *int_of_orig; // caller element -> internal id
  const int32_t *orig_of_int; // internal id -> caller element
  const MixedTet *mtet;       // mixed-precision records of the TET4 (internal ids < n_tet)
  const MixedHex *mhex;       // mixed meshes: mixed-precision records of the HEX8 (id - n_tet)
  int64_t n_tet;              // internal ids below n_tet are TET4, the rest HEX8
  const SplitRec *split;      // split interfaces of mixed meshes (neighbour code split * 8 + 7)
  // point location: uniform grid over the mesh bounding box, CSR lists of
  // the (internal) elements whose bounding box meets each cell
  double grid_lo[3], grid_inv[3];
  int32_t grid_n[3];
  int32_t pad_grid;
  const int32_t *cell_start;
  const int32_t *cell_elem;
};

struct XSView {
  const int64_t *grid_off;   // per nuclide: first point in rec
  const int32_t *n_points;
  const double4 *rec;        // per point: (E, total, absorption, fission), fission 0 when absent
  const double *scatter;     // per point (only scatter scores read it)
  const double4 *fis;        // per nuclide: (nu0, nu1, Watt a, Watt b), nu(E) = nu0 + nu1 E
  const int32_t *fissionable; // per nuclide: 1 when it has fission data
  const double *watt_t;      // per nuclide: r1 r2 above this means the first Watt attempt is
                             // below e_max (-a ln(r1 r2) < the w that reaches e_max)
  const double *awr;
  const int32_t *nuc_of;     // per nuclide instance (one temperature's tables): the caller's nuclide
  const double *kappa;       // per nuclide instance: energy released per fission [eV]
  const int32_t *mat_fissionable; // per device material: 1 when a nuclide of it is fissionable
  // unionized grid: every nuclide's points merged, sorted, unique
  const double *U;
  const int32_t *uidx;       // n_union x n_nuc: largest j with E_n[j] <= U[k], -1 if none
  const int32_t *hash;       // n_hash + 1: largest k with U[k] <= lower edge of log bucket b
  int32_t n_union;
  int32_t n_nuc;             // nuclide instances: the caller's nuclides first, then other temperatures
  int32_t n_hash;
  int32_t n_mat;
  double logU0, inv_dl;      // bucket b = floor((log E - logU0) * inv_dl)
  const int32_t *mat_off;
  const int32_t *mat_nuc;
  const double *mat_dens;
  double e_min, e_max;
};

constexpr int MAX_HASH = 8190; // LDS table of MAX_HASH + 1 ints: five 256-thread workgroups fit in 160 KB

struct SourceView {
  int32_t space, angle, energy, only_fissionable;
  double point[3];
  double box_lo[3], box_hi[3];
  double axis[3];
  double beta, E0, watt_a, watt_b;
  int64_t point_elem;
};

struct TallyView {
  int32_t estimator;
  int32_t n_scores;
  int32_t n_nuclides;
  int32_t n_score_bins;
  int32_t simple_flux;       // one flux score, one nuclide bin
  int32_t needs_xs;
  int32_t n_edges;           // energy filter edges (0 = none)
  int32_t pad;
  int64_t energy_stride;
  const int64_t *elem_base;  // spatial filter index per element (-1 = no bin), nullptr = element id
  const double *edges;
  const int32_t *scores;
  const int32_t *nuclides;
  double *value;
};

// Structure-of-arrays particle bank of the event-based path.
struct BankView {
  double *x, *y, *z, *u, *v, *w, *E, *sigt;
  uint64_t *seed;
  int64_t *hid;      // history index within the batch (0-based)
  int32_t *elem, *inface, *mat, *k, *nseg;
  uint64_t *digest;  // record mode only
  int32_t *ncoll;    // record mode only
};

// Source sites sampled ahead of transport (one per history of a chunk).
struct SourceBank {
  double *x = nullptr, *y = nullptr, *z = nullptr, *u = nullptr, *v = nullptr, *w = nullptr, *E = nullptr;
  uint64_t *seed = nullptr; // tracking-stream seed
  int32_t *elem = nullptr;
  int32_t *req = nullptr;   // request-server rounds: slice index of each site, else nullptr
  int32_t *slot = nullptr;  // sorted banks: the site's slot before sorting, else nullptr
};

// One request's slice of a request-server round: histories hist0 ..
// hist0 + count - 1 of the request occupy source-bank slots first .. first +
// count - 1.  Seeds follow OpenMC's formula with batch 1, the request's
// particle count and a master seed derived from the request id, so a result
// does not depend on what else runs beside it.
struct ReqSlice {
  int64_t first;
  int64_t count;
  int64_t hist0;
  int64_t n_particles;
  uint64_t master_seed;
  double *tally;                   // per-element track-length flux (internal order)
  unsigned long long *counts;      // per-request segments, collisions
  SourceView src;
};

// One element visit of a traced history (record mode of the persistent
// kernel): from entering the element (or the start, or a reflection) to
// leaving it (crossing, reflection, leak, loss) or dying in it (absorption,
// energy cutoff).  Energy and tracking seed are the values at the visit's end.
struct TraceRec {
  double x0, y0, z0, x1, y1, z1;
  double E;
  uint64_t seed;
  int64_t hid;    // history index within the batch
  int32_t elem;   // caller element id
  int32_t visit;  // visit index within the history, from 0
  int32_t event;  // event that ended the visit (EV_*)
  int32_t pad;
};

struct RecordView {
  uint64_t *digest, *seed;
  double *E, *x, *y, *z;
  int32_t *nseg, *ncoll, *elem, *fate;
};

// Fission bank of an eigenvalue batch (structure of arrays, unsorted: slots
// are taken in arrival order; `key` gives OpenMC's order) and the
// per-history keff scores, reduced in a fixed order after the batch.
struct FissionView {
  double *x, *y, *z, *u, *v, *w, *E;
  int32_t *elem;           // internal element id
  uint64_t *key;           // history << PROGENY_BITS | progeny
  uint64_t *seed;          // tracking-stream state before the site's first draw
  int32_t *nuc;            // nuclide that fissioned
  int64_t cap;             // slots of the bank: 3 n_particles plus room for partly used chunks
  double *k_hist;          // 3 x n_particles: track-length, collision, absorption scores
  double keff;             // divisor of the fission yield
};

constexpr int PROGENY_BITS = 24;

// Staged per-element flux tally.  Instead of one device atomic per segment
// (every FP64 atomic is performed memory-side, at a fixed rate), a wavefront
// of the persistent kernel stages its track-length scores in LDS by bucket of
// 2^STG_BITS consecutive tally bins and writes a full bucket of STG_S records
// with plain stores to a chunk of its own slice of a device pool.  After the
// launch the chunks are listed by bucket (k_stage_count, k_stage_scatter) and
// k_stage_reduce sums each bucket's records in an LDS array of the bucket's
// bins and adds the sums to the tally.
constexpr int STG_BITS = 14;           // bins per bucket: 2^14 (the reduce's 128 KB FP64 LDS array)
constexpr int STG_NB = 64;             // buckets: tallies of up to 2^20 bins
constexpr int STG_S = 8;               // records per bucket staged in LDS
constexpr int STG_CH = 256;            // records per pool chunk (a multiple of STG_S)
constexpr int STG_HASH = MAX_HASH / 2; // energy-bucket LDS table of the staging kernels (every other bucket)
constexpr uint8_t STG_FREE = 0xff;     // owner of a chunk no wavefront took

struct StageView {
  double *sc;                  // per record: the score
  uint16_t *el;                // per record: the bin within its bucket
  uint8_t *owner;              // per chunk: its bucket, STG_FREE before a wavefront takes it
  uint16_t *fill;              // per chunk: records written
  uint32_t *list;              // chunks listed by bucket (k_stage_scatter)
  unsigned long long *bcount;  // per bucket: its chunks; then per bucket: list cursor
  int64_t wave_chunks;         // chunks of each wavefront's slice
  int64_t n_chunks;            // chunks of the pool: wavefronts x wave_chunks
  int32_t n_buckets;           // 0: staging off
  int32_t pad;
};

// Global counters, accumulated with one atomic per wavefront.
struct Counters {
  unsigned long long segments, collisions, absorptions, leaks, cutoffs, lost, reflections, rechecks;
  unsigned long long hex_segments; // segments in HEX8 elements
  unsigned long long trace_count;  // trace records emitted (may exceed the buffer)
  unsigned long long fission_sites;     // sites written to the bank
  unsigned long long fission_requested; // sites the collisions asked for
  unsigned long long fission_slots;     // bank slots handed out in chunks (some stay empty)
  double track;
  unsigned long long next_hist;   // persistent kernel work queue
  unsigned long long q_count[4];  // event queue lengths
};

struct Params {
  MeshView mesh;
  XSView xs;
  SourceView src;
  TallyView tally[MAX_TALLIES];
  int32_t n_tallies;
  int32_t n_tl;            // tallies with the track-length estimator
  int32_t n_coll;          // tallies with the collision estimator
  int32_t roundtrip;
  int32_t record;
  int32_t lanes_min;       // persistent kernel: keep stepping segments while this many lanes are in flight
  int32_t tl_fast;         // the one track-length tally is a plain flux per element bin (no energy filter)
  int32_t eigen;           // eigenvalue batch: keff scores and fission sites (fb)
  int32_t exact_refill;    // persistent kernel: refills take exactly the idle lanes' count (no chunks)
  int32_t lds_tally;       // persistent kernel: bins of the tl_fast tally summed in LDS first (0: off)
  int32_t lds_coll;        // and of the one collision tally, a plain flux, after them (0: off)
  int32_t pad_lds;
  int64_t batch;
  int64_t n_particles;     // particles in the batch (seed id stride)
  int64_t hist_begin;      // first history index handled by this launch
  int64_t hist_end;        // one past the last
  int64_t max_seg;
  uint64_t master_seed;
  Counters *cnt;
  RecordView rec;
  const ReqSlice *slices;  // request-server rounds, else nullptr
  TraceRec *trace;         // record mode: visit records of histories hid < trace_n
  int64_t trace_n;
  int64_t trace_cap;
  FissionView fb;
  StageView stg;           // staged flux tally (persistent kernel, STG variants)
};

} // namespace mcw
