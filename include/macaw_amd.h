/*
 * macaw_amd.h -- C ABI of the MI355X particle-tracking library.
 *
 * The library replaces the hot path of a MaCaw transport study: the MOOSE
 * TraceRay element walk, MaCaw's CollisionKernel::onSegment and the OpenMC
 * routines it calls (cross sections, random numbers, collisions, tallies).
 * Every entry point is `extern "C"`, takes plain pointers and sizes, returns
 * 0 on success and non-zero on error, and leaves a message readable with
 * mcw_last_error() -- the contract of OpenMC's C API
 * (`int openmc_init(...)` + `openmc_err_msg`, used at
 * src/userobjects/OpenMCStudy.C:127-131,213-215).
 *
 * Host arrays passed in are caller-owned and copied; the context owns every
 * device allocation.  One context per host thread; calls on one context are
 * serialised by the caller.
 *
 * Conventions that are part of the interface:
 *   - element node order and side numbering are libMesh's (Tet4: sides
 *     {0,2,1},{0,1,3},{1,2,3},{2,0,3}; Hex8: sides {0,3,2,1},{0,1,5,4},
 *     {1,2,6,5},{2,3,7,6},{3,0,4,7},{4,5,6,7}); exit-face ties go to the
 *     lowest side index;
 *   - blocks map to material *ids* (CollisionKernel `blocks`/`materials`,
 *     src/raykernels/CollisionKernel.C:63-69), and every mapped id must exist
 *     (the check at CollisionKernel.C:85-90).  By default the material used
 *     is the one with that id.  MaCaw itself uses material *index* id - 1
 *     into OpenMC's materials list (CollisionKernel.C:110), which is the
 *     same material only when ids are 1..N in list order;
 *     options.material_index = MCW_MATERIAL_BY_INDEX reproduces that;
 *   - random numbers follow OpenMC 0.13: 64-bit LCG with PCG-RXS-M-XS output,
 *     per-particle streams (0 tracking, 1 source, 2 URR), particle seed
 *     (batch-1)*n_particles + id, stride 152917;
 *   - tally results are laid out as OpenMC's results_(filter, score, kind)
 *     with kind VALUE=0, SUM=1, SUM_SQ=2 (src/auxkernels/OpenMCTallyAux.C:110-203).
 */
#ifndef MACAW_AMD_H
#define MACAW_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enums --------------------------------------------------------------- */

enum { MCW_TET4 = 4, MCW_HEX8 = 8 };

/* RayBC kinds: KillRayBC (leakage) and ReflectRayBC (specular). */
enum { MCW_BC_VACUUM = 0, MCW_BC_REFLECT = 1 };

/* Estimators, numbered as OpenMC's TallyEstimator / OpenMCTally's MooseEnum
 * "ANALOG TRACKLENGTH COLLISION" (src/userobjects/OpenMCTally.C:39). */
enum { MCW_EST_ANALOG = 0, MCW_EST_TRACKLENGTH = 1, MCW_EST_COLLISION = 2 };

/* Tally result kinds (OpenMC TallyResult). */
enum { MCW_RESULT_VALUE = 0, MCW_RESULT_SUM = 1, MCW_RESULT_SUM_SQ = 2 };

/* Tally filters (OpenMCTally.C:125-191) and scores. */
enum { MCW_FILTER_UNIVERSE = 1, MCW_FILTER_CELL = 2, MCW_FILTER_ENERGY = 3, MCW_FILTER_PARTICLE = 4 };
enum { MCW_SCORE_FLUX = 0, MCW_SCORE_TOTAL = 1, MCW_SCORE_ABSORPTION = 2, MCW_SCORE_SCATTER = 3,
       MCW_SCORE_FISSION = 4, MCW_SCORE_KAPPA_FISSION = 5 /* kappa sigma_f: fission energy [eV] */ };

/* Source distributions. */
enum { MCW_SPACE_POINT = 0, MCW_SPACE_BOX = 1, MCW_SPACE_MESH_VOLUME = 2 };
enum { MCW_ANGLE_ISOTROPIC = 0, MCW_ANGLE_LINEAR_MU = 1 };
enum { MCW_ENERGY_MONO = 0, MCW_ENERGY_WATT = 1 };

/* Transport algorithms.  PLAIN: event kernels (xs, segment, collide,
 * compact) driven from the host.  BATCHED: one persistent fused kernel per
 * bank chunk.  Both give bit-identical histories. */
enum { MCW_MODE_PLAIN = 0, MCW_MODE_BATCHED = 1 };

/* Exit-side arithmetic.  FP64 everywhere, or (BATCHED, TET4 meshes) FP32
 * candidate distances from one 128-byte record per element with an FP64
 * re-solve of near ties: the exit side then always equals FP64's choice,
 * segment lengths carry FP32 rounding; positions, cross sections and
 * tallies stay FP64.  Other meshes and the PLAIN mode run FP64. */
enum { MCW_PREC_FP64 = 0, MCW_PREC_MIXED = 1 };

/* Run modes (settings.xml <run_mode>).  EIGENVALUE: batch 1 starts from the
 * external source, each later batch from the fission sites of the one before
 * it, resampled to n_particles as OpenMCStudy::synchronizeBanks does
 * (src/userobjects/OpenMCStudy.C:454-607); keff from the track-length
 * estimator; tallies score in active batches only. */
enum { MCW_RUN_FIXED_SOURCE = 0, MCW_RUN_EIGENVALUE = 1 };

/* How a block's material id selects a material (mcw_options.material_index).
 * BY_ID: the material whose id matches.  BY_INDEX: MaCaw's
 * `_block_to_openmc_materials.at(subdomain) - 1` (CollisionKernel.C:110),
 * the (id-1)-th material of mcw_materials_set's list; an id outside
 * 1..n_materials is an error at the next batch. */
enum { MCW_MATERIAL_BY_ID = 0, MCW_MATERIAL_BY_INDEX = 1 };

/* Per-history fates. */
enum { MCW_FATE_ALIVE = 0, MCW_FATE_LEAKED = 1, MCW_FATE_ABSORBED = 2, MCW_FATE_CUTOFF = 3,
       MCW_FATE_LOST = 4 };

/* ---- plain-data inputs ---------------------------------------------------- */

typedef struct mcw_ctx mcw_ctx;

typedef struct {
  int32_t device;           /* HIP device ordinal */
  int32_t mode;             /* MCW_MODE_* */
  int32_t precision;        /* MCW_PREC_* */
  int32_t seed_roundtrip;   /* 1: round the tracking seed through a double at the end of
                               every segment, as MaCaw's Real ray aux data does
                               (CollisionKernel.C:132-134,157-159,217-219) */
  int32_t record_histories; /* 1: keep per-history records (mcw_history_get) */
  int32_t block_size;       /* threads per workgroup, 0 = default */
  int64_t bank_capacity;    /* particles resident per chunk, 0 = default */
  int64_t max_segments;     /* per-history guard; beyond it the history is lost; 0 = default */
  uint64_t master_seed;     /* OpenMC settings::seed, default 1 */
  int32_t sort_cadence;     /* order particles by (material, element): BATCHED, > 0 sorts each
                               chunk's source sites before transport; PLAIN, event cycles
                               between sorts of the live bank (rounded up to a multiple of
                               8); 0 = never */
  int32_t lanes_min;        /* BATCHED: a wavefront keeps stepping segments while at least this
                               many of its 64 lanes are in flight (1..64, 0 = default) */
  int32_t waves_per_eu;     /* BATCHED: 5 = kernel variant built for 5 wavefronts per SIMD
                               (fewer registers, some spills), 0 = compiler's choice */
  int32_t use_graph;        /* PLAIN: 1 = replay the event cycle (8 segment + collision
                               cycles) as one captured HIP graph */
  int32_t trace_histories;  /* BATCHED with record_histories: keep one record per element
                               visit of the first trace_histories histories of a batch
                               (mcw_trace_get), as RayTracingExodus writes them */
  int32_t run_mode;         /* MCW_RUN_* (EIGENVALUE needs mode BATCHED) */
  int32_t n_inactive;       /* EIGENVALUE: batches 1..n_inactive are inactive (settings
                               <inactive>): no tally scores, keff not averaged */
  int32_t material_index;   /* MCW_MATERIAL_* (default BY_ID) */
  int32_t tally_stage;      /* BATCHED fixed-source runs on TET4 meshes (256-thread workgroups,
                               waves_per_eu 0; not mcw_run_batch_noalloc) whose only
                               track-length tally is a plain flux of more than 1024 and at most
                               2^20 bins:
                               0 = stage its scores in LDS and sum them in a second kernel
                               (no device atomic per segment), -1 = one device atomic per
                               segment, n > 0 = stage into a pool of at most n records per
                               launch (the rest go to device atomics) */
} mcw_options;

/* One synthetic continuous-energy nuclide: pointwise lin-lin tables on an
 * increasing energy grid [eV]; microscopic cross sections in barns; elastic
 * scattering isotropic in the centre-of-mass frame off a target at rest of
 * mass ratio awr.  A nuclide with a fission table is fissionable: fission is
 * part of absorption, and in EIGENVALUE runs a collision with it banks
 * fission sites (nu(E) = nu[0] + nu[1] E neutrons per fission, isotropic,
 * energies from a Watt spectrum with a = watt[0] eV, b = watt[1] 1/eV). */
typedef struct {
  int32_t n_points;
  double awr;
  const double *energy;
  const double *total;
  const double *absorption;
  const double *scatter;
  const double *fission; /* may be NULL (not fissionable) */
  double nu[2];
  double watt[2];
  double temperature;    /* K, the temperature of this table (0: 293.6) */
  double kappa;          /* energy released per fission [eV], for kappa-fission scores */
} mcw_nuclide;

/* A material: id (as in materials.xml) and atom densities [1/(b cm)]. */
typedef struct {
  int32_t id;
  int32_t n_nuclides;
  const int32_t *nuclide; /* indices into the nuclide list */
  const double *density;
} mcw_material;

typedef struct {
  int32_t space;  /* MCW_SPACE_* */
  int32_t angle;  /* MCW_ANGLE_* */
  int32_t energy; /* MCW_ENERGY_* */
  int32_t only_fissionable; /* BOX: keep only sites in fissionable materials
                               (OpenMC <space type="fission">); a zero-width
                               box dimension is allowed */
  double point[3];
  double box_lo[3];
  double box_hi[3];
  double axis[3]; /* LINEAR_MU reference direction (unit) */
  double beta;    /* LINEAR_MU: p(mu) ~ 1 + beta*mu, |beta| <= 1 */
  double E0;      /* MONO energy [eV] */
  double watt_a;  /* WATT a [eV] */
  double watt_b;  /* WATT b [1/eV] */
} mcw_source;

typedef struct {
  int32_t type;         /* MCW_FILTER_* */
  int32_t n_bins;       /* CELL/UNIVERSE: number of ids (0 = all); ENERGY: number of edges */
  const int64_t *ids;   /* CELL: element ids; UNIVERSE: block ids */
  const double *edges;  /* ENERGY: increasing edges [eV] */
} mcw_filter;

typedef struct {
  int64_t histories;
  int64_t segments;
  int64_t collisions;
  int64_t absorptions;
  int64_t leaks;
  int64_t cutoffs;
  int64_t lost;
  int64_t reflections;
  double track_length;  /* total weighted track length [cm] */
  double time_ms;       /* device time of the batch, transport only */
  double kernel_ms;     /* device time of the dominant kernel(s) */
  int64_t kernel_launches;
  int64_t fp64_rechecks; /* precision MIXED: segments whose exit side was re-solved in FP64 */
  int64_t hex_segments;  /* segments in HEX8 elements (the rest are in TET4) */
  int64_t trace_records; /* visit records emitted (trace_histories); beyond the buffer they are dropped */
  double stage_ms;       /* device time of the staged tally's sums (tally_stage), part of kernel_ms */
} mcw_batch_stats;

/* Per-history record, for parity checks against a history-based reference. */
typedef struct {
  int64_t id;          /* particle seed id (batch-1)*n + i + 1 */
  uint64_t digest;     /* hash chain over (element, event) of every segment */
  uint64_t seed;       /* final tracking seed */
  double E;            /* final energy */
  double x, y, z;      /* final position */
  int32_t n_segments;
  int32_t n_collisions;
  int32_t elem;        /* final element */
  int32_t fate;        /* MCW_FATE_* */
} mcw_history;

/* ---- context -------------------------------------------------------------- */

/* Replaces openmc_init + openmc_simulation_init (OpenMCStudy.C:127-147). */
int mcw_create(mcw_ctx **ctx, const mcw_options *opts);
/* Replaces openmc_finalize (OpenMCStudy.C:213-215). */
int mcw_destroy(mcw_ctx *ctx);
/* Change the schedule and tuning fields of a live context (mode, precision,
 * block_size, bank_capacity, lanes_min, waves_per_eu, sort_cadence,
 * use_graph, trace_histories, record_histories, max_segments) without re-uploading the problem; device,
 * master_seed and seed_roundtrip cannot change.  Used by the autotuner. */
int mcw_options_set(mcw_ctx *ctx, const mcw_options *opts);
/* Replaces openmc_err_msg. */
const char *mcw_last_error(const mcw_ctx *ctx);
void mcw_default_options(mcw_options *opts);

/* ---- problem definition -------------------------------------------------- */

/* The mesh the study traces (libMesh mesh of OpenMCStudy / TraceRay;
 * OpenMCStudy.C:152-187 maps each element to a cell and each subdomain to a
 * universe).  conn_offsets has n_elem+1 entries.  face_boundary, optional,
 * gives for every side slot (element-major, sides of the element's type) the
 * boundary id of an external side; NULL puts every external side on
 * boundary 0.  Neighbours and outward face planes are derived here. */
int mcw_mesh_set(mcw_ctx *ctx, int64_t n_nodes, const double *xyz, int64_t n_elem,
                 const uint8_t *elem_type, const int64_t *conn_offsets, const int32_t *conn,
                 const int32_t *subdomain, const int32_t *face_boundary);

/* RayBCs: ReflectRayBC / KillRayBC on a boundary id (openmc_study.i:49-54). */
int mcw_bc_set(mcw_ctx *ctx, int32_t boundary_id, int32_t kind);

/* Nuclear data (replaces the HDF5 library OpenMC loads in openmc_init). */
int mcw_xs_set(mcw_ctx *ctx, int32_t n_nuclides, const mcw_nuclide *nuclides);

/* A further table of nuclide `nuclide` (index into mcw_xs_set's list) at
 * table->temperature, with fission cross sections if and only if the first
 * table has them; awr, nu and the Watt spectrum stay the first table's (one
 * set of reaction products per nuclide, as in OpenMC's data files).
 * mcw_xs_set drops every added table.  A nuclide with several tables is read at each element's
 * temperature by OpenMC's default "nearest" method: the table whose kT is
 * closest to the element's kT (CollisionKernel.C:109 sets the particle's
 * sqrt(kT) from the element's temperature variable; ties go to the lower
 * temperature). */
int mcw_xs_add_temperature(mcw_ctx *ctx, int32_t nuclide, const mcw_nuclide *table);

/* Boltzmann constant [eV/K] (OpenMC K_BOLTZMANN): kT of a table is
 * MCW_K_BOLTZMANN * temperature. */
#define MCW_K_BOLTZMANN 8.617333262e-5

/* Element temperatures [K], caller order (MaCaw's CONSTANT MONOMIAL
 * temperature variable); n_elem = 0 clears them.  A nuclide whose nearest
 * table is further than `tolerance` K from an element's temperature is an
 * error at the next batch; INFINITY reads the nearest table however far, as
 * OpenMC does at run time (its temperature_tolerance only decides which
 * tables are loaded).  Without temperatures, every element reads each
 * nuclide's first table. */
int mcw_temperature_set(mcw_ctx *ctx, int64_t n_elem, const double *kelvin, double tolerance);
int mcw_materials_set(mcw_ctx *ctx, int32_t n_materials, const mcw_material *materials);

/* CollisionKernel `blocks` / `materials` parameters (CollisionKernel.C:63-69). */
int mcw_block_material_map(mcw_ctx *ctx, int32_t n, const int32_t *blocks,
                           const int32_t *materials);

/* External source (settings.xml <source>; sampled as OpenMC's
 * sample_external_source on stream 1). */
int mcw_source_set(mcw_ctx *ctx, const mcw_source *src);

/* OpenMCTally::initialize (OpenMCTally.C:103-225): a tally with filters,
 * scores and nuclide bins (-1 = "total").  Filters are ordered as
 * OpenMCTally sorts them (universe < cell < energy). */
int mcw_tally_add(mcw_ctx *ctx, int32_t id, int32_t estimator, int32_t n_filters,
                  const mcw_filter *filters, int32_t n_scores, const int32_t *scores,
                  int32_t n_nuclides, const int32_t *nuclides);

/* Shape of a tally's results array: n_filter_bins x n_score_bins. */
int mcw_tally_shape(mcw_ctx *ctx, int32_t id, int64_t *n_filter_bins, int64_t *n_score_bins);

/* ---- execution ------------------------------------------------------------ */

/* One batch: generateRays + execute + postExecuteStudy (OpenMCStudy.C:218-402)
 * for particles 1..n_particles of `batch` (1-based), then VALUE -> SUM,
 * SUM_SQ as OpenMC's finalize_batch.  EIGENVALUE runs go batch 1, 2, ... with
 * one n_particles (batch 1 restarts the run); each batch ends with the
 * fission bank sorted and resampled into the next batch's source. */
int mcw_run_batch(mcw_ctx *ctx, int64_t batch, int64_t n_particles, mcw_batch_stats *out);

/* No counterpart in the reference.  mcw_run_batch for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same batch, with the same histories
 * and tallies, and no hipMalloc, hipFree or allocating library call.  Its
 * counters and source bank live in a static array of 64 MiB per device that
 * the library holds until the process ends; histories go through it in
 * chunks of at most MCW_NOALLOC_CHUNK (an explicit bank_capacity above that
 * is an error).  Everything else must exist already: the context must have
 * run mcw_run_batch since its problem or options last changed (set-up
 * allocates), and history records need a record buffer that an earlier
 * mcw_run_batch of at least n_particles histories allocated.  Fixed-source
 * runs in mode BATCHED only, without sorting or visit records.  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (from
 * different threads or contexts) race. */
#define MCW_NOALLOC_CHUNK 986624
int mcw_run_batch_noalloc(mcw_ctx *ctx, int64_t batch, int64_t n_particles, mcw_batch_stats *out);

/* ---- eigenvalue runs ------------------------------------------------------ */

/* A fission source site (OpenMC SourceSite): position, direction, energy,
 * the history that banked it (index within its batch, from 0), its index
 * among that history's sites, and the caller element it lies in. */
typedef struct {
  double r[3];
  double u[3];
  double E;
  int64_t parent;
  int32_t progeny;
  int32_t element;
} mcw_site;

/* The keff state after the last EIGENVALUE batch (OpenMC
 * calculate_generation_keff / calculate_average_keff, called from
 * OpenMCStudy::finalizeGeneration, OpenMCStudy.C:404-452). */
typedef struct {
  int64_t batch;          /* last batch run */
  int64_t fission_sites;  /* sites banked in it (the fission bank holds at most 3 n_particles) */
  int64_t sites_dropped;  /* sites that found the bank full (OpenMC warns and drops them) */
  int32_t n_active;       /* active batches so far */
  int32_t pad;
  double k_generation;    /* the batch's track-length estimate: sum w l nuSigma_f / n_particles */
  double k_collision;     /* the batch's collision estimate: sum w nuSigma_f / Sigma_t / n */
  double k_absorption;    /* the batch's absorption estimate: sum w nu sigma_f / sigma_a / n */
  double keff;            /* mean of k_generation over active batches (inactive: the last
                             k_generation); used for the next batch's fission yield */
  double keff_std;        /* standard deviation of that mean (0 below two active batches) */
  double keff_used;       /* the keff the batch divided the fission yield by */
  double close_ms;        /* wall time of the batch close: keff reduction, bank sort, resampling */
} mcw_keff_result;

int mcw_keff_get(mcw_ctx *ctx, mcw_keff_result *out);

/* The last EIGENVALUE batch's fission bank in OpenMC's sorted order (parent,
 * then progeny).  *n_out gets the number of sites; at most `max` are copied
 * (out may be NULL). */
int mcw_fission_bank_get(mcw_ctx *ctx, mcw_site *out, size_t max, int64_t *n_out);

/* The source of the next EIGENVALUE batch: n_particles sites resampled from
 * the fission bank, each with the parent and progeny of the site it copies. */
int mcw_source_bank_get(mcw_ctx *ctx, mcw_site *out, size_t max, int64_t *n_out);

/* Read a tally's results kind (MCW_RESULT_*) into out[n_filter_bins*n_score_bins]. */
int mcw_tally_get(mcw_ctx *ctx, int32_t id, int32_t kind, double *out, size_t n);

/* Zero every tally (VALUE, SUM, SUM_SQ). */
int mcw_tally_reset(mcw_ctx *ctx);

/* Per-history records of the last batch (options.record_histories = 1). */
int mcw_history_get(mcw_ctx *ctx, mcw_history *out, size_t n);

/* Diagnostics (no reference counterpart): evaluate the device log,
 * sin/cos(2 pi frac(x)), sqrt, one random draw from each seed and a seed
 * skip-ahead future_seed(i*152917 + 7, seed_i) on n host inputs, so tests can
 * compare the device's arithmetic with the CPU's bit for bit. */
int mcw_debug_math_probe(mcw_ctx *ctx, const double *x, int64_t n, double *log_out, double *sin_out,
                         double *cos_out, double *sqrt_out, const uint64_t *seeds, double *prn_out,
                         uint64_t *seed_out);

/* One element visit of a traced history, as a record of MOOSE's
 * RayTracingExodus output (test/tests/study/openmc_study.i:73-81): the
 * segment from where the history entered the element (or started, or was
 * reflected) to where it left it (or died), with the energy and tracking seed
 * at the visit's end.  event: MCW_EVENT_*. */
enum { MCW_EVENT_CROSS = 0, MCW_EVENT_REFLECT = 1, MCW_EVENT_LEAK = 2, MCW_EVENT_ABSORB = 4,
       MCW_EVENT_CUTOFF = 5, MCW_EVENT_LOST = 6 };
typedef struct {
  int64_t history; /* index within the batch, from 0 */
  int32_t visit;   /* visit index within the history, from 0 */
  int32_t element; /* caller element id */
  double start[3];
  double end[3];
  double energy;
  uint64_t seed_tracking;
  int32_t event;
  int32_t pad;
} mcw_trace_record;

/* Visit records of the last batch, ordered by (history, visit).  *n_out gets
 * the number available; at most `max` are copied (out may be NULL). */
int mcw_trace_get(mcw_ctx *ctx, mcw_trace_record *out, size_t max, int64_t *n_out);

/* Diagnostics (no reference counterpart): the device's union-grid lookup of
 * nuclide `nuc` at n host energies: union index and (total, absorption,
 * scatter, fission) per energy, for comparison with a per-nuclide search. */
int mcw_debug_xs_probe(mcw_ctx *ctx, int32_t nuc, const double *E, int64_t n, int32_t *k_out,
                       double *xs_out);

/* Diagnostics (no reference counterpart): the device's FP64 atomic-add rate
 * (adds per second) for pseudo-random adds, with the tallies' instruction,
 * into an array of n_bins doubles (a power of two, 2^10 .. 2^28); sum_out,
 * if not NULL, receives the array's sum (the number of adds made, after a
 * warm-up launch of 8 adds per thread).  The tally kernels' roofline: every
 * segment of a per-element flux tally makes one such add. */
int mcw_debug_atomic_rate(mcw_ctx *ctx, int64_t n_bins, double *adds_per_s, double *sum_out);

/* Diagnostics (no reference counterpart): the device's rate of reads of
 * pseudo-random 128-byte lines (lines per second) from an array of n_lines
 * lines (a power of two, 2^10 .. 2^28), eight independent reads in flight
 * per thread.  The gather kernels' roofline: a segment reads one (mixed
 * precision) or two (FP64) element lines. */
int mcw_debug_gather_rate(mcw_ctx *ctx, int64_t n_lines, double *lines_per_s);

/* ---- serving (continuous batching of fixed-source requests) -------------- */

/* A request: a source, a particle count and a per-element flux tally
 * (track-length) on one resident mesh (mesh_slot: index into the contexts
 * given to mcw_server_create). */
typedef struct {
  int64_t request_id;
  int32_t mesh_slot;
  int32_t pad;
  int64_t n_particles;
  mcw_source source;
} mcw_request;

typedef struct {
  int64_t request_id;
  int64_t histories;
  int64_t segments;
  int64_t collisions;
  double t_submit_ms;
  double t_admit_ms;
  double t_done_ms;
} mcw_request_status;

/* Build a server over `n_meshes` resident contexts (each a mesh + XS +
 * materials + block map, set as for mcw_run_batch; a context's own source and
 * tallies are not used).  max_inflight bounds concurrently admitted requests;
 * bank_bytes (0: 16 GiB) bounds their memory, each request counting its flux
 * tally plus min(n_particles, 2^18) x 76 B of particle bank (see
 * mcw_server_tune).  A request's
 * histories are those of batch 1 with n_particles histories and master seed
 * request_id + 1, whatever else is in flight. */
typedef struct mcw_server mcw_server;
int mcw_server_create(mcw_server **srv, mcw_ctx *const *meshes, int32_t n_meshes,
                      int32_t max_inflight, int64_t bank_bytes);
int mcw_server_destroy(mcw_server *srv);
/* No counterpart in the reference.  Scheduling of a server with no request
 * queued: histories per round (1024 .. 2^24, default 2^18; a request's bank
 * share counts min(n_particles, round_histories) x 76 B) and the age after
 * which an admitted request is served before shorter ones (default 250 ms). */
int mcw_server_tune(mcw_server *srv, int64_t round_histories, double age_ms);
const char *mcw_server_last_error(const mcw_server *srv);
int mcw_submit(mcw_server *srv, const mcw_request *req);
/* Advance the device work; returns the number of finished requests waiting. */
int mcw_poll(mcw_server *srv, int32_t *n_done);
/* Pop one finished request (in completion order); flux receives n_elem
 * values of its mesh (track-length flux per source particle, caller element
 * order), or is NULL to drop them. */
int mcw_result(mcw_server *srv, mcw_request_status *status, double *flux, size_t n);

#ifdef __cplusplus
}
#endif

#endif /* MACAW_AMD_H */
