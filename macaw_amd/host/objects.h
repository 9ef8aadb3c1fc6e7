// objects.h -- MOOSE-shaped host objects over the C ABI of libmacaw_amd.
//
// Each class keeps the name, the parameters and the error messages of the
// MaCaw / MOOSE object it stands in for, so an input written for the
// reference (test/tests/study/openmc_study.i, test/tests/tally/openmc_tally.i)
// reads the same here.  The device does the tracing: the study's execute()
// is one mcw_run_batch, where MaCaw's RayTracingStudy calls
// CollisionKernel::onSegment once per segment.
#pragma once

#include <cmath>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/macaw_amd.h"
#include "hit.h"

namespace mcwhost {

struct MooseError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Typed view of one input block.
class InputParameters {
public:
  explicit InputParameters(const HitNode *n) : node_(n) {}
  const std::string &path() const { return node_->path; }
  bool isParamValid(const std::string &k) const { return node_->has(k); }
  std::string getString(const std::string &k) const;
  std::string getString(const std::string &k, const std::string &dflt) const;
  std::vector<std::string> getWords(const std::string &k) const;
  long long getInt(const std::string &k) const;
  long long getInt(const std::string &k, long long dflt) const;
  double getReal(const std::string &k) const;
  double getReal(const std::string &k, double dflt) const;
  bool getBool(const std::string &k, bool dflt) const;
  std::vector<long long> getInts(const std::string &k) const;
  std::vector<double> getReals(const std::string &k) const;

private:
  const HitNode *node_;
};

[[noreturn]] void paramError(const InputParameters &p, const std::string &param, const std::string &msg);
[[noreturn]] void mooseError(const std::string &msg);

// Mesh built from the [Mesh] block (libMesh node order and side numbering).
struct MeshData {
  std::vector<double> xyz;
  std::vector<uint8_t> type;
  std::vector<int64_t> conn_off{0};
  std::vector<int32_t> conn;
  std::vector<int32_t> subdomain;
  std::vector<int32_t> face_boundary; // per side slot; -1: no boundary id
  std::map<std::string, int32_t> boundary_names;
  int dim = 3;                        // 2: QUAD4 elements (MESH_QUAD4), to be extruded
  int64_t n_nodes() const { return (int64_t)xyz.size() / 3; }
  int64_t n_elem() const { return (int64_t)type.size(); }
};

// Driver-side element type of a 2D mesh before extrusion (never reaches the library).
constexpr uint8_t MESH_QUAD4 = 200;

// `input_dir`: where relative mesh file names are looked up (the input file's directory).
MeshData build_mesh(const HitNode &mesh_block, const std::string &input_dir);

struct NuclideData {
  std::string name;
  double awr;
  std::vector<double> E, t, a, s;
  std::vector<double> f;              // fission (part of absorption); empty: not fissionable
  double nu[2] = {0.0, 0.0};          // nu(E) = nu[0] + nu[1] E
  double watt[2] = {0.988e6, 2.249e-6}; // fission spectrum a [eV], b [1/eV]
  double temperature = 293.6;         // K
  double kappa = 0.0;                 // eV released per fission (kappa-fission scores)
  std::vector<NuclideData> temps;     // tables at further temperatures
};

struct MaterialData {
  int32_t id;
  std::vector<int32_t> nuclides;
  std::vector<double> densities;
};

class Problem;

// MaCaw CollisionKernel (src/raykernels/CollisionKernel.C:24-96).
class CollisionKernel {
public:
  CollisionKernel(const InputParameters &p, Problem &fe);
  void initialSetup();
  const std::map<int, int> &blockToMaterial() const { return _block_to_openmc_materials; }
  const std::string &temperatureVariable() const { return _temperature; }

private:
  InputParameters _params;
  Problem &_fe;
  std::map<int, int> _block_to_openmc_materials;
  std::string _temperature;
  bool _verbose;
};

// MOOSE ReflectRayBC / KillRayBC.
class RayBC {
public:
  RayBC(const InputParameters &p, Problem &fe, int kind);
  void apply(mcw_ctx *ctx) const;
  const std::vector<int32_t> &ids() const { return _ids; }
  int kind() const { return _kind; }

private:
  std::vector<int32_t> _ids;
  int _kind;
};

// MaCaw OpenMCTally (src/userobjects/OpenMCTally.C:30-225).
class OpenMCTally {
public:
  OpenMCTally(const InputParameters &p, Problem &fe);
  void initialize();
  int id() const { return _id; }
  const std::vector<std::string> &filters() const { return _filters; }
  const std::vector<std::string> &scores() const { return _scores; }
  const std::vector<std::string> &nuclides() const { return _nuclides; }
  const std::vector<int64_t> &strides() const { return _strides; }
  const std::vector<int64_t> &filterBins() const { return _filter_bins; }
  const std::vector<int64_t> &universeIds() const { return _universe_ids; }
  int64_t cellBin(int64_t elem) const; // bin of an element in the cell filter, -1 if absent
  int64_t nScoreBins() const { return (int64_t)_scores.size() * (int64_t)_nuclide_idx.size(); }
  const std::vector<int32_t> &nuclideIndices() const { return _nuclide_idx; }

private:
  InputParameters _params;
  Problem &_fe;
  int _id;
  int _estimator;
  std::vector<std::string> _scores, _filters, _nuclides;
  std::vector<double> _energy_bins;
  std::vector<long long> _cell_bins, _block_bins;
  std::vector<int64_t> _strides, _filter_bins, _universe_ids;
  std::map<int64_t, int64_t> _cell_pos;
  std::vector<int32_t> _nuclide_idx;
};

// MaCaw OpenMCTallyAux (src/auxkernels/OpenMCTallyAux.C:26-215): one value
// per element from a tally's SUM.
class OpenMCTallyAux {
public:
  OpenMCTallyAux(const InputParameters &p, Problem &fe);
  double computeValue(int64_t elem, const std::vector<double> &sum) const;
  const std::string &variable() const { return _variable; }
  int tallyId() const { return _tally_id; }

private:
  InputParameters _params;
  Problem &_fe;
  int _tally_id;
  int _granularity; // 0 global, 1 universe, 2 cell
  std::string _score, _nuclide, _variable;
  int _energy_bin;
  bool _score_major; // score_index_layout = macaw: score_bin * n_nuclides + nuc_bin (OpenMCTallyAux.C:116-119)
};

// MaCaw OpenMCStudy (src/userobjects/OpenMCStudy.C:49-402).
class OpenMCStudy {
public:
  OpenMCStudy(const InputParameters &p, Problem &fe);
  ~OpenMCStudy();
  void initialSetup(); // creates the device context and uploads the problem
  void generateRays();
  void execute();
  void postExecuteStudy();
  mcw_ctx *ctx() const { return _ctx; }
  const mcw_batch_stats &lastStats() const { return _stats; }
  const mcw_keff_result &lastKeff() const { return _keff; }

private:
  InputParameters _params;
  Problem &_fe;
  mcw_ctx *_ctx = nullptr;
  mcw_batch_stats _stats{};
  mcw_keff_result _keff{};
  bool _verbose;
  int64_t _t_step = 0;
};

// The problem: mesh, Monte Carlo settings and data, objects, batches.
class Problem {
public:
  Problem(std::unique_ptr<HitNode> root, const std::string &input);
  ~Problem();
  void run();
  void dump(const std::string &path) const;
  std::map<int, int> blockMap() const;

  MeshData mesh;
  std::vector<NuclideData> nuclides;
  std::vector<MaterialData> materials;
  mcw_source source{};
  mcw_options options{};
  int64_t particles = 0;
  int n_batches = 0;
  std::string file_base;
  std::map<std::string, int> aux_vars;   // name -> order (0 = CONSTANT)
  std::map<std::string, std::string> aux_family;
  std::map<std::string, double> aux_initial;  // AuxVariables initial_condition
  std::vector<double> elem_temperature;       // K per element: the collision kernel's temperature variable
  double temperature_tolerance = INFINITY;    // K; [MonteCarlo] temperature_tolerance
  std::vector<std::unique_ptr<OpenMCTally>> tallies;
  std::vector<std::unique_ptr<OpenMCTallyAux>> auxkernels;
  std::vector<std::unique_ptr<CollisionKernel>> raykernels;
  std::vector<std::unique_ptr<RayBC>> raybcs;
  std::unique_ptr<OpenMCStudy> study;
  bool csv = true;
  std::vector<std::string> ray_outputs; // [Outputs/<name>] type = RayTracingExodus (execute_on = final)
  const OpenMCTally *tally(int id) const;

private:
  std::unique_ptr<HitNode> _root;
  std::string _input;
  void read_monte_carlo(const HitNode &mc);
  std::vector<double> element_values(const std::string &var) const;
  void write_outputs(const std::vector<std::vector<double>> &rows);
};

} // namespace mcwhost
