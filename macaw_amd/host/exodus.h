// exodus.h -- traced rays as MOOSE's RayTracingExodus output writes them.
//
// One EDGE2 element per element visit (consecutive visits of a ray share a
// node) with the Ray data as element variables: MaCaw's eight aux data
// (src/userobjects/OpenMCStudy.C:190-197) plus `intersections` (visit index)
// and `ray_id`, in alphabetical order, in an Exodus II file (NetCDF classic
// with 64-bit offsets), laid out as the reference's gold file
// test/tests/study/gold/openmc_study_rays.e.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/macaw_amd.h"

namespace mcwhost {

// OpenMC's seed of particle `id`'s stream (0 tracking, 1 source, 2 URR).
uint64_t init_seed(uint64_t id, uint64_t stream, uint64_t master_seed);

// An Exodus II mesh as stored: coordinates (z = 0 in 2D), element blocks with
// 0-based connectivity, side sets of (0-based element, 1-based Exodus side).
struct ExodusMesh {
  struct Block {
    int64_t id = 0;
    std::string name, elem_type;
    int nodes_per_elem = 0;
    std::vector<int64_t> conn;
  };
  struct SideSet {
    int64_t id = 0;
    std::string name;
    std::vector<int64_t> elem, side;
  };
  int dim = 3;
  std::vector<double> xyz; // 3 per node
  std::vector<Block> blocks;
  std::vector<SideSet> sidesets;
};

// Read an Exodus II file (NetCDF classic, 64-bit offset or CDF-5).  Throws
// std::runtime_error with the reason when it cannot.
ExodusMesh read_exodus(const std::string &path);

// Write the visit records of batch `batch` (ordered by history, visit) to
// `path`.  Ray ids are (batch - 1) * n_particles + history.  Throws
// std::runtime_error when the file cannot be written.
void write_rays_exodus(const std::string &path, const std::vector<mcw_trace_record> &recs, int64_t batch,
                       int64_t n_particles, uint64_t master_seed, const std::string &title);

} // namespace mcwhost
