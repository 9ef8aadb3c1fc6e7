// mcw_mesh.h -- host mesh container and derivation (planes, neighbours, CDF).
#pragma once

#include <stdint.h>

#include <set>
#include <string>
#include <vector>

namespace mcw {

struct HostMesh {
  int64_t n_nodes = 0, n_elem = 0;
  std::vector<double> xyz;
  std::vector<uint8_t> type;
  std::vector<int64_t> conn_off;
  std::vector<int32_t> conn;
  std::vector<int32_t> subdomain;
  // derived
  int uniform_sides = 0;            // 4 all-TET4, 6 all-HEX8, 0 mixed
  std::vector<int64_t> face_off;    // n_elem + 1
  std::vector<double> plane;        // 4 per side slot
  std::vector<int32_t> nbr;         // neighbour*8 + side, split*8 + 7, or -1 - boundary id
  struct Split {                    // HEX8 side covered by two TET4 sides
    double m[4];                    // plane through the diagonal: m.x - c
    int32_t elem[2];                // tet taking m.x - c <= 0, tet taking the rest
    int32_t side[2];                // their side on the hex side
  };
  std::vector<Split> splits;
  std::set<int32_t> boundary_ids;   // ids of external sides
  std::vector<double> cdf;          // element volume CDF (6x volumes)
  // guide table of the CDF: cdf_guide[j] = first e with cdf[e] > j * total / G,
  // j = 0 .. G (cdf_guide[G] = n_elem), G = min(n_elem, 2^22) buckets
  std::vector<int32_t> cdf_guide;
  // device order: TET4 then HEX8, each along a Morton curve of the centroids
  int64_t n_tet = 0;                // internal ids below n_tet are TET4
  std::vector<int32_t> orig_of_int; // internal position -> caller element
  std::vector<int32_t> int_of_orig; // caller element -> internal position
  // point locator (built on demand): uniform grid, CSR of internal ids
  bool have_grid = false;
  double grid_lo[3] = {0, 0, 0}, grid_inv[3] = {0, 0, 0};
  int32_t grid_n[3] = {1, 1, 1};
  std::vector<int32_t> cell_start, cell_elem;
};

// Build the uniform-grid point locator (about one element per cell).
void build_locator(HostMesh &m);

extern const int HEX_KUHN_TETS[6][4];

// Returns an error message, or "" on success.
std::string derive_mesh(HostMesh &m, const int32_t *face_boundary);

// Lowest-index element whose side planes all contain p within tol, or -1
// (through the grid locator once it is built, else by a scan).
int64_t locate_point(const HostMesh &m, const double *p, double tol);

double tet_volume6(const double *a, const double *b, const double *c, const double *d);

} // namespace mcw
