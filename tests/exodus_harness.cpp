// Test harness: write a RayTracingExodus-layout file from raw visit records
// with the driver's C++ writer (macaw_amd/host/exodus.cpp), no GPU needed.
//   exodus_harness records.bin out.e batch n_particles master_seed
// records.bin: an int64 count, then that many mcw_trace_record.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "../macaw_amd/host/exodus.h"

int main(int argc, char **argv)
{
  if (argc != 6) return 2;
  std::ifstream f(argv[1], std::ios::binary);
  int64_t n = 0;
  f.read((char *)&n, 8);
  std::vector<mcw_trace_record> r((size_t)n);
  f.read((char *)r.data(), (std::streamsize)(n * sizeof(mcw_trace_record)));
  if (!f) return 3;
  mcwhost::write_rays_exodus(argv[2], r, std::atoll(argv[3]), std::atoll(argv[4]), std::strtoull(argv[5], 0, 10),
                             "harness");
  return 0;
}
