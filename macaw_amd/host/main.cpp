// main.cpp -- macaw-amd-opt: run a MaCaw-style input on the GPU library.
//
//   macaw-amd-opt -i input.i [Block/sub/key=value ...] [--check-input]
//                 [--dump-problem file]
//
// --check-input builds and validates every object without touching the GPU;
// --dump-problem writes the problem as the library receives it (mesh, data,
// boundary conditions, block map, source, settings) and exits.
#include <fstream>
#include <iostream>
#include <sstream>

#include "hit.h"
#include "objects.h"

int main(int argc, char **argv)
{
  std::string input, dump;
  bool check_only = false;
  std::vector<std::string> overrides;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "-i" && i + 1 < argc) input = argv[++i];
    else if (a == "--check-input") check_only = true;
    else if (a == "--dump-problem" && i + 1 < argc) dump = argv[++i];
    else if (a.find('=') != std::string::npos) overrides.push_back(a);
    else {
      std::cerr << "usage: macaw-amd-opt -i input.i [Block/key=value ...] [--check-input] [--dump-problem file]\n";
      return 2;
    }
  }
  if (input.empty()) {
    std::cerr << "usage: macaw-amd-opt -i input.i\n";
    return 2;
  }
  try {
    std::ifstream f(input);
    if (!f) throw std::runtime_error("cannot open " + input);
    std::stringstream ss;
    ss << f.rdbuf();
    auto root = mcwhost::parse_hit(ss.str(), input);
    for (auto &o : overrides) mcwhost::apply_override(*root, o);
    mcwhost::Problem problem(std::move(root), input);
    if (!dump.empty()) {
      problem.dump(dump);
      return 0;
    }
    if (check_only) {
      std::cout << "input ok: " << problem.mesh.n_elem() << " elements, " << problem.tallies.size() << " tallies, "
                << problem.auxkernels.size() << " aux kernels" << std::endl;
      return 0;
    }
    problem.run();
  } catch (const std::exception &e) {
    std::cerr << "\n*** ERROR ***\n" << e.what() << "\n" << std::endl;
    return 1;
  }
  return 0;
}
