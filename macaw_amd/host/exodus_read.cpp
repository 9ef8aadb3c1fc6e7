// exodus_read.cpp -- Exodus II meshes as MOOSE's FileMeshGenerator reads them.
//
// An Exodus II file is a NetCDF file of fixed variables.  This reads the
// NetCDF classic formats (CDF-1, 64-bit offset CDF-2 and 64-bit data CDF-5:
// a big-endian header of dimensions, attributes and variables, then each
// fixed variable's values at its `begin` offset) and from them the Exodus
// mesh: coordinates (coordx/coordy/coordz or coord), element blocks
// (connect<i> with its elem_type attribute, ids from eb_prop1), and side sets
// (elem_ss<i>, side_ss<i>, ids from ss_prop1; names from ss_names).
// NetCDF-4 (HDF5) files are refused with a message.
#include "exodus.h"

#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>

namespace mcwhost {

namespace {

enum { NC_DIMENSION = 10, NC_VARIABLE = 11, NC_ATTRIBUTE = 12 };
enum { NC_BYTE = 1, NC_CHAR = 2, NC_SHORT = 3, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6, NC_UBYTE = 7,
       NC_USHORT = 8, NC_UINT = 9, NC_INT64 = 10, NC_UINT64 = 11 };

int type_size(int t)
{
  switch (t) {
    case NC_BYTE: case NC_CHAR: case NC_UBYTE: return 1;
    case NC_SHORT: case NC_USHORT: return 2;
    case NC_INT: case NC_FLOAT: case NC_UINT: return 4;
    case NC_DOUBLE: case NC_INT64: case NC_UINT64: return 8;
    default: return 0;
  }
}

struct Var {
  std::string name;
  std::vector<int64_t> dims;           // dimension ids
  std::map<std::string, std::string> text_atts;
  int type = 0;
  uint64_t begin = 0;
};

struct NcFile {
  std::vector<uint8_t> b;
  size_t pos = 0;
  int version = 1;
  std::vector<std::pair<std::string, int64_t>> dims;
  std::map<std::string, Var> vars;
  std::string path;

  [[noreturn]] void bad(const std::string &m) const { throw std::runtime_error(path + ": " + m); }
  uint64_t u(int n)
  {
    if (pos + n > b.size()) bad("truncated NetCDF header");
    uint64_t v = 0;
    for (int k = 0; k < n; ++k) v = (v << 8) | b[pos + k];
    pos += n;
    return v;
  }
  // NON_NEG: a count or length; negative (a 64-bit CDF-5 value above
  // INT64_MAX) or larger than the file (every counted item takes at least a
  // byte) is refused before it is used in any size computation
  int64_t count()
  {
    const int64_t v = (int64_t)u(version == 5 ? 8 : 4);
    if (v < 0 || (uint64_t)v > b.size()) bad("bad count in NetCDF header");
    return v;
  }
  // bytes left in the file from `at`
  uint64_t left(uint64_t at) const { return at <= b.size() ? b.size() - at : 0; }
  std::string name()
  {
    const int64_t n = count();
    if ((uint64_t)n > left(pos)) bad("bad name in NetCDF header");
    std::string s((const char *)&b[pos], (size_t)n);
    pos += (size_t)((n + 3) & ~3);
    return s;
  }
  // attribute list; text attributes kept, others skipped
  void atts(std::map<std::string, std::string> *out)
  {
    const uint64_t tag = u(4);
    const int64_t n = count();
    if (tag == 0) return;
    if (tag != NC_ATTRIBUTE) bad("bad attribute list");
    for (int64_t i = 0; i < n; ++i) {
      const std::string nm = name();
      const int t = (int)u(4);
      const int64_t ne = count();
      const int sz = type_size(t);
      if (!sz) bad("unknown attribute type");
      if ((uint64_t)ne > left(pos) / (uint64_t)sz) bad("truncated attribute");
      const size_t bytes = (size_t)ne * sz;
      if (out && t == NC_CHAR) (*out)[nm] = std::string((const char *)&b[pos], bytes);
      pos += (bytes + 3) & ~(size_t)3;
    }
  }
  void parse()
  {
    if (b.size() < 8) bad("not a NetCDF file");
    if (b[0] == 0x89 && b[1] == 'H' && b[2] == 'D' && b[3] == 'F')
      bad("a NetCDF-4 (HDF5) file; write the mesh as NetCDF classic or 64-bit offset (e.g. ncks -3)");
    if (b[0] != 'C' || b[1] != 'D' || b[2] != 'F') bad("not a NetCDF file");
    version = b[3];
    if (version != 1 && version != 2 && version != 5) bad("unknown NetCDF version");
    pos = 4;
    count(); // numrecs
    // dimensions
    uint64_t tag = u(4);
    int64_t n = count();
    if (tag == NC_DIMENSION) {
      for (int64_t i = 0; i < n; ++i) {
        std::string nm = name();
        dims.push_back({nm, count()});
      }
    } else if (tag != 0) {
      bad("bad dimension list");
    }
    atts(nullptr); // global attributes
    tag = u(4);
    n = count();
    if (tag == 0) return;
    if (tag != NC_VARIABLE) bad("bad variable list");
    for (int64_t i = 0; i < n; ++i) {
      Var v;
      v.name = name();
      const int64_t nd = count();
      for (int64_t k = 0; k < nd; ++k) v.dims.push_back(count());
      atts(&v.text_atts);
      v.type = (int)u(4);
      count(); // vsize
      v.begin = u(version == 1 ? 4 : 8);
      for (int64_t d : v.dims)
        if (d < 0 || d >= (int64_t)dims.size()) bad("variable " + v.name + " names an unknown dimension");
      vars[v.name] = v;
    }
  }
  int64_t dim(const std::string &nm) const
  {
    for (auto &d : dims)
      if (d.first == nm) return d.second;
    return -1;
  }
  const Var *var(const std::string &nm) const
  {
    auto it = vars.find(nm);
    return it == vars.end() ? nullptr : &it->second;
  }
  // values of a fixed variable; a product of dimensions that cannot fit in
  // the file is refused before it can overflow
  int64_t n_values(const Var &v) const
  {
    uint64_t n = 1;
    for (int64_t d : v.dims) {
      const uint64_t len = (uint64_t)dims[d].second;
      if (len == 0) bad("variable " + v.name + " is a record variable");
      if (n > b.size() / len) bad("variable " + v.name + " is larger than the file");
      n *= len;
    }
    return (int64_t)n;
  }
  // the variable's n values of sz bytes lie inside the file
  void check_extent(const Var &v, int64_t n, int sz) const
  {
    if (v.begin > b.size() || (uint64_t)n > left(v.begin) / (uint64_t)sz)
      bad(v.name + " lies past the end of the file");
  }
  std::vector<double> reals(const Var &v) const
  {
    const int64_t n = n_values(v);
    const int sz = type_size(v.type);
    if (v.type != NC_DOUBLE && v.type != NC_FLOAT) bad(v.name + " is not floating point");
    check_extent(v, n, sz);
    std::vector<double> out((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
      uint64_t w = 0;
      for (int k = 0; k < sz; ++k) w = (w << 8) | b[v.begin + (uint64_t)i * sz + k];
      if (sz == 8) {
        double d;
        memcpy(&d, &w, 8);
        out[i] = d;
      } else {
        uint32_t w4 = (uint32_t)w;
        float f;
        memcpy(&f, &w4, 4);
        out[i] = f;
      }
    }
    return out;
  }
  std::vector<int64_t> ints(const Var &v) const
  {
    const int64_t n = n_values(v);
    const int sz = type_size(v.type);
    if (v.type != NC_INT && v.type != NC_INT64 && v.type != NC_UINT && v.type != NC_SHORT)
      bad(v.name + " is not integer");
    check_extent(v, n, sz);
    std::vector<int64_t> out((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
      uint64_t w = 0;
      for (int k = 0; k < sz; ++k) w = (w << 8) | b[v.begin + (uint64_t)i * sz + k];
      if (v.type == NC_INT) out[i] = (int32_t)(uint32_t)w;
      else if (v.type == NC_SHORT) out[i] = (int16_t)(uint16_t)w;
      else out[i] = (int64_t)w;
    }
    return out;
  }
  std::vector<std::string> names(const Var &v) const
  {
    if (v.type != NC_CHAR || v.dims.size() != 2) bad(v.name + " is not a name table");
    const int64_t rows = dims[v.dims[0]].second, width = dims[v.dims[1]].second;
    check_extent(v, n_values(v), 1);
    std::vector<std::string> out;
    for (int64_t r = 0; r < rows; ++r) {
      std::string s((const char *)&b[v.begin + (uint64_t)(r * width)], (size_t)width);
      s = s.substr(0, s.find('\0'));
      out.push_back(s);
    }
    return out;
  }
};

} // namespace

ExodusMesh read_exodus(const std::string &path)
{
  NcFile f;
  f.path = path;
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open mesh file '" + path + "'");
  f.b.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
  f.parse();
  ExodusMesh m;
  m.dim = (int)f.dim("num_dim");
  const int64_t nn = f.dim("num_nodes");
  if (m.dim < 2 || m.dim > 3 || nn <= 0) f.bad("not an Exodus II mesh (num_dim, num_nodes)");
  m.xyz.assign((size_t)(3 * nn), 0.0);
  if (const Var *c = f.var("coord")) { // older single coordinate array, (num_dim, num_nodes)
    const std::vector<double> v = f.reals(*c);
    if ((int64_t)v.size() < m.dim * nn) f.bad("coord holds fewer than num_dim x num_nodes values");
    for (int d = 0; d < m.dim; ++d)
      for (int64_t i = 0; i < nn; ++i) m.xyz[3 * i + d] = v[(size_t)(d * nn + i)];
  } else {
    const char *cn[3] = {"coordx", "coordy", "coordz"};
    for (int d = 0; d < m.dim; ++d) {
      const Var *c = f.var(cn[d]);
      if (!c) f.bad(std::string("no ") + cn[d]);
      const std::vector<double> v = f.reals(*c);
      if ((int64_t)v.size() < nn) f.bad(std::string(cn[d]) + " holds fewer than num_nodes values");
      for (int64_t i = 0; i < nn; ++i) m.xyz[3 * i + d] = v[(size_t)i];
    }
  }
  const int64_t nb = f.dim("num_el_blk");
  std::vector<int64_t> ids;
  if (const Var *p = f.var("eb_prop1")) ids = f.ints(*p);
  std::vector<std::string> bnames;
  if (const Var *p = f.var("eb_names")) bnames = f.names(*p);
  for (int64_t k = 0; k < nb; ++k) {
    const Var *c = f.var("connect" + std::to_string(k + 1));
    if (!c) continue; // an empty block has no connectivity
    ExodusMesh::Block B;
    B.id = k < (int64_t)ids.size() ? ids[k] : k + 1;
    B.name = k < (int64_t)bnames.size() ? bnames[k] : "";
    auto t = c->text_atts.find("elem_type");
    B.elem_type = t == c->text_atts.end() ? "" : t->second.substr(0, t->second.find('\0'));
    if (c->dims.size() != 2) f.bad("connect" + std::to_string(k + 1) + " is not (elements, nodes)");
    B.nodes_per_elem = (int)f.dims[c->dims[1]].second;
    B.conn = f.ints(*c);
    for (int64_t &q : B.conn) {
      if (q < 1 || q > nn) f.bad("block " + std::to_string(B.id) + " names a node outside 1.." + std::to_string(nn));
      --q; // to 0-based
    }
    m.blocks.push_back(std::move(B));
  }
  const int64_t ns = f.dim("num_side_sets");
  std::vector<int64_t> sids;
  if (const Var *p = f.var("ss_prop1")) sids = f.ints(*p);
  std::vector<std::string> snames;
  if (const Var *p = f.var("ss_names")) snames = f.names(*p);
  for (int64_t k = 0; k < ns; ++k) {
    const Var *e = f.var("elem_ss" + std::to_string(k + 1));
    const Var *s = f.var("side_ss" + std::to_string(k + 1));
    ExodusMesh::SideSet S;
    S.id = k < (int64_t)sids.size() ? sids[k] : k + 1;
    S.name = k < (int64_t)snames.size() ? snames[k] : "";
    if (e && s) {
      S.elem = f.ints(*e);
      S.side = f.ints(*s);
      if (S.elem.size() != S.side.size()) f.bad("side set " + std::to_string(S.id) + ": elem and side lengths differ");
      for (int64_t &q : S.elem) --q;
    }
    m.sidesets.push_back(std::move(S));
  }
  return m;
}

} // namespace mcwhost
