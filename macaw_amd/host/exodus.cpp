// exodus.cpp -- RayTracingExodus-layout writer (see exodus.h).
//
// NetCDF classic, 64-bit offset variant ("CDF\2"): a header of dimensions,
// global attributes and variables (name, dimension ids, attributes, type,
// size, file offset), then the fixed-size variables' data, then the record
// variables' data record by record.  Everything is big-endian and padded to
// 4 bytes.
#include "exodus.h"

#include <cmath>
#include <cstring>
#include <fstream>
#include <stdexcept>

namespace mcwhost {

uint64_t init_seed(uint64_t id, uint64_t stream, uint64_t master_seed)
{
  // Brown (1994) arbitrary-stride skip-ahead of the LCG s <- M s + C, over
  // id * 152917 steps from master_seed + stream (OpenMC's future_seed)
  const uint64_t M = 6364136223846793005ull, C = 1442695040888963407ull;
  uint64_t n = id * 152917ull, g = M, c = C, g_new = 1, c_new = 0;
  while (n) {
    if (n & 1) {
      g_new *= g;
      c_new = c_new * g + c;
    }
    c = (g + 1) * c;
    g *= g;
    n >>= 1;
  }
  return g_new * (master_seed + stream) + c_new;
}

namespace {

enum { NC_CHAR = 2, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6 };
enum { TAG_DIM = 10, TAG_VAR = 11, TAG_ATT = 12 };

struct Att {
  std::string name;
  int type;
  std::vector<uint8_t> bytes; // big-endian values
  int64_t n;
};

struct Var {
  std::string name;
  std::vector<int> dims;
  std::vector<Att> atts;
  int type;
  std::vector<uint8_t> data; // big-endian, one record for record variables
  bool record = false;
  int64_t begin = 0;
};

void be32(std::vector<uint8_t> &b, uint32_t v)
{
  for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
}
void be64(std::vector<uint8_t> &b, uint64_t v)
{
  for (int s = 56; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
}
void pad4(std::vector<uint8_t> &b)
{
  while (b.size() % 4) b.push_back(0);
}
void name(std::vector<uint8_t> &b, const std::string &s)
{
  be32(b, (uint32_t)s.size());
  b.insert(b.end(), s.begin(), s.end());
  pad4(b);
}

Att att_text(const std::string &n, const std::string &v)
{
  return Att{n, NC_CHAR, std::vector<uint8_t>(v.begin(), v.end()), (int64_t)v.size()};
}
Att att_int(const std::string &n, int32_t v)
{
  Att a{n, NC_INT, {}, 1};
  be32(a.bytes, (uint32_t)v);
  return a;
}
Att att_float(const std::string &n, float v)
{
  Att a{n, NC_FLOAT, {}, 1};
  uint32_t u;
  memcpy(&u, &v, 4);
  be32(a.bytes, u);
  return a;
}

std::vector<uint8_t> doubles(const std::vector<double> &v)
{
  std::vector<uint8_t> b;
  b.reserve(v.size() * 8);
  for (double x : v) {
    uint64_t u;
    memcpy(&u, &x, 8);
    be64(b, u);
  }
  return b;
}
std::vector<uint8_t> ints(const std::vector<int32_t> &v)
{
  std::vector<uint8_t> b;
  b.reserve(v.size() * 4);
  for (int32_t x : v) be32(b, (uint32_t)x);
  return b;
}
std::vector<uint8_t> chars(const std::vector<std::string> &rows, size_t width)
{
  std::vector<uint8_t> b;
  for (auto &r : rows) {
    std::string s = r.substr(0, width - 1);
    s.resize(width, '\0');
    b.insert(b.end(), s.begin(), s.end());
  }
  return b;
}

void put_atts(std::vector<uint8_t> &h, const std::vector<Att> &atts)
{
  if (atts.empty()) {
    be32(h, 0);
    be32(h, 0);
    return;
  }
  be32(h, TAG_ATT);
  be32(h, (uint32_t)atts.size());
  for (auto &a : atts) {
    name(h, a.name);
    be32(h, (uint32_t)a.type);
    be32(h, (uint32_t)a.n);
    h.insert(h.end(), a.bytes.begin(), a.bytes.end());
    pad4(h);
  }
}

} // namespace

void write_rays_exodus(const std::string &path, const std::vector<mcw_trace_record> &recs, int64_t batch,
                       int64_t n_particles, uint64_t master_seed, const std::string &title)
{
  const int64_t ne = (int64_t)recs.size();
  // nodes: every visit's start, plus the end of each ray's last visit
  std::vector<double> X, Y, Z;
  std::vector<int32_t> conn;
  int64_t n_rays = 0;
  for (int64_t i = 0; i < ne; ++i) {
    const mcw_trace_record &r = recs[i];
    const bool first = i == 0 || recs[i - 1].history != r.history;
    const bool last = i + 1 == ne || recs[i + 1].history != r.history;
    n_rays += first;
    X.push_back(r.start[0]);
    Y.push_back(r.start[1]);
    Z.push_back(r.start[2]);
    const int32_t node = (int32_t)X.size(); // 1-based id of this visit's start
    conn.push_back(node);
    conn.push_back(node + 1);
    if (last) {
      X.push_back(r.end[0]);
      Y.push_back(r.end[1]);
      Z.push_back(r.end[2]);
    }
  }
  const int64_t nn = (int64_t)X.size();
  static const char *vars[10] = {"aux_energy", "aux_id", "aux_n_progeny", "aux_particle_type", "aux_seed_URR",
                                 "aux_seed_source", "aux_seed_tracking", "aux_weight", "intersections", "ray_id"};
  std::vector<std::vector<double>> vals(10, std::vector<double>((size_t)ne));
  for (int64_t i = 0; i < ne; ++i) {
    const mcw_trace_record &r = recs[i];
    const uint64_t pid = (uint64_t)((batch - 1) * n_particles + r.history);
    vals[0][i] = r.energy;
    vals[1][i] = (double)(r.history + 1);
    vals[2][i] = 0.0;
    vals[3][i] = 0.0;
    vals[4][i] = (double)init_seed(pid + 1, 2, master_seed);
    vals[5][i] = (double)init_seed(pid + 1, 1, master_seed);
    vals[6][i] = (double)r.seed_tracking;
    vals[7][i] = 1.0;
    vals[8][i] = (double)r.visit;
    vals[9][i] = (double)pid;
  }

  // dimensions (the record dimension first)
  const std::vector<std::pair<std::string, int64_t>> dims = {
      {"time_step", 0}, {"len_string", 33}, {"len_line", 81}, {"four", 4}, {"len_name", 33}, {"num_dim", 3},
      {"num_nodes", nn}, {"num_elem", ne}, {"num_el_blk", 1}, {"num_el_in_blk1", ne}, {"num_nod_per_el1", 2},
      {"num_elem_var", 10}};
  auto dim = [&](const char *n) {
    for (size_t k = 0; k < dims.size(); ++k)
      if (dims[k].first == n) return (int)k;
    throw std::runtime_error(std::string("no dimension ") + n);
  };
  std::vector<Var> vs;
  auto add = [&](const std::string &n, std::vector<int> d, int type, std::vector<uint8_t> data, bool rec = false,
                 std::vector<Att> atts = {}) {
    vs.push_back(Var{n, d, atts, type, std::move(data), rec, 0});
  };
  add("time_whole", {dim("time_step")}, NC_DOUBLE, doubles({(double)batch}), true);
  add("eb_status", {dim("num_el_blk")}, NC_INT, ints({1}));
  add("eb_prop1", {dim("num_el_blk")}, NC_INT, ints({0}), false, {att_text("name", "ID")});
  add("coordx", {dim("num_nodes")}, NC_DOUBLE, doubles(X));
  add("coordy", {dim("num_nodes")}, NC_DOUBLE, doubles(Y));
  add("coordz", {dim("num_nodes")}, NC_DOUBLE, doubles(Z));
  add("eb_names", {dim("num_el_blk"), dim("len_name")}, NC_CHAR, chars({""}, 33));
  add("coor_names", {dim("num_dim"), dim("len_name")}, NC_CHAR, chars({"", "", ""}, 33));
  std::vector<int32_t> seq((size_t)std::max(nn, ne));
  for (size_t k = 0; k < seq.size(); ++k) seq[k] = (int32_t)k + 1;
  add("node_num_map", {dim("num_nodes")}, NC_INT, ints(std::vector<int32_t>(seq.begin(), seq.begin() + nn)));
  add("connect1", {dim("num_el_in_blk1"), dim("num_nod_per_el1")}, NC_INT, ints(conn), false,
      {att_text("elem_type", "EDGE2")});
  add("elem_num_map", {dim("num_elem")}, NC_INT, ints(std::vector<int32_t>(seq.begin(), seq.begin() + ne)));
  add("name_elem_var", {dim("num_elem_var"), dim("len_name")}, NC_CHAR,
      chars(std::vector<std::string>(vars, vars + 10), 33));
  for (int k = 0; k < 10; ++k)
    add("vals_elem_var" + std::to_string(k + 1) + "eb1", {dim("time_step"), dim("num_el_in_blk1")}, NC_DOUBLE,
        doubles(vals[k]), true);
  add("elem_var_tab", {dim("num_el_blk"), dim("num_elem_var")}, NC_INT, ints(std::vector<int32_t>(10, 1)));

  const std::vector<Att> gatts = {att_float("api_version", 5.22f), att_float("version", 5.22f),
                                  att_int("floating_point_word_size", 8), att_int("file_size", 1),
                                  att_int("int64_status", 0), att_text("title", title),
                                  att_int("maximum_name_length", 32)};

  // header, with the offsets laid out after it
  auto header = [&]() {
    std::vector<uint8_t> h = {'C', 'D', 'F', 2};
    be32(h, 1); // one record (time step)
    be32(h, TAG_DIM);
    be32(h, (uint32_t)dims.size());
    for (auto &d : dims) {
      name(h, d.first);
      be32(h, (uint32_t)d.second);
    }
    put_atts(h, gatts);
    be32(h, TAG_VAR);
    be32(h, (uint32_t)vs.size());
    for (auto &v : vs) {
      name(h, v.name);
      be32(h, (uint32_t)v.dims.size());
      for (int d : v.dims) be32(h, (uint32_t)d);
      put_atts(h, v.atts);
      be32(h, (uint32_t)v.type);
      be32(h, (uint32_t)((v.data.size() + 3) / 4 * 4));
      be64(h, (uint64_t)v.begin);
    }
    return h;
  };
  const int64_t hsize = (int64_t)header().size();
  int64_t off = hsize;
  for (auto &v : vs)
    if (!v.record) {
      v.begin = off;
      off += (int64_t)((v.data.size() + 3) / 4 * 4);
    }
  for (auto &v : vs)
    if (v.record) {
      v.begin = off;
      off += (int64_t)((v.data.size() + 3) / 4 * 4);
    }
  std::vector<uint8_t> out = header();
  for (int pass = 0; pass < 2; ++pass)
    for (auto &v : vs)
      if (v.record == (pass == 1)) {
        out.insert(out.end(), v.data.begin(), v.data.end());
        pad4(out);
      }
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + path);
  f.write((const char *)out.data(), (std::streamsize)out.size());
  if (!f) throw std::runtime_error("cannot write " + path);
}

} // namespace mcwhost
