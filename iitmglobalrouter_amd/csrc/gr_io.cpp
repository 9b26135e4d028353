// Host side of the router's inputs and outputs: the design database (grid
// capacities, layers, named nets), the ISPD .cap / .net reader, a writer of
// those files, and the route (guide) writer.  No device work happens here.
//
// Reader semantics follow READ_DATA (read_data.hpp): the same tokenizer rules
// and the same fp32 digit accumulation, so capacities come out bit-identical;
// pins of a net are deduplicated and sorted by cell (remove_redundant_pins,
// :413-432) and nets left with one pin are set aside (remove_single_deg_nets,
// :366-412).  The parse runs in two phases: a sequential scan that moves over
// the text exactly as the tokenizer does, recording where every block of
// tokens starts, then a parallel conversion of the blocks.
//
// Writer semantics follow ROUTE_WRITER_UTIL (writer_files/writer_util.hpp):
// the per-edge path of write_route (:115-165), the segment merge of
// mrg_route_write (:166-205, :290-364), the pin patch (:234-289), and the
// two line formats (grid coordinates, or DBU 4200*g+2100 with metal names).
// Nets are written in schedule order, one-pin nets after the first batch
// (writer.hpp:38-45), each net as "name\n(\n<segments>)\n".
#include <algorithm>
#include <charconv>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/iitmgr.h"

namespace grh {
extern std::string g_err;
}

struct gr_design {
  int32_t X = 0, Y = 0, Z = 0;
  float unit_length_cost = 0.0f, unit_via_cost = 0.0f;
  std::vector<float> short_cost;    // per layer
  std::vector<float> header_extra;  // the four extra header numbers of the new format
  bool new_fmt = false;
  std::vector<int64_t> hlen, vlen;  // gcell lengths, X-1 and Y-1 values
  std::vector<std::string> layer_name;
  std::vector<int32_t> layer_dir, layer_min_len;
  std::vector<float> cap;  // X*Y*Z, WIRE_ENC order
  // every net in file order: raw name text and deduplicated pins
  std::vector<std::string> raw_name;
  std::vector<int64_t> all_pfx;
  std::vector<int32_t> all_pins;
  // multi-pin view (router net ids) and one-pin nets, each in file order
  std::vector<int32_t> multi, single;
  std::vector<int64_t> pin_pfx;
  std::vector<int32_t> pins;
};

namespace {

int fail(int code, const std::string& msg) {
  grh::g_err = msg;
  return code;
}

int pick_threads(int32_t t) {
  if (t > 0) return std::min(t, 256);
  unsigned h = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(h ? h : 1u, 16u));
}

template <class F>
void parallel_for(int threads, int64_t n, F f) {
  if (n <= 0) return;
  threads = (int)std::min<int64_t>(threads, n);
  if (threads <= 1) { f(0, (int64_t)0, n); return; }
  std::vector<std::thread> ts;
  for (int t = 0; t < threads; ++t) {
    const int64_t lo = n * t / threads, hi = n * (t + 1) / threads;
    ts.emplace_back([=]() { f(t, lo, hi); });
  }
  for (auto& th : ts) th.join();
}

inline int64_t enc_of(int64_t X, int64_t Y, int64_t z, int64_t x, int64_t y) {
  return (z & 1) ? z * X * Y + x * Y + y : z * X * Y + y * X + x;
}
inline void dec_of(int64_t X, int64_t Y, int64_t v, int& z, int& x, int& y) {
  z = (int)(v / (X * Y));
  const int64_t r = v - (int64_t)z * X * Y;
  if (z & 1) { x = (int)(r / Y); y = (int)(r % Y); }
  else       { y = (int)(r / X); x = (int)(r % X); }
}

bool load_file(const char* path, std::vector<char>& out) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  out.assign((size_t)std::max(0L, n) + 1, 0);  // trailing 0 ends every scan
  const size_t got = n > 0 ? std::fread(out.data(), 1, (size_t)n, f) : 0;
  std::fclose(f);
  return got == (size_t)std::max(0L, n);
}

inline bool dig(char c) { return c >= '0' && c <= '9'; }

// ---- tokenizer of READ_DATA (read_data.hpp:15-93) -------------------------
struct Text {
  const char* p;
  int64_t n;  // p[n] == 0
};

// get_long / get_gcd_dt: skip to a digit, read digits.
bool get_long(const Text& t, int64_t& i, int64_t& v) {
  v = 0;
  while (i < t.n && !dig(t.p[i])) ++i;
  if (i >= t.n) return false;
  while (dig(t.p[i])) {
    if (v > (INT64_MAX - 9) / 10) return false;
    v = v * 10 + (t.p[i++] - '0');
  }
  return true;
}

// get_string: skip ' ', '\n', ','; read up to the next of them.
bool get_string(const Text& t, int64_t& i, std::string& s) {
  while (i < t.n && (t.p[i] == ' ' || t.p[i] == '\n' || t.p[i] == ',')) ++i;
  const int64_t b = i;
  while (i < t.n && t.p[i] != ' ' && t.p[i] != '\n' && t.p[i] != ',') ++i;
  s.assign(t.p + b, t.p + i);
  return i > b;
}

// Moves over one number exactly as get_float/get_double do, without converting.
bool skip_float(const Text& t, int64_t& i) {
  while (i < t.n && t.p[i] != '-' && !dig(t.p[i])) ++i;
  if (i >= t.n) return false;
  if (t.p[i] == '-') ++i;
  while (dig(t.p[i])) ++i;
  if (t.p[i] == '.') { ++i; while (dig(t.p[i])) ++i; }
  if (t.p[i] == 'e') {
    ++i;
    if (t.p[i] == '-') ++i;
    if (t.p[i] == '+') ++i;
    while (dig(t.p[i])) ++i;
  }
  return true;
}

// get_float / get_double (identical for the fp32 gcb_dt): integer digits
// accumulate in double and round to fp32 per digit, fraction digits add
// digit * mul in fp32 with mul = 0.1f scaled by 0.1 (double) per digit, and
// each exponent step multiplies or divides by 10 in double.
bool get_float(const Text& t, int64_t& i, float& val) {
  val = 0.0f;
  while (i < t.n && t.p[i] != '-' && !dig(t.p[i])) ++i;
  if (i >= t.n) return false;
  bool neg = false;
  if (t.p[i] == '-') { neg = true; ++i; }
  while (dig(t.p[i])) {
    const char ch = t.p[i++];
    val = (float)((double)val * 10.0 + (double)ch - (double)'0');
  }
  if (t.p[i] == '.') {
    float mul = (float)0.1;
    ++i;
    while (dig(t.p[i])) {
      val = val + (float)(t.p[i++] - '0') * mul;
      mul = (float)((double)mul * 0.1);
    }
  }
  if (t.p[i] == 'e') {
    ++i;
    bool eneg = false;
    int pw = 0;
    if (t.p[i] == '-') { eneg = true; ++i; }
    if (t.p[i] == '+') ++i;
    while (dig(t.p[i])) {
      pw = pw * 10 + (t.p[i++] - '0');
      if (pw > 100000) return false;
    }
    for (int k = 0; k < pw; ++k) val = eneg ? (float)((double)val * 1.0 / 10.0) : (float)((double)val * 10.0);
  }
  val = neg ? -val : val;
  return true;
}

// ---- .cap (read_data.hpp:94-183) -----------------------------------------
int read_cap(const char* path, bool new_fmt, int threads, gr_design& d) {
  std::vector<char> buf;
  if (!load_file(path, buf)) return fail(GR_EINVAL, std::string("cannot read ") + path);
  const Text t{buf.data(), (int64_t)buf.size() - 1};
  int64_t i = 0, v = 0;
  int64_t dims[3];
  for (int k = 0; k < 3; ++k) {
    if (!get_long(t, i, v) || v > (1 << 20)) return fail(GR_EINVAL, ".cap: bad grid header");
    dims[k] = v;
  }
  // read_net_metadata: Z, X, Y in this order
  d.Z = (int32_t)dims[0]; d.X = (int32_t)dims[1]; d.Y = (int32_t)dims[2];
  if (d.X < 1 || d.Y < 1 || d.Z < 1) return fail(GR_EINVAL, ".cap: grid sizes must be positive");
  if ((int64_t)d.X * d.Y * d.Z >= (1ll << 31)) return fail(GR_ERANGE, ".cap: grid has 2^31 or more cells");
  d.new_fmt = new_fmt;
  if (!get_float(t, i, d.unit_length_cost) || !get_float(t, i, d.unit_via_cost))
    return fail(GR_EINVAL, ".cap: bad unit costs");
  d.short_cost.assign(d.Z, 0.0f);
  for (int z = 0; z < d.Z; ++z)
    if (!get_float(t, i, d.short_cost[z])) return fail(GR_EINVAL, ".cap: bad per-layer cost");
  d.header_extra.clear();
  if (new_fmt) {
    d.header_extra.assign(4, 0.0f);
    for (int k = 0; k < 4; ++k)
      if (!get_float(t, i, d.header_extra[k])) return fail(GR_EINVAL, ".cap: bad header");
  }
  d.hlen.assign(d.X - 1, 0);
  d.vlen.assign(d.Y - 1, 0);
  for (auto& h : d.hlen) if (!get_long(t, i, h)) return fail(GR_EINVAL, ".cap: bad horizontal gcell lengths");
  for (auto& h : d.vlen) if (!get_long(t, i, h)) return fail(GR_EINVAL, ".cap: bad vertical gcell lengths");
  // phase 1: layer headers in order; every layer's X*Y capacities are
  // skipped token by token and block starts are recorded every kBlk tokens
  constexpr int64_t kBlk = 4096;
  const int64_t per_layer = (int64_t)d.X * d.Y;
  const int64_t blocks_per_layer = (per_layer + kBlk - 1) / kBlk;
  std::vector<int64_t> starts((size_t)d.Z * blocks_per_layer);
  d.layer_name.assign(d.Z, "");
  d.layer_dir.assign(d.Z, 0);
  d.layer_min_len.assign(d.Z, 0);
  for (int z = 0; z < d.Z; ++z) {
    get_string(t, i, d.layer_name[z]);
    int64_t dir = 0, ml = 0;
    if (!get_long(t, i, dir) || !get_long(t, i, ml)) return fail(GR_EINVAL, ".cap: bad layer header");
    d.layer_dir[z] = (int32_t)dir;
    d.layer_min_len[z] = (int32_t)ml;
    for (int64_t k = 0; k < per_layer; ++k) {
      if (k % kBlk == 0) starts[(size_t)z * blocks_per_layer + k / kBlk] = i;
      if (!skip_float(t, i)) return fail(GR_EINVAL, ".cap: file ends inside the capacity data");
    }
  }
  // phase 2: convert the blocks in parallel; capacity k of layer z is at
  // (x, y) = (k % X, k / X) (read_layer_data: y outer, x inner)
  d.cap.assign((size_t)per_layer * d.Z, 0.0f);
  const int64_t nblk = (int64_t)starts.size();
  parallel_for(threads, nblk, [&](int, int64_t lo, int64_t hi) {
    for (int64_t b = lo; b < hi; ++b) {
      const int z = (int)(b / blocks_per_layer);
      const int64_t k0 = (b % blocks_per_layer) * kBlk, k1 = std::min(per_layer, k0 + kBlk);
      int64_t pos = starts[b];
      for (int64_t k = k0; k < k1; ++k) {
        float v2 = 0.0f;
        get_float(t, pos, v2);
        const int64_t x = k % d.X, y = k / d.X;
        d.cap[enc_of(d.X, d.Y, z, x, y)] = v2;
      }
    }
  });
  return 0;
}

// ---- .net (read_data.hpp:194-345) -----------------------------------------
// digits up to the next stop character (',' or ')'); the stop is consumed.
bool read_cd(const Text& t, int64_t& i, char stop, int32_t& out) {
  int64_t v = 0;
  int nd = 0;
  while (i < t.n && t.p[i] != stop) {
    if (dig(t.p[i])) {
      v = v * 10 + (t.p[i] - '0');
      ++nd;
      if (v > INT32_MAX) return false;  // stoi's range
    }
    ++i;
  }
  ++i;
  if (nd == 0) return false;
  out = (int32_t)v;
  return true;
}

// One net's pins from its span; mode 0: one access point per pin
// (read_ispd24_net_single, :314-333), 1: first access point of each pin
// (read_ispd24_net, :284-313), 2: new format (read_nt, :262-283).
bool parse_net(const Text& t, int64_t st, int64_t ed, int mode, std::vector<int32_t>& zxy) {
  zxy.clear();
  while (true) {
    if (mode == 2) {
      for (int c = 0; c < 2; ++c) {
        while (st < ed && t.p[st] != ',') ++st;
        ++st;
      }
      if (st >= ed) break;
      int32_t z, x, y;
      if (!read_cd(t, st, ',', z) || !read_cd(t, st, ',', x) || !read_cd(t, st, ')', y)) return false;
      zxy.insert(zxy.end(), {z, x, y});
      continue;
    }
    while (st < ed && t.p[st] != '(') ++st;
    ++st;
    if (st + 5 >= ed) break;
    bool first = true;
    while (true) {
      int32_t z, x, y;
      if (!read_cd(t, st, ',', z) || !read_cd(t, st, ',', x) || !read_cd(t, st, ')', y)) return false;
      if (first) zxy.insert(zxy.end(), {z, x, y});
      first = false;
      if (t.p[st] == ']') break;
      if (mode == 0) return false;  // a second access point
      ++st;
    }
  }
  return true;
}

int read_net(const char* path, int mode, int threads, gr_design& d) {
  std::vector<char> buf;
  if (!load_file(path, buf)) return fail(GR_EINVAL, std::string("cannot read ") + path);
  const Text t{buf.data(), (int64_t)buf.size() - 1};
  // get_num_nets_temp: a net spans a top-level '(' ... ')' pair
  std::vector<int64_t> open, close;
  int depth = 0;
  for (int64_t i = 0; i < t.n; ++i) {
    if (t.p[i] == '(') {
      if (depth == 0) open.push_back(i);
      ++depth;
    } else if (t.p[i] == ')') {
      if (--depth == 0) close.push_back(i);
      if (depth < 0) return fail(GR_EINVAL, ".net: unbalanced ')'");
    }
  }
  if (depth != 0) return fail(GR_EINVAL, ".net: unbalanced '('");
  const int64_t nn = (int64_t)close.size();
  if (nn >= INT32_MAX) return fail(GR_ERANGE, ".net: too many nets");
  std::vector<std::vector<int32_t>> cells(nn);
  d.raw_name.assign(nn, "");
  std::vector<int64_t> bad(threads, -1);
  parallel_for(threads, nn, [&](int tid, int64_t lo, int64_t hi) {
    std::vector<int32_t> zxy;
    for (int64_t k = lo; k < hi; ++k) {
      const int64_t st = k == 0 ? 0 : close[k - 1] + 1;
      d.raw_name[k].assign(t.p + st, t.p + open[k]);
      if (!parse_net(t, st, close[k], mode, zxy)) { if (bad[tid] < 0) bad[tid] = k; continue; }
      std::vector<int32_t>& c = cells[k];
      c.clear();
      for (size_t q = 0; q + 2 < zxy.size(); q += 3) {
        const int32_t z = zxy[q], x = zxy[q + 1], y = zxy[q + 2];
        if (z >= d.Z || x >= d.X || y >= d.Y) { if (bad[tid] < 0) bad[tid] = k; break; }
        c.push_back((int32_t)enc_of(d.X, d.Y, z, x, y));
      }
      // remove_redundant_pins: std::map by cell -> ascending, unique
      std::sort(c.begin(), c.end());
      c.erase(std::unique(c.begin(), c.end()), c.end());
    }
  });
  for (int64_t b : bad)
    if (b >= 0) return fail(GR_EINVAL, ".net: cannot parse the pins of net " + std::to_string(b) + " (or a pin is outside the grid)");
  d.all_pfx.assign(nn + 1, 0);
  for (int64_t k = 0; k < nn; ++k) {
    if (cells[k].empty()) return fail(GR_EINVAL, ".net: net " + std::to_string(k) + " has no pins");
    d.all_pfx[k + 1] = d.all_pfx[k] + (int64_t)cells[k].size();
  }
  d.all_pins.resize(d.all_pfx[nn]);
  parallel_for(threads, nn, [&](int, int64_t lo, int64_t hi) {
    for (int64_t k = lo; k < hi; ++k) std::copy(cells[k].begin(), cells[k].end(), d.all_pins.begin() + d.all_pfx[k]);
  });
  return 0;
}

// remove_single_deg_nets (read_data.hpp:366-412): one-pin nets leave the
// routed set; both keep file order.
int finish(gr_design& d) {
  const int64_t nn = (int64_t)d.all_pfx.size() - 1;
  d.multi.clear(); d.single.clear();
  d.pin_pfx.assign(1, 0);
  d.pins.clear();
  for (int64_t k = 0; k < nn; ++k) {
    const int64_t deg = d.all_pfx[k + 1] - d.all_pfx[k];
    if (deg == 1) { d.single.push_back((int32_t)k); continue; }
    if (deg > 32767) return fail(GR_ERANGE, "net " + std::to_string(k) + " has more than 32767 pins");
    d.multi.push_back((int32_t)k);
    d.pins.insert(d.pins.end(), d.all_pins.begin() + d.all_pfx[k], d.all_pins.begin() + d.all_pfx[k + 1]);
    d.pin_pfx.push_back((int64_t)d.pins.size());
  }
  return 0;
}

std::string clean_name(const std::string& raw) {
  // write_net_name (writer_util.hpp:9-13): drop '\n' and ' '
  std::string s;
  s.reserve(raw.size());
  for (char ch : raw)
    if (ch != '\n' && ch != ' ') s.push_back(ch);
  return s;
}

// ---- route writer ---------------------------------------------------------
struct Out {
  std::string s;
  void num(int64_t v) {
    char tmp[24];
    auto r = std::to_chars(tmp, tmp + sizeof tmp, v);
    s.append(tmp, r.ptr);
  }
  void sp() { s.push_back(' '); }
  void nl() { s.push_back('\n'); }
};

inline int64_t dbu(int64_t g) { return 4200 * g + 2100; }  // get_gran_cd, writer_util.hpp:17-19

// "x y metal(z+1) x y metal(z+2)" or "z x y z x y" (add_vias_cpu, :20-29)
void line_via(Out& o, bool D, int z, int x, int y) {
  if (D) {
    o.num(dbu(x)); o.sp(); o.num(dbu(y)); o.s += " metal"; o.num(z + 1); o.sp();
    o.num(dbu(x)); o.sp(); o.num(dbu(y)); o.s += " metal"; o.num(z + 2); o.nl();
  } else {
    o.num(z); o.sp(); o.num(x); o.sp(); o.num(y); o.sp(); o.num(z); o.sp(); o.num(x); o.sp(); o.num(y); o.nl();
  }
}
// Unmerged wire, add_horizontal_cpu / add_vertical_cpu (:30-53): in DBU form
// the second end carries metal(z+2), as there.
void line_wire(Out& o, bool D, int z, int xa, int ya, int xb, int yb) {
  if (D) {
    o.num(dbu(xa)); o.sp(); o.num(dbu(ya)); o.s += " metal"; o.num(z + 1); o.sp();
    o.num(dbu(xb)); o.sp(); o.num(dbu(yb)); o.s += " metal"; o.num(z + 2); o.nl();
  } else {
    o.num(z); o.sp(); o.num(xa); o.sp(); o.num(ya); o.sp(); o.num(z); o.sp(); o.num(xb); o.sp(); o.num(yb); o.nl();
  }
}
// Merged segment, add_segment (:206-220): a via point or a wire on one layer.
void line_seg(Out& o, bool D, int z1, int x1, int y1, int z2, int x2, int y2) {
  if (D) {
    const bool point = x1 == x2 && y1 == y2 && z1 == z2;
    o.num(dbu(x1)); o.sp(); o.num(dbu(y1)); o.s += " metal"; o.num(z1 + 1); o.sp();
    o.num(dbu(x2)); o.sp(); o.num(dbu(y2)); o.s += " metal"; o.num(point ? z1 + 2 : z1 + 1); o.nl();
  } else {
    o.num(z1); o.sp(); o.num(x1); o.sp(); o.num(y1); o.sp(); o.num(z2); o.sp(); o.num(x2); o.sp(); o.num(y2); o.nl();
  }
}

struct Seg { int64_t a, b; };  // minterval_t: cells a <= b; a == b is one via layer

struct Walker {
  const gr_design& d;
  // One tree edge's path (write_route, writer_util.hpp:115-165): via stack
  // to layer a, wire on a (layer 0 carries none), second via stack and wire
  // on b' only when both coordinates differ, last via stack to the pin.
  // via(x, y, za, zb) gets each via stack, wire(z, xa, ya, xb, yb) each wire
  // (possibly of zero length); false if the path does not end at dst.
  template <class V, class W>
  bool path(int32_t src, int32_t dst, int32_t p, V via, W wire) const {
    const int Z2 = (d.Z + 1) / 2;
    int z1, x1, y1, z2, x2, y2;
    dec_of(d.X, d.Y, src, z1, x1, y1);
    dec_of(d.X, d.Y, dst, z2, x2, y2);
    if (p < 0 || p >= d.Z * Z2) return false;
    const int a = p / Z2;
    int b = p % Z2;
    const bool both = x1 != x2 && y1 != y2;
    if (z1 != a) { via(x1, y1, z1, a); z1 = a; }
    auto run = [&](int z) {
      if (z & 1) { wire(z, x1, y1, x1, y2); y1 = y2; }
      else       { wire(z, x1, y1, x2, y1); x1 = x2; }
    };
    if (z1 > 0) run(z1);
    b = (z1 & 1) ? 2 * b : 2 * b + 1;
    if (both && b != 0 && z1 != b) {
      if (b >= d.Z) return false;
      via(x1, y1, z1, b);
      z1 = b;
      run(z1);
    }
    if (z1 != z2) { via(x1, y1, z1, z2); z1 = z2; }
    return x1 == x2 && y1 == y2;
  }
};

// Writes one routed net; false if an edge does not decode to a valid path.
bool write_net(const gr_design& d, int flags, int32_t net, int32_t local, const int64_t* edges,
               const int32_t* lpat, Out& o) {
  const bool D = flags & GR_WRITE_DBU;
  const int32_t file_id = d.multi[net];
  o.s += clean_name(d.raw_name[file_id]);
  o.s += "\n(\n";
  const int64_t p0 = d.pin_pfx[net], deg = d.pin_pfx[net + 1] - p0;
  const int32_t* pins = d.pins.data() + p0;
  Walker w{d};
  std::vector<Seg> segs;
  for (int64_t j = 0; j < deg; ++j) {
    const int64_t e = edges[j];
    if (e == -1) continue;
    int n1 = (int)(e & 0x7fff), n2 = (int)((e >> 15) & 0x7fff);
    if ((int32_t)(e >> 30) != local || n1 >= deg || n2 >= deg) return false;
    if (n1 < n2) std::swap(n1, n2);
    bool ok;
    if (flags & GR_WRITE_MERGE) {
      ok = w.path(pins[n1], pins[n2], lpat[j],
                  [&](int x, int y, int za, int zb) {  // add_vias_cpu_vec
                    if (za > zb) std::swap(za, zb);
                    for (int l = za; l < zb; ++l) {
                      const int64_t c = enc_of(d.X, d.Y, l, x, y);
                      segs.push_back({c, c});
                    }
                  },
                  [&](int z, int xa, int ya, int xb, int yb) {  // add_{horizontal,vertical}_cpu_vec
                    if (xa == xb && ya == yb) return;
                    int64_t c1 = enc_of(d.X, d.Y, z, xa, ya), c2 = enc_of(d.X, d.Y, z, xb, yb);
                    if (c1 > c2) std::swap(c1, c2);
                    segs.push_back({c1, c2});
                  });
    } else {
      ok = w.path(pins[n1], pins[n2], lpat[j],
                  [&](int x, int y, int za, int zb) {
                    if (za > zb) std::swap(za, zb);
                    for (int l = za; l < zb; ++l) line_via(o, D, l, x, y);
                  },
                  [&](int z, int xa, int ya, int xb, int yb) {
                    if (xa == xb && ya == yb) return;
                    if (xa > xb) std::swap(xa, xb);
                    if (ya > yb) std::swap(ya, yb);
                    line_wire(o, D, z, xa, ya, xb, yb);
                  });
    }
    if (!ok) return false;
  }
  if (flags & GR_WRITE_MERGE) {
    // merge_st_edges_nnew (:166-205): wires sorted and fused when they touch,
    // then the distinct via cells in ascending order
    std::vector<Seg> wires;
    std::vector<int64_t> pts;
    for (const Seg& s : segs) (s.a == s.b ? pts.push_back(s.a) : wires.push_back(s));
    std::sort(wires.begin(), wires.end(), [](const Seg& u, const Seg& v) { return u.a != v.a ? u.a < v.a : u.b < v.b; });
    std::vector<Seg> merged;
    for (const Seg& s : wires) {
      if (!merged.empty() && merged.back().b >= s.a) merged.back().b = std::max(merged.back().b, s.b);
      else merged.push_back(s);
    }
    std::sort(pts.begin(), pts.end());
    pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
    std::vector<Seg> order;
    if (flags & GR_WRITE_PATCH) {
      // oprd_ppatch (:234-289): a pin above layer 0 that is no segment end
      // splits the wire that contains it; via cells first, then wires in order
      auto cmp = [](const Seg& u, const Seg& v) { return u.a != v.a ? u.a < v.a : u.b < v.b; };
      std::set<Seg, decltype(cmp)> cst(cmp);
      std::set<int64_t> ends;
      for (int64_t c : pts) {
        int z, x, y;
        dec_of(d.X, d.Y, c, z, x, y);
        ends.insert(c);
        ends.insert(enc_of(d.X, d.Y, z + 1, x, y));
      }
      for (const Seg& s : merged) { ends.insert(s.a); ends.insert(s.b); cst.insert(s); }
      for (int64_t k = 0; k < deg; ++k) {
        const int64_t c = pins[k];
        if (c < (int64_t)d.X * d.Y || ends.count(c)) continue;
        auto it = cst.upper_bound(Seg{c, c});
        if (it == cst.begin()) return false;
        --it;
        const Seg s = *it;
        if (!(s.a < c && c < s.b)) return false;
        cst.erase(it);
        cst.insert(Seg{s.a, c});
        cst.insert(Seg{c, s.b});
      }
      for (int64_t c : pts) order.push_back({c, c});
      order.insert(order.end(), cst.begin(), cst.end());
    } else {
      order = merged;
      for (int64_t c : pts) order.push_back({c, c});
    }
    for (const Seg& s : order) {
      int z1, x1, y1, z2, x2, y2;
      dec_of(d.X, d.Y, s.a, z1, x1, y1);
      dec_of(d.X, d.Y, s.b, z2, x2, y2);
      line_seg(o, D, z1, x1, y1, z2, x2, y2);
    }
  }
  o.s += ")\n";
  return true;
}

// write_odeg_route (writer_util.hpp:101-114): one via up from the pin.
void write_single(const gr_design& d, int flags, int32_t k, Out& o) {
  const int32_t file_id = d.single[k];
  o.s += clean_name(d.raw_name[file_id]);
  o.s += "\n(\n";
  int z, x, y;
  dec_of(d.X, d.Y, d.all_pins[d.all_pfx[file_id]], z, x, y);
  line_via(o, flags & GR_WRITE_DBU, z, x, y);
  o.s += ")\n";
}

}  // namespace

extern "C" {

int gr_design_read(const char* cap_path, const char* net_path, int32_t flags, int32_t threads, gr_design** out) {
  if (!cap_path || !net_path || !out) return fail(GR_EINVAL, "null argument");
  *out = nullptr;
  const int mode = (flags & GR_FMT_NEW) ? 2 : ((flags & GR_FMT_FIRST_ACCESS_POINT) ? 1 : 0);
  gr_design* d = new gr_design();
  const int nt = pick_threads(threads);
  int rc = read_cap(cap_path, flags & GR_FMT_NEW, nt, *d);
  if (!rc) rc = read_net(net_path, mode, nt, *d);
  if (!rc) rc = finish(*d);
  if (rc) { delete d; return rc; }
  *out = d;
  return 0;
}

int gr_design_create(int32_t X, int32_t Y, int32_t Z, const float* cap, int32_t n_nets, const int64_t* pin_pfx,
                     const int32_t* pin_enc, gr_design** out) {
  if (!out || X < 1 || Y < 1 || Z < 1 || n_nets < 0 || !pin_pfx || (n_nets && !pin_enc))
    return fail(GR_EINVAL, "bad design arrays");
  *out = nullptr;
  const int64_t cs = (int64_t)X * Y * Z;
  if (cs >= (1ll << 31)) return fail(GR_ERANGE, "grid has 2^31 or more cells");
  if (pin_pfx[0] != 0) return fail(GR_EINVAL, "pin_pfx[0] must be 0");
  gr_design* d = new gr_design();
  d->X = X; d->Y = Y; d->Z = Z;
  d->cap.assign(cs, 0.0f);
  if (cap) std::copy(cap, cap + cs, d->cap.begin());
  d->short_cost.assign(Z, 0.0f);
  d->hlen.assign(X - 1, 1);
  d->vlen.assign(Y - 1, 1);
  d->layer_name.resize(Z);
  d->layer_dir.resize(Z);
  d->layer_min_len.assign(Z, 0);
  for (int z = 0; z < Z; ++z) { d->layer_name[z] = "metal" + std::to_string(z + 1); d->layer_dir[z] = z & 1; }
  d->raw_name.resize(n_nets);
  d->all_pfx.assign(n_nets + 1, 0);
  for (int32_t k = 0; k < n_nets; ++k) {
    if (pin_pfx[k + 1] <= pin_pfx[k]) { delete d; return fail(GR_EINVAL, "every net needs a pin"); }
    std::vector<int32_t> c(pin_enc + pin_pfx[k], pin_enc + pin_pfx[k + 1]);
    for (int32_t v : c)
      if (v < 0 || v >= cs) { delete d; return fail(GR_EINVAL, "pin outside the grid"); }
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
    d->all_pins.insert(d->all_pins.end(), c.begin(), c.end());
    d->all_pfx[k + 1] = (int64_t)d->all_pins.size();
    d->raw_name[k] = "net" + std::to_string(k);
  }
  const int rc = finish(*d);
  if (rc) { delete d; return rc; }
  *out = d;
  return 0;
}

void gr_design_free(gr_design* d) { delete d; }

int gr_design_info(const gr_design* d, gr_design_desc* info) {
  if (!d || !info) return fail(GR_EINVAL, "null argument");
  info->X = d->X; info->Y = d->Y; info->Z = d->Z;
  info->n_nets = (int32_t)d->multi.size();
  info->n_pins = (int64_t)d->pins.size();
  info->n_single = (int32_t)d->single.size();
  info->unit_length_cost = d->unit_length_cost;
  info->unit_via_cost = d->unit_via_cost;
  return 0;
}

int gr_design_capacity(const gr_design* d, float* cap) {
  if (!d || !cap) return fail(GR_EINVAL, "null argument");
  std::copy(d->cap.begin(), d->cap.end(), cap);
  return 0;
}

int gr_design_layers(const gr_design* d, int32_t* dir, int32_t* min_len, float* unit_short_cost) {
  if (!d) return fail(GR_EINVAL, "null argument");
  for (int z = 0; z < d->Z; ++z) {
    if (dir) dir[z] = d->layer_dir[z];
    if (min_len) min_len[z] = d->layer_min_len[z];
    if (unit_short_cost) unit_short_cost[z] = d->short_cost[z];
  }
  return 0;
}

int gr_design_layer_name(const gr_design* d, int32_t z, char* buf, int32_t size) {
  if (!d || z < 0 || z >= d->Z || !buf || size < 1) return fail(GR_EINVAL, "bad layer name request");
  const std::string& s = d->layer_name[z];
  const int32_t n = (int32_t)std::min<size_t>(s.size(), (size_t)size - 1);
  std::memcpy(buf, s.data(), n);
  buf[n] = 0;
  return (int32_t)s.size();
}

int gr_design_gcell_lengths(const gr_design* d, int64_t* hlen, int64_t* vlen) {
  if (!d) return fail(GR_EINVAL, "null argument");
  if (hlen) std::copy(d->hlen.begin(), d->hlen.end(), hlen);
  if (vlen) std::copy(d->vlen.begin(), d->vlen.end(), vlen);
  return 0;
}

int gr_design_nets(const gr_design* d, int64_t* pin_pfx, int32_t* pin_enc) {
  if (!d) return fail(GR_EINVAL, "null argument");
  if (pin_pfx) std::copy(d->pin_pfx.begin(), d->pin_pfx.end(), pin_pfx);
  if (pin_enc) std::copy(d->pins.begin(), d->pins.end(), pin_enc);
  return 0;
}

int gr_design_single_nets(const gr_design* d, int32_t* pin_enc) {
  if (!d) return fail(GR_EINVAL, "null argument");
  if (pin_enc)
    for (size_t k = 0; k < d->single.size(); ++k) pin_enc[k] = d->all_pins[d->all_pfx[d->single[k]]];
  return 0;
}

int gr_design_net_name(const gr_design* d, int32_t net, char* buf, int32_t size) {
  if (!d || !buf || size < 1) return fail(GR_EINVAL, "bad name request");
  const int64_t nm = (int64_t)d->multi.size(), ns = (int64_t)d->single.size();
  if (net < 0 || net >= nm + ns) return fail(GR_EINVAL, "net id out of range");
  const std::string s = clean_name(d->raw_name[net < nm ? d->multi[net] : d->single[net - nm]]);
  const int32_t n = (int32_t)std::min<size_t>(s.size(), (size_t)size - 1);
  std::memcpy(buf, s.data(), n);
  buf[n] = 0;
  return (int32_t)s.size();
}

int gr_design_save(const gr_design* d, const char* cap_path, const char* net_path) {
  if (!d || (!cap_path && !net_path)) return fail(GR_EINVAL, "null argument");
  if (cap_path) {
    FILE* f = std::fopen(cap_path, "wb");
    if (!f) return fail(GR_EINVAL, std::string("cannot write ") + cap_path);
    Out o;
    o.num(d->Z); o.sp(); o.num(d->X); o.sp(); o.num(d->Y); o.nl();
    auto flt = [&](float v) {
      char tmp[48];
      auto r = std::to_chars(tmp, tmp + sizeof tmp, v);  // shortest round-trip text
      o.s.append(tmp, r.ptr);
    };
    flt(d->unit_length_cost); o.sp(); flt(d->unit_via_cost); o.nl();
    for (int z = 0; z < d->Z; ++z) { if (z) o.sp(); flt(d->short_cost[z]); }
    o.nl();
    if (d->new_fmt) { for (int k = 0; k < 4; ++k) { if (k) o.sp(); flt(d->header_extra[k]); } o.nl(); }
    for (size_t k = 0; k < d->hlen.size(); ++k) { if (k) o.sp(); o.num(d->hlen[k]); }
    o.nl();
    for (size_t k = 0; k < d->vlen.size(); ++k) { if (k) o.sp(); o.num(d->vlen[k]); }
    o.nl();
    bool ok = true;
    for (int z = 0; z < d->Z && ok; ++z) {
      o.s += d->layer_name[z]; o.sp(); o.num(d->layer_dir[z]); o.sp(); o.num(d->layer_min_len[z]); o.nl();
      for (int y = 0; y < d->Y; ++y) {
        for (int x = 0; x < d->X; ++x) { if (x) o.sp(); flt(d->cap[enc_of(d->X, d->Y, z, x, y)]); }
        o.nl();
      }
      ok = std::fwrite(o.s.data(), 1, o.s.size(), f) == o.s.size();
      o.s.clear();
    }
    ok = ok && std::fwrite(o.s.data(), 1, o.s.size(), f) == o.s.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(GR_EINVAL, std::string("write failed: ") + cap_path);
  }
  if (net_path) {
    // write_ispd24_nets (read_data.hpp:349-359): "name(" then one
    // "[(z, x, y)]" line per pin, then ")"; pins as deduplicated here
    FILE* f = std::fopen(net_path, "wb");
    if (!f) return fail(GR_EINVAL, std::string("cannot write ") + net_path);
    Out o;
    bool ok = true;
    const int64_t nn = (int64_t)d->raw_name.size();
    for (int64_t k = 0; k < nn && ok; ++k) {
      o.s += d->raw_name[k];
      o.s += "(\n";
      for (int64_t q = d->all_pfx[k]; q < d->all_pfx[k + 1]; ++q) {
        int z, x, y;
        dec_of(d->X, d->Y, d->all_pins[q], z, x, y);
        o.s += "[("; o.num(z); o.s += ", "; o.num(x); o.s += ", "; o.num(y); o.s += ")]\n";
      }
      o.s += ")";
      if (o.s.size() > (1u << 22)) { ok = std::fwrite(o.s.data(), 1, o.s.size(), f) == o.s.size(); o.s.clear(); }
    }
    ok = ok && std::fwrite(o.s.data(), 1, o.s.size(), f) == o.s.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(GR_EINVAL, std::string("write failed: ") + net_path);
  }
  return 0;
}

int gr_write_routes(const gr_design* d, const char* path, int32_t flags, int32_t n_batches, const int32_t* batch_ofst,
                    const int32_t* net_ids, const int64_t* edges, const int32_t* lpat, int32_t threads,
                    int64_t* bytes) {
  if (!d || !path || n_batches < 0 || !batch_ofst || (batch_ofst[n_batches] && (!net_ids || !edges || !lpat)))
    return fail(GR_EINVAL, "null argument");
  if (batch_ofst[0] != 0) return fail(GR_EINVAL, "batch_ofst[0] must be 0");
  const int32_t npos = batch_ofst[n_batches];
  const int32_t nm = (int32_t)d->multi.size();
  std::vector<int64_t> slot(npos + 1, 0);
  std::vector<int32_t> local(npos);
  for (int32_t b = 0; b < n_batches; ++b) {
    if (batch_ofst[b + 1] < batch_ofst[b]) return fail(GR_EINVAL, "batch offsets must be non-decreasing");
    for (int32_t q = batch_ofst[b]; q < batch_ofst[b + 1]; ++q) {
      if (net_ids[q] < 0 || net_ids[q] >= nm) return fail(GR_EINVAL, "net id out of range");
      local[q] = q - batch_ofst[b];
    }
  }
  for (int32_t q = 0; q < npos; ++q) slot[q + 1] = slot[q] + (d->pin_pfx[net_ids[q] + 1] - d->pin_pfx[net_ids[q]]);
  // each thread formats a contiguous range of units (routed nets in schedule
  // order, the one-pin nets placed after the first batch); the texts are
  // then written in unit order
  const int32_t ns = (int32_t)d->single.size();
  const int32_t split = n_batches > 0 ? batch_ofst[1] : 0;
  const int64_t units = (int64_t)npos + ns;
  const int nt = pick_threads(threads);
  const int64_t nchunks = std::min<int64_t>(units, (int64_t)nt * 8);
  std::vector<Out> outs(std::max<int64_t>(nchunks, 1));
  std::vector<int64_t> bad(outs.size(), -1);
  parallel_for(nt, nchunks, [&](int, int64_t lo, int64_t hi) {
    for (int64_t ch = lo; ch < hi; ++ch) {
      const int64_t u0 = units * ch / nchunks, u1 = units * (ch + 1) / nchunks;
      Out& o = outs[ch];
      for (int64_t u = u0; u < u1; ++u) {
        if (u >= split && u < split + ns) { write_single(*d, flags, (int32_t)(u - split), o); continue; }
        const int32_t q = (int32_t)(u < split ? u : u - ns);
        if (!write_net(*d, flags, net_ids[q], local[q], edges + slot[q], lpat + slot[q], o)) {
          bad[ch] = q;
          break;
        }
      }
    }
  });
  for (int64_t b : bad)
    if (b >= 0) return fail(GR_EINVAL, "edges of schedule position " + std::to_string(b) + " do not form valid paths");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(GR_EINVAL, std::string("cannot write ") + path);
  int64_t total = 0;
  bool ok = true;
  for (const Out& o : outs) {
    ok = ok && std::fwrite(o.s.data(), 1, o.s.size(), f) == o.s.size();
    total += (int64_t)o.s.size();
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(GR_EINVAL, std::string("write failed: ") + path);
  if (bytes) *bytes = total;
  return 0;
}

}  // extern "C"
