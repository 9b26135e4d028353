// CPU oracle for the per-batch routing hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is a plain C++ restatement of the
// reference router's per-batch algorithm, used by tests/ as the checker, by
// __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline leg
// as the timed CPU baseline.  The product library (iitmglobalrouter_amd/) never
// links, loads or calls it.
//
// What it restates (reference files are cited file:line):
//   * grid encoding                      util.hpp:37-68
//   * cost of a cell f(r)                bit_util.hpp:13-18, util.hpp:119-132
//   * Fenwick run trees + range query    bit_util.hpp:19-55, 91-117
//   * L-pattern enumeration/selection    bit_util.hpp:118-191 (CPU twin order:
//                                        lpattern.hpp:98-160, lowest p wins)
//   * pattern code p = a*Z2 + b          lpattern_util.hpp:43-53
//   * pair order o = j(j-1)/2 + k        test_lpat.hpp:311-331, gmst_util.hpp:192-194
//   * root-pair Boruvka                  gmst_util.hpp:164-341, mst_util.hpp:289-314
//     (true-MST mode records the minimum pairs: the tree of the CPU twin
//     MST_UTIL::t_get_mst, mst_util.hpp:34-102)
//   * BIT commit cell marking            mst_bit_commit.hpp:17-67, 81-235
//   * per-batch demand update            bit_util.hpp:75-90
//
// Canonical arithmetic (the reference is nondeterministic here; these choices
// make every output reproducible and are shared with the HIP path):
//   * cell cost c = LUT[clamp(cap - dem, -256, 255)], LUT entry = fp32 rounding
//     of the model evaluated in double, saturated at 1e30; with fp32
//     capacities c = cost_f32(model, r) with r lowered by 1.0f per commit;
//   * Fenwick node covering an aligned block = pairwise fp32 sum of its cells,
//     recomputed from integer demand after every batch (the reference builds
//     and updates nodes with float atomicAdd, bit_util.hpp:44,51);
//   * prefix query Q(x) accumulates nodes in the reference's descent order
//     (res = 0; res += node[i]; i -= i & -i), bit_util.hpp:19-35;
//   * ties: lowest pattern p, lowest pair index o, candidates in ascending root.
//
// Pinning of each function is listed in tests/test_oracle.py (known-answer
// tests, brute-force cell walks in double, exhaustive small cases).
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace {

constexpr int kRMin = -256;
constexpr int kRMax = 255;
constexpr int kLut = kRMax - kRMin + 1;
constexpr float kSat = 1e30f;

// Cost models: 0 = reference BIT f(r) = exp(-r/2) (bit_util.hpp:13-15),
// 1 = reference incremental f(r-1) - f(r) (util.hpp:127-129),
// 2 = literal exp(d + 1 - c) = exp(1 - r) (README.md:10).
void build_lut(int model, float* lut) {
  for (int i = 0; i < kLut; ++i) {
    int r = i + kRMin;
    float v;
    if (model == 1) {
      float f0 = (float)std::exp(-0.5 * (double)(r - 1));
      float f1 = (float)std::exp(-0.5 * (double)r);
      v = f0 - f1;
    } else if (model == 2) {
      v = (float)std::exp(1.0 - (double)r);
    } else {
      v = (float)std::exp(-0.5 * (double)r);
    }
    if (!(v < kSat)) v = kSat;
    lut[i] = v;
  }
}

// Grids with fp32 capacities: e^x by Cody-Waite reduction x = k ln2 + t
// (Cody and Waite, Software Manual for the Elementary Functions, 1980) and the
// degree-13 Taylor polynomial of e^t in Horner form with fused multiply-adds,
// scaled by 2^k; rounded once to fp32 by the caller.  The cost functions are
// those of the table above with r in fp32 (util.hpp:127-129: r - 1 in fp32).
double cw_exp(double x) {
  if (x > 100.0) return HUGE_VAL;
  if (x < -110.0) return 0.0;
  const double k = std::rint(x * 1.4426950408889634);
  double t = std::fma(-k, 0.6931471805599453, x);
  t = std::fma(-k, 2.3190468138462996e-17, t);
  static const double inv_fact[14] = {1.0, 1.0, 1.0 / 2, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720, 1.0 / 5040,
                                      1.0 / 40320, 1.0 / 362880, 1.0 / 3628800, 1.0 / 39916800,
                                      1.0 / 479001600, 1.0 / 6227020800.0};
  double p = inv_fact[13];
  for (int n = 12; n >= 0; --n) p = std::fma(p, t, inv_fact[n]);
  return std::ldexp(p, (int)k);
}

float cost_f32(int model, float r) {
  float v;
  if (model == 1) {
    float r1 = r - 1.0f;
    v = (float)cw_exp(-0.5 * (double)r1) - (float)cw_exp(-0.5 * (double)r);
  } else if (model == 2) {
    v = (float)cw_exp(1.0 - (double)r);
  } else {
    v = (float)cw_exp(-0.5 * (double)r);
  }
  if (!(v < kSat)) v = kSat;
  return v;
}

// Increment of the base cost f(r) = e^(-r/2) when r drops by d:
// f(r - d) - f(r), r - d in fp32 (get_gpu_of_cost_inc, util.hpp:127-129);
// each f rounded to fp32 from the double cw_exp; saturated at 1e30 (an
// infinite or NaN difference also saturates).  d = 1 is the wire cost of the
// non-BIT mode, d = 0.5 half a via (util.hpp:149-193).
float inc_cost(float r, float d) {
  const float rd = r - d;
  float v = (float)cw_exp(-0.5 * (double)rd) - (float)cw_exp(-0.5 * (double)r);
  if (!(v < kSat)) v = kSat;
  return v;
}

inline uint32_t ord_key(float f) {
  uint32_t u;
  f = f + 0.0f;  // -0 -> +0 so that -0 and +0 compare equal as in float ==
  std::memcpy(&u, &f, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

struct Ctx {
  int X, Y, Z, Z2, model, commit_mode;
  bool true_mst = false;
  int64_t cs;
  std::vector<int32_t> cap0, rem, stamp;
  bool fmode = false;       // fp32 capacities: remf is the state, rem counts -demand
  std::vector<float> remf;
  // non-BIT mode (commit_mode 2): remf is the state; pfx holds the wire
  // costs' exclusive run prefix P, vterm the via cost per cell
  bool prefix = false;
  std::vector<float> vterm;
  int chunk = 1;  // scan chunk: ceil(max(X, Y) / 128)
  std::vector<float> pfx;  // materialised Q(x) per cell, WIRE_ENC order
  float lut[kLut];
  int32_t bstamp = 0;
  // nets
  std::vector<int64_t> pin_pfx;
  std::vector<int32_t> pins;
  // last batch pair results
  std::vector<float> last_cost;
  std::vector<int16_t> last_pat;

  // util.hpp:38-44
  inline int64_t enc(int z, int x, int y) const {
    return (z & 1) ? (int64_t)z * X * Y + (int64_t)x * Y + y
                   : (int64_t)z * X * Y + (int64_t)y * X + x;
  }
  // util.hpp:55-59
  inline void dec(int64_t v, int& z, int& x, int& y) const {
    z = (int)(v / ((int64_t)X * Y));
    if (z & 1) { x = (int)(v / Y % X); y = (int)(v % Y); }
    else       { x = (int)(v % X);     y = (int)(v / X % Y); }
  }
  inline int run_len(int z) const { return (z & 1) ? Y : X; }
  // Q(z, line, pos): sum of cells 0..pos on the run, pos = -1 gives 0.
  inline float Q(int z, int line, int pos) const {
    if (pos < 0) return 0.0f;
    int L = run_len(z);
    int64_t base = (int64_t)z * X * Y + (int64_t)line * L;
    return pfx[base + pos];
  }
  inline float cost_of_rem(int32_t r) const {
    int i = r < kRMin ? kRMin : (r > kRMax ? kRMax : r);
    return lut[i - kRMin];
  }
  inline float cell_cost(int64_t cell) const {
    return fmode ? cost_f32(model, remf[cell]) : cost_of_rem(rem[cell]);
  }

  // Non-BIT cost state of one run (z, line) from its remaining capacities,
  // rebuilt before every batch in the reference (test_lpat.hpp:281-289):
  //   w(c) = inc_cost(r_c, 1)                        get_gpu_cost_cube, util.hpp:145-146
  //   P    = exclusive prefix of w along the run, in the chunked order of
  //          d_get_epfx_cost_cube_inplace (util.hpp:277-353, launched with
  //          128 scanning threads over max(X, Y) cells, test_lpat.hpp:270-273):
  //          chunk t covers [tT, tT + T), T = ceil(max(X, Y) / 128); its local
  //          scan starts from the raw cost of the cell before the chunk, and
  //          P[i] = P[tT - 1] + local(i) for t >= 1 (local(i) alone for t = 0)
  //   V(c) = via cost of the layer at c.  The reference writes the two halves
  //          of a via's charge to VIA_ENC slots vtid - 1 and vtid
  //          (util.hpp:147-193); VIA_ENC puts z fastest (:41-42), so vtid - 1
  //          is another layer's slot at a transposed cell, written by racing
  //          threads.  Contract: V(c) is the increase in cost of exactly the
  //          capacity the commit's via charge takes (mst_commit.hpp:17-61):
  //          position 0 of the run: inc_cost(r_c, 1); last position:
  //          inc_cost(r_(c-1), 1); otherwise inc_cost(r_(c-1), 0.5) +
  //          inc_cost(r_c, 0.5) in fp32, left half first.
  void refresh_run_prefix(int z, int line) {
    const int L = run_len(z);
    const int64_t base = (int64_t)z * X * Y + (int64_t)line * L;
    std::vector<float> w(L);
    for (int i = 0; i < L; ++i) w[i] = inc_cost(remf[base + i], 1.0f);
    for (int i = 0; i < L; ++i) {
      float v;
      if (i == 0) v = inc_cost(remf[base], 1.0f);
      else if (i == L - 1) v = inc_cost(remf[base + i - 1], 1.0f);
      else v = inc_cost(remf[base + i - 1], 0.5f) + inc_cost(remf[base + i], 0.5f);
      vterm[base + i] = v;
    }
    const int T = chunk;
    for (int t = 0; t * T < L; ++t) {
      float local = t ? w[t * T - 1] : 0.0f;
      const float carry = t ? pfx[base + t * T - 1] : 0.0f;
      for (int i = t * T; i < std::min(L, (t + 1) * T); ++i) {
        pfx[base + i] = t ? carry + local : local;
        local += w[i];
      }
    }
  }

  // Rebuild the materialised prefix of one run (z, line) from integer state.
  void refresh_run(int z, int line) {
    if (prefix) { refresh_run_prefix(z, line); return; }
    int L = run_len(z);
    int64_t base = (int64_t)z * X * Y + (int64_t)line * L;
    // level k holds the aligned block sums of 2^k cells, levels stored one
    // after another in a per-thread buffer (level k starts at ofs[k])
    static thread_local std::vector<float> lv;
    int ofs[32];
    size_t total = 0;
    int nlev = 0;
    for (int k = 0; (L >> k) > 0; ++k) { ofs[k] = (int)total; total += (size_t)(L >> k); nlev = k + 1; }
    if (lv.size() < total) lv.resize(total);
    for (int i = 0; i < L; ++i) lv[i] = cell_cost(base + i);
    for (int k = 1; k < nlev; ++k) {
      const int n = L >> k;
      for (int j = 0; j < n; ++j) lv[ofs[k] + j] = lv[ofs[k - 1] + 2 * j] + lv[ofs[k - 1] + 2 * j + 1];
    }
    for (int x = 0; x < L; ++x) {
      // bit_util.hpp:19-35: res = 0; while (i > 0) { res += node[i]; i -= i & -i; }
      uint32_t i = (uint32_t)x + 1;
      float res = 0.0f;
      while (i > 0) {
        int k = __builtin_ctz(i);
        res += lv[ofs[k] + (i >> k) - 1];
        i &= i - 1;
      }
      pfx[base + x] = res;
    }
  }
  // Rebuild the listed runs (run id z * max(X, Y) + line), in parallel.
  void refresh_runs(const std::vector<int>& runs) {
    const int M = std::max(X, Y);
#pragma omp parallel for schedule(dynamic, 4)
    for (size_t i = 0; i < runs.size(); ++i) refresh_run(runs[i] / M, runs[i] % M);
  }
  void refresh_all() {
    std::vector<int> runs;
    const int M = std::max(X, Y);
    for (int z = 0; z < Z; ++z) {
      int lines = (z & 1) ? X : Y;
      for (int l = 0; l < lines; ++l) runs.push_back(z * M + l);
    }
    refresh_runs(runs);
  }

  // bit_util.hpp:108-117: each via layer i in [min, max) adds Q(pos) - Q(pos-1).
  inline void via(int za, int zb, int x, int y, float& cost) const {
    int lo = std::min(za, zb), hi = std::max(za, zb);
    for (int i = lo; i < hi; ++i) {
      if (i & 1) cost += Q(i, x, y) - Q(i, x, y - 1);
      else       cost += Q(i, y, x) - Q(i, y, x - 1);
    }
  }
  // bit_util.hpp:91-107: inclusive wire along layer z's direction.
  inline void wire(int z, int& xs, int& ys, int x2, int y2, float& cost) const {
    if (z & 1) {
      if (ys < y2)      cost += Q(z, xs, y2) - Q(z, xs, ys - 1);
      else if (ys > y2) cost += Q(z, xs, ys) - Q(z, xs, y2 - 1);
      ys = y2;
    } else {
      if (xs < x2)      cost += Q(z, ys, x2) - Q(z, ys, xs - 1);
      else if (xs > x2) cost += Q(z, ys, xs) - Q(z, ys, x2 - 1);
      xs = x2;
    }
  }
  // bit_util.hpp:131-179 for one (pair, pattern).
  float pattern_cost(int64_t src, int64_t dst, int p, bool& valid) const {
    int z1, x1, y1, z2, x2, y2;
    dec(src, z1, x1, y1);
    dec(dst, z2, x2, y2);
    valid = false;
    int a = p / Z2, b = p % Z2;
    if (a == 0) return 0.0f;
    bool D = (x1 != x2 && y1 != y2);
    int zs = z1, xs = x1, ys = y1;
    float cost = 0.0f;
    if (zs != a) { via(zs, a, xs, ys, cost); zs = a; }
    wire(zs, xs, ys, x2, y2, cost);
    int bp = (zs & 1) ? 2 * b : 2 * b + 1;
    if (D && bp != 0) {
      if (bp >= Z) return 0.0f;  // b' < Z enforced (the reference reads past the top layer)
      via(zs, bp, xs, ys, cost);
      zs = bp;
      wire(zs, xs, ys, x2, y2, cost);
    }
    if (zs != z2) { via(zs, z2, xs, ys, cost); zs = z2; }
    if (xs != x2 || ys != y2) return 0.0f;
    valid = true;
    return cost;
  }
  // Non-BIT pattern cost, LPATTERN::get_lpat_dist (lpattern.hpp:98-160; the
  // GPU kernel :13-97 walks the same route).  Wires add |P(end) - P(start)|
  // (add_wires, lpattern_util.hpp:72-77): the cells [min, max) of the run.
  // The first via stack adds V over layers [min, max) (add_vias :64-68 called
  // with (zs, via1) or (via1, zs)), the corner and final stacks over the open
  // interval (min, max) (called with zs + 1 / via2 + 1).  b' >= Z is invalid
  // even for a straight pair, and the second leg is taken whenever the first
  // wire does not end at the destination cell, so b matters for straight
  // pairs too.  cost = wires + vias in fp32 (:155).
  float pattern_cost_prefix(int64_t src, int64_t dst, int p, bool& valid) const {
    int z1, x1, y1, z2, x2, y2;
    dec(src, z1, x1, y1);
    dec(dst, z2, x2, y2);
    valid = false;
    const int a = p / Z2;
    int bp = p % Z2;
    if (a == 0 || a >= Z) return 0.0f;
    float wc = 0.0f, vc = 0.0f;
    int zs = z1, xs = x1, ys = y1;
    auto V = [&](int i, int x, int y) { return vterm[enc(i, x, y)]; };
    auto Pat = [&](int z, int x, int y) { return pfx[enc(z, x, y)]; };
    auto wire = [&](int z) {
      if (z & 1) { wc += std::fabs(Pat(z, xs, y2) - Pat(z, xs, ys)); ys = y2; }
      else       { wc += std::fabs(Pat(z, x2, ys) - Pat(z, xs, ys)); xs = x2; }
    };
    if (zs != a) {
      for (int i = std::min(zs, a); i < std::max(zs, a); ++i) vc += V(i, xs, ys);
      zs = a;
    }
    wire(zs);
    bp = (zs & 1) ? 2 * bp : 2 * bp + 1;
    if (bp >= Z) return 0.0f;
    if (zs != z2 || xs != x2 || ys != y2) {
      if (bp > 0) {
        if (bp != zs) {
          for (int i = std::min(zs, bp) + 1; i < std::max(zs, bp); ++i) vc += V(i, xs, ys);
          zs = bp;
        }
        wire(zs);
      }
      if (zs != z2) {
        for (int i = std::min(zs, z2) + 1; i < std::max(zs, z2); ++i) vc += V(i, xs, ys);
        zs = z2;
      }
    }
    if (xs != x2 || ys != y2 || zs != z2) return 0.0f;
    valid = true;
    return wc + vc;
  }

  // Minimum over p of (cost, p), ascending p with strict <, lpattern.hpp:144-157.
  void best_pattern(int64_t src, int64_t dst, float& best, int& bp) const {
    bool have = false;
    best = 0.0f; bp = -1;
    for (int p = 0; p < Z * Z2; ++p) {
      bool v;
      float c = prefix ? pattern_cost_prefix(src, dst, p, v) : pattern_cost(src, dst, p, v);
      if (!v) continue;
      if (!have || ord_key(c) < ord_key(best)) { best = c; bp = p; have = true; }
    }
  }

  // Non-BIT commit of one edge (MST_COMMIT, mst_commit.hpp:114-251): wire
  // cells [min, max) of each wire (gpu_commit_wire :68-113, the difference
  // array's +1 / -1), the first wire on a, the second on b' whenever b' > 0
  // (gpu_add_wires :188-251); then via charges in half units along the route
  // (gpu_add_vias :114-186): a stack from z1 to a at the source, a to b' at
  // the position after the first wire when b' > 0, then to z2, each over
  // layers [min, max) (gpu_commit_via :17-61): 2 halves to the cell at run
  // position 0, 2 to the previous cell at the last position, else 1 to the
  // previous cell and 1 to the cell.
  void edge_cells_prefix(int64_t src, int64_t dst, int p, std::vector<int64_t>& wires,
                         std::vector<std::pair<int64_t, int>>& halves) const {
    int z1, x1, y1, z2, x2, y2;
    dec(src, z1, x1, y1);
    dec(dst, z2, x2, y2);
    const int a = p / Z2, b = p % Z2;
    int x = x1, y = y1;
    auto wire = [&](int z) {
      if (z & 1) { for (int t = std::min(y, y2); t < std::max(y, y2); ++t) wires.push_back(enc(z, x, t)); y = y2; }
      else       { for (int t = std::min(x, x2); t < std::max(x, x2); ++t) wires.push_back(enc(z, t, y)); x = x2; }
    };
    wire(a);
    const int bp = (a & 1) ? 2 * b : 2 * b + 1;
    if (bp > 0) wire(bp);
    auto via = [&](int za, int zb) {
      for (int i = std::min(za, zb); i < std::max(za, zb); ++i) {
        const int pos = (i & 1) ? y : x, L = run_len(i);
        const int64_t c = enc(i, x, y);
        if (pos == 0) halves.emplace_back(c, 2);
        else if (pos == L - 1) halves.emplace_back(c - 1, 2);
        else { halves.emplace_back(c - 1, 1); halves.emplace_back(c, 1); }
      }
    };
    x = x1; y = y1;
    int z = z1;
    via(z, a);
    z = a;
    if (z & 1) y = y2; else x = x2;
    if (bp > 0) {
      via(z, bp);
      z = bp;
      if (z & 1) y = y2; else x = x2;
    }
    if (z != z2) via(z, z2);
  }

  // mst_bit_commit.hpp:36-67,147-235: cells marked for one committed edge.
  void edge_cells(int64_t src, int64_t dst, int p, std::vector<int64_t>& out) const {
    int z1, x1, y1, z2, x2, y2;
    dec(src, z1, x1, y1);
    dec(dst, z2, x2, y2);
    int a = p / Z2, b = p % Z2;
    bool D = (x1 != x2 && y1 != y2);
    int z = z1, x = x1, y = y1;
    int k = 0;
    auto interior = [&](int za, int zb) {
      int lo = std::min(za, zb), hi = std::max(za, zb);
      for (int l = lo + 1; l < hi; ++l) out.push_back(enc(l, x, y));
      return hi - lo - 1;
    };
    auto wire_cells = [&](int zw, bool first) {
      if (zw & 1) {
        int lo = std::min(y, y2), hi = std::max(y, y2);
        for (int t = lo; t <= hi; ++t) out.push_back(enc(zw, x, t));
        if (first && commit_mode == 0)  // quirk cells below enc 0 (X*Y < Z only) are outside the grid: skipped
          for (int t = 1; t <= k && enc(zw, x, lo) - t >= 0; ++t) out.push_back(enc(zw, x, lo) - t);
        y = y2;
      } else {
        int lo = std::min(x, x2), hi = std::max(x, x2);
        for (int t = lo; t <= hi; ++t) out.push_back(enc(zw, t, y));
        if (first && commit_mode == 0)
          for (int t = 1; t <= k && enc(zw, lo, y) - t >= 0; ++t) out.push_back(enc(zw, lo, y) - t);
        x = x2;
      }
    };
    if (z != a) { k = interior(z, a); z = a; }
    wire_cells(z, true);
    int bp = (z & 1) ? 2 * b : 2 * b + 1;
    if (D && bp != 0) {
      interior(z, bp);
      z = bp;
      wire_cells(z, false);
    }
    if (z != z2) { interior(z, z2); z = z2; }
  }
};

// Root-pair Boruvka on one net, gmst_util.hpp:164-341 with the lowest-index
// tie contract.  costs[o] for o = j(j-1)/2 + k.  Emits (on1, on2) pairs where
// on1 is the selecting component's root at round start and on2 the partner's;
// with true_mst the selecting component's minimum pair itself, its endpoint
// in that component first (the rounds and unions are the same).
int boruvka(int n, const float* costs, std::vector<std::pair<int, int>>& edges,
            std::vector<int>& eo, float& net_cost, bool true_mst = false) {
  std::vector<int> dsu(n), partner(n), end_in(n), end_out(n);
  std::vector<uint64_t> best(n);
  for (int i = 0; i < n; ++i) dsu[i] = i;
  auto find = [&](int v) {
    int r = v;
    while (dsu[r] != r) r = dsu[r];
    dsu[v] = r;  // gmst_util.hpp:68-82 compresses only the queried node
    return r;
  };
  edges.clear(); eo.clear();
  net_cost = 0.0f;
  int rounds = 0;
  while ((int)edges.size() < n - 1) {
    ++rounds;
    std::fill(best.begin(), best.end(), ~0ull);
    int64_t P = (int64_t)n * (n - 1) / 2;
    int64_t o = 0;
    for (int j = 1; j < n; ++j) {
      for (int k = 0; k < j; ++k, ++o) {
        int rj = find(j), rk = find(k);
        if (rj == rk) continue;
        uint64_t key = ((uint64_t)ord_key(costs[o]) << 32) | (uint64_t)o;
        if (key < best[rj]) best[rj] = key;
        if (key < best[rk]) best[rk] = key;
      }
    }
    (void)P;
    // partners at round start (gmst_util.hpp:245-262)
    for (int r = 0; r < n; ++r) {
      partner[r] = -1;
      if (dsu[r] != r || best[r] == ~0ull) continue;
      int64_t bo = (int64_t)(best[r] & 0xffffffffu);
      int j = (int)((1.0 + std::sqrt(1.0 + 8.0 * (double)bo)) / 2.0);
      while ((int64_t)j * (j - 1) / 2 > bo) --j;
      while ((int64_t)(j + 1) * j / 2 <= bo) ++j;
      int k = (int)(bo - (int64_t)j * (j - 1) / 2);
      int rj = find(j), rk = find(k);
      partner[r] = (rj == r) ? rk : rj;
      end_in[r] = (rj == r) ? j : k;
      end_out[r] = (rj == r) ? k : j;
    }
    // serial union in ascending root order (gmst_util.hpp:296-341)
    for (int r = 0; r < n; ++r) {
      if (partner[r] < 0) continue;
      int n1 = find(r), n2 = find(partner[r]);
      if (n1 == n2) continue;
      if (n1 < n2) dsu[n1] = n2; else dsu[n2] = n1;
      net_cost += costs[best[r] & 0xffffffffu];
      if (true_mst) edges.emplace_back(end_in[r], end_out[r]);
      else edges.emplace_back(r, partner[r]);
      eo.push_back((int)(best[r] & 0xffffffffu));
    }
  }
  return rounds;
}

inline int64_t pair_index(int hi, int lo) { return (int64_t)hi * (hi - 1) / 2 + lo; }

}  // namespace

extern "C" {

void orc_lut(int model, float* out) { build_lut(model, out); }
int orc_lut_rmin() { return kRMin; }
int orc_lut_size() { return kLut; }

void* orc_create(int X, int Y, int Z, int cost_model, int commit_mode, const int32_t* cap) {
  if (X <= 0 || Y <= 0 || Z < 2 || !cap) return nullptr;
  Ctx* c = new Ctx();
  c->X = X; c->Y = Y; c->Z = Z; c->Z2 = (Z + 1) / 2;
  c->model = cost_model; c->commit_mode = commit_mode;
  c->cs = (int64_t)X * Y * Z;
  c->cap0.assign(cap, cap + c->cs);
  c->rem = c->cap0;
  c->stamp.assign(c->cs, 0);
  c->pfx.assign(c->cs, 0.0f);
  build_lut(cost_model, c->lut);
  if (commit_mode == 2) {  // non-BIT mode: fp32 remaining capacity is the state
    c->prefix = true;
    c->fmode = true;
    c->remf.assign(cap, cap + c->cs);
    c->vterm.assign(c->cs, 0.0f);
    c->chunk = (std::max(X, Y) + 127) / 128;
  }
  c->refresh_all();
  return c;
}

// fp32 capacities (read_data.hpp:136-143): the integer state starts at 0 and
// counts commits downwards, so cap0 - rem is still the demand.
void* orc_create_f32(int X, int Y, int Z, int cost_model, int commit_mode, const float* cap) {
  if (X <= 0 || Y <= 0 || Z < 2 || !cap) return nullptr;
  std::vector<int32_t> zero((size_t)X * Y * Z, 0);
  Ctx* c = (Ctx*)orc_create(X, Y, Z, cost_model, commit_mode, zero.data());
  c->fmode = true;
  c->remf.assign(cap, cap + c->cs);
  c->refresh_all();
  return c;
}

int orc_get_remf(void* h, float* out) {
  Ctx* c = (Ctx*)h;
  if (!c->fmode) return -1;
  std::memcpy(out, c->remf.data(), sizeof(float) * c->cs);
  return 0;
}

int orc_set_remf(void* h, const float* remf, const int32_t* demand) {
  Ctx* c = (Ctx*)h;
  if (!c->fmode) return -1;
  c->remf.assign(remf, remf + c->cs);
  for (int64_t i = 0; i < c->cs; ++i) c->rem[i] = -demand[i];
  c->refresh_all();
  return 0;
}

// Route evaluation (main.cpp:103-135 over the non-BIT commit's counters).
// For every batch, the wires of its routes go into a difference array
// (+1 at the run's lower end cell, -1 at the upper, gpu_commit_wire
// mst_commit.hpp:68-113), every run is scanned inclusively
// (get_gpu_pat_ipfx :252-328) and each cell with a positive sum counts one
// wire (update_cap_cube :329-341); vias charge their layer's run neighbours
// (gpu_commit_via :17-61).  Routes are walked as the writer prints them
// (writer_util.hpp:115-165).  Then the report loop of main.cpp:109-131 runs
// per layer in its x-then-y order, summing in double.  layer_out gets 9
// doubles per layer: wire, via, max_wire, max_via, max_cap, max_overflow,
// of_cost, demand, overflow.
int orc_evaluate(int X, int Y, int Z, const float* cap, const int32_t* demand, int n_batches,
                 const int32_t* batch_ofst, const int32_t* net_ids, const int64_t* pin_pfx, const int32_t* pins,
                 const int64_t* edges, const int32_t* lpat, int32_t* wcnt, int32_t* vcnt, double* layer_out) {
  const int64_t cs = (int64_t)X * Y * Z;
  const int Z2 = (Z + 1) / 2;
  auto enc = [&](int z, int x, int y) -> int64_t {
    return (z & 1) ? (int64_t)z * X * Y + (int64_t)x * Y + y : (int64_t)z * X * Y + (int64_t)y * X + x;
  };
  std::fill(wcnt, wcnt + cs, 0);
  std::fill(vcnt, vcnt + cs, 0);
  std::vector<int32_t> diff(cs);
  int64_t slot = 0;
  for (int b = 0; b < n_batches; ++b) {
    std::fill(diff.begin(), diff.end(), 0);
    for (int q = batch_ofst[b]; q < batch_ofst[b + 1]; ++q) {
      const int net = net_ids[q];
      const int64_t deg = pin_pfx[net + 1] - pin_pfx[net];
      const int32_t* np = pins + pin_pfx[net];
      for (int64_t j = 0; j < deg; ++j) {
        const int64_t e = edges[slot + j];
        if (e == -1) continue;
        int hi = (int)(e & 0x7fff), lo = (int)((e >> 15) & 0x7fff);
        if (hi < lo) std::swap(hi, lo);
        int z1, x1, y1, z2, x2, y2;
        auto dec = [&](int64_t v, int& z, int& x, int& y) {
          z = (int)(v / ((int64_t)X * Y));
          const int64_t r = v - (int64_t)z * X * Y;
          if (z & 1) { x = (int)(r / Y); y = (int)(r % Y); } else { y = (int)(r / X); x = (int)(r % X); }
        };
        dec(np[hi], z1, x1, y1);
        dec(np[lo], z2, x2, y2);
        const int p = lpat[slot + j];
        const int via1 = p / Z2;
        int via2 = p % Z2;
        const bool both = x1 != x2 && y1 != y2;
        auto vias = [&](int za, int zb) {
          for (int i = std::min(za, zb); i < std::max(za, zb); ++i) {
            const int64_t c = enc(i, x1, y1);
            const int pos = (i & 1) ? y1 : x1, L = (i & 1) ? Y : X;
            if (pos == 0) vcnt[c] += 2;
            else if (pos == L - 1) vcnt[c - 1] += 2;
            else { vcnt[c - 1] += 1; vcnt[c] += 1; }
          }
        };
        auto wire = [&](int z) {
          int64_t c1, c2;
          if (z & 1) { c1 = enc(z, x1, std::min(y1, y2)); c2 = enc(z, x1, std::max(y1, y2)); y1 = y2; }
          else       { c1 = enc(z, std::min(x1, x2), y1); c2 = enc(z, std::max(x1, x2), y1); x1 = x2; }
          diff[c1] += 1;
          diff[c2] -= 1;
        };
        if (z1 != via1) { vias(z1, via1); z1 = via1; }
        if (z1 > 0) wire(z1);
        via2 = (z1 & 1) ? 2 * via2 : 2 * via2 + 1;
        if (both && via2 != 0 && z1 != via2) { vias(z1, via2); z1 = via2; wire(z1); }
        if (z1 != z2) vias(z1, z2);
      }
      slot += deg;
    }
    for (int z = 0; z < Z; ++z) {
      const int L = (z & 1) ? Y : X, lines = (z & 1) ? X : Y;
      for (int l = 0; l < lines; ++l) {
        int32_t run = 0;
        const int64_t base = (int64_t)z * X * Y + (int64_t)l * L;
        for (int t = 0; t < L; ++t) {
          run += diff[base + t];
          if (run > 0) wcnt[base + t] += 1;
        }
      }
    }
  }
  for (int z = 0; z < Z; ++z) {
    double w = 0, v = 0, mw = 0, mv = 0, mc = 0, mo = 0, oc = 0, dm = 0, ov = 0;
    for (int x = 0; x < X; ++x)
      for (int y = 0; y < Y; ++y) {
        const int64_t c = enc(z, x, y);
        const double cw = wcnt[c], cv = vcnt[c], cc = cap[c];
        mw = std::max(mw, cw);
        mv = std::max(mv, cv);
        mc = std::max(mc, cc);
        const double of = 2 * cw + cv - 2 * cc;
        mo = std::max(mo, of / 2);
        oc += (double)(float)cw_exp((double)(float)(of * 0.25));
        w += cw;
        v += cv;
        dm += demand[c];
        ov += std::max(0.0, (double)demand[c] - cc);
      }
    double* o = layer_out + 9 * z;
    o[0] = w; o[1] = v; o[2] = mw; o[3] = mv; o[4] = mc; o[5] = mo; o[6] = oc; o[7] = dm; o[8] = ov;
  }
  return 0;
}

void orc_set_tree_mode(void* h, int true_mst) { ((Ctx*)h)->true_mst = true_mst != 0; }

double orc_cw_exp(double x) { return cw_exp(x); }
float orc_cost_f32(int model, float r) { return cost_f32(model, r); }

void orc_destroy(void* h) { delete (Ctx*)h; }

int orc_set_nets(void* h, int n_nets, const int64_t* pin_pfx, const int32_t* pins) {
  Ctx* c = (Ctx*)h;
  c->pin_pfx.assign(pin_pfx, pin_pfx + n_nets + 1);
  c->pins.assign(pins, pins + pin_pfx[n_nets]);
  return 0;
}

int orc_get_demand(void* h, int32_t* out) {
  Ctx* c = (Ctx*)h;
  for (int64_t i = 0; i < c->cs; ++i) out[i] = c->cap0[i] - c->rem[i];
  return 0;
}

int orc_get_pfx(void* h, float* out) {
  Ctx* c = (Ctx*)h;
  std::memcpy(out, c->pfx.data(), sizeof(float) * c->cs);
  return 0;
}

int orc_set_rem(void* h, const int32_t* rem) {
  Ctx* c = (Ctx*)h;
  c->rem.assign(rem, rem + c->cs);
  c->refresh_all();
  return 0;
}

int orc_get_vterm(void* h, float* out) {
  Ctx* c = (Ctx*)h;
  if (!c->prefix) return -1;
  std::memcpy(out, c->vterm.data(), sizeof(float) * c->cs);
  return 0;
}

float orc_inc_cost(float r, float d) { return inc_cost(r, d); }

// Non-BIT commit of one edge: wire cells (with repeats) and (cell, halves).
int orc_edge_commit_prefix(void* h, int32_t src, int32_t dst, int p, int64_t* wires, int wcap, int64_t* hcell,
                           int32_t* hcount, int hcap, int* n_halves) {
  Ctx* c = (Ctx*)h;
  std::vector<int64_t> w;
  std::vector<std::pair<int64_t, int>> hv;
  c->edge_cells_prefix(src, dst, p, w, hv);
  for (int i = 0; i < (int)std::min<size_t>(w.size(), (size_t)wcap); ++i) wires[i] = w[i];
  for (int i = 0; i < (int)std::min<size_t>(hv.size(), (size_t)hcap); ++i) { hcell[i] = hv[i].first; hcount[i] = hv[i].second; }
  *n_halves = (int)hv.size();
  return (int)w.size();
}

float orc_pattern_cost(void* h, int32_t src, int32_t dst, int p, int* valid) {
  bool v;
  Ctx* c = (Ctx*)h;
  float r = c->prefix ? c->pattern_cost_prefix(src, dst, p, v) : c->pattern_cost(src, dst, p, v);
  *valid = v;
  return r;
}

// Brute-force restatement of one pattern's cost: walk every cell of the
// multiset of A.4 explicitly and sum the cell costs in double.  Used only to
// pin the Fenwick evaluation (the two agree up to fp32 rounding).
double orc_pattern_cost_walk(void* h, int32_t src, int32_t dst, int p, int* valid) {
  Ctx* c = (Ctx*)h;
  int z1, x1, y1, z2, x2, y2;
  c->dec(src, z1, x1, y1);
  c->dec(dst, z2, x2, y2);
  *valid = 0;
  int a = p / c->Z2, b = p % c->Z2;
  if (a == 0) return 0.0;
  bool D = (x1 != x2 && y1 != y2);
  int z = z1, x = x1, y = y1;
  double s = 0.0;
  auto cell = [&](int zz, int xx, int yy) { return (double)c->cell_cost(c->enc(zz, xx, yy)); };
  auto via = [&](int za, int zb) {
    for (int l = std::min(za, zb); l < std::max(za, zb); ++l) s += cell(l, x, y);
  };
  auto wire = [&](int zw) {
    if (zw & 1) { if (y != y2) for (int t = std::min(y, y2); t <= std::max(y, y2); ++t) s += cell(zw, x, t); y = y2; }
    else        { if (x != x2) for (int t = std::min(x, x2); t <= std::max(x, x2); ++t) s += cell(zw, t, y); x = x2; }
  };
  if (z != a) { via(z, a); z = a; }
  wire(z);
  int bp = (z & 1) ? 2 * b : 2 * b + 1;
  if (D && bp != 0) {
    if (bp >= c->Z) return 0.0;
    via(z, bp); z = bp; wire(z);
  }
  if (z != z2) { via(z, z2); z = z2; }
  if (x != x2 || y != y2) return 0.0;
  *valid = 1;
  return s;
}

int orc_best_pattern(void* h, int32_t src, int32_t dst, float* cost) {
  float b; int p;
  ((Ctx*)h)->best_pattern(src, dst, b, p);
  *cost = b;
  return p;
}

int orc_edge_cells(void* h, int32_t src, int32_t dst, int p, int64_t* out, int cap) {
  std::vector<int64_t> v;
  ((Ctx*)h)->edge_cells(src, dst, p, v);
  int n = (int)std::min<size_t>(v.size(), (size_t)cap);
  for (int i = 0; i < n; ++i) out[i] = v[i];
  return (int)v.size();
}

// Boruvka on an explicit cost array (local net 0).  out_edges gets n slots:
// n-1 encoded edges then -1; returns the number of rounds.
int orc_boruvka(int n, const float* costs, int64_t* out_edges, float* out_cost, int true_mst) {
  std::vector<std::pair<int, int>> e;
  std::vector<int> eo;
  float nc;
  int rounds = boruvka(n, costs, e, eo, nc, true_mst != 0);
  for (size_t i = 0; i < e.size(); ++i)
    out_edges[i] = ((int64_t)e[i].first << 15) | (int64_t)e[i].second;
  out_edges[n - 1] = -1;
  *out_cost = nc;
  return rounds;
}

// One batch of nets (disjoint bounding boxes).  Slot layout: net i of the
// batch owns deg(i) consecutive slots; n-1 edges then a -1 sentinel
// (gmst_util.hpp:322-339).  out_lpat gets each edge's pattern, -1 on the
// sentinel.  Returns the number of distinct cells whose demand was bumped.
int64_t orc_route_batch(void* h, const int32_t* net_ids, int n, int64_t* out_edges,
                        int32_t* out_lpat, float* out_net_cost, int nthreads) {
  Ctx* c = (Ctx*)h;
  std::vector<int64_t> slot_pfx(n + 1, 0), pair_pfx(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    int64_t d = c->pin_pfx[net_ids[i] + 1] - c->pin_pfx[net_ids[i]];
    slot_pfx[i + 1] = slot_pfx[i] + d;
    pair_pfx[i + 1] = pair_pfx[i] + d * (d - 1) / 2;
  }
  c->last_cost.assign(pair_pfx[n], 0.0f);
  c->last_pat.assign(pair_pfx[n], -1);
  std::vector<std::vector<int64_t>> cells(n);
  std::vector<std::vector<std::pair<int64_t, int>>> halves(n);  // non-BIT mode via charges
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; ++i) {
    const int32_t* pins = c->pins.data() + c->pin_pfx[net_ids[i]];
    int d = (int)(slot_pfx[i + 1] - slot_pfx[i]);
    float* pc = c->last_cost.data() + pair_pfx[i];
    int16_t* pp = c->last_pat.data() + pair_pfx[i];
    int64_t o = 0;
    for (int j = 1; j < d; ++j)
      for (int k = 0; k < j; ++k, ++o) {
        float b; int p;
        c->best_pattern(pins[j], pins[k], b, p);
        pc[o] = b; pp[o] = (int16_t)p;
      }
    std::vector<std::pair<int, int>> e;
    std::vector<int> eo;
    float nc;
    boruvka(d, pc, e, eo, nc, c->true_mst);
    out_net_cost[i] = nc;
    int64_t base = slot_pfx[i];
    for (size_t t = 0; t < e.size(); ++t) {
      int on1 = e[t].first, on2 = e[t].second;
      out_edges[base + t] = ((int64_t)i << 30) | ((int64_t)on1 << 15) | (int64_t)on2;
      int hi = std::max(on1, on2), lo = std::min(on1, on2);
      int p = pp[pair_index(hi, lo)];
      out_lpat[base + t] = p;
      if (c->prefix) c->edge_cells_prefix(pins[hi], pins[lo], p, cells[i], halves[i]);
      else c->edge_cells(pins[hi], pins[lo], p, cells[i]);
    }
    out_edges[base + d - 1] = -1;
    out_lpat[base + d - 1] = -1;
  }
  if (c->prefix) {
    // mst_commit.hpp:342-364: every cell a wire of the batch covers loses 1.0
    // once (the run scan of the difference array and update_cap_cube,
    // :252-341); then the via charges (:114-186), summed per cell in half
    // units and subtracted in one fp32 step (exact for half-integer states;
    // the reference adds them with float atomics in any order).
    c->bstamp++;
    std::vector<char> dirty((size_t)c->Z * std::max(c->X, c->Y), 0);
    auto mark_dirty = [&](int64_t cell) {
      int z, x, y;
      c->dec(cell, z, x, y);
      dirty[(size_t)z * std::max(c->X, c->Y) + ((z & 1) ? x : y)] = 1;
    };
    int64_t touched = 0;
    for (int i = 0; i < n; ++i)
      for (int64_t cell : cells[i]) {
        if (c->stamp[cell] == c->bstamp) continue;
        c->stamp[cell] = c->bstamp;
        c->remf[cell] -= 1.0f;
        ++touched;
        mark_dirty(cell);
      }
    std::vector<std::pair<int64_t, int>> all;
    for (int i = 0; i < n; ++i) all.insert(all.end(), halves[i].begin(), halves[i].end());
    std::sort(all.begin(), all.end());
    for (size_t k = 0; k < all.size();) {
      const int64_t cell = all[k].first;
      int h = 0;
      for (; k < all.size() && all[k].first == cell; ++k) h += all[k].second;
      c->remf[cell] -= 0.5f * (float)h;
      mark_dirty(cell);
    }
    std::vector<int> runs;
    for (int z = 0; z < c->Z; ++z) {
      int lines = (z & 1) ? c->X : c->Y;
      for (int l = 0; l < lines; ++l)
        if (dirty[(size_t)z * std::max(c->X, c->Y) + l]) runs.push_back(z * std::max(c->X, c->Y) + l);
    }
    c->refresh_runs(runs);
    return touched;
  }
  // bit_util.hpp:75-90: each flagged cell loses one unit of capacity once.
  c->bstamp++;
  int64_t touched = 0;
  std::vector<char> dirty((size_t)c->Z * std::max(c->X, c->Y), 0);
  for (int i = 0; i < n; ++i)
    for (int64_t cell : cells[i]) {
      if (c->stamp[cell] == c->bstamp) continue;
      c->stamp[cell] = c->bstamp;
      c->rem[cell] -= 1;
      if (c->fmode) c->remf[cell] -= 1.0f;  // bit_util.hpp:87
      ++touched;
      int z, x, y;
      c->dec(cell, z, x, y);
      int line = (z & 1) ? x : y;
      dirty[(size_t)z * std::max(c->X, c->Y) + line] = 1;
    }
  std::vector<int> runs;
  for (int z = 0; z < c->Z; ++z) {
    int lines = (z & 1) ? c->X : c->Y;
    for (int l = 0; l < lines; ++l)
      if (dirty[(size_t)z * std::max(c->X, c->Y) + l]) runs.push_back(z * std::max(c->X, c->Y) + l);
  }
  c->refresh_runs(runs);
  return touched;
}

// ---------------------------------------------------------------------------
// Batch generation: NET_PART::part_nets in its batch mode
// (part_net/net_part.hpp:236-262: nets first sorted by (minx, maxx); then per
// batch part_xdim, :167-235, and the compaction of :326-337).
//
// One batch:
//   * targ = ceil(remaining / parts)                                   :168
//   * the remaining nets are sorted by maxx                            :186-188
//   * walking them in that order, a net joins the current x-bucket when its
//     minx lies right of the bucket cut (cut = maxx of the net that closed
//     the previous bucket, -1 at first); a bucket closes once it holds targ
//     nets; nets left of the cut join no bucket this batch                :191-202
//   * each bucket lists its nets in reverse walking order (the fill of
//     pick_ids from the back, :207-212), then is sorted by half perimeter
//     (W + H, :178-182, :215-218)
//   * in that order a net is taken when its closed 2-D box meets no box
//     already taken in its bucket (Boost rtree "intersects", :158-166, 219-226)
//   * the taken nets form the batch; the others keep their relative order
//     at the front of the list for the next batch (:326-337)
// Ties of the three sorts follow std::sort on the same sequence with the same
// comparators (unstable: the tie order is the library's, as in the reference
// built with the same g++).  The compaction loop of :326-337 keeps the nets not
// taken in their relative order (its scan holds every taken net between the
// write and read positions), which is std::stable_partition.  The batch's own
// order is (pin count, net id): the reference appends under omp critical and
// sorts by pin count with std::sort (:227-230, :347-349), which fixes no order.
//
// Box tests here use a uniform grid of 32x32-GCELL tiles listing the boxes
// taken so far that overlap each tile, and the exact closed-box test on them;
// brute force (every taken box) when tile == 0.
namespace {
struct PBox { int minx, maxx, miny, maxy; };
inline bool boxes_meet(const PBox& a, const PBox& b) {
  return !(a.maxx < b.minx || b.maxx < a.minx || a.maxy < b.miny || b.maxy < a.miny);
}
struct TakenBoxes {
  int tile, tx, ty;
  std::vector<PBox> all;
  std::vector<std::vector<int>> bins;
  TakenBoxes(int X, int Y, int t) : tile(t) {
    tx = t ? (X + t - 1) / t : 1;
    ty = t ? (Y + t - 1) / t : 1;
    bins.assign((size_t)tx * ty, {});
  }
  bool meets(const PBox& b) const {
    if (!tile) {
      for (const PBox& o : all) if (boxes_meet(o, b)) return true;
      return false;
    }
    for (int j = b.miny / tile; j <= b.maxy / tile; ++j)
      for (int i = b.minx / tile; i <= b.maxx / tile; ++i)
        for (int k : bins[(size_t)j * tx + i]) if (boxes_meet(all[k], b)) return true;
    return false;
  }
  void add(const PBox& b) {
    const int k = (int)all.size();
    all.push_back(b);
    if (!tile) return;
    for (int j = b.miny / tile; j <= b.maxy / tile; ++j)
      for (int i = b.minx / tile; i <= b.maxx / tile; ++i) bins[(size_t)j * tx + i].push_back(k);
  }
};
}  // namespace

// Returns the number of batches; batch_ofst[nb+1], net_ids[n_nets].  With
// trace_batch >= 0, trace_bucket[n_nets] gets each net's bucket in that batch
// (-1: in no bucket, -2: taken by an earlier batch).
int orc_partition(int X, int Y, int n_nets, const int64_t* pin_pfx, const int32_t* pins, int parts, int tile,
                  int32_t* batch_ofst, int32_t* net_ids, int trace_batch, int32_t* trace_bucket) {
  if (X <= 0 || Y <= 0 || n_nets < 0 || parts < 1) return -1;
  std::vector<PBox> bx(n_nets);
  std::vector<int> deg(n_nets);
  const int64_t XY = (int64_t)X * Y;
  for (int i = 0; i < n_nets; ++i) {
    PBox b{X, -1, Y, -1};
    for (int64_t p = pin_pfx[i]; p < pin_pfx[i + 1]; ++p) {
      const int64_t v = pins[p], z = v / XY, r = v - z * XY;
      const int x = (z & 1) ? (int)(r / Y) : (int)(r % X);
      const int y = (z & 1) ? (int)(r % Y) : (int)(r / X);
      b.minx = std::min(b.minx, x); b.maxx = std::max(b.maxx, x);
      b.miny = std::min(b.miny, y); b.maxy = std::max(b.maxy, y);
    }
    bx[i] = b;
    deg[i] = (int)(pin_pfx[i + 1] - pin_pfx[i]);
  }
  std::vector<int> order(n_nets);
  for (int i = 0; i < n_nets; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) {  // net_part.hpp:244-249, 259-262
    if (bx[a].minx == bx[b].minx) return bx[a].maxx < bx[b].maxx;
    return bx[a].minx < bx[b].minx;
  });
  std::vector<char> done(n_nets, 0);
  int remaining = n_nets, nb = 0, out = 0;
  batch_ofst[0] = 0;
  while (remaining > 0) {
    const int targ = (remaining + parts - 1) / parts;
    std::sort(order.begin(), order.begin() + remaining, [&](int a, int b) { return bx[a].maxx < bx[b].maxx; });
    std::vector<std::vector<int>> bucket(1);
    int cut = -1;
    for (int i = 0; i < remaining; ++i) {
      const int id = order[i];
      if (bx[id].minx > cut) bucket.back().push_back(id);
      if ((int)bucket.back().size() == targ) {
        cut = bx[id].maxx;
        bucket.emplace_back();
      }
    }
    if (trace_batch == nb)
      for (int i = 0; i < n_nets; ++i) trace_bucket[i] = done[i] ? -2 : -1;
    std::vector<int> batch;
    for (size_t k = 0; k < bucket.size(); ++k) {
      std::vector<int>& list = bucket[k];
      std::reverse(list.begin(), list.end());
      std::sort(list.begin(), list.end(), [&](int a, int b) {
        return (bx[a].maxx - bx[a].minx + 1) + (bx[a].maxy - bx[a].miny + 1) <
               (bx[b].maxx - bx[b].minx + 1) + (bx[b].maxy - bx[b].miny + 1);
      });
      TakenBoxes taken(X, Y, tile);
      for (int id : list) {
        if (trace_batch == nb) trace_bucket[id] = (int32_t)k;
        if (taken.meets(bx[id])) continue;
        taken.add(bx[id]);
        batch.push_back(id);
        done[id] = 1;
      }
    }
    std::stable_partition(order.begin(), order.begin() + remaining, [&](int id) { return !done[id]; });
    remaining -= (int)batch.size();
    std::sort(batch.begin(), batch.end(), [&](int a, int b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
    for (int id : batch) net_ids[out++] = id;
    batch_ofst[++nb] = out;
  }
  return nb;
}

int64_t orc_last_pairs(void* h) { return (int64_t)((Ctx*)h)->last_cost.size(); }

int orc_get_pair_costs(void* h, float* cost, int16_t* pat) {
  Ctx* c = (Ctx*)h;
  std::memcpy(cost, c->last_cost.data(), sizeof(float) * c->last_cost.size());
  std::memcpy(pat, c->last_pat.data(), sizeof(int16_t) * c->last_pat.size());
  return 0;
}

}  // extern "C"
