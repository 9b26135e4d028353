// Host side of the routing path: seeded synthetic workloads and batch
// generation.  No device work happens here.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/iitmgr.h"

namespace {

// ---- PRNG: splitmix64 seeding xoshiro256** ------------------------------
inline uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t t = seed;
    for (auto& w : s) w = splitmix64(t);
  }
  static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  // uniform integer in [lo, hi]
  int64_t uni(int64_t lo, int64_t hi) {
    uint64_t range = (uint64_t)(hi - lo + 1);
    return lo + (int64_t)(((unsigned __int128)next() * range) >> 64);
  }
  // uniform double in [0, 1)
  double unif() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  // standard normal (Box-Muller, one value per call)
  double normal() {
    double u1 = 1.0 - unif();  // (0, 1]
    double u2 = unif();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

constexpr uint64_t kGridSalt = 0x67726964ull;  // "grid"
constexpr uint64_t kNetSalt = 0x6e657473ull;   // "nets"
constexpr uint64_t kHotSalt = 0x686f7473ull;   // "hots"

inline int64_t wenc(int X, int Y, int z, int x, int y) {
  return (z & 1) ? (int64_t)z * X * Y + (int64_t)x * Y + y : (int64_t)z * X * Y + (int64_t)y * X + x;
}

// Pin count per workload family.
int draw_pins(int kind, Rng& r, const std::vector<double>& zipf_cdf) {
  if (kind == 1) {
    double u = r.unif();
    if (u < 0.60) return 2;
    if (u < 0.90) return (int)r.uni(3, 5);
    return (int)r.uni(6, 16);
  }
  if (kind == 2 || kind == 4) {
    double u = r.unif();
    if (u < 0.55) return 2;
    if (u < 0.83) return (int)r.uni(3, 5);
    if (u < 0.96) return (int)r.uni(6, 20);
    return (int)r.uni(21, 200);
  }
  // kinds 3, 5: Zipf(2.2) on [2, 1000]
  double u = r.unif();
  size_t i = (size_t)(std::upper_bound(zipf_cdf.begin(), zipf_cdf.end(), u) - zipf_cdf.begin());
  if (i >= zipf_cdf.size()) i = zipf_cdf.size() - 1;
  return 2 + (int)i;
}

int draw_side(int kind, Rng& r, int limit) {
  int s;
  if (kind == 1) {
    // geometric, support >= 1, mean 8: 1 + floor(log(U) / log(1 - 1/8))
    double u = 1.0 - r.unif();
    s = 1 + (int)std::floor(std::log(u) / std::log(1.0 - 1.0 / 8.0));
  } else {
    // log-normal, median 6, sigma 1.2, capped at 400
    double v = std::exp(std::log(6.0) + 1.2 * r.normal());
    s = (int)std::lround(v);
    if (s > 400) s = 400;
  }
  if (s < 1) s = 1;
  if (s > limit) s = limit;
  return s;
}

struct BBox { int minx, maxx, miny, maxy; };

}  // namespace

extern "C" {

int gr_gen_grid(int32_t kind, uint64_t seed, int32_t X, int32_t Y, int32_t Z, int32_t* cap) {
  if (kind < 1 || kind > 5 || X <= 0 || Y <= 0 || Z < 3 || !cap) return GR_EINVAL;
  Rng r(seed ^ (kGridSalt << 32));
  int64_t cs = (int64_t)X * Y * Z;
  if (kind == 1) {
    for (int64_t i = 0; i < cs; ++i) cap[i] = (int32_t)r.uni(2, 8);
    return 0;
  }
  for (int64_t i = 0; i < cs; ++i) {
    int32_t c = (int32_t)r.uni(4, 16);
    if (r.unif() < 0.05) c = 0;  // blockage
    cap[i] = c;
  }
  if (kind == 5) {
    for (int64_t i = 0; i < cs; ++i) cap[i] = (int32_t)std::floor(cap[i] * 0.6);
    Rng h(seed ^ (kHotSalt << 32));
    const double sigma = 40.0;
    const double rad = sigma * std::sqrt(2.0 * std::log(2.0));  // weight >= 0.5
    for (int k = 0; k < 30; ++k) {
      int cx = (int)h.uni(0, X - 1), cy = (int)h.uni(0, Y - 1);
      int r0 = (int)std::ceil(rad);
      for (int y = std::max(0, cy - r0); y <= std::min(Y - 1, cy + r0); ++y)
        for (int x = std::max(0, cx - r0); x <= std::min(X - 1, cx + r0); ++x) {
          double d2 = (double)(x - cx) * (x - cx) + (double)(y - cy) * (y - cy);
          if (std::exp(-d2 / (2.0 * sigma * sigma)) < 0.5) continue;
          for (int z = 0; z < Z; ++z) cap[wenc(X, Y, z, x, y)] = (int32_t)h.uni(1, 2);
        }
    }
  }
  return 0;
}

int gr_gen_nets(int32_t kind, uint64_t seed, int32_t X, int32_t Y, int32_t Z, int32_t n_nets_req,
                int32_t* n_nets, int64_t* n_pins, int64_t* pin_pfx, int32_t* pin_enc) {
  if (kind < 1 || kind > 5 || X <= 0 || Y <= 0 || Z < 3 || n_nets_req < 0 || !n_nets || !n_pins)
    return GR_EINVAL;
  std::vector<double> zipf;
  if (kind == 3 || kind == 5) {
    double s = 0;
    for (int n = 2; n <= 1000; ++n) s += std::pow((double)n, -2.2);
    double c = 0;
    for (int n = 2; n <= 1000; ++n) { c += std::pow((double)n, -2.2) / s; zipf.push_back(c); }
  }
  Rng r(seed ^ (kNetSalt << 32));
  std::vector<int64_t> buf;
  int32_t kept = 0;
  int64_t pins_total = 0;
  if (pin_pfx) pin_pfx[0] = 0;
  for (int32_t t = 0; t < n_nets_req; ++t) {
    int np = draw_pins(kind, r, zipf);
    int w = draw_side(kind, r, X), h = draw_side(kind, r, Y);
    int x0 = (int)r.uni(0, X - w), y0 = (int)r.uni(0, Y - h);
    buf.resize(np);
    for (int i = 0; i < np; ++i) {
      int x = (int)r.uni(x0, x0 + w - 1);
      int y = (int)r.uni(y0, y0 + h - 1);
      int z = (int)r.uni(0, 1);
      buf[i] = wenc(X, Y, z, x, y);
    }
    // read_data.hpp:413-432: one pin per cell, ascending enc
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
    // read_data.hpp:366-412: single-pin nets leave the routing set
    if (buf.size() < 2) continue;
    if (pin_enc)
      for (size_t i = 0; i < buf.size(); ++i) pin_enc[pins_total + (int64_t)i] = (int32_t)buf[i];
    pins_total += (int64_t)buf.size();
    ++kept;
    if (pin_pfx) pin_pfx[kept] = pins_total;
  }
  *n_nets = kept;
  *n_pins = pins_total;
  return 0;
}

// Batch generation, restating NET_PART::part_nets with -batch_part
// (part_net/net_part.hpp:236-379) and part_xdim (:167-235).  The per-bucket
// Boost R*-tree "intersects" query (:158-166, closed boxes) is replaced by a row
// bitmap of the cells covered by accepted boxes: boxes of one batch are
// pairwise disjoint and buckets are disjoint in x, so a box intersects an
// accepted box exactly when one of its cells is covered.
int gr_partition_nets(int32_t X, int32_t Y, int32_t Z, int32_t n_nets, const int64_t* pin_pfx,
                      const int32_t* pin_enc, int32_t num_xdim_parts,
                      int32_t* batch_ofst, int32_t* net_ids_out) {
  (void)Z;
  if (X <= 0 || Y <= 0 || n_nets < 0 || num_xdim_parts < 1 || !pin_pfx || !pin_enc || !batch_ofst || !net_ids_out)
    return GR_EINVAL;
  std::vector<BBox> bx(n_nets);
  std::vector<int> deg(n_nets);
  for (int i = 0; i < n_nets; ++i) {
    BBox b{X, -1, Y, -1};
    for (int64_t p = pin_pfx[i]; p < pin_pfx[i + 1]; ++p) {
      int64_t v = pin_enc[p];
      int z = (int)(v / ((int64_t)X * Y));
      int x = (z & 1) ? (int)(v / Y % X) : (int)(v % X);
      int y = (z & 1) ? (int)(v % Y) : (int)(v / X % Y);
      b.minx = std::min(b.minx, x); b.maxx = std::max(b.maxx, x);
      b.miny = std::min(b.miny, y); b.maxy = std::max(b.maxy, y);
    }
    bx[i] = b;
    deg[i] = (int)(pin_pfx[i + 1] - pin_pfx[i]);
  }
  const int words = (X + 63) / 64;
  std::vector<uint64_t> occ((size_t)words * Y, 0);
  auto row_hit = [&](int y, int x0, int x1) {
    const uint64_t* row = occ.data() + (size_t)y * words;
    int w0 = x0 >> 6, w1 = x1 >> 6;
    for (int w = w0; w <= w1; ++w) {
      uint64_t m = ~0ull;
      if (w == w0) m &= ~0ull << (x0 & 63);
      if (w == w1) m &= ~0ull >> (63 - (x1 & 63));
      if (row[w] & m) return true;
    }
    return false;
  };
  auto row_set = [&](int y, int x0, int x1, bool on) {
    uint64_t* row = occ.data() + (size_t)y * words;
    int w0 = x0 >> 6, w1 = x1 >> 6;
    for (int w = w0; w <= w1; ++w) {
      uint64_t m = ~0ull;
      if (w == w0) m &= ~0ull << (x0 & 63);
      if (w == w1) m &= ~0ull >> (63 - (x1 & 63));
      if (on) row[w] |= m; else row[w] &= ~m;
    }
  };

  std::vector<int> ids(n_nets);
  for (int i = 0; i < n_nets; ++i) ids[i] = i;
  std::vector<char> vis(n_nets, 0);
  // net_part.hpp:244-262 (cmp1: minx, then maxx)
  std::sort(ids.begin(), ids.end(), [&](int a, int b) {
    if (bx[a].minx == bx[b].minx) return bx[a].maxx < bx[b].maxx;
    return bx[a].minx < bx[b].minx;
  });
  int rem = n_nets, nb = 0, gp = 0;
  batch_ofst[0] = 0;
  std::vector<int> taken;
  while (rem > 0) {
    taken.clear();
    // part_xdim, net_part.hpp:167-235
    int targ = (rem + num_xdim_parts - 1) / num_xdim_parts;
    std::sort(ids.begin(), ids.begin() + rem, [&](int a, int b) { return bx[a].maxx < bx[b].maxx; });
    std::vector<int> xid(rem, -1);
    std::vector<int> part_len(num_xdim_parts + 1, 0);
    int cmax = -1, cur = 0, npn = 0;
    for (int i = 0; i < rem; ++i) {
      int id = ids[i];
      if (cmax < bx[id].minx) {
        xid[i] = cur; ++npn;
        ++part_len[cur];
      }
      if (npn == targ) {
        npn = 0; ++cur; cmax = bx[id].maxx;
      }
    }
    std::vector<std::vector<int>> pick(cur + 1);
    for (int i = 0; i <= cur; ++i) pick[i].assign(part_len[i], 0);
    for (int i = 0; i < rem; ++i)
      if (xid[i] != -1) {
        int p = xid[i], f = part_len[p]--;
        pick[p][f - 1] = ids[i];
      }
    for (int p = 0; p <= cur; ++p) {
      // cmp3: half perimeter (net_part.hpp:178-182)
      std::sort(pick[p].begin(), pick[p].end(), [&](int a, int b) {
        int a1 = (bx[a].maxx - bx[a].minx + 1) + (bx[a].maxy - bx[a].miny + 1);
        int a2 = (bx[b].maxx - bx[b].minx + 1) + (bx[b].maxy - bx[b].miny + 1);
        return a1 < a2;
      });
      for (int nid : pick[p]) {
        const BBox& b = bx[nid];
        bool hit = false;
        for (int y = b.miny; y <= b.maxy && !hit; ++y) hit = row_hit(y, b.minx, b.maxx);
        if (hit) continue;
        for (int y = b.miny; y <= b.maxy; ++y) row_set(y, b.minx, b.maxx, true);
        taken.push_back(nid);
        vis[nid] = 1;
      }
    }
    for (int nid : taken)
      for (int y = bx[nid].miny; y <= bx[nid].maxy; ++y) row_set(y, bx[nid].minx, bx[nid].maxx, false);
    // net_part.hpp:326-337: move the nets still to route to the front
    {
      int id = 0, fid = 0;
      while (id < rem) {
        if (vis[ids[fid]] == 0) fid++;
        if (id > fid && vis[ids[id]] == 0) {
          std::swap(ids[fid], ids[id]);
          fid++;
        }
        id++;
      }
    }
    rem -= (int)taken.size();
    // in-batch order: (pin count, net id) - the reference's std::sort by pin
    // count after an omp critical append is not reproducible (:347-349)
    std::sort(taken.begin(), taken.end(), [&](int a, int b) {
      if (deg[a] != deg[b]) return deg[a] < deg[b];
      return a < b;
    });
    for (int nid : taken) net_ids_out[gp++] = nid;
    ++nb;
    batch_ofst[nb] = gp;
  }
  return nb;
}

}  // extern "C"
