// Device-side building blocks of the routing hot path (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace grd {

constexpr int kRMin = -256;
constexpr int kLut = 512;  // r in [-256, 255]
constexpr int kSmallDeg = 64;  // nets up to this many pins: one wavefront per net
constexpr int kLargeDeg = 4096;  // LDS-resident tree (36 B per pin) for nets up to this many pins
constexpr int kRowTaskPins = 16;  // pins per workgroup of the first-round row-minimum kernel
constexpr int kMaxDeg = 32767;   // 15-bit pin index of the edge encoding (mst_common.h:25)

struct Geo {
  int X, Y, Z, Z2;
  uint32_t XY;
};

// Grid encoding, util.hpp:38-44 (cells fit in int32: CSIZE < 2^31).
__device__ __forceinline__ uint32_t enc(const Geo& g, int z, int x, int y) {
  return (z & 1) ? (uint32_t)z * g.XY + (uint32_t)x * g.Y + y : (uint32_t)z * g.XY + (uint32_t)y * g.X + x;
}
__device__ __forceinline__ void dec(const Geo& g, uint32_t v, int& z, int& x, int& y) {
  z = (int)(v / g.XY);
  const uint32_t r = v - (uint32_t)z * g.XY;
  // one division by the run length; the results are selected by value (a
  // branch assigning through the references makes the compiler keep them in
  // scratch memory)
  const uint32_t L = (z & 1) ? g.Y : g.X;
  const uint32_t q = r / L, m = r - q * L;
  x = (int)((z & 1) ? q : m);
  y = (int)((z & 1) ? m : q);
}

// Packed pin cell of a pair endpoint table: z in bits 28..31, x in bits
// 14..27, y in bits 0..13 (grids of at most 16,384 x 16,384 GCELLs and 16
// layers).
__host__ __device__ inline bool ends_fit(const Geo& g) { return g.X <= 16384 && g.Y <= 16384 && g.Z <= 16; }
__device__ __forceinline__ uint32_t pack_cell(int z, int x, int y) {
  return ((uint32_t)z << 28) | ((uint32_t)x << 14) | (uint32_t)y;
}
__device__ __forceinline__ void unpack_cell(uint32_t v, int& z, int& x, int& y) {
  z = (int)(v >> 28);
  x = (int)((v >> 14) & 0x3fffu);
  y = (int)(v & 0x3fffu);
}

// e^x in double for cell costs of grids with fp32 capacities.  Cody-Waite
// reduction x = k ln2 + t, |t| <= ln2/2 (W. J. Cody, W. Waite, Software Manual
// for the Elementary Functions, 1980; ln2 split into its nearest double and the
// remainder), then the degree-13 Taylor polynomial of e^t in Horner form with
// fused multiply-adds (truncation < 1e-17 relative), scaled by 2^k.  Every
// step is one IEEE double operation, so host and device give the same bits.
// Arguments above 100 give +inf and below -110 give 0 (both outside fp32's
// finite non-zero range after rounding).
__host__ __device__ inline double cw_exp(double x) {
  if (x > 100.0) return __builtin_inf();
  if (x < -110.0) return 0.0;
  const double kInvLn2 = 1.4426950408889634;
  const double kLn2Hi = 0.6931471805599453;         // nearest double to ln 2
  const double kLn2Lo = 2.3190468138462996e-17;     // ln 2 - kLn2Hi
  const double k = __builtin_rint(x * kInvLn2);
  double t = __builtin_fma(-k, kLn2Hi, x);
  t = __builtin_fma(-k, kLn2Lo, t);
  double p = 1.0 / 6227020800.0;                    // 1/13!
  p = __builtin_fma(p, t, 1.0 / 479001600.0);
  p = __builtin_fma(p, t, 1.0 / 39916800.0);
  p = __builtin_fma(p, t, 1.0 / 3628800.0);
  p = __builtin_fma(p, t, 1.0 / 362880.0);
  p = __builtin_fma(p, t, 1.0 / 40320.0);
  p = __builtin_fma(p, t, 1.0 / 5040.0);
  p = __builtin_fma(p, t, 1.0 / 720.0);
  p = __builtin_fma(p, t, 1.0 / 120.0);
  p = __builtin_fma(p, t, 1.0 / 24.0);
  p = __builtin_fma(p, t, 1.0 / 6.0);
  p = __builtin_fma(p, t, 0.5);
  p = __builtin_fma(p, t, 1.0);
  p = __builtin_fma(p, t, 1.0);
  return __builtin_ldexp(p, (int)k);
}

// Cost of a cell with fp32 remaining capacity r under cost model `model`
// (0: exp(-r/2), bit_util.hpp:13-15; 1: f(r - 1) - f(r) with r - 1 in fp32,
// util.hpp:127-129; 2: exp(1 - r), README.md:10), saturated at 1e30 like the
// integer-capacity table.  For integral r it equals that table's entry.
__host__ __device__ inline float cost_f32(int model, float r) {
  float v;
  if (model == 1) {
    const float r1 = r - 1.0f;
    v = (float)cw_exp(-0.5 * (double)r1) - (float)cw_exp(-0.5 * (double)r);
  } else if (model == 2) {
    v = (float)cw_exp(1.0 - (double)r);
  } else {
    v = (float)cw_exp(-0.5 * (double)r);
  }
  if (!(v < 1e30f)) v = 1e30f;
  return v;
}

// Increment of the base cost f(r) = e^(-r/2) when r drops by d, for the
// non-BIT mode: f(r - d) - f(r) with r - d in fp32 (get_gpu_of_cost_inc,
// util.hpp:127-129), each f rounded to fp32 from cw_exp, saturated at 1e30
// (an infinite or NaN difference saturates too).  d = 1 is a wire cell's
// cost, d = 0.5 half a via charge.
__host__ __device__ inline float inc_cost(float r, float d) {
  const float rd = r - d;
  float v = (float)cw_exp(-0.5 * (double)rd) - (float)cw_exp(-0.5 * (double)r);
  if (!(v < 1e30f)) v = 1e30f;
  return v;
}

// Orderable key of an fp32 cost: -0 and +0 map to the same key, negatives
// below positives, so (key << 32 | index) orders lexicographically by
// (cost, index).
__device__ __forceinline__ uint32_t ord_key(float f) {
  f = f + 0.0f;
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_cost(uint32_t k) {
  uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(u);
}

// Materialised Fenwick prefix Q(z, line, pos) (bit_util.hpp:19-35); pos = -1 -> 0.
__device__ __forceinline__ float Q(const float* __restrict__ pfx, const Geo& g, int z, int line, int pos) {
  if (pos < 0) return 0.0f;
  uint32_t L = (z & 1) ? g.Y : g.X;
  return pfx[(uint32_t)z * g.XY + (uint32_t)line * L + pos];
}

// Single-cell range query Q(pos) - Q(pos - 1) of the via cost at cell
// (z, x, y) (bit_util.hpp:108-117), Q(-1) = 0: both prefix values come from
// one 8-byte load at the cell before (at run position 0 that value belongs to
// another run, or to the padding float before the first cell, and is not
// used; x - 0.0f is x).
typedef float f32x2_a4 __attribute__((ext_vector_type(2), aligned(4)));
__device__ __forceinline__ float via_term(const float* __restrict__ pfx, const Geo& g, int z, int x, int y) {
  const int pos = (z & 1) ? y : x;
  const f32x2_a4 t = *reinterpret_cast<const f32x2_a4*>(pfx + enc(g, z, x, y) - 1);
  return t.y - (pos ? t.x : 0.0f);
}

// Via stack cost, bit_util.hpp:108-117: each layer i in [min, max) adds the
// single-cell range query Q(pos) - Q(pos - 1), in ascending layer order.
__device__ __forceinline__ void via_cost(const float* __restrict__ pfx, const Geo& g, int za, int zb, int x,
                                         int y, float& cost) {
  int lo = za < zb ? za : zb, hi = za < zb ? zb : za;
  for (int i = lo; i < hi; ++i) {
    if (i & 1) cost += Q(pfx, g, i, x, y) - Q(pfx, g, i, x, y - 1);
    else       cost += Q(pfx, g, i, y, x) - Q(pfx, g, i, y, x - 1);
  }
}

// Wire cost along layer z's direction, inclusive of both ends, bit_util.hpp:91-107.
__device__ __forceinline__ void wire_cost(const float* __restrict__ pfx, const Geo& g, int z, int& xs, int& ys,
                                          int x2, int y2, float& cost) {
  if (z & 1) {
    if (ys < y2)      cost += Q(pfx, g, z, xs, y2) - Q(pfx, g, z, xs, ys - 1);
    else if (ys > y2) cost += Q(pfx, g, z, xs, ys) - Q(pfx, g, z, xs, y2 - 1);
    ys = y2;
  } else {
    if (xs < x2)      cost += Q(pfx, g, z, ys, x2) - Q(pfx, g, z, ys, xs - 1);
    else if (xs > x2) cost += Q(pfx, g, z, ys, xs) - Q(pfx, g, z, ys, x2 - 1);
    xs = x2;
  }
}

// Cost of pattern p for the pair src -> dst (bit_util.hpp:131-179).  Returns
// false for an invalid pattern (a = 0, b' >= Z, or not ending at dst).
__device__ __forceinline__ bool pattern_cost(const float* __restrict__ pfx, const Geo& g, int z1, int x1, int y1,
                                             int z2, int x2, int y2, int p, float& out) {
  int a = p / g.Z2, b = p - a * g.Z2;
  if (a == 0) return false;
  bool D = (x1 != x2) && (y1 != y2);
  int zs = z1, xs = x1, ys = y1;
  float cost = 0.0f;
  if (zs != a) { via_cost(pfx, g, zs, a, xs, ys, cost); zs = a; }
  wire_cost(pfx, g, zs, xs, ys, x2, y2, cost);
  int bp = (zs & 1) ? 2 * b : 2 * b + 1;
  if (D && bp != 0) {
    if (bp >= g.Z) return false;
    via_cost(pfx, g, zs, bp, xs, ys, cost);
    zs = bp;
    wire_cost(pfx, g, zs, xs, ys, x2, y2, cost);
  }
  if (zs != z2) { via_cost(pfx, g, zs, z2, xs, ys, cost); zs = z2; }
  if (xs != x2 || ys != y2) return false;
  out = cost;
  return true;
}

// Pair index o = j(j-1)/2 + k, k < j (test_lpat.hpp:316-331) -> (j, k).
__device__ __forceinline__ void pair_decode(uint32_t o, int& j, int& k) {
  int jj = (int)((1.0f + sqrtf(1.0f + 8.0f * (float)o)) * 0.5f);
  while ((uint64_t)jj * (jj - 1) / 2 > o) --jj;
  while ((uint64_t)(jj + 1) * jj / 2 <= o) ++jj;
  j = jj;
  k = (int)(o - (uint32_t)((uint64_t)jj * (jj - 1) / 2));
}

// Single-cell via term at (x, y) on layer i: Q(pos) - Q(pos - 1), bit_util.hpp:108-117.
__device__ __forceinline__ float dterm(const float* __restrict__ pfx, const Geo& g, int i, int x, int y) {
  return (i & 1) ? Q(pfx, g, i, x, y) - Q(pfx, g, i, x, y - 1) : Q(pfx, g, i, y, x) - Q(pfx, g, i, y, x - 1);
}
// Inclusive wire u..v on a run: Q(hi) - Q(lo - 1), bit_util.hpp:91-107.
__device__ __forceinline__ float wterm(const float* __restrict__ pfx, const Geo& g, int z, int line, int u, int v) {
  const int lo = u < v ? u : v, hi = u < v ? v : u;
  return Q(pfx, g, z, line, hi) - Q(pfx, g, z, line, lo - 1);
}

// Pair w of the batch -> batch position of the net that owns it, with that
// position's first pair and first pin.  chunk_net[c] is the owner of pair
// 64*c; a net has >= 1 pair, so the owner of any pair of chunk c is among the
// 64 positions starting there.  The wavefront of chunk c loads that window of
// pair and pin offsets once (coalesced) and each lane binary searches it
// through lane shuffles; the owner's offsets come from the window as well, so
// no load waits on the search.  Every lane of the wave must call this.
struct PairOwner {
  int pos;
  int64_t first_pair, first_pin;
};
__device__ __forceinline__ int64_t shfl_i64(int64_t v, int src) {
  return (int64_t)(((uint64_t)(uint32_t)__shfl((int)(v >> 32), src, 64) << 32) | (uint32_t)__shfl((int)v, src, 64));
}
__device__ __forceinline__ PairOwner pair_owner_wave(const int64_t* __restrict__ bpair,
                                                     const int64_t* __restrict__ pinofs,
                                                     const int32_t* __restrict__ chunk_net, int n, int64_t w) {
  const int lane = threadIdx.x & 63;
  const int base = chunk_net[w >> 6];
  const bool in = base + lane < n;
  const int64_t v = in ? bpair[base + lane] : INT64_MAX;
  const int64_t pv = in ? pinofs[base + lane] : 0;
  int lo = 0, hi = 64;  // v[lo] <= w < v[hi]
#pragma unroll
  for (int it = 0; it < 6; ++it) {
    const int mid = (lo + hi) >> 1;
    if (shfl_i64(v, mid) <= w) lo = mid; else hi = mid;
  }
  return PairOwner{base + lo, shfl_i64(v, lo), shfl_i64(pv, lo)};
}

// Lexicographic minimum key (ord_key(cost) << 32 | p) over every pattern of
// the pair (z1, x1, y1) -> (z2, x2, y2), for one lane; g.Z must equal ZT.
// All via and wire terms of the pair are loaded once into registers; patterns
// are then evaluated in ascending p with exactly the oracle's fp32 operation
// sequence (bit_util.hpp:131-179):
//   cost = 0; += via cells z1..a (ascending); += wire on a;
//   [D and b' != 0: += via cells a..b' at the corner (ascending); += wire on b'];
//   += via cells to z2 at the destination (ascending).
// The source via sum for a > z1 is a running sum shared across a.  For
// straight pairs (not D) all b tie, so only b = 0 is evaluated.  Selection is
// strict-less in ascending p, i.e. the lexicographic minimum of (cost, p).
//
// SEL: only patterns whose first wire layer a has parity par count (the
// caller takes the other parity in another lane), only candidates among them
// (bit q = p - Z2 of mask, q < 64) unless full, and only the terms those
// patterns use are loaded (a lane's skipped loads make no memory request).  A
// straight pair's candidates (a, b) all cost what (a, 0) costs, so the lowest
// candidate p of that a stands for them.
template <int ZT>
__device__ __forceinline__ constexpr uint64_t parity_patterns(int par) {
  uint64_t m = 0;
  for (int q = 0; q < (ZT - 1) * ((ZT + 1) / 2); ++q)
    if (((q / ((ZT + 1) / 2) + 1) & 1) == par) m |= 1ull << q;
  return m;
}

// The terms of one pin pair that every pattern is built from: single-cell
// via terms at the source, destination and the two corners (layers 0..Z-2),
// and the first / second wire term on each layer 1..Z-1.
template <int ZT>
struct PairTerms {
  float ds[ZT - 1], dd[ZT - 1], dc1[ZT - 1], dc2[ZT - 1], w1[ZT], w2[ZT];
};

// The pattern loop over loaded terms (dc1, dc2 and w2 are read only for bent
// pairs, w1 only on layers the first wire moves on); mask as in lane_best.
template <int ZT, bool SEL>
__device__ __forceinline__ uint64_t lane_select(const PairTerms<ZT>& t, int z1, int x1, int y1, int z2, int x2,
                                                int y2, uint64_t mask) {
  constexpr int Z2T = (ZT + 1) / 2;
  const bool D = (x1 != x2) && (y1 != y2);
  // Destination via stack from the last wire layer zs (a compile-time value
  // in the unrolled loops below) to z2: cells z2..zs-1 going down, zs..z2-1
  // going up, added in ascending layer order.  ddn[i] = t.dd[i] for i >= z2 and
  // dup[i] = t.dd[i] for i < z2, +0.0f elsewhere, so each pattern adds ddn[0..zs)
  // and, only when z2 > zs, dup[zs..).  The running cost is never -0.0, so
  // adding +0.0f leaves it unchanged and the sums equal the cell-by-cell ones.
  float ddn[ZT - 1], dup[ZT - 1];
#pragma unroll
  for (int i = 0; i < ZT - 1; ++i) {
    ddn[i] = i >= z2 ? t.dd[i] : 0.f;
    dup[i] = i < z2 ? t.dd[i] : 0.f;
  }
  // (fixed trip counts with conditions on i, so that every loop unrolls
  // fully once zs is known: no register-indexed loops)
  auto to_dst = [&](float& c, const int zs) {
#pragma unroll
    for (int i = 0; i < ZT - 1; ++i)
      if (i < zs) c += ddn[i];
    if (z2 > zs) {
#pragma unroll
      for (int i = 0; i < ZT - 1; ++i)
        if (i >= zs) c += dup[i];
    }
  };

  // running minimum: strict-less on the cost in ascending p keeps the lowest
  // p among equal costs, the order of (ord_key(cost), p) (costs are finite
  // and not NaN; -0.0 and +0.0 compare equal in both)
  float bc = __builtin_inff();
  int bpat = -1;
  float vs = 0.f;  // running source via sum for a > z1
#pragma unroll
  for (int a = 1; a < ZT; ++a) {
    float ca = 0.f;
    if (a > z1) {
      vs = (a == z1 + 1) ? t.ds[a - 1] : vs + t.ds[a - 1];
      ca = vs;
    } else if (a < z1) {
#pragma unroll
      for (int i = 0; i < ZT - 1; ++i)
        if (i >= a && i < z1) ca += t.ds[i];
    }
    const bool moved1 = (a & 1) ? (y1 != y2) : (x1 != x2);
    if (moved1) ca += t.w1[a];
    if (D) {
      // second wire on b' (b' = 0 leaves both coordinates to cover: invalid)
#pragma unroll
      for (int b = 0; b < Z2T; ++b) {
        const int bp = (a & 1) ? 2 * b : 2 * b + 1;
        if (bp == 0 || bp >= ZT) continue;
        if (SEL && !((mask >> ((a - 1) * Z2T + b)) & 1)) continue;  // also skips the other parity
        float c = ca;
        const int lo = a < bp ? a : bp, hi = a < bp ? bp : a;
#pragma unroll
        for (int i = 0; i < ZT - 1; ++i)
          if (i >= lo && i < hi) c += (a & 1) ? t.dc2[i] : t.dc1[i];
        c += t.w2[bp];
        to_dst(c, bp);
        const bool take = c < bc;
        bc = take ? c : bc;
        bpat = take ? a * Z2T + b : bpat;
      }
    } else {
      // straight pair: every b gives the same route, b = 0 is the lowest p
      const bool at_dst = (a & 1) ? (x1 == x2) : (y1 == y2);
      int b0 = 0;
      if (SEL) {
        const uint64_t sub = (mask >> ((a - 1) * Z2T)) & ((1ull << Z2T) - 1);
        b0 = sub ? __builtin_ctzll(sub) : -1;
      }
      if (at_dst && b0 >= 0) {
        float c = ca;
        to_dst(c, a);
        const bool take = c < bc;
        bc = take ? c : bc;
        bpat = take ? a * Z2T + b0 : bpat;
      }
    }
  }
  return bpat < 0 ? ~0ull : (((uint64_t)ord_key(bc) << 32) | (uint32_t)bpat);
}

// vt: the via terms stored per cell (the refresh writes them while a learned
// proposal is set), or null to form them from the prefix.
template <int ZT, bool SEL = false>
__device__ __forceinline__ uint64_t lane_best(const Geo& g, const float* __restrict__ pfx, int z1, int x1, int y1, int z2,
                                              int x2, int y2, uint64_t mask = ~0ull, int par = 0,
                                              bool full = true, const float* __restrict__ vt = nullptr) {
  auto via_term_at = [&](int z, int x, int y) { return vt ? vt[enc(g, z, x, y)] : via_term(pfx, g, z, x, y); };
  constexpr int Z2T = (ZT + 1) / 2;
  static_assert(!SEL || (ZT - 1) * Z2T <= 64, "candidate masks hold 64 patterns");
  const bool D = (x1 != x2) && (y1 != y2);
  if (SEL) mask = (full ? ~0ull : mask) & (par ? parity_patterns<ZT>(1) : parity_patterns<ZT>(0));

  // terms the patterns use: bit i of nds / ndd / ndc1 / ndc2 (via cells of
  // layer i), bit l of nw1 / nw2 (wires on layer l)
  uint32_t nds = ~0u, ndd = ~0u, ndc1 = ~0u, ndc2 = ~0u, nw1 = ~0u, nw2 = ~0u;
  if (SEL && full) {  // every pattern of the parity
    constexpr uint32_t odd = 0xAAAAAAAAu;
    nw1 = par ? odd : ~odd;
    nw2 = par ? ~odd : odd;
    ndc1 = par ? 0u : ~0u;
    ndc2 = par ? ~0u : 0u;
  } else if (SEL) {
    nds = ndd = ndc1 = ndc2 = nw1 = nw2 = 0;
    auto rng = [](int u, int v) {  // bits [min, max)
      const int lo = u < v ? u : v, hi = u < v ? v : u;
      return ((1u << hi) - 1u) & ~((1u << lo) - 1u);
    };
    for (uint64_t m = mask; m; m &= m - 1) {
      const int q = __builtin_ctzll(m), a = q / Z2T + 1, b = q - (a - 1) * Z2T;
      int zl = a;
      if (D) {
        const int bp = (a & 1) ? 2 * b : 2 * b + 1;
        if (bp == 0 || bp >= ZT) continue;
        if (a & 1) ndc2 |= rng(a, bp); else ndc1 |= rng(a, bp);
        nw2 |= 1u << bp;
        zl = bp;
      } else if ((a & 1) ? (x1 != x2) : (y1 != y2)) {
        continue;  // the first wire does not reach the destination
      }
      nds |= rng(z1, a);
      nw1 |= 1u << a;
      ndd |= rng(zl, z2);
    }
  }

  PairTerms<ZT> t;
#pragma unroll
  for (int i = 0; i < ZT - 1; ++i) {
    t.ds[i] = ((nds >> i) & 1) ? via_term_at(i, x1, y1) : 0.f;
    t.dd[i] = ((ndd >> i) & 1) ? via_term_at(i, x2, y2) : 0.f;
  }
#pragma unroll
  for (int l = 1; l < ZT; ++l) {
    // first wire on layer l from the source: rows y1 on even layers, columns x1 on odd
    const bool moved = (l & 1) ? (y1 != y2) : (x1 != x2);
    t.w1[l] = (moved && ((nw1 >> l) & 1)) ? ((l & 1) ? wterm(pfx, g, l, x1, y1, y2) : wterm(pfx, g, l, y1, x1, x2))
                                        : 0.f;
  }
  if (D) {
#pragma unroll
    for (int i = 0; i < ZT - 1; ++i) {
      t.dc1[i] = ((ndc1 >> i) & 1) ? via_term_at(i, x2, y1) : 0.f;  // corner after a horizontal first wire
      t.dc2[i] = ((ndc2 >> i) & 1) ? via_term_at(i, x1, y2) : 0.f;  // corner after a vertical first wire
    }
#pragma unroll
    for (int l = 1; l < ZT; ++l) {
      // second wire into the destination: columns x2 on odd layers, rows y2 on even
      t.w2[l] = ((nw2 >> l) & 1) ? ((l & 1) ? wterm(pfx, g, l, x2, y1, y2) : wterm(pfx, g, l, y2, x1, x2)) : 0.f;
    }
  }
  return lane_select<ZT, SEL>(t, z1, x1, y1, z2, x2, y2, mask);
}

// The prefix pair (Q(pos - 1), Q(pos)) at cell (z, x, y), one 8-byte load;
// Q(pos - 1) is never used at pos = 0 (it belongs to another run, or to the
// padding float before the first cell).
__device__ __forceinline__ f32x2_a4 qpair(const float* __restrict__ pfx, const Geo& g, int z, int x, int y) {
  return *reinterpret_cast<const f32x2_a4*>(pfx + enc(g, z, x, y) - 1);
}
// Single-cell via term Q(pos) - Q(pos - 1), Q(-1) = 0, from its prefix pair.
__device__ __forceinline__ float via_of(f32x2_a4 t, int pos) { return t.y - (pos ? t.x : 0.0f); }
// Inclusive wire Q(hi) - Q(lo - 1), Q(-1) = 0, between run positions pu != pv
// from the prefix pairs at its two ends.
__device__ __forceinline__ float wire_of(f32x2_a4 u, int pu, f32x2_a4 v, int pv) {
  return pu < pv ? v.y - (pu ? u.x : 0.0f) : u.y - (pv ? v.x : 0.0f);
}

// lane_best<ZT> (every pattern, prefix in pfx) with the terms loaded per
// point instead of per term.  Every term of the pair lies at one of four
// GCELLs, source s = (x1, y1), destination d = (x2, y2) and the corners
// c1 = (x2, y1), c2 = (x1, y2), and each needs Q(pos) and Q(pos - 1) at one
// cell of it on some layer:
//   via terms at s, d, c1, c2 (layers 0..Z-2);
//   the first wire on layer l runs from s to c1 (even l, row y1) or to c2
//     (odd l, column x1): its ends are s's and that corner's cells on l;
//   the second wire on layer l runs from c2 (even l, row y2) or c1 (odd l,
//     column x2) to d.
// So one 8-byte load per point and layer, 4Z for a bent pair and 2Z for a
// straight one (whose moving wires end at d), replaces 8(Z-1) loads of single
// terms; every term is formed from the same two prefix values by the same
// fp32 subtraction, so the results are those of lane_best.
//
// BUF: the pairs are read through a buffer resource whose byte 0 is the
// padding float, so cell e's pair starts at byte 4e: the point's run offset
// is a VGPR and the layer's offset l*XY*4 an SGPR (prefixes of fewer than
// 2^30 cells).
template <int ZT, bool BUF = false>
__device__ __forceinline__ uint64_t lane_best_pts(const Geo& g, const float* __restrict__ pfx, int z1, int x1, int y1,
                                                  int z2, int x2, int y2) {
  const bool mx = x1 != x2, my = y1 != y2, D = mx && my;
  constexpr bool top_odd = ((ZT - 1) & 1) != 0;
  // the top layer carries no via term: its cells are needed for wires only
  const bool top_moves = top_odd ? my : mx;
  __amdgpu_buffer_rsrc_t rs;
  if constexpr (BUF)
    rs = __builtin_amdgcn_make_buffer_rsrc((void*)(pfx - 1), (short)0, (int)((g.XY * (uint32_t)g.Z + 1u) * 4u),
                                           0x00020000);
  auto ld = [&](int l, int x, int y) -> f32x2_a4 {
    if constexpr (BUF) {
      typedef float f32x2 __attribute__((ext_vector_type(2)));
      const uint32_t off = (l & 1) ? ((uint32_t)x * g.Y + (uint32_t)y) * 4u : ((uint32_t)y * g.X + (uint32_t)x) * 4u;
      const f32x2 v = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)off, (int)((uint32_t)l * g.XY * 4u), 0);
      f32x2_a4 t;
      t.x = v.x;
      t.y = v.y;
      return t;
    } else {
      return qpair(pfx, g, l, x, y);
    }
  };
  f32x2_a4 s[ZT], d[ZT], c1[ZT], c2[ZT];
#pragma unroll
  for (int l = 0; l < ZT; ++l) {
    if (l < ZT - 1 || top_moves) s[l] = ld(l, x1, y1);
    if (l < ZT - 1 || top_moves || D) d[l] = ld(l, x2, y2);
  }
  if (D) {
#pragma unroll
    for (int l = 0; l < ZT; ++l) {
      c1[l] = ld(l, x2, y1);
      c2[l] = ld(l, x1, y2);
    }
  }
  PairTerms<ZT> t;
#pragma unroll
  for (int i = 0; i < ZT - 1; ++i) {
    t.ds[i] = via_of(s[i], (i & 1) ? y1 : x1);
    t.dd[i] = via_of(d[i], (i & 1) ? y2 : x2);
  }
#pragma unroll
  for (int l = 1; l < ZT; ++l) {
    if (l & 1) t.w1[l] = my ? wire_of(s[l], y1, D ? c2[l] : d[l], y2) : 0.f;
    else       t.w1[l] = mx ? wire_of(s[l], x1, D ? c1[l] : d[l], x2) : 0.f;
  }
  if (D) {
#pragma unroll
    for (int i = 0; i < ZT - 1; ++i) {
      t.dc1[i] = via_of(c1[i], (i & 1) ? y1 : x2);
      t.dc2[i] = via_of(c2[i], (i & 1) ? y2 : x1);
    }
#pragma unroll
    for (int l = 1; l < ZT; ++l) {
      if (l & 1) t.w2[l] = wire_of(c1[l], y1, d[l], y2);  // column x2
      else       t.w2[l] = wire_of(c2[l], x1, d[l], x2);  // row y2
    }
  }
  return lane_select<ZT, false>(t, z1, x1, y1, z2, x2, y2, ~0ull);
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  lo = __shfl_xor(lo, m, 64);
  hi = __shfl_xor(hi, m, 64);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    uint64_t o = shfl_xor_u64(v, m);
    v = o < v ? o : v;
  }
  return v;
}

// Net classes handled by one k_tree_small launch: class 0 = 2 pins (one lane
// per net, 256 nets per block); classes 1..4 = up to 8/16/32/64 pins (a group
// of G lanes per net, 256/G nets per block).  Block ranges per class.
struct SmallClasses {
  const int32_t* list[5];
  int n[5];
  int block0[6];  // first block of each class; block0[5] = grid size
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace grd
