// Kernels of the non-BIT ("prefix") mode of the routing hot path: the
// reference's path with -bit_pmodel off.  Costing reads an exclusive prefix
// of per-cell wire costs along each run (LPATTERN::get_gpu_lpat_dist,
// l_pattern_files/lpattern.hpp:13-97, and its CPU twin :98-160) instead of
// Fenwick trees; the commit takes 1.0 of capacity from every cell a wire of
// the batch covers and half units for vias (MST_COMMIT::gpu_commit_edges,
// mst_files/mst_commit.hpp:342-364).
//
// Device state (WIRE_ENC order, as the BIT mode's):
//   remf[cell]   fp32  remaining capacity r (the state; via charges make it
//                      half-integral)
//   pfx[cell]    fp32  P: exclusive prefix of the wire costs w = f(r-1) - f(r)
//                      along the run, in the chunked order of the reference's
//                      scan kernel (d_get_epfx_cost_cube_inplace,
//                      util.hpp:277-353, 128 scanning threads over max(X, Y)
//                      cells): chunk t = [tT, tT + T), T = ceil(max(X,Y)/128),
//                      starts its local scan from the raw cost of the cell
//                      before it, and P[i] = P[tT - 1] + local(i) for t >= 1
//   vterm[cell]  fp32  V: via cost of the layer at the cell, the increase of f
//                      over exactly the capacity the commit's via charge takes
//                      there (run position 0: f(r-1) - f(r); last position:
//                      the same of the previous cell; otherwise half of each,
//                      previous cell first).  The reference's via cube is
//                      written by racing threads at transposed VIA_ENC slots
//                      (util.hpp:147-193 with :41-42) and has no defined
//                      content; this is the contract that replaces it.
//   mark[cell]   uint8 1 when a wire of this batch covers the cell
//   vhalf[cell]  int32 via half charges of this batch
//   run_stamp          batch stamp of the last commit that touched a run
#pragma once
#include "gr_kernels.h"

namespace grd {

// Cost state of the runs the batch touched (all != 0: every run, no
// decrements).  One workgroup per run at a time; LDS holds the run's
// remaining capacities, then its local scans (A), the wire costs (W), the
// half-via costs (H) and the chunk-end prefixes (C).
__global__ __launch_bounds__(256) void k_refresh_prefix(Geo g, int M, float* __restrict__ remf,
                                                        uint8_t* __restrict__ mark, int32_t* __restrict__ vhalf,
                                                        float* __restrict__ pfx, float* __restrict__ vterm,
                                                        const int32_t* __restrict__ run_stamp,
                                                        const int32_t* __restrict__ ctr, int all,
                                                        const int32_t* __restrict__ rl, int nrl) {
  extern __shared__ float sh[];
  float* A = sh;
  float* W = A + M;
  float* H = W + M;
  float* C = H + M;  // 128 chunk-end prefixes
  const int32_t bstamp = all ? 0 : *ctr;
  const int T = (M + 127) / 128;
  const int nr = rl ? nrl : g.Z * M;  // every run, or the batch's candidate list
  for (int ri = blockIdx.x; ri < nr; ri += gridDim.x) {
    const int r = rl ? rl[ri] : ri;
    const int z = r / M, line = r - z * M;
    const int lines = (z & 1) ? g.X : g.Y;
    if (line >= lines) continue;
    if (!all && run_stamp[r] != bstamp) continue;
    const int L = (z & 1) ? g.Y : g.X;
    const uint32_t base = (uint32_t)z * g.XY + (uint32_t)line * L;
    // this batch's decrements: 1.0 under a wire, then the via halves in one
    // fp32 step (mst_commit.hpp:329-341, then :17-61)
    for (int i = threadIdx.x; i < L; i += blockDim.x) {
      float rv = remf[base + i];
      if (!all) {
        const bool wired = mark[base + i] != 0;
        const int32_t h = vhalf[base + i];
        if (wired) rv -= 1.0f;
        if (h) rv -= 0.5f * (float)h;
        if (wired || h) {
          remf[base + i] = rv;
          mark[base + i] = 0;
          vhalf[base + i] = 0;
        }
      }
      A[i] = rv;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < L; i += blockDim.x) {
      const float rv = A[i];
      W[i] = inc_cost(rv, 1.0f);
      H[i] = inc_cost(rv, 0.5f);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < L; i += blockDim.x)
      vterm[base + i] = i == 0 ? W[0] : (i == L - 1 ? W[i - 1] : H[i - 1] + H[i]);
    // local scans: thread t takes chunk t (A is free again)
    if ((int)threadIdx.x < 128) {
      const int t = threadIdx.x, lo = t * T, hi = min(L, lo + T);
      if (lo < L) {
        float local = t ? W[lo - 1] : 0.0f;
        for (int i = lo; i < hi; ++i) {
          A[i] = local;
          local += W[i];
        }
      }
    }
    __syncthreads();
    // chunk ends in order: P[(t+1)T - 1] = P[tT - 1] + local
    if (threadIdx.x == 0) {
      float e = 0.0f;
      for (int t = 0; t * T < L; ++t) {
        const int last = min(L, (t + 1) * T) - 1;
        e = t ? e + A[last] : A[last];
        C[t] = e;
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < L; i += blockDim.x) {
      const int t = i / T;
      pfx[base + i] = t ? C[t - 1] + A[i] : A[i];
    }
    __syncthreads();
  }
}

// Pattern cost of the non-BIT mode for one (pair, pattern), read from global
// memory (lpattern.hpp:98-160): first via stack over layers [min, max), wires
// |P(end) - P(start)|, corner and final stacks over the open interval; b' >= Z
// invalid; the second leg whenever the first wire does not end at the
// destination cell; cost = wires + vias in fp32.
__device__ inline bool pattern_cost_prefix(const float* __restrict__ P, const float* __restrict__ V, const Geo& g,
                                           int z1, int x1, int y1, int z2, int x2, int y2, int p, float& out) {
  const int a = p / g.Z2;
  int bp = p - a * g.Z2;
  if (a == 0) return false;
  float wc = 0.0f, vc = 0.0f;
  int zs = z1, xs = x1, ys = y1;
  auto wire = [&](int z) {
    if (z & 1) { wc += fabsf(P[enc(g, z, xs, y2)] - P[enc(g, z, xs, ys)]); ys = y2; }
    else       { wc += fabsf(P[enc(g, z, x2, ys)] - P[enc(g, z, xs, ys)]); xs = x2; }
  };
  if (zs != a) {
    const int lo = min(zs, a), hi = max(zs, a);
    for (int i = lo; i < hi; ++i) vc += V[enc(g, i, xs, ys)];
    zs = a;
  }
  wire(zs);
  bp = (zs & 1) ? 2 * bp : 2 * bp + 1;
  if (bp >= g.Z) return false;
  if (zs != z2 || xs != x2 || ys != y2) {
    if (bp > 0) {
      const int lo = min(zs, bp), hi = max(zs, bp);
      for (int i = lo + 1; i < hi; ++i) vc += V[enc(g, i, xs, ys)];
      zs = bp;
      wire(zs);
    }
    if (zs != z2) {
      const int lo = min(zs, z2), hi = max(zs, z2);
      for (int i = lo + 1; i < hi; ++i) vc += V[enc(g, i, xs, ys)];
      zs = z2;
    }
  }
  if (xs != x2 || ys != y2 || zs != z2) return false;
  out = wc + vc;
  return true;
}

// Lexicographic (cost, p) minimum over every pattern of one pair, one lane:
// all terms the patterns use are loaded once into registers (via costs at
// the source, both corners and the destination; both ends of every layer's
// first and second wire), then patterns are evaluated in ascending p with the
// fp32 operation sequence of pattern_cost_prefix.  g.Z must equal ZT.
template <int ZT>
__device__ __forceinline__ uint64_t lane_best_prefix(const Geo& g, const float* __restrict__ P,
                                                     const float* __restrict__ V, int z1, int x1, int y1, int z2,
                                                     int x2, int y2) {
  constexpr int Z2T = (ZT + 1) / 2;
  float vs[ZT], vce[ZT], vco[ZT], vd[ZT], w1[ZT], w2[ZT];
  // the first stack spans layers [min, max) <= Z-2; corner and final stacks
  // the open interval, layers 1..Z-2
#pragma unroll
  for (int i = 0; i < ZT; ++i) {
    const bool inner = i >= 1 && i < ZT - 1;
    vs[i] = i < ZT - 1 ? V[enc(g, i, x1, y1)] : 0.f;
    vce[i] = inner ? V[enc(g, i, x2, y1)] : 0.f;  // after a first wire along x (even layer)
    vco[i] = inner ? V[enc(g, i, x1, y2)] : 0.f;  // after a first wire along y (odd layer)
    vd[i] = inner ? V[enc(g, i, x2, y2)] : 0.f;
  }
  w1[0] = w2[0] = 0.f;
#pragma unroll
  for (int l = 1; l < ZT; ++l) {
    if (l & 1) {
      w1[l] = y1 != y2 ? fabsf(P[enc(g, l, x1, y2)] - P[enc(g, l, x1, y1)]) : 0.f;
      w2[l] = y1 != y2 ? fabsf(P[enc(g, l, x2, y2)] - P[enc(g, l, x2, y1)]) : 0.f;
    } else {
      w1[l] = x1 != x2 ? fabsf(P[enc(g, l, x2, y1)] - P[enc(g, l, x1, y1)]) : 0.f;
      w2[l] = x1 != x2 ? fabsf(P[enc(g, l, x2, y2)] - P[enc(g, l, x1, y2)]) : 0.f;
    }
  }
  uint64_t best = ~0ull;
#pragma unroll
  for (int a = 1; a < ZT; ++a) {
    // first via stack [min(z1, a), max) at the source, ascending
    float v0 = 0.0f;
#pragma unroll
    for (int i = 0; i < ZT; ++i)
      if (i >= min(z1, a) && i < max(z1, a)) v0 += vs[i];
    const bool odd = a & 1;
    // position after the first wire
    const int px = odd ? x1 : x2, py = odd ? y2 : y1;
    const bool at_xy = px == x2 && py == y2;
#pragma unroll
    for (int b = 0; b < Z2T; ++b) {
      const int bp = odd ? 2 * b : 2 * b + 1;
      if (bp >= ZT) continue;
      float wc = 0.0f + w1[a];
      float vc = v0;
      int zs = a;
      bool at = at_xy;
      if (!(at_xy && a == z2)) {
        if (bp > 0) {
#pragma unroll
          for (int i = 0; i < ZT; ++i)
            if (i > min(a, bp) && i < max(a, bp)) vc += odd ? vco[i] : vce[i];
          zs = bp;
          wc += w2[bp];
          at = true;  // the second wire covers the remaining coordinate
        }
        if (zs != z2) {
#pragma unroll
          for (int i = 0; i < ZT; ++i)
            if (i > min(zs, z2) && i < max(zs, z2)) vc += vd[i];
        }
      }
      if (!at) continue;
      const float c = wc + vc;
      const uint64_t key = ((uint64_t)ord_key(c) << 32) | (uint32_t)(a * Z2T + b);
      best = key < best ? key : best;
    }
  }
  return best;
}

template <int ZT>
__global__ __launch_bounds__(256) void k_lpat_prefix(Geo g, const float* __restrict__ P, const float* __restrict__ V,
                                                     const int32_t* __restrict__ pins,
                                                     const int64_t* __restrict__ pinofs,
                                                     const int64_t* __restrict__ bpair,
                                                     const int32_t* __restrict__ chunk_net, int n, int64_t Pn,
                                                     float* __restrict__ out_cost, int16_t* __restrict__ out_pat,
                                                     int32_t* __restrict__ ctr) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *ctr += 1;  // this batch's stamp
  const int64_t w0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t w = w0 < Pn ? w0 : Pn - 1;
  const PairOwner own = pair_owner_wave(bpair, pinofs, chunk_net, n, w);
  if (w0 >= Pn) return;
  int j, k;
  pair_decode((uint32_t)(w - own.first_pair), j, k);
  int z1, x1, y1, z2, x2, y2;
  dec(g, (uint32_t)pins[own.first_pin + j], z1, x1, y1);
  dec(g, (uint32_t)pins[own.first_pin + k], z2, x2, y2);
  const uint64_t best = lane_best_prefix<ZT>(g, P, V, z1, x1, y1, z2, x2, y2);
  out_cost[w] = key_cost((uint32_t)(best >> 32));
  out_pat[w] = (int16_t)(uint32_t)best;
}

// Any Z: one wavefront per pair, one lane per pattern.
__global__ __launch_bounds__(256) void k_lpat_prefix_wave(Geo g, const float* __restrict__ P,
                                                          const float* __restrict__ V,
                                                          const int32_t* __restrict__ pins,
                                                          const int64_t* __restrict__ pinofs,
                                                          const int64_t* __restrict__ bpair,
                                                          const int32_t* __restrict__ chunk_net, int n, int64_t Pn,
                                                          float* __restrict__ out_cost, int16_t* __restrict__ out_pat,
                                                          int32_t* __restrict__ ctr) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *ctr += 1;
  const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= Pn) return;
  int pos = chunk_net[w >> 6];
  while (pos + 1 < n && bpair[pos + 1] <= w) ++pos;
  int j, k;
  pair_decode((uint32_t)(w - bpair[pos]), j, k);
  const int64_t pb = pinofs[pos];
  int z1, x1, y1, z2, x2, y2;
  dec(g, (uint32_t)pins[pb + j], z1, x1, y1);
  dec(g, (uint32_t)pins[pb + k], z2, x2, y2);
  uint64_t best = ~0ull;
  for (int p = lane; p < g.Z * g.Z2; p += 64) {
    float c;
    if (pattern_cost_prefix(P, V, g, z1, x1, y1, z2, x2, y2, p, c)) {
      const uint64_t key = ((uint64_t)ord_key(c) << 32) | (uint32_t)p;
      best = key < best ? key : best;
    }
  }
  best = wave_min_u64(best);
  if (lane == 0) {
    out_cost[w] = key_cost((uint32_t)(best >> 32));
    out_pat[w] = (int16_t)(uint32_t)best;
  }
}

}  // namespace grd
