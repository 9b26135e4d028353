// Route evaluator (gfx950): per-cell wire and via counts of a routed
// schedule and the per-layer overflow figures of the reference's end-of-run
// report (main.cpp:103-135).
//
// Counting rules are those of the reference's non-BIT commit, which is where
// it accumulates the report's counters (mst_files/mst_commit.hpp), applied to
// each tree edge's route as the writer emits it (writer_util.hpp:115-165):
//   * a wire on layer z from u to v counts the cells [min, max) of its run
//     (difference array +1/-1 and a run scan, :68-113, 252-328), once per
//     batch however many wires of the batch cover a cell (count > 0,
//     update_cap_cube :329-341);
//   * a via between layers l and l+1 at (x, y) charges layer l's run
//     neighbours of the cell: 2 to the cell at the run's first position, 2 to
//     the previous cell at its last, otherwise 1 to each (gpu_commit_via
//     :17-61).
// Report (main.cpp:109-131), per layer z: demand = 2 wire + via,
// of = demand - 2 cap, max overflow = max(of / 2) (starting at 0), cost =
// sum of e^(of / 4), the exponent rounded to fp32 as expf's argument and e^x
// evaluated by cw_exp.  Totals add layers 1..Z-1 as the report does.
#include <algorithm>
#include <cstring>

#include "gr_internal.h"

using namespace grd;
using namespace grh;

namespace {

constexpr int kRedBlocks = 64;  // partial sums per layer

struct EvalPart {
  long long wire, via, demand;
  int max_wire, max_via;
  float max_cap;
  double max_of, of_cost, overflow;
};

// Via between layers l and l+1 at (x, y).
__device__ __forceinline__ void via_charge(const Geo& g, int32_t* __restrict__ vcnt, int l, int x, int y) {
  const int c = (l & 1) ? y : x, L = (l & 1) ? g.Y : g.X;
  const uint32_t cell = enc(g, l, x, y);
  if (c == 0) atomicAdd(&vcnt[cell], 2);
  else if (c == L - 1) atomicAdd(&vcnt[cell - 1], 2);
  else { atomicAdd(&vcnt[cell - 1], 1); atomicAdd(&vcnt[cell], 1); }
}

// One launch per batch; one lane per output slot of the batch.  Each lane
// walks its edge's route: via charges go out at once (no-return atomics), the
// up to two wires are kept as cell ranges.  The wavefront then spreads all
// its wire cells over its 64 lanes (prefix of lengths, owner found by a
// shuffle binary search), so every step issues 64 independent exchanges.  A
// wire cell is counted by the lane whose exchange of the batch stamp finds
// another value; launches of one stream run in order, so each (cell, batch)
// counts once and no two lanes of a launch update the same wire count.
__global__ __launch_bounds__(256) void k_eval_paths(Geo g, const int64_t* __restrict__ edges,
                                                    const int32_t* __restrict__ lpat, int64_t s0, int64_t ns,
                                                    const int64_t* __restrict__ pinofs,
                                                    const int32_t* __restrict__ deg, int32_t npos,
                                                    const int32_t* __restrict__ pins, int32_t* __restrict__ wcnt,
                                                    int32_t* __restrict__ vcnt, int32_t* __restrict__ estamp,
                                                    int32_t stamp) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int lane = threadIdx.x & 63;
  uint32_t st0 = 0, st1 = 0;  // first cell of each wire
  int len0 = 0, len1 = 0;     // cells of each wire, [min, max) of its run
  bool ok = j < ns;
  int64_t e = ok ? edges[s0 + j] : -1;
  ok = ok && e >= 0;
  const int local = ok ? (int)(e >> 30) : 0;
  ok = ok && local < npos;
  int n1 = (int)(e & 0x7fff), n2 = (int)((e >> 15) & 0x7fff);
  if (n1 < n2) { const int t = n1; n1 = n2; n2 = t; }
  ok = ok && n1 < deg[local];
  const int p = ok ? lpat[s0 + j] : -1;
  ok = ok && p >= 0 && p < g.Z * g.Z2;
  if (ok) {
    const int64_t po = pinofs[local];
    int z1, x1, y1, z2, x2, y2;
    dec(g, (uint32_t)pins[po + n1], z1, x1, y1);
    dec(g, (uint32_t)pins[po + n2], z2, x2, y2);
    const int a = p / g.Z2;
    int b = p - a * g.Z2;
    const bool both = x1 != x2 && y1 != y2;
    auto vias = [&](int za, int zb) {
      const int lo = za < zb ? za : zb, hi = za < zb ? zb : za;
      for (int l = lo; l < hi; ++l) via_charge(g, vcnt, l, x1, y1);
    };
    auto wire = [&](int z, uint32_t& st, int& len) {
      const bool odd = z & 1;
      const int u = odd ? y1 : x1, v = odd ? y2 : x2;
      const int line = odd ? x1 : y1, L = odd ? g.Y : g.X;
      st = (uint32_t)z * g.XY + (uint32_t)line * L + (uint32_t)min(u, v);
      len = max(u, v) - min(u, v);
      if (odd) y1 = y2; else x1 = x2;
    };
    if (z1 != a) { vias(z1, a); z1 = a; }
    if (z1 > 0) wire(z1, st0, len0);
    b = (z1 & 1) ? 2 * b : 2 * b + 1;
    if (both && b != 0 && z1 != b && b < g.Z) {
      vias(z1, b);
      z1 = b;
      wire(z1, st1, len1);
    }
    if (z1 != z2) vias(z1, z2);
  }
  // inclusive prefix of wire cells over the wavefront
  int incl = len0 + len1;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int v = __shfl_up(incl, d, 64);
    if (lane >= d) incl += v;
  }
  const int total = __shfl(incl, 63, 64);
  for (int k0 = 0; k0 < total; k0 += 64) {
    const int k = k0 + lane;
    // owner: the first lane whose inclusive prefix exceeds k
    int lo = 0, hi = 63;
#pragma unroll
    for (int it = 0; it < 6; ++it) {
      const int mid = (lo + hi) >> 1;
      if (__shfl(incl, mid, 64) > k) hi = mid; else lo = mid + 1;
    }
    const int o = lo;
    const int before = __shfl(incl, o, 64) - __shfl(len0 + len1, o, 64);
    const int l0 = __shfl(len0, o, 64);
    const uint32_t a0 = (uint32_t)__shfl((int)st0, o, 64), a1 = (uint32_t)__shfl((int)st1, o, 64);
    if (k < total) {
      const int off = k - before;
      const uint32_t cell = off < l0 ? a0 + off : a1 + (off - l0);
      if (atomicExch(&estamp[cell], stamp) != stamp) wcnt[cell] += 1;
    }
  }
}

__device__ __forceinline__ void merge(EvalPart& a, const EvalPart& b) {
  a.wire += b.wire; a.via += b.via; a.demand += b.demand;
  a.max_wire = max(a.max_wire, b.max_wire); a.max_via = max(a.max_via, b.max_via);
  a.max_cap = fmaxf(a.max_cap, b.max_cap);
  a.max_of = fmax(a.max_of, b.max_of);
  a.of_cost += b.of_cost; a.overflow += b.overflow;
}

// Fixed-shape tree over the block's 256 partials (deterministic order).
__device__ void block_reduce(EvalPart& v, EvalPart* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) merge(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  v = sh[0];
}

// grid (kRedBlocks, Z): block (k, z) reduces a strided share of layer z.
__global__ __launch_bounds__(256) void k_eval_reduce(Geo g, const int32_t* __restrict__ wcnt,
                                                     const int32_t* __restrict__ vcnt,
                                                     const int32_t* __restrict__ cap0, const float* __restrict__ capf0,
                                                     const int32_t* __restrict__ rem, EvalPart* __restrict__ part) {
  __shared__ EvalPart sh[256];
  const int z = blockIdx.y;
  const uint32_t base = (uint32_t)z * g.XY;
  EvalPart acc{0, 0, 0, 0, 0, 0.0f, 0.0, 0.0, 0.0};
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < g.XY; i += gridDim.x * 256) {
    const int w = wcnt[base + i], v = vcnt[base + i];
    const float cap = capf0 ? capf0[base + i] : (float)cap0[base + i];
    const int dem = cap0[base + i] - rem[base + i];
    acc.wire += w; acc.via += v; acc.demand += dem;
    acc.max_wire = max(acc.max_wire, w); acc.max_via = max(acc.max_via, v);
    acc.max_cap = fmaxf(acc.max_cap, cap);
    const double of = (2.0 * (double)w + (double)v) - 2.0 * (double)cap;
    acc.max_of = fmax(acc.max_of, of / 2.0);
    acc.of_cost += (double)(float)cw_exp((double)(float)(of * 0.25));
    const double over = (double)dem - (double)cap;
    acc.overflow += over > 0.0 ? over : 0.0;
  }
  block_reduce(acc, sh);
  if (threadIdx.x == 0) part[z * gridDim.x + blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void k_eval_final(const EvalPart* __restrict__ part, int nb,
                                                    EvalPart* __restrict__ out) {
  __shared__ EvalPart sh[256];
  EvalPart acc{0, 0, 0, 0, 0, 0.0f, 0.0, 0.0, 0.0};
  for (int k = threadIdx.x; k < nb; k += 256) merge(acc, part[blockIdx.x * nb + k]);
  block_reduce(acc, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = acc;
}

}  // namespace

extern "C" {

int gr_evaluate(gr_ctx* c, gr_eval_layer* layers, gr_eval_total* total) {
  if (!c) return fail(c, GR_EINVAL, "null context");
  if (c->prefix) return fail(c, GR_ESTATE, "the evaluator reports BIT-mode passes (integer demand)");
  Sched& s = c->sched;
  if (s.meta.empty()) return fail(c, GR_ESTATE, "gr_set_schedule and a routed pass must come first");
  (void)hipSetDevice(c->dev);
  int rc;
  const int Z = c->g.Z;
  if (!c->d_wcnt) {
    if ((rc = dalloc(c, &c->d_wcnt, c->cs)) || (rc = dalloc(c, &c->d_vcnt, c->cs)) ||
        (rc = dalloc(c, &c->d_estamp, c->cs)))
      return rc;
    EvalPart* p = nullptr;
    EvalPart* o = nullptr;
    if ((rc = dalloc(c, &p, (size_t)Z * kRedBlocks)) || (rc = dalloc(c, &o, (size_t)Z))) return rc;
    c->d_eval_part = p;
    c->d_eval_out = o;
  }
  HIPC(c, hipMemsetAsync(c->d_wcnt, 0, sizeof(int32_t) * c->cs, c->stream));
  HIPC(c, hipMemsetAsync(c->d_vcnt, 0, sizeof(int32_t) * c->cs, c->stream));
  HIPC(c, hipMemsetAsync(c->d_estamp, 0, sizeof(int32_t) * c->cs, c->stream));
  for (size_t b = 0; b < s.meta.size(); ++b) {
    const BatchMeta& m = s.meta[b];
    if (!m.nslots) continue;
    hipLaunchKernelGGL(k_eval_paths, dim3((unsigned)((m.nslots + 255) / 256)), dim3(256), 0, c->stream, c->g,
                       s.d_edges, s.d_lpat, m.slot0, m.nslots, s.d_pinofs + m.begin, s.d_deg + m.begin, m.n,
                       c->d_pins, c->d_wcnt,
                       c->d_vcnt, c->d_estamp, (int32_t)b + 1);
  }
  hipLaunchKernelGGL(k_eval_reduce, dim3(kRedBlocks, Z), dim3(256), 0, c->stream, c->g, c->d_wcnt, c->d_vcnt,
                     c->d_cap0, c->d_capf0, c->d_rem, (EvalPart*)c->d_eval_part);
  hipLaunchKernelGGL(k_eval_final, dim3(Z), dim3(256), 0, c->stream, (const EvalPart*)c->d_eval_part, kRedBlocks,
                     (EvalPart*)c->d_eval_out);
  HIPC(c, hipGetLastError());
  std::vector<EvalPart> out(Z);
  HIPC(c, hipMemcpyAsync(out.data(), c->d_eval_out, sizeof(EvalPart) * Z, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  gr_eval_total t{};
  for (int z = 0; z < Z; ++z) {
    const EvalPart& e = out[z];
    if (layers) {
      gr_eval_layer& L = layers[z];
      L.wire = e.wire; L.via = e.via; L.max_wire = e.max_wire; L.max_via = e.max_via;
      L.max_cap = e.max_cap; L.max_overflow = e.max_of; L.of_cost = e.of_cost;
      L.demand = e.demand; L.overflow = e.overflow;
    }
    if (z >= 1) { t.wire += e.wire; t.via += e.via; t.of_cost += e.of_cost; }
    t.demand += e.demand;
    t.overflow += e.overflow;
  }
  if (total) *total = t;
  return 0;
}

int gr_eval_counts(gr_ctx* c, int32_t* wire, int32_t* via) {
  if (!c) return fail(c, GR_EINVAL, "null context");
  if (!c->d_wcnt) return fail(c, GR_ESTATE, "gr_evaluate must come first");
  HIPC(c, hipStreamSynchronize(c->stream));
  if (wire) HIPC(c, hipMemcpy(wire, c->d_wcnt, sizeof(int32_t) * c->cs, hipMemcpyDeviceToHost));
  if (via) HIPC(c, hipMemcpy(via, c->d_vcnt, sizeof(int32_t) * c->cs, hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
