// MI355X (gfx950) routing hot path: L-pattern costing, root-pair Boruvka
// trees, Fenwick cost commit; and the C ABI declared in include/iitmgr.h.
//
// Device state (WIRE_ENC cell order, util.hpp:38-44; every layer is a set of
// contiguous runs: rows of length X on even layers, columns of length Y on odd):
//   rem[cell]   int32  remaining capacity cap - demand (the integer state)
//   pfx[cell]   fp32   Q(pos) of the run's Fenwick tree, materialised per cell,
//                      so a reference range query Q(hi) - Q(lo-1)
//                      (bit_util.hpp:36-38) is two loads
//   stamp[cell] int32  last batch that bumped this cell's demand
//   run_stamp   int32  last batch that touched a run (dirty-run list builder)
// Per batch: k_lpat (one wavefront per pin pair, one lane per pattern) ->
// k_tree (root-pair Boruvka and commit, one wavefront per small net, one
// workgroup per large net) -> k_refresh (one workgroup per dirty run rebuilds
// the canonical Fenwick nodes in LDS and rewrites the run's pfx).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/iitmgr.h"
#include "gr_device.h"

using namespace grd;

namespace {

// ------------------------------------------------------------------------
// Kernels
// ------------------------------------------------------------------------

// Rebuild pfx for the runs in list[0 .. *count).  Packed run id: z << 16 | line.
// Fenwick node i (1-based, lowbit 2^k) covers cells [i - 2^k, i); its value is
// the pairwise sum of that aligned block, built level by level in LDS.  Q(x)
// then accumulates nodes in the reference's descent order (bit_util.hpp:19-35).
__global__ __launch_bounds__(256) void k_refresh(Geo g, const int32_t* __restrict__ rem, float* __restrict__ pfx,
                                                 const float* __restrict__ lut, const int32_t* __restrict__ list,
                                                 const int32_t* __restrict__ count, int32_t* __restrict__ reset_cnt) {
  extern __shared__ float sh[];
  float* slut = sh;
  float* lv = sh + kLut;
  __shared__ int offk[16];
  for (int i = threadIdx.x; i < kLut; i += blockDim.x) slut[i] = lut[i];
  if (reset_cnt && blockIdx.x == 0 && threadIdx.x == 0) *reset_cnt = 0;
  const int n = *count;
  __syncthreads();
  for (int w = blockIdx.x; w < n; w += gridDim.x) {
    const int packed = list[w];
    const int z = packed >> 16, line = packed & 0xffff;
    const int L = (z & 1) ? g.Y : g.X;
    const uint32_t base = (uint32_t)z * g.XY + (uint32_t)line * L;
    for (int i = threadIdx.x; i < L; i += blockDim.x) {
      int r = rem[base + i];
      r = r < kRMin ? kRMin : (r > kRMin + kLut - 1 ? kRMin + kLut - 1 : r);
      lv[i] = slut[r - kRMin];
    }
    if (threadIdx.x == 0) {
      int off = 0;
      for (int k = 0; (L >> k) > 0; ++k) { offk[k] = off; off += L >> k; }
    }
    __syncthreads();
    for (int k = 1; (L >> k) > 0; ++k) {
      const int nk = L >> k, o = offk[k], op = offk[k - 1];
      for (int j = threadIdx.x; j < nk; j += blockDim.x) lv[o + j] = lv[op + 2 * j] + lv[op + 2 * j + 1];
      __syncthreads();
    }
    for (int x = threadIdx.x; x < L; x += blockDim.x) {
      uint32_t i = (uint32_t)x + 1;
      float res = 0.0f;
      while (i) {
        const int k = __builtin_ctz(i);
        res += lv[offk[k] + (int)(i >> k) - 1];
        i &= i - 1;
      }
      pfx[base + x] = res;
    }
    __syncthreads();
  }
}

// L-pattern costing: wavefront w handles pair w of the batch; lane l
// evaluates patterns p = l, l + 64, ...; a wave-wide lexicographic min of
// (cost, p) selects the pattern (bit_util.hpp:118-191; ties -> lowest p).
__global__ __launch_bounds__(256) void k_lpat(Geo g, const float* __restrict__ pfx,
                                              const int64_t* __restrict__ pin_pfx, const int32_t* __restrict__ pins,
                                              const int32_t* __restrict__ bnet, const int64_t* __restrict__ bpair,
                                              int n, int64_t P, float* __restrict__ out_cost,
                                              int16_t* __restrict__ out_pat) {
  const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= P) return;
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (bpair[mid] <= w) lo = mid; else hi = mid;
  }
  const uint32_t o = (uint32_t)(w - bpair[lo]);
  int j, k;
  pair_decode(o, j, k);
  const int64_t pb = pin_pfx[bnet[lo]];
  int z1, x1, y1, z2, x2, y2;
  dec(g, (uint32_t)pins[pb + j], z1, x1, y1);
  dec(g, (uint32_t)pins[pb + k], z2, x2, y2);
  uint64_t best = ~0ull;
  const int np = g.Z * g.Z2;
  for (int p = lane; p < np; p += 64) {
    float c;
    if (pattern_cost(pfx, g, z1, x1, y1, z2, x2, y2, p, c)) {
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

struct TreeArgs {
  Geo g;
  const int32_t* net;       // batch position -> net id
  const int64_t* slot;      // batch position -> first output slot
  const int64_t* pair;      // batch position -> first pair of the net in the batch
  const int32_t* list;      // batch positions handled by this launch
  int n_list;
  int64_t cost_base;        // out_cost index of batch position 0
  const int64_t* pin_pfx;
  const int32_t* pins;
  const float* pair_cost;
  const int16_t* pair_pat;
  int64_t* out_edges;
  int32_t* out_lpat;
  float* out_cost;
  int32_t* rem;
  int32_t* stamp;
  int32_t* run_stamp;
  int32_t* dirty_list;
  int32_t* dirty_cnt;
  int32_t bstamp;
  int32_t quirk;
  int32_t M;                // max(X, Y): run_stamp row stride
};

// Demand bump of one cell, at most once per batch (bit_util.hpp:75-90 applies
// cap -= 1 to every flagged cell).  The first thread to stamp the cell owns
// the decrement; the first to stamp its run appends the run to the dirty list.
__device__ __forceinline__ void mark_cell(const TreeArgs& A, uint32_t cell) {
  if (atomicExch(&A.stamp[cell], A.bstamp) == A.bstamp) return;
  A.rem[cell] -= 1;
  int z, x, y;
  dec(A.g, cell, z, x, y);
  const int line = (z & 1) ? x : y;
  if (atomicExch(&A.run_stamp[z * A.M + line], A.bstamp) != A.bstamp) {
    const int idx = atomicAdd(A.dirty_cnt, 1);
    A.dirty_list[idx] = (z << 16) | line;
  }
}

// Cells of one committed edge (mst_bit_commit.hpp:36-67, 147-235): interior
// via cells, inclusive wires; with the quirk, the k threads that marked the
// first via stack also mark k cells before the first wire's low end (:216).
__device__ void commit_edge(const TreeArgs& A, uint32_t src, uint32_t dst, int p) {
  const Geo& g = A.g;
  if (p < g.Z2 || p >= g.Z * g.Z2) return;  // no valid pattern (cannot happen for Z >= 3)
  int z1, x1, y1, z2, x2, y2;
  dec(g, src, z1, x1, y1);
  dec(g, dst, z2, x2, y2);
  const int a = p / g.Z2, b = p - a * g.Z2;
  const bool D = (x1 != x2) && (y1 != y2);
  int z = z1, x = x1, y = y1, k = 0;
  if (z != a) {
    const int lo = min(z, a), hi = max(z, a);
    for (int l = lo + 1; l < hi; ++l) mark_cell(A, enc(g, l, x, y));
    k = hi - lo - 1;
    z = a;
  }
  // first wire
  if (z & 1) {
    const int lo = min(y, y2), hi = max(y, y2);
    for (int t = lo; t <= hi; ++t) mark_cell(A, enc(g, z, x, t));
    if (A.quirk) for (int t = 1; t <= k; ++t) mark_cell(A, enc(g, z, x, lo) - t);
    y = y2;
  } else {
    const int lo = min(x, x2), hi = max(x, x2);
    for (int t = lo; t <= hi; ++t) mark_cell(A, enc(g, z, t, y));
    if (A.quirk) for (int t = 1; t <= k; ++t) mark_cell(A, enc(g, z, lo, y) - t);
    x = x2;
  }
  const int bp = (z & 1) ? 2 * b : 2 * b + 1;
  if (D && bp != 0) {
    const int lo = min(z, bp), hi = max(z, bp);
    for (int l = lo + 1; l < hi; ++l) mark_cell(A, enc(g, l, x, y));
    z = bp;
    if (z & 1) {
      const int l2 = min(y, y2), h2 = max(y, y2);
      for (int t = l2; t <= h2; ++t) mark_cell(A, enc(g, z, x, t));
      y = y2;
    } else {
      const int l2 = min(x, x2), h2 = max(x, x2);
      for (int t = l2; t <= h2; ++t) mark_cell(A, enc(g, z, t, y));
      x = x2;
    }
  }
  if (z != z2) {
    const int lo = min(z, z2), hi = max(z, z2);
    for (int l = lo + 1; l < hi; ++l) mark_cell(A, enc(g, l, x, y));
  }
}

// Root-pair Boruvka (gmst_util.hpp:164-341) and commit for one net, run by a
// group of gsize threads (one wavefront or one workgroup).
//   min phase    : per component root, min (cost, o) over outgoing pairs
//   partner phase: root of the other endpoint of that pair, at round start
//   union phase  : one thread, candidates in ascending root order, smaller
//                  root hooked under larger; edge recorded as (root, partner)
// Rounds repeat until the net has deg-1 edges (the reference's feedback loop,
// mst_util.hpp:298-314, ends in the same state).
template <bool WAVE>
__device__ void tree_commit(const TreeArgs& A, int pos, int lane, int gsize, int* dsu, unsigned long long* best,
                            int* partner, int* tmp, int2* edges, int* sh_ne, float* sh_cost) {
  auto sync = [&]() {
    if (WAVE) wave_sync(); else __syncthreads();
  };
  const int net = A.net[pos];
  const int64_t pb = A.pin_pfx[net];
  const int n = (int)(A.pin_pfx[net + 1] - pb);
  const float* cost = A.pair_cost + A.pair[pos];
  const int16_t* pat = A.pair_pat + A.pair[pos];
  const uint32_t P = (uint32_t)((int64_t)n * (n - 1) / 2);
  for (int i = lane; i < n; i += gsize) dsu[i] = i;
  if (lane == 0) { *sh_ne = 0; *sh_cost = 0.0f; }
  sync();
  while (*sh_ne < n - 1) {
    for (int i = lane; i < n; i += gsize) best[i] = ~0ull;
    sync();
    for (uint32_t o = lane; o < P; o += gsize) {
      int j, k;
      pair_decode(o, j, k);
      int rj = j, rk = k;
      while (dsu[rj] != rj) rj = dsu[rj];
      while (dsu[rk] != rk) rk = dsu[rk];
      if (rj == rk) continue;
      const unsigned long long key = ((unsigned long long)ord_key(cost[o]) << 32) | o;
      atomicMin(&best[rj], key);
      atomicMin(&best[rk], key);
    }
    sync();
    for (int r = lane; r < n; r += gsize) {
      int pr = -1;
      if (dsu[r] == r && best[r] != ~0ull) {
        int j, k;
        pair_decode((uint32_t)best[r], j, k);
        int rj = j, rk = k;
        while (dsu[rj] != rj) rj = dsu[rj];
        while (dsu[rk] != rk) rk = dsu[rk];
        pr = (rj == r) ? rk : rj;
      }
      partner[r] = pr;
    }
    sync();
    if (lane == 0) {
      int ne = *sh_ne;
      float nc = *sh_cost;
      for (int r = 0; r < n; ++r) {
        const int q = partner[r];
        if (q < 0) continue;
        int n1 = r, n2 = q;
        while (dsu[n1] != n1) n1 = dsu[n1];
        while (dsu[n2] != n2) n2 = dsu[n2];
        for (int v = r; dsu[v] != n1;) { const int nx = dsu[v]; dsu[v] = n1; v = nx; }
        for (int v = q; dsu[v] != n2;) { const int nx = dsu[v]; dsu[v] = n2; v = nx; }
        if (n1 == n2) continue;
        if (n1 < n2) dsu[n1] = n2; else dsu[n2] = n1;
        nc += cost[(uint32_t)best[r]];
        edges[ne++] = make_int2(r, q);
      }
      *sh_ne = ne;
      *sh_cost = nc;
    }
    sync();
    for (int i = lane; i < n; i += gsize) {
      int r = i;
      while (dsu[r] != r) r = dsu[r];
      tmp[i] = r;
    }
    sync();
    for (int i = lane; i < n; i += gsize) dsu[i] = tmp[i];
    sync();
  }
  const int64_t sb = A.slot[pos];
  const int32_t* npins = A.pins + pb;
  for (int e = lane; e < n - 1; e += gsize) {
    const int on1 = edges[e].x, on2 = edges[e].y;
    A.out_edges[sb + e] = ((int64_t)pos << 30) | ((int64_t)on1 << 15) | (int64_t)on2;
    const int hi = max(on1, on2), lo = min(on1, on2);
    const int p = pat[(uint32_t)((int64_t)hi * (hi - 1) / 2 + lo)];
    A.out_lpat[sb + e] = p;
    commit_edge(A, (uint32_t)npins[hi], (uint32_t)npins[lo], p);
  }
  if (lane == 0) {
    A.out_edges[sb + n - 1] = -1;
    A.out_lpat[sb + n - 1] = -1;
    A.out_cost[A.cost_base + pos] = *sh_cost;
  }
}

// One wavefront per net (nets of at most kSmallDeg pins), four per workgroup.
__global__ __launch_bounds__(256) void k_tree_small(TreeArgs A) {
  __shared__ int s_dsu[4][kSmallDeg];
  __shared__ unsigned long long s_best[4][kSmallDeg];
  __shared__ int s_partner[4][kSmallDeg];
  __shared__ int s_tmp[4][kSmallDeg];
  __shared__ int2 s_edges[4][kSmallDeg];
  __shared__ int s_ne[4];
  __shared__ float s_cost[4];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int idx = blockIdx.x * 4 + wv;
  if (idx >= A.n_list) return;
  tree_commit<true>(A, A.list[idx], lane, 64, s_dsu[wv], s_best[wv], s_partner[wv], s_tmp[wv], s_edges[wv],
                    &s_ne[wv], &s_cost[wv]);
}

// One workgroup per net (nets of more than kSmallDeg pins); LDS sized by the
// largest such net of the batch.
__global__ __launch_bounds__(256) void k_tree_large(TreeArgs A, int cap) {
  extern __shared__ unsigned long long shl[];
  unsigned long long* best = shl;
  int2* edges = (int2*)(best + cap);
  int* dsu = (int*)(edges + cap);
  int* partner = dsu + cap;
  int* tmp = partner + cap;
  __shared__ int s_ne;
  __shared__ float s_cost;
  tree_commit<false>(A, A.list[blockIdx.x], threadIdx.x, blockDim.x, dsu, best, partner, tmp, edges, &s_ne,
                     &s_cost);
}

// ------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------

void build_lut(int model, float* lut) {
  for (int i = 0; i < kLut; ++i) {
    const int r = i + kRMin;
    float v;
    if (model == GR_COST_REF_INC) {
      v = (float)std::exp(-0.5 * (double)(r - 1)) - (float)std::exp(-0.5 * (double)r);
    } else if (model == GR_COST_LIT_EXP) {
      v = (float)std::exp(1.0 - (double)r);
    } else {
      v = (float)std::exp(-0.5 * (double)r);
    }
    if (!(v < 1e30f)) v = 1e30f;  // saturation keeps every run sum finite
    lut[i] = v;
  }
}

struct BatchMeta {
  int32_t begin, n;          // schedule positions [begin, begin+n)
  int64_t P;                 // pin pairs in the batch
  int32_t small_ofs, n_small, large_ofs, n_large, max_large;
};

struct Sched {
  std::vector<BatchMeta> meta;
  int32_t n_pos = 0;
  int64_t slots = 0, max_P = 0;
  int32_t* d_net = nullptr;
  int64_t* d_slot = nullptr;
  int64_t* d_pair = nullptr;
  int32_t* d_list = nullptr;
  int64_t* d_edges = nullptr;
  int32_t* d_lpat = nullptr;
  float* d_cost = nullptr;
  size_t cap_pos = 0, cap_slots = 0;
};

struct EvPair { hipEvent_t a, b; int kind; };

std::string g_err;

}  // namespace

struct gr_ctx {
  int dev = 0;
  hipStream_t stream = nullptr;
  Geo g{};
  int model = 0, quirk = 1, M = 0, nruns = 0;
  int64_t cs = 0;
  float lut[kLut];
  int32_t *d_rem = nullptr, *d_cap0 = nullptr, *d_stamp = nullptr, *d_run_stamp = nullptr;
  float *d_pfx = nullptr, *d_lut = nullptr;
  int32_t *d_all_runs = nullptr, *d_dirty[2] = {nullptr, nullptr}, *d_cnt = nullptr;  // d_cnt: [2] + all-count
  int32_t bstamp = 0;
  int32_t n_nets = 0;
  std::vector<int64_t> h_pin_pfx;
  int64_t* d_pin_pfx = nullptr;
  int32_t* d_pins = nullptr;
  Sched sched, single;
  float* d_pair_cost = nullptr;
  int16_t* d_pair_pat = nullptr;
  int64_t pair_cap = 0, last_P = 0;
  bool timing = false;
  std::vector<EvPair> evs;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::string err;
};

namespace {

int fail(gr_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg; else g_err = msg;
  return code;
}

#define HIPC(c, x)                                                                      \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess)                                                               \
      return fail((c), GR_ENODEV, std::string(#x " failed: ") + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int dalloc(gr_ctx* c, T** p, size_t n) {
  if (*p) { (void)hipFree(*p); *p = nullptr; }
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) { *p = nullptr; return fail(c, GR_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e)); }
  return 0;
}

hipEvent_t take_event(gr_ctx* c) {
  if (c->ev_used == c->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->ev_pool.push_back(e);
  }
  return c->ev_pool[c->ev_used++];
}

struct KTimer {
  gr_ctx* c; int kind; hipEvent_t b = nullptr;
  KTimer(gr_ctx* c_, int k) : c(c_), kind(k) {
    if (c->timing) { b = take_event(c); if (b) (void)hipEventRecord(b, c->stream); }
  }
  ~KTimer() {
    if (c->timing && b) {
      hipEvent_t e = take_event(c);
      if (e) { (void)hipEventRecord(e, c->stream); c->evs.push_back({b, e, kind}); }
    }
  }
};

int refresh_all(gr_ctx* c) {
  const size_t lds = (kLut + 2 * (size_t)std::max(c->g.X, c->g.Y)) * sizeof(float);
  {
    KTimer t(c, GR_K_REFRESH);
    hipLaunchKernelGGL(k_refresh, dim3(std::min(c->nruns, 4096)), dim3(256), lds, c->stream, c->g, c->d_rem,
                       c->d_pfx, c->d_lut, c->d_all_runs, c->d_cnt + 2, (int32_t*)nullptr);
  }
  HIPC(c, hipGetLastError());
  return 0;
}

// Build the per-position arrays of a schedule on the host and upload them.
int build_sched(gr_ctx* c, Sched& s, int32_t n_batches, const int32_t* batch_ofst, const int32_t* net_ids) {
  const int32_t npos = batch_ofst[n_batches];
  std::vector<int64_t> slot(npos), pair(npos);
  std::vector<int32_t> list(npos);
  s.meta.assign(n_batches, BatchMeta{});
  int64_t slots = 0, maxP = 0;
  int32_t lo = 0;
  for (int b = 0; b < n_batches; ++b) {
    BatchMeta& m = s.meta[b];
    m.begin = batch_ofst[b];
    m.n = batch_ofst[b + 1] - batch_ofst[b];
    if (m.n < 0) return fail(c, GR_EINVAL, "batch offsets must be non-decreasing");
    int64_t P = 0;
    m.small_ofs = lo;
    for (int i = 0; i < m.n; ++i) {
      const int32_t id = net_ids[m.begin + i];
      if (id < 0 || id >= c->n_nets) return fail(c, GR_EINVAL, "net id out of range");
      const int64_t d = c->h_pin_pfx[id + 1] - c->h_pin_pfx[id];
      if (d > kLargeDeg) return fail(c, GR_ERANGE, "net has more pins than the device tree supports (2048)");
      slot[m.begin + i] = slots;
      pair[m.begin + i] = P;
      slots += d;
      P += d * (d - 1) / 2;
      if (d <= kSmallDeg) list[lo++] = i;
    }
    m.n_small = lo - m.small_ofs;
    m.large_ofs = lo;
    m.max_large = 0;
    for (int i = 0; i < m.n; ++i) {
      const int32_t id = net_ids[m.begin + i];
      const int64_t d = c->h_pin_pfx[id + 1] - c->h_pin_pfx[id];
      if (d > kSmallDeg) { list[lo++] = i; m.max_large = std::max<int>(m.max_large, (int)d); }
    }
    m.n_large = lo - m.large_ofs;
    m.P = P;
    maxP = std::max(maxP, P);
  }
  s.n_pos = npos;
  s.slots = slots;
  s.max_P = maxP;
  if ((size_t)npos > s.cap_pos) {
    int rc;
    if ((rc = dalloc(c, &s.d_net, npos)) || (rc = dalloc(c, &s.d_slot, npos)) || (rc = dalloc(c, &s.d_pair, npos)) ||
        (rc = dalloc(c, &s.d_list, npos)) || (rc = dalloc(c, &s.d_cost, npos)))
      return rc;
    s.cap_pos = npos;
  }
  if ((size_t)slots > s.cap_slots) {
    int rc;
    if ((rc = dalloc(c, &s.d_edges, slots)) || (rc = dalloc(c, &s.d_lpat, slots))) return rc;
    s.cap_slots = slots;
  }
  if (maxP > c->pair_cap) {
    int rc;
    if ((rc = dalloc(c, &c->d_pair_cost, maxP)) || (rc = dalloc(c, &c->d_pair_pat, maxP))) return rc;
    c->pair_cap = maxP;
  }
  if (npos) {
    HIPC(c, hipMemcpyAsync(s.d_net, net_ids, sizeof(int32_t) * npos, hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(s.d_slot, slot.data(), sizeof(int64_t) * npos, hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(s.d_pair, pair.data(), sizeof(int64_t) * npos, hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(s.d_list, list.data(), sizeof(int32_t) * npos, hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));  // host vectors go out of scope
  }
  return 0;
}

int enqueue_batch(gr_ctx* c, Sched& s, int b) {
  const BatchMeta& m = s.meta[b];
  if (m.n == 0) return 0;
  const int32_t st = ++c->bstamp;
  const int par = st & 1;
  if (m.P > 0) {
    KTimer t(c, GR_K_LPAT);
    const int64_t blocks = (m.P + 3) / 4;
    hipLaunchKernelGGL(k_lpat, dim3((uint32_t)blocks), dim3(256), 0, c->stream, c->g, c->d_pfx, c->d_pin_pfx,
                       c->d_pins, s.d_net + m.begin, s.d_pair + m.begin, m.n, m.P, c->d_pair_cost, c->d_pair_pat);
  }
  TreeArgs A;
  A.g = c->g;
  A.net = s.d_net + m.begin;
  A.slot = s.d_slot + m.begin;
  A.pair = s.d_pair + m.begin;
  A.cost_base = m.begin;
  A.pin_pfx = c->d_pin_pfx;
  A.pins = c->d_pins;
  A.pair_cost = c->d_pair_cost;
  A.pair_pat = c->d_pair_pat;
  A.out_edges = s.d_edges;
  A.out_lpat = s.d_lpat;
  A.out_cost = s.d_cost;
  A.rem = c->d_rem;
  A.stamp = c->d_stamp;
  A.run_stamp = c->d_run_stamp;
  A.dirty_list = c->d_dirty[par];
  A.dirty_cnt = c->d_cnt + par;
  A.bstamp = st;
  A.quirk = c->quirk;
  A.M = c->M;
  {
    KTimer t(c, GR_K_TREE);
    if (m.n_small) {
      A.list = s.d_list + m.small_ofs;
      A.n_list = m.n_small;
      hipLaunchKernelGGL(k_tree_small, dim3((m.n_small + 3) / 4), dim3(256), 0, c->stream, A);
    }
    if (m.n_large) {
      A.list = s.d_list + m.large_ofs;
      A.n_list = m.n_large;
      const int cap = m.max_large;
      const size_t lds = (size_t)cap * (8 + 8 + 4 + 4 + 4);
      hipLaunchKernelGGL(k_tree_large, dim3(m.n_large), dim3(256), lds, c->stream, A, cap);
    }
  }
  {
    KTimer t(c, GR_K_REFRESH);
    const size_t lds = (kLut + 2 * (size_t)std::max(c->g.X, c->g.Y)) * sizeof(float);
    hipLaunchKernelGGL(k_refresh, dim3(std::min(c->nruns, 2048)), dim3(256), lds, c->stream, c->g, c->d_rem,
                       c->d_pfx, c->d_lut, c->d_dirty[par], c->d_cnt + par, c->d_cnt + (par ^ 1));
  }
  HIPC(c, hipGetLastError());
  c->last_P = m.P;
  return 0;
}

void free_sched(Sched& s) {
  (void)hipFree(s.d_net); (void)hipFree(s.d_slot); (void)hipFree(s.d_pair); (void)hipFree(s.d_list);
  (void)hipFree(s.d_edges); (void)hipFree(s.d_lpat); (void)hipFree(s.d_cost);
  s = Sched();
}

}  // namespace

extern "C" {

const char* gr_version(void) { return "iitmglobalrouter_amd 0.1 (gfx950)"; }

int gr_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  int k = 0;
  for (int i = 0; i < n; ++i) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, i) == hipSuccess && std::strncmp(p.gcnArchName, "gfx950", 6) == 0) ++k;
  }
  return k;
}

int gr_lut_rmin(void) { return kRMin; }
int gr_lut_size(void) { return kLut; }

const char* gr_last_error(const gr_ctx* c) { return c ? c->err.c_str() : g_err.c_str(); }

int gr_create(const gr_grid_desc* d, const int32_t* cap, gr_ctx** out) {
  if (!d || !cap || !out) return fail(nullptr, GR_EINVAL, "null argument");
  *out = nullptr;
  if (d->X < 1 || d->Y < 1 || d->X > 8192 || d->Y > 8192 || d->Z < 3 || d->Z > 32)
    return fail(nullptr, GR_ERANGE, "grid must have 1 <= X, Y <= 8192 and 3 <= Z <= 32");
  const int64_t cs = (int64_t)d->X * d->Y * d->Z;
  if (cs >= (1ll << 31)) return fail(nullptr, GR_ERANGE, "grid has 2^31 or more cells");
  if (d->cost_model < 0 || d->cost_model > 2 || d->commit_mode < 0 || d->commit_mode > 1)
    return fail(nullptr, GR_EINVAL, "unknown cost model or commit mode");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= d->device || d->device < 0)
    return fail(nullptr, GR_ENODEV, "no HIP device");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, d->device) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(nullptr, GR_ENODEV, "device is not gfx950");
  gr_ctx* c = new gr_ctx();
  c->dev = d->device;
  if (hipSetDevice(c->dev) != hipSuccess) { delete c; return fail(nullptr, GR_ENODEV, "hipSetDevice failed"); }
  c->g.X = d->X; c->g.Y = d->Y; c->g.Z = d->Z; c->g.Z2 = (d->Z + 1) / 2;
  c->g.XY = (uint32_t)d->X * (uint32_t)d->Y;
  c->model = d->cost_model;
  c->quirk = d->commit_mode == GR_COMMIT_REF_QUIRK;
  c->cs = cs;
  c->M = std::max(d->X, d->Y);
  build_lut(c->model, c->lut);
  std::vector<int32_t> runs;
  for (int z = 0; z < d->Z; ++z) {
    const int lines = (z & 1) ? d->X : d->Y;
    for (int l = 0; l < lines; ++l) runs.push_back((z << 16) | l);
  }
  c->nruns = (int)runs.size();
  int rc = 0;
  auto bail = [&](int code) { gr_destroy(c); return code; };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail(fail(nullptr, GR_ENODEV, "stream"));
  if ((rc = dalloc(c, &c->d_rem, cs)) || (rc = dalloc(c, &c->d_cap0, cs)) || (rc = dalloc(c, &c->d_stamp, cs)) ||
      (rc = dalloc(c, &c->d_pfx, cs)) || (rc = dalloc(c, &c->d_lut, kLut)) ||
      (rc = dalloc(c, &c->d_run_stamp, (size_t)d->Z * c->M)) || (rc = dalloc(c, &c->d_all_runs, c->nruns)) ||
      (rc = dalloc(c, &c->d_dirty[0], (size_t)d->Z * c->M)) || (rc = dalloc(c, &c->d_dirty[1], (size_t)d->Z * c->M)) ||
      (rc = dalloc(c, &c->d_cnt, 4))) {
    g_err = c->err;
    return bail(rc);
  }
  const size_t lds_max = (kLut + 2 * (size_t)c->M) * sizeof(float);
  int32_t cnt[4] = {0, 0, c->nruns, 0};
  if (hipFuncSetAttribute((const void*)k_refresh, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_max) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_tree_large, hipFuncAttributeMaxDynamicSharedMemorySize,
                          kLargeDeg * 28) != hipSuccess ||
      hipMemcpy(c->d_cap0, cap, sizeof(int32_t) * cs, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_rem, cap, sizeof(int32_t) * cs, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(c->d_stamp, 0, sizeof(int32_t) * cs) != hipSuccess ||
      hipMemset(c->d_run_stamp, 0, sizeof(int32_t) * (size_t)d->Z * c->M) != hipSuccess ||
      hipMemcpy(c->d_lut, c->lut, sizeof(float) * kLut, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_all_runs, runs.data(), sizeof(int32_t) * runs.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_cnt, cnt, sizeof(cnt), hipMemcpyHostToDevice) != hipSuccess) {
    g_err = "device set-up failed";
    return bail(GR_ENODEV);
  }
  if ((rc = refresh_all(c)) || hipStreamSynchronize(c->stream) != hipSuccess) {
    g_err = c->err.empty() ? "initial cost build failed" : c->err;
    return bail(rc ? rc : GR_ENODEV);
  }
  *out = c;
  return 0;
}

void gr_destroy(gr_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->dev);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  (void)hipFree(c->d_rem); (void)hipFree(c->d_cap0); (void)hipFree(c->d_stamp); (void)hipFree(c->d_run_stamp);
  (void)hipFree(c->d_pfx); (void)hipFree(c->d_lut); (void)hipFree(c->d_all_runs);
  (void)hipFree(c->d_dirty[0]); (void)hipFree(c->d_dirty[1]); (void)hipFree(c->d_cnt);
  (void)hipFree(c->d_pin_pfx); (void)hipFree(c->d_pins);
  (void)hipFree(c->d_pair_cost); (void)hipFree(c->d_pair_pat);
  free_sched(c->sched);
  free_sched(c->single);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int gr_set_nets(gr_ctx* c, int32_t n_nets, const int64_t* pin_pfx, const int32_t* pin_enc) {
  if (!c || n_nets < 0 || !pin_pfx || (!pin_enc && n_nets)) return fail(c, GR_EINVAL, "bad net arrays");
  if (pin_pfx[0] != 0) return fail(c, GR_EINVAL, "pin_pfx[0] must be 0");
  for (int32_t i = 0; i < n_nets; ++i) {
    const int64_t d = pin_pfx[i + 1] - pin_pfx[i];
    if (d < 2) return fail(c, GR_EINVAL, "every net needs at least 2 distinct pins");
    if (d > 32767) return fail(c, GR_ERANGE, "more than 32767 pins in a net (15-bit pin index)");
    for (int64_t p = pin_pfx[i]; p < pin_pfx[i + 1]; ++p) {
      if (pin_enc[p] < 0 || pin_enc[p] >= c->cs) return fail(c, GR_EINVAL, "pin outside the grid");
      if (p > pin_pfx[i] && pin_enc[p] <= pin_enc[p - 1])
        return fail(c, GR_EINVAL, "pins of a net must be strictly ascending by enc");
    }
  }
  (void)hipSetDevice(c->dev);
  int rc;
  const int64_t np = pin_pfx[n_nets];
  if ((rc = dalloc(c, &c->d_pin_pfx, n_nets + 1)) || (rc = dalloc(c, &c->d_pins, np))) return rc;
  HIPC(c, hipMemcpy(c->d_pin_pfx, pin_pfx, sizeof(int64_t) * (n_nets + 1), hipMemcpyHostToDevice));
  if (np) HIPC(c, hipMemcpy(c->d_pins, pin_enc, sizeof(int32_t) * np, hipMemcpyHostToDevice));
  c->h_pin_pfx.assign(pin_pfx, pin_pfx + n_nets + 1);
  c->n_nets = n_nets;
  free_sched(c->sched);
  return 0;
}

int gr_route_batch(gr_ctx* c, const int32_t* net_ids, int32_t n, int64_t* out_edges, int32_t* out_lpat,
                   float* out_net_cost) {
  if (!c || n < 0 || (n && !net_ids)) return fail(c, GR_EINVAL, "bad batch");
  if (!c->d_pins) return fail(c, GR_ESTATE, "gr_set_nets must come first");
  (void)hipSetDevice(c->dev);
  int32_t ofs[2] = {0, n};
  int rc;
  if ((rc = build_sched(c, c->single, 1, ofs, net_ids))) return rc;
  if ((rc = enqueue_batch(c, c->single, 0))) return rc;
  const int64_t slots = c->single.slots;
  if (out_edges && slots)
    HIPC(c, hipMemcpyAsync(out_edges, c->single.d_edges, sizeof(int64_t) * slots, hipMemcpyDeviceToHost, c->stream));
  if (out_lpat && slots)
    HIPC(c, hipMemcpyAsync(out_lpat, c->single.d_lpat, sizeof(int32_t) * slots, hipMemcpyDeviceToHost, c->stream));
  if (out_net_cost && n)
    HIPC(c, hipMemcpyAsync(out_net_cost, c->single.d_cost, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return 0;
}

int gr_set_schedule(gr_ctx* c, int32_t n_batches, const int32_t* batch_ofst, const int32_t* net_ids) {
  if (!c || n_batches < 0 || !batch_ofst || (batch_ofst[n_batches] && !net_ids)) return fail(c, GR_EINVAL, "bad schedule");
  if (!c->d_pins) return fail(c, GR_ESTATE, "gr_set_nets must come first");
  if (batch_ofst[0] != 0) return fail(c, GR_EINVAL, "batch_ofst[0] must be 0");
  (void)hipSetDevice(c->dev);
  return build_sched(c, c->sched, n_batches, batch_ofst, net_ids);
}

int gr_route_schedule(gr_ctx* c, int32_t first, int32_t count) {
  if (!c) return fail(c, GR_EINVAL, "null context");
  if (first < 0 || count < 0 || (size_t)first + count > c->sched.meta.size())
    return fail(c, GR_EINVAL, "batch range outside the schedule");
  (void)hipSetDevice(c->dev);
  for (int32_t b = first; b < first + count; ++b) {
    const int rc = enqueue_batch(c, c->sched, b);
    if (rc) return rc;
  }
  return 0;
}

int gr_sync(gr_ctx* c) {
  if (!c) return fail(c, GR_EINVAL, "null context");
  HIPC(c, hipStreamSynchronize(c->stream));
  return 0;
}

int64_t gr_schedule_slots(const gr_ctx* c) { return c ? c->sched.slots : 0; }

int gr_get_schedule_outputs(gr_ctx* c, int64_t* out_edges, int32_t* out_lpat, float* out_net_cost) {
  if (!c) return fail(c, GR_EINVAL, "null context");
  const Sched& s = c->sched;
  HIPC(c, hipStreamSynchronize(c->stream));
  if (out_edges && s.slots) HIPC(c, hipMemcpy(out_edges, s.d_edges, sizeof(int64_t) * s.slots, hipMemcpyDeviceToHost));
  if (out_lpat && s.slots) HIPC(c, hipMemcpy(out_lpat, s.d_lpat, sizeof(int32_t) * s.slots, hipMemcpyDeviceToHost));
  if (out_net_cost && s.n_pos) HIPC(c, hipMemcpy(out_net_cost, s.d_cost, sizeof(float) * s.n_pos, hipMemcpyDeviceToHost));
  return 0;
}

int gr_reset_demand(gr_ctx* c) {
  if (!c) return fail(c, GR_EINVAL, "null context");
  (void)hipSetDevice(c->dev);
  HIPC(c, hipMemcpyAsync(c->d_rem, c->d_cap0, sizeof(int32_t) * c->cs, hipMemcpyDeviceToDevice, c->stream));
  HIPC(c, hipMemsetAsync(c->d_cnt, 0, 2 * sizeof(int32_t), c->stream));
  return refresh_all(c);
}

int gr_get_demand(gr_ctx* c, int32_t* out) {
  if (!c || !out) return fail(c, GR_EINVAL, "null argument");
  (void)hipSetDevice(c->dev);
  std::vector<int32_t> cap0(c->cs);
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipMemcpy(out, c->d_rem, sizeof(int32_t) * c->cs, hipMemcpyDeviceToHost));
  HIPC(c, hipMemcpy(cap0.data(), c->d_cap0, sizeof(int32_t) * c->cs, hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < c->cs; ++i) out[i] = cap0[i] - out[i];
  return 0;
}

int gr_get_prefix(gr_ctx* c, float* out) {
  if (!c || !out) return fail(c, GR_EINVAL, "null argument");
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipMemcpy(out, c->d_pfx, sizeof(float) * c->cs, hipMemcpyDeviceToHost));
  return 0;
}

int64_t gr_last_pairs(const gr_ctx* c) { return c ? c->last_P : 0; }

int gr_get_pair_costs(gr_ctx* c, float* cost, int16_t* pat) {
  if (!c) return fail(c, GR_EINVAL, "null context");
  HIPC(c, hipStreamSynchronize(c->stream));
  if (cost && c->last_P) HIPC(c, hipMemcpy(cost, c->d_pair_cost, sizeof(float) * c->last_P, hipMemcpyDeviceToHost));
  if (pat && c->last_P) HIPC(c, hipMemcpy(pat, c->d_pair_pat, sizeof(int16_t) * c->last_P, hipMemcpyDeviceToHost));
  return 0;
}

int gr_get_lut(const gr_ctx* c, float* out) {
  if (!c || !out) return GR_EINVAL;
  std::memcpy(out, c->lut, sizeof(float) * kLut);
  return 0;
}

int gr_enable_timing(gr_ctx* c, int on) {
  if (!c) return fail(c, GR_EINVAL, "null context");
  HIPC(c, hipStreamSynchronize(c->stream));
  c->timing = on != 0;
  c->evs.clear();
  c->ev_used = 0;
  return 0;
}

int gr_kernel_stats(gr_ctx* c, int64_t* launches, double* total_ms) {
  if (!c || !launches || !total_ms) return fail(c, GR_EINVAL, "null argument");
  HIPC(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < GR_K_COUNT; ++k) { launches[k] = 0; total_ms[k] = 0.0; }
  for (const EvPair& e : c->evs) {
    float ms = 0.f;
    HIPC(c, hipEventElapsedTime(&ms, e.a, e.b));
    launches[e.kind] += 1;
    total_ms[e.kind] += ms;
  }
  return 0;
}

}  // extern "C"
