// MI355X (gfx950) routing hot path: the C ABI declared in include/iitmgr.h
// over the kernels of gr_kernels.h.
//
// A context owns the grid state (integer remaining capacity, materialised
// Fenwick prefixes, stamps) and the nets; a schedule is a sequence of batches
// whose per-position descriptors live on the device, so routing a pass only
// enqueues kernels: no per-batch allocation, copy or host synchronisation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/iitmgr.h"
#include "gr_internal.h"
#include "gr_kernels.h"
#include "gr_prefix.h"

using namespace grd;

namespace {

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

}  // namespace

namespace grh {
std::string g_err;

// Allocation-free routing (gr_route_batch_static): one static device array per
// GPU holds every device buffer a batch needs beyond the context's grid state
// and nets.  Host code never uses the array's name as an address: the device
// address comes from hipGetSymbolAddress, once per device, kept below.
constexpr size_t kScratchBytes = (size_t)GR_STATIC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's array; the device must be current.
int scratch_address(int dev, void** out) {
  *out = nullptr;
  if (dev < 0 || dev >= kMaxDevices) return fail(nullptr, GR_ERANGE, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(nullptr, GR_ENODEV, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

void drop_graph(gr_ctx* c) {
  Sched& s = c->sched;
  if (s.exec) { (void)hipGraphExecDestroy(s.exec); s.exec = nullptr; }
  s.graph_failed = false;
}
}  // namespace grh

using namespace grh;

namespace {

hipEvent_t take_event(gr_ctx* c) {
  if (c->ev_used == c->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->ev_pool.push_back(e);
  }
  return c->ev_pool[c->ev_used++];
}

// Brackets the launches of one kernel kind with events on the launch stream.
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

size_t refresh_lds(const gr_ctx* c, bool with_vterm) {
  return (kLut + (with_vterm ? 2 : 1) * (size_t)refresh_stride(c->M)) * sizeof(float);
}
size_t refresh_prefix_lds(const gr_ctx* c) { return (3 * (size_t)c->M + 128) * sizeof(float); }
// Refresh grid: the tuned size, or fewer workgroups when the batch can touch
// fewer runs (runs < 0: every run, as at set-up and reset).  Workgroups scan
// the runs with a stride, so any size visits every stamped run.
int refresh_grid(const gr_ctx* c, int runs) {
  int g = std::min(c->g.Z * c->M, c->tune.refresh_blocks);
  if (runs >= 0) g = std::min(g, std::max(runs, 64));
  return std::max(1, g);
}

// rl (nrl entries): the batch's candidate runs, one workgroup each up to the
// tuned grid; null: every run, with the grid bounded by runs.
void launch_refresh(gr_ctx* c, int all, int runs = -1, const int32_t* rl = nullptr, int nrl = 0,
                    int32_t* zero8 = nullptr) {
  const int grid = rl ? std::max(1, std::min(nrl, c->tune.refresh_blocks)) : refresh_grid(c, runs);
  if (c->prefix)
    hipLaunchKernelGGL(k_refresh_prefix, dim3(grid), dim3(256), refresh_prefix_lds(c), c->stream,
                       c->g, c->M, c->d_remf, c->d_mark, c->d_vhalf, c->d_pfx, c->d_vterm, c->d_run_stamp, c->d_ctr,
                       all, rl, nrl);
  else if (c->d_remf)
    hipLaunchKernelGGL(k_refresh<true>, dim3(grid), dim3(256), refresh_lds(c, false), c->stream,
                       c->g, c->M, c->d_rem, c->d_mark, c->d_pfx, c->d_lut, c->d_run_stamp, c->d_ctr, all, c->d_remf,
                       c->model, (int16_t*)nullptr, (float*)nullptr, rl, nrl, zero8);
  else
    hipLaunchKernelGGL((c->vterm_live ? k_refresh<false, true> : k_refresh<false, false>), dim3(grid),
                       dim3(256), refresh_lds(c, c->vterm_live), c->stream, c->g, c->M, c->d_rem, c->d_mark, c->d_pfx,
                       c->d_lut, c->d_run_stamp, c->d_ctr, all, c->d_remf, c->model,
                       c->remz_live ? c->d_remz : (int16_t*)nullptr, c->vterm_live ? c->d_vterm : (float*)nullptr,
                       rl, nrl, zero8);
}

int refresh_all(gr_ctx* c) {
  {
    KTimer t(c, GR_K_REFRESH);
    launch_refresh(c, 1);
  }
  HIPC(c, hipGetLastError());
  return 0;
}

// The workgroup trees of a batch keep the pair keys in LDS when the largest
// large net's fit next to its node arrays; otherwise every large net of the
// batch streams them from global memory.
bool tree_streams(const gr_ctx* c, const BatchMeta& m) {
  if (!m.cls_n[kClsLarge]) return false;
  const int64_t budget = (int64_t)c->lds_max - 2048;
  return (int64_t)tree_lds_bytes(m.max_large, 0, (int)std::min<int64_t>(m.max_large_P, 1 << 30)) > budget;
}

// Candidate runs of one batch (appended to h.rlist, m.rlist_ofs / m.nrl set)
// when they are fewer than three quarters of all runs: the rows of each net's
// box on even layers and its columns on odd ones, each range with one line
// more below for the quirk cells just before a first wire's low end (they
// fall in the previous run of the layer), and the last run of every layer
// when a box touches line 0 (the previous run of line 0 is the previous
// layer's last one).  Grids with a side of at most Z + 1 cells, where quirk
// cells could reach further back, refresh every run.
void batch_runs(const gr_ctx* c, HostSched& h, BatchMeta& m, const int32_t* ids) {
  m.rlist_ofs = -1;
  m.nrl = 0;
  const Geo& g = c->g;
  const int total = g.Z * c->M;
  if ((int)c->h_net_box.size() < 4 * c->n_nets || std::min(g.X, g.Y) <= g.Z + 1 || m.runs >= total) return;
  std::vector<uint8_t> rows(g.Y, 0), cols(g.X, 0);
  bool wrap = false;
  for (int i = 0; i < m.n; ++i) {
    const int32_t* b = &c->h_net_box[4 * (size_t)ids[i]];
    for (int y = std::max(b[2] - 1, 0); y <= b[3]; ++y) rows[y] = 1;
    for (int x = std::max(b[0] - 1, 0); x <= b[1]; ++x) cols[x] = 1;
    wrap = wrap || b[0] == 0 || b[2] == 0;
  }
  const size_t at = h.rlist.size();
  for (int z = 0; z < g.Z; ++z) {
    const int lines = (z & 1) ? g.X : g.Y;
    const uint8_t* cov = (z & 1) ? cols.data() : rows.data();
    for (int l = 0; l < lines; ++l)
      if (cov[l] || (wrap && l == lines - 1)) h.rlist.push_back(z * c->M + l);
  }
  const size_t n = h.rlist.size() - at;
  if (4 * n >= 3 * (size_t)total) {
    h.rlist.resize(at);
    return;
  }
  m.rlist_ofs = (int32_t)at;
  m.nrl = (int32_t)n;
}

int host_sched(gr_ctx* c, Sched& s, HostSched& h, int32_t n_batches, const int32_t* batch_ofst,
               const int32_t* net_ids, bool lists = false) {
  const int32_t npos = batch_ofst[n_batches];
  h.rlist.clear();
  h.slot.assign(npos, 0);
  h.sqofs.assign(npos, -1);
  h.rowtask.clear();
  s.sq_words = 0;
  s.rowmin_slots = 0;
  h.pair.assign(npos, 0);
  h.pinofs.assign(npos, 0);
  h.degv.assign(npos, 0);
  h.list.assign(npos, 0);
  h.chunk.clear();
  s.meta.assign(n_batches, BatchMeta{});
  s.giant_bytes = 0;
  int64_t slots = 0, maxP = 0, total_P = 0;
  int32_t lo = 0;
  for (int b = 0; b < n_batches; ++b) {
    BatchMeta& m = s.meta[b];
    m.begin = batch_ofst[b];
    m.n = batch_ofst[b + 1] - batch_ofst[b];
    if (m.n < 0) return fail(c, GR_EINVAL, "batch offsets must be non-decreasing");
    int64_t P = 0, runs = 0;
    m.max_large = 0;
    m.max_large_P = 0;
    m.max_giant = 0;
    m.slot0 = slots;
    for (int i = 0; i < m.n; ++i) {
      const int32_t id = net_ids[m.begin + i];
      if (id < 0 || id >= c->n_nets) return fail(c, GR_EINVAL, "net id out of range");
      const int64_t d = c->h_pin_pfx[id + 1] - c->h_pin_pfx[id];
      if (d > kMaxDeg)
        return fail(c, GR_ERANGE, "net has more pins than the 15-bit pin index allows (" + std::to_string(kMaxDeg) + ")");
      h.slot[m.begin + i] = slots;
      h.pair[m.begin + i] = P;
      h.pinofs[m.begin + i] = c->h_pin_pfx[id];
      h.degv[m.begin + i] = (int32_t)d;
      slots += d;
      P += d * (d - 1) / 2;
      runs += c->h_net_runs[id];
    }
    m.pair_base = total_P;
    total_P += P;
    m.runs = (int32_t)std::min<int64_t>(runs, (int64_t)c->g.Z * c->M);
    if (lists) batch_runs(c, h, m, net_ids + m.begin);
    m.nslots = slots - m.slot0;
    for (int cls = 0; cls < kNumCls; ++cls) {
      m.cls_ofs[cls] = lo;
      for (int i = 0; i < m.n; ++i) {
        const int32_t id = net_ids[m.begin + i];
        const int64_t d = c->h_pin_pfx[id + 1] - c->h_pin_pfx[id];
        const int k = net_class(d, c->tune.group_max);
        if (k != cls) continue;
        h.list[lo++] = i;
        if (k == kClsLarge) {
          m.max_large = std::max<int>(m.max_large, (int)d);
          m.max_large_P = std::max<int64_t>(m.max_large_P, d * (d - 1) / 2);
        }
        if (k == kClsGiant) m.max_giant = std::max<int>(m.max_giant, (int)d);
      }
      m.cls_n[cls] = lo - m.cls_ofs[cls];
      // workgroup trees by descending pin count: the batch's longest trees
      // are dispatched first
      if (cls == kClsLarge || cls == kClsGiant)
        std::stable_sort(h.list.begin() + m.cls_ofs[cls], h.list.begin() + lo,
                         [&](int32_t a, int32_t b2) { return h.degv[m.begin + a] > h.degv[m.begin + b2]; });
    }
    // owner of pair 64*c for every 64-pair chunk of the batch
    m.chunk_ofs = (int32_t)h.chunk.size();
    int owner = 0;
    for (int64_t w = 0; w < P; w += 64) {
      while (owner + 1 < m.n && h.pair[m.begin + owner + 1] <= w) ++owner;
      h.chunk.push_back(owner);
    }
    m.P = P;
    maxP = std::max(maxP, P);
    if (m.max_giant) s.giant_bytes = std::max<int64_t>(s.giant_bytes, m.nslots * 44);
    m.streams = tree_streams(c, m);
    m.rowtask_ofs = (int32_t)h.rowtask.size();
    if (m.streams) {
      int64_t words = 0;
      for (int q = m.cls_ofs[kClsLarge]; q < m.cls_ofs[kClsLarge] + m.cls_n[kClsLarge]; ++q) {
        const int32_t i = h.list[q];
        const int64_t d = h.degv[m.begin + i];
        h.sqofs[m.begin + i] = words;
        words += d * square_stride(d);
        for (int r0 = 0; r0 < d; r0 += kRowTaskPins) h.rowtask.push_back(make_int2(i, r0));
      }
      s.sq_words = std::max(s.sq_words, words);
      s.rowmin_slots = std::max(s.rowmin_slots, m.nslots);
    }
    m.n_rowtask = (int32_t)h.rowtask.size() - m.rowtask_ofs;
  }
  s.n_pos = npos;
  s.slots = slots;
  s.max_P = maxP;
  s.total_P = total_P;
  h.slots = slots;
  h.maxP = maxP;
  return 0;
}

// Copies the host descriptors into the schedule's device arrays.
int upload_sched(gr_ctx* c, Sched& s, const HostSched& h) {
  const size_t npos = h.slot.size();
  if (npos) {
    HIPC(c, hipMemcpyAsync(s.d_pinofs, h.pinofs.data(), sizeof(int64_t) * npos, hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(s.d_deg, h.degv.data(), sizeof(int32_t) * npos, hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(s.d_slot, h.slot.data(), sizeof(int64_t) * npos, hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(s.d_pair, h.pair.data(), sizeof(int64_t) * npos, hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(s.d_list, h.list.data(), sizeof(int32_t) * npos, hipMemcpyHostToDevice, c->stream));
    if (s.d_sqofs)
      HIPC(c, hipMemcpyAsync(s.d_sqofs, h.sqofs.data(), sizeof(int64_t) * npos, hipMemcpyHostToDevice, c->stream));
  }
  if (s.d_rowtask && !h.rowtask.empty())
    HIPC(c, hipMemcpyAsync(s.d_rowtask, h.rowtask.data(), sizeof(int2) * h.rowtask.size(), hipMemcpyHostToDevice,
                           c->stream));
  if (!h.chunk.empty())
    HIPC(c, hipMemcpyAsync(s.d_chunk, h.chunk.data(), sizeof(int32_t) * h.chunk.size(), hipMemcpyHostToDevice,
                           c->stream));
  if (s.d_rlist && !h.rlist.empty())
    HIPC(c, hipMemcpyAsync(s.d_rlist, h.rlist.data(), sizeof(int32_t) * h.rlist.size(), hipMemcpyHostToDevice,
                           c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));  // the host vectors may change after this
  return 0;
}

// Endpoint table of the resident schedule (k_pair_ends per batch), for grids
// that fit its packing; other schedules cost without one.
int build_ends(gr_ctx* c, Sched& s) {
  s.ends_ok = false;
  if (&s != &c->sched || s.ext || !ends_fit(c->g) || s.total_P <= 0) return 0;
  int rc;
  if ((size_t)s.total_P > s.cap_ends) {
    if ((rc = dalloc(c, &s.d_ends, (size_t)s.total_P))) return rc;
    s.cap_ends = (size_t)s.total_P;
  }
  for (const BatchMeta& m : s.meta) {
    if (m.P <= 0) continue;
    hipLaunchKernelGGL(k_pair_ends, dim3((uint32_t)((m.P + 255) / 256)), dim3(256), 0, c->stream, c->g, c->d_pins,
                       s.d_pinofs + m.begin, s.d_pair + m.begin, s.d_chunk + m.chunk_ofs, m.n, m.P,
                       s.d_ends + m.pair_base);
  }
  HIPC(c, hipGetLastError());
  HIPC(c, hipStreamSynchronize(c->stream));
  s.ends_ok = true;
  return 0;
}

// Per-position descriptors of a schedule, built on the host and uploaded once
// into device arrays the schedule owns (grown as needed).
int build_sched(gr_ctx* c, Sched& s, int32_t n_batches, const int32_t* batch_ofst, const int32_t* net_ids) {
  // The captured whole-pass graph holds the resident schedule's batch
  // boundaries and device pointers: rebuilding that schedule invalidates it.
  if (&s == &c->sched) drop_graph(c);
  HostSched h;
  int rc = host_sched(c, s, h, n_batches, batch_ofst, net_ids, true);
  if (rc) return rc;
  s.h_list = h.list;
  s.h_deg = h.degv;
  s.h_pair = h.pair;
  s.fused_built = false;
  if (h.rlist.size() > s.cap_rlist) {
    if ((rc = dalloc(c, &s.d_rlist, h.rlist.size()))) return rc;
    s.cap_rlist = h.rlist.size();
  }
  const size_t npos = h.slot.size();
  const int64_t slots = h.slots, maxP = h.maxP;
  if (npos > s.cap_pos) {
    if ((rc = dalloc(c, &s.d_pinofs, npos)) || (rc = dalloc(c, &s.d_deg, npos)) ||
        (rc = dalloc(c, &s.d_slot, npos)) || (rc = dalloc(c, &s.d_pair, npos)) ||
        (rc = dalloc(c, &s.d_list, npos)) || (rc = dalloc(c, &s.d_cost, npos)) ||
        (rc = dalloc(c, &s.d_sqofs, npos)))
      return rc;
    s.cap_pos = npos;
  }
  if ((size_t)s.sq_words > s.cap_sq) {
    if ((rc = dalloc(c, &s.d_sq, (size_t)s.sq_words))) return rc;
    s.cap_sq = (size_t)s.sq_words;
  }
  if (h.rowtask.size() > s.cap_rowtask) {
    if ((rc = dalloc(c, &s.d_rowtask, h.rowtask.size()))) return rc;
    s.cap_rowtask = h.rowtask.size();
  }
  if ((size_t)s.rowmin_slots > s.cap_rowmin) {
    if ((rc = dalloc(c, &s.d_rowmin, (size_t)s.rowmin_slots))) return rc;
    s.cap_rowmin = (size_t)s.rowmin_slots;
  }
  if ((size_t)slots > s.cap_slots) {
    if ((rc = dalloc(c, &s.d_edges, slots)) || (rc = dalloc(c, &s.d_lpat, slots))) return rc;
    s.cap_slots = slots;
  }
  if (h.chunk.size() > s.cap_chunk) {
    if ((rc = dalloc(c, &s.d_chunk, h.chunk.size()))) return rc;
    s.cap_chunk = h.chunk.size();
  }
  if ((size_t)s.giant_bytes > s.cap_giant) {
    if ((rc = dalloc(c, &s.d_giant, (size_t)s.giant_bytes))) return rc;
    s.cap_giant = (size_t)s.giant_bytes;
  }
  if (maxP > c->pair_cap) {
    // the pair buffers are the context's, shared by every schedule: the
    // captured pass must not keep their old addresses
    drop_graph(c);
    if ((rc = dalloc(c, &c->d_pair_cost, maxP)) || (rc = dalloc(c, &c->d_pair_pat, maxP)) ||
        (rc = dalloc(c, &c->d_pair_fb, maxP)))
      return rc;
    c->pair_cap = maxP;
  }
  if ((rc = upload_sched(c, s, h))) return rc;
  return build_ends(c, s);
}

typedef void (*LpatFn)(Geo, const float*, const int32_t*, const int64_t*, const int64_t*, const int32_t*, int,
                       int64_t, float*, int16_t*, int32_t*, SquareOut);

// Workgroup size of the lane costing kernel for a batch of P pairs: the
// tuned size, or with 0 the largest of 256, 128, 64 that still gives every
// CU several workgroups (at least 2048 in all, 8 per CU).
constexpr int64_t kLpatMinBlocks = 2048;
int lpat_block_for(int tuned, int64_t P) {
  if (tuned) return tuned;
  for (int b : {256, 128})
    if ((P + b - 1) / b >= kLpatMinBlocks) return b;
  return 64;
}

typedef void (*PrefixFn)(Geo, const float*, const float*, const int32_t*, const int64_t*, const int64_t*,
                         const int32_t*, int, int64_t, float*, int16_t*, int32_t*);
PrefixFn prefix_kernel(int Z) {
  switch (Z) {
    case 3: return k_lpat_prefix<3>;
    case 4: return k_lpat_prefix<4>;
    case 5: return k_lpat_prefix<5>;
    case 6: return k_lpat_prefix<6>;
    case 7: return k_lpat_prefix<7>;
    case 8: return k_lpat_prefix<8>;
    case 9: return k_lpat_prefix<9>;
    case 10: return k_lpat_prefix<10>;
    case 11: return k_lpat_prefix<11>;
    case 12: return k_lpat_prefix<12>;
    default: return nullptr;
  }
}

template <bool BUF>
LpatFn lane_kernel_for(int Z) {
  switch (Z) {
    case 3: return k_lpat_lane<3, BUF>;
    case 4: return k_lpat_lane<4, BUF>;
    case 5: return k_lpat_lane<5, BUF>;
    case 6: return k_lpat_lane<6, BUF>;
    case 7: return k_lpat_lane<7, BUF>;
    case 8: return k_lpat_lane<8, BUF>;
    case 9: return k_lpat_lane<9, BUF>;
    case 10: return k_lpat_lane<10, BUF>;
    case 11: return k_lpat_lane<11, BUF>;
    case 12: return k_lpat_lane<12, BUF>;
    default: return nullptr;
  }
}
typedef void (*EndsFn)(Geo, const float*, const uint2*, const int64_t*, const int64_t*, const int32_t*, int, int64_t,
                       float*, int16_t*, int32_t*, SquareOut);
template <bool BUF>
EndsFn ends_kernel_for(int Z) {
  switch (Z) {
    case 3: return k_lpat_ends<3, BUF>;
    case 4: return k_lpat_ends<4, BUF>;
    case 5: return k_lpat_ends<5, BUF>;
    case 6: return k_lpat_ends<6, BUF>;
    case 7: return k_lpat_ends<7, BUF>;
    case 8: return k_lpat_ends<8, BUF>;
    case 9: return k_lpat_ends<9, BUF>;
    case 10: return k_lpat_ends<10, BUF>;
    case 11: return k_lpat_ends<11, BUF>;
    case 12: return k_lpat_ends<12, BUF>;
    default: return nullptr;
  }
}
typedef void (*BatchFn)(TreeArgs, FusedArgs, int, int, int);
BatchFn batch_kernel_for(int Z) {
  switch (Z) {
    case 3: return k_batch<3, true>;
    case 4: return k_batch<4, true>;
    case 5: return k_batch<5, true>;
    case 6: return k_batch<6, true>;
    case 7: return k_batch<7, true>;
    case 8: return k_batch<8, true>;
    case 9: return k_batch<9, true>;
    case 10: return k_batch<10, true>;
    case 11: return k_batch<11, true>;
    case 12: return k_batch<12, true>;
    default: return nullptr;
  }
}
// Buffer-resource loads when the prefix (one padding float and a float per
// cell) spans fewer than 2^31 bytes, flat loads otherwise.
bool prefix_buf(const Geo& g) { return ((uint64_t)g.XY * (uint64_t)g.Z + 1) * 4 <= 0x7fffffffull; }
LpatFn lane_kernel(const Geo& g) {
  return prefix_buf(g) ? lane_kernel_for<true>(g.Z) : lane_kernel_for<false>(g.Z);
}
EndsFn ends_kernel(const Geo& g) {
  return prefix_buf(g) ? ends_kernel_for<true>(g.Z) : ends_kernel_for<false>(g.Z);
}

// Tree stage of one batch.  Small nets (class lists 0..4) and large nets are
// independent.  With both present, tuning fork selects: 0 two launches in
// sequence; 1 the large-net launch on a second stream (a parallel branch of
// the graph, joined before the refresh); 2 one k_tree_mixed launch when two
// of its workgroups fit a CU's LDS (the small-net blocks then run W = 8
// wavefronts each, as two 4-wavefront blocks would), else as 1; 3 also one
// mixed launch at one workgroup per CU, the small-net blocks using as many
// group areas as the large tree's LDS already holds.
int launch_trees(gr_ctx* c, const Sched& s, const BatchMeta& m, TreeArgs A) {
  if (m.cls_n[kClsGiant]) {
    // nets above kLargeDeg pins: node arrays in the schedule's global scratch
    TreeArgs G = A;
    G.list = s.d_list + m.cls_ofs[kClsGiant];
    G.n_list = m.cls_n[kClsGiant];
    const int mcap = (int)(((int64_t)c->lds_max - 4096) / 8);
    hipLaunchKernelGGL(k_tree_giant, dim3(m.cls_n[kClsGiant]), dim3(1024), (size_t)mcap * 8, c->stream, G, mcap,
                       s.d_giant, m.nslots, m.slot0);
  }
  SmallClasses C;
  int blocks = 0;
  for (int k = 0; k < 5; ++k) {
    C.list[k] = s.d_list + m.cls_ofs[k];
    C.n[k] = m.cls_n[k];
    C.block0[k] = blocks;
    const int per_block = k == 0 ? 256 : 256 / kClsMax[k];
    blocks += (C.n[k] + per_block - 1) / per_block;
  }
  C.block0[5] = blocks;
  const int n_large = m.cls_n[kClsLarge];
  const int cap = m.max_large;
  // LDS of the workgroup trees: node arrays, then the pair keys when the
  // largest net's fit; otherwise keys are streamed from global memory and
  // the rest of the LDS holds the component matrix (the matrix rounds start
  // once a round leaves few enough components)
  int scap = 0, mcap = 0;
  bool mat = false;
  if (n_large) {
    const int64_t budget = (int64_t)c->lds_max - 2048;
    if (!m.streams) {
      scap = (int)m.max_large_P;
    } else {
      mat = true;
      int64_t mc = (budget - (int64_t)tree_lds_bytes(cap, 0, 0)) / 8;
      while (mc > 0 && (int64_t)tree_lds_bytes(cap, (int)mc, 0) > budget) mc -= 16;
      mcap = (int)std::max<int64_t>(0, mc);
    }
  }
  const size_t large_lds = n_large ? tree_lds_bytes(cap, mcap, scap) : 0;
  TreeArgs L = A;
  L.list = s.d_list + m.cls_ofs[kClsLarge];
  L.n_list = n_large;
  if (c->tune.fork >= 2 && blocks && n_large) {
    const int Wmax = c->tune.large_threads / 64;
    int W = std::min(8, Wmax);
    size_t lds = std::max(large_lds, (size_t)W * sizeof(GroupLds));
    bool fits = 2 * (lds + 1024) <= (size_t)c->lds_max;
    if (!fits && c->tune.fork >= 3) {
      W = (int)std::min<size_t>((size_t)Wmax, std::max<size_t>(8, large_lds / sizeof(GroupLds)));
      lds = std::max(large_lds, (size_t)W * sizeof(GroupLds));
      fits = lds + 1024 <= (size_t)c->lds_max;
    }
    if (fits) {
      SmallClasses Cm = C;
      int mb = 0;
      for (int k = 0; k < 5; ++k) {
        Cm.block0[k] = mb;
        const int per_block = k == 0 ? 64 * W : 64 * W / kClsMax[k];
        mb += (Cm.n[k] + per_block - 1) / per_block;
      }
      Cm.block0[5] = mb;
      hipLaunchKernelGGL(mat ? k_tree_mixed<true> : k_tree_mixed<false>, dim3(n_large + mb),
                         dim3(c->tune.large_threads), lds, c->stream, L, Cm, n_large, cap, mcap, scap, W);
      return 0;
    }
  }
  const bool fork = c->tune.fork && blocks && n_large && !c->timing;
  hipStream_t lst = c->stream;
  if (fork) {
    HIPC(c, hipEventRecord(c->ev_fork, c->stream));
    HIPC(c, hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
    lst = c->stream2;
  }
  if (blocks) hipLaunchKernelGGL(k_tree_small, dim3(blocks), dim3(256), 0, c->stream, A, C);
  if (n_large)
    hipLaunchKernelGGL(mat ? k_tree_large<true> : k_tree_large<false>, dim3(n_large), dim3(c->tune.large_threads),
                       large_lds, lst, L, cap, mcap, scap);
  if (fork) {
    HIPC(c, hipEventRecord(c->ev_join, c->stream2));
    HIPC(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
  }
  return 0;
}

// Work items of every batch's fused launch (k_batch, tuning fork = 4) for
// the current tuning, uploaded once; a batch the fused launch does not cover
// (square key matrices, nets above kLargeDeg pins, no LDS fit) keeps the
// separate costing and tree launches.  Each batch's work is split into one
// queue per XCD, balanced by estimated workgroup time (a chunk of costing
// about 10 us, a large tree 5 + 0.12 n us for n pins, a small-net block about
// 13 us):
//   small nets: 8 runs of consecutive positions of about equal estimated
//   time; XCD x costs its run's pairs in chunks (one pair per thread of the
//   workgroup) that do not cross a large net's pairs, and runs the tree blocks
//   of the run's nets (the class lists restricted to the run);
//   large nets: by descending pin count, each to the XCD with the least time
//   so far; the XCD costs the net's pairs in chunks and runs its tree.
// A queue is [chunks of its largest net][that net's tree][the small nets'
// chunks][the other large nets' chunks][their trees][the small tree blocks]:
// the largest tree, the batch's critical path, starts once its own chunks are
// done, and every other tree is claimed after all chunks of the queue, so it
// seldom waits holding a CU.
int build_fused(gr_ctx* c, Sched& s) {
  s.fused.assign(s.meta.size(), FusedBatch{});
  s.fused_tune = c->tune;
  s.fused_built = true;
  if (c->tune.fork != 4 || &s != &c->sched || !s.ends_ok || c->prefix || c->tune.lpat_wave || !prefix_buf(c->g) ||
      !batch_kernel_for(c->g.Z))
    return 0;
  const int T = c->tune.large_threads;
  const double kChunkUs = 10.0, kBlockUs = 13.0;
  std::vector<int4> items;
  std::vector<SmallClasses> cxs;
  int max_large = 0;
  for (size_t b = 0; b < s.meta.size(); ++b) {
    const BatchMeta& m = s.meta[b];
    FusedBatch& f = s.fused[b];
    if (m.n == 0 || m.P <= 0 || m.P > INT32_MAX || m.streams || m.cls_n[kClsGiant]) continue;
    const int n_large = m.cls_n[kClsLarge];
    const int scap = n_large ? (int)m.max_large_P : 0;
    const size_t large_lds = n_large ? tree_lds_bytes(m.max_large, 0, scap) : 0;
    const int Wmax = T / 64;
    int W = std::min(8, Wmax);
    size_t lds = std::max(large_lds, (size_t)W * sizeof(GroupLds));
    if (2 * (lds + 1024) > (size_t)c->lds_max) {
      W = (int)std::min<size_t>((size_t)Wmax, std::max<size_t>(8, large_lds / sizeof(GroupLds)));
      lds = std::max(large_lds, (size_t)W * sizeof(GroupLds));
      if (lds + 1024 > (size_t)c->lds_max) continue;
    }
    auto npairs = [&](int i) {
      const int64_t d = s.h_deg[m.begin + i];
      return d * (d - 1) / 2;
    };
    std::vector<uint8_t> large(m.n, 0);
    for (int q = 0; q < n_large; ++q) large[s.h_list[m.cls_ofs[kClsLarge] + q]] = 1;
    // estimated workgroup time of a small net: its share of a costing chunk
    // and of its class's tree block
    auto small_us = [&](int i) {
      const int64_t d = s.h_deg[m.begin + i];
      const int k = net_class(d, c->tune.group_max);
      const int per_block = k == 0 ? 64 * W : 64 * W / kClsMax[k];
      return (double)npairs(i) * kChunkUs / T + kBlockUs / per_block;
    };
    // small nets: 8 position runs [cut[x], cut[x+1]) of about equal time
    double small_total = 0;
    for (int i = 0; i < m.n; ++i)
      if (!large[i]) small_total += small_us(i);
    int cut[9];
    cut[0] = 0;
    {
      double acc = 0;
      int x = 1;
      for (int i = 0; i < m.n && x < 8; ++i) {
        if (!large[i]) acc += small_us(i);
        while (x < 8 && acc * 8 >= small_total * x) cut[x++] = i + 1;
      }
      while (x < 8) cut[x++] = m.n;
    }
    cut[8] = m.n;
    double load[8];
    for (int x = 0; x < 8; ++x) {
      load[x] = 0;
      for (int i = cut[x]; i < cut[x + 1]; ++i)
        if (!large[i]) load[x] += small_us(i);
    }
    auto chunks = [&](std::vector<int4>& dst, int64_t p0, int64_t np, int dep) {
      int k = 0;
      for (int64_t o = 0; o < np; o += T, ++k)
        dst.push_back(make_int4(0, (int)(p0 + o), (int)std::min<int64_t>(T, np - o), dep));
      return k;
    };
    // large nets to XCDs, largest first
    std::vector<int32_t> order(s.h_list.begin() + m.cls_ofs[kClsLarge],
                               s.h_list.begin() + m.cls_ofs[kClsLarge] + n_large);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b2) { return s.h_deg[m.begin + a] > s.h_deg[m.begin + b2]; });
    std::vector<int> lx[8];  // large-net indices (into order) of each XCD, largest first
    for (int k = 0; k < n_large; ++k) {
      const int32_t i = order[k];
      int x = 0;
      for (int y = 1; y < 8; ++y)
        if (load[y] < load[x]) x = y;
      const int64_t np = npairs(i);
      load[x] += (double)((np + T - 1) / T) * kChunkUs + 5.0 + 0.12 * (double)s.h_deg[m.begin + i];
      lx[x].push_back(k);
    }
    std::vector<int4> q[8];
    int n_tree = n_large;
    const size_t cx0 = cxs.size();
    for (int x = 0; x < 8; ++x) {
      std::vector<int> nch(lx[x].size(), 0);
      // the largest net's chunks and tree
      if (!lx[x].empty()) {
        const int k = lx[x][0];
        nch[0] = chunks(q[x], s.h_pair[m.begin + order[k]], npairs(order[k]), k);
        q[x].push_back(make_int4(1, order[k], nch[0], k));
      }
      // the small nets' chunks over the run's pairs between large nets
      int nsc = 0;
      int64_t run0 = -1, run_n = 0;
      for (int i = cut[x]; i <= cut[x + 1]; ++i) {
        if (i == cut[x + 1] || large[i]) {
          if (run_n) nsc += chunks(q[x], run0, run_n, -1);
          run0 = -1;
          run_n = 0;
          continue;
        }
        if (run0 < 0) run0 = s.h_pair[m.begin + i];
        run_n += npairs(i);
      }
      // the other large nets' chunks, then their trees
      for (size_t j = 1; j < lx[x].size(); ++j) {
        const int k = lx[x][j];
        nch[j] = chunks(q[x], s.h_pair[m.begin + order[k]], npairs(order[k]), k);
      }
      for (size_t j = 1; j < lx[x].size(); ++j) {
        const int k = lx[x][j];
        q[x].push_back(make_int4(1, order[k], nch[j], k));
      }
      // the small tree blocks over the class lists restricted to the run
      SmallClasses Cx;
      int mb = 0;
      for (int k = 0; k < 5; ++k) {
        const int32_t* lb = s.h_list.data() + m.cls_ofs[k];
        const int32_t* le = lb + m.cls_n[k];
        const int32_t* a = std::lower_bound(lb, le, cut[x]);
        const int32_t* e = std::lower_bound(lb, le, cut[x + 1]);
        Cx.list[k] = s.d_list + m.cls_ofs[k] + (a - lb);
        Cx.n[k] = (int)(e - a);
        Cx.block0[k] = mb;
        const int per_block = k == 0 ? 64 * W : 64 * W / kClsMax[k];
        mb += (Cx.n[k] + per_block - 1) / per_block;
      }
      Cx.block0[5] = mb;
      cxs.push_back(Cx);
      for (int bq = 0; bq < mb; ++bq) q[x].push_back(make_int4(2, bq, nsc, mb));
      n_tree += mb;
      if (mb == 0 && nsc) return fail(c, GR_ESTATE, "fused batch: small-net chunks without tree blocks");
    }
    // queue x at [at + x * len, at + (x + 1) * len), padded to one length
    size_t len = 0;
    for (int x = 0; x < 8; ++x) len = std::max(len, q[x].size());
    const size_t at = items.size();
    for (int x = 0; x < 8; ++x)
      for (size_t j = 0; j < len; ++j) items.push_back(j < q[x].size() ? q[x][j] : make_int4(3, 0, 0, 0));
    // the chunks tile the batch's pairs exactly once
    std::vector<std::pair<int32_t, int32_t>> tiles;
    for (size_t j = at; j < items.size(); ++j)
      if (items[j].x == 0) tiles.emplace_back(items[j].y, items[j].z);
    std::sort(tiles.begin(), tiles.end());
    int64_t next = 0;
    for (const auto& t : tiles) {
      if (t.first != next || t.second <= 0) return fail(c, GR_ESTATE, "fused batch items do not tile the pairs");
      next += t.second;
    }
    if (next != m.P) return fail(c, GR_ESTATE, "fused batch items do not tile the pairs");
    f.item_ofs = (int32_t)at;
    f.n_items = (int32_t)(items.size() - at);
    f.qlen = (int32_t)len;
    f.n_tree = n_tree;
    f.cx_ofs = (int32_t)cx0;
    f.W = W;
    f.cap = m.max_large;
    f.scap = scap;
    f.lds = lds;
    f.ok = n_tree > 0;
    max_large = std::max(max_large, n_large);
  }
  int rc;
  if (items.size() > s.cap_items) {
    if ((rc = dalloc(c, &s.d_items, items.size()))) return rc;
    s.cap_items = items.size();
  }
  if (cxs.size() > s.cap_cx) {
    if ((rc = dalloc(c, &s.d_cx, cxs.size()))) return rc;
    s.cap_cx = cxs.size();
  }
  if ((size_t)max_large + 26 > s.cap_fcnt) {
    if ((rc = dalloc(c, &s.d_fcnt, (size_t)max_large + 26))) return rc;
    s.cap_fcnt = (size_t)max_large + 26;
  }
  if (!items.empty())
    HIPC(c, hipMemcpy(s.d_items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice));
  if (!cxs.empty())
    HIPC(c, hipMemcpy(s.d_cx, cxs.data(), sizeof(SmallClasses) * cxs.size(), hipMemcpyHostToDevice));
  HIPC(c, hipMemset(s.d_fcnt, 0, sizeof(int32_t) * s.cap_fcnt));
  return 0;
}

bool fused_applies(const gr_ctx* c, const Sched& s, int b) {
  return c->tune.fork == 4 && s.fused_built && !c->prop && !c->collect && (size_t)b < s.fused.size() && s.fused[b].ok;
}

int enqueue_batch(gr_ctx* c, Sched& s, int b) {
  const BatchMeta& m = s.meta[b];
  if (m.n == 0) return 0;
  float* pcost = s.ext ? s.x_pair_cost : c->d_pair_cost;
  int16_t* ppat = s.ext ? s.x_pair_pat : c->d_pair_pat;
  if (fused_applies(c, s, b)) {
    const FusedBatch& f = s.fused[b];
    TreeArgs A;
    A.g = c->g;
    A.list = nullptr;
    A.n_list = 0;
    A.pinofs = s.d_pinofs + m.begin;
    A.deg = s.d_deg + m.begin;
    A.slot = s.d_slot + m.begin;
    A.pair = s.d_pair + m.begin;
    A.cost_base = m.begin;
    A.pins = c->d_pins;
    A.pair_cost = pcost;
    A.pair_pat = ppat;
    A.out_edges = s.d_edges;
    A.out_lpat = s.d_lpat;
    A.out_cost = s.d_cost;
    A.mark = c->d_mark;
    A.run_stamp = c->d_run_stamp;
    A.ctr = c->d_ctr;
    A.quirk = c->quirk;
    A.M = c->M;
    A.true_mst = c->true_mst;
    A.prefix = 0;
    A.vhalf = nullptr;
    A.sqofs = nullptr;
    A.sq = nullptr;
    A.rowmin = nullptr;
    A.rowmin_base = 0;
    FusedArgs F;
    F.items = s.d_items + f.item_ofs;
    F.qlen = f.qlen;
    F.heads = s.d_fcnt + 18;
    F.cx = s.d_cx + f.cx_ofs;
    F.cnt = s.d_fcnt;
    F.n_tree = f.n_tree;
    F.ends = s.d_ends + m.pair_base;
    F.pfx = c->d_pfx;
    F.out_cost = pcost;
    F.out_pat = ppat;
    F.ctr = c->d_ctr;
    {
      KTimer t(c, GR_K_TREE);
      hipLaunchKernelGGL(batch_kernel_for(c->g.Z), dim3(f.n_items), dim3(c->tune.large_threads), f.lds, c->stream, A,
                         F, f.cap, f.scap, f.W);
    }
    {
      KTimer t(c, GR_K_REFRESH);
      if (m.rlist_ofs >= 0 && s.d_rlist) launch_refresh(c, 0, m.runs, s.d_rlist + m.rlist_ofs, m.nrl, F.heads);
      else launch_refresh(c, 0, m.runs, nullptr, 0, F.heads);
    }
    HIPC(c, hipGetLastError());
    c->last_P = m.P;
    c->last_pair_cost = pcost;
    c->last_pair_pat = ppat;
    return 0;
  }
  // square key matrices: written by the exhaustive BIT costing kernels only
  SquareOut sq{nullptr, nullptr, nullptr};
  if (m.streams && s.d_sq && !c->prop && !c->prefix) sq = SquareOut{s.d_sqofs + m.begin, s.d_deg + m.begin, s.d_sq};
  if (m.P > 0 && c->prop) {
    KTimer t(c, GR_K_LPAT);
    const int rc = mlp_enqueue_costing(c, s, m);
    if (rc) return rc;
  } else if (m.P > 0 && c->prefix) {
    KTimer t(c, GR_K_LPAT);
    PrefixFn fn = c->tune.lpat_wave ? nullptr : prefix_kernel(c->g.Z);
    if (fn) {
      const int lb = lpat_block_for(c->tune.lpat_block, m.P);
      hipLaunchKernelGGL(fn, dim3((uint32_t)((m.P + lb - 1) / lb)), dim3(lb), 0, c->stream, c->g, c->d_pfx,
                         c->d_vterm, c->d_pins, s.d_pinofs + m.begin, s.d_pair + m.begin, s.d_chunk + m.chunk_ofs, m.n,
                         m.P, pcost, ppat, c->d_ctr);
    } else {
      hipLaunchKernelGGL(k_lpat_prefix_wave, dim3((uint32_t)((m.P + 3) / 4)), dim3(256), 0, c->stream, c->g, c->d_pfx,
                         c->d_vterm, c->d_pins, s.d_pinofs + m.begin, s.d_pair + m.begin, s.d_chunk + m.chunk_ofs, m.n,
                         m.P, pcost, ppat, c->d_ctr);
    }
  } else if (m.P > 0) {
    KTimer t(c, GR_K_LPAT);
    LpatFn fn = c->tune.lpat_wave ? nullptr : lane_kernel(c->g);
    EndsFn efn = (fn && s.ends_ok) ? ends_kernel(c->g) : nullptr;
    if (efn) {
      const int lb = lpat_block_for(c->tune.lpat_block, m.P);
      hipLaunchKernelGGL(efn, dim3((uint32_t)((m.P + lb - 1) / lb)), dim3(lb), 0, c->stream, c->g, c->d_pfx,
                         s.d_ends + m.pair_base, s.d_pinofs + m.begin, s.d_pair + m.begin, s.d_chunk + m.chunk_ofs,
                         m.n, m.P, pcost, ppat, c->d_ctr, sq);
    } else if (fn) {
      const int lb = lpat_block_for(c->tune.lpat_block, m.P);
      hipLaunchKernelGGL(fn, dim3((uint32_t)((m.P + lb - 1) / lb)), dim3(lb), 0, c->stream, c->g, c->d_pfx,
                         c->d_pins, s.d_pinofs + m.begin, s.d_pair + m.begin, s.d_chunk + m.chunk_ofs, m.n, m.P,
                         pcost, ppat, c->d_ctr, sq);
    } else {
      hipLaunchKernelGGL(k_lpat_wave, dim3((uint32_t)((m.P + 3) / 4)), dim3(256), 0, c->stream, c->g, c->d_pfx,
                         c->d_pins, s.d_pinofs + m.begin, s.d_pair + m.begin, s.d_chunk + m.chunk_ofs, m.n, m.P,
                         pcost, ppat, c->d_ctr, sq);
    }
  }
  if (m.P > 0 && c->collect) {
    const int rc = mlp_enqueue_collect(c, s, m);
    if (rc) return rc;
  }
  TreeArgs A;
  A.g = c->g;
  A.pinofs = s.d_pinofs + m.begin;
  A.deg = s.d_deg + m.begin;
  A.slot = s.d_slot + m.begin;
  A.pair = s.d_pair + m.begin;
  A.cost_base = m.begin;
  A.pins = c->d_pins;
  A.pair_cost = pcost;
  A.pair_pat = ppat;
  A.out_edges = s.d_edges;
  A.out_lpat = s.d_lpat;
  A.out_cost = s.d_cost;
  A.sqofs = sq.ofs;
  A.sq = sq.key;
  A.rowmin = nullptr;
  A.rowmin_base = 0;
  A.mark = c->d_mark;
  A.run_stamp = c->d_run_stamp;
  A.ctr = c->d_ctr;
  A.quirk = c->quirk;
  A.M = c->M;
  A.true_mst = c->true_mst;
  A.prefix = c->prefix;
  A.vhalf = c->d_vhalf;
  {
    KTimer t(c, GR_K_TREE);
    if (sq.key && m.n_rowtask && s.d_rowmin) {
      // first Boruvka round of the square-matrix nets over the whole GPU:
      // every pin's minimum over its row (no component is merged yet)
      hipLaunchKernelGGL(k_row_min, dim3(m.n_rowtask), dim3(64 * kRowTaskPins), 0, c->stream,
                         s.d_rowtask + m.rowtask_ofs, s.d_deg + m.begin, s.d_slot + m.begin, s.d_sqofs + m.begin,
                         (const uint32_t*)s.d_sq, s.d_rowmin, m.slot0);
      A.rowmin = s.d_rowmin;
      A.rowmin_base = m.slot0;
    }
    const int rc = launch_trees(c, s, m, A);
    if (rc) return rc;
  }
  {
    KTimer t(c, GR_K_REFRESH);
    if (m.rlist_ofs >= 0 && s.d_rlist) launch_refresh(c, 0, m.runs, s.d_rlist + m.rlist_ofs, m.nrl);
    else launch_refresh(c, 0, m.runs);
  }
  HIPC(c, hipGetLastError());
  c->last_P = m.P;
  c->last_pair_cost = pcost;
  c->last_pair_pat = ppat;
  return 0;
}

void free_sched(Sched& s) {
  if (s.exec) (void)hipGraphExecDestroy(s.exec);
  (void)hipFree(s.d_pinofs); (void)hipFree(s.d_deg); (void)hipFree(s.d_slot); (void)hipFree(s.d_pair); (void)hipFree(s.d_list);
  (void)hipFree(s.d_chunk); (void)hipFree(s.d_edges); (void)hipFree(s.d_lpat); (void)hipFree(s.d_cost);
  (void)hipFree(s.d_giant); (void)hipFree(s.d_sqofs); (void)hipFree(s.d_sq); (void)hipFree(s.d_rowtask);
  (void)hipFree(s.d_rowmin); (void)hipFree(s.d_ends); (void)hipFree(s.d_rlist);
  (void)hipFree(s.d_items); (void)hipFree(s.d_fcnt); (void)hipFree(s.d_cx);
  s = Sched();
}

}  // namespace

namespace grh {
int refresh_all_runs(gr_ctx* c) { return refresh_all(c); }
}  // namespace grh


extern "C" {

const char* gr_version(void) { return "iitmglobalrouter_amd 0.2 (gfx950)"; }

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

namespace {

// Shared set-up of gr_create / gr_create_f32: exactly one of cap, capf is set.
int create_ctx(const gr_grid_desc* d, const int32_t* cap, const float* capf, gr_ctx** out) {
  if (!d || (!cap && !capf) || !out) return fail(nullptr, GR_EINVAL, "null argument");
  *out = nullptr;
  if (d->X < 1 || d->Y < 1 || d->X > 8192 || d->Y > 8192 || d->Z < 3 || d->Z > 32)
    return fail(nullptr, GR_ERANGE, "grid must have 1 <= X, Y <= 8192 and 3 <= Z <= 32");
  const int64_t cs = (int64_t)d->X * d->Y * d->Z;
  if (cs >= (1ll << 31)) return fail(nullptr, GR_ERANGE, "grid has 2^31 or more cells");
  if (d->cost_model < 0 || d->cost_model > 2 || d->commit_mode < 0 || d->commit_mode > 2)
    return fail(nullptr, GR_EINVAL, "unknown cost model or commit mode");
  if (d->commit_mode == GR_COMMIT_PREFIX && d->cost_model != GR_COST_REF_INC)
    return fail(nullptr, GR_EINVAL, "the non-BIT mode costs the reference's increments: use GR_COST_REF_INC");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= d->device || d->device < 0)
    return fail(nullptr, GR_ENODEV, "no HIP device");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, d->device) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(nullptr, GR_ENODEV, "device is not gfx950");
  gr_ctx* c = new gr_ctx();
  c->dev = d->device;
  if (hipSetDevice(c->dev) != hipSuccess) { delete c; return fail(nullptr, GR_ENODEV, "hipSetDevice failed"); }
  {
    const int rc0 = scratch_address(c->dev, &c->scratch);
    if (rc0) { delete c; return rc0; }
  }
  c->g.X = d->X; c->g.Y = d->Y; c->g.Z = d->Z; c->g.Z2 = (d->Z + 1) / 2;
  c->g.XY = (uint32_t)d->X * (uint32_t)d->Y;
  c->model = d->cost_model;
  c->quirk = d->commit_mode == GR_COMMIT_REF_QUIRK;
  c->prefix = d->commit_mode == GR_COMMIT_PREFIX;
  c->cs = cs;
  c->M = std::max(d->X, d->Y);
  build_lut(c->model, c->lut);
  int rc = 0;
  auto bail = [&](int code) { gr_destroy(c); return code; };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess)
    return bail(fail(nullptr, GR_ENODEV, "stream"));
  const size_t nr = (size_t)d->Z * c->M;
  if ((rc = dalloc(c, &c->d_rem, cs)) || (rc = dalloc(c, &c->d_cap0, cs)) || (rc = dalloc(c, &c->d_mark, cs)) ||
      (rc = dalloc(c, &c->d_pfx_mem, cs + 4)) || (rc = dalloc(c, &c->d_lut, kLut)) || (rc = dalloc(c, &c->d_run_stamp, nr)) ||
      (rc = dalloc(c, &c->d_ctr, 1)) ||
      ((capf || c->prefix) && ((rc = dalloc(c, &c->d_remf, cs)) || (rc = dalloc(c, &c->d_capf0, cs)))) ||
      (c->prefix && ((rc = dalloc(c, &c->d_vhalf, cs)) || (rc = dalloc(c, &c->d_vterm, cs))))) {
    g_err = c->err;
    return bail(rc);
  }
  c->d_pfx = c->d_pfx_mem + 4;  // 16-byte aligned: the refresh writes aligned quads
  c->lds_max = prop.sharedMemPerBlock;
  bool ok = hipFuncSetAttribute((const void*)k_refresh<false, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)refresh_lds(c, false)) == hipSuccess &&
            hipFuncSetAttribute((const void*)k_refresh<false, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)refresh_lds(c, true)) == hipSuccess &&
            hipFuncSetAttribute((const void*)k_refresh<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)refresh_lds(c, false)) == hipSuccess &&
            hipFuncSetAttribute((const void*)k_refresh_prefix, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)refresh_prefix_lds(c)) == hipSuccess;
  if (c->prefix && refresh_prefix_lds(c) > c->lds_max) ok = false;
  if (cap) {
    ok = ok && hipMemcpy(c->d_cap0, cap, sizeof(int32_t) * cs, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(c->d_rem, cap, sizeof(int32_t) * cs, hipMemcpyHostToDevice) == hipSuccess;
    if (c->prefix) {  // the non-BIT state is the fp32 remaining capacity
      std::vector<float> f(cap, cap + cs);
      ok = ok && hipMemcpy(c->d_capf0, f.data(), sizeof(float) * cs, hipMemcpyHostToDevice) == hipSuccess &&
           hipMemcpy(c->d_remf, f.data(), sizeof(float) * cs, hipMemcpyHostToDevice) == hipSuccess;
    }
  } else {
    // fp32 grid: the integer state starts at 0 and counts commits downwards,
    // so cap0 - rem is still the demand
    ok = ok && hipMemset(c->d_cap0, 0, sizeof(int32_t) * cs) == hipSuccess &&
         hipMemset(c->d_rem, 0, sizeof(int32_t) * cs) == hipSuccess &&
         hipMemcpy(c->d_capf0, capf, sizeof(float) * cs, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(c->d_remf, capf, sizeof(float) * cs, hipMemcpyHostToDevice) == hipSuccess;
  }
  if (!ok ||
      hipFuncSetAttribute((const void*)k_tree_large<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)c->lds_max - 2048) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_tree_large<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)c->lds_max - 2048) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_tree_mixed<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)c->lds_max - 2048) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_tree_mixed<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)c->lds_max - 2048) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_tree_giant, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)c->lds_max - 4096) != hipSuccess ||
      (batch_kernel_for(c->g.Z) &&
       hipFuncSetAttribute((const void*)batch_kernel_for(c->g.Z), hipFuncAttributeMaxDynamicSharedMemorySize,
                           (int)c->lds_max - 2048) != hipSuccess) ||
      hipMemset(c->d_ctr, 0, sizeof(int32_t)) != hipSuccess ||
      hipMemset(c->d_mark, 0, sizeof(uint8_t) * cs) != hipSuccess ||
      hipMemset(c->d_run_stamp, 0, sizeof(int32_t) * nr) != hipSuccess ||
      hipMemset(c->d_pfx_mem, 0, 4 * sizeof(float)) != hipSuccess ||
      (c->d_vhalf && hipMemset(c->d_vhalf, 0, sizeof(int32_t) * cs) != hipSuccess) ||
      hipMemcpy(c->d_lut, c->lut, sizeof(float) * kLut, hipMemcpyHostToDevice) != hipSuccess) {
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

}  // namespace

int gr_create(const gr_grid_desc* d, const int32_t* cap, gr_ctx** out) {
  if (!cap) return fail(nullptr, GR_EINVAL, "null argument");
  return create_ctx(d, cap, nullptr, out);
}

int gr_create_f32(const gr_grid_desc* d, const float* cap, gr_ctx** out) {
  if (!cap) return fail(nullptr, GR_EINVAL, "null argument");
  return create_ctx(d, nullptr, cap, out);
}

void gr_destroy(gr_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->dev);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  mlp_detach_all(c);
  (void)hipFree(c->d_rem); (void)hipFree(c->d_cap0); (void)hipFree(c->d_mark); (void)hipFree(c->d_run_stamp);
  (void)hipFree(c->d_pfx_mem); (void)hipFree(c->d_vterm); (void)hipFree(c->d_lut); (void)hipFree(c->d_ctr);
  (void)hipFree(c->d_remf); (void)hipFree(c->d_capf0); (void)hipFree(c->d_vhalf);
  (void)hipFree(c->d_wcnt); (void)hipFree(c->d_vcnt); (void)hipFree(c->d_estamp);
  (void)hipFree(c->d_eval_part); (void)hipFree(c->d_eval_out);
  (void)hipFree(c->d_pin_pfx); (void)hipFree(c->d_pins);
  (void)hipFree(c->d_pair_cost); (void)hipFree(c->d_pair_pat); (void)hipFree(c->d_pair_fb);
  (void)hipFree(c->d_remz); (void)hipFree(c->d_remz0);
  free_sched(c->sched);
  free_sched(c->single);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
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
  HIPC(c, hipStreamSynchronize(c->stream));
  if ((rc = dalloc(c, &c->d_pin_pfx, n_nets + 1)) || (rc = dalloc(c, &c->d_pins, np))) return rc;
  HIPC(c, hipMemcpy(c->d_pin_pfx, pin_pfx, sizeof(int64_t) * (n_nets + 1), hipMemcpyHostToDevice));
  if (np) HIPC(c, hipMemcpy(c->d_pins, pin_enc, sizeof(int32_t) * np, hipMemcpyHostToDevice));
  c->h_pin_pfx.assign(pin_pfx, pin_pfx + n_nets + 1);
  // runs a net's routes can touch: the rows of its box on even layers and
  // the columns on odd ones, each with one more for the quirk cells that fall
  // just before a first wire's low end (their enc arithmetic can reach the
  // previous run)
  c->h_net_runs.assign(n_nets, 0);
  c->h_net_box.assign(4 * (size_t)n_nets, 0);
  for (int32_t i = 0; i < n_nets; ++i) {
    int x0 = c->g.X, x1 = -1, y0 = c->g.Y, y1 = -1;
    for (int64_t p = pin_pfx[i]; p < pin_pfx[i + 1]; ++p) {
      const uint32_t v = (uint32_t)pin_enc[p];
      const int z = (int)(v / c->g.XY);
      const uint32_t r = v - (uint32_t)z * c->g.XY;
      const int x = (z & 1) ? (int)(r / c->g.Y) : (int)(r % c->g.X);
      const int y = (z & 1) ? (int)(r % c->g.Y) : (int)(r / c->g.X);
      x0 = std::min(x0, x); x1 = std::max(x1, x);
      y0 = std::min(y0, y); y1 = std::max(y1, y);
    }
    const int64_t even = (c->g.Z + 1) / 2, odd = c->g.Z / 2;
    c->h_net_runs[i] = (int32_t)std::min<int64_t>(even * (y1 - y0 + 2) + odd * (x1 - x0 + 2), (int64_t)c->g.Z * c->M);
    int32_t* box = &c->h_net_box[4 * (size_t)i];
    box[0] = x0; box[1] = x1; box[2] = y0; box[3] = y1;
  }
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

int gr_route_batch_static(gr_ctx* c, const int32_t* net_ids, int32_t n, int64_t* out_edges, int32_t* out_lpat,
                          float* out_net_cost) {
  if (!c || n < 0 || (n && !net_ids)) return fail(c, GR_EINVAL, "bad batch");
  if (!c->d_pins) return fail(c, GR_ESTATE, "gr_set_nets must come first");
  if (c->prop || c->collect)
    return fail(c, GR_ESTATE, "the allocation-free entry point routes with exhaustive costing and no sample capture");
  if (!c->scratch) return fail(c, GR_ESTATE, "no static scratch address for this device");
  (void)hipSetDevice(c->dev);
  Sched st;
  HostSched& h = c->hst;
  int32_t ofs[2] = {0, n};
  int rc = host_sched(c, st, h, 1, ofs, net_ids);
  if (rc) return rc;
  if (st.giant_bytes)
    return fail(c, GR_ERANGE, "nets of more than 4096 pins need device memory for their trees (use gr_route_batch)");
  // carve the static array: descriptors, chunk owners, outputs, pair results
  const size_t npos = h.slot.size(), nch = h.chunk.size();
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += (bytes + 255) & ~(size_t)255;
    return at;
  };
  const size_t o_pinofs = take(8 * npos), o_slot = take(8 * npos), o_pair = take(8 * npos);
  const size_t o_deg = take(4 * npos), o_list = take(4 * npos), o_cost = take(4 * npos);
  const size_t o_chunk = take(4 * nch);
  const size_t o_edges = take(8 * (size_t)h.slots), o_lpat = take(4 * (size_t)h.slots);
  const size_t o_pc = take(4 * (size_t)h.maxP), o_pp = take(2 * (size_t)h.maxP);
  if (off > kScratchBytes)
    return fail(c, GR_ERANGE,
                "batch needs " + std::to_string(off) + " bytes of device scratch; the allocation-free entry point has " +
                    std::to_string(kScratchBytes) + " (use gr_route_batch)");
  unsigned char* base = (unsigned char*)c->scratch;
  st.ext = true;
  st.d_pinofs = (int64_t*)(base + o_pinofs);
  st.d_slot = (int64_t*)(base + o_slot);
  st.d_pair = (int64_t*)(base + o_pair);
  st.d_deg = (int32_t*)(base + o_deg);
  st.d_list = (int32_t*)(base + o_list);
  st.d_cost = (float*)(base + o_cost);
  st.d_chunk = (int32_t*)(base + o_chunk);
  st.d_edges = (int64_t*)(base + o_edges);
  st.d_lpat = (int32_t*)(base + o_lpat);
  st.x_pair_cost = (float*)(base + o_pc);
  st.x_pair_pat = (int16_t*)(base + o_pp);
  if ((rc = upload_sched(c, st, h))) return rc;
  // per-kernel timing records events, which may allocate: not on this path
  const bool timing = c->timing;
  c->timing = false;
  rc = enqueue_batch(c, st, 0);
  c->timing = timing;
  if (rc) return rc;
  const int64_t slots = st.slots;
  if (out_edges && slots)
    HIPC(c, hipMemcpyAsync(out_edges, st.d_edges, sizeof(int64_t) * slots, hipMemcpyDeviceToHost, c->stream));
  if (out_lpat && slots)
    HIPC(c, hipMemcpyAsync(out_lpat, st.d_lpat, sizeof(int32_t) * slots, hipMemcpyDeviceToHost, c->stream));
  if (out_net_cost && n)
    HIPC(c, hipMemcpyAsync(out_net_cost, st.d_cost, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return 0;
}

int64_t gr_static_scratch_bytes(void) { return (int64_t)kScratchBytes; }

int gr_set_schedule(gr_ctx* c, int32_t n_batches, const int32_t* batch_ofst, const int32_t* net_ids) {
  if (!c || n_batches < 0 || !batch_ofst || (batch_ofst[n_batches] && !net_ids)) return fail(c, GR_EINVAL, "bad schedule");
  if (!c->d_pins) return fail(c, GR_ESTATE, "gr_set_nets must come first");
  if (batch_ofst[0] != 0) return fail(c, GR_EINVAL, "batch_ofst[0] must be 0");
  (void)hipSetDevice(c->dev);
  const int rc = build_sched(c, c->sched, n_batches, batch_ofst, net_ids);
  if (rc) return rc;
  c->sched.h_ofst.assign(batch_ofst, batch_ofst + n_batches + 1);
  c->sched.h_ids.assign(net_ids, net_ids + batch_ofst[n_batches]);
  return 0;
}

int gr_route_schedule(gr_ctx* c, int32_t first, int32_t count) {
  if (!c) return fail(c, GR_EINVAL, "null context");
  if (first < 0 || count < 0 || (size_t)first + count > c->sched.meta.size())
    return fail(c, GR_EINVAL, "batch range outside the schedule");
  (void)hipSetDevice(c->dev);
  Sched& s = c->sched;
  if (!s.fused_built) {
    const int rc = build_fused(c, s);
    if (rc) return rc;
  }
  const bool whole = first == 0 && count == (int32_t)s.meta.size() && count > 0;
  if (whole && !c->timing && (!c->prop || mlp_fused(c->prop)) && !c->collect && !s.graph_failed) {
    if (!s.exec) {
      // capture the whole pass once: kernels only (no memset nodes), fixed
      // arguments; the batch stamp comes from the device counter
      hipGraph_t graph = nullptr;
      bool ok = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
      int rc = 0;
      if (ok) {
        for (int32_t b = 0; b < count && !rc; ++b) rc = enqueue_batch(c, s, b);
        ok = hipStreamEndCapture(c->stream, &graph) == hipSuccess && !rc;
      }
      if (ok) ok = hipGraphInstantiate(&s.exec, graph, nullptr, nullptr, 0) == hipSuccess;
      if (graph) (void)hipGraphDestroy(graph);
      if (!ok) {
        s.exec = nullptr;
        s.graph_failed = true;
        (void)hipGetLastError();
      }
    }
    if (s.exec) {
      HIPC(c, hipGraphLaunch(s.exec, c->stream));
      c->last_P = s.meta.back().P;
      c->last_pair_cost = c->d_pair_cost;
      c->last_pair_pat = c->d_pair_pat;
      return 0;
    }
  }
  for (int32_t b = first; b < first + count; ++b) {
    const int rc = enqueue_batch(c, s, b);
    if (rc) return rc;
  }
  return 0;
}

int gr_sync(gr_ctx* c) {
  if (!c) return fail(c, GR_EINVAL, "null context");
  HIPC(c, hipStreamSynchronize(c->stream));
  if (c->sched.d_fcnt) {
    int32_t err = 0;
    HIPC(c, hipMemcpy(&err, c->sched.d_fcnt + 1, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (err) {
      (void)hipMemset(c->sched.d_fcnt, 0, sizeof(int32_t) * c->sched.cap_fcnt);
      return fail(c, GR_ESTATE, "a fused batch launch timed out waiting for its costing chunks; its routes are invalid");
    }
  }
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
  if (c->remz_live)
    HIPC(c, hipMemcpyAsync(c->d_remz, c->d_remz0, sizeof(int16_t) * 16 * c->g.XY, hipMemcpyDeviceToDevice,
                           c->stream));
  if (c->d_remf)
    HIPC(c, hipMemcpyAsync(c->d_remf, c->d_capf0, sizeof(float) * c->cs, hipMemcpyDeviceToDevice, c->stream));
  return refresh_all(c);
}

int gr_get_remaining_f32(gr_ctx* c, float* out) {
  if (!c || !out) return fail(c, GR_EINVAL, "null argument");
  if (!c->d_remf) return fail(c, GR_ESTATE, "the grid has integer capacities (gr_get_demand)");
  (void)hipSetDevice(c->dev);
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipMemcpy(out, c->d_remf, sizeof(float) * c->cs, hipMemcpyDeviceToHost));
  return 0;
}

int gr_get_demand(gr_ctx* c, int32_t* out) {
  if (!c || !out) return fail(c, GR_EINVAL, "null argument");
  if (c->prefix)
    return fail(c, GR_ESTATE, "the non-BIT mode's state is fp32 remaining capacity in half units: gr_get_remaining_f32");
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

int gr_get_via_terms(gr_ctx* c, float* out) {
  if (!c || !out) return fail(c, GR_EINVAL, "null argument");
  HIPC(c, hipStreamSynchronize(c->stream));
  if (c->prefix) {
    HIPC(c, hipMemcpy(out, c->d_vterm, sizeof(float) * c->cs, hipMemcpyDeviceToHost));
    return 0;
  }
  // BIT modes: Q(pos) - Q(pos - 1) of the stored prefix, the fp32
  // subtraction the costing kernels make
  std::vector<float> q((size_t)c->cs);
  HIPC(c, hipMemcpy(q.data(), c->d_pfx, sizeof(float) * c->cs, hipMemcpyDeviceToHost));
  for (int z = 0; z < c->g.Z; ++z) {
    const int L = (z & 1) ? c->g.Y : c->g.X, lines = (z & 1) ? c->g.X : c->g.Y;
    for (int l = 0; l < lines; ++l) {
      const size_t b = (size_t)z * c->g.XY + (size_t)l * L;
      out[b] = q[b];
      for (int x = 1; x < L; ++x) out[b + x] = q[b + x] - q[b + x - 1];
    }
  }
  return 0;
}

int64_t gr_last_pairs(const gr_ctx* c) { return c ? c->last_P : 0; }

int gr_get_pair_costs(gr_ctx* c, float* cost, int16_t* pat) {
  if (!c) return fail(c, GR_EINVAL, "null context");
  HIPC(c, hipStreamSynchronize(c->stream));
  if (cost && c->last_P)
    HIPC(c, hipMemcpy(cost, c->last_pair_cost, sizeof(float) * c->last_P, hipMemcpyDeviceToHost));
  if (pat && c->last_P)
    HIPC(c, hipMemcpy(pat, c->last_pair_pat, sizeof(int16_t) * c->last_P, hipMemcpyDeviceToHost));
  return 0;
}

int gr_get_lut(const gr_ctx* c, float* out) {
  if (!c || !out) return GR_EINVAL;
  std::memcpy(out, c->lut, sizeof(float) * kLut);
  return 0;
}

int gr_get_tuning(const gr_ctx* c, gr_tuning* t) {
  if (!c || !t) return GR_EINVAL;
  *t = c->tune;
  return 0;
}

int gr_set_tuning(gr_ctx* c, const gr_tuning* t) {
  if (!c || !t) return fail(c, GR_EINVAL, "null argument");
  if ((t->large_threads != 256 && t->large_threads != 512 && t->large_threads != 1024) ||
      (t->group_max != 32 && t->group_max != 64) || t->refresh_blocks < 1 || t->fork < 0 || t->fork > 4 ||
      (t->lpat_wave != 0 && t->lpat_wave != 1) ||
      (t->lpat_block != 0 && t->lpat_block != 64 && t->lpat_block != 128 && t->lpat_block != 256))
    return fail(c, GR_EINVAL, "tuning out of range");
  (void)hipSetDevice(c->dev);
  HIPC(c, hipStreamSynchronize(c->stream));
  const bool reclass = t->group_max != c->tune.group_max;
  c->tune = *t;
  c->tune.fork = t->fork;
  Sched& s = c->sched;
  if (s.exec) { (void)hipGraphExecDestroy(s.exec); s.exec = nullptr; }
  s.graph_failed = false;
  s.fused_built = false;
  if (reclass && !s.meta.empty()) {
    const std::vector<int32_t> ofst = s.h_ofst, ids = s.h_ids;
    return build_sched(c, s, (int32_t)ofst.size() - 1, ofst.data(), ids.data());
  }
  return 0;
}

int gr_set_tree_mode(gr_ctx* c, int32_t mode) {
  if (!c) return fail(c, GR_EINVAL, "null context");
  if (mode != GR_TREE_ROOT_PAIR && mode != GR_TREE_TRUE_MST) return fail(c, GR_EINVAL, "unknown tree mode");
  (void)hipSetDevice(c->dev);
  HIPC(c, hipStreamSynchronize(c->stream));
  c->true_mst = mode == GR_TREE_TRUE_MST;
  Sched& s = c->sched;
  if (s.exec) { (void)hipGraphExecDestroy(s.exec); s.exec = nullptr; }  // the graph holds the old mode
  s.graph_failed = false;
  return 0;
}

int gr_autotune(gr_ctx* c, int32_t reps, gr_tuning* best, double* best_ms) {
  if (!c || reps < 1) return fail(c, GR_EINVAL, "bad autotune request");
  if (c->sched.meta.empty()) return fail(c, GR_ESTATE, "gr_set_schedule must come first");
  (void)hipSetDevice(c->dev);
  const bool timing = c->timing;
  c->timing = false;
  hipEvent_t e0, e1;
  HIPC(c, hipEventCreate(&e0));
  HIPC(c, hipEventCreate(&e1));
  const gr_tuning start = c->tune;
  gr_tuning win = start;
  double win_ms = 1e30;
  const int nb = (int)c->sched.meta.size();
  int rc = 0;
  // one untimed pass captures the graph for the candidate, then reps timed passes
  auto time_one = [&](const gr_tuning& t) -> int {
    int r0 = gr_set_tuning(c, &t);
    if (r0 || (r0 = gr_reset_demand(c)) || (r0 = gr_route_schedule(c, 0, nb))) return r0;
    float ms = 0.f;
    double tot = 0.0;
    for (int r = 0; r < reps; ++r) {
      if ((r0 = gr_reset_demand(c))) return r0;
      if (hipEventRecord(e0, c->stream) != hipSuccess) return fail(c, GR_ENODEV, "event");
      if ((r0 = gr_route_schedule(c, 0, nb))) return r0;
      if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
        return fail(c, GR_ENODEV, "event timing failed");
      tot += ms;
    }
    if (tot / reps < win_ms) { win_ms = tot / reps; win = t; }
    return 0;
  };
  for (int gm : {64, 32})
    for (int lt : {1024, 512, 256})
      for (int fk : {4, 3, 1, 0})
        for (int rb : {1024, 2048, 4096, 8192})
          for (int lw : {0, 1}) {
            if (lw && !lane_kernel(c->g)) continue;  // only the wave kernel exists for this Z
            if ((rc = time_one(gr_tuning{lt, gm, fk, rb, lw, 0}))) goto done;
          }
  // then the lane kernel's workgroup size, around the winner
  if (!win.lpat_wave)
    for (int lb : {256, 128, 64}) {
      gr_tuning t = win;
      t.lpat_block = lb;
      if ((rc = time_one(t))) goto done;
    }
done:
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->timing = timing;
  if (rc) { gr_set_tuning(c, &start); return rc; }
  if ((rc = gr_set_tuning(c, &win)) || (rc = gr_reset_demand(c))) return rc;
  HIPC(c, hipStreamSynchronize(c->stream));
  if (best) *best = win;
  if (best_ms) *best_ms = win_ms;
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
