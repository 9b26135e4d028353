// Internal state shared by the routing (gr_route.hip) and learned-proposal
// (gr_mlp.hip) translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/iitmgr.h"
#include "gr_device.h"

namespace grh {

using grd::Geo;
using grd::kLut;

// Net size classes of the tree stage: 2 pins; up to 8, 16, 32, 64 pins
// (one group of that many lanes per net); more (one workgroup per net, node
// arrays in LDS); more than kLargeDeg (one workgroup per net, node arrays in
// global memory).
constexpr int kNumCls = 7;
constexpr int kClsLarge = 5;
constexpr int kClsGiant = 6;
constexpr int kClsMax[kNumCls] = {2, 8, 16, 32, 64, grd::kLargeDeg, 1 << 30};
inline int net_class(int64_t d, int group_max) {
  if (d > grd::kLargeDeg) return kClsGiant;
  if (d > group_max) return kClsLarge;
  int k = 0;
  while (d > kClsMax[k]) ++k;
  return k;
}

struct BatchMeta {
  int32_t begin, n;       // schedule positions [begin, begin+n)
  int64_t P;              // pin pairs in the batch
  int32_t chunk_ofs;      // first entry of the batch in the pair-chunk table
  int32_t cls_ofs[kNumCls], cls_n[kNumCls];
  int64_t slot0, nslots;  // output slots of the batch
  int32_t max_large;      // most pins of a large net in the batch
  int64_t max_large_P;    // most pin pairs of a large net in the batch
  int32_t max_giant;      // most pins of a net above kLargeDeg in the batch (0: none)
  int32_t runs;           // bound on the runs the batch's commit can touch (refresh grid)
  int32_t rlist_ofs = -1; // first entry of the batch's candidate-run list (-1: no list, every run)
  int32_t nrl = 0;        // entries of that list
  bool streams;           // the large nets' pair keys do not fit in the trees' LDS
  int32_t rowtask_ofs, n_rowtask;  // first-round row-minimum tasks (16 pins of a square-matrix net each)
  int64_t pair_base;      // first pair of the batch in the schedule's pair order
};

// One batch's fused launch (k_batch): whether it applies, its items in the
// schedule's item array, and the mixed-tree geometry it shares.
struct FusedBatch {
  bool ok = false;
  int32_t item_ofs = 0, n_items = 0, qlen = 0, n_tree = 0;
  int32_t cx_ofs = 0;  // first of the batch's 8 per-XCD small-net classes in d_cx
  int32_t W = 0, cap = 0, scap = 0;
  size_t lds = 0;
};

struct Sched {
  std::vector<BatchMeta> meta;
  int32_t n_pos = 0;
  int64_t slots = 0, max_P = 0;
  int64_t* d_pinofs = nullptr;  // position -> first pin of the net
  int32_t* d_deg = nullptr;     // position -> pin count
  int64_t* d_slot = nullptr;
  int64_t* d_pair = nullptr;
  int32_t* d_list = nullptr;
  int32_t* d_chunk = nullptr;
  int64_t* d_edges = nullptr;
  int32_t* d_lpat = nullptr;
  float* d_cost = nullptr;
  uint8_t* d_giant = nullptr;  // node arrays of the nets above kLargeDeg pins: 44 B per slot of the batch
  // square key matrices of the large nets of batches whose keys stream: key
  // of pair (i, k) at d_sq[d_sqofs[pos] + i*n + k] and [.. + k*n + i]
  // (d_sqofs[pos] = -1: none), written by the costing kernel
  int64_t* d_sqofs = nullptr;
  uint32_t* d_sq = nullptr;
  int2* d_rowtask = nullptr;          // (batch position, first pin) of each row-minimum task
  unsigned long long* d_rowmin = nullptr;  // per output slot: the pin's first-round minimum (key << 32 | pair)
  int32_t* d_rlist = nullptr;  // candidate-run lists of the batches (BatchMeta::rlist_ofs)
  size_t cap_pos = 0, cap_slots = 0, cap_chunk = 0, cap_giant = 0, cap_sq = 0, cap_rowtask = 0, cap_rowmin = 0;
  size_t cap_rlist = 0;
  int64_t giant_bytes = 0;     // largest need over the batches with such nets
  int64_t sq_words = 0;        // largest square-matrix need over the batches
  int64_t rowmin_slots = 0;    // most output slots of a batch with square matrices
  int64_t total_P = 0;         // pin pairs of all batches
  // packed endpoints of every pin pair of the schedule (pair_base order,
  // pack_cell layout), built once by build_sched for the resident schedule
  // when the grid fits the packing; null: the costing kernel finds each
  // pair's net and pins itself
  uint2* d_ends = nullptr;
  size_t cap_ends = 0;
  bool ends_ok = false;
  std::vector<int32_t> h_ofst, h_ids;  // host copy of the schedule (rebuilt when the tuning changes)
  std::vector<int32_t> h_list, h_deg;  // host copies of the class lists and pin counts per position
  std::vector<int64_t> h_pair;         // and of each position's first pair in its batch
  hipGraphExec_t exec = nullptr;  // whole schedule as one graph (built on first full replay)
  bool graph_failed = false;
  // fused batch launches (tuning fork = 4): per-batch geometry and the work
  // items of every batch, built for the tuning in fused_tune
  std::vector<FusedBatch> fused;
  gr_tuning fused_tune{};
  bool fused_built = false;
  int4* d_items = nullptr;
  size_t cap_items = 0;
  int32_t* d_fcnt = nullptr;  // the fused launch's counters and error word (FusedArgs::cnt)
  size_t cap_fcnt = 0;
  grd::SmallClasses* d_cx = nullptr;  // per batch, per XCD: the small-net classes of its fused launch
  size_t cap_cx = 0;
  // ext: every device array above, and the pair buffers below, point into the
  // allocation-free entry point's static array (never freed, never grown)
  bool ext = false;
  float* x_pair_cost = nullptr;
  int16_t* x_pair_pat = nullptr;
};

// Per-position descriptors of a schedule, computed on the host before upload.
struct HostSched {
  std::vector<int64_t> slot, pair, pinofs, sqofs;
  std::vector<int32_t> degv, list, chunk;
  std::vector<int2> rowtask;
  std::vector<int32_t> rlist;  // per batch: the runs its commit can stamp, when fewer than all
  int64_t slots = 0, maxP = 0;
};

struct EvPair { hipEvent_t a, b; int kind; };

}  // namespace grh

struct gr_ctx {
  int dev = 0;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // large-net trees, joined back before the refresh
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  grd::Geo g{};
  int model = 0, quirk = 1, M = 0;
  int prefix = 0;  // GR_COMMIT_PREFIX: the reference's non-BIT mode (prefix costing, difference-array commit)
  int32_t* d_vhalf = nullptr;  // non-BIT mode: via half charges per cell of the current batch
  int true_mst = 0;  // GR_TREE_TRUE_MST
  int64_t cs = 0;
  float lut[grd::kLut];
  int32_t *d_rem = nullptr, *d_cap0 = nullptr, *d_run_stamp = nullptr;
  uint8_t* d_mark = nullptr;  // per cell: marked by this batch's commit
  float *d_pfx = nullptr, *d_lut = nullptr;
  float* d_pfx_mem = nullptr;  // the prefix's allocation: four padding floats (zero), then d_pfx
  float* d_vterm = nullptr;  // non-BIT mode: per cell via cost of the layer, written by the refresh
  int32_t* d_ctr = nullptr;  // device batch counter: value = current batch stamp
  float *d_remf = nullptr, *d_capf0 = nullptr;  // fp32 grids: remaining and initial capacity
  // route evaluator: per-cell wire and via counts, per-batch wire stamps, partial sums
  int32_t *d_wcnt = nullptr, *d_vcnt = nullptr, *d_estamp = nullptr;
  void* d_eval_part = nullptr;
  void* d_eval_out = nullptr;
  size_t lds_max = 65536;
  gr_tuning tune = {1024, 64, 3, 2048, 0, 0};
  int32_t n_nets = 0;
  std::vector<int64_t> h_pin_pfx;
  std::vector<int32_t> h_net_runs;  // per net: bound on the runs its routes can touch
  std::vector<int32_t> h_net_box;   // per net: x0, x1, y0, y1 of its pins' bounding box
  int64_t* d_pin_pfx = nullptr;
  int32_t* d_pins = nullptr;
  grh::Sched sched, single;
  float* d_pair_cost = nullptr;
  int16_t* d_pair_pat = nullptr;
  uint8_t* d_pair_fb = nullptr;  // per pair of the batch: 1 when the proposal fell back to every pattern
  int64_t pair_cap = 0, last_P = 0;
  const float* last_pair_cost = nullptr;  // where the last batch's per-pair results are
  const int16_t* last_pair_pat = nullptr;
  void* scratch = nullptr;  // this device's static array of the allocation-free entry point
  grh::HostSched hst;       // host descriptors of that entry point, reused across calls
  // learned proposal: when set, costing evaluates only the model's top-k patterns
  gr_mlp* prop = nullptr;
  int prop_k = 8;
  float prop_thr = 0.6f;
  int64_t* d_prop_stats = nullptr;  // [pairs costed, pairs that fell back to full enumeration]
  // while a fused proposal is set: every layer's remaining capacity, clamped
  // to the cost table's range, per GCELL (16 int16 at (y*X + x)*16, layer z
  // at +z), kept by the refresh on the cells it decrements; d_remz0 is the
  // copy for the initial capacity (the reset's)
  int16_t* d_remz = nullptr;
  int16_t* d_remz0 = nullptr;
  bool remz_live = false;
  bool vterm_live = false;  // BIT modes: the refresh also writes d_vterm (a fused proposal reads it)
  // training-data capture after each batch's costing
  gr_mlp* collect = nullptr;
  int collect_per_batch = 0;
  // every live model created on this context; gr_destroy detaches them
  std::vector<gr_mlp*> models;
  bool timing = false;
  std::vector<grh::EvPair> evs;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::string err;
};


namespace grh {

extern std::string g_err;

inline int fail(gr_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg; else g_err = msg;
  return code;
}

#define HIPC(c, x)                                                                      \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess)                                                               \
      return grh::fail((c), GR_ENODEV, std::string(#x " failed: ") + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int dalloc(gr_ctx* c, T** p, size_t n) {
  if (*p) { (void)hipFree(*p); *p = nullptr; }
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) { *p = nullptr; return fail(c, GR_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e)); }
  return 0;
}

// Learned-proposal hooks (gr_mlp.hip): costing of one batch by the model's
// top-k patterns, and capture of (features, chosen pattern) training samples.
int mlp_enqueue_costing(gr_ctx* c, Sched& s, const BatchMeta& m);
int mlp_enqueue_collect(gr_ctx* c, Sched& s, const BatchMeta& m);
// Called by gr_destroy: the context's models lose it (their device buffers are
// still freed by gr_mlp_destroy; every other model call then fails).
void mlp_detach_all(gr_ctx* c);
// True when the model's costing is the single fused kernel (no allocation at
// enqueue time, so a pass with it can be captured as a graph).
bool mlp_fused(const gr_mlp* m);
// Drops the captured whole-schedule graph (it holds the old costing set-up).
void drop_graph(gr_ctx* c);
int refresh_all_runs(gr_ctx* c);  // rebuild every run's cost state (and the stored via terms, when live)

}  // namespace grh
