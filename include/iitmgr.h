/*
 * iitmgr.h - C ABI of the MI355X (gfx950) routing hot path.
 *
 * The reference router has no API for this path: its per-batch body lives in
 * TEST_LPAT::test_gpu_lpat (l_pattern_files/test_lpat.hpp:280-450) and talks to
 * its callers through namespace globals (common.h:313-349, mst_common.h:76-88).
 * Each entry point below replaces one piece of that body or of its one-time
 * set-up, as cited per function.  Only plain C types cross this boundary.
 *
 * Conventions
 *   - Every int-returning call returns 0 (or a non-negative count) on success
 *     and a negative GR_E* code on failure; gr_last_error() explains it.
 *     Nothing aborts the process (the reference asserts, common.h:140-146).
 *   - Grid cells use the reference's WIRE_ENC order (util.hpp:38-44):
 *       enc(z,x,y) = z*X*Y + (z odd ? x*Y + y : y*X + x)
 *     Layer 0 is horizontal; layers alternate H/V.
 *   - Pins of a net are deduplicated and ascending by enc
 *     (read_data.hpp:413-432); nets need at least 2 pins.
 *   - A batch is a list of nets whose closed 2-D pin bounding boxes are
 *     pairwise disjoint (part_net/net_part.hpp:167-235).  Batches must be
 *     routed in order: batch b+1 reads the demand committed by batch b.
 *   - Output slots: net i of a batch owns deg(i) consecutive slots; the first
 *     deg-1 hold tree edges encoded (i << 30) | (on1 << 15) | on2 and the last
 *     holds -1 (gmst_util.hpp:322-339).  on1 is the root of the component that
 *     chose the edge, on2 the root of the component it reached; both are pin
 *     indices of the net.  The routed pair is (max(on1,on2), min(on1,on2)).
 *   - Pattern code p = a*Z2 + b, Z2 = ceil(Z/2) (lpattern_util.hpp:43-53):
 *     first wire on layer a, second on b' = (a odd ? 2b : 2b+1).
 *   - One context is driven by one host thread.  The library owns all device
 *     memory and reuses it across batches.
 */
#ifndef IITMGR_H
#define IITMGR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define GR_OK 0
#define GR_EINVAL -1    /* bad argument */
#define GR_ENODEV -2    /* no usable gfx950 device / HIP runtime error */
#define GR_ENOMEM -3    /* device allocation failed */
#define GR_ESTATE -4    /* call out of order (e.g. routing before gr_set_nets) */
#define GR_ERANGE -5    /* size beyond a documented limit */

/* cost models: all map remaining capacity r = cap - demand to fp32 */
#define GR_COST_REF_BIT 0   /* exp(-r/2): bit_util.hpp:13-15 (paper -bit_pmodel run) */
#define GR_COST_REF_INC 1   /* exp(-(r-1)/2) - exp(-r/2): util.hpp:127-129 */
#define GR_COST_LIT_EXP 2   /* exp(d + 1 - c) = exp(1 - r): README.md:10 */

/* commit modes */
#define GR_COMMIT_REF_QUIRK 0 /* reproduce the unguarded first wire, mst_bit_commit.hpp:216 */
#define GR_COMMIT_FIXED 1     /* mark only the cells of the route */
/* The reference's non-BIT mode (-bit_pmodel off), with GR_COST_REF_INC only:
 * costing from an exclusive prefix of per-cell wire costs f(r-1) - f(r)
 * along each run and a per-cell via cost (LPATTERN::get_gpu_lpat_dist,
 * l_pattern_files/lpattern.hpp:13-160, util.hpp:133-353), and the
 * difference-array commit: -1.0 on every cell a wire of the batch covers
 * ([min, max) of its run), -0.5 per via half charge (MST_COMMIT::
 * gpu_commit_edges, mst_files/mst_commit.hpp:17-364).  The state is fp32
 * remaining capacity (gr_get_remaining_f32; gr_get_demand and gr_evaluate
 * give GR_ESTATE).  The reference's via-cost cube is written by racing
 * threads at transposed slots; here the via cost of a layer at a cell is the
 * increase of f over exactly the capacity the commit's via charge takes. */
#define GR_COMMIT_PREFIX 2

/* tree modes */
#define GR_TREE_ROOT_PAIR 0  /* record the component-root pair (gmst_util.hpp:245, 312, 322), default */
#define GR_TREE_TRUE_MST 1   /* record each component's minimum pair: a true MST */

typedef struct gr_ctx gr_ctx;

typedef struct {
  int32_t X, Y, Z;      /* grid; Z >= 3; X, Y <= 8192 */
  int32_t cost_model;   /* GR_COST_* */
  int32_t commit_mode;  /* GR_COMMIT_* */
  int32_t device;       /* HIP device ordinal */
} gr_grid_desc;

/* Library version string. */
const char* gr_version(void);

/* Number of gfx950 devices visible (0 without a GPU); never fails. */
int gr_device_count(void);

/*
 * Create a context and upload the capacity cube (CSIZE = X*Y*Z int32 values,
 * WIRE_ENC order; demand starts at 0).  Builds the per-run Fenwick cost state.
 * Replaces: read_cap_file's cube upload (read_data.hpp:554-562) and the
 * cost-tree build BIT_UTIL_SPACE::init_bit_data (bit_util.hpp:192-200),
 * called from test_lpat.hpp:239-242.
 */
int gr_create(const gr_grid_desc* desc, const int32_t* cap, gr_ctx** out);

/*
 * Same for a grid with fp32 capacities (the .cap file's values,
 * read_data.hpp:136-143).  The remaining capacity r is kept per cell in fp32
 * and lowered by 1.0f when a batch commits the cell (bit_util.hpp:87); a cell
 * costs f(r) evaluated in double by a fixed Cody-Waite/Taylor sequence and
 * rounded to fp32 (equal to the integer table for integral r).  gr_get_demand
 * still returns the integer demand.  The learned proposal needs integer
 * capacities and refuses such a context.
 * Replaces: the float cube upload and init_bit_data as for gr_create.
 */
int gr_create_f32(const gr_grid_desc* desc, const float* cap, gr_ctx** out);

/* Remaining fp32 capacity per cell of a gr_create_f32 context (GR_ESTATE otherwise). */
int gr_get_remaining_f32(gr_ctx* ctx, float* out);

/* Free every device and host resource of the context. */
void gr_destroy(gr_ctx* ctx);

/* Message for the last failing call on this context (or a global one if ctx is NULL). */
const char* gr_last_error(const gr_ctx* ctx);

/*
 * Upload all nets: pin_pfx[n_nets+1] (int64 exclusive prefix of pin counts)
 * and pin_enc[pin_pfx[n_nets]] (per net ascending, deduplicated, >= 2 pins,
 * <= 32767 pins).  Replaces READ_DATA::add_net_gpu_data (read_data.hpp:486-524).
 */
int gr_set_nets(gr_ctx* ctx, int32_t n_nets, const int64_t* pin_pfx, const int32_t* pin_enc);

/*
 * Route one batch synchronously: L-pattern costing of every pin pair, root-pair
 * Boruvka tree per net, commit of the tree's routes into demand, refresh of
 * the cost state.  out_edges / out_lpat have sum(deg) slots, out_net_cost n
 * entries (sum of the chosen minimum edge costs, reference net_dist).  Any of
 * the outputs may be NULL.  Returns 0.
 * Replaces one iteration of the batch loop, test_lpat.hpp:280-450: pair build
 * (:293-346), BIT_UTIL_SPACE::get_gpu_lpat_dist (bit_util.hpp:118-191),
 * MST_UTIL::get_block_net_map (mst_util.hpp:152-320) and
 * MST_BIT_COMMIT::bit_commit_edges (mst_bit_commit.hpp:236-255), plus the
 * D2H hand-off to the writer (mst_util.hpp:316-319).
 */
int gr_route_batch(gr_ctx* ctx, const int32_t* net_ids, int32_t n,
                   int64_t* out_edges, int32_t* out_lpat, float* out_net_cost);

/*
 * Allocation-free variant of gr_route_batch.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, outputs and results as gr_route_batch, bit for bit, but it
 * calls no hipMalloc/hipFree: every device buffer the batch needs beyond the
 * context's grid state and nets (per-net descriptors, the pair-owner table,
 * per-pair minimum costs and patterns, output slots) lives in one static
 * device array of GR_STATIC_SCRATCH_BYTES per GPU, whose address the library
 * takes once per device when the first context is created there.  A batch
 * whose buffers do not fit (6 B per pin pair + 4 B per 64 pairs + 12 B per
 * pin + 36 B per net, each part rounded up to 256 B) gets GR_ERANGE before any device work.
 * Exhaustive costing only: GR_ESTATE while a learned proposal or sample
 * capture is set; nets of more than 4096 pins give GR_ERANGE (their trees
 * keep node arrays in device memory).  Per-kernel timing is not recorded for
 * these calls.
 * Every call on one GPU uses that GPU's one array, which keeps the last
 * call's contents (nothing in it needs clearing: a call writes every byte it
 * reads); calls that overlap on one GPU, from any context, race.
 */
#define GR_STATIC_SCRATCH_BYTES (64ll << 20)
int gr_route_batch_static(gr_ctx* ctx, const int32_t* net_ids, int32_t n,
                          int64_t* out_edges, int32_t* out_lpat, float* out_net_cost);
/* Size of that array in bytes (GR_STATIC_SCRATCH_BYTES). */
int64_t gr_static_scratch_bytes(void);

/*
 * Keep a whole batch schedule resident on the device: batch b covers
 * net_ids[batch_ofst[b] .. batch_ofst[b+1]).  Replaces the batcher's per-batch
 * uploads and scans (part_net/net_part.hpp:345-369).  Nets of up to 4096 pins
 * build their trees in one workgroup's LDS; larger ones (up to the 32767 the
 * 15-bit pin index allows, mst_common.h:25) keep their node arrays in device
 * memory the schedule owns (44 B per output slot of the batch).
 */
int gr_set_schedule(gr_ctx* ctx, int32_t n_batches, const int32_t* batch_ofst,
                    const int32_t* net_ids);

/*
 * Enqueue batches [first, first+count) of the resident schedule on the
 * context's stream without any host synchronisation.  Outputs stay on the
 * device until gr_get_schedule_outputs.  Same per-batch semantics as
 * gr_route_batch.  Replaces the router thread's while loop, test_lpat.hpp:280.
 */
int gr_route_schedule(gr_ctx* ctx, int32_t first, int32_t count);

/*
 * Tree mode for the following batches.  GR_TREE_ROOT_PAIR reproduces the
 * reference: a component that picks its minimum outgoing pair records the
 * pair of component roots, and that pair's L-pattern is committed.
 * GR_TREE_TRUE_MST records the minimum pair itself (same Boruvka rounds,
 * same lowest-index ties), so each net's tree is its minimum spanning tree
 * over the pair costs, the tree the reference's CPU Kruskal twin builds
 * (MST_UTIL::t_get_mst, mst_util.hpp:34-102).  net_cost is the same sum in
 * both modes; the committed routes differ.
 */
int gr_set_tree_mode(gr_ctx* ctx, int32_t mode);

/* Wait for all enqueued work; returns a HIP error as GR_ENODEV. */
int gr_sync(gr_ctx* ctx);

/*
 * Copy schedule outputs to the host: edges/lpat have sum over the schedule of
 * deg slots (schedule order), net_cost one entry per scheduled net.
 * Replaces the writer hand-off (mst_util.hpp:316-319).
 */
int gr_get_schedule_outputs(gr_ctx* ctx, int64_t* out_edges, int32_t* out_lpat, float* out_net_cost);

/* Total slots (sum of deg) and nets of the resident schedule. */
int64_t gr_schedule_slots(const gr_ctx* ctx);

/* Restore demand to zero and rebuild the cost state (start a new routing pass). */
int gr_reset_demand(gr_ctx* ctx);

/* Integer demand per cell (CSIZE int32), the parity state. */
int gr_get_demand(gr_ctx* ctx, int32_t* out);

/* Materialised Fenwick prefix Q per cell (CSIZE fp32), for parity tests; in
 * the non-BIT mode the exclusive run prefix P of the wire costs. */
int gr_get_prefix(gr_ctx* ctx, float* out);

/* Per-cell via cost term (CSIZE fp32): Q(pos) - Q(pos-1) of the Fenwick
 * prefix in the BIT modes; the via cost of the layer at the cell in the
 * non-BIT mode.  For parity tests. */
int gr_get_via_terms(gr_ctx* ctx, float* out);

/* Number of pin pairs in the last routed batch. */
int64_t gr_last_pairs(const gr_ctx* ctx);

/* Per-pair minimum cost and pattern of the last routed batch, batch pair order. */
int gr_get_pair_costs(gr_ctx* ctx, float* cost, int16_t* pat);

/* Cost LUT of the context: fills out[gr_lut_size()] for r = gr_lut_rmin() + i. */
int gr_get_lut(const gr_ctx* ctx, float* out);
int gr_lut_rmin(void);
int gr_lut_size(void);

/*
 * Schedule parameters.  They change speed, never results (every choice is
 * bit-identical).  Defaults: {1024, 64, 3, 2048, 0, 0}.
 *   large_threads   workgroup size for nets above group_max pins (256, 512 or 1024)
 *   group_max       largest pin count handled by a wavefront group of <= 64 lanes (32 or 64)
 *   fork            how a batch with small and large nets runs its trees:
 *                   0: two launches in sequence; 1: the large-net launch on a second
 *                   stream, parallel to the small-net one; 2: one launch for both when
 *                   two of its workgroups fit a CU's LDS, else as 1; 3: one launch
 *                   for both whenever the large tree fits in LDS; 4: one launch per
 *                   batch for costing, trees and commit (k_batch: per-XCD queues of
 *                   costing chunks and trees, each tree waiting on its own XCD's
 *                   chunks), on the resident schedule in the exhaustive BIT modes,
 *                   else as 3.  gr_sync reports a fused launch whose bounded wait
 *                   timed out (GR_ESTATE)
 *   refresh_blocks  persistent grid of the cost refresh (>= 1)
 *   lpat_wave       1: one wavefront per pin pair instead of one lane per pair
 *   lpat_block      workgroup size of the lane-per-pair costing (64, 128 or 256), or
 *                   0: per batch, the largest of these that still gives >= 2048
 *                   workgroups (64 below that)
 */
typedef struct {
  int32_t large_threads, group_max, fork, refresh_blocks, lpat_wave, lpat_block;
} gr_tuning;
int gr_set_tuning(gr_ctx* ctx, const gr_tuning* t);
int gr_get_tuning(const gr_ctx* ctx, gr_tuning* t);
/*
 * Autotuner: time every candidate tuning on the resident schedule (reps
 * passes each, demand reset before each pass; HIP events on the library
 * stream), keep the fastest and return it with its ms per pass.  Demand is
 * reset afterwards.  No counterpart in the reference.
 */
int gr_autotune(gr_ctx* ctx, int32_t reps, gr_tuning* best, double* best_ms);

/*
 * Kernel timing: when enabled, every launch of the listed kernels is bracketed
 * by HIP events on the launch stream.  gr_kernel_stats returns, per kernel
 * id, the launch count and total milliseconds since the last enable.
 * Kernel ids: 0 L-pattern costing, 1 tree+commit, 2 cost refresh.
 */
#define GR_K_LPAT 0
#define GR_K_TREE 1
#define GR_K_REFRESH 2
#define GR_K_COUNT 3
int gr_enable_timing(gr_ctx* ctx, int on);
int gr_kernel_stats(gr_ctx* ctx, int64_t* launches, double* total_ms);

/* ---- learned proposal (no counterpart in the reference) ----
 *
 * A 3-layer MLP (n_in -> hidden -> hidden -> n_out, ReLU) scores the L-pattern
 * candidates of a pin pair so that only its top-k are costed.  n_in = 4 + 25*Z
 * features per pair: dx/64, dy/64, z_src, z_dst, then for every layer the log
 * cell cost at a 5x5 lattice spanning the pair's bounding box.  n_out =
 * (Z-1)*Z2 candidates (patterns with first layer a >= 1).  Weights are bf16 on
 * MFMA with fp32 accumulation and an fp32 master copy; training runs on the
 * GPU with the library's own forward and backward kernels on samples captured
 * from exhaustive routing (the chosen pattern is the label), or online from
 * routing with the proposal, where only the pairs that fell back to every
 * pattern carry a label.
 */
typedef struct gr_mlp gr_mlp;

int gr_mlp_create(gr_ctx* ctx, int32_t hidden, uint64_t seed, int32_t max_samples, gr_mlp** out);
void gr_mlp_destroy(gr_mlp* m);
/* n_in, k_in (n_in padded to 32), hidden, n_out, n_out padded (64), parameter count */
int gr_mlp_shape(const gr_mlp* m, int32_t* shape6);
/* fp32 master parameters: W1 [k_in x hidden], b1 [hidden], W2 [hidden x hidden], b2,
 * W3 [hidden x n_out_pad], b3 [n_out_pad], row-major, concatenated. */
int gr_mlp_get_params(gr_mlp* m, float* out);
int gr_mlp_set_params(gr_mlp* m, const float* in);
/* While per_batch > 0, every routed batch appends up to per_batch samples
 * (features at the state before its commit, label = the chosen pattern).
 * Under a proposal (hidden width 256) a sample's label is kept only when its
 * pair fell back to every pattern (the exhaustive choice); the others are
 * stored with label -1 and training skips them. */
int gr_mlp_collect(gr_ctx* ctx, gr_mlp* m, int32_t per_batch);
int64_t gr_mlp_samples(const gr_mlp* m);
int gr_mlp_clear_samples(gr_mlp* m);
/* SGD with momentum 0.9 over the labelled samples; *loss = mean cross-entropy of the last epoch. */
int gr_mlp_train(gr_mlp* m, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* loss);
/* Features of explicit pairs (src/dst cell encodings) at the current state: n x k_in. */
int gr_mlp_features(gr_ctx* ctx, gr_mlp* m, const int32_t* src, const int32_t* dst, int32_t n, float* out);
/* Forward pass on host rows x (n x k_in, rounded to bf16 on upload) -> logits (n x n_out_pad). */
int gr_mlp_forward(gr_mlp* m, const float* x, int32_t n, float* logits);
/* Gradient of the mean cross-entropy over (x, y) w.r.t. all parameters; no update. */
int gr_mlp_gradient(gr_mlp* m, const float* x, const int32_t* y, int32_t n, float* grad, float* loss);
/* Route with the model's top-k candidates (1 <= k <= 8), falling back to full
 * enumeration when the top-1 softmax probability is below threshold.  m = NULL
 * restores exhaustive costing.  The chosen pattern is the lexicographic
 * minimum of (cost, p) over the costed candidates. */
int gr_set_proposal(gr_ctx* ctx, gr_mlp* m, int32_t k, float threshold);
/* Pairs costed and pairs that fell back since the last call (counters reset). */
int gr_proposal_stats(gr_ctx* ctx, int64_t* pairs, int64_t* fallbacks);

/* ---- route evaluation ---- */

typedef struct {
  int64_t wire;         /* sum over the layer's cells of the wire count */
  int64_t via;          /* sum of the via count */
  int32_t max_wire;     /* largest wire count of a cell (report starts at 0) */
  int32_t max_via;      /* largest via count (the reference log prints it halved) */
  float max_cap;        /* largest capacity (starts at 0) */
  double max_overflow;  /* largest (2 wire + via - 2 cap) / 2, starting at 0 */
  double of_cost;       /* sum of e^((2 wire + via - 2 cap) / 4), exponent rounded to fp32 */
  int64_t demand;       /* sum of the committed demand (cells the router used) */
  double overflow;      /* sum of max(0, demand - cap) */
} gr_eval_layer;

typedef struct {
  int64_t wire, via;    /* layers 1..Z-1, as the report sums them */
  double of_cost;       /* layers 1..Z-1 */
  int64_t demand;       /* all layers */
  double overflow;      /* all layers */
} gr_eval_total;

/*
 * Evaluate the routes held in the resident schedule's outputs (after
 * gr_route_schedule over the schedule) on the device.  Wire and via counts
 * per cell follow the reference's non-BIT commit counters
 * (mst_files/mst_commit.hpp:17-67 via charges, :68-113 and :329-341 wire
 * cells [min, max) of the run, once per batch) along each edge's written
 * route; the per-layer figures are those of the end-of-run report
 * (main.cpp:103-135).  layers gets Z entries (may be NULL).
 * Replaces the EVAL_OFCOST counters and report (main.cpp:103-135,
 * read_data.hpp:162-168).
 */
int gr_evaluate(gr_ctx* ctx, gr_eval_layer* layers, gr_eval_total* total);

/* Per-cell wire and via counts (CSIZE int32 each, either may be NULL) of the last gr_evaluate. */
int gr_eval_counts(gr_ctx* ctx, int32_t* wire, int32_t* via);

/* ---- host side of the path (no device work) ---- */

/*
 * Synthetic workload generator (seeded, deterministic; the distributions of
 * each family are fixed in gr_host.cpp).
 * kind: 1..5 selects the workload family.  First call with NULL buffers to
 * learn sizes (n_nets, n_pins); then call again with buffers.
 * cap gets X*Y*Z values; pin_pfx n_nets+1; pin_enc n_pins.
 */
int gr_gen_grid(int32_t kind, uint64_t seed, int32_t X, int32_t Y, int32_t Z, int32_t* cap);
int gr_gen_nets(int32_t kind, uint64_t seed, int32_t X, int32_t Y, int32_t Z, int32_t n_nets_req,
                int32_t* n_nets, int64_t* n_pins, int64_t* pin_pfx, int32_t* pin_enc);

/*
 * Batch generation: greedy packing of nets with pairwise disjoint closed
 * bounding boxes into batches, num_xdim_parts x-interval buckets per batch.
 * Replaces NET_PART::part_nets / part_xdim (part_net/net_part.hpp:167-379).
 * Outputs batch_ofst[n_batches+1] and net_ids[n_nets] (each batch sorted by
 * (pin count, net id)).  Returns the number of batches.
 */
int gr_partition_nets(int32_t X, int32_t Y, int32_t Z, int32_t n_nets, const int64_t* pin_pfx,
                      const int32_t* pin_enc, int32_t num_xdim_parts,
                      int32_t* batch_ofst, int32_t* net_ids);

/* ---- design database, ISPD files and route writer (host only) ---- */

typedef struct gr_design gr_design;

typedef struct {
  int32_t X, Y, Z;
  int32_t n_nets;      /* nets with >= 2 distinct pins: the routed nets, file order */
  int64_t n_pins;      /* their pins */
  int32_t n_single;    /* nets with one distinct pin (written, never routed) */
  float unit_length_cost, unit_via_cost;
} gr_design_desc;

/* .net dialects (read_data.hpp:540-546): one access point per pin
 * "[(z, x, y)]" (read_ispd24_net_single, :314-333, default); the first of
 * several access points (read_ispd24_net, :284-313); the -new_fmt dialect
 * (read_nt, :262-283, whose .cap header also carries four more numbers). */
#define GR_FMT_FIRST_ACCESS_POINT 1
#define GR_FMT_NEW 2

/*
 * Read an ISPD .cap / .net pair.  Capacities are parsed with the reference's
 * fp32 digit accumulation (bit-identical values), pins of a net are
 * deduplicated and sorted by cell, nets with one pin are set aside, the rest
 * keep file order and are the router's net ids 0..n_nets-1.  threads <= 0
 * picks min(hardware threads, 16).
 * Replaces READ_DATA::read_cap_file and read_net_file (read_data.hpp:94-195,
 * 196-345, 366-432, 534-565).
 */
int gr_design_read(const char* cap_path, const char* net_path, int32_t flags, int32_t threads, gr_design** out);

/* A design from arrays (fp32 capacities, may be NULL for zeros; nets named
 * "net<i>"); pins are deduplicated and one-pin nets set aside as when reading. */
int gr_design_create(int32_t X, int32_t Y, int32_t Z, const float* cap, int32_t n_nets, const int64_t* pin_pfx,
                     const int32_t* pin_enc, gr_design** out);
void gr_design_free(gr_design* d);
int gr_design_info(const gr_design* d, gr_design_desc* out);
/* X*Y*Z fp32 capacities, WIRE_ENC order (read_layer_data, read_data.hpp:136-143). */
int gr_design_capacity(const gr_design* d, float* cap);
/* Per layer: direction flag, minimum length, unit short cost (any may be NULL). */
int gr_design_layers(const gr_design* d, int32_t* dir, int32_t* min_len, float* unit_short_cost);
/* Layer name into buf (NUL-terminated, truncated to size-1); returns its length. */
int gr_design_layer_name(const gr_design* d, int32_t z, char* buf, int32_t size);
/* GCell lengths: hlen X-1 values, vlen Y-1 values (read_data.hpp:116-131). */
int gr_design_gcell_lengths(const gr_design* d, int64_t* hlen, int64_t* vlen);
/* Routed nets as gr_set_nets takes them: pin_pfx n_nets+1, pin_enc n_pins. */
int gr_design_nets(const gr_design* d, int64_t* pin_pfx, int32_t* pin_enc);
/* The pin cell of each one-pin net (n_single values). */
int gr_design_single_nets(const gr_design* d, int32_t* pin_enc);
/* Name of routed net i (0..n_nets-1) or one-pin net n_nets+j, with '\n' and
 * ' ' removed as the writer prints it; returns its length. */
int gr_design_net_name(const gr_design* d, int32_t net, char* buf, int32_t size);
/* Write the design back as .cap / .net text (either path may be NULL); the
 * .net side is write_ispd24_nets' format (read_data.hpp:349-359). */
int gr_design_save(const gr_design* d, const char* cap_path, const char* net_path);

/* route text formats (writer_util.hpp; the reference's -rseg_fmt, -wrt_mrg, -oprd_ppatch) */
#define GR_WRITE_DBU 1    /* "4200g+2100 ... metalN" lines instead of grid "z x y z x y" */
#define GR_WRITE_MERGE 2  /* per net: fuse touching wires, distinct via cells */
#define GR_WRITE_PATCH 4  /* with MERGE: split wires at pins that are no segment end */

/*
 * Write the routes of a schedule: for every batch in order and every net of
 * the batch in order, the net's tree edges (gr_get_schedule_outputs layout:
 * deg slots per net, n-1 edges and a -1 sentinel) expanded to segments along
 * each edge's L-pattern; one-pin nets (one via each) follow the first batch.
 * Each net is "name\n(\n" + one line per segment + ")\n".  Output is
 * deterministic; threads <= 0 picks min(hardware threads, 16).  *bytes (may
 * be NULL) gets the file size.
 * Replaces ROUTE_WRITER::write_route (writer_files/writer.hpp:12-51) and
 * ROUTE_WRITER_UTIL::write_route / mrg_route_write / write_odeg_route
 * (writer_files/writer_util.hpp:101-165, 290-364).
 */
int gr_write_routes(const gr_design* d, const char* path, int32_t flags, int32_t n_batches, const int32_t* batch_ofst,
                    const int32_t* net_ids, const int64_t* edges, const int32_t* lpat, int32_t threads,
                    int64_t* bytes);

#ifdef __cplusplus
}
#endif

#endif /* IITMGR_H */
