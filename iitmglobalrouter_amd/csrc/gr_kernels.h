// Kernels of the routing hot path (gfx950, wave64).
//
// Device state (WIRE_ENC cell order, util.hpp:38-44; every layer is a set of
// contiguous runs: rows of length X on even layers, columns of length Y on odd):
//   rem[cell]        int32  remaining capacity cap - demand (the integer state);
//                           with fp32 capacities: minus the demand, and
//   remf[cell]       fp32   remaining capacity r (fp32 grids only)
//   pfx[cell]        fp32   Q(pos) of the run's Fenwick tree, materialised per
//                           cell, so a reference range query Q(hi) - Q(lo-1)
//                           (bit_util.hpp:36-38) costs two loads
//                           (four floats of padding precede the first cell,
//                           so a run of a multiple of 4 cells starts 16-byte
//                           aligned when its offset is a multiple of 4)
//   mark[cell]       uint8  1 from the commit that used the cell until the
//                           same batch's refresh consumes it
//   run_stamp[z*M+l] int32  batch stamp of the last commit that touched run l of layer z
// Per batch:
//   k_lpat_lane  one lane per pin pair evaluates every L-pattern serially
//   k_tree_small/_large/_mixed  root-pair Boruvka + commit: one lane per
//                2-pin net, one wavefront per net up to 64 pins, one workgroup
//                above; _mixed runs both kinds of one batch in one launch
//   k_refresh    per dirty run: demand -1 on marked cells, canonical Fenwick
//                nodes rebuilt in LDS, run's pfx rewritten
#pragma once
#include "gr_device.h"

namespace grd {

// Phase markers of the workgroup tree and of the fused batch launch's items
// (no-ops unless a profiling build defines them).
#ifndef GR_TREE_MARK
#define GR_TREE_MARK(id)
#define GR_TREE_MARK_INIT
#endif
#ifndef GR_ITEM_MARK
#define GR_ITEM_MARK(phase, kind, xcd)
#endif

// ---------------------------------------------------------------------------
// Cost refresh
// ---------------------------------------------------------------------------

// Runs are scanned by a persistent grid; run id r = z*M + line, every run or
// the batch's candidate list rl[0, nrl) (runs its commit can stamp).  all != 0
// rebuilds every run without touching demand (set-up and reset).  Otherwise a
// run is processed only when the batch's commit stamped it: its marked cells
// lose one unit of capacity (bit_util.hpp:75-90 applies cap -= 1 once per
// flagged cell) and their marks are cleared, then Fenwick node i (1-based,
// lowbit 2^k, covering cells [i - 2^k, i)) is the pairwise fp32 sum of its
// aligned block, and Q(x) accumulates nodes in the reference's descent order
// (bit_util.hpp:19-35).
//
// Layout: the run's tree lives in LDS in place, node i at F[i - 1] (each
// index holds its highest node, as the reference's array does).  Each thread
// holds 4 consecutive cells and forms the nodes of 2 and 4 cells in
// registers; a wavefront covers 256 cells and forms the nodes of 8..256 cells
// by lane shuffles (the lane whose quad ends the block adds the left half
// from the lane below), so only nodes of 512 cells and more need LDS
// barriers.  Every node is (left child) + (right child) in fp32, as in a
// level-by-level build, and a node exists only when its whole block lies in
// the run.  The loads of two 1024-cell chunks are issued together.
//
// Descent per quad: for x = 4t + q, Q(x) adds the nodes on the path of
// i = x + 1 from its lowest set bit up (i -= lowbit(i)): the node of set bit
// b of i is F[((i >> b) << b) - 1].  For q = 0..2 the path is the quad's own
// nodes (F[4t] for q = 0, F[4t+1] for q = 1, F[4t+2] then F[4t+1] for q = 2)
// followed by the path of 4t; for q = 3 it is the path of 4t + 4.  So a quad
// reads the path of 4t once for three of its cells; an unset bit adds +0.0,
// an exact identity on the non-negative sums.  A quad is written with one
// 16-byte store when the run is aligned: values below the lowest marked cell
// are rewritten with the bits they hold (their nodes cover unchanged cells).
//
// FLT: grids with fp32 capacities.  remf holds r per cell; a marked cell's r
// is lowered by 1.0f in fp32 (bit_util.hpp:87, d_cube_data_bit[tid]--) and the
// integer rem counts the commits; the cell cost is cost_f32(model, r).
//
// remz (integer grids, or null): the per-GCELL copy of every layer's clamped
// remaining capacity that the learned proposal's lattice gather reads; a
// marked cell's entry is rewritten with its new value.  VT: vterm holds the
// per-cell via terms the proposal's candidate costing reads, rewritten with
// the prefix (LDS then holds the run's prefix too).
//
// LDS: the cost table, then F and R, each M rounded up to a multiple of 4
// floats (refresh_lds).
__host__ __device__ inline int refresh_stride(int M) { return (M + 3) & ~3; }

template <bool FLT, bool VT = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FLT ? 6 : 8))) void k_refresh(Geo g, int M, int32_t* __restrict__ rem,
                                                 uint8_t* __restrict__ mark, float* __restrict__ pfx,
                                                 const float* __restrict__ lut,
                                                 const int32_t* __restrict__ run_stamp,
                                                 const int32_t* __restrict__ ctr, int all,
                                                 float* __restrict__ remf, int model,
                                                 int16_t* __restrict__ remz, float* __restrict__ vterm,
                                                 const int32_t* __restrict__ rl, int nrl,
                                                 int32_t* __restrict__ zero8) {
  extern __shared__ float sh[];
  // the claim counters of the batch's fused launch (null: none)
  if (zero8 && blockIdx.x == 0 && threadIdx.x < 8) zero8[threadIdx.x] = 0;
  // lowest marked position of the run, in two slots that alternate between
  // the runs a workgroup processes: a slot is cleared one run before its use
  __shared__ int first_hit[2];
  const int32_t bstamp = all ? 0 : *ctr;
  float* slut = sh;
  float* F = sh + kLut;
  float* R = F + refresh_stride(M);  // the run's prefix, for stored via terms (vterm non-null)
  if (threadIdx.x < 2) first_hit[threadIdx.x] = 0x7fffffff;
  // the cost table goes to LDS at the first run the workgroup processes,
  // after that run's first cell loads are issued (a workgroup whose runs
  // are all unchanged loads nothing)
  bool ready = false;
  int slot = 0;
  const int lane = threadIdx.x & 63;
  const int nr = rl ? nrl : g.Z * M;
  for (int ri = blockIdx.x; ri < nr; ri += gridDim.x) {
    const int r = rl ? rl[ri] : ri;
    const int z = r / M, line = r - z * M;
    const int lines = (z & 1) ? g.X : g.Y;
    if (line >= lines) continue;
    if (!all && run_stamp[r] != bstamp) continue;
    const int L = (z & 1) ? g.Y : g.X;
    const uint32_t base = (uint32_t)z * g.XY + (uint32_t)line * L;
    if (threadIdx.x == 0) first_hit[slot ^ 1] = 0x7fffffff;  // read last by the previous run
    int my_first = all ? 0 : 0x7fffffff;
    // a run that starts at a multiple of 4 cells is read in quads (one
    // 16-byte load of the state, one 4-byte load of the marks per thread)
    const bool quads = (base & 3u) == 0;
    for (int c0 = 0; c0 < L; c0 += 1024) {
      const int t4 = c0 + 4 * (int)threadIdx.x;
      float v[4];
      bool hits[4];
      int rv[4];
      float rfv[4];
      if (quads && t4 + 3 < L) {
        const uchar4 mk = *reinterpret_cast<const uchar4*>(mark + base + t4);
        hits[0] = !all && mk.x;
        hits[1] = !all && mk.y;
        hits[2] = !all && mk.z;
        hits[3] = !all && mk.w;
        if constexpr (FLT) {
          const float4 t = *reinterpret_cast<const float4*>(remf + base + t4);
          rfv[0] = t.x; rfv[1] = t.y; rfv[2] = t.z; rfv[3] = t.w;
        } else {
          const int4 t = *reinterpret_cast<const int4*>(rem + base + t4);
          rv[0] = t.x; rv[1] = t.y; rv[2] = t.z; rv[3] = t.w;
        }
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = t4 + q;
          hits[q] = i < L && !all && mark[base + i];
          if constexpr (FLT) rfv[q] = i < L ? remf[base + i] : 0.0f;
          else rv[q] = i < L ? rem[base + i] : 0;
        }
      }
      if (!ready) {
        if (!FLT)
          for (int i = threadIdx.x; i < kLut; i += blockDim.x) slut[i] = lut[i];
        __syncthreads();
        ready = true;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = t4 + q;
        float c = 0.0f;
        if (i < L) {
          const bool hit = hits[q];
          if constexpr (FLT) {
            float rf = rfv[q];
            if (hit) {
              rf -= 1.0f;
              remf[base + i] = rf;
              rem[base + i] -= 1;
            }
            c = cost_f32(model, rf);
          } else {
            int vv = rv[q];
            if (hit) {
              vv -= 1;
              rem[base + i] = vv;
            }
            vv = vv < kRMin ? kRMin : (vv > kRMin + kLut - 1 ? kRMin + kLut - 1 : vv);
            c = slut[vv - kRMin];
            if (hit && remz) remz[((z & 1) ? (uint32_t)i * g.X + line : (uint32_t)line * g.X + i) * 16 + z] = (int16_t)vv;
          }
          if (hit) {
            mark[base + i] = 0;
            my_first = min(my_first, i);
          }
        }
        v[q] = c;
      }
      // F[t4] = node(t4+1) = cell; F[t4+1] = node(t4+2) = 2 cells; F[t4+2] =
      // cell; F[t4+3] = node(t4+4), 4 cells or more: built up below
      const float p0 = v[0] + v[1], p1 = v[2] + v[3];
      float q = p0 + p1;
      float top = q;
#pragma unroll
      for (int k = 3; k <= 8; ++k) {
        const int s2 = 1 << (k - 3);  // lanes per half block
        const float left = __shfl_up(q, s2, 64);
        q = left + q;
        // this lane ends a block of 2^k cells: node(t4 + 4) gets it
        if (((lane + 1) & ((2 << (k - 3)) - 1)) == 0) top = q;
      }
      if (t4 < L) F[t4] = v[0];
      if (t4 + 1 < L) F[t4 + 1] = p0;
      if (t4 + 2 < L) F[t4 + 2] = v[2];
      if (t4 + 3 < L) F[t4 + 3] = top;
    }
    my_first = min(my_first, __shfl_xor(my_first, 1, 64));
    my_first = min(my_first, __shfl_xor(my_first, 2, 64));
    my_first = min(my_first, __shfl_xor(my_first, 4, 64));
    my_first = min(my_first, __shfl_xor(my_first, 8, 64));
    my_first = min(my_first, __shfl_xor(my_first, 16, 64));
    my_first = min(my_first, __shfl_xor(my_first, 32, 64));
    if (lane == 0 && my_first < 0x7fffffff) atomicMin(&first_hit[slot], my_first);
    __syncthreads();
    // nodes of 512 cells and more: F[i-1] = F[i - 2^(k-1) - 1] + F[i-1],
    // in place, one level per barrier
    for (int k = 9; (1 << k) <= L; ++k) {
      const int h = 1 << (k - 1);
      for (int i = (threadIdx.x + 1) << k; i <= L; i += (int)blockDim.x << k) F[i - 1] = F[i - 1 - h] + F[i - 1];
      __syncthreads();
    }
    // Q(x) changes only from the lowest marked cell on: positions below it
    // keep their stored values
    const int f = first_hit[slot];
    slot ^= 1;
    // with stored via terms, Q(f - 1) is formed again for the first of them
    const int f0 = (VT && f > 0) ? f - 1 : f;
    for (int t4 = 4 * (int)threadIdx.x; t4 < L; t4 += 4 * (int)blockDim.x) {
      if (t4 + 3 < f0) continue;
      const float4 o = *reinterpret_cast<const float4*>(F + t4);  // entries past L are never used
      // runs are at most 8192 cells, so bits 2..13 cover every path
      const uint32_t u = (uint32_t)t4, w = (uint32_t)t4 + 4u;
      float hv[12], gv[12];
#pragma unroll
      for (int b = 2; b < 14; ++b) {
        hv[b - 2] = ((u >> b) & 1u) ? F[((u >> b) << b) - 1] : 0.0f;
        gv[b - 2] = ((w >> b) & 1u) ? F[((w >> b) << b) - 1] : 0.0f;
      }
      float q0 = o.x, q1 = o.y, q2 = o.z + o.y, q3 = 0.0f;
#pragma unroll
      for (int b = 0; b < 12; ++b) {
        q0 += hv[b];
        q1 += hv[b];
        q2 += hv[b];
        q3 += gv[b];
      }
      if (quads && t4 + 3 < L) {
        *reinterpret_cast<float4*>(pfx + base + t4) = make_float4(q0, q1, q2, q3);
      } else {
        const float qv[4] = {q0, q1, q2, q3};
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (t4 + q >= f0 && t4 + q < L) pfx[base + t4 + q] = qv[q];
      }
      if (VT) {
        const float qv[4] = {q0, q1, q2, q3};
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (t4 + q < L) R[t4 + q] = qv[q];
      }
    }
    if constexpr (VT) {
      // Q(pos) - Q(pos - 1), Q(-1) = 0 (bit_util.hpp:108-117): the fp32
      // subtraction the costing kernels make from the prefix
      __syncthreads();
      for (int x = f + threadIdx.x; x < L; x += blockDim.x) vterm[base + x] = R[x] - (x ? R[x - 1] : 0.0f);
    }
    __syncthreads();  // F (and R) are rewritten by the next run
  }
}

// ---------------------------------------------------------------------------
// L-pattern costing
// ---------------------------------------------------------------------------

// Square key matrices of the nets whose trees stream their pair keys from
// global memory: the pair (j, k) key at key[ofs[pos] + j*n4 + k] and at
// [ofs[pos] + k*n4 + j], n4 = n rounded up to a multiple of 4, so that every
// pin's keys form one contiguous 16-byte aligned row (ofs[pos] < 0 or key
// null: the net has none).  The diagonal and the padding are never written.
__host__ __device__ inline int64_t square_stride(int64_t n) { return (n + 3) & ~(int64_t)3; }
struct SquareOut {
  const int64_t* ofs;
  const int32_t* deg;
  uint32_t* key;
};
__device__ __forceinline__ void square_store(const SquareOut& sq, int pos, int j, int k, uint32_t key) {
  if (!sq.key) return;
  const int64_t so = sq.ofs[pos];
  if (so < 0) return;
  const int64_t n4 = square_stride(sq.deg[pos]);
  sq.key[so + j * n4 + k] = key;
  sq.key[so + k * n4 + j] = key;
}

// One lane per pin pair of the batch (workgroups of 64, 128 or 256 lanes);
// BUF: prefix pairs through a buffer resource (prefixes of fewer than 2^30 cells).
template <int ZT, bool BUF = true>
__global__ __launch_bounds__(256) void k_lpat_lane(Geo g, const float* __restrict__ pfx,
                                                   const int32_t* __restrict__ pins,
                                                   const int64_t* __restrict__ pinofs,
                                                   const int64_t* __restrict__ bpair,
                                                   const int32_t* __restrict__ chunk_net, int n, int64_t P,
                                                   float* __restrict__ out_cost, int16_t* __restrict__ out_pat,
                                                   int32_t* __restrict__ ctr, SquareOut sq) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *ctr += 1;  // this batch's stamp, read by k_tree*/k_refresh
  const int64_t w0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t w = w0 < P ? w0 : P - 1;  // tail lanes shadow the last pair until after the shuffles
  const PairOwner own = pair_owner_wave(bpair, pinofs, chunk_net, n, w);
  if (w0 >= P) return;
  int j, k;
  pair_decode((uint32_t)(w - own.first_pair), j, k);
  const int64_t pb = own.first_pin;
  int z1, x1, y1, z2, x2, y2;
  dec(g, (uint32_t)pins[pb + j], z1, x1, y1);
  dec(g, (uint32_t)pins[pb + k], z2, x2, y2);
  const uint64_t best = lane_best_pts<ZT, BUF>(g, pfx, z1, x1, y1, z2, x2, y2);
  out_cost[w] = key_cost((uint32_t)(best >> 32));
  out_pat[w] = (int16_t)(uint32_t)best;
  square_store(sq, own.pos, j, k, (uint32_t)(best >> 32));
}

// Endpoint table of one batch: ends[w] = (pin j, pin k) of pair w, packed;
// the nets and pins do not change between passes, so this is built once per
// schedule and the costing kernel reads one coalesced 8-byte record per pair
// instead of finding the pair's net and loading and decoding its two pins.
__global__ __launch_bounds__(256) void k_pair_ends(Geo g, const int32_t* __restrict__ pins,
                                                   const int64_t* __restrict__ pinofs,
                                                   const int64_t* __restrict__ bpair,
                                                   const int32_t* __restrict__ chunk_net, int n, int64_t P,
                                                   uint2* __restrict__ ends) {
  const int64_t w0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t w = w0 < P ? w0 : P - 1;
  const PairOwner own = pair_owner_wave(bpair, pinofs, chunk_net, n, w);
  if (w0 >= P) return;
  int j, k;
  pair_decode((uint32_t)(w - own.first_pair), j, k);
  int z1, x1, y1, z2, x2, y2;
  dec(g, (uint32_t)pins[own.first_pin + j], z1, x1, y1);
  dec(g, (uint32_t)pins[own.first_pin + k], z2, x2, y2);
  ends[w] = make_uint2(pack_cell(z1, x1, y1), pack_cell(z2, x2, y2));
}

// k_lpat_lane with the pair's endpoints from the schedule's endpoint table.
// The pair's net is looked up only when the batch writes square key
// matrices (sq.key set: the same for every lane of the launch).
template <int ZT, bool BUF = true>
__global__ __launch_bounds__(256) void k_lpat_ends(Geo g, const float* __restrict__ pfx,
                                                   const uint2* __restrict__ ends,
                                                   const int64_t* __restrict__ pinofs,
                                                   const int64_t* __restrict__ bpair,
                                                   const int32_t* __restrict__ chunk_net, int n, int64_t P,
                                                   float* __restrict__ out_cost, int16_t* __restrict__ out_pat,
                                                   int32_t* __restrict__ ctr, SquareOut sq) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *ctr += 1;  // this batch's stamp, read by k_tree*/k_refresh
  const int64_t w0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t w = w0 < P ? w0 : P - 1;  // tail lanes shadow the last pair until after the shuffles
  const uint2 e = ends[w];
  PairOwner own{0, 0, 0};
  if (sq.key) own = pair_owner_wave(bpair, pinofs, chunk_net, n, w);
  if (w0 >= P) return;
  int z1, x1, y1, z2, x2, y2;
  unpack_cell(e.x, z1, x1, y1);
  unpack_cell(e.y, z2, x2, y2);
  const uint64_t best = lane_best_pts<ZT, BUF>(g, pfx, z1, x1, y1, z2, x2, y2);
  out_cost[w] = key_cost((uint32_t)(best >> 32));
  out_pat[w] = (int16_t)(uint32_t)best;
  if (sq.key) {
    int j, k;
    pair_decode((uint32_t)(w - own.first_pair), j, k);
    square_store(sq, own.pos, j, k, (uint32_t)(best >> 32));
  }
}

// Any Z (up to 32): one wavefront per pair, one lane per pattern.
__global__ __launch_bounds__(256) void k_lpat_wave(Geo g, const float* __restrict__ pfx,
                                                   const int32_t* __restrict__ pins,
                                                   const int64_t* __restrict__ pinofs,
                                                   const int64_t* __restrict__ bpair,
                                                   const int32_t* __restrict__ chunk_net, int n, int64_t P,
                                                   float* __restrict__ out_cost, int16_t* __restrict__ out_pat,
                                                   int32_t* __restrict__ ctr, SquareOut sq) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *ctr += 1;  // this batch's stamp, read by k_tree*/k_refresh
  const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= P) return;
  int pos = chunk_net[w >> 6];
  while (pos + 1 < n && bpair[pos + 1] <= w) ++pos;
  int j, k;
  pair_decode((uint32_t)(w - bpair[pos]), j, k);
  const int64_t pb = pinofs[pos];
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
    square_store(sq, pos, j, k, (uint32_t)(best >> 32));
  }
}

// ---------------------------------------------------------------------------
// Tree + commit
// ---------------------------------------------------------------------------

struct TreeArgs {
  Geo g;
  const int64_t* pinofs;  // batch position -> first pin of the net in pins[]
  const int32_t* deg;     // batch position -> pin count
  const int64_t* slot;  // batch position -> first output slot
  const int64_t* pair;  // batch position -> first pair of the net in the batch
  const int32_t* list;  // batch positions handled by this launch
  int n_list;
  int64_t cost_base;    // out_cost index of batch position 0
  const int32_t* pins;
  const float* pair_cost;
  const int16_t* pair_pat;
  int64_t* out_edges;
  int32_t* out_lpat;
  float* out_cost;
  uint8_t* mark;
  int32_t* run_stamp;
  const int32_t* ctr;  // device batch counter; its value is the batch stamp
  int32_t quirk;
  int32_t M;  // run_stamp row stride: max(X, Y)
  int32_t true_mst;  // record each accepted component's minimum pair instead of the root pair
  int32_t prefix;    // non-BIT commit: half-open wire marks and via half charges
  int32_t* vhalf;    // non-BIT commit: via half charges per cell of this batch
  const int64_t* sqofs;  // batch position -> square key matrix (-1: none; null: no matrices)
  const uint32_t* sq;
  const unsigned long long* rowmin;  // per output slot - rowmin_base: first-round pin minima (null: none)
  int64_t rowmin_base;
};

// The edge an accepted candidate of component root r records: the root pair
// (r, partner root) as the reference does (gmst_util.hpp:245, 312, 322), or
// in true-MST mode the minimum pair itself, its endpoint in r's component
// first.  root[] must still hold the round's starting roots.
__device__ __forceinline__ int2 recorded_edge(const TreeArgs& A, const int* root, int r, int pr, uint32_t o) {
  if (!A.true_mst) return make_int2(r, pr);
  int j, k;
  pair_decode(o, j, k);
  return root[j] == r ? make_int2(j, k) : make_int2(k, j);
}

// Commit marks are idempotent plain stores: the cell's mark byte and its
// run's stamp.  The demand update itself happens once per marked cell in
// k_refresh, which also clears the mark.
__device__ __forceinline__ void mark_run(const TreeArgs& A, int32_t bs, int z, int line, int pos) {
  (void)pos;
  A.run_stamp[z * A.M + line] = bs;
}
__device__ __forceinline__ void mark_any(const TreeArgs& A, int32_t bs, uint32_t cell) {
  A.mark[cell] = 1;
  int z, x, y;
  dec(A.g, cell, z, x, y);
  mark_run(A, bs, z, (z & 1) ? x : y, (z & 1) ? y : x);
}
__device__ __forceinline__ void mark_wire(const TreeArgs& A, int32_t bs, int z, int line, int lo, int hi) {
  const uint32_t L = (z & 1) ? A.g.Y : A.g.X;
  const uint32_t base = (uint32_t)z * A.g.XY + (uint32_t)line * L;
  for (int t = lo; t <= hi; ++t) A.mark[base + t] = 1;
  mark_run(A, bs, z, line, lo);
}
__device__ __forceinline__ void mark_via(const TreeArgs& A, int32_t bs, int za, int zb, int x, int y) {
  const int lo = za < zb ? za : zb, hi = za < zb ? zb : za;
  for (int l = lo + 1; l < hi; ++l) {
    A.mark[enc(A.g, l, x, y)] = 1;
    mark_run(A, bs, l, (l & 1) ? x : y, (l & 1) ? y : x);
  }
}

// Cells of one committed edge (mst_bit_commit.hpp:36-67, 147-235): interior
// via cells, inclusive wires.  With the quirk, the k threads that marked the
// first via stack also mark the k cells just before the first wire's low end,
// by enc arithmetic (:213-216: the first wire call is not guarded by tid >= 0).
// A quirk cell whose enc would be negative (possible only when X*Y < Z) lies
// outside the grid: the reference writes before its array there; it is
// skipped here.
// Non-BIT commit of one edge (MST_COMMIT, mst_commit.hpp:114-251): each
// wire marks the cells [min, max) of its run (the difference array's +1 at
// the low end and -1 at the high end, gpu_commit_wire :68-113), the first on
// a, the second on b' whenever b' > 0 (gpu_add_wires :188-251); via stacks
// z1 -> a at the source, a -> b' after the first wire when b' > 0, then to
// z2, each over layers [min, max) (gpu_add_vias :114-186), charge half units
// (gpu_commit_via :17-61): 2 to the cell at run position 0, 2 to the
// previous cell at the last position, else 1 to each.  The refresh applies
// them: -1.0 once per wired cell, then -0.5 per half.
__device__ __forceinline__ void prefix_wire(const TreeArgs& A, int32_t bs, int z, int line, int lo, int hi) {
  if (lo >= hi) return;
  const uint32_t L = (z & 1) ? A.g.Y : A.g.X;
  const uint32_t base = (uint32_t)z * A.g.XY + (uint32_t)line * L;
  for (int t = lo; t < hi; ++t) A.mark[base + t] = 1;
  A.run_stamp[z * A.M + line] = bs;
}
__device__ __forceinline__ void prefix_via(const TreeArgs& A, int32_t bs, int za, int zb, int x, int y) {
  const int lo = za < zb ? za : zb, hi = za < zb ? zb : za;
  for (int i = lo; i < hi; ++i) {
    const int pos = (i & 1) ? y : x, L = (i & 1) ? A.g.Y : A.g.X;
    const uint32_t c = enc(A.g, i, x, y);
    if (pos == 0) atomicAdd(&A.vhalf[c], 2);
    else if (pos == L - 1) atomicAdd(&A.vhalf[c - 1], 2);
    else { atomicAdd(&A.vhalf[c - 1], 1); atomicAdd(&A.vhalf[c], 1); }
    A.run_stamp[i * A.M + ((i & 1) ? x : y)] = bs;
  }
}
__device__ __forceinline__ void commit_edge_prefix(const TreeArgs& A, int32_t bs, uint32_t src, uint32_t dst, int p) {
  const Geo& g = A.g;
  int z1, x1, y1, z2, x2, y2;
  dec(g, src, z1, x1, y1);
  dec(g, dst, z2, x2, y2);
  const int a = p / g.Z2, b = p - a * g.Z2;
  const int bp = (a & 1) ? 2 * b : 2 * b + 1;
  // wires
  int x = x1, y = y1;
  if (a & 1) { prefix_wire(A, bs, a, x, min(y, y2), max(y, y2)); y = y2; }
  else       { prefix_wire(A, bs, a, y, min(x, x2), max(x, x2)); x = x2; }
  if (bp > 0) {
    if (bp & 1) prefix_wire(A, bs, bp, x, min(y, y2), max(y, y2));
    else        prefix_wire(A, bs, bp, y, min(x, x2), max(x, x2));
  }
  // vias
  x = x1; y = y1;
  prefix_via(A, bs, z1, a, x, y);
  int z = a;
  if (z & 1) y = y2; else x = x2;
  if (bp > 0) {
    prefix_via(A, bs, z, bp, x, y);
    z = bp;
    if (z & 1) y = y2; else x = x2;
  }
  if (z != z2) prefix_via(A, bs, z, z2, x, y);
}

__device__ __forceinline__ void commit_edge(const TreeArgs& A, int32_t bs, uint32_t src, uint32_t dst, int p) {
  const Geo& g = A.g;
  if (p < g.Z2 || p >= g.Z * g.Z2) return;  // no valid pattern (cannot happen for Z >= 3)
  if (A.prefix) {
    commit_edge_prefix(A, bs, src, dst, p);
    return;
  }
  int z1, x1, y1, z2, x2, y2;
  dec(g, src, z1, x1, y1);
  dec(g, dst, z2, x2, y2);
  const int a = p / g.Z2, b = p - a * g.Z2;
  const bool D = (x1 != x2) && (y1 != y2);
  int z = z1, x = x1, y = y1, k = 0;
  if (z != a) {
    mark_via(A, bs, z, a, x, y);
    k = (z < a ? a - z : z - a) - 1;
    z = a;
  }
  if (z & 1) {
    const int lo = min(y, y2), hi = max(y, y2);
    mark_wire(A, bs, z, x, lo, hi);
    if (A.quirk) {
      const uint32_t e0 = enc(g, z, x, lo);
      for (int t = 1; t <= k && (uint32_t)t <= e0; ++t) mark_any(A, bs, e0 - t);
    }
    y = y2;
  } else {
    const int lo = min(x, x2), hi = max(x, x2);
    mark_wire(A, bs, z, y, lo, hi);
    if (A.quirk) {
      const uint32_t e0 = enc(g, z, lo, y);
      for (int t = 1; t <= k && (uint32_t)t <= e0; ++t) mark_any(A, bs, e0 - t);
    }
    x = x2;
  }
  const int bp = (z & 1) ? 2 * b : 2 * b + 1;
  if (D && bp != 0) {
    mark_via(A, bs, z, bp, x, y);
    z = bp;
    if (z & 1) { mark_wire(A, bs, z, x, min(y, y2), max(y, y2)); y = y2; }
    else       { mark_wire(A, bs, z, y, min(x, x2), max(x, x2)); x = x2; }
  }
  if (z != z2) mark_via(A, bs, z, z2, x, y);
}

// One of the five parts of commit_edge's cell sequence for a BIT-mode edge
// (the same cells, so a net's edges can be marked part by part on separate
// threads; marks and stamps are idempotent stores):
//   0 the via stack z1 -> a at the source, with the quirk cells before the
//     first wire's low end;  1 the first wire on layer a;  2 the via stack
//   a -> b' at the corner;  3 the second wire on b';  4 the via stack to z2.
__device__ __forceinline__ void commit_part(const TreeArgs& A, int32_t bs, int part, int z1, int x1, int y1, int z2,
                                            int x2, int y2, int p) {
  const Geo& g = A.g;
  if (p < g.Z2 || p >= g.Z * g.Z2) return;
  const int a = p / g.Z2, b = p - a * g.Z2;
  const bool D = (x1 != x2) && (y1 != y2);
  if (part == 0) {
    if (z1 == a) return;
    mark_via(A, bs, z1, a, x1, y1);
    if (A.quirk) {
      const int k = (z1 < a ? a - z1 : z1 - a) - 1;
      const uint32_t e0 = (a & 1) ? enc(g, a, x1, min(y1, y2)) : enc(g, a, min(x1, x2), y1);
      for (int t = 1; t <= k && (uint32_t)t <= e0; ++t) mark_any(A, bs, e0 - t);
    }
    return;
  }
  if (part == 1) {
    if (a & 1) mark_wire(A, bs, a, x1, min(y1, y2), max(y1, y2));
    else       mark_wire(A, bs, a, y1, min(x1, x2), max(x1, x2));
    return;
  }
  // after the first wire
  const int x = (a & 1) ? x1 : x2, y = (a & 1) ? y2 : y1;
  const int bp = (a & 1) ? 2 * b : 2 * b + 1;
  const bool two = D && bp != 0;
  if (part == 2) {
    if (two) mark_via(A, bs, a, bp, x, y);
    return;
  }
  if (part == 3) {
    if (!two) return;
    if (bp & 1) mark_wire(A, bs, bp, x, min(y, y2), max(y, y2));
    else        mark_wire(A, bs, bp, y, min(x, x2), max(x, x2));
    return;
  }
  int z = a, xf = x, yf = y;
  if (two) {
    z = bp;
    if (bp & 1) yf = y2; else xf = x2;
  }
  if (z != z2) mark_via(A, bs, z, z2, xf, yf);
}

// Net cost: the n-1 accepted minimum costs added in slot order (the
// reference's serial net_dist += cost), the LDS loads issued in groups of 8
// ahead of the dependent adds.
__device__ __forceinline__ float sum_edge_costs(const float* ecost, int cnt) {
  float nc = 0.0f;
  int e = 0;
  for (; e + 8 <= cnt; e += 8) {
    float v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = ecost[e + q];
#pragma unroll
    for (int q = 0; q < 8; ++q) nc += v[q];
  }
  for (; e < cnt; ++e) nc += ecost[e];
  return nc;
}

__device__ __forceinline__ void emit_edge(const TreeArgs& A, int32_t bs, int pos, int64_t sb, int e, int on1, int on2,
                                          const int16_t* pat, const int32_t* npins) {
  A.out_edges[sb + e] = ((int64_t)pos << 30) | ((int64_t)on1 << 15) | (int64_t)on2;
  const int hi = on1 > on2 ? on1 : on2, lo = on1 > on2 ? on2 : on1;
  const int p = pat[(uint32_t)((int64_t)hi * (hi - 1) / 2 + lo)];
  A.out_lpat[sb + e] = p;
  commit_edge(A, bs, (uint32_t)npins[hi], (uint32_t)npins[lo], p);
}

// 2-pin nets: the only pair is (1, 0); Boruvka's single round
// has the mutual candidates 0 -> 1 and 1 -> 0, and root 0 is processed first,
// so the edge is (0, 1) and the tree cost is 0 + cost (in true-MST mode the
// minimum pair's endpoints are the roots themselves: the same edge).
__device__ __forceinline__ void tree2_commit(const TreeArgs& A, int32_t bs, int pos) {
  const int64_t pb = A.pinofs[pos];
  const int64_t sb = A.slot[pos];
  const int16_t* pat = A.pair_pat + A.pair[pos];
  emit_edge(A, bs, pos, sb, 0, 0, 1, pat, A.pins + pb);
  A.out_edges[sb + 1] = -1;
  A.out_lpat[sb + 1] = -1;
  A.out_cost[A.cost_base + pos] = 0.0f + A.pair_cost[A.pair[pos]];
}

// Root of x in a round's merge forest with path halving: each visited
// non-root is re-pointed to its grandparent, an ancestor, so concurrent finds
// stay valid and the chains shrink as they run.
__device__ __forceinline__ int dsu_find(volatile int* vp, int x) {
  while (true) {
    const int p = vp[x];
    if (p == x) return x;
    const int pp = vp[p];
    if (pp != p) vp[x] = pp;
    x = pp;
  }
}

// Root-pair Boruvka (gmst_util.hpp:164-341) with the lowest-index tie contract
// for one net on one workgroup (node r handled by thread r mod gsize):
//   min phase    per component root, min (cost, o) over outgoing pairs
//   partner      root of the other endpoint of that pair
//   accept       the reference processes candidates serially in ascending
//                root order and drops one whose roots already joined.  With
//                distinct (cost, o) keys the candidates form a forest plus
//                duplicated mutual pairs, so exactly the larger root of each
//                mutual pair is dropped: the serial outcome, computed in
//                parallel.  Accepted edges (root, partner) are written in
//                ascending root order.
//   merge        the accepted edges form a forest (each root has at most one,
//                to its partner root): one plain pointer store per root, then
//                each tree's largest old root, i.e. its largest member, is
//                the new root of all its members (a find with path halving
//                and an atomic maximum), which is the root that hooking the
//                smaller root under the larger gives (gmst_util.hpp:314-325).
// Rounds repeat until deg-1 edges (the reference's feedback loop,
// mst_util.hpp:298-314, ends in the same state).
//
// The min phase reads pair keys while many components remain: staged in LDS
// as ord_key values when they fit (scap), else streamed from global memory.
// Once the C current components fit a triangular matrix in LDS (mcap 64-bit
// entries, with room for the next round's matrix of at most C/2
// components), the keys are read one last time to build M[a][b] = the
// minimum (key << 32 | o) over the pairs joining components a and b, and
// every later round works on that matrix: component a's minimum outgoing
// pair is the minimum of its row and column, and after the round's merges
// M' [a'][b'] is the minimum of the entries of the old components that
// formed a' and b'.  A minimum over a union of pair sets is the minimum of
// the minima, so each round selects exactly the pairs the key scan selects;
// the matrix holds C(C-1)/2 entries instead of the n(n-1)/2 keys, and C at
// least halves every round.
//
// LDS: node arrays root, partner (also the compact index of each root, cid),
// par_a, par_b, ecost, best and edges (n each), croot_a / croot_b (the root of
// each component), the matrix (mcap), the staged keys (scap).
__device__ __forceinline__ uint32_t tri_index(int a, int b) {  // a > b
  return (uint32_t)a * (uint32_t)(a - 1) / 2 + (uint32_t)b;
}
struct TreeLds {
  int* root;
  unsigned long long* best;
  int* partner;
  int* par_a;
  int* par_b;
  int2* edges;
  float* ecost;
  int* cid;
  int* croot_a;
  int* croot_b;
  unsigned long long* mat;
  int mcap;
  int* wcnt;
  int* sh;
  uint32_t* skey;
  int scap;
};

// Workgroup barrier of the tree; with the node arrays in global memory (GN)
// it is also an agent-scope release and acquire, so stores are in the L2
// before the barrier and the L1 holds no stale line after it (the atomics of
// the min phase and the merge act in the L2).
template <bool GN>
__device__ __forceinline__ void tree_sync() {
  if (GN) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (GN) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Compact indices of the current roots in ascending node order: cid[r] for
// every root r, croot[c] = the root of component c; returns the count.
template <bool GN>
__device__ __forceinline__ int tree_ids(const TreeLds& L, int n, int tid, int gsize, int* croot) {
  const int lane = tid & 63, wv = tid >> 6, nwv = (gsize + 63) >> 6;
  const int iters = (n + gsize - 1) / gsize;
  for (int it = 0; it < iters; ++it) {
    const int i = it * gsize + tid;
    const bool isr = i < n && L.root[i] == i;
    const unsigned long long m = __ballot(isr);
    if (lane == 0) L.wcnt[it * nwv + wv] = __popcll(m);
  }
  tree_sync<GN>();
  for (int it = 0; it < iters; ++it) {
    const int i = it * gsize + tid;
    const bool isr = i < n && L.root[i] == i;
    const unsigned long long m = __ballot(isr);
    if (isr) {
      int off = __popcll(m & ((1ull << lane) - 1ull));
      for (int q = 0; q < it * nwv + wv; ++q) off += L.wcnt[q];
      L.cid[i] = off;
      croot[off] = i;
    }
  }
  int total = 0;
  for (int q = 0; q < iters * nwv; ++q) total += L.wcnt[q];
  tree_sync<GN>();
  return total;
}

// Minimum (key << 32 | pair index) over row i of a square key matrix (n
// pins, rows of square_stride(n) keys) among the columns k with other(k),
// for one wavefront: 16 bytes per lane and four loads in flight per chunk of
// 1024 keys, a per-lane minimum, then a wavefront minimum (~0: none).  A lane
// visits increasing k, and the pair index grows with k along a row (pairs
// (i, k < i) come before pairs (k > i, i)), so a strict-less key compare
// keeps the lowest index among equal keys.
template <class Other>
__device__ __forceinline__ unsigned long long square_row_min(const uint32_t* rowp, int n, int i, int lane, Other other) {
  const int n4 = (int)square_stride(n);
  const uint4* row = reinterpret_cast<const uint4*>(rowp);
  uint32_t mk = ~0u;
  int mkk = -1;
  for (int c0 = 0; c0 < n4; c0 += 1024) {
    uint4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k4 = c0 + 4 * (64 * u + lane);
      v[u] = k4 < n4 ? row[k4 >> 2] : make_uint4(~0u, ~0u, ~0u, ~0u);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k4 = c0 + 4 * (64 * u + lane);
      const uint32_t kv[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k = k4 + q;
        const bool take = (k < n) && (kv[q] < mk) && other(k < n ? k : 0);
        mk = take ? kv[q] : mk;
        mkk = take ? k : mkk;
      }
    }
  }
  unsigned long long m = ~0ull;
  if (mkk >= 0) m = ((unsigned long long)mk << 32) | (mkk < i ? tri_index(i, mkk) : tri_index(mkk, i));
  return wave_min_u64(m);
}

// First Boruvka round of the nets with square key matrices, over the whole
// GPU: no component is merged yet, so pin i's candidate is its minimum over
// the whole row (k != i), exactly as the workgroup tree's round computes it.
// One wavefront per pin, kRowTaskPins pins of one net per workgroup.
__global__ __launch_bounds__(64 * kRowTaskPins) void k_row_min(const int2* __restrict__ task,
                                                               const int32_t* __restrict__ deg,
                                                               const int64_t* __restrict__ slot,
                                                               const int64_t* __restrict__ sqofs,
                                                               const uint32_t* __restrict__ sq,
                                                               unsigned long long* __restrict__ rowmin,
                                                               int64_t slot0) {
  const int2 t = task[blockIdx.x];
  const int pos = t.x, i = t.y + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int n = deg[pos];
  if (i >= n) return;  // the whole wavefront
  const unsigned long long m = square_row_min(sq + sqofs[pos] + (size_t)i * (size_t)square_stride(n), n, i, lane,
                                              [&](int k) { return k != i; });
  if (lane == 0) rowmin[slot[pos] - slot0 + i] = m;
}

template <bool MAT, bool GN = false>
__device__ __forceinline__ void tree_commit(const TreeArgs& A, int32_t bs, int pos, int tid, int gsize, const TreeLds& L) {
  GR_TREE_MARK_INIT
  GR_TREE_MARK(0);
  int* root = L.root;
  unsigned long long* best = L.best;
  int* partner = L.partner;
  int2* edges = L.edges;
  float* ecost = L.ecost;
  int* wcnt = L.wcnt;
  int* sh = L.sh;
  const int64_t pb = A.pinofs[pos];
  const int n = A.deg[pos];
  const float* gcost = A.pair_cost + A.pair[pos];
  const int16_t* pat = A.pair_pat + A.pair[pos];
  const uint32_t P = (uint32_t)((int64_t)n * (n - 1) / 2);
  const int lane = tid & 63, wv = tid >> 6, nwv = (gsize + 63) >> 6;
  // every key round reads all P costs: keep them in LDS when they fit, as the
  // orderable keys the minimum compares (ord_key), converted once here
  const float* cost = gcost;
  const bool staged = P <= (uint32_t)L.scap;
  // keys that stream from global memory: the net's square matrix when the
  // costing kernel wrote one (every pin's keys in one contiguous row)
  const uint32_t* sqk = (MAT && !staged && A.sq && A.sqofs[pos] >= 0) ? A.sq + A.sqofs[pos] : nullptr;
  const int n4 = (int)square_stride(n);
  uint32_t* skey = L.skey;
  if (MAT && staged) {
    // launches with matrix rounds are those whose keys stream: nothing to stage
    for (uint32_t o = tid; o < P; o += gsize) skey[o] = ord_key(gcost[o]);
  } else if (staged) {
    // 16-byte loads from the first 16-byte boundary on, four per thread in
    // flight before any is converted; the head and tail singly
    const uint32_t mis = (uint32_t)(reinterpret_cast<uintptr_t>(gcost) >> 2) & 3u;
    const uint32_t head = min(P, (4u - mis) & 3u);
    const uint32_t n4 = (P - head) >> 2;
    const float4* g4 = reinterpret_cast<const float4*>(gcost + head);
    for (uint32_t o = tid; o < head; o += gsize) skey[o] = ord_key(gcost[o]);
    for (uint32_t q0 = tid; q0 < n4; q0 += 4u * gsize) {
      float4 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t q = q0 + (uint32_t)u * gsize;
        if (q < n4) v[u] = g4[q];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t q = q0 + (uint32_t)u * gsize;
        if (q < n4) {
          uint32_t* d = skey + head + 4u * q;
          d[0] = ord_key(v[u].x);
          d[1] = ord_key(v[u].y);
          d[2] = ord_key(v[u].z);
          d[3] = ord_key(v[u].w);
        }
      }
    }
    for (uint32_t o = head + 4u * n4 + tid; o < P; o += gsize) skey[o] = ord_key(gcost[o]);
  }
  for (int i = tid; i < n; i += gsize) {
    root[i] = i;
    best[i] = ~0ull;
    L.par_a[i] = i;
  }
  if (tid == 0) sh[0] = 0;
  tree_sync<GN>();
  GR_TREE_MARK(1);
  // matrix rounds: C components, matrix in buffer mb (the other buffer gets
  // the next round's), croot_cur / croot_nxt the components' roots
  bool matrix = false;
  int C = n;
  unsigned long long* mcur = L.mat;
  unsigned long long* mnxt = L.mat;
  int* croot_cur = L.croot_a;
  int* croot_nxt = L.croot_b;
  // Each round adds at least one edge (the globally smallest key is a mutual
  // pick), so n rounds always suffice; the bound only guards termination.
  // The DSU parents alternate between two arrays: the flatten of one round
  // reads this round's while it prepares the next round's (and resets best,
  // which nothing reads after the accept), saving a barrier per round.
  for (int round = 0; round < n; ++round) {
    const int ne = sh[0];
    if (ne >= n - 1) break;
    int* par = (round & 1) ? L.par_b : L.par_a;
    int* par_next = (round & 1) ? L.par_a : L.par_b;
    if (MAT && matrix) {
      // component a: minimum over its row (b < a) and column (b > a); S
      // adjacent threads per component, the result written by the first
      const int S = (C * 8 <= gsize) ? 8 : (C * 4 <= gsize) ? 4 : (C * 2 <= gsize) ? 2 : 1;
      const int per = gsize / S;
      for (int b0 = 0; b0 < C; b0 += per) {
        const int a = b0 + tid / S, part = tid & (S - 1);
        unsigned long long m = ~0ull;
        if (a < C) {
          const uint32_t rb = tri_index(a, 0);
          for (int b = part; b < a; b += S) {
            const unsigned long long v = mcur[rb + b];
            m = v < m ? v : m;
          }
          for (int b = a + 1 + ((part - (a + 1)) & (S - 1)); b < C; b += S) {
            const unsigned long long v = mcur[tri_index(b, a)];
            m = v < m ? v : m;
          }
        }
        for (int sh2 = 1; sh2 < S; sh2 <<= 1) {
          const unsigned long long t = shfl_xor_u64(m, sh2);
          m = t < m ? t : m;
        }
        if (a < C && part == 0) best[croot_cur[a]] = m;
      }
    } else if (sqk && round == 0 && A.rowmin) {
      // the first round's pin minima came from k_row_min (every root is its
      // own pin, so each minimum is its root's)
      const unsigned long long* rm = A.rowmin + (A.slot[pos] - A.rowmin_base);
      for (int i = tid; i < n; i += gsize) best[i] = rm[i];
    } else if (sqk) {
      // node-centric over the square matrix: one wavefront per pin scans the
      // pin's row (square_row_min), then one atomic per pin
      for (int i = wv; i < n; i += nwv) {
        const int ri = root[i];
        const unsigned long long m =
            square_row_min(sqk + (size_t)i * (size_t)n4, n, i, lane, [&](int k) { return root[k] != ri; });
        if (lane == 0 && m != ~0ull) atomicMin(&best[ri], m);
      }
    } else if (!staged) {
      // costs did not fit in LDS: each thread streams a contiguous chunk of
      // pairs from global memory.  The row root's running minimum and that of
      // the current run of equal column roots stay in registers and go to LDS
      // (atomicMin) only when the root changes.
      const uint32_t chunk = (P + gsize - 1) / gsize;
      uint32_t o = (uint32_t)tid * chunk;
      const uint32_t oe = o + chunk < P ? o + chunk : P;
      if (o < oe) {
        int j, k;
        pair_decode(o, j, k);
        int cj = root[j], ck = -1;
        unsigned long long mj = ~0ull, mk = ~0ull;
        for (; o < oe; ++o) {
          const int rk = root[k];
          if (rk != cj) {
            const unsigned long long key = ((unsigned long long)ord_key(cost[o]) << 32) | o;
            mj = key < mj ? key : mj;
            if (rk != ck) {
              if (mk != ~0ull) atomicMin(&best[ck], mk);
              ck = rk;
              mk = key;
            } else {
              mk = key < mk ? key : mk;
            }
          }
          if (++k == j) {
            if (mj != ~0ull) atomicMin(&best[cj], mj);
            mj = ~0ull;
            ++j;
            k = 0;
            if (j < n) cj = root[j];
          }
        }
        if (mj != ~0ull) atomicMin(&best[cj], mj);
        if (mk != ~0ull) atomicMin(&best[ck], mk);
      }
    } else {
      // node-centric: S adjacent threads scan node i's n-1 pairs (rows and
      // columns of the triangle) into a register min; one atomic per node.
      // A thread visits its pairs in increasing index (its row part, then
      // its column part), so a strict-less 32-bit key compare keeps the
      // lowest index among equal keys: the (key, index) minimum.  Column
      // indices advance by additions only.
      const int S = (n * 8 <= gsize) ? 8 : (n * 4 <= gsize) ? 4 : (n * 2 <= gsize) ? 2 : 1;
      const int lgS = S == 8 ? 3 : S == 4 ? 2 : S == 2 ? 1 : 0;
      const int per = gsize / S;
      for (int b0 = 0; b0 < n; b0 += per) {
        const int i = b0 + tid / S, part = tid & (S - 1);
        uint32_t mk = ~0u, mo = ~0u;
        int ri = -1;
        if (i < n) {
          ri = root[i];
          // row part: pairs (i, k), k < i, index i(i-1)/2 + k
          uint32_t o = (uint32_t)i * (uint32_t)(i - 1) / 2 + (uint32_t)part;
#pragma unroll 4
          for (int k = part; k < i; k += S, o += S) {
            const uint32_t key = skey[o];
            const bool take = (root[k] != ri) & (key < mk);
            mk = take ? key : mk;
            mo = take ? o : mo;
          }
          // column part: pairs (k, i), k > i, index k(k-1)/2 + i
          int k = i + 1 + ((part - (i + 1)) & (S - 1));
          o = (uint32_t)k * (uint32_t)(k - 1) / 2 + (uint32_t)i;
          const uint32_t cS = (uint32_t)(S * (S - 1) / 2);
#pragma unroll 4
          for (; k < n; k += S) {
            const uint32_t key = skey[o];
            const bool take = (root[k] != ri) & (key < mk);
            mk = take ? key : mk;
            mo = take ? o : mo;
            o += ((uint32_t)k << lgS) + cS;
          }
        }
        unsigned long long m = mk == ~0u ? ~0ull : (((unsigned long long)mk << 32) | mo);
        for (int sh2 = 1; sh2 < S; sh2 <<= 1) {
          const unsigned long long t = shfl_xor_u64(m, sh2);
          m = t < m ? t : m;
        }
        if (i < n && part == 0 && m != ~0ull) atomicMin(&best[ri], m);
      }
    }
    tree_sync<GN>();
    GR_TREE_MARK(2);
    for (int r = tid; r < n; r += gsize) {
      int pr = -1;
      if (root[r] == r && best[r] != ~0ull) {
        int j, k;
        pair_decode((uint32_t)best[r], j, k);
        pr = (root[j] == r) ? root[k] : root[j];
      }
      partner[r] = pr;
      par[r] = r;       // this round's forest pointers
      par_next[r] = r;  // the largest old root of each tree, below
    }
    tree_sync<GN>();
    GR_TREE_MARK(3);
    // accept, ordered slots, union
    const int iters = (n + gsize - 1) / gsize;
    for (int it = 0; it < iters; ++it) {
      const int r = it * gsize + tid;
      bool acc = false;
      int pr = -1;
      if (r < n) {
        pr = partner[r];
        acc = pr >= 0 && !(partner[pr] == r && r > pr);
      }
      const unsigned long long m = __ballot(acc);
      if (lane == 0) wcnt[it * nwv + wv] = __popcll(m);
      tree_sync<GN>();
      if (acc) {
        int off = ne;
        for (int q = 0; q < it * nwv + wv; ++q) off += wcnt[q];
        off += __popcll(m & ((1ull << lane) - 1ull));
        edges[off] = recorded_edge(A, root, r, pr, (uint32_t)best[r]);
        ecost[off] = key_cost((uint32_t)(best[r] >> 32));  // pair costs are never -0.0
        par[r] = pr;
      }
    }
    tree_sync<GN>();
    GR_TREE_MARK(4);
    int tot = ne;
    for (int q = 0; q < iters * nwv; ++q) tot += wcnt[q];
    if (tid == 0) sh[0] = tot;
    {
      // the accepted edges (root -> partner root) form a forest, one
      // pointer per root, written without compare-and-swap; every old root
      // finds its tree's root (path halving), and the tree's largest old root,
      // i.e. its largest member, becomes the component's root, as hooking
      // the smaller root under the larger makes it
      for (int r = tid; r < n; r += gsize)
        if (root[r] == r) atomicMax(&par_next[dsu_find(par, r)], r);
      tree_sync<GN>();
      for (int i = tid; i < n; i += gsize) {
        root[i] = par_next[dsu_find(par, root[i])];
        best[i] = ~0ull;
      }
    }
    tree_sync<GN>();
    GR_TREE_MARK(5);
    if (!MAT) continue;
    const int Cn = n - tot;  // components after this round's unions
    if (Cn <= 1) continue;
    const uint32_t Tn = tri_index(Cn, 0), Th = tri_index(Cn / 2, 0);
    if (matrix) {
      // M' from M: the minimum over the old components forming each new pair
      const int C2 = tree_ids<GN>(L, n, tid, gsize, croot_nxt);
      (void)C2;
      unsigned long long* mout = (mcur == L.mat) ? L.mat + L.mcap - Tn : L.mat;
      for (uint32_t e = tid; e < Tn; e += gsize) mout[e] = ~0ull;
      tree_sync<GN>();
      for (int a = 1 + wv; a < C; a += nwv) {
        const int na = L.cid[root[croot_cur[a]]];
        const uint32_t rb = tri_index(a, 0);
        for (int b = lane; b < a; b += 64) {
          const int nb = L.cid[root[croot_cur[b]]];
          if (na != nb) {
            const unsigned long long v = mcur[rb + b];
            atomicMin(&mout[na > nb ? tri_index(na, nb) : tri_index(nb, na)], v);
          }
        }
      }
      tree_sync<GN>();
      GR_TREE_MARK(8);
      mcur = mout;
      int* t = croot_cur;
      croot_cur = croot_nxt;
      croot_nxt = t;
      C = Cn;
    } else if ((uint64_t)Tn + Th <= (uint64_t)L.mcap) {
      // switch to matrix rounds: one last pass over the keys builds M
      C = tree_ids<GN>(L, n, tid, gsize, croot_cur);
      mcur = L.mat;
      for (uint32_t e = tid; e < Tn; e += gsize) mcur[e] = ~0ull;
      // every node's component index, read once per pair by the build
      // (par_a is free until the next round's partner phase rewrites it)
      int* comp = L.par_a;
      if (sqk)
        for (int k = tid; k < n; k += gsize) comp[k] = L.cid[root[k]];
      tree_sync<GN>();
      const uint32_t chunk = (P + gsize - 1) / gsize;
      uint32_t o = (uint32_t)tid * chunk;
      const uint32_t oe = o + chunk < P ? o + chunk : P;
      if (sqk) {
        // the lower half of the square, row by row: 16-byte loads (4 keys
        // per lane, 256 columns per wavefront step, two steps in flight);
        // a lane's consecutive keys into the same component are combined
        // before the atomic (the minimum over a union is the minimum of the
        // minima, and within a run the pair index grows with k)
        for (int j = 1 + wv; j < n; j += nwv) {
          const int cj = comp[j];
          const uint4* row = reinterpret_cast<const uint4*>(sqk + (size_t)j * (size_t)n4);
          const uint32_t ob = tri_index(j, 0);
          for (int k0 = 0; k0 < j; k0 += 512) {
            uint4 v[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
              const int k4 = k0 + 256 * u + 4 * lane;
              v[u] = k4 < j ? row[k4 >> 2] : make_uint4(0u, 0u, 0u, 0u);
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
              const int k4 = k0 + 256 * u + 4 * lane;
              const uint32_t kv[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
              int pc = -1;
              unsigned long long pm = ~0ull;
#pragma unroll
              for (int q = 0; q < 4; ++q) {
                const int k = k4 + q;
                const int ck = k < j ? comp[k] : cj;
                if (ck != pc) {
                  if (pc >= 0 && pc != cj) atomicMin(&mcur[cj > pc ? tri_index(cj, pc) : tri_index(pc, cj)], pm);
                  pc = ck;
                  pm = ~0ull;
                }
                const unsigned long long key = ((unsigned long long)kv[q] << 32) | (ob + (uint32_t)k);
                pm = key < pm ? key : pm;
              }
              if (pc >= 0 && pc != cj) atomicMin(&mcur[cj > pc ? tri_index(cj, pc) : tri_index(pc, cj)], pm);
            }
          }
        }
      } else if (o < oe) {
        int j, k;
        pair_decode(o, j, k);
        int cj = L.cid[root[j]];
        for (; o < oe; ++o) {
          const int ck = L.cid[root[k]];
          if (ck != cj) {
            const uint32_t kv = staged ? skey[o] : ord_key(cost[o]);
            atomicMin(&mcur[cj > ck ? tri_index(cj, ck) : tri_index(ck, cj)], ((unsigned long long)kv << 32) | o);
          }
          if (++k == j) {
            ++j;
            k = 0;
            if (j < n) cj = L.cid[root[j]];
          }
        }
      }
      tree_sync<GN>();
      GR_TREE_MARK(7);
      matrix = true;
    }
  }
  const int64_t sb = A.slot[pos];
  const int32_t* npins = A.pins + pb;
  if (A.prefix) {
    for (int e = tid; e < n - 1; e += gsize) emit_edge(A, bs, pos, sb, e, edges[e].x, edges[e].y, pat, npins);
  } else {
    // outputs and each edge's endpoints and pattern (in best and partner,
    // free after the rounds), then the commit marks as (part, edge) tasks
    // over every thread: a wire no longer waits for the vias and the other
    // wire of its edge on the same thread
    for (int e = tid; e < n - 1; e += gsize) {
      const int on1 = edges[e].x, on2 = edges[e].y;
      A.out_edges[sb + e] = ((int64_t)pos << 30) | ((int64_t)on1 << 15) | (int64_t)on2;
      const int hi = on1 > on2 ? on1 : on2, lo = on1 > on2 ? on2 : on1;
      const int p = pat[(uint32_t)((int64_t)hi * (hi - 1) / 2 + lo)];
      A.out_lpat[sb + e] = p;
      int z1, x1, y1, z2, x2, y2;
      dec(A.g, (uint32_t)npins[hi], z1, x1, y1);
      dec(A.g, (uint32_t)npins[lo], z2, x2, y2);
      best[e] = (unsigned long long)(uint32_t)x1 | ((unsigned long long)(uint32_t)y1 << 16) |
                ((unsigned long long)(uint32_t)x2 << 32) | ((unsigned long long)(uint32_t)y2 << 48);
      partner[e] = z1 | (z2 << 8) | (p << 16);
    }
    tree_sync<GN>();
    const int ne = n - 1;
    for (int t = tid; t < 5 * ne; t += gsize) {
      const int part = t / ne, e = t - part * ne;
      const unsigned long long xy = best[e];
      const int zz = partner[e];
      commit_part(A, bs, part, zz & 0xff, (int)(xy & 0xffff), (int)((xy >> 16) & 0xffff), (zz >> 8) & 0xff,
                  (int)((xy >> 32) & 0xffff), (int)(xy >> 48), zz >> 16);
    }
  }
  // the net cost on a thread without an edge when there is one
  if (tid == (n - 1) % gsize) {
    A.out_edges[sb + n - 1] = -1;
    A.out_lpat[sb + n - 1] = -1;
    A.out_cost[A.cost_base + pos] = sum_edge_costs(ecost, n - 1);
  }
  GR_TREE_MARK(6);
}

// The same root-pair Boruvka for nets of at most G pins (G in {8,16,32,64}):
// a group of G lanes per net, node i on lane i, 64/G nets per wavefront.
// Groups of one wavefront advance in lockstep (wave-wide syncs); a group whose
// net is complete idles until every group of the wave is.  Phases and their
// justification are those of tree_commit.
struct GroupLds {
  int root[64];
  unsigned long long best[64];
  int partner[64];
  int link[64];
  int2 edges[64];
  float ecost[64];
  uint32_t key[kSmallDeg * (kSmallDeg - 1) / 2];  // per group: G(G-1)/2 pair keys (ord_key of the cost)
};

template <int G>
__device__ __forceinline__ void tree_group(const TreeArgs& A, int32_t bs, const int32_t* list, int n_list, int idx,
                                           GroupLds& L) {
  const int lane = threadIdx.x & 63;
  const int gb = lane & ~(G - 1), tid = lane - gb;
  const unsigned long long gbits = (G == 64) ? ~0ull : (((1ull << (G & 63)) - 1ull) << gb);
  const bool valid = idx < n_list;
  const int pos = valid ? list[idx] : 0;
  const int n = valid ? A.deg[pos] : 0;
  const int64_t pb = valid ? A.pinofs[pos] : 0;
  const uint32_t P = (uint32_t)(n * (n - 1) / 2);
  int* root = L.root + gb;
  unsigned long long* best = L.best + gb;
  int* partner = L.partner + gb;
  int* link = L.link + gb;
  int2* edges = L.edges + gb;
  float* ecost = L.ecost + gb;
  uint32_t* key = L.key + (gb / G) * (G * (G - 1) / 2);
  if (valid) {
    const float* gcost = A.pair_cost + A.pair[pos];
    for (uint32_t o0 = tid; o0 < P; o0 += 4 * G) {  // four loads in flight per lane
      float v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t o = o0 + (uint32_t)(u * G);
        v[u] = o < P ? gcost[o] : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t o = o0 + (uint32_t)(u * G);
        if (o < P) key[o] = ord_key(v[u]);
      }
    }
  }
  if (tid < n) root[tid] = tid;
  int ne = 0;
  wave_sync();
  for (int round = 0; round < G; ++round) {
    const bool act = ne < n - 1;
    if (__ballot(act) == 0ull) break;
    const bool me = act && tid < n;
    if (me) {
      best[tid] = ~0ull;
      link[tid] = tid;  // DSU parent for this round's unions
    }
    wave_sync();
    if (me) {
      // pairs of node tid in increasing index (row, then column), strict-less
      // on the 32-bit key: the (key, index) minimum, as in tree_commit
      const int ri = root[tid];
      uint32_t mk = ~0u, mo = ~0u;
      uint32_t o = (uint32_t)tid * (uint32_t)(tid - 1) / 2;
#pragma unroll 4
      for (int k = 0; k < tid; ++k, ++o) {
        const uint32_t kv = key[o];
        const bool take = (root[k] != ri) & (kv < mk);
        mk = take ? kv : mk;
        mo = take ? o : mo;
      }
      o = (uint32_t)(tid + 1) * (uint32_t)tid / 2 + (uint32_t)tid;
#pragma unroll 4
      for (int k = tid + 1; k < n; ++k) {
        const uint32_t kv = key[o];
        const bool take = (root[k] != ri) & (kv < mk);
        mk = take ? kv : mk;
        mo = take ? o : mo;
        o += (uint32_t)k;
      }
      if (mk != ~0u) atomicMin(&best[ri], ((unsigned long long)mk << 32) | mo);
    }
    wave_sync();
    int pr = -1;
    if (me && root[tid] == tid && best[tid] != ~0ull) {
      int j, k;
      pair_decode((uint32_t)best[tid], j, k);
      pr = (root[j] == tid) ? root[k] : root[j];
    }
    if (me) partner[tid] = pr;
    wave_sync();
    const bool acc = me && pr >= 0 && !(partner[pr] == tid && tid > pr);
    const unsigned long long gm = __ballot(acc) & gbits;
    if (acc) {
      const int off = ne + __popcll(gm & ((1ull << lane) - 1ull));
      edges[off] = recorded_edge(A, root, tid, pr, (uint32_t)best[tid]);
      ecost[off] = key_cost((uint32_t)(best[tid] >> 32));  // pair costs are never -0.0
      link[tid] = pr;  // the accepted edges form a forest: one pointer per root
    }
    ne += __popcll(gm);
    wave_sync();
    // each tree's largest old root (its largest member) becomes the
    // component's root; partner is free after the accept and holds the maxima
    if (me) partner[tid] = tid;
    wave_sync();
    if (me && root[tid] == tid) atomicMax(&partner[dsu_find(link, tid)], tid);
    wave_sync();
    if (me) root[tid] = partner[dsu_find(link, root[tid])];
    wave_sync();
  }
  if (!valid) return;
  const int64_t sb = A.slot[pos];
  const int16_t* pat = A.pair_pat + A.pair[pos];
  const int32_t* npins = A.pins + pb;
  for (int e = tid; e < n - 1; e += G) emit_edge(A, bs, pos, sb, e, edges[e].x, edges[e].y, pat, npins);
  if (tid == n - 1) {  // a lane without an edge
    A.out_edges[sb + n - 1] = -1;
    A.out_lpat[sb + n - 1] = -1;
    A.out_cost[A.cost_base + pos] = sum_edge_costs(ecost, n - 1);
  }
}


template <int G>
__device__ __forceinline__ void run_group_class(const TreeArgs& A, int32_t bs, const SmallClasses& C, int cls, int b,
                                                int W, GroupLds* lds) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int idx = ((b - C.block0[cls]) * W + wv) * (64 / G) + lane / G;
  tree_group<G>(A, bs, C.list[cls], C.n[cls], idx, lds[wv]);
}

// Small-net work of block b (block0 ranges of C): W wavefronts, lds[W].
__device__ __forceinline__ void small_block(const TreeArgs& A, int32_t bs, const SmallClasses& C, int b, int W, GroupLds* lds) {
  if (b < C.block0[1]) {
    const int idx = b * 64 * W + threadIdx.x;
    if (idx < C.n[0]) tree2_commit(A, bs, C.list[0][idx]);
  } else if (b < C.block0[2]) {
    run_group_class<8>(A, bs, C, 1, b, W, lds);
  } else if (b < C.block0[3]) {
    run_group_class<16>(A, bs, C, 2, b, W, lds);
  } else if (b < C.block0[4]) {
    run_group_class<32>(A, bs, C, 3, b, W, lds);
  } else {
    run_group_class<64>(A, bs, C, 4, b, W, lds);
  }
}

__global__ __launch_bounds__(256) void k_tree_small(TreeArgs A0, SmallClasses C) {
  const TreeArgs& A = A0;
  const int32_t bs = *A0.ctr;  // this batch's stamp
  __shared__ GroupLds lds[4];
  small_block(A, bs, C, blockIdx.x, 4, lds);
}

// LDS carve-up of a workgroup tree for nets of up to cap pins, mcap matrix
// entries and scap staged keys (the host sizes the launch with
// tree_lds_bytes).  The compact ids of the roots share the partner array:
// they are written after the flatten and read by the matrix merge or build
// of the same round, and partner is written again only in the next round.
// The component root lists hold the most components a matrix of mcap
// entries can describe.
// Components a matrix of mcap entries can describe: the smallest c with
// c(c-1)/2 > mcap, at most cap + 1.  Closed form with an exact correction
// (every wavefront of a tree computes it at entry; a counting loop there
// ran about 200 scalar iterations per wavefront, all of a CU's wavefronts
// sharing its scalar unit, and delayed the first barrier by up to 16 us).
__host__ __device__ inline int tree_ccap(int cap, int mcap) {
  if (mcap <= 0) return cap + 1 < 2 ? cap + 1 : 2;  // no matrix: c(c-1)/2 > 0 first at c = 2
  int64_t c = (int64_t)((1.0 + sqrt(1.0 + 8.0 * (double)mcap)) * 0.5);
  if (c < 1) c = 1;
  while (c > 1 && (c - 1) * (c - 2) / 2 > mcap) --c;
  while (c * (c - 1) / 2 <= mcap) ++c;
  return (int)(c < (int64_t)cap + 1 ? c : (int64_t)cap + 1);
}
__host__ __device__ inline size_t tree_lds_bytes(int cap, int mcap, int scap) {
  return (size_t)cap * 16 + (size_t)mcap * 8 + (size_t)cap * 20 + (size_t)tree_ccap(cap, mcap) * 8 +
         (size_t)scap * 4;
}
__device__ inline TreeLds tree_lds(unsigned long long* shl, int cap, int mcap, int scap, int* wcnt, int* sh) {
  TreeLds L;
  const int ccap = tree_ccap(cap, mcap);
  L.best = shl;
  L.edges = (int2*)(L.best + cap);
  L.mat = (unsigned long long*)(L.edges + cap);
  L.mcap = mcap;
  L.root = (int*)(L.mat + mcap);
  L.partner = L.root + cap;
  L.par_a = L.partner + cap;
  L.par_b = L.par_a + cap;
  L.ecost = (float*)(L.par_b + cap);
  L.cid = L.partner;
  L.croot_a = (int*)(L.ecost + cap);
  L.croot_b = L.croot_a + ccap;
  L.skey = (uint32_t*)(L.croot_b + ccap);
  L.scap = scap;
  L.wcnt = wcnt;
  L.sh = sh;
  return L;
}

// One workgroup per net of more than kSmallDeg pins; LDS sized by the largest
// such net of the batch (cap nodes).  MAT: matrix rounds (launches whose
// largest net's keys do not fit in LDS); without them the code is leaner.
template <bool MAT>
__global__ __launch_bounds__(1024) void k_tree_large(TreeArgs A0, int cap, int mcap, int scap) {
  const TreeArgs& A = A0;
  const int32_t bs = *A0.ctr;  // this batch's stamp
  extern __shared__ unsigned long long shl[];
  __shared__ int wcnt[((kLargeDeg + 1023) / 1024) * 16];
  __shared__ int s_sh[2];
  const TreeLds L = tree_lds(shl, cap, mcap, scap, wcnt, s_sh);
  tree_commit<MAT>(A, bs, A.list[blockIdx.x], threadIdx.x, blockDim.x, L);
}

// One workgroup per net of more than kLargeDeg pins: the node arrays live in
// global memory (44 B per slot of the batch, at the net's slot offset), the
// LDS holds the component matrix; keys are always streamed.  Global memory
// written by the workgroup's own threads is seen by them after each barrier,
// and the merge's atomic maxima and volatile reads go to the L2.
__global__ __launch_bounds__(1024) void k_tree_giant(TreeArgs A0, int mcap, uint8_t* scratch, int64_t nslots,
                                                     int64_t slot0) {
  const TreeArgs& A = A0;
  const int32_t bs = *A0.ctr;  // this batch's stamp
  extern __shared__ unsigned long long shl[];
  __shared__ int wcnt[((kMaxDeg + 1023) / 1024) * 16];
  __shared__ int s_sh[2];
  const int pos = A.list[blockIdx.x];
  const int64_t o = A.slot[pos] - slot0;
  TreeLds L;
  L.best = (unsigned long long*)scratch + o;
  L.edges = (int2*)(scratch + nslots * 8) + o;
  int* ib = (int*)(scratch + nslots * 16);
  L.root = ib + o;
  L.partner = ib + nslots + o;
  L.par_a = ib + 2 * nslots + o;
  L.par_b = ib + 3 * nslots + o;
  L.ecost = (float*)(ib + 4 * nslots) + o;
  L.cid = L.partner;
  L.croot_a = ib + 5 * nslots + o;
  L.croot_b = ib + 6 * nslots + o;
  L.mat = shl;
  L.mcap = mcap;
  L.skey = nullptr;
  L.scap = 0;
  L.wcnt = wcnt;
  L.sh = s_sh;
  tree_commit<true, true>(A, bs, pos, threadIdx.x, blockDim.x, L);
}

// Small and large nets of one batch in one launch (no second stream, so no
// fork and join between the costing and the refresh): workgroups
// [0, n_large) take one large net each with all blockDim threads, the others
// run small-net work on their first W wavefronts (the rest exit at once).
// The dynamic LDS holds either the large tree or W group areas; the launch
// takes the larger of the two.
template <bool MAT>
__global__ __launch_bounds__(1024) void k_tree_mixed(TreeArgs A0, SmallClasses C, int n_large, int cap, int mcap,
                                                     int scap, int W) {
  const TreeArgs& A = A0;
  const int32_t bs = *A0.ctr;  // this batch's stamp
  extern __shared__ unsigned long long shl[];
  __shared__ int wcnt[((kLargeDeg + 1023) / 1024) * 16];
  __shared__ int s_sh[2];
  if ((int)blockIdx.x < n_large) {
    const TreeLds L = tree_lds(shl, cap, mcap, scap, wcnt, s_sh);
    tree_commit<MAT>(A, bs, A.list[blockIdx.x], threadIdx.x, blockDim.x, L);
  } else {
    if ((int)(threadIdx.x >> 6) >= W) return;
    small_block(A, bs, C, (int)blockIdx.x - n_large, W, (GroupLds*)shl);
  }
}


// ---------------------------------------------------------------------------
// One launch per batch: costing, trees and commit (tuning fork = 4)
// ---------------------------------------------------------------------------

// Work items of a fused batch launch.  The batch's work is split into 8
// queues, one per XCD; a workgroup reads the XCC_ID of the XCD it runs on and
// claims the next item of that XCD's queue (the dispatcher deals the grid's
// workgroups to the XCDs in turn, so a grid of 8 * qlen gives every XCD qlen
// workgroups).  Each item depends only on items of its own queue, so a
// dependency is met through that XCD's L2 without writing the L2 back:
//   kind 0  costing chunk: pairs [a, a + b) of the batch, then one count on
//           large-net counter c, or on this XCD's small-net counter (c < 0)
//   kind 1  large-net tree: batch position a, once its b chunks are counted
//           on counter c
//   kind 2  small-net tree block a of this XCD's classes, once the XCD's b
//           small-net chunks are counted; c = the XCD's small blocks
//   kind 3  nothing (pads the queues to one length)
// A tree waits only for chunks claimed before it from its own queue, and
// chunks wait for nothing, so every wait ends: the chunks it waits for are
// held by running workgroups.  The wait is bounded all the same (a time-out
// sets the error word, which the host reports).  The claim counters are
// cleared by the batch's refresh launch.
struct FusedArgs {
  const int4* items;       // queue x at items[x * qlen]
  int32_t qlen;
  int32_t* heads;          // per XCD: items claimed so far
  const SmallClasses* cx;  // the small-net classes of each XCD
  int32_t* cnt;            // [0] trees finished, [1] error word, [2 + x] XCD x's small chunks finished,
                           // [10 + x] XCD x's small blocks past their wait, [18 + x] XCD x claim counter (heads), [26 + k] large net k's chunks finished
  int32_t n_tree;          // tree items of the batch (kinds 1 and 2)
  const uint2* ends;       // the batch's endpoint table
  const float* pfx;
  float* out_cost;         // the batch's pair costs and patterns (the trees read them)
  int16_t* out_pat;
  int32_t* ctr;            // the device batch counter
};
constexpr int kSpinLimit = 1 << 22;  // s_sleep(2) rounds: about a quarter of a second
// s_waitcnt operand (gfx9 encoding): vmcnt 0, expcnt and lgkmcnt not waited on
constexpr int kWaitVm0 = 0x0F70;
// s_getreg operand: HW_REG_XCC_ID (20), bits [3:0]
constexpr int kHwRegXccId = 20 | (3 << 11);

// Costing of one chunk: every wavefront waits for its stores to reach the
// L2 (vmcnt 0: the workgroup barrier does not wait for global stores), then
// one count (a relaxed agent-scope atomic) publishes them to the XCD.
template <int ZT, bool BUF>
__device__ __forceinline__ void fused_chunk(const Geo& g, const FusedArgs& F, int4 it, int xcd) {
  for (int t = threadIdx.x; t < it.y; t += blockDim.x) {
    const int64_t w = (int64_t)it.x + t;
    const uint2 e = F.ends[w];
    int z1, x1, y1, z2, x2, y2;
    unpack_cell(e.x, z1, x1, y1);
    unpack_cell(e.y, z2, x2, y2);
    const uint64_t best = lane_best_pts<ZT, BUF>(g, F.pfx, z1, x1, y1, z2, x2, y2);
    F.out_cost[w] = key_cost((uint32_t)(best >> 32));
    F.out_pat[w] = (int16_t)(uint32_t)best;
  }
  __builtin_amdgcn_s_waitcnt(kWaitVm0);
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_fetch_add(it.z >= 0 ? &F.cnt[26 + it.z] : &F.cnt[2 + xcd], 1, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

template <int ZT, bool BUF>
__global__ __launch_bounds__(1024) void k_batch(TreeArgs A0, FusedArgs F, int cap, int scap, int W) {
  const TreeArgs& A = A0;
  extern __shared__ unsigned long long shl[];
  __shared__ int wcnt[((kLargeDeg + 1023) / 1024) * 16];
  __shared__ int s_sh[2];
  __shared__ int32_t s_bs;
  __shared__ int s_claim[2];
  if (threadIdx.x == 0) {
    const int x = (int)(__builtin_amdgcn_s_getreg(kHwRegXccId) & 7);
    const int j = __hip_atomic_fetch_add(&F.heads[x], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_claim[0] = x;
    s_claim[1] = j < F.qlen ? x * F.qlen + j : -1;
  }
  __syncthreads();
  const int xcd = s_claim[0];
  if (s_claim[1] < 0) return;  // more workgroups than items on this XCD
  const int4 item = F.items[s_claim[1]];
  const int kind = item.x;
  const int4 it = make_int4(item.y, item.z, item.w, 0);
  if (kind == 3) return;
  GR_ITEM_MARK(0, kind, xcd);
  if (kind == 0) {
    fused_chunk<ZT, BUF>(A.g, F, it, xcd);
    GR_ITEM_MARK(2, kind, xcd);
    return;
  }
  if (threadIdx.x == 0) {
    // relaxed polls, then an L1 invalidation: this CU's L1 may hold lines of
    // the pair arrays read before the chunks wrote them
    int32_t* cp = kind == 1 ? &F.cnt[26 + it.z] : &F.cnt[2 + xcd];
    int spins = 0;
    while (__hip_atomic_load(cp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < it.y) {
      if (++spins > kSpinLimit) {
        __hip_atomic_store(&F.cnt[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    // the chunks' stores are in this XCD's L2; invalidate only this CU's L1
    // (an agent-scope acquire would also invalidate the L2 the XCD's other
    // workgroups are costing from).  The memory clobber keeps the pair loads
    // below it.
    asm volatile("buffer_inv sc0" ::: "memory");
    // every chunk this item waited for is in: its counter can start over
    // for the next batch (small-net: once every small block of the XCD is past)
    if (kind == 1) {
      __hip_atomic_store(cp, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (__hip_atomic_fetch_add(&F.cnt[10 + xcd], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
               it.z - 1) {
      __hip_atomic_store(cp, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&F.cnt[10 + xcd], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    s_bs = __hip_atomic_load(F.ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;  // published below
  }
  __syncthreads();
  GR_ITEM_MARK(1, kind, xcd);
  const int32_t bs = s_bs;
  if (kind == 1) {
    const TreeLds L = tree_lds(shl, cap, 0, scap, wcnt, s_sh);
    tree_commit<false>(A, bs, it.x, threadIdx.x, blockDim.x, L);
  } else if ((int)(threadIdx.x >> 6) < W) {
    small_block(A, bs, F.cx[xcd], it.x, W, (GroupLds*)shl);
  }
  // the batch's last tree publishes its stamp for the refresh (every tree
  // read the old counter before counting itself)
  __syncthreads();
  GR_ITEM_MARK(2, kind, xcd);
  if (threadIdx.x == 0 &&
      __hip_atomic_fetch_add(&F.cnt[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == F.n_tree - 1) {
    __hip_atomic_store(F.ctr, bs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&F.cnt[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

}  // namespace grd
