// Learned proposal for L-pattern costing (no counterpart in the reference):
// a 3-layer MLP scores the candidate patterns of each pin pair so that only
// its top-k are costed exactly.  Hand-written gfx950 kernels, all on bf16 MFMA
// (v_mfma_f32_16x16x32_bf16, fp32 accumulation): the fused routing kernel
// k_prop and its persistent form k_prop_tiles (feature gather, three layers
// in LDS, top-k, candidate costing; one launch per batch), the training GEMMs (k_gemm forward and data gradients,
// k_gemm_tn split-K weight and bias gradients), softmax cross-entropy,
// momentum SGD on fp32 master weights, and the feature / top-k kernels of the
// general path (hidden widths other than 256, more than 64 candidates).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "gr_internal.h"

using namespace grd;
using namespace grh;

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));

namespace {

constexpr int kMaxTopK = 8;
constexpr int kSplit = 8;  // row chunks of the weight-gradient products

// bf16 <-> fp32 by bit manipulation (round to nearest even); identical on host and device.
__host__ __device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u;
#ifdef __HIP_DEVICE_COMPILE__
  u = __float_as_uint(x);
#else
  std::memcpy(&u, &x, 4);
#endif
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
#ifdef __HIP_DEVICE_COMPILE__
  return __uint_as_float(u);
#else
  float f;
  std::memcpy(&f, &u, 4);
  return f;
#endif
}

// ---------------------------------------------------------------------------
// Features
// ---------------------------------------------------------------------------

// Row of k_in bf16 features for the pair src -> dst:
//   [0] dx/64, [1] dy/64, [2] z_src, [3] z_dst,
//   [4 + l*25 + u*5 + v] log cost of cell (l, x1 + round(u*dx/4), y1 + round(v*dy/4)),
//   zero padding to k_in.
__device__ void feat_row(const Geo& g, const int32_t* __restrict__ rem, const float* __restrict__ loglut,
                         uint32_t src, uint32_t dst, uint16_t* __restrict__ row, int k_in) {
  int z1, x1, y1, z2, x2, y2;
  dec(g, src, z1, x1, y1);
  dec(g, dst, z2, x2, y2);
  const int dx = x2 - x1, dy = y2 - y1;
  row[0] = f2bf((float)dx * (1.0f / 64.0f));
  row[1] = f2bf((float)dy * (1.0f / 64.0f));
  row[2] = f2bf((float)z1);
  row[3] = f2bf((float)z2);
  int xs[5], ys[5];
#pragma unroll
  for (int u = 0; u < 5; ++u) {
    xs[u] = x1 + (dx * u + (dx >= 0 ? 2 : -2)) / 4;
    ys[u] = y1 + (dy * u + (dy >= 0 ? 2 : -2)) / 4;
  }
  int f = 4;
  for (int l = 0; l < g.Z; ++l)
    for (int u = 0; u < 5; ++u)
      for (int v = 0; v < 5; ++v) {
        int r = rem[enc(g, l, xs[u], ys[v])];
        r = r < kRMin ? kRMin : (r > kRMin + kLut - 1 ? kRMin + kLut - 1 : r);
        row[f++] = f2bf(loglut[r - kRMin]);
      }
  for (; f < k_in; ++f) row[f] = 0;
}

// Features of every pair of a batch (costing) or of a strided sample of them
// (training capture: sample i is pair i*stride; its label is that pair's
// chosen pattern minus Z2, or -1 when fb is given and the pair's choice came
// from the proposal's candidates rather than from every pattern).
__global__ __launch_bounds__(256) void k_feat_batch(Geo g, const int32_t* __restrict__ rem,
                                                    const float* __restrict__ loglut,
                                                    const int32_t* __restrict__ pins,
                                                    const int64_t* __restrict__ pinofs,
                                                    const int64_t* __restrict__ bpair, int n, int64_t count,
                                                    int64_t stride, uint16_t* __restrict__ X, int k_in,
                                                    const int16_t* __restrict__ pat, int32_t* __restrict__ y,
                                                    const uint8_t* __restrict__ fb) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  const int64_t w = i * stride;
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (bpair[mid] <= w) lo = mid; else hi = mid;
  }
  int j, k;
  pair_decode((uint32_t)(w - bpair[lo]), j, k);
  const int64_t pb = pinofs[lo];
  feat_row(g, rem, loglut, (uint32_t)pins[pb + j], (uint32_t)pins[pb + k], X + i * k_in, k_in);
  if (y) y[i] = (fb && !fb[w]) ? -1 : (int32_t)pat[w] - g.Z2;
}

__global__ __launch_bounds__(256) void k_feat_pairs(Geo g, const int32_t* __restrict__ rem,
                                                    const float* __restrict__ loglut,
                                                    const int32_t* __restrict__ src,
                                                    const int32_t* __restrict__ dst, int n,
                                                    uint16_t* __restrict__ X, int k_in) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  feat_row(g, rem, loglut, (uint32_t)src[i], (uint32_t)dst[i], X + (int64_t)i * k_in, k_in);
}

// ---------------------------------------------------------------------------
// bf16 MFMA GEMM
// ---------------------------------------------------------------------------

// C[M x N] = epilogue(A[M x K] * B[K x N]); A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn]; K % 32 == 0.  Epilogue: + bias[n], ReLU,
// zero where mask(m,n) <= 0, store bf16 or fp32.  A 256-thread block
// computes a 64x64 tile as 2x2 wavefronts of 32x32, each 2x2 MFMA
// 16x16x32 fragments: lane l holds A[l%16][8*(l/16) .. +8), B[8*(l/16) ..
// +8)[l%16] and D[4*(l/16) + i][l%16], i < 4.
template <bool RELU, bool OUT_F32>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const uint16_t* __restrict__ A, int64_t sam,
                                              int64_t sak, const uint16_t* __restrict__ B, int64_t sbk, int64_t sbn,
                                              const float* __restrict__ bias, const uint16_t* __restrict__ mask,
                                              int64_t ldm, uint16_t* __restrict__ Cb, float* __restrict__ Cf,
                                              int64_t ldc) {
  __shared__ __attribute__((aligned(16))) uint16_t As[64][40];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[64][40];
  const int t = threadIdx.x, wv = t >> 6, lane = t & 63;
  const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
  v4f acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = (v4f){0.f, 0.f, 0.f, 0.f};
  const int lr = t >> 2, lk = (t & 3) * 8;
  const int gm = m0 + lr, gn = n0 + lr;
  for (int k0 = 0; k0 < K; k0 += 32) {
    if (gm < M && sak == 1) {
      *(uint4*)&As[lr][lk] = *(const uint4*)(A + (int64_t)gm * sam + k0 + lk);
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) As[lr][lk + i] = gm < M ? A[(int64_t)gm * sam + (int64_t)(k0 + lk + i) * sak] : 0;
    }
    if (gn < N && sbk == 1) {
      *(uint4*)&Bs[lr][lk] = *(const uint4*)(B + (int64_t)gn * sbn + k0 + lk);
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) Bs[lr][lk + i] = gn < N ? B[(int64_t)(k0 + lk + i) * sbk + (int64_t)gn * sbn] : 0;
    }
    __syncthreads();
    v8bf fa[2], fb[2];
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      fa[f] = __builtin_bit_cast(v8bf, *(const uint4*)&As[wm + f * 16 + (lane & 15)][8 * (lane >> 4)]);
      fb[f] = __builtin_bit_cast(v8bf, *(const uint4*)&Bs[wn + f * 16 + (lane & 15)][8 * (lane >> 4)]);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = m0 + wm + a * 16 + 4 * (lane >> 4) + i;
        const int col = n0 + wn + b * 16 + (lane & 15);
        if (row >= M || col >= N) continue;
        float v = acc[a][b][i];
        if (bias) v += bias[col];
        if (RELU) v = v > 0.f ? v : 0.f;
        if (mask && !(bf2f(mask[(int64_t)row * ldm + col]) > 0.f)) v = 0.f;
        if (OUT_F32) Cf[(int64_t)row * ldc + col] = v;
        else Cb[(int64_t)row * ldc + col] = f2bf(v);
      }
}

// ---------------------------------------------------------------------------
// Loss, bias gradients, update
// ---------------------------------------------------------------------------

// Softmax cross-entropy over the first n_out logits of each row; labels < 0
// mark padding rows.  dL = (softmax - onehot) / n_real as bf16 (padding
// columns 0).  loss_sum += -log p[y] (float atomics over rows: the sum's last
// bits depend on their order; the gradient does not).  One wavefront per row,
// lane j holding columns j and j + 64 (n_out <= 128).
__global__ __launch_bounds__(256) void k_softmax_ce(int rows, int n_out, int ld, const float* __restrict__ L,
                                                    const int32_t* __restrict__ y, float inv_n,
                                                    uint16_t* __restrict__ dL, float* __restrict__ loss_sum) {
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (r >= rows) return;
  const float* l = L + (int64_t)r * ld;
  uint16_t* d = dL + (int64_t)r * ld;
  const int lab = y[r];
  const bool v0 = lane < n_out, v1 = lane + 64 < n_out;
  const float x0 = v0 ? l[lane] : -INFINITY, x1 = v1 ? l[lane + 64] : -INFINITY;
  float mx = fmaxf(x0, x1);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
  const float e0 = v0 ? expf(x0 - mx) : 0.f, e1 = v1 ? expf(x1 - mx) : 0.f;
  float sm = e0 + e1;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) sm += __shfl_xor(sm, o, 64);
  const float inv = 1.f / sm;
  if (lane < ld) d[lane] = lab < 0 || !v0 ? (uint16_t)0 : f2bf((e0 * inv - (lane == lab ? 1.f : 0.f)) * inv_n);
  if (lane + 64 < ld) d[lane + 64] = lab < 0 || !v1 ? (uint16_t)0 : f2bf((e1 * inv - (lane + 64 == lab ? 1.f : 0.f)) * inv_n);
  if (lane == 0 && lab >= 0) atomicAdd(loss_sum, -(l[lab] - mx - logf(sm)));
}

// Weight and bias gradients of one layer, split over the rows (split-K):
// part z of C [(M + 1) x N] = sum over rows r in [z*rz, (z+1)*rz) of
// A'[r][m] B[r][n], A' = [A | 1] (A: [R x M] row-major, ld lda; B: [R x N],
// ld ldb; rz % 32 == 0), so row M of C is the bias gradient.  Tiles of 32
// rows are loaded coalesced (8 consecutive columns per thread) and stored
// transposed in LDS, where the MFMA fragments (8 consecutive rows of one
// column) are 16-byte reads.  64 x 64 output tile per 256 threads, 2 x 2
// wavefronts of 32 x 32, as k_gemm.
__global__ __launch_bounds__(256) void k_gemm_tn(int M, int N, int rz, const uint16_t* __restrict__ A, int64_t lda,
                                                 const uint16_t* __restrict__ B, int64_t ldb, float* __restrict__ Cp,
                                                 int64_t ldc) {
  __shared__ __attribute__((aligned(16))) uint16_t As[64][40];  // [m][r]
  __shared__ __attribute__((aligned(16))) uint16_t Bs[64][40];  // [n][r]
  const int t = threadIdx.x, wv = t >> 6, lane = t & 63;
  const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
  const int64_t rbase = (int64_t)blockIdx.z * rz;
  float* C = Cp + (int64_t)blockIdx.z * (M + 1) * ldc;
  v4f acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = (v4f){0.f, 0.f, 0.f, 0.f};
  const int lr = t >> 3, lc = (t & 7) * 8;
  const int gm = m0 + lc, gn = n0 + lc;
  for (int r0 = 0; r0 < rz; r0 += 32) {
    const int64_t row = rbase + r0 + lr;
    uint16_t av[8], bv[8];
    if (gm + 8 <= M) {
      const uint4 v = *reinterpret_cast<const uint4*>(A + row * lda + gm);
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int i = 0; i < 8; ++i) av[i] = (uint16_t)(w[i >> 1] >> (16 * (i & 1)));
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) av[i] = gm + i < M ? A[row * lda + gm + i] : (gm + i == M ? (uint16_t)0x3F80 : 0);
    }
    if (gn + 8 <= N) {
      const uint4 v = *reinterpret_cast<const uint4*>(B + row * ldb + gn);
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int i = 0; i < 8; ++i) bv[i] = (uint16_t)(w[i >> 1] >> (16 * (i & 1)));
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) bv[i] = gn + i < N ? B[row * ldb + gn + i] : 0;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      As[lc + i][lr] = av[i];
      Bs[lc + i][lr] = bv[i];
    }
    __syncthreads();
    v8bf fa[2], fb[2];
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      fa[f] = __builtin_bit_cast(v8bf, *(const uint4*)&As[wm + f * 16 + (lane & 15)][8 * (lane >> 4)]);
      fb[f] = __builtin_bit_cast(v8bf, *(const uint4*)&Bs[wn + f * 16 + (lane & 15)][8 * (lane >> 4)]);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = m0 + wm + a * 16 + 4 * (lane >> 4) + i;
        const int col = n0 + wn + b * 16 + (lane & 15);
        if (r <= M && col < N) C[(int64_t)r * ldc + col] = acc[a][b][i];
      }
}

// out[i] = sum over z < nz of parts[z][i], in z order.
__global__ __launch_bounds__(256) void k_sum_parts(int64_t n, int nz, const float* __restrict__ parts,
                                                   float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float s = parts[i];
  for (int z = 1; z < nz; ++z) s += parts[(int64_t)z * n + i];
  out[i] = s;
}

// Momentum SGD on the fp32 master parameters.
__global__ __launch_bounds__(256) void k_sgd(int n, float* __restrict__ w, float* __restrict__ mom,
                                             const float* __restrict__ g, float lr) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float m = 0.9f * mom[i] + g[i];
  mom[i] = m;
  w[i] -= lr * m;
}

// bf16 copies of a [K x N] fp32 weight: row-major Wb and transposed Wt.
__global__ __launch_bounds__(256) void k_to_bf16(int K, int N, const float* __restrict__ W, uint16_t* __restrict__ Wb,
                                                 uint16_t* __restrict__ Wt) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= K * N) return;
  const int k = i / N, n = i - k * N;
  const uint16_t h = f2bf(W[i]);
  Wb[i] = h;
  Wt[(int64_t)n * K + k] = h;
}

__global__ __launch_bounds__(256) void k_gather_rows(int rows, int k_in, const int32_t* __restrict__ perm,
                                                     const uint16_t* __restrict__ X, const int32_t* __restrict__ y,
                                                     uint16_t* __restrict__ Xo, int32_t* __restrict__ yo) {
  const int r = blockIdx.x;
  if (r >= rows) return;
  const int64_t src = perm[r];
  for (int f = threadIdx.x; f < k_in; f += 256) Xo[(int64_t)r * k_in + f] = X[src * k_in + f];
  if (threadIdx.x == 0) yo[r] = y[src];
}

// ---------------------------------------------------------------------------
// Top-k costing
// ---------------------------------------------------------------------------

// One lane per pair: the k largest logits (ties: lower index first) name the
// candidate patterns p = q + Z2; each is costed exactly and the lexicographic
// minimum of (cost, p) wins.  When the top-1 softmax probability is below
// thr, or no candidate is a valid route, every pattern is costed.
__global__ __launch_bounds__(256) void k_lpat_topk(Geo g, const float* __restrict__ pfx,
                                                   const int32_t* __restrict__ pins,
                                                   const int64_t* __restrict__ pinofs,
                                                   const int64_t* __restrict__ bpair,
                                                   const int32_t* __restrict__ chunk_net, int n, int64_t P,
                                                   const float* __restrict__ logits, int ld, int n_out, int topk,
                                                   float thr, float* __restrict__ out_cost,
                                                   int16_t* __restrict__ out_pat, int32_t* __restrict__ ctr,
                                                   unsigned long long* __restrict__ stats) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *ctr += 1;  // this batch's stamp
  const int64_t w0 = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t w = w0 < P ? w0 : P - 1;
  const PairOwner own = pair_owner_wave(bpair, pinofs, chunk_net, n, w);
  bool fb = false;
  if (w0 < P) {
    int j, k;
    pair_decode((uint32_t)(w - own.first_pair), j, k);
    const int64_t pb = own.first_pin;
    int z1, x1, y1, z2, x2, y2;
    dec(g, (uint32_t)pins[pb + j], z1, x1, y1);
    dec(g, (uint32_t)pins[pb + k], z2, x2, y2);
    const float* l = logits + w * ld;
    float tv[kMaxTopK];
    int ti[kMaxTopK];
#pragma unroll
    for (int t = 0; t < kMaxTopK; ++t) { tv[t] = -INFINITY; ti[t] = -1; }
    float mx = -INFINITY;
    for (int q = 0; q < n_out; ++q) {
      float v = l[q];
      mx = fmaxf(mx, v);
      int vi = q;
#pragma unroll
      for (int t = 0; t < kMaxTopK; ++t) {  // insertion keeps (value desc, index asc)
        if (t < topk && (v > tv[t] || ti[t] < 0)) {
          const float sv = tv[t];
          const int si = ti[t];
          tv[t] = v; ti[t] = vi;
          v = sv; vi = si;
        }
      }
    }
    float s = 0.f;
    for (int q = 0; q < n_out; ++q) s += expf(l[q] - mx);
    uint64_t best = ~0ull;
    fb = (1.0f / s) < thr;
    if (!fb) {
#pragma unroll
      for (int t = 0; t < kMaxTopK; ++t) {
        if (t >= topk || ti[t] < 0) continue;
        const int p = ti[t] + g.Z2;
        float c;
        if (pattern_cost(pfx, g, z1, x1, y1, z2, x2, y2, p, c)) {
          const uint64_t key = ((uint64_t)ord_key(c) << 32) | (uint32_t)p;
          best = key < best ? key : best;
        }
      }
      fb = best == ~0ull;
    }
    if (fb) {
      for (int p = g.Z2; p < g.Z * g.Z2; ++p) {
        float c;
        if (pattern_cost(pfx, g, z1, x1, y1, z2, x2, y2, p, c)) {
          const uint64_t key = ((uint64_t)ord_key(c) << 32) | (uint32_t)p;
          best = key < best ? key : best;
        }
      }
    }
    out_cost[w] = key_cost((uint32_t)(best >> 32));
    out_pat[w] = (int16_t)(uint32_t)best;
  }
  const unsigned long long act = __ballot(w0 < P), fbm = __ballot(fb && w0 < P);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&stats[0], (unsigned long long)__popcll(act));
    atomicAdd(&stats[1], (unsigned long long)__popcll(fbm));
  }
}

// Per-GCELL copy of every layer's remaining capacity for the lattice gather:
// dst[(y*X + x)*16 + z] = src at (z, x, y), clamped to the cost table's range
// (z < Z; the rest 0).  One thread per GCELL, two 16-byte stores.
__global__ __launch_bounds__(256) void k_remz_build(Geo g, const int32_t* __restrict__ src, int16_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)g.XY) return;
  const int y = (int)(i / g.X), x = (int)(i - (int64_t)y * g.X);
  uint32_t w[8];
#pragma unroll
  for (int h = 0; h < 8; ++h) {
    uint32_t pk = 0;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int z = 2 * h + e;
      int v = 0;
      if (z < g.Z) {
        v = src[enc(g, z, x, y)];
        v = v < kRMin ? kRMin : (v > kRMin + kLut - 1 ? kRMin + kLut - 1 : v);
      }
      pk |= (uint32_t)(uint16_t)(int16_t)v << (16 * e);
    }
    w[h] = pk;
  }
  uint4* d = reinterpret_cast<uint4*>(dst + i * 16);
  d[0] = make_uint4(w[0], w[1], w[2], w[3]);
  d[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// ---------------------------------------------------------------------------
// Fused proposal: features -> three MFMA layers -> top-k costing, per tile
// ---------------------------------------------------------------------------

// A workgroup of NW = TT/16 wavefronts takes TT consecutive pairs (or input
// rows); TT = 128 is used: two 8-wave workgroups per CU (at most 128 VGPRs,
// 74 KB of LDS each), and each weight fragment serves 128 pairs.  Activations stay in LDS as one [TT][RS] bf16 buffer (features, then
// H1, then H2), so only the grid cells, the weights and the results cross the
// memory hierarchy.  Layers run transposed: for neurons n
// and pairs r, D[n][r] = sum_k W^T[n][k] X[r][k] with W^T in MFMA fragment
// order (see k_pack) as the A operand from global memory (L2-resident, 280 KB
// at the 256-wide model) and X from LDS as the B operand, so each lane's four
// results are four consecutive neurons of one pair: one 8-byte LDS store.
// Hidden layers: wave w owns neurons [16 NFH w, 16 NFH (w + 1)) for all TT
// pairs (NFH = 16/NW fragments of 16 neurons).  Output layer: wave f owns the
// output fragment f for all pairs; the logits then go through LDS to the
// top-k, where wave w takes pairs [16w, 16w + 16).  The K order of every sum and
// the bias, ReLU and bf16 rounding points are those of k_gemm's forward pass.
constexpr int kHid = 256;  // hidden width served by the fused kernel
constexpr int kTile = 128;  // rows per workgroup of the fused kernel (8 wavefronts; two fit a CU)

template <int ZT>
struct PropShape {
  static constexpr int NFEAT = 25 * ZT;                         // lattice features
  static constexpr int KIN = (4 + NFEAT + 31) / 32 * 32;         // padded input width
  static constexpr int RS = (KIN > kHid ? KIN : kHid) + 8;       // LDS row stride (bf16), +16 B against bank conflicts
  static constexpr int NOUT = (ZT - 1) * ((ZT + 1) / 2);        // candidate patterns
  static constexpr int NF3 = (NOUT + 15) / 16;                  // output fragments routing ranks
  static constexpr int NFO = (NOUT + 63) / 64 * 4;              // all padded output columns (forward pass)
};

// Weight [K x N] (fp32, row-major) -> bf16 in MFMA A-fragment order: block
// (f, s) of 64 lanes x 8 values holds W^T rows f*16 + (l % 16), columns
// s*32 + 8*(l / 16) + e; rows >= N are zero.  One fragment load is then one
// contiguous 1 KB read.
__global__ __launch_bounds__(256) void k_pack(int K, int N, int NF, const float* __restrict__ W,
                                              uint16_t* __restrict__ Wp) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int KS = K / 32;
  if (i >= (int64_t)NF * KS * 512) return;
  const int e = (int)(i & 7), l = (int)((i >> 3) & 63);
  const int fs = (int)(i >> 9), s = fs % KS, f = fs / KS;
  const int n = f * 16 + (l & 15), k = s * 32 + 8 * (l >> 4) + e;
  Wp[i] = n < N ? f2bf(W[(int64_t)k * N + n]) : (uint16_t)0;
}

// W^T fragments (f, 0) of the first K step, loaded ahead of mma_tile.
template <int KS, int NF>
__device__ __forceinline__ void mma_first(const uint16_t* __restrict__ wp, uint4 (&a0)[NF], int lane) {
  const uint4* wv = reinterpret_cast<const uint4*>(wp) + lane;
#pragma unroll
  for (int f = 0; f < NF; ++f) a0[f] = wv[f * KS * 64];
}

// acc[f][q] = sum over the K steps of W^T fragment (f, s) x X rows
// prow0 + 16q .. + 16; a0 holds the fragments of step 0 (mma_first).
template <int KS, int NF, int PF>
__device__ __forceinline__ void mma_tile(const uint16_t* __restrict__ wp, const uint4 (&a0)[NF], const uint16_t* act,
                                         int RS, int prow0, v4f (&acc)[NF][PF], int lane) {
  const uint4* wv = reinterpret_cast<const uint4*>(wp) + lane;
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int q = 0; q < PF; ++q) acc[f][q] = (v4f){0.f, 0.f, 0.f, 0.f};
  const uint16_t* bp = act + (prow0 + (lane & 15)) * RS + 8 * (lane >> 4);
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    uint4 a[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) a[f] = s == 0 ? a0[f] : wv[(f * KS + s) * 64];
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      const v8bf b = __builtin_bit_cast(v8bf, *(const uint4*)(bp + q * 16 * RS + s * 32));
#pragma unroll
      for (int f = 0; f < NF; ++f)
        acc[f][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a[f]), b, acc[f][q], 0, 0, 0);
    }
  }
}

// Hidden-layer epilogue: + bias, ReLU, bf16, into act[pair][neuron].
template <int NF, int PF>
__device__ __forceinline__ void store_hidden(const v4f (&acc)[NF][PF], const float* __restrict__ bias, int n0,
                                             uint16_t* act, int RS, int lane) {
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const int n = n0 + 16 * f + 4 * (lane >> 4);
    const float4 bb = *reinterpret_cast<const float4*>(bias + n);
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      float v0 = acc[f][q][0] + bb.x, v1 = acc[f][q][1] + bb.y, v2 = acc[f][q][2] + bb.z, v3 = acc[f][q][3] + bb.w;
      v0 = v0 > 0.f ? v0 : 0.f;
      v1 = v1 > 0.f ? v1 : 0.f;
      v2 = v2 > 0.f ? v2 : 0.f;
      v3 = v3 > 0.f ? v3 : 0.f;
      // two values per v_cvt_pk_bf16_f32, round to nearest even like f2bf
      uint2 pk;
      pk.x = __builtin_bit_cast(uint32_t, __builtin_convertvector((v2f){v0, v1}, v2bf));
      pk.y = __builtin_bit_cast(uint32_t, __builtin_convertvector((v2f){v2, v3}, v2bf));
      *reinterpret_cast<uint2*>(act + (16 * q + (lane & 15)) * RS + n) = pk;
    }
  }
}

struct PropArgs {
  Geo g;
  const int32_t* rem;
  const int16_t* remz;  // per-GCELL layer copy of rem (k_remz_build; kept by the refresh)
  const float* vterm;   // per-cell via terms (kept by the refresh while the proposal is set)
  const float* loglut;
  const float* pfx;
  const int32_t* pins;
  const int64_t* pinofs;
  const int64_t* bpair;
  const int32_t* chunk_net;
  const uint2* ends;  // the batch's packed pair endpoints (pack_cell), or null: found from the nets
  int n;
  int64_t P;          // pairs of the batch, or input rows
  const uint16_t* X;  // input rows [P][k_in] (forward only)
  const uint16_t *W1p, *W2p, *W3p;
  const float *b1, *b2, *b3;
  int n_op;
  float* logits;  // [P][n_op] (forward only)
  int topk;
  float thr;
  float* out_cost;
  int16_t* out_pat;
  int32_t* ctr;
  unsigned long long* stats;
  uint8_t* fbout;  // per pair: fell back to every pattern (sample capture), or null
};

// Hidden layers of one tile: act holds the tile's inputs (features) on entry
// and H2 on return; a0 holds W1's first K-step fragments of this wave
// (mma_first), loaded ahead by the caller.  Wave w owns neurons
// [16 NFH w, 16 NFH (w + 1)) of both layers for all TT rows.
template <int ZT, int TT, int NW>
__device__ __forceinline__ void prop_hidden(const PropArgs& A, uint16_t* act, const uint4 (&a0)[16 / NW], int wv,
                                            int lane) {
  constexpr int NFH = 16 / NW;
  using S = PropShape<ZT>;
  constexpr int RS = S::RS, KIN = S::KIN;
  const uint16_t* w1 = A.W1p + (size_t)wv * NFH * (KIN / 32) * 512;
  const uint16_t* w2 = A.W2p + (size_t)wv * NFH * (kHid / 32) * 512;
  v4f acc[NFH][TT / 16];
  mma_tile<KIN / 32, NFH, TT / 16>(w1, a0, act, RS, 0, acc, lane);
  uint4 a2[NFH];
  mma_first<kHid / 32, NFH>(w2, a2, lane);
  __syncthreads();
  store_hidden<NFH, TT / 16>(acc, A.b1, wv * 16 * NFH, act, RS, lane);
  __syncthreads();
  mma_tile<kHid / 32, NFH, TT / 16>(w2, a2, act, RS, 0, acc, lane);
  __syncthreads();
  store_hidden<NFH, TT / 16>(acc, A.b2, wv * 16 * NFH, act, RS, lane);
  __syncthreads();
}

// Output layer and top-k of one routed tile (act holds H2): wave f < NF3
// computes output neurons [16f, 16f + 16) for all TT pairs, so each W3
// fragment is loaded once per tile; the fp32 logits go to LDS over the
// activations, then each pair's 4 lanes take k rounds of the best (value
// desc, index asc) logit not yet taken, and the softmax top-1 probability
// comes from the 4-lane maximum and sum.  Writes smask and sfb for the
// tile's TT pairs; ends with a barrier.
template <int ZT, int TT, int NW>
__device__ __forceinline__ void prop_out_topk(const PropArgs& A, uint16_t* act, unsigned long long* smask,
                                              int8_t* sfb, int wv, int lane) {
  using S = PropShape<ZT>;
  constexpr int NF3 = S::NF3, NOUT = S::NOUT;
  constexpr int QW = TT / 16 / NW;
  static_assert(NF3 <= NW, "one output fragment per wave");
  const bool has3 = wv < NF3;  // wave-uniform
  v4f acc[1][TT / 16];
  float4 b3v = make_float4(0.f, 0.f, 0.f, 0.f);
  if (has3) {
    const uint16_t* w3 = A.W3p + (size_t)wv * (kHid / 32) * 512;
    uint4 a3[1];
    mma_first<kHid / 32, 1>(w3, a3, lane);
    mma_tile<kHid / 32, 1, TT / 16>(w3, a3, act, S::RS, 0, acc, lane);
    b3v = *reinterpret_cast<const float4*>(A.b3 + 16 * wv + 4 * (lane >> 4));
  }
  // logits (fp32) into LDS over the activations, once every wave is done with them
  constexpr int LS = NF3 * 16 + 4;  // row stride in floats
  static_assert(TT * LS * 4 <= TT * S::RS * 2, "logits fit the activation buffer");
  float* lg = reinterpret_cast<float*>(act);
  __syncthreads();
  if (has3) {
    const int n = 16 * wv + 4 * (lane >> 4);
#pragma unroll
    for (int q = 0; q < TT / 16; ++q)
      *reinterpret_cast<float4*>(lg + (16 * q + (lane & 15)) * LS + n) =
          make_float4(acc[0][q][0] + b3v.x, acc[0][q][1] + b3v.y, acc[0][q][2] + b3v.z, acc[0][q][3] + b3v.w);
  }
  __syncthreads();
  const int gq = lane >> 4;
#pragma unroll
  for (int q = 0; q < QW; ++q) {
    float vals[NF3][4];
    float mxl = -INFINITY;
    const int prq = 16 * QW * wv + 16 * q + (lane & 15);
#pragma unroll
    for (int f = 0; f < NF3; ++f) {
      const int nb = 16 * f + 4 * gq;
      const float4 lv4 = *reinterpret_cast<const float4*>(lg + prq * LS + nb);
      vals[f][0] = lv4.x;
      vals[f][1] = lv4.y;
      vals[f][2] = lv4.z;
      vals[f][3] = lv4.w;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (nb + i < NOUT) mxl = fmaxf(mxl, vals[f][i]);
    }
    float mx = fmaxf(mxl, __shfl_xor(mxl, 16, 64));
    mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
    float sl = 0.f;
#pragma unroll
    for (int f = 0; f < NF3; ++f)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (16 * f + 4 * gq + i < NOUT) sl += expf(vals[f][i] - mx);
    float sm = sl + __shfl_xor(sl, 16, 64);
    sm = sm + __shfl_xor(sm, 32, 64);
    uint32_t taken = 0;  // bit f*4 + i: this lane's logit already chosen
    uint64_t msk = 0;
    for (int t = 0; t < A.topk; ++t) {
      float bv = -INFINITY;
      int bi = 127;  // none
#pragma unroll
      for (int f = 0; f < NF3; ++f)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int n = 16 * f + 4 * gq + i;
          // ascending n, strict >: the lowest index wins among equal values
          if (n < NOUT && !((taken >> (f * 4 + i)) & 1) && (vals[f][i] > bv || bi == 127)) {
            bv = vals[f][i];
            bi = n;
          }
        }
#pragma unroll
      for (int d = 16; d <= 32; d <<= 1) {
        const float ov = __shfl_xor(bv, d, 64);
        const int oi = __shfl_xor(bi, d, 64);
        if (oi != 127 && (bi == 127 || ov > bv || (ov == bv && oi < bi))) { bv = ov; bi = oi; }
      }
      if (bi < NOUT) {
        msk |= 1ull << bi;
        if (((bi >> 2) & 3) == gq) taken |= 1u << ((bi >> 4) * 4 + (bi & 3));  // this lane's logit
      }
    }
    if (gq == 0) {
      smask[prq] = msk;
      sfb[prq] = (int8_t)((1.0f / sm) < A.thr);
    }
  }
  __syncthreads();
}

// Candidate costing of pair pr of a routed tile for one parity of the first
// wire layer (uniform per wave, so the loads do not diverge by parity); only
// the terms of the pair's candidates are loaded.  Waves 2h and 2h + 1 take
// pairs 64h .. 64h + 63; the two parity keys meet in skey[2][TT].
template <int ZT, int TT>
__device__ __forceinline__ void prop_cost_pair(const PropArgs& A, int nrow, int pr, int par, const int16_t (*sc)[6],
                                               const unsigned long long* smask, const int8_t* sfb,
                                               unsigned long long* skey) {
  const int pc = pr < nrow ? pr : nrow - 1;
  const int z1 = sc[pc][0], x1 = sc[pc][1], y1 = sc[pc][2], z2 = sc[pc][3], x2 = sc[pc][4], y2 = sc[pc][5];
  skey[par * TT + pr] =
      lane_best<ZT, true>(A.g, A.pfx, z1, x1, y1, z2, x2, y2, smask[pc], par, sfb[pc] != 0, A.vterm);
}

// Result of pair pr of the tile at t0: the smaller of its two parity keys;
// no valid candidate -> every pattern.  Returns whether the pair fell back.
template <int ZT, int TT>
__device__ __forceinline__ bool prop_finish_pair(const PropArgs& A, int64_t t0, int pr, const int16_t (*sc)[6],
                                                 const int8_t* sfb, const unsigned long long* skey) {
  const uint64_t k0 = skey[pr], k1 = skey[TT + pr];
  uint64_t best = k0 < k1 ? k0 : k1;
  bool fb = sfb[pr] != 0;
  if (best == ~0ull) {
    fb = true;
    best = lane_best<ZT>(A.g, A.pfx, sc[pr][0], sc[pr][1], sc[pr][2], sc[pr][3], sc[pr][4], sc[pr][5], ~0ull, 0, true,
                         A.vterm);
  }
  A.out_cost[t0 + pr] = key_cost((uint32_t)(best >> 32));
  A.out_pat[t0 + pr] = (int16_t)(uint32_t)best;
  if (A.fbout) A.fbout[t0 + pr] = fb ? 1 : 0;
  return fb;
}

// ROUTE: features of the batch's pairs gathered from the grid, top-k selection
// from the output fragments, then candidate costing with the full-enumeration
// fallback (k_lpat_topk's rules).  !ROUTE: rows of A.X in, logits out (the
// inference forward pass).
template <int ZT, bool ROUTE, int TT, int NW = TT / 16>
__global__ __launch_bounds__(64 * NW, 4) void k_prop(PropArgs A) {
  constexpr int NT = 64 * NW;
  constexpr int NFH = 16 / NW;
  using S = PropShape<ZT>;
  constexpr int RS = S::RS, KIN = S::KIN, NFEAT = S::NFEAT;
  __shared__ __attribute__((aligned(16))) uint16_t act[TT * RS];
  __shared__ uint16_t slut[ROUTE ? kLut : 1];         // log-cost table as bf16
  __shared__ int16_t sx[ROUTE ? TT : 1][5], sy[ROUTE ? TT : 1][5];  // lattice columns / rows
  __shared__ int16_t sc[ROUTE ? TT : 1][6];           // z1 x1 y1 z2 x2 y2
  __shared__ unsigned long long smask[ROUTE ? TT : 1];  // candidate patterns q = p - Z2
  __shared__ int8_t sfb[ROUTE ? TT : 1];              // low confidence: enumerate every pattern
  __shared__ int nfb;
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int64_t t0 = (int64_t)blockIdx.x * TT;
  const int nrow = (int)(A.P - t0 < TT ? A.P - t0 : TT);
  const Geo& g = A.g;
  uint4 a0[NFH];
  mma_first<KIN / 32, NFH>(A.W1p + (size_t)wv * NFH * (KIN / 32) * 512, a0, lane);

  if (ROUTE) {
    if (blockIdx.x == 0 && tid == 0) *A.ctr += 1;  // this batch's stamp
    if (tid == 0) nfb = 0;
    for (int i = tid; i < kLut; i += NT) slut[i] = f2bf(A.loglut[i]);
    if (tid < TT) {  // waves 0 and 1 whole: the owner search is wave-collective
      const int64_t w = t0 + (tid < nrow ? tid : nrow - 1);
      int z1, x1, y1, z2, x2, y2;
      if (A.ends) {  // the same for every lane
        const uint2 e = A.ends[w];
        unpack_cell(e.x, z1, x1, y1);
        unpack_cell(e.y, z2, x2, y2);
      } else {
        const PairOwner own = pair_owner_wave(A.bpair, A.pinofs, A.chunk_net, A.n, w);
        int j, k;
        pair_decode((uint32_t)(w - own.first_pair), j, k);
        const int64_t pb = own.first_pin;
        dec(g, (uint32_t)A.pins[pb + j], z1, x1, y1);
        dec(g, (uint32_t)A.pins[pb + k], z2, x2, y2);
      }
      sc[tid][0] = (int16_t)z1; sc[tid][1] = (int16_t)x1; sc[tid][2] = (int16_t)y1;
      sc[tid][3] = (int16_t)z2; sc[tid][4] = (int16_t)x2; sc[tid][5] = (int16_t)y2;
      const int dx = x2 - x1, dy = y2 - y1;
#pragma unroll
      for (int u = 0; u < 5; ++u) {
        sx[tid][u] = (int16_t)(x1 + (dx * u + (dx >= 0 ? 2 : -2)) / 4);
        sy[tid][u] = (int16_t)(y1 + (dy * u + (dy >= 0 ? 2 : -2)) / 4);
      }
      uint16_t* row = act + tid * RS;
      row[0] = f2bf((float)dx * (1.0f / 64.0f));
      row[1] = f2bf((float)dy * (1.0f / 64.0f));
      row[2] = f2bf((float)z1);
      row[3] = f2bf((float)z2);
      for (int f = 4 + NFEAT; f < KIN; ++f) row[f] = 0;
    }
    __syncthreads();
    // lattice features: one lattice point (pair, u, v) per thread and step,
    // all ZT layers of it (feature 4 + l*25 + u*5 + v), UG points in flight
    constexpr int NG = TT * 25, UG = 7;
    for (int gb = tid; gb < NG; gb += NT * UG) {
      uint32_t xy[UG];
      int dst[UG];
#pragma unroll
      for (int u = 0; u < UG; ++u) {
        const int gi0 = gb + NT * u, gi = gi0 < NG ? gi0 : NG - 1;
        const int pr = gi / 25, uv = gi - 25 * pr, iu = uv / 5, iv = uv - 5 * iu;
        xy[u] = (uint32_t)sy[pr][iv] * (uint32_t)g.X + (uint32_t)sx[pr][iu];
        dst[u] = gi0 < NG ? pr * RS + 4 + uv : -1;
      }
      // all layers of a lattice point in one or two 16-byte loads (remz:
      // values already clamped to the table's range)
      uint4 q[UG][2];
#pragma unroll
      for (int u = 0; u < UG; ++u) {
        const uint4* src = reinterpret_cast<const uint4*>(A.remz + (size_t)xy[u] * 16);
        q[u][0] = src[0];
        if (ZT > 8) q[u][1] = src[1];
      }
#pragma unroll
      for (int u = 0; u < UG; ++u) {
        if (dst[u] < 0) continue;
        const uint32_t w[8] = {q[u][0].x, q[u][0].y, q[u][0].z, q[u][0].w,
                               ZT > 8 ? q[u][1].x : 0u, ZT > 8 ? q[u][1].y : 0u, ZT > 8 ? q[u][1].z : 0u,
                               ZT > 8 ? q[u][1].w : 0u};
#pragma unroll
        for (int l = 0; l < ZT; ++l) {
          const int v = (int)(int16_t)(uint16_t)(w[l >> 1] >> (16 * (l & 1)));
          act[dst[u] + 25 * l] = slut[v - kRMin];
        }
      }
    }
  } else {
    constexpr int C8 = KIN / 8;
    for (int i = tid; i < TT * C8; i += NT) {
      const int pr = i / C8, c = i - pr * C8;
      const int64_t row = t0 + (pr < nrow ? pr : nrow - 1);
      *reinterpret_cast<uint4*>(act + pr * RS + c * 8) =
          *reinterpret_cast<const uint4*>(A.X + row * KIN + c * 8);
    }
  }
  __syncthreads();

  prop_hidden<ZT, TT, NW>(A, act, a0, wv, lane);
  if (!ROUTE) {
    // output layer, split by neuron fragment: wave f < NFO computes output
    // neurons [16f, 16f + 16) for all TT rows; lane l then holds, for row
    // 16q + (l & 15), the logits of neurons 16f + 4(l >> 4) + i (i < 4)
    static_assert(S::NFO <= NW, "one output fragment per wave");
    if (wv < S::NFO) {
      v4f acc[1][TT / 16];
      const uint16_t* w3 = A.W3p + (size_t)wv * (kHid / 32) * 512;
      uint4 a3[1];
      mma_first<kHid / 32, 1>(w3, a3, lane);
      mma_tile<kHid / 32, 1, TT / 16>(w3, a3, act, RS, 0, acc, lane);
      const float4 b3v = *reinterpret_cast<const float4*>(A.b3 + 16 * wv + 4 * (lane >> 4));
      const int n = 16 * wv + 4 * (lane >> 4);
#pragma unroll
      for (int q = 0; q < TT / 16; ++q) {
        const int pr = 16 * q + (lane & 15);
        if (pr < nrow)
          *reinterpret_cast<float4*>(A.logits + (t0 + pr) * A.n_op + n) =
              make_float4(acc[0][q][0] + b3v.x, acc[0][q][1] + b3v.y, acc[0][q][2] + b3v.z, acc[0][q][3] + b3v.w);
      }
    }
    return;
  }
  prop_out_topk<ZT, TT, NW>(A, act, smask, sfb, wv, lane);
  unsigned long long* skey = reinterpret_cast<unsigned long long*>(act);  // [2][TT]
  if (64 * (wv >> 1) < TT) prop_cost_pair<ZT, TT>(A, nrow, 64 * (wv >> 1) + lane, wv & 1, sc, smask, sfb, skey);
  __syncthreads();
  if (tid < nrow && prop_finish_pair<ZT, TT>(A, t0, tid, sc, sfb, skey)) atomicAdd(&nfb, 1);
  __syncthreads();
  if (tid == 0) {
    atomicAdd(&A.stats[0], (unsigned long long)nrow);
    atomicAdd(&A.stats[1], (unsigned long long)nfb);
  }
}

// Features of one routed tile (pairs t0 .. t0 + nrow - 1, from the endpoint
// table) by TT * tpp threads: thread j works on pair j / tpp, decoding it
// itself, and gathers its lattice points uv = j % tpp, + tpp, ... < 25, all
// ZT layers of each from the per-GCELL layer copy (feature 4 + l*25 + uv,
// uv = 5u + v for lattice column u and row v).  No thread reads what
// another wrote, so no barrier separates the decode from the gather.  The
// first thread of each pair writes its coordinates (sc) and the row's
// non-lattice features.
// The ZT layer values of one GCELL of the per-GCELL layer copy (16 int16
// per cell, 32-byte aligned) as 32-bit words w[l / 2], loading only the
// words that hold layers below ZT.
template <int ZT>
__device__ __forceinline__ void load_remz(const int16_t* cell, uint32_t (&w)[8]) {
  const uint4 a = *reinterpret_cast<const uint4*>(cell);
  w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w;
  w[4] = w[5] = w[6] = w[7] = 0u;
  if constexpr (ZT > 12) {
    const uint4 b = *reinterpret_cast<const uint4*>(cell + 8);
    w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
  } else if constexpr (ZT > 10) {
    const uint2 b = *reinterpret_cast<const uint2*>(cell + 8);
    w[4] = b.x; w[5] = b.y;
  } else if constexpr (ZT > 8) {
    w[4] = *reinterpret_cast<const uint32_t*>(cell + 8);
  }
}

// Features of one routed tile (pairs t0 .. t0 + nrow - 1, from the endpoint
// table) by TT * TPP threads: thread j works on pair j / TPP, decoding it
// itself, and gathers its lattice points uv = j % TPP, + TPP, ... < 25 (all
// of them in flight at once), all ZT layers of each from the per-GCELL layer
// copy (feature 4 + l*25 + uv, uv = 5u + v for lattice column u and row v).
// No thread reads what another wrote, so no barrier separates the decode
// from the gather.  The first thread of each pair writes its coordinates
// (sc) and the row's non-lattice features.
template <int ZT, int TT, int TPP>
__device__ __forceinline__ void prop_tile_features(const PropArgs& A, int64_t t0, int nrow, int j, uint16_t* act,
                                                   const uint16_t* slut, int16_t (*sc)[6]) {
  using S = PropShape<ZT>;
  constexpr int RS = S::RS, KIN = S::KIN, NFEAT = S::NFEAT, UG = (25 + TPP - 1) / TPP;
  const int pr = j / TPP, h = j - pr * TPP;
  const uint2 e = A.ends[t0 + (pr < nrow ? pr : nrow - 1)];
  int z1, x1, y1, z2, x2, y2;
  unpack_cell(e.x, z1, x1, y1);
  unpack_cell(e.y, z2, x2, y2);
  const int dx = x2 - x1, dy = y2 - y1;
  uint16_t* row = act + pr * RS;
  if (h == 0) {
    sc[pr][0] = (int16_t)z1; sc[pr][1] = (int16_t)x1; sc[pr][2] = (int16_t)y1;
    sc[pr][3] = (int16_t)z2; sc[pr][4] = (int16_t)x2; sc[pr][5] = (int16_t)y2;
    row[0] = f2bf((float)dx * (1.0f / 64.0f));
    row[1] = f2bf((float)dy * (1.0f / 64.0f));
    row[2] = f2bf((float)z1);
    row[3] = f2bf((float)z2);
    for (int f = 4 + NFEAT; f < KIN; ++f) row[f] = 0;
  }
  uint32_t w[UG][8];
#pragma unroll
  for (int u = 0; u < UG; ++u) {
    const int uv = h + TPP * u < 25 ? h + TPP * u : 24, iu = uv / 5, iv = uv - 5 * iu;
    const int cx = x1 + (dx * iu + (dx >= 0 ? 2 : -2)) / 4, cy = y1 + (dy * iv + (dy >= 0 ? 2 : -2)) / 4;
    load_remz<ZT>(A.remz + ((size_t)cy * (size_t)A.g.X + (size_t)cx) * 16, w[u]);
  }
#pragma unroll
  for (int u = 0; u < UG; ++u) {
    const int uv = h + TPP * u;
    if (uv >= 25) continue;
#pragma unroll
    for (int l = 0; l < ZT; ++l) {
      const int v = (int)(int16_t)(uint16_t)(w[u][l >> 1] >> (16 * (l & 1)));
      row[4 + uv + 25 * l] = slut[v - kRMin];
    }
  }
}

// The routing kernel as a persistent loop over tiles t, t + grid, ... of a
// batch with the endpoint table: while the first NC = TT/32 waves cost tile
// t's candidates (pair half x parity), the other NC decode the next tile's
// pairs and gather its features into the activation buffer, which the
// costing does not touch (its keys have their own LDS array, the coordinates
// two buffers).  Per tile the phases are the layers and the top-k (all
// waves), then costing beside the next tile's features, then the results;
// each tile's arithmetic is k_prop's.  TT = 64 is used: four 4-wave
// workgroups per CU.
template <int ZT, int TT, int NW = TT / 16>
__global__ __launch_bounds__(64 * NW, 4) void k_prop_tiles(PropArgs A) {
  constexpr int NT = 64 * NW;
  constexpr int NFH = 16 / NW;
  using S = PropShape<ZT>;
  constexpr int RS = S::RS, KIN = S::KIN;
  constexpr int NC = TT / 32;  // costing wavefronts: (pair half, parity) each
  static_assert(NW == TT / 16 && 2 * NC == NW, "as many costing as feature wavefronts");
  __shared__ __attribute__((aligned(16))) uint16_t act[TT * RS];
  __shared__ uint16_t slut[kLut];                  // log-cost table as bf16
  __shared__ int16_t sc[2][TT][6];                 // z1 x1 y1 z2 x2 y2, this tile's and the next's
  __shared__ unsigned long long smask[TT];         // candidate patterns q = p - Z2
  __shared__ int8_t sfb[TT];                       // low confidence: enumerate every pattern
  __shared__ unsigned long long skey[2 * TT];      // the two parity keys of each pair
  __shared__ int nfb;
  const int tid = threadIdx.x, wv0 = tid >> 6, lane0 = tid & 63;
  const int64_t ntile = (A.P + TT - 1) / TT;
  if (blockIdx.x == 0 && tid == 0) *A.ctr += 1;  // this batch's stamp
  if (tid == 0) nfb = 0;
  for (int i = tid; i < kLut; i += NT) slut[i] = f2bf(A.loglut[i]);
  __syncthreads();
  int64_t t = blockIdx.x;
  int cur = 0;
  int64_t rows = 0;
  uint4 a0[NFH];
  mma_first<KIN / 32, NFH>(A.W1p + (size_t)wv0 * NFH * (KIN / 32) * 512, a0, lane0);
  prop_tile_features<ZT, TT, NT / TT>(A, t * TT, (int)(A.P - t * TT < TT ? A.P - t * TT : TT), tid, act, slut, sc[0]);
  __syncthreads();
  for (;;) {
    // the thread index as opaque per iteration: the per-lane LDS addresses
    // of the layers are then formed inside the loop, not held across it
    int tl = tid;
    asm volatile("" : "+v"(tl));
    const int wv = __builtin_amdgcn_readfirstlane(tl >> 6), lane = tl & 63;  // wv: a scalar, so the phase branch is uniform
    const int64_t t0 = t * TT, tn = t + gridDim.x;
    const int nrow = (int)(A.P - t0 < TT ? A.P - t0 : TT);
    rows += nrow;
    prop_hidden<ZT, TT, NW>(A, act, a0, wv, lane);
    prop_out_topk<ZT, TT, NW>(A, act, smask, sfb, wv, lane);
    if (wv < NC) {
      prop_cost_pair<ZT, TT>(A, nrow, 64 * (wv >> 1) + lane, wv & 1, sc[cur], smask, sfb, skey);
    } else if (tn < ntile) {
      const int64_t n0 = tn * TT;
      prop_tile_features<ZT, TT, 2>(A, n0, (int)(A.P - n0 < TT ? A.P - n0 : TT), tl - NC * 64, act, slut, sc[cur ^ 1]);
    }
    __syncthreads();
    if (tl < nrow && prop_finish_pair<ZT, TT>(A, t0, tl, sc[cur], sfb, skey)) atomicAdd(&nfb, 1);
    if (tn >= ntile) break;
    mma_first<KIN / 32, NFH>(A.W1p + (size_t)wv * NFH * (KIN / 32) * 512, a0, lane);  // the next tile's
    t = tn;
    cur ^= 1;
  }
  __syncthreads();
  if (tid == 0) {
    atomicAdd(&A.stats[0], (unsigned long long)rows);
    atomicAdd(&A.stats[1], (unsigned long long)nfb);
  }
}

typedef void (*PropFn)(PropArgs);
template <bool ROUTE, int TT>
PropFn prop_kernel_t(int Z) {
  switch (Z) {
    case 3: return k_prop<3, ROUTE, TT>;
    case 4: return k_prop<4, ROUTE, TT>;
    case 5: return k_prop<5, ROUTE, TT>;
    case 6: return k_prop<6, ROUTE, TT>;
    case 7: return k_prop<7, ROUTE, TT>;
    case 8: return k_prop<8, ROUTE, TT>;
    case 9: return k_prop<9, ROUTE, TT>;
    case 10: return k_prop<10, ROUTE, TT>;
    case 11: return k_prop<11, ROUTE, TT>;  // 60 candidate patterns; Z = 12 has 66 (unfused path)
    default: return nullptr;
  }
}
template <bool ROUTE>
PropFn prop_kernel(int Z) { return prop_kernel_t<ROUTE, kTile>(Z); }

constexpr int kTilesTT = 64;  // pairs per tile of the persistent routing kernel (4 wavefronts)
PropFn prop_tiles_kernel(int Z) {
  switch (Z) {
    case 3: return k_prop_tiles<3, kTilesTT>;
    case 4: return k_prop_tiles<4, kTilesTT>;
    case 5: return k_prop_tiles<5, kTilesTT>;
    case 6: return k_prop_tiles<6, kTilesTT>;
    case 7: return k_prop_tiles<7, kTilesTT>;
    case 8: return k_prop_tiles<8, kTilesTT>;
    case 9: return k_prop_tiles<9, kTilesTT>;
    case 10: return k_prop_tiles<10, kTilesTT>;
    case 11: return k_prop_tiles<11, kTilesTT>;
    default: return nullptr;
  }
}

int blocks_for(int64_t n, int per) { return (int)((n + per - 1) / per); }

}  // namespace

struct gr_mlp {
  gr_ctx* ctx = nullptr;
  int n_in = 0, k_in = 0, H = 0, n_out = 0, n_op = 0;  // n_op: n_out padded to 64
  int n_params = 0;
  size_t oW1, ob1, oW2, ob2, oW3, ob3;
  float *d_master = nullptr, *d_mom = nullptr, *d_grad = nullptr, *d_loglut = nullptr, *d_loss = nullptr;
  uint16_t *W1b = nullptr, *W1t = nullptr, *W2b = nullptr, *W2t = nullptr, *W3b = nullptr, *W3t = nullptr;
  // fused inference (hidden width kHid): weights in MFMA fragment order
  uint16_t *W1p = nullptr, *W2p = nullptr, *W3p = nullptr;
  int nf3 = 0;  // packed output fragments of 16 neurons (n_op / 16)
  int tiles_grid = 0;  // workgroups of the persistent routing kernel: as many as are resident at once
  int64_t cap_samples = 0, n_samples = 0;
  uint16_t* d_X = nullptr;
  int32_t* d_y = nullptr;
  int64_t act_cap = 0;  // rows of the inference / training activation buffers
  uint16_t *a_X = nullptr, *a_H1 = nullptr, *a_H2 = nullptr, *a_dL = nullptr, *a_dZ2 = nullptr, *a_dZ1 = nullptr;
  float* a_L = nullptr;
  int32_t *a_y = nullptr, *d_perm = nullptr;
  float* a_part = nullptr;  // split-K partial gradients [kSplit][largest layer's (M + 1) x N]
};

namespace {

int mlp_reserve(gr_mlp* m, int64_t rows) {
  if (rows <= m->act_cap) return 0;
  gr_ctx* c = m->ctx;
  int rc;
  if ((rc = dalloc(c, &m->a_X, rows * m->k_in)) || (rc = dalloc(c, &m->a_H1, rows * m->H)) ||
      (rc = dalloc(c, &m->a_H2, rows * m->H)) || (rc = dalloc(c, &m->a_L, rows * m->n_op)) ||
      (rc = dalloc(c, &m->a_dL, rows * m->n_op)) || (rc = dalloc(c, &m->a_dZ2, rows * m->H)) ||
      (rc = dalloc(c, &m->a_dZ1, rows * m->H)) || (rc = dalloc(c, &m->a_y, rows)) || (rc = dalloc(c, &m->d_perm, rows)) ||
      (rc = dalloc(c, &m->a_part, (size_t)kSplit * (std::max(m->k_in, m->H) + 1) * std::max(m->H, m->n_op))))
    return rc;
  m->act_cap = rows;
  return 0;
}

template <bool RELU, bool OUT_F32>
void gemm(hipStream_t st, int M, int N, int K, const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B,
          int64_t sbk, int64_t sbn, const float* bias, const uint16_t* mask, int64_t ldm, uint16_t* Cb, float* Cf,
          int64_t ldc) {
  hipLaunchKernelGGL((k_gemm<RELU, OUT_F32>), dim3(blocks_for(M, 64), blocks_for(N, 64)), dim3(256), 0, st, M, N, K,
                     A, sam, sak, B, sbk, sbn, bias, mask, ldm, Cb, Cf, ldc);
}

// bf16 copies of the master weights for the training GEMMs; with pack, also
// the fragment-order copies the fused kernel reads (routing, forward).
void refresh_bf16(gr_mlp* m, hipStream_t st, bool pack = true) {
  const float* w = m->d_master;
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for((int64_t)m->k_in * m->H, 256)), dim3(256), 0, st, m->k_in, m->H,
                     w + m->oW1, m->W1b, m->W1t);
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for((int64_t)m->H * m->H, 256)), dim3(256), 0, st, m->H, m->H,
                     w + m->oW2, m->W2b, m->W2t);
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for((int64_t)m->H * m->n_op, 256)), dim3(256), 0, st, m->H, m->n_op,
                     w + m->oW3, m->W3b, m->W3t);
  if (m->W1p && pack) {
    hipLaunchKernelGGL(k_pack, dim3(blocks_for((int64_t)m->k_in * m->H, 256)), dim3(256), 0, st, m->k_in, m->H,
                       m->H / 16, w + m->oW1, m->W1p);
    hipLaunchKernelGGL(k_pack, dim3(blocks_for((int64_t)m->H * m->H, 256)), dim3(256), 0, st, m->H, m->H, m->H / 16,
                       w + m->oW2, m->W2p);
    hipLaunchKernelGGL(k_pack, dim3(blocks_for((int64_t)m->H * m->nf3 * 16, 256)), dim3(256), 0, st, m->H, m->n_op,
                       m->nf3, w + m->oW3, m->W3p);
  }
}

// Arguments of the fused kernel common to routing and the forward pass.
PropArgs prop_args(gr_mlp* m) {
  gr_ctx* c = m->ctx;
  PropArgs A{};
  A.g = c->g;
  A.W1p = m->W1p;
  A.W2p = m->W2p;
  A.W3p = m->W3p;
  A.b1 = m->d_master + m->ob1;
  A.b2 = m->d_master + m->ob2;
  A.b3 = m->d_master + m->ob3;
  A.n_op = m->n_op;
  return A;
}

// Forward pass of rows x a_X -> a_H1, a_H2, a_L (fp32 logits).
void forward(gr_mlp* m, hipStream_t st, int64_t rows, const uint16_t* X) {
  const float* w = m->d_master;
  gemm<true, false>(st, (int)rows, m->H, m->k_in, X, m->k_in, 1, m->W1t, 1, m->k_in, w + m->ob1, nullptr, 0, m->a_H1,
                    nullptr, m->H);
  gemm<true, false>(st, (int)rows, m->H, m->H, m->a_H1, m->H, 1, m->W2t, 1, m->H, w + m->ob2, nullptr, 0, m->a_H2,
                    nullptr, m->H);
  gemm<false, true>(st, (int)rows, m->n_op, m->H, m->a_H2, m->H, 1, m->W3t, 1, m->H, w + m->ob3, nullptr, 0, nullptr,
                    m->a_L, m->n_op);
}

// [dW; db] of one layer into g (the bias follows its weight in the parameter
// layout, as row M of an (M + 1) x N matrix): split-K over the rows, then the
// parts summed in order.
void grad_tn(gr_mlp* m, hipStream_t st, int rows, int M, int N, const uint16_t* A, const uint16_t* B, float* g) {
  int nz = kSplit;
  while (nz > 1 && (rows % (32 * nz) || rows / nz < 64)) nz >>= 1;
  hipLaunchKernelGGL(k_gemm_tn, dim3(blocks_for(M + 1, 64), blocks_for(N, 64), nz), dim3(256), 0, st, M, N, rows / nz,
                     A, (int64_t)M, B, (int64_t)N, nz > 1 ? m->a_part : g, (int64_t)N);
  if (nz > 1)
    hipLaunchKernelGGL(k_sum_parts, dim3(blocks_for((int64_t)(M + 1) * N, 256)), dim3(256), 0, st,
                       (int64_t)(M + 1) * N, nz, m->a_part, g);
}

// Loss and gradient of rows (multiple of 32) in a_X / a_y (padding labels -1);
// gradient into d_grad, loss sum into d_loss.
void backward(gr_mlp* m, hipStream_t st, int rows, int n_real) {
  float* gr = m->d_grad;
  hipLaunchKernelGGL(k_softmax_ce, dim3(blocks_for(rows, 4)), dim3(256), 0, st, rows, m->n_out, m->n_op, m->a_L,
                     m->a_y, 1.0f / (float)n_real, m->a_dL, m->d_loss);
  // layer 3: [dW3; db3] = [H2 | 1]^T dL, dZ2 = (dL W3^T) * [H2 > 0]
  grad_tn(m, st, rows, m->H, m->n_op, m->a_H2, m->a_dL, gr + m->oW3);
  gemm<false, false>(st, rows, m->H, m->n_op, m->a_dL, m->n_op, 1, m->W3b, 1, m->n_op, nullptr, m->a_H2, m->H,
                     m->a_dZ2, nullptr, m->H);
  // layer 2
  grad_tn(m, st, rows, m->H, m->H, m->a_H1, m->a_dZ2, gr + m->oW2);
  gemm<false, false>(st, rows, m->H, m->H, m->a_dZ2, m->H, 1, m->W2b, 1, m->H, nullptr, m->a_H1, m->H, m->a_dZ1,
                     nullptr, m->H);
  // layer 1
  grad_tn(m, st, rows, m->k_in, m->H, m->a_X, m->a_dZ1, gr + m->oW1);
}

}  // namespace

namespace grh {

bool mlp_fused(const gr_mlp* m) { return m->W1p != nullptr; }

void mlp_detach_all(gr_ctx* c) {
  for (gr_mlp* m : c->models) m->ctx = nullptr;
  c->models.clear();
  c->prop = c->collect = nullptr;
  c->remz_live = false;
  c->vterm_live = false;
}

// Exhaustive costing already ran; capture a strided sample of the batch's
// pairs with their chosen patterns as labels.
int mlp_enqueue_collect(gr_ctx* c, Sched& s, const BatchMeta& bm) {
  gr_mlp* m = c->collect;
  const int64_t room = m->cap_samples - m->n_samples;
  if (room <= 0 || bm.P <= 0) return 0;
  int64_t want = std::min<int64_t>({(int64_t)c->collect_per_batch, bm.P, room});
  const int64_t stride = std::max<int64_t>(1, bm.P / want);
  want = std::min<int64_t>(want, (bm.P + stride - 1) / stride);
  hipLaunchKernelGGL(k_feat_batch, dim3(blocks_for(want, 256)), dim3(256), 0, c->stream, c->g, c->d_rem, m->d_loglut,
                     c->d_pins, s.d_pinofs + bm.begin, s.d_pair + bm.begin, bm.n, want, stride,
                     m->d_X + m->n_samples * m->k_in, m->k_in, (const int16_t*)c->d_pair_pat, m->d_y + m->n_samples,
                     c->prop ? (const uint8_t*)c->d_pair_fb : nullptr);
  m->n_samples += want;
  HIPC(c, hipGetLastError());
  return 0;
}

// Learned costing of one batch: features -> MLP -> top-k exact costing.
int mlp_enqueue_costing(gr_ctx* c, Sched& s, const BatchMeta& bm) {
  gr_mlp* m = c->prop;
  if (m->W1p) {
    PropArgs A = prop_args(m);
    A.rem = c->d_rem;
    A.remz = c->d_remz;
    A.loglut = m->d_loglut;
    A.pfx = c->d_pfx;
    A.vterm = c->vterm_live ? c->d_vterm : nullptr;
    A.pins = c->d_pins;
    A.pinofs = s.d_pinofs + bm.begin;
    A.bpair = s.d_pair + bm.begin;
    A.chunk_net = s.d_chunk + bm.chunk_ofs;
    A.ends = s.ends_ok ? s.d_ends + bm.pair_base : nullptr;
    A.n = bm.n;
    A.P = bm.P;
    A.topk = c->prop_k;
    A.thr = c->prop_thr;
    A.out_cost = c->d_pair_cost;
    A.out_pat = c->d_pair_pat;
    A.ctr = c->d_ctr;
    A.stats = (unsigned long long*)c->d_prop_stats;
    A.fbout = c->collect ? c->d_pair_fb : nullptr;
    if (A.ends && m->tiles_grid > 0)
      hipLaunchKernelGGL(prop_tiles_kernel(c->g.Z), dim3(std::min(blocks_for(bm.P, kTilesTT), m->tiles_grid)),
                         dim3(kTilesTT * 4), 0, c->stream, A);
    else
      hipLaunchKernelGGL(prop_kernel<true>(c->g.Z), dim3(blocks_for(bm.P, kTile)), dim3(kTile * 4), 0, c->stream, A);
    HIPC(c, hipGetLastError());
    return 0;
  }
  int rc;
  if ((rc = mlp_reserve(m, bm.P))) return rc;
  hipLaunchKernelGGL(k_feat_batch, dim3(blocks_for(bm.P, 256)), dim3(256), 0, c->stream, c->g, c->d_rem, m->d_loglut,
                     c->d_pins, s.d_pinofs + bm.begin, s.d_pair + bm.begin, bm.n, bm.P, (int64_t)1, m->a_X, m->k_in,
                     (const int16_t*)nullptr, (int32_t*)nullptr, (const uint8_t*)nullptr);
  forward(m, c->stream, bm.P, m->a_X);
  hipLaunchKernelGGL(k_lpat_topk, dim3(blocks_for(bm.P, 256)), dim3(256), 0, c->stream, c->g, c->d_pfx, c->d_pins,
                     s.d_pinofs + bm.begin, s.d_pair + bm.begin, s.d_chunk + bm.chunk_ofs, bm.n, bm.P, m->a_L, m->n_op,
                     m->n_out, c->prop_k, c->prop_thr, c->d_pair_cost, c->d_pair_pat, c->d_ctr,
                     (unsigned long long*)c->d_prop_stats);
  HIPC(c, hipGetLastError());
  return 0;
}

}  // namespace grh

extern "C" {

int gr_mlp_create(gr_ctx* c, int32_t hidden, uint64_t seed, int32_t max_samples, gr_mlp** out) {
  if (!c || !out || hidden < 32 || hidden % 32 || max_samples < 0) return fail(c, GR_EINVAL, "bad model shape");
  *out = nullptr;
  if (c->d_remf) return fail(c, GR_ESTATE, "the learned proposal needs a grid with integer capacities");
  if ((c->g.Z - 1) * c->g.Z2 > 128) return fail(c, GR_EINVAL, "the learned proposal scores at most 128 patterns (Z <= 16)");
  if (c->g.X > 32767 || c->g.Y > 32767)  // pair coordinates and lattice points are int16 in LDS
    return fail(c, GR_EINVAL, "the learned proposal supports grids of at most 32767 GCELLs per side");
  (void)hipSetDevice(c->dev);
  gr_mlp* m = new gr_mlp();
  m->ctx = c;
  m->n_in = 4 + 25 * c->g.Z;
  m->k_in = (m->n_in + 31) / 32 * 32;
  m->H = hidden;
  m->n_out = (c->g.Z - 1) * c->g.Z2;
  m->n_op = (m->n_out + 63) / 64 * 64;
  size_t o = 0;
  m->oW1 = o; o += (size_t)m->k_in * m->H;
  m->ob1 = o; o += m->H;
  m->oW2 = o; o += (size_t)m->H * m->H;
  m->ob2 = o; o += m->H;
  m->oW3 = o; o += (size_t)m->H * m->n_op;
  m->ob3 = o; o += m->n_op;
  m->n_params = (int)o;
  // He-uniform initialisation of the weights (fixed seed), zero biases; padded
  // input rows and output columns are zero
  std::vector<float> w(o, 0.f);
  std::mt19937_64 rng(seed);
  auto init = [&](size_t off, int K, int N, int Kreal, int Nreal) {
    const float a = std::sqrt(6.0f / (float)Kreal);
    std::uniform_real_distribution<float> u(-a, a);
    for (int k = 0; k < K; ++k)
      for (int n = 0; n < N; ++n) w[off + (size_t)k * N + n] = (k < Kreal && n < Nreal) ? u(rng) : 0.f;
  };
  init(m->oW1, m->k_in, m->H, m->n_in, m->H);
  init(m->oW2, m->H, m->H, m->H, m->H);
  init(m->oW3, m->H, m->n_op, m->H, m->n_out);
  std::vector<float> loglut(kLut);
  for (int i = 0; i < kLut; ++i) {
    const float v = std::log(std::max(c->lut[i], 1e-30f));
    loglut[i] = std::min(std::max(v, -40.f), 40.f) * 0.25f;
  }
  m->cap_samples = max_samples;
  int rc;
  auto bail = [&](int code) { gr_mlp_destroy(m); return code; };
  if ((rc = dalloc(c, &m->d_master, o)) || (rc = dalloc(c, &m->d_mom, o)) || (rc = dalloc(c, &m->d_grad, o)) ||
      (rc = dalloc(c, &m->d_loglut, kLut)) || (rc = dalloc(c, &m->d_loss, 1)) ||
      (rc = dalloc(c, &m->W1b, (size_t)m->k_in * m->H)) || (rc = dalloc(c, &m->W1t, (size_t)m->k_in * m->H)) ||
      (rc = dalloc(c, &m->W2b, (size_t)m->H * m->H)) || (rc = dalloc(c, &m->W2t, (size_t)m->H * m->H)) ||
      (rc = dalloc(c, &m->W3b, (size_t)m->H * m->n_op)) || (rc = dalloc(c, &m->W3t, (size_t)m->H * m->n_op)) ||
      (rc = dalloc(c, &m->d_X, (size_t)std::max<int64_t>(1, max_samples) * m->k_in)) ||
      (rc = dalloc(c, &m->d_y, (size_t)std::max<int64_t>(1, max_samples))))
    return bail(rc);
  if (hipMemcpy(m->d_master, w.data(), o * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->d_mom, 0, o * 4) != hipSuccess ||
      hipMemcpy(m->d_loglut, loglut.data(), kLut * 4, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(c, GR_ENODEV, "model upload failed"));
  if (m->H == kHid && prop_kernel<true>(c->g.Z)) {
    m->nf3 = m->n_op / 16;  // every padded output column (the forward pass returns them all)
    int ncu = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, prop_tiles_kernel(c->g.Z), kTilesTT * 4, 0) != hipSuccess)
      return bail(fail(c, GR_ENODEV, "device query for the routing kernel failed"));
    m->tiles_grid = ncu * per_cu;
    if ((rc = dalloc(c, &m->W1p, (size_t)m->k_in * m->H)) || (rc = dalloc(c, &m->W2p, (size_t)m->H * m->H)) ||
        (rc = dalloc(c, &m->W3p, (size_t)m->H * m->n_op)))
      return bail(rc);
  }
  refresh_bf16(m, c->stream);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(fail(c, GR_ENODEV, "model set-up failed"));
  c->models.push_back(m);
  *out = m;
  return 0;
}

void gr_mlp_destroy(gr_mlp* m) {
  if (!m) return;
  gr_ctx* c = m->ctx;
  if (c) {
    (void)hipSetDevice(c->dev);
    (void)hipStreamSynchronize(c->stream);
    if (c->prop == m) {
      c->prop = nullptr;
      c->remz_live = false;
      c->vterm_live = false;
      drop_graph(c);
    }
    if (c->collect == m) c->collect = nullptr;
    c->models.erase(std::remove(c->models.begin(), c->models.end(), m), c->models.end());
  }
  void* ps[] = {m->d_master, m->d_mom, m->d_grad, m->d_loglut, m->d_loss, m->W1b, m->W1t, m->W2b, m->W2t, m->W3b,
                m->W3t, m->d_X, m->d_y, m->a_X, m->a_H1, m->a_H2, m->a_L, m->a_dL, m->a_dZ2, m->a_dZ1, m->a_y,
                m->d_perm, m->W1p, m->W2p, m->W3p, m->a_part};
  for (void* p : ps) (void)hipFree(p);
  delete m;
}

int gr_mlp_shape(const gr_mlp* m, int32_t* s) {
  if (!m || !s) return GR_EINVAL;
  s[0] = m->n_in; s[1] = m->k_in; s[2] = m->H; s[3] = m->n_out; s[4] = m->n_op; s[5] = m->n_params;
  return 0;
}

int gr_mlp_get_params(gr_mlp* m, float* out) {
  if (!m || !out) return GR_EINVAL;
  if (!m->ctx) return GR_ESTATE;
  gr_ctx* c = m->ctx;
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipMemcpy(out, m->d_master, (size_t)m->n_params * 4, hipMemcpyDeviceToHost));
  return 0;
}

int gr_mlp_set_params(gr_mlp* m, const float* in) {
  if (!m || !in) return GR_EINVAL;
  if (!m->ctx) return GR_ESTATE;
  gr_ctx* c = m->ctx;
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipMemcpy(m->d_master, in, (size_t)m->n_params * 4, hipMemcpyHostToDevice));
  refresh_bf16(m, c->stream);
  HIPC(c, hipStreamSynchronize(c->stream));
  return 0;
}

int gr_mlp_collect(gr_ctx* c, gr_mlp* m, int32_t per_batch) {
  if (!c || (per_batch > 0 && (!m || m->ctx != c)) || per_batch < 0)
    return fail(c, GR_EINVAL, "bad collect request (the model must belong to this context)");
  if (per_batch > 0 && c->prop && !mlp_fused(c->prop))
    return fail(c, GR_ESTATE, "capture under a proposal needs the fused kernel's fallback flags (hidden width 256)");
  c->collect = per_batch > 0 ? m : nullptr;
  c->collect_per_batch = per_batch;
  return 0;
}

int64_t gr_mlp_samples(const gr_mlp* m) { return m ? m->n_samples : 0; }

int gr_mlp_clear_samples(gr_mlp* m) {
  if (!m) return GR_EINVAL;
  m->n_samples = 0;
  return 0;
}

int gr_mlp_train(gr_mlp* m, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* loss) {
  if (!m || epochs < 0 || batch < 32 || batch % 32) return GR_EINVAL;
  if (!m->ctx) return GR_ESTATE;
  gr_ctx* c = m->ctx;
  (void)hipSetDevice(c->dev);
  int rc;
  if ((rc = mlp_reserve(m, batch))) return rc;
  // samples labelled -1 (capture under a proposal, choice not from every
  // pattern) are left out
  std::vector<int32_t> lab(m->n_samples), keep;
  HIPC(c, hipStreamSynchronize(c->stream));
  if (m->n_samples) HIPC(c, hipMemcpy(lab.data(), m->d_y, 4 * lab.size(), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < m->n_samples; ++i)
    if (lab[i] >= 0) keep.push_back((int32_t)i);
  const int64_t S = (int64_t)keep.size();
  if (S < batch) return fail(c, GR_ESTATE, "fewer labelled samples than one minibatch");
  std::mt19937_64 rng(seed);
  std::vector<int32_t> perm(S);
  float last = 0.f;
  const int steps = (int)(S / batch);
  for (int e = 0; e < epochs; ++e) {
    perm = keep;
    std::shuffle(perm.begin(), perm.end(), rng);
    HIPC(c, hipMemsetAsync(m->d_loss, 0, 4, c->stream));
    for (int st = 0; st < steps; ++st) {
      HIPC(c, hipMemcpyAsync(m->d_perm, perm.data() + (int64_t)st * batch, sizeof(int32_t) * batch,
                             hipMemcpyHostToDevice, c->stream));
      hipLaunchKernelGGL(k_gather_rows, dim3(batch), dim3(256), 0, c->stream, batch, m->k_in, m->d_perm, m->d_X,
                         m->d_y, m->a_X, m->a_y);
      forward(m, c->stream, batch, m->a_X);
      backward(m, c->stream, batch, batch);
      hipLaunchKernelGGL(k_sgd, dim3(blocks_for(m->n_params, 256)), dim3(256), 0, c->stream, m->n_params,
                         m->d_master, m->d_mom, m->d_grad, lr);
      refresh_bf16(m, c->stream, false);
      HIPC(c, hipStreamSynchronize(c->stream));  // perm chunk is a host buffer
    }
    float ls = 0.f;
    HIPC(c, hipMemcpy(&ls, m->d_loss, 4, hipMemcpyDeviceToHost));
    last = ls / (float)((int64_t)steps * batch);
  }
  refresh_bf16(m, c->stream);  // fragment-order copies for routing
  HIPC(c, hipStreamSynchronize(c->stream));
  if (loss) *loss = last;
  return 0;
}

int gr_mlp_features(gr_ctx* c, gr_mlp* m, const int32_t* src, const int32_t* dst, int32_t n, float* out) {
  if (!c || !m || m->ctx != c || n < 0 || (n && (!src || !dst || !out))) return fail(c, GR_EINVAL, "bad feature request");
  if (!n) return 0;
  (void)hipSetDevice(c->dev);
  int32_t *ds = nullptr, *dd = nullptr;
  uint16_t* X = nullptr;
  int rc;
  if ((rc = dalloc(c, &ds, n)) || (rc = dalloc(c, &dd, n)) || (rc = dalloc(c, &X, (size_t)n * m->k_in))) {
    (void)hipFree(ds); (void)hipFree(dd); (void)hipFree(X);
    return rc;
  }
  std::vector<uint16_t> h((size_t)n * m->k_in);
  bool ok = hipMemcpy(ds, src, 4 * (size_t)n, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(dd, dst, 4 * (size_t)n, hipMemcpyHostToDevice) == hipSuccess &&
            hipStreamSynchronize(c->stream) == hipSuccess;
  if (ok) {
    hipLaunchKernelGGL(k_feat_pairs, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, c->g, c->d_rem, m->d_loglut,
                       ds, dd, n, X, m->k_in);
    ok = hipStreamSynchronize(c->stream) == hipSuccess &&
         hipMemcpy(h.data(), X, h.size() * 2, hipMemcpyDeviceToHost) == hipSuccess;
  }
  (void)hipFree(ds); (void)hipFree(dd); (void)hipFree(X);
  if (!ok) return fail(c, GR_ENODEV, "feature kernel failed");
  for (size_t i = 0; i < h.size(); ++i) out[i] = bf2f(h[i]);
  return 0;
}

int gr_mlp_forward(gr_mlp* m, const float* x, int32_t n, float* logits) {
  if (!m || n < 0 || (n && (!x || !logits))) return GR_EINVAL;
  if (!m->ctx) return GR_ESTATE;
  if (!n) return 0;
  gr_ctx* c = m->ctx;
  (void)hipSetDevice(c->dev);
  int rc;
  if ((rc = mlp_reserve(m, n))) return rc;
  std::vector<uint16_t> h((size_t)n * m->k_in);
  for (size_t i = 0; i < h.size(); ++i) h[i] = f2bf(x[i]);
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipMemcpy(m->a_X, h.data(), h.size() * 2, hipMemcpyHostToDevice));
  if (m->W1p) {  // the fused kernel routing uses, so its logits are the ones routing ranks
    PropArgs A = prop_args(m);
    A.X = m->a_X;
    A.P = n;
    A.logits = m->a_L;
    hipLaunchKernelGGL(prop_kernel<false>(c->g.Z), dim3(blocks_for(n, kTile)), dim3(kTile * 4), 0, c->stream, A);
    HIPC(c, hipGetLastError());
  } else {
    forward(m, c->stream, n, m->a_X);
  }
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipMemcpy(logits, m->a_L, (size_t)n * m->n_op * 4, hipMemcpyDeviceToHost));
  return 0;
}

int gr_mlp_gradient(gr_mlp* m, const float* x, const int32_t* y, int32_t n, float* grad, float* loss) {
  if (!m || n <= 0 || !x || !y || !grad) return GR_EINVAL;
  if (!m->ctx) return GR_ESTATE;
  gr_ctx* c = m->ctx;
  (void)hipSetDevice(c->dev);
  const int rows = (n + 31) / 32 * 32;
  int rc;
  if ((rc = mlp_reserve(m, rows))) return rc;
  std::vector<uint16_t> h((size_t)rows * m->k_in, 0);
  for (size_t i = 0; i < (size_t)n * m->k_in; ++i) h[i] = f2bf(x[i]);
  std::vector<int32_t> yy(rows, -1);
  for (int i = 0; i < n; ++i) yy[i] = y[i];
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipMemcpy(m->a_X, h.data(), h.size() * 2, hipMemcpyHostToDevice));
  HIPC(c, hipMemcpy(m->a_y, yy.data(), 4 * (size_t)rows, hipMemcpyHostToDevice));
  HIPC(c, hipMemset(m->d_loss, 0, 4));
  forward(m, c->stream, rows, m->a_X);
  backward(m, c->stream, rows, n);
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipMemcpy(grad, m->d_grad, (size_t)m->n_params * 4, hipMemcpyDeviceToHost));
  float ls = 0.f;
  HIPC(c, hipMemcpy(&ls, m->d_loss, 4, hipMemcpyDeviceToHost));
  if (loss) *loss = ls / (float)n;
  return 0;
}

int gr_set_proposal(gr_ctx* c, gr_mlp* m, int32_t k, float threshold) {
  if (!c) return fail(c, GR_EINVAL, "null context");
  if (m && (k < 1 || k > kMaxTopK || m->ctx != c)) return fail(c, GR_EINVAL, "k must be 1..8 and the model this context's");
  if (m && c->collect && !mlp_fused(m))
    return fail(c, GR_ESTATE, "capture under a proposal needs the fused kernel's fallback flags (hidden width 256)");
  (void)hipSetDevice(c->dev);
  if (!c->d_prop_stats) {
    int rc;
    if ((rc = dalloc(c, &c->d_prop_stats, 2))) return rc;
    HIPC(c, hipMemset(c->d_prop_stats, 0, 16));
  }
  HIPC(c, hipStreamSynchronize(c->stream));
  drop_graph(c);  // the captured pass holds the old costing kernel and its arguments
  c->remz_live = false;
  c->vterm_live = false;
  if (m && mlp_fused(m)) {  // the lattice gather's per-GCELL layer copy, from here on kept by the refresh
    int rc;
    if (!c->d_remz && ((rc = dalloc(c, &c->d_remz, (size_t)16 * c->g.XY)) ||
                       (rc = dalloc(c, &c->d_remz0, (size_t)16 * c->g.XY))))
      return rc;
    hipLaunchKernelGGL(k_remz_build, dim3(blocks_for((int64_t)c->g.XY, 256)), dim3(256), 0, c->stream, c->g, c->d_cap0,
                       c->d_remz0);
    hipLaunchKernelGGL(k_remz_build, dim3(blocks_for((int64_t)c->g.XY, 256)), dim3(256), 0, c->stream, c->g, c->d_rem,
                       c->d_remz);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(c->stream));
    c->remz_live = true;
    // the candidate costing reads stored via terms (one 4-byte load each):
    // from here on the refresh writes them too
    if (!c->prefix) {
      if (!c->d_vterm && (rc = dalloc(c, &c->d_vterm, (size_t)c->cs))) return rc;
      c->vterm_live = true;
      if ((rc = refresh_all_runs(c))) return rc;
      HIPC(c, hipStreamSynchronize(c->stream));
    }
  }
  c->prop = m;
  c->prop_k = k;
  c->prop_thr = threshold;
  return 0;
}

int gr_proposal_stats(gr_ctx* c, int64_t* pairs, int64_t* fallbacks) {
  if (!c || !pairs || !fallbacks) return fail(c, GR_EINVAL, "null argument");
  *pairs = *fallbacks = 0;
  if (!c->d_prop_stats) return 0;
  HIPC(c, hipStreamSynchronize(c->stream));
  int64_t v[2];
  HIPC(c, hipMemcpy(v, c->d_prop_stats, 16, hipMemcpyDeviceToHost));
  HIPC(c, hipMemset(c->d_prop_stats, 0, 16));
  *pairs = v[0];
  *fallbacks = v[1];
  return 0;
}

}  // extern "C"
