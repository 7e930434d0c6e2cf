// Learned pivot / phase scheduler: a 16 -> 64 -> 64 -> 2 MLP evaluated and
// trained with BF16 matrix-core (MFMA) instructions, FP32 accumulation.
//
// No reference counterpart: the reference picks the first vertex of a colour
// as pivot (src/scc_color.cc:64-93) and runs a fixed phase order
// (src/scc_main.cc:272-351).  Here each live vertex gets 16 features (its
// degrees, its partition's size, and round-level counts from the phase trace:
// live fraction, what the last trim / forward-backward / colouring steps
// resolved and cost, and which kind of partition step came last).  Head 0
// estimates how large the vertex's SCC is relative to its partition (the
// partition's pivot is its best-scoring vertex); head 1 is the logit of "taking
// a colouring step now, then alternating, finishes sooner than taking a
// forward-backward step now, then alternating", averaged over the live set to
// choose the next partition step.  Labels stay exact whatever the network
// says: it only orders work.
//
// Network layout (FP32 master copy, row-major): W1[16][64] b1[64] W2[64][64]
// b2[64] W3[64][2] b3[2].  One wave evaluates 16 samples at a time in
// transposed form: Z1^T = W1^T X^T, Z2^T = W2^T relu(Z1^T), Z3^T = W3^T relu(Z2^T),
// with v_mfma_f32_16x16x16_bf16; the accumulator layout of one product is the
// B-operand layout of the next, so activations stay in registers.  The
// backward pass runs the same chain backwards (dH2^T = W3 dZ3^T, dH1^T = W2 dZ2^T)
// and forms the weight gradients as products over the 16-sample dimension,
// with operands transposed through LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "common.h"
#include "engine.h"
#include "kernels.h"

namespace sccmi {

namespace {

constexpr int NF = SCC_SCHED_FEATURES;  // 16
constexpr int NH = SCC_SCHED_HIDDEN;    // 64
constexpr int NO = SCC_SCHED_HEADS;     // 2
constexpr int OFF_W1 = 0, OFF_B1 = NF * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NO, NPARAMS = OFF_B3 + NO;
static_assert(NPARAMS == SCC_SCHED_NPARAMS, "parameter count");

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (inputs are finite)
__device__ __forceinline__ short f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  return (short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(short h) { return __uint_as_float(((uint32_t)(uint16_t)h) << 16); }

__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// ---- weight fragments (A operands), built from the FP32 master copy --------------
// MFMA 16x16x16 operand layouts (lane l, q = 0..3):
//   A[i][k]: i = l % 16, k = 4 (l / 16) + q
//   B[k][j]: j = l % 16, k = 4 (l / 16) + q
//   D[i][j]: j = l % 16, i = 4 (l / 16) + q
struct Frags {
  s4 a1[4];      // W1^T tile t: rows 16t.., k = feature
  s4 a2[4][4];   // W2^T tile (t, s): rows 16t.. (out), k = 16s.. (in)
  s4 a3[4];      // W3^T (rows = head, padded to 16), k = 16s..
  f4 b1[4], b2[4];
  float b3[2];
};

__device__ __forceinline__ void load_frags(const float* __restrict__ P, Frags& F) {
  const int l = lane_id(), li = l & 15, lk = 4 * (l >> 4);
#pragma unroll
  for (int t = 0; t < 4; t++) {
    s4 a;
#pragma unroll
    for (int q = 0; q < 4; q++) a[q] = f2bf(P[OFF_W1 + (lk + q) * NH + 16 * t + li]);
    F.a1[t] = a;
#pragma unroll
    for (int s = 0; s < 4; s++) {
#pragma unroll
      for (int q = 0; q < 4; q++) a[q] = f2bf(P[OFF_W2 + (16 * s + lk + q) * NH + 16 * t + li]);
      F.a2[t][s] = a;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) a[q] = li < NO ? f2bf(P[OFF_W3 + (16 * t + lk + q) * NO + li]) : (short)0;
    F.a3[t] = a;
    f4 b, c;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      b[q] = P[OFF_B1 + 16 * t + lk + q];
      c[q] = P[OFF_B2 + 16 * t + lk + q];
    }
    F.b1[t] = b;
    F.b2[t] = c;
  }
  F.b3[0] = P[OFF_B3 + 0];
  F.b3[1] = P[OFF_B3 + 1];
}

// Forward of one 16-sample tile.  xb: lane l holds features 4(l/16).. of sample
// l%16.  Keeps pre-activations (for the backward pass) and returns the heads in
// lanes 0..15 (sample = lane).
struct Acts {
  f4 z1[4], z2[4];
  s4 h1[4], h2[4];  // bf16 activations in B-operand layout
};

__device__ __forceinline__ void forward_tile(const Frags& F, s4 xb, Acts& A, float& o0, float& o1) {
  const f4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < 4; t++) {
    f4 z = mfma(F.a1[t], xb, zero);
    s4 h;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      z[q] += F.b1[t][q];
      h[q] = f2bf(fmaxf(z[q], 0.f));
    }
    A.z1[t] = z;
    A.h1[t] = h;
  }
#pragma unroll
  for (int t = 0; t < 4; t++) {
    f4 z = zero;
#pragma unroll
    for (int s = 0; s < 4; s++) z = mfma(F.a2[t][s], A.h1[s], z);
    s4 h;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      z[q] += F.b2[t][q];
      h[q] = f2bf(fmaxf(z[q], 0.f));
    }
    A.z2[t] = z;
    A.h2[t] = h;
  }
  f4 z3 = zero;
#pragma unroll
  for (int s = 0; s < 4; s++) z3 = mfma(F.a3[s], A.h2[s], z3);
  o0 = z3[0] + F.b3[0];
  o1 = z3[1] + F.b3[1];
}

__device__ __forceinline__ float sigmoidf_(float x) { return 1.f / (1.f + __expf(-x)); }

// B operand of a 16-sample tile from FP32 rows (feature-major per sample).
__device__ __forceinline__ s4 load_x(const float* __restrict__ X, int64_t row, bool valid) {
  const int lk = 4 * (lane_id() >> 4);
  s4 x;
#pragma unroll
  for (int q = 0; q < 4; q++) x[q] = valid ? f2bf(X[row * NF + lk + q]) : (short)0;
  return x;
}

// ---- dataset forward ---------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_mlp_forward(const float* __restrict__ P, const float* __restrict__ X,
                                                       int64_t n, float* __restrict__ out) {
  Frags F;
  load_frags(P, F);
  const int l = lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * BLOCK) >> 6;
  const int64_t tiles = (n + 15) / 16;
  for (int64_t tile = wave; tile < tiles; tile += nwaves) {
    const int64_t row = tile * 16 + (l & 15);
    Acts A;
    float o0, o1;
    forward_tile(F, load_x(X, row, row < n), A, o0, o1);
    if (l < 16 && row < n) {
      out[row * NO + 0] = o0;
      out[row * NO + 1] = o1;
    }
  }
}

// ---- training: forward + backward of a batch, gradients summed into G --------------
// Loss per sample (mean over the batch, masks select the heads that have targets):
//   m0 (o0 - y0)^2 + m1 BCE(sigmoid(o1), y1).
// Each wave accumulates weight gradients of its tiles in registers and adds them
// to G (FP32 atomics) at the end; loss goes to G[NPARAMS].
constexpr int TW = 4;  // waves per block
__global__ __launch_bounds__(BLOCK) void k_mlp_grad(const float* __restrict__ P, const float* __restrict__ X,
                                                    const float* __restrict__ Y, const float* __restrict__ M,
                                                    const int32_t* __restrict__ idx, int64_t n, float inv_n,
                                                    float* __restrict__ G) {
  // per-wave LDS tiles, [unit][sample] bf16
  __shared__ short s_x[TW][NF * 16];
  __shared__ short s_h1[TW][NH * 16], s_h2[TW][NH * 16];
  __shared__ short s_d1[TW][NH * 16], s_d2[TW][NH * 16];
  __shared__ short s_d3[TW][NO * 16];
  Frags F;
  load_frags(P, F);
  const int l = lane_id(), li = l & 15, lg = l >> 4, lk = 4 * lg;
  const int w = threadIdx.x >> 6;
  // W2 (not transposed) and W3 as A operands of the backward chain
  s4 a2n[4][4], a3n[4];
  for (int t = 0; t < 4; t++) {
    for (int s = 0; s < 4; s++) {
      s4 a;
      for (int q = 0; q < 4; q++) a[q] = f2bf(P[OFF_W2 + (16 * t + li) * NH + 16 * s + lk + q]);
      a2n[t][s] = a;
    }
    s4 a;
    for (int q = 0; q < 4; q++) a[q] = (lk + q) < NO ? f2bf(P[OFF_W3 + (16 * t + li) * NO + lk + q]) : (short)0;
    a3n[t] = a;
  }
  const f4 zero = {0.f, 0.f, 0.f, 0.f};
  f4 gw1[4], gw2[4][4], gw3[4], gb1[4], gb2[4];
  for (int t = 0; t < 4; t++) {
    gw1[t] = gw3[t] = gb1[t] = gb2[t] = zero;
    for (int s = 0; s < 4; s++) gw2[t][s] = zero;
  }
  float gb3_0 = 0.f, gb3_1 = 0.f, loss = 0.f;
  const int64_t tiles = (n + 15) / 16;
  for (int64_t t0 = (int64_t)blockIdx.x * TW; t0 < tiles; t0 += (int64_t)gridDim.x * TW) {
    const int64_t tile = t0 + w;
    const bool tv = tile < tiles;
    const int64_t pos = tile * 16 + li;
    const bool valid = tv && pos < n;
    const int64_t row = valid ? (int64_t)idx[pos] : 0;
    const s4 xb = load_x(X, row, valid);
    Acts A;
    float o0, o1;
    forward_tile(F, xb, A, o0, o1);
    // head gradients (lanes 0..15 own sample li)
    float d0 = 0.f, d1 = 0.f;
    if (l < 16 && valid) {
      const float m0 = M[row * NO + 0], m1 = M[row * NO + 1];
      const float y0 = Y[row * NO + 0], y1 = Y[row * NO + 1];
      const float e = o0 - y0;
      const float p = sigmoidf_(o1);
      loss += (m0 * e * e + m1 * (fmaxf(o1, 0.f) - o1 * y1 + log1pf(__expf(-fabsf(o1))))) * inv_n;
      d0 = m0 * 2.f * e * inv_n;
      d1 = m1 * (p - y1) * inv_n;
    }
    gb3_0 += d0;
    gb3_1 += d1;
    // dZ3^T as a B operand: k = head (lanes 0..15, q = 0/1), j = sample
    s4 db3;
    db3[0] = l < 16 ? f2bf(d0) : (short)0;
    db3[1] = l < 16 ? f2bf(d1) : (short)0;
    db3[2] = 0;
    db3[3] = 0;
    // dH2^T = W3 dZ3^T ; dZ2^T = dH2^T * relu'(Z2^T)
    f4 dz2[4];
    s4 bz2[4];
    for (int t = 0; t < 4; t++) {
      f4 d = mfma(a3n[t], db3, zero);
      s4 b;
      for (int q = 0; q < 4; q++) {
        d[q] = A.z2[t][q] > 0.f ? d[q] : 0.f;
        b[q] = f2bf(d[q]);
      }
      dz2[t] = d;
      bz2[t] = b;
      gb2[t] += d;
    }
    // dH1^T = W2 dZ2^T ; dZ1^T = dH1^T * relu'(Z1^T)
    s4 bz1[4];
    for (int t = 0; t < 4; t++) {
      f4 d = zero;
      for (int s = 0; s < 4; s++) d = mfma(a2n[t][s], bz2[s], d);
      s4 b;
      for (int q = 0; q < 4; q++) {
        d[q] = A.z1[t][q] > 0.f ? d[q] : 0.f;
        b[q] = f2bf(d[q]);
      }
      bz1[t] = b;
      gb1[t] += d;
    }
    // stage [unit][sample] tiles in LDS for the weight-gradient products
    for (int t = 0; t < 4; t++)
      for (int q = 0; q < 4; q++) {
        const int u = 16 * t + lk + q;
        s_h1[w][u * 16 + li] = A.h1[t][q];
        s_h2[w][u * 16 + li] = A.h2[t][q];
        s_d1[w][u * 16 + li] = bz1[t][q];
        s_d2[w][u * 16 + li] = bz2[t][q];
      }
    for (int q = 0; q < 4; q++) s_x[w][(lk + q) * 16 + li] = xb[q];
    if (l < 16) {
      s_d3[w][li] = f2bf(d0);
      s_d3[w][16 + li] = f2bf(d1);
    }
    __syncthreads();
    // fragments along the sample dimension: lane reads unit (16t + li), samples lk..lk+3
    auto frag = [&](const short* base, int unit) {
      s4 v;
      for (int q = 0; q < 4; q++) v[q] = base[unit * 16 + lk + q];
      return v;
    };
    // dW1 (16 x 64) += X^T (16 x S) . dZ1 (S x 64)
    {
      const s4 ax = frag(s_x[w], li);
      for (int t = 0; t < 4; t++) gw1[t] = mfma(ax, frag(s_d1[w], 16 * t + li), gw1[t]);
    }
    // dW2 (64 in x 64 out) += H1^T . dZ2
    for (int t = 0; t < 4; t++) {
      const s4 ah = frag(s_h1[w], 16 * t + li);
      for (int s = 0; s < 4; s++) gw2[t][s] = mfma(ah, frag(s_d2[w], 16 * s + li), gw2[t][s]);
    }
    // dW3 (64 x heads) += H2^T . dZ3 (heads padded to 16 with zeros)
    {
      s4 bd;
      for (int q = 0; q < 4; q++) bd[q] = li < NO ? s_d3[w][li * 16 + lk + q] : (short)0;
      for (int t = 0; t < 4; t++) gw3[t] = mfma(frag(s_h2[w], 16 * t + li), bd, gw3[t]);
    }
    __syncthreads();
  }
  // bias gradients: sum the per-lane partials over the 16 sample lanes
  for (int t = 0; t < 4; t++)
    for (int q = 0; q < 4; q++) {
      float a = gb1[t][q], b = gb2[t][q];
      for (int o = 1; o < 16; o <<= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
      }
      if (li == 0) {
        atomicAdd(&G[OFF_B1 + 16 * t + lk + q], a);
        atomicAdd(&G[OFF_B2 + 16 * t + lk + q], b);
      }
    }
  for (int o = 1; o < 64; o <<= 1) {
    gb3_0 += __shfl_xor(gb3_0, o);
    gb3_1 += __shfl_xor(gb3_1, o);
    loss += __shfl_xor(loss, o);
  }
  if (l == 0) {
    atomicAdd(&G[OFF_B3 + 0], gb3_0);
    atomicAdd(&G[OFF_B3 + 1], gb3_1);
    atomicAdd(&G[NPARAMS], loss);
  }
  // weight gradients: D layout row = 4(l/16)+q (+16t), col = l%16 (+16s)
  for (int t = 0; t < 4; t++)
    for (int q = 0; q < 4; q++) {
      atomicAdd(&G[OFF_W1 + (lk + q) * NH + 16 * t + li], gw1[t][q]);
      for (int s = 0; s < 4; s++) atomicAdd(&G[OFF_W2 + (16 * t + lk + q) * NH + 16 * s + li], gw2[t][s][q]);
      if (li < NO) atomicAdd(&G[OFF_W3 + (16 * t + lk + q) * NO + li], gw3[t][q]);
    }
}

// Adam on the FP32 master copy; G[NPARAMS] (the loss) is left alone.
__global__ __launch_bounds__(BLOCK) void k_adam(float* __restrict__ P, const float* __restrict__ G,
                                                float* __restrict__ m1, float* __restrict__ m2, float lr,
                                                float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= NPARAMS) return;
  const float g = G[i];
  const float a = b1 * m1[i] + (1.f - b1) * g;
  const float b = b2 * m2[i] + (1.f - b2) * g * g;
  m1[i] = a;
  m2[i] = b;
  P[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// ---- features of live vertices ----------------------------------------------------
// Round-level entries come from the host (RoundFeat); per-vertex entries
// 0..5 are filled here.  Used identically for training samples and inference.
struct RoundFeat {
  float f[NF];
};

__device__ __forceinline__ void vertex_features(const GraphPtrs& G, const int32_t* __restrict__ color,
                                               const int32_t* __restrict__ pcount, float nlive, int32_t v,
                                               const RoundFeat& rf, float* x) {
  const float od = (float)(G.end[v] - G.begin[v]);
  const float id = (float)(G.rend[v] - G.rbegin[v]);
  const float ps = (float)pcount[slot_of(color[v], G.n)];
#pragma unroll
  for (int k = 0; k < NF; k++) x[k] = rf.f[k];
  x[0] = __log2f(1.f + od) * (1.f / 16.f);
  x[1] = __log2f(1.f + id) * (1.f / 16.f);
  x[2] = __log2f(1.f + od * id) * (1.f / 32.f);
  x[3] = __log2f(1.f + ps) * (1.f / 24.f);
  x[4] = ps / fmaxf(nlive, 1.f);
  x[5] = (od - id) / (od + id + 1.f);
}

// Partition sizes: pcount[slot] = live members.  Each workgroup counts its
// entries in an LDS table (open addressing, a few probes) and adds each
// bucket once at the end; an entry whose probes find no room adds directly.
// Large partitions (the global one, a giant) then cost one atomic per
// workgroup instead of one per wave and iteration on a single address.
constexpr int PC_HASH = 1024;  // LDS buckets (a power of two)
constexpr int PC_PROBES = 8;

__global__ __launch_bounds__(BLOCK) void k_part_count(const DevCtl* __restrict__ dc, int32_t n,
                                                      const int32_t* __restrict__ color,
                                                      int32_t* __restrict__ pcount, const int32_t* __restrict__ L0,
                                                      const int32_t* __restrict__ L1) {
  __shared__ int32_t hk[PC_HASH], hc[PC_HASH];
  for (int i = threadIdx.x; i < PC_HASH; i += BLOCK) {
    hk[i] = -1;
    hc[i] = 0;
  }
  __syncthreads();
  const int32_t* __restrict__ list = LIVE_IN(dc, L0, L1);
  const int64_t n_in = (int64_t)dc->live.n;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n_in; i += (int64_t)gridDim.x * BLOCK) {
    const int32_t c = color[list[i]];
    if (c == DONE) continue;
    const int32_t s = slot_of(c, n);
    int h = (int)(((uint32_t)s * 2654435761u) >> 22) & (PC_HASH - 1);
    bool added = false;
    for (int p = 0; p < PC_PROBES; p++, h = (h + 1) & (PC_HASH - 1)) {
      int32_t k = hk[h];
      if (k == -1) k = atomicCAS(&hk[h], -1, s);
      if (k == -1 || k == s) {
        atomicAdd(&hc[h], 1);
        added = true;
        break;
      }
    }
    if (!added) atomicAdd(&pcount[s], 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < PC_HASH; i += BLOCK)
    if (hc[i] > 0) atomicAdd(&pcount[hk[i]], hc[i]);
}

__global__ __launch_bounds__(BLOCK) void k_part_clear(const DevCtl* __restrict__ dc, int32_t n,
                                                      const int32_t* __restrict__ color,
                                                      int32_t* __restrict__ pcount, const int32_t* __restrict__ L0,
                                                      const int32_t* __restrict__ L1) {
  const int32_t* __restrict__ list = LIVE_IN(dc, L0, L1);
  const int64_t n_in = (int64_t)dc->live.n;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n_in; i += (int64_t)gridDim.x * BLOCK)
    pcount[slot_of(color[list[i]], n)] = 0;
}

// The phase vote needs the partition sizes of a sample of live vertices only.
// Sampled entries list[e * stride] (e < m, m <= VOTE_SAMPLES) name their
// partition slots; one pass over the live list counts the members of those
// slots in an LDS table per workgroup (open addressing), flushed with one
// atomic per slot and workgroup.
constexpr int VOTE_SAMPLES = 1024;
constexpr int VOTE_HASH = 2048;  // LDS buckets, a power of two >= 2 * VOTE_SAMPLES

__device__ __forceinline__ int vote_bucket(int32_t slot) {
  return (int)(((uint32_t)slot * 2654435761u) >> 21) & (VOTE_HASH - 1);
}

__global__ __launch_bounds__(BLOCK) void k_sample_slots(const DevCtl* __restrict__ dc, int32_t n,
                                                        const int32_t* __restrict__ color,
                                                        const int32_t* __restrict__ L0, const int32_t* __restrict__ L1,
                                                        int32_t stride, int32_t m, int32_t* __restrict__ sslots) {
  const int32_t* __restrict__ list = LIVE_IN(dc, L0, L1);
  const int e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= m) return;
  const int32_t c = color[list[(int64_t)e * stride]];
  sslots[e] = c != DONE ? slot_of(c, n) : -1;
}

__global__ __launch_bounds__(BLOCK) void k_part_count_sampled(const DevCtl* __restrict__ dc, int32_t n,
                                                              const int32_t* __restrict__ color,
                                                              const int32_t* __restrict__ L0,
                                                              const int32_t* __restrict__ L1,
                                                              const int32_t* __restrict__ sslots, int32_t m,
                                                              int32_t* __restrict__ pcount) {
  __shared__ int32_t hk[VOTE_HASH], hc[VOTE_HASH];
  for (int i = threadIdx.x; i < VOTE_HASH; i += BLOCK) {
    hk[i] = -1;
    hc[i] = 0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < m; e += BLOCK) {
    const int32_t slot = sslots[e];
    if (slot < 0) continue;
    for (int h = vote_bucket(slot);; h = (h + 1) & (VOTE_HASH - 1)) {
      const int32_t old = atomicCAS(&hk[h], -1, slot);
      if (old == -1 || old == slot) break;
    }
  }
  __syncthreads();
  const int32_t* __restrict__ list = LIVE_IN(dc, L0, L1);
  const int64_t n_in = (int64_t)dc->live.n;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n_in; i += (int64_t)gridDim.x * BLOCK) {
    const int32_t c = color[list[i]];
    if (c == DONE) continue;
    const int32_t slot = slot_of(c, n);
    for (int h = vote_bucket(slot); hk[h] != -1; h = (h + 1) & (VOTE_HASH - 1)) {
      if (hk[h] == slot) {
        atomicAdd(&hc[h], 1);
        break;
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < VOTE_HASH; i += BLOCK)
    if (hc[i] > 0) atomicAdd(&pcount[hk[i]], hc[i]);
}

__global__ __launch_bounds__(BLOCK) void k_part_clear_sampled(const int32_t* __restrict__ sslots, int32_t m,
                                                              int32_t* __restrict__ pcount) {
  const int e = blockIdx.x * BLOCK + threadIdx.x;
  if (e < m && sslots[e] >= 0) pcount[sslots[e]] = 0;
}

// Inference over the live list: pivot scores for every live vertex and the sum
// of sigmoid(head 1) (phase vote) into acc[0], count into acc[1].
__global__ __launch_bounds__(BLOCK) void k_sched_infer(const DevCtl* __restrict__ dc, GraphPtrs G,
                                                       const int32_t* __restrict__ color,
                                                       const int32_t* __restrict__ pcount, RoundFeat rf,
                                                       const float* __restrict__ P, uint32_t* __restrict__ scores,
                                                       float* __restrict__ acc, const int32_t* __restrict__ L0,
                                                       const int32_t* __restrict__ L1, int32_t stride) {
  const int32_t* __restrict__ list = LIVE_IN(dc, L0, L1);
  const int64_t n_in = (int64_t)dc->live.n;
  const int64_t m = (n_in + stride - 1) / stride;  // evaluated entries: every stride-th
  Frags F;
  load_frags(P, F);
  const int l = lane_id(), lk = 4 * (l >> 4);
  const int64_t wave = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * BLOCK) >> 6;
  const float nlive = (float)n_in;
  float vote = 0.f, cnt = 0.f;
  for (int64_t tile = wave; tile < (m + 15) / 16; tile += nwaves) {
    const int64_t e = tile * 16 + (l & 15);
    const bool valid = e < m;
    int32_t v = 0;
    s4 xb;
    if (valid) {
      v = list[e * stride];
      float x[NF];
      vertex_features(G, color, pcount, nlive, v, rf, x);
#pragma unroll
      for (int q = 0; q < 4; q++) xb[q] = f2bf(x[lk + q]);
    } else {
      xb = s4{0, 0, 0, 0};
    }
    Acts A;
    float o0, o1;
    forward_tile(F, xb, A, o0, o1);
    if (l < 16 && valid) {
      if (scores) {
        const float s = fminf(fmaxf(o0, 0.f), 1.f);
        scores[v] = (uint32_t)(s * 4294967040.f);
      }
      vote += sigmoidf_(o1);
      cnt += 1.f;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    vote += __shfl_xor(vote, o);
    cnt += __shfl_xor(cnt, o);
  }
  if (acc && l == 0 && cnt > 0.f) {
    atomicAdd(&acc[0], vote);
    atomicAdd(&acc[1], cnt);
  }
}

// Features of listed vertices into rows of X (training samples).
__global__ __launch_bounds__(BLOCK) void k_sample_features(GraphPtrs G, const int32_t* __restrict__ color,
                                                           const int32_t* __restrict__ pcount, RoundFeat rf,
                                                           float nlive, const int32_t* __restrict__ verts, int32_t k,
                                                           float* __restrict__ X) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= k) return;
  float x[NF];
  vertex_features(G, color, pcount, nlive, verts[i], rf, x);
  for (int q = 0; q < NF; q++) X[(int64_t)i * NF + q] = x[q];
}

#define LAUNCH(kern, grid, st, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), 0, st, __VA_ARGS__)

// ---- host state ------------------------------------------------------------------------
struct Sched {
  bool have_weights = false;
  float* P = nullptr;       // device master weights (NPARAMS)
  float* G = nullptr;       // gradients + loss (NPARAMS + 1)
  float* m1 = nullptr;      // Adam moments
  float* m2 = nullptr;
  float* acc = nullptr;     // inference vote accumulators (2) ...
  float* h_acc = nullptr;   // pinned mirror
  int32_t* pcount = nullptr;  // partition sizes (n + 1)
  int32_t* sslots = nullptr;  // partition slots of the vote's samples (VOTE_SAMPLES)
  int32_t pcount_n = -1;
  int64_t adam_t = 0;
  // dataset buffers for training / evaluation
  float *dX = nullptr, *dY = nullptr, *dM = nullptr, *dO = nullptr;
  int32_t* didx = nullptr;
  int64_t cap = 0;
  int32_t* dverts = nullptr;
  float* dsx = nullptr;
  int32_t vcap = 0;
  bool vote_pending = false;
};

Sched* sched_of(scc_ctx* c) { return (Sched*)c->sched; }

int ensure_sched(scc_ctx* c) {
  Sched* s = sched_of(c);
  if (!s) {
    s = new Sched();
    c->sched = s;
  }
  int rc;
  if (!s->P) {
    if ((rc = dev_alloc(c, (void**)&s->P, 4ull * NPARAMS, "scheduler weights")) ||
        (rc = dev_alloc(c, (void**)&s->G, 4ull * (NPARAMS + 1), "scheduler gradients")) ||
        (rc = dev_alloc(c, (void**)&s->m1, 4ull * NPARAMS, "scheduler moments")) ||
        (rc = dev_alloc(c, (void**)&s->m2, 4ull * NPARAMS, "scheduler moments")) ||
        (rc = dev_alloc(c, (void**)&s->acc, 64, "scheduler vote")) ||
        (rc = dev_alloc(c, (void**)&s->sslots, 4ull * VOTE_SAMPLES, "vote sample slots")))
      return rc;
    if ((rc = hip_check(c, hipHostMalloc((void**)&s->h_acc, 64, 0), "pinned vote"))) return rc;
    if ((rc = hip_check(c, hipMemsetAsync(s->m1, 0, 4ull * NPARAMS, c->stream), "moments")) ||
        (rc = hip_check(c, hipMemsetAsync(s->m2, 0, 4ull * NPARAMS, c->stream), "moments")))
      return rc;
  }
  return SCC_OK;
}

// Per-vertex buffers of the current graph: partition sizes and pivot scores,
// both (re)sized together whenever the vertex count changes.
int ensure_pcount(scc_ctx* c, Sched* s) {
  if (s->pcount_n == c->n && s->pcount && c->sched_scores) return SCC_OK;
  dev_free(c, s->pcount);
  dev_free(c, c->sched_scores);
  s->pcount = nullptr;
  c->sched_scores = nullptr;
  s->pcount_n = -1;
  int rc;
  if ((rc = dev_alloc(c, (void**)&s->pcount, 4ull * ((uint64_t)c->n + 1), "partition sizes")) ||
      (rc = dev_alloc(c, (void**)&c->sched_scores, 4ull * ((uint64_t)c->n + 1), "pivot scores")))
    return rc;
  launch_fill_i32(c->stream, s->pcount, 0, c->n + 1);
  s->pcount_n = c->n;
  return SCC_OK;
}

int ensure_data(scc_ctx* c, Sched* s, int64_t n) {
  if (n <= s->cap) return SCC_OK;
  float** fs[] = {&s->dX, &s->dY, &s->dM, &s->dO};
  for (auto p : fs) {
    dev_free(c, *p);
    *p = nullptr;
  }
  dev_free(c, s->didx);
  s->didx = nullptr;
  int rc;
  if ((rc = dev_alloc(c, (void**)&s->dX, 4ull * n * NF, "scheduler samples")) ||
      (rc = dev_alloc(c, (void**)&s->dY, 4ull * n * NO, "scheduler targets")) ||
      (rc = dev_alloc(c, (void**)&s->dM, 4ull * n * NO, "scheduler masks")) ||
      (rc = dev_alloc(c, (void**)&s->dO, 4ull * n * NO, "scheduler outputs")) ||
      (rc = dev_alloc(c, (void**)&s->didx, 4ull * n, "scheduler batch index")))
    return rc;
  s->cap = n;
  return SCC_OK;
}

// Round-level features from the phase trace.
RoundFeat round_features(const scc_ctx* c) {
  RoundFeat rf;
  for (auto& x : rf.f) x = 0.f;
  const double nl = (double)c->n_live, n = std::max<double>(1.0, (double)c->n);
  double trim = 0, fb = 0, col = 0, fbms = 0, colms = 0, last = 0.5;
  bool got_trim = false, got_fb = false, got_col = false, got_last = false;
  for (auto it = c->trace.rbegin(); it != c->trace.rend(); ++it) {
    if (!got_last && (it->kind == SCC_PHASE_FB_ROUND || it->kind == SCC_PHASE_COLOR_ROUND)) {
      last = it->kind == SCC_PHASE_COLOR_ROUND ? 1.0 : 0.0;
      got_last = true;
    }
    const double before = std::max<double>(1.0, (double)it->live_before);
    const double res = (double)(it->live_before - it->live_after) / before;
    if (!got_trim && it->kind == SCC_PHASE_TRIM1) {
      trim = res;
      got_trim = true;
    } else if (!got_fb && (it->kind == SCC_PHASE_FB_ROUND || it->kind == SCC_PHASE_FB_GLOBAL)) {
      fb = res;
      fbms = it->ms * 1e6 / before;
      got_fb = true;
    } else if (!got_col && it->kind == SCC_PHASE_COLOR_ROUND) {
      col = res;
      colms = it->ms * 1e6 / before;
      got_col = true;
    }
  }
  rf.f[6] = (float)(std::log2(1.0 + nl) / 24.0);
  rf.f[7] = (float)(nl / n);
  rf.f[8] = c->classes ? 1.f : 0.f;
  rf.f[9] = (float)trim;
  rf.f[10] = (float)fb;
  rf.f[11] = (float)col;
  rf.f[12] = (float)(std::log2(1.0 + fbms) / 16.0);
  rf.f[13] = (float)(std::log2(1.0 + colms) / 16.0);
  rf.f[14] = (float)last;
  rf.f[15] = (float)((double)c->m / n / 64.0);
  return rf;
}

int count_partitions(scc_ctx* c, Sched* s) {
  int rc;
  if ((rc = ensure_pcount(c, s))) return rc;
  const unsigned grid = grid_for(c->n_live);
  LAUNCH(k_part_count, grid, c->stream, c->dc, c->n, c->color, s->pcount, c->live[0], c->live[1]);
  return hip_check(c, hipGetLastError(), "partition sizes");
}

int clear_partitions(scc_ctx* c, Sched* s) {
  LAUNCH(k_part_clear, grid_for(c->n_live), c->stream, c->dc, c->n, c->color, s->pcount, c->live[0], c->live[1]);
  return hip_check(c, hipGetLastError(), "partition sizes clear");
}

// Run inference over the live set (every stride-th vertex for the vote; all
// vertices when scores are wanted).  Returns the mean phase vote.
int infer(scc_ctx* c, Sched* s, bool scores, double* vote) {
  int rc;
  // scores: every live vertex, with every partition's size; the vote alone:
  // at most VOTE_SAMPLES evenly spaced entries, with their partitions' sizes
  const int32_t stride =
      scores ? 1 : (int32_t)std::max<int64_t>(1, (c->n_live + VOTE_SAMPLES - 1) / VOTE_SAMPLES);
  const int64_t m = (c->n_live + stride - 1) / stride;
  if (scores) {
    if ((rc = count_partitions(c, s))) return rc;
  } else {
    if ((rc = ensure_pcount(c, s))) return rc;
    const unsigned sg = (unsigned)std::max<int64_t>(1, (m + BLOCK - 1) / BLOCK);
    LAUNCH(k_sample_slots, sg, c->stream, c->dc, c->n, c->color, c->live[0], c->live[1], stride, (int32_t)m,
           s->sslots);
    LAUNCH(k_part_count_sampled, grid_for(c->n_live, (unsigned)c->tune.grid_cap), c->stream, c->dc, c->n, c->color,
           c->live[0], c->live[1], s->sslots, (int32_t)m, s->pcount);
  }
  const RoundFeat rf = round_features(c);
  // (scores alone: no vote, no read-back)
  if (vote) (void)hipMemsetAsync(s->acc, 0, 8, c->stream);
  const unsigned grid = grid_for(std::max<int64_t>(1, (m + 15) / 16) * 64, 2048);
  LAUNCH(k_sched_infer, grid, c->stream, c->dc, gp(c), c->color, s->pcount, rf, s->P,
         scores ? c->sched_scores : nullptr, vote ? s->acc : nullptr, c->live[0], c->live[1], stride);
  if (scores) {
    if ((rc = clear_partitions(c, s))) return rc;
  } else {
    LAUNCH(k_part_clear_sampled, (unsigned)std::max<int64_t>(1, (m + BLOCK - 1) / BLOCK), c->stream, s->sslots,
           (int32_t)m, s->pcount);
  }
  if (!vote) return hip_check(c, hipGetLastError(), "pivot scores");
  if ((rc = hip_check(c, hipMemcpyAsync(s->h_acc, s->acc, 8, hipMemcpyDeviceToHost, c->stream), "vote")) ||
      (rc = hip_check(c, hipStreamSynchronize(c->stream), "vote sync")))
    return rc;
  if (vote) *vote = s->h_acc[1] > 0 ? s->h_acc[0] / s->h_acc[1] : 0.0;
  return SCC_OK;
}

// ---- snapshot / restore of the decomposition state (counterfactual rounds) ------------
struct Snap {
  std::vector<int32_t> color, label, live, W;
  int64_t n_live = 0;
  bool colors_split = false, fresh = false, multi_colors = false, classes = false, last_color_round = false,
       w_fresh = false;
  size_t trace_len = 0;
  int32_t lv_hint[3] = {0, 0, 0}, it_hint = 0, trim_hint[2] = {0, 0};  // launch-batch hints: both branches start alike
  int32_t fix_idx = 0, search_idx = 0;
};

int take_snap(scc_ctx* c, Snap& sn) {
  sn.color.resize(c->n);
  sn.label.resize(c->n);
  sn.live.resize(std::max<int64_t>(1, c->n_live));
  int rc;
  const int32_t* live = c->live[c->cur];
  if ((rc = hip_check(c, hipMemcpyAsync(sn.color.data(), c->color, 4ull * c->n, hipMemcpyDeviceToHost, c->stream), "snapshot")) ||
      (rc = hip_check(c, hipMemcpyAsync(sn.label.data(), c->label, 4ull * c->n, hipMemcpyDeviceToHost, c->stream), "snapshot")) ||
      (c->n_live && (rc = hip_check(c, hipMemcpyAsync(sn.live.data(), live, 4ull * c->n_live, hipMemcpyDeviceToHost, c->stream), "snapshot"))) ||
      (rc = hip_check(c, hipStreamSynchronize(c->stream), "snapshot sync")))
    return rc;
  // the WCC labels a first colouring step reuses (both branches start alike)
  sn.w_fresh = c->w_fresh;
  if (sn.w_fresh) {
    sn.W.resize(c->n);
    if ((rc = hip_check(c, hipMemcpyAsync(sn.W.data(), c->W, 4ull * c->n, hipMemcpyDeviceToHost, c->stream), "snapshot")) ||
        (rc = hip_check(c, hipStreamSynchronize(c->stream), "snapshot sync")))
      return rc;
  }
  sn.n_live = c->n_live;
  sn.colors_split = c->colors_split;
  sn.fresh = c->fresh;
  sn.multi_colors = c->multi_colors;
  sn.classes = c->classes;
  sn.last_color_round = c->last_color_round;
  sn.trace_len = c->trace.size();
  for (int i = 0; i < 3; i++) sn.lv_hint[i] = c->lv_hint[i];
  sn.it_hint = c->it_hint;
  sn.fix_idx = c->fix_idx;
  sn.search_idx = c->search_idx;
  for (int i = 0; i < 2; i++) sn.trim_hint[i] = c->trim_hint[i];
  return SCC_OK;
}

int restore_snap(scc_ctx* c, const Snap& sn) {
  int rc;
  if ((rc = hip_check(c, hipMemcpyAsync(c->color, sn.color.data(), 4ull * c->n, hipMemcpyHostToDevice, c->stream), "restore")) ||
      (rc = hip_check(c, hipMemcpyAsync(c->label, sn.label.data(), 4ull * c->n, hipMemcpyHostToDevice, c->stream), "restore")) ||
      (sn.n_live && (rc = hip_check(c, hipMemcpyAsync(c->live[0], sn.live.data(), 4ull * sn.n_live, hipMemcpyHostToDevice, c->stream), "restore"))))
    return rc;
  if (sn.w_fresh &&
      (rc = hip_check(c, hipMemcpyAsync(c->W, sn.W.data(), 4ull * c->n, hipMemcpyHostToDevice, c->stream), "restore")))
    return rc;
  c->w_fresh = sn.w_fresh;
  c->cut_out = c->cut_in = false;  // (rows cut in the other branch may lack edges of this state)
  if ((rc = set_live(c, 0, sn.n_live))) return rc;
  c->colors_split = sn.colors_split;
  c->fresh = sn.fresh;
  c->multi_colors = sn.multi_colors;
  c->classes = sn.classes;
  c->last_color_round = sn.last_color_round;
  c->trace.resize(sn.trace_len);
  for (int i = 0; i < 3; i++) c->lv_hint[i] = sn.lv_hint[i];
  c->it_hint = sn.it_hint;
  c->fix_idx = sn.fix_idx;
  c->search_idx = sn.search_idx;
  for (int i = 0; i < 2; i++) c->trim_hint[i] = sn.trim_hint[i];
  return hip_check(c, hipStreamSynchronize(c->stream), "restore sync");
}

// One partition step: 0 = forward-backward (after WCC when not yet in classes),
// 1 = colouring.  Plain pivots when the network is not used for them.
int run_action(scc_ctx* c, int action, bool learned_pivots) {
  int rc;
  if (action == 0) {
    if (!c->classes && (rc = do_wcc(c, nullptr))) return rc;
    const int policy = learned_pivots ? SCC_PIVOT_LEARNED : SCC_PIVOT_DEGREE_PRODUCT;
    return do_fb_round(c, policy, true, -1, nullptr, nullptr, nullptr, false, nullptr);
  }
  return do_color_round(c, nullptr);
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

bool sched_ready(const scc_ctx* c) {
  const Sched* s = (const Sched*)c->sched;
  return s && s->have_weights;
}

// The network overrides the plain rule (a colouring step, the default
// partition mode's) only when its mean vote is far from even.
constexpr double SCHED_MARGIN = 0.25;

// A step kind repeated right after it resolved less than this share of the
// live set is replaced by the other kind (a progress guard: a network that
// keeps choosing forward-backward steps on a graph of many small partitions
// would otherwise approach the reference's per-partition FB, orders of
// magnitude slower).
constexpr double SCHED_MIN_PROGRESS = 0.02;

int sched_choose(scc_ctx* c, int* action) {
  Sched* s = sched_of(c);
  double vote = 0;
  int rc = infer(c, s, false, &vote);
  if (rc) return rc;
  const int plain = 1;
  int last = -1;
  double progress = 1.0;
  for (auto it = c->trace.rbegin(); it != c->trace.rend(); ++it) {
    if (it->kind == SCC_PHASE_FB_ROUND || it->kind == SCC_PHASE_COLOR_ROUND) {
      last = it->kind == SCC_PHASE_FB_ROUND ? 0 : 1;
      progress = (double)(it->live_before - it->live_after) / (double)std::max<int64_t>(1, it->live_before);
      break;
    }
  }
  *action = vote > 0.5 + SCHED_MARGIN ? 1 : (vote < 0.5 - SCHED_MARGIN ? 0 : plain);
  if (*action == last && progress < SCHED_MIN_PROGRESS) *action = 1 - last;
  return SCC_OK;
}

int sched_pivots(scc_ctx* c) {
  Sched* s = sched_of(c);
  if (!s || !s->have_weights) return set_error(c, SCC_ERR_STATE, "no scheduler weights");
  return infer(c, s, true, nullptr);
}

void destroy_sched(scc_ctx* c) {
  Sched* s = sched_of(c);
  if (!s) return;
  void* ptrs[] = {s->P, s->G, s->m1, s->m2, s->acc, s->sslots, s->pcount, s->dX, s->dY, s->dM, s->dO, s->didx,
                  s->dverts, s->dsx, c->sched_scores};
  for (void* p : ptrs) dev_free(c, p);
  c->sched_scores = nullptr;
  if (s->h_acc) (void)hipHostFree(s->h_acc);
  delete s;
  c->sched = nullptr;
}

}  // namespace sccmi

using namespace sccmi;

extern "C" {

int scc_phase_trace(scc_ctx* c, scc_phase_rec* out, int64_t cap, int64_t* n) {
  if (!c || !n || cap < 0 || (cap > 0 && !out)) return SCC_ERR_ARG;
  *n = (int64_t)c->trace.size();
  const int64_t k = std::min<int64_t>(cap, *n);
  if (k > 0) std::memcpy(out, c->trace.data(), sizeof(scc_phase_rec) * (size_t)k);
  return SCC_OK;
}

int scc_sched_set_weights(scc_ctx* c, const float* params) {
  if (!c) return SCC_ERR_ARG;
  (void)hipSetDevice(c->device);
  int rc;
  if ((rc = ensure_sched(c))) return rc;
  Sched* s = sched_of(c);
  if (!params) {
    s->have_weights = false;
    return SCC_OK;
  }
  for (int i = 0; i < NPARAMS; i++)
    if (!std::isfinite(params[i])) return set_error(c, SCC_ERR_ARG, "non-finite scheduler weight");
  if ((rc = hip_check(c, hipMemcpyAsync(s->P, params, 4ull * NPARAMS, hipMemcpyHostToDevice, c->stream), "weights")) ||
      (rc = hip_check(c, hipMemsetAsync(s->m1, 0, 4ull * NPARAMS, c->stream), "moments")) ||
      (rc = hip_check(c, hipMemsetAsync(s->m2, 0, 4ull * NPARAMS, c->stream), "moments")) ||
      (rc = hip_check(c, hipStreamSynchronize(c->stream), "weights sync")))
    return rc;
  s->adam_t = 0;
  s->have_weights = true;
  return SCC_OK;
}

int scc_sched_get_weights(scc_ctx* c, float* params) {
  if (!c || !params) return SCC_ERR_ARG;
  Sched* s = sched_of(c);
  if (!s || !s->have_weights) return set_error(c, SCC_ERR_STATE, "no scheduler weights");
  (void)hipSetDevice(c->device);
  int rc;
  if ((rc = hip_check(c, hipMemcpyAsync(params, s->P, 4ull * NPARAMS, hipMemcpyDeviceToHost, c->stream), "weights")))
    return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "weights sync");
}

int scc_sched_forward(scc_ctx* c, const float* X, int64_t n, float* out) {
  if (!c || n < 0 || (n > 0 && (!X || !out))) return SCC_ERR_ARG;
  Sched* s = sched_of(c);
  if (!s || !s->have_weights) return set_error(c, SCC_ERR_STATE, "no scheduler weights");
  if (n == 0) return SCC_OK;
  (void)hipSetDevice(c->device);
  int rc;
  if ((rc = ensure_data(c, s, n))) return rc;
  if ((rc = hip_check(c, hipMemcpyAsync(s->dX, X, 4ull * n * NF, hipMemcpyHostToDevice, c->stream), "samples")))
    return rc;
  LAUNCH(k_mlp_forward, grid_for((n + 15) / 16 * 64, 1024), c->stream, s->P, s->dX, n, s->dO);
  if ((rc = hip_check(c, hipGetLastError(), "forward")) ||
      (rc = hip_check(c, hipMemcpyAsync(out, s->dO, 4ull * n * NO, hipMemcpyDeviceToHost, c->stream), "outputs")))
    return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "forward sync");
}

int scc_sched_gradients(scc_ctx* c, const float* X, const float* Y, const float* M, int64_t n, float* loss,
                        float* grads) {
  if (!c || n <= 0 || !X || !Y || !M || !grads) return SCC_ERR_ARG;
  Sched* s = sched_of(c);
  if (!s || !s->have_weights) return set_error(c, SCC_ERR_STATE, "no scheduler weights");
  (void)hipSetDevice(c->device);
  int rc;
  if ((rc = ensure_data(c, s, n))) return rc;
  std::vector<int32_t> idx(n);
  for (int64_t i = 0; i < n; i++) idx[i] = (int32_t)i;
  if ((rc = hip_check(c, hipMemcpyAsync(s->dX, X, 4ull * n * NF, hipMemcpyHostToDevice, c->stream), "samples")) ||
      (rc = hip_check(c, hipMemcpyAsync(s->dY, Y, 4ull * n * NO, hipMemcpyHostToDevice, c->stream), "targets")) ||
      (rc = hip_check(c, hipMemcpyAsync(s->dM, M, 4ull * n * NO, hipMemcpyHostToDevice, c->stream), "masks")) ||
      (rc = hip_check(c, hipMemcpyAsync(s->didx, idx.data(), 4ull * n, hipMemcpyHostToDevice, c->stream), "index")) ||
      (rc = hip_check(c, hipMemsetAsync(s->G, 0, 4ull * (NPARAMS + 1), c->stream), "grad clear")))
    return rc;
  const int64_t tiles = (n + 15) / 16;
  LAUNCH(k_mlp_grad, (unsigned)std::min<int64_t>(1024, (tiles + TW - 1) / TW), c->stream, s->P, s->dX, s->dY,
         s->dM, s->didx, n, 1.f / (float)n, s->G);
  std::vector<float> g(NPARAMS + 1);
  if ((rc = hip_check(c, hipGetLastError(), "gradients")) ||
      (rc = hip_check(c, hipMemcpyAsync(g.data(), s->G, 4ull * (NPARAMS + 1), hipMemcpyDeviceToHost, c->stream), "grads")) ||
      (rc = hip_check(c, hipStreamSynchronize(c->stream), "grads sync")))
    return rc;
  std::memcpy(grads, g.data(), 4ull * NPARAMS);
  if (loss) *loss = g[NPARAMS];
  return SCC_OK;
}

int scc_sched_train(scc_ctx* c, const float* X, const float* Y, const float* M, int64_t n, int32_t epochs,
                    int32_t batch, float lr, uint64_t seed, float* last_loss) {
  if (!c || n <= 0 || !X || !Y || !M || epochs < 1 || batch < 16 || !(lr > 0)) return SCC_ERR_ARG;
  Sched* s = sched_of(c);
  if (!s || !s->have_weights) return set_error(c, SCC_ERR_STATE, "set initial weights first");
  (void)hipSetDevice(c->device);
  int rc;
  if ((rc = ensure_data(c, s, n))) return rc;
  if ((rc = hip_check(c, hipMemcpyAsync(s->dX, X, 4ull * n * NF, hipMemcpyHostToDevice, c->stream), "samples")) ||
      (rc = hip_check(c, hipMemcpyAsync(s->dY, Y, 4ull * n * NO, hipMemcpyHostToDevice, c->stream), "targets")) ||
      (rc = hip_check(c, hipMemcpyAsync(s->dM, M, 4ull * n * NO, hipMemcpyHostToDevice, c->stream), "masks")))
    return rc;
  std::mt19937_64 rng(seed);
  std::vector<int32_t> perm(n);
  for (int64_t i = 0; i < n; i++) perm[i] = (int32_t)i;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float loss = 0.f;
  for (int ep = 0; ep < epochs; ep++) {
    std::shuffle(perm.begin(), perm.end(), rng);
    if ((rc = hip_check(c, hipMemcpyAsync(s->didx, perm.data(), 4ull * n, hipMemcpyHostToDevice, c->stream), "index")))
      return rc;
    // the permutation buffer is pageable host memory: wait before reshuffling it
    if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "index sync"))) return rc;
    for (int64_t b0 = 0; b0 < n; b0 += batch) {
      const int64_t nb = std::min<int64_t>(batch, n - b0);
      (void)hipMemsetAsync(s->G, 0, 4ull * (NPARAMS + 1), c->stream);
      const int64_t tiles = (nb + 15) / 16;
      LAUNCH(k_mlp_grad, (unsigned)std::min<int64_t>(256, (tiles + TW - 1) / TW), c->stream, s->P, s->dX, s->dY,
             s->dM, s->didx + b0, nb, 1.f / (float)nb, s->G);
      s->adam_t++;
      const float c1 = 1.f - std::pow(b1, (float)s->adam_t), c2 = 1.f - std::pow(b2, (float)s->adam_t);
      LAUNCH(k_adam, (NPARAMS + BLOCK - 1) / BLOCK, c->stream, s->P, s->G, s->m1, s->m2, lr, b1, b2, eps, c1, c2);
    }
    if ((rc = hip_check(c, hipGetLastError(), "train step"))) return rc;
  }
  if ((rc = hip_check(c, hipMemcpyAsync(&loss, s->G + NPARAMS, 4, hipMemcpyDeviceToHost, c->stream), "loss")) ||
      (rc = hip_check(c, hipStreamSynchronize(c->stream), "train sync")))
    return rc;
  if (last_loss) *last_loss = loss;
  return SCC_OK;
}

int scc_sched_rollout(scc_ctx* c, const scc_run_opts* opts_in, uint64_t seed, float explore,
                      int32_t samples_per_round, float* X, float* Y, float* M, int64_t cap, int64_t* n_out,
                      scc_rollout_stats* rs) {
  if (!c || !n_out || cap < 0 || samples_per_round < 0 || (cap > 0 && (!X || !Y || !M))) return SCC_ERR_ARG;
  if (int r0 = require_graph(c)) return r0;
  int rc;
  if ((rc = ensure_sched(c))) return rc;
  Sched* s = sched_of(c);
  if ((rc = ensure_pcount(c, s))) return rc;
  scc_run_opts opts;
  if (opts_in) opts = *opts_in;
  else scc_run_opts_default(&opts);
  const bool learned = s->have_weights;
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<float> U(0.f, 1.f);
  scc_rollout_stats st;
  std::memset(&st, 0, sizeof(st));
  auto t_all = std::chrono::steady_clock::now();
  // a batch of graphs starts from its per-graph partitions (scc_batch_run)
  if ((rc = start_run(c))) return rc;
  // method-2 prefix: trim1 -> global FB -> trim1 / trim2 / trim1
  if ((rc = do_trim1(c, opts.trim_stop, nullptr))) return rc;
  if (c->n_live &&
      (rc = do_fb_round(c, opts.pivot_policy, c->colors_split, -1, nullptr, nullptr, nullptr, false, nullptr)))
    return rc;
  if ((rc = do_trim1(c, opts.trim_stop, nullptr))) return rc;
  if (opts.use_trim2 && c->n_live) {
    if ((rc = do_trim2(c, nullptr)) || (rc = do_trim1(c, opts.trim_stop, nullptr))) return rc;
  }
  if (opts.use_wcc && c->n_live && (rc = do_wcc(c, nullptr))) return rc;
  std::vector<int32_t> sample_v;       // sampled vertex per row
  std::vector<float> sample_part;      // its partition size when sampled
  int64_t rows = 0;
  int32_t r = 0;
  Snap sn;
  while (c->n_live > 0) {
    if ((rc = check_cancel(c))) return rc;
    c->round_idx = r;
    if (opts.trim_between && r > 0) {
      if ((rc = do_trim1(c, 0, nullptr))) return rc;
      if (c->n_live == 0) break;
    }
    // samples of this state
    const int32_t k = (int32_t)std::min<int64_t>({(int64_t)samples_per_round, c->n_live, cap - rows});
    int64_t base_row = rows;
    if (k > 0) {
      std::vector<int32_t> live(c->n_live);
      if ((rc = hip_check(c, hipMemcpyAsync(live.data(), c->live[c->cur], 4ull * c->n_live, hipMemcpyDeviceToHost, c->stream), "live")) ||
          (rc = hip_check(c, hipStreamSynchronize(c->stream), "live sync")))
        return rc;
      std::vector<int32_t> pick(k);
      for (int32_t i = 0; i < k; i++) pick[i] = live[(size_t)(U(rng) * (float)c->n_live) % c->n_live];
      if (k > s->vcap) {
        dev_free(c, s->dverts);
        dev_free(c, s->dsx);
        s->dverts = nullptr;
        s->dsx = nullptr;
        if ((rc = dev_alloc(c, (void**)&s->dverts, 4ull * k, "sample vertices")) ||
            (rc = dev_alloc(c, (void**)&s->dsx, 4ull * k * NF, "sample features")))
          return rc;
        s->vcap = k;
      }
      if ((rc = count_partitions(c, s))) return rc;
      const RoundFeat rf = round_features(c);
      std::vector<int32_t> pc(k);
      if ((rc = hip_check(c, hipMemcpyAsync(s->dverts, pick.data(), 4ull * k, hipMemcpyHostToDevice, c->stream), "samples")))
        return rc;
      LAUNCH(k_sample_features, (k + BLOCK - 1) / BLOCK, c->stream, gp(c), c->color, s->pcount, rf,
             (float)c->n_live, s->dverts, k, s->dsx);
      if ((rc = hip_check(c, hipMemcpyAsync(X + rows * NF, s->dsx, 4ull * k * NF, hipMemcpyDeviceToHost, c->stream), "features")) ||
          (rc = hip_check(c, hipStreamSynchronize(c->stream), "features sync")))
        return rc;
      if ((rc = clear_partitions(c, s))) return rc;
      for (int32_t i = 0; i < k; i++) {
        // partition size back from feature 3 (log2(1 + size) / 24)
        const float ps = std::exp2(X[(rows + i) * NF + 3] * 24.f) - 1.f;
        sample_v.push_back(pick[i]);
        sample_part.push_back(ps);
      }
      rows += k;
    }
    // counterfactual: from the same state, each step followed by the
    // continuation policy until nothing is live (the installed network's
    // choices, else colouring steps, the plain rule); the better step is the
    // one whose decomposition finishes first
    if ((rc = take_snap(c, sn))) return rc;
    double total[2];
    const int first = U(rng) < 0.5f ? 0 : 1;  // random order: neither branch always runs first
    for (int k = 0; k < 2; k++) {
      const int a = k == 0 ? first : 1 - first;
      auto t0 = std::chrono::steady_clock::now();
      int step = a;
      for (int64_t guard = 0; c->n_live > 0; guard++) {
        if ((rc = check_cancel(c))) return rc;
        if (guard > 4 * (int64_t)c->n + 64) return set_error(c, SCC_ERR_LIMIT, "rollout did not converge");
        if ((rc = run_action(c, step, false))) return rc;
        step = 1;
        if (learned && c->n_live > 0 && (rc = sched_choose(c, &step))) return rc;
      }
      total[a] = ms_since(t0);
      if ((rc = restore_snap(c, sn))) return rc;
    }
    const int better = total[1] < total[0] ? 1 : 0;
    for (int64_t i = base_row; i < rows; i++) {
      Y[i * NO + 1] = (float)better;
      M[i * NO + 1] = 1.f;
    }
    int act = better;
    if (learned) {
      int la = 0;
      if ((rc = sched_choose(c, &la))) return rc;
      act = la;
      st.agree += la == better;
    }
    if (U(rng) < explore) act = 1 - act;
    if (st.rounds == 0) st.ms_best = std::min(total[0], total[1]);
    st.rounds++;
    if ((rc = run_action(c, act, learned))) return rc;
    r++;
  }
  c->round_idx = -1;
  int64_t nsccs = 0;
  if ((rc = do_canon(c, &nsccs))) return rc;
  // pivot targets: log(1 + |SCC(v)|) / log(1 + |partition(v)| at sampling)
  std::vector<int32_t> canon(c->n);
  if ((rc = hip_check(c, hipMemcpyAsync(canon.data(), c->canon, 4ull * c->n, hipMemcpyDeviceToHost, c->stream), "labels")) ||
      (rc = hip_check(c, hipStreamSynchronize(c->stream), "labels sync")))
    return rc;
  std::vector<int32_t> size(c->n, 0);
  for (int32_t v = 0; v < c->n; v++) size[canon[v]]++;
  for (int64_t i = 0; i < rows; i++) {
    const double sc = (double)size[canon[sample_v[i]]];
    const double ps = std::max(1.0, (double)sample_part[i]);
    Y[i * NO + 0] = (float)(std::log2(1.0 + sc) / std::log2(1.0 + std::max(ps, sc)));
    M[i * NO + 0] = 1.f;
  }
  *n_out = rows;
  st.num_sccs = nsccs;
  st.ms_total = ms_since(t_all);
  if (rs) *rs = st;
  return SCC_OK;
}

}  // extern "C"
