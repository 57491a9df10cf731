// Learned candidate prefilter: an MLP 8 -> 256 -> 256 -> 1 (ReLU, BF16 operands, FP32 accumulation) over the
// index-hit features of vh_k1_features, trained on the GPU and used to choose which of the index top-K
// candidates reach exact rescoring. No reference counterpart: the reference rescoring package
// (vulcan/rescore/fast.py:198-254 GpuHGB, static.py:25-82) scores PSMs after K2; this network runs before K2.
//
// Kernels (all gfx950 MFMA v_mfma_f32_16x16x32_bf16, wave64):
//   mlp_tile           one wave evaluates 16*NB candidates. Layer 1 and 2 are computed transposed (neurons are
//                      the MFMA M dimension, candidates N), so each layer's accumulator tile is already the next
//                      layer's B operand: lane l holds neurons 4*(l/16)+r of candidate l%16, and two 16-neuron
//                      blocks fill the 8 bf16 of a 32-deep K slice. The K order this implies is folded into the
//                      packed W2 layout (vh_mlp_pack), so no data moves between layers. Layer 3 (256 -> 1) is
//                      fused into layer 2's epilogue: H2 never exists in memory or registers as a whole.
//   prefilter_score    persistent: W2 (128 KB bf16) + W1 + biases staged once per workgroup in LDS, then waves
//                      stream 64-candidate tiles; features are computed in registers from the candidate lists.
//   mlp_fwd            the same tile on a materialised feature matrix, optionally storing the activations
//                      needed by the backward pass.
//   gemm_bf16          64x64-tile MFMA GEMM (strided bf16 operands through LDS, split-K, optional ReLU mask
//                      epilogue) for the backward products dW2 = dZ2^T A1 and dZ1 = (dZ2 W2) * [A1 > 0].
//   head / l1 reductions, adam, select: VALU.
#include "vh_common.h"

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8_t;
typedef __attribute__((ext_vector_type(8))) unsigned short u16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2_t;
typedef __attribute__((ext_vector_type(2))) short i16x2;

// parameter vector (f32 master weights): W1 [256][8] (out, in), b1, W2 [256][256] (out, in), b2, w3, b3
#define TH_W1 0
#define TH_B1 2048
#define TH_W2 2304
#define TH_B2 67840
#define TH_W3 68096
#define TH_B3 68352
#define TH_N 68353

// packed model blob (bytes); [0, BL_LDS) is what the kernels stage in LDS
#define BL_W2P 0          // bf16 [16 mb][8 kslice][64 lane][8]
#define BL_W1P 131072     // bf16 [16 mb][64 lane][8] (lanes 16..63 zero: the K padding of layer 1)
#define BL_B1 147456      // f32 [256]
#define BL_B2 148480      // f32 [256]
#define BL_W3 149504      // f32 [256]
#define BL_B3 150528      // f32 (16 B slot)
#define BL_MU 150544      // f32 [8]
#define BL_ISD 150576     // f32 [8]
#define BL_LDS 150608
#define BL_W2B 150656     // bf16 [256][256] row-major (out, in), for the backward GEMM
#define BL_TOTAL 281728

static_assert(BL_TOTAL == VH_MLP_BLOB_BYTES, "blob size");
static_assert(TH_N == VH_MLP_N_PARAMS, "parameter count");

__device__ __forceinline__ unsigned short f2bf(float f) {   // round to nearest even (finite inputs)
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }

// two f32 -> packed bf16 (lo = a), round to nearest even: one v_cvt_pk_bf16_f32 on gfx950 (same bits as f2bf)
__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b) {
  typedef __attribute__((ext_vector_type(2))) float f32x2;
  const f32x2 v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
}

// max(x, 0.f) on the bit pattern: a float with the sign bit set is a negative int32 (one v_max_i32, no NaN
// canonicalisation; a positive NaN passes through)
__device__ __forceinline__ float relu_f32(float x) {
  const int b = __float_as_int(x);
  return __int_as_float(b > 0 ? b : 0);
}

// ReLU of two packed bf16: a bf16 with the sign bit set is a negative int16, so max(x, 0) as int16 pairs
// (v_pk_max_i16) zeroes it and keeps the rest. Rounding f32 -> bf16 keeps the sign, so this equals rounding
// max(x, 0.f).
__device__ __forceinline__ uint32_t relu_bf16x2(uint32_t w) {
  const i16x2 z = {0, 0};
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(i16x2, w), z));
}

__device__ __forceinline__ f32x4 mfma_bf16(u16x8 a, u16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b),
                                                 c, 0, 0, 0);
}

// ------------------------------------------------------------------------------------------------ features
struct FeatIn {
  int32_t stride;
  const int32_t *n_cand, *c_row, *c_hyp, *c_cnt, *s_maxcnt;
  const int16_t* c_nb;
  const uint32_t* c_iq;
  const int8_t* h_z;
  const float *h_M, *mass32;
  const uint8_t* length;
  const uint64_t* mask;
};

// the 8 raw features of slot c = s * stride + i of spectrum s (a filled slot); vh_prefilter_features documents
// them
__device__ __forceinline__ void slot_features(const FeatIn& F, int64_t s, int64_t c, float* f) {
  const int row = F.c_row[c], h = F.c_hyp[c], cnt = F.c_cnt[c], mx = F.s_maxcnt[s];
  const float m32 = F.mass32[row], M = F.h_M[h];
  const uint32_t iq = F.c_iq[c];
  f[0] = (float)cnt;
  f[1] = mx > 0 ? (float)cnt / (float)mx : 0.f;
  f[2] = (float)F.c_nb[c];
  f[3] = (float)(iq >> 16) * 0.00390625f;
  f[4] = (float)(iq & 0xFFFFu) * 0.00390625f;
  f[5] = ((M - m32) / m32) * 1000000.0f;
  f[6] = (float)F.length[row];
  f[7] = (float)F.h_z[h];
}

// the 8 raw features of candidate slot c (false for an empty slot)
__device__ __forceinline__ bool cand_features(const FeatIn& F, int64_t c, float* f) {
  const int64_t s = c / F.stride;
  const int i = (int)(c - s * F.stride);
  if (i >= F.n_cand[s]) return false;
  slot_features(F, s, c, f);
  return true;
}

__global__ __launch_bounds__(256) void features_kernel(FeatIn F, int64_t n_slots, float* out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_slots) return;
  float f[8];
  if (!cand_features(F, c, f))
    for (int j = 0; j < 8; ++j) f[j] = 0.f;
  for (int j = 0; j < 8; ++j) out[c * 8 + j] = f[j];
}

// normalised bf16 B operand of layer 1: lanes 0-15 carry the 8 features of their candidate, other lanes zero
__device__ __forceinline__ u16x8 feature_operand(const unsigned char* lds, const float* f, bool valid) {
  const float* mu = (const float*)(lds + BL_MU);
  const float* isd = (const float*)(lds + BL_ISD);
  u32x4 x = {0u, 0u, 0u, 0u};
  if (valid) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
      x[j] = pack_bf16x2((f[2 * j] - mu[2 * j]) * isd[2 * j], (f[2 * j + 1] - mu[2 * j + 1]) * isd[2 * j + 1]);
  }
  return __builtin_bit_cast(u16x8, x);
}

// ------------------------------------------------------------------------------------------------ MLP tile
// xb[nb]: B operand of candidates c0 + 16*nb + (lane & 15). score[nb] (every lane): that candidate's output.
// STORE: a1 (bf16 [n][256]) and a2 (f32 [n][256]) rows of candidates < n.
template <int NB, bool STORE>
__device__ __forceinline__ void mlp_tile(const unsigned char* lds, const u16x8 (&xb)[NB], float (&score)[NB],
                                         int64_t c0, int64_t n, unsigned short* a1, float* a2) {
  const int lane = lane_id(), g = lane >> 4, cl = lane & 15;
  const u16x8* w1p = (const u16x8*)(lds + BL_W1P);
  const u16x8* w2p = (const u16x8*)(lds + BL_W2P);
  const float* b1 = (const float*)(lds + BL_B1);
  const float* b2 = (const float*)(lds + BL_B2);
  const float* w3 = (const float*)(lds + BL_W3);
  const float b3 = *(const float*)(lds + BL_B3);
  // h1[nb][c] = bf16 pairs of K slice c: words 0-1 from neuron block 2c, words 2-3 from block 2c+1
  uint32_t h1[NB][8][4];
  // layer 1 in neuron blocks of 16: the MFMAs of block mb + 1 are issued before the epilogue (ReLU, bf16
  // packing) of block mb; the scheduling fences keep two blocks of accumulators live at most
  f32x4 la[2][NB];
  auto l1_mfma = [&](int mb, f32x4(&acc)[NB]) {
    const u16x8 a = w1p[mb * 64 + lane];
    const f32x4 bb = *(const f32x4*)(b1 + 16 * mb + 4 * g);   // the bias enters as the accumulator's start
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[nb] = mfma_bf16(a, xb[nb], bb);
  };
  l1_mfma(0, la[0]);
#pragma unroll
  for (int mb = 0; mb < 16; ++mb) {
    if (mb + 1 < 16) l1_mfma(mb + 1, la[(mb + 1) & 1]);
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const f32x4 acc = la[mb & 1][nb];
      const uint32_t w0 = relu_bf16x2(pack_bf16x2(acc[0], acc[1]));
      const uint32_t w1 = relu_bf16x2(pack_bf16x2(acc[2], acc[3]));
      h1[nb][mb >> 1][2 * (mb & 1)] = w0;
      h1[nb][mb >> 1][2 * (mb & 1) + 1] = w1;
      if (STORE) {
        const int64_t c = c0 + 16 * nb + cl;
        if (c < n) {
          uint32_t* d = (uint32_t*)(a1 + c * 256 + 16 * mb + 4 * g);
          d[0] = w0;
          d[1] = w1;
        }
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  float part[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) part[nb] = 0.f;
  // layer 2 in neuron blocks of 16, software-pipelined: the MFMAs of block mb2 + 1 are issued before the
  // epilogue (ReLU, layer-3 dot product) of block mb2, so the VALU work runs while the matrix core is busy
  auto l2_mfma = [&](int mb2, f32x4(&acc)[NB]) {
    const f32x4 bb = *(const f32x4*)(b2 + 16 * mb2 + 4 * g);   // the bias enters as the accumulator's start
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const u16x8 a = w2p[(mb2 * 8 + c) * 64 + lane];
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const u32x4 hw = {h1[nb][c][0], h1[nb][c][1], h1[nb][c][2], h1[nb][c][3]};
        acc[nb] = mfma_bf16(a, __builtin_bit_cast(u16x8, hw), c == 0 ? bb : acc[nb]);
      }
    }
  };
  auto l2_epi = [&](int mb2, const f32x4(&acc)[NB]) {
    const f32x4 ww = *(const f32x4*)(w3 + 16 * mb2 + 4 * g);
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      float v4[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float v = relu_f32(acc[nb][r]);
        v4[r] = v;
        part[nb] = fmaf(ww[r], v, part[nb]);
      }
      if (STORE) {
        const int64_t c = c0 + 16 * nb + cl;
        if (c < n) {
          float* d = a2 + c * 256 + 16 * mb2 + 4 * g;
          d[0] = v4[0]; d[1] = v4[1]; d[2] = v4[2]; d[3] = v4[3];
        }
      }
    }
  };
  f32x4 acc0[NB], acc1[NB];
  l2_mfma(0, acc0);
#pragma unroll 1
  for (int mb2 = 0; mb2 < 14; mb2 += 2) {
    l2_mfma(mb2 + 1, acc1);
    l2_epi(mb2, acc0);
    l2_mfma(mb2 + 2, acc0);
    l2_epi(mb2 + 1, acc1);
  }
  l2_mfma(15, acc1);
  l2_epi(14, acc0);
  l2_epi(15, acc1);
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    float p = part[nb];
    p = p + __shfl_xor(p, 16);
    p = p + __shfl_xor(p, 32);
    score[nb] = p + b3;
  }
}

__device__ __forceinline__ void stage_blob(unsigned char* lds, const unsigned char* blob) {
  const uint4* src = (const uint4*)blob;
  uint4* dst = (uint4*)lds;
  for (int i = threadIdx.x; i < BL_LDS / 16; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
}

#define MLP_NB 4
#define MLP_WAVES 8

template <bool STORE>
__global__ __launch_bounds__(64 * MLP_WAVES) void mlp_fwd_kernel(const unsigned char* __restrict__ blob, int64_t n,
                                                                 const float* __restrict__ X, float* s,
                                                                 unsigned short* a1, float* a2) {
  extern __shared__ __align__(16) unsigned char lds[];
  stage_blob(lds, blob);
  const int w = threadIdx.x >> 6, lane = lane_id(), g = lane >> 4, cl = lane & 15;
  const int64_t tile = 16 * MLP_NB, n_tiles = (n + tile - 1) / tile;
  for (int64_t t = (int64_t)blockIdx.x * MLP_WAVES + w; t < n_tiles; t += (int64_t)gridDim.x * MLP_WAVES) {
    const int64_t c0 = t * tile;
    u16x8 xb[MLP_NB];
#pragma unroll
    for (int nb = 0; nb < MLP_NB; ++nb) {
      const int64_t c = c0 + 16 * nb + cl;
      float f[8];
      const bool v = g == 0 && c < n;
      if (v)
        for (int j = 0; j < 8; ++j) f[j] = X[c * 8 + j];
      xb[nb] = feature_operand(lds, f, v);
    }
    float sc[MLP_NB];
    mlp_tile<MLP_NB, STORE>(lds, xb, sc, c0, n, a1, a2);
    if (g == 0) {
#pragma unroll
      for (int nb = 0; nb < MLP_NB; ++nb) {
        const int64_t c = c0 + 16 * nb + cl;
        if (c < n) s[c] = sc[nb];
      }
    }
  }
}

// Scoring runs over the filled slots only: off = exclusive prefix of the per-spectrum counts (off[n_spec] = the
// total), tile_spec[t] = the spectrum holding candidate 64 t of that compact order.
__global__ __launch_bounds__(256) void score_count_kernel(const int32_t* __restrict__ n_cand, int n_spec,
                                                          int stride, uint32_t* off) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > n_spec) return;
  const int v = s < n_spec ? n_cand[s] : 0;
  off[s] = (uint32_t)(v < 0 ? 0 : (v > stride ? stride : v));
}

__global__ __launch_bounds__(256) void score_tiles_kernel(const uint32_t* __restrict__ off, int n_spec,
                                                          uint32_t* tile_spec) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_spec) return;
  const uint32_t a = off[s], b = off[s + 1];
  for (uint32_t t = (a + 63u) >> 6; (t << 6) < b; ++t) tile_spec[t] = (uint32_t)s;
}

// Thread per filled slot in the compact order q: its slot index (cidx[q] = s * stride + i) and its normalised
// bf16 layer-1 operand (feat[q], 16 B: the 8 features as vh_prefilter_features, (f - mu) * isd).
__global__ __launch_bounds__(256) void score_feat_kernel(const unsigned char* __restrict__ blob, FeatIn F,
                                                         const uint32_t* __restrict__ off,
                                                         const uint32_t* __restrict__ tile_spec, int n_spec,
                                                         u32x4* feat, uint32_t* cidx) {
  const uint32_t total = off[n_spec];
  const int64_t q64 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q64 >= total) return;
  const uint32_t q = (uint32_t)q64;
  int s = (int)tile_spec[q >> 6];
  while (off[s + 1] <= q) ++s;   // spectra without candidates are passed over; off[n_spec] > q ends it
  const int64_t c = (int64_t)s * F.stride + (q - off[s]);
  float f[8];
  slot_features(F, s, c, f);
  feat[q] = __builtin_bit_cast(u32x4, feature_operand(blob, f, true));
  cidx[q] = (uint32_t)c;
}

// Persistent: waves take 64-candidate tiles of the compact order; lanes 0-15 load the four 16-candidate B
// operands of layer 1 (lanes 16-63 carry its zero K padding), and lane l writes the score of candidate 64 t + l.
// The next tile's operands are loaded before the current tile's MFMA work.
__global__ __launch_bounds__(64 * MLP_WAVES) void prefilter_score_kernel(const unsigned char* __restrict__ blob,
                                                                         const uint32_t* __restrict__ off,
                                                                         const u32x4* __restrict__ feat,
                                                                         const uint32_t* __restrict__ cidx,
                                                                         int n_spec, float* score) {
  extern __shared__ __align__(16) unsigned char lds[];
  stage_blob(lds, blob);
  const int w = threadIdx.x >> 6, lane = lane_id(), g = lane >> 4, cl = lane & 15;
  const uint32_t total = off[n_spec];
  const int64_t n_tiles = ((int64_t)total + 63) >> 6;
  const int64_t step = (int64_t)gridDim.x * MLP_WAVES;
  u32x4 nx[MLP_NB];
  uint32_t nc = 0;
  auto load = [&](int64_t t) {
    const uint32_t q0 = (uint32_t)(t << 6);
    const u32x4 zero4 = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int nb = 0; nb < MLP_NB; ++nb) {
      const uint32_t q = q0 + 16u * nb + (uint32_t)cl;
      nx[nb] = (g == 0 && q < total) ? feat[q] : zero4;
    }
    nc = q0 + (uint32_t)lane < total ? cidx[q0 + lane] : 0u;
  };
  int64_t t = (int64_t)blockIdx.x * MLP_WAVES + w;
  if (t < n_tiles) load(t);
  for (; t < n_tiles; t += step) {
    u16x8 xb[MLP_NB];
#pragma unroll
    for (int nb = 0; nb < MLP_NB; ++nb) xb[nb] = __builtin_bit_cast(u16x8, nx[nb]);
    const uint32_t c = nc;
    const bool ok = (uint32_t)(t << 6) + (uint32_t)lane < total;
    if (t + step < n_tiles) load(t + step);
    float sc[MLP_NB];
    mlp_tile<MLP_NB, false>(lds, xb, sc, 0, 0, nullptr, nullptr);
    const float mine = g == 0 ? sc[0] : g == 1 ? sc[1] : g == 2 ? sc[2] : sc[3];   // candidate 16 g + cl = lane
    if (ok) score[c] = mine;
  }
}

// ------------------------------------------------------------------------------------------------ packing
__global__ __launch_bounds__(256) void pack_kernel(const float* __restrict__ th, const float* __restrict__ mu,
                                                   const float* __restrict__ isd, unsigned char* blob) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;   // 0 .. 65535
  if (i >= 65536) return;
  {   // W2 in the MFMA A-operand order of mlp_tile
    const int mb = i >> 12, c = (i >> 9) & 7, l = (i >> 3) & 63, j = i & 7;
    const int row = 16 * mb + (l & 15);
    const int col = 32 * c + 16 * (j >> 2) + 4 * (l >> 4) + (j & 3);
    ((unsigned short*)(blob + BL_W2P))[i] = f2bf(th[TH_W2 + row * 256 + col]);
    ((unsigned short*)(blob + BL_W2B))[i] = f2bf(th[TH_W2 + i]);
  }
  if (i < 8192) {
    const int mb = i >> 9, l = (i >> 3) & 63, j = i & 7;
    ((unsigned short*)(blob + BL_W1P))[i] = l < 16 ? f2bf(th[TH_W1 + (16 * mb + l) * 8 + j]) : (unsigned short)0;
  }
  if (i < 256) {
    ((float*)(blob + BL_B1))[i] = th[TH_B1 + i];
    ((float*)(blob + BL_B2))[i] = th[TH_B2 + i];
    ((float*)(blob + BL_W3))[i] = th[TH_W3 + i];
  }
  if (i < 4) ((float*)(blob + BL_B3))[i] = i == 0 ? th[TH_B3] : 0.f;
  if (i < 8) {
    ((float*)(blob + BL_MU))[i] = mu[i];
    ((float*)(blob + BL_ISD))[i] = isd[i];
  }
  if (i < (BL_W2B - BL_LDS) / 4) ((float*)(blob + BL_LDS))[i] = 0.f;
}

// ------------------------------------------------------------------------------------------------ GEMM
struct GemmArgs {
  const unsigned short* A;   // A(m, k) = A[m * sam + k * sak]
  int64_t sam, sak;
  const unsigned short* B;   // B(k, n) = B[k * sbk + n * sbn]
  int64_t sbk, sbn;
  int64_t M, N, K, kchunk;
  float* C;                  // C[z * split_stride + m * ldc + n]
  int64_t ldc, split_stride;
  const unsigned short* mask;   // optional: C(m, n) = 0 where bf16 mask[m * ldm + n] <= 0
  int64_t ldm;
};

__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs G) {
  __shared__ __align__(16) unsigned short As[64][40];
  __shared__ __align__(16) unsigned short Bs[64][40];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, g = lane >> 4, cl = lane & 15;
  const int64_t m0 = (int64_t)blockIdx.x * 64, n0 = (int64_t)blockIdx.y * 64;
  const int64_t kb = (int64_t)blockIdx.z * G.kchunk;
  const int64_t ke = kb + G.kchunk < G.K ? kb + G.kchunk : G.K;
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  f32x4 acc[2][2] = {{zero, zero}, {zero, zero}};
  for (int64_t k0 = kb; k0 < ke; k0 += 32) {
    for (int e = tid; e < 2048; e += 256) {
      int mm, kk;
      if (G.sak == 1) { mm = e >> 5; kk = e & 31; } else { mm = e & 63; kk = e >> 6; }
      const int64_t gm = m0 + mm, gk = k0 + kk;
      As[mm][kk] = (gm < G.M && gk < ke) ? G.A[gm * G.sam + gk * G.sak] : (unsigned short)0;
      int nn, kq;
      if (G.sbn == 1) { nn = e & 63; kq = e >> 6; } else { nn = e >> 5; kq = e & 31; }
      const int64_t gn = n0 + nn, gq = k0 + kq;
      Bs[nn][kq] = (gn < G.N && gq < ke) ? G.B[gq * G.sbk + gn * G.sbn] : (unsigned short)0;
    }
    __syncthreads();
    u16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i] = *(const u16x8*)&As[32 * (w >> 1) + 16 * i + cl][8 * g];
#pragma unroll
    for (int j = 0; j < 2; ++j) b[j] = *(const u16x8*)&Bs[32 * (w & 1) + 16 * j + cl][8 * g];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = mfma_bf16(a[i], b[j], acc[i][j]);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + 32 * (w >> 1) + 16 * i + 4 * g + r;
        const int64_t col = n0 + 32 * (w & 1) + 16 * j + cl;
        if (row < G.M && col < G.N) {
          float v = acc[i][j][r];
          if (G.mask && !(bf2f(G.mask[row * G.ldm + col]) > 0.f)) v = 0.f;
          G.C[(int64_t)blockIdx.z * G.split_stride + row * G.ldc + col] = v;
        }
      }
}

// sum of `parts` partial arrays of length len (fixed order), written at out[o_index(i)]
__global__ __launch_bounds__(256) void reduce_parts_kernel(const float* __restrict__ part, int parts, int64_t len,
                                                           float* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  float acc = 0.f;
  for (int z = 0; z < parts; ++z) acc = acc + part[(int64_t)z * len + i];
  out[i] = acc;
}

// ------------------------------------------------------------------------------------------------ backward
#define BW_ROWS 256

// per chunk of BW_ROWS candidates, thread m: dz2 (bf16 out), partial dw3[m], db2[m], db3
__global__ __launch_bounds__(256) void bwd_head_kernel(const unsigned char* __restrict__ blob, int64_t n,
                                                       const float* __restrict__ s, const float* __restrict__ y,
                                                       const float* __restrict__ a2, unsigned short* dz2,
                                                       float* part) {
  const int m = threadIdx.x;
  const float w3m = ((const float*)(blob + BL_W3))[m];
  const float inv_n = 1.f / (float)n;
  const int64_t r0 = (int64_t)blockIdx.x * BW_ROWS;
  const int64_t r1 = r0 + BW_ROWS < n ? r0 + BW_ROWS : n;
  float sw3 = 0.f, sb2 = 0.f, sb3 = 0.f;
  for (int64_t r = r0; r < r1; ++r) {
    const float ds = (2.f * (s[r] - y[r])) * inv_n;
    const float a = a2[r * 256 + m];
    sw3 = sw3 + ds * a;
    const float d = a > 0.f ? ds * w3m : 0.f;
    sb2 = sb2 + d;
    dz2[r * 256 + m] = f2bf(d);
    sb3 = sb3 + ds;
  }
  float* p = part + (int64_t)blockIdx.x * 513;
  p[m] = sw3;
  p[256 + m] = sb2;
  if (m == 0) p[512] = sb3;
}

// per chunk, thread m: partial dW1[m][0..7] = sum dz1[r][m] * xb[r][j], db1[m] = sum dz1[r][m]
__global__ __launch_bounds__(256) void bwd_l1_kernel(const unsigned char* __restrict__ blob, int64_t n,
                                                     const float* __restrict__ X, const float* __restrict__ dz1,
                                                     float* part) {
  __shared__ float xs[BW_ROWS][8];
  const int m = threadIdx.x;
  const float* mu = (const float*)(blob + BL_MU);
  const float* isd = (const float*)(blob + BL_ISD);
  const int64_t r0 = (int64_t)blockIdx.x * BW_ROWS;
  const int64_t r1 = r0 + BW_ROWS < n ? r0 + BW_ROWS : n;
  for (int e = threadIdx.x; e < BW_ROWS * 8; e += 256) {
    const int64_t r = r0 + (e >> 3);
    const int j = e & 7;
    xs[e >> 3][j] = r < r1 ? bf2f(f2bf((X[r * 8 + j] - mu[j]) * isd[j])) : 0.f;
  }
  __syncthreads();
  float acc[9];
  for (int j = 0; j < 9; ++j) acc[j] = 0.f;
  for (int64_t r = r0; r < r1; ++r) {
    const float d = dz1[r * 256 + m];
    for (int j = 0; j < 8; ++j) acc[j] = acc[j] + d * xs[r - r0][j];
    acc[8] = acc[8] + d;
  }
  float* p = part + (int64_t)blockIdx.x * 2304;
  for (int j = 0; j < 8; ++j) p[m * 8 + j] = acc[j];
  p[2048 + m] = acc[8];
}

// fixed-order sums of the chunk partials into the gradient vector
__global__ __launch_bounds__(256) void bwd_finish_kernel(const float* __restrict__ ph, const float* __restrict__ pl,
                                                         int chunks, float* grad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 513) {
    float acc = 0.f;
    for (int z = 0; z < chunks; ++z) acc = acc + ph[(int64_t)z * 513 + i];
    if (i < 256) grad[TH_W3 + i] = acc;
    else if (i < 512) grad[TH_B2 + i - 256] = acc;
    else grad[TH_B3] = acc;
  }
  if (i < 2304) {
    float acc = 0.f;
    for (int z = 0; z < chunks; ++z) acc = acc + pl[(int64_t)z * 2304 + i];
    if (i < 2048) grad[TH_W1 + i] = acc;
    else grad[TH_B1 + i - 2048] = acc;
  }
}

__global__ __launch_bounds__(256) void adam_kernel(float* th, const float* __restrict__ gr, float* m1, float* m2,
                                                   int64_t n, float lr, float b1, float b2, float eps, float c1,
                                                   float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = gr[i];
  const float m = b1 * m1[i] + (1.f - b1) * g;
  const float v = b2 * m2[i] + (1.f - b2) * (g * g);
  m1[i] = m;
  m2[i] = v;
  const float mh = m * c1, vh = v * c2;
  th[i] = th[i] - lr * (mh / (sqrtf(vh) + eps));
}

// ------------------------------------------------------------------------------------------------ selection
__device__ __forceinline__ uint32_t score_ord(float x) {   // order-preserving u32 of an f32; NaN lowest
  if (x != x) return 0u;
  const uint32_t u = __float_as_uint(x);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

struct SelArgs {
  int32_t n_spec, stride_in, keep_min, keep_max, closure, out_stride, sort_n, clos_n, wave_bytes;
  float margin;
  const int32_t *n_cand, *c_row, *c_hyp, *c_cnt;
  const float* score;
  const int64_t* partner;
  int32_t *out_n, *out_row, *out_hyp, *out_cnt;
};

__global__ __launch_bounds__(256) void select_kernel(SelArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int w = threadIdx.x >> 6, lane = lane_id();
  const int s = blockIdx.x * (blockDim.x >> 6) + w;
  if (s >= A.n_spec) return;
  unsigned char* wb = smem + (size_t)w * A.wave_bytes;
  uint64_t* keys = (uint64_t*)wb;
  uint64_t* ck = keys + A.sort_n;
  int32_t* eh = (int32_t*)(ck + A.clos_n);
  int32_t* ec = eh + A.clos_n;
  const int64_t base = (int64_t)s * A.stride_in;
  int n = A.n_cand[s];
  n = n < A.stride_in ? n : A.stride_in;
  // keys (unsorted, slot order): score_ord << 32 | ~slot, so key order = score desc (NaN last), earlier slot first
  uint64_t mx = 0;
  for (int i = lane; i < n; i += 64) {
    const uint64_t k = ((uint64_t)score_ord(A.score[base + i]) << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)i);
    keys[i] = k;
    mx = k > mx ? k : mx;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t v = __shfl_xor(mx, o);
    mx = v > mx ? v : mx;
  }
  wave_sync();
  // kept: the best keep_min, then every candidate scoring within margin of the best, at most keep_max. Those
  // scoring >= best - margin are a prefix of the key order, so the kept set is the top nt keys.
  const float thr = (n > 0 ? A.score[base + (0xFFFFFFFFu - (uint32_t)mx)] : 0.f) - A.margin;
  int within = 0;
  for (int i0 = 0; i0 < n; i0 += 64) within += __popcll(ballot(i0 + lane < n && A.score[base + i0 + lane] >= thr));
  const int lim = n < A.keep_max ? n : A.keep_max;
  const int cnt = within < lim ? within : lim;
  const int lo = n < A.keep_min ? n : A.keep_min;
  const int nt = cnt > lo ? cnt : lo;
  // the nt-th largest score_ord T (bitwise search), then the lowest slots among the keys scoring exactly T
  uint32_t T = 0;
  int above = 0;   // keys with score_ord > T
  const int mode = nt >= n ? 0 : (nt == within ? 1 : 2);   // 0: all, 1: the prefix {score >= thr}, 2: search
  if (mode == 2) {
    constexpr int SR = 4;   // lists of up to 256 slots: the score_ords in registers, counted by ballot
    if (n <= 64 * SR) {
      uint32_t kr[SR];   // 0 past the fill: never counted (every candidate c below has a bit set)
#pragma unroll
      for (int j = 0; j < SR; ++j) kr[j] = j * 64 + lane < n ? (uint32_t)(keys[j * 64 + lane] >> 32) : 0u;
      for (int bit = 31; bit >= 0; --bit) {
        const uint32_t c = T | (1u << bit);
        int k = 0;
#pragma unroll
        for (int j = 0; j < SR; ++j) k += __popcll(ballot(kr[j] >= c));
        if (k >= nt) T = c;
      }
#pragma unroll
      for (int j = 0; j < SR; ++j) above += __popcll(ballot(kr[j] > T));
    } else {
      for (int bit = 31; bit >= 0; --bit) {
        const uint32_t c = T | (1u << bit);
        int k = 0;
        for (int i = lane; i < n; i += 64) k += (uint32_t)(keys[i] >> 32) >= c ? 1 : 0;
        if (wave_sum(k) >= nt) T = c;
      }
      int k = 0;
      for (int i = lane; i < n; i += 64) k += (uint32_t)(keys[i] >> 32) > T ? 1 : 0;
      above = wave_sum(k);
    }
  }
  // compaction of the kept slots (slot order) into closure entries
  int nk = 0, ties = 0;
  for (int i0 = 0; i0 < n; i0 += 64) {
    const int i = i0 + lane;
    bool keep = false;
    if (i < n) {
      if (mode == 0) keep = true;
      else if (mode == 1) keep = A.score[base + i] >= thr;
      else {
        const uint32_t o = (uint32_t)(keys[i] >> 32);
        keep = o > T;
        const bool tie = o == T;
        const uint64_t tb = ballot(tie);
        if (tie) keep = ties + __popcll(tb & lanemask_lt()) < nt - above;
        ties += __popcll(tb);
      }
    }
    const uint64_t kb = ballot(keep);
    if (keep) {
      const int e = nk + __popcll(kb & lanemask_lt());
      const int32_t row = A.c_row[base + i];
      eh[e] = A.c_hyp[base + i];
      ec[e] = A.c_cnt[base + i];
      ck[e] = ((uint64_t)(uint32_t)row << 32) | (uint64_t)e;
      if (A.closure) {
        eh[nt + e] = A.c_hyp[base + i];
        ec[nt + e] = -1;
        ck[nt + e] = ((uint64_t)(uint32_t)A.partner[row] << 32) | (uint64_t)(nt + e);
      }
    }
    nk += __popcll(kb);
  }
  const int ncl = A.closure ? 2 * nt : nt;
  int np2 = 1;
  while (np2 < ncl) np2 <<= 1;
  for (int i = ncl + lane; i < np2; i += 64) ck[i] = ~0ULL;
  wave_sync();
  wave_bitonic_u64(ck, np2, false);
  int nout = 0;
  const int64_t ob = (int64_t)s * A.out_stride;
  for (int i0 = 0; i0 < ncl; i0 += 64) {
    const int i = i0 + lane;
    bool first = false;
    uint32_t row = 0;
    int32_t hy = 0, cc = 0;
    if (i < ncl) {
      const uint64_t v = ck[i];
      row = (uint32_t)(v >> 32);
      first = i == 0 || (uint32_t)(ck[i - 1] >> 32) != row;
      if (first) {
        hy = eh[v & 0xFFFFFFFFu];
        cc = ec[v & 0xFFFFFFFFu];
        for (int j = i + 1; j < ncl && (uint32_t)(ck[j] >> 32) == row; ++j) {
          const uint32_t q = (uint32_t)(ck[j] & 0xFFFFFFFFu);
          hy = eh[q] < hy ? eh[q] : hy;
          cc = ec[q] > cc ? ec[q] : cc;
        }
      }
    }
    const uint64_t bal = ballot(first);
    if (first) {
      const int pos = nout + __popcll(bal & lanemask_lt());
      A.out_row[ob + pos] = (int32_t)row;
      A.out_hyp[ob + pos] = hy;
      A.out_cnt[ob + pos] = cc;
    }
    nout += __popcll(bal);
  }
  if (lane == 0) A.out_n[s] = nout;
}

// ------------------------------------------------------------------------------------------------ launchers
static int pow2_ge(int x) {
  int n = 1;
  while (n < x) n <<= 1;
  return n;
}

static int mlp_grid(int64_t tiles) {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  int64_t need = (tiles + MLP_WAVES - 1) / MLP_WAVES;
  int64_t g = need < cus ? need : cus;
  return g < 1 ? 1 : (int)g;
}

// the MLP kernels stage 135 KB of LDS per workgroup (gfx950: 160 KB per CU)
static int allow_big_lds() {
  static int done = 0;
  if (done) return 0;
  hipError_t e = hipFuncSetAttribute((const void*)mlp_fwd_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     BL_LDS);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)mlp_fwd_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize, BL_LDS);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)prefilter_score_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            BL_LDS);
  if (e != hipSuccess) return VH_E_LIMIT;
  done = 1;
  return 0;
}

static FeatIn make_feat(const VhIndexView* ix, int32_t stride, const int32_t* n_cand, const int32_t* c_row,
                        const int32_t* c_hyp, const int32_t* c_cnt, const int16_t* c_nb, const uint32_t* c_iq,
                        const int32_t* s_maxcnt, const int8_t* h_z, const float* h_M) {
  FeatIn F;
  F.stride = stride;
  F.n_cand = n_cand; F.c_row = c_row; F.c_hyp = c_hyp; F.c_cnt = c_cnt; F.s_maxcnt = s_maxcnt;
  F.c_nb = c_nb; F.c_iq = c_iq;
  F.h_z = h_z; F.h_M = h_M;
  F.mass32 = ix->mass32; F.length = ix->length; F.mask = (const uint64_t*)ix->mask;
  return F;
}

static int64_t bw_chunks(int64_t n) { return (n + BW_ROWS - 1) / BW_ROWS; }
static int bw_splits(int64_t n) {
  int64_t s = (n + 1023) / 1024;
  return (int)(s < 1 ? 1 : (s > 16 ? 16 : s));
}

// Count-only features of index top-K lists: zero b-ion counts and intensity pseudo-logs of the filled slots and
// the best count per spectrum (wave per spectrum).
__global__ __launch_bounds__(256) void count_features_kernel(int n_spec, int stride, const int32_t* __restrict__ n_cand,
                                                             const int32_t* __restrict__ cnt, int16_t* nb,
                                                             uint32_t* iq, int32_t* maxcnt) {
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6), lane = lane_id();
  if (s >= n_spec) return;
  const int nc = n_cand[s] < stride ? n_cand[s] : stride;
  const int64_t b = (int64_t)s * stride;
  int m = 0;
  for (int i = lane; i < nc; i += 64) {
    const int c = cnt[b + i];
    m = c > m ? c : m;
    nb[b + i] = 0;
    iq[b + i] = 0u;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int x = __shfl_xor(m, o);
    m = x > m ? x : m;
  }
  if (lane == 0) maxcnt[s] = m;
}

extern "C" {

int vh_mlp_pack(const float* theta, const float* mu, const float* isd, void* blob, void* stream) {
  if (!theta || !mu || !isd || !blob) return VH_E_ARG;
  pack_kernel<<<256, 256, 0, (hipStream_t)stream>>>(theta, mu, isd, (unsigned char*)blob);
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_mlp_forward(const void* blob, int64_t n, const float* x, float* s, uint16_t* a1, float* a2, void* stream) {
  if (!blob || n < 0 || (n && (!x || !s)) || ((a1 == nullptr) != (a2 == nullptr))) return VH_E_ARG;
  if (n == 0) return 0;
  if (allow_big_lds()) return VH_E_LIMIT;
  const int64_t tiles = (n + 16 * MLP_NB - 1) / (16 * MLP_NB);
  const int g = mlp_grid(tiles);
  hipStream_t st = (hipStream_t)stream;
  if (a1)
    mlp_fwd_kernel<true><<<g, 64 * MLP_WAVES, BL_LDS, st>>>((const unsigned char*)blob, n, x, s, a1, a2);
  else
    mlp_fwd_kernel<false><<<g, 64 * MLP_WAVES, BL_LDS, st>>>((const unsigned char*)blob, n, x, s, nullptr, nullptr);
  VH_CHECK_LAUNCH();
  return 0;
}

size_t vh_mlp_workspace_bytes(int64_t n) {
  if (n < 0) return 0;
  const int64_t ch = bw_chunks(n);
  size_t b = 0;
  b += (size_t)n * 256 * 2;                      // dz2 bf16
  b = (b + 255) & ~(size_t)255;
  b += (size_t)n * 256 * 4;                      // dz1 f32
  b += (size_t)bw_splits(n) * 65536 * 4;         // dW2 split-K partials
  b += (size_t)ch * 513 * 4 + (size_t)ch * 2304 * 4;
  return b + 256;
}

int vh_mlp_backward(const void* blob, int64_t n, const float* x, const uint16_t* a1, const float* a2, const float* s,
                    const float* y, float* grad, void* ws, size_t ws_bytes, void* stream) {
  if (!blob || n < 1 || !x || !a1 || !a2 || !s || !y || !grad || !ws) return VH_E_ARG;
  if (ws_bytes < vh_mlp_workspace_bytes(n)) return VH_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const unsigned char* B = (const unsigned char*)blob;
  unsigned char* p = (unsigned char*)ws;
  unsigned short* dz2 = (unsigned short*)p;
  size_t off = ((size_t)n * 256 * 2 + 255) & ~(size_t)255;
  float* dz1 = (float*)(p + off);
  off += (size_t)n * 256 * 4;
  const int ks = bw_splits(n);
  float* pw2 = (float*)(p + off);
  off += (size_t)ks * 65536 * 4;
  const int64_t ch = bw_chunks(n);
  float* ph = (float*)(p + off);
  off += (size_t)ch * 513 * 4;
  float* pl = (float*)(p + off);
  if (ch > 0x7fffffff) return VH_E_LIMIT;
  bwd_head_kernel<<<(unsigned)ch, 256, 0, st>>>(B, n, s, y, a2, dz2, ph);
  VH_CHECK_LAUNCH();
  // dW2[m][k] = sum_r dz2[r][m] * a1[r][k]
  GemmArgs G = {};
  G.A = dz2; G.sam = 1; G.sak = 256;
  G.B = a1; G.sbk = 256; G.sbn = 1;
  G.M = 256; G.N = 256; G.K = n;
  G.kchunk = (((n + ks - 1) / ks) + 31) / 32 * 32;
  G.C = pw2; G.ldc = 256; G.split_stride = 65536;
  gemm_bf16_kernel<<<dim3(4, 4, ks), 256, 0, st>>>(G);
  VH_CHECK_LAUNCH();
  reduce_parts_kernel<<<256, 256, 0, st>>>(pw2, ks, 65536, grad + TH_W2);
  VH_CHECK_LAUNCH();
  // dz1[r][k] = (sum_m dz2[r][m] * W2[m][k]) * [a1[r][k] > 0]
  GemmArgs H = {};
  H.A = dz2; H.sam = 256; H.sak = 1;
  H.B = (const unsigned short*)(B + BL_W2B); H.sbk = 256; H.sbn = 1;
  H.M = n; H.N = 256; H.K = 256; H.kchunk = 256;
  H.C = dz1; H.ldc = 256; H.split_stride = 0;
  H.mask = a1; H.ldm = 256;
  gemm_bf16_kernel<<<dim3((unsigned)((n + 63) / 64), 4, 1), 256, 0, st>>>(H);
  VH_CHECK_LAUNCH();
  bwd_l1_kernel<<<(unsigned)ch, 256, 0, st>>>(B, n, x, dz1, pl);
  VH_CHECK_LAUNCH();
  bwd_finish_kernel<<<9, 256, 0, st>>>(ph, pl, (int)ch, grad);
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_mlp_adam(float* theta, const float* grad, float* m1, float* m2, int64_t n, float lr, float beta1,
                float beta2, float eps, float c1, float c2, void* stream) {
  if (!theta || !grad || !m1 || !m2 || n < 0) return VH_E_ARG;
  if (n == 0) return 0;
  adam_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(theta, grad, m1, m2, n, lr, beta1, beta2,
                                                                          eps, c1, c2);
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_prefilter_count_features(int32_t n_spec, int32_t stride, const int32_t* n_cand, const int32_t* c_cnt,
                                int16_t* c_nb, uint32_t* c_iq, int32_t* s_maxcnt, void* stream) {
  if (n_spec < 0 || stride < 1) return VH_E_ARG;
  if (n_spec == 0) return 0;
  if (!n_cand || !c_cnt || !c_nb || !c_iq || !s_maxcnt) return VH_E_ARG;
  count_features_kernel<<<(unsigned)((n_spec + 3) / 4), 256, 0, (hipStream_t)stream>>>(n_spec, stride, n_cand, c_cnt,
                                                                                       c_nb, c_iq, s_maxcnt);
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_prefilter_features(const VhIndexView* ix, int32_t n_spec, int32_t stride, const int32_t* n_cand,
                          const int32_t* c_row, const int32_t* c_hyp, const int32_t* c_cnt, const int16_t* c_nb,
                          const uint32_t* c_iq, const int32_t* s_maxcnt, const int8_t* h_z, const float* h_M,
                          float* out, void* stream) {
  if (!ix || n_spec < 0 || stride < 1 || !out) return VH_E_ARG;
  if (n_spec == 0) return 0;
  if (!n_cand || !c_row || !c_hyp || !c_cnt || !c_nb || !c_iq || !s_maxcnt || !h_z || !h_M) return VH_E_ARG;
  const int64_t slots = (int64_t)n_spec * stride;
  FeatIn F = make_feat(ix, stride, n_cand, c_row, c_hyp, c_cnt, c_nb, c_iq, s_maxcnt, h_z, h_M);
  features_kernel<<<(unsigned)((slots + 255) / 256), 256, 0, (hipStream_t)stream>>>(F, slots, out);
  VH_CHECK_LAUNCH();
  return 0;
}

static size_t score_a256(size_t b) { return (b + 255) & ~(size_t)255; }

// workspace: off [n_spec + 1], tile_spec [tiles], feat [slots] (16 B), cidx [slots], scan scratch
size_t vh_prefilter_score_workspace_bytes(int32_t n_spec, int32_t stride) {
  if (n_spec < 0 || stride < 1) return 0;
  const int64_t slots = (int64_t)n_spec * stride, tiles = (slots + 63) / 64 + 1;
  return score_a256((size_t)(n_spec + 1) * 4) + score_a256((size_t)tiles * 4) + score_a256((size_t)slots * 16) +
         score_a256((size_t)slots * 4) + vh_scan_ws_u32(n_spec + 1) + 256;
}

int vh_prefilter_score(const VhIndexView* ix, const void* blob, int32_t n_spec, int32_t stride,
                       const int32_t* n_cand, const int32_t* c_row, const int32_t* c_hyp, const int32_t* c_cnt,
                       const int16_t* c_nb, const uint32_t* c_iq, const int32_t* s_maxcnt, const int8_t* h_z,
                       const float* h_M, float* score, void* ws, size_t ws_bytes, void* stream) {
  if (!ix || !blob || n_spec < 0 || stride < 1 || !score) return VH_E_ARG;
  if (n_spec == 0) return 0;
  if (!n_cand || !c_row || !c_hyp || !c_cnt || !c_nb || !c_iq || !s_maxcnt || !h_z || !h_M || !ws) return VH_E_ARG;
  if ((int64_t)n_spec * stride >= ((int64_t)1 << 32) - 64) return VH_E_LIMIT;
  if (ws_bytes < vh_prefilter_score_workspace_bytes(n_spec, stride)) return VH_E_WORKSPACE;
  FeatIn F = make_feat(ix, stride, n_cand, c_row, c_hyp, c_cnt, c_nb, c_iq, s_maxcnt, h_z, h_M);
  if (allow_big_lds()) return VH_E_LIMIT;
  hipStream_t st = (hipStream_t)stream;
  unsigned char* p = (unsigned char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  const int64_t slots = (int64_t)n_spec * stride, tiles = (slots + 63) / 64 + 1;
  uint32_t* off = (uint32_t*)p;
  p += score_a256((size_t)(n_spec + 1) * 4);
  uint32_t* tile_spec = (uint32_t*)p;
  p += score_a256((size_t)tiles * 4);
  u32x4* feat = (u32x4*)p;
  p += score_a256((size_t)slots * 16);
  uint32_t* cidx = (uint32_t*)p;
  p += score_a256((size_t)slots * 4);
  void* sws = p;
  score_count_kernel<<<(unsigned)((n_spec + 256) / 256), 256, 0, st>>>(n_cand, n_spec, stride, off);
  VH_CHECK_LAUNCH();
  int rc = vh_scan_exclusive_u32(off, (int64_t)n_spec + 1, sws, vh_scan_ws_u32(n_spec + 1), st);
  if (rc) return rc;
  score_tiles_kernel<<<(unsigned)((n_spec + 255) / 256), 256, 0, st>>>(off, n_spec, tile_spec);
  VH_CHECK_LAUNCH();
  score_feat_kernel<<<(unsigned)((slots + 255) / 256), 256, 0, st>>>((const unsigned char*)blob, F, off, tile_spec,
                                                                     n_spec, feat, cidx);
  VH_CHECK_LAUNCH();
  const int g = mlp_grid(tiles);
  prefilter_score_kernel<<<g, 64 * MLP_WAVES, BL_LDS, st>>>((const unsigned char*)blob, off, feat, cidx, n_spec,
                                                            score);
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_prefilter_select(const VhIndexView* ix, int32_t n_spec, int32_t stride_in, const int32_t* n_cand,
                        const int32_t* c_row, const int32_t* c_hyp, const int32_t* c_cnt, const float* score,
                        int32_t keep_min, int32_t keep_max, float margin, int32_t pair_closure, int32_t out_stride,
                        int32_t* out_n, int32_t* out_row, int32_t* out_hyp, int32_t* out_cnt, void* stream) {
  if (!ix || n_spec < 0 || stride_in < 1 || stride_in > 1024 || keep_min < 1 || keep_max < keep_min ||
      keep_max > 512 || !(margin >= 0.f))
    return VH_E_ARG;
  const int keep = keep_max;
  if (pair_closure && !ix->partner) return VH_E_ARG;
  if (out_stride < (pair_closure ? 2 * keep : keep)) return VH_E_CAPACITY;
  if (n_spec == 0) return 0;
  if (!n_cand || !c_row || !c_hyp || !c_cnt || !score || !out_n || !out_row || !out_hyp || !out_cnt)
    return VH_E_ARG;
  SelArgs A;
  A.n_spec = n_spec; A.stride_in = stride_in; A.keep_min = keep_min; A.keep_max = keep_max; A.margin = margin;
  A.closure = pair_closure ? 1 : 0;
  A.out_stride = out_stride;
  A.sort_n = pow2_ge(stride_in < 64 ? 64 : stride_in);
  const int k = keep < stride_in ? keep : stride_in;
  A.clos_n = pow2_ge((pair_closure ? 2 * k : k) < 64 ? 64 : (pair_closure ? 2 * k : k));
  A.wave_bytes = A.sort_n * 8 + A.clos_n * 16;
  A.n_cand = n_cand; A.c_row = c_row; A.c_hyp = c_hyp; A.c_cnt = c_cnt; A.score = score;
  A.partner = ix->partner;
  A.out_n = out_n; A.out_row = out_row; A.out_hyp = out_hyp; A.out_cnt = out_cnt;
  // one wave (spectrum) per workgroup, as the index top-K select: a workgroup holds its CU slot until its
  // slowest spectrum ends
  const int waves = 1;
  if (A.wave_bytes > 65536) return VH_E_LIMIT;
  const unsigned g = (unsigned)((n_spec + waves - 1) / waves);
  select_kernel<<<g, 64 * waves, (size_t)A.wave_bytes * waves, (hipStream_t)stream>>>(A);
  VH_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
