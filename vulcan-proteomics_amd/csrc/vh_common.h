// Shared device helpers for the gfx950 search kernels (wave64, no FMA contraction: built with
// -ffp-contract=off so every f32 operation rounds exactly like the reference's Sage arithmetic).
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vulcan_hip.h"

#define VH_PROTON 1.0072764f    // host.py:10
#define VH_NEUTRON 1.00335f     // host.py:11
#define VH_WAVE 64

// Per-call device parameters (by value). Residue masses are pre-added on the host in f32
// (mono + static, mono + var: the same single rounding as kernels.cu:47-51).
struct DevP {
  float rm[2][26];
  float nterm;
  int tol_da;
  float tol_lo, tol_hi, fmin, fmax;
  int min_ion_index;
  float prec_lo, prec_hi;
  int iso_lo, iso_hi, z_lo, z_hi, max_fragment_charge;
  int max_peaks, min_peaks, min_matched_peaks, topk, k1_threshold, pair_closure, tile, xcd_group;
  int n_offsets;
  float offsets[VH_MAX_OFFSETS];
  int open_search;
  float open_lo, open_hi;
  int k1_threshold_open;
  int peak_window_k;       // > 0: K0 window view (top peak_window_k per floor(m/z / peak_window_th))
  float peak_window_th;
};

struct DevIdx {
  int64_t n_rows;
  const int64_t* seq_off;
  const uint8_t* length;
  const uint64_t* mask;
  const float* mass32;
  const uint8_t* residues;
  const int64_t* partner;
  int bucket_log2, bin_shift;
  uint32_t bin_base;
  int n_bins;
  int64_t n_buckets;
  const uint32_t* offsets;
  const uint32_t* postings;
};

DevP vh_make_devp(const VhParams* p);
DevIdx vh_make_devidx(const VhIndexView* ix);
int vh_scan_inclusive_i64(int64_t* a, int64_t n, void* ws, size_t ws_bytes, hipStream_t st);
int vh_scan_exclusive_u32(uint32_t* a, int64_t n, void* ws, size_t ws_bytes, hipStream_t st);
size_t vh_scan_ws_u32(int64_t n);
size_t vh_scan_ws_i64(int64_t n);
size_t vh_radix_ws_bytes(int64_t n);
int vh_radix_sort_u64(uint64_t* a, uint64_t* b, int64_t n, int bit_lo, int bit_hi, void* ws, size_t ws_bytes,
                      hipStream_t st, uint64_t** result);

// Blocks are dispatched to the 8 XCDs round-robin (block b -> XCD b % 8). Work items (spectra in precursor-mass
// order) are handed out in groups of 8*c blocks: inside a group, XCD x takes the contiguous sub-range
// [x*c, (x+1)*c) of the group's positions. Each XCD's L2 then serves a narrow band of index buckets and table
// rows, while all XCDs move through the mass range together (no XCD is left with only the heaviest spectra).
// c = 0 keeps the plain block order.
#define VH_XCDS 8
__device__ __forceinline__ int xcd_band_position(int c) {
  if (c <= 0) return blockIdx.x;
  const int G = VH_XCDS * c;
  const int b = blockIdx.x;
  const int j = b % G;
  return (b / G) * G + (j % VH_XCDS) * c + j / VH_XCDS;
}
static inline unsigned xcd_band_grid(int64_t n_blocks, int c) {
  if (c <= 0) return (unsigned)n_blocks;
  const int64_t G = (int64_t)VH_XCDS * c;
  return (unsigned)(((n_blocks + G - 1) / G) * G);
}

#define VH_CHECK_LAUNCH() do { hipError_t e_ = hipGetLastError(); if (e_ != hipSuccess) return (int)e_; } while (0)

// x / 1e6 rounded to nearest, as the IEEE division gives it, without the division's instruction sequence
// (scale, reciprocal, four fmas, fixup): q0 = x * RN(1e-6), the residual x - q0 * 1e6 exactly by an fma, and one
// correction q0 + residual * RN(1e-6) (Markstein's correction step). Compared on the host with the division for
// every f32 x (tests/div1e6_check.c): equal for every finite |x| >= 2^-90; smaller |x| (quotients near the
// subnormal range), zeros, infinities and NaN take the division.
__device__ __forceinline__ float div_1e6(float x) {
  const float ax = fabsf(x);
  if (!(ax >= 0x1p-90f && ax <= 0x1.fffffep127f)) return x / 1000000.0f;
  const float r = 1e-6f;
  const float q0 = x * r;
  return fmaf(fmaf(-q0, 1000000.0f, x), r, q0);
}

// kernels.cu:36-38: ppm bound = m + (m*tol)/1e6 (each op rounded), Da bound = m + tol
__device__ __forceinline__ float tol_bound(float m, float tol, int da) {
  return da ? (m + tol) : (m + div_1e6(m * tol));
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Order all LDS traffic of this wave before/after (waves own private LDS regions; no block barrier needed).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave-wide bitonic sort of a[0, n) in LDS (n a power of two, wave-private): every stage visits its n / 2
// compare pairs (i with bit j clear, i | j), so no lane idles on the upper element of a pair.
__device__ __forceinline__ void wave_bitonic_u64(uint64_t* a, int n, bool descending) {
  const int lane = lane_id();
  const int half = n >> 1;
  for (int k = 2; k <= n; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int p = lane; p < half; p += 64) {
        const int i = ((p & ~(j - 1)) << 1) | (p & (j - 1));
        const uint64_t x = a[i], y = a[i | j];
        const bool up = ((i & k) == 0) != descending;
        if (up ? (x > y) : (x < y)) {
          a[i] = y;
          a[i | j] = x;
        }
      }
      wave_sync();
    }
}

__device__ __forceinline__ uint64_t ballot(bool b) { return __ballot(b); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ULL << lane_id()) - 1ULL; }

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ uint32_t fbits(float x) { return __float_as_uint(x); }

// Residue table in LDS: rm[bit][letter] (bit = variable-mod bit of the mask).
__device__ __forceinline__ void load_rm(float* lds_rm, const DevP& P) {
  for (int i = threadIdx.x; i < 52; i += blockDim.x) lds_rm[i] = (&P.rm[0][0])[i];
}
