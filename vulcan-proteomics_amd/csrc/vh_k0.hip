// K0 peak view and precursor hypotheses on the device.
//
// K0 restates k0_preprocess (vulcan/gpu/engine.py:171-210; the window view is at k0w_*): keep fmin <= m/z <= fmax (not NaN),
// mass = m/z - PROTON, the top max_peaks peaks by (intensity desc with NaN first, mass desc, input order),
// then mass ascending (ties keep that order); TIC = f64 sum; valid = kept >= min_peaks.
// One wave64 per spectrum: in-range peaks are compacted (ballot prefix, order preserved) and held in registers,
// up to 8 per lane. The selection order is the 64-bit key (intensity key << 32 | mass bits) descending, ties
// by input position, where the intensity key maps NaN above every number (torch's descending sort puts NaN
// first) and -0 onto +0, and the mass key is the order-preserving u32 image of the float. A peak's rank = number
// of peaks before it; candidates are broadcast with v_readlane (no LDS traffic in the O(n^2) passes). The kept
// peaks (rank < max_peaks) then get their mass-ordered slot as #(kept with (mass bits, rank) smaller). Spectra
// with more than 512 in-range peaks rank straight from global memory. A spectrum whose in-range masses are
// strictly ascending (m/z-sorted input) takes a fast path: a 64-step bitwise search for the max_peaks-th key and
// a ballot compaction in input order give the same kept set and slots without the O(n^2) passes.
// Hypotheses restate make_hypotheses (vulcan/gpu/host.py:64-91) with the same f32 operation order.
#include <mutex>

#include "vh_common.h"

#define K0_WAVES 4
#define K0_LDS 512       // in-range peaks staged per wave
#define K0_SEL 512       // max_peaks limit of the device path (kept peaks are staged in the same LDS arrays)

__device__ __forceinline__ bool inrange(float m, const DevP& P) { return !isnan(m) && m >= P.fmin && m <= P.fmax; }
__device__ __forceinline__ bool gt_nan_first(float a, float b) { return (isnan(a) && !isnan(b)) || a > b; }
__device__ __forceinline__ bool eq_nan(float a, float b) { return a == b || (isnan(a) && isnan(b)); }
// j precedes i in the selection order
__device__ __forceinline__ bool sel_before(float ij, float mj, int j, float ii, float mi, int i) {
  if (!eq_nan(ij, ii)) return gt_nan_first(ij, ii);
  if (mj != mi) return mj > mi;
  return j < i;
}

__global__ __launch_bounds__(256) void k0_count_kernel(DevP P, int n_spec, const float* __restrict__ mz,
                                                       const int64_t* __restrict__ raw_off, int64_t* pk_off) {
  int s = blockIdx.x * K0_WAVES + (threadIdx.x >> 6);
  if (blockIdx.x == 0 && threadIdx.x == 0) pk_off[0] = 0;
  if (s >= n_spec) return;
  int64_t a = raw_off[s], b = raw_off[s + 1];
  int c = 0;
  for (int64_t i0 = a; i0 < b; i0 += 4 * 64) {   // four chunks' loads in flight together
    float m[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t i = i0 + u * 64 + lane_id();
      m[u] = i < b ? mz[i] : __builtin_nanf("");
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) c += inrange(m[u], P);
  }
  c = wave_sum(c);
  if (lane_id() == 0) pk_off[s + 1] = c < P.max_peaks ? c : P.max_peaks;
}

__device__ __forceinline__ uint32_t int_key(float x) {
  if (isnan(x)) return 0xFFFFFFFFu;
  uint32_t b = __float_as_uint(x == 0.f ? 0.f : x);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// order-preserving u32 image of a finite float (and its inverse), for the mass keys
__device__ __forceinline__ uint32_t ord_key(float x) {
  uint32_t b = __float_as_uint(x == 0.f ? 0.f : x);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ord_inv(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

__device__ __forceinline__ uint32_t rdl(uint32_t v, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, l); }

#define K0_REG 8   // 8 x 64 = 512 in-range peaks in registers

__global__ __launch_bounds__(256) void k0_emit_kernel(DevP P, int n_spec, const float* __restrict__ mz,
                                                      const float* __restrict__ it, const int64_t* __restrict__ raw_off,
                                                      const int64_t* __restrict__ pk_off, float* pk_mass, float* pk_int,
                                                      float* tic, uint8_t* valid) {
  __shared__ float s_m[K0_WAVES][K0_LDS], s_i[K0_WAVES][K0_LDS];
  __shared__ uint32_t s_h[K0_WAVES][256];
  const int w = threadIdx.x >> 6, lane = lane_id();
  int s = blockIdx.x * K0_WAVES + w;
  if (s >= n_spec) return;
  float* lm = s_m[w];
  float* li = s_i[w];
  const int64_t base = raw_off[s];
  const int n = (int)(raw_off[s + 1] - base);
  // compact in-range peaks (order preserved); the loads of up to K0_REG chunks are issued together
  int n_in = 0;
  for (int i00 = 0; i00 < n; i00 += 64 * K0_REG) {
    float mv[K0_REG], iv[K0_REG];
#pragma unroll
    for (int c = 0; c < K0_REG; ++c) {
      const int i = i00 + c * 64 + lane;
      mv[c] = i < n ? mz[base + i] : 0.f;
      iv[c] = i < n ? it[base + i] : 0.f;
    }
#pragma unroll
    for (int c = 0; c < K0_REG; ++c) {
      const int i = i00 + c * 64 + lane;
      const bool ok = i < n && inrange(mv[c], P);
      const uint64_t bal = ballot(ok);
      const int pos = n_in + __popcll(bal & lanemask_lt());
      if (ok && pos < K0_LDS) { lm[pos] = mv[c] - VH_PROTON; li[pos] = iv[c]; }
      n_in += __popcll(bal);
    }
  }
  wave_sync();
  const int K = n_in < P.max_peaks ? n_in : P.max_peaks;
  const int64_t o = pk_off[s];
  double t = 0.0;
  if (n_in <= K0_LDS) {
    const int nch = (n_in + 63) >> 6;
    uint32_t mb[K0_REG], ik[K0_REG], rk[K0_REG];
    float fi[K0_REG];
#pragma unroll
    for (int c = 0; c < K0_REG; ++c) {
      int d = c * 64 + lane;
      bool v = d < n_in;
      float m = v ? lm[d] : 0.f;
      fi[c] = v ? li[d] : 0.f;
      mb[c] = ord_key(m);
      ik[c] = int_key(fi[c]);
      rk[c] = 0;
    }
    // Fast path: masses strictly ascending in input order (the usual m/z-sorted spectrum), so the mass order
    // of the kept peaks is their input order. The kept set {rank < K} is found by a bitwise search for the
    // K-th largest selection key T: kept = key > T, plus the first K - #(key > T) peaks with key == T.
    bool unsorted = false;
    for (int d = lane; d + 1 < n_in; d += 64) unsorted |= !(lm[d] < lm[d + 1]);
    if (!ballot(unsorted)) {
      bool kept[K0_REG];
      if (K == n_in) {
#pragma unroll
        for (int c = 0; c < K0_REG; ++c) kept[c] = c * 64 + lane < n_in;
      } else {
        // the intensity half of T first (32 passes); its low (mass) half is needed only when the peaks whose
        // intensity key equals that half are more than the places left: otherwise T is the smallest of them
        // Ti = the K-th largest intensity key, by a radix select over its four bytes (high first): per byte, an
        // LDS histogram of the keys that share the bytes fixed so far, then the largest byte t whose keys >= t
        // (plus those already above the prefix) number K or more
        uint32_t Ti = 0;
        int above = 0;
        uint32_t* hist = s_h[w];
        for (int d = 3; d >= 0; --d) {
          const int sh = 8 * d;
          const uint32_t hm = d == 3 ? 0u : (0xFFFFFFFFu << (sh + 8));
#pragma unroll
          for (int q = 0; q < 4; ++q) hist[q * 64 + lane] = 0u;
          wave_sync();
#pragma unroll
          for (int c = 0; c < K0_REG; ++c)
            if (c < nch && c * 64 + lane < n_in && (ik[c] & hm) == Ti) atomicAdd(&hist[(ik[c] >> sh) & 255u], 1u);
          wave_sync();
          uint32_t h4[4];   // this lane's bytes 4 lane .. 4 lane + 3
#pragma unroll
          for (int q = 0; q < 4; ++q) h4[q] = hist[4 * lane + q];
          uint32_t suf = h4[0] + h4[1] + h4[2] + h4[3];   // -> keys with byte >= 4 lane
#pragma unroll
          for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = (uint32_t)__shfl_down((int)suf, o);
            if (lane + o < 64) suf += y;
          }
          int tb = -1;   // largest byte of this lane with above + #(byte >= it) >= K (the condition is monotone)
          uint32_t run = suf;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            if (above + (int)run >= K) tb = 4 * lane + q;
            run -= h4[q];
          }
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) {
            const int y = __shfl_xor(tb, o);
            tb = y > tb ? y : tb;
          }
          int gt = 0;   // keys of this prefix with byte > tb
#pragma unroll
          for (int q = 0; q < 4; ++q) gt += 4 * lane + q > tb ? (int)h4[q] : 0;
          above += wave_sum(gt);
          Ti |= (uint32_t)tb << sh;
          wave_sync();   // the next byte's histogram reuses the bins
        }
        int gti = 0, eqi = 0;
        uint32_t mlo = 0xFFFFFFFFu;
#pragma unroll
        for (int c = 0; c < K0_REG; ++c) {
          if (c < nch) {
            const bool v = c * 64 + lane < n_in;
            gti += __popcll(ballot(v && ik[c] > Ti));
            eqi += __popcll(ballot(v && ik[c] == Ti));
            if (v && ik[c] == Ti && mb[c] < mlo) mlo = mb[c];
          }
        }
        for (int off = 32; off > 0; off >>= 1) {
          const uint32_t x = (uint32_t)__shfl_xor((int)mlo, off);
          mlo = x < mlo ? x : mlo;
        }
        uint64_t T = ((uint64_t)Ti << 32) | mlo;   // every tie at Ti is kept
        if (gti + eqi > K) {                       // some ties at Ti are not: the mass half by bitwise search
          T = (uint64_t)Ti << 32;
          for (int bit = 31; bit >= 0; --bit) {
            const uint64_t cand = T | (1ULL << bit);
            int cnt = 0;
#pragma unroll
            for (int c = 0; c < K0_REG; ++c)
              if (c < nch)
                cnt += __popcll(ballot(c * 64 + lane < n_in && (((uint64_t)ik[c] << 32) | mb[c]) >= cand));
            if (cnt >= K) T = cand;
          }
        }
        int gt = 0;
#pragma unroll
        for (int c = 0; c < K0_REG; ++c)
          if (c < nch) gt += __popcll(ballot(c * 64 + lane < n_in && (((uint64_t)ik[c] << 32) | mb[c]) > T));
        const int need = K - gt;
        int eq_before = 0;
#pragma unroll
        for (int c = 0; c < K0_REG; ++c) {
          const uint64_t key = ((uint64_t)ik[c] << 32) | mb[c];
          const bool v = c < nch && c * 64 + lane < n_in;
          const uint64_t eb = ballot(v && key == T);
          kept[c] = v && (key > T || (key == T && eq_before + __popcll(eb & lanemask_lt()) < need));
          eq_before += __popcll(eb);
        }
      }
      int slot = 0;
#pragma unroll
      for (int c = 0; c < K0_REG; ++c) {
        const uint64_t kb = ballot(c < nch && kept[c]);
        if (c < nch && kept[c]) {
          const int ps = slot + __popcll(kb & lanemask_lt());
          pk_mass[o + ps] = ord_inv(mb[c]);
          pk_int[o + ps] = fi[c];
          t += (double)fi[c];
        }
        slot += __popcll(kb);
      }
      t = wave_sum(t);
      if (lane == 0) {
        tic[s] = (float)t;
        valid[s] = K >= P.min_peaks;
      }
      return;
    }
    // rank = #(j before i): key_j > key_i, or equal keys and j < i
#pragma unroll
    for (int cj = 0; cj < K0_REG; ++cj) {
      if (cj < nch) {
        const int lim = n_in - cj * 64 < 64 ? n_in - cj * 64 : 64;
        for (int l = 0; l < lim; ++l) {
          const uint64_t kj = ((uint64_t)rdl(ik[cj], l) << 32) | rdl(mb[cj], l);
          const int j = cj * 64 + l;
#pragma unroll
          for (int ci = 0; ci < K0_REG; ++ci) {
            if (ci < nch) {
              const uint64_t ki = ((uint64_t)ik[ci] << 32) | mb[ci];
              rk[ci] += (kj > ki) || (kj == ki && j < ci * 64 + lane);
            }
          }
        }
      }
    }
    // slot among the kept peaks: (mass bits, rank) ascending
    uint32_t ps[K0_REG];
#pragma unroll
    for (int c = 0; c < K0_REG; ++c) ps[c] = 0;
#pragma unroll
    for (int cj = 0; cj < K0_REG; ++cj) {
      if (cj < nch) {
        const int lim = n_in - cj * 64 < 64 ? n_in - cj * 64 : 64;
        for (int l = 0; l < lim; ++l) {
          const uint32_t rj = rdl(rk[cj], l);
          if (rj >= (uint32_t)K) continue;   // uniform: only kept peaks count
          const uint64_t kj = ((uint64_t)rdl(mb[cj], l) << 32) | rj;
#pragma unroll
          for (int ci = 0; ci < K0_REG; ++ci)
            if (ci < nch) ps[ci] += kj < (((uint64_t)mb[ci] << 32) | rk[ci]);
        }
      }
    }
#pragma unroll
    for (int c = 0; c < K0_REG; ++c) {
      if (c < nch && c * 64 + lane < n_in && rk[c] < (uint32_t)K) {
        pk_mass[o + ps[c]] = ord_inv(mb[c]);
        pk_int[o + ps[c]] = fi[c];
        t += (double)fi[c];
      }
    }
  } else {
    // more than K0_LDS in-range peaks: rank over the raw peaks from global memory, kept ones staged by rank
    float* sm = lm;
    float* si = li;
    for (int i = lane; i < n; i += 64) {
      float mzi = mz[base + i];
      if (!inrange(mzi, P)) continue;
      float mi = mzi - VH_PROTON, ii = it[base + i];
      int r = 0;
      for (int j = 0; j < n; ++j) {
        float mzj = mz[base + j];
        if (inrange(mzj, P)) r += sel_before(it[base + j], mzj - VH_PROTON, j, ii, mi, i);
      }
      if (r < K) { sm[r] = mi; si[r] = ii; }
    }
    wave_sync();
    for (int i = lane; i < K; i += 64) {
      float mi = sm[i];
      int pos = 0;
      for (int j = 0; j < K; ++j) {
        float mj = sm[j];
        pos += (mj < mi) || (mj == mi && j < i);
      }
      pk_mass[o + pos] = mi;
      pk_int[o + pos] = si[i];
      t += (double)si[i];
    }
  }
  t = wave_sum(t);
  if (lane == 0) {
    tic[s] = (float)t;
    valid[s] = K >= P.min_peaks;
  }
}

// ------------------------------------------------------------------------------------------------ hypotheses
__device__ __forceinline__ int n_hyp_of(const DevP& P, double pmz, int8_t z, uint8_t ok) {
  if (!ok || !isfinite(pmz)) return 0;
  int niso = P.iso_hi - P.iso_lo + 1;
  int nz = z > 0 ? 1 : (P.z_hi - P.z_lo + 1);
  int extra = (P.iso_lo <= 0 && P.iso_hi >= 0) ? nz * (P.n_offsets + (P.open_search ? 1 : 0)) : 0;
  return niso * nz + extra;
}

// an empty batch still gets its leading CSR offset (hyp_off[0] / pk_off[0] = 0): a kernel, not a memset, so the
// call stays capturable in a graph
__global__ void k0_zero_i64_kernel(int64_t* p) { *p = 0; }

__global__ __launch_bounds__(256) void hyp_count_kernel(DevP P, int n_spec, const double* __restrict__ pmz,
                                                        const int8_t* __restrict__ charge,
                                                        const uint8_t* __restrict__ valid, int64_t* hyp_off) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0) hyp_off[0] = 0;
  if (s >= n_spec) return;
  hyp_off[s + 1] = n_hyp_of(P, pmz[s], charge[s], valid[s]);
}

// Window bounds of every hypothesis, thread per hypothesis: hyp_emit_kernel leaves the f32 mass bounds in h_lo /
// h_hi (bit patterns); they become lower_bound(lo) / upper_bound(hi) over mass32, the two searches advancing in
// lockstep so their loads are in flight together.
__global__ __launch_bounds__(256) void hyp_bounds_kernel(int n_spec, const int64_t* __restrict__ hyp_off,
                                                         const float* __restrict__ mass32, int64_t n_rows,
                                                         int64_t cap, int64_t* h_lo, int64_t* h_hi) {
  const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= cap || h >= hyp_off[n_spec]) return;
  const float lo = __int_as_float((int)h_lo[h]), hi = __int_as_float((int)h_hi[h]);
  int64_t L0 = 0, R0 = n_rows, L1 = 0, R1 = n_rows;
  while (L0 < R0 || L1 < R1) {
    if (L0 < R0) {
      const int64_t M = (L0 + R0) >> 1;
      if (mass32[M] < lo) L0 = M + 1; else R0 = M;
    }
    if (L1 < R1) {
      const int64_t M = (L1 + R1) >> 1;
      if (!(hi < mass32[M])) L1 = M + 1; else R1 = M;
    }
  }
  h_lo[h] = L0;
  h_hi[h] = L1;
}

__global__ __launch_bounds__(256) void hyp_emit_kernel(DevP P, int n_spec, const double* __restrict__ pmz,
                                                       const int8_t* __restrict__ charge,
                                                       const int64_t* __restrict__ hyp_off,
                                                       const float* __restrict__ mass32, int64_t n_rows,
                                                       int32_t* h_spec, int8_t* h_z, int8_t* h_iso, float* h_pm,
                                                       float* h_M, int64_t* h_lo, int64_t* h_hi, int8_t* h_maxfz,
                                                       uint8_t* h_mode, float* h_off) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_spec) return;
  int64_t o = hyp_off[s];
  if (hyp_off[s + 1] == o) return;
  int zc = charge[s];
  int zlo = zc > 0 ? zc : P.z_lo, zhi = zc > 0 ? zc : P.z_hi;
  float mz = (float)pmz[s] - VH_PROTON;
  for (int z = zlo; z <= zhi; ++z) {
    for (int iso = P.iso_lo; iso <= P.iso_hi; ++iso, ++o) {
      float pm = mz * (float)z;
      float ie = (float)iso * VH_NEUTRON;
      float M = pm - ie;
      float lo = M + (M * P.prec_lo) / 1000000.0f;
      float hi = M + (M * P.prec_hi) / 1000000.0f;
      int mf = P.max_fragment_charge ? (z < P.max_fragment_charge + 1 ? z : P.max_fragment_charge + 1) : z;
      h_spec[o] = s;
      h_z[o] = (int8_t)z;
      h_iso[o] = (int8_t)iso;
      h_pm[o] = pm;
      h_M[o] = M;
      h_lo[o] = (int64_t)__float_as_int(lo);   // mass bounds; rows by hyp_bounds_kernel
      h_hi[o] = (int64_t)__float_as_int(hi);
      h_maxfz[o] = (int8_t)(mf > 2 ? mf : 2);
      if (h_mode) { h_mode[o] = VH_MODE_CLOSED; h_off[o] = 0.f; }
    }
  }
  if (!(P.iso_lo <= 0 && P.iso_hi >= 0)) return;
  // mass-offset then open hypotheses of the isotope-0 base (vulcan/ptm/search.py:56-76)
  const int nmode = P.n_offsets + (P.open_search ? 1 : 0);
  for (int q = 0; q < nmode; ++q) {
    const bool open = q == P.n_offsets;
    for (int z = zlo; z <= zhi; ++z, ++o) {
      float pm = mz * (float)z;
      float M0 = pm;   // isotope 0: pm - 0 * NEUTRON
      float lo, hi;
      if (!open) {
        float Mo = M0 - P.offsets[q];
        lo = Mo + (Mo * P.prec_lo) / 1000000.0f;
        hi = Mo + (Mo * P.prec_hi) / 1000000.0f;
      } else {
        lo = M0 - P.open_hi;
        hi = M0 - P.open_lo;
      }
      int mf = P.max_fragment_charge ? (z < P.max_fragment_charge + 1 ? z : P.max_fragment_charge + 1) : z;
      h_spec[o] = s;
      h_z[o] = (int8_t)z;
      h_iso[o] = 0;
      h_pm[o] = pm;
      h_M[o] = M0;
      h_lo[o] = (int64_t)__float_as_int(lo);
      h_hi[o] = (int64_t)__float_as_int(hi);
      h_maxfz[o] = (int8_t)(mf > 2 ? mf : 2);
      h_mode[o] = open ? VH_MODE_OPEN : VH_MODE_OFFSET;
      h_off[o] = open ? __int_as_float(0x7fc00000) : P.offsets[q];
    }
  }
}

// ------------------------------------------------------------------------------------------------ spectrum order
// Mass order for locality only (results do not depend on it): spectra by their first window row in
// 2^ORDER_BITS equal row bands (two radix passes), spectra without hypotheses last; the sort is stable, so a band
// keeps spectrum order.
#define ORDER_BITS 16
__global__ __launch_bounds__(256) void order_keys_kernel(int n_spec, int64_t n_rows, const int64_t* __restrict__ hyp_off,
                                                         const int64_t* __restrict__ h_lo, uint64_t* keys) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_spec) return;
  int64_t h0 = hyp_off[s];
  const uint64_t top = (1ULL << ORDER_BITS) - 1ULL;
  uint64_t band = top;
  if (hyp_off[s + 1] > h0 && n_rows > 0) {
    const int64_t lo = h_lo[h0] < 0 ? 0 : (h_lo[h0] >= n_rows ? n_rows - 1 : h_lo[h0]);
    band = ((uint64_t)lo << ORDER_BITS) / (uint64_t)n_rows;   // lo < n_rows < 2^47
  }
  keys[s] = (band << 32) | (uint64_t)(uint32_t)s;
}

// work-ordered key: mass band of the first window, then estimated K1 work descending, then the spectrum
__global__ __launch_bounds__(256) void order_work_keys_kernel(int n_spec, const int64_t* __restrict__ hyp_off,
                                                              const int64_t* __restrict__ h_lo,
                                                              const int64_t* __restrict__ h_hi,
                                                              const int8_t* __restrict__ h_maxfz,
                                                              const int64_t* __restrict__ pk_off, int band_log2,
                                                              uint64_t* keys) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_spec) return;
  const int64_t h0 = hyp_off[s], h1 = hyp_off[s + 1];
  const int64_t np = pk_off[s + 1] - pk_off[s];
  int64_t zsum = 0, rows = 0;
  for (int64_t h = h0; h < h1; ++h) {
    zsum += h_maxfz[h] > 1 ? h_maxfz[h] - 1 : 1;
    rows += h_hi[h] - h_lo[h];
  }
  int64_t est = np * zsum + rows / 4;
  est = est < 0xFFFFF ? est : 0xFFFFF;
  const uint64_t band = h1 > h0 ? ((uint64_t)h_lo[h0] >> band_log2) : 0xFFFFFULL;
  keys[s] = ((band < 0xFFFFFULL ? band : 0xFFFFFULL) << 44) | ((uint64_t)(0xFFFFF - est) << 24) | (uint64_t)s;
}

// the spectrum sits in the low `bits` bits of a sorted key (32 for the mass order, 24 for the work order)
__global__ __launch_bounds__(256) void order_out_kernel(int n_spec, const uint64_t* __restrict__ keys, int bits,
                                                        int32_t* order) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_spec) order[i] = (int32_t)(keys[i] & ((1ULL << bits) - 1ULL));
}

static size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

// K0 = count kernel, inclusive scan of the per-spectrum kept counts into pk_off (scan partials in `ws`), emit
// kernel. Shared by vh_k0 (caller workspace) and vh_k0_noalloc (the static scratch below).
// ---- window view (peak_window_k > 0; engine.py:182-196 with peak_mode "window"): a peak is kept when fewer than
// peak_window_k in-range peaks of its window floor(m/z / peak_window_th) (f32 division) precede it in the selection
// order of the global view; there is no total cap. The kept peaks are then ordered by mass, ties by (window,
// selection order), the order the reference's stable sorts leave. One wave per spectrum, O(n^2) comparisons over
// its raw peaks (the global view's register paths do not apply: the rank counts only the peak's own window). The
// emit pass keeps the kept flags of the first K0W_BITS raw peaks in LDS and recomputes the flag of later ones.
#define K0W_BITS 16384

__device__ __forceinline__ int k0w_window(float m, const DevP& P) { return (int)floorf(m / P.peak_window_th); }

__device__ bool k0w_kept(const DevP& P, const float* __restrict__ mz, const float* __restrict__ it, int n, int i) {
  const float mi = mz[i];
  if (!inrange(mi, P)) return false;
  const int wi = k0w_window(mi, P);
  const float ii = it[i], ai = mi - VH_PROTON;
  int r = 0;
  for (int j = 0; j < n && r < P.peak_window_k; ++j) {
    const float mj = mz[j];
    if (j == i || !inrange(mj, P) || k0w_window(mj, P) != wi) continue;
    r += sel_before(it[j], mj - VH_PROTON, j, ii, ai, i) ? 1 : 0;
  }
  return r < P.peak_window_k;
}

__global__ __launch_bounds__(256) void k0w_count_kernel(DevP P, int n_spec, const float* __restrict__ mz,
                                                        const float* __restrict__ it,
                                                        const int64_t* __restrict__ raw_off, int64_t* pk_off) {
  const int s = blockIdx.x * K0_WAVES + (threadIdx.x >> 6);
  if (blockIdx.x == 0 && threadIdx.x == 0) pk_off[0] = 0;
  if (s >= n_spec) return;
  const int64_t a = raw_off[s];
  const int n = (int)(raw_off[s + 1] - a);
  int c = 0;
  for (int i = lane_id(); i < n; i += 64) c += k0w_kept(P, mz + a, it + a, n, i) ? 1 : 0;
  c = wave_sum(c);
  if (lane_id() == 0) pk_off[s + 1] = c < P.max_peaks ? c : P.max_peaks;
}

__global__ __launch_bounds__(256) void k0w_emit_kernel(DevP P, int n_spec, const float* __restrict__ mz,
                                                       const float* __restrict__ it,
                                                       const int64_t* __restrict__ raw_off,
                                                       const int64_t* __restrict__ pk_off, float* pk_mass,
                                                       float* pk_int, float* tic, uint8_t* valid) {
  __shared__ uint32_t s_kept[K0_WAVES][K0W_BITS / 32];
  const int w = threadIdx.x >> 6, lane = lane_id();
  const int s = blockIdx.x * K0_WAVES + w;
  if (s >= n_spec) return;
  const int64_t a = raw_off[s];
  const int n = (int)(raw_off[s + 1] - a);
  const float* m = mz + a;
  const float* v = it + a;
  uint32_t* kb = s_kept[w];
  const int nb = n < K0W_BITS ? n : K0W_BITS;
  for (int i0 = 0; i0 < nb; i0 += 64) {   // kept flags of the first K0W_BITS peaks, one ballot per 64
    const int i = i0 + lane;
    const uint64_t bal = ballot(i < nb && k0w_kept(P, m, v, n, i));
    if (lane == 0) {
      kb[i0 >> 5] = (uint32_t)bal;
      if ((i0 >> 5) + 1 < K0W_BITS / 32) kb[(i0 >> 5) + 1] = (uint32_t)(bal >> 32);
    }
  }
  wave_sync();
  auto kept = [&](int j) -> bool { return j < K0W_BITS ? ((kb[j >> 5] >> (j & 31)) & 1u) : k0w_kept(P, m, v, n, j); };
  const int64_t o = pk_off[s];
  const int K = (int)(pk_off[s + 1] - o);
  double t = 0.0;
  for (int i = lane; i < n; i += 64) {
    if (!kept(i)) continue;
    const float mi = m[i], ai = mi - VH_PROTON, ii = v[i];
    const int wi = k0w_window(mi, P);
    int slot = 0;   // kept peaks before it in (mass, window, selection order)
    for (int j = 0; j < n; ++j) {
      if (j == i || !kept(j)) continue;
      const float aj = m[j] - VH_PROTON;
      if (aj < ai) ++slot;
      else if (aj == ai) {
        const int wj = k0w_window(m[j], P);
        slot += (wj < wi || (wj == wi && sel_before(v[j], aj, j, ii, ai, i))) ? 1 : 0;
      }
    }
    if (slot < K) {
      pk_mass[o + slot] = ai;
      pk_int[o + slot] = ii;
      t += (double)ii;
    }
  }
  t = wave_sum(t);
  if (lane == 0) {
    tic[s] = (float)t;
    valid[s] = K >= P.min_peaks;
  }
}

static int k0_launch(const VhParams* p, int32_t n_spec, const float* mz, const float* intensity,
                     const int64_t* raw_off, int64_t* pk_off, float* pk_mass, float* pk_int, float* tic,
                     uint8_t* valid, int64_t pk_capacity, void* ws, size_t ws_bytes, hipStream_t st) {
  if (!p || n_spec < 0 || !raw_off || !pk_off || !tic || !valid) return VH_E_ARG;
  if (p->max_peaks < 0 || p->max_peaks > K0_SEL) return VH_E_LIMIT;
  if (pk_capacity < (int64_t)n_spec * p->max_peaks) return VH_E_CAPACITY;
  if (n_spec == 0) {
    k0_zero_i64_kernel<<<1, 1, 0, st>>>(pk_off);
    VH_CHECK_LAUNCH();
    return 0;
  }
  if (p->peak_window_k < 0 || (p->peak_window_k > 0 && !(p->peak_window_th > 0.f))) return VH_E_ARG;
  DevP P = vh_make_devp(p);
  unsigned g = (unsigned)((n_spec + K0_WAVES - 1) / K0_WAVES);
  if (p->peak_window_k > 0) {
    k0w_count_kernel<<<g, 64 * K0_WAVES, 0, st>>>(P, n_spec, mz, intensity, raw_off, pk_off);
    VH_CHECK_LAUNCH();
    int rc = vh_scan_inclusive_i64(pk_off + 1, n_spec, ws, ws_bytes, st);
    if (rc) return rc;
    k0w_emit_kernel<<<g, 64 * K0_WAVES, 0, st>>>(P, n_spec, mz, intensity, raw_off, pk_off, pk_mass, pk_int, tic,
                                                 valid);
    VH_CHECK_LAUNCH();
    return 0;
  }
  k0_count_kernel<<<g, 64 * K0_WAVES, 0, st>>>(P, n_spec, mz, raw_off, pk_off);
  VH_CHECK_LAUNCH();
  int rc = vh_scan_inclusive_i64(pk_off + 1, n_spec, ws, ws_bytes, st);
  if (rc) return rc;
  k0_emit_kernel<<<g, 64 * K0_WAVES, 0, st>>>(P, n_spec, mz, intensity, raw_off, pk_off, pk_mass, pk_int, tic,
                                              valid);
  VH_CHECK_LAUNCH();
  return 0;
}

// Scratch of vh_k0_noalloc: the scan partials of up to VH_K0_NOALLOC_MAX_SPECTRA spectra (one i64 per 2,048
// spectra + 256 bytes of alignment slack, the size vh_scan_workspace_bytes returns for that count). One copy per
// GPU, alive until the process ends; every call on a GPU reuses it (the partials are fully rewritten by each
// scan, so nothing needs clearing), hence calls that overlap on one GPU race.
#define K0_SCRATCH_BYTES ((size_t)(VH_K0_NOALLOC_MAX_SPECTRA / 2048) * 8 + 256)
static __device__ __align__(256) unsigned char scratch[K0_SCRATCH_BYTES];

#define K0_MAX_DEVICES 64
static std::mutex k0_scratch_mu;
static void* k0_scratch_ptr[K0_MAX_DEVICES];

// Device address of `scratch` on the current device, looked up once per device with that device current. The
// host-side symbol is only a stand-in: kernels and scans get this address, never the array's name.
static int k0_scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= K0_MAX_DEVICES) return VH_E_LIMIT;
  std::lock_guard<std::mutex> lock(k0_scratch_mu);
  if (!k0_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    if (!ptr) return VH_E_ARG;
    k0_scratch_ptr[dev] = ptr;
  }
  *out = k0_scratch_ptr[dev];
  return 0;
}

extern "C" {

size_t vh_spectrum_order_workspace_bytes(int32_t n_spec) {
  size_t n = (size_t)(n_spec > 0 ? n_spec : 1);
  return 2 * a256(n * 8) + vh_radix_ws_bytes((int64_t)n);
}

int vh_spectrum_order(int32_t n_spec, int64_t n_rows, const int64_t* hyp_off, const int64_t* h_lo, int32_t* order,
                      void* ws, size_t ws_bytes, void* stream) {
  if (n_spec < 0 || n_rows < 0 || (n_spec && (!hyp_off || !h_lo || !order || !ws))) return VH_E_ARG;
  if (n_rows >= ((int64_t)1 << 47)) return VH_E_LIMIT;
  if (n_spec == 0) return 0;
  if (ws_bytes < vh_spectrum_order_workspace_bytes(n_spec)) return VH_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  size_t kb = a256((size_t)n_spec * 8);
  uint64_t* a = (uint64_t*)ws;
  uint64_t* b = (uint64_t*)((char*)ws + kb);
  unsigned g = (unsigned)((n_spec + 255) / 256);
  order_keys_kernel<<<g, 256, 0, st>>>(n_spec, n_rows, hyp_off, h_lo, a);
  VH_CHECK_LAUNCH();
  uint64_t* res = a;
  int rc = vh_radix_sort_u64(a, b, n_spec, 32, 32 + ORDER_BITS, (char*)ws + 2 * kb, ws_bytes - 2 * kb, st, &res);
  if (rc) return rc;
  order_out_kernel<<<g, 256, 0, st>>>(n_spec, res, 32, order);
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_spectrum_order_work(int32_t n_spec, const int64_t* hyp_off, const int64_t* h_lo, const int64_t* h_hi,
                           const int8_t* h_maxfz, const int64_t* pk_off, int32_t band_log2, int32_t* order, void* ws,
                           size_t ws_bytes, void* stream) {
  if (n_spec < 0 || band_log2 < 0 || band_log2 > 40) return VH_E_ARG;
  if (n_spec && (!hyp_off || !h_lo || !h_hi || !h_maxfz || !pk_off || !order || !ws)) return VH_E_ARG;
  if (n_spec == 0) return 0;
  if (n_spec >= (1 << 24)) return VH_E_LIMIT;
  if (ws_bytes < vh_spectrum_order_workspace_bytes(n_spec)) return VH_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  size_t kb = a256((size_t)n_spec * 8);
  uint64_t* a = (uint64_t*)ws;
  uint64_t* b = (uint64_t*)((char*)ws + kb);
  unsigned g = (unsigned)((n_spec + 255) / 256);
  order_work_keys_kernel<<<g, 256, 0, st>>>(n_spec, hyp_off, h_lo, h_hi, h_maxfz, pk_off, band_log2, a);
  VH_CHECK_LAUNCH();
  uint64_t* res = a;
  int rc = vh_radix_sort_u64(a, b, n_spec, 24, 64, (char*)ws + 2 * kb, ws_bytes - 2 * kb, st, &res);
  if (rc) return rc;
  order_out_kernel<<<g, 256, 0, st>>>(n_spec, res, 24, order);
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_k0(const VhParams* p, int32_t n_spec, const float* mz, const float* intensity, const int64_t* raw_off,
          int64_t* pk_off, float* pk_mass, float* pk_int, float* tic, uint8_t* valid, int64_t pk_capacity,
          void* ws, size_t ws_bytes, void* stream) {
  return k0_launch(p, n_spec, mz, intensity, raw_off, pk_off, pk_mass, pk_int, tic, valid, pk_capacity, ws,
                   ws_bytes, (hipStream_t)stream);
}

int vh_k0_noalloc(const VhParams* p, int32_t n_spec, const float* mz, const float* intensity,
                  const int64_t* raw_off, int64_t* pk_off, float* pk_mass, float* pk_int, float* tic, uint8_t* valid,
                  int64_t pk_capacity, void* stream) {
  if (n_spec < 0) return VH_E_ARG;
  if (n_spec > VH_K0_NOALLOC_MAX_SPECTRA) return VH_E_LIMIT;
  void* ws = nullptr;
  const int rc = k0_scratch_address(&ws);
  if (rc) return rc;
  return k0_launch(p, n_spec, mz, intensity, raw_off, pk_off, pk_mass, pk_int, tic, valid, pk_capacity, ws,
                   K0_SCRATCH_BYTES, (hipStream_t)stream);
}

int vh_hypotheses(const VhParams* p, int32_t n_spec, const double* precursor_mz, const int8_t* charge,
                  const uint8_t* valid, const float* mass32, int64_t n_rows, int64_t* hyp_off, int32_t* h_spec,
                  int8_t* h_z, int8_t* h_iso, float* h_pm, float* h_M, int64_t* h_lo, int64_t* h_hi,
                  int8_t* h_maxfz, uint8_t* h_mode, float* h_off, int64_t hyp_capacity, void* ws, size_t ws_bytes,
                  void* stream) {
  if (!p || n_spec < 0 || !hyp_off) return VH_E_ARG;
  int niso = p->iso_hi - p->iso_lo + 1, nz = p->z_hi - p->z_lo + 1;
  if (niso < 1 || nz < 1) return VH_E_ARG;
  if (p->n_offsets < 0 || p->n_offsets > VH_MAX_OFFSETS) return VH_E_ARG;
  const bool modes = p->n_offsets > 0 || p->open_search;
  if (modes && (!h_mode || !h_off)) return VH_E_ARG;
  int maxh = niso * nz + nz * (p->n_offsets + (p->open_search ? 1 : 0));
  if (maxh > VH_MAX_HYP_PER_SPEC) return VH_E_LIMIT;
  if (hyp_capacity < (int64_t)n_spec * maxh) return VH_E_CAPACITY;
  if (n_spec == 0) {
    k0_zero_i64_kernel<<<1, 1, 0, (hipStream_t)stream>>>(hyp_off);
    VH_CHECK_LAUNCH();
    return 0;
  }
  hipStream_t st = (hipStream_t)stream;
  DevP P = vh_make_devp(p);
  unsigned g = (unsigned)((n_spec + 255) / 256);
  hyp_count_kernel<<<g, 256, 0, st>>>(P, n_spec, precursor_mz, charge, valid, hyp_off);
  VH_CHECK_LAUNCH();
  int rc = vh_scan_inclusive_i64(hyp_off + 1, n_spec, ws, ws_bytes, st);
  if (rc) return rc;
  hyp_emit_kernel<<<g, 256, 0, st>>>(P, n_spec, precursor_mz, charge, hyp_off, mass32, n_rows, h_spec, h_z, h_iso,
                                     h_pm, h_M, h_lo, h_hi, h_maxfz, h_mode, h_off);
  VH_CHECK_LAUNCH();
  const int64_t cap = (int64_t)n_spec * maxh;   // <= hyp_capacity
  hyp_bounds_kernel<<<(unsigned)((cap + 255) / 256), 256, 0, st>>>(n_spec, hyp_off, mass32, n_rows, cap, h_lo, h_hi);
  VH_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
