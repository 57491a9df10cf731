// Per-spectrum top-K machinery shared by the K1 kernels (tile K1 and the select stage of the bucket-join K1):
// wave-wide bitonic sorts, the ballot-compacted top-K buffer, the feature-mode intensity quantisation and the
// final top-K / pair closure / dedupe stage. Top-K key = count (clamped at 32767) << 48 | row << 20 |
// (63 - local hypothesis) << 14, so the buffer order is count desc, row desc, lower hypothesis first
// (select_topk of the reference, vulcan/gpu/engine.py:284-312).
#pragma once
#include "vh_common.h"

__device__ void bitonic_sort_u64(uint64_t* a, int n, bool descending) {
  const int lane = lane_id();
  for (int k = 2; k <= n; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < n; i += 64) {
        int ixj = i ^ j;
        if (ixj > i) {
          uint64_t x = a[i], y = a[ixj];
          bool up = ((i & k) == 0) != descending;   // ascending run in this block
          if (up ? (x > y) : (x < y)) {
            a[i] = y;
            a[ixj] = x;
          }
        }
      }
      wave_sync();
    }
  }
}

// bitonic sort of keys a (descending) carrying the payloads b along
__device__ void bitonic_sort_u64_pair(uint64_t* a, uint64_t* b, int n) {
  const int lane = lane_id();
  for (int k = 2; k <= n; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < n; i += 64) {
        int ixj = i ^ j;
        if (ixj > i) {
          uint64_t x = a[i], y = a[ixj];
          bool up = (i & k) != 0;
          if (up ? (x > y) : (x < y)) {
            a[i] = y;
            a[ixj] = x;
            uint64_t t = b[i];
            b[i] = b[ixj];
            b[ixj] = t;
          }
        }
      }
      wave_sync();
    }
  }
}

// sort the buffer descending, keep the best K; returns the new fill and updates the pruning threshold.
// pay (feature mode): payload array moved with the keys
__device__ __forceinline__ int topk_compact(uint64_t* buf, uint64_t* pay, int capb, int nbuf, int K,
                                            uint64_t& thresh) {
  for (int i = nbuf + lane_id(); i < capb; i += 64) buf[i] = 0ULL;
  wave_sync();
  if (pay)
    bitonic_sort_u64_pair(buf, pay, capb);
  else
    bitonic_sort_u64(buf, capb, true);
  int n = nbuf < K ? nbuf : K;
  thresh = n == K ? buf[K - 1] : 0ULL;
  wave_sync();
  return n;
}

// Intensity quantisation of the feature mode: each peak's intensity scaled by the power of two that maps the
// spectrum's largest finite intensity into [2^15, 2^16), truncated to an integer (non-finite, negative, denormal
// and NaN intensities count 0). vh_k1_features documents the same rule.
__device__ __forceinline__ int iq_shift(float maxI) {
  if (!(maxI >= 1.17549435e-38f) || maxI > 3.40282347e38f) return -1000;
  int x;
  frexpf(maxI, &x);
  return 16 - x;
}

__device__ __forceinline__ uint32_t iq_quant(float I, int e) {
  if (e == -1000 || !(I >= 1.17549435e-38f)) return 0u;
  const float v = ldexpf(I, e);
  return v >= 1.f ? (uint32_t)v : 0u;
}

// 14-bit pseudo-log2 of 1 + sum: exponent << 8 | the 8 bits below the leading one
__device__ __forceinline__ uint32_t iq_log(uint64_t sum) {
  const uint64_t x = sum + 1ULL;
  const int e = 63 - __builtin_clzll(x);
  const uint32_t m = e >= 8 ? (uint32_t)((x >> (e - 8)) & 0xFFULL) : (uint32_t)((x << (8 - e)) & 0xFFULL);
  return ((uint32_t)e << 8) | m;
}

struct K1Out {
  // append mode
  int64_t cap;
  unsigned long long* n_out;
  int32_t *out_hyp, *out_row, *out_cnt;
  // top-K mode
  int cand_stride, capb, clos_n;
  int32_t *n_cand, *c_row, *c_hyp, *c_cnt;
  uint32_t* hyp_scored;
  unsigned long long* hyp_matched;
  int32_t* status;
  int wave_bytes, hist_bytes;
  uint32_t* queue;   // persistent mode work counter (NULL: grid mode)
  // feature mode
  const float* pk_int;
  int16_t* c_nb;
  uint32_t* c_iq;
  int32_t* s_maxcnt;
};

// Offer one (count c, row, local hypothesis hl) per lane to the wave's top-K buffer; `flag` = passes the K1
// threshold. Ballot-compacted append, pruned by the current K-th key; a full buffer is compacted first.
template <int MODE>
__device__ __forceinline__ void topk_offer(bool flag, uint32_t c, int row, int hl, uint64_t pl, uint64_t* buf,
                                           uint64_t* pay, int capb, int K, int& nbuf, uint64_t& thresh,
                                           bool& clamped) {
  uint64_t key = 0;
  if (flag) {
    uint32_t cc = c < 32767u ? c : 32767u;
    clamped |= c > 32767u;
    key = ((uint64_t)cc << 48) | ((uint64_t)(uint32_t)row << 20) | ((uint64_t)(63 - hl) << 14);
  }
  bool ins = flag && key > thresh;
  uint64_t bal = ballot(ins);
  int cnt = __popcll(bal);
  if (cnt) {
    if (nbuf + cnt > capb) {
      nbuf = topk_compact(buf, pay, capb, nbuf, K, thresh);
      ins = flag && key > thresh;
      bal = ballot(ins);
      cnt = __popcll(bal);
    }
    if (ins) {
      const int at = nbuf + __popcll(bal & lanemask_lt());
      buf[at] = key;
      if (MODE == 2) pay[at] = pl;
    }
    nbuf += cnt;
    wave_sync();
  }
}

// Feature-mode payload of one slot from its two per-ion counters (count in the low 24 bits, quantised
// intensity sum above): b count (clamped at 32767) << 28 | iq_log(b sum) << 14 | iq_log(y sum); the count is
// returned in c.
__device__ __forceinline__ uint64_t feat_payload(unsigned long long cb, unsigned long long cy, uint32_t& c) {
  const uint32_t nb = (uint32_t)(cb & 0xFFFFFFULL);
  c = nb + (uint32_t)(cy & 0xFFFFFFULL);
  return ((uint64_t)(nb < 32767u ? nb : 32767u) << 28) | ((uint64_t)iq_log(cb >> 24) << 14) |
         (uint64_t)iq_log(cy >> 24);
}

// Final stage of one spectrum: top-K of the buffer, pair closure (MODE 1 with pair_closure: every winner's
// partner row with count -1), dedupe by row (lowest hypothesis, highest count), candidates in row order at
// s * cand_stride. ck: clos_n u64 followed by clos_n i32 of wave-private LDS.
template <int MODE>
__device__ __forceinline__ void k1_finish(const DevP& P, const DevIdx& X, int s, int64_t h0, uint64_t* buf,
                                          uint64_t* pay, int nbuf, uint64_t thresh, bool clamped, uint64_t* ck,
                                          const K1Out& O) {
  const int lane = lane_id();
  const int K = P.topk;
  if (ballot(clamped)) {
    if (lane == 0) atomicOr(O.status, 2);
  }
  const int ntop = topk_compact(buf, pay, O.capb, nbuf, K, thresh);
  if (MODE == 2 && lane == 0) O.s_maxcnt[s] = ntop ? (int32_t)(buf[0] >> 48) : 0;
  int32_t* cv = (int32_t*)(ck + O.clos_n);
  const int ncl = (MODE == 1 && P.pair_closure) ? 2 * ntop : ntop;
  for (int i = lane; i < O.clos_n; i += 64) {
    uint64_t v = ~0ULL;
    if (i < ntop) {
      uint64_t key = buf[i];
      uint32_t row = (uint32_t)((key >> 20) & 0xFFFFFFFULL);
      uint32_t hl = 63u - (uint32_t)((key >> 14) & 63ULL);
      cv[i] = (int32_t)(key >> 48);
      v = ((uint64_t)row << 32) | ((uint64_t)hl << 24) | (uint64_t)i;
    } else if (i < ncl) {
      uint64_t key = buf[i - ntop];
      uint32_t row = (uint32_t)((key >> 20) & 0xFFFFFFFULL);
      uint32_t hl = 63u - (uint32_t)((key >> 14) & 63ULL);
      uint32_t pr = (uint32_t)X.partner[row];
      cv[i] = -1;
      v = ((uint64_t)pr << 32) | ((uint64_t)hl << 24) | (uint64_t)i;
    }
    ck[i] = v;
  }
  wave_sync();
  bitonic_sort_u64(ck, O.clos_n, false);
  int nout = 0;
  int64_t ob = (int64_t)s * O.cand_stride;
  for (int i0 = 0; i0 < ncl; i0 += 64) {
    const int i = i0 + lane;
    bool first = false;
    uint32_t row = 0, hl = 0;
    int32_t cc = -1;
    uint32_t best = 0;
    if (i < ncl) {
      uint64_t v = ck[i];
      row = (uint32_t)(v >> 32);
      hl = (uint32_t)((v >> 24) & 0xFFu);
      first = i == 0 || (uint32_t)(ck[i - 1] >> 32) != row;
      if (first) {
        best = (uint32_t)(v & 0xFFFFFFu);
        cc = cv[best];
        for (int j = i + 1; j < ncl && (uint32_t)(ck[j] >> 32) == row; ++j) {
          const uint32_t q = (uint32_t)(ck[j] & 0xFFFFFFu);
          // the duplicate with the larger count (ties: lower hypothesis) gives the feature payload
          if (cv[q] > cc) {
            cc = cv[q];
            best = q;
          }
        }
      }
    }
    uint64_t bal = ballot(first);
    if (first) {
      int pos = nout + __popcll(bal & lanemask_lt());
      O.c_row[ob + pos] = (int32_t)row;
      O.c_hyp[ob + pos] = (int32_t)(h0 + hl);
      O.c_cnt[ob + pos] = cc;
      if (MODE == 2) {
        const uint64_t pv = pay[best];
        O.c_nb[ob + pos] = (int16_t)((pv >> 28) & 0x7FFFULL);
        O.c_iq[ob + pos] = (uint32_t)(((pv >> 14) & 0x3FFFULL) << 16) | (uint32_t)(pv & 0x3FFFULL);
      }
    }
    nout += __popcll(bal);
  }
  if (lane == 0) O.n_cand[s] = nout;
}
