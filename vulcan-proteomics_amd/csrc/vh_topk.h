// Per-spectrum top-K machinery shared by the K1 kernels (tile K1 and the select stage of the bucket-join K1):
// the ballot-compacted top-K buffer, the feature-mode intensity quantisation and the
// final top-K / pair closure / dedupe stage. Top-K key = count (clamped at 32767) << 48 | row << 20 |
// (63 - local hypothesis) << 14, so the buffer order is count desc, row desc, lower hypothesis first
// (select_topk of the reference, vulcan/gpu/engine.py:284-312).
#pragma once
#include "vh_common.h"

// Keep the best K keys of the buffer (keys are unique: (row, local hypothesis) occurs once per spectrum), return
// the new fill and set the pruning threshold T (once K are kept: a value with exactly K keys >= it, at most the
// K-th largest key; else 0); pay (feature mode): payloads moved with the keys. The kept keys stay in buffer order
// (unsorted): T is found by a bitwise search over key bits 62..14 (bits 13..0 and 63 are zero; each step counts
// the keys >= the candidate: from registers for buffers of up to 512 keys, else with one LDS pass), then keys >= T
// are ballot-compacted in place (every read of a chunk happens before its writes, and writes never pass the read
// position).
__device__ __forceinline__ int topk_compact(uint64_t* buf, uint64_t* pay, int capb, int nbuf, int K,
                                            uint64_t& thresh) {
  const int lane = lane_id();
  if (nbuf <= K) {
    uint64_t mn = ~0ULL;
    for (int i = lane; i < nbuf; i += 64) mn = buf[i] < mn ? buf[i] : mn;
    for (int o = 32; o > 0; o >>= 1) {
      const uint64_t v = __shfl_xor(mn, o);
      mn = v < mn ? v : mn;
    }
    thresh = nbuf == K ? mn : 0ULL;
    return nbuf;
  }
  // The search starts at the highest bit set in any key (the bits above are zero in all of them) and stops at the
  // first candidate with exactly K keys >= it: keys are unique, so keeping the keys >= T keeps exactly the K
  // largest for any such T (the largest of them is the K-th key itself; a smaller one only prunes later offers
  // less tightly).
  uint64_t T = 0;
  constexpr int KR = 8;   // keys per lane held in registers (buffers of up to 512 keys)
  if (capb <= 64 * KR) {
    // the buffer's keys in registers; each pass counts the keys >= the candidate with one compare per register
    // (the compare mask is the ballot) and a scalar popcount
    const int nk = (nbuf + 63) >> 6;
    uint64_t k[KR];
    uint64_t any = 0;
#pragma unroll
    for (int j = 0; j < KR; ++j) {
      k[j] = (j < nk && j * 64 + lane < nbuf) ? buf[j * 64 + lane] : 0ULL;
      any |= k[j];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) any |= __shfl_xor(any, o);
    const int top = any ? 63 - __builtin_clzll(any) : 14;
    for (int bit = top < 62 ? top : 62; bit >= 14; --bit) {
      const uint64_t c = T | (1ULL << bit);
      int cnt = 0;
#pragma unroll
      for (int j = 0; j < KR; ++j)
        if (j < nk) cnt += __popcll(ballot(k[j] >= c));
      if (cnt >= K) T = c;
      if (cnt == K) break;   // uniform: cnt is a wave-wide count
    }
  } else {
    for (int bit = 62; bit >= 14; --bit) {
      const uint64_t c = T | (1ULL << bit);
      int cnt = 0;
      for (int i = lane; i < nbuf; i += 64) cnt += buf[i] >= c ? 1 : 0;
      cnt = wave_sum(cnt);
      if (cnt >= K) T = c;
      if (cnt == K) break;
    }
  }
  int n = 0;
  for (int i0 = 0; i0 < nbuf; i0 += 64) {
    const int i = i0 + lane;
    const uint64_t k = i < nbuf ? buf[i] : 0ULL;
    const uint64_t pv = (pay && i < nbuf) ? pay[i] : 0ULL;
    const bool keep = k >= T;   // T > 0, so lanes past the fill never keep
    const uint64_t bal = ballot(keep);
    wave_sync();
    if (keep) {
      const int at = n + __popcll(bal & lanemask_lt());
      buf[at] = k;
      if (pay) pay[at] = pv;
    }
    n += __popcll(bal);
    wave_sync();
  }
  thresh = T;
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
// s * cand_stride. ck: clos_n u64 of wave-private LDS; the sorted entries go back into buf (capb >= 2K).
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
  if (MODE == 2) {
    uint64_t mx = 0;
    for (int i = lane; i < ntop; i += 64) mx = buf[i] > mx ? buf[i] : mx;
    for (int o = 32; o > 0; o >>= 1) {
      const uint64_t v = __shfl_xor(mx, o);
      mx = v > mx ? v : mx;
    }
    if (lane == 0) O.s_maxcnt[s] = (int32_t)(mx >> 48);
  }
  // closure entries: row << 36 | hypothesis << 30 | (count + 1) << 11 | entry (count -1 = partner row)
  const int ncl = (MODE == 1 && P.pair_closure) ? 2 * ntop : ntop;
  for (int i = lane; i < ncl; i += 64) {
    const uint64_t key = buf[i < ntop ? i : i - ntop];
    uint32_t row = (uint32_t)((key >> 20) & 0xFFFFFFFULL);
    const uint32_t hl = 63u - (uint32_t)((key >> 14) & 63ULL);
    uint32_t c1 = (uint32_t)(key >> 48) + 1u;
    if (i >= ntop) {
      row = (uint32_t)X.partner[row];
      c1 = 0u;
    }
    ck[i] = ((uint64_t)row << 36) | ((uint64_t)hl << 30) | ((uint64_t)c1 << 11) | (uint64_t)i;
  }
  // ascending sort, padded to a power of two with all-ones entries (they sort last; ck holds clos_n >= that)
  int np2 = 1;
  while (np2 < ncl) np2 <<= 1;
  for (int i = ncl + lane; i < np2; i += 64) ck[i] = ~0ULL;
  wave_sync();
  wave_bitonic_u64(ck, np2, false);
  const uint64_t* cs = ck;
  // dedupe by row: the first entry (lowest hypothesis) names the hypothesis, the largest count wins (ties: the
  // earlier entry, which also gives the feature payload)
  int nout = 0;
  const int64_t ob = (int64_t)s * O.cand_stride;
  for (int i0 = 0; i0 < ncl; i0 += 64) {
    const int i = i0 + lane;
    bool first = false;
    uint32_t row = 0, hl = 0, best = 0;
    int32_t cc = -1;
    if (i < ncl) {
      const uint64_t v = cs[i];
      row = (uint32_t)(v >> 36);
      hl = (uint32_t)((v >> 30) & 63ULL);
      first = i == 0 || (uint32_t)(cs[i - 1] >> 36) != row;
      if (first) {
        best = (uint32_t)(v & 0x7FFULL);
        cc = (int32_t)((v >> 11) & 0x7FFFFULL) - 1;
        for (int j = i + 1; j < ncl && (uint32_t)(cs[j] >> 36) == row; ++j) {
          const int32_t cj = (int32_t)((cs[j] >> 11) & 0x7FFFFULL) - 1;
          if (cj > cc) {
            cc = cj;
            best = (uint32_t)(cs[j] & 0x7FFULL);
          }
        }
      }
    }
    const uint64_t bal = ballot(first);
    if (first) {
      const int pos = nout + __popcll(bal & lanemask_lt());
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
