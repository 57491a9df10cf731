// Family top-2 retention for multi-replicate tables (reference vulcan/gpu/family.py:90-203): the per-family
// shortlist on the bucket-join K1's per-row window counts, and the top-2 table after K2.
//
// Shortlist (vh_family_select), wave per spectrum, in three calls (mode 0 counts each spectrum's output entries,
// mode 2 scans them into offsets, mode 1 writes them):
//   1. the rows of the spectrum's hypothesis windows with count >= k1_threshold, ballot-compacted into LDS as
//      keys (K1 family, count desc, row desc, hypothesis asc) and bitonic-sorted;
//   2. rank inside the K1 family = position - first position of the family (binary search); the ones ranked
//      below top-K are kept and each adds its pair partner in every decoy set (count -1, not shortlisted);
//   3. those entries are sorted by (row, hypothesis) and reduced per row: lowest source hypothesis, highest
//      count, shortlisted if any source was; the canonical hypothesis is the spectrum's lowest hypothesis whose
//      window holds the row, else the lowest source. Output in row order.
// A spectrum with more than FS_E rows over the threshold or FS_C shortlist + partner entries does not fit the
// wave's LDS: mode 0 lists it in `wide` and the wide path (vh_family_wide_*) computes its entries with device
// radix sorts over global memory, keys packed with field widths fitted to the batch.
// Top-2 table (vh_family_top2), wave per spectrum: per retention group the two best admissible candidates by
// (hyperscore desc, row asc), found with LDS atomics over order-preserving integer keys.
#include "vh_common.h"

namespace {

constexpr int FS_E = 256;    // rows over the K1 threshold per spectrum (fast path)
constexpr int FS_C = 1024;   // shortlist + partner entries per spectrum (fast path)
constexpr int FS_WAVES = 4;  // 10 KB of LDS per wave
constexpr uint64_t FS_ROWMASK = (1ull << 28) - 1;

struct FsArgs {
  int n_spec;
  const int64_t* hyp_off;
  const int64_t* h_lo;
  const int64_t* h_hi;
  const int64_t* h_slot;
  int64_t slot_cap;
  const uint32_t* counts;
  const uint8_t* kfam;
  int64_t n_rows;
  const int64_t* partners;
  int J;
  int topk;
  int thr;
  int64_t* out_off;
  int32_t* o_row;
  int32_t* o_hyp;
  int32_t* o_cnt;
  uint8_t* o_top;
  int32_t* wide;     // [0] = number of spectra listed, [1 + i] = the spectra (mode 0)
  int32_t* status;
};

// a window whose slots lie beyond slot_cap was flagged by vh_k1j_pairs: it is skipped (status bit 1)
__device__ __forceinline__ bool window_ok(const FsArgs& A, int64_t h, int len) {
  const int64_t sl = A.h_slot[h];
  return sl >= 0 && sl + len <= A.slot_cap;
}

template <int MODE>
__global__ __launch_bounds__(64 * FS_WAVES) void family_select_kernel(FsArgs A) {
  __shared__ uint64_t s_a[FS_WAVES][FS_E];
  __shared__ uint64_t s_c[FS_WAVES][FS_C];
  const int w = threadIdx.x >> 6, lane = lane_id();
  const int s = blockIdx.x * FS_WAVES + w;
  if (s >= A.n_spec) return;   // wave-uniform
  uint64_t* a = s_a[w];
  uint64_t* cl = s_c[w];
  const int64_t h0 = A.hyp_off[s];
  const int nh_all = (int)(A.hyp_off[s + 1] - h0);
  const int nh = nh_all < 64 ? nh_all : 64;
  bool bad = nh_all > 64;
  // 1. rows over the threshold
  int n = 0;
  for (int k = 0; k < nh; ++k) {
    const int64_t lo = A.h_lo[h0 + k], hi = A.h_hi[h0 + k];
    const int len = hi > lo ? (int)(hi - lo) : 0;
    if (len == 0) continue;
    if (!window_ok(A, h0 + k, len)) {   // uniform
      bad = true;
      continue;
    }
    const int64_t sl = A.h_slot[h0 + k];
    for (int i0 = 0; i0 < len; i0 += 64) {
      const int i = i0 + lane;
      uint32_t c = 0;
      if (i < len) c = A.counts[sl + i];
      const bool take = i < len && c >= (uint32_t)A.thr && lo + i < A.n_rows;
      const uint64_t m = ballot(take);
      const int pos = n + __popcll(m & lanemask_lt());
      if (take && pos < FS_E) {
        const uint64_t row = (uint64_t)(lo + i);
        const uint64_t kf = A.kfam[row];
        const uint64_t inv_c = 32767u - (c < 32767u ? c : 32767u);
        a[pos] = (kf << 49) | (inv_c << 34) | ((FS_ROWMASK - row) << 6) | (uint64_t)k;
      }
      n += __popcll(m);
    }
  }
  bool wide = n > FS_E;
  int nu = 0;
  if (!wide) {
    int np2 = 1;
    while (np2 < n) np2 <<= 1;
    for (int i = n + lane; i < np2; i += 64) a[i] = ~0ull;
    wave_sync();
    if (np2 > 1) wave_bitonic_u64(a, np2, false);
    wave_sync();
    // 2. top-K per K1 family + partners in every decoy set
    const int J1 = A.J + 1;
    int nk = 0;
    for (int i0 = 0; i0 < n; i0 += 64) {
      const int i = i0 + lane;
      bool keep = false;
      uint64_t e = 0;
      if (i < n) {
        e = a[i];
        const uint64_t fam = e >> 49;
        int lo = 0, hi = i;   // first position of the family
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if ((a[mid] >> 49) < fam) lo = mid + 1;
          else hi = mid;
        }
        keep = (i - lo) < A.topk;
      }
      const uint64_t m = ballot(keep);
      const int pos = nk + __popcll(m & lanemask_lt());
      if (keep && (pos + 1) * J1 <= FS_C) {
        const uint64_t k = e & 63ull;
        const uint64_t row = FS_ROWMASK - ((e >> 6) & FS_ROWMASK);
        const uint64_t c = 32767ull - ((e >> 34) & 32767ull);
        cl[pos * J1] = (row << 24) | (k << 18) | ((c + 1) << 1) | 1ull;
        for (int j = 0; j < A.J; ++j) {
          const uint64_t pr = (uint64_t)A.partners[(int64_t)j * A.n_rows + (int64_t)row];
          cl[pos * J1 + 1 + j] = ((pr & FS_ROWMASK) << 24) | (k << 18);
        }
      }
      nk += __popcll(m);
    }
    const int ncl = nk * J1;
    wide = ncl > FS_C;
    if (!wide) {
      int np3 = 1;
      while (np3 < ncl) np3 <<= 1;
      for (int i = ncl + lane; i < np3; i += 64) cl[i] = ~0ull;
      wave_sync();
      if (np3 > 1) wave_bitonic_u64(cl, np3, false);
      wave_sync();
      // 3. one entry per row
      const int64_t o0 = MODE == 1 ? A.out_off[s] : 0;
      for (int i0 = 0; i0 < ncl; i0 += 64) {
        const int i = i0 + lane;
        bool start = false;
        uint64_t e = 0;
        if (i < ncl) {
          e = cl[i];
          start = i == 0 || (cl[i - 1] >> 24) != (e >> 24);
        }
        const uint64_t m = ballot(start);
        if (MODE == 1 && start) {
          const int pos = nu + __popcll(m & lanemask_lt());
          const int64_t row = (int64_t)(e >> 24);
          const int ksrc = (int)((e >> 18) & 63ull);
          uint32_t cmax = (uint32_t)((e >> 1) & 0xFFFFull), top = (uint32_t)(e & 1ull);
          for (int j = i + 1; j < ncl && (cl[j] >> 24) == (e >> 24); ++j) {
            const uint32_t cj = (uint32_t)((cl[j] >> 1) & 0xFFFFull);
            cmax = cj > cmax ? cj : cmax;
            top |= (uint32_t)(cl[j] & 1ull);
          }
          int hc = -1;
          for (int k = 0; k < nh; ++k) {
            if (A.h_lo[h0 + k] <= row && row < A.h_hi[h0 + k]) {
              hc = k;
              break;
            }
          }
          const int64_t q = o0 + pos;
          A.o_row[q] = (int32_t)row;
          A.o_hyp[q] = (int32_t)(h0 + (hc >= 0 ? hc : ksrc));
          A.o_cnt[q] = (int32_t)cmax - 1;
          A.o_top[q] = (uint8_t)top;
        }
        nu += __popcll(m);
      }
    }
  }
  if (lane == 0) {
    if (MODE == 0) {
      A.out_off[s + 1] = wide ? 0 : nu;   // a wide spectrum's count comes from vh_family_wide
      if (s == 0) A.out_off[0] = 0;
      if (wide) {
        const int at = atomicAdd(&A.wide[0], 1);
        A.wide[1 + at] = s;
      }
    }
    if (bad) atomicOr(A.status, 1);
  }
}

__global__ void fs_zero_kernel(int64_t* p) { *p = 0; }

// ------------------------------------------------------------------------------------------------ wide path
struct WideArgs {
  int n_wide;
  const int32_t* wide_spec;
  const int64_t* hyp_off;
  const int64_t* h_lo;
  const int64_t* h_hi;
  const int64_t* h_slot;
  int64_t slot_cap;
  const uint32_t* counts;
  const uint8_t* kfam;
  int64_t n_rows;
  const int64_t* partners;
  int J;
  int topk;
  int thr;
};

// per wide spectrum q: rows over the threshold (w_off[q + 1]), the lowest such row (rmin[q]); over all: the
// largest count (clamped at 32767) and the largest row span rmax - rmin + 1 (gstats[0], gstats[1])
__global__ __launch_bounds__(256) void wide_stats_kernel(WideArgs A, int64_t* w_off, int64_t* rmin,
                                                         unsigned long long* gstats) {
  const int q = blockIdx.x * 4 + (threadIdx.x >> 6), lane = lane_id();
  if (q >= A.n_wide) return;
  const int s = A.wide_spec[q];
  const int64_t h0 = A.hyp_off[s];
  int nh = (int)(A.hyp_off[s + 1] - h0);
  nh = nh < 64 ? nh : 64;
  int64_t n = 0, lo_row = INT64_MAX, hi_row = -1;
  uint32_t cmax = 0;
  for (int k = 0; k < nh; ++k) {
    const int64_t lo = A.h_lo[h0 + k], hi = A.h_hi[h0 + k];
    const int len = hi > lo ? (int)(hi - lo) : 0;
    const int64_t sl = A.h_slot[h0 + k];
    if (len == 0 || sl < 0 || sl + len > A.slot_cap) continue;
    for (int i = lane; i < len; i += 64) {
      const uint32_t c = A.counts[sl + i];
      if (c >= (uint32_t)A.thr && lo + i < A.n_rows) {
        ++n;
        lo_row = lo + i < lo_row ? lo + i : lo_row;
        hi_row = lo + i > hi_row ? lo + i : hi_row;
        cmax = c > cmax ? c : cmax;
      }
    }
  }
  n = wave_sum(n);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int64_t a = __shfl_xor(lo_row, o), b = __shfl_xor(hi_row, o);
    const uint32_t c = __shfl_xor(cmax, o);
    lo_row = a < lo_row ? a : lo_row;
    hi_row = b > hi_row ? b : hi_row;
    cmax = c > cmax ? c : cmax;
  }
  if (lane == 0) {
    w_off[q + 1] = n;
    rmin[q] = n ? lo_row : 0;
    if (q == 0) w_off[0] = 0;
    if (n) {
      atomicMax(&gstats[0], (unsigned long long)(cmax < 32767u ? cmax : 32767u));
      atomicMax(&gstats[1], (unsigned long long)(hi_row - lo_row + 1));
    }
  }
}

__global__ void wide_stats_zero_kernel(unsigned long long* g) {
  if (threadIdx.x < 2) g[threadIdx.x] = 0ull;
}

// Field widths of the first sort key, low to high: hypothesis (6), row offset from rmin (rb, descending),
// count (cb, descending), K1 family (fb), spectrum of the batch (the rest).
struct KeyW {
  int rb, cb, fb;
  __host__ __device__ int low() const { return 6 + rb + cb + fb; }
};

// wave per spectrum of the batch [q0, q0 + nq): its over-threshold rows as first-stage keys at w_off[q] - base
__global__ __launch_bounds__(256) void wide_emit_kernel(WideArgs A, int q0, int nq, const int64_t* w_off,
                                                        const int64_t* rmin, KeyW W, uint64_t* keys) {
  const int qb = blockIdx.x * 4 + (threadIdx.x >> 6), lane = lane_id();
  if (qb >= nq) return;
  const int q = q0 + qb;
  const int s = A.wide_spec[q];
  const int64_t h0 = A.hyp_off[s];
  int nh = (int)(A.hyp_off[s + 1] - h0);
  nh = nh < 64 ? nh : 64;
  const int64_t base = w_off[q] - w_off[q0];
  const int64_t cap = w_off[q + 1] - w_off[q];
  const uint64_t cmask = (1ull << W.cb) - 1ull, rmask = (1ull << W.rb) - 1ull;
  int64_t n = 0;
  for (int k = 0; k < nh; ++k) {
    const int64_t lo = A.h_lo[h0 + k], hi = A.h_hi[h0 + k];
    const int len = hi > lo ? (int)(hi - lo) : 0;
    const int64_t sl = A.h_slot[h0 + k];
    if (len == 0 || sl < 0 || sl + len > A.slot_cap) continue;
    for (int i0 = 0; i0 < len; i0 += 64) {
      const int i = i0 + lane;
      uint32_t c = 0;
      if (i < len) c = A.counts[sl + i];
      const bool take = i < len && c >= (uint32_t)A.thr && lo + i < A.n_rows;
      const uint64_t m = ballot(take);
      const int64_t pos = n + __popcll(m & lanemask_lt());
      if (take && pos < cap) {
        const int64_t row = lo + i;
        const uint64_t cc = c < 32767u ? c : 32767u;
        const uint64_t rrel = (uint64_t)(row - rmin[q]);
        keys[base + pos] = ((uint64_t)qb << W.low()) | ((uint64_t)A.kfam[row] << (6 + W.rb + W.cb)) |
                           ((cmask - cc) << (6 + W.rb)) | ((rmask - rrel) << 6) | (uint64_t)k;
      }
      n += __popcll(m);
    }
  }
}

// Second-stage entries: batch spectrum << 51 | row << 23 | source hypothesis << 17 | (count + 1) << 1 |
// shortlisted; a key that is not in its family's top-K yields no entries (~0, sorted last).
constexpr uint64_t W2_NONE = ~0ull;

__global__ __launch_bounds__(256) void wide_rank_kernel(WideArgs A, int q0, int nq, const int64_t* w_off,
                                                        const int64_t* rmin, KeyW W, const uint64_t* keys,
                                                        int64_t n1, uint64_t* ent) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n1) return;
  const int J1 = A.J + 1;
  const uint64_t e = keys[i];
  const int qb = (int)(e >> W.low());
  const int q = q0 + qb;
  const int sh = 6 + W.rb + W.cb;       // family and spectrum bits start here
  const uint64_t grp = e >> sh;
  int64_t lo = w_off[q] - w_off[q0], hi = i;   // first key of (spectrum, family): keys of q are contiguous
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((keys[mid] >> sh) < grp) lo = mid + 1;
    else hi = mid;
  }
  const bool keep = (i - lo) < A.topk;
  if (!keep) {
    for (int j = 0; j < J1; ++j) ent[i * J1 + j] = W2_NONE;
    return;
  }
  const uint64_t cmask = (1ull << W.cb) - 1ull, rmask = (1ull << W.rb) - 1ull;
  const uint64_t k = e & 63ull;
  const int64_t row = rmin[q] + (int64_t)(rmask - ((e >> 6) & rmask));
  const uint64_t c = cmask - ((e >> (6 + W.rb)) & cmask);
  const uint64_t hiq = (uint64_t)qb << 51;
  ent[i * J1] = hiq | ((uint64_t)row << 23) | (k << 17) | ((c + 1) << 1) | 1ull;
  for (int j = 0; j < A.J; ++j) {
    const uint64_t pr = (uint64_t)A.partners[(int64_t)j * A.n_rows + row] & FS_ROWMASK;
    ent[i * J1 + 1 + j] = hiq | (pr << 23) | (k << 17);
  }
}

// starts of (spectrum, row) runs among the sorted entries: flag[i] = 1, and one more entry for the spectrum
__global__ __launch_bounds__(256) void wide_flag_kernel(const uint64_t* ent, int64_t n2, int64_t* flag,
                                                        int64_t* ucnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n2) return;
  const uint64_t e = ent[i];
  const bool start = e != W2_NONE && (i == 0 || (ent[i - 1] >> 23) != (e >> 23));
  flag[i] = start ? 1 : 0;
  if (start) atomicAdd((unsigned long long*)&ucnt[e >> 51], 1ull);
}

__global__ __launch_bounds__(256) void wide_zero_kernel(int64_t* a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// per batch spectrum: its unique entries start at `ubase` + (unique entries of the batch's earlier spectra);
// the spectrum's count also goes to out_off[spectrum + 1]
__global__ __launch_bounds__(256) void wide_counts_kernel(const WideArgs A, int q0, int nq, const int64_t* ucnt_incl,
                                                          int64_t ubase, int64_t* uniq_off, int64_t* uniq_cnt,
                                                          int64_t* out_off) {
  const int qb = blockIdx.x * blockDim.x + threadIdx.x;
  if (qb >= nq) return;
  const int64_t before = qb ? ucnt_incl[qb - 1] : 0;
  const int64_t c = ucnt_incl[qb] - before;
  uniq_off[q0 + qb] = ubase + before;
  uniq_cnt[q0 + qb] = c;
  out_off[A.wide_spec[q0 + qb] + 1] = c;
}

// one packed entry per (spectrum, row) run: row << 23 | canonical hypothesis (local) << 17 | (max count + 1) << 1
// | shortlisted by any source
__global__ __launch_bounds__(256) void wide_reduce_kernel(WideArgs A, int q0, const uint64_t* ent, int64_t n2,
                                                          const int64_t* flag_incl, int64_t ubase,
                                                          uint64_t* uniq) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n2) return;
  const uint64_t e = ent[i];
  if (e == W2_NONE || (i > 0 && (ent[i - 1] >> 23) == (e >> 23))) return;
  const int s = A.wide_spec[q0 + (int)(e >> 51)];
  const int64_t row = (int64_t)((e >> 23) & FS_ROWMASK);
  const uint64_t ksrc = (e >> 17) & 63ull;   // runs are sorted by source hypothesis: the first is the lowest
  uint64_t cmax = (e >> 1) & 0xFFFFull, top = e & 1ull;
  for (int64_t j = i + 1; j < n2 && (ent[j] >> 23) == (e >> 23); ++j) {
    const uint64_t cj = (ent[j] >> 1) & 0xFFFFull;
    cmax = cj > cmax ? cj : cmax;
    top |= ent[j] & 1ull;
  }
  const int64_t h0 = A.hyp_off[s];
  int nh = (int)(A.hyp_off[s + 1] - h0);
  nh = nh < 64 ? nh : 64;
  uint64_t hc = ksrc;
  for (int k = 0; k < nh; ++k)
    if (A.h_lo[h0 + k] <= row && row < A.h_hi[h0 + k]) {
      hc = (uint64_t)k;
      break;
    }
  uniq[ubase + flag_incl[i] - 1] = ((uint64_t)row << 23) | (hc << 17) | (cmax << 1) | top;
}

// wave per wide spectrum: its packed entries into the shortlist outputs at out_off[spectrum]
__global__ __launch_bounds__(256) void wide_copy_kernel(int n_wide, const int32_t* wide_spec, const int64_t* hyp_off,
                                                        const int64_t* uniq_off, const int64_t* uniq_cnt,
                                                        const uint64_t* uniq, const int64_t* out_off,
                                                        int32_t* o_row, int32_t* o_hyp, int32_t* o_cnt,
                                                        uint8_t* o_top) {
  const int q = blockIdx.x * 4 + (threadIdx.x >> 6), lane = lane_id();
  if (q >= n_wide) return;
  const int s = wide_spec[q];
  const int64_t u0 = uniq_off[q], nu = uniq_cnt[q], o0 = out_off[s], h0 = hyp_off[s];
  for (int64_t j = lane; j < nu; j += 64) {
    const uint64_t e = uniq[u0 + j];
    o_row[o0 + j] = (int32_t)((e >> 23) & FS_ROWMASK);
    o_hyp[o0 + j] = (int32_t)(h0 + (int64_t)((e >> 17) & 63ull));
    o_cnt[o0 + j] = (int32_t)((e >> 1) & 0xFFFFull) - 1;
    o_top[o0 + j] = (uint8_t)(e & 1ull);
  }
}

int bits_for(uint64_t v) {   // bits to hold 0..v (at least 1)
  int b = 1;
  while (b < 64 && (v >> b)) ++b;
  return b;
}

size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

// ------------------------------------------------------------------------------------------------ top-2 table
constexpr int T2_WAVES = 4;
constexpr int T2_GROUPS = 256;

__device__ __forceinline__ uint64_t f64_order(double x) {   // order-preserving u64 of a double (no NaN)
  const uint64_t b = (uint64_t)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | (1ull << 63));
}

struct T2Args {
  int n_spec;
  const int64_t* cand_off;
  const int32_t* c_row;
  const int32_t* c_hyp;
  const uint8_t* c_top;
  const double* hs;
  const int16_t* mb;
  const int16_t* my;
  const int8_t* fwd_fam;
  const int8_t* side;
  int64_t n_rows;
  const int64_t* partners;
  int J;
  int n_groups;
  int min_matched;
  const int8_t* h_z;
  const int8_t* h_iso;
  int64_t* out_off;
  VhFamilyCols o;
};

// candidate i of the spectrum (rows sorted) in context j: its retention group, or -1 when it is not admissible
// (fewer than min_matched matches; neither it nor its set-j partner shortlisted) or not on either side of j
__device__ __forceinline__ int t2_group(const T2Args& A, int64_t c0, int64_t c1, int64_t i, int j) {
  if ((int)A.mb[i] + (int)A.my[i] < A.min_matched) return -1;
  const int64_t row = A.c_row[i];
  const int sd = A.side[row];
  if (sd >= 0 && sd != j) return -1;
  bool adm = A.c_top[i] != 0;
  if (!adm) {
    const int64_t pr = A.partners[(int64_t)j * A.n_rows + row];
    int64_t lo = c0, hi = c1;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (A.c_row[mid] < pr) lo = mid + 1;
      else hi = mid;
    }
    adm = lo < c1 && A.c_row[lo] == pr && A.c_top[lo] != 0;
  }
  if (!adm) return -1;
  const int g = ((int)A.fwd_fam[row] * A.J + j) * 2 + (sd >= 0 ? 1 : 0);
  return g >= 0 && g < A.n_groups ? g : -1;   // a family code beyond n_fwd_fam names no group
}

template <int MODE>
__global__ __launch_bounds__(64 * T2_WAVES) void family_top2_kernel(T2Args A) {
  __shared__ unsigned long long s_k1[T2_WAVES][T2_GROUPS], s_k2[T2_WAVES][T2_GROUPS];
  __shared__ int s_r1[T2_WAVES][T2_GROUPS], s_r2[T2_WAVES][T2_GROUPS];
  const int w = threadIdx.x >> 6, lane = lane_id();
  const int s = blockIdx.x * T2_WAVES + w;
  if (s >= A.n_spec) return;   // wave-uniform
  unsigned long long* k1 = s_k1[w];
  unsigned long long* k2 = s_k2[w];
  int* r1 = s_r1[w];
  int* r2 = s_r2[w];
  const int G = A.n_groups;
  for (int g = lane; g < G; g += 64) {
    k1[g] = 0ull;   // f64_order of every finite score is > 0
    k2[g] = 0ull;
    r1[g] = INT32_MAX;
    r2[g] = INT32_MAX;
  }
  wave_sync();
  const int64_t c0 = A.cand_off[s], c1 = A.cand_off[s + 1];
  // pass 1: best score per group; pass 2: lowest row at it; pass 3: second best (the first excluded); pass 4:
  // lowest row at the second
  for (int pass = 0; pass < 4; ++pass) {
    for (int64_t i = c0 + lane; i < c1; i += 64) {
      const uint64_t key = f64_order(A.hs[i]);
      const int row = A.c_row[i];
      for (int j = 0; j < A.J; ++j) {
        const int g = t2_group(A, c0, c1, i, j);
        if (g < 0) continue;
        const bool first = key == k1[g] && row == r1[g];
        if (pass == 0) atomicMax(&k1[g], key);
        else if (pass == 1) {
          if (key == k1[g]) atomicMin(&r1[g], row);
        } else if (pass == 2) {
          if (!first) atomicMax(&k2[g], key);
        } else if (key == k2[g] && !first) atomicMin(&r2[g], row);
      }
    }
    wave_sync();
  }
  // filled (group, rank) slots in group order
  int nu = 0;
  const int64_t o0 = MODE == 1 ? A.out_off[s] : 0;
  for (int g0 = 0; g0 < G; g0 += 32) {
    const int g = g0 + (lane >> 1), rk = lane & 1;
    const bool full = g < G && (rk == 0 ? r1[g] != INT32_MAX : r2[g] != INT32_MAX);
    const uint64_t m = ballot(full);
    if (MODE == 1 && full) {
      const int row = rk == 0 ? r1[g] : r2[g];
      int64_t lo = c0, hi = c1;   // the candidate holding the row (rows are unique in the spectrum's list)
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (A.c_row[mid] < row) lo = mid + 1;
        else hi = mid;
      }
      const int64_t q = o0 + nu + __popcll(m & lanemask_lt());
      const int h = A.c_hyp[lo];
      A.o.idx[q] = lo;
      A.o.group[q] = (int16_t)g;
      A.o.rank[q] = (int8_t)(rk + 1);
      A.o.row[q] = row;
      A.o.hyp[q] = h;
      A.o.hs[q] = A.hs[lo];
      A.o.mb[q] = A.mb[lo];
      A.o.my[q] = A.my[lo];
      A.o.z[q] = A.h_z[h];
      A.o.iso[q] = A.h_iso[h];
    }
    nu += __popcll(m);
  }
  if (MODE == 0 && lane == 0) {
    A.out_off[s + 1] = nu;
    if (s == 0) A.out_off[0] = 0;
  }
}

}  // namespace

extern "C" size_t vh_family_select_workspace_bytes(int32_t n_spec) {
  return vh_scan_ws_i64(n_spec > 0 ? n_spec : 1);
}

extern "C" int vh_family_select(const VhParams* p, int32_t n_spec, const int64_t* hyp_off, const int64_t* h_lo,
                                const int64_t* h_hi, const int64_t* h_slot, int64_t slot_cap,
                                const uint32_t* counts, const uint8_t* kfam, int64_t n_rows,
                                const int64_t* partners, int32_t n_sets, int32_t mode, int64_t* out_off,
                                int32_t* o_row, int32_t* o_hyp, int32_t* o_cnt, uint8_t* o_top, int32_t* wide,
                                int32_t* status, void* ws, size_t ws_bytes, void* stream) {
  if (!p || n_spec < 0 || !out_off || !status || mode < 0 || mode > 2 || n_sets < 1 || n_sets > 8 ||
      n_rows < 0 || p->topk < 1)
    return VH_E_ARG;
  if (n_rows >= ((int64_t)1 << 28)) return VH_E_LIMIT;
  hipStream_t st = (hipStream_t)stream;
  if (n_spec == 0) {
    if (mode == 0) {   // out_off[0] = 0 by a one-thread launch (no memset: the call stays capturable)
      hipLaunchKernelGGL(fs_zero_kernel, dim3(1), dim3(1), 0, st, out_off);
      VH_CHECK_LAUNCH();
    }
    return 0;
  }
  if (mode == 2) {
    if (!ws || ws_bytes < vh_family_select_workspace_bytes(n_spec)) return VH_E_WORKSPACE;
    return vh_scan_inclusive_i64(out_off + 1, n_spec, ws, ws_bytes, st);
  }
  if (!hyp_off || !h_lo || !h_hi || !h_slot || !kfam || !partners || (slot_cap > 0 && !counts)) return VH_E_ARG;
  if (mode == 1 && (!o_row || !o_hyp || !o_cnt || !o_top)) return VH_E_ARG;
  if (mode == 0 && !wide) return VH_E_ARG;
  FsArgs A{n_spec, hyp_off, h_lo, h_hi, h_slot, slot_cap, counts, kfam, n_rows, partners, n_sets, p->topk,
           p->k1_threshold, out_off, o_row, o_hyp, o_cnt, o_top, wide, status};
  const unsigned g = (unsigned)((n_spec + FS_WAVES - 1) / FS_WAVES);
  if (mode == 0) hipLaunchKernelGGL(family_select_kernel<0>, dim3(g), dim3(64 * FS_WAVES), 0, st, A);
  else hipLaunchKernelGGL(family_select_kernel<1>, dim3(g), dim3(64 * FS_WAVES), 0, st, A);
  VH_CHECK_LAUNCH();
  return 0;
}

extern "C" size_t vh_family_wide_stats_workspace_bytes(int32_t n_wide) {
  return vh_scan_ws_i64(n_wide > 0 ? n_wide : 1);
}

extern "C" int vh_family_wide_stats(const VhParams* p, int32_t n_wide, const int32_t* wide_spec,
                                    const int64_t* hyp_off, const int64_t* h_lo, const int64_t* h_hi,
                                    const int64_t* h_slot, int64_t slot_cap, const uint32_t* counts,
                                    int64_t n_rows, int64_t* w_off, int64_t* rmin, uint64_t* gstats, void* ws,
                                    size_t ws_bytes, void* stream) {
  if (!p || n_wide < 1 || !wide_spec || !hyp_off || !h_lo || !h_hi || !h_slot || !w_off || !rmin || !gstats ||
      (slot_cap > 0 && !counts))
    return VH_E_ARG;
  if (!ws || ws_bytes < vh_family_wide_stats_workspace_bytes(n_wide)) return VH_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  WideArgs A{n_wide, wide_spec, hyp_off, h_lo, h_hi, h_slot, slot_cap, counts, nullptr, n_rows, nullptr, 1, p->topk,
             p->k1_threshold};
  hipLaunchKernelGGL(wide_stats_zero_kernel, dim3(1), dim3(64), 0, st, (unsigned long long*)gstats);
  hipLaunchKernelGGL(wide_stats_kernel, dim3((unsigned)((n_wide + 3) / 4)), dim3(256), 0, st, A, w_off, rmin,
                     (unsigned long long*)gstats);
  VH_CHECK_LAUNCH();
  return vh_scan_inclusive_i64(w_off + 1, n_wide, ws, ws_bytes, st);
}

// batch of wide spectra whose first-stage keys fit 64 bits and second-stage keys 13 spectrum bits
static int wide_batch(int low_bits) {
  const int qb = 64 - low_bits;
  return qb >= 13 ? 8192 : (1 << qb);
}

static KeyW wide_widths(int32_t n_kfam, uint64_t cmax, uint64_t span) {
  KeyW W;
  W.fb = bits_for((uint64_t)(n_kfam > 1 ? n_kfam - 1 : 1));
  W.cb = bits_for(cmax < 32767 ? cmax : 32767);
  W.rb = bits_for(span > 1 ? span - 1 : 1);
  return W;
}

extern "C" size_t vh_family_wide_workspace_bytes(int64_t n_entries, int32_t n_sets) {
  const int64_t n1 = n_entries > 0 ? n_entries : 1;
  const int64_t n2 = n1 * (int64_t)(n_sets + 1);
  return 2 * a256((size_t)n1 * 8) + 2 * a256((size_t)n2 * 8) + a256((size_t)n2 * 8) + a256((size_t)8193 * 8) +
         vh_radix_ws_bytes(n2) + vh_scan_ws_i64(n2 > 8192 ? n2 : 8192);
}

extern "C" int vh_family_wide(const VhParams* p, int32_t n_wide, const int32_t* wide_spec, const int64_t* hyp_off,
                              const int64_t* h_lo, const int64_t* h_hi, const int64_t* h_slot, int64_t slot_cap,
                              const uint32_t* counts, const uint8_t* kfam, int32_t n_kfam, int64_t n_rows,
                              const int64_t* partners, int32_t n_sets, const int64_t* w_off,
                              const int64_t* w_off_host, const int64_t* rmin, uint64_t cmax, uint64_t span,
                              uint64_t* uniq, int64_t* uniq_off, int64_t* uniq_cnt, int64_t* out_off, void* ws,
                              size_t ws_bytes, void* stream) {
  if (!p || n_wide < 1 || !wide_spec || !hyp_off || !h_lo || !h_hi || !h_slot || !kfam || !partners || !w_off ||
      !w_off_host || !rmin || !uniq || !uniq_off || !uniq_cnt || !out_off || n_sets < 1 || n_sets > 8 ||
      n_kfam < 1 || n_kfam > 256 || n_rows < 0 || p->topk < 1 || (slot_cap > 0 && !counts))
    return VH_E_ARG;
  if (n_rows >= ((int64_t)1 << 28)) return VH_E_LIMIT;
  const int64_t total = w_off_host[n_wide];
  if (total < 0 || ws_bytes < vh_family_wide_workspace_bytes(total, n_sets) || !ws) return VH_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const KeyW W = wide_widths(n_kfam, cmax, span);
  if (W.low() > 57) return VH_E_LIMIT;   // never for rows < 2^28, counts <= 32767, <= 256 families
  const int J1 = n_sets + 1;
  const int64_t n1max = total > 0 ? total : 1, n2max = n1max * J1;
  char* w = (char*)ws;
  uint64_t* ka = (uint64_t*)w;
  w += a256((size_t)n1max * 8);
  uint64_t* kb = (uint64_t*)w;
  w += a256((size_t)n1max * 8);
  uint64_t* ea = (uint64_t*)w;
  w += a256((size_t)n2max * 8);
  uint64_t* eb = (uint64_t*)w;
  w += a256((size_t)n2max * 8);
  int64_t* flag = (int64_t*)w;
  w += a256((size_t)n2max * 8);
  int64_t* ucnt = (int64_t*)w;
  w += a256((size_t)8193 * 8);
  void* rws = w;
  const size_t rws_bytes = vh_radix_ws_bytes(n2max);
  w += rws_bytes;
  void* sws = w;
  const size_t sws_bytes = vh_scan_ws_i64(n2max > 8192 ? n2max : 8192);   // also scans a batch's 8,192 counts
  WideArgs A{n_wide, wide_spec, hyp_off, h_lo, h_hi, h_slot, slot_cap, counts, kfam, n_rows, partners, n_sets,
             p->topk, p->k1_threshold};
  const int batch = wide_batch(W.low());
  for (int q0 = 0; q0 < n_wide; q0 += batch) {
    const int nq = n_wide - q0 < batch ? n_wide - q0 : batch;
    const int64_t n1 = w_off_host[q0 + nq] - w_off_host[q0];
    const int64_t n2 = n1 * J1;
    const int64_t ubase = w_off_host[q0] * J1;   // the batch's unique entries: at most J1 per first-stage key
    hipLaunchKernelGGL(wide_zero_kernel, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, st, ucnt, (int64_t)nq);
    VH_CHECK_LAUNCH();
    if (n1 > 0) {
      hipLaunchKernelGGL(wide_emit_kernel, dim3((unsigned)((nq + 3) / 4)), dim3(256), 0, st, A, q0, nq, w_off, rmin,
                         W, ka);
      VH_CHECK_LAUNCH();
      uint64_t* k1s = ka;
      const int bits1 = W.low() + bits_for((uint64_t)(nq > 1 ? nq - 1 : 1));
      int rc = vh_radix_sort_u64(ka, kb, n1, 0, bits1, rws, rws_bytes, st, &k1s);
      if (rc) return rc;
      hipLaunchKernelGGL(wide_rank_kernel, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, st, A, q0, nq, w_off,
                         rmin, W, k1s, n1, ea);
      VH_CHECK_LAUNCH();
      uint64_t* es = ea;
      rc = vh_radix_sort_u64(ea, eb, n2, 0, 64, rws, rws_bytes, st, &es);
      if (rc) return rc;
      hipLaunchKernelGGL(wide_flag_kernel, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st, es, n2, flag, ucnt);
      VH_CHECK_LAUNCH();
      rc = vh_scan_inclusive_i64(flag, n2, sws, sws_bytes, st);
      if (rc) return rc;
      hipLaunchKernelGGL(wide_reduce_kernel, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st, A, q0, es, n2,
                         flag, ubase, uniq);
      VH_CHECK_LAUNCH();
    }
    int rc = vh_scan_inclusive_i64(ucnt, nq, sws, sws_bytes, st);
    if (rc) return rc;
    hipLaunchKernelGGL(wide_counts_kernel, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, st, A, q0, nq, ucnt,
                       ubase, uniq_off, uniq_cnt, out_off);
    VH_CHECK_LAUNCH();
  }
  return 0;
}

extern "C" int vh_family_wide_copy(int32_t n_wide, const int32_t* wide_spec, const int64_t* hyp_off,
                                   const int64_t* uniq_off, const int64_t* uniq_cnt, const uint64_t* uniq,
                                   const int64_t* out_off, int32_t* o_row, int32_t* o_hyp, int32_t* o_cnt,
                                   uint8_t* o_top, void* stream) {
  if (n_wide < 0) return VH_E_ARG;
  if (n_wide == 0) return 0;
  if (!wide_spec || !hyp_off || !uniq_off || !uniq_cnt || !uniq || !out_off || !o_row || !o_hyp || !o_cnt || !o_top)
    return VH_E_ARG;
  hipLaunchKernelGGL(wide_copy_kernel, dim3((unsigned)((n_wide + 3) / 4)), dim3(256), 0, (hipStream_t)stream, n_wide,
                     wide_spec, hyp_off, uniq_off, uniq_cnt, uniq, out_off, o_row, o_hyp, o_cnt, o_top);
  VH_CHECK_LAUNCH();
  return 0;
}

extern "C" size_t vh_family_top2_workspace_bytes(int32_t n_spec) {
  return vh_scan_ws_i64(n_spec > 0 ? n_spec : 1);
}

extern "C" int vh_family_top2(const VhParams* p, int32_t n_spec, const int64_t* cand_off, const int32_t* c_row,
                              const int32_t* c_hyp, const uint8_t* c_top, const double* hs, const int16_t* mb,
                              const int16_t* my, const int8_t* h_z, const int8_t* h_iso, const int8_t* fwd_fam,
                              const int8_t* side, int64_t n_rows, const int64_t* partners, int32_t n_sets,
                              int32_t n_fwd_fam, int32_t mode, int64_t* out_off, const VhFamilyCols* cols, void* ws,
                              size_t ws_bytes, void* stream) {
  if (!p || n_spec < 0 || !out_off || mode < 0 || mode > 2 || n_sets < 1 || n_sets > 8 || n_fwd_fam < 1 ||
      n_rows < 0)
    return VH_E_ARG;
  const int G = n_fwd_fam * n_sets * 2;
  if (G > T2_GROUPS) return VH_E_LIMIT;
  hipStream_t st = (hipStream_t)stream;
  if (n_spec == 0) {
    if (mode == 0) {
      hipLaunchKernelGGL(fs_zero_kernel, dim3(1), dim3(1), 0, st, out_off);
      VH_CHECK_LAUNCH();
    }
    return 0;
  }
  if (mode == 2) {
    if (!ws || ws_bytes < vh_family_top2_workspace_bytes(n_spec)) return VH_E_WORKSPACE;
    return vh_scan_inclusive_i64(out_off + 1, n_spec, ws, ws_bytes, st);
  }
  if (!cand_off || !c_row || !c_hyp || !c_top || !hs || !mb || !my || !h_z || !h_iso || !fwd_fam || !side ||
      !partners)
    return VH_E_ARG;
  if (mode == 1 && (!cols || !cols->idx || !cols->group || !cols->rank || !cols->row || !cols->hyp || !cols->hs ||
                    !cols->mb || !cols->my || !cols->z || !cols->iso))
    return VH_E_ARG;
  T2Args A{n_spec, cand_off, c_row, c_hyp, c_top, hs, mb, my, fwd_fam, side, n_rows, partners, n_sets, G,
           p->min_matched_peaks, h_z, h_iso, out_off, mode == 1 ? *cols : VhFamilyCols{}};
  const unsigned g = (unsigned)((n_spec + T2_WAVES - 1) / T2_WAVES);
  if (mode == 0) hipLaunchKernelGGL(family_top2_kernel<0>, dim3(g), dim3(64 * T2_WAVES), 0, st, A);
  else hipLaunchKernelGGL(family_top2_kernel<1>, dim3(g), dim3(64 * T2_WAVES), 0, st, A);
  VH_CHECK_LAUNCH();
  return 0;
}
