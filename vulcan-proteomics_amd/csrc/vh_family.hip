// Family shortlist for multi-replicate tables (reference vulcan/gpu/family.py:90-136, select_family), on the
// bucket-join K1's per-row window counts.
//
// Wave per spectrum, in two passes (mode 0 counts each spectrum's output entries and scans them into offsets,
// mode 1 writes them):
//   1. the rows of the spectrum's hypothesis windows with count >= k1_threshold, ballot-compacted into LDS as
//      keys (K1 family, count desc, row desc, hypothesis asc) and bitonic-sorted;
//   2. rank inside the K1 family = position - first position of the family (binary search); the ones ranked
//      below top-K are kept and each adds its pair partner in every decoy set (count -1, not shortlisted);
//   3. those entries are sorted by (row, hypothesis) and reduced per row: lowest source hypothesis, highest
//      count, shortlisted if any source was; the canonical hypothesis is the spectrum's lowest hypothesis whose
//      window holds the row, else the lowest source. Output in row order.
// Limits (status bit 1, the spectrum's list is then incomplete): 64 hypotheses per spectrum, FS_E rows over
// the threshold, FS_C shortlist + partner entries; a window beyond slot_cap (flagged by vh_k1j_pairs) is
// skipped.
#include "vh_common.h"

namespace {

constexpr int FS_E = 256;    // rows over the K1 threshold per spectrum
constexpr int FS_C = 1024;   // shortlist + partner entries per spectrum
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
  int32_t* status;
};

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
    const int64_t sl = A.h_slot[h0 + k];
    const int len = hi > lo ? (int)(hi - lo) : 0;
    if (len == 0) continue;
    if (sl < 0 || sl + len > A.slot_cap) {   // uniform
      bad = true;
      continue;
    }
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
  if (n > FS_E) {
    bad = true;
    n = FS_E;
  }
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
  int ncl = nk * J1;
  if (ncl > FS_C) {
    bad = true;
    ncl = (FS_C / J1) * J1;
  }
  int np3 = 1;
  while (np3 < ncl) np3 <<= 1;
  for (int i = ncl + lane; i < np3; i += 64) cl[i] = ~0ull;
  wave_sync();
  if (np3 > 1) wave_bitonic_u64(cl, np3, false);
  wave_sync();
  // 3. one entry per row
  int nu = 0;
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
  if (lane == 0) {
    if (MODE == 0) {
      A.out_off[s + 1] = nu;
      if (s == 0) A.out_off[0] = 0;
    }
    if (bad) atomicOr(A.status, 1);
  }
}

__global__ void fs_zero_kernel(int64_t* p) { *p = 0; }

}  // namespace

extern "C" size_t vh_family_select_workspace_bytes(int32_t n_spec) {
  return vh_scan_ws_i64(n_spec > 0 ? n_spec : 1);
}

extern "C" int vh_family_select(const VhParams* p, int32_t n_spec, const int64_t* hyp_off, const int64_t* h_lo,
                                const int64_t* h_hi, const int64_t* h_slot, int64_t slot_cap,
                                const uint32_t* counts, const uint8_t* kfam, int64_t n_rows,
                                const int64_t* partners, int32_t n_sets, int32_t mode, int64_t* out_off,
                                int32_t* o_row, int32_t* o_hyp, int32_t* o_cnt, uint8_t* o_top, int32_t* status,
                                void* ws, size_t ws_bytes, void* stream) {
  if (!p || n_spec < 0 || !out_off || !status || (mode != 0 && mode != 1) || n_sets < 1 || n_sets > 8 ||
      n_rows < 0 || p->topk < 1)
    return VH_E_ARG;
  if (n_rows >= ((int64_t)1 << 28)) return VH_E_LIMIT;
  if (n_spec == 0) {
    if (mode == 0) {   // out_off[0] = 0 by a one-thread launch (no memset: the call stays capturable)
      hipLaunchKernelGGL(fs_zero_kernel, dim3(1), dim3(1), 0, (hipStream_t)stream, out_off);
      VH_CHECK_LAUNCH();
    }
    return 0;
  }
  if (!hyp_off || !h_lo || !h_hi || !h_slot || !kfam || !partners || (slot_cap > 0 && !counts)) return VH_E_ARG;
  if (mode == 1 && (!o_row || !o_hyp || !o_cnt || !o_top)) return VH_E_ARG;
  if (mode == 0 && (!ws || ws_bytes < vh_family_select_workspace_bytes(n_spec))) return VH_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  FsArgs A{n_spec, hyp_off, h_lo, h_hi, h_slot, slot_cap, counts, kfam, n_rows, partners, n_sets, p->topk,
           p->k1_threshold, out_off, o_row, o_hyp, o_cnt, o_top, status};
  const unsigned g = (unsigned)((n_spec + FS_WAVES - 1) / FS_WAVES);
  if (mode == 0) {
    hipLaunchKernelGGL(family_select_kernel<0>, dim3(g), dim3(64 * FS_WAVES), 0, st, A);
    VH_CHECK_LAUNCH();
    return vh_scan_inclusive_i64(out_off + 1, n_spec, ws, ws_bytes, st);
  }
  hipLaunchKernelGGL(family_select_kernel<1>, dim3(g), dim3(64 * FS_WAVES), 0, st, A);
  VH_CHECK_LAUNCH();
  return 0;
}
