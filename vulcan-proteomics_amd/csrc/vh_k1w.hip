// Wide-window K1 (open search), fused with the per-spectrum top-K and pair closure: the same candidate lists as
// vh_k1_topk (reference k1_kernel, vulcan/gpu/kernels.cu:112-190, select_topk, engine.py:284-312; open windows
// of vulcan/ptm/search.py:119-175), for spectra whose precursor windows span 10^5 - 10^6 rows.
//
// Design: one workgroup (4 waves) per spectrum walks the union of its windows one index bucket at a time, from
// the highest bucket down (buckets of 2^14 rows of a bin-major view built for this kernel). A slot is one
// (peak, fragment charge, directory bin) triple: the peak mass times the charge gives an f32 key range
// [bits(lo), bits(hi)] that touches one or more bins, and the bin-major layout keeps the postings of a
// (bin, bucket) cell contiguous, so a slot's cell in bucket b is [D(b), D(b + 1)) of its directory row D. Each
// thread holds up to KW_SPT slots in registers with the bounds of the current bucket and the start of the next
// one below (loaded one bucket ahead, so the directory reads overlap the counting of the bucket before). Per
// bucket:
//   * counting: the threads append their non-empty cells to an LDS list with one 64-bit LDS atomic per cell
//     (count << 32 | postings so far), which gives every cell its first position in the bucket's posting
//     sequence; the workgroup then takes that sequence in turn (posting m goes to thread m mod 256, which
//     finds its cell by a branchless search of the positions) and adds 1 to the posting's row in an LDS
//     histogram of the bucket when the posting's key bits lie in the slot's range inside the bin (two u16
//     counts per u32 word; a bucket whose cells hold 2^16 postings or more could wrap a u16 count and is
//     counted again in u32 sub-chunks). Whole cells are read: a cell holds about twice the matches of its
//     slot, which costs less than narrowing it by dependent loads;
//   * fragment charges: hypotheses of different maxfz overlap a bucket only for spectra without a known
//     charge; charges are added level by level (z < m for the smallest maxfz m first), and after each level the
//     hypotheses with that maxfz read their counts;
//   * offers: the rows of each overlapping hypothesis' window in the bucket, in strips of 256 words dealt to
//     the four waves (four words a lane), go to that wave's top-K buffer (threshold-pruned, vh_topk.h) when
//     count >= the hypothesis' K1 threshold; per-hypothesis scored / matched sums are LDS atomics.
// At the end the four buffers (each a superset of its wave's top K) are merged and one wave runs the shared
// final stage (top-K, closure, dedupe by row). Results do not depend on the bucket, cell or strip order.
#include <type_traits>

#include "vh_topk.h"

#define KW_WAVES 4
#define KW_THREADS (64 * KW_WAVES)
#define KW_SPT 4      // slots a thread keeps in registers (1024 a workgroup)
#define KW_WORK 512   // cells in the LDS list of one counting pass (further cells are counted by their owner)

namespace {

typedef unsigned short kw_u16x2 __attribute__((ext_vector_type(2)));

struct KwArgs {
  int n_spec;
  const int32_t* order;
  const int64_t* pk_off;
  const float* pk_mass;
  const int64_t* hyp_off;
  const int8_t* h_maxfz;
  const uint8_t* h_mode;
  const int64_t* h_lo;
  const int64_t* h_hi;
  uint32_t n_postings;
  int hist_words;   // u32 words of the LDS histogram
  int region_bytes; // bytes of the histogram / merge region
};

// One lookup (peak mass x fragment charge): its f32 key range and the directory bins it touches (bh < bl: none).
struct Lookup {
  uint32_t ka, kb;
  int bl, bh;
};

__device__ __forceinline__ Lookup kw_lookup(const DevP& P, const DevIdx& X, float pmass, int z) {
  Lookup L;
  const float m = pmass * (float)z;
  const float a = tol_bound(m, P.tol_lo, P.tol_da);
  const float b = tol_bound(m, P.tol_hi, P.tol_da);
  L.ka = a > 0.f ? fbits(a) : 0u;
  L.kb = fbits(b);
  const int S = X.bin_shift, base = (int)X.bin_base;
  L.bl = (int)(L.ka >> S) - base;
  L.bh = (int)(L.kb >> S) - base;
  L.bl = L.bl < 0 ? 0 : L.bl;
  L.bh = L.bh > X.n_bins - 1 ? X.n_bins - 1 : L.bh;
  if (b < P.fmin || a > P.fmax) L.bh = L.bl - 1;
  return L;
}

// The lookup's key range inside bin bn as bounds on a posting's key bits (posting >> 17 = f32 key & (2^S - 1)):
// lo | hi << 16.
__device__ __forceinline__ uint32_t kw_lohi(const DevIdx& X, const Lookup& L, int bn) {
  const int S = X.bin_shift;
  const uint32_t khi = (uint32_t)(bn + (int)X.bin_base) << S;
  const uint32_t all = (1u << S) - 1u;
  const uint32_t lo = L.ka > khi ? L.ka - khi : 0u;
  const uint32_t hi = L.kb < khi + all ? L.kb - khi : all;
  return lo | (hi << 16);
}

__device__ __forceinline__ bool kw_in(uint32_t pv, uint32_t lohi) {
  const uint32_t k = pv >> 17;
  return k >= (lohi & 0xFFFFu) && k <= (lohi >> 16);
}

// The count-of-counts bound: hcc[c] = keys (row, hypothesis) offered with count c (c >= 63 in hcc[63]) among the
// keys whose count exceeded the bound in force when their strip was read; the largest c with at least K tallied
// keys of count >= c (0 if none). The bound only rises, so the tally is complete for every c above it (and the
// bound itself had K keys), and a key of count below it has K keys of larger count: not in the spectrum's top K.
__device__ __forceinline__ int kw_tstar(const uint32_t* hcc, int K, int lane) {
  uint32_t x = hcc[63 - lane];   // suffix sums over c = 63 .. 0
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  const uint64_t ok = ballot(x >= (uint32_t)K);   // lane l: keys of count >= 63 - l
  return ok ? 63 - (int)__builtin_ctzll(ok) : 0;
}

__global__ __launch_bounds__(KW_THREADS, 3) void k1w_kernel(DevP P, DevIdx X, KwArgs A, K1Out O) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint32_t* hist = (uint32_t*)smem;
  uint64_t* bufs = (uint64_t*)(smem + A.region_bytes);
  unsigned long long* hmt = (unsigned long long*)(bufs + (size_t)KW_WAVES * O.capb);
  unsigned long long* wl_n = hmt + 64;                    // cells listed << 32 | their postings
  // the largest K-th key of any wave's buffer: K keys of the spectrum are at least this large, so every wave
  // prunes with it (a lower bound of the spectrum's K-th key)
  uint64_t* gthr = (uint64_t*)(wl_n + 1);
  uint32_t* hsc = (uint32_t*)(wl_n + 2);
  uint32_t* hcc = hsc + 64;                               // keys tallied by count (63: 63 or more)
  int* sh = (int*)(hcc + 64);                             // [1] clamped, [2..5] buffer fills, [6..7] slots
  float* pm = (float*)(sh + 8);                           // the spectrum's peak masses
  uint32_t* wl_pos = (uint32_t*)(pm + VH_K2_MAX_PEAKS);   // cell list: first position (KW_WORK + 1: end mark)
  uint32_t* wl_e0 = wl_pos + KW_WORK + 1;                 //   postings [e0, e1)
  uint32_t* wl_e1 = wl_e0 + KW_WORK;
  uint32_t* wl_lohi = wl_e1 + KW_WORK;                    //   key bits range
  const int tid = threadIdx.x, w = tid >> 6, lane = lane_id();
  const int wv = xcd_band_position(P.xcd_group);
  if (wv >= A.n_spec) return;   // the whole workgroup leaves
  const int s = A.order ? A.order[wv] : wv;
  const int64_t h0 = A.hyp_off[s];
  const int nh = (int)(A.hyp_off[s + 1] - h0);
  if (nh == 0 || nh > 64) {
    if (tid == 0) {
      O.n_cand[s] = 0;
      if (nh > 64) atomicOr(O.status, 1);
    }
    return;
  }
  const int64_t p0 = A.pk_off[s];
  int np = (int)(A.pk_off[s + 1] - p0);
  np = np < VH_K2_MAX_PEAKS ? np : VH_K2_MAX_PEAKS;   // K0 keeps at most max_peaks (<= 512)
  for (int i = tid; i < np; i += KW_THREADS) pm[i] = A.pk_mass[p0 + i];
  if (tid < 64) {
    hsc[tid] = 0u;
    hmt[tid] = 0ull;
    hcc[tid] = 0u;
  }
  if (tid < 8) sh[tid] = 0;
  if (tid == 0) *gthr = 0ull;
  // lane j of every wave holds hypothesis j: window, maxfz, K1 threshold
  int my_lo = 0, my_hi = 0, my_mfz = 0, my_thr = 0;
  if (lane < nh) {
    my_lo = (int)A.h_lo[h0 + lane];
    my_hi = (int)A.h_hi[h0 + lane];
    my_mfz = A.h_maxfz[h0 + lane];
    my_thr = (A.h_mode && A.h_mode[h0 + lane] == VH_MODE_OPEN) ? P.k1_threshold_open : P.k1_threshold;
  }
  const bool live = lane < nh && my_hi > my_lo;
  int rlo = live ? my_lo : 0x7fffffff, rhi = live ? my_hi : 0, zt = live ? my_mfz : 1;
  for (int o = 32; o > 0; o >>= 1) {
    const int a = __shfl_xor(rlo, o), b = __shfl_xor(rhi, o), c = __shfl_xor(zt, o);
    rlo = a < rlo ? a : rlo;
    rhi = b > rhi ? b : rhi;
    zt = c > zt ? c : zt;
  }
  const int nzt = zt - 1 > 0 ? zt - 1 : 1;   // fragment charges counted: 1 .. max maxfz - 1
  const int LB = X.bucket_log2, CH = 1 << LB;
  const int K = P.topk;
  uint64_t* mybuf = bufs + (size_t)w * O.capb;
  int nbuf = 0;
  uint64_t thresh = 0;
  int tstar = 0;   // count bound: K keys of the spectrum have a count >= tstar
  bool clamped = false;
  __syncthreads();   // peak masses in LDS
  // ---- slots: lookups i = (peak i / nzt, fragment charge 1 + i % nzt), each with its bins in order, numbered
  // by a block scan; slot g goes to thread g mod 256 (register j = g / 256) when it belongs to one of the first
  // lookups whose slots all fit the registers; the later lookups (from i_over on) read the directory at every
  // bucket, by the thread that owns the lookup (i mod 256). Staged through the histogram region.
  const uint32_t nb = (uint32_t)X.n_buckets;
  const int nlk = np * nzt;
  uint32_t* st_dir = hist;
  uint32_t* st_lohi = st_dir + KW_SPT * KW_THREADS;
  int8_t* st_z = (int8_t*)(st_lohi + KW_SPT * KW_THREADS);
  int* scan_tmp = (int*)wl_pos;   // the cell list is unused until the walk
  if (tid == 0) {
    sh[6] = nlk;                    // i_over
    sh[7] = 0x7fffffff;             // register slots (when some lookup is over)
  }
  int n_slots = 0;
  for (int i0 = 0; i0 < nlk; i0 += KW_THREADS) {
    const int i = i0 + tid;
    int z = 0, nbn = 0;
    Lookup L = {0u, 0u, 0, -1};
    if (i < nlk) {
      const int p = i / nzt;
      z = 1 + (i - p * nzt);
      L = kw_lookup(P, X, pm[p], z);
      nbn = L.bh >= L.bl ? L.bh - L.bl + 1 : 0;
    }
    int x = nbn;   // block exclusive scan of the bin counts
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    __syncthreads();   // scan_tmp free; sh[6..7] set
    if (lane == 63) scan_tmp[w] = x;
    __syncthreads();
    int before = 0, tot = 0;
#pragma unroll
    for (int j = 0; j < KW_WAVES; ++j) {
      before += j < w ? scan_tmp[j] : 0;
      tot += scan_tmp[j];
    }
    const int g = n_slots + before + x - nbn;
    if (nbn) {
      if (g + nbn > KW_SPT * KW_THREADS) {
        atomicMin(&sh[6], i);
        atomicMin(&sh[7], g);
      } else {
        for (int bn = L.bl; bn <= L.bh; ++bn) {
          st_dir[g + bn - L.bl] = (uint32_t)bn * nb;
          st_lohi[g + bn - L.bl] = kw_lohi(X, L, bn);
          st_z[g + bn - L.bl] = (int8_t)z;
        }
      }
    }
    n_slots += tot;
  }
  __syncthreads();
  const int i_over = sh[6];
  const int n_reg = sh[7] < n_slots ? sh[7] : n_slots;
  uint32_t s_dir[KW_SPT], s_lohi[KW_SPT], s_lo[KW_SPT], s_hi[KW_SPT], s_nx[KW_SPT];
  int s_z[KW_SPT];
  int ns = 0;   // registers in use: slots tid + 256 j, j < ns
#pragma unroll
  for (int j = 0; j < KW_SPT; ++j) {
    const int g = tid + j * KW_THREADS;
    const bool in = g < n_reg;
    s_dir[j] = in ? st_dir[g] : 0u;
    s_lohi[j] = in ? st_lohi[g] : 0u;
    s_z[j] = in ? st_z[g] : 0;
    s_lo[j] = s_hi[j] = s_nx[j] = 0u;
    ns += in ? 1 : 0;
  }
  __syncthreads();   // staging read before the histogram is cleared
  const bool packed_ok = A.hist_words >= CH / 2;
  const int sub_rows = A.hist_words < CH ? A.hist_words : CH;   // u32 sub-chunk rows
  const int bk_top = rhi > rlo ? (rhi - 1) >> LB : -1, bk_bot = rlo >> LB;
  if (bk_top >= 0) {
#pragma unroll
    for (int j = 0; j < KW_SPT; ++j)
      if (j < ns) {
        s_lo[j] = X.offsets[s_dir[j] + bk_top];
        s_hi[j] = X.offsets[s_dir[j] + bk_top + 1];
        s_nx[j] = bk_top > 0 ? X.offsets[s_dir[j] + bk_top - 1] : 0u;
      }
  }
  for (int bk = bk_top; bk >= 0 && bk >= bk_bot; --bk) {
    if (bk < bk_top) {   // the slots' cells of this bucket; the start of the next one below is loaded ahead
#pragma unroll
      for (int j = 0; j < KW_SPT; ++j)
        if (j < ns) {
          s_hi[j] = s_lo[j];
          s_lo[j] = s_nx[j];
          s_nx[j] = bk > 0 ? X.offsets[s_dir[j] + bk - 1] : 0u;
        }
    }
    const int c0 = bk << LB;
    const uint64_t ov = ballot(live && my_lo < c0 + CH && my_hi > c0);
    if (!ov) continue;   // uniform
    const int rb = c0;
    // Count fragment charges [z0, z1) of rows [r0, r0 + nr) of this bucket into hist; returns the postings of
    // the listed cells (an upper bound of the matches). Called by the whole workgroup.
    auto count = [&](int z0, int z1, int r0, int nr, bool packed) -> uint32_t {
      auto visit = [&](uint32_t pv) {
        const int rr = rb + (int)(pv & 0xFFFFu) - r0;
        if (rr < 0 || rr >= nr) return;
        if (packed) atomicAdd(&hist[rr >> 1], 1u << ((rr & 1) << 4));
        else atomicAdd(&hist[rr], 1u);
      };
      auto add_cell = [&](uint32_t e0, uint32_t e1, uint32_t lh) {
        const uint32_t groups = (((e1 + 3u) & ~3u) - (e0 & ~3u)) >> 2;   // 16-byte groups holding [e0, e1)
        const unsigned long long old = atomicAdd(wl_n, (1ull << 32) | (unsigned long long)groups);
        const uint32_t idx = (uint32_t)(old >> 32), pos = (uint32_t)old;
        if (idx < KW_WORK) {
          wl_pos[idx] = pos;
          wl_e0[idx] = e0;
          wl_e1[idx] = e1;
          wl_lohi[idx] = lh;
        } else {   // list full: the owner counts its cell (the first such cell marks the list's end)
          if (idx == KW_WORK) wl_pos[KW_WORK] = pos;
          for (uint32_t e = e0; e < e1; ++e) {
            const uint32_t pv = X.postings[e];
            if (kw_in(pv, lh)) visit(pv);
          }
        }
      };
      // (A) list the non-empty cells
#pragma unroll
      for (int j = 0; j < KW_SPT; ++j)
        if (j < ns && s_z[j] >= z0 && s_z[j] < z1 && s_hi[j] > s_lo[j]) add_cell(s_lo[j], s_hi[j], s_lohi[j]);
      if (i_over < nlk) {   // lookups beyond the registers: directory read at this bucket
        for (int i = tid; i < nlk; i += KW_THREADS) {
          if (i < i_over) continue;
          const int p = i / nzt, z = 1 + (i - p * nzt);
          if (z < z0 || z >= z1) continue;
          const Lookup L = kw_lookup(P, X, pm[p], z);
          for (int bn = L.bl; bn <= L.bh; ++bn) {
            const uint32_t d = (uint32_t)bn * nb + (uint32_t)bk;
            const uint32_t e0 = X.offsets[d], e1 = X.offsets[d + 1];
            if (e1 > e0) add_cell(e0, e1, kw_lohi(X, L, bn));
          }
        }
      }
      __syncthreads();
      // (B) 16-byte group m of the listed cells' sequence to thread m mod 256: one 128-bit load of four postings
      // (aligned to four entries; entries outside the cell are skipped)
      const unsigned long long wn = *wl_n;
      const uint32_t n_all = (uint32_t)(wn >> 32);
      const uint32_t n = n_all < KW_WORK ? n_all : KW_WORK;
      const uint32_t total = n_all <= KW_WORK ? (uint32_t)wn : wl_pos[KW_WORK];
      for (uint32_t m0 = tid; m0 < total; m0 += 4 * KW_THREADS) {
        uint32_t at[4] = {0u, 0u, 0u, 0u};   // last cell whose first position is <= m
        for (uint32_t len = n; len > 1;) {
          const uint32_t half = len >> 1;
#pragma unroll
          for (int u = 0; u < 4; ++u)
            at[u] = wl_pos[at[u] + half] <= m0 + (uint32_t)(u * KW_THREADS) ? at[u] + half : at[u];
          len -= half;
        }
        uint4 pv[4];
        uint32_t lh[4], eb[4], ea[4], ee[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t m = m0 + (uint32_t)(u * KW_THREADS);
          lh[u] = wl_lohi[at[u]];
          ea[u] = wl_e0[at[u]];
          ee[u] = m < total ? wl_e1[at[u]] : 0u;   // an empty range past the sequence
          eb[u] = (ea[u] & ~3u) + 4u * (m - wl_pos[at[u]]);
          if (m >= total) pv[u] = make_uint4(0u, 0u, 0u, 0u);
          else if (eb[u] + 4u <= A.n_postings) pv[u] = *(const uint4*)&X.postings[eb[u]];
          else {   // the index's last group: entries past its end are not read
            pv[u].x = X.postings[eb[u]];
            pv[u].y = eb[u] + 1u < ee[u] ? X.postings[eb[u] + 1u] : 0u;
            pv[u].z = eb[u] + 2u < ee[u] ? X.postings[eb[u] + 2u] : 0u;
            pv[u].w = eb[u] + 3u < ee[u] ? X.postings[eb[u] + 3u] : 0u;
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t v[4] = {pv[u].x, pv[u].y, pv[u].z, pv[u].w};
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const uint32_t e = eb[u] + (uint32_t)t;
            if (e >= ea[u] && e < ee[u] && kw_in(v[t], lh[u])) visit(v[t]);
          }
        }
      }
      return (uint32_t)wn;
    };
    // the bucket as one packed chunk when it fits and cannot wrap a u16 count, else u32 sub-chunks; done_m: the
    // hypotheses with maxfz <= done_m already made their offers for this bucket (before a switch to u32)
    bool packed = packed_ok;
    int done_m = 1;
    for (int r0 = c0; r0 < c0 + CH;) {
      const int nr = packed ? CH : sub_rows;
      const uint64_t ovc = ballot(live && my_lo < r0 + nr && my_hi > r0);
      if (!ovc) {
        r0 += nr;
        continue;
      }
      const int words = packed ? nr / 2 : nr;
      int mprev = 1;
      bool redo = false;
      uint32_t posts = 0;   // postings of the cells listed for this chunk so far
      for (;;) {
        // next maxfz level among the hypotheses over this chunk
        const int cand = ((ovc >> lane) & 1ull) && my_mfz > mprev ? my_mfz : 0x7fff;
        int m = cand;
        for (int o = 32; o > 0; o >>= 1) {
          const int x = __shfl_xor(m, o);
          m = x < m ? x : m;
        }
        const bool first = mprev == 1;
        if (m == 0x7fff && !first) break;
        if (m == 0x7fff) m = 1;   // no hypothesis with maxfz > 1: nothing to count, still offer (counts 0)
        if (first) {
          uint4* h4 = (uint4*)hist;
          for (int i = tid; i < words / 4; i += KW_THREADS) h4[i] = make_uint4(0u, 0u, 0u, 0u);
          for (int i = (words & ~3) + tid; i < words; i += KW_THREADS) hist[i] = 0u;
        }
        if (tid == 0) *wl_n = 0ull;
        __syncthreads();   // cleared, or the previous level's offers and list reads are done
        if (m > mprev) posts += count(mprev, m, r0, nr, packed);
        __syncthreads();
        if (packed && posts >= 65536u) {   // a u16 count may have wrapped: count this bucket again in u32
          redo = true;
          done_m = mprev;
          break;
        }
        // offers of the hypotheses with maxfz == m (not yet made for this bucket): strips of 256 words (512 rows
        // packed, 256 rows in u32) dealt to the waves, four words a lane, rows descending. The fast path sums
        // scored / matched and tests whether any of the lane's rows could beat the wave's current K-th key
        // (count above its count, or equal count from a higher row); only a strip where some lane could goes
        // through the per-row offers.
        uint64_t hs = ballot(((ovc >> lane) & 1ull) && my_mfz == m && m > done_m);
        while (hs) {
          const int q = (int)__builtin_ctzll(hs);
          hs &= hs - 1ull;
          const int qlo = __shfl(my_lo, q), qhi = __shfl(my_hi, q), qthr = __shfl(my_thr, q);
          const int a = qlo > r0 ? qlo : r0, b = qhi < r0 + nr ? qhi : r0 + nr;
          const bool whole = a == r0 && b == r0 + nr;   // the window covers the chunk: no row masks
          uint32_t sc = 0, mt = 0;   // per lane: < 2^32 (a chunk holds fewer postings)
          auto strips = [&](auto PK, auto WH) {   // specialised on the packed and whole flags (uniform)
          constexpr bool pk = decltype(PK)::value, wh = decltype(WH)::value;
          constexpr int rpw = pk ? 2 : 1;                               // rows per word
          const int wa = (a - r0) / rpw, wb = (b - 1 - r0) / rpw + 1;  // words holding rows [a, b)
          const int W0 = wa & ~3, W1 = (wb + 3) & ~3;                   // in whole 16-byte groups
          const uint64_t qlow = (uint64_t)(63 - q) << 14;
          const int t_tally = tstar + 1 > qthr ? tstar + 1 : (qthr > 1 ? qthr : 1);
          for (int k = w; W1 - 256 * k > W0; k += KW_WAVES) {
            const int wl = W1 - 256 * (k + 1) + 4 * lane;   // the lane's words wl .. wl + 3
            const bool act = wl >= W0;
            uint4 v4 = act ? *(const uint4*)&hist[wl] : make_uint4(0u, 0u, 0u, 0u);
            uint32_t word[4] = {v4.x, v4.y, v4.z, v4.w};
            if (!wh) {   // rows outside [a, b) count 0 here
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                const int rl = r0 + (wl + u) * rpw;   // the word's lower row
                if (pk)
                  word[u] &= ((rl >= a && rl < b) ? 0x0000FFFFu : 0u) | ((rl + 1 >= a && rl + 1 < b) ? 0xFFFF0000u : 0u);
                else
                  word[u] = (rl >= a && rl < b) ? word[u] : 0u;
              }
            }
            uint32_t cmax = 0;
            if (pk) {   // per word: non-zero halves, their sum and maximum, as packed u16 operations
              kw_u16x2 cm = {0, 0};
              const uint32_t ones = 0x00010001u;
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                uint32_t nz;   // min(half, 1) per half (the compiler splits the vector min into compares)
                asm("v_pk_min_u16 %0, %1, %2" : "=v"(nz) : "v"(word[u]), "v"(ones));
                sc = __builtin_amdgcn_sad_u16(nz, 0u, sc);
                mt = __builtin_amdgcn_sad_u16(word[u], 0u, mt);
                cm = __builtin_elementwise_max(cm, __builtin_bit_cast(kw_u16x2, word[u]));
              }
              cmax = cm.x > cm.y ? cm.x : cm.y;
            } else {
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                const uint32_t x = word[u];
                sc += x != 0u;
                mt += x;
                cmax = x > cmax ? x : cmax;
              }
            }
            const uint64_t g = *gthr;   // the largest K-th key of any wave so far
            thresh = g > thresh ? g : thresh;
            const uint32_t tc = (uint32_t)(thresh >> 48);
            const int rtop = r0 + (wl + 3) * rpw + rpw - 1;   // the lane's highest row
            const bool could = act && (int)cmax >= qthr &&
                               (cmax > tc || (cmax == tc && (((uint64_t)(uint32_t)rtop << 20) | qlow) >
                                                                (thresh & 0xFFFFFFFFFFFFull)));
            clamped |= (int)cmax >= qthr && cmax > 32767u;   // rows outside the window count 0 here
            if (!ballot(act && ((int)cmax >= t_tally || qthr <= 0))) continue;
            // keys of count >= the tally bound go into the count-of-counts (rows outside the window read 0)
            if (act && (int)cmax >= t_tally) {
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                const uint32_t cA = pk ? word[u] >> 16 : word[u], cB = word[u] & 0xFFFFu;
                if ((int)cA >= t_tally) atomicAdd(&hcc[cA < 63u ? cA : 63u], 1u);
                if (pk && (int)cB >= t_tally) atomicAdd(&hcc[cB < 63u ? cB : 63u], 1u);
              }
            }
            if (!ballot(could)) continue;
            // the lane's rows that pass the threshold and beat the K-th key (bit 2 u: the word's higher row,
            // bit 2 u + 1: its lower row), offered one a lane per round
            uint32_t cand = 0;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const int wi = wl + u;
              const int rA = r0 + wi * rpw + rpw - 1, rB = r0 + wi * rpw;   // the word's higher and lower row
              const uint32_t cA = pk ? word[u] >> 16 : word[u], cB = word[u] & 0xFFFFu;
              const uint64_t kA = ((uint64_t)(cA < 32767u ? cA : 32767u) << 48) | ((uint64_t)(uint32_t)rA << 20) | qlow;
              const uint64_t kB = ((uint64_t)(cB < 32767u ? cB : 32767u) << 48) | ((uint64_t)(uint32_t)rB << 20) | qlow;
              if (act && rA >= a && rA < b && (int)cA >= qthr && kA > thresh) cand |= 1u << (2 * u);
              if (pk && act && rB >= a && rB < b && (int)cB >= qthr && kB > thresh) cand |= 2u << (2 * u);
            }
            for (;;) {
              const bool has = cand != 0u;
              if (!ballot(has)) break;
              const int bit = has ? __builtin_ctz(cand) : 0;
              cand &= cand - 1u;
              const int u = bit >> 1;
              uint32_t x = word[0];
#pragma unroll
              for (int v = 1; v < 4; ++v) x = u == v ? word[v] : x;
              const bool hi_row = (bit & 1) == 0;
              const uint32_t c = hi_row ? (pk ? x >> 16 : x) : (x & 0xFFFFu);
              const int row = r0 + (wl + u) * rpw + (hi_row ? rpw - 1 : 0);
              topk_offer<1>(has, c, row, q, 0, mybuf, nullptr, O.capb, K, nbuf, thresh, clamped);
            }
            if (lane == 0 && thresh > g) atomicMax((unsigned long long*)gthr, (unsigned long long)thresh);
          }
          };
          if (packed) {
            if (whole) strips(std::true_type(), std::true_type());
            else strips(std::true_type(), std::false_type());
          } else {
            if (whole) strips(std::false_type(), std::true_type());
            else strips(std::false_type(), std::false_type());
          }
          sc = wave_sum(sc);
          const unsigned long long mts = wave_sum((unsigned long long)mt);
          if (lane == 0 && (sc || mts)) {
            atomicAdd(&hsc[q], sc);
            atomicAdd(&hmt[q], mts);
          }
        }
        mprev = m;
        if (m == 1) break;
      }
      __syncthreads();   // offers done before the histogram is cleared again
      tstar = kw_tstar(hcc, K, lane);
      if (((uint64_t)tstar << 48) > thresh) thresh = (uint64_t)tstar << 48;
      if (redo) {
        packed = false;   // same bucket, u32 sub-chunks
        continue;
      }
      r0 += nr;
    }
  }
  // ---- merge: every wave keeps its best K, wave 0 selects over their union
  nbuf = topk_compact(mybuf, nullptr, O.capb, nbuf, K, thresh);
  const bool any_clamped = ballot(clamped) != 0ull;
  if (lane == 0) {
    sh[2 + w] = nbuf;
    if (any_clamped) sh[1] = 1;
  }
  __syncthreads();
  int at = 0, tot = 0;
  for (int j = 0; j < KW_WAVES; ++j) {
    at += j < w ? sh[2 + j] : 0;
    tot += sh[2 + j];
  }
  uint64_t* merged = (uint64_t*)smem;   // the histogram region is free now
  for (int i = lane; i < nbuf; i += 64) merged[at + i] = mybuf[i];
  __syncthreads();
  if (tid < nh) {
    O.hyp_scored[h0 + tid] = hsc[tid];
    O.hyp_matched[h0 + tid] = hmt[tid];
  }
  if (w == 0) k1_finish<1>(P, X, s, h0, merged, nullptr, tot, 0ull, sh[1] != 0, merged + 4 * O.capb, O);
}

}  // namespace

extern "C" int vh_k1w_topk(const VhParams* p, const VhIndexView* ix, int32_t n_spec, const int32_t* order,
                           const int64_t* pk_off, const float* pk_mass, const int64_t* hyp_off,
                           const int8_t* h_maxfz, const uint8_t* h_mode, const int64_t* h_lo, const int64_t* h_hi,
                           int32_t cand_stride, int32_t* n_cand, int32_t* c_row, int32_t* c_hyp, int32_t* c_cnt,
                           uint32_t* hyp_scored, unsigned long long* hyp_matched, int32_t* status, void* stream) {
  if (!p || !ix || !ix->offsets || (ix->n_postings && !ix->postings)) return VH_E_ARG;
  if (n_spec < 0 || !n_cand || !c_row || !c_hyp || !c_cnt || !hyp_scored || !hyp_matched || !status) return VH_E_ARG;
  if (ix->bucket_log2 != VH_K1W_BUCKET_LOG2 || ix->bin_shift < 4 || ix->bin_shift > 15) return VH_E_ARG;
  if ((uintptr_t)ix->postings % 16u || ix->n_postings >= (1LL << 32)) return VH_E_ARG;   // 128-bit group loads
  if (ix->n_rows >= (1LL << 28)) return VH_E_LIMIT;
  if (p->topk < 1 || p->topk > 1024 || (p->pair_closure && !ix->partner)) return VH_E_ARG;
  if (p->max_peaks > VH_K2_MAX_PEAKS) return VH_E_LIMIT;
  if (cand_stride < (p->pair_closure ? 2 * p->topk : p->topk)) return VH_E_CAPACITY;
  if (n_spec == 0) return 0;
  K1Out O = {};
  O.cand_stride = cand_stride;
  O.n_cand = n_cand;
  O.c_row = c_row;
  O.c_hyp = c_hyp;
  O.c_cnt = c_cnt;
  O.hyp_scored = hyp_scored;
  O.hyp_matched = hyp_matched;
  O.status = status;
  int K = p->topk;
  int capb = 1;
  while (capb < (K + 64 > 2 * K ? K + 64 : 2 * K)) capb <<= 1;
  int clos = 1;
  while (clos < (2 * K > 64 ? 2 * K : 64)) clos <<= 1;
  O.capb = capb;
  O.clos_n = clos;
  // region: the histogram during the walk, then the merged buffers (4 capb) and the closure keys (clos)
  const int merge_bytes = (KW_WAVES * capb + clos) * 8;
  const int tail = KW_WAVES * capb * 8 + 64 * 8 + 16 + 64 * 4 + 8 * 4 + VH_K2_MAX_PEAKS * 4 + (4 * KW_WORK + 1) * 4 + 64 * 4;
  int dev = 0, lds_max = 65536;
  if (hipGetDevice(&dev) == hipSuccess)
    hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  const int full = (1 << VH_K1W_BUCKET_LOG2) / 2 * 4;   // a whole bucket of packed u16 counts (32 KB)
  int region = full;
  while (region + tail > lds_max && region > 4096) region >>= 1;   // u32 sub-chunks in a smaller region
  if (region < merge_bytes) region = merge_bytes;
  const int stage_bytes = (KW_SPT * KW_THREADS * 9 + 15) & ~15;   // slot staging (dir, key bits, charge)
  if (region < stage_bytes) region = stage_bytes;
  if (region + tail > lds_max) return VH_E_LIMIT;
  KwArgs A{n_spec, order, pk_off, pk_mass, hyp_off, h_maxfz, h_mode, h_lo, h_hi, (uint32_t)ix->n_postings, region / 4,
           region};
  const size_t lds = (size_t)region + (size_t)tail;
  hipStream_t st = (hipStream_t)stream;
  if (lds > 65536) hipFuncSetAttribute((const void*)k1w_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const unsigned g = xcd_band_grid(n_spec, p->xcd_group);
  k1w_kernel<<<g, KW_THREADS, lds, st>>>(vh_make_devp(p), vh_make_devidx(ix), A, O);
  VH_CHECK_LAUNCH();
  return 0;
}
