// Bucket-join K1: the K1 candidate lists (count definition of the reference's k1_kernel,
// vulcan/gpu/kernels.cu:112-190; top-K of select_topk, engine.py:284-312) computed block-major.
//
// Why: the spectrum-major K1 (vh_k1.hip) looks up every (peak, fragment charge) in every index bucket of the
// spectrum's windows with random global loads (a directory line and a posting line per lookup), so it runs at
// the latency of scattered 128-byte lines and fetches ~12x the bytes it uses. Here every block of consecutive
// rows (a few thousand postings, key-sorted; VhBlockView) is read once per search step, coalesced, into LDS by
// one workgroup, and all hypotheses whose precursor windows overlap the block are matched against it there:
//   pairs   per hypothesis, the blocks of its window -> (block, hypothesis) pairs grouped by block (atomic
//           counting sort: count per block, exclusive scan, scatter) and a slot per window row;
//   count   workgroup per block: stage keys + rows, build a directory with about one posting per cell (cell t =
//           first posting whose key offset >> tsh reaches t), then each wave takes a pair: lanes take peaks; for
//           each fragment charge the exact f32 key range [bits(lo bound), bits(hi bound)] starts at its low
//           bound's cell and the postings from there are examined four at a time, branch-free, until one passes
//           the high bound; those in range whose row lies in the window add to a wave-private LDS histogram,
//           which is written to the pair's slots (coalesced);
//   select  wave per spectrum: reads its slots (rows descending, as vh_k1.hip scans), per-hypothesis scored /
//           matched, ballot top-K buffer, pair closure and dedupe -- the code of vh_k1.hip (vh_topk.h).
// Counts are exact integers, so the lists are bit-identical to vh_k1_topk / vh_k1_features.
#include "vh_topk.h"

#define J_THREADS 512
#define J_WAVES (J_THREADS / 64)
#define J_PCH 64    // pairs whose metadata is staged at once

// One (block, hypothesis) pair as the count kernel reads it (32 bytes, written by vh_k1j_pairs): the spectrum's
// first peak and peak count, the hypothesis' first slot, window and fragment-charge bound, and (feature mode) the
// spectrum's intensity shift, so a workgroup gathers its pairs with one load each.
struct PairMeta {
  int64_t p0;     // first peak of the spectrum
  int64_t slot;   // first slot of the hypothesis (row lo)
  int32_t lo, hi; // window rows
  int32_t np;     // peak count
  int32_t mq;     // maxfz | (iq shift + 2048) << 16
};

// slot word: u32 count (mode 1); mode 2: count << 43 | b count (clamped 32767) << 28 | iq_log(b sum) << 14 |
// iq_log(y sum), the count below 2^20
template <int MODE>
struct SlotType {
  typedef uint32_t T;
};
template <>
struct SlotType<2> {
  typedef unsigned long long T;
};

struct JArgs {
  // block layout
  const int32_t* row0;
  const uint32_t* off;
  const uint32_t* key;
  const uint16_t* row;
  int maxp;        // LDS posting capacity (>= largest block + 4 sentinels, multiple of 8)
  int tcells;      // directory cells (power of two)
  int hist_rows;   // histogram slots per wave (>= largest block's rows)
  // pairs and hypotheses
  const uint32_t* blk_pairs;
  const PairMeta* pairs;
  int64_t pair_cap;
  // peaks
  const float* pk_mass;
  const float* pk_int;
  // output
  void* counts;
  int32_t* status;
};

// ------------------------------------------------------------------------------------------------ pairs
__device__ __forceinline__ int64_t block_of_row(const int32_t* __restrict__ row0, int64_t n_blocks, int64_t r) {
  int64_t a = 0, b = n_blocks;   // row0[a] <= r < row0[b]
  while (b - a > 1) {
    const int64_t m = (a + b) >> 1;
    if (row0[m] <= r)
      a = m;
    else
      b = m;
  }
  return a;
}

__global__ __launch_bounds__(256) void pj_zero_kernel(uint32_t* a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0u;
}

__global__ __launch_bounds__(256) void pj_count_kernel(const int32_t* __restrict__ row0, int64_t n_blocks,
                                                       int32_t n_spec, const int64_t* __restrict__ hyp_off,
                                                       const int64_t* __restrict__ h_lo,
                                                       const int64_t* __restrict__ h_hi, int64_t max_hyp,
                                                       int64_t* h_slot, uint32_t* blk_pairs) {
  const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= max_hyp) return;
  const int64_t n_hyp = hyp_off[n_spec] < max_hyp ? hyp_off[n_spec] : max_hyp;
  if (h == 0) h_slot[0] = 0;
  int64_t len = 0;
  if (h < n_hyp) {
    const int64_t lo = h_lo[h], hi = h_hi[h];
    if (hi > lo) {
      len = hi - lo;
      const int64_t b0 = block_of_row(row0, n_blocks, lo), b1 = block_of_row(row0, n_blocks, hi - 1);
      for (int64_t b = b0; b <= b1; ++b) atomicAdd(&blk_pairs[b], 1u);
    }
  }
  h_slot[h + 1] = len;
}

__global__ __launch_bounds__(256) void pj_scatter_kernel(const int32_t* __restrict__ row0, int64_t n_blocks,
                                                         int32_t n_spec, const int64_t* __restrict__ hyp_off,
                                                         const int32_t* __restrict__ h_spec,
                                                         const int64_t* __restrict__ h_lo,
                                                         const int64_t* __restrict__ h_hi,
                                                         const int8_t* __restrict__ h_maxfz,
                                                         const int64_t* __restrict__ pk_off,
                                                         const int32_t* __restrict__ s_qe, int64_t max_hyp,
                                                         const int64_t* __restrict__ h_slot, int64_t slot_cap,
                                                         const uint32_t* __restrict__ blk_pairs, uint32_t* cursor,
                                                         PairMeta* pairs, int64_t pair_cap, int32_t* status) {
  const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= max_hyp) return;
  const int64_t n_hyp = hyp_off[n_spec] < max_hyp ? hyp_off[n_spec] : max_hyp;
  if (h == 0 && ((int64_t)blk_pairs[n_blocks] > pair_cap || h_slot[n_hyp] > slot_cap)) atomicOr(status, 8);
  if (h >= n_hyp) return;
  const int64_t lo = h_lo[h], hi = h_hi[h];
  if (hi <= lo) return;
  const int s = h_spec[h];
  PairMeta M;
  M.p0 = pk_off[s];
  M.np = (int32_t)(pk_off[s + 1] - M.p0);
  M.slot = h_slot[h];
  M.lo = (int32_t)lo;
  M.hi = (int32_t)hi;
  M.mq = (int32_t)(uint8_t)h_maxfz[h] | ((s_qe ? s_qe[s] : 0) + 2048) << 16;
  if (M.slot + (hi - lo) > slot_cap) M.hi = M.lo;   // overflow (flagged above): the count kernel skips it
  const int64_t b0 = block_of_row(row0, n_blocks, lo), b1 = block_of_row(row0, n_blocks, hi - 1);
  for (int64_t b = b0; b <= b1; ++b) {
    const int64_t pos = (int64_t)blk_pairs[b] + (int64_t)atomicAdd(&cursor[b], 1u);
    if (pos < pair_cap) pairs[pos] = M;
  }
}

// ------------------------------------------------------------------------------------------------ count
// peaks [pc + u*64 + lane] of a pair, u < U (mass -1 past the peak count; intensities only in feature mode)
template <int MODE, int U>
__device__ __forceinline__ void k1j_load_peaks(const JArgs& A, const PairMeta& M, int pc, float (&pm)[U],
                                               float (&pi)[U]) {
  const int lane = lane_id();
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int p = pc + u * 64 + lane;
    pm[u] = p < M.np ? A.pk_mass[M.p0 + p] : -1.f;
    pi[u] = (MODE == 2 && p < M.np) ? A.pk_int[M.p0 + p] : 0.f;
  }
}

// Staged block in LDS (shared by the workgroup's waves)
struct JBlock {
  const uint32_t* keys;
  const uint16_t* rows;
  const uint16_t* table;
  int n, tsh, rb0;
  uint32_t kmin, kmax;
};

// peaks per lane held in registers (the first 256 of a spectrum; more are loaded on demand, four per lane at a
// time): four measured faster than eight (count 0.745 -> 0.728 ms on the proteome-scale batch)
#define J_PU 4

// One (peak, fragment charge) range lookup: the exact f32 key range [bits(lo bound), bits(hi bound)] starts at
// its low bound's directory cell (cells hold about one posting each); postings are examined four at a time
// (the staged keys end with four all-ones sentinels, so the loads need no bounds test) until one passes the
// high bound, and those in range whose row lies in [ra, rb) add to the histogram.
template <int MODE>
__device__ __forceinline__ void k1j_lookup(const DevP& P, const JBlock& B, float pmu, unsigned long long inc, int z,
                                           int ra, int rb, uint32_t* hist) {
  const float m = pmu * (float)z;
  const float a = tol_bound(m, P.tol_lo, P.tol_da);
  const float bd = tol_bound(m, P.tol_hi, P.tol_da);
  if (bd < P.fmin || a > P.fmax) return;   // (about half of the charge-2 lookups: above the fragment range)
  const uint32_t ka = a > 0.f ? fbits(a) : 0u;
  const uint32_t kb = fbits(bd);
  if (kb < B.kmin || ka > B.kmax) return;
  int e = B.table[ka <= B.kmin ? 0u : (ka - B.kmin) >> B.tsh];   // first posting of ka's cell
  for (;;) {
    uint32_t k4[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) k4[v] = B.keys[e + v];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      if (k4[v] >= ka && k4[v] <= kb) {
        const uint32_t rr = B.rows[e + v];
        const int row = B.rb0 + (int)(rr & 0x7FFFu);
        if (row >= ra && row < rb) {
          if (MODE == 2)
            atomicAdd(&((unsigned long long*)hist)[2 * (row - ra) + (int)(rr >> 15)], inc);
          else
            atomicAdd(&hist[row - ra], 1u);
        }
      }
    }
    if (k4[3] > kb) break;   // (past the end the sentinels stop the scan)
    e += 4;
  }
}

// Counts of one pair for the rows [ra, rb) into the wave's histogram (zeroed here): lanes take peaks (pm0 / pi0:
// the first 64 * J_PU, already in registers), each (peak, fragment charge) range is looked up in the staged block.
template <int MODE>
__device__ __forceinline__ void k1j_pair_rows(const DevP& P, const JArgs& A, const PairMeta& M, int ra, int rb,
                                              const float (&pm0)[J_PU], const float (&pi0)[J_PU], const JBlock& B,
                                              uint32_t* hist) {
  const int lane = lane_id();
  const int len = rb - ra;
  const int mfz = M.mq & 0xFF;
  const int qe = (M.mq >> 16) - 2048;
  unsigned long long* hist64 = (unsigned long long*)hist;
  if (MODE == 2) {
    for (int j = lane; j < 2 * len; j += 64) hist64[j] = 0ULL;
  } else {
    for (int j = lane; j < len; j += 64) hist[j] = 0u;
  }
  wave_sync();
  if (B.n > 0) {
#pragma unroll
    for (int u = 0; u < J_PU; ++u) {
      const float pmu = pm0[u];
      if (!(pmu > 0.f)) continue;   // past the peak count (kept peak masses are positive: fmin > proton)
      const unsigned long long inc = MODE == 2 ? 1ULL | ((unsigned long long)iq_quant(pi0[u], qe) << 24) : 0ULL;
      for (int z = 1; z < mfz; ++z) k1j_lookup<MODE>(P, B, pmu, inc, z, ra, rb, hist);
    }
    for (int pc = 64 * J_PU; pc < M.np; pc += 256) {   // spectra beyond 64 * J_PU peaks
      float pm[4], pi[4];
      k1j_load_peaks<MODE, 4>(A, M, pc, pm, pi);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (!(pm[u] > 0.f)) continue;
        const unsigned long long inc = MODE == 2 ? 1ULL | ((unsigned long long)iq_quant(pi[u], qe) << 24) : 0ULL;
        for (int z = 1; z < mfz; ++z) k1j_lookup<MODE>(P, B, pm[u], inc, z, ra, rb, hist);
      }
    }
  }
  wave_sync();
}

// Slot words of the histogram rows j = u*64 + lane (u < 8, j < len): counts (mode 1) or the packed feature word
// (mode 2; a count >= 2^20 is clamped and flagged)
template <int MODE, class WordT>
__device__ __forceinline__ void k1j_hist_words(const uint32_t* hist, int len, WordT (&v)[8], bool& big) {
  const int lane = lane_id();
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int j = u * 64 + lane;
    v[u] = 0;
    if (j < len) {
      if (MODE == 2) {
        const unsigned long long* h64 = (const unsigned long long*)hist;
        uint32_t c;
        const uint64_t pl = feat_payload(h64[2 * j], h64[2 * j + 1], c);
        big |= c >= (1u << 20);
        v[u] = ((unsigned long long)(c < (1u << 20) ? c : (1u << 20) - 1u) << 43) | pl;
      } else {
        v[u] = hist[j];
      }
    }
  }
}

template <int MODE, class WordT>
__device__ __forceinline__ void k1j_store_words(const JArgs& A, int64_t sb, int len, const WordT (&v)[8]) {
  const int lane = lane_id();
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int j = u * 64 + lane;
    if (j < len) {
      if (MODE == 2)
        ((unsigned long long*)A.counts)[sb + j] = v[u];
      else
        ((uint32_t*)A.counts)[sb + j] = (uint32_t)v[u];
    }
  }
}

// Workgroup per block. The block's keys / rows and the first J_PCH pair records are loaded together (all of a
// thread's loads in flight at once), the LDS directory is built, then the waves take pairs (the next pair's
// first peaks are loaded before the current pair is processed); a pair whose row range exceeds the histogram is
// counted in several row passes.
template <int MODE>
__global__ __launch_bounds__(J_THREADS) void k1j_count_kernel(DevP P, JArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int b = blockIdx.x;
  const uint32_t pb = A.blk_pairs[b];
  uint32_t pe = A.blk_pairs[b + 1];
  if ((int64_t)pe > A.pair_cap) pe = (uint32_t)A.pair_cap;   // overflow is flagged by vh_k1j_pairs
  if (pb >= pe) return;                                       // uniform over the workgroup
  uint32_t* keys = (uint32_t*)smem;
  uint16_t* rows = (uint16_t*)(keys + A.maxp);
  uint16_t* table = rows + A.maxp;
  PairMeta* meta = (PairMeta*)(table + A.tcells + 8);
  unsigned char* hbase = (unsigned char*)(meta + J_PCH);
  const int tid = threadIdx.x;
  const uint32_t e0 = A.off[b];
  const int n = (int)(A.off[b + 1] - e0);
  const int rb0 = A.row0[b];
  PairMeta m0;
  const int cnt0 = (int)(pe - pb < (uint32_t)J_PCH ? pe - pb : (uint32_t)J_PCH);
  if (tid < cnt0) m0 = A.pairs[pb + tid];
  for (int i0 = 0; i0 < n; i0 += 8 * J_THREADS) {
    uint32_t kv[8];
    uint16_t rv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * J_THREADS + tid;
      kv[u] = i < n ? A.key[e0 + i] : 0u;
      rv[u] = i < n ? A.row[e0 + i] : (uint16_t)0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * J_THREADS + tid;
      if (i < n) {
        keys[i] = kv[u];
        rows[i] = rv[u];
      }
    }
  }
  if (tid < cnt0) meta[tid] = m0;
  if (tid < 4) keys[n + tid] = 0xFFFFFFFFu;   // scan sentinels (maxp leaves room for them)
  __syncthreads();
  const uint32_t kmin = n ? keys[0] : 0u, kmax = n ? keys[n - 1] : 0u;
  int tsh = 0;
  while (((kmax - kmin) >> tsh) >= (uint32_t)A.tcells) ++tsh;
  // table[t] = first posting whose cell (key - kmin) >> tsh is >= t, t = 0..tcells
  for (int i = tid; i < n; i += J_THREADS) {
    const int c = (int)((keys[i] - kmin) >> tsh);
    const int cp = i ? (int)((keys[i - 1] - kmin) >> tsh) : -1;
    for (int t = cp + 1; t <= c; ++t) table[t] = (uint16_t)i;
  }
  {
    const int cl = n ? (int)((kmax - kmin) >> tsh) : -1;
    for (int t = tid; t <= A.tcells; t += J_THREADS)
      if (t > cl) table[t] = (uint16_t)n;
  }
  const int w = tid >> 6;
  const int rb1 = A.row0[b + 1];
  uint32_t* hist = (uint32_t*)(hbase + (size_t)w * A.hist_rows * (MODE == 2 ? 16 : 4));
  JBlock B;
  B.keys = keys;
  B.rows = rows;
  B.table = table;
  B.n = n;
  B.tsh = tsh;
  B.rb0 = rb0;
  B.kmin = kmin;
  B.kmax = kmax;
  for (uint32_t c0 = pb; c0 < pe; c0 += J_PCH) {
    const int cnt = (int)(pe - c0 < (uint32_t)J_PCH ? pe - c0 : (uint32_t)J_PCH);
    if (c0 != pb) {
      if (tid < cnt) meta[tid] = A.pairs[c0 + tid];
    }
    __syncthreads();
    // the next pair's peaks are loaded while the current pair is counted; a pair's slot words are stored at the
    // start of the next iteration, after its peak wait, so no store stands between a prefetch and its use
    float pmn[J_PU], pin[J_PU];
    if (w < cnt) k1j_load_peaks<MODE, J_PU>(A, meta[w], 0, pmn, pin);
    typedef typename SlotType<MODE>::T WordT;
    WordT pend[8];
    int64_t pend_sb = 0;
    int pend_len = 0;
    bool big = false;
    for (int q = w; q < cnt; q += J_WAVES) {
      const PairMeta M = meta[q];
      float pm0[J_PU], pi0[J_PU];
#pragma unroll
      for (int u = 0; u < J_PU; ++u) {
        pm0[u] = pmn[u];
        pi0[u] = pin[u];
      }
      if (pend_len) k1j_store_words<MODE>(A, pend_sb, pend_len, pend);
      pend_len = 0;
      if (q + J_WAVES < cnt) k1j_load_peaks<MODE, J_PU>(A, meta[q + J_WAVES], 0, pmn, pin);
      const int r0 = M.lo > rb0 ? M.lo : rb0, r1 = M.hi < rb1 ? M.hi : rb1;
      for (int ra = r0; ra < r1; ra += A.hist_rows) {
        const int rb = ra + A.hist_rows < r1 ? ra + A.hist_rows : r1;
        k1j_pair_rows<MODE>(P, A, M, ra, rb, pm0, pi0, B, hist);
        WordT v[8];
        k1j_hist_words<MODE>(hist, rb - ra, v, big);
        wave_sync();
        if (rb < r1) {
          k1j_store_words<MODE>(A, M.slot + (ra - M.lo), rb - ra, v);   // an earlier row pass
        } else {
#pragma unroll
          for (int u = 0; u < 8; ++u) pend[u] = v[u];
          pend_sb = M.slot + (ra - M.lo);
          pend_len = rb - ra;
        }
      }
    }
    if (pend_len) k1j_store_words<MODE>(A, pend_sb, pend_len, pend);
    if (MODE == 2 && ballot(big) && lane_id() == 0) atomicOr(A.status, 16);
    __syncthreads();
  }
}

// feature mode: per spectrum the intensity scale shift of its kept peaks (iq_shift of the largest intensity)
__global__ __launch_bounds__(256) void k1j_iq_shift_kernel(int32_t n_spec, const int64_t* __restrict__ pk_off,
                                                           const float* __restrict__ pk_int, int32_t* s_qe) {
  const int s = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (s >= n_spec) return;
  const int lane = lane_id();
  const int64_t p0 = pk_off[s];
  const int np = (int)(pk_off[s + 1] - p0);
  float mx = 0.f;
  for (int p = lane; p < np; p += 64) mx = fmaxf(mx, pk_int[p0 + p]);
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  if (lane == 0) s_qe[s] = iq_shift(mx);
}

// ------------------------------------------------------------------------------------------------ select
struct JSel {
  const int64_t* hyp_off;
  const uint8_t* h_mode;
  const int64_t* h_lo;
  const int64_t* h_hi;
  const int64_t* h_slot;
  int64_t slot_cap;
  const void* counts;
};

// slots of hypothesis rows [len - 1 - (b0 + u*64 + lane)] for u < JU (0 past the window)
template <class SlotT, int JU>
__device__ __forceinline__ void k1j_load_batch(const JSel& J, int64_t slot, int len, int b0, bool ok,
                                               SlotT (&v)[JU]) {
  const SlotT* c = (const SlotT*)J.counts;
  const int lane = lane_id();
#pragma unroll
  for (int u = 0; u < JU; ++u) {
    // every lane loads (a clamped, in-window index: no exec-masked load), the value past the window is zeroed
    const int jd = b0 + u * 64 + lane;
    const int jr = len - 1 - jd;
    const SlotT x = c[slot + (jr > 0 ? jr : 0)];
    v[u] = (ok && jd < len) ? x : (SlotT)0;
  }
}

template <int MODE>
__device__ void k1j_spectrum(const DevP& P, const DevIdx& X, int s, unsigned char* wb, const JSel& J,
                             const K1Out& O) {
  const int lane = lane_id();
  uint64_t* buf = (uint64_t*)wb;
  uint64_t* pay = MODE == 2 ? buf + O.capb : nullptr;
  uint64_t* ck = buf + (MODE == 2 ? 2 : 1) * O.capb;
  const int64_t h0 = J.hyp_off[s];
  const int nh = (int)(J.hyp_off[s + 1] - h0);
  if (nh == 0 || nh > 64) {
    if (lane == 0) O.n_cand[s] = 0;
    if (MODE == 2 && lane == 0) O.s_maxcnt[s] = 0;
    if (nh > 64 && lane == 0) atomicOr(O.status, 1);
    return;
  }
  int my_lo = 0, my_hi = 0, my_open = 0;
  int64_t my_slot = 0;
  if (lane < nh) {
    my_lo = (int)J.h_lo[h0 + lane];
    my_hi = (int)J.h_hi[h0 + lane];
    my_open = (J.h_mode && J.h_mode[h0 + lane] == VH_MODE_OPEN) ? 1 : 0;
    my_slot = J.h_slot[h0 + lane];
  }
  int my_rank = 0;   // windows by descending start (the scan order of vh_k1.hip; the result does not depend on it)
  for (int h2 = 0; h2 < nh; ++h2) {
    const int lo2 = __shfl(my_lo, h2);
    my_rank += (lo2 > my_lo || (lo2 == my_lo && h2 < lane)) ? 1 : 0;
  }
  using SlotT = typename SlotType<MODE>::T;
  constexpr int JU = 4;
  const int K = P.topk;
  int nbuf = 0;
  uint64_t thresh = 0;
  bool clamped = false;
  for (int it = 0; it < nh; ++it) {
    const int hc = (int)__builtin_ctzll(ballot(lane < nh && my_rank == it));
    const int lo = __shfl(my_lo, hc), hi = __shfl(my_hi, hc);
    const int thr = __shfl(my_open, hc) ? P.k1_threshold_open : P.k1_threshold;
    const int len = hi > lo ? hi - lo : 0;
    const int64_t slot = __shfl(my_slot, hc);
    const bool ok = slot + len <= J.slot_cap;   // overflow is flagged by vh_k1j_pairs
    uint32_t sc = 0;
    unsigned long long mt = 0;
    // rows descending, JU chunks of 64 slots per batch; the next batch's loads are issued before the current
    // batch is processed
    if (len == 0 || !ok) {   // uniform: nothing to read (an overflowing window is flagged by vh_k1j_pairs)
      if (lane == 0) {
        O.hyp_scored[h0 + hc] = 0;
        O.hyp_matched[h0 + hc] = 0;
      }
      continue;
    }
    SlotT cur[JU], nxt[JU];
    k1j_load_batch(J, slot, len, 0, ok, cur);
    for (int b0 = 0; b0 < len; b0 += 64 * JU) {
      if (b0 + 64 * JU < len) k1j_load_batch(J, slot, len, b0 + 64 * JU, ok, nxt);
#pragma unroll
      for (int u = 0; u < JU; ++u) {
        const int jd = b0 + u * 64 + lane;
        const int j = len - 1 - jd;
        const bool act = ok && jd < len;
        uint32_t c;
        uint64_t pl = 0;
        if (MODE == 2) {
          c = (uint32_t)(cur[u] >> 43);
          pl = cur[u] & ((1ULL << 43) - 1ULL);
        } else {
          c = (uint32_t)cur[u];
        }
        sc += c > 0u;
        mt += c;
        const bool flag = act && (int)c >= thr;
        topk_offer<MODE>(flag, c, lo + j, hc, pl, buf, pay, O.capb, K, nbuf, thresh, clamped);
      }
#pragma unroll
      for (int u = 0; u < JU; ++u) cur[u] = nxt[u];
    }
    sc = wave_sum(sc);
    mt = wave_sum(mt);
    if (lane == 0) {
      O.hyp_scored[h0 + hc] = sc;
      O.hyp_matched[h0 + hc] = mt;
    }
  }
  k1_finish<MODE>(P, X, s, h0, buf, pay, nbuf, thresh, clamped, ck, O);
}

template <int MODE, bool PERSIST>
__global__ __launch_bounds__(256) void k1j_select_kernel(DevP P, DevIdx X, int n_spec,
                                                         const int32_t* __restrict__ order, JSel J, K1Out O) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int w = threadIdx.x >> 6, lane = lane_id();
  unsigned char* wb = smem + (size_t)w * O.wave_bytes;
  if (!PERSIST) {
    const int wv = xcd_band_position(P.xcd_group) * (blockDim.x >> 6) + w;
    if (wv >= n_spec) return;
    k1j_spectrum<MODE>(P, X, order ? order[wv] : wv, wb, J, O);
    return;
  }
  for (;;) {
    int wv = 0;
    if (lane == 0) wv = (int)atomicAdd(O.queue, 1u);
    wv = __shfl(wv, 0);
    if (wv >= n_spec) return;   // every wave reaches this exit once the counter passes n_spec
    k1j_spectrum<MODE>(P, X, order ? order[wv] : wv, wb, J, O);
    wave_sync();
  }
}

__global__ void k1j_zero_u32_kernel(uint32_t* p) { *p = 0u; }

// ------------------------------------------------------------------------------------------------ launchers
static int j_pow2_at_least(int x) {
  int n = 1;
  while (n < x) n <<= 1;
  return n;
}

static size_t align256j(size_t x) { return (x + 255) & ~(size_t)255; }

// LDS of a count workgroup: keys (maxp u32) + rows (maxp u16) + directory (tcells + 8 u16) + per-wave histograms
// LDS of a count workgroup: keys (maxp u32) + rows (maxp u16) + directory (tcells + 8 u16) + pair metadata +
// per-wave histograms of hist_rows slots (u32; mode 2: two u64), at most 512 (mode 2: 256) rows per pass
static size_t count_lds(const VhBlockView* bv, int mode, int& maxp, int& tcells, int& hist_rows) {
  maxp = (bv->max_postings + 4 + 7) & ~7;   // + 4 sentinel keys
  tcells = 256;   // about one posting per directory cell, at most 4096 cells
  while (tcells < 4096 && tcells * 2 <= bv->max_postings) tcells *= 2;
  // one pass holds at most 512 rows (its slot words travel in 8 registers per lane); feature mode 256 (LDS)
  const int cap = mode == 2 ? 256 : 512;
  hist_rows = bv->max_rows < cap ? (bv->max_rows + 3) & ~3 : cap;
  return (size_t)maxp * 6 + (size_t)(tcells + 8) * 2 + (size_t)J_PCH * sizeof(PairMeta) +
         (size_t)J_WAVES * hist_rows * (mode == 2 ? 16 : 4);
}

static int select_geometry(const VhParams* p, int mode, K1Out& O, int& waves) {
  const int K = p->topk > 0 ? p->topk : 1;
  O.capb = j_pow2_at_least(K + 64 > 2 * K ? K + 64 : 2 * K);
  O.clos_n = j_pow2_at_least((2 * K) > 64 ? 2 * K : 64);
  O.hist_bytes = 0;
  O.wave_bytes = (int)(((size_t)O.capb * (mode == 2 ? 16 : 8) + (size_t)O.clos_n * 8 + 15) & ~(size_t)15);
  // one wave (spectrum) per workgroup: a workgroup keeps its CU slot until its slowest wave ends, and the
  // spectra's select times vary, so single-wave workgroups free the slot as each spectrum finishes (measured:
  // 0.352 -> 0.315 ms on the proteome batch at K = 50, 0.986 -> 0.775 ms on the phospho batch at K = 200,
  // against four waves per workgroup)
  waves = 1;
  return O.wave_bytes > 65536 ? VH_E_LIMIT : 0;
}

static unsigned select_grid(const void* kern, int waves, size_t lds, int32_t n_spec, const VhParams* p,
                            bool persistent) {
  const unsigned blocks = (unsigned)((n_spec + waves - 1) / waves);
  if (!persistent) return xcd_band_grid(blocks, p->xcd_group);
  int dev = 0, cus = 256, per = 1;
  if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, 64 * waves, lds) != hipSuccess || per < 1) per = 1;
  const unsigned g = (unsigned)cus * (unsigned)per;
  return g < blocks ? g : blocks;
}

template <int MODE>
static int select_launch(const VhParams* p, const VhIndexView* ix, int32_t n_spec, const int32_t* order,
                         const JSel& J, K1Out& O, hipStream_t st) {
  int waves;
  int rc = select_geometry(p, MODE, O, waves);
  if (rc) return rc;
  const size_t lds = (size_t)O.wave_bytes * waves;
  if (O.queue) {
    const void* kern = (const void*)k1j_select_kernel<MODE, true>;
    unsigned g = select_grid(kern, waves, lds, n_spec, p, true);
    k1j_zero_u32_kernel<<<1, 1, 0, st>>>(O.queue);
    k1j_select_kernel<MODE, true><<<g, 64 * waves, lds, st>>>(vh_make_devp(p), vh_make_devidx(ix), n_spec, order,
                                                              J, O);
  } else {
    const void* kern = (const void*)k1j_select_kernel<MODE, false>;
    unsigned g = select_grid(kern, waves, lds, n_spec, p, false);
    k1j_select_kernel<MODE, false><<<g, 64 * waves, lds, st>>>(vh_make_devp(p), vh_make_devidx(ix), n_spec,
                                                               order, J, O);
  }
  VH_CHECK_LAUNCH();
  return 0;
}

static int block_view_ok(const VhBlockView* bv) {
  if (!bv || bv->n_blocks < 1 || !bv->row0 || !bv->off) return VH_E_ARG;
  if (bv->max_postings < 0 || bv->max_postings > 65535 || bv->max_rows < 1 || bv->max_rows > 32768)
    return VH_E_LIMIT;
  if (bv->max_postings && (!bv->key || !bv->row)) return VH_E_ARG;
  if (bv->n_blocks >= (1LL << 31)) return VH_E_LIMIT;
  return 0;
}

extern "C" {

size_t vh_k1j_workspace_bytes(int64_t n_blocks, int64_t max_hyp) {
  const size_t a = vh_scan_ws_i64(max_hyp > 0 ? max_hyp : 1), b = vh_scan_ws_u32(n_blocks + 1);
  return align256j((size_t)(n_blocks > 0 ? n_blocks : 1) * 4) + align256j(a > b ? a : b);
}

int vh_k1j_pairs(const VhBlockView* bv, int32_t n_spec, const int64_t* hyp_off, const int32_t* h_spec,
                 const int64_t* h_lo, const int64_t* h_hi, const int8_t* h_maxfz, const int64_t* pk_off,
                 const int32_t* s_qe, int64_t max_hyp, int64_t* h_slot, int64_t slot_cap, uint32_t* blk_pairs,
                 void* pairs, int64_t pair_cap, int32_t* status, void* ws, size_t ws_bytes, void* stream) {
  int rc = block_view_ok(bv);
  if (rc) return rc;
  if (n_spec < 0 || max_hyp < 1 || !hyp_off || !h_spec || !h_lo || !h_hi || !h_maxfz || !pk_off || !h_slot ||
      !blk_pairs || !status || !ws || slot_cap < 0 || pair_cap < 0 || (pair_cap && !pairs))
    return VH_E_ARG;
  if (ws_bytes < vh_k1j_workspace_bytes(bv->n_blocks, max_hyp)) return VH_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const int64_t nb = bv->n_blocks;
  uint32_t* cursor = (uint32_t*)ws;
  void* sws = (char*)ws + align256j((size_t)nb * 4);
  const size_t sws_bytes = ws_bytes - align256j((size_t)nb * 4);
  const unsigned gz = (unsigned)((nb + 1 + 255) / 256 < 4096 ? (nb + 1 + 255) / 256 : 4096);
  pj_zero_kernel<<<gz, 256, 0, st>>>(blk_pairs, nb + 1);
  pj_zero_kernel<<<gz, 256, 0, st>>>(cursor, nb);
  const unsigned gh = (unsigned)((max_hyp + 255) / 256);
  pj_count_kernel<<<gh, 256, 0, st>>>(bv->row0, nb, n_spec, hyp_off, h_lo, h_hi, max_hyp, h_slot, blk_pairs);
  VH_CHECK_LAUNCH();
  rc = vh_scan_inclusive_i64(h_slot + 1, max_hyp, sws, sws_bytes, st);
  if (rc) return rc;
  rc = vh_scan_exclusive_u32(blk_pairs, nb + 1, sws, sws_bytes, st);
  if (rc) return rc;
  pj_scatter_kernel<<<gh, 256, 0, st>>>(bv->row0, nb, n_spec, hyp_off, h_spec, h_lo, h_hi, h_maxfz, pk_off, s_qe,
                                        max_hyp, h_slot, slot_cap, blk_pairs, cursor, (PairMeta*)pairs, pair_cap,
                                        status);
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_k1j_count(const VhParams* p, const VhBlockView* bv, int32_t mode, const uint32_t* blk_pairs,
                 const void* pairs, int64_t pair_cap, const float* pk_mass, const float* pk_int, void* counts,
                 int32_t* status, void* stream) {
  int rc = block_view_ok(bv);
  if (rc) return rc;
  if (!p || (mode != 1 && mode != 2) || !blk_pairs || !pk_mass || (mode == 2 && !pk_int) || pair_cap < 0 ||
      !status || (pair_cap && (!pairs || !counts)))
    return VH_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  JArgs A;
  A.row0 = bv->row0;
  A.off = bv->off;
  A.key = bv->key;
  A.row = bv->row;
  const size_t lds = count_lds(bv, mode, A.maxp, A.tcells, A.hist_rows);
  if (lds > 160 * 1024) return VH_E_LIMIT;
  A.blk_pairs = blk_pairs;
  A.pairs = (const PairMeta*)pairs;
  A.pair_cap = pair_cap;
  A.pk_mass = pk_mass;
  A.pk_int = pk_int;
  A.counts = counts;
  A.status = status;
  const DevP P = vh_make_devp(p);
  const unsigned g = (unsigned)bv->n_blocks;
  if (mode == 1) {
    if (lds > 65536) hipFuncSetAttribute((const void*)k1j_count_kernel<1>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    k1j_count_kernel<1><<<g, J_THREADS, lds, st>>>(P, A);
  } else {
    if (lds > 65536) hipFuncSetAttribute((const void*)k1j_count_kernel<2>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    k1j_count_kernel<2><<<g, J_THREADS, lds, st>>>(P, A);
  }
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_k1j_iq_shift(int32_t n_spec, const int64_t* pk_off, const float* pk_int, int32_t* s_qe, void* stream) {
  if (n_spec < 0 || (n_spec && (!pk_off || !pk_int || !s_qe))) return VH_E_ARG;
  if (n_spec == 0) return 0;
  const unsigned g = (unsigned)(((int64_t)n_spec * 64 + 255) / 256);
  k1j_iq_shift_kernel<<<g, 256, 0, (hipStream_t)stream>>>(n_spec, pk_off, pk_int, s_qe);
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_k1j_topk(const VhParams* p, const VhIndexView* ix, int32_t n_spec, const int32_t* order,
                const int64_t* hyp_off, const uint8_t* h_mode, const int64_t* h_lo, const int64_t* h_hi,
                const int64_t* h_slot, int64_t slot_cap, const uint32_t* counts, int32_t cand_stride,
                int32_t* n_cand, int32_t* c_row, int32_t* c_hyp, int32_t* c_cnt, uint32_t* hyp_scored,
                unsigned long long* hyp_matched, int32_t* status, uint32_t* queue, void* stream) {
  if (!p || !ix || n_spec < 0 || !hyp_off || !h_lo || !h_hi || !h_slot || slot_cap < 0 || (slot_cap && !counts) ||
      !n_cand || !c_row || !c_hyp || !c_cnt || !hyp_scored || !hyp_matched || !status)
    return VH_E_ARG;
  if (p->topk < 1 || p->topk > 1024 || (p->pair_closure && !ix->partner)) return VH_E_ARG;
  if (ix->n_rows >= (1LL << 28)) return VH_E_LIMIT;
  if (cand_stride < (p->pair_closure ? 2 * p->topk : p->topk)) return VH_E_CAPACITY;
  if (n_spec == 0) return 0;
  JSel J = {hyp_off, h_mode, h_lo, h_hi, h_slot, slot_cap, counts};
  K1Out O = {};
  O.cand_stride = cand_stride;
  O.n_cand = n_cand;
  O.c_row = c_row;
  O.c_hyp = c_hyp;
  O.c_cnt = c_cnt;
  O.hyp_scored = hyp_scored;
  O.hyp_matched = hyp_matched;
  O.status = status;
  O.queue = queue;
  return select_launch<1>(p, ix, n_spec, order, J, O, (hipStream_t)stream);
}

int vh_k1j_features(const VhParams* p, const VhIndexView* ix, int32_t n_spec, const int32_t* order,
                    const int64_t* hyp_off, const uint8_t* h_mode, const int64_t* h_lo, const int64_t* h_hi,
                    const int64_t* h_slot, int64_t slot_cap, const unsigned long long* counts, int32_t cand_stride,
                    int32_t* n_cand, int32_t* c_row, int32_t* c_hyp, int32_t* c_cnt, int16_t* c_nb,
                    uint32_t* c_iq, int32_t* s_maxcnt, uint32_t* hyp_scored, unsigned long long* hyp_matched,
                    int32_t* status, uint32_t* queue, void* stream) {
  if (!p || !ix || n_spec < 0 || !hyp_off || !h_lo || !h_hi || !h_slot || slot_cap < 0 || (slot_cap && !counts) ||
      !n_cand || !c_row || !c_hyp || !c_cnt || !c_nb || !c_iq || !s_maxcnt || !hyp_scored || !hyp_matched ||
      !status)
    return VH_E_ARG;
  if (p->topk < 1 || p->topk > 1024) return VH_E_ARG;
  if (ix->n_rows >= (1LL << 28)) return VH_E_LIMIT;
  if (cand_stride < p->topk) return VH_E_CAPACITY;
  if (n_spec == 0) return 0;
  JSel J = {hyp_off, h_mode, h_lo, h_hi, h_slot, slot_cap, counts};
  K1Out O = {};
  O.cand_stride = cand_stride;
  O.n_cand = n_cand;
  O.c_row = c_row;
  O.c_hyp = c_hyp;
  O.c_cnt = c_cnt;
  O.c_nb = c_nb;
  O.c_iq = c_iq;
  O.s_maxcnt = s_maxcnt;
  O.hyp_scored = hyp_scored;
  O.hyp_matched = hyp_matched;
  O.status = status;
  O.queue = queue;
  return select_launch<2>(p, ix, n_spec, order, J, O, (hipStream_t)stream);
}

}  // extern "C"
