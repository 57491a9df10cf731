// K1 candidate generation on the fragment index, fused with the per-spectrum top-K and pair closure.
//
// Semantics (the reference's k1_kernel, vulcan/gpu/kernels.cu:112-190, and select_topk, engine.py:284-312):
//   count(h, r) = #(indexed fragment f of row r, peak p, fragment charge z < maxfz(h)) with
//                 tol_bound(mass_p*z, lo) <= f <= tol_bound(mass_p*z, hi), for r in h's precursor window;
//   appended if count >= k1_threshold; scored[h] = #(r: count > 0); matched[h] = sum of counts.
//   Top-K per spectrum over its appended entries by (count clamped at 32767 desc, row desc; identical keys from
//   two hypotheses: lower hypothesis first), + partner rows (count -1), dedupe by row (lowest hypothesis,
//   highest count), row order.
// Design: one wave64 per spectrum. All (hypothesis, row) pairs of the spectrum's windows get a slot in an LDS
// u32 histogram (tiled when they exceed `tile` slots). A tile's segments (hypothesis, row range) are grouped
// by index bucket (ascending). Each lane takes peaks; for each (peak, z) the exact f32 key range [bits(lo bound),
// bits(hi bound)] is looked up per bin it touches and per bucket of the tile: the bin-major directory holds the
// entries of consecutive buckets side by side (one line for a group of four), each bucket's postings in that bin
// are key-sorted and scanned until the key passes the high bound, and postings whose row falls in a segment of
// that bucket increment their slot with an LDS atomic. The scan of the histogram feeds a
// wave-private top-K buffer (ballot-compacted inserts, pruned by the current K-th key, compacted by a bitwise
// K-th-key search when full). Closure and dedupe follow one bitonic sort. Nothing leaves the wave but the final candidate list.
//
// Feature mode (MODE 2, the learned prefilter's input): each slot has two u64 LDS counters, one per ion type
// (the posting's ion bit), count in the low 24 bits and the sum of the matching peaks' quantised intensities in
// the high 40 bits (one u64 atomic per match). A winner's per-ion counts and intensity pseudo-logs travel in a
// payload array beside the top-K key buffer (moved with the keys by the compaction); the keys themselves
// are those of MODE 1, so the top-K set and order are exactly MODE 1's; no closure.
#include "vh_topk.h"

#define K1_MAXSEG 32
#define K1_MAXBK 32
#define K1_PREF 4   // postings per bucket run loaded in the first batch

struct Seg {   // one (hypothesis, row range) piece of a tile: slots [base, base + r1 - r0)
  int r0, r1, base, hl_mfz;   // hl_mfz = local hypothesis | maxfz << 8 | open hypothesis << 16
};

struct BkEnt {   // a bucket touched by a tile: its segments (bit mask), their row hull and max maxfz
  int bk, zmax, rlo, rhi;
  uint32_t mask;
  int pad0, pad1, pad2;
};

// one spectrum s, processed by the calling wave in its LDS area wb
template <int MODE>
__device__ void k1_spectrum(const DevP& P, const DevIdx& X, int s, unsigned char* wb,
                            const int64_t* __restrict__ pk_off, const float* __restrict__ pk_mass,
                            const int64_t* __restrict__ hyp_off, const int8_t* __restrict__ h_maxfz,
                            const uint8_t* __restrict__ h_mode, const int64_t* __restrict__ h_lo,
                            const int64_t* __restrict__ h_hi, const K1Out& O) {
  const int lane = lane_id();
  uint32_t* hist = (uint32_t*)wb;
  unsigned long long* hist64 = (unsigned long long*)wb;   // feature mode
  Seg* segs = (Seg*)(wb + O.hist_bytes);
  BkEnt* bks = (BkEnt*)(segs + K1_MAXSEG);
  uint32_t* hsc = (uint32_t*)(bks + K1_MAXBK);
  unsigned long long* hmt = (unsigned long long*)(hsc + 64);
  uint64_t* buf = (uint64_t*)(hmt + 64);
  uint64_t* pay = MODE == 2 ? buf + O.capb : nullptr;   // feature mode payloads beside the keys

  const int64_t h0 = hyp_off[s];
  const int nh = (int)(hyp_off[s + 1] - h0);
  if (nh == 0 || nh > 64) {
    if (MODE >= 1 && lane == 0) O.n_cand[s] = 0;
    if (MODE == 2 && lane == 0) O.s_maxcnt[s] = 0;
    if (nh > 64 && lane == 0) atomicOr(O.status, 1);
    return;
  }
  int my_lo = 0, my_hi = 0, my_mfz = 0;   // my_mfz also carries the open-hypothesis flag at bit 8
  if (lane < nh) {
    my_lo = (int)h_lo[h0 + lane];
    my_hi = (int)h_hi[h0 + lane];
    my_mfz = h_maxfz[h0 + lane] | ((h_mode && h_mode[h0 + lane] == VH_MODE_OPEN) ? 256 : 0);
  }
  hsc[lane] = 0;
  hmt[lane] = 0;
  const int64_t p0 = pk_off[s];
  const int np = (int)(pk_off[s + 1] - p0);
  const int T = P.tile;
  const int S = X.bin_shift, LB = X.bucket_log2;
  const int base = (int)X.bin_base, n_bins = X.n_bins;
  const int B = 1 << LB;
  const int K = P.topk;
  int qe = -1000;
  if (MODE == 2) {
    float mx = 0.f;
    for (int p = lane; p < np; p += 64) mx = fmaxf(mx, O.pk_int[p0 + p]);
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    qe = iq_shift(mx);
  }
  int nbuf = 0;
  uint64_t thresh = 0;
  bool clamped = false;

  // Windows are scanned from the top row down (hypotheses by descending window start, rows descending): the
  // top-K key breaks count ties by the larger row, so once the buffer is full an equal-count row met later has
  // the smaller key and is pruned by the threshold instead of forcing another compaction. The selected set does
  // not depend on the scan order.
  int my_rank = 0;
  for (int h2 = 0; h2 < nh; ++h2) {
    const int lo2 = __shfl(my_lo, h2);
    my_rank += (lo2 > my_lo || (lo2 == my_lo && h2 < lane)) ? 1 : 0;
  }
  int it = 0;
  int hc = (int)__builtin_ctzll(ballot(lane < nh && my_rank == 0));
  int rc = __shfl(my_hi, hc);
  for (;;) {
    // ---- plan one tile of (hypothesis, row-range) segments (wave-uniform); lane j holds bucket entry j
    int used = 0, nseg = 0, zmax = 0, nbk = 0;
    int my_bk = -1, my_bz = 0, my_rlo = 0x7fffffff, my_rhi = 0;
    uint32_t my_mask = 0;
    while (it < nh && used < T && nseg < K1_MAXSEG) {
      const int lo_c = __shfl(my_lo, hc);
      if (rc <= lo_c) {
        ++it;
        if (it < nh) {
          hc = (int)__builtin_ctzll(ballot(lane < nh && my_rank == it));
          rc = __shfl(my_hi, hc);
        }
        continue;
      }
      int take = rc - lo_c < T - used ? rc - lo_c : T - used;
      // bucket entries this piece may add must fit: cap its span to the free entries
      int free_bk = K1_MAXBK - nbk;
      if (free_bk <= 0) break;
      const int cap_rows = (((rc - 1) & (B - 1)) + 1) + ((free_bk - 1) << LB);
      take = take < cap_rows ? take : cap_rows;
      const int mfo = __shfl(my_mfz, hc);
      const int mf = mfo & 0xFF;
      const int r0 = rc - take;
      if (lane == 0) {
        Seg g;
        g.r0 = r0; g.r1 = rc; g.base = used; g.hl_mfz = hc | (mfo << 8);
        segs[nseg] = g;
      }
      for (int bk = r0 >> LB; bk <= (rc - 1) >> LB; ++bk) {
        uint64_t hit = ballot(lane < nbk && my_bk == bk);
        int j = hit ? (int)__builtin_ctzll(hit) : nbk;
        if (!hit) ++nbk;
        if (lane == j) {
          my_bk = bk;
          my_mask |= 1u << nseg;
          my_bz = mf > my_bz ? mf : my_bz;
          my_rlo = r0 < my_rlo ? r0 : my_rlo;
          my_rhi = rc > my_rhi ? rc : my_rhi;
        }
      }
      zmax = mf > zmax ? mf : zmax;
      ++nseg;
      used += take;
      rc = r0;
    }
    if (nseg == 0) break;
    {   // bucket entries in ascending bucket order: groups of consecutive buckets share directory lines
      int rank = 0;
      for (int j = 0; j < nbk; ++j) rank += __shfl(my_bk, j) < my_bk ? 1 : 0;
      if (lane < nbk) {
        BkEnt e;
        e.bk = my_bk; e.zmax = my_bz; e.rlo = my_rlo; e.rhi = my_rhi; e.mask = my_mask;
        e.pad0 = e.pad1 = e.pad2 = 0;
        bks[rank] = e;
      }
    }
    if (MODE == 2) {
      for (int j = lane; j < 2 * used; j += 64) hist64[j] = 0ULL;
    } else {
      for (int j = lane; j < used; j += 64) hist[j] = 0;
    }
    wave_sync();
    // ---- accumulate shared-peak counts
    for (int p = lane; p < np; p += 64) {
      const float pm = pk_mass[p0 + p];
      const unsigned long long inc = MODE == 2 ? 1ULL | ((unsigned long long)iq_quant(O.pk_int[p0 + p], qe) << 24)
                                               : 0ULL;
      for (int z = 1; z < zmax; ++z) {
        float m = pm * (float)z;
        float a = tol_bound(m, P.tol_lo, P.tol_da);
        float b = tol_bound(m, P.tol_hi, P.tol_da);
        if (b < P.fmin || a > P.fmax) continue;
        uint32_t ka = a > 0.f ? fbits(a) : 0u;
        uint32_t kb = fbits(b);
        int bl = (int)(ka >> S) - base;
        int bh = (int)(kb >> S) - base;
        bl = bl < 0 ? 0 : bl;
        bh = bh > n_bins - 1 ? n_bins - 1 : bh;
        for (int bn = bl; bn <= bh; ++bn) {
          const uint32_t khi = (uint32_t)(bn + base) << S;
          const uint32_t* dirb = X.offsets + (size_t)bn * (size_t)X.n_buckets;
          // directory entries of up to 4 buckets (consecutive buckets: one line) are loaded together, then the
          // buckets' key-sorted posting runs of this bin are scanned
          for (int j0 = 0; j0 < nbk; j0 += 4) {
            uint32_t e0[4], e1[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              e0[u] = e1[u] = 0;
              if (j0 + u < nbk && z < bks[j0 + u].zmax) {
                const int bk = bks[j0 + u].bk;
                e0[u] = dirb[bk];
                e1[u] = dirb[bk + 1];
              }
            }
            // the first K1_PREF postings of each bucket's run are loaded in one batch (independent loads), so a
            // typical lookup costs two memory round trips; longer runs continue one posting at a time
            uint32_t pw[4][K1_PREF];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
              for (int t = 0; t < K1_PREF; ++t) pw[u][t] = e0[u] + t < e1[u] ? X.postings[e0[u] + t] : 0u;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              if (e0[u] == e1[u]) continue;
              const BkEnt E = bks[j0 + u];
              const int rb = E.bk << LB;
              // one posting: false once the key passes the high bound (the run is key-sorted)
              auto visit = [&](uint32_t w) -> bool {
                const uint32_t key = khi | (w >> 17);
                if (key > kb) return false;
                const int row = rb + (int)(w & 0xFFFFu);
                if (key < ka || row < E.rlo || row >= E.rhi) return true;
                uint32_t msk = E.mask;
                while (msk) {
                  const int q = __builtin_ctz(msk);
                  msk &= msk - 1u;
                  const Seg g = segs[q];
                  if (z < ((g.hl_mfz >> 8) & 0xFF) && row >= g.r0 && row < g.r1) {
                    if (MODE == 2)
                      atomicAdd(&hist64[2 * (g.base + row - g.r0) + ((w >> 16) & 1u)], inc);
                    else
                      atomicAdd(&hist[g.base + row - g.r0], 1u);
                  }
                }
                return true;
              };
              bool more = true;
#pragma unroll
              for (int t = 0; t < K1_PREF; ++t)
                if (more) more = e0[u] + t < e1[u] && visit(pw[u][t]);
              if (more)
                for (uint32_t e = e0[u] + K1_PREF; e < e1[u]; ++e)
                  if (!visit(X.postings[e])) break;
            }
          }
        }
      }
    }
    wave_sync();
    // ---- histogram scan: per-hypothesis statistics, append or top-K insert
    for (int q = 0; q < nseg; ++q) {
      const Seg g = segs[q];
      const int g_hl = g.hl_mfz & 0xFF;
      const int thr = (g.hl_mfz >> 16) & 1 ? P.k1_threshold_open : P.k1_threshold;
      const int len = g.r1 - g.r0;
      uint32_t sc = 0;
      unsigned long long mt = 0;
      for (int j0 = 0; j0 < len; j0 += 64) {
        const int jd = j0 + lane;          // rows descending: slot j holds row g.r0 + j
        const int j = len - 1 - jd;
        const bool act = jd < len;
        uint32_t c;
        uint64_t pl = 0;
        if (MODE == 2) {
          const unsigned long long cb = act ? hist64[2 * (g.base + j)] : 0ULL;
          const unsigned long long cy = act ? hist64[2 * (g.base + j) + 1] : 0ULL;
          pl = feat_payload(cb, cy, c);
        } else {
          c = act ? hist[g.base + j] : 0u;
        }
        sc += c > 0u;
        mt += c;
        const bool flag = act && (int)c >= thr;
        if (MODE == 0) {
          uint64_t bal = ballot(flag);
          if (bal) {
            unsigned long long bp = 0;
            if (lane == 0) bp = atomicAdd(O.n_out, (unsigned long long)__popcll(bal));
            bp = __shfl(bp, 0);
            if (flag) {
              unsigned long long pos = bp + (unsigned long long)__popcll(bal & lanemask_lt());
              if ((int64_t)pos < O.cap) {
                O.out_hyp[pos] = (int32_t)(h0 + g_hl);
                O.out_row[pos] = g.r0 + j;
                O.out_cnt[pos] = (int32_t)c;
              }
            }
          }
        } else {
          topk_offer<MODE>(flag, c, g.r0 + j, g_hl, pl, buf, pay, O.capb, K, nbuf, thresh, clamped);
        }
      }
      sc = wave_sum(sc);
      mt = wave_sum(mt);
      if (lane == 0) {
        hsc[g_hl] += sc;
        hmt[g_hl] += mt;
      }
    }
    wave_sync();
  }
  if (lane < nh) {
    O.hyp_scored[h0 + lane] = hsc[lane];
    O.hyp_matched[h0 + lane] = hmt[lane];
  }
  if (MODE == 0) return;

  // ---- final top-K, pair closure, dedupe by row (the closure keys reuse the histogram area)
  k1_finish<MODE>(P, X, s, h0, buf, pay, nbuf, thresh, clamped, (uint64_t*)wb, O);
}

// Grid mode: wave wv takes order[wv] (XCD-banded block order). Persistent mode (O.queue set): every wave takes the
// next position of `order` from an atomic counter until the list is exhausted, so a wave that finishes a short
// spectrum immediately starts another instead of idling until its block's slowest wave is done.
template <int MODE, bool PERSIST>
__global__ __launch_bounds__(256) void k1_kernel(DevP P, DevIdx X, int n_spec, const int32_t* __restrict__ order,
                                                 const int64_t* __restrict__ pk_off,
                                                 const float* __restrict__ pk_mass,
                                                 const int64_t* __restrict__ hyp_off,
                                                 const int8_t* __restrict__ h_maxfz,
                                                 const uint8_t* __restrict__ h_mode,
                                                 const int64_t* __restrict__ h_lo,
                                                 const int64_t* __restrict__ h_hi, K1Out O) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int w = threadIdx.x >> 6, lane = lane_id();
  unsigned char* wb = smem + (size_t)w * O.wave_bytes;
  if (!PERSIST) {
    const int wv = xcd_band_position(P.xcd_group) * (blockDim.x >> 6) + w;
    if (wv >= n_spec) return;   // a whole wave leaves; no block-level barrier is used
    k1_spectrum<MODE>(P, X, order ? order[wv] : wv, wb, pk_off, pk_mass, hyp_off, h_maxfz, h_mode, h_lo, h_hi, O);
    return;
  }
  for (;;) {
    int wv = 0;
    if (lane == 0) wv = (int)atomicAdd(O.queue, 1u);
    wv = __shfl(wv, 0);
    if (wv >= n_spec) return;   // every wave reaches this exit once the counter passes n_spec
    k1_spectrum<MODE>(P, X, order ? order[wv] : wv, wb, pk_off, pk_mass, hyp_off, h_maxfz, h_mode, h_lo, h_hi, O);
    wave_sync();
  }
}

// ------------------------------------------------------------------------------------------------ brute force
// k1_brute_kernel (kernels.cu:195-232): one thread per (hypothesis, row) pair, fragments generated on the fly,
// every peak and charge compared with plain f32 bounds.
__global__ __launch_bounds__(256) void k1_brute_kernel(DevP P, DevIdx X, int64_t n, const int32_t* __restrict__ c_hyp,
                                                       const int32_t* __restrict__ c_row,
                                                       const int32_t* __restrict__ h_spec,
                                                       const int8_t* __restrict__ h_maxfz,
                                                       const int64_t* __restrict__ pk_off,
                                                       const float* __restrict__ pk_mass, int32_t* out_cnt) {
  __shared__ float rm[52];
  load_rm(rm, P);
  __syncthreads();
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int h = c_hyp[i], r = c_row[i];
  int s = h_spec[h], maxfz = h_maxfz[h];
  int64_t q0 = pk_off[s], q1 = pk_off[s + 1];
  const uint8_t* sq = X.residues + X.seq_off[r];
  int L = X.length[r];
  uint64_t mk = X.mask[r];
  float nt = (mk >> 63) ? P.nterm : 0.f;
  float b = nt, y = X.mass32[r] - nt;
  int cnt = 0;
  for (int k = 0; k < L - 1; ++k) {
    float rmv = rm[((mk >> k) & 1ULL) * 26 + (sq[k] - 'A')];
    b = b + rmv;
    y = y - rmv;
    for (int t = 0; t < 2; ++t) {
      float f = t == 0 ? b : y;
      bool ok = t == 0 ? (k + 1 > P.min_ion_index) : (L - 1 - k > P.min_ion_index);
      if (!ok || f < P.fmin || f > P.fmax) continue;
      for (int64_t q = q0; q < q1; ++q)
        for (int z = 1; z < maxfz; ++z) {
          float m = pk_mass[q] * (float)z;
          float a = tol_bound(m, P.tol_lo, P.tol_da);
          float bd = tol_bound(m, P.tol_hi, P.tol_da);
          if (f >= a && f <= bd) ++cnt;
        }
    }
  }
  out_cnt[i] = cnt;
}

__global__ void zero_u64_kernel(unsigned long long* p) { *p = 0ULL; }
__global__ void zero_u32_kernel(uint32_t* p) { *p = 0u; }

// grid of a K1 launch: grid mode = one wave per spectrum (XCD-banded); persistent mode = as many blocks as stay
// resident on the device at once (occupancy query), each wave looping over the work counter
static unsigned k1_launch_grid(const void* kern, int waves, size_t lds, int32_t n_spec, const VhParams* p,
                               bool persistent) {
  const unsigned blocks = (unsigned)((n_spec + waves - 1) / waves);
  if (!persistent) return xcd_band_grid(blocks, p->xcd_group);
  int dev = 0, cus = 256, per = 1;
  if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, 64 * waves, lds) != hipSuccess || per < 1) per = 1;
  const unsigned g = (unsigned)cus * (unsigned)per;
  return g < blocks ? g : blocks;
}

// ------------------------------------------------------------------------------------------------ launchers
static int pow2_at_least(int x) {
  int n = 1;
  while (n < x) n <<= 1;
  return n;
}

static int k1_check(const VhParams* p, const VhIndexView* ix) {
  if (!p || !ix || !ix->offsets || (ix->n_postings && !ix->postings)) return VH_E_ARG;
  if (p->tile < 64 || p->tile > 8192 || (p->tile & (p->tile - 1))) return VH_E_ARG;
  if (ix->n_rows >= (1LL << 28)) return VH_E_LIMIT;
  if (ix->bin_shift < 4 || ix->bin_shift > 16 || ix->bucket_log2 < 0 || ix->bucket_log2 > 16) return VH_E_ARG;
  return 0;
}

static void k1_geometry(const VhParams* p, int mode, K1Out& O, int& waves) {
  int K = p->topk > 0 ? p->topk : 1;
  O.capb = mode ? pow2_at_least(K + 64 > 2 * K ? K + 64 : 2 * K) : 0;
  O.clos_n = mode ? pow2_at_least((2 * K) > 64 ? 2 * K : 64) : 0;
  int hb = p->tile * (mode == 2 ? 16 : 4);
  int cb = O.clos_n * 8;
  O.hist_bytes = ((hb > cb ? hb : cb) + 15) & ~15;
  O.wave_bytes = O.hist_bytes + K1_MAXSEG * (int)sizeof(Seg) + K1_MAXBK * (int)sizeof(BkEnt) + 64 * 4 + 64 * 8 +
                 O.capb * (mode == 2 ? 16 : 8);
  waves = 65536 / O.wave_bytes;
  waves = waves < 1 ? 1 : (waves > 4 ? 4 : waves);
}

extern "C" {

int vh_k1(const VhParams* p, const VhIndexView* ix, int32_t n_spec, const int32_t* order, const int64_t* pk_off,
          const float* pk_mass, const int64_t* hyp_off, const int8_t* h_maxfz, const uint8_t* h_mode,
          const int64_t* h_lo, const int64_t* h_hi, int64_t cap,
          unsigned long long* n_out, int32_t* out_hyp, int32_t* out_row, int32_t* out_cnt, uint32_t* hyp_scored,
          unsigned long long* hyp_matched, void* stream) {
  int rc = k1_check(p, ix);
  if (rc) return rc;
  if (n_spec < 0 || !n_out || !hyp_scored || !hyp_matched || cap < 0) return VH_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  zero_u64_kernel<<<1, 1, 0, st>>>(n_out);
  if (n_spec == 0) return 0;
  K1Out O = {};
  O.cap = cap;
  O.n_out = n_out;
  O.out_hyp = out_hyp;
  O.out_row = out_row;
  O.out_cnt = out_cnt;
  O.hyp_scored = hyp_scored;
  O.hyp_matched = hyp_matched;
  int waves;
  k1_geometry(p, 0, O, waves);
  if (O.wave_bytes > 65536) return VH_E_LIMIT;
  unsigned g = xcd_band_grid((n_spec + waves - 1) / waves, p->xcd_group);
  k1_kernel<0, false><<<g, 64 * waves, (size_t)O.wave_bytes * waves, st>>>(
      vh_make_devp(p), vh_make_devidx(ix), n_spec, order, pk_off, pk_mass, hyp_off, h_maxfz, h_mode, h_lo, h_hi, O);
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_k1_topk(const VhParams* p, const VhIndexView* ix, int32_t n_spec, const int32_t* order,
               const int64_t* pk_off, const float* pk_mass, const int64_t* hyp_off, const int8_t* h_maxfz,
               const uint8_t* h_mode, const int64_t* h_lo, const int64_t* h_hi, int32_t cand_stride,
               int32_t* n_cand, int32_t* c_row, int32_t* c_hyp, int32_t* c_cnt, uint32_t* hyp_scored,
               unsigned long long* hyp_matched, int32_t* status, uint32_t* queue, void* stream) {
  int rc = k1_check(p, ix);
  if (rc) return rc;
  if (n_spec < 0 || !n_cand || !c_row || !c_hyp || !c_cnt || !hyp_scored || !hyp_matched || !status)
    return VH_E_ARG;
  if (p->topk < 1 || p->topk > 1024 || (p->pair_closure && !ix->partner)) return VH_E_ARG;
  if (cand_stride < (p->pair_closure ? 2 * p->topk : p->topk)) return VH_E_CAPACITY;
  if (n_spec == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
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
  int waves;
  k1_geometry(p, 1, O, waves);
  if (O.wave_bytes > 65536) return VH_E_LIMIT;
  const size_t lds = (size_t)O.wave_bytes * waves;
  if (queue) {
    unsigned g = k1_launch_grid((const void*)k1_kernel<1, true>, waves, lds, n_spec, p, true);
    zero_u32_kernel<<<1, 1, 0, st>>>(queue);
    k1_kernel<1, true><<<g, 64 * waves, lds, st>>>(vh_make_devp(p), vh_make_devidx(ix), n_spec, order, pk_off,
                                                   pk_mass, hyp_off, h_maxfz, h_mode, h_lo, h_hi, O);
  } else {
    unsigned g = k1_launch_grid((const void*)k1_kernel<1, false>, waves, lds, n_spec, p, false);
    k1_kernel<1, false><<<g, 64 * waves, lds, st>>>(vh_make_devp(p), vh_make_devidx(ix), n_spec, order, pk_off,
                                                    pk_mass, hyp_off, h_maxfz, h_mode, h_lo, h_hi, O);
  }
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_k1_features(const VhParams* p, const VhIndexView* ix, int32_t n_spec, const int32_t* order,
                   const int64_t* pk_off, const float* pk_mass, const float* pk_int, const int64_t* hyp_off,
                   const int8_t* h_maxfz, const uint8_t* h_mode, const int64_t* h_lo, const int64_t* h_hi,
                   int32_t cand_stride, int32_t* n_cand, int32_t* c_row, int32_t* c_hyp, int32_t* c_cnt,
                   int16_t* c_nb, uint32_t* c_iq, int32_t* s_maxcnt, uint32_t* hyp_scored,
                   unsigned long long* hyp_matched, int32_t* status, uint32_t* queue, void* stream) {
  int rc = k1_check(p, ix);
  if (rc) return rc;
  if (n_spec < 0 || !n_cand || !c_row || !c_hyp || !c_cnt || !c_nb || !c_iq || !s_maxcnt || !hyp_scored ||
      !hyp_matched || !status || !pk_int)
    return VH_E_ARG;
  if (p->topk < 1 || p->topk > 1024) return VH_E_ARG;
  if (cand_stride < p->topk) return VH_E_CAPACITY;
  if (n_spec == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  K1Out O = {};
  O.cand_stride = cand_stride;
  O.n_cand = n_cand;
  O.c_row = c_row;
  O.c_hyp = c_hyp;
  O.c_cnt = c_cnt;
  O.hyp_scored = hyp_scored;
  O.hyp_matched = hyp_matched;
  O.status = status;
  O.pk_int = pk_int;
  O.c_nb = c_nb;
  O.c_iq = c_iq;
  O.s_maxcnt = s_maxcnt;
  O.queue = queue;
  int waves;
  k1_geometry(p, 2, O, waves);
  if (O.wave_bytes > 65536) return VH_E_LIMIT;
  const size_t lds = (size_t)O.wave_bytes * waves;
  if (queue) {
    unsigned g = k1_launch_grid((const void*)k1_kernel<2, true>, waves, lds, n_spec, p, true);
    zero_u32_kernel<<<1, 1, 0, st>>>(queue);
    k1_kernel<2, true><<<g, 64 * waves, lds, st>>>(vh_make_devp(p), vh_make_devidx(ix), n_spec, order, pk_off,
                                                   pk_mass, hyp_off, h_maxfz, h_mode, h_lo, h_hi, O);
  } else {
    unsigned g = k1_launch_grid((const void*)k1_kernel<2, false>, waves, lds, n_spec, p, false);
    k1_kernel<2, false><<<g, 64 * waves, lds, st>>>(vh_make_devp(p), vh_make_devidx(ix), n_spec, order, pk_off,
                                                    pk_mass, hyp_off, h_maxfz, h_mode, h_lo, h_hi, O);
  }
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_k1_brute(const VhParams* p, const VhIndexView* ix, int64_t n, const int32_t* c_hyp, const int32_t* c_row,
                const int32_t* h_spec, const int8_t* h_maxfz, const int64_t* pk_off, const float* pk_mass,
                int32_t* out_cnt, void* stream) {
  if (!p || !ix || n < 0) return VH_E_ARG;
  if (n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  unsigned g = (unsigned)((n + 255) / 256);
  k1_brute_kernel<<<g, 256, 0, st>>>(vh_make_devp(p), vh_make_devidx(ix), n, c_hyp, c_row, h_spec, h_maxfz, pk_off,
                                     pk_mass, out_cnt);
  VH_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
