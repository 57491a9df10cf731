/* C ABI of libvulcan_hip.so: the MI355X (gfx950) DDA search hot path.
 *
 * Drop-in boundary for the reference's ctypes surface vulcan/gpu/_lib.py:22-35 (launchers in
 * vulcan/gpu/kernels.cu:379-474). Differences by design:
 *   - no process-global state: the residue tables, N-term delta and tolerance unit that the reference sets
 *     with vk_set_params / vk_set_nterm / vk_set_tol_da (kernels.cu:381-389, cudaMemcpyToSymbol) travel in a
 *     VhParams passed by pointer to every call, so concurrent callers with different settings are safe;
 *   - every buffer is caller-owned device memory (torch tensors in the Python layer); scratch memory comes
 *     from a caller workspace whose size a *_workspace_bytes query returns; nothing here allocates;
 *   - every call is asynchronous on the given stream (a hipStream_t passed as void*) and never synchronizes,
 *     so a whole search step can be captured into a hipGraph;
 *   - return value: 0 = ok, > 0 = hipError_t, < 0 = VH_E_* argument/overflow error (reference: cudaError_t,
 *     _lib.py:74-76 turns nonzero into RuntimeError).
 * u32/u64 data travel as raw bit patterns (the reference passes u32 as int32 tensors, engine.py:38-41).
 */
#ifndef VULCAN_HIP_H
#define VULCAN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VH_ABI_VERSION 19

#define VH_E_ARG (-1)        /* bad argument (null pointer, negative size, unsupported setting) */
#define VH_E_CAPACITY (-2)   /* an output buffer is smaller than the call needs */
#define VH_E_WORKSPACE (-3)  /* workspace smaller than the *_workspace_bytes query */
#define VH_E_LIMIT (-4)      /* a documented size limit is exceeded (rows, bins, hypotheses per spectrum) */

#define VH_MAX_HYP_PER_SPEC 64   /* hypotheses of one spectrum handled by one K1 wave */
#define VH_MAX_OFFSETS 8         /* mass-offset hypotheses per (spectrum, charge) */
#define VH_MODE_CLOSED 0         /* hypothesis kinds (vulcan/ptm/search.py:35) */
#define VH_MODE_OFFSET 1
#define VH_MODE_OPEN 2
#define VH_K2_MAX_PEAKS 512      /* K2 looks at the first 512 mass-sorted peaks (kernels.cu:260, 278) */
#define VH_MAX_PEPTIDE_LEN 63    /* residues per peptidoform (mask bits 0..62; bit 63 = N-term mod) */
#define VH_MLP_N_PARAMS 68353    /* prefilter MLP 8 -> 256 -> 256 -> 1: W1 [256][8], b1, W2 [256][256], b2, w3, b3 */
#define VH_MLP_BLOB_BYTES 281728 /* packed bf16 model (vh_mlp_pack) */

/* Search parameters. Reference: SearchParams (vulcan/gpu/host.py:15-44) plus the constant-memory state of
 * kernels.cu:19-38 (FragParams, c_nterm, c_tol_da). */
typedef struct VhParams {
  float mono[26];             /* Sage f32 residue masses by letter (engine.py:30-32) */
  float stat[26];             /* static mod per residue, f32 (engine.py:74-76) */
  float var[26];              /* variable mod per residue, f32 (engine.py:78-82) */
  float nterm;                /* variable peptide-N-term delta, mask bit 63 (kernels.cu:24-30) */
  int32_t tol_da;             /* fragment unit: 0 ppm, 1 Da (kernels.cu:31-38) */
  float tol_lo, tol_hi;       /* fragment tolerance bounds (fragment_ppm) */
  float fmin, fmax;           /* fragment_min_mz / fragment_max_mz */
  int32_t min_ion_index;      /* index filter of K1 (kernels.cu:67-68) */
  float prec_lo, prec_hi;     /* precursor_ppm */
  int32_t iso_lo, iso_hi;     /* isotope_errors */
  int32_t z_lo, z_hi;         /* precursor_charge (tried when the spectrum has none) */
  int32_t max_fragment_charge;/* 0 = None (fragment charges up to z-1) */
  int32_t max_peaks;          /* K0 top peaks per spectrum */
  int32_t min_peaks;          /* K0 validity */
  int32_t min_matched_peaks;  /* ranking filter (engine.py:368) */
  int32_t topk;               /* K1 top-K per spectrum (engine.py:291-296) */
  int32_t k1_threshold;       /* K1 append threshold (kernels.cu:171) */
  int32_t pair_closure;       /* add partner rows (engine.py:297-301) */
  int32_t tile;               /* K1 LDS histogram slots per wave: 64..8192, power of two */
  int32_t xcd_group;          /* blocks per XCD per scheduling group (0: plain block order), see vh_common.h */
  /* search modes (ModeParams, vulcan/ptm/search.py:38-45): per (spectrum, charge) at isotope 0, one hypothesis
   * per mass offset (window M - offset +- precursor ppm) and one open hypothesis (calc mass in
   * [M - open_hi, M - open_lo]); they follow the closed hypotheses of the spectrum in that order. */
  int32_t n_offsets;
  float offsets[VH_MAX_OFFSETS];
  int32_t open_search;
  float open_lo, open_hi;
  int32_t k1_threshold_open;  /* K1 append threshold of open hypotheses */
  /* K0 "window" peak view (engine.py:182-183, 195-196; host.py:33-35): peak_window_k > 0 keeps the top
   * peak_window_k peaks of each floor(m/z / peak_window_th) window (f32 division) by the global view's selection
   * order, with no total cap; max_peaks is then the per-spectrum capacity of the output (at most
   * peak_window_k x the number of windows in [fmin, fmax], which the caller must allow; the kept peaks are
   * clamped to it). 0 = the global view (top max_peaks). */
  int32_t peak_window_k;
  float peak_window_th;
  int32_t reserved[6];
} VhParams;

/* Device view of the peptidoform table (vulcan/index/build.py:13-21) and the fragment index.
 * Fragment index: rows are cut into buckets of 2^bucket_log2 consecutive rows (the table is sorted by
 * precursor mass, so a bucket is a precursor-mass interval). A fragment's key is the IEEE-754 bit pattern of
 * its f32 mass (order-preserving for positive floats, so the exact value is the key: no lossy quantisation) and
 * its bin is (key >> bin_shift) - bin_base. Postings are ordered bin-major: by (bin, bucket, key, row), so the
 * postings of consecutive buckets in one bin are contiguous. postings[i] = (key & (2^bin_shift - 1)) << 17 |
 * ion << 16 | (row - bucket start), ion 0 = b, 1 = y (the bin gives the key's high bits). The directory offsets[bin * n_buckets + bucket] =
 * first posting of (bin, bucket); offsets has n_bins * n_buckets + 1 entries (the last = n_postings). */
typedef struct VhIndexView {
  int64_t n_rows;
  const int64_t* seq_off;
  const uint8_t* length;
  const uint64_t* mask;
  const float* mass32;
  const uint8_t* residues;
  const int64_t* partner;
  int32_t bucket_log2;
  int32_t bin_shift;
  uint32_t bin_base;
  int32_t n_bins;
  int64_t n_buckets;
  const uint32_t* offsets;
  const uint32_t* postings;
  int64_t n_postings;
} VhIndexView;

/* Block layout of the same fragment index (bucket-join K1): rows are cut into blocks of consecutive rows, a row
 * starting a new block when floor(post_off[row] / 2^q_log2) or floor(row / 2^row_log2) changes, so a block holds
 * fewer than 2^q_log2 + 63 postings and at most 2^row_log2 rows. Inside a block the postings are sorted by key
 * (f32 bit pattern of the fragment mass). key[i] = that bit pattern; row[i] = ion << 15 | (row - row0[block]). */
typedef struct VhBlockView {
  int64_t n_blocks;
  int32_t max_rows;        /* rows of the largest block */
  int32_t max_postings;    /* postings of the largest block */
  const int32_t* row0;     /* n_blocks + 1: first row of each block; row0[n_blocks] = n_rows */
  const uint32_t* off;     /* n_blocks + 1: first posting of each block; off[n_blocks] = n_postings */
  const uint32_t* key;
  const uint16_t* row;
} VhBlockView;

/* ---------------------------------------------------------------------------------------------- misc */
int vh_version(void);
/* sizeof(VhParams) (which = 0), VhIndexView (1), VhBlockView (2), VhK2Cols (3), VhPsmCols (4), VhGbmNode (5) or
 * VhFamilyCols (6), to check a binding's struct layout. */
size_t vh_abi_sizeof(int which);
/* Scratch bytes of the prefix scans used inside vh_k0 / vh_hypotheses / vh_frag_count for n items. */
size_t vh_scan_workspace_bytes(int64_t n);
/* Plain copy kernel (n u32 words), used to calibrate HBM byte counters. No reference counterpart. */
int vh_copy_u32(const uint32_t* src, uint32_t* dst, int64_t n, void* stream);

/* ---------------------------------------------------------------------------------------------- index */
/* Per-row indexed-fragment count, then inclusive scan: post_off[0] = 0, post_off[i+1] = sum of counts of
 * rows 0..i. Replaces vk_frag_count (kernels.cu:53-71, 391-399) and the cumsum of engine.py:109-128. */
int vh_frag_count(const VhParams* p, const VhIndexView* ix, int64_t* post_off, void* ws, size_t ws_bytes,
                  void* stream);
/* Scratch bytes of vh_index_build for n_postings keys. */
size_t vh_index_workspace_bytes(int64_t n_postings, int64_t n_buckets, int32_t n_bins);
/* Emit 64-bit keys (bin << (17 + bin_shift + bucket bits) | bucket << (17 + bin_shift) | (f32 bits &
 * (2^bin_shift - 1)) << 17 | ion << 16 | row in bucket), LSD radix sort them on the device over the used bits
 * from 17 up (stable, so equal keys keep emission order), write postings (n_postings) and the directory offsets
 * (n_bins * n_buckets + 1). Replaces vk_frag_emit (kernels.cu:73-101, 401-411) and the torch.sort / bincount /
 * cumsum CSR assembly of engine.py:129-148. Needs bucket_log2 <= 16, bin_shift <= 15 and the key to fit 64 bits. */
int vh_index_build(const VhParams* p, const VhIndexView* ix, const int64_t* post_off, uint32_t* offsets,
                   uint32_t* postings, void* ws, size_t ws_bytes, void* stream);
/* Block layout (VhBlockView) of the index, built in two calls: vh_index_block_count writes stats[0] = n_blocks,
 * stats[1] = rows of the largest block, stats[2] = postings of the largest block (device u64[3]); the caller
 * allocates row0 / off (n_blocks + 1) and key / row (n_postings) and calls vh_index_build_blocks with the same
 * q_log2 (6..15) and row_log2 (4..15). Same fragments as vh_index_build (vk_frag_emit, kernels.cu:73-101; CSR
 * assembly of engine.py:129-148), a second ordering of them: by (block, key). Workspace: one query for both. */
size_t vh_index_blocks_workspace_bytes(int64_t n_rows, int64_t n_postings);
int vh_index_block_count(const VhIndexView* ix, const int64_t* post_off, int32_t q_log2, int32_t row_log2,
                         unsigned long long* stats, void* ws, size_t ws_bytes, void* stream);
int vh_index_build_blocks(const VhParams* p, const VhIndexView* ix, const int64_t* post_off, int32_t q_log2,
                          int32_t row_log2, int64_t n_blocks, int32_t* row0, uint32_t* off, uint32_t* key,
                          uint16_t* row, void* ws, size_t ws_bytes, void* stream);
/* Row-range slices (engine.py:59-65, 113-148: an index of more postings than one slice holds is cut into
 * consecutive row ranges, all resident): a slice's block layout is built on the view of its rows (n_rows = r1 -
 * r0, the per-row pointers advanced by r0; post_off from vh_frag_count on that view), then its row0 (n_blocks +
 * 1 entries) is moved to table rows by adding r0. The join K1 runs vh_k1j_pairs and vh_k1j_count once per slice
 * into the same per-window slot array (slices partition the rows, so every slot is written once) and one
 * vh_k1j_topk over it. */
int vh_index_rebase_rows(int32_t* row0, int64_t n, int32_t base, void* stream);

/* ---------------------------------------------------------------------------------------------- K0 */
/* Peak view (engine.py:171-210): keep fmin <= m/z <= fmax (not NaN), mass = m/z - PROTON (f32), top
 * max_peaks by intensity desc (ties: mass desc, then input order), then mass asc; TIC = f64 sum -> f32;
 * valid = kept >= min_peaks. Output is CSR: pk_off (n_spec+1); capacity = pk_capacity peaks. */
int vh_k0(const VhParams* p, int32_t n_spec, const float* mz, const float* intensity, const int64_t* raw_off,
          int64_t* pk_off, float* pk_mass, float* pk_int, float* tic, uint8_t* valid, int64_t pk_capacity,
          void* ws, size_t ws_bytes, void* stream);

/* vh_k0 without a workspace argument and without any allocation: the scan scratch is a static device array
 * of the library (65,792 bytes per GPU, held until the process ends), so n_spec is limited to
 * VH_K0_NOALLOC_MAX_SPECTRA (VH_E_LIMIT beyond). Same outputs as vh_k0. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under hipGraph capture, tight serving loops, hosts
 * that forbid hipMalloc). Every call on one GPU shares that GPU's scratch: calls that overlap on one GPU (two
 * streams, two threads) race and must be serialised by the caller. */
#define VH_K0_NOALLOC_MAX_SPECTRA (1 << 24)
int vh_k0_noalloc(const VhParams* p, int32_t n_spec, const float* mz, const float* intensity,
                  const int64_t* raw_off, int64_t* pk_off, float* pk_mass, float* pk_int, float* tic, uint8_t* valid,
                  int64_t pk_capacity, void* stream);

/* ---------------------------------------------------------------------------------------------- hypotheses */
/* Precursor hypotheses (host.py:64-91) on the device: order (spectrum, charge, isotope), f32 arithmetic
 * mz = pmz - PROTON, pm = mz*z, M = pm - iso*NEUTRON, window = M + M*ppm/1e6, rows = searchsorted left/right
 * on mass32; maxfz = max(min(z, max_fragment_charge+1), 2). hyp_off (n_spec+1) groups them by spectrum.
 * With n_offsets / open_search set, each spectrum's closed hypotheses are followed by its mass-offset and open
 * hypotheses (make_mode_hypotheses, vulcan/ptm/search.py:48-77; h_mode = VH_MODE_*, h_off = offset or NaN for
 * open, M of the isotope-0 hypothesis). h_mode / h_off may be NULL for closed-only parameters. */
int vh_hypotheses(const VhParams* p, int32_t n_spec, const double* precursor_mz, const int8_t* charge,
                  const uint8_t* valid, const float* mass32, int64_t n_rows, int64_t* hyp_off, int32_t* h_spec,
                  int8_t* h_z, int8_t* h_iso, float* h_pm, float* h_M, int64_t* h_lo, int64_t* h_hi,
                  int8_t* h_maxfz, uint8_t* h_mode, float* h_off, int64_t hyp_capacity, void* ws, size_t ws_bytes,
                  void* stream);

/* Order of the spectra for the K1/K2 launches: by the band of the first hypothesis' window start row (65,536
 * equal bands of the n_rows index rows, i.e. of precursor mass; spectra without hypotheses last; spectrum order
 * inside a band), by two device radix passes. n_rows < 2^47. Waves then visit the index in mass order and the kernels map
 * contiguous ranges of this order onto each XCD, so an XCD's L2 serves a narrow band of rows. Results do not
 * depend on the order. No reference counterpart (the reference launches work items in hypothesis order,
 * engine.py:213-229). */
size_t vh_spectrum_order_workspace_bytes(int32_t n_spec);
int vh_spectrum_order(int32_t n_spec, int64_t n_rows, const int64_t* hyp_off, const int64_t* h_lo, int32_t* order,
                      void* ws, size_t ws_bytes, void* stream);
/* Work-ordered variant for uneven batches: spectra by (band = first window row >> band_log2 ascending, estimated
 * K1 work descending, spectrum), with work = n_peaks * sum_h max(maxfz_h - 1, 1) + (window rows) / 4 saturated
 * at 2^20 - 1. band_log2 >= 28 gives one band (pure longest-first order, the classic list-scheduling rule for a
 * persistent launch); smaller values keep precursor-mass locality. n_spec < 2^24. Same workspace as
 * vh_spectrum_order. No reference counterpart. */
int vh_spectrum_order_work(int32_t n_spec, const int64_t* hyp_off, const int64_t* h_lo, const int64_t* h_hi,
                           const int8_t* h_maxfz, const int64_t* pk_off, int32_t band_log2, int32_t* order, void* ws,
                           size_t ws_bytes, void* stream);

/* ---------------------------------------------------------------------------------------------- K1 */
/* Candidate generation, append form (run_k1, engine.py:250-281; k1_kernel, kernels.cu:112-190): for every
 * hypothesis h and row r of its window, count(h, r) = #(indexed fragment of r, peak, z < maxfz) with
 * tol_bound(peak*z, lo) <= fragment <= tol_bound(peak*z, hi); append (h, r, count) when count >= threshold
 * (unordered; *n_out may exceed cap, then only cap entries are written); hyp_scored[h] = #rows with
 * count > 0, hyp_matched[h] = sum of counts (written, not accumulated). One wave64 per spectrum; `order`
 * (NULL = 0..n_spec-1) is the processing order of the spectra (vh_spectrum_order); h_mode (NULL = all closed)
 * selects k1_threshold_open for open hypotheses. */
int vh_k1(const VhParams* p, const VhIndexView* ix, int32_t n_spec, const int32_t* order, const int64_t* pk_off,
          const float* pk_mass, const int64_t* hyp_off, const int8_t* h_maxfz, const uint8_t* h_mode,
          const int64_t* h_lo, const int64_t* h_hi, int64_t cap,
          unsigned long long* n_out, int32_t* out_hyp, int32_t* out_row, int32_t* out_cnt, uint32_t* hyp_scored,
          unsigned long long* hyp_matched, void* stream);

/* Fused candidate generation + top-K + pair closure (run_k1 + select_topk, engine.py:250-312): per spectrum,
 * the top-K (count desc, row desc, hypothesis asc) of its appended entries, their partner rows (count -1),
 * deduplicated by row (lowest hypothesis, highest count), written in row order at s*cand_stride.
 * n_cand[s] = number of candidates (<= cand_stride). status[0] is set nonzero if a spectrum exceeded a
 * limit (its list is then empty). queue (NULL = grid launch, one wave per position of `order`): a device u32
 * that makes the launch persistent -- as many waves as stay resident loop over `order` through this atomic
 * counter (reset by the call itself on the stream), which keeps waves busy when spectra differ widely in work
 * (uneven batches). Results do not depend on the mode. */
int vh_k1_topk(const VhParams* p, const VhIndexView* ix, int32_t n_spec, const int32_t* order,
               const int64_t* pk_off, const float* pk_mass, const int64_t* hyp_off, const int8_t* h_maxfz,
               const uint8_t* h_mode, const int64_t* h_lo, const int64_t* h_hi, int32_t cand_stride,
               int32_t* n_cand, int32_t* c_row, int32_t* c_hyp, int32_t* c_cnt, uint32_t* hyp_scored,
               unsigned long long* hyp_matched, int32_t* status, uint32_t* queue, void* stream);

/* Index top-K with per-candidate features for the learned prefilter (no reference counterpart; the top-K set and
 * order are those of vh_k1_topk without closure, engine.py:284-312 at pair_closure off). Per spectrum s: the
 * topk best (count, row) entries of its hypotheses, deduplicated by row (lowest hypothesis, highest count) and
 * written in row order at s*cand_stride. Besides row / hypothesis / count (clamped at 32767): c_nb[c] = the
 * count's b-ion part (clamped at 32767); c_iq[c] = iq(sum over its b-ion matches) << 16 | iq(sum over its y-ion
 * matches), where a match adds the peak's quantised intensity trunc(I * 2^(16 - x)) (x = frexp exponent of the
 * spectrum's largest finite peak intensity; 0 for non-finite, negative or denormal I) and iq(s) = e << 8 | the 8
 * bits below the leading one of 1 + s (e = its leading bit position), a 14-bit pseudo-log2. s_maxcnt[s] = the
 * largest count of the spectrum's entries (0 if none). Among duplicates of a row the entry with the larger count
 * (ties: lower hypothesis) gives the features. */
int vh_k1_features(const VhParams* p, const VhIndexView* ix, int32_t n_spec, const int32_t* order,
                   const int64_t* pk_off, const float* pk_mass, const float* pk_int, const int64_t* hyp_off,
                   const int8_t* h_maxfz, const uint8_t* h_mode, const int64_t* h_lo, const int64_t* h_hi,
                   int32_t cand_stride, int32_t* n_cand, int32_t* c_row, int32_t* c_hyp, int32_t* c_cnt,
                   int16_t* c_nb, uint32_t* c_iq, int32_t* s_maxcnt, uint32_t* hyp_scored,
                   unsigned long long* hyp_matched, int32_t* status, uint32_t* queue, void* stream);

/* The per-candidate features of vh_k1_features (c_nb, c_iq, s_maxcnt; same definitions) computed after the
 * selection for the listed candidates only: for each candidate row the b / y matches of its indexed fragments with
 * the spectrum's kept peaks (tol_bound(mass * z, lo) <= f <= tol_bound(mass * z, hi), z < maxfz), counted and
 * intensity-summed (quantised as in vh_k1_features) against the mass-sorted peaks; among the spectrum's hypotheses
 * whose window holds the row, the largest count (ties: lower hypothesis) gives the features. The lists are those of
 * vh_k1_topk / vh_k1j_topk without pair closure (cand_stride slots per spectrum); status |= 32 if a recomputed
 * count differs from c_cnt (clamped at 32767). */
int vh_k1_candidate_features(const VhParams* p, const VhIndexView* ix, int32_t n_spec, const int32_t* order,
                             const int64_t* pk_off, const float* pk_mass, const float* pk_int,
                             const int64_t* hyp_off, const int8_t* h_maxfz, const int64_t* h_lo, const int64_t* h_hi,
                             int32_t cand_stride, const int32_t* n_cand, const int32_t* c_row, const int32_t* c_cnt,
                             int16_t* c_nb, uint32_t* c_iq, int32_t* s_maxcnt, int32_t* status, void* stream);

/* Brute-force counts for given (hypothesis, row) pairs without the index (k1_brute_kernel,
 * kernels.cu:195-232): the definition K1 must equal. */
int vh_k1_brute(const VhParams* p, const VhIndexView* ix, int64_t n, const int32_t* c_hyp, const int32_t* c_row,
                const int32_t* h_spec, const int8_t* h_maxfz, const int64_t* pk_off, const float* pk_mass,
                int32_t* out_cnt, void* stream);

/* Bucket-join K1: the candidate lists of vh_k1_topk / vh_k1_features (same definition, bit-identical output)
 * computed block-major instead of spectrum-major, for closed and mass-offset searches (precursor windows of
 * a few hundred rows). Three calls on the block layout (VhBlockView):
 *  1. vh_k1j_pairs: for every hypothesis h < hyp_off[n_spec]: its slots (one per window row; h_slot, max_hyp + 1
 *     entries, h_slot[h] = first slot of h, h_slot[n_hyp] = total) and one (block, h) pair per block its window
 *     overlaps, grouped by block: pairs[blk_pairs[b] .. blk_pairs[b + 1]) are the pairs of block b (blk_pairs:
 *     n_blocks + 1), each a 32-byte record {i64 first peak of the spectrum, i64 h_slot[h], i32 lo, i32 hi,
 *     i32 peak count, i32 maxfz | (s_qe + 2048) << 16}. status |= 8 when pair_cap or slot_cap is too small
 *     (results then incomplete; slots past slot_cap are never written).
 *  2. vh_k1j_count: per block, the block's postings are staged once in LDS and every pair's window rows in the
 *     block get their shared-peak count (the K1 count of vh_k1) at counts[h_slot[h] + row - lo]: u32 (mode 1)
 *     or, mode 2, one u64 per slot: count << 43 | min(b-ion count, 32767) << 28 | iq_log(b intensity sum) << 14 |
 *     iq_log(y intensity sum) (the per-ion terms of vh_k1_features); a count >= 2^20 is clamped and sets
 *     status bit 16.
 *  3. vh_k1j_topk (mode 1 counts) / vh_k1j_features (mode 2 counts): one wave per spectrum reads its slots,
 *     writes scored / matched per hypothesis and the top-K lists exactly as vh_k1_topk / vh_k1_features.
 * The reference computes the same lists with k1_kernel + select_topk (kernels.cu:112-190, engine.py:284-312). */
size_t vh_k1j_workspace_bytes(int64_t n_blocks, int64_t max_hyp);
int vh_k1j_pairs(const VhBlockView* bv, int32_t n_spec, const int64_t* hyp_off, const int32_t* h_spec,
                 const int64_t* h_lo, const int64_t* h_hi, const int8_t* h_maxfz, const int64_t* pk_off,
                 const int32_t* s_qe, int64_t max_hyp, int64_t* h_slot, int64_t slot_cap, uint32_t* blk_pairs,
                 void* pairs, int64_t pair_cap, int32_t* status, void* ws, size_t ws_bytes, void* stream);
int vh_k1j_count(const VhParams* p, const VhBlockView* bv, int32_t mode, const uint32_t* blk_pairs,
                 const void* pairs, int64_t pair_cap, const float* pk_mass, const float* pk_int, void* counts,
                 int32_t* status, void* stream);
/* Feature mode helper, called before vh_k1j_pairs: s_qe[s] = the intensity scale shift of spectrum s's kept
 * peaks (the power of two that brings its largest finite intensity into [2^15, 2^16), -1000 if none), the
 * quantisation of vh_k1_features; the pair records carry it (s_qe may be NULL for mode 1). */
int vh_k1j_iq_shift(int32_t n_spec, const int64_t* pk_off, const float* pk_int, int32_t* s_qe, void* stream);
int vh_k1j_topk(const VhParams* p, const VhIndexView* ix, int32_t n_spec, const int32_t* order,
                const int64_t* hyp_off, const uint8_t* h_mode, const int64_t* h_lo, const int64_t* h_hi,
                const int64_t* h_slot, int64_t slot_cap, const uint32_t* counts, int32_t cand_stride,
                int32_t* n_cand, int32_t* c_row, int32_t* c_hyp, int32_t* c_cnt, uint32_t* hyp_scored,
                unsigned long long* hyp_matched, int32_t* status, uint32_t* queue, void* stream);
int vh_k1j_features(const VhParams* p, const VhIndexView* ix, int32_t n_spec, const int32_t* order,
                    const int64_t* hyp_off, const uint8_t* h_mode, const int64_t* h_lo, const int64_t* h_hi,
                    const int64_t* h_slot, int64_t slot_cap, const unsigned long long* counts, int32_t cand_stride,
                    int32_t* n_cand, int32_t* c_row, int32_t* c_hyp, int32_t* c_cnt, int16_t* c_nb,
                    uint32_t* c_iq, int32_t* s_maxcnt, uint32_t* hyp_scored, unsigned long long* hyp_matched,
                    int32_t* status, uint32_t* queue, void* stream);

/* Wide-window K1 for open search: the lists of vh_k1_topk (same count, threshold, top-K, closure and dedupe
 * rules, same outputs) by one workgroup per spectrum walking its windows one index bucket at a time with an LDS
 * histogram of the bucket; each (peak, fragment charge) lookup reads one directory cell per bin. `ix` must be a
 * bin-major view with buckets of 2^VH_K1W_BUCKET_LOG2 rows (vh_index_build with that bucket_log2), bin_shift
 * <= 15, fewer than 2^32 postings and `ix->postings` 16-byte aligned (postings are read in aligned groups of
 * four; nothing past n_postings is read). topk <= 1024, max_peaks <= 512. VH_E_ARG otherwise. No reference
 * counterpart beyond the semantics (the reference runs its K1 on open windows in per-spectrum batches,
 * vulcan/ptm/search.py:119-175). */
#define VH_K1W_BUCKET_LOG2 14
int vh_k1w_topk(const VhParams* p, const VhIndexView* ix, int32_t n_spec, const int32_t* order, const int64_t* pk_off,
                const float* pk_mass, const int64_t* hyp_off, const int8_t* h_maxfz, const uint8_t* h_mode,
                const int64_t* h_lo, const int64_t* h_hi, int32_t cand_stride, int32_t* n_cand, int32_t* c_row,
                int32_t* c_hyp, int32_t* c_cnt, uint32_t* hyp_scored, unsigned long long* hyp_matched,
                int32_t* status, void* stream);

/* Candidates sorted by spectrum (non-decreasing h_spec[c_hyp[c]]): o_spec[c] = the candidate's spectrum (o_spec
 * may be NULL) and o_off[s] = the first candidate of spectrum s, s = 0..n_spec (o_off[n_spec] = n_cand), the
 * grouping vh_k2 / vh_rank take with grp_count NULL. Reference: the per-spectrum grouping run_k2 and psm_table
 * derive from the sorted candidate list (engine.py:315-351, 366-385). */
int vh_candidate_spectra(int64_t n_cand, const int32_t* c_hyp, const int32_t* h_spec, int32_t n_spec,
                         int32_t* o_spec, int64_t* o_off, void* stream);

/* ---------------------------------------------------------------------------------------------- K2 */
/* Exact b/y rescoring (k2_kernel, kernels.cu:263-376): all b_k / y_k (k = 0..L-2) at fragment charges
 * 1..maxfz-1 against the spectrum's first 512 mass-sorted peaks; most intense peak in the tolerance window
 * (ties: the later peak); matched counts, summed intensities, intensity-weighted ppm, longest consecutive
 * ladders; hyperscore = ln((sb+1)(sy+1)) + lnfact(mb) + lnfact(my) (255 if not finite).
 * Groups of candidates share a spectrum: group g has spectrum grp_spec[g] (NULL: g), first candidate
 * grp_start[g] (NULL: g*stride) and grp_count[g] candidates (grp_count NULL: grp_start holds n_groups + 1
 * offsets and group g has grp_start[g + 1] - grp_start[g]). maxfz of a candidate = h_maxfz[c_hyp[c]].
 * c_shift (NULL = closed search) = observed - candidate mass: a candidate with a nonzero shift scores shifted +
 * unshifted ions and the delta is placed on the residue that maximises the hyperscore (vk_k2_shift,
 * kernels.cu:291-363, 461-472; peptides up to 63 residues). One wave64 per group; `order` (NULL =
 * 0..n_groups-1) is the processing order of the groups; queue as in vh_k1_topk (persistent launch). */
int vh_k2(const VhParams* p, const VhIndexView* ix, int32_t n_groups, const int32_t* order, const int32_t* grp_spec,
          const int64_t* grp_start, const int32_t* grp_count, int32_t stride, const int32_t* c_row,
          const int32_t* c_hyp, const int8_t* h_maxfz, const float* c_shift, const int64_t* pk_off,
          const float* pk_mass, const float* pk_int, double* o_hs, float* o_sb, float* o_sy, float* o_ppm,
          int16_t* o_mb, int16_t* o_my, int16_t* o_lb, int16_t* o_ly, uint32_t* queue, void* stream);

/* Per-candidate precursor delta for vh_k2's shifted scoring: c_shift[c] = h_M[h] - mass32[row] when h is an
 * offset or open hypothesis, 0 for a closed one (search_modes, vulcan/ptm/search.py:187-189). Groups as in
 * vh_k2. */
int vh_candidate_shift(int32_t n_groups, const int64_t* grp_start, const int32_t* grp_count, int32_t stride,
                       const int32_t* c_row, const int32_t* c_hyp, const uint8_t* h_mode, const float* h_M,
                       const float* mass32, float* c_shift, void* stream);

/* ---------------------------------------------------------------------------------------------- ranking */
/* PSM ranking (psm_table, engine.py:366-385): per group, candidates with mb + my >= min_matched_peaks
 * ordered by (hyperscore desc, row asc); o_sel[g*report + i] = candidate index of rank i+1 (-1 if none),
 * o_next[g*report + i] = hyperscore of the following candidate in that order (0 if none). */
int vh_rank(const VhParams* p, int32_t n_groups, const int64_t* grp_start, const int32_t* grp_count,
            int32_t stride, const int32_t* c_row, const double* hs, const int16_t* mb, const int16_t* my,
            int32_t report, int64_t* o_sel, double* o_next, void* stream);

/* Per-candidate outputs of vh_k2 (inputs of vh_psm_features). */
typedef struct VhK2Cols {
  const double* hs;
  const float* sb;
  const float* sy;
  const float* ppm;
  const int16_t* mb;
  const int16_t* my;
  const int16_t* lb;
  const int16_t* ly;
} VhK2Cols;

/* PSM table columns (PSM_SCHEMA.md section 1), one slot per (group, rank), dense [n_groups * report]; a slot
 * whose vh_rank selection is -1 gets rank 0 and no other column is written. */
typedef struct VhPsmCols {
  int64_t* spec_idx;
  int64_t* pf_row;
  int32_t* rank;
  int8_t* charge;
  int8_t* isotope_error;
  float* expmass;
  float* calcmass;
  float* delta_mass_ppm;
  double* hyperscore;
  double* delta_next;
  double* delta_best;
  int16_t* matched_b;
  int16_t* matched_y;
  int16_t* matched_peaks;
  float* summed_intensity;
  float* matched_intensity_pct;
  float* average_ppm;
  int16_t* longest_b;
  int16_t* longest_y;
  float* longest_y_pct;
  int16_t* peptide_len;
  int32_t* k1_count;
  int32_t* scored_candidates;
  double* poisson;
  int32_t* hyp;               /* the PSM's hypothesis (index into the hypothesis arrays) */
} VhPsmCols;

/* The feature columns of psm_table (vulcan/gpu/engine.py:361-432) for the vh_rank selections sel / next_hs of
 * n_groups candidate groups (group g = spectrum grp_spec[g], or g when grp_spec is NULL): candidate columns
 * gathered (c_row, c_hyp, c_cnt, k2), hypothesis columns (h_z, h_iso, h_pm), mass error
 * |pm - calc - iso*NEUTRON| * 2e6 / (pm - iso*NEUTRON + calc) (f32), delta_next = hs - next, delta_best =
 * rank-1 hs - hs, matched_intensity_pct = 100 * (sb + sy) / tic (f32), longest_y_pct = longest_y / length,
 * scored_candidates = sum of h_scored over the spectrum's hypotheses, poisson = log10(lam^k e^-lam / e^lnfact(k))
 * with lam = sum h_matched / sum h_scored (f64; an infinite probability becomes 0, Sage's 1e-325). */
int vh_psm_features(const VhIndexView* ix, int32_t n_groups, const int32_t* grp_spec, int32_t report,
                    const int64_t* sel, const double* next_hs, const int32_t* c_row, const int32_t* c_hyp,
                    const int32_t* c_cnt, const VhK2Cols* k2, const int64_t* hyp_off, const int8_t* h_z,
                    const int8_t* h_iso, const float* h_pm, const uint32_t* h_scored,
                    const unsigned long long* h_matched, const float* tic, const VhPsmCols* out, void* stream);

/* ---------------------------------------------------------------------------------------------- learned prefilter */
/* No reference counterpart (the reference's learned models, vulcan/rescore/fast.py:198-254 and static.py:25-82,
 * rescore PSMs after K2). An MLP s = w3 . relu(W2 relu(W1 x + b1) + b2) + b3 over the 8 normalised features
 * x_j = (f_j - mu_j) * isd_j of an index top-K candidate chooses which candidates reach K2. Weights live as an
 * f32 parameter vector theta[VH_MLP_N_PARAMS] (W1 row-major [out][in] at 0, b1 at 2048, W2 [out][in] at 2304,
 * b2 at 67840, w3 at 68096, b3 at 68352); the kernels read a packed copy (bf16 W1/W2, f32 biases). Arithmetic:
 * x, W1, relu(layer 1) and W2 are rounded to bf16 (nearest even), products accumulate in f32 (MFMA), biases,
 * relu(layer 2) and layer 3 are f32. */

/* Pack theta + feature normalisation (mu[8], isd[8]) into blob (VH_MLP_BLOB_BYTES, device). */
int vh_mlp_pack(const float* theta, const float* mu, const float* isd, void* blob, void* stream);

/* Forward over a feature matrix x [n][8] (raw features): s[n]. With a1 (bf16 bits [n][256], relu of layer 1)
 * and a2 (f32 [n][256], relu of layer 2) the activations are stored for vh_mlp_backward (both or neither). */
int vh_mlp_forward(const void* blob, int64_t n, const float* x, float* s, uint16_t* a1, float* a2, void* stream);

/* Gradient of L = (1/n) sum_r (s_r - y_r)^2 with respect to theta, from the activations of vh_mlp_forward on the
 * same x: grad[VH_MLP_N_PARAMS]. dZ2 is rounded to bf16 for the two layer-2 products (MFMA, f32 accumulation);
 * the layer-1 and bias gradients are f32 sums in a fixed order (deterministic). */
size_t vh_mlp_workspace_bytes(int64_t n);
int vh_mlp_backward(const void* blob, int64_t n, const float* x, const uint16_t* a1, const float* a2, const float* s,
                    const float* y, float* grad, void* ws, size_t ws_bytes, void* stream);

/* Adam step on theta (f32): m1 = b1 m1 + (1-b1) g; m2 = b2 m2 + (1-b2) g^2;
 * theta -= lr * (m1 c1) / (sqrt(m2 c2) + eps), with c1 = 1/(1-b1^t), c2 = 1/(1-b2^t) given by the caller. */
int vh_mlp_adam(float* theta, const float* grad, float* m1, float* m2, int64_t n, float lr, float beta1,
                float beta2, float eps, float c1, float c2, void* stream);

/* Count-only feature source: for the filled slots of index top-K lists (stride slots per spectrum, n_cand
 * filled) zero the b-ion counts c_nb and intensity pseudo-logs c_iq, and s_maxcnt[s] = the spectrum's best count
 * (0 for an empty list). No reference counterpart (the learned prefilter is new). */
int vh_prefilter_count_features(int32_t n_spec, int32_t stride, const int32_t* n_cand, const int32_t* c_cnt,
                                int16_t* c_nb, uint32_t* c_iq, int32_t* s_maxcnt, void* stream);

/* Raw features of every candidate slot of vh_k1_features' lists (stride slots per spectrum): out[slot][8] =
 * (count, count / s_maxcnt (0 if s_maxcnt is 0), b-ion count c_nb, b-ion intensity pseudo-log (c_iq >> 16) / 256,
 * y-ion intensity pseudo-log (c_iq & 0xFFFF) / 256, ((h_M - mass32) / mass32) * 1e6, length, charge h_z), f32;
 * zeros for empty slots. */
int vh_prefilter_features(const VhIndexView* ix, int32_t n_spec, int32_t stride, const int32_t* n_cand,
                          const int32_t* c_row, const int32_t* c_hyp, const int32_t* c_cnt, const int16_t* c_nb,
                          const uint32_t* c_iq, const int32_t* s_maxcnt, const int8_t* h_z, const float* h_M,
                          float* out, void* stream);

/* MLP score of every non-empty candidate slot (features as vh_prefilter_features, computed in registers):
 * score[slot]; empty slots are not written. Only the filled slots are evaluated (their compact order is built
 * in the workspace: vh_prefilter_score_workspace_bytes). n_spec * stride < 2^32 - 64. */
size_t vh_prefilter_score_workspace_bytes(int32_t n_spec, int32_t stride);
int vh_prefilter_score(const VhIndexView* ix, const void* blob, int32_t n_spec, int32_t stride,
                       const int32_t* n_cand, const int32_t* c_row, const int32_t* c_hyp, const int32_t* c_cnt,
                       const int16_t* c_nb, const uint32_t* c_iq, const int32_t* s_maxcnt, const int8_t* h_z,
                       const float* h_M, float* score, void* ws, size_t ws_bytes, void* stream);

/* Per spectrum: the kept slots by (score desc, NaN last; ties: earlier slot) -- the best keep_min, then every
 * slot whose score is >= best score - margin (f32), at most keep_max (keep_min = keep_max: a fixed top-keep) --
 * then the reference's pair closure and dedupe (partner rows with count -1; by row: lowest hypothesis, highest
 * count; engine.py:284-312), written in row order at s*out_stride. stride_in <= 1024, keep_max <= 512. */
int vh_prefilter_select(const VhIndexView* ix, int32_t n_spec, int32_t stride_in, const int32_t* n_cand,
                        const int32_t* c_row, const int32_t* c_hyp, const int32_t* c_cnt, const float* score,
                        int32_t keep_min, int32_t keep_max, float margin, int32_t pair_closure, int32_t out_stride,
                        int32_t* out_n, int32_t* out_row, int32_t* out_hyp, int32_t* out_cnt, void* stream);

/* ---- gradient-boosted tree scorer (replaces the torch GpuHGB of vulcan/rescore/fast.py:198-254) ---- */
#define VH_GBM_CHUNK_NODES 1024   /* LDS node chunk: a tree has at most this many nodes */
#define VH_GBM_MAX_FEATURES 63

/* One tree node, 16 B. Internal: v = numeric threshold, feature = the split feature, children at left and
 * left + 1 (left: x <= v, right: x > v; a NaN goes left iff missing_left). Leaf: v = the leaf value, left = its
 * own index, feature = -1 (0xFFFF), is_leaf = 1. left is an index inside the node's chunk (node
 * chunk_node0[c] + left). */
typedef struct {
  double v;
  int32_t left;
  int16_t feature;
  uint8_t missing_left;
  uint8_t is_leaf;
} VhGbmNode;

/* sklearn HistGradientBoosting{Classifier,Regressor} raw prediction, one tree per iteration (binary or
 * regression; no categorical splits), in f64: out[i] = ((0 + baseline) + v_0(x_i)) + v_1(x_i) + ... in tree
 * order, v_t the leaf reached from tree_root[t] -- bit-identical to sklearn's decision_function / raw predict.
 * X: n rows of n_features f64 (row-major). The trees are cut into n_chunks runs of whole trees: chunk c holds
 * trees [chunk_tree0[c], chunk_tree0[c+1]) whose nodes are [chunk_node0[c], chunk_node0[c+1]), at most
 * VH_GBM_CHUNK_NODES nodes. max_depth: the deepest leaf's depth (walks stop after max_depth + 1 nodes).
 * status (zeroed by the caller): bit 1 = a walk did not end on a leaf within max_depth + 1 steps (a malformed
 * model: indexes beyond the chunk / features beyond n_features are clamped, so every read stays in bounds, but
 * the sample's sum is then wrong); bit 2 = the chunk table itself is malformed (a chunk of more than
 * VH_GBM_CHUNK_NODES nodes, node ranges outside [0, n_nodes) or not increasing, tree ranges outside
 * [0, n_trees)): it is clamped into bounds, so every read stays in bounds, and the sums are then wrong. */
int vh_gbm_decision(const VhGbmNode* nodes, int64_t n_nodes, const int32_t* tree_root, int32_t n_trees,
                    const int32_t* chunk_node0, const int32_t* chunk_tree0, int32_t n_chunks, int32_t max_depth,
                    double baseline, const double* X, int64_t n, int32_t n_features, double* out, int32_t* status,
                    void* stream);

/* ---- family top-2 retention (replaces the torch select_family / family_table of vulcan/gpu/family.py:90-203) -- */
/* Shortlist per spectrum, from the bucket-join K1's window counts (vh_k1j_pairs / vh_k1j_count, mode 1: u32 per
 * slot): the window rows with count >= p->k1_threshold; the best p->topk of every K1 family kfam[row] by (count
 * desc, row desc, hypothesis asc); each kept row plus its pair partner in each of the n_sets decoy sets
 * (partners[j * n_rows + row], count -1, not shortlisted); one entry per row in row order: the spectrum's lowest
 * hypothesis whose window holds the row (else the lowest source hypothesis), the highest count (clamped at
 * 32,767), shortlisted if any source was.
 * mode 0: out_off[s + 1] = the spectrum's entry count (out_off[0] = 0); a spectrum whose rows over the threshold
 *   (more than 256) or shortlist + partner entries (more than 1,024) do not fit one wave's LDS gets count 0 and
 *   is appended to `wide` (wide[0], zeroed by the caller, = the number listed; wide[1 + i] = the spectra): the
 *   vh_family_wide_* calls below compute those spectra and write their counts;
 * mode 2: out_off[1..n_spec] scanned into exclusive offsets (the total at out_off[n_spec]); needs ws;
 * mode 1: the entries of the spectra that fit at out_off[s] (o_row, o_hyp, o_cnt, o_top); vh_family_wide_copy
 *   writes the wide ones.
 * status (zeroed by the caller) bit 1: a spectrum has more than 64 hypotheses, or a window lay beyond slot_cap
 * (its list is then incomplete). n_rows < 2^28, 1 <= n_sets <= 8. */
size_t vh_family_select_workspace_bytes(int32_t n_spec);
int vh_family_select(const VhParams* p, int32_t n_spec, const int64_t* hyp_off, const int64_t* h_lo,
                     const int64_t* h_hi, const int64_t* h_slot, int64_t slot_cap, const uint32_t* counts,
                     const uint8_t* kfam, int64_t n_rows, const int64_t* partners, int32_t n_sets, int32_t mode,
                     int64_t* out_off, int32_t* o_row, int32_t* o_hyp, int32_t* o_cnt, uint8_t* o_top, int32_t* wide,
                     int32_t* status, void* ws, size_t ws_bytes, void* stream);
/* Wide path, for the n_wide spectra wide_spec[0..n_wide) that mode 0 listed. 1. vh_family_wide_stats: w_off
 * (n_wide + 1) = exclusive offsets of their rows over the threshold, rmin[q] = the lowest such row, gstats[0] =
 * the largest count (clamped at 32,767), gstats[1] = the largest row span. 2. vh_family_wide: the same shortlist
 * rule as vh_family_select through two device radix sorts (keys packed with field widths fitted to cmax / span /
 * n_kfam, the number of K1 families; batches of at most 8,192 spectra); w_off_host = a host copy of w_off;
 * writes each spectrum's entries packed into uniq (capacity (n_sets + 1) * w_off[n_wide]) at uniq_off[q],
 * uniq_cnt[q] of them, and the count to out_off[wide_spec[q] + 1] (before vh_family_select mode 2).
 * 3. vh_family_wide_copy (after mode 2): the packed entries into o_row / o_hyp / o_cnt / o_top at out_off. */
size_t vh_family_wide_stats_workspace_bytes(int32_t n_wide);
int vh_family_wide_stats(const VhParams* p, int32_t n_wide, const int32_t* wide_spec, const int64_t* hyp_off,
                         const int64_t* h_lo, const int64_t* h_hi, const int64_t* h_slot, int64_t slot_cap,
                         const uint32_t* counts, int64_t n_rows, int64_t* w_off, int64_t* rmin, uint64_t* gstats,
                         void* ws, size_t ws_bytes, void* stream);
size_t vh_family_wide_workspace_bytes(int64_t n_entries, int32_t n_sets);
int vh_family_wide(const VhParams* p, int32_t n_wide, const int32_t* wide_spec, const int64_t* hyp_off,
                   const int64_t* h_lo, const int64_t* h_hi, const int64_t* h_slot, int64_t slot_cap,
                   const uint32_t* counts, const uint8_t* kfam, int32_t n_kfam, int64_t n_rows,
                   const int64_t* partners, int32_t n_sets, const int64_t* w_off, const int64_t* w_off_host,
                   const int64_t* rmin, uint64_t cmax, uint64_t span, uint64_t* uniq, int64_t* uniq_off,
                   int64_t* uniq_cnt, int64_t* out_off, void* ws, size_t ws_bytes, void* stream);
int vh_family_wide_copy(int32_t n_wide, const int32_t* wide_spec, const int64_t* hyp_off, const int64_t* uniq_off,
                        const int64_t* uniq_cnt, const uint64_t* uniq, const int64_t* out_off, int32_t* o_row,
                        int32_t* o_hyp, int32_t* o_cnt, uint8_t* o_top, void* stream);
/* Top-2 table (family_table, family.py:160-203): per spectrum s, over its shortlist [cand_off[s],
 * cand_off[s + 1]) (rows ascending, vh_family_select's output order; c_hyp = the entries' hypotheses) with K2
 * results hs / mb / my per entry: an entry with mb + my >= p->min_matched_peaks is admissible in decoy context j
 * when it is shortlisted or the set-j partner of its row is a shortlisted entry of the spectrum; forward rows
 * (side < 0) join group (fwd_fam * n_sets + j) * 2, rows of decoy set j group (fwd_fam * n_sets + j) * 2 + 1; per
 * group the best two by (hyperscore desc, row asc). mode 0 / 2 / 1 as vh_family_select: counts, offsets, then
 * the table rows in (group, rank) order into `cols`: the entry's index, group, rank (1 or 2), row, hypothesis,
 * hyperscore, matched b / y and the hypothesis' charge / isotope (h_z / h_iso). n_fwd_fam * n_sets * 2 <= 256. */
typedef struct VhFamilyCols {
  int64_t* idx;
  int16_t* group;
  int8_t* rank;
  int32_t* row;
  int32_t* hyp;
  double* hs;
  int16_t* mb;
  int16_t* my;
  int8_t* z;
  int8_t* iso;
} VhFamilyCols;
size_t vh_family_top2_workspace_bytes(int32_t n_spec);
int vh_family_top2(const VhParams* p, int32_t n_spec, const int64_t* cand_off, const int32_t* c_row,
                   const int32_t* c_hyp, const uint8_t* c_top, const double* hs, const int16_t* mb, const int16_t* my,
                   const int8_t* h_z, const int8_t* h_iso, const int8_t* fwd_fam, const int8_t* side, int64_t n_rows,
                   const int64_t* partners, int32_t n_sets, int32_t n_fwd_fam, int32_t mode, int64_t* out_off,
                   const VhFamilyCols* cols, void* ws, size_t ws_bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* VULCAN_HIP_H */
