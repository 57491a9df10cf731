// Fragment inverted index on the device: per-row fragment count, key emission, LSD radix sort, directory.
//
// Replaces frag_count_kernel / frag_emit_kernel (vulcan/gpu/kernels.cu:53-101) and the torch.sort + bincount +
// cumsum assembly of GpuIndex.__init__ (vulcan/gpu/engine.py:101-148). A fragment's key is its f32 bit pattern
// (order-preserving for positive floats: exact mass order), its bin the key's high bits. Emitted sort key =
// bin << (17 + S + nbk) | bucket << (17 + S) | (key & (2^S - 1)) << 17 | ion << 16 | row in bucket (S =
// bin_shift, nbk = bucket bits, bucket = row >> bucket_log2 with rows sorted by precursor mass, ion 0 = b,
// 1 = y). Bin-major order puts the postings of consecutive buckets of one bin next to each other, so a K1 lookup
// over a window's run of buckets reads one directory line and one contiguous posting range. Only the used bits
// from 17 up are sorted (stable LSD, 8-bit digits): equal keys keep emission order (row, then ladder order).
#include <string.h>

#include "vh_common.h"

// ------------------------------------------------------------------------------------------------ count / emit
__global__ __launch_bounds__(256) void frag_count_kernel(DevP P, DevIdx X, int64_t* post_off) {
  __shared__ float rm[52];
  load_rm(rm, P);
  __syncthreads();
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) post_off[0] = 0;
  if (i >= X.n_rows) return;
  const uint8_t* s = X.residues + X.seq_off[i];
  int L = X.length[i];
  uint64_t mk = X.mask[i];
  float nt = (mk >> 63) ? P.nterm : 0.f;
  float b = nt, y = X.mass32[i] - nt;
  int c = 0;
  for (int k = 0; k < L - 1; ++k) {
    float r = rm[((mk >> k) & 1ULL) * 26 + (s[k] - 'A')];
    b = b + r;
    y = y - r;
    if (k + 1 > P.min_ion_index && b >= P.fmin && b <= P.fmax) ++c;
    if (L - 1 - k > P.min_ion_index && y >= P.fmin && y <= P.fmax) ++c;
  }
  post_off[i + 1] = c;
}

struct KeyGeom {
  int S;      // bin_shift
  int nbk;    // bucket bits
  uint32_t bin_base;
  // block layout (vh_index_build_blocks): sort key = block << (16 + nd) | (key - fbase) << 16 | ion << 15 | row
  // in block, with the block of each row in bid and the first row of each block in brow0
  int block_mode;
  uint32_t fbase;     // f32 bits of fragment_min_mz
  int nd;             // bits of bits(fmax) - bits(fmin)
  const int64_t* bid;
  const int32_t* brow0;
  int64_t n_blocks;
};

__device__ __forceinline__ uint64_t posting_key(float f, uint64_t bucket, uint32_t local, uint32_t ion,
                                                const KeyGeom& G) {
  const uint32_t k = fbits(f);
  if (G.block_mode)
    return (bucket << (16 + G.nd)) | ((uint64_t)(k - G.fbase) << 16) | ((uint64_t)ion << 15) | (uint64_t)local;
  const uint64_t bin = (uint64_t)((k >> G.S) - G.bin_base);
  return (bin << (17 + G.S + G.nbk)) | (bucket << (17 + G.S)) | ((uint64_t)(k & ((1u << G.S) - 1u)) << 17) |
         ((uint64_t)ion << 16) | (uint64_t)local;
}

__global__ __launch_bounds__(256) void frag_emit_kernel(DevP P, DevIdx X, KeyGeom G, const int64_t* post_off,
                                                       uint64_t* keys) {
  __shared__ float rm[52];
  load_rm(rm, P);
  __syncthreads();
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= X.n_rows) return;
  const uint8_t* s = X.residues + X.seq_off[i];
  int L = X.length[i];
  uint64_t mk = X.mask[i];
  float nt = (mk >> 63) ? P.nterm : 0.f;
  float b = nt, y = X.mass32[i] - nt;
  uint64_t bucket = (uint64_t)i >> X.bucket_log2;
  uint32_t local = (uint32_t)(i & ((1LL << X.bucket_log2) - 1));
  if (G.block_mode) {   // (block ids beyond the caller's count are clamped: no out-of-bounds access)
    const int64_t bk = G.bid[i] - 1;
    bucket = (uint64_t)(bk < G.n_blocks ? bk : G.n_blocks - 1);
    local = (uint32_t)(i - G.brow0[bucket]) & 0x7FFFu;
  }
  int64_t o = post_off[i];
  for (int k = 0; k < L - 1; ++k) {
    float r = rm[((mk >> k) & 1ULL) * 26 + (s[k] - 'A')];
    b = b + r;
    y = y - r;
    if (k + 1 > P.min_ion_index && b >= P.fmin && b <= P.fmax) keys[o++] = posting_key(b, bucket, local, 0u, G);
    if (L - 1 - k > P.min_ion_index && y >= P.fmin && y <= P.fmax) keys[o++] = posting_key(y, bucket, local, 1u, G);
  }
}

// ------------------------------------------------------------------------------------------------ radix sort
// Tile = 4 waves x 16 steps x 64 lanes = 4096 keys; wave w owns positions [w*1024, (w+1)*1024) of the tile and
// processes them step-major, so per-wave running digit counters give stable ranks.
#define RS_THREADS 256
#define RS_STEPS 16
#define RS_TILE (RS_THREADS * RS_STEPS)

__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool ok) {
  uint64_t peers = ballot(ok);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    bool bit = (d >> b) & 1u;
    uint64_t m = ballot(bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

__global__ __launch_bounds__(RS_THREADS) void rs_upsweep_kernel(const uint64_t* __restrict__ keys, int64_t n,
                                                                 int shift, uint32_t* hist, int64_t n_tiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  int w = threadIdx.x >> 6;
  int64_t base = (int64_t)blockIdx.x * RS_TILE + w * (RS_STEPS * 64);
  for (int st = 0; st < RS_STEPS; ++st) {
    int64_t j = base + st * 64 + lane_id();
    bool ok = j < n;
    uint32_t d = ok ? (uint32_t)(keys[j] >> shift) & 255u : 0u;
    uint64_t peers = match_digit(d, ok);
    if (ok && __popcll(peers & lanemask_lt()) == 0) atomicAdd(&h[d], (uint32_t)__popcll(peers));
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * n_tiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(RS_THREADS) void rs_downsweep_kernel(const uint64_t* __restrict__ kin,
                                                                   uint64_t* __restrict__ kout, int64_t n, int shift,
                                                                   const uint32_t* __restrict__ hist_scan,
                                                                   int64_t n_tiles) {
  __shared__ uint32_t wcnt[4][256];
  __shared__ uint32_t goff[256];
  const int w = threadIdx.x >> 6, lane = lane_id();
  for (int i = threadIdx.x; i < 4 * 256; i += RS_THREADS) (&wcnt[0][0])[i] = 0;
  goff[threadIdx.x] = hist_scan[(int64_t)threadIdx.x * n_tiles + blockIdx.x];
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * RS_TILE + w * (RS_STEPS * 64);
  uint64_t k[RS_STEPS];
  uint32_t rank[RS_STEPS];
#pragma unroll
  for (int st = 0; st < RS_STEPS; ++st) {
    int64_t j = base + st * 64 + lane;
    bool ok = j < n;
    uint64_t key = ok ? kin[j] : 0ULL;
    uint32_t d = (uint32_t)(key >> shift) & 255u;
    uint64_t peers = match_digit(d, ok);
    uint32_t before = wcnt[w][d];
    wave_sync();
    if (ok && __popcll(peers & lanemask_lt()) == 0) wcnt[w][d] = before + (uint32_t)__popcll(peers);
    wave_sync();
    k[st] = key;
    rank[st] = before + (uint32_t)__popcll(peers & lanemask_lt());
  }
  __syncthreads();
  {
    int d = threadIdx.x;
    uint32_t acc = 0;
    for (int ww = 0; ww < 4; ++ww) {
      uint32_t c = wcnt[ww][d];
      wcnt[ww][d] = acc;
      acc += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int st = 0; st < RS_STEPS; ++st) {
    int64_t j = base + st * 64 + lane;
    if (j < n) {
      uint32_t d = (uint32_t)(k[st] >> shift) & 255u;
      kout[(int64_t)goff[d] + wcnt[w][d] + rank[st]] = k[st];
    }
  }
}

// ------------------------------------------------------------------------------------------------ directory
__device__ __forceinline__ int64_t dir_of(uint64_t key, const KeyGeom& G, const DevIdx& X) {
  const int64_t bin = (int64_t)(key >> (17 + G.S + G.nbk));
  const int64_t bucket = (int64_t)((key >> (17 + G.S)) & ((1ULL << G.nbk) - 1ULL));
  return bin * X.n_buckets + bucket;
}

__global__ __launch_bounds__(256) void directory_kernel(const uint64_t* __restrict__ keys, int64_t n, DevIdx X,
                                                        KeyGeom G, uint32_t* offsets, uint32_t* postings,
                                                        int64_t n_dir) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n == 0) {
    for (int64_t b = i; b <= n_dir; b += (int64_t)gridDim.x * blockDim.x) offsets[b] = 0;
    return;
  }
  if (i >= n) return;
  uint64_t kk = keys[i];
  int64_t d = dir_of(kk, G, X);
  int64_t prev = i ? dir_of(keys[i - 1], G, X) : -1;
  for (int64_t b = prev + 1; b <= d; ++b) offsets[b] = (uint32_t)i;
  postings[i] = (uint32_t)(((kk >> 17) & ((1ULL << G.S) - 1ULL)) << 17) | (uint32_t)(kk & 0x1FFFFULL);
  if (i == n - 1)
    for (int64_t b = d + 1; b <= n_dir; ++b) offsets[b] = (uint32_t)n;
}

// ------------------------------------------------------------------------------------------------ blocks
// Block boundaries: row r starts a block when floor(post_off[r] / 2^q) or floor(r / 2^rb) differs from row r-1's,
// so a block holds fewer than 2^q + (postings of one row) postings and at most 2^rb rows.
__global__ __launch_bounds__(256) void block_flag_kernel(const int64_t* __restrict__ post_off, int64_t n, int q,
                                                         int rb, int64_t* flag) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  flag[r] = r == 0 || (post_off[r] >> q) != (post_off[r - 1] >> q) || (r >> rb) != ((r - 1) >> rb) ? 1 : 0;
}

// bid = inclusive scan of the flags (block of row r = bid[r] - 1): first row of each block, the row count as the
// sentinel, and (stats != NULL) the block count, largest row count and largest posting count
// (row0 has cap + 1 entries: writes beyond are dropped, so a wrong block count cannot write out of bounds)
__global__ __launch_bounds__(256) void block_rows_kernel(const int64_t* __restrict__ post_off,
                                                         const int64_t* __restrict__ bid, int64_t n, int64_t cap,
                                                         int32_t* row0) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  if ((r == 0 || bid[r] != bid[r - 1]) && bid[r] - 1 <= cap) row0[bid[r] - 1] = (int32_t)r;
  if (r == n - 1 && bid[r] <= cap) row0[bid[r]] = (int32_t)n;
}

__global__ __launch_bounds__(256) void block_stats_kernel(const int64_t* __restrict__ post_off,
                                                          const int32_t* __restrict__ row0,
                                                          const int64_t* __restrict__ bid, int64_t n,
                                                          unsigned long long* stats) {
  const int64_t nb = bid[n - 1];
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b == 0) stats[0] = (unsigned long long)nb;
  if (b >= nb) return;
  atomicMax(&stats[1], (unsigned long long)(row0[b + 1] - row0[b]));
  atomicMax(&stats[2], (unsigned long long)(post_off[row0[b + 1]] - post_off[row0[b]]));
}

__global__ void block_stats_zero_kernel(unsigned long long* stats) {
  if (threadIdx.x < 3) stats[threadIdx.x] = 0ULL;
}

// sorted block-layout keys -> first posting of each block, key (f32 bits) and ion << 15 | row in block
__global__ __launch_bounds__(256) void block_split_kernel(const uint64_t* __restrict__ keys, int64_t n, KeyGeom G,
                                                          int64_t n_blocks, uint32_t* off, uint32_t* key,
                                                          uint16_t* row) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n == 0) {
    for (int64_t b = i; b <= n_blocks; b += (int64_t)gridDim.x * blockDim.x) off[b] = 0;
    return;
  }
  if (i >= n) return;
  const uint64_t kk = keys[i];
  const int64_t d = (int64_t)(kk >> (16 + G.nd));
  const int64_t prev = i ? (int64_t)(keys[i - 1] >> (16 + G.nd)) : -1;
  for (int64_t b = prev + 1; b <= d && b <= n_blocks; ++b) off[b] = (uint32_t)i;
  key[i] = (uint32_t)((kk >> 16) & ((1ULL << G.nd) - 1ULL)) + G.fbase;
  row[i] = (uint16_t)(kk & 0xFFFFULL);
  if (i == n - 1)
    for (int64_t b = d + 1; b <= n_blocks; ++b) off[b] = (uint32_t)n;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static int64_t n_tiles_of(int64_t n) { return (n + RS_TILE - 1) / RS_TILE; }

size_t vh_radix_ws_bytes(int64_t n) {
  int64_t nt = n_tiles_of(n > 0 ? n : 1);
  size_t hist = (size_t)256 * nt;
  return align256(hist * 4) + align256(vh_scan_ws_u32(hist));
}

// Stable LSD radix sort of n u64 keys on bits [bit_lo, bit_hi) (8-bit digits), ping-ponging between a and b.
// *result receives the buffer that holds the sorted keys.
int vh_radix_sort_u64(uint64_t* a, uint64_t* b, int64_t n, int bit_lo, int bit_hi, void* ws, size_t ws_bytes,
                      hipStream_t st, uint64_t** result) {
  *result = a;
  if (n <= 0) return 0;
  if (ws_bytes < vh_radix_ws_bytes(n)) return VH_E_WORKSPACE;
  int64_t nt = n_tiles_of(n);
  size_t hist_n = (size_t)256 * nt;
  uint32_t* hist = (uint32_t*)ws;
  void* sws = (char*)ws + align256(hist_n * 4);
  size_t sws_bytes = ws_bytes - align256(hist_n * 4);
  uint64_t* src = a;
  uint64_t* dst = b;
  for (int shift = bit_lo; shift < bit_hi; shift += 8) {
    rs_upsweep_kernel<<<(unsigned)nt, RS_THREADS, 0, st>>>(src, n, shift, hist, nt);
    int rc = vh_scan_exclusive_u32(hist, (int64_t)hist_n, sws, sws_bytes, st);
    if (rc) return rc;
    rs_downsweep_kernel<<<(unsigned)nt, RS_THREADS, 0, st>>>(src, dst, n, shift, hist, nt);
    VH_CHECK_LAUNCH();
    uint64_t* t = src;
    src = dst;
    dst = t;
  }
  *result = src;
  return 0;
}

extern "C" {

size_t vh_index_workspace_bytes(int64_t n_postings, int64_t n_buckets, int32_t n_bins) {
  (void)n_buckets;
  (void)n_bins;
  return 2 * align256((size_t)(n_postings > 0 ? n_postings : 1) * 8) + vh_radix_ws_bytes(n_postings);
}

int vh_frag_count(const VhParams* p, const VhIndexView* ix, int64_t* post_off, void* ws, size_t ws_bytes,
                  void* stream) {
  if (!p || !ix || !post_off || ix->n_rows < 0) return VH_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  DevP P = vh_make_devp(p);
  DevIdx X = vh_make_devidx(ix);
  int64_t n = ix->n_rows;
  unsigned grid = (unsigned)((n + 255) / 256);
  frag_count_kernel<<<grid ? grid : 1, 256, 0, st>>>(P, X, post_off);
  VH_CHECK_LAUNCH();
  return vh_scan_inclusive_i64(post_off + 1, n, ws, ws_bytes, st);
}

int vh_index_build(const VhParams* p, const VhIndexView* ix, const int64_t* post_off, uint32_t* offsets,
                   uint32_t* postings, void* ws, size_t ws_bytes, void* stream) {
  if (!p || !ix || !post_off || !offsets || (ix->n_postings && !postings)) return VH_E_ARG;
  if (ix->bin_shift < 4 || ix->bin_shift > 15 || ix->bucket_log2 < 0 || ix->bucket_log2 > 16) return VH_E_ARG;
  if (!(p->fmin > 0.f) || !(p->fmax >= p->fmin)) return VH_E_ARG;
  if (ix->n_postings < 0 || ix->n_postings >= (1LL << 32)) return VH_E_LIMIT;
  int64_t n_dir = ix->n_buckets * (int64_t)ix->n_bins;
  if (n_dir <= 0 || n_dir >= (1LL << 32)) return VH_E_LIMIT;
  if (ix->n_buckets < ((ix->n_rows + (1LL << ix->bucket_log2) - 1) >> ix->bucket_log2)) return VH_E_ARG;
  if (ws_bytes < vh_index_workspace_bytes(ix->n_postings, ix->n_buckets, ix->n_bins)) return VH_E_WORKSPACE;
  uint32_t fb0, fb1;
  memcpy(&fb0, &p->fmin, 4);
  memcpy(&fb1, &p->fmax, 4);
  KeyGeom G;
  G.S = ix->bin_shift;
  G.bin_base = ix->bin_base;
  G.block_mode = 0;
  G.fbase = fb0;
  G.nd = 1;
  G.bid = nullptr;
  G.brow0 = nullptr;
  G.n_blocks = 0;
  G.nbk = 0;
  while (G.nbk < 32 && ((uint64_t)(ix->n_buckets - 1) >> G.nbk) != 0) ++G.nbk;
  int nbn = 1;
  while (nbn < 32 && ((uint64_t)(ix->n_bins - 1) >> nbn) != 0) ++nbn;
  if (17 + G.S + G.nbk + nbn > 64) return VH_E_LIMIT;
  if ((fb0 >> ix->bin_shift) != ix->bin_base || (int64_t)((fb1 >> ix->bin_shift) - ix->bin_base) >= ix->n_bins)
    return VH_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  DevP P = vh_make_devp(p);
  DevIdx X = vh_make_devidx(ix);
  int64_t n = ix->n_postings;
  char* w = (char*)ws;
  size_t kb = align256((size_t)(n > 0 ? n : 1) * 8);
  uint64_t* ka = (uint64_t*)w;
  uint64_t* kb2 = (uint64_t*)(w + kb);
  unsigned grid = (unsigned)((ix->n_rows + 255) / 256);
  if (ix->n_rows) frag_emit_kernel<<<grid, 256, 0, st>>>(P, X, G, post_off, ka);
  VH_CHECK_LAUNCH();
  uint64_t* src = ka;
  int rc = vh_radix_sort_u64(ka, kb2, n, 17, 17 + G.S + G.nbk + nbn, w + 2 * kb, ws_bytes - 2 * kb, st, &src);
  if (rc) return rc;
  unsigned g2 = (unsigned)((n > 0 ? n : 1024) + 255) / 256;
  directory_kernel<<<g2, 256, 0, st>>>(src, n, X, G, offsets, postings, n_dir);
  VH_CHECK_LAUNCH();
  return 0;
}

size_t vh_index_blocks_workspace_bytes(int64_t n_rows, int64_t n_postings) {
  const size_t n = (size_t)(n_rows > 0 ? n_rows : 1);
  return 2 * align256((size_t)(n_postings > 0 ? n_postings : 1) * 8) + vh_radix_ws_bytes(n_postings) +
         align256(n * 8) + align256((n + 1) * 4) + align256(vh_scan_ws_i64(n_rows)) + 256;
}

// flags + scan: bid (n_rows i64) and row0 (n_blocks + 1, at most n_rows + 1 entries)
static int block_layout(const int64_t* post_off, int64_t n, int q, int rb, int64_t* bid, int64_t cap, int32_t* row0,
                        void* sws, size_t sws_bytes, hipStream_t st) {
  const unsigned g = (unsigned)((n + 255) / 256);
  block_flag_kernel<<<g, 256, 0, st>>>(post_off, n, q, rb, bid);
  VH_CHECK_LAUNCH();
  int rc = vh_scan_inclusive_i64(bid, n, sws, sws_bytes, st);
  if (rc) return rc;
  block_rows_kernel<<<g, 256, 0, st>>>(post_off, bid, n, cap, row0);
  VH_CHECK_LAUNCH();
  return 0;
}

static int block_args(const VhIndexView* ix, const int64_t* post_off, int32_t q_log2, int32_t row_log2) {
  if (!ix || !post_off || ix->n_rows < 1) return VH_E_ARG;
  if (q_log2 < 6 || q_log2 > 15 || row_log2 < 4 || row_log2 > 15) return VH_E_ARG;
  if (ix->n_rows >= (1LL << 31) || ix->n_postings < 0 || ix->n_postings >= (1LL << 32)) return VH_E_LIMIT;
  return 0;
}

int vh_index_block_count(const VhIndexView* ix, const int64_t* post_off, int32_t q_log2, int32_t row_log2,
                         unsigned long long* stats, void* ws, size_t ws_bytes, void* stream) {
  int rc = block_args(ix, post_off, q_log2, row_log2);
  if (rc) return rc;
  if (!stats) return VH_E_ARG;
  if (ws_bytes < vh_index_blocks_workspace_bytes(ix->n_rows, 0)) return VH_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = ix->n_rows;
  char* w = (char*)ws;
  int64_t* bid = (int64_t*)w;
  int32_t* row0 = (int32_t*)(w + align256((size_t)n * 8));
  void* sws = w + align256((size_t)n * 8) + align256((size_t)(n + 1) * 4);
  rc = block_layout(post_off, n, q_log2, row_log2, bid, n, row0, sws, vh_scan_ws_i64(n), st);
  if (rc) return rc;
  block_stats_zero_kernel<<<1, 64, 0, st>>>(stats);
  block_stats_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(post_off, row0, bid, n, stats);
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_index_build_blocks(const VhParams* p, const VhIndexView* ix, const int64_t* post_off, int32_t q_log2,
                          int32_t row_log2, int64_t n_blocks, int32_t* row0, uint32_t* off, uint32_t* key,
                          uint16_t* row, void* ws, size_t ws_bytes, void* stream) {
  int rc = block_args(ix, post_off, q_log2, row_log2);
  if (rc) return rc;
  if (!p || !row0 || !off || (ix->n_postings && (!key || !row)) || n_blocks < 1 || n_blocks > ix->n_rows)
    return VH_E_ARG;
  if (!(p->fmin > 0.f) || !(p->fmax >= p->fmin)) return VH_E_ARG;
  if (ws_bytes < vh_index_blocks_workspace_bytes(ix->n_rows, ix->n_postings)) return VH_E_WORKSPACE;
  uint32_t fb0, fb1;
  memcpy(&fb0, &p->fmin, 4);
  memcpy(&fb1, &p->fmax, 4);
  KeyGeom G;
  G.S = 0;
  G.bin_base = 0;
  G.block_mode = 1;
  G.fbase = fb0;
  G.nd = 1;
  while (G.nd < 32 && ((uint64_t)(fb1 - fb0) >> G.nd) != 0) ++G.nd;
  G.nbk = 1;
  while (G.nbk < 40 && ((uint64_t)n_blocks >> G.nbk) != 0) ++G.nbk;
  if (16 + G.nd + G.nbk > 64) return VH_E_LIMIT;
  hipStream_t st = (hipStream_t)stream;
  const int64_t nr = ix->n_rows;
  const int64_t n = ix->n_postings;
  char* w = (char*)ws;
  const size_t kb = align256((size_t)(n > 0 ? n : 1) * 8);
  uint64_t* ka = (uint64_t*)w;
  uint64_t* kb2 = (uint64_t*)(w + kb);
  char* w2 = w + 2 * kb + vh_radix_ws_bytes(n);
  int64_t* bid = (int64_t*)w2;
  void* sws = w2 + align256((size_t)nr * 8) + align256((size_t)(nr + 1) * 4);
  // the caller's row0 has n_blocks + 1 entries: n_blocks must be the count vh_index_block_count reported
  rc = block_layout(post_off, nr, q_log2, row_log2, bid, n_blocks, row0, sws, vh_scan_ws_i64(nr), st);
  if (rc) return rc;
  G.bid = bid;
  G.brow0 = row0;
  G.n_blocks = n_blocks;
  DevP P = vh_make_devp(p);
  DevIdx X = vh_make_devidx(ix);
  frag_emit_kernel<<<(unsigned)((nr + 255) / 256), 256, 0, st>>>(P, X, G, post_off, ka);
  VH_CHECK_LAUNCH();
  uint64_t* src = ka;
  rc = vh_radix_sort_u64(ka, kb2, n, 16, 16 + G.nd + G.nbk, w + 2 * kb, vh_radix_ws_bytes(n), st, &src);
  if (rc) return rc;
  const unsigned g2 = (unsigned)((n > 0 ? n : 1024) + 255) / 256;
  block_split_kernel<<<g2, 256, 0, st>>>(src, n, G, n_blocks, off, key, row);
  VH_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
