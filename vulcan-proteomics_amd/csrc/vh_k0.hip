// K0 peak view and precursor hypotheses on the device.
//
// K0 restates k0_preprocess (vulcan/gpu/engine.py:171-210, global view): keep fmin <= m/z <= fmax (not NaN),
// mass = m/z - PROTON, the top max_peaks peaks by (intensity desc with NaN first, mass desc, input order),
// then mass ascending (ties keep that order); TIC = f64 sum; valid = kept >= min_peaks.
// One wave64 per spectrum: in-range peaks are compacted into LDS (ballot prefix), each lane ranks its peaks by
// counting the peaks that precede them (broadcast LDS reads), the kept peaks are placed in rank order, and a
// second counting pass gives each its mass-ordered slot. Spectra with more than K0_LDS in-range peaks rank
// straight from global memory.
// Hypotheses restate make_hypotheses (vulcan/gpu/host.py:64-91) with the same f32 operation order.
#include "vh_common.h"

#define K0_WAVES 4
#define K0_LDS 512       // in-range peaks staged per wave
#define K0_SEL 512       // max_peaks limit of the device path

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
  for (int64_t i = a + lane_id(); i < b; i += 64) c += inrange(mz[i], P);
  c = wave_sum(c);
  if (lane_id() == 0) pk_off[s + 1] = c < P.max_peaks ? c : P.max_peaks;
}

__global__ __launch_bounds__(256) void k0_emit_kernel(DevP P, int n_spec, const float* __restrict__ mz,
                                                      const float* __restrict__ it, const int64_t* __restrict__ raw_off,
                                                      const int64_t* __restrict__ pk_off, float* pk_mass, float* pk_int,
                                                      float* tic, uint8_t* valid) {
  __shared__ float s_m[K0_WAVES][K0_LDS], s_i[K0_WAVES][K0_LDS];
  __shared__ float s_sm[K0_WAVES][K0_SEL], s_si[K0_WAVES][K0_SEL];
  const int w = threadIdx.x >> 6, lane = lane_id();
  int s = blockIdx.x * K0_WAVES + w;
  if (s >= n_spec) return;
  float* lm = s_m[w];
  float* li = s_i[w];
  float* sm = s_sm[w];
  float* si = s_si[w];
  const int64_t base = raw_off[s];
  const int n = (int)(raw_off[s + 1] - base);
  // compact in-range peaks (order preserved)
  int n_in = 0;
  for (int i0 = 0; i0 < n; i0 += 64) {
    int i = i0 + lane;
    float m = i < n ? mz[base + i] : 0.f;
    bool ok = i < n && inrange(m, P);
    uint64_t bal = ballot(ok);
    int pos = n_in + __popcll(bal & lanemask_lt());
    if (ok && pos < K0_LDS) { lm[pos] = m - VH_PROTON; li[pos] = it[base + i]; }
    n_in += __popcll(bal);
  }
  wave_sync();
  const int K = n_in < P.max_peaks ? n_in : P.max_peaks;
  if (n_in <= K0_LDS) {
    for (int i = lane; i < n_in; i += 64) {
      float mi = lm[i], ii = li[i];
      int r = 0;
      for (int j = 0; j < n_in; ++j) r += sel_before(li[j], lm[j], j, ii, mi, i);
      if (r < K) { sm[r] = mi; si[r] = ii; }
    }
  } else {
    // rank over the raw peaks from global memory (indices in raw order preserve the input order)
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
  }
  wave_sync();
  // mass ascending, ties keep the selection order
  const int64_t o = pk_off[s];
  double t = 0.0;
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
  return niso * nz;
}

__global__ __launch_bounds__(256) void hyp_count_kernel(DevP P, int n_spec, const double* __restrict__ pmz,
                                                        const int8_t* __restrict__ charge,
                                                        const uint8_t* __restrict__ valid, int64_t* hyp_off) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0) hyp_off[0] = 0;
  if (s >= n_spec) return;
  hyp_off[s + 1] = n_hyp_of(P, pmz[s], charge[s], valid[s]);
}

__device__ __forceinline__ int64_t lower_bound_f(const float* a, int64_t n, float x) {
  int64_t L = 0, R = n;
  while (L < R) {
    int64_t M = (L + R) >> 1;
    if (a[M] < x) L = M + 1; else R = M;
  }
  return L;
}

__device__ __forceinline__ int64_t upper_bound_f(const float* a, int64_t n, float x) {
  int64_t L = 0, R = n;
  while (L < R) {
    int64_t M = (L + R) >> 1;
    if (!(x < a[M])) L = M + 1; else R = M;
  }
  return L;
}

__global__ __launch_bounds__(256) void hyp_emit_kernel(DevP P, int n_spec, const double* __restrict__ pmz,
                                                       const int8_t* __restrict__ charge,
                                                       const int64_t* __restrict__ hyp_off,
                                                       const float* __restrict__ mass32, int64_t n_rows,
                                                       int32_t* h_spec, int8_t* h_z, int8_t* h_iso, float* h_pm,
                                                       float* h_M, int64_t* h_lo, int64_t* h_hi, int8_t* h_maxfz) {
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
      h_lo[o] = lower_bound_f(mass32, n_rows, lo);
      h_hi[o] = upper_bound_f(mass32, n_rows, hi);
      h_maxfz[o] = (int8_t)(mf > 2 ? mf : 2);
    }
  }
}

extern "C" {

int vh_k0(const VhParams* p, int32_t n_spec, const float* mz, const float* intensity, const int64_t* raw_off,
          int64_t* pk_off, float* pk_mass, float* pk_int, float* tic, uint8_t* valid, int64_t pk_capacity,
          void* ws, size_t ws_bytes, void* stream) {
  if (!p || n_spec < 0 || !raw_off || !pk_off || !tic || !valid) return VH_E_ARG;
  if (p->max_peaks < 0 || p->max_peaks > K0_SEL) return VH_E_LIMIT;
  if (pk_capacity < (int64_t)n_spec * p->max_peaks) return VH_E_CAPACITY;
  if (n_spec == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  DevP P = vh_make_devp(p);
  unsigned g = (unsigned)((n_spec + K0_WAVES - 1) / K0_WAVES);
  k0_count_kernel<<<g, 64 * K0_WAVES, 0, st>>>(P, n_spec, mz, raw_off, pk_off);
  VH_CHECK_LAUNCH();
  int rc = vh_scan_inclusive_i64(pk_off + 1, n_spec, ws, ws_bytes, st);
  if (rc) return rc;
  k0_emit_kernel<<<g, 64 * K0_WAVES, 0, st>>>(P, n_spec, mz, intensity, raw_off, pk_off, pk_mass, pk_int, tic,
                                              valid);
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_hypotheses(const VhParams* p, int32_t n_spec, const double* precursor_mz, const int8_t* charge,
                  const uint8_t* valid, const float* mass32, int64_t n_rows, int64_t* hyp_off, int32_t* h_spec,
                  int8_t* h_z, int8_t* h_iso, float* h_pm, float* h_M, int64_t* h_lo, int64_t* h_hi,
                  int8_t* h_maxfz, int64_t hyp_capacity, void* ws, size_t ws_bytes, void* stream) {
  if (!p || n_spec < 0 || !hyp_off) return VH_E_ARG;
  int niso = p->iso_hi - p->iso_lo + 1, nz = p->z_hi - p->z_lo + 1;
  if (niso < 1 || nz < 1) return VH_E_ARG;
  int maxh = niso * nz > niso ? niso * nz : niso;
  if (maxh > VH_MAX_HYP_PER_SPEC) return VH_E_LIMIT;
  if (hyp_capacity < (int64_t)n_spec * maxh) return VH_E_CAPACITY;
  if (n_spec == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  DevP P = vh_make_devp(p);
  unsigned g = (unsigned)((n_spec + 255) / 256);
  hyp_count_kernel<<<g, 256, 0, st>>>(P, n_spec, precursor_mz, charge, valid, hyp_off);
  VH_CHECK_LAUNCH();
  int rc = vh_scan_inclusive_i64(hyp_off + 1, n_spec, ws, ws_bytes, st);
  if (rc) return rc;
  hyp_emit_kernel<<<g, 256, 0, st>>>(P, n_spec, precursor_mz, charge, hyp_off, mass32, n_rows, h_spec, h_z, h_iso,
                                     h_pm, h_M, h_lo, h_hi, h_maxfz);
  VH_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
