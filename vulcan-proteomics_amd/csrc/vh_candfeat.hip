// Per-candidate K1 features after the index top-K: the b-ion share of the shared-peak count and the per-ion
// intensity pseudo-logs of vh_k1_features, computed for the selected candidates only (no per-slot feature
// counters while counting).
//
// Definition (the one vh_k1_features / the feature-mode K1 uses): for candidate (row r, hypothesis h), every
// indexed fragment f of r (b and y ions, vk_frag_emit's f32 ladder, fragment range and ion-index filter) and
// every kept peak p and fragment charge z < maxfz(h) with tol_bound(mass_p * z, lo) <= f <= tol_bound(mass_p * z,
// hi) is a match; the b-ion matches are counted and each ion type's matches sum the peaks' quantised intensities
// (iq_quant with the spectrum's iq_shift). Both bounds grow with the peak mass, so the matching peaks of one
// (f, z) are a contiguous run of the mass-sorted peaks: it starts at or after a mass-table cell below
// (f - tolerance) / z and ends at the first peak whose low bound passes f.
// Among hypotheses of the spectrum whose window holds the row, the one with the largest count (ties: the lower
// hypothesis) gives the features, as the feature-mode dedupe does; the total count is checked against c_cnt.
#include "vh_topk.h"

#define CF_CELLS 1024

struct CFArgs {
  const int32_t* order;
  const int64_t* pk_off;
  const float* pk_mass;
  const float* pk_int;
  const int64_t* hyp_off;
  const int8_t* h_maxfz;
  const int64_t* h_lo;
  const int64_t* h_hi;
  int32_t stride;
  const int32_t* n_cand;
  const int32_t* c_row;
  const int32_t* c_cnt;
  int16_t* c_nb;
  uint32_t* c_iq;
  int32_t* s_maxcnt;
  int32_t* status;
};

struct CFPeaks {
  const float* pm;
  const uint32_t* pre;   // prefix sums of the quantised intensities (np + 1)
  const uint16_t* tb;    // first peak of each mass cell
  int np;
  float inv_w;
};

__device__ __forceinline__ int cf_cell(float x, float inv_w) {
  const float f = floorf(x * inv_w);
  return f < 0.f ? 0 : (f > (float)(CF_CELLS - 1) ? CF_CELLS - 1 : (int)f);
}

// matches of fragment f at charge z: count and intensity sum
__device__ __forceinline__ void cf_match(const DevP& P, const CFPeaks& V, float f, int z, float margin,
                                         uint32_t& cnt, uint64_t& isum) {
  const float x = ((f - margin) / (float)z) * 0.99999f - 1e-3f;   // below every matching peak mass
  int i = V.tb[cf_cell(x, V.inv_w)];
  int first = -1, last = -1;
  for (; i < V.np; ++i) {
    const float m = V.pm[i] * (float)z;
    if (tol_bound(m, P.tol_lo, P.tol_da) > f) break;   // low bounds only grow from here
    if (tol_bound(m, P.tol_hi, P.tol_da) >= f) {
      if (first < 0) first = i;
      last = i;
    }
  }
  if (first >= 0) {
    cnt += (uint32_t)(last - first + 1);
    isum += V.pre[last + 1] - V.pre[first];
  }
}

// b count, b sum, y sum and total count of (row, maxfz) against the staged peaks
__device__ __forceinline__ void cf_row(const DevP& P, const DevIdx& X, const float* rm, const CFPeaks& V, int r,
                                       int maxfz, float margin_da, float margin_ppm, uint32_t& nb, uint64_t& sb,
                                       uint64_t& sy, uint32_t& tot) {
  const uint8_t* sq = X.residues + X.seq_off[r];
  const int L = X.length[r];
  const uint64_t mk = X.mask[r];
  const float nt = (mk >> 63) ? P.nterm : 0.f;
  float b = nt, y = X.mass32[r] - nt;
  uint32_t ny = 0;
  nb = sb = sy = 0;
  for (int k = 0; k < L - 1; ++k) {
    const float rv = rm[((mk >> k) & 1ULL) * 26 + (sq[k] - 'A')];
    b = b + rv;
    y = y - rv;
    if (k + 1 > P.min_ion_index && b >= P.fmin && b <= P.fmax) {
      const float mg = margin_da + b * margin_ppm;
      for (int z = 1; z < maxfz; ++z) cf_match(P, V, b, z, mg, nb, sb);
    }
    if (L - 1 - k > P.min_ion_index && y >= P.fmin && y <= P.fmax) {
      const float mg = margin_da + y * margin_ppm;
      for (int z = 1; z < maxfz; ++z) cf_match(P, V, y, z, mg, ny, sy);
    }
  }
  tot = nb + ny;
}

__global__ __launch_bounds__(256) void cand_features_kernel(DevP P, DevIdx X, int n_spec, CFArgs A) {
  __shared__ float rm[52];
  __shared__ float s_pm[4][VH_K2_MAX_PEAKS];
  __shared__ uint32_t s_pre[4][VH_K2_MAX_PEAKS + 1];
  __shared__ uint16_t s_tb[4][CF_CELLS];
  load_rm(rm, P);
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = lane_id();
  const int wv = xcd_band_position(P.xcd_group) * (blockDim.x >> 6) + w;
  if (wv >= n_spec) return;   // a whole wave leaves; no block barrier follows
  const int s = A.order ? A.order[wv] : wv;
  const int nc = A.n_cand[s];
  const int64_t cb = (int64_t)s * A.stride;
  int mx = 0;
  for (int i = lane; i < nc; i += 64) mx = A.c_cnt[cb + i] > mx ? A.c_cnt[cb + i] : mx;
  for (int o = 32; o > 0; o >>= 1) {
    const int v = __shfl_xor(mx, o);
    mx = v > mx ? v : mx;
  }
  if (lane == 0) A.s_maxcnt[s] = nc ? mx : 0;
  if (nc == 0) return;
  // peaks (mass-sorted) and the prefix sums of their quantised intensities
  const int64_t p0 = A.pk_off[s];
  int np = (int)(A.pk_off[s + 1] - p0);
  np = np < VH_K2_MAX_PEAKS ? np : VH_K2_MAX_PEAKS;   // K0 keeps at most 512
  float* pm = s_pm[w];
  uint32_t* pre = s_pre[w];
  uint16_t* tb = s_tb[w];
  float imax = 0.f;
  for (int p = lane; p < np; p += 64) imax = fmaxf(imax, A.pk_int[p0 + p]);
  for (int o = 32; o > 0; o >>= 1) imax = fmaxf(imax, __shfl_xor(imax, o));
  const int qe = iq_shift(imax);
  uint32_t run = 0;
  for (int p0c = 0; p0c < np; p0c += 64) {
    const int p = p0c + lane;
    const float m = p < np ? A.pk_mass[p0 + p] : 0.f;
    uint32_t q = p < np ? iq_quant(A.pk_int[p0 + p], qe) : 0u;
    if (p < np) pm[p] = m;
    uint32_t inc = q;   // wave inclusive scan
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t v = __shfl_up(inc, o);
      if (lane >= o) inc += v;
    }
    if (p < np) pre[p + 1] = run + inc;
    run += __shfl(inc, 63);
  }
  if (lane == 0) pre[0] = 0u;
  wave_sync();
  const float top = np ? pm[np - 1] : 1.f;
  const float inv_w = (float)CF_CELLS / (top > 1.f ? top : 1.f);
  // tb[c] = first peak whose cell is >= c
  for (int p = lane; p < np; p += 64) {
    const int cj = cf_cell(pm[p], inv_w);
    const int cp = p ? cf_cell(pm[p - 1], inv_w) : -1;
    for (int c = cp + 1; c <= cj; ++c) tb[c] = (uint16_t)p;
  }
  {
    const int cl = np ? cf_cell(pm[np - 1], inv_w) : -1;
    for (int c = lane; c < CF_CELLS; c += 64)
      if (c > cl) tb[c] = (uint16_t)np;
  }
  wave_sync();
  const CFPeaks V = {pm, pre, tb, np, inv_w};
  // a conservative tolerance span (Da part and ppm part) below the fragment
  const float ta = fabsf(P.tol_lo) + fabsf(P.tol_hi);
  const float margin_da = P.tol_da ? ta : 0.f;
  const float margin_ppm = P.tol_da ? 0.f : ta * 1e-6f;
  const int64_t h0 = A.hyp_off[s];
  const int nh = (int)(A.hyp_off[s + 1] - h0);
  bool bad = false;
  for (int i = lane; i < nc; i += 64) {
    const int r = A.c_row[cb + i];
    uint32_t bnb = 0, btot = 0;
    uint64_t bsb = 0, bsy = 0;
    bool any = false;
    for (int h = 0; h < nh; ++h) {
      if (r < A.h_lo[h0 + h] || r >= A.h_hi[h0 + h]) continue;
      uint32_t nb, tot;
      uint64_t sb, sy;
      cf_row(P, X, rm, V, r, A.h_maxfz[h0 + h], margin_da, margin_ppm, nb, sb, sy, tot);
      if (!any || tot > btot) {
        bnb = nb;
        bsb = sb;
        bsy = sy;
        btot = tot;
        any = true;
      }
    }
    const uint32_t cc = (uint32_t)A.c_cnt[cb + i];
    bad |= (btot < 32767u ? btot : 32767u) != cc;
    A.c_nb[cb + i] = (int16_t)(bnb < 32767u ? bnb : 32767u);
    A.c_iq[cb + i] = (iq_log(bsb) << 16) | iq_log(bsy);
  }
  if (ballot(bad) && lane == 0) atomicOr(A.status, 32);
}

extern "C" {

int vh_k1_candidate_features(const VhParams* p, const VhIndexView* ix, int32_t n_spec, const int32_t* order,
                             const int64_t* pk_off, const float* pk_mass, const float* pk_int,
                             const int64_t* hyp_off, const int8_t* h_maxfz, const int64_t* h_lo, const int64_t* h_hi,
                             int32_t cand_stride, const int32_t* n_cand, const int32_t* c_row, const int32_t* c_cnt,
                             int16_t* c_nb, uint32_t* c_iq, int32_t* s_maxcnt, int32_t* status, void* stream) {
  if (!p || !ix || n_spec < 0 || !pk_off || !pk_mass || !pk_int || !hyp_off || !h_maxfz || !h_lo || !h_hi ||
      !n_cand || !c_row || !c_cnt || !c_nb || !c_iq || !s_maxcnt || !status || cand_stride < 1)
    return VH_E_ARG;
  if (n_spec == 0) return 0;
  CFArgs A = {order, pk_off, pk_mass, pk_int, hyp_off, h_maxfz, h_lo, h_hi, cand_stride, n_cand, c_row, c_cnt,
              c_nb, c_iq, s_maxcnt, status};
  const unsigned g = xcd_band_grid((n_spec + 3) / 4, p->xcd_group);
  cand_features_kernel<<<g, 256, 0, (hipStream_t)stream>>>(vh_make_devp(p), vh_make_devidx(ix), n_spec, A);
  VH_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
