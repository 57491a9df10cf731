// CPU ORACLE for the search hot path -- test infrastructure, not product code.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library, as the checker and as
// the multi-core CPU baseline. It restates the reference semantics function by function (citations are
// files of pedrobeltrao/vulcan-proteomics):
//   or_k0           vulcan/gpu/engine.py:171-210        (k0_preprocess, global and window peak views)
//   or_hypotheses   vulcan/gpu/host.py:64-91             (make_hypotheses, numpy f32 arithmetic)
//   or_k1_brute     vulcan/gpu/kernels.cu:195-232        (k1_brute_kernel: the definition of the K1 count)
//   or_k1_counts    vulcan/gpu/kernels.cu:112-190        (k1_kernel counts, scored, matched) computed per row
//                   from the fragment ladder with two binary searches per (fragment, charge) instead of an index
//   or_topk         vulcan/gpu/engine.py:284-312         (select_topk + pair closure)
//   or_topk_feat    vh_k1_features (no reference counterpart): the same top-K without closure, plus the
//                   quantised matched-intensity sum of each entry, computed per row like or_k1_counts
//   or_k2           vulcan/gpu/kernels.cu:237-376        (k2_kernel incl. the shifted branch; f64 hyperscore)
//   or_rank         vulcan/gpu/engine.py:361-386         (psm_table ordering: spectrum, hyperscore desc, row asc)
//   or_search       vulcan/gpu/engine.py:435-464         (the whole per-spectrum pipeline, OpenMP over spectra)
// Pinned by: K1 fast == K1 brute force; closed-form known-answer spectra (exact b/y ladders -> known counts,
// runs and hyperscores); lnfact closed form; brute-force Python orderings on tiny inputs (K0, top-K, ranking);
// the reference's own invariants (closure symmetry, window containment, label blindness).
// Build: oracle/Makefile (g++ -O3 -fopenmp -ffp-contract=off), no GPU.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "vulcan_hip.h"

namespace {

const float PROTON = 1.0072764f;   // host.py:10
const float NEUTRON = 1.00335f;    // host.py:11

struct Ctx {
  const VhParams* p;
  float rm[2][26];   // residue mass: mono + static (bit 0) / mono + var (bit 1), f32 (kernels.cu:47-51)
  explicit Ctx(const VhParams* pp) : p(pp) {
    for (int i = 0; i < 26; ++i) {
      rm[0][i] = pp->mono[i] + pp->stat[i];
      rm[1][i] = pp->mono[i] + pp->var[i];
    }
  }
  float nterm_of(uint64_t mask) const { return (mask >> 63) ? p->nterm : 0.0f; }
  float res(uint8_t r, int k, uint64_t mask) const { return rm[(mask >> k) & 1ULL][r - 'A']; }
};

// kernels.cu:36-38
inline float tol_bound(float m, float tol, int da) {
  if (da) return m + tol;
  float t = m * tol;
  t = t / 1000000.0f;
  return m + t;
}

inline bool gt_nan_first(float a, float b) { return (isnan(a) && !isnan(b)) || a > b; }
inline bool eq_nan(float a, float b) { return a == b || (isnan(a) && isnan(b)); }

// ------------------------------------------------------------------------------------------------ K0
// Returns number of kept peaks; writes mass-sorted (mass, intensity) and the f64 TIC.
int k0_one(const Ctx& c, const float* mz, const float* it, int64_t n, float* om, float* oi, double* tic) {
  const VhParams& p = *c.p;
  std::vector<int64_t> idx;
  std::vector<float> mass(n);
  idx.reserve(n);
  for (int64_t i = 0; i < n; ++i) {
    float m = mz[i];
    if (!isnan(m) && m >= p.fmin && m <= p.fmax) idx.push_back(i);
    mass[i] = m - PROTON;
  }
  // selection order: intensity desc (NaN first), mass desc, input order (engine.py:187-189, stable sorts); the
  // window view groups by floor(m/z / window_th) (f32 division) first and keeps the first window_k of each group
  // (engine.py:185-193), the global view the first max_peaks
  const bool win = p.peak_window_k > 0;
  auto wnd = [&](int64_t i) -> int64_t { return win ? (int64_t)floorf(mz[i] / p.peak_window_th) : 0; };
  std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
    if (wnd(a) != wnd(b)) return wnd(a) < wnd(b);
    if (!eq_nan(it[a], it[b])) return gt_nan_first(it[a], it[b]);
    if (mass[a] != mass[b]) return mass[a] > mass[b];
    return a < b;
  });
  if (win) {
    std::vector<int64_t> kept;
    for (size_t j = 0, first = 0; j < idx.size(); ++j) {
      if (j > 0 && wnd(idx[j]) != wnd(idx[j - 1])) first = j;
      if ((int64_t)(j - first) < p.peak_window_k) kept.push_back(idx[j]);
    }
    idx.swap(kept);
    if ((int64_t)idx.size() > p.max_peaks) idx.resize(p.max_peaks);   // the output capacity (a caller limit)
  } else if ((int64_t)idx.size() > p.max_peaks) {
    idx.resize(p.max_peaks);
  }
  const int64_t k = (int64_t)idx.size();
  // final order: mass asc, ties keep the selection order (engine.py:196-197)
  std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return mass[a] < mass[b]; });
  double t = 0.0;
  for (int64_t j = 0; j < k; ++j) {
    om[j] = mass[idx[j]];
    oi[j] = it[idx[j]];
    t += (double)it[idx[j]];
  }
  *tic = t;
  return (int)k;
}

// ------------------------------------------------------------------------------------------------ hypotheses
int n_hyp_of(const VhParams& p, double pmz, int8_t charge, bool valid) {
  if (!valid || !std::isfinite(pmz)) return 0;
  int niso = p.iso_hi - p.iso_lo + 1;
  int nz = charge > 0 ? 1 : (p.z_hi - p.z_lo + 1);
  int extra = (p.iso_lo <= 0 && p.iso_hi >= 0) ? nz * (p.n_offsets + (p.open_search ? 1 : 0)) : 0;
  return niso * nz + extra;
}

int8_t maxfz_of(const VhParams& p, int z) {
  int m = p.max_fragment_charge ? std::min(z, p.max_fragment_charge + 1) : z;
  return (int8_t)std::max(m, 2);
}

// closed hypotheses (host.py:64-91), then per mass offset and for the open window one hypothesis per charge at
// isotope 0 (make_mode_hypotheses, vulcan/ptm/search.py:48-77), all of one spectrum together
void hyp_one(const VhParams& p, int32_t s, double pmz, int8_t charge, const float* mass32, int64_t n_rows,
             int64_t o, int32_t* h_spec, int8_t* h_z, int8_t* h_iso, float* h_pm, float* h_M, int64_t* h_lo,
             int64_t* h_hi, int8_t* h_maxfz, uint8_t* h_mode, float* h_off) {
  int zlo = charge > 0 ? charge : p.z_lo, zhi = charge > 0 ? charge : p.z_hi;
  float mz = (float)pmz - PROTON;
  for (int z = zlo; z <= zhi; ++z) {
    for (int iso = p.iso_lo; iso <= p.iso_hi; ++iso, ++o) {
      float pm = mz * (float)z;
      float ie = (float)iso * NEUTRON;
      float M = pm - ie;
      float lo = M + (M * p.prec_lo) / 1000000.0f;
      float hi = M + (M * p.prec_hi) / 1000000.0f;
      h_spec[o] = s;
      h_z[o] = (int8_t)z;
      h_iso[o] = (int8_t)iso;
      h_pm[o] = pm;
      h_M[o] = M;
      h_lo[o] = std::lower_bound(mass32, mass32 + n_rows, lo) - mass32;
      h_hi[o] = std::upper_bound(mass32, mass32 + n_rows, hi) - mass32;
      h_maxfz[o] = maxfz_of(p, z);
      if (h_mode) { h_mode[o] = VH_MODE_CLOSED; h_off[o] = 0.f; }
    }
  }
  if (!(p.iso_lo <= 0 && p.iso_hi >= 0)) return;
  int nmode = p.n_offsets + (p.open_search ? 1 : 0);
  for (int q = 0; q < nmode; ++q) {
    bool open = q == p.n_offsets;
    for (int z = zlo; z <= zhi; ++z, ++o) {
      float pm = mz * (float)z;
      float M0 = pm - 0.0f * NEUTRON;
      float lo, hi;
      if (!open) {
        float Mo = M0 - p.offsets[q];
        lo = Mo + (Mo * p.prec_lo) / 1000000.0f;
        hi = Mo + (Mo * p.prec_hi) / 1000000.0f;
      } else {
        lo = M0 - p.open_hi;
        hi = M0 - p.open_lo;
      }
      h_spec[o] = s;
      h_z[o] = (int8_t)z;
      h_iso[o] = 0;
      h_pm[o] = pm;
      h_M[o] = M0;
      h_lo[o] = std::lower_bound(mass32, mass32 + n_rows, lo) - mass32;
      h_hi[o] = std::upper_bound(mass32, mass32 + n_rows, hi) - mass32;
      h_maxfz[o] = maxfz_of(p, z);
      h_mode[o] = open ? VH_MODE_OPEN : VH_MODE_OFFSET;
      h_off[o] = open ? NAN : p.offsets[q];
    }
  }
}

// ------------------------------------------------------------------------------------------------ K1
// Tolerance bounds of every (peak, charge) for one spectrum: a[z][p], b[z][p], nondecreasing in p.
struct PeakBounds {
  int np = 0, nz = 0;
  std::vector<float> a, b;
  void build(const Ctx& c, const float* pm, int n, int zmax) {
    np = n;
    nz = std::max(zmax - 1, 0);
    a.resize((size_t)nz * n);
    b.resize((size_t)nz * n);
    for (int z = 1; z < zmax; ++z)
      for (int q = 0; q < n; ++q) {
        float m = pm[q] * (float)z;
        a[(size_t)(z - 1) * n + q] = tol_bound(m, c.p->tol_lo, c.p->tol_da);
        b[(size_t)(z - 1) * n + q] = tol_bound(m, c.p->tol_hi, c.p->tol_da);
      }
  }
  // #(p : a_p <= f <= b_p) = #(a_p <= f) - #(b_p < f)
  int count(int z, float f) const {
    const float* A = a.data() + (size_t)(z - 1) * np;
    const float* Bv = b.data() + (size_t)(z - 1) * np;
    int ua = (int)(std::upper_bound(A, A + np, f) - A);
    int lb = (int)(std::lower_bound(Bv, Bv + np, f) - Bv);
    return ua - lb;
  }
};

int k1_count_row(const Ctx& c, const VhIndexView& ix, int64_t r, int maxfz, const PeakBounds& pb) {
  const VhParams& p = *c.p;
  const uint8_t* sq = ix.residues + ix.seq_off[r];
  int L = ix.length[r];
  uint64_t mk = ix.mask[r];
  float b = c.nterm_of(mk), y = ix.mass32[r] - c.nterm_of(mk);
  int cnt = 0;
  for (int k = 0; k < L - 1; ++k) {
    float rm = c.res(sq[k], k, mk);
    b = b + rm;
    y = y - rm;
    if (k + 1 > p.min_ion_index && b >= p.fmin && b <= p.fmax)
      for (int z = 1; z < maxfz; ++z) cnt += pb.count(z, b);
    if (L - 1 - k > p.min_ion_index && y >= p.fmin && y <= p.fmax)
      for (int z = 1; z < maxfz; ++z) cnt += pb.count(z, y);
  }
  return cnt;
}

// literal restatement of k1_brute_kernel (kernels.cu:195-232)
int k1_brute_pair(const Ctx& c, const VhIndexView& ix, int64_t r, int maxfz, const float* pm, int64_t np) {
  const VhParams& p = *c.p;
  const uint8_t* sq = ix.residues + ix.seq_off[r];
  int L = ix.length[r];
  uint64_t mk = ix.mask[r];
  float b = c.nterm_of(mk), y = ix.mass32[r] - c.nterm_of(mk);
  int cnt = 0;
  for (int k = 0; k < L - 1; ++k) {
    float rm = c.res(sq[k], k, mk);
    b = b + rm;
    y = y - rm;
    for (int t = 0; t < 2; ++t) {
      float f = t == 0 ? b : y;
      bool ok = t == 0 ? (k + 1 > p.min_ion_index) : (L - 1 - k > p.min_ion_index);
      if (!ok || f < p.fmin || f > p.fmax) continue;
      for (int64_t q = 0; q < np; ++q)
        for (int z = 1; z < maxfz; ++z) {
          float m = pm[q] * (float)z;
          float a = tol_bound(m, p.tol_lo, p.tol_da);
          float bd = tol_bound(m, p.tol_hi, p.tol_da);
          if (f >= a && f <= bd) ++cnt;
        }
    }
  }
  return cnt;
}

// ------------------------------------------------------------------------------------------------ top-K
struct Ent { int64_t row; int32_t hl; int32_t cnt; };

// select_topk for one spectrum (engine.py:284-312). entries: (row, local hyp, count >= threshold).
// Order: count (clamped at 32767) desc, row desc; identical keys (same row and count from two hypotheses):
// lower hypothesis first. Then pair closure (partner rows, count -1) and dedupe by row (min hyp, max count).
int topk_one(const VhParams& p, const int64_t* partner, std::vector<Ent>& e, Ent* out) {
  std::sort(e.begin(), e.end(), [](const Ent& a, const Ent& b) {
    int ca = std::min(a.cnt, 32767), cb = std::min(b.cnt, 32767);
    if (ca != cb) return ca > cb;
    if (a.row != b.row) return a.row > b.row;
    return a.hl < b.hl;
  });
  size_t k = std::min<size_t>(e.size(), (size_t)std::max(p.topk, 0));
  std::vector<Ent> sel(e.begin(), e.begin() + k);
  if (p.pair_closure)
    for (size_t i = 0; i < k; ++i) sel.push_back({partner[e[i].row], e[i].hl, -1});
  std::sort(sel.begin(), sel.end(), [](const Ent& a, const Ent& b) {
    if (a.row != b.row) return a.row < b.row;
    return a.hl < b.hl;
  });
  int n = 0;
  for (size_t i = 0; i < sel.size(); ++i) {
    if (n && out[n - 1].row == sel[i].row) {
      out[n - 1].cnt = std::max(out[n - 1].cnt, sel[i].cnt);
      continue;
    }
    out[n++] = sel[i];
  }
  return n;
}

// ------------------------------------------------------------------------------------------------ K2
double lnfact(int n) {   // kernels.cu:246-250 (Stirling; n == 0 -> 1.0 as in Sage)
  if (n == 0) return 1.0;
  double x = (double)n;
  return x * log(x) - x + 0.5 * log(x) + 0.5 * log(2.0 * 3.141592653589793 * x);
}

struct Run { int start, length, last, longest; };
inline void run_matched(Run& r, int idx) {   // kernels.cu:252-258
  if (r.last == idx) return;
  if (r.start + r.length == idx) { r.length += 1; r.longest = std::max(r.longest, r.length); }
  else { r.start = idx; r.length = 1; r.longest = std::max(r.longest, r.length); }
  r.last = idx;
}

const int K2_MAXPK = 512;
const int K2_MAXLEN = 64;

struct K2Out { double hs; float sb, sy, ppm; int mb, my, lb, ly; };

K2Out k2_one(const Ctx& c, const VhIndexView& ix, int64_t r, int maxfz, const float* sm, const float* si, int np,
             const float* c_shift_ptr, float shift_val) {
  const VhParams& p = *c.p;
  if (np > K2_MAXPK) np = K2_MAXPK;
  const uint8_t* sq = ix.residues + ix.seq_off[r];
  int L = ix.length[r];
  uint64_t mk = ix.mask[r];
  float sb = 0.f, sy = 0.f, ppm = 0.f;
  int mb = 0, my = 0;
  Run rb = {0, 0, 0, 0}, ry = {0, 0, 0, 0};
  float shift = c_shift_ptr ? shift_val : 0.f;
  int n_alt = (c_shift_ptr && shift != 0.f) ? 2 : 1;
  float iu[2][K2_MAXLEN], is_[2][K2_MAXLEN];
  unsigned char nu[2][K2_MAXLEN], ns[2][K2_MAXLEN];
  if (n_alt == 2) {
    for (int k = 0; k < K2_MAXLEN; ++k) {
      iu[0][k] = iu[1][k] = is_[0][k] = is_[1][k] = 0.f;
      nu[0][k] = nu[1][k] = ns[0][k] = ns[1][k] = 0;
    }
  }
  for (int kind = 0; kind < 2; ++kind) {
    float cum = kind == 0 ? c.nterm_of(mk) : ix.mass32[r] - c.nterm_of(mk);
    for (int k = 0; k < L - 1; ++k) {
      float rmv = c.res(sq[k], k, mk);
      cum = kind == 0 ? cum + rmv : cum - rmv;
      for (int z = 1; z < maxfz; ++z) {
        int best = -1;
        float mx = 0.f, mz = 0.f, best_mz = 0.f;
        for (int alt = 0; alt < n_alt; ++alt) {
          int best_a = -1;
          float mx_a = 0.f;
          float base = alt == 0 ? cum : cum + shift;
          mz = base / (float)z;
          float a = tol_bound(mz, p.tol_lo, p.tol_da);
          float bd = tol_bound(mz, p.tol_hi, p.tol_da);
          int j = (int)(std::lower_bound(sm, sm + np, a) - sm);
          for (; j < np && sm[j] <= bd; ++j) {
            if (si[j] >= mx) { mx = si[j]; best = j; best_mz = mz; }
            if (si[j] >= mx_a) { mx_a = si[j]; best_a = j; }
          }
          if (n_alt == 2 && best_a >= 0 && k < K2_MAXLEN) {
            if (alt == 0) { iu[kind][k] = iu[kind][k] + si[best_a]; nu[kind][k] += 1; }
            else { is_[kind][k] = is_[kind][k] + si[best_a]; ns[kind][k] += 1; }
          }
        }
        mz = best_mz;
        if (best >= 0) {
          float pmass = sm[best], pint = si[best];
          float num = (pint * fabsf(mz - pmass)) * 2.0e6f;
          ppm = ppm + num / (mz + pmass);
          if (kind == 0) { mb += 1; sb = sb + pint; run_matched(rb, k); }
          else { my += 1; sy = sy + pint; run_matched(ry, k); }
        }
      }
    }
  }
  if (n_alt == 2) {
    int Lk = std::min(L - 1, K2_MAXLEN);
    float b_u = 0.f, y_s = 0.f, b_s = 0.f, y_u = 0.f;
    int nb_u = 0, ny_s = 0, nb_s = 0, ny_u = 0;
    for (int k = 0; k < Lk; ++k) {
      b_s = b_s + is_[0][k]; nb_s += ns[0][k];
      y_u = y_u + iu[1][k]; ny_u += nu[1][k];
    }
    double best_hs = -1e300;
    for (int st = 0; st < L; ++st) {
      if (st > 0 && st - 1 < Lk) {
        int k = st - 1;
        b_u = b_u + iu[0][k]; nb_u += nu[0][k]; b_s = b_s - is_[0][k]; nb_s -= ns[0][k];
        y_s = y_s + is_[1][k]; ny_s += ns[1][k]; y_u = y_u - iu[1][k]; ny_u -= nu[1][k];
      }
      float sb2 = b_u + b_s, sy2 = y_u + y_s;
      int mb2 = nb_u + nb_s, my2 = ny_u + ny_s;
      double h2 = log((double)(fmaxf(sb2, 0.f) + 1.0f) * (double)(fmaxf(sy2, 0.f) + 1.0f)) + lnfact(mb2) +
                  lnfact(my2);
      if (h2 > best_hs) { best_hs = h2; mb = mb2; my = my2; sb = fmaxf(sb2, 0.f); sy = fmaxf(sy2, 0.f); }
    }
  }
  double i2 = (double)(sb + 1.0f) * (double)(sy + 1.0f);
  double hs = log(i2) + lnfact(mb) + lnfact(my);
  if (!std::isfinite(hs)) hs = 255.0;
  K2Out o;
  o.hs = hs; o.sb = sb; o.sy = sy; o.ppm = ppm / (sb + sy);
  o.mb = mb; o.my = my; o.lb = rb.longest; o.ly = ry.longest;
  return o;
}

// ------------------------------------------------------------------------------------------------ ranking
// candidates [0, n) of one group: indices with mb+my >= min_matched ordered by (hs desc, row asc)
void rank_one(const VhParams& p, int64_t start, int n, const int32_t* c_row, const double* hs, const int16_t* mb,
              const int16_t* my, int report, int64_t* o_sel, double* o_next) {
  std::vector<int64_t> ok;
  for (int i = 0; i < n; ++i) {
    int64_t c = start + i;
    if ((int)mb[c] + (int)my[c] >= p.min_matched_peaks) ok.push_back(c);
  }
  std::stable_sort(ok.begin(), ok.end(), [&](int64_t a, int64_t b) {
    if (hs[a] != hs[b]) return hs[a] > hs[b];
    return c_row[a] < c_row[b];
  });
  for (int i = 0; i < report; ++i) {
    o_sel[i] = i < (int)ok.size() ? ok[i] : -1;
    o_next[i] = i + 1 < (int)ok.size() ? hs[ok[i + 1]] : 0.0;
  }
}

// ---- feature mode of K1 (vh_k1_features): integer intensity quantisation and its 14-bit pseudo-log2
int iq_shift(float maxI) {
  if (!(maxI >= 1.17549435e-38f) || maxI > 3.40282347e38f) return -1000;
  int x;
  frexpf(maxI, &x);
  return 16 - x;
}

uint32_t iq_quant(float I, int e) {
  if (e == -1000 || !(I >= 1.17549435e-38f)) return 0u;
  const float v = ldexpf(I, e);
  return v >= 1.f ? (uint32_t)v : 0u;
}

uint32_t iq_log(uint64_t sum) {
  const uint64_t x = sum + 1ULL;
  const int e = 63 - __builtin_clzll(x);
  const uint32_t m = e >= 8 ? (uint32_t)((x >> (e - 8)) & 0xFFULL) : (uint32_t)((x << (8 - e)) & 0xFFULL);
  return ((uint32_t)e << 8) | m;
}

// per-ion (b = 0, y = 1) counts and intensity sums of one row: the peaks matching fragment f at charge z are the
// contiguous index range [lower_bound(hi bounds, f), upper_bound(lo bounds, f)) of the mass-sorted peaks
void k1_feat_row(const Ctx& c, const VhIndexView& ix, int64_t r, int maxfz, const PeakBounds& pb,
                 const std::vector<uint64_t>& qpre, int cnt_out[2], uint64_t isum_out[2]) {
  const VhParams& p = *c.p;
  const uint8_t* sq = ix.residues + ix.seq_off[r];
  int L = ix.length[r];
  uint64_t mk = ix.mask[r];
  float b = c.nterm_of(mk), y = ix.mass32[r] - c.nterm_of(mk);
  int cnt[2] = {0, 0};
  uint64_t isum[2] = {0, 0};
  auto add = [&](float f, int ion) {
    for (int z = 1; z < maxfz; ++z) {
      const float* A = pb.a.data() + (size_t)(z - 1) * pb.np;
      const float* Bv = pb.b.data() + (size_t)(z - 1) * pb.np;
      int ua = (int)(std::upper_bound(A, A + pb.np, f) - A);
      int lb = (int)(std::lower_bound(Bv, Bv + pb.np, f) - Bv);
      if (ua > lb) {
        cnt[ion] += ua - lb;
        isum[ion] += qpre[ua] - qpre[lb];
      }
    }
  };
  for (int k = 0; k < L - 1; ++k) {
    float rm = c.res(sq[k], k, mk);
    b = b + rm;
    y = y - rm;
    if (k + 1 > p.min_ion_index && b >= p.fmin && b <= p.fmax) add(b, 0);
    if (L - 1 - k > p.min_ion_index && y >= p.fmin && y <= p.fmax) add(y, 1);
  }
  for (int t = 0; t < 2; ++t) {
    cnt_out[t] = cnt[t];
    isum_out[t] = isum[t];
  }
}

}  // namespace

extern "C" {

int or_version(void) { return VH_ABI_VERSION; }

// ------------------------------------------------------------------------------------------------ K0
int or_k0(const VhParams* p, int32_t n_spec, const float* mz, const float* it, const int64_t* raw_off,
          int64_t* pk_off, float* pk_mass, float* pk_int, float* tic, uint8_t* valid, int64_t cap) {
  Ctx c(p);
  std::vector<int64_t> cnt(n_spec);
  if (p->peak_window_k < 0 || (p->peak_window_k > 0 && !(p->peak_window_th > 0.f))) return VH_E_ARG;
#pragma omp parallel for schedule(dynamic, 256)
  for (int32_t s = 0; s < n_spec; ++s) {
    if (p->peak_window_k > 0) {   // kept count of the window view: sum over windows of min(in window, k)
      std::vector<int64_t> w;
      for (int64_t i = raw_off[s]; i < raw_off[s + 1]; ++i) {
        float m = mz[i];
        if (!isnan(m) && m >= p->fmin && m <= p->fmax) w.push_back((int64_t)floorf(m / p->peak_window_th));
      }
      std::sort(w.begin(), w.end());
      int64_t c = 0;
      for (size_t j = 0, first = 0; j < w.size(); ++j) {
        if (j > 0 && w[j] != w[j - 1]) first = j;
        c += (int64_t)(j - first) < p->peak_window_k ? 1 : 0;
      }
      cnt[s] = std::min<int64_t>(c, p->max_peaks);
      continue;
    }
    int64_t n_in = 0;
    for (int64_t i = raw_off[s]; i < raw_off[s + 1]; ++i) {
      float m = mz[i];
      if (!isnan(m) && m >= p->fmin && m <= p->fmax) ++n_in;
    }
    cnt[s] = std::min<int64_t>(n_in, p->max_peaks);
  }
  pk_off[0] = 0;
  for (int32_t s = 0; s < n_spec; ++s) pk_off[s + 1] = pk_off[s] + cnt[s];
  if (pk_off[n_spec] > cap) return VH_E_CAPACITY;
#pragma omp parallel for schedule(dynamic, 256)
  for (int32_t s = 0; s < n_spec; ++s) {
    double t = 0.0;
    int k = k0_one(c, mz + raw_off[s], it + raw_off[s], raw_off[s + 1] - raw_off[s], pk_mass + pk_off[s],
                   pk_int + pk_off[s], &t);
    tic[s] = (float)t;
    valid[s] = k >= p->min_peaks;
  }
  return 0;
}

// ------------------------------------------------------------------------------------------------ hypotheses
int or_hypotheses(const VhParams* p, int32_t n_spec, const double* pmz, const int8_t* charge, const uint8_t* valid,
                  const float* mass32, int64_t n_rows, int64_t* hyp_off, int32_t* h_spec, int8_t* h_z,
                  int8_t* h_iso, float* h_pm, float* h_M, int64_t* h_lo, int64_t* h_hi, int8_t* h_maxfz,
                  uint8_t* h_mode, float* h_off, int64_t cap) {
  if ((p->n_offsets || p->open_search) && (!h_mode || !h_off)) return VH_E_ARG;
  hyp_off[0] = 0;
  for (int32_t s = 0; s < n_spec; ++s) hyp_off[s + 1] = hyp_off[s] + n_hyp_of(*p, pmz[s], charge[s], valid[s]);
  if (hyp_off[n_spec] > cap) return VH_E_CAPACITY;
#pragma omp parallel for schedule(dynamic, 256)
  for (int32_t s = 0; s < n_spec; ++s)
    if (hyp_off[s + 1] > hyp_off[s])
      hyp_one(*p, s, pmz[s], charge[s], mass32, n_rows, hyp_off[s], h_spec, h_z, h_iso, h_pm, h_M, h_lo, h_hi,
              h_maxfz, h_mode, h_off);
  return 0;
}

// ------------------------------------------------------------------------------------------------ K1
// counts[win_off[h] + (r - h_lo[h])] for every row r of every hypothesis window; scored/matched per hyp.
int or_k1_counts(const VhParams* p, const VhIndexView* ix, int32_t n_spec, const int64_t* pk_off,
                 const float* pk_mass, const int64_t* hyp_off, const int8_t* h_maxfz, const int64_t* h_lo,
                 const int64_t* h_hi, const int64_t* win_off, int32_t* counts, uint32_t* scored, uint64_t* matched) {
  Ctx c(p);
#pragma omp parallel for schedule(dynamic, 16)
  for (int32_t s = 0; s < n_spec; ++s) {
    int64_t h0 = hyp_off[s], h1 = hyp_off[s + 1];
    if (h0 == h1) continue;
    int zmax = 2;
    for (int64_t h = h0; h < h1; ++h) zmax = std::max(zmax, (int)h_maxfz[h]);
    PeakBounds pb;
    pb.build(c, pk_mass + pk_off[s], (int)(pk_off[s + 1] - pk_off[s]), zmax);
    for (int64_t h = h0; h < h1; ++h) {
      uint32_t sc = 0;
      uint64_t mt = 0;
      for (int64_t r = h_lo[h]; r < h_hi[h]; ++r) {
        int k = k1_count_row(c, *ix, r, h_maxfz[h], pb);
        counts[win_off[h] + (r - h_lo[h])] = k;
        sc += k > 0;
        mt += (uint64_t)k;
      }
      scored[h] = sc;
      matched[h] = mt;
    }
  }
  return 0;
}

int or_k1_brute(const VhParams* p, const VhIndexView* ix, int64_t n, const int32_t* c_hyp, const int32_t* c_row,
                const int32_t* h_spec, const int8_t* h_maxfz, const int64_t* pk_off, const float* pk_mass,
                int32_t* out_cnt) {
  Ctx c(p);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    int h = c_hyp[i];
    int s = h_spec[h];
    out_cnt[i] = k1_brute_pair(c, *ix, c_row[i], h_maxfz[h], pk_mass + pk_off[s], pk_off[s + 1] - pk_off[s]);
  }
  return 0;
}

// ------------------------------------------------------------------------------------------------ top-K
int or_topk(const VhParams* p, const int64_t* partner, int32_t n_spec, const int64_t* hyp_off, const uint8_t* h_mode,
            const int64_t* h_lo, const int64_t* h_hi, const int64_t* win_off, const int32_t* counts, int32_t stride,
            int32_t* n_cand, int32_t* c_row, int32_t* c_hyp, int32_t* c_cnt) {
  int need = p->pair_closure ? 2 * p->topk : p->topk;
  if (stride < need) return VH_E_CAPACITY;
#pragma omp parallel for schedule(dynamic, 64)
  for (int32_t s = 0; s < n_spec; ++s) {
    std::vector<Ent> e;
    int64_t h0 = hyp_off[s];
    for (int64_t h = h0; h < hyp_off[s + 1]; ++h) {
      int thr = (h_mode && h_mode[h] == VH_MODE_OPEN) ? p->k1_threshold_open : p->k1_threshold;
      for (int64_t r = h_lo[h]; r < h_hi[h]; ++r) {
        int k = counts[win_off[h] + (r - h_lo[h])];
        if (k >= thr) e.push_back({r, (int32_t)(h - h0), k});
      }
    }
    std::vector<Ent> out(2 * e.size() + 1);
    int n = topk_one(*p, partner, e, out.data());
    n_cand[s] = n;
    for (int i = 0; i < n; ++i) {
      int64_t o = (int64_t)s * stride + i;
      c_row[o] = (int32_t)out[i].row;
      c_hyp[o] = (int32_t)(h0 + out[i].hl);
      c_cnt[o] = out[i].cnt;
    }
  }
  return 0;
}

// vh_k1_features: per spectrum the top-K (count clamped at 32767 desc, row desc, lower hypothesis first) of the
// entries with count >= threshold, deduplicated by row (lowest hypothesis; features of the entry with the larger
// count, ties: lower hypothesis), row order; c_nb = b-ion count, c_iq = iq_log(b sum) << 16 | iq_log(y sum),
// s_maxcnt = best count.
int or_topk_feat(const VhParams* p, const VhIndexView* ix, int32_t n_spec, const int64_t* pk_off,
                 const float* pk_mass, const float* pk_int, const int64_t* hyp_off, const int8_t* h_maxfz,
                 const uint8_t* h_mode, const int64_t* h_lo, const int64_t* h_hi, int32_t stride, int32_t* n_cand,
                 int32_t* c_row, int32_t* c_hyp, int32_t* c_cnt, int16_t* c_nb, uint32_t* c_iq,
                 int32_t* s_maxcnt) {
  if (stride < p->topk) return VH_E_CAPACITY;
  Ctx c(p);
#pragma omp parallel for schedule(dynamic, 16)
  for (int32_t s = 0; s < n_spec; ++s) {
    const int64_t h0 = hyp_off[s], h1 = hyp_off[s + 1];
    n_cand[s] = 0;
    s_maxcnt[s] = 0;
    if (h0 == h1) continue;
    const int np = (int)(pk_off[s + 1] - pk_off[s]);
    const float* pm = pk_mass + pk_off[s];
    const float* pi = pk_int + pk_off[s];
    float mx = 0.f;
    for (int q = 0; q < np; ++q) mx = fmaxf(mx, pi[q]);
    const int e = iq_shift(mx);
    std::vector<uint64_t> qpre(np + 1, 0);
    for (int q = 0; q < np; ++q) qpre[q + 1] = qpre[q] + iq_quant(pi[q], e);
    int zmax = 2;
    for (int64_t h = h0; h < h1; ++h) zmax = std::max(zmax, (int)h_maxfz[h]);
    PeakBounds pb;
    pb.build(c, pm, np, zmax);
    struct FE { int64_t row; int32_t hl; int32_t cnt; int32_t nb; uint32_t iq; };
    std::vector<FE> ents;
    for (int64_t h = h0; h < h1; ++h) {
      int thr = (h_mode && h_mode[h] == VH_MODE_OPEN) ? p->k1_threshold_open : p->k1_threshold;
      for (int64_t r = h_lo[h]; r < h_hi[h]; ++r) {
        int k[2];
        uint64_t is[2];
        k1_feat_row(c, *ix, r, h_maxfz[h], pb, qpre, k, is);
        const int kk = k[0] + k[1];
        if (kk >= thr)
          ents.push_back({r, (int32_t)(h - h0), std::min(kk, 32767), std::min(k[0], 32767),
                          (iq_log(is[0]) << 16) | iq_log(is[1])});
      }
    }
    std::sort(ents.begin(), ents.end(), [](const FE& a, const FE& b) {
      if (a.cnt != b.cnt) return a.cnt > b.cnt;
      if (a.row != b.row) return a.row > b.row;
      return a.hl < b.hl;
    });
    size_t k = std::min<size_t>(ents.size(), (size_t)std::max(p->topk, 0));
    if (k) s_maxcnt[s] = ents[0].cnt;
    std::vector<FE> sel(ents.begin(), ents.begin() + k);
    std::sort(sel.begin(), sel.end(), [](const FE& a, const FE& b) {
      if (a.row != b.row) return a.row < b.row;
      return a.hl < b.hl;
    });
    int n = 0;
    for (size_t i = 0; i < sel.size(); ++i) {
      const int64_t o = (int64_t)s * stride + n;
      if (n && c_row[o - 1] == (int32_t)sel[i].row) {
        if (sel[i].cnt > c_cnt[o - 1]) {
          c_cnt[o - 1] = sel[i].cnt;
          c_nb[o - 1] = (int16_t)sel[i].nb;
          c_iq[o - 1] = sel[i].iq;
        }
        continue;
      }
      c_row[o] = (int32_t)sel[i].row;
      c_hyp[o] = (int32_t)(h0 + sel[i].hl);
      c_cnt[o] = sel[i].cnt;
      c_nb[o] = (int16_t)sel[i].nb;
      c_iq[o] = sel[i].iq;
      ++n;
    }
    n_cand[s] = n;
  }
  return 0;
}

extern "C" uint32_t or_iq_log(uint64_t sum) { return iq_log(sum); }
extern "C" uint32_t or_iq_quant(float I, float maxI) { return iq_quant(I, iq_shift(maxI)); }

// ------------------------------------------------------------------------------------------------ K2
int or_k2(const VhParams* p, const VhIndexView* ix, int32_t n_groups, const int32_t* grp_spec,
          const int64_t* grp_start, const int32_t* grp_count, int32_t stride, const int32_t* c_row,
          const int32_t* c_hyp, const int8_t* h_maxfz, const float* c_shift, const int64_t* pk_off,
          const float* pk_mass, const float* pk_int, double* o_hs, float* o_sb, float* o_sy, float* o_ppm,
          int16_t* o_mb, int16_t* o_my, int16_t* o_lb, int16_t* o_ly) {
  Ctx c(p);
#pragma omp parallel for schedule(dynamic, 16)
  for (int32_t g = 0; g < n_groups; ++g) {
    int s = grp_spec ? grp_spec[g] : g;
    int64_t c0 = grp_start ? grp_start[g] : (int64_t)g * stride;
    int nc = grp_count[g];
    const float* sm = pk_mass + pk_off[s];
    const float* si = pk_int + pk_off[s];
    int np = (int)(pk_off[s + 1] - pk_off[s]);
    for (int64_t q = c0; q < c0 + nc; ++q) {
      K2Out o = k2_one(c, *ix, c_row[q], h_maxfz[c_hyp[q]], sm, si, np, c_shift, c_shift ? c_shift[q] : 0.f);
      o_hs[q] = o.hs; o_sb[q] = o.sb; o_sy[q] = o.sy; o_ppm[q] = o.ppm;
      o_mb[q] = (int16_t)o.mb; o_my[q] = (int16_t)o.my; o_lb[q] = (int16_t)o.lb; o_ly[q] = (int16_t)o.ly;
    }
  }
  return 0;
}

double or_lnfact(int n) { return lnfact(n); }

// ------------------------------------------------------------------------------------------------ ranking
int or_rank(const VhParams* p, int32_t n_groups, const int64_t* grp_start, const int32_t* grp_count, int32_t stride,
            const int32_t* c_row, const double* hs, const int16_t* mb, const int16_t* my, int32_t report,
            int64_t* o_sel, double* o_next) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int32_t g = 0; g < n_groups; ++g) {
    int64_t c0 = grp_start ? grp_start[g] : (int64_t)g * stride;
    rank_one(*p, c0, grp_count[g], c_row, hs, mb, my, report, o_sel + (int64_t)g * report,
             o_next + (int64_t)g * report);
  }
  return 0;
}

// ------------------------------------------------------------------------------------------------ search
// Whole closed search per spectrum (engine.py:435-464 without the feature columns): K0 -> hypotheses ->
// K1 counts -> top-K + closure -> K2 -> rank 1. OpenMP over spectra; this is the CPU baseline.
// Outputs per spectrum: best row (-1 if none), its hyperscore, delta_next, the number of rescored candidates;
// stats[0] = total rescored candidates, stats[1] = total K1 window rows, stats[2] = total matched (sum counts).
int or_search(const VhParams* p, const VhIndexView* ix, int32_t n_spec, const float* mz, const float* it,
              const int64_t* raw_off, const double* pmz, const int8_t* charge, int32_t* o_row, double* o_hs,
              double* o_dn, int32_t* o_ncand, int64_t* stats) {
  if (p->n_offsets || p->open_search) return VH_E_ARG;   // closed search only (modes: component functions)
  Ctx c(p);
  int64_t st_resc = 0, st_win = 0, st_match = 0;
  int niso = p->iso_hi - p->iso_lo + 1;
  int maxh = niso * std::max(1, p->z_hi - p->z_lo + 1);
#pragma omp parallel reduction(+ : st_resc, st_win, st_match)
  {
    std::vector<float> pm(p->max_peaks + 1), pi(p->max_peaks + 1);
    std::vector<int32_t> hs_spec(maxh);
    std::vector<int8_t> hz(maxh), hiso(maxh), hmfz(maxh);
    std::vector<float> hpm(maxh), hM(maxh);
    std::vector<int64_t> hlo(maxh), hhi(maxh);
    std::vector<Ent> ent, out;
    PeakBounds pb;
#pragma omp for schedule(dynamic, 32)
    for (int32_t s = 0; s < n_spec; ++s) {
      o_row[s] = -1; o_hs[s] = 0.0; o_dn[s] = 0.0; o_ncand[s] = 0;
      double tic;
      int np = k0_one(c, mz + raw_off[s], it + raw_off[s], raw_off[s + 1] - raw_off[s], pm.data(), pi.data(), &tic);
      int nh = n_hyp_of(*p, pmz[s], charge[s], np >= p->min_peaks);
      if (!nh) continue;
      hyp_one(*p, s, pmz[s], charge[s], ix->mass32, ix->n_rows, 0, hs_spec.data(), hz.data(), hiso.data(),
              hpm.data(), hM.data(), hlo.data(), hhi.data(), hmfz.data(), nullptr, nullptr);
      int zmax = 2;
      for (int h = 0; h < nh; ++h) zmax = std::max(zmax, (int)hmfz[h]);
      pb.build(c, pm.data(), np, zmax);
      ent.clear();
      for (int h = 0; h < nh; ++h) {
        st_win += hhi[h] - hlo[h];
        for (int64_t r = hlo[h]; r < hhi[h]; ++r) {
          int k = k1_count_row(c, *ix, r, hmfz[h], pb);
          st_match += k;
          if (k >= p->k1_threshold) ent.push_back({r, h, k});
        }
      }
      out.resize(2 * ent.size() + 1);
      int n = topk_one(*p, ix->partner, ent, out.data());
      o_ncand[s] = n;
      st_resc += n;
      double best = -1e300, second = 0.0;
      int64_t brow = -1;
      bool have2 = false;
      for (int i = 0; i < n; ++i) {
        K2Out o = k2_one(c, *ix, out[i].row, hmfz[out[i].hl], pm.data(), pi.data(), np, nullptr, 0.f);
        if (o.mb + o.my < p->min_matched_peaks) continue;
        bool better = brow < 0 || o.hs > best || (o.hs == best && out[i].row < brow);
        if (better) {
          if (brow >= 0) { second = best; have2 = true; }
          best = o.hs; brow = out[i].row;
        } else if (!have2 || o.hs > second) {
          second = o.hs; have2 = true;
        }
      }
      if (brow >= 0) {
        o_row[s] = (int32_t)brow;
        o_hs[s] = best;
        o_dn[s] = best - (have2 ? second : 0.0);
      }
    }
  }
  stats[0] = st_resc; stats[1] = st_win; stats[2] = st_match;
  return 0;
}

}  // extern "C"

extern "C" size_t or_abi_sizeof(int which) { return which == 0 ? sizeof(VhParams) : sizeof(VhIndexView); }

#ifdef _OPENMP
#include <omp.h>
extern "C" int or_max_threads(void) { return omp_get_max_threads(); }
#else
extern "C" int or_max_threads(void) { return 1; }
#endif
