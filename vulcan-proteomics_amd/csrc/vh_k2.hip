// K2 exact b/y rescoring and PSM ranking.
//
// K2 restates k2_kernel (vulcan/gpu/kernels.cu:237-376): for every candidate, all b_k / y_k (k = 0..L-2, no
// index filter) at fragment charges 1..maxfz-1, window [tol_bound(f/z, lo), tol_bound(f/z, hi)] over the
// spectrum's mass-sorted peaks (first 512), the most intense peak wins (ties -> later peak, `>=`);
// ppm += I*|mz - m|*2e6/(mz + m); b ions first (k asc, z asc), then y ions; longest consecutive ladder with the
// reference's run bookkeeping; hyperscore = ln(f64(sb+1) * f64(sy+1)) + lnfact(mb) + lnfact(my), 255 if not
// finite. One wave64 per spectrum group: peaks are staged in LDS once with a coarse lookup table (cell c of
// width w holds the first peak whose cell index is >= c; the cell index is a monotone function of the mass, so
// the table entry never passes the first peak >= the window start and a short forward skip finds it), lanes take
// candidates. The log-factorial (Stirling, lnfact(0) = 1) comes from an f64 LDS table built once per block.
// Ranking restates psm_table's order (engine.py:366-385): mb + my >= min_matched_peaks, hyperscore desc,
// row asc (then candidate index, the stable-sort tie order).
#include "vh_common.h"

// waves per workgroup: two measured faster than four and than one (cfg1 K2 0.839 / 0.830 -> 0.813 ms)
#define K2_WAVES 2
#define K2_LNF 256
#define K2_CELLS 1024   // coarse peak lookup table per spectrum (u16 first-peak index per mass cell)

__device__ __forceinline__ double lnfact_dev(int n) {   // kernels.cu:246-250
  if (n == 0) return 1.0;
  double x = (double)n;
  return x * log(x) - x + 0.5 * log(x) + 0.5 * log(2.0 * 3.141592653589793 * x);
}

struct Run {
  int start, length, last, longest;
};

__device__ __forceinline__ void run_matched(Run& r, int idx) {   // kernels.cu:252-258
  if (r.last == idx) return;
  if (r.start + r.length == idx) {
    r.length += 1;
    r.longest = r.longest > r.length ? r.longest : r.length;
  } else {
    r.start = idx;
    r.length = 1;
    r.longest = r.longest > 1 ? r.longest : 1;
  }
  r.last = idx;
}

struct K2Buf {
  double* hs;
  float *sb, *sy, *ppm;
  int16_t *mb, *my, *lb, *ly;
  uint32_t* queue;   // persistent mode work counter (NULL: grid mode)
};

// Lookup cell of mass x: clamp((x - base) * inv_w, 0, K2_CELLS - 1) truncated (= floor on the clamped range;
// NaN goes to cell 0). Monotone in x, so the table built with it never passes the first peak >= x.
__device__ __forceinline__ int k2_cell(float x, float base, float inv_w) {
  return (int)fminf(fmaxf((x - base) * inv_w, 0.f), (float)(K2_CELLS - 1));
}

// f / z for a fragment charge z: exact for z = 1, 2 (a power-of-two scale rounds like the division)
__device__ __forceinline__ float div_charge(float f, int z) {
  if (z == 1) return f;
  if (z == 2) return f * 0.5f;
  return f / (float)z;
}

// Mass-sorted peaks of one spectrum in LDS as (mass, intensity) pairs, pk[np] = (NaN, 0) ends every scan, with
// the coarse lookup table (tb[c] = first peak whose cell is >= c).
struct PeakView {
  const float2* pk;
  const uint16_t* tb;
  int np;
  float base, inv_w;

  __device__ __forceinline__ int cell_start(float a) const { return tb[k2_cell(a, base, inv_w)]; }
};

// One tolerance window at m/z `mz`: the most intense peak (ties -> later peak) updates the running (mx, best)
// shared across the unshifted/shifted windows, with its mass best_x, and the window's own (mx_a, best_a)
// (kernels.cu:312-329). mx / mx_a are the best peaks' intensities, so callers need not read them again.
__device__ __forceinline__ void window_scan(const PeakView& V, const DevP& P, float mz, float& mx, int& best,
                                            float& best_mz, float& best_x, float& mx_a, int& best_a) {
  const float a = tol_bound(mz, P.tol_lo, P.tol_da);
  const float bd = tol_bound(mz, P.tol_hi, P.tol_da);
  for (int j = V.cell_start(a);; ++j) {
    const float2 q = V.pk[j];
    if (!(q.x <= bd)) break;
    if (q.x < a) continue;
    if (q.y >= mx) { mx = q.y; best = j; best_mz = mz; best_x = q.x; }
    if (q.y >= mx_a) { mx_a = q.y; best_a = j; }
  }
}

// Per-ion unshifted (u) and shifted (s) matched intensity and counts over the fragment charges (the iu/nu/is/ns
// arrays of kernels.cu:300-305, one ion at a time).
__device__ __forceinline__ void ion_alt_sums(const PeakView& V, const DevP& P, float cum, float shift, int maxfz,
                                             float& iu, int& nu, float& is, int& ns) {
  iu = 0.f; is = 0.f; nu = 0; ns = 0;
  for (int z = 1; z < maxfz; ++z) {
    float mx = 0.f, bmz = 0.f, bx = 0.f;
    int best = -1;
    for (int alt = 0; alt < 2; ++alt) {
      float mx_a = 0.f;
      int best_a = -1;
      const float base = alt == 0 ? cum : cum + shift;
      window_scan(V, P, div_charge(base, z), mx, best, bmz, bx, mx_a, best_a);
      if (best_a >= 0) {
        if (alt == 0) { iu = iu + mx_a; nu += 1; }
        else { is = is + mx_a; ns += 1; }
      }
    }
  }
}

template <bool SHIFT>
__global__ __launch_bounds__(64 * K2_WAVES) void k2_kernel(DevP P, DevIdx X, int n_groups, const int32_t* __restrict__ order,
                                                 const int32_t* __restrict__ grp_spec,
                                                 const int64_t* __restrict__ grp_start,
                                                 const int32_t* __restrict__ grp_count, int stride,
                                                 const int32_t* __restrict__ c_row, const int32_t* __restrict__ c_hyp,
                                                 const int8_t* __restrict__ h_maxfz,
                                                 const float* __restrict__ c_shift,
                                                 const int64_t* __restrict__ pk_off,
                                                 const float* __restrict__ pk_mass,
                                                 const float* __restrict__ pk_int, K2Buf O) {
  __shared__ float rm[52];
  __shared__ double lnf[K2_LNF];
  __shared__ float2 s_p[K2_WAVES][VH_K2_MAX_PEAKS + 1];
  __shared__ uint16_t s_t[K2_WAVES][K2_CELLS];
  load_rm(rm, P);
  for (int i = threadIdx.x; i < K2_LNF; i += blockDim.x) lnf[i] = lnfact_dev(i);
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = lane_id();
  // grid mode: one group per wave (XCD-banded); persistent mode: groups from the atomic counter until exhausted
  int wv = O.queue ? 0 : xcd_band_position(P.xcd_group) * K2_WAVES + w;
  for (;;) {
  if (O.queue) {
    if (lane == 0) wv = (int)atomicAdd(O.queue, 1u);
    wv = __shfl(wv, 0);
  }
  if (wv >= n_groups) return;   // every wave of a persistent grid reaches this once the counter passes n_groups
  const int g = order ? order[wv] : wv;
  const int s = grp_spec ? grp_spec[g] : g;
  const int64_t c0 = grp_start ? grp_start[g] : (int64_t)g * stride;
  const int nc = grp_count ? grp_count[g] : (int)(grp_start[g + 1] - c0);
  const int64_t q0 = pk_off[s];
  int np = (int)(pk_off[s + 1] - q0);
  np = np < VH_K2_MAX_PEAKS ? np : VH_K2_MAX_PEAKS;
  float2* pk = s_p[w];
  uint16_t* tb = s_t[w];
  // the peaks' chunk loads in flight before their stores: all of them at once in the closed kernel, four chunks
  // at a time in the shifted one (short of registers); pk[np] = (NaN, 0) ends every scan
  if constexpr (!SHIFT) {
    constexpr int PC = (VH_K2_MAX_PEAKS + 63) / 64;
    float2 v[PC];
#pragma unroll
    for (int u = 0; u < PC; ++u) {
      const int j = u * 64 + lane;
      v[u] = j < np ? make_float2(pk_mass[q0 + j], pk_int[q0 + j]) : make_float2(__builtin_nanf(""), 0.f);
    }
#pragma unroll
    for (int u = 0; u < PC; ++u) {
      const int j = u * 64 + lane;
      if (j <= np) pk[j] = v[u];
    }
    if (np == PC * 64 && lane == 0) pk[np] = make_float2(__builtin_nanf(""), 0.f);
  } else {
    for (int j0 = 0; j0 <= np; j0 += 4 * 64) {
      float2 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = j0 + u * 64 + lane;
        v[u] = j < np ? make_float2(pk_mass[q0 + j], pk_int[q0 + j]) : make_float2(__builtin_nanf(""), 0.f);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = j0 + u * 64 + lane;
        if (j <= np) pk[j] = v[u];
      }
    }
  }
  wave_sync();
  // coarse table over [first peak, last peak]
  const float base = np ? pk[0].x : 0.f;
  const float span = np ? pk[np - 1].x - base : 0.f;
  const float inv_w = span > 0.f ? (float)(K2_CELLS - 2) / span : 1.f;
  for (int j = lane; j <= np; j += 64) {
    const int cj = j < np ? k2_cell(pk[j].x, base, inv_w) : K2_CELLS;   // j == np: past the last cell
    const int cp = j > 0 ? k2_cell(pk[j - 1].x, base, inv_w) : -1;
    for (int c = cp + 1; c <= cj && c < K2_CELLS; ++c) tb[c] = (uint16_t)j;
  }
  wave_sync();
  const PeakView V = {pk, tb, np, base, inv_w};
  for (int64_t c = c0 + lane; c < c0 + nc; c += 64) {
    const int r = c_row[c];
    const int maxfz = h_maxfz[c_hyp[c]];
    const uint8_t* sq = X.residues + X.seq_off[r];
    const int L = X.length[r];
    const uint64_t mk = X.mask[r];
    const float nt = (mk >> 63) ? P.nterm : 0.f;
    const float shift = SHIFT ? c_shift[c] : 0.f;
    const int n_alt = (SHIFT && shift != 0.f) ? 2 : 1;
    float sb = 0.f, sy = 0.f, ppm = 0.f;
    int mb = 0, my = 0;
    Run rb = {0, 0, 0, 0}, ry = {0, 0, 0, 0};
    float bs_init = 0.f, yu_init = 0.f;   // sum_k is[0][k] and sum_k iu[1][k] (k ascending)
    int nbs_init = 0, nyu_init = 0;
    // one fragment window [a, bd] at m/z mz, its scan starting at peak j (= q): the best peak is matched (ppm,
    // count, intensity, ladder run); its mass and intensity are kept from the scan (no second LDS read)
    auto window = [&](float mz, float a, float bd, int j, float2 q, int kind, int k) {
      float mx = 0.f, bx = 0.f;
      bool hit = false;
      while (q.x <= bd) {   // the NaN sentinel ends every scan
        if (q.x >= a && q.y >= mx) { mx = q.y; bx = q.x; hit = true; }
        q = pk[++j];
      }
      if (hit) {
        const float num = (mx * fabsf(mz - bx)) * 2.0e6f;
        ppm = ppm + num / (mz + bx);
        if (kind == 0) { mb += 1; sb = sb + mx; run_matched(rb, k); }
        else { my += 1; sy = sy + mx; run_matched(ry, k); }
      }
    };
    auto match = [&](float mz, int kind, int k) {
      const float a = tol_bound(mz, P.tol_lo, P.tol_da), bd = tol_bound(mz, P.tol_hi, P.tol_da);
      const int j = V.cell_start(a);
      window(mz, a, bd, j, pk[j], kind, k);
    };
    for (int kind = 0; kind < 2; ++kind) {
      float cum = kind == 0 ? nt : X.mass32[r] - nt;
      int nres = L > 1 ? sq[0] : 'A';   // residue k + 1 is loaded while fragment k is matched
      for (int k = 0; k < L - 1; ++k) {
        const int res = nres;
        if (k + 2 < L) nres = sq[k + 1];
        const float rmv = rm[((mk >> k) & 1ULL) * 26 + (res - 'A')];
        cum = kind == 0 ? cum + rmv : cum - rmv;
        if (!SHIFT || n_alt == 1) {   // charges ascending: z = 1, 2 without a division
          const float m2 = cum * 0.5f;
          const float a1 = tol_bound(cum, P.tol_lo, P.tol_da), b1 = tol_bound(cum, P.tol_hi, P.tol_da);
          const float a2 = tol_bound(m2, P.tol_lo, P.tol_da), b2 = tol_bound(m2, P.tol_hi, P.tol_da);
          // both charges' lookups (cell start, first peak) are issued before either window is scanned; a charge
          // at or above maxfz gets an empty window (high bound -inf: its scan stops at the first peak), so the
          // lookups take no branch on maxfz and the two chains of LDS reads overlap
          const int j1 = V.cell_start(a1), j2 = V.cell_start(a2);
          const float2 q1 = pk[j1], q2 = pk[j2];
          window(cum, a1, maxfz > 1 ? b1 : -INFINITY, j1, q1, kind, k);
          window(m2, a2, maxfz > 2 ? b2 : -INFINITY, j2, q2, kind, k);
          for (int z = 3; z < maxfz; ++z) match(cum / (float)z, kind, k);
          continue;
        }
        float iu = 0.f, is = 0.f;
        int nu = 0, ns = 0;
        for (int z = 1; z < maxfz; ++z) {
          int best = -1;
          float mx = 0.f, mz = 0.f, bx = 0.f;
          for (int alt = 0; alt < n_alt; ++alt) {
            float mx_a = 0.f;
            int best_a = -1;
            const float b = alt == 0 ? cum : cum + shift;
            window_scan(V, P, div_charge(b, z), mx, best, mz, bx, mx_a, best_a);
            if (best_a >= 0) {
              if (alt == 0) { iu = iu + mx_a; nu += 1; }
              else { is = is + mx_a; ns += 1; }
            }
          }
          if (best >= 0) {
            const float pmass = bx, pint = mx;
            const float num = (pint * fabsf(mz - pmass)) * 2.0e6f;
            ppm = ppm + num / (mz + pmass);
            if (kind == 0) { mb += 1; sb = sb + pint; run_matched(rb, k); }
            else { my += 1; sy = sy + pint; run_matched(ry, k); }
          }
        }
        if (SHIFT && n_alt == 2) {
          if (kind == 0) { bs_init = bs_init + is; nbs_init += ns; }
          else { yu_init = yu_init + iu; nyu_init += nu; }
        }
      }
    }
    if (SHIFT && n_alt == 2) {
      // site st in [0, L): b_k carries the delta iff k >= st, y_k iff k < st (kernels.cu:344-363); the per-ion
      // sums of ion st-1 are recomputed from the same f32 ladder instead of being stored
      float b_u = 0.f, y_s = 0.f, b_s = bs_init, y_u = yu_init;
      int nb_u = 0, ny_s = 0, nb_s = nbs_init, ny_u = nyu_init;
      float bcum = nt, ycum = X.mass32[r] - nt;
      double best_hs = -1e300;
      for (int st = 0; st < L; ++st) {
        if (st > 0) {
          const int k = st - 1;
          const float rmv = rm[((mk >> k) & 1ULL) * 26 + (sq[k] - 'A')];
          bcum = bcum + rmv;
          ycum = ycum - rmv;
          float iu0, is0, iu1, is1;
          int nu0, ns0, nu1, ns1;
          ion_alt_sums(V, P, bcum, shift, maxfz, iu0, nu0, is0, ns0);
          ion_alt_sums(V, P, ycum, shift, maxfz, iu1, nu1, is1, ns1);
          b_u = b_u + iu0; nb_u += nu0; b_s = b_s - is0; nb_s -= ns0;
          y_s = y_s + is1; ny_s += ns1; y_u = y_u - iu1; ny_u -= nu1;
        }
        const float sb2 = b_u + b_s, sy2 = y_u + y_s;
        const int mb2 = nb_u + nb_s, my2 = ny_u + ny_s;
        const double h2 = log((double)(fmaxf(sb2, 0.f) + 1.0f) * (double)(fmaxf(sy2, 0.f) + 1.0f)) +
                          (mb2 < K2_LNF ? lnf[mb2] : lnfact_dev(mb2)) + (my2 < K2_LNF ? lnf[my2] : lnfact_dev(my2));
        if (h2 > best_hs) { best_hs = h2; mb = mb2; my = my2; sb = fmaxf(sb2, 0.f); sy = fmaxf(sy2, 0.f); }
      }
    }
    const double i2 = (double)(sb + 1.0f) * (double)(sy + 1.0f);
    double hs = log(i2) + (mb < K2_LNF ? lnf[mb] : lnfact_dev(mb)) + (my < K2_LNF ? lnf[my] : lnfact_dev(my));
    if (!isfinite(hs)) hs = 255.0;
    O.hs[c] = hs;
    O.sb[c] = sb;
    O.sy[c] = sy;
    O.ppm[c] = ppm / (sb + sy);
    O.mb[c] = (int16_t)mb;
    O.my[c] = (int16_t)my;
    O.lb[c] = (int16_t)rb.longest;
    O.ly[c] = (int16_t)ry.longest;
  }
  if (!O.queue) return;
  wave_sync();   // the next group reloads this wave's LDS peaks
  }
}

// ------------------------------------------------------------------------------------------------ shifts
// c_shift = M(h) - mass32(row) for offset / open hypotheses, 0 for closed ones (vulcan/ptm/search.py:187-189)
__global__ __launch_bounds__(64) void cand_shift_kernel(int n_groups, const int64_t* __restrict__ grp_start,
                                                        const int32_t* __restrict__ grp_count, int stride,
                                                        const int32_t* __restrict__ c_row,
                                                        const int32_t* __restrict__ c_hyp,
                                                        const uint8_t* __restrict__ h_mode,
                                                        const float* __restrict__ h_M,
                                                        const float* __restrict__ mass32, float* c_shift) {
  const int g = blockIdx.x;
  const int64_t c0 = grp_start ? grp_start[g] : (int64_t)g * stride;
  const int nc = grp_count ? grp_count[g] : (int)(grp_start[g + 1] - c0);
  for (int j = threadIdx.x; j < nc; j += blockDim.x) {
    const int64_t c = c0 + j;
    const int h = c_hyp[c];
    c_shift[c] = h_mode[h] == VH_MODE_CLOSED ? 0.f : h_M[h] - mass32[c_row[c]];
  }
}

// ------------------------------------------------------------------------------------------------ ranking
struct RKey {
  double hs;
  int row;
  long long idx;   // -1 = none
};

// a before b in (hyperscore desc, row asc, index asc)
__device__ __forceinline__ bool rk_before(const RKey& a, const RKey& b) {
  if (a.idx < 0) return false;
  if (b.idx < 0) return true;
  if (a.hs != b.hs) return a.hs > b.hs;
  if (a.row != b.row) return a.row < b.row;
  return a.idx < b.idx;
}

__global__ __launch_bounds__(256) void rank_kernel(DevP P, int n_groups, const int64_t* __restrict__ grp_start,
                                                   const int32_t* __restrict__ grp_count, int stride,
                                                   const int32_t* __restrict__ c_row, const double* __restrict__ hs,
                                                   const int16_t* __restrict__ mb, const int16_t* __restrict__ my,
                                                   int report, int64_t* o_sel, double* o_next) {
  const int w = threadIdx.x >> 6, lane = lane_id();
  const int g = blockIdx.x * (blockDim.x >> 6) + w;
  if (g >= n_groups) return;
  const int64_t c0 = grp_start ? grp_start[g] : (int64_t)g * stride;
  const int nc = grp_count ? grp_count[g] : (int)(grp_start[g + 1] - c0);
  RKey prev = {0.0, 0, -2};   // idx -2: nothing taken yet
  for (int i = 0; i <= report; ++i) {
    RKey best = {0.0, 0, -1};
    for (int j = lane; j < nc; j += 64) {
      int64_t c = c0 + j;
      if ((int)mb[c] + (int)my[c] < P.min_matched_peaks) continue;
      RKey k = {hs[c], c_row[c], (long long)c};
      if (prev.idx != -2 && !rk_before(prev, k)) continue;   // only candidates after the previous pick
      if (rk_before(k, best)) best = k;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      RKey x;
      x.hs = __shfl_xor(best.hs, o);
      x.row = __shfl_xor(best.row, o);
      x.idx = __shfl_xor(best.idx, o);
      if (rk_before(x, best)) best = x;
    }
    if (lane == 0) {
      if (i < report) o_sel[(int64_t)g * report + i] = best.idx;
      if (i > 0) o_next[(int64_t)g * report + i - 1] = best.idx >= 0 ? best.hs : 0.0;
    }
    if (best.idx < 0) {
      if (lane == 0)
        for (int k = i + 1; k <= report; ++k) {
          if (k < report) o_sel[(int64_t)g * report + k] = -1;
          o_next[(int64_t)g * report + k - 1] = 0.0;
        }
      break;
    }
    prev = best;
  }
}

__global__ void k2_zero_u32_kernel(uint32_t* p) { *p = 0u; }

// ------------------------------------------------------------------------------------------------ PSM features
// Thread per (group, rank) slot: the psm_table columns of the selected candidate (engine.py:361-432), in the
// f32 / f64 operation order of the reference's numpy expressions.
__global__ __launch_bounds__(256) void psm_kernel(DevIdx X, int n_groups, const int32_t* __restrict__ grp_spec,
                                                  int report, const int64_t* __restrict__ sel,
                                                  const double* __restrict__ next_hs,
                                                  const int32_t* __restrict__ c_row,
                                                  const int32_t* __restrict__ c_hyp,
                                                  const int32_t* __restrict__ c_cnt, VhK2Cols K,
                                                  const int64_t* __restrict__ hyp_off,
                                                  const int8_t* __restrict__ h_z, const int8_t* __restrict__ h_iso,
                                                  const float* __restrict__ h_pm,
                                                  const uint32_t* __restrict__ h_scored,
                                                  const unsigned long long* __restrict__ h_matched,
                                                  const float* __restrict__ tic, VhPsmCols O) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)n_groups * report) return;
  const int64_t g = q / report;
  const int i = (int)(q - g * report);
  const int64_t c = sel[q];
  if (c < 0) {
    O.rank[q] = 0;
    return;
  }
  const int s = grp_spec ? grp_spec[g] : (int)g;
  const int row = c_row[c], h = c_hyp[c];
  const double hs = K.hs[c];
  const int mb = K.mb[c], my = K.my[c], ly = K.ly[c];
  const float sb = K.sb[c], sy = K.sy[c];
  // hypothesis and mass columns (f32, the reference's order of operations)
  const int iso = h_iso[h];
  const float pm = h_pm[h], calc = X.mass32[row];
  const float ie = (float)iso * VH_NEUTRON;
  const float dm = (fabsf((pm - calc) - ie) * 2.0e6f) / ((pm - ie) + calc);
  const int len = X.length[row];
  // per-spectrum K1 statistics summed over its hypotheses
  int64_t sc = 0, mt = 0;
  for (int64_t hh = hyp_off[s]; hh < hyp_off[s + 1]; ++hh) {
    sc += h_scored[hh];
    mt += (int64_t)h_matched[hh];
  }
  const int kk = mb + my;
  const double lam = (double)mt / (double)sc;
  double pois = pow(lam, (double)kk) * exp(-lam) / exp(lnfact_dev(kk));
  if (isinf(pois)) pois = 0.0;   // Sage's 1e-325, which is 0 in f64
  const float summed = sb + sy;
  O.spec_idx[q] = s;
  O.pf_row[q] = row;
  O.rank[q] = i + 1;
  O.charge[q] = h_z[h];
  O.isotope_error[q] = (int8_t)iso;
  O.expmass[q] = pm;
  O.calcmass[q] = calc;
  O.delta_mass_ppm[q] = dm;
  O.hyperscore[q] = hs;
  O.delta_next[q] = hs - next_hs[q];
  O.delta_best[q] = K.hs[sel[g * report]] - hs;
  O.matched_b[q] = (int16_t)mb;
  O.matched_y[q] = (int16_t)my;
  O.matched_peaks[q] = (int16_t)kk;
  O.summed_intensity[q] = summed;
  O.matched_intensity_pct[q] = (100.0f * summed) / tic[s];
  O.average_ppm[q] = K.ppm[c];
  O.longest_b[q] = K.lb[c];
  O.longest_y[q] = (int16_t)ly;
  O.longest_y_pct[q] = (float)((double)ly / (double)len);
  O.peptide_len[q] = (int16_t)len;
  O.k1_count[q] = c_cnt[c];
  O.scored_candidates[q] = (int32_t)sc;
  O.poisson[q] = log10(pois);
  O.hyp[q] = h;
}

// candidates sorted by spectrum: each candidate's spectrum, and per spectrum the first candidate (lower bound)
__global__ __launch_bounds__(256) void cand_spec_kernel(int64_t n, const int32_t* __restrict__ c_hyp,
                                                        const int32_t* __restrict__ h_spec, int32_t* o_spec) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) o_spec[c] = h_spec[c_hyp[c]];
}

__global__ __launch_bounds__(256) void spec_off_kernel(int64_t n, const int32_t* __restrict__ c_hyp,
                                                       const int32_t* __restrict__ h_spec, int32_t n_spec,
                                                       int64_t* o_off) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > n_spec) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (h_spec[c_hyp[mid]] < s) lo = mid + 1;
    else hi = mid;
  }
  o_off[s] = lo;
}

extern "C" {

int vh_candidate_spectra(int64_t n_cand, const int32_t* c_hyp, const int32_t* h_spec, int32_t n_spec,
                         int32_t* o_spec, int64_t* o_off, void* stream) {
  if (n_cand < 0 || n_spec < 0 || !o_off || (n_cand && (!c_hyp || !h_spec))) return VH_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  if (o_spec && n_cand)
    cand_spec_kernel<<<(unsigned)((n_cand + 255) / 256), 256, 0, st>>>(n_cand, c_hyp, h_spec, o_spec);
  spec_off_kernel<<<(unsigned)((n_spec + 256) / 256), 256, 0, st>>>(n_cand, c_hyp, h_spec, n_spec, o_off);
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_k2(const VhParams* p, const VhIndexView* ix, int32_t n_groups, const int32_t* order, const int32_t* grp_spec,
          const int64_t* grp_start, const int32_t* grp_count, int32_t stride, const int32_t* c_row,
          const int32_t* c_hyp, const int8_t* h_maxfz, const float* c_shift, const int64_t* pk_off,
          const float* pk_mass, const float* pk_int, double* o_hs, float* o_sb, float* o_sy, float* o_ppm,
          int16_t* o_mb, int16_t* o_my, int16_t* o_lb, int16_t* o_ly, uint32_t* queue, void* stream) {
  if (!p || !ix || n_groups < 0 || (!grp_count && !grp_start)) return VH_E_ARG;
  if (!grp_start && stride <= 0) return VH_E_ARG;
  if (n_groups == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  K2Buf O = {o_hs, o_sb, o_sy, o_ppm, o_mb, o_my, o_lb, o_ly, queue};
  const unsigned blocks = (unsigned)((n_groups + K2_WAVES - 1) / K2_WAVES);
  unsigned g = xcd_band_grid(blocks, p->xcd_group);
  if (queue) {
    int dev = 0, cus = 256, per = 1;
    const void* kern = c_shift ? (const void*)k2_kernel<true> : (const void*)k2_kernel<false>;
    if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, 64 * K2_WAVES, 0) != hipSuccess || per < 1) per = 1;
    g = (unsigned)cus * (unsigned)per;
    g = g < blocks ? g : blocks;
    k2_zero_u32_kernel<<<1, 1, 0, st>>>(queue);
  }
  if (c_shift)
    k2_kernel<true><<<g, 64 * K2_WAVES, 0, st>>>(vh_make_devp(p), vh_make_devidx(ix), n_groups, order, grp_spec,
                                                 grp_start, grp_count, stride, c_row, c_hyp, h_maxfz, c_shift,
                                                 pk_off, pk_mass, pk_int, O);
  else
    k2_kernel<false><<<g, 64 * K2_WAVES, 0, st>>>(vh_make_devp(p), vh_make_devidx(ix), n_groups, order, grp_spec,
                                                  grp_start, grp_count, stride, c_row, c_hyp, h_maxfz, nullptr,
                                                  pk_off, pk_mass, pk_int, O);
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_candidate_shift(int32_t n_groups, const int64_t* grp_start, const int32_t* grp_count, int32_t stride,
                       const int32_t* c_row, const int32_t* c_hyp, const uint8_t* h_mode, const float* h_M,
                       const float* mass32, float* c_shift, void* stream) {
  if (n_groups < 0 || (!grp_count && !grp_start) || (!grp_start && stride <= 0)) return VH_E_ARG;
  if (n_groups == 0) return 0;
  if (!c_row || !c_hyp || !h_mode || !h_M || !mass32 || !c_shift) return VH_E_ARG;
  cand_shift_kernel<<<(unsigned)n_groups, 64, 0, (hipStream_t)stream>>>(n_groups, grp_start, grp_count, stride, c_row,
                                                                         c_hyp, h_mode, h_M, mass32, c_shift);
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_psm_features(const VhIndexView* ix, int32_t n_groups, const int32_t* grp_spec, int32_t report,
                    const int64_t* sel, const double* next_hs, const int32_t* c_row, const int32_t* c_hyp,
                    const int32_t* c_cnt, const VhK2Cols* k2, const int64_t* hyp_off, const int8_t* h_z,
                    const int8_t* h_iso, const float* h_pm, const uint32_t* h_scored,
                    const unsigned long long* h_matched, const float* tic, const VhPsmCols* out, void* stream) {
  if (!ix || n_groups < 0 || report < 1 || !k2 || !out) return VH_E_ARG;
  if (n_groups == 0) return 0;
  if (!sel || !next_hs || !c_row || !c_hyp || !c_cnt || !hyp_off || !h_z || !h_iso || !h_pm || !h_scored ||
      !h_matched || !tic)
    return VH_E_ARG;
  const VhK2Cols K = *k2;
  const VhPsmCols O = *out;
  if (!K.hs || !K.sb || !K.sy || !K.ppm || !K.mb || !K.my || !K.lb || !K.ly) return VH_E_ARG;
  if (!O.spec_idx || !O.pf_row || !O.rank || !O.charge || !O.isotope_error || !O.expmass || !O.calcmass ||
      !O.delta_mass_ppm || !O.hyperscore || !O.delta_next || !O.delta_best || !O.matched_b || !O.matched_y ||
      !O.matched_peaks || !O.summed_intensity || !O.matched_intensity_pct || !O.average_ppm || !O.longest_b ||
      !O.longest_y || !O.longest_y_pct || !O.peptide_len || !O.k1_count || !O.scored_candidates || !O.poisson)
    return VH_E_ARG;
  const int64_t slots = (int64_t)n_groups * report;
  psm_kernel<<<(unsigned)((slots + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      vh_make_devidx(ix), n_groups, grp_spec, report, sel, next_hs, c_row, c_hyp, c_cnt, K, hyp_off, h_z, h_iso,
      h_pm, h_scored, h_matched, tic, O);
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_rank(const VhParams* p, int32_t n_groups, const int64_t* grp_start, const int32_t* grp_count, int32_t stride,
            const int32_t* c_row, const double* hs, const int16_t* mb, const int16_t* my, int32_t report,
            int64_t* o_sel, double* o_next, void* stream) {
  if (!p || n_groups < 0 || (!grp_count && !grp_start) || report < 1 || !o_sel || !o_next) return VH_E_ARG;
  if (!grp_start && stride <= 0) return VH_E_ARG;
  if (n_groups == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  unsigned g = (unsigned)((n_groups + 3) / 4);
  rank_kernel<<<g, 256, 0, st>>>(vh_make_devp(p), n_groups, grp_start, grp_count, stride, c_row, hs, mb, my, report,
                                 o_sel, o_next);
  VH_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
