// Host-side peptidoform expansion for the L1 peptidoform table (input producer of the search path).
//
// Restates vulcan/rust/vulcan_core/src/lib.rs:283-399 (`expand_pairs`) without Rust or sage-core:
//   * variable-mod configurations = every subset of the modifiable positions of the forward sequence with at
//     most `max_var` sites (optionally also at most `group_cap[g]` sites per modification group), plus the
//     optional peptide-N-term variable mod carried in mask bit 63 (lib.rs:180-190, 327-344);
//   * the decoy mask is the forward mask carried through the permutation dec[j] = fwd[perm[j]] (lib.rs:349-354);
//   * the f64 filter mass `form_mass` (lib.rs:256-269), Sage's f32 mass `sage_mass32` (lib.rs:238-254);
//   * ids = xxHash64 of the canonical peptidoform string (lib.rs:216-234), seed `id_seed`.
// Enumeration order is irrelevant to the table: rows are content-sorted afterwards (build.py:168).
//
// Plain C ABI (ctypes): two passes, count then fill, so the caller owns every output buffer.
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ------------------------------------------------------------------------------------------------ xxHash64
// Published algorithm (Yann Collet, xxHash specification v0.8): 4 lanes of 64-bit multiply-rotate rounds.
const uint64_t P1 = 11400714785074694791ULL;
const uint64_t P2 = 14029467366897019727ULL;
const uint64_t P3 = 1609587929392839161ULL;
const uint64_t P4 = 9650029242287828579ULL;
const uint64_t P5 = 2870177450012600261ULL;

inline uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
inline uint64_t rd64(const uint8_t* p) { uint64_t v; memcpy(&v, p, 8); return v; }
inline uint32_t rd32(const uint8_t* p) { uint32_t v; memcpy(&v, p, 4); return v; }
inline uint64_t round1(uint64_t acc, uint64_t in) { acc += in * P2; acc = rotl(acc, 31); return acc * P1; }
inline uint64_t merge(uint64_t acc, uint64_t v) { acc ^= round1(0, v); return acc * P1 + P4; }

uint64_t xxh64(const uint8_t* p, size_t len, uint64_t seed) {
  const uint8_t* end = p + len;
  uint64_t h;
  if (len >= 32) {
    uint64_t v1 = seed + P1 + P2, v2 = seed + P2, v3 = seed, v4 = seed - P1;
    const uint8_t* lim = end - 32;
    do {
      v1 = round1(v1, rd64(p)); v2 = round1(v2, rd64(p + 8));
      v3 = round1(v3, rd64(p + 16)); v4 = round1(v4, rd64(p + 24));
      p += 32;
    } while (p <= lim);
    h = rotl(v1, 1) + rotl(v2, 7) + rotl(v3, 12) + rotl(v4, 18);
    h = merge(h, v1); h = merge(h, v2); h = merge(h, v3); h = merge(h, v4);
  } else {
    h = seed + P5;
  }
  h += (uint64_t)len;
  while (p + 8 <= end) { h ^= round1(0, rd64(p)); h = rotl(h, 27) * P1 + P4; p += 8; }
  if (p + 4 <= end) { h ^= (uint64_t)rd32(p) * P1; h = rotl(h, 23) * P2 + P3; p += 4; }
  while (p < end) { h ^= (uint64_t)(*p) * P5; h = rotl(h, 11) * P1; ++p; }
  h ^= h >> 33; h *= P2; h ^= h >> 29; h *= P3; h ^= h >> 32;
  return h;
}

const uint64_t NTERM_BIT = 1ULL << 63;

struct Cfg {
  const double* static_res;   // [26] f64 static delta per residue
  const double* var_res;      // [26] f64 variable delta per residue (0 = none)
  const int32_t* var_group;   // [26] group id of the residue's variable mod (caps), -1 = none
  const int32_t* group_cap;   // [n_groups]
  int n_groups;
  int max_var;
  double nterm_var;           // 0 = no variable peptide-N-term mod
  double min_mass, max_mass;
  const float* mono32;        // [26] Sage f32 residue masses
  float h2o32;                // Sage f32 H2O
  double h2o64;               // f64 H2O (form_mass)
  const double* res64;        // [26] f64 residue masses (form_mass)
  uint64_t id_seed;
};

double form_mass(const uint8_t* s, int n, uint64_t mask, const Cfg& c) {
  double m = c.h2o64;
  for (int k = 0; k < n; ++k) {
    int i = s[k] - 'A';
    m += c.res64[i] + c.static_res[i];
    if ((mask >> k) & 1ULL) m += c.var_res[i];
  }
  if (mask & NTERM_BIT) m += c.nterm_var;
  return m;
}

float sage_mass32(const uint8_t* s, int n, uint64_t mask, const Cfg& c) {
  float m = c.h2o32;
  for (int k = 0; k < n; ++k) m += c.mono32[s[k] - 'A'];
  float ms = 0.0f;
  for (int k = 0; k < n; ++k) {
    int i = s[k] - 'A';
    float d = ((mask >> k) & 1ULL) ? (float)c.var_res[i] : (float)c.static_res[i];
    ms += d;
  }
  if (mask & NTERM_BIT) ms += (float)c.nterm_var;
  return m + ms;
}

std::string canonical(const uint8_t* s, int n, uint64_t mask, const Cfg& c) {
  std::string out;
  out.reserve(n + 24);
  char buf[48];
  if (mask & NTERM_BIT) { snprintf(buf, sizeof buf, "[%+.4f]-", c.nterm_var); out += buf; }
  for (int k = 0; k < n; ++k) {
    out.push_back((char)s[k]);
    if ((mask >> k) & 1ULL) { snprintf(buf, sizeof buf, "[%+.4f]", c.var_res[s[k] - 'A']); out += buf; }
  }
  return out;
}

// Enumerate the variable-mod masks of one forward sequence (deterministic DFS over modifiable positions).
template <class F>
void enumerate(const uint8_t* s, int n, const Cfg& c, F&& emit) {
  int pos[64], grp[64], np = 0;
  for (int k = 0; k < n; ++k) {
    int i = s[k] - 'A';
    if (c.var_res[i] != 0.0) { pos[np] = k; grp[np] = c.var_group[i]; ++np; }
  }
  int n_nt = c.nterm_var != 0.0 ? 2 : 1;
  for (int nt = 0; nt < n_nt; ++nt) {
    uint64_t base = nt ? NTERM_BIT : 0ULL;
    // iterative DFS over subsets of `pos` in increasing position order; frame = (next candidate, mask, #sites)
    struct Fr { int next; uint64_t mask; int cnt; };
    Fr st[65];
    int sp = 0;
    st[sp++] = {0, base, nt};
    emit(base);
    while (sp) {
      Fr& f = st[sp - 1];
      if (f.next >= np || f.cnt >= c.max_var) { --sp; continue; }
      int j = f.next++;
      int g = grp[j];
      if (g >= 0 && g < c.n_groups) {   // per-group cap: count the group's sites already in the mask
        int gu = 0;
        for (int q = 0; q < j; ++q) if (((f.mask >> pos[q]) & 1ULL) && grp[q] == g) ++gu;
        if (gu >= c.group_cap[g]) continue;
      }
      uint64_t m2 = f.mask | (1ULL << pos[j]);
      int c2 = f.cnt + 1;
      emit(m2);
      st[sp++] = {j + 1, m2, c2};
    }
  }
}

}  // namespace

extern "C" {

uint64_t vl_xxh64(const uint8_t* data, int64_t len, uint64_t seed) { return xxh64(data, (size_t)len, seed); }

// Pass 1: number of peptidoforms (within [min_mass, max_mass]) per forward sequence.
int vl_expand_count(int64_t n_pep, const uint8_t* seq, const int64_t* seq_off, const double* static_res,
                    const double* var_res, const int32_t* var_group, const int32_t* group_cap, int n_groups,
                    int max_var, double nterm_var, double min_mass, double max_mass, const double* res64,
                    double h2o64, int64_t* counts) {
  Cfg c{static_res, var_res, var_group, group_cap, n_groups, max_var, nterm_var, min_mass, max_mass,
        nullptr, 0.0f, h2o64, res64, 0};
  if (n_groups > 64) return -1;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n_pep; ++i) {
    const uint8_t* s = seq + seq_off[i];
    int n = (int)(seq_off[i + 1] - seq_off[i]);
    int64_t k = 0;
    enumerate(s, n, c, [&](uint64_t mask) {
      double m = form_mass(s, n, mask, c);
      if (!(m < c.min_mass || m > c.max_mass)) ++k;
    });
    counts[i] = k;
  }
  return 0;
}

// Pass 2: fill one row per pair of peptidoforms. out_off = exclusive prefix sum of the counts.
// perm_flat/perm_off: dec[j] = fwd[perm[j]]; an empty slice means the pseudo-reverse with the C-term fixed.
int vl_expand_fill(int64_t n_pep, const uint8_t* seq, const int64_t* seq_off, const uint8_t* perm_flat,
                   const int64_t* perm_off, const double* static_res, const double* var_res,
                   const int32_t* var_group, const int32_t* group_cap, int n_groups, int max_var, double nterm_var,
                   double min_mass, double max_mass, const double* res64, double h2o64, const float* mono32,
                   float h2o32, uint64_t id_seed, const int64_t* out_off, uint32_t* o_seq, uint64_t* o_mask_f,
                   uint64_t* o_mask_d, double* o_mass, uint8_t* o_nvar, uint64_t* o_id_f, uint64_t* o_id_d,
                   float* o_m32f, float* o_m32d) {
  Cfg c{static_res, var_res, var_group, group_cap, n_groups, max_var, nterm_var, min_mass, max_mass,
        mono32, h2o32, h2o64, res64, id_seed};
  if (n_groups > 64) return -1;
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : bad)
  for (int64_t i = 0; i < n_pep; ++i) {
    const uint8_t* s = seq + seq_off[i];
    int n = (int)(seq_off[i + 1] - seq_off[i]);
    if (n > 63 || n < 2) { bad |= 1; continue; }
    int perm[64];
    int64_t a = perm_off[i], b = perm_off[i + 1];
    if (a == b) {
      for (int j = 0; j < n; ++j) perm[j] = j + 1 < n ? n - 2 - j : n - 1;
    } else {
      if (b - a != n) { bad |= 2; continue; }
      for (int j = 0; j < n; ++j) perm[j] = perm_flat[a + j];
    }
    uint8_t d[64];
    for (int j = 0; j < n; ++j) d[j] = s[perm[j]];
    int64_t o = out_off[i];
    enumerate(s, n, c, [&](uint64_t mask) {
      double m = form_mass(s, n, mask, c);
      if (m < c.min_mass || m > c.max_mass) return;
      uint64_t md = mask & NTERM_BIT;
      for (int j = 0; j < n; ++j)
        if ((mask >> perm[j]) & 1ULL) md |= 1ULL << j;
      std::string cf = canonical(s, n, mask, c), cd = canonical(d, n, md, c);
      o_seq[o] = (uint32_t)i;
      o_mask_f[o] = mask;
      o_mask_d[o] = md;
      o_mass[o] = m;
      o_nvar[o] = (uint8_t)__builtin_popcountll(mask);
      o_id_f[o] = xxh64((const uint8_t*)cf.data(), cf.size(), c.id_seed);
      o_id_d[o] = xxh64((const uint8_t*)cd.data(), cd.size(), c.id_seed);
      o_m32f[o] = sage_mass32(s, n, mask, c);
      o_m32d[o] = sage_mass32(d, n, md, c);
      ++o;
    });
    if (o != out_off[i + 1]) bad |= 4;
  }
  return bad ? -bad : 0;
}

}  // extern "C"
