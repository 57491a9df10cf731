"""ctypes mirror of include/vulcan_hip.h (VhParams, VhIndexView, VhBlockView, VhK2Cols, VhPsmCols) and conversion
from SearchParams.

Torch-free. The struct layouts are checked against the C side with vh_abi_sizeof (tests).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .host import SAGE_MONO, SearchParams

F32, I32, I64, U32, P = C.c_float, C.c_int32, C.c_int64, C.c_uint32, C.c_void_p


class VhParams(C.Structure):
    _fields_ = [("mono", F32 * 26), ("stat", F32 * 26), ("var", F32 * 26), ("nterm", F32), ("tol_da", I32),
                ("tol_lo", F32), ("tol_hi", F32), ("fmin", F32), ("fmax", F32), ("min_ion_index", I32),
                ("prec_lo", F32), ("prec_hi", F32), ("iso_lo", I32), ("iso_hi", I32), ("z_lo", I32), ("z_hi", I32),
                ("max_fragment_charge", I32), ("max_peaks", I32), ("min_peaks", I32), ("min_matched_peaks", I32),
                ("topk", I32), ("k1_threshold", I32), ("pair_closure", I32), ("tile", I32), ("xcd_group", I32),
                ("n_offsets", I32), ("offsets", F32 * 8), ("open_search", I32), ("open_lo", F32), ("open_hi", F32),
                ("k1_threshold_open", I32), ("peak_window_k", I32), ("peak_window_th", F32), ("reserved", I32 * 6)]


class VhIndexView(C.Structure):
    _fields_ = [("n_rows", I64), ("seq_off", P), ("length", P), ("mask", P), ("mass32", P), ("residues", P),
                ("partner", P), ("bucket_log2", I32), ("bin_shift", I32), ("bin_base", U32), ("n_bins", I32),
                ("n_buckets", I64), ("offsets", P), ("postings", P), ("n_postings", I64)]


class VhBlockView(C.Structure):
    _fields_ = [("n_blocks", I64), ("max_rows", I32), ("max_postings", I32), ("row0", P), ("off", P), ("key", P),
                ("row", P)]


class VhK2Cols(C.Structure):
    _fields_ = [(k, P) for k in ("hs", "sb", "sy", "ppm", "mb", "my", "lb", "ly")]


PSM_COLS = ("spec_idx", "pf_row", "rank", "charge", "isotope_error", "expmass", "calcmass", "delta_mass_ppm",
            "hyperscore", "delta_next", "delta_best", "matched_b", "matched_y", "matched_peaks", "summed_intensity",
            "matched_intensity_pct", "average_ppm", "longest_b", "longest_y", "longest_y_pct", "peptide_len",
            "k1_count", "scored_candidates", "poisson", "hyp")


class VhPsmCols(C.Structure):
    _fields_ = [(k, P) for k in PSM_COLS]


FAMILY_COLS = ("idx", "group", "rank", "row", "hyp", "hs", "mb", "my", "z", "iso")


class VhFamilyCols(C.Structure):
    """Output columns of vh_family_top2 (mode 1)."""
    _fields_ = [(k, P) for k in FAMILY_COLS]


GBM_NODE = np.dtype([("v", "<f8"), ("left", "<i4"), ("feature", "<i2"), ("missing_left", "u1"), ("is_leaf", "u1")])
GBM_CHUNK_NODES = 1024    # VH_GBM_CHUNK_NODES
GBM_MAX_FEATURES = 63     # VH_GBM_MAX_FEATURES


def mod_arrays(static_mods: dict, var_mods: dict):
    """f32 static/variable tables and the N-term delta as the reference sets them (engine.py:71-82)."""
    stat = np.zeros(26, np.float32)
    var = np.zeros(26, np.float32)
    nterm = 0.0
    for r, m in static_mods.items():
        stat[ord(r) - 65] = np.float32(m)
    for r, m in var_mods.items():
        if r == "^":
            nterm = float(np.float32(m))
            continue
        var[ord(r) - 65] = np.float32(m)
    return stat, var, nterm


def make_params(p: SearchParams, static_mods: dict, var_mods: dict) -> VhParams:
    if p.fragment_unit not in ("ppm", "da"):
        raise ValueError("fragment_unit must be 'ppm' or 'da'")
    if p.peak_mode not in ("global", "window"):
        raise ValueError("peak_mode must be 'global' or 'window'")
    stat, var, nterm = mod_arrays(static_mods, var_mods)
    v = VhParams()
    v.mono[:] = SAGE_MONO.tolist()
    v.stat[:] = stat.tolist()
    v.var[:] = var.tolist()
    v.nterm = nterm
    v.tol_da = int(p.fragment_unit == "da")
    v.tol_lo, v.tol_hi = float(p.fragment_ppm[0]), float(p.fragment_ppm[1])
    v.fmin, v.fmax = float(p.fragment_min_mz), float(p.fragment_max_mz)
    v.min_ion_index = int(p.min_ion_index)
    v.prec_lo, v.prec_hi = float(p.precursor_ppm[0]), float(p.precursor_ppm[1])
    v.iso_lo, v.iso_hi = int(p.isotope_errors[0]), int(p.isotope_errors[1])
    v.z_lo, v.z_hi = int(p.precursor_charge[0]), int(p.precursor_charge[1])
    v.max_fragment_charge = int(p.max_fragment_charge or 0)
    v.max_peaks, v.min_peaks = peak_capacity(p), int(p.min_peaks)
    if p.peak_mode == "window":
        v.peak_window_k, v.peak_window_th = int(p.window_k), float(np.float32(p.window_th))
    v.min_matched_peaks = int(p.min_matched_peaks)
    v.topk = int(p.topk)
    v.k1_threshold = int(p.k1_threshold)
    v.pair_closure = int(bool(p.pair_closure))
    v.tile = int(p.tile)
    v.xcd_group = int(p.xcd_group)
    v.k1_threshold_open = int(p.k1_threshold)
    return v


def peak_capacity(p: SearchParams) -> int:
    """Kept peaks per spectrum at most: max_peaks in the global view; in the window view (no total cap,
    engine.py:182-198) window_k per floor(m/z / window_th) window of [fragment_min_mz, fragment_max_mz]
    (f32 division, as K0 computes it). The device K0 and K2 hold at most 512 peaks per spectrum."""
    if p.peak_mode != "window":
        return int(p.max_peaks)
    th = np.float32(p.window_th)
    if not (p.window_k > 0 and th > 0):
        raise ValueError("the window view needs window_k > 0 and window_th > 0")
    w0 = int(np.floor(np.float32(p.fragment_min_mz) / th))
    w1 = int(np.floor(np.float32(p.fragment_max_mz) / th))
    if w1 >= 1000:
        raise ValueError("window view: fragment_max_mz / window_th >= 1000 would mix spectra in the reference's "
                         "grouping key (spectrum * 1000 + window, engine.py:185)")
    cap = (w1 - w0 + 1) * int(p.window_k)
    if cap > 512:
        raise ValueError(f"window view keeps up to {cap} peaks per spectrum; at most 512 are supported")
    return cap


def set_modes(v: VhParams, offsets=(), open_da=None, k1_threshold_open: int = 1) -> VhParams:
    """Mass-offset / open hypotheses (ModeParams, vulcan/ptm/search.py:38-45) into a VhParams."""
    if len(offsets) > 8:
        raise ValueError("at most 8 mass offsets")
    v.n_offsets = len(offsets)
    for i, o in enumerate(offsets):
        v.offsets[i] = float(np.float32(o))
    v.open_search = int(open_da is not None)
    if open_da is not None:
        v.open_lo, v.open_hi = float(np.float32(open_da[0])), float(np.float32(open_da[1]))
    v.k1_threshold_open = int(k1_threshold_open)
    return v


def max_hyp_per_spectrum(p: SearchParams, n_offsets: int = 0, open_search: bool = False) -> int:
    niso = p.isotope_errors[1] - p.isotope_errors[0] + 1
    nz = max(1, p.precursor_charge[1] - p.precursor_charge[0] + 1)
    return niso * nz + nz * (n_offsets + int(bool(open_search)))


def f32_bits(x: float) -> int:
    return int(np.array([x], np.float32).view(np.uint32)[0])


DIRECTORY_MAX = 2 ** 30     # directory entries (4 GiB of u32 offsets)


def index_geometry(p: SearchParams, n_rows: int):
    """(bucket_log2, bin_shift, bin_base, n_bins, n_buckets) of the fragment index for `n_rows` rows."""
    c = int(p.chunk_size)
    if c <= 0 or c & (c - 1) or c > 2 ** 16:
        raise ValueError("chunk_size must be a power of two <= 65536")
    if not 0 < p.fragment_min_mz <= p.fragment_max_mz:
        raise ValueError("need 0 < fragment_min_mz <= fragment_max_mz")
    if f32_bits(p.fragment_max_mz) - f32_bits(p.fragment_min_mz) >= 2 ** 26:
        raise ValueError("fragment_max_mz / fragment_min_mz spans more than 8 binary octaves")
    s = p.bin_shift()
    base = f32_bits(p.fragment_min_mz) >> s
    n_bins = (f32_bits(p.fragment_max_mz) >> s) - base + 1
    # a larger table gets larger buckets so that the directory stays under DIRECTORY_MAX entries
    while c < 2 ** 16 and max(1, (n_rows + c - 1) // c) * n_bins > DIRECTORY_MAX:
        c *= 2
    lb = c.bit_length() - 1
    n_buckets = max(1, (n_rows + c - 1) // c)
    if n_buckets * n_bins >= 2 ** 32:
        raise ValueError("bucket x bin directory exceeds 2^32 entries; raise chunk_size or bin_ppm")
    return lb, s, base, n_bins, n_buckets
