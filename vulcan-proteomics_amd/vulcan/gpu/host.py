"""Torch-free host-side pieces of the closed search: constants, SearchParams, Spectra, precursor hypotheses.

Same names, fields and arithmetic as the reference's vulcan/gpu/host.py:10-105, so decode workers can build
search-ready arrays without importing torch. The device path computes the same hypotheses with the
`vh_hypotheses` HIP kernel (include/vulcan_hip.h); this numpy version serves callers that precompute hypotheses
on the host (`pre_hypotheses`).
"""
from __future__ import annotations

import math
from dataclasses import asdict, dataclass

import numpy as np

# host.py:10-11 (Sage f32 constants)
PROTON = np.float32(1.0072764)
NEUTRON = np.float32(1.00335)
# Sage f32 H2O (mirrored at vulcan/gates/a2_identity.py:35 of the reference)
H2O32 = np.float32(18.010565)
# Sage MONOISOTOPIC residue masses, f32, index = ASCII - 'A' (engine.py:30-32)
SAGE_MONO = np.array([71.03711, 0.0, 103.00919, 115.02694, 129.04259, 147.0684, 57.02146, 137.05891, 113.08406,
                      0.0, 128.09496, 113.08406, 131.0405, 114.04293, 237.14774, 97.05276, 128.05858, 156.1011,
                      87.03203, 101.04768, 150.95363, 99.06841, 186.07932, 0.0, 163.06332, 0.0], np.float32)


@dataclass
class SearchParams:
    """Search parameters (host.py:15-44). Fields after `max_slice_entries` are new to this implementation."""
    precursor_ppm: tuple = (-10.0, 10.0)
    fragment_ppm: tuple = (-20.0, 20.0)
    fragment_unit: str = "ppm"              # "ppm" | "da" (fragment_ppm then holds Da bounds)
    isotope_errors: tuple = (0, 2)
    precursor_charge: tuple = (2, 4)        # tried when the spectrum gives none
    max_peaks: int = 150
    min_peaks: int = 15
    fragment_min_mz: float = 150.0
    fragment_max_mz: float = 2000.0
    min_ion_index: int = 2
    max_fragment_charge: int | None = 1
    min_matched_peaks: int = 4
    topk: int = 500
    k1_threshold: int = 1
    pair_closure: bool = True
    report_psms: int = 1
    peak_mode: str = "global"               # "global" (Sage: top max_peaks) | "window" (top window_k per window_th)
    window_th: float = 100.0
    window_k: int = 15
    # index layout
    chunk_size: int = 512                   # rows per precursor-mass bucket (power of two, <= 2^16)
    bin_ppm: float = 80.0                   # directory bin width; rounded down to a power-of-two f32 bit bin
    tile: int = 512                         # K1 LDS histogram slots per wave (rows per tile)
    k2_threads: int = 128                   # kept for signature compatibility (K2 runs one wave64 per spectrum)
    max_slice_entries: int = 2 ** 31 - 1    # postings per row-range slice of the join K1's block index
    xcd_group: int = 32                     # K1/K2 blocks per XCD per scheduling group (0 = plain block order)
    k1_impl: str = "auto"                   # K1: "tile" (spectrum-major), "join" (block-major), "wide"
                                            # (workgroup per spectrum over 2^14-row buckets), "auto" (join, wide
                                            # when an open-search window is on)
    block_postings_log2: int = 12           # block layout: ~2^q postings per block (LDS-staged by the join K1)
    block_rows_log2: int = 10               # block layout: at most 2^rb rows per block

    def to_dict(self):
        return asdict(self)

    def bin_shift(self) -> int:
        """f32 bit-pattern bin: key >> shift; a bin spans 2^(shift-23) relative (shift 7 = 15.3 ppm)."""
        s = int(math.floor(23 + math.log2(self.bin_ppm * 1e-6)))
        return max(4, min(15, s))

    def maxfz_of(self, z):
        mfc = self.max_fragment_charge
        z = np.asarray(z, np.int64)
        return np.maximum(np.minimum(z, mfc + 1) if mfc else z, 2).astype(np.int8)


@dataclass
class Spectra:
    """MS2 spectra of one run (host.py:48-61)."""
    run: str
    scan: np.ndarray
    index: np.ndarray
    rt: np.ndarray
    precursor_mz: np.ndarray       # f64
    charge: np.ndarray             # i8, 0 = unknown
    mz: np.ndarray                 # f32 flat
    intensity: np.ndarray          # f32 flat
    offsets: np.ndarray            # i64 n+1

    def __len__(self):
        return len(self.scan)


def _hypothesis_grid(spec: np.ndarray, charge: np.ndarray, p: SearchParams):
    """(spectrum, charge, isotope) of every precursor hypothesis, grouped by spectrum, charge ascending, then
    isotope ascending: a spectrum with a known charge (> 0) tries that charge only, the others every charge of
    p.precursor_charge; every charge takes every isotope error of p.isotope_errors."""
    z_lo, z_hi = (int(x) for x in p.precursor_charge)
    i_lo, i_hi = (int(x) for x in p.isotope_errors)
    n_iso = max(i_hi - i_lo + 1, 0)
    z_known = charge[spec].astype(np.int64)
    per = np.where(z_known > 0, 1, max(z_hi - z_lo + 1, 0)) * n_iso
    total = int(per.sum())
    if total == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0, np.int64)
    start = np.cumsum(per) - per
    j = np.arange(total) - np.repeat(start, per)           # position inside the spectrum's block
    zk = np.repeat(z_known, per)
    z = np.where(zk > 0, zk, z_lo + j // n_iso)
    return np.repeat(spec, per), z, i_lo + j % n_iso


def _precursor_window(pmz: np.ndarray, z: np.ndarray, iso: np.ndarray, p: SearchParams):
    """Sage's f32 precursor arithmetic, one rounding per operation: m/z - proton, times charge, minus the isotope
    shift, then the ppm bounds m + (m * ppm) / 1e6."""
    f32 = np.float32
    neutral = pmz.astype(f32) - PROTON                       # f32 - f32
    pm = neutral * z.astype(f32)
    M = pm - iso.astype(f32) * NEUTRON
    bounds = [M + (M * f32(t)) / f32(1e6) for t in p.precursor_ppm]
    return pm.astype(f32), M.astype(f32), bounds[0].astype(f32), bounds[1].astype(f32)


def make_hypotheses(sp: Spectra, valid: np.ndarray, idx_mass32: np.ndarray, p: SearchParams):
    """Precursor hypotheses and their candidate row windows (host.py:64-91): every valid spectrum with a finite
    precursor, grouped by spectrum, charge, isotope; the window is the rows whose mass32 lies in [lo, hi]
    (searchsorted left / right on the mass-sorted table)."""
    spec = np.flatnonzero(np.asarray(valid, bool) & np.isfinite(sp.precursor_mz))
    s, z, iso = _hypothesis_grid(spec, np.asarray(sp.charge), p)
    pm, M, lo, hi = _precursor_window(np.asarray(sp.precursor_mz, np.float64)[s], z, iso, p)
    return {"spec": s.astype(np.int32), "z": z.astype(np.int8), "iso": iso.astype(np.int8), "pm": pm, "M": M,
            "lo": np.searchsorted(idx_mass32, lo, side="left").astype(np.int64),
            "hi": np.searchsorted(idx_mass32, hi, side="right").astype(np.int64), "maxfz": p.maxfz_of(z)}


def pre_hypotheses(sp: Spectra, gi, p: SearchParams):
    """(valid, hypotheses) computable on the host before K0, e.g. while a run is loaded (host.py:94-105): with
    the global peak view a spectrum is valid when min(#peaks with fragment_min_mz <= m/z <= fragment_max_mz,
    max_peaks) >= min_peaks. `engine.search(pre_hyp=...)` uses them only when this mask equals K0's."""
    if p.peak_mode != "global":
        raise ValueError("pre_hypotheses needs the global peak view")
    mz = np.asarray(sp.mz)
    kept = np.flatnonzero((mz >= p.fragment_min_mz) & (mz <= p.fragment_max_mz))   # NaN compares false
    n_in = np.diff(np.searchsorted(kept, np.asarray(sp.offsets)))
    valid = np.minimum(n_in, p.max_peaks) >= p.min_peaks
    return valid, make_hypotheses(sp, valid, gi.mass32_np, p)
