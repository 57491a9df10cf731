"""Torch-free host-side pieces of the closed search: constants, SearchParams, Spectra, precursor hypotheses.

Mirrors vulcan/gpu/host.py:10-105 of the reference (same names, fields and arithmetic), so decode workers can
build search-ready arrays without importing torch. The device path computes the same hypotheses with the
`vh_hypotheses` HIP kernel (include/vulcan_hip.h); this numpy version is kept for callers that precompute
hypotheses on the host (reference `pre_hypotheses`, host.py:94-105).
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
    peak_mode: str = "global"               # only "global" (Sage) is implemented on the device
    window_th: float = 100.0
    window_k: int = 15
    # index layout
    chunk_size: int = 512                   # rows per precursor-mass bucket (power of two, <= 2^16)
    bin_ppm: float = 80.0                   # directory bin width; rounded down to a power-of-two f32 bit bin
    tile: int = 512                         # K1 LDS histogram slots per wave (rows per tile)
    k2_threads: int = 128                   # kept for signature compatibility (K2 runs one wave64 per spectrum)
    max_slice_entries: int = 2 ** 32 - 1    # postings per index (u32 offsets)
    xcd_group: int = 32                     # K1/K2 blocks per XCD per scheduling group (0 = plain block order)
    k1_impl: str = "auto"                   # K1: "tile" (spectrum-major), "join" (block-major), "auto" (join
                                            # unless an open-search window is on)
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


def make_hypotheses(sp: Spectra, valid: np.ndarray, idx_mass32: np.ndarray, p: SearchParams):
    """Precursor hypotheses (spectrum, charge, isotope) and candidate row windows, Sage f32 arithmetic
    (host.py:64-91): order (spectrum, charge, isotope); window rows by searchsorted left/right on mass32."""
    spec = np.nonzero(np.asarray(valid, bool) & np.isfinite(sp.precursor_mz))[0]
    zi = sp.charge[spec].astype(np.int64)
    known = zi > 0
    zr = np.arange(p.precursor_charge[0], p.precursor_charge[1] + 1)
    isos = np.arange(p.isotope_errors[0], p.isotope_errors[1] + 1)
    parts_s, parts_z, parts_i = [], [], []
    for iso in isos:
        parts_s.append(spec[known]); parts_z.append(zi[known]); parts_i.append(np.full(known.sum(), iso))
        for z in zr:
            parts_s.append(spec[~known]); parts_z.append(np.full((~known).sum(), z))
            parts_i.append(np.full((~known).sum(), iso))
    s = np.concatenate(parts_s) if parts_s else np.zeros(0, np.int64)
    z = np.concatenate(parts_z) if parts_z else np.zeros(0, np.int64)
    iso = np.concatenate(parts_i) if parts_i else np.zeros(0, np.int64)
    o = np.lexsort((iso, z, s))
    s, z, iso = s[o], z[o], iso[o]
    mz = (sp.precursor_mz[s].astype(np.float32) - PROTON).astype(np.float32)
    pm = (mz * z.astype(np.float32)).astype(np.float32)
    M = (pm - (iso.astype(np.float32) * NEUTRON).astype(np.float32)).astype(np.float32)
    plo, phi = np.float32(p.precursor_ppm[0]), np.float32(p.precursor_ppm[1])
    lo = (M + (M * plo) / np.float32(1e6)).astype(np.float32)
    hi = (M + (M * phi) / np.float32(1e6)).astype(np.float32)
    rlo = np.searchsorted(idx_mass32, lo, side="left")
    rhi = np.searchsorted(idx_mass32, hi, side="right")
    return {"spec": s.astype(np.int32), "z": z.astype(np.int8), "iso": iso.astype(np.int8), "pm": pm, "M": M,
            "lo": rlo.astype(np.int64), "hi": rhi.astype(np.int64), "maxfz": p.maxfz_of(z)}


def pre_hypotheses(sp: Spectra, gi, p: SearchParams):
    """Host hypotheses computable before K0 for the global peak view (host.py:94-105)."""
    assert p.peak_mode == "global"
    mz = sp.mz
    ok = (mz >= p.fragment_min_mz) & (mz <= p.fragment_max_mz) & ~np.isnan(mz)
    cs = np.zeros(len(ok) + 1, np.int64)
    np.cumsum(ok, out=cs[1:])
    n_in = cs[sp.offsets[1:]] - cs[sp.offsets[:-1]]
    valid = np.minimum(n_in, p.max_peaks) >= p.min_peaks
    return valid, make_hypotheses(sp, valid, gi.mass32_np, p)
