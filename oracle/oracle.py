"""CPU ORACLE wrapper -- test infrastructure, not product code.

Loads oracle/_build/liboracle.so (oracle/vh_oracle.cpp, built by oracle/Makefile). Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, as the checker and as the CPU
baseline; the product package never does. Arrays are host numpy with the same layouts as the device path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_build", "liboracle.so")

P = C.c_void_p
I32, I64, SZ, F64 = C.c_int32, C.c_int64, C.c_size_t, C.c_double

_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(SO):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []), check=True)
    return SO


def lib():
    global _lib
    if _lib is not None:
        return _lib
    from vulcan.gpu._abi import VhIndexView, VhParams
    L = C.CDLL(build())
    PP, PIX = C.POINTER(VhParams), C.POINTER(VhIndexView)
    sigs = {
        "or_version": ([], I32),
        "or_max_threads": ([], I32),
        "or_abi_sizeof": ([I32], SZ),
        "or_lnfact": ([I32], F64),
        "or_k0": ([PP, I32, P, P, P, P, P, P, P, P, I64], I32),
        "or_hypotheses": ([PP, I32, P, P, P, P, I64, P, P, P, P, P, P, P, P, P, P, P, I64], I32),
        "or_k1_counts": ([PP, PIX, I32, P, P, P, P, P, P, P, P, P, P], I32),
        "or_k1_brute": ([PP, PIX, I64, P, P, P, P, P, P, P], I32),
        "or_topk": ([PP, P, I32, P, P, P, P, P, P, I32, P, P, P, P], I32),
        "or_k2": ([PP, PIX, I32, P, P, P, I32, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P], I32),
        "or_rank": ([PP, I32, P, P, I32, P, P, P, P, I32, P, P], I32),
        "or_search": ([PP, PIX, I32, P, P, P, P, P, P, P, P, P, P], I32),
        "or_topk_feat": ([PP, PIX, I32, P, P, P, P, P, P, P, P, I32, P, P, P, P, P, P, P], I32),
        "or_iq_log": ([C.c_uint64], C.c_uint32),
        "or_iq_quant": ([C.c_float, C.c_float], C.c_uint32),
    }
    for name, (args, res) in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    if L.or_abi_sizeof(0) != C.sizeof(VhParams) or L.or_abi_sizeof(1) != C.sizeof(VhIndexView):
        raise RuntimeError("oracle ABI layout mismatch")
    _lib = L
    return L


def _p(a):
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"], type(a)
    return a.ctypes.data


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what}: error {rc}")


class HostIndex:
    """Host view of an L1 table for the oracle (no fragment index: the oracle's K1 works per row)."""

    def __init__(self, l1):
        from vulcan.gpu._abi import VhIndexView
        self.l1 = l1
        self.seq_off = np.ascontiguousarray(l1.seq_off, np.int64)
        self.length = np.ascontiguousarray(l1.length, np.uint8)
        self.mask = np.ascontiguousarray(l1.mask, np.uint64)
        self.mass32 = np.ascontiguousarray(l1.mass32, np.float32)
        self.residues = np.ascontiguousarray(l1.residues, np.uint8)
        self.partner = np.ascontiguousarray(l1.partner, np.int64)
        v = VhIndexView()
        v.n_rows = len(self.mass32)
        v.seq_off, v.length, v.mask = _p(self.seq_off), _p(self.length), _p(self.mask)
        v.mass32, v.residues, v.partner = _p(self.mass32), _p(self.residues), _p(self.partner)
        self.view = v


def k0(vp, sp):
    L = lib()
    n = len(sp)
    cap = max(n * vp.max_peaks, 1)
    o = {"off": np.zeros(n + 1, np.int64), "mass": np.zeros(cap, np.float32), "int": np.zeros(cap, np.float32),
         "tic": np.zeros(max(n, 1), np.float32), "valid": np.zeros(max(n, 1), np.uint8)}
    mz = np.ascontiguousarray(sp.mz, np.float32)
    it = np.ascontiguousarray(sp.intensity, np.float32)
    off = np.ascontiguousarray(sp.offsets, np.int64)
    _chk(L.or_k0(C.byref(vp), n, _p(mz), _p(it), _p(off), _p(o["off"]), _p(o["mass"]), _p(o["int"]), _p(o["tic"]),
                 _p(o["valid"]), cap), "k0")
    m = int(o["off"][-1])
    o["mass"], o["int"], o["tic"], o["valid"] = o["mass"][:m], o["int"][:m], o["tic"][:n], o["valid"][:n]
    return o


def hypotheses(vp, sp, valid, mass32, maxh: int):
    L = lib()
    n = len(sp)
    cap = max(n * maxh, 1)
    o = {"hyp_off": np.zeros(n + 1, np.int64), "spec": np.zeros(cap, np.int32), "z": np.zeros(cap, np.int8),
         "iso": np.zeros(cap, np.int8), "pm": np.zeros(cap, np.float32), "M": np.zeros(cap, np.float32),
         "lo": np.zeros(cap, np.int64), "hi": np.zeros(cap, np.int64), "maxfz": np.zeros(cap, np.int8),
         "mode": np.zeros(cap, np.uint8), "off": np.zeros(cap, np.float32)}
    pmz = np.ascontiguousarray(sp.precursor_mz, np.float64)
    z = np.ascontiguousarray(sp.charge, np.int8)
    va = np.ascontiguousarray(valid, np.uint8)
    m32 = np.ascontiguousarray(mass32, np.float32)
    _chk(L.or_hypotheses(C.byref(vp), n, _p(pmz), _p(z), _p(va), _p(m32), len(m32), _p(o["hyp_off"]), _p(o["spec"]),
                         _p(o["z"]), _p(o["iso"]), _p(o["pm"]), _p(o["M"]), _p(o["lo"]), _p(o["hi"]),
                         _p(o["maxfz"]), _p(o["mode"]), _p(o["off"]), cap), "hypotheses")
    nh = int(o["hyp_off"][-1])
    for k in ("spec", "z", "iso", "pm", "M", "lo", "hi", "maxfz", "mode", "off"):
        o[k] = o[k][:nh]
    return o


def k1_counts(vp, hix: HostIndex, peaks, hyp):
    """Counts of every window row (flat, window order) + scored/matched per hypothesis."""
    L = lib()
    n_spec = len(hyp["hyp_off"]) - 1
    nh = len(hyp["spec"])
    w = np.maximum(hyp["hi"] - hyp["lo"], 0)
    win_off = np.zeros(nh + 1, np.int64)
    np.cumsum(w, out=win_off[1:])
    counts = np.zeros(max(int(win_off[-1]), 1), np.int32)
    scored = np.zeros(max(nh, 1), np.uint32)
    matched = np.zeros(max(nh, 1), np.uint64)
    _chk(L.or_k1_counts(C.byref(vp), C.byref(hix.view), n_spec, _p(peaks["off"]), _p(peaks["mass"]),
                        _p(hyp["hyp_off"]), _p(np.ascontiguousarray(hyp["maxfz"])),
                        _p(np.ascontiguousarray(hyp["lo"])), _p(np.ascontiguousarray(hyp["hi"])), _p(win_off),
                        _p(counts), _p(scored), _p(matched)), "k1_counts")
    return {"win_off": win_off, "counts": counts[:int(win_off[-1])], "scored": scored[:nh], "matched": matched[:nh]}


def k1_append(hyp, k1, thr: int = 1):
    """(hyp, row, count) entries with count >= thr, sorted by (hyp, row)."""
    nh = len(hyp["spec"])
    w = np.maximum(hyp["hi"] - hyp["lo"], 0)
    hh = np.repeat(np.arange(nh), w)
    rr = np.repeat(hyp["lo"], w) + (np.arange(w.sum()) - np.repeat(k1["win_off"][:-1], w))
    c = k1["counts"]
    m = c >= thr
    return hh[m], rr[m], c[m]


def k1_brute(vp, hix: HostIndex, h, r, hyp, peaks):
    L = lib()
    h = np.ascontiguousarray(h, np.int32)
    r = np.ascontiguousarray(r, np.int32)
    out = np.zeros(max(len(h), 1), np.int32)
    _chk(L.or_k1_brute(C.byref(vp), C.byref(hix.view), len(h), _p(h), _p(r), _p(np.ascontiguousarray(hyp["spec"])),
                       _p(np.ascontiguousarray(hyp["maxfz"])), _p(peaks["off"]), _p(peaks["mass"]), _p(out)),
         "k1_brute")
    return out[:len(h)]


def topk(vp, hix: HostIndex, hyp, k1, stride: int):
    L = lib()
    n_spec = len(hyp["hyp_off"]) - 1
    n = max(n_spec, 1)
    o = {"n_cand": np.zeros(n, np.int32), "row": np.zeros(n * stride, np.int32),
         "hyp": np.zeros(n * stride, np.int32), "cnt": np.zeros(n * stride, np.int32)}
    mode = np.ascontiguousarray(hyp["mode"], np.uint8) if "mode" in hyp else None
    _chk(L.or_topk(C.byref(vp), _p(hix.partner), n_spec, _p(hyp["hyp_off"]), _p(mode),
                   _p(np.ascontiguousarray(hyp["lo"])),
                   _p(np.ascontiguousarray(hyp["hi"])), _p(k1["win_off"]), _p(np.ascontiguousarray(k1["counts"])),
                   stride, _p(o["n_cand"]), _p(o["row"]), _p(o["hyp"]), _p(o["cnt"])), "topk")
    o["n_cand"] = o["n_cand"][:n_spec]
    return o


def topk_feat(vp, hix: HostIndex, peaks, hyp, stride: int):
    """vh_k1_features restated: top-K lists (no closure) with count, b-ion count, per-ion intensity
    pseudo-log2s and the best count."""
    L = lib()
    n_spec = len(hyp["hyp_off"]) - 1
    n = max(n_spec, 1)
    o = {"n_cand": np.zeros(n, np.int32), "row": np.zeros(n * stride, np.int32),
         "hyp": np.zeros(n * stride, np.int32), "cnt": np.zeros(n * stride, np.int32),
         "nb": np.zeros(n * stride, np.int16), "iq": np.zeros(n * stride, np.uint32),
         "maxcnt": np.zeros(n, np.int32)}
    mode = np.ascontiguousarray(hyp["mode"], np.uint8) if "mode" in hyp else None
    _chk(L.or_topk_feat(C.byref(vp), C.byref(hix.view), n_spec, _p(peaks["off"]), _p(peaks["mass"]),
                        _p(peaks["int"]), _p(hyp["hyp_off"]), _p(np.ascontiguousarray(hyp["maxfz"])), _p(mode),
                        _p(np.ascontiguousarray(hyp["lo"])), _p(np.ascontiguousarray(hyp["hi"])), stride,
                        _p(o["n_cand"]), _p(o["row"]), _p(o["hyp"]), _p(o["cnt"]), _p(o["nb"]), _p(o["iq"]),
                        _p(o["maxcnt"])),
         "topk_feat")
    o["n_cand"] = o["n_cand"][:n_spec]
    o["maxcnt"] = o["maxcnt"][:n_spec]
    return o


def iq_log(s: int) -> int:
    return int(lib().or_iq_log(s))


def iq_quant(intensity: float, max_intensity: float) -> int:
    return int(lib().or_iq_quant(intensity, max_intensity))


def k2(vp, hix: HostIndex, peaks, row, hyp_idx, maxfz, grp_spec=None, grp_start=None, grp_count=None, stride=0,
       shift=None):
    """K2 outputs for candidates grouped by spectrum (same grouping arguments as vh_k2)."""
    L = lib()
    row = np.ascontiguousarray(row, np.int32)
    hyp_idx = np.ascontiguousarray(hyp_idx, np.int32)
    n = len(row)
    o = {"hyperscore": np.zeros(max(n, 1), np.float64), "summed_b": np.zeros(max(n, 1), np.float32),
         "summed_y": np.zeros(max(n, 1), np.float32), "ppm": np.zeros(max(n, 1), np.float32),
         "matched_b": np.zeros(max(n, 1), np.int16), "matched_y": np.zeros(max(n, 1), np.int16),
         "longest_b": np.zeros(max(n, 1), np.int16), "longest_y": np.zeros(max(n, 1), np.int16)}
    gc = np.ascontiguousarray(grp_count, np.int32)
    gs = None if grp_spec is None else np.ascontiguousarray(grp_spec, np.int32)
    g0 = None if grp_start is None else np.ascontiguousarray(grp_start, np.int64)
    _chk(L.or_k2(C.byref(vp), C.byref(hix.view), len(gc), _p(gs), _p(g0), _p(gc), stride, _p(row), _p(hyp_idx),
                 _p(np.ascontiguousarray(maxfz, np.int8)),
                 None if shift is None else _p(np.ascontiguousarray(shift, np.float32)), _p(peaks["off"]),
                 _p(peaks["mass"]), _p(peaks["int"]),
                 *[_p(o[k]) for k in ("hyperscore", "summed_b", "summed_y", "ppm", "matched_b", "matched_y",
                                      "longest_b", "longest_y")]), "k2")
    return {k: v[:n] for k, v in o.items()}


def rank(vp, grp_start, grp_count, row, hs, mb, my, report: int, stride: int = 0):
    L = lib()
    gc = np.ascontiguousarray(grp_count, np.int32)
    g0 = None if grp_start is None else np.ascontiguousarray(grp_start, np.int64)
    sel = np.zeros(max(len(gc) * report, 1), np.int64)
    nxt = np.zeros(max(len(gc) * report, 1), np.float64)
    _chk(L.or_rank(C.byref(vp), len(gc), _p(g0), _p(gc), stride, _p(np.ascontiguousarray(row, np.int32)),
                   _p(np.ascontiguousarray(hs, np.float64)), _p(np.ascontiguousarray(mb, np.int16)),
                   _p(np.ascontiguousarray(my, np.int16)), report, _p(sel), _p(nxt)), "rank")
    return sel[:len(gc) * report], nxt[:len(gc) * report]


def search(vp, hix: HostIndex, sp):
    """Whole closed search on the CPU (OpenMP over spectra): rank-1 row / hyperscore / delta_next per spectrum."""
    L = lib()
    n = len(sp)
    o = {"row": np.zeros(max(n, 1), np.int32), "hs": np.zeros(max(n, 1), np.float64),
         "dn": np.zeros(max(n, 1), np.float64), "n_cand": np.zeros(max(n, 1), np.int32)}
    stats = np.zeros(3, np.int64)
    mz = np.ascontiguousarray(sp.mz, np.float32)
    it = np.ascontiguousarray(sp.intensity, np.float32)
    off = np.ascontiguousarray(sp.offsets, np.int64)
    pmz = np.ascontiguousarray(sp.precursor_mz, np.float64)
    z = np.ascontiguousarray(sp.charge, np.int8)
    _chk(L.or_search(C.byref(vp), C.byref(hix.view), n, _p(mz), _p(it), _p(off), _p(pmz), _p(z), _p(o["row"]),
                     _p(o["hs"]), _p(o["dn"]), _p(o["n_cand"]), _p(stats)), "search")
    out = {k: v[:n] for k, v in o.items()}
    out["stats"] = {"n_rescored": int(stats[0]), "n_window": int(stats[1]), "matched": int(stats[2])}
    return out


def lnfact(n: int) -> float:
    return float(lib().or_lnfact(int(n)))
