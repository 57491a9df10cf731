"""ctypes binding of libvulcan_l1.so (csrc/l1_expand.cpp): peptidoform expansion and xxHash64.

Replaces the Rust/PyO3 `vulcan_core` crate of the reference (vulcan/rust/vulcan_core/src/lib.rs) for the parts
the L1 builder needs; Rust and sage-core are not available to this build.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _build

# Monoisotopic residue masses, f64, index = ASCII - 'A' (lib.rs:33-60); 0.0 = not a standard residue
RESIDUE_F64 = np.array([71.037113805, 0.0, 103.009184505, 115.026943065, 129.042593135, 147.068413945,
                        57.021463721, 137.058911875, 113.084064015, 0.0, 128.094963050, 113.084064015,
                        131.040484645, 114.042927470, 0.0, 97.052763875, 128.058577540, 156.101111050,
                        87.032028435, 101.047678505, 0.0, 99.068413945, 186.079312980, 0.0, 163.063328575, 0.0],
                       np.float64)
H2O_F64 = 18.010564684          # lib.rs:61
NTERM_BIT = np.uint64(1) << np.uint64(63)
STD_AA = set("ACDEFGHIKLMNPQRSTVWY")

_lib = None
P = C.c_void_p
I32, I64, U64, F32, F64 = C.c_int, C.c_int64, C.c_uint64, C.c_float, C.c_double


def lib():
    global _lib
    if _lib is None:
        so = _build.build_l1(force=False)
        L = C.CDLL(so)
        L.vl_xxh64.argtypes = [P, I64, U64]
        L.vl_xxh64.restype = U64
        L.vl_expand_count.argtypes = [I64, P, P, P, P, P, P, I32, I32, F64, F64, F64, P, F64, P]
        L.vl_expand_count.restype = I32
        L.vl_expand_fill.argtypes = [I64, P, P, P, P, P, P, P, P, I32, I32, F64, F64, F64, P, F64, P, F32, U64, P,
                                     P, P, P, P, P, P, P, P, P]
        L.vl_expand_fill.restype = I32
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def xxh64(data: bytes, seed: int = 0) -> int:
    buf = np.frombuffer(data, np.uint8) if data else np.zeros(1, np.uint8)
    return int(lib().vl_xxh64(_p(buf), len(data), seed))


def mod_tables(static_mods: dict, var_mods: dict, var_caps: dict | None = None):
    """Per-residue f64 static/variable deltas, variable-mod group ids/caps and the N-term delta ('^')."""
    stat = np.zeros(26, np.float64)
    var = np.zeros(26, np.float64)
    group = np.full(26, -1, np.int32)
    caps: list[int] = []
    nterm = 0.0
    for r, m in static_mods.items():
        stat[ord(r) - 65] = float(m)
    gid: dict[float, int] = {}
    for r, m in var_mods.items():
        if r == "^":
            nterm = float(m)
            continue
        i = ord(r) - 65
        if var[i] != 0.0 or stat[i] != 0.0:
            raise ValueError("at most one variable mod per residue, and not on a statically modified residue")
        var[i] = float(m)
        key = round(float(m), 4)
        if key not in gid:
            gid[key] = len(caps)
            caps.append(int((var_caps or {}).get(r, 64)))
        group[i] = gid[key]
        if var_caps and r in var_caps:
            caps[gid[key]] = min(caps[gid[key]], int(var_caps[r]))
    return stat, var, group, np.array(caps or [64], np.int32), nterm


def expand_pairs(fwd: list[str], dec_perm: list, static_mods: dict, var_mods: dict, max_var: int,
                 min_mass: float, max_mass: float, id_seed: int = 0, var_caps: dict | None = None,
                 mono32: np.ndarray | None = None, h2o32: float | None = None):
    """Pairs of peptidoforms (forward, decoy) as in lib.rs:283-399. dec_perm[i] = permutation array or None
    (pseudo-reverse). Returns dict of numpy arrays (seq, mask_f, mask_d, mass, n_var, id_f, id_d, mass32_f,
    mass32_d)."""
    from ..gpu.host import H2O32, SAGE_MONO
    mono32 = SAGE_MONO if mono32 is None else mono32
    h2o32 = float(H2O32) if h2o32 is None else h2o32
    L = lib()
    n = len(fwd)
    seq = np.frombuffer("".join(fwd).encode(), np.uint8).copy() if n else np.zeros(1, np.uint8)
    off = np.zeros(n + 1, np.int64)
    if n:
        np.cumsum([len(s) for s in fwd], out=off[1:])
    pl = [np.asarray(q, np.uint8) if q is not None else np.zeros(0, np.uint8) for q in dec_perm]
    poff = np.zeros(n + 1, np.int64)
    if n:
        np.cumsum([len(q) for q in pl], out=poff[1:])
    pflat = np.concatenate(pl) if sum(len(q) for q in pl) else np.zeros(1, np.uint8)
    stat, var, group, caps, nterm = mod_tables(static_mods, var_mods, var_caps)
    mono = np.ascontiguousarray(mono32, np.float32)
    cnt = np.zeros(max(n, 1), np.int64)
    rc = L.vl_expand_count(n, _p(seq), _p(off), _p(stat), _p(var), _p(group), _p(caps), len(caps), max_var, nterm,
                           min_mass, max_mass, _p(RESIDUE_F64), H2O_F64, _p(cnt))
    if rc != 0:
        raise RuntimeError(f"vl_expand_count: error {rc}")
    oo = np.zeros(n + 1, np.int64)
    np.cumsum(cnt[:n], out=oo[1:])
    tot = int(oo[-1])
    out = {"seq": np.empty(tot, np.uint32), "mask_f": np.empty(tot, np.uint64), "mask_d": np.empty(tot, np.uint64),
           "mass": np.empty(tot, np.float64), "n_var": np.empty(tot, np.uint8), "id_f": np.empty(tot, np.uint64),
           "id_d": np.empty(tot, np.uint64), "mass32_f": np.empty(tot, np.float32),
           "mass32_d": np.empty(tot, np.float32)}
    bufs = {k: (v if tot else np.empty(1, v.dtype)) for k, v in out.items()}
    rc = L.vl_expand_fill(n, _p(seq), _p(off), _p(pflat), _p(poff), _p(stat), _p(var), _p(group), _p(caps),
                          len(caps), max_var, nterm, min_mass, max_mass, _p(RESIDUE_F64), H2O_F64, _p(mono), h2o32,
                          id_seed, _p(oo), *[_p(bufs[k]) for k in ("seq", "mask_f", "mask_d", "mass", "n_var",
                                                                   "id_f", "id_d", "mass32_f", "mass32_d")])
    if rc != 0:
        raise RuntimeError(f"vl_expand_fill: error {rc}")
    return out
