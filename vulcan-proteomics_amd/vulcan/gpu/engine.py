"""Closed search on the MI355X: K0 (peak view) -> hypotheses -> K1 (fragment index, top-K, pair closure) -> K2
(exact rescoring) -> ranking. Call surface of the reference's vulcan/gpu/engine.py (GpuIndex, k0_preprocess,
run_k1, select_topk, run_k2, lnfact, psm_table, search, window_candidates, k1_brute, search_exhaustive, Stats,
Timer) so callers switch without changes.

Every compute step is a HIP kernel of libvulcan_hip.so; torch tensors are only device buffers. The fused
device pipeline (`DeviceSearch`) runs K0 -> rank 1 with no host synchronisation and no allocation per call, so
it can be captured in a hipGraph. Nothing here reads labels: candidates are rows of the content-sorted table.
"""
from __future__ import annotations

import ctypes as C
import resource
import time
from dataclasses import dataclass, field

import numpy as np
import pyarrow as pa
import torch

from . import _lib as K
from ._abi import (VhBlockView, VhIndexView, index_geometry, make_params, max_hyp_per_spectrum, peak_capacity,
                   set_modes)
from .host import NEUTRON, PROTON, SAGE_MONO, SearchParams, Spectra, make_hypotheses, pre_hypotheses  # noqa: F401

ROW_BITS = 28   # peptidoform rows < 2^28 (engine.py:35 of the reference)
WIDE_BUCKET_LOG2 = 14   # VH_K1W_BUCKET_LOG2: bucket rows of the wide-window K1's index view


def _u8(t):
    return t.view(torch.uint8) if t.dtype != torch.uint8 else t


@dataclass
class Slice:
    """One resident posting directory (the whole index is one slice here; 288 GB HBM holds it)."""
    chunk0: int
    chunk1: int
    offsets: torch.Tensor    # int32 bit pattern of u32, n_bins * n_buckets + 1 (bin-major)
    postings: torch.Tensor   # int32 bit pattern of (f32 key & (2^bin_shift - 1)) << 16 | row in bucket


@dataclass
class Stats:
    h2d_bytes: int = 0
    gpu_ms: dict = field(default_factory=dict)
    wall_s: dict = field(default_factory=dict)
    counts: dict = field(default_factory=dict)


class GpuIndex:
    """Device-resident peptidoform table + fragment index (reference engine.py:59-153).

    Buckets of `chunk_size` consecutive rows (precursor-mass intervals); inside a bucket the postings are
    sorted by the exact f32 fragment mass (its bit pattern) by a device LSD radix sort (vh_index_build), with a
    directory of f32-bit-pattern bins for the first posting of a mass range."""

    def __init__(self, l1, params: SearchParams, static_mods: dict, var_mods: dict, device="cuda"):
        t0 = time.perf_counter()
        self.p = params
        self.dev = torch.device(device)
        L = K.lib()
        self.vp = make_params(params, static_mods, var_mods)
        if params.fragment_min_mz <= 0:
            raise ValueError("fragment_min_mz must be positive")
        self.n = len(l1)
        d = self.dev
        self.h2d = 0

        def up(a, dtype=None):
            a = np.ascontiguousarray(a)
            t = torch.from_numpy(a if a.flags.writeable else a.copy()).to(d)   # arrays of a loaded table are read-only
            self.h2d += t.numel() * t.element_size()
            return t
        self.mass32_np = np.ascontiguousarray(l1.mass32, np.float32)
        self.seq_off = up(l1.seq_off.astype(np.int64))
        self.length = up(l1.length.astype(np.uint8))
        self.mask = up(l1.mask.view(np.int64))
        self.mass32 = up(self.mass32_np)
        self.residues = up(l1.residues.astype(np.uint8))
        self.partner = up(l1.partner.astype(np.int64))
        if self.n >= 2 ** ROW_BITS:
            raise ValueError("more than 2^28 peptidoform rows")
        lb, s, base, n_bins, n_buckets = index_geometry(params, self.n)
        self.bucket_log2, self.bin_shift, self.bin_base, self.n_bins, self.n_buckets = lb, s, base, n_bins, n_buckets
        self.view = VhIndexView()
        v = self.view
        v.n_rows = self.n
        v.seq_off, v.length, v.mask = K.ptr(self.seq_off), K.ptr(self.length), K.ptr(self.mask)
        v.mass32, v.residues, v.partner = K.ptr(self.mass32), K.ptr(self.residues), K.ptr(self.partner)
        v.bucket_log2, v.bin_shift, v.bin_base, v.n_bins, v.n_buckets = lb, s, base, n_bins, n_buckets
        st = K.stream()
        self.post_off = torch.empty(self.n + 1, dtype=torch.int64, device=d)
        ws = torch.empty(int(L.vh_scan_workspace_bytes(self.n)), dtype=torch.uint8, device=d)
        K.check(L.vh_frag_count(C.byref(self.vp), C.byref(v), K.ptr(self.post_off), K.ptr(ws), ws.numel(), st),
                "frag_count")
        self.n_postings = int(self.post_off[-1].item())
        v.n_postings = self.n_postings
        torch.cuda.synchronize()
        self._bin = None
        self._blocks = None
        self.inv_logw = None
        self.build_s = time.perf_counter() - t0
        self.device_bytes = sum(t.numel() * t.element_size() for t in
                                [self.seq_off, self.length, self.mask, self.mass32, self.residues, self.partner,
                                 self.post_off])
        if params.k1_impl == "tile":
            self.tile_view()

    def _u32_postings(self, what: str):
        if self.n_postings >= 2 ** 32:
            raise ValueError(f"{what}: {self.n_postings} postings; the bin-major views address postings with u32 "
                             "offsets (the join K1's row-range slices hold any number)")

    def tile_view(self) -> VhIndexView:
        """The index view with the bin-major postings and directory the spectrum-major (tile) K1 reads
        (vh_index_build), built on first use: the bucket-join K1 reads the block layout instead (blocks())."""
        if self._bin is None:
            self._u32_postings("tile view")
            t0 = time.perf_counter()
            L = K.lib()
            v, d, st = self.view, self.dev, K.stream()
            nb, nn = self.n_buckets, self.n_bins
            off = torch.empty(nb * nn + 1, dtype=torch.int32, device=d)
            post = torch.empty(max(self.n_postings, 1), dtype=torch.int32, device=d)
            wsb = int(L.vh_index_workspace_bytes(self.n_postings, nb, nn))
            ws = torch.empty(wsb, dtype=torch.uint8, device=d)
            K.check(L.vh_index_build(C.byref(self.vp), C.byref(v), K.ptr(self.post_off), K.ptr(off), K.ptr(post),
                                     K.ptr(ws), wsb, st), "index_build")
            v.offsets, v.postings = K.ptr(off), K.ptr(post)
            torch.cuda.synchronize()
            del ws
            self._bin = {"offsets": off, "postings": post, "build_s": time.perf_counter() - t0}
            self.device_bytes += off.numel() * 4 + post.numel() * 4
        return self.view

    def wide_view(self) -> VhIndexView:
        """Bin-major view with buckets of 2^VH_K1W_BUCKET_LOG2 rows for the wide-window (open search) K1
        (vh_k1w_topk), built on first use with the same bins as the tile view."""
        if getattr(self, "_wide", None) is None:
            self._u32_postings("wide view")
            t0 = time.perf_counter()
            L = K.lib()
            d, st = self.dev, K.stream()
            v = VhIndexView()
            C.memmove(C.byref(v), C.byref(self.view), C.sizeof(VhIndexView))
            v.bucket_log2 = WIDE_BUCKET_LOG2
            v.n_buckets = max(1, (self.n + (1 << WIDE_BUCKET_LOG2) - 1) >> WIDE_BUCKET_LOG2)
            nb, nn = v.n_buckets, self.n_bins
            off = torch.empty(nb * nn + 1, dtype=torch.int32, device=d)
            post = torch.empty(max(self.n_postings, 1), dtype=torch.int32, device=d)
            wsb = int(L.vh_index_workspace_bytes(self.n_postings, nb, nn))
            ws = torch.empty(wsb, dtype=torch.uint8, device=d)
            K.check(L.vh_index_build(C.byref(self.vp), C.byref(v), K.ptr(self.post_off), K.ptr(off), K.ptr(post),
                                     K.ptr(ws), wsb, st), "index_build (wide)")
            v.offsets, v.postings = K.ptr(off), K.ptr(post)
            torch.cuda.synchronize()
            del ws
            self._wide = {"view": v, "offsets": off, "postings": post, "build_s": time.perf_counter() - t0}
            self.device_bytes += off.numel() * 4 + post.numel() * 4
        return self._wide["view"]

    @property
    def offsets(self) -> torch.Tensor:
        self.tile_view()
        return self._bin["offsets"]

    @property
    def postings(self) -> torch.Tensor:
        self.tile_view()
        return self._bin["postings"]

    @property
    def n_slices(self) -> int:
        """Row-range slices of the join K1's block index (reference engine.py:59-65, 113-148)."""
        return len(self.block_slices())

    @property
    def slices(self):
        self.tile_view()
        return [Slice(0, self.n_buckets, self._bin["offsets"], self._bin["postings"])]

    def _row_view(self, r0: int, r1: int) -> VhIndexView:
        """The index view of rows [r0, r1): per-row pointers advanced by r0 (residues are addressed through
        seq_off, partner rows stay table rows)."""
        v = VhIndexView()
        C.memmove(C.byref(v), C.byref(self.view), C.sizeof(VhIndexView))
        v.n_rows = r1 - r0
        v.seq_off = K.ptr(self.seq_off) + 8 * r0
        v.length = K.ptr(self.length) + r0
        v.mask = K.ptr(self.mask) + 8 * r0
        v.mass32 = K.ptr(self.mass32) + 4 * r0
        v.partner = K.ptr(self.partner) + 8 * r0
        return v

    def block_slices(self) -> list:
        """Block layout of the index for the bucket-join K1 (vh_index_build_blocks), built on first use:
        consecutive rows cut into blocks of ~2^block_postings_log2 postings (at most 2^block_rows_log2 rows),
        postings key-sorted inside a block. Rows are cut into consecutive ranges (slices) of at most
        max_slice_entries postings, each with its own block layout (u32 offsets) and its block starts in table
        rows (vh_index_rebase_rows); all slices stay resident."""
        if self._blocks is not None:
            return self._blocks
        t0 = time.perf_counter()
        L = K.lib()
        d, st = self.dev, K.stream()
        q, rb = int(self.p.block_postings_log2), int(self.p.block_rows_log2)
        cap = min(int(self.p.max_slice_entries), 2 ** 32 - 1)
        if cap < 1:
            raise ValueError("max_slice_entries must be positive")
        bounds = [0]
        while bounds[-1] < self.n:
            r0 = bounds[-1]
            target = torch.tensor([int(self.post_off[r0].item()) + cap], dtype=torch.int64, device=d)
            r1 = int(torch.searchsorted(self.post_off, target, right=True).item()) - 1
            bounds.append(min(max(r1, r0 + 1), self.n))   # a row holds < 2^16 postings: one row always fits
        out = []
        for r0, r1 in zip(bounds[:-1], bounds[1:]):
            if r0 == 0 and r1 == self.n:
                sv, po = self.view, self.post_off
            else:
                sv = self._row_view(r0, r1)
                po = torch.empty(r1 - r0 + 1, dtype=torch.int64, device=d)
                ws = torch.empty(int(L.vh_scan_workspace_bytes(r1 - r0)), dtype=torch.uint8, device=d)
                K.check(L.vh_frag_count(C.byref(self.vp), C.byref(sv), K.ptr(po), K.ptr(ws), ws.numel(), st),
                        "frag_count (slice)")
                sv.n_postings = int(po[-1].item())
            npost = int(sv.n_postings)
            wsb = int(L.vh_index_blocks_workspace_bytes(r1 - r0, npost))
            ws = torch.empty(wsb, dtype=torch.uint8, device=d)
            stats = torch.zeros(3, dtype=torch.int64, device=d)
            K.check(L.vh_index_block_count(C.byref(sv), K.ptr(po), q, rb, K.ptr(stats), K.ptr(ws), wsb, st),
                    "index_block_count")
            nb, max_rows, max_post = (int(x) for x in stats.tolist())
            b = {"row0": torch.empty(nb + 1, dtype=torch.int32, device=d),
                 "off": torch.empty(nb + 1, dtype=torch.int32, device=d),
                 "key": torch.empty(max(npost, 1), dtype=torch.int32, device=d),
                 "row": torch.empty(max(npost, 1), dtype=torch.int16, device=d)}
            K.check(L.vh_index_build_blocks(C.byref(self.vp), C.byref(sv), K.ptr(po), q, rb, nb, K.ptr(b["row0"]),
                                            K.ptr(b["off"]), K.ptr(b["key"]), K.ptr(b["row"]), K.ptr(ws), wsb, st),
                    "index_build_blocks")
            K.check(L.vh_index_rebase_rows(K.ptr(b["row0"]), nb + 1, r0, st), "index_rebase_rows")
            torch.cuda.synchronize()
            del ws
            v = VhBlockView()
            v.n_blocks, v.max_rows, v.max_postings = nb, max_rows, max_post
            v.row0, v.off, v.key, v.row = (K.ptr(b[k]) for k in ("row0", "off", "key", "row"))
            b.update(view=v, r0=r0, r1=r1, n_postings=npost)
            self.device_bytes += sum(b[k].numel() * b[k].element_size() for k in ("row0", "off", "key", "row"))
            out.append(b)
        for b in out:
            b["build_s"] = time.perf_counter() - t0
        self._blocks = out
        return out

    def blocks(self) -> VhBlockView:
        """The block view of a one-slice index (block_slices() for any number of slices)."""
        sl = self.block_slices()
        if len(sl) != 1:
            raise ValueError(f"the index has {len(sl)} row-range slices: use block_slices()")
        return sl[0]["view"]

    def block_arrays(self) -> dict:
        self.blocks()
        return self._blocks[0]

    def params_for(self, p: SearchParams):
        """VhParams for a SearchParams that shares this index's residue tables (topk / tile / tolerance may
        differ; the fragment range and min_ion_index must match the index)."""
        if (p.fragment_min_mz, p.fragment_max_mz, p.min_ion_index) != (self.p.fragment_min_mz,
                                                                       self.p.fragment_max_mz, self.p.min_ion_index):
            raise ValueError("fragment range / min_ion_index differ from the index")
        v = make_params(p, {}, {})
        for f in ("mono", "stat", "var"):
            getattr(v, f)[:] = getattr(self.vp, f)[:]
        v.nterm = self.vp.nterm
        return v


# ------------------------------------------------------------------------------------------------ spectra
def load_store(path: str, out: dict | None = None, threads: int = 4) -> Spectra:
    """The MS2 spectra of a run's store (`<run>.vspec.arrow`, reference engine.py:157-168): MS1 rows dropped,
    an unknown charge becomes 0, the run name from the store's metadata. out: optional preallocated arrays
    {"mz", "int", "off", "pmz", "z"} the peaks and precursors are written into (e.g. pinned host memory);
    threads: host threads copying the peak record batches."""
    from ..io import store
    t = store.read_table(path)
    keep = t.column("ms_level").to_numpy() == 2
    idx = None if keep.all() else np.flatnonzero(keep)
    pk = store.peaks_numpy(t, None if idx is None else keep, out, threads)

    def col(name, dtype, fill=None):
        a = t.column(name).to_numpy(zero_copy_only=False)
        if fill is not None:
            a = np.nan_to_num(np.asarray(a, np.float64), nan=fill)
        return np.array(a if idx is None else a[idx], dtype)   # writable copies of the per-spectrum columns

    pmz, z = col("precursor_mz", np.float64), col("precursor_charge", np.int8, fill=0)
    if out is not None:
        n = len(z)
        out["pmz"][:n], out["z"][:n] = pmz, z
        pmz, z = out["pmz"][:n], out["z"][:n]
    name = store.run_meta(path).get("run_name", path)
    return Spectra(name, col("scan", np.uint32), col("index", np.uint32), col("rt_min", np.float64), pmz, z,
                   pk.mz, pk.intensity, pk.offsets)


def write_store(sp: Spectra, path: str, run_meta: dict | None = None) -> dict:
    """Write MS2 spectra as a run store (the inverse of load_store; for synthetic runs and tests)."""
    from ..io import store
    w = store.StoreWriter(path, {"run_name": sp.run, **(run_meta or {})})
    n = len(sp)
    w.write({"offsets": sp.offsets, "mz": sp.mz, "intensity": sp.intensity, "scan": sp.scan, "rt_min": sp.rt,
             "precursor_mz": sp.precursor_mz, "precursor_charge": sp.charge, "ms_level": np.full(n, 2, np.uint8),
             "activation": np.full(n, 2, np.uint8)})
    return w.close()


def upload_spectra(sp: Spectra, dev, stats: Stats | None = None) -> dict:
    """Host run -> device buffers (raw peaks CSR, precursor m/z f64, charge i8)."""
    d = {"mz": torch.from_numpy(np.ascontiguousarray(sp.mz, np.float32)).to(dev),
         "int": torch.from_numpy(np.ascontiguousarray(sp.intensity, np.float32)).to(dev),
         "off": torch.from_numpy(np.ascontiguousarray(sp.offsets, np.int64)).to(dev),
         "pmz": torch.from_numpy(np.ascontiguousarray(sp.precursor_mz, np.float64)).to(dev),
         "z": torch.from_numpy(np.ascontiguousarray(sp.charge, np.int8)).to(dev),
         "n": len(sp)}
    if stats is not None:
        stats.h2d_bytes += sum(v.numel() * v.element_size() for k, v in d.items() if k != "n")
    return d


def _k0_dev(vp, raw, p: SearchParams, dev, out=None):
    L = K.lib()
    n = raw["n"]
    cap = max(n * peak_capacity(p), 1)
    o = out or {"off": torch.empty(n + 1, dtype=torch.int64, device=dev),
                "mass": torch.empty(cap, dtype=torch.float32, device=dev),
                "int": torch.empty(cap, dtype=torch.float32, device=dev),
                "tic": torch.empty(max(n, 1), dtype=torch.float32, device=dev),
                "valid_d": torch.empty(max(n, 1), dtype=torch.uint8, device=dev),
                "ws": torch.empty(int(L.vh_scan_workspace_bytes(max(n, 1))), dtype=torch.uint8, device=dev)}
    K.check(L.vh_k0(C.byref(vp), n, K.ptr(raw["mz"]), K.ptr(raw["int"]), K.ptr(raw["off"]), K.ptr(o["off"]),
                    K.ptr(o["mass"]), K.ptr(o["int"]), K.ptr(o["tic"]), K.ptr(o["valid_d"]), cap, K.ptr(o["ws"]),
                    o["ws"].numel(), K.stream()), "k0")
    return o


def k0_preprocess(sp: Spectra, p: SearchParams, dev, stats: Stats, gi: GpuIndex | None = None):
    """K0 on the device (reference engine.py:171-210). Returns device peaks (CSR) + host validity mask."""
    raw = sp if isinstance(sp, dict) else upload_spectra(sp, dev, stats)
    vp = gi.vp if gi is not None else make_params(p, {}, {})
    o = _k0_dev(vp, raw, p, dev)
    n = raw["n"]
    o["valid"] = o["valid_d"][:n].cpu().numpy().astype(bool)
    o["raw"] = raw
    return o


def device_hypotheses(gi: GpuIndex, raw: dict, peaks: dict, p: SearchParams, vp=None, out=None):
    """Precursor hypotheses on the device (vh_hypotheses; host.py:64-91 semantics). Device dict; `hyp_off`
    groups them by spectrum."""
    L = K.lib()
    dev = gi.dev
    n = raw["n"]
    vp = vp or gi.vp
    cap = max(n * max_hyp_per_spectrum(p, vp.n_offsets, vp.open_search), 1)
    o = out or {"hyp_off": torch.empty(n + 1, dtype=torch.int64, device=dev),
                "mode": torch.empty(cap, dtype=torch.uint8, device=dev),
                "off": torch.empty(cap, dtype=torch.float32, device=dev),
                "spec": torch.empty(cap, dtype=torch.int32, device=dev),
                "z": torch.empty(cap, dtype=torch.int8, device=dev),
                "iso": torch.empty(cap, dtype=torch.int8, device=dev),
                "pm": torch.empty(cap, dtype=torch.float32, device=dev),
                "M": torch.empty(cap, dtype=torch.float32, device=dev),
                "lo": torch.empty(cap, dtype=torch.int64, device=dev),
                "hi": torch.empty(cap, dtype=torch.int64, device=dev),
                "maxfz": torch.empty(cap, dtype=torch.int8, device=dev),
                "ws": torch.empty(int(L.vh_scan_workspace_bytes(max(n, 1))), dtype=torch.uint8, device=dev)}
    K.check(L.vh_hypotheses(C.byref(vp), n, K.ptr(raw["pmz"]), K.ptr(raw["z"]), K.ptr(peaks["valid_d"]),
                            K.ptr(gi.mass32), gi.n, K.ptr(o["hyp_off"]), K.ptr(o["spec"]), K.ptr(o["z"]),
                            K.ptr(o["iso"]), K.ptr(o["pm"]), K.ptr(o["M"]), K.ptr(o["lo"]), K.ptr(o["hi"]),
                            K.ptr(o["maxfz"]), K.ptr(o["mode"]), K.ptr(o["off"]), cap, K.ptr(o["ws"]), o["ws"].numel(),
                            K.stream()), "hypotheses")
    return o


def hyp_to_host(hd: dict, n_spec: int | None = None) -> dict:
    """Host copy of device hypotheses; n_spec: spectra of the batch when hd's buffers were sized for more."""
    off = hd["hyp_off"] if n_spec is None else hd["hyp_off"][:n_spec + 1]
    n = int(off[-1].item())
    out = {k: hd[k][:n].cpu().numpy() for k in ("spec", "z", "iso", "pm", "M", "lo", "hi", "maxfz", "mode", "off")
           if k in hd}
    out["hyp_off"] = off.cpu().numpy()
    return out


def _hyp_device(gi, hyp, dev):
    """Accept a host hypothesis dict (reference layout) or a device one; return device tensors + hyp_off."""
    if isinstance(hyp.get("spec"), torch.Tensor):
        return hyp
    s = np.asarray(hyp["spec"])
    n_spec = int(s.max()) + 1 if len(s) else 0
    off = np.zeros(n_spec + 1, np.int64)
    if len(s):
        if np.any(np.diff(s) < 0):
            raise ValueError("hypotheses must be grouped by spectrum (make_hypotheses order)")
        np.cumsum(np.bincount(s, minlength=n_spec), out=off[1:])
    t = {k: torch.from_numpy(np.ascontiguousarray(hyp[k])).to(dev) for k in ("spec", "maxfz", "lo", "hi", "mode")
         if k in hyp}
    t["lo"], t["hi"] = t["lo"].to(torch.int64), t["hi"].to(torch.int64)
    t["spec"] = t["spec"].to(torch.int32)
    t["maxfz"] = t["maxfz"].to(torch.int8)
    t["hyp_off"] = torch.from_numpy(off).to(dev)
    return t


class Timer:
    """Per-phase GPU events + wall clock (reference engine.py:232-247)."""

    def __init__(self, stats: Stats, name: str):
        self.stats, self.name = stats, name

    def __enter__(self):
        self.e0 = torch.cuda.Event(enable_timing=True)
        self.e1 = torch.cuda.Event(enable_timing=True)
        self.e0.record()
        self.t0 = time.perf_counter()
        return self

    def __exit__(self, *a):
        self.e1.record()
        self.e1.synchronize()
        self.stats.gpu_ms[self.name] = self.stats.gpu_ms.get(self.name, 0.0) + self.e0.elapsed_time(self.e1)
        self.stats.wall_s[self.name] = self.stats.wall_s.get(self.name, 0.0) + time.perf_counter() - self.t0


def run_k1(gi: GpuIndex, hyp, peaks, p: SearchParams, stats: Stats, k1_threshold_open: int | None = None):
    """K1 append form (reference engine.py:250-281): (o_h, o_r, o_c, scored, matched, d_spec, d_maxfz)."""
    L = K.lib()
    dev = gi.dev
    hd = _hyp_device(gi, hyp, dev)
    n_spec = hd["hyp_off"].numel() - 1
    n_h = int(hd["hyp_off"][-1].item())
    n_window = int((hd["hi"][:n_h] - hd["lo"][:n_h]).clamp(min=0).sum().item()) if n_h else 0
    cap = max(1, n_window)
    scored = torch.zeros(max(n_h, 1), dtype=torch.int32, device=dev)
    matched = torch.zeros(max(n_h, 1), dtype=torch.int64, device=dev)
    n_out = torch.zeros(1, dtype=torch.int64, device=dev)
    o_h = torch.empty(cap, dtype=torch.int32, device=dev)
    o_r = torch.empty(cap, dtype=torch.int32, device=dev)
    o_c = torch.empty(cap, dtype=torch.int32, device=dev)
    vp = gi.params_for(p)
    n_pk = peaks["off"].numel() - 1
    if n_pk < n_spec:
        raise ValueError("hypotheses reference spectra beyond the peak view")
    if "mode" in hd:
        vp.k1_threshold_open = int(p.k1_threshold if k1_threshold_open is None else k1_threshold_open)
    K.check(L.vh_k1(C.byref(vp), C.byref(gi.tile_view()), n_spec, None, K.ptr(peaks["off"]), K.ptr(peaks["mass"]),
                    K.ptr(hd["hyp_off"]), K.ptr(hd["maxfz"]), K.ptr(hd.get("mode")), K.ptr(hd["lo"]),
                    K.ptr(hd["hi"]), cap, K.ptr(n_out),
                    K.ptr(o_h), K.ptr(o_r), K.ptr(o_c), K.ptr(scored), K.ptr(matched), K.stream()), "k1")
    n = int(n_out.item())
    assert n <= cap
    stats.counts["k1_window_candidates"] = stats.counts.get("k1_window_candidates", 0) + n_window
    stats.counts["k1_appended"] = stats.counts.get("k1_appended", 0) + n
    return o_h[:n], o_r[:n], o_c[:n], scored[:n_h], matched[:n_h], hd["spec"][:n_h], hd["maxfz"][:n_h]


def select_topk(o_h, o_r, o_c, d_spec, gi: GpuIndex, p: SearchParams):
    """Top-K per spectrum + pair closure over an appended K1 list (reference engine.py:284-312), as torch
    device ops for callers that hold an appended list; `search` uses the fused vh_k1_topk kernel instead.
    Identical keys (same row and count reached from two hypotheses) order the lower hypothesis first."""
    dev = gi.dev
    spec = d_spec[o_h.long()].long()
    key = (spec << 43) | ((32767 - o_c.long().clamp(max=32767)) << ROW_BITS) | ((2 ** ROW_BITS - 1) - o_r.long())
    o = torch.argsort(o_h.long(), stable=True)
    o = o[torch.argsort(key[o], stable=True)]
    ss = spec[o]
    first = torch.searchsorted(ss, ss, side="left")
    rank = torch.arange(len(ss), device=dev) - first
    sel = o[rank < p.topk]
    h, r, c = o_h[sel].long(), o_r[sel].long(), o_c[sel].long()
    if p.pair_closure:
        pr = gi.partner[r]
        h = torch.cat([h, h])
        r = torch.cat([r, pr])
        c = torch.cat([c, torch.full_like(c, -1)])
    sp_ = d_spec[h].long()
    k = (sp_ << ROW_BITS) | r
    uk, inv = torch.unique(k, sorted=True, return_inverse=True)
    hh = torch.full((len(uk),), 2 ** 62, dtype=torch.int64, device=dev).scatter_reduce_(0, inv, h, reduce="amin")
    cc = torch.full((len(uk),), -1, dtype=torch.int64, device=dev)
    cc.scatter_reduce_(0, inv, c, reduce="amax")
    rr = uk & (2 ** ROW_BITS - 1)
    return hh, rr, cc


def _k2_out(n, dev):
    out = {"hyperscore": torch.empty(max(n, 1), dtype=torch.float64, device=dev)}
    for k in ("summed_b", "summed_y", "ppm"):
        out[k] = torch.empty(max(n, 1), dtype=torch.float32, device=dev)
    for k in ("matched_b", "matched_y", "longest_b", "longest_y"):
        out[k] = torch.empty(max(n, 1), dtype=torch.int16, device=dev)
    return out


def _k2_call(gi, vp, n_groups, grp_spec, grp_start, grp_count, stride, row, hyp, maxfz, peaks, out, order=None,
             shift=None, queue=None):
    L = K.lib()
    K.check(L.vh_k2(C.byref(vp), C.byref(gi.view), n_groups, K.ptr(order), K.ptr(grp_spec), K.ptr(grp_start),
                    K.ptr(grp_count),
                    stride, K.ptr(row), K.ptr(hyp), K.ptr(maxfz), K.ptr(shift), K.ptr(peaks["off"]), K.ptr(peaks["mass"]),
                    K.ptr(peaks["int"]), K.ptr(out["hyperscore"]), K.ptr(out["summed_b"]), K.ptr(out["summed_y"]),
                    K.ptr(out["ppm"]), K.ptr(out["matched_b"]), K.ptr(out["matched_y"]), K.ptr(out["longest_b"]),
                    K.ptr(out["longest_y"]), K.ptr(queue), K.stream()), "k2")


def candidate_spectra(h: torch.Tensor, d_spec: torch.Tensor, n_spec: int):
    """Candidates sorted by spectrum (hypothesis h[c] of spectrum d_spec[h[c]]): (spectrum per candidate, i32;
    first candidate per spectrum, i64 n_spec + 1) by vh_candidate_spectra."""
    dev = h.device
    n = len(h)
    spec = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    off = torch.empty(n_spec + 1, dtype=torch.int64, device=dev)
    K.check(K.lib().vh_candidate_spectra(n, K.ptr(h), K.ptr(d_spec), n_spec, K.ptr(spec), K.ptr(off), K.stream()),
            "candidate_spectra")
    return spec[:n], off


def run_k2(gi: GpuIndex, h, r, d_spec, d_maxfz, peaks, p: SearchParams, shift: torch.Tensor | None = None):
    """K2 over candidates sorted by spectrum (reference engine.py:315-351). shift (optional, f32 per candidate):
    observed - candidate mass; nonzero entries get shifted + unshifted ions with the delta localized on the best
    residue (open / mass-offset modes). Returns a dict of device tensors, one value per candidate, plus "spec"
    (its spectrum) and "off" (first candidate per spectrum)."""
    n = len(r)
    n_spec = peaks["off"].numel() - 1
    hh = h.to(torch.int32).contiguous()
    spec, off = candidate_spectra(hh, d_spec.to(torch.int32).contiguous(), n_spec)
    out = _k2_out(n, gi.dev)
    if n:
        sh = None if shift is None else shift.to(torch.float32).contiguous()
        _k2_call(gi, gi.params_for(p), n_spec, None, off, None, 0, r.to(torch.int32).contiguous(), hh,
                 d_maxfz.contiguous(), peaks, out, shift=sh)
    out = {k: v[:n] for k, v in out.items()}
    out["spec"], out["off"] = spec, off
    return out


def lnfact(n):
    """Stirling log-factorial with lnfact(0) = 1.0 (reference engine.py:354-358)."""
    n = np.asarray(n, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        v = n * np.log(n) - n + 0.5 * np.log(n) + 0.5 * np.log(2 * np.pi * n)
    return np.where(n == 0, 1.0, v)


def rank_groups(gi, p: SearchParams, n_groups, grp_start, grp_count, stride, row, hs, mb, my, report):
    """vh_rank: per group the top `report` candidate indices and the next hyperscore (device). grp_count None:
    grp_start holds n_groups + 1 offsets."""
    L = K.lib()
    dev = gi.dev
    sel = torch.empty(max(n_groups * report, 1), dtype=torch.int64, device=dev)
    nxt = torch.empty(max(n_groups * report, 1), dtype=torch.float64, device=dev)
    if n_groups:
        K.check(L.vh_rank(C.byref(gi.params_for(p)), n_groups, K.ptr(grp_start), K.ptr(grp_count), stride,
                          K.ptr(row), K.ptr(hs), K.ptr(mb), K.ptr(my), report, K.ptr(sel), K.ptr(nxt), K.stream()),
                "rank")
    return sel[:n_groups * report], nxt[:n_groups * report]


_PSM_TYPES = {"spec_idx": torch.int64, "pf_row": torch.int64, "rank": torch.int32, "charge": torch.int8,
              "isotope_error": torch.int8, "hyperscore": torch.float64, "delta_next": torch.float64,
              "delta_best": torch.float64, "poisson": torch.float64, "matched_b": torch.int16,
              "matched_y": torch.int16, "matched_peaks": torch.int16, "longest_b": torch.int16,
              "longest_y": torch.int16, "peptide_len": torch.int16, "k1_count": torch.int32,
              "scored_candidates": torch.int32, "hyp": torch.int32}


def psm_features(gi: GpuIndex, n_groups: int, grp_spec, sel, nxt, report: int, c_hyp, c_row, c_cnt, k2: dict,
                 hyp_off, h_z, h_iso, h_pm, scored, matched, tic) -> dict:
    """vh_psm_features: the psm_table feature columns (engine.py:361-432) of vh_rank's selections, computed on
    the device, dense [n_groups * report] (rank 0 = empty slot). Candidate arrays (c_hyp / c_row / c_cnt and the
    K2 columns) are indexed by the selections; the hypothesis arrays by c_hyp. All device tensors."""
    from ._abi import PSM_COLS, VhK2Cols, VhPsmCols
    dev = gi.dev
    q = max(n_groups * report, 1)
    cols = {k: torch.empty(q, dtype=_PSM_TYPES.get(k, torch.float32), device=dev) for k in PSM_COLS}
    if n_groups == 0:
        return {k: v[:0] for k, v in cols.items()}
    kc = {"hs": k2["hyperscore"], "sb": k2["summed_b"], "sy": k2["summed_y"], "ppm": k2["ppm"],
          "mb": k2["matched_b"], "my": k2["matched_y"], "lb": k2["longest_b"], "ly": k2["longest_y"]}
    k2c = VhK2Cols(**{k: K.ptr(v) for k, v in kc.items()})
    oc = VhPsmCols(**{k: K.ptr(v) for k, v in cols.items()})
    K.check(K.lib().vh_psm_features(C.byref(gi.view), n_groups, K.ptr(grp_spec), report, K.ptr(sel), K.ptr(nxt),
                                    K.ptr(c_row), K.ptr(c_hyp), K.ptr(c_cnt), C.byref(k2c), K.ptr(hyp_off),
                                    K.ptr(h_z), K.ptr(h_iso), K.ptr(h_pm), K.ptr(scored), K.ptr(matched),
                                    K.ptr(tic), C.byref(oc), K.stream()), "psm_features")
    return {k: v[:n_groups * report] for k, v in cols.items()}


def string_column(value: str, n: int) -> pa.Array:
    """n copies of one string as an Arrow string column, built from buffers (a Python list of n strings is
    orders of magnitude slower)."""
    b = value.encode()
    offsets = np.arange(n + 1, dtype=np.int64) * len(b)
    if offsets[-1] < 2 ** 31:
        return pa.StringArray.from_buffers(n, pa.py_buffer(offsets.astype(np.int32)), pa.py_buffer(b * n))
    return pa.array([value] * n)


def _to_host(cols: dict) -> dict:
    """Device columns -> numpy arrays with one device-to-host copy (the columns packed as bytes)."""
    names = sorted(cols, key=lambda k: -cols[k].element_size())    # widest first: every column stays aligned
    flat = [cols[k].contiguous().view(-1) for k in names]
    if not flat or sum(t.numel() for t in flat) == 0:
        return {k: t.cpu().numpy() for k, t in zip(names, flat)}
    buf = torch.cat([t.view(torch.uint8) for t in flat]).cpu().numpy()
    out, o = {}, 0
    for k, t in zip(names, flat):
        nb = t.numel() * t.element_size()
        out[k] = buf[o:o + nb].view(_NP_OF[t.dtype])
        o += nb
    return out


_NP_OF = {torch.int64: np.int64, torch.int32: np.int32, torch.int16: np.int16, torch.int8: np.int8,
          torch.uint8: np.uint8, torch.float64: np.float64, torch.float32: np.float32}


_PSM_ORDER = ("charge", "isotope_error", "expmass", "calcmass", "delta_mass_ppm", "hyperscore", "delta_next",
              "delta_best", "matched_b", "matched_y", "matched_peaks", "summed_intensity", "matched_intensity_pct",
              "average_ppm", "longest_b", "longest_y", "longest_y_pct", "peptide_len")


def _table_from_dense(sp: Spectra, l1, dense: dict):
    """PSM table (PSM_SCHEMA.md column order) from the dense device feature columns: one packed copy to the
    host, empty rank slots dropped there. Returns (table, hypothesis per PSM row)."""
    d = _to_host(dense)
    keep = d["rank"] > 0
    d = {k: v[keep] for k, v in d.items()}
    s, rows = d["spec_idx"], d["pf_row"]
    cols = {"run": string_column(sp.run, len(s)), "spec_idx": s, "scan": sp.scan[s], "rt": sp.rt[s],
            "rank": d["rank"], "pf_row": rows, "pf_id": l1.pf_id[rows]}
    for k in _PSM_ORDER:
        cols[k] = d[k]
    cols["missed_cleavages"] = l1.missed[rows]
    cols["n_var"] = l1.n_var[rows]
    for k in ("k1_count", "scored_candidates", "poisson"):
        cols[k] = d[k]
    return pa.table(cols), d["hyp"].astype(np.int64)


def _hyp_offsets(hyp: dict, n_spec: int, dev) -> torch.Tensor:
    if "hyp_off" in hyp:
        off = hyp["hyp_off"]
        return off[:n_spec + 1] if isinstance(off, torch.Tensor) else torch.as_tensor(np.asarray(off)[:n_spec + 1],
                                                                                    device=dev)
    spec = hyp["spec"].cpu().numpy() if isinstance(hyp["spec"], torch.Tensor) else np.asarray(hyp["spec"])
    off = np.zeros(n_spec + 1, np.int64)
    np.cumsum(np.bincount(spec, minlength=n_spec)[:n_spec], out=off[1:])
    return torch.as_tensor(off, device=dev)


def psm_table(sp: Spectra, gi: GpuIndex, l1, hyp, h, r, k1c, k2, scored, matched, peaks, p: SearchParams,
              report: int | None = None, return_hyp: bool = False):
    """Rank the candidates of each spectrum (vh_rank: hyperscore desc, row asc), compute the feature columns on
    the device (vh_psm_features) and build the PSM table with the reference's columns (engine.py:361-432,
    PSM_SCHEMA.md section 1); the host only drops empty slots and gathers scan / rt / pf_id / missed / n_var.
    Candidates (h, r, k1c and the K2 outputs) are sorted by spectrum; hyp: device or host hypotheses."""
    report = p.report_psms if report is None else report
    dev = gi.dev
    n_spec = len(sp)
    hd = {k: (hyp[k] if isinstance(hyp[k], torch.Tensor) else torch.as_tensor(np.ascontiguousarray(hyp[k]),
                                                                               device=dev))
          for k in ("spec", "z", "iso", "pm")}
    hd = {k: v.to(t).contiguous() for (k, v), t in zip(hd.items(), (torch.int32, torch.int8, torch.int8,
                                                                    torch.float32))}
    hh = h.to(torch.int32).contiguous()
    off = k2["off"] if "off" in k2 else candidate_spectra(hh, hd["spec"], n_spec)[1]
    row32 = r.to(torch.int32).contiguous()
    sel, nxt = rank_groups(gi, p, n_spec, off, None, 0, row32, k2["hyperscore"].contiguous(),
                           k2["matched_b"].contiguous(), k2["matched_y"].contiguous(), report)
    dense = psm_features(gi, n_spec, None, sel, nxt, report, hh, row32, k1c.to(torch.int32).contiguous(),
                         {k: v.contiguous() for k, v in k2.items()}, _hyp_offsets(hyp, n_spec, dev), hd["z"],
                         hd["iso"], hd["pm"], scored.to(torch.int32).contiguous(),
                         matched.to(torch.int64).contiguous(), peaks["tic"].to(torch.float32).contiguous())
    tbl, hyp_of_row = _table_from_dense(sp, l1, dense)
    return (tbl, hyp_of_row) if return_hyp else tbl


def write_psms(tbl: pa.Table, out_dir: str, run: str | None = None) -> str:
    """Write a PSM table as <out_dir>/<run>.psms.parquet (the reference's file naming, vulcan/gpu/run.py:66)."""
    import os

    import pyarrow.parquet as pq
    os.makedirs(out_dir, exist_ok=True)
    if run is None:
        run = tbl.column("run")[0].as_py() if tbl.num_rows else "run"
    path = os.path.join(out_dir, f"{run}.psms.parquet")
    pq.write_table(tbl, path)
    return path


def _marker(ev: dict | None):
    def mark(name):
        if ev is None:
            return None
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        ev.setdefault(name, []).append(e)
        return e
    return mark


class DeviceSearch:
    """The fused closed-search step on preallocated device buffers: vh_k0 -> vh_hypotheses -> vh_spectrum_order
    -> vh_k1_topk -> vh_k2 -> vh_rank (report 1). `run(raw)` enqueues the launches on the current stream and
    returns the per-spectrum rank-1 result; no host synchronisation, no allocation, hipGraph-capturable."""

    def __init__(self, gi: GpuIndex, p: SearchParams, n_spec: int, report: int = 1, offsets=(), open_da=None,
                 k1_threshold_open: int = 1, persistent: bool = False, order_mode: str = "mass",
                 band_log2: int = 28, k1_impl: str | None = None, join_features: bool = False):
        self.gi, self.p, self.n_spec, self.report = gi, p, n_spec, report
        # k1_impl: "tile" (vh_k1_topk, spectrum-major), "join" (vh_k1j_*, block-major) or "wide" (vh_k1w_topk,
        # workgroup per spectrum over 2^14-row buckets); "auto" picks join, and wide when an open-search window is
        # on (its windows span 10^5 - 10^6 rows: per-row slots would not pay)
        impl = k1_impl or p.k1_impl
        if impl == "auto":
            impl = "wide" if open_da is not None else "join"
        if impl not in ("tile", "join", "wide"):
            raise ValueError(impl)
        if impl == "wide" and persistent:
            raise ValueError("the wide-window K1 runs as a grid launch (no persistent mode)")
        self.k1_impl = impl
        if impl == "tile":
            gi.tile_view()     # the bin-major layout is built here, never inside a (captured) step
        if impl == "wide":
            gi.wide_view()
        # persistent: K1 and K2 run as resident waves pulling spectra from an atomic work counter.
        # order_mode: "mass" (first window row, vh_spectrum_order), "work" (mass band of 2^band_log2 rows, then
        # estimated work descending, vh_spectrum_order_work) or "fixed" (self.order is set by the caller)
        if order_mode not in ("mass", "work", "fixed"):
            raise ValueError(order_mode)
        self.order_mode, self.band_log2 = order_mode, band_log2
        self.vp = gi.params_for(p)
        self.modes = bool(len(offsets) or open_da is not None)
        if self.modes:
            set_modes(self.vp, offsets, open_da, k1_threshold_open)
        dev = gi.dev
        L = K.lib()
        n = max(n_spec, 1)
        self.stride = 2 * p.topk if p.pair_closure else p.topk
        maxh = max_hyp_per_spectrum(p, len(offsets), open_da is not None)
        cap_h = n * maxh
        cap_c = n * self.stride
        ws = int(L.vh_scan_workspace_bytes(n))
        e = lambda m, dt: torch.empty(m, dtype=dt, device=dev)  # noqa: E731
        self.peaks = {"off": e(n + 1, torch.int64), "mass": e(n * peak_capacity(p), torch.float32),
                      "int": e(n * peak_capacity(p), torch.float32), "tic": e(n, torch.float32),
                      "valid_d": e(n, torch.uint8), "ws": e(ws, torch.uint8)}
        self.hyp = {"hyp_off": e(n + 1, torch.int64), "spec": e(cap_h, torch.int32), "z": e(cap_h, torch.int8),
                    "iso": e(cap_h, torch.int8), "pm": e(cap_h, torch.float32), "M": e(cap_h, torch.float32),
                    "lo": e(cap_h, torch.int64), "hi": e(cap_h, torch.int64), "maxfz": e(cap_h, torch.int8),
                    "mode": e(cap_h, torch.uint8), "off": e(cap_h, torch.float32), "ws": e(ws, torch.uint8)}
        self.c_shift = e(cap_c, torch.float32) if self.modes else None
        self.order = e(n, torch.int32)
        self.order_ws = e(int(L.vh_spectrum_order_workspace_bytes(n)), torch.uint8)
        self.n_cand = e(n, torch.int32)
        self.c_row, self.c_hyp, self.c_cnt = e(cap_c, torch.int32), e(cap_c, torch.int32), e(cap_c, torch.int32)
        self.scored, self.matched = e(cap_h, torch.int32), e(cap_h, torch.int64)
        self.status = torch.zeros(1, dtype=torch.int32, device=dev)
        self.queue = torch.zeros(1, dtype=torch.int32, device=dev) if persistent else None
        self.k2 = _k2_out(cap_c, dev)
        self.sel = e(n * report, torch.int64)
        self.nxt = e(n * report, torch.float64)
        # join_features: the join K1 keeps per-ion counts and intensity sums per slot (vh_k1j_features)
        self.join = self._join_buffers(cap_h, maxh, join_features) if impl == "join" else None

    # ---------------------------------------------------------------------------------------- bucket-join K1
    def _join_buffers(self, cap_h: int, maxh: int, feat: bool = False) -> dict:
        """Pair / slot buffers of the join K1: one per-window slot array for the batch, pair lists per row-range
        slice of the index. Capacities start from the index's mean precursor window (rows within the precursor
        tolerance of each row's mass, averaged over the table) and are fitted to the batch on the first run (one
        host read of the totals)."""
        gi, dev = self.gi, self.gi.dev
        sl = gi.block_slices()
        L = K.lib()
        m = gi.mass32
        p = self.p
        lo = m * (1.0 + p.precursor_ppm[0] * 1e-6)
        hi = m * (1.0 + p.precursor_ppm[1] * 1e-6)
        rows = (torch.searchsorted(m, hi, right=True) - torch.searchsorted(m, lo)).float()
        mean_rows = float(rows.mean().item()) if len(m) else 1.0
        nb_tot = max(sum(b["view"].n_blocks for b in sl), 1)
        rows_per_block = max(gi.n / nb_tot, 1.0)
        n_h = max(self.n_spec, 1) * min(maxh, 4)       # typical hypotheses per spectrum; fitted on the first run
        slot_cap = int(n_h * (1.5 * mean_rows + 8)) + 4096
        pair_cap = int(n_h * (1.5 * mean_rows / rows_per_block + 2)) + 4096
        j = {"feat": feat, "cap_h": cap_h,
             "s_qe": torch.empty(max(self.n_spec, 1), dtype=torch.int32, device=dev) if feat else None,
             "h_slot": torch.empty(cap_h + 1, dtype=torch.int64, device=dev),
             "slices": [{"bv": b["view"], "blk_pairs": torch.empty(b["view"].n_blocks + 1, dtype=torch.int32,
                                                                   device=dev),
                         "ws": torch.empty(int(L.vh_k1j_workspace_bytes(b["view"].n_blocks, cap_h)),
                                           dtype=torch.uint8, device=dev)} for b in sl],
             "fitted": False}
        self._join_alloc(j, slot_cap, [max(pair_cap * b["view"].n_blocks // nb_tot, 4096) for b in sl])
        return j

    def _join_alloc(self, j: dict, slot_cap: int, pair_cap):
        """(Re)allocate the slot array and the pair lists (pair_cap: one capacity for every slice, or a list)."""
        dev = self.gi.dev
        caps = pair_cap if isinstance(pair_cap, (list, tuple)) else [pair_cap] * len(j["slices"])
        j["slot_cap"] = slot_cap
        for sj, c in zip(j["slices"], caps):
            sj["pair_cap"] = c
            sj["pairs"] = torch.empty((max(c, 1), 4), dtype=torch.int64, device=dev)    # 32-byte records
        j["counts"] = torch.empty(max(slot_cap, 1), dtype=torch.int64 if j["feat"] else torch.int32, device=dev)
        s0 = j["slices"][0]   # the one-slice names
        j["pair_cap"], j["pairs"], j["blk_pairs"] = s0["pair_cap"], s0["pairs"], s0["blk_pairs"]

    def _join_pairs(self, n: int):
        L = K.lib()
        j, hy, pk = self.join, self.hyp, self.peaks
        for sj in j["slices"]:   # every call writes the same window slot offsets (h_slot)
            K.check(L.vh_k1j_pairs(C.byref(sj["bv"]), n, K.ptr(hy["hyp_off"]), K.ptr(hy["spec"]), K.ptr(hy["lo"]),
                                   K.ptr(hy["hi"]), K.ptr(hy["maxfz"]), K.ptr(pk["off"]), K.ptr(j["s_qe"]),
                                   j["cap_h"], K.ptr(j["h_slot"]), j["slot_cap"], K.ptr(sj["blk_pairs"]),
                                   K.ptr(sj["pairs"]), sj["pair_cap"], K.ptr(self.status), K.ptr(sj["ws"]),
                                   sj["ws"].numel(), K.stream()), "k1j_pairs")

    def join_front(self, n: int, mark):
        """Pairs + per-block counts of the join K1, slice by slice into one slot array (fits the pair / slot
        capacities on the first call)."""
        L = K.lib()
        j, pk = self.join, self.peaks
        mark("k1j_pairs")
        if j["feat"]:
            K.check(L.vh_k1j_iq_shift(n, K.ptr(pk["off"]), K.ptr(pk["int"]), K.ptr(j["s_qe"]), K.stream()),
                    "k1j_iq_shift")
        self._join_pairs(n)
        if not j["fitted"]:
            tot_slots = int(j["h_slot"][j["cap_h"]].item())
            tot_pairs = [int(sj["blk_pairs"][sj["bv"].n_blocks].item()) for sj in j["slices"]]
            if tot_slots > j["slot_cap"] or any(t > sj["pair_cap"] for t, sj in zip(tot_pairs, j["slices"])):
                self._join_alloc(j, max(tot_slots * 5 // 4, j["slot_cap"]),
                                 [max(t * 5 // 4, sj["pair_cap"]) for t, sj in zip(tot_pairs, j["slices"])])
                self.status.zero_()
                self._join_pairs(n)
            j["fitted"] = True
        mark("k1j_pairs")
        mark("k1j_count")
        for sj in j["slices"]:
            K.check(L.vh_k1j_count(C.byref(self.vp), C.byref(sj["bv"]), 2 if j["feat"] else 1,
                                   K.ptr(sj["blk_pairs"]), K.ptr(sj["pairs"]), sj["pair_cap"], K.ptr(pk["mass"]),
                                   K.ptr(pk["int"]), K.ptr(j["counts"]), K.ptr(self.status), K.stream()),
                    "k1j_count")
        mark("k1j_count")

    def run(self, raw: dict, ev: dict | None = None):
        """ev (optional): stage name -> list, receives a (start, end) pair of torch.cuda.Event per stage,
        recorded on the launch stream (timing of the individual kernels)."""
        n = raw["n"]
        assert n <= self.n_spec
        mark = _marker(ev)
        self.front(raw, n, mark)
        self.candidate_stage(n, mark)
        return self.back(n, mark)

    def front(self, raw: dict, n: int, mark):
        """K0 peak view, precursor hypotheses, processing order."""
        self.peak_view(raw, n, mark)
        self.hypotheses(raw, n, mark)
        self.processing_order(n, mark)

    def peak_view(self, raw: dict, n: int, mark):
        """vh_k0 into self.peaks."""
        L = K.lib()
        pk = self.peaks
        mark("k0")
        K.check(L.vh_k0(C.byref(self.vp), n, K.ptr(raw["mz"]), K.ptr(raw["int"]), K.ptr(raw["off"]),
                        K.ptr(pk["off"]), K.ptr(pk["mass"]), K.ptr(pk["int"]), K.ptr(pk["tic"]),
                        K.ptr(pk["valid_d"]), pk["mass"].numel(), K.ptr(pk["ws"]), pk["ws"].numel(), K.stream()), "k0")
        mark("k0")

    def hypotheses(self, raw: dict, n: int, mark):
        """vh_hypotheses into self.hyp (closed, mass-offset and open hypotheses of the step's parameters)."""
        L = K.lib()
        gi, pk, hy = self.gi, self.peaks, self.hyp
        mark("hyp")
        K.check(L.vh_hypotheses(C.byref(self.vp), n, K.ptr(raw["pmz"]), K.ptr(raw["z"]), K.ptr(pk["valid_d"]),
                                K.ptr(gi.mass32), gi.n, K.ptr(hy["hyp_off"]), K.ptr(hy["spec"]), K.ptr(hy["z"]),
                                K.ptr(hy["iso"]), K.ptr(hy["pm"]), K.ptr(hy["M"]), K.ptr(hy["lo"]), K.ptr(hy["hi"]),
                                K.ptr(hy["maxfz"]), K.ptr(hy["mode"]), K.ptr(hy["off"]), hy["spec"].numel(),
                                K.ptr(hy["ws"]), hy["ws"].numel(), K.stream()), "hypotheses")
        mark("hyp")

    def load_hypotheses(self, hyp: dict, n: int):
        """Closed hypotheses computed on the host (host.make_hypotheses / pre_hypotheses layout, grouped by
        spectrum) copied into self.hyp instead of vh_hypotheses."""
        if self.modes:
            raise ValueError("host hypotheses are closed hypotheses; this step searches mass-offset / open modes")
        s = np.asarray(hyp["spec"], np.int64)
        nh = len(s)
        if nh > self.hyp["spec"].numel():
            raise ValueError("more host hypotheses than the step's capacity")
        if nh and (np.any(np.diff(s) < 0) or s[0] < 0 or s[-1] >= n):
            raise ValueError("host hypotheses must be grouped by spectrum and name spectra of this batch")
        off = np.zeros(n + 1, np.int64)
        np.cumsum(np.bincount(s, minlength=n)[:n], out=off[1:])
        hy, dev = self.hyp, self.gi.dev
        hy["hyp_off"][:n + 1].copy_(torch.from_numpy(off))
        if nh:
            for k, dt in (("spec", np.int32), ("z", np.int8), ("iso", np.int8), ("pm", np.float32),
                          ("M", np.float32), ("lo", np.int64), ("hi", np.int64), ("maxfz", np.int8)):
                hy[k][:nh].copy_(torch.from_numpy(np.ascontiguousarray(hyp[k], dt)).to(dev))
            hy["mode"][:nh].zero_()
            hy["off"][:nh].fill_(float("nan"))

    def processing_order(self, n: int, mark):
        """The spectra's processing order for the K1 / K2 launches (self.order)."""
        L = K.lib()
        hy, pk, st = self.hyp, self.peaks, K.stream()
        mark("order")
        if self.order_mode == "mass":
            K.check(L.vh_spectrum_order(n, self.gi.n, K.ptr(hy["hyp_off"]), K.ptr(hy["lo"]), K.ptr(self.order),
                                        K.ptr(self.order_ws), self.order_ws.numel(), st), "spectrum_order")
        elif self.order_mode == "work":
            K.check(L.vh_spectrum_order_work(n, K.ptr(hy["hyp_off"]), K.ptr(hy["lo"]), K.ptr(hy["hi"]),
                                             K.ptr(hy["maxfz"]), K.ptr(pk["off"]), self.band_log2, K.ptr(self.order),
                                             K.ptr(self.order_ws), self.order_ws.numel(), st), "spectrum_order_work")
        mark("order")

    def candidate_stage(self, n: int, mark):
        """Index top-K + pair closure into the candidate lists (n_cand / c_row / c_hyp / c_cnt, self.stride)."""
        L = K.lib()
        gi, vp = self.gi, self.vp
        pk, hy = self.peaks, self.hyp
        if self.k1_impl == "join":
            self.join_front(n, mark)
            j = self.join
            mark("k1j_select")
            K.check(L.vh_k1j_topk(C.byref(vp), C.byref(gi.view), n, K.ptr(self.order), K.ptr(hy["hyp_off"]),
                                  K.ptr(hy["mode"]), K.ptr(hy["lo"]), K.ptr(hy["hi"]), K.ptr(j["h_slot"]),
                                  j["slot_cap"], K.ptr(j["counts"]), self.stride, K.ptr(self.n_cand),
                                  K.ptr(self.c_row), K.ptr(self.c_hyp), K.ptr(self.c_cnt), K.ptr(self.scored),
                                  K.ptr(self.matched), K.ptr(self.status), K.ptr(self.queue), K.stream()),
                    "k1j_topk")
            mark("k1j_select")
            return
        if self.k1_impl == "wide":
            mark("k1w_topk")
            K.check(L.vh_k1w_topk(C.byref(vp), C.byref(gi.wide_view()), n, K.ptr(self.order), K.ptr(pk["off"]),
                                  K.ptr(pk["mass"]), K.ptr(hy["hyp_off"]), K.ptr(hy["maxfz"]), K.ptr(hy["mode"]),
                                  K.ptr(hy["lo"]), K.ptr(hy["hi"]), self.stride, K.ptr(self.n_cand),
                                  K.ptr(self.c_row), K.ptr(self.c_hyp), K.ptr(self.c_cnt), K.ptr(self.scored),
                                  K.ptr(self.matched), K.ptr(self.status), K.stream()), "k1w_topk")
            mark("k1w_topk")
            return
        mark("k1_topk")
        K.check(L.vh_k1_topk(C.byref(vp), C.byref(gi.tile_view()), n, K.ptr(self.order), K.ptr(pk["off"]),
                             K.ptr(pk["mass"]), K.ptr(hy["hyp_off"]), K.ptr(hy["maxfz"]), K.ptr(hy["mode"]),
                             K.ptr(hy["lo"]), K.ptr(hy["hi"]), self.stride, K.ptr(self.n_cand), K.ptr(self.c_row),
                             K.ptr(self.c_hyp), K.ptr(self.c_cnt), K.ptr(self.scored), K.ptr(self.matched),
                             K.ptr(self.status), K.ptr(self.queue), K.stream()), "k1_topk")
        mark("k1_topk")

    def back(self, n: int, mark):
        """Exact rescoring of the candidate lists and the rank-`report` selection."""
        L = K.lib()
        gi, vp = self.gi, self.vp
        pk, hy = self.peaks, self.hyp
        st = K.stream()
        mark("k2")
        if self.modes:
            K.check(L.vh_candidate_shift(n, None, K.ptr(self.n_cand), self.stride, K.ptr(self.c_row),
                                         K.ptr(self.c_hyp), K.ptr(hy["mode"]), K.ptr(hy["M"]), K.ptr(gi.mass32),
                                         K.ptr(self.c_shift), st), "candidate_shift")
        _k2_call(gi, vp, n, None, None, self.n_cand, self.stride, self.c_row, self.c_hyp, hy["maxfz"], pk, self.k2,
                 order=self.order, shift=self.c_shift, queue=self.queue)
        mark("k2")
        mark("rank")
        K.check(L.vh_rank(C.byref(vp), n, None, K.ptr(self.n_cand), self.stride, K.ptr(self.c_row),
                          K.ptr(self.k2["hyperscore"]), K.ptr(self.k2["matched_b"]), K.ptr(self.k2["matched_y"]),
                          self.report, K.ptr(self.sel), K.ptr(self.nxt), st), "rank")
        mark("rank")
        return {"sel": self.sel[:n * self.report], "next": self.nxt[:n * self.report], "n_cand": self.n_cand[:n]}

    def rank1_rows(self, n: int) -> np.ndarray:
        """Host copy of the rank-1 row per spectrum (-1: no PSM)."""
        sel = self.sel[:n * self.report].view(n, self.report)[:, 0]
        rows = self.c_row[sel.clamp(min=0)].long()        # gathered on the device: n values cross to the host
        return torch.where(sel >= 0, rows, -1).cpu().numpy()

    def n_rescored(self, n: int) -> int:
        return int(self.n_cand[:n].sum().item())


class SearchPipeline:
    """Several closed-search steps in flight at once, for a stream of batches (runs, or slices of one run).

    `depth` DeviceSearch slots, each with a stream of its own (default priority); step i runs on slot
    i mod depth. A step touches only its slot's buffers and the read-only index, so consecutive steps overlap
    on the device: one step's short launches (K0, hypotheses, order, pair lists, rank: about 40 kernels whose
    grids drain in microseconds) and the ramp and tail of its count / select / K2 grids run beside another
    step's long launches. Each step's results are those of DeviceSearch.run on the same batch. The reference's
    cohort runner overlaps runs the same way, one GPU-stage thread per CUDA stream (`--search-threads`,
    vulcan/cohort.py:665-745); here one host thread enqueues every slot, since a step never waits on the host.

    Ordering contract: a step starts after the work already enqueued on the caller's current stream when
    `run` is called (inputs written there are visible to it); its inputs must stay unchanged, and its results
    stay in its slot, until the step ends (`wait(step)` / `synchronize()`) -- the slot is rewritten by step
    + depth. No device memory is allocated per step."""

    def __init__(self, gi: GpuIndex, p: SearchParams, n_spec: int, depth: int = 3, make=None, **kw):
        """make: builds one slot (a DeviceSearch or a subclass: ClassSchedule, LearnedSearch); default
        DeviceSearch(gi, p, n_spec, **kw). Three slots measured fastest on the proteome-scale batch (two and
        four are slower: the device runs four hardware queues per process)."""
        if depth < 1:
            raise ValueError("depth must be at least 1")
        self.depth = depth
        make = make or (lambda: DeviceSearch(gi, p, n_spec, **kw))
        self.slots = [make() for _ in range(depth)]
        self.streams = [torch.cuda.Stream(device=gi.dev) for _ in range(depth)]
        self.done = [None] * depth
        self.n_steps = 0

    def run(self, raw: dict, ev: dict | None = None) -> int:
        """Enqueue one step over `raw` on the next slot; returns its step number."""
        i = self.n_steps
        k = i % self.depth
        ready = torch.cuda.Event()
        ready.record()
        s = self.streams[k]
        s.wait_event(ready)
        with torch.cuda.stream(s):
            self.slots[k].run(raw, ev)
            done = torch.cuda.Event()
            done.record()
        self.done[k] = done
        self.n_steps += 1
        return i

    def slot(self, step: int) -> DeviceSearch:
        """The DeviceSearch holding `step`'s results (the step must still be resident)."""
        if not (self.n_steps - self.depth <= step < self.n_steps):
            raise ValueError(f"step {step} is not resident (steps {max(self.n_steps - self.depth, 0)}"
                             f"..{self.n_steps - 1} are)")
        return self.slots[step % self.depth]

    def wait(self, step: int) -> DeviceSearch:
        """Make the caller's current stream wait for `step`'s end; returns its slot."""
        ds = self.slot(step)
        torch.cuda.current_stream().wait_event(self.done[step % self.depth])
        return ds

    def synchronize(self):
        for s in self.streams:
            s.synchronize()


def flat_candidates(ds: DeviceSearch, n: int) -> dict:
    """Host copy of a DeviceSearch's candidate lists flattened in (spectrum, row) order: "h", "r", "k1c",
    "spec" and the K2 columns (for checks and debugging; one packed copy, flattened with numpy)."""
    cols = {"h": ds.c_hyp, "r": ds.c_row, "k1c": ds.c_cnt, **ds.k2}
    d = _to_host({k: v[:n * ds.stride] for k, v in cols.items()})
    nc = ds.n_cand[:n].cpu().numpy().astype(np.int64)
    m = (np.arange(ds.stride)[None, :] < nc[:, None]).ravel()
    out = {k: v[m] for k, v in d.items()}
    out["spec"] = np.repeat(np.arange(n), nc)
    return out


def device_table(ds: DeviceSearch, sp: Spectra, gi: GpuIndex, l1, p: SearchParams, report: int,
                 host_hyp: bool = True):
    """PSM table from a DeviceSearch that has run on `sp`, straight from its per-spectrum candidate lists (no
    flattening): vh_rank for `report` ranks (the step's own rank-1 selection when report is 1), vh_psm_features,
    empty slots dropped on the host. Returns (table, hypothesis per PSM row, debug dict with the host hypotheses
    unless host_hyp=False and the rescored candidate count)."""
    n = len(sp)
    stt = int(ds.status.item())
    if stt & 1:
        raise RuntimeError("k1_topk: a spectrum has more hypotheses than one wave handles")
    if stt & 8:
        raise RuntimeError("k1j: pair / slot capacity exceeded")
    if report == ds.report:
        sel, nxt = ds.sel[:n * report], ds.nxt[:n * report]
    else:
        sel, nxt = rank_groups(gi, p, n, None, ds.n_cand, ds.stride, ds.c_row, ds.k2["hyperscore"],
                               ds.k2["matched_b"], ds.k2["matched_y"], report)
    hy = ds.hyp
    dense = psm_features(gi, n, None, sel, nxt, report, ds.c_hyp, ds.c_row, ds.c_cnt, ds.k2, hy["hyp_off"],
                         hy["z"], hy["iso"], hy["pm"], ds.scored, ds.matched, ds.peaks["tic"])
    tbl, hh = _table_from_dense(sp, l1, dense)
    return tbl, hh, {"hyp": hyp_to_host(hy, n) if host_hyp else None, "n_rescored": ds.n_rescored(n)}


def search(sp: Spectra, gi: GpuIndex, l1, p: SearchParams, report: int | None = None, pre_hyp=None):
    """Full closed search of one run (reference engine.py:435-464). Returns (PSM table, stats, debug).
    pre_hyp: optional (valid, hypotheses) from host.pre_hypotheses; they replace the device hypotheses when
    `valid` equals K0's validity mask (counts["pre_hyp_used"] = 1; -1 when given but mismatched, 0 when not
    given), so the results are the same either way."""
    st = Stats()
    report = p.report_psms if report is None else report
    ru0 = resource.getrusage(resource.RUSAGE_SELF)
    t0 = time.perf_counter()
    dev = gi.dev
    n = len(sp)
    raw = upload_spectra(sp, dev, st)
    ds = DeviceSearch(gi, p, n, report=1)
    mark = _marker(None)
    with Timer(st, "search_device"):
        ds.peak_view(raw, n, mark)
        used = 0
        if pre_hyp is not None:
            valid_k0 = ds.peaks["valid_d"][:n].cpu().numpy().astype(bool)
            used = 1 if np.array_equal(np.asarray(pre_hyp[0], bool), valid_k0) else -1
        if used == 1:
            ds.load_hypotheses(pre_hyp[1], n)
        else:
            ds.hypotheses(raw, n, mark)
        ds.processing_order(n, mark)
        ds.candidate_stage(n, mark)
        ds.back(n, mark)
    with Timer(st, "features"):
        tbl, _, dbg = device_table(ds, sp, gi, l1, p, report)
    ru1 = resource.getrusage(resource.RUSAGE_SELF)
    st.wall_s["total"] = time.perf_counter() - t0
    valid = ds.peaks["valid_d"][:n].cpu().numpy().astype(bool)
    st.counts.update(n_spectra=n, n_searched=int(valid.sum()), n_hypotheses=len(dbg["hyp"]["spec"]),
                     n_rescored=dbg["n_rescored"], n_psms=tbl.num_rows, pre_hyp_used=used)
    st.counts["cpu_core_s"] = (ru1.ru_utime - ru0.ru_utime) + (ru1.ru_stime - ru0.ru_stime)
    peaks = dict(ds.peaks)
    peaks["valid"] = valid
    dbg.update(peaks=peaks, device=ds)
    return tbl, st, dbg


# ------------------------------------------------------------------------------------------------ references
def window_candidates(hyp, dev, spec_subset: np.ndarray | None = None):
    """Every (hypothesis, row) in the precursor windows, sorted by (spectrum, row) (engine.py:468-480)."""
    if isinstance(hyp.get("spec"), torch.Tensor):
        hyp = hyp_to_host(hyp)
    lo, hi = np.asarray(hyp["lo"]), np.asarray(hyp["hi"])
    hsel = np.arange(len(lo))
    if spec_subset is not None:
        hsel = hsel[np.isin(hyp["spec"], spec_subset)]
    n = np.maximum(hi - lo, 0)[hsel]
    hh = np.repeat(hsel, n)
    rr = np.repeat(lo[hsel], n) + (np.arange(n.sum()) - np.repeat(np.cumsum(n) - n, n))
    s = np.asarray(hyp["spec"])[hh]
    o = np.lexsort((rr, s))
    return torch.from_numpy(hh[o].astype(np.int64)).to(dev), torch.from_numpy(rr[o].astype(np.int64)).to(dev)


def k1_brute(gi: GpuIndex, hyp, peaks, h: torch.Tensor, r: torch.Tensor, p: SearchParams) -> torch.Tensor:
    """Brute-force K1 counts for (hyp, row) pairs without the index (vh_k1_brute, kernels.cu:195-232)."""
    L = K.lib()
    dev = gi.dev
    hd = _hyp_device(gi, hyp, dev)
    out = torch.empty(max(len(h), 1), dtype=torch.int32, device=dev)
    hh = h.to(torch.int32).contiguous()
    rr = r.to(torch.int32).contiguous()
    if len(h):
        K.check(L.vh_k1_brute(C.byref(gi.params_for(p)), C.byref(gi.view), len(h), K.ptr(hh), K.ptr(rr),
                              K.ptr(hd["spec"]), K.ptr(hd["maxfz"]), K.ptr(peaks["off"]), K.ptr(peaks["mass"]),
                              K.ptr(out), K.stream()), "k1_brute")
    return out[:len(h)]


def search_exhaustive(sp: Spectra, gi: GpuIndex, l1, p: SearchParams, report: int | None = None,
                      spec_subset: np.ndarray | None = None, batch: int = 20_000_000):
    """K2 over EVERY candidate in the precursor windows (no K1 shortlist): the recall reference
    (engine.py:500-523). Candidates are rescored in spectrum-aligned batches of at most `batch` rows into one
    output buffer."""
    st = Stats()
    dev = gi.dev
    peaks = k0_preprocess(sp, p, dev, st, gi)
    hd = device_hypotheses(gi, peaks["raw"], peaks, p)
    hyp = hyp_to_host(hd, len(sp))
    o_h, o_r, o_c, scored, matched, d_spec, d_maxfz = run_k1(gi, hd, peaks, p, st)
    h, r = window_candidates(hyp, dev, spec_subset)
    n = len(h)
    hh, rr = h.to(torch.int32).contiguous(), r.to(torch.int32).contiguous()
    n_spec = len(sp)
    spec, off = candidate_spectra(hh, d_spec.to(torch.int32).contiguous(), n_spec)
    k2 = _k2_out(n, dev)
    off_h = off.cpu().numpy()
    vp = gi.params_for(p)
    s0 = 0
    while s0 < n_spec:                                   # spectra [s0, s1) hold at most `batch` candidates
        s1 = max(int(np.searchsorted(off_h, off_h[s0] + batch, side="right")) - 1, s0 + 1)
        s1 = min(s1, n_spec)
        if off_h[s1] > off_h[s0]:       # group g = spectrum s0 + g: offsets and peak CSR viewed from s0
            pk = {"off": peaks["off"][s0:], "mass": peaks["mass"], "int": peaks["int"]}
            _k2_call(gi, vp, s1 - s0, None, off[s0:], None, 0, rr, hh, d_maxfz.contiguous(), pk, k2)
        s0 = s1
    k2 = {k: v[:n] for k, v in k2.items()}
    k2["spec"], k2["off"] = spec, off
    c = torch.full((n,), -1, dtype=torch.int32, device=dev)
    tbl = psm_table(sp, gi, l1, hyp, hh, rr, c, k2, scored, matched, peaks, p, report)
    return tbl, st, {"peaks": peaks, "hyp": hyp, "h": h, "r": r, "k2": k2}
