"""ctypes binding of libvulcan_hip.so (csrc/*.hip, C ABI in include/vulcan_hip.h).

Same surface as the reference's vulcan/gpu/_lib.py:22-76: `lib()` (re)builds the library when the recorded
source hash differs and loads it with ctypes, `ptr()` checks device + contiguity, `stream()` is torch's
current stream (a hipStream_t on ROCm), `check()` raises RuntimeError on a nonzero return.
There is no CPU fallback: without the library or a ROCm device every call fails loudly.
"""
from __future__ import annotations

import ctypes as C

import torch

from .. import _build
from ._abi import VhBlockView, VhFamilyCols, VhIndexView, VhK2Cols, VhParams, VhPsmCols

P = C.c_void_p
I32, I64, SZ, F32, F64 = C.c_int32, C.c_int64, C.c_size_t, C.c_float, C.c_double
PP = C.POINTER(VhParams)
PIX = C.POINTER(VhIndexView)
PBV = C.POINTER(VhBlockView)
PK2 = C.POINTER(VhK2Cols)
PPSM = C.POINTER(VhPsmCols)
PFAM = C.POINTER(VhFamilyCols)

_SIGS = {
    "vh_version": ([], I32),
    "vh_abi_sizeof": ([I32], SZ),
    "vh_scan_workspace_bytes": ([I64], SZ),
    "vh_copy_u32": ([P, P, I64, P], I32),
    "vh_frag_count": ([PP, PIX, P, P, SZ, P], I32),
    "vh_index_workspace_bytes": ([I64, I64, I32], SZ),
    "vh_index_build": ([PP, PIX, P, P, P, P, SZ, P], I32),
    "vh_index_blocks_workspace_bytes": ([I64, I64], SZ),
    "vh_index_block_count": ([PIX, P, I32, I32, P, P, SZ, P], I32),
    "vh_index_build_blocks": ([PP, PIX, P, I32, I32, I64, P, P, P, P, P, SZ, P], I32),
    "vh_index_rebase_rows": ([P, I64, I32, P], I32),
    "vh_k0": ([PP, I32, P, P, P, P, P, P, P, P, I64, P, SZ, P], I32),
    "vh_k0_noalloc": ([PP, I32, P, P, P, P, P, P, P, P, I64, P], I32),
    "vh_hypotheses": ([PP, I32, P, P, P, P, I64, P, P, P, P, P, P, P, P, P, P, P, I64, P, SZ, P], I32),
    "vh_spectrum_order_workspace_bytes": ([I32], SZ),
    "vh_spectrum_order": ([I32, I64, P, P, P, P, SZ, P], I32),
    "vh_spectrum_order_work": ([I32, P, P, P, P, P, I32, P, P, SZ, P], I32),
    "vh_k1": ([PP, PIX, I32, P, P, P, P, P, P, P, P, I64, P, P, P, P, P, P, P], I32),
    "vh_k1_topk": ([PP, PIX, I32, P, P, P, P, P, P, P, P, I32, P, P, P, P, P, P, P, P, P], I32),
    "vh_k1w_topk": ([PP, PIX, I32, P, P, P, P, P, P, P, P, I32, P, P, P, P, P, P, P, P], I32),
    "vh_k1_brute": ([PP, PIX, I64, P, P, P, P, P, P, P, P], I32),
    "vh_k2": ([PP, PIX, I32, P, P, P, P, I32, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P], I32),
    "vh_candidate_shift": ([I32, P, P, I32, P, P, P, P, P, P, P], I32),
    "vh_candidate_spectra": ([I64, P, P, I32, P, P, P], I32),
    "vh_rank": ([PP, I32, P, P, I32, P, P, P, P, I32, P, P, P], I32),
    "vh_psm_features": ([PIX, I32, P, I32, P, P, P, P, P, PK2, P, P, P, P, P, P, P, PPSM, P], I32),
    "vh_k1_features": ([PP, PIX, I32, P, P, P, P, P, P, P, P, P, I32, P, P, P, P, P, P, P, P, P, P, P, P], I32),
    "vh_k1j_workspace_bytes": ([I64, I64], SZ),
    "vh_k1j_pairs": ([PBV, I32, P, P, P, P, P, P, P, I64, P, I64, P, P, I64, P, P, SZ, P], I32),
    "vh_k1j_count": ([PP, PBV, I32, P, P, I64, P, P, P, P, P], I32),
    "vh_k1j_iq_shift": ([I32, P, P, P, P], I32),
    "vh_k1_candidate_features": ([PP, PIX, I32, P, P, P, P, P, P, P, P, I32, P, P, P, P, P, P, P, P], I32),
    "vh_k1j_topk": ([PP, PIX, I32, P, P, P, P, P, P, I64, P, I32, P, P, P, P, P, P, P, P, P], I32),
    "vh_k1j_features": ([PP, PIX, I32, P, P, P, P, P, P, I64, P, I32, P, P, P, P, P, P, P, P, P, P, P, P], I32),
    "vh_mlp_pack": ([P, P, P, P, P], I32),
    "vh_mlp_forward": ([P, I64, P, P, P, P, P], I32),
    "vh_mlp_workspace_bytes": ([I64], SZ),
    "vh_mlp_backward": ([P, I64, P, P, P, P, P, P, P, SZ, P], I32),
    "vh_mlp_adam": ([P, P, P, P, I64, F32, F32, F32, F32, F32, F32, P], I32),
    "vh_prefilter_count_features": ([I32, I32, P, P, P, P, P, P], I32),
    "vh_prefilter_features": ([PIX, I32, I32, P, P, P, P, P, P, P, P, P, P, P], I32),
    "vh_prefilter_score_workspace_bytes": ([I32, I32], SZ),
    "vh_prefilter_score": ([PIX, P, I32, I32, P, P, P, P, P, P, P, P, P, P, P, SZ, P], I32),
    "vh_prefilter_select": ([PIX, I32, I32, P, P, P, P, P, I32, I32, F32, I32, I32, P, P, P, P, P], I32),
    "vh_gbm_decision": ([P, I64, P, I32, P, P, I32, I32, F64, P, I64, I32, P, P, P], I32),
    "vh_family_select_workspace_bytes": ([I32], SZ),
    "vh_family_select": ([PP, I32, P, P, P, P, I64, P, P, I64, P, I32, I32, P, P, P, P, P, P, P, P, SZ, P], I32),
    "vh_family_wide_stats_workspace_bytes": ([I32], SZ),
    "vh_family_wide_stats": ([PP, I32, P, P, P, P, P, I64, P, I64, P, P, P, P, SZ, P], I32),
    "vh_family_wide_workspace_bytes": ([I64, I32], SZ),
    "vh_family_wide": ([PP, I32, P, P, P, P, P, I64, P, P, I32, I64, P, I32, P, P, P, C.c_uint64, C.c_uint64, P, P,
                        P, P, P, SZ, P], I32),
    "vh_family_wide_copy": ([I32, P, P, P, P, P, P, P, P, P, P, P], I32),
    "vh_family_top2_workspace_bytes": ([I32], SZ),
    "vh_family_top2": ([PP, I32, P, P, P, P, P, P, P, P, P, P, P, I64, P, I32, I32, I32, P, PFAM, P, SZ, P], I32),
}

ERRORS = {-1: "bad argument", -2: "output capacity too small", -3: "workspace too small", -4: "size limit exceeded"}

_lib = None


class _Strict:
    """ctypes accepts surplus arguments silently (they would land in the wrong stack slots, e.g. a truncated
    stream handle); every call here must pass exactly the declared parameters."""

    def __init__(self, cdll):
        self._cdll = cdll
        for name, (args, res) in _SIGS.items():
            f = getattr(cdll, name)
            f.argtypes = args
            f.restype = res
            setattr(self, name, self._wrap(name, f, len(args)))

    @staticmethod
    def _wrap(name, f, n):
        def call(*a):
            if len(a) != n:
                raise TypeError(f"{name} takes {n} arguments, {len(a)} given")
            return f(*a)
        call.__name__ = name
        return call


def load(path: str | None = None):
    """Load the library without touching the GPU (symbol/ABI checks on CPU-only hosts)."""
    return _Strict(C.CDLL(path or _build.build_hip(force=False)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not torch.cuda.is_available():
        raise RuntimeError("libvulcan_hip needs a ROCm device (torch.cuda.is_available() is False)")
    torch.cuda.init()
    _lib = load()
    if _lib.vh_abi_sizeof(0) != C.sizeof(VhParams) or _lib.vh_abi_sizeof(1) != C.sizeof(VhIndexView):
        raise RuntimeError("VhParams / VhIndexView layout differs between Python and libvulcan_hip.so")
    return _lib


def ptr(t: torch.Tensor | None):
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous(), "kernels need contiguous device tensors"
    return t.data_ptr()


def stream():
    return torch.cuda.current_stream().cuda_stream


def check(code: int, what: str):
    if code != 0:
        msg = ERRORS.get(code, f"HIP error {code}")
        raise RuntimeError(f"{what}: {msg} ({code})")
