"""The C-ABI library loads on a CPU-only host and exports every symbol include/vulcan_hip.h declares;
the ctypes struct mirror matches the C layout (no GPU calls)."""
import ctypes as C
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "vulcan_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|size_t)\s+(vh_\w+)\(", text, flags=re.M)))


def test_header_declares_the_hot_path():
    names = declared_symbols()
    for need in ("vh_frag_count", "vh_index_build", "vh_k0", "vh_hypotheses", "vh_k1", "vh_k1_topk", "vh_k1_brute",
                 "vh_k2", "vh_rank", "vh_version", "vh_abi_sizeof", "vh_k1_features", "vh_mlp_pack",
                 "vh_mlp_forward", "vh_mlp_backward", "vh_mlp_adam", "vh_prefilter_features", "vh_prefilter_score",
                 "vh_prefilter_score_workspace_bytes", "vh_psm_features",
                 "vh_prefilter_select", "vh_index_block_count", "vh_index_build_blocks", "vh_k1j_pairs",
                 "vh_k1j_count", "vh_k1j_topk", "vh_k1j_features", "vh_gbm_decision",
                 "vh_family_select"):
        assert need in names


def test_library_exports_every_declared_symbol():
    from vulcan.gpu import _lib
    L = _lib.load()
    for name in declared_symbols():
        assert hasattr(L._cdll, name), name
    assert set(declared_symbols()) <= set(_lib._SIGS), "binding lacks a signature"


def test_abi_layout_and_version():
    from vulcan.gpu import _lib
    from vulcan.gpu._abi import VhBlockView, VhIndexView, VhK2Cols, VhParams, VhPsmCols
    L = _lib.load()
    assert L.vh_abi_sizeof(0) == C.sizeof(VhParams)
    assert L.vh_abi_sizeof(1) == C.sizeof(VhIndexView)
    assert L.vh_abi_sizeof(2) == C.sizeof(VhBlockView)
    assert L.vh_abi_sizeof(3) == C.sizeof(VhK2Cols) == 64
    assert L.vh_abi_sizeof(4) == C.sizeof(VhPsmCols) == 25 * 8
    from vulcan.gpu._abi import GBM_NODE
    assert L.vh_abi_sizeof(5) == GBM_NODE.itemsize == 16
    from vulcan.gpu._abi import VhFamilyCols
    assert L.vh_abi_sizeof(6) == C.sizeof(VhFamilyCols) == 80
    assert L.vh_abi_sizeof(7) == 0
    m = re.search(r"#define VH_ABI_VERSION (\d+)", open(HEADER).read())
    assert L.vh_version() == int(m.group(1))


def test_argument_errors_are_reported_without_a_gpu():
    """Null / inconsistent arguments are rejected on the host before any launch (negative VH_E_* codes)."""
    from vulcan.gpu import _lib
    from vulcan.gpu._abi import VhBlockView, VhIndexView, make_params
    from vulcan.gpu.host import SearchParams
    L = _lib.load()
    vp = make_params(SearchParams(), {}, {})
    ix = VhIndexView()
    assert L.vh_k1_topk(C.byref(vp), C.byref(ix), 1, None, None, None, None, None, None, None, None, 2, None, None,
                        None, None, None, None, None, None, None) == -1
    assert L.vh_k2(C.byref(vp), C.byref(ix), 1, None, None, None, None, 0, None, None, None, None, None, None, None,
                   None, None, None, None, None, None, None, None, None, None) == -1
    assert L.vh_spectrum_order(3, 10, None, None, None, None, 0, None) == -1
    assert L.vh_mlp_pack(None, None, None, None, None) == -1
    assert L.vh_mlp_forward(None, 10, None, None, None, None, None) == -1
    assert L.vh_mlp_forward(C.c_void_p(8), 10, C.c_void_p(8), C.c_void_p(8), C.c_void_p(8), None, None) == -1
    assert L.vh_mlp_backward(C.c_void_p(8), 100, *[C.c_void_p(8)] * 7, 16, None) == -3
    assert L.vh_mlp_workspace_bytes(-1) == 0 and L.vh_mlp_workspace_bytes(1000) > 1000 * 256 * 6
    assert L.vh_psm_features(C.byref(ix), 1, None, 0, *[None] * 6, *[None] * 7, None, None) == -1  # report 0
    assert L.vh_psm_features(C.byref(ix), 1, None, 1, *[None] * 6, *[None] * 7, None, None) == -1  # no columns
    assert L.vh_prefilter_score_workspace_bytes(-1, 8) == 0 and L.vh_prefilter_score_workspace_bytes(10, 0) == 0
    assert L.vh_prefilter_score_workspace_bytes(1000, 200) >= 1001 * 4 + (1000 * 200 // 64) * 4
    ws_small = L.vh_prefilter_score_workspace_bytes(1000, 200) - 1
    # pointers that are never dereferenced: both calls return before any launch
    assert L.vh_prefilter_score(C.byref(ix), 1, 1000, 200, *[1] * 10, 1, ws_small, None) == -3   # workspace short
    assert L.vh_prefilter_score(C.byref(ix), 1, 1000, 200, *[1] * 10, None, 10 ** 9, None) == -1  # no workspace
    assert L.vh_prefilter_select(C.byref(ix), 1, 2000, *[None] * 5, 16, 16, 0.0, 1, 32, *[None] * 4, None) == -1
    assert L.vh_prefilter_select(C.byref(ix), 1, 200, *[None] * 5, 0, 16, 0.0, 1, 32, *[None] * 4, None) == -1
    # GBM scorer: no status word / too many features / negative depth; pointers never dereferenced
    assert L.vh_gbm_decision(1, 10, 1, 1, 1, 1, 1, 3, 0.0, 1, 5, 7, 1, None, None) == -1
    assert L.vh_gbm_decision(1, 10, 1, 1, 1, 1, 1, 3, 0.0, 1, 5, 64, 1, 1, None) == -4
    assert L.vh_gbm_decision(1, 10, 1, 1, 1, 1, 1, -1, 0.0, 1, 5, 7, 1, 1, None) == -1
    # family select: bad mode / decoy-set count / table too large / scan workspace short / no wide list;
    # pointers are never dereferenced (every call returns before a launch)
    assert L.vh_family_select(C.byref(vp), 4, *[1] * 4, 10, 1, 1, 100, 1, 2, 3, *[1] * 7, 1, 0, None) == -1
    assert L.vh_family_select(C.byref(vp), 4, *[1] * 4, 10, 1, 1, 100, 1, 0, 0, *[1] * 7, 1, 0, None) == -1
    assert L.vh_family_select(C.byref(vp), 4, *[1] * 4, 10, 1, 1, 1 << 28, 1, 2, 0, *[1] * 7, 1, 0, None) == -4
    assert L.vh_family_select(C.byref(vp), 4, *[1] * 4, 10, 1, 1, 100, 1, 2, 2, *[1] * 7, 1, 0, None) == -3
    assert L.vh_family_select(C.byref(vp), 4, *[1] * 4, 10, 1, 1, 100, 1, 2, 0, *[1] * 5, None, 1, 1, 0,
                              None) == -1
    assert L.vh_family_select_workspace_bytes(1000) > 0
    # wide path: workspace short for the listed rows; top-2 table: too many retention groups
    import numpy as np
    w_host = np.array([0, 5000], np.int64)
    assert L.vh_family_wide(C.byref(vp), 1, *[1] * 5, 10, 1, 1, 9, 100, 1, 2, 1, w_host.ctypes.data, 1, 10, 10,
                            *[1] * 4, 1, 16, None) == -3
    assert L.vh_family_wide_workspace_bytes(5000, 2) > 2 * 5000 * 8 + 2 * 15000 * 8
    assert L.vh_family_top2(C.byref(vp), 4, *[1] * 11, 100, 1, 8, 17, 0, 1, None, 1, 0, None) == -4
    assert L.vh_family_top2(C.byref(vp), 4, *[1] * 11, 100, 1, 2, 3, 3, 1, None, 1, 0, None) == -1
    assert L.vh_family_top2(C.byref(vp), 4, *[1] * 11, 100, 1, 2, 3, 1, 1, None, 1, 0, None) == -1   # no columns
    assert L.vh_prefilter_select(C.byref(ix), 1, 200, *[None] * 5, 16, 8, 0.0, 1, 32, *[None] * 4, None) == -1
    assert L.vh_prefilter_select(C.byref(ix), 1, 200, *[None] * 5, 8, 16, -1.0, 1, 32, *[None] * 4, None) == -1
    ixp = VhIndexView()
    ixp.partner = 8
    assert L.vh_prefilter_select(C.byref(ixp), 1, 200, *[None] * 5, 16, 16, 0.0, 1, 31, *[None] * 4, None) == -2
    assert L.vh_k1_features(C.byref(vp), C.byref(ix), 1, *[None] * 9, 10, *[None] * 11, None) == -1
    # block layout / join K1
    ix.n_rows = 100
    assert L.vh_index_block_count(C.byref(ix), C.c_void_p(8), 5, 10, C.c_void_p(8), None, 0, None) == -1   # q < 6
    assert L.vh_index_block_count(C.byref(ix), C.c_void_p(8), 12, 10, C.c_void_p(8), None, 0, None) == -3
    assert L.vh_index_blocks_workspace_bytes(1000, 5000) > 2 * 5000 * 8
    bv = VhBlockView()
    assert L.vh_k1j_pairs(None, 1, *[None] * 7, 10, None, 0, None, None, 0, None, None, 0, None) == -1
    assert L.vh_k1j_pairs(C.byref(bv), 1, *[None] * 7, 10, None, 0, None, None, 0, None, None, 0, None) == -1
    bv.n_blocks, bv.max_rows, bv.max_postings, bv.row0, bv.off = 4, 64, 70000, 8, 8
    assert L.vh_k1j_count(C.byref(vp), C.byref(bv), 1, None, None, 0, *[None] * 4, None) == -4
    bv.max_postings = 4096
    assert L.vh_k1j_count(C.byref(vp), C.byref(bv), 3, None, None, 0, *[None] * 4, None) == -1
    assert L.vh_k1j_iq_shift(2, None, None, None, None) == -1
    assert L.vh_k1j_workspace_bytes(1000, 5000) >= 4000
    assert L.vh_k1j_topk(C.byref(vp), C.byref(ix), 1, *[None] * 6, 0, None, 2, *[None] * 8, None) == -1
    assert L.vh_k1j_features(C.byref(vp), C.byref(ix), 1, *[None] * 6, 0, None, 2, *[None] * 11, None) == -1
    vp.max_peaks = 4096
    assert L.vh_k0(C.byref(vp), 1, None, None, C.c_void_p(8), C.c_void_p(8), None, None, C.c_void_p(8),
                   C.c_void_p(8), 10, None, 0, None) == -4
    # the allocation-free K0 refuses batches beyond its static scratch before touching the GPU
    assert L.vh_k0_noalloc(C.byref(vp), (1 << 24) + 1, *[None] * 8, 0, None) == -4
    assert L.vh_k0_noalloc(C.byref(vp), -1, *[None] * 8, 0, None) == -1


def _declarations():
    """name -> list of parameter type strings, parsed from the header."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"^(?:int|size_t)\s+(vh_\w+)\(([^)]*)\);", text, flags=re.M):
        params = [x.strip() for x in m.group(2).replace("\n", " ").split(",") if x.strip()]
        out[m.group(1)] = [] if params == ["void"] else params
    return out


def test_binding_argtypes_match_header_arity_and_kind():
    """Every ctypes signature has the header's parameter count, and pointer/integer/float kinds line up (a
    mismatch silently corrupts trailing stack arguments such as the stream handle)."""
    from vulcan.gpu import _lib
    decl = _declarations()
    assert set(decl) == set(_lib._SIGS)
    for name, params in decl.items():
        args = _lib._SIGS[name][0]
        assert len(args) == len(params), (name, len(args), len(params))
        for a, prm in zip(args, params):
            is_ptr = "*" in prm
            if is_ptr:
                assert a is C.c_void_p or hasattr(a, "_type_") and not isinstance(a._type_, str), (name, prm, a)
            elif prm.startswith(("int64_t", "size_t")):
                assert a in (C.c_int64, C.c_size_t), (name, prm, a)
            elif prm.startswith(("int32_t", "int ")):
                assert a is C.c_int32, (name, prm, a)


def test_binding_rejects_wrong_argument_count():
    import pytest

    from vulcan.gpu import _lib
    L = _lib.load()
    with pytest.raises(TypeError):
        L.vh_copy_u32(None, None, 0, None, None)
    assert L.vh_copy_u32(None, None, 0, None) == 0
