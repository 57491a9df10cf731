"""vh_k0_noalloc (K0 with its scan scratch in the library's static device array, no allocation) against vh_k0
(caller workspace) through the C ABI: bit-identical peak views on two calls in a row, the second on a smaller
batch, so the second call runs on scratch the first one left behind; both also equal the oracle's K0."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)


def _spectra(n, seed):
    from vulcan.gpu.host import Spectra
    rng = np.random.default_rng(seed)
    npk = rng.integers(0, 700, n)                  # empty spectra, > 512 in-range peaks, everything between
    npk[::17] = 0
    off = np.concatenate([[0], np.cumsum(npk)]).astype(np.int64)
    m = int(off[-1])
    mz = rng.uniform(100.0, 2100.0, m).astype(np.float32)
    it = rng.lognormal(0.0, 1.0, m).astype(np.float32)
    it[rng.random(m) < 0.05] = 1.0                 # intensity ties
    mz[rng.random(m) < 0.01] = np.nan
    return Spectra("noalloc", np.arange(n, dtype=np.uint32), np.arange(n, dtype=np.uint32), np.zeros(n),
                   np.full(n, 500.0), np.full(n, 2, np.int8), mz, it, off)


def _k0_both(L, K, vp, sp, dev):
    n = len(sp)
    cap = max(n * vp.max_peaks, 1)
    raw = {"mz": torch.from_numpy(sp.mz).to(dev), "int": torch.from_numpy(sp.intensity).to(dev),
           "off": torch.from_numpy(sp.offsets).to(dev)}
    outs = []
    for noalloc in (False, True):
        o = {"off": torch.full((n + 1,), -7, dtype=torch.int64, device=dev),
             "mass": torch.full((cap,), -1.0, dtype=torch.float32, device=dev),
             "int": torch.full((cap,), -1.0, dtype=torch.float32, device=dev),
             "tic": torch.full((max(n, 1),), -1.0, dtype=torch.float32, device=dev),
             "valid": torch.full((max(n, 1),), 9, dtype=torch.uint8, device=dev)}
        head = (C.byref(vp), n, K.ptr(raw["mz"]), K.ptr(raw["int"]), K.ptr(raw["off"]), K.ptr(o["off"]),
                K.ptr(o["mass"]), K.ptr(o["int"]), K.ptr(o["tic"]), K.ptr(o["valid"]), cap)
        if noalloc:
            rc = L.vh_k0_noalloc(*head, K.stream())
        else:
            ws = torch.empty(int(L.vh_scan_workspace_bytes(max(n, 1))), dtype=torch.uint8, device=dev)
            rc = L.vh_k0(*head, K.ptr(ws), ws.numel(), K.stream())
        K.check(rc, "vh_k0_noalloc" if noalloc else "vh_k0")
        torch.cuda.synchronize()
        m = int(o["off"][n].item())
        outs.append({"off": o["off"].cpu().numpy(), "mass": o["mass"][:m].cpu().numpy(),
                     "int": o["int"][:m].cpu().numpy(), "tic": o["tic"][:n].cpu().numpy(),
                     "valid": o["valid"][:n].cpu().numpy()})
    return outs


def test_k0_noalloc_equals_k0_on_two_calls():
    from oracle import oracle as O
    from vulcan.gpu import _lib as K
    from vulcan.gpu._abi import make_params
    from vulcan.gpu.host import SearchParams
    L = K.lib()
    dev = torch.device("cuda:0")
    vp = make_params(SearchParams(), {"C": 57.021464}, {})
    for n, seed in ((9000, 1), (2500, 2)):         # the second call is smaller: reuses the first call's scratch
        print(f"[gpu] vh_k0 vs vh_k0_noalloc: {n} spectra, seed {seed}", flush=True)
        sp = _spectra(n, seed)
        a, b = _k0_both(L, K, vp, sp, dev)
        for k in a:
            assert np.array_equal(a[k], b[k], equal_nan=True), k
        ref = O.k0(vp, sp)
        assert np.array_equal(b["off"], ref["off"])
        assert np.array_equal(b["mass"], ref["mass"]) and np.array_equal(b["int"], ref["int"])
        assert np.array_equal(b["valid"].astype(bool), ref["valid"].astype(bool))
        assert np.allclose(b["tic"], ref["tic"], rtol=1e-6)


def test_k0_noalloc_limits():
    from vulcan.gpu import _lib as K
    from vulcan.gpu._abi import make_params
    from vulcan.gpu.host import SearchParams
    L = K.lib()
    vp = make_params(SearchParams(), {}, {})
    too_many = (1 << 24) + 1
    assert L.vh_k0_noalloc(C.byref(vp), too_many, None, None, None, None, None, None, None, None, 0,
                           K.stream()) == -4
    assert L.vh_k0_noalloc(C.byref(vp), -1, None, None, None, None, None, None, None, None, 0, K.stream()) == -1
    off = torch.full((1,), 5, dtype=torch.int64, device="cuda:0")
    dummy = torch.zeros(1, device="cuda:0")
    u8 = torch.zeros(1, dtype=torch.uint8, device="cuda:0")
    K.check(L.vh_k0_noalloc(C.byref(vp), 0, K.ptr(dummy), K.ptr(dummy), K.ptr(off), K.ptr(off), K.ptr(dummy),
                            K.ptr(dummy), K.ptr(dummy), K.ptr(u8), 0, K.stream()), "empty batch")
    torch.cuda.synchronize()
    assert int(off[0].item()) == 0
