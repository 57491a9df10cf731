"""Precursor hypotheses against vectors written by the reference's own host rule (tests/golden/
make_hypothesis_vectors.py -> hypothesis_vectors.json; reference vulcan/gpu/host.py:64-91): spectrum, charge,
isotope error, f32 precursor / neutral masses, row windows and max fragment charge, for unknown charges, NaN
precursors, invalid spectra and masses on the table's ends and duplicates (CPU)."""
import json
import os

import numpy as np
import pytest

from vulcan.gpu.host import SearchParams, Spectra, make_hypotheses

VEC = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "hypothesis_vectors.json")))


def test_make_hypotheses_equals_the_reference_rule():
    mass32 = np.array(VEC["mass32"], np.float32)
    pmz = np.array([np.nan if x is None else x for x in VEC["precursor_mz"]], np.float64)
    n = len(pmz)
    sp = Spectra("golden", np.arange(n), np.arange(n), np.zeros(n), pmz, np.array(VEC["charge"], np.int8),
                 np.zeros(0, np.float32), np.zeros(0, np.float32), np.zeros(n + 1, np.int64))
    valid = np.array(VEC["valid"], bool)
    for case in VEC["cases"]:
        kw = {k: tuple(v) if isinstance(v, list) else v for k, v in case["params"].items()}
        got = make_hypotheses(sp, valid, mass32, SearchParams(**kw))
        want = case["out"]
        assert len(got["spec"]) == len(want["spec"]) > 0
        for k, v in want.items():
            w = np.asarray(v)
            g = np.asarray(got[k])
            if k in ("pm", "M"):                     # f32 values, compared bit for bit
                assert np.array_equal(g.astype(np.float32), w.astype(np.float32)), k
            else:
                assert np.array_equal(g.astype(np.int64), w.astype(np.int64)), k


@pytest.mark.gpu
def test_device_hypotheses_equal_the_reference_rule():
    """vh_hypotheses (device) on the same inputs gives the reference's vectors."""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("needs a ROCm GPU")
    from types import SimpleNamespace

    from vulcan.gpu import engine as E
    from vulcan.gpu._abi import make_params
    dev = torch.device("cuda:0")
    mass32 = np.array(VEC["mass32"], np.float32)
    pmz = np.array([np.nan if x is None else x for x in VEC["precursor_mz"]], np.float64)
    n = len(pmz)
    for case in VEC["cases"]:
        print(f"[gpu] golden hypotheses {case['params']}", flush=True)
        kw = {k: tuple(v) if isinstance(v, list) else v for k, v in case["params"].items()}
        p = SearchParams(**kw)
        gi = SimpleNamespace(dev=dev, mass32=torch.from_numpy(mass32).to(dev), n=len(mass32),
                             vp=make_params(p, {}, {}))
        raw = {"pmz": torch.from_numpy(pmz).to(dev), "z": torch.tensor(VEC["charge"], dtype=torch.int8, device=dev),
               "n": n}
        peaks = {"valid_d": torch.tensor(VEC["valid"], dtype=torch.uint8, device=dev)}
        hd = E.device_hypotheses(gi, raw, peaks, p)
        got = E.hyp_to_host(hd, n)
        want = case["out"]
        assert len(got["spec"]) == len(want["spec"])
        for k, v in want.items():
            w = np.asarray(v)
            g = np.asarray(got[k])
            if k in ("pm", "M"):
                assert np.array_equal(g.astype(np.float32), w.astype(np.float32)), k
            else:
                assert np.array_equal(g.astype(np.int64), w.astype(np.int64)), k



def test_pre_hypotheses_validity_equals_k0(small_plain):
    """pre_hypotheses' host validity rule (global peak view: min(#in-range peaks, max_peaks) >= min_peaks) equals
    the oracle K0's mask, NaN m/z and boundary peaks included, and its hypotheses are make_hypotheses' on it."""
    from oracle import oracle as O
    from vulcan.gpu._abi import make_params
    from vulcan.gpu.host import pre_hypotheses
    d = small_plain
    sp = d["sp"]
    rng = np.random.default_rng(5)
    mz = sp.mz.copy()
    spec_of = np.repeat(np.arange(len(sp)), np.diff(sp.offsets))
    drop = rng.uniform(0.0, 1.0, len(sp))[spec_of]                 # per-spectrum NaN share: some fall below min_peaks
    mz[rng.random(len(mz)) < drop] = np.nan
    mz[rng.random(len(mz)) < 0.01] = np.float32(d["p"].fragment_min_mz)
    mz[rng.random(len(mz)) < 0.01] = np.float32(d["p"].fragment_max_mz)
    sp2 = Spectra(sp.run, sp.scan, sp.index, sp.rt, sp.precursor_mz, sp.charge, mz, sp.intensity, sp.offsets)

    class _Gi:
        mass32_np = d["l1"].mass32
    valid, hyp = pre_hypotheses(sp2, _Gi, d["p"])
    pk = O.k0(make_params(d["p"], d["static"], d["var"]), sp2)
    assert np.array_equal(valid, pk["valid"].astype(bool))
    assert 0 < valid.sum() < len(valid)
    ref = make_hypotheses(sp2, valid, d["l1"].mass32, d["p"])
    for k in ref:
        assert np.array_equal(hyp[k], ref[k]), k
