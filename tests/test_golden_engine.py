"""K0, top-K + closure and the PSM table against vectors written by the reference's own functions
(tests/golden/make_engine_vectors.py ran the reference's vulcan/gpu/engine.py k0_preprocess, select_topk and
psm_table on CPU torch): the oracle on the CPU, the HIP path on the GPU.

The TIC is compared to 1e-6 relative (the reference sums a zero-padded f64 matrix row; the device sums per lane
then across the wave: the f64 summation order is not pinned). Everything else is compared bit for bit, except
the Poisson column: 1e-12 relative, with its -inf (overflow replaced by 1e-325 = 0) and NaN (no scored
candidate) entries at the same places."""
import hashlib
import os
import types

import numpy as np
import pytest

from conftest import STATIC

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "engine_vectors.npz")
K0_CASES = [("global150", {}), ("global50", {"max_peaks": 50, "min_peaks": 5}),
            ("window100", {"peak_mode": "window"}),
            ("window37", {"peak_mode": "window", "window_th": 37.5, "window_k": 4, "min_peaks": 10})]
TK_CASES = [("k5c", 5, True), ("k3", 3, False)]
PSM_F32 = ("expmass", "calcmass", "delta_mass_ppm", "summed_intensity", "matched_intensity_pct", "average_ppm",
           "longest_y_pct")
PSM_EXACT = ("spec_idx", "scan", "rank", "pf_row", "pf_id", "charge", "isotope_error", "matched_b", "matched_y",
             "matched_peaks", "longest_b", "longest_y", "peptide_len", "missed_cleavages", "n_var", "k1_count",
             "scored_candidates")


@pytest.fixture(scope="module")
def G():
    return dict(np.load(GOLD))


def _O():
    from oracle import oracle as O
    return O


def _params(**kw):
    from vulcan.gpu.host import SearchParams
    return SearchParams(**kw)


def _spectra(G):
    from vulcan.gpu.host import Spectra
    off = G["k0_off"]
    n = len(off) - 1
    return Spectra("golden", np.arange(n) + 1000, np.arange(n), np.linspace(0.0, 90.0, n), np.full(n, 600.0),
                   np.zeros(n, np.int8), G["k0_mz"], G["k0_int"], off)


def _check_k0(G, name, off, mass, inten, tic, valid):
    assert np.array_equal(off, G[f"k0_{name}_off"])
    assert np.array_equal(np.asarray(mass, np.float32).view(np.uint32), G[f"k0_{name}_mass"].view(np.uint32))
    assert np.array_equal(np.asarray(inten, np.float32).view(np.uint32), G[f"k0_{name}_int"].view(np.uint32))
    assert np.array_equal(np.asarray(valid, bool), G[f"k0_{name}_valid"])
    np.testing.assert_allclose(tic, G[f"k0_{name}_tic"], rtol=1e-6, atol=0)


@pytest.mark.parametrize("name,kw", K0_CASES)
def test_k0_oracle_equals_reference(G, name, kw):
    from vulcan.gpu._abi import make_params
    O = _O()
    r = O.k0(make_params(_params(**kw), {}, {}), _spectra(G))
    _check_k0(G, name, r["off"], r["mass"], r["int"], r["tic"], r["valid"].astype(bool))


def _tk_hyp(G):
    lo, hi = G["tk_lo"], G["tk_hi"]
    w = np.maximum(hi - lo, 0)
    win_off = np.concatenate([[0], np.cumsum(w)]).astype(np.int64)
    counts = np.zeros(max(int(win_off[-1]), 1), np.int32)
    h, r, c = G["tk_h"], G["tk_r"], G["tk_c"]
    counts[win_off[h] + r - lo[h]] = c
    hyp = {"hyp_off": G["tk_hyp_off"], "lo": lo, "hi": hi, "spec": G["tk_hyp_spec"]}
    return hyp, {"win_off": win_off, "counts": counts[:int(win_off[-1])]}


def _ref_entries(G, name):
    """The reference's (spectrum, row, hypothesis, count) entries, sorted by (spectrum, row)."""
    h, r, c = G[f"tk_{name}_h"], G[f"tk_{name}_r"], G[f"tk_{name}_c"]
    return np.stack([G["tk_hyp_spec"][h], r, h, c], 1)


def _list_entries(n_cand, row, hyp, cnt, stride):
    e = [(s, row[s * stride + j], hyp[s * stride + j], cnt[s * stride + j])
         for s in range(len(n_cand)) for j in range(n_cand[s])]
    e = np.array(sorted(e), np.int64).reshape(-1, 4)
    return e


def _check_topk(G, name, got):
    """Entries equal to the reference's; the hypothesis may differ only where the reference's unstable
    torch.argsort chose between identical top-K keys (one row reached with the same count from two hypotheses
    of the spectrum, cut at the K-th place): this build keeps the lower hypothesis, the reference whichever its
    sort put first. Both must hold the entry's source row with that count in the appended list."""
    ref = _ref_entries(G, name)
    assert got.shape == ref.shape
    assert np.array_equal(got[:, [0, 1, 3]], ref[:, [0, 1, 3]])
    app = {(int(h), int(r)): int(c) for h, r, c in zip(G["tk_h"], G["tk_r"], G["tk_c"])}
    diff = np.flatnonzero(got[:, 2] != ref[:, 2])
    for i in diff:
        s, r, _, c = got[i]
        src = r if c >= 0 else G["tk_partner"][r]
        cg, cr = app.get((int(got[i, 2]), int(src))), app.get((int(ref[i, 2]), int(src)))
        assert cg is not None and cg == cr, (got[i], ref[i])
        assert got[i, 2] < ref[i, 2]
    return len(diff)


@pytest.mark.parametrize("name,topk,closure", TK_CASES)
def test_select_topk_oracle_equals_reference(G, name, topk, closure):
    """The oracle's top-K + closure over the window counts of the reference's appended list."""
    from vulcan.gpu._abi import make_params
    O = _O()
    hyp, k1 = _tk_hyp(G)
    vp = make_params(_params(topk=topk, pair_closure=closure), {}, {})
    stride = 2 * topk if closure else topk
    part = types.SimpleNamespace(partner=np.ascontiguousarray(G["tk_partner"], np.int64))
    t = O.topk(vp, part, hyp, k1, stride)
    got = _list_entries(t["n_cand"], t["row"], t["hyp"], t["cnt"], stride)
    _check_topk(G, name, got)


def _tk_table(G):
    from vulcan.index import synth
    l1, lab = synth.build_index(0, n_proteins=500)
    sha = np.frombuffer(hashlib.sha256(np.asarray(l1.mass32, np.float32).tobytes()).digest(), np.uint8)
    assert np.array_equal(sha, G["tk_mass32_sha"]), "the synthetic table of the vectors changed"
    return l1, lab


def _psm_inputs(G):
    k2 = {k: G[f"ps_k2_{k}"] for k in ("hyperscore", "summed_b", "summed_y", "ppm", "matched_b", "matched_y",
                                       "longest_b", "longest_y")}
    hyp = {k: G[f"ps_hyp_{k}"] for k in ("spec", "z", "iso", "pm")}
    n_spec = len(G["ps_tic"])
    hyp_off = np.searchsorted(hyp["spec"], np.arange(n_spec + 1)).astype(np.int64)
    return k2, hyp, hyp_off, n_spec


def _check_psm(G, cols):
    n = len(G["psm_spec_idx"])
    for k in PSM_EXACT:
        assert np.array_equal(np.asarray(cols[k]).astype(np.int64), G[f"psm_{k}"].astype(np.int64)), k
        assert len(cols[k]) == n
    for k in PSM_F32:
        assert np.array_equal(np.asarray(cols[k], np.float32).view(np.uint32), G[f"psm_{k}"].view(np.uint32)), k
    for k in ("hyperscore", "delta_next", "delta_best", "rt"):
        assert np.array_equal(np.asarray(cols[k], np.float64), G[f"psm_{k}"]), k
    got, ref = np.asarray(cols["poisson"], np.float64), G["psm_poisson"]
    assert np.array_equal(np.isnan(got), np.isnan(ref)) and np.array_equal(np.isneginf(got), np.isneginf(ref))
    fin = np.isfinite(ref)
    np.testing.assert_allclose(got[fin], ref[fin], rtol=1e-12, atol=0)
    assert np.isneginf(ref).sum() > 0 and np.isnan(ref).sum() > 0


def test_psm_table_oracle_equals_reference(G):
    """The oracle's ranking (or_rank) and feature columns (oracle/psm.py) on the reference's psm_table inputs."""
    from oracle import psm as OP
    from vulcan.gpu._abi import make_params
    O = _O()
    l1, _ = _tk_table(G)
    k2, hyp, hyp_off, n_spec = _psm_inputs(G)
    h, r = G["ps_h"], G["ps_r"]
    c_spec = hyp["spec"][h]
    g0 = np.searchsorted(c_spec, np.arange(n_spec)).astype(np.int64)
    gc = np.diff(np.searchsorted(c_spec, np.arange(n_spec + 1))).astype(np.int32)
    report = 3
    vp = make_params(_params(), {}, {})
    sel, nxt = O.rank(vp, g0, gc, r, k2["hyperscore"], k2["matched_b"], k2["matched_y"], report)
    kc = {"hs": k2["hyperscore"], "sb": k2["summed_b"], "sy": k2["summed_y"], "ppm": k2["ppm"],
          "mb": k2["matched_b"], "my": k2["matched_y"], "lb": k2["longest_b"], "ly": k2["longest_y"]}
    d = OP.features(sel, nxt, None, r, h, G["ps_k1c"], kc, hyp_off, hyp["z"], hyp["iso"], hyp["pm"],
                    G["ps_scored"], G["ps_matched"], G["ps_tic"], l1.mass32, l1.length, report)
    keep = d["rank"] > 0
    cols = {k: v[keep] for k, v in d.items()}
    rows = cols["pf_row"]
    cols.update(scan=cols["spec_idx"] + 1000, rt=np.linspace(0.0, 90.0, n_spec)[cols["spec_idx"]],
                pf_id=l1.pf_id[rows], missed_cleavages=l1.missed[rows], n_var=l1.n_var[rows])
    _check_psm(G, cols)


# ------------------------------------------------------------------------------------------------ GPU
@pytest.mark.gpu
@pytest.mark.parametrize("name,kw", K0_CASES)
def test_k0_device_equals_reference(G, name, kw):
    import torch
    from vulcan.gpu import engine as E
    from vulcan.gpu._abi import make_params
    p = _params(**kw)
    raw = E.upload_spectra(_spectra(G), torch.device("cuda"))
    o = E._k0_dev(make_params(p, {}, {}), raw, p, torch.device("cuda"))
    torch.cuda.synchronize()
    n = raw["n"]
    off = o["off"].cpu().numpy()
    m = int(off[-1])
    _check_k0(G, name, off, o["mass"][:m].cpu().numpy(), o["int"][:m].cpu().numpy(), o["tic"][:n].cpu().numpy(),
              o["valid_d"][:n].cpu().numpy().astype(bool))


@pytest.mark.gpu
@pytest.mark.parametrize("impl", ["join", "tile"])
@pytest.mark.parametrize("name,topk,closure", TK_CASES)
def test_select_topk_device_equals_reference(G, name, topk, closure, impl):
    """The fused K1 + top-K + closure of the device step on the table and spectra of the vectors equals the
    reference's select_topk over the appended list of the same windows."""
    import torch
    from vulcan.gpu import engine as E
    from vulcan.index import synth
    l1, lab = _tk_table(G)
    sp, _, _ = synth.spectra(l1, lab, 0, n=60, seed=7)
    w = synth.WORKLOADS[0]
    p = w.search_params()
    p.precursor_ppm, p.isotope_errors, p.k1_threshold = (-1000.0, 1000.0), (0, 2), 1
    p.topk, p.pair_closure = topk, closure
    gi = E.GpuIndex(l1, p, STATIC, dict(w.var_mods))
    ds = E.DeviceSearch(gi, p, len(sp), k1_impl=impl)
    ds.run(E.upload_spectra(sp, gi.dev))
    torch.cuda.synchronize()
    assert int(ds.status.item()) == 0
    n = len(sp)
    nh = int(ds.hyp["hyp_off"][n].item())
    assert np.array_equal(ds.hyp["spec"][:nh].cpu().numpy(), G["tk_hyp_spec"])
    got = _list_entries(ds.n_cand[:n].cpu().numpy(), ds.c_row.cpu().numpy(), ds.c_hyp.cpu().numpy(),
                        ds.c_cnt.cpu().numpy(), ds.stride)
    _check_topk(G, name, got)


@pytest.mark.gpu
def test_psm_table_device_equals_reference(G):
    """engine.psm_table (vh_rank + vh_psm_features) on the reference's psm_table inputs."""
    import torch
    from vulcan.gpu import engine as E
    from vulcan.gpu.host import Spectra
    from vulcan.index import synth
    l1, _ = _tk_table(G)
    k2, hyp, hyp_off, n_spec = _psm_inputs(G)
    p = _params()
    gi = E.GpuIndex(l1, p, STATIC, dict(synth.WORKLOADS[0].var_mods))
    d = gi.dev
    sp = Spectra("golden", np.arange(n_spec) + 1000, np.arange(n_spec), np.linspace(0.0, 90.0, n_spec),
                 np.full(n_spec, 600.0), np.zeros(n_spec, np.int8), np.zeros(0, np.float32),
                 np.zeros(0, np.float32), np.zeros(n_spec + 1, np.int64))
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=d)   # noqa: E731
    tbl = E.psm_table(sp, gi, l1, dict(hyp, hyp_off=hyp_off), t(G["ps_h"]), t(G["ps_r"]), t(G["ps_k1c"]),
                      {k: t(v) for k, v in k2.items()}, t(G["ps_scored"]), t(G["ps_matched"]),
                      {"tic": t(G["ps_tic"])}, p, report=3)
    _check_psm(G, {k: tbl.column(k).to_numpy() for k in tbl.column_names if k != "run"})
