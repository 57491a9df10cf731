"""CPU checks of the learned-prefilter oracle (oracle/prefilter.py, or_topk_feat in oracle/vh_oracle.cpp):
finite differences for the analytic gradient, closed forms for the intensity quantisation, brute force for the
matched-intensity sums, invariants for the selection. No GPU."""
import numpy as np
import pytest

from conftest import STATIC


def _O():
    from oracle import oracle as O
    return O


def _P():
    from oracle import prefilter as P
    return P


# ------------------------------------------------------------------------------------------------ MLP maths
def _mlp_case(seed=0, n=48):
    P = _P()
    rng = np.random.default_rng(seed)
    th = (rng.standard_normal(P.N_PARAMS) * 0.08).astype(np.float32)
    mu = rng.standard_normal(8).astype(np.float32)
    isd = (0.5 + rng.random(8)).astype(np.float32)
    X = (rng.standard_normal((n, 8)) * 3).astype(np.float32)
    y = rng.standard_normal(n).astype(np.float32)
    return th, mu, isd, X, y


def test_exact_gradient_matches_central_differences():
    P = _P()
    th, mu, isd, X, y = _mlp_case()
    g = P.grad(th, mu, isd, X, y, exact=True)
    rng = np.random.default_rng(1)
    th64 = th.astype(np.float64)
    # every parameter block is probed
    idx = np.concatenate([rng.choice(np.arange(o, o + int(np.prod(sh))), 6, replace=int(np.prod(sh)) < 6)
                          for o, sh in P.OFF.values()])
    for i in idx:
        e = 1e-6
        tp, tm = th64.copy(), th64.copy()
        tp[i] += e
        tm[i] -= e
        fd = (P.loss(tp, mu, isd, X, y) - P.loss(tm, mu, isd, X, y)) / (2 * e)
        assert abs(fd - g[i]) <= 1e-5 * (abs(fd) + abs(g[i])) + 1e-9, (i, fd, g[i])


def test_gradient_of_a_linear_network_is_the_least_squares_gradient():
    """With ReLU never active the output is b3 alone: dL/db3 = 2 mean(b3 - y), everything else 0 except w3 = 0."""
    P = _P()
    th = np.zeros(P.N_PARAMS, np.float32)
    th[P.OFF["b1"][0]:P.OFF["b1"][0] + 256] = -1.0   # layer 1 dead
    th[P.OFF["b3"][0]] = 0.75
    X = np.random.default_rng(2).standard_normal((10, 8)).astype(np.float32)
    y = np.arange(10, dtype=np.float32)
    g = P.grad(th, np.zeros(8, np.float32), np.ones(8, np.float32), X, y, exact=True)
    assert g[P.OFF["b3"][0]] == pytest.approx(2 * np.mean(0.75 - y))
    g2 = g.copy()
    g2[P.OFF["b3"][0]] = 0
    assert not g2.any()


def test_emulated_rounding_stays_within_bf16_error():
    P = _P()
    th, mu, isd, X, y = _mlp_case(3, 256)
    s_ex = P.forward(th, mu, isd, X, exact=True)["s"]
    s_em = P.forward(th, mu, isd, X, exact=False)["s"]
    assert np.max(np.abs(s_em - s_ex)) <= 0.05 * (np.max(np.abs(s_ex)) + 1)
    g_ex = P.grad(th, mu, isd, X, y, exact=True)
    g_em = P.grad(th, mu, isd, X, y, exact=False)
    assert np.linalg.norm(g_em - g_ex) <= 0.05 * np.linalg.norm(g_ex)


def test_adam_restatement_closed_form_first_step():
    """Step 1 from zero moments: m = (1-b1) g, v = (1-b2) g^2, bias-corrected -> theta - lr * g / (|g| + eps)."""
    P = _P()
    g = np.array([0.5, -2.0, 1e-3, 0.0], np.float32)
    th = np.ones(4, np.float32)
    z = np.zeros(4, np.float32)
    b1, b2, eps, lr = 0.9, 0.999, 1e-8, 0.01
    c1, c2 = np.float32(1 / (1 - b1)), np.float32(1 / (1 - b2))
    t, m, v = P.adam(th, g, z, z, lr, b1, b2, eps, c1, c2)
    np.testing.assert_allclose(t, 1 - lr * g / (np.abs(g) + eps), rtol=1e-5)
    np.testing.assert_allclose(m, 0.1 * g, rtol=1e-6)


# ------------------------------------------------------------------------------------------------ quantisation
def test_iq_log_is_the_piecewise_linear_log2():
    O = _O()
    rng = np.random.default_rng(5)
    vals = [0, 1, 2, 3, 254, 255, 256, 257, 511, 512, 2 ** 40 - 1] + rng.integers(0, 2 ** 40, 200).tolist()
    for s in vals:
        x = s + 1
        e = x.bit_length() - 1
        frac = (x - (1 << e)) * 256 // (1 << e)   # 8 fraction bits below the leading one, truncated
        assert O.iq_log(s) == (e << 8) | frac, s
    assert O.iq_log(0) == 0 and O.iq_log(1) == 256 and O.iq_log(2) == 384


def test_iq_quant_maps_the_largest_peak_into_16_bits():
    O = _O()
    assert O.iq_quant(1.0, 1.0) == 32768          # 1.0 = 0.5 * 2^1 -> scaled by 2^15
    assert O.iq_quant(0.5, 1.0) == 16384
    assert O.iq_quant(0.99, 0.99) == int(np.float32(0.99) * np.float32(65536))
    for mx in (1e-30, 3.7, 1234.5, 6.0e7, 3e38):
        q = O.iq_quant(mx, mx)
        assert 32768 <= q <= 65535
        assert O.iq_quant(mx * 0.25, mx) in (q // 4, q // 4 - 1, q // 4 + 1)
    assert O.iq_quant(float("nan"), 1.0) == 0
    assert O.iq_quant(-1.0, 1.0) == 0
    assert O.iq_quant(1.0, float("inf")) == 0
    assert O.iq_quant(1.0, 0.0) == 0
    assert O.iq_quant(1e-45, 1.0) == 0               # denormal intensity


# ------------------------------------------------------------------------------------------------ K1 features
@pytest.fixture(scope="module")
def feat_env(small_plain):
    from vulcan.gpu._abi import make_params, max_hyp_per_spectrum
    O = _O()
    sp = small_plain["sp"]
    p = type(small_plain["p"])(**small_plain["p"].to_dict())
    p.topk = 40
    p.pair_closure = False
    vp = make_params(p, STATIC, {})
    hix = O.HostIndex(small_plain["l1"])
    pk = O.k0(vp, sp)
    hyp = O.hypotheses(vp, sp, pk["valid"], small_plain["l1"].mass32, max_hyp_per_spectrum(p))
    k1 = O.k1_counts(vp, hix, pk, hyp)
    return {"vp": vp, "hix": hix, "pk": pk, "hyp": hyp, "k1": k1, "p": p, "l1": small_plain["l1"]}


def test_topk_feat_lists_equal_the_plain_topk_without_closure(feat_env):
    O = _O()
    e = feat_env
    tf = O.topk_feat(e["vp"], e["hix"], e["pk"], e["hyp"], 40)
    tk = O.topk(e["vp"], e["hix"], e["hyp"], e["k1"], 40)
    assert np.array_equal(tf["n_cand"], tk["n_cand"])
    n = len(tf["n_cand"])
    m = (np.arange(40)[None, :] < tf["n_cand"][:, None]).ravel()
    assert m.sum() > 1000
    for k in ("row", "hyp", "cnt"):
        assert np.array_equal(tf[k][:n * 40][m], tk[k][:n * 40][m]), k
    best = np.array([tf["cnt"][s * 40:s * 40 + c].max() if c else 0 for s, c in enumerate(tf["n_cand"])])
    assert np.array_equal(tf["maxcnt"], best)


def _ladder(vp, l1, r, ion):
    rm0 = np.array([np.float32(vp.mono[i]) + np.float32(vp.stat[i]) for i in range(26)], np.float32)
    rm1 = np.array([np.float32(vp.mono[i]) + np.float32(vp.var[i]) for i in range(26)], np.float32)
    seq = l1.residues[l1.seq_off[r]:l1.seq_off[r] + l1.length[r]]
    mk = int(l1.mask[r])
    nt = np.float32(vp.nterm) if mk >> 63 else np.float32(0)
    b, y = nt, np.float32(l1.mass32[r]) - nt
    L = int(l1.length[r])
    out = []
    for k in range(L - 1):
        rm = (rm1 if (mk >> k) & 1 else rm0)[seq[k] - 65]
        b = np.float32(b + rm)
        y = np.float32(y - rm)
        if ion == 0 and k + 1 > vp.min_ion_index and vp.fmin <= b <= vp.fmax:
            out.append(b)
        if ion == 1 and L - 1 - k > vp.min_ion_index and vp.fmin <= y <= vp.fmax:
            out.append(y)
    return np.array(out, np.float32)


def test_topk_feat_intensity_sums_by_brute_force(feat_env):
    """Each entry's b-ion count and per-ion c_iq = iq_log(sum over matched (fragment, peak, z) of the peak's
    quantised intensity), recomputed by brute force over all peaks (Da tolerance)."""
    O = _O()
    e = feat_env
    vp, pk, hyp, l1 = e["vp"], e["pk"], e["hyp"], e["l1"]
    assert vp.tol_da == 1
    tf = O.topk_feat(vp, e["hix"], pk, hyp, 40)
    checked = 0
    skipped = 0
    for s in range(0, len(tf["n_cand"]), 13):
        p0, p1 = pk["off"][s], pk["off"][s + 1]
        pm, pi = pk["mass"][p0:p1], pk["int"][p0:p1]
        if len(pm) == 0:
            continue
        mx = float(np.max(np.where(np.isnan(pi), 0, np.maximum(pi, 0))))
        q = np.array([O.iq_quant(float(v), mx) for v in pi], np.uint64)
        for j in range(tf["n_cand"][s]):
            o = s * 40 + j
            r, h = int(tf["row"][o]), int(tf["hyp"][o])
            cnt, tot = [0, 0], [0, 0]
            for ion in (0, 1):
                fr = _ladder(vp, l1, r, ion)
                for z in range(1, int(hyp["maxfz"][h])):
                    m = (pm * np.float32(z)).astype(np.float32)
                    lo = (m + np.float32(vp.tol_lo)).astype(np.float32)
                    hi = (m + np.float32(vp.tol_hi)).astype(np.float32)
                    hit = (fr[:, None] >= lo[None, :]) & (fr[:, None] <= hi[None, :])
                    cnt[ion] += int(hit.sum())
                    tot[ion] += int((hit * q[None, :]).sum())
            # a row in two hypotheses' windows keeps the lower hypothesis but the larger count (and its intensity)
            if sum(cnt) != tf["cnt"][o]:
                skipped += 1
                continue
            assert tf["nb"][o] == cnt[0], (s, r)
            assert tf["iq"][o] == (O.iq_log(tot[0]) << 16) | O.iq_log(tot[1]), (s, r)
            checked += 1
    assert checked > 200 and skipped <= checked // 20, (checked, skipped)


def test_select_oracle_invariants():
    P = _P()
    rng = np.random.default_rng(7)
    n_spec, stride, keep = 30, 50, 6
    partner = rng.permutation(1000)
    partner[partner] = np.arange(1000)   # an involution
    n_cand = rng.integers(0, stride + 1, n_spec).astype(np.int32)
    row = np.zeros(n_spec * stride, np.int32)
    for s in range(n_spec):
        row[s * stride:(s + 1) * stride] = np.sort(rng.choice(1000, stride, replace=False))
    hyp = rng.integers(0, 5, n_spec * stride).astype(np.int32)
    cnt = rng.integers(1, 20, n_spec * stride).astype(np.int32)
    score = rng.standard_normal(n_spec * stride).astype(np.float32)
    score[::17] = np.nan
    out = P.select(n_cand, row, hyp, cnt, score, stride, keep, True, partner, 2 * keep)
    for s in range(n_spec):
        n = n_cand[s]
        got = out["row"][s * 2 * keep:s * 2 * keep + out["n"][s]]
        assert np.all(np.diff(got) > 0)
        sc = np.where(np.isnan(score[s * stride:s * stride + n]), -np.inf, score[s * stride:s * stride + n])
        top = row[s * stride:s * stride + n][np.argsort(-sc, kind="stable")[:keep]]
        assert set(top) <= set(got)
        assert set(partner[top]) <= set(got)
        assert len(got) == len(set(top) | set(partner[top]))


def test_select_oracle_adaptive_keep():
    """keep_min / margin: per spectrum exactly max(min(keep_min, n), #{score >= best - margin} among the first keep)
    slots are kept, and they are the best-scoring ones (no closure)."""
    P = _P()
    rng = np.random.default_rng(8)
    n_spec, stride = 40, 60
    n_cand = rng.integers(0, stride + 1, n_spec).astype(np.int32)
    row = np.zeros(n_spec * stride, np.int32)
    for s in range(n_spec):
        row[s * stride:(s + 1) * stride] = np.sort(rng.choice(5000, stride, replace=False))
    hyp = np.zeros(n_spec * stride, np.int32)
    cnt = np.ones(n_spec * stride, np.int32)
    score = (rng.standard_normal(n_spec * stride) * 2).astype(np.float32)
    score[::23] = np.nan
    for kmin, keep, margin in ((3, 30, 1.0), (1, 5, 10.0), (10, 10, 0.0), (2, 60, 0.0)):
        out = P.select(n_cand, row, hyp, cnt, score, stride, keep, False, None, keep, keep_min=kmin, margin=margin)
        for s in range(n_spec):
            n = int(n_cand[s])
            sc = score[s * stride:s * stride + n].astype(np.float64)
            scn = np.where(np.isnan(sc), -np.inf, sc)
            o = np.argsort(-scn, kind="stable")
            within = int(np.sum(scn[o[:min(keep, n)]] >= (np.float32(scn[o[0]] - np.float32(margin)) if n else 0)))
            want_n = max(min(kmin, n), within)
            assert out["n"][s] == want_n, (kmin, keep, margin, s)
            got = set(out["row"][s * keep:s * keep + want_n].tolist())
            assert got == set(row[s * stride:s * stride + n][o[:want_n]].tolist())


def test_features_restatement_known_values():
    P = _P()
    mass32 = np.array([1000.0, 2000.0], np.float32)
    length = np.array([10, 12], np.uint8)
    mask = np.array([0b101, (1 << 63) | 1], np.uint64)
    f = P.features(mass32, length, n_cand=np.array([2, 0], np.int32), row=np.array([0, 1, 0, 0], np.int32),
                   hyp=np.array([0, 1, 0, 0], np.int32), cnt=np.array([6, 3, 0, 0], np.int32),
                   nb=np.array([4, 0, 0, 0], np.int16),
                   iq=np.array([(512 << 16) | 256, 1000, 0, 0], np.uint32), maxcnt=np.array([6, 0], np.int32),
                   h_z=np.array([2, 3], np.int8), h_M=np.array([1000.01, 2000.0], np.float32), stride=2)
    assert f.shape == (4, 8)
    assert f[0].tolist()[:5] == [6.0, 1.0, 4.0, 2.0, 1.0]
    assert f[0][5] == pytest.approx(10.0, rel=1e-3)
    assert f[0][6:].tolist() == [10.0, 2.0]
    assert f[1].tolist() == [3.0, 0.5, 0.0, 0.0, 1000 / 256, 0.0, 12.0, 3.0]
    assert not f[2:].any()
