"""HIP path vs the CPU oracle on the same seeded inputs (bit-exact for counts, candidate lists, matched ions and
f32 sums; hyperscore within 1e-12 relative: the device and host f64 log may differ in the last ulp)."""
import numpy as np
import pytest

from conftest import STATIC

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

HS_RTOL = 1e-12


def _O():
    from oracle import oracle as O
    return O


class Env:
    def __init__(self, d, **over):
        from vulcan.gpu import engine as E
        from vulcan.gpu._abi import make_params
        self.d = d
        self.p = d["p"]
        for k, v in over.items():
            setattr(self.p, k, v)
        self.l1 = d["l1"]
        self.gi = E.GpuIndex(self.l1, self.p, d["static"], d["var"])
        self.vp = make_params(self.p, d["static"], d["var"])
        O = _O()
        self.hix = O.HostIndex(self.l1)
        self.pk = O.k0(self.vp, d["sp"])
        from vulcan.gpu._abi import max_hyp_per_spectrum
        self.hyp = O.hypotheses(self.vp, d["sp"], self.pk["valid"], self.l1.mass32, max_hyp_per_spectrum(self.p))
        self.k1 = O.k1_counts(self.vp, self.hix, self.pk, self.hyp)


@pytest.fixture(scope="module")
def env0(small_plain):
    print("[gpu] building small_plain env", flush=True)
    return Env(small_plain)


@pytest.fixture(scope="module")
def env1(proteome_small):
    print("[gpu] building proteome-small env", flush=True)
    return Env(proteome_small)


def _np(t):
    return t.detach().cpu().numpy()


# ------------------------------------------------------------------------------------------------ index
def _postings_numpy(l1, vp, gi):
    """(bucket, f32 bits, row in bucket, ion 0 = b / 1 = y) of every posting, by f32 accumulation in emission
    order (np.add/subtract.accumulate are sequential)."""
    rm0 = np.array([np.float32(vp.mono[i]) + np.float32(vp.stat[i]) for i in range(26)], np.float32)
    rm1 = np.array([np.float32(vp.mono[i]) + np.float32(vp.var[i]) for i in range(26)], np.float32)
    out = []
    mii, fmin, fmax = vp.min_ion_index, np.float32(vp.fmin), np.float32(vp.fmax)
    for i in range(len(l1)):
        L = int(l1.length[i])
        s = l1.residues[l1.seq_off[i]:l1.seq_off[i] + L].astype(np.int64) - 65
        bits = (int(l1.mask[i]) >> np.arange(L)) & 1
        rm = np.where(bits[:L - 1] == 1, rm1[s[:L - 1]], rm0[s[:L - 1]]).astype(np.float32)
        b = np.add.accumulate(rm, dtype=np.float32)
        y = np.subtract.accumulate(np.concatenate([[l1.mass32[i]], rm]).astype(np.float32), dtype=np.float32)[1:]
        for j in range(L - 1):
            for f, ok, ion in ((b[j], j + 1 > mii, 0), (y[j], L - 1 - j > mii, 1)):
                if ok and fmin <= f <= fmax:
                    fb = int(np.array([f], np.float32).view(np.uint32)[0])
                    out.append((i >> gi.bucket_log2, fb, i & ((1 << gi.bucket_log2) - 1), ion))
    return np.array(out, np.int64).reshape(-1, 4)


def test_index_build_matches_numpy(env0):
    """Device emission + LSD radix sort + directory == numpy emission + stable sort by (bin, bucket, f32 bits), on the
    first 3,000 rows (a sub-table keeps the Python loop short)."""
    from vulcan.gpu import engine as E
    from vulcan.index.build import L1Index
    l1 = env0.l1
    n = 3000
    sub = L1Index(arrays={"pf_id": l1.pf_id[:n], "mass32": l1.mass32[:n], "seq_off": l1.seq_off[:n],
                          "length": l1.length[:n], "mask": l1.mask[:n], "n_var": l1.n_var[:n],
                          "missed": l1.missed[:n], "partner": np.arange(n), "residues": l1.residues})
    p = env0.p
    p2 = type(p)(**{**p.to_dict(), "chunk_size": 512})
    gi = E.GpuIndex(sub, p2, STATIC, {})
    post = _postings_numpy(sub, gi.vp, gi)
    assert gi.n_postings == len(post)
    S = gi.bin_shift
    bins = (post[:, 1] >> S) - gi.bin_base
    o = np.lexsort((post[:, 1], post[:, 0], bins))        # stable: equal (bin, bucket, key) keep emission order
    ps, pb = post[o], bins[o]
    want_post = ((ps[:, 1] & ((1 << S) - 1)) << 17) | (ps[:, 3] << 16) | ps[:, 2]
    assert np.array_equal(_np(gi.postings).view(np.uint32)[:len(ps)].astype(np.int64), want_post)
    d = pb * gi.n_buckets + ps[:, 0]
    want = np.searchsorted(d, np.arange(gi.n_bins * gi.n_buckets + 1), side="left")
    assert np.array_equal(_np(gi.offsets).view(np.uint32).astype(np.int64), want)


@pytest.mark.parametrize("q_rb", [(12, 10), (7, 4)])
def test_block_layout_matches_numpy(env0, q_rb):
    """vh_index_block_count + vh_index_build_blocks == numpy: block cuts where floor(post_off / 2^q) or
    floor(row / 2^rb) changes, postings of a block stably sorted by f32 bits (emission order inside equal keys),
    row field ion << 15 | row - block start, per-block posting offsets and the size statistics."""
    from vulcan.gpu import engine as E
    from vulcan.index.build import L1Index
    l1 = env0.l1
    n = 3000
    sub = L1Index(arrays={"pf_id": l1.pf_id[:n], "mass32": l1.mass32[:n], "seq_off": l1.seq_off[:n],
                          "length": l1.length[:n], "mask": l1.mask[:n], "n_var": l1.n_var[:n],
                          "missed": l1.missed[:n], "partner": np.arange(n), "residues": l1.residues})
    q, rb = q_rb
    p2 = type(env0.p)(**{**env0.p.to_dict(), "chunk_size": 512, "block_postings_log2": q, "block_rows_log2": rb})
    gi = E.GpuIndex(sub, p2, STATIC, {})
    bv = gi.blocks()
    b = gi.block_arrays()
    post = _postings_numpy(sub, gi.vp, gi)
    rows = (post[:, 0] << gi.bucket_log2) + post[:, 2]
    po = _np(gi.post_off)
    r = np.arange(n)
    flag = np.ones(n, bool)
    flag[1:] = ((po[1:n] >> q) != (po[:n - 1] >> q)) | ((r[1:] >> rb) != (r[:-1] >> rb))
    row0 = np.concatenate([np.flatnonzero(flag), [n]])
    nb = len(row0) - 1
    assert bv.n_blocks == nb
    assert np.array_equal(_np(b["row0"]), row0)
    assert np.array_equal(_np(b["off"]).view(np.uint32).astype(np.int64), po[row0])
    blk = np.searchsorted(row0, rows, side="right") - 1
    o = np.lexsort((post[:, 1], blk))                       # stable
    assert np.array_equal(_np(b["key"]).view(np.uint32)[:len(o)].astype(np.int64), post[o, 1])
    want_row = (post[o, 3] << 15) | (rows[o] - row0[blk[o]])
    assert np.array_equal(_np(b["row"]).view(np.uint16)[:len(o)].astype(np.int64), want_row)
    sizes = np.diff(po[row0])
    assert bv.max_rows == np.diff(row0).max() and bv.max_postings == sizes.max()
    assert np.diff(row0).max() <= 2 ** rb and sizes.max() < 2 ** q + 64


# ------------------------------------------------------------------------------------------------ K0 / hyps
def test_k0_parity(env0):
    from vulcan.gpu import engine as E
    st = E.Stats()
    pk = E.k0_preprocess(env0.d["sp"], env0.p, env0.gi.dev, st, env0.gi)
    o = env0.pk
    assert np.array_equal(_np(pk["off"]), o["off"])
    m = int(o["off"][-1])
    assert np.array_equal(_np(pk["mass"])[:m], o["mass"])
    assert np.array_equal(_np(pk["int"])[:m], o["int"])
    assert np.array_equal(pk["valid"], o["valid"].astype(bool))
    assert np.allclose(_np(pk["tic"]), o["tic"], rtol=1e-6)


def test_k0_edge_cases_parity(env0):
    """Empty spectra, NaN m/z and intensities, ties, bounds, and spectra with > 512 in-range peaks (global path)."""
    from vulcan.gpu import engine as E
    from vulcan.gpu.host import SearchParams, Spectra
    rng = np.random.default_rng(11)
    peaks = []
    for n in (0, 3, 150, 151, 513, 900, 1200):
        mz = rng.uniform(140, 2010, n).astype(np.float32)
        it = rng.integers(1, 4, n).astype(np.float32)
        if n > 10:
            mz[:5] = mz[5:10]
            mz[3] = np.nan
            it[4] = np.nan
            mz[6], mz[7] = 150.0, 2000.0
        peaks.append((mz, it))
    off = np.zeros(len(peaks) + 1, np.int64)
    np.cumsum([len(m) for m, _ in peaks], out=off[1:])
    sp = Spectra("e", np.arange(len(peaks), dtype=np.uint32), np.arange(len(peaks), dtype=np.uint32),
                 np.zeros(len(peaks)), np.full(len(peaks), 600.0), np.full(len(peaks), 2, np.int8),
                 np.concatenate([m for m, _ in peaks]), np.concatenate([i for _, i in peaks]), off)
    p = SearchParams(max_peaks=150, min_peaks=10)
    from vulcan.gpu._abi import make_params
    vp = make_params(p, STATIC, {})
    o = _O().k0(vp, sp)
    pk = E.k0_preprocess(sp, p, env0.gi.dev, E.Stats(), None)
    m = int(o["off"][-1])
    assert np.array_equal(_np(pk["off"]), o["off"])
    assert np.array_equal(_np(pk["mass"])[:m], o["mass"])
    assert np.array_equal(_np(pk["int"])[:m], o["int"], equal_nan=True)
    assert np.array_equal(pk["valid"], o["valid"].astype(bool))


def test_k0_sorted_fast_path_parity(env0):
    """m/z-sorted spectra (the K0 fast path: bitwise top-K search + ballot compaction) with more in-range peaks
    than max_peaks, many tied intensities (NaN included), exact-threshold ties, and unsorted twins of the same
    spectra (the general path), against the oracle."""
    from vulcan.gpu import engine as E
    from vulcan.gpu._abi import make_params
    from vulcan.gpu.host import SearchParams, Spectra
    rng = np.random.default_rng(12)
    peaks = []
    for n in (0, 1, 20, 149, 150, 151, 200, 300, 500, 512):
        for variant in range(4):
            mz = np.sort(rng.uniform(140, 2010, n)).astype(np.float32)
            mz = np.unique(mz) if variant < 3 else mz          # strictly ascending for the fast path
            k = len(mz)
            if variant == 0:
                it = rng.lognormal(0, 1, k).astype(np.float32)
            elif variant == 1:
                it = rng.integers(1, 6, k).astype(np.float32)   # heavy ties at the selection threshold
            elif variant == 2:
                it = rng.integers(1, 4, k).astype(np.float32)
                if k > 3:
                    it[rng.choice(k, 3, replace=False)] = np.nan
            else:
                it = rng.integers(1, 4, k).astype(np.float32)
                mz = mz[rng.permutation(k)]                     # unsorted: the general path
            peaks.append((mz, it))
    off = np.zeros(len(peaks) + 1, np.int64)
    np.cumsum([len(m) for m, _ in peaks], out=off[1:])
    n_sp = len(peaks)
    sp = Spectra("f", np.arange(n_sp, dtype=np.uint32), np.arange(n_sp, dtype=np.uint32), np.zeros(n_sp),
                 np.full(n_sp, 600.0), np.full(n_sp, 2, np.int8), np.concatenate([m for m, _ in peaks]),
                 np.concatenate([i for _, i in peaks]), off)
    for max_peaks in (150, 64):
        p = SearchParams(max_peaks=max_peaks, min_peaks=10)
        vp = make_params(p, STATIC, {})
        o = _O().k0(vp, sp)
        pk = E.k0_preprocess(sp, p, env0.gi.dev, E.Stats(), None)
        m = int(o["off"][-1])
        assert np.array_equal(_np(pk["off"]), o["off"])
        assert np.array_equal(_np(pk["mass"])[:m], o["mass"])
        assert np.array_equal(_np(pk["int"])[:m], o["int"], equal_nan=True)
        assert np.array_equal(pk["valid"], o["valid"].astype(bool))
        assert np.allclose(_np(pk["tic"])[~np.isnan(o["tic"])], o["tic"][~np.isnan(o["tic"])], rtol=1e-6)


def test_hypotheses_parity(env1):
    from vulcan.gpu import engine as E
    st = E.Stats()
    pk = E.k0_preprocess(env1.d["sp"], env1.p, env1.gi.dev, st, env1.gi)
    hd = E.device_hypotheses(env1.gi, pk["raw"], pk, env1.p)
    h = E.hyp_to_host(hd)
    for k in ("hyp_off", "spec", "z", "iso", "pm", "M", "lo", "hi", "maxfz"):
        assert np.array_equal(h[k], env1.hyp[k]), k


# ------------------------------------------------------------------------------------------------ K1
def _device_peaks(env):
    from vulcan.gpu import engine as E
    return E.k0_preprocess(env.d["sp"], env.p, env.gi.dev, E.Stats(), env.gi)


@pytest.mark.parametrize("which", ["env0", "env1"])
def test_k1_append_equals_oracle(which, request):
    from vulcan.gpu import engine as E
    env = request.getfixturevalue(which)
    pk = _device_peaks(env)
    o_h, o_r, o_c, scored, matched, d_spec, d_maxfz = E.run_k1(env.gi, env.hyp, pk, env.p, E.Stats())
    O = _O()
    hh, rr, cc = O.k1_append(env.hyp, env.k1, env.p.k1_threshold)
    a = sorted(zip(_np(o_h).tolist(), _np(o_r).tolist(), _np(o_c).tolist()))
    b = sorted(zip(hh.tolist(), rr.tolist(), cc.tolist()))
    assert len(a) > 100 and a == b
    assert np.array_equal(_np(scored).astype(np.uint32), env.k1["scored"])
    assert np.array_equal(_np(matched).astype(np.uint64), env.k1["matched"])


def test_k1_brute_device_equals_oracle(env1):
    from vulcan.gpu import engine as E
    pk = _device_peaks(env1)
    hh, rr, cc = _O().k1_append(env1.hyp, env1.k1, 0)
    sel = np.random.default_rng(4).choice(len(hh), min(len(hh), 20000), replace=False)
    dev = env1.gi.dev
    bc = E.k1_brute(env1.gi, env1.hyp, pk, torch.from_numpy(hh[sel]).to(dev), torch.from_numpy(rr[sel]).to(dev),
                    env1.p)
    assert np.array_equal(_np(bc), cc[sel])


def _topk_device(env, **over):
    from vulcan.gpu import engine as E
    p = type(env.p)(**{**env.p.to_dict(), **over})
    gi = env.gi
    if (p.block_postings_log2, p.block_rows_log2) != (env.p.block_postings_log2, env.p.block_rows_log2):
        gi = E.GpuIndex(env.l1, p, env.d["static"], env.d["var"])     # another block layout
    ds = E.DeviceSearch(gi, p, len(env.d["sp"]))
    raw = E.upload_spectra(env.d["sp"], env.gi.dev)
    ds.run(raw)
    torch.cuda.synchronize()
    assert int(ds.status.item()) == 0
    return ds, p


@pytest.mark.parametrize("over", [{"k1_impl": "tile"}, {"k1_impl": "tile", "topk": 3},
                                  {"k1_impl": "tile", "topk": 3, "tile": 64}, {"k1_impl": "tile", "topk": 200, "tile": 256},
                                  {"k1_impl": "join"}, {"k1_impl": "join", "topk": 3}, {"k1_impl": "join", "topk": 200},
                                  {"k1_impl": "join", "block_postings_log2": 8, "block_rows_log2": 5}])
@pytest.mark.parametrize("which", ["env0", "env1"])
def test_k1_topk_fused_equals_oracle(which, over, request):
    from vulcan.gpu._abi import make_params
    env = request.getfixturevalue(which)
    ds, p = _topk_device(env, **over)
    vp = make_params(p, env.d["static"], env.d["var"])
    tk = _O().topk(vp, env.hix, env.hyp, env.k1, ds.stride)
    n = len(env.d["sp"])
    nc = _np(ds.n_cand)[:n]
    assert np.array_equal(nc, tk["n_cand"])
    m = (np.arange(ds.stride)[None, :] < nc[:, None]).ravel()
    for dv, ov in ((ds.c_row, tk["row"]), (ds.c_hyp, tk["hyp"]), (ds.c_cnt, tk["cnt"])):
        assert np.array_equal(_np(dv)[:n * ds.stride][m], ov[m])
    nh = len(env.hyp["spec"])
    assert np.array_equal(_np(ds.scored)[:nh].astype(np.uint32), env.k1["scored"])
    assert np.array_equal(_np(ds.matched)[:nh].astype(np.uint64), env.k1["matched"])


@pytest.mark.parametrize("which", ["env0", "env1"])
def test_join_pairs_and_counts_equal_oracle(which, request):
    """vh_k1j_pairs: slots = window rows in hypothesis order (the oracle's window layout), every block's pair
    list = the hypotheses whose window overlaps it; vh_k1j_count: every slot == the oracle's K1 count."""
    from vulcan.gpu import engine as E
    env = request.getfixturevalue(which)
    ds, p = _topk_device(env, k1_impl="join")
    j = ds.join
    nh = len(env.hyp["spec"])
    hs = _np(j["h_slot"])
    assert np.array_equal(hs[:nh + 1], env.k1["win_off"])
    assert np.all(hs[nh:] == env.k1["win_off"][-1])
    b = env.gi.block_arrays()
    row0 = _np(b["row0"]).astype(np.int64)
    nb = len(row0) - 1
    bp = _np(j["blk_pairs"]).view(np.uint32).astype(np.int64)
    rec = _np(j["pairs"])
    lo, hi = env.hyp["lo"], env.hyp["hi"]
    # records: (first peak, slot, lo | hi << 32, peak count | maxfz << 32); the slot names the hypothesis
    pairs = np.searchsorted(hs[:nh + 1], rec[:, 1], side="right") - 1
    w32 = rec[:, 2:].copy().view(np.int32)
    po = _np(ds.peaks["off"])
    spec = env.hyp["spec"][pairs[:bp[nb]]]
    assert np.array_equal(rec[:bp[nb], 0], po[spec])
    assert np.array_equal(w32[:bp[nb], 0], lo[pairs[:bp[nb]]]) and np.array_equal(w32[:bp[nb], 1], hi[pairs[:bp[nb]]])
    assert np.array_equal(w32[:bp[nb], 2], po[spec + 1] - po[spec])
    assert np.array_equal(w32[:bp[nb], 3] & 0xFF, env.hyp["maxfz"][pairs[:bp[nb]]])
    want = [[] for _ in range(nb)]
    for h in range(nh):
        if hi[h] > lo[h]:
            b0 = np.searchsorted(row0, lo[h], side="right") - 1
            b1 = np.searchsorted(row0, hi[h] - 1, side="right") - 1
            for bb in range(b0, b1 + 1):
                want[bb].append(h)
    assert bp[nb] == sum(len(w) for w in want)
    for bb in range(nb):
        assert sorted(pairs[bp[bb]:bp[bb + 1]].tolist()) == want[bb]
    tot = int(env.k1["win_off"][-1])
    assert tot > 1000
    assert np.array_equal(_np(j["counts"])[:tot], env.k1["counts"])


def test_join_row_range_slices_equal_oracle(env1):
    """An index cut into row-range slices (max_slice_entries a fraction of its postings, as an index beyond
    2^32 postings is cut): every slice's block starts are table rows and its blocks cover its rows once; the
    join K1 over the slices gives every window slot the oracle's count and the oracle's top-K lists; the
    capacities are fitted per slice."""
    from vulcan.gpu import engine as E
    env = env1
    p = type(env.p)(**env.p.to_dict())
    p.max_slice_entries = env.gi.n_postings // 3 + 1000
    gi = E.GpuIndex(env.l1, p, env.d["static"], env.d["var"])
    sl = gi.block_slices()
    assert gi.n_slices >= 3
    r = [(b["r0"], b["r1"]) for b in sl]
    assert r[0][0] == 0 and r[-1][1] == gi.n and all(a[1] == b_[0] for a, b_ in zip(r[:-1], r[1:]))
    po = _np(gi.post_off)
    for b in sl:
        row0 = _np(b["row0"]).astype(np.int64)
        assert row0[0] == b["r0"] and row0[-1] == b["r1"] and np.all(np.diff(row0) > 0)
        assert b["n_postings"] == po[b["r1"]] - po[b["r0"]] <= p.max_slice_entries
    with pytest.raises(ValueError, match="use block_slices"):
        gi.blocks()
    n = len(env.d["sp"])
    ds = E.DeviceSearch(gi, p, n, k1_impl="join")
    ds._join_alloc(ds.join, 16, 16)   # too small for every slice: fitted on the first run
    ds.run(E.upload_spectra(env.d["sp"], gi.dev))
    torch.cuda.synchronize()
    assert int(ds.status.item()) == 0
    tot = int(env.k1["win_off"][-1])
    assert np.array_equal(_np(ds.join["counts"])[:tot], env.k1["counts"])
    tk = _O().topk(env.vp, env.hix, env.hyp, env.k1, ds.stride)
    nc = _np(ds.n_cand)[:n]
    assert np.array_equal(nc, tk["n_cand"])
    m = (np.arange(ds.stride)[None, :] < nc[:, None]).ravel()
    for dv, k in ((ds.c_row, "row"), (ds.c_hyp, "hyp"), (ds.c_cnt, "cnt")):
        assert np.array_equal(_np(dv)[:n * ds.stride][m], tk[k][m]), k


def test_join_capacity_fit_and_overflow_flag(env1):
    """Pair / slot buffers far too small: the first run grows them to the batch (status clean, lists equal to
    the tile K1); a later batch that needs more than the fitted capacity raises status bit 8."""
    from vulcan.gpu import engine as E
    n = len(env1.d["sp"])
    raw = E.upload_spectra(env1.d["sp"], env1.gi.dev)
    ref = E.DeviceSearch(env1.gi, env1.p, n, k1_impl="tile")
    ref.run(raw)
    ds = E.DeviceSearch(env1.gi, env1.p, n, k1_impl="join")
    ds._join_alloc(ds.join, 16, 16)
    ds.run(raw)
    torch.cuda.synchronize()
    assert int(ds.status.item()) == 0
    assert ds.join["slot_cap"] >= int(env1.k1["win_off"][-1])
    nc = _np(ds.n_cand)[:n]
    assert np.array_equal(nc, _np(ref.n_cand)[:n])
    m = (np.arange(ds.stride)[None, :] < nc[:, None]).ravel()
    for a, c in ((ds.c_row, ref.c_row), (ds.c_hyp, ref.c_hyp), (ds.c_cnt, ref.c_cnt)):
        assert np.array_equal(_np(a)[:n * ds.stride][m], _np(c)[:n * ds.stride][m])
    ds._join_alloc(ds.join, 16, 16)                    # fitted already: no growth, overflow is flagged
    ds.run(raw)
    torch.cuda.synchronize()
    assert int(ds.status.item()) & 8


def test_select_topk_torch_equals_oracle(env1):
    from vulcan.gpu import engine as E
    pk = _device_peaks(env1)
    o_h, o_r, o_c, scored, matched, d_spec, d_maxfz = E.run_k1(env1.gi, env1.hyp, pk, env1.p, E.Stats())
    h, r, c = E.select_topk(o_h, o_r, o_c, d_spec, env1.gi, env1.p)
    stride = 2 * env1.p.topk
    tk = _O().topk(env1.vp, env1.hix, env1.hyp, env1.k1, stride)
    n = len(env1.d["sp"])
    m = (np.arange(stride)[None, :] < tk["n_cand"][:, None]).ravel()
    assert np.array_equal(_np(r), tk["row"][m])
    assert np.array_equal(_np(h), tk["hyp"][m])
    assert np.array_equal(_np(c), tk["cnt"][m])
    del n


# ------------------------------------------------------------------------------------------------ K2 / rank
@pytest.mark.parametrize("which", ["env0", "env1"])
def test_k2_and_rank_equal_oracle(which, request):
    env = request.getfixturevalue(which)
    ds, p = _topk_device(env)
    O = _O()
    n = len(env.d["sp"])
    nc = _np(ds.n_cand)[:n]
    m = (np.arange(ds.stride)[None, :] < nc[:, None]).ravel()
    rows = _np(ds.c_row)[:n * ds.stride][m]
    hyps = _np(ds.c_hyp)[:n * ds.stride][m]
    start = np.concatenate([[0], np.cumsum(nc)[:-1]])
    ref = O.k2(env.vp, env.hix, env.pk, rows, hyps, env.hyp["maxfz"], grp_start=start, grp_count=nc)
    assert len(rows) > 1000
    for k in ("matched_b", "matched_y", "longest_b", "longest_y", "summed_b", "summed_y"):
        assert np.array_equal(_np(ds.k2[k])[:n * ds.stride][m], ref[k]), k
    assert np.array_equal(_np(ds.k2["ppm"])[:n * ds.stride][m], ref["ppm"], equal_nan=True)
    hs = _np(ds.k2["hyperscore"])[:n * ds.stride][m]
    assert np.allclose(hs, ref["hyperscore"], rtol=HS_RTOL, atol=0)
    sel, nxt = O.rank(env.vp, start, nc, rows, ref["hyperscore"], ref["matched_b"], ref["matched_y"], 1)
    dsel = _np(ds.sel)[:n]
    flat_of = np.flatnonzero(m)          # oracle index -> strided device index
    want = np.where(sel >= 0, flat_of[np.maximum(sel, 0)], -1)
    assert np.array_equal(dsel, want)
    assert np.allclose(_np(ds.nxt)[:n], nxt, rtol=HS_RTOL, atol=0)


def test_search_end_to_end(env0):
    from vulcan.gpu import engine as E
    O = _O()
    tbl, st, dbg = E.search(env0.d["sp"], env0.gi, env0.l1, env0.p)
    ref = O.search(env0.vp, env0.hix, env0.d["sp"])
    got = dict(zip(tbl.column("spec_idx").to_pylist(), tbl.column("pf_row").to_pylist()))
    want = {i: int(r) for i, r in enumerate(ref["row"]) if r >= 0}
    assert got == want
    hs = dict(zip(tbl.column("spec_idx").to_pylist(), tbl.column("hyperscore").to_pylist()))
    assert all(abs(hs[i] - ref["hs"][i]) <= HS_RTOL * abs(ref["hs"][i]) for i in want)
    truth = env0.d["truth"]
    sig = truth >= 0
    assert np.mean([got.get(i) == truth[i] for i in np.flatnonzero(sig)]) > 0.9
    assert st.counts["n_rescored"] == int(ref["n_cand"].sum())


@pytest.mark.parametrize("impl", ["join", "tile"])
def test_search_window_peak_view_equals_oracle(env0, impl):
    """The K0 window view (top 4 peaks per 37.5 Th window, no total cap: up to 200 peaks a spectrum) through
    the whole search: rank-1 rows and hyperscores equal the oracle's search with the same view."""
    from vulcan.gpu import engine as E
    from vulcan.gpu._abi import make_params
    O = _O()
    p = type(env0.p)(**env0.p.to_dict())
    p.peak_mode, p.window_th, p.window_k, p.min_peaks, p.k1_impl = "window", 37.5, 4, 10, impl
    vp = make_params(p, env0.d["static"], env0.d["var"])
    assert vp.max_peaks == 200 and vp.peak_window_k == 4
    gi = E.GpuIndex(env0.l1, p, env0.d["static"], env0.d["var"])
    tbl, st, _ = E.search(env0.d["sp"], gi, env0.l1, p)
    ref = O.search(vp, env0.hix, env0.d["sp"])
    got = dict(zip(tbl.column("spec_idx").to_pylist(), tbl.column("pf_row").to_pylist()))
    want = {i: int(r) for i, r in enumerate(ref["row"]) if r >= 0}
    assert got == want and len(want) > 100
    hs = dict(zip(tbl.column("spec_idx").to_pylist(), tbl.column("hyperscore").to_pylist()))
    assert all(abs(hs[i] - ref["hs"][i]) <= HS_RTOL * abs(ref["hs"][i]) for i in want)
    assert st.counts["n_rescored"] == int(ref["n_cand"].sum())


def test_search_with_host_hypotheses(env0):
    """search(pre_hyp=...): host hypotheses whose validity mask equals K0's replace the device ones
    (pre_hyp_used 1), a mismatched mask falls back to the device hypotheses (-1); the table is the same."""
    from vulcan.gpu import engine as E
    from vulcan.gpu.host import pre_hypotheses
    sp = env0.d["sp"]
    base, st0, _ = E.search(sp, env0.gi, env0.l1, env0.p)
    assert st0.counts["pre_hyp_used"] == 0
    pre = pre_hypotheses(sp, env0.gi, env0.p)
    tbl, st, _ = E.search(sp, env0.gi, env0.l1, env0.p, pre_hyp=pre)
    assert st.counts["pre_hyp_used"] == 1 and tbl.equals(base)
    bad = (~pre[0], pre[1])
    tbl2, st2, _ = E.search(sp, env0.gi, env0.l1, env0.p, pre_hyp=bad)
    assert st2.counts["pre_hyp_used"] == -1 and tbl2.equals(base)


def test_search_report5_matches_oracle_ranking(env1):
    from vulcan.gpu import engine as E
    tbl, _, dbg = E.search(env1.d["sp"], env1.gi, env1.l1, env1.p, report=5)
    O = _O()
    k2 = E.flat_candidates(dbg["device"], len(env1.d["sp"]))
    spec = k2["spec"]
    us, cnt = np.unique(spec, return_counts=True)
    start = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    rows = k2["r"].astype(np.int32)
    sel, nxt = O.rank(env1.vp, start, cnt, rows, k2["hyperscore"], k2["matched_b"], k2["matched_y"], 5)
    want = [(int(spec[i]), int(rows[i])) for i in sel if i >= 0]
    got = list(zip(tbl.column("spec_idx").to_pylist(), tbl.column("pf_row").to_pylist()))
    assert got == want


def test_psm_features_equal_oracle(env1):
    """vh_psm_features (the PSM table's feature columns, computed on the device) against the numpy restatement
    of psm_table's arithmetic on the same K2 outputs and oracle rank selections; the table's column order and
    types follow the PSM schema."""
    import pyarrow as pa
    from oracle import psm as OP
    from vulcan.gpu import engine as E
    report = 5
    tbl, _, dbg = E.search(env1.d["sp"], env1.gi, env1.l1, env1.p, report=report)
    O = _O()
    k2 = E.flat_candidates(dbg["device"], len(env1.d["sp"]))
    spec = k2["spec"]
    us, cnt = np.unique(spec, return_counts=True)
    start = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    rows = k2["r"].astype(np.int32)
    sel, nxt = O.rank(env1.vp, start, cnt, rows, k2["hyperscore"], k2["matched_b"], k2["matched_y"], report)
    hyp = dbg["hyp"]
    n = len(env1.d["sp"])
    n_h = len(hyp["spec"])
    hyp_off = np.concatenate([[0], np.cumsum(np.bincount(np.asarray(hyp["spec"]), minlength=n)[:n])])
    ds = dbg["device"]
    ref = OP.features(sel, nxt, us, rows, k2["h"], k2["k1c"],
                      {"hs": k2["hyperscore"], "sb": k2["summed_b"], "sy": k2["summed_y"], "ppm": k2["ppm"],
                       "mb": k2["matched_b"], "my": k2["matched_y"], "lb": k2["longest_b"], "ly": k2["longest_y"]},
                      hyp_off, hyp["z"], hyp["iso"], hyp["pm"], _np(ds.scored)[:n_h], _np(ds.matched)[:n_h],
                      _np(dbg["peaks"]["tic"])[:n], env1.l1.mass32, env1.l1.length, report)
    ok = ref["rank"] > 0
    assert tbl.num_rows == int(ok.sum()) > 0
    for k, v in ref.items():
        got, want = tbl.column(k).to_numpy(), v[ok]
        if k == "poisson":
            fin = np.isfinite(want)
            assert np.array_equal(np.isfinite(got), fin) and np.array_equal(np.isnan(got), np.isnan(want))
            assert np.allclose(got[fin], want[fin], rtol=HS_RTOL, atol=0), k
        else:
            assert got.dtype == want.dtype and np.array_equal(got, want), k
    names = ["run", "spec_idx", "scan", "rt", "rank", "pf_row", "pf_id", "charge", "isotope_error", "expmass",
             "calcmass", "delta_mass_ppm", "hyperscore", "delta_next", "delta_best", "matched_b", "matched_y",
             "matched_peaks", "summed_intensity", "matched_intensity_pct", "average_ppm", "longest_b", "longest_y",
             "longest_y_pct", "peptide_len", "missed_cleavages", "n_var", "k1_count", "scored_candidates", "poisson"]
    assert tbl.column_names == names
    types = {"spec_idx": pa.int64(), "rank": pa.int32(), "pf_row": pa.int64(), "charge": pa.int8(),
             "isotope_error": pa.int8(), "expmass": pa.float32(), "hyperscore": pa.float64(),
             "matched_b": pa.int16(), "average_ppm": pa.float32(), "k1_count": pa.int32(), "poisson": pa.float64()}
    for k, t in types.items():
        assert tbl.schema.field(k).type == t, k


def test_deterministic_rerun(env1):
    a, _ = _topk_device(env1)
    ra = [_np(t).copy() for t in (a.n_cand, a.c_row, a.c_hyp, a.c_cnt, a.k2["hyperscore"], a.sel)]
    b, _ = _topk_device(env1)
    rb = [_np(t) for t in (b.n_cand, b.c_row, b.c_hyp, b.c_cnt, b.k2["hyperscore"], b.sel)]
    n = len(env1.d["sp"])
    nc = ra[0][:n]
    m = (np.arange(a.stride)[None, :] < nc[:, None]).ravel()
    assert np.array_equal(ra[0][:n], rb[0][:n])
    for x, y in zip(ra[1:5], rb[1:5]):
        assert np.array_equal(x[:n * a.stride][m], y[:n * a.stride][m])
    assert np.array_equal(ra[5][:n], rb[5][:n])


def _spectra_range(sp, a: int, b: int):
    from vulcan.gpu.host import Spectra
    off = sp.offsets[a:b + 1]
    return Spectra(sp.run, sp.scan[a:b], sp.index[a:b], sp.rt[a:b], sp.precursor_mz[a:b], sp.charge[a:b],
                   sp.mz[off[0]:off[-1]], sp.intensity[off[0]:off[-1]], off - off[0])


@pytest.mark.parametrize("depth", [1, 2, 3])
def test_pipeline_steps_equal_oracle(env1, depth):
    """Two batches of different sizes dealt alternately to a SearchPipeline, no synchronisation between
    submissions: every resident step's rank 1 and candidate counts equal the oracle's whole search of its batch,
    and its candidate lists equal one DeviceSearch run of the same batch."""
    from vulcan.gpu import engine as E
    sp = env1.d["sp"]
    n = len(sp)
    cut = n * 3 // 5
    parts = [(0, cut), (cut, n)]
    ref = _O().search(env1.vp, env1.hix, sp)
    raws = [E.upload_spectra(_spectra_range(sp, a, b), env1.gi.dev) for a, b in parts]
    solo = []
    for raw in raws:
        ds = E.DeviceSearch(env1.gi, env1.p, n)
        ds.run(raw)
        solo.append(ds)
    pipe = E.SearchPipeline(env1.gi, env1.p, n, depth=depth)
    n_steps = 2 * depth + 1
    for i in range(n_steps):
        pipe.run(raws[i % 2])
    pipe.synchronize()
    torch.cuda.synchronize()
    for i in range(n_steps - depth, n_steps):
        a, b = parts[i % 2]
        ds = pipe.slot(i)
        assert int(ds.status.item()) == 0
        m = b - a
        assert np.array_equal(ds.rank1_rows(m), ref["row"][a:b]), i
        nc = _np(ds.n_cand)[:m]
        assert np.array_equal(nc, ref["n_cand"][a:b]), i
        one = solo[i % 2]
        mk = (np.arange(ds.stride)[None, :] < nc[:, None]).ravel()
        for x, y in ((ds.c_row, one.c_row), (ds.c_hyp, one.c_hyp), (ds.c_cnt, one.c_cnt),
                     (ds.k2["hyperscore"], one.k2["hyperscore"])):
            assert np.array_equal(_np(x)[:m * ds.stride][mk], _np(y)[:m * ds.stride][mk]), i
    with pytest.raises(ValueError, match="not resident"):
        pipe.slot(n_steps - depth - 1)


def test_label_blindness_gpu():
    from vulcan.gpu import engine as E
    from vulcan.index import synth
    from vulcan.index.build import build_tables
    w = synth.WORKLOADS[1]
    prots = synth.proteome(300, 77)
    a, la = build_tables(prots, w.index_params())
    b, _ = build_tables(prots, w.index_params(), shuffle_rows_seed=3, swap_labels=True)
    sp, _, _ = synth.spectra(a, la, 1, n=300, seed=9)
    p = w.search_params()
    outs = []
    for l1 in (a, b):
        gi = E.GpuIndex(l1, p, STATIC, w.var_mods)
        outs.append(E.search(sp, gi, l1, p, report=5)[0])
    assert outs[0].equals(outs[1])


def test_spectrum_order_sorted_permutation(env1):
    """vh_spectrum_order: a permutation of the spectra by the band of the first-hypothesis window row (65,536
    equal bands of the index rows; spectra without hypotheses last), stable."""
    ds, _ = _topk_device(env1)
    n = len(env1.d["sp"])
    order = _np(ds.order)[:n]
    assert np.array_equal(np.sort(order), np.arange(n))
    off = env1.hyp["hyp_off"]
    n_rows = int(ds.gi.n)
    lo = env1.hyp["lo"][np.minimum(off[:-1], max(len(env1.hyp["lo"]) - 1, 0))].astype(np.int64)
    key = np.where(off[1:] > off[:-1], (np.clip(lo, 0, n_rows - 1) << 16) // n_rows, 65535)
    assert len(np.unique(key)) > 8                   # the bands spread the spectra
    k = key[order]
    assert np.all(np.diff(k) >= 0)
    same = np.diff(k) == 0
    assert np.all(np.diff(order)[same] > 0)          # stable: equal keys keep spectrum order


# ------------------------------------------------------------------------------------------------ search modes
@pytest.fixture(scope="module")
def env_modes():
    """300-protein workload-1-semantics table, spectra with 40% precursor-only shifts (the open-search workload's
    generator), open window (-150, 500) Da and two mass offsets."""
    from vulcan.gpu import engine as E
    from vulcan.gpu._abi import make_params, max_hyp_per_spectrum, set_modes
    from vulcan.index import synth
    from vulcan.index.build import build_tables
    from vulcan.ptm.search import ModeParams, mode_search_params
    print("[gpu] building modes env", flush=True)
    w1 = synth.WORKLOADS[1]
    l1, lab = build_tables(synth.proteome(300, 103), w1.index_params())
    sp, truth, meta = synth.spectra(l1, lab, 3, n=300, seed=33, index_params=w1.index_params())
    mp = ModeParams(offsets=(15.994915, 42.010565), open_da=(-150.0, 500.0), open_tile=2048)
    p = mode_search_params(w1.search_params(), mp)
    p.topk = 20
    gi = E.GpuIndex(l1, p, STATIC, w1.var_mods)
    vp = set_modes(make_params(p, STATIC, w1.var_mods), mp.offsets, mp.open_da)
    O = _O()
    hix = O.HostIndex(l1)
    pk = O.k0(vp, sp)
    hyp = O.hypotheses(vp, sp, pk["valid"], l1.mass32, max_hyp_per_spectrum(p, 2, True))
    k1 = O.k1_counts(vp, hix, pk, hyp)
    return {"l1": l1, "sp": sp, "truth": truth, "shift": meta["shift"], "mp": mp, "p": p, "gi": gi, "vp": vp,
            "hix": hix, "pk": pk, "hyp": hyp, "k1": k1, "var": w1.var_mods}


@pytest.fixture(scope="module")
def env_open500():
    """The open-search workload's parameters at full width: 1,000-protein workload-1-semantics table, 400
    spectra with 40% precursor-only shifts U(-150, +400) Da, open window +-500 Da, top-K 100 + closure."""
    from vulcan.gpu import engine as E
    from vulcan.gpu._abi import make_params, max_hyp_per_spectrum, set_modes
    from vulcan.index import synth
    from vulcan.index.build import build_tables
    from vulcan.ptm.search import ModeParams, mode_search_params
    print("[gpu] building open +-500 Da env", flush=True)
    w1, w3 = synth.WORKLOADS[1], synth.WORKLOADS[3]
    l1, lab = build_tables(synth.proteome(1000, 103), w1.index_params())
    sp, truth, meta = synth.spectra(l1, lab, 3, n=400, seed=35, index_params=w1.index_params())
    mp = ModeParams(open_da=w3.open_da)
    p = mode_search_params(w3.search_params(), mp)
    assert p.topk == 100 and mp.open_da == (-500.0, 500.0)
    gi = E.GpuIndex(l1, p, STATIC, w1.var_mods)
    vp = set_modes(make_params(p, STATIC, w1.var_mods), mp.offsets, mp.open_da)
    O = _O()
    hix = O.HostIndex(l1)
    pk = O.k0(vp, sp)
    hyp = O.hypotheses(vp, sp, pk["valid"], l1.mass32, max_hyp_per_spectrum(p, 0, True))
    k1 = O.k1_counts(vp, hix, pk, hyp)
    print(f"[gpu] open env: {len(l1)} rows, {int(k1['win_off'][-1])} window rows", flush=True)
    return {"l1": l1, "sp": sp, "truth": truth, "shift": meta["shift"], "mp": mp, "p": p, "gi": gi, "vp": vp,
            "hix": hix, "pk": pk, "hyp": hyp, "k1": k1, "var": w1.var_mods}


@pytest.mark.parametrize("which", ["env_modes", "env_open500"])
def test_modes_device_pipeline_equals_oracle(which, request):
    """Mode hypotheses, K1 top-K over closed + offset + open windows, shifted K2 and rank 1 == the oracle."""
    from vulcan.gpu import engine as E
    e = request.getfixturevalue(which)
    O = _O()
    ds = E.DeviceSearch(e["gi"], e["p"], len(e["sp"]), offsets=e["mp"].offsets, open_da=e["mp"].open_da)
    ds.run(E.upload_spectra(e["sp"], e["gi"].dev))
    torch.cuda.synchronize()
    assert int(ds.status.item()) == 0
    hd = E.hyp_to_host(ds.hyp)
    for k in ("hyp_off", "spec", "z", "iso", "pm", "M", "lo", "hi", "maxfz", "mode"):
        assert np.array_equal(hd[k], e["hyp"][k]), k
    assert np.array_equal(hd["off"], e["hyp"]["off"], equal_nan=True)
    tk = O.topk(e["vp"], e["hix"], e["hyp"], e["k1"], ds.stride)
    n = len(e["sp"])
    nc = _np(ds.n_cand)[:n]
    assert np.array_equal(nc, tk["n_cand"])
    m = (np.arange(ds.stride)[None, :] < nc[:, None]).ravel()
    rows = _np(ds.c_row)[:n * ds.stride][m]
    hyps = _np(ds.c_hyp)[:n * ds.stride][m]
    assert np.array_equal(rows, tk["row"][m]) and np.array_equal(hyps, tk["hyp"][m])
    assert np.array_equal(_np(ds.c_cnt)[:n * ds.stride][m], tk["cnt"][m])
    hm, hM = e["hyp"]["mode"], e["hyp"]["M"]
    shift = np.where(hm[hyps] == 0, np.float32(0), (hM[hyps] - e["l1"].mass32[rows]).astype(np.float32))
    assert np.array_equal(_np(ds.c_shift)[:n * ds.stride][m], shift.astype(np.float32))
    start = np.concatenate([[0], np.cumsum(nc)[:-1]])
    ref = O.k2(e["vp"], e["hix"], e["pk"], rows, hyps, e["hyp"]["maxfz"], grp_start=start, grp_count=nc,
               shift=shift)
    assert (shift != 0).sum() > 100
    for k in ("matched_b", "matched_y", "longest_b", "longest_y", "summed_b", "summed_y"):
        assert np.array_equal(_np(ds.k2[k])[:n * ds.stride][m], ref[k]), k
    assert np.array_equal(_np(ds.k2["ppm"])[:n * ds.stride][m], ref["ppm"], equal_nan=True)
    assert np.allclose(_np(ds.k2["hyperscore"])[:n * ds.stride][m], ref["hyperscore"], rtol=HS_RTOL, atol=0)
    sel, _ = O.rank(e["vp"], start, nc, rows, ref["hyperscore"], ref["matched_b"], ref["matched_y"], 1)
    flat_of = np.flatnonzero(m)
    assert np.array_equal(_np(ds.sel)[:n], np.where(sel >= 0, flat_of[np.maximum(sel, 0)], -1))


def test_search_modes_recovers_shifted_spectra_and_is_label_blind(env_modes):
    """Shifted spectra are found through offset / open hypotheses (reference test_ptm_modes.py:114-117 bar:
    > 50% of rank-1 PSMs of shifted spectra come from mode > 0), and role swap + pair permutation give identical
    tables (test_ptm_modes.py:94-117)."""
    import pyarrow as pa

    from vulcan.gpu import engine as E
    from vulcan.index import synth
    from vulcan.index.build import build_tables
    from vulcan.ptm.search import search_modes
    e = env_modes
    tbl, st = search_modes(e["sp"], e["gi"], e["l1"], e["p"], e["mp"], report=5)
    mode = tbl.column("mode").to_numpy()
    r1 = tbl.column("rank").to_numpy() == 1
    spec = tbl.column("spec_idx").to_numpy()
    shifted = set(np.flatnonzero(np.abs(e["shift"]) > 1.0).tolist())
    sel = r1 & np.isin(spec, list(shifted))
    assert sel.sum() > 20 and (mode[sel] > 0).mean() > 0.5
    w1 = synth.WORKLOADS[1]
    b, _ = build_tables(synth.proteome(300, 103), w1.index_params(), shuffle_rows_seed=5, swap_labels=True)
    gib = E.GpuIndex(b, e["p"], STATIC, e["var"])
    tb, _ = search_modes(e["sp"], gib, b, e["p"], e["mp"], report=5)
    ho = lambda t: t.set_column(t.column_names.index("hyp_offset"), "hyp_offset",  # noqa: E731
                                pa.array(t.column("hyp_offset").to_numpy().view(np.uint32)))
    assert ho(tbl).equals(ho(tb))


@pytest.mark.parametrize("impl", ["join", "tile"])
@pytest.mark.parametrize("which", ["env_modes", "env_open500"])
def test_modes_join_equals_oracle(which, impl, request):
    """The join K1 (each open hypothesis spans many blocks) and the tile K1 forced on the mode search (mass
    offsets + a 650 Da open window, or the +-500 Da window) give the oracle's candidate lists, as the default
    wide-window K1 does."""
    from vulcan.gpu import engine as E
    e = request.getfixturevalue(which)
    ds = E.DeviceSearch(e["gi"], e["p"], len(e["sp"]), offsets=e["mp"].offsets, open_da=e["mp"].open_da,
                        k1_impl=impl)
    ds.run(E.upload_spectra(e["sp"], e["gi"].dev))
    torch.cuda.synchronize()
    assert int(ds.status.item()) == 0
    tk = _O().topk(e["vp"], e["hix"], e["hyp"], e["k1"], ds.stride)
    n = len(e["sp"])
    nc = _np(ds.n_cand)[:n]
    assert np.array_equal(nc, tk["n_cand"])
    m = (np.arange(ds.stride)[None, :] < nc[:, None]).ravel()
    for dv, k in ((ds.c_row, "row"), (ds.c_hyp, "hyp"), (ds.c_cnt, "cnt")):
        assert np.array_equal(_np(dv)[:n * ds.stride][m], tk[k][m]), k
    nh = len(e["hyp"]["spec"])
    assert np.array_equal(_np(ds.scored)[:nh].astype(np.uint32), e["k1"]["scored"])
    assert np.array_equal(_np(ds.matched)[:nh].astype(np.uint64), e["k1"]["matched"])


@pytest.mark.parametrize("case", ["unknown_charges", "u32_fallback"])
def test_wide_k1_edge_cases_equal_oracle(env_open500, case):
    """Wide-window K1 vs the oracle where its bucket walk has extra paths: spectra without a known charge (open
    and closed hypotheses of several maxfz over one bucket: counts added charge level by level), and a 1.5 Da
    fragment tolerance that pushes a bucket's matches past 2^16 (the bucket is counted again in u32 sub-chunks
    after the first level's offers)."""
    from vulcan.gpu import engine as E
    from vulcan.gpu._abi import make_params, max_hyp_per_spectrum, set_modes
    from vulcan.gpu.host import Spectra
    e = env_open500
    O = _O()
    sp = e["sp"]
    n = 120
    off = sp.offsets[:n + 1]
    z = sp.charge[:n].copy()
    p = type(e["p"])(**e["p"].to_dict())
    if case == "unknown_charges":
        z[::2] = 0
    else:
        p.fragment_ppm = (-1.5, 1.5)
    sub = Spectra(sp.run, sp.scan[:n], sp.index[:n], sp.rt[:n], sp.precursor_mz[:n], z, sp.mz[:off[-1]],
                  sp.intensity[:off[-1]], off.copy())
    gi = e["gi"] if case == "unknown_charges" else E.GpuIndex(e["l1"], p, STATIC, e["var"])
    vp = set_modes(make_params(p, STATIC, e["var"]), e["mp"].offsets, e["mp"].open_da)
    pk = O.k0(vp, sub)
    hyp = O.hypotheses(vp, sub, pk["valid"], e["l1"].mass32, max_hyp_per_spectrum(p, 0, True))
    k1 = O.k1_counts(vp, e["hix"], pk, hyp)
    print(f"[gpu] wide K1 {case}: {int(k1['counts'].sum())} matches over {len(hyp['spec'])} hypotheses", flush=True)
    ds = E.DeviceSearch(gi, p, n, open_da=e["mp"].open_da, k1_impl="wide")
    ds.run(E.upload_spectra(sub, gi.dev))
    torch.cuda.synchronize()
    assert int(ds.status.item()) == 0
    tk = O.topk(vp, e["hix"], hyp, k1, ds.stride)
    nc = _np(ds.n_cand)[:n]
    assert np.array_equal(nc, tk["n_cand"])
    m = (np.arange(ds.stride)[None, :] < nc[:, None]).ravel()
    for dv, k in ((ds.c_row, "row"), (ds.c_hyp, "hyp"), (ds.c_cnt, "cnt")):
        assert np.array_equal(_np(dv)[:n * ds.stride][m], tk[k][m]), k
    nh = len(hyp["spec"])
    assert np.array_equal(_np(ds.scored)[:nh].astype(np.uint32), k1["scored"])
    assert np.array_equal(_np(ds.matched)[:nh].astype(np.uint64), k1["matched"])
    if case == "u32_fallback":     # some bucket of some spectrum really held 2^16 matches or more
        w = np.maximum(hyp["hi"] - hyp["lo"], 0)
        rows = np.repeat(hyp["lo"], w) + (np.arange(int(w.sum())) - np.repeat(k1["win_off"][:-1], w))
        spec = np.repeat(hyp["spec"], w)
        open_h = np.repeat(hyp["mode"] == 2, w)
        tot = np.bincount(spec[open_h] * 100000 + (rows[open_h] >> 15), weights=k1["counts"][open_h])
        assert tot.max() >= 65536
