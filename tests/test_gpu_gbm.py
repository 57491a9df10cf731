"""vh_gbm_decision (through GpuHGB) vs sklearn's decision_function and the oracle: bit-identical f64 on the
reference's random-model case, feature counts across the three row-tile sizes, multi-chunk models, ragged and
empty inputs; a malformed node array is reported through the status word."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)
sk = pytest.importorskip("sklearn.ensemble")

from test_gbm_oracle import _clf, _test_rows  # noqa: E402


def test_reference_random_model_bit_identical():
    from oracle import gbm as O
    from vulcan.rescore.fast import GpuHGB
    print("[gpu] gbm: reference random model, 30001 rows", flush=True)
    m = _clf()
    Xt = _test_rows(m, 30001, 7)
    got = GpuHGB(m).decision_function(Xt, chunk=7000)
    assert np.array_equal(got, m.decision_function(Xt))
    assert np.array_equal(got, O.raw_predict(O.trees_of(m), O.baseline_of(m), Xt))


@pytest.mark.parametrize("nf,n", [(1, 1), (15, 257), (31, 1000), (40, 129), (63, 5003)])
def test_feature_counts_and_ragged_sizes(nf, n):
    from vulcan.rescore.fast import GpuHGB
    print(f"[gpu] gbm: nf={nf} n={n}", flush=True)
    rng = np.random.default_rng(nf)
    X = rng.normal(size=(3000, nf))
    X[rng.random(X.shape) < 0.03] = np.nan
    y = np.nan_to_num(X).sum(1) + rng.normal(size=3000)
    m = sk.HistGradientBoostingRegressor(max_iter=30, max_leaf_nodes=63, random_state=0).fit(X, y)
    Xt = _test_rows(m, max(n, 330), nf, seed=nf + 1)[:n] if n >= 330 else rng.normal(size=(n, nf))
    assert np.array_equal(GpuHGB(m).decision_function(Xt), m.predict(Xt))


def test_many_chunks_and_large_batch():
    """300 trees of up to 63 nodes (about 19 chunks); 2M rows: a sample checked against sklearn, duplicated
    rows give identical scores."""
    from vulcan.rescore.fast import GpuHGB
    print("[gpu] gbm: 300 trees, 2M rows", flush=True)
    rng = np.random.default_rng(9)
    X = rng.normal(size=(20000, 24))
    y = (X[:, 0] * X[:, 1] + np.sin(X[:, 2] * 3) + rng.normal(size=20000) * 0.3 > 0).astype(int)
    m = sk.HistGradientBoostingClassifier(max_iter=300, max_leaf_nodes=32, early_stopping=False,
                                          random_state=0).fit(X, y)
    g = GpuHGB(m)
    assert g.n_chunks > 5
    base = rng.normal(size=(1000, 24))
    idx = rng.integers(0, 1000, 2_000_000)
    Xd = torch.from_numpy(base[idx]).cuda()
    out = g.decision_device(Xd).cpu().numpy()
    g.check()
    ref = m.decision_function(base)
    assert np.array_equal(out, ref[idx])


def _node_view(g):
    return g.nodes.cpu().numpy().copy().view(np.dtype([("v", "<f8"), ("left", "<i4"), ("feature", "<i2"),
                                                       ("ml", "u1"), ("leaf", "u1")]))


def test_empty_input_and_malformed_models():
    from vulcan.rescore.fast import GpuHGB
    m = _clf(iters=5)
    g = GpuHGB(m)
    assert g.decision_function(np.zeros((0, 7))).shape == (0,)
    ref = g.decision_function(np.zeros((10, 7)))
    # leaves not marked as leaves: every walk ends on a node that is not a leaf -> status word
    nodes = _node_view(g)
    nodes["leaf"] = 0
    g.nodes = torch.from_numpy(nodes.view(np.uint8)).cuda()
    with pytest.raises(RuntimeError, match="malformed"):
        g.decision_function(np.zeros((10, 7)))
    # children and features far outside the chunk / the feature count: clamped in the kernel (every read stays
    # inside the staged arrays); the call returns, flagged or not
    g2 = GpuHGB(m)
    nodes = _node_view(g2)
    nodes["left"][nodes["leaf"] == 0] = 1 << 20
    nodes["feature"][nodes["leaf"] == 0] = 30000
    g2.nodes = torch.from_numpy(nodes.view(np.uint8)).cuda()
    try:
        g2.decision_function(np.zeros((10, 7)))
    except RuntimeError as e:
        assert "malformed" in str(e)
    assert np.array_equal(GpuHGB(m).decision_function(np.zeros((10, 7))), ref)


def test_trees_near_the_chunk_limit():
    """512-leaf trees (about 1,000 nodes: one tree per LDS chunk), bit-identical to sklearn."""
    from vulcan.rescore.fast import GpuHGB
    print("[gpu] gbm: 512-leaf trees", flush=True)
    rng = np.random.default_rng(8)
    X = rng.normal(size=(20000, 6))
    y = np.sin(X[:, 0] * 3) + X[:, 1] * X[:, 2] + rng.normal(size=20000) * 0.1
    m = sk.HistGradientBoostingRegressor(max_iter=6, max_leaf_nodes=512, min_samples_leaf=2, max_depth=None,
                                         random_state=0).fit(X, y)
    g = GpuHGB(m)
    assert g.n_chunks >= 5
    Xt = _test_rows(m, 7001, 6)
    assert np.array_equal(g.decision_function(Xt), m.predict(Xt))


def test_malformed_chunk_table_is_clamped_and_reported():
    """A chunk table whose chunk exceeds the LDS chunk size, runs past the node array or names trees outside the
    model: the kernel clamps it into bounds (every read stays inside the arrays) and sets status bit 2."""
    from vulcan.rescore.fast import GpuHGB
    m = _clf(iters=20)
    X = np.zeros((64, 7))
    for bad in ("oversized", "past_end", "trees"):
        g = GpuHGB(m)
        cn0 = g.chunk_node0.cpu().numpy().copy()
        ct0 = g.chunk_tree0.cpu().numpy().copy()
        if bad == "oversized":
            cn0[-1] = cn0[-2] + 5000                 # last chunk: 5,000 nodes (limit 1,024)
        elif bad == "past_end":
            cn0[-1] = g.M + 700                      # beyond the node array
        else:
            ct0[-1] = g.T + 40                       # trees beyond the model
        g.chunk_node0 = torch.from_numpy(cn0).cuda()
        g.chunk_tree0 = torch.from_numpy(ct0).cuda()
        print(f"[gpu] gbm malformed chunk table: {bad}", flush=True)
        with pytest.raises(RuntimeError, match="chunk table"):
            g.decision_function(X)
    assert np.array_equal(GpuHGB(m).decision_function(X), m.decision_function(X))
