"""Gradient-boosted tree scorer, CPU side: the oracle (oracle/gbm.py) against sklearn's own predictions and a
model computed by hand; the packed node layout (breadth-first renumbering, chunks of whole trees) walked on the
host against sklearn (no GPU)."""
import numpy as np
import pytest

sk = pytest.importorskip("sklearn.ensemble")


def _O():
    from oracle import gbm
    return gbm


def _clf(seed=0, n=20000, nf=7, iters=60, depth=5):
    """The reference's random-model case (tests/rescore/test_fast_equal.py:15-26)."""
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, nf))
    X[rng.random(X.shape) < 0.05] = np.nan
    y = (np.nan_to_num(X[:, 0]) + 0.5 * np.nan_to_num(X[:, 1]) ** 2 + rng.normal(size=len(X)) > 0.3).astype(int)
    return sk.HistGradientBoostingClassifier(max_iter=iters, max_depth=depth, random_state=0).fit(X, y)


def _test_rows(m, n, nf, seed=1):
    rng = np.random.default_rng(seed)
    Xt = rng.normal(size=(n, nf))
    Xt[rng.random(Xt.shape) < 0.05] = np.nan
    Xt[:50, 2 % nf] = m._predictors[0][0].nodes["num_threshold"][0]      # values exactly on a threshold
    thr = np.concatenate([it[0].nodes["num_threshold"] for it in m._predictors])
    Xt[50:300] = rng.choice(thr, size=(250, nf))
    Xt[300:310, 0] = np.inf
    Xt[310:320, 0] = -np.inf
    return Xt


def walk_packed(nodes, roots, cn0, ct0, depth, baseline, X):
    """Host walk of the packed records (VhGbmNode layout: children chunk-local, leaves point to themselves),
    for the packing test."""
    out = np.zeros(len(X))
    chunk_of = np.searchsorted(ct0, np.arange(len(roots)), side="right") - 1
    for i, x in enumerate(X):
        acc = 0.0
        acc += baseline
        for t in range(len(roots)):
            base = int(cn0[chunk_of[t]])
            k = int(roots[t]) - base
            for _ in range(depth + 1):
                nd = nodes[base + k]
                if nd["is_leaf"]:
                    assert nd["left"] == k and nd["feature"] == -1
                    break
                v = x[nd["feature"]]
                go_left = bool(nd["missing_left"]) if np.isnan(v) else v <= nd["v"]
                k = int(nd["left"]) + (0 if go_left else 1)
            assert nodes[base + k]["is_leaf"]
            acc += nodes[base + k]["v"]
        out[i] = acc
    return out


def test_oracle_equals_sklearn_classifier():
    O = _O()
    m = _clf()
    Xt = _test_rows(m, 30001, 7)
    assert np.array_equal(O.raw_predict(O.trees_of(m), O.baseline_of(m), Xt), m.decision_function(Xt))


def test_oracle_equals_sklearn_regressor_and_single_leaf_trees():
    O = _O()
    rng = np.random.default_rng(3)
    X = rng.normal(size=(5000, 12))
    X[rng.random(X.shape) < 0.1] = np.nan
    y = np.nan_to_num(X[:, 3]) * 2 - np.nan_to_num(X[:, 7]) + rng.normal(size=5000) * 0.1
    m = sk.HistGradientBoostingRegressor(max_iter=40, max_leaf_nodes=15, random_state=0).fit(X, y)
    Xt = _test_rows(m, 4000, 12, seed=4)
    assert np.array_equal(O.raw_predict(O.trees_of(m), O.baseline_of(m), Xt), m.predict(Xt))
    # a constant target: every tree is a single leaf (depth 0)
    mc = sk.HistGradientBoostingRegressor(max_iter=5).fit(X, np.full(5000, 1.25))
    assert all(len(t) == 1 for t in _O().trees_of(mc))
    assert np.array_equal(O.raw_predict(O.trees_of(mc), O.baseline_of(mc), Xt), mc.predict(Xt))


def test_oracle_on_a_model_written_by_hand():
    """Two trees: t0 splits x0 <= 1 (NaN right) into leaves 0.5 / -0.25; t1 splits x1 <= -2 (NaN left) into a
    leaf 1.0 and a split on x0 <= 3 (NaN left) with leaves 0.125 / 2.0. Baseline 0.75."""
    from sklearn.ensemble._hist_gradient_boosting.common import PREDICTOR_RECORD_DTYPE as D
    t0 = np.zeros(3, D)
    t0[0] = (0, 0, 0, 1.0, 0, 1, 2, 0, 0, 0, 0, 0, 0)
    t0[1]["value"], t0[1]["is_leaf"], t0[1]["depth"] = 0.5, 1, 1
    t0[2]["value"], t0[2]["is_leaf"], t0[2]["depth"] = -0.25, 1, 1
    t1 = np.zeros(5, D)
    t1[0]["feature_idx"], t1[0]["num_threshold"], t1[0]["missing_go_to_left"] = 1, -2.0, 1
    t1[0]["left"], t1[0]["right"] = 3, 1
    t1[3]["value"], t1[3]["is_leaf"], t1[3]["depth"] = 1.0, 1, 1
    t1[1]["feature_idx"], t1[1]["num_threshold"], t1[1]["missing_go_to_left"], t1[1]["depth"] = 0, 3.0, 1, 1
    t1[1]["left"], t1[1]["right"] = 4, 2
    t1[4]["value"], t1[4]["is_leaf"], t1[4]["depth"] = 0.125, 1, 2
    t1[2]["value"], t1[2]["is_leaf"], t1[2]["depth"] = 2.0, 1, 2
    X = np.array([[0.0, -5.0], [1.0, 0.0], [2.0, 0.0], [np.nan, np.nan], [5.0, -2.0], [3.0, 1.0]])
    want = np.array([0.75 + 0.5 + 1.0, 0.75 + 0.5 + 0.125, 0.75 - 0.25 + 0.125, 0.75 - 0.25 + 1.0,
                     0.75 - 0.25 + 1.0, 0.75 - 0.25 + 0.125])
    assert np.array_equal(_O().raw_predict([t0, t1], 0.75, X), want)
    from vulcan.rescore.fast import pack_trees
    packed = pack_trees([t0, t1])
    assert np.array_equal(walk_packed(*packed, 0.75, X), want)


def test_packed_layout_equals_sklearn():
    from vulcan.rescore.fast import pack_trees
    rng = np.random.default_rng(5)
    X = rng.normal(size=(4000, 5))
    y = (X[:, 0] * X[:, 1] + rng.normal(size=4000) > 0).astype(int)
    m = sk.HistGradientBoostingClassifier(max_iter=80, max_leaf_nodes=31, random_state=0).fit(X, y)
    nodes, roots, cn0, ct0, depth = pack_trees(_O().trees_of(m))
    assert len(cn0) > 2                                    # more than one chunk
    assert ct0[0] == 0 and ct0[-1] == len(roots) and np.all(np.diff(cn0) <= 1024) and cn0[-1] == len(nodes)
    for c in range(len(cn0) - 1):                          # whole trees: every root of a chunk lies in it
        r = roots[ct0[c]:ct0[c + 1]]
        assert np.all((r >= cn0[c]) & (r < cn0[c + 1]))
    idx = np.arange(len(nodes))
    base = cn0[np.searchsorted(cn0, idx, side="right") - 1]       # chunk start of every node
    inner = nodes["is_leaf"] == 0
    assert np.all(nodes["left"][inner] + base[inner] > idx[inner])  # breadth-first: children after the parent
    assert np.all(nodes["left"][~inner] + base[~inner] == idx[~inner])    # a leaf points to itself
    assert np.all(nodes["left"][inner] + 1 < np.diff(cn0)[np.searchsorted(cn0, idx[inner], side="right") - 1])
    Xt = _test_rows(m, 300, 5)
    assert np.array_equal(walk_packed(nodes, roots, cn0, ct0, depth, _O().baseline_of(m), Xt),
                          m.decision_function(Xt))
    with pytest.raises(ValueError, match="categorical"):
        bad = [t.copy() for t in _O().trees_of(m)]
        bad[0]["is_categorical"][0] = 1
        pack_trees(bad)


def test_large_trees_fill_one_chunk_each_and_oversized_trees_are_refused():
    from vulcan.rescore.fast import pack_trees
    rng = np.random.default_rng(8)
    X = rng.normal(size=(20000, 6))
    y = np.sin(X[:, 0] * 3) + X[:, 1] * X[:, 2] + rng.normal(size=20000) * 0.1
    m = sk.HistGradientBoostingRegressor(max_iter=4, max_leaf_nodes=512, min_samples_leaf=2, max_depth=None,
                                         random_state=0).fit(X, y)
    trees = _O().trees_of(m)
    sizes = [len(t) for t in trees]
    assert max(sizes) > 900                                  # close to the 1,024-node chunk
    nodes, roots, cn0, ct0, depth = pack_trees(trees)
    assert np.all(np.diff(cn0) <= 1024)
    Xt = _test_rows(m, 400, 6)
    assert np.array_equal(walk_packed(nodes, roots, cn0, ct0, depth, _O().baseline_of(m), Xt), m.predict(Xt))
    with pytest.raises(ValueError, match="nodes"):            # a tree of 1,030 nodes cannot be staged
        pack_trees([np.resize(trees[0], 1030)])
