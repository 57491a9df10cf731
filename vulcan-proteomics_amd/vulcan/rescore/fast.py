"""Production scoring on the device: the gradient-boosted tree scorer (reference vulcan/rescore/fast.py:198-254,
`GpuHGB`), evaluated by one HIP kernel (vh_gbm_decision) instead of per-level torch gathers over [n, trees].

GpuHGB(est).decision_function(X) equals est.decision_function(X) (classifier) / the raw prediction (regressor)
bit for bit: the same comparisons (x <= threshold, NaN by missing_go_to_left) and the same f64 summation order
(0 + baseline, then tree by tree).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ..gpu import _lib as K
from ..gpu._abi import GBM_CHUNK_NODES, GBM_MAX_FEATURES, GBM_NODE


def pack_trees(trees: list):
    """sklearn predictor node arrays -> (VhGbmNode records, tree roots, chunk node / tree starts, max depth).
    Each tree is renumbered breadth-first so an internal node's children are adjacent (left, left + 1); trees
    are kept whole and cut into chunks of at most GBM_CHUNK_NODES nodes; `left` is an index inside the chunk
    (a leaf's is its own, with feature -1)."""
    recs, roots = [], []
    chunk_node0, chunk_tree0 = [0], [0]
    base = chunk_base = 0
    max_depth = 0
    for t, nodes in enumerate(trees):
        if (nodes["is_categorical"] != 0).any():
            raise ValueError("categorical splits are not supported")
        m = len(nodes)
        if m > GBM_CHUNK_NODES:
            raise ValueError(f"tree {t} has {m} nodes (at most {GBM_CHUNK_NODES})")
        if base + m - chunk_base > GBM_CHUNK_NODES:       # start a new chunk at this tree
            chunk_node0.append(base)
            chunk_tree0.append(t)
            chunk_base = base
        order, pos = [0], {0: 0}
        i = 0
        while i < len(order):                            # breadth-first, children appended in pairs
            k = order[i]
            if not nodes["is_leaf"][k]:
                for ch in (int(nodes["left"][k]), int(nodes["right"][k])):
                    if ch in pos or not 0 <= ch < m:
                        raise ValueError(f"tree {t}: node {k} has an invalid child {ch}")
                    pos[ch] = len(order)
                    order.append(ch)
            i += 1
        r = np.zeros(len(order), GBM_NODE)
        src = nodes[np.array(order)]
        leaf = src["is_leaf"] != 0
        r["v"] = np.where(leaf, src["value"], src["num_threshold"])
        loc = base - chunk_base                          # the tree's first node inside its chunk
        r["left"] = [loc + (i if lf else pos[int(c)]) for i, (c, lf) in enumerate(zip(src["left"], leaf))]
        r["feature"] = np.where(leaf, -1, src["feature_idx"])
        r["missing_left"] = src["missing_go_to_left"]
        r["is_leaf"] = leaf
        recs.append(r)
        roots.append(base)
        base += len(order)
        max_depth = max(max_depth, int(src["depth"].max()))
    chunk_node0.append(base)
    chunk_tree0.append(len(trees))
    nodes_all = np.concatenate(recs) if recs else np.zeros(0, GBM_NODE)
    return nodes_all, np.array(roots, np.int32), np.array(chunk_node0, np.int32), np.array(chunk_tree0, np.int32), \
        max_depth


class GpuHGB:
    """sklearn HistGradientBoosting{Classifier,Regressor} decision_function on the GPU, bit-identical (float64)."""

    def __init__(self, est, dev="cuda"):
        dev = torch.device(dev)
        preds = est._predictors
        if any(len(it) != 1 for it in preds):
            raise ValueError("only one tree per iteration (binary / regression) is supported")
        trees = [it[0].nodes for it in preds]
        nodes, roots, cn0, ct0, depth = pack_trees(trees)
        if int(est.n_features_in_) > GBM_MAX_FEATURES:
            raise ValueError(f"{est.n_features_in_} features (at most {GBM_MAX_FEATURES})")
        if len(nodes) and int(nodes["feature"][nodes["is_leaf"] == 0].max(initial=0)) >= int(est.n_features_in_):
            raise ValueError("a split uses a feature index beyond n_features_in_")
        up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
        self.nodes = up(nodes.view(np.uint8)) if len(nodes) else torch.zeros(16, dtype=torch.uint8, device=dev)
        self.roots, self.chunk_node0, self.chunk_tree0 = up(roots), up(cn0), up(ct0)
        self.T, self.M = len(trees), len(nodes)
        self.n_chunks = len(cn0) - 1
        self.depth = depth
        self.baseline = float(np.asarray(est._baseline_prediction).ravel()[0])
        self.n_features = int(est.n_features_in_)
        self.dev = dev
        self.status = torch.zeros(1, dtype=torch.int32, device=dev)

    def decision_device(self, X: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """X: device f64 [n, n_features] (contiguous) -> device f64 [n], on the current stream."""
        if X.dtype != torch.float64 or X.dim() != 2 or X.shape[1] != self.n_features or not X.is_contiguous():
            raise ValueError("feature count mismatch" if X.dim() == 2 and X.shape[1] != self.n_features else
                             "X must be a contiguous f64 [n, n_features] device tensor")
        n = X.shape[0]
        out = torch.empty(n, dtype=torch.float64, device=self.dev) if out is None else out
        self.status.zero_()
        K.check(K.lib().vh_gbm_decision(K.ptr(self.nodes), self.M, K.ptr(self.roots), self.T,
                                        K.ptr(self.chunk_node0), K.ptr(self.chunk_tree0), self.n_chunks,
                                        self.depth, C.c_double(self.baseline), K.ptr(X), n, self.n_features,
                                        K.ptr(out), K.ptr(self.status), K.stream()), "gbm_decision")
        return out

    def check(self) -> None:
        st = int(self.status.item())
        if st & 2:
            raise RuntimeError("gbm_decision: malformed chunk table (a chunk beyond the node array, larger than "
                               f"{GBM_CHUNK_NODES} nodes, or trees outside the model)")
        if st:
            raise RuntimeError("gbm_decision: malformed tree (a walk left its chunk or found no leaf)")

    def decision_function(self, X, chunk=100_000) -> np.ndarray:
        """Host rows -> host f64 scores; `chunk` bounds the rows resident on the device at once."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        if X.ndim != 2 or X.shape[1] != self.n_features:
            raise ValueError("feature count mismatch")
        out = np.empty(len(X), np.float64)
        step = max(int(chunk), 1)
        for b in range(0, len(X), step):
            xd = torch.from_numpy(X[b:b + step]).to(self.dev)
            out[b:b + len(xd)] = self.decision_device(xd).cpu().numpy()
            self.check()
        return out
