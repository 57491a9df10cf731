"""CPU ORACLE of the learned prefilter -- test infrastructure, not product code.

Only tests/ (and __graft_entry__.smoke()) import this module, as the checker of the HIP prefilter kernels
(vulcan-proteomics_amd/csrc/vh_mlp.hip). The prefilter has no counterpart in the reference; what is pinned:
  features  f32 restatement of vh_prefilter_features (bit-exact target);
  forward / grad (exact=True)   the MLP 8 -> 256 -> 256 -> 1 and the gradient of L = mean((s - y)^2) in float64
            with no rounding -- the mathematical definition; its analytic gradient is pinned by central finite
            differences in the tests;
  forward / grad (exact=False)  the same with the device's rounding points emulated (bf16 operands rounded to
            nearest even, f32 bias adds and ReLU), sums in float64: the GPU must agree within f32 accumulation
            error, which a wrong layout, index or mask would exceed by orders of magnitude;
  adam      f32 restatement of vh_mlp_adam (bit-exact target);
  select    vh_prefilter_select: the best keep_min by (score desc, NaN last, earlier slot first), then every slot
            within margin of the best score (at most keep), then the reference's pair closure / dedupe
            (vulcan/gpu/engine.py:284-312) (bit-exact target).
"""
from __future__ import annotations

import numpy as np

N_PARAMS = 68353
OFF = {"W1": (0, (256, 8)), "b1": (2048, (256,)), "W2": (2304, (256, 256)), "b2": (67840, (256,)),
       "w3": (68096, (256,)), "b3": (68352, (1,))}


def bf16(x) -> np.ndarray:
    """Round f32 values to bf16 (nearest even), returned as f32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFFFFFF
    return ((u >> 16) << 16).astype(np.uint32).view(np.float32)


def unpack(theta):
    t = np.asarray(theta)
    return {k: t[o:o + int(np.prod(sh))].reshape(sh) for k, (o, sh) in OFF.items()}


def features(mass32, length, n_cand, row, hyp, cnt, nb, iq, maxcnt, h_z, h_M, stride: int):
    """f32 restatement of vh_prefilter_features: [n_spec * stride, 8], zeros in empty slots."""
    n_spec = len(n_cand)
    slots = n_spec * stride
    out = np.zeros((slots, 8), np.float32)
    pos = np.arange(slots) % stride
    spec = np.arange(slots) // stride
    ok = pos < np.asarray(n_cand)[spec]
    r = np.asarray(row)[:slots][ok]
    h = np.asarray(hyp)[:slots][ok]
    c = np.asarray(cnt)[:slots][ok].astype(np.float32)
    mx = np.asarray(maxcnt)[spec[ok]].astype(np.float32)
    m32 = np.asarray(mass32, np.float32)[r]
    M = np.asarray(h_M, np.float32)[h]
    f = np.zeros((len(r), 8), np.float32)
    f[:, 0] = c
    with np.errstate(divide="ignore", invalid="ignore"):
        f[:, 1] = np.where(mx > 0, c / np.where(mx > 0, mx, np.float32(1)), np.float32(0))
    q = np.asarray(iq).view(np.uint32)[:slots][ok]
    f[:, 2] = np.asarray(nb)[:slots][ok].astype(np.float32)
    f[:, 3] = (q >> 16).astype(np.float32) * np.float32(0.00390625)
    f[:, 4] = (q & 0xFFFF).astype(np.float32) * np.float32(0.00390625)
    f[:, 5] = ((M - m32) / m32) * np.float32(1000000.0)
    f[:, 6] = np.asarray(length)[r].astype(np.float32)
    f[:, 7] = np.asarray(h_z)[h].astype(np.float32)
    out[ok] = f
    return out


def normalise(X, mu, isd, exact: bool):
    if exact:
        return (np.asarray(X, np.float64) - np.asarray(mu, np.float64)) * np.asarray(isd, np.float64)
    xn = (np.asarray(X, np.float32) - np.asarray(mu, np.float32)) * np.asarray(isd, np.float32)
    return bf16(xn).astype(np.float64)


def forward(theta, mu, isd, X, exact: bool = True):
    P = unpack(theta)
    xb = normalise(X, mu, isd, exact)
    if exact:
        z1 = xb @ P["W1"].astype(np.float64).T + P["b1"]
        a1 = np.maximum(z1, 0.0)
        z2 = a1 @ P["W2"].astype(np.float64).T + P["b2"]
        a2 = np.maximum(z2, 0.0)
    else:
        W1b, W2b = bf16(P["W1"]).astype(np.float64), bf16(P["W2"]).astype(np.float64)
        z1 = ((xb @ W1b.T).astype(np.float32) + P["b1"].astype(np.float32)).astype(np.float64)
        a1 = bf16(np.maximum(z1, 0.0)).astype(np.float64)
        z2 = ((a1 @ W2b.T).astype(np.float32) + P["b2"].astype(np.float32)).astype(np.float64)
        a2 = np.maximum(z2, 0.0)
    s = a2 @ P["w3"].astype(np.float64) + float(P["b3"][0])
    return {"s": s, "xb": xb, "a1": a1, "a2": a2, "z1": z1, "z2": z2}


def loss(theta, mu, isd, X, y, exact: bool = True) -> float:
    s = forward(theta, mu, isd, X, exact)["s"]
    return float(np.mean((s - np.asarray(y, np.float64)) ** 2))


def grad(theta, mu, isd, X, y, exact: bool = True):
    """Gradient of mean((s - y)^2) w.r.t. theta (float64 vector of N_PARAMS)."""
    P = unpack(theta)
    F = forward(theta, mu, isd, X, exact)
    n = len(X)
    y = np.asarray(y, np.float64)
    g = np.zeros(N_PARAMS, np.float64)
    G = unpack(g)
    if exact:
        ds = 2.0 * (F["s"] - y) / n
        d2 = np.where(F["a2"] > 0, ds[:, None] * P["w3"][None, :], 0.0)
        dz2 = d2
        W2 = P["W2"].astype(np.float64)
        da1 = dz2 @ W2
        dz1 = np.where(F["a1"] > 0, da1, 0.0)
    else:
        s32 = F["s"].astype(np.float32)
        inv_n = np.float32(1.0) / np.float32(n)
        ds = ((np.float32(2.0) * (s32 - np.asarray(y, np.float32))) * inv_n).astype(np.float64)
        d2 = np.where(F["a2"] > 0, (ds.astype(np.float32)[:, None] * P["w3"].astype(np.float32)[None, :]), 0.0)
        dz2 = bf16(d2).astype(np.float64)
        W2 = bf16(P["W2"]).astype(np.float64)
        da1 = (dz2 @ W2).astype(np.float32).astype(np.float64)
        dz1 = np.where(F["a1"] > 0, da1, 0.0)
    G["w3"][:] = F["a2"].T @ ds
    G["b3"][:] = ds.sum()
    G["b2"][:] = np.asarray(d2, np.float64).sum(0)
    G["W2"][:] = dz2.T @ F["a1"]
    G["W1"][:] = dz1.T @ F["xb"]
    G["b1"][:] = dz1.sum(0)
    return g


def adam(theta, g, m1, m2, lr, b1, b2, eps, c1, c2):
    """f32 restatement of vh_mlp_adam; returns (theta, m1, m2) (new arrays)."""
    f = np.float32
    th, g, m1, m2 = (np.asarray(a, np.float32) for a in (theta, g, m1, m2))
    lr, b1, b2, eps, c1, c2 = f(lr), f(b1), f(b2), f(eps), f(c1), f(c2)
    m = b1 * m1 + (f(1) - b1) * g
    v = b2 * m2 + (f(1) - b2) * (g * g)
    mh, vh = m * c1, v * c2
    return th - lr * (mh / (np.sqrt(vh) + eps)), m, v


def score_ord(x: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    o = np.where(u & 0x80000000, (~u) & 0xFFFFFFFF, u | 0x80000000)
    return np.where(np.isnan(x), 0, o).astype(np.uint64)


def select(n_cand, row, hyp, cnt, score, stride_in: int, keep: int, closure: bool, partner, out_stride: int,
           keep_min: int | None = None, margin: float = 0.0):
    """keep_min (default keep): the best keep_min by score, then every slot scoring >= best - margin (f32), at most
    keep."""
    keep_min = keep if keep_min is None else keep_min
    n_spec = len(n_cand)
    on = np.zeros(n_spec, np.int32)
    orow = np.zeros(n_spec * out_stride, np.int32)
    ohyp = np.zeros_like(orow)
    ocnt = np.zeros_like(orow)
    for s in range(n_spec):
        n = min(int(n_cand[s]), stride_in)
        b = s * stride_in
        sc = np.asarray(score[b:b + n], np.float32)
        key = score_ord(sc)
        order = np.lexsort((np.arange(n), -key.astype(np.float64)))
        lim = min(keep, n)
        if n:
            thr = np.float32(sc[order[0]] - np.float32(margin))
            c = int(np.sum(sc[order[:lim]] >= thr))
        else:
            c = 0
        idx = order[:max(c, min(keep_min, n))]
        ent = [(int(row[b + i]), int(hyp[b + i]), int(cnt[b + i])) for i in idx]
        if closure:
            ent += [(int(partner[r]), h, -1) for r, h, _ in ent[:len(idx)]]
        by = {}
        for r, h, c in ent:
            if r in by:
                h0, c0 = by[r]
                by[r] = (min(h0, h), max(c0, c))
            else:
                by[r] = (h, c)
        rs = sorted(by)
        on[s] = len(rs)
        for j, r in enumerate(rs):
            orow[s * out_stride + j] = r
            ohyp[s * out_stride + j], ocnt[s * out_stride + j] = by[r]
    return {"n": on, "row": orow, "hyp": ohyp, "cnt": ocnt}
