"""Learned candidate prefilter: which of a spectrum's index top-K candidates reach exact rescoring.

No reference counterpart. The reference shortlists by K1 shared-peak count alone (select_topk,
vulcan/gpu/engine.py:284-312) and rescores the whole shortlist with K2 (engine.py:315-351); its learned models
(vulcan/rescore/fast.py:198-254, static.py:25-82) act after K2. Here an MLP 8 -> 256 -> 256 -> 1 scores the
index top-K (default 200) from index-hit features and only its best `keep` (default 16), pair-closed as the
reference closes its shortlist, are rescored. The index and the exact scores are never approximated: K1 and K2
are the plain kernels, and the plain path (DeviceSearch at the same index top-K) stays the judge of recall.

Device pipeline (LearnedSearch.run, one stream, no host synchronisation):
  vh_k0 -> vh_hypotheses -> vh_spectrum_order -> vh_k1_features (top-K lists + count + matched-intensity
  pseudo-log) -> vh_prefilter_score (features in registers, BF16 MFMA MLP) -> vh_prefilter_select (top keep +
  closure) -> vh_k2 -> vh_rank.

Training (train): on the GPU with the project's kernels -- vh_mlp_forward (activations stored),
vh_mlp_backward (BF16 MFMA products, FP32 accumulation), vh_mlp_adam -- on labels from the exact rescorer's own
output: the target of a candidate is its K2 hyperscore minus the best hyperscore of its spectrum's top-K
(clipped at -clip), so the network learns the rescorer's order within a spectrum. Features and labels never
look at target/decoy labels.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from ..gpu import _lib as K
from ..gpu._abi import VhParams
from ..gpu.engine import DeviceSearch, GpuIndex, _k2_call, _k2_out, _marker
from ..gpu.host import SearchParams

N_PARAMS = 68353
BLOB_BYTES = 281728
FEATURES = ("k1_count", "count_over_best", "b_ion_count", "log2_b_intensity", "log2_y_intensity", "precursor_ppm",
            "length", "charge")
_OFF = {"W1": (0, (256, 8)), "b1": (2048, (256,)), "W2": (2304, (256, 256)), "b2": (67840, (256,)),
        "w3": (68096, (256,)), "b3": (68352, (1,))}


@dataclass
class PrefilterParams:
    index_topk: int = 200      # K1 shortlist scored by the network
    keep: int = 64             # at most this many candidates kept per spectrum before pair closure
    keep_min: int = 8          # always the best keep_min by score ...
    margin: float = 3.0        # ... then every candidate scoring within margin of the best (up to keep)
    features: str = "counts"   # "counts": count features only (b-ion count and intensity fields zero);
                               # "full": + per-candidate b-ion count and per-ion intensity pseudo-logs
    pair_closure: bool = True
    steps: int = 3000
    batch: int = 8192
    lr: float = 2e-3
    lr_final: float = 1e-4     # cosine decay to this
    beta1: float = 0.9
    beta2: float = 0.999
    eps: float = 1e-8
    clip: float = 30.0         # targets below best - clip are clipped
    seed: int = 0


@dataclass
class PrefilterModel:
    theta: np.ndarray                      # f32 [N_PARAMS]
    mu: np.ndarray                         # f32 [8] feature mean
    isd: np.ndarray                        # f32 [8] 1 / feature std
    history: list = field(default_factory=list)

    @classmethod
    def init(cls, seed: int = 0, mu=None, isd=None) -> "PrefilterModel":
        """He-normal weights, zero biases."""
        rng = np.random.default_rng(seed)
        th = np.zeros(N_PARAMS, np.float32)
        o, sh = _OFF["W1"]
        th[o:o + 2048] = rng.standard_normal(2048) * np.sqrt(2.0 / 8)
        o, sh = _OFF["W2"]
        th[o:o + 65536] = rng.standard_normal(65536) * np.sqrt(2.0 / 256)
        o, sh = _OFF["w3"]
        th[o:o + 256] = rng.standard_normal(256) * np.sqrt(1.0 / 256)
        mu = np.zeros(8, np.float32) if mu is None else np.asarray(mu, np.float32)
        isd = np.ones(8, np.float32) if isd is None else np.asarray(isd, np.float32)
        return cls(th, mu, isd)

    def save(self, path: str):
        np.savez(path, theta=self.theta, mu=self.mu, isd=self.isd)

    @classmethod
    def load(cls, path: str) -> "PrefilterModel":
        z = np.load(path, allow_pickle=False)
        return cls(z["theta"].astype(np.float32), z["mu"].astype(np.float32), z["isd"].astype(np.float32))

    def blob(self, dev) -> torch.Tensor:
        """Packed bf16 model on the device (vh_mlp_pack)."""
        th = torch.from_numpy(self.theta).to(dev)
        return pack(th, torch.from_numpy(self.mu).to(dev), torch.from_numpy(self.isd).to(dev))


def pack(theta: torch.Tensor, mu: torch.Tensor, isd: torch.Tensor, out: torch.Tensor | None = None):
    b = torch.empty(BLOB_BYTES, dtype=torch.uint8, device=theta.device) if out is None else out
    K.check(K.lib().vh_mlp_pack(K.ptr(theta), K.ptr(mu), K.ptr(isd), K.ptr(b), K.stream()), "mlp_pack")
    return b


def forward(blob: torch.Tensor, x: torch.Tensor, store: bool = False):
    """MLP scores of a feature matrix [n, 8] (f32, raw features); store: also the backward activations."""
    n = x.shape[0]
    dev = x.device
    s = torch.empty(max(n, 1), dtype=torch.float32, device=dev)
    a1 = torch.empty((max(n, 1), 256), dtype=torch.int16, device=dev) if store else None
    a2 = torch.empty((max(n, 1), 256), dtype=torch.float32, device=dev) if store else None
    K.check(K.lib().vh_mlp_forward(K.ptr(blob), n, K.ptr(x), K.ptr(s), K.ptr(a1), K.ptr(a2), K.stream()),
            "mlp_forward")
    return (s[:n], a1, a2) if store else s[:n]


def backward(blob, x, a1, a2, s, y, ws: torch.Tensor | None = None):
    """Gradient of mean((s - y)^2) w.r.t. the parameter vector (f32 [N_PARAMS], device)."""
    n = x.shape[0]
    L = K.lib()
    need = int(L.vh_mlp_workspace_bytes(n))
    if ws is None or ws.numel() < need:
        ws = torch.empty(need, dtype=torch.uint8, device=x.device)
    g = torch.empty(N_PARAMS, dtype=torch.float32, device=x.device)
    K.check(L.vh_mlp_backward(K.ptr(blob), n, K.ptr(x), K.ptr(a1), K.ptr(a2), K.ptr(s), K.ptr(y), K.ptr(g),
                              K.ptr(ws), ws.numel(), K.stream()), "mlp_backward")
    return g


def adam_step(theta, g, m1, m2, t: int, lr: float, beta1: float, beta2: float, eps: float):
    c1 = float(np.float32(1.0 / (1.0 - beta1 ** t)))
    c2 = float(np.float32(1.0 / (1.0 - beta2 ** t)))
    K.check(K.lib().vh_mlp_adam(K.ptr(theta), K.ptr(g), K.ptr(m1), K.ptr(m2), theta.numel(), lr, beta1, beta2, eps,
                                c1, c2, K.stream()), "mlp_adam")


def train(model: PrefilterModel, x: torch.Tensor, y: torch.Tensor, pp: PrefilterParams,
          log_every: int = 0) -> PrefilterModel:
    """Adam on mean((s - y)^2) over random minibatches of (x, y) (device tensors), all on the GPU. Sets the
    feature normalisation from x first. Returns a new model; the loss history is kept in model.history."""
    dev = x.device
    xf = x.float()
    mu = xf.mean(0)
    sd = xf.std(0)
    isd = torch.where(sd > 1e-6, 1.0 / sd, torch.ones_like(sd)).float()
    mu, isd = mu.float().contiguous(), isd.contiguous()
    th = torch.from_numpy(model.theta.copy()).to(dev)
    m1 = torch.zeros_like(th)
    m2 = torch.zeros_like(th)
    blob = torch.empty(BLOB_BYTES, dtype=torch.uint8, device=dev)
    ws = torch.empty(int(K.lib().vh_mlp_workspace_bytes(pp.batch)), dtype=torch.uint8, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(pp.seed)
    n = x.shape[0]
    hist = []
    t0 = time.perf_counter()
    for t in range(1, pp.steps + 1):
        idx = torch.randint(0, n, (pp.batch,), generator=gen, device=dev)
        xb = x[idx].contiguous()
        yb = y[idx].contiguous()
        pack(th, mu, isd, blob)
        s, a1, a2 = forward(blob, xb, store=True)
        g = backward(blob, xb, a1, a2, s, yb, ws)
        lr = pp.lr_final + 0.5 * (pp.lr - pp.lr_final) * (1 + np.cos(np.pi * (t - 1) / max(pp.steps - 1, 1)))
        adam_step(th, g, m1, m2, t, float(lr), pp.beta1, pp.beta2, pp.eps)
        if log_every and (t % log_every == 0 or t == pp.steps):
            hist.append((t, float(((s - yb) ** 2).mean().item())))
    torch.cuda.synchronize()
    out = PrefilterModel(th.cpu().numpy(), mu.cpu().numpy(), isd.cpu().numpy(), hist)
    out.history.append(("train_s", time.perf_counter() - t0))
    return out


def _feature_params(vp: VhParams, topk: int) -> VhParams:
    q = VhParams()
    C.memmove(C.byref(q), C.byref(vp), C.sizeof(VhParams))
    q.topk = topk
    q.pair_closure = 0
    return q


class FeatureLists:
    """Device buffers of vh_k1_features' per-spectrum lists (stride = index top-K). source "full": b-ion count and
    per-ion intensity pseudo-logs per candidate; "counts": those fields are zero (the network sees count, count /
    best count, precursor ppm, length and charge)."""

    def __init__(self, n_spec: int, topk: int, dev, source: str = "counts"):
        if source not in ("counts", "full"):
            raise ValueError(source)
        self.source = source
        n = max(n_spec, 1)
        self.stride = topk
        e = lambda m, dt: torch.empty(m, dtype=dt, device=dev)  # noqa: E731
        self.n_cand = e(n, torch.int32)
        self.row, self.hyp, self.cnt = e(n * topk, torch.int32), e(n * topk, torch.int32), e(n * topk, torch.int32)
        self.nb = e(n * topk, torch.int16)
        self.iq = e(n * topk, torch.int32)
        self.maxcnt = e(n, torch.int32)
        self.score = e(n * topk, torch.float32)
        self.score_ws = e(int(K.lib().vh_prefilter_score_workspace_bytes(n, topk)), torch.uint8)

    def k1(self, ds: DeviceSearch, vpf: VhParams, n: int):
        pk, hy = ds.peaks, ds.hyp
        if ds.k1_impl == "join":
            # block-major counts, index top-K without closure, then the features of the selected candidates
            ds.join_front(n, _marker(None))
            j = ds.join
            L = K.lib()
            K.check(L.vh_k1j_topk(C.byref(vpf), C.byref(ds.gi.view), n, K.ptr(ds.order), K.ptr(hy["hyp_off"]),
                                  K.ptr(hy["mode"]), K.ptr(hy["lo"]), K.ptr(hy["hi"]), K.ptr(j["h_slot"]),
                                  j["slot_cap"], K.ptr(j["counts"]), self.stride, K.ptr(self.n_cand),
                                  K.ptr(self.row), K.ptr(self.hyp), K.ptr(self.cnt), K.ptr(ds.scored),
                                  K.ptr(ds.matched), K.ptr(ds.status), K.ptr(ds.queue), K.stream()), "k1j_topk")
            if self.source == "counts":
                self._count_only(n)
                return
            K.check(L.vh_k1_candidate_features(C.byref(vpf), C.byref(ds.gi.view), n, K.ptr(ds.order),
                                               K.ptr(pk["off"]), K.ptr(pk["mass"]), K.ptr(pk["int"]),
                                               K.ptr(hy["hyp_off"]), K.ptr(hy["maxfz"]), K.ptr(hy["lo"]),
                                               K.ptr(hy["hi"]), self.stride, K.ptr(self.n_cand), K.ptr(self.row),
                                               K.ptr(self.cnt), K.ptr(self.nb), K.ptr(self.iq), K.ptr(self.maxcnt),
                                               K.ptr(ds.status), K.stream()), "k1_candidate_features")
            return
        K.check(K.lib().vh_k1_features(C.byref(vpf), C.byref(ds.gi.tile_view()), n, K.ptr(ds.order), K.ptr(pk["off"]),
                                       K.ptr(pk["mass"]), K.ptr(pk["int"]), K.ptr(hy["hyp_off"]),
                                       K.ptr(hy["maxfz"]), K.ptr(hy["mode"]), K.ptr(hy["lo"]), K.ptr(hy["hi"]),
                                       self.stride, K.ptr(self.n_cand), K.ptr(self.row), K.ptr(self.hyp),
                                       K.ptr(self.cnt), K.ptr(self.nb), K.ptr(self.iq), K.ptr(self.maxcnt),
                                       K.ptr(ds.scored),
                                       K.ptr(ds.matched), K.ptr(ds.status), K.ptr(ds.queue), K.stream()),
                "k1_features")
        if self.source == "counts":
            self._count_only(n)

    def _count_only(self, n: int):
        """Count-only features: zero b-ion counts / intensity fields, best count per spectrum
        (vh_prefilter_count_features)."""
        K.check(K.lib().vh_prefilter_count_features(n, self.stride, K.ptr(self.n_cand), K.ptr(self.cnt),
                                                    K.ptr(self.nb), K.ptr(self.iq), K.ptr(self.maxcnt),
                                                    K.stream()), "prefilter_count_features")

    def features(self, ds: DeviceSearch, n: int) -> torch.Tensor:
        out = torch.empty((max(n, 1) * self.stride, 8), dtype=torch.float32, device=ds.gi.dev)
        hy = ds.hyp
        K.check(K.lib().vh_prefilter_features(C.byref(ds.gi.view), n, self.stride, K.ptr(self.n_cand),
                                              K.ptr(self.row), K.ptr(self.hyp), K.ptr(self.cnt), K.ptr(self.nb),
                                              K.ptr(self.iq), K.ptr(self.maxcnt), K.ptr(hy["z"]), K.ptr(hy["M"]),
                                              K.ptr(out), K.stream()), "prefilter_features")
        return out[: n * self.stride]

    def score_with(self, ds: DeviceSearch, blob: torch.Tensor, n: int):
        hy = ds.hyp
        K.check(K.lib().vh_prefilter_score(C.byref(ds.gi.view), K.ptr(blob), n, self.stride, K.ptr(self.n_cand),
                                           K.ptr(self.row), K.ptr(self.hyp), K.ptr(self.cnt), K.ptr(self.nb),
                                           K.ptr(self.iq), K.ptr(self.maxcnt), K.ptr(hy["z"]), K.ptr(hy["M"]),
                                           K.ptr(self.score), K.ptr(self.score_ws), self.score_ws.numel(),
                                           K.stream()), "prefilter_score")


class LearnedSearch(DeviceSearch):
    """DeviceSearch whose candidate stage is index top-K -> MLP -> top `keep` + closure (module docstring)."""

    def __init__(self, gi: GpuIndex, p: SearchParams, n_spec: int, model: PrefilterModel,
                 pp: PrefilterParams | None = None, report: int = 1, persistent: bool = False):
        self.pp = pp or PrefilterParams()
        q = SearchParams(**p.to_dict())
        q.topk = self.pp.keep
        q.pair_closure = self.pp.pair_closure
        super().__init__(gi, q, n_spec, report=report, persistent=persistent)
        self.vpf = _feature_params(self.vp, self.pp.index_topk)
        self.lists = FeatureLists(n_spec, self.pp.index_topk, gi.dev, self.pp.features)
        self.model = model
        self.blob = model.blob(gi.dev)

    def candidate_stage(self, n: int, mark):
        L = K.lib()
        fl = self.lists
        mark("k1_features")
        fl.k1(self, self.vpf, n)
        mark("k1_features")
        mark("mlp")
        fl.score_with(self, self.blob, n)
        mark("mlp")
        mark("select")
        K.check(L.vh_prefilter_select(C.byref(self.gi.view), n, fl.stride, K.ptr(fl.n_cand), K.ptr(fl.row),
                                      K.ptr(fl.hyp), K.ptr(fl.cnt), K.ptr(fl.score),
                                      min(self.pp.keep_min, self.pp.keep), self.pp.keep, float(self.pp.margin),
                                      int(self.pp.pair_closure), self.stride, K.ptr(self.n_cand), K.ptr(self.c_row),
                                      K.ptr(self.c_hyp), K.ptr(self.c_cnt), K.stream()), "prefilter_select")
        mark("select")


def training_data(gi: GpuIndex, p: SearchParams, raw: dict, pp: PrefilterParams):
    """Features and rescorer labels of every index top-K candidate of the spectra in `raw` (device tensors):
    returns (x [m, 8] f32, y [m] f32, spectrum [m] i64, hyperscore [m] f64) over the non-empty slots."""
    n = raw["n"]
    ds = DeviceSearch(gi, p, n)
    mark = _marker(None)
    ds.front(raw, n, mark)
    vpf = _feature_params(ds.vp, pp.index_topk)
    fl = FeatureLists(n, pp.index_topk, gi.dev, pp.features)
    fl.k1(ds, vpf, n)
    x = fl.features(ds, n)
    out = _k2_out(n * fl.stride, gi.dev)
    _k2_call(gi, ds.vp, n, None, None, fl.n_cand, fl.stride, fl.row, fl.hyp, ds.hyp["maxfz"], ds.peaks, out,
             order=ds.order)
    hs = out["hyperscore"][: n * fl.stride].view(n, fl.stride)
    valid = torch.arange(fl.stride, device=gi.dev)[None, :] < fl.n_cand[:n, None].long()
    best = torch.where(valid, hs, torch.full_like(hs, -1e30)).max(1, keepdim=True).values
    y = torch.clamp(hs - best, min=-pp.clip).float()
    m = valid.reshape(-1)
    spec = torch.arange(n, device=gi.dev)[:, None].expand(n, fl.stride).reshape(-1)
    return x[m].contiguous(), y.reshape(-1)[m].contiguous(), spec[m], hs.reshape(-1)[m]


def agreement(plain_rows: np.ndarray, learned_rows: np.ndarray) -> float:
    """Top-1 agreement: over spectra with a plain rank-1 PSM, the fraction whose rank-1 row is the same."""
    m = plain_rows >= 0
    return float(np.mean(learned_rows[m] == plain_rows[m])) if m.any() else 1.0
