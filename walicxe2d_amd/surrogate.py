"""Cooling surrogate for coolingh (src/coolingh.f90:171-259): data harvest
from plain runs, training on the GPU with the library's own HIP kernels
(wx_sur_train: bf16 matrix-core forward/backward, Adam), evaluation, and the
model file.

    python -m walicxe2d_amd.surrogate train [--out walicxe2d_amd/data/surrogate_jet.npz]

The network (4 -> 64 -> 64 -> 64 -> 2, SiLU) maps the features
    log10 n, log10 T, x_HII = 1 - n_HI/n, log10(dt / t_cool)
to (c, dx): c = log10(T1 / T1_table) corrects the routine's relaxation law
T1 = 10 + (T - 10) exp(-dt / t_cool) evaluated with a device table of the
loss function, and dx = x_HII(new) - x_HII(old).  Training pairs come from
plain coolingh runs of seeded jet variants (velocity 150-250 km/s, density
30-300 cm^-3; the jet workload with seed 777 + k): cell states (n, T, y0,
dt) of leaf interiors, sampled every few hundred steps, and their exact
update by the device's coolingh (wx_sur_pairs).  The last variants are held
out for evaluation.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

from . import params as P
from .capi import NP, Simulation, WxSurModel, _check, load_library

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "data", "surrogate_jet.npz")
NPARAM = 8770


class SurrogateModel:
    """Parameters, standardisation, training hull and residual tolerance."""

    def __init__(self, theta=None, seed: int = 0):
        self.theta = np.asarray(theta, dtype=np.float32) if theta is not None else init_theta(seed)
        assert self.theta.shape == (NPARAM,)
        self.in_mu = np.zeros(4, np.float32)
        self.in_sig = np.ones(4, np.float32)
        self.out_mu = np.zeros(2, np.float32)
        self.out_sig = np.ones(2, np.float32)
        self.lo = np.full(4, -np.inf, np.float32)
        self.hi = np.full(4, np.inf, np.float32)
        self.tol = np.float32(1e-3)
        self.meta = {}

    def to_c(self) -> WxSurModel:
        m = WxSurModel()
        th = np.ascontiguousarray(self.theta, dtype=np.float32)   # kept alive across the copy
        ctypes.memmove(m.theta, th.ctypes.data, NPARAM * 4)
        for name in ("in_mu", "in_sig", "out_mu", "out_sig", "lo", "hi"):
            v = getattr(self, name).astype(np.float32)
            arr = getattr(m, name)
            for k in range(len(v)):
                arr[k] = float(v[k])
        m.tol = float(self.tol)
        return m

    def save(self, path: str):
        np.savez(path, theta=self.theta, in_mu=self.in_mu, in_sig=self.in_sig, out_mu=self.out_mu,
                 out_sig=self.out_sig, lo=self.lo, hi=self.hi, tol=np.float32(self.tol),
                 meta=np.frombuffer(json.dumps(self.meta).encode(), dtype=np.uint8))

    @classmethod
    def load(cls, path: str = MODEL_PATH) -> "SurrogateModel":
        d = np.load(path, allow_pickle=False)
        m = cls(d["theta"])
        for name in ("in_mu", "in_sig", "out_mu", "out_sig", "lo", "hi"):
            setattr(m, name, d[name].astype(np.float32))
        m.tol = np.float32(d["tol"])
        m.meta = json.loads(bytes(d["meta"]).decode()) if "meta" in d else {}
        return m


def init_theta(seed: int = 0) -> np.ndarray:
    """Uniform initialisation scaled by fan-in (biases zero)."""
    rng = np.random.default_rng(seed)
    parts = []
    for o, i in ((64, 4), (64, 64), (64, 64), (2, 64)):
        lim = np.sqrt(6.0 / i) * 0.5
        parts.append(rng.uniform(-lim, lim, size=o * i))
        parts.append(np.zeros(o))
    return np.concatenate(parts).astype(np.float32)


# ---------------------------------------------------------------------------
# device entry points
def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def exact_pairs(params, states: np.ndarray) -> np.ndarray:
    """(T1, y1) of the exact device coolingh for states (n, T, y0, dt[s])."""
    L = load_library()
    s = np.ascontiguousarray(states, dtype=np.float64)
    out = np.zeros((s.shape[0], 2))
    _check(L.wx_sur_pairs(ctypes.byref(params), _dp(s), s.shape[0], _dp(out)), "wx_sur_pairs")
    return out


def features(params, states: np.ndarray):
    """Raw features [n][4] and ce*dt [n] as the cooling kernel computes them."""
    L = load_library()
    s = np.ascontiguousarray(states, dtype=np.float64)
    f = np.zeros((s.shape[0], 4), np.float32)
    c = np.zeros(s.shape[0], np.float32)
    _check(L.wx_sur_features(ctypes.byref(params), _dp(s), s.shape[0], _fp(f), _fp(c)), "wx_sur_features")
    return f, c


def infer(model: SurrogateModel, x: np.ndarray) -> np.ndarray:
    L = load_library()
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros((x.shape[0], 2), np.float32)
    _check(L.wx_sur_infer(ctypes.byref(model.to_c()), _fp(x), x.shape[0], _fp(y)), "wx_sur_infer")
    return y


def grad(model: SurrogateModel, x, t, w=(1.0, 1.0)):
    L = load_library()
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    g = np.zeros(NPARAM, np.float32)
    loss = np.zeros(1, np.float32)
    _check(L.wx_sur_grad(ctypes.byref(model.to_c()), _fp(x), _fp(t), x.shape[0], w[0], w[1], _fp(loss), _fp(g)),
           "wx_sur_grad")
    return float(loss[0]), g


def train_steps(model: SurrogateModel, x, t, batch: int, steps: int, lr0: float, lr1: float, w=(1.0, 1.0)):
    """`steps` Adam steps on the GPU; updates model.theta; returns the loss per step."""
    L = load_library()
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    c = model.to_c()
    hist = np.zeros(max(steps, 1), np.float32)
    _check(L.wx_sur_train(ctypes.byref(c), _fp(x), _fp(t), x.shape[0], batch, steps, lr0, lr1, w[0], w[1],
                          _fp(hist)), "wx_sur_train")
    model.theta = np.frombuffer(bytes(c.theta), dtype=np.float32).copy()
    return hist[:steps]


# ---------------------------------------------------------------------------
# data
def relax_table(T, cedt):
    """The relaxation law T1 = 10 + (T - 10) exp(-ce dt) with the routine's
    clamps (src/coolingh.f90:249-254), as the cooling kernel evaluates it."""
    T = np.asarray(T, dtype=np.float64)
    tt = 10.0 + (T - 10.0) * np.exp(-np.asarray(cedt, dtype=np.float32).astype(np.float64))
    tt = np.maximum(tt, 0.1 * T)
    tt = np.minimum(tt, 10.0 * T)
    return np.maximum(tt, 10.0)


def harvest(cfg: P.Config, nsteps: int, every: int, per_snap: int, rng) -> np.ndarray:
    """Cell states (n, T, y0, dt[s]) of leaf interiors of a plain run,
    sampled every `every` steps (at most per_snap cells per snapshot)."""
    sc = cfg.sc
    sim = Simulation(cfg)
    sim.basegrid()
    for _ in range(4):
        sim.admesh()
    out = []
    done = 0
    while done < nsteps:
        k = min(every, nsteps - done)
        _, dt, _ = sim.step(k)
        done += k
        _, pr = sim.download(prim=True)
        leaf = sim.tree()[1]
        ids = np.nonzero(leaf)[0]
        q = pr[ids][:, :, 2:NP - 2, 2:NP - 2]
        rho = q[:, 0].ravel()
        pres = q[:, 3].ravel()
        nhi = q[:, 4].ravel()
        T = pres * sc["Tempsc"] / (2.0 * rho - nhi)
        st = np.stack([rho, T, nhi / rho, np.full(rho.shape, dt * sc["tsc"])], axis=1)
        if st.shape[0] > per_snap:
            st = st[rng.choice(st.shape[0], per_snap, replace=False)]
        out.append(st)
    sim.close()
    return np.concatenate(out)


def stratify(states: np.ndarray, cedt: np.ndarray, target: int, rng) -> np.ndarray:
    """Indices of about `target` states, capped per (log T, log ce dt) cell
    of 0.1 x 0.5 dex so that rare (hot, strongly cooling) states are kept."""
    lt = np.floor(np.log10(states[:, 1]) / 0.1).astype(np.int64)
    lc = np.floor(np.log10(np.maximum(cedt, 1e-30)) / 0.5).astype(np.int64)
    key = (lt - lt.min()) * 1000 + (lc - lc.min())
    uniq, inv, counts = np.unique(key, return_inverse=True, return_counts=True)
    # largest per-cell cap that keeps the total near the target
    lo, hi = 1, int(counts.max())
    while lo < hi:
        mid = (lo + hi + 1) // 2
        if np.minimum(counts, mid).sum() <= target:
            lo = mid
        else:
            hi = mid - 1
    cap = lo
    # a random order, then each state's rank within its cell
    order = rng.permutation(states.shape[0])
    b = inv[order]
    idx = np.argsort(b, kind="stable")
    sb = b[idx]
    rank = np.arange(sb.shape[0]) - np.searchsorted(sb, sb, side="left")
    return np.sort(order[idx[rank < cap]])


def dataset(params, states: np.ndarray):
    """Raw features, ce*dt and targets (c, dx) of states."""
    out = exact_pairs(params, states)
    f, cedt = features(params, states)
    T, y0 = states[:, 1], states[:, 2]
    T1, y1 = out[:, 0], out[:, 1]
    tt = relax_table(T, cedt)
    tgt = np.stack([np.log10(T1 / tt), y0 - y1], axis=1)
    return f, cedt, tgt


def errors(model: SurrogateModel, f, tgt):
    """RMS and max errors of the network's (c, dx) against the exact ones,
    beside those of the same method without the network (c = 0, dx = 0),
    over the states inside the hull."""
    inside = np.all((f >= model.lo) & (f <= model.hi), axis=1)
    x = (f[inside] - model.in_mu) / model.in_sig
    y = infer(model, x) * model.out_sig + model.out_mu
    e = y - tgt[inside]
    base = tgt[inside]
    return {"in_hull": float(inside.mean()), "n": int(inside.sum()),
            "rms_c": float(np.sqrt(np.mean(e[:, 0] ** 2))), "rms_dx": float(np.sqrt(np.mean(e[:, 1] ** 2))),
            "max_c": float(np.abs(e[:, 0]).max()), "max_dx": float(np.abs(e[:, 1]).max()),
            "rms_c_without": float(np.sqrt(np.mean(base[:, 0] ** 2))),
            "rms_dx_without": float(np.sqrt(np.mean(base[:, 1] ** 2)))}


def fit(f, tgt, steps: int, batch: int, lr0: float, lr1: float, seed: int = 0, tol: float = 1e-3):
    """Standardises, trains on the GPU and returns (model, loss history)."""
    model = SurrogateModel(seed=seed)
    model.in_mu = f.mean(axis=0).astype(np.float32)
    model.in_sig = np.maximum(f.std(axis=0), 1e-6).astype(np.float32)
    model.out_mu = tgt.mean(axis=0).astype(np.float32)
    model.out_sig = np.maximum(tgt.std(axis=0), 1e-12).astype(np.float32)
    model.lo = f.min(axis=0).astype(np.float32)
    model.hi = f.max(axis=0).astype(np.float32)
    model.tol = np.float32(tol)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(f.shape[0])
    x = ((f[perm] - model.in_mu) / model.in_sig).astype(np.float32)
    t = ((tgt[perm] - model.out_mu) / model.out_sig).astype(np.float32)
    hist = train_steps(model, x, t, batch, steps, lr0, lr1)
    return model, hist


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--variants", type=int, default=8, help="jet variants (seeds 777 + k)")
    ap.add_argument("--holdout", type=int, default=2, help="last variants kept out of training")
    ap.add_argument("--steps-run", type=int, default=12000, help="plain steps per variant (about 300 yr)")
    ap.add_argument("--every", type=int, default=400)
    ap.add_argument("--per-snap", type=int, default=120000)
    ap.add_argument("--pairs", type=int, default=10_000_000)
    ap.add_argument("--train-steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr0", type=float, default=2e-3)
    ap.add_argument("--lr1", type=float, default=2e-5)
    ap.add_argument("--out", default=MODEL_PATH)
    a = ap.parse_args(argv)
    rng = np.random.default_rng(2026)
    t0 = time.time()
    train_states, hold_states = [], []
    for k in range(a.variants):
        cfg = P.config_c5(k)
        st = harvest(cfg, a.steps_run, a.every, a.per_snap, rng)
        (hold_states if k >= a.variants - a.holdout else train_states).append(st)
        print(f"variant {k}: {st.shape[0]} states ({time.time() - t0:.0f} s)", flush=True)
    pp = P.config_c5(0).params()
    S = np.concatenate(train_states)
    f_all, cedt_all, tgt_all = dataset(pp, S)
    keep = stratify(S, cedt_all, a.pairs, rng)
    f, tgt = f_all[keep], tgt_all[keep]
    print(f"training pairs: {f.shape[0]} of {S.shape[0]} harvested ({time.time() - t0:.0f} s)", flush=True)
    model, hist = fit(f, tgt, a.train_steps, a.batch, a.lr0, a.lr1)
    print(f"loss: first {hist[:50].mean():.4e}, last {hist[-200:].mean():.4e} ({time.time() - t0:.0f} s)", flush=True)
    H = np.concatenate(hold_states)
    fh, _, th = dataset(pp, H)
    ev_train = errors(model, f, tgt)
    ev_hold = errors(model, fh, th)
    model.meta = {"variants_train": a.variants - a.holdout, "variants_holdout": a.holdout,
                  "pairs_train": int(f.shape[0]), "harvested": int(S.shape[0]), "train_steps": a.train_steps,
                  "batch": a.batch, "lr": [a.lr0, a.lr1], "loss_first": float(hist[:50].mean()),
                  "loss_last": float(hist[-200:].mean()), "train": ev_train, "holdout": ev_hold}
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    model.save(a.out)
    print(json.dumps(model.meta), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
