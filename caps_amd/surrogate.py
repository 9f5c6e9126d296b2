"""Planner surrogate: an MLP 5 -> 64 -> 64 -> 1 (tanh, FP32) predicting
log|logdetD| from (log R/L, log ξ(L+R)/c, m/ldim, log ωp/eV, log γ/eV).

Training data are blocks computed by the GPU path itself (caps_logdetD_batch)
on 300 configurations sampled from seed 1 (R/L log-uniform in [20, 1000],
ldim = 6 R/L, about 2e5 blocks); the network is trained by the library's
HIP forward/backward/Adam kernels (caps_amd/csrc/mlp.hip).  The free-energy
driver uses it only to plan which blocks to batch (mode "surrogate"); every
planned frequency is verified against the cutoff rules, so the energy equals
the plain form's.

Material encoding: perfect reflector -> (log ωp, log γ) = (PR_OMEGAP,
NO_GAMMA); plasma (γ = 0) -> (log ωp, NO_GAMMA); Drude -> (log ωp, log γ).

    python -m caps_amd.surrogate --out caps_amd/surrogate_weights.json
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import caps_amd  # noqa: E402
from caps_amd import lib, _dptr  # noqa: E402

PR_OMEGAP = math.log(1000.0)   # "infinite" plasma frequency
NO_GAMMA = math.log(1e-4)      # no dissipation
DEFAULT_WEIGHTS = os.path.join(HERE, "surrogate_weights.json")
NIN = 5


def _sig():
    L = lib()
    p, i, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
    fp = ctypes.POINTER(ctypes.c_float)
    for name, res, args in [
        ("capsx_mlp_new", p, [ctypes.c_ulonglong]), ("capsx_mlp_free", None, [p]),
        ("capsx_mlp_nparams", i, []), ("capsx_mlp_get_params", None, [p, fp]),
        ("capsx_mlp_set_params", None, [p, fp]), ("capsx_mlp_set_norm", None, [p, fp, fp, f, f]),
        ("capsx_mlp_get_norm", None, [p, fp, fp, fp, fp]), ("capsx_mlp_predict", None, [p, fp, i, fp]),
        ("capsx_mlp_grad", None, [p, fp, fp, i, fp]),
        ("capsx_mlp_train", f, [p, fp, fp, i, i, i, f, ctypes.c_ulonglong, fp]),
        ("capsx_set_surrogate", None, [p, p, f, f]), ("capsx_plan_batches", ctypes.c_long, [p, i]),
    ]:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def material_features(omegap_eV=None, gamma_eV=None):
    if omegap_eV is None:
        return PR_OMEGAP, NO_GAMMA
    return math.log(omegap_eV), (math.log(gamma_eV) if gamma_eV else NO_GAMMA)


class MLP:
    """The surrogate network (device parameters owned by the library)."""

    def __init__(self, seed=1):
        self.L = _sig()
        self.ptr = self.L.capsx_mlp_new(seed)
        self.n = self.L.capsx_mlp_nparams()

    def __del__(self):
        if getattr(self, "ptr", None):
            self.L.capsx_mlp_free(self.ptr)
            self.ptr = None

    def params(self):
        out = np.empty(self.n, dtype=np.float32)
        self.L.capsx_mlp_get_params(self.ptr, _fptr(out))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self.L.capsx_mlp_set_params(self.ptr, _fptr(p))

    def norm(self):
        xm, xs = np.empty(NIN, np.float32), np.empty(NIN, np.float32)
        ym, ys = ctypes.c_float(), ctypes.c_float()
        self.L.capsx_mlp_get_norm(self.ptr, _fptr(xm), _fptr(xs), ctypes.byref(ym), ctypes.byref(ys))
        return xm, xs, ym.value, ys.value

    def set_norm(self, xm, xs, ym, ys):
        xm = np.ascontiguousarray(xm, np.float32)
        xs = np.ascontiguousarray(xs, np.float32)
        self.L.capsx_mlp_set_norm(self.ptr, _fptr(xm), _fptr(xs), ym, ys)

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.empty(len(X), dtype=np.float32)
        self.L.capsx_mlp_predict(self.ptr, _fptr(X), len(X), _fptr(Y))
        return Y

    def grad(self, Xn, Tn):
        """Gradient of the standardised MSE (inputs already standardised)."""
        Xn = np.ascontiguousarray(Xn, dtype=np.float32)
        Tn = np.ascontiguousarray(Tn, dtype=np.float32)
        g = np.empty(self.n, dtype=np.float32)
        self.L.capsx_mlp_grad(self.ptr, _fptr(Xn), _fptr(Tn), len(Xn), _fptr(g))
        return g

    def train(self, X, Y, epochs=300, batch=512, lr=2e-3, seed=1):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        curve = np.empty(epochs, dtype=np.float32)
        loss = self.L.capsx_mlp_train(self.ptr, _fptr(X), _fptr(Y), len(X), epochs, batch, lr, seed, _fptr(curve))
        return loss, curve

    def save(self, path):
        xm, xs, ym, ys = self.norm()
        with open(path, "w") as f:
            json.dump({"layout": "W1[64][5] b1[64] W2[64][64] b2[64] W3[64] b3",
                       "features": "log R/L, log xi(L+R)/c, m/ldim, log omegap/eV, log gamma/eV",
                       "pr_omegap_feature": PR_OMEGAP, "no_gamma_feature": NO_GAMMA,
                       "xmean": xm.tolist(), "xstd": xs.tolist(), "ymean": ym, "ystd": ys,
                       "params": self.params().tolist()}, f)

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS):
        with open(path) as f:
            d = json.load(f)
        m = cls()
        m.set_params(np.array(d["params"], dtype=np.float32))
        m.set_norm(d["xmean"], d["xstd"], d["ymean"], d["ystd"])
        return m

    def attach(self, caps, omegap_eV=None, gamma_eV=None):
        """Let caps.free_energy(mode="surrogate") plan with this network."""
        fw, fg = material_features(omegap_eV, gamma_eV)
        self.L.capsx_set_surrogate(caps.ptr, self.ptr, fw, fg)
        caps._surrogate = self


def sample_configs(n_configs=300, seed=1):
    """(R/L, material) pairs from seed 1: R/L log-uniform in [20, 1000];
    perfect reflector, plasma and Drude in turn, with ωp log-U[1, 20] eV and
    γ log-U[0.005, 0.1] eV."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_configs):
        RbyL = float(math.exp(rng.uniform(math.log(20), math.log(1000))))
        kind = ["pr", "plasma", "drude"][k % 3]
        wp = float(math.exp(rng.uniform(0, math.log(20))))
        gm = float(math.exp(rng.uniform(math.log(0.005), math.log(0.1))))
        out.append({"RbyL": RbyL, "kind": kind, "omegap": None if kind == "pr" else wp,
                    "gamma": gm if kind == "drude" else None})
    return out


def blocks_per_config(RbyL, full=1100, knee=150.0, floor=40):
    """Blocks drawn for one configuration: `full` up to R/L = knee, then
    (knee/(R/L))^3 of that (a block costs ~ ldim^3), at least `floor`."""
    return max(floor, int(full * min(1.0, (knee / RbyL) ** 3)))


def block_sample(RbyL, ldim, nblocks, rng):
    """(ξ(L+R)/c, m) of the blocks of one configuration: ξ(L+R)/c
    log-uniform in [0.02, 40 (1 + R/L)] (where the Matsubara terms and most
    of the T = 0 integral lie; beyond it logdetD(m = 0) < e^-80) for 70% of
    the blocks, in [1e-4, 120 (1 + R/L)] (the T = 0 quadrature nodes up to
    the frequency cutoff) for the rest; m uniform in [0, 16 sqrt(R/L) + 16]
    (about twice the m-cutoff 1e-9)."""
    wide = rng.random(nblocks) < 0.3
    lo = np.where(wide, math.log(1e-4), math.log(0.02))
    hi = np.where(wide, math.log(120 * (1 + RbyL)), math.log(40 * (1 + RbyL)))
    x = np.exp(lo + (hi - lo) * rng.random(nblocks))
    mhi = min(ldim - 1, int(16 * math.sqrt(RbyL)) + 16)
    m = rng.integers(0, mhi + 1, size=nblocks)
    return x, m


def generate(n_configs=300, seed=1, ldim_factor=6, full=1100, log=print):
    """Blocks of the GPU path (caps_logdetD_batch, one batch per
    configuration) on the sampled configurations with ldim = 6 R/L; returns
    X (n x 5) and Y = log|logdetD| (blocks with logdetD = 0 dropped)."""
    rng = np.random.default_rng(seed + 1000)
    X, Y = [], []
    t0 = time.time()
    for k, cfg in enumerate(sample_configs(n_configs, seed)):
        L_ = 1e-6
        R_ = cfg["RbyL"] * L_
        ldim = max(20, int(math.ceil(ldim_factor * cfg["RbyL"])))
        c = caps_amd.Caps(R_, L_, ldim=ldim)
        if cfg["omegap"] is not None:
            c.set_drude(cfg["omegap"], cfg["gamma"] or 0.0)
        xs, ms = block_sample(cfg["RbyL"], ldim, blocks_per_config(cfg["RbyL"], full), rng)
        vals = c.logdetD_batch(xs, ms)
        fw, fg = material_features(cfg["omegap"], cfg["gamma"])
        keep = (vals != 0) & np.isfinite(vals)
        n = int(keep.sum())
        X.append(np.column_stack([np.full(n, math.log(cfg["RbyL"])), np.log(xs[keep]), ms[keep] / ldim,
                                  np.full(n, fw), np.full(n, fg)]))
        Y.append(np.log(np.abs(vals[keep])))
        if k % 10 == 0 or k == n_configs - 1:
            log("surrogate data: config %d/%d R/L=%.1f %s ldim=%d blocks=%d (total %d, %.0fs)" % (
                k + 1, n_configs, cfg["RbyL"], cfg["kind"], ldim, len(xs), sum(len(y) for y in Y),
                time.time() - t0), flush=True)
        del c
    return np.concatenate(X).astype(np.float32), np.concatenate(Y).astype(np.float32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--configs", type=int, default=300)
    ap.add_argument("--blocks", type=int, default=1100, help="blocks per configuration up to R/L = 150")
    ap.add_argument("--epochs", type=int, default=600)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--data", default=None, help="also save the data set (npz)")
    ap.add_argument("--load-data", default=None, help="train on a saved data set instead of generating one")
    args = ap.parse_args()
    caps_amd.lib()
    if args.load_data:
        d = np.load(args.load_data)
        X, Y = d["X"], d["Y"]
    else:
        X, Y = generate(args.configs, full=args.blocks)
    if args.data:
        np.savez(args.data, X=X, Y=Y)
    rng = np.random.default_rng(7)
    idx = rng.permutation(len(X))
    nval = max(1, len(X) // 10)
    val, tr = idx[:nval], idx[nval:]
    net = MLP(seed=1)
    t0 = time.time()
    loss, curve = net.train(X[tr], Y[tr], epochs=args.epochs, batch=args.batch)
    pv = net.predict(X[val])
    err = pv - Y[val]
    rmse = float(np.sqrt(np.mean(err ** 2)))
    print("surrogate: %d samples, train loss %.4g (standardised), validation RMSE %.3f in log|logdetD|, "
          "training %.1fs" % (len(X), loss, rmse, time.time() - t0), flush=True)
    rl = np.exp(X[val, 0])
    for lo, hi in [(20, 50), (50, 150), (150, 400), (400, 1000)]:
        sel = (rl >= lo) & (rl < hi)
        if sel.any():
            print("surrogate: validation R/L in [%d, %d): %d samples, RMSE %.3f, max |error| %.3f" % (
                lo, hi, int(sel.sum()), float(np.sqrt(np.mean(err[sel] ** 2))), float(np.max(np.abs(err[sel])))),
                flush=True)
    net.save(args.out)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
