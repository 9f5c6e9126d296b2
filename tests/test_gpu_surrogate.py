"""The planner surrogate (MLP 5-64-64-1, tanh, FP32) on the GPU: forward and
backward kernels against a float64 NumPy restatement of the same network,
Adam training, and what the trained network gains the free-energy driver
(C2 workload) over the same driver without it."""
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

H, NIN = 64, 5


def unpack(p):
    p = p.astype(np.float64)
    o = 0
    W1 = p[o:o + H * NIN].reshape(H, NIN); o += H * NIN
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def np_forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ W3 + b3, h1, h2


def np_grad(p, X, T):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    y, h1, h2 = np_forward(p, X)
    n = len(X)
    dy = 2 * (y - T) / n
    gW3 = dy @ h2
    gb3 = dy.sum()
    dh2 = np.outer(dy, W3) * (1 - h2 ** 2)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2) * (1 - h1 ** 2)
    gW1 = dh1.T @ X
    gb1 = dh1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])


@pytest.fixture(scope="module")
def sur():
    import caps_amd
    caps_amd.lib()
    from caps_amd import surrogate
    return surrogate


def test_forward_matches_numpy(sur):
    net = sur.MLP(seed=3)
    rng = np.random.default_rng(0)
    X = rng.normal(size=(300, NIN)).astype(np.float32)
    y = net.predict(X)          # identity standardisation
    ref, _, _ = np_forward(net.params(), X.astype(np.float64))
    assert np.max(np.abs(y - ref)) < 1e-5 * max(1, np.max(np.abs(ref)))


def test_backward_matches_numpy(sur):
    net = sur.MLP(seed=5)
    rng = np.random.default_rng(1)
    for n in [1, 64, 200]:
        X = rng.normal(size=(n, NIN)).astype(np.float32)
        T = rng.normal(size=n).astype(np.float32)
        g = net.grad(X, T).astype(np.float64)
        ref = np_grad(net.params(), X.astype(np.float64), T.astype(np.float64))
        err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
        assert err < 1e-5, (n, err)
        # and against a central difference of the loss on a few parameters
        p0 = net.params().astype(np.float64)
        for k in [0, 400, 2000, 4608]:
            e = np.zeros_like(p0)
            e[k] = 1e-3
            lp = np.mean((np_forward(p0 + e, X)[0] - T) ** 2)
            lm = np.mean((np_forward(p0 - e, X)[0] - T) ** 2)
            assert abs((lp - lm) / 2e-3 - g[k]) < 2e-3 * max(1e-3, abs(g[k])) + 1e-6, (k, g[k])


def test_adam_fits_smooth_function(sur):
    rng = np.random.default_rng(2)
    X = rng.uniform(-2, 2, size=(4096, NIN)).astype(np.float32)
    Y = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.3 * X[:, 3] ** 2 + 0.1 * X[:, 4]).astype(np.float32)
    net = sur.MLP(seed=1)
    loss, curve = net.train(X, Y, epochs=150, batch=256, lr=3e-3)
    assert curve[-1] < 0.05 * curve[0] and loss < 0.02, (curve[0], curve[-1])


WEIGHTS = os.path.join(os.path.dirname(GOLDEN), "..", "caps_amd", "surrogate_weights.json")


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="no trained surrogate weights committed")
def test_surrogate_planned_free_energy_c2(sur):
    """C2 free energy with the surrogate-planned batch: same energy as the
    reference run and as the plain form's block set, in fewer launch
    sequences than the driver without the surrogate."""
    import caps_amd
    with open(os.path.join(GOLDEN, "ref_free_energy_c2.json")) as f:
        ref = json.load(f)["data"]
    c = caps_amd.Caps(200e-6, 1e-6, ldim=1200)
    c.set_drude(9.0, 0.035)
    L = sur._sig()
    L.capsx_plan_batches(c.ptr, 1)
    F1, comp1, used1 = c.free_energy(300.0, mode="batched")
    b1 = L.capsx_plan_batches(c.ptr, 1)
    net = sur.MLP.load(WEIGHTS)
    net.attach(c, 9.0, 0.035)
    F2, comp2, used2 = c.free_energy(300.0, mode="surrogate")
    b2 = L.capsx_plan_batches(c.ptr, 1)
    rep = c.surrogate_report()
    assert used1 == used2 == ref["determinants"]
    assert abs(F2 / ref["F"] - 1) < 1e-8 and abs(F1 - F2) <= 1e-13 * abs(F1)
    print("without surrogate: %d launch sequences, %d blocks computed; with: %d, %d" % (b1, comp1, b2, comp2))
    print("surrogate vs computed blocks:", rep)
    assert b2 < b1
    # every planned frequency is scored, and the m-cutoff is predicted to a
    # few m (the plan adds a margin of max(4, m/25))
    assert rep["frequencies"] >= 5 and rep["blocks"] >= used2
    assert rep["rms_log_error"] < 1.0 and rep["mean_mcut_error"] < 8


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="no trained surrogate weights committed")
@pytest.mark.parametrize("RbyL,ldim,material,epsrel", [(50, 300, None, 1e-8), (100, 600, (9.0, 0.035), 1e-6)],
                         ids=["pr_RbyL50", "drude_RbyL100"])
def test_surrogate_planned_free_energy_T0(sur, RbyL, ldim, material, epsrel):
    """T = 0 free energy with every quadrature node's m-range planned by the
    surrogate: the same energy and the same blocks in the sums as the batched
    driver that guesses m-ranges from neighbouring nodes (which itself equals
    the plain form, test_free_energy_T0_plain_equals_batched), in no more
    launch sequences."""
    import caps_amd
    c = caps_amd.Caps(RbyL * 1e-6, 1e-6, ldim=ldim)
    if material:
        c.set_drude(*material)
    L = sur._sig()
    L.capsx_plan_batches(c.ptr, 1)
    F1, comp1, used1 = c.free_energy(0.0, cutoff=1e-9, epsrel=epsrel, mode="batched")
    b1 = L.capsx_plan_batches(c.ptr, 1)
    net = sur.MLP.load(WEIGHTS)
    net.attach(c, *(material or (None, None)))
    c.surrogate_report(reset=True)
    F2, comp2, used2 = c.free_energy(0.0, cutoff=1e-9, epsrel=epsrel, mode="surrogate")
    b2 = L.capsx_plan_batches(c.ptr, 1)
    rep = c.surrogate_report()
    print("T=0 R/L=%d: F %.16g / %.16g; without surrogate %d launch sequences, %d blocks; with %d, %d; %s" % (
        RbyL, F1, F2, b1, comp1, b2, comp2, rep))
    assert used1 == used2 and abs(F1 - F2) <= 1e-13 * abs(F1)
    assert rep["frequencies"] > 0 and rep["rms_log_error"] < 1.0
    assert b2 <= b1
