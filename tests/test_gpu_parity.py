"""GPU parity of the HIP path (caps_amd / libcaps_amd.so) against the
reference's values (fixtures under tests/golden/, written from the reference
compiled from its sources and from its own test files) and against the CPU
oracle on the same inputs.

Tolerances:
  * logdetD per block: |Δ| <= 1e-9 |v_ref| + 1e-12.  The absolute floor is
    the reference's own determinant accuracy: its HODLR tolerance is
    max(1e-13, 1e-13 tr M) (src/matrix.c:136), and a dense FP64 LU of 1 - M
    carries N·eps absolute rounding; blocks near the m-cutoff have
    |logdetD| ~ 1e-9, where no FP64 method holds 1e-9 relative.
  * K_ν, I and matrix elements: relative 1e-9 against the reference
    (QUADPACK at epsrel 1e-8) and 1e-11 against the oracle (adaptive
    quadrature at 1e-12), both limited by FP64 noise on log K ~ 1e4.
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_golden, relerr

pytestmark = pytest.mark.gpu

HBAR_EV = 6.582119514e-16


def _ref(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)["data"]


@pytest.fixture(scope="module")
def ca():
    import caps_amd
    caps_amd.lib()
    return caps_amd


def make(ca, R, L, ldim, omegap=None, gamma=None, detalg=None):
    c = ca.Caps(R, L, ldim=ldim, detalg=detalg)
    if omegap is not None:
        c.set_drude(omegap * HBAR_EV, gamma * HBAR_EV)
    return c


def close_block(v, ref, rtol=1e-9, atol=1e-12):
    return abs(v - ref) <= rtol * abs(ref) + atol


def test_mfma_f64_layout(ca):
    rng = np.random.default_rng(1)
    A = rng.normal(size=(16, 4))
    B = rng.normal(size=(4, 16))
    D = ca.mfma_probe(A, B)
    assert np.allclose(D, A @ B, rtol=1e-14, atol=1e-14), np.abs(D - A @ B).max()


@pytest.mark.parametrize("N", [5, 64, 100, 129, 240, 700])
def test_lu_random_vs_numpy(ca, N):
    rng = np.random.default_rng(N)
    mats = [rng.normal(size=(N, N)) for _ in range(3)]
    mats.append(np.eye(N) - 0.4 * (lambda X: (X + X.T) / 2)(rng.normal(size=(N, N)) / math.sqrt(N)))
    out = ca.lu_logdet_batch(np.array(mats))
    for A, v in zip(mats, out):
        ref = np.linalg.slogdet(A)[1]
        assert abs(v - ref) < 1e-11 * max(1.0, abs(ref)) * N, (N, v, ref)


def test_lu_pivoting_needed(ca):
    # zero leading diagonal entries force row interchanges in every panel
    rng = np.random.default_rng(7)
    N = 300
    A = rng.normal(size=(N, N))
    A[np.arange(N), np.arange(N)] = 0.0
    P = np.eye(N)[rng.permutation(N)]
    mats = np.array([A, P @ (np.eye(N) + 0.1 * A)])
    out = ca.lu_logdet_batch(mats)
    for M, v in zip(mats, out):
        assert abs(v - np.linalg.slogdet(M)[1]) < 1e-10 * N


def test_lu_large_dim(ca):
    rng = np.random.default_rng(3)
    N = 2400
    X = rng.normal(size=(N, N)) / math.sqrt(N)
    A = np.eye(N) - 0.3 * (X + X.T) / 2
    v = ca.lu_logdet_batch(A[None])[0]
    ref = np.linalg.slogdet(A)[1]
    assert abs(v - ref) < 1e-10 * abs(ref), (v, ref)


@pytest.mark.parametrize("fixture", ["mie_perf.json", "mie_drude.json"])
def test_mie_golden_gpu(ca, fixture):
    rows = load_golden(fixture)
    bad = []
    cache = {}
    for r in rows[:: max(1, len(rows) // 150)]:
        key = (r["R"], r["L"], r["omegap"], r["gamma"])
        if key not in cache:
            c = ca.Caps(r["R"], r["L"])
            if r["omegap"] is not None:
                c.set_drude(r["omegap"] * HBAR_EV, r["gamma"] * HBAR_EV)
            cache[key] = c
        lna, lnb = cache[key].mie(r["xi_"], r["l"])
        v = lna if r["which"] == "lna" else lnb
        if relerr(v, r["expect"]) >= r["tol"]:
            bad.append((r["xi_"], r["l"], r["which"], v, r["expect"]))
    assert not bad, bad[:5]


def test_fresnel_golden_gpu(ca):
    rows = load_golden("fresnel.json")[::23]
    bad = []
    for r in rows:
        c = ca.Caps(r["R"], r["L"])
        c.set_drude(r["omegap"] * HBAR_EV, r["gamma"] * HBAR_EV)
        te, tm = c.fresnel(r["xi_"], r["k_"])
        v = te if r["which"] == "rTE" else tm
        if relerr(v, r["expect"]) >= r["tol"]:
            bad.append((r, v))
    assert not bad, bad[:3]


TABLES = _ref("ref_tables.json")


@pytest.mark.parametrize("case", TABLES, ids=lambda c: "%s_xi%g_m%d" % (c["kind"], c["xi_"], c["m"]))
def test_K_I_M_vs_reference(ca, case):
    c = make(ca, case["R"], case["L"], case["ldim"], case["omegap"], case["gamma"])
    t = c.tables(case["xi_"], case["m"])
    mu, lo = t["mu"], t["lo"]
    for nu, p, v, s in case["K"]:
        if nu - mu >= t["lnK"].shape[1]:
            continue
        assert relerr(t["lnK"][p, nu - mu], v) < 1e-12 or abs(math.exp(t["lnK"][p, nu - mu] - v) - 1) < 1e-9, \
            ("K", nu, p, t["lnK"][p, nu - mu], v)
    for l1, l2, p, v, s in case["I"]:
        gv = t["logI"][p, l1 - lo, l2 - lo]
        gs = t["sgnI"][p, l1 - lo, l2 - lo]
        assert gs == s, ("I sign", l1, l2, p, gs, s)
        assert abs(math.exp(gv - v) - 1) < 1e-9, ("I", l1, l2, p, gv, v)
    M, lmin, obj = c.matrix(case["xi_"], case["m"])
    assert lmin == case["lmin"]
    for l1, l2, p1, p2, v in case["M"]:
        g = obj.elem(l1, l2, p1, p2)
        assert abs(g - v) <= 1e-9 * abs(v) + 1e-300, ("M", l1, l2, p1, p2, g, v)


BLOCKS = _ref("ref_blocks.json")


@pytest.mark.parametrize("tag", ["C1", "C3s", "drude50"])
def test_logdetD_vs_reference_small(ca, tag):
    rows = [r for r in BLOCKS if r["set"] == tag]
    r0 = rows[0]
    c = make(ca, r0["R"], r0["L"], r0["ldim"], r0["omegap"], r0["gamma"])
    bad = []
    for r in rows:
        if (r["R"], r["L"], r["ldim"]) != (r0["R"], r0["L"], r0["ldim"]):
            c = make(ca, r["R"], r["L"], r["ldim"], r["omegap"], r["gamma"])
            r0 = r
        v = c.logdetD(r["xi_"], r["m"])
        if not close_block(v, r["logdetD"]):
            bad.append((r["xi_"], r["m"], v, r["logdetD"]))
    assert not bad, bad[:5]


def test_logdetD_C1_batch_equals_single(ca):
    rows = [r for r in BLOCKS if r["set"] == "C1"]
    c = make(ca, 20e-6, 1e-6, 120)
    xs = np.array([r["xi_"] for r in rows])
    ms = np.array([r["m"] for r in rows])
    vb = c.logdetD_batch(xs, ms)
    for r, v in zip(rows, vb):
        assert close_block(v, r["logdetD"]), (r["m"], v, r["logdetD"])
    for i in [0, 7, 60, 120]:
        assert c.logdetD(xs[i], int(ms[i])) == vb[i]


@pytest.mark.parametrize("schedule", [None, (3, 1, 0, 128), (1, 1, 0, 64)], ids=["tuned", "levels3", "levels1"])
def test_mixed_aspect_batch_vs_reference(ca, schedule):
    """64 perfect-reflector blocks of 64 geometries (R/L 50..500, ldim =
    ceil(6 R/L), dimensions ~600..5500) in one batch, against the reference
    run block by block on the same inputs; with the tuned LU schedule and
    with fixed ones."""
    from caps_amd import autotune as at
    from caps_amd.workloads import mixed_aspect_blocks, make_objects
    rows = _ref("ref_mixed_aspect.json")
    blocks = mixed_aspect_blocks()
    assert len(rows) == len(blocks) == 64
    for r, b in zip(rows, blocks):
        assert (r["ldim"], r["m"], r["xi_"]) == (b["ldim"], b["m"], b["xi_"])
    objs = make_objects(blocks)
    xs = [b["xi_"] for b in blocks]
    ms = [b["m"] for b in blocks]
    try:
        if schedule:
            at.set_config(*schedule)
        out = ca.logdetD_multi(objs, xs, ms)
    finally:
        at.set_config(0)
    errs = [abs(v - r["logdetD"]) / max(abs(r["logdetD"]), 1e-300) for v, r in zip(out, rows)]
    print("mixed-aspect batch: max rel err vs reference %.2e, dims %d..%d" % (
        max(errs), min(2 * b["ldim"] for b in blocks), max(2 * b["ldim"] for b in blocks)))
    bad = [(i, xs[i], ms[i], v, r["logdetD"]) for i, (v, r) in enumerate(zip(out, rows))
           if not close_block(v, r["logdetD"])]
    assert not bad, bad[:5]
    # the same blocks alone (other panel widths, same arithmetic up to rounding)
    for i in (0, 2, 21):
        v = objs[i].logdetD(xs[i], ms[i])
        assert abs(v - out[i]) <= 1e-12 * abs(out[i]) + 1e-15, (i, v, out[i])


def test_mixed_batch_rejects_mixed_modes(ca):
    a = make(ca, 50e-6, 1e-6, 100)
    b = make(ca, 60e-6, 1e-6, 120, detalg="LU")
    with pytest.raises(ValueError):
        ca.logdetD_multi([a, b], [1.0, 1.0], [0, 0])


def test_logdetD_C2_blocks(ca):
    rows = [r for r in BLOCKS if r["set"] == "C2"]
    r0 = rows[0]
    c = make(ca, r0["R"], r0["L"], r0["ldim"], r0["omegap"], r0["gamma"])
    xs = np.array([r["xi_"] for r in rows])
    ms = np.array([r["m"] for r in rows])
    vb = c.logdetD_batch(xs, ms)
    bad = [(r["xi_"], r["m"], v, r["logdetD"]) for r, v in zip(rows, vb) if not close_block(v, r["logdetD"])]
    assert not bad, bad


GOLD = load_golden("logdetD.json")


@pytest.mark.parametrize("row", [r for r in GOLD if r["ldim"] == 500], ids=lambda r: "xi%g_m%d" % (r["xi_"], r["m"]))
def test_logdetD_reference_goldens(ca, row):
    c = ca.Caps(row["R"], row["L"], ldim=row["ldim"])
    v = c.logdetD(row["xi_"], row["m"])
    assert relerr(v, row["expect"]) < row["tol"], (v, row["expect"])


def test_logdetD_vs_oracle_lu_mode(ca):
    """LU-mode shortcuts (Mercator) and a Drude block against the oracle."""
    import pyoracle as po
    oc = po.caps(20e-6, 1e-6, ldim=40, detalg=1, drude=(9 / HBAR_EV, 0.035 / HBAR_EV))
    c = make(ca, 20e-6, 1e-6, 40, 9 / HBAR_EV, 0.035 / HBAR_EV, detalg="LU")
    for xi, m in [(0.7, 0), (5.0, 3), (60.0, 20), (2.0, 35)]:
        v = c.logdetD(xi, m)
        ref = po.logdetD(oc, xi, m)
        assert close_block(v, ref), (xi, m, v, ref)


CLI = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "caps_logdetD_amd")


@pytest.mark.skipif(not os.path.exists(CLI), reason="reference CLI relinked against libcaps_amd.so not built here")
@pytest.mark.parametrize("args,expect", [
    (["-R", "100e-6", "-L", "1e-6", "--xi", "1", "-m", "1", "-l", "501"], -6.463973151333012),
    (["-R", "100e-6", "-L", "1e-6", "--xi", "2", "-m", "1", "-l", "501", "-d", "LU"], -6.349155228127988),
])
def test_reference_cli_relinked(args, expect):
    """The reference's block CLI (src/caps_logdetD.c), linked unchanged
    against libcaps_amd.so, reproduces the manual's single-block answers
    (docs/manual/index.rst:866,888)."""
    import subprocess
    out = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    last = [ln for ln in out.stdout.splitlines() if ln and not ln.startswith("#")][-1]
    v = float(last.split(",")[4])
    assert abs(v - expect) <= 1e-9 * abs(expect), (v, expect, out.stdout)


def test_lu_big_panel_path(ca):
    """N above the 512-lane panel's 6144 rows: the 1024-lane panel kernel."""
    rng = np.random.default_rng(11)
    N = 6600
    X = rng.normal(size=(N, N)) / math.sqrt(N)
    A = np.eye(N) - 0.3 * (X + X.T) / 2
    A[:5, :5] = rng.normal(size=(5, 5))     # force row interchanges in the first panel
    v = ca.lu_logdet_batch(A[None])[0]
    ref = np.linalg.slogdet(A)[1]
    assert abs(v - ref) < 1e-9 * abs(ref), (v, ref)


@pytest.mark.parametrize("row", [r for r in GOLD if r["ldim"] == 10000 and (r["xi_"], r["m"]) in
                                 [(500.5, 1), (1.0, 0), (1000.0, 10)]],
                         ids=lambda r: "RbyL1000_xi%g_m%d" % (r["xi_"], r["m"]))
def test_logdetD_reference_goldens_ldim10000(ca, row):
    """The reference's R/L = 1000, ldim = 10000 goldens (N = 20000)."""
    c = ca.Caps(row["R"], row["L"], ldim=row["ldim"])
    v = c.logdetD(row["xi_"], row["m"])
    assert relerr(v, row["expect"]) < row["tol"], (v, row["expect"])


T0_CASES = _ref("ref_free_energy_t0.json") if os.path.exists(os.path.join(GOLDEN, "ref_free_energy_t0.json")) else []


@pytest.mark.parametrize("case", T0_CASES, ids=lambda c: "T0_%s" % c["kind"])
def test_free_energy_T0_vs_reference(ca, case):
    """T = 0 free energy at R/L = 10, ldim = 60 against the reference driver's
    integrand integrated by the reference's dqagi at epsrel 1e-10: the
    Fourier-Chebyshev levels of the batched form converge to the same
    integral within 1e-8.  The integrand itself is checked at the frequencies
    dqagi sampled (m-sum of the GPU blocks vs the reference's)."""
    c = make(ca, case["R"], case["L"], case["ldim"], case["omegap"], case["gamma"])
    F, computed, used = c.free_energy(0.0, cutoff=case["cutoff"], epsrel=1e-10, mode="batched")
    print("T=0 %s: F=%.16g reference %.16g (rel %.2e), %d blocks computed, %d used" % (
        case["kind"], F, case["F"], abs(F / case["F"] - 1), computed, used))
    assert abs(F / case["F"] - 1) < 1e-8, (F, case["F"])
    for xi, mstop, v in case["samples"][::4]:
        vals = c.logdetD_batch(np.full(mstop + 1, xi), np.arange(mstop + 1))
        got = math.fsum([vals[0] / 2] + list(vals[1:]))
        assert abs(got - v) <= 1e-9 * abs(v) + 1e-12, (xi, got, v)


def test_free_energy_T0_plain_equals_batched(ca):
    c = make(ca, 10e-6, 1e-6, 40)
    Fb, cb, ub = c.free_energy(0.0, cutoff=1e-9, epsrel=1e-6, mode="batched")
    Fp, cp, up = c.free_energy(0.0, cutoff=1e-9, epsrel=1e-6, mode="plain")
    assert ub == up and cp == up
    assert abs(Fb - Fp) <= 1e-12 * abs(Fp), (Fb, Fp)


def test_lu_schedules_agree(ca):
    """Every LU schedule (levels, streams, tile order, tile) gives the same
    determinants, with and without row interchanges."""
    import itertools
    from caps_amd import autotune as at
    rng = np.random.default_rng(5)
    # 300: a partial second 256-column step; 517: two full steps and 5
    # columns; 129: one column past a 128-column step
    for N in (300, 517, 129):
        mats = [np.eye(N) - 0.3 * rng.normal(size=(N, N)) / math.sqrt(N) for _ in range(7)]
        A = rng.normal(size=(N, N))
        A[np.arange(N), np.arange(N)] = 0.0
        mats.append(A)
        mats = np.array(mats)
        ref = [np.linalg.slogdet(M)[1] for M in mats]
        try:
            for cfg in itertools.product((1, 2, 3), (1, 2), (0, 1), (64, 128, 256)):
                at.set_config(*cfg)
                out = ca.lu_logdet_batch(mats)
                for v, r in zip(out, ref):
                    assert abs(v - r) < 1e-10 * N, (N, cfg, v, r)
        finally:
            at.set_config(0)


def test_lu_kernel_variants_bitwise(ca):
    """Every kernel variant of the LU (trailing update with operand stages
    straight into LDS, 8 or 16 deep, or through registers; panel leaves of
    8 or 4 columns; narrow tiles on or off) gives the same determinants bit
    for bit under every schedule level: each matrix entry receives the same
    products in the same order."""
    import ctypes
    import itertools
    from caps_amd import autotune as at
    L = ca.lib()
    rng = np.random.default_rng(11)
    prev = (ctypes.c_int * 3)()
    assert L.capsx_lu_kernels(-1, -1, -1, prev) == 0
    default = tuple(prev)
    assert L.capsx_lu_kernels(3, -1, -1, None) == 1
    try:
        # 2700 and 3100: the 512-lane panels (6 and 12 rows per lane)
        for N, nmat in ((517, 6), (300, 6), (129, 6), (2700, 2), (3100, 2)):
            mats = np.array([np.eye(N) - 0.3 * rng.normal(size=(N, N)) / math.sqrt(N) for _ in range(nmat)])
            for levels in (1, 2, 3):
                at.set_config(levels, 1, 0, 128)
                L.capsx_lu_kernels(*default, None)
                ref = ca.lu_logdet_batch(mats)
                for gk, ps, nt in itertools.product((0, 1, 2), (0, 1), (0, 1)):
                    assert L.capsx_lu_kernels(gk, ps, nt, None) == 0
                    out = ca.lu_logdet_batch(mats)
                    assert np.array_equal(out, ref), (N, levels, gk, ps, nt, np.abs(out - ref).max())
    finally:
        L.capsx_lu_kernels(*default, None)
        at.set_config(0)


def test_autotune_table_and_timer(ca, tmp_path):
    from caps_amd import autotune as at
    p = tmp_path / "lu_tuning.txt"
    p.write_text("# test table\n600 16 1 2 1 64\n2400 640 2 1 0 128\n")
    try:
        assert at.load_table(str(p)) == 2
        assert at.config_for(600, 16) == (1, 2, 1, 64)
        assert at.config_for(500, 12) == (1, 2, 1, 64)
        assert at.config_for(2400, 700) == (2, 1, 0, 128)
        ms = at.lu_time(240, 8, (2, 2, 0, 64), reps=1)
        assert 0 < ms < 10000
    finally:
        shipped = os.path.join(os.path.dirname(GOLDEN), "..", "caps_amd", "lu_tuning.txt")
        at.load_table(shipped if os.path.exists(shipped) else str(tmp_path / "none.txt"))


@pytest.mark.parametrize("streams", [1, 2])
def test_graph_replay_matches_eager(ca, streams):
    """A batch captured into a HIP graph and replayed gives the eager values
    bit for bit, over several replays with eager work and synchronisation in
    between, and for other contents of the same shape (other ξ)."""
    from caps_amd import autotune as at
    xs = np.full(40, 1.0)
    ms = np.arange(40)
    at.set_config(2, streams, 0, 128)
    try:
        eager = make(ca, 20e-6, 1e-6, 120)
        ref1 = eager.logdetD_batch(xs, ms)
        ref2 = eager.logdetD_batch(2 * xs, ms)
        g = make(ca, 20e-6, 1e-6, 120)
        g.set_graph(True)
        for rep in range(3):
            v1 = g.logdetD_batch(xs, ms)
            assert np.array_equal(v1, ref1), (rep, np.abs(v1 - ref1).max())
            other = eager.logdetD_batch(xs[:7] * 3, ms[:7])        # eager work between replays
            assert np.all(np.isfinite(other))
            v2 = g.logdetD_batch(2 * xs, ms)
            assert np.array_equal(v2, ref2), (rep, np.abs(v2 - ref2).max())
        g.set_graph(False)
        assert np.array_equal(g.logdetD_batch(xs, ms), ref1)
    finally:
        at.set_config(0)


def test_lu_grid_panel_small_threshold(ca):
    """The grid-wide column panel (used for tall panels) on small matrices,
    with and without row interchanges, both LU schedules."""
    from caps_amd import autotune as at
    rng = np.random.default_rng(9)
    N = 300
    mats = [np.eye(N) - 0.3 * rng.normal(size=(N, N)) / math.sqrt(N) for _ in range(3)]
    A = rng.normal(size=(N, N))
    A[np.arange(N), np.arange(N)] = 0.0
    mats.append(A)
    mats = np.array(mats)
    ref = [np.linalg.slogdet(M)[1] for M in mats]
    old = ca.lib().capsx_lu_grid_panel_rows(100)
    try:
        for levels in (1, 2):
            at.set_config(levels, 1, 0, 128)
            out = ca.lu_logdet_batch(mats)
            for v, r in zip(out, ref):
                assert abs(v - r) < 1e-10 * N, (levels, v, r)
    finally:
        ca.lib().capsx_lu_grid_panel_rows(old)
        at.set_config(0)


def test_c5_blocks_vs_reference(ca):
    """R/L = 2000, ldim = 12000 (N = 24000, 4.6 GB per matrix): eight blocks
    in one batch against the reference (its default HODLR route) on the same
    inputs; the LU's tall panels take the grid-wide panel kernels."""
    rows = _ref("ref_c5_blocks.json")
    c = ca.Caps(2e-3, 1e-6, ldim=12000)
    out = c.logdetD_batch([r["xi_"] for r in rows], [r["m"] for r in rows])
    t = c.last_timing()
    print("C5: 8 blocks N=24000 in %.2f s (LU %.2f s), max rel err %.2e" % (
        t["total"] / 1e3, t["lu"] / 1e3, max(abs(v / r["logdetD"] - 1) for v, r in zip(out, rows))))
    bad = [(r["xi_"], r["m"], v, r["logdetD"]) for v, r in zip(out, rows) if not close_block(v, r["logdetD"])]
    assert not bad, bad


XI0_ROWS = load_golden("logdetD0.json") if os.path.exists(os.path.join(GOLDEN, "logdetD0.json")) else []


@pytest.mark.parametrize("row", [r for r in XI0_ROWS if r["fn"] == "ht_plasma"],
                         ids=lambda r: "omegap%.3g" % r["omegap"])
def test_ht_plasma_vs_reference(ca, row):
    """ξ = 0 plasma-model free energy (reference test_logdetD0: R/L = 100,
    ldim = 500, ωp (L+R)/c = 0.01 .. 1000): plasma integrals on the host,
    blocks filled and factored on the GPU, m-sum to 1e-13."""
    c = ca.Caps(row["R"], row["L"], ldim=row["ldim"])
    v = c.ht_plasma(row["omegap"], row["eps"])
    print("ht_plasma omegap=%.4g: %.16g reference %.16g (rel %.1e)" % (
        row["omegap"], v, row["expect"], abs(v / row["expect"] - 1)))
    assert relerr(v, row["expect"]) < row["tol"], (v, row["expect"])


def test_logdetD0_m0_vs_reference(ca):
    """ξ = 0, m = 0 EE and MM blocks at R/L = 1000, ldim = 12000 (12000 x
    12000 matrices filled and factored on the GPU) against the reference's
    test_logdetD0 values."""
    rows = {r["fn"]: r for r in XI0_ROWS if r["fn"].startswith("logdetD0_")}
    r = rows["logdetD0_EE"]
    c = ca.Caps(r["R"], r["L"], ldim=r["ldim"])
    ee, mm = c.logdetD0(0)
    print("logdetD0 m=0 ldim=12000: EE %.16g (%.16g), MM %.16g (%.16g)" % (
        ee, r["expect"], mm, rows["logdetD0_MM"]["expect"]))
    assert relerr(ee, r["expect"]) < r["tol"], (ee, r["expect"])
    assert relerr(mm, rows["logdetD0_MM"]["expect"]) < rows["logdetD0_MM"]["tol"]


def test_logdetD0_blocks_vs_oracle(ca):
    """ξ = 0 EE, MM and plasma MM blocks of the GPU path against the oracle
    (dense CPU LU, adaptive quadrature) at R/L = 50, ldim = 150, for several
    m and plasma frequencies; and caps_ht_perf against the oracle's."""
    import ctypes
    import pyoracle as po
    L = po.lib()
    R, Lx = 50e-6, 1e-6
    c = ca.Caps(R, Lx, ldim=150)
    oc = po.caps(R, Lx, ldim=150, epsrel=1e-12)
    for m in (0, 1, 7, 30):
        for w in (0.1, 10.0, 500.0):
            omegap = w * po.CAPS_C / (R + Lx)
            ee, mm, mp = c.logdetD0(m, omegap)
            oe, om, op = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
            L.orc_logdetD0(ctypes.byref(oc), m, omegap, ctypes.byref(oe), ctypes.byref(om), ctypes.byref(op))
            for got, want in [(ee, oe.value), (mm, om.value), (mp, op.value)]:
                # absolute floor: a log det of a near-identity block carries
                # ~N ulp of rounding (these blocks have |log det| down to 1e-4)
                assert abs(got - want) <= 1e-10 * abs(want) + 1e-13, (m, w, got, want)
    oc2 = po.caps(10e-6, 1e-6, ldim=60, epsrel=1e-12)
    c2 = ca.Caps(10e-6, 1e-6, ldim=60)
    assert abs(c2.ht_perf(1e-13) / L.orc_ht_perf(ctypes.byref(oc2), 1e-13) - 1) < 1e-12


def test_caps_dump_matrix(ca, tmp_path, monkeypatch):
    """CAPS_DUMP with a dense determinant algorithm: caps_logdetD writes the
    block's M in the reference's file layout; it is the matrix caps_M_init
    builds, the oracle's block matrix to 1e-10, and log det(1 - M) of it is
    the returned value.  caps_logdetD0 writes its last block (MM), whose
    elements follow the closed form."""
    import pyoracle as po
    path = tmp_path / "M.npy"
    monkeypatch.setenv("CAPS_DUMP", str(path))
    c = make(ca, 20e-6, 1e-6, 40, detalg="LU")
    v = c.logdetD(1.0, 2)
    M = ca.npy_load(path)
    Mb, lmin, obj = c.matrix(1.0, 2)
    assert np.array_equal(M, Mb)
    N = M.shape[0]
    assert abs(np.linalg.slogdet(np.eye(N) - M)[1] - v) < 1e-12 * max(1, abs(v))
    oc = po.caps(20e-6, 1e-6, ldim=40, detalg=1)
    B = po.Block(oc, 1.0, 2)
    Mo = B.matrix()
    assert np.max(np.abs(M - Mo)) < 1e-10 * np.max(np.abs(Mo))
    # xi = 0: the MM block of m = 3 is the file's last matrix
    ee, mm = c.logdetD0(3)
    M0 = ca.npy_load(path)
    lmin0 = c.lminmax(3)[0]
    l = lmin0 + np.arange(c.ldim)
    from math import lgamma
    lf = np.vectorize(lambda n: lgamma(n + 1.0))
    y = -math.log(2) - math.log1p(1 / 20)
    L1, L2 = np.meshgrid(l, l, indexing="ij")
    ref = np.exp((L1 + L2 + 1) * y + lf(L1 + L2) - 0.5 * (lf(L1 + 3) + lf(L1 - 3) + lf(L2 + 3) + lf(L2 - 3)))
    ref *= np.sqrt(L1 * L2 / ((L1 + 1.0) * (L2 + 1.0)))
    assert np.max(np.abs(M0 - ref) / ref) < 1e-11
    assert abs(np.linalg.slogdet(np.eye(c.ldim) - M0)[1] - mm) < 1e-12 * max(1, abs(mm))


def test_free_energy_psd_equals_matsubara(ca):
    """T > 0 by the Padé spectrum decomposition (the driver's automatic
    order) and by the Matsubara sum give the same free energy to the
    requested accuracy (Drude, R/L = 50, ldim = 300, T = 300 K, epsrel
    1e-8), with fewer frequencies; the plain and batched PSD forms sum the
    same blocks."""
    c = make(ca, 50e-6, 1e-6, 300, 9 / HBAR_EV, 0.035 / HBAR_EV)
    Fm, cm, um = c.free_energy_ex(300.0, epsrel=1e-8)
    Fp, cp, up = c.free_energy_ex(300.0, epsrel=1e-8, psd_order=-1)
    print("MSD F=%.16g (%d blocks), PSD F=%.16g (%d blocks), rel %.2e" % (Fm, um, Fp, up, abs(Fp / Fm - 1)))
    assert abs(Fp / Fm - 1) < 1e-7 and up < um
    Fp0, cp0, up0 = c.free_energy_ex(300.0, epsrel=1e-8, mode="plain", psd_order=-1)
    assert up0 == up and cp0 == up0 and abs(Fp0 - Fp) <= 1e-12 * abs(Fp)


def test_free_energy_xi0_models(ca):
    """The driver's ξ = 0 terms for perfect reflectors and the plasma model
    (Drude EE term plus Σ'_m of the MM blocks up to the m-cutoff): the
    difference to the Drude-term energy is T_scaled/(2π) Σ'_m MM, checked
    against the oracle's blocks (R/L = 10, ldim = 60)."""
    import ctypes
    import pyoracle as po
    L = po.lib()
    R, Lx, T, cutoff = 10e-6, 1e-6, 300.0, 1e-9
    c = ca.Caps(R, Lx, ldim=60)
    oc = po.caps(R, Lx, ldim=60, epsrel=1e-12)
    T_scaled = 2 * math.pi * 1.38064852e-23 * (R + Lx) * T / (1.0545718e-34 * 299792458.0)
    Fd, _, _ = c.free_energy_ex(T, cutoff=cutoff, xi0="drude")
    wp = 9 / HBAR_EV
    for model, omegap in [("pr", 0.0), ("plasma", wp)]:
        F, _, _ = c.free_energy_ex(T, cutoff=cutoff, xi0=model, omegap=omegap)
        got = (F - Fd) * math.pi / T_scaled * 2
        v = []
        for m in range(1000):
            x, y = ctypes.c_double(), ctypes.c_double()
            if model == "pr":
                L.orc_logdetD0(ctypes.byref(oc), m, 0.0, None, ctypes.byref(x), None)
            else:
                L.orc_logdetD0(ctypes.byref(oc), m, omegap, None, None, ctypes.byref(x))
            v.append(x.value)
            if x.value == 0 or x.value / v[0] < cutoff:
                break
        want = math.fsum([v[0] / 2] + v[1:])
        print("xi=0 %s: %.16g oracle %.16g (%d m)" % (model, got, want, len(v)))
        assert abs(got - want) <= 1e-9 * abs(want), (model, got, want)


def test_tabulated_material_block(ca, tmp_path):
    """A tabulated material (the reference's file format) that lists the
    block's frequency reproduces the analytic Drude block to rounding; on a
    coarse table the block follows the interpolated ε (bounded change)."""
    import numpy as np
    from test_cpu_abi import _write_drude_table
    R, Lx = 100e-6, 1e-6
    xi_ = 3.0
    xi = xi_ * 299792458.0 / (R + Lx)          # rad/s of the block
    grid = np.unique(np.concatenate([np.logspace(11, 17, 300), [xi]]))
    p = tmp_path / "drude.csv"
    _write_drude_table(p, grid, 9.0, 0.035)
    a = make(ca, R, Lx, 600, 9 / HBAR_EV, 0.035 / HBAR_EV)
    b = ca.Caps(R, Lx, ldim=600)
    b.set_material(p)
    for m in (0, 5, 40):
        va, vb = a.logdetD(xi_, m), b.logdetD(xi_, m)
        assert abs(va - vb) <= 1e-12 * abs(va), (m, va, vb)
    q = tmp_path / "coarse.csv"
    _write_drude_table(q, np.logspace(11, 17, 13), 9.0, 0.035)
    cc = ca.Caps(R, Lx, ldim=600)
    cc.set_material(q)
    v, ref = cc.logdetD(xi_, 5), a.logdetD(xi_, 5)
    assert v != ref and abs(v / ref - 1) < 1e-2
