"""GPU parity of the HIP path (caps_amd / libcaps_amd.so) against the
reference's values (fixtures under tests/golden/, written from the reference
compiled from its sources and from its own test files) and against the CPU
oracle on the same inputs.

Tolerances:
  * logdetD per block (conftest.check_blocks): trace-shortcut blocks
    (|tr M| < 1e-8, both sides return -tr M) relative 1e-9; factored blocks
    |Δ| <= 1e-9 |v_ref| + 2 max(1e-13, 1e-13 |tr M|), the second term being
    twice the reference's own HODLR tolerance (src/matrix.c:135), which bounds
    the reference's error on small blocks (measured: up to 6.8e-7 relative at
    |logdetD| = 1.9e-8, and up to 1.47 times that tolerance against the exact
    log det of the same matrix).  The GPU's own error is checked separately against
    the oracle (dense delta-form LU, exact to ~1e-16 relative) at 1e-10
    relative with no absolute floor.
  * K_ν, I and matrix elements: relative 1e-9 against the reference
    (QUADPACK at epsrel 1e-8) and 1e-11 against the oracle (adaptive
    quadrature at 1e-12), both limited by FP64 noise on log K ~ 1e4.
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, check_blocks, load_golden, relerr

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
    """Every Mie golden of the reference's tests (test_mie.c, test_mie_drude.c)
    through the device function, one batched launch per material/geometry."""
    rows = load_golden(fixture)
    groups = {}
    for r in rows:
        groups.setdefault((r["R"], r["L"], r["omegap"], r["gamma"]), []).append(r)
    bad, worst = [], 0.0
    for (R, L, wp, g), rs in groups.items():
        c = ca.Caps(R, L)
        if wp is not None:
            c.set_drude(wp * HBAR_EV, g * HBAR_EV)
        lna, lnb = c.mie_points([r["xi_"] for r in rs], [r["l"] for r in rs])
        for r, a, b in zip(rs, lna, lnb):
            v = a if r["which"] == "lna" else b
            e = relerr(v, r["expect"])
            worst = max(worst, e / r["tol"])
            if e >= r["tol"]:
                bad.append((r["xi_"], r["l"], r["which"], v, r["expect"]))
    print("%s: %d goldens on the device, largest error / tolerance %.3f" % (fixture, len(rows), worst))
    assert not bad, bad[:5]


def test_fresnel_golden_gpu(ca):
    """All 3430 Fresnel goldens (test_fresnel.c) in one launch per material."""
    rows = load_golden("fresnel.json")
    groups = {}
    for r in rows:
        groups.setdefault((r["R"], r["L"], r["omegap"], r["gamma"]), []).append(r)
    bad, worst = [], 0.0
    for (R, L, wp, g), rs in groups.items():
        c = ca.Caps(R, L)
        c.set_drude(wp * HBAR_EV, g * HBAR_EV)
        te, tm = c.fresnel_points([r["xi_"] for r in rs], [r["k_"] for r in rs])
        for r, a, b in zip(rs, te, tm):
            v = a if r["which"] == "rTE" else b
            e = relerr(v, r["expect"])
            worst = max(worst, e / r["tol"])
            if e >= r["tol"]:
                bad.append((r, v))
    print("fresnel: %d goldens on the device, largest error / tolerance %.3f" % (len(rows), worst))
    assert not bad, bad[:3]


def test_bessel_golden_gpu(ca):
    """The half-integer Bessel goldens (log I_{n+1/2}, log K_{n+1/2}) and the
    ratio I_ν/I_{ν+1} of test_bessels.c through the device functions the
    Mie coefficients use, in one launch (2643 + 396 values)."""
    rows = [r for r in load_golden("bessel.json") if r["fn"] in ("logIn_half", "logKn_half", "ratioI")]
    orders = [int(r["args"][0]) if r["fn"] != "ratioI" else -1 for r in rows]
    nus = [r["args"][0] if r["fn"] == "ratioI" else -1.0 for r in rows]
    xs = [r["args"][1] for r in rows]
    lI, lK, rat = ca.bessel_points(orders, nus, xs)
    bad, worst = [], 0.0
    for r, a, b, q in zip(rows, lI, lK, rat):
        v = {"logIn_half": a, "logKn_half": b, "ratioI": q}[r["fn"]]
        e = relerr(v, r["expect"]) if r["expect"] != 0 else abs(v)
        worst = max(worst, e / r["tol"])
        if not e < r["tol"]:
            bad.append((r["fn"], r["args"], v, r["expect"]))
    print("bessel: %d goldens on the device, largest error / tolerance %.3f" % (len(rows), worst))
    assert not bad, bad[:5]


TABLES = _ref("ref_tables.json") + (_ref("ref_tables_c2.json")
                                    if os.path.exists(os.path.join(GOLDEN, "ref_tables_c2.json")) else [])


@pytest.mark.parametrize("case", TABLES, ids=lambda c: "%s_ldim%d_xi%g_m%d" % (c["kind"], c["ldim"], c["xi_"], c["m"]))
def test_K_I_M_vs_reference(ca, case):
    """K_ν, I_{l1 l2} and matrix elements against the reference compiled from
    its sources.  K: relative 1e-9.  I: an alternating q-sum Σ_q a_q K_{ν-2q}
    whose terms can cancel (far from the diagonal: 2e6-fold at l1 = 610,
    l2 = 10 of the R/L = 200, m = 10 block, where the reference and the
    oracle differ by 3.6e-8 relative); its error scales with Σ_q |a_q K|
    (from the oracle), so |ΔI| <= 1e-9 Σ_q |a_q K|, i.e. 1e-9 relative
    where the terms do not cancel.  Elements: 1e-9 relative plus
    1e-14 sqrt|M_ii M_jj| (the cancelling pairs' elements are ~1e-21 of
    that scale; an error there cannot reach the determinant)."""
    import pyoracle as po
    c = make(ca, case["R"], case["L"], case["ldim"], case["omegap"], case["gamma"])
    t = c.tables(case["xi_"], case["m"])
    mu, lo = t["mu"], t["lo"]
    for nu, p, v, s in case["K"]:
        if nu - mu >= t["lnK"].shape[1]:
            continue
        assert relerr(t["lnK"][p, nu - mu], v) < 1e-12 or abs(math.exp(t["lnK"][p, nu - mu] - v) - 1) < 1e-9, \
            ("K", nu, p, t["lnK"][p, nu - mu], v)
    oc = po.caps(case["R"], case["L"], ldim=case["ldim"], epsrel=1e-12,
                 drude=None if case["omegap"] is None else (case["omegap"], case["gamma"]))
    B = po.Block(oc, case["xi_"], case["m"])
    worst_cond = 1.0
    for l1, l2, p, v, s in case["I"]:
        gv = t["logI"][p, l1 - lo, l2 - lo]
        gs = t["sgnI"][p, l1 - lo, l2 - lo]
        assert gs == s, ("I sign", l1, l2, p, gs, s)
        cond = max(1.0, math.exp(B.I_abs(l1, l2, p) - v))
        worst_cond = max(worst_cond, cond)
        assert abs(math.exp(gv - v) - 1) < 1e-9 * cond, ("I", l1, l2, p, gv, v, cond)
    M, lmin, obj = c.matrix(case["xi_"], case["m"])
    assert lmin == case["lmin"]
    for l1, l2, p1, p2, v in case["M"]:
        g = obj.elem(l1, l2, p1, p2)
        scale = math.sqrt(abs(obj.elem(l1, l1, p1, p1) * obj.elem(l2, l2, p2, p2)))
        assert abs(g - v) <= 1e-9 * abs(v) + 1e-14 * scale + 1e-300, ("M", l1, l2, p1, p2, g, v, scale)
    print("%s ldim=%d xi=%g m=%d: %d K, %d I (largest q-sum condition %.1e), %d elements" % (
        case["kind"], case["ldim"], case["xi_"], case["m"], len(case["K"]), len(case["I"]), worst_cond,
        len(case["M"])))


@pytest.mark.parametrize("case", _ref("ref_tables.json"),
                         ids=lambda c: "%s_ldim%d_xi%g_m%d" % (c["kind"], c["ldim"], c["xi_"], c["m"]))
def test_integrate_exports_vs_reference(ca, case):
    """caps_integrate_init / K / I / A / B / C / D (src/include/integration.h)
    read from the device tables, against the reference's own functions.  K
    and I as in test_K_I_M_vs_reference; A..D are sums c_i I_i of I values
    at l ± 1: each I carries up to 1e-9 of its q-sum Σ_q |a_q K| (the
    oracle's I_abs), so |ΔV| <= 1e-9 Σ_i |c_i| I_abs_i."""
    import ctypes
    import pyoracle as po
    L = ca.lib()
    c = make(ca, case["R"], case["L"], case["ldim"], case["omegap"], case["gamma"])
    t = c.tables(case["xi_"], case["m"])
    lo, nl, mu, nK = t["lo"], t["logI"].shape[1], t["mu"], t["lnK"].shape[1]
    m = case["m"]
    integ = L.caps_integrate_init(c.ptr, case["xi_"], m, 1e-8)
    assert integ
    s = ctypes.c_byte()
    oc = po.caps(case["R"], case["L"], ldim=case["ldim"], epsrel=1e-12,
                 drude=None if case["omegap"] is None else (case["omegap"], case["gamma"]))
    B = po.Block(oc, case["xi_"], m)
    try:
        nk = 0
        for nu, p, v, sg in case["K"]:
            if not mu <= nu < mu + nK:
                continue
            g = L.caps_integrate_K(integ, nu, p, ctypes.byref(s))
            assert s.value == sg and abs(math.exp(g - v) - 1) < 1e-9, ("K", nu, p, g, v, s.value, sg)
            nk += 1
        for l1, l2, p, v, sg in case["I"]:
            g = L.caps_integrate_I(integ, l1, l2, p, ctypes.byref(s))
            cond = max(1.0, math.exp(B.I_abs(l1, l2, p) - v))
            assert s.value == sg and abs(math.exp(g - v) - 1) < 1e-9 * cond, ("I", l1, l2, p, g, v, cond)

        def iabs(l1, l2, p):
            # I_abs / |V| (logs: the values span far beyond double range)
            return 0.0 if l1 < m or l2 < m else math.exp(B.I_abs(l1, l2, p) - v)

        def uv(l):
            return (l + 1.0) * (l + m) / (2 * l + 1.0), l * (l - m + 1.0) / (2 * l + 1.0)
        worst, nabcd = 0.0, 0
        for l1, l2, p, f, v, sg in case["ABCD"]:
            if min(l1, l2) - 1 < lo or max(l1, l2) + 1 >= lo + nl:
                continue
            nabcd += 1
            g = getattr(L, "caps_integrate_" + f)(integ, l1, l2, p, ctypes.byref(s))
            assert s.value == sg, (f, l1, l2, p, s.value, sg)
            if sg == 0:
                assert g == v == -math.inf
                continue
            (u1, v1), (u2, v2) = uv(l1), uv(l2)
            # bound: sum of |c_i| I_abs_i relative to |V|
            if f == "A":
                bound = m * m * iabs(l1, l2, p)
            elif f == "B":
                bound = iabs(l1, l2, p) if m == 0 else (
                    u1 * u2 * iabs(l1 - 1, l2 - 1, p) + v1 * u2 * iabs(l1 + 1, l2 - 1, p)
                    + u1 * v2 * iabs(l1 - 1, l2 + 1, p) + v1 * v2 * iabs(l1 + 1, l2 + 1, p))
            elif f == "C":
                bound = m * (u2 * iabs(l1, l2 - 1, p) + v2 * iabs(l1, l2 + 1, p))
            else:
                bound = m * (u1 * iabs(l1 - 1, l2, p) + v1 * iabs(l1 + 1, l2, p))
            err = abs(math.exp(g - v) - 1)
            tol = 1e-9 * max(1.0, bound)
            worst = max(worst, err / tol)
            assert err < tol, (f, l1, l2, p, g, v, tol)
    finally:
        L.caps_integrate_free(integ)
    print("%s m=%d: %d K, %d I, %d A..D against the reference; largest A..D error / tolerance %.3f"
          % (case["kind"], m, nk, len(case["I"]), nabcd, worst))
    assert nabcd >= 100


BLOCKS = _ref("ref_blocks.json")


@pytest.mark.parametrize("tag", ["C1", "C3s", "drude50"])
def test_logdetD_vs_reference_small(ca, tag):
    rows = [r for r in BLOCKS if r["set"] == tag]
    vals, trs, fls = [], [], []
    for r in rows:
        c = make(ca, r["R"], r["L"], r["ldim"], r["omegap"], r["gamma"])
        v, t, f = c.logdetD_batch([r["xi_"]], [r["m"]], info=True)
        vals.append(v[0]), trs.append(t[0]), fls.append(f[0])
    print(check_blocks(vals, trs, fls, [r["logdetD"] for r in rows], tag))


def test_logdetD_C1_batch_equals_single(ca):
    """All 121 blocks of the R/L = 20 configuration (PR, ξ = 1, ldim = 120) in
    one batch against the reference, each factored block in delta form; the
    same blocks one at a time give the same values bit for bit."""
    rows = [r for r in BLOCKS if r["set"] == "C1"]
    c = make(ca, 20e-6, 1e-6, 120)
    xs = np.array([r["xi_"] for r in rows])
    ms = np.array([r["m"] for r in rows])
    vb, tr, fl = c.logdetD_batch(xs, ms, info=True)
    print(check_blocks(vb, tr, fl, [r["logdetD"] for r in rows], "C1"))
    assert all(f & 4 for f in fl if (f & 3) == 0)   # every factored block: delta form
    for i in [0, 7, 25, 60, 120]:
        assert c.logdetD(xs[i], int(ms[i])) == vb[i]


def test_logdetD_C1_vs_oracle(ca):
    """The GPU's error on its own (no HODLR in the comparison): all 121 blocks
    of the R/L = 20 configuration against the oracle (adaptive quadrature at
    1e-12, dense delta-form LU) at 1e-10 relative, without an absolute floor,
    down to |logdetD| = 1e-33."""
    import pyoracle as po
    c = make(ca, 20e-6, 1e-6, 120)
    ms = np.arange(121)
    vb = c.logdetD_batch(np.ones(121), ms)
    oc = po.caps(20e-6, 1e-6, ldim=120, epsrel=1e-12)
    ref = po.logdetD_many(oc, np.ones(121), ms, threads=16)
    rel = np.abs(vb / ref - 1)
    print("C1 vs oracle: max rel %.2e (at m=%d, |v|=%.2e)" % (rel.max(), rel.argmax(), abs(ref[rel.argmax()])))
    assert rel.max() < 1e-10, (rel.argmax(), vb[rel.argmax()], ref[rel.argmax()])


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
        out, tr, fl = ca.logdetD_multi(objs, xs, ms, info=True)
    finally:
        at.set_config(0)
    print(check_blocks(out, tr, fl, [r["logdetD"] for r in rows],
                       "mixed-aspect batch (dims %d..%d)" % (min(2 * b["ldim"] for b in blocks),
                                                             max(2 * b["ldim"] for b in blocks))))
    # the same blocks alone (other panel widths, same arithmetic up to rounding)
    for i in (0, 2, 21):
        v = objs[i].logdetD(xs[i], ms[i])
        assert abs(v - out[i]) <= 1e-12 * abs(out[i]) + 1e-15, (i, v, out[i])


def test_mixed_aspect_small_blocks_exact(ca):
    """The GPU's own error on the factored mixed-aspect blocks with |logdetD|
    < 1e-3 (where the reference's HODLR tolerance dominates the comparison
    above): against the Mercator series of the block's own matrix M
    (caps_M_init), which never forms 1 - M, at 1e-12 relative."""
    from conftest import mercator_logdet
    from caps_amd.workloads import mixed_aspect_blocks, make_objects
    blocks = mixed_aspect_blocks()
    objs = make_objects(blocks)
    xs = [b["xi_"] for b in blocks]
    ms = [b["m"] for b in blocks]
    out, tr, fl = ca.logdetD_multi(objs, xs, ms, info=True)
    small = [i for i in range(64) if (fl[i] & 3) == 0 and abs(out[i]) < 1e-3]
    assert small, "no small factored block in the batch"
    worst = 0.0
    for i in small:
        M, _, _ = objs[i].matrix(xs[i], ms[i])
        ref = mercator_logdet(M)
        e = abs(out[i] / ref - 1)
        worst = max(worst, e)
        assert fl[i] & 4 and e < 1e-12, (i, out[i], ref, fl[i])
    print("mixed-aspect: %d small factored blocks (|v| %.1e..%.1e) vs Mercator of their own M: max rel %.1e" % (
        len(small), min(abs(out[i]) for i in small), max(abs(out[i]) for i in small), worst))


def test_mixed_batch_rejects_mixed_modes(ca):
    a = make(ca, 50e-6, 1e-6, 100)
    b = make(ca, 60e-6, 1e-6, 120, detalg="LU")
    with pytest.raises(ValueError):
        ca.logdetD_multi([a, b], [1.0, 1.0], [0, 0])


def test_logdetD_C2_blocks(ca):
    """The R/L = 200 Drude blocks (N = 2400) against the reference, and the
    same blocks against the oracle (adaptive quadrature at 1e-12, dense
    delta-form LU) at 1e-10 relative with no absolute floor."""
    import pyoracle as po
    rows = [r for r in BLOCKS if r["set"] == "C2"]
    r0 = rows[0]
    c = make(ca, r0["R"], r0["L"], r0["ldim"], r0["omegap"], r0["gamma"])
    xs = np.array([r["xi_"] for r in rows])
    ms = np.array([r["m"] for r in rows])
    vb, tr, fl = c.logdetD_batch(xs, ms, info=True)
    print(check_blocks(vb, tr, fl, [r["logdetD"] for r in rows], "C2 blocks"))
    oc = po.caps(r0["R"], r0["L"], ldim=r0["ldim"], epsrel=1e-12, drude=(r0["omegap"], r0["gamma"]))
    ref = po.logdetD_many(oc, xs, ms, threads=16)
    rel = np.abs(vb / ref - 1)
    print("C2 blocks vs oracle: max rel %.2e (|v| down to %.1e)" % (rel.max(), np.abs(ref).min()))
    assert rel.max() < 1e-10, (rel.argmax(), vb[rel.argmax()], ref[rel.argmax()])


GOLD = load_golden("logdetD.json")


@pytest.mark.parametrize("row", [r for r in GOLD if r["ldim"] == 500], ids=lambda r: "xi%g_m%d" % (r["xi_"], r["m"]))
def test_logdetD_reference_goldens(ca, row):
    c = ca.Caps(row["R"], row["L"], ldim=row["ldim"])
    v = c.logdetD(row["xi_"], row["m"])
    assert relerr(v, row["expect"]) < row["tol"], (v, row["expect"])


def test_logdetD_vs_oracle_lu_mode(ca):
    """LU-mode rules (src/matrix.c:489-505) against the oracle on Drude
    blocks: each of the trace shortcut, the Mercator series and the LU is
    taken where the reference takes it (flags of capsx_logdetD_multi_info),
    values to 1e-10 relative."""
    import pyoracle as po
    oc = po.caps(20e-6, 1e-6, ldim=40, detalg=1, drude=(9 / HBAR_EV, 0.035 / HBAR_EV))
    c = make(ca, 20e-6, 1e-6, 40, 9 / HBAR_EV, 0.035 / HBAR_EV, detalg="LU")
    cases = [(0.7, 0), (5.0, 3), (60.0, 20), (2.0, 35), (1.0, 12), (0.5, 8)]
    vals, tr, fl = c.logdetD_batch([x for x, _ in cases], [m for _, m in cases], info=True)
    kinds = set()
    for (xi, m), v, f in zip(cases, vals, fl):
        B = po.Block(oc, xi, m)
        M = B.matrix()
        t = math.fsum(np.diag(M))
        nrm = np.linalg.norm(M)
        merc = -t - float(np.sum(M * M.T)) / 2
        expect = 1 if abs(t) < 1e-8 else (2 if nrm < 1 and abs((nrm ** 2 / 2 + nrm + math.log1p(-nrm)) / merc) < 1e-8
                                          else 0)
        assert (f & 3) == expect, (xi, m, f, expect)
        kinds.add(expect)
        ref = po.logdetD(oc, xi, m)
        assert abs(v - ref) <= 1e-10 * abs(ref), (xi, m, v, ref)
    assert kinds == {0, 1, 2}, kinds


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


def test_logdetD_reference_goldens_ldim10000(ca):
    """All 19 of the reference's R/L = 1000, ldim = 10000 goldens (N = 20000,
    src/tests/test_logdetD.c:72-102) in one batch, at the test's own 1e-10."""
    rows = [r for r in GOLD if r["ldim"] == 10000]
    assert len(rows) == 19
    c = ca.Caps(rows[0]["R"], rows[0]["L"], ldim=10000)
    out = c.logdetD_batch([r["xi_"] for r in rows], [r["m"] for r in rows])
    t = c.last_timing()
    errs = [relerr(v, r["expect"]) for v, r in zip(out, rows)]
    print("R/L=1000 ldim=10000: 19 blocks in %.2f s, max rel err %.2e" % (t["total"] / 1e3, max(errs)))
    bad = [(r["xi_"], r["m"], v, r["expect"]) for v, r, e in zip(out, rows, errs) if not e < r["tol"]]
    assert not bad, bad


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
            for levels in (1, 2, 3, 4):
                at.set_config(levels, 1, 0, 128)
                L.capsx_lu_kernels(*default, None)
                ref = ca.lu_logdet_batch(mats)
                for gk, ps, nt in itertools.product((0, 1, 2), (0, 1), (0, 1)):
                    assert L.capsx_lu_kernels(gk, ps, nt, None) == 0
                    out = ca.lu_logdet_batch(mats)
                    assert np.array_equal(out, ref), (N, levels, gk, ps, nt, np.abs(out - ref).max())
        # the delta form (column dominant 1 - M: C tiles preloaded into the
        # accumulators), short grids: split slices, narrow and wide tiles
        for N, nmat in ((517, 3), (2700, 2)):
            Ms = np.array([0.3 * rng.normal(size=(N, N)) / N for _ in range(nmat)])
            for levels in (1, 2, 3, 4):
                at.set_config(levels, 1, 0, 128)
                L.capsx_lu_kernels(*default, None)
                ref, fl = ca.logdet_1mM_batch(Ms)
                assert fl.all()
                for gk, nt in itertools.product((0, 1, 2), (0, 1)):
                    assert L.capsx_lu_kernels(gk, -1, nt, None) == 0
                    out, _ = ca.logdet_1mM_batch(Ms)
                    assert np.array_equal(out, ref), ("delta", N, levels, gk, nt, np.abs(out - ref).max())
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


def test_graph_replay_uneven_batches(ca):
    """Batches of blocks of unequal dimension (two geometries) run their LU in
    cost-balanced stream groups; captured and replayed they give the eager
    values bit for bit, also for a second batch of the same size and largest
    dimension whose groups split differently, and with replays interleaved."""
    def objs():
        return make(ca, 20e-6, 1e-6, 120), make(ca, 10e-6, 1e-6, 60)
    big, small = objs()
    gbig, gsmall = objs()
    b1 = [big] * 8 + [small] * 12
    b2 = [big] * 12 + [small] * 8
    xs = np.linspace(0.5, 3.0, 20)
    ms = np.arange(20) % 7
    ref1 = ca.logdetD_multi(b1, xs, ms)
    ref2 = ca.logdetD_multi(b2, xs, ms)
    g1 = [gbig if o is big else gsmall for o in b1]
    g2 = [gbig if o is big else gsmall for o in b2]
    gbig.set_graph(True)
    try:
        for rep in range(3):
            v1 = ca.logdetD_multi(g1, xs, ms)
            assert np.array_equal(v1, ref1), (rep, np.abs(v1 - ref1).max())
            v2 = ca.logdetD_multi(g2, xs, ms)
            assert np.array_equal(v2, ref2), (rep, np.abs(v2 - ref2).max())
    finally:
        gbig.set_graph(False)


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
    out, tr, fl = c.logdetD_batch([r["xi_"] for r in rows], [r["m"] for r in rows], info=True)
    t = c.last_timing()
    print("C5: 8 blocks N=24000 in %.2f s (LU %.2f s)" % (t["total"] / 1e3, t["lu"] / 1e3))
    print(check_blocks(out, tr, fl, [r["logdetD"] for r in rows], "C5 blocks"))


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


def test_profiled_update_statistics(ca):
    """With per-launch profiling on, the statistics count the k_lu_gemm_dma
    launches (time, algorithmic flops, launches, bytes) and time the other
    trailing-update launches apart; the update flops stay below the LU's."""
    c = make(ca, 60e-6, 1e-6, 360)
    c.profile(True)
    try:
        c.reset_stats()
        c.logdetD_batch(np.full(24, 0.5), np.arange(24))
        st = c.stats(reset=True)
    finally:
        c.profile(False)
    assert st["blocks_factored"] > 0
    assert st["gemm_launches"] > 0 and st["gemm_ms"] > 0 and st["gemm_flops"] > 0
    assert st["other_update_ms"] >= 0
    assert st["gemm_flops"] < st["lu_flops"]
    tflops = st["gemm_flops"] / (st["gemm_ms"] * 1e-3) / 1e12
    print("profiled updates: %d k_lu_gemm_dma launches, %.2f TF/s, other updates %.3f ms"
          % (st["gemm_launches"], tflops, st["other_update_ms"]))
    assert 0 < tflops < 78.6


def test_K_segments_bitwise(ca):
    """The K quadrature's ν loop split over several workgroups per wave (short
    grids) runs the same recurrence steps: the K and I tables are the same
    bits for 1, 3 and the automatic number of segments."""
    c = make(ca, 200e-6, 1e-6, 1200, 9.0 / HBAR_EV, 0.035 / HBAR_EV)
    out = []
    try:
        for seg in ("1", "3", None):
            if seg is None:
                os.environ.pop("CAPS_AMD_K_SEGMENTS", None)
            else:
                os.environ["CAPS_AMD_K_SEGMENTS"] = seg
            out.append(c.tables(165.4563560875888, 125))
    finally:
        os.environ.pop("CAPS_AMD_K_SEGMENTS", None)
    for t in out[1:]:
        assert np.array_equal(t["lnK"], out[0]["lnK"])
        assert np.array_equal(t["logI"], out[0]["logI"]) and np.array_equal(t["sgnI"], out[0]["sgnI"])
