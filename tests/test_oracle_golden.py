"""Pin the CPU oracle against the golden vectors held by the reference's own
tests (transcribed by tests/golden/extract_reference_vectors.py), at the
tolerance each reference test uses, and against closed forms.

Relative error is the reference's criterion: x == y or |1 - x/y| < tol
(src/tests/unittest.c:64-78)."""
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_golden, relerr
import pyoracle as po


def test_lnPlm_golden():
    L = po.lib()
    bad = []
    for r in load_golden("lnPlm.json"):
        l, m, x = int(r["args"][0]), int(r["args"][1]), r["args"][2]
        v = L.orc_lnPlm(l, m, x)
        if relerr(v, r["expect"]) >= r["tol"]:
            bad.append((l, m, x, v, r["expect"], r["src"]))
    assert not bad, bad[:10]


def test_lnLambda_golden():
    L = po.lib()
    bad = []
    for r in load_golden("lnLambda.json"):
        l1, l2, m = map(int, r["args"])
        v = L.orc_lnLambda(l1, l2, m)
        if relerr(v, r["expect"]) >= r["tol"]:
            bad.append((l1, l2, m, v, r["expect"]))
    assert not bad, bad[:10]


# The oracle's Bessel functions cover what the path uses: half-integer
# orders (Mie), the ratio I_ν/I_{ν+1}, and integer orders for arguments
# below 30 (the small-argument Legendre expansion).
@pytest.mark.parametrize("fn", ["logIn_half", "logKn_half", "ratioI", "logIn", "logI0", "logI1", "I0", "I1"])
def test_bessel_golden(fn):
    L = po.lib()
    rows = [r for r in load_golden("bessel.json") if r["fn"] == fn]
    assert rows
    bad, n = [], 0
    for r in rows:
        a = r["args"]
        if fn == "logIn_half":
            v = L.orc_bessel_logIn_half(int(a[0]), a[1])
        elif fn == "logKn_half":
            v = L.orc_bessel_logKn_half(int(a[0]), a[1])
        elif fn == "ratioI":
            v = L.orc_bessel_ratioI(a[0], a[1])
        else:
            if fn == "logIn":
                order, x = int(a[0]), a[1]
            else:
                order, x = (0 if "0" in fn else 1), a[0]
            if x > 30 or (order < 100 and x == 0):
                continue
            v = L.orc_bessel_logIn(order, x)
            if not fn.startswith("log"):
                v = math.exp(v)
        n += 1
        if relerr(v, r["expect"]) >= r["tol"]:
            bad.append((fn, a, v, r["expect"], r["src"]))
    assert n > 20
    assert not bad, bad[:10]


@pytest.mark.parametrize("fixture", ["mie_perf.json", "mie_drude.json"])
def test_mie_golden(fixture):
    bad = []
    for r in load_golden(fixture):
        drude = None if r["omegap"] is None else (r["omegap"], r["gamma"])
        c = po.caps(r["R"], r["L"], drude=drude)
        lna, lnb = po.mie(c, r["xi_"], r["l"])
        v = lna if r["which"] == "lna" else lnb
        if relerr(v, r["expect"]) >= r["tol"]:
            bad.append((r["xi_"], r["l"], r["which"], v, r["expect"], r["src"]))
    assert not bad, bad[:10]


def test_fresnel_golden():
    bad = []
    for r in load_golden("fresnel.json"):
        c = po.caps(r["R"], r["L"], drude=(r["omegap"], r["gamma"]))
        epsm1 = po.lib().orc_epsm1_plate(c, r["xi_"]) if False else None
        # ε(iξ)-1 of the plate at the frequency of the call
        xi = r["xi_"] * po.CAPS_C / (r["R"] + r["L"])
        epsm1 = r["omegap"] ** 2 / (xi * (xi + r["gamma"]))
        rTE, rTM = po.fresnel(epsm1, r["xi_"], r["k_"])
        v = rTE if r["which"] == "rTE" else rTM
        if relerr(v, r["expect"]) >= r["tol"]:
            bad.append((r["xi_"], r["k_"], r["which"], v, r["expect"], r["src"]))
    assert not bad, bad[:10]


def test_lfac_logi_closed_form():
    L = po.lib()
    for n in list(range(0, 3000)) + [10**5, 10**6, 10**8]:
        assert relerr(L.orc_lfac(n), math.lgamma(n + 1)) < 1e-15 or n < 2
        if n > 0:
            assert relerr(L.orc_logi(n), math.log(n)) < 1e-15
    # n!! for even and odd n
    for n in range(1, 60):
        exact = math.log(math.prod(range(n, 0, -2)))
        assert abs(L.orc_lfac2(n) - exact) < 1e-12 * max(1, abs(exact))


def test_plm_routes_agree():
    """The three lnPlm routes (upward recurrence, continued fraction +
    downward recurrence, Bogaert expansions) agree where they overlap: a
    planted index error in any one of them breaks this."""
    L = po.lib()
    for l, m, x in [(150, 3, 1.2), (300, 20, 1.05), (250, 100, 3.0), (400, 2, 1.001), (260, 40, 50.0)]:
        up = L.orc_lnPlm_upwards(l, m, x)
        dn = L.orc_lnPlm_downwards(l, m, x)
        assert relerr(up, dn) < 1e-12, (l, m, x, up, dn)
    # m = 0: Bonnet recurrence (l < 100) vs. P_l^0 from the upward route
    for l, x in [(50, 1.3), (99, 1.01), (20, 7.0)]:
        assert relerr(L.orc_lnPl(l, x), L.orc_lnPlm_upwards(l, 0, x)) < 1e-13
    # l >= 100 uses the Bogaert expansions; compare with the exact recurrence
    for l, x in [(100, 1.0001), (150, 1.00001), (500, 1.2), (2000, 3.0), (120, 1.002)]:
        assert relerr(L.orc_lnPl(l, x), L.orc_lnPlm_upwards(l, 0, x)) < 1e-12


def test_dlnPlm_finite_difference():
    """d/dx log P_l^m and d²/dx² log P_l^m against central differences."""
    L = po.lib()
    import ctypes
    for l, m, x in [(40, 4, 1.5), (300, 10, 2.0), (600, 200, 1.3), (20, 2, 4.0)]:
        d2 = ctypes.c_double()
        d1 = L.orc_dlnPlm(l, m, x, ctypes.byref(d2))
        h = 1e-5 * x
        f = lambda t: L.orc_lnPlm(l, m, t)
        fd1 = (f(x + h) - f(x - h)) / (2 * h)
        fd2 = (f(x + h) - 2 * f(x) + f(x - h)) / h**2
        assert relerr(d1, fd1) < 1e-7, (l, m, x, d1, fd1)
        assert relerr(d2.value, fd2) < 1e-3, (l, m, x, d2.value, fd2)


def test_K_closed_forms_perfect_reflector():
    """Perfect-reflector K_ν against the closed forms of the reference's
    disabled analytic branch (src/integration.c:305-360):
      m = 0:   K = 2e^{-α} + sqrt(2/(πα)) [(ν+1)(ν+2) K_{ν+1/2}(α) - 2α K_{ν+3/2}(α)]
      m = 1:   K = α^{3/2} sqrt(2/π) K_{ν+1/2}(α) - e^{-α}(α + ν(ν+1)/2)
      ν = 2m:  log K = -log(π)/2 + (m-1/2) log(2/α) + log (m-1)! + log (4m-1)!! + log K_{m-1/2}(α)
    """
    L = po.lib()
    lKh = lambda n, x: L.orc_bessel_logKn_half(n, x)
    for xi_ in [0.5, 3.0, 20.0]:
        a = 2 * xi_
        c = po.caps(1.0, 0.1, ldim=30)
        B0 = po.Block(c, xi_, 0)
        for nu in [2, 5, 17, 40]:
            exact = 2 * math.exp(-a) + math.sqrt(2 / (math.pi * a)) * (
                (nu + 1) * (nu + 2) * math.exp(lKh(nu, a)) - 2 * a * math.exp(lKh(nu + 1, a)))
            v, s = B0.K(nu, 1)
            assert s == 1 and relerr(math.exp(v), exact) < 1e-11, (xi_, nu, math.exp(v), exact)
        B1 = po.Block(c, xi_, 1)
        for nu in [2, 3, 8, 25]:
            exact = a**1.5 * math.sqrt(2 / math.pi) * math.exp(lKh(nu, a)) - math.exp(-a) * (a + nu * (nu + 1) / 2)
            v, s = B1.K(nu, 0)
            assert s == -1 and relerr(math.exp(v), exact) < 1e-10, (xi_, nu, math.exp(v), exact)
        for m in [2, 5, 9]:
            Bm = po.Block(c, xi_, m)
            logK = (-0.5 * math.log(math.pi) + (m - 0.5) * math.log(2 / a) + math.lgamma(m)
                    + L.orc_lfac2(4 * m - 1) + lKh(m - 1, a))
            v, s = Bm.K(2 * m, 1)
            assert abs(v - logK) < 1e-11 * max(1, abs(logK)), (xi_, m, v, logK)


def test_logdet_lu_closed_forms():
    """Dense LU log|det| against diagonal, 2x2 and numpy slogdet; the HODLR-mode
    and LU-mode shortcuts of log det(1-M)."""
    rng = np.random.default_rng(0)
    d = rng.uniform(0.1, 3, 50)
    A = np.diag(d)
    assert abs(po.lib().orc_logdet_lu(np.ascontiguousarray(A).ctypes.data_as(
        __import__("ctypes").POINTER(__import__("ctypes").c_double)), 50) - np.log(d).sum()) < 1e-13
    for n in [2, 7, 64, 130]:
        M = rng.normal(size=(n, n)) * 0.3 / math.sqrt(n)
        M = (M + M.T) / 2
        sign, ref = np.linalg.slogdet(np.eye(n) - M)
        assert sign > 0
        assert abs(po.logdet_1mM(M, 0) - ref) < 1e-12 * max(1, abs(ref))
    # tiny matrix: trace shortcut
    M = np.diag([1e-10, 2e-10])
    assert po.logdet_1mM(M, 0) == -3e-10
    # LU mode: Mercator shortcut for ||M||_F small
    M = np.diag([1e-5, 2e-5]) + 1e-6
    tr = np.trace(M)
    tr2 = np.trace(M @ M)
    assert abs(po.logdet_1mM(M, 1) - (-tr - tr2 / 2)) < 1e-20


def test_ht_drude_limit():
    """ξ = 0 Drude term: finite, negative, and approaching the PFA limit
    -ζ(3) R/(8L) ... as R/L grows (ratio -> 1)."""
    L = po.lib()
    for x in [1.0, 0.1, 0.01, 0.001]:
        v = L.orc_ht_drude(x)
        pfa = -1.2020569031595942 / (8 * x)
        assert v < 0
        if x <= 0.01:
            assert abs(v / pfa - 1) < 5 * x ** 0.5, (x, v, pfa)


# The R/L = 100, ldim = 500 block values of the reference's test_logdetD
# (N = 1000); the R/L = 1000, ldim = 10000 values (N = 20000) are checked by
# the GPU tests.
LOGDET_ROWS = [r for r in load_golden("logdetD.json") if r["ldim"] == 500]


@pytest.mark.parametrize("row", LOGDET_ROWS[::4], ids=lambda r: "xi%g_m%d" % (r["xi_"], r["m"]))
def test_logdetD_golden(row):
    c = po.caps(row["R"], row["L"], ldim=row["ldim"])
    v = po.logdetD(c, row["xi_"], row["m"])
    assert relerr(v, row["expect"]) < row["tol"], (v, row["expect"])


# ξ = 0 blocks (reference test_logdetD0): the plasma-model free energies at
# R/L = 100, ldim = 500; the m = 0 EE/MM values at ldim = 12000 are checked
# by the GPU tests (a 12000 x 12000 LU is too long for the CPU oracle).
XI0_ROWS = load_golden("logdetD0.json")


@pytest.mark.parametrize("row", [r for r in XI0_ROWS if r["fn"] == "ht_plasma"],
                         ids=lambda r: "omegap%.3g" % r["omegap"])
def test_ht_plasma_golden(row):
    import ctypes
    c = po.caps(row["R"], row["L"], ldim=row["ldim"], epsrel=1e-12)
    v = po.lib().orc_ht_plasma(ctypes.byref(c), row["omegap"], row["eps"])
    assert relerr(v, row["expect"]) < row["tol"], (v, row["expect"])


def test_xi0_blocks_closed_forms():
    """ξ = 0 kernels against closed forms at R/L = 10 (ldim = 160, where the
    window has converged): the m = 0 MM block equals twice Bimonte's series
    Σ_l ½ log(1 - Z^{2l+3}) (the m = 0 term of caps_ht_perf, which is halved
    in the m-sum), and Σ'_m of the EE blocks equals the closed-form Drude
    term (Bimonte-Emig)."""
    import ctypes
    L = po.lib()
    c = po.caps(10e-6, 1e-6, ldim=160, epsrel=1e-12)
    Z = 1 / (1 + c.LbyR + math.sqrt(c.LbyR * (2 + c.LbyR)))
    series = sum(0.5 * math.log1p(-Z ** (2 * l + 3)) for l in range(2000))
    EE, MM = ctypes.c_double(), ctypes.c_double()
    L.orc_logdetD0(ctypes.byref(c), 0, 0.0, ctypes.byref(EE), ctypes.byref(MM), None)
    assert abs(MM.value / (2 * series) - 1) < 1e-12, (MM.value, 2 * series)
    s = 0.0
    for m in range(200):
        L.orc_logdetD0(ctypes.byref(c), m, 0.0, ctypes.byref(EE), None, None)
        s += EE.value / 2 if m == 0 else EE.value
        if m > 0 and abs(EE.value / s) < 1e-15:
            break
    drude = L.orc_ht_drude(c.LbyR)
    assert abs(s / drude - 1) < 1e-12, (s, drude)


def test_plasma_integral_limits():
    """ξ = 0 plasma integral (1/ν!) ∫ -r_TE e^{-t} t^ν dt: tends to
    1 - (ν+1)/ω as ω -> ∞ (-r_TE ≈ 1 - t/ω) and to ω²/(ν(ν-1)) as ω -> 0
    (-r_TE ≈ ω²/t²), ω = ωp (L+R)/c."""
    L = po.lib()
    for nu in (2, 7, 40, 300):
        big = L.orc_integrate_plasma(nu, 1e9, 1e-13)
        assert abs(big - (1 - (nu + 1) / 1e9)) < 1e-12, (nu, big)
        # the next order is O(ω) relative at ν = 2 (t < ω, where -r_TE ≈ 1),
        # O(ω²/ν²) above
        w = 1e-8 if nu == 2 else 1e-4
        small = L.orc_integrate_plasma(nu, w, 1e-13)
        assert abs(small / (w * w / (nu * (nu - 1))) - 1) < 1e-6, (nu, small)


def test_logdet_delta_form_vs_mercator():
    """The oracle's delta-form determinant (no pivot search on u = 1 + d,
    log1p(d)) of near-identity 1 - M against the Mercator series
    -Σ tr(M^k)/k of M itself (closed form, never forming 1 - M), down to
    |log det| ~ 1e-12; a plain LU of 1 - M is shown to lose the small ones
    (the reason for the delta form).  A physical block (R/L = 20, m = 24)
    too."""
    from conftest import mercator_logdet
    rng = np.random.default_rng(4)
    for N in (5, 40, 200):
        for scale in (1e-1, 1e-3, 1e-6, 1e-10):
            X = rng.normal(size=(N, N))
            M = scale * (X + X.T) / (2 * N) + np.diag(np.full(N, scale / N))
            assert po.dominant_1mM(M)
            ref = mercator_logdet(M)
            assert abs(po.logdet_1mM_delta(M) / ref - 1) < 1e-14, (N, scale)
    c = po.caps(20e-6, 1e-6, ldim=120, epsrel=1e-12)
    M = po.Block(c, 1.0, 24).matrix()
    ref = mercator_logdet(M)
    assert abs(po.logdet_1mM_delta(M) / ref - 1) < 1e-14
    plain = np.linalg.slogdet(np.eye(M.shape[0]) - M)[1]
    assert abs(plain / ref - 1) > 1e-10        # the plain LU's rounding of 1 - M_ii


def test_dominance_test_cases():
    """orc_dominant_1mM: strict column dominance of 1 - M (row-major M),
    Σ_{i≠j} |M_ij| < |1 - M_jj|, checked on hand-made cases."""
    M = np.array([[0.2, 0.3], [0.4, 0.1]])
    assert po.dominant_1mM(M)                      # columns: 0.4 < 0.8, 0.3 < 0.9
    M = np.array([[0.2, 0.95], [0.4, 0.1]])
    assert not po.dominant_1mM(M)                  # column 1: 0.95 >= 0.9
    M = np.array([[1.5, 0.1], [0.1, 0.0]])
    assert po.dominant_1mM(M)                      # |1 - 1.5| = 0.5 > 0.1
    M = np.array([[0.0, 0.5], [0.5, 0.5]])
    assert not po.dominant_1mM(M)                  # column 1: 0.5 = |1 - 0.5|, not strict


def test_I_abs_bounds_I():
    """log Σ_q |a_q K| >= log |I| for every pair, with equality where the
    terms have one sign; the condition number of the far-off-diagonal pairs
    of a Drude block at R/L = 200 is large (up to ~1e6)."""
    c = po.caps(200e-6, 1e-6, ldim=200, epsrel=1e-10, drude=(9 / po.HBAR_EV, 0.035 / po.HBAR_EV))
    B = po.Block(c, 1489.1072047882992, 10)
    conds = []
    for l1, l2 in [(10, 10), (60, 10), (150, 10), (209, 10), (120, 100), (209, 209)]:
        for p in (0, 1):
            v, _ = B.I(l1, l2, p)
            a = B.I_abs(l1, l2, p)
            assert a >= v - 1e-12, (l1, l2, p, a, v)
            conds.append(math.exp(a - v))
    assert min(conds) < 1 + 1e-12 and max(conds) > 10, conds


def test_manual_runs_fixture():
    """The transcribed manual runs: one T = 0 run, four T = 300 K runs (18,
    17, 18, 18 printed frequencies starting at ξ = 0) and the high-temperature
    limit; frequencies on the Matsubara grid 2π k_B T (L+R)/(ħc)."""
    import json
    import os
    from conftest import GOLDEN
    runs = json.load(open(os.path.join(GOLDEN, "manual_runs.json")))["rows"]
    assert len(runs) == 6
    T300 = [r for r in runs if r["header"].get("T") == 300.0]
    assert [len(r["xi"]) for r in T300] == [18, 17, 18, 18]
    for r in T300:
        Ts = 2 * math.pi * 1.38064852e-23 * 50.5e-6 * 300 / (1.0545718e-34 * 299792458.0)
        assert r["xi"][0]["xi_"] == 0
        for k, row in enumerate(r["xi"][1:], 1):
            assert abs(row["xi_"] / (k * Ts) - 1) < 1e-13


def _tables():
    rows = []
    for name in ("ref_tables.json", "ref_tables_c2.json"):
        path = os.path.join(GOLDEN, name)
        if os.path.exists(path):
            with open(path) as f:
                rows += json.load(f)["data"]
    return rows


@pytest.mark.parametrize("case", _tables(), ids=lambda c: "%s_ldim%d_xi%g_m%d" % (c["kind"], c["ldim"], c["xi_"], c["m"]))
def test_oracle_K_I_M_vs_reference_tables(case):
    """The oracle's K_ν, I_{l1 l2} and matrix elements against the tables the
    reference compiled from its sources wrote (oracle/build_ref.sh,
    tests/golden/gen_ref_vectors.py), with the tolerances of the device test:
    K 1e-9 relative; I 1e-9 times the q-sum's condition Σ_q|a_q K| / |I|
    (orc_block_I_abs); elements 1e-9 relative plus 1e-14 sqrt|M_ii M_jj|."""
    import pyoracle as po
    oc = po.caps(case["R"], case["L"], ldim=case["ldim"], epsrel=1e-12,
                 drude=None if case["omegap"] is None else (case["omegap"], case["gamma"]))
    B = po.Block(oc, case["xi_"], case["m"])
    for nu, p, v, s in case["K"]:
        g, gs = B.K(nu, p)
        assert gs == s and abs(math.exp(g - v) - 1) < 1e-9, ("K", nu, p, g, v, gs, s)
    for l1, l2, p, v, s in case["I"]:
        g, gs = B.I(l1, l2, p)
        cond = max(1.0, math.exp(B.I_abs(l1, l2, p) - v))
        assert gs == s and abs(math.exp(g - v) - 1) < 1e-9 * cond, ("I", l1, l2, p, g, v, cond)
    assert B.lmin == case["lmin"]
    for l1, l2, p1, p2, v in case["M"]:
        g = B.elem(l1, l2, p1, p2)
        scale = math.sqrt(abs(B.elem(l1, l1, p1, p1) * B.elem(l2, l2, p2, p2)))
        assert abs(g - v) <= 1e-9 * abs(v) + 1e-14 * scale + 1e-300, ("M", l1, l2, p1, p2, g, v, scale)
