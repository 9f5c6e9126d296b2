"""CPU-side checks of the drop-in boundary: libcaps_amd.so loads and exports
every function include/libcaps.h declares (no compute calls: no GPU here)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def declared_functions():
    text = open(os.path.join(ROOT, "include", "libcaps.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b(\w+)\s*\([^;{]*\)\s*;", text, flags=re.M)
    return sorted(set(n for n in names if not n.startswith("__")))


def test_header_declares_reference_api():
    names = declared_functions()
    for n in ["caps_init", "caps_free", "caps_logdetD", "caps_M_init", "caps_M_elem", "caps_M_free",
              "caps_kernel_M", "kernel_logdet", "caps_set_ldim", "caps_set_epsilonm1", "caps_mie",
              "caps_fresnel", "caps_lnLambda", "caps_estimate_lminmax", "caps_ht_drude",
              "caps_logdetD_batch", "caps_free_energy"]:
        assert n in names, n


def test_library_exports_every_declared_symbol():
    import caps_amd
    L = ctypes.CDLL(caps_amd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_caps_struct_layout_matches_reference():
    """Public fields of caps_t sit at the reference's offsets
    (src/include/libcaps.h:41-78): doubles L, R, calL, LbyR, y, then the two
    callback/userdata pairs, ldim, epsrel, detalg."""
    import caps_amd
    L = caps_amd.lib()
    p = L.caps_init(2.0, 0.5)
    d = ctypes.cast(p, ctypes.POINTER(ctypes.c_double))
    assert (d[0], d[1], d[2], d[3]) == (0.5, 2.0, 2.5, 0.25)
    ldim = ctypes.cast(p + 9 * 8, ctypes.POINTER(ctypes.c_int))[0]
    assert ldim == 20
    epsrel = ctypes.cast(p + 10 * 8, ctypes.POINTER(ctypes.c_double))[0]
    assert epsrel == 1e-8
    L.caps_set_ldim(p, 77)
    assert ctypes.cast(p + 9 * 8, ctypes.POINTER(ctypes.c_int))[0] == 77
    assert L.caps_set_ldim(p, 0) == 0 and L.caps_set_epsrel(p, -1.0) == 0
    L.caps_free(p)
    assert L.caps_init(-1.0, 1.0) is None


def test_host_only_entry_points():
    import caps_amd
    L = caps_amd.lib()
    import pyoracle as po
    for l1, l2, m in [(3, 5, 1), (100, 120, 40), (2000, 2100, 5)]:
        assert abs(L.caps_lnLambda(l1, l2, m) - po.lib().orc_lnLambda(l1, l2, m)) < 1e-12 * max(1, abs(L.caps_lnLambda(l1, l2, m)))
    c = caps_amd.Caps(200e-6, 1e-6, ldim=1200)
    assert c.lminmax(0) == (1, 1201)
    assert c.lminmax(100) == (400, 1600)
    assert c.lminmax(200) == (1400, 2600)
    assert abs(c.ht_drude() - po.lib().orc_ht_drude(1e-6 / 200e-6)) < 1e-14 * abs(c.ht_drude())


def test_plasma_integrals_host():
    """The library's ξ = 0 plasma integrals (trapezoid rule in log t, host
    code) and Bessel ratios (Lentz continued fraction) against the oracle's
    adaptive quadrature and the reference-style continued fraction."""
    import caps_amd
    import pyoracle as po
    L = po.lib()
    R, Lx = 100e-6, 1e-6
    c = caps_amd.Caps(R, Lx, ldim=500)
    for w in (0.01, 1.0, 100.0, 1000.0):
        omegap = w * po.CAPS_C / (R + Lx)
        alpha = w / (1 + Lx / R)
        for l1, l2 in [(1, 1), (3, 7), (50, 60), (300, 500), (1000, 1200)]:
            I, r1, r2 = c.plasma_integrals(omegap, l1, l2)
            assert abs(I / L.orc_integrate_plasma(l1 + l2, w, 1e-13) - 1) < 1e-12, (w, l1, l2)
            assert abs(r1 / L.orc_bessel_ratioI(l1 - 0.5, alpha) - 1) < 1e-13, (w, l1)
            assert abs(r2 / L.orc_bessel_ratioI(l2 - 0.5, alpha) - 1) < 1e-13, (w, l2)


def test_npy_files(tmp_path):
    """Matrix files in the reference's layout (src/matrix.c:211-335): the
    header bytes, a round trip, and numpy's reading of the same bytes (the
    header says fortran order, so numpy returns the transpose); files of
    another layout are refused."""
    import numpy as np
    import caps_amd
    rng = np.random.default_rng(3)
    M = rng.normal(size=(7, 7))
    p = tmp_path / "m.npy"
    caps_amd.npy_save(p, M)
    raw = p.read_bytes()
    hdr = b"{'descr': '<f8', 'fortran_order': True, 'shape': (7, 7), }"
    assert raw[:8] == b"\x93NUMPY\x01\x00" and raw[8:10] == len(hdr).to_bytes(2, "little")
    assert raw[10:10 + len(hdr)] == hdr and len(raw) == 10 + len(hdr) + 8 * 49
    assert np.array_equal(caps_amd.npy_load(p), M)
    assert np.array_equal(np.load(p, allow_pickle=False), M.T)
    q = tmp_path / "c.npy"
    np.save(q, np.ones((3, 3)))   # C order: not the reference's layout
    with pytest.raises(ValueError):
        caps_amd.npy_load(q)


def test_psd_poles_and_residues():
    """Padé spectrum decomposition (host code): the poles are -2/w of the
    negative eigenvalues w of the tridiagonal matrix (numpy's eigvalsh as
    the independent solver), the lowest ones sit at the Matsubara
    frequencies 2πj with residue 1, and the expansion reproduces the Bose
    function 1/(1 - e^{-x}) = 1/x + 1/2 + Σ 2η x/(x² + ξ²) to rounding for
    x well inside the spectrum."""
    import math
    import numpy as np
    import caps_amd
    for N in (5, 20, 60):
        xi, eta = caps_amd.psd(N)
        n = 2 * N
        T = np.zeros((n, n))
        for j in range(1, n):
            T[j - 1, j] = T[j, j - 1] = 1 / math.sqrt((2 * j + 1) * (2 * j + 3))
        w = np.sort(np.linalg.eigvalsh(T))[:N]
        assert np.max(np.abs(xi / (-2 / w) - 1)) < 1e-13
        assert np.all(np.diff(xi) > 0)
        if N >= 20:
            assert np.max(np.abs(xi[:3] / (2 * math.pi * np.arange(1, 4)) - 1)) < 1e-13
            assert np.max(np.abs(eta[:3] - 1)) < 1e-13
            for x in (0.3, 2.0, 8.0, 20.0):
                bose = 1 / (-math.expm1(-x))
                assert abs((1 / x + 0.5 + np.sum(2 * eta * x / (x * x + xi * xi))) / bose - 1) < 1e-14


def _write_drude_table(path, xis, wp, gm, wp_hi=None, gm_hi=None):
    hev = 6.582119514e-16
    with open(path, "w") as f:
        f.write("# synthetic Drude table\n#\n# omegap_low = %.17geV\n# gamma_low = %.17geV\n" % (wp, gm))
        f.write("# omegap_high = %.17geV\n# gamma_high = %.17geV\n#\n# xi in rad/s\tepsilon(i*xi)\n" % (
            wp if wp_hi is None else wp_hi, gm if gm_hi is None else gm_hi))
        for x in xis:
            W, G = wp / hev, gm / hev
            f.write("%.17g\t%.17g\n" % (x, 1 + W * W / (x * (x + G))))


def test_material_table(tmp_path):
    """Tabulated ε(iξ) (host code, the reference's material file format):
    tabulated values at the nodes, linear interpolation between them, Drude
    extrapolation with the header parameters outside; a table whose ξ do
    not ascend is refused."""
    import numpy as np
    import caps_amd
    hev = 6.582119514e-16
    xis = np.logspace(12, 16, 200)
    p = tmp_path / "drude.csv"
    _write_drude_table(p, xis, 9.0, 0.035, 50.0, 200.0)
    c = caps_amd.Caps(200e-6, 1e-6, ldim=100)
    m = c.set_material(p)
    W, G = 9.0 / hev, 0.035 / hev
    drude = lambda x, w, g: w * w / (x * (x + g))
    for k in (0, 17, 199):
        assert abs(m.epsilonm1(xis[k]) / drude(xis[k], W, G) - 1) < 1e-14
    x = 0.3 * xis[40] + 0.7 * xis[41]
    lin = drude(xis[40], W, G) + (x - xis[40]) * (drude(xis[41], W, G) - drude(xis[40], W, G)) / (xis[41] - xis[40])
    assert abs(m.epsilonm1(x) / lin - 1) < 1e-13
    assert abs(m.epsilonm1(1e10) / drude(1e10, W, G) - 1) < 1e-14
    assert abs(m.epsilonm1(1e18) / drude(1e18, 50.0 / hev, 200.0 / hev) - 1) < 1e-14
    ext = m.extrapolation()
    assert np.allclose(ext, (W, G, 50.0 / hev, 200.0 / hev), rtol=1e-14)
    q = tmp_path / "bad.csv"
    _write_drude_table(q, xis[::-1], 9.0, 0.035)
    with pytest.raises(ValueError):
        c.set_material(q)
