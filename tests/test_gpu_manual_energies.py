"""The driver's sums against the free energies the reference's manual prints
(docs/manual/index.rst, transcribed to tests/golden/manual_runs.json): sphere
R = 50 µm, plate at L = 500 nm, ldim = 701, T = 300 K for perfect reflectors,
the plasma model, the Drude model and gold (optical data,
tests/golden/gold.csv), T = 0 for perfect reflectors, and the
high-temperature limit at R/L = 1000 (ldim = 7001).

Each printed per-frequency value is the driver's F_xi (src/caps.c:307-396):
Σ'_m logdetD(ξ, m), m = 0 halved, stopped at the first m with v/v_0 <
cutoff.  The manual's runs used 8 MPI processes (7 workers): when the master
sees a block below the cutoff, the blocks still in flight are summed too, up
to 6 further m (src/caps.c:353-376).  Those terms are all below the cutoff
and of one sign, so each printed value must lie between the sequential
m-sum (the library's semantics) and the sum extended by 7 terms, up to the
blocks' own errors (the reference's HODLR tolerance, max(1e-13, 1e-13 tr M)
per block, plus 1e-11 relative).  The free energy F = T_scaled/π Σ'_n F_xi
is bracketed the same way, and the library's capsx_free_energy must equal
the sequential bracket end.
"""
import ctypes
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

HBAR_EV = 6.582119514e-16
EXTRA = 7      # workers of the manual's 8-process runs: at most 6 more terms in flight

with open(os.path.join(GOLDEN, "manual_runs.json")) as f:
    RUNS = json.load(f)["rows"]


def run_of(src_line):
    return [r for r in RUNS if r["src"].endswith(":%d" % src_line)][0]


@pytest.fixture(scope="module")
def ca():
    import caps_amd
    caps_amd.lib()
    return caps_amd


def caps_for(ca, run):
    o = run["options"]
    R, L = float(o["R"]), float(o["L"])
    c = ca.Caps(R, L, ldim=int(run["header"]["ldim"]))
    if "material" in o:
        c.set_material(os.path.join(GOLDEN, "gold.csv"))
    elif "omegap" in o:
        c.set_drude(float(o["omegap"]), float(o.get("gamma", 0.0)))
    return c, R, L


def kahan(v):
    s = c = 0.0
    for t in v:
        y = t - c
        u = s + y
        c = (u - s) - y
        s = u
    return s


def bracket_sums(vals, cutoff):
    """(sequential m-sum, sum extended by EXTRA terms, m-cutoff) of the
    values v_0, v_1, ... (enough of them)."""
    for m, v in enumerate(vals):
        if v == 0 or v / vals[0] < cutoff:
            t = list(vals[:m + 1])
            t[0] /= 2
            seq = kahan(t)
            ext = kahan(t + list(vals[m + 1:m + 1 + EXTRA]))
            assert len(vals) >= m + 1 + EXTRA
            return seq, ext, m
    raise AssertionError("cutoff not reached")


def xi_sums(c, xi_, cutoff, m0=96):
    """F_xi brackets at ξ > 0 from one batch, grown until the cutoff and the
    extra terms are covered."""
    n = m0
    while True:
        vals = c.logdetD_batch(np.full(n, xi_), np.arange(n))
        try:
            return bracket_sums(list(vals), cutoff)
        except AssertionError:
            n *= 2
            assert n < 20000


def xi0_sums(ca, c, cutoff, model, omegap=0.0):
    """F_xi at ξ = 0 (src/caps.c:323-335, 392-393): the Drude term plus
    Σ'_m of the MM blocks of perfect reflectors (model "pr") or the plasma
    model (omegap in rad/s)."""
    L = ca.lib()
    vals, m = [], 0
    while True:
        x = ctypes.c_double()
        if model == "pr":
            L.caps_logdetD0(c.ptr, m, 0.0, None, ctypes.byref(x), None)
        else:
            L.caps_logdetD0(c.ptr, m, omegap, None, None, ctypes.byref(x))
        vals.append(x.value)
        m += 1
        try:
            seq, ext, mc = bracket_sums(vals, cutoff)
            break
        except AssertionError:
            continue
    drude = c.ht_drude()
    return drude + seq, drude + ext, mc


def within(value, seq, ext, nblocks):
    lo, hi = min(seq, ext), max(seq, ext)
    slack = 1e-11 * abs(value) + 1e-13 * nblocks
    return lo - slack <= value <= hi + slack


def check_T300(ca, src_line, xi0_model):
    run = run_of(src_line)
    c, R, L = caps_for(ca, run)
    cutoff, epsrel = run["header"]["cutoff"], run["header"]["epsrel"]
    T = run["header"]["T"]
    T_scaled = ca.T_scaled(R, L, T)
    seqs, exts, report = [], [], []
    for k, row in enumerate(run["xi"]):
        xi_ = row["xi_"]
        if k == 0:
            assert xi_ == 0
            if xi0_model == "drude":
                v = c.ht_drude()
                assert abs(v / row["value"] - 1) < 1e-13, (v, row["value"])
                seq = ext = v
                mc = 0
            else:
                wp = float(run["options"]["omegap"]) / HBAR_EV if xi0_model == "plasma" else 0.0
                seq, ext, mc = xi0_sums(ca, c, cutoff, xi0_model, wp)
        else:
            assert abs(xi_ / (k * T_scaled) - 1) < 1e-14, (xi_, k * T_scaled)
            seq, ext, mc = xi_sums(c, xi_, cutoff)
        assert within(row["value"], seq, ext, mc + EXTRA), (src_line, k, xi_, row["value"], seq, ext)
        report.append(abs(row["value"] / seq - 1))
        seqs.append(seq)
        exts.append(ext)
    # the Matsubara sum stops at the same frequency
    assert abs(seqs[-1] / seqs[0]) < epsrel <= abs(seqs[-2] / seqs[0])
    F_seq = T_scaled / math.pi * kahan([seqs[0] / 2] + seqs[1:])
    F_ext = T_scaled / math.pi * kahan([exts[0] / 2] + exts[1:])
    F_man = run["final"][-1]
    wp = float(run["options"].get("omegap", 0)) / HBAR_EV
    F_lib, computed, used = c.free_energy_ex(T, cutoff=cutoff, epsrel=epsrel,
                                             xi0={"drude": "drude", "pr": "pr", "plasma": "plasma"}[xi0_model],
                                             omegap=wp)
    print("%s: F manual %.16g, library %.16g (rel %.1e), bracket [%.16g, %.16g]; per-frequency sums "
          "within their brackets, largest |printed/sequential - 1| %.1e; %d blocks used (manual: %d computed)" % (
              run["src"], F_man, F_lib, abs(F_lib / F_man - 1), F_seq, F_ext, max(report), used,
              run["determinants"]))
    assert abs(F_lib - F_seq) <= 1e-13 * abs(F_seq), (F_lib, F_seq)
    assert within(F_man, F_seq, F_ext, len(seqs) * 100), (F_man, F_seq, F_ext)
    assert abs(F_lib / F_man - 1) < 1e-8
    return c


def test_manual_T300_perfect_reflectors(ca):
    check_T300(ca, 412, "pr")


def test_manual_T300_plasma(ca):
    check_T300(ca, 561, "plasma")


def test_manual_T300_drude(ca):
    check_T300(ca, 619, "drude")


def test_manual_T300_gold(ca):
    """Gold from tabulated optical data; the ξ = 0 term is the Drude one
    (the table's low-frequency Drude extrapolation has γ > 0); the printed
    plasma-model ξ = 0 term of the same run is checked too."""
    run = run_of(712)
    c = check_T300(ca, 712, "drude")
    mt = c._materials[-1]
    omegap_low = mt.extrapolation()[0]
    seq, ext, mc = xi0_sums(ca, c, run["header"]["cutoff"], "plasma", omegap_low)
    assert within(run["plasma_xi0"], seq, ext, mc + EXTRA), (run["plasma_xi0"], seq, ext)


def test_manual_T0_perfect_reflectors(ca):
    """T = 0: the printed integrand values at the frequencies the reference's
    dqagi sampled, and F = ∫ F_xi dx/(απ) (the manual's run: epsrel 1e-6,
    its own error estimate 4.2e-7 relative) against the library's
    Fourier-Chebyshev integral at epsrel 1e-8."""
    run = run_of(314)
    c, R, L = caps_for(ca, run)
    cutoff = run["header"]["cutoff"]
    for row in run["xi"]:
        seq, ext, mc = xi_sums(c, row["xi_"], cutoff)
        assert within(row["value"], seq, ext, mc + EXTRA), (row, seq, ext)
    F, computed, used = c.free_energy(0.0, cutoff=cutoff, epsrel=1e-8, mode="batched")
    F_man = run["final"][-1]
    print("T=0 R/L=100: F manual %.16g, library %.16g (rel %.1e), %d blocks (manual: %d)" % (
        F_man, F, abs(F / F_man - 1), used, run["determinants"]))
    assert abs(F / F_man - 1) < 1e-6


def test_manual_high_temperature_limit(ca):
    """R/L = 1000, ldim = 7001: the Drude ξ = 0 term (closed form) and the
    perfect-reflector one (Drude term + Σ'_m of the 7001 x 7001 MM blocks,
    filled and factored on the GPU)."""
    run = run_of(499)
    o = run["options"]
    c = ca.Caps(float(o["R"]), float(o["L"]), ldim=int(run["header"]["ldim"]))
    E_drude, E_pr = run["final"][-2], run["final"][-1]
    d = c.ht_drude()
    assert abs(d / E_drude - 1) < 1e-13, (d, E_drude)
    seq, ext, mc = xi0_sums(ca, c, run["header"]["cutoff"], "pr")
    print("high-temperature limit R/L=1000: PR %.16g, bracket [%.16g, %.16g] (m-cutoff %d)" % (E_pr, seq, ext, mc))
    assert within(E_pr, seq, ext, mc + EXTRA), (E_pr, seq, ext)
