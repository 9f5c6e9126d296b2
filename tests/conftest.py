import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long CPU test")


def load_golden(name):
    import json
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)["rows"]


def relerr(x, y):
    if x == y:
        return 0.0
    return abs(1 - x / y)


# Per-block parity criteria against the reference's caps_logdetD values.
#  * trace-shortcut blocks (|tr M| < 1e-8; both sides return -tr M from
#    elements that agree to ~1e-11): relative 1e-9;
#  * factored blocks: |Δ| <= 1e-9 |v_ref| + 2 max(1e-13, 1e-13 |tr M|).
#    max(1e-13, 1e-13 tr M) is the reference's HODLR tolerance
#    (src/matrix.c:135), which bounds the low-rank approximation of each
#    off-diagonal block, not the determinant: against the exact log det of
#    the same matrix (its Mercator series) the reference deviates by up to
#    1.47 times it on these fixtures (mixed-aspect block 0: 1.47e-13 at
#    |logdetD| = 4.4e-8, where the GPU value equals the series to 2e-16), so
#    the factor 2.  The GPU's own error is checked without this term against
#    the oracle and the Mercator series (1e-10 and 1e-12 relative).
def block_tolerance(ref, trace, shortcut):
    if shortcut:
        return 1e-9 * abs(ref)
    return 1e-9 * abs(ref) + 2 * max(1e-13, 1e-13 * abs(trace))


def check_blocks(values, traces, flags, refs, label=""):
    """Assert every block against block_tolerance; return a summary (counts
    and largest errors per kind) to print with the test."""
    bad, worst_short, worst_lu, worst_frac = [], 0.0, 0.0, 0.0
    nshort = nlu = 0
    for i, (v, tr, f, ref) in enumerate(zip(values, traces, flags, refs)):
        shortcut = (int(f) & 3) == 1
        tol = block_tolerance(ref, tr, shortcut)
        err = abs(v - ref)
        rel = err / abs(ref) if ref != 0 else err
        if shortcut:
            nshort += 1
            worst_short = max(worst_short, rel)
            if v != -tr:
                bad.append((i, "shortcut value is not -tr", v, tr))
        else:
            nlu += 1
            worst_lu = max(worst_lu, rel)
        worst_frac = max(worst_frac, err / tol if tol > 0 else float(err > 0))
        if not err <= tol:
            bad.append((i, v, ref, tr, int(f)))
    summary = ("%s: %d trace-shortcut blocks (max rel %.1e), %d factored (max rel %.1e), "
               "max |err|/tol %.3f" % (label, nshort, worst_short, nlu, worst_lu, worst_frac))
    assert not bad, (summary, bad[:5])
    return summary


def mercator_logdet(M, tol=1e-16):
    """log det(1 - M) = -Σ_k tr(M^k)/k for ||M||_F < 1/2, summed until the
    tail bound ||M||_F^(k+1)/(1 - ||M||_F) falls below tol |sum|: no entry
    near 1 is ever formed, so the result carries the relative accuracy of M
    itself (a closed form to check a determinant of a near-identity 1 - M)."""
    import math
    import numpy as np
    nrm = float(np.linalg.norm(M))
    assert nrm < 0.5, nrm
    terms, P, k = [], M, 1
    while True:
        terms.append(float(np.trace(P)) / k)
        s = abs(math.fsum(terms))
        if nrm ** (k + 1) / (1 - nrm) <= tol * s or k > 200:
            return -math.fsum(terms)
        P = P @ M
        k += 1
