"""The determinant's two forms on the GPU, through the C ABI.

Delta form: for a column diagonally dominant 1 - M the matrix holds -M (no
unit diagonal), the LU runs without pivot search on u = 1 + d and sums
log1p(d); partial pivoting would exchange no rows there, so it is the same
factorization, but no pivot near 1 is rounded to a double near 1.  Checked
against the Mercator series of M (conftest.mercator_logdet, which never
forms 1 - M) at 1e-13 relative.  Explicit form (1 - M not dominant): partial
pivoting, checked against numpy's slogdet.  kernel_logdet with a caller's
callback (src/capc.cpp:110-114) under every detalg, against the oracle and
the reference's rules (src/matrix.c:57-149, 489-544).
"""
import math

import numpy as np
import pytest

from conftest import mercator_logdet

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ca():
    import caps_amd
    caps_amd.lib()
    return caps_amd


def sym_matrix(rng, N, scale):
    X = rng.normal(size=(N, N))
    return scale * (X + X.T) / (2 * N)


@pytest.mark.parametrize("N", [7, 64, 130, 300, 700])
def test_delta_form_small_norm(ca, N):
    """Near-identity 1 - M (||M|| from 1e-2 down to 1e-9): every block runs in
    delta form and matches the Mercator series of M to 1e-13 relative; a
    plain LU of 1 - M would lose up to sqrt(N) eps / |log det| (5e-8
    relative at |log det| = 1e-9)."""
    rng = np.random.default_rng(N)
    mats = [sym_matrix(rng, N, s) + np.diag(np.full(N, s / N)) for s in (1e-2, 1e-4, 1e-6, 1e-9)]
    out, flags = ca.logdet_1mM_batch(np.array(mats))
    assert list(flags) == [1, 1, 1, 1], flags
    for M, v in zip(mats, out):
        ref = mercator_logdet(M)
        assert abs(v / ref - 1) < 1e-13, (N, v, ref)


def test_delta_form_physical_block(ca):
    """A round-trip block (R/L = 20, ξ = 1, m = 24, |log det| = 8e-7) from the
    oracle: delta form, against its Mercator series."""
    import pyoracle as po
    oc = po.caps(20e-6, 1e-6, ldim=120, epsrel=1e-12)
    M = po.Block(oc, 1.0, 24).matrix()
    out, flags = ca.logdet_1mM_batch(M[None])
    ref = mercator_logdet(M)
    assert flags[0] == 1 and abs(out[0] / ref - 1) < 1e-13, (out[0], ref)


@pytest.mark.parametrize("N", [50, 300, 517])
def test_explicit_form_when_not_dominant(ca, N):
    """Column sums of |M| above |1 - M_jj|: the unit diagonal is written back
    and the LU pivots (slogdet at 1e-10 N); a dominant matrix of the same
    batch stays in delta form."""
    rng = np.random.default_rng(100 + N)
    big = rng.normal(size=(N, N)) * (3.0 / math.sqrt(N))       # column sums ~ 2.4 sqrt(N)
    small = sym_matrix(rng, N, 0.3)
    out, flags = ca.logdet_1mM_batch(np.array([big, small]))
    assert list(flags) == [0, 1], flags
    for M, v in zip([big, small], out):
        ref = np.linalg.slogdet(np.eye(N) - M)[1]
        assert abs(v - ref) < 1e-10 * N * max(1, abs(ref)), (N, v, ref)


def test_lu_nan_entry_gives_nan(ca):
    """A NaN entry in a matrix of the pivoted LU: no candidate row wins a
    column whose entries are all NaN; the pivot falls back to the diagonal
    row (indices stay in range) and the determinant is NaN."""
    rng = np.random.default_rng(3)
    N = 200
    A = np.eye(N) + 0.1 * rng.normal(size=(N, N)) / math.sqrt(N)
    B = A.copy()
    B[:, 5] = np.nan
    out = ca.lu_logdet_batch(np.array([A, B]))
    assert abs(out[0] - np.linalg.slogdet(A)[1]) < 1e-10
    assert np.isnan(out[1]), out[1]


@pytest.mark.parametrize("detalg", ["HODLR", "LU", "QR", "CHOLESKY"])
def test_kernel_logdet_callback(ca, detalg):
    """kernel_logdet with a caller's callback, as the reference's cylinder
    program calls it (src/capc.cpp:110-114): a block-sized symmetric matrix
    (LU), a small one (LU-mode Mercator: -tr M - tr M²/2, taken only in the
    dense modes; CHOLESKY fills and tests the lower triangle only, as
    src/matrix.c:93-102 does) and a tiny one (trace shortcut)."""
    import pyoracle as po
    rng = np.random.default_rng(8)
    N = 160
    big = sym_matrix(rng, N, 4.0)                    # ||M||_F ~ 3, eigenvalues < 1/2: LU
    small = sym_matrix(rng, N, 1e-4) + np.diag(np.full(N, 1e-5))   # ||M||_F ~ 1.5e-4: Mercator
    tiny = sym_matrix(rng, N, 1e-9)
    for M in (big, small, tiny):
        v = ca.kernel_logdet(N, lambda i, j: float(M[i, j]), detalg)
        tr = math.fsum(np.diag(M))
        if abs(tr) < 1e-8:
            expect = -tr
        elif detalg == "HODLR":
            ld = po.logdet_1mM(M, 0)            # dense LU with the HODLR-mode rule
            expect = ld
        else:
            H = np.tril(M) if detalg == "CHOLESKY" else M
            nrm = np.linalg.norm(H)
            merc = -tr - float(np.sum(H * H.T)) / 2
            if nrm < 1 and abs((nrm ** 2 / 2 + nrm + math.log1p(-nrm)) / merc) < 1e-8:
                expect = merc
            else:
                expect = po.logdet_1mM_delta(M) if po.dominant_1mM(M) else np.linalg.slogdet(np.eye(N) - M)[1]
        assert abs(v - expect) <= 1e-12 * abs(expect), (detalg, v, expect)


def test_pivoted_fallback_on_round_trip_blocks(ca):
    """Round-trip blocks whose 1 - M is not column dominant are left out by
    the delta-form LU and factored again with the pivoted LU.  The test hook
    capsx_lu_delta_mode(0) reports every block as not dominant: all blocks
    of the R/L = 20 configuration and of the mixed-aspect batch then take
    that path, and agree with the delta-form values to the pivoted LU's
    rounding: each of the N pivots u_ii = fl(1 - M_ii - ...) near 1 and its
    log carry up to an ulp of 1, so 1e-12 relative plus N eps absolute."""
    from caps_amd.workloads import mixed_aspect_blocks, make_objects
    L = ca.lib()
    c = ca.Caps(20e-6, 1e-6, ldim=120)
    ms = np.arange(121)
    vd, _, fd = c.logdetD_batch(np.ones(121), ms, info=True)
    blocks = mixed_aspect_blocks()
    objs = make_objects(blocks)
    xs, mm = [b["xi_"] for b in blocks], [b["m"] for b in blocks]
    wd, _, gd = ca.logdetD_multi(objs, xs, mm, info=True)
    old = L.capsx_lu_delta_mode(0)
    try:
        vp, _, fp = c.logdetD_batch(np.ones(121), ms, info=True)
        wp, _, gp = ca.logdetD_multi(objs, xs, mm, info=True)
    finally:
        L.capsx_lu_delta_mode(old)
    for vals_d, vals_p, fl_d, fl_p, dims in ((vd, vp, fd, fp, [240] * 121),
                                              (wd, wp, gd, gp, [2 * b["ldim"] for b in blocks])):
        factored = [(f & 3) == 0 for f in fl_d]
        assert all(f & 4 for f, x in zip(fl_d, factored) if x)          # delta form by default
        assert not any(f & 4 for f, x in zip(fl_p, factored) if x)      # pivoted when forced
        for a, b, x, n in zip(vals_d, vals_p, factored, dims):
            if x:
                assert abs(a - b) <= 1e-12 * abs(a) + n * 2.2e-16, (a, b, n)
            else:
                assert a == b


def test_symmetric_form_matches_general(ca):
    """Round-trip blocks (1 - M symmetric) factored in the symmetric delta
    form (U12 = D11 L21^T, lower-triangle updates; the default) and in the
    general delta form (CAPS_AMD_LU_SYM=0) agree to the rounding of the
    factorization (N eps relative; the two forms round the symmetric pairs
    differently), for the R/L = 20 blocks and for a batch of blocks of
    unequal dimension.  The symmetric form leaves the mirror tiles outside
    the 64 x 64 diagonal blocks unassembled and takes their column sums for
    the dominance test from the lower tiles' row sums: the blocks' flags
    (which blocks took the delta form) agree too."""
    import os
    from caps_amd.workloads import mixed_aspect_blocks, make_objects
    c = ca.Caps(20e-6, 1e-6, ldim=120)
    ms = np.arange(121)
    blocks = mixed_aspect_blocks()[:24]
    objs = make_objects(blocks)
    xs, mm = [b["xi_"] for b in blocks], [b["m"] for b in blocks]
    res = {}
    try:
        for sym in ("1", "0"):
            os.environ["CAPS_AMD_LU_SYM"] = sym
            v, _, flags = ca.logdetD_multi(objs, xs, mm, info=True)
            res[sym] = (c.logdetD_batch(np.ones(121), ms), v, flags)
    finally:
        os.environ.pop("CAPS_AMD_LU_SYM", None)
    assert np.array_equal(res["1"][2], res["0"][2]), (res["1"][2], res["0"][2])
    worst = 0.0
    for a, b, N in ((res["1"][0], res["0"][0], 240), (res["1"][1], res["0"][1], 2 * max(b_["ldim"] for b_ in blocks))):
        rel = np.abs(a - b) / np.abs(b)
        worst = max(worst, rel.max())
        assert rel.max() < N * 2.3e-16 * 8, (rel.max(), N)
    print("symmetric vs general delta form: largest relative difference %.2e" % worst)
