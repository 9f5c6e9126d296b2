"""Resumed free energies (the reference driver's --resume, src/caps.c:72-101,
265, 316-327): a run writes one line per frequency; a second run that reads
part of that output computes only the frequencies it lacks and gives the same
free energy; one that reads all of it computes no blocks."""
import pytest

pytestmark = pytest.mark.gpu

R, L, LDIM = 20e-6, 1e-6, 60


@pytest.fixture(scope="module")
def ca():
    import caps_amd
    caps_amd.lib()
    return caps_amd


def _run(ca, T, out=None, resume=None, mode="batched"):
    c = ca.Caps(R, L, ldim=LDIM)
    if out:
        c.set_output(out)
    if resume:
        c.resume_load(resume)
    F, nb, nu = c.free_energy_ex(T, xi0="pr", mode=mode) if T > 0 else c.free_energy(0, mode=mode)
    if out:
        c.set_output(None)
    return F, nb, nu


@pytest.mark.parametrize("T", [300.0, 0.0])
def test_resume_partial_and_full(ca, tmp_path, T):
    full = tmp_path / "full.out"
    F1, nb1, _ = _run(ca, T, out=str(full))
    lines = [ln for ln in full.read_text().splitlines() if ln.startswith("# xi*(L+R)/c=")]
    assert len(lines) >= 4
    # every frequency's line: the value the driver used (printed to 16 digits)
    part = tmp_path / "part.out"
    part.write_text("\n".join(lines[: len(lines) // 2]) + "\n")
    F2, nb2, _ = _run(ca, T, resume=str(part))
    F3, nb3, _ = _run(ca, T, resume=str(full))
    print("T=%g: %d frequencies; F %.16g, resumed from half %.16g (%d of %d blocks), from all %.16g (%d blocks)"
          % (T, len(lines), F1, F2, nb2, nb1, F3, nb3))
    assert 0 < nb2 < nb1
    assert nb3 == 0
    # 16 printed digits: the resumed terms carry a relative rounding of 1e-16
    assert abs(F2 - F1) <= 1e-14 * abs(F1)
    assert abs(F3 - F1) <= 1e-14 * abs(F1)


def test_resume_plain_mode_matches(ca, tmp_path):
    full = tmp_path / "full.out"
    F1, nb1, _ = _run(ca, 300.0, out=str(full), mode="plain")
    F3, nb3, _ = _run(ca, 300.0, resume=str(full), mode="plain")
    assert nb3 == 0 and abs(F3 - F1) <= 1e-14 * abs(F1)


def test_checkpointed_T0_runs_one_node_per_call(ca, tmp_path):
    """With one quadrature node per batch and a time limit already passed,
    every call computes one node, writes its line and returns NaN; calls that
    resume from the growing output file end with the uninterrupted F."""
    import math
    F1, nb1, _ = _run(ca, 0.0)
    out = tmp_path / "ck.out"
    calls, nb_sum, F = 0, 0, float("nan")
    while math.isnan(F):
        calls += 1
        assert calls <= 200
        c = ca.Caps(R, L, ldim=LDIM)
        c.set_output(str(out))
        if out.exists():
            c.resume_load(str(out))
        c.set_checkpoint(nodes=1, seconds=1e-9)
        F, nb, _ = c.free_energy(0, mode="batched")
        c.set_output(None)
        nb_sum += nb
        lines = [ln for ln in out.read_text().splitlines() if ln.startswith("# xi*(L+R)/c=")]
        assert len(lines) == calls
    print("T=0 checkpointed: %d calls, %d blocks (uninterrupted %d), F %.16g vs %.16g" % (calls, nb_sum, nb1, F, F1))
    assert calls >= 4
    assert abs(F - F1) <= 1e-14 * abs(F1)

