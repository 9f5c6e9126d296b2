"""The free-energy drivers' resume cache (the reference driver's --resume,
src/caps.c:72-101, 316-327): parsing of the output lines and the frequency
matching rule, on the CPU (host code of libcaps_amd.so, no GPU calls)."""
import json
import os

import pytest

from conftest import GOLDEN


def _lines(rows):
    # the reference's output line (src/caps.c:265), one per frequency
    out = ["# R = 5e-05", "# some other comment, logdetD=ignored"]
    for r in rows:
        out.append("# xi*(L+R)/c=%.16g, logdetD=%.16g, t=%g" % (r["xi_"], r["value"], 1.5))
    return "\n".join(out) + "\n"


def test_resume_parse_and_match(tmp_path):
    import caps_amd
    with open(os.path.join(GOLDEN, "manual_runs.json")) as f:
        rows = json.load(f)["rows"][0]["xi"]
    p = tmp_path / "run.out"
    p.write_text(_lines(rows))
    c = caps_amd.Caps(50e-6, 500e-9, ldim=30)
    assert c.resume_load(str(p)) == len(rows)
    for r in rows:
        assert c.resume_lookup(r["xi_"]) == r["value"]
        # the reference's rule: equal, or |1 - xi_cached/xi| < 1e-11
        assert c.resume_lookup(r["xi_"] * (1 + 1e-12)) == r["value"]
        assert c.resume_lookup(r["xi_"] * (1 + 1e-10)) is None
    assert c.resume_lookup(123.456) is None
    # loading appends; None clears
    assert c.resume_load(str(p)) == len(rows)
    c.resume_load(None)
    assert c.resume_lookup(rows[0]["xi_"]) is None


def test_resume_errors(tmp_path):
    import caps_amd
    c = caps_amd.Caps(50e-6, 500e-9, ldim=30)
    with pytest.raises(ValueError):
        c.resume_load(str(tmp_path / "missing.out"))
    bad = tmp_path / "bad.out"
    bad.write_text("# xi*(L+R)/c=1.5 logdetD=-2.0\n")   # no commas: the reference aborts here
    with pytest.raises(ValueError):
        c.resume_load(str(bad))


def test_output_line_format(tmp_path):
    """set_output opens the file for appending (lines are written by the
    drivers on the GPU; here only the open/close path)."""
    import caps_amd
    c = caps_amd.Caps(50e-6, 500e-9, ldim=30)
    p = tmp_path / "out.txt"
    c.set_output(str(p))
    c.set_output(None)
    assert p.exists()
    with pytest.raises(ValueError):
        c.set_output(str(tmp_path / "no" / "such" / "dir.txt"))


def test_checkpoint_arguments():
    """capsx_set_checkpoint rejects negative node counts and times (host code)."""
    import caps_amd
    c = caps_amd.Caps(50e-6, 500e-9, ldim=30)
    with pytest.raises(ValueError):
        c.set_checkpoint(nodes=-1)
    with pytest.raises(ValueError):
        c.set_checkpoint(seconds=-1.0)
    with pytest.raises(ValueError):
        c.set_checkpoint(seconds=float("nan"))
    c.set_checkpoint(nodes=3, seconds=0.0)
    c.set_checkpoint()
