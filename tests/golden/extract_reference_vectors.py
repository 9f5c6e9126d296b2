"""Transcribe the reference's golden test vectors into JSON fixtures.

Reads the reference test sources (``src/tests/test_*.c`` of CaPS v0.5, under
``$CAPS_REFERENCE`` or ``/root/reference``) and writes, for each assertion,
only data: the function, its arguments, the expected value, the tolerance the
reference test uses and the source line it came from.  No reference source
text is stored.  Run once in a container that holds the reference; the JSON
files are committed and are what the tests read.

    python tests/golden/extract_reference_vectors.py
"""
import json
import os
import re
import sys

REF = os.environ.get("CAPS_REFERENCE", "/root/reference")
TESTS = os.path.join(REF, "src", "tests")
OUT = os.path.dirname(os.path.abspath(__file__))

CAPS_C = 299792458.0
HBAR_EV = 6.582119514e-16

NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"
SAFE_EXPR = re.compile(r"^[0-9eE.+\-*/() ]+$")


def evaluate(expr, env):
    """Evaluate a C arithmetic initialiser such as ``50000.*CAPS_c/(L+R)``.
    Only numbers, + - * / and parentheses and the names in ``env`` are
    accepted; anything else raises."""
    e = expr.strip().rstrip(";")
    for name in sorted(env, key=len, reverse=True):
        e = re.sub(r"\b%s\b" % re.escape(name), repr(env[name]), e)
    if not SAFE_EXPR.match(e):
        raise ValueError("unexpected expression %r" % expr)
    return float(eval(e, {"__builtins__": {}}, {}))


def read_lines(fname, strip_comments=True):
    """Lines of a reference test file; block and line comments are blanked
    (newlines kept, so line numbers stay valid): commented-out assertions
    are not vectors."""
    text = open(os.path.join(TESTS, fname)).read()
    if strip_comments:
        text = re.sub(r"/\*.*?\*/", lambda m: "\n" * m.group(0).count("\n"), text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
    return text.splitlines()


def tolerance_by_line(lines):
    """Map line number -> tolerance of the enclosing unittest_init."""
    tol, out = None, {}
    for i, ln in enumerate(lines, 1):
        m = re.search(r"unittest_init\([^,]+,\s*\"([^\"]*)\"\s*,\s*\"[^\"]*\"\s*,\s*(%s)\)" % NUM, ln)
        if m:
            tol = (m.group(1), float(m.group(2)))
        out[i] = tol
    return out


def simple_calls(fname, func_re, nargs):
    """Assertions of the form AssertAlmostEqual(&t, func(a,b,...), value)."""
    lines = read_lines(fname)
    tols = tolerance_by_line(lines)
    pat = re.compile(r"AssertAlmostEqual\(&\w+,\s*(%s)\(([^()]*)\),\s*(%s)\s*\)" % (func_re, NUM))
    rows = []
    for i, ln in enumerate(lines, 1):
        m = pat.search(ln)
        if not m:
            continue
        args = [float(a) for a in m.group(2).split(",")]
        assert len(args) == nargs, ln
        rows.append({"fn": m.group(1), "args": args, "expect": float(m.group(3)),
                     "tol": tols[i][1], "suite": tols[i][0], "src": "src/tests/%s:%d" % (fname, i)})
    return rows


def mie_rows(fname):
    lines = read_lines(fname)
    tols = tolerance_by_line(lines)
    env = {"CAPS_c": CAPS_C, "CAPS_hbar_eV": HBAR_EV}
    R = L = None
    omegap = gamma = None
    call = None
    rows = []
    for i, ln in enumerate(lines, 1):
        for m in re.finditer(r"double\s+R\s*=\s*(%s)\s*,\s*L\s*=\s*(%s)" % (NUM, NUM), ln):
            env["R"], env["L"] = float(m.group(1)), float(m.group(2))
        m = re.search(r"caps_init\(([^,()]+),([^,()]+)\)", ln)
        if m:
            R, L = evaluate(m.group(1), env), evaluate(m.group(2), env)
            omegap = gamma = None
        for m in re.finditer(r"userdata\[(\d)\]\s*=\s*([^;]+);", ln):
            v = evaluate(m.group(2), env)
            if m.group(1) == "0":
                omegap = v
            else:
                gamma = v
        m = re.search(r"caps_mie(?:_perf)?\(\w+,\s*(%s),\s*(\d+),\s*&lna,\s*&lnb\)" % NUM, ln)
        if m:
            call = (float(m.group(1)), int(m.group(2)))
        m = re.search(r"AssertAlmostEqual\(&\w+,\s*(lna|lnb),\s*(%s)\s*\)" % NUM, ln)
        if m:
            rows.append({"R": R, "L": L, "omegap": omegap, "gamma": gamma,
                         "xi_": call[0], "l": call[1], "which": m.group(1),
                         "expect": float(m.group(2)), "tol": tols[i][1],
                         "src": "src/tests/%s:%d" % (fname, i)})
    return rows


def fresnel_rows(fname):
    lines = read_lines(fname)
    tols = tolerance_by_line(lines)
    env = {"CAPS_c": CAPS_C, "CAPS_hbar_eV": HBAR_EV}
    R = L = omegap = gamma = call = None
    rows = []
    for i, ln in enumerate(lines, 1):
        m = re.search(r"caps_init\(([^,()]+),([^,()]+)\)", ln)
        if m:
            R, L = evaluate(m.group(1), env), evaluate(m.group(2), env)
        for m in re.finditer(r"userdata\[(\d)\]\s*=\s*([^;]+);", ln):
            v = evaluate(m.group(2), env)
            if m.group(1) == "0":
                omegap = v
            else:
                gamma = v
        m = re.search(r"caps_fresnel\(\w+,\s*(%s),\s*(%s),\s*&rTE,\s*&rTM\)" % (NUM, NUM), ln)
        if m:
            call = (float(m.group(1)), float(m.group(2)))
        m = re.search(r"AssertAlmostEqual\(&\w+,\s*(rTE|rTM),\s*(%s)\s*\)" % NUM, ln)
        if m:
            rows.append({"R": R, "L": L, "omegap": omegap, "gamma": gamma,
                         "xi_": call[0], "k_": call[1], "which": m.group(1),
                         "expect": float(m.group(2)), "tol": tols[i][1],
                         "src": "src/tests/%s:%d" % (fname, i)})
    return rows


def logdetD_rows(fname):
    """Block values of test_logdetD: arrays v1,v2,v3 (R/L=100, ldim=500 at
    xi 1, 0.1, 10, m = index) and the explicit R/L=1000 calls."""
    lines = read_lines(fname, strip_comments=False)
    rows = []
    arrays = {}
    cur, start = None, None
    for i, ln in enumerate(lines, 1):
        m = re.search(r"double\s+(v\d)\[\]\s*=\s*\{", ln)
        if m:
            cur, arrays[m.group(1)] = m.group(1), []
            continue
        if cur:
            m = re.match(r"\s*(%s)\s*,?\s*/\*\s*m=(\d+)" % NUM, ln)
            if m:
                arrays[cur].append((float(m.group(1)), int(m.group(2)), i))
            if "};" in ln:
                cur = None
    state = {}
    for i, ln in enumerate(lines, 1):
        m = re.search(r"caps_init\((%s),\s*(%s)\)" % (NUM, NUM), ln)
        if m:
            state = {"R": float(m.group(1)), "L": float(m.group(2))}
        m = re.search(r"caps_set_ldim\(\w+,\s*(\d+)\)", ln)
        if m:
            state["ldim"] = int(m.group(1))
        m = re.search(r"caps_logdetD\(\w+,\s*(%s),\s*m\)\s*,\s*(v\d)\[m\]" % NUM, ln)
        if m:
            xi = float(m.group(1))
            for v, mm, src in arrays[m.group(2)]:
                rows.append(dict(state, xi_=xi, m=mm, expect=v, tol=1e-10,
                                 src="src/tests/%s:%d" % (fname, src)))
        m = re.search(r"AssertAlmostEqual\(&\w+,\s*caps_logdetD\(\w+,\s*(%s),\s*(\d+)\),\s*(%s)\)" % (NUM, NUM), ln)
        if m:
            rows.append(dict(state, xi_=float(m.group(1)), m=int(m.group(2)),
                             expect=float(m.group(3)), tol=1e-10,
                             src="src/tests/%s:%d" % (fname, i)))
    return rows


def logdetD0_rows(fname):
    """ξ = 0 assertions of test_logdetD0: caps_ht_plasma(omegap) values
    (R/L = 100, ldim = 500) and the m = 0 EE, MM blocks (R/L = 1000,
    ldim = 12000)."""
    lines = read_lines(fname)
    tols = tolerance_by_line(lines)
    rows = []
    state, env, m0 = {}, {"CAPS_c": CAPS_C}, None
    for i, ln in enumerate(lines, 1):
        m = re.search(r"const double R\s*=\s*(%s)\s*,\s*L\s*=\s*(%s)\s*;" % (NUM, NUM), ln)
        if m:
            env["R"], env["L"] = float(m.group(1)), float(m.group(2))
        m = re.search(r"caps_init\(\s*(\w+|%s)\s*,\s*(\w+|%s)\s*\)" % (NUM, NUM), ln)
        if m:
            state = {"R": evaluate(m.group(1), env), "L": evaluate(m.group(2), env)}
        m = re.search(r"caps_set_ldim\(\w+,\s*(\d+)\)", ln)
        if m:
            state["ldim"] = int(m.group(1))
        m = re.search(r"caps_logdetD0\(\w+,\s*(\d+),", ln)
        if m:
            m0 = int(m.group(1))
        m = re.search(r"AssertAlmostEqual\(&\w+,\s*caps_ht_plasma\(\w+,\s*([^,]+),\s*eps\),\s*(%s)\)" % NUM, ln)
        if m and tols[i] and tols[i][0] == "caps_logdetD0":
            rows.append(dict(state, fn="ht_plasma", omegap=evaluate(m.group(1), env), eps=1e-13,
                             expect=float(m.group(2)), tol=tols[i][1], src="src/tests/%s:%d" % (fname, i)))
        m = re.search(r"AssertAlmostEqual\(&\w+,\s*(EE|MM),\s*(%s)\)" % NUM, ln)
        if m and tols[i] and tols[i][0] == "caps_logdetD0":
            rows.append(dict(state, fn="logdetD0_" + m.group(1), m=m0, expect=float(m.group(2)), tol=tols[i][1],
                             src="src/tests/%s:%d" % (fname, i)))
    return rows


def manual_runs():
    """The free-energy runs the reference's manual prints
    (docs/manual/index.rst): per run its command-line options, the parameters
    of its header (ldim, cutoff, epsrel, cores), every printed per-frequency
    value "xi*(L+R)/c=..., logdetD=..." (the driver's F_xi, src/caps.c:
    307-396), the "# plasma = " value where printed, the number of
    determinants and the final line's values; each with its line number."""
    path = os.path.join(REF, "docs", "manual", "index.rst")
    lines = open(path).read().splitlines()
    runs, run = [], None
    for i, ln in enumerate(lines, 1):
        t = ln.strip()
        m = re.match(r"\$ mpirun -n (\d+) \./caps (.*)$", t)
        if m:
            if run:
                runs.append(run)
            opts, toks = {}, m.group(2).split()
            for a, b in zip(toks[::2], toks[1::2]):
                opts[a.lstrip("-")] = b
            run = {"src": "docs/manual/index.rst:%d" % i, "cores": int(m.group(1)), "options": opts,
                   "xi": [], "header": {}}
            continue
        if run is None:
            continue
        m = re.match(r"# (ldim|cutoff|epsrel|iepsrel|T|cores) = (%s)$" % NUM, t)
        if m:
            run["header"][m.group(1)] = float(m.group(2))
        m = re.match(r"# xi\*\(L\+R\)/c=(%s), logdetD=(%s), t=" % (NUM, NUM), t)
        if m:
            run["xi"].append({"xi_": float(m.group(1)), "value": float(m.group(2)), "line": i})
        m = re.match(r"# plasma = (%s) " % NUM, t)
        if m:
            run["plasma_xi0"] = float(m.group(1))
        m = re.match(r"# (\d+) determinants computed", t)
        if m:
            run["determinants"] = int(m.group(1))
        if re.match(r"(%s), (%s), (%s)" % (NUM, NUM, NUM), t) and not t.startswith("#"):
            run["final"] = [float(x) for x in t.split(",")[:3]] + [float(x) for x in t.split(",")[3:]]
            run["final_line"] = i
            runs.append(run)
            run = None
    return [r for r in runs if "final" in r]


def main():
    if not os.path.isdir(TESTS):
        sys.exit("reference tests not found at %s" % TESTS)
    fixtures = {
        "lnPlm.json": simple_calls("test_lnPlm.c", r"lnPlm", 3),
        "lnLambda.json": simple_calls("test_lnLambda.c", r"caps_lnLambda", 3),
        "mie_perf.json": mie_rows("test_mie.c"),
        "mie_drude.json": mie_rows("test_mie_drude.c"),
        "fresnel.json": fresnel_rows("test_fresnel.c"),
        "logdetD.json": logdetD_rows("test_logdetD.c"),
        "logdetD0.json": logdetD0_rows("test_logdetD.c"),
    }
    # Bessel assertions take one or two arguments depending on the function.
    lines = read_lines("test_bessels.c")
    tols = tolerance_by_line(lines)
    pat = re.compile(r"AssertAlmostEqual\(&\w+,\s*bessel_(\w+)\(([^()]*)\),\s*(%s)\s*\)" % NUM)
    rows = []
    for i, ln in enumerate(lines, 1):
        m = pat.search(ln)
        if m:
            rows.append({"fn": m.group(1), "args": [float(a) for a in m.group(2).split(",")],
                         "expect": float(m.group(3)), "tol": tols[i][1],
                         "src": "src/tests/test_bessels.c:%d" % i})
    fixtures["bessel.json"] = rows
    for name, rows in fixtures.items():
        with open(os.path.join(OUT, name), "w") as f:
            json.dump({"source": "CaPS v0.5 src/tests (transcribed by extract_reference_vectors.py)",
                       "rows": rows}, f, indent=0)
        print("%-16s %5d rows" % (name, len(rows)))
    runs = manual_runs()
    with open(os.path.join(OUT, "manual_runs.json"), "w") as f:
        json.dump({"source": "CaPS v0.5 docs/manual/index.rst (transcribed by extract_reference_vectors.py)",
                   "rows": runs}, f, indent=1)
    print("%-16s %5d runs" % ("manual_runs.json", len(runs)))
    # the optical data of gold the manual's material run reads (a data
    # table: xi in rad/s, epsilon(i xi), with its Drude extrapolation header)
    with open(os.path.join(REF, "materials", "gold.csv")) as f, open(os.path.join(OUT, "gold.csv"), "w") as g:
        g.write(f.read())


if __name__ == "__main__":
    main()
