"""Benchmark: logdetD blocks/s of the round-trip-matrix path on one MI355X.

Workload (one "step"): the T = 300 K free energy of a Drude sphere (ωp = 9 eV,
γ = 35 meV) above a Drude plate, R = 200 µm, L = 1 µm (R/L = 200),
ldim = 1200 (blocks 2400 x 2400, FP64), full Matsubara sum with the
reference driver's cutoffs (m-cutoff 1e-9, Matsubara epsrel 1e-6).  Every
block entering the energy is computed on the GPU; the batched form runs all
m of a Matsubara frequency in one launch sequence.  The same step in the
plain one-block-at-a-time form is timed once beside it.

Prints one JSON line (the driver's contract), with a roofline entry for the
dominant kernel (HIP events on the library's stream) and a CPU baseline: the
CPU oracle on a bounded sample of the workload's blocks on this host's cores,
with the reference library's own timing on the build container beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

R, L, LDIM, T = 200e-6, 1e-6, 1200, 300.0
OMEGAP_EV, GAMMA_EV = 9.0, 0.035
CUTOFF, EPSREL = 1e-9, 1e-6
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def gpu_step(c, mode):
    t0 = time.perf_counter()
    F, ncomputed, nused = c.free_energy(T, cutoff=CUTOFF, epsrel=EPSREL, mode=mode)
    return F, nused, ncomputed, time.perf_counter() - t0


def c2_block_sample(k):
    """k blocks spread evenly over the 1120 (ξ_n, m) blocks of the workload's
    sequential sums (ξ_1 .. ξ_9, m up to each m-cutoff, from
    tests/golden/ref_free_energy_c2.json), so the sample has the energy's
    mix of frequencies and m."""
    with open(os.path.join(ROOT, "tests", "golden", "ref_free_energy_c2.json")) as f:
        per_xi = json.load(f)["data"]["per_xi"]
    blocks = [(r["xi_"], m) for r in per_xi for m in range(r["mstop"] + 1)]
    return [blocks[(i * len(blocks)) // k] for i in range(k)]


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sample_blocks=32):
    """The CPU oracle (this repository's FP64 C restatement of the path:
    adaptive quadrature at the reference's epsrel 1e-8, dense delta-form LU)
    on a sample of the workload's blocks spread over all its Matsubara terms,
    one block per OpenMP thread on the host cores available (at most 16,
    the GPU host's share), as the reference's one-block-per-process farm;
    and the reference library itself, timed on the build container (it does
    not go to the GPU host), from profiles/cpu_reference_c2_sample.txt."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    cores = min(len(os.sched_getaffinity(0)), 16)
    jobs = c2_block_sample(sample_blocks)
    hb = po.HBAR_EV
    oc = po.caps(R, L, ldim=LDIM, epsrel=1e-8, drude=(OMEGAP_EV / hb, GAMMA_EV / hb))
    t0 = time.perf_counter()
    vals = po.logdetD_many(oc, [x for x, _ in jobs], [m for _, m in jobs], threads=cores)
    dt = time.perf_counter() - t0
    ref = {}
    path = os.path.join(ROOT, "profiles", "cpu_reference_c2_sample.txt")
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                if "=" in line and not line.startswith("#"):
                    k, v = line.split("=", 1)
                    ref[k.strip()] = v.split("#")[0].strip()
    return {"value": len(jobs) / dt, "unit": "blocks/s", "cores": cores, "kind": "port",
            "sample": "%d of the workload's 1120 blocks (every 35th (xi_n, m), N=2400), CPU oracle (epsrel 1e-8, "
                      "dense delta-form LU), one block per thread" % len(jobs),
            "seconds": dt, "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "reference_other_host": ref or None, "values": list(vals)}


def block_errors(ca):
    """Largest deviations of the GPU blocks from the reference's values held
    in tests/golden/ref_blocks.json (C1: all 121 blocks at R/L = 20; C2: 17
    blocks of the workload), per kind: trace-shortcut blocks relative,
    factored blocks relative and as a fraction of the tolerance of
    tests/conftest.py (1e-9 |v| + 2 max(1e-13, 1e-13 |tr M|))."""
    with open(os.path.join(ROOT, "tests", "golden", "ref_blocks.json")) as f:
        rows = json.load(f)["data"]
    out = {}
    for tag in ("C1", "C2"):
        rs = [r for r in rows if r["set"] == tag]
        c = ca.Caps(rs[0]["R"], rs[0]["L"], ldim=rs[0]["ldim"])
        if rs[0]["omegap"]:
            c.set_drude(rs[0]["omegap"] * ca.HBAR_EV, rs[0]["gamma"] * ca.HBAR_EV)
        v, tr, fl = c.logdetD_batch([r["xi_"] for r in rs], [r["m"] for r in rs], info=True)
        short = [abs(x / r["logdetD"] - 1) for x, r, f in zip(v, rs, fl) if (f & 3) == 1]
        lu = [(abs(x / r["logdetD"] - 1), abs(x - r["logdetD"]) / (1e-9 * abs(r["logdetD"]) + 2 * max(1e-13, 1e-13 * abs(t))))
              for x, r, t, f in zip(v, rs, tr, fl) if (f & 3) != 1]
        out[tag] = {"blocks": len(rs), "trace_shortcut": len(short), "max_rel_trace_shortcut": max(short, default=0.0),
                    "factored": len(lu), "max_rel_factored": max((a for a, _ in lu), default=0.0),
                    "max_err_over_tolerance": max((b for _, b in lu), default=0.0)}
    return out


def mixed_aspect(ca, steps):
    """The 64-block mixed-geometry batch (perfect reflectors, R/L log-U[50,
    500], ldim = ceil(6 R/L), seed 0) as one batched call; blocks/s, LU
    TF/s, and the largest deviation from the reference values held in
    tests/golden/ref_mixed_aspect.json."""
    from caps_amd.workloads import mixed_aspect_blocks, make_objects
    blocks = mixed_aspect_blocks()
    objs = make_objects(blocks)
    xs = [b["xi_"] for b in blocks]
    ms = [b["m"] for b in blocks]
    _, tr, fl = ca.logdetD_multi(objs, xs, ms, info=True)
    objs[0].reset_stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = ca.logdetD_multi(objs, xs, ms)
    dt = (time.perf_counter() - t0) / steps
    st = objs[0].stats(reset=True)
    lu_s = st["stage_ms"]["lu"] / 1e3
    # the same batch captured once into a HIP graph and replayed
    objs[0].set_graph(True)
    out_g = ca.logdetD_multi(objs, xs, ms)
    t0 = time.perf_counter()
    for _ in range(steps):
        out_g = ca.logdetD_multi(objs, xs, ms)
    dt_graph = (time.perf_counter() - t0) / steps
    objs[0].set_graph(False)
    objs[0].reset_stats()
    same = bool((out_g == out).all())
    err = None
    ref = os.path.join(ROOT, "tests", "golden", "ref_mixed_aspect.json")
    if os.path.exists(ref):
        with open(ref) as f:
            rows = json.load(f)["data"]
        # tolerance of tests/conftest.py: trace-shortcut blocks 1e-9
        # relative, factored blocks 1e-9 |v| + 2 max(1e-13, 1e-13 |tr M|)
        tol = [1e-9 * abs(r["logdetD"]) + (0 if (f & 3) == 1 else 2 * max(1e-13, 1e-13 * abs(t)))
               for r, t, f in zip(rows, tr, fl)]
        err = {"max_abs": max(abs(v - r["logdetD"]) for v, r in zip(out, rows)),
               "max_rel": max(abs(v - r["logdetD"]) / abs(r["logdetD"]) for v, r in zip(out, rows)),
               "max_err_over_tolerance": max(abs(v - r["logdetD"]) / t for v, r, t in zip(out, rows, tol))}
    print("bench: mixed-aspect batch %.3f s for %d blocks" % (dt, len(blocks)), file=sys.stderr, flush=True)
    return {"workload": "64 perfect-reflector blocks, R/L log-U[50,500], ldim=ceil(6 R/L), xi_ U[0.1,300], "
                        "m U[0,ldim/4], seed 0, one batch",
            "blocks_per_s": len(blocks) / dt, "seconds_per_batch": dt,
            "graph_replay": {"blocks_per_s": len(blocks) / dt_graph, "seconds_per_batch": dt_graph,
                             "identical_to_eager": same},
            "dims": [min(2 * b["ldim"] for b in blocks), max(2 * b["ldim"] for b in blocks)],
            "stage_ms_per_batch": {k: v / steps for k, v in st["stage_ms"].items()},
            "lu_tflops": st["lu_flops"] / lu_s / 1e12 if lu_s > 0 else None,
            "blocks_factored": st["blocks_factored"] / steps,
            "error_vs_reference": err}


def hbm_probe(nbytes=1 << 30, reps=3):
    """A plain copy of known size: per dispatch it reads and writes nbytes.
    Profiled with FETCH_SIZE / WRITE_SIZE it gives the factor between those
    counters and bytes."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"hbm_probe_bytes_read_per_dispatch": nbytes, "bytes_written_per_dispatch": nbytes,
                      "dispatches": reps}))


def gemm_traffic():
    """HBM bytes per k_lu_gemm_dma launch measured with FETCH_SIZE and
    WRITE_SIZE (profiles/k_lu_gemm_dma_c2_traffic.txt, scaled by the copy
    calibration), or None when that summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_lu_gemm_dma_c2_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("hbm_bytes_per_launch"):
                return float(line.split("=")[1])
    return None


def large_aspect(ca):
    """Eight perfect-reflector blocks at R/L = 2000, ldim = 12000 (N = 24000,
    the large-aspect configuration) in one batch: blocks/s, LU TF/s and the
    largest deviation from the reference values in
    tests/golden/ref_c5_blocks.json."""
    path = os.path.join(ROOT, "tests", "golden", "ref_c5_blocks.json")
    with open(path) as f:
        rows = json.load(f)["data"]
    c = ca.Caps(2e-3, 1e-6, ldim=12000)
    xs = [r["xi_"] for r in rows]
    ms = [r["m"] for r in rows]
    t0 = time.perf_counter()
    c.logdetD_batch(xs, ms)          # first call: the object's HBM arena (~66 GB) is allocated
    dt_first = time.perf_counter() - t0
    c.reset_stats()
    t0 = time.perf_counter()
    out = c.logdetD_batch(xs, ms)
    dt = time.perf_counter() - t0
    st = c.stats(reset=True)
    lu_s = st["stage_ms"]["lu"] / 1e3
    err = max(abs(v - r["logdetD"]) / abs(r["logdetD"]) for v, r in zip(out, rows))
    print("bench: large-aspect batch %.2f s for %d blocks" % (dt, len(rows)), file=sys.stderr, flush=True)
    return {"workload": "8 perfect-reflector blocks, R/L=2000 (R=2mm, L=1um), ldim=12000 (N=24000), one batch",
            "blocks_per_s": len(rows) / dt, "seconds": dt, "seconds_first_call_with_allocation": dt_first,
            "stage_ms": {k: v for k, v in st["stage_ms"].items()},
            "lu_tflops": st["lu_flops"] / lu_s / 1e12 if lu_s > 0 else None,
            "max_rel_err_vs_reference": err}


LA_R, LA_L, LA_LDIM, LA_EPSREL = 2e-3, 1e-6, 12000, 5e-5


def fcqs_nodes(M, alpha):
    """(ξ(L+R)/c, weight) of the Fourier-Chebyshev rule on [0, ∞) with M - 1
    nodes (Boyd 1987, eq. 3.2e; the library's T = 0 rule), x = α ξ."""
    out = []
    for i in range(1, M):
        t = math.pi * i / M
        s = sum(math.sin(j * t) / j for j in range(1, M, 2))
        w = 8 * math.sin(t) / (1 - math.cos(t)) ** 2 / M * s
        out.append(((math.cos(t / 2) / math.sin(t / 2)) ** 2 / alpha, w))
    return out


def fcqs_levels(path, LbyR):
    """The rule's integral F (L+R)/(ħc) per level M = 5, 10, 20, ... from the
    F_ξ lines of an output file (frequencies above the driver's PFA cutoff
    count as 0); stops at the first level with a node missing."""
    vals = {}
    with open(path) as f:
        for line in f:
            if line.startswith("# xi*(L+R)/c="):
                a, b = line[13:].split(", logdetD=")
                vals[float(a)] = float(b.split(",")[0])
    alpha = 2 * LbyR / (1 + LbyR)
    xi_cut = -(1 + 1 / LbyR) * math.log(2 * LbyR * 1e-100) / 2
    levels, M = {}, 5
    while M <= 2560:
        I = 0.0
        for xi, w in fcqs_nodes(M, alpha):
            if xi > xi_cut:
                continue
            v = next((vals[k] for k in vals if k == xi or abs(1 - k / xi) < 1e-11), None)
            if v is None:
                return levels, len(vals)
            I += w * v
        levels[M] = I / alpha / math.pi
        M *= 2
    return levels, len(vals)


def pfa_rule_error(M):
    """Relative error of the M-level rule on the PFA shape of the T = 0
    integrand for perfect reflectors, f(x) = Σ_n n^-3 [E_2(n x) + n x E_1(n x)]
    (∫_0^∞ f = ζ(4) = π^4/90): the quadrature error expected at large R/L."""
    import numpy as np
    from scipy.special import expn
    n = np.arange(1, 400001, dtype=float)
    I = 0.0
    for x, w in fcqs_nodes(M, 1.0):
        I += w * float(np.sum((expn(2, n * x) + n * x * expn(1, n * x)) / n ** 3))
    return I / (math.pi ** 4 / 90) - 1


def large_aspect_energy(ca, out, budget):
    """The T = 0 free energy of the large-aspect configuration (perfect
    reflectors, R/L = 2000, ldim = 12000, blocks up to 24000 x 24000),
    surrogate-planned, one quadrature node per batch, resumed from and
    appended to the output file `out`; stops after `budget` seconds (F is
    then NaN and a later run continues).  When complete: the rule's levels,
    F against the proximity-force value and its first correction, and the
    largest-frequency node recomputed one block at a time (plain form)."""
    os.environ["CAPS_AMD_PROGRESS"] = "1"
    from caps_amd import surrogate
    c = ca.Caps(LA_R, LA_L, ldim=LA_LDIM)
    net = surrogate.MLP.load()
    net.attach(c)
    resumed = c.resume_load(out) if os.path.exists(out) else 0
    c.set_output(out)
    c.set_checkpoint(nodes=1, seconds=budget)
    t0 = time.perf_counter()
    F, nb, nu = c.free_energy(0, cutoff=CUTOFF, epsrel=LA_EPSREL, mode="surrogate")
    dt = time.perf_counter() - t0
    c.set_output(None)
    sur = c.surrogate_report()
    LbyR = LA_L / LA_R
    levels, nlines = fcqs_levels(out, LbyR)
    res = {"workload": "T=0 free energy, perfect reflectors, R/L=2000 (R=2mm, L=1um), ldim=12000 (N<=24000), "
                       "Fourier-Chebyshev xi-rule with epsrel %g, surrogate-planned m-cutoffs" % LA_EPSREL,
           "complete": bool(math.isfinite(F)), "F_scaled": F if math.isfinite(F) else None,
           "frequencies_resumed": resumed, "frequencies_in_file": nlines, "blocks_this_run": nb,
           "seconds_this_run": dt, "surrogate": sur,
           "levels": {str(k): v for k, v in levels.items()}}
    print("bench: large-aspect energy %s after %.0f s (%d blocks this run, %d frequencies in %s)"
          % ("complete" if res["complete"] else "incomplete", dt, nb, nlines, out), file=sys.stderr, flush=True)
    if not res["complete"]:
        return res
    F_pfa = -math.pi ** 3 / 720 * (1 / LbyR) * (1 + 1 / LbyR)
    theta = 1 / 3 - 20 / math.pi ** 2   # derivative expansion, electromagnetic perfect reflectors
    Ms = sorted(levels)
    res.update({"F_PFA_scaled": F_pfa, "F_over_PFA": F / F_pfa, "derivative_expansion": 1 + theta * LbyR,
                "F_over_PFA_minus_DE": F / F_pfa - (1 + theta * LbyR),
                "last_two_levels_rel_diff": abs(1 - levels[Ms[-2]] / levels[Ms[-1]]) if len(Ms) > 1 else None,
                "rule_error_on_pfa_shape": pfa_rule_error(Ms[-1])})
    # plain-form identity: the computed frequency nearest ξ(L+R)/c = 1e4
    # recomputed one block at a time, every other F_ξ from the file
    xs = []
    with open(out) as f:
        for line in f:
            if line.startswith("# xi*(L+R)/c="):
                a, b = line[13:].split(", logdetD=")
                xs.append((float(a), float(b.split(",")[0]), line))
    pick = min(xs, key=lambda r: abs(math.log(r[0] / 1e4)))
    rest, plain_out = out + ".rest", out + ".plain"
    with open(rest, "w") as f:
        f.writelines(r[2] for r in xs if r[0] != pick[0])
    c2 = ca.Caps(LA_R, LA_L, ldim=LA_LDIM)
    c2.resume_load(rest)
    c2.set_output(plain_out)
    t0 = time.perf_counter()
    Fp, nbp, _ = c2.free_energy(0, cutoff=CUTOFF, epsrel=LA_EPSREL, mode="plain")
    dtp = time.perf_counter() - t0
    c2.set_output(None)
    with open(plain_out) as f:
        fxi = [float(line.split(", logdetD=")[1].split(",")[0]) for line in f if line.startswith("# xi*(L+R)/c=")]
    res["plain_check"] = {"xi_": pick[0], "blocks": nbp, "seconds": dtp, "F_xi_batched": pick[1],
                          "F_xi_plain": fxi[-1] if fxi else None,
                          "F_xi_rel_diff": abs(fxi[-1] - pick[1]) / abs(pick[1]) if fxi else None,
                          "F_plain": Fp, "F_rel_diff": abs(Fp - F) / abs(F)}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="launch the timed steps eagerly (default: HIP graph replay)")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--hbm-probe", action="store_true",
                    help="only run a 1 GiB device-to-device copy three times (calibrates HBM byte counters)")
    ap.add_argument("--large-energy", metavar="OUT",
                    help="only run the T=0 free energy at R/L=2000, ldim=12000, resumed from and appended to OUT")
    ap.add_argument("--budget", type=float, default=900.0,
                    help="seconds after which --large-energy stops before its next node (continue with a new run)")
    args = ap.parse_args()
    if args.hbm_probe:
        hbm_probe()
        return
    if args.large_energy:
        import caps_amd
        caps_amd.lib()
        print(json.dumps({"large_aspect_energy": large_aspect_energy(caps_amd, args.large_energy, args.budget)}),
              flush=True)
        return
    if args.gpus != 1:
        print("bench.py: the library runs on one GPU only (every (xi, m) block fits one MI355X); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import caps_amd
    caps_amd.lib()
    c = caps_amd.Caps(R, L, ldim=LDIM)
    c.set_drude(OMEGAP_EV, GAMMA_EV)
    form = "batched"
    from caps_amd import surrogate
    if os.path.exists(surrogate.DEFAULT_WEIGHTS) and not args.no_surrogate:
        net = surrogate.MLP.load()
        net.attach(c, OMEGAP_EV, GAMMA_EV)
        form = "surrogate"

    # warmup eagerly, then (default) with graph capture on: the step's batch
    # shape is captured once and the timed steps replay it
    for _ in range(args.warmup):
        gpu_step(c, form)
    if not args.no_graph:
        c.set_graph(True)
        gpu_step(c, form)                  # captures the batch's launch sequence
    c.reset_stats()
    times, F, nb, ncomp = [], None, 0, 0
    for _ in range(args.steps):
        F, nb, ncomp, dt = gpu_step(c, form)
        times.append(dt)
        print("bench: %s step %.3f s, %d blocks in the energy (%d computed), F=%.16g%s" % (
            form, dt, nb, ncomp, F, "" if args.no_graph else " (graph replay)"), file=sys.stderr, flush=True)
    total = sum(times)
    value = nb * args.steps / total
    sur_acc = c.surrogate_report(reset=True) if form == "surrogate" else None
    # one eager step: stage times and LU rate (a replayed graph is timed as
    # a whole), and the same F bit for bit
    c.set_graph(False)
    c.reset_stats()
    Fe, _, _, dte = gpu_step(c, form)
    st = c.stats(reset=True)
    eager = {"seconds": dte, "F_identical_to_timed": Fe == F}
    print("bench: eager step %.3f s" % dte, file=sys.stderr, flush=True)
    st_steps = 1

    # one more step with HIP events around every launch of the dominant
    # kernel (the LU's trailing update); its launches then run one at a time
    c.profile(True)
    gpu_step(c, form)
    prof = c.stats(reset=True)
    c.profile(False)

    heuristic = None
    if form == "surrogate":
        Fh, nbh, nch, dth = gpu_step(c, "batched")
        heuristic = {"blocks_per_s": nbh / dth, "seconds": dth, "F": Fh, "blocks_computed": nch}
        print("bench: batched (no surrogate) step %.3f s" % dth, file=sys.stderr, flush=True)
    plain = None
    if not args.no_plain:
        Fp, nbp, _, dtp = gpu_step(c, "plain")
        plain = {"blocks_per_s": nbp / dtp, "seconds": dtp, "F": Fp, "blocks": nbp,
                 "speedup_vs_plain": (nb / (total / args.steps)) / (nbp / dtp)}
        print("bench: plain step %.3f s, %d blocks" % (dtp, nbp), file=sys.stderr, flush=True)

    # dominant kernel: the LU's trailing update (k_lu_gemm_dma), FP64 MFMA;
    # its launches only (the trailing updates of regions one panel wide or
    # high run on k_lu_gemm tiles and are reported beside it)
    gemm_s = prof["gemm_ms"] / 1e3
    achieved = prof["gemm_flops"] / gemm_s / 1e12 if gemm_s > 0 else None
    roof = {"kernel": "k_lu_gemm_dma", "bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS if achieved else None, "traffic": gemm_traffic(),
            "traffic_source": "stored: profiles/k_lu_gemm_dma_c2_traffic.txt (FETCH_SIZE + WRITE_SIZE runs)",
            "launches": prof["gemm_launches"],
            "avg_launch_us": 1e3 * prof["gemm_ms"] / max(1, prof["gemm_launches"]),
            "flops_per_launch": prof["gemm_flops"] / max(1, prof["gemm_launches"]),
            "algorithmic_bytes_per_launch": prof["gemm_bytes"] / max(1, prof["gemm_launches"]),
            "arithmetic_intensity_flop_per_byte": prof["gemm_flops"] / max(1.0, prof["gemm_bytes"]),
            "other_update_launches_ms": prof["other_update_ms"]}
    lu_s = st["stage_ms"]["lu"] / 1e3
    mixed = None if args.no_mixed else mixed_aspect(caps_amd, args.steps)
    large = None if args.no_large else large_aspect(caps_amd)
    errors = block_errors(caps_amd)
    ref_F = None
    with open(os.path.join(ROOT, "tests", "golden", "ref_free_energy_c2.json")) as f:
        ref_F = json.load(f)["data"]["F"]
    errors["F_rel_vs_reference"] = abs(F / ref_F - 1)
    cpu = None if args.no_cpu else cpu_baseline()
    line = {
        "metric": "logdetD blocks/s @R/L=200 FP64, batched+surrogate vs plain; FP64 MFMA TF/s",
        "value": value, "unit": "blocks/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1000 * total / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (analytic Drude material, no dataset)",
        "config": {"workload": "free energy T=300K, Drude 9eV/35meV both bodies, R/L=200 (R=200um, L=1um), "
                               "ldim=1200 (N=2400), m-cutoff 1e-9, Matsubara epsrel 1e-6",
                   "blocks_per_step": nb, "blocks_computed_per_step": ncomp, "F": F, "form": form},
        "plain": plain,
        "batched_without_surrogate": heuristic,
        "surrogate_vs_computed": sur_acc,
        "timed_steps": "HIP graph replay" if not args.no_graph else "eager",
        "eager_step": eager,
        "stage_ms_per_step": {k: v / st_steps for k, v in st["stage_ms"].items()},
        "lu_tflops": st["lu_flops"] / lu_s / 1e12 if lu_s > 0 else None,
        "blocks_factored_per_step": st["blocks_factored"] / st_steps,
        "roofline": roof,
        "errors_vs_reference": errors,
        "mixed_aspect_batch": mixed,
        "large_aspect_batch": large,
        "cpu_baseline": cpu and {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample", "nproc", "cpu_model",
                                                    "reference_other_host")},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
