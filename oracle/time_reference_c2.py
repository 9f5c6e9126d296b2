"""Time the reference library (oracle/_ref/libcaps_ref.so, compiled from its
own sources by oracle/build_ref.sh; default HODLR determinant) and the CPU
oracle on a stratified sample of the R/L = 200 Drude free energy's blocks,
one block per process / thread, on the host it runs on, and write the result
to profiles/cpu_reference_c2_sample.txt.  Runs only where the reference tree
is present (the build container): the reference never goes to the GPU host.

    python oracle/time_reference_c2.py [sample] [processes]
"""
import json
import multiprocessing as mp
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)
sys.path.insert(0, ROOT)
from bench import R, L, LDIM, OMEGAP_EV, GAMMA_EV, c2_block_sample, cpu_model  # noqa: E402

HBAR_EV = 6.582119514e-16


def _ref_block(job):
    import ctypes
    xi, m = job
    lib = ctypes.CDLL(os.path.join(HERE, "_ref", "libcaps_ref.so"), mode=os.RTLD_LAZY)
    lib.caps_init.restype = ctypes.c_void_p
    lib.caps_init.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.caps_logdetD.restype = ctypes.c_double
    lib.caps_logdetD.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
    lib.caps_set_ldim.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.caps_set_epsilonm1.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    ud = (ctypes.c_double * 2)(OMEGAP_EV / HBAR_EV, GAMMA_EV / HBAR_EV)
    c = lib.caps_init(R, L)
    lib.caps_set_ldim(c, LDIM)
    lib.caps_set_epsilonm1(c, ctypes.cast(lib.caps_epsilonm1_drude, ctypes.c_void_p), ctypes.cast(ud, ctypes.c_void_p))
    t = time.perf_counter()
    v = lib.caps_logdetD(c, xi, m)
    return v, time.perf_counter() - t


def main():
    sample = int(sys.argv[1]) if len(sys.argv) > 1 else 32
    procs = int(sys.argv[2]) if len(sys.argv) > 2 else len(os.sched_getaffinity(0))
    if not os.path.exists(os.path.join(HERE, "_ref", "libcaps_ref.so")):
        sys.exit("oracle/_ref/libcaps_ref.so not built (needs the reference tree)")
    jobs = c2_block_sample(sample)
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(procs) as pool:
        res = pool.map(_ref_block, jobs, chunksize=1)
    dt = time.perf_counter() - t0
    import pyoracle as po
    oc = po.caps(R, L, ldim=LDIM, epsrel=1e-8, drude=(OMEGAP_EV / HBAR_EV, GAMMA_EV / HBAR_EV))
    t1 = time.perf_counter()
    po.logdetD_many(oc, [x for x, _ in jobs], [m for _, m in jobs], threads=procs)
    dto = time.perf_counter() - t1
    out = os.path.join(ROOT, "profiles", "cpu_reference_c2_sample.txt")
    with open(out, "w") as f:
        f.write("# CPU timing of the R/L = 200 Drude workload's blocks on the build container (no GPU)\n")
        f.write("# command: python oracle/time_reference_c2.py %d %d\n" % (sample, procs))
        f.write("host = %s, %d processes / threads\n" % (cpu_model(), procs))
        f.write("sample = %d of the 1120 blocks (every k-th (xi_n, m) of the sequential sums), N = 2400\n" % sample)
        f.write("reference_blocks_per_s = %.4f   # reference library, default HODLR, one block per process\n"
                % (sample / dt))
        f.write("reference_seconds_per_block_mean = %.3f\n" % (sum(t for _, t in res) / len(res)))
        f.write("oracle_blocks_per_s = %.4f   # CPU oracle (epsrel 1e-8, dense delta-form LU), OpenMP\n"
                % (sample / dto))
        f.write("whole_energy_reference = 1120 blocks in 415.6 s on 8 processes (2.70 blocks/s; "
                "tests/golden/ref_free_energy_c2.json seconds_8_workers)\n")
    print(open(out).read())


if __name__ == "__main__":
    main()
