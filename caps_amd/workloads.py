"""Synthetic block workloads shared by the benchmark and the parity tests.

mixed_aspect_blocks: the batched-launch workload of perfect-reflector blocks
over several geometries, drawn from numpy's default_rng(seed):
    R/L    log-uniform in [50, 500]   (L = 1 µm, R = R/L · L)
    ldim   ceil(6 R/L)                (block dimension 2 ldim, uneven)
    ξ(L+R)/c  uniform in [0.1, 300]
    m      uniform integer in [0, ldim/4]
Every block is its own geometry, so one batch holds blocks of 64 different
dimensions between about 600 and 6000.
"""
import math

import numpy as np


def mixed_aspect_blocks(n=64, seed=0, rbyl=(50.0, 500.0), ldim_factor=6.0, xi_range=(0.1, 300.0),
                        m_frac=0.25, L=1e-6):
    """List of dicts {R, L, RbyL, ldim, xi_, m} (perfect reflectors)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        RbyL = float(math.exp(rng.uniform(math.log(rbyl[0]), math.log(rbyl[1]))))
        ldim = int(math.ceil(ldim_factor * RbyL))
        xi = float(rng.uniform(*xi_range))
        m = int(rng.integers(0, int(m_frac * ldim) + 1))
        out.append({"R": RbyL * L, "L": L, "RbyL": RbyL, "ldim": ldim, "xi_": xi, "m": m})
    return out


def make_objects(blocks):
    """One caps object per block (the batch's geometries)."""
    import caps_amd
    return [caps_amd.Caps(b["R"], b["L"], ldim=b["ldim"]) for b in blocks]
