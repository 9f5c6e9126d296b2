"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: tests, ``__graft_entry__.smoke()`` and the CPU
baseline leg of ``bench.py`` use it as the checker.  The product path
(``caps_amd``) never imports it.
"""
import ctypes
import math
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

CAPS_C = 299792458.0
HBAR_EV = 6.582119514e-16


class OrcCaps(ctypes.Structure):
    _fields_ = [
        ("R", ctypes.c_double), ("L", ctypes.c_double), ("LbyR", ctypes.c_double),
        ("calL", ctypes.c_double), ("y", ctypes.c_double),
        ("ldim", ctypes.c_int), ("epsrel", ctypes.c_double), ("detalg", ctypes.c_int),
        ("omegap_plate", ctypes.c_double), ("gamma_plate", ctypes.c_double),
        ("omegap_sphere", ctypes.c_double), ("gamma_sphere", ctypes.c_double),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        d, i, p = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        sp = ctypes.POINTER(ctypes.c_byte)
        sigs = {
            "orc_logi": (d, [ctypes.c_uint]), "orc_lfac": (d, [ctypes.c_uint]),
            "orc_lfac2": (d, [ctypes.c_uint]),
            "orc_bessel_logIn_half": (d, [i, d]), "orc_bessel_logKn_half": (d, [i, d]),
            "orc_bessel_ratioI": (d, [d, d]), "orc_bessel_logIn": (d, [i, d]),
            "orc_bessel_logInu_asymp": (d, [d, d]), "orc_bessel_logKnu_asymp": (d, [d, d]),
            "orc_lnPl": (d, [i, d]), "orc_lnPlm": (d, [i, i, d]),
            "orc_lnPlm_upwards": (d, [i, i, d]), "orc_lnPlm_downwards": (d, [i, i, d]),
            "orc_Plm_cf": (d, [ctypes.c_long, ctypes.c_long, d]),
            "orc_dlnPlm": (d, [i, i, d, dp]),
            "orc_caps_init": (None, [ctypes.POINTER(OrcCaps), d, d]),
            "orc_lnLambda": (d, [i, i, i]),
            "orc_estimate_lmin": (i, [ctypes.POINTER(OrcCaps), i]),
            "orc_fresnel": (None, [d, d, d, dp, dp]),
            "orc_mie": (None, [ctypes.POINTER(OrcCaps), d, i, dp, dp]),
            "orc_K_estimate": (d, [i, i, d, d, dp, dp, dp]),
            "orc_block_new": (p, [ctypes.POINTER(OrcCaps), d, i]),
            "orc_block_free": (None, [p]), "orc_block_lmin": (i, [p]),
            "orc_block_K": (d, [p, i, i, sp]), "orc_block_I": (d, [p, i, i, i, sp]),
            "orc_block_I_abs": (d, [p, i, i, i]),
            "orc_block_elem": (d, [p, i, i, ctypes.c_char, ctypes.c_char]),
            "orc_block_matrix": (None, [p, dp]),
            "orc_block_counters": (ctypes.c_long, [p] + [ctypes.POINTER(ctypes.c_long)] * 4),
            "orc_logdet_lu": (d, [dp, i]), "orc_logdet_1mM": (d, [dp, i, i]),
            "orc_dominant_1mM": (i, [dp, i]), "orc_logdet_1mM_delta": (d, [dp, i]),
            "orc_logdetD": (d, [ctypes.POINTER(OrcCaps), d, i]),
            "orc_logdetD_many": (None, [ctypes.POINTER(OrcCaps), dp, ctypes.POINTER(ctypes.c_int), i, dp, i]),
            "orc_ht_drude": (d, [d]),
            "orc_integrate_plasma": (d, [i, d, d]),
            "orc_logdetD0": (None, [ctypes.POINTER(OrcCaps), i, d, dp, dp, dp]),
            "orc_ht_perf": (d, [ctypes.POINTER(OrcCaps), d]),
            "orc_ht_plasma": (d, [ctypes.POINTER(OrcCaps), d, d]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def caps(R, L, ldim=None, epsrel=1e-12, detalg=0, drude=None, drude_plate=None, drude_sphere=None):
    """Oracle geometry/material record.  ``drude=(omegap, gamma)`` in rad/s
    sets both bodies; None means perfect reflectors."""
    c = OrcCaps()
    lib().orc_caps_init(ctypes.byref(c), R, L)
    if ldim is not None:
        c.ldim = ldim
    c.epsrel = epsrel
    c.detalg = detalg
    dp_ = drude_plate if drude_plate is not None else drude
    ds_ = drude_sphere if drude_sphere is not None else drude
    if dp_ is not None:
        c.omegap_plate, c.gamma_plate = dp_
    if ds_ is not None:
        c.omegap_sphere, c.gamma_sphere = ds_
    return c


def logdetD(c, xi_, m):
    return lib().orc_logdetD(ctypes.byref(c), xi_, m)


def logdetD_many(c, xis, ms, threads=1):
    import numpy as np
    xis = np.ascontiguousarray(xis, dtype=np.float64)
    ms = np.ascontiguousarray(ms, dtype=np.int32)
    out = np.empty(len(xis))
    lib().orc_logdetD_many(ctypes.byref(c), xis.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                           ms.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(xis),
                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), threads)
    return out


class Block:
    """One (ξ, m) block of the oracle: K, I, elements, matrix."""

    def __init__(self, c, xi_, m):
        self.c = c
        self.ptr = lib().orc_block_new(ctypes.byref(c), xi_, m)
        if not self.ptr:
            raise ValueError("invalid block")
        self.lmin = lib().orc_block_lmin(self.ptr)
        self.m = m

    def K(self, nu, p):
        s = ctypes.c_byte()
        v = lib().orc_block_K(self.ptr, nu, p, ctypes.byref(s))
        return v, s.value

    def I(self, l1, l2, p):
        s = ctypes.c_byte()
        v = lib().orc_block_I(self.ptr, l1, l2, p, ctypes.byref(s))
        return v, s.value

    def I_abs(self, l1, l2, p):
        """log of the sum of the q-sum's term magnitudes of I_{l1 l2}^p."""
        return lib().orc_block_I_abs(self.ptr, l1, l2, p)

    def elem(self, l1, l2, p1, p2):
        return lib().orc_block_elem(self.ptr, l1, l2, p1.encode(), p2.encode())

    def matrix(self):
        import numpy as np
        N = 2 * self.c.ldim
        M = np.empty((N, N))
        lib().orc_block_matrix(self.ptr, M.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return M

    def counters(self):
        v = [ctypes.c_long() for _ in range(4)]
        lib().orc_block_counters(self.ptr, *[ctypes.byref(x) for x in v])
        return dict(zip(["n_K", "n_Kevals", "n_I", "n_qterms"], [x.value for x in v]))

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc_block_free(self.ptr)
            self.ptr = None


def logdet_1mM(M, detalg=0):
    import numpy as np
    M = np.ascontiguousarray(M, dtype=np.float64)
    return lib().orc_logdet_1mM(M.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), M.shape[0], detalg)


def dominant_1mM(M):
    import numpy as np
    M = np.ascontiguousarray(M, dtype=np.float64)
    return bool(lib().orc_dominant_1mM(M.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), M.shape[0]))


def logdet_1mM_delta(M):
    import numpy as np
    M = np.ascontiguousarray(M, dtype=np.float64)
    return lib().orc_logdet_1mM_delta(M.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), M.shape[0])


def mie(c, xi_, l):
    a, b = ctypes.c_double(), ctypes.c_double()
    lib().orc_mie(ctypes.byref(c), xi_, l, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def fresnel(epsm1, xi_, k_):
    a, b = ctypes.c_double(), ctypes.c_double()
    lib().orc_fresnel(epsm1, xi_, k_, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value
