"""caps_amd: the CaPS plane-sphere round-trip-matrix path on MI355X (gfx950).

Python view of ``libcaps_amd.so`` (the libcaps C ABI; see include/libcaps.h).
Every numerical call runs the library's HIP kernels; importing this module
without the built library, or calling it without a usable gfx950 GPU, fails
loudly (the library aborts with a message, as the reference's TERMINATE
does).  There is no CPU fallback.

    from caps_amd import Caps
    c = Caps(R=200e-6, L=1e-6, ldim=1200)
    c.set_drude(9.0, 0.035)          # eV, both bodies
    c.logdetD(165.4564, 0)
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcaps_amd.so")

HBAR_EV = 6.582119514e-16   # ħ in eV s (reference src/include/constants.h:46)
CAPS_C = 299792458.0
KB, HBAR = 1.38064852e-23, 1.0545718e-34

DETALG = {"HODLR": 0, "LU": 1, "QR": 2, "CHOLESKY": 3}

_lib = None


def lib():
    """The loaded libcaps_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libcaps_amd.so is not built; run `python caps_amd/build.py` "
                               "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        d, i, p = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        sigs = {
            "caps_init": (p, [d, d]), "caps_free": (None, [p]),
            "caps_set_ldim": (i, [p, i]), "caps_get_ldim": (i, [p]),
            "caps_set_epsrel": (i, [p, d]), "caps_get_epsrel": (d, [p]),
            "caps_set_detalg": (i, [p, i]), "caps_get_detalg": (i, [p]),
            "caps_set_epsilonm1": (None, [p, p, p]),
            "caps_set_epsilonm1_plate": (None, [p, p, p]),
            "caps_set_epsilonm1_sphere": (None, [p, p, p]),
            "caps_epsilonm1_plate": (d, [p, d]), "caps_epsilonm1_sphere": (d, [p, d]),
            "caps_lnLambda": (d, [i, i, i]),
            "caps_estimate_lminmax": (i, [p, i, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
            "caps_mie": (None, [p, d, i, dp, dp]), "caps_mie_perf": (None, [p, d, i, dp, dp]),
            "caps_fresnel": (None, [p, d, d, dp, dp]),
            "caps_logdetD": (d, [p, d, i]),
            "caps_logdetD_batch": (i, [p, dp, ip, i, dp]),
            "capsx_logdetD_multi": (i, [ctypes.POINTER(p), dp, ip, i, dp]),
            "capsx_logdetD_multi_info": (i, [ctypes.POINTER(p), dp, ip, i, dp, dp, ip]),
            "capsx_mie_points": (i, [p, i, dp, ip, dp, dp]),
            "capsx_fresnel_points": (i, [p, i, dp, dp, dp, dp]),
            "capsx_bessel_points": (i, [i, ip, dp, dp, dp, dp, dp]),
            "caps_last_timing": (None, [p, dp]),
            "caps_M_init": (p, [p, i, d]), "caps_M_elem": (d, [p, i, i, ctypes.c_char, ctypes.c_char]),
            "caps_M_free": (None, [p]), "caps_kernel_M": (d, [i, i, p]),
            "caps_integrate_init": (p, [p, d, i, d]), "caps_integrate_free": (None, [p]),
            "caps_integrate_K": (d, [p, i, i, ctypes.POINTER(ctypes.c_byte)]),
            "caps_integrate_I": (d, [p, i, i, i, ctypes.POINTER(ctypes.c_byte)]),
            "caps_integrate_A": (d, [p, i, i, i, ctypes.POINTER(ctypes.c_byte)]),
            "caps_integrate_B": (d, [p, i, i, i, ctypes.POINTER(ctypes.c_byte)]),
            "caps_integrate_C": (d, [p, i, i, i, ctypes.POINTER(ctypes.c_byte)]),
            "caps_integrate_D": (d, [p, i, i, i, ctypes.POINTER(ctypes.c_byte)]),
            "caps_ht_drude": (d, [p]),
            "caps_free_energy": (d, [p, d, d, d, i, i, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]),
            "caps_logdetD0": (None, [p, i, d, dp, dp, dp]),
            "caps_ht_perf": (d, [p, d]),
            "capsx_lu_logdet_batch": (i, [dp, i, i, dp]),
            "capsx_logdet_1mM_batch": (i, [dp, i, i, dp, ip]),
            "capsx_mfma_probe": (i, [dp, dp, dp]),
            "kernel_logdet": (d, [i, p, p, i, i]),
            "capsx_lu_delta_mode": (i, [i]),
            "capsx_lu_time_form": (i, [i, i, i, i, i, i, i, i, dp]),
            "capsx_block_tables": (i, [p, d, i, ip, ip, dp, ip, ip, dp, ctypes.POINTER(ctypes.c_byte), i, i]),
            "capsx_block_nodes": (i, [p, d, i]),
            "capsx_M_copy": (i, [p, dp]),
            "capsx_stats": (None, [p, dp]), "capsx_stats_reset": (None, [p]),
            "capsx_profile": (None, [p, i]),
            "capsx_set_graph": (None, [p, i]),
            "capsx_lu_grid_panel_rows": (i, [i]),
            "capsx_lu_kernels": (i, [i, i, i, ip]),
            "capsx_surrogate_report": (None, [p, dp, i]),
            "caps_ht_plasma": (d, [p, d, d]),
            "capsx_material_init": (p, [ctypes.c_char_p, d]),
            "capsx_material_free": (None, [p]),
            "capsx_material_epsilonm1": (d, [d, p]),
            "capsx_material_get_extrapolation": (None, [p, dp, dp, dp, dp]),
            "capsx_resume_load": (i, [p, ctypes.c_char_p]),
            "capsx_resume_lookup": (i, [p, d, dp]),
            "capsx_set_output": (i, [p, ctypes.c_char_p]),
            "capsx_set_checkpoint": (i, [p, i, d]),
            "capsx_free_energy": (d, [p, d, d, d, i, i, d, i, ctypes.POINTER(ctypes.c_long),
                                      ctypes.POINTER(ctypes.c_long)]),
            "capsx_psd": (i, [i, dp, dp]),
            "capsx_npy_save": (i, [ctypes.c_char_p, dp, i]),
            "capsx_npy_load": (dp, [ctypes.c_char_p, ip]),
            "capsx_npy_free": (None, [dp]),
            "caps_integrate_plasma_init": (p, [p, d, d]),
            "caps_integrate_plasma": (d, [p, i, i, i, dp, dp]),
            "caps_integrate_plasma_free": (None, [p]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Caps:
    """A caps_t object (reference: src/include/libcaps.h:41-78)."""

    def __init__(self, R, L, ldim=None, epsrel=None, detalg=None):
        self.ptr = lib().caps_init(R, L)
        if not self.ptr:
            raise ValueError("caps_init: R and L must be positive")
        self.R, self.L = R, L
        self._ud = None
        if ldim is not None:
            self.set_ldim(ldim)
        if epsrel is not None:
            lib().caps_set_epsrel(self.ptr, epsrel)
        if detalg is not None:
            self.set_detalg(detalg)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().caps_free(self.ptr)
            self.ptr = None

    @property
    def ldim(self):
        return lib().caps_get_ldim(self.ptr)

    def set_ldim(self, ldim):
        if not lib().caps_set_ldim(self.ptr, ldim):
            raise ValueError("ldim must be >= 1")

    def set_detalg(self, name):
        lib().caps_set_detalg(self.ptr, DETALG[name] if isinstance(name, str) else int(name))

    def set_drude(self, omegap_eV, gamma_eV, body="both"):
        """Drude metal ε(iξ)-1 = ωp²/(ξ(ξ+γ)) (the reference's
        caps_epsilonm1_drude, userdata {ωp, γ} in rad/s)."""
        ud = (ctypes.c_double * 2)(omegap_eV / HBAR_EV, gamma_eV / HBAR_EV)
        self._ud = getattr(self, "_ud_list", []) + [ud]
        self._ud_list = self._ud
        fn = ctypes.cast(lib().caps_epsilonm1_drude, ctypes.c_void_p)
        setter = {"both": lib().caps_set_epsilonm1, "plate": lib().caps_set_epsilonm1_plate,
                  "sphere": lib().caps_set_epsilonm1_sphere}[body]
        setter(self.ptr, fn, ctypes.cast(ud, ctypes.c_void_p))

    def set_material(self, path, body="both"):
        """Tabulated ε(iξ) from a file in the reference's material format
        (capsx_material_init); linear interpolation, Drude extrapolation."""
        h = lib().capsx_material_init(str(path).encode(), self.R + self.L)
        if not h:
            raise ValueError("cannot read material file %s" % path)
        self._materials = getattr(self, "_materials", []) + [_Material(h)]
        fn = ctypes.cast(lib().capsx_material_epsilonm1, ctypes.c_void_p)
        setter = {"both": lib().caps_set_epsilonm1, "plate": lib().caps_set_epsilonm1_plate,
                  "sphere": lib().caps_set_epsilonm1_sphere}[body]
        setter(self.ptr, fn, ctypes.c_void_p(h))
        return self._materials[-1]

    def lminmax(self, m):
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        lib().caps_estimate_lminmax(self.ptr, m, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def logdetD(self, xi_, m):
        return lib().caps_logdetD(self.ptr, float(xi_), int(m))

    def logdetD_batch(self, xis, ms, info=False):
        """logdetD of the blocks (xis[i], ms[i]) in one batch.  info=True
        also returns each block's tr M and flags (capsx_logdetD_multi_info:
        bits 0-1 shortcut 0 LU / 1 trace / 2 Mercator / 3 non-finite, bit 2
        delta-form LU, bit 3 HODLR-mode rule gave -tr M)."""
        xis = np.ascontiguousarray(xis, dtype=np.float64)
        ms = np.ascontiguousarray(ms, dtype=np.int32)
        if xis.shape != ms.shape:
            raise ValueError("xi_ and m must have the same length")
        if info:
            return logdetD_multi([self] * len(xis), xis, ms, info=True)
        out = np.empty(len(xis))
        rc = lib().caps_logdetD_batch(self.ptr, _dptr(xis), ms.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                      len(xis), _dptr(out))
        if rc != 0:
            raise ValueError("caps_logdetD_batch rejected its input")
        return out

    def last_timing(self):
        t = np.zeros(6)
        lib().caps_last_timing(self.ptr, _dptr(t))
        return dict(zip(["mie", "K", "I", "assemble", "lu", "total"], t))

    def stats(self, reset=False):
        """Accumulated stage times and work counters of this object's batches
        (the Mie kernel runs beside the K quadrature: "K" is the time the K
        stage adds after the Mie kernel ends)."""
        a = np.zeros(16)
        lib().capsx_stats(self.ptr, _dptr(a))
        if reset:
            lib().capsx_stats_reset(self.ptr)
        names = ["mie", "K", "I", "assemble", "lu", "total"]
        return {"stage_ms": dict(zip(names, a[:6])), "blocks": int(a[6]), "blocks_factored": int(a[7]),
                "lu_flops": a[8], "batches": int(a[9]), "I_integrals": a[10],
                "gemm_ms": a[11], "gemm_flops": a[12], "gemm_launches": int(a[13]), "gemm_bytes": a[14],
                "other_update_ms": a[15]}

    def profile(self, on=True):
        """Time every launch of the LU's trailing-update kernel with HIP events."""
        lib().capsx_profile(self.ptr, 1 if on else 0)

    def set_graph(self, on=True):
        """Capture each batch shape into a HIP graph and replay it (on), or
        launch eagerly (off, the default)."""
        lib().capsx_set_graph(self.ptr, 1 if on else 0)

    def reset_stats(self):
        lib().capsx_stats_reset(self.ptr)

    def mie(self, xi_, l):
        a, b = ctypes.c_double(), ctypes.c_double()
        lib().caps_mie(self.ptr, xi_, l, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def fresnel(self, xi_, k_):
        a, b = ctypes.c_double(), ctypes.c_double()
        lib().caps_fresnel(self.ptr, xi_, k_, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def mie_points(self, xis, ls):
        """log|a_l|, log|b_l| at the points (xis[i], ls[i]) in one launch."""
        xis = np.ascontiguousarray(xis, dtype=np.float64)
        ls = np.ascontiguousarray(ls, dtype=np.int32)
        a, b = np.empty(len(xis)), np.empty(len(xis))
        if lib().capsx_mie_points(self.ptr, len(xis), _dptr(xis), ls.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                  _dptr(a), _dptr(b)) != 0:
            raise ValueError("capsx_mie_points rejected its input")
        return a, b

    def fresnel_points(self, xis, ks):
        """r_TE, r_TM at the points (xis[i], ks[i]) in one launch."""
        xis = np.ascontiguousarray(xis, dtype=np.float64)
        ks = np.ascontiguousarray(ks, dtype=np.float64)
        te, tm = np.empty(len(xis)), np.empty(len(xis))
        if lib().capsx_fresnel_points(self.ptr, len(xis), _dptr(xis), _dptr(ks), _dptr(te), _dptr(tm)) != 0:
            raise ValueError("capsx_fresnel_points rejected its input")
        return te, tm

    def matrix(self, xi_, m):
        """The round-trip matrix M (N x N, N = 2 ldim) and lmin, via caps_M_init."""
        M = lib().caps_M_init(self.ptr, int(m), float(xi_))
        N = 2 * self.ldim
        out = np.empty(N * N)
        lib().capsx_M_copy(M, _dptr(out))
        lmin = ctypes.cast(M, ctypes.POINTER(ctypes.c_int))[3]   # caps_M_t: caps*, m, lmin
        obj = _MObj(M)
        return out.reshape(N, N, order="F"), lmin, obj

    def tables(self, xi_, m):
        """K table (log|K_ν^p|, ν = mu..) and I table (log|I|, sign) of a block."""
        mu, nK, lo, nl = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib().capsx_block_tables(self.ptr, xi_, m, ctypes.byref(mu), ctypes.byref(nK), None, ctypes.byref(lo),
                                 ctypes.byref(nl), None, None, 0, 0)
        lnK = np.empty(2 * nK.value)
        logI = np.empty(2 * nl.value * nl.value)
        sgn = np.empty(2 * nl.value * nl.value, dtype=np.int8)
        rc = lib().capsx_block_tables(self.ptr, xi_, m, ctypes.byref(mu), ctypes.byref(nK), _dptr(lnK),
                                      ctypes.byref(lo), ctypes.byref(nl), _dptr(logI),
                                      sgn.ctypes.data_as(ctypes.POINTER(ctypes.c_byte)), nK.value, nl.value)
        assert rc == 0
        return {"mu": mu.value, "lnK": lnK.reshape(2, nK.value), "lo": lo.value,
                "logI": logI.reshape(2, nl.value, nl.value), "sgnI": sgn.reshape(2, nl.value, nl.value)}

    def nodes(self, xi_, m):
        return lib().capsx_block_nodes(self.ptr, xi_, m)

    def ht_drude(self):
        return lib().caps_ht_drude(self.ptr)

    def free_energy_ex(self, T, cutoff=1e-9, epsrel=1e-6, mode="batched", xi0="drude", omegap=0.0, psd_order=0):
        """Free energy with the reference driver's options (capsx_free_energy):
        xi0 in ("drude", "pr", "plasma") selects the ξ = 0 term (plasma:
        omegap in rad/s); psd_order 0 = Matsubara sum, > 0 = Padé spectrum
        decomposition of that order, < 0 = automatic order.  Returns (F,
        determinants computed, determinants entering the sums)."""
        nb, nu = ctypes.c_long(), ctypes.c_long()
        F = lib().capsx_free_energy(self.ptr, T, cutoff, epsrel, {"plain": 0, "batched": 1, "surrogate": 2}[mode],
                                    {"drude": 1, "pr": 2, "plasma": 3}[xi0], omegap, psd_order,
                                    ctypes.byref(nb), ctypes.byref(nu))
        return F, nb.value, nu.value

    def free_energy(self, T, cutoff=1e-9, epsrel=1e-6, mode="batched", drude_xi0=True):
        """F (L+R)/(ħc) at temperature T with the reference driver's sums
        (src/caps.c:235-396, 722-850); T = 0 integrates over ξ with the
        Fourier-Chebyshev rule on [0, ∞) to relative accuracy epsrel.
        Returns (F, determinants computed, determinants entering the sums).
        mode: "plain" (one block at a time), "batched" (planned from the
        first Matsubara terms, or one batch per quadrature level) or
        "surrogate" (planned by an attached caps_amd.surrogate.MLP: at T > 0
        the m-cutoffs and end of the Matsubara sum, at T = 0 the m-cutoff of
        every quadrature node)."""
        nb, nu = ctypes.c_long(), ctypes.c_long()
        F = lib().caps_free_energy(self.ptr, T, cutoff, epsrel, {"plain": 0, "batched": 1, "surrogate": 2}[mode],
                                   1 if drude_xi0 else 0, ctypes.byref(nb), ctypes.byref(nu))
        return F, nb.value, nu.value

    def resume_load(self, path):
        """Read the (ξ(L+R)/c, F_ξ) lines of an earlier run's output into the
        resume cache (path None clears it); returns the number read
        (capsx_resume_load; ValueError for a missing or malformed file)."""
        n = lib().capsx_resume_load(self.ptr, None if path is None else os.fsencode(path))
        if n < 0:
            raise ValueError("%s: %s" % (path, "cannot open" if n == -1 else "wrong format"))
        return n

    def resume_lookup(self, xi_):
        """F_ξ from the resume cache, or None."""
        v = ctypes.c_double()
        return v.value if lib().capsx_resume_lookup(self.ptr, xi_, ctypes.byref(v)) else None

    def set_output(self, path):
        """Append a "# xi*(L+R)/c=..., logdetD=..., t=..." line per computed
        F_ξ to path (None closes it)."""
        if lib().capsx_set_output(self.ptr, None if path is None else os.fsencode(path)) != 0:
            raise ValueError("cannot open %s" % path)

    def set_checkpoint(self, nodes=0, seconds=0.0):
        """T = 0 free energy in batches of at most `nodes` quadrature nodes
        (0: a refinement level's new nodes at once), each F_ξ written to the
        output file as its batch completes; with seconds > 0 the driver stops
        before its next batch once that time has passed and free_energy
        returns NaN (continue with resume_load of the output file)."""
        if lib().capsx_set_checkpoint(self.ptr, int(nodes), float(seconds)) != 0:
            raise ValueError("nodes and seconds must be >= 0")

    def surrogate_report(self, reset=True):
        """Surrogate accuracy on the blocks computed for its plans (mode
        "surrogate") since the last reset."""
        out = np.zeros(6)
        lib().capsx_surrogate_report(self.ptr, _dptr(out), 1 if reset else 0)
        return {"blocks": int(out[0]), "rms_log_error": out[1], "max_log_error": out[2],
                "frequencies": int(out[3]), "mean_mcut_error": out[4], "max_mcut_error": int(out[5])}

    def logdetD0(self, m, omegap=None):
        """ξ = 0 EE and MM blocks (perfect reflector / Drude); with omegap
        (rad/s) also the plasma-model MM block: (EE, MM, MM_plasma)."""
        ee, mm, mp = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        lib().caps_logdetD0(self.ptr, m, omegap or 0.0, ctypes.byref(ee), ctypes.byref(mm),
                            ctypes.byref(mp) if omegap else None)
        return (ee.value, mm.value, mp.value) if omegap else (ee.value, mm.value)

    def ht_perf(self, eps=1e-13):
        return lib().caps_ht_perf(self.ptr, eps)

    def ht_plasma(self, omegap, eps=1e-13):
        """ξ = 0 plasma-model free energy (k_B T units), omegap in rad/s."""
        return lib().caps_ht_plasma(self.ptr, omegap, eps)

    def plasma_integrals(self, omegap, l1, l2, m=0):
        """(I(l1 + l2), I_{l1-1/2}/I_{l1+1/2}, I_{l2-1/2}/I_{l2+1/2}) at α =
        ωp (L+R)/c/(1 + L/R) (host computation)."""
        L = lib()
        h = L.caps_integrate_plasma_init(self.ptr, omegap, 1e-8)
        r1, r2 = ctypes.c_double(), ctypes.c_double()
        v = L.caps_integrate_plasma(h, l1, l2, m, ctypes.byref(r1), ctypes.byref(r2))
        L.caps_integrate_plasma_free(h)
        return v, r1.value, r2.value


class _Material:
    """A loaded material table (freed with the object that holds it)."""

    def __init__(self, ptr):
        self.ptr = ptr

    def epsilonm1(self, xi):
        return lib().capsx_material_epsilonm1(xi, self.ptr)

    def extrapolation(self):
        v = [ctypes.c_double() for _ in range(4)]
        lib().capsx_material_get_extrapolation(self.ptr, *[ctypes.byref(x) for x in v])
        return tuple(x.value for x in v)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().capsx_material_free(self.ptr)
            self.ptr = None


class _MObj:
    def __init__(self, ptr):
        self.ptr = ptr

    def elem(self, l1, l2, p1, p2):
        return lib().caps_M_elem(self.ptr, l1, l2, p1.encode(), p2.encode())

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().caps_M_free(self.ptr)
            self.ptr = None


def psd(N):
    """Poles ξ_j and residues η_j of the Padé spectrum decomposition of
    order N (capsx_psd)."""
    xi, eta = np.empty(N), np.empty(N)
    if lib().capsx_psd(N, _dptr(xi), _dptr(eta)) != 0:
        raise ValueError("invalid order %d" % N)
    return xi, eta


def npy_save(path, M):
    """Write a square matrix as .npy in the reference's byte layout
    (element (i, j) at i*dim + j under a fortran_order header)."""
    M = np.ascontiguousarray(M, dtype=np.float64)
    if M.ndim != 2 or M.shape[0] != M.shape[1]:
        raise ValueError("square matrix required")
    if lib().capsx_npy_save(str(path).encode(), _dptr(M), M.shape[0]) != 0:
        raise OSError("cannot write %s" % path)


def npy_load(path):
    """Read a matrix file of that layout; returns the dim x dim array with
    element (i, j) = file element i*dim + j (numpy.load gives its
    transpose)."""
    dim = ctypes.c_int()
    p = lib().capsx_npy_load(str(path).encode(), ctypes.byref(dim))
    if not p:
        raise ValueError("%s is not a square '<f8' fortran-order .npy file" % path)
    n = dim.value
    out = np.ctypeslib.as_array(p, shape=(n * n,)).copy().reshape(n, n)
    lib().capsx_npy_free(p)
    return out


def bessel_points(orders, nus, xs):
    """log I_{n+1/2}(x), log K_{n+1/2}(x) (orders[i] = n, < 0 skipped) and
    I_ν(x)/I_{ν+1}(x) (nus[i] = ν, < 0 skipped) at the points xs, in one
    launch of the device functions."""
    orders = np.ascontiguousarray(orders, dtype=np.int32)
    nus = np.ascontiguousarray(nus, dtype=np.float64)
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    n = len(xs)
    lI, lK, r = np.empty(n), np.empty(n), np.empty(n)
    if lib().capsx_bessel_points(n, orders.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dptr(nus), _dptr(xs),
                                 _dptr(lI), _dptr(lK), _dptr(r)) != 0:
        raise ValueError("capsx_bessel_points rejected its input")
    return lI, lK, r


def logdetD_multi(objs, xis, ms, info=False):
    """logdetD of blocks i = (objs[i], xis[i], ms[i]) in one batched launch
    sequence; the objects may differ in R/L, ldim and material (uneven
    dimensions 2 ldim).  Runs on objs[0]'s device context, whose stage
    timings (last_timing) then cover the whole batch."""
    xis = np.ascontiguousarray(xis, dtype=np.float64)
    ms = np.ascontiguousarray(ms, dtype=np.int32)
    if not (len(objs) == len(xis) == len(ms)):
        raise ValueError("objs, xi_ and m must have the same length")
    ptrs = (ctypes.c_void_p * len(objs))(*[o.ptr for o in objs])
    out = np.empty(len(xis))
    if info:
        tr = np.empty(len(xis))
        flags = np.zeros(len(xis), dtype=np.int32)
        rc = lib().capsx_logdetD_multi_info(ptrs, _dptr(xis), ms.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                            len(xis), _dptr(out), _dptr(tr),
                                            flags.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    else:
        rc = lib().capsx_logdetD_multi(ptrs, _dptr(xis), ms.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                       len(xis), _dptr(out))
    if rc != 0:
        raise ValueError("capsx_logdetD_multi rejected its input")
    return (out, tr, flags) if info else out


def lu_logdet_batch(mats):
    """log|det A| of square matrices on the GPU's batched LU (test hook)."""
    mats = np.asarray(mats, dtype=np.float64)
    n, N, _ = mats.shape
    buf = np.ascontiguousarray(np.transpose(mats, (0, 2, 1)))   # column-major per matrix
    out = np.empty(n)
    if lib().capsx_lu_logdet_batch(_dptr(buf), N, n, _dptr(out)) != 0:
        raise ValueError("bad LU batch")
    return out


def logdet_1mM_batch(Ms):
    """log|det(1 - M)| of square matrices M through the device's column
    dominance test and LU (test hook): returns (values, delta flags), flag 1
    where 1 - M is column diagonally dominant and the LU ran in delta form
    (no pivot search, log1p of the pivots' deviations from 1)."""
    Ms = np.asarray(Ms, dtype=np.float64)
    n, N, _ = Ms.shape
    buf = np.ascontiguousarray(np.transpose(Ms, (0, 2, 1)))   # column-major per matrix
    out = np.empty(n)
    flags = np.zeros(n, dtype=np.int32)
    if lib().capsx_logdet_1mM_batch(_dptr(buf), N, n, _dptr(out), flags.ctypes.data_as(ctypes.POINTER(ctypes.c_int))) != 0:
        raise ValueError("bad batch")
    return out, flags


KERNEL_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p)


def kernel_logdet(dim, kernel, detalg="HODLR", sym_spd=1):
    """log det(1 - M) of the dim x dim matrix M_ij = kernel(i, j) through the
    library's kernel_logdet (host callback, as src/capc.cpp:110-114 calls it)."""
    cb = KERNEL_FN(lambda i, j, _a: kernel(i, j))
    return lib().kernel_logdet(dim, ctypes.cast(cb, ctypes.c_void_p), None, sym_spd,
                               DETALG[detalg] if isinstance(detalg, str) else int(detalg))


def mfma_probe(A, B):
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    D = np.empty((16, 16))
    lib().capsx_mfma_probe(_dptr(A), _dptr(B), _dptr(D))
    return D


def T_scaled(R, L, T):
    """2π k_B (R+L) T/(ħc): the Matsubara step in units of c/(L+R)."""
    import math
    return 2 * math.pi * KB * (R + L) * T / (HBAR * CAPS_C)
