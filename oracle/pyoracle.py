"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py for its cpu_baseline leg.  Never used by the
product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class Geom(C.Structure):
    _fields_ = [
        ("L", C.c_int), ("LT", C.c_int), ("V", C.c_int), ("half", C.c_int),
        ("lex2f", C.POINTER(C.c_int)), ("f2lex", C.POINTER(C.c_int)),
        ("coords", C.POINTER(C.c_int)), ("fpp", C.POINTER(C.c_int)),
        ("fmm", C.POINTER(C.c_int)), ("eta_pp", C.POINTER(C.c_double)),
        ("eta_mm", C.POINTER(C.c_double)), ("eta_pbc", C.POINTER(C.c_double)),
    ]


class Rng(C.Structure):
    _fields_ = [("st", C.c_uint64 * ((1269 + 1) * 2)), ("idx", C.c_int)]


class CGStat(C.Structure):
    _fields_ = [("iterations", C.c_double), ("cg_call", C.c_int), ("last_iters", C.c_int)]


class Kit(C.Structure):
    _fields_ = [
        ("V", C.c_int), ("work", C.POINTER(C.c_double)),
        ("chi", C.POINTER(C.c_double) * 2), ("phi", C.POINTER(C.c_double) * 2),
        ("inv_phi", C.POINTER(C.c_double) * 2), ("inv_phi_obs", C.POINTER(C.c_double) * 2),
        ("ddag", C.POINTER(C.c_double)), ("vec", C.POINTER(C.c_double)),
        ("st", CGStat),
    ]


class Traj(C.Structure):
    _fields_ = [("dH", C.c_double), ("H0", C.c_double), ("H1", C.c_double),
                ("accepted", C.c_int), ("ok", C.c_int)]


def build():
    subprocess.run(["make", "-C", _HERE, "liboracle.so"], check=True,
                   stdout=subprocess.DEVNULL)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        G = C.POINTER(Geom)
        R = C.POINTER(Rng)
        K = C.POINTER(Kit)
        L.orc_geom_init.argtypes = [G, C.c_int, C.c_int]
        L.orc_geom_free.argtypes = [G]
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        L.orc_setup_tables_by_search.argtypes = [G, ip, ip]
        L.orc_rng_init.argtypes = [R, C.c_uint32]
        L.orc_casuale.argtypes = [R]
        L.orc_casuale.restype = C.c_double
        L.orc_gauss1.argtypes = [R]
        L.orc_gauss1.restype = C.c_double
        L.orc_gauss2mv.argtypes = [R, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double, C.c_double]
        L.orc_fill_casuale.argtypes = [R, _dp, C.c_long]
        L.orc_draw_sigma.argtypes = [R, C.c_int, C.c_double, _dp]
        L.orc_draw_z2_complex.argtypes = [R, C.c_int, _dp]
        L.orc_draw_z2_real.argtypes = [R, C.c_int, _dp]
        for n in ("orc_dirac", "orc_adj_dirac", "orc_sub_dirac", "orc_sub_adj_dirac"):
            getattr(L, n).argtypes = [G, _dp, _dp, _dp]
        L.orc_sigma_mass.argtypes = [G, _dp, _dp]
        L.orc_norm.argtypes = [C.c_int, _dp]
        L.orc_norm.restype = C.c_double
        for n in ("orc_cg_null", "orc_cg_imposed"):
            getattr(L, n).argtypes = [G, _dp, _dp, _dp, C.c_double, C.c_long, _dp, C.POINTER(CGStat)]
            getattr(L, n).restype = C.c_bool
        L.orc_kit_alloc.argtypes = [K, C.c_int]
        L.orc_kit_free.argtypes = [K]
        L.orc_sigma_force.argtypes = [G, _dp, _dp, C.c_double, C.c_double, C.c_long, K, _dp]
        L.orc_sigma_force.restype = C.c_bool
        L.orc_sigma_force_from_solutions.argtypes = [G, _dp, C.c_double, C.POINTER(C.c_double) * 2,
                                                     C.POINTER(C.c_double) * 2, _dp]
        L.orc_energy.argtypes = [G, _dp, _dp, C.c_double, C.POINTER(C.c_double) * 2, C.POINTER(C.c_double) * 2]
        L.orc_energy.restype = C.c_double
        L.orc_lfmn2vf.argtypes = [G, R, _dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int,
                                  C.c_double, C.c_double, C.c_long, K, C.POINTER(Traj)]
        L.orc_condensate.argtypes = [G, R, _dp, C.c_double, C.c_long, K, C.POINTER(C.c_double), _dp]
        L.orc_condensate_sigma.argtypes = [G, _dp, _dp]
        for n in ("orc_link_dirac", "orc_link_adj_dirac", "orc_link_sub_dirac"):
            getattr(L, n).argtypes = [G, _dp, _dp, _dp, _dp]
        for n in ("orc_link_cg_null", "orc_link_cg_imposed"):
            getattr(L, n).argtypes = [G, _dp, _dp, _dp, _dp, C.c_double, C.c_long, _dp, C.POINTER(CGStat)]
            getattr(L, n).restype = C.c_bool
        L.orc_plaquette_sum.argtypes = [G, _dp]
        L.orc_plaquette_sum.restype = C.c_double
        L.orc_gauge_force.argtypes = [G, _dp, C.c_double, _dp]
        L.orc_link_force_from_solutions.argtypes = [G, _dp, C.POINTER(C.c_double) * 2, C.POINTER(C.c_double) * 2,
                                                    _dp]
        L.orc_link_force.argtypes = [G, _dp, _dp, C.c_double, C.c_int, C.c_int, C.c_double, C.c_long, K, _dp]
        L.orc_link_force.restype = C.c_bool
        L.orc_link_energy.argtypes = [G, _dp, _dp, C.c_double, C.c_int, C.POINTER(C.c_double) * 2,
                                      C.POINTER(C.c_double) * 2]
        L.orc_link_energy.restype = C.c_double
        L.orc_leapfrog2link.argtypes = [G, R, _dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int, C.c_double,
                                        C.c_long, C.c_int, C.c_int, K, C.POINTER(Traj)]
        _lib = L
    return _lib


class Lattice:
    """Geometry of an L^2 x LT lattice (closed-form tables)."""

    def __init__(self, L, LT=None):
        self.g = Geom()
        rc = lib().orc_geom_init(C.byref(self.g), L, LT or L)
        if rc != 0:
            raise ValueError(f"bad lattice {L}x{LT}")
        self.L, self.LT, self.V = self.g.L, self.g.LT, self.g.V

    def __del__(self):
        try:
            lib().orc_geom_free(C.byref(self.g))
        except Exception:
            pass

    def table(self, name, n):
        p = getattr(self.g, name)
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    def tables_by_search(self):
        """The reference's O(V^2) set-up of the neighbour tables (fpp, fmm)."""
        fpp = np.zeros(3 * self.V, np.int32)
        fmm = np.zeros(3 * self.V, np.int32)
        lib().orc_setup_tables_by_search(C.byref(self.g), fpp, fmm)
        return fpp, fmm

    def sigma_mass(self, sigma):
        sm = np.empty(self.V)
        lib().orc_sigma_mass(C.byref(self.g), np.ascontiguousarray(sigma, np.float64), sm)
        return sm

    def _op(self, name, sm, v, out=None):
        v = as_pairs(v)
        if out is None:
            out = np.zeros(2 * self.V)
        else:
            out = as_pairs(out).copy()
        getattr(lib(), name)(C.byref(self.g), np.ascontiguousarray(sm, np.float64), v, out)
        return out

    def M(self, sm, v):
        return self._op("orc_dirac", sm, v)

    def Mdag(self, sm, v):
        return self._op("orc_adj_dirac", sm, v)

    def sub_M(self, sm, v, target):
        return self._op("orc_sub_dirac", sm, v, target)

    def sub_Mdag(self, sm, v, target):
        return self._op("orc_sub_adj_dirac", sm, v, target)

    def cg(self, sm, b, x0=None, tol=1e-10, max_iter=None, start="null"):
        b = as_pairs(b)
        x = np.zeros(2 * self.V) if x0 is None else as_pairs(x0).copy()
        work = np.zeros(8 * self.V)
        st = CGStat()
        fn = lib().orc_cg_null if start == "null" else lib().orc_cg_imposed
        ok = fn(C.byref(self.g), np.ascontiguousarray(sm, np.float64), b, x, tol,
                max_iter if max_iter is not None else self.V + 1, work, C.byref(st))
        return x, bool(ok), st


    # -- link sector (link-coupled operator of copy_dirac.c; U lex [V][3]) --
    def _lop(self, name, sm, U, v, out=None):
        v = as_pairs(v)
        out = np.zeros(2 * self.V) if out is None else as_pairs(out).copy()
        getattr(lib(), name)(C.byref(self.g), np.ascontiguousarray(sm, np.float64), as_pairs(U), v, out)
        return out

    def link_M(self, sm, U, v):
        return self._lop("orc_link_dirac", sm, U, v)

    def link_Mdag(self, sm, U, v):
        return self._lop("orc_link_adj_dirac", sm, U, v)

    def link_sub_M(self, sm, U, v, target):
        return self._lop("orc_link_sub_dirac", sm, U, v, target)

    def link_cg(self, sm, U, b, x0=None, tol=1e-10, max_iter=None, start="null"):
        b = as_pairs(b)
        x = np.zeros(2 * self.V) if x0 is None else as_pairs(x0).copy()
        work = np.zeros(8 * self.V)
        st = CGStat()
        fn = lib().orc_link_cg_null if start == "null" else lib().orc_link_cg_imposed
        ok = fn(C.byref(self.g), np.ascontiguousarray(sm, np.float64), as_pairs(U), b, x, tol,
                max_iter if max_iter is not None else self.V + 1, work, C.byref(st))
        return x, bool(ok), st

    def plaquette_sum(self, U):
        return lib().orc_plaquette_sum(C.byref(self.g), as_pairs(U))

    def gauge_force(self, U, beta):
        F = np.zeros(3 * self.V)
        lib().orc_gauge_force(C.byref(self.g), as_pairs(U), beta, F)
        return F

    def link_force_from_solutions(self, U, X, Y, force=None):
        F = np.zeros(3 * self.V) if force is None else force
        dp = C.POINTER(C.c_double)
        Xs = [as_pairs(a) for a in X]
        Ys = [as_pairs(a) for a in Y]
        lib().orc_link_force_from_solutions(C.byref(self.g), as_pairs(U),
                                            (dp * 2)(*[a.ctypes.data_as(dp) for a in Xs]),
                                            (dp * 2)(*[a.ctypes.data_as(dp) for a in Ys]), F)
        return F


def random_links(V, seed):
    """U(1) links exp(i theta), theta uniform in [0, 2 pi), lex order [V][3]."""
    th = np.random.default_rng(seed).uniform(0.0, 2 * np.pi, 3 * V)
    return np.exp(1j * th)


def as_pairs(v):
    """complex128[V] or float64[2V] -> contiguous float64[2V] (re, im interleaved)."""
    a = np.asarray(v)
    if np.iscomplexobj(a):
        a = np.ascontiguousarray(a, np.complex128).view(np.float64)
    return np.ascontiguousarray(a, np.float64).reshape(-1)


def as_complex(v):
    return np.ascontiguousarray(v, np.float64).view(np.complex128)


class Random:
    """Host dSFMT-132049 stream with the reference's draw helpers."""

    def __init__(self, seed):
        self.r = Rng()
        lib().orc_rng_init(C.byref(self.r), seed)

    def casuale(self, n=None):
        if n is None:
            return lib().orc_casuale(C.byref(self.r))
        out = np.empty(n)
        lib().orc_fill_casuale(C.byref(self.r), out, n)
        return out

    def gauss1(self):
        return lib().orc_gauss1(C.byref(self.r))

    def gauss2mv(self, mean, var):
        a, b = C.c_double(), C.c_double()
        lib().orc_gauss2mv(C.byref(self.r), C.byref(a), C.byref(b), mean, var)
        return a.value, b.value

    def sigma(self, V, var):
        s = np.empty(V)
        lib().orc_draw_sigma(C.byref(self.r), V, var, s)
        return s

    def z2_complex(self, V):
        v = np.empty(2 * V)
        lib().orc_draw_z2_complex(C.byref(self.r), V, v)
        return v

    def z2_real(self, V):
        v = np.empty(2 * V)
        lib().orc_draw_z2_real(C.byref(self.r), V, v)
        return v


class LinkHMC:
    """Serial link HMC (LeapFrog2Link) at fixed sigma, driving the oracle."""

    def __init__(self, lat, seed, sm, U, beta=1.0, pure_gauge=True, linked=True):
        self.lat = lat
        self.rng = Random(seed)
        self.sm = np.ascontiguousarray(sm, np.float64)
        self.U = as_pairs(np.asarray(U, np.complex128)).copy()
        self.beta = beta
        self.pure_gauge = int(pure_gauge)
        self.linked = int(linked)
        self.mom = np.zeros(3 * lat.V)
        self.force = np.zeros(3 * lat.V)
        self.kit = Kit()
        lib().orc_kit_alloc(C.byref(self.kit), lat.V)

    def __del__(self):
        try:
            lib().orc_kit_free(C.byref(self.kit))
        except Exception:
            pass

    def trajectory(self, eps, steps, tol, max_iter=None):
        t = Traj()
        lib().orc_leapfrog2link(C.byref(self.lat.g), C.byref(self.rng.r), self.sm, self.U, self.mom, self.force,
                                self.beta, eps, steps, tol, max_iter if max_iter is not None else self.lat.V + 1,
                                self.pure_gauge, self.linked, C.byref(self.kit), C.byref(t))
        return t

    def links(self):
        return as_complex(self.U).copy()


def sigma_variance(beta_s, nflav=4):
    """Variance 2/(N_f beta_s) of the Gaussian sigma draw used by every workload."""
    return 2.0 / (nflav * beta_s)


class HMC:
    """Serial sigma-HMC (LFMN2VF) with measurements, driving the oracle."""

    def __init__(self, lat, seed, beta_s=0.85):
        self.lat = lat
        self.rng = Random(seed)
        self.beta_s = beta_s
        V = lat.V
        self.sigma = self.rng.sigma(V, sigma_variance(beta_s))
        self.sm = lat.sigma_mass(self.sigma)
        self.mom = np.zeros(V)
        self.force = np.zeros(V)
        self.kit = Kit()
        lib().orc_kit_alloc(C.byref(self.kit), V)

    def __del__(self):
        try:
            lib().orc_kit_free(C.byref(self.kit))
        except Exception:
            pass

    def trajectory(self, eps, steps, tol_high, tol_medium, max_iter=None):
        t = Traj()
        lib().orc_lfmn2vf(C.byref(self.lat.g), C.byref(self.rng.r), self.sigma, self.sm, self.mom,
                          self.force, self.beta_s, eps, steps, tol_high, tol_medium,
                          max_iter if max_iter is not None else self.lat.V + 1,
                          C.byref(self.kit), C.byref(t))
        return t

    def condensate(self, tol, max_iter=None):
        out = np.zeros(11)
        lib().orc_condensate(C.byref(self.lat.g), C.byref(self.rng.r), self.sm, tol,
                             max_iter if max_iter is not None else self.lat.V + 1,
                             C.byref(self.kit), self.kit.inv_phi_obs[0], out)
        return out

    def condensate_sigma(self):
        out = np.zeros(8)
        lib().orc_condensate_sigma(C.byref(self.lat.g), self.sigma, out)
        return out
