"""Python host binding of libgnx (include/gnx.h) over ctypes.

Device memory and streams come from PyTorch (ROCm build); every call passes
raw device pointers to the C ABI.  There is no CPU fallback: if the HIP
library is missing or no GPU is visible, construction raises.

Complex vectors are torch.float64 tensors of 2*V elements (re, im
interleaved) or torch.complex128 tensors of V elements, in the reference's
f-order (coordinate.c:301-313).
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgnx.so")

OP_M, OP_MDAG, OP_MMDAG, OP_MDAGM, OP_SUB_M, OP_SUB_MDAG = range(6)
START_NULL, START_IMPOSED = 0, 1
PREC_F64, PREC_F32, PREC_BF16, PREC_F16 = 0, 1, 2, 3
ABI_VERSION = 9  # GNX_ABI_VERSION of include/gnx.h this binding follows

STATUS = {0: "OK", 1: "ERR_ARG", 2: "ERR_HIP", 3: "ERR_NOCONV", 4: "ERR_SIZE", 5: "ERR_STATE"}


class GnxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"gnx status {STATUS.get(status, status)}: {msg}")
        self.status = status


class CGResult(C.Structure):
    _fields_ = [("converged", C.c_int), ("early_exit", C.c_int), ("iterations", C.c_int),
                ("reliable_updates", C.c_int), ("residual", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TrajResult(C.Structure):
    _fields_ = [("dH", C.c_double), ("H0", C.c_double), ("H1", C.c_double), ("ok", C.c_int),
                ("cg_solves", C.c_int), ("cg_iterations", C.c_long)]


_lib = None


def lib():
    """Load libgnx.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `make -C {_HERE}` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.gnx_abi_version.restype = C.c_int
        if L.gnx_abi_version() != ABI_VERSION:
            raise RuntimeError(f"{LIB_PATH} implements ABI version {L.gnx_abi_version()}, this binding "
                               f"needs {ABI_VERSION}: rebuild it with `make -C {_HERE}`")
        L.gnx_rng_create.argtypes = [C.c_uint, C.POINTER(vp)]
        L.gnx_rng_destroy.argtypes = [vp]
        L.gnx_rng_casuale.argtypes = [vp]
        L.gnx_rng_casuale.restype = C.c_double
        L.gnx_rng_gauss1.argtypes = [vp]
        L.gnx_rng_gauss1.restype = C.c_double
        L.gnx_rng_gauss2mv.argtypes = [vp, C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.gnx_rng_fill.argtypes = [vp, C.c_int, C.c_long, C.c_double, C.POINTER(C.c_double)]
        L.gnx_set_autotune.argtypes = [vp, C.c_int]
        L.gnx_tune_runs.argtypes = [vp, C.POINTER(C.c_long), C.POINTER(C.c_long)]
        L.gnx_replay_stats.argtypes = [vp, C.POINTER(C.c_long), C.POINTER(C.c_long)]
        L.gnx_last_error.restype = C.c_char_p
        L.gnx_create.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(vp)]
        L.gnx_destroy.argtypes = [vp]
        L.gnx_set_stream.argtypes = [vp, vp]
        L.gnx_synchronize.argtypes = [vp]
        L.gnx_volume.argtypes = [vp]
        L.gnx_volume.restype = C.c_long
        L.gnx_geometry_tables.argtypes = [vp, vp, vp, vp, vp, vp]
        L.gnx_sigma_mass.argtypes = [vp, vp, vp]
        L.gnx_dirac.argtypes = [vp, C.c_int, vp, vp, vp]
        L.gnx_dirac_batch.argtypes = [vp, C.c_int, vp, vp, vp, C.c_int]
        L.gnx_cg.argtypes = [vp, vp, vp, vp, C.c_double, C.c_long, C.c_int, C.c_int, C.POINTER(CGResult)]
        L.gnx_cg_noalloc.argtypes = [vp, vp, vp, vp, C.c_double, C.c_long, C.c_int, C.POINTER(CGResult)]
        L.gnx_cg_batch.argtypes = [vp, vp, vp, vp, C.c_int, C.c_double, C.c_long, C.c_int, C.c_int,
                                   C.POINTER(CGResult)]
        L.gnx_set_reliable_delta.argtypes = [vp, C.c_double]
        L.gnx_condensate_batch.argtypes = [vp, vp, vp, C.c_int, C.c_double, C.c_long, C.c_int, vp,
                                           C.POINTER(C.c_double), C.POINTER(CGResult)]
        L.gnx_cg_shape.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.gnx_cg_mixed_shape.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                         C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.gnx_sigma_force.argtypes = [vp, vp, vp, C.c_double, C.c_int, C.c_double, C.c_long,
                                      vp * 2, vp * 2, vp * 2, vp, CGResult * 2]
        L.gnx_sigma_force_from_solutions.argtypes = [vp, vp, C.c_double, C.c_int, vp, vp, vp, vp, vp]
        L.gnx_evolve_momenta.argtypes = [vp, vp, vp, C.c_double]
        L.gnx_evolve_sigma.argtypes = [vp, vp, vp, vp, C.c_double]
        L.gnx_energy.argtypes = [vp, vp, vp, C.c_double, C.c_int, vp * 2, vp * 2, C.POINTER(C.c_double)]
        L.gnx_dot.argtypes = [vp, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.gnx_norm2.argtypes = [vp, vp, C.POINTER(C.c_double)]
        L.gnx_scale_by_field.argtypes = [vp, vp, vp, vp]
        dp = C.POINTER(C.c_double)
        L.gnx_hmc_create.argtypes = [vp, C.c_double, C.c_int, C.POINTER(vp)]
        L.gnx_hmc_destroy.argtypes = [vp]
        L.gnx_hmc_set_sigma.argtypes = [vp, dp]
        L.gnx_hmc_get_sigma.argtypes = [vp, dp, dp]
        L.gnx_hmc_device_fields.argtypes = [vp] + [C.POINTER(vp)] * 7
        L.gnx_hmc_trajectory.argtypes = [vp, dp, dp, dp, C.c_double, C.c_int, C.c_double, C.c_double,
                                         C.c_long, C.POINTER(TrajResult)]
        L.gnx_hmc_accept.argtypes = [vp, C.c_int]
        L.gnx_hmc_set_force_precision.argtypes = [vp, C.c_int]
        L.gnx_hmc_condensate.argtypes = [vp, dp, C.c_double, C.c_long, dp]
        L.gnx_hmc_condensate_sigma.argtypes = [vp, dp]
        L.gnx_hmc_cg_stats.argtypes = [vp, dp, C.POINTER(C.c_int)]
        L.gnx_links_from_lex.argtypes = [vp, vp, vp]
        L.gnx_links_to_lex.argtypes = [vp, vp, vp]
        L.gnx_link_dirac.argtypes = [vp, C.c_int, vp, vp, vp, vp]
        L.gnx_link_cg_batch.argtypes = [vp, vp, vp, vp, vp, C.c_int, C.c_double, C.c_long, C.c_int,
                                        C.POINTER(CGResult)]
        L.gnx_gauge_force.argtypes = [vp, vp, C.c_double, vp]
        L.gnx_link_force_from_solutions.argtypes = [vp, vp, vp, vp, vp, vp, vp, C.c_int]
        L.gnx_plaquette_sum.argtypes = [vp, vp, dp]
        L.gnx_link_hmc_create.argtypes = [vp, C.c_double, C.c_int, C.c_int, C.POINTER(vp)]
        L.gnx_link_hmc_destroy.argtypes = [vp]
        L.gnx_link_hmc_set_fields.argtypes = [vp, dp, dp]
        L.gnx_link_hmc_set_solutions.argtypes = [vp, dp, dp]
        L.gnx_link_hmc_get_links.argtypes = [vp, dp]
        L.gnx_link_hmc_device_fields.argtypes = [vp] + [C.POINTER(vp)] * 7
        L.gnx_link_hmc_compute_force.argtypes = [vp, C.c_double, C.c_long, C.POINTER(C.c_int)]
        L.gnx_link_hmc_trajectory.argtypes = [vp, dp, dp, dp, C.c_double, C.c_int, C.c_double, C.c_long,
                                              C.POINTER(TrajResult)]
        L.gnx_link_hmc_accept.argtypes = [vp, C.c_int]
        L.gnx_link_hmc_energy.argtypes = [vp, dp]
        L.gnx_link_hmc_cg_stats.argtypes = [vp, dp, C.POINTER(C.c_int)]
        ip = C.POINTER(C.c_int)
        L.gnx_multi_create.argtypes = [C.c_int, ip, ip, C.c_int, C.POINTER(vp)]
        L.gnx_multi_destroy.argtypes = [vp]
        L.gnx_multi_set_stream.argtypes = [vp, vp]
        L.gnx_multi_cg.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.c_double, C.c_long,
                                   C.c_int, C.POINTER(CGResult)]
        L.gnx_multi_link_cg.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.c_double,
                                        C.c_long, C.c_int, C.POINTER(CGResult)]
        L.gnx_multi_cg_prec.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.c_double, C.c_long,
                                        C.c_int, C.c_int, C.POINTER(CGResult)]
        L.gnx_hmc_ensemble_trajectory.argtypes = [vp, C.c_int, C.POINTER(vp), C.POINTER(dp), C.POINTER(dp),
                                                  C.POINTER(dp), C.c_double, C.c_int, C.c_double, C.c_double,
                                                  C.c_long, C.POINTER(TrajResult)]
        L.gnx_link_hmc_ensemble_trajectory.argtypes = [vp, C.c_int, C.POINTER(vp), C.POINTER(dp), C.POINTER(dp),
                                                       C.POINTER(dp), C.c_double, C.c_int, C.c_double, C.c_long,
                                                       C.POINTER(TrajResult)]
        _lib = L
    return _lib


def _ptr(t):
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def _check(status):
    if status != 0:
        raise GnxError(status, lib().gnx_last_error().decode())


def as_real(t):
    """complex128[V] -> float64[2V] view; float64 passes through."""
    import torch
    if t.dtype == torch.complex128:
        return torch.view_as_real(t).reshape(-1)
    return t


class Context:
    """One L x L x LT lattice on one GPU (gnx_create)."""

    def __init__(self, L, LT=None, device=0, stream=None):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("gnx needs a ROCm GPU; no device is visible")
        LT = LT or L
        h = C.c_void_p()
        _check(lib().gnx_create(L, LT, device, C.byref(h)))
        self.h = h
        self.L, self.LT, self.V = L, LT, L * L * LT
        self.device = torch.device("cuda", device)
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        self.stream = s
        _check(lib().gnx_set_stream(self.h, C.c_void_p(s.cuda_stream)))

    def close(self):
        if getattr(self, "h", None):
            lib().gnx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(lib().gnx_synchronize(self.h))

    # -- allocation helpers (torch) ------------------------------------
    def vec(self, n=1):
        import torch
        return torch.zeros(n * 2 * self.V, dtype=torch.float64, device=self.device)

    def field(self):
        import torch
        return torch.zeros(self.V, dtype=torch.float64, device=self.device)

    # -- A1/A2/A3-A5 ------------------------------------------------------
    def geometry_tables(self):
        import torch
        V = self.V
        fpp = torch.zeros(3 * V, dtype=torch.int32, device=self.device)
        fmm = torch.zeros_like(fpp)
        epp = torch.zeros(3 * V, dtype=torch.float64, device=self.device)
        emm = torch.zeros_like(epp)
        f2l = torch.zeros(V, dtype=torch.int32, device=self.device)
        _check(lib().gnx_geometry_tables(self.h, _ptr(fpp), _ptr(fmm), _ptr(epp), _ptr(emm), _ptr(f2l)))
        self.synchronize()
        return fpp, fmm, epp, emm, f2l

    def sigma_mass(self, sigma, sm):
        _check(lib().gnx_sigma_mass(self.h, _ptr(sigma), _ptr(sm)))

    def dirac(self, op, sm, vin, vout):
        _check(lib().gnx_dirac(self.h, op, _ptr(sm), _ptr(as_real(vin)), _ptr(as_real(vout))))

    def dirac_batch(self, op, sm, vin, vout, nvec):
        assert as_real(vin).numel() == 2 * self.V * nvec and as_real(vout).numel() == 2 * self.V * nvec
        _check(lib().gnx_dirac_batch(self.h, op, _ptr(sm), _ptr(as_real(vin)), _ptr(as_real(vout)), nvec))

    # -- A6/A7 --------------------------------------------------------------
    def cg(self, sm, b, x, tol, max_iter=0, start=START_NULL, prec=PREC_F64, check=True):
        r = CGResult()
        st = lib().gnx_cg(self.h, _ptr(sm), _ptr(as_real(b)), _ptr(as_real(x)), tol, max_iter, start, prec,
                          C.byref(r))
        if check and st not in (0, 3):
            _check(st)
        return st, r

    def cg_noalloc(self, sm, b, x, tol, max_iter=0, start=START_NULL):
        """gnx_cg_noalloc: the FP64 single-RHS solve with no allocation in
        the call (buffers in the library's static device array)."""
        r = CGResult()
        st = lib().gnx_cg_noalloc(self.h, _ptr(sm), _ptr(as_real(b)), _ptr(as_real(x)), tol, max_iter, start,
                                  C.byref(r))
        if st not in (0, 3):
            _check(st)
        return st, r

    def cg_batch(self, sm, b, x, nrhs, tol, max_iter=0, start=START_NULL, prec=PREC_F64):
        res = (CGResult * nrhs)()
        st = lib().gnx_cg_batch(self.h, _ptr(sm), _ptr(as_real(b)), _ptr(as_real(x)), nrhs, tol, max_iter,
                                start, prec, res)
        if st not in (0, 3):
            _check(st)
        return st, list(res)

    def condensate_batch(self, sm, noise, nsrc, tol, X, max_iter=0, prec=PREC_F64):
        """Re(v_k^dag M^-1 v_k)/V for nsrc noise vectors in one batched solve."""
        res = (CGResult * nsrc)()
        out = (C.c_double * nsrc)()
        st = lib().gnx_condensate_batch(self.h, _ptr(sm), _ptr(as_real(noise)), nsrc, tol, max_iter, prec,
                                        _ptr(as_real(X)), out, res)
        if st not in (0, 3):
            _check(st)
        return st, list(out), list(res)

    # -- A11: link sector (links f-order [V][3] complex on the device) --
    def links_from_lex(self, ulex, uf):
        _check(lib().gnx_links_from_lex(self.h, _ptr(as_real(ulex)), _ptr(as_real(uf))))

    def links_to_lex(self, uf, ulex):
        _check(lib().gnx_links_to_lex(self.h, _ptr(as_real(uf)), _ptr(as_real(ulex))))

    def link_dirac(self, op, sm, uf, vin, vout):
        _check(lib().gnx_link_dirac(self.h, op, _ptr(sm), _ptr(as_real(uf)), _ptr(as_real(vin)), _ptr(as_real(vout))))

    def link_cg_batch(self, sm, uf, b, x, nrhs, tol, max_iter=0, start=START_NULL):
        res = (CGResult * nrhs)()
        st = lib().gnx_link_cg_batch(self.h, _ptr(sm), _ptr(as_real(uf)), _ptr(as_real(b)), _ptr(as_real(x)), nrhs,
                                     tol, max_iter, start, res)
        if st not in (0, 3):
            _check(st)
        return st, list(res)

    def gauge_force(self, uf, beta, force):
        _check(lib().gnx_gauge_force(self.h, _ptr(as_real(uf)), beta, _ptr(force)))

    def link_force_from_solutions(self, uf, X0, Y0, X1, Y1, force, accumulate=False):
        _check(lib().gnx_link_force_from_solutions(self.h, _ptr(as_real(uf)), _ptr(as_real(X0)), _ptr(as_real(Y0)),
                                                   _ptr(as_real(X1)), _ptr(as_real(Y1)), _ptr(force),
                                                   int(accumulate)))

    def plaquette_sum(self, uf):
        o = C.c_double()
        _check(lib().gnx_plaquette_sum(self.h, _ptr(as_real(uf)), C.byref(o)))
        return o.value

    def set_reliable_delta(self, delta):
        _check(lib().gnx_set_reliable_delta(self.h, delta))

    def set_autotune(self, on=True):
        _check(lib().gnx_set_autotune(self.h, int(bool(on))))

    def tune_runs(self):
        a, b = C.c_long(), C.c_long()
        _check(lib().gnx_tune_runs(self.h, C.byref(a), C.byref(b)))
        return {"fp64": a.value, "mixed": b.value}

    def replay_stats(self):
        a, b = C.c_long(), C.c_long()
        _check(lib().gnx_replay_stats(self.h, C.byref(a), C.byref(b)))
        return {"slots": a.value, "polls": b.value}

    def cg_shape(self, nrhs=1):
        spt, xcd = C.c_int(), C.c_int()
        st = lib().gnx_cg_shape(self.h, nrhs, C.byref(spt), C.byref(xcd))
        return None if st != 0 else {"sites_per_thread": spt.value, "xcd_slab": bool(xcd.value & 1),
                                     "one_reduction": bool(xcd.value & 2), "deferred": bool(xcd.value & 4),
                                     "fused": bool(xcd.value & 8), "one_launch": bool(xcd.value & 16)}

    def cg_mixed_shape(self, nrhs, prec):
        spt, xcd, pair, early = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        st = lib().gnx_cg_mixed_shape(self.h, nrhs, prec, C.byref(spt), C.byref(xcd), C.byref(pair),
                                      C.byref(early))
        return None if st != 0 else {"sites_per_thread": spt.value, "xcd_slab": bool(xcd.value & 1),
                                     "pair": bool(pair.value), "loads_first": bool(early.value),
                                     "consecutive": bool(xcd.value & 2)}

    # -- A9/A10 -----------------------------------------------------------
    def sigma_force(self, sigma, sm, beta_s, tol, phi, X, Y, force, max_iter=0, nflav=4):
        vp = C.c_void_p
        res = (CGResult * 2)()
        st = lib().gnx_sigma_force(self.h, _ptr(sigma), _ptr(sm), beta_s, nflav, tol, max_iter,
                                   (vp * 2)(*[_ptr(as_real(p)) for p in phi]),
                                   (vp * 2)(*[_ptr(as_real(p)) for p in X]),
                                   (vp * 2)(*[_ptr(as_real(p)) for p in Y]), _ptr(force), res)
        if st not in (0, 3):
            _check(st)
        return st, list(res)

    def sigma_force_from_solutions(self, sigma, beta_s, X0, Y0, X1, Y1, force, nflav=4):
        _check(lib().gnx_sigma_force_from_solutions(self.h, _ptr(sigma), beta_s, nflav, _ptr(as_real(X0)),
                                                    _ptr(as_real(Y0)), _ptr(as_real(X1)),
                                                    _ptr(as_real(Y1)), _ptr(force)))

    def evolve_momenta(self, mom, force, eps):
        _check(lib().gnx_evolve_momenta(self.h, _ptr(mom), _ptr(force), eps))

    def evolve_sigma(self, sigma, sm, mom, eps):
        _check(lib().gnx_evolve_sigma(self.h, _ptr(sigma), _ptr(sm), _ptr(mom), eps))

    def energy(self, mom, sigma, beta_s, phi, X, nflav=4):
        vp = C.c_void_p
        H = C.c_double()
        _check(lib().gnx_energy(self.h, _ptr(mom), _ptr(sigma), beta_s, nflav,
                                (vp * 2)(*[_ptr(as_real(p)) for p in phi]),
                                (vp * 2)(*[_ptr(as_real(p)) for p in X]), C.byref(H)))
        return H.value

    def dot(self, a, b):
        re, im = C.c_double(), C.c_double()
        _check(lib().gnx_dot(self.h, _ptr(as_real(a)), _ptr(as_real(b)), C.byref(re), C.byref(im)))
        return complex(re.value, im.value)

    def norm2(self, a):
        o = C.c_double()
        _check(lib().gnx_norm2(self.h, _ptr(as_real(a)), C.byref(o)))
        return o.value

    def scale_by_field(self, field, x, y):
        _check(lib().gnx_scale_by_field(self.h, _ptr(field), _ptr(as_real(x)), _ptr(as_real(y))))


def _dptr(a):
    import numpy as np
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


class HMC:
    """Device-resident sigma-HMC state (gnx_hmc_*).  Draws come from the caller."""

    def __init__(self, ctx, beta_s=0.85, nflav=4):
        self.ctx = ctx
        h = C.c_void_p()
        _check(lib().gnx_hmc_create(ctx.h, beta_s, nflav, C.byref(h)))
        self.h = h
        self.beta_s = beta_s

    def close(self):
        if getattr(self, "h", None):
            lib().gnx_hmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_sigma(self, sigma):
        a, p = _dptr(sigma)
        _check(lib().gnx_hmc_set_sigma(self.h, p))

    def get_sigma(self):
        import numpy as np
        s = np.zeros(self.ctx.V)
        m = np.zeros(self.ctx.V)
        _check(lib().gnx_hmc_get_sigma(self.h, s.ctypes.data_as(C.POINTER(C.c_double)),
                                       m.ctypes.data_as(C.POINTER(C.c_double))))
        return s, m

    def trajectory(self, mom, chi0, chi1, eps, steps, tol_high, tol_medium, max_iter=0):
        r = TrajResult()
        a, pa = _dptr(mom)
        b, pb = _dptr(chi0)
        c, pc = _dptr(chi1)
        st = lib().gnx_hmc_trajectory(self.h, pa, pb, pc, eps, steps, tol_high, tol_medium, max_iter, C.byref(r))
        if st not in (0, 3):
            _check(st)
        return st, r

    def accept(self, accepted):
        _check(lib().gnx_hmc_accept(self.h, int(accepted)))

    def set_force_precision(self, prec):
        """Arithmetic of the MEDIUM-tolerance force solves (PREC_F64 default)."""
        _check(lib().gnx_hmc_set_force_precision(self.h, prec))

    def condensate(self, z2, tol, max_iter=0):
        import numpy as np
        out = np.zeros(11)
        a, p = _dptr(z2)
        _check(lib().gnx_hmc_condensate(self.h, p, tol, max_iter, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def condensate_sigma(self):
        import numpy as np
        out = np.zeros(8)
        _check(lib().gnx_hmc_condensate_sigma(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def cg_stats(self):
        it = C.c_double()
        calls = C.c_int()
        _check(lib().gnx_hmc_cg_stats(self.h, C.byref(it), C.byref(calls)))
        return it.value, calls.value


class Rng:
    """The reference's host random stream (include/gnx_rng.h): Initrand(seed)
    then Casuale / Gauss1 / Gauss2MeanVariance (random.c), and bulk draws in
    the reference's order.  Host memory, numpy arrays."""

    DRAW_CASUALE, DRAW_GAUSS1, DRAW_SIGMA, DRAW_Z2_COMPLEX, DRAW_Z2_REAL, DRAW_GAUSS2_PAIRS = range(6)

    def __init__(self, seed):
        h = C.c_void_p()
        _check(lib().gnx_rng_create(seed, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().gnx_rng_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _fill(self, kind, n, out_len, var=0.0):
        import numpy as np
        out = np.empty(out_len)
        _check(lib().gnx_rng_fill(self.h, kind, n, var, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def casuale(self, n=None):
        if n is None:
            return lib().gnx_rng_casuale(self.h)
        return self._fill(self.DRAW_CASUALE, n, n)

    def gauss1(self, n=None):
        if n is None:
            return lib().gnx_rng_gauss1(self.h)
        return self._fill(self.DRAW_GAUSS1, n, n)

    def gauss2mv(self, mean, var):
        a, b = C.c_double(), C.c_double()
        lib().gnx_rng_gauss2mv(self.h, mean, var, C.byref(a), C.byref(b))
        return a.value, b.value

    def sigma(self, V, var):
        """V values: Gauss2MeanVariance(0, var) pairs into (x, x + V/2) (sigma.c:29-30)."""
        return self._fill(self.DRAW_SIGMA, V, V, var)

    def gauss2_pairs(self, n, var=1.0):
        """n values: Gauss2MeanVariance(0, var) pairs into (2r, 2r+1) (link momenta)."""
        return self._fill(self.DRAW_GAUSS2_PAIRS, n, n, var)

    def z2_complex(self, V):
        """2V values: per site re then im, each +-1."""
        return self._fill(self.DRAW_Z2_COMPLEX, V, 2 * V)

    def z2_real(self, V):
        """V values +-1 (condensate.c:52-60)."""
        return self._fill(self.DRAW_Z2_REAL, V, V)


class Multi:
    """Uneven batch (gnx_multi_*): n systems on lattices of different sizes,
    solved together by one graph-replayed FP64 CG."""

    def __init__(self, sizes, device=0, stream=None):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("gnx needs a ROCm GPU; no device is visible")
        self.sizes = [(L, LT or L) for L, LT in ((s if isinstance(s, tuple) else (s, s)) for s in sizes)]
        n = len(self.sizes)
        Ls = (C.c_int * n)(*[L for L, _ in self.sizes])
        LTs = (C.c_int * n)(*[LT for _, LT in self.sizes])
        h = C.c_void_p()
        _check(lib().gnx_multi_create(n, Ls, LTs, device, C.byref(h)))
        self.h = h
        self.n = n
        self.V = [L * L * LT for L, LT in self.sizes]
        self.device = torch.device("cuda", device)
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        self.stream = s
        _check(lib().gnx_multi_set_stream(self.h, C.c_void_p(s.cuda_stream)))

    def close(self):
        if getattr(self, "h", None):
            lib().gnx_multi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cg(self, sms, bs, xs, tol, max_iter=0, start=START_NULL, links=None, prec=PREC_F64):
        """sms[k] (V_k), bs[k], xs[k] (2 V_k) device tensors; links[k] (6 V_k
        doubles, f-order, three complex per site) selects the link-coupled
        operator, prec a mixed-precision solve (live operator only); returns
        (status, [CGResult])."""
        vp = C.c_void_p
        n = self.n
        assert len(sms) == len(bs) == len(xs) == n
        for k in range(n):
            assert sms[k].numel() == self.V[k] and as_real(bs[k]).numel() == 2 * self.V[k]
            assert as_real(xs[k]).numel() == 2 * self.V[k]
            assert links is None or as_real(links[k]).numel() == 6 * self.V[k]
        res = (CGResult * n)()
        a_sm = (vp * n)(*[_ptr(t) for t in sms])
        a_b = (vp * n)(*[_ptr(as_real(t)) for t in bs])
        a_x = (vp * n)(*[_ptr(as_real(t)) for t in xs])
        if links is None and prec != PREC_F64:
            st = lib().gnx_multi_cg_prec(self.h, a_sm, a_b, a_x, tol, max_iter, start, prec, res)
        elif links is None:
            st = lib().gnx_multi_cg(self.h, a_sm, a_b, a_x, tol, max_iter, start, res)
        else:
            assert prec == PREC_F64, "the link-coupled uneven solve is FP64" 
            st = lib().gnx_multi_link_cg(self.h, a_sm, (vp * n)(*[_ptr(as_real(t)) for t in links]), a_b, a_x, tol,
                                         max_iter, start, res)
        if st not in (0, 3):
            _check(st)
        return st, list(res)


class LinkHMC:
    """Device-resident link HMC (gnx_link_hmc_*) at fixed sigma mass."""

    def __init__(self, ctx, beta=1.0, pure_gauge=True, linked=True):
        self.ctx = ctx
        h = C.c_void_p()
        _check(lib().gnx_link_hmc_create(ctx.h, beta, int(pure_gauge), int(linked), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().gnx_link_hmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_fields(self, ulex=None, sm=None):
        """ulex: complex128[3V] (lex order), sm: float64[V], host arrays."""
        import numpy as np
        a = b = None
        pa = pb = None
        if ulex is not None:
            a, pa = _dptr(np.ascontiguousarray(ulex, np.complex128).view(np.float64))
        if sm is not None:
            b, pb = _dptr(sm)
        _check(lib().gnx_link_hmc_set_fields(self.h, pa, pb))

    def set_solutions(self, X0, X1):
        a, pa = _dptr(X0)
        b, pb = _dptr(X1)
        _check(lib().gnx_link_hmc_set_solutions(self.h, pa, pb))

    def get_links(self):
        import numpy as np
        out = np.zeros(6 * self.ctx.V)
        _check(lib().gnx_link_hmc_get_links(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out.view(np.complex128)

    def device_fields(self):
        ps = [C.c_void_p() for _ in range(7)]
        _check(lib().gnx_link_hmc_device_fields(self.h, *[C.byref(p) for p in ps]))
        return [p.value for p in ps]

    def compute_force(self, tol, max_iter=0):
        ok = C.c_int()
        _check(lib().gnx_link_hmc_compute_force(self.h, tol, max_iter, C.byref(ok)))
        return bool(ok.value)

    def trajectory(self, mom, chi0, chi1, eps, steps, tol, max_iter=0):
        r = TrajResult()
        a, pa = _dptr(mom)
        b, pb = _dptr(chi0)
        c, pc = _dptr(chi1)
        st = lib().gnx_link_hmc_trajectory(self.h, pa, pb, pc, eps, steps, tol, max_iter, C.byref(r))
        if st not in (0, 3):
            _check(st)
        return st, r

    def accept(self, accepted):
        _check(lib().gnx_link_hmc_accept(self.h, int(accepted)))

    def energy(self):
        o = C.c_double()
        _check(lib().gnx_link_hmc_energy(self.h, C.byref(o)))
        return o.value

    def cg_stats(self):
        it = C.c_double()
        calls = C.c_int()
        _check(lib().gnx_link_hmc_cg_stats(self.h, C.byref(it), C.byref(calls)))
        return it.value, calls.value


class Ensemble:
    """Chains on lattices of uneven size stepped in lockstep: every solve point
    of the integrator is one uneven-batch launch per CG kernel over all
    chains (gnx_hmc_ensemble_trajectory / gnx_link_hmc_ensemble_trajectory).
    chains: HMC or LinkHMC objects (one kind per ensemble)."""

    def __init__(self, chains, device=0):
        self.chains = list(chains)
        sizes = []
        for c in self.chains:
            sizes += [(c.ctx.L, c.ctx.LT)] * 2
        self.multi = Multi(sizes, device)

    def close(self):
        self.multi.close()

    def _draws(self, arrs):
        keep = [_dptr(a) for a in arrs]
        return keep, (C.POINTER(C.c_double) * len(keep))(*[p for _, p in keep])

    def sigma_trajectory(self, moms, chi0s, chi1s, eps, steps, tol_high, tol_medium, max_iter=0):
        """One gnx_hmc_trajectory per chain, in lockstep; returns (status, [TrajResult])."""
        n = len(self.chains)
        k0, m = self._draws(moms)
        k1, a = self._draws(chi0s)
        k2, b = self._draws(chi1s)
        res = (TrajResult * n)()
        hs = (C.c_void_p * n)(*[c.h for c in self.chains])
        st = lib().gnx_hmc_ensemble_trajectory(self.multi.h, n, hs, m, a, b, eps, steps, tol_high, tol_medium,
                                               max_iter, res)
        if st not in (0, 3):
            _check(st)
        return st, list(res)

    def link_trajectory(self, moms, chi0s, chi1s, eps, steps, tol, max_iter=0):
        """One gnx_link_hmc_trajectory per chain (live operator), in lockstep."""
        n = len(self.chains)
        k0, m = self._draws(moms)
        k1, a = self._draws(chi0s)
        k2, b = self._draws(chi1s)
        res = (TrajResult * n)()
        hs = (C.c_void_p * n)(*[c.h for c in self.chains])
        st = lib().gnx_link_hmc_ensemble_trajectory(self.multi.h, n, hs, m, a, b, eps, steps, tol, max_iter, res)
        if st not in (0, 3):
            _check(st)
        return st, list(res)
