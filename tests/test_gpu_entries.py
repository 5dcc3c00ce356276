"""GPU parity of the small public entry points the solvers and trajectories
use internally, called directly through the C ABI at 8^3 and 64^3:

  gnx_evolve_momenta  p += F eps            (moleculardynamics.c:843-848)  bit-exact
  gnx_evolve_sigma    sigma += eps p, then the sigma mass
                                            (moleculardynamics.c:824-836)  bit-exact
  gnx_energy          sum p^2/2 + sum_f 0.5 Re(phi_f^dag X_f)
                      + (N_f beta_s / 4) sum sigma^2
                                            (moleculardynamics.c:299-319)  1e-12 relative
  gnx_norm2           sum |a|^2             (fermion.h:57-67, squared)     1e-13 relative
  gnx_dot             sum conj(a) b                                        1e-13 relative
  gnx_scale_by_field  y = x * field         (condensate.c:221,225)         bit-exact
  gnx_set_reliable_delta: argument range; the factor is the drop of the
                      iterated residual since the last FP64 reliable update
                      that triggers the next one, so a larger factor forces
                      more updates, every solve still reaching the absolute
                      tolerance on its true residual.

Sums are taken in a tree order on the GPU and sequentially in the oracle,
so they agree to rounding, not bit for bit.  Also here: the launch-shape
choice is a fixed rule by default (no timing runs), and with autotuning on
each key is timed once even past the cache's first eight keys.
"""
import ctypes as C

import numpy as np
import pytest

from gpu_util import banner, dev, host, rel

pytestmark = pytest.mark.gpu


def fields(P, L, seed):
    lat = P.Lattice(L)
    r = P.Random(seed)
    V = lat.V
    sigma = r.sigma(V, P.sigma_variance(0.85))
    mom = np.array([r.gauss1() for _ in range(V)]) if V <= 4096 else np.random.default_rng(seed).standard_normal(V)
    force = np.random.default_rng(seed + 1).standard_normal(V)
    a = r.z2_complex(V) * np.random.default_rng(seed + 2).standard_normal(2 * V)
    b = np.random.default_rng(seed + 3).standard_normal(2 * V)
    return lat, sigma, mom, force, a, b


@pytest.mark.parametrize("L", [8, 64])
def test_evolve_and_scale_bit_exact(P, gnx_mod, L):
    banner(f"evolve_momenta, evolve_sigma, scale_by_field {L}^3")
    import torch
    lat, sigma, mom, force, a, _ = fields(P, L, 40 + L)
    ctx = gnx_mod.Context(L)
    eps = 0.0193
    m = dev(mom)
    ctx.evolve_momenta(m, dev(force), eps)
    s, sm = dev(sigma), ctx.field()
    ctx.evolve_sigma(s, sm, dev(mom), eps)
    y = ctx.vec()
    ctx.scale_by_field(dev(lat.sigma_mass(sigma)), dev(a), y)
    torch.cuda.synchronize()
    assert np.array_equal(host(m), mom + force * eps)
    s_new = sigma + eps * mom
    assert np.array_equal(host(s), s_new)
    assert np.array_equal(host(sm), lat.sigma_mass(s_new))
    f = lat.sigma_mass(sigma)
    assert np.array_equal(host(y), (a.reshape(-1, 2) * f[:, None]).ravel())


@pytest.mark.parametrize("L", [8, 64])
def test_energy_norm_dot(P, gnx_mod, L):
    banner(f"energy, norm2, dot {L}^3")
    lat, sigma, mom, _, a, b = fields(P, L, 70 + L)
    V = lat.V
    rng = np.random.default_rng(L)
    phi = [rng.standard_normal(2 * V) for _ in range(2)]
    X = [rng.standard_normal(2 * V) for _ in range(2)]
    ctx = gnx_mod.Context(L)
    H = ctx.energy(dev(mom), dev(sigma), 0.85, [dev(p) for p in phi], [dev(x) for x in X])
    dp = C.POINTER(C.c_double)
    keep = [np.ascontiguousarray(v) for v in phi + X]
    ph = (dp * 2)(*[v.ctypes.data_as(dp) for v in keep[:2]])
    xh = (dp * 2)(*[v.ctypes.data_as(dp) for v in keep[2:]])
    H_ref = P.lib().orc_energy(C.byref(lat.g), mom, sigma, 0.85, ph, xh)
    assert abs(H - H_ref) <= 1e-12 * abs(H_ref), (H, H_ref)
    n2 = ctx.norm2(dev(a))
    assert abs(n2 - float(np.dot(a, a))) <= 1e-13 * float(np.dot(a, a))
    assert abs(np.sqrt(n2) - P.lib().orc_norm(V, a)) <= 1e-13 * np.sqrt(n2)
    d = ctx.dot(dev(a), dev(b))
    ac, bc = a.view(np.complex128), b.view(np.complex128)
    d_ref = np.vdot(ac, bc)
    assert abs(d - d_ref) <= 1e-13 * np.sqrt(np.dot(a, a) * np.dot(b, b))


@pytest.mark.parametrize("prec", ["f32", "bf16", "f16"])
def test_reliable_delta(P, gnx_mod, prec):
    banner(f"set_reliable_delta {prec} 16^3")
    import torch
    pc = {"f32": gnx_mod.PREC_F32, "bf16": gnx_mod.PREC_BF16, "f16": gnx_mod.PREC_F16}[prec]
    L = 16
    lat = P.Lattice(L)
    r = P.Random(33)
    sigma = r.sigma(lat.V, P.sigma_variance(0.85))
    sm = lat.sigma_mass(sigma)
    b = r.z2_complex(lat.V)
    ctx = gnx_mod.Context(L)
    for bad in (-0.1, 1.0, 2.0):
        with pytest.raises(gnx_mod.GnxError):
            ctx.set_reliable_delta(bad)
    ups = {}
    for delta in (0.5, 0.05, 0.005):
        ctx.set_reliable_delta(delta)
        x = ctx.vec()
        st, res = ctx.cg(dev(sm), dev(b), x, 1e-10, start=gnx_mod.START_IMPOSED, prec=pc)
        torch.cuda.synchronize()
        assert st == 0 and res.converged
        tr = np.linalg.norm(b - lat.M(sm, lat.Mdag(sm, host(x))))
        assert tr < 1.5e-10, (delta, tr)
        ups[delta] = res.reliable_updates
    ctx.set_reliable_delta(0.0)  # back to the per-precision default
    banner(f"  reliable updates by delta: {ups}")
    assert ups[0.5] > ups[0.05] > ups[0.005]


def test_shape_rule_is_default_and_deterministic(P, gnx_mod):
    """Without autotuning no shape is timed: the FP64 and mixed solvers use
    the fixed rule, and two contexts produce the same bits."""
    banner("launch shape rule: no timing runs, same bits in two contexts")
    import torch
    L = 64
    lat = P.Lattice(L)
    r = P.Random(21)
    sm = dev(lat.sigma_mass(r.sigma(lat.V, P.sigma_variance(0.85))))
    b = dev(r.z2_complex(lat.V))
    out = []
    for _ in range(2):
        ctx = gnx_mod.Context(L)
        x = ctx.vec()
        ctx.cg(sm, b, x, 1e-10, start=gnx_mod.START_IMPOSED)
        xm = ctx.vec()
        ctx.cg(sm, b, xm, 1e-10, start=gnx_mod.START_IMPOSED, prec=gnx_mod.PREC_F16)
        torch.cuda.synchronize()
        assert ctx.tune_runs() == {"fp64": 0, "mixed": 0}
        assert ctx.cg_shape(1) == {"sites_per_thread": 2, "xcd_slab": True, "one_reduction": True,
                                   "deferred": False, "fused": False, "one_launch": False}
        assert ctx.cg_mixed_shape(1, gnx_mod.PREC_F16) == {"sites_per_thread": 2, "xcd_slab": True, "pair": False,
                                                           "loads_first": True, "consecutive": False}
        out.append((host(x), host(xm)))
        ctx.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_autotune_many_keys_tuned_once(P, gnx_mod):
    """Ten batch sizes on one context with autotuning on: each key is timed
    once (the cache holds sixteen, least recently used out), and solving
    every key again times nothing."""
    banner("autotune: 10 keys, each timed once")
    import torch
    L = 8
    lat = P.Lattice(L)
    r = P.Random(22)
    sm = dev(lat.sigma_mass(r.sigma(lat.V, P.sigma_variance(0.85))))
    B = dev(np.concatenate([r.z2_complex(lat.V) for _ in range(10)]))
    ctx = gnx_mod.Context(L)
    ctx.set_autotune(True)
    for rep in range(2):
        for n in range(1, 11):
            X = ctx.vec(n)
            st, res = ctx.cg_batch(sm, B[:2 * lat.V * n], X, n, 1e-10, start=gnx_mod.START_IMPOSED)
            assert st == 0 and all(q.converged for q in res)
            assert ctx.cg_shape(n) is not None
        torch.cuda.synchronize()
        assert ctx.tune_runs()["fp64"] == 10, (rep, ctx.tune_runs())
    for n in (1, 2):
        X = ctx.vec(n)
        ctx.cg_batch(sm, B[:2 * lat.V * n], X, n, 1e-10, start=gnx_mod.START_IMPOSED, prec=gnx_mod.PREC_F32)
        ctx.cg_batch(sm, B[:2 * lat.V * n], X, n, 1e-10, start=gnx_mod.START_IMPOSED, prec=gnx_mod.PREC_F32)
    assert ctx.tune_runs()["mixed"] == 2
