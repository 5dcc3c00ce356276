"""gnx_cg_noalloc: the FP64 single-RHS solve whose call allocates nothing
(every solver buffer in the library's static device array), against gnx_cg,
the entry point it is a variant of, through the C ABI.

Bars (written here): with the default launch shape both run the same kernels
in the same order, so every result is bit-identical to gnx_cg's -- the
solution, the iteration count, the exit flags -- on calls in a row on one
device, each on a smaller lattice than the last (the array keeps the
previous call's contents), for both starts, the prologue exit and the
iteration cap;
both also agree with the oracle to the solver tolerance (iterations within
+-1, ||x - x_ref|| <= 1e-7 ||x_ref||).  A lattice too large for the array
is refused with GNX_ERR_SIZE before any work.
"""
import numpy as np
import pytest

from gpu_util import banner, dev, host, rel

pytestmark = pytest.mark.gpu


def _inputs(P, L, seed):
    lat = P.Lattice(L)
    r = P.Random(seed)
    sigma = r.sigma(lat.V, P.sigma_variance(0.85))
    sm = lat.sigma_mass(sigma)
    return lat, sm, r.z2_complex(lat.V), 0.01 * r.z2_complex(lat.V)


def test_noalloc_matches_gnx_cg(P, gnx_mod, monkeypatch):
    monkeypatch.delenv("GNX_CG_SHAPE", raising=False)
    monkeypatch.delenv("GNX_AUTOTUNE", raising=False)
    import torch
    # each call on a smaller lattice than the one before; 48^3 runs the fused
    # form (one kernel per iteration, the rule's choice there), 32^3 and
    # 16^3 the one-reduction pair, 8^3 the one-launch solve
    for L, seed in ((48, 70), (32, 71), (16, 72), (8, 74)):
        for start in (gnx_mod.START_IMPOSED, gnx_mod.START_NULL):
            banner(f"cg_noalloc {L}^3 start={start} vs gnx_cg")
            lat, sm, b, x0 = _inputs(P, L, seed)
            ctx = gnx_mod.Context(L)
            smd, bd = dev(sm), dev(b)
            x_ref, x = dev(x0), dev(x0)
            s1, r1 = ctx.cg(smd, bd, x_ref, 1e-10, start=start)
            s2, r2 = ctx.cg_noalloc(smd, bd, x, 1e-10, start=start)
            torch.cuda.synchronize()
            assert s1 == 0 and s2 == 0
            assert (r2.converged, r2.early_exit, r2.iterations) == (r1.converged, r1.early_exit, r1.iterations)
            assert r2.residual == r1.residual
            assert np.array_equal(host(x), host(x_ref))
            xo, ok, st = lat.cg(sm, b, x0.copy(), tol=1e-10, start="null" if start == gnx_mod.START_NULL else "imposed")
            assert ok and abs(r2.iterations - st.last_iters) <= 1
            assert rel(host(x), xo) < 1e-7
            # from the converged x the null start exits in the prologue
            xc = host(x).copy()
            s3, r3 = ctx.cg_noalloc(smd, bd, x, 1e-10, start=gnx_mod.START_NULL)
            assert s3 == 0 and r3.early_exit == 1 and r3.iterations == 0
            assert np.array_equal(host(x), xc)
            ctx.close()


def test_noalloc_iteration_cap(P, gnx_mod, monkeypatch):
    banner("cg_noalloc iteration cap")
    monkeypatch.delenv("GNX_CG_SHAPE", raising=False)
    monkeypatch.delenv("GNX_AUTOTUNE", raising=False)
    lat, sm, b, x0 = _inputs(P, 24, 73)
    ctx = gnx_mod.Context(24)
    smd, bd = dev(sm), dev(b)
    x_ref, x = ctx.vec(), ctx.vec()
    s1, r1 = ctx.cg(smd, bd, x_ref, 1e-10, max_iter=7, start=gnx_mod.START_IMPOSED, check=False)
    s2, r2 = ctx.cg_noalloc(smd, bd, x, 1e-10, max_iter=7, start=gnx_mod.START_IMPOSED)
    assert s1 == 3 and s2 == 3 and r2.converged == 0 and r2.iterations == 7 == r1.iterations
    assert np.array_equal(host(x), host(x_ref))


def test_noalloc_refuses_oversize(gnx_mod):
    banner("cg_noalloc refuses a lattice larger than its static array")
    import torch
    ctx = gnx_mod.Context(128)
    v = torch.zeros(2 * ctx.V, dtype=torch.float64, device=ctx.device)
    sm = torch.zeros(ctx.V, dtype=torch.float64, device=ctx.device)
    with pytest.raises(gnx_mod.GnxError) as e:
        ctx.cg_noalloc(sm, v, torch.zeros_like(v), 1e-10)
    assert e.value.status == 4 and "static array" in str(e.value)
