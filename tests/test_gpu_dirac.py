"""GPU parity of geometry, sigma mass and the Dirac operator (bit-exact)."""
import numpy as np
import pytest

from gpu_util import banner, dev, host

pytestmark = pytest.mark.gpu

SIZES = [(4, 4), (6, 10), (8, 8), (12, 12), (16, 16), (32, 32), (64, 64)]


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return torch


def _fields(P, lat, seed):
    r = P.Random(seed)
    sigma = r.sigma(lat.V, P.sigma_variance(0.85))
    b = r.z2_complex(lat.V)
    c = r.z2_complex(lat.V)
    return sigma, b, c


@pytest.mark.parametrize("L,LT", SIZES)
def test_geometry_tables(torch_dev, P, gnx_mod, L, LT):
    banner(f"geometry {L}x{L}x{LT}")
    lat = P.Lattice(L, LT)
    ctx = gnx_mod.Context(L, LT)
    fpp, fmm, epp, emm, f2l = ctx.geometry_tables()
    V = lat.V
    assert np.array_equal(host(fpp), lat.table("fpp", 3 * V))
    assert np.array_equal(host(fmm), lat.table("fmm", 3 * V))
    assert np.array_equal(host(epp), lat.table("eta_pp", 3 * V))
    assert np.array_equal(host(emm), lat.table("eta_mm", 3 * V))
    assert np.array_equal(host(f2l), lat.table("f2lex", V))


def test_golden_L8(torch_dev, golden, gnx_mod):
    banner("golden 8^3 operators")
    ctx = gnx_mod.Context(8)
    sigma = dev(golden["op_sigma"])
    sm = ctx.field()
    ctx.sigma_mass(sigma, sm)
    ctx.synchronize()
    assert np.array_equal(host(sm), golden["op_sigmamass"])
    b = dev(golden["op_b"])
    out = ctx.vec()
    for op, key in [(gnx_mod.OP_M, "op_Mb"), (gnx_mod.OP_MDAG, "op_Mdagb")]:
        ctx.dirac(op, sm, b, out)
        ctx.synchronize()
        assert np.array_equal(host(out), golden[key]), key
    for op, key in [(gnx_mod.OP_SUB_M, "op_c_minus_Mb"), (gnx_mod.OP_SUB_MDAG, "op_c_minus_Mdagb")]:
        out = dev(golden["op_c"])
        ctx.dirac(op, sm, b, out)
        ctx.synchronize()
        assert np.array_equal(host(out), golden[key]), key


@pytest.mark.parametrize("L,LT", SIZES)
def test_ops_vs_oracle(torch_dev, P, gnx_mod, L, LT):
    banner(f"operators {L}x{L}x{LT}")
    lat = P.Lattice(L, LT)
    sigma, b, c = _fields(P, lat, 1000 + L)
    sm_ref = lat.sigma_mass(sigma)
    ctx = gnx_mod.Context(L, LT)
    sm = ctx.field()
    ctx.sigma_mass(dev(sigma), sm)
    ctx.synchronize()
    assert np.array_equal(host(sm), sm_ref)
    bd = dev(b)
    out = ctx.vec()
    ctx.dirac(gnx_mod.OP_M, sm, bd, out)
    ctx.synchronize()
    assert np.array_equal(host(out), lat.M(sm_ref, b))
    ctx.dirac(gnx_mod.OP_MDAG, sm, bd, out)
    ctx.synchronize()
    assert np.array_equal(host(out), lat.Mdag(sm_ref, b))
    out = dev(c)
    ctx.dirac(gnx_mod.OP_SUB_M, sm, bd, out)
    ctx.synchronize()
    assert np.array_equal(host(out), lat.sub_M(sm_ref, b, c))
    out = dev(c)
    ctx.dirac(gnx_mod.OP_SUB_MDAG, sm, bd, out)
    ctx.synchronize()
    assert np.array_equal(host(out), lat.sub_Mdag(sm_ref, b, c))
    # normal operators are compositions of the bit-exact stages
    out = ctx.vec()
    ctx.dirac(gnx_mod.OP_MMDAG, sm, bd, out)
    ctx.synchronize()
    assert np.array_equal(host(out), lat.M(sm_ref, lat.Mdag(sm_ref, b)))
    ctx.dirac(gnx_mod.OP_MDAGM, sm, bd, out)
    ctx.synchronize()
    assert np.array_equal(host(out), lat.Mdag(sm_ref, lat.M(sm_ref, b)))


@pytest.mark.parametrize("L,LT", [(8, 8), (6, 10)])
def test_dirac_batch_matches_single(torch_dev, P, gnx_mod, L, LT):
    """gnx_dirac_batch: each of 5 vectors bit-identical to its own gnx_dirac,
    for the four single-operator forms (the subtracting ones read out), and
    the batch against the oracle."""
    banner(f"dirac batch {L}x{L}x{LT}")
    lat = P.Lattice(L, LT)
    V = lat.V
    r = P.Random(71)
    smh = lat.sigma_mass(r.sigma(V, P.sigma_variance(0.85)))
    sm = dev(smh)
    n = 5
    vs = [r.z2_complex(V) for _ in range(n)]
    ts = [r.z2_complex(V) for _ in range(n)]
    ctx = gnx_mod.Context(L, LT)
    for op, ref in ((gnx_mod.OP_M, lambda v, t: lat.M(smh, v)), (gnx_mod.OP_MDAG, lambda v, t: lat.Mdag(smh, v)),
                    (gnx_mod.OP_SUB_M, lambda v, t: lat.sub_M(smh, v, t)),
                    (gnx_mod.OP_SUB_MDAG, lambda v, t: lat.sub_Mdag(smh, v, t))):
        vin = dev(np.concatenate(vs))
        out = dev(np.concatenate(ts))
        ctx.dirac_batch(op, sm, vin, out, n)
        outh = host(out)
        for k in range(n):
            o1 = dev(ts[k])
            ctx.dirac(op, sm, dev(vs[k]), o1)
            seg = outh[2 * V * k:2 * V * (k + 1)]
            assert np.array_equal(seg, host(o1)), (op, k)
            assert np.array_equal(seg, ref(vs[k], ts[k])), (op, k)


@pytest.mark.parametrize("L,LT", [(256, 1024), (382, 382)])
def test_ops_at_the_largest_volumes(torch_dev, P, gnx_mod, L, LT):
    """The largest lattices the library accepts: 256^2 x 1024 is the 2^26-site
    cap (powers of two), 382^3 (55.7 M sites) the largest cube whose in-kernel
    decode (umulhi by ceil(2^32/L)) is still exact; there the vectors' 32-bit
    byte offsets and the decode are closest to their limits.  sigma mass, M
    and M^dag bit-exact against the oracle; at the cap also three FP64 CG
    iterations (iteration cap 3, the rule's launch shape) against the
    oracle's three, within 1e-12 relative."""
    banner(f"operators at {L}x{L}x{LT}")
    import torch
    lat = P.Lattice(L, LT)
    sigma, b, _ = _fields(P, lat, 77)
    sm_ref = lat.sigma_mass(sigma)
    ctx = gnx_mod.Context(L, LT)
    sm = ctx.field()
    ctx.sigma_mass(dev(sigma), sm)
    ctx.synchronize()
    assert np.array_equal(host(sm), sm_ref)
    bd = dev(b)
    out = ctx.vec()
    ctx.dirac(gnx_mod.OP_M, sm, bd, out)
    ctx.synchronize()
    assert np.array_equal(host(out), lat.M(sm_ref, b))
    ctx.dirac(gnx_mod.OP_MDAG, sm, bd, out)
    ctx.synchronize()
    assert np.array_equal(host(out), lat.Mdag(sm_ref, b))
    if lat.V == 1 << 26:
        x = ctx.vec()
        st, res = ctx.cg(sm, bd, x, 1e-10, max_iter=3, start=gnx_mod.START_IMPOSED, check=False)
        assert st == 3 and res.iterations == 3 and res.converged == 0
        xo, ok, _ = lat.cg(sm_ref, b, None, tol=1e-10, max_iter=3, start="imposed")
        xg = host(x)
        assert np.linalg.norm(xg - xo) <= 1e-12 * np.linalg.norm(xo)
    del out, bd, sm
    ctx.close()
    torch.cuda.empty_cache()


@pytest.mark.parametrize("L,LT", [(16, 16), (8, 16), (32, 32)])
def test_normal_operators_one_launch(torch_dev, P, gnx_mod, L, LT):
    """M M^dag and M^dag M in one launch (the intermediate in LDS, on
    lattices of 4 x 8 x 16 tiles; 8^2 x 16 is a single tile along x1 and
    x2, so its halo wraps onto itself): bit-exact against the oracle's
    composition, and a batch of 3 (gnx_dirac_batch, which accepts the normal
    operators on such lattices) bit-identical to single applications."""
    banner(f"normal operators in one launch {L}x{L}x{LT}")
    import torch
    lat = P.Lattice(L, LT)
    sigma, b, c = _fields(P, lat, 500 + L + LT)
    sm_ref = lat.sigma_mass(sigma)
    ctx = gnx_mod.Context(L, LT)
    sm = ctx.field()
    ctx.sigma_mass(dev(sigma), sm)
    vs = [b, c, 0.5 * b - c]
    for op, ref in ((gnx_mod.OP_MMDAG, lambda v: lat.M(sm_ref, lat.Mdag(sm_ref, v))),
                    (gnx_mod.OP_MDAGM, lambda v: lat.Mdag(sm_ref, lat.M(sm_ref, v)))):
        outs = []
        for v in vs:
            out = ctx.vec()
            ctx.dirac(op, sm, dev(v), out)
            ctx.synchronize()
            assert np.array_equal(host(out), ref(v)), op
            outs.append(host(out))
        O = ctx.vec(3)
        ctx.dirac_batch(op, sm, dev(np.concatenate(vs)), O, 3)
        ctx.synchronize()
        assert np.array_equal(host(O), np.concatenate(outs)), op
    ctx.close()
