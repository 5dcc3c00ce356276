"""Mixed-precision CG (FP32 and BF16-storage inner iterations, FP64 reliable
updates) against the FP64 oracle.

Bars (written here): the solver only stops on the FP64 true residual, so
  - the true residual ||b - M M^dag x|| recomputed by the oracle is < tol;
  - ||x - x_ref|| <= 1e-6 ||x_ref|| against the oracle's FP64 solution at the
    same absolute tolerance 1e-10 (both within tol/lambda_min of the exact
    solution);
  - batched solves equal single solves bit-for-bit.
"""
import numpy as np
import pytest

from gpu_util import banner, dev, host, rel

pytestmark = pytest.mark.gpu


def true_res(lat, sm, b, x):
    return np.linalg.norm(b - lat.M(sm, lat.Mdag(sm, x)))


PRECS = ["f32", "bf16", "f16"]


def _prec(gnx_mod, prec):
    return {"f32": gnx_mod.PREC_F32, "bf16": gnx_mod.PREC_BF16, "f16": gnx_mod.PREC_F16}[prec]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("L", [8, 16, 24])
def test_mixed_vs_oracle(P, gnx_mod, prec, L):
    banner(f"mixed {prec} {L}^3")
    import torch
    lat = P.Lattice(L)
    r = P.Random(300 + L)
    sigma = r.sigma(lat.V, P.sigma_variance(0.85))
    sm = lat.sigma_mass(sigma)
    b = r.z2_complex(lat.V)
    xr, ok, st = lat.cg(sm, b, None, tol=1e-10, start="imposed")
    ctx = gnx_mod.Context(L)
    x = ctx.vec()
    p = _prec(gnx_mod, prec)
    s, res = ctx.cg(dev(sm), dev(b), x, 1e-10, start=gnx_mod.START_IMPOSED, prec=p)
    torch.cuda.synchronize()
    assert s == 0 and res.converged and res.reliable_updates >= 1
    xg = host(x)
    assert true_res(lat, sm, b, xg) < 1e-10
    assert rel(xg, xr) < 1e-6
    banner(f"  iterations fp64={st.last_iters} {prec}={res.iterations} updates={res.reliable_updates}")


SPT_OF_SHAPE = [1, 1, 2, 2, 4, 4, 1, 1, 2, 2, 1, 1, 2, 2, 4, 4, 4]


@pytest.mark.parametrize("shape", list(range(17)))
@pytest.mark.parametrize("prec", PRECS)
def test_mixed_batch_and_null_start(P, gnx_mod, prec, shape, monkeypatch):
    """Every launch shape: 0..5 one system per block (1/2/4 sites per thread,
    with and without the XCD slab mapping), 6..9 two systems per block,
    10..15 as 0..5 with the loads-first kernels (which also defer the
    correction update to every other slot), 16 the loads-first kernels on
    groups of four consecutive sites (4-byte storage, LT a multiple of 8: run
    at 16^3, two groups per parity row).  A batch of 3 (the last pair half
    empty) is bit-identical to 3 single solves with the same steps: a paired
    batch to the unpaired kernels of its site mapping, the others to their own
    shape; each solution meets the FP64 bar."""
    if prec == "f16" and shape < 10:
        pytest.skip("FP16 storage runs the loads-first shapes only (they carry the scale)")
    if prec == "f32" and shape == 16:
        pytest.skip("the four-site group kernels exist for the 4-byte storage formats only")
    banner(f"mixed {prec} shape {shape}: batch of 3 and null-start early exit")
    import torch
    L = 16 if shape == 16 else 12
    lat = P.Lattice(L)
    r = P.Random(41)
    sigma = r.sigma(lat.V, P.sigma_variance(0.85))
    smh = lat.sigma_mass(sigma)
    sm = dev(smh)
    bs = [r.z2_complex(lat.V) for _ in range(3)]
    ctx = gnx_mod.Context(L)
    p = _prec(gnx_mod, prec)
    monkeypatch.setenv("GNX_MIX_SHAPE", str(shape))
    X = ctx.vec(3)
    st, res = ctx.cg_batch(sm, dev(np.concatenate(bs)), X, 3, 1e-10, start=gnx_mod.START_IMPOSED, prec=p)
    assert st == 0
    pair = 6 <= shape < 10
    early = shape >= 10
    assert ctx.cg_mixed_shape(3, p) == {"sites_per_thread": SPT_OF_SHAPE[shape], "xcd_slab": shape % 2 == 1 or shape == 16,
                                        "pair": pair, "loads_first": early, "consecutive": shape == 16}
    # single solves: a paired batch against the unpaired kernels of the same
    # site mapping, the other shapes against themselves
    monkeypatch.setenv("GNX_MIX_SHAPE", str(shape - 6 if pair else shape))
    for k in range(3):
        x = ctx.vec()
        s1, r1 = ctx.cg(sm, dev(bs[k]), x, 1e-10, start=gnx_mod.START_IMPOSED, prec=p)
        assert r1.iterations == res[k].iterations
        xk = host(X[k * 2 * lat.V:(k + 1) * 2 * lat.V])
        assert np.array_equal(host(x), xk)
        assert true_res(lat, smh, bs[k], xk) < 1e-10
    # null start from the converged solution: the prologue exits, x untouched
    x = X[:2 * lat.V].clone()
    before = host(x).copy()
    s2, r2 = ctx.cg(sm, dev(bs[0]), x, 1e-10, start=gnx_mod.START_NULL, prec=p)
    torch.cuda.synchronize()
    assert s2 == 0 and r2.early_exit == 1 and np.array_equal(host(x), before)
    # null start from a non-converged x: restarts from zero and converges
    x = ctx.vec()
    x[:] = 1.0
    s3, r3 = ctx.cg(sm, dev(bs[1]), x, 1e-10, start=gnx_mod.START_NULL, prec=p)
    assert s3 == 0 and r3.converged and r3.early_exit == 0
    assert np.array_equal(host(x), host(X[2 * lat.V:4 * lat.V]))


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("cap", [5, 6, 13])
def test_mixed_iteration_cap_folds_pending(P, gnx_mod, prec, cap, monkeypatch):
    """Iteration cap reached on an odd and on an even slot.  The loads-first
    kernels keep up to two step lengths pending between their correction
    flushes; the cap's closing reliable update must fold them into x.  Checked
    three ways: the returned x agrees with the two-buffer kernels' (which
    update the correction every slot) to FP32 rounding, the reported residual
    is the true residual of the returned x (oracle, FP64), and the cap is
    reported as not converged after exactly `cap` iterations."""
    banner(f"mixed {prec} iteration cap {cap}")
    import torch
    L = 8
    lat = P.Lattice(L)
    r = P.Random(77)
    sigma = r.sigma(lat.V, P.sigma_variance(0.85))
    smh = lat.sigma_mass(sigma)
    sm = dev(smh)
    b = r.z2_complex(lat.V)
    p = _prec(gnx_mod, prec)
    xs = {}
    # 16: the four-site groups, one per parity row at LT = 8 (the x2 hops of
    # both row ends inside one group)
    for shape in (0, 10) if prec == "f32" else (0, 10, 16):
        monkeypatch.setenv("GNX_MIX_SHAPE", str(shape))
        c = gnx_mod.Context(L)
        x = c.vec()
        # FP16 runs only the loads-first form: its reference is FP32's two-buffer form
        pp = gnx_mod.PREC_F32 if (prec == "f16" and shape == 0) else p
        st, res = c.cg(sm, dev(b), x, 1e-30, max_iter=cap, start=gnx_mod.START_IMPOSED, prec=pp)
        torch.cuda.synchronize()
        assert st == 3 and not res.converged and res.iterations == cap, (st, res.as_dict())
        xh = host(x)
        tr = true_res(lat, smh, b, xh)
        assert abs(tr - res.residual) <= 1e-9 * np.linalg.norm(b), (tr, res.residual)
        xs[shape] = xh
        assert c.cg_mixed_shape(1, p if shape else pp)["consecutive"] == (shape == 16)
    for shape in xs:
        assert rel(xs[shape], xs[0]) < (2e-2 if prec == "f16" else 1e-4), (shape, rel(xs[shape], xs[0]))


@pytest.mark.parametrize("prec", ["bf16", "f16"])
@pytest.mark.parametrize("dims", [(12, 8), (6, 24), (16, 32)])
def test_mixed_four_site_groups(P, gnx_mod, prec, dims):
    """The four-site group kernels on L^2 x LT lattices: 12^2 x 8 (V / 2 = 576,
    the last block partly past the last group, one group per parity row),
    6^2 x 24 (three groups per row) and 16^2 x 32 (x0, x1 hops across row
    blocks of 16 groups).  Two systems: each meets the FP64 bar against the
    oracle's operator, and the solution matches the oracle's FP64 solve."""
    L, LT = dims
    banner(f"mixed {prec} four-site groups {L}^2x{LT}")
    import os
    import torch
    lat = P.Lattice(L, LT)
    r = P.Random(500 + L + LT)
    smh = lat.sigma_mass(r.sigma(lat.V, P.sigma_variance(0.85)))
    bs = [r.z2_complex(lat.V) for _ in range(2)]
    ctx = gnx_mod.Context(L, LT)
    p = _prec(gnx_mod, prec)
    os.environ["GNX_MIX_SHAPE"] = "16"
    try:
        X = ctx.vec(2)
        st, res = ctx.cg_batch(dev(smh), dev(np.concatenate(bs)), X, 2, 1e-10, start=gnx_mod.START_IMPOSED,
                               prec=p)
        torch.cuda.synchronize()
    finally:
        del os.environ["GNX_MIX_SHAPE"]
    assert st == 0 and ctx.cg_mixed_shape(2, p)["consecutive"]
    for k in range(2):
        assert res[k].converged and res[k].reliable_updates >= 1
        xk = host(X[k * 2 * lat.V:(k + 1) * 2 * lat.V])
        assert true_res(lat, smh, bs[k], xk) < 1e-10
        xr, ok, _ = lat.cg(smh, bs[k], None, tol=1e-10, start="imposed")
        assert rel(xk, xr) < 1e-6


def test_mixed_four_site_groups_not_pinned_where_unfit(P, gnx_mod):
    """A pinned four-site shape that does not fit (FP32 storage, or LT not a
    multiple of 8) is ignored: the rule's shape runs and the solve converges."""
    banner("mixed: four-site shape ignored where it does not fit")
    import os
    L = 12
    lat = P.Lattice(L)
    r = P.Random(9)
    smh = lat.sigma_mass(r.sigma(lat.V, P.sigma_variance(0.85)))
    b = r.z2_complex(lat.V)
    os.environ["GNX_MIX_SHAPE"] = "16"
    try:
        for prec in (gnx_mod.PREC_F32, gnx_mod.PREC_F16):
            ctx = gnx_mod.Context(L)
            x = ctx.vec()
            st, res = ctx.cg(dev(smh), dev(b), x, 1e-10, start=gnx_mod.START_IMPOSED, prec=prec)
            assert st == 0 and res.converged
            assert not ctx.cg_mixed_shape(1, prec)["consecutive"]
            assert true_res(lat, smh, b, host(x)) < 1e-10
            ctx.close()
    finally:
        del os.environ["GNX_MIX_SHAPE"]


@pytest.mark.parametrize("mag", [1e-8, 1e6])
def test_mixed_f16_scaled_storage(P, gnx_mod, mag):
    """FP16 holds 6e-5 .. 65504: a right-hand side with elements of 1e-8 or
    1e6 would underflow or overflow unscaled.  The per-system scale (set at
    the start and at every reliable update) keeps the stored r, p, t O(1):
    the solve converges to the same relative accuracy as at magnitude 1 and
    matches the FP64 solve, in a batch with a magnitude-1 system."""
    banner(f"mixed f16 with |b| elements {mag:g}")
    import torch
    L = 12
    lat = P.Lattice(L)
    r = P.Random(123)
    smh = lat.sigma_mass(r.sigma(lat.V, P.sigma_variance(0.85)))
    sm = dev(smh)
    b1 = r.z2_complex(lat.V)
    b2 = mag * r.z2_complex(lat.V)
    ctx = gnx_mod.Context(L)
    tol = 1e-10 * mag
    X = ctx.vec(2)
    st, res = ctx.cg_batch(sm, dev(np.concatenate([b2, b1])), X, 2, tol, start=gnx_mod.START_IMPOSED,
                           prec=gnx_mod.PREC_F16)
    torch.cuda.synchronize()
    assert st == 3 or st == 0
    # system 0 (magnitude mag) to tol * mag; system 1 to tol * mag as well,
    # which for mag = 1e6 is loose and for 1e-8 is below FP64 reach: check
    # system 0 only, and system 1 with its own solve at 1e-10 below
    assert res[0].converged
    x0 = host(X[:2 * lat.V])
    assert true_res(lat, smh, b2, x0) < tol
    xr = ctx.vec()
    ctx.cg(sm, dev(b2), xr, tol, start=gnx_mod.START_IMPOSED)
    assert rel(x0, host(xr)) < 1e-6
    x1 = ctx.vec()
    s1, r1 = ctx.cg(sm, dev(b1), x1, 1e-10, start=gnx_mod.START_IMPOSED, prec=gnx_mod.PREC_F16)
    assert s1 == 0 and r1.converged and true_res(lat, smh, b1, host(x1)) < 1e-10
    # the iteration count does not depend on the magnitude
    assert abs(res[0].iterations - r1.iterations) <= 3, (res[0].iterations, r1.iterations)


def test_mixed_four_site_groups_streaming_stores(P, gnx_mod, monkeypatch):
    """40 FP16 systems at 64^3: the working set (r, t, two p, xs: 24 B per
    site and system, 252 MB) passes the Infinity Cache, so the kernels store
    non-temporally.  The four-site groups (shape 16) against the site-strided
    loads-first kernels (shape 15) on the same batch: every system converges,
    the two solutions agree to the solver tolerance, and two systems meet
    the FP64 bar against the oracle's operator."""
    banner("mixed f16 x40 at 64^3: four-site groups with streaming stores")
    import torch
    L, n = 64, 40
    lat = P.Lattice(L)
    r = P.Random(640)
    smh = lat.sigma_mass(r.sigma(lat.V, P.sigma_variance(0.85)))
    bs = [r.z2_complex(lat.V) for _ in range(n)]
    B = dev(np.concatenate(bs))
    sm = dev(smh)
    xs, its = {}, {}
    for shape in (15, 16):
        monkeypatch.setenv("GNX_MIX_SHAPE", str(shape))
        ctx = gnx_mod.Context(L)
        X = ctx.vec(n)
        st, res = ctx.cg_batch(sm, B, X, n, 1e-10, start=gnx_mod.START_IMPOSED, prec=gnx_mod.PREC_F16)
        torch.cuda.synchronize()
        assert st == 0 and all(q.converged for q in res)
        assert ctx.cg_mixed_shape(n, gnx_mod.PREC_F16)["consecutive"] == (shape == 16)
        xs[shape] = X.clone()
        its[shape] = [q.iterations for q in res]
        ctx.close()
    banner(f"  iterations shape 15 {min(its[15])}..{max(its[15])}, shape 16 {min(its[16])}..{max(its[16])}")
    assert max(abs(a - b) for a, b in zip(its[15], its[16])) <= 10
    d = float(torch.linalg.norm(xs[16] - xs[15]) / torch.linalg.norm(xs[15]))
    assert d < 1e-6, d
    for k in (0, n - 1):
        xk = host(xs[16][k * 2 * lat.V:(k + 1) * 2 * lat.V])
        assert true_res(lat, smh, bs[k], xk) < 1e-10
