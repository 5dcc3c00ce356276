"""Uneven batches (gnx_multi_cg): systems on lattices of different sizes,
each with its own sigma mass, solved together in one launch per kernel.

Bars (written here):
  - every system's solution bit-identical, and its iteration count and exit
    identical, to its own gnx_cg solve with one site per thread in the XCD
    slab mapping (GNX_CG_SHAPE=1: the same partition, arithmetic and order
    of sums);
  - against the oracle (the reference's CG restated): iterations within +-1,
    solution within 1e-8 relative at tol 1e-10; both starts;
  - null start: a system handed its own solution leaves in the prologue
    (x untouched) while the others iterate to their single-solve results;
  - iteration cap: every system stops at max_iter, reports no convergence,
    and holds the same iterate as its single solve.
"""
import numpy as np
import pytest

from gpu_util import banner, dev, host, rel

pytestmark = pytest.mark.gpu

SIZES = [(4, 4), (6, 10), (8, 8), (12, 12), (8, 8), (10, 6)]


def _systems(P, seed):
    out = []
    for i, (L, LT) in enumerate(SIZES):
        lat = P.Lattice(L, LT)
        r = P.Random(seed + i)
        sm = lat.sigma_mass(r.sigma(lat.V, P.sigma_variance(0.85))) + 0.25 * (i % 3)
        out.append((lat, sm, r.z2_complex(lat.V), r.z2_complex(lat.V)))
    return out


def _single(gnx_mod, lat, sm, b, x0, tol, start, max_iter=0):
    ctx = gnx_mod.Context(lat.L, lat.LT)
    x = dev(x0)
    st, res = ctx.cg(dev(sm), dev(b), x, tol, max_iter=max_iter, start=start, check=False)
    return st, res, host(x)


@pytest.mark.parametrize("start", ["imposed", "null"])
def test_multi_matches_single_and_oracle(P, gnx_mod, start, monkeypatch):
    banner(f"uneven batch {SIZES} {start}")
    monkeypatch.setenv("GNX_CG_SHAPE", "1")
    sy = _systems(P, 700)
    code = gnx_mod.START_IMPOSED if start == "imposed" else gnx_mod.START_NULL
    m = gnx_mod.Multi(SIZES)
    xs = [dev(x0) for _, _, _, x0 in sy]
    st, res = m.cg([dev(sm) for _, sm, _, _ in sy], [dev(b) for _, _, b, _ in sy], xs, 1e-10, start=code)
    assert st == 0 and all(r.converged for r in res)
    for k, (lat, sm, b, x0) in enumerate(sy):
        s1, r1, x1 = _single(gnx_mod, lat, sm, b, x0, 1e-10, code)
        assert s1 == 0
        assert (r1.iterations, r1.early_exit, r1.converged) == (res[k].iterations, res[k].early_exit, 1), k
        assert np.array_equal(host(xs[k]), x1), k
        xr, ok, ost = lat.cg(sm, b, x0, tol=1e-10, start=start)
        assert ok and abs(res[k].iterations - ost.last_iters) <= 1, (k, res[k].iterations, ost.last_iters)
        assert rel(host(xs[k]), xr) < 1e-8, k


def test_multi_null_start_prologue_exit(P, gnx_mod, monkeypatch):
    banner("uneven batch: one system starts from its solution")
    monkeypatch.setenv("GNX_CG_SHAPE", "1")
    sy = _systems(P, 800)
    m = gnx_mod.Multi(SIZES)
    sms = [dev(sm) for _, sm, _, _ in sy]
    bs = [dev(b) for _, _, b, _ in sy]
    xs = [dev(x0) for _, _, _, x0 in sy]
    st, _ = m.cg(sms, bs, xs, 1e-10, start=gnx_mod.START_IMPOSED)
    assert st == 0
    solved = host(xs[2]).copy()
    xs2 = [dev(x0) for _, _, _, x0 in sy]
    xs2[2] = dev(solved)
    # the prologue tests the true residual of the handed x: a looser tolerance
    # than the one it was solved to, so that the early exit is certain
    st, res = m.cg(sms, bs, xs2, 1e-8, start=gnx_mod.START_NULL)
    assert st == 0
    assert res[2].early_exit == 1 and res[2].iterations == 0
    assert np.array_equal(host(xs2[2]), solved)
    for k, (lat, sm, b, x0) in enumerate(sy):
        if k == 2:
            continue
        _, r1, x1 = _single(gnx_mod, lat, sm, b, x0, 1e-8, gnx_mod.START_NULL)
        assert res[k].early_exit == 0 and res[k].iterations == r1.iterations, k
        assert np.array_equal(host(xs2[k]), x1), k


def test_multi_iteration_cap(P, gnx_mod, monkeypatch):
    banner("uneven batch: iteration cap")
    monkeypatch.setenv("GNX_CG_SHAPE", "1")
    sy = _systems(P, 900)
    m = gnx_mod.Multi(SIZES)
    xs = [dev(x0) for _, _, _, x0 in sy]
    st, res = m.cg([dev(sm) for _, sm, _, _ in sy], [dev(b) for _, _, b, _ in sy], xs, 1e-10,
                   max_iter=5, start=gnx_mod.START_IMPOSED)
    assert st == 3  # GNX_ERR_NOCONV
    for k, (lat, sm, b, x0) in enumerate(sy):
        assert res[k].converged == 0 and res[k].iterations == 5, k
        s1, r1, x1 = _single(gnx_mod, lat, sm, b, x0, 1e-10, gnx_mod.START_IMPOSED, max_iter=5)
        assert s1 == 3 and r1.iterations == 5
        assert np.array_equal(host(xs[k]), x1), k


@pytest.mark.parametrize("start", ["imposed", "null"])
def test_multi_link_matches_single_and_oracle(P, gnx_mod, start, monkeypatch):
    """The link-coupled operator (gnx_multi_link_cg): every system with its
    own random U(1) links; bit-identical to its gnx_link_cg_batch solve with
    GNX_CG_SHAPE=1, and against the oracle's link-coupled CG (iterations
    +-1, solution within 1e-8)."""
    import torch
    banner(f"uneven batch, link-coupled {SIZES} {start}")
    monkeypatch.setenv("GNX_CG_SHAPE", "1")
    # random U(1) links: the mass shifted by 0.5 keeps M_U M_U^dag well
    # conditioned (as in the single-lattice link tests)
    sy = [(lat, sm + 0.5, b, x0) for lat, sm, b, x0 in _systems(P, 760)]
    links = [P.random_links(lat.V, 770 + k) for k, (lat, _, _, _) in enumerate(sy)]
    code = gnx_mod.START_IMPOSED if start == "imposed" else gnx_mod.START_NULL
    ufs = []
    for (lat, _, _, _), U in zip(sy, links):
        ctx = gnx_mod.Context(lat.L, lat.LT)
        uf = torch.zeros(6 * lat.V, dtype=torch.float64, device="cuda")
        ctx.links_from_lex(dev(U.view(np.float64)), uf)
        ufs.append((ctx, uf))
    m = gnx_mod.Multi(SIZES)
    xs = [dev(x0) for _, _, _, x0 in sy]
    st, res = m.cg([dev(sm) for _, sm, _, _ in sy], [dev(b) for _, _, b, _ in sy], xs, 1e-10, max_iter=20000,
                   start=code, links=[uf for _, uf in ufs])
    assert st == 0 and all(r.converged for r in res)
    for k, (lat, sm, b, x0) in enumerate(sy):
        ctx, uf = ufs[k]
        x1 = dev(x0)
        s1, r1 = ctx.link_cg_batch(dev(sm), uf, dev(b), x1, 1, 1e-10, max_iter=20000, start=code)
        assert s1 == 0
        assert (r1[0].iterations, r1[0].early_exit) == (res[k].iterations, res[k].early_exit), k
        assert np.array_equal(host(xs[k]), host(x1)), k
        xr, ok, ost = lat.link_cg(sm, links[k], b, x0=x0.copy(), tol=1e-10, start=start, max_iter=20000)
        assert ok and abs(res[k].iterations - ost.last_iters) <= 1, (k, res[k].iterations, ost.last_iters)
        assert rel(host(xs[k]), xr) < 1e-8, k
    # and the live operator again on the same solver (the other graph)
    st, res2 = m.cg([dev(sm) for _, sm, _, _ in sy], [dev(b) for _, _, b, _ in sy], [dev(x0) for *_, x0 in sy],
                    1e-10, max_iter=20000, start=code)
    assert st == 0 and all(r.converged for r in res2)
    m.close()


@pytest.mark.parametrize("prec", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("start", ["imposed", "null"])
def test_multi_mixed_matches_single_and_oracle(P, gnx_mod, prec, start, monkeypatch):
    """Mixed precision over an uneven batch (gnx_multi_cg_prec): every
    system bit-identical, with the same iterations and reliable updates, to
    its own gnx_cg solve in that storage with GNX_MIX_SHAPE=11 (loads-first
    kernels, one site per thread, slab mapping); against the oracle's FP64
    solve: solution within 1e-7 and the true residual (the oracle's
    operator) below 2 tol."""
    banner(f"uneven batch, mixed {prec} {SIZES} {start}")
    monkeypatch.setenv("GNX_MIX_SHAPE", "11")
    code = {"f32": gnx_mod.PREC_F32, "bf16": gnx_mod.PREC_BF16, "f16": gnx_mod.PREC_F16}[prec]
    scode = gnx_mod.START_IMPOSED if start == "imposed" else gnx_mod.START_NULL
    sy = _systems(P, 780)
    m = gnx_mod.Multi(SIZES)
    xs = [dev(x0) for _, _, _, x0 in sy]
    st, res = m.cg([dev(sm) for _, sm, _, _ in sy], [dev(b) for _, _, b, _ in sy], xs, 1e-10, start=scode,
                   prec=code)
    assert st == 0 and all(r.converged for r in res)
    for k, (lat, sm, b, x0) in enumerate(sy):
        ctx = gnx_mod.Context(lat.L, lat.LT)
        x1 = dev(x0)
        s1, r1 = ctx.cg(dev(sm), dev(b), x1, 1e-10, start=scode, prec=code)
        assert s1 == 0
        assert (r1.iterations, r1.reliable_updates, r1.early_exit) == \
            (res[k].iterations, res[k].reliable_updates, res[k].early_exit), k
        assert np.array_equal(host(xs[k]), host(x1)), k
        xr, ok, _ = lat.cg(sm, b, x0.copy(), tol=1e-10, start=start)
        xk = host(xs[k])
        assert ok and rel(xk, xr) < 1e-7, k
        assert np.linalg.norm(b - lat.M(sm, lat.Mdag(sm, xk))) < 2e-10, k
    m.close()


@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_multi_mixed_prologue_exit_and_cap(P, gnx_mod, prec, monkeypatch):
    """Mixed precision over an uneven batch, the exits: (1) null start with
    one system handed its own solution -- it leaves in the prologue, x
    untouched, while the others iterate to their single-solve results bit
    for bit; (2) the iteration cap (5) -- every system stops unconverged
    with the iterate of its single mixed solve under the same cap."""
    banner(f"uneven batch, mixed {prec}: prologue exit and cap")
    monkeypatch.setenv("GNX_MIX_SHAPE", "11")
    code = {"f32": gnx_mod.PREC_F32, "f16": gnx_mod.PREC_F16}[prec]
    sy = _systems(P, 820)
    m = gnx_mod.Multi(SIZES)
    sms = [dev(sm) for _, sm, _, _ in sy]
    bs = [dev(b) for _, _, b, _ in sy]
    xs = [dev(x0) for _, _, _, x0 in sy]
    st, _ = m.cg(sms, bs, xs, 1e-10, start=gnx_mod.START_IMPOSED, prec=code)
    assert st == 0
    solved = host(xs[1]).copy()
    xs2 = [dev(x0) for _, _, _, x0 in sy]
    xs2[1] = dev(solved)
    st, res = m.cg(sms, bs, xs2, 1e-8, start=gnx_mod.START_NULL, prec=code)
    assert st == 0
    assert res[1].early_exit == 1 and res[1].iterations == 0
    assert np.array_equal(host(xs2[1]), solved)
    for k, (lat, sm, b, x0) in enumerate(sy):
        if k == 1:
            continue
        ctx = gnx_mod.Context(lat.L, lat.LT)
        x1 = dev(x0)
        s1, r1 = ctx.cg(dev(sm), dev(b), x1, 1e-8, start=gnx_mod.START_NULL, prec=code)
        assert s1 == 0 and res[k].early_exit == 0 and res[k].iterations == r1.iterations, k
        assert np.array_equal(host(xs2[k]), host(x1)), k
    xs3 = [dev(x0) for _, _, _, x0 in sy]
    st, res = m.cg(sms, bs, xs3, 1e-10, max_iter=5, start=gnx_mod.START_IMPOSED, prec=code)
    assert st == 3
    for k, (lat, sm, b, x0) in enumerate(sy):
        assert res[k].converged == 0 and res[k].iterations == 5, k
        ctx = gnx_mod.Context(lat.L, lat.LT)
        x1 = dev(x0)
        s1, r1 = ctx.cg(dev(sm), dev(b), x1, 1e-10, max_iter=5, start=gnx_mod.START_IMPOSED, prec=code,
                        check=False)
        assert s1 == 3 and r1.iterations == 5
        assert np.array_equal(host(xs3[k]), host(x1)), k
    m.close()
