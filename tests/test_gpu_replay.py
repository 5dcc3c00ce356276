"""Replay of the graph-captured solvers: whole graphs of 32 iterations and
tails of 8, 16 or 24, launched per solve about as many as the last solve of
the same kind (graph, batch size, tolerance, start) needed, then one poll.

Bars (written here):
  - a solve replayed through tails is bit-identical to the same solve
    replayed through whole graphs, for every form whose buffers turn over
    with the slot index (two-reduction p parity, the one-reduction form,
    the deferred form's ring of 8 residuals, the fused form's state halves
    and ring), and meets the FP64 bar against the oracle;
  - a repeated solve of a kind polls once and launches at most the slots it
    used plus one tail (iterations + 2 + 7 for FP64, iterations + reliable
    updates + 2 + 7 for the mixed solver), also when two kinds alternate.
"""
import numpy as np
import pytest

from gpu_util import banner, dev, host

pytestmark = pytest.mark.gpu


def true_res(lat, sm, b, x):
    return np.linalg.norm(b - lat.M(sm, lat.Mdag(sm, x)))


def _system(P, L, seed):
    lat = P.Lattice(L)
    r = P.Random(seed)
    smh = lat.sigma_mass(r.sigma(lat.V, P.sigma_variance(0.85)))
    return lat, smh, r.z2_complex(lat.V)


def _solve_counted(ctx, fn):
    s0 = ctx.replay_stats()
    out = fn()
    s1 = ctx.replay_stats()
    return out, s1["slots"] - s0["slots"], s1["polls"] - s0["polls"]


@pytest.mark.parametrize("shape", [1, 9, 15, 18, 19, 21])
def test_tails_bit_identical_and_one_poll(P, gnx_mod, shape, monkeypatch):
    """Shapes: 1 two-reduction, 9 one-reduction, 15 deferred one-reduction,
    18/19 fused on 4 x 8 x 16 tiles (plain, deferred), 21 fused deferred on
    8 x 8 x 8 tiles.  The first solve of a context starts with one whole
    graph and misses (then a tail of 16, then whole graphs); the second
    launches whole graphs and one tail to its predicted need."""
    banner(f"replay tails, FP64 shape {shape}")
    import torch
    monkeypatch.setenv("GNX_CG_SHAPE", str(shape))
    L = 16
    lat, smh, b = _system(P, L, 61)
    ctx = gnx_mod.Context(L)
    sm = dev(smh)
    xs, its = [], []
    for k in range(2):
        x = ctx.vec()
        (st, res), slots, polls = _solve_counted(
            ctx, lambda: ctx.cg(sm, dev(b), x, 1e-10, start=gnx_mod.START_IMPOSED))
        torch.cuda.synchronize()
        assert st == 0 and res.converged
        assert ctx.cg_shape(1)["fused"] == (shape >= 18)
        xs.append(host(x))
        its.append(res.iterations)
        banner(f"  solve {k}: {res.iterations} iterations, {slots} slots, {polls} polls")
        if k == 1:
            assert polls == 1, polls
            assert res.iterations + 1 <= slots <= res.iterations + 2 + 7, (slots, res.iterations)
    assert its[0] == its[1]
    assert np.array_equal(xs[0], xs[1])
    assert true_res(lat, smh, b, xs[1]) < 1e-10
    ctx.close()


def test_prediction_per_kind(P, gnx_mod, monkeypatch):
    """Two kinds alternating on one context (a loose and a tight tolerance,
    and a batch of two): after one solve of each, every solve polls once and
    launches at most one tail more than it used."""
    banner("replay prediction per kind")
    import torch
    monkeypatch.delenv("GNX_CG_SHAPE", raising=False)
    L = 24
    lat, smh, b = _system(P, L, 62)
    b2 = P.Random(63).z2_complex(lat.V)
    ctx = gnx_mod.Context(L)
    sm = dev(smh)
    B = dev(np.concatenate([b, b2]))

    def tight():
        x = ctx.vec()
        return ctx.cg(sm, dev(b), x, 1e-10, start=gnx_mod.START_IMPOSED)

    def loose():
        x = ctx.vec()
        return ctx.cg(sm, dev(b), x, 1e-4, start=gnx_mod.START_IMPOSED)

    def pair():
        X = ctx.vec(2)
        st, res = ctx.cg_batch(sm, B, X, 2, 1e-8, start=gnx_mod.START_IMPOSED)
        return st, res[0] if res[0].iterations >= res[1].iterations else res[1]

    for fn in (tight, loose, pair):
        fn()
    for rnd in range(2):
        for name, fn in (("tight", tight), ("loose", loose), ("pair", pair)):
            (st, res), slots, polls = _solve_counted(ctx, fn)
            torch.cuda.synchronize()
            banner(f"  {name}: {res.iterations} iterations, {slots} slots, {polls} polls")
            assert st == 0 and res.converged
            assert polls == 1, (name, polls)
            assert res.iterations + 1 <= slots <= res.iterations + 2 + 7, (name, slots, res.iterations)
    ctx.close()


@pytest.mark.parametrize("prec", ["bf16", "f16"])
def test_mixed_repeat_one_poll(P, gnx_mod, prec, monkeypatch):
    """The mixed solver counts a slot pair per iteration and per reliable
    update: a repeated solve polls once, launches at most one tail more than
    it used, and returns the same bits."""
    banner(f"replay prediction, mixed {prec}")
    import torch
    monkeypatch.delenv("GNX_MIX_SHAPE", raising=False)
    L = 16
    lat, smh, b = _system(P, L, 64)
    ctx = gnx_mod.Context(L)
    sm = dev(smh)
    p = {"bf16": gnx_mod.PREC_BF16, "f16": gnx_mod.PREC_F16}[prec]
    xs = []
    for k in range(2):
        x = ctx.vec()
        (st, res), slots, polls = _solve_counted(
            ctx, lambda: ctx.cg(sm, dev(b), x, 1e-10, start=gnx_mod.START_IMPOSED, prec=p))
        torch.cuda.synchronize()
        assert st == 0 and res.converged
        xs.append(host(x))
        used = res.iterations + res.reliable_updates + 1
        banner(f"  solve {k}: {res.iterations} iterations, {res.reliable_updates} updates, "
               f"{slots} slots, {polls} polls")
        if k == 1:
            assert polls == 1, polls
            assert used <= slots <= used + 1 + 7, (slots, used)
    assert np.array_equal(xs[0], xs[1])
    assert true_res(lat, smh, b, xs[1]) < 1e-10
    ctx.close()


def test_prologue_exit_does_not_spoil_prediction(P, gnx_mod, monkeypatch):
    """The HMC's force solves alternate within one kind (same tolerance,
    null start) between solves that stop in their prologue (the incoming x
    already meets the tolerance) and solves that iterate.  After a prologue
    exit the next iterating solve launches one short burst, misses, and
    then launches what the kind's last iterating solve used: two polls,
    within one tail of its need."""
    banner("replay prediction after a prologue exit")
    import torch
    monkeypatch.delenv("GNX_CG_SHAPE", raising=False)
    L = 16
    lat, smh, b = _system(P, L, 65)
    ctx = gnx_mod.Context(L)
    sm = dev(smh)
    tol = 1e-8

    def from_zero():
        x = ctx.vec()
        return ctx.cg(sm, dev(b), x, tol, start=gnx_mod.START_NULL), x

    (st, res), x_conv = from_zero()
    from_zero()
    for rnd in range(2):
        xc = x_conv.clone()
        (st0, r0), slots0, polls0 = _solve_counted(
            ctx, lambda: ctx.cg(sm, dev(b), xc, tol, start=gnx_mod.START_NULL))
        torch.cuda.synchronize()
        assert st0 == 0 and r0.early_exit == 1 and r0.iterations == 0 and polls0 == 1
        ((st1, r1), _), slots1, polls1 = _solve_counted(ctx, from_zero)
        torch.cuda.synchronize()
        banner(f"  round {rnd}: prologue exit {slots0} slots; then {r1.iterations} iterations, "
               f"{slots1} slots, {polls1} polls")
        assert st1 == 0 and r1.converged and r1.iterations == res.iterations
        assert polls1 == 2, polls1
        assert r1.iterations + 1 <= slots1 <= r1.iterations + 2 + 7, (slots1, r1.iterations)
    ctx.close()
