"""GPU sigma-HMC: 20 LFMN2VF trajectories at 8^3 with measurements, against
the reference's own run (golden) and the oracle: eps 0.1 and 0.2 (the
README's command, README.md:24) with 2 steps, and eps 0.05 with 4 steps.

Draws come from the host RNG in the reference's order (momenta, chi_0, chi_1,
Metropolis uniform after the trajectory, then 6V Z2 draws for the condensate).
Bars (written here):
  - Delta-H per trajectory within 1e-9 absolute of the reference's;
  - accept/reject identical trajectory by trajectory;
  - sigma after each trajectory within 1e-11 relative;
  - condensate outputs within 1e-8 relative (CG to 1e-9 absolute);
  - CondensateSigmaPi within 1e-12 relative (pure sums of sigma).
"""
import numpy as np
import pytest

from gpu_util import banner

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


@pytest.mark.parametrize("shape", [None, "7"])
@pytest.mark.parametrize("tag,eps,steps,ntraj", [("hmc_e0.1", 0.1, 2, 20), ("hmc_e0.2", 0.2, 2, 20),
                                                   ("hmc_e0.05s4", 0.05, 4, 5)])
def test_hmc_golden(golden, P, gnx_mod, tag, eps, steps, ntraj, shape, monkeypatch):
    """The default CG shapes and, pinned, the one-reduction kernels (shape 7)
    in every solve of the trajectory."""
    banner(f"hmc {tag} shape={shape}")
    if shape is not None:
        monkeypatch.setenv("GNX_CG_SHAPE", shape)
    import torch
    assert torch.cuda.is_available()
    L = 8
    V = L ** 3
    rng = P.Random(1231)
    sigma0 = rng.sigma(V, P.sigma_variance(0.85))
    assert np.array_equal(sigma0, golden[tag + "_sigma0"])
    ctx = gnx_mod.Context(L)
    h = gnx_mod.HMC(ctx, 0.85)
    h.set_sigma(sigma0)
    dh, acc, cond, cs = [], [], [], []
    for t in range(ntraj):
        mom = np.array([rng.gauss1() for _ in range(V)])
        chi0 = np.array([rng.gauss1() for _ in range(V)])
        chi1 = np.array([rng.gauss1() for _ in range(V)])
        st, r = h.trajectory(mom, chi0, chi1, eps, steps, 1e-9, 1e-3)
        assert st == 0 and r.ok
        u = rng.casuale()
        a = u <= np.exp(-r.dH)
        h.accept(a)
        dh.append(r.dH)
        acc.append(int(a))
        z2 = np.concatenate([rng.z2_real(V)[0::2] for _ in range(6)])
        cond.append(h.condensate(z2, 1e-9))
        cs.append(h.condensate_sigma())
        if t == 0:
            s1, _ = h.get_sigma()
            assert relerr(s1, golden[tag + "_sigma1"]) < 1e-11
    assert np.max(np.abs(np.array(dh) - golden[tag + "_dh"])) < 1e-9
    assert np.array_equal(acc, golden[tag + "_acc"])
    sN, _ = h.get_sigma()
    assert relerr(sN, golden[tag + "_sigmaN"]) < 1e-10
    cond = np.array(cond).ravel()
    gc = golden[tag + "_cond"]
    assert np.max(np.abs(cond - gc) / np.maximum(np.abs(gc), 1e-6)) < 1e-7
    cs = np.array(cs).ravel()
    gs = golden[tag + "_csig"]
    assert np.max(np.abs(cs - gs) / np.maximum(np.abs(gs), 1e-12)) < 1e-10
    it, calls = h.cg_stats()
    assert calls == golden[tag + "_calls"][-1]
    assert abs(it - golden[tag + "_iters"][-1]) <= 0.01 * golden[tag + "_iters"][-1]


def test_sigma_force_golden(golden, gnx_mod):
    """Force kernel alone on the reference's own solutions: bit-exact."""
    banner("sigma force from golden solutions")
    import torch
    from gpu_util import dev, host
    ctx = gnx_mod.Context(8)
    F = ctx.field()
    ctx.sigma_force_from_solutions(dev(golden["op_sigma"]), 0.85, dev(golden["force_X0"]), dev(golden["force_Y0"]),
                                   dev(golden["force_X1"]), dev(golden["force_Y1"]), F)
    torch.cuda.synchronize()
    assert np.array_equal(host(F), golden["force_F"])


def test_sigma_force_full(golden, gnx_mod):
    """CG + M^dag + force from phi (X starts at 0): force to the solver tolerance."""
    banner("sigma force with solves")
    import torch
    from gpu_util import dev, host
    ctx = gnx_mod.Context(8)
    V = 512
    sigma = dev(golden["op_sigma"])
    sm = ctx.field()
    ctx.sigma_mass(sigma, sm)
    phi = dev(np.concatenate([golden["force_phi0"], golden["force_phi1"]]))
    X = ctx.vec(2)
    Y = ctx.vec(2)
    F = ctx.field()
    st, res = ctx.sigma_force(sigma, sm, 0.85, 1e-10, [phi[:2 * V], phi[2 * V:]], [X[:2 * V], X[2 * V:]],
                              [Y[:2 * V], Y[2 * V:]], F)
    torch.cuda.synchronize()
    assert st == 0
    gF = golden["force_F"]
    assert np.max(np.abs(host(F) - gF)) < 1e-9 * np.max(np.abs(gF))
    assert relerr(host(X[:2 * V]), golden["force_X0"]) < 1e-6 or \
        np.linalg.norm(host(X[:2 * V]) - golden["force_X0"]) < 1e-8 * np.linalg.norm(golden["force_X0"])


@pytest.mark.parametrize("prec", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("tag,eps,steps,ntraj", [("hmc_e0.1", 0.1, 2, 20), ("hmc_e0.2", 0.2, 2, 20)])
def test_hmc_golden_mixed_force_solves(golden, P, gnx_mod, tag, eps, steps, ntraj, prec):
    """The MEDIUM-tolerance force solves (1e-3) in mixed precision, the
    HIGH-tolerance ones (energies, last force) in FP64: the force solves
    stop anywhere below 1e-3 on the FP64 true residual instead of at the
    reference's FP64 iterate, so the trajectories move by a part of that
    tolerance.  Bars: the same accept decisions as the reference's; Delta-H
    within 1e-6 of the reference's with FP32 storage and within 5e-4 with
    the 2-byte storage formats (whose iterates leave the FP64 ones sooner,
    so the 1e-3 stop lands elsewhere; Delta-H itself is 1e-4..1e-2 here);
    sigma after the run within 1e-6 (FP32) and 1e-4 (2-byte storage)."""
    banner(f"hmc {tag} with {prec} force solves")
    pc = {"f32": gnx_mod.PREC_F32, "bf16": gnx_mod.PREC_BF16, "f16": gnx_mod.PREC_F16}[prec]
    L = 8
    V = L ** 3
    rng = P.Random(1231)
    sigma0 = rng.sigma(V, P.sigma_variance(0.85))
    h = gnx_mod.HMC(gnx_mod.Context(L), 0.85)
    h.set_sigma(sigma0)
    h.set_force_precision(pc)
    dh, acc = [], []
    for t in range(ntraj):
        mom = np.array([rng.gauss1() for _ in range(V)])
        chi0 = np.array([rng.gauss1() for _ in range(V)])
        chi1 = np.array([rng.gauss1() for _ in range(V)])
        st, r = h.trajectory(mom, chi0, chi1, eps, steps, 1e-9, 1e-3)
        assert st == 0 and r.ok
        a = rng.casuale() <= np.exp(-r.dH)
        h.accept(a)
        dh.append(r.dH)
        acc.append(int(a))
        rng.z2_real(6 * V)  # the measurement's draws, not used here
    err = np.max(np.abs(np.array(dh) - golden[tag + "_dh"]))
    sN, _ = h.get_sigma()
    serr = np.max(np.abs(sN - golden[tag + "_sigmaN"]))
    banner(f"  max |dH - dH_ref| {err:.2e}, max |sigma - sigma_ref| {serr:.2e}")
    assert np.array_equal(acc, golden[tag + "_acc"])
    assert err < (1e-6 if prec == "f32" else 5e-4)
    assert serr < (1e-6 if prec == "f32" else 1e-4)
    _, calls = h.cg_stats()
    assert 0 < calls <= ntraj * (4 * steps + 4)  # prologue exits are not counted (conjugategradient.c:22-27)
