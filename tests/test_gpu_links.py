"""GPU link sector (A11) against the oracle's link sector and the reference.

No driver path of the reference reaches its link functions, but they are
compiled and callable: the golden file holds the reference's own
AddForceFromFermionSectorRespectToLink and three LeapFrog2Link trajectories
(live-operator solves, as shipped), which the oracle reproduces bit for bit.
The link-coupled operator (copy_dirac.c, commented out in the reference) is
pinned by the invariants of tests/test_oracle_links.py.
Bars here (written):
  - links lex <-> f-order round trip exact;
  - M_U, M_U^dag, r -= M_U v bit-exact against the oracle (same products,
    same add order, no FMA);
  - gauge force and link fermion force (from the same X, Y) bit-exact;
    plaquette sum within 1e-13 relative (tree vs serial sum);
  - link CG: iterations within +-1 and solution within 1e-8 relative;
  - LeapFrog2Link driven with the oracle's draws: Delta-H within 1e-8
    (absolute, |Delta-H| < 1) and identical accept decisions over three
    trajectories, both for the reference's live-operator semantics
    (linked = 0) and the link-coupled theory (linked = 1); links after the
    trajectories within 1e-9;
  - against the reference's golden LeapFrog2Link: Delta-H within 1e-8
    relative, identical accept decisions, links within 1e-10.
"""
import numpy as np
import pytest

from gpu_util import banner, dev, host, rel

pytestmark = pytest.mark.gpu


def setup(P, L, LT, seed, shift=0.5):
    lat = P.Lattice(L, LT)
    r = P.Random(seed)
    sm = lat.sigma_mass(r.sigma(lat.V, P.sigma_variance(0.85))) + shift
    U = P.random_links(lat.V, seed)
    return lat, r, sm, U


@pytest.mark.parametrize("L,LT", [(8, 8), (6, 10)])
def test_link_ops_vs_oracle(P, gnx_mod, L, LT):
    banner(f"link ops {L}^2x{LT}")
    import torch
    lat, r, sm, U = setup(P, L, LT, 31, shift=0.0)
    V = lat.V
    ctx = gnx_mod.Context(L, LT)
    ulex = dev(U.view(np.float64))
    uf = torch.zeros_like(ulex)
    ctx.links_from_lex(ulex, uf)
    back = torch.zeros_like(ulex)
    ctx.links_to_lex(uf, back)
    assert np.array_equal(host(back), U.view(np.float64))
    v, t = r.z2_complex(V), r.z2_complex(V)
    smd = dev(sm)
    out = ctx.vec()
    ctx.link_dirac(gnx_mod.OP_M, smd, uf, dev(v), out)
    assert np.array_equal(host(out), lat.link_M(sm, U, v))
    ctx.link_dirac(gnx_mod.OP_MDAG, smd, uf, dev(v), out)
    assert np.array_equal(host(out), lat.link_Mdag(sm, U, v))
    out = dev(t)
    ctx.link_dirac(gnx_mod.OP_SUB_M, smd, uf, dev(v), out)
    assert np.array_equal(host(out), lat.link_sub_M(sm, U, v, t))


def test_link_forces_vs_oracle(P, gnx_mod):
    banner("gauge and link fermion forces 8^3")
    import torch
    lat, r, sm, U = setup(P, 8, 8, 32)
    V = lat.V
    ctx = gnx_mod.Context(8)
    uf = torch.zeros(6 * V, dtype=torch.float64, device="cuda")
    ctx.links_from_lex(dev(U.view(np.float64)), uf)
    F = torch.zeros(3 * V, dtype=torch.float64, device="cuda")
    ctx.gauge_force(uf, 1.0, F)
    Fo = lat.gauge_force(U, 1.0)
    assert np.array_equal(host(F), Fo)
    X = [r.z2_complex(V) for _ in range(2)]
    Y = [lat.link_Mdag(sm, U, x) for x in X]
    ctx.link_force_from_solutions(uf, dev(X[0]), dev(Y[0]), dev(X[1]), dev(Y[1]), F, accumulate=True)
    Fo2 = lat.link_force_from_solutions(U, X, Y, force=Fo.copy())
    assert np.array_equal(host(F), Fo2)
    p = ctx.plaquette_sum(uf)
    po = lat.plaquette_sum(U)
    assert abs(p - po) <= 1e-13 * max(1.0, abs(po))


@pytest.mark.parametrize("L,shape", [(8, None), (8, "7"), (8, "20"), (8, "21"), (16, "18"), (16, "19")])
@pytest.mark.parametrize("start", ["null", "imposed"])
def test_link_cg_vs_oracle(P, gnx_mod, start, L, shape, monkeypatch):
    """The default shape and, pinned, the one-reduction kernels (shape 7)
    and the fused kernels (one per iteration; 20/21 on one 8 x 8 x 8 tile
    that is the whole 8^3 lattice, 18/19 on 4 x 8 x 16 tiles at 16^3)."""
    banner(f"link cg {L}^3 {start} shape={shape}")
    if shape is not None:
        monkeypatch.setenv("GNX_CG_SHAPE", shape)
    import torch
    lat, r, sm, U = setup(P, L, L, 33)
    V = lat.V
    b = r.z2_complex(V)
    x0 = r.z2_complex(V)
    xr, ok, st = lat.link_cg(sm, U, b, x0=x0, tol=1e-10, start=start, max_iter=20000)
    assert ok
    ctx = gnx_mod.Context(L)
    uf = torch.zeros(6 * V, dtype=torch.float64, device="cuda")
    ctx.links_from_lex(dev(U.view(np.float64)), uf)
    x = dev(x0)
    s, res = ctx.link_cg_batch(dev(sm), uf, dev(b), x, 1, 1e-10, max_iter=20000,
                               start=gnx_mod.START_NULL if start == "null" else gnx_mod.START_IMPOSED)
    assert s == 0 and res[0].converged
    assert abs(res[0].iterations - st.last_iters) <= 1
    assert rel(host(x), xr) < 1e-8


@pytest.mark.parametrize("L,linked,spread", [(8, 0, None), (8, 1, None), (12, 1, 0.05), (16, 1, 0.05)])
def test_link_trajectories_vs_oracle(P, gnx_mod, L, linked, spread):
    """LeapFrog2Link against the oracle's: uniform random links at 8^3
    (both operator semantics), and the consistent link-coupled theory
    (linked = 1) at 12^3 and 16^3 from links near unity (phases
    N(0, spread^2)), where trajectories are accepted; Delta-H within 1e-8
    relative, identical accept decisions, links within 1e-9."""
    banner(f"LeapFrog2Link {L}^3 linked={linked} spread={spread}")
    lat, r, sm, U = setup(P, L, L, 34)
    if spread is not None:
        U = np.exp(1j * spread * np.random.default_rng(L).normal(size=3 * lat.V))
    V = lat.V
    eps, steps, tol, beta = 0.02, 10, 1e-10, 1.0
    ref = P.LinkHMC(lat, 777, sm, U, beta=beta, pure_gauge=True, linked=linked)
    draws = P.Random(777)  # the same stream, drawn here for the GPU
    ctx = gnx_mod.Context(L)
    h = gnx_mod.LinkHMC(ctx, beta=beta, pure_gauge=True, linked=linked)
    h.set_fields(U, sm)
    for k in range(3):
        t_ref = ref.trajectory(eps, steps, tol, max_iter=20000)
        assert t_ref.ok
        mom = np.empty(3 * V)
        for i in range(3 * V // 2):
            mom[2 * i], mom[2 * i + 1] = draws.gauss2mv(0.0, 1.0)
        chi0 = np.array([draws.gauss1() for _ in range(V)])
        chi1 = np.array([draws.gauss1() for _ in range(V)])
        st, t = h.trajectory(mom, chi0, chi1, eps, steps, tol, max_iter=20000)
        assert st == 0 and t.ok
        acc = draws.casuale() <= np.exp(-t.dH)
        h.accept(acc)
        assert abs(t.dH - t_ref.dH) <= 1e-8 * max(1.0, abs(t_ref.dH)), (k, t.dH, t_ref.dH)
        assert acc == bool(t_ref.accepted)
    np.testing.assert_allclose(h.get_links(), ref.links(), rtol=0, atol=1e-9)


def test_link_trajectories_vs_reference_golden(golden, P, gnx_mod):
    """The reference's own LeapFrog2Link (as shipped: live-operator solves,
    PURE_GAUGE off) recorded at 8^3: Delta-H within 1e-8, identical accept
    decisions, links after three trajectories within 1e-10, CG counters
    within one iteration per solve."""
    banner("LeapFrog2Link vs reference golden 8^3")
    lat = P.Lattice(8)
    V = lat.V
    U = golden["link_U"].view(np.complex128)
    sm = golden["link_sigmamass"]
    draws = P.Random(4444)
    ctx = gnx_mod.Context(8)
    h = gnx_mod.LinkHMC(ctx, beta=1.0, pure_gauge=False, linked=False)
    h.set_fields(U, sm)
    for k in range(3):
        mom = np.empty(3 * V)
        for i in range(3 * V // 2):
            mom[2 * i], mom[2 * i + 1] = draws.gauss2mv(0.0, 1.0)
        chi0 = np.array([draws.gauss1() for _ in range(V)])
        chi1 = np.array([draws.gauss1() for _ in range(V)])
        st, t = h.trajectory(mom, chi0, chi1, 0.02, 10, 1e-9)
        assert st == 0 and t.ok
        acc = draws.casuale() <= np.exp(-t.dH)
        h.accept(acc)
        g = golden["link_traj_dh"][k]
        assert abs(t.dH - g) <= 1e-8 * max(1.0, abs(g)), (k, t.dH, g)
        assert acc == bool(golden["link_traj_acc"][k])
    np.testing.assert_allclose(h.get_links().view(np.float64), golden["link_traj_U"], rtol=0, atol=1e-10)
    it, calls = h.cg_stats()
    assert calls == golden["link_traj_stat"][1]
    assert abs(it - golden["link_traj_stat"][0]) <= calls
