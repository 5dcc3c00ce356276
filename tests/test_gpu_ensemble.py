"""Chains on lattices of uneven size in lockstep (gnx_hmc_ensemble_trajectory,
gnx_link_hmc_ensemble_trajectory): every solve point of the integrator is
one uneven-batch launch over all chains.

Bars (written here):
  - with the one-site-per-thread slab CG shape pinned (GNX_CG_SHAPE=1, the
    partition and order of sums of the uneven solver), every chain's
    trajectory is bit-identical to its own single-chain trajectory on the
    same draws: dH, H0, H1, the CG counters, and the field after it;
  - a chain whose solve reaches the iteration cap leaves the lockstep as the
    single-chain trajectory does (ok = 0, sigma restored) while the others
    finish with their single-chain results.
"""
import numpy as np
import pytest

from gpu_util import banner, dev, host

pytestmark = pytest.mark.gpu

CHAINS = [(4, 0.6), (8, 0.85), (6, 1.0), (10, 0.75)]


def _draws(seed, V, nmom):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(nmom), rng.standard_normal(V), rng.standard_normal(V)


def _sigma(P, L, bs, seed):
    r = P.Random(seed)
    return r.sigma(L ** 3, P.sigma_variance(bs))


def _same(a, b, what):
    assert (a.dH == b.dH and a.H0 == b.H0 and a.H1 == b.H1 and a.ok == b.ok and a.cg_solves == b.cg_solves
            and a.cg_iterations == b.cg_iterations), (what, (a.dH, a.H0, a.ok, a.cg_iterations),
                                                      (b.dH, b.H0, b.ok, b.cg_iterations))


@pytest.mark.parametrize("fprec", [None, "f32", "f16"])
def test_sigma_ensemble_matches_single_chains(P, gnx_mod, fprec, monkeypatch):
    """FP64 solves, and the MEDIUM-tolerance force solves in mixed precision
    (gnx_hmc_set_force_precision on every chain: the lockstep's uneven mixed
    solve against each chain's own mixed batch of its two flavours, both in
    the loads-first one-site-per-thread shape, GNX_MIX_SHAPE=11)."""
    banner(f"sigma ensemble (4, 8, 6, 10)^3 vs single chains, force solves {fprec or 'f64'}")
    monkeypatch.setenv("GNX_CG_SHAPE", "1")
    monkeypatch.setenv("GNX_MIX_SHAPE", "11")
    codes = {"f32": gnx_mod.PREC_F32, "f16": gnx_mod.PREC_F16}
    single, ens = [], []
    for i, (L, bs) in enumerate(CHAINS):
        sig = _sigma(P, L, bs, 900 + i)
        for lst in (single, ens):
            h = gnx_mod.HMC(gnx_mod.Context(L), beta_s=bs)
            h.set_sigma(sig)
            if fprec is not None:
                h.set_force_precision(codes[fprec])
            lst.append(h)
    E = gnx_mod.Ensemble(ens)
    for t in range(2):
        draws = [_draws(100 * t + i, L ** 3, L ** 3) for i, (L, _) in enumerate(CHAINS)]
        st, res = E.sigma_trajectory([d[0] for d in draws], [d[1] for d in draws], [d[2] for d in draws],
                                     0.1, 3, 1e-10, 1e-3)
        assert st == 0
        for i, h in enumerate(single):
            s1, r1 = h.trajectory(*draws[i], 0.1, 3, 1e-10, 1e-3)
            assert s1 == 0 and r1.ok
            _same(res[i], r1, (t, CHAINS[i]))
            accept = (t + i) % 2 == 0
            h.accept(accept)
            ens[i].accept(accept)
            sa, ma = h.get_sigma()
            sb, mb = ens[i].get_sigma()
            assert np.array_equal(sa, sb) and np.array_equal(ma, mb), (t, CHAINS[i])
            assert h.cg_stats() == ens[i].cg_stats()
    E.close()


@pytest.mark.parametrize("linked", [False, True])
def test_link_ensemble_matches_single_chains(P, gnx_mod, linked, monkeypatch):
    """Both operator semantics: the live operator (linked = 0) and the
    link-coupled one (linked = 1, each chain's solves on M_U M_U^dag with its
    own links in the shared uneven-batch launch)."""
    banner(f"link ensemble (4, 8, 6, 10)^3 linked={linked} vs single chains")
    monkeypatch.setenv("GNX_CG_SHAPE", "1")
    single, ens = [], []
    for i, (L, bs) in enumerate(CHAINS):
        V = L ** 3
        # with random links, M_U M_U^dag needs the mass shifted to stay well
        # conditioned (as in the single-lattice link tests)
        sm = P.Lattice(L).sigma_mass(_sigma(P, L, bs, 950 + i)) + (0.5 if linked else 0.0)
        U = P.random_links(V, 960 + i)
        for lst in (single, ens):
            h = gnx_mod.LinkHMC(gnx_mod.Context(L), beta=1.0, pure_gauge=True, linked=linked)
            h.set_fields(U, sm)
            lst.append(h)
    E = gnx_mod.Ensemble(ens)
    for t in range(2):
        draws = [_draws(300 + 100 * t + i, L ** 3, 3 * L ** 3) for i, (L, _) in enumerate(CHAINS)]
        st, res = E.link_trajectory([d[0] for d in draws], [d[1] for d in draws], [d[2] for d in draws],
                                    0.02, 4, 1e-10, max_iter=20000)
        assert st == 0
        for i, h in enumerate(single):
            s1, r1 = h.trajectory(*draws[i], 0.02, 4, 1e-10, max_iter=20000)
            assert s1 == 0 and r1.ok
            _same(res[i], r1, (t, CHAINS[i]))
            accept = (t + i) % 2 == 1
            h.accept(accept)
            ens[i].accept(accept)
            assert np.array_equal(h.get_links(), ens[i].get_links()), (t, CHAINS[i])
            assert h.cg_stats() == ens[i].cg_stats()
    E.close()


def test_ensemble_chain_leaves_on_iteration_cap(P, gnx_mod, monkeypatch):
    """The cap lies between the first (imposed-start, tol 1e-10) solve of the
    small chain and that of the large one: the large chain fails there,
    restores sigma and leaves; the small chain's trajectory is unchanged."""
    banner("sigma ensemble, one chain at the iteration cap")
    import torch
    monkeypatch.setenv("GNX_CG_SHAPE", "1")
    chains = [(4, 0.85), (12, 0.85)]
    sig = [_sigma(P, L, bs, 990 + i) for i, (L, bs) in enumerate(chains)]
    draws = [_draws(17 + i, L ** 3, L ** 3) for i, (L, _) in enumerate(chains)]
    # iterations of each chain's first solve: phi_f = M chi_f from x = 0 to 1e-10
    first = []
    for i, (L, bs) in enumerate(chains):
        ctx = gnx_mod.Context(L)
        sm = ctx.field()
        ctx.sigma_mass(dev(sig[i]), sm)
        its = 0
        for f in (1, 2):
            chi = torch.zeros(2 * L ** 3, dtype=torch.float64, device=sm.device)
            chi[0::2] = dev(draws[i][f])
            phi = ctx.vec()
            ctx.dirac(gnx_mod.OP_M, sm, chi, phi)
            x = ctx.vec()
            _, r = ctx.cg(sm, phi, x, 1e-10, start=gnx_mod.START_IMPOSED)
            its = max(its, r.iterations)
        first.append(its)
    assert first[0] + 4 < first[1], first
    cap = (first[0] + first[1]) // 2
    single, ens = [], []
    for i, (L, bs) in enumerate(chains):
        for lst in (single, ens):
            h = gnx_mod.HMC(gnx_mod.Context(L), beta_s=bs)
            h.set_sigma(sig[i])
            lst.append(h)
    E = gnx_mod.Ensemble(ens)
    st, res = E.sigma_trajectory([d[0] for d in draws], [d[1] for d in draws], [d[2] for d in draws],
                                 0.1, 3, 1e-10, 1e-3, max_iter=cap)
    assert st == 3  # GNX_ERR_NOCONV: a chain failed
    s0, r0 = single[0].trajectory(*draws[0], 0.1, 3, 1e-10, 1e-3, max_iter=cap)
    assert s0 == 0 and r0.ok
    _same(res[0], r0, "small chain")
    s1, r1 = single[1].trajectory(*draws[1], 0.1, 3, 1e-10, 1e-3, max_iter=cap)
    assert s1 == 3 and not r1.ok and not res[1].ok
    for h in (single[1], ens[1]):
        assert np.array_equal(h.get_sigma()[0], sig[1])
    assert np.array_equal(single[0].get_sigma()[0], ens[0].get_sigma()[0])
    E.close()
