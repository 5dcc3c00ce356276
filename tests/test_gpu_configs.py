"""GPU parity at the workloads' own sizes: every configuration of the
benchmark plan runs here at the lattice sizes, batch sizes and tolerances it
names, with inputs drawn from the reference's own random stream (dSFMT-132049
through Initrand(seed), random.c:52-66): sigma by Gauss2MeanVariance pairs
(sigma.c:29-30, random.c:129-143) with variance 2/(N_f beta_s), complex Z2
sources by Casuale() <= 0.5 per component (condensate.c:57-58).

Bars (written here):
  configuration 1, 8^3, tol_high 1e-10 (the literal reading, eps 0.1 x 2
      steps) against the oracle's trajectories and measurements: Delta-H
      within 1e-9, identical accept decisions, condensates within 1e-7
      relative, CondensateSigmaPi within 1e-10 relative.  (The README's
      eps 0.2 series at the shipped tol 1e-9 runs against the reference's
      own golden records in test_gpu_hmc.py.)
  configuration 2, FP64 CG at 32^3 and 64^3, both starts, the default
      shape and the two-reduction (1) and one-reduction (7, 9) kernels
      pinned: iterations within +-1 of the oracle's, ||x - x_ref|| <= 1e-7
      ||x_ref||, true residual ||b - M M^dag x|| from the oracle's operators
      below 1.5e-10 (the absolute tolerance is 1e-10 on the recursive
      residual, conjugategradient.c:87-89); the null start from the
      converged x exits in the prologue with x untouched.
  configuration 3, mixed precision (FP32, BF16, FP16 storage) at 64^3 with
      1 and 32 right-hand sides and at 96^3 with 16: every solution's true
      residual from the oracle's operators below 1.5e-10; error against the
      oracle's FP64 solution (64^3, first source) and against the GPU's FP64
      batch (every source) within 1e-9 relative.
  configuration 4, 48^3 with 64 complex Z2 sources, batched condensate in
      FP64, FP32 and FP16 storage: every per-source term within 1e-9
      relative of the term from a single FP64 solve of that source, two of
      them within 1e-9 of the oracle's, and the estimate within 1e-6 of its
      statistical error of the single-solve estimate.
  configuration 5, 16 chains on {16, 24, 32, 48}^3 x beta_s {0.6, 0.75,
      0.85, 1.0}, links uniform in [0, 2 pi), beta 1, pure gauge on, 10-step
      sigma (LFMN2VF) and link (LeapFrog2Link) trajectories: with the
      one-site-per-thread slab CG shape pinned, every chain in lockstep
      bit-identical to its single-chain run; the 16^3 chains' sigma and link
      trajectories against the oracle's: Delta-H within 1e-8 (relative for
      the link sector's O(V) Delta-H), identical accept decisions.
"""
import numpy as np
import pytest

from gpu_util import banner, dev, host, rel

pytestmark = pytest.mark.gpu

TOL = 1e-10
TRUE_RES = 1.5e-10


def inputs(P, L, seed=1231, nsrc=1, beta_s=0.85):
    """sigma, sigma mass and nsrc complex Z2 sources from one dSFMT stream."""
    lat = P.Lattice(L)
    r = P.Random(seed)
    sigma = r.sigma(lat.V, P.sigma_variance(beta_s))
    bs = [r.z2_complex(lat.V) for _ in range(nsrc)]
    return lat, sigma, lat.sigma_mass(sigma), bs


def true_residual(lat, sm, b, x):
    return float(np.linalg.norm(b - lat.M(sm, lat.Mdag(sm, x))))


_oracle_cache = {}


def oracle_solution(P, L):
    """The oracle's imposed-start FP64 solve of configuration 2's system."""
    if L not in _oracle_cache:
        lat, sigma, sm, (b,) = inputs(P, L)
        xr, ok, st = lat.cg(sm, b, None, tol=TOL, start="imposed")
        assert ok
        _oracle_cache[L] = (lat, sm, b, xr, st.last_iters, true_residual(lat, sm, b, xr))
    return _oracle_cache[L]


# ---------------------------------------------------------------- configuration 2
@pytest.mark.parametrize("shape", [None, "1", "7", "9", "15"])
@pytest.mark.parametrize("L", [32, 64])
def test_config2_cg(P, gnx_mod, L, shape, monkeypatch):
    banner(f"config 2: FP64 CG {L}^3 shape={shape}")
    if shape is not None:
        monkeypatch.setenv("GNX_CG_SHAPE", shape)
    lat, sm, b, xr, iters_ref, tr_ref = oracle_solution(P, L)
    ctx = gnx_mod.Context(L)
    smd, bd = dev(sm), dev(b)
    x = ctx.vec()
    st, res = ctx.cg(smd, bd, x, TOL, start=gnx_mod.START_IMPOSED)
    assert st == 0 and res.converged and res.residual < TOL
    xg = host(x)
    tr = true_residual(lat, sm, b, xg)
    err = rel(xg, xr)
    banner(f"  iterations {res.iterations} (oracle {iters_ref}), true residual {tr:.3e} "
           f"(oracle {tr_ref:.3e}), rel err {err:.2e}")
    assert abs(res.iterations - iters_ref) <= 1
    assert err <= 1e-7
    assert tr < TRUE_RES
    # null start from the converged solution: the prologue exits, x untouched
    st, r2 = ctx.cg(smd, bd, x, TOL, start=gnx_mod.START_NULL)
    assert st == 0 and r2.early_exit == 1 and r2.iterations == 0
    assert np.array_equal(host(x), xg)
    # null start from a nonzero guess that is not converged: restarts from 0
    # (conjugategradient.c:31-33), so the same iterate sequence as imposed
    x2 = dev(0.5 * xg)
    st, r3 = ctx.cg(smd, bd, x2, TOL, start=gnx_mod.START_NULL)
    assert st == 0 and r3.early_exit == 0 and r3.iterations == res.iterations
    assert np.array_equal(host(x2), xg)


# ---------------------------------------------------------------- configuration 3
@pytest.mark.parametrize("prec", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("L,nrhs", [(64, 1), (64, 32), (96, 16)])
def test_config3_mixed(P, gnx_mod, L, nrhs, prec):
    banner(f"config 3: mixed {prec} {L}^3 x{nrhs}")
    import torch
    pc = {"f32": gnx_mod.PREC_F32, "bf16": gnx_mod.PREC_BF16, "f16": gnx_mod.PREC_F16}[prec]
    lat, sigma, sm, bs = inputs(P, L, nsrc=nrhs)
    V = lat.V
    ctx = gnx_mod.Context(L)
    smd = dev(sm)
    B = dev(np.concatenate(bs))
    X = ctx.vec(nrhs)
    st, res = ctx.cg_batch(smd, B, X, nrhs, TOL, start=gnx_mod.START_IMPOSED, prec=pc)
    assert st == 0 and all(r.converged for r in res)
    X64 = ctx.vec(nrhs)
    st, res64 = ctx.cg_batch(smd, B, X64, nrhs, TOL, start=gnx_mod.START_IMPOSED)
    assert st == 0
    torch.cuda.synchronize()
    Xh, X64h = host(X), host(X64)
    worst_tr, worst_err = 0.0, 0.0
    for k in range(nrhs):
        xk = Xh[2 * V * k:2 * V * (k + 1)]
        worst_tr = max(worst_tr, true_residual(lat, sm, bs[k], xk))
        worst_err = max(worst_err, rel(xk, X64h[2 * V * k:2 * V * (k + 1)]))
    its = [r.iterations for r in res]
    ups = [r.reliable_updates for r in res]
    banner(f"  iterations {min(its)}..{max(its)} (FP64 {max(r.iterations for r in res64)}), reliable updates "
           f"{max(ups)}, worst true residual {worst_tr:.3e}, worst rel err vs GPU FP64 {worst_err:.2e}")
    assert worst_tr < TRUE_RES
    assert worst_err < 1e-9
    if L == 64:
        _, _, b0, xr, _, _ = oracle_solution(P, 64)
        assert np.array_equal(b0, bs[0])
        e0 = rel(Xh[:2 * V], xr)
        banner(f"  rel err vs oracle FP64 (first source) {e0:.2e}")
        assert e0 < 1e-9


# ---------------------------------------------------------------- configuration 4
def test_config4_condensate(P, gnx_mod):
    banner("config 4: batched condensate 48^3 x 64 sources")
    import torch
    L, n = 48, 64
    lat, sigma, sm, noise = inputs(P, L, seed=4848, nsrc=n)
    V = lat.V
    ctx = gnx_mod.Context(L)
    smd = dev(sm)
    N = dev(np.concatenate(noise))
    # single FP64 solves, one source at a time (condensate.c:127-145 per vector)
    x1, y1 = ctx.vec(), ctx.vec()
    single = []
    for k in range(n):
        v = N[2 * V * k:2 * V * (k + 1)]
        st, r = ctx.cg(smd, v, x1, TOL, start=gnx_mod.START_IMPOSED)
        assert st == 0 and r.converged
        ctx.dirac(gnx_mod.OP_MDAG, smd, x1, y1)
        single.append(ctx.dot(v, y1).real / V)
    single = np.array(single)
    # two sources through the oracle: X = (M M^dag)^-1 v, Re(v^dag M^dag X) / V
    for k in (0, n - 1):
        xr, ok, _ = lat.cg(sm, noise[k], None, tol=TOL, start="imposed")
        assert ok
        term = float(np.dot(noise[k], lat.Mdag(sm, xr))) / V
        assert abs(single[k] - term) <= 1e-9 * abs(term), (k, single[k], term)
    est1 = single.mean()
    sem = single.std(ddof=1) / np.sqrt(n)
    X = ctx.vec(n)
    for name, pc in (("f64", gnx_mod.PREC_F64), ("f32", gnx_mod.PREC_F32), ("f16", gnx_mod.PREC_F16)):
        st, per, res = ctx.condensate_batch(smd, N, n, TOL, X, prec=pc)
        torch.cuda.synchronize()
        assert st == 0 and all(r.converged for r in res), name
        per = np.array(per)
        worst = float(np.max(np.abs(per - single) / np.abs(single)))
        banner(f"  {name}: worst per-source rel diff {worst:.2e}, estimate {per.mean():.12e} vs "
               f"{est1:.12e} (sem {sem:.3e})")
        assert worst <= 1e-9, name
        assert abs(per.mean() - est1) <= 1e-6 * sem, name


# ---------------------------------------------------------------- configuration 5
SIZES = [16, 24, 32, 48]
BETAS = [0.6, 0.75, 0.85, 1.0]
EPS, STEPS = 0.02, 10


def _chain_fields(P, L, bs, i):
    V = L ** 3
    r = P.Random(5000 + i)
    sigma = r.sigma(V, P.sigma_variance(bs))
    ang = 2.0 * np.pi * r.casuale(3 * V)
    return sigma, np.exp(1j * ang)


def test_config5_ensemble_lockstep(P, gnx_mod, monkeypatch):
    banner("config 5: 16 chains {16,24,32,48}^3, sigma + link trajectories, lockstep vs single")
    monkeypatch.setenv("GNX_CG_SHAPE", "1")
    chains = [(L, bs) for L in SIZES for bs in BETAS]
    single_s, ens_s, single_l, ens_l = [], [], [], []
    for i, (L, bs) in enumerate(chains):
        sigma, U = _chain_fields(P, L, bs, i)
        sm = P.Lattice(L).sigma_mass(sigma)
        for ls, ll in ((single_s, single_l), (ens_s, ens_l)):
            ctx = gnx_mod.Context(L)
            h = gnx_mod.HMC(ctx, beta_s=bs)
            h.set_sigma(sigma)
            ls.append(h)
            lh = gnx_mod.LinkHMC(ctx, beta=1.0, pure_gauge=True, linked=False)
            lh.set_fields(U, sm)
            ll.append(lh)
    Es, El = gnx_mod.Ensemble(ens_s), gnx_mod.Ensemble(ens_l)
    rng = np.random.default_rng(55)
    ds = [(rng.standard_normal(L ** 3), rng.standard_normal(L ** 3), rng.standard_normal(L ** 3))
          for L, _ in chains]
    st, res = Es.sigma_trajectory([d[0] for d in ds], [d[1] for d in ds], [d[2] for d in ds],
                                  EPS, STEPS, TOL, 1e-3)
    assert st == 0
    for i, h in enumerate(single_s):
        s1, r1 = h.trajectory(*ds[i], EPS, STEPS, TOL, 1e-3)
        assert s1 == 0 and r1.ok
        assert (res[i].dH, res[i].H0, res[i].H1, res[i].cg_iterations) == (r1.dH, r1.H0, r1.H1, r1.cg_iterations), \
            chains[i]
        h.accept(True)
        ens_s[i].accept(True)
        assert np.array_equal(h.get_sigma()[0], ens_s[i].get_sigma()[0]), chains[i]
    dl = [(rng.standard_normal(3 * L ** 3), rng.standard_normal(L ** 3), rng.standard_normal(L ** 3))
          for L, _ in chains]
    st, resl = El.link_trajectory([d[0] for d in dl], [d[1] for d in dl], [d[2] for d in dl], EPS, STEPS, TOL)
    assert st == 0
    for i, lh in enumerate(single_l):
        s1, r1 = lh.trajectory(*dl[i], EPS, STEPS, TOL)
        assert s1 == 0 and r1.ok
        assert (resl[i].dH, resl[i].cg_iterations) == (r1.dH, r1.cg_iterations), chains[i]
        lh.accept(False)
        ens_l[i].accept(False)
        assert np.array_equal(lh.get_links(), ens_l[i].get_links()), chains[i]
    banner(f"  sigma dH {[round(r.dH, 5) for r in res]}")
    Es.close()
    El.close()


@pytest.mark.parametrize("bs", BETAS)
def test_config5_small_chains_vs_oracle(P, gnx_mod, bs):
    """The 16^3 chains against the oracle's LFMN2VF and LeapFrog2Link,
    drawing from the oracle's own streams in the reference's order."""
    banner(f"config 5: 16^3 beta_s={bs} sigma and link trajectories vs oracle")
    L = 16
    lat = P.Lattice(L)
    V = lat.V
    seed = 6000 + int(100 * bs)
    ref = P.HMC(lat, seed, beta_s=bs)          # draws sigma from its stream first
    draws = P.Random(seed)
    sigma = draws.sigma(V, P.sigma_variance(bs))
    assert np.array_equal(sigma, ref.sigma)
    ctx = gnx_mod.Context(L)
    h = gnx_mod.HMC(ctx, beta_s=bs)
    h.set_sigma(sigma)
    for t in range(2):
        t_ref = ref.trajectory(EPS, STEPS, TOL, 1e-3)
        assert t_ref.ok
        mom = np.array([draws.gauss1() for _ in range(V)])
        chi0 = np.array([draws.gauss1() for _ in range(V)])
        chi1 = np.array([draws.gauss1() for _ in range(V)])
        st, r = h.trajectory(mom, chi0, chi1, EPS, STEPS, TOL, 1e-3)
        assert st == 0 and r.ok
        acc = draws.casuale() <= np.exp(-r.dH)
        h.accept(acc)
        banner(f"  sigma dH {r.dH:.10e} oracle {t_ref.dH:.10e}")
        assert abs(r.dH - t_ref.dH) <= 1e-8, (t, r.dH, t_ref.dH)
        assert acc == bool(t_ref.accepted)
    s_gpu, sm_gpu = h.get_sigma()
    np.testing.assert_allclose(s_gpu, ref.sigma, rtol=0, atol=1e-9)
    # link trajectories at the chain's sigma mass
    U = np.exp(1j * 2.0 * np.pi * P.Random(seed + 1).casuale(3 * V))
    lref = P.LinkHMC(lat, seed + 2, ref.sm, U, beta=1.0, pure_gauge=True, linked=False)
    ldraws = P.Random(seed + 2)
    lh = gnx_mod.LinkHMC(ctx, beta=1.0, pure_gauge=True, linked=False)
    lh.set_fields(U, ref.sm)
    for t in range(2):
        t_ref = lref.trajectory(EPS, STEPS, TOL)
        assert t_ref.ok
        mom = np.empty(3 * V)
        for i in range(3 * V // 2):
            mom[2 * i], mom[2 * i + 1] = ldraws.gauss2mv(0.0, 1.0)
        chi0 = np.array([ldraws.gauss1() for _ in range(V)])
        chi1 = np.array([ldraws.gauss1() for _ in range(V)])
        st, r = lh.trajectory(mom, chi0, chi1, EPS, STEPS, TOL)
        assert st == 0 and r.ok
        acc = ldraws.casuale() <= np.exp(-r.dH)
        lh.accept(acc)
        banner(f"  link dH {r.dH:.10e} oracle {t_ref.dH:.10e}")
        assert abs(r.dH - t_ref.dH) <= 1e-8 * max(1.0, abs(t_ref.dH)), (t, r.dH, t_ref.dH)
        assert acc == bool(t_ref.accepted)


# ---------------------------------------------------------------- configuration 1
def test_config1_hmc_tol_1e10_vs_oracle(P, gnx_mod):
    banner("config 1: 8^3 sigma HMC, eps 0.1 x 2 steps, tol_high 1e-10, 20 trajectories vs oracle")
    L, V = 8, 512
    lat = P.Lattice(L)
    ref = P.HMC(lat, 1231, beta_s=0.85)
    rng = P.Random(1231)
    sigma0 = rng.sigma(V, P.sigma_variance(0.85))
    assert np.array_equal(sigma0, ref.sigma)
    ctx = gnx_mod.Context(L)
    h = gnx_mod.HMC(ctx, 0.85)
    h.set_sigma(sigma0)
    for t in range(20):
        t_ref = ref.trajectory(0.1, 2, 1e-10, 1e-3)
        c_ref = ref.condensate(1e-10)
        s_ref = ref.condensate_sigma()
        mom = np.array([rng.gauss1() for _ in range(V)])
        chi0 = np.array([rng.gauss1() for _ in range(V)])
        chi1 = np.array([rng.gauss1() for _ in range(V)])
        st, r = h.trajectory(mom, chi0, chi1, 0.1, 2, 1e-10, 1e-3)
        assert st == 0 and r.ok
        acc = rng.casuale() <= np.exp(-r.dH)
        h.accept(acc)
        z2 = np.concatenate([rng.z2_real(V)[0::2] for _ in range(6)])
        cond = np.array(h.condensate(z2, 1e-10))
        csig = np.array(h.condensate_sigma())
        assert abs(r.dH - t_ref.dH) <= 1e-9, (t, r.dH, t_ref.dH)
        assert acc == bool(t_ref.accepted), t
        assert np.max(np.abs(cond - c_ref) / np.maximum(np.abs(c_ref), 1e-6)) < 1e-7, t
        assert np.max(np.abs(csig - s_ref) / np.maximum(np.abs(s_ref), 1e-12)) < 1e-10, t
