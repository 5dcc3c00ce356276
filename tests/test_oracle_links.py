"""Pins of the oracle's link sector (no reference output exists for it).

The reference's live operator ignores the links and no live path reaches
the link HMC (observables.c:127,183 are commented out), so the link-coupled
operator of copy_dirac.c:46-123, the link forces and LeapFrog2Link are
parity-unpinned against the reference.  These invariants fix them instead,
each chosen so that a plausible mistake (a conjugate dropped, a link taken
at the wrong site, a sign, a missing 1/2) fails:

  * U = 1 reduces M_U to the live M bit for bit (same add order), and r -= M_U v
    to r -= M v; M_U^dag to M^dag within rounding (gather vs scatter order);
  * dense M_U^dag equals the conjugate transpose of dense M_U bit for bit;
  * gauge covariance: with U^g_d(x) = g(x) U_d(x) conj g(x+d),
    M_{U^g} (g v) = g (M_U v);
  * the link CG returns the dense solution, both starts;
  * gauge force = -dS_G/dtheta and fermion link force = -dS_PF/dtheta by
    central finite differences, S_G = -beta sum Re Pi, S_PF = sum_f Re phi^dag (M M^dag)^-1 phi;
  * LeapFrog2Link is a second-order integrator: Delta-H shrinks ~4x when the
    step is halved at fixed trajectory length.
"""
import numpy as np
import pytest


def cplx(v):
    return np.ascontiguousarray(v, np.float64).view(np.complex128)


def dense(lat, fn, *args):
    V = lat.V
    A = np.zeros((V, V), dtype=np.complex128)
    for j in range(V):
        e = np.zeros(V, dtype=np.complex128)
        e[j] = 1.0
        A[:, j] = cplx(fn(*args, e))
    return A


def fields(P, L, LT, seed, shift=0.0):
    lat = P.Lattice(L, LT)
    r = P.Random(seed)
    sm = lat.sigma_mass(r.sigma(lat.V, P.sigma_variance(0.85))) + shift
    U = P.random_links(lat.V, seed)
    return lat, r, sm, U


@pytest.mark.parametrize("L,LT", [(4, 4), (4, 6)])
def test_unit_links_reduce_to_live_operator(P, L, LT):
    lat, r, sm, _ = fields(P, L, LT, 2)
    U1 = np.ones(3 * lat.V, dtype=np.complex128)
    v, t = r.z2_complex(lat.V), r.z2_complex(lat.V)
    np.testing.assert_array_equal(lat.link_M(sm, U1, v), lat.M(sm, v))
    np.testing.assert_array_equal(lat.link_sub_M(sm, U1, v, t), lat.sub_M(sm, v, t))
    np.testing.assert_allclose(lat.link_Mdag(sm, U1, v), lat.Mdag(sm, v), rtol=0, atol=1e-14)


@pytest.mark.parametrize("L,LT", [(4, 4), (4, 6)])
def test_link_adjoint_is_conjugate_transpose(P, L, LT):
    lat, r, sm, U = fields(P, L, LT, 3)
    M = dense(lat, lat.link_M, sm, U)
    Md = dense(lat, lat.link_Mdag, sm, U)
    np.testing.assert_array_equal(Md, M.conj().T)
    np.testing.assert_array_equal(np.diag(M).real, sm)


@pytest.mark.parametrize("L,LT", [(4, 4), (6, 4)])
def test_gauge_covariance(P, L, LT):
    lat, r, sm, U = fields(P, L, LT, 4)
    V = lat.V
    g = np.exp(1j * np.random.default_rng(9).uniform(0, 2 * np.pi, V))  # lex order
    f2lex = lat.table("f2lex", V)
    lex2f = lat.table("lex2f", V)
    fpp = lat.table("fpp", 3 * V).reshape(V, 3)
    Ug = U.copy()
    for x in range(V):
        for d in range(3):
            xp = f2lex[fpp[lex2f[x], d]]
            Ug[3 * x + d] = g[x] * U[3 * x + d] * np.conj(g[xp])
    gf = g[f2lex]  # f order
    v = cplx(r.z2_complex(V))
    for fn in (lat.link_M, lat.link_Mdag):
        lhs = cplx(fn(sm, Ug, gf * v))
        rhs = gf * cplx(fn(sm, U, v))
        np.testing.assert_allclose(lhs, rhs, rtol=0, atol=1e-13)


@pytest.mark.parametrize("start", ["null", "imposed"])
def test_link_cg_dense_solution(P, start):
    lat, r, sm, U = fields(P, 4, 6, 5, shift=0.5)
    b = cplx(r.z2_complex(lat.V))
    x0 = cplx(r.z2_complex(lat.V))
    M = dense(lat, lat.link_M, sm, U)
    xd = np.linalg.solve(M @ M.conj().T, b)
    x, ok, st = lat.link_cg(sm, U, b, x0=x0, tol=1e-12, start=start)
    assert ok and st.last_iters > 0
    np.testing.assert_allclose(cplx(x), xd, rtol=0, atol=1e-10)


def test_gauge_force_finite_difference(P):
    lat, r, sm, U = fields(P, 4, 4, 6)
    V, beta, h = lat.V, 1.3, 1e-6
    lex2f = lat.table("lex2f", V)
    F = lat.gauge_force(U, beta)
    for x, d in [(0, 0), (5, 1), (17, 2), (42, 0), (63, 2)]:
        up, um = U.copy(), U.copy()
        up[3 * x + d] *= np.exp(1j * h)
        um[3 * x + d] *= np.exp(-1j * h)
        fd = -(-beta * lat.plaquette_sum(up) + beta * lat.plaquette_sum(um)) / (2 * h)
        assert abs(F[3 * lex2f[x] + d] - fd) < 1e-7 * max(1.0, abs(fd))


def test_link_fermion_force_finite_difference(P):
    lat, r, sm, U = fields(P, 4, 4, 7, shift=1.0)
    V, h = lat.V, 1e-6
    lex2f = lat.table("lex2f", V)
    rng = np.random.default_rng(1)
    phis = [rng.normal(size=V) + 0j for _ in range(2)]

    def action(Uc):
        M = dense(lat, lat.link_M, sm, Uc)
        A = M @ M.conj().T
        return sum(np.vdot(p, np.linalg.solve(A, p)).real for p in phis)

    M = dense(lat, lat.link_M, sm, U)
    A = M @ M.conj().T
    X = [np.linalg.solve(A, p) for p in phis]
    Y = [M.conj().T @ x for x in X]
    F = lat.link_force_from_solutions(U, X, Y)
    for x, d in [(0, 0), (5, 1), (17, 2), (33, 0), (60, 1)]:
        up, um = U.copy(), U.copy()
        up[3 * x + d] *= np.exp(1j * h)
        um[3 * x + d] *= np.exp(-1j * h)
        fd = -(action(up) - action(um)) / (2 * h)
        assert abs(F[3 * lex2f[x] + d] - fd) < 1e-6 * max(1.0, abs(fd))


@pytest.mark.parametrize("linked", [1, 0])
def test_leapfrog2link_second_order(P, linked):
    """linked = 1: the consistent theory, Delta-H ~ eps^2.  linked = 0 is the
    reference as written: the solves use the live M (no links) while the
    force differentiates M_U, so the fermion 'force' does work that no term
    of H accounts for and Delta-H is O(tau), independent of the step."""
    lat, r, sm, U = fields(P, 6, 6, 8, shift=0.5)
    dh = []
    for eps, steps in ((0.04, 5), (0.02, 10)):
        hmc = P.LinkHMC(lat, 21, sm, U, beta=1.0, pure_gauge=True, linked=linked)
        t = hmc.trajectory(eps, steps, 1e-12, max_iter=10000)
        assert t.ok
        dh.append(abs(t.dH))
    if linked:
        assert 2.5 < dh[0] / dh[1] < 6.0, dh
    else:
        assert dh[0] > 1.0 and abs(dh[0] / dh[1] - 1.0) < 0.1, dh
