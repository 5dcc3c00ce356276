"""Mathematical pins of the oracle beyond the 8^3 golden vectors.

The golden file pins every oracle routine bit-exactly at the size the
reference ships (8^3).  These checks cover other lattice shapes and the
properties a plausible mistake would break (a transposed operand, a wrong
sign, a dropped corner, a swapped L/LT):

  geometry    : closed forms of coordinates, f-index, neighbours, eta phases
                and the antiperiodic time boundary (coordinate.c:25-65,
                133-224, 301-374) on non-cubic shapes;
  M, M^dag    : dense M^dag equals the transpose of dense M entry by entry
                (D is real antisymmetric, dirac.c:46-123), M + M^dag = 2 sm;
  M M^dag, CG : free-field spectrum at constant sigma mass c: plane waves with
                p_space = 2 pi n / L, p_time = (2n+1) pi / LT are eigenvectors
                with eigenvalue c^2 + sum_mu sin^2 p_mu, so CG returns
                b / lambda (conjugategradient.c:13-182);
  sigma force : central finite difference of the action
                S = 1/2 sum_f phi_f^dag (M M^dag)^-1 phi_f + (N_f beta_s / 4) sum sigma^2
                computed with dense linear algebra equals the oracle force,
                F = -dS/dsigma (moleculardynamics.c:576-647, 299-319).
"""
import ctypes as C

import numpy as np
import pytest


def closed_geometry(L, LT):
    V = L * L * LT
    lex = np.arange(V)
    x0, x1, x2 = lex // (L * LT), (lex // LT) % L, lex % LT
    par = (x0 + x1 + x2) & 1
    f_of_lex = lex // 2 + par * (V // 2)
    lex_of_f = np.empty(V, dtype=np.int64)
    lex_of_f[f_of_lex] = lex

    def lexc(a, b, c):
        return ((a % L) * L + (b % L)) * LT + (c % LT)

    pp = np.stack([lexc(x0 + 1, x1, x2), lexc(x0, x1 + 1, x2), lexc(x0, x1, x2 + 1)], 1)
    mm = np.stack([lexc(x0 - 1, x1, x2), lexc(x0, x1 - 1, x2), lexc(x0, x1, x2 - 1)], 1)
    eta = np.stack([np.ones(V), (-1.0) ** x0, (-1.0) ** (x0 + x1)], 1)
    eta_pp = eta.copy()
    eta_mm = eta.copy()
    eta_pp[x2 == LT - 1, 2] *= -1
    eta_mm[x2 == 0, 2] *= -1
    # f-order tables
    o = lex_of_f
    return dict(V=V, f_of_lex=f_of_lex, lex_of_f=lex_of_f, coords=np.stack([x0, x1, x2], 1),
                fpp=f_of_lex[pp[o]], fmm=f_of_lex[mm[o]], eta_pp=eta_pp[o], eta_mm=eta_mm[o], eta_pbc=eta[o])


@pytest.mark.parametrize("L,LT", [(4, 4), (4, 6), (6, 4), (6, 10), (8, 8)])
def test_geometry_closed_form(P, L, LT):
    lat = P.Lattice(L, LT)
    c = closed_geometry(L, LT)
    V = c["V"]
    assert lat.V == V
    np.testing.assert_array_equal(lat.table("lex2f", V), c["f_of_lex"])
    np.testing.assert_array_equal(lat.table("f2lex", V), c["lex_of_f"])
    np.testing.assert_array_equal(lat.table("coords", 3 * V).reshape(V, 3), c["coords"])
    fpp = lat.table("fpp", 3 * V).reshape(V, 3)
    fmm = lat.table("fmm", 3 * V).reshape(V, 3)
    np.testing.assert_array_equal(fpp, c["fpp"])
    np.testing.assert_array_equal(fmm, c["fmm"])
    np.testing.assert_array_equal(lat.table("eta_pp", 3 * V).reshape(V, 3), c["eta_pp"])
    np.testing.assert_array_equal(lat.table("eta_mm", 3 * V).reshape(V, 3), c["eta_mm"])
    np.testing.assert_array_equal(lat.table("eta_pbc", 3 * V).reshape(V, 3), c["eta_pbc"])
    # pp(mm(x)) = x and every hop flips the f-parity half
    for d in range(3):
        np.testing.assert_array_equal(fpp[fmm[:, d], d], np.arange(V))
        assert np.all((fpp[:, d] >= V // 2) == (np.arange(V) < V // 2))


def dense(lat, fn, sm):
    V = lat.V
    A = np.zeros((V, V))
    for j in range(V):
        e = np.zeros(V, dtype=np.complex128)
        e[j] = 1.0
        A[:, j] = P_as_complex(fn(sm, e)).real
    return A


def P_as_complex(v):
    return np.ascontiguousarray(v, np.float64).view(np.complex128)


@pytest.mark.parametrize("L,LT", [(4, 4), (4, 6)])
def test_adjoint_is_transpose(P, L, LT):
    lat = P.Lattice(L, LT)
    r = P.Random(3)
    sm = lat.sigma_mass(r.sigma(lat.V, 0.5))
    M = dense(lat, lat.M, sm)
    Md = dense(lat, lat.Mdag, sm)
    # every entry is sm or +-1/2 exactly, so the transpose holds bit for bit
    np.testing.assert_array_equal(Md, M.T)
    np.testing.assert_array_equal(np.diag(M), sm)
    np.testing.assert_array_equal(M + Md, np.diag(2 * sm))
    # r - Mv and r - M^dag v are the subtracting forms of the same operators
    v = r.z2_complex(lat.V)
    t = r.z2_complex(lat.V)
    np.testing.assert_allclose(lat.sub_M(sm, v, t), t - lat.M(sm, v), rtol=0, atol=1e-14)
    np.testing.assert_allclose(lat.sub_Mdag(sm, v, t), t - lat.Mdag(sm, v), rtol=0, atol=1e-14)


@pytest.mark.parametrize("L,LT", [(4, 4), (4, 6), (6, 4)])
def test_free_field_spectrum_and_cg(P, L, LT):
    lat = P.Lattice(L, LT)
    V = lat.V
    c = 0.375  # dyadic: sm = c exactly
    sm = np.full(V, c)
    coords = lat.table("coords", 3 * V).reshape(V, 3)[lat.table("f2lex", V)]
    rng = np.random.default_rng(0)
    for _ in range(4):
        n = rng.integers(0, [L, L, LT])
        p = np.array([2 * np.pi * n[0] / L, 2 * np.pi * n[1] / L, (2 * n[2] + 1) * np.pi / LT])
        b = np.exp(1j * coords @ p)
        lam = c * c + np.sum(np.sin(p) ** 2)
        y = P_as_complex(lat.M(sm, lat.Mdag(sm, b)))
        np.testing.assert_allclose(y, lam * b, rtol=0, atol=1e-13)
        x, ok, st = lat.cg(sm, b, tol=1e-12, start="imposed")
        assert ok
        np.testing.assert_allclose(P_as_complex(x), b / lam, rtol=0, atol=1e-11)


def test_sigma_force_finite_difference(P):
    L, LT = 4, 4
    lat = P.Lattice(L, LT)
    V = lat.V
    beta_s = 0.85
    r = P.Random(11)
    sigma = r.sigma(V, P.sigma_variance(beta_s)) + 0.6
    phis = [P_as_complex(r.z2_complex(V)) for _ in range(2)]

    def action(s):
        sm = lat.sigma_mass(s)
        M = dense(lat, lat.M, sm)
        A = M @ M.T
        val = sum(0.5 * np.vdot(ph, np.linalg.solve(A, ph)).real for ph in phis)
        return val + (4 * beta_s / 4.0) * np.sum(s * s)

    sm = lat.sigma_mass(sigma)
    M = dense(lat, lat.M, sm)
    A = M @ M.T
    Xs = [np.linalg.solve(A, ph) for ph in phis]
    Ys = [M.T @ X for X in Xs]
    Xp = [np.ascontiguousarray(X).view(np.float64) for X in Xs]
    Yp = [np.ascontiguousarray(Y).view(np.float64) for Y in Ys]
    dp = C.POINTER(C.c_double)
    F = np.zeros(V)
    P.lib().orc_sigma_force_from_solutions(
        C.byref(lat.g), sigma, beta_s,
        (dp * 2)(*[X.ctypes.data_as(dp) for X in Xp]),
        (dp * 2)(*[Y.ctypes.data_as(dp) for Y in Yp]), F)
    h = 1e-5
    for x in [0, 5, V // 2, V // 2 + 7, V - 1]:
        e = np.zeros(V)
        e[x] = h
        fd = -(action(sigma + e) - action(sigma - e)) / (2 * h)
        assert abs(F[x] - fd) < 1e-6 * max(1.0, abs(fd)), (x, F[x], fd)
