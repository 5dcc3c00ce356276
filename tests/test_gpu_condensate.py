"""Batched stochastic condensate (gnx_condensate_batch) against the reference.

The reference estimates Tr M^-1 / V one noise vector at a time
(condensate.c:127-157): for each real Z2 vector v it solves M M^dag X = v,
forms M^dag X and accumulates Re(conj v . M^dag X) / V.  The batched path
solves all vectors in one launch.

Bars (written here):
  - against the reference's own outputs: after the first trajectory of the
    golden 8^3 run (seed 1231, eps 0.1, 2 steps) the four Z2 vectors the
    reference draws next give tr1, tr2 and tr4 (condensate.c:147-157) within
    1e-7 relative of the golden values (solver tolerance 1e-9, as the
    reference's HIGH_PRECISION_INVERSION);
  - against the oracle per vector: |per_source - oracle| <= 1e-9 |oracle| in
    FP64 and FP32-mixed at tol 1e-10 (complex Z2 sources, 12^3).
"""
import numpy as np
import pytest

from gpu_util import banner, dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [None, "7"])
def test_condensate_batch_golden(golden, P, gnx_mod, shape, monkeypatch):
    """The default CG shapes and, pinned, the one-reduction kernels (shape 7)."""
    banner(f"condensate batch vs golden 8^3 (trajectory 1 of hmc_e0.1), shape={shape}")
    if shape is not None:
        monkeypatch.setenv("GNX_CG_SHAPE", shape)
    lat = P.Lattice(8)
    V = lat.V
    h = P.HMC(lat, 1231)
    tr = h.trajectory(0.1, 2, 1e-9, 1e-3)
    assert tr.ok
    assert np.array_equal(h.sigma, golden["hmc_e0.1_sigma1"])
    vs = [h.rng.z2_real(V) for _ in range(4)]  # Z2NoiseFullRandomVec, condensate.c:52-60
    ctx = gnx_mod.Context(8)
    X = ctx.vec(4)
    st, per, res = ctx.condensate_batch(dev(h.sm), dev(np.concatenate(vs)), 4, 1e-9, X)
    assert st == 0 and all(r.converged for r in res)
    t0, t1, t2, t3 = per
    tr1 = (t0 + t1 + t2 + t3) / 4.0
    tr2 = (t0 * t1 + t0 * t2 + t0 * t3 + t1 * t2 + t1 * t3 + t2 * t3) / 6.0
    tr4 = t0 * t1 * t2 * t3
    g = golden["hmc_e0.1_cond"][:11]
    assert abs(tr1 - g[0]) <= 1e-7 * abs(g[0])
    assert abs(tr2 - g[2]) <= 1e-7 * abs(g[2])
    assert abs(tr4 - g[3]) <= 1e-7 * abs(g[3])


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_condensate_batch_vs_oracle(P, gnx_mod, prec):
    banner(f"condensate batch vs oracle 12^3 {prec}")
    L, n = 12, 8
    lat = P.Lattice(L)
    V = lat.V
    r = P.Random(404)
    sm = lat.sigma_mass(r.sigma(V, P.sigma_variance(0.85)))
    vs = [r.z2_complex(V) for _ in range(n)]
    ref = []
    for v in vs:
        x, ok, _ = lat.cg(sm, v, None, tol=1e-10, start="imposed")
        y = P.as_complex(lat.Mdag(sm, x))
        ref.append(np.sum((np.conj(P.as_complex(v)) * y).real) / V)
    ctx = gnx_mod.Context(L)
    X = ctx.vec(n)
    p = gnx_mod.PREC_F64 if prec == "f64" else gnx_mod.PREC_F32
    st, per, res = ctx.condensate_batch(dev(sm), dev(np.concatenate(vs)), n, 1e-10, X, prec=p)
    assert st == 0 and all(q.converged for q in res)
    for a, b in zip(per, ref):
        assert abs(a - b) <= 1e-9 * abs(b), (a, b)
