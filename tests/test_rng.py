"""The library's host random stream (include/gnx_rng.h, gnx.Rng) against
the reference's own stream recorded at Initrand(1231) and Initrand(77)
(tests/golden, written by the reference compiled from its sources) and
against the oracle's draws in bulk: bit-exact.  Host code only (no GPU)."""
import numpy as np
import pytest


def test_stream_matches_reference_golden(golden, gnx_mod):
    r = gnx_mod.Rng(1231)
    assert np.array_equal(r.casuale(6000), golden["rng_casuale_1231"])
    g = gnx_mod.Rng(77)
    vals = [g.gauss1() for _ in range(200)]
    for _ in range(100):
        vals += list(g.gauss2mv(0.0, 0.5))
    assert np.array_equal(np.array(vals), golden["rng_gauss_77"])
    # single draws and bulk draws continue one stream
    a, b = gnx_mod.Rng(5), gnx_mod.Rng(5)
    assert np.array_equal(np.array([a.casuale() for _ in range(3000)]), b.casuale(3000))
    assert np.array_equal(np.array([a.gauss1() for _ in range(500)]), b.gauss1(500))


@pytest.mark.parametrize("L,seed", [(8, 1231), (16, 4848), (64, 1231)])
def test_bulk_draws_match_oracle(P, gnx_mod, L, seed):
    V = L ** 3
    r, o = gnx_mod.Rng(seed), P.Random(seed)
    var = P.sigma_variance(0.85)
    assert np.array_equal(r.sigma(V, var), o.sigma(V, var))
    assert np.array_equal(r.z2_complex(V), o.z2_complex(V))
    assert np.array_equal(r.z2_real(V), o.z2_real(V)[0::2])
    assert np.array_equal(r.gauss1(1000), np.array([o.gauss1() for _ in range(1000)]))
    assert np.array_equal(r.casuale(10), o.casuale(10))
    pairs = np.array([v for _ in range(3 * V // 2) for v in o.gauss2mv(0.0, 1.0)]) if V <= 4096 else None
    if pairs is not None:
        assert np.array_equal(r.gauss2_pairs(3 * V), pairs)


def test_bad_arguments(gnx_mod):
    r = gnx_mod.Rng(3)
    with pytest.raises(gnx_mod.GnxError):
        r._fill(r.DRAW_SIGMA, 7, 7, 1.0)   # sigma pairs need an even count
    with pytest.raises(gnx_mod.GnxError):
        r._fill(99, 4, 4)
