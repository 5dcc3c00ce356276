"""The reference's own driver flow, alone and over the drop-in.

oracle/driver_main.c runs the reference's main (test.c) with fixed
arguments (beta 0, beta_s 0.85, mass 0, sigma HMC eps 0.1 x 2 steps) and the
clock seed replaced by Initrand(20241), through UpdateLattice / TakeMeasures
(observables.c:178-310), and stops after the first measurement block: 100
LFMN2VF trajectories, each followed by ComputeCondensateDirac and
CondensateSigmaPi, then one observables line and one terma checkpoint.
tests/golden/ref_driver_L8.npz holds the three files the reference alone
wrote (oracle/_ref/ref_driver).

CPU: the oracle replays the same 100 updates and measurements from the same
seed and reproduces the acceptance lines (Delta-H as printed, accept flags)
and the observables line's averages bit for bit.
GPU: oracle/_ref/dropin_driver -- the same reference objects with every
function the drop-in exports bound to libgnx_ref.so -- writes the three
files through the reference's own output code; they match the reference's:
titles and parameters as text, Delta-H within 1e-8 with identical accept
flags, averages within 1e-6 relative (condensates) / 1e-9 (sigma
condensates), the average CG iteration count within 1, the checkpoint's
links byte for byte, sigma within 1e-9 and its counters byte for byte.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

DRIVER = os.path.join(ROOT, "oracle", "_ref", "dropin_driver")
SEED = 20241


@pytest.fixture(scope="module")
def ref_files():
    d = np.load(os.path.join(ROOT, "tests", "golden", "ref_driver_L8.npz"))
    return {k: d[k].tobytes() for k in d.files}


def _lines(b):
    return b.decode().rstrip("\n").split("\n")


def test_driver_files_replayed_by_oracle(ref_files, P):
    lat = P.Lattice(8)
    V = lat.V
    h = P.HMC(lat, 1)
    h.rng = P.Random(SEED)           # StartSeed -> Initrand(20241)
    h.sigma[:] = 0.3                  # START_WEAK_SIGMA (macro.h:65, sigma.c:34-37)
    h.sm[:] = lat.sigma_mass(h.sigma)
    acc_lines = _lines(ref_files["acceptance"])
    assert len(acc_lines) == 3 + 100
    cond, csig = [], []
    for t in range(100):
        tr = h.trajectory(0.1, 2, 1e-9, 1e-3)
        assert tr.ok
        f = acc_lines[3 + t].split("\t")
        assert f[11] == "%.16f" % tr.dH, t
        assert float(f[12]) == float(tr.accepted), t
        cond.append(h.condensate(1e-9))
        csig.append(h.condensate_sigma())
    obs = _lines(ref_files["data"])
    assert len(obs) == 4
    fields = obs[3].split("\t")
    # Mean (inputoutput.c:307-315): sum of a[i]/n in order
    def mean(vals):
        m = 0.0
        for v in vals:
            m += v / 100
        return m
    cond = np.array(cond)
    csig = np.array(csig)
    want = ["%.16f" % mean(cond[:, k]) for k in range(11)] + ["%.16f" % mean(csig[:, k]) for k in range(8)]
    assert fields[7:26] == want
    assert fields[0:5] == ["0", "8", "0.0000000000000000", "0.8500000000000000", "0.0000000000000000"]
    # the checkpoint holds the sigma after the 100 updates
    sigma_file = np.frombuffer(ref_files["terma"], np.float64, V, 48 * V)
    assert np.array_equal(sigma_file, h.sigma)


@pytest.mark.gpu
def test_driver_over_dropin(ref_files, tmp_path):
    if not os.path.exists(DRIVER):
        pytest.skip("oracle/_ref/dropin_driver is not built: it is linked from the reference's sources "
                    "(make -C oracle drivers where they are present)")
    print(f"[driver] running {DRIVER}", flush=True)
    subprocess.run([DRIVER, str(tmp_path)], check=True, timeout=300, stdout=subprocess.DEVNULL)
    base = tmp_path / "dati_simulation3d"
    got = {}
    for kind in ("data", "acceptance", "terma"):
        (d,) = [p for p in base.iterdir() if p.name.startswith(kind + "_")]
        (f,) = list(d.iterdir())
        got[kind] = f.read_bytes()
    # acceptance file: titles and every field as text but Delta-H
    ra, ga = _lines(ref_files["acceptance"]), _lines(got["acceptance"])
    assert len(ga) == len(ra) == 103 and ga[:3] == ra[:3]
    for t in range(3, 103):
        rf, gf = ra[t].split("\t"), ga[t].split("\t")
        assert gf[:11] == rf[:11] and gf[12] == rf[12], t
        assert abs(float(gf[11]) - float(rf[11])) <= 1e-8, (t, gf[11], rf[11])
    # observables file
    ro, go = _lines(ref_files["data"]), _lines(got["data"])
    assert len(go) == 4 and go[:3] == ro[:3]
    rf, gf = ro[3].split("\t"), go[3].split("\t")
    assert len(gf) == len(rf) == 26
    assert gf[:5] == rf[:5] and gf[6] == rf[6]          # parameters, distance per trajectory
    assert abs(float(gf[5]) - float(rf[5])) <= 1.0       # average CG iterations
    r = np.array([float(v) for v in rf[7:]])
    g = np.array([float(v) for v in gf[7:]])
    np.testing.assert_allclose(g[:11], r[:11], rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(g[11:], r[11:], rtol=1e-9, atol=1e-14)
    # checkpoint
    V = 512
    rt, gt = ref_files["terma"], got["terma"]
    assert len(gt) == len(rt) == 28748
    assert gt[:48 * V] == rt[:48 * V]                    # links (identity, untouched)
    np.testing.assert_allclose(np.frombuffer(gt, np.float64, V, 48 * V),
                               np.frombuffer(rt, np.float64, V, 48 * V), rtol=0, atol=1e-9)
    assert abs(np.frombuffer(gt, np.float64, 1, 56 * V)[0] - np.frombuffer(rt, np.float64, 1, 56 * V)[0]) <= 1e-6
    assert gt[56 * V + 8:] == rt[56 * V + 8:]            # counters and measure
