"""Pin the CPU oracle against golden vectors produced by the reference itself
(its own sources compiled at 8^3, oracle/ref_harness.c).  Every comparison is
bit-exact: the oracle keeps the reference's expression shapes and order."""
import numpy as np
import pytest


def test_rng_stream(golden, P):
    r = P.Random(1231)
    assert np.array_equal(r.casuale(6000), golden["rng_casuale_1231"])
    g = P.Random(77)
    vals = [g.gauss1() for _ in range(200)]
    for _ in range(100):
        vals += list(g.gauss2mv(0.0, 0.5))
    assert np.array_equal(np.array(vals), golden["rng_gauss_77"])


@pytest.mark.parametrize("name,key,mult", [
    ("fpp", "fcor_fermionpp", 3), ("fmm", "fcor_fermionmm", 3),
    ("eta_pp", "fcor_eta_pp", 3), ("eta_mm", "fcor_eta_mm", 3), ("eta_pbc", "fcor_eta_pbc", 3),
    ("f2lex", "fcor_site", 1), ("lex2f", "lat_fermionsite", 1), ("coords", "lat_lattice_site", 3)])
def test_geometry_tables(golden, P, name, key, mult):
    lat = P.Lattice(8)
    assert np.array_equal(lat.table(name, mult * lat.V), golden[key])


@pytest.fixture(scope="module")
def fields(golden, P):
    lat = P.Lattice(8)
    r = P.Random(1231)
    sigma = r.sigma(lat.V, P.sigma_variance(0.85))
    sm = lat.sigma_mass(sigma)
    b = r.z2_complex(lat.V)
    c = r.z2_complex(lat.V)
    return lat, sigma, sm, b, c


def test_inputs(golden, fields):
    lat, sigma, sm, b, c = fields
    assert np.array_equal(sigma, golden["op_sigma"])
    assert np.array_equal(b, golden["op_b"]) and np.array_equal(c, golden["op_c"])


def test_sigma_mass(golden, fields):
    lat, sigma, sm, b, c = fields
    assert np.array_equal(sm, golden["op_sigmamass"])


def test_dirac_ops(golden, fields):
    lat, sigma, sm, b, c = fields
    assert np.array_equal(lat.M(sm, b), golden["op_Mb"])
    assert np.array_equal(lat.Mdag(sm, b), golden["op_Mdagb"])
    assert np.array_equal(lat.sub_M(sm, b, c), golden["op_c_minus_Mb"])
    assert np.array_equal(lat.sub_Mdag(sm, b, c), golden["op_c_minus_Mdagb"])


@pytest.mark.parametrize("tol", ["1e-10", "1e-6", "1e-3"])
def test_cg(golden, fields, tol):
    lat, sigma, sm, b, c = fields
    x, ok, st = lat.cg(sm, b, c, tol=float(tol))
    ref = golden["cg_null_stat_" + tol]
    assert np.array_equal(x, golden["cg_null_x_" + tol])
    assert ok == bool(ref[0]) and st.iterations == ref[1] and st.cg_call == ref[2]
    # second call from the converged x exits in the prologue, uncounted
    x2, ok2, st2 = lat.cg(sm, b, x, tol=float(tol))
    ref2 = golden["cg_null_again_stat_" + tol]
    assert ok2 and st2.iterations == ref2[1] == 0 and st2.cg_call == ref2[2] == 0
    assert np.array_equal(x2, golden["cg_null_again_x_" + tol])
    x3, ok3, st3 = lat.cg(sm, b, c, tol=float(tol), start="imposed")
    ref3 = golden["cg_imp_stat_" + tol]
    assert np.array_equal(x3, golden["cg_imp_x_" + tol])
    assert st3.iterations == ref3[1]


def test_sigma_force(golden, P):
    import ctypes as C
    lat = P.Lattice(8)
    r = P.Random(1231)
    sigma = r.sigma(lat.V, P.sigma_variance(0.85))
    sm = lat.sigma_mass(sigma)
    assert np.array_equal(sm, golden["op_sigmamass"])
    h = P.HMC(lat, 1)
    h.sigma[:] = sigma
    h.sm[:] = sm
    V = lat.V
    for f in range(2):
        dst = np.ctypeslib.as_array(h.kit.phi[f], shape=(2 * V,))
        dst[:] = golden[f"force_phi{f}"]
        np.ctypeslib.as_array(h.kit.inv_phi[f], shape=(2 * V,))[:] = 0.0
    F = np.zeros(V)
    ok = P.lib().orc_sigma_force(C.byref(lat.g), h.sigma, h.sm, 0.85, 1e-10, V + 1, C.byref(h.kit), F)
    assert ok
    assert np.array_equal(F, golden["force_F"])
    for f in range(2):
        assert np.array_equal(np.ctypeslib.as_array(h.kit.inv_phi[f], shape=(2 * V,)), golden[f"force_X{f}"])
        assert np.array_equal(np.ctypeslib.as_array(h.kit.inv_phi_obs[f], shape=(2 * V,)), golden[f"force_Y{f}"])


@pytest.mark.parametrize("tag,eps,steps,ntraj", [("hmc_e0.1", 0.1, 2, 20), ("hmc_e0.2", 0.2, 2, 20),
                                                 ("hmc_e0.05s4", 0.05, 4, 5)])
def test_hmc_and_measurements(golden, P, tag, eps, steps, ntraj):
    """20 LFMN2VF trajectories (HIGH 1e-9, MEDIUM 1e-3, as macro.h:135-136),
    each followed by ComputeCondensateDirac and CondensateSigmaPi."""
    lat = P.Lattice(8)
    h = P.HMC(lat, 1231)
    assert np.array_equal(h.sigma, golden[tag + "_sigma0"])
    dh, acc, cond, cs = [], [], [], []
    for t in range(ntraj):
        tr = h.trajectory(eps, steps, 1e-9, 1e-3)
        assert tr.ok
        dh.append(tr.dH)
        acc.append(tr.accepted)
        cond.append(h.condensate(1e-9))
        cs.append(h.condensate_sigma())
        if t == 0:
            assert np.array_equal(h.sigma, golden[tag + "_sigma1"])
            assert np.array_equal(h.mom, golden[tag + "_mom1"])
    # Delta-H is printed by the reference with %.16f (moleculardynamics.c:1426)
    assert ["%.16f" % v for v in dh] == ["%.16f" % v for v in golden[tag + "_dh"]]
    assert np.array_equal(acc, golden[tag + "_acc"])
    assert np.array_equal(np.array(cond).ravel(), golden[tag + "_cond"])
    assert np.array_equal(np.array(cs).ravel(), golden[tag + "_csig"])
    assert np.array_equal(h.sigma, golden[tag + "_sigmaN"])
    assert h.kit.st.iterations == golden[tag + "_iters"][-1]


def test_link_force_golden(golden, P):
    """AddForceFromFermionSectorRespectToLink as the reference computes it
    (moleculardynamics.c:164-193: live-operator solves, force from the links)."""
    lat = P.Lattice(8)
    U = golden["link_U"].view(np.complex128)
    sm = golden["link_sigmamass"]
    h = P.LinkHMC(lat, 1, sm, U, beta=1.0, pure_gauge=False, linked=False)
    for f in range(2):
        np.ctypeslib.as_array(h.kit.phi[f], shape=(2 * lat.V,))[:] = golden[f"link_phi{f}"]
    F = np.zeros(3 * lat.V)
    ok = P.lib().orc_link_force(P.C.byref(lat.g), sm, h.U, 1.0, 0, 0, 1e-9, lat.V + 1, P.C.byref(h.kit), F)
    assert ok
    assert np.array_equal(F, golden["link_force_F"])
    for f in range(2):
        assert np.array_equal(np.ctypeslib.as_array(h.kit.inv_phi[f], shape=(2 * lat.V,)), golden[f"link_force_X{f}"])
        assert np.array_equal(np.ctypeslib.as_array(h.kit.inv_phi_obs[f], shape=(2 * lat.V,)),
                              golden[f"link_force_Y{f}"])
    assert h.kit.st.iterations == golden["link_force_stat"][1]
    assert h.kit.st.cg_call == golden["link_force_stat"][2]


def test_leapfrog2link_golden(golden, P):
    """LeapFrog2Link (moleculardynamics.c:419-542) as shipped (PURE_GAUGE off),
    3 trajectories of 10 steps, eps 0.02, seed 4444."""
    lat = P.Lattice(8)
    U = golden["link_U"].view(np.complex128)
    sm = golden["link_sigmamass"]
    h = P.LinkHMC(lat, 4444, sm, U, beta=1.0, pure_gauge=False, linked=False)
    dh, acc = [], []
    for t in range(3):
        tr = h.trajectory(0.02, 10, 1e-9)
        assert tr.ok
        dh.append(tr.dH)
        acc.append(tr.accepted)
        if t == 0:
            assert np.array_equal(h.mom, golden["link_traj_mom1"])
            assert np.array_equal(h.force, golden["link_traj_force1"])
    assert ["%.16f" % v for v in dh] == ["%.16f" % v for v in golden["link_traj_dh"]]
    assert np.array_equal(acc, golden["link_traj_acc"])
    assert np.array_equal(h.U, golden["link_traj_U"])
    assert h.kit.st.iterations == golden["link_traj_stat"][0]
    assert h.kit.st.cg_call == golden["link_traj_stat"][1]


def test_terma_checkpoint_format_golden(golden, P):
    """The terma checkpoint the reference wrote (simulation.c:129-185), parsed
    with the byte layout the drop-in writes: the links of seed 555 in lex
    order, the sigma of seed 4343 in f-order, the check value
    sum_x sum_d (carg U_d(x) + sm[x]/3) recomputed here bit for bit from the
    oracle's sigma mass, then the counters, the int story fields in 8-byte
    slots (each carrying the next field, the last one the zeroed tail padding
    of struct Info).  The second write made the first file the copy."""
    import math
    import struct
    lat = P.Lattice(8)
    V = lat.V
    b = golden["terma_first"].tobytes()
    assert len(b) == 48 * V + 8 * V + 8 + 68
    U = np.frombuffer(b, np.complex128, 3 * V, 0)
    assert np.array_equal(U.view(np.float64), golden["link_U"])
    sigma = np.frombuffer(b, np.float64, V, 48 * V)
    assert np.array_equal(sigma, P.Random(4343).sigma(V, P.sigma_variance(0.85)))
    sm = lat.sigma_mass(sigma)
    assert np.array_equal(sm, golden["link_sigmamass"])
    total = 0.0
    for x in range(V):
        for d in range(3):
            u = U[3 * x + d]
            total += math.atan2(u.imag, u.real)
            total += sm[x] / 3
    tail = struct.unpack_from("<ddiiiiiiiidiiiii", b, 56 * V)
    assert tail[0] == total == golden["terma_test"][0]
    assert tail[1:] == (0.02, 3, 11, 5, 7, 11, 19, 13, 17, 0.05, 19, 0, 23, 29, 31)
    m = golden["terma_main"].tobytes()
    assert m[:56 * V + 8] == b[:56 * V + 8]
    assert struct.unpack_from("<ddiiiiiiiidiiiii", m, 56 * V)[1:] == \
        (0.02, 3, 11, 5, 7, 11, 19, 13, 18, 0.05, 19, 0, 23, 29, 32)
    assert golden["terma_copy"].tobytes() == b
    # what the reference's reader restored from each file
    assert np.array_equal(golden["terma_read_sigma"], sigma)
    assert np.array_equal(golden["terma_read_sigmamass"], sm)
    np.testing.assert_array_equal(golden["terma_read_info"],
                                  [0.02, 3, 5, 7, 5 / 7, 11, 13, 18, 13 / 18, 0.05, 19, 23, 29, 23 / 29, 32])
    for case in ("fallback", "missing"):
        np.testing.assert_array_equal(golden[f"terma_{case}_info"],
                                      [0.02, 3, 5, 7, 5 / 7, 11, 13, 17, 13 / 17, 0.05, 19, 23, 29, 23 / 29, 31])
        assert np.array_equal(golden[f"terma_{case}_U"], golden["link_U"])


def test_setup_by_search_matches_reference_tables(golden, P):
    """The oracle's restatement of the reference's O(V^2) table set-up (timed
    by the benchmark's CPU baseline) builds the reference's own tables at
    8^3 and the closed-form tables at non-cubic shapes."""
    lat = P.Lattice(8)
    fpp, fmm = lat.tables_by_search()
    assert np.array_equal(fpp, golden["fcor_fermionpp"])
    assert np.array_equal(fmm, golden["fcor_fermionmm"])
    for L, LT in ((4, 6), (6, 4), (6, 10)):
        lat = P.Lattice(L, LT)
        fpp, fmm = lat.tables_by_search()
        assert np.array_equal(fpp, lat.table("fpp", 3 * lat.V))
        assert np.array_equal(fmm, lat.table("fmm", 3 * lat.V))
