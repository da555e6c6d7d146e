"""Lattice layer on the GPU: CPT against the U=0 closed form, the VCA functional against
the reference's recorded values, the CDMFT host against a dense restatement, and the
plain and batched CDMFT loops against the reference's recorded fixed point."""
import itertools

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

_names = itertools.count(1)


def _name(base):
    return "%s_%d" % (base, next(_names))


def chain_model(n, n_bath=0):
    from pyqcm_amd import lattice as LT
    return LT.LatticeModel(_name("chain%d" % n), LT.chain_tiling(n), n_bath=n_bath)


def chain2_4b_model():
    """examples/1D_2_4b/model_1D_2_4b.py (normal-state operators only)."""
    from pyqcm_amd import lattice as LT
    M = LT.LatticeModel(_name("L2_4b"), LT.chain_tiling(2), n_bath=4)
    M.add_cluster_operator("tb1", [(1, 3, -1.0), (2, 4, -1.0), (7, 9, -1.0), (8, 10, -1.0)])
    M.add_cluster_operator("tb2", [(1, 5, -1.0), (2, 6, -1.0), (7, 11, -1.0), (8, 12, -1.0)])
    M.add_cluster_operator("eb1", [(3, 3, 1.0), (4, 4, 1.0), (9, 9, 1.0), (10, 10, 1.0)])
    M.add_cluster_operator("eb2", [(5, 5, 1.0), (6, 6, 1.0), (11, 11, 1.0), (12, 12, 1.0)])
    return M


# ------------------------------------------------------------------------- CPT

@pytest.mark.parametrize("shape", ["chain4", "square2x2"])
def test_cpt_exact_at_U0(gpu_lib, shape):
    """At U=0 the cluster G is exact and CPT is exact: G_cpt(k, z) = (z - t(k))^-1."""
    from pyqcm_amd import lattice as LT
    if shape == "chain4":
        M = chain_model(4)
        vals = {"t": 1.0, "tp": -0.25, "mu": 0.3, "U": 0.0}
        sectors = "R0:N4:S0"
        k = np.array([[0.1, 0, 0], [0.37, 0, 0], [0.5, 0, 0]])
    else:
        M = LT.LatticeModel(_name("sq"), LT.square_tiling(2, 2))
        vals = {"t": 1.0, "tp": -0.3, "mu": 0.2, "U": 0.0}
        sectors = "R0:N4:S0"
        k = np.array([[0.1, 0.2, 0], [0.37, 0.8, 0], [0.5, 0.5, 0]])
    inst = M.instance(vals, sectors)
    z = np.array([0.3j, 1.0 + 0.1j, -0.7 + 0.05j, 2.5j])
    G = inst.cpt_green(k, z, "full")
    tk = M.tiling.tk(vals, k)
    n = M.n
    for a in range(len(z)):
        for b in range(len(k)):
            ref = np.linalg.inv(z[a] * np.eye(n) - tk[b])
            np.testing.assert_allclose(G[a, b], ref, atol=1e-11)
    tr = inst.cpt_green(k, z, "trace")
    np.testing.assert_allclose(tr, np.trace(G, axis1=2, axis2=3), atol=1e-11)
    av = inst.cpt_green(k, z, "average")
    np.testing.assert_allclose(av, G.mean(1), atol=1e-11)


def test_cpt_interacting_matches_dense_restatement(gpu_lib):
    """U=4 chain: G_cpt = G_c (1 - V G_c)^-1 with V = t(k) - t_c, against numpy."""
    M = chain_model(4)
    vals = {"t": 1.0, "mu": 2.0, "U": 4.0}
    inst = M.instance(vals, "R0:N4:S0")
    z = np.array([0.2j, 1.5j, 0.4 + 0.1j])
    k = np.array([[0.13, 0, 0], [0.71, 0, 0]])
    G = inst.cpt_green(k, z)
    Gc = inst.green(z)
    tk = M.tiling.tk(vals, k)
    tc = M.tiling.cluster_matrix(vals)
    for a in range(len(z)):
        for b in range(len(k)):
            np.testing.assert_allclose(G[a, b], O.cpt_green(Gc[a], tk[b] - tc), atol=1e-11)


# ------------------------------------------------------------------------- VCA

def test_potthoff_functional_matches_reference_sef(gpu_lib, golden):
    """examples/1D_4/sef.tsv: Omega of the 4-site chain (U=4, mu=2) at cdw_1 = 0 and
    +-0.01. The reference integrates adaptively; the grid here is fine enough that the
    quadrature error is below the 8 printed digits."""
    rows = [r for r in golden["chain4_sef"]["rows"] if r["mu"] == "2"]
    M = chain_model(4)
    got = {}
    for r in rows:
        cdw1 = float(r["cdw_1"])
        vals = {"t": 1.0, "mu": 2.0, "U": 4.0, "cdw_1": cdw1}
        inst = M.instance(vals, "R0:N4:S0")
        E0, sec = inst.ground_state()
        assert abs(E0 - float(r["E0_1"])) < 2e-7
        om = inst.potthoff_functional(grid=(1, 10, 30, 30, 30), kside=128)
        got[cdw1] = om
        assert abs(om - float(r["omega"])) < 2e-7, (cdw1, om, r["omega"])
    # the Weiss-field response (second order in cdw_1)
    d = got[max(got)] - got[min(got, key=abs)]
    assert abs(d - (-2.5526019 + 2.5525862)) < 1e-7


# ------------------------------------------------------------------------ CDMFT

def test_cdmft_host_matches_dense_restatement(gpu_lib):
    """G_host = [avg_k G_cpt]^-1 + Sigma_c + t_c - i w, with the k-sum done in numpy."""
    from pyqcm_amd import qcm
    M = chain2_4b_model()
    vals = {"t": 1.0, "mu": 2.0, "U": 4.0, "eb1_1": 1.0, "eb2_1": -1.0, "tb1_1": 0.5, "tb2_1": 0.5}
    inst = M.instance(vals, "R0:N6:S0")
    wr = np.array([0.1, 0.5, 2.0])
    side = 8
    Gh = inst.cdmft_host(wr, side)
    z = 1j * wr
    Gc = qcm.Green_function(z, 0, inst.label)
    Gam = qcm.hybridization_functionC(z, 0, inst.label)
    S = qcm.self_energyC(z, 0, inst.label)
    tc = M.tiling.cluster_matrix({"t": 1.0, "mu": 2.0})
    kg = M.tiling.kgrid(side, 0.5)
    tk = M.tiling.tk({"t": 1.0, "mu": 2.0}, kg)
    for a in range(len(wr)):
        acc = np.zeros((2, 2), complex)
        for b in range(len(kg)):
            acc += O.cpt_green(Gc[a], tk[b] - tc - Gam[a])
        ref = np.linalg.inv(acc / len(kg)) + S[a] + tc - z[a] * np.eye(2)
        np.testing.assert_allclose(Gh[a], ref, atol=1e-10)


def test_cdmft_distance_of_problem_matches_restatement(gpu_lib):
    from pyqcm_amd import cdmft as C
    M = chain2_4b_model()
    P = C.CDMFTProblem(M, ["eb1_1", "eb2_1", "tb1_1", "tb2_1"], "R0:N6:S0",
                       grid=C.frequency_grid("matsubara", (2, 50)))
    inst = P.solve({"t": 1.0, "mu": 2.0, "U": 4.0}, [1.0, -1.0, 0.5, 0.5])
    Gh = P.host(inst)
    X = np.array([[1.0, -1.0, 0.5, 0.5], [1.2, -0.9, 0.6, 0.4]])
    d = P.distance(Gh, X)
    for i, x in enumerate(X):
        eb = np.array([x[0], x[0], x[1], x[1]])
        tcb = np.zeros((2, 4))
        tcb[0, 0] = tcb[1, 1] = -x[2]
        tcb[0, 2] = tcb[1, 3] = -x[3]
        ref = O.cdmft_distance(P.grid.wr, P.grid.cdmft_weight, Gh, eb, tcb)
        assert abs(d[i] - ref) <= 1e-12 * max(1.0, abs(ref))


def test_cdmft_chain2_4b_reaches_reference_fixed_point(gpu_lib, golden):
    """examples/1D_2_4b/cdmft.tsv (mu=2): converged bath and E0 of the reference loop,
    from the starting values of examples/1D_2_4b/QCM_params.def."""
    from pyqcm_amd import cdmft as C
    row = golden["chain2_4b_cdmft"]["rows"][0]
    assert row["dist_function"] == "sharp_wc_2.0_b_50" and row["mu"] == "2"
    M = chain2_4b_model()
    P = C.CDMFTProblem(M, ["eb1_1", "eb2_1", "tb1_1", "tb2_1"], "R0:N6:S0",
                       grid=C.frequency_grid("matsubara", (2, 50)))
    vals = {"t": 1.0, "mu": 2.0, "U": 4.0}
    run = C.CDMFT(P, vals, [1.0, -1.0, 0.5, 0.5])
    ref = np.array([float(row[k]) for k in ("eb1_1", "eb2_1", "tb1_1", "tb2_1")])
    print("converged", run.x, "reference", ref, "iterations", run.niter, "distance", run.dist)
    np.testing.assert_allclose(run.x, ref, atol=2e-4)
    # run.dist (6.0e-5 here) is the distance of the last fit, against a host that the
    # final step of the loop moved; the reference recorded 1.20e-4 at its own stopping
    # point. The distance function itself is compared at the recorded baths in
    # test_cdmft_distance_at_reference_baths.
    inst = P.solve(vals, ref)
    assert abs(inst.ground_state()[0] - float(row["E0_1"])) < 2e-7


def test_cdmft_distance_at_reference_baths(gpu_lib, golden):
    """The distance function (weights, frequency grid, the hybridization of the bath, the
    normalisation) against the values recorded along the mu scan of
    examples/1D_2_4b/cdmft.tsv: at each converged bath x, d(x; host(x)) over the recorded
    distance is one constant for all rows. The file was written by pyqcm v1.0.7 (its
    version column), whose recorded distance is twice the current definition
    (src_qcm/CPT.cpp:519-555: factor 2 at mixing 0, times the frequency step, over dim^2),
    which this package follows: the ratio is 1/2 to the three digits the file keeps."""
    from pyqcm_amd import cdmft as C
    rows = golden["chain2_4b_cdmft"]["rows"]
    M = chain2_4b_model()
    P = C.CDMFTProblem(M, ["eb1_1", "eb2_1", "tb1_1", "tb2_1"], "R0:N6:S0",
                       grid=C.frequency_grid("matsubara", (2, 50)))
    ratio = []
    for row in rows[::10]:
        vals = {"t": float(row["t"]), "mu": float(row["mu"]), "U": float(row["U"])}
        x = np.array([float(row[k]) for k in ("eb1_1", "eb2_1", "tb1_1", "tb2_1")])
        d = float(P.distance(P.host(P.solve(vals, x)), x[None, :])[0])
        ratio.append(d / float(row["distance"]))
        print("mu", row["mu"], "recorded", row["distance"], "here %.4e" % d, "ratio %.4f" % ratio[-1])
    assert all(0.497 < r < 0.507 for r in ratio), ratio


def test_batched_cdmft_equals_plain_loop(gpu_lib):
    """Two instances advanced in lockstep reach the plain loop's fixed points. The ED start
    vectors come from one process-wide random stream (as in the reference), so solving the
    instances in another order changes the Lanczos vectors at round-off level; the
    converged baths agree to well within the loop's tolerance."""
    from pyqcm_amd import cdmft as C
    M = chain2_4b_model()
    P = C.CDMFTProblem(M, ["eb1_1", "eb2_1", "tb1_1", "tb2_1"], "R0:N6:S0",
                       grid=C.frequency_grid("matsubara", (2, 50)))
    cases = [({"t": 1.0, "mu": 2.0, "U": 4.0}, [1.0, -1.0, 0.5, 0.5]),
             ({"t": 1.0, "mu": 1.5, "U": 3.0}, [1.2, -0.8, 0.6, 0.4])]
    plain = [C.CDMFT(P, v, x) for v, x in cases]
    batch = C.BatchedCDMFT(P, [v for v, _ in cases], [x for _, x in cases])
    for i, p in enumerate(plain):
        assert batch.converged[i]
        np.testing.assert_allclose(batch.x[i], p.x, atol=1e-5)
        assert abs(int(batch.niter[i]) - p.niter) <= 1


@pytest.mark.parametrize("model", ["chain2_4b", "plaquette_8b"])
def test_device_nelder_mead_matches_host_coroutine(gpu_lib, model):
    """qed_cdmft_fit_batch (one wavefront per fit) follows the host form of scipy's
    adaptive Nelder-Mead evaluation for evaluation: same minimizer, value and number of
    evaluations, bit for bit, with several hosts in one launch."""
    from pyqcm_amd import cdmft as C
    rng = np.random.default_rng(11)
    if model == "chain2_4b":
        M = chain2_4b_model()
        P = C.CDMFTProblem(M, ["eb1_1", "eb2_1", "tb1_1", "tb2_1"], "R0:N6:S0",
                           grid=C.frequency_grid("matsubara", (2, 50)))
        x0 = np.array([1.0, -1.0, 0.5, 0.5])
        cases = [({"t": 1.0, "mu": 2.0, "U": 4.0}, x0), ({"t": 1.0, "mu": 1.4, "U": 3.0}, x0 * 1.1)]
    else:
        M = C.plaquette_8b_model(_name("plaq"))
        P = C.CDMFTProblem(M, C.PLAQUETTE_VARIA, "R0:N12:S0")
        x0 = np.concatenate([rng.uniform(-3, 3, 8), rng.uniform(0.1, 1, 8)])
        cases = [({"t": 1.0, "tp": -0.3, "U": 8.0, "mu": 4.0}, x0),
                 ({"t": 1.0, "tp": 0.0, "U": 4.0, "mu": 2.0}, x0)]
    hosts = np.array([P.host(P.solve(v, x)) for v, x in cases])
    X0 = np.array([x + rng.normal(0, 0.05, len(x)) for x in [cases[0][1], cases[1][1]] * 2])
    host_of = np.array([0, 1, 0, 1])
    xd, fd, nd, okd = P.fit(hosts, host_of, X0, where="device")
    xh, fh, nh, okh = P.fit(hosts, host_of, X0, where="host")
    assert okd.all() and okh.all()
    np.testing.assert_array_equal(nd, nh)
    np.testing.assert_array_equal(xd, xh)
    np.testing.assert_array_equal(fd, fh)
