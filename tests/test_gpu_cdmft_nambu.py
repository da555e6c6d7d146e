"""CDMFT with the Nambu spinor: the bath fit over one-body and pairing bath parameters
(the spinor bath blocks, pyqcm's CDMFT_distance with its mixing factors,
src_qcm/CPT.cpp:519-555), the Nambu host and the plain loop, against

* the cluster's own hybridization function: Gamma(z) from the fit's spinor bath blocks
  equals the solver's Gamma(z) of the same instance (one-body, bath pairing and
  cluster-bath pairing terms), whose matrices the Sigma -> 0 tests pin;
* the normal mixing: with a vanishing pairing field the Nambu host and distance are the
  normal ones (distance x 4: twice the matrix elements over four times dim^2, and no
  spin factor), and the Nambu loop reaches the recorded fixed point of
  examples/1D_2_4b/cdmft.tsv;
* a planted bath: a host equal to -Gamma of a paired bath is fitted back to that bath.
"""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_names = itertools.count(1)
L = 6  # 2 sites + 4 bath orbitals


def model():
    """examples/1D_2_4b/model_1D_2_4b.py with on-site lattice pairing S, on-site bath
    pairing db and cluster-bath pairing dh."""
    from pyqcm_amd import lattice as LT
    T = LT.chain_tiling(2).pairing("S", (0, 0, 0))
    M = LT.LatticeModel("nL2_4b_%d" % next(_names), T, n_bath=4, lattice_ops=("t", "mu", "S"))
    M.add_cluster_operator("tb1", [(1, 3, -1.0), (2, 4, -1.0), (7, 9, -1.0), (8, 10, -1.0)])
    M.add_cluster_operator("tb2", [(1, 5, -1.0), (2, 6, -1.0), (7, 11, -1.0), (8, 12, -1.0)])
    M.add_cluster_operator("eb1", [(3, 3, 1.0), (4, 4, 1.0), (9, 9, 1.0), (10, 10, 1.0)])
    M.add_cluster_operator("eb2", [(5, 5, 1.0), (6, 6, 1.0), (11, 11, 1.0), (12, 12, 1.0)])
    M.add_cluster_operator("db", [(b, b + L, 1.0) for b in range(3, 7)], kind="anomalous")
    M.add_cluster_operator("dh", [(1, 3 + L, 1.0), (2, 4 + L, 1.0), (3, 1 + L, 1.0), (4, 2 + L, 1.0)],
                           kind="anomalous")
    return M


VARIA = ["eb1", "eb2", "tb1", "tb2", "db", "dh"]


def test_nambu_bath_blocks_give_the_solver_hybridization(gpu_lib):
    from pyqcm_amd import qcm
    M = model()
    x = np.array([0.9, -1.3, 0.55, 0.4, 0.25, 0.15])
    vals = {"t": 1.0, "mu": 1.5, "U": 4.0, "S": 0.1}
    vals.update({k + "_1": v for k, v in zip(VARIA, x)})
    inst = M.instance(vals, "R0:S0")
    assert inst.mixing == 1
    z = np.array([0.3j, 1.1j, 0.5 + 0.2j])
    eb, tcb = M.diagonal_bath(VARIA, x[None], nambu=True)
    Gam = np.array([(tcb[0] / (zz - eb[0])) @ tcb[0].T for zz in z])
    ref = qcm.hybridization_functionC(z, 0, inst.label)
    print("max |Gamma(fit blocks) - Gamma(solver)| = %.1e" % np.abs(Gam - ref).max())
    np.testing.assert_allclose(Gam, ref, atol=1e-12)


def _problems(nambu_sectors="R0:S0"):
    from pyqcm_amd import cdmft as C
    grid = C.frequency_grid("matsubara", (2, 50))
    # two cluster models: a model keeps the mixing of its first instance
    # (src_ed/model_instance_base.cpp:69-74)
    M = model()
    P_n = C.CDMFTProblem(M, VARIA[:4], nambu_sectors, grid=grid)
    P_0 = C.CDMFTProblem(model(), VARIA[:4], "R0:N6:S0", grid=grid)
    return M, P_n, P_0


def test_nambu_distance_without_pairing_is_the_normal_one(gpu_lib):
    M, P_n, P_0 = _problems()
    base = {"t": 1.0, "mu": 2.0, "U": 4.0}
    x0 = [1.0, -1.0, 0.5, 0.5]
    # a vanishing lattice pairing field turns the instance into the Nambu mixing
    Ih_n = P_n.solve(dict(base, S=1e-12), x0)
    Ih_0 = P_0.solve(base, x0)
    assert P_n.nambu and not P_0.nambu
    assert abs(Ih_n.ground_state()[0] - Ih_0.ground_state()[0]) < 1e-9
    Gh_n, Gh_0 = P_n.host(Ih_n), P_0.host(Ih_0)
    # particle block: the normal host; hole block: -G_host(-i w)^* = -conj(G_host) on the axis
    np.testing.assert_allclose(Gh_n[:, :2, :2], Gh_0, atol=1e-9)
    np.testing.assert_allclose(Gh_n[:, 2:, 2:], -np.conj(Gh_0), atol=1e-9)
    X = np.array([x0, [1.2, -0.9, 0.6, 0.4], [0.7, -1.4, 0.3, 0.65]])
    d_n, d_0 = P_n.distance(Gh_n, X), P_0.distance(Gh_0, X)
    print("Nambu distance x 4 / normal:", 4 * d_n / d_0)
    np.testing.assert_allclose(4 * d_n, d_0, rtol=1e-8)


def test_nambu_cdmft_loop_reaches_the_reference_fixed_point(gpu_lib, golden):
    """The loop in the Nambu spinor (host Nelder-Mead over the spinor bath, device
    distances) with a vanishing pairing field: the fixed point of cdmft.tsv (mu = 2)."""
    from pyqcm_amd import cdmft as C
    row = golden["chain2_4b_cdmft"]["rows"][0]
    assert row["mu"] == "2"
    M, P_n, _ = _problems()
    run = C.CDMFT(P_n, {"t": 1.0, "mu": 2.0, "U": 4.0, "S": 1e-12}, [1.0, -1.0, 0.5, 0.5])
    ref = np.array([float(row[k]) for k in ("eb1_1", "eb2_1", "tb1_1", "tb2_1")])
    print("Nambu loop", run.x, "reference", ref, "iterations", run.niter)
    assert P_n.nambu
    np.testing.assert_allclose(run.x, ref, atol=2e-4)


def test_nambu_fit_recovers_a_planted_paired_bath(gpu_lib):
    from pyqcm_amd import cdmft as C
    M = model()
    P = C.CDMFTProblem(M, VARIA, "R0:S0", grid=C.frequency_grid("matsubara", (2, 50)), nambu=True)
    xs = np.array([0.9, -1.3, 0.55, 0.4, 0.25, 0.15])
    eb, tcb = M.diagonal_bath(VARIA, xs[None], nambu=True)
    z = 1j * P.grid.wr
    Gh = -np.array([(tcb[0] / (zz - eb[0])) @ tcb[0].T for zz in z])
    assert P.distance(Gh, xs[None])[0] < 1e-28
    x0 = xs + np.array([0.05, -0.04, 0.03, -0.02, 0.03, -0.02])
    X, F, NF, OK = P.fit(Gh[None], [0], x0[None], initial_step=0.05, xatol=1e-9, fatol=1e-16)
    print("planted", xs, "fitted", X[0], "distance %.2e" % F[0], "evaluations", NF[0])
    assert OK[0]
    np.testing.assert_allclose(X[0], xs, atol=1e-5)


def test_batched_nambu_cdmft_equals_plain_loop(gpu_lib):
    """A paired problem (lattice on-site pairing field S, variational bath pairing db):
    two instances advanced in lockstep by the batched engine reach the plain loop's fixed
    points (the spinor instances are solved one by one inside the batch), and those fixed
    points carry bath pairing."""
    from pyqcm_amd import cdmft as C
    grid = C.frequency_grid("matsubara", (2, 50))
    varia = ["eb1", "eb2", "tb1", "tb2", "db"]
    P = C.CDMFTProblem(model(), varia, "R0:S0", grid=grid)
    cases = [({"t": 1.0, "mu": 2.0, "U": 4.0, "S": 0.1}, [1.0, -1.0, 0.5, 0.5, 0.1]),
             ({"t": 1.0, "mu": 1.5, "U": 3.0, "S": 0.1}, [1.2, -0.8, 0.6, 0.4, 0.1])]
    plain = [C.CDMFT(P, v, x) for v, x in cases]
    assert P.nambu
    batch = C.BatchedCDMFT(P, [v for v, _ in cases], [x for _, x in cases])
    for i, p in enumerate(plain):
        print("plain", p.x, p.niter, "batched", batch.x[i], batch.niter[i])
        assert batch.converged[i]
        # the same fixed point within the loop's tolerance (accur = 1e-4 on the parameters):
        # the batch draws its ED start vectors in another order, and a round-off-level
        # difference can end the loop one iteration earlier
        np.testing.assert_allclose(batch.x[i], p.x, atol=2e-4)
        assert abs(p.x[4]) > 1e-3


def test_spin_flip_cdmft_without_field_reaches_the_reference_fixed_point(gpu_lib, golden):
    """The loop in the spin-flip spinor (c_up, c_down): the bath fit over the spinor's
    blocks (LatticeModel.spin_flip_bath_blocks), distance factor 1. With a vanishing
    field along x the loop reaches the recorded fixed point of cdmft.tsv (mu = 2), and
    Gamma(z) from the fit's blocks equals the solver's."""
    from pyqcm_amd import cdmft as C
    from pyqcm_amd import lattice as LT
    from pyqcm_amd import qcm
    row = golden["chain2_4b_cdmft"]["rows"][0]
    T = LT.chain_tiling(2).spin_flip("hx")
    M = LT.LatticeModel("sfL2_4b_%d" % next(_names), T, n_bath=4, lattice_ops=("t", "mu", "hx"))
    M.add_cluster_operator("tb1", [(1, 3, -1.0), (2, 4, -1.0), (7, 9, -1.0), (8, 10, -1.0)])
    M.add_cluster_operator("tb2", [(1, 5, -1.0), (2, 6, -1.0), (7, 11, -1.0), (8, 12, -1.0)])
    M.add_cluster_operator("eb1", [(3, 3, 1.0), (4, 4, 1.0), (9, 9, 1.0), (10, 10, 1.0)])
    M.add_cluster_operator("eb2", [(5, 5, 1.0), (6, 6, 1.0), (11, 11, 1.0), (12, 12, 1.0)])
    P = C.CDMFTProblem(M, VARIA[:4], "R0:N6", grid=C.frequency_grid("matsubara", (2, 50)))
    vals = {"t": 1.0, "mu": 2.0, "U": 4.0, "hx": 1e-12}
    x0 = np.array([1.0, -1.0, 0.5, 0.5])
    inst = P.solve(vals, x0)
    assert P.nambu and P.spinor_mixing == 2
    z = np.array([0.3j, 1.1j])
    eb, tcb = M.diagonal_bath(VARIA[:4], x0[None], mixing=2)
    Gam = np.array([(tcb[0] / (zz - eb[0])) @ tcb[0].T for zz in z])
    np.testing.assert_allclose(Gam, qcm.hybridization_functionC(z, 0, inst.label), atol=1e-12)
    run = C.CDMFT(P, vals, x0)
    ref = np.array([float(row[k]) for k in ("eb1_1", "eb2_1", "tb1_1", "tb2_1")])
    print("spin-flip loop", run.x, "reference", ref, "iterations", run.niter)
    np.testing.assert_allclose(run.x, ref, atol=2e-4)


def test_up_down_cdmft_reaches_the_reference_fixed_point(gpu_lib, golden):
    """The up/down mixing (a spin-dependent bath term, here a vanishing bath field along z):
    both spin blocks' hosts side by side, the bath fit over both spins' blocks, the
    distance summing both (src_qcm/CPT.cpp:519-552, no spin factor). Its distance equals
    the normal one for spin-symmetric baths, and the loop reaches the recorded fixed point
    of cdmft.tsv (mu = 2)."""
    from pyqcm_amd import cdmft as C
    from pyqcm_amd import lattice as LT
    row = golden["chain2_4b_cdmft"]["rows"][0]

    def mk():
        M = LT.LatticeModel("udL2_4b_%d" % next(_names), LT.chain_tiling(2), n_bath=4, lattice_ops=("t", "mu"))
        M.add_cluster_operator("tb1", [(1, 3, -1.0), (2, 4, -1.0), (7, 9, -1.0), (8, 10, -1.0)])
        M.add_cluster_operator("tb2", [(1, 5, -1.0), (2, 6, -1.0), (7, 11, -1.0), (8, 12, -1.0)])
        M.add_cluster_operator("eb1", [(3, 3, 1.0), (4, 4, 1.0), (9, 9, 1.0), (10, 10, 1.0)])
        M.add_cluster_operator("eb2", [(5, 5, 1.0), (6, 6, 1.0), (11, 11, 1.0), (12, 12, 1.0)])
        M.add_cluster_operator("hzb", [(b, b, 1.0) for b in range(3, 7)] + [(b + L, b + L, -1.0) for b in range(3, 7)])
        return M
    grid = C.frequency_grid("matsubara", (2, 50))
    P4 = C.CDMFTProblem(mk(), VARIA[:4], "R0:N6:S0", grid=grid)
    P0 = C.CDMFTProblem(mk(), VARIA[:4], "R0:N6:S0", grid=grid)
    base = {"t": 1.0, "mu": 2.0, "U": 4.0}
    x0 = np.array([1.0, -1.0, 0.5, 0.5])
    i4 = P4.solve(dict(base, hzb_1=1e-12), x0)
    i0 = P0.solve(base, x0)
    assert P4.spinor_mixing == 4 and P0.spinor_mixing == 0
    X = np.array([x0, [1.2, -0.9, 0.6, 0.4]])
    d4, d0 = P4.distance(P4.host(i4), X), P0.distance(P0.host(i0), X)
    print("up/down distance / normal:", d4 / d0)
    np.testing.assert_allclose(d4, d0, rtol=1e-8)
    run = C.CDMFT(P4, dict(base, hzb_1=1e-12), x0)
    ref = np.array([float(row[k]) for k in ("eb1_1", "eb2_1", "tb1_1", "tb2_1")])
    print("up/down loop", run.x, "reference", ref, "iterations", run.niter)
    np.testing.assert_allclose(run.x, ref, atol=2e-4)
