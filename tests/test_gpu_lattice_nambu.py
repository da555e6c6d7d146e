"""Lattice functions with pairing terms (the Nambu mixing): CPT and the Potthoff functional
in the spinor (c_up(k), c+_down(-k)) for a chain with on-site and nearest-neighbour
singlet pairing (pyqcm anomalous_operator, src_qcm/lattice_model.cpp:308-341,1101-1133;
the Nambu GF matrix elements, src_qcm/lattice_model.cpp:595-690; the functional,
src_qcm/lattice_model_instance.cpp:503-584), against

* the cluster's own spinor matrices: the lattice's intra-cluster Nambu matrix equals t_c
  of the cluster instance (V0 = 0), whose convention the dense Jordan-Wigner tests pin;
* a 4-site ring solved by dense Jordan-Wigner ED (2 cells of the 2-site cluster, pairing
  across the cells included): at U -> 0, CPT is exact, so the ring's Nambu G between
  sites (i, R) and (j, R') is (1/2) sum_k exp(2 pi i k (R - R')) G_cpt(k)_ij over
  k = 0, 1/2 -- this fixes the inter-cluster pairing and hole blocks and their phases;
* the BdG ground-state energy at U -> 0: the functional reduces to
  (1/n) <sum_{E_m(k) < 0} E_m(k)> + Tr T_0 / n on the same k-grid (and, with the
  pairing off, to the normal mixing's (2/n) <sum_{e(k) < 0} e(k)>).
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

_names = itertools.count(1)


def nambu_chain(n):
    from pyqcm_amd import lattice as LT
    T = LT.chain_tiling(n).pairing("D", (1, 0, 0)).pairing("S", (0, 0, 0))
    return LT.LatticeModel("nchain%d_%d" % (n, next(_names)), T, lattice_ops=("t", "tp", "mu", "D", "S"))


ZS = np.array([0.4j, 1.3j, 0.5 + 0.2j, -0.8 + 0.1j])


def test_lattice_intra_cluster_nambu_matrix_is_the_cluster_one(gpu_lib):
    M = nambu_chain(4)
    vals = {"t": 1.0, "tp": -0.3, "mu": 0.7, "D": 0.25, "S": 0.15, "U": 2.0}
    inst = M.instance(vals, "R0:S0")
    assert inst.mixing == 1
    tc = inst.v0() + M.tiling.cluster_matrix(vals, nambu=True)
    print("max |V0| = %.1e" % np.abs(inst.v0()).max())
    np.testing.assert_allclose(inst.v0(), 0.0, atol=1e-14)
    # the blocks: particle t_c, hole -t_c^T, pairing -D on the bonds and -S on the sites
    n = 4
    assert tc[0, 1] == -1.0 and tc[n, n + 1] == 1.0 and tc[0, n] == -0.15 and tc[0, n + 1] == -0.25
    np.testing.assert_allclose(tc, tc.T, atol=0)


def _ring_model(vals):
    """The 4-site ring (NN hopping, chemical potential, NN singlet and on-site pairing,
    vanishing U) as Jordan-Wigner operators, sites 0..3, spin down at i + 4."""
    L = 4
    m = O.ClusterModel(L)
    bonds = [(i, (i + 1) % L) for i in range(L)]
    hops = [(min(i, j), max(i, j)) for i, j in bonds]  # one-body elements r < c stand for both directions
    m.add("t", "one-body", [(i, j, -1.0) for i, j in hops] + [(i + L, j + L, -1.0) for i, j in hops])
    m.add("mu", "one-body", [(i, i, -1.0) for i in range(L)] + [(i + L, i + L, -1.0) for i in range(L)])
    m.add("D", "anomalous", [(i, j + L, 1.0) for i, j in bonds] + [(j, i + L, 1.0) for i, j in bonds])
    m.add("S", "anomalous", [(i, i + L, 1.0) for i in range(L)])
    m.add("U", "interaction", [(i, i + L, 1.0) for i in range(L)])
    return m


def _s_blocks(L, S):
    return [(nu, nu - S) for nu in range(L + 1) if 0 <= nu - S <= L]


def test_nambu_cpt_is_exact_at_U0_on_a_ring(gpu_lib):
    vals = {"t": 1.0, "mu": 0.6, "D": 0.3, "S": 0.2, "U": 1e-9}
    M = nambu_chain(2)
    inst = M.instance(vals, "R0:S0")
    k = np.array([[0.0, 0, 0], [0.5, 0, 0]])
    G = inst.cpt_green(k, ZS, "full")  # (nz, 2, 4, 4): particle 0, 1, hole 0, 1
    # U -> 0: the device CPT is the inverse of z - h(k) from the tiling's Nambu terms
    hk = M.tiling.tk(vals, k, nambu=True)
    for a, z in enumerate(ZS):
        for b in range(len(k)):
            np.testing.assert_allclose(G[a, b], np.linalg.inv(z * np.eye(4) - hk[b]), atol=1e-8)
    # the ring's Nambu G from dense ED, spinor (c_i up, c+_i down), i = 0..3
    m = _ring_model(vals)
    L = 4
    spinor = [(i, False) for i in range(L)] + [(i + L, True) for i in range(L)]
    E0, Gd = O.dense_spinor_gf(m, vals, O.block_states(L, _s_blocks(L, 0)), spinor,
                               O.block_states(L, _s_blocks(L, -1)), O.block_states(L, _s_blocks(L, 1)), ZS)
    e0 = np.linalg.eigvalsh(O.dense_hamiltonian(m, vals, 0, 0, states=O.block_states(L, _s_blocks(L, 0)))[0])
    assert e0[1] - e0[0] > 1e-6
    # ring site i + 2R <- cluster site i in cell R; spinor index c (particle) or c + 2 (hole)
    Gr = np.zeros_like(Gd)
    for s1, s2 in itertools.product(range(L), repeat=2):
        i, R = s1 % 2, s1 // 2
        j, Rp = s2 % 2, s2 // 2
        for h1, h2 in itertools.product((0, 1), repeat=2):
            ph = np.exp(2j * np.pi * k[:, 0] * (R - Rp))
            val = 0.5 * np.einsum("b,zb->z", ph, G[:, :, i + 2 * h1, j + 2 * h2])
            Gr[:, s1 + L * h1, s2 + L * h2] = val
    print("max |G_ring(CPT) - G_ring(dense)| = %.2e" % np.abs(Gr - Gd).max())
    np.testing.assert_allclose(Gr, Gd, atol=5e-8)  # U = 1e-9 in both: O(U) apart


def _bdg_omega(tiling, vals, kside, nambu):
    """T = 0 grand potential per site of the noninteracting lattice on the k-grid
    i/kside: Nambu (1/n) <sum_{E<0} E + Tr t_down(k)>, normal (2/n) <sum_{e<0} e>."""
    k = tiling.kgrid(kside, shift=0.0)
    h = tiling.tk(vals, k, nambu=nambu)
    e = np.linalg.eigvalsh(h)
    n = tiling.n
    neg = np.where(e < 0, e, 0.0).sum(1).mean()
    if nambu:
        tdn = np.trace(tiling.tk(vals, k), axis1=1, axis2=2).real.mean()
        return (neg + tdn) / n
    return 2.0 * neg / n


@pytest.mark.parametrize("pairing", [True, False])
def test_nambu_potthoff_functional_is_the_bdg_energy_at_U0(gpu_lib, pairing):
    vals = {"t": 1.0, "tp": -0.2, "mu": 0.5, "U": 1e-9}
    if pairing:
        vals.update({"D": 0.3, "S": 0.1})
    M = nambu_chain(4)
    # the free cluster has three negative levels: its ground state holds N = 6
    inst = M.instance(vals, "R0:S0" if pairing else "R0:N6:S0")
    assert inst.mixing == (1 if pairing else 0)
    om = inst.potthoff_functional(grid=(1, 10, 30, 30, 30), kside=128)
    ref = _bdg_omega(M.tiling, vals, 128, nambu=pairing)
    print("pairing", pairing, "Omega %.10f, BdG energy per site %.10f" % (om, ref))
    assert abs(om - ref) < 2e-6
    if pairing:
        # the Nambu form of the pairing-free model is the normal one
        v0 = {k: v for k, v in vals.items() if k not in ("D", "S")}
        assert abs(_bdg_omega(M.tiling, v0, 128, True) - _bdg_omega(M.tiling, v0, 128, False)) < 1e-12


@pytest.mark.parametrize("bath_pairing", [False, True])
def test_nambu_potthoff_functional_with_a_bath_at_U0(gpu_lib, bath_pairing):
    """2-site cluster + 2 bath orbitals, lattice pairing D and S: at U -> 0 the functional
    is the lattice's BdG energy whatever the bath, so the bath's own terms (SEF_bath and
    log det G_c - log det G_eff) must cancel."""
    from pyqcm_amd import lattice as LT
    T = LT.chain_tiling(2).pairing("D", (1, 0, 0)).pairing("S", (0, 0, 0))
    M = LT.LatticeModel("nbath_%d" % next(_names), T, n_bath=2, lattice_ops=("t", "mu", "D", "S"))
    L = 4
    M.add_cluster_operator("tb", [(1, 3, 1.0), (2, 4, 1.0), (1 + L, 3 + L, 1.0), (2 + L, 4 + L, 1.0)])
    M.add_cluster_operator("eb", [(3, 3, 1.0), (4, 4, -1.0), (3 + L, 3 + L, 1.0), (4 + L, 4 + L, -1.0)])
    vals = {"t": 1.0, "mu": 0.4, "D": 0.3, "S": 0.1, "U": 1e-9, "tb_1": 0.5, "eb_1": 0.7}
    if bath_pairing:
        M.add_cluster_operator("db", [(3, 3 + L, 1.0), (4, 4 + L, 1.0)], kind="anomalous")
        vals["db_1"] = 0.35
    inst = M.instance(vals, "R0:S0")
    assert inst.mixing == 1
    om = inst.potthoff_functional(grid=(1, 10, 30, 30, 30), kside=128)
    ref = _bdg_omega(M.tiling, {k: v for k, v in vals.items() if not k.endswith("_1")}, 128, nambu=True)
    print("bath pairing", bath_pairing, "Omega %.10f, BdG energy per site %.10f" % (om, ref))
    assert abs(om - ref) < 2e-6


def test_nambu_cpt_interacting_matches_dense_restatement(gpu_lib):
    """U = 4 with a Nambu bath: G_cpt = (G_c^-1 - V(k, z))^-1 with V = h(k) - t_c - Gamma(z)
    (the bath's hybridization replaced by the lattice hopping, src_qcm/CPT.cpp:57-61),
    from the cluster's own G_c and Gamma and the tiling's Nambu h(k), against numpy."""
    from pyqcm_amd import lattice as LT
    from pyqcm_amd import qcm
    T = LT.chain_tiling(2).pairing("D", (1, 0, 0)).pairing("S", (0, 0, 0))
    M = LT.LatticeModel("nint_%d" % next(_names), T, n_bath=2, lattice_ops=("t", "mu", "D", "S"))
    L = 4
    M.add_cluster_operator("tb", [(1, 3, 1.0), (2, 4, 1.0), (1 + L, 3 + L, 1.0), (2 + L, 4 + L, 1.0)])
    M.add_cluster_operator("eb", [(3, 3, 1.0), (4, 4, -1.0), (3 + L, 3 + L, 1.0), (4 + L, 4 + L, -1.0)])
    M.add_cluster_operator("db", [(3, 3 + L, 1.0), (4, 4 + L, 1.0)], kind="anomalous")
    vals = {"t": 1.0, "mu": 1.2, "D": 0.2, "S": 0.05, "U": 4.0, "tb_1": 0.5, "eb_1": 1.1, "db_1": 0.15}
    inst = M.instance(vals, "R0:S0")
    k = np.array([[0.13, 0, 0], [0.5, 0, 0], [0.81, 0, 0]])
    G = inst.cpt_green(k, ZS)
    Gc = inst.green(ZS)
    Gam = qcm.hybridization_functionC(ZS, 0, inst.label)
    tc = qcm.hopping_matrix(0, 0, inst.label).real
    hk = M.tiling.tk({k_: v for k_, v in vals.items() if not k_.endswith("_1")}, k, nambu=True)
    for a in range(len(ZS)):
        for b in range(len(k)):
            ref = np.linalg.inv(np.linalg.inv(Gc[a]) - (hk[b] - tc - Gam[a]))
            np.testing.assert_allclose(G[a, b], ref, atol=1e-10)


def test_spin_flip_lattice_cpt_and_potthoff_at_U0(gpu_lib):
    """A field along x on every site (pyqcm hopping_operator with tau = 0, sigma = 1): the
    instance has the spin-flip mixing, and CPT and the functional work in the spinor
    (c_up(k), c_down(k)). The lattice's intra-cluster spinor matrix is the cluster's t_c;
    at U -> 0 CPT is (z - h(k))^-1 and the functional is the lattice ground-state energy
    per site, (1/n) <sum_{E(k) < 0} E(k)> on the same k-grid."""
    from pyqcm_amd import lattice as LT
    T = LT.chain_tiling(4).spin_flip("hx")
    M = LT.LatticeModel("sfchain_%d" % next(_names), T, lattice_ops=("t", "mu", "hx"))
    vals = {"t": 1.0, "mu": 0.0, "hx": 0.3, "U": 1e-9}
    # the cluster's spinor levels e +- 0.3 have four below zero: N = 4
    inst = M.instance(vals, "R0:N4")
    assert inst.mixing == 2
    np.testing.assert_allclose(inst.v0(), 0.0, atol=1e-14)
    k = np.array([[0.1, 0, 0], [0.37, 0, 0], [0.5, 0, 0]])
    G = inst.cpt_green(k, ZS, "full")
    hk = M.tiling.tk(vals, k, mixing=2)
    for a, z in enumerate(ZS):
        for b in range(len(k)):
            np.testing.assert_allclose(G[a, b], np.linalg.inv(z * np.eye(8) - hk[b]), atol=1e-8)
    om = inst.potthoff_functional(grid=(1, 10, 30, 30, 30), kside=128)
    kg = M.tiling.kgrid(128, shift=0.0)
    e = np.linalg.eigvalsh(M.tiling.tk(vals, kg, mixing=2))
    ref = np.where(e < 0, e, 0.0).sum(1).mean() / 4
    print("spin flip: Omega %.10f, lattice energy per site %.10f" % (om, ref))
    assert abs(om - ref) < 2e-6
