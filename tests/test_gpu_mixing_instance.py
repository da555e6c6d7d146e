"""The spin-flip, anomalous (Nambu) and full mixings through the instance layer:
anomalous operators and spin-flip one-body terms, sector strings without N or S
("R0:S0", "R0:N4", "R0"), the spinor Green function of dimension n_mixed x n_sites
(src_ed/model_instance.hpp:710-817, src_ed/symmetry_group.cpp:616-645), and the spinor
one-body matrices (one_body_operator.hpp:326-411, anomalous_operator.hpp:241-284), against

* dense Jordan-Wigner Hamiltonians of the same states (E0, G of the spinor);
* the noninteracting limit: with U -> 0 the self-energy G0^-1 - G^-1 vanishes only if the
  spinor hopping matrices (cluster, bath and hybridization blocks, one-body and pairing
  terms) are those of the Hamiltonian the solver diagonalized;
* the reference's errors for sector strings that contradict the mixing.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O
from tests import models as TM

pytestmark = pytest.mark.gpu

_counter = itertools.count(1)


def register(model, values, sectors, n_sites):
    from pyqcm_amd import qcm
    name = "mix%d" % next(_counter)
    qcm.new_model(name, n_sites, model.L - n_sites)
    for opname, op in model.operators.items():
        qcm.new_operator(name, opname, op.kind, [(r + 1, c + 1, v) for (r, c, v) in op.elements])
    label = 9500 + next(_counter)
    qcm.new_model_instanceC(name, values, sectors, label)
    return label


def chain(hx=0.0, d=0.0, s=0.0, U=3.0):
    n = 4
    m, vals = TM.chain(n, tp=-0.3, U=U, mu=1.1)
    L = m.L
    if hx:
        m.add("hx", "one-body", [(i, i + L, 1.0) for i in range(n)])
        vals["hx"] = hx
    if d:
        m.add("D", "anomalous", [(i, i + 1 + L, 1.0) for i in range(n - 1)] + [(i + 1, i + L, 1.0) for i in range(n - 1)])
        vals["D"] = d
    if s:
        m.add("S", "anomalous", [(i, i + L, 1.0) for i in range(n)])
        vals["S"] = s
    return m, vals, n


def blocks(L, N=None, S=None, parity=None):
    if N is not None:
        return [(nu, N - nu) for nu in range(L + 1) if 0 <= N - nu <= L]
    if S is not None:
        return [(nu, nu - S) for nu in range(L + 1) if 0 <= nu - S <= L]
    return [(nu, nd) for nu in range(L + 1) for nd in range(L + 1) if (nu + nd) % 2 == parity]


def spinor(L, n, nambu, flip):
    out = [(i, False) for i in range(n)]
    if flip:
        out += [(i + L, False) for i in range(n)]
    if nambu:
        out += [(i, True) for i in range(n)] + [(i + L, True) for i in range(n)] if flip else [(i + L, True) for i in range(n)]
    return out


ZS = np.array([0.3j, 1.0j, 2.5j, 0.6 + 0.2j])


@pytest.mark.parametrize("case", ["nambu", "spin_flip", "full"])
def test_mixing_instance_ground_state_and_green_function(gpu_lib, case):
    from pyqcm_amd import qcm
    if case == "nambu":
        m, vals, n = chain(d=0.5, s=0.2)
        sec, b0, bm, bp = "R0:S0", blocks(m.L, S=0), blocks(m.L, S=-1), blocks(m.L, S=1)
        sp, mixing = spinor(m.L, n, True, False), 1
    elif case == "spin_flip":
        m, vals, n = chain(hx=0.4)
        sec, b0, bm, bp = "R0:N4", blocks(m.L, N=4), blocks(m.L, N=3), blocks(m.L, N=5)
        sp, mixing = spinor(m.L, n, False, True), 2
    else:
        m, vals, n = chain(hx=0.3, d=0.4)
        sec, b0, bm, bp = "R0", blocks(m.L, parity=0), blocks(m.L, parity=1), blocks(m.L, parity=1)
        sp, mixing = spinor(m.L, n, True, True), 3
    lab = register(m, vals, sec, n)
    E0, gs = qcm.ground_state_solve(lab)
    assert qcm.mixingC(lab) == mixing
    assert gs == sec + ":1"
    assert qcm.Green_function_dimensionC(lab) == len(sp)
    E0d, Gd = O.dense_spinor_gf(m, vals, O.block_states(m.L, b0), sp, O.block_states(m.L, bm),
                                O.block_states(m.L, bp), ZS)
    G = qcm.Green_function(ZS, 0, lab)
    print(case, "E0 %.12f dense %.12f" % (E0, E0d), "max |dG| %.2e" % np.abs(G - Gd).max())
    assert abs(E0 - E0d) <= 1e-10 * abs(E0d)
    np.testing.assert_allclose(G, Gd, atol=1e-9)
    # G(z) = G(z*)^* and the particle-hole structure leaves G real-symmetric
    np.testing.assert_allclose(G, np.transpose(G, (0, 2, 1)), atol=1e-12)


def nambu_bath_model(U):
    """2 sites + 2 bath orbitals with every kind of term: site hopping and mu, bath
    energies, site-bath hybridization, site pairing (cluster), bath pairing (bath) and
    site-bath pairing (hybridization)."""
    ns, nb = 2, 2
    L = ns + nb
    m = O.ClusterModel(L)
    m.add("t", "one-body", [(0, 1, -1.0), (L, L + 1, -1.0)])
    m.add("U", "interaction", [(i, i + L, 1.0) for i in range(ns)])
    m.add("mu", "one-body", [(i, i, -1.0) for i in range(ns)] + [(i + L, i + L, -1.0) for i in range(ns)])
    m.add("eb", "one-body", [(2, 2, 1.0), (2 + L, 2 + L, 1.0), (3, 3, -1.0), (3 + L, 3 + L, -1.0)])
    m.add("tb", "one-body", [(0, 2, 1.0), (0 + L, 2 + L, 1.0), (1, 3, 1.0), (1 + L, 3 + L, 1.0)])
    m.add("D", "anomalous", [(0, 1 + L, 1.0), (1, 0 + L, 1.0)])
    m.add("Db", "anomalous", [(2, 2 + L, 1.0), (3, 3 + L, 1.0)])
    m.add("Dh", "anomalous", [(0, 2 + L, 1.0), (1, 3 + L, 0.5)])
    vals = {"t": 1.0, "U": U, "mu": 0.4, "eb": 0.8, "tb": 0.6, "D": 0.3, "Db": 0.25, "Dh": 0.2}
    return m, vals, ns


def test_nambu_self_energy_vanishes_without_interaction(gpu_lib):
    """U = 1e-9: Sigma = z - t_c - Gamma(z) - G^-1 vanishes to O(U) only if the spinor
    matrices t_c, t_b, t_cb (and the pairing blocks of each) match the Hamiltonian the
    solver diagonalized; the ED G also equals the dense one."""
    from pyqcm_amd import qcm
    m, vals, ns = nambu_bath_model(1e-9)
    lab = register(m, vals, "R0:S0", ns)
    E0, gs = qcm.ground_state_solve(lab)
    assert qcm.mixingC(lab) == 1
    sp = [(i, False) for i in range(ns)] + [(i + m.L, True) for i in range(ns)]
    _, Gd = O.dense_spinor_gf(m, vals, O.block_states(m.L, blocks(m.L, S=0)), sp,
                              O.block_states(m.L, blocks(m.L, S=-1)), O.block_states(m.L, blocks(m.L, S=1)), ZS)
    G = qcm.Green_function(ZS, 0, lab)
    np.testing.assert_allclose(G, Gd, atol=1e-9)
    Sig = qcm.self_energyC(ZS, 0, lab)
    print("max |Sigma| at U = 1e-9: %.2e" % np.abs(Sig).max())
    assert np.abs(Sig).max() < 1e-7
    # with the interaction on, Sigma is no longer small (the check above has teeth)
    m2, vals2, _ = nambu_bath_model(2.0)
    lab2 = register(m2, vals2, "R0:S0", ns)
    assert np.abs(qcm.self_energyC(ZS, 0, lab2)).max() > 1e-3


def test_full_mixing_self_energy_vanishes_without_interaction(gpu_lib):
    """The same model with a transverse field on sites and bath (spin flip and pairing:
    mixing 3, spinor of 4 n_sites): Sigma vanishes at U = 1e-9 and G is the dense one."""
    from pyqcm_amd import qcm
    m, vals, ns = nambu_bath_model(1e-9)
    L = m.L
    m.add("hx", "one-body", [(i, i + L, 1.0) for i in range(ns)])
    m.add("hxb", "one-body", [(i, i + L, 1.0) for i in range(ns, L)])
    vals["hx"] = 0.3
    vals["hxb"] = 0.15
    lab = register(m, vals, "R0", ns)
    qcm.ground_state_solve(lab)
    assert qcm.mixingC(lab) == 3
    sp = ([(i, False) for i in range(ns)] + [(i + L, False) for i in range(ns)] + [(i, True) for i in range(ns)]
          + [(i + L, True) for i in range(ns)])
    _, Gd = O.dense_spinor_gf(m, vals, O.block_states(L, blocks(L, parity=0)), sp,
                              O.block_states(L, blocks(L, parity=1)), O.block_states(L, blocks(L, parity=1)), ZS)
    np.testing.assert_allclose(qcm.Green_function(ZS, 0, lab), Gd, atol=1e-9)
    Sig = qcm.self_energyC(ZS, 0, lab)
    print("full: max |Sigma| at U = 1e-9: %.2e" % np.abs(Sig).max())
    assert np.abs(Sig).max() < 1e-7


def test_spin_flip_self_energy_vanishes_without_interaction(gpu_lib):
    from pyqcm_amd import qcm
    m, vals, n = chain(hx=0.35, U=1e-9)
    lab = register(m, vals, "R0:N4", n)
    qcm.ground_state_solve(lab)
    Sig = qcm.self_energyC(ZS, 0, lab)
    print("spin flip: max |Sigma| at U = 1e-9: %.2e" % np.abs(Sig).max())
    assert np.abs(Sig).max() < 1e-7


def test_mixing_sector_errors(gpu_lib):
    from pyqcm_amd import qcm
    m, vals, ns = nambu_bath_model(1.0)
    m.add("bad", "one-body", [(0, 0, 1.0), (2, 2, 1.0)])  # a site and a bath orbital
    with pytest.raises(RuntimeError, match="no unique target"):
        register(m, vals, "R0:S0", ns)
    m, vals, n = chain(d=0.5)
    with pytest.raises(RuntimeError, match="non conserved particle number|conserved particle number"):
        register(m, vals, "R0:N4:S0", n)
    m, vals, n = chain()
    with pytest.raises(RuntimeError, match="non conserved particle number"):
        register(m, vals, "R0:S0", n)
    m, vals, n = chain(hx=0.3)
    with pytest.raises(RuntimeError, match="conserved spin"):
        register(m, vals, "R0:N4:S0", n)
    m, vals, n = chain(d=0.5)
    lab = register(m, vals, "R0:S0", n)
    with pytest.raises(RuntimeError, match="mixing"):
        qcm.Green_function(ZS, 1, lab)


@pytest.mark.parametrize("case", ["nambu", "spin_flip", "full"])
def test_mixing_cluster_averages_match_dense(gpu_lib, case):
    """Ground-state averages and variances of every operator, the pairing and spin-flip
    ones included (model_instance::cluster_averages, src_ed/model_instance.hpp:152-163),
    against <psi|O|psi> and <psi|O^2|psi> - <O>^2 of the dense ground state; site
    operators are divided by the number of sites, mu with the opposite sign."""
    from pyqcm_amd import qcm
    if case == "nambu":
        m, vals, n = chain(d=0.5, s=0.2)
        sec, b0 = "R0:S0", blocks(m.L, S=0)
    elif case == "spin_flip":
        m, vals, n = chain(hx=0.4)
        sec, b0 = "R0:N4", blocks(m.L, N=4)
    else:
        m, vals, n = chain(hx=0.3, d=0.4)
        sec, b0 = "R0", blocks(m.L, parity=0)
    lab = register(m, vals, sec, n)
    qcm.ground_state_solve(lab)
    ave = qcm.cluster_averages(lab)
    states = O.block_states(m.L, b0)
    e, U = np.linalg.eigh(O.dense_hamiltonian(m, vals, 0, 0, states=states)[0])
    assert e[1] - e[0] > 1e-6
    psi = U[:, 0]
    for name, op in m.operators.items():
        if vals.get(name, 0.0) == 0.0:
            continue
        one = O.ClusterModel(m.L)
        one.add(name, op.kind, op.elements)
        Om = O.dense_hamiltonian(one, {name: 1.0}, 0, 0, states=states)[0]
        a = psi @ Om @ psi
        v = psi @ Om @ (Om @ psi) - a * a
        sgn = -1.0 if name == "mu" else 1.0
        print(case, name, "average %.12f dense %.12f" % (ave[name][0], sgn * a / n))
        assert abs(ave[name][0] - sgn * a / n) < 1e-9, name
        assert abs(ave[name][1] - v / n ** 2) < 1e-9, name


@pytest.mark.parametrize("case", ["nambu", "spin_flip", "full"])
def test_mixing_continued_fraction_green_function(gpu_lib, case):
    """GF_method 'F' (one Lanczos continued fraction per pair of spinor components,
    model_instance<double>::build_cf, src_ed/model_instance.cpp:34-99) against the dense
    spinor G and against the instance's own Lehmann form; the stop criterion is tightened
    so that the fractions are exhaustive on these sectors. The integrated G (the hole
    fractions' weights) equals the Lehmann form's."""
    from pyqcm_amd import qcm
    if case == "nambu":
        m, vals, n = chain(d=0.5, s=0.2)
        sec, b0, bm, bp = "R0:S0", blocks(m.L, S=0), blocks(m.L, S=-1), blocks(m.L, S=1)
        sp = spinor(m.L, n, True, False)
    elif case == "spin_flip":
        m, vals, n = chain(hx=0.4)
        sec, b0, bm, bp = "R0:N4", blocks(m.L, N=4), blocks(m.L, N=3), blocks(m.L, N=5)
        sp = spinor(m.L, n, False, True)
    else:
        m, vals, n = chain(hx=0.3, d=0.4)
        sec, b0, bm, bp = "R0", blocks(m.L, parity=0), blocks(m.L, parity=1), blocks(m.L, parity=1)
        sp = spinor(m.L, n, True, True)
    lab_q = register(m, vals, sec, n)
    Gq = qcm.Green_function(ZS, 0, lab_q)
    qcm.set_global_parameter("GF_method", "F")
    qcm.set_global_parameter("accur_continued_fraction", 1e-12)
    try:
        lab = register(m, vals, sec, n)
        G = qcm.Green_function(ZS, 0, lab)
    finally:
        qcm.set_global_parameter("GF_method", "L")
        qcm.set_global_parameter("accur_continued_fraction", 0.01)
    _, Gd = O.dense_spinor_gf(m, vals, O.block_states(m.L, b0), sp, O.block_states(m.L, bm),
                              O.block_states(m.L, bp), ZS)
    print(case, "max |G_cf - G_dense| %.2e, |G_cf - G_lehmann| %.2e" % (np.abs(G - Gd).max(), np.abs(G - Gq).max()))
    np.testing.assert_allclose(G, Gd, atol=1e-9)
    np.testing.assert_allclose(G, Gq, atol=1e-9)


@pytest.mark.parametrize("case", ["nambu", "spin_flip", "full"])
def test_mixing_green_function_average_and_density(gpu_lib, case):
    """The integrated spinor G (model_instance::Green_function_average,
    src_ed/model_instance.hpp:643-654: the Lehmann poles below zero) equals
    <A_b psi, A_a psi> of the dense ground state for the spinor components A (c or c+),
    and the density (src_ed/model_instance.hpp:662-675: 1 - <c c+> on the Nambu
    components) equals <n> per site."""
    from pyqcm_amd import qcm
    if case == "nambu":
        m, vals, n = chain(d=0.5, s=0.2)
        sec, b0 = "R0:S0", blocks(m.L, S=0)
        sp = spinor(m.L, n, True, False)
    elif case == "spin_flip":
        m, vals, n = chain(hx=0.4)
        sec, b0 = "R0:N4", blocks(m.L, N=4)
        sp = spinor(m.L, n, False, True)
    else:
        m, vals, n = chain(hx=0.3, d=0.4)
        sec, b0 = "R0", blocks(m.L, parity=0)
        sp = spinor(m.L, n, True, True)
    lab = register(m, vals, sec, n)
    qcm.ground_state_solve(lab)
    M = qcm.Green_function_average(lab)
    H, (c, cd) = O.dense_hamiltonian(m, vals, 0, 0, states=O.block_states(m.L, b0))
    psi = np.zeros(1 << (2 * m.L))
    psi[O.block_states(m.L, b0)] = np.linalg.eigh(H)[1][:, 0]
    A = [(cd[a] if nambu else c[a]) @ psi for a, nambu in sp]
    ref = np.array([[A[b] @ A[a] for b in range(len(sp))] for a in range(len(sp))])
    print(case, "max |M - dense| %.1e" % np.abs(M - ref).max())
    np.testing.assert_allclose(M, ref, atol=1e-10)
    dens = sum(psi @ (cd[a] @ (c[a] @ psi)) for a in list(range(n)) + list(range(m.L, m.L + n))) / n
    assert abs(qcm.Green_function_density(lab) - dens) < 1e-10


def test_nambu_tr_sigma_inf_and_hybridization_lehmann(gpu_lib):
    """tr_sigma_inf with the Nambu spinor (src_ed/model_instance.hpp:1279-1311: hole
    components with the opposite sign) is the Hartree term U <N> of the dense ground
    state; the Lehmann form of the hybridization (the bath diagonalized in place,
    model_instance::hybridization) rebuilds the solver's Gamma(z), the paired bath
    included."""
    from pyqcm_amd import qcm
    m, vals, ns = nambu_bath_model(2.0)
    lab = register(m, vals, "R0:S0", ns)
    qcm.ground_state_solve(lab)
    L = m.L
    states = O.block_states(L, blocks(L, S=0))
    H, (c, cd) = O.dense_hamiltonian(m, vals, 0, 0, states=states)
    psi = np.zeros(1 << (2 * L))
    psi[states] = np.linalg.eigh(H)[1][:, 0]
    N = sum(psi @ (cd[a] @ (c[a] @ psi)) for a in list(range(ns)) + list(range(L, L + ns)))
    t = qcm.tr_sigma_inf(lab)
    print("Nambu tr_sigma_inf %.8f, U <N> %.8f" % (t, vals["U"] * N))
    assert abs(t - vals["U"] * N) < 1e-5
    E, T = qcm.hybridization(lab)
    assert E.shape == (2 * (L - ns),) and T.shape == (2 * ns, 2 * (L - ns))
    tcb = T.reshape(-1).reshape(2 * (L - ns), 2 * ns).T.real  # column-major 2ns x 2nb
    z = np.array([0.4j, 1.7j, 0.3 + 0.2j])
    Gam = np.array([(tcb / (zz - E)) @ tcb.T for zz in z])
    np.testing.assert_allclose(Gam, qcm.hybridization_functionC(z, 0, lab), atol=1e-13)


def test_fidelity_of_nambu_and_symmetric_instances(gpu_lib):
    """|<psi_A|psi_B>|^2 of two instances of one model in one sector: the Nambu chain at
    two pairing amplitudes, and the plaquette's mirror basis at two values of U, against
    the dense ground states."""
    from pyqcm_amd import qcm
    from tests import test_gpu_symmetry_instance as SI
    m, va, n = chain(d=0.5, s=0.2)
    vb = dict(va, D=0.35)
    name = "fid%d" % next(_counter)
    qcm.new_model(name, n, m.L - n)
    for opname, op in m.operators.items():
        qcm.new_operator(name, opname, op.kind, [(r + 1, c + 1, v) for (r, c, v) in op.elements])
    qcm.new_model_instanceC(name, va, "R0:S0", 9801)
    qcm.new_model_instanceC(name, vb, "R0:S0", 9802)
    states = O.block_states(m.L, blocks(m.L, S=0))
    ga = np.linalg.eigh(O.dense_hamiltonian(m, va, 0, 0, states=states)[0])[1][:, 0]
    gb = np.linalg.eigh(O.dense_hamiltonian(m, vb, 0, 0, states=states)[0])[1][:, 0]
    f = qcm.fidelity(9801, 9802)
    print("Nambu fidelity %.12f dense %.12f" % (f, (ga @ gb) ** 2))
    assert abs(f - (ga @ gb) ** 2) < 1e-9
    m, vals, ns, gens, bi = SI._plaquette4()
    name = "fids%d" % next(_counter)
    qcm.new_model(name, ns, m.L - ns, gens, bi)
    for opname, op in m.operators.items():
        qcm.new_operator(name, opname, op.kind, [(r + 1, c + 1, v) for (r, c, v) in op.elements])
    v2 = dict(vals, U=vals["U"] + 0.5)
    qcm.new_model_instanceC(name, vals, "R0:N8:S0", 9803)
    qcm.new_model_instanceC(name, v2, "R0:N8:S0", 9804)
    g1 = np.linalg.eigh(O.dense_hamiltonian(m, vals, 4, 4)[0])[1][:, 0]
    g2 = np.linalg.eigh(O.dense_hamiltonian(m, v2, 4, 4)[0])[1][:, 0]
    f = qcm.fidelity(9803, 9804)
    print("mirror-basis fidelity %.12f dense %.12f" % (f, (g1 @ g2) ** 2))
    assert abs(f - (g1 @ g2) ** 2) < 1e-9


@pytest.mark.parametrize("case", ["nambu", "spin_flip", "full"])
def test_mixing_uncorrelated_solution(gpu_lib, case):
    """No interaction: the one-body solution of the spinor (model_instance::one_body_solve,
    src_ed/model_instance.hpp:1164-1256; Nambu correction of the energy) against the
    lowest state of the whole Fock space and its spinor G from dense ED."""
    from pyqcm_amd import qcm
    if case == "nambu":
        m, vals, n = chain(d=0.5, s=0.2)
        sec, sp = "R0:S0", spinor(m.L, n, True, False)
    elif case == "spin_flip":
        m, vals, n = chain(hx=0.4)
        sec, sp = "R0:N4", spinor(m.L, n, False, True)
    else:
        m, vals, n = chain(hx=0.3, d=0.4)
        sec, sp = "R0", spinor(m.L, n, True, True)
    del vals["U"]
    lab = register(m, vals, sec, n)
    E0, gs = qcm.ground_state_solve(lab)
    assert gs == "uncorrelated"
    allst = O.block_states(m.L, [(nu, nd) for nu in range(m.L + 1) for nd in range(m.L + 1)])
    e = np.linalg.eigvalsh(O.dense_hamiltonian(m, vals, 0, 0, states=allst)[0])
    assert e[1] - e[0] > 1e-6
    Ed, Gd = O.dense_spinor_gf(m, vals, allst, sp, allst, allst, ZS)
    G = qcm.Green_function(ZS, 0, lab)
    print(case, "E0 %.12f dense %.12f, max |dG| %.1e" % (E0, Ed, np.abs(G - Gd).max()))
    assert abs(E0 - Ed) < 1e-10
    np.testing.assert_allclose(G, Gd, atol=1e-10)


@pytest.mark.parametrize("case", ["nambu", "spin_flip", "full"])
def test_mixing_matrix_continued_fraction(gpu_lib, case):
    """GF_method 'M' with the spinor's components as the start block (block Lanczos per
    shifted sector, model_instance::build_mcf, src_ed/model_instance.hpp:846-946),
    against the oracle's block Lanczos and matrix continued fraction on the same start
    blocks of the dense ground state. The spin-flip blocks deflate at the second floor
    (the Krylov space grows by 4, not 8, per step: S_x is conserved) and the fraction stops
    there, as the reference's does, so only the Nambu and full cases are also the exact G;
    with combine_mcf the combined fraction gives the same G."""
    from pyqcm_amd import qcm
    if case == "nambu":
        m, vals, n = chain(d=0.5, s=0.2)
        sec, b0, bm, bp = "R0:S0", blocks(m.L, S=0), blocks(m.L, S=-1), blocks(m.L, S=1)
        sp = spinor(m.L, n, True, False)
    elif case == "spin_flip":
        m, vals, n = chain(hx=0.4)
        sec, b0, bm, bp = "R0:N4", blocks(m.L, N=4), blocks(m.L, N=3), blocks(m.L, N=5)
        sp = spinor(m.L, n, False, True)
    else:
        m, vals, n = chain(hx=0.3, d=0.4)
        sec, b0, bm, bp = "R0", blocks(m.L, parity=0), blocks(m.L, parity=1), blocks(m.L, parity=1)
        sp = spinor(m.L, n, True, True)
    _, Gd = O.dense_spinor_gf(m, vals, O.block_states(m.L, b0), sp, O.block_states(m.L, bm),
                              O.block_states(m.L, bp), ZS)
    qcm.set_global_parameter("GF_method", "M")
    try:
        G = qcm.Green_function(ZS, 0, register(m, vals, sec, n))
        qcm.set_global_parameter("combine_mcf", True)
        Gc = qcm.Green_function(ZS, 0, register(m, vals, sec, n))
    finally:
        qcm.set_global_parameter("GF_method", "L")
        qcm.set_global_parameter("combine_mcf", False)
    st0 = O.block_states(m.L, b0)
    H, (c, cd) = O.dense_hamiltonian(m, vals, 0, 0, states=st0)
    e, U = np.linalg.eigh(H)
    psi = np.zeros(1 << (2 * m.L))
    psi[st0] = U[:, 0]
    Go = np.zeros_like(Gd)
    for pm, bt in ((-1, bm), (1, bp)):
        stt = O.block_states(m.L, bt)
        Ht = O.dense_hamiltonian(m, vals, 0, 0, states=stt)[0]
        # A_a = c_a (c+_a for a Nambu component), A+_a its conjugate
        phi = [(((cd[a] if nambu else c[a]) if pm < 0 else (c[a] if nambu else cd[a])) @ psi)[stt]
               for a, nambu in sp]
        q, W = O.mgs_weight(phi)
        A, B = O.block_lanczos(lambda x: Ht @ x, q)
        A = O.mcf_shift(A, e[0], pm == 1)
        for iz, z in enumerate(ZS):
            g = O.mcf_evaluate(A, B, W, z)
            Go[iz] += g if pm == 1 else g.T
    print(case, "max |G_mcf - G_oracle| %.1e, |G_mcf - G_dense| %.1e, combined - oracle %.1e"
          % (np.abs(G - Go).max(), np.abs(G - Gd).max(), np.abs(Gc - Go).max()))
    np.testing.assert_allclose(G, Go, atol=1e-9)
    np.testing.assert_allclose(Gc, Go, atol=1e-9)
    if case != "spin_flip":
        np.testing.assert_allclose(G, Gd, atol=1e-9)


@pytest.mark.parametrize("method", ["L", "M"])
def test_nambu_instance_file_round_trip(gpu_lib, tmp_path, method):
    """An instance of the Nambu mixing with a paired bath written to an instance file and
    read into a fresh instance of the model: the sector string with its code for N
    ("R0:S0"), G(z), Sigma(z) and the Potthoff functional bit for bit, for the Lehmann
    form and the matrix continued fraction."""
    from pyqcm_amd import qcm
    m, vals, ns = nambu_bath_model(2.0)
    qcm.set_global_parameter("GF_method", method)
    try:
        a = register(m, vals, "R0:S0", ns)
        E0, sec = qcm.ground_state_solve(a)
        G, S = qcm.Green_function(ZS, 0, a), qcm.self_energyC(ZS, 0, a)
        path = str(tmp_path / "nambu.h5")
        qcm.write_instance_to_hdf5(path, "cluster_0", a)
        b = register(m, vals, "R0:S0", ns)
        qcm.read_instance_from_hdf5(path, "cluster_0", b)
        E0b, secb = qcm.ground_state_solve(b)
        np.testing.assert_array_equal(qcm.Green_function(ZS, 0, b), G)
        np.testing.assert_array_equal(qcm.self_energyC(ZS, 0, b), S)
        assert qcm.Potthoff_functional(b) == qcm.Potthoff_functional(a)
    finally:
        qcm.set_global_parameter("GF_method", "L")
    assert sec == secb == "R0:S0:1"
    assert abs(E0b - E0) <= 1e-12 * abs(E0)


@pytest.mark.parametrize("case", ["nambu", "spin_flip", "full"])
def test_mixing_uncorrelated_averages(gpu_lib, case):
    """Operator averages of an uncorrelated spinor instance from its integrated Green
    function and the operator's spinor matrix, with the Nambu correction
    (src_ed/model_instance.hpp:144-163), against the lowest state of the whole Fock space."""
    from pyqcm_amd import qcm
    if case == "nambu":
        m, vals, n = chain(d=0.5, s=0.2)
        sec = "R0:S0"
    elif case == "spin_flip":
        m, vals, n = chain(hx=0.4)
        sec = "R0:N4"
    else:
        m, vals, n = chain(hx=0.3, d=0.4)
        sec = "R0"
    del vals["U"]
    lab = register(m, vals, sec, n)
    qcm.ground_state_solve(lab)
    ave = qcm.cluster_averages(lab)
    allst = O.block_states(m.L, [(nu, nd) for nu in range(m.L + 1) for nd in range(m.L + 1)])
    psi = np.linalg.eigh(O.dense_hamiltonian(m, vals, 0, 0, states=allst)[0])[1][:, 0]
    for name, op in m.operators.items():
        if vals.get(name, 0.0) == 0.0:
            continue
        one = O.ClusterModel(m.L)
        one.add(name, op.kind, op.elements)
        a = psi @ O.dense_hamiltonian(one, {name: 1.0}, 0, 0, states=allst)[0] @ psi
        sgn = -1.0 if name == "mu" else 1.0
        print(case, name, "average %.12f dense %.12f" % (ave[name][0], sgn * a / n))
        assert abs(ave[name][0] - sgn * a / n) < 1e-10, name


def test_hopping_matrix_full_with_a_paired_bath(gpu_lib):
    """hopping_matrix(full=1) of a Nambu instance with a paired bath: the cluster block is
    t_c, the bath block as the operators give it (pairing off the diagonal) or, with diag,
    the bath energies of the Lehmann form of the hybridization; both forms have the same
    spectrum (the bath is rotated by an orthogonal matrix)."""
    from pyqcm_amd import qcm
    m, vals, ns = nambu_bath_model(2.0)
    lab = register(m, vals, "R0:S0", ns)
    qcm.ground_state_solve(lab)
    H0 = qcm.hopping_matrix(0, 0, lab, 1)
    Hd = qcm.hopping_matrix(0, 1, lab, 1)
    d = 2 * ns
    tc = qcm.hopping_matrix(0, 0, lab)
    np.testing.assert_array_equal(H0[:d, :d], tc)
    np.testing.assert_array_equal(Hd[:d, :d], tc)
    E, _ = qcm.hybridization(lab)
    np.testing.assert_allclose(np.diag(Hd[d:, d:]).real, E, atol=1e-14)
    assert np.abs(H0[d:, d:] - np.diag(np.diag(H0[d:, d:]))).max() > 0.1
    np.testing.assert_allclose(np.linalg.eigvalsh(H0), np.linalg.eigvalsh(Hd), atol=1e-12)


def test_nambu_print_wavefunction(gpu_lib):
    """print_wavefunction of a Nambu instance: the sector string with its code ("R0:S0"),
    and per basis state of the mixed-number sector |psi|^2, psi and |up,down> (the mixed
    basis' PrintBinaryDouble, src_ed/ED_basis.cpp:391-399), against the dense ground state
    of the same states in the same order."""
    from pyqcm_amd import qcm
    m, vals, n = chain(d=0.5, s=0.2)
    lab = register(m, vals, "R0:S0", n)
    qcm.set_global_parameter("max_dim_print", 100)
    try:
        lines = qcm.print_wavefunction(lab).strip("\n").split("\n")
    finally:
        qcm.set_global_parameter("max_dim_print", 64)
    states = O.block_states(m.L, blocks(m.L, S=0))
    psi = np.linalg.eigh(O.dense_hamiltonian(m, vals, 0, 0, states=states)[0])[1][:, 0]
    assert lines[0] == "state" and lines[1].split("\t")[0] == "R0:S0"
    assert len(lines) == 2 + len(states)
    L = m.L
    bits = lambda v: "".join("1" if (v >> k) & 1 else "0" for k in range(L))
    for i, ln in enumerate(lines[2:]):
        p2, amp, st = ln.split("\t")
        up, dn = int(states[i]) & ((1 << L) - 1), int(states[i]) >> L
        assert st == "|%s,%s>" % (bits(up), bits(dn))
        assert abs(float(p2) - psi[i] ** 2) < 1e-5
