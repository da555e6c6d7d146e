"""Green-function averages and the remaining instance queries of namespace ED:
Green_function_average / Green_function_density (model_instance.hpp:643-675), fidelity
(:1261-1274), tr_sigma_inf (:1277-1312) and the Lehmann form of the hybridization
(:1455-1466), against brute force and closed forms."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import models as TM
from tests.test_gpu_ed import register, with_sites

pytestmark = pytest.mark.gpu


def _plaquette4():
    """2x2 plaquette + 4 bath orbitals (L = 8) with a particle-hole symmetric bath at
    mu = U/2: half filling (N = 8, S = 0) is the ground state of every particle number, so
    the Lehmann form's poles below zero are exactly the hole poles. (The Lehmann average
    integrates every pole below zero, as the reference's does; for a sector that is not
    the ground state of all particle numbers that includes electron poles.)"""
    p = {"U": 4.0, "mu": 2.0, "eb1": 0.6, "tb1": 0.5, "eb2": -0.6, "tb2": 0.5, "eb3": 1.1, "tb3": 0.4,
         "eb4": -1.1, "tb4": 0.4}
    return TM.plaquette_bath(p, n_bath=4)


def _bruteforce_average(m, vals, nup, ndn, n_sites):
    """<c^dagger_a c_b> (spin up) from the oracle's ground state: the Gram matrix of c_a|psi>."""
    E0, gs, *_ = O.Sector(m, vals, nup, ndn).lanczos_gs()
    phi = np.array([O.create_destroy(m.L, nup, ndn, a, -1, gs) for a in range(n_sites)])
    return E0, gs, phi @ phi.T


@pytest.mark.parametrize("method", ["L", "F", "M"])
def test_green_function_average_and_density(gpu_lib, method):
    """2x2 plaquette + 4 bath orbitals, N = 8, S = 0 (D = 4,900): the integrated Green
    function in each representation is the one-body density matrix of the ground state
    (the hole spectral weight is complete in all three), and the density its trace."""
    from pyqcm_amd import qcm
    m, vals = _plaquette4()
    qcm.set_global_parameter("GF_method", method)
    try:
        lab = register(with_sites(m, 4), vals, "R0:N8:S0")
        M = qcm.Green_function_average(lab)
        Md = qcm.Green_function_average(lab, 1)
        n = qcm.Green_function_density(lab)
    finally:
        qcm.set_global_parameter("GF_method", "L")
    _, _, ref = _bruteforce_average(m, vals, 4, 4, 4)
    err = np.abs(M - ref).max()
    print("%s: max |M - <c+c>| = %.3g, density %.12f" % (method, err, n))
    assert err < 1e-7
    np.testing.assert_array_equal(M, Md)  # mixing 0: one spin
    assert abs(n - 2 * np.trace(ref) / 4) < 1e-7


def test_fidelity(gpu_lib):
    """An instance with itself: 1. Two instances at U = 4 and U = 4.5: |<psi1|psi2>|^2
    from the oracle's ground states."""
    from pyqcm_amd import qcm
    m, vals = _plaquette4()
    a = register(with_sites(m, 4), vals, "R0:N8:S0")
    v2 = dict(vals, U=vals["U"] + 0.5)
    b = register(with_sites(m, 4), v2, "R0:N8:S0")
    assert abs(qcm.fidelity(a, a) - 1.0) < 1e-12
    f = qcm.fidelity(a, b)
    _, g1, _ = _bruteforce_average(m, vals, 4, 4, 4)
    _, g2, _ = _bruteforce_average(m, v2, 4, 4, 4)
    ref = float(np.dot(g1, g2)) ** 2
    print("fidelity %.12f, oracle %.12f" % (f, ref))
    assert abs(f - ref) < 1e-7


def test_tr_sigma_inf_is_the_hartree_term(gpu_lib):
    """Sigma(z -> infinity) of the Hubbard interaction is U <n_i, other spin> on each
    site: the trace over both spins is 2 U sum_i <n_i up> (mixing 0)."""
    from pyqcm_amd import qcm
    m, vals = _plaquette4()
    lab = register(with_sites(m, 4), vals, "R0:N8:S0")
    t = qcm.tr_sigma_inf(lab)
    _, _, ref = _bruteforce_average(m, vals, 4, 4, 4)
    hartree = 2 * vals["U"] * np.trace(ref)
    print("tr_sigma_inf %.10f, Hartree %.10f" % (t, hartree))
    assert abs(t - hartree) < 1e-6 * max(1.0, abs(hartree))


def test_hybridization_lehmann_form(gpu_lib):
    """The bath energies and the cluster-bath matrix of the model's parameters, in the
    reference binding's layout; a model without bath raises the reference's message."""
    from pyqcm_amd import qcm
    m, vals = _plaquette4()
    lab = register(with_sites(m, 4), vals, "R0:N8:S0")
    E, T = qcm.hybridization(lab)
    np.testing.assert_allclose(E, [vals["eb%d" % x] for x in range(1, 5)], atol=1e-15)
    tcb = T.reshape(-1).reshape(4, 4).T.real  # the column-major n_sites x n_bath matrix
    ref = np.array([[vals["tb%d" % x] * TM.PLAQ_SEQ[x][i] for x in range(1, 5)] for i in range(4)])
    np.testing.assert_allclose(tcb, ref, atol=1e-15)
    mc, vc = TM.chain(4)
    lc = register(with_sites(mc, 4), vc, "R0:N4:S0")
    with pytest.raises(Exception, match="System has no bath. Hybridization not defined."):
        qcm.hybridization(lc)


def test_model_and_instance_queries(gpu_lib):
    """matrix_elements, parametersC, interactions, mixingC, ED_complex_HS and the refusal of
    complex operator elements (src_python/qcm_ED_wrap.hpp:27-30,108-135,219-273)."""
    from pyqcm_amd import qcm
    m, vals = TM.chain(4, U=4.0, mu=2.0)
    lab = register(with_sites(m, 4), vals, "R0:N5:S1")
    name = qcm._instance_model[lab]
    t, els = qcm.matrix_elements(name, "U")
    assert t == "interaction" and els == [(i, i + 4, 1.0) for i in range(4)]
    p, mname = qcm.parametersC(lab)
    assert mname == name and p == dict(sorted(vals.items()))
    assert qcm.interactions(lab) == [(i, i + 4, 4.0) for i in range(4)]
    qcm.ground_state_solve(lab)
    assert qcm.mixingC(lab) == 4  # N = 5, S = 1: spin up and down apart
    assert qcm.ED_complex_HS(lab) == 0
    qcm.new_model("cplx_ops", 2, 0)  # operators are added before the first instance
    qcm.new_operator_complex("cplx_ops", "h_real", "one-body", [(1, 2, 0.5 + 0.0j)])
    assert qcm.matrix_elements("cplx_ops", "h_real") == ("one-body", [(0, 1, 0.5)])
    with pytest.raises(RuntimeError, match="complex matrix elements"):
        qcm.new_operator_complex("cplx_ops", "h_cplx", "one-body", [(1, 2, 0.5j)])
    with pytest.raises(RuntimeError, match="The label 424242 is out of range."):
        qcm.parametersC(424242)


def _one_body_average(m, op, M, n_sites):
    """<O> for a cluster one-body operator from the spin-up density matrix M(b, a) =
    <c+_a c_b>, both spins (normal mixing), times the reference's norm (1/n_sites, -1 for mu)."""
    a = 0.0
    for r, c, v in m.operators[op].elements:
        if r >= m.L or c >= m.L:
            continue
        a += v * M[r, c] + (v * M[c, r] if r != c else 0.0)
    return 2 * a / n_sites * (-1.0 if op == "mu" else 1.0)


def test_cluster_averages_from_green_function(gpu_lib, tmp_path):
    """An instance read from a file and an uncorrelated instance take their averages from
    the Green function average (one-body cluster operators; NaN otherwise): equal to the
    averages of the solved instance's ground state, and to the Slater determinant at U = 0."""
    from pyqcm_amd import qcm
    m, vals = TM.chain(4, U=4.0, mu=2.0)
    a = register(with_sites(m, 4), vals, "R0:N4:S0")
    ref = qcm.cluster_averages(a)
    path = str(tmp_path / "avg.h5")
    qcm.write_instance_to_hdf5(path, "cluster_0", a)
    b = register(with_sites(m, 4), vals, "R0:N4:S0")
    qcm.read_instance_from_hdf5(path, "cluster_0", b)
    got = qcm.cluster_averages(b)
    for op in ("t", "mu"):
        assert abs(got[op][0] - ref[op][0]) < 1e-9, op
        assert np.isnan(got[op][1])
    assert np.isnan(got["U"][0]) and np.isnan(got["U"][1])
    # U = 0: one-body solve, the grand-canonical Slater determinant (levels below zero)
    v0 = dict(vals, U=0.0)
    c = register(with_sites(m, 4), v0, "R0:N4:S0")
    got0 = qcm.cluster_averages(c)
    h = np.zeros((4, 4))
    for op in ("t", "mu"):
        for r, cc, v in m.operators[op].elements:
            if r < 4 and cc < 4:
                h[r, cc] += v * v0[op]
                if r != cc:
                    h[cc, r] += v * v0[op]
    e, U = np.linalg.eigh(h)
    occ = U[:, e < 0]
    Mslater = occ @ occ.T
    for op in ("t", "mu"):
        assert abs(got0[op][0] - _one_body_average(m, op, Mslater, 4)) < 1e-12, op
    qcm.set_global_parameter("Hamiltonian_format", "F")
    try:
        assert qcm.density_matrix([0, 1], a)[0].shape == (0, 0)
    finally:
        qcm.set_global_parameter("Hamiltonian_format", "E")


def test_reduced_density_matrix(gpu_lib):
    """ED::density_matrix with the default (operator) format: the reduced density matrix of
    the ground state on sites {0, 1} of the 4-site chain (density_matrix_mixed,
    src_ed/model_instance.hpp:1547-1625: subsystem states numbered by first appearance in
    the basis order, collapsed labels, no fermionic sign) against the dense ground state
    restated the same way; unit trace, symmetric, positive."""
    from pyqcm_amd import qcm
    m, vals = TM.chain(4, U=4.0, mu=2.0, tp=-0.2)
    a = register(with_sites(m, 4), vals, "R0:N4:S0")
    rho, st = qcm.density_matrix([0, 1], a)
    H = O.dense_hamiltonian(m, vals, 2, 2)[0]
    psi = np.linalg.eigh(H)[1][:, 0]
    labels = O.sector_states(4, 2, 2)  # up | down << 4, the factorized basis order
    b64 = [(int(x) & 15) | ((int(x) >> 4) << 32) for x in labels]
    mask = (3 | (3 << 32))
    ia, ib, binA, members = {}, {}, [], []
    for k, b in enumerate(b64):
        A, B = b & mask, b & ~mask
        if A not in ia:
            ia[A] = len(binA)
            binA.append(A)
        if B not in ib:
            ib[B] = len(members)
            members.append([])
        members[ib[B]].append((ia[A], k))
    ref = np.zeros((len(binA), len(binA)))
    for v in members:
        for a_, i in v:
            for b_, j in v:
                ref[a_, b_] += psi[i] * psi[j]
    # the sign of the eigenvector is arbitrary: rho is quadratic in it
    np.testing.assert_allclose(rho.real, ref, atol=1e-10)
    coll = [(A & 3) | (((A >> 32) & 3) << 32) for A in binA]
    np.testing.assert_array_equal(st, np.array(coll, np.uint64))
    assert abs(np.trace(rho).real - 1.0) < 1e-10
    assert np.linalg.eigvalsh(rho.real).min() > -1e-12


def test_print_wavefunction(gpu_lib):
    """4-site chain, N = 4, S = 0 (D = 36 <= max_dim_print): the header, and per basis
    state |psi|^2, psi and, for the factorized format 'F', |down,up> with the lowest
    orbital first (the operator formats print |up,down>, the mixed basis' order), against
    the oracle's ground state (printed with six significant digits) and its half bases in
    rank order. Above max_dim_print only the header is printed."""
    from pyqcm_amd import qcm
    m, vals = TM.chain(4, U=4.0, mu=2.0)
    lab = register(with_sites(m, 4), vals, "R0:N4:S0")
    text_e = qcm.print_wavefunction(lab)
    qcm.set_global_parameter("Hamiltonian_format", "F")
    try:
        text = qcm.print_wavefunction(lab)
    finally:
        qcm.set_global_parameter("Hamiltonian_format", "E")
    for le, lf in zip(text_e.strip("\n").split("\n")[2:], text.strip("\n").split("\n")[2:]):
        se, sf = le.split("\t")[2], lf.split("\t")[2]
        dn, up = sf[1:-1].split(",")
        assert se == "|%s,%s>" % (up, dn)
    lines = text.strip("\n").split("\n")
    assert lines[0] == "state"
    sec, e, w = lines[1].split("\t")
    E0, gs, *_ = O.Sector(m, vals, 2, 2).lanczos_gs()
    assert sec == "R0:N4:S0" and abs(float(e) - E0) < 1e-5 * abs(E0) and float(w) == 1.0
    hb = O.halfbasis(4, 2)
    assert len(lines) == 2 + 36
    sign = None
    for i, ln in enumerate(lines[2:]):
        p2, amp, st = ln.split("\t")
        up, dn = hb[i % len(hb)], hb[i // len(hb)]
        bits = lambda v: "".join("1" if (v >> k) & 1 else "0" for k in range(4))
        assert st == "|%s,%s>" % (bits(dn), bits(up))
        assert abs(float(p2) - gs[i] ** 2) < 1e-5
        if abs(gs[i]) > 1e-3:
            s = np.sign(float(amp) * gs[i])
            sign = s if sign is None else sign
            assert s == sign and abs(abs(float(amp)) - abs(gs[i])) < 1e-5
    qcm.set_global_parameter("max_dim_print", 16)
    try:
        lab2 = register(with_sites(m, 4), vals, "R0:N4:S0")
        assert len(qcm.print_wavefunction(lab2).strip("\n").split("\n")) == 2
    finally:
        qcm.set_global_parameter("max_dim_print", 64)


def test_green_function_lookup_table(gpu_lib):
    """Frequencies within 1e-8 of one already evaluated come from the instance's table of
    recent evaluations, as the reference's look-up (model_instance.hpp:566-582,624-632);
    GF_lookup_depth = 0 turns the table off."""
    from pyqcm_amd import qcm
    m, vals = TM.chain(4, U=4.0, mu=2.0)
    z = np.array([0.3j, 1.1j, 0.5 + 0.2j])
    a = register(with_sites(m, 4), vals, "R0:N4:S0")
    G1 = qcm.Green_function(z, 0, a)
    np.testing.assert_array_equal(qcm.Green_function(z + 1e-10, 0, a), G1)
    G_mixed = qcm.Green_function(np.array([2.0j, z[1] + 1e-10]), 0, a)
    np.testing.assert_array_equal(G_mixed[1], G1[1])
    qcm.set_global_parameter("GF_lookup_depth", 0)
    try:
        b = register(with_sites(m, 4), vals, "R0:N4:S0")
        G2 = qcm.Green_function(z, 0, b)
        G3 = qcm.Green_function(z + 1e-10, 0, b)
    finally:
        qcm.set_global_parameter("GF_lookup_depth", 64)
    assert not np.array_equal(G2, G3)
    np.testing.assert_allclose(G2, G3, atol=1e-8)
    np.testing.assert_allclose(G2, G1, atol=1e-10)
    np.testing.assert_allclose(G_mixed[0], qcm.Green_function(np.array([2.0j]), 0, b)[0], atol=1e-10)
