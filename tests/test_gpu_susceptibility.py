"""Dynamic susceptibility of an operator (ED::susceptibility, susceptibility_poles;
src_ed/model_instance.hpp:1024-1113) against brute force: dense sector Hamiltonian and
operator matrices, chi(w) = sum_n |<n|O|psi0>|^2 / (w - (E_n - E0))."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import models as TM
from tests.test_gpu_ed import register, with_sites

pytestmark = pytest.mark.gpu

WS = np.array([0.3 + 0.1j, 1.5 + 0.05j, -2.0 + 0.2j, 4.0j, 7.0 + 0.5j])


def _dense(m, vals, nup, ndn):
    S = O.Sector(m, vals, nup, ndn)
    eye = np.eye(S.dim)
    return np.array([S.hv(eye[k]) for k in range(S.dim)]).T


def _lehmann(m, vals, op, nup, ndn):
    e, U = np.linalg.eigh(_dense(m, vals, nup, ndn))
    b = _dense(m, {op: 1.0}, nup, ndn) @ U[:, 0]
    return e - e[0], U.T @ b


@pytest.mark.parametrize("op", ["t", "U"])
def test_susceptibility_full_diagonalization_is_exact(gpu_lib, op):
    """4-site chain, half filling (D = 36 < max_dim_full: full diagonalization), kinetic and
    double-occupancy operators: equal to the brute-force sum."""
    from pyqcm_amd import qcm
    m, vals = TM.chain(4, U=4.0, mu=2.0, tp=-0.2)
    lab = register(with_sites(m, 4), vals, "R0:N4:S0")
    chi = qcm.susceptibility(op, WS, lab)
    de, Q = _lehmann(m, vals, op, 2, 2)
    ref = np.array([np.sum(Q ** 2 / (w - de)) for w in WS])
    err = np.abs(chi - ref).max() / np.abs(ref).max()
    print("op=%s: max relative error %.3g" % (op, err))
    assert err < 1e-12


@pytest.mark.parametrize("op", ["t", "U"])
def test_susceptibility_band_lanczos_matches_oracle(gpu_lib, op):
    """6-site chain, half filling (D = 400: band Lanczos with one start vector). The
    recurrence stops once its lowest pole has converged (accur_band_lanczos), as in the
    reference, so the Lehmann form is truncated: against the oracle's restatement of the
    same steps to 1e-9, against the exact sum only to the truncation (1e-2)."""
    from pyqcm_amd import qcm
    m, vals = TM.chain(6, U=4.0, mu=2.0, tp=-0.2)
    lab = register(with_sites(m, 6), vals, "R0:N6:S0")
    chi = qcm.susceptibility(op, WS, lab)
    S = O.Sector(m, vals, 3, 3)
    E0, gs, *_ = S.lanczos_gs()
    b = O.Sector(m, {op: 1.0}, 3, 3).hv(gs)
    e, Q = O.band_lanczos(S.hv, [b])
    ref = np.array([np.sum(Q[0] ** 2 / (w - (e - E0))) for w in WS])
    err = np.abs(chi - ref).max() / np.abs(ref).max()
    de, Qx = _lehmann(m, vals, op, 3, 3)
    exact = np.array([np.sum(Qx ** 2 / (w - de)) for w in WS])
    err_x = np.abs(chi - exact).max() / np.abs(exact).max()
    print("op=%s: vs oracle %.3g, vs exact %.3g (%d poles of %d)" % (op, err, err_x, len(e), S.dim))
    assert err < 1e-9
    assert err_x < 1e-2


def test_susceptibility_poles_match_bruteforce(gpu_lib):
    """Poles above 1e-8, residues above 1e-6, poles closer than 1e-4 merged (the level
    degeneracies of the chain), against the same rules applied to the dense spectrum."""
    from pyqcm_amd import qcm
    m, vals = TM.chain(4, U=4.0, mu=2.0)
    lab = register(with_sites(m, 4), vals, "R0:N4:S0")
    got = qcm.susceptibility_poles("t", lab)
    de, Q = _lehmann(m, vals, "t", 2, 2)
    ref = []
    for e, q in zip(de, Q):
        if e < 1e-8 or q * q <= 1e-6:
            continue
        for p in ref:
            if abs(p[0] - e) < 1e-4:
                p[1] += q * q
                break
        else:
            ref.append([e, q * q])
    ref = np.array(sorted(ref))
    got = got[np.argsort(got[:, 0])]
    assert got.shape == ref.shape
    np.testing.assert_allclose(got, ref, atol=1e-10)
    # the sum rule: all residues (pole 0 excluded) = <b|b> - <psi|b>^2
    assert abs(got[:, 1].sum() - (np.sum(Q ** 2) - Q[0] ** 2)) < 1e-8


def test_susceptibility_errors(gpu_lib):
    from pyqcm_amd import qcm
    m, vals = TM.chain(4, U=4.0, mu=2.0)
    lab = register(with_sites(m, 4), {"t": 1.0, "U": 4.0}, "R0:N2:S0")
    with pytest.raises(Exception, match="Operator nope is not defined in the model."):
        qcm.susceptibility("nope", WS, lab)
    with pytest.raises(Exception, match="operator mu is not activated in instance"):
        qcm.susceptibility_poles("mu", lab)


def test_susceptibility_of_symmetric_and_nambu_instances(gpu_lib):
    """The operator applied within the state's own sector (HS_operator.at(sec)): on the
    chain's mirror-symmetric basis (irrep of the ground state; t and U are symmetric, so
    O|psi> stays in it) the result equals the trivial group's brute-force sum; on the
    Nambu sector S = 0 of a paired chain, for the pairing operator and U, it equals the
    sum over the dense Jordan-Wigner Hamiltonian of the same states. Both sectors are
    below max_dim_full (exact Lehmann forms)."""
    from pyqcm_amd import qcm
    from tests import test_gpu_mixing_instance as MI
    from tests import test_gpu_symmetry_instance as SI
    m, vals, ns, gens, bi = SI._chain4()
    r, _ = SI._lowest_irrep(m, vals, ns, gens, bi, 4, 0)
    lab = SI.register(m, vals, "R%d:N4:S0" % r, ns, gens, bi)
    for op in ("t", "U"):
        chi = qcm.susceptibility(op, WS, lab)
        de, Q = _lehmann(m, vals, op, 2, 2)
        ref = np.array([np.sum(Q ** 2 / (w - de)) for w in WS])
        print("mirror basis", op, "max relative error %.2e" % (np.abs(chi - ref).max() / np.abs(ref).max()))
        assert np.abs(chi - ref).max() <= 1e-12 * np.abs(ref).max()
    m, vals, n = MI.chain(d=0.5, s=0.2)
    lab = MI.register(m, vals, "R0:S0", n)
    states = O.block_states(m.L, MI.blocks(m.L, S=0))
    e, U = np.linalg.eigh(O.dense_hamiltonian(m, vals, 0, 0, states=states)[0])
    for op in ("D", "U"):
        chi = qcm.susceptibility(op, WS, lab)
        one = O.ClusterModel(m.L)
        one.add(op, m.operators[op].kind, m.operators[op].elements)
        Q = U.T @ (O.dense_hamiltonian(one, {op: 1.0}, 0, 0, states=states)[0] @ U[:, 0])
        ref = np.array([np.sum(Q ** 2 / (w - (e - e[0]))) for w in WS])
        print("Nambu", op, "max relative error %.2e" % (np.abs(chi - ref).max() / np.abs(ref).max()))
        assert np.abs(chi - ref).max() <= 1e-12 * np.abs(ref).max()
