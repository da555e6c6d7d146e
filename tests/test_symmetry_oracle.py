"""The oracle's restatement of the Abelian point-group reduction (oracle/symmetry.py) pinned
by what the mathematics fixes: the irreducible sectors partition the Hilbert space (their
dimensions add up to the sector's), the union of their spectra is the spectrum of the
dense Jordan-Wigner Hamiltonian of the whole sector, each irrep's Hamiltonian is
symmetric, and the group and characters are those of the generators."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import symmetry as Y
from tests import models as TM

# the 2x2 plaquette (sites 0..3 at (0,0), (1,0), (0,1), (1,1)) with the two mirrors of the
# reference's tests/test_diagonalization/S_format.ref (generators [2,1,4,3] and [3,4,1,2],
# 1-based) and bath orbitals labelled by their parity under each mirror (bath_irrep)
MIRRORS = [[1, 0, 3, 2], [2, 3, 0, 1]]


def plaquette4(params):
    m, vals = TM.plaquette_bath(params, n_bath=4)
    # bath x couples with the sign pattern PLAQ_SEQ[x]: its parity under each mirror
    labels = [[0, 0, 0, 0], [0, 0, 2, 2]]
    gens = [MIRRORS[k] + labels[k] for k in range(2)]
    return m, vals, gens


def elements(m, vals):
    one, inter = [], []
    for name, op in m.operators.items():
        z = vals.get(name, 0.0)
        if z == 0.0:
            continue
        for (r, c, v) in op.elements:
            (one if op.kind == "one-body" else inter).append((r, c, z * v))
    return one, inter


def test_group_of_two_mirrors():
    G = Y.SymmetryGroup(8, 4, [MIRRORS[0] + [0, 0, 0, 0], MIRRORS[1] + [0, 0, 2, 2]], bath_irrep=True)
    assert G.g == 4
    # characters: rows orthogonal, real (phi in {0, 2} in units of 2 pi / g)
    chi = np.cos(2 * np.pi * G.phi / G.g)
    np.testing.assert_allclose(chi @ chi.T, G.g * np.eye(G.g), atol=1e-12)
    # the element of index 3 is the product of the two mirrors (the inversion)
    assert G.e[3][:4] == [3, 2, 1, 0]
    assert G.e[3][4:] == [0, 0, 2, 2]


@pytest.mark.parametrize("nup,ndn", [(4, 4), (3, 4), (2, 1)])
def test_irrep_spectra_are_the_sector_spectrum(nup, ndn):
    rng = np.random.default_rng(3)
    p = TM.random_bath_params(rng, n_bath=4)
    m, vals, gens = plaquette4(p)
    G = Y.SymmetryGroup(8, 4, gens, bath_irrep=True)
    one, inter = elements(m, vals)
    full = np.linalg.eigvalsh(O.dense_hamiltonian(m, vals, nup, ndn)[0])
    parts, dims = [], []
    for r in range(G.g):
        B = Y.mixed_basis(G, nup, ndn, r)
        dims.append(len(B))
        if not B:
            continue
        H = Y.mixed_hamiltonian(G, B, r, one, inter)
        assert np.abs(H - H.T).max() == 0.0
        parts.append(np.linalg.eigvalsh(H))
    assert sum(dims) == len(full)
    np.testing.assert_allclose(np.sort(np.concatenate(parts)), full, atol=1e-10)


def test_plain_permutation_group_without_bath_labels():
    """A 4-site chain with its reflection (no bath): a permutation of all orbitals, so the
    fermionic signs of the permutation carry the whole phase."""
    m, vals = TM.chain(4, tp=-0.3, U=3.0, mu=1.1)
    G = Y.SymmetryGroup(4, 4, [[3, 2, 1, 0]], bath_irrep=False)
    assert G.g == 2
    one, inter = elements(m, vals)
    for nup, ndn in [(2, 2), (3, 1), (1, 2)]:
        full = np.linalg.eigvalsh(O.dense_hamiltonian(m, vals, nup, ndn)[0])
        parts = []
        for r in range(2):
            B = Y.mixed_basis(G, nup, ndn, r)
            if B:
                parts.append(np.linalg.eigvalsh(Y.mixed_hamiltonian(G, B, r, one, inter)))
        np.testing.assert_allclose(np.sort(np.concatenate(parts)), full, atol=1e-10)
