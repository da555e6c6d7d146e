"""The bench's 4x4 cluster Hamiltonian, built from the square-lattice tiling of the package,
is the same operator as the oracle's Hubbard cluster with the reference's conventions
(hopping -1 per bond, src_qcm/lattice_model.cpp:968-1019; mu as -n, lattice_model.cpp:286-295;
Hubbard (i, i+L, 1), lattice_model.cpp:929-933)."""
import os
import sys

import numpy as np

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _dense(flat, L):
    t, c, n, r, cc, v = flat
    H = np.zeros((2 * L, 2 * L))
    U = np.zeros((2 * L, 2 * L))
    k = 0
    for i in range(len(t)):
        for _ in range(n[i]):
            if t[i] == 0:
                H[r[k], cc[k]] += c[i] * v[k]
                if r[k] != cc[k]:
                    H[cc[k], r[k]] += c[i] * v[k]
            else:
                U[r[k], cc[k]] += c[i] * v[k]
            k += 1
    return H, U


def test_bench_hamiltonian_matches_oracle_cluster():
    sys.path.insert(0, ROOT)
    import bench
    vals = {"t": 1.0, "tp": -0.3, "U": 8.0, "mu": 4.0}
    H, U = _dense(bench.cluster_hamiltonian_4x4(vals), 16)
    nn, nnn = O.square_bonds(4, 4)
    m, ov = O.hubbard_cluster(16, nn, tp_bonds=nnn, tp=-0.3, U=8.0, mu=4.0)
    Ho, Uo = _dense(m.flat(ov), 16)
    assert np.count_nonzero(H) == 2 * (2 * 24 + 2 * 18) + 32  # NN and NNN bonds both ways, mu
    np.testing.assert_array_equal(H, Ho)
    np.testing.assert_array_equal(U, Uo)
