"""The headline 16-site workload end to end: the 4x4 Hubbard cluster (t=1, t'=-0.3, U=8,
mu=4) on its half-filled N=16, Sz=0 sector (D = 165,636,900).

* Lanczos ground state against the oracle's run from the same start vector
  (tests/golden/c2_lanczos.json, written by tests/golden/make_c2_lanczos.py), with a
  residual check of the returned eigenvector.
* Band-Lanczos Green function with 16 start vectors per excitation kind (N=15 and N=17
  sectors, D = 147,232,800 each): Q-matrix sum rule, symmetry, high-frequency tail.
* The ground state in the symmetry-adapted basis of the two mirrors (irrep 0).
* VCA k-integral on the Legendre grid (20 frequencies) and a 64^2 k-grid: the device
  integrand against the dense restatement on sampled k points, and Omega finite.
"""
import itertools
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import models as TM

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
_names = itertools.count(1)


def c2_values():
    return {"t": 1.0, "tp": -0.3, "U": 8.0, "mu": 4.0}


@pytest.fixture(scope="module")
def c2_instance(gpu_lib):
    from pyqcm_amd import lattice as LT
    M = LT.LatticeModel("c2_4x4_%d" % next(_names), LT.square_tiling(4, 4))
    inst = M.instance(c2_values(), "R0:N16:S0")
    yield M, inst
    inst.erase()


def test_lanczos_16_sites_matches_oracle_fixture(gpu_lib):
    from pyqcm_amd import _lib
    from pyqcm_amd._lib import DeviceArray
    from pyqcm_amd.sector import Sector
    with open(os.path.join(HERE, "golden", "c2_lanczos.json")) as f:
        fx = json.load(f)
    m, vals = TM.square(4, 4, tp=-0.3, U=8.0, mu=4.0)
    S = Sector(m.L, 8, 8, m.flat(vals))
    assert S.dim == fx["dim"]
    _lib.call("qed_reset_start_vector_state")
    E0, psi, n, a, b = S.lanczos_gs(seed=0, accur=fx["accur"], max_iter=600)
    print("E0 %.14f (oracle %.14f), iterations %d (oracle %d)" % (E0, fx["E0"], n, fx["iterations"]))
    assert abs(E0 - fx["E0"]) <= 1e-10 * abs(fx["E0"])
    # the first recurrence coefficients agree to rounding (the products sum in other orders)
    k = len(fx["alpha"])
    np.testing.assert_allclose(a[:k], fx["alpha"], rtol=1e-10)
    np.testing.assert_allclose(b[:k], fx["beta"], rtol=1e-9)
    assert abs(n - fx["iterations"]) <= 3
    # the eigenvector: normalized, and H psi = E0 psi to the Lanczos accuracy
    hp = DeviceArray(S.dim).zeros()
    S.hv_dev(psi, hp)
    g = psi.to_host()
    h = hp.to_host()
    hp.free()
    assert abs(np.dot(g, g) - 1.0) < 1e-10
    r = h - E0 * g
    res2 = float(np.dot(r, r))
    print("residual |H psi - E0 psi|^2 = %.3e" % res2)
    assert res2 <= 1e-10 * E0 * E0


def test_ground_state_16_sites_in_the_mirror_basis(gpu_lib):
    """The same sector in the symmetry-adapted basis of the cluster's two mirrors
    (x -> 3 - x, y -> 3 - y; C2v, 4 irreps): irrep 0 holds D = 41,412,900 of the
    165,636,900 states, and its Lanczos ground state is the fixture's E0 (the ground state
    is fully symmetric), with H psi = E0 psi in that basis."""
    from pyqcm_amd._lib import DeviceArray
    from pyqcm_amd.sector import MixedSector
    with open(os.path.join(HERE, "golden", "c2_lanczos.json")) as f:
        fx = json.load(f)
    m, vals = TM.square(4, 4, tp=-0.3, U=8.0, mu=4.0)
    mx = [(3 - i % 4) + 4 * (i // 4) for i in range(16)]
    my = [i % 4 + 4 * (3 - i // 4) for i in range(16)]
    S = MixedSector(m.L, 8, 8, m.flat(vals), 16, [mx, my], 0, False)
    assert S.dim == 41412900
    E0, psi, n, *_ = S.lanczos_gs(seed=0, accur=1e-12, max_iter=600)
    print("irrep 0: E0 %.14f (fixture %.14f), %d iterations" % (E0, fx["E0"], n))
    assert abs(E0 - fx["E0"]) <= 1e-10 * abs(fx["E0"])
    hp = DeviceArray(S.dim).zeros()
    S.hv_dev(psi, hp)
    g, h = psi.to_host(), hp.to_host()
    hp.free()
    psi.free()
    S.close()
    r = h - E0 * g
    assert abs(np.dot(g, g) - 1.0) < 1e-10
    assert float(np.dot(r, r)) <= 1e-10 * E0 * E0


def test_band_lanczos_green_function_16_sites(c2_instance):
    """16 start vectors c_i|0> (N=15) and c+_i|0> (N=17) per spin-up orbital: the Lehmann
    Q-matrix rows of every orbital pair sum to the anticommutator (sum_k Q_ik Q_jk = delta_ij,
    the reference's check_norm, src_ed/Q_matrix.hpp:235-260), G(iw) is symmetric, and
    i w G(i w) tends to 1 at large w."""
    from pyqcm_amd import qcm
    M, inst = c2_instance
    E0, sec = inst.ground_state()
    print("E0", E0, sec)
    e, Q = qcm.qmatrix(inst.label)
    Q = np.asarray(Q)
    n = M.n
    assert Q.shape[1] == n
    W = Q.T @ Q.conj()
    err = np.abs(W - np.eye(n)).max()
    print("poles", len(e), "sum-rule error %.2e" % err)
    assert err < 1e-5
    for w in (0.1, 1.0, 4.0):
        G = inst.green(np.array([1j * w]))[0]
        assert np.abs(G - G.T).max() < 1e-10
    big = 1.0e6
    G = inst.green(np.array([1j * big]))[0]
    assert np.abs(1j * big * G - np.eye(n)).max() < 1e-3


def test_vca_integrand_16_sites_matches_dense(c2_instance):
    """The device (1/nk) sum_k log|det(V(k) G_c - 1)| on sampled k points against numpy, and
    the full functional on the Legendre (1,10,5,10,5) grid with a 64^2 k-grid."""
    from pyqcm_amd import lattice as LT
    M, inst = c2_instance
    inst.ground_state()
    ws, _ = LT.legendre_frequency_grid(1, 10, 5, 10, 5)
    assert len(ws) == 20
    k = M.tiling.kgrid(64, shift=0.0)
    rng = np.random.default_rng(5)
    pick = rng.choice(len(k), 24, replace=False)
    z = 1j * ws[::4]
    got = np.array([inst.sef_integrand(z, k[[q]]) for q in pick])  # per k point
    Gc = inst.green(z)
    tk = M.tiling.tk(inst.values, k[pick])
    tc = M.tiling.cluster_matrix(inst.values) - inst.v0()
    for i in range(len(pick)):
        for a in range(len(z)):
            ref = O.sef_integrand(Gc[a], tk[i] - tc, factor=1.0)
            assert abs(got[i, a] - ref) <= 1e-10 * max(1.0, abs(ref)), (i, a, got[i, a], ref)
    om = inst.potthoff_functional(grid=(1, 10, 5, 10, 5), kside=64)
    print("Omega", om)
    assert np.isfinite(om)
