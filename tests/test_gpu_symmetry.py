"""Symmetry-adapted (mixed) basis on the device (qed_mixed_sector_create) against the
oracle's restatement of the reference's point-group reduction (oracle/symmetry.py, itself
pinned by the union of the irreps' spectra being the sector's dense spectrum):

* the basis (orbit representatives in each irrep) bit-exact;
* H.v against the oracle's dense mixed-basis Hamiltonian;
* the Lanczos ground state of each irrep against the oracle's lowest eigenvalue, and the
  lowest over the irreps against the factorized sector's;
* the reference's recorded plaquette + 8 bath ground state
  (tests/test_diagonalization/S_format.ref: GS_energy -75.0353 in R0:N12:S0 with the two
  mirrors as generators and bath_irrep) in the irrep 0 sector.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import symmetry as Y
from tests import models as TM

pytestmark = pytest.mark.gpu

MIRRORS = [[1, 0, 3, 2], [2, 3, 0, 1]]


def _elements(m, vals):
    one, inter = [], []
    for name, op in m.operators.items():
        z = vals.get(name, 0.0)
        if z == 0.0:
            continue
        for (r, c, v) in op.elements:
            (one if op.kind == "one-body" else inter).append((r, c, z * v))
    return one, inter


def _plaquette4():
    rng = np.random.default_rng(3)
    p = TM.random_bath_params(rng, n_bath=4)
    m, vals = TM.plaquette_bath(p, n_bath=4)
    gens = [MIRRORS[0] + [0, 0, 0, 0], MIRRORS[1] + [0, 0, 2, 2]]
    return m, vals, 4, gens, True


def _chain4():
    m, vals = TM.chain(4, tp=-0.3, U=3.0, mu=1.1)
    return m, vals, 4, [[3, 2, 1, 0]], False


CASES = {"plaquette4": _plaquette4, "chain4": _chain4}


@pytest.mark.parametrize("case,nup,ndn", [("plaquette4", 4, 4), ("plaquette4", 3, 4), ("plaquette4", 2, 1),
                                          ("chain4", 2, 2), ("chain4", 3, 1)])
def test_mixed_basis_and_hv_match_oracle(gpu_lib, case, nup, ndn):
    from pyqcm_amd.sector import MixedSector
    m, vals, n_sites, gens, bath_irrep = CASES[case]()
    G = Y.SymmetryGroup(m.L, n_sites, gens, bath_irrep=bath_irrep)
    one, inter = _elements(m, vals)
    rng = np.random.default_rng(7)
    for r in range(G.g):
        ref = Y.mixed_basis(G, nup, ndn, r)
        S = MixedSector(m.L, nup, ndn, m.flat(vals), n_sites, gens, r, bath_irrep)
        got = S.basis()
        assert S.dim == len(ref)
        np.testing.assert_array_equal(got, np.array(ref, np.uint64))
        if not ref:
            S.close()
            continue
        H = Y.mixed_hamiltonian(G, ref, r, one, inter)
        x = rng.standard_normal(len(ref))
        y0 = rng.standard_normal(len(ref))
        hy = S.hv(x, y0)
        err = np.abs(hy - (y0 + H @ x)).max() / np.abs(H @ x).max()
        print(case, nup, ndn, "irrep", r, "dim", len(ref), "H.v rel err %.1e" % err)
        assert err < 1e-13
        S.close()


@pytest.mark.parametrize("case,nup,ndn", [("plaquette4", 4, 4), ("chain4", 2, 2)])
def test_mixed_lanczos_per_irrep(gpu_lib, case, nup, ndn):
    from pyqcm_amd.sector import MixedSector, Sector
    m, vals, n_sites, gens, bath_irrep = CASES[case]()
    G = Y.SymmetryGroup(m.L, n_sites, gens, bath_irrep=bath_irrep)
    one, inter = _elements(m, vals)
    lows = []
    for r in range(G.g):
        ref = Y.mixed_basis(G, nup, ndn, r)
        if len(ref) < 2:
            continue
        e_ref = np.linalg.eigvalsh(Y.mixed_hamiltonian(G, ref, r, one, inter))[0]
        S = MixedSector(m.L, nup, ndn, m.flat(vals), n_sites, gens, r, bath_irrep)
        E0 = S.lanczos_gs(accur=1e-12)[0]
        print(case, "irrep", r, "E0 %.12f (oracle %.12f)" % (E0, e_ref))
        assert abs(E0 - e_ref) <= 1e-10 * max(1.0, abs(e_ref))
        lows.append(E0)
        S.close()
    F = Sector(m.L, nup, ndn, m.flat(vals))
    E_full = F.lanczos_gs(accur=1e-12)[0]
    assert abs(min(lows) - E_full) <= 1e-10 * abs(E_full)


def test_trivial_group_is_the_factorized_sector(gpu_lib):
    """No generators: one irrep, every factorized state its own representative; the
    mixed-basis product is the factorized one."""
    from pyqcm_amd.sector import MixedSector, Sector
    m, vals = TM.square(3, 3, tp=-0.3)
    S = MixedSector(m.L, 5, 4, m.flat(vals), m.L, np.zeros((0, m.L), np.int32), 0, False)
    F = Sector(m.L, 5, 4, m.flat(vals))
    assert S.dim == F.dim
    b = S.basis()
    up, dn = O.halfbasis(m.L, 5), O.halfbasis(m.L, 4)
    np.testing.assert_array_equal(b, (np.tile(up.astype(np.uint64), len(dn)) |
                                      (np.repeat(dn.astype(np.uint64), len(up)) << np.uint64(32))))
    rng = np.random.default_rng(2)
    x = rng.standard_normal(F.dim)
    y = rng.standard_normal(F.dim)
    np.testing.assert_allclose(S.hv(x, y), F.hv(x, y), rtol=0, atol=1e-12)


def test_plaquette_8bath_ground_state_in_irrep_0(gpu_lib, golden):
    """tests/test_diagonalization/S_format.ref: new_cluster_model('clus', 4, 8,
    generators=[[2,1,4,3,0,0,0,0,2,2,2,2],[3,4,1,2,0,0,2,2,0,0,2,2]], bath_irrep=True);
    the recorded GS_energy -75.0353 lies in R0:N12:S0 (about a quarter of the 853,776
    factorized states)."""
    from pyqcm_amd.sector import MixedSector, Sector
    g = golden["plaquette_8bath"]
    m, vals = TM.plaquette_bath(g["params"])
    gens = [[1, 0, 3, 2, 0, 0, 0, 0, 2, 2, 2, 2], [2, 3, 0, 1, 0, 0, 2, 2, 0, 0, 2, 2]]
    S = MixedSector(m.L, 6, 6, m.flat(vals), 4, gens, 0, True)
    E0 = S.lanczos_gs(accur=1e-12)[0]
    F = Sector(m.L, 6, 6, m.flat(vals))
    E_full = F.lanczos_gs(accur=1e-12)[0]
    print("irrep 0: dim %d, E0 %.10f; factorized: dim %d, E0 %.10f; recorded %.4f"
          % (S.dim, E0, F.dim, E_full, g["E0"]))
    assert abs(E0 - g["E0"]) < 6e-5
    assert abs(E0 - E_full) <= 1e-10 * abs(E_full)
    assert S.dim < F.dim / 3


@pytest.mark.parametrize("case,nup,ndn", [("plaquette4", 4, 4), ("chain4", 2, 2)])
def test_mixed_green_function_matches_dense(gpu_lib, case, nup, ndn):
    """The Green function of the irrep-0 ground state through symmetric orbitals, the
    creation and destruction operators between irreps and a band Lanczos per orbital
    irrep, turned to the site basis, against the exact G of the dense Jordan-Wigner
    Hamiltonian of the whole sector (the ground state is unique and in irrep 0 here)."""
    from pyqcm_amd import symmetry as SY
    m, vals, n_sites, gens, bath_irrep = CASES[case]()
    zs = np.array([0.3j, 1.0j, 2.5j, 0.7 + 0.2j])
    E0, G = SY.green_function(m.L, n_sites, m.flat(vals), gens, bath_irrep, nup, ndn, 0, zs)
    E0d, Gd = O.dense_gf(m, vals, nup, ndn, list(range(n_sites)), zs)
    print(case, "E0 %.12f dense %.12f" % (E0, E0d), "max |dG| %.2e" % np.abs(G - Gd).max())
    assert abs(E0 - E0d) <= 1e-10 * abs(E0d)
    np.testing.assert_allclose(G, Gd, atol=1e-9)


def test_mixed_lanczos_4x4_quarter_filling(gpu_lib):
    """The 4x4 cluster at N = 8 (N_up = N_down = 4, 3.3 M factorized states) in the
    two-mirror basis: the irreps' dimensions add up to the factorized sector's, and the
    lowest Lanczos E0 over the irreps is the factorized sector's."""
    from pyqcm_amd.sector import MixedSector, Sector
    m, vals = TM.square(4, 4, tp=-0.3, U=8.0, mu=2.0)
    mx = [(3 - i % 4) + 4 * (i // 4) for i in range(16)]
    my = [i % 4 + 4 * (3 - i // 4) for i in range(16)]
    lows, dims = [], []
    for r in range(4):
        S = MixedSector(m.L, 4, 4, m.flat(vals), 16, [mx, my], r, False)
        dims.append(S.dim)
        lows.append(S.lanczos_gs(accur=1e-12)[0])
        S.close()
    F = Sector(m.L, 4, 4, m.flat(vals))
    E_full = F.lanczos_gs(accur=1e-12)[0]
    print("irrep dims", dims, "of", F.dim, "E0 per irrep", lows, "factorized", E_full)
    # an orbit of length l survives in the l irreps that are trivial on its stabilizer
    assert sum(dims) == F.dim
    assert abs(min(lows) - E_full) <= 1e-10 * abs(E_full)


def _davidson(gpu_lib, S, K, accur=1e-5):
    import ctypes
    from pyqcm_amd._lib import DeviceArray
    psi = DeviceArray(K * S.dim).zeros()
    ev = np.zeros(K)
    nit = ctypes.c_int()
    gpu_lib.call("qed_davidson", S.h, K, accur, 1000, gpu_lib.dp(ev), psi.ptr, ctypes.byref(nit))
    V = psi.to_host().reshape(K, S.dim)
    psi.free()
    return ev, V, nit.value


@pytest.mark.parametrize("case,nup,ndn", [("plaquette4", 4, 4), ("chain4", 2, 2)])
def test_mixed_diagonal_and_davidson(gpu_lib, case, nup, ndn):
    """The diagonal of H in the symmetry-adapted basis (the factorized state's diagonal plus
    the hops that stay inside its own orbit, which the plaquette's mirrors produce) against
    the dense oracle matrix, and Davidson-Liu's 4 lowest states of each irrep (the true
    diagonal as its preconditioner) against the dense spectrum: values, residuals
    <= accur_Davidson, orthonormal vectors."""
    from pyqcm_amd.sector import MixedSector
    m, vals, n_sites, gens, bath_irrep = CASES[case]()
    G = Y.SymmetryGroup(m.L, n_sites, gens, bath_irrep=bath_irrep)
    one, inter = _elements(m, vals)
    K = 4
    for r in range(G.g):
        ref = Y.mixed_basis(G, nup, ndn, r)
        if len(ref) < 2 * K:
            continue
        H = Y.mixed_hamiltonian(G, ref, r, one, inter)
        S = MixedSector(m.L, nup, ndn, m.flat(vals), n_sites, gens, r, bath_irrep)
        d = S.diag()
        np.testing.assert_allclose(d, np.diag(H), rtol=0, atol=1e-12 * np.abs(H).max())
        ev, V, nit = _davidson(gpu_lib, S, K)
        e = np.linalg.eigvalsh(H)[:K]
        print(case, "irrep", r, "dim", len(ref), "Davidson %d iterations" % nit, ev, "dense", e)
        np.testing.assert_allclose(ev, e, rtol=0, atol=1e-8)
        np.testing.assert_allclose(V @ V.T, np.eye(K), atol=1e-10)
        for k in range(K):
            assert np.linalg.norm(H @ V[k] - ev[k] * V[k]) <= 1.1e-5
        S.close()
