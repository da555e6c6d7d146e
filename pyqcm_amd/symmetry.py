"""Green function of a symmetry-adapted ground state (a sector of the Abelian point-group
basis, `sector.MixedSector`), in the site basis.

As the reference does for a model with generators (src_ed/model_instance.hpp:710-817 with
the destruction operators of src_ed/Operators/destruction_operator.cpp and
symmetry_group::to_site_basis, src_ed/symmetry_group.cpp:581-608): the sites are combined
into symmetric orbitals (one set per irrep, src_ed/symmetry_group.cpp:300-323); for each
irrep r_o of orbitals, c_o|0> and c^dagger_o|0> live in the sectors of irrep
r (x) r_o with one particle less or more; a band Lanczos there gives the Lehmann form of
that irrep's block of G; G in the site basis is S^T G_sym S. Spin up (the spin-down
function of an S = 0 state is the same)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import DeviceArray, call
from .sector import MixedSector


def group_of(S: MixedSector):
    """(g, elements (g x L), phi (g x g)) of a symmetry-adapted sector's group"""
    g = ctypes.c_int()
    call("qed_mixed_group", S.h, ctypes.byref(g), None, None)
    el = np.zeros(g.value * S.L, np.int32)
    phi = np.zeros(g.value * g.value, np.int32)
    call("qed_mixed_group", S.h, ctypes.byref(g), _lib.ip(el), _lib.ip(phi))
    return g.value, el.reshape(g.value, S.L), phi.reshape(g.value, g.value)


def symmetric_orbitals(g, elements, phi, n_sites, tiny=1e-12):
    """Rows of S (n_sites x n_sites) and the irrep of each row: for each irrep, each orbit
    of sites gives sum_i chi(r, e_i) |e_i(site)>, kept if it does not vanish, normalized
    (src_ed/symmetry_group.cpp:300-323; real characters)."""
    rows, irreps = [], []
    for r in range(g):
        skip = np.zeros(n_sites, bool)
        for site in range(n_sites):
            if skip[site]:
                continue
            skip[site] = True
            orb = np.zeros(n_sites)
            for i in range(g):
                img = int(elements[i][site])
                skip[img] = True
                orb[img] += np.cos(2 * np.pi * phi[r, i] / g)
            nrm = np.linalg.norm(orb)
            if nrm > tiny:
                rows.append(orb / nrm)
                irreps.append(r)
    if len(rows) != n_sites:
        raise RuntimeError("Problem with construction of symmetric orbitals: %d found instead of %d"
                           % (len(rows), n_sites))
    return np.array(rows), np.array(irreps)


def irrep_product(g, phi, r1, r2):
    """the irrep whose characters are the products of r1's and r2's"""
    want = (phi[r1] + phi[r2]) % g
    for r in range(g):
        if np.array_equal(phi[r] % g, want):
            return r
    raise RuntimeError("no irrep product")


def green_function(L, n_sites, flat, generators, bath_irrep, nup, ndn, irrep, zs, accur=1e-12, max_iter=1000,
                   seed=0, max_iter_BL=600, accur_deflation=1e-7, accur_band_lanczos=1e-12):
    """(E0, G): the ground state of sector (nup, ndn, irrep) and its spin-up Green function
    G[z] (n_sites x n_sites) at the complex frequencies zs, in the site basis."""
    zs = np.atleast_1d(np.asarray(zs, complex))
    B = MixedSector(L, nup, ndn, flat, n_sites, generators, irrep, bath_irrep)
    E0, psi, *_ = B.lanczos_gs(seed=seed, accur=accur, max_iter=max_iter)
    g, el, phi = group_of(B)
    Srows, orb_irrep = symmetric_orbitals(g, el, phi, n_sites)
    G = np.zeros((len(zs), n_sites, n_sites), complex)
    for ro in range(g):
        sel = np.where(orb_irrep == ro)[0]
        p = len(sel)
        if p == 0:
            continue
        rt = irrep_product(g, phi, irrep, ro)
        Gs = np.zeros((len(zs), p, p), complex)
        for pm in (-1, 1):
            if not (0 <= nup + pm <= L):
                continue
            T = MixedSector(L, nup + pm, ndn, flat, n_sites, generators, rt, bath_irrep)
            if T.dim == 0:
                T.close()
                continue
            phis = DeviceArray(p * T.dim).zeros()
            for k, o in enumerate(sel):
                coef = np.ascontiguousarray(Srows[o], np.float64)
                call("qed_mixed_apply_c", B.h, T.h, 0, pm, n_sites, _lib.dp(coef), psi.ptr,
                     phis.offset(k * T.dim))
            e, Q = T.band_lanczos(phis, p, max_iter_BL=max_iter_BL, accur_deflation=accur_deflation,
                                  accur_band_lanczos=accur_band_lanczos)
            phis.free()
            T.close()
            e = e - E0
            if pm < 0:
                e = -e
            for i, z in enumerate(zs):
                Gs[i] += (Q / (z - e)) @ Q.T
        S = Srows[sel]
        for i in range(len(zs)):
            G[i] += S.T @ Gs[i] @ S
    psi.free()
    B.close()
    return E0, G
