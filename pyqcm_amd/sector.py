"""Sector-level Python view of the device Hamiltonian (the reference's Hamiltonian<T>
seam: src_ed/Hamiltonian/Hamiltonian_base.hpp:43-82)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import DeviceArray, call, dp, ip


def flatten_ops(ops, values):
    """ops: iterable of (name, kind, elements) with 0-based spin-orbital elements (r <= c);
    values: {name: coefficient}. Returns the flat arrays qed_sector_create expects."""
    types, coefs, nels, er, ec, ev = [], [], [], [], [], []
    for name, kind, elements in ops:
        if name not in values or values[name] == 0.0:
            continue
        if kind not in ("one-body", "interaction", "anomalous"):
            raise ValueError("operator type %r is not supported" % kind)
        types.append({"one-body": 0, "interaction": 1, "anomalous": 2}[kind])
        coefs.append(float(values[name]))
        nels.append(len(elements))
        for (r, c, v) in elements:
            er.append(int(r))
            ec.append(int(c))
            ev.append(float(v))
    return (np.array(types, np.int32), np.array(coefs, np.float64), np.array(nels, np.int32),
            np.array(er, np.int32), np.array(ec, np.int32), np.array(ev, np.float64))


class Sector:
    """H on one factorized sector (n_up, n_dn), resident on the GPU."""

    def __init__(self, L, nup, ndn, flat):
        self.L, self.nup, self.ndn = L, nup, ndn
        t, c, n, r, cc, v = flat
        self._flat = flat
        h = ctypes.c_void_p()
        call("qed_sector_create", L, nup, ndn, len(t), ip(t), dp(c), ip(n), ip(r), ip(cc), dp(v),
             ctypes.byref(h))
        self.h = h
        dim, du, dd = ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_uint32()
        call("qed_sector_dims", h, ctypes.byref(dim), ctypes.byref(du), ctypes.byref(dd))
        self.dim, self.du, self.dd = dim.value, du.value, dd.value

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            call("qed_sector_destroy", self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- H.v ----------------------------------------------------------------------------
    def hv_dev(self, x: DeviceArray, y: DeviceArray, variant=None):
        if variant is None:
            call("qed_hv", self.h, x.ptr, y.ptr)
        else:
            call("qed_hv_variant", self.h, int(variant), x.ptr, y.ptr)

    def hv(self, x, y=None, variant=None):
        x = np.ascontiguousarray(x, np.float64)
        dx = DeviceArray.from_host(x)
        dy = DeviceArray.from_host(np.zeros(self.dim) if y is None else y)
        self.hv_dev(dx, dy, variant)
        return dy.to_host()

    def diag(self):
        d = DeviceArray(self.dim).zeros()
        call("qed_diag", self.h, d.ptr)
        return d.to_host()

    # --- solvers -----------------------------------------------------------------------
    def start_vector(self, seed=0):
        x = DeviceArray(self.dim)
        call("qed_start_vector", self.h, int(seed), x.ptr)
        return x

    def lanczos_gs(self, seed=0, accur=1e-12, max_iter=1000, start=None):
        psi = self.start_vector(seed) if start is None else DeviceArray.from_host(start)
        E0 = np.zeros(1)
        nit = np.zeros(1, np.int32)
        a = np.zeros(max_iter + 2)
        b = np.zeros(max_iter + 2)
        call("qed_lanczos_gs", self.h, float(accur), int(max_iter), psi.ptr, dp(E0), ip(nit), dp(a), dp(b))
        n = int(nit[0])
        return float(E0[0]), psi, n, a[:n].copy(), b[:n].copy()

    def lanczos_green(self, psi: DeviceArray, accur_cf=0.01, max_iter=400):
        a = np.zeros(max_iter + 2)
        b = np.zeros(max_iter + 3)
        n = np.zeros(1, np.int32)
        call("qed_lanczos_green", self.h, psi.ptr, float(accur_cf), int(max_iter), dp(a), dp(b), ip(n))
        k = int(n[0])
        return a[:k].copy(), b[:k + 1].copy()

    def band_lanczos(self, phi: DeviceArray, p, M0=0, max_iter_BL=600, accur_deflation=1e-7,
                     accur_band_lanczos=1e-12):
        cap = max_iter_BL + 2 * p + 2
        e = np.zeros(cap)
        Q = np.zeros(cap * p)
        M = np.zeros(1, np.int32)
        call("qed_band_lanczos", self.h, phi.ptr, int(p), int(M0), int(max_iter_BL),
             float(accur_deflation), float(accur_band_lanczos), dp(e), dp(Q), int(cap), ip(M))
        m = int(M[0])
        return e[:m].copy(), Q[:m * p].reshape(m, p).T.copy()


class MixedSector(Sector):
    """H on one sector of the symmetry-adapted basis: orbit representatives of an Abelian
    point group in one irreducible representation (qed_mixed_sector_create). generators:
    lists of L 0-based entries (site images; with bath_irrep, bath phase labels)."""

    def __init__(self, L, nup, ndn, flat, n_sites, generators, irrep, bath_irrep=False):
        self.L, self.nup, self.ndn = L, nup, ndn
        t, c, n, r, cc, v = flat
        self._flat = flat
        g = np.ascontiguousarray(np.asarray(generators, np.int32).reshape(-1, L))
        h = ctypes.c_void_p()
        call("qed_mixed_sector_create", L, nup, ndn, len(t), ip(t), dp(c), ip(n), ip(r), ip(cc), dp(v),
             int(n_sites), len(g), ip(g.ravel()), int(bool(bath_irrep)), int(irrep), ctypes.byref(h))
        self.h = h
        dim, du, dd = ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_uint32()
        call("qed_sector_dims", h, ctypes.byref(dim), ctypes.byref(du), ctypes.byref(dd))
        self.dim, self.du, self.dd = dim.value, du.value, dd.value
        self.irrep = irrep

    def basis(self):
        """the representatives as 64-bit states (up | down << 32), ascending"""
        n = ctypes.c_uint64()
        call("qed_mixed_basis", self.h, None, ctypes.byref(n))
        out = np.zeros(n.value, np.uint64)
        call("qed_mixed_basis", self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(n))
        return out


class GeneralSector(Sector):
    """H on a sector that conserves N or S but not both (spin-flip and anomalous mixings,
    trivial group): the union of the (N_up, N_down) blocks in pairs, 64-bit states
    ascending (qed_general_sector_create). flat: flatten_ops output; operator type 2 is
    anomalous."""

    def __init__(self, L, pairs, flat):
        self.L = L
        self.nup = self.ndn = -1
        t, c, n, r, cc, v = flat
        self._flat = flat
        pr = np.ascontiguousarray(np.asarray(pairs, np.int32).reshape(-1, 2))
        h = ctypes.c_void_p()
        call("qed_general_sector_create", L, len(pr), ip(pr.ravel()), len(t), ip(t), dp(c), ip(n), ip(r), ip(cc),
             dp(v), ctypes.byref(h))
        self.h = h
        dim = ctypes.c_uint64()
        call("qed_sector_dims", h, ctypes.byref(dim), None, None)
        self.dim, self.du, self.dd = dim.value, 0, 0

    def basis(self):
        """the basis as 64-bit states (up | down << 32), ascending"""
        n = ctypes.c_uint64()
        call("qed_general_basis", self.h, None, ctypes.byref(n))
        out = np.zeros(n.value, np.uint64)
        call("qed_general_basis", self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(n))
        return out

    def apply(self, target, orbital, pm, x, y, z=1.0):
        """y += z c_a x (pm = -1) or z c+_a x (pm = +1), x in this sector, y in target's
        (DeviceArray or device address)"""
        call("qed_general_apply", self.h, target.h, int(orbital), int(pm), float(z), getattr(x, "ptr", x),
             getattr(y, "ptr", y))


def apply_c(L, nup, ndn, a, pm, x: DeviceArray, y: DeviceArray, z=1.0):
    """y += z c_a x (pm=-1) or z c†_a x (pm=+1), device vectors."""
    call("qed_apply_c", int(L), int(nup), int(ndn), int(a), int(pm), float(z), x.ptr, y.ptr)


def halfbasis(L, N):
    dim = ctypes.c_uint32()
    call("qed_halfbasis", L, N, None, ctypes.byref(dim))
    out = np.zeros(dim.value, np.uint32)
    if dim.value:
        call("qed_halfbasis", L, N, out.ctypes.data_as(_lib.u32ptr), ctypes.byref(dim))
    return out


def halfbasis_rank(L, N, states):
    states = np.ascontiguousarray(states, np.uint32)
    ranks = np.zeros(len(states), np.uint32)
    if len(states):
        call("qed_halfbasis_rank", L, N, states.ctypes.data_as(_lib.u32ptr), len(states),
             ranks.ctypes.data_as(_lib.u32ptr))
    return ranks
