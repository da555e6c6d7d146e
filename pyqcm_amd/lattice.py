"""Lattice side of the quantum cluster methods on this path (CPT, VCA, CDMFT).

A cluster of sites tiles a lattice through a superlattice; lattice operators (hoppings,
on-site terms, density waves) decompose into an intra-cluster part and inter-cluster
Fourier terms T_n, t(k) = sum_n T_n exp(2 pi i k.n), with k in the superdual basis and n
the integer superlattice translation of the second site (src_qcm/CPT.cpp:22-64 set_V).

For one cluster instance (ED model with bath, solved through ``qcm``):

  V(k, w) = t(k) - t_c - Gamma_c(w)                        (src_qcm/CPT.cpp:57-61)
  G_cpt   = G_c (1 - V G_c)^-1                             (src_qcm/CPT.cpp:76-86)

The frequency-dependent Gamma_c is folded into an effective cluster function
G_eff = (G_c^-1 + Gamma_c)^-1, so that the device kernels see a k-dependence only through
V(k) = V0 + sum_n T_n e^{2 pi i k.n} with V0 = T_0(lattice) - t_c. Identities used:
G_eff^-1 - (t(k) - t_c) = G_c^-1 - V(k, w), and
log|det(1 - V G_c)| = log|det(1 - (V+Gamma) G_eff)| + log|det G_c| - log|det G_eff|.

All k-sums run on the device (qed_cpt_batch, qed_sef_batch, qed_cdmft_distance_batch);
this module only assembles the small per-frequency matrices.
"""
from __future__ import annotations

import itertools

import numpy as np

from . import _lib, qcm


def _cm(a):
    """complex (..., n, n) row-major array -> flat float64 buffer, column-major per matrix."""
    a = np.ascontiguousarray(np.swapaxes(np.asarray(a, np.complex128), -1, -2))
    return a.view(np.float64).ravel()


def _from_cm(buf, shape):
    n = shape[-1]
    a = buf.view(np.complex128).reshape(*shape[:-2], n, n)
    return np.swapaxes(a, -1, -2).copy()


# ------------------------------------------------------------------------- tiling

class Tiling:
    """Cluster sites (integer positions) repeated by a superlattice of dimension 1-3.

    Operators are registered as real one-body matrices per spin (spin up and down equal):
      hop(name, link, amplitude)   amplitude on c+_i c_j and c+_j c_i, j = i + link
                                   (pyqcm hopping_operator);
      onsite(name, amplitude)      amplitude on every site (mu uses -1);
      density_wave(name, Q)        cos(pi Q.r_i) on site i, both spins
                                   (src_qcm/lattice_model.cpp:1149-1215, type 'N');
    and as real pairing matrices P^n, P^n_ij (c_{i up, 0} c_{j down, n} + h.c.):
      pairing(name, link, amplitude)  singlet pairing, amplitude on
                                   c_i up c_j down - c_i down c_j up + h.c., j = i + link,
                                   and amplitude on c_i up c_i down + h.c. for link 0
                                   (pyqcm anomalous_operator, type 'singlet',
                                   src_qcm/lattice_model.cpp:308-341,1101-1133).
    and as real spin-flip matrices F^n, F^n_ij (c+_{i up, 0} c_{j down, n} + h.c.):
      spin_flip(name, amplitude)   amplitude on c+_i up c_i down + h.c. on every site (pyqcm
                                   hopping_operator with tau = 0, sigma = 1: a field along x).
    Operators with pairing terms need the Nambu spinor (c_up(k), c+_down(-k)), spin-flip
    operators the spinor (c_up(k), c_down(k)) (mixing 2).
    """

    def __init__(self, sites, superlattice):
        self.sites = np.array(sites, np.int64).reshape(-1, 3)
        self.n = len(self.sites)
        S = np.array(superlattice, np.float64).reshape(-1, 3)
        self.dim = S.shape[0]
        full = np.eye(3)
        full[: self.dim] = S
        # complete the basis with unit vectors independent of the superlattice
        k = self.dim
        for e in np.eye(3):
            if k == 3:
                break
            trial = full.copy()
            trial[k] = e
            if abs(np.linalg.det(trial)) > 1e-9:
                full = trial
                k += 1
        self.S = full
        self.Sinv = np.linalg.inv(full.T)
        self.terms = {}  # name -> {n (tuple of dim ints): n x n real matrix}
        self.pairs = {}  # name -> {n: n x n real pairing matrix P^n}
        self.flips = {}  # name -> {n: n x n real spin-flip matrix F^n}
        self.n_orb = self.n

    def _zero(self):
        return tuple([0] * self.dim)

    def _locate(self, pos):
        for j, r in enumerate(self.sites):
            c = self.Sinv @ (np.asarray(pos, float) - r)
            cr = np.round(c)
            if np.allclose(c, cr, atol=1e-9) and np.all(cr[self.dim:] == 0):
                return j, tuple(int(x) for x in cr[: self.dim])
        return None

    def _mat(self, name, nvec):
        d = self.terms.setdefault(name, {})
        if nvec not in d:
            d[nvec] = np.zeros((self.n, self.n))
        return d[nvec]

    def hop(self, name, link, amplitude=-1.0):
        link = np.array(list(link) + [0] * (3 - len(link)), np.int64)
        for i, r in enumerate(self.sites):
            loc = self._locate(r + link)
            if loc is None:
                raise ValueError("site %d + link %s is not a lattice site" % (i, link))
            j, nvec = loc
            neg = tuple(-x for x in nvec)
            self._mat(name, nvec)[i, j] += amplitude
            self._mat(name, neg)[j, i] += amplitude
        return self

    def pairing(self, name, link, amplitude=1.0):
        link = np.array(list(link) + [0] * (3 - len(link)), np.int64)
        d = self.pairs.setdefault(name, {})

        def mat(nvec):
            if nvec not in d:
                d[nvec] = np.zeros((self.n, self.n))
            return d[nvec]
        for i, r in enumerate(self.sites):
            loc = self._locate(r + link)
            if loc is None:
                raise ValueError("site %d + link %s is not a lattice site" % (i, link))
            j, nvec = loc
            mat(nvec)[i, j] += amplitude
            if np.any(link):
                # c_i up c_j down - c_i down c_j up = c_i up c_j down + c_j up c_i down
                mat(tuple(-x for x in nvec))[j, i] += amplitude
        return self

    def spin_flip(self, name, amplitude=1.0):
        d = self.flips.setdefault(name, {})
        z = self._zero()
        if z not in d:
            d[z] = np.zeros((self.n, self.n))
        d[z][np.diag_indices(self.n)] += amplitude
        return self

    def onsite(self, name, amplitude=1.0):
        M = self._mat(name, self._zero())
        M[np.diag_indices(self.n)] += amplitude
        return self

    def density_wave(self, name, Q, amplitude=1.0, phase=0.0):
        Q = np.array(list(Q) + [0] * (3 - len(Q)), float)
        M = self._mat(name, self._zero())
        for i, r in enumerate(self.sites):
            M[i, i] += amplitude * np.cos(np.pi * (Q @ r) + np.pi * phase)
        return self

    # ---- decompositions

    def spinor_terms(self, values, mixing):
        """The lattice terms in the spinor of a mixing: 1 Nambu (nambu_terms), 2 spin flip
        (spin_flip_terms)."""
        return self.nambu_terms(values) if mixing == 1 else self.spin_flip_terms(values)

    def cluster_matrix(self, values, nambu=False, mixing=None):
        """Intra-cluster one-body matrix T_0 = sum_op value_op T_0^op (with nambu, or a
        spinor mixing 1 or 2, the 2n x 2n spinor matrix of the translation n = 0)."""
        mixing = (1 if nambu else 0) if mixing is None else mixing
        if mixing:
            return self.spinor_terms(values, mixing).get(self._zero(), np.zeros((2 * self.n, 2 * self.n))).real.copy()
        T0 = np.zeros((self.n, self.n))
        for name, d in self.terms.items():
            z = values.get(name, 0.0)
            if z and self._zero() in d:
                T0 += z * d[self._zero()]
        return T0

    def nambu_terms(self, values):
        """{n: 2n x 2n matrix} of the lattice Hamiltonian in the Nambu spinor
        Psi_k = (c_up(k), c+_down(-k)), h(k) = sum_n H^n exp(2 pi i k.n) with
          H^n = [[T^n, -P^n], [-(P^-n)^T, -(T^-n)^T]]
        (the hole block from c+_down c_down = 1 - c_down c+_down; the pairing blocks from
        v c_up c_down = -v c_down c_up), the convention of the cluster's spinor matrices
        (src_ed/Operators/anomalous_operator.hpp:241-284)."""
        n = self.n
        T, P = {}, {}
        for src, dst in ((self.terms, T), (self.pairs, P)):
            for name, d in src.items():
                z = values.get(name, 0.0)
                if not z:
                    continue
                for nvec, M in d.items():
                    dst[nvec] = dst.get(nvec, 0.0) + z * M
        out = {}
        for nvec in set(T) | set(P) | {tuple(-x for x in v) for v in list(T) + list(P)}:
            neg = tuple(-x for x in nvec)
            H = np.zeros((2 * n, 2 * n))
            if nvec in T:
                H[:n, :n] = T[nvec]
            if neg in T:
                H[n:, n:] = -np.asarray(T[neg]).T
            if nvec in P:
                H[:n, n:] = -P[nvec]
            if neg in P:
                H[n:, :n] = -np.asarray(P[neg]).T
            if np.any(H):
                out[nvec] = H
        return out

    def cluster_operator_elements(self, name):
        """Intra-cluster part of operator `name` as 1-based (r, c, v), r <= c, over both
        spins (site i -> label i+1 for spin up, i+1+n_orb for spin down), the form
        qcm.new_operator takes; for a pairing operator, the anomalous elements
        (i up, j down, P^0_ij), each standing for P^0_ij (c_i up c_j down + h.c.)."""
        if name in self.flips:
            F0 = self.flips[name].get(self._zero())
            if F0 is None:
                return []
            return [(i + 1, j + 1 + self.n_orb, float(F0[i, j])) for i in range(self.n) for j in range(self.n)
                    if F0[i, j] != 0.0]
        if name in self.pairs:
            P0 = self.pairs[name].get(self._zero())
            if P0 is None:
                return []
            return [(i + 1, j + 1 + self.n_orb, float(P0[i, j])) for i in range(self.n) for j in range(self.n)
                    if P0[i, j] != 0.0]
        T0 = self.terms.get(name, {}).get(self._zero())
        out = []
        if T0 is None:
            return out
        for s in (0, self.n_orb):
            for i in range(self.n):
                for j in range(i, self.n):
                    if T0[i, j] != 0.0:
                        out.append((i + 1 + s, j + 1 + s, float(T0[i, j])))
        return out

    def spin_flip_terms(self, values):
        """{n: 2n x 2n matrix} of the lattice Hamiltonian in the spinor
        Psi_k = (c_up(k), c_down(k)), h(k) = sum_n H^n exp(2 pi i k.n) with
          H^n = [[T^n, F^n], [(F^-n)^T, T^n]]
        (the one-body matrices of the spin-flip mixing, one_body_operator.hpp:326-411)."""
        n = self.n
        T, F = {}, {}
        for src, dst in ((self.terms, T), (self.flips, F)):
            for name, d in src.items():
                z = values.get(name, 0.0)
                if not z:
                    continue
                for nvec, M in d.items():
                    dst[nvec] = dst.get(nvec, 0.0) + z * M
        out = {}
        for nvec in set(T) | set(F) | {tuple(-x for x in v) for v in F}:
            neg = tuple(-x for x in nvec)
            H = np.zeros((2 * n, 2 * n))
            if nvec in T:
                H[:n, :n] = T[nvec]
                H[n:, n:] = T[nvec]
            if nvec in F:
                H[:n, n:] = F[nvec]
            if neg in F:
                H[n:, :n] = np.asarray(F[neg]).T
            if np.any(H):
                out[nvec] = H
        return out

    def fourier_terms(self, values, nambu=False, mixing=None):
        """Inter-cluster terms: (nvecs (m x 3) float, mats (m x n x n) complex), n != 0
        (with nambu, or a spinor mixing 1 or 2, the 2n x 2n spinor matrices)."""
        acc = {}
        zero = self._zero()
        mixing = (1 if nambu else 0) if mixing is None else mixing
        if mixing:
            acc = {k: v for k, v in self.spinor_terms(values, mixing).items() if k != zero}
            keys = list(acc)
            nv = np.array([list(k) + [0] * (3 - self.dim) for k in keys], float).reshape(-1, 3)
            mats = np.array([acc[k] for k in keys], complex).reshape(-1, 2 * self.n, 2 * self.n)
            return nv, mats
        for name, d in self.terms.items():
            z = values.get(name, 0.0)
            if not z:
                continue
            for nvec, T in d.items():
                if nvec == zero:
                    continue
                acc[nvec] = acc.get(nvec, 0.0) + z * T
        keys = [k for k in acc if np.any(acc[k] != 0.0)]
        nv = np.array([list(k) + [0] * (3 - self.dim) for k in keys], float).reshape(-1, 3)
        mats = np.array([acc[k] for k in keys], complex).reshape(-1, self.n, self.n)
        return nv, mats

    def tk(self, values, k, nambu=False, mixing=None):
        """t(k) for wavevectors k (nk x 3), dense host evaluation (with nambu or a spinor
        mixing, the spinor h(k))."""
        k = np.atleast_2d(np.asarray(k, float))
        nv, mats = self.fourier_terms(values, nambu, mixing)
        T0 = self.cluster_matrix(values, nambu, mixing).astype(complex)
        if not len(nv):
            return np.repeat(T0[None], len(k), 0)
        ph = np.exp(2j * np.pi * (k @ nv.T))
        return T0[None] + np.einsum("km,mij->kij", ph, mats)

    def kgrid(self, side, shift=0.5):
        """Regular grid in the superdual basis, k_a = (i_a + shift)/side (nk x 3)."""
        axes = [(np.arange(side) + shift) / side for _ in range(self.dim)]
        mesh = np.meshgrid(*axes, indexing="ij")
        k = np.zeros((side ** self.dim, 3))
        for a in range(self.dim):
            k[:, a] = mesh[a].ravel()
        return k


def square_tiling(nx, ny):
    """nx x ny cluster (site index x + nx*y) tiling the square lattice, with the
    operators of the Hubbard model: t (NN), tp (NNN), mu."""
    sites = [[x, y, 0] for y in range(ny) for x in range(nx)]
    T = Tiling(sites, [[nx, 0, 0], [0, ny, 0]])
    T.hop("t", (1, 0, 0)).hop("t", (0, 1, 0))
    T.hop("tp", (1, 1, 0)).hop("tp", (-1, 1, 0))
    T.onsite("mu", -1.0)
    return T


def chain_tiling(n):
    """n-site cluster of a 1D chain: t (NN), tp (NNN), mu, cdw (Q = pi)."""
    T = Tiling([[x, 0, 0] for x in range(n)], [[n, 0, 0]])
    T.hop("t", (1, 0, 0)).hop("tp", (2, 0, 0))
    T.onsite("mu", -1.0)
    T.density_wave("cdw", (1, 0, 0))
    return T


# --------------------------------------------------------------- device k-sums

def _lattice_call(fn, n, nw, Geff, V0, nv, mats, k, *tail):
    k = np.ascontiguousarray(np.asarray(k, float).reshape(-1, 3))
    nt = len(nv)
    Vt = _cm(mats) if nt else np.zeros(2)
    nvb = np.ascontiguousarray(nv, float).ravel() if nt else np.zeros(3)
    _lib.call(fn, n, nw, _lib.dp(_cm(Geff)), _lib.dp(_cm(V0)), nt, _lib.dp(Vt), _lib.dp(nvb),
              len(k), _lib.dp(k.ravel()), *tail)


def cpt_batch(Geff, V0, nv, mats, k, mode="full"):
    """G_cpt for every (w, k): Geff (nw x n x n), V(k) = V0 + sum_t mats[t] e^{2 pi i k.nv[t]}.
    mode 'full' -> (nw, nk, n, n); 'trace' -> (nw, nk); 'average' -> (nw, n, n)."""
    Geff = np.asarray(Geff, np.complex128)
    nw, n = Geff.shape[0], Geff.shape[1]
    nk = len(np.asarray(k).reshape(-1, 3))
    m = {"full": 0, "trace": 1, "average": 2}[mode]
    size = {0: nw * nk * n * n, 1: nw * nk, 2: nw * n * n}[m]
    out = np.zeros(2 * size)
    _lattice_call("qed_cpt_batch", n, nw, Geff, np.asarray(V0, complex), nv, mats, k, m,
                  _lib.dp(out))
    if m == 0:
        return _from_cm(out, (nw, nk, n, n))
    if m == 1:
        return out.view(np.complex128).reshape(nw, nk).copy()
    return _from_cm(out, (nw, n, n))


def sef_batch(Geff, V0, nv, mats, k):
    """(1/nk) sum_k log|det(V(k) Geff(w) - 1)| per frequency."""
    Geff = np.asarray(Geff, np.complex128)
    nw, n = Geff.shape[0], Geff.shape[1]
    out = np.zeros(nw)
    _lattice_call("qed_sef_batch", n, nw, Geff, np.asarray(V0, complex), nv, mats, k,
                  _lib.dp(out))
    return out


def _distance_scale(mixing, wr, n):
    """The distance's factor (src_qcm/CPT.cpp:548-552): 2 at mixing 0, 1/2 at mixing 3,
    the frequency step, over dim^2; at mixing 4 the two spin blocks come side by side
    (n = 2 dim), the reference adding their sums over its dim x dim matrices."""
    f = {0: 2.0, 3: 0.5}.get(mixing, 1.0)
    dim = n // 2 if mixing == 4 else n
    return f * (wr[1] - wr[0]) / (dim * dim)


def cdmft_distance_batch(wr, weights, Ghost, eb, tcb, mixing=0):
    """Distance function for B bath parameter sets at once (src_qcm/CPT.cpp:519-555):
    d = f dw / dim^2 sum_w weight(w) ||Gamma(i w) + G_host(i w)||_F^2, f = 2 at mixing 0,
    dw = wr[1] - wr[0]. eb: (B, nb) bath energies, tcb: (B, n, nb) cluster-bath couplings
    in the basis that diagonalizes the bath."""
    wr = np.ascontiguousarray(wr, float)
    weights = np.ascontiguousarray(weights, float)
    eb = np.ascontiguousarray(np.atleast_2d(eb), float)
    tcb = np.asarray(tcb, float)
    if tcb.ndim == 2:
        tcb = tcb[None]
    B, n, nb = tcb.shape
    if eb.shape != (B, nb) or np.shape(Ghost) != (len(wr), n, n):
        raise ValueError("cdmft_distance_batch: inconsistent shapes")
    scale = _distance_scale(mixing, wr, n)
    d = np.zeros(B)
    tcm = np.ascontiguousarray(np.swapaxes(tcb, 1, 2)).ravel()  # (i, m) at i + n m
    _lib.call("qed_cdmft_distance_batch", n, nb, len(wr), _lib.dp(wr), _lib.dp(weights),
              _lib.dp(_cm(Ghost)), B, _lib.dp(eb.ravel()), _lib.dp(tcm), float(scale),
              _lib.dp(d))
    return d


def cdmft_distance_multi(wr, weights, hosts, host_of, eb, tcb, mixing=0):
    """Distance for B parameter sets, set b against hosts[host_of[b]] (nhost x nw x n x n)."""
    wr = np.ascontiguousarray(wr, float)
    weights = np.ascontiguousarray(weights, float)
    hosts = np.asarray(hosts, complex)
    eb = np.ascontiguousarray(np.atleast_2d(eb), float)
    tcb = np.asarray(tcb, float)
    B, n, nb = tcb.shape
    host_of = np.ascontiguousarray(host_of, np.int32)
    if hosts.shape[1:] != (len(wr), n, n) or eb.shape != (B, nb) or host_of.shape != (B,):
        raise ValueError("cdmft_distance_multi: inconsistent shapes")
    scale = _distance_scale(mixing, wr, n)
    d = np.zeros(B)
    tcm = np.ascontiguousarray(np.swapaxes(tcb, 1, 2)).ravel()
    _lib.call("qed_cdmft_distance_multi", n, nb, len(wr), _lib.dp(wr), _lib.dp(weights),
              len(hosts), _lib.dp(_cm(hosts)), B, _lib.ip(host_of), _lib.dp(eb.ravel()),
              _lib.dp(tcm), float(scale), _lib.dp(d))
    return d


def cdmft_fit_batch(wr, weights, hosts, host_of, E, T, X0, initial_step=0.1, xatol=1e-6,
                    fatol=1e-8, maxfev=1_000_000, mixing=0):
    """Adaptive Nelder-Mead bath fits of B instances in one device launch.
    E: (nvar, nb) bath-energy coefficients, T: (nvar, n, nb) coupling coefficients,
    X0: (B, nvar) starting points. Returns (x (B, nvar), f (B), nfev (B), ok (B))."""
    wr = np.ascontiguousarray(wr, float)
    weights = np.ascontiguousarray(weights, float)
    hosts = np.asarray(hosts, complex)
    E = np.asarray(E, float)
    T = np.asarray(T, float)
    nvar, n, nb = T.shape
    X0 = np.ascontiguousarray(np.atleast_2d(X0), float)
    B = len(X0)
    host_of = np.ascontiguousarray(host_of, np.int32)
    if hosts.shape[1:] != (len(wr), n, n) or E.shape != (nvar, nb) or X0.shape[1] != nvar:
        raise ValueError("cdmft_fit_batch: inconsistent shapes")
    scale = _distance_scale(mixing, wr, n)
    Tcm = np.ascontiguousarray(np.swapaxes(T, 1, 2)).ravel()
    x = np.zeros((B, nvar))
    f = np.zeros(B)
    nfev = np.zeros(B, np.int32)
    status = np.zeros(B, np.int32)
    _lib.call("qed_cdmft_fit_batch", n, nb, len(wr), _lib.dp(wr), _lib.dp(weights), len(hosts),
              _lib.dp(_cm(hosts)), B, _lib.ip(host_of), nvar,
              _lib.dp(np.ascontiguousarray(E).ravel()), _lib.dp(Tcm), _lib.dp(X0.ravel()),
              float(initial_step), float(xatol), float(fatol), int(maxfev), float(scale),
              _lib.dp(x), _lib.dp(f), _lib.ip(nfev), _lib.ip(status))
    return x, f, nfev, status == 0


# -------------------------------------------------------------- frequency grids

def legendre_frequency_grid(w1, w2, n1, n2, n3):
    """Gauss-Legendre grid on the positive imaginary axis with n1, n2, n3 points in
    [0, w1], [w1, w2] and [w2, inf) (the last interval mapped by u = 1/w); the weights
    include the 1/pi of the frequency integral (pyqcm/__init__.py:3040-3066)."""
    w = np.zeros(n1 + n2 + n3)
    p = np.zeros(n1 + n2 + n3)
    x, q = np.polynomial.legendre.leggauss(n1)
    w[:n1] = 0.5 * w1 * (x + 1.0)
    p[:n1] = 0.5 * w1 * q
    x, q = np.polynomial.legendre.leggauss(n2)
    w[n1:n1 + n2] = w1 + 0.5 * (w2 - w1) * (x + 1.0)
    p[n1:n1 + n2] = 0.5 * (w2 - w1) * q
    x, q = np.polynomial.legendre.leggauss(n3)
    u = 0.5 * (1.0 / w2) * (x[::-1] + 1.0)
    w[n1 + n2:] = 1.0 / u
    p[n1 + n2:] = 0.5 * (1.0 / w2) * q[::-1] * w[n1 + n2:] ** 2
    return w, p / np.pi


# ---------------------------------------------------------------- cluster side

# instance labels shared by every lattice model, so that two models never reuse one
_instance_labels = itertools.count(1)


class LatticeModel:
    """A lattice model: a tiling plus the ED cluster model (cluster sites first, then
    bath orbitals) registered through qcm. `lattice_ops` are the tiling operators whose
    intra-cluster parts are also cluster operators of the same name."""

    def __init__(self, name, tiling, n_bath=0, lattice_ops=("t", "tp", "mu", "cdw"),
                 interaction="U"):
        self.name = name
        self.tiling = tiling
        self.n = tiling.n
        self.n_bath = n_bath
        tiling.n_orb = self.n + n_bath
        qcm.new_model(name, self.n, n_bath)
        self.cluster_ops = []
        for op in lattice_ops:
            el = tiling.cluster_operator_elements(op)
            if el:
                qcm.new_operator(name, op, "anomalous" if op in tiling.pairs else "one-body", el)
                self.cluster_ops.append(op)
        L = self.L
        qcm.new_operator(name, interaction, "interaction",
                         [(i + 1, i + 1 + L, 1.0) for i in range(self.n)])
        self.cluster_ops.append(interaction)
        self.bath_ops = {}  # name -> (B (nb x nb), C (n x nb)) spin-up linear blocks
        self.bath_elements = {}  # name -> (1-based elements, kind) of add_cluster_operator

    @property
    def L(self):
        return self.n + self.n_bath

    def add_cluster_operator(self, name, elements, kind="one-body"):
        """Cluster operator with 1-based (r, c, v) elements over 2L spin orbitals (kind
        "one-body" or "anomalous"); a one-body operator's bath-bath and cluster-bath spin-up
        blocks are recorded for the bath fit."""
        qcm.new_operator(self.name, name, kind, elements)
        self.bath_elements[name] = (list(elements), kind)
        if kind != "one-body":
            self.bath_ops[name] = (np.zeros((self.n_bath, self.n_bath)), np.zeros((self.n, self.n_bath)))
            self.cluster_ops.append(name)
            return
        B = np.zeros((self.n_bath, self.n_bath))
        C = np.zeros((self.n, self.n_bath))
        for r, c, v in elements:
            r, c = r - 1, c - 1
            if r >= self.L or c >= self.L:
                continue
            if r >= self.n and c >= self.n:
                B[r - self.n, c - self.n] += v
                if r != c:
                    B[c - self.n, r - self.n] += v
            elif r < self.n <= c:
                C[r, c - self.n] += v
        self.bath_ops[name] = (B, C)
        self.cluster_ops.append(name)

    def up_down_bath_blocks(self, name):
        """(B (2nb x 2nb), C (2n x 2nb)): the spin-up and spin-down bath-bath and
        cluster-bath blocks of a one-body cluster operator at unit value, side by side
        (the up/down mixing: the two spins solved apart, src_ed/model_instance.hpp:424-475)."""
        el, kind = self.bath_elements[name]
        n, nb, L = self.n, self.n_bath, self.L
        Bn = np.zeros((2 * nb, 2 * nb))
        Cn = np.zeros((2 * n, 2 * nb))
        if kind != "one-body":
            return Bn, Cn
        for r, c, v in el:
            r, c = r - 1, c - 1
            if (r >= L) != (c >= L):
                raise RuntimeError("operator %s couples the two spins" % name)
            sp = 1 if r >= L else 0
            r, c = r - sp * L, c - sp * L
            for a, b in ((r, c), (c, r)) if r != c else ((r, c),):
                if a >= n and b >= n:
                    Bn[sp * nb + a - n, sp * nb + b - n] += v
                elif a < n <= b:
                    Cn[sp * n + a, sp * nb + b - n] += v
        return Bn, Cn

    def spin_flip_bath_blocks(self, name):
        """(B (2nb x 2nb), C (2n x 2nb)): the bath-bath and cluster-bath blocks of the spinor
        (c_up, c_down) matrices of a one-body cluster operator at unit value, the spin-flip
        mixing's rule (the operator's 2L x 2L matrix restricted to the orbitals,
        one_body_operator.hpp:326-411)."""
        el, kind = self.bath_elements[name]
        if kind != "one-body":
            raise RuntimeError("operator %s: the spin-flip mixing takes one-body bath operators" % name)
        n, nb, L = self.n, self.n_bath, self.L
        F = np.zeros((2 * L, 2 * L))
        for r, c, v in el:
            r, c = r - 1, c - 1
            if r < c:
                F[r, c] += v
                F[c, r] += v
            elif r == c:
                F[r, r] += v
        S = list(range(n)) + list(range(L, L + n))
        Bt = list(range(n, L)) + list(range(L + n, 2 * L))
        return F[np.ix_(Bt, Bt)].copy(), F[np.ix_(S, Bt)].copy()

    def nambu_bath_blocks(self, name):
        """(B (2nb x 2nb), C (2n x 2nb)): the bath-bath and cluster-bath blocks of the Nambu
        spinor matrices of a cluster operator at unit value, by the rules of the cluster's
        spinor one-body matrices (one_body_operator.hpp:326-411, anomalous_operator.hpp:
        241-284): with M the operator over the 2L spin orbitals (one-body: v on (r, c) and
        (c, r); anomalous: v/2 on (r, c) and -v/2 on (c, r), the binding's halving),
          one-body   bath  diag(M_uu, -M_dd^T)    cluster-bath  diag(M_uu, -M_dd)
          anomalous  bath  2 M_du below, 2 M_du^T above
                     cluster-bath  2 M_du below, -2 M_ud above."""
        el, kind = self.bath_elements[name]
        n, nb, L = self.n, self.n_bath, self.L
        F = np.zeros((2 * L, 2 * L))
        for r, c, v in el:
            r, c = r - 1, c - 1
            if kind == "anomalous":
                if r < c:
                    F[r, c] += 0.5 * v
                    F[c, r] -= 0.5 * v
            elif r < c:
                F[r, c] += v
                F[c, r] += v
            elif r == c:
                F[r, r] += v
        S, Bt = slice(0, n), slice(n, L)
        uu, dd = F[:L, :L], F[L:, L:]
        du, ud = F[L:, :L], F[:L, L:]
        Bn = np.zeros((2 * nb, 2 * nb))
        Cn = np.zeros((2 * n, 2 * nb))
        if kind == "anomalous":
            Bn[nb:, :nb] = 2.0 * du[Bt, Bt]
            Bn[:nb, nb:] = 2.0 * du[Bt, Bt].T
            Cn[n:, :nb] = 2.0 * du[S, Bt]
            Cn[:n, nb:] = -2.0 * ud[S, Bt]
        else:
            Bn[:nb, :nb] = uu[Bt, Bt]
            Bn[nb:, nb:] = -dd[Bt, Bt].T
            Cn[:n, :nb] = uu[S, Bt]
            Cn[n:, nb:] = -dd[S, Bt]
        return Bn, Cn

    def diagonal_bath(self, names, X, nambu=False, mixing=None):
        """(eb (B x nb), tcb (B x n x nb)) for rows X (B x len(names)) of bath parameters;
        a non-diagonal bath matrix is diagonalized as in the reference's hybridization
        function (src_ed/model_instance.hpp:1004-1018). With nambu, the spinor forms:
        eb (B x 2nb), tcb (B x 2n x 2nb)."""
        X = np.atleast_2d(np.asarray(X, float))
        mixing = (1 if nambu else 0) if mixing is None else mixing
        if mixing == 4:
            # both spin blocks side by side (up, then down)
            blocks = [self.up_down_bath_blocks(nm) for nm in names]
            E = np.einsum("bi,ijk->bjk", X, np.array([b[0] for b in blocks]))
            T = np.einsum("bi,ijk->bjk", X, np.array([b[1] for b in blocks]))
            eb, U = np.linalg.eigh(E)
            return eb, T @ U
        if mixing in (1, 2):
            blocks = [(self.nambu_bath_blocks if mixing == 1 else self.spin_flip_bath_blocks)(nm) for nm in names]
            E = np.einsum("bi,ijk->bjk", X, np.array([b[0] for b in blocks]))
            T = np.einsum("bi,ijk->bjk", X, np.array([b[1] for b in blocks]))
            eb, U = np.linalg.eigh(E)
            return eb, T @ U
        E = np.zeros((len(X), self.n_bath, self.n_bath))
        T = np.zeros((len(X), self.n, self.n_bath))
        for i, nm in enumerate(names):
            Bm, Cm = self.bath_ops[nm]
            E += X[:, i, None, None] * Bm
            T += X[:, i, None, None] * Cm
        off = E - np.einsum("bii->bi", E)[..., None] * np.eye(self.n_bath)
        if not np.any(off):
            return np.einsum("bii->bi", E).copy(), T
        eb, U = np.linalg.eigh(E)
        return eb, T @ U

    def instance(self, values, sectors, label=None):
        """Solves nothing yet: registers the cluster instance. As in pyqcm, a name with the
        suffix '_1' is a cluster parameter (e.g. a Weiss field 'cdw_1' or bath 'eb1_1');
        other names are lattice parameters, inherited by the cluster operator of the same
        name unless that has its own '_1' value."""
        if label is None:
            label = 7000 + next(_instance_labels)
        lattice_vals = {k: v for k, v in values.items() if not k.endswith("_1")}
        cluster_vals = {k: v for k, v in lattice_vals.items() if k in self.cluster_ops}
        for k, v in values.items():
            if k.endswith("_1"):
                if k[:-2] not in self.cluster_ops:
                    raise ValueError("%s is not a cluster operator of %s" % (k[:-2], self.name))
                cluster_vals[k[:-2]] = v
        qcm.new_model_instanceC(self.name, cluster_vals, sectors, label)
        return LatticeInstance(self, lattice_vals, label)


class LatticeInstance:
    """One cluster instance on the lattice (model_instance + lattice_model_instance)."""

    def __init__(self, model, values, label):
        self.model = model
        self.values = values
        self.label = label
        self._gs = None

    def erase(self):
        """Releases the cluster instance (its device vectors) in the solver."""
        qcm.erase_model_instance(self.label)

    def ground_state(self):
        if self._gs is None:
            self._gs = qcm.ground_state_solve(self.label)
        return self._gs

    @property
    def mixing(self):
        """The instance's mixing: 0 normal, 1 Nambu (pairing terms), 4 up/down (spin
        asymmetric cluster terms), the lattice model's mixing being the instance's
        (src_qcm/lattice_model.cpp:458-468)."""
        self.ground_state()
        m = qcm.mixingC(self.label)
        if m not in (0, 1, 2, 4):
            raise RuntimeError("the lattice functions take the normal, Nambu, spin-flip or up/down mixing "
                               "only (instance mixing %d)" % m)
        return m

    @property
    def spinor(self):
        """The spinor mixing of the lattice functions: 1 Nambu, 2 spin flip, 0 none."""
        return self.mixing if self.mixing in (1, 2) else 0

    @property
    def nambu(self):
        return self.mixing == 1

    def spins(self):
        """The spin blocks the lattice functions sum over: both for the up/down mixing."""
        return (0, 1) if self.mixing == 4 else (0,)

    def green(self, z, spin_down=0):
        self.mixing
        return qcm.Green_function(np.atleast_1d(z), spin_down, self.label)

    def _geff(self, z, spin_down=0):
        G = self.green(z, spin_down)
        if self.model.n_bath == 0:
            return G, G
        Gam = qcm.hybridization_functionC(np.atleast_1d(z), spin_down, self.label)
        return G, np.linalg.inv(np.linalg.inv(G) + Gam)

    def v0(self, spin_down=0):
        """V0 = T_0(lattice) - t_c, the k-independent part of V (Nambu: 2n x 2n)."""
        tc = qcm.hopping_matrix(spin_down, 0, self.label).real
        return self.model.tiling.cluster_matrix(self.values, mixing=self.spinor) - tc

    def cpt_green(self, k, z, mode="full", spin_down=0):
        """CPT Green function at frequencies z and wavevectors k (QCM::CPT_Green_function,
        src_qcm/QCM.cpp:328-344); with the Nambu mixing, of the spinor
        (c_up(k), c+_down(-k)), 2n x 2n."""
        _, Geff = self._geff(z, spin_down)
        nv, mats = self.model.tiling.fourier_terms(self.values, mixing=self.spinor)
        return cpt_batch(Geff, self.v0(spin_down), nv, mats, k, mode)

    def sef_integrand(self, z, k, spin_down=0):
        """(1/nk) sum_k log|det(V(k, z) G_c(z) - 1)| per frequency (one spin block)."""
        G, Geff = self._geff(z, spin_down)
        nv, mats = self.model.tiling.fourier_terms(self.values, mixing=self.spinor)
        I = sef_batch(Geff, self.v0(spin_down), nv, mats, k)
        if self.model.n_bath:
            I = I + np.log(np.abs(np.linalg.det(G))) - np.log(np.abs(np.linalg.det(Geff)))
        return I

    def potthoff_functional(self, grid=(1, 10, 10, 10, 10), kside=64):
        """Self-energy functional on a fixed (w, k) grid, the grid path of
        lattice_model_instance::Potthoff_functional (src_qcm/lattice_model_instance.cpp:503-584):
          Omega = [Omega_c - sum_w weight_w f (1/nk) sum_k log|det(V G - 1)| - Omega_diag] / n_sites,
          Omega_diag = (1/2) f sum_i s_i (H_ii - Hc_ii),
        with f = 2 for the normal mixing (two equal spins) and 1 otherwise (the up/down
        mixing sums both spin blocks), s_i = +1 on the Nambu (hole) components and -1 on
        the others, and the k-grid k = i/side (src_util/integrate.cpp:177-239)."""
        ws, wt = legendre_frequency_grid(*grid)
        k = self.model.tiling.kgrid(kside, shift=0.0)
        f = 2.0 if self.mixing == 0 else 1.0
        Iw = f * sum(self.sef_integrand(1j * ws, k, sp) for sp in self.spins())
        omega_trace = -np.sum(Iw * wt)
        omega_diag = 0.0
        for sp in self.spins():
            d = np.diag(self.v0(sp))
            if self.nambu:
                n = self.model.n
                omega_diag += d[n:].sum() - d[:n].sum()
            else:
                omega_diag -= d.sum()
        omega_diag *= 0.5 * f
        omega_clus = qcm.Potthoff_functional(self.label)
        return (omega_trace - omega_diag + omega_clus) / self.model.n

    def projected_green(self, wr, kside=32):
        """(1/nk) sum_k G_cpt(k, i w) on k = ((i+1/2)/side, ...) (src_qcm/CPT.cpp:283-345)."""
        k = self.model.tiling.kgrid(kside, shift=0.5)
        return self.cpt_green(k, 1j * np.asarray(wr, float), mode="average")

    def cdmft_host(self, wr, kside=32, spin_down=0):
        """G_host(i w) = Gproj^-1 + Sigma_c + t_c - i w (src_qcm/CPT.cpp:356-424), of one
        spin block."""
        z = 1j * np.asarray(wr, float)
        k = self.model.tiling.kgrid(kside, shift=0.5)
        Gp = self.cpt_green(k, z, mode="average", spin_down=spin_down)
        S = qcm.self_energyC(z, spin_down, self.label)
        tc = qcm.hopping_matrix(spin_down, 0, self.label)
        n = Gp.shape[1]
        return np.linalg.inv(Gp) + S + tc[None] - z[:, None, None] * np.eye(n)[None]
