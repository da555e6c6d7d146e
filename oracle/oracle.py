"""ORACLE — test infrastructure only (never imported by the product path).

Python half of the CPU oracle for the ED impurity-solver hot path:

* ctypes access to ``liboracle.so`` (the C++ restatement in ``qed_oracle.cpp``);
* an independent Jordan-Wigner brute force (dense fermion matrices) used to pin the
  restatement on small clusters;
* numpy restatements of the reference's band Lanczos, Q-matrix post-processing,
  continued fractions and the small dense lattice kernels (CPT, SEF, CDMFT distance).

Reference provenance of each function is given in its docstring. Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.check_call(["make", "-s", "-f", os.path.join(_HERE, "Makefile")])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        dptr = P(ctypes.c_double)
        iptr = P(ctypes.c_int)
        L.or_halfbasis.argtypes = [ctypes.c_int, ctypes.c_int, P(ctypes.c_uint32)]
        L.or_halfbasis.restype = ctypes.c_int
        L.or_sector_new.argtypes = [ctypes.c_int] * 4 + [iptr, dptr, iptr, iptr, iptr, dptr]
        L.or_sector_new.restype = ctypes.c_void_p
        L.or_sector_free.argtypes = [ctypes.c_void_p]
        L.or_sector_dim.argtypes = [ctypes.c_void_p]
        L.or_sector_dim.restype = ctypes.c_uint64
        L.or_hv.argtypes = [ctypes.c_void_p, dptr, dptr]
        L.or_hv_omp.argtypes = [ctypes.c_void_p, dptr, dptr, ctypes.c_int]
        L.or_start_vector.argtypes = [ctypes.c_uint64, ctypes.c_int, dptr]
        L.or_reset_normal.argtypes = []
        L.or_lanczos_gs.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, dptr, dptr, iptr,
                                    dptr, dptr]
        L.or_lanczos_gs.restype = ctypes.c_int
        L.or_lanczos_gs_threads.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, dptr, dptr, iptr,
                                            dptr, dptr, ctypes.c_int, ctypes.c_int]
        L.or_lanczos_gs_threads.restype = ctypes.c_int
        L.or_lanczos_steps.argtypes = [ctypes.c_void_p, dptr, ctypes.c_int, ctypes.c_int, dptr, dptr]
        L.or_lanczos_steps.restype = ctypes.c_int
        L.or_lanczos_green.argtypes = [ctypes.c_void_p, dptr, ctypes.c_double, ctypes.c_int, dptr, dptr]
        L.or_lanczos_green.restype = ctypes.c_int
        L.or_create_destroy.argtypes = [ctypes.c_int] * 5 + [dptr, dptr, ctypes.c_double]
        L.or_create_destroy.restype = ctypes.c_int
        L.or_dstev.argtypes = [ctypes.c_int, dptr, dptr, dptr]
        L.or_dstev.restype = ctypes.c_int
        L.or_last_error.restype = ctypes.c_char_p
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


# ----------------------------------------------------------------------------------------
# model description shared by oracle and tests


@dataclass
class Operator:
    """A cluster operator as ED::new_operator receives it: 0-based spin-orbital element list
    (down orbital = orbital + L) with r <= c, and its type ('one-body' or 'interaction')."""
    name: str
    kind: str
    elements: list  # [(r, c, v)]


@dataclass
class ClusterModel:
    L: int                      # orbitals per spin (sites + bath)
    operators: dict = field(default_factory=dict)

    def add(self, name, kind, elements):
        self.operators[name] = Operator(name, kind, list(elements))

    def flat(self, values):
        """Flattened element arrays for the active operators (reference order = insertion)."""
        types, coefs, nels, er, ec, ev = [], [], [], [], [], []
        for name, op in self.operators.items():
            if name not in values or values[name] == 0.0:
                continue
            types.append(0 if op.kind == "one-body" else 1)
            coefs.append(float(values[name]))
            nels.append(len(op.elements))
            for (r, c, v) in op.elements:
                er.append(r)
                ec.append(c)
                ev.append(v)
        return (np.array(types, np.int32), np.array(coefs, np.float64), np.array(nels, np.int32),
                np.array(er, np.int32), np.array(ec, np.int32), np.array(ev, np.float64))


def hubbard_cluster(sites, bonds, t=1.0, tp_bonds=(), tp=0.0, U=0.0, mu=0.0, n_bath=0):
    """Cluster model with pyqcm's lattice conventions: hopping amplitude -1 per bond
    (src_qcm/lattice_model.cpp:968-1019 with amplitude -1), mu as -n per orbital
    (src_qcm/lattice_model.cpp:286-295), Hubbard (i, i+L, 1) (lattice_model.cpp:929-933)."""
    L = sites + n_bath
    m = ClusterModel(L)
    el = []
    for (i, j) in bonds:
        a, b = min(i, j), max(i, j)
        el += [(a, b, -1.0), (a + L, b + L, -1.0)]
    m.add("t", "one-body", el)
    if tp_bonds:
        el = []
        for (i, j) in tp_bonds:
            a, b = min(i, j), max(i, j)
            el += [(a, b, -1.0), (a + L, b + L, -1.0)]
        m.add("tp", "one-body", el)
    m.add("U", "interaction", [(i, i + L, 1.0) for i in range(sites)])
    m.add("mu", "one-body", [(i, i, -1.0) for i in range(sites)] + [(i + L, i + L, -1.0) for i in range(sites)])
    vals = {"t": t, "U": U, "mu": mu}
    if tp_bonds:
        vals["tp"] = tp
    return m, vals


def square_bonds(nx, ny):
    """Open-boundary nearest and next-nearest neighbour bonds of an nx x ny cluster,
    site index x + nx*y."""
    nn, nnn = [], []
    idx = lambda x, y: x + nx * y
    for y in range(ny):
        for x in range(nx):
            if x + 1 < nx:
                nn.append((idx(x, y), idx(x + 1, y)))
            if y + 1 < ny:
                nn.append((idx(x, y), idx(x, y + 1)))
            if x + 1 < nx and y + 1 < ny:
                nnn.append((idx(x, y), idx(x + 1, y + 1)))
            if x - 1 >= 0 and y + 1 < ny:
                nnn.append((idx(x, y), idx(x - 1, y + 1)))
    return nn, nnn


# ----------------------------------------------------------------------------------------
# C++ restatement access


def halfbasis(L, N):
    n = lib().or_halfbasis(L, N, None)
    out = np.zeros(n, np.uint32)
    lib().or_halfbasis(L, N, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return out


class Sector:
    """One (N_up, N_dn) sector of a cluster model in the factorized basis."""

    def __init__(self, model: ClusterModel, values: dict, nup: int, ndn: int):
        self.model, self.values, self.nup, self.ndn = model, values, nup, ndn
        arrs = model.flat(values)
        self._keep = arrs
        t, c, n, r, cc, v = arrs
        self.h = lib().or_sector_new(model.L, nup, ndn, len(t), _ip(t), _dp(c), _ip(n), _ip(r), _ip(cc), _dp(v))
        if not self.h:
            raise RuntimeError(lib().or_last_error().decode())
        self.dim = int(lib().or_sector_dim(self.h))
        self.du = len(halfbasis(model.L, nup))
        self.dd = len(halfbasis(model.L, ndn))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_sector_free(self.h)

    def hv(self, x, y=None):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.dim) if y is None else np.array(y, np.float64, copy=True)
        lib().or_hv(self.h, _dp(x), _dp(y))
        return y

    def hv_omp(self, x, y=None, threads=0):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.dim) if y is None else np.array(y, np.float64, copy=True)
        lib().or_hv_omp(self.h, _dp(x), _dp(y), threads)
        return y

    def start_vector(self, seed=0):
        x = np.zeros(self.dim)
        lib().or_start_vector(self.dim, seed, _dp(x))
        return x

    def lanczos_gs(self, seed=0, accur=1e-12, max_iter=1000, start=None):
        gs = self.start_vector(seed) if start is None else np.array(start, np.float64, copy=True)
        E0 = np.zeros(1)
        nit = np.zeros(1, np.int32)
        a = np.zeros(max_iter + 2)
        b = np.zeros(max_iter + 2)
        rc = lib().or_lanczos_gs(self.h, accur, max_iter, _dp(gs), _dp(E0), _ip(nit), _dp(a), _dp(b))
        if rc != 0:
            raise RuntimeError(lib().or_last_error().decode())
        n = int(nit[0])
        return float(E0[0]), gs, n, a[:n].copy(), b[:n].copy()

    def lanczos_green(self, psi, accur_cf=0.01, max_iter=400):
        psi = np.ascontiguousarray(psi, np.float64)
        a = np.zeros(max_iter + 2)
        b = np.zeros(max_iter + 2)
        n = lib().or_lanczos_green(self.h, _dp(psi), accur_cf, max_iter, _dp(a), _dp(b))
        return a[:n].copy(), b[:n + 1].copy()


def create_destroy(L, nup, ndn, a, pm, x, y=None, z=1.0):
    """y += z c_a x (pm=-1) or z c†_a x (pm=+1); a is a spin-orbital index."""
    tup, tdn = nup, ndn
    if a >= L:
        tdn += pm
    else:
        tup += pm
    dt = len(halfbasis(L, tup)) * len(halfbasis(L, tdn))
    y = np.zeros(dt) if y is None else y
    x = np.ascontiguousarray(x, np.float64)
    lib().or_create_destroy(L, nup, ndn, a, pm, _dp(x), _dp(y), z)
    return y


def dstev(d, e, vectors=True):
    d = np.array(d, np.float64, copy=True)
    n = len(d)
    e2 = np.zeros(max(n, 1))
    e2[:n - 1] = e[:n - 1]
    z = np.zeros((n, n)) if vectors else None
    info = lib().or_dstev(n, _dp(d), _dp(e2), _dp(z) if vectors else None)
    if info != 0:
        raise RuntimeError("dstev info %d" % info)
    # LAPACK writes column-major, so row k of the C-ordered array is eigenvector k
    return (d, z) if vectors else d


# ----------------------------------------------------------------------------------------
# independent brute force (Jordan-Wigner), small clusters only


def _fock_ops(L):
    """Annihilation operators c_a (a < 2L) on the 2^(2L) Fock space, standard Jordan-Wigner
    string over lower orbital indices (up orbitals 0..L-1 before down orbitals L..2L-1)."""
    import scipy.sparse as sp
    n = 2 * L
    dimF = 1 << n
    states = np.arange(dimF, dtype=np.int64)
    ops = []
    for a in range(n):
        occ = (states >> a) & 1
        src = states[occ == 1]
        tgt = src ^ (1 << a)
        below = src & ((1 << a) - 1)
        par = np.array([bin(int(s)).count("1") & 1 for s in below])
        val = np.where(par, -1.0, 1.0)
        ops.append(sp.csr_matrix((val, (tgt, src)), shape=(dimF, dimF)))
    return ops


def sector_states(L, nup, ndn):
    """Fock-space labels of the factorized sector basis, in the factorized index order
    I = i_up + dim_up * i_dn (i_up, i_dn ranks in ascending order)."""
    bu = [s for s in range(1 << L) if bin(s).count("1") == nup]
    bd = [s for s in range(1 << L) if bin(s).count("1") == ndn]
    return np.array([u | (d << L) for d in bd for u in bu], np.int64)


def dense_hamiltonian(model: ClusterModel, values, nup, ndn, states=None):
    """Dense H on one sector (or on the Fock states listed in states), built from
    Jordan-Wigner matrices (independent of the restatement's half-operator tables)."""
    L = model.L
    c = _fock_ops(L)
    cd = [m.T.tocsr() for m in c]
    dimF = 1 << (2 * L)
    import scipy.sparse as sp
    H = sp.csr_matrix((dimF, dimF))
    for name, op in model.operators.items():
        if name not in values or values[name] == 0.0:
            continue
        z = values[name]
        for (r, cc, v) in op.elements:
            if op.kind == "one-body":
                if r < cc:
                    H = H + z * v * (cd[r] @ c[cc] + cd[cc] @ c[r])
                elif r == cc:
                    H = H + z * v * (cd[r] @ c[r])
            elif op.kind == "anomalous":
                # (r, c, v) with r < c as the user passes it to new_operator: the binding
                # halves v (src_python/qcm_ED_wrap.hpp:179), anomalous_operator stores
                # (r, c, v/2) and (c, r, -v/2) (src_ed/Operators/anomalous_operator.hpp:54-58),
                # and HS_anomalous_operator (HS_anomalous_operator.hpp:20-36) puts
                # X = 2 (-v/2) pauli at <J|.|I> for J = I with bits r and c removed by
                # pair_annihilate(mask(c), mask(r)) (src_ed/binary_state.cpp:36-43, pauli =
                # (-1)^(bits strictly between)), and X at <I|.|J>. With Jordan-Wigner
                # operators c_c c_r |I> = (-1)^(bits strictly between) |J>, so the operator
                # is v (c_r c_c + h.c.).
                if r < cc:
                    P = c[r] @ c[cc]
                    H = H + z * v * (P + P.T)
            else:
                a, b = min(r, cc), max(r, cc)
                H = H + z * v * ((cd[a] @ c[a]) @ (cd[b] @ c[b]))
    idx = sector_states(L, nup, ndn) if states is None else np.asarray(states, np.int64)
    return H[idx][:, idx].toarray(), (c, cd)


def block_states(L, pairs):
    """Fock-space labels (up | down << L) of the union of the (N_up, N_down) blocks in
    pairs, ascending: the order of the reference's mixed basis for sectors that conserve
    only N or only S (src_ed/ED_basis.cpp:185-361, trivial group), whose binary states
    (up | down << 32) sort the same way."""
    out = []
    for (nu, nd) in pairs:
        out.extend(int(x) for x in sector_states(L, nu, nd))
    return np.array(sorted(out), np.int64)


def dense_spinor_gf(model, values, states0, spinor, states_minus, states_plus, zs):
    """Brute-force G_ab(z) = <A_a (z-H+E0)^-1 A_b^+> + <A_b^+ (z+H-E0)^-1 A_a> for the
    spinor components A_a = c_a (nambu False) or c_a^+ (nambu True), spinor =
    [(spin orbital a, nambu)], from the ground state of the basis states0 (labels);
    A_a |0> lies in states_minus and A_a^+ |0> in states_plus. T = 0, nondegenerate
    ground state."""
    L = model.L
    H0, (c, cd) = dense_hamiltonian(model, values, 0, 0, states=states0)
    e, U = np.linalg.eigh(H0)
    E0, psi = e[0], U[:, 0]
    dimF = 1 << (2 * L)
    full = np.zeros(dimF)
    full[np.asarray(states0)] = psi
    A = [(cd[a] if nb else c[a]) for (a, nb) in spinor]
    out = np.zeros((len(zs), len(spinor), len(spinor)), complex)
    for part, st in ((-1, states_minus), (1, states_plus)):
        if len(st) == 0:
            continue
        Hs, _ = dense_hamiltonian(model, values, 0, 0, states=st)
        es, Us = np.linalg.eigh(Hs)
        st = np.asarray(st)
        if part < 0:
            phi = np.array([(Aa @ full)[st] for Aa in A])
            Q = phi @ Us
            for i, z in enumerate(zs):
                out[i] += (Q / (z + (es - E0))) @ Q.T
        else:
            phi = np.array([(Aa.T @ full)[st] for Aa in A])
            Q = phi @ Us
            for i, z in enumerate(zs):
                out[i] += (Q / (z - (es - E0))) @ Q.T
    return E0, out


def dense_gf(model, values, nup, ndn, orbitals, zs):
    """Brute-force cluster Green function G_ab(z) = <c_a (z-H+E0)^-1 c†_b> + <c†_b (z+H-E0)^-1 c_a>
    for spin-up orbitals, from dense sector Hamiltonians (T=0, nondegenerate ground state)."""
    L = model.L
    H0, (c, cd) = dense_hamiltonian(model, values, nup, ndn)
    e, U = np.linalg.eigh(H0)
    E0, psi = e[0], U[:, 0]
    i0 = sector_states(L, nup, ndn)
    dimF = 1 << (2 * L)
    full = np.zeros(dimF)
    full[i0] = psi
    Hp, _ = dense_hamiltonian(model, values, nup + 1, ndn)
    Hm, _ = dense_hamiltonian(model, values, nup - 1, ndn)
    ip = sector_states(L, nup + 1, ndn)
    im = sector_states(L, nup - 1, ndn)
    phip = np.array([(cd[a] @ full)[ip] for a in orbitals])
    phim = np.array([(c[a] @ full)[im] for a in orbitals])
    ep, Up = np.linalg.eigh(Hp)
    em, Um = np.linalg.eigh(Hm)
    Qp = phip @ Up          # (n_orb, dim+)
    Qm = phim @ Um
    out = []
    for z in zs:
        G = (Qp / (z - (ep - E0))) @ Qp.T + (Qm / (z + (em - E0))) @ Qm.T
        out.append(G)
    return E0, np.array(out)


# ----------------------------------------------------------------------------------------
# band Lanczos and Q-matrix (numpy restatement)


def band_lanczos(hv, phi, accur_deflation=1e-7, accur_band_lanczos=1e-12, max_iter_BL=600, M0=None):
    """Freund band Lanczos as in src_ed/Hamiltonian/Lanczos.hpp:214-395 with the
    iteration budget of src_ed/Hamiltonian/Hamiltonian_base.hpp:237. Returns (e, Q) with
    Q = P U (p x M), or None when no iteration ran."""
    phi = [np.asarray(v, np.float64) for v in phi]
    p0 = len(phi)
    dim = len(phi[0])
    if M0 is None:
        M0 = int(14 * p0 * np.log(1.0 * dim))
    pc = p0
    if M0 > max_iter_BL:
        M0 = max_iter_BL
    M0p = (M0 // p0 + 1) * p0
    M0 = M0p - pc
    v = [None] * (M0p + 1)
    for i in range(p0):
        v[i] = phi[i].copy()
    t = np.zeros((M0, M0))
    P = np.zeros((p0, M0))
    defl = np.zeros(M0, np.int32)
    I = []
    ev_old = 1e12
    j = 0
    while j < M0:
        z = np.linalg.norm(v[j])
        if z < accur_deflation:
            if j >= pc:
                I.append(j - pc)
                defl[j - pc] = 1
            pc -= 1
            if pc == 0:
                break
            for k in range(j, j + pc):
                v[k] = v[k + 1]
            v[j + pc] = None
            continue
        v[j] = v[j] / z
        if j >= pc:
            t[j, j - pc] = z
        for k in range(p0):
            P[k, j] = phi[k] @ v[j]
        for k in range(j + 1, j + pc):
            zz = v[j] @ v[k]
            v[k] = v[k] - zz * v[j]
            if k >= pc:
                t[j, k - pc] = zz
        w = hv(v[j])
        k0 = j - pc if j > pc else 0
        for k in range(k0, j):
            t[k, j] = t[j, k]
            w = w - t[k, j] * v[k]
        for k in I:
            ww = v[k] @ w
            w = w - ww * v[k]
            t[k, j] = ww
        ww = v[j] @ w
        w = w - ww * v[j]
        t[j, j] = ww
        for k in I:
            t[j, k] = t[k, j]
        v[j + pc] = w
        if j >= 4 * p0 and j % p0 == 0:
            evs = np.linalg.eigvalsh(t[:j, :j])
            test = abs(evs[0] - ev_old)
            ev_old = evs[0]
            if test < accur_band_lanczos:
                break
        j += 1
    if j == 0:
        return None
    e, U = np.linalg.eigh(t[:j, :j])
    return e, P[:, :j] @ U


def dense_qmatrix(hv, dim, phi):
    """Full-diagonalization Lehmann form used below max_dim_full
    (src_ed/Hamiltonian/Hamiltonian_Dense.hpp:157-176): Q = phi . U."""
    H = np.array([hv(np.eye(dim)[i]) for i in range(dim)]).T
    e, U = np.linalg.eigh(H)
    return e, np.array(phi) @ U


def streamline(e, Q, wtol=1e-5, vtol=1e-10):
    """Merges poles closer than wtol and drops contributions below vtol
    (src_ed/Q_matrix.hpp:180-223, 285-318)."""
    if vtol < 1e-12:
        return e, Q
    L, M = Q.shape
    e2, cols = [], []
    i = 0
    while i < M:
        j = i + 1
        while j < M and abs(e[j] - e[i]) <= wtol:
            j += 1
        n = j - i
        if n == 1:
            col = Q[:, i]
            if np.sum(np.abs(col) ** 2) >= vtol:
                cols.append(col.reshape(L, 1))
                e2.append(e[i])
        else:
            blk = Q[:, i:j]
            QQ = blk @ blk.conj().T
            d, Uq = np.linalg.eigh(QQ)
            keep = [k for k in range(L) if d[k] > vtol]
            for k in keep:
                cols.append((Uq[:, k] * np.sqrt(d[k])).reshape(L, 1))
                e2.append(e[i])
        i = j
    if not cols:
        return np.zeros(0), np.zeros((L, 0))
    return np.array(e2), np.hstack(cols)


def check_norm(Q, threshold=1e-5, norm=1.0):
    """Unitarity error of a Q-matrix (src_ed/Q_matrix.hpp:235-260)."""
    S = Q @ Q.conj().T
    D = S - norm * np.eye(S.shape[0])
    return float(np.sum(np.abs(D) ** 2).real)


def qmatrix_gf(e, Q, z):
    """G_ab(z) = sum_i Q_ai conj(Q_bi)/(z - e_i) (src_ed/Q_matrix.hpp:133-151)."""
    return (Q / (z - e)) @ Q.conj().T


def cf_coefficients(a, b, e0, norm, create):
    """src_ed/continued_fraction.cpp:33-39."""
    a = np.array(a, np.float64, copy=True)
    b = np.array(b, np.float64, copy=True) ** 2
    b[0] = norm
    if create:
        a = a - e0
    else:
        a = -a + e0
    return a, b


def cf_evaluate(a, b, z):
    """src_ed/continued_fraction.cpp:52-57 (bottom-up)."""
    G = 0.0 + 0.0j
    for i in range(len(a) - 1, -1, -1):
        G = b[i] / (z - a[i] - G)
    return G


# ----------------------------------------------------------------------------------------
# block Lanczos and the matrix continued fraction (GF_method 'M')


def mgs_weight(phi, accur_deflation=1e-7):
    """Modified Gram-Schmidt of the start block with its upper-triangular factor W
    (phi[l] = sum_{k<=l} q[k] W[k, l]); stops at the first column whose norm falls below
    accur_deflation (src_ed/model_instance.hpp:885-905). Returns (q, W[:p_actual, :])."""
    q = [np.array(v, np.float64, copy=True) for v in phi]
    p = len(q)
    W = np.zeros((p, p))
    pa = p
    for l in range(p):
        for k in range(l):
            z = q[k] @ q[l]
            W[k, l] = z
            q[l] = q[l] - z * q[k]
        nrm = np.linalg.norm(q[l])
        if nrm < accur_deflation:
            pa = l
            break
        W[l, l] = nrm
        q[l] = q[l] / nrm
    return q[:pa], W[:pa, :]


def _block_lowest(A, B, j, p):
    sz = (j + 1) * p
    T = np.zeros((sz, sz))
    for jj in range(j + 1):
        T[jj * p:(jj + 1) * p, jj * p:(jj + 1) * p] = A[jj]
    for jj in range(j):
        T[(jj + 1) * p:(jj + 2) * p, jj * p:(jj + 1) * p] = B[jj]
        T[jj * p:(jj + 1) * p, (jj + 1) * p:(jj + 2) * p] = B[jj].T
    return np.linalg.eigvalsh(T)[0]


def block_lanczos(hv, q0, qr=False, accur_deflation=1e-7, accur_band_lanczos=1e-12, max_iter_BL=600,
                  M0=None):
    """Block Lanczos on p start vectors: H Q_j = Q_{j-1} B_{j-1}^T + Q_j A_j + Q_{j+1} B_j.
    qr=True: B_j upper triangular from modified Gram-Schmidt of the residual block
    (blockLanczos, src_ed/Hamiltonian/Lanczos.hpp:426-603); qr=False: B_j the symmetric
    square root of the residual Gram matrix (polar decomposition, blockLanczosSVD,
    src_ed/Hamiltonian/Lanczos.hpp:604-779). The lowest eigenvalue of the block-tridiagonal
    projection is tested every p steps from step 2p. Returns the lists (A, B)."""
    p = len(q0)
    dim = len(q0[0])
    if M0 is None:
        M0 = int(14 * p * np.log(1.0 * dim))  # src_ed/model_instance.hpp:929
    M0 = min(M0, max_iter_BL)
    Qc = [np.array(v, np.float64, copy=True) for v in q0]
    for l in range(p):  # the start block is orthonormalized again
        for k in range(l):
            Qc[l] = Qc[l] - (Qc[k] @ Qc[l]) * Qc[k]
        nrm = np.linalg.norm(Qc[l])
        if nrm < accur_deflation:
            raise ValueError("block Lanczos: rank-deficient start block")
        Qc[l] = Qc[l] / nrm
    Qp = None
    A, B = [], []
    ev_old = 1e12
    for j in range(M0):
        Z = [hv(Qc[k]) for k in range(p)]
        Aj = np.zeros((p, p))
        for l in range(p):
            for k in range(l + 1):
                Aj[k, l] = Aj[l, k] = Qc[k] @ Z[l]
        A.append(Aj)
        for l in range(p):
            for k in range(p):
                Z[l] = Z[l] - Aj[k, l] * Qc[k]
        if j > 0:
            Bp = B[-1]
            for l in range(p):
                for k in range(p):
                    Z[l] = Z[l] - Bp[l, k] * Qp[k]
        Bj = np.zeros((p, p))
        singular = False
        if qr:
            for l in range(p):
                for k in range(l):
                    z = Z[k] @ Z[l]
                    Bj[k, l] = z
                    Z[l] = Z[l] - z * Z[k]
                nrm = np.linalg.norm(Z[l])
                if nrm < accur_deflation:
                    for ll in range(l, p):
                        Z[ll] = np.zeros(dim)
                    singular = True
                    break
                Bj[l, l] = nrm
                Z[l] = Z[l] / nrm
            B.append(Bj)
            Qn = Z
        else:
            G = np.array([[Z[k] @ Z[l] for l in range(p)] for k in range(p)])
            D, V = np.linalg.eigh(G)
            singular = bool(np.any(D < accur_deflation ** 2))
            Bj = (V * np.sqrt(np.maximum(D, 0.0))) @ V.T
            B.append(Bj)
            if singular:
                break
            Binv = (V / np.sqrt(D)) @ V.T
            Qn = [sum(Binv[k, l] * Z[k] for k in range(p)) for l in range(p)]
        if (j + 1) >= 2 * p and (j + 1) % p == 0:
            e0 = _block_lowest(A, B, j, p)
            delta = abs(e0 - ev_old)
            ev_old = e0
            if delta < accur_band_lanczos:
                break
        if singular:
            break
        Qp, Qc = Qc, Qn
    return A, B


def mcf_shift(A, e0, create):
    """Energy shift of the diagonal blocks: A - e0 for particles, e0 - A for holes
    (src_ed/matrix_continued_fraction.hpp:93-115)."""
    I = np.eye(A[0].shape[0])
    return [(a - e0 * I) if create else (e0 * I - a) for a in A]


def mcf_evaluate(A, B, W, z):
    """G(z) = W^T F_0 W with F_M = 0, F_j = (z - A_j - B_j^T F_{j+1} B_j)^-1; the last B is
    the truncation residual and is not used (src_ed/matrix_continued_fraction.hpp:139-182)."""
    M = len(A)
    p = A[0].shape[0]
    F = np.zeros((p, p), complex)
    for j in range(M - 1, -1, -1):
        S = B[j].T @ F @ B[j] if j < M - 1 else 0.0
        F = np.linalg.inv(z * np.eye(p) - A[j] - S)
    return W.T @ F @ W


# ----------------------------------------------------------------------------------------
# lattice kernels (dense numpy, one (k, w) at a time)


def cpt_green(Gc, V):
    """G_cpt = G_c (1 - V G_c)^-1 (src_qcm/CPT.cpp:76-86)."""
    n = Gc.shape[0]
    return Gc @ np.linalg.inv(np.eye(n) - V @ Gc)


def sef_integrand(Gc, V, factor=2.0):
    """log|det(V G - 1)| times 2 for spin-degenerate mixing (src_qcm/lattice_model_instance.cpp:624-643)."""
    n = Gc.shape[0]
    return factor * np.log(abs(np.linalg.det(V @ Gc - np.eye(n))))


def hybridization(w, eb, tcb):
    """Gamma_ab(w) = sum_i tcb(a,i) conj(tcb(b,i)) / (w - eb_i) (src_ed/model_instance.hpp:1004-1018)."""
    return (tcb / (w - eb)) @ tcb.conj().T


def cdmft_distance(freqs, weights, G_host, eb, tcb, mixing_factor=2.0):
    """src_qcm/CPT.cpp:519-555 for one cluster, spin-degenerate (mixing 0)."""
    dist = 0.0
    dim = G_host[0].shape[0]
    for i, wv in enumerate(freqs):
        g = hybridization(1j * wv, eb, tcb) + G_host[i]
        dist += np.sum(np.abs(g) ** 2) * weights[i]
    dist *= mixing_factor
    dist *= freqs[1] - freqs[0]
    dist /= dim * dim
    return dist
