"""Test infrastructure (the oracle): a plain-Python restatement of the reference's Abelian
point-group reduction of the Hilbert space, used only by tests/ to check the device's
symmetry-adapted (mixed) basis and its Hamiltonian.

Follows:
  * symmetry_group: src_ed/symmetry_group.cpp:18-31 (constructor), :36-100 (elements,
    products, validity), :115-149 (apply to a bit string: signed permutation, bath
    phases, fermionic sign), :183-352 (build: generator lengths, elements from the
    mixed-radix Index, characters), :384-412 (apply to a two-spin state), :428-454
    (Representative); the Index struct, src_ed/index.hpp:16-95.
  * ED_mixed_basis, unmixed case (fixed N_up, N_down): src_ed/ED_basis.cpp:185-361.
  * HS_one_body_operator: src_ed/Operators/HS_one_body_operator.hpp:8-40, with the
    Hermitian completion of HS_nondiagonal_operator::multiply_add
    (src_ed/Operators/HS_nondiagonal_operator.hpp:93-112), and the density-density
    interaction on the diagonal.

Phases are integers in units of pi/g (as in the reference); only groups whose characters
are real (every generator of order <= 2) are restated here.
"""
import numpy as np


class Index:
    """Mixed-radix index over the generators' orders (src_ed/index.hpp:16-95)."""

    def __init__(self, dim):
        self.dim = list(dim)
        self.tot = int(np.prod(self.dim)) if self.dim else 1

    def digits(self, label):
        out = []
        for d in self.dim:
            out.append(label % d)
            label //= d
        return out

    def dot(self, a, b):
        """the 'hermitian product' of two labels: sum_j a_j b_j tot / dim_j mod tot"""
        da, db = self.digits(a), self.digits(b)
        s = 0
        for j, d in enumerate(self.dim):
            s += (da[j] * db[j] * self.tot) // d
        return s % self.tot


class SymmetryGroup:
    """N orbitals per spin, the first n_sites of them sites. generators: lists of N
    0-based entries: site images for i < L, and for i >= L (bath_irrep) the phase label
    of bath orbital i (in units of 2 pi / g); L = n_sites if bath_irrep else N."""

    def __init__(self, N, n_sites, generators, bath_irrep=False):
        self.N = N
        self.n_sites = n_sites
        self.bath_irrep = bool(bath_irrep) and len(generators) > 0
        self.L = n_sites if self.bath_irrep else N
        gens = [list(map(int, g)) for g in generators]
        if not gens:
            gens = [list(range(N))]
        self.generators = gens
        self._build()

    # group algebra on element vectors (src_ed/symmetry_group.cpp:36-100)
    def identity(self):
        return [i for i in range(self.L)] + [0] * (self.N - self.L)

    def product(self, x, y, g):
        v = [x[y[i]] for i in range(self.L)]
        v += [(x[i] + y[i]) % g if g else 0 for i in range(self.L, self.N)]
        return v

    def is_identity(self, v, full=False):
        if any(v[i] != i for i in range(self.L)):
            return False
        if full and any(v[i] != 0 for i in range(self.L, self.N)):
            return False
        return True

    def _build(self):
        L = self.L
        for k, gen in enumerate(self.generators):
            if len(gen) != self.N or sorted(gen[:L]) != list(range(L)):
                raise ValueError("generator %d is not valid!" % (k + 1))
        # orders of the generators (the bath labels add modulo g, unknown yet: 0)
        dims = []
        for gen in self.generators:
            p = self.identity()
            length = 1
            while True:
                p = [p[gen[i]] for i in range(L)] + p[L:]
                if self.is_identity(p):
                    break
                length += 1
            dims.append(length)
        self.g = int(np.prod(dims))
        if any(d > 2 for d in dims):
            raise ValueError("groups with complex characters are not restated here")
        self.index = Index(dims)
        g = self.g
        elems = []
        for i in range(g):
            p = self.identity()
            for j, gen in enumerate(self.generators):
                for _ in range(self.index.digits(i)[j]):
                    p = self.product(p, gen, g)
            if i != 0 and self.is_identity(p, True):
                raise ValueError("The identity is generated as the non trivial element")
            elems.append(p)
        self.e = elems
        # phi(irrep, elem) = irrep . elem  (units of 2 pi / g)
        self.phi = np.array([[self.index.dot(r, i) for i in range(g)] for r in range(g)], int)

    # action on one spin's bit string (src_ed/symmetry_group.cpp:115-149)
    def apply_word(self, v, b):
        bp = 0
        q = []
        phase = 0
        for i in range(self.L):
            if b >> i & 1:
                bp |= 1 << v[i]
                q.append(v[i])
        for i in range(self.L, self.N):
            if b >> i & 1:
                phase += 2 * v[i]
                bp |= 1 << i
        inv = sum(1 for a in range(len(q)) for c in range(a + 1, len(q)) if q[a] > q[c])
        phase += (inv % 2) * self.g
        return bp, phase

    def apply(self, elem, state):
        """state = up | down << 32 (the reference's binary_state order)"""
        up, dn = state & 0xFFFFFFFF, state >> 32
        v = self.e[elem]
        r, pr = self.apply_word(v, up)
        l, pl = self.apply_word(v, dn)
        return r | (l << 32), pr + pl

    def representative(self, state, irrep):
        """(representative, phase in units of pi/g, orbit length; 0 if the state vanishes)
        (src_ed/symmetry_group.cpp:428-454)"""
        g = self.g
        h = 1
        best, best_phase = state, 0
        for elem in range(1, g):
            s, ph = self.apply(elem, state)
            ph = (ph + 2 * self.phi[elem, irrep]) % (2 * g)
            if s == state:
                h += 1
                if ph != 0 and (2 * g) % ph == 0:
                    return state, 0, 0
            if s < best:
                best, best_phase = s, ph
        return best, best_phase, g // h


def half_states(N, n):
    return [b for b in range(1 << N) if bin(b).count("1") == n]


def mixed_basis(G, nup, ndn, irrep):
    """sorted representatives of the (nup, ndn) sector in irrep (src_ed/ED_basis.cpp:318-359)"""
    out = []
    for up in half_states(G.N, nup):
        for dn in half_states(G.N, ndn):
            s = up | dn << 32
            r, _, length = G.representative(s, irrep)
            if length and r == s:
                out.append(s)
    out.sort()
    return out


def _mask(orb, N):
    return 1 << orb if orb < N else 1 << (orb - N + 32)


def _one_body(s, i, j):
    """the reference's binary_state::one_body: moves the electron at mask i to mask j
    (c^dagger_j c_i) with the sign of the bits strictly between (src_ed/binary_state.cpp:23-28)"""
    if s & j or not s & i:
        return None, 0
    lo, hi = min(i, j), max(i, j)
    inside = hi - (lo << 1)
    sign = -1 if bin(s & inside).count("1") & 1 else 1
    return s - i + j, sign


def mixed_hamiltonian(G, basis, irrep, one_body, interaction):
    """Dense H in the mixed basis. one_body = [(r, c, v)]: 0-based spin orbitals
    (0..2N-1, up first), r <= c, the operators' elements times their coefficients;
    interaction = [(i, j, v)]: density-density terms n_i n_j. As the reference: each
    off-diagonal element is stored in both directions (one_body_operator.hpp:67-76);
    applied to a basis state I (c^dagger_r c_c) it gives an element X of row J; X is kept
    where J > I (lower triangle) or J = I (diagonal) and dropped otherwise
    (HS_nondiagonal_operator::insert), and the lower triangle is completed by its
    transpose (multiply_add)."""
    N = G.N
    index = {s: k for k, s in enumerate(basis)}
    d = len(basis)
    A = np.zeros((d, d))
    H = np.zeros((d, d))
    elems = []
    for (r, c, v) in one_body:
        if r < c:
            elems += [(r, c, v), (c, r, v)]
        elif r == c:
            elems.append((r, c, v))
    for I, s in enumerate(basis):
        _, _, len_in = G.representative(s, irrep)
        for (r, c, v) in elems:
            if r != c:
                sp, sign = _one_body(s, _mask(c, N), _mask(r, N))
                if sign == 0:
                    continue
                rep, phase, len_out = G.representative(sp, irrep)
                J = index.get(rep)
                if J is None or len_out == 0 or J < I:
                    continue
                ph = 1.0 if phase == 0 else -1.0
                X = ph * v * sign * np.sqrt(len_in / len_out)
                if J == I:
                    H[I, I] += X
                else:
                    A[J, I] += X
            elif s & _mask(r, N):
                H[I, I] += v
        for (i, j, v) in interaction:
            if s & _mask(i, N) and s & _mask(j, N):
                H[I, I] += v
    return H + A + A.T
