"""The ``qcm`` functions pyqcm's Python layer calls on the ED path, backed by libqed.so.

Names, argument meaning, defaults and errors follow the reference binding
(src_python/qcm_ED_wrap.hpp:27-381, src_python/qcm_wrap.hpp:292-910): operator element
labels are 1-based with r <= c, anomalous elements are scaled by 1/2 as there, failures
raise RuntimeError with the library's message.

Beyond the reference's one-frequency calls, every Green-function entry point also takes
an array of frequencies and evaluates them in one device batch.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import call, dp, ip

_instance_values: dict[int, dict] = {}
_instance_model: dict[int, str] = {}
_model_sizes: dict[str, tuple] = {}
_model_ops: dict[str, dict] = {}  # model -> {operator: (type, [(r, c, v)] 0-based, as stored)}


def _b(s):
    return s.encode() if isinstance(s, str) else s


# ---------------------------------------------------------------- models and operators

def new_model(name, n_sites, n_bath, generators=None, bath_irrep=False):
    """ED::new_model (src_ed/qcm_ED.cpp:39-58). generators: a two-dimensional array of
    n_sites + n_bath entries per generator, site images 1-based, then bath entries (phase
    labels if bath_irrep, else 1-based orbital images), converted to 0-based as the
    reference's binding does (src_python/qcm_ED_wrap.hpp:145-163). A model with a
    nontrivial group needs a Hamiltonian_format other than 'F'
    (src_ed/model_instance_base.cpp:86-88)."""
    n_orb = int(n_sites) + int(n_bath)
    if generators is None or len(generators) == 0:
        call("qed_ed_new_model", _b(name), int(n_sites), int(n_bath))
    else:
        gen = np.asarray(generators)
        if gen.ndim != 2:
            raise RuntimeError("generators or positions should form a two-dimensional array!")
        gen = gen.astype(np.int32)
        for j in range(len(gen)):
            if gen.shape[1] != n_orb:
                raise RuntimeError("generator %d should have %d elements" % (j + 1, n_orb))
        gen[:, :int(n_sites)] -= 1
        if not bath_irrep:
            gen[:, int(n_sites):] -= 1
        gen = np.ascontiguousarray(gen)
        call("qed_ed_new_model_generators", _b(name), int(n_sites), int(n_bath), len(gen), ip(gen.ravel()),
             int(bool(bath_irrep)))
    _model_sizes[name] = (int(n_sites), int(n_bath))


def model_sizeC(name):
    ns, nb = ctypes.c_int(), ctypes.c_int()
    call("qed_ed_model_size", _b(name), ctypes.byref(ns), ctypes.byref(nb))
    return ns.value, nb.value, 1


def new_operator(model, name, type, elements):
    """ED::new_operator with the binding's label conversion: elements (r, c, v) are
    1-based, r <= c (src_python/qcm_ED_wrap.hpp:174-215)."""
    r, c, v = [], [], []
    fac = 0.5 if type == "anomalous" else 1.0
    check_upper = type != "general_interaction"
    for i, el in enumerate(elements):
        if len(el) != 3:
            raise RuntimeError("element %d of argument 4 of 'new_operator' should be a 3-tuple" % i)
        a, b, val = int(el[0]), int(el[1]), float(el[2])
        if a > b and check_upper:
            raise RuntimeError("the first index of element %d of argument 4 of 'new_operator' cannot be "
                               "bigger than the second index" % i)
        if a == 0 or b == 0:
            raise RuntimeError("indices in matrix elements of operators are labelled starting at 1, not at 0.")
        r.append(a - 1)
        c.append(b - 1)
        v.append(fac * val)
    r = np.array(r, np.int32)
    c = np.array(c, np.int32)
    v = np.array(v, np.float64)
    call("qed_ed_new_operator", _b(model), _b(name), _b(type), len(v), ip(r), ip(c), dp(v))
    _model_ops.setdefault(model, {})[name] = (type, [(int(a), int(b), float(x)) for a, b, x in zip(r, c, v)])


def new_operator_complex(model, name, type, elements):
    """ED::new_operator with complex elements (src_python/qcm_ED_wrap.hpp:219-261). The
    solver has a real Hilbert space: elements with a nonzero imaginary part are refused."""
    els = []
    for i, el in enumerate(elements):
        if len(el) != 3:
            raise RuntimeError("element %d of argument 4 of 'new_operator' should be a 3-tuple" % i)
        z = complex(el[2])
        if z.imag != 0.0:
            raise RuntimeError("operator %s has complex matrix elements: this solver has a real Hilbert space" % name)
        els.append((el[0], el[1], z.real))
    new_operator(model, name, type, els)


def matrix_elements(model, op):
    """(type, [(r, c, v)]) of an operator, 0-based and as stored (anomalous elements
    halved) (src_python/qcm_ED_wrap.hpp:121-130)."""
    if model not in _model_ops:
        raise RuntimeError("model %s does not exist!" % model)
    if op not in _model_ops[model]:
        raise RuntimeError("operator %s does not exist in model %s" % (op, model))
    t, els = _model_ops[model][op]
    return t, list(els)


def print_models():
    """Prints the registered models and their operators (src_python/qcm_ED_wrap.hpp:275-276)."""
    for name in sorted(_model_ops):
        ns, nb = _model_sizes.get(name, (0, 0))
        print("model %s: %d sites, %d bath orbitals" % (name, ns, nb))
        for op, (t, els) in _model_ops[name].items():
            print("  %s (%s): %d elements" % (op, t, len(els)))


def new_model_instanceC(name, values, sectors, label=0):
    """ED::new_model_instance (src_ed/qcm_ED.cpp:127-158)."""
    names = list(values.keys())
    arr = (ctypes.c_char_p * len(names))(*[_b(n) for n in names])
    vals = np.array([float(values[n]) for n in names], np.float64)
    call("qed_ed_new_model_instance", _b(name), len(names), arr, dp(vals), _b(sectors), int(label))
    _instance_values[int(label)] = dict(values)
    _instance_model[int(label)] = name


def parametersC(label=0):
    """(dict of parameters, model name) of an instance (src_python/qcm_ED_wrap.hpp:263-273)."""
    if int(label) not in _instance_values:
        raise RuntimeError("The label %d is out of range." % int(label))
    return dict(sorted(_instance_values[int(label)].items())), _instance_model[int(label)]


def mixingC(label=0):
    """The Hilbert-space mixing of an instance (src_python/qcm_ED_wrap.hpp:132-135)."""
    m = ctypes.c_int()
    call("qed_ed_mixing", int(label), ctypes.byref(m))
    return m.value


def ED_complex_HS(label=0):
    """0: the Hilbert space of this solver is real (src_python/qcm_ED_wrap.hpp:27-30)."""
    return 0


def interactions(label=0):
    """Density-density interaction elements (r, c, value x parameter) of the instance, in
    the order of the parameter names (model_instance.hpp:1528-1541)."""
    if int(label) not in _instance_values:
        raise RuntimeError("The label %d is out of range." % int(label))
    ops = _model_ops.get(_instance_model[int(label)], {})
    out = []
    for name, val in sorted(_instance_values[int(label)].items()):
        t, els = ops.get(name, (None, []))
        if t == "interaction":
            out.extend((r, c, v * float(val)) for r, c, v in els)
    return out


def erase_model_instance(label=0):
    call("qed_ed_erase_model_instance", int(label))
    _instance_values.pop(int(label), None)


# ---------------------------------------------------------------- solving

def ground_state_solve(label=0):
    """Returns (energy, Hilbert space sector string)."""
    e = np.zeros(1)
    buf = ctypes.create_string_buffer(4096)
    call("qed_ed_ground_state_solve", int(label), dp(e), buf, 4096)
    return float(e[0]), buf.value.decode()


def batch_solve(labels, graph=True, chunk=16):
    """Ground states and Green functions of several instances in one batched device solve
    (same results as solving them one after the other). Returns the solver statistics
    (lockstep Lanczos iterations, band Lanczos steps, H.v block size, H.v grid)."""
    labels = np.ascontiguousarray(labels, np.uint64)
    st = np.zeros(4, np.int32)
    call("qed_ed_batch_solve_instances", len(labels),
         labels.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), int(bool(graph)), int(chunk), ip(st))
    return {"lanczos_iterations": int(st[0]), "band_steps": int(st[1]), "hv_block": int(st[2]),
            "hv_grid": int(st[3])}


def Green_function_solveC(label=0):
    call("qed_ed_green_function_solve", int(label))


def Green_function_dimensionC(label=0):
    d = ctypes.c_int()
    call("qed_ed_dim_gf", int(label), ctypes.byref(d))
    return d.value


def _zarr(z):
    z = np.atleast_1d(np.asarray(z, np.complex128))
    return z, np.ascontiguousarray(np.stack([z.real, z.imag], -1).ravel())


def _batched(fn, z, spin_down, label):
    zs, zz = _zarr(z)
    d = Green_function_dimensionC(label)
    out = np.zeros(2 * len(zs) * d * d)
    call(fn, int(label), len(zs), dp(zz), int(spin_down), dp(out))
    g = out.view(np.complex128).reshape(len(zs), d, d)
    # column-major per matrix -> (row, col) arrays
    g = np.transpose(g, (0, 2, 1)).copy()
    return g[0] if np.ndim(z) == 0 else g


def Green_function(z, spin_down=0, label=0, blocks=0):
    """G(z) (dim_GF x dim_GF); z may be an array of frequencies (batched)."""
    return _batched("qed_ed_green_function", z, spin_down, label)


def self_energyC(z, spin_down=0, label=0):
    return _batched("qed_ed_self_energy", z, spin_down, label)


def hybridization_functionC(z, spin_down=0, label=0):
    return _batched("qed_ed_hybridization_function", z, spin_down, label)


def hopping_matrix(spin_down=0, diag=0, label=0, full=0):
    """t_c (dim_GF x dim_GF); with full, the matrix of cluster and bath with the bath
    diagonalized (diag) or as given (src_python/qcm_ED_wrap.hpp:94-106)."""
    d = Green_function_dimensionC(label)
    if full:
        ns, nb = model_sizeC(_instance_model[int(label)])[:2]
        d += nb * (d // ns)
        out = np.zeros(2 * d * d)
        call("qed_ed_hopping_matrix_full", int(label), int(spin_down), int(diag), dp(out))
        return out.view(np.complex128).reshape(d, d).T.copy()
    out = np.zeros(2 * d * d)
    call("qed_ed_hopping_matrix", int(label), int(spin_down), dp(out))
    return out.view(np.complex128).reshape(d, d).T.copy()


def bath(label=0, spin_down=0):
    """Diagonalized bath: energies (nb) and cluster-bath hybridization (d x nb); with the
    spin-flip or anomalous mixings both count the spinor components (d = n_mixed x n_sites,
    nb = n_mixed x n_bath)."""
    d = Green_function_dimensionC(label)
    ns, nb, _ = model_sizeC(_instance_model[int(label)])
    nb *= d // ns
    eb = np.zeros(max(nb, 1))
    t = np.zeros(max(d * nb, 1))
    call("qed_ed_bath", int(label), int(spin_down), dp(eb), dp(t))
    return eb[:nb], t[:d * nb].reshape(nb, d).T.copy()


def qmatrix(label=0, spin_down=0):
    """Lehmann representation: (eigenvalues M, M x L matrix) as the reference binding."""
    M = ctypes.c_int()
    call("qed_ed_qmatrix", int(label), int(spin_down), ctypes.byref(M), None, None)
    m = M.value
    d = Green_function_dimensionC(label)
    e = np.zeros(max(m, 1))
    Q = np.zeros(max(m * d, 1))
    call("qed_ed_qmatrix", int(label), int(spin_down), ctypes.byref(M), dp(e), dp(Q))
    return e[:m].copy(), Q[:m * d].reshape(m, d).astype(np.complex128)


def Potthoff_functional(label=0):
    """ED part of the Potthoff functional: E0 + SEF_bath (src_ed/qcm_ED.cpp:330-334)."""
    w = np.zeros(1)
    call("qed_ed_potthoff_functional", int(label), dp(w))
    return float(w[0])


def cluster_averages(label=0):
    """dict name -> (average, variance) over the active operators."""
    names = list(_instance_values.get(int(label), {}).keys())
    if not names:
        return {}
    arr = (ctypes.c_char_p * len(names))(*[_b(n) for n in names])
    ave = np.zeros(len(names))
    var = np.zeros(len(names))
    call("qed_ed_cluster_averages", int(label), len(names), arr, dp(ave), dp(var))
    return {n: (float(a), float(v)) for n, a, v in zip(names, ave, var)}


def Green_function_average(label=0, spin_down=0):
    """<c^dagger_i c_j> from the Green-function representation, d x d
    (src_python/qcm_ED_wrap.hpp:44-50, src_ed/model_instance.hpp:643-654)."""
    d = Green_function_dimensionC(label)
    out = np.zeros(2 * d * d)
    call("qed_ed_green_function_average", int(label), int(spin_down), dp(out))
    # column-major M, returned with the reference binding's row-major reading of it
    return out.view(np.complex128).reshape(d, d).copy()


def Green_function_density(label=0):
    """Cluster density from the trace of the Green function average
    (src_python/qcm_ED_wrap.hpp:52-55, src_ed/model_instance.hpp:661-675)."""
    out = np.zeros(1)
    call("qed_ed_green_function_density", int(label), dp(out))
    return float(out[0])


def fidelity(label1, label2):
    """Overlap of the low-energy states of two instances (src_python/qcm_ED_wrap.hpp:378-381)."""
    out = np.zeros(1)
    call("qed_ed_fidelity", int(label1), int(label2), dp(out))
    return float(out[0])


def tr_sigma_inf(label=0):
    """Trace of the self-energy extrapolated to infinite frequency
    (src_ed/model_instance.hpp:1277-1312)."""
    out = np.zeros(1)
    call("qed_ed_tr_sigma_inf", int(label), dp(out))
    return float(out[0])


def hybridization(label=0, spin_down=0):
    """Lehmann form of the hybridization function as the reference binding returns it:
    (bath energies, the column-major n_sites x n_bath cluster-bath matrix in the bath's
    eigenbasis read row-major into shape (n_sites, n_bath)), of the spinor's dimensions
    for the spin-flip and anomalous mixings (src_python/qcm_ED_wrap.hpp:351-362)."""
    d = Green_function_dimensionC(label)
    ns, nb = model_sizeC(_instance_model[int(label)])[:2]
    nb *= d // ns  # the spinor's bath orbitals for the spin-flip and anomalous mixings
    E = np.zeros(max(nb, 1))
    T = np.zeros(max(d * nb, 1))
    call("qed_ed_hybridization_lehmann", int(label), int(spin_down), dp(E), dp(T))
    return E[:nb].copy(), T[:d * nb].astype(np.complex128).reshape(d, nb)


def combined_mcf(label=0, k=None):
    """(W, [A_j], [B_j]) of the combined matrix continued fraction of the cluster Green
    function, G(z) = W^T F_0(z) W (src_python/qcm_wrap.hpp:502-522). The lattice form (k
    given) is not provided."""
    if k is not None:
        raise RuntimeError("combined_mcf at a lattice wavevector is not provided by this solver")
    p, M, q = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    call("qed_ed_get_combined_mcf", int(label), 0, ctypes.byref(p), ctypes.byref(M), ctypes.byref(q), None, None, None)
    pp = p.value * p.value
    A = np.zeros(max(M.value * pp, 1))
    B = np.zeros(max(M.value * pp, 1))
    W = np.zeros(max(p.value * q.value, 1))
    call("qed_ed_get_combined_mcf", int(label), 0, ctypes.byref(p), ctypes.byref(M), ctypes.byref(q), dp(A), dp(B), dp(W))
    blk = lambda X, j: X[j * pp:(j + 1) * pp].reshape(p.value, p.value).T.astype(np.complex128)
    Wm = W[:p.value * q.value].reshape(q.value, p.value).T.astype(np.complex128)
    return Wm, [blk(A, j) for j in range(M.value)], [blk(B, j) for j in range(M.value)]


def print_wavefunction(label=0):
    """Text of the low-energy states: header, and amplitudes with basis states when the
    sector is at most max_dim_print (src_python/qcm_ED_wrap.hpp:335-337)."""
    n = ctypes.c_int()
    call("qed_ed_print_wavefunction", int(label), None, 0, ctypes.byref(n))
    buf = ctypes.create_string_buffer(n.value)
    call("qed_ed_print_wavefunction", int(label), buf, n.value, ctypes.byref(n))
    return buf.value.decode()


def density_matrix(sites, label=0):
    """(density matrix, basis) of the subsystem `sites` (0-based) in the first state: empty
    for the factorized format 'F', the reduced density matrix otherwise
    (src_python/qcm_ED_wrap.hpp:32-42, src_ed/model_instance.hpp:1547-1633)."""
    s = np.ascontiguousarray(sites, np.int32)
    dim = ctypes.c_int()
    call("qed_ed_density_matrix", int(label), len(s), ip(s), ctypes.byref(dim), None, None)
    d = dim.value
    if d == 0:
        return np.zeros((0, 0), np.complex128), np.zeros(0, np.uint64)
    rho = np.zeros(2 * d * d)
    st = np.zeros(d, np.uint64)
    call("qed_ed_density_matrix", int(label), len(s), ip(s), ctypes.byref(dim), dp(rho),
         st.ctypes.data_as(ctypes.c_void_p))
    return rho.view(np.complex128).reshape(d, d).T.copy(), st


def susceptibility(op, w, label=0):
    """Dynamic susceptibility of operator `op` at the complex frequencies w
    (src_python/qcm_ED_wrap.hpp:327-333, src_ed/model_instance.hpp:1024-1058)."""
    w = np.ascontiguousarray(np.atleast_1d(np.asarray(w, dtype=np.complex128)))
    out = np.zeros(len(w), np.complex128)
    call("qed_ed_susceptibility", int(label), _b(op), len(w), dp(w.view(np.float64)), dp(out.view(np.float64)))
    return out


def susceptibility_poles(op, label=0):
    """Lehmann form of the dynamic susceptibility: an (n, 2) array of (pole, residue) rows
    (src_python/qcm_ED_wrap.hpp:318-325, src_ed/model_instance.hpp:1066-1113)."""
    n = ctypes.c_int()
    cap = 1024
    while True:
        out = np.zeros(2 * cap)
        call("qed_ed_susceptibility_poles", int(label), _b(op), cap, ctypes.byref(n), dp(out))
        if n.value <= cap:
            return out[:2 * n.value].reshape(n.value, 2).copy()
        cap = n.value


def write_instance_to_hdf5(filename, group, label=0):
    """Writes the solved instance to group `group` of an HDF5 file (created if absent; the
    group is replaced if present), in the reference's layout
    (src_python/qcm_ED_wrap.hpp:364-369, src_ed/qcm_ED.cpp:373-389)."""
    call("qed_ed_write_instance_hdf5", _b(str(filename)), int(label), _b(group))


def read_instance_from_hdf5(filename, group, label=0):
    """Replaces the solution of instance `label` by the one stored in group `group`
    (src_python/qcm_ED_wrap.hpp:371-376, src_ed/qcm_ED.cpp:392-403)."""
    call("qed_ed_read_instance_hdf5", _b(str(filename)), int(label), _b(group))


# ---------------------------------------------------------------- global parameters

_BOOL = {"verb_ED", "no_degenerate_BL", "continued_fraction", "block_Lanczos_QR", "combine_mcf"}


def set_global_parameter(name, value=None):
    """Same typing rules as the reference binding: None -> bool true; int -> bool or int;
    float -> double; str -> char."""
    if value is None:
        call("qed_ed_set_global_bool", _b(name), 1)
    elif isinstance(value, bool) or (isinstance(value, int) and name in _BOOL):
        call("qed_ed_set_global_bool", _b(name), int(bool(value)))
    elif isinstance(value, int):
        call("qed_ed_set_global_int", _b(name), int(value))
    elif isinstance(value, float):
        call("qed_ed_set_global_double", _b(name), float(value))
    elif isinstance(value, str):
        call("qed_ed_set_global_char", _b(name), ctypes.c_char(value.encode()[0]))
    else:
        raise RuntimeError("unknown type of global_parameter")


set_global_parameterC = set_global_parameter


def get_global_double(name):
    v = np.zeros(1)
    call("qed_ed_get_global_double", _b(name), dp(v))
    return float(v[0])
