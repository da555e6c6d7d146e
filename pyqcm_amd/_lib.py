"""ctypes binding of ``libqed.so`` (the C ABI declared in ``include/qed.h``).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C pyqcm_amd/csrc``).
There is no fallback: if the HIP library is missing or no GPU is present, every call
raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqed.so")

c_int, c_double, c_char, c_char_p, c_void_p = (ctypes.c_int, ctypes.c_double, ctypes.c_char,
                                               ctypes.c_char_p, ctypes.c_void_p)
c_u32, c_u64, c_float = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float
P = ctypes.POINTER
dptr, fptr, iptr, u32ptr = P(c_double), P(c_float), P(c_int), P(c_u32)

# (name, restype, argtypes) for every entry point of include/qed.h
SIGNATURES = [
    ("qed_init", c_int, [c_int]),
    ("qed_last_error", c_char_p, []),
    ("qed_sync", c_int, []),
    ("qed_alloc", c_int, [P(c_void_p), c_u64]),
    ("qed_free", c_int, [c_void_p]),
    ("qed_memcpy_h2d", c_int, [c_void_p, c_void_p, c_u64]),
    ("qed_memcpy_d2h", c_int, [c_void_p, c_void_p, c_u64]),
    ("qed_memcpy_d2d", c_int, [c_void_p, c_void_p, c_u64]),
    ("qed_memset", c_int, [c_void_p, c_int, c_u64]),
    ("qed_halfbasis", c_int, [c_int, c_int, u32ptr, P(c_u32)]),
    ("qed_halfbasis_rank", c_int, [c_int, c_int, u32ptr, c_u64, u32ptr]),
    ("qed_sector_create", c_int, [c_int, c_int, c_int, c_int, iptr, dptr, iptr, iptr, iptr, dptr,
                                  P(c_void_p)]),
    ("qed_sector_destroy", c_int, [c_void_p]),
    ("qed_sector_dims", c_int, [c_void_p, P(c_u64), P(c_u32), P(c_u32)]),
    ("qed_hv", c_int, [c_void_p, c_void_p, c_void_p]),
    ("qed_hv_host", c_int, [c_void_p, dptr, dptr]),
    ("qed_diag", c_int, [c_void_p, c_void_p]),
    ("qed_start_vector", c_int, [c_void_p, c_int, c_void_p]),
    ("qed_lanczos_gs", c_int, [c_void_p, c_double, c_int, c_void_p, dptr, iptr, dptr, dptr]),
    ("qed_lanczos_gs_noalloc", c_int, [c_void_p, c_double, c_int, c_void_p, dptr, iptr, dptr, dptr]),
    ("qed_mixed_sector_create", c_int, [c_int, c_int, c_int, c_int, iptr, dptr, iptr, iptr, iptr, dptr, c_int,
                                        c_int, iptr, c_int, c_int, P(c_void_p)]),
    ("qed_mixed_basis", c_int, [c_void_p, P(c_u64), P(c_u64)]),
    ("qed_general_sector_create", c_int, [c_int, c_int, iptr, c_int, iptr, dptr, iptr, iptr, iptr, dptr,
                                          P(c_void_p)]),
    ("qed_general_basis", c_int, [c_void_p, P(c_u64), P(c_u64)]),
    ("qed_general_apply", c_int, [c_void_p, c_void_p, c_int, c_int, c_double, c_void_p, c_void_p]),
    ("qed_mixed_apply_c", c_int, [c_void_p, c_void_p, c_int, c_int, c_int, dptr, c_void_p, c_void_p]),
    ("qed_mixed_group", c_int, [c_void_p, iptr, iptr, iptr]),
    ("qed_lanczos_green", c_int, [c_void_p, c_void_p, c_double, c_int, dptr, dptr, iptr]),
    ("qed_band_lanczos", c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_double, c_double, dptr,
                                 dptr, c_int, iptr]),
    ("qed_davidson", c_int, [c_void_p, c_int, c_double, c_int, dptr, c_void_p, iptr]),
    ("qed_apply_c", c_int, [c_int, c_int, c_int, c_int, c_int, c_double, c_void_p, c_void_p]),
    ("qed_qmatrix_gf", c_int, [c_int, c_int, dptr, dptr, c_int, dptr, dptr]),
    ("qed_hv_batch", c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, dptr]),
    ("qed_block_lanczos", c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_double, c_double, c_int, dptr,
                                  dptr, c_int, iptr]),
    ("qed_mcf_eval", c_int, [c_int, c_int, c_int, dptr, dptr, dptr, c_int, dptr, c_int, dptr]),
    ("qed_cf_eval", c_int, [c_int, iptr, iptr, dptr, dptr, c_int, dptr, dptr]),
    ("qed_cpt_batch", c_int, [c_int, c_int, dptr, dptr, c_int, dptr, dptr, c_int, dptr, c_int, dptr]),
    ("qed_sef_batch", c_int, [c_int, c_int, dptr, dptr, c_int, dptr, dptr, c_int, dptr, dptr]),
    ("qed_cdmft_distance_batch", c_int, [c_int, c_int, c_int, dptr, dptr, dptr, c_int, dptr, dptr,
                                         c_double, dptr]),
    ("qed_cdmft_distance_multi", c_int, [c_int, c_int, c_int, dptr, dptr, c_int, dptr, c_int, iptr,
                                         dptr, dptr, c_double, dptr]),
    ("qed_cdmft_fit_batch", c_int, [c_int, c_int, c_int, dptr, dptr, c_int, dptr, c_int, iptr, c_int,
                                    dptr, dptr, dptr, c_double, c_double, c_double, c_int, c_double,
                                    dptr, dptr, iptr, iptr]),
    ("qed_ed_new_model", c_int, [c_char_p, c_int, c_int]),
    ("qed_ed_new_model_generators", c_int, [c_char_p, c_int, c_int, c_int, iptr, c_int]),
    ("qed_ed_new_operator", c_int, [c_char_p, c_char_p, c_char_p, c_int, iptr, iptr, dptr]),
    ("qed_ed_new_model_instance", c_int, [c_char_p, c_int, P(c_char_p), dptr, c_char_p, c_u64]),
    ("qed_ed_erase_model_instance", c_int, [c_u64]),
    ("qed_ed_ground_state_solve", c_int, [c_u64, dptr, c_char_p, c_int]),
    ("qed_ed_green_function_solve", c_int, [c_u64]),
    ("qed_ed_green_function", c_int, [c_u64, c_int, dptr, c_int, dptr]),
    ("qed_ed_self_energy", c_int, [c_u64, c_int, dptr, c_int, dptr]),
    ("qed_ed_hybridization_function", c_int, [c_u64, c_int, dptr, c_int, dptr]),
    ("qed_ed_hopping_matrix", c_int, [c_u64, c_int, dptr]),
    ("qed_ed_hopping_matrix_full", c_int, [c_u64, c_int, c_int, dptr]),
    ("qed_ed_bath", c_int, [c_u64, c_int, dptr, dptr]),
    ("qed_ed_qmatrix", c_int, [c_u64, c_int, iptr, dptr, dptr]),
    ("qed_ed_potthoff_functional", c_int, [c_u64, dptr]),
    ("qed_ed_model_size", c_int, [c_char_p, iptr, iptr]),
    ("qed_ed_dim_gf", c_int, [c_u64, iptr]),
    ("qed_ed_cluster_averages", c_int, [c_u64, c_int, P(c_char_p), dptr, dptr]),
    ("qed_ed_mixing", c_int, [c_u64, P(c_int)]),
    ("qed_ksum_last_ms", c_int, [dptr]),
    ("qed_ed_density_matrix", c_int, [c_u64, c_int, iptr, P(c_int), c_void_p, c_void_p]),
    ("qed_ed_print_wavefunction", c_int, [c_u64, c_char_p, c_int, P(c_int)]),
    ("qed_ed_get_combined_mcf", c_int, [c_u64, c_int, P(c_int), P(c_int), P(c_int), dptr, dptr, dptr]),
    ("qed_ed_green_function_average", c_int, [c_u64, c_int, dptr]),
    ("qed_ed_green_function_density", c_int, [c_u64, dptr]),
    ("qed_ed_fidelity", c_int, [c_u64, c_u64, dptr]),
    ("qed_ed_tr_sigma_inf", c_int, [c_u64, dptr]),
    ("qed_ed_hybridization_lehmann", c_int, [c_u64, c_int, dptr, dptr]),
    ("qed_ed_susceptibility", c_int, [c_u64, c_char_p, c_int, dptr, dptr]),
    ("qed_ed_susceptibility_poles", c_int, [c_u64, c_char_p, c_int, P(c_int), dptr]),
    ("qed_ed_write_instance_hdf5", c_int, [c_char_p, c_u64, c_char_p]),
    ("qed_ed_read_instance_hdf5", c_int, [c_char_p, c_u64, c_char_p]),
    ("qed_ed_set_global_int", c_int, [c_char_p, c_int]),
    ("qed_ed_set_global_double", c_int, [c_char_p, c_double]),
    ("qed_ed_set_global_bool", c_int, [c_char_p, c_int]),
    ("qed_ed_set_global_char", c_int, [c_char_p, c_char]),
    ("qed_ed_get_global_double", c_int, [c_char_p, dptr]),
    ("qed_ed_batch_solve_instances", c_int, [c_int, P(c_u64), c_int, c_int, iptr]),
    ("qed_batch_solve", c_int, [c_char_p, c_int, P(c_char_p), c_int, dptr, P(c_char_p), c_int, c_int,
                                dptr, iptr, dptr, dptr]),
    ("qed_mlp_create", c_int, [c_int, c_int, c_int, c_u64, P(c_void_p)]),
    ("qed_mlp_destroy", c_int, [c_void_p]),
    ("qed_mlp_get_params", c_int, [c_void_p, fptr, iptr]),
    ("qed_mlp_set_params", c_int, [c_void_p, fptr, c_int]),
    ("qed_mlp_forward", c_int, [c_void_p, c_int, fptr, fptr]),
    ("qed_mlp_forward_bf16", c_int, [c_void_p, c_int, fptr, fptr]),
    ("qed_mlp_loss_grad", c_int, [c_void_p, c_int, fptr, fptr, fptr, fptr]),
    ("qed_mlp_train", c_int, [c_void_p, c_int, fptr, fptr, c_int, c_int, c_float, c_u64, fptr]),
]

# internal entry points used by tests and the bench (not part of include/qed.h)
EXTRA = [
    ("qed_hv_variant", c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    ("qed_reset_start_vector_state", c_int, []),
    ("qed_blas_dot", c_int, [c_void_p, c_void_p, c_u64, dptr]),
    ("qed_util_tridiag_eigen", c_int, [c_int, dptr, dptr, dptr]),
    ("qed_util_sym_eigen", c_int, [c_int, dptr, dptr, c_int]),
    ("qed_hv_timed", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, dptr]),
    ("qed_lanczos_steps_timed", c_int, [c_void_p, c_void_p, c_int, dptr]),
    ("qed_batch_last_timings", c_int, [dptr, c_int]),
    ("qed_stream_rmw_timed", c_int, [c_void_p, c_void_p, c_u64, c_int, dptr]),
    ("qed_util_normal_stream", c_int, [c_int, c_u64, c_int, c_int, dptr]),
]

_lib = None


class QEDError(RuntimeError):
    pass


def load():
    """Loads libqed.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise QEDError("libqed.so is not built; run __graft_entry__.build() or make -C pyqcm_amd/csrc")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES + EXTRA:
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def call(name, *args):
    L = load()
    rc = getattr(L, name)(*args)
    if rc != 0:
        raise QEDError(L.qed_last_error().decode(errors="replace"))
    return rc


def dp(a):
    return a.ctypes.data_as(dptr)


def fp(a):
    return a.ctypes.data_as(fptr)


def ip(a):
    return a.ctypes.data_as(iptr)


class DeviceArray:
    """A device (HBM) buffer of float64 owned by Python."""

    def __init__(self, n, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        self.ptr = c_void_p()
        call("qed_alloc", ctypes.byref(self.ptr), self.nbytes)

    @classmethod
    def from_host(cls, a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype)
        d = cls(a.size, dtype)
        if a.size:
            call("qed_memcpy_h2d", d.ptr, a.ctypes.data_as(c_void_p), d.nbytes)
        return d

    def zeros(self):
        if self.nbytes:
            call("qed_memset", self.ptr, 0, self.nbytes)
        return self

    def to_host(self):
        a = np.empty(self.n, self.dtype)
        if self.n:
            call("qed_memcpy_d2h", a.ctypes.data_as(c_void_p), self.ptr, self.nbytes)
        return a

    def offset(self, k):
        """Address of element k (for views into stacked vectors)."""
        return c_void_p(self.ptr.value + k * self.dtype.itemsize)

    def free(self):
        if self.ptr and self.ptr.value:
            call("qed_free", self.ptr)
            self.ptr = c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
