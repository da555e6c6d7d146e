"""pyqcm_amd — MI355X-native exact-diagonalization impurity solver for pyqcm's cluster
methods (CPT, VCA, CDMFT).

Layers:
  _lib      ctypes binding of libqed.so (C ABI in include/qed.h)
  sector    the Hamiltonian-on-a-sector seam: H.v, Lanczos, band Lanczos, c/c†
  qcm       the qcm.* functions pyqcm's Python layer calls on this path
"""
__version__ = "0.1.0"
