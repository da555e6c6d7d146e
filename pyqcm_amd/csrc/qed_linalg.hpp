// Small dense FP64 linear algebra on the host: symmetric tridiagonal and dense symmetric
// eigensolvers (Householder reduction + implicit QL with Wilkinson shifts, the EISPACK
// tred2/tql2 method of Bowdler, Martin, Reinsch and Wilkinson), and complex LU.
// These replace the reference's LAPACK dstev (src_ed/Hamiltonian/Lanczos.cpp:25-52) and
// Eigen eigensystem/inverse/determinant (src_util/matrix.cpp:24-141) on small matrices.
#pragma once

#include <complex>
#include <vector>

namespace qed {

// Eigen-decomposition of the symmetric tridiagonal matrix (d: diagonal n, e: off-diagonal
// n-1). On exit d holds ascending eigenvalues; if z != nullptr it receives the
// eigenvectors column-major (column k = eigenvector k, n x n).
void tridiag_eigen(int n, double* d, const double* e, double* z);

// Dense real symmetric eigensystem: a (n x n, column-major, overwritten). Ascending
// eigenvalues in w; eigenvectors (columns) in a if vectors.
void sym_eigen(int n, double* a, double* w, bool vectors);

// Complex dense inverse in place (n x n column-major); returns false if singular.
bool cinverse(int n, std::complex<double>* a);
// log |det a| (a is overwritten by its LU factors).
double clogabsdet(int n, std::complex<double>* a);

}  // namespace qed
