// Small dense FP64 linear algebra on the host: symmetric tridiagonal and dense symmetric
// eigensolvers (Householder reduction + implicit QL with Wilkinson shifts, the EISPACK
// tred2/tql2 method of Bowdler, Martin, Reinsch and Wilkinson), and complex LU.
// These replace the reference's LAPACK dstev (src_ed/Hamiltonian/Lanczos.cpp:25-52) and
// Eigen eigensystem/inverse/determinant (src_util/matrix.cpp:24-141) on small matrices.
#pragma once

#include <algorithm>
#include <cmath>
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


// Lowest eigenvalue of the leading n x n block of a symmetric band matrix (half-bandwidth
// w; T(i, j) is read for |i - j| <= w only) by inverse iteration with the shift sigma,
// which must lie below the spectrum: then T - sigma is positive definite and its banded
// Cholesky factor exists, and a breakdown means it does not (the caller falls back).
// x: start vector (previous eigenvector, zero-padded), overwritten by the eigenvector.
// Accepted when the residual of the Rayleigh quotient after `iters` solves is below
// 1e-10 max(1, max |T_ii|).
template <class TF>
bool band_lowest(TF T, int n, int w, double sigma, int iters, std::vector<double>& x, double& lam) {
  std::vector<double> Lb((size_t)n * (w + 1), 0.0);  // L(i, i - k) at Lb[i (w+1) + k]
  auto Lget = [&](int i, int k) { return Lb[(size_t)i * (w + 1) + (i - k)]; };
  for (int i = 0; i < n; ++i)
    for (int k = std::max(0, i - w); k <= i; ++k) {
      double sum = T(i, k) - (i == k ? sigma : 0.0);
      for (int m = std::max(std::max(0, i - w), k - w); m < k; ++m) sum -= Lget(i, m) * Lget(k, m);
      if (i == k) {
        if (!(sum > 0.0)) return false;
        Lb[(size_t)i * (w + 1)] = std::sqrt(sum);
      } else {
        Lb[(size_t)i * (w + 1) + (i - k)] = sum / Lget(k, k);
      }
    }
  x.resize(n, 0.0);
  bool any = false;
  for (double v : x) any = any || v != 0.0;
  if (!any) std::fill(x.begin(), x.end(), 1.0);
  std::vector<double> z(n);
  for (int it = 0; it < iters; ++it) {
    for (int i = 0; i < n; ++i) {  // L z = x
      double sum = x[i];
      for (int m = std::max(0, i - w); m < i; ++m) sum -= Lget(i, m) * z[m];
      z[i] = sum / Lget(i, i);
    }
    for (int i = n - 1; i >= 0; --i) {  // L^T y = z
      double sum = z[i];
      for (int m = i + 1; m <= std::min(n - 1, i + w); ++m) sum -= Lget(m, i) * x[m];
      x[i] = sum / Lget(i, i);
    }
    double nrm = 0.0;
    for (double v : x) nrm += v * v;
    nrm = std::sqrt(nrm);
    if (!(nrm > 0.0) || !std::isfinite(nrm)) return false;
    for (double& v : x) v /= nrm;
  }
  double rq = 0.0, res = 0.0, scale = 0.0;
  std::vector<double> tx(n, 0.0);
  for (int i = 0; i < n; ++i) {
    for (int j = std::max(0, i - w); j <= std::min(n - 1, i + w); ++j) tx[i] += T(i, j) * x[j];
    rq += x[i] * tx[i];
    scale = std::max(scale, std::fabs(T(i, i)));
  }
  for (int i = 0; i < n; ++i) res += (tx[i] - rq * x[i]) * (tx[i] - rq * x[i]);
  if (std::sqrt(res) > 1e-10 * std::max(1.0, scale)) return false;
  lam = rq;
  return true;
}

}  // namespace qed
