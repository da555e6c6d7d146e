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


// Lowest eigenvalue of the leading n x n block of a symmetric matrix that is banded
// (half-bandwidth w) apart from a few border indices whose rows and columns may be full
// (T(i, j) is read for |i - j| <= w and for rows/columns in `border`, sorted ascending),
// by inverse iteration with the shift sigma, which must lie below the spectrum: then
// T - sigma is positive definite, and a failed factorization means it is not (the caller
// falls back). The factorization is block Cholesky with the border ordered last: banded
// Cholesky of the core, the border columns solved through it, Cholesky of the border's
// Schur complement. x: start vector (previous eigenvector, zero-padded), overwritten by
// the eigenvector. Accepted when the residual of the Rayleigh quotient after `iters`
// solves is below 1e-10 max(1, max |T_ii|).
template <class TF>
bool band_lowest(TF T, int n, int w, double sigma, int iters, std::vector<double>& x, double& lam,
                 const std::vector<int>& border = {}) {
  const int nb = (int)border.size();
  std::vector<char> isb(n, 0);
  for (int k : border) {
    if (k < 0 || k >= n) return false;
    isb[k] = 1;
  }
  std::vector<int> core;
  core.reserve(n);
  for (int i = 0; i < n; ++i)
    if (!isb[i]) core.push_back(i);
  const int m = (int)core.size();
  // core entries keep their half-bandwidth in the compressed numbering
  auto A = [&](int i, int j) { return T(core[i], core[j]) - (i == j ? sigma : 0.0); };
  auto inband = [&](int i, int j) { return std::abs(core[i] - core[j]) <= w; };
  std::vector<double> Lb((size_t)m * (w + 1), 0.0);  // L(i, i - k) at Lb[i (w+1) + k]
  auto Lget = [&](int i, int k) { return Lb[(size_t)i * (w + 1) + (i - k)]; };
  for (int i = 0; i < m; ++i)
    for (int k = std::max(0, i - w); k <= i; ++k) {
      double sum = inband(i, k) ? A(i, k) : 0.0;
      for (int q = std::max(std::max(0, i - w), k - w); q < k; ++q) sum -= Lget(i, q) * Lget(k, q);
      if (i == k) {
        if (!(sum > 0.0)) return false;
        Lb[(size_t)i * (w + 1)] = std::sqrt(sum);
      } else {
        Lb[(size_t)i * (w + 1) + (i - k)] = sum / Lget(k, k);
      }
    }
  auto lsolve = [&](std::vector<double>& v) {  // v <- L^-1 v (core)
    for (int i = 0; i < m; ++i) {
      double sum = v[i];
      for (int q = std::max(0, i - w); q < i; ++q) sum -= Lget(i, q) * v[q];
      v[i] = sum / Lget(i, i);
    }
  };
  auto ltsolve = [&](std::vector<double>& v) {  // v <- L^-T v (core)
    for (int i = m - 1; i >= 0; --i) {
      double sum = v[i];
      for (int q = i + 1; q <= std::min(m - 1, i + w); ++q) sum -= Lget(q, i) * v[q];
      v[i] = sum / Lget(i, i);
    }
  };
  // Y = L^-1 C (m x nb), S = E - Y^T Y, Cholesky S = Ls Ls^T
  std::vector<std::vector<double>> Y(nb, std::vector<double>(m));
  for (int b = 0; b < nb; ++b) {
    for (int i = 0; i < m; ++i) Y[b][i] = T(core[i], border[b]);
    lsolve(Y[b]);
  }
  std::vector<double> Ls((size_t)nb * nb, 0.0);
  for (int b1 = 0; b1 < nb; ++b1)
    for (int b2 = 0; b2 <= b1; ++b2) {
      double sum = T(border[b1], border[b2]) - (b1 == b2 ? sigma : 0.0);
      for (int i = 0; i < m; ++i) sum -= Y[b1][i] * Y[b2][i];
      for (int q = 0; q < b2; ++q) sum -= Ls[b1 * nb + q] * Ls[b2 * nb + q];
      if (b1 == b2) {
        if (!(sum > 0.0)) return false;
        Ls[b1 * nb + b1] = std::sqrt(sum);
      } else {
        Ls[b1 * nb + b2] = sum / Ls[b2 * nb + b2];
      }
    }
  x.resize(n, 0.0);
  bool any = false;
  for (double v : x) any = any || v != 0.0;
  if (!any) std::fill(x.begin(), x.end(), 1.0);
  std::vector<double> u(m), ub(nb);
  for (int it = 0; it < iters; ++it) {
    for (int i = 0; i < m; ++i) u[i] = x[core[i]];
    lsolve(u);
    for (int b = 0; b < nb; ++b) {
      double sum = x[border[b]];
      for (int i = 0; i < m; ++i) sum -= Y[b][i] * u[i];
      for (int q = 0; q < b; ++q) sum -= Ls[b * nb + q] * ub[q];
      ub[b] = sum / Ls[b * nb + b];
    }
    for (int b = nb - 1; b >= 0; --b) {
      double sum = ub[b];
      for (int q = b + 1; q < nb; ++q) sum -= Ls[q * nb + b] * ub[q];
      ub[b] = sum / Ls[b * nb + b];
    }
    for (int b = 0; b < nb; ++b)
      for (int i = 0; i < m; ++i) u[i] -= Y[b][i] * ub[b];
    ltsolve(u);
    for (int i = 0; i < m; ++i) x[core[i]] = u[i];
    for (int b = 0; b < nb; ++b) x[border[b]] = ub[b];
    double nrm = 0.0;
    for (double v : x) nrm += v * v;
    nrm = std::sqrt(nrm);
    if (!(nrm > 0.0) || !std::isfinite(nrm)) return false;
    for (double& v : x) v /= nrm;
  }
  // T x: band entries between core indices, full rows and columns of the border
  double rq = 0.0, res = 0.0, scale = 0.0;
  std::vector<double> tx(n, 0.0);
  for (int i = 0; i < n; ++i) {
    if (isb[i]) continue;
    for (int j = std::max(0, i - w); j <= std::min(n - 1, i + w); ++j)
      if (!isb[j]) tx[i] += T(i, j) * x[j];
  }
  for (int k : border)
    for (int j = 0; j < n; ++j) {
      tx[k] += T(k, j) * x[j];
      if (!isb[j]) tx[j] += T(j, k) * x[k];
    }
  for (int i = 0; i < n; ++i) {
    rq += x[i] * tx[i];
    scale = std::max(scale, std::fabs(T(i, i)));
  }
  for (int i = 0; i < n; ++i) res += (tx[i] - rq * x[i]) * (tx[i] - rq * x[i]);
  if (std::sqrt(res) > 1e-10 * std::max(1.0, scale)) return false;
  lam = rq;
  return true;
}

}  // namespace qed
