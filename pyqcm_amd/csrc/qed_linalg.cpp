#include "qed_linalg.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace qed {

namespace {

inline double sign_of(double a, double b) { return b >= 0.0 ? std::fabs(a) : -std::fabs(a); }

// Implicit QL iterations on a symmetric tridiagonal matrix with diagonal d[0..n) and
// subdiagonal e[0..n-1) stored in e[1..n) on entry of the classic formulation; here e is
// laid out as e[i] = T(i, i+1), e[n-1] = 0. Rotations are accumulated into z (column-major)
// when z != nullptr.
void ql_implicit(int n, double* d, double* e, double* z, int ldz) {
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    int m;
    do {
      for (m = l; m < n - 1; ++m) {
        double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m != l) {
        if (++iter > 60) throw std::runtime_error("tridiagonal eigensolver did not converge");
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + sign_of(r, g));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        bool underflow = false;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i];
          double b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            underflow = true;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          if (z) {
            double* zi = z + (size_t)i * ldz;
            double* zi1 = z + (size_t)(i + 1) * ldz;
            for (int k = 0; k < n; ++k) {
              double t = zi1[k];
              zi1[k] = s * zi[k] + c * t;
              zi[k] = c * zi[k] - s * t;
            }
          }
        }
        if (underflow) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
}

void sort_eigen(int n, double* d, double* z, int ldz) {
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
  std::vector<double> d2(n);
  for (int i = 0; i < n; ++i) d2[i] = d[idx[i]];
  std::copy(d2.begin(), d2.end(), d);
  if (z) {
    std::vector<double> z2((size_t)n * ldz);
    for (int i = 0; i < n; ++i)
      std::copy(z + (size_t)idx[i] * ldz, z + (size_t)idx[i] * ldz + n, z2.begin() + (size_t)i * ldz);
    std::copy(z2.begin(), z2.end(), z);
  }
}

}  // namespace

void tridiag_eigen(int n, double* d, const double* e_in, double* z) {
  if (n <= 0) return;
  std::vector<double> e(n, 0.0);
  for (int i = 0; i + 1 < n; ++i) e[i] = e_in[i];
  if (z) {
    std::fill(z, z + (size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) z[(size_t)i * n + i] = 1.0;
  }
  ql_implicit(n, d, e.data(), z, n);
  sort_eigen(n, d, z, n);
}

// Householder reduction of a symmetric matrix to tridiagonal form (lower triangle used),
// accumulating the orthogonal transformation in a when vectors is set.
void sym_eigen(int n, double* a, double* w, bool vectors) {
  if (n <= 0) return;
  // work on row-major view A(i,j) = a[i + j*n] (symmetric, so layout is immaterial for input)
  auto A = [&](int i, int j) -> double& { return a[(size_t)j * n + i]; };
  std::vector<double> d(n), e(n, 0.0);
  for (int i = n - 1; i > 0; --i) {
    int l = i - 1;
    double h = 0.0, scale = 0.0;
    if (l > 0) {
      for (int k = 0; k <= l; ++k) scale += std::fabs(A(i, k));
      if (scale == 0.0) {
        e[i] = A(i, l);
      } else {
        for (int k = 0; k <= l; ++k) {
          A(i, k) /= scale;
          h += A(i, k) * A(i, k);
        }
        double f = A(i, l);
        double g = (f >= 0.0) ? -std::sqrt(h) : std::sqrt(h);
        e[i] = scale * g;
        h -= f * g;
        A(i, l) = f - g;
        f = 0.0;
        for (int j = 0; j <= l; ++j) {
          if (vectors) A(j, i) = A(i, j) / h;
          double gg = 0.0;
          for (int k = 0; k <= j; ++k) gg += A(j, k) * A(i, k);
          for (int k = j + 1; k <= l; ++k) gg += A(k, j) * A(i, k);
          e[j] = gg / h;
          f += e[j] * A(i, j);
        }
        double hh = f / (h + h);
        for (int j = 0; j <= l; ++j) {
          double ff = A(i, j);
          double gg = e[j] - hh * ff;
          e[j] = gg;
          for (int k = 0; k <= j; ++k) A(j, k) -= (ff * e[k] + gg * A(i, k));
        }
      }
    } else {
      e[i] = A(i, l);
    }
    d[i] = h;
  }
  d[0] = 0.0;
  e[0] = 0.0;
  for (int i = 0; i < n; ++i) {
    if (vectors) {
      int l = i - 1;
      if (d[i] != 0.0) {
        for (int j = 0; j <= l; ++j) {
          double g = 0.0;
          for (int k = 0; k <= l; ++k) g += A(i, k) * A(k, j);
          for (int k = 0; k <= l; ++k) A(k, j) -= g * A(k, i);
        }
      }
      d[i] = A(i, i);
      A(i, i) = 1.0;
      for (int j = 0; j <= l; ++j) A(j, i) = A(i, j) = 0.0;
    } else {
      d[i] = A(i, i);
    }
  }
  // subdiagonal: e[i] = T(i-1, i) -> shift to e[i-1]
  std::vector<double> sub(n, 0.0);
  for (int i = 1; i < n; ++i) sub[i - 1] = e[i];
  if (vectors) {
    // eigenvectors of T rotated into A: A currently holds Q (rows = components).
    // The QL rotations act on columns of Z = Q^T stored such that Z(k, i) = A(i, k)... use
    // explicit transpose for clarity.
    std::vector<double> z((size_t)n * n);
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < n; ++k) z[(size_t)i * n + k] = A(k, i);
    ql_implicit(n, d.data(), sub.data(), z.data(), n);
    sort_eigen(n, d.data(), z.data(), n);
    std::copy(z.begin(), z.end(), a);
  } else {
    ql_implicit(n, d.data(), sub.data(), nullptr, n);
    std::sort(d.begin(), d.end());
  }
  std::copy(d.begin(), d.end(), w);
}

bool cinverse(int n, std::complex<double>* a) {
  using C = std::complex<double>;
  std::vector<int> piv(n);
  auto A = [&](int i, int j) -> C& { return a[(size_t)j * n + i]; };
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::abs(A(k, k));
    for (int i = k + 1; i < n; ++i)
      if (std::abs(A(i, k)) > best) {
        best = std::abs(A(i, k));
        p = i;
      }
    if (best == 0.0) return false;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(A(k, j), A(p, j));
    C inv = 1.0 / A(k, k);
    A(k, k) = 1.0;
    for (int j = 0; j < n; ++j) A(k, j) *= inv;
    for (int i = 0; i < n; ++i) {
      if (i == k) continue;
      C f = A(i, k);
      if (f == C(0.0)) continue;
      A(i, k) = 0.0;
      for (int j = 0; j < n; ++j) A(i, j) -= f * A(k, j);
    }
  }
  for (int k = n - 1; k >= 0; --k)
    if (piv[k] != k)
      for (int i = 0; i < n; ++i) std::swap(A(i, k), A(i, piv[k]));
  return true;
}

double clogabsdet(int n, std::complex<double>* a) {
  auto A = [&](int i, int j) -> std::complex<double>& { return a[(size_t)j * n + i]; };
  double s = 0.0;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::abs(A(k, k));
    for (int i = k + 1; i < n; ++i)
      if (std::abs(A(i, k)) > best) {
        best = std::abs(A(i, k));
        p = i;
      }
    if (best == 0.0) return -INFINITY;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(A(k, j), A(p, j));
    s += std::log(std::abs(A(k, k)));
    for (int i = k + 1; i < n; ++i) {
      std::complex<double> f = A(i, k) / A(k, k);
      for (int j = k + 1; j < n; ++j) A(i, j) -= f * A(k, j);
    }
  }
  return s;
}

}  // namespace qed
