// Small dense FP64 linear algebra on the host for the k x k problems of the
// subspace iterations (k <= 256): Cholesky, triangular inverse and the cyclic
// Jacobi eigensolver (Golub & Van Loan, Matrix Computations, 4th ed., 8.5).
#pragma once

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace rfx {

// Lower Cholesky factor of a symmetric positive definite k x k (row-major).
// Returns false if a pivot is not positive.
inline bool cholesky_lower(std::vector<double>& a, int k) {
  for (int j = 0; j < k; ++j) {
    double d = a[j * k + j];
    for (int q = 0; q < j; ++q) d -= a[j * k + q] * a[j * k + q];
    if (!(d > 0.0)) return false;
    d = std::sqrt(d);
    a[j * k + j] = d;
    for (int i = j + 1; i < k; ++i) {
      double s = a[i * k + j];
      for (int q = 0; q < j; ++q) s -= a[i * k + q] * a[j * k + q];
      a[i * k + j] = s / d;
    }
    for (int q = j + 1; q < k; ++q) a[j * k + q] = 0.0;
  }
  return true;
}

// Inverse of the transpose of a lower-triangular L: returns R^{-1} with R = L^T
// (upper), so that V R^{-1} has orthonormal columns when V^T V = R^T R.
inline std::vector<double> inv_upper_from_lower(const std::vector<double>& L, int k) {
  // R = L^T upper; solve R X = I column by column (back substitution).
  std::vector<double> X(static_cast<size_t>(k) * k, 0.0);
  for (int c = 0; c < k; ++c) {
    for (int i = k - 1; i >= 0; --i) {
      double s = (i == c) ? 1.0 : 0.0;
      for (int q = i + 1; q < k; ++q) s -= L[q * k + i] * X[q * k + c];
      X[i * k + c] = s / L[i * k + i];
    }
  }
  return X;
}

// Eigen-decomposition of a symmetric k x k (row-major) by cyclic Jacobi.
// Returns eigenvalues in descending order and eigenvectors as columns of Z.
inline void jacobi_eigh(std::vector<double> a, int k, std::vector<double>& w, std::vector<double>& Z) {
  Z.assign(static_cast<size_t>(k) * k, 0.0);
  for (int i = 0; i < k; ++i) Z[i * k + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) {
        const double v = a[i * k + j] * a[i * k + j];
        tot += v;
        if (i != j) off += v;
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < k - 1; ++p)
      for (int q = p + 1; q < k; ++q) {
        const double apq = a[p * k + q];
        if (apq == 0.0) continue;
        const double app = a[p * k + p], aqq = a[q * k + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int r = 0; r < k; ++r) {
          const double arp = a[r * k + p], arq = a[r * k + q];
          a[r * k + p] = c * arp - s * arq;
          a[r * k + q] = s * arp + c * arq;
        }
        for (int r = 0; r < k; ++r) {
          const double apr = a[p * k + r], aqr = a[q * k + r];
          a[p * k + r] = c * apr - s * aqr;
          a[q * k + r] = s * apr + c * aqr;
        }
        for (int r = 0; r < k; ++r) {
          const double zrp = Z[r * k + p], zrq = Z[r * k + q];
          Z[r * k + p] = c * zrp - s * zrq;
          Z[r * k + q] = s * zrp + c * zrq;
        }
      }
  }
  std::vector<int> ord(k);
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int x, int y) { return a[x * k + x] > a[y * k + y]; });
  w.resize(k);
  std::vector<double> Zs(static_cast<size_t>(k) * k);
  for (int c = 0; c < k; ++c) {
    w[c] = a[ord[c] * k + ord[c]];
    for (int r = 0; r < k; ++r) Zs[r * k + c] = Z[r * k + ord[c]];
  }
  Z.swap(Zs);
}

}  // namespace rfx
