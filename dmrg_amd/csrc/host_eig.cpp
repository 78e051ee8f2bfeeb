// Small dense symmetric / Hermitian eigensolvers on the host: the projected
// Lanczos matrices (at most ncv x ncv) and the tiny exact-diagonalisation
// branch.  Replaces LAPACK ?syev / ?heev of the reference
// (src/tensor.cpp:803-840) for these sizes.
#include "host_eig.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace dmrgx {

namespace {

// Householder reduction of the symmetric z (n x n, row-major, lower triangle
// used) to tridiagonal form: on return d holds the diagonal, e[1..n) the
// subdiagonal (e[0] = 0) and z the orthogonal Q with A = Q T Q^T.
void tridiagonalise(int n, double* z, double* d, double* e) {
  auto Z = [&](int i, int j) -> double& { return z[(size_t)i * n + j]; };
  for (int i = n - 1; i > 0; --i) {
    const int l = i - 1;
    double h = 0;
    if (l > 0) {
      double scale = 0;
      for (int k = 0; k <= l; ++k) scale += std::fabs(Z(i, k));
      if (scale == 0) {
        e[i] = Z(i, l);
      } else {
        // reflector v = row i scaled, with v_l shifted so that H x = -sign(x_l)|x| e_l
        for (int k = 0; k <= l; ++k) {
          Z(i, k) /= scale;
          h += Z(i, k) * Z(i, k);
        }
        double f = Z(i, l);
        const double g = f >= 0 ? -std::sqrt(h) : std::sqrt(h);
        e[i] = scale * g;
        h -= f * g;
        Z(i, l) = f - g;
        // p = A v / h (stored in e[0..l]), K = v^T p / 2h, q = p - K v
        f = 0;
        for (int j = 0; j <= l; ++j) {
          Z(j, i) = Z(i, j) / h;
          double s = 0;
          for (int k = 0; k <= j; ++k) s += Z(j, k) * Z(i, k);
          for (int k = j + 1; k <= l; ++k) s += Z(k, j) * Z(i, k);
          e[j] = s / h;
          f += e[j] * Z(i, j);
        }
        const double hh = f / (h + h);
        // A <- A - v q^T - q v^T on the lower triangle
        for (int j = 0; j <= l; ++j) {
          const double fj = Z(i, j);
          const double gj = e[j] - hh * fj;
          e[j] = gj;
          for (int k = 0; k <= j; ++k) Z(j, k) -= fj * e[k] + gj * Z(i, k);
        }
      }
    } else {
      e[i] = Z(i, l);
    }
    d[i] = h;
  }
  d[0] = 0;
  e[0] = 0;
  // accumulate Q from the stored reflectors
  for (int i = 0; i < n; ++i) {
    const int l = i - 1;
    if (d[i] != 0) {
      for (int j = 0; j <= l; ++j) {
        double g = 0;
        for (int k = 0; k <= l; ++k) g += Z(i, k) * Z(k, j);
        for (int k = 0; k <= l; ++k) Z(k, j) -= g * Z(k, i);
      }
    }
    d[i] = Z(i, i);
    Z(i, i) = 1;
    for (int j = 0; j <= l; ++j) Z(j, i) = Z(i, j) = 0;
  }
}

// Implicit QL with Wilkinson-type shifts on the symmetric tridiagonal (d, e),
// e[1..n) the subdiagonal; rotations accumulated into the columns of z.
void tridiagonal_ql(int n, double* d, double* e, double* z) {
  auto Z = [&](int i, int j) -> double& { return z[(size_t)i * n + j]; };
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  if (n > 0) e[n - 1] = 0;
  const double eps = 2.220446049250313e-16;
  for (int l = 0; l < n; ++l) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m != l) {
        if (iter++ >= 100) throw std::runtime_error("dmrgx: tridiagonal QL did not converge");
        double g = (d[l + 1] - d[l]) / (2 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
        double s = 1, c = 1, p = 0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i];
          const double b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0) {
            d[i + 1] -= p;
            e[m] = 0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          for (int k = 0; k < n; ++k) {
            f = Z(k, i + 1);
            Z(k, i + 1) = s * Z(k, i) + c * f;
            Z(k, i) = c * Z(k, i) - s * f;
          }
        }
        if (r == 0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0;
      }
    } while (m != l);
  }
}

}  // namespace

void host_symeig(int n, std::vector<double>& a, std::vector<double>& w, std::vector<double>& v) {
  if (n < 0 || (long long)a.size() < (long long)n * n) throw std::runtime_error("dmrgx: host_symeig: matrix too small");
  // symmetrise (the lower triangle drives the reduction)
  std::vector<double> z((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) z[(size_t)i * n + j] = z[(size_t)j * n + i] =
        0.5 * (a[(size_t)i * n + j] + a[(size_t)j * n + i]);
  std::vector<double> d(n), e(n);
  if (n > 0) {
    tridiagonalise(n, z.data(), d.data(), e.data());
    tridiagonal_ql(n, d.data(), e.data(), z.data());
  }
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
  w.resize(n);
  v.assign((size_t)n * n, 0.0);
  for (int c = 0; c < n; ++c) {
    w[c] = d[idx[c]];
    for (int k = 0; k < n; ++k) v[(size_t)k * n + c] = z[(size_t)k * n + idx[c]];
  }
}

void host_heeig(int n, std::vector<cplx>& a, std::vector<double>& w, std::vector<cplx>& v) {
  // embed the Hermitian matrix H = X + iY as the real symmetric [[X, -Y], [Y, X]]
  int m = 2 * n;
  std::vector<double> r((size_t)m * m);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      cplx h = 0.5 * (a[(size_t)i * n + j] + std::conj(a[(size_t)j * n + i]));
      r[(size_t)i * m + j] = h.real();
      r[(size_t)(i + n) * m + j + n] = h.real();
      r[(size_t)i * m + j + n] = -h.imag();
      r[(size_t)(i + n) * m + j] = h.imag();
    }
  std::vector<double> wr, vr;
  host_symeig(m, r, wr, vr);
  // eigenvalues come in pairs; take every other one
  w.resize(n);
  v.assign((size_t)n * n, 0.0);
  for (int c = 0; c < n; ++c) {
    w[c] = wr[2 * c];
    double nrm = 0;
    for (int k = 0; k < n; ++k) {
      cplx z(vr[(size_t)k * m + 2 * c], vr[(size_t)(k + n) * m + 2 * c]);
      v[(size_t)k * n + c] = z;
      nrm += std::norm(z);
    }
    nrm = std::sqrt(nrm);
    for (int k = 0; k < n; ++k) v[(size_t)k * n + c] /= nrm;
  }
}

void host_hermitian_eig(int n, std::vector<cplx>& a, bool real, std::vector<double>& w, std::vector<cplx>& v) {
  if (!real) {
    host_heeig(n, a, w, v);
    return;
  }
  std::vector<double> ar((size_t)n * n), vr;
  for (size_t i = 0; i < ar.size(); ++i) ar[i] = a[i].real();
  host_symeig(n, ar, w, vr);
  v.assign(vr.begin(), vr.end());
}

}  // namespace dmrgx
