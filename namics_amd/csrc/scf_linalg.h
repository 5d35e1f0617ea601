// Small dense solve for the DIIS step, usable on host and device.
//
// DIIS (reference SFNewton::DIIS / Ax, src/sfnewton.cpp:793-880) solves
// A c = (1,...,1) for the Gram matrix A of the last k error vectors through a
// singular value decomposition.  The SVD here is the Golub-Reinsch algorithm
// (Householder bidiagonalisation followed by implicitly shifted QR), in the
// formulation published in Numerical Recipes in C (2nd ed., section 2.6),
// with the matrix held transposed (a[col][row]) as the reference stores it
// (src/tools.cpp:1076-1349), so the solve returns the same coefficients.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace nscf {

constexpr int kMaxDIIS = 16;

__host__ __device__ inline double la_pythag(double a, double b) {
  double at = fabs(a), bt = fabs(b), ct;
  if (at > bt) {
    ct = bt / at;
    return at * sqrt(1.0 + ct * ct);
  }
  if (bt > 0.0) {
    ct = at / bt;
    return bt * sqrt(1.0 + ct * ct);
  }
  return 0.0;
}

__host__ __device__ inline double la_sign(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

// SVD of the n x n matrix a (transposed storage).  Returns false on NaN.
__host__ __device__ inline bool la_svd(double (*a)[kMaxDIIS], int n, double *w, double (*v)[kMaxDIIS]) {
  const int m = n;
  int flag, i, its, j, jj, k, l = 0, nm = 0;
  double c, f, h, s, x, y, z, anorm = 0.0, g = 0.0, scale = 0.0;
  double rv1[kMaxDIIS];
  for (i = 0; i < n; i++) rv1[i] = 0.0;
  for (i = 0; i < n; i++) {
    l = i + 1;
    rv1[i] = scale * g;
    g = s = scale = 0.0;
    for (k = i; k < m; k++) scale += fabs(a[i][k]);
    if (scale != 0.0) {
      for (k = i; k < m; k++) {
        a[i][k] = a[i][k] / scale;
        s += a[i][k] * a[i][k];
      }
      f = a[i][i];
      g = -la_sign(sqrt(s), f);
      h = f * g - s;
      a[i][i] = f - g;
      if (i != n - 1) {
        for (j = l; j < n; j++) {
          for (s = 0.0, k = i; k < m; k++) s += a[i][k] * a[j][k];
          f = s / h;
          for (k = i; k < m; k++) a[j][k] += f * a[i][k];
        }
      }
      for (k = i; k < m; k++) a[i][k] = a[i][k] * scale;
    }
    w[i] = scale * g;
    g = s = scale = 0.0;
    if (i != n - 1) {
      for (k = l; k < n; k++) scale += fabs(a[k][i]);
      if (scale != 0.0) {
        for (k = l; k < n; k++) {
          a[k][i] = a[k][i] / scale;
          s += a[k][i] * a[k][i];
        }
        f = a[l][i];
        g = -la_sign(sqrt(s), f);
        h = f * g - s;
        a[l][i] = f - g;
        for (k = l; k < n; k++) rv1[k] = a[k][i] / h;
        if (i != m - 1) {
          for (j = l; j < m; j++) {
            for (s = 0.0, k = l; k < n; k++) s += a[k][j] * a[k][i];
            for (k = l; k < n; k++) a[k][j] += s * rv1[k];
          }
        }
        for (k = l; k < n; k++) a[k][i] = a[k][i] * scale;
      }
    }
    double t = fabs(w[i]) + fabs(rv1[i]);
    anorm = anorm > t ? anorm : t;
  }
  for (i = n - 1; i >= 0; i--) {
    if (i < n - 1) {
      if (g != 0.0) {
        for (j = l; j < n; j++) v[j][i] = (a[j][i] / a[l][i]) / g;
        for (j = l; j < n; j++) {
          for (s = 0.0, k = l; k < n; k++) s += a[k][i] * v[k][j];
          for (k = l; k < n; k++) v[k][j] += s * v[k][i];
        }
      }
      for (j = l; j < n; j++) v[i][j] = v[j][i] = 0.0;
    }
    v[i][i] = 1.0;
    g = rv1[i];
    l = i;
  }
  for (i = n - 1; i >= 0; i--) {
    l = i + 1;
    g = w[i];
    if (i < n - 1)
      for (j = l; j < n; j++) a[j][i] = 0.0;
    if (g != 0.0) {
      g = 1.0 / g;
      if (i != n - 1) {
        for (j = l; j < n; j++) {
          for (s = 0.0, k = l; k < m; k++) s += a[i][k] * a[j][k];
          f = (s / a[i][i]) * g;
          for (k = i; k < m; k++) a[j][k] += f * a[i][k];
        }
      }
      for (j = i; j < m; j++) a[i][j] = a[i][j] * g;
    } else {
      for (j = i; j < m; j++) a[i][j] = 0.0;
    }
    a[i][i] += 1.0;
  }
  for (i = 0; i < n; i++)
    for (j = 0; j < n; j++)
      if (a[i][j] != a[i][j]) return false;
  for (k = n - 1; k >= 0; k--) {
    for (its = 0; its < 30; its++) {
      flag = 1;
      for (l = k; l >= 0; l--) {
        nm = l - 1;
        if (fabs(rv1[l]) + anorm == anorm) {
          flag = 0;
          break;
        }
        if (fabs(w[nm]) + anorm == anorm) break;
      }
      if (flag) {
        c = 0.0;
        s = 1.0;
        for (i = l; i <= k; i++) {
          f = s * rv1[i];
          if (fabs(f) + anorm != anorm) {
            g = w[i];
            h = la_pythag(f, g);
            w[i] = h;
            h = 1.0 / h;
            c = g * h;
            s = -f * h;
            for (j = 0; j < m; j++) {
              y = a[nm][j];
              z = a[i][j];
              a[nm][j] = y * c + z * s;
              a[i][j] = z * c - y * s;
            }
          }
        }
      }
      z = w[k];
      if (l == k) {
        if (z < 0.0) {
          w[k] = -z;
          for (j = 0; j < n; j++) v[j][k] = -v[j][k];
        }
        break;
      }
      if (its >= 30) return true;
      x = w[l];
      nm = k - 1;
      y = w[nm];
      g = rv1[nm];
      h = rv1[k];
      f = ((y - z) * (y + z) + (g - h) * (g + h)) / (2.0 * h * y);
      g = la_pythag(f, 1.0);
      f = ((x - z) * (x + z) + h * ((y / (f + la_sign(g, f))) - h)) / x;
      c = s = 1.0;
      for (j = l; j <= nm; j++) {
        i = j + 1;
        g = rv1[i];
        y = w[i];
        h = s * g;
        g = c * g;
        z = la_pythag(f, h);
        rv1[j] = z;
        c = f / z;
        s = h / z;
        f = x * c + g * s;
        g = g * c - x * s;
        h = y * s;
        y = y * c;
        for (jj = 0; jj < n; jj++) {
          x = v[jj][j];
          z = v[jj][i];
          v[jj][j] = x * c + z * s;
          v[jj][i] = z * c - x * s;
        }
        z = la_pythag(f, h);
        w[j] = z;
        if (z != 0.0) {
          z = 1.0 / z;
          c = f * z;
          s = h * z;
        }
        f = (c * g) + (s * y);
        x = (c * y) - (s * g);
        for (jj = 0; jj < m; jj++) {
          y = a[j][jj];
          z = a[i][jj];
          a[j][jj] = y * c + z * s;
          a[i][jj] = z * c - y * s;
        }
      }
      rv1[l] = 0.0;
      rv1[k] = f;
      w[k] = x;
    }
  }
  return true;
}

// c = A^-1 (1..1) through the SVD, A given row-major with leading dimension n;
// then normalised to sum 1 (the DIIS coefficients).  Returns false on NaN.
__host__ __device__ inline bool la_diis_coefficients(const double *A, int n, double *c) {
  for (int i = 0; i < n * n; i++)
    if (A[i] != A[i]) return false;
  if (n == 1) {
    c[0] = 1.0;
    return true;
  }
  double U[kMaxDIIS][kMaxDIIS], V[kMaxDIIS][kMaxDIIS], S[kMaxDIIS];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      U[i][j] = A[i * n + j];
      V[i][j] = 0.0;
    }
  if (!la_svd(U, n, S, V)) return false;
  double X[kMaxDIIS];
  for (int i = 0; i < n; i++) {
    X[i] = 0;
    for (int j = 0; j < n; j++) X[i] += U[i][j];
    S[i] = X[i] / S[i];
  }
  for (int i = 0; i < n; i++) {
    X[i] = 0;
    for (int j = 0; j < n; j++) X[i] += V[i][j] * S[j];
  }
  double normC = 0;
  for (int i = 0; i < n; i++) normC += X[i];
  for (int i = 0; i < n; i++) c[i] = X[i] / normC;
  return true;
}

}  // namespace nscf
