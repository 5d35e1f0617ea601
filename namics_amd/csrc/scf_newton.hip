// Newton solver for `newton : method : pseudohessian | hessian` inputs.
//
// The reference solves these inputs with a dense quasi-Newton method
// (SFNewton::iterate, src/sfnewton.cpp:636-747, selected in
// src/solve_scf.cpp:191-231).  This is a different, self-contained Newton
// method with the same fixed point: a Jacobian-free inexact Newton iteration
// (Knoll & Keyes, J. Comput. Phys. 193 (2004) 357) --
//
//   * the Newton system J d = -F(x) is solved by GMRES (Saad & Schultz 1986)
//     with modified Gram-Schmidt applied twice and Givens rotations, up to
//     kRestarts cycles of kKrylov vectors;
//   * J v is the forward difference (F(x + h v) - F(x)) / h with
//     h = sqrt(eps (1 + ||x||)) / ||v||;
//   * the GMRES tolerance follows Eisenstat & Walker's choice 2
//     (SIAM J. Sci. Comput. 17 (1996) 16) with their safeguards, kept in
//     [kEtaMin, kEtaMax] and no tighter than the final tolerance asks;
//   * the step is capped at kMaxStep in max-norm and shortened by halving
//     until ||F|| drops (Armijo condition, Dennis & Schnabel 6.3).
//
// F is the residual of the SCF equations (the hot path: every evaluation
// runs on the device, fused or grid path); the Krylov vectors and the small
// least-squares problem live on the host, as the reference keeps its Newton
// on the host.  Masked variables (frozen sites, ghosts) have F = 0 for every
// x, so J v vanishes there and the iteration never moves them.  The stop
// test is the system's own: ||g||_2 (or max|g|) <= tolerance.  `iterations`
// counts Newton steps, bounded by the input's iterationlimit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <vector>

#include "scf_internal.h"

namespace nscf {
namespace {

constexpr int kKrylov = 40;        // Krylov vectors per GMRES cycle
constexpr int kRestarts = 3;       // GMRES cycles per Newton step
constexpr int kHalvings = 30;      // step halvings before the step is given up
constexpr double kMaxStep = 2.0;   // max |dx_i| of one Newton step
constexpr double kEtaMax = 0.1;
constexpr double kEtaMin = 1e-8;
constexpr double kArmijo = 1e-4;

double norm2(const std::vector<double> &v) {
  double s = 0;
  for (double a : v) s += a * a;
  return std::sqrt(s);
}

double dot(const std::vector<double> &a, const std::vector<double> &b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
  return s;
}

// Residual of the SCF equations at host fields x on the device.  Returns
// SCF_OK or a HIP error; *status is the evaluation's own status (DS_OK or a
// failure such as a vanishing partition function).  With `thermo`, the grid
// path also computes F, Omega and mu (the fused path always does).
int evaluate(scf_system *S, const double *x, double *g, double *res, int *status, bool thermo) {
  const DevSys &h = S->h;
  const size_t nvar = (size_t)h.nit * h.M;
  hipStream_t st = S->stream;
  HIP_TRY(hipMemcpyAsync(h.x, x, nvar * sizeof(double), hipMemcpyHostToDevice, st));
  if (S->effective_mode() == SCF_MODE_FUSED) {
    HIP_TRY(launch_fused(S->d_list, 1, FUSED_RESIDUAL, h.threads, S->lds_bytes, st));
  } else {
    HIP_TRY(grid_residual(h, S->d, h.partial, S->nb, st));
    if (thermo) HIP_TRY(grid_thermo(h, S->d, h.partial, S->nb, st));
  }
  DevScalars sc;
  HIP_TRY(hipMemcpyAsync(&sc, h.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, st));
  if (g) HIP_TRY(hipMemcpyAsync(g, h.g, nvar * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  *res = sc.residual;
  *status = sc.status;
  return SCF_OK;
}

struct Newton {
  scf_system *S;
  size_t n;
  std::vector<double> x, F, xt, Ft, w;
  double fnorm = 0;
  int hip_rc = SCF_OK;

  // J v by a forward difference; false when the perturbed evaluation fails
  bool jv(const std::vector<double> &v, std::vector<double> &out) {
    const double vn = norm2(v);
    if (!(vn > 0)) {
      std::fill(out.begin(), out.end(), 0.0);
      return true;
    }
    const double h = std::sqrt(2.220446049250313e-16 * (1.0 + norm2(x))) / vn;
    for (size_t i = 0; i < n; i++) xt[i] = x[i] + h * v[i];
    double r;
    int status;
    hip_rc = evaluate(S, xt.data(), Ft.data(), &r, &status, false);
    if (hip_rc != SCF_OK || status != DS_OK) return false;
    for (size_t i = 0; i < n; i++) out[i] = (Ft[i] - F[i]) / h;
    return true;
  }

  // GMRES for J d = -F to relative residual eta; d holds the best iterate
  bool gmres(double eta, std::vector<double> &d) {
    std::fill(d.begin(), d.end(), 0.0);
    std::vector<std::vector<double>> V(kKrylov + 1, std::vector<double>(n));
    std::vector<double> H((kKrylov + 1) * kKrylov), cs(kKrylov), sn(kKrylov), gv(kKrylov + 1), y(kKrylov);
    std::vector<double> r(n);
    const double target = eta * fnorm;
    for (int cycle = 0; cycle < kRestarts; cycle++) {
      if (cycle == 0) {
        for (size_t i = 0; i < n; i++) r[i] = -F[i];
      } else {
        if (!jv(d, w)) return false;
        for (size_t i = 0; i < n; i++) r[i] = -F[i] - w[i];
      }
      const double beta = norm2(r);
      if (beta <= target || !(beta > 0)) return true;
      for (size_t i = 0; i < n; i++) V[0][i] = r[i] / beta;
      std::fill(gv.begin(), gv.end(), 0.0);
      gv[0] = beta;
      int k = 0;
      for (; k < kKrylov; k++) {
        if (!jv(V[k], w)) return false;
        for (int i = 0; i <= k; i++) H[i * kKrylov + k] = 0.0;
        for (int pass = 0; pass < 2; pass++)
          for (int i = 0; i <= k; i++) {
            const double c = dot(w, V[i]);
            H[i * kKrylov + k] += c;
            for (size_t j = 0; j < n; j++) w[j] -= c * V[i][j];
          }
        const double hn = norm2(w);
        H[(k + 1) * kKrylov + k] = hn;
        if (hn > 0)
          for (size_t j = 0; j < n; j++) V[k + 1][j] = w[j] / hn;
        for (int i = 0; i < k; i++) {
          const double a = H[i * kKrylov + k], b = H[(i + 1) * kKrylov + k];
          H[i * kKrylov + k] = cs[i] * a + sn[i] * b;
          H[(i + 1) * kKrylov + k] = -sn[i] * a + cs[i] * b;
        }
        const double a = H[k * kKrylov + k], b = H[(k + 1) * kKrylov + k];
        const double rr = std::hypot(a, b);
        cs[k] = rr > 0 ? a / rr : 1.0;
        sn[k] = rr > 0 ? b / rr : 0.0;
        H[k * kKrylov + k] = rr;
        H[(k + 1) * kKrylov + k] = 0.0;
        gv[k + 1] = -sn[k] * gv[k];
        gv[k] = cs[k] * gv[k];
        if (std::fabs(gv[k + 1]) <= target || !(hn > 0)) {
          k++;
          break;
        }
      }
      // back substitution and update of d
      for (int i = k - 1; i >= 0; i--) {
        double s = gv[i];
        for (int j = i + 1; j < k; j++) s -= H[i * kKrylov + j] * y[j];
        y[i] = H[i * kKrylov + i] != 0.0 ? s / H[i * kKrylov + i] : 0.0;
      }
      for (int i = 0; i < k; i++)
        for (size_t j = 0; j < n; j++) d[j] += y[i] * V[i][j];
      if (std::fabs(gv[k]) <= target) return true;
    }
    return true;
  }
};

}  // namespace

int newton_solve(scf_system *S, scf_solve_info *info) {
  const auto t0 = std::chrono::steady_clock::now();
  const DevSys &h = S->h;
  const size_t n = (size_t)h.nit * h.M;
  Newton N;
  N.S = S;
  N.n = n;
  N.x.resize(n);
  N.F.resize(n);
  N.xt.resize(n);
  N.Ft.resize(n);
  N.w.resize(n);
  HIP_TRY(hipMemcpyAsync(N.x.data(), h.x, n * sizeof(double), hipMemcpyDeviceToHost, S->stream));
  HIP_TRY(hipStreamSynchronize(S->stream));
  const double tol = h.tol;
  const int itlim = h.itlim;
  double res;
  int status;
  int rc = evaluate(S, N.x.data(), N.F.data(), &res, &status, false);
  if (rc != SCF_OK) return rc;
  N.fnorm = norm2(N.F);
  int it = 0;
  double eta = kEtaMax, fprev = N.fnorm;
  std::vector<double> d(n), xbest(n);
  while (status == DS_OK && res > tol && it < itlim) {
    // forcing term (Eisenstat-Walker choice 2, safeguarded)
    if (it > 0) {
      const double ratio = N.fnorm / fprev;
      double e = 0.9 * ratio * ratio;
      const double safe = 0.9 * eta * eta;
      if (safe > 0.1) e = std::max(e, safe);
      eta = std::min(kEtaMax, std::max(kEtaMin, e));
      // no tighter than what the remaining distance to the tolerance needs
      eta = std::max(eta, std::min(kEtaMax, 0.5 * tol / N.fnorm));
    }
    if (!N.gmres(eta, d)) {
      if (N.hip_rc != SCF_OK) return N.hip_rc;
      break;  // a perturbed evaluation failed: stop where the iterate is
    }
    double dmax = 0;
    for (double v : d) dmax = std::max(dmax, std::fabs(v));
    double lambda = dmax > kMaxStep ? kMaxStep / dmax : 1.0;
    bool accepted = false;
    double rt = 0;
    int st = DS_OK;
    for (int k = 0; k < kHalvings; k++, lambda *= 0.5) {
      for (size_t i = 0; i < n; i++) N.xt[i] = N.x[i] + lambda * d[i];
      rc = evaluate(S, N.xt.data(), N.Ft.data(), &rt, &st, false);
      if (rc != SCF_OK) return rc;
      if (st != DS_OK) continue;
      const double ft = norm2(N.Ft);
      if (ft <= (1.0 - kArmijo * lambda) * N.fnorm) {
        accepted = true;
        break;
      }
    }
    if (!accepted) break;  // no decrease along d: stop where the iterate is
    fprev = N.fnorm;
    N.x.swap(N.xt);
    N.F.swap(N.Ft);
    N.fnorm = norm2(N.F);
    res = rt;
    status = st;
    it++;
  }
  // leave the device state (fields, densities, F, Omega, mu) at the final x
  rc = evaluate(S, N.x.data(), nullptr, &res, &status, true);
  if (rc != SCF_OK) return rc;
  info->iterations = it;
  info->residual = res;
  info->status = status != DS_OK ? status : (res <= tol ? SCF_OK : SCF_NOT_CONVERGED);
  info->fp32_iterations = 0;
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  info->kernel_ms = ms;
  info->solve_ms = ms;
  S->have_results = status == DS_OK;
  return SCF_OK;
}


// ---------------------------------------------------------------------------
// LBFGS solver for `newton : method : LBFGS` inputs.
//
// The reference runs an adaptation of the LBFGSpp solver (Y. Qiu, MIT
// licence; src/LBFGS.h, src/LBFGSpp/BFGSMat.h, driven from
// src/solve_scf.cpp:526-556,630-644) with the SCF residual g(x) in the role of
// the gradient and ||g||_2 as the value; this restates its steps:
//
//   * the direction is -H g, H the limited-memory inverse-Hessian estimate of
//     the last m corrections (s = x_k - x_{k-1}, y = g_k - g_{k-1}), applied
//     by the two-loop recursion (Nocedal & Wright, Numerical Optimization,
//     2nd ed., Algorithm 7.4) from H0 = (s.y / y.y) I of the newest pair;
//   * the step x_new = x + alpha d has alpha = min(1, r / ||d||) for a trust
//     radius r: before each step r shrinks to 2 alpha_prev ||d|| when it
//     exceeds that, is multiplied by the factor the previous step earned,
//     and is kept in [deltamin, deltamax]; a step's factor grows when the
//     new residual is small or turned away from the old one and shrinks when
//     it grew (`step_factor`), times alpha / (r / ||d||) of the step taken;
//   * no line search: every step is taken; the run stops at ||g||_2 <=
//     tolerance or after iterationlimit steps.
//
// Vectors live on the host (m correction pairs, 2 m + 4 field-sized
// vectors), as for the Newton solver; every residual is the device's.
namespace {

// factor of the trust radius earned by a step from residual g0 to g
double step_factor(const std::vector<double> &g, const std::vector<double> &g0) {
  const double n0 = norm2(g0);
  double ratio = dot(g, g0) / n0 / n0;
  double q = norm2(g) / n0;
  q = q * q;
  if ((ratio > 1 || q > 1) && q - ratio * std::fabs(ratio) < 0.2) q = 1.5 * q;
  if (ratio < 0 && q < 2) return 1;
  if (q > 10) return 0.01;
  return 0.4 * (1 + 0.75 * q) / (q - ratio * std::fabs(ratio) + 0.1);
}

}  // namespace

int lbfgs_solve(scf_system *S, scf_solve_info *info) {
  const auto t0 = std::chrono::steady_clock::now();
  const DevSys &h = S->h;
  const size_t n = (size_t)h.nit * h.M;
  const int m = S->spec.solver.m;
  const double tol = h.tol, dmax = h.deltamax, dmin = S->spec.solver.deltamin;
  const int itlim = h.itlim;
  std::vector<double> x(n), g(n), xp(n), gp(n), d(n);
  std::vector<std::vector<double>> sv(m, std::vector<double>(n)), yv(m, std::vector<double>(n));
  std::vector<double> ys(m), al(m);
  HIP_TRY(hipMemcpyAsync(x.data(), h.x, n * sizeof(double), hipMemcpyDeviceToHost, S->stream));
  HIP_TRY(hipStreamSynchronize(S->stream));
  double res;
  int status;
  int rc = evaluate(S, x.data(), g.data(), &res, &status, false);
  if (rc != SCF_OK) return rc;
  double gnorm = norm2(g);
  int it = 1;
  if (status == DS_OK && gnorm > tol) {
    double radius = dmax, factor = 1, alpha = 1, theta = 1;
    int npairs = 0, next = 0;
    for (size_t i = 0; i < n; i++) d[i] = -g[i];
    for (;;) {
      xp = x;
      gp = g;
      const double dn = norm2(d);
      if (dn > 0 && radius > 2 * alpha * dn) radius = 2 * alpha * dn;
      radius *= factor;
      factor = 1;
      radius = std::min(dmax, std::max(dmin, radius));
      const double bound = radius / dn;
      alpha = bound > 1 ? 1 : bound;
      for (size_t i = 0; i < n; i++) x[i] = xp[i] + alpha * d[i];
      rc = evaluate(S, x.data(), g.data(), &res, &status, false);
      if (rc != SCF_OK) return rc;
      if (status != DS_OK) break;
      gnorm = norm2(g);
      factor *= step_factor(g, gp);
      factor *= alpha / bound;
      if (alpha < tol) alpha *= 2;
      if (gnorm <= tol || it >= itlim) break;
      // newest correction pair
      std::vector<double> &s = sv[next], &y = yv[next];
      for (size_t i = 0; i < n; i++) {
        s[i] = x[i] - xp[i];
        y[i] = g[i] - gp[i];
      }
      ys[next] = dot(s, y);
      theta = dot(y, y) / ys[next];
      next = (next + 1) % m;
      if (npairs < m) npairs++;
      // d = -H g by the two-loop recursion, newest pair first
      for (size_t i = 0; i < n; i++) d[i] = -g[i];
      int j = next;
      for (int q = 0; q < npairs; q++) {
        j = (j + m - 1) % m;
        al[j] = dot(sv[j], d) / ys[j];
        for (size_t i = 0; i < n; i++) d[i] -= al[j] * yv[j][i];
      }
      for (size_t i = 0; i < n; i++) d[i] /= theta;
      for (int q = 0; q < npairs; q++) {
        const double beta = dot(yv[j], d) / ys[j];
        for (size_t i = 0; i < n; i++) d[i] += (al[j] - beta) * sv[j][i];
        j = (j + 1) % m;
      }
      it++;
    }
  }
  // leave the device state (fields, densities, F, Omega, mu) at the final x
  rc = evaluate(S, x.data(), nullptr, &res, &status, true);
  if (rc != SCF_OK) return rc;
  info->iterations = it;
  info->residual = res;
  info->status = status != DS_OK ? status : (gnorm <= tol ? SCF_OK : SCF_NOT_CONVERGED);
  info->fp32_iterations = 0;
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  info->kernel_ms = ms;
  info->solve_ms = ms;
  S->have_results = status == DS_OK;
  return SCF_OK;
}

// ---------------------------------------------------------------------------
// Broyden rank-reduction solver for `newton : method : BRR`
// (SFNewton::iterate_BRR src/sfnewton.cpp:929-1035; Broyden's inverse update
// kept as a sum of at most m rank-one terms C_j D_j^T, reduced to its
// dominant terms when full).  Per iteration:
//
//   * s = -g + sum_{j=1..k} C_j (D_j . g), and the step x = x0 + alpha s with
//     alpha = deltamax while ||g|| > 1, else deltamax - it (1 - deltamax)
//     ln(||g||) / 1e4, at most 1 (a negative alpha becomes the tolerance);
//   * with y = g0 - g, a new term: C_k = (s - c) / (s . c) ||d||, D_k = d /
//     ||d||, where c = sum_{j<k} C_j (D_j . y) - y and d = sum_{j<k} D_j
//     (C_j . s) - s (the unscaled direction s, as the reference has it);
//   * when the m terms are in use, D = Q R (thin QR), C <- C R^T, D <- Q,
//     then C = U S V^T (thin SVD), C <- U S, D <- D V: the same operator,
//     its terms ordered by singular value; the last term is dropped, and one
//     more for every singular value (of the 2nd..(m-1)th) below tolerance
//     times the largest.  Term 0 takes part in the reduction but not in the
//     updates (the reference's blocks start at column 1).
//
// The reference takes the QR and SVD from Eigen (HouseholderQR, JacobiSVD);
// here a Householder QR that leaves a zero column unreflected (as Eigen's
// makeHouseholder does) and a one-sided Jacobi SVD.  Both give the same
// terms up to signs and rounding, and the operator C D^T does not depend on
// the signs.  Vectors live on the host (2 (m + 1) field-sized columns), as
// for the Newton and LBFGS solvers; every residual is the device's.
namespace {

using Cols = std::vector<std::vector<double>>;

// thin Householder QR of the n x p matrix A (columns): Q (n x p, columns)
// and R (p x p, R[i][j] for i <= j)
void thin_qr(const Cols &A, Cols &Q, std::vector<std::vector<double>> &R) {
  const int p = (int)A.size();
  const size_t n = A[0].size();
  Cols W = A;  // reflected in place
  std::vector<double> tau(p, 0.0);
  std::vector<std::vector<double>> ess(p);
  R.assign(p, std::vector<double>(p, 0.0));
  for (int j = 0; j < p; j++) {
    std::vector<double> &c = W[j];
    double tail = 0;
    for (size_t i = j + 1; i < n; i++) tail += c[i] * c[i];
    const double c0 = c[j];
    double beta;
    if (tail <= std::numeric_limits<double>::min()) {
      tau[j] = 0;
      beta = c0;
    } else {
      beta = std::sqrt(c0 * c0 + tail);
      if (c0 >= 0) beta = -beta;
      ess[j].assign(n, 0.0);
      for (size_t i = j + 1; i < n; i++) ess[j][i] = c[i] / (c0 - beta);
      tau[j] = (beta - c0) / beta;
    }
    R[j][j] = beta;
    // apply H_j = I - tau w w^T (w = e_j + essential) to the later columns
    if (tau[j] != 0)
      for (int l = j + 1; l < p; l++) {
        std::vector<double> &a = W[l];
        double t = a[j];
        for (size_t i = j + 1; i < n; i++) t += ess[j][i] * a[i];
        t *= tau[j];
        a[j] -= t;
        for (size_t i = j + 1; i < n; i++) a[i] -= t * ess[j][i];
      }
    for (int l = j + 1; l < p; l++) R[j][l] = W[l][j];
  }
  // Q = H_0 H_1 ... H_{p-1} applied to the first p unit vectors
  Q.assign(p, std::vector<double>(n, 0.0));
  for (int c = 0; c < p; c++) Q[c][c] = 1.0;
  for (int j = p - 1; j >= 0; j--) {
    if (tau[j] == 0) continue;
    for (int c = 0; c < p; c++) {
      std::vector<double> &q = Q[c];
      double t = q[j];
      for (size_t i = j + 1; i < n; i++) t += ess[j][i] * q[i];
      t *= tau[j];
      q[j] -= t;
      for (size_t i = j + 1; i < n; i++) q[i] -= t * ess[j][i];
    }
  }
}

// one-sided Jacobi SVD of the n x p matrix A (columns): A V = U S; returns
// U S (columns, by decreasing singular value), V (p x p, V[row][col]) and S
void thin_svd(const Cols &A, Cols &US, std::vector<std::vector<double>> &V, std::vector<double> &S) {
  const int p = (int)A.size();
  Cols W = A;
  std::vector<std::vector<double>> Vr(p, std::vector<double>(p, 0.0));
  for (int c = 0; c < p; c++) Vr[c][c] = 1.0;
  for (int sweep = 0; sweep < 60; sweep++) {
    bool rotated = false;
    for (int a = 0; a < p - 1; a++)
      for (int b = a + 1; b < p; b++) {
        const double al = dot(W[a], W[a]), be = dot(W[b], W[b]), ga = dot(W[a], W[b]);
        if (ga == 0 || std::fabs(ga) <= 1e-15 * std::sqrt(al * be)) continue;
        rotated = true;
        const double zeta = (be - al) / (2 * ga);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
        const double c = 1 / std::sqrt(1 + t * t), s = c * t;
        for (size_t i = 0; i < W[a].size(); i++) {
          const double wa = W[a][i], wb = W[b][i];
          W[a][i] = c * wa - s * wb;
          W[b][i] = s * wa + c * wb;
        }
        for (int r = 0; r < p; r++) {
          const double va = Vr[r][a], vb = Vr[r][b];
          Vr[r][a] = c * va - s * vb;
          Vr[r][b] = s * va + c * vb;
        }
      }
    if (!rotated) break;
  }
  std::vector<double> sig(p);
  for (int c = 0; c < p; c++) sig[c] = norm2(W[c]);
  std::vector<int> order(p);
  for (int c = 0; c < p; c++) order[c] = c;
  std::stable_sort(order.begin(), order.end(), [&](int u, int v) { return sig[u] > sig[v]; });
  US.assign(p, {});
  V.assign(p, std::vector<double>(p, 0.0));
  S.assign(p, 0.0);
  for (int j = 0; j < p; j++) {
    US[j] = W[order[j]];
    S[j] = sig[order[j]];
    for (int r = 0; r < p; r++) V[r][j] = Vr[r][order[j]];
  }
}

}  // namespace

int brr_solve(scf_system *S, scf_solve_info *info) {
  const auto t0 = std::chrono::steady_clock::now();
  const DevSys &h = S->h;
  const size_t n = (size_t)h.nit * h.M;
  const int m = S->spec.solver.m, p = m + 1;
  const double tol = h.tol, dmax = h.deltamax;
  const int itlim = h.itlim;
  std::vector<double> x(n), g(n), x0(n), g0(n), sdir(n), y(n), w(n);
  Cols C(p, std::vector<double>(n, 0.0)), D(p, std::vector<double>(n, 0.0));
  HIP_TRY(hipMemcpyAsync(x.data(), h.x, n * sizeof(double), hipMemcpyDeviceToHost, S->stream));
  HIP_TRY(hipStreamSynchronize(S->stream));
  double res;
  int status;
  int rc = evaluate(S, x.data(), g.data(), &res, &status, false);
  if (rc != SCF_OK) return rc;
  double error = norm2(g);
  int k = 0, it = 0;
  bool broke = false;
  // sum_{j=1..kk} A_j (B_j . v) into out (out starts at zero)
  auto apply = [&](const Cols &A, const Cols &B, int kk, const std::vector<double> &v, std::vector<double> &out) {
    std::fill(out.begin(), out.end(), 0.0);
    for (int j = 1; j <= kk; j++) {
      const double c = dot(B[j], v);
      for (size_t i = 0; i < n; i++) out[i] += A[j][i] * c;
    }
  };
  while (status == DS_OK && it < itlim && error > tol) {
    x0 = x;
    apply(C, D, k, g, w);
    for (size_t i = 0; i < n; i++) sdir[i] = -g[i] + w[i];
    g0 = g;
    double alpha = error > 1 ? dmax : dmax - it * (1.0 - dmax) / 1e4 * std::log(error);
    if (alpha > 1) alpha = 1;
    if (alpha < 0) alpha = tol;
    for (size_t i = 0; i < n; i++) x[i] = x0[i] + alpha * sdir[i];
    rc = evaluate(S, x.data(), g.data(), &res, &status, false);
    if (rc != SCF_OK) return rc;
    if (status != DS_OK) break;
    error = norm2(g);
    for (size_t i = 0; i < n; i++) y[i] = g0[i] - g[i];
    it++;
    if (k == m) {
      // rank reduction: D = Q R, C <- C R^T, D <- Q; C = U S V^T, C <- U S, D <- D V
      Cols Q;
      std::vector<std::vector<double>> R, V;
      thin_qr(D, Q, R);
      Cols CR(p, std::vector<double>(n, 0.0));
      for (int j = 0; j < p; j++)
        for (int l = j; l < p; l++) {
          const double r = R[j][l];
          if (r != 0)
            for (size_t i = 0; i < n; i++) CR[j][i] += C[l][i] * r;
        }
      std::vector<double> Sv;
      Cols US;
      thin_svd(CR, US, V, Sv);
      Cols DV(p, std::vector<double>(n, 0.0));
      for (int j = 0; j < p; j++)
        for (int l = 0; l < p; l++) {
          const double v = V[l][j];
          if (v != 0)
            for (size_t i = 0; i < n; i++) DV[j][i] += Q[l][i] * v;
        }
      C = US;
      D = DV;
      std::fill(C[k].begin(), C[k].end(), 0.0);
      std::fill(D[k].begin(), D[k].end(), 0.0);
      k--;
      for (int i = m - 1; i > 1; i--)
        if (Sv[i] / Sv[0] < tol) {
          std::fill(C[k].begin(), C[k].end(), 0.0);
          std::fill(D[k].begin(), D[k].end(), 0.0);
          k--;
        }
    }
    k++;
    // new rank-one term from s and y
    apply(C, D, k - 1, y, C[k]);
    for (size_t i = 0; i < n; i++) C[k][i] -= y[i];
    const double sthy = dot(sdir, C[k]);
    apply(D, C, k - 1, sdir, D[k]);
    for (size_t i = 0; i < n; i++) D[k][i] -= sdir[i];
    const double nhts = norm2(D[k]);
    if (sthy == 0 || nhts == 0) {
      broke = true;  // the reference gives up here
      break;
    }
    for (size_t i = 0; i < n; i++) {
      C[k][i] = (sdir[i] - C[k][i]) / sthy * nhts;
      D[k][i] = D[k][i] / nhts;
    }
  }
  // leave the device state (fields, densities, F, Omega, mu) at the final x
  rc = evaluate(S, x.data(), nullptr, &res, &status, true);
  if (rc != SCF_OK) return rc;
  info->iterations = it;
  info->residual = res;
  info->status = status != DS_OK ? status : (!broke && error <= tol ? SCF_OK : SCF_NOT_CONVERGED);
  info->fp32_iterations = 0;
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  info->kernel_ms = ms;
  info->solve_ms = ms;
  S->have_results = status == DS_OK;
  return SCF_OK;
}

}  // namespace nscf
