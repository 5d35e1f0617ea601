// Fused SCF kernel: one workgroup owns one SCF system and runs, without
// returning to the host, either one residual evaluation or a complete DIIS
// solve (all iterations), followed by the thermodynamic summary.
//
// Why one workgroup per system: the propagator recursion over the chain
// contour s is strictly sequential, and for 1- and 2-gradient lattices
// (M <= a few 10^4 sites) one step is ~1 us of work for one CU; splitting a
// system over CUs would pay a grid-wide barrier per step.  Keeping the
// previous end-point distribution in LDS and synchronising with s_barrier
// makes a step a few hundred cycles, and many independent systems (the
// batched and served workloads) fill the GPU one workgroup each.
//
// Reference path replaced per system (one call = one Solve_scf::Solve):
//   SFNewton::iterate_DIIS        src/sfnewton.cpp:1037-1130
//   System::Classical_residual    src/system.cpp:2087-2189
//   Molecule::ComputePhi          src/mol_linear.cpp:131-149, src/molecule.cpp:2300-2615
//   System::CheckResults          src/system.cpp:2719-2808
#include <hip/hip_runtime.h>

#include "scf_device.h"
#include "scf_kernels.h"

namespace nscf {

// ------------------------------------------------------------- reductions
// Workgroup sum; every thread gets the same value (summed in a fixed order).
__device__ __forceinline__ double wg_sum(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; w++) t += red[w];
  return t;
}

__device__ __forceinline__ double wg_max(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; w++) t = fmax(t, red[w]);
  return t;
}

// ------------------------------------------------------------- propagation
// Forward and backward recursion of one linear chain with the composition
// law fused into the backward step.  rows0/rows1 are the two LDS row
// buffers (size*M each) or null, in which case the rows are read from the
// end-point store in HBM (Gf) and from the Gb ping-pong.
struct BlockWalk {
  int b, j;  // current block and position inside it
};

__device__ void chain_markov1(const DevSys &S, int mo, double *rows0, double *rows1, double *red) {
  const int M = S.M, W = n_work_sites(S);
  const int tid = threadIdx.x, nt = blockDim.x;
  const int b0 = S.mol_block_begin[mo], b1 = S.mol_block_begin[mo + 1];
  const int N = S.mol_N[mo];
  double *Gf = S.Gf;
  const bool lds = rows0 != nullptr;
  // s = 0: G(.,1) = G1 (Lattice::Initiate)
  {
    const double *g1 = S.G1 + S.block_seg[b0] * M;
    for (int w = tid; w < W; w += nt) {
      int x, y, z;
      const int i = work_site(S, w, x, y, z);
      const double v = g1[i];
      Gf[i] = v;
      if (lds) rows0[i] = v;
    }
  }
  __syncthreads();
  const double *prev = lds ? rows0 : Gf;
  double *next = rows1;
  int s = 1;
  for (int b = b0; b < b1; b++) {
    const double *g1 = S.G1 + S.block_seg[b] * M;
    for (int j = (b == b0 ? 1 : 0); j < S.block_n[b]; j++, s++) {
      double *gs = Gf + (size_t)s * M;
      for (int w = tid; w < W; w += nt) {
        int x, y, z;
        const int i = work_site(S, w, x, y, z);
        const double v = stencil_prop<double>(S, prev, i, x, y, z) * g1[i];
        gs[i] = v;
        if (lds) next[i] = v;
      }
      __syncthreads();
      if (lds) {
        double *t = (double *)prev;
        prev = next;
        next = t;
      } else {
        prev = gs;
      }
    }
  }
  // GN = weighted sum of the last end-point distribution (Lattice::ComputeGN)
  {
    const double *gl = Gf + (size_t)(N - 1) * M;
    double acc = 0;
    for (int w = tid; w < W; w += nt) {
      int x, y, z;
      const int i = work_site(S, w, x, y, z);
      if (is_interior(S, x, y, z)) acc += site_weight(S, i) * gl[i];
    }
    const double gn = wg_sum(acc, red);
    if (tid == 0) S.sc->GN[mo] = gn;
  }
  // backward with composition: rho_mm += Gf(s) * Gb(s)
  double *gbA = lds ? rows0 : S.Gb, *gbB = lds ? rows1 : S.Gb + M;
  {
    const int bl = b1 - 1;
    const double *g1 = S.G1 + S.block_seg[bl] * M;
    double *rho = S.rho + S.block_mm[bl] * M;
    const double *gf = Gf + (size_t)(N - 1) * M;
    for (int w = tid; w < W; w += nt) {
      int x, y, z;
      const int i = work_site(S, w, x, y, z);
      const double v = g1[i];
      gbA[i] = v;
      rho[i] += gf[i] * v;
    }
  }
  __syncthreads();
  s = N - 2;
  for (int b = b1 - 1; b >= b0; b--) {
    const double *g1 = S.G1 + S.block_seg[b] * M;
    double *rho = S.rho + S.block_mm[b] * M;
    for (int j = (b == b1 - 1 ? S.block_n[b] - 2 : S.block_n[b] - 1); j >= 0; j--, s--) {
      const double *gf = Gf + (size_t)s * M;
      for (int w = tid; w < W; w += nt) {
        int x, y, z;
        const int i = work_site(S, w, x, y, z);
        const double v = stencil_prop<double>(S, gbA, i, x, y, z) * g1[i];
        gbB[i] = v;
        rho[i] += gf[i] * v;
      }
      __syncthreads();
      double *t = gbA;
      gbA = gbB;
      gbB = t;
    }
  }
}

// Second-order Markov chain (1 gradient, three direction components).
__device__ void chain_markov2(const DevSys &S, int mo, double *rows0, double *rows1, double *red) {
  const int M = S.M, W = S.M;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int b0 = S.mol_block_begin[mo], b1 = S.mol_block_begin[mo + 1];
  const int N = S.mol_N[mo];
  const int R = 3 * M;
  const double *P = S.mol_P[mo];
  double *Gf = S.Gf;
  const bool lds = rows0 != nullptr;
  {
    const double *g1 = S.G1 + S.block_seg[b0] * M;
    for (int i = tid; i < W; i += nt) {
      const double v = g1[i];
      for (int c = 0; c < 3; c++) {
        Gf[c * M + i] = v;
        if (lds) rows0[c * M + i] = v;
      }
    }
  }
  __syncthreads();
  const double *prev = lds ? rows0 : Gf;
  double *next = rows1;
  int s = 1;
  for (int b = b0; b < b1; b++) {
    const double *g1 = S.G1 + S.block_seg[b] * M;
    for (int j = (b == b0 ? 1 : 0); j < S.block_n[b]; j++, s++) {
      double *gs = Gf + (size_t)s * R;
      for (int i = tid; i < W; i += nt) {
        double o[3];
        m2_step<double>(S, prev, P, true, i, o);
        const double w = g1[i];
        for (int c = 0; c < 3; c++) {
          const double v = o[c] * w;
          gs[c * M + i] = v;
          if (lds) next[c * M + i] = v;
        }
      }
      __syncthreads();
      if (lds) {
        double *t = (double *)prev;
        prev = next;
        next = t;
      } else {
        prev = gs;
      }
    }
  }
  {
    const double *gl = Gf + (size_t)(N - 1) * R;
    double a0 = 0, a1 = 0, a2 = 0;
    for (int i = tid; i < W; i += nt) {
      if (i >= 1 && i <= S.MX) {
        const double wt = site_weight(S, i);
        a0 += wt * gl[i];
        a1 += wt * gl[M + i];
        a2 += wt * gl[2 * M + i];
      }
    }
    double gn = wg_sum(a0, red);
    gn += 4.0 * wg_sum(a1, red);
    gn += wg_sum(a2, red);
    if (tid == 0) S.sc->GN[mo] = gn / 6.0;
  }
  double *gbA = lds ? rows0 : S.Gb, *gbB = lds ? rows1 : S.Gb + R;
  {
    const int bl = b1 - 1;
    const double *g1 = S.G1 + S.block_seg[bl] * M;
    double *rho = S.rho + S.block_mm[bl] * M;
    const double *gf = Gf + (size_t)(N - 1) * R;
    for (int i = tid; i < W; i += nt) {
      const double v = g1[i];
      double r = rho[i];
      for (int c = 0; c < 3; c++) gbA[c * M + i] = v;
      r += (1.0 / 6.0) * gf[i] * v;
      r += (4.0 / 6.0) * gf[M + i] * v;
      r += (1.0 / 6.0) * gf[2 * M + i] * v;
      rho[i] = r;
    }
  }
  __syncthreads();
  s = N - 2;
  for (int b = b1 - 1; b >= b0; b--) {
    const double *g1 = S.G1 + S.block_seg[b] * M;
    double *rho = S.rho + S.block_mm[b] * M;
    for (int j = (b == b1 - 1 ? S.block_n[b] - 2 : S.block_n[b] - 1); j >= 0; j--, s--) {
      const double *gf = Gf + (size_t)s * R;
      for (int i = tid; i < W; i += nt) {
        double o[3];
        m2_step<double>(S, gbA, P, false, i, o);
        const double w = g1[i];
        double r = rho[i];
        for (int c = 0; c < 3; c++) {
          const double v = o[c] * w;
          gbB[c * M + i] = v;
          r += (c == 1 ? 4.0 / 6.0 : 1.0 / 6.0) * gf[c * M + i] * v;
        }
        rho[i] = r;
      }
      __syncthreads();
      double *t = gbA;
      gbA = gbB;
      gbB = t;
    }
  }
}

// ------------------------------------------------------------- residual
// One Classical_residual: fields x -> g, densities, norms.  Returns ||g||
// (or max|g|) to every thread, or a negative value on a status error.
__device__ double fused_residual(const DevSys &S, double *rows0, double *rows1, double *red) {
  const int M = S.M, W = n_work_sites(S);
  const int tid = threadIdx.x, nt = blockDim.x;
  DevScalars &sc = *S.sc;
  for (int w = tid; w < W; w += nt) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    site_prepare(S, i, x, y, z);
    for (int mm = 0; mm < S.nmm; mm++) S.rho[mm * M + i] = 0.0;
  }
  __syncthreads();
  for (int mo = 0; mo < S.nmol; mo++) {
    if (S.mol_markov[mo] == 2)
      chain_markov2(S, mo, rows0, rows1, red);
    else
      chain_markov1(S, mo, rows0, rows1, red);
    __syncthreads();
  }
  if (tid == 0) {
    for (int i = 0; i < S.nmol; i++) {
      if (S.mol_freedom[i] == 0) sc.phibulk[i] = S.mol_phibulk_in[i];
      if (S.mol_freedom[i] == 1) sc.n[i] = S.mol_n_in[i];
    }
    scalars_norms(S);
  }
  __syncthreads();
  if (sc.status != DS_OK) return -1.0;
  for (int w = tid; w < W; w += nt) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    site_densities(S, i);
  }
  __syncthreads();
  double acc = 0;
  for (int w = tid; w < W; w += nt) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    if (!is_interior(S, x, y, z)) continue;
    if (S.max_g) {
      const double g2 = site_residual(S, i, x, y, z);
      (void)g2;
      for (int it = 0; it < S.nit; it++) acc = fmax(acc, fabs(S.g[it * M + i]));
    } else {
      acc += site_residual(S, i, x, y, z);
    }
  }
  const double r = S.max_g ? wg_max(acc, red) : sqrt(wg_sum(acc, red));
  if (tid == 0) sc.residual = r;
  return r;
}

// Thermodynamic summary at the state of the last residual.
__device__ void fused_thermo(const DevSys &S, double *red) {
  const int W = n_work_sites(S);
  const int tid = threadIdx.x, nt = blockDim.x;
  DevScalars &sc = *S.sc;
  double logc[kMaxMol];
  for (int mo = 0; mo < S.nmol; mo++) {
    const double N = S.mol_N[mo];
    logc[mo] = log(N * sc.n[mo] / sc.GN[mo]) / N;
  }
  double f = 0, gp = 0, th[kMaxMol];
  for (int mo = 0; mo < kMaxMol; mo++) th[mo] = 0;
  for (int w = tid; w < W; w += nt) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    if (!is_interior(S, x, y, z)) continue;
    double F, GP;
    site_thermo(S, i, logc, F, GP);
    const double wt = site_weight(S, i);
    f += wt * F;
    gp += wt * GP;
    for (int mo = 0; mo < S.nmol; mo++) th[mo] += wt * S.mol_phitot[mo * S.M + i];
  }
  f = wg_sum(f, red);
  gp = wg_sum(gp, red);
  for (int mo = 0; mo < S.nmol; mo++) th[mo] = wg_sum(th[mo], red);
  if (tid == 0) {
    sc.F = f;
    sc.GP = gp;
    for (int mo = 0; mo < S.nmol; mo++) sc.theta[mo] = th[mo];
    scalars_mu(S);
  }
  __syncthreads();
}

// ------------------------------------------------------------- DIIS solve
__device__ void fused_solve(const DevSys &S, double *rows0, double *rows1, double *red) {
  const int nvar = S.nit * S.M;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int m = S.m;
  const double dmax = S.deltamax;
  DevScalars &sc = *S.sc;
  double *x = S.x, *g = S.g, *x0 = S.x0, *xR = S.xR, *xx0 = S.xx0;
  for (int v = tid; v < nvar; v += nt) x0[v] = x[v];
  __syncthreads();
  double res = fused_residual(S, rows0, rows1, red);
  if (res < 0) return;
  for (int v = tid; v < nvar; v += nt) {
    const double xv = x[v] + (-dmax) * g[v];
    x[v] = xv;
    xx0[v] = xv - x0[v];
    xR[v] = xv;
  }
  int it = 0, k_diis = 0;
  __syncthreads();
  while (res > S.tol && it < S.itlim) {
    it++;
    for (int v = tid; v < nvar; v += nt) x0[v] = x[v];
    __syncthreads();
    res = fused_residual(S, rows0, rows1, red);
    if (res < 0) break;
    const int k = it % m;
    k_diis++;
    double *xRk = xR + (size_t)k * nvar, *xxk = xx0 + (size_t)k * nvar;
    for (int v = tid; v < nvar; v += nt) {
      const double xv = x[v] + (-dmax) * g[v];
      x[v] = xv;
      xRk[v] = xv;
      xxk[v] = xv - x0[v];
    }
    __syncthreads();
    const int kd = k_diis > m ? m : k_diis;
    double dots[kMaxDIIS];
    for (int i = 0; i < kd; i++) {
      int posi = k - kd + 1 + i;
      if (posi < 0) posi += m;
      const double *a = xx0 + (size_t)posi * nvar;
      double acc = 0;
      for (int v = tid; v < nvar; v += nt) acc += a[v] * xxk[v];
      dots[i] = wg_sum(acc, red);
    }
    if (tid == 0) {
      int kdc;
      if (!diis_coefficients(sc, m, k_diis, dots, kdc)) sc.status = DS_NAN;
    }
    __syncthreads();
    if (sc.status != DS_OK) break;
    for (int v = tid; v < nvar; v += nt) {
      int posi = k - kd + 1;
      if (posi < 0) posi += m;
      double xv = 0.0 + sc.C[0] * xR[(size_t)posi * nvar + v];
      for (int i = 1; i < kd; i++) {
        posi = k - kd + 1 + i;
        if (posi < 0) posi += m;
        xv += sc.C[i] * xR[(size_t)posi * nvar + v];
      }
      x[v] = xv;
    }
    __syncthreads();
  }
  if (tid == 0) {
    sc.iterations = it;
    sc.converged = (sc.status == DS_OK && res <= S.tol) ? 1 : 0;
  }
  __syncthreads();
}

// ------------------------------------------------------------- kernel
__global__ void __launch_bounds__(1024) k_fused(DevSys *const *systems, int mode, int use_lds) {
  const DevSys &S = *systems[blockIdx.x];
  extern __shared__ double lds[];
  __shared__ double red[32];
  double *rows0 = nullptr, *rows1 = nullptr;
  if (use_lds && S.use_lds) {
    int R = 0;
    for (int mo = 0; mo < S.nmol; mo++) R = R > S.mol_size[mo] * S.M ? R : S.mol_size[mo] * S.M;
    rows0 = lds;
    rows1 = lds + R;
  }
  if (threadIdx.x == 0) S.sc->status = DS_OK;
  __syncthreads();
  if (mode == FUSED_RESIDUAL) {
    const double r = fused_residual(S, rows0, rows1, red);
    if (r >= 0) fused_thermo(S, red);
  } else {
    fused_solve(S, rows0, rows1, red);
    if (S.sc->status == DS_OK) fused_thermo(S, red);
  }
}

hipError_t launch_fused(DevSys *const *d_systems, int nsys, int mode, int threads, size_t lds_bytes,
                        hipStream_t stream) {
  if (lds_bytes > 0) {
    hipError_t e = hipFuncSetAttribute((const void *)k_fused, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds_bytes);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_fused, dim3(nsys), dim3(threads), lds_bytes, stream, d_systems, mode,
                     lds_bytes > 0 ? 1 : 0);
  return hipGetLastError();
}

}  // namespace nscf
