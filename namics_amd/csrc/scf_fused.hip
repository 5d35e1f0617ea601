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

// Work is distributed over the system's own thread count (S.threads), not
// the launch's: in a batch launched with more threads, the extra threads
// take no sites and add exact zeros to the reductions, so a system does the
// same arithmetic in the same order as when it is launched alone.
__device__ __forceinline__ int work_tid(const DevSys &S) {
  return (int)threadIdx.x < S.threads ? (int)threadIdx.x : (1 << 30);
}

// ------------------------------------------------------------- propagation
constexpr int kStripX = 8;    // sites per thread along x in the 2-gradient step
constexpr int kMaxLamX = 1026;  // radial layers + 2 held in the LDS weight table

// Forward and backward recursion of one linear chain with the composition
// law fused into the backward step.  rows0/rows1 are the two LDS row
// buffers (size*M each) or null, in which case the rows are read from the
// end-point store in HBM (Gf) and from the Gb ping-pong.
struct BlockWalk {
  int b, j;  // current block and position inside it
};

__device__ __noinline__ void chain_markov1(const DevSys &S, int mo, double *rows0, double *rows1, double *red) {
  const int M = S.M, W = n_work_sites(S);
  const int tid = work_tid(S), nt = S.threads;
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

// 2-gradient Markov-1 chain, register-blocked.  Thread t owns the strip of
// SX sites (x0..x0+SX-1, y) with y = t % MY + 1: lanes of a wavefront walk the
// contiguous y axis, so every LDS row read is unit-stride across lanes.  A
// step reads the three rows y-1, y, y+1 over x0-1..x0+SX (3(SX+2) LDS reads
// for SX sites instead of 9 per site), multiplies by the Boltzmann weights
// held in registers for the current block, writes the end-point
// distribution to HBM and the next LDS row.  The density of the current
// block accumulates in registers and is flushed at block boundaries.
// The 9-point weights are those of LGrad2::propagate (src/LGrad2.cpp:1058-1073);
// the planar full stencil (src/LG2Planar.cpp:1085-1096) is the same form with
// lambda- = lambda+ = 1/6.
// One lattice step of a strip from row `src`: v = W(src) * g1 (rolling
// 3x3 window over x, rows y-1, y, y+1).
// Keeps a value opaque to the optimiser so address arithmetic derived from
// it is recomputed inside the step loop instead of being hoisted into (and
// spilling) registers.
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

template <int SX>
__device__ __forceinline__ void strip_step(const double *__restrict__ src, int first, int JX, int last, int ym,
                                           int y, int yp, int cnt, int x0, const double2 *lam, double C1,
                                           double C2, double C3, const double (&g1)[SX], double (&v)[SX]) {
  // columns: first (mapped x0-1), x0 .. x0+SX-1 at stride JX, last (mapped x0+cnt)
  const int dm = opaque(ym - y), dp = opaque(yp - y);
  const double *c0p = src + first + y;
  double am = c0p[dm], bm = c0p[0], cm = c0p[dp];
  const double *cp_ = src + x0 * JX + y;
  double a0 = cp_[dm], b0 = cp_[0], c0 = cp_[dp];
#pragma unroll
  for (int k = 1; k <= SX; k++) {
    const double *q = (k < SX) ? cp_ + k * JX : src + last + y;
    const double ap = q[dm], bp = q[0], cp = q[dp];
    const double2 lw = lam[opaque(x0 + k - 1)];
    const double lmk = lw.x, lpk = lw.y;
    double t = C1 * b0;
    t += lmk * bm;
    t += lpk * bp;
    t += C2 * a0;
    t += C2 * c0;
    t *= C3;
    t += lmk * am;
    t += lmk * cm;
    t += lpk * ap;
    t += lpk * cp;
    t *= C2;
    v[k - 1] = t * g1[k - 1];
    am = a0;
    a0 = ap;
    bm = b0;
    b0 = bp;
    cm = c0;
    c0 = cp;
  }
}

template <int SX>
__device__ __noinline__ void chain2d_m1(const DevSys &S, int mo, double *rows0, double *rows1, double *red,
                                        double2 *lam) {
  const int MX = S.MX, MY = S.MY, JX = S.JX;
  const int tid = threadIdx.x;
  const int nsx = (MX + SX - 1) / SX;
  const bool active = tid < nsx * MY;
  const int y = active ? tid % MY + 1 : 1;
  const int x0 = active ? (tid / MY) * SX + 1 : 1;
  const int cnt = SX;  // the caller guarantees MX % SX == 0
  const int ym = map_y(S, y - 1), yp = map_y(S, y + 1);
  const bool cyl = S.geometry != 0;
  const bool sc = S.lattice_type == 0;
  const double C1 = sc ? 4.0 / 6.0 : 0.5, C2 = sc ? 1.0 / 6.0 : 0.25, C3 = sc ? 4.0 : 2.0;
  const int first = map_x(S, x0 - 1) * JX, last = map_x(S, x0 + cnt) * JX;
  const int b0 = S.mol_block_begin[mo], b1 = S.mol_block_begin[mo + 1];
  const int N = S.mol_N[mo];
  const size_t M = S.M;
  const bool lds = rows0 != nullptr;
  double *Gf = S.Gf;
  // radial weights per x (wave-uniform reads from LDS): (lambda-, lambda+)
  for (int x = tid; x < MX + 2; x += blockDim.x)
    lam[x] = cyl ? make_double2(S.lam_m[x * JX + 1], S.lam_p[x * JX + 1]) : make_double2(1.0 / 6.0, 1.0 / 6.0);
  __syncthreads();
  const int base = x0 * JX + y;  // site of k = 0; site k is base + k*JX
  double g1[SX], v[SX], racc[SX];

#define LOAD_G1(seg)                                              \
  {                                                               \
    const double *p_ = S.G1 + (seg) * M + base;                   \
    _Pragma("unroll") for (int k = 0; k < SX; k++) g1[k] = p_[k * JX]; \
  }
#define STORE_V(row)                                              \
  {                                                               \
    double *r_ = (row) + opaque(base);                            \
    _Pragma("unroll") for (int k = 0; k < SX; k++) r_[k * JX] = v[k]; \
  }

  // ---- forward
  LOAD_G1(S.block_seg[b0]);
#pragma unroll
  for (int k = 0; k < SX; k++) v[k] = g1[k];
  if (active) {
    STORE_V(Gf);
    if (lds) STORE_V(rows0);
  }
  __syncthreads();
  const double *prev = lds ? rows0 : Gf;
  double *next = rows1;
  int s = 1;
  for (int b = b0; b < b1; b++) {
    LOAD_G1(S.block_seg[b]);
    for (int j = (b == b0 ? 1 : 0); j < S.block_n[b]; j++, s++) {
      double *gs = Gf + (size_t)s * M;
      if (active) {
        strip_step<SX>(prev, first, JX, last, ym, y, yp, cnt, x0, lam, C1, C2, C3, g1, v);
        STORE_V(gs);
        if (lds) STORE_V(next);
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
  // GN: weighted sum of the last distribution (held in v)
  {
    double acc = 0;
    if (active) {
      const double *Lw = S.L + base;
#pragma unroll
      for (int k = 0; k < SX; k++)
        acc += (cyl ? Lw[k * JX] : 1.0) * v[k];
    }
    const double gn = wg_sum(acc, red);
    if (tid == 0) S.sc->GN[mo] = gn;
  }
  // ---- backward with composition; rho of the current block in registers
  double *ga = lds ? rows0 : S.Gb, *gb = lds ? rows1 : S.Gb + M;
  LOAD_G1(S.block_seg[b1 - 1]);
#pragma unroll
  for (int k = 0; k < SX; k++) {
    v[k] = g1[k];
    racc[k] = 0.0;
  }
  if (active) {
    const double *gf = Gf + (size_t)(N - 1) * M + base;
#pragma unroll
    for (int k = 0; k < SX; k++)
      racc[k] += gf[k * JX] * v[k];
    STORE_V(ga);
  }
  __syncthreads();
  s = N - 2;
  for (int b = b1 - 1; b >= b0; b--) {
    if (b != b1 - 1) {
      LOAD_G1(S.block_seg[b]);
#pragma unroll
      for (int k = 0; k < SX; k++) racc[k] = 0.0;
    }
    for (int j = (b == b1 - 1 ? S.block_n[b] - 2 : S.block_n[b] - 1); j >= 0; j--, s--) {
      if (active) {
        strip_step<SX>(ga, first, JX, last, ym, y, yp, cnt, x0, lam, C1, C2, C3, g1, v);
        const double *gf = Gf + (size_t)s * M + base;
#pragma unroll
        for (int k = 0; k < SX; k++) racc[k] += gf[k * JX] * v[k];
        STORE_V(gb);
      }
      __syncthreads();
      double *t = ga;
      ga = gb;
      gb = t;
    }
    if (active) {
      double *rho = S.rho + S.block_mm[b] * M + base;
#pragma unroll
      for (int k = 0; k < SX; k++) rho[k * JX] += racc[k];
    }
  }
#undef LOAD_G1
#undef STORE_V
}

// Second-order Markov chain (1 gradient, three direction components).
__device__ __noinline__ void chain_markov2(const DevSys &S, int mo, double *rows0, double *rows1, double *red) {
  const int M = S.M, W = S.M;
  const int tid = work_tid(S), nt = S.threads;
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
  __shared__ double2 lam2d[kMaxLamX];
  const int M = S.M, W = n_work_sites(S);
  const int tid = work_tid(S), nt = S.threads;
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
    else if (S.gradients == 2 && (S.geometry != 0 || S.stencil_full) && S.MX + 2 <= kMaxLamX &&
             S.MX % 8 == 0 && (S.MX / 8) * S.MY <= S.threads)
      chain2d_m1<8>(S, mo, rows0, rows1, red, lam2d);
    else if (S.gradients == 2 && (S.geometry != 0 || S.stencil_full) && S.MX + 2 <= kMaxLamX &&
             S.MX % 4 == 0 && (S.MX / 4) * S.MY <= S.threads)
      chain2d_m1<4>(S, mo, rows0, rows1, red, lam2d);
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
__device__ __noinline__ void fused_thermo(const DevSys &S, double *red) {
  const int W = n_work_sites(S);
  const int tid = work_tid(S), nt = S.threads;
  DevScalars &sc = *S.sc;
  double logc[kMaxMol];
#pragma unroll
  for (int mo = 0; mo < kMaxMol; mo++) {
    logc[mo] = 0.0;
    if (mo < S.nmol) {
      const double N = S.mol_N[mo];
      logc[mo] = log(N * sc.n[mo] / sc.GN[mo]) / N;
    }
  }
  double f = 0, gp = 0, th[kMaxMol];
#pragma unroll
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
#pragma unroll
    for (int mo = 0; mo < kMaxMol; mo++)
      if (mo < S.nmol) th[mo] += wt * S.mol_phitot[mo * S.M + i];
  }
  f = wg_sum(f, red);
  gp = wg_sum(gp, red);
#pragma unroll
  for (int mo = 0; mo < kMaxMol; mo++)
    if (mo < S.nmol) th[mo] = wg_sum(th[mo], red);
  if (tid == 0) {
    sc.F = f;
    sc.GP = gp;
#pragma unroll
    for (int mo = 0; mo < kMaxMol; mo++)
      if (mo < S.nmol) sc.theta[mo] = th[mo];
    scalars_mu(S);
  }
  __syncthreads();
}

// ------------------------------------------------------------- DIIS solve
__device__ void fused_solve(const DevSys &S, double *rows0, double *rows1, double *red) {
  __shared__ double diis_dots[kMaxDIIS];
  __shared__ double diis_A[kMaxDIIS * kMaxDIIS];
  const int nvar = S.nit * S.M;
  const int tid = work_tid(S), nt = S.threads;
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
    for (int i = 0; i < kd; i++) {
      int posi = k - kd + 1 + i;
      if (posi < 0) posi += m;
      const double *a = xx0 + (size_t)posi * nvar;
      double acc = 0;
      for (int v = tid; v < nvar; v += nt) acc += a[v] * xxk[v];
      acc = wg_sum(acc, red);
      if (tid == 0) diis_dots[i] = acc;
    }
    __syncthreads();
    if (tid == 0) diis_gram_update(sc, m, k_diis, diis_dots, diis_A);
    __syncthreads();
    if (tid < 64) {
      const bool ok = wave_diis_coefficients(diis_A, kd, sc.C);
      if (tid == 0 && !ok) sc.status = DS_NAN;
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
  if (threadIdx.x == 0) {
    S.sc->status = DS_OK;
    S.sc->t_begin = wall_clock64();
  }
  __syncthreads();
  if (mode == FUSED_RESIDUAL) {
    const double r = fused_residual(S, rows0, rows1, red);
    if (r >= 0) fused_thermo(S, red);
  } else {
    fused_solve(S, rows0, rows1, red);
    if (S.sc->status == DS_OK) fused_thermo(S, red);
  }
  if (threadIdx.x == 0) S.sc->t_end = wall_clock64();
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
