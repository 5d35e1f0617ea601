// Fused SCF kernel: one workgroup owns one SCF system and runs, without
// returning to the host, either one residual evaluation or a complete DIIS
// solve (all iterations), followed by the thermodynamic summary.
//
// Why one workgroup per system: the propagator recursion over the chain
// contour s is strictly sequential, and for 1- and 2-gradient lattices
// (M <= a few 10^4 sites) one step is well under a microsecond of work for
// one CU; splitting a system over CUs would pay a cross-CU synchronisation
// per step.  Keeping the previous end-point distribution in LDS and
// synchronising with s_barrier makes a step a few hundred cycles, and many
// independent systems (the batched and served workloads) fill the GPU one
// workgroup each.
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

// ------------------------------------------------------------- LDS helpers
__device__ __forceinline__ ldbl *lds_row(int offset) {
  extern __shared__ double lds_dyn[];
  return (ldbl *)(lds_dyn + offset);
}

// Barrier that orders LDS only.  HBM stores issued before it may still be in
// flight: inside the step loops they are read back only by the thread that
// wrote them, and every chain ends with a full __syncthreads().
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Propagator rows in LDS (kLds) or, when they do not fit, in HBM.
template <bool kLds>
struct Rows;
template <>
struct Rows<true> {
  typedef ldbl T;
  static __device__ __forceinline__ void barrier() { lds_barrier(); }
};
template <>
struct Rows<false> {
  typedef gdbl T;
  static __device__ __forceinline__ void barrier() { __syncthreads(); }
};

constexpr int kMaxLamX = 1026;  // radial layers + 2 held in the LDS weight table

__shared__ double s_red[16];                                // one value per wave
__shared__ double s_lam[2 * kMaxLamX];                      // (lambda-, lambda+) per x
__shared__ double s_diis[kMaxDIIS + kMaxDIIS * kMaxDIIS];  // dots, Gram matrix

// Work is distributed over the system's own thread count (S.threads), not
// the launch's: in a batch launched with more threads, the extra threads
// take no sites and add exact zeros to the reductions, so a system does the
// same arithmetic in the same order as when it is launched alone.
template <typename Sys>
__device__ __forceinline__ int work_tid(const Sys &S) {
  return (int)threadIdx.x < S.threads ? (int)threadIdx.x : (1 << 30);
}

// ------------------------------------------------------------- reductions
// Workgroup sum; every thread gets the same value (summed in a fixed order).
__device__ __forceinline__ double wg_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  ldbl *red = (ldbl *)s_red;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  lds_barrier();
  if (lane == 0) red[wave] = v;
  lds_barrier();
  double t = 0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; w++) t += red[w];
  return t;
}

__device__ __forceinline__ double wg_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  ldbl *red = (ldbl *)s_red;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  lds_barrier();
  if (lane == 0) red[wave] = v;
  lds_barrier();
  double t = 0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; w++) t = fmax(t, red[w]);
  return t;
}

// ------------------------------------------------------------- propagation
// Forward recursion of one linear chain, then the backward recursion with
// the composition law fused into it: rho_block += Gf(s) * Gb(s)
// (Molecule::propagate_forward/backward src/molecule.cpp:2300-2446,
// mol_linear::ComputePhi src/mol_linear.cpp:131-149).  Gf keeps all N rows in
// HBM; the row being read and the row being written live in LDS when they
// fit.

// 1-gradient Markov-1 chain, one site per thread (M <= threads): the
// neighbour indices and lattice weights of the site stay in registers for
// the whole chain, and the density of a block is accumulated in a register.
// LG1Planar::propagate (src/LG1Planar.cpp:146-149) and LGrad1::propagate
// (src/LGrad1.cpp:441-445); a neighbour that falls on a ghost reads the site
// set_bounds copies there, a missing neighbour (the stencil at the outermost
// sites) has weight zero.
__device__ __noinline__ void chain1d_m1(const DevSys &S0, int mo) {
  CSys &S = uniform_sys(S0);
  const int M = S.M;
  const int i = threadIdx.x;
  const bool active = i < M;
  const int ii = active ? i : 0;
  const int i0 = map_x(S, ii);
  const int im = ii >= 1 ? map_x(S, ii - 1) : i0;
  const int ip = ii <= M - 2 ? map_x(S, ii + 1) : i0;
  const bool planar = S.geometry == 0;
  double wm, w0, wp;
  if (planar) {
    wm = S.lambda;
    w0 = 1.0 - 2.0 * S.lambda;
    wp = S.lambda;
  } else {
    wm = hbm(S.lam_m)[ii];
    w0 = hbm(S.lam_0)[ii];
    wp = hbm(S.lam_p)[ii];
  }
  if (ii < 1) wm = 0.0;
  if (ii > M - 2) wp = 0.0;
  const int b0 = S.mol_block_begin[mo], b1 = S.mol_block_begin[mo + 1];
  const int N = S.mol_N[mo];
  gdbl *Gf = hbm(S.Gf);
  const gdbl *G1 = hbm(S.G1);
  ldbl *ra = lds_row(0), *rb = lds_row(S.row_stride);

  double g1 = active ? G1[(size_t)S.block_seg[b0] * M + ii] : 0.0;
  double v = g1;
  if (active) {
    Gf[ii] = v;
    ra[ii] = v;
  }
  lds_barrier();
  ldbl *prev = ra, *next = rb;
  int s = 1;
  for (int b = b0; b < b1; b++) {
    if (b != b0 && active) g1 = G1[(size_t)S.block_seg[b] * M + ii];
    const int n = S.block_n[b] - (b == b0 ? 1 : 0);
    for (int j = 0; j < n; j++, s++) {
      if (active) {
        double t;
        if (planar) {
          t = wm * prev[im];
          t += w0 * prev[i0];
        } else {
          t = w0 * prev[i0];
          t += wm * prev[im];
        }
        t += wp * prev[ip];
        v = t * g1;
        Gf[(size_t)s * M + ii] = v;
        next[ii] = v;
      }
      lds_barrier();
      ldbl *tmp = prev;
      prev = next;
      next = tmp;
    }
  }
  // GN = weighted sum of the last end-point distribution (Lattice::ComputeGN)
  {
    double acc = 0.0;
    if (active && ii >= 1 && ii <= S.MX) acc = (planar ? 1.0 : hbm(S.L)[ii]) * v;
    const double gn = wg_sum(acc);
    if (threadIdx.x == 0) S0.sc->GN[mo] = gn;
  }
  // backward with composition, Gf(s-1) fetched one step ahead.  On a ghost
  // the composition reads Gf(s) as the reference leaves it: the value
  // set_bounds copied there when Gf(s) propagated to s+1 (the site i0), and
  // zero for the last row, whose ghosts ComputeGN's remove_bounds cleared
  // (src/LGrad1.cpp:213-221,863-874).  Only phitot on the ghosts sees this.
  ldbl *ga = ra, *gb = rb;
  const int bl = b1 - 1;
  const bool ghost = ii < 1 || ii > S.MX;
  g1 = active ? G1[(size_t)S.block_seg[bl] * M + ii] : 0.0;
  v = g1;
  double racc = 0.0, gf = 0.0;
  if (active) {
    racc = (ghost ? 0.0 : Gf[(size_t)(N - 1) * M + ii]) * v;
    ga[ii] = v;
    if (N >= 2) gf = Gf[(size_t)(N - 2) * M + i0];
  }
  lds_barrier();
  s = N - 2;
  for (int b = bl; b >= b0; b--) {
    if (b != bl) {
      if (active) g1 = G1[(size_t)S.block_seg[b] * M + ii];
      racc = 0.0;
    }
    const int n = S.block_n[b] - (b == bl ? 1 : 0);
    for (int j = 0; j < n; j++, s--) {
      if (active) {
        double t;
        if (planar) {
          t = wm * ga[im];
          t += w0 * ga[i0];
        } else {
          t = w0 * ga[i0];
          t += wm * ga[im];
        }
        t += wp * ga[ip];
        v = t * g1;
        racc += gf * v;
        gb[ii] = v;
        if (s >= 1) gf = Gf[(size_t)(s - 1) * M + i0];
      }
      lds_barrier();
      ldbl *tmp = ga;
      ga = gb;
      gb = tmp;
    }
    if (active) hbm(S.rho)[(size_t)S.block_mm[b] * M + ii] += racc;
  }
  __syncthreads();
}

// Ring of a first-order chain on a 1-gradient lattice (fjc = 1), one site
// per thread (mol_linear::ComputePhi src/mol_linear.cpp:35-111).  The first
// and last blocks are one segment each of the same type, merged: for every
// interior site x where that segment may sit, a chain starts there
// (row 0 = delta_x G1), runs through the middle blocks (rows 1 .. N-1, kept
// in HBM) and ends on x (row N, needed on x only); GN sums the end weights
// over x.  The merged segment's density on x is G1 times that end weight;
// the middle blocks' densities come from the backward pass started at
// delta_x G1, with the same stencil, ghost reads and composition order as
// chain1d_m1.  Densities and GN collect over the starts in registers.
__device__ __noinline__ void chain1d_ring(const DevSys &S0, int mo) {
  CSys &S = uniform_sys(S0);
  const int M = S.M;
  const int i = threadIdx.x;
  const bool active = i < M;
  const int ii = active ? i : 0;
  const int i0 = map_x(S, ii);
  const int im = ii >= 1 ? map_x(S, ii - 1) : i0;
  const int ip = ii <= M - 2 ? map_x(S, ii + 1) : i0;
  const bool planar = S.geometry == 0;
  double wm, w0, wp;
  if (planar) {
    wm = S.lambda;
    w0 = 1.0 - 2.0 * S.lambda;
    wp = S.lambda;
  } else {
    wm = hbm(S.lam_m)[ii];
    w0 = hbm(S.lam_0)[ii];
    wp = hbm(S.lam_p)[ii];
  }
  if (ii < 1) wm = 0.0;
  if (ii > M - 2) wp = 0.0;
  auto step = [&](const ldbl *row) {
    double t;
    if (planar) {
      t = wm * row[im];
      t += w0 * row[i0];
    } else {
      t = w0 * row[i0];
      t += wm * row[im];
    }
    t += wp * row[ip];
    return t;
  };
  const int b0 = S.mol_block_begin[mo], b1 = S.mol_block_begin[mo + 1];
  const int N = S.mol_N[mo];
  gdbl *Gf = hbm(S.Gf);
  const gdbl *G1 = hbm(S.G1);
  ldbl *ra = lds_row(0), *rb = lds_row(S.row_stride);
  const int sa = S.block_seg[b0];
  const double g1a = active ? G1[(size_t)sa * M + ii] : 0.0;
  const double wsite = active && ii >= 1 && ii <= S.MX ? (planar ? 1.0 : hbm(S.L)[ii]) : 0.0;
  double gnacc = 0.0, racc_a = 0.0;
  for (int x = 1; x <= S.MX; x++) {
    if (!(G1[(size_t)sa * M + x] > 0)) continue;  // the same on every thread
    if (active) ra[ii] = ii == x ? g1a : 0.0;
    lds_barrier();
    ldbl *prev = ra, *next = rb;
    int s = 1;
    for (int b = b0 + 1; b < b1 - 1; b++) {
      const double g1 = active ? G1[(size_t)S.block_seg[b] * M + ii] : 0.0;
      for (int j = 0; j < S.block_n[b]; j++, s++) {
        if (active) {
          const double v = step(prev) * g1;
          Gf[(size_t)s * M + ii] = v;
          next[ii] = v;
        }
        lds_barrier();
        ldbl *tmp = prev;
        prev = next;
        next = tmp;
      }
    }
    // the merged end segment on x (the row masked to x)
    double last = 0.0;
    if (active && ii == x) last = step(prev) * g1a;
    gnacc += wsite * last;
    racc_a += g1a * last;
    lds_barrier();  // every thread has read prev before it is overwritten
    // backward from delta_x G1 through the middle blocks
    ldbl *ga = prev, *gb = next;
    if (active) ga[ii] = ii == x ? g1a : 0.0;
    double gf = 0.0;
    if (active && N - 1 >= 1) gf = Gf[(size_t)(N - 1) * M + i0];
    lds_barrier();
    s = N - 1;
    for (int b = b1 - 2; b >= b0 + 1; b--) {
      const double g1 = active ? G1[(size_t)S.block_seg[b] * M + ii] : 0.0;
      double racc = 0.0;
      for (int j = 0; j < S.block_n[b]; j++, s--) {
        if (active) {
          const double v = step(ga) * g1;
          racc += gf * v;
          gb[ii] = v;
          if (s - 1 >= 1) gf = Gf[(size_t)(s - 1) * M + i0];
        }
        lds_barrier();
        ldbl *tmp = ga;
        ga = gb;
        gb = tmp;
      }
      if (active) hbm(S.rho)[(size_t)S.block_mm[b] * M + ii] += racc;
    }
  }
  if (active) hbm(S.rho)[(size_t)S.block_mm[b0] * M + ii] += racc_a;
  const double gn = wg_sum(gnacc);
  if (threadIdx.x == 0) S0.sc->GN[mo] = gn;
  __syncthreads();
}

// Second-order (Markov 2) chain on a 1-gradient simple-cubic lattice, one
// site per thread (M <= threads), rows in LDS: three direction components per
// site (+x, lateral, -x), the stencil of m2_step, composition weights
// (1/6, 4/6, 1/6) (LGrad1::AddPhiS src/LGrad1.cpp:876-910),
// GN = (G0 + 4 G1 + G2) / 6 (LGrad1::ComputeGN src/LGrad1.cpp:943-955).
// The arithmetic is chain_markov2's, in the same order; what differs is that
// no global access sits on a step's critical path: the Boltzmann weight of a
// block and its density stay in registers (the density is read once at the
// block's start and written once at its end), and Gf of the backward pass is
// fetched one step ahead.
template <int kSPT>
__device__ __noinline__ void chain1d_m2(const DevSys &S0, int mo) {
  CSys &S = uniform_sys(S0);
  const int M = S.M;
  const int tid = work_tid(S), nt = S.threads;
  const int R = 3 * M;
  double P[kMaxComp];
#pragma unroll
  for (int k = 0; k < kMaxComp; k++) P[k] = S.mol_P[mo][k];
  constexpr double wc[3] = {1.0 / 6.0, 4.0 / 6.0, 1.0 / 6.0};
  int ii[kSPT];
  bool act[kSPT];
#pragma unroll
  for (int k = 0; k < kSPT; k++) {
    const int w = tid + k * nt;
    act[k] = w < M;
    ii[k] = act[k] ? w : 0;
  }
  const int b0 = S.mol_block_begin[mo], b1 = S.mol_block_begin[mo + 1];
  const int N = S.mol_N[mo];
  gdbl *Gf = hbm(S.Gf);
  const gdbl *G1 = hbm(S.G1);
  ldbl *ra = lds_row(0), *rb = lds_row(S.row_stride);
  double g1[kSPT], o[kSPT][3];
#pragma unroll
  for (int k = 0; k < kSPT; k++) {
    g1[k] = act[k] ? G1[(size_t)S.block_seg[b0] * M + ii[k]] : 0.0;
    if (act[k])
#pragma unroll
      for (int c = 0; c < 3; c++) {
        Gf[c * M + ii[k]] = g1[k];
        ra[c * M + ii[k]] = g1[k];
      }
  }
  lds_barrier();
  ldbl *prev = ra, *next = rb;
  int s = 1;
  for (int b = b0; b < b1; b++) {
    if (b != b0) {
#pragma unroll
      for (int k = 0; k < kSPT; k++)
        if (act[k]) g1[k] = G1[(size_t)S.block_seg[b] * M + ii[k]];
    }
    for (int j = (b == b0 ? 1 : 0); j < S.block_n[b]; j++, s++) {
#pragma unroll
      for (int k = 0; k < kSPT; k++) {
        m2_step<double>(S, (const ldbl *)prev, P, true, ii[k], o[k]);
#pragma unroll
        for (int c = 0; c < 3; c++) o[k][c] *= g1[k];
      }
      gdbl *gs = Gf + (size_t)s * R;
#pragma unroll
      for (int k = 0; k < kSPT; k++)
        if (act[k])
#pragma unroll
          for (int c = 0; c < 3; c++) {
            gs[c * M + ii[k]] = o[k][c];
            next[c * M + ii[k]] = o[k][c];
          }
      lds_barrier();
      ldbl *t = prev;
      prev = next;
      next = t;
    }
  }
  {
    double a[3] = {0.0, 0.0, 0.0};
    const gdbl *gl = Gf + (size_t)(N - 1) * R;
#pragma unroll
    for (int k = 0; k < kSPT; k++)
      if (act[k] && ii[k] >= 1 && ii[k] <= S.MX) {
        const double wt = site_weight(S0, ii[k]);
#pragma unroll
        for (int c = 0; c < 3; c++) a[c] += wt * gl[c * M + ii[k]];
      }
    double gn = wg_sum(a[0]);
    gn += 4.0 * wg_sum(a[1]);
    gn += wg_sum(a[2]);
    if (threadIdx.x == 0) S0.sc->GN[mo] = gn / 6.0;
  }
  // backward with composition; Gf(s) of every component fetched one step
  // ahead, the block density in registers
  ldbl *gbA = ra, *gbB = rb;
  const int bl = b1 - 1;
  double racc[kSPT], gf[kSPT][3];
  {
    const gdbl *gfl = Gf + (size_t)(N - 1) * R;
#pragma unroll
    for (int k = 0; k < kSPT; k++) {
      g1[k] = act[k] ? G1[(size_t)S.block_seg[bl] * M + ii[k]] : 0.0;
      racc[k] = act[k] ? hbm(S.rho)[(size_t)S.block_mm[bl] * M + ii[k]] : 0.0;
      // ghosts: the last row reads zero (ComputeGN's remove_bounds), earlier
      // rows the components set_bounds(X, Y) paired there (m2_val)
      const bool ghost = ii[k] < 1 || ii[k] > S.MX;
      if (act[k]) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
          gbA[c * M + ii[k]] = g1[k];
          racc[k] += wc[c] * (ghost ? 0.0 : gfl[c * M + ii[k]]) * g1[k];
        }
      }
#pragma unroll
      for (int c = 0; c < 3; c++)
        gf[k][c] = (act[k] && N >= 2) ? m2_val<double>(S, Gf + (size_t)(N - 2) * R, c, ii[k]) : 0.0;
    }
  }
  lds_barrier();
  s = N - 2;
  for (int b = bl; b >= b0; b--) {
    if (b != bl) {
#pragma unroll
      for (int k = 0; k < kSPT; k++) {
        if (act[k]) {
          hbm(S.rho)[(size_t)S.block_mm[b + 1] * M + ii[k]] = racc[k];
          g1[k] = G1[(size_t)S.block_seg[b] * M + ii[k]];
          racc[k] = hbm(S.rho)[(size_t)S.block_mm[b] * M + ii[k]];
        }
      }
    }
    for (int j = (b == bl ? S.block_n[b] - 2 : S.block_n[b] - 1); j >= 0; j--, s--) {
#pragma unroll
      for (int k = 0; k < kSPT; k++) {
        m2_step<double>(S, (const ldbl *)gbA, P, false, ii[k], o[k]);
#pragma unroll
        for (int c = 0; c < 3; c++) {
          o[k][c] *= g1[k];
          racc[k] += wc[c] * gf[k][c] * o[k][c];
        }
      }
#pragma unroll
      for (int k = 0; k < kSPT; k++)
        if (act[k])
#pragma unroll
          for (int c = 0; c < 3; c++) {
            gbB[c * M + ii[k]] = o[k][c];
            if (s >= 1) gf[k][c] = m2_val<double>(S, Gf + (size_t)(s - 1) * R, c, ii[k]);
          }
      lds_barrier();
      ldbl *t = gbA;
      gbA = gbB;
      gbB = t;
    }
  }
#pragma unroll
  for (int k = 0; k < kSPT; k++)
    if (act[k]) hbm(S.rho)[(size_t)S.block_mm[b0] * M + ii[k]] = racc[k];
  __syncthreads();
}

// Markov-1 chain on any lattice, sites strided over the threads, rows in LDS
// (kLds) or HBM.
template <bool kLds>
__device__ __noinline__ void chain_markov1(const DevSys &S0, int mo) {
  typedef typename Rows<kLds>::T RowT;
  CSys &S = uniform_sys(S0);
  const int M = S.M, W = n_work_sites(S);
  const int tid = work_tid(S), nt = S.threads;
  const int b0 = S.mol_block_begin[mo], b1 = S.mol_block_begin[mo + 1];
  const int N = S.mol_N[mo];
  gdbl *Gf = hbm(S.Gf);
  const gdbl *G1 = hbm(S.G1);
  RowT *ra, *rb;
  if constexpr (kLds) {
    ra = lds_row(0);
    rb = lds_row(S.row_stride);
  } else {
    ra = Gf;
    rb = Gf;
  }
  // s = 0: G(.,1) = G1 (Lattice::Initiate)
  {
    const gdbl *g1 = G1 + (size_t)S.block_seg[b0] * M;
    for (int w = tid; w < W; w += nt) {
      int x, y, z;
      const int i = work_site(S, w, x, y, z);
      const double v = g1[i];
      Gf[i] = v;
      if constexpr (kLds) ra[i] = v;
    }
  }
  Rows<kLds>::barrier();
  const RowT *prev = ra;
  RowT *next = rb;
  int s = 1;
  for (int b = b0; b < b1; b++) {
    const gdbl *g1 = G1 + (size_t)S.block_seg[b] * M;
    for (int j = (b == b0 ? 1 : 0); j < S.block_n[b]; j++, s++) {
      gdbl *gs = Gf + (size_t)s * M;
      for (int w = tid; w < W; w += nt) {
        int x, y, z;
        const int i = work_site(S, w, x, y, z);
        const double v = stencil_prop<double>(S, prev, i, x, y, z) * g1[i];
        gs[i] = v;
        if constexpr (kLds) next[i] = v;
      }
      Rows<kLds>::barrier();
      if constexpr (kLds) {
        RowT *t = (RowT *)prev;
        prev = next;
        next = t;
      } else {
        prev = gs;
      }
    }
  }
  {
    const gdbl *gl = Gf + (size_t)(N - 1) * M;
    double acc = 0;
    for (int w = tid; w < W; w += nt) {
      int x, y, z;
      const int i = work_site(S, w, x, y, z);
      if (is_interior(S, x, y, z)) acc += site_weight(S0, i) * gl[i];
    }
    const double gn = wg_sum(acc);
    if (threadIdx.x == 0) S0.sc->GN[mo] = gn;
  }
  RowT *gbA, *gbB;
  if constexpr (kLds) {
    gbA = ra;
    gbB = rb;
  } else {
    gbA = hbm(S.Gb);
    gbB = hbm(S.Gb) + M;
  }
  {
    const int bl = b1 - 1;
    const gdbl *g1 = G1 + (size_t)S.block_seg[bl] * M;
    gdbl *rho = hbm(S.rho) + (size_t)S.block_mm[bl] * M;
    const gdbl *gf = Gf + (size_t)(N - 1) * M;
    for (int w = tid; w < W; w += nt) {
      int x, y, z;
      const int i = work_site(S, w, x, y, z);
      const double v = g1[i];
      gbA[i] = v;
      // a 1-gradient ghost (the only ghosts visited) reads the last row as
      // ComputeGN's remove_bounds left it: zero (src/LGrad1.cpp:213-221)
      rho[i] += (is_interior(S, x, y, z) ? gf[i] : 0.0) * v;
    }
  }
  Rows<kLds>::barrier();
  s = N - 2;
  for (int b = b1 - 1; b >= b0; b--) {
    const gdbl *g1 = G1 + (size_t)S.block_seg[b] * M;
    gdbl *rho = hbm(S.rho) + (size_t)S.block_mm[b] * M;
    for (int j = (b == b1 - 1 ? S.block_n[b] - 2 : S.block_n[b] - 1); j >= 0; j--, s--) {
      const gdbl *gf = Gf + (size_t)s * M;
      for (int w = tid; w < W; w += nt) {
        int x, y, z;
        const int i = work_site(S, w, x, y, z);
        const double v = stencil_prop<double>(S, gbA, i, x, y, z) * g1[i];
        gbB[i] = v;
        // on a ghost, Gf(s) as set_bounds left it when it propagated to s+1
        rho[i] += gf[S.gradients == 1 ? map_x1(S, i) : i] * v;
      }
      Rows<kLds>::barrier();
      RowT *t = gbA;
      gbA = gbB;
      gbB = t;
    }
  }
  __syncthreads();
}

// Ring of a first-order chain on a 2- or 3-gradient lattice (fjc = 1), the
// sites strided over the workgroup as in chain_markov1: the same chain as
// chain1d_ring (a start on every interior site where the merged end segment
// may sit, mol_linear::ComputePhi src/mol_linear.cpp:35-111), with the
// lattice's own stencil.  A site belongs to one thread throughout, so the
// merged segment's density on the start site is added by its owner.
template <bool kLds>
__device__ __noinline__ void chain_ring(const DevSys &S0, int mo) {
  typedef typename Rows<kLds>::T RowT;
  CSys &S = uniform_sys(S0);
  const int M = S.M, W = n_work_sites(S);
  const int tid = work_tid(S), nt = S.threads;
  const int b0 = S.mol_block_begin[mo], b1 = S.mol_block_begin[mo + 1];
  const int N = S.mol_N[mo];
  gdbl *Gf = hbm(S.Gf);
  const gdbl *G1 = hbm(S.G1);
  const gdbl *g1a = G1 + (size_t)S.block_seg[b0] * M;
  gdbl *rho_a = hbm(S.rho) + (size_t)S.block_mm[b0] * M;
  RowT *ra, *rb;
  if constexpr (kLds) {
    ra = lds_row(0);
    rb = lds_row(S.row_stride);
  } else {
    ra = hbm(S.Gb);
    rb = hbm(S.Gb) + M;
  }
  double gnacc = 0.0;
  for (int wx = 0; wx < W; wx++) {
    int sx, sy, sz;
    const int ix = work_site(S, wx, sx, sy, sz);
    if (!is_interior(S, sx, sy, sz) || !(g1a[ix] > 0)) continue;  // the same on every thread
    // row 0 = delta_x G1
    RowT *prev;
    if constexpr (kLds)
      prev = ra;
    else
      prev = Gf;
    for (int w = tid; w < W; w += nt) {
      int x, y, z;
      const int i = work_site(S, w, x, y, z);
      prev[i] = i == ix ? g1a[i] : 0.0;
    }
    Rows<kLds>::barrier();
    RowT *next = rb;
    int s = 1;
    for (int b = b0 + 1; b < b1 - 1; b++) {
      const gdbl *g1 = G1 + (size_t)S.block_seg[b] * M;
      for (int j = 0; j < S.block_n[b]; j++, s++) {
        gdbl *gs = Gf + (size_t)s * M;
        for (int w = tid; w < W; w += nt) {
          int x, y, z;
          const int i = work_site(S, w, x, y, z);
          const double v = stencil_prop<double>(S, prev, i, x, y, z) * g1[i];
          gs[i] = v;
          if constexpr (kLds) next[i] = v;
        }
        Rows<kLds>::barrier();
        if constexpr (kLds) {
          RowT *t = prev;
          prev = next;
          next = t;
        } else {
          prev = gs;
        }
      }
    }
    // the merged end segment, on the start site only
    for (int w = tid; w < W; w += nt) {
      int x, y, z;
      const int i = work_site(S, w, x, y, z);
      if (i != ix) continue;
      const double last = stencil_prop<double>(S, prev, i, x, y, z) * g1a[i];
      gnacc += site_weight(S0, i) * last;
      rho_a[i] += g1a[i] * last;
    }
    Rows<kLds>::barrier();
    // backward from delta_x G1 through the middle blocks
    RowT *gbA = ra, *gbB = rb;
    for (int w = tid; w < W; w += nt) {
      int x, y, z;
      const int i = work_site(S, w, x, y, z);
      gbA[i] = i == ix ? g1a[i] : 0.0;
    }
    Rows<kLds>::barrier();
    s = N - 1;
    for (int b = b1 - 2; b >= b0 + 1; b--) {
      const gdbl *g1 = G1 + (size_t)S.block_seg[b] * M;
      gdbl *rho = hbm(S.rho) + (size_t)S.block_mm[b] * M;
      for (int j = 0; j < S.block_n[b]; j++, s--) {
        const gdbl *gf = Gf + (size_t)s * M;
        for (int w = tid; w < W; w += nt) {
          int x, y, z;
          const int i = work_site(S, w, x, y, z);
          const double v = stencil_prop<double>(S, gbA, i, x, y, z) * g1[i];
          gbB[i] = v;
          rho[i] += gf[i] * v;
        }
        Rows<kLds>::barrier();
        RowT *t = gbA;
        gbA = gbB;
        gbB = t;
      }
    }
  }
  const double gn = wg_sum(gnacc);
  if (threadIdx.x == 0) S0.sc->GN[mo] = gn;
  __syncthreads();
}

// Chains of a charged 2- or 3-gradient system over every site, ghosts
// included, in the reference's array form (Markov 1, fjc = 1): before each
// step the previous row's ghosts take their bounded values in place
// (set_bounds), the step is the reference's passes over the flattened array
// (each pass a shifted copy over the index range the array allows, so the
// ghost sites get the partial sums the reference gives them), and the
// composition adds Gf Gb on every site.  The last forward row's ghosts are
// cleared before GN, as ComputeGN's remove_bounds clears them.  The charge
// and dielectric constant read the total density on ghosts, which is why
// these systems need the ghosts computed the reference's way.  One
// workgroup, rows in HBM.
// LGrad3::propagate (src/LGrad3.cpp:731-743), LG2Planar::propagate
// (src/LG2Planar.cpp:1039-1127, full simple-cubic stencil), LGrad2::propagate
// (src/LGrad2.cpp:1053-1073), Molecule::propagate_forward / backward
// (src/molecule.cpp:2300-2446).
template <typename Sys>
__device__ __forceinline__ double literal_step(const Sys &S, const gdbl *g, int i) {
  const int M = S.M, JX = S.JX;
  if (S.gradients == 3) {
    const int JY = S.JY;
    double t = 0.0;
    if (i >= JX) t += g[i - JX];
    if (i < M - JX) t += g[i + JX];
    if (i >= JY) t += g[i - JY];
    if (i < M - JY) t += g[i + JY];
    if (i >= 1) t += g[i - 1];
    if (i < M - 1) t += g[i + 1];
    return t * (1.0 / 6.0);
  }
  if (S.geometry == 0) {
    const double C1 = 16.0 / 36.0, C2 = 4.0 / 36.0, C3 = 1.0 / 36.0;
    double t = C1 * g[i];
    if (i >= 1) t += C2 * g[i - 1];
    if (i < M - 1) t += C2 * g[i + 1];
    if (i >= JX) t += C2 * g[i - JX];
    if (i < M - JX) t += C2 * g[i + JX];
    if (i >= JX + 1) t += C3 * g[i - JX - 1];
    if (i >= JX) t += C3 * g[i - JX + 1];
    if (i >= 1 && i < M - JX + 1) t += C3 * g[i + JX - 1];
    if (i < M - JX - 1) t += C3 * g[i + JX + 1];
    return t;
  }
  const double lm = hbm(S.lam_m)[i], lp = hbm(S.lam_p)[i];
  double t = (4.0 / 6.0) * g[i];
  if (i >= JX) t += g[i - JX] * lm;
  if (i < M - JX) t += g[i + JX] * lp;
  if (i >= 1) t += (1.0 / 6.0) * g[i - 1];
  if (i < M - 1) t += (1.0 / 6.0) * g[i + 1];
  t = t * 4.0;
  if (i >= JX + 1) t += g[i - JX - 1] * lm;
  if (i >= JX && i < M - 1) t += g[i - JX + 1] * lm;
  if (i >= 1 && i < M - JX) t += g[i + JX - 1] * lp;
  if (i < M - JX - 1) t += g[i + JX + 1] * lp;
  return t * (1.0 / 6.0);
}

// ghosts of a row take their bounded values (set_bounds).  Every source is
// an interior site or a surface ghost that keeps its value, so the copies do
// not depend on each other.
template <typename Sys>
__device__ __forceinline__ void literal_bounds(const Sys &S, gdbl *row, int tid, int nt) {
  for (int i = tid; i < S.M; i += nt) {
    int x, y, z;
    site_coords(S, i, x, y, z);
    if (is_interior(S, x, y, z)) continue;
    const int b = bound_src(S, x, y, z);
    if (b != i) row[i] = row[b];
  }
}

__device__ __noinline__ void chain_literal(const DevSys &S0, int mo) {
  CSys &S = uniform_sys(S0);
  const int M = S.M;
  const int tid = work_tid(S), nt = S.threads;
  const int b0 = S.mol_block_begin[mo], b1 = S.mol_block_begin[mo + 1];
  const int N = S.mol_N[mo];
  gdbl *Gf = hbm(S.Gf);
  const gdbl *G1 = hbm(S.G1);
  {
    const gdbl *g1 = G1 + (size_t)S.block_seg[b0] * M;
    for (int i = tid; i < M; i += nt) Gf[i] = g1[i];
  }
  __syncthreads();
  int s = 1;
  for (int b = b0; b < b1; b++) {
    const gdbl *g1 = G1 + (size_t)S.block_seg[b] * M;
    for (int j = (b == b0 ? 1 : 0); j < S.block_n[b]; j++, s++) {
      gdbl *prev = Gf + (size_t)(s - 1) * M, *gs = Gf + (size_t)s * M;
      literal_bounds(S, prev, tid, nt);
      __syncthreads();
      for (int i = tid; i < M; i += nt) gs[i] = literal_step(S, prev, i) * g1[i];
      __syncthreads();
    }
  }
  {
    gdbl *gl = Gf + (size_t)(N - 1) * M;
    double acc = 0;
    for (int i = tid; i < M; i += nt) {
      int x, y, z;
      site_coords(S, i, x, y, z);
      if (is_interior(S, x, y, z))
        acc += site_weight(S0, i) * gl[i];
      else
        gl[i] = 0.0;  // remove_bounds
    }
    const double gn = wg_sum(acc);
    if (threadIdx.x == 0) S0.sc->GN[mo] = gn;
  }
  gdbl *gbA = hbm(S.Gb), *gbB = hbm(S.Gb) + M;
  {
    const int bl = b1 - 1;
    const gdbl *g1 = G1 + (size_t)S.block_seg[bl] * M;
    gdbl *rho = hbm(S.rho) + (size_t)S.block_mm[bl] * M;
    const gdbl *gf = Gf + (size_t)(N - 1) * M;
    for (int i = tid; i < M; i += nt) {
      const double v = g1[i];
      gbA[i] = v;
      rho[i] += gf[i] * v;
    }
  }
  __syncthreads();
  s = N - 2;
  for (int b = b1 - 1; b >= b0; b--) {
    const gdbl *g1 = G1 + (size_t)S.block_seg[b] * M;
    gdbl *rho = hbm(S.rho) + (size_t)S.block_mm[b] * M;
    for (int j = (b == b1 - 1 ? S.block_n[b] - 2 : S.block_n[b] - 1); j >= 0; j--, s--) {
      const gdbl *gf = Gf + (size_t)s * M;
      literal_bounds(S, gbA, tid, nt);
      __syncthreads();
      for (int i = tid; i < M; i += nt) {
        const double v = literal_step(S, gbA, i) * g1[i];
        gbB[i] = v;
        rho[i] += gf[i] * v;
      }
      __syncthreads();
      gdbl *t = gbA;
      gbA = gbB;
      gbB = t;
    }
  }
  __syncthreads();
}

// 2-gradient Markov-1 chain, register-blocked.  Thread t < nact owns the
// strip of SX sites (x0 .. x0+SX-1, y) with x0 = (t / MY) SX + 1 and
// y = t % MY + 1, and keeps its own previous values in registers.  The rows
// live in LDS in a strip layout: slot(k, t) = k STRIDE + t holds site
// (x0+k, y) of thread t, so the y-1 / y+1 neighbours of a column are the
// adjacent slots (unit stride across the lanes of a wave, constant offsets
// for the unrolled columns) and ghost sites are never stored.  Gf rows in HBM
// use the same layout with the dense column stride nact, so each store or
// load instruction of a wave covers one contiguous span.
//
// The 9-point weights of LGrad2::propagate (src/LGrad2.cpp:1058-1073) and of
// the planar full stencil (src/LG2Planar.cpp:1085-1096) factor into an axial
// and a radial pass: with column sums q(x) = C2 G(x,y-1) + C1 G(x,y) +
// C2 G(x,y+1), the step is G'(x,y) = [C1 q(x) + lambda-(x) q(x-1) +
// lambda+(x) q(x+1)] G1(x,y), C1 = 1 - 2 lambda (4/6 simple cubic, 1/2
// hexagonal), C2 = lambda (1/6, 1/4); the planar stencil has lambda- =
// lambda+ = C2.  The factored form takes 62 instead of 88 operations per
// strip of 8 sites; it equals the reference's sum up to rounding.
//
// T is the propagator type.  With T = float the Boltzmann weights carry a
// gauge factor e^c (c per molecule), which keeps the end-point
// distributions near unity: G'(s) = G(s) e^{c s}.  It cancels in the
// composition up to the constant e^{c(N+1)}, which is divided out in FP64
// together with GN' = GN e^{c N}; the composition sums are FP64.
template <typename T, int SX, int STRIDE>
__device__ __forceinline__ void strip_step(const __attribute__((address_space(3))) T *__restrict__ p, int am0,
                                           int ap0, int sl, int sr, int dm, int dp,
                                           const __attribute__((address_space(3))) T *__restrict__ lamx, T C1,
                                           T C2, const T (&g1)[SX], T (&v)[SX]) {
  // column sums of columns x0-1 (slot sl), x0 .. x0+SX-1 (own), x0+SX (slot sr)
  T qm = C1 * p[sl];
  qm += C2 * p[sl + dm];
  qm += C2 * p[sl + dp];
  T q0 = C1 * v[0];
  q0 += C2 * p[am0];
  q0 += C2 * p[ap0];
#pragma unroll
  for (int k = 0; k < SX; k++) {
    T qp;
    if (k + 1 < SX) {
      qp = C1 * v[k + 1];
      qp += C2 * p[am0 + (k + 1) * STRIDE];
      qp += C2 * p[ap0 + (k + 1) * STRIDE];
    } else {
      qp = C1 * p[sr];
      qp += C2 * p[sr + dm];
      qp += C2 * p[sr + dp];
    }
    T t = C1 * q0;
    t += lamx[2 * k] * qm;
    t += lamx[2 * k + 1] * qp;
    v[k] = t * g1[k];
    qm = q0;
    q0 = qp;
  }
}

// Per-thread geometry of the strip layout: thread t < nact owns the SX sites
// (x0 .. x0+SX-1, y) of strip column t, held in LDS slots k*STRIDE + t.
// Neighbours after set_bounds: y +- 1 as slot offsets inside the strip,
// x0-1 and x0+SX as slots of the strips that own them.  site0 is the site of
// column 0; column k is site0 + k JX.
#define STRIP_GEOM                                                     \
  const int nact = (S.MX / SX) * S.MY;                                 \
  const bool active = (int)threadIdx.x < nact;                         \
  const int xs_ = active ? (int)threadIdx.x / S.MY : 0;                \
  const int y_ = active ? (int)threadIdx.x - xs_ * S.MY + 1 : 1;       \
  const int x0 = xs_ * SX + 1;                                         \
  const int dm = map_y(S, y_ - 1) - y_, dp = map_y(S, y_ + 1) - y_;    \
  const int xl_ = map_x(S, x0 - 1) - 1, xr_ = map_x(S, x0 + SX) - 1;   \
  const int sl = (xl_ % SX) * STRIDE + (xl_ / SX) * S.MY + (y_ - 1);   \
  const int sr = (xr_ % SX) * STRIDE + (xr_ / SX) * S.MY + (y_ - 1);   \
  const int tt = active ? (int)threadIdx.x : 0;                        \
  const int am0 = tt + dm, ap0 = tt + dp;                              \
  const size_t site0 = (size_t)x0 * S.JX + y_;

// Row access of the strip recursion (local names: g1, v, gfv, G1, Gf, M,
// site0, JX, eshift, tt, nact, RG)
#define LOAD_G1(seg)                                                                  \
  {                                                                                   \
    const gdbl *p_ = G1 + (size_t)(seg) * M + site0;                                  \
    _Pragma("unroll") for (int k = 0; k < SX; k++) g1[k] = (T)(p_[k * JX] * eshift); \
  }
#define STORE_ROW(row)                                                            \
  {                                                                               \
    _Pragma("unroll") for (int k = 0; k < SX; k++) (row)[k * STRIDE + tt] = v[k]; \
  }
#define STORE_ROW_FROM(row, a)                                                    \
  {                                                                               \
    _Pragma("unroll") for (int k = 0; k < SX; k++) (row)[k * STRIDE + tt] = a[k]; \
  }
#define STORE_GF(s_)                                                                \
  {                                                                                 \
    GT *q_ = Gf + (size_t)(s_) * RG;                                                \
    _Pragma("unroll") for (int k = 0; k < SX; k++) q_[(size_t)k * nact] = v[k];     \
  }
#define LOAD_GF(s_)                                                                 \
  {                                                                                 \
    const GT *q_ = Gf + (size_t)(s_) * RG;                                          \
    _Pragma("unroll") for (int k = 0; k < SX; k++) gfv[k] = q_[(size_t)k * nact];   \
  }

// Checkpointed backward pass with composition (save_memory, the
// reference's src/molecule.cpp:2375-2410): Gf holds rows 0, K, 2K, ... of
// the forward recursion; the rows of segment q (qK+1 .. qK+K-1) are
// recomputed from checkpoint q into a buffer of K-1 rows (Gf rows nck ..
// nck+K-2) just before the backward pass needs them.  The recomputation
// repeats the forward arithmetic exactly, so the densities are bit-identical
// to full storage.  It borrows gfv as its state, g1 for its weights and both
// LDS rows; afterwards every thread rewrites its backward values (v) into the
// current row.  Expects the LDS weights (s_lam) of the forward pass.
template <typename T, int SX, int STRIDE>
__device__ __noinline__ void strip_backward_ckpt(const DevSys &S0, int mo, double shift) {
  typedef __attribute__((address_space(3))) T LT;
  typedef __attribute__((address_space(1))) T GT;
  CSys &S = uniform_sys(S0);
  const int JX = S.JX;
  STRIP_GEOM;
  const bool sc = S.lattice_type == 0;
  const T C1 = sc ? (T)(4.0 / 6.0) : (T)0.5, C2 = sc ? (T)(1.0 / 6.0) : (T)0.25;
  const int b0 = S.mol_block_begin[mo], b1 = S.mol_block_begin[mo + 1];
  const int N = S.mol_N[mo];
  const size_t M = S.M;
  const size_t RG = (size_t)SX * nact;
  GT *Gf = (GT *)S.Gf + tt;
  const gdbl *G1 = hbm(S.G1);
  LT *ra = (LT *)lds_row(0), *rb = ra + SX * STRIDE;
  const LT *lamx = (LT *)s_lam + 2 * x0;
  const double eshift = exp(shift);
  const double unshift = exp(-shift * (N + 1));
  T g1[SX], v[SX], gfv[SX];
  double racc[SX];
  const int K = S.ckpt;
  LT *ga = ra, *gb = rb;
  const int bl = b1 - 1;
  const int nck = (N + K - 1) / K;
  GT *Bf = Gf + (size_t)nck * RG;
  // row r of the forward recursion: a checkpoint or a recomputed row
#define LOAD_ROWF(r_)                                                                   \
  {                                                                                     \
    const int rr_ = (r_);                                                               \
    const GT *q_ = (rr_ % K == 0) ? Gf + (size_t)(rr_ / K) * RG                         \
                                  : Bf + (size_t)(rr_ % K - 1) * RG;                    \
    _Pragma("unroll") for (int k = 0; k < SX; k++) gfv[k] = q_[(size_t)k * nact];      \
  }
  int b = bl, bstart = N - S.block_n[bl], q = -1;
  if (active) LOAD_G1(S.block_seg[bl]);
#pragma unroll
  for (int k = 0; k < SX; k++) {
    racc[k] = 0.0;
    v[k] = g1[k];
  }
  for (int s2 = N - 1; s2 >= 0; s2--) {
    if (s2 < bstart) {
      if (active) {
        gdbl *rho = hbm(S.rho) + (size_t)S.block_mm[b] * M + site0;
#pragma unroll
        for (int k = 0; k < SX; k++) rho[k * JX] += racc[k] * unshift;
      }
      b--;
      bstart -= S.block_n[b];
      if (active) LOAD_G1(S.block_seg[b]);
#pragma unroll
      for (int k = 0; k < SX; k++) racc[k] = 0.0;
    }
    if (s2 / K != q) {
      q = s2 / K;
      const int a0 = q * K, e0 = (a0 + K < N ? a0 + K : N) - 1;
      if (e0 > a0) {
        // forward state from checkpoint q, weights of the block of row a0+1
        int rb0 = b0, rst = 0;
        while (rst + S.block_n[rb0] <= a0 + 1) {
          rst += S.block_n[rb0];
          rb0++;
        }
        if (active) {
          const GT *c_ = Gf + (size_t)q * RG;
#pragma unroll
          for (int k = 0; k < SX; k++) gfv[k] = c_[(size_t)k * nact];
          LOAD_G1(S.block_seg[rb0]);
        }
        lds_barrier();
        if (active) {
#pragma unroll
          for (int k = 0; k < SX; k++) ra[k * STRIDE + tt] = gfv[k];
        }
        lds_barrier();
        LT *pr = ra, *nx = rb;
        for (int r = a0 + 1; r <= e0; r++) {
          if (r >= rst + S.block_n[rb0]) {
            rst += S.block_n[rb0];
            rb0++;
            if (active) LOAD_G1(S.block_seg[rb0]);
          }
          if (active) {
            strip_step<T, SX, STRIDE>(pr, am0, ap0, sl, sr, dm, dp, lamx, C1, C2, g1, gfv);
            GT *w_ = Bf + (size_t)(r - a0 - 1) * RG;
#pragma unroll
            for (int k = 0; k < SX; k++) {
              w_[(size_t)k * nact] = gfv[k];
              nx[k * STRIDE + tt] = gfv[k];
            }
          }
          lds_barrier();
          LT *tmp = pr;
          pr = nx;
          nx = tmp;
        }
        // back to the backward pass: its current row and weights
        if (active) {
          STORE_ROW(ga);
          LOAD_G1(S.block_seg[b]);
        }
        lds_barrier();
      }
      if (active) LOAD_ROWF(s2);
    }
    if (active) {
      if (s2 == N - 1) {
#pragma unroll
        for (int k = 0; k < SX; k++) v[k] = g1[k];
      } else {
        strip_step<T, SX, STRIDE>(ga, am0, ap0, sl, sr, dm, dp, lamx, C1, C2, g1, v);
      }
#pragma unroll
      for (int k = 0; k < SX; k++) racc[k] += (double)gfv[k] * (double)v[k];
      if (s2 == N - 1) {
        STORE_ROW(ga);
      } else {
        STORE_ROW(gb);
      }
      if (s2 >= 1 && (s2 - 1) / K == q) LOAD_ROWF(s2 - 1);
    }
    lds_barrier();
    if (s2 != N - 1) {
      LT *tmp = ga;
      ga = gb;
      gb = tmp;
    }
  }
  if (active) {
    gdbl *rho = hbm(S.rho) + (size_t)S.block_mm[b] * M + site0;
#pragma unroll
    for (int k = 0; k < SX; k++) rho[k * JX] += racc[k] * unshift;
  }
#undef LOAD_ROWF
}

// Checkpointed backward pass with composition for K = 2: Gf holds the even
// rows of the forward recursion (row 2q at index q).  An odd row 2q+1 is
// recomputed on chip just before the backward step that composes it: its
// base row 2q goes into the free LDS row and one forward step (the same
// arithmetic, bit-identical) turns it into row 2q+1 in registers.  The HBM
// traffic of the end-point rows halves for one extra propagation step per
// two rows; every Gf load is issued one step ahead.  Expects the LDS weights
// (s_lam) of the forward pass.
template <typename T, int SX, int STRIDE>
__device__ __noinline__ void strip_backward_pair(const DevSys &S0, int mo, double shift) {
  typedef __attribute__((address_space(3))) T LT;
  typedef __attribute__((address_space(1))) T GT;
  CSys &S = uniform_sys(S0);
  const int JX = S.JX;
  STRIP_GEOM;
  const bool sc = S.lattice_type == 0;
  const T C1 = sc ? (T)(4.0 / 6.0) : (T)0.5, C2 = sc ? (T)(1.0 / 6.0) : (T)0.25;
  const int b1 = S.mol_block_begin[mo + 1];
  const int N = S.mol_N[mo];
  const size_t M = S.M;
  const size_t RG = (size_t)SX * nact;
  GT *Gf = (GT *)S.Gf + tt;
  const gdbl *G1 = hbm(S.G1);
  LT *ga = (LT *)lds_row(0), *gb = ga + SX * STRIDE;
  const LT *lamx = (LT *)s_lam + 2 * x0;
  const double eshift = exp(shift);
  const double unshift = exp(-shift * (N + 1));
  T g1[SX], v[SX], gfv[SX];
  double racc[SX];
  int b = b1 - 1, bstart = N - S.block_n[b1 - 1];
  if (active) {
    LOAD_G1(S.block_seg[b]);
    LOAD_GF((N - 1) / 2);  // row N-1 itself (even) or the base of its pair (odd)
  }
#pragma unroll
  for (int k = 0; k < SX; k++) racc[k] = 0.0;
  for (int s2 = N - 1; s2 >= 0; s2--) {
    if (s2 < bstart) {
      if (active) {
        gdbl *rho = hbm(S.rho) + (size_t)S.block_mm[b] * M + site0;
#pragma unroll
        for (int k = 0; k < SX; k++) rho[k * JX] += racc[k] * unshift;
      }
      b--;
      bstart -= S.block_n[b];
      if (active) LOAD_G1(S.block_seg[b]);
#pragma unroll
      for (int k = 0; k < SX; k++) racc[k] = 0.0;
    }
    if (s2 & 1) {
      // gfv holds row s2-1: one forward step through the free LDS row
      if (active) STORE_ROW_FROM(gb, gfv);
      lds_barrier();
      if (active) strip_step<T, SX, STRIDE>(gb, am0, ap0, sl, sr, dm, dp, lamx, C1, C2, g1, gfv);
      lds_barrier();
    }
    if (active) {
      if (s2 == N - 1) {
#pragma unroll
        for (int k = 0; k < SX; k++) v[k] = g1[k];
      } else {
        strip_step<T, SX, STRIDE>(ga, am0, ap0, sl, sr, dm, dp, lamx, C1, C2, g1, v);
      }
#pragma unroll
      for (int k = 0; k < SX; k++) racc[k] += (double)gfv[k] * (double)v[k];
      if (s2 == N - 1) {
        STORE_ROW(ga);
      } else {
        STORE_ROW(gb);
      }
      // row s2-1 (even) or the base of the pair ending at s2-1 (odd)
      if (s2 >= 1) LOAD_GF((s2 - 1) / 2);
    }
    lds_barrier();
    if (s2 != N - 1) {
      LT *tmp = ga;
      ga = gb;
      gb = tmp;
    }
  }
  if (active) {
    gdbl *rho = hbm(S.rho) + (size_t)S.block_mm[b] * M + site0;
#pragma unroll
    for (int k = 0; k < SX; k++) rho[k * JX] += racc[k] * unshift;
  }
}

template <typename T, int SX, int STRIDE, bool kCk>
__device__ __noinline__ void chain2d_strip(const DevSys &S0, int mo, double shift) {
  typedef __attribute__((address_space(3))) T LT;
  typedef __attribute__((address_space(1))) T GT;
  CSys &S = uniform_sys(S0);
  const int MX = S.MX, JX = S.JX;
  STRIP_GEOM;
  const bool cyl = S.geometry != 0;
  const bool sc = S.lattice_type == 0;
  const T C1 = sc ? (T)(4.0 / 6.0) : (T)0.5, C2 = sc ? (T)(1.0 / 6.0) : (T)0.25;
  const int b0 = S.mol_block_begin[mo], b1 = S.mol_block_begin[mo + 1];
  const int N = S.mol_N[mo];
  const size_t M = S.M;
  const size_t RG = (size_t)SX * nact;  // Gf row length (interior sites)
  GT *Gf = (GT *)S.Gf + tt;
  const gdbl *G1 = hbm(S.G1);
  LT *ra = (LT *)lds_row(0), *rb = ra + SX * STRIDE;
  LT *lam = (LT *)s_lam;
  for (int x = threadIdx.x; x < MX + 2; x += blockDim.x) {
    lam[2 * x] = cyl ? (T)hbm(S.lam_m)[x * JX + 1] : C2;
    lam[2 * x + 1] = cyl ? (T)hbm(S.lam_p)[x * JX + 1] : C2;
  }
  lds_barrier();
  const LT *lamx = lam + 2 * x0;
  const double eshift = exp(shift);
  T g1[SX], v[SX], gfv[SX];
  double racc[SX];

  // ---- forward (kCk: keep only every K-th row, K = S.ckpt > 1)
  const int K = kCk ? S.ckpt : 1;
  if (active) LOAD_G1(S.block_seg[b0]);
#pragma unroll
  for (int k = 0; k < SX; k++) v[k] = active ? g1[k] : (T)0;
  if (active) {
    STORE_GF(0);
    STORE_ROW(ra);
  }
  lds_barrier();
  LT *prev = ra, *next = rb;
  int s = 1;
  for (int b = b0; b < b1; b++) {
    if (b != b0 && active) LOAD_G1(S.block_seg[b]);
    const int n = S.block_n[b] - (b == b0 ? 1 : 0);
    for (int j = 0; j < n; j++, s++) {
      if (active) {
        strip_step<T, SX, STRIDE>(prev, am0, ap0, sl, sr, dm, dp, lamx, C1, C2, g1, v);
        if constexpr (!kCk) {
          STORE_GF(s);
        } else if (s % K == 0) {
          STORE_GF(s / K);
        }
        STORE_ROW(next);
      }
      lds_barrier();
      LT *tmp = prev;
      prev = next;
      next = tmp;
    }
  }
  // GN: weighted sum of the last distribution (held in v)
  {
    double acc = 0;
    if (active) {
      const gdbl *Lw = hbm(S.L) + site0;
#pragma unroll
      for (int k = 0; k < SX; k++) acc += (cyl ? Lw[k * JX] : 1.0) * (double)v[k];
    }
    const double gn = wg_sum(acc);
    if (threadIdx.x == 0) S0.sc->GN[mo] = gn * exp(-shift * N);
  }
  // ---- backward with composition; the density of the current block in
  // registers, Gf(s-1) fetched one step ahead
  const double unshift = exp(-shift * (N + 1));
  LT *ga = ra, *gb = rb;
  const int bl = b1 - 1;
  if constexpr (kCk) {
    if (S.ckpt == 2)
      strip_backward_pair<T, SX, STRIDE>(S0, mo, shift);
    else
      strip_backward_ckpt<T, SX, STRIDE>(S0, mo, shift);
  } else {
  if (active) {
    LOAD_G1(S.block_seg[bl]);
    LOAD_GF(N - 1);
#pragma unroll
    for (int k = 0; k < SX; k++) {
      v[k] = g1[k];
      racc[k] = (double)gfv[k] * (double)v[k];
    }
    STORE_ROW(ga);
    if (N >= 2) LOAD_GF(N - 2);
  }
  lds_barrier();
  s = N - 2;
  for (int b = bl; b >= b0; b--) {
    if (b != bl) {
      if (active) LOAD_G1(S.block_seg[b]);
#pragma unroll
      for (int k = 0; k < SX; k++) racc[k] = 0.0;
    }
    const int n = S.block_n[b] - (b == bl ? 1 : 0);
    for (int j = 0; j < n; j++, s--) {
      if (active) {
        strip_step<T, SX, STRIDE>(ga, am0, ap0, sl, sr, dm, dp, lamx, C1, C2, g1, v);
#pragma unroll
        for (int k = 0; k < SX; k++) racc[k] += (double)gfv[k] * (double)v[k];
        STORE_ROW(gb);
        if (s >= 1) LOAD_GF(s - 1);
      }
      lds_barrier();
      LT *tmp = ga;
      ga = gb;
      gb = tmp;
    }
    if (active) {
      gdbl *rho = hbm(S.rho) + (size_t)S.block_mm[b] * M + site0;
#pragma unroll
      for (int k = 0; k < SX; k++) rho[k * JX] += racc[k] * unshift;
    }
  }
  }
#undef STRIP_GEOM
#undef LOAD_G1
#undef STORE_ROW
#undef STORE_ROW_FROM
#undef STORE_GF
#undef LOAD_GF
  __syncthreads();
}

// Second-order Markov chain, rows of C M values (C direction components) in
// LDS (kLds) or HBM.  kKind picks the lattice at compile time, so each
// instance holds only its own step and a component count fixed where it
// can be: M2_SC1 1-gradient simple cubic (3 components), M2_HEX1
// 1-gradient hexagonal (FJC components, at most kMaxComp), M2_D2
// 2-gradient (5), M2_D3 3-gradient (6).
enum M2Kind : int { M2_SC1 = 0, M2_HEX1 = 1, M2_D2 = 2, M2_D3 = 3 };
template <bool kLds, int kKind>
__device__ __noinline__ void chain_markov2(const DevSys &S0, int mo) {
  constexpr int kC = kKind == M2_SC1 ? 3 : (kKind == M2_D2 ? 5 : (kKind == M2_D3 ? 6 : kMaxComp));
  typedef typename Rows<kLds>::T RowT;
  CSys &S = uniform_sys(S0);
  const int M = S.M, W = n_work_sites(S);
  const int tid = work_tid(S), nt = S.threads;
  const int b0 = S.mol_block_begin[mo], b1 = S.mol_block_begin[mo + 1];
  const int N = S.mol_N[mo];
  // direction components: 3 on the simple-cubic lattice, FJC = 2 fjc + 1 on
  // the hexagonal one
  const int C = S.mol_size[mo];
  const int R = C * M;
  constexpr bool hex = kKind == M2_HEX1;
  constexpr bool two = kKind == M2_D2;    // 5 components: +x, +y, out of plane, -y, -x
  constexpr bool three = kKind == M2_D3;  // 6 components: +x, +y, +z, -z, -y, -x
  double P[kMaxComp];
#pragma unroll
  for (int k = 0; k < kMaxComp; k++) P[k] = S.mol_P[mo][k];
  // composition weights (LGrad1::AddPhiS src/LGrad1.cpp:876-910): ends and
  // inner components
  const int F = C;
  const double wend = hex ? (S.fjc == 1 ? 1.0 / 4.0 : 0.5 / (F - 1.0)) : 1.0 / 6.0;
  const double wmid = hex ? (S.fjc == 1 ? 2.0 / 4.0 : 1.0 / (F - 1.0)) : 4.0 / 6.0;
  // composition weight of component c (LGrad2::AddPhiS src/LGrad2.cpp:1849-1882:
  // 1/3 for the out-of-plane component, 1/6 for the others)
  auto wcomp = [&](int c) {
    return three ? 1.0 / 6.0 : (two ? (c == 2 ? 1.0 / 3.0 : 1.0 / 6.0) : (c == 0 || c == C - 1 ? wend : wmid));
  };
  gdbl *Gf = hbm(S.Gf);
  const gdbl *G1 = hbm(S.G1);
  RowT *ra, *rb;
  if constexpr (kLds) {
    ra = lds_row(0);
    rb = lds_row(S.row_stride);
  } else {
    ra = Gf;
    rb = Gf;
  }
  {
    const gdbl *g1 = G1 + (size_t)S.block_seg[b0] * M;
    for (int w = tid; w < W; w += nt) {
      int x, y, z;
      const int i = work_site(S, w, x, y, z);
      const double v = g1[i];
      for (int c = 0; c < kC && c < C; c++) {
        Gf[c * M + i] = v;
        if constexpr (kLds) ra[c * M + i] = v;
      }
    }
  }
  Rows<kLds>::barrier();
  const RowT *prev = ra;
  RowT *next = rb;
  int s = 1;
  for (int b = b0; b < b1; b++) {
    const gdbl *g1 = G1 + (size_t)S.block_seg[b] * M;
    for (int j = (b == b0 ? 1 : 0); j < S.block_n[b]; j++, s++) {
      gdbl *gs = Gf + (size_t)s * R;
      for (int w = tid; w < W; w += nt) {
        int x, y, z;
        const int i = work_site(S, w, x, y, z);
        double o[kC];
        if constexpr (three)
          m3_step<double>(S, prev, P, true, x, y, z, o);
        else if constexpr (two)
          m2d_step<double>(S, prev, P, true, i, x, y, o);
        else if constexpr (hex)
          m2h_step<double>(S, prev, P, true, i, o);
        else
          m2_step<double>(S, prev, P, true, i, o);
        const double wg = g1[i];
#pragma unroll
        for (int c = 0; c < kC; c++) {
          if (c >= C) break;
          const double v = o[c] * wg;
          gs[c * M + i] = v;
          if constexpr (kLds) next[c * M + i] = v;
        }
      }
      Rows<kLds>::barrier();
      if constexpr (kLds) {
        RowT *t = (RowT *)prev;
        prev = next;
        next = t;
      } else {
        prev = gs;
      }
    }
  }
  // GN (LGrad1::ComputeGN src/LGrad1.cpp:943-955): simple cubic
  // (G0 + 4 G1 + G2) / 6, hexagonal (G0 + 2 sum G_inner + G_last) / (4 fjc)
  {
    const gdbl *gl = Gf + (size_t)(N - 1) * R;
    double a[kC];
#pragma unroll
    for (int c = 0; c < kC; c++) a[c] = 0;
    for (int w = tid; w < W; w += nt) {
      int x, y, z;
      const int i = work_site(S, w, x, y, z);
      if (is_interior(S, x, y, z)) {
        const double wt = site_weight(S0, i);
#pragma unroll
        for (int c = 0; c < kC; c++)
          if (c < C) a[c] += wt * gl[c * M + i];
      }
    }
    double gn = wg_sum(a[0]);
    if constexpr (three) {
      // LGrad3::ComputeGN (src/LGrad3.cpp:1883-1895): the mean of the components
#pragma unroll
      for (int c = 1; c < 6; c++) gn += wg_sum(a[c]);
      gn = gn / (1.0 * 6);
    } else if constexpr (two) {
      // LGrad2::ComputeGN (src/LGrad2.cpp:1814-1840): (G0 + G1 + 2 G2 + G3 + G4) / 6
      gn += wg_sum(a[1]);
      gn += 2 * wg_sum(a[2]);
      gn += wg_sum(a[3]);
      gn += wg_sum(a[4]);
    } else {
      const double inner = hex ? 2.0 : 4.0;
#pragma unroll
      for (int c = 1; c < kC - 1; c++)
        if (c < C - 1) gn += inner * wg_sum(a[c]);
#pragma unroll
      for (int c = 2; c < kC; c++)
        if (c == C - 1) gn += wg_sum(a[c]);
    }
    if (threadIdx.x == 0) S0.sc->GN[mo] = three ? gn : (hex ? gn / (4.0 * S.fjc) : gn / 6.0);
  }
  RowT *gbA, *gbB;
  if constexpr (kLds) {
    gbA = ra;
    gbB = rb;
  } else {
    gbA = hbm(S.Gb);
    gbB = hbm(S.Gb) + R;
  }
  {
    const int bl = b1 - 1;
    const gdbl *g1 = G1 + (size_t)S.block_seg[bl] * M;
    gdbl *rho = hbm(S.rho) + (size_t)S.block_mm[bl] * M;
    const gdbl *gf = Gf + (size_t)(N - 1) * R;
    for (int w = tid; w < W; w += nt) {
      int x, y, z;
      const int i = work_site(S, w, x, y, z);
      const double v = g1[i];
      double r = rho[i];
      const bool in = is_interior(S, x, y, z);  // 1-gradient ghosts read zero (ComputeGN's remove_bounds)
#pragma unroll
      for (int c = 0; c < kC; c++) {
        if (c >= C) break;
        gbA[c * M + i] = v;
        r += wcomp(c) * (in ? gf[c * M + i] : 0.0) * v;
      }
      rho[i] = r;
    }
  }
  Rows<kLds>::barrier();
  s = N - 2;
  for (int b = b1 - 1; b >= b0; b--) {
    const gdbl *g1 = G1 + (size_t)S.block_seg[b] * M;
    gdbl *rho = hbm(S.rho) + (size_t)S.block_mm[b] * M;
    for (int j = (b == b1 - 1 ? S.block_n[b] - 2 : S.block_n[b] - 1); j >= 0; j--, s--) {
      const gdbl *gf = Gf + (size_t)s * R;
      for (int w = tid; w < W; w += nt) {
        int x, y, z;
        const int i = work_site(S, w, x, y, z);
        double o[kC];
        if constexpr (three)
          m3_step<double>(S, gbA, P, false, x, y, z, o);
        else if constexpr (two)
          m2d_step<double>(S, gbA, P, false, i, x, y, o);
        else if constexpr (hex)
          m2h_step<double>(S, gbA, P, false, i, o);
        else
          m2_step<double>(S, gbA, P, false, i, o);
        const double wg = g1[i];
        double r = rho[i];
#pragma unroll
        for (int c = 0; c < kC; c++) {
          if (c >= C) break;
          const double v = o[c] * wg;
          gbB[c * M + i] = v;
          // 1-gradient ghosts: the components set_bounds(X, Y) paired there
          double f;
          if constexpr (kKind == M2_SC1)
            f = m2_val<double>(S, gf, c, i);
          else if constexpr (kKind == M2_HEX1)
            f = m2h_val<double>(S, gf, c, i);
          else
            f = gf[c * M + i];
          r += wcomp(c) * f * v;
        }
        rho[i] = r;
      }
      Rows<kLds>::barrier();
      RowT *t = gbA;
      gbA = gbB;
      gbB = t;
    }
  }
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ void chain_strip_select(const DevSys &S, int mo, double shift) {
  const int sx = S.strip_sx, st = S.strip_stride;
  if (S.ckpt > 1) {
    if (sx == 8) {
      if (st == 256)
        chain2d_strip<T, 8, 256, true>(S, mo, shift);
      else if (st == 512)
        chain2d_strip<T, 8, 512, true>(S, mo, shift);
      else
        chain2d_strip<T, 8, 1024, true>(S, mo, shift);
    } else {
      if (st == 256)
        chain2d_strip<T, 4, 256, true>(S, mo, shift);
      else if (st == 512)
        chain2d_strip<T, 4, 512, true>(S, mo, shift);
      else
        chain2d_strip<T, 4, 1024, true>(S, mo, shift);
    }
    return;
  }
  if (sx == 8) {
    if (st == 256)
      chain2d_strip<T, 8, 256, false>(S, mo, shift);
    else if (st == 512)
      chain2d_strip<T, 8, 512, false>(S, mo, shift);
    else
      chain2d_strip<T, 8, 1024, false>(S, mo, shift);
  } else {
    if (st == 256)
      chain2d_strip<T, 4, 256, false>(S, mo, shift);
    else if (st == 512)
      chain2d_strip<T, 4, 512, false>(S, mo, shift);
    else
      chain2d_strip<T, 4, 1024, false>(S, mo, shift);
  }
}

__device__ __noinline__ void chain_strip_dispatch(const DevSys &S, int mo, bool fp32, double shift) {
  if (fp32)
    chain_strip_select<float>(S, mo, shift);
  else
    chain_strip_select<double>(S, mo, 0.0);
}

// ------------------------------------------------------ electrostatics
// The potential block with fixed surface potentials (frozen segments with
// e.psi0/kT), as the reference computes it.  Segment::UpdateValence
// (src/segment.cpp:1972-1979) overwrites psi with psi0 on the segment's
// mask and ZERO everywhere else (OverwriteC, src/tools_host.h:514-517), and
// UpdateQ (src/LG1Planar.cpp:226-244, src/LGrad1.cpp:733-767) sets the
// charge on the mask from that psi, g = -q there; UpdatePsi
// (src/LG1Planar.cpp:200-222, src/LGrad1.cpp:687-729) then takes Jacobi
// sweeps of the last segment's overwritten psi off the masks, and g = x -
// the swept psi there.  So the potential the sweep reads is psi0 on the
// last segment's mask and zero elsewhere, not the iterate; this is kept
// for parity.  Each sweep is a separate pass (elA, elB), as every one of
// the reference's loops reads the values of the loop before it.
template <typename Sys>
__device__ __forceinline__ double fixed_psi_at(const Sys &S, int k, int x) {
  return hbm(S.seg_mask)[k * S.M + el_mirror(S, x)] == 1.0 ? S.seg_psi0[k] : 0.0;
}

__device__ __noinline__ double electro_fixed(const DevSys &S0, double acc) {
  CSys &S = uniform_sys(S0);
  const int M = S.M, f = S.fjc, MX = S.MX;
  const int tid = work_tid(S), nt = S.threads;
  const bool max_g = S.max_g, cyl = S.geometry == 1;
  gdbl *gp = hbm(S.g) + S.nitm * M;
  const gdbl *xp = hbm(S.x) + S.nitm * M;
  gdbl *J1 = hbm(S.elA), *J2 = cyl ? hbm(S.elB) : hbm(S.elA);
  gdbl *q = hbm(S.q), *P = hbm(S.psi);
  const gdbl *E = hbm(S.eps);
  const int L = S.el_last_fixed;
  const double C = S.el_C0, Cq = -S.el_C0, ff = (double)f * f;
  auto on_mask = [&](int x) {
    bool m = false;
    for (int k = 0; k < S.nseg; k++)
      if (S.seg_fixpsi[k] && hbm(S.seg_mask)[k * M + x] != 0.0) m = true;
    return m;
  };
  for (int x = tid; x < M; x += nt) {
    if (x < f || x >= MX + f) {  // ghosts keep the last overwrite, mirrored (psi[0] = psi[1])
      J1[x] = fixed_psi_at(S, L, x == 0 ? 1 : x);
      continue;
    }
    for (int k = 0; k < S.nseg; k++) {
      if (!S.seg_fixpsi[k] || hbm(S.seg_mask)[k * M + x] != 1.0) continue;
      const double a = fixed_psi_at(S, k, x - 1), b = fixed_psi_at(S, k, x), c = fixed_psi_at(S, k, x + 1);
      double qx = -0.5 * (a - 2 * b + c) * ff * E[x] / Cq;
      if (cyl) qx -= (c - a) / (2.0 * (S.el_r0 + x - f + 0.5)) * f * E[x] / Cq;
      if (S.grad_eps) qx -= 0.25 * (E[x + 1] - E[x - 1]) * (c - a) * ff / Cq;
      q[x] = qx;
      gp[x] = -qx;
    }
    const double pL = fixed_psi_at(S, L, x);
    J1[x] = on_mask(x) ? pL : 0.5 * (fixed_psi_at(S, L, x - 1) + fixed_psi_at(S, L, x + 1)) + q[x] * C / E[x];
  }
  __syncthreads();
  if (cyl) {
    for (int x = tid; x < M; x += nt) {
      const bool sweep = x >= f && x < MX + f && !on_mask(x);
      J2[x] = sweep ? J1[x] + (J1[x + 1] - J1[x - 1]) / (2.0 * (S.el_r0 + x - f + 0.5)) * f : J1[x];
    }
    __syncthreads();
  }
  for (int x = tid; x < M; x += nt) {
    const bool in = x >= f && x < MX + f;
    const bool sweep = in && !on_mask(x);
    double v = J2[x];
    if (sweep && S.grad_eps) v = v + 0.25 * (E[x + 1] - E[x - 1]) * (J2[x + 1] - J2[x - 1]) / E[x] * ff;
    P[x] = v;
    if (sweep) gp[x] = xp[x] - v;
    if (in) {
      const double gv = gp[x];
      acc = max_g ? fmax(acc, fabs(gv)) : acc + gv * gv;
    }
  }
  return acc;
}

// The potential sweep of a charged 2- or 3-gradient system, after the
// charge, dielectric constant, psi and EE of every site are stored.
// * 2-gradient cylindrical (LGrad2::UpdatePsi src/LGrad2.cpp:1428-1445):
//   Jacobi, one site per thread;
// * 2-gradient planar (LG2Planar::UpdatePsi src/LG2Planar.cpp:1163-1181):
//   upwind in x and in y, X[x,y] reads X[x-1,y] and X[x,y-1] of the same
//   sweep, so it runs as a wavefront over the anti-diagonals (one barrier
//   each), which gives every site the values the reference's row-by-row loop
//   gives it;
// * 3-gradient (LGrad3::UpdatePsi src/LGrad3.cpp:1150-1224): a symmetric
//   Gauss-Seidel sweep in place, forward in lexicographic order and then
//   backward, g -= psi in the backward one.  Each sweep runs as a wavefront
//   over the planes x + y + z = k: a site's lower neighbours lie on plane
//   k - 1 (already updated, as in the lexicographic loop) and its upper ones
//   on plane k + 1 (not yet), so the result is the reference's.
// g = x - X (2 gradients) or x - the swept psi (3 gradients) on the interior.
__device__ __noinline__ double electro_residual_xyz(const DevSys &S0, double acc) {
  CSys &S = uniform_sys(S0);
  const int M = S.M, MX = S.MX, MY = S.MY, MZ = S.MZ, JX = S.JX, JY = S.JY;
  const int tid = work_tid(S), nt = S.threads;
  const bool max_g = S.max_g;
  gdbl *gp = hbm(S.g) + S.nitm * M;
  const gdbl *xp = hbm(S.x) + S.nitm * M;
  const gdbl *E = hbm(S.eps), *q = hbm(S.q);
  gdbl *P = hbm(S.psi);
  const double C = S.el_C;
  auto add = [&](int i, double X) {
    const double gv = xp[i] - X;
    gp[i] = gv;
    acc = max_g ? fmax(acc, fabs(gv)) : acc + gv * gv;
  };
  if (S.gradients == 2 && S.geometry != 0) {
    for (int w = tid; w < MX * MY; w += nt) {
      const int x = 1 + w / MY, y = 1 + w % MY, i = x * JX + y;
      const double r = S.el_r0 + x;  // offset_first_layer fjc + (x - fjc + 1), fjc = 1
      const double Li = hbm(S.L)[i];
      const double epsXmin = 2 * 3.14159265 * (r - 1) * (E[i] + E[i - JX]) / Li;
      const double epsXplus = 2 * 3.14159265 * r * (E[i] + E[i + JX]) / Li;
      const double epsYmin = E[i] + E[i - 1], epsYplus = E[i] + E[i + 1];
      const double X = (C * q[i] + epsXmin * P[i - JX] + epsXplus * P[i + JX] + epsYmin * P[i - 1] +
                        epsYplus * P[i + 1]) /
                       (epsXmin + epsXplus + epsYmin + epsYplus);
      add(i, X);
    }
    return acc;
  }
  if (S.gradients == 2) {
    gdbl *X = hbm(S.elA);
    for (int d = 0; d <= MX + MY - 2; d++) {
      const int x0 = d - (MY - 1) > 0 ? d - (MY - 1) : 0;  // x - 1 on the diagonal x + y - 2 = d
      const int x1 = d < MX - 1 ? d : MX - 1;
      for (int t = tid; t <= x1 - x0; t += nt) {
        const int x = 1 + x0 + t, y = 1 + d - (x0 + t), i = x * JX + y;
        const double epsXmin = E[i] + E[i - JX], epsXplus = E[i] + E[i + JX];
        const double epsYmin = E[i] + E[i - 1], epsYplus = E[i] + E[i + 1];
        const double ax = x == 1 ? P[i - JX] : X[i - JX];
        const double ay = y == 1 ? P[i - 1] : X[i - 1];
        const double v = (C * q[i] + epsXmin * ax + epsXplus * P[i + JX] + epsYmin * ay + epsYplus * P[i + 1]) /
                         (epsXmin + epsXplus + epsYmin + epsYplus);
        X[i] = v;
        add(i, v);
      }
      __syncthreads();
    }
    return acc;
  }
  // 3 gradients: symmetric Gauss-Seidel by planes
  for (int pass = 0; pass < 2; pass++) {
    for (int kk = 0; kk <= MX + MY + MZ - 3; kk++) {
      const int k = pass == 0 ? 3 + kk : MX + MY + MZ - kk;
      for (int w = tid; w < MX * MY; w += nt) {
        const int x = 1 + w / MY, y = 1 + w % MY, z = k - x - y;
        if (z < 1 || z > MZ) continue;
        const int i = x * JX + y * JY + z;
        const double epsZmin = E[i] + E[i - 1], epsZplus = E[i] + E[i + 1];
        const double epsYmin = E[i] + E[i - JY], epsYplus = E[i] + E[i + JY];
        const double epsXmin = E[i] + E[i - JX], epsXplus = E[i] + E[i + JX];
        const double v = (epsXmin * P[i - JX] + epsXplus * P[i + JX] + epsYmin * P[i - JY] + epsYplus * P[i + JY] +
                          epsZmin * P[i - 1] + epsZplus * P[i + 1] + C * q[i]) /
                         (epsXmin + epsXplus + epsYmin + epsYplus + epsZmin + epsZplus);
        P[i] = v;
        if (pass == 1) add(i, v);
      }
      __syncthreads();
    }
  }
  return acc;
}

// Fixed potentials on 2 and 3 gradients, as the reference computes them
// (the same overwrite as on 1 gradient: psi0 on a plate's mask, zero
// elsewhere, then set_M_bounds).
// * 2 gradients: UpdateQ (src/LGrad2.cpp:1447-1484, src/LG2Planar.cpp:
//   1211-1236) sets q on each plate's mask from its overwritten psi, g = -q
//   there (the radial term of LGrad2 reads eps at the flat index x, kept);
//   UpdatePsi (src/LGrad2.cpp:1448-1482, src/LG2Planar.cpp:1182-1206) takes
//   one Jacobi sweep of the last plate's overwritten psi off the masks.
// * 3 gradients: UpdateQ (src/LGrad3.cpp:1314-1347) writes, on the mask and
//   one layer short of each upper face, the discrete Laplacian over C into
//   psi in place (lexicographic order: run as a plane wavefront), and g = -q
//   with the q of DoElectrostatics; UpdatePsi (src/LGrad3.cpp:1239-1282) is
//   the symmetric Gauss-Seidel sweep off the masks.
// The working psi is S.psi; elA holds the Jacobi values.
template <typename Sys>
__device__ __forceinline__ double fixed_psi_xyz(const Sys &S, int k, int x, int y, int z) {
  const gdbl *mk = hbm(S.seg_mask) + (size_t)k * S.M;
  auto mir = [](int c, int n) { return c < 1 ? 1 : (c > n ? n : c); };
  int i;
  if (S.gradients == 2) {
    i = mir(x, S.MX) * S.JX + mir(y, S.MY);
  } else {
    const int nghost = (x < 1 || x > S.MX) + (y < 1 || y > S.MY) + (z < 1 || z > S.MZ);
    i = nghost >= 2 ? x * S.JX + y * S.JY + z : mir(x, S.MX) * S.JX + mir(y, S.MY) * S.JY + mir(z, S.MZ);
  }
  return mk[i] == 1.0 ? S.seg_psi0[k] : 0.0;
}

__device__ __noinline__ double electro_fixed_xyz(const DevSys &S0, double acc) {
  CSys &S = uniform_sys(S0);
  const int M = S.M, MX = S.MX, MY = S.MY, MZ = S.MZ, JX = S.JX, JY = S.JY;
  const int tid = work_tid(S), nt = S.threads;
  const bool max_g = S.max_g, cyl = S.geometry != 0;
  gdbl *gp = hbm(S.g) + S.nitm * M;
  const gdbl *xp = hbm(S.x) + S.nitm * M;
  const gdbl *E = hbm(S.eps);
  gdbl *q = hbm(S.q), *P = hbm(S.psi), *X = hbm(S.elA);
  const int L = S.el_last_fixed;
  const double C = S.el_C0, Cq = -S.el_C0;
  auto on_mask = [&](int i) {
    bool m = false;
    for (int k = 0; k < S.nseg; k++)
      if (S.seg_fixpsi[k] && hbm(S.seg_mask)[(size_t)k * M + i] != 0.0) m = true;
    return m;
  };
  // the last plate's overwritten, bounded psi; on 2 gradients each plate's
  // charge from its own overwrite
  for (int i = tid; i < M; i += nt) {
    int x, y, z;
    site_coords(S, i == 0 ? 1 : i, x, y, z);
    P[i] = fixed_psi_xyz(S, L, x, y, z);
    site_coords(S, i, x, y, z);
    if (!is_interior(S, x, y, z)) continue;
    gp[i] = xp[i];  // g starts as the potential block of x (DoElectrostatics)
    if (S.gradients != 2) continue;
    for (int k = 0; k < S.nseg; k++) {
      if (!S.seg_fixpsi[k] || hbm(S.seg_mask)[(size_t)k * M + i] != 1.0) continue;
      const double pm = fixed_psi_xyz(S, k, x - 1, y, 0), pp = fixed_psi_xyz(S, k, x + 1, y, 0);
      const double ym = fixed_psi_xyz(S, k, x, y - 1, 0), yp = fixed_psi_xyz(S, k, x, y + 1, 0);
      const double p0 = fixed_psi_xyz(S, k, x, y, 0);
      double qx = -0.5 * (pm + pp + ym + yp - 4 * p0) * E[i] / Cq;
      if (cyl) qx -= (pp - pm) / (2.0 * (S.el_r0 + x - 1 + 0.5)) * E[x] / Cq;
      if (S.grad_eps) qx -= 0.25 * (E[i + JX] - E[i - JX]) * (pp - pm + E[i + 1] - E[i - 1]) * (yp - ym) / Cq;
      q[i] = qx;
      gp[i] = -qx;
    }
  }
  __syncthreads();
  if (S.gradients == 2) {
    // one Jacobi sweep off the masks, from the last plate's overwrite
    for (int w = tid; w < MX * MY; w += nt) {
      const int x = 1 + w / MY, y = 1 + w % MY, i = x * JX + y;
      if (on_mask(i)) continue;
      double v = 0.25 * (P[i - JX] + P[i + JX] + P[i - 1] + P[i + 1]) + 0.5 * q[i] * C / E[i];
      if (cyl) v += (P[i + JX] - P[i - JX]) / (2.0 * (S.el_r0 + x - 1 + 0.5));
      if (S.grad_eps)
        v += 0.25 * (E[i + JX] - E[i - JX]) * (P[i + JX] - P[i - JX] + E[i + 1] - E[i - 1]) * (P[i + 1] - P[i - 1]) / E[i];
      X[i] = v;
    }
    __syncthreads();
    for (int w = tid; w < MX * MY; w += nt) {
      const int x = 1 + w / MY, y = 1 + w % MY, i = x * JX + y;
      if (!on_mask(i)) {
        P[i] = X[i];
        gp[i] = xp[i] - X[i];
      }
      const double gv = gp[i];
      acc = max_g ? fmax(acc, fabs(gv)) : acc + gv * gv;
    }
    return acc;
  }
  // 3 gradients: UpdateQ in place on the last plate's mask (the earlier
  // plates' psi is overwritten by the later ones), g = -q on every plate's
  // mask inside its loop range
  for (int w = tid; w < MX * MY * MZ; w += nt) {
    const int x = 1 + w / (MY * MZ), r = w % (MY * MZ), y = 1 + r / MZ, z = 1 + r % MZ;
    if (x > MX - 1 || y > MY - 1 || z > MZ - 1) continue;
    const int i = x * JX + y * JY + z;
    for (int k = 0; k < S.nseg; k++)
      if (S.seg_fixpsi[k] && hbm(S.seg_mask)[(size_t)k * M + i] == 1.0) gp[i] = -q[i];
  }
  for (int kk = 0; kk <= MX + MY + MZ - 6; kk++) {
    const int k = 3 + kk;
    for (int w = tid; w < (MX - 1) * (MY - 1); w += nt) {
      const int x = 1 + w / (MY - 1), y = 1 + w % (MY - 1), z = k - x - y;
      if (z < 1 || z > MZ - 1) continue;
      const int i = x * JX + y * JY + z;
      if (hbm(S.seg_mask)[(size_t)L * M + i] != 1.0) continue;
      const double epsZmin = E[i] + E[i - 1], epsZplus = E[i] + E[i + 1];
      const double epsYmin = E[i] + E[i - JY], epsYplus = E[i] + E[i + JY];
      const double epsXmin = E[i] + E[i - JX], epsXplus = E[i] + E[i + JX];
      P[i] = (epsXmin * P[i - JX] + epsXplus * P[i + JX] + epsYmin * P[i - JY] + epsYplus * P[i + JY] +
              epsZmin * P[i - 1] + epsZplus * P[i + 1] -
              (epsXmin + epsXplus + epsYmin + epsYplus + epsZmin + epsZplus) * P[i]) /
             Cq;
    }
    __syncthreads();
  }
  // set_M_bounds again: the faces take the (changed) interior values
  for (int i = tid; i < M; i += nt) {
    int x, y, z;
    site_coords(S, i, x, y, z);
    const int nghost = (x < 1 || x > MX) + (y < 1 || y > MY) + (z < 1 || z > MZ);
    if (nghost != 1) continue;
    const int xs = x < 1 ? 1 : (x > MX ? MX : x), ys = y < 1 ? 1 : (y > MY ? MY : y),
              zs = z < 1 ? 1 : (z > MZ ? MZ : z);
    P[i] = P[xs * JX + ys * JY + zs];
  }
  __syncthreads();
  for (int pass = 0; pass < 2; pass++) {
    for (int kk = 0; kk <= MX + MY + MZ - 3; kk++) {
      const int k = pass == 0 ? 3 + kk : MX + MY + MZ - kk;
      for (int w = tid; w < MX * MY; w += nt) {
        const int x = 1 + w / MY, y = 1 + w % MY, z = k - x - y;
        if (z < 1 || z > MZ) continue;
        const int i = x * JX + y * JY + z;
        if (on_mask(i)) continue;
        const double epsZmin = E[i] + E[i - 1], epsZplus = E[i] + E[i + 1];
        const double epsYmin = E[i] + E[i - JY], epsYplus = E[i] + E[i + JY];
        const double epsXmin = E[i] + E[i - JX], epsXplus = E[i] + E[i + JX];
        const double v = (epsXmin * P[i - JX] + epsXplus * P[i + JX] + epsYmin * P[i - JY] + epsYplus * P[i + JY] +
                          epsZmin * P[i - 1] + epsZplus * P[i + 1] + C * q[i]) /
                         (epsXmin + epsXplus + epsYmin + epsYplus + epsZmin + epsZplus);
        P[i] = v;
        if (pass == 1) gp[i] = xp[i] - v;
      }
      __syncthreads();
    }
  }
  for (int w = tid; w < MX * MY * MZ; w += nt) {
    const int x = 1 + w / (MY * MZ), r = w % (MY * MZ), y = 1 + r / MZ, z = 1 + r % MZ;
    const double gv = gp[x * JX + y * JY + z];
    acc = max_g ? fmax(acc, fabs(gv)) : acc + gv * gv;
  }
  return acc;
}

// The potential block of the residual of a charged 1-gradient system
// (System::Classical_residual src/system.cpp:2161-2169): charge density,
// dielectric constant, psi and EE at every site, then one sweep of the
// Poisson equation X and g = psi - X on the interior (zero on ghosts).
// The upwind sweep of planar and cylindrical lattices is a first-order
// linear recurrence; the first wave runs it as a scan of affine maps (each
// lane composes a run of consecutive sites, the lanes combine in six
// shuffle steps, then each lane replays its run from its entry value).
// acc is the thread's running sum of g^2 (or max |g|); returns it updated.
__device__ __noinline__ double electro_residual(const DevSys &S0, double acc) {
  CSys &S = uniform_sys(S0);
  const int M = S.M, f = S.fjc, MX = S.MX;
  const int tid = work_tid(S), nt = S.threads;
  const bool max_g = S.max_g;
  gdbl *gp = hbm(S.g) + S.nitm * M;
  const gdbl *xp = hbm(S.x) + S.nitm * M;
  for (int i = tid; i < M; i += nt) {
    site_electro(S, i);
    if (S.gradients == 1) {
      if (i < f || i >= MX + f) gp[i] = 0.0;
    } else {
      int x, y, z;
      site_coords(S, i, x, y, z);
      if (!is_interior(S, x, y, z)) gp[i] = 0.0;
    }
  }
  __syncthreads();
  if (S.gradients >= 2) return S.fixed_psi0 ? electro_fixed_xyz(S0, acc) : electro_residual_xyz(S0, acc);
  if (S.fixed_psi0) return electro_fixed(S0, acc);
  const bool upwind = S.geometry != 2;
  for (int x = f + tid; x < MX + f; x += nt) {
    const double X = site_psi_sweep(S, x);
    if (!upwind) {
      const double gv = xp[x] - X;
      gp[x] = gv;
      acc = max_g ? fmax(acc, fabs(gv)) : acc + gv * gv;
    }
  }
  if (!upwind) return acc;
  __syncthreads();
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    const int chunk = (MX + 63) / 64;
    const int lo = f + lane * chunk, hi = min(MX + f, lo + chunk);
    const gdbl *A = hbm(S.elA), *B = hbm(S.elB);
    double a = 1.0, b = 0.0;  // this lane's run as X_out = a X_in + b
    for (int x = lo; x < hi; x++) {
      const double Ax = A[x];
      b = Ax * b + B[x];
      a = Ax * a;
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const double ap = __shfl_up(a, off), bp = __shfl_up(b, off);
      if (lane >= off) {
        b = a * bp + b;
        a = a * ap;
      }
    }
    const double ain = __shfl_up(a, 1), bin = __shfl_up(b, 1);
    const double X0 = el_psi(S, f - 1);
    double X = lane == 0 ? X0 : ain * X0 + bin;
    for (int x = lo; x < hi; x++) {
      X = A[x] * X + B[x];
      const double gv = xp[x] - X;
      gp[x] = gv;
      acc = max_g ? fmax(acc, fabs(gv)) : acc + gv * gv;
    }
  }
  return acc;
}

// ------------------------------------------------------------- residual
// One Classical_residual: fields x -> g, densities, norms.  Returns ||g||
// (or max|g|) to every thread, or a negative value on a status error.
__device__ double fused_residual(const DevSys &S0) {
  CSys &S = uniform_sys(S0);
  const int M = S.M, W = n_work_sites(S);
  const int tid = work_tid(S), nt = S.threads;
  DevScalars &sc = *S0.sc;
  const bool fp32 = S.precision != 0 && sc.fp32 != 0;
  const unsigned long long t0 = wall_clock64();
  gdbl *rho = hbm(S.rho);
  const int nmm = S.nmm;
  for (int w = tid; w < W; w += nt) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    site_prepare(S, i, x, y, z);
    for (int mm = 0; mm < nmm; mm++) rho[mm * M + i] = 0.0;
  }
  __syncthreads();
  const unsigned long long t1 = wall_clock64();
  for (int mo = 0; mo < S.nmol; mo++) {
    if (S.mol_markov[mo] == 2) {
      const int kind = S.gradients == 3 ? M2_D3 : (S.gradients == 2 ? M2_D2 : (S.lattice_type != 0 ? M2_HEX1 : M2_SC1));
      if (S.use_lds) {
        if (kind == M2_SC1 && S.M <= S.threads)
          chain1d_m2<1>(S0, mo);
        else if (kind == M2_SC1)
          chain_markov2<true, M2_SC1>(S0, mo);
        else if (kind == M2_HEX1)
          chain_markov2<true, M2_HEX1>(S0, mo);
        else if (kind == M2_D2)
          chain_markov2<true, M2_D2>(S0, mo);
        else
          chain_markov2<true, M2_D3>(S0, mo);
      } else {
        if (kind == M2_SC1)
          chain_markov2<false, M2_SC1>(S0, mo);
        else if (kind == M2_HEX1)
          chain_markov2<false, M2_HEX1>(S0, mo);
        else if (kind == M2_D2)
          chain_markov2<false, M2_D2>(S0, mo);
        else
          chain_markov2<false, M2_D3>(S0, mo);
      }
    } else if (S.mol_ring[mo]) {
      if (S.gradients == 1)
        chain1d_ring(S0, mo);
      else if (S.use_lds)
        chain_ring<true>(S0, mo);
      else
        chain_ring<false>(S0, mo);
    } else if (S.all_sites) {
      chain_literal(S0, mo);
    } else if (S.strip_sx > 0) {
      chain_strip_dispatch(S0, mo, fp32, fp32 ? sc.lnshift[mo] : 0.0);
    } else if (S.gradients == 1 && S.fjc == 1 && S.use_lds && S.M <= S.threads) {
      chain1d_m1(S0, mo);
    } else if (S.use_lds) {
      chain_markov1<true>(S0, mo);
    } else {
      chain_markov1<false>(S0, mo);
    }
  }
  const unsigned long long t2 = wall_clock64();
  if (threadIdx.x == 0) {
    for (int i = 0; i < S0.nmol; i++) {
      if (S0.mol_freedom[i] == 0) sc.phibulk[i] = S0.mol_phibulk_in[i];
      if (S0.mol_freedom[i] == 1) sc.n[i] = S0.mol_n_in[i];
    }
    scalars_norms(S0);
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
  const bool max_g = S.max_g;
  const gdbl *g = hbm(S.g);
  for (int w = tid; w < W; w += nt) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    if (!is_interior(S, x, y, z)) continue;
    if (max_g) {
      const double g2 = site_residual(S, i, x, y, z);
      (void)g2;
      for (int it = 0; it < S.nitm; it++) acc = fmax(acc, fabs(g[it * M + i]));
    } else {
      acc += site_residual(S, i, x, y, z);
    }
  }
  if (S.charged) acc = electro_residual(S0, acc);
  const double r = max_g ? wg_max(acc) : sqrt(wg_sum(acc));
  if (threadIdx.x == 0) {
    sc.residual = r;
    sc.t_phase[0] += t1 - t0;
    sc.t_phase[1] += t2 - t1;
    sc.t_phase[2] += wall_clock64() - t2;
  }
  __syncthreads();
  return r;
}

// Thermodynamic summary at the state of the last residual.
__device__ __noinline__ void fused_thermo(const DevSys &S0) {
  CSys &S = uniform_sys(S0);
  const int W = n_work_sites(S);
  const int tid = work_tid(S), nt = S.threads;
  DevScalars &sc = *S0.sc;
  const int nmol = S.nmol;
  const gdbl *mphi = hbm(S.mol_phitot);
  double logc[kMaxMol];
#pragma unroll
  for (int mo = 0; mo < kMaxMol; mo++) {
    logc[mo] = 0.0;
    if (mo < nmol) {
      const double N = S.mol_NF[mo];  // a tagged segment does not count
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
      if (mo < nmol) th[mo] += wt * mphi[mo * S.M + i];
  }
  f = wg_sum(f);
  gp = wg_sum(gp);
#pragma unroll
  for (int mo = 0; mo < kMaxMol; mo++)
    if (mo < nmol) th[mo] = wg_sum(th[mo]);
  if (threadIdx.x == 0) {
    sc.F = f;
    sc.GP = gp;
#pragma unroll
    for (int mo = 0; mo < kMaxMol; mo++)
      if (mo < nmol) sc.theta[mo] = th[mo];
    scalars_mu(S0);
  }
  __syncthreads();
}

// ------------------------------------------------------------- DIIS solve

// SFNewton::iterate_DIIS (src/sfnewton.cpp:1037-1130).  The residual never
// changes x, so the iteration's x0 is x itself at the update: one pass per
// iteration reads x and g, writes x, xR_k and e_k = x_new - x0, and
// accumulates all kd Gram dots <e_i, e_k> (each thread sums its own entries
// in index order, then the waves' sums are added in wave order, as for a
// single dot); a second pass forms x = sum_i c_i xR_i.
__shared__ double s_dred[kMaxDIIS * 16];  // per-wave partial dots
__shared__ double s_coef[kMaxDIIS];        // DIIS coefficients of the combine

// Mixed precision: after an evaluation with residual res, decide whether the
// next ones use FP32 propagators (while res is above the switch level) and
// set each molecule's gauge shift c = -ln(GN / V) / N from the latest GN, so
// that the shifted end-point distributions average about one per site.
__device__ __noinline__ void mixed_update(const DevSys &S, double res, bool first) {
  if (threadIdx.x == 0) {
    DevScalars &sc = *S.sc;
    // FP32 from the first evaluation on while the residual is above the
    // switch level and still falling (best value improved within the last
    // kMixedPatience evaluations); never again once the solve has gone over
    // to FP64.  At the switch the DIIS history restarts, so the FP64
    // iterations combine FP64 residuals only.
    if (first) {
      sc.mixed_best = res;
      sc.mixed_since = 0;
    } else if (res < sc.mixed_best) {
      sc.mixed_best = res;
      sc.mixed_since = 0;
    } else {
      sc.mixed_since++;
    }
    bool ok = res > S.mixed_switch && (first || sc.fp32 != 0) && sc.mixed_since < kMixedPatience;
    for (int mo = 0; mo < S.nmol; mo++) {
      const double gn = sc.GN[mo];
      if (!(gn > 0) || !isfinite(gn)) ok = false;
      else sc.lnshift[mo] = -log(gn / S.volume) / S.mol_N[mo];
    }
    if (!ok && sc.fp32) sc.mixed_restart = 1;
    sc.fp32 = ok ? 1 : 0;
  }
  __syncthreads();
}

// One DIIS update (iteration `it`, ring slot k, kd vectors in use): x -= dmax
// g, xR_k = x, e_k = x_new - x_old, and the kd Gram dots <e_i, e_k> into
// dots[0..kd-1] (each thread sums its entries in index order, the waves'
// sums are added in wave order).
__device__ __noinline__ void diis_update_dots(const DevSys &S0, int k, int kd, double *dots) {
  CSys &S = uniform_sys(S0);
  k = __builtin_amdgcn_readfirstlane(k);
  kd = __builtin_amdgcn_readfirstlane(kd);
  const int nvar = S.nit * S.M, m = S.m;
  const int tid = work_tid(S), nt = S.threads;
  const double dmax = S.deltamax;
  gdbl *x = hbm(S.x), *xR = hbm(S.xR), *xx0 = hbm(S.xx0);
  const gdbl *g = hbm(S.g);
  ldbl *dred = (ldbl *)s_dred;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  gdbl *xRk = xR + (size_t)k * nvar, *xxk = xx0 + (size_t)k * nvar;
  // e vectors of the earlier kd-1 iterations, oldest first
  const gdbl *ei[kMaxDIIS - 1];
#pragma unroll
  for (int i = 0; i < kMaxDIIS - 1; i++) {
    int posi = k - kd + 1 + i;
    if (posi < 0) posi += m;
    ei[i] = xx0 + (size_t)(i < kd - 1 ? posi : 0) * nvar;
  }
  double acc[kMaxDIIS - 1], self = 0.0;
#pragma unroll
  for (int i = 0; i < kMaxDIIS - 1; i++) acc[i] = 0.0;
  for (int v = tid; v < nvar; v += nt) {
    double a[kMaxDIIS - 1];
#pragma unroll
    for (int i = 0; i < kMaxDIIS - 1; i++) a[i] = i < kd - 1 ? ei[i][v] : 0.0;
    const double xo = x[v];
    const double xv = xo + (-dmax) * g[v];
    const double e = xv - xo;
    x[v] = xv;
    xRk[v] = xv;
    xxk[v] = e;
#pragma unroll
    for (int i = 0; i < kMaxDIIS - 1; i++)
      if (i < kd - 1) acc[i] += a[i] * e;
    self += e * e;
  }
#pragma unroll
  for (int i = 0; i < kMaxDIIS; i++) {
    if (i < kd) {
      double t = i < kd - 1 ? acc[i < kMaxDIIS - 1 ? i : 0] : self;
      for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
      if (lane == 0) dred[i * 16 + wave] = t;
    }
  }
  lds_barrier();
  if ((int)threadIdx.x < kd) {
    double t = 0;
    for (int w = 0; w < nw; w++) t += dred[threadIdx.x * 16 + w];
    dots[threadIdx.x] = t;
  }
  __syncthreads();
}

// x = sum_i c_i xR_i (Xr_times_ci src/tools_host.h:405-416); the
// coefficients are broadcast from LDS, two entries per trip with their loads
// issued together.
__device__ __noinline__ void diis_combine(const DevSys &S0, int k, int kd) {
  CSys &S = uniform_sys(S0);
  k = __builtin_amdgcn_readfirstlane(k);
  kd = __builtin_amdgcn_readfirstlane(kd);
  const int nvar = S.nit * S.M, m = S.m;
  const int tid = work_tid(S), nt = S.threads;
  gdbl *x = hbm(S.x);
  const gdbl *xR = hbm(S.xR);
  ldbl *c = (ldbl *)s_coef;
  if ((int)threadIdx.x < kd) c[threadIdx.x] = S0.sc->C[threadIdx.x];
  lds_barrier();
  int p0 = k - kd + 1;
  if (p0 < 0) p0 += m;
  for (int v0 = tid; v0 < nvar; v0 += 2 * nt) {
    const int v1 = v0 + nt;
    const bool has1 = v1 < nvar;
    const gdbl *r = xR + (size_t)p0 * nvar;
    double s0 = 0.0 + c[0] * r[v0];
    double s1 = has1 ? 0.0 + c[0] * r[v1] : 0.0;
    for (int i = 1; i < kd; i++) {
      int posi = k - kd + 1 + i;
      if (posi < 0) posi += m;
      r = xR + (size_t)posi * nvar;
      const double ci = c[i];
      s0 += ci * r[v0];
      if (has1) s1 += ci * r[v1];
    }
    x[v0] = s0;
    if (has1) x[v1] = s1;
  }
  __syncthreads();
}

// SFNewton::iterate_DIIS (src/sfnewton.cpp:1037-1130).  The residual never
// changes x, so the iteration's x0 is x itself at the update: one pass per
// iteration reads x and g, writes x, xR_k and e_k = x_new - x0, and
// accumulates all kd Gram dots; a second pass forms x = sum_i c_i xR_i.
__shared__ int s_slice_stop;  // time-sliced solve: stop flag broadcast by thread 0

// Returns true when the solve has finished (converged, iteration limit or
// error).  With a budget (device-clock ticks) it may instead stop between
// iterations once the budget is used, with its state (iteration count,
// DIIS ring position and Gram matrix, residual, precision flags) in the
// system's scalars, and return false; called again with fresh = false it
// resumes exactly where it stopped.
__device__ bool fused_solve(const DevSys &S0, bool fresh, unsigned long long budget) {
  CSys &S = uniform_sys(S0);
  double *diis_dots = s_diis;
  double *diis_A = s_diis + kMaxDIIS;
  const int nvar = S.nit * S.M;
  const int tid = work_tid(S), nt = S.threads;
  const int m = S.m;
  const double dmax = S.deltamax, tol = S.tol;
  const int itlim = S.itlim;
  DevScalars &sc = *S0.sc;
  const bool mixed = S.precision != 0;
  double res;
  int it, k_diis;
  if (fresh) {
    res = fused_residual(S0);
    if (res < 0) return true;
    if (mixed) mixed_update(S0, res, true);
    gdbl *x = hbm(S.x), *xR = hbm(S.xR), *xx0 = hbm(S.xx0);
    const gdbl *g = hbm(S.g);
    for (int v = tid; v < nvar; v += nt) {
      const double xo = x[v];
      const double xv = xo + (-dmax) * g[v];
      x[v] = xv;
      xx0[v] = xv - xo;
      xR[v] = xv;
    }
    it = 0;
    k_diis = 0;
  } else {
    res = sc.residual;
    it = sc.iterations;
    k_diis = sc.k_diis;
  }
  const unsigned long long t_slice = wall_clock64();
  bool stopped = false;
  __syncthreads();
  while (res > tol && it < itlim) {
    if (budget) {
      if (threadIdx.x == 0) s_slice_stop = (wall_clock64() - t_slice) > budget ? 1 : 0;
      __syncthreads();
      if (s_slice_stop) {
        stopped = true;
        break;
      }
    }
    it++;
    if (mixed && sc.fp32) {
      if (threadIdx.x == 0) sc.fp32_iterations++;
      res = fused_residual(S0);
      if (res < 0) {
        // an FP32 evaluation that failed (e.g. GN underflow) is redone in FP64
        if (threadIdx.x == 0) {
          sc.status = DS_OK;
          sc.fp32 = 0;
        }
        __syncthreads();
        res = fused_residual(S0);
      }
    } else {
      res = fused_residual(S0);
    }
    if (res < 0) break;
    if (mixed) {
      mixed_update(S0, res, false);
      if (sc.mixed_restart) {
        k_diis = 0;
        __syncthreads();
        if (threadIdx.x == 0) sc.mixed_restart = 0;
      }
    }
    const int k = it % m;
    k_diis++;
    const int kd = k_diis > m ? m : k_diis;
    diis_update_dots(S0, k, kd, diis_dots);
    diis_gram_update_block(sc, m, k_diis, diis_dots, diis_A);
    if (threadIdx.x < 64) {
      const bool ok = wave_diis_coefficients(diis_A, kd, sc.C);
      if (threadIdx.x == 0 && !ok) sc.status = DS_NAN;
    }
    __syncthreads();
    if (sc.status != DS_OK) break;
    diis_combine(S0, k, kd);
  }
  if (threadIdx.x == 0) {
    sc.iterations = it;
    sc.k_diis = k_diis;
    sc.converged = (sc.status == DS_OK && res <= tol) ? 1 : 0;
  }
  __syncthreads();
  return !stopped;
}

// ------------------------------------------------------------- kernels
// One residual evaluation or one full solve of system S by this workgroup,
// with the thermodynamic summary and the device-clock bookkeeping.
__device__ void run_system(const DevSys &S, int mode, unsigned long long budget = 0) {
  DevScalars &sc = *S.sc;
  const bool fresh = mode != FUSED_SLICE || sc.phase == 0;
  if (fresh && threadIdx.x == 0) {
    sc.status = DS_OK;
    sc.fp32 = 0;
    sc.fp32_iterations = 0;
    sc.mixed_restart = 0;
    for (int p = 0; p < 4; p++) sc.t_phase[p] = 0;
    sc.t_begin = wall_clock64();
  }
  __syncthreads();
  bool fin = true;
  if (mode == FUSED_RESIDUAL) {
    const double r = fused_residual(S);
    if (r >= 0) fused_thermo(S);
  } else {
    fin = fused_solve(S, fresh, mode == FUSED_SLICE ? budget : 0);
    if (fin && sc.status == DS_OK) fused_thermo(S);
  }
  if (threadIdx.x == 0) {
    sc.phase = fin ? 2 : 1;
    if (fin) {
      sc.t_end = wall_clock64();
      const unsigned long long in_res = sc.t_phase[0] + sc.t_phase[1] + sc.t_phase[2];
      const unsigned long long total = sc.t_end - sc.t_begin;
      sc.t_phase[3] = total > in_res ? total - in_res : 0;
    }
  }
  __syncthreads();
}

__global__ void __launch_bounds__(1024) k_fused(DevSys *const *systems, int mode, unsigned long long budget,
                                                int *phase_out) {
  const DevSys &S = *systems[blockIdx.x];
  run_system(S, mode, budget);
  if (phase_out && threadIdx.x == 0) phase_out[blockIdx.x] = S.sc->phase;
}

// Fields of the first system of a job from its seed (nvar values).
__device__ __noinline__ void seed_fields(const DevSys &D0, const double *seed) {
  CSys &D = uniform_sys(D0);
  const int nvar = D.nit * D.M;
  gdbl *x = hbm(D.x);
  const gdbl *sd = hbm(seed);
  for (int v = threadIdx.x; v < nvar; v += blockDim.x) x[v] = sd[v];
  __syncthreads();
}

// Fields of a chained system from the preceding one, block by block as the
// host mapped them by segment name (initial_guess : previous_result);
// blocks without a counterpart start at zero.
__device__ __noinline__ void carry_fields(const DevSys &D0, const DevSys &S0) {
  CSys &D = uniform_sys(D0);
  const int M = D.M;
  gdbl *x = hbm(D.x);
  const gdbl *xs = hbm(S0.x);
  for (int j = 0; j < D.nit; j++) {
    const int sj = D.carry_src[j];
    for (int i = threadIdx.x; i < M; i += blockDim.x)
      x[(size_t)j * M + i] = sj >= 0 ? xs[(size_t)sj * M + i] : 0.0;
  }
  __syncthreads();
}

// Persistent job queue.  Workgroup w owns slot w: `chain` systems solved one
// after another, the first from the job's seed fields, each later one from
// its predecessor's fields.  A workgroup takes the next job number from the
// shared counter until the counter passes njobs, so uneven jobs balance over
// the slots; every workgroup leaves the loop once the queue is empty.
__global__ void __launch_bounds__(1024) k_fused_jobs(DevSys *const *slots, int chain, const double *seeds,
                                                     int seed_stride, int njobs, int *counter, JobRecord *rec) {
  __shared__ int job_s;
  DevSys *const *mine = slots + (size_t)blockIdx.x * chain;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) job_s = atomicAdd(counter, 1);
    __syncthreads();
    const int job = job_s;
    if (job >= njobs) break;
    for (int c = 0; c < chain; c++) {
      const DevSys &S = *mine[c];
      if (c == 0)
        seed_fields(S, seeds + (size_t)job * seed_stride);
      else
        carry_fields(S, *mine[c - 1]);
      run_system(S, FUSED_SOLVE);
      if (threadIdx.x == 0) {
        const DevScalars &sc = *S.sc;
        JobRecord &r = rec[(size_t)job * chain + c];
        r.iterations = sc.iterations;
        r.status = sc.status;
        r.fp32_iterations = sc.fp32_iterations;
        r.converged = sc.converged;
        r.residual = sc.residual;
        r.F = sc.F;
        r.GP = sc.GP;
        r.t_begin = sc.t_begin;
        r.t_end = sc.t_end;
        for (int q = 0; q < 4; q++) r.t_phase[q] = sc.t_phase[q];
      }
    }
  }
}

hipError_t launch_fused_jobs(DevSys *const *slots, int nslots, int chain, const double *seeds, int seed_stride,
                             int njobs, int *counter, JobRecord *rec, int threads, size_t lds_bytes,
                             hipStream_t stream) {
  if (lds_bytes > 0) {
    hipError_t e = hipFuncSetAttribute((const void *)k_fused_jobs, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds_bytes);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_fused_jobs, dim3(nslots), dim3(threads), lds_bytes, stream, slots, chain, seeds,
                     seed_stride, njobs, counter, rec);
  return hipGetLastError();
}

hipError_t launch_fused(DevSys *const *d_systems, int nsys, int mode, int threads, size_t lds_bytes,
                        hipStream_t stream, unsigned long long budget_ticks, int *phase_out) {
  if (lds_bytes > 0) {
    hipError_t e = hipFuncSetAttribute((const void *)k_fused, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds_bytes);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_fused, dim3(nsys), dim3(threads), lds_bytes, stream, d_systems, mode, budget_ticks,
                     phase_out);
  return hipGetLastError();
}

size_t fused_static_lds() {
  hipFuncAttributes a;
  if (hipFuncGetAttributes(&a, (const void *)k_fused) != hipSuccess) return 0;
  return a.sharedSizeBytes;
}

}  // namespace nscf
