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

// ------------------------------------------------------------- memory spaces
// Pointers are cast to their address space so that the compiler emits
// global_* (HBM) and ds_* (LDS) instructions rather than flat ones: a flat
// access counts against both the vector-memory and the LDS wait counters,
// so every wait for an LDS read would also wait for HBM traffic in flight.
typedef __attribute__((address_space(1))) double gdbl;
typedef __attribute__((address_space(3))) double ldbl;
// The system descriptor is read-only during a launch and the same for the
// whole workgroup; read through the constant address space it is fetched
// with scalar loads.
typedef const __attribute__((address_space(4))) DevSys CSys;

__device__ __forceinline__ CSys &uniform_sys(const DevSys &S) {
  const unsigned long long p = (unsigned long long)&S;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)p);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(p >> 32));
  return *(CSys *)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ gdbl *hbm(double *p) { return (gdbl *)p; }
__device__ __forceinline__ const gdbl *hbm(const double *p) { return (const gdbl *)p; }
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
  // backward with composition, Gf(s-1) fetched one step ahead
  ldbl *ga = ra, *gb = rb;
  const int bl = b1 - 1;
  g1 = active ? G1[(size_t)S.block_seg[bl] * M + ii] : 0.0;
  v = g1;
  double racc = 0.0, gf = 0.0;
  if (active) {
    racc = Gf[(size_t)(N - 1) * M + ii] * v;
    ga[ii] = v;
    if (N >= 2) gf = Gf[(size_t)(N - 2) * M + ii];
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
        if (s >= 1) gf = Gf[(size_t)(s - 1) * M + ii];
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
      rho[i] += gf[i] * v;
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
        rho[i] += gf[i] * v;
      }
      Rows<kLds>::barrier();
      RowT *t = gbA;
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
// use the same layout with the dense column stride nact.
// The 9-point weights are those of LGrad2::propagate (src/LGrad2.cpp:1058-1073);
// the planar full stencil (src/LG2Planar.cpp:1085-1096) is the same form with
// lambda- = lambda+ = 1/6.
template <int SX, int STRIDE>
__device__ __forceinline__ void strip_step(const ldbl *__restrict__ p, int am0, int ap0, int sl, int sr, int dm,
                                           int dp, const ldbl *__restrict__ lamx, double C1, double C2, double C3,
                                           const double (&g1)[SX], double (&v)[SX]) {
  // columns x0-1 (slot sl), x0 .. x0+SX-1 (own), x0+SX (slot sr); rows y-1 (a), y (b), y+1 (c)
  double am = p[sl + dm], bm = p[sl], cm = p[sl + dp];
  double a0 = p[am0], b0 = v[0], c0 = p[ap0];
#pragma unroll
  for (int k = 0; k < SX; k++) {
    double ap, bp, cp;
    if (k + 1 < SX) {
      ap = p[am0 + (k + 1) * STRIDE];
      bp = v[k + 1];
      cp = p[ap0 + (k + 1) * STRIDE];
    } else {
      ap = p[sr + dm];
      bp = p[sr];
      cp = p[sr + dp];
    }
    const double lmk = lamx[2 * k], lpk = lamx[2 * k + 1];
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
    v[k] = t * g1[k];
    am = a0;
    a0 = ap;
    bm = b0;
    b0 = bp;
    cm = c0;
    c0 = cp;
  }
}

template <int SX, int STRIDE>
__device__ __noinline__ void chain2d_strip(const DevSys &S0, int mo) {
  CSys &S = uniform_sys(S0);
  const int MX = S.MX, MY = S.MY, JX = S.JX;
  const int nact = (MX / SX) * MY;
  const int t = threadIdx.x;
  const bool active = t < nact;
  const int xs = active ? t / MY : 0;
  const int y = active ? t - xs * MY + 1 : 1;
  const int x0 = xs * SX + 1;
  // neighbours after set_bounds: y +- 1 as slot offsets inside the strip,
  // x0-1 and x0+SX as slots of the strips that own them
  const int dm = map_y(S, y - 1) - y, dp = map_y(S, y + 1) - y;
  const int xl = map_x(S, x0 - 1) - 1, xr = map_x(S, x0 + SX) - 1;
  const int sl = (xl % SX) * STRIDE + (xl / SX) * MY + (y - 1);
  const int sr = (xr % SX) * STRIDE + (xr / SX) * MY + (y - 1);
  const int tt = active ? t : 0;
  const int am0 = tt + dm, ap0 = tt + dp;
  const bool cyl = S.geometry != 0;
  const bool sc = S.lattice_type == 0;
  const double C1 = sc ? 4.0 / 6.0 : 0.5, C2 = sc ? 1.0 / 6.0 : 0.25, C3 = sc ? 4.0 : 2.0;
  const int b0 = S.mol_block_begin[mo], b1 = S.mol_block_begin[mo + 1];
  const int N = S.mol_N[mo];
  const size_t M = S.M;
  const size_t RG = (size_t)SX * nact;  // Gf row length (interior sites)
  gdbl *Gf = hbm(S.Gf) + tt;
  const gdbl *G1 = hbm(S.G1);
  ldbl *ra = lds_row(0), *rb = lds_row(SX * STRIDE);
  ldbl *lam = (ldbl *)s_lam;
  for (int x = threadIdx.x; x < MX + 2; x += blockDim.x) {
    lam[2 * x] = cyl ? hbm(S.lam_m)[x * JX + 1] : 1.0 / 6.0;
    lam[2 * x + 1] = cyl ? hbm(S.lam_p)[x * JX + 1] : 1.0 / 6.0;
  }
  lds_barrier();
  const ldbl *lamx = lam + 2 * x0;
  const size_t site0 = (size_t)x0 * JX + y;  // site of column 0; column k is site0 + k JX
  double g1[SX], v[SX], racc[SX], gfv[SX];

#define LOAD_G1(seg)                                                   \
  {                                                                    \
    const gdbl *p_ = G1 + (size_t)(seg) * M + site0;                   \
    _Pragma("unroll") for (int k = 0; k < SX; k++) g1[k] = p_[k * JX]; \
  }
#define STORE_ROW(row)                                                            \
  {                                                                               \
    _Pragma("unroll") for (int k = 0; k < SX; k++) (row)[k * STRIDE + tt] = v[k]; \
  }
#define STORE_GF(s_)                                                            \
  {                                                                             \
    gdbl *q_ = Gf + (size_t)(s_) * RG;                                          \
    _Pragma("unroll") for (int k = 0; k < SX; k++) q_[(size_t)k * nact] = v[k]; \
  }
#define LOAD_GF(s_)                                                               \
  {                                                                               \
    const gdbl *q_ = Gf + (size_t)(s_) * RG;                                      \
    _Pragma("unroll") for (int k = 0; k < SX; k++) gfv[k] = q_[(size_t)k * nact]; \
  }

  // ---- forward
  if (active) LOAD_G1(S.block_seg[b0]);
#pragma unroll
  for (int k = 0; k < SX; k++) v[k] = active ? g1[k] : 0.0;
  if (active) {
    STORE_GF(0);
    STORE_ROW(ra);
  }
  lds_barrier();
  ldbl *prev = ra, *next = rb;
  int s = 1;
  for (int b = b0; b < b1; b++) {
    if (b != b0 && active) LOAD_G1(S.block_seg[b]);
    const int n = S.block_n[b] - (b == b0 ? 1 : 0);
    for (int j = 0; j < n; j++, s++) {
      if (active) {
        strip_step<SX, STRIDE>(prev, am0, ap0, sl, sr, dm, dp, lamx, C1, C2, C3, g1, v);
        STORE_GF(s);
        STORE_ROW(next);
      }
      lds_barrier();
      ldbl *tmp = prev;
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
      for (int k = 0; k < SX; k++) acc += (cyl ? Lw[k * JX] : 1.0) * v[k];
    }
    const double gn = wg_sum(acc);
    if (threadIdx.x == 0) S0.sc->GN[mo] = gn;
  }
  // ---- backward with composition; the density of the current block in
  // registers, Gf(s-1) fetched one step ahead
  ldbl *ga = ra, *gb = rb;
  const int bl = b1 - 1;
  if (active) {
    LOAD_G1(S.block_seg[bl]);
    LOAD_GF(N - 1);
#pragma unroll
    for (int k = 0; k < SX; k++) {
      v[k] = g1[k];
      racc[k] = gfv[k] * v[k];
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
        strip_step<SX, STRIDE>(ga, am0, ap0, sl, sr, dm, dp, lamx, C1, C2, C3, g1, v);
#pragma unroll
        for (int k = 0; k < SX; k++) racc[k] += gfv[k] * v[k];
        STORE_ROW(gb);
        if (s >= 1) LOAD_GF(s - 1);
      }
      lds_barrier();
      ldbl *tmp = ga;
      ga = gb;
      gb = tmp;
    }
    if (active) {
      gdbl *rho = hbm(S.rho) + (size_t)S.block_mm[b] * M + site0;
#pragma unroll
      for (int k = 0; k < SX; k++) rho[k * JX] += racc[k];
    }
  }
#undef LOAD_G1
#undef STORE_ROW
#undef STORE_GF
#undef LOAD_GF
  __syncthreads();
}

// Second-order Markov chain (1 gradient, three direction components), rows
// of 3M in LDS (kLds) or HBM.
template <bool kLds>
__device__ __noinline__ void chain_markov2(const DevSys &S0, int mo) {
  typedef typename Rows<kLds>::T RowT;
  CSys &S = uniform_sys(S0);
  const int M = S.M, W = S.M;
  const int tid = work_tid(S), nt = S.threads;
  const int b0 = S.mol_block_begin[mo], b1 = S.mol_block_begin[mo + 1];
  const int N = S.mol_N[mo];
  const int R = 3 * M;
  const double P[3] = {S.mol_P[mo][0], S.mol_P[mo][1], S.mol_P[mo][2]};
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
    for (int i = tid; i < W; i += nt) {
      const double v = g1[i];
      for (int c = 0; c < 3; c++) {
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
      for (int i = tid; i < W; i += nt) {
        double o[3];
        m2_step<double>(S, prev, P, true, i, o);
        const double w = g1[i];
        for (int c = 0; c < 3; c++) {
          const double v = o[c] * w;
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
  {
    const gdbl *gl = Gf + (size_t)(N - 1) * R;
    double a0 = 0, a1 = 0, a2 = 0;
    for (int i = tid; i < W; i += nt) {
      if (i >= 1 && i <= S.MX) {
        const double wt = site_weight(S0, i);
        a0 += wt * gl[i];
        a1 += wt * gl[M + i];
        a2 += wt * gl[2 * M + i];
      }
    }
    double gn = wg_sum(a0);
    gn += 4.0 * wg_sum(a1);
    gn += wg_sum(a2);
    if (threadIdx.x == 0) S0.sc->GN[mo] = gn / 6.0;
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
  Rows<kLds>::barrier();
  s = N - 2;
  for (int b = b1 - 1; b >= b0; b--) {
    const gdbl *g1 = G1 + (size_t)S.block_seg[b] * M;
    gdbl *rho = hbm(S.rho) + (size_t)S.block_mm[b] * M;
    for (int j = (b == b1 - 1 ? S.block_n[b] - 2 : S.block_n[b] - 1); j >= 0; j--, s--) {
      const gdbl *gf = Gf + (size_t)s * R;
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
      Rows<kLds>::barrier();
      RowT *t = gbA;
      gbA = gbB;
      gbB = t;
    }
  }
  __syncthreads();
}

__device__ __noinline__ void chain_strip_dispatch(const DevSys &S, int mo) {
  const int sx = S.strip_sx, st = S.strip_stride;
  if (sx == 8) {
    if (st == 256)
      chain2d_strip<8, 256>(S, mo);
    else if (st == 512)
      chain2d_strip<8, 512>(S, mo);
    else
      chain2d_strip<8, 1024>(S, mo);
  } else {
    if (st == 256)
      chain2d_strip<4, 256>(S, mo);
    else if (st == 512)
      chain2d_strip<4, 512>(S, mo);
    else
      chain2d_strip<4, 1024>(S, mo);
  }
}

// ------------------------------------------------------------- residual
// One Classical_residual: fields x -> g, densities, norms.  Returns ||g||
// (or max|g|) to every thread, or a negative value on a status error.
__device__ double fused_residual(const DevSys &S) {
  const int M = S.M, W = n_work_sites(S);
  const int tid = work_tid(S), nt = S.threads;
  DevScalars &sc = *S.sc;
  const unsigned long long t0 = wall_clock64();
  for (int w = tid; w < W; w += nt) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    site_prepare(S, i, x, y, z);
    for (int mm = 0; mm < S.nmm; mm++) S.rho[mm * M + i] = 0.0;
  }
  __syncthreads();
  const unsigned long long t1 = wall_clock64();
  for (int mo = 0; mo < S.nmol; mo++) {
    if (S.mol_markov[mo] == 2) {
      if (S.use_lds)
        chain_markov2<true>(S, mo);
      else
        chain_markov2<false>(S, mo);
    } else if (S.strip_sx > 0) {
      chain_strip_dispatch(S, mo);
    } else if (S.gradients == 1 && S.use_lds && S.M <= S.threads) {
      chain1d_m1(S, mo);
    } else if (S.use_lds) {
      chain_markov1<true>(S, mo);
    } else {
      chain_markov1<false>(S, mo);
    }
  }
  const unsigned long long t2 = wall_clock64();
  if (threadIdx.x == 0) {
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
  const double r = S.max_g ? wg_max(acc) : sqrt(wg_sum(acc));
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
__device__ __noinline__ void fused_thermo(const DevSys &S) {
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
  f = wg_sum(f);
  gp = wg_sum(gp);
#pragma unroll
  for (int mo = 0; mo < kMaxMol; mo++)
    if (mo < S.nmol) th[mo] = wg_sum(th[mo]);
  if (threadIdx.x == 0) {
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
__device__ void fused_solve(const DevSys &S0) {
  CSys &S = uniform_sys(S0);
  double *diis_dots = s_diis;
  double *diis_A = s_diis + kMaxDIIS;
  const int nvar = S.nit * S.M;
  const int tid = work_tid(S), nt = S.threads;
  const int m = S.m;
  const double dmax = S.deltamax, tol = S.tol;
  const int itlim = S.itlim;
  DevScalars &sc = *S0.sc;
  gdbl *x = hbm(S.x), *x0 = hbm(S.x0), *xR = hbm(S.xR), *xx0 = hbm(S.xx0);
  const gdbl *g = hbm(S.g);
  for (int v = tid; v < nvar; v += nt) x0[v] = x[v];
  __syncthreads();
  double res = fused_residual(S0);
  if (res < 0) return;
  for (int v = tid; v < nvar; v += nt) {
    const double xv = x[v] + (-dmax) * g[v];
    x[v] = xv;
    xx0[v] = xv - x0[v];
    xR[v] = xv;
  }
  int it = 0, k_diis = 0;
  __syncthreads();
  while (res > tol && it < itlim) {
    it++;
    for (int v = tid; v < nvar; v += nt) x0[v] = x[v];
    __syncthreads();
    res = fused_residual(S0);
    if (res < 0) break;
    const int k = it % m;
    k_diis++;
    gdbl *xRk = xR + (size_t)k * nvar, *xxk = xx0 + (size_t)k * nvar;
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
      const gdbl *a = xx0 + (size_t)posi * nvar;
      double acc = 0;
      for (int v = tid; v < nvar; v += nt) acc += a[v] * xxk[v];
      acc = wg_sum(acc);
      if (threadIdx.x == 0) diis_dots[i] = acc;
    }
    __syncthreads();
    if (threadIdx.x == 0) diis_gram_update(sc, m, k_diis, diis_dots, diis_A);
    __syncthreads();
    if (threadIdx.x < 64) {
      const bool ok = wave_diis_coefficients(diis_A, kd, sc.C);
      if (threadIdx.x == 0 && !ok) sc.status = DS_NAN;
    }
    __syncthreads();
    if (sc.status != DS_OK) break;
    const gdbl *C = hbm((const double *)sc.C);
    for (int v = tid; v < nvar; v += nt) {
      int posi = k - kd + 1;
      if (posi < 0) posi += m;
      double xv = 0.0 + C[0] * xR[(size_t)posi * nvar + v];
      for (int i = 1; i < kd; i++) {
        posi = k - kd + 1 + i;
        if (posi < 0) posi += m;
        xv += C[i] * xR[(size_t)posi * nvar + v];
      }
      x[v] = xv;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    sc.iterations = it;
    sc.converged = (sc.status == DS_OK && res <= tol) ? 1 : 0;
  }
  __syncthreads();
}

// ------------------------------------------------------------- kernel
__global__ void __launch_bounds__(1024) k_fused(DevSys *const *systems, int mode) {
  const DevSys &S = *systems[blockIdx.x];
  if (threadIdx.x == 0) {
    S.sc->status = DS_OK;
    for (int p = 0; p < 4; p++) S.sc->t_phase[p] = 0;
    S.sc->t_begin = wall_clock64();
  }
  __syncthreads();
  if (mode == FUSED_RESIDUAL) {
    const double r = fused_residual(S);
    if (r >= 0) fused_thermo(S);
  } else {
    fused_solve(S);
    if (S.sc->status == DS_OK) fused_thermo(S);
  }
  if (threadIdx.x == 0) {
    DevScalars &sc = *S.sc;
    sc.t_end = wall_clock64();
    const unsigned long long in_res = sc.t_phase[0] + sc.t_phase[1] + sc.t_phase[2];
    const unsigned long long total = sc.t_end - sc.t_begin;
    sc.t_phase[3] = total > in_res ? total - in_res : 0;
  }
}

hipError_t launch_fused(DevSys *const *d_systems, int nsys, int mode, int threads, size_t lds_bytes,
                        hipStream_t stream) {
  if (lds_bytes > 0) {
    hipError_t e = hipFuncSetAttribute((const void *)k_fused, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds_bytes);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_fused, dim3(nsys), dim3(threads), lds_bytes, stream, d_systems, mode);
  return hipGetLastError();
}

size_t fused_static_lds() {
  hipFuncAttributes a;
  if (hipFuncGetAttributes(&a, (const void *)k_fused) != hipSuccess) return 0;
  return a.sharedSizeBytes;
}

}  // namespace nscf
