// Device-side description of one SCF system and the per-site lattice
// operators of the hot path (propagator step, site average, densities,
// residual, free-energy densities).
//
// Layout in HBM: every lattice array is x-major with one ghost layer per side
// (M = (MX+2)(MY+2)(MZ+2) sites, last index contiguous), the reference's own
// indexing (src/LGrad1.cpp:182, src/LGrad2.cpp:109, src/LGrad3.cpp:19), so
// fields and profiles cross the boundary without reshuffling.  Ghost layers
// are never filled in a separate pass: a neighbour read that falls on a ghost
// is remapped to the site the reference's set_bounds would copy there
// (mirror: first/last layer, periodic: opposite layer; corners map in every
// direction).  With a `surface` boundary (1 gradient only) the ghost keeps
// its own value and is recomputed every step with the truncated stencil, as
// the reference does (src/LG1Planar.cpp:138-159).
#pragma once
#include <hip/hip_runtime.h>

#include "scf_linalg.h"

namespace nscf {

constexpr int kMaxSeg = 8;
constexpr int kMaxMol = 8;
constexpr int kMixedPatience = 25;  // mixed precision: FP32 evaluations without a new best residual before FP64
constexpr int kMaxBlocks = 16;
constexpr int kMaxMM = 16;
constexpr int kMaxFjc = 4;   // lattice refinement up to FJC_choices 9
constexpr int kMaxComp = 2 * kMaxFjc + 1;  // Markov-2 direction components (FJC)

enum DevStatus : int {
  DS_OK = 0,
  DS_GN = -1,        // GN <= 0 (reference throws -1, src/system.cpp:2398-2402) or not finite
  DS_NAN = -2,       // NaN in the DIIS matrix (-2, src/sfnewton.cpp:808-809)
  DS_SVD = -3,       // NaN in the SVD (-3, src/tools.cpp:1225-1228)
  DS_PHIB = -4,      // negative solvent phibulk (-4, src/system.cpp:2521-2525)
};

// Scalars of one system, kept in device memory next to its fields.
struct DevScalars {
  double GN[kMaxMol], norm[kMaxMol], n[kMaxMol], phibulk[kMaxMol], theta[kMaxMol], Mu[kMaxMol];
  double seg_phibulk[kMaxSeg];
  double F, GP, residual, alpha_unused;
  double A[kMaxDIIS * kMaxDIIS];  // DIIS Gram matrix, leading dimension m
  double C[kMaxDIIS];
  double red[kMaxDIIS + 4];       // reduction results handed between phases
  int status, iterations, k_diis, converged;
  unsigned long long t_begin, t_end;  // device wall clock at solve start / end
  unsigned long long t_phase[4];      // wall-clock ticks: prepare, chains, densities+residual, DIIS
  double lnshift[kMaxMol];            // FP32 propagators: gauge shift c per molecule (G1' = G1 e^c)
  int fp32;                           // 1 while the solve evaluates propagators in FP32
  int fp32_iterations;                // iterations of the last solve done with FP32 propagators
  int mixed_since, mixed_restart;     // evaluations since the best FP32 residual; DIIS restart flag
  int phase;                          // time-sliced solve: 0 not started, 1 running, 2 finished
  double mixed_best;                  // best residual of the FP32 phase
};

// Result of one solve of a queued job (persistent job queue).
struct JobRecord {
  int iterations, status, fp32_iterations, converged;
  double residual, F, GP;
  unsigned long long t_begin, t_end;
  unsigned long long t_phase[4];  // prepare, chains, densities+residual, rest (device ticks)
};

struct DevSys {
  // ---- lattice
  int gradients, geometry, lattice_type, stencil_full;
  int MX, MY, MZ, M, JX, JY;
  int bx1, bxm, by1, bym, bz1, bzm;
  int surface_lo, surface_hi;  // 1-gradient surface boundaries
  int fjc;                     // lattice refinement (grid layers per bond; > 1 only with 1 gradient)
  int bx1f[kMaxFjc], bxmf[kMaxFjc];  // refined ghost sources along x
  int by1f[kMaxFjc], bymf[kMaxFjc];  // along y (2 and 3 gradients)
  int bz1f[kMaxFjc], bzmf[kMaxFjc];  // along z (3 gradients)
  double wplanar;              // site weight of planar sums: 1 / fjc^(gradients)
  const double *LAM;           // refined 1-gradient weights: (2 fjc + 1) rows of M
  double lambda;
  const double *L, *lam_m, *lam_0, *lam_p;  // M each (1 and 2 gradients)
  const double *l1, *l11, *l_1, *l_11;      // Markov-2 curved weights (1 gradient)
  // ---- segments
  int nseg, nit;
  int seg_block[kMaxSeg];   // field block of x, -1 for frozen segments
  int seg_frozen[kMaxSeg], seg_pinned[kMaxSeg];
  int seg_tagged[kMaxSeg];  // tagged: weight and density are the mask, no field
  int itmon[kMaxSeg];
  int sysmon[kMaxSeg];      // 1 if the segment is used by a molecule
  double chi[kMaxSeg * kMaxSeg];
  const double *seg_mask;   // nseg * M (1.0 on frozen / pinned sites)
  const double *ksam;       // M (0 on frozen sites)
  // ---- molecules (linear chains: a sequence of blocks)
  int nmol, solvent, nblocks, nmm;
  int mol_N[kMaxMol], mol_freedom[kMaxMol], mol_pinned[kMaxMol], mol_markov[kMaxMol], mol_size[kMaxMol];
  int mol_tagged[kMaxMol], mol_NF[kMaxMol];  // holds a tagged segment; N less the tagged segment
  int mol_ring[kMaxMol];                      // ring: first and last one-segment blocks merged
  int mol_block_begin[kMaxMol + 1];
  int mol_mm_begin[kMaxMol + 1];
  int block_seg[kMaxBlocks], block_n[kMaxBlocks], block_mm[kMaxBlocks];
  int mm_seg[kMaxMM];
  double mol_P[kMaxMol][kMaxComp];  // Markov-2 bending weights, FJC entries
  double mol_phibulk_in[kMaxMol], mol_n_in[kMaxMol];
  double mol_charge[kMaxMol];  // net valence per segment (Molecule::Charge)
  int neutralizer;             // molecule whose bulk fraction neutralizes the bulk, or -1
  double frac[kMaxMol][kMaxSeg];
  double mol_constF[kMaxMol];
  // ---- work arrays (device)
  double *x, *g;
  double *G1;         // nseg * M
  float *G1f;         // nseg * M (3-gradient grid path): FP32 segment weights times the gauge factor e^c
  // host only: the row schedule of a branched molecule (a BranchPlan, see
  // scf_spec.h), null for a linear chain; never read on the device
  const void *branch[kMaxMol];
  double *Gf;         // max over molecules of N * size * M (a branched molecule: its schedule's rows)
  double *Gb;         // 2 * size * M
  double *rho;        // nmm * M (molecule-segment densities)
  double *phi;        // nseg * M
  double *phi_side;   // nseg * M
  double *phitot, *alpha;  // M
  double *mol_phitot; // nmol * M
  // ---- electrostatics (charged systems, 1 gradient): x and g end with the
  // potential block; nitm counts the segment field blocks in front of it
  int charged, nitm, el_ri;  // el_ri: integer radius offset of the field energy (curved)
  int fixed_psi0, grad_eps;  // frozen segments fix the potential; segments differ in epsilon
  int seg_fixpsi[kMaxSeg], el_last_fixed;  // frozen segments with a fixed potential; the last of them
  double seg_val[kMaxSeg], seg_eps[kMaxSeg], seg_psi0[kMaxSeg];
  double el_C0;              // e^2 / (eps0 kT b), the constant of the fixed-potential sweeps
  double el_pf, el_C;        // field-energy prefactor and Poisson constant (geometry factors included)
  double el_r0;              // offset_first_layer * fjc
  double *psi, *EE, *q, *eps;  // M each: potential (mirrored ghosts), field energy, charge, dielectric constant
  double *elA, *elB;           // M each: the potential sweep as X[x] = A[x] X[x-1] + B[x]
  // ---- DIIS state
  double *xR, *xx0, *x0;  // m * nvar, m * nvar, nvar
  int m, itlim, max_g;
  double deltamax, tol;
  // ---- execution
  int use_lds;            // propagator rows fit the LDS budget of the launch
  int all_sites;          // 2 and 3 gradients: phases visit ghosts too (charged systems, literal chains)
  int row_stride;         // doubles between the two LDS row buffers
  int strip_sx;           // 2-gradient strip path: sites per thread along x (0: not used)
  int strip_stride;       // 2-gradient strip path: LDS slots per strip column
  int threads;            // preferred workgroup size for the fused kernel
  int precision;          // 0: FP64 throughout; 1: FP32 propagators, FP64 refinement
  int carry_src[kMaxSeg + 1];  // job queue: field block of the preceding system for each block (-1: zero)
  double mixed_switch;    // residual below which a mixed solve continues in FP64
  double volume;          // weighted interior volume (sum of site weights)
  int ckpt;               // 2-gradient strip path: keep every ckpt-th end-point row (0: all rows)
  // ---- scalars and reduction workspace
  DevScalars *sc;
  double *partial;        // per-workgroup partial sums (multi-workgroup path)
};

// ------------------------------------------------------------ address spaces
// Device arrays are accessed through pointers cast to the global address
// space (global_* instructions) and LDS rows through the LDS address space
// (ds_*), rather than generic (flat) pointers: a flat access counts against
// both the vector-memory and the LDS wait counters, so every wait for an
// LDS value would also wait for HBM traffic in flight.
typedef __attribute__((address_space(1))) double gdbl;
typedef __attribute__((address_space(3))) double ldbl;
// The system descriptor is read-only during a launch and the same for the
// whole workgroup; read through the constant address space it is fetched
// with scalar loads and kept in scalar registers.
typedef const __attribute__((address_space(4))) DevSys CSys;

__device__ __forceinline__ gdbl *hbm(double *p) { return (gdbl *)p; }
__device__ __forceinline__ const gdbl *hbm(const double *p) { return (const gdbl *)p; }
__device__ __forceinline__ CSys &uniform_sys(const DevSys &S) {
  const unsigned long long p = (unsigned long long)&S;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)p);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(p >> 32));
  return *(CSys *)(((unsigned long long)hi << 32) | lo);
}

// ------------------------------------------------------------ index helpers
template <typename Sys>
__device__ __forceinline__ int map_x(const Sys &S, int x) {
  if (S.fjc > 1) {
    if (x < S.fjc) return S.bx1f[x];
    if (x >= S.MX + S.fjc) return S.bxmf[x - S.MX - S.fjc];
    return x;
  }
  return x < 1 ? S.bx1 : (x > S.MX ? S.bxm : x);
}
// 1-gradient site x (0..M-1) -> the site set_bounds copies there; on a
// refined lattice the ghosts are fjc layers per side (src/LGrad1.cpp:823-835,
// src/LGrad2.cpp:1674-1687)
template <typename Sys>
__device__ __forceinline__ int map_x1(const Sys &S, int x) {
  return map_x(S, x);
}
template <typename Sys>
__device__ __forceinline__ int map_y(const Sys &S, int y) {
  if (S.fjc > 1) {
    if (y < S.fjc) return S.by1f[y];
    if (y >= S.MY + S.fjc) return S.bymf[y - S.MY - S.fjc];
    return y;
  }
  return y < 1 ? S.by1 : (y > S.MY ? S.bym : y);
}
template <typename Sys>
__device__ __forceinline__ int map_z(const Sys &S, int z) {
  if (S.fjc > 1) {
    if (z < S.fjc) return S.bz1f[z];
    if (z >= S.MZ + S.fjc) return S.bzmf[z - S.MZ - S.fjc];
    return z;
  }
  return z < 1 ? S.bz1 : (z > S.MZ ? S.bzm : z);
}

// number of sites a phase visits: all sites in 1 gradient (ghosts included,
// for surface boundaries) and for charged systems (whose potential reads
// densities on ghosts), interior sites otherwise
template <typename Sys>
__device__ __forceinline__ int n_work_sites(const Sys &S) {
  return S.gradients == 1 || S.all_sites ? S.M : S.MX * S.MY * (S.gradients == 3 ? S.MZ : 1);
}

// work index -> lattice site (and its coordinates)
template <typename Sys>
__device__ __forceinline__ int work_site(const Sys &S, int w, int &x, int &y, int &z) {
  if (S.gradients == 1) {
    x = w;
    y = z = 0;
    return w;
  }
  if (S.all_sites) {  // every site, ghosts included, in memory order
    x = w / S.JX;
    const int r = w - x * S.JX;
    if (S.gradients == 2) {
      y = r;
      z = 0;
    } else {
      y = r / S.JY;
      z = r - y * S.JY;
    }
    return w;
  }
  if (S.gradients == 2) {
    const int xs = w / S.MY;
    x = xs + S.fjc;
    y = w - xs * S.MY + S.fjc;
    z = 0;
    return x * S.JX + y;
  }
  const int yz = S.MY * S.MZ, f = S.fjc;
  const int xs = w / yz;
  const int r = w - xs * yz;
  const int ys = r / S.MZ;
  x = xs + f;
  y = ys + f;
  z = r - ys * S.MZ + f;
  return x * S.JX + y * S.JY + z;
}

template <typename Sys>
__device__ __forceinline__ bool is_interior(const Sys &S, int x, int y, int z) {
  if (S.gradients == 1) return x >= S.fjc && x < S.MX + S.fjc;
  if (S.gradients == 2) return x >= S.fjc && x < S.MX + S.fjc && y >= S.fjc && y < S.MY + S.fjc;
  if (S.fjc > 1)
    return x >= S.fjc && x < S.MX + S.fjc && y >= S.fjc && y < S.MY + S.fjc && z >= S.fjc && z < S.MZ + S.fjc;
  if (x < 1 || x > S.MX) return false;
  if (S.gradients >= 2 && (y < 1 || y > S.MY)) return false;
  if (S.gradients == 3 && (z < 1 || z > S.MZ)) return false;
  return true;
}

// The site set_bounds copies into site (x, y, z) of a 2- or 3-gradient
// lattice (fjc = 1), or the site itself.  2 gradients follow LGrad2::set_bounds
// (src/LGrad2.cpp:1652-1673): the y ghosts of the interior rows, then the x
// ghost rows, then each corner from its own ghost row at the y source, so a
// corner next to a y surface keeps its value.  3 gradients (mirror and
// periodic boundaries only) map coordinate by coordinate.
template <typename Sys>
__device__ __forceinline__ int bound_src(const Sys &S, int x, int y, int z) {
  if (S.gradients == 3) return map_x(S, x) * S.JX + map_y(S, y) * S.JY + map_z(S, z);
  const bool gx = x < 1 || x > S.MX, gy = y < 1 || y > S.MY;
  const int yb = map_y(S, y);
  if (gx && gy && (yb < 1 || yb > S.MY)) return x * S.JX + y;
  return map_x(S, x) * S.JX + yb;
}


// ----------------------------------------------------- Markov-1 step / Side
// Lattice average of a bounded array at site i=(x,y,z) for the propagator
// (without the Boltzmann factor).  Arithmetic order follows the reference's
// element-wise passes (LG1Planar::propagate src/LG1Planar.cpp:146-149,
// LGrad1::propagate src/LGrad1.cpp:441-445, LGrad2::propagate
// src/LGrad2.cpp:1058-1073, LG2Planar::propagate src/LG2Planar.cpp:1085-1096,
// LGrad3::propagate src/LGrad3.cpp:733-743).
// 3-gradient stencils over the 26 neighbours (LGrad3::propagate
// src/LGrad3.cpp:686-729 without stencil_full: faces, edges and corners with
// the weights 16 : 4 : 1 / 152 on the simple-cubic lattice and 4 : 2 : 1 / 56
// on the hexagonal one; the hexagonal stencil_full form, src/LGrad3.cpp:
// 746-782, adds the site twice and sums to 64/40, as the reference does).
// Neighbours are summed in the reference's pass order; ghosts map
// coordinate by coordinate.
template <typename T, typename Sys, typename Ptr>
__device__ __forceinline__ T stencil3_wide(const Sys &S, Ptr g, int x, int y, int z) {
  const int JX = S.JX, JY = S.JY;
  const int xm = map_x(S, x - 1) * JX, x0 = x * JX, xp = map_x(S, x + 1) * JX;
  const int ym = map_y(S, y - 1) * JY, y0 = y * JY, yp = map_y(S, y + 1) * JY;
  const int zm = map_z(S, z - 1), z0 = z, zp = map_z(S, z + 1);
  const bool hexfull = S.stencil_full;  // only the hexagonal lattice reaches here with stencil_full
  const T f1 = hexfull ? (T)2.0 : (S.lattice_type == 0 ? (T)4.0 : (T)2.0);
  T v = hexfull ? (T)2.0 * g[x0 + y0 + z0] : (T)0;
  v += g[xm + y0 + z0];
  v += g[xp + y0 + z0];
  v += g[x0 + ym + z0];
  v += g[x0 + yp + z0];
  v += g[x0 + y0 + zm];
  v += g[x0 + y0 + zp];
  v *= f1;
  v += g[xm + ym + z0];
  v += g[xp + yp + z0];
  if (hexfull) {
    v += g[xm + yp + z0];
    v += g[xp + ym + z0];
    v += g[xm + y0 + zm];
    v += g[xp + y0 + zm];
    v += g[x0 + ym + zm];
    v += g[x0 + yp + zm];
    v += g[xm + y0 + zp];
    v += g[xp + y0 + zp];
    v += g[x0 + ym + zp];
    v += g[x0 + yp + zp];
  } else {
    v += g[xp + ym + z0];
    v += g[xm + yp + z0];
    v += g[xm + y0 + zm];
    v += g[xp + y0 + zp];
    v += g[xm + y0 + zp];
    v += g[xp + y0 + zm];
    v += g[x0 + ym + zm];
    v += g[x0 + yp + zp];
    v += g[x0 + ym + zp];
    v += g[x0 + yp + zm];
  }
  v *= f1;
  v += g[xm + ym + zm];
  v += g[xp + yp + zp];
  v += g[xm + ym + zp];
  v += g[xp + yp + zm];
  v += g[xm + yp + zm];
  v += g[xp + ym + zp];
  v += g[xp + ym + zm];
  v += g[xm + yp + zp];
  const T c = hexfull ? (T)(1.0 / 40.0) : (S.lattice_type == 0 ? (T)(1.0 / 152.0) : (T)(1.0 / 56.0));
  return v * c;
}

// Refined 1-gradient stencil (fjc > 1): 2 fjc + 1 taps weighted by the
// LAMBDA rows, summed in the order of LG1Planar::propagate / LGrad1::
// propagate / Side (src/LG1Planar.cpp:151-156, src/LGrad1.cpp:447-452,
// src/LG1Planar.cpp:47-52): for distance fjc down to 1 the lower then the
// upper neighbour, the site itself last.  A neighbour outside 0..M-1 is
// missing (the reference's shifted passes stop at the array ends).
template <typename T, typename Sys, typename Ptr>
__device__ __forceinline__ T refined_taps(const Sys &S, Ptr g, int x) {
  const int M = S.M, f = S.fjc, F = 2 * f + 1;
  const gdbl *LA = hbm(S.LAM);
  T v = 0;
  for (int j = 0; j < f; j++) {
    const int kk = f - j;
    if (x - kk >= 0) v += (T)LA[j * M + x] * g[map_x1(S, x - kk)];
    if (x + kk <= M - 1) v += (T)LA[(F - 1 - j) * M + x] * g[map_x1(S, x + kk)];
  }
  v += (T)LA[f * M + x] * g[map_x1(S, x)];
  return v;
}

// Refined 2-gradient stencils (fjc > 1) at an interior site, block sums
// over the (2 fjc + 1)^2 offsets in the reference's pass order.  Cylindrical
// (LGrad2::propagate / Side src/LGrad2.cpp:1091-1106, 264-278): offsets
// with |dy| < fjc weighted by the LAMBDA row of dx, times 2, then those with
// |dy| = fjc, times 1 / (2 (fjc + 2)).  Planar (LG2Planar src/LG2Planar.cpp:
// 1110-1124, 117-130): plain sums of the offsets with none, one, two
// coordinates at +-fjc, times 2, 2 and 1 / (4 (FJC - 2) FJC + 1).
template <typename T, typename Sys, typename Ptr>
__device__ __forceinline__ T refined2(const Sys &S, Ptr g, int i, int x, int y) {
  const int f = S.fjc, F = 2 * S.fjc + 1, M = S.M, JX = S.JX;
  const bool planar = S.geometry == 0;
  const gdbl *LA = hbm(S.LAM);
  const int nblocks = planar ? 3 : 2;
  T v = 0;
  for (int block = 0; block < nblocks; block++) {
    for (int dx = -f; dx <= f; dx++) {
      const int sx = map_x(S, x + dx) * JX;
      for (int dy = -f; dy <= f; dy++) {
        const int bk = (planar && (dx == -f || dx == f) ? 1 : 0) + (dy == -f || dy == f ? 1 : 0);
        if (bk != block) continue;
        const T gv = g[sx + map_y(S, y + dy)];
        if (planar)
          v += gv;
        else
          v += (T)LA[(f + dx) * M + i] * gv;
      }
    }
    if (planar)
      v *= block != 2 ? (T)2.0 : (T)(1.0 / (4.0 * (F - 2) * F + 1));
    else
      v *= block != 1 ? (T)2.0 : (T)(1.0 / (2 * (f + 2)));
  }
  return v;
}

// Refined 3-gradient hexagonal stencil (fjc = 2; LGrad3::propagate / Side
// src/LGrad3.cpp:808-830, 206-226): plain sums of the (2 fjc + 1)^3 offsets
// with none, one, two, three coordinates at +-fjc, times 2, 2, 2 and
// 1/8/((FJC-2)^3 + 3 (FJC-2)^2 + 3 (FJC-2) + 1), in the reference's order.
template <typename T, typename Sys, typename Ptr>
__device__ __forceinline__ T refined3(const Sys &S, Ptr g, int x, int y, int z) {
  const int f = S.fjc, F = 2 * S.fjc + 1, JX = S.JX, JY = S.JY;
  T v = 0;
  for (int block = 0; block < 4; block++) {
    for (int dx = -f; dx <= f; dx++) {
      const int sx = map_x(S, x + dx) * JX;
      for (int dy = -f; dy <= f; dy++) {
        const int sy = sx + map_y(S, y + dy) * JY;
        for (int dz = -f; dz <= f; dz++) {
          const int bk = (dx == -f || dx == f) + (dy == -f || dy == f) + (dz == -f || dz == f);
          if (bk == block) v += g[sy + map_z(S, z + dz)];
        }
      }
    }
    v *= block != 3 ? (T)2.0
                    : (T)(1.0 / 8.0 / ((F - 2) * (F - 2) * (F - 2) + 3 * (F - 2) * (F - 2) + 3 * (F - 2) + 1));
  }
  return v;
}

template <typename T, typename Sys, typename Ptr>
__device__ __forceinline__ T stencil_prop(const Sys &S, Ptr g, int i, int x, int y, int z) {
  if (S.gradients == 1) {
    const int M = S.M;
    if (S.fjc > 1) return refined_taps<T>(S, g, x);
    if (S.geometry == 0) {
      const T lam = (T)S.lambda, c0 = (T)(1.0 - 2.0 * S.lambda);
      T v = 0;
      if (x >= 1) v = lam * g[map_x(S, x - 1)];
      v += c0 * g[map_x(S, x)];
      if (x <= M - 2) v += lam * g[map_x(S, x + 1)];
      return v;
    }
    T v = (T)hbm(S.lam_0)[i] * g[map_x(S, x)];
    if (x >= 1) v += (T)hbm(S.lam_m)[i] * g[map_x(S, x - 1)];
    if (x <= M - 2) v += (T)hbm(S.lam_p)[i] * g[map_x(S, x + 1)];
    return v;
  }
  if (S.gradients == 2 && S.fjc > 1) return refined2<T>(S, g, i, x, y);
  if (S.gradients == 2) {
    const int JX = S.JX;
    const int xm = map_x(S, x - 1) * JX, x0 = x * JX, xp = map_x(S, x + 1) * JX;
    const int ym = map_y(S, y - 1), yp = map_y(S, y + 1);
    // diagonal neighbours; a corner ghost next to a y surface keeps its own
    // value (bound_src), which the coordinate-wise map does not give
    int dmm = xm + ym, dmp = xm + yp, dpm = xp + ym, dpp = xp + yp;
    if (S.all_sites) {
      dmm = bound_src(S, x - 1, y - 1, 0);
      dmp = bound_src(S, x - 1, y + 1, 0);
      dpm = bound_src(S, x + 1, y - 1, 0);
      dpp = bound_src(S, x + 1, y + 1, 0);
    }
    if (S.geometry == 0) {
      if (S.stencil_full && S.lattice_type != 0) {
        // hexagonal 9-point, LG2Planar::propagate src/LG2Planar.cpp:1092-1108:
        // (2 (2 centre + 4 nearest) + 4 diagonals) / 16
        T v = (T)2.0 * g[x0 + y];
        v += g[xm + y];
        v += g[xp + y];
        v += g[x0 + ym];
        v += g[x0 + yp];
        v *= (T)2.0;
        v += g[dmm];
        v += g[dpp];
        v += g[dmp];
        v += g[dpm];
        return v * (T)(1.0 / 16.0);
      }
      if (S.stencil_full) {
        const T C1 = (T)(16.0 / 36.0), C2 = (T)(4.0 / 36.0), C3 = (T)(1.0 / 36.0);
        T v = C1 * g[x0 + y];
        v += C2 * g[x0 + ym];
        v += C2 * g[x0 + yp];
        v += C2 * g[xm + y];
        v += C2 * g[xp + y];
        v += C3 * g[dmm];
        v += C3 * g[dmp];
        v += C3 * g[dpm];
        v += C3 * g[dpp];
        return v;
      }
      T v = g[xm + y];
      v += g[xp + y];
      v += g[x0 + ym];
      v += g[x0 + yp];
      v *= (T)0.5;
      v += g[x0 + y];
      return v * (T)(1.0 / 3.0);
    }
    const bool sc = S.lattice_type == 0;
    const T C1 = (T)(sc ? 4.0 / 6.0 : 0.5), C2 = (T)(sc ? 1.0 / 6.0 : 0.25), C3 = (T)(sc ? 4.0 : 2.0);
    const T lm = (T)hbm(S.lam_m)[i], lp = (T)hbm(S.lam_p)[i];
    T v = C1 * g[x0 + y];
    v += lm * g[xm + y];
    v += lp * g[xp + y];
    v += C2 * g[x0 + ym];
    v += C2 * g[x0 + yp];
    v *= C3;
    v += lm * g[dmm];
    v += lm * g[dmp];
    v += lp * g[dpm];
    v += lp * g[dpp];
    return v * C2;
  }
  const int JX = S.JX, JY = S.JY;
  if (S.fjc > 1) return refined3<T>(S, g, x, y, z);
  if (S.stencil_full && S.lattice_type == 0) {
    const int bx = x * JX, by = y * JY;
    T v = g[map_x(S, x - 1) * JX + by + z];
    v += g[map_x(S, x + 1) * JX + by + z];
    v += g[bx + map_y(S, y - 1) * JY + z];
    v += g[bx + map_y(S, y + 1) * JY + z];
    v += g[bx + by + map_z(S, z - 1)];
    v += g[bx + by + map_z(S, z + 1)];
    return v * (T)(1.0 / 6.0);
  }
  return stencil3_wide<T>(S, g, x, y, z);
}

// Site average <X> for the Flory-Huggins term (Lattice::Side,
// src/LG1Planar.cpp:28-55, src/LGrad1.cpp:251-281, src/LGrad2.cpp:207-240,
// src/LG2Planar.cpp:27-100, src/LGrad3.cpp:156-166).  Identical weights to
// the propagator; the 1-gradient planar form sums in a different order.
template <typename Sys, typename Ptr>
__device__ __forceinline__ double stencil_side(const Sys &S, Ptr X, int i, int x, int y, int z) {
  if (S.gradients == 1 && S.fjc > 1) return refined_taps<double>(S, X, x);
  if (S.gradients == 1 && S.geometry == 0) {
    const int M = S.M;
    const double lam = S.lambda;
    double v = 0;
    if (x >= 1) v = lam * X[map_x(S, x - 1)];
    if (x <= M - 2) v += lam * X[map_x(S, x + 1)];
    v += (1.0 - 2.0 * lam) * X[map_x(S, x)];
    return v;
  }
  return stencil_prop<double>(S, X, i, x, y, z);
}

// ----------------------------------------------------- Markov-2 (1 gradient)
// Bounded value of direction component c (0: +x, 1: lateral, 2: -x) of a
// Markov-2 row at position x.  The reference bounds components 0 and 2 as a
// pair, so a mirror reflection turns an outgoing walker into an incoming one
// (LGrad1::set_bounds(X,Y) src/LGrad1.cpp:787-804).
template <typename T, typename Sys, typename Ptr>
__device__ __forceinline__ T m2_val(const Sys &S, Ptr row, int c, int x) {
  const int M = S.M;
  if (x >= 1 && x <= S.MX) return row[c * M + x];
  const bool lo = x < 1;
  const int b = lo ? S.bx1 : S.bxm;
  const bool surf = lo ? S.surface_lo : S.surface_hi;
  if (c == 1) return row[M + b];
  if (c == 0) return row[2 * M + b];
  return surf ? row[2 * M + x] : row[b];
}

// Forward (F) or backward (B) Markov-2 step at site x: writes the three
// direction components (without the Boltzmann factor) to out[0..2].
// Follows LG1Planar::propagateF/B (src/LG1Planar.cpp:57-136) and
// LGrad1::propagateF/B (src/LGrad1.cpp:285-429), simple-cubic lattice.
template <typename T, typename Sys, typename Ptr>
__device__ __forceinline__ void m2_step(const Sys &S, Ptr g, const double *P, bool forward, int x, T out[3]) {
  const int M = S.M;
  const T P0 = (T)P[0], P1 = (T)P[1], P1x4 = (T)(4 * P[1]), Pm = (T)(2 * P[1] + P[0]);
  const bool hasm = x >= 1, hasp = x <= M - 2;
  T a = 0, b = 0, c = 0;
  if (S.geometry == 0) {
    if (forward) {
      if (hasm) {
        a = P0 * m2_val<T>(S, g, 0, x - 1);
        a += P1x4 * m2_val<T>(S, g, 1, x - 1);
      }
      b = P1 * m2_val<T>(S, g, 0, x);
      b += Pm * m2_val<T>(S, g, 1, x);
      b += P1 * m2_val<T>(S, g, 2, x);
      if (hasp) {
        c = P1x4 * m2_val<T>(S, g, 1, x + 1);
        c += P0 * m2_val<T>(S, g, 2, x + 1);
      }
    } else {
      if (hasp) a = P0 * m2_val<T>(S, g, 0, x + 1);
      a += P1x4 * m2_val<T>(S, g, 1, x);
      if (hasp) b = P1 * m2_val<T>(S, g, 0, x + 1);
      b += Pm * m2_val<T>(S, g, 1, x);
      if (hasm) b += P1 * m2_val<T>(S, g, 2, x - 1);
      c = P1x4 * m2_val<T>(S, g, 1, x);
      if (hasm) c += P0 * m2_val<T>(S, g, 2, x - 1);
    }
  } else {
    const T l1 = (T)hbm(S.l1)[x], l11 = (T)hbm(S.l11)[x], l_1 = (T)hbm(S.l_1)[x], l_11 = (T)hbm(S.l_11)[x];
    if (forward) {
      if (hasm) {
        T h = l_1 * m2_val<T>(S, g, 0, x - 1);
        h += l_11 * m2_val<T>(S, g, 2, x);
        a = P0 * h;
        h = l_1 * m2_val<T>(S, g, 1, x - 1);
        h += l_11 * m2_val<T>(S, g, 1, x);
        a += P1x4 * h;
      }
      b = P1 * m2_val<T>(S, g, 0, x);
      b += Pm * m2_val<T>(S, g, 1, x);
      b += P1 * m2_val<T>(S, g, 2, x);
      if (hasp) {
        T h = l1 * m2_val<T>(S, g, 1, x + 1);
        h += l11 * m2_val<T>(S, g, 1, x);
        c = P1x4 * h;
        h = l1 * m2_val<T>(S, g, 2, x + 1);
        h += l11 * m2_val<T>(S, g, 0, x);
        c += P0 * h;
      }
    } else {
      T hm = 0, hp = 0;
      if (hasm) {
        hm = l_1 * m2_val<T>(S, g, 2, x - 1);
        hm += l_11 * m2_val<T>(S, g, 0, x);
      }
      if (hasp) {
        hp = l1 * m2_val<T>(S, g, 0, x + 1);
        hp += l11 * m2_val<T>(S, g, 2, x);
      }
      b = P1 * hm;
      c = P0 * hm;
      a = P1x4 * m2_val<T>(S, g, 1, x);
      b += Pm * m2_val<T>(S, g, 1, x);
      c += P1x4 * m2_val<T>(S, g, 1, x);
      a += P0 * hp;
      b += P1 * hp;
    }
  }
  out[0] = a;
  out[1] = b;
  out[2] = c;
}

// Bounded value of direction component c of a Markov-2 row at position x on
// a planar hexagonal lattice with FJC = 2 fjc + 1 components.  Components c
// and FJC-1-c (c < fjc) are bounded as a pair and the lateral component fjc
// alone (LG1Planar::propagateF/B src/LG1Planar.cpp:64-65, LGrad1::set_bounds
// src/LGrad1.cpp:787-821): a ghost of component c < fjc takes the partner's
// value at the source site, a ghost of c > fjc the partner's value there
// unless the boundary is a surface, where it keeps its own.
template <typename T, typename Sys, typename Ptr>
__device__ __forceinline__ T m2h_val(const Sys &S, Ptr row, int c, int x) {
  const int M = S.M, f = S.fjc, F = 2 * S.fjc + 1;
  if (x >= f && x < S.MX + f) return row[c * M + x];
  const int b = map_x1(S, x);
  const bool surf = x < f ? S.surface_lo : S.surface_hi;
  if (c == f) return row[c * M + b];
  if (c < f) return row[(F - 1 - c) * M + b];
  return surf ? row[c * M + x] : row[(F - 1 - c) * M + b];
}

// Forward or backward Markov-2 step at site x on a planar hexagonal lattice
// (LG1Planar::propagateF/B src/LG1Planar.cpp:82-94,122-134): component p
// advances fjc - p sites per bond; the weight from component q is
// P[|p-q|] + P[FJC-1-|FJC-1-p-q|] (the second term only for q off the x
// axis).  Forward, G'_p(x) = sum_q w G_q(x - fjc + p), q ascending; backward,
// G'_p(x) = sum_q w G_q(x + fjc - q), q descending.  Writes FJC components
// (without the Boltzmann factor) to out.
template <typename T, typename Sys, typename Ptr>
__device__ __forceinline__ void m2h_step(const Sys &S, Ptr g, const double *P, bool forward, int x,
                                         T out[kMaxComp]) {
  const int M = S.M, f = S.fjc, F = 2 * S.fjc + 1;
  if (S.geometry != 0) {
    // curved, fjc = 1 (LGrad1::propagateF / propagateB src/LGrad1.cpp:298-321,
    // 374-398): a component that steps off the axis mixes the incoming
    // component with its reflection, weighted by l_1 / l_11 (towards the
    // centre) or l1 / l11 (outwards) at x
    const T l1 = (T)hbm(S.l1)[x], l11 = (T)hbm(S.l11)[x], l_1 = (T)hbm(S.l_1)[x], l_11 = (T)hbm(S.l_11)[x];
    for (int p = 0; p < F; p++) {
      T acc = 0;
      for (int qq = 0; qq < F; qq++) {
        const int q = forward ? qq : F - 1 - qq;
        // the step's direction: towards lower x (d < 0) for p < f forward
        // and for q > f backward
        const int d = forward ? p - f : f - q;
        double c = P[p > q ? p - q : q - p];
        if (q > 0 && q < F - 1) c += P[F - 1 - abs(F - 1 - p - q)];
        if (c == 0) continue;
        T h;
        if (d < 0) {  // reads x - 1
          if (x < 1) continue;
          h = l_1 * m2h_val<T>(S, g, q, x - 1);
          h += l_11 * m2h_val<T>(S, g, F - 1 - q, x);
        } else if (d > 0) {  // reads x + 1
          if (x > M - 2) continue;
          h = l1 * m2h_val<T>(S, g, q, x + 1);
          h += l11 * m2h_val<T>(S, g, F - 1 - q, x);
        } else {
          h = m2h_val<T>(S, g, q, x);
        }
        acc += (T)c * h;
      }
      out[p] = acc;
    }
    return;
  }
  for (int p = 0; p < F; p++) {
    T acc = 0;
    if (forward) {
      const int src = x - f + p;
      if (src >= 0 && src <= M - 1)
        for (int q = 0; q < F; q++) {
          double c = P[p > q ? p - q : q - p];
          if (q > 0 && q < F - 1) c += P[F - 1 - abs(F - 1 - p - q)];
          if (c != 0) acc += (T)c * m2h_val<T>(S, g, q, src);
        }
    } else {
      for (int q = F - 1; q >= 0; q--) {
        const int src = x + f - q;
        if (src < 0 || src > M - 1) continue;
        double c = P[p > q ? p - q : q - p];
        if (q > 0 && q < F - 1) c += P[F - 1 - abs(F - 1 - p - q)];
        if (c != 0) acc += (T)c * m2h_val<T>(S, g, q, src);
      }
    }
    out[p] = acc;
  }
}

// Bounded value of direction component c (0: +x, 1: +y, 2: out of plane,
// 3: -y, 4: -x) of a 2-gradient Markov-2 row at (x, y): at a reflecting x
// boundary components 0 and 4 swap, at a reflecting y boundary 1 and 3, the
// rest (and everything at a periodic boundary) takes its own value at the
// source site (LGrad2::set_bounds_x/y src/LGrad2.cpp:1555-1650).  Ghost
// corners are never read by the axis-aligned Markov-2 steps.
template <typename T, typename Sys, typename Ptr>
__device__ __forceinline__ T m2d_val(const Sys &S, Ptr row, int c, int x, int y) {
  const int M = S.M, JX = S.JX;
  if (x < 1 || x > S.MX) {
    const int b = x < 1 ? S.bx1 : S.bxm;
    const int cc = (S.bx1 <= S.bxm && (c == 0 || c == 4)) ? 4 - c : c;
    return row[cc * M + b * JX + y];
  }
  if (y < 1 || y > S.MY) {
    const int b = y < 1 ? S.by1 : S.bym;
    const int cc = (S.by1 <= S.bym && (c == 1 || c == 3)) ? 4 - c : c;
    return row[cc * M + x * JX + b];
  }
  return row[c * M + x * JX + y];
}

// Forward or backward 2-gradient Markov-2 step at interior site (x, y),
// simple cubic (LGrad2::propagateF/B src/LGrad2.cpp:645-681,1003-1039 with
// the curved reflection weights; LG2Planar::propagateF/B
// src/LG2Planar.cpp:445-481,930-966 without them).  Straight continuation
// has weight P0, a turn P1, the out-of-plane component counts both
// perpendicular directions (2 P1).  Writes the five components without the
// Boltzmann factor, each summed in the reference's pass order.
template <typename T, typename Sys, typename Ptr>
__device__ __forceinline__ void m2d_step(const Sys &S, Ptr g, const double *P, bool forward, int i, int x, int y,
                                         T out[5]) {
  const T P0 = (T)P[0], P1 = (T)P[1], P2 = (T)(2 * P[1]);
  const bool cyl = S.geometry != 0;
  auto z = [&](int c, int xx, int yy) { return m2d_val<T>(S, g, c, xx, yy); };
  if (forward) {
    T hl = z(0, x - 1, y), hu = z(4, x + 1, y);
    if (cyl) {
      hl = (T)hbm(S.l_1)[i] * z(0, x - 1, y);
      hl += (T)hbm(S.l_11)[i] * z(4, x, y);
      hu = (T)hbm(S.l1)[i] * z(4, x + 1, y);
      hu += (T)hbm(S.l11)[i] * z(0, x, y);
    }
    T o = P0 * hl;
    o += P1 * z(1, x - 1, y);
    o += P2 * z(2, x - 1, y);
    o += P1 * z(3, x - 1, y);
    out[0] = o;
    o = P1 * z(0, x, y - 1);
    o += P0 * z(1, x, y - 1);
    o += P2 * z(2, x, y - 1);
    o += P1 * z(4, x, y - 1);
    out[1] = o;
    o = P1 * z(0, x, y);
    o += P1 * z(1, x, y);
    o += P0 * z(2, x, y);
    o += P1 * z(3, x, y);
    o += P1 * z(4, x, y);
    out[2] = o;
    o = P1 * z(0, x, y + 1);
    o += P2 * z(2, x, y + 1);
    o += P0 * z(3, x, y + 1);
    o += P1 * z(4, x, y + 1);
    out[3] = o;
    o = P1 * z(1, x + 1, y);
    o += P2 * z(2, x + 1, y);
    o += P1 * z(3, x + 1, y);
    o += P0 * hu;
    out[4] = o;
  } else {
    T hl = z(4, x - 1, y), hu = z(0, x + 1, y);
    if (cyl) {
      hl = (T)hbm(S.l_1)[i] * z(4, x - 1, y);
      hl += (T)hbm(S.l_11)[i] * z(0, x, y);
      hu = (T)hbm(S.l1)[i] * z(0, x + 1, y);
      hu += (T)hbm(S.l11)[i] * z(4, x, y);
    }
    const T z3m = z(3, x, y - 1), z2 = z(2, x, y), z1p = z(1, x, y + 1);
    T o = P1 * z3m;
    o += P2 * z2;
    o += P1 * z1p;
    o += P0 * hu;
    out[0] = o;
    o = P1 * hl;
    o += P2 * z2;
    o += P0 * z1p;
    o += P1 * hu;
    out[1] = o;
    o = P1 * hl;
    o += P1 * z3m;
    o += P0 * z2;
    o += P1 * z1p;
    o += P1 * hu;
    out[2] = o;
    o = P1 * hl;
    o += P0 * z3m;
    o += P2 * z2;
    o += P1 * hu;
    out[3] = o;
    o = P0 * hl;
    o += P1 * z3m;
    o += P2 * z2;
    o += P1 * z1p;
    out[4] = o;
  }
}

// Bounded value of direction component c (0: +x, 1: +y, 2: +z, 3: -z,
// 4: -y, 5: -x) of a 3-gradient Markov-2 row at (x, y, z): at a reflecting
// boundary of an axis that axis' pair of directions swaps, other components
// and periodic boundaries take their own values (LGrad3::set_bounds_x/y/z
// src/LGrad3.cpp:1376-1522).  Only face ghosts are read.
template <typename T, typename Sys, typename Ptr>
__device__ __forceinline__ T m3_val(const Sys &S, Ptr row, int c, int x, int y, int z) {
  const int M = S.M, JX = S.JX, JY = S.JY;
  if (x < 1 || x > S.MX) {
    const int b = x < 1 ? S.bx1 : S.bxm;
    const int cc = (S.bx1 <= S.bxm && (c == 0 || c == 5)) ? 5 - c : c;
    return row[cc * M + b * JX + y * JY + z];
  }
  if (y < 1 || y > S.MY) {
    const int b = y < 1 ? S.by1 : S.bym;
    const int cc = (S.by1 <= S.bym && (c == 1 || c == 4)) ? 5 - c : c;
    return row[cc * M + x * JX + b * JY + z];
  }
  if (z < 1 || z > S.MZ) {
    const int b = z < 1 ? S.bz1 : S.bzm;
    const int cc = (S.bz1 <= S.bzm && (c == 2 || c == 3)) ? 5 - c : c;
    return row[cc * M + x * JX + y * JY + b];
  }
  return row[c * M + x * JX + y * JY + z];
}

// Forward or backward 3-gradient Markov-2 step at interior site (x, y, z),
// simple cubic (LGrad3::propagateF/B src/LGrad3.cpp:399-446, 605-663): a
// walker never reverses, straight on weighs P0 and a turn P1.  Forward,
// component p arrives from the site behind it and sums the incoming
// components q != 5 - p in increasing q; backward, the components are
// summed in decreasing q from the site ahead of each incoming direction.
template <typename T, typename Sys, typename Ptr>
__device__ __forceinline__ void m3_step(const Sys &S, Ptr g, const double *P, bool forward, int x, int y, int z,
                                        T out[6]) {
  const T P0 = (T)P[0], P1 = (T)P[1];
  // displacement of direction d
  const int dx[6] = {1, 0, 0, 0, 0, -1}, dy[6] = {0, 1, 0, 0, -1, 0}, dz[6] = {0, 0, 1, -1, 0, 0};
#pragma unroll
  for (int p = 0; p < 6; p++) {
    T o = 0;
    if (forward) {
#pragma unroll
      for (int q = 0; q < 6; q++) {
        if (q == 5 - p) continue;
        o += (q == p ? P0 : P1) * m3_val<T>(S, g, q, x - dx[p], y - dy[p], z - dz[p]);
      }
    } else {
#pragma unroll
      for (int q = 5; q >= 0; q--) {
        if (p == 5 - q) continue;
        o += (q == p ? P0 : P1) * m3_val<T>(S, g, q, x + dx[q], y + dy[q], z + dz[q]);
      }
    }
    out[p] = o;
  }
}

// ----------------------------------------------------- electrostatics
// Charged systems on 1-gradient lattices.  The potential psi is the last
// block of x; its ghosts are mirrored whatever the boundary conditions
// (LGrad1::set_M_bounds src/LGrad1.cpp:820-832).
template <typename Sys>
__device__ __forceinline__ int el_mirror(const Sys &S, int x) {
  const int f = S.fjc;
  if (x < f) return 2 * f - x - 1;
  if (x >= S.MX + f) return 2 * (S.MX + f) - x - 1;
  return x;
}

template <typename Sys>
__device__ __forceinline__ double el_psi(const Sys &S, int x) {
  return hbm(S.x)[S.nitm * S.M + el_mirror(S, x)];
}

// coordinates of memory index i (x-major, z contiguous)
template <typename Sys>
__device__ __forceinline__ void site_coords(const Sys &S, int i, int &x, int &y, int &z) {
  x = i / S.JX;
  const int r = i - x * S.JX;
  if (S.gradients == 2) {
    y = r;
    z = 0;
  } else {
    y = r / S.JY;
    z = r - y * S.JY;
  }
}

// psi at site (x, y, z) of a 2- or 3-gradient lattice with the ghosts of
// set_M_bounds (src/LGrad2.cpp:1702-1723, src/LGrad3.cpp:1679-1702, fjc = 1):
// faces mirrored; 2 gradients also the corners, 3 gradients leave edges and
// corners as the field has them.
template <typename Sys>
__device__ __forceinline__ double el_psi_xyz(const Sys &S, int x, int y, int z) {
  const gdbl *P = hbm(S.x) + S.nitm * S.M;
  auto mir = [](int c, int n) { return c < 1 ? 1 : (c > n ? n : c); };
  if (S.gradients == 2) return P[mir(x, S.MX) * S.JX + mir(y, S.MY)];
  const int nghost = (x < 1 || x > S.MX) + (y < 1 || y > S.MY) + (z < 1 || z > S.MZ);
  if (nghost >= 2) return P[x * S.JX + y * S.JY + z];
  return P[mir(x, S.MX) * S.JX + mir(y, S.MY) * S.JY + mir(z, S.MZ)];
}

// Field energy density at an interior site of a 2- or 3-gradient lattice
// (LG2Planar::UpdateEE src/LG2Planar.cpp:1130-1152, LGrad2::UpdateEE
// src/LGrad2.cpp:1392-1417 with its integer radius, LGrad3::UpdateEE
// src/LGrad3.cpp:1131-1141: squared differences along each axis).
template <typename Sys>
__device__ __forceinline__ double el_EE_xyz(const Sys &S, int i, int x, int y, int z) {
  const double p0 = el_psi_xyz(S, x, y, z);
  const double xm = el_psi_xyz(S, x - 1, y, z), xp = el_psi_xyz(S, x + 1, y, z);
  const double ym = el_psi_xyz(S, x, y - 1, z), yp = el_psi_xyz(S, x, y + 1, z);
  if (S.gradients == 3) {
    const double zm = el_psi_xyz(S, x, y, z - 1), zp = el_psi_xyz(S, x, y, z + 1);
    double e = 0.0;
    e += (zm - p0) * (zm - p0) + (p0 - zp) * (p0 - zp);
    e += (xm - p0) * (xm - p0) + (p0 - xp) * (p0 - xp);
    e += (ym - p0) * (ym - p0) + (p0 - yp) * (p0 - yp);
    return e * S.el_pf;
  }
  if (S.geometry == 0) {
    double a = p0 - xm, b = p0 - xp, c = p0 - ym, d = p0 - yp;
    a *= a;
    b *= b;
    c *= c;
    d *= d;
    return S.el_pf * (a + b + c + d);
  }
  const int r = S.el_ri + (x - S.fjc + 1);
  double a = p0 - xm;
  a *= (r - 0.5) * a * 2 * 3.14159265;
  double b = p0 - xp;
  b *= (r + 0.5) * b * 2 * 3.14159265;
  double c = (p0 - ym) * S.fjc;
  c *= c;
  double d = (p0 - yp) * S.fjc;
  d *= d;
  return S.el_pf * ((a + b) / hbm(S.L)[i] + c + d);
}

// Field energy density at site x, zero on ghosts (LG1Planar::UpdateEE
// src/LG1Planar.cpp:162-178; LGrad1::UpdateEE src/LGrad1.cpp:592-651, where
// the cylindrical and spherical radii are integers).
template <typename Sys>
__device__ __forceinline__ double el_EE(const Sys &S, int x) {
  const int f = S.fjc;
  if (x < f || x >= S.MX + f) return 0.0;
  const double pm = el_psi(S, x - 1), p0 = el_psi(S, x), pp = el_psi(S, x + 1);
  if (S.geometry == 0) {
    double em = pm - p0;
    em *= em;
    double ep = p0 - pp;
    ep *= ep;
    return S.el_pf * (em + ep);
  }
  if (S.geometry == 1) {
    const int r = S.el_ri + (x - f + 1);
    double em = p0 - pm;
    em *= (r - 1) * em;
    double ep = p0 - pp;
    ep *= r * ep;
    return S.el_pf * (em + ep) / hbm(S.L)[x];
  }
  const int r = S.el_ri + (x - f);
  double ep = r * (p0 - pp);
  ep *= ep;
  if (x == f) return S.el_pf * ep / hbm(S.L)[x];
  double em = (r - 1) * (pm - p0);
  em *= em;
  return S.el_pf * (em + ep) / hbm(S.L)[x];
}

// Charge density and dielectric constant at site i (all sites, ghosts
// included) from the segment densities with their ghosts bounded, divided
// by phitot (zero where phitot is zero); also stores psi and EE for the
// thermodynamics (System::DoElectrostatics src/system.cpp:1920-1953; the
// residual sets psi[0] = psi[1], src/system.cpp:2164).
template <typename Sys>
__device__ __forceinline__ void site_electro(const Sys &S, int i) {
  const int M = S.M;
  int b = map_x1(S, i);
  if (S.gradients >= 2) {  // the site set_bounds copies here
    int x, y, z;
    site_coords(S, i, x, y, z);
    b = bound_src(S, x, y, z);
  }
  double q = 0, e = 0;
#pragma unroll
  for (int k = 0; k < kMaxSeg; k++) {
    if (k >= S.nseg) break;
    const double p = hbm(S.phi)[k * M + b];
    q = q + S.seg_val[k] * p;
    e = e + S.seg_eps[k] * p;
  }
  const double pt = hbm(S.phitot)[i];
  hbm(S.q)[i] = pt != 0 ? q / pt : 0.0;
  hbm(S.eps)[i] = pt != 0 ? e / pt : 0.0;
  if (S.gradients == 1) {
    hbm(S.psi)[i] = el_psi(S, i == 0 ? 1 : i);
    hbm(S.EE)[i] = el_EE(S, i);
    return;
  }
  int x, y, z;
  site_coords(S, i == 0 ? 1 : i, x, y, z);
  hbm(S.psi)[i] = el_psi_xyz(S, x, y, z);
  if (i == 0) site_coords(S, i, x, y, z);
  hbm(S.EE)[i] = is_interior(S, x, y, z) ? el_EE_xyz(S, i, x, y, z) : 0.0;
}

// One site of the potential sweep (LG1Planar::UpdatePsi src/LG1Planar.cpp:181-199,
// LGrad1::UpdatePsi src/LGrad1.cpp:654-686) at interior x.  Planar and
// cylindrical sweeps are upwind, X[x] = (epsXmin X[x-1] + C q + epsXplus
// psi[x+1]) / (epsXmin + epsXplus) starting from psi at the lower ghost:
// stored as the affine map X[x] = A X[x-1] + B.  The spherical sweep reads
// psi on both sides and returns X directly.
template <typename Sys>
__device__ __forceinline__ double site_psi_sweep(const Sys &S, int x) {
  const int f = S.fjc;
  const double Em = hbm(S.eps)[x - 1], E0 = hbm(S.eps)[x], Ep = hbm(S.eps)[x + 1];
  const double q = hbm(S.q)[x];
  const double pp = el_psi(S, x + 1);
  double wm = 1.0, wp = 1.0;  // radial weights of the two bonds
  if (S.geometry == 1) {
    wm = S.el_r0 + (x - f);
    wp = S.el_r0 + (x - f + 1);
  } else if (S.geometry == 2) {
    const double rm = S.el_r0 + (x - f), rp = S.el_r0 + (x - f + 1);
    wm = rm * rm;
    wp = rp * rp;
  }
  const double epsXmin = wm * (Em + E0), epsXplus = wp * (E0 + Ep);
  const double den = epsXmin + epsXplus;
  if (S.geometry == 2) {
    const double pm = el_psi(S, x - 1);
    return (epsXmin * pm + S.el_C * q * hbm(S.L)[x] + epsXplus * pp) / den;
  }
  hbm(S.elA)[x] = epsXmin / den;
  hbm(S.elB)[x] = (S.el_C * q + epsXplus * pp) / den;
  return 0.0;
}

// ----------------------------------------------------- per-site phases
// Boltzmann weights of every segment at site i (Segment::PrepareForCalculations
// src/segment.cpp:835-879): u is the segment's field block with its ghosts
// bounded, G1 = exp(-u), times the pinning mask, times KSAM.
template <typename Sys>
__device__ __forceinline__ void site_prepare(const Sys &S, int i, int x, int y, int z) {
  int b;  // bounded source site for u
  if (S.gradients == 1) {
    b = map_x1(S, x);
  } else if (S.all_sites) {
    b = bound_src(S, x, y, z);
  } else if (S.gradients == 2) {
    b = map_x(S, x) * S.JX + map_y(S, y);
  } else {
    b = map_x(S, x) * S.JX + map_y(S, y) * S.JY + map_z(S, z);
  }
  const int M = S.M;
  // charged systems: u also carries -epsilon EE + valence psi at the source
  // site (System::PutU src/system.cpp:2001-2040)
  double EEb = 0.0, psib = 0.0;
  if (S.charged && S.gradients == 1) {
    EEb = el_EE(S, b);
    psib = el_psi(S, b);
  } else if (S.charged) {  // the source is an interior site (mirror and periodic boundaries)
    int bx, by, bz;
    site_coords(S, b, bx, by, bz);
    EEb = el_EE_xyz(S, b, bx, by, bz);
    psib = el_psi_xyz(S, bx, by, bz);
  }
  for (int k = 0; k < S.nseg; k++) {
    const int blk = S.seg_block[k];
    double u = blk >= 0 ? hbm(S.x)[blk * M + b] : 0.0;
    if (S.charged && blk >= 0) {
      u = u + (-1.0 * S.seg_eps[k]) * EEb;
      if (S.seg_val[k] != 0) u = u + S.seg_val[k] * psib;
    }
    double w = exp(-u);
    if (S.seg_pinned[k]) w = w * hbm(S.seg_mask)[k * M + i];
    // a tagged segment's weight is its mask, without KSAM
    // (Segment::PrepareForCalculations src/segment.cpp:859-879)
    if (S.seg_tagged[k])
      w = hbm(S.seg_mask)[k * M + i];
    else
      w = w * hbm(S.ksam)[i];
    hbm(S.G1)[k * M + i] = w;
  }
}

// Densities at site i once GN and the normalisations are known
// (System::ComputePhis src/system.cpp:2345-2610): molecule-segment densities
// rho/G1*norm, segment densities, total density.
template <typename Sys>
__device__ __forceinline__ void site_densities(const Sys &S, int i) {
  const int M = S.M;
  const DevScalars &sc = *S.sc;
  double phitot = 0;
#pragma unroll
  for (int k = 0; k < kMaxSeg; k++)
    if (k < S.nseg && S.seg_frozen[k]) phitot += hbm(S.seg_mask)[k * M + i];
  for (int mo = 0; mo < S.nmol; mo++) {
    const double nrm = sc.norm[mo];
    double mtot = 0;
    for (int mm = S.mol_mm_begin[mo]; mm < S.mol_mm_begin[mo + 1]; mm++) {
      const double g1 = hbm(S.G1)[S.mm_seg[mm] * M + i];
      double p = hbm(S.rho)[mm * M + i];
      p = g1 != 0 ? p / g1 : 0.0;
      if (nrm > 0) p = p * nrm;
      hbm(S.rho)[mm * M + i] = p;
      phitot += p;
      mtot += p;
    }
    hbm(S.mol_phitot)[mo * M + i] = mtot;
  }
  // segment densities: frozen mask, then molecule contributions in order
#pragma unroll
  for (int k = 0; k < kMaxSeg; k++) {
    if (k >= S.nseg) break;
    double p = S.seg_frozen[k] ? hbm(S.seg_mask)[k * M + i] : 0.0;
    for (int mm = 0; mm < S.nmm; mm++)
      if (S.mm_seg[mm] == k) p += hbm(S.rho)[mm * M + i];
    // a tagged segment's density is its mask (src/system.cpp:2596-2601)
    if (S.seg_tagged[k]) p = hbm(S.seg_mask)[k * M + i];
    hbm(S.phi)[k * M + i] = p;
  }
  hbm(S.phitot)[i] = phitot;
}

// Residual at an interior site (System::Classical_residual
// src/system.cpp:2099-2157); also stores the site averages phi_side.
// Returns g_i^2 summed over the field blocks (for the norm).
template <typename Sys>
__device__ __forceinline__ double site_residual(const Sys &S, int i, int x, int y, int z) {
  const int M = S.M;
  const DevScalars &sc = *S.sc;
  // every load first (site averages of all segments, fields, phitot, KSAM),
  // then the stores: a store between them would keep the later loads from
  // being issued early
  double side[kMaxSeg], xin[kMaxSeg];
#pragma unroll
  for (int k = 0; k < kMaxSeg; k++) {
    side[k] = 0.0;
    if (k < S.nseg) side[k] = stencil_side(S, hbm(S.phi) + k * M, i, x, y, z);
  }
#pragma unroll
  for (int it = 0; it < kMaxSeg; it++) xin[it] = it < S.nitm ? hbm(S.x)[it * M + i] : 0.0;
  const double pt = hbm(S.phitot)[i];
  const double ks = hbm(S.ksam)[i];
#pragma unroll
  for (int k = 0; k < kMaxSeg; k++)
    if (k < S.nseg) hbm(S.phi_side)[k * M + i] = side[k];
  double gv[kMaxSeg];
  double alpha = 0;
#pragma unroll
  for (int it = 0; it < kMaxSeg; it++) {
    gv[it] = 0.0;
    if (it < S.nitm) {
      const int si = S.itmon[it];
      double gi = xin[it];
#pragma unroll
      for (int k = 0; k < kMaxSeg; k++) {
        if (k < S.nseg) {
          const double chi = S.chi[si * kMaxSeg + k];
          if (chi != 0 && pt > 0) gi = gi - chi * (side[k] / pt - sc.seg_phibulk[k]);
        }
      }
      gv[it] = gi;
      alpha += gi;
    }
  }
  alpha = alpha * (1.0 / S.nitm);
  hbm(S.alpha)[i] = alpha;
  double g2 = 0;
#pragma unroll
  for (int it = 0; it < kMaxSeg; it++) {
    if (it < S.nitm) {
      double gi = pt > 0 ? gv[it] - alpha + 1 / pt - 1.0 : 0.0;
      gi = gi * ks;
      hbm(S.g)[it * M + i] = gi;
      g2 += gi * gi;
    }
  }
  return g2;
}

// Free-energy and grand-potential densities at an interior site
// (System::GetFreeEnergy src/system.cpp:2809-3029, GetGrandPotential
// src/system.cpp:3038-3244).  Returns both through F and GP.
template <typename Sys>
__device__ __forceinline__ void site_thermo(const Sys &S, int i, double logc[kMaxMol], double &F, double &GP) {
  const int M = S.M;
  const DevScalars &sc = *S.sc;
  double f = 0;
#pragma unroll
  for (int mo = 0; mo < kMaxMol; mo++)
    if (mo < S.nmol) f += logc[mo] * hbm(S.mol_phitot)[mo * M + i];
  for (int j = 0; j < S.nseg; j++) {
    if (!S.sysmon[j]) continue;
    const double g1 = hbm(S.G1)[j * M + i];
    double t = g1 > 0 ? -log(g1) : 0.0;
    f += -(hbm(S.phi)[j * M + i] * t);
  }
  // frozen and tagged partners count in full, their own pairs not at all
  for (int j = 0; j < S.nseg; j++) {
    if (S.seg_frozen[j] || S.seg_tagged[j]) continue;
    const double pj = hbm(S.phi)[j * M + i];
    for (int k = 0; k < S.nseg; k++) {
      double chi = S.chi[j * kMaxSeg + k];
      if (!(S.seg_frozen[k] || S.seg_tagged[k])) chi = chi / 2;
      if (chi != 0) f += pj * hbm(S.phi_side)[k * M + i] * chi;
    }
  }
#pragma unroll
  for (int mo = 0; mo < kMaxMol; mo++)
    if (mo < S.nmol) f += S.mol_constF[mo] * hbm(S.mol_phitot)[mo * M + i];
  const double ks = hbm(S.ksam)[i];
  F = f * ks;
  double qpsi = 0.0;
  if (S.charged) {
    // q psi / 2, frozen sites included (System::GetFreeEnergy src/system.cpp:3019-3027)
    qpsi = hbm(S.q)[i] * hbm(S.psi)[i];
    F = F + qpsi * 0.5;
  }
  double gp = 0;
  for (int mo = 0; mo < S.nmol; mo++)
    gp += (hbm(S.mol_phitot)[mo * M + i] - sc.phibulk[mo]) * (1.0 / S.mol_NF[mo]);
  gp += hbm(S.alpha)[i];
  for (int j = 0; j < S.nseg; j++) {
    if (S.seg_frozen[j] || S.seg_tagged[j]) continue;
    const double pj = hbm(S.phi)[j * M + i], pbA = sc.seg_phibulk[j];
    for (int k = 0; k < S.nseg; k++) {
      if (S.seg_frozen[k] || S.seg_tagged[k]) continue;
      const double chi = S.chi[j * kMaxSeg + k] / 2;
      if (chi != 0) gp += (pj * hbm(S.phi_side)[k * M + i] - pbA * sc.seg_phibulk[k]) * chi;
    }
  }
  if (S.charged) {
    // eps EE - q psi / 2 on accessible sites, q psi / 2 on the others
    // (System::GetGrandPotential src/system.cpp:3223-3238)
    const double q = hbm(S.q)[i];
    double t = hbm(S.EE)[i] * hbm(S.eps)[i];
    t = t * -2.0;
    t = t + qpsi;
    t = t * (-1.0 / 2.0);
    GP = (-gp + t) * ks;
    GP = GP + (q - q * ks) * hbm(S.psi)[i] * 0.5;
  } else {
    GP = -gp * ks;
  }
}

// site weight of sums over the lattice (Lattice::WeightedSum): 1 / fjc on
// planar lattices, the site volume L on curved ones; ghosts are excluded by
// the callers (interior work sites only).
template <typename Sys>
__device__ __forceinline__ double site_weight(const Sys &S, int i) {
  if (S.gradients == 3 || S.geometry == 0) return S.wplanar;
  return hbm(S.L)[i];
}

// Normalisations after the propagators (System::ComputePhis
// src/system.cpp:2366-2558).  Single thread.
__device__ inline void scalars_norms(const DevSys &S) {
  DevScalars &sc = *S.sc;
  double A = 0, B = 0;  // bulk charge and volume fraction of free and restricted molecules
  for (int i = 0; i < S.nmol; i++) {
    // A partition function that overflowed (or came out NaN) ends the
    // evaluation as GN <= 0 does.  The reference carries on: its densities
    // turn NaN, Classical_residual sets g = 0 wherever phi_tot > 0 fails
    // (src/system.cpp:2114-2142), and ||g|| = 0 passes the stop test, so a
    // blown-up iterate would be reported as converged with NaN results.
    if (!isfinite(sc.GN[i])) {
      sc.status = DS_GN;
      return;
    }
    double nrm = 0;
    const int fr = S.mol_freedom[i];
    if (fr == 0) {  // free
      nrm = sc.phibulk[i] / S.mol_N[i];
      sc.n[i] = nrm * sc.GN[i];
      A += sc.phibulk[i] * S.mol_charge[i];
      B += sc.phibulk[i];
    }
    if (fr == 1) {  // restricted
      if (sc.GN[i] > 0) {
        nrm = sc.n[i] / sc.GN[i];
        if (S.mol_pinned[i]) {
          sc.phibulk[i] = 0;
        } else {
          sc.phibulk[i] = S.mol_N[i] * nrm;
          A += sc.phibulk[i] * S.mol_charge[i];
          B += sc.phibulk[i];
        }
      } else {
        sc.status = DS_GN;
        return;
      }
    }
    if (S.mol_pinned[i] || S.mol_tagged[i]) {
      if (sc.GN[i] > 0) {
        if (sc.n[i] == 0) sc.n[i] = 1;
        nrm = sc.n[i] / sc.GN[i];
      } else {
        nrm = 0;
      }
      sc.phibulk[i] = 0;
    }
    sc.norm[i] = nrm;
  }
  // the neutralizer's bulk fraction makes the bulk neutral
  // (System::ComputePhis src/system.cpp:2487-2516)
  if (S.charged && S.neutralizer >= 0) {
    const int nz = S.neutralizer, sv = S.solvent;
    sc.phibulk[nz] = ((B - 1.0) * S.mol_charge[sv] - A) / (S.mol_charge[nz] - S.mol_charge[sv]);
    B += sc.phibulk[nz];
    const double nrm = sc.phibulk[nz] / S.mol_N[nz];
    sc.n[nz] = nrm * sc.GN[nz];
    sc.norm[nz] = nrm;
  }
  if (S.solvent >= 0) {
    const int sv = S.solvent;
    sc.phibulk[sv] = 1.0 - B;
    if (sc.phibulk[sv] < 0) {
      sc.status = DS_PHIB;
      return;
    }
    const double nrm = sc.phibulk[sv] / S.mol_N[sv];
    sc.n[sv] = nrm * sc.GN[sv];
    sc.norm[sv] = nrm;
  }
  for (int k = 0; k < S.nseg; k++) sc.seg_phibulk[k] = 0;
  for (int i = 0; i < S.nmol; i++)
    for (int mm = S.mol_mm_begin[i]; mm < S.mol_mm_begin[i + 1]; mm++) {
      const int sg = S.mm_seg[mm];
      sc.seg_phibulk[sg] += S.frac[i][sg] * sc.phibulk[i];
    }
}

// Chemical potentials (System::CreateMu src/system.cpp:3246-3428), single
// thread.  A molecule with a tagged segment counts N - 1 segments; the tagged
// segment has no fraction and the others are scaled by N / (N - 1).
__device__ inline void scalars_mu(const DevSys &S) {
  DevScalars &sc = *S.sc;
  for (int i = 0; i < S.nmol; i++) {
    const double NA = S.mol_NF[i];
    double mu = log(NA * sc.n[i] / sc.GN[i]) + 1;
    double constant = 0;
    for (int k = 0; k < S.nmol; k++) constant += sc.phibulk[k] / S.mol_NF[k];
    mu = mu - NA * constant;
    for (int j = 0; j < S.nseg; j++) {
      const double pbA = sc.seg_phibulk[j];
      double FA = S.seg_tagged[j] ? 0.0 : S.frac[i][j];
      if (S.mol_tagged[i]) FA *= (NA + 1) / (NA);
      for (int k = 0; k < S.nseg; k++) {
        const double chi = S.chi[j * kMaxSeg + k] / 2;
        double FB = S.seg_tagged[k] ? 0.0 : S.frac[i][k];
        if (S.mol_tagged[i]) FB *= (NA + 1) / (NA);
        mu = mu - NA * chi * (pbA - FA) * (sc.seg_phibulk[k] - FB);
      }
    }
    sc.Mu[i] = mu;
  }
}

// Lane exchanges for the DIIS solve.  dpp_f64/dpp_i32 move a value across
// lanes within rows of 16 by DPP (a VALU operand modifier, no LDS round
// trip): 0xB1 = quad_perm [1,0,3,2] (lane ^ 1), 0x4E = quad_perm [2,3,0,1]
// (lane ^ 2), 0x124 / 0x128 = row_ror 4 / 8.  readlane_f64 reads one lane's
// value (the lane index must be the same on every lane).
template <int kCtrl>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, kCtrl, 0xF, 0xF, false);
}
template <int kCtrl>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_i32<kCtrl>((int)b), hi = dpp_i32<kCtrl>((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane), hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// DIIS coefficients on the device by one wavefront: c = A^-1 (1..1) / sum,
// A is the kd x kd Gram matrix (leading dimension kd, kd <= kMaxDIIS = 16).
// Lane r holds row r; Gauss-Jordan elimination with partial pivoting (pivot
// = largest |a_rk| among rows not yet used, lowest lane on ties).
// Mathematically the reference's SVD solve (SFNewton::Ax
// src/sfnewton.cpp:793-835) computes the same vector.  The pivot search is a
// max-reduction over lanes 0..15 by DPP (the winner is the unique largest
// (|a|, -lane) pair, whatever order the lanes are combined in), and the pivot
// row is read by readlane, so an elimination step has no LDS round trip.
// Returns false on NaN / zero pivot.
static_assert(kMaxDIIS <= 16, "wave_diis_coefficients: rows live in lanes 0..15");
__device__ __forceinline__ bool wave_diis_coefficients(const double *A, int n, double *C) {
  const int lane = threadIdx.x & 63;
  double row[kMaxDIIS];
  double rhs = 1.0;
  bool used = false, bad = false;
  int pc = -1;  // column this lane's row became the pivot of
#pragma unroll
  for (int c = 0; c < kMaxDIIS; c++) row[c] = (lane < n && c < n) ? A[lane * n + c] : 0.0;
  for (int k = 0; k < n; k++) {
    double ak = 0.0;
#pragma unroll
    for (int c = 0; c < kMaxDIIS; c++)
      if (c == k) ak = row[c];
    // pivot: largest |a_rk| over rows not yet used (lowest lane on ties)
    double cand = (lane < n && !used) ? fabs(ak) : -1.0;
    if (cand != cand) cand = 1e308;  // a NaN is picked and reported
    int who = lane;
    auto take = [&](double oc, int ow) {
      if (oc > cand || (oc == cand && ow < who)) {
        cand = oc;
        who = ow;
      }
    };
    take(dpp_f64<0xB1>(cand), dpp_i32<0xB1>(who));
    take(dpp_f64<0x4E>(cand), dpp_i32<0x4E>(who));
    take(dpp_f64<0x124>(cand), dpp_i32<0x124>(who));
    take(dpp_f64<0x128>(cand), dpp_i32<0x128>(who));
    who = __builtin_amdgcn_readfirstlane(who);  // lane 0's row of 16 holds every candidate
    if (lane == who) {
      used = true;
      pc = k;
    }
    const double piv = readlane_f64(ak, who);
    const double prhs = readlane_f64(rhs, who);
    if (!(piv != 0.0) || piv != piv) bad = true;
    const bool upd = lane < n && lane != who && piv != 0.0;
    const double f = upd ? ak / piv : 0.0;
#pragma unroll
    for (int c = 0; c < kMaxDIIS; c++) {
      if (c >= n) break;  // n is the same on every lane
      const double pcv = readlane_f64(row[c], who);
      if (upd) row[c] = (c == k) ? 0.0 : row[c] - f * pcv;
    }
    if (upd) rhs -= f * prhs;
  }
  double x = 0.0;
  if (lane < n && pc >= 0) {
    double d = 0.0;
#pragma unroll
    for (int c = 0; c < kMaxDIIS; c++)
      if (c == pc) d = row[c];
    x = rhs / d;
  }
  double sum = lane < n ? x : 0.0;
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
  int anybad = (bad || (lane < n && (pc < 0 || x != x))) ? 1 : 0;
  for (int off = 32; off > 0; off >>= 1) anybad |= __shfl_xor(anybad, off, 64);
  if (lane < n && pc >= 0) C[pc] = x / sum;
  return anybad == 0 && sum == sum && sum != 0.0;
}

// Record the new Gram row (SFNewton::DIIS src/sfnewton.cpp:837-867): shift
// the ring when it is full, write row/column kd-1 from dots (+1e-9 on every
// entry, as the reference does), and pack the kd x kd matrix into Ap.
// Single thread.  Returns kd.
__device__ __host__ inline int diis_gram_update(DevScalars &sc, int m, int k_diis, const double *dots, double *Ap) {
  int kd = k_diis;
  double *Aij = sc.A;
  if (kd > m) {
    kd = m;
    for (int i = 1; i < m; i++)
      for (int j = 1; j < m; j++) Aij[m * (i - 1) + j - 1] = Aij[m * i + j];
  }
  for (int i = 0; i < kd; i++) Aij[i + m * (kd - 1)] = Aij[kd - 1 + m * i] = dots[i] + 1e-9;
  for (int i = 0; i < kd; i++)
    for (int j = 0; j < kd; j++) Ap[j + kd * i] = Aij[j + m * i];
  return kd;
}

// Workgroup-parallel form of diis_gram_update for device code; every thread
// of the workgroup calls it.  Thread t owns Gram entries t, t + blockDim.x:
// it reads the entry's new value (the shifted one, or the new row/column from
// dots) before a barrier and stores it after, which equals the in-place
// row-major shift above (each source entry lies after its target and is read
// before it is overwritten).  dots must be visible to the workgroup on entry;
// sc.A and Ap are on return.  Returns kd.
__device__ inline int diis_gram_update_block(DevScalars &sc, int m, int k_diis, const double *dots, double *Ap) {
  constexpr int kPer = (kMaxDIIS * kMaxDIIS + 63) / 64;
  const int kd = k_diis > m ? m : k_diis;
  const bool shift = k_diis > m;
  double v[kPer];
#pragma unroll
  for (int r = 0; r < kPer; r++) {
    const int t = threadIdx.x + r * blockDim.x;
    v[r] = 0.0;
    if (t < m * m) {
      const int i = t / m, j = t - (t / m) * m;
      v[r] = sc.A[(shift && i < m - 1 && j < m - 1) ? t + m + 1 : t];
      if (i == kd - 1 && j < kd) v[r] = dots[j] + 1e-9;
      if (j == kd - 1 && i < kd) v[r] = dots[i] + 1e-9;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kPer; r++) {
    const int t = threadIdx.x + r * blockDim.x;
    if (t < m * m) {
      const int i = t / m, j = t - (t / m) * m;
      sc.A[t] = v[r];
      if (i < kd && j < kd) Ap[j + kd * i] = v[r];
    }
  }
  __syncthreads();
  return kd;
}

}  // namespace nscf
