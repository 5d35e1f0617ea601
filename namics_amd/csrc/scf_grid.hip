// Multi-workgroup SCF path for lattices too large for one workgroup
// (3-gradient boxes such as 96^3): each phase of a residual evaluation is a
// grid-stride kernel, each propagation step one launch over all sites, and
// the DIIS loop is driven from the host with one small readback (status and
// residual) per iteration.  The per-site arithmetic is shared with the fused
// kernel (scf_device.h).  Reductions are deterministic: per-workgroup
// partial sums in HBM, summed in workgroup order by one thread.
#include <hip/hip_runtime.h>

#include <vector>

#include "scf_device.h"
#include "scf_kernels.h"
#include "scf_spec.h"

namespace nscf {
namespace {

constexpr int kGridThreads = 256;

__device__ __forceinline__ double block_sum(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
  return t;
}

// block_sum of acc[0..ncols-1] at once, in the same order (lane tree, then
// waves 0, 1, ... in turn) behind one barrier; the workgroup's sum of column
// i goes to partial[i * gridDim.x + blockIdx.x].  red needs kMaxDIIS * waves
// doubles.
__device__ __forceinline__ void block_sums_store(const double *acc, int ncols, double *red, double *partial) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int i = 0; i < kMaxDIIS; i++) {
    if (i < ncols) {
      double v = acc[i];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      if (lane == 0) red[i * nw + wave] = v;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < ncols) {
    double t = 0;
    for (int w = 0; w < nw; w++) t += red[threadIdx.x * nw + w];
    partial[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = t;
  }
}

__device__ __forceinline__ double block_max(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) t = fmax(t, red[w]);
  return t;
}

// Sum (or max) of one column of per-workgroup partials by one workgroup:
// thread t adds entries t, t + blockDim, ... in order, then the threads'
// sums are combined by block_sum's fixed tree, so the result does not depend
// on timing.
__device__ __forceinline__ double ordered_sum(const double *partial, int nparts, int stride, int col,
                                              double *red) {
  double t = 0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) t += partial[(size_t)b * stride + col];
  return block_sum(t, red);
}

__device__ __forceinline__ double ordered_max(const double *partial, int nparts, int stride, int col,
                                              double *red) {
  double t = 0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) t = fmax(t, partial[(size_t)b * stride + col]);
  return block_max(t, red);
}

// Column sums of the column-major partial[kMaxDIIS][nparts] (entry (b, i) at
// i * nparts + b), columns 0..ncols-1, in a fixed order: thread t adds rows
// t, t + blockDim.x, ... (loads of two rows in flight, consecutive threads on
// consecutive addresses); the per-thread sums go to LDS, 16 threads per
// column add 16 of them each, and one thread per column adds those 16.  No
// lane-exchange chains: the whole reduction is two rounds of LDS reads.
// Needs blockDim.x == 16 * kMaxDIIS and red of kMaxDIIS * blockDim.x
// doubles.  The sums land in out[] and are visible to the workgroup on
// return.
static_assert(kGridThreads == 16 * kMaxDIIS, "ordered_sums: 16 threads per column");
__device__ __forceinline__ void ordered_sums(const double *partial, int nparts, int ncols, double *out,
                                             double *red) {
  double acc[kMaxDIIS];
#pragma unroll
  for (int i = 0; i < kMaxDIIS; i++) acc[i] = 0;
  const int t = threadIdx.x, nt = blockDim.x;
  int b = t;
  for (; b + nt < nparts; b += 2 * nt) {
    double a0[kMaxDIIS], a1[kMaxDIIS];
#pragma unroll
    for (int i = 0; i < kMaxDIIS; i++) {
      a0[i] = i < ncols ? partial[(size_t)i * nparts + b] : 0.0;
      a1[i] = i < ncols ? partial[(size_t)i * nparts + b + nt] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < kMaxDIIS; i++) acc[i] = (acc[i] + a0[i]) + a1[i];
  }
  if (b < nparts)
#pragma unroll
    for (int i = 0; i < kMaxDIIS; i++)
      if (i < ncols) acc[i] += partial[(size_t)i * nparts + b];
#pragma unroll
  for (int i = 0; i < kMaxDIIS; i++)
    if (i < ncols) red[i * nt + t] = acc[i];
  __syncthreads();
  const int c = t >> 4, q = t & 15, per = nt >> 4;
  double v = 0;
  if (c < ncols)
    for (int j = 0; j < per; j++) v += red[c * nt + j * 16 + q];
  __syncthreads();
  if (c < ncols) red[c * 16 + q] = v;
  __syncthreads();
  if (t < ncols) {
    double u = 0;
    for (int j = 0; j < 16; j++) u += red[t * 16 + j];
    out[t] = u;
  }
  __syncthreads();
}

#define GRID_LOOP(S) for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < n_work_sites(S); w += gridDim.x * blockDim.x)

__global__ void __launch_bounds__(kGridThreads) k_prepare(const DevSys *__restrict__ d) {
  CSys &S = uniform_sys(*d);
  GRID_LOOP(S) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    site_prepare(S, i, x, y, z);
    for (int mm = 0; mm < S.nmm; mm++) hbm(S.rho)[mm * S.M + i] = 0.0;
  }
}

__global__ void __launch_bounds__(kGridThreads) k_row_init(const DevSys *__restrict__ d, int seg) {
  CSys &S = uniform_sys(*d);
  const double *g1 = S.G1 + seg * S.M;
  GRID_LOOP(S) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    hbm(S.Gf)[i] = g1[i];
  }
}

// one forward step: Gf[dst] = W(Gf[src]) * G1 (src, dst: end-point storage
// slots, see grid_residual_body)
__global__ void __launch_bounds__(kGridThreads) k_step_fwd(const DevSys *__restrict__ d, int src, int dst,
                                                           int seg) {
  CSys &S = uniform_sys(*d);
  const int M = S.M;
  const double *__restrict__ prev = S.Gf + (size_t)src * M;
  double *__restrict__ gs = S.Gf + (size_t)dst * M;
  const double *__restrict__ g1 = S.G1 + seg * M;
  GRID_LOOP(S) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    gs[i] = stencil_prop<double>(S, prev, i, x, y, z) * g1[i];
  }
}

// One operation of a branched molecule's row schedule (BranchOpKind in
// scf_spec.h) over the sites a phase visits: rows are Gf rows of M sites.
__global__ void __launch_bounds__(kGridThreads) k_branch_op(const DevSys *__restrict__ d, int kind, int a, int b,
                                                            int c, int seg, double f) {
  CSys &S = uniform_sys(*d);
  const int M = S.M;
  gdbl *ra = hbm(S.Gf) + (size_t)a * M;
  const gdbl *rb = hbm(S.Gf) + (size_t)b * M;
  const gdbl *rc = hbm(S.Gf) + (size_t)c * M;
  const gdbl *g1 = hbm(S.G1) + (size_t)(seg < 0 ? 0 : seg) * M;
  GRID_LOOP(S) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    if (kind == BR_INIT) {
      ra[i] = g1[i];
    } else if (kind == BR_PROP) {
      const double v = stencil_prop<double>(S, S.Gf + (size_t)b * M, i, x, y, z);
      ra[i] = seg < 0 ? v : v * g1[i];
    } else if (kind == BR_COPY) {
      ra[i] = rb[i];
    } else if (kind == BR_TIMES) {
      ra[i] = rb[i] * rc[i];
    } else if (kind == BR_ONE) {
      ra[i] = 1.0;
    } else if (kind == BR_TIMES_G1) {
      ra[i] = rb[i] * g1[i];
    } else if (kind == BR_DIV) {
      ra[i] = rb[i] != 0 ? ra[i] / rb[i] : 0.0;
    } else if (kind == BR_DIV_G1) {
      ra[i] = g1[i] != 0 ? ra[i] / g1[i] : 0.0;
    } else {  // BR_ADDPHI: (f Gf) Gb, f = 1 or a dendrimer block's degeneracy
      hbm(S.rho)[(size_t)c * M + i] += f * ra[i] * rb[i];
    }
  }
}

// first backward row and its composition
__global__ void __launch_bounds__(kGridThreads) k_bwd_first(const DevSys *__restrict__ d, int slot, int seg,
                                                            int mm) {
  CSys &S = uniform_sys(*d);
  const int M = S.M;
  const double *g1 = S.G1 + seg * M;
  const double *gf = S.Gf + (size_t)slot * M;
  double *rho = S.rho + mm * M;
  GRID_LOOP(S) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    const double v = g1[i];
    hbm(S.Gb)[i] = v;
    rho[i] += gf[i] * v;
  }
}

// backward step with fused composition: Gb[dst] = W(Gb[src]) * G1,
// rho += Gf[slot] Gb[dst]
__global__ void __launch_bounds__(kGridThreads) k_step_bwd(const DevSys *__restrict__ d, int slot, int seg, int mm,
                                                           int src) {
  CSys &S = uniform_sys(*d);
  const int M = S.M;
  const double *__restrict__ gb = S.Gb + (size_t)src * M;
  double *__restrict__ gbn = S.Gb + (size_t)(1 - src) * M;
  const double *__restrict__ g1 = S.G1 + seg * M;
  const double *__restrict__ gf = S.Gf + (size_t)slot * M;
  double *__restrict__ rho = S.rho + mm * M;
  GRID_LOOP(S) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    const double v = stencil_prop<double>(S, gb, i, x, y, z) * g1[i];
    gbn[i] = v;
    rho[i] += gf[i] * v;
  }
}

// 3-gradient steps (LGrad3::propagate src/LGrad3.cpp:731-743): one thread
// per site of an x-plane, (y, z) flattened with z fastest, so every load is
// coalesced; a workgroup covers part of one plane.
__device__ __forceinline__ double stencil3(CSys &S, const gdbl *__restrict__ g, int x, int y, int z) {
  const int JX = S.JX, JY = S.JY;
  const int bx = x * JX, by = y * JY;
  double v = g[map_x(S, x - 1) * JX + by + z];
  v += g[map_x(S, x + 1) * JX + by + z];
  v += g[bx + map_y(S, y - 1) * JY + z];
  v += g[bx + map_y(S, y + 1) * JY + z];
  v += g[bx + by + map_z(S, z - 1)];
  v += g[bx + by + map_z(S, z + 1)];
  return v * (1.0 / 6.0);
}

// Workgroups are handed to the 8 XCDs round-robin by dispatch order.  The
// 3-gradient steps renumber them so that each XCD takes a contiguous run of
// x-planes: the x-1 / x+1 planes a workgroup reads are then mostly read by
// workgroups of the same XCD and hit its L2 instead of the MALL.
__device__ __forceinline__ int xcd_block_id() {
  const int nb = gridDim.x * gridDim.y;
  const int hw = blockIdx.y * gridDim.x + blockIdx.x;
  const int q = nb >> 3, r = nb & 7, c = hw & 7, k = hw >> 3;
  return (c < r ? c * (q + 1) : r * (q + 1) + (c - r) * q) + k;
}

__device__ __forceinline__ void xcd_plane_block(int &xplane, int &bx) {
  const int L = xcd_block_id();
  xplane = L / gridDim.x;
  bx = L - xplane * gridDim.x;
}

__global__ void __launch_bounds__(kGridThreads) k_step3_fwd(const DevSys *__restrict__ d, int src, int dst,
                                                            int seg) {
  CSys &S = uniform_sys(*d);
  const int MZ = S.MZ, M = S.M;
  int xp, bx;
  xcd_plane_block(xp, bx);
  const int p = bx * kGridThreads + threadIdx.x;
  if (p >= S.MY * MZ) return;
  const int yy = p / MZ;
  const int x = xp + 1, y = yy + 1, z = p - yy * MZ + 1;
  const int i = x * S.JX + y * S.JY + z;
  const gdbl *prev = hbm(S.Gf) + (size_t)src * M;
  hbm(S.Gf)[(size_t)dst * M + i] = stencil3(S, prev, x, y, z) * hbm(S.G1)[(size_t)seg * M + i];
}

__global__ void __launch_bounds__(kGridThreads) k_step3_bwd(const DevSys *__restrict__ d, int slot, int seg, int mm,
                                                            int src) {
  CSys &S = uniform_sys(*d);
  const int MZ = S.MZ, M = S.M;
  int xp, bx;
  xcd_plane_block(xp, bx);
  const int p = bx * kGridThreads + threadIdx.x;
  if (p >= S.MY * MZ) return;
  const int yy = p / MZ;
  const int x = xp + 1, y = yy + 1, z = p - yy * MZ + 1;
  const int i = x * S.JX + y * S.JY + z;
  const gdbl *gb = hbm(S.Gb) + (size_t)src * M;
  const double v = stencil3(S, gb, x, y, z) * hbm(S.G1)[(size_t)seg * M + i];
  hbm(S.Gb)[(size_t)(1 - src) * M + i] = v;
  gdbl *rho = hbm(S.rho) + (size_t)mm * M;
  rho[i] += hbm(S.Gf)[(size_t)slot * M + i] * v;
}

// The forward pass's first and last 6-point steps with the neighbouring
// one-row phases folded in: the first step reads row 0 from the segment
// weights and stores it (k_row_init), the last one also starts the
// backward pass (k_bwd_first: Gb = G1, rho += G(N-1) G1).  Same arithmetic
// as the separate launches, one launch less each.
template <bool kFirst, bool kLast>
__global__ void __launch_bounds__(kGridThreads) k_step3_fwd_x(const DevSys *__restrict__ d, int src, int dst,
                                                              int seg, int seg0, int slot0, int mm) {
  CSys &S = uniform_sys(*d);
  const int MZ = S.MZ, M = S.M;
  int xp, bx;
  xcd_plane_block(xp, bx);
  const int p = bx * kGridThreads + threadIdx.x;
  if (p >= S.MY * MZ) return;
  const int yy = p / MZ;
  const int x = xp + 1, y = yy + 1, z = p - yy * MZ + 1;
  const int i = x * S.JX + y * S.JY + z;
  const gdbl *prev = kFirst ? hbm(S.G1) + (size_t)seg0 * M : hbm(S.Gf) + (size_t)src * M;
  const double g = hbm(S.G1)[(size_t)seg * M + i];
  const double v = stencil3(S, prev, x, y, z) * g;
  hbm(S.Gf)[(size_t)dst * M + i] = v;
  if (kFirst) hbm(S.Gf)[(size_t)slot0 * M + i] = prev[i];
  if (kLast) {
    hbm(S.Gb)[i] = g;
    hbm(S.rho)[(size_t)mm * M + i] += v * g;
  }
}

// ---- FP32 end points on the 6-point 3-gradient steps (mixed precision).
// Rows hold G(s) e^{c (s+1)} in FP32 (forward) and G(s) e^{c (N-s)}
// (backward), where c = -ln(GN / V) / N is the molecule's gauge shift from
// the latest GN: the segment weights are G1' = G1 e^c (k_g1_shift32), so
// the shifted end points average about one per site whatever N and the
// fields.  GN and the composition are taken back to FP64 and unshifted
// (GN = GN' e^{-cN}; rho += Gf' Gb' e^{-c(N+1)}).  The steps move 12 B per
// site (forward) and 28 B (backward, rho in FP64) instead of 24 and 48.
typedef __attribute__((address_space(1))) float gflt;
__device__ __forceinline__ gflt *hbmf(float *p) { return (gflt *)p; }
__device__ __forceinline__ const gflt *hbmf(const float *p) { return (const gflt *)p; }

__device__ __forceinline__ float stencil3f(CSys &S, const gflt *__restrict__ g, int x, int y, int z) {
  const int JX = S.JX, JY = S.JY;
  const int bx = x * JX, by = y * JY;
  float v = g[map_x(S, x - 1) * JX + by + z];
  v += g[map_x(S, x + 1) * JX + by + z];
  v += g[bx + map_y(S, y - 1) * JY + z];
  v += g[bx + map_y(S, y + 1) * JY + z];
  v += g[bx + by + map_z(S, z - 1)];
  v += g[bx + by + map_z(S, z + 1)];
  return v * (1.0f / 6.0f);
}

// G1' = G1 e^c of every segment, for molecule mo's chain
__global__ void __launch_bounds__(kGridThreads) k_g1_shift32(const DevSys *__restrict__ d, int mo) {
  CSys &S = uniform_sys(*d);
  const double es = exp(S.sc->lnshift[mo]);
  const int n = S.nseg * S.M;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    hbmf(S.G1f)[i] = (float)(hbm(S.G1)[i] * es);
}

__global__ void __launch_bounds__(kGridThreads) k_row_init32(const DevSys *__restrict__ d, int seg) {
  CSys &S = uniform_sys(*d);
  const gflt *g1 = hbmf(S.G1f) + (size_t)seg * S.M;
  gflt *gf = hbmf((float *)S.Gf);
  GRID_LOOP(S) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    gf[i] = g1[i];
  }
}

__global__ void __launch_bounds__(kGridThreads) k_step3_fwd32(const DevSys *__restrict__ d, int src, int dst,
                                                              int seg) {
  CSys &S = uniform_sys(*d);
  const int MZ = S.MZ, M = S.M;
  int xp, bx;
  xcd_plane_block(xp, bx);
  const int p = bx * kGridThreads + threadIdx.x;
  if (p >= S.MY * MZ) return;
  const int yy = p / MZ;
  const int x = xp + 1, y = yy + 1, z = p - yy * MZ + 1;
  const int i = x * S.JX + y * S.JY + z;
  const gflt *prev = hbmf((const float *)S.Gf) + (size_t)src * M;
  hbmf((float *)S.Gf)[(size_t)dst * M + i] = stencil3f(S, prev, x, y, z) * hbmf(S.G1f)[(size_t)seg * M + i];
}

// k_step3_fwd_x with FP32 end points
template <bool kFirst, bool kLast>
__global__ void __launch_bounds__(kGridThreads) k_step3_fwd32_x(const DevSys *__restrict__ d, int src, int dst,
                                                                int seg, int seg0, int slot0, int mm, int mo) {
  CSys &S = uniform_sys(*d);
  const int MZ = S.MZ, M = S.M;
  int xp, bx;
  xcd_plane_block(xp, bx);
  const int p = bx * kGridThreads + threadIdx.x;
  if (p >= S.MY * MZ) return;
  const int yy = p / MZ;
  const int x = xp + 1, y = yy + 1, z = p - yy * MZ + 1;
  const int i = x * S.JX + y * S.JY + z;
  const gflt *prev = kFirst ? hbmf(S.G1f) + (size_t)seg0 * M : hbmf((const float *)S.Gf) + (size_t)src * M;
  const float g = hbmf(S.G1f)[(size_t)seg * M + i];
  const float v = stencil3f(S, prev, x, y, z) * g;
  hbmf((float *)S.Gf)[(size_t)dst * M + i] = v;
  if (kFirst) hbmf((float *)S.Gf)[(size_t)slot0 * M + i] = prev[i];
  if (kLast) {
    const double unshift = exp(-S.sc->lnshift[mo] * (S.mol_N[mo] + 1));
    hbmf((float *)S.Gb)[i] = g;
    hbm(S.rho)[(size_t)mm * M + i] += (double)v * (double)g * unshift;
  }
}

// first backward row (G1') and its composition, unshifted
__global__ void __launch_bounds__(kGridThreads) k_bwd_first32(const DevSys *__restrict__ d, int slot, int seg,
                                                              int mm, int mo) {
  CSys &S = uniform_sys(*d);
  const int M = S.M;
  const double unshift = exp(-S.sc->lnshift[mo] * (S.mol_N[mo] + 1));
  const gflt *g1 = hbmf(S.G1f) + (size_t)seg * M;
  const gflt *gf = hbmf((const float *)S.Gf) + (size_t)slot * M;
  gdbl *rho = hbm(S.rho) + (size_t)mm * M;
  GRID_LOOP(S) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    const float v = g1[i];
    hbmf((float *)S.Gb)[i] = v;
    rho[i] += (double)gf[i] * (double)v * unshift;
  }
}

__global__ void __launch_bounds__(kGridThreads) k_step3_bwd32(const DevSys *__restrict__ d, int slot, int seg,
                                                              int mm, int src, int mo) {
  CSys &S = uniform_sys(*d);
  const int MZ = S.MZ, M = S.M;
  int xp, bx;
  xcd_plane_block(xp, bx);
  const int p = bx * kGridThreads + threadIdx.x;
  if (p >= S.MY * MZ) return;
  const int yy = p / MZ;
  const int x = xp + 1, y = yy + 1, z = p - yy * MZ + 1;
  const int i = x * S.JX + y * S.JY + z;
  const double unshift = exp(-S.sc->lnshift[mo] * (S.mol_N[mo] + 1));
  const gflt *gb = hbmf((const float *)S.Gb) + (size_t)src * M;
  const float v = stencil3f(S, gb, x, y, z) * hbmf(S.G1f)[(size_t)seg * M + i];
  hbmf((float *)S.Gb)[(size_t)(1 - src) * M + i] = v;
  gdbl *rho = hbm(S.rho) + (size_t)mm * M;
  rho[i] += (double)hbmf((const float *)S.Gf)[(size_t)slot * M + i] * (double)v * unshift;
}

__global__ void __launch_bounds__(kGridThreads) k_gn_partial32(const DevSys *__restrict__ d, int slot,
                                                               double *partial) {
  CSys &S = uniform_sys(*d);
  __shared__ double red[kGridThreads / 64];
  const gflt *gl = hbmf((const float *)S.Gf) + (size_t)slot * S.M;
  double acc = 0;
  GRID_LOOP(S) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    if (is_interior(S, x, y, z)) acc += site_weight(S, i) * (double)gl[i];
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// gauge shifts for the next evaluation, c = -ln(GN / V) / N from this GN
// (0 where GN is not positive and finite; such an FP32 evaluation fails)
__global__ void k_gauge_shifts(const DevSys *__restrict__ d) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const DevSys &S = *d;
  DevScalars &sc = *S.sc;
  for (int mo = 0; mo < S.nmol; mo++) {
    const double gn = sc.GN[mo];
    sc.lnshift[mo] = (gn > 0 && isfinite(gn)) ? -log(gn / S.volume) / S.mol_N[mo] : 0.0;
  }
}

__global__ void k_diis_restart(const DevSys *__restrict__ d) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    d->sc->k_diis = 0;
    d->sc->status = DS_OK;
  }
}

__global__ void __launch_bounds__(kGridThreads) k_gn_partial(const DevSys *__restrict__ d, int slot,
                                                             double *partial) {
  CSys &S = uniform_sys(*d);
  __shared__ double red[kGridThreads / 64];
  const double *gl = S.Gf + (size_t)slot * S.M;
  double acc = 0;
  GRID_LOOP(S) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    if (is_interior(S, x, y, z)) acc += site_weight(S, i) * gl[i];
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// GN of molecule mo from the per-workgroup partials
__global__ void __launch_bounds__(kGridThreads) k_finalize_gn(const DevSys *__restrict__ d, const double *partial,
                                                               int nparts, int mo) {
  __shared__ double red[kGridThreads / 64];
  const double t = ordered_sum(partial, nparts, 1, 0, red);
  if (threadIdx.x == 0) d->sc->GN[mo] = t;
}

// the same from FP32 end points: GN = GN' e^{-cN}
__global__ void __launch_bounds__(kGridThreads) k_finalize_gn32(const DevSys *__restrict__ d,
                                                                 const double *partial, int nparts, int mo) {
  __shared__ double red[kGridThreads / 64];
  const double t = ordered_sum(partial, nparts, 1, 0, red);
  if (threadIdx.x == 0) d->sc->GN[mo] = t * exp(-d->sc->lnshift[mo] * d->mol_N[mo]);
}

__global__ void k_norms(const DevSys *__restrict__ d) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const DevSys &S = *d;
  DevScalars &sc = *S.sc;
  for (int i = 0; i < S.nmol; i++) {
    if (S.mol_freedom[i] == 0) sc.phibulk[i] = S.mol_phibulk_in[i];
    if (S.mol_freedom[i] == 1) sc.n[i] = S.mol_n_in[i];
  }
  scalars_norms(S);
}

__global__ void __launch_bounds__(kGridThreads) k_densities(const DevSys *__restrict__ d) {
  CSys &S = uniform_sys(*d);
  if (S.sc->status != DS_OK) return;
  GRID_LOOP(S) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    site_densities(S, i);
  }
}

__global__ void __launch_bounds__(kGridThreads) k_residual_partial(const DevSys *__restrict__ d, double *partial) {
  CSys &S = uniform_sys(*d);
  __shared__ double red[kGridThreads / 64];
  if (S.sc->status != DS_OK) return;
  double acc = 0;
  GRID_LOOP(S) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    if (!is_interior(S, x, y, z)) continue;
    const double g2 = site_residual(S, i, x, y, z);
    if (S.max_g) {
      for (int it = 0; it < S.nit; it++) acc = fmax(acc, fabs(hbm(S.g)[it * S.M + i]));
    } else {
      acc += g2;
    }
  }
  acc = S.max_g ? block_max(acc, red) : block_sum(acc, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ void __launch_bounds__(kGridThreads) k_finalize_residual(const DevSys *__restrict__ d,
                                                                     const double *partial, int nparts) {
  __shared__ double red[kGridThreads / 64];
  CSys &S = uniform_sys(*d);
  if (S.sc->status != DS_OK) return;
  const double t = S.max_g ? ordered_max(partial, nparts, 1, 0, red) : ordered_sum(partial, nparts, 1, 0, red);
  if (threadIdx.x == 0) S.sc->residual = S.max_g ? t : sqrt(t);
}

__global__ void __launch_bounds__(kGridThreads) k_thermo_partial(const DevSys *__restrict__ d, double *partial) {
  CSys &S = uniform_sys(*d);
  __shared__ double red[kGridThreads / 64];
  const DevScalars &sc = *S.sc;
  double logc[kMaxMol];
#pragma unroll
  for (int mo = 0; mo < kMaxMol; mo++) {
    logc[mo] = 0.0;
    if (mo < S.nmol) {
      const double N = S.mol_NF[mo];  // a tagged segment does not count
      logc[mo] = log(N * sc.n[mo] / sc.GN[mo]) / N;
    }
  }
  double f = 0, gp = 0, th[kMaxMol];
#pragma unroll
  for (int mo = 0; mo < kMaxMol; mo++) th[mo] = 0;
  GRID_LOOP(S) {
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
      if (mo < S.nmol) th[mo] += wt * hbm(S.mol_phitot)[mo * S.M + i];
  }
  const int nc = 2 + S.nmol;
  f = block_sum(f, red);
  gp = block_sum(gp, red);
  if (threadIdx.x == 0) {
    partial[blockIdx.x * nc + 0] = f;
    partial[blockIdx.x * nc + 1] = gp;
  }
#pragma unroll
  for (int mo = 0; mo < kMaxMol; mo++) {
    if (mo < S.nmol) {
      const double t = block_sum(th[mo], red);
      if (threadIdx.x == 0) partial[blockIdx.x * nc + 2 + mo] = t;
    }
  }
}

__global__ void __launch_bounds__(kGridThreads) k_finalize_thermo(const DevSys *__restrict__ d,
                                                                   const double *partial, int nparts) {
  __shared__ double red[kGridThreads / 64];
  const DevSys &S = *d;
  DevScalars &sc = *S.sc;
  const int nc = 2 + S.nmol;
  const double f = ordered_sum(partial, nparts, nc, 0, red);
  const double gp = ordered_sum(partial, nparts, nc, 1, red);
  if (threadIdx.x == 0) {
    sc.F = f;
    sc.GP = gp;
  }
  for (int mo = 0; mo < S.nmol; mo++) {
    const double t = ordered_sum(partial, nparts, nc, 2 + mo, red);
    if (threadIdx.x == 0) sc.theta[mo] = t;
  }
  if (threadIdx.x == 0) scalars_mu(S);
}

// ---- DIIS element-wise parts (nvar = nit * M entries, all sites)
#define VAR_LOOP(S) for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < S.nit * S.M; v += gridDim.x * blockDim.x)

__global__ void __launch_bounds__(kGridThreads) k_copy_x0(const DevSys *__restrict__ d) {
  CSys &S = uniform_sys(*d);
  VAR_LOOP(S) hbm(S.x0)[v] = hbm(S.x)[v];
}

__global__ void __launch_bounds__(kGridThreads) k_axpy_first(const DevSys *__restrict__ d) {
  CSys &S = uniform_sys(*d);
  const double dmax = S.deltamax;
  VAR_LOOP(S) {
    const double xo = hbm(S.x)[v];
    const double xv = xo + (-dmax) * hbm(S.g)[v];
    hbm(S.x)[v] = xv;
    hbm(S.xx0)[v] = xv - xo;
    hbm(S.xR)[v] = xv;
  }
}

// x -= dmax g; xR[k] = x; xx0[k] = x_new - x_old; partial dots <xx0[posi_i], xx0[k]>.
// The residual never changes x, so x_old is the iteration's x0.
__global__ void __launch_bounds__(kGridThreads) k_diis_step(const DevSys *__restrict__ d, int k, int kd,
                                                             double *partial) {
  CSys &S = uniform_sys(*d);
  __shared__ double red[kMaxDIIS * (kGridThreads / 64)];
  if (S.sc->status != DS_OK) return;
  const int nvar = S.nit * S.M, m = S.m;
  const double dmax = S.deltamax;
  gdbl *xRk = hbm(S.xR) + (size_t)k * nvar, *xxk = hbm(S.xx0) + (size_t)k * nvar;
  double acc[kMaxDIIS];
#pragma unroll
  for (int i = 0; i < kMaxDIIS; i++) acc[i] = 0;
  VAR_LOOP(S) {
    const double xo = hbm(S.x)[v];
    const double xv = xo + (-dmax) * hbm(S.g)[v];
    hbm(S.x)[v] = xv;
    xRk[v] = xv;
    const double e = xv - xo;
    xxk[v] = e;
#pragma unroll
    for (int i = 0; i < kMaxDIIS; i++) {
      if (i < kd) {
        int posi = k - kd + 1 + i;
        if (posi < 0) posi += m;
        const double a = posi == k ? e : hbm(S.xx0)[(size_t)posi * nvar + v];
        acc[i] += a * e;
      }
    }
  }
  block_sums_store(acc, kd, red, partial);
}

__global__ void __launch_bounds__(kGridThreads) k_diis_solve(const DevSys *__restrict__ d, const double *partial,
                                                              int nparts, int k_diis, int kd) {
  __shared__ double dots[kMaxDIIS];
  __shared__ double Ap[kMaxDIIS * kMaxDIIS];
  __shared__ double red[kMaxDIIS * kGridThreads];
  const DevSys &S = *d;
  DevScalars &sc = *S.sc;
  const int st = sc.status;
  if (st != DS_OK) return;
  ordered_sums(partial, nparts, kd, dots, red);
  diis_gram_update_block(sc, S.m, k_diis, dots, Ap);
  if (threadIdx.x < 64) {
    const bool ok = wave_diis_coefficients(Ap, kd, sc.C);
    if (threadIdx.x == 0 && !ok) sc.status = DS_NAN;
  }
}

__global__ void __launch_bounds__(kGridThreads) k_diis_combine(const DevSys *__restrict__ d, int k, int kd) {
  CSys &S = uniform_sys(*d);
  if (S.sc->status != DS_OK) return;
  const int nvar = S.nit * S.M, m = S.m;
  VAR_LOOP(S) {
    int posi = k - kd + 1;
    if (posi < 0) posi += m;
    double xv = 0.0 + S.sc->C[0] * hbm(S.xR)[(size_t)posi * nvar + v];
    for (int i = 1; i < kd; i++) {
      posi = k - kd + 1 + i;
      if (posi < 0) posi += m;
      xv += S.sc->C[i] * hbm(S.xR)[(size_t)posi * nvar + v];
    }
    hbm(S.x)[v] = xv;
  }
}

__global__ void k_reset_status(const DevSys *__restrict__ d) {
  if (threadIdx.x == 0 && blockIdx.x == 0) d->sc->status = DS_OK;
}

__global__ void k_grid_init(const DevSys *__restrict__ d) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    d->sc->status = DS_OK;
    d->sc->iterations = 0;
    d->sc->k_diis = 0;
  }
}

// start of one DIIS iteration: iterations++, k_diis++ (SFNewton::iterate_DIIS),
// and a clear status for the evaluation that follows
__global__ void k_iter_begin(const DevSys *__restrict__ d) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (d->sc->status == DS_OK) {
      d->sc->iterations++;
      d->sc->k_diis++;
    }
    d->sc->status = DS_OK;
  }
}

template <typename Sys>
__device__ __forceinline__ void diis_indices(const Sys &S, int &k, int &kd, int &k_diis) {
  k = S.sc->iterations % S.m;
  k_diis = S.sc->k_diis;
  kd = k_diis > S.m ? S.m : k_diis;
}

__global__ void __launch_bounds__(kGridThreads) k_diis_step_dev(const DevSys *__restrict__ d, double *partial) {
  CSys &S = uniform_sys(*d);
  __shared__ double red[kMaxDIIS * (kGridThreads / 64)];
  if (S.sc->status != DS_OK) return;
  int k, kd, k_diis;
  diis_indices(S, k, kd, k_diis);
  const int nvar = S.nit * S.M, m = S.m;
  const double dmax = S.deltamax;
  gdbl *xRk = hbm(S.xR) + (size_t)k * nvar, *xxk = hbm(S.xx0) + (size_t)k * nvar;
  double acc[kMaxDIIS];
#pragma unroll
  for (int i = 0; i < kMaxDIIS; i++) acc[i] = 0;
  VAR_LOOP(S) {
    const double xo = hbm(S.x)[v];
    const double xv = xo + (-dmax) * hbm(S.g)[v];
    hbm(S.x)[v] = xv;
    xRk[v] = xv;
    const double e = xv - xo;
    xxk[v] = e;
#pragma unroll
    for (int i = 0; i < kMaxDIIS; i++) {
      if (i < kd) {
        int posi = k - kd + 1 + i;
        if (posi < 0) posi += m;
        const double a = posi == k ? e : hbm(S.xx0)[(size_t)posi * nvar + v];
        acc[i] += a * e;
      }
    }
  }
  block_sums_store(acc, kd, red, partial);
}

__global__ void __launch_bounds__(kGridThreads) k_diis_solve_dev(const DevSys *__restrict__ d,
                                                                  const double *partial, int nparts) {
  __shared__ double dots[kMaxDIIS];
  __shared__ double Ap[kMaxDIIS * kMaxDIIS];
  __shared__ double red[kMaxDIIS * kGridThreads];
  const DevSys &S = *d;
  DevScalars &sc = *S.sc;
  if (sc.status != DS_OK) return;
  int k, kd, k_diis;
  diis_indices(S, k, kd, k_diis);
  ordered_sums(partial, nparts, kd, dots, red);
  diis_gram_update_block(sc, S.m, k_diis, dots, Ap);
  if (threadIdx.x < 64) {
    const bool ok = wave_diis_coefficients(Ap, kd, sc.C);
    if (threadIdx.x == 0 && !ok) sc.status = DS_NAN;
  }
}

__global__ void __launch_bounds__(kGridThreads) k_diis_combine_dev(const DevSys *__restrict__ d) {
  CSys &S = uniform_sys(*d);
  if (S.sc->status != DS_OK) return;
  int k, kd, k_diis;
  diis_indices(S, k, kd, k_diis);
  const int nvar = S.nit * S.M, m = S.m;
  VAR_LOOP(S) {
    int posi = k - kd + 1;
    if (posi < 0) posi += m;
    double xv = 0.0 + S.sc->C[0] * hbm(S.xR)[(size_t)posi * nvar + v];
    for (int i = 1; i < kd; i++) {
      posi = k - kd + 1 + i;
      if (posi < 0) posi += m;
      xv += S.sc->C[i] * hbm(S.xR)[(size_t)posi * nvar + v];
    }
    hbm(S.x)[v] = xv;
  }
}

}  // namespace

hipError_t grid_residual(const DevSys &h, DevSys *d, double *partial, int nb, hipStream_t st) {
  hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, st, d);
  return grid_residual_body(h, d, partial, nb, st);
}

// End-point storage of the grid path.  With every row kept (ckpt <= 1) row
// s of the forward recursion is Gf slot s.  With checkpoints every K rows
// (the reference's save_memory, src/molecule.cpp:2300-2446) the slots hold
// rows 0, K, 2K, ... (slot s/K) and one buffer of K-1 rows (slots nck ..
// nck+K-2) for the rows between two checkpoints: the forward pass writes
// each such row into the buffer, so after it the buffer holds the last
// segment; the backward pass recomputes the rows of every earlier segment
// from its checkpoint, with the forward arithmetic, just before it composes
// them.  Densities and partition functions are bit-identical to full
// storage.
struct RowSlots {
  int K, nck;
  int operator()(int r) const { return K <= 1 ? r : (r % K == 0 ? r / K : nck + r % K - 1); }
};

hipError_t grid_residual_body(const DevSys &h, DevSys *d, double *partial, int nb, hipStream_t st, bool fp32) {
  const dim3 B(kGridThreads), G(nb);
  const dim3 G3((h.MY * h.MZ + kGridThreads - 1) / kGridThreads, h.MX);
  // the plane-mapped 6-point kernels serve the default simple-cubic 3-D
  // stencil; the 26-neighbour stencils go through the generic step
  const bool six = h.gradients == 3 && h.stencil_full && h.lattice_type == 0;
  hipLaunchKernelGGL(k_prepare, G, B, 0, st, d);
  for (int mo = 0; mo < h.nmol; mo++) {
    if (h.branch[mo]) {
      // a branched molecule: its schedule, GN over the row it names
      const BranchPlan &plan = *(const BranchPlan *)h.branch[mo];
      for (const BranchOp &o : plan.forward)
        hipLaunchKernelGGL(k_branch_op, G, B, 0, st, d, o.kind, o.a, o.b, o.c, o.seg, o.f);
      hipLaunchKernelGGL(k_gn_partial, G, B, 0, st, d, plan.gn_row, partial);
      hipLaunchKernelGGL(k_finalize_gn, dim3(1), dim3(kGridThreads), 0, st, d, partial, nb, mo);
      for (const BranchOp &o : plan.backward)
        hipLaunchKernelGGL(k_branch_op, G, B, 0, st, d, o.kind, o.a, o.b, o.c, o.seg, o.f);
      continue;
    }
    if (h.mol_markov[mo] != 1) return hipErrorInvalidValue;
    if (fp32 && !(six && h.G1f)) return hipErrorInvalidValue;
    const int b0 = h.mol_block_begin[mo], b1 = h.mol_block_begin[mo + 1], N = h.mol_N[mo];
    // block (segment, molecule-segment pair) of every row of the chain
    std::vector<int> rb(N);
    for (int b = b0, r = 0; b < b1; b++)
      for (int j = 0; j < h.block_n[b]; j++) rb[r++] = b;
    const RowSlots loc{h.ckpt, h.ckpt > 1 ? (N + h.ckpt - 1) / h.ckpt : N};
    auto fwd = [&](int r) {
      const int seg = h.block_seg[rb[r]];
      if (fp32)
        hipLaunchKernelGGL(k_step3_fwd32, G3, B, 0, st, d, loc(r - 1), loc(r), seg);
      else if (six)
        hipLaunchKernelGGL(k_step3_fwd, G3, B, 0, st, d, loc(r - 1), loc(r), seg);
      else
        hipLaunchKernelGGL(k_step_fwd, G, B, 0, st, d, loc(r - 1), loc(r), seg);
    };
    // 6-point chains fold the row init into the first step and the first
    // backward row into the last one
    const bool fold = six && N >= 2;
    const int seg0 = h.block_seg[b0], segN = h.block_seg[rb[N - 1]], mmN = h.block_mm[rb[N - 1]];
    auto fwd_folded = [&](int r) {
      const bool first = r == 1, last = r == N - 1;
      const int seg = h.block_seg[rb[r]];
      if (fp32) {
        if (first && last)
          hipLaunchKernelGGL((k_step3_fwd32_x<true, true>), G3, B, 0, st, d, loc(r - 1), loc(r), seg, seg0, loc(0),
                             mmN, mo);
        else if (first)
          hipLaunchKernelGGL((k_step3_fwd32_x<true, false>), G3, B, 0, st, d, loc(r - 1), loc(r), seg, seg0, loc(0),
                             mmN, mo);
        else if (last)
          hipLaunchKernelGGL((k_step3_fwd32_x<false, true>), G3, B, 0, st, d, loc(r - 1), loc(r), seg, seg0,
                             loc(0), mmN, mo);
        else
          fwd(r);
      } else {
        if (first && last)
          hipLaunchKernelGGL((k_step3_fwd_x<true, true>), G3, B, 0, st, d, loc(r - 1), loc(r), seg, seg0, loc(0),
                             mmN);
        else if (first)
          hipLaunchKernelGGL((k_step3_fwd_x<true, false>), G3, B, 0, st, d, loc(r - 1), loc(r), seg, seg0, loc(0),
                             mmN);
        else if (last)
          hipLaunchKernelGGL((k_step3_fwd_x<false, true>), G3, B, 0, st, d, loc(r - 1), loc(r), seg, seg0, loc(0),
                             mmN);
        else
          fwd(r);
      }
    };
    if (fp32) hipLaunchKernelGGL(k_g1_shift32, G, B, 0, st, d, mo);
    if (!fold) {
      if (fp32)
        hipLaunchKernelGGL(k_row_init32, G, B, 0, st, d, seg0);
      else
        hipLaunchKernelGGL(k_row_init, G, B, 0, st, d, seg0);
    }
    for (int r = 1; r < N; r++) {
      if (fold)
        fwd_folded(r);
      else
        fwd(r);
    }
    if (fp32) {
      hipLaunchKernelGGL(k_gn_partial32, G, B, 0, st, d, loc(N - 1), partial);
      hipLaunchKernelGGL(k_finalize_gn32, dim3(1), dim3(kGridThreads), 0, st, d, partial, nb, mo);
      if (!fold) hipLaunchKernelGGL(k_bwd_first32, G, B, 0, st, d, loc(N - 1), segN, mmN, mo);
    } else {
      hipLaunchKernelGGL(k_gn_partial, G, B, 0, st, d, loc(N - 1), partial);
      hipLaunchKernelGGL(k_finalize_gn, dim3(1), dim3(kGridThreads), 0, st, d, partial, nb, mo);
      if (!fold) hipLaunchKernelGGL(k_bwd_first, G, B, 0, st, d, loc(N - 1), segN, mmN);
    }
    int src = 0, seg_in_buffer = h.ckpt > 1 ? (N - 1) / h.ckpt : -1;
    for (int s = N - 2; s >= 0; s--) {
      if (h.ckpt > 1 && s / h.ckpt != seg_in_buffer) {
        // rows of segment q (qK+1 .. qK+K-1) back into the buffer
        seg_in_buffer = s / h.ckpt;
        const int a0 = seg_in_buffer * h.ckpt;
        for (int r = a0 + 1; r < a0 + h.ckpt && r < N; r++) fwd(r);
      }
      const int b = rb[s];
      if (fp32)
        hipLaunchKernelGGL(k_step3_bwd32, G3, B, 0, st, d, loc(s), h.block_seg[b], h.block_mm[b], src, mo);
      else if (six)
        hipLaunchKernelGGL(k_step3_bwd, G3, B, 0, st, d, loc(s), h.block_seg[b], h.block_mm[b], src);
      else
        hipLaunchKernelGGL(k_step_bwd, G, B, 0, st, d, loc(s), h.block_seg[b], h.block_mm[b], src);
      src = 1 - src;
    }
  }
  // a mixed-precision solve keeps its gauge shifts current in both precisions
  if (h.precision != 0) hipLaunchKernelGGL(k_gauge_shifts, dim3(1), dim3(1), 0, st, d);
  hipLaunchKernelGGL(k_norms, dim3(1), dim3(64), 0, st, d);
  hipLaunchKernelGGL(k_densities, G, B, 0, st, d);
  hipLaunchKernelGGL(k_residual_partial, G, B, 0, st, d, partial);
  hipLaunchKernelGGL(k_finalize_residual, dim3(1), dim3(kGridThreads), 0, st, d, partial, nb);
  return hipGetLastError();
}

hipError_t grid_thermo(const DevSys &h, DevSys *d, double *partial, int nb, hipStream_t st) {
  (void)h;
  hipLaunchKernelGGL(k_thermo_partial, dim3(nb), dim3(kGridThreads), 0, st, d, partial);
  hipLaunchKernelGGL(k_finalize_thermo, dim3(1), dim3(kGridThreads), 0, st, d, partial, nb);
  return hipGetLastError();
}

hipError_t grid_init(DevSys *d, hipStream_t st) {
  hipLaunchKernelGGL(k_grid_init, dim3(1), dim3(1), 0, st, d);
  return hipGetLastError();
}

// One DIIS iteration of the grid path with every index on the device, so the
// launch sequence is identical from one iteration to the next and can be
// captured once into a hipGraph.
hipError_t grid_iteration(const DevSys &h, DevSys *d, double *partial, int nb, hipStream_t st, bool fp32) {
  hipLaunchKernelGGL(k_iter_begin, dim3(1), dim3(1), 0, st, d);
  hipError_t e = grid_residual_body(h, d, partial, nb, st, fp32);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_diis_step_dev, dim3(nb), dim3(kGridThreads), 0, st, d, partial);
  hipLaunchKernelGGL(k_diis_solve_dev, dim3(1), dim3(kGridThreads), 0, st, d, partial, nb);
  hipLaunchKernelGGL(k_diis_combine_dev, dim3(nb), dim3(kGridThreads), 0, st, d);
  return hipGetLastError();
}

hipError_t grid_diis_restart(DevSys *d, hipStream_t st) {
  hipLaunchKernelGGL(k_diis_restart, dim3(1), dim3(1), 0, st, d);
  return hipGetLastError();
}

hipError_t grid_copy_x0(const DevSys &h, DevSys *d, int nb, hipStream_t st) {
  (void)h;
  hipLaunchKernelGGL(k_copy_x0, dim3(nb), dim3(kGridThreads), 0, st, d);
  return hipGetLastError();
}

hipError_t grid_axpy_first(const DevSys &h, DevSys *d, int nb, hipStream_t st) {
  (void)h;
  hipLaunchKernelGGL(k_axpy_first, dim3(nb), dim3(kGridThreads), 0, st, d);
  return hipGetLastError();
}

hipError_t grid_diis_step(const DevSys &h, DevSys *d, double *partial, int nb, int k, int k_diis,
                          hipStream_t st) {
  const int kd = k_diis > h.m ? h.m : k_diis;
  hipLaunchKernelGGL(k_diis_step, dim3(nb), dim3(kGridThreads), 0, st, d, k, kd, partial);
  hipLaunchKernelGGL(k_diis_solve, dim3(1), dim3(kGridThreads), 0, st, d, partial, nb, k_diis, kd);
  return hipGetLastError();
}

hipError_t grid_diis_combine(const DevSys &h, DevSys *d, int nb, int k, int k_diis, hipStream_t st) {
  const int kd = k_diis > h.m ? h.m : k_diis;
  hipLaunchKernelGGL(k_diis_combine, dim3(nb), dim3(kGridThreads), 0, st, d, k, kd);
  return hipGetLastError();
}

// ---- unit-test propagation: out[0] = G1, out[s] = W(out[s-1]) * G1
namespace {
__global__ void __launch_bounds__(kGridThreads) k_chain_step(const DevSys *__restrict__ d, const double *G1,
                                                             const double *prev, double *out) {
  CSys &S = uniform_sys(*d);
  if (S.gradients == 1) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < S.M; i += gridDim.x * blockDim.x)
      out[i] = stencil_prop<double>(S, prev, i, i, 0, 0) * G1[i];
    return;
  }
  GRID_LOOP(S) {
    int x, y, z;
    const int i = work_site(S, w, x, y, z);
    out[i] = stencil_prop<double>(S, prev, i, x, y, z) * G1[i];
  }
}
}  // namespace

hipError_t launch_propagate_chain(DevSys *d, const double *G1, int N, double *out, hipStream_t st) {
  DevSys h;
  hipError_t e = hipMemcpyAsync(&h, d, sizeof(DevSys), hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(st);
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(out, G1, sizeof(double) * h.M, hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return e;
  const int nb = (h.M + kGridThreads - 1) / kGridThreads;
  for (int s = 1; s < N; s++)
    hipLaunchKernelGGL(k_chain_step, dim3(nb), dim3(kGridThreads), 0, st, d, G1, out + (size_t)(s - 1) * h.M,
                       out + (size_t)s * h.M);
  return hipGetLastError();
}

}  // namespace nscf
