// Host side of the SCF library: builds the device description of a system
// from its SystemSpec, owns its HBM arena, and implements the C ABI of
// include/namics_scf.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "scf_internal.h"

using namespace nscf;

namespace nscf {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char *where) {
  return fail(SCF_E_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

namespace {

// Bump allocator over one device allocation.
struct Arena {
  size_t used = 0;
  size_t take(size_t bytes) {
    size_t off = used;
    used += (bytes + 255) & ~size_t(255);
    return off;
  }
};

size_t g_lds_limit = 0;  // bytes of dynamic LDS one workgroup may use

}  // namespace

size_t lds_limit() {
  if (g_lds_limit == 0) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) == hipSuccess) {
      g_lds_limit = p.sharedMemPerBlock;
      // leave room for the kernel's static shared memory
      const size_t st = fused_static_lds() + 256;
      g_lds_limit = g_lds_limit > st ? g_lds_limit - st : 0;
    } else {
      g_lds_limit = 64 * 1024 - 1024;
    }
  }
  return g_lds_limit;
}


double g_wall_khz = 0;

// frequency of the device wall clock (wall_clock64) in kHz
double wall_clock_khz() {
  if (g_wall_khz == 0) {
    int dev = 0, khz = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
      g_wall_khz = khz;
    else
      g_wall_khz = 100000.0;
  }
  return g_wall_khz;
}

int validate(const SystemSpec &s, std::string &why) {
  const LatticeSpec &l = s.lat;
  if ((int)s.seg.size() > kMaxSeg) return why = "more than 8 segment types", SCF_E_UNSUPPORTED;
  if ((int)s.mol.size() > kMaxMol) return why = "more than 8 molecule types", SCF_E_UNSUPPORTED;
  int nb = 0, nmm = 0;
  for (auto &m : s.mol) {
    if (!m.scheduled()) nb += (int)m.blocks.size();  // branched molecules and dendrimers run from a schedule
    nmm += (int)m.molmon.size();
  }
  for (auto &m : s.mol) {
    if (!m.scheduled()) continue;
    if (m.markov != 1)
      return why = "branched molecules and dendrimers need first-order (Markov 1) chains", SCF_E_UNSUPPORTED;
    if (m.save_memory)
      return why = "save_memory with branched molecules or dendrimers is not supported", SCF_E_UNSUPPORTED;
    BranchPlan plan;
    std::string e;
    if (!build_branch_plan(m, plan, e)) return why = "mol '" + m.name + "': " + e, SCF_E_UNSUPPORTED;
  }
  if (nb > kMaxBlocks) return why = "more than 16 blocks over all molecules", SCF_E_UNSUPPORTED;
  if (nmm > kMaxMM) return why = "more than 16 molecule-segment pairs", SCF_E_UNSUPPORTED;
  if (s.solver.method == SOLVER_LBFGS) {
    if (s.solver.m < 1) return why = "LBFGS needs at least one correction pair (m >= 1)", SCF_E_UNSUPPORTED;
  } else if (s.solver.method == SOLVER_BRR) {
    if (s.solver.m < 2) return why = "BRR needs at least two rank-one terms (m >= 2)", SCF_E_UNSUPPORTED;
  } else if (s.solver.m > kMaxDIIS || s.solver.m < 1) {
    return why = "DIIS history m must be 1..16", SCF_E_UNSUPPORTED;
  }
  if (save_memory_interval(s) > 1)
    for (auto &m : s.mol)
      if (m.markov == 2)
        return why = "save_memory with second-order (Markov 2) chains is not supported", SCF_E_UNSUPPORTED;
  bool surface = false;
  for (int d = 0; d < 6; d++) surface = surface || l.bc[d] == BC_SURFACE;
  if (l.gradients >= 2 && (surface || s.charged)) {
    // the chains over every site in the reference's array form (chain_literal)
    for (auto &m : s.mol)
      if (m.markov != 1 || m.save_memory || m.scheduled())
        return why = "2- and 3-gradient systems with surface boundaries or charges need first-order linear chains "
                     "without save_memory",
               SCF_E_UNSUPPORTED;
    if (l.lattice_type != LAT_SIMPLE_CUBIC || !l.stencil_full || l.fjc != 1)
      return why = "2- and 3-gradient systems with surface boundaries or charges need the full simple-cubic "
                   "stencil without refinement",
             SCF_E_UNSUPPORTED;
  }
  if (l.gradients == 2 && l.geometry == GEOM_PLANAR && !l.stencil_full && l.lattice_type != LAT_SIMPLE_CUBIC)
    return why = "2-gradient planar hexagonal stencil without stencil_full is not supported", SCF_E_UNSUPPORTED;
  if (s.charged) {
    // the potential runs in the one-workgroup kernels; on 2 and 3 gradients
    // the chains run over every site in the reference's array form
    for (auto &m : s.mol)
      if (m.scheduled())
        return why = "charged systems with branched molecules are not supported", SCF_E_UNSUPPORTED;

    if ((int)s.itmon.size() + 1 > kMaxSeg + 1) return why = "too many field blocks", SCF_E_UNSUPPORTED;
  }
  return SCF_OK;
}

int ckpt_rows(int N, int K) { return K <= 1 ? N : (N + K - 1) / K + std::min(K, N) - 1; }

int save_memory_interval(const SystemSpec &s) {
  int n = 0;
  for (auto &m : s.mol)
    if (m.save_memory) n = std::max(n, m.chainlength);
  if (n < 3) return 0;
  const int kept = std::max(1, (int)std::lround(std::sqrt(2.0 * n)));
  const int K = (n + kept - 1) / kept;
  return K > 1 ? K : 0;
}

int check_layout(const scf_system *S, int mode, int strip_sx) {
  if (S->branched() && mode == SCF_MODE_FUSED)
    return fail(SCF_E_UNSUPPORTED, "branched molecules and dendrimers run on the grid path");
  if ((S->spec.charged || S->literal()) && mode != SCF_MODE_FUSED)
    return fail(SCF_E_UNSUPPORTED, "charged systems and 2- or 3-gradient surface boundaries run on the fused "
                                   "(one workgroup per system) path");
  if (S->has_ring() && (mode != SCF_MODE_FUSED || S->literal() ||
                        (S->spec.lat.gradients == 1 && (S->spec.lat.M > 1024 ||
                                                        (size_t)2 * S->spec.lat.M * sizeof(double) > lds_limit()))))
    return fail(SCF_E_UNSUPPORTED, "ring molecules run on the fused (one workgroup per system) path, on 1-gradient "
                                   "lattices up to 1022 layers, and without surface boundaries or charges in 2 and 3 "
                                   "gradients");
  bool markov1 = true;
  for (auto &m : S->spec.mol) markov1 = markov1 && m.markov == 1;
  const bool strip = mode == SCF_MODE_FUSED && strip_sx > 0;
  // FP32 chains on the grid path: the 6-point simple-cubic 3-gradient steps
  const SystemSpec &sp = S->spec;
  const bool grid32 = mode == SCF_MODE_GRID && sp.lat.gradients == 3 && sp.lat.stencil_full &&
                      sp.lat.lattice_type == LAT_SIMPLE_CUBIC && sp.lat.fjc == 1 && markov1 && !S->branched();
  if (S->h.precision != SCF_PREC_FP64 && !strip && !grid32)
    return fail(SCF_E_UNSUPPORTED, "FP32 propagators need the 2-gradient register-strip path or the 6-point "
                                   "3-gradient grid path (first-order linear chains)");
  if (S->h.ckpt > 1 && !strip && !(mode == SCF_MODE_GRID && markov1))
    return fail(SCF_E_UNSUPPORTED, "checkpointed propagators (save_memory) need first-order chains on the grid "
                                   "path or the 2-gradient register-strip path");
  return SCF_OK;
}

// share of segment seg in molecule m (Molecule::fraction src/molecule.cpp:2686-2697,
// with the blocks' degeneracy for dendrimers, mol_dend::fraction src/mol_dendrimer.cpp:12-22)
double fraction(const MoleculeSpec &m, int seg) {
  // a ring's first segment is merged with its last and not counted
  // (Molecule::fraction src/molecule.cpp:2686-2697)
  int n = 0;
  for (size_t b = m.ring ? 1 : 0; b < m.blocks.size(); b++)
    if (m.blocks[b].seg == seg) n += m.blocks[b].n * (b < m.block_deg.size() ? m.block_deg[b] : 1);
  return 1.0 * n / m.chainlength;
}

// ---------------------------------------------------------------- layouts
// Launch layout of the fused path: threads per system, whether the
// propagator rows live in LDS, and on 2-gradient lattices the register-strip
// width SX (8 or 4 sites per thread along x; 0 = the generic site-strided
// chain).  want_sx < 0 picks the first strip width that fits.
void plan_layout(scf_system *S, int want_sx) {
  const LatticeSpec &l = S->spec.lat;
  DevSys &h = S->h;
  const size_t D = sizeof(double);
  const int M = l.M;
  const int W = l.gradients == 1 ? M : l.MX * l.MY * (l.gradients == 3 ? l.MZ : 1);
  int maxR = 1;
  for (auto &m : S->spec.mol) maxR = std::max(maxR, m.size);
  const int work = W;
  h.threads = work <= 64 ? 64 : (work <= 128 ? 128 : (work <= 256 ? 256 : (work <= 512 ? 512 : 1024)));
  h.strip_sx = 0;
  h.strip_stride = 0;
  size_t rows = (size_t)2 * maxR * M * D;
  h.row_stride = maxR * M;
  // 2-gradient strip path: every thread owns SX sites along x; needs
  // Markov-1 chains, the 9-point stencil, ghosts that map onto interior
  // sites, and the rows of one strip layout in LDS
  const bool strip_ok = l.gradients == 2 && l.fjc == 1 && maxR == 1 && (l.geometry != 0 || l.stencil_full) &&
                        !S->literal() && !S->has_ring() &&
                        l.MX + 2 <= 1026 && h.bx1 >= 1 && h.bx1 <= l.MX && h.bxm >= 1 && h.bxm <= l.MX &&
                        h.by1 >= 1 && h.by1 <= l.MY && h.bym >= 1 && h.bym <= l.MY;
  for (int sx : {8, 4}) {
    if (!strip_ok || h.strip_sx || l.MX % sx || (want_sx >= 0 && want_sx != sx)) continue;
    const int nact = (l.MX / sx) * l.MY;
    if (nact > 1024) continue;
    const int stride = nact <= 256 ? 256 : (nact <= 512 ? 512 : 1024);
    const size_t srows = (size_t)2 * sx * stride * D;
    if (srows > lds_limit()) continue;
    h.strip_sx = sx;
    h.strip_stride = stride;
    h.row_stride = sx * stride;
    rows = srows;
    h.threads = std::max(h.threads, stride);
  }
  h.use_lds = rows <= lds_limit() ? 1 : 0;
  S->lds_bytes = h.use_lds ? rows : 0;
}

// Layout choices the autotuner made, per lattice shape, for this process:
// systems built later with the same shape start from the tuned layout.
namespace {
struct TunedLayout {
  int mode;      // SCF_MODE_FUSED or SCF_MODE_GRID
  int strip_sx;  // fused path: 8, 4 or 0
};
std::mutex g_tune_mu;
std::map<std::string, TunedLayout> g_tuned;
}  // namespace

std::string shape_key(const scf_system *S) {
  const LatticeSpec &l = S->spec.lat;
  std::string k = std::to_string(l.gradients) + "/" + std::to_string(l.geometry) + "/" +
                  std::to_string(l.lattice_type) + "/" + std::to_string(l.MX) + "x" + std::to_string(l.MY) + "x" +
                  std::to_string(l.MZ) + "/f" + std::to_string(l.fjc);
  for (auto &m : S->spec.mol) k += "/" + std::to_string(m.chainlength) + "m" + std::to_string(m.markov);
  return k;
}

int tuned_strip(scf_system *S) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tuned.find(shape_key(S));
  if (it == g_tuned.end()) return -1;
  S->tuned_mode = it->second.mode;
  return it->second.strip_sx;
}

int build(scf_system *S, const Placement *pl) {
  const SystemSpec &s = S->spec;
  const LatticeSpec &l = s.lat;
  DevSys &h = S->h;
  std::memset(&h, 0, sizeof(DevSys));
  const int M = l.M;
  const int nseg = (int)s.seg.size(), nmol = (int)s.mol.size(), nit = (int)s.itmon.size();
  const size_t nvar = (size_t)s.nvar();
  h.gradients = l.gradients;
  h.geometry = l.geometry;
  h.lattice_type = l.lattice_type;
  h.stencil_full = l.stencil_full ? 1 : 0;
  h.MX = l.MX;
  h.MY = l.MY;
  h.MZ = l.MZ;
  h.M = M;
  h.JX = l.JX;
  h.JY = l.JY;
  h.bx1 = l.BX1;
  h.bxm = l.BXM;
  h.by1 = l.BY1;
  h.bym = l.BYM;
  h.bz1 = l.BZ1;
  h.bzm = l.BZM;
  h.surface_lo = l.bc[0] == BC_SURFACE;
  h.surface_hi = l.bc[3] == BC_SURFACE;
  static_assert(kMaxFjc >= kMaxFjcSpec, "refinement table too small");
  h.fjc = l.fjc;
  for (int k = 0; k < l.fjc; k++) {
    h.bx1f[k] = l.bx1f[k];
    h.bxmf[k] = l.bxmf[k];
    h.by1f[k] = l.by1f[k];
    h.bymf[k] = l.bymf[k];
    h.bz1f[k] = l.bz1f[k];
    h.bzmf[k] = l.bzmf[k];
  }
  // planar weighted sums (src/LGrad1.cpp:213-221, src/LGrad2.cpp:147-157,
  // src/LGrad3.cpp:40-46): divided by fjc, fjc^2, and not at all in 3-D
  h.wplanar = l.gradients == 1 ? 1.0 / l.fjc : (l.gradients == 2 ? 1.0 / (l.fjc * l.fjc) : 1.0);
  h.lambda = l.lambda;
  h.nseg = nseg;
  h.nit = s.nblocks_x();  // field blocks, plus the potential of a charged system
  h.nitm = nit;
  h.charged = s.charged ? 1 : 0;
  h.all_sites = S->literal() ? 1 : 0;
  h.neutralizer = s.charged ? s.neutralizer : -1;
  if (s.charged) {
    // constants of the reference (src/namics.cpp:43-49); the bond length is
    // `bondlength` over fjc (src/lattice.cpp:332-337)
    const double e = 1.60217e-19, T = 298.15, k_B = 1.38065e-23, eps0 = 8.85418e-12;
    const double k_BT = k_B * T, b = l.bond_length / l.fjc, f = l.fjc;
    // UpdateEE (src/LG1Planar.cpp:163-168, src/LGrad1.cpp:593-652)
    double pf = 0.5 * eps0 * b / k_BT * (k_BT / e) * (k_BT / e);
    // UpdatePsi (src/LG1Planar.cpp:185-188, src/LGrad1.cpp:659-680)
    double C = e * e / (eps0 * k_BT * b);
    if (l.gradients == 3) {
      // LGrad3::UpdateEE / UpdatePsi (src/LGrad3.cpp:1131-1141,1150-1224)
      h.el_ri = 0;
    } else if (l.gradients == 2 && l.geometry == GEOM_PLANAR) {
      // LG2Planar::UpdateEE / UpdatePsi (src/LG2Planar.cpp:1130-1181)
      pf = pf / 2.0;
      C = C * 2.0 / f / f;
      h.el_ri = 0;
    } else if (l.gradients == 2) {
      // LGrad2::UpdateEE / UpdatePsi (src/LGrad2.cpp:1392-1445)
      pf = pf / 2;
      C = C / 2 / f / f;
      h.el_ri = (int)l.offset_first_layer;
    } else if (l.geometry == GEOM_PLANAR) {
      pf = pf / 2.0 * f * f;
      C = C * 2.0 / f / f;
      h.el_ri = 0;
    } else if (l.geometry == GEOM_CYLINDRICAL) {
      pf = pf * kPIE;
      C = C / kPIE;
      h.el_ri = (int)l.offset_first_layer;
    } else {
      pf = pf * kPIE * 2 / f;
      C = C / (2.0 * kPIE) * f;
      h.el_ri = (int)(l.offset_first_layer + 1.0);
    }
    h.el_pf = pf;
    h.el_C = C;
    h.el_r0 = l.offset_first_layer;
    h.el_C0 = e * e / (eps0 * k_BT * b);
    h.grad_eps = s.grad_epsilon ? 1 : 0;
    h.fixed_psi0 = s.fixed_psi0 ? 1 : 0;
    h.el_last_fixed = -1;
    for (int k = 0; k < nseg; k++) {
      h.seg_fixpsi[k] = s.seg[k].freedom == SEG_FROZEN && s.seg[k].fixed_psi0 ? 1 : 0;
      h.seg_psi0[k] = s.seg[k].psi0;
      if (h.seg_fixpsi[k]) h.el_last_fixed = k;
    }
  }
  for (int k = 0; k < nseg; k++) {
    h.seg_val[k] = s.seg[k].valence;
    h.seg_eps[k] = s.seg[k].epsilon;
    h.seg_block[k] = s.seg[k].field_block;
    h.seg_frozen[k] = s.seg[k].freedom == SEG_FROZEN;
    h.seg_pinned[k] = s.seg[k].freedom == SEG_PINNED;
    h.seg_tagged[k] = s.seg[k].freedom == SEG_TAGGED;
    for (int j = 0; j < nseg; j++) h.chi[k * kMaxSeg + j] = s.seg[k].chi[j];
  }
  for (int i = 0; i < nit; i++) h.itmon[i] = s.itmon[i];
  h.nmol = nmol;
  S->plans.assign(nmol, BranchPlan());
  h.solvent = s.solvent;
  int nb = 0, nmm = 0, maxNR = 0, maxR = 0;
  for (int i = 0; i < nmol; i++) {
    const MoleculeSpec &m = s.mol[i];
    h.mol_N[i] = m.chainlength;
    h.mol_freedom[i] = m.freedom;
    h.mol_pinned[i] = m.pinned;
    h.mol_tagged[i] = m.tagged;
    h.mol_ring[i] = m.ring;
    h.mol_NF[i] = m.chainlength - (m.tagged ? 1 : 0);
    h.mol_markov[i] = m.markov;
    h.mol_size[i] = m.size;
    h.mol_block_begin[i] = nb;
    h.mol_mm_begin[i] = nmm;
    h.branch[i] = nullptr;
    if (m.scheduled()) {
      std::string e;
      if (!build_branch_plan(m, S->plans[i], e)) return fail(SCF_E_UNSUPPORTED, e);
      h.branch[i] = &S->plans[i];
    }
    for (size_t b = 0; b < m.blocks.size() && !m.scheduled(); b++) {
      h.block_seg[nb] = m.blocks[b].seg;
      h.block_n[nb] = m.blocks[b].n;
      h.block_mm[nb] = nmm + m.block_molmon[b];
      nb++;
    }
    for (int sg : m.molmon) {
      h.mm_seg[nmm++] = sg;
      h.sysmon[sg] = s.seg[sg].freedom != SEG_TAGGED;  // SysMonList leaves tagged segments out
    }
    for (int k = 0; k < kMaxComp && k < (int)m.P.size(); k++) h.mol_P[i][k] = m.P[k];
    h.mol_phibulk_in[i] = m.phibulk;
    h.mol_n_in[i] = m.n;
    h.mol_charge[i] = s.mol_charge[i];
    // fractions of a molecule with a tagged segment are scaled by N / (N - 1)
    // (System::GetFreeEnergy src/system.cpp:2988-3012)
    double cf = 0;
    for (int a : m.molmon)
      for (int b : m.molmon) {
        double fA = fraction(m, a), fB = fraction(m, b);
        if (m.tagged) {
          fA = m.chainlength > 1 ? fA * m.chainlength / (m.chainlength - 1) : 0.0;
          fB = m.chainlength > 1 ? fB * m.chainlength / (m.chainlength - 1) : 0.0;
        }
        cf -= fA * fB * (s.seg[a].chi[b] / 2);
      }
    h.mol_constF[i] = cf;
    for (int k = 0; k < nseg; k++) h.frac[i][k] = fraction(m, k);
    maxR = std::max(maxR, m.size);
  }
  // end-point storage: every row, or checkpoints with the input's save_memory
  h.ckpt = save_memory_interval(s);
  for (int i = 0; i < nmol; i++) {
    const MoleculeSpec &m = s.mol[i];
    maxNR = std::max(maxNR, m.scheduled() ? S->plans[i].rows : ckpt_rows(m.chainlength, h.ckpt) * m.size);
  }
  S->gf_rows = maxNR;
  h.mol_block_begin[nmol] = nb;
  h.mol_mm_begin[nmol] = nmm;
  h.nblocks = nb;
  h.nmm = nmm;
  // DIIS history of the device solvers (the job queue and the server solve
  // every input by DIIS); an LBFGS input's m counts its correction pairs
  h.m = s.solver.method == SOLVER_LBFGS || s.solver.method == SOLVER_BRR ? std::min(s.solver.m, kMaxDIIS)
                                                                           : s.solver.m;
  h.itlim = s.solver.iterationlimit;
  h.max_g = s.solver.max_g;
  h.deltamax = s.solver.deltamax;
  h.tol = s.solver.tolerance;

  // ---- arena layout
  Arena a;
  const size_t D = sizeof(double);
  const bool curved = l.gradients < 3;
  const bool m2curved = !l.l1.empty();
  size_t o_L = curved ? a.take(4 * M * D) : 0;
  size_t o_l1 = m2curved ? a.take(4 * M * D) : 0;
  const bool has_LAM = l.fjc > 1 && l.LAMBDA.size() == (size_t)l.FJC * M;  // 1 and 2 gradients
  size_t o_LAM = has_LAM ? a.take((size_t)l.FJC * M * D) : 0;
  size_t o_mask = a.take((size_t)nseg * M * D);
  size_t o_ksam = a.take(M * D);
  size_t o_x = a.take(nvar * D);
  size_t o_g = a.take(nvar * D);
  size_t o_G1 = a.take((size_t)nseg * M * D);
  // FP32 segment weights of the grid path's mixed-precision chains (3 gradients)
  size_t o_G1f = l.gradients == 3 ? a.take((size_t)nseg * M * sizeof(float)) : 0;
  size_t o_Gf = a.take((size_t)maxNR * M * D);
  size_t o_Gb = a.take((size_t)2 * maxR * M * D);
  size_t o_rho = a.take((size_t)nmm * M * D);
  size_t o_phi = a.take((size_t)nseg * M * D);
  size_t o_side = a.take((size_t)nseg * M * D);
  size_t o_pt = a.take(M * D);
  size_t o_alpha = a.take(M * D);
  size_t o_mpt = a.take((size_t)nmol * M * D);
  size_t o_el = s.charged ? a.take((size_t)6 * M * D) : 0;
  S->m_alloc = h.m;
  size_t o_xR = a.take((size_t)h.m * nvar * D);
  size_t o_xx0 = a.take((size_t)h.m * nvar * D);
  size_t o_x0 = a.take(nvar * D);
  size_t o_sc = a.take(sizeof(DevScalars));
  // grid path: partial sums (thermo needs 2+nmol columns, DIIS kMaxDIIS)
  const int W = l.gradients == 1 ? M : l.MX * l.MY * (l.gradients == 3 ? l.MZ : 1);
  S->nb = std::max(1, std::min((W + 255) / 256, 2048));
  size_t o_part = a.take((size_t)S->nb * std::max(kMaxDIIS, 2 + kMaxMol) * D);
  size_t o_dev = a.take(sizeof(DevSys));
  size_t o_list = a.take(sizeof(DevSys *));
  S->arena_bytes = a.used;
  if (pl && pl->dry) return SCF_OK;
  char *base = nullptr;
  if (pl && pl->region) {
    if (a.used > pl->capacity) return fail(SCF_E_OOM, "system needs " + std::to_string(a.used) + " bytes");
    base = S->region = pl->region;
  } else {
    hipError_t e = hipMalloc((void **)&S->arena, a.used);
    if (e != hipSuccess) return fail(SCF_E_OOM, "device allocation of " + std::to_string(a.used) + " bytes failed");
    base = S->arena;
  }
  hipStream_t st = S->stream;
  HIP_TRY(hipMemsetAsync(base, 0, a.used, st));
  auto P = [&](size_t off) { return (double *)(base + off); };
  if (curved) {
    h.L = P(o_L);
    h.lam_m = P(o_L) + M;
    h.lam_0 = P(o_L) + 2 * M;
    h.lam_p = P(o_L) + 3 * M;
    HIP_TRY(hipMemcpyAsync((void *)h.L, l.L.data(), M * D, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync((void *)h.lam_m, l.lambda_1.data(), M * D, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync((void *)h.lam_0, l.lambda0.data(), M * D, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync((void *)h.lam_p, l.lambda1.data(), M * D, hipMemcpyHostToDevice, st));
  }
  if (has_LAM) {
    h.LAM = P(o_LAM);
    HIP_TRY(hipMemcpyAsync((void *)h.LAM, l.LAMBDA.data(), (size_t)l.FJC * M * D, hipMemcpyHostToDevice, st));
  }
  if (m2curved) {
    h.l1 = P(o_l1);
    h.l11 = P(o_l1) + M;
    h.l_1 = P(o_l1) + 2 * M;
    h.l_11 = P(o_l1) + 3 * M;
    HIP_TRY(hipMemcpyAsync((void *)h.l1, l.l1.data(), M * D, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync((void *)h.l11, l.l11.data(), M * D, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync((void *)h.l_1, l.l_1.data(), M * D, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync((void *)h.l_11, l.l_11.data(), M * D, hipMemcpyHostToDevice, st));
  }
  {
    std::vector<double> mask((size_t)nseg * M, 0.0), ks(M, 0.0);
    for (int k = 0; k < nseg; k++)
      for (int i = 0; i < M; i++) mask[(size_t)k * M + i] = s.seg[k].mask[i];
    for (int i = 0; i < M; i++) ks[i] = s.ksam[i];
    h.seg_mask = P(o_mask);
    h.ksam = P(o_ksam);
    HIP_TRY(hipMemcpyAsync((void *)h.seg_mask, mask.data(), mask.size() * D, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync((void *)h.ksam, ks.data(), M * D, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));  // the host vectors end here
  }
  h.x = P(o_x);
  h.g = P(o_g);
  h.G1 = P(o_G1);
  h.G1f = l.gradients == 3 ? (float *)P(o_G1f) : nullptr;
  h.Gf = P(o_Gf);
  h.Gb = P(o_Gb);
  h.rho = P(o_rho);
  h.phi = P(o_phi);
  h.phi_side = P(o_side);
  h.phitot = P(o_pt);
  h.alpha = P(o_alpha);
  h.mol_phitot = P(o_mpt);
  if (s.charged) {
    h.psi = P(o_el);
    h.EE = P(o_el) + M;
    h.q = P(o_el) + 2 * M;
    h.eps = P(o_el) + 3 * M;
    h.elA = P(o_el) + 4 * M;
    h.elB = P(o_el) + 5 * M;
  }
  h.xR = P(o_xR);
  h.xx0 = P(o_xx0);
  h.x0 = P(o_x0);
  h.sc = (DevScalars *)(base + o_sc);
  h.partial = P(o_part);
  plan_layout(S, tuned_strip(S));
  h.precision = SCF_PREC_FP64;
  h.mixed_switch = 1e-4;
  {
    // weighted interior volume (Lattice::WeightedSum of ones)
    double vol = 0;
    if (l.fjc > 1) vol = l.volume;
    for (int x = 1; l.fjc == 1 && x <= l.MX; x++)
      for (int y = 1; y <= (l.gradients >= 2 ? l.MY : 1); y++)
        for (int z = 1; z <= (l.gradients == 3 ? l.MZ : 1); z++) {
          const int i = l.gradients == 1 ? x : (l.gradients == 2 ? x * h.JX + y : x * h.JX + y * h.JY + z);
          vol += (l.gradients == 3 || l.geometry == GEOM_PLANAR) ? 1.0 : l.L[i];
        }
    h.volume = vol;
  }
  S->d = (DevSys *)(base + o_dev);
  S->d_list = (DevSys **)(base + o_list);
  HIP_TRY(hipMemcpyAsync(S->d, &h, sizeof(DevSys), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(S->d_list, &S->d, sizeof(DevSys *), hipMemcpyHostToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));
  return SCF_OK;
}

int upload_solver(scf_system *S) {
  HIP_TRY(hipMemcpyAsync(S->d, &S->h, sizeof(DevSys), hipMemcpyHostToDevice, S->stream));
  HIP_TRY(hipStreamSynchronize(S->stream));
  return SCF_OK;
}

int read_scalars(const scf_system *S, DevScalars &sc) {
  HIP_TRY(hipMemcpyAsync(&sc, S->h.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, S->stream));
  HIP_TRY(hipStreamSynchronize(S->stream));
  return SCF_OK;
}

// host-driven DIIS over the grid path (SFNewton::iterate_DIIS
// src/sfnewton.cpp:1037-1130).  The loop body (one iteration: residual, DIIS
// step, combine) is captured once into a hipGraph and replayed; the host
// reads back the scalars after each replay to test convergence.
int grid_solve(scf_system *S, scf_solve_info *info) {
  const DevSys &h = S->h;
  hipStream_t st = S->stream;
  DevScalars sc;
  HIP_TRY(grid_init(S->d, st));
  HIP_TRY(grid_residual(h, S->d, h.partial, S->nb, st));
  HIP_TRY(hipMemcpyAsync(&sc, h.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  int it = 0, fp32_its = 0;
  double res = sc.residual;
  // Mixed precision: FP32 chains from the first iteration on while the
  // residual is above the switch level and still falling (a new best within
  // kMixedPatience evaluations); then FP64 for good, with the DIIS history
  // restarted, so the converged state is an FP64 one.
  bool fp32 = h.precision != SCF_PREC_FP64 && sc.status == DS_OK && res > h.mixed_switch;
  double best = res;
  int since = 0;
  auto leave_fp32 = [&]() -> int {
    fp32 = false;
    HIP_TRY(grid_diis_restart(S->d, st));
    if (res <= h.tol) res = 2 * h.tol;  // the last residual was an FP32 one: evaluate again in FP64
    return SCF_OK;
  };
  if (sc.status == DS_OK) {
    HIP_TRY(grid_axpy_first(h, S->d, S->nb, st));
    while (res > h.tol && it < h.itlim) {
      hipGraphExec_t &ge = fp32 ? S->graph_exec32 : S->graph_exec;
      if (!ge && S->use_graph) {
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t ce = grid_iteration(h, S->d, h.partial, S->nb, st, fp32);
        hipError_t ee = hipStreamEndCapture(st, &graph);
        if (ce != hipSuccess || ee != hipSuccess) {
          if (graph) (void)hipGraphDestroy(graph);
          return ce != hipSuccess ? hip_fail(ce, "grid_iteration capture") : hip_fail(ee, "hipStreamEndCapture");
        }
        hipError_t ie = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) {
          ge = nullptr;
          return hip_fail(ie, "hipGraphInstantiate");
        }
      }
      if (ge)
        HIP_TRY(hipGraphLaunch(ge, st));
      else
        HIP_TRY(grid_iteration(h, S->d, h.partial, S->nb, st, fp32));
      HIP_TRY(hipMemcpyAsync(&sc, h.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      it = sc.iterations;
      if (fp32) fp32_its++;
      if (sc.status != DS_OK) {
        if (!fp32) break;
        // an FP32 evaluation that failed (GN out of range): the iteration
        // changed nothing; FP64 from here
        int rc = leave_fp32();
        if (rc != SCF_OK) return rc;
        continue;
      }
      res = sc.residual;
      if (fp32) {
        if (res < best) {
          best = res;
          since = 0;
        } else {
          since++;
        }
        if (!(res > h.mixed_switch) || since >= kMixedPatience) {
          int rc = leave_fp32();
          if (rc != SCF_OK) return rc;
        }
      }
    }
  }
  if (sc.status == DS_OK) {
    HIP_TRY(grid_thermo(h, S->d, h.partial, S->nb, st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  info->iterations = it;
  info->residual = res;
  info->fp32_iterations = fp32_its;
  info->status = sc.status != DS_OK ? sc.status : (res <= h.tol ? SCF_OK : SCF_NOT_CONVERGED);
  return SCF_OK;
}

}  // namespace

extern "C" {

const char *scf_last_error(void) { return g_last_error.c_str(); }

int scf_system_create(const char *input_text, int start, scf_system **out, char *err, size_t errlen) {
  if (!input_text || !out) return fail(SCF_E_ARG, "null argument");
  *out = nullptr;
  auto report = [&](int code, const std::string &msg) {
    if (err && errlen > 0) snprintf(err, errlen, "%s", msg.c_str());
    return fail(code, msg);
  };
  scf_system *S = new scf_system();
  std::string why;
  if (!parse_input(input_text, start, S->spec, why)) {
    delete S;
    return report(SCF_E_PARSE, why);
  }
  int v = validate(S->spec, why);
  if (v != SCF_OK) {
    delete S;
    return report(v, why);
  }
  if (hipStreamCreate(&S->stream) != hipSuccess || hipEventCreate(&S->ev0) != hipSuccess ||
      hipEventCreate(&S->ev1) != hipSuccess) {
    std::string msg = "hipStreamCreate / hipEventCreate failed";
    scf_system_destroy(S);
    return report(SCF_E_HIP, msg);
  }
  int rc = build(S);
  if (rc != SCF_OK) {
    std::string msg = g_last_error;
    scf_system_destroy(S);
    return report(rc, msg);
  }
  *out = S;
  return SCF_OK;
}

void scf_system_destroy(scf_system *S) {
  if (!S) return;
  if (S->stream) hipStreamSynchronize(S->stream);
  S->drop_graphs();
  if (S->arena) hipFree(S->arena);
  if (S->ev0) hipEventDestroy(S->ev0);
  if (S->ev1) hipEventDestroy(S->ev1);
  if (S->stream && S->owns_stream) hipStreamDestroy(S->stream);
  delete S;
}

int scf_system_describe(const scf_system *S, char *buf, size_t len) {
  if (!S) return fail(SCF_E_ARG, "null system");
  // the system's current end-point checkpoint interval first (0: every row)
  std::string d = "{\"checkpoint\":" + std::to_string(S->h.ckpt) + "," + describe_json(S->spec).substr(1);
  if (buf && len > 0) snprintf(buf, len, "%s", d.c_str());
  return (int)d.size() + 1;
}

int scf_describe_input(const char *input_text, int start, char *buf, size_t len) {
  if (!input_text) return fail(SCF_E_ARG, "null input");
  SystemSpec spec;
  std::string why;
  if (!parse_input(input_text, start, spec, why)) return fail(SCF_E_PARSE, why);
  std::string d = describe_json(spec);
  if (buf && len > 0) snprintf(buf, len, "%s", d.c_str());
  return (int)d.size() + 1;
}

int64_t scf_system_nvar(const scf_system *S) { return S ? (int64_t)S->spec.nvar() : -1; }
int64_t scf_system_sites(const scf_system *S) { return S ? (int64_t)S->spec.lat.M : -1; }
size_t scf_system_device_bytes(const scf_system *S) { return S ? S->arena_bytes : 0; }

int scf_system_set_mode(scf_system *S, int mode) {
  if (!S || mode < SCF_MODE_AUTO || mode > SCF_MODE_GRID) return fail(SCF_E_ARG, "bad mode");
  if (mode == SCF_MODE_GRID)
    for (auto &m : S->spec.mol)
      if (m.markov == 2) return fail(SCF_E_UNSUPPORTED, "second-order chains run on the fused path only");
  const int saved = S->mode;
  S->mode = mode;
  int rc = check_layout(S, S->effective_mode(), S->h.strip_sx);
  if (rc != SCF_OK) S->mode = saved;
  return rc;
}

int scf_system_set_precision(scf_system *S, int mode, double switch_residual) {
  if (!S) return fail(SCF_E_ARG, "null system");
  if (mode != SCF_PREC_FP64 && mode != SCF_PREC_MIXED) return fail(SCF_E_ARG, "unknown precision mode");
  if (mode == SCF_PREC_MIXED && !(switch_residual > S->h.tol))
    return fail(SCF_E_ARG, "the FP64 switch level must lie above the tolerance");
  const int saved = S->h.precision;
  S->h.precision = mode;
  int rc = check_layout(S, S->effective_mode(), S->h.strip_sx);
  if (rc != SCF_OK) {
    S->h.precision = saved;
    return rc;
  }
  if (mode == SCF_PREC_MIXED) S->h.mixed_switch = switch_residual;
  if (mode != saved && (S->graph_exec || S->graph_exec32)) {
    // a mixed solve's captured iterations also refresh the gauge shifts
    HIP_TRY(hipStreamSynchronize(S->stream));
    S->drop_graphs();
  }
  return upload_solver(S);
}

int scf_system_set_graph(scf_system *S, int on) {
  if (!S) return fail(SCF_E_ARG, "null system");
  S->use_graph = on != 0;
  if (!S->use_graph && (S->graph_exec || S->graph_exec32)) {
    HIP_TRY(hipStreamSynchronize(S->stream));
    S->drop_graphs();
  }
  return SCF_OK;
}

int scf_system_set_checkpoint(scf_system *S, int interval) {
  if (!S) return fail(SCF_E_ARG, "null system");
  if (interval < 0) return fail(SCF_E_ARG, "negative checkpoint interval");
  const int K = interval <= 1 ? 0 : interval;
  for (auto &m : S->spec.mol)
    if (ckpt_rows(m.chainlength, K) * m.size > S->gf_rows)
      return fail(SCF_E_ARG, "the arena holds " + std::to_string(S->gf_rows) +
                                 " end-point rows (the input's save_memory); interval " + std::to_string(interval) +
                                 " needs " + std::to_string(ckpt_rows(m.chainlength, K) * m.size));
  const int saved = S->h.ckpt;
  S->h.ckpt = K;
  int rc = check_layout(S, S->effective_mode(), S->h.strip_sx);
  if (rc != SCF_OK) {
    S->h.ckpt = saved;
    return rc;
  }
  if (K != saved && (S->graph_exec || S->graph_exec32)) {
    // the captured grid-path iterations hold the old step schedule
    HIP_TRY(hipStreamSynchronize(S->stream));
    S->drop_graphs();
  }
  return upload_solver(S);
}

int scf_system_layout(const scf_system *S, int *mode, int *strip_sx, int *threads) {
  if (!S) return fail(SCF_E_ARG, "null system");
  if (mode) *mode = S->effective_mode();
  if (strip_sx) *strip_sx = S->h.strip_sx;
  if (threads) *threads = S->h.threads;
  return SCF_OK;
}

void scf_autotune_forget(void) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned.clear();
}

int scf_system_autotune(scf_system *S, int evals, scf_tune_report *rep) {
  if (!S) return fail(SCF_E_ARG, "null system");
  if (evals < 1) evals = 1;
  scf_tune_report tmp;
  if (!rep) rep = &tmp;
  std::memset(rep, 0, sizeof(*rep));
  const LatticeSpec &l = S->spec.lat;
  // candidate layouts: those the system's options (precision, checkpointed
  // storage, second-order chains) run on
  bool markov2 = false;
  for (auto &m : S->spec.mol) markov2 = markov2 || m.markov == 2;
  std::vector<TunedLayout> cand;
  auto consider = [&](int mode, int sx) {
    if (check_layout(S, mode, sx) == SCF_OK) cand.push_back({mode, sx});
  };
  if (S->mode == SCF_MODE_AUTO && !markov2) consider(SCF_MODE_GRID, 0);
  if (S->mode != SCF_MODE_GRID) {
    if (l.gradients == 2) {
      for (int sx : {8, 4}) {
        plan_layout(S, sx);
        if (S->h.strip_sx == sx) consider(SCF_MODE_FUSED, sx);
      }
      consider(SCF_MODE_FUSED, 0);
    } else {
      consider(SCF_MODE_FUSED, 0);
    }
  }
  if (cand.empty()) {
    plan_layout(S, -1);
    return check_layout(S, S->effective_mode(), S->h.strip_sx);
  }
  const int saved_mode = S->mode, saved_tuned = S->tuned_mode;
  double best = 1e300;
  int besti = 0;
  DevSys &h = S->h;
  for (size_t c = 0; c < cand.size() && c < SCF_TUNE_MAX; c++) {
    S->tuned_mode = cand[c].mode;
    plan_layout(S, cand[c].mode == SCF_MODE_FUSED && l.gradients == 2 ? cand[c].strip_sx : -1);
    int rc = upload_solver(S);
    if (rc != SCF_OK) return rc;
    auto one = [&]() -> hipError_t {
      if (S->effective_mode() == SCF_MODE_FUSED)
        return launch_fused(S->d_list, 1, FUSED_RESIDUAL, h.threads, S->lds_bytes, S->stream);
      hipError_t e = grid_residual(h, S->d, h.partial, S->nb, S->stream);
      return e != hipSuccess ? e : grid_thermo(h, S->d, h.partial, S->nb, S->stream);
    };
    HIP_TRY(one());  // warm-up (and the cost bound for slow layouts)
    HIP_TRY(hipEventRecord(S->ev0, S->stream));
    HIP_TRY(one());
    HIP_TRY(hipEventRecord(S->ev1, S->stream));
    HIP_TRY(hipEventSynchronize(S->ev1));
    float ms1 = 0;
    HIP_TRY(hipEventElapsedTime(&ms1, S->ev0, S->ev1));
    double ms = ms1;
    if (ms1 < 4 * best && evals > 1) {  // a clearly slower layout is not timed further
      HIP_TRY(hipEventRecord(S->ev0, S->stream));
      for (int e = 0; e < evals; e++) HIP_TRY(one());
      HIP_TRY(hipEventRecord(S->ev1, S->stream));
      HIP_TRY(hipEventSynchronize(S->ev1));
      HIP_TRY(hipEventElapsedTime(&ms1, S->ev0, S->ev1));
      ms = ms1 / evals;
    }
    rep->mode[c] = cand[c].mode;
    rep->strip_sx[c] = l.gradients == 2 && cand[c].mode == SCF_MODE_FUSED ? h.strip_sx : 0;
    rep->ms_per_eval[c] = ms;
    rep->candidates = (int)c + 1;
    if (ms < best) {
      best = ms;
      besti = (int)c;
    }
  }
  rep->chosen = besti;
  TunedLayout pick{rep->mode[besti], rep->strip_sx[besti]};
  S->mode = saved_mode;
  S->tuned_mode = saved_mode == SCF_MODE_AUTO ? pick.mode : saved_tuned;
  plan_layout(S, pick.mode == SCF_MODE_FUSED && l.gradients == 2 ? pick.strip_sx : -1);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[shape_key(S)] = pick;
  }
  return upload_solver(S);
}

int scf_autotune_batch(scf_system *const *systems, int n, int evals, scf_batch_tune_report *rep) {
  if (!systems || n <= 0) return fail(SCF_E_ARG, "null argument");
  if (evals < 1) evals = 1;
  scf_batch_tune_report tmp;
  if (!rep) rep = &tmp;
  std::memset(rep, 0, sizeof(*rep));
  bool strip = true;
  int threads = 64;
  size_t lds = 0;
  for (int i = 0; i < n; i++) {
    scf_system *S = systems[i];
    if (!S) return fail(SCF_E_ARG, "null system");
    if (S->effective_mode() != SCF_MODE_FUSED || S->newton())
      return fail(SCF_E_UNSUPPORTED, "batch tuning times the one-workgroup-per-system launch");
    strip = strip && S->h.strip_sx > 0;
    threads = std::max(threads, S->h.threads);
    lds = std::max(lds, S->lds_bytes);
    HIP_TRY(hipStreamSynchronize(S->stream));
  }
  // end-point storages every system can hold (the input's save_memory sizes
  // the arena for checkpoints only)
  std::vector<int> cand;
  for (int K : {1, 2}) {
    bool fits = K == 1 || strip;
    for (int i = 0; i < n && fits; i++)
      for (auto &m : systems[i]->spec.mol) fits = fits && ckpt_rows(m.chainlength, K) * m.size <= systems[i]->gf_rows;
    if (fits) cand.push_back(K);
  }
  if (cand.empty()) return fail(SCF_E_UNSUPPORTED, "no end-point storage fits every system");
  scf_system *S0 = systems[0];
  hipStream_t st = S0->stream;
  std::vector<DevSys *> list;
  for (int i = 0; i < n; i++) list.push_back(systems[i]->d);
  DevSys **d_list = nullptr;
  HIP_TRY(hipMalloc((void **)&d_list, n * sizeof(DevSys *)));
  auto body = [&]() -> int {
    HIP_TRY(hipMemcpy(d_list, list.data(), n * sizeof(DevSys *), hipMemcpyHostToDevice));
    double best = 1e300;
    for (size_t c = 0; c < cand.size(); c++) {
      for (int i = 0; i < n; i++) {
        int rc = scf_system_set_checkpoint(systems[i], cand[c]);
        if (rc != SCF_OK) return rc;
      }
      HIP_TRY(launch_fused(d_list, n, FUSED_RESIDUAL, threads, lds, st));  // warm-up
      HIP_TRY(hipEventRecord(S0->ev0, st));
      for (int e = 0; e < evals; e++) HIP_TRY(launch_fused(d_list, n, FUSED_RESIDUAL, threads, lds, st));
      HIP_TRY(hipEventRecord(S0->ev1, st));
      HIP_TRY(hipEventSynchronize(S0->ev1));
      float ms = 0;
      HIP_TRY(hipEventElapsedTime(&ms, S0->ev0, S0->ev1));
      rep->checkpoint[c] = cand[c];
      rep->ms_per_launch[c] = ms / evals;
      rep->candidates = (int)c + 1;
      if (ms < best) {
        best = ms;
        rep->chosen = (int)c;
      }
    }
    for (int i = 0; i < n; i++) {
      int rc = scf_system_set_checkpoint(systems[i], rep->checkpoint[rep->chosen]);
      if (rc != SCF_OK) return rc;
    }
    return SCF_OK;
  };
  const int rc = body();
  hipFree(d_list);
  return rc;
}

int scf_system_set_solver(scf_system *S, double tolerance, double deltamax, int m, int iterationlimit) {
  if (!S) return fail(SCF_E_ARG, "null system");
  const bool lbfgs = S->spec.solver.method == SOLVER_LBFGS || S->spec.solver.method == SOLVER_BRR;
  if (m > S->m_alloc && !lbfgs) return fail(SCF_E_UNSUPPORTED, "DIIS history m above the input's setting");
  if (tolerance > 0) S->h.tol = S->spec.solver.tolerance = tolerance;
  if (deltamax > 0) S->h.deltamax = S->spec.solver.deltamax = deltamax;
  if (m > 0) {
    S->spec.solver.m = m;  // LBFGS: correction pairs; the device DIIS keeps what its arena holds
    S->h.m = lbfgs ? std::min(m, S->m_alloc) : m;
  }
  if (iterationlimit >= 0) S->h.itlim = S->spec.solver.iterationlimit = iterationlimit;
  return upload_solver(S);
}

int scf_residual(scf_system *S, const double *x, double *g, double *residual) {
  if (!S || !x) return fail(SCF_E_ARG, "null argument");
  const DevSys &h = S->h;
  const size_t nvar = (size_t)h.nit * h.M;
  HIP_TRY(hipMemcpyAsync(h.x, x, nvar * sizeof(double), hipMemcpyHostToDevice, S->stream));
  if (S->effective_mode() == SCF_MODE_FUSED) {
    HIP_TRY(launch_fused(S->d_list, 1, FUSED_RESIDUAL, h.threads, S->lds_bytes, S->stream));
  } else {
    HIP_TRY(grid_residual(h, S->d, h.partial, S->nb, S->stream));
    HIP_TRY(grid_thermo(h, S->d, h.partial, S->nb, S->stream));
  }
  DevScalars sc;
  HIP_TRY(hipMemcpyAsync(&sc, h.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, S->stream));
  if (g) HIP_TRY(hipMemcpyAsync(g, h.g, nvar * sizeof(double), hipMemcpyDeviceToHost, S->stream));
  HIP_TRY(hipStreamSynchronize(S->stream));
  if (residual) *residual = sc.residual;
  S->have_results = sc.status == DS_OK;
  if (sc.status != DS_OK) return fail(sc.status, "residual evaluation failed with status " + std::to_string(sc.status));
  return SCF_OK;
}

int scf_solve(scf_system *S, double *x, scf_solve_info *info) {
  scf_solve_info tmp;
  if (!info) info = &tmp;
  return scf_solve_batch(&S, 1, &x, info);
}

int scf_upload_fields(scf_system *S, const double *x) {
  if (!S || !x) return fail(SCF_E_ARG, "null argument");
  const size_t nvar = (size_t)S->h.nit * S->h.M;
  HIP_TRY(hipMemcpyAsync(S->h.x, x, nvar * sizeof(double), hipMemcpyHostToDevice, S->stream));
  HIP_TRY(hipStreamSynchronize(S->stream));
  return SCF_OK;
}

int scf_download_fields(const scf_system *S, double *x) {
  if (!S || !x) return fail(SCF_E_ARG, "null argument");
  const size_t nvar = (size_t)S->h.nit * S->h.M;
  HIP_TRY(hipMemcpyAsync(x, S->h.x, nvar * sizeof(double), hipMemcpyDeviceToHost, S->stream));
  HIP_TRY(hipStreamSynchronize(S->stream));
  return SCF_OK;
}

int scf_zero_fields(scf_system *S) {
  if (!S) return fail(SCF_E_ARG, "null system");
  const size_t nvar = (size_t)S->h.nit * S->h.M;
  HIP_TRY(hipMemsetAsync(S->h.x, 0, nvar * sizeof(double), S->stream));
  return SCF_OK;
}

int scf_carry_fields(scf_system *dst, const scf_system *src) {
  if (!dst || !src) return fail(SCF_E_ARG, "null argument");
  const int M = dst->h.M;
  if (src->h.M != M) return fail(SCF_E_ARG, "systems differ in lattice size");
  HIP_TRY(hipStreamSynchronize(src->stream));
  for (size_t j = 0; j < dst->spec.itmon.size(); j++) {
    const std::string &name = dst->spec.seg[dst->spec.itmon[j]].name;
    int found = -1;
    for (size_t i = 0; i < src->spec.itmon.size(); i++)
      if (src->spec.seg[src->spec.itmon[i]].name == name) found = (int)i;
    double *d = dst->h.x + j * (size_t)M;
    if (found >= 0)
      HIP_TRY(hipMemcpyAsync(d, src->h.x + found * (size_t)M, M * sizeof(double), hipMemcpyDeviceToDevice,
                             dst->stream));
    else
      HIP_TRY(hipMemsetAsync(d, 0, M * sizeof(double), dst->stream));
  }
  if (dst->spec.charged) {  // the potential block (Solve_scf::Guess src/solve_scf.cpp:519-521)
    double *d = dst->h.x + dst->spec.itmon.size() * (size_t)M;
    if (src->spec.charged)
      HIP_TRY(hipMemcpyAsync(d, src->h.x + src->spec.itmon.size() * (size_t)M, M * sizeof(double),
                             hipMemcpyDeviceToDevice, dst->stream));
    else
      HIP_TRY(hipMemsetAsync(d, 0, M * sizeof(double), dst->stream));
  }
  HIP_TRY(hipStreamSynchronize(dst->stream));
  return SCF_OK;
}

}  // extern "C"

// carry map of `dst` from `src` by segment name (initial_guess : previous_result)
static int set_carry_map(scf_system *dst, const scf_system *src) {
  bool changed = false;
  for (size_t j = 0; j < dst->spec.itmon.size(); j++) {
    const std::string &name = dst->spec.seg[dst->spec.itmon[j]].name;
    int found = -1;
    for (size_t i = 0; i < src->spec.itmon.size(); i++)
      if (src->spec.seg[src->spec.itmon[i]].name == name) found = (int)i;
    if (dst->h.carry_src[j] != found) changed = true;
    dst->h.carry_src[j] = found;
  }
  if (dst->spec.charged) {
    const int j = (int)dst->spec.itmon.size();
    const int found = src->spec.charged ? (int)src->spec.itmon.size() : -1;
    if (dst->h.carry_src[j] != found) changed = true;
    dst->h.carry_src[j] = found;
  }
  if (changed) return upload_solver(dst);
  return SCF_OK;
}

namespace nscf {

int prepare_jobs(scf_system *const *slots, int nslots, int chain, int *threads, size_t *lds) {
  for (int s = 0; s < nslots; s++)
    for (int c = 0; c < chain; c++) {
      scf_system *S = slots[(size_t)s * chain + c];
      if (!S) return fail(SCF_E_ARG, "null system");
      if (S->effective_mode() != SCF_MODE_FUSED) return fail(SCF_E_UNSUPPORTED, "job queue needs fused systems");
      scf_system *R = slots[c];  // slot 0 fixes the layout of position c
      if (S->h.nit != R->h.nit || S->h.M != R->h.M) return fail(SCF_E_ARG, "slots differ in field layout");
      if (c > 0) {
        scf_system *P = slots[(size_t)s * chain + c - 1];
        if (P->h.M != S->h.M) return fail(SCF_E_ARG, "chained systems differ in lattice size");
        int rc = set_carry_map(S, P);
        if (rc != SCF_OK) return rc;
      }
      HIP_TRY(hipStreamSynchronize(S->stream));
      *threads = std::max(*threads, S->h.threads);
      *lds = std::max(*lds, S->lds_bytes);
    }
  return SCF_OK;
}

void convert_job_records(const JobRecord *h, size_t n, unsigned long long t0, scf_job_result *results) {
  const double khz = wall_clock_khz();
  for (size_t i = 0; i < n; i++) {
    scf_job_result &r = results[i];
    r.iterations = h[i].iterations;
    r.status = h[i].status != DS_OK ? h[i].status : (h[i].converged ? SCF_OK : SCF_NOT_CONVERGED);
    r.fp32_iterations = h[i].fp32_iterations;
    r.residual = h[i].residual;
    r.free_energy = h[i].F;
    r.grand_potential = h[i].GP;
    r.solve_ms = (double)(h[i].t_end - h[i].t_begin) / khz;
    for (int q = 0; q < 4; q++) r.phase_ms[q] = (double)h[i].t_phase[q] / khz;
    r.start_ms = (double)(h[i].t_begin - t0) / khz;
  }
}

}  // namespace nscf

extern "C" {

int scf_solve_jobs(scf_system *const *slots, int nslots, int chain, const double *seeds_dev, int njobs,
                   scf_job_result *results, double *kernel_ms) {
  if (!slots || nslots <= 0 || chain <= 0 || !seeds_dev || njobs < 0 || !results)
    return fail(SCF_E_ARG, "bad argument");
  int threads = 64;
  size_t lds = 0;
  int rc = prepare_jobs(slots, nslots, chain, &threads, &lds);
  if (rc != SCF_OK) return rc;
  if (njobs == 0) return SCF_OK;
  scf_system *S0 = slots[0];
  hipStream_t st = S0->stream;
  const size_t nsys = (size_t)nslots * chain, nrec = (size_t)njobs * chain;
  const size_t bytes = 256 + nsys * sizeof(DevSys *) + nrec * sizeof(JobRecord);
  char *tmp = nullptr;
  HIP_TRY(hipMalloc((void **)&tmp, bytes));
  int *counter = (int *)tmp;
  DevSys **d_slots = (DevSys **)(tmp + 256);
  JobRecord *rec = (JobRecord *)(tmp + 256 + nsys * sizeof(DevSys *));
  std::vector<DevSys *> list(nsys);
  for (size_t i = 0; i < nsys; i++) list[i] = slots[i]->d;
  float ms = 0.f;
  hipError_t e = hipMemcpy(d_slots, list.data(), nsys * sizeof(DevSys *), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(counter, 0, sizeof(int));
  if (e == hipSuccess) e = hipEventRecord(S0->ev0, st);
  if (e == hipSuccess)
    e = launch_fused_jobs(d_slots, nslots, chain, seeds_dev, S0->h.nit * S0->h.M, njobs, counter, rec, threads,
                          lds, st);
  if (e == hipSuccess) e = hipEventRecord(S0->ev1, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, S0->ev0, S0->ev1);
  std::vector<JobRecord> h(nrec);
  if (e == hipSuccess) e = hipMemcpy(h.data(), rec, nrec * sizeof(JobRecord), hipMemcpyDeviceToHost);
  hipFree(tmp);
  if (e != hipSuccess) return hip_fail(e, "job queue");
  unsigned long long t0 = ~0ull;
  for (size_t i = 0; i < nrec; i++) t0 = h[i].t_begin < t0 ? h[i].t_begin : t0;
  convert_job_records(h.data(), nrec, t0, results);
  if (kernel_ms) *kernel_ms = ms;
  for (size_t i = 0; i < nsys; i++) slots[i]->have_results = 1;
  return rc;
}

int scf_solve_resident(scf_system *const *systems, int n, scf_solve_info *info) {
  if (!systems || n <= 0 || !info) return fail(SCF_E_ARG, "null argument");
  std::vector<int> fused;
  size_t lds = 0;
  int threads = 64;
  for (int i = 0; i < n; i++) {
    scf_system *S = systems[i];
    if (!S) return fail(SCF_E_ARG, "null system");
    HIP_TRY(hipStreamSynchronize(S->stream));
    if (S->newton()) continue;
    if (S->effective_mode() == SCF_MODE_FUSED) {
      fused.push_back(i);
      threads = std::max(threads, S->h.threads);
      lds = std::max(lds, S->lds_bytes);
    }
  }
  float fused_ms = 0.f;
  // one launch for all fused systems
  if (!fused.empty()) {
    scf_system *S0 = systems[fused[0]];
    hipStream_t st = S0->stream;
    std::vector<DevSys *> list;
    for (int i : fused) list.push_back(systems[i]->d);
    DevSys **d_list = S0->d_list;
    if (list.size() > 1) {
      HIP_TRY(hipMalloc((void **)&d_list, list.size() * sizeof(DevSys *)));
      HIP_TRY(hipMemcpy(d_list, list.data(), list.size() * sizeof(DevSys *), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipEventRecord(S0->ev0, st));
    hipError_t le = launch_fused(d_list, (int)list.size(), FUSED_SOLVE, threads, lds, st);
    hipError_t re = hipEventRecord(S0->ev1, st);
    hipError_t se = hipStreamSynchronize(st);
    if (list.size() > 1) hipFree(d_list);
    if (le != hipSuccess) return hip_fail(le, "launch_fused");
    if (re != hipSuccess) return hip_fail(re, "hipEventRecord");
    if (se != hipSuccess) return hip_fail(se, "hipStreamSynchronize");
    HIP_TRY(hipEventElapsedTime(&fused_ms, S0->ev0, S0->ev1));
  }
  for (int i = 0; i < n; i++) {
    scf_system *S = systems[i];
    if (S->newton()) {
      const int method = S->spec.solver.method;
      int rc = method == SOLVER_LBFGS ? lbfgs_solve(S, &info[i])
                                      : (method == SOLVER_BRR ? brr_solve(S, &info[i]) : newton_solve(S, &info[i]));
      if (rc != SCF_OK) return rc;
    } else if (S->effective_mode() == SCF_MODE_GRID) {
      HIP_TRY(hipEventRecord(S->ev0, S->stream));
      int rc = grid_solve(S, &info[i]);
      if (rc != SCF_OK) return rc;
      HIP_TRY(hipEventRecord(S->ev1, S->stream));
      HIP_TRY(hipStreamSynchronize(S->stream));
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, S->ev0, S->ev1));
      info[i].kernel_ms = ms;
      info[i].solve_ms = ms;
    } else {
      DevScalars sc;
      int rc = read_scalars(S, sc);
      if (rc != SCF_OK) return rc;
      info[i].iterations = sc.iterations;
      info[i].residual = sc.residual;
      info[i].status = sc.status != DS_OK ? sc.status : (sc.converged ? SCF_OK : SCF_NOT_CONVERGED);
      info[i].kernel_ms = fused_ms;
      info[i].solve_ms = (double)(sc.t_end - sc.t_begin) / wall_clock_khz();
      info[i].fp32_iterations = sc.fp32_iterations;
    }
    S->have_results = info[i].status >= 0;
  }
  return SCF_OK;
}

int scf_solve_batch(scf_system *const *systems, int n, double *const *x, scf_solve_info *info) {
  if (!systems || n <= 0 || !x || !info) return fail(SCF_E_ARG, "null argument");
  for (int i = 0; i < n; i++) {
    if (!systems[i] || !x[i]) return fail(SCF_E_ARG, "null system or field buffer");
    int rc = scf_upload_fields(systems[i], x[i]);
    if (rc != SCF_OK) return rc;
  }
  int rc = scf_solve_resident(systems, n, info);
  if (rc != SCF_OK) return rc;
  for (int i = 0; i < n; i++) {
    rc = scf_download_fields(systems[i], x[i]);
    if (rc != SCF_OK) return rc;
  }
  return SCF_OK;
}

int scf_get_results(const scf_system *S, scf_results *out) {
  if (!S || !out) return fail(SCF_E_ARG, "null argument");
  DevScalars sc;
  int rc = read_scalars(S, sc);
  if (rc != SCF_OK) return rc;
  std::memset(out, 0, sizeof(*out));
  out->free_energy = sc.F;
  out->grand_potential = sc.GP;
  out->residual = sc.residual;
  out->n_mol = S->h.nmol;
  out->n_seg = S->h.nseg;
  for (int i = 0; i < S->h.nmol; i++) {
    out->GN[i] = sc.GN[i];
    out->n[i] = sc.n[i];
    out->theta[i] = sc.theta[i];
    out->phibulk[i] = sc.phibulk[i];
    out->mu[i] = sc.Mu[i];
  }
  for (int k = 0; k < S->h.nseg; k++) out->seg_phibulk[k] = sc.seg_phibulk[k];
  return SCF_OK;
}

int scf_get_profiles(const scf_system *S, double *seg_phi, double *mol_phi) {
  if (!S) return fail(SCF_E_ARG, "null system");
  const size_t M = S->h.M;
  if (seg_phi) HIP_TRY(hipMemcpy(seg_phi, S->h.phi, S->h.nseg * M * sizeof(double), hipMemcpyDeviceToHost));
  if (mol_phi) HIP_TRY(hipMemcpy(mol_phi, S->h.mol_phitot, S->h.nmol * M * sizeof(double), hipMemcpyDeviceToHost));
  return SCF_OK;
}

int scf_get_electrostatics(const scf_system *S, double *psi, double *q, double *eps) {
  if (!S) return fail(SCF_E_ARG, "null system");
  if (!S->spec.charged) return fail(SCF_E_ARG, "the system carries no charge");
  const size_t bytes = S->h.M * sizeof(double);
  if (psi) HIP_TRY(hipMemcpy(psi, S->h.psi, bytes, hipMemcpyDeviceToHost));
  if (q) HIP_TRY(hipMemcpy(q, S->h.q, bytes, hipMemcpyDeviceToHost));
  if (eps) HIP_TRY(hipMemcpy(eps, S->h.eps, bytes, hipMemcpyDeviceToHost));
  return SCF_OK;
}

int scf_buffer_alloc(size_t bytes, void **dev) {
  if (!dev || bytes == 0) return fail(SCF_E_ARG, "bad argument");
  hipError_t e = hipMalloc(dev, bytes);
  if (e != hipSuccess) return fail(SCF_E_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed");
  return SCF_OK;
}

int scf_buffer_upload(void *dev, const void *host, size_t bytes) {
  if (!dev || !host) return fail(SCF_E_ARG, "null argument");
  HIP_TRY(hipMemcpy(dev, host, bytes, hipMemcpyHostToDevice));
  return SCF_OK;
}

int scf_buffer_free(void *dev) {
  if (dev) HIP_TRY(hipFree(dev));
  return SCF_OK;
}

int scf_device_info(int *compute_units, size_t *hbm_bytes) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t p;
  HIP_TRY(hipGetDeviceProperties(&p, dev));
  if (compute_units) *compute_units = p.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = p.totalGlobalMem;
  return SCF_OK;
}

namespace {
// Plain streaming copy, 16 bytes per thread per trip (HBM counter calibration).
__global__ void __launch_bounds__(256) k_stream_copy(double2 *__restrict__ dst, const double2 *__restrict__ src,
                                                      size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace

int scf_device_copy(void *dst, const void *src, size_t bytes) {
  if (!dst || !src || bytes % 16) return fail(SCF_E_ARG, "bad argument (16-byte multiples)");
  const size_t n = bytes / 16;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, 0, (double2 *)dst, (const double2 *)src, n);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  return SCF_OK;
}

int scf_device_synchronize(void) {
  HIP_TRY(hipDeviceSynchronize());
  return SCF_OK;
}

int scf_get_phase_ms(const scf_system *S, double *ms) {
  if (!S || !ms) return fail(SCF_E_ARG, "null argument");
  DevScalars sc;
  if (read_scalars(S, sc) != SCF_OK) return SCF_E_HIP;
  for (int p = 0; p < 4; p++) ms[p] = (double)sc.t_phase[p] / wall_clock_khz();
  return SCF_OK;
}

int scf_propagate_chain(scf_system *S, const double *G1, int N, double *out) {
  if (!S || !G1 || !out || N < 1) return fail(SCF_E_ARG, "bad argument");
  const size_t M = S->h.M;
  double *dG1 = nullptr, *dout = nullptr;
  HIP_TRY(hipMalloc((void **)&dG1, M * sizeof(double)));
  HIP_TRY(hipMalloc((void **)&dout, (size_t)N * M * sizeof(double)));
  HIP_TRY(hipMemset(dout, 0, (size_t)N * M * sizeof(double)));
  HIP_TRY(hipMemcpy(dG1, G1, M * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(launch_propagate_chain(S->d, dG1, N, dout, S->stream));
  HIP_TRY(hipStreamSynchronize(S->stream));
  HIP_TRY(hipMemcpy(out, dout, (size_t)N * M * sizeof(double), hipMemcpyDeviceToHost));
  hipFree(dG1);
  hipFree(dout);
  return SCF_OK;
}

}  // extern "C"
