// Host-side internals shared by the C ABI (scf_api.hip) and the request
// server (scf_server.hip): the system handle and the helpers that build it.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/namics_scf.h"
#include "scf_device.h"
#include "scf_kernels.h"
#include "scf_spec.h"

// One SCF system: its parsed input, device description and HBM arena.
struct scf_system {
  nscf::SystemSpec spec;
  nscf::DevSys h;              // host copy holding device pointers
  nscf::DevSys *d = nullptr;   // device copy
  nscf::DevSys **d_list = nullptr;
  char *arena = nullptr;
  size_t arena_bytes = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraphExec_t graph_exec = nullptr;    // captured grid-path iteration
  hipGraphExec_t graph_exec32 = nullptr;  // the same with FP32 propagators (mixed precision)
  bool use_graph = true;
  int mode = SCF_MODE_AUTO;
  int tuned_mode = SCF_MODE_AUTO;  // the autotuner's choice for this shape (used under AUTO)
  int nb = 0;            // workgroups of the grid path
  int gf_rows = 0;       // end-point rows (of M sites) the arena holds: chain length x components, or
                         // fewer with the input's save_memory
  int m_alloc = 0;       // DIIS history the arena was sized for
  size_t lds_bytes = 0;  // dynamic LDS the fused path asks for (0: rows in HBM)
  bool have_results = false;
  bool owns_stream = true;  // false when a server lends its stream
  char *region = nullptr;   // arena placed in a caller's device region (not owned)
  std::vector<nscf::BranchPlan> plans;  // row schedules of the branched molecules (h.branch points here)

  bool newton() const { return spec.solver.method != nscf::SOLVER_DIIS; }
  // destroys the captured grid-path iterations (both precisions)
  void drop_graphs() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph_exec32) (void)hipGraphExecDestroy(graph_exec32);
    graph_exec = graph_exec32 = nullptr;
  }
  bool branched() const {
    for (auto &m : spec.mol)
      if (m.scheduled()) return true;
    return false;
  }

  // 2- and 3-gradient systems whose results depend on ghost values the
  // reference computes over its flattened arrays: charged systems (the
  // potential reads densities on ghosts) and surface boundaries (a surface
  // ghost keeps its own value).  They run one workgroup per system with the
  // chains over every site (chain_literal).
  bool has_ring() const {
    for (auto &m : spec.mol)
      if (m.ring) return true;
    return false;
  }
  bool literal() const {
    if (spec.lat.gradients < 2) return false;
    if (spec.charged) return true;
    for (int d = 0; d < 6; d++)
      if (spec.lat.bc[d] == nscf::BC_SURFACE) return true;
    return false;
  }

  int effective_mode() const {
    if (mode != SCF_MODE_AUTO) return mode;
    // branched molecules and dendrimers run their row schedules on the grid path
    if (branched()) return SCF_MODE_GRID;
    // charged systems solve the potential in the one-workgroup kernels
    if (spec.charged || literal() || has_ring()) return SCF_MODE_FUSED;
    // checkpointed end-point storage runs on the strip chains or the grid path
    if (h.ckpt > 1 && h.strip_sx == 0) return SCF_MODE_GRID;
    if (tuned_mode != SCF_MODE_AUTO) return tuned_mode;
    bool markov2 = false;  // second-order chains run on the fused path only
    for (auto &m : spec.mol) markov2 = markov2 || m.markov == 2;
    return spec.lat.gradients == 3 && !markov2 ? SCF_MODE_GRID : SCF_MODE_FUSED;
  }
};

namespace nscf {

int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *where);

#define HIP_TRY(expr)                                 \
  do {                                                \
    hipError_t _e = (expr);                           \
    if (_e != hipSuccess) return hip_fail(_e, #expr); \
  } while (0)

// Where build() places a system's arena: its own allocation (default), a
// caller's device region of `capacity` bytes, or nowhere (dry: compute the
// footprint only, into scf_system::arena_bytes).
struct Placement {
  char *region = nullptr;
  size_t capacity = 0;
  bool dry = false;
};

int validate(const SystemSpec &s, std::string &why);
int build(scf_system *S, const Placement *pl = nullptr);
int upload_solver(scf_system *S);
int read_scalars(const scf_system *S, DevScalars &sc);
int grid_solve(scf_system *S, scf_solve_info *info);
// Newton-Krylov solve of a pseudohessian / hessian input (scf_newton.hip)
int newton_solve(scf_system *S, scf_solve_info *info);
// LBFGS solve of an input with `method : LBFGS` (scf_newton.hip)
int lbfgs_solve(scf_system *S, scf_solve_info *info);
// Broyden rank-reduction solve of an input with `method : BRR` (scf_newton.hip)
int brr_solve(scf_system *S, scf_solve_info *info);
double wall_clock_khz();
size_t lds_limit();
void plan_layout(scf_system *S, int want_sx);
int tuned_strip(scf_system *S);
std::string shape_key(const scf_system *S);
// End-point rows kept with checkpoint interval K (K <= 1: all N rows): the
// checkpoints 0, K, 2K, ... plus a buffer for the rows of one segment.
int ckpt_rows(int N, int K);
// Checkpoint interval for the input's save_memory (0: every row): about
// sqrt(2 N) rows kept for the longest chain that asks for it, as the
// reference plans them (src/molecule.cpp:174-191), on a uniform grid.
int save_memory_interval(const SystemSpec &s);
// SCF_OK when the system's options (precision, checkpoints) run on the
// given layout, else SCF_E_UNSUPPORTED with the reason.
int check_layout(const scf_system *S, int mode, int strip_sx);
// Job queue: check that `slots` (nslots x chain systems) can share one
// launch, set the carry maps between chain positions, and return the
// workgroup size and dynamic LDS the launch needs.
int prepare_jobs(scf_system *const *slots, int nslots, int chain, int *threads, size_t *lds);
// Device job records -> caller results; t0 = the earliest solve start.
void convert_job_records(const JobRecord *h, size_t n, unsigned long long t0, scf_job_result *results);

}  // namespace nscf
