/*
 * namics_scf.h — C ABI of the MI355X lattice-SCF hot path.
 *
 * One `scf_system` is one Scheutjens-Fleer SCF problem resident on the GPU:
 * its fields, Boltzmann weights, end-point distributions, densities and
 * DIIS history live in one device arena owned by the library.  Callers own
 * every host buffer they pass; no pointer ownership crosses this boundary.
 * All functions return an int status (SCF_OK = 0, negative = error,
 * SCF_NOT_CONVERGED = 1); no C++ exception crosses the ABI.
 *
 * The reference (leermakers/Namics) has no plugin API; its seams are C++
 * virtuals.  Each entry point names the reference interface it replaces, so a
 * maintainer can route the reference's own classes through this library.
 */
#ifndef NAMICS_SCF_H
#define NAMICS_SCF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct scf_system scf_system;

/* status codes; the negative ones 1..4 mirror the reference's `throw int`
 * (src/system.cpp:2398-2402, src/sfnewton.cpp:808-809, src/tools.cpp:1225-1228,
 * src/system.cpp:2521-2525) */
enum {
  SCF_OK = 0,
  SCF_NOT_CONVERGED = 1,     /* iteration limit reached (src/sfnewton.cpp:614-617) */
  SCF_E_GN = -1,             /* partition function GN <= 0, or not finite (FP64 overflow) */
  SCF_E_NAN = -2,            /* NaN in the DIIS Gram matrix */
  SCF_E_SVD = -3,            /* NaN in the SVD */
  SCF_E_PHIB = -4,           /* solvent bulk volume fraction < 0 */
  SCF_E_ARG = -10,           /* bad argument */
  SCF_E_PARSE = -11,         /* input text rejected */
  SCF_E_UNSUPPORTED = -12,   /* valid input outside this build's scope */
  SCF_E_OOM = -13,           /* device allocation failed */
  SCF_E_HIP = -14            /* HIP runtime error */
};

/* execution strategy for one system */
enum {
  SCF_MODE_AUTO = 0,   /* fused for 1-2 gradients, grid for 3 gradients */
  SCF_MODE_FUSED = 1,  /* one workgroup runs the whole residual / solve */
  SCF_MODE_GRID = 2    /* one grid-wide launch per phase and propagation step */
};

#define SCF_MAX_MOL 8
#define SCF_MAX_SEG 8

typedef struct {
  int iterations;
  int status;      /* SCF_OK, SCF_NOT_CONVERGED or a negative error */
  double residual; /* ||g||_2 (or max|g| with stop_criterion max_of_element_of_|g|) */
  double kernel_ms; /* GPU time of the launch(es) that solved this system (HIP events) */
  double solve_ms;  /* this system's own solve time on the device (device wall clock;
                       equals kernel_ms for the grid path) */
  int fp32_iterations; /* iterations evaluated with FP32 propagators (SCF_PREC_MIXED) */
} scf_solve_info;

typedef struct {
  double free_energy;      /* System::GetFreeEnergy (src/system.cpp:2809-3029) */
  double grand_potential;  /* System::GetGrandPotential (src/system.cpp:3038-3244) */
  double residual;
  int n_mol;
  int n_seg;
  double GN[SCF_MAX_MOL];       /* single-chain partition functions */
  double n[SCF_MAX_MOL];        /* number of chains */
  double theta[SCF_MAX_MOL];    /* weighted sum of the molecule density */
  double phibulk[SCF_MAX_MOL];
  double mu[SCF_MAX_MOL];       /* System::CreateMu (src/system.cpp:3246-3428) */
  double seg_phibulk[SCF_MAX_SEG];
} scf_results;

/* Parse `.in` text (the reference's input format) at start block `start`
 * (1-based) and allocate the system on the current HIP device.
 * Compositions may be linear, `(A)10(B)10`, branched with square
 * brackets, `(C)16(N)1[(N)1][(N)1](N)1`, dendrimers,
 * `@dend(A,(B)2,3; A,(B)2,3)` or `@dend(A,(B)6,3,(B)5(C)1,1)`, or combs,
 * `@comb((A)5; A,(B)3,(A)6,4; (A)7)` (the last three: first-order chains
 * without save_memory, run on the grid path); aliases, repeat groups and
 * @water are refused with SCF_E_PARSE and a message naming them.  Charged
 * systems (`mon : <name> : valence`, `epsilon`, `lat : <name> : bondlength`,
 * a molecule with freedom `neutralizer`) solve for the electrostatic
 * potential as one more field block on the fused path, on 1-gradient
 * lattices and on simple-cubic 2- and 3-gradient lattices with first-order
 * chains (one workgroup per system), fixed surface potentials on frozen
 * segments (`e.psi0/kT`) included, as the reference computes them; charged
 * branched molecules are refused.  2-gradient surface boundaries run on the
 * same one-workgroup kernels.  A linear molecule may hold one tagged
 * segment (`mon : <name> : freedom : tagged`, `tagged_range`): its weight
 * and density are the mask, and the molecule holds one chain per tagged
 * site.  Linear rings (`mol : <name> : ring : true`, first and last blocks
 * one segment of the same type) run with first-order chains on lattices
 * with fjc 1 (not with 2- or 3-gradient surfaces or charges).
 * Replaces: Input + Lattice/Segment/Molecule/System/Solve_scf::CheckInput and
 * AllocateMemory (src/namics.cpp:153-417, src/lattice.cpp:267-601), the
 * composition parser Molecule::Decomposition / GenerateTree / Interpret
 * (src/molecule.cpp:1155-1280,1383-1767). */
int scf_system_create(const char *input_text, int start, scf_system **out, char *err, size_t errlen);

void scf_system_destroy(scf_system *sys);

/* JSON description: names of segments, molecules and field blocks, lattice
 * sizes (M, MX, MY, MZ), nvar, solver settings, and the current end-point
 * checkpoint interval ("checkpoint", 0 = every row kept); a branched
 * molecule also lists its blocks, generations and row schedule ("branch").
 * Returns bytes needed. */
int scf_system_describe(const scf_system *sys, char *buf, size_t len);

/* The same description for input text without creating a system (no
 * device work): the field layout a caller needs to prepare starting fields. */
int scf_describe_input(const char *input_text, int start, char *buf, size_t len);

/* Number of field values (n_field_blocks * M), the length of x and g. */
int64_t scf_system_nvar(const scf_system *sys);

/* Lattice sites M including ghost layers. */
int64_t scf_system_sites(const scf_system *sys);

/* Device bytes held by this system (its HBM footprint). */
size_t scf_system_device_bytes(const scf_system *sys);

/* Force an execution strategy (SCF_MODE_*). */
int scf_system_set_mode(scf_system *sys, int mode);

/* propagator precision of a solve */
enum {
  SCF_PREC_FP64 = 0,  /* every evaluation in FP64 (the reference's arithmetic) */
  SCF_PREC_MIXED = 1  /* FP32 propagators while ||g|| > switch, then FP64 */
};

/* Precision of the solves of this system.  SCF_PREC_MIXED evaluates the
 * propagator recursion in FP32 (per-molecule gauge shift, FP64 composition
 * sums, partition functions, DIIS and thermodynamics) until the residual
 * drops below `switch_residual` (> the tolerance) or stops falling for 25
 * evaluations, then continues in FP64 to the tolerance with a fresh DIIS
 * history, so the converged state is an FP64 one.  Applies to the
 * 2-gradient strip path and to the 6-point simple-cubic 3-gradient grid path
 * (first-order linear chains); elsewhere it returns SCF_E_UNSUPPORTED.  No
 * counterpart in the reference (its arithmetic is FP64 throughout). */
int scf_system_set_precision(scf_system *sys, int mode, double switch_residual);

/* Checkpointed propagators (the reference's `save_memory`, src/molecule.cpp
 * :2300-2446): the forward pass keeps only every `interval`-th end-point
 * distribution; the backward pass recomputes the rows between two
 * checkpoints from the earlier one just before it composes them.  Densities,
 * partition functions and every iterate are bit-identical to full storage;
 * the end-point traffic falls about interval-fold for up to 50% more
 * propagation steps.  interval 0 (or 1) stores every row.  Runs on the
 * 2-gradient strip path and, for first-order chains, on the grid path (a
 * system off the strip path moves to the grid path); second-order chains
 * get SCF_E_UNSUPPORTED.  An input with `mol : <name> : save_memory : true`
 * starts with the interval that keeps about sqrt(2N) rows of the longest
 * such chain and an arena sized for them; a larger storage than that arena
 * holds is refused with SCF_E_ARG.  The schedule is uniform rather than the
 * reference's per-block round(sqrt(2 N_b)) plan, which changes no output. */
int scf_system_set_checkpoint(scf_system *sys, int interval);

/* Grid-path iteration replay: on (default), the first grid-path DIIS
 * iteration of a system is captured into a hipGraph and replayed; off, every
 * iteration launches its kernels one by one (same kernels, same results).
 * No counterpart in the reference. */
int scf_system_set_graph(scf_system *sys, int on);

/* Launch-layout autotuner.  Times `evals` residual evaluations (at the
 * system's current device fields) for every layout that applies to the
 * lattice -- on 2-gradient lattices the register-strip widths SX = 8 and 4
 * and the generic site-strided chain; on 3-gradient lattices one workgroup
 * per system against one launch per phase -- keeps the fastest for this
 * system, and remembers it for the lattice shape, so systems built later in
 * the process with the same shape start from it.  Every layout computes the
 * same residual (to rounding).  A layout the system's options exclude (FP32
 * propagators need a strip layout, checkpointing a strip layout or the grid
 * path; an explicit scf_system_set_mode) is not tried.  No counterpart in the reference, which
 * has one code path per lattice. */
#define SCF_TUNE_MAX 4
typedef struct {
  int candidates;                   /* layouts timed */
  int chosen;                       /* index of the fastest */
  int mode[SCF_TUNE_MAX];           /* SCF_MODE_FUSED / SCF_MODE_GRID */
  int strip_sx[SCF_TUNE_MAX];       /* fused path: 8, 4, or 0 for the generic chain */
  double ms_per_eval[SCF_TUNE_MAX]; /* device time of one residual evaluation */
} scf_tune_report;
int scf_system_autotune(scf_system *sys, int evals, scf_tune_report *report);
/* Concurrency autotuner for systems solved together (one launch, one
 * workgroup per system, as scf_solve_resident and the job queue run them).
 * Times `evals` residual evaluations of all n systems in one launch for each
 * end-point storage that applies -- every row (interval 1) and, when every
 * system is on the 2-gradient strip path, every other row with the odd rows
 * recomputed on chip (interval 2, scf_system_set_checkpoint) -- and applies
 * the fastest to all n.  Alone, a system's propagation is bound by its own
 * LDS traffic and recomputing rows costs time; with one system per CU the
 * stored rows saturate HBM and halving them pays.  Results are bit-identical
 * either way.  No counterpart in the reference. */
typedef struct {
  int candidates;
  int chosen;
  int checkpoint[SCF_TUNE_MAX];      /* end-point row interval */
  double ms_per_launch[SCF_TUNE_MAX]; /* one residual evaluation of all n systems */
} scf_batch_tune_report;
int scf_autotune_batch(scf_system *const *systems, int n, int evals, scf_batch_tune_report *report);
/* Forget the layouts remembered per lattice shape (tests, benchmarks). */
void scf_autotune_forget(void);
/* The system's current launch layout: execution mode, strip width (0 when
 * not on the strip path) and threads per system. */
int scf_system_layout(const scf_system *sys, int *mode, int *strip_sx, int *threads);

/* ---- output and format layer (the reference's files) ----
 * The `kal` / `pro` requests of the system's input (`output : kal : ...`,
 * `kal : mol : pol : GN`, `pro : mon : A : phi`, names `*` expanded) are
 * written in the reference's formats.  kal quantities: sys free_energy,
 * grand_potential, start; newton iterations, residual, tolerance,
 * iterationlimit, m, delta_max, method, stop_criterion; mol GN, n, theta,
 * theta_exc, phibulk, Mu, chainlength, composition, freedom, Markov,
 * k_stiff (1-D also phi1, phiM, Dphi); mon freedom, valence, phibulk, theta;
 * lat gradients, volume, lattice_type, FJC_choices, offset_first_layer
 * (when > 0), bond_length, n_layers, n_layers_y, n_layers_z.  pro profiles: mon and mol phi.  A quantity the
 * library does not know prints "NiN" in a kal row and is left out of a pro
 * file, as the reference does for unknown names.
 *
 * Append this system's kal row to `path` (header first when the file is new
 * or `append` is 0).  Replaces: Output::WriteOutput for kal
 * (src/output.cpp:511-557). */
int scf_write_kal(scf_system *sys, const char *path, int append, const scf_solve_info *info);
/* Write the requested profiles.  Replaces: Output::WriteOutput for pro and
 * Lattice::PutProfiles (src/output.cpp:487-510, src/LGrad1.cpp:228-249,
 * src/LGrad2.cpp:185-205, src/LGrad3.cpp:81-102). */
int scf_write_pro(scf_system *sys, const char *path);
/* Write the first requested `vtk` profile (2 and 3 gradients; the reference
 * writes no 1-gradient vtk file).  Replaces: LGrad2::vtk / LGrad3::vtk
 * (src/LGrad2.cpp:159-183, src/LGrad3.cpp:48-79). */
int scf_write_vtk(scf_system *sys, const char *path);
/* Append the requested `vec` profiles to `path`, one line each: the name,
 * then every site's value (ghosts included) as %e.  Replaces:
 * Output::WriteOutput for vec (src/output.cpp:452-473). */
int scf_write_vec(scf_system *sys, const char *path);
/* Write the system's fields as a guess file.  Replaces: Lattice::StoreGuess
 * (src/lattice.cpp:971-1019). */
int scf_store_guess(scf_system *sys, const char *path);
/* Read a guess file and upload the field blocks it names (same lattice
 * size).  Replaces: Lattice::ReadGuess + Solve_scf::Guess
 * (src/lattice.cpp:926-968, src/solve_scf.cpp:489-524). */
int scf_read_guess(scf_system *sys, const char *path);
/* Run every start block of an input file like the reference program: initial
 * guess (file, previous_result or none), solve with the input's method, kal /
 * pro / final-guess files in the output folder (./output/ unless
 * `out_info : folder : path` is given, relative to the working directory).
 * A `var : <kind>-<name> : scan : <property>` block (mol theta, n, phibulk;
 * mon chi_X) solves every scan point in turn, each from the previous one.
 * A `var : mol-<name> : search : theta | n | phibulk` block with a target
 * (`var : sys-<name> : grand_potential | free_energy : value`, or
 * `var : mol-<other> : theta | n | phibulk | mu : value`) varies the searched
 * property, each solve from the previous one, until the target's error is
 * within newton's super_tolerance (10 x tolerance), by this library's own
 * secant / Illinois regula-falsi root finder, and writes the final state.
 * *n_solves counts the solves (start blocks times scan points), *n_converged
 * those that converged.  Replaces: the start and scan loops of main
 * (src/namics.cpp:163-737, Variate::PutVarScan src/variate.cpp:361-400,
 * Solve_scf::SuperIterate src/solve_scf.cpp:717-750). */
int scf_run_input_file(const char *path, int *n_solves, int *n_converged);

/* Override solver settings: tolerance, deltamax, DIIS history m, iteration
 * limit (values <= 0 keep the input's setting). */
int scf_system_set_solver(scf_system *sys, double tolerance, double deltamax, int m, int iterationlimit);

/* One residual evaluation g = F(x) at the host fields x (nvar values).
 * Replaces: Solve_scf::residuals -> System::Classical_residual
 * (src/solve_scf.cpp:901-906, src/system.cpp:2087-2189), including
 * Molecule::ComputePhi (src/mol_linear.cpp:131-149) and the thermodynamic
 * summary of System::CheckResults (src/system.cpp:2719-2808). */
int scf_residual(scf_system *sys, const double *x, double *g, double *residual);

/* Solve the SCF equations from the host fields x (updated in place with the
 * final fields), by the input's `newton : method`.
 * Replaces: Solve_scf::Solve (src/solve_scf.cpp:558-651) with
 *   - DIIS -> SFNewton::iterate_DIIS (src/sfnewton.cpp:1037-1130), restated
 *     on the device (same iterates up to rounding);
 *   - pseudohessian / hessian (the reference's default) -> SFNewton::iterate
 *     (src/sfnewton.cpp:636-747): solved by a Jacobian-free Newton-Krylov
 *     method of this library instead (GMRES on forward-difference J v,
 *     backtracking line search) that reaches the same fixed point by a
 *     different path; `iterations` counts its Newton steps;
 *   - LBFGS -> the reference's LBFGSpp-based solver (src/LBFGS.h,
 *     src/solve_scf.cpp:526-556,630-644), its steps restated over device
 *     residuals (two-loop recursion, trust-radius step, m = 6 pairs by
 *     default);
 *   - BRR -> SFNewton::iterate_BRR (src/sfnewton.cpp:929-1035), Broyden's
 *     inverse update as m = 10 rank-one terms with the reference's rank
 *     reduction (QR, SVD), restated over device residuals.
 * The Newton, LBFGS and BRR solvers keep their vectors in host memory:
 * about 47, 2 m + 5 and 2 m + 9 vectors of nvar doubles (a 96^3 box with
 * three field blocks: about 1.1 GB, 0.38 GB and 0.66 GB with the default m).
 * The same rule holds for scf_solve_batch and scf_solve_resident (Newton,
 * LBFGS and BRR systems are solved one after another, DIIS systems in one launch);
 * the job queue and the request server solve every input by DIIS. */
int scf_solve(scf_system *sys, double *x, scf_solve_info *info);

/* Solve n independent systems in one launch (one workgroup per system);
 * x[i] are host field buffers updated in place, info[i] per system.
 * Systems in SCF_MODE_GRID are solved one after another. */
int scf_solve_batch(scf_system *const *systems, int n, double *const *x, scf_solve_info *info);

/* Device-resident fields: copy host fields in / out of the system's arena. */
int scf_upload_fields(scf_system *sys, const double *x);
int scf_download_fields(const scf_system *sys, double *x);

/* Set the device-resident fields to zero (the reference's default start). */
int scf_zero_fields(scf_system *sys);

/* Start `dst` from the fields of `src` (same lattice), block by segment
 * name, zero where `src` has no block of that name, device to device.
 * Replaces the chaining of `start` blocks with initial_guess
 * previous_result (Solve_scf::Guess, src/solve_scf.cpp:489-524). */
int scf_carry_fields(scf_system *dst, const scf_system *src);

/* Solve n systems from their device-resident fields, in one launch for the
 * fused ones; no host copies of fields.  info[i] per system. */
int scf_solve_resident(scf_system *const *systems, int n, scf_solve_info *info);

/* Result of one solve of a queued job (scf_solve_jobs). */
typedef struct {
  int iterations;
  int status;          /* SCF_OK, SCF_NOT_CONVERGED or a negative error */
  int fp32_iterations;
  double residual;
  double free_energy;
  double grand_potential;
  double solve_ms;     /* device time of this solve */
  double phase_ms[4];  /* of which: Boltzmann weights, propagator chains, densities + residual, the rest */
  double start_ms;     /* device time from the earliest solve start of the launch to this solve's start */
} scf_job_result;

/* Persistent job queue (parameter sweeps, many seeds of one problem).
 * `slots` holds nslots x chain systems: slot s owns systems
 * slots[s*chain .. s*chain+chain-1], all slots alike (same inputs up to
 * parameters; every system of chain position c has the same field layout).
 * Job j solves the chain on whichever slot takes it: position 0 from the
 * seed fields `seeds_dev + j*nvar0` (DEVICE memory, nvar0 = fields of
 * position 0), position c > 0 from position c-1's converged fields mapped
 * by segment name (initial_guess : previous_result).  One launch, one
 * workgroup per slot, jobs taken from a counter until none is left, so
 * uneven jobs balance across the GPU.  results: njobs x chain (host).
 * Replaces: running the reference once per parameter set / start block
 * (src/namics.cpp:407-470 `start` loop). */
int scf_solve_jobs(scf_system *const *slots, int nslots, int chain, const double *seeds_dev, int njobs,
                   scf_job_result *results, double *kernel_ms);

/* The job queue without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (serving loops
 * and harnesses that forbid hipMalloc).  Same arguments, kernels and results
 * as scf_solve_jobs, but the queue's scratch (counter, slot list, one record
 * per solve: 256 B + nslots*chain*8 B rounded up to 256 B + njobs*chain*88 B)
 * lives in a fixed 4 MiB array in device memory that the library holds on
 * every GPU it runs on until the process ends; larger calls are refused with
 * SCF_E_ARG before anything runs.  The call allocates nothing and returns
 * when the launch has finished.  All calls on one GPU share that GPU's
 * array: this library serialises its own calls per device, and calls from
 * several threads that bypass it (other copies of the library in the
 * process) would race on the array. */
int scf_solve_jobs_noalloc(scf_system *const *slots, int nslots, int chain, const double *seeds_dev, int njobs,
                           scf_job_result *results, double *kernel_ms);

/* Scalars of the last residual / solve. */
int scf_get_results(const scf_system *sys, scf_results *out);

/* Densities of the last residual / solve: segment densities (n_seg x M) and
 * molecule total densities (n_mol x M), x-major with ghost layers.
 * Replaces: the `pro` output of mon/mol phi (src/output.cpp:385-708). */
int scf_get_profiles(const scf_system *sys, double *seg_phi, double *mol_phi);

/* Electrostatic profiles of the last residual / solve of a charged system
 * (M each, x-major with ghost layers; any pointer may be NULL): the
 * dimensionless potential e psi / kT, the charge density q (valence-weighted
 * volume fractions over phi_tot) and the relative dielectric constant eps.
 * SCF_E_ARG when the system carries no charge.
 * Replaces: the `pro : sys : <name> : psi | q | eps` profiles
 * (System::PushOutput / GetPointer src/system.cpp:1610-1650). */
int scf_get_electrostatics(const scf_system *sys, double *psi, double *q, double *eps);

/* Device time of the last fused residual / solve by phase, in ms:
 * [0] Boltzmann weights (Segment::PrepareForCalculations), [1] propagator
 * chains with composition (Molecule::ComputePhi), [2] normalisation,
 * densities and residual (System::ComputePhis, Classical_residual),
 * [3] the rest of the solve (DIIS update and thermodynamics).
 * Profiling aid without a counterpart in the reference. */
int scf_get_phase_ms(const scf_system *sys, double *ms);

/* Unit-test entry: N-1 Markov-1 steps from G(1) = G1 (host arrays of M),
 * out = N x M end-point distributions.  Replaces Lattice::propagate
 * (src/lattice.h:133; src/LG1Planar.cpp:138-159, src/LGrad2.cpp:1053-1107,
 * src/LGrad3.cpp:673-743). */
int scf_propagate_chain(scf_system *sys, const double *G1, int N, double *out);

/* ---------------------------------------------------------------------
 * Request server: continuous batching of independent solves.
 * Requests are parsed and sized on submission and admitted first come,
 * first served while their HBM footprint fits in the server's pool (one
 * device allocation, sub-allocated, so finished requests return memory
 * without synchronising the device).  Admitted 1- and 2-gradient requests
 * are solved in time slices, one workgroup each: a launch runs the active
 * requests of one workgroup shape for a device-clock budget, unfinished
 * solves save their state and resume in the next launch together with newly
 * admitted requests (iteration-level continuous batching); several streams
 * carry launches at once.  3-gradient requests run the grid path in a
 * worker thread.  scf_server_progress() drives admission,
 * launches and completions without blocking; scf_server_results() hands
 * over finished requests.  Latency = device completion time of the
 * request's own solve (device clock mapped to host time) - submission time.
 * No counterpart in the reference (one system per process run). */
typedef struct scf_server scf_server;

typedef struct {
  long long id;
  int status;          /* SCF_OK, SCF_NOT_CONVERGED or a negative error */
  int iterations;
  int gradients;
  double residual;
  double free_energy;
  double grand_potential;
  double latency_ms;   /* submission -> end of the request's solve */
  double queue_ms;     /* submission -> admission (waiting for HBM) */
  double solve_ms;     /* the solve on the device */
  size_t device_bytes; /* HBM footprint */
} scf_served;

int scf_server_create(size_t pool_bytes, int streams, scf_server **out);
/* Device-clock budget of one time slice of the fused-path launches
 * (default 20 ms): the longest a newly admitted request waits for a start. */
int scf_server_set_slice(scf_server *srv, double slice_ms);
/* x0: starting fields (nx0 values, the system's nvar) or NULL for zeros. */
int scf_server_submit(scf_server *srv, const char *input_text, const double *x0, size_t nx0, long long *id);
int scf_server_progress(scf_server *srv, int *queued, int *running);
int scf_server_results(scf_server *srv, scf_served *out, int max, int *n);
void scf_server_destroy(scf_server *srv);

/* Device plumbing for callers without their own GPU runtime (the job
 * queue's seed fields): allocate / fill / free device memory, the current
 * device's compute-unit count and HBM size, and a device-wide
 * synchronisation.  No counterpart in the reference (CPU only). */
int scf_buffer_alloc(size_t bytes, void **dev);
int scf_buffer_upload(void *dev, const void *host, size_t bytes);
int scf_buffer_free(void *dev);
int scf_device_info(int *compute_units, size_t *hbm_bytes);
int scf_device_synchronize(void);
/* A plain streaming copy kernel of `bytes` (a multiple of 16) between device
 * buffers: reads and writes exactly `bytes`, the reference point for
 * hardware memory-traffic counters. */
int scf_device_copy(void *dst, const void *src, size_t bytes);

/* Last error message of this thread (static storage). */
const char *scf_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* NAMICS_SCF_H */
