// Kernel launchers of the SCF hot path (host-callable).
#pragma once
#include <hip/hip_runtime.h>

#include "scf_device.h"

namespace nscf {

// FUSED_SLICE: a solve that stops after a device-clock budget with its state
// saved (phase 1) and resumes on the next launch, until it finishes (phase 2).
enum FusedMode : int { FUSED_RESIDUAL = 0, FUSED_SOLVE = 1, FUSED_SLICE = 2 };

// One workgroup per system: residual evaluation or full DIIS solve.
// lds_bytes > 0 stages the propagator rows in LDS (2 rows of size*M doubles).
hipError_t launch_fused(DevSys *const *d_systems, int nsys, int mode, int threads, size_t lds_bytes,
                        hipStream_t stream, unsigned long long budget_ticks = 0, int *phase_out = nullptr);
// Persistent job queue over nslots slots of `chain` systems (see k_fused_jobs).
hipError_t launch_fused_jobs(DevSys *const *slots, int nslots, int chain, const double *seeds, int seed_stride,
                             int njobs, int *counter, JobRecord *rec, int threads, size_t lds_bytes,
                             hipStream_t stream);
// Static LDS of the fused kernel (bytes), to size the dynamic part.
size_t fused_static_lds();

// Multi-workgroup path for large lattices (one launch per phase / per
// propagation step).  Host-driven; see scf_grid.hip.
struct GridPlan;
hipError_t grid_residual(const DevSys &h, DevSys *d, double *d_partial, int nblocks, hipStream_t stream);
hipError_t grid_thermo(const DevSys &h, DevSys *d, double *d_partial, int nblocks, hipStream_t stream);
hipError_t grid_diis_step(const DevSys &h, DevSys *d, double *d_partial, int nblocks, int k, int kd,
                          hipStream_t stream);
hipError_t grid_diis_combine(const DevSys &h, DevSys *d, int nblocks, int k, int kd, hipStream_t stream);
hipError_t grid_axpy_first(const DevSys &h, DevSys *d, int nblocks, hipStream_t stream);
hipError_t grid_copy_x0(const DevSys &h, DevSys *d, int nblocks, hipStream_t stream);
// fp32: chains with FP32 end points and the molecules' gauge shifts (the
// mixed-precision phase of a solve; 6-point 3-gradient steps only).  With
// precision MIXED every evaluation also refreshes the gauge shifts from GN.
hipError_t grid_residual_body(const DevSys &h, DevSys *d, double *d_partial, int nblocks, hipStream_t stream,
                              bool fp32 = false);
hipError_t grid_init(DevSys *d, hipStream_t stream);
// One DIIS iteration with all indices on the device (capturable in a hipGraph).
hipError_t grid_iteration(const DevSys &h, DevSys *d, double *d_partial, int nblocks, hipStream_t stream,
                          bool fp32 = false);
// Restarts the DIIS history (k_diis = 0) and clears the status, on the device.
hipError_t grid_diis_restart(DevSys *d, hipStream_t stream);

// Plain chain propagation for unit tests: out[s*M+i] = G_s(i) for a field
// whose Boltzmann weight is G1 (Lattice::propagate repeated N-1 times).
hipError_t launch_propagate_chain(DevSys *d, const double *G1, int N, double *out, hipStream_t stream);

}  // namespace nscf
