// Job queue without device allocation (scf_solve_jobs_noalloc).
//
// scf_solve_jobs allocates its scratch (the queue counter, the device list
// of slot systems and one record per solve) with hipMalloc on every call.
// This variant takes that scratch from `scratch`, a fixed array in device
// memory that the library holds for the life of the process, so callers
// that may not allocate device memory (tight serving loops, code that
// forbids hipMalloc) can run the same persistent launch.  Everything else
// -- the systems' arenas, the caller's seed fields and result buffer -- is
// the caller's, as for scf_solve_jobs.
//
// The array's name is a host-side handle, not a device address: its device
// address is looked up once per device with hipGetSymbolAddress and only
// that address is handed to kernels and copies.  Each GPU holds its own copy
// at its own address.
#include <hip/hip_runtime.h>

#include <mutex>

#include "scf_internal.h"

namespace nscf {
namespace {

constexpr size_t kJobScratchBytes = size_t(4) << 20;  // 4 MiB per device
constexpr size_t kSlotListOffset = 256;               // after the queue counter
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kJobScratchBytes];

std::mutex g_scratch_mu[kMaxDevices];
void *g_scratch_ptr[kMaxDevices];  // device address of `scratch`, per device

__global__ void k_zero_counter(int *counter) { *counter = 0; }

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Device address of this device's copy of `scratch` (looked up on the first
// call on the device, with that device current).  Called with the device's
// mutex held.
int scratch_address(int dev, void **out) {
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SCF_OK;
}

}  // namespace

size_t job_scratch_bytes(size_t nsys, size_t nrec) {
  return kSlotListOffset + align256(nsys * sizeof(DevSys *)) + nrec * sizeof(JobRecord);
}

}  // namespace nscf

using namespace nscf;

extern "C" int scf_solve_jobs_noalloc(scf_system *const *slots, int nslots, int chain, const double *seeds_dev,
                                      int njobs, scf_job_result *results, double *kernel_ms) {
  if (!slots || nslots <= 0 || chain <= 0 || !seeds_dev || njobs < 0 || !results)
    return fail(SCF_E_ARG, "bad argument");
  const size_t nsys = (size_t)nslots * chain, nrec = (size_t)njobs * chain;
  const size_t need = job_scratch_bytes(nsys, nrec);
  if (need > kJobScratchBytes)
    return fail(SCF_E_ARG, "job queue needs " + std::to_string(need) + " bytes of scratch; the fixed scratch holds " +
                               std::to_string(kJobScratchBytes) + " (fewer jobs or slots per call)");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(SCF_E_ARG, "device index out of range");
  // one call at a time per device in this process: every call shares the
  // device's scratch
  std::lock_guard<std::mutex> lk(g_scratch_mu[dev]);
  void *base = nullptr;
  int rc = scratch_address(dev, &base);
  if (rc != SCF_OK) return rc;
  int threads = 64;
  size_t lds = 0;
  rc = prepare_jobs(slots, nslots, chain, &threads, &lds);
  if (rc != SCF_OK) return rc;
  if (njobs == 0) return SCF_OK;
  scf_system *S0 = slots[0];
  hipStream_t st = S0->stream;
  char *sbase = (char *)base;
  int *counter = (int *)sbase;
  DevSys **d_slots = (DevSys **)(sbase + kSlotListOffset);
  JobRecord *rec = (JobRecord *)(sbase + kSlotListOffset + align256(nsys * sizeof(DevSys *)));
  // device list of the slot systems, staged through a stack buffer
  DevSys *chunk[256];
  for (size_t i = 0; i < nsys; i += 256) {
    const size_t n = nsys - i < 256 ? nsys - i : 256;
    for (size_t k = 0; k < n; k++) chunk[k] = slots[i + k]->d;
    HIP_TRY(hipMemcpyAsync(d_slots + i, chunk, n * sizeof(DevSys *), hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  // the counter starts at zero (a fill kernel: the previous call left it at
  // njobs + nslots)
  hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, st, counter);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(S0->ev0, st));
  HIP_TRY(launch_fused_jobs(d_slots, nslots, chain, seeds_dev, S0->h.nit * S0->h.M, njobs, counter, rec, threads,
                            lds, st));
  HIP_TRY(hipEventRecord(S0->ev1, st));
  HIP_TRY(hipStreamSynchronize(st));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, S0->ev0, S0->ev1));
  // records back through a stack buffer: the earliest start first, then the
  // conversion
  JobRecord h[64];
  unsigned long long t0 = ~0ull;
  for (size_t i = 0; i < nrec; i += 64) {
    const size_t n = nrec - i < 64 ? nrec - i : 64;
    HIP_TRY(hipMemcpyAsync(h, rec + i, n * sizeof(JobRecord), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (size_t k = 0; k < n; k++) t0 = h[k].t_begin < t0 ? h[k].t_begin : t0;
  }
  for (size_t i = 0; i < nrec; i += 64) {
    const size_t n = nrec - i < 64 ? nrec - i : 64;
    HIP_TRY(hipMemcpyAsync(h, rec + i, n * sizeof(JobRecord), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    convert_job_records(h, n, t0, results + i);
  }
  if (kernel_ms) *kernel_ms = ms;
  for (size_t i = 0; i < nsys; i++) slots[i]->have_results = 1;
  return SCF_OK;
}
