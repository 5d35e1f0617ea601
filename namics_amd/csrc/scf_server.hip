// Request server: continuous batching of independent SCF solves on one GPU.
//
// Requests (input text + optional starting fields) are parsed and sized when
// they are submitted and wait in a FIFO until their HBM footprint fits in the
// server's pool (one device allocation, first-fit sub-allocation, so freeing
// a finished request never synchronises the device).  Admitted 1- and
// 2-gradient requests are solved in time slices (iteration-level continuous
// batching): a launch of the fused kernel runs every active request of one
// workgroup shape (one workgroup each) for a device-clock budget, after which
// each unfinished solve saves its state (DIIS ring and Gram matrix,
// iteration count, residual) in its scalars and leaves; the next launch
// resumes it together with the requests admitted in the meantime.  Several
// non-blocking streams carry launches of different shapes at once.  So a new
// request waits at most about one slice, long solves cannot hold others back,
// and a solve's arithmetic -- hence its result -- is the same as alone.
// 3-gradient requests run the grid path (host-driven DIIS) in worker threads
// (a few at once), each on its own stream.  A request's completion time comes from the device
// clock at the end of its own solve, converted to host time.
//
// No counterpart in the reference, which solves one system per process.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "scf_internal.h"

using namespace nscf;

namespace {

constexpr int kGridWorkers = 3;  // 3-gradient requests solved at once (each a grid-wide launch sequence)

double host_now() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

__global__ void k_read_clock(unsigned long long *out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = wall_clock64();
}

// First-fit allocator over [0, capacity) with coalescing.
class PoolAllocator {
 public:
  explicit PoolAllocator(size_t capacity) { free_[0] = capacity; }
  static constexpr size_t npos = ~size_t(0);
  size_t alloc(size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        const size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > bytes) free_[off + bytes] = len - bytes;
        return off;
      }
    }
    return npos;
  }
  void release(size_t off, size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<size_t, size_t> free_;
};

struct Request {
  long long id = 0;
  std::unique_ptr<scf_system> S;
  std::vector<double> x0;
  double t_arrive = 0, t_admit = 0;
  size_t bytes = 0, offset = 0;
  scf_served out{};
};

struct Launch {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  DevSys **d_list = nullptr;  // device array of the launch's systems
  int *d_phase = nullptr;     // phase of each system after the slice
  DevSys **h_list = nullptr;  // pinned host copies (asynchronous copies from pageable
  int *h_phase = nullptr;     // memory can wait for the device)
  std::vector<Request *> reqs;
  bool busy = false;
};

}  // namespace

struct scf_server {
  char *pool = nullptr;
  size_t capacity = 0;
  std::unique_ptr<PoolAllocator> alloc;
  int max_batch = 1024;
  std::vector<Launch> launches;
  hipStream_t aux = nullptr;  // building systems, reading results
  std::deque<Request *> pending;
  std::deque<Request *> active;  // admitted fused-path requests not in a launch
  double slice_ms = 20.0;
  std::vector<scf_served> done;
  long long next_id = 1;
  double clock_offset = 0;  // host seconds = device ticks / (khz * 1000) + clock_offset
  double khz = 0;
  // grid-path worker
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Request *> grid_queue;
  int grid_busy = 0;
  bool stop = false;
  std::vector<std::thread> workers;  // grid-path solves, each on its own stream
};

namespace {

void finish(scf_server *srv, Request *r, const DevScalars &sc, double t_done_host, double solve_ms) {
  scf_served &o = r->out;
  o.id = r->id;
  o.status = sc.status != DS_OK ? sc.status : (sc.converged ? SCF_OK : SCF_NOT_CONVERGED);
  o.iterations = sc.iterations;
  o.gradients = r->S->h.gradients;
  o.residual = sc.residual;
  o.free_energy = sc.F;
  o.grand_potential = sc.GP;
  o.queue_ms = 1e3 * (r->t_admit - r->t_arrive);
  o.latency_ms = 1e3 * (t_done_host - r->t_arrive);
  o.solve_ms = solve_ms;
  o.device_bytes = r->bytes;
  srv->done.push_back(o);
  srv->alloc->release(r->offset, r->bytes);
  delete r;
}

void grid_worker(scf_server *srv) {
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return;
  for (;;) {
    Request *r = nullptr;
    {
      std::unique_lock<std::mutex> lk(srv->mu);
      srv->cv.wait(lk, [&] { return srv->stop || !srv->grid_queue.empty(); });
      if (srv->grid_queue.empty()) break;
      r = srv->grid_queue.front();
      srv->grid_queue.pop_front();
      srv->grid_busy++;
    }
    r->S->stream = st;
    scf_solve_info info{};
    const double t0 = host_now();
    int rc = grid_solve(r->S.get(), &info);
    DevScalars sc{};
    if (rc == SCF_OK) rc = read_scalars(r->S.get(), sc);
    const double t1 = host_now();
    if (rc != SCF_OK) {
      sc.status = rc;
    } else {
      sc.iterations = info.iterations;
      sc.residual = info.residual;
      sc.converged = info.status == SCF_OK;
      if (info.status < 0) sc.status = info.status;
    }
    r->S->drop_graphs();
    std::lock_guard<std::mutex> lk(srv->mu);
    finish(srv, r, sc, t1, 1e3 * (t1 - t0));
    srv->grid_busy--;
  }
  hipStreamDestroy(st);
}

}  // namespace

extern "C" {

int scf_server_create(size_t pool_bytes, int streams, scf_server **out) {
  if (!out || pool_bytes == 0 || streams < 1 || streams > 8) return fail(SCF_E_ARG, "bad argument");
  auto *srv = new scf_server();
  if (hipMalloc((void **)&srv->pool, pool_bytes) != hipSuccess) {
    delete srv;
    return fail(SCF_E_OOM, "server pool of " + std::to_string(pool_bytes) + " bytes");
  }
  srv->capacity = pool_bytes;
  srv->alloc.reset(new PoolAllocator(pool_bytes));
  int rc = SCF_OK;
  hipError_t e = hipStreamCreateWithFlags(&srv->aux, hipStreamNonBlocking);
  srv->launches.resize(streams);
  for (auto &L : srv->launches) {
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&L.done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc((void **)&L.d_list, srv->max_batch * sizeof(DevSys *));
    if (e == hipSuccess) e = hipMalloc((void **)&L.d_phase, srv->max_batch * sizeof(int));
    if (e == hipSuccess) e = hipHostMalloc((void **)&L.h_list, srv->max_batch * sizeof(DevSys *), 0);
    if (e == hipSuccess) e = hipHostMalloc((void **)&L.h_phase, srv->max_batch * sizeof(int), 0);
  }
  // device clock -> host clock
  unsigned long long *dclk = nullptr, hclk = 0;
  if (e == hipSuccess) e = hipMalloc((void **)&dclk, sizeof(unsigned long long));
  if (e == hipSuccess) {
    const double t0 = host_now();
    hipLaunchKernelGGL(k_read_clock, dim3(1), dim3(64), 0, srv->aux, dclk);
    e = hipStreamSynchronize(srv->aux);
    const double t1 = host_now();
    if (e == hipSuccess) e = hipMemcpy(&hclk, dclk, sizeof(hclk), hipMemcpyDeviceToHost);
    srv->khz = wall_clock_khz();
    srv->clock_offset = 0.5 * (t0 + t1) - (double)hclk / (srv->khz * 1e3);
  }
  if (dclk) hipFree(dclk);
  if (e != hipSuccess) rc = hip_fail(e, "server setup");
  if (rc != SCF_OK) {
    scf_server_destroy(srv);
    return rc;
  }
  for (int w = 0; w < kGridWorkers; w++) srv->workers.emplace_back(grid_worker, srv);
  *out = srv;
  return SCF_OK;
}

int scf_server_set_slice(scf_server *srv, double slice_ms) {
  if (!srv || !(slice_ms > 0)) return fail(SCF_E_ARG, "bad argument");
  srv->slice_ms = slice_ms;
  return SCF_OK;
}

int scf_server_submit(scf_server *srv, const char *input_text, const double *x0, size_t nx0, long long *id) {
  if (!srv || !input_text) return fail(SCF_E_ARG, "null argument");
  std::unique_ptr<Request> r(new Request());
  r->S.reset(new scf_system());
  std::string why;
  if (!parse_input(input_text, 1, r->S->spec, why)) return fail(SCF_E_PARSE, why);
  int v = validate(r->S->spec, why);
  if (v != SCF_OK) return fail(v, why);
  r->S->owns_stream = false;
  Placement dry;
  dry.dry = true;
  int rc = build(r->S.get(), &dry);
  if (rc != SCF_OK) return rc;
  r->bytes = r->S->arena_bytes;
  if (r->bytes > srv->capacity) return fail(SCF_E_OOM, "request larger than the server pool");
  const size_t nvar = (size_t)r->S->spec.itmon.size() * r->S->spec.lat.M;
  if (x0) {
    if (nx0 != nvar) return fail(SCF_E_ARG, "x0 has " + std::to_string(nx0) + " values, the system " +
                                                std::to_string(nvar));
    r->x0.assign(x0, x0 + nx0);
  }
  r->t_arrive = host_now();
  std::lock_guard<std::mutex> lk(srv->mu);
  r->id = srv->next_id++;
  if (id) *id = r->id;
  srv->pending.push_back(r.release());
  return SCF_OK;
}

int scf_server_progress(scf_server *srv, int *queued, int *running) {
  if (!srv) return fail(SCF_E_ARG, "null server");
  // 1. admission, first come first served, while the footprint fits
  for (;;) {
    Request *r = nullptr;
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      if (srv->pending.empty()) break;
      r = srv->pending.front();
      const size_t off = srv->alloc->alloc(r->bytes);
      if (off == PoolAllocator::npos) break;
      r->offset = off;
      srv->pending.pop_front();
    }
    scf_system *S = r->S.get();
    S->stream = srv->aux;
    Placement pl;
    pl.region = srv->pool + r->offset;
    pl.capacity = r->bytes;
    int rc = build(S, &pl);
    if (rc == SCF_OK && !r->x0.empty()) {
      hipError_t e = hipMemcpyAsync(S->h.x, r->x0.data(), r->x0.size() * sizeof(double), hipMemcpyHostToDevice,
                                    srv->aux);
      if (e == hipSuccess) e = hipStreamSynchronize(srv->aux);
      if (e != hipSuccess) rc = hip_fail(e, "request fields");
    }
    r->t_admit = host_now();
    if (rc != SCF_OK) {
      DevScalars sc{};
      sc.status = rc;
      std::lock_guard<std::mutex> lk(srv->mu);
      finish(srv, r, sc, host_now(), 0.0);
      continue;
    }
    if (S->effective_mode() == SCF_MODE_GRID) {
      std::lock_guard<std::mutex> lk(srv->mu);
      srv->grid_queue.push_back(r);
      srv->cv.notify_one();
    } else {
      srv->active.push_back(r);
    }
  }
  // 2. slices: every free stream takes the active requests of one workgroup
  //    shape (threads, LDS), in queue order, for one time slice
  const unsigned long long budget = (unsigned long long)(srv->slice_ms * srv->khz);
  for (auto &L : srv->launches) {
    if (L.busy || srv->active.empty()) continue;
    const int threads = srv->active.front()->S->h.threads;
    const size_t lds = srv->active.front()->S->lds_bytes;
    L.reqs.clear();
    std::deque<Request *> rest;
    for (Request *r : srv->active) {
      if (r->S->h.threads == threads && r->S->lds_bytes == lds && (int)L.reqs.size() < srv->max_batch) {
        L.h_list[L.reqs.size()] = r->S->d;
        L.reqs.push_back(r);
      } else {
        rest.push_back(r);
      }
    }
    srv->active.swap(rest);
    HIP_TRY(hipMemcpyAsync(L.d_list, L.h_list, L.reqs.size() * sizeof(DevSys *), hipMemcpyHostToDevice,
                           L.stream));
    for (Request *r : L.reqs) r->S->stream = L.stream;
    HIP_TRY(launch_fused(L.d_list, (int)L.reqs.size(), FUSED_SLICE, threads, lds, L.stream, budget, L.d_phase));
    HIP_TRY(hipMemcpyAsync(L.h_phase, L.d_phase, L.reqs.size() * sizeof(int), hipMemcpyDeviceToHost,
                           L.stream));
    HIP_TRY(hipEventRecord(L.done, L.stream));
    L.busy = true;
  }
  // 3. finished slices: completed requests hand back results and memory,
  //    the others rejoin the active queue
  for (auto &L : srv->launches) {
    if (!L.busy) continue;
    hipError_t q = hipEventQuery(L.done);
    if (q == hipErrorNotReady) continue;
    if (q != hipSuccess) return hip_fail(q, "launch");
    for (size_t i = 0; i < L.reqs.size(); i++) {
      Request *r = L.reqs[i];
      if (L.h_phase[i] != 2) {
        srv->active.push_back(r);
        continue;
      }
      DevScalars sc;
      HIP_TRY(hipMemcpyAsync(&sc, r->S->h.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, L.stream));
      HIP_TRY(hipStreamSynchronize(L.stream));
      const double t_done = (double)sc.t_end / (srv->khz * 1e3) + srv->clock_offset;
      std::lock_guard<std::mutex> lk(srv->mu);
      finish(srv, r, sc, t_done, (double)(sc.t_end - sc.t_begin) / srv->khz);
    }
    L.reqs.clear();
    L.busy = false;
  }
  std::lock_guard<std::mutex> lk(srv->mu);
  int run = srv->grid_busy + (int)srv->grid_queue.size() + (int)srv->active.size();
  for (auto &L : srv->launches) run += (int)L.reqs.size();
  if (queued) *queued = (int)srv->pending.size();
  if (running) *running = run;
  return SCF_OK;
}

int scf_server_results(scf_server *srv, scf_served *out, int max, int *n) {
  if (!srv || !n) return fail(SCF_E_ARG, "null argument");
  std::lock_guard<std::mutex> lk(srv->mu);
  const int k = std::min(max, (int)srv->done.size());
  for (int i = 0; i < k; i++) out[i] = srv->done[i];
  srv->done.erase(srv->done.begin(), srv->done.begin() + k);
  *n = k;
  return SCF_OK;
}

void scf_server_destroy(scf_server *srv) {
  if (!srv) return;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
    srv->cv.notify_all();
  }
  for (auto &w : srv->workers)
    if (w.joinable()) w.join();
  for (auto &L : srv->launches) {
    if (L.stream) hipStreamSynchronize(L.stream);
    for (Request *r : L.reqs) delete r;
    if (L.d_list) hipFree(L.d_list);
    if (L.d_phase) hipFree(L.d_phase);
    if (L.h_list) hipHostFree(L.h_list);
    if (L.h_phase) hipHostFree(L.h_phase);
    if (L.done) hipEventDestroy(L.done);
    if (L.stream) hipStreamDestroy(L.stream);
  }
  for (Request *r : srv->active) delete r;
  for (Request *r : srv->pending) delete r;
  for (Request *r : srv->grid_queue) delete r;
  if (srv->aux) {
    hipStreamSynchronize(srv->aux);
    hipStreamDestroy(srv->aux);
  }
  if (srv->pool) hipFree(srv->pool);
  delete srv;
}

}  // extern "C"
