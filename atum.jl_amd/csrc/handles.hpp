// Library-internal definitions shared by the translation units of
// libatumhip.so: the context and operator handles and the error helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "measures.hpp"
#include "resolve.hpp"

using namespace atm;

struct atm_ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  size_t max_lds = 65536;
  std::string err;      // last error message (any thread; guarded by err_mu)
  std::mutex err_mu;
  hipEvent_t events[16] = {};
  // snapshots: a second stream (default priority) for device-to-host copies
  hipStream_t copy = nullptr;
  hipEvent_t snap_ready[16] = {}, snap_done[16] = {};
  // device address of the library's fixed no-allocation scratch array on this
  // context's device (null if it could not be resolved)
  double* noalloc_scratch = nullptr;
};

struct atm_dg {
  atm_ctx* ctx = nullptr;
  DevDG dev{};
  LaunchFn fn = nullptr;
  BatchLaunchFn bfn = nullptr;
  size_t lds = 0;
  int epw = 1;
  size_t lds_fixed = 0, lds_elem = 0;  // LDS bytes of a workgroup = lds_fixed + epw * lds_elem
  int max_threads = 512;               // largest workgroup of the element kernels
  BatchLaunchFn mbfn = nullptr;         // one-launch batched kernel across Nq (or null)
  int32_t* sched = nullptr;             // device copy of the element processing order (or null)
  int law = 0, dim = 0, Nq = 0, Np = 0, Ns = 0;
  int64_t Ntot = 0;
  MeasureFns mfn;
  double* points = nullptr;  // node coordinates (dim blocks of Ntot) when the geometry was built on the device
  double* red = nullptr;  // reduction partials, relaxation scalars, results (allocated on first use)
  // the last two-step LSRK graph and the arguments it was captured with
  hipGraphExec_t lsrk_exec = nullptr;
  int lsrk_scheme = -1;
  const double* lsrk_q = nullptr;
  const double* lsrk_dq = nullptr;
  const double* lsrk_scratch = nullptr;
  double lsrk_dt = 0.0;
  std::vector<void*> allocs;
  size_t dev_bytes = 0;  // device memory the operator holds (allocs, reduction buffer, schedule)
};


namespace atm {

// Stream override of the calling thread: while set, every launch, copy and
// wait of the library on this thread uses it instead of the context's stream
// (the serving scheduler runs its requests on its own streams this way,
// leaving the context's stream to the caller's thread).
extern thread_local hipStream_t tl_stream;
inline hipStream_t stream_of(const atm_ctx* c) { return tl_stream ? tl_stream : c->stream; }

inline int fail(atm_ctx* ctx, int code, const std::string& msg) {
  if (ctx) {
    std::lock_guard<std::mutex> lk(ctx->err_mu);
    ctx->err = msg;
  }
  return code;
}

inline int hipfail(atm_ctx* ctx, hipError_t e, const char* what) {
  return fail(ctx, ATM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(ctx, call)                               \
  do {                                                  \
    hipError_t _e = (call);                             \
    if (_e != hipSuccess) return hipfail(ctx, _e, #call); \
  } while (0)


// Validates everything in d except the geometry arrays, resolves the kernels
// and creates the operator with its device description filled in (geometry,
// aux and boundary-state pointers still null).
int dg_prepare(atm_ctx* ctx, const atm_dg_desc* d, atm_dg** out);

// The linear operator a banded factorisation was built from.
atm_dg* banded_lin(const atm_banded* b);

// A copy of the context's last error message (threads may set it concurrently).
inline std::string last_error(atm_ctx* ctx) {
  std::lock_guard<std::mutex> lk(ctx->err_mu);
  return ctx->err;
}

// Device allocation owned by the operator (freed by atm_dg_destroy).
int dg_alloc(atm_dg* dg, size_t bytes, void** out);

// Regroups a batch for a new member list, keeping its table memory (the serving
// scheduler changes members whenever requests arrive or finish).  Calls on one
// batch must be ordered on one stream.
int batch_set_members(atm_batch* b, int32_t n, atm_dg* const* members);

// Allocation owned by the operator, filled from host memory (count == 0: null).
template <class T>
int upload_copy(atm_dg* dg, const T* host, size_t count, const T** out) {
  *out = nullptr;
  if (count == 0) return ATM_OK;
  if (!host) return fail(dg->ctx, ATM_ERR_ARG, "missing array");
  void* p = nullptr;
  int rc = dg_alloc(dg, count * sizeof(T), &p);
  if (rc) return rc;
  HIPCHK(dg->ctx, hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice));
  *out = static_cast<const T*>(p);
  return ATM_OK;
}

}  // namespace atm
