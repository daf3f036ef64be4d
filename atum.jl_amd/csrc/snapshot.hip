// Snapshot stream: the output hook of a time loop (the after_step callback of
// solve!, src/odesolvers.jl:9-26) without stalling the compute stream.  The
// state is copied device-to-device into a staging buffer on the compute stream
// (microseconds), then from the staging buffer into pinned host memory on a
// second stream while the compute stream moves on; an event per slot tells
// the host when the copy has landed.
#include <hip/hip_runtime.h>

#include "handles.hpp"

namespace {

int ensure_copy_stream(atm_ctx* ctx, int slot) {
  if (!ctx->copy) HIPCHK(ctx, hipStreamCreateWithFlags(&ctx->copy, hipStreamNonBlocking));
  if (!ctx->snap_ready[slot]) HIPCHK(ctx, hipEventCreateWithFlags(&ctx->snap_ready[slot], hipEventDisableTiming));
  if (!ctx->snap_done[slot]) HIPCHK(ctx, hipEventCreateWithFlags(&ctx->snap_done[slot], hipEventDisableTiming));
  return ATM_OK;
}

}  // namespace

extern "C" {

int atm_host_alloc(atm_ctx* ctx, size_t bytes, void** host) {
  if (!ctx || !host) return ATM_ERR_ARG;
  *host = nullptr;
  hipSetDevice(ctx->device);
  HIPCHK(ctx, hipHostMalloc(host, bytes > 0 ? bytes : 8, hipHostMallocDefault));
  return ATM_OK;
}

int atm_host_free(atm_ctx* ctx, void* host) {
  if (!ctx) return ATM_ERR_ARG;
  if (host) HIPCHK(ctx, hipHostFree(host));
  return ATM_OK;
}

int atm_snapshot(atm_ctx* ctx, int slot, void* host_dst, void* staging, const void* src, size_t bytes) {
  if (!ctx || !host_dst || !src || slot < 0 || slot >= 16)
    return fail(ctx, ATM_ERR_ARG, "atm_snapshot: bad argument");
  hipSetDevice(ctx->device);
  int rc = ensure_copy_stream(ctx, slot);
  if (rc) return rc;
  // the slot's previous copy must have left the staging buffer before it is overwritten
  HIPCHK(ctx, hipStreamWaitEvent(stream_of(ctx), ctx->snap_done[slot], 0));
  const void* from = src;
  if (staging) {
    HIPCHK(ctx, hipMemcpyAsync(staging, src, bytes, hipMemcpyDeviceToDevice, stream_of(ctx)));
    from = staging;
  }
  HIPCHK(ctx, hipEventRecord(ctx->snap_ready[slot], stream_of(ctx)));
  HIPCHK(ctx, hipStreamWaitEvent(ctx->copy, ctx->snap_ready[slot], 0));
  HIPCHK(ctx, hipMemcpyAsync(host_dst, from, bytes, hipMemcpyDeviceToHost, ctx->copy));
  HIPCHK(ctx, hipEventRecord(ctx->snap_done[slot], ctx->copy));
  // without staging, later work on the compute stream may not modify src before the copy
  if (!staging) HIPCHK(ctx, hipStreamWaitEvent(stream_of(ctx), ctx->snap_done[slot], 0));
  return ATM_OK;
}

int atm_snapshot_query(atm_ctx* ctx, int slot) {
  if (!ctx || slot < 0 || slot >= 16) return ATM_ERR_ARG;
  if (!ctx->snap_done[slot]) return 1;
  hipError_t e = hipEventQuery(ctx->snap_done[slot]);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return hipfail(ctx, e, "hipEventQuery");
}

int atm_snapshot_wait(atm_ctx* ctx, int slot) {
  if (!ctx || slot < 0 || slot >= 16) return ATM_ERR_ARG;
  if (!ctx->snap_done[slot]) return ATM_OK;
  hipError_t e = hipEventSynchronize(ctx->snap_done[slot]);
  if (e != hipSuccess) return hipfail(ctx, e, "hipEventSynchronize");
  return ATM_OK;
}

}  // extern "C"
