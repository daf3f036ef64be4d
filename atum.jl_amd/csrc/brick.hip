// Brick meshes built on the device: the geometry a DGSEM operator derives from
// Bennu's brickgrid (LobattoCell, vertex coordinates per dimension, periodic
// flags, Cartesian or stacked element order), src/dgsem.jl:38-56.
//
// One thread per volume node computes the node coordinates, MJ = w_i w_j w_k J,
// MJI = 1/MJ and the affine metric g = diag(2/h); one thread per face point
// computes the face tables (neighbour node, boundary tag), faceMJ and the unit
// normal.  Every value is formed with the same operations in the same order as
// the host builder (atum_hip/grid.py), so both produce identical bits.
#include <hip/hip_runtime.h>

#include <cstring>

#include "handles.hpp"

namespace {

struct BrickGeo {
  int dim, Nq, Np, Nfpf, Nfp, stacked;
  int Kd[3], periodic[3];
  int64_t K, Ntot, NF;
  const double* coords[3];
  const double* xi;
  const double* w;
};

__device__ void elem_multi(const BrickGeo& b, int64_t e, int em[3]) {
  em[0] = em[1] = em[2] = 0;
  if (!b.stacked) {
    for (int d = 0; d < b.dim; ++d) {
      em[d] = (int)(e % b.Kd[d]);
      e /= b.Kd[d];
    }
  } else {
    const int V = b.dim - 1;
    em[V] = (int)(e % b.Kd[V]);
    int64_t h = e / b.Kd[V];
    for (int d = 0; d < V; ++d) {
      em[d] = (int)(h % b.Kd[d]);
      h /= b.Kd[d];
    }
  }
}

__device__ int64_t elem_id(const BrickGeo& b, const int em[3]) {
  if (!b.stacked) {
    int64_t e = 0;
    for (int d = b.dim - 1; d >= 0; --d) e = e * b.Kd[d] + em[d];
    return e;
  }
  const int V = b.dim - 1;
  int64_t h = 0, stride = 1;
  for (int d = 0; d < V; ++d) {
    h += em[d] * stride;
    stride *= b.Kd[d];
  }
  return em[V] + (int64_t)b.Kd[V] * h;
}

__global__ void brick_volume_kernel(BrickGeo b, double* __restrict__ points, double* __restrict__ g,
                                    double* __restrict__ MJ, double* __restrict__ MJI) {
  for (int64_t node = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; node < b.Ntot;
       node += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = node / b.Np;
    const int ijk = (int)(node - e * b.Np);
    int em[3];
    elem_multi(b, e, em);
    int idx[3] = {ijk % b.Nq, (ijk / b.Nq) % b.Nq, ijk / (b.Nq * b.Nq)};
    double h[3], J = 1.0, wp = 1.0;
    for (int d = 0; d < b.dim; ++d) {
      const double lo = b.coords[d][em[d]], hi = b.coords[d][em[d] + 1];
      h[d] = hi - lo;
      const double t = b.xi[idx[d]];
      points[d * b.Ntot + node] = (1 - t) / 2 * lo + (1 + t) / 2 * hi;
      J = (d == 0) ? h[d] / 2 : J * (h[d] / 2);
      wp = wp * b.w[idx[d]];
    }
    const double mj = wp * J;
    MJ[node] = mj;
    MJI[node] = 1 / mj;
    for (int a = 0; a < b.dim; ++a)
      for (int c = 0; c < b.dim; ++c) g[(int64_t)(a + b.dim * c) * b.Ntot + node] = (a == c) ? 2 / h[a] : 0.0;
  }
}

__global__ void brick_face_kernel(BrickGeo b, int64_t* __restrict__ faceix_p, double* __restrict__ faceMJ,
                                  double* __restrict__ normal, int32_t* __restrict__ bctag) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < b.NF; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = t / b.Nfp;
    const int f = (int)(t - e * b.Nfp);
    const int face = f / b.Nfpf, fp = f - face * b.Nfpf, r = face >> 1, side = face & 1;
    int em[3];
    elem_multi(b, e, em);
    int im[3] = {0, 0, 0};
    int rem = fp;
    for (int d = 0; d < b.dim; ++d) {
      if (d == r) continue;
      im[d] = rem % b.Nq;
      rem /= b.Nq;
    }
    int ip[3] = {im[0], im[1], im[2]};
    im[r] = side == 0 ? 0 : b.Nq - 1;
    ip[r] = side == 0 ? b.Nq - 1 : 0;
    const int64_t nm = im[0] + b.Nq * (im[1] + (int64_t)b.Nq * im[2]);
    const int64_t npn = ip[0] + b.Nq * (ip[1] + (int64_t)b.Nq * ip[2]);
    int nb[3] = {em[0], em[1], em[2]};
    nb[r] += side == 0 ? -1 : 1;
    const bool out = nb[r] < 0 || nb[r] >= b.Kd[r];
    bool bnd;
    if (b.periodic[r]) {
      nb[r] = ((nb[r] % b.Kd[r]) + b.Kd[r]) % b.Kd[r];
      bnd = false;
    } else {
      bnd = out;
      if (out) nb[r] = em[r];
    }
    const int64_t en = elem_id(b, nb);
    faceix_p[t] = bnd ? e * b.Np + nm : en * b.Np + npn;
    double wf = 1.0, fJ = 1.0;
    for (int d = 0; d < b.dim; ++d) {
      if (d == r) continue;
      const double h = b.coords[d][em[d] + 1] - b.coords[d][em[d]];
      wf = wf * b.w[im[d]];
      fJ = fJ * (h / 2);
    }
    faceMJ[t] = wf * fJ;
    for (int d = 0; d < b.dim; ++d) normal[d * b.NF + t] = (d == r) ? (side == 0 ? -1.0 : 1.0) : 0.0;
    if (fp == 0) bctag[e * (2 * b.dim) + face] = bnd ? face + 1 : 0;
  }
}

inline int grid_blocks(int64_t n) {
  int64_t g = (n + 255) / 256;
  return (int)(g > 65535 * 16 ? 65535 * 16 : (g < 1 ? 1 : g));
}

}  // namespace

extern "C" {

int atm_dg_create_brick(atm_ctx* ctx, const atm_dg_desc* d, const atm_brick_desc* bd, atm_dg** out) {
  if (!ctx || !d || !bd || !out) return fail(ctx, ATM_ERR_ARG, "atm_dg_create_brick: null argument");
  *out = nullptr;
  if (d->dim < 1 || d->dim > 3) return fail(ctx, ATM_ERR_ARG, "dim must be 1, 2 or 3");
  int64_t K = 1;
  for (int i = 0; i < d->dim; ++i) {
    if (bd->Kd[i] < 1) return fail(ctx, ATM_ERR_ARG, "at least one element per dimension");
    if (!bd->coords[i]) return fail(ctx, ATM_ERR_ARG, "missing vertex coordinates");
    K *= bd->Kd[i];
  }
  if (K != d->K) return fail(ctx, ATM_ERR_ARG, "K differs from the product of elements per dimension");
  if (!d->D || !bd->xi || !bd->w) return fail(ctx, ATM_ERR_ARG, "missing LGL nodes, weights or D");
  atm_dg* dg = nullptr;
  int rc = dg_prepare(ctx, d, &dg);
  if (rc) return rc;
  DevDG& p = dg->dev;
  const int dim = d->dim, Nq = dg->Nq, Np = dg->Np;
  const int64_t Ntot = dg->Ntot, NF = p.NF;
  BrickGeo b{};
  b.dim = dim;
  b.Nq = Nq;
  b.Np = Np;
  b.Nfpf = p.Nfpf;
  b.Nfp = p.Nfp;
  b.stacked = bd->stacked ? 1 : 0;
  for (int i = 0; i < 3; ++i) {
    b.Kd[i] = i < dim ? bd->Kd[i] : 1;
    b.periodic[i] = i < dim ? (bd->periodic[i] ? 1 : 0) : 0;
  }
  b.K = K;
  b.Ntot = Ntot;
  b.NF = NF;
  const double* cdev[3] = {nullptr, nullptr, nullptr};
  void *pts = nullptr, *g = nullptr, *MJ = nullptr, *MJI = nullptr, *fMJ = nullptr, *nrm = nullptr, *fxp = nullptr,
       *tag = nullptr;
  for (int i = 0; i < dim && rc == ATM_OK; ++i) rc = upload_copy(dg, bd->coords[i], (size_t)bd->Kd[i] + 1, &cdev[i]);
  if (rc == ATM_OK) rc = upload_copy(dg, bd->xi, (size_t)Nq, &b.xi);
  if (rc == ATM_OK) rc = upload_copy(dg, bd->w, (size_t)Nq, &b.w);
  if (rc == ATM_OK) rc = upload_copy(dg, d->D, (size_t)Nq * Nq, &p.D);
  const size_t S = sizeof(double);
  if (rc == ATM_OK && Ntot > 0 &&
      ((rc = dg_alloc(dg, (size_t)dim * Ntot * S, &pts)) || (rc = dg_alloc(dg, (size_t)dim * dim * Ntot * S, &g)) ||
       (rc = dg_alloc(dg, (size_t)Ntot * S, &MJ)) || (rc = dg_alloc(dg, (size_t)Ntot * S, &MJI)) ||
       (rc = dg_alloc(dg, (size_t)NF * S, &fMJ)) || (rc = dg_alloc(dg, (size_t)dim * NF * S, &nrm)) ||
       (rc = dg_alloc(dg, (size_t)NF * sizeof(int64_t), &fxp)) ||
       (rc = dg_alloc(dg, (size_t)2 * dim * K * sizeof(int32_t), &tag)))) {
  }
  if (rc) {
    atm_dg_destroy(dg);
    return rc;
  }
  for (int i = 0; i < 3; ++i) b.coords[i] = cdev[i];
  if (Ntot > 0) {
    hipStream_t s = stream_of(ctx);
    hipLaunchKernelGGL(brick_volume_kernel, dim3(grid_blocks(Ntot)), dim3(256), 0, s, b, (double*)pts, (double*)g,
                       (double*)MJ, (double*)MJI);
    hipLaunchKernelGGL(brick_face_kernel, dim3(grid_blocks(NF)), dim3(256), 0, s, b, (int64_t*)fxp, (double*)fMJ,
                       (double*)nrm, (int32_t*)tag);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      atm_dg_destroy(dg);
      return hipfail(ctx, e, "brick geometry kernels");
    }
  }
  dg->points = (double*)pts;
  p.g = (const double*)g;
  p.MJ = (const double*)MJ;
  p.MJI = (const double*)MJI;
  p.faceMJ = (const double*)fMJ;
  p.normal = (const double*)nrm;
  p.faceix_p = (const int64_t*)fxp;
  p.bctag = (const int32_t*)tag;
  *out = dg;
  return ATM_OK;
}

int atm_dg_geometry(atm_dg* dg, int which, void* host) {
  if (!dg || !host) return fail(dg ? dg->ctx : nullptr, ATM_ERR_ARG, "atm_dg_geometry: null argument");
  const DevDG& p = dg->dev;
  const int dim = dg->dim;
  const void* src = nullptr;
  size_t bytes = 0;
  switch (which) {
    case ATM_GEOM_POINTS: src = dg->points; bytes = (size_t)dim * dg->Ntot * sizeof(double); break;
    case ATM_GEOM_G: src = p.g; bytes = (size_t)dim * dim * dg->Ntot * sizeof(double); break;
    case ATM_GEOM_MJ: src = p.MJ; bytes = (size_t)dg->Ntot * sizeof(double); break;
    case ATM_GEOM_MJI: src = p.MJI; bytes = (size_t)dg->Ntot * sizeof(double); break;
    case ATM_GEOM_FACEMJ: src = p.faceMJ; bytes = (size_t)p.NF * sizeof(double); break;
    case ATM_GEOM_NORMAL: src = p.normal; bytes = (size_t)dim * p.NF * sizeof(double); break;
    case ATM_GEOM_FACEIX_P: src = p.faceix_p; bytes = (size_t)p.NF * sizeof(int64_t); break;
    case ATM_GEOM_BCTAG: src = p.bctag; bytes = (size_t)2 * dim * p.K * sizeof(int32_t); break;
    default: return fail(dg->ctx, ATM_ERR_ARG, "atm_dg_geometry: unknown array");
  }
  if (bytes == 0) return ATM_OK;
  if (!src) return fail(dg->ctx, ATM_ERR_UNSUPPORTED, "atm_dg_geometry: array not held on the device");
  hipSetDevice(dg->ctx->device);
  HIPCHK(dg->ctx, hipStreamSynchronize(stream_of(dg->ctx)));
  HIPCHK(dg->ctx, hipMemcpy(host, src, bytes, hipMemcpyDeviceToHost));
  return ATM_OK;
}

}  // extern "C"
