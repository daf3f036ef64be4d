// Entropy / norm reductions and the relaxation-LSRK device kernels: law
// instantiations and the law-independent kernels.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "measures.hpp"

namespace atm {

template <template <int> class LAW>
MeasureFns measures_dims(int dim) {
  if (dim == 1) return measure_entry<LAW<1>>();
  if (dim == 2) return measure_entry<LAW<2>>();
  if (dim == 3) return measure_entry<LAW<3>>();
  return MeasureFns{};
}

MeasureFns resolve_measures(int law, int dim) {
  switch (law) {
    case ATM_LAW_ADVECTION: return measures_dims<Advection>(dim);
    case ATM_LAW_EULER: return measures_dims<Euler>(dim);
    case ATM_LAW_EULER_GRAVITY: return measures_dims<EulerGravity>(dim);
    case ATM_LAW_SHALLOW_WATER: return measures_dims<ShallowWater>(dim);
    default: return MeasureFns{};
  }
}

namespace {

ATM_DEV double nanmax(double a, double b) { return (a > b || a != a) ? a : b; }

// src/measures.jl:6-21, per component: sum MJ*|f|^p, or max |f| for p = inf
template <int NS>
__global__ void __launch_bounds__(RED_THREADS)
    wnorm_kernel(const double* __restrict__ q, const double* __restrict__ MJ, int64_t Ntot, double pw,
                 double* __restrict__ part) {
  const bool inf = pw == (double)INFINITY;
  DD acc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) acc[s] = DD{0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * RED_THREADS;
  for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < Ntot; i += stride) {
    const double m = MJ[i];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const double a = fabs(q[s * Ntot + i]);
      if (inf) {
        acc[s].hi = nanmax(acc[s].hi, a);
      } else {
        const double ap = (pw == 2.0) ? a * a : ((pw == 1.0) ? a : pow(a, pw));
        acc[s] = dd_add(acc[s], m * ap);
      }
    }
  }
  if (!inf) {
    block_store<NS>(acc, part);
    return;
  }
  __shared__ double sh[NS][RED_THREADS];
  const int t = threadIdx.x;
#pragma unroll
  for (int s = 0; s < NS; ++s) sh[s][t] = acc[s].hi;
  __syncthreads();
  for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int s = 0; s < NS; ++s) sh[s][t] = nanmax(sh[s][t], sh[s][t + w]);
    }
    __syncthreads();
  }
  if (t == 0) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      part[(blockIdx.x * NS + s) * 2 + 0] = sh[s][0];
      part[(blockIdx.x * NS + s) * 2 + 1] = 0.0;
    }
  }
}

__global__ void __launch_bounds__(RED_THREADS)
    finalize_kernel(const double* __restrict__ part, int nblk, int nq, int mode, double coef, double* __restrict__ st,
                    double* __restrict__ out) {
  __shared__ double sh[RED_MAXQ][2][RED_THREADS];
  const int t = threadIdx.x;
  if (mode == FIN_NEWTON && st[RX_DONE] != 0.0) return;
  const bool mx = mode == FIN_MAX;
  for (int j = 0; j < nq; ++j) {
    DD a{0.0, 0.0};
    for (int b = t; b < nblk; b += RED_THREADS) {
      const DD v{part[(b * nq + j) * 2 + 0], part[(b * nq + j) * 2 + 1]};
      if (mx) a.hi = nanmax(a.hi, v.hi);
      else a = dd_add(a, v);
    }
    sh[j][0][t] = a.hi;
    sh[j][1][t] = a.lo;
  }
  __syncthreads();
  for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
    if (t < w) {
      for (int j = 0; j < nq; ++j) {
        if (mx) {
          sh[j][0][t] = nanmax(sh[j][0][t], sh[j][0][t + w]);
        } else {
          DD c = dd_add(DD{sh[j][0][t], sh[j][1][t]}, DD{sh[j][0][t + w], sh[j][1][t + w]});
          sh[j][0][t] = c.hi;
          sh[j][1][t] = c.lo;
        }
      }
    }
    __syncthreads();
  }
  if (t != 0) return;
  // the double-double is normalized, so hi is the correctly rounded sum
  const double S0 = sh[0][0][0], S1 = nq > 1 ? sh[1][0][0] : 0.0;
  switch (mode) {
    case FIN_OUT:
    case FIN_MAX:
      for (int j = 0; j < nq; ++j) out[j] = sh[j][0][0];
      break;
    case FIN_ETA0:
      st[RX_ETA0] = S0;
      st[RX_DETA] = -0.0;
      st[RX_DONE] = 0.0;
      st[RX_ITER] = 0.0;
      break;
    case FIN_DETA:
      st[RX_DETA] = st[RX_DETA] + coef * S0;
      break;
    case FIN_NEWTON: {
      // src/odesolvers.jl:121-131: r(γ) = η(q0 + γk) - η0 - γ dη, r'(γ) = <v(q0 + γk), k> - dη;
      // Newton from the previous step's γ, stopping at |r| <= 10 eps or at a
      // step of round-off size (r(0) = 0 is a trivial root to stay away from)
      const double ftol = 10 * DBL_EPSILON;
      double gam = st[RX_GAMMA];
      const double r = (S0 - st[RX_ETA0]) - gam * st[RX_DETA];
      st[RX_R] = r;
      if (fabs(r) <= ftol) {
        st[RX_DONE] = 1.0;
        break;
      }
      const double J = S1 - st[RX_DETA];
      st[RX_J] = J;
      const double step = r / J;
      if (!isfinite(step)) {
        st[RX_DONE] = 1.0;
        break;
      }
      gam -= step;
      st[RX_GAMMA] = gam;
      st[RX_ITER] = st[RX_ITER] + 1.0;
      if (fabs(step) <= 4 * DBL_EPSILON * fabs(gam) || st[RX_ITER] >= 50.0) st[RX_DONE] = 1.0;
      break;
    }
    default:
      break;
  }
}

__global__ void diff_kernel(double* __restrict__ y, const double* __restrict__ x, const double* __restrict__ z,
                            int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = x[i] - z[i];
}

__global__ void relax_apply_kernel(double* __restrict__ q, const double* __restrict__ q0,
                                   const double* __restrict__ k, const double* __restrict__ st, int64_t n) {
  const double gam = st[RX_GAMMA];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    q[i] = q0[i] + gam * k[i];
}

__global__ void set_slot_kernel(double* st, int i, double v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) st[i] = v;
}

inline int ew_grid(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (int)(b > 65535 * 16 ? 65535 * 16 : (b < 1 ? 1 : b));
}

}  // namespace

hipError_t launch_wnorm(hipStream_t s, const double* q, const double* MJ, int Ns, int64_t Ntot, double pw,
                        double* part) {
  const dim3 g(red_blocks(Ntot)), b(RED_THREADS);
  switch (Ns) {
    case 1: hipLaunchKernelGGL(wnorm_kernel<1>, g, b, 0, s, q, MJ, Ntot, pw, part); break;
    case 2: hipLaunchKernelGGL(wnorm_kernel<2>, g, b, 0, s, q, MJ, Ntot, pw, part); break;
    case 3: hipLaunchKernelGGL(wnorm_kernel<3>, g, b, 0, s, q, MJ, Ntot, pw, part); break;
    case 4: hipLaunchKernelGGL(wnorm_kernel<4>, g, b, 0, s, q, MJ, Ntot, pw, part); break;
    case 5: hipLaunchKernelGGL(wnorm_kernel<5>, g, b, 0, s, q, MJ, Ntot, pw, part); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_finalize(hipStream_t s, const double* part, int nblk, int nq, int mode, double coef, double* st,
                           double* out) {
  if (nq < 1 || nq > RED_MAXQ || nblk < 1 || nblk > RED_MAXBLK) return hipErrorInvalidValue;
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(RED_THREADS), 0, s, part, nblk, nq, mode, coef, st, out);
  return hipGetLastError();
}

hipError_t launch_diff(hipStream_t s, double* y, const double* x, const double* z, int64_t n) {
  hipLaunchKernelGGL(diff_kernel, dim3(ew_grid(n)), dim3(256), 0, s, y, x, z, n);
  return hipGetLastError();
}

hipError_t launch_relax_apply(hipStream_t s, double* q, const double* q0, const double* k, const double* st,
                              int64_t n) {
  hipLaunchKernelGGL(relax_apply_kernel, dim3(ew_grid(n)), dim3(256), 0, s, q, q0, k, st, n);
  return hipGetLastError();
}

hipError_t launch_set_slot(hipStream_t s, double* st, int i, double v) {
  hipLaunchKernelGGL(set_slot_kernel, dim3(1), dim3(64), 0, s, st, i, v);
  return hipGetLastError();
}

}  // namespace atm
