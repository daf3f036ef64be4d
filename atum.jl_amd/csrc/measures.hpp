// Reductions over the nodes of a DGSEM field (src/measures.jl:6-33) and the
// device side of the relaxation LSRK step (src/odesolvers.jl:102-136).
//
// Sums are accumulated in double-double arithmetic (error-free two-sum) with a
// fixed decomposition: the node range is split into at most RED_MAXBLK
// workgroups of RED_THREADS threads, a function of the node count only, each
// thread sums a strided subset in node order, workgroups combine by a fixed
// LDS tree, and one workgroup combines the partials in block order.  Results
// are therefore reproducible bit for bit, and accurate to about one rounding
// of the exact sum of the rounded per-node terms.
#pragma once
#include "dg_kernels.hpp"

namespace atm {

constexpr int RED_THREADS = 256;
constexpr int RED_MAXBLK = 1024;
constexpr int RED_MAXQ = 8;  // quantities per reduction (components of a field)

// relaxation state, device scalars
enum RelaxSlot { RX_ETA0 = 0, RX_DETA, RX_GAMMA, RX_DONE, RX_ITER, RX_R, RX_J, RX_NSLOTS };

// reduce modes
enum RedMode {
  RED_ENTROPY = 0,  // sum MJ*eta(q)
  RED_PRODUCT,      // sum MJ*v(q).k
  RED_RELAX,        // with q_g = q0 + gamma*k: sum MJ*eta(q_g), sum MJ*v(q_g).k (skipped once done)
};

// finalize modes
enum FinMode {
  FIN_OUT = 0,  // out[j] = sum j
  FIN_MAX,      // out[j] = max j (partials hold maxima)
  FIN_ETA0,     // st[ETA0] = sum 0; st[DETA] = -0.0; done = iter = 0
  FIN_DETA,     // st[DETA] += coef * sum 0
  FIN_NEWTON,   // one Newton step on r(gamma) = eta(q_g) - eta0 - gamma*deta
};

struct DD {
  double hi, lo;
};

ATM_DEV DD two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  return DD{s, (a - (s - bb)) + (b - bb)};
}

ATM_DEV DD fast_two_sum(double a, double b) {
  const double s = a + b;
  return DD{s, b - (s - a)};
}

ATM_DEV DD dd_add(DD x, double y) {
  DD s = two_sum(x.hi, y);
  return fast_two_sum(s.hi, s.lo + x.lo);
}

ATM_DEV DD dd_add(DD x, DD y) {
  DD s = two_sum(x.hi, y.hi);
  DD t = two_sum(x.lo, y.lo);
  s = fast_two_sum(s.hi, s.lo + t.hi);
  return fast_two_sum(s.hi, s.lo + t.lo);
}

inline int red_blocks(int64_t n) {
  int64_t b = (n + RED_THREADS - 1) / RED_THREADS;
  return (int)(b > RED_MAXBLK ? RED_MAXBLK : (b < 1 ? 1 : b));
}

// Workgroup tree combination of NQ double-doubles per thread; thread 0 writes
// part[(blockIdx.x * NQ + j) * 2 + {0, 1}].
template <int NQ>
ATM_DEV void block_store(DD* v, double* __restrict__ part) {
  __shared__ double sh[NQ][2][RED_THREADS];
  const int t = threadIdx.x;
#pragma unroll
  for (int j = 0; j < NQ; ++j) {
    sh[j][0][t] = v[j].hi;
    sh[j][1][t] = v[j].lo;
  }
  __syncthreads();
  for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int j = 0; j < NQ; ++j) {
        DD a{sh[j][0][t], sh[j][1][t]}, b{sh[j][0][t + w], sh[j][1][t + w]};
        DD c = dd_add(a, b);
        sh[j][0][t] = c.hi;
        sh[j][1][t] = c.lo;
      }
    }
    __syncthreads();
  }
  if (t == 0) {
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      part[(blockIdx.x * NQ + j) * 2 + 0] = sh[j][0][0];
      part[(blockIdx.x * NQ + j) * 2 + 1] = sh[j][1][0];
    }
  }
}

template <class L>
ATM_DEV double vdot(const LawConst& c, const double* qn, const double* an, const double* kn) {
  double v[L::Ns];
  L::entropyvariables(c, qn, an, v);
  double vq = v[0] * kn[0];
#pragma unroll
  for (int s = 1; s < L::Ns; ++s) vq += v[s] * kn[s];
  return vq;
}

// RED_ENTROPY / RED_PRODUCT / RED_RELAX over all nodes
template <class L>
__global__ void __launch_bounds__(RED_THREADS)
    entropy_reduce_kernel(const DevDG p, int mode, const double* __restrict__ q, const double* __restrict__ k,
                          const double* __restrict__ q0, const double* __restrict__ st, double* __restrict__ part) {
  constexpr int Ns = L::Ns, NA = L::NA;
  if (mode == RED_RELAX && st[RX_DONE] != 0.0) return;
  const double gam = (mode == RED_RELAX) ? st[RX_GAMMA] : 0.0;
  DD acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
  const int64_t stride = (int64_t)gridDim.x * RED_THREADS;
  for (int64_t node = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; node < p.Ntot; node += stride) {
    double qn[Ns], an[NA], kn[Ns];
    load_aux<L>(p, node, an);
    const double MJ = p.MJ[node];
    if (mode == RED_RELAX) {
#pragma unroll
      for (int s = 0; s < Ns; ++s) {
        kn[s] = k[s * p.Ntot + node];
        qn[s] = q0[s * p.Ntot + node] + gam * kn[s];
      }
      acc[0] = dd_add(acc[0], MJ * L::entropy(p.c, qn, an));
      acc[1] = dd_add(acc[1], MJ * vdot<L>(p.c, qn, an, kn));
    } else {
#pragma unroll
      for (int s = 0; s < Ns; ++s) qn[s] = q[s * p.Ntot + node];
      if (mode == RED_ENTROPY) {
        acc[0] = dd_add(acc[0], MJ * L::entropy(p.c, qn, an));
      } else {
#pragma unroll
        for (int s = 0; s < Ns; ++s) kn[s] = k[s * p.Ntot + node];
        acc[0] = dd_add(acc[0], MJ * vdot<L>(p.c, qn, an, kn));
      }
    }
  }
  if (mode == RED_RELAX) block_store<2>(acc, part);
  else block_store<1>(acc, part);
}

// One RLSRK stage after k = L(q) (src/odesolvers.jl:108-117): the entropy
// production MJ*v(q).k of the stage (partials), dq = a*dq + k, q += bdt*dq.
template <class L>
__global__ void __launch_bounds__(RED_THREADS)
    rlsrk_stage_kernel(const DevDG p, double* __restrict__ q, double* __restrict__ dq, const double* __restrict__ k,
                       double a, double bdt, double* __restrict__ part) {
  constexpr int Ns = L::Ns, NA = L::NA;
  DD acc[1] = {{0.0, 0.0}};
  const int64_t stride = (int64_t)gridDim.x * RED_THREADS;
  for (int64_t node = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; node < p.Ntot; node += stride) {
    double qn[Ns], an[NA], kn[Ns];
    load_aux<L>(p, node, an);
#pragma unroll
    for (int s = 0; s < Ns; ++s) {
      qn[s] = q[s * p.Ntot + node];
      kn[s] = k[s * p.Ntot + node];
    }
    acc[0] = dd_add(acc[0], p.MJ[node] * vdot<L>(p.c, qn, an, kn));
#pragma unroll
    for (int s = 0; s < Ns; ++s) {
      const int64_t i = s * p.Ntot + node;
      const double d = a * dq[i];
      const double dn = d + kn[s];
      dq[i] = dn;
      q[i] = qn[s] + bdt * dn;
    }
  }
  block_store<1>(acc, part);
}

typedef hipError_t (*ReduceFn)(const DevDG& p, hipStream_t s, int mode, const double* q, const double* k,
                               const double* q0, const double* st, double* part);
typedef hipError_t (*RlsrkStageFn)(const DevDG& p, hipStream_t s, double* q, double* dq, const double* k, double a,
                                   double bdt, double* part);

struct MeasureFns {
  ReduceFn reduce = nullptr;      // nullptr: the law defines no entropy
  RlsrkStageFn stage = nullptr;
};

template <class L>
hipError_t launch_entropy_reduce(const DevDG& p, hipStream_t s, int mode, const double* q, const double* k,
                                 const double* q0, const double* st, double* part) {
  hipLaunchKernelGGL(entropy_reduce_kernel<L>, dim3(red_blocks(p.Ntot)), dim3(RED_THREADS), 0, s, p, mode, q, k,
                     q0, st, part);
  return hipGetLastError();
}

template <class L>
hipError_t launch_rlsrk_stage(const DevDG& p, hipStream_t s, double* q, double* dq, const double* k, double a,
                              double bdt, double* part) {
  hipLaunchKernelGGL(rlsrk_stage_kernel<L>, dim3(red_blocks(p.Ntot)), dim3(RED_THREADS), 0, s, p, q, dq, k, a, bdt,
                     part);
  return hipGetLastError();
}

template <class L>
MeasureFns measure_entry() {
  MeasureFns m;
  if constexpr (L::HAS_ENTROPY) {
    m.reduce = launch_entropy_reduce<L>;
    m.stage = launch_rlsrk_stage<L>;
  }
  return m;
}

MeasureFns resolve_measures(int law, int dim);

// law-independent pieces (inst_measures.hip)
// weighted p-norm partials per component: sum MJ*|f|^p (p > 0) or max |f| (p = inf)
hipError_t launch_wnorm(hipStream_t s, const double* q, const double* MJ, int Ns, int64_t Ntot, double pw,
                        double* part);
// combine nblk partials of nq quantities (see FinMode)
hipError_t launch_finalize(hipStream_t s, const double* part, int nblk, int nq, int mode, double coef, double* st,
                           double* out);
// y = x - z over n doubles
hipError_t launch_diff(hipStream_t s, double* y, const double* x, const double* z, int64_t n);
// q = q0 + gamma*k with gamma = st[RX_GAMMA]
hipError_t launch_relax_apply(hipStream_t s, double* q, const double* q0, const double* k, const double* st,
                              int64_t n);
// st[i] = v
hipError_t launch_set_slot(hipStream_t s, double* st, int i, double v);

}  // namespace atm
