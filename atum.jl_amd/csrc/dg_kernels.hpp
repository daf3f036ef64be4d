// DGSEM element kernels for gfx950: volume term + surface term (+ optional
// low-storage RK update) fused into one launch, one workgroup per `epw`
// elements, one thread per volume node.
//
// Summation order follows the reference kernels so that, compiled without FMA
// contraction, results match the reference to round-off of log/sqrt only:
//   weak form:          src/dgsem.jl:217-306 (volumeterm!)
//   flux differencing:  src/dgsem.jl:470-561 (esvolumeterm_dir!, one direction
//                       after the other, each summed over n, then s, then d)
//   surface:            src/dgsem.jl:683-732 (surfaceterm!, directions in order,
//                       side 1 then 2; every node sits on at most one face per
//                       direction, so node-centric evaluation keeps that order)
// Epilogue: dq = (assign ? 0 : alpha*dq) + volume, then the surface terms are
// subtracted, then optionally q_out = q + bdt*dq (src/odesolvers.jl:51-53).
#pragma once
#include "fast_flux.hpp"
#include "pairflux.hpp"

namespace atm {

struct DevDG {
  LawConst c;
  int K, Nq, Np, Nfpf, Nfp, naux, dirmask, vflux, sflux;
  int fdk;          // flux-differencing summation order: ATM_FD_PER_DIR / _SYMMETRIC / _NAIVE
  int64_t Ntot, NF;
  int epw, lds;     // elements per workgroup and LDS bytes of a single-operator launch
  const int32_t* sched;  // processing order of the elements (null: element-id order)
  const double* D;  // Nq*Nq column major
  const double* g;
  const double* MJ;
  const double* MJI;
  const double* faceMJ;
  const double* normal;
  const double* aux;
  const int64_t* faceix_p;
  const int32_t* bctag;
};

// aux components the two-point fluxes read from neighbour nodes: Φ, ρ_ref, p_ref
template <class L>
constexpr int staged_aux() {
  return (L::LAW == ATM_LAW_EULER_GRAVITY || L::LAW == ATM_LAW_LINEAR_EULER_GRAVITY) ? 3 : 0;
}

template <class L>
ATM_DEV void load_aux(const DevDG& p, int64_t node, double* a) {
#pragma unroll
  for (int s = 0; s < L::NA; ++s) a[s] = (s < p.naux) ? p.aux[s * p.Ntot + node] : 0.0;
}

template <class L, int SF>
ATM_DEV void dispatch_sflux(const DevDG& p, const double* n, const double* qm, const double* am,
                            const double* qp, const double* ap, double* nf) {
  if constexpr (SF >= 0) {
    surfaceflux<L, SF>(p.c, n, qm, am, qp, ap, nf);
  } else {
    switch (p.sflux) {
      case ATM_FLUX_CENTRAL:
        if constexpr (L::has_sflux(ATM_FLUX_CENTRAL)) surfaceflux<L, ATM_FLUX_CENTRAL>(p.c, n, qm, am, qp, ap, nf);
        break;
      case ATM_FLUX_RUSANOV:
        if constexpr (L::has_sflux(ATM_FLUX_RUSANOV)) surfaceflux<L, ATM_FLUX_RUSANOV>(p.c, n, qm, am, qp, ap, nf);
        break;
      case ATM_FLUX_EC:
        if constexpr (L::has_sflux(ATM_FLUX_EC) && L::has_vflux(ATM_FLUX_EC))
          surfaceflux<L, ATM_FLUX_EC>(p.c, n, qm, am, qp, ap, nf);
        break;
      case ATM_FLUX_KG:
        if constexpr (L::has_sflux(ATM_FLUX_KG) && L::has_vflux(ATM_FLUX_KG))
          surfaceflux<L, ATM_FLUX_KG>(p.c, n, qm, am, qp, ap, nf);
        break;
      case ATM_FLUX_ROE:
        if constexpr (L::has_sflux(ATM_FLUX_ROE) && L::LAW != ATM_LAW_ADVECTION &&
                      L::LAW != ATM_LAW_LINEAR_EULER_GRAVITY)
          surfaceflux<L, ATM_FLUX_ROE>(p.c, n, qm, am, qp, ap, nf);
        break;
      case ATM_FLUX_MATRIX:
        if constexpr (L::has_sflux(ATM_FLUX_MATRIX) && (L::LAW == ATM_LAW_EULER || L::LAW == ATM_LAW_EULER_GRAVITY))
          surfaceflux<L, ATM_FLUX_MATRIX>(p.c, n, qm, am, qp, ap, nf);
        break;
      default:
        break;
    }
  }
}

// Surface contributions of the faces this node lies on, subtracted from acc.
template <class L, int SF>
ATM_DEV void surface_node(const DevDG& p, int Nq, int64_t e, const int idx[3], const double* q,
                          const double* qn, const double* an, double MJIn, double* acc) {
  constexpr int D = L::D, Ns = L::Ns, NA = L::NA;
#pragma unroll
  for (int r = 0; r < D; ++r) {
    if (!((p.dirmask >> r) & 1)) continue;
    const int ir = idx[r];
    const int side = (ir == 0) ? 0 : ((ir == Nq - 1) ? 1 : -1);
    if (side < 0) continue;
    const int face = 2 * r + side;
    int fp = 0;
    if (D == 2) fp = (r == 0) ? idx[1] : idx[0];
    if (D == 3) fp = (r == 0) ? idx[1] + Nq * idx[2] : ((r == 1) ? idx[0] + Nq * idx[2] : idx[0] + Nq * idx[1]);
    const int64_t jj = e * p.Nfp + (int64_t)face * p.Nfpf + fp;
    // loads unconditional (a boundary face reads the element's first node as the
    // neighbour and replaces it by the boundary state), so that none waits on a
    // conditional path
    const int tag = p.bctag[e * (2 * D) + face];
    const int64_t idx_p = p.faceix_p[jj];
    const int64_t idp = tag == 0 ? idx_p : e * p.Np;
    double n[D];
#pragma unroll
    for (int d = 0; d < D; ++d) n[d] = p.normal[d * p.NF + jj];
    const double fMJ = p.faceMJ[jj];
    double qp[Ns], ap[NA], nf[Ns];
#pragma unroll
    for (int s = 0; s < Ns; ++s) qp[s] = q[s * p.Ntot + idp];
    load_aux<L>(p, idp, ap);
    if (tag != 0) {
      boundarystate<L>(p.c, n, qn, tag, qp);
#pragma unroll
      for (int s = 0; s < NA; ++s) ap[s] = an[s];
    }
    dispatch_sflux<L, SF>(p, n, qn, an, qp, ap, nf);
#pragma unroll
    for (int s = 0; s < Ns; ++s) acc[s] -= fMJ * nf[s] * MJIn;
  }
}

template <int D>
ATM_DEV void node_index(int ijk, int Nq, int idx[3]) {
  idx[0] = ijk % Nq;
  idx[1] = (D > 1) ? (ijk / Nq) % Nq : 0;
  idx[2] = (D > 2) ? ijk / (Nq * Nq) : 0;
}

template <class L>
ATM_DEV void epilogue(const DevDG& p, int64_t node, const double* q1, const double* acc,
                      double* __restrict__ dq, double* __restrict__ qout, double bdt) {
#pragma unroll
  for (int s = 0; s < L::Ns; ++s) {
    dq[s * p.Ntot + node] = acc[s];
    if (qout) qout[s * p.Ntot + node] = q1[s] + bdt * acc[s];
  }
}

// ---------------------------------------------------------------------------
// Flux-differencing form, all directions of the mask in one launch, in one of
// the reference's three summation orders (p.fdk; the Nq-specialised kernels are
// built for :per_dir only, the others run the generic runtime-Nq kernels):
//   :per_dir            src/dgsem.jl:470-561, direction after direction, each
//                       summed over n, then s, then d;
//   :per_dir_symmetric  src/dgsem.jl:563-681, per direction two sums with
//                       halved terms, n ascending and n descending, added;
//   :naive              src/dgsem.jl:308-468, one sum over n with the
//                       directions interleaved inside it (all directions).
// Per-node primitives of the two-point flux live in LDS for the whole element;
// the metric row (MJ*g_r) of the direction being summed is staged in LDS one
// direction at a time (the naive order reads the neighbours' rows from memory).
template <class L, int NQ, int VF, int SF>
ATM_DEV void fd_body(const DevDG& p, double* lds, int64_t blk, const double* __restrict__ q, double* __restrict__ dq, double alpha,
          int assign, double* __restrict__ qout, double bdt, int epw) {
  using PF = PairFlux<L, VF>;
  constexpr int D = L::D, Ns = L::Ns, NA = L::NA, NP = PF::NP;
  constexpr int FD_PARTNER_UNROLL = D == 3 ? 1 : 8;
  const int Nq = NQ > 0 ? NQ : p.Nq;
  const int Np = NQ > 0 ? (D == 1 ? NQ : (D == 2 ? NQ * NQ : NQ * NQ * NQ)) : p.Np;
  double* sD = lds;
  const int stride = (NP + D) * Np;
  const int tid = threadIdx.x;
  const int el = tid / Np;
  const int ijk = tid - el * Np;
  const int64_t le = blk * epw + el;  // position in the processing order
  const bool active = el < epw && le < p.K;
  const int64_t e = (p.sched && active) ? (int64_t)p.sched[le] : le;
  double* sP = lds + Nq * Nq + el * stride;  // [k][node]
  double* slg = sP + NP * Np;                // [d][node], current direction
  for (int t = tid; t < Nq * Nq; t += blockDim.x) sD[t] = p.D[t];
  double q1[Ns], a1[NA], P1[NP];
  const int64_t node = e * Np + ijk;
  // every load first and unconditionally (a lane past the last element reads node 0 and
  // discards it): state, aux, MJ, MJI, the metric rows of every direction and the stage's
  // dq in flight together (one memory latency), then the primitives
  const int64_t nl = active ? node : 0;
#pragma unroll
  for (int s = 0; s < Ns; ++s) q1[s] = q[s * p.Ntot + nl];
  load_aux<L>(p, nl, a1);
  const double MJn = p.MJ[nl];
  const double MJIn = p.MJI[nl];
  double gall[D][D];
#pragma unroll
  for (int r = 0; r < D; ++r)
#pragma unroll
    for (int d = 0; d < D; ++d) gall[r][d] = p.g[(int64_t)(r + D * d) * p.Ntot + nl];
  double dq0[Ns];
#pragma unroll
  for (int s = 0; s < Ns; ++s) dq0[s] = 0.0;
  if (!assign) {  // uniform over the launch
#pragma unroll
    for (int s = 0; s < Ns; ++s) dq0[s] = dq[s * p.Ntot + nl];
  }
  PF::prim(p.c, q1, a1, P1);
  if (active) {
#pragma unroll
    for (int k = 0; k < NP; ++k) sP[k * Np + ijk] = P1[k];
  }
  int idx[3];
  node_index<D>(ijk, Nq, idx);
  double acc[Ns];
  const int fdk = (NQ > 0) ? ATM_FD_PER_DIR : p.fdk;
  if (fdk == ATM_FD_NAIVE) {
    // src/dgsem.jl:308-468: sources of all directions, then for each n the pair
    // of every direction (requires all directions; checked at creation)
    __syncthreads();  // primitives written
    if (!active) return;
    double dl[Ns];
#pragma unroll
    for (int s = 0; s < Ns; ++s) dl[s] = -0.0;
    L::sources(p.c, dl, q1, a1, p.dirmask, false);
    for (int n = 0; n < Nq; ++n) {
#pragma unroll
      for (int r = 0; r < D; ++r) {
        const int id = idx[r];
        const int lstride = (r == 0) ? 1 : ((r == 1) ? Nq : Nq * Nq);
        const int ild = ijk - id * lstride + n * lstride;
        double P2[NP];
#pragma unroll
        for (int k = 0; k < NP; ++k) P2[k] = sP[k * Np + ild];
        double F[D][Ns];
        PF::pair(p.c, P1, P2, F);
        const double Din = MJIn * sD[id + Nq * n];
        const double Dni = MJIn * sD[n + Nq * id];
        const int64_t nn = e * Np + ild;
        const double MJ2 = p.MJ[nn];
        double lg1[D], lg2[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
          lg1[d] = MJn * gall[r][d];
          lg2[d] = MJ2 * p.g[(int64_t)(r + D * d) * p.Ntot + nn];
        }
#pragma unroll
        for (int s = 0; s < Ns; ++s) {
#pragma unroll
          for (int d = 0; d < D; ++d) {
            dl[s] -= Din * lg1[d] * F[d][s];
            dl[s] += F[d][s] * lg2[d] * Dni;
          }
        }
      }
    }
#pragma unroll
    for (int s = 0; s < Ns; ++s) acc[s] = assign ? dl[s] : alpha * dq0[s] + dl[s];
    surface_node<L, SF>(p, Nq, e, idx, q, q1, a1, MJIn, acc);
    epilogue<L>(p, node, q1, acc, dq, qout, bdt);
    return;
  }
  const bool sym = fdk == ATM_FD_PER_DIR_SYMMETRIC;
  bool first = true;
#pragma unroll
  for (int r = 0; r < D; ++r) {
    if (!((p.dirmask >> r) & 1)) continue;
    __syncthreads();  // primitives written / previous direction's metric row consumed
    if (active) {
#pragma unroll
      for (int d = 0; d < D; ++d) slg[d * Np + ijk] = MJn * gall[r][d];
    }
    __syncthreads();
    if (!active) continue;
    double dl[Ns];
#pragma unroll
    for (int s = 0; s < Ns; ++s) dl[s] = -0.0;
    L::sources(p.c, dl, q1, a1, 1 << r, false);
    const int id = idx[r];
    const int lstride = (r == 0) ? 1 : ((r == 1) ? Nq : Nq * Nq);
    const int lbase = ijk - id * lstride;
    double lgi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) lgi[d] = slg[d * Np + ijk];
    // the partner loop of the 3-D kernels is not unrolled: unrolled, the small-Nq kernels hold
    // several partners' primitives at once and drop to 2 waves per SIMD (Nq = 4: 210 VGPRs)
#pragma unroll FD_PARTNER_UNROLL
    for (int n = 0; n < Nq; ++n) {
      const int ild = lbase + n * lstride;
      double P2[NP];
#pragma unroll
      for (int k = 0; k < NP; ++k) P2[k] = sP[k * Np + ild];
      double F[D][Ns];
      PF::pair(p.c, P1, P2, F);
      const double Ddn = MJIn * sD[id + Nq * n];
      const double Dnd = MJIn * sD[n + Nq * id];
      double lgn[D];
#pragma unroll
      for (int d = 0; d < D; ++d) lgn[d] = slg[d * Np + ild];
      if (sym) {
#pragma unroll
        for (int s = 0; s < Ns; ++s) {
#pragma unroll
          for (int d = 0; d < D; ++d) {
            dl[s] -= Ddn * lgi[d] * F[d][s] / 2;
            dl[s] += F[d][s] * lgn[d] * Dnd / 2;
          }
        }
      } else {
#pragma unroll
        for (int s = 0; s < Ns; ++s) {
#pragma unroll
          for (int d = 0; d < D; ++d) {
            dl[s] -= Ddn * lgi[d] * F[d][s];
            dl[s] += F[d][s] * lgn[d] * Dnd;
          }
        }
      }
    }
    if (sym) {
      // the second sum of the symmetric order: n descending, from -0.0
      double dl2[Ns];
#pragma unroll
      for (int s = 0; s < Ns; ++s) dl2[s] = -0.0;
      for (int n = Nq - 1; n >= 0; --n) {
        const int ild = lbase + n * lstride;
        double P2[NP];
#pragma unroll
        for (int k = 0; k < NP; ++k) P2[k] = sP[k * Np + ild];
        double F[D][Ns];
        PF::pair(p.c, P1, P2, F);
        const double Ddn = MJIn * sD[id + Nq * n];
        const double Dnd = MJIn * sD[n + Nq * id];
        double lgn[D];
#pragma unroll
        for (int d = 0; d < D; ++d) lgn[d] = slg[d * Np + ild];
#pragma unroll
        for (int s = 0; s < Ns; ++s) {
#pragma unroll
          for (int d = 0; d < D; ++d) {
            dl2[s] -= Ddn * lgi[d] * F[d][s] / 2;
            dl2[s] += F[d][s] * lgn[d] * Dnd / 2;
          }
        }
      }
#pragma unroll
      for (int s = 0; s < Ns; ++s) dl[s] = dl[s] + dl2[s];
    }
    if (first) {
#pragma unroll
      for (int s = 0; s < Ns; ++s) acc[s] = assign ? dl[s] : alpha * dq0[s] + dl[s];
      first = false;
    } else {
#pragma unroll
      for (int s = 0; s < Ns; ++s) acc[s] = acc[s] + dl[s];
    }
  }
  if (!active) return;
  if (first) {
#pragma unroll
    for (int s = 0; s < Ns; ++s) acc[s] = assign ? 0.0 : alpha * dq0[s];
  }
  surface_node<L, SF>(p, Nq, e, idx, q, q1, a1, MJIn, acc);
  epilogue<L>(p, node, q1, acc, dq, qout, bdt);
}

// ---------------------------------------------------------------------------
// Weak form (volumeterm! order).
template <class L, int NQ, int SF>
ATM_DEV void weak_body(const DevDG& p, double* lds, int64_t blk, const double* __restrict__ q, double* __restrict__ dq, double alpha,
            int assign, double* __restrict__ qout, double bdt, int epw) {
  constexpr int D = L::D, Ns = L::Ns, NA = L::NA;
  const int Nq = NQ > 0 ? NQ : p.Nq;
  const int Np = NQ > 0 ? (D == 1 ? NQ : (D == 2 ? NQ * NQ : NQ * NQ * NQ)) : p.Np;
  double* sD = lds;
  const int stride = D * Ns * Np;
  const int tid = threadIdx.x;
  const int el = tid / Np;
  const int ijk = tid - el * Np;
  const int64_t le = blk * epw + el;  // position in the processing order
  const bool active = el < epw && le < p.K;
  const int64_t e = (p.sched && active) ? (int64_t)p.sched[le] : le;
  double* slF = lds + Nq * Nq + el * stride;  // [dir slot][s][node]
  for (int t = tid; t < Nq * Nq; t += blockDim.x) sD[t] = p.D[t];
  double q1[Ns], a1[NA];
  const int64_t node = e * Np + ijk;
  // every load first and unconditionally (a lane past the last element reads node 0 and
  // discards it): state, aux, metric rows, MJ, MJI and the stage's dq in flight together
  const int64_t nl = active ? node : 0;
#pragma unroll
  for (int s = 0; s < Ns; ++s) q1[s] = q[s * p.Ntot + nl];
  load_aux<L>(p, nl, a1);
  double gl[D][D];
#pragma unroll
  for (int r = 0; r < D; ++r)
#pragma unroll
    for (int dd = 0; dd < D; ++dd) gl[r][dd] = p.g[(int64_t)(r + D * dd) * p.Ntot + nl];
  const double MJn = p.MJ[nl];
  const double MJIn = p.MJI[nl];
  double dq0[Ns];
#pragma unroll
  for (int s = 0; s < Ns; ++s) dq0[s] = 0.0;
  if (!assign) {  // uniform over the launch
#pragma unroll
    for (int s = 0; s < Ns; ++s) dq0[s] = dq[s * p.Ntot + nl];
  }
  if (active) {
    double F[D][Ns];
    L::flux(p.c, q1, a1, F);
#pragma unroll
    for (int s = 0; s < Ns; ++s) {
      int slot = 0;
#pragma unroll
      for (int r = 0; r < D; ++r) {
        if (!((p.dirmask >> r) & 1)) continue;
        double v = 0;
#pragma unroll
        for (int dd = 0; dd < D; ++dd) v += gl[r][dd] * F[dd][s];
        v *= MJn;
        slF[(slot * Ns + s) * Np + ijk] = v;
        ++slot;
      }
    }
  }
  __syncthreads();
  if (!active) return;
  int idx[3];
  node_index<D>(ijk, Nq, idx);
  double dl[Ns];
#pragma unroll
  for (int s = 0; s < Ns; ++s) dl[s] = -0.0;
  L::sources(p.c, dl, q1, a1, p.dirmask, true);
  const int i = idx[0], j = idx[1], k = idx[2];
  for (int n = 0; n < Nq; ++n) {
    const double Dni = sD[n + Nq * i] * MJIn;
    const double Dnj = sD[n + Nq * j] * MJIn;
    const double Dnk = sD[n + Nq * k] * MJIn;
#pragma unroll
    for (int s = 0; s < Ns; ++s) {
      int slot = 0;
      if (p.dirmask & 1) {
        dl[s] += Dni * slF[(slot * Ns + s) * Np + (ijk - i + n)];
        ++slot;
      }
      if (D > 1 && (p.dirmask & 2)) {
        dl[s] += Dnj * slF[(slot * Ns + s) * Np + (ijk + (n - j) * Nq)];
        ++slot;
      }
      if (D > 2 && (p.dirmask & 4)) {
        dl[s] += Dnk * slF[(slot * Ns + s) * Np + (ijk + (n - k) * Nq * Nq)];
      }
    }
  }
  double acc[Ns];
#pragma unroll
  for (int s = 0; s < Ns; ++s) acc[s] = assign ? dl[s] : alpha * dq0[s] + dl[s];
  surface_node<L, SF>(p, Nq, e, idx, q, q1, a1, MJIn, acc);
  epilogue<L>(p, node, q1, acc, dq, qout, bdt);
}

// ---------------------------------------------------------------------------
// Batched launches: one launch advances many operators that share a kernel.
// The device table lists, per instance, its operator, its state buffers and its
// time step, and the first workgroup of its element range (a flattened element
// table); a workgroup finds its instance by binary search.
struct BatchEntry {
  DevDG p;
  double* buf[2];  // stage reads buf[parity], writes buf[1 - parity]
  double* dq;
  double dt;
  int64_t block_start;
};

// Workgroups are dealt to the 8 XCDs round-robin (workgroup w runs on XCD w % 8),
// each XCD with its own L2.  The logical block index gives XCD x a contiguous
// range of blocks (consecutive elements, whose faces neighbour each other), so a
// neighbour's face states are read from the same L2.  A bijection of [0, G).
ATM_DEV int64_t xcd_block() {
  const int64_t G = gridDim.x, w = blockIdx.x;
  const int64_t x = w & 7, i = w >> 3, base = G >> 3, rem = G & 7;
  return x * base + (x < rem ? x : rem) + i;
}

ATM_DEV int batch_find(const BatchEntry* __restrict__ tab, int n, int64_t b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab[mid].block_start <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

#define ATM_KERNEL_PAIR(NAME, BODY, TPARAMS, TARGS)                                                      \
  template <TPARAMS>                                                                                      \
  __global__ void __launch_bounds__(512)                                                                  \
      NAME(const DevDG p, const double* __restrict__ q, double* __restrict__ dq, double alpha, int assign, \
           double* __restrict__ qout, double bdt, int epw) {                                              \
    extern __shared__ double lds[];                                                                       \
    BODY<TARGS>(p, lds, xcd_block(), q, dq, alpha, assign, qout, bdt, epw);                               \
  }                                                                                                       \
  template <TPARAMS>                                                                                      \
  __global__ void __launch_bounds__(512)                                                                  \
      NAME##_batch(const BatchEntry* __restrict__ tab, int n, int parity, double alpha, int assign,      \
                   double b, int epw) {                                                                   \
    extern __shared__ double lds[];                                                                       \
    const int64_t lb = xcd_block();                                                                       \
    const int i = batch_find(tab, n, lb);                                                                 \
    const DevDG p = tab[i].p;                                                                             \
    BODY<TARGS>(p, lds, lb - tab[i].block_start, tab[i].buf[parity], tab[i].dq, alpha,                    \
                assign, tab[i].buf[1 - parity], b * tab[i].dt, epw);                                      \
  }

#define ATM_COMMA ,
ATM_KERNEL_PAIR(fd_kernel, fd_body, class L ATM_COMMA int NQ ATM_COMMA int VF ATM_COMMA int SF, L ATM_COMMA NQ ATM_COMMA VF ATM_COMMA SF)
ATM_KERNEL_PAIR(weak_kernel, weak_body, class L ATM_COMMA int NQ ATM_COMMA int SF, L ATM_COMMA NQ ATM_COMMA SF)

typedef hipError_t (*BatchLaunchFn)(const BatchEntry* tab, int n, int64_t blocks, int threads, size_t lds,
                                    int epw, hipStream_t st, int parity, double alpha, int assign, double b);

template <class L, int NQ, int VF, int SF>
hipError_t launch_fd_batch(const BatchEntry* tab, int n, int64_t blocks, int threads, size_t lds, int epw,
                           hipStream_t st, int parity, double alpha, int assign, double b) {
  hipLaunchKernelGGL((fd_kernel_batch<L, NQ, VF, SF>), dim3(blocks), dim3(threads), lds, st, tab, n, parity,
                     alpha, assign, b, epw);
  return hipGetLastError();
}

template <class L, int NQ, int SF>
hipError_t launch_weak_batch(const BatchEntry* tab, int n, int64_t blocks, int threads, size_t lds, int epw,
                             hipStream_t st, int parity, double alpha, int assign, double b) {
  hipLaunchKernelGGL((weak_kernel_batch<L, NQ, SF>), dim3(blocks), dim3(threads), lds, st, tab, n, parity,
                     alpha, assign, b, epw);
  return hipGetLastError();
}

// One launch per stage for weak-form members of different degrees (N = 3..7,
// Nq = 4..8): each workgroup finds its member in the flattened table and runs
// that member's Nq-specialised body with the member's own elements per
// workgroup (p.epw) and LDS share (p.lds <= the launch's LDS); threads beyond
// a member's epw * Np are inactive, as in every element kernel.
template <class L, int SF>
__global__ void __launch_bounds__(512)
    weak_multi_kernel_batch(const BatchEntry* __restrict__ tab, int n, int parity, double alpha, int assign,
                            double b) {
  extern __shared__ double lds[];
  const int64_t lb = xcd_block();
  const int i = batch_find(tab, n, lb);
  const DevDG p = tab[i].p;
  const int64_t blk = lb - tab[i].block_start;
  const double* qin = tab[i].buf[parity];
  double* qo = tab[i].buf[1 - parity];
  const double bdt = b * tab[i].dt;
  switch (p.Nq) {
    case 4: weak_body<L, 4, SF>(p, lds, blk, qin, tab[i].dq, alpha, assign, qo, bdt, p.epw); break;
    case 5: weak_body<L, 5, SF>(p, lds, blk, qin, tab[i].dq, alpha, assign, qo, bdt, p.epw); break;
    case 6: weak_body<L, 6, SF>(p, lds, blk, qin, tab[i].dq, alpha, assign, qo, bdt, p.epw); break;
    case 7: weak_body<L, 7, SF>(p, lds, blk, qin, tab[i].dq, alpha, assign, qo, bdt, p.epw); break;
    case 8: weak_body<L, 8, SF>(p, lds, blk, qin, tab[i].dq, alpha, assign, qo, bdt, p.epw); break;
    default: break;
  }
}

// `epw` is unused: every member carries its own in the table.
template <class L, int SF>
hipError_t launch_weak_multi_batch(const BatchEntry* tab, int n, int64_t blocks, int threads, size_t lds, int epw,
                                   hipStream_t st, int parity, double alpha, int assign, double b) {
  (void)epw;
  hipLaunchKernelGGL((weak_multi_kernel_batch<L, SF>), dim3(blocks), dim3(threads), lds, st, tab, n, parity, alpha,
                     assign, b);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
typedef hipError_t (*LaunchFn)(const DevDG& p, hipStream_t st, const double* q, double* dq,
                               double alpha, int assign, double* qout, double bdt);

inline int elements_per_workgroup(int Np) { return Np >= 256 ? 1 : 256 / Np; }

template <class L, int NQ, int VF, int SF>
hipError_t launch_fd(const DevDG& p, hipStream_t st, const double* q, double* dq, double alpha,
                     int assign, double* qout, double bdt) {
  const int epw = p.epw;
  const int threads = ((epw * p.Np + 63) / 64) * 64;
  if (epw < 1 || threads > 512) return hipErrorInvalidConfiguration;
  const size_t lds = p.lds;
  const int blocks = (p.K + epw - 1) / epw;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL((fd_kernel<L, NQ, VF, SF>), dim3(blocks), dim3(threads), lds, st, p, q, dq,
                     alpha, assign, qout, bdt, epw);
  return hipGetLastError();
}

template <class L, int NQ, int SF>
hipError_t launch_weak(const DevDG& p, hipStream_t st, const double* q, double* dq, double alpha,
                       int assign, double* qout, double bdt) {
  const int epw = p.epw;
  const int threads = ((epw * p.Np + 63) / 64) * 64;
  if (epw < 1 || threads > 512) return hipErrorInvalidConfiguration;
  const size_t lds = p.lds;
  const int blocks = (p.K + epw - 1) / epw;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL((weak_kernel<L, NQ, SF>), dim3(blocks), dim3(threads), lds, st, p, q, dq,
                     alpha, assign, qout, bdt, epw);
  return hipGetLastError();
}

// LDS bytes per element a configuration needs (the workgroup adds Nq*Nq doubles
// for D); the total is checked against the device limit whenever epw is chosen
template <class L>
size_t lds_elem_bytes(int form, int vflux, int Np) {
  constexpr int D = L::D, Ns = L::Ns;
  if (form == ATM_FORM_WEAK) return sizeof(double) * (size_t)D * Ns * Np;
  int np = D * Ns;  // central: physical flux per node
  if (vflux != ATM_FLUX_CENTRAL) np = (L::LAW == ATM_LAW_SHALLOW_WATER) ? 2 * D + 3 : D + 4;
  return sizeof(double) * (size_t)(np + D) * Np;
}

// A kernel family for an operator: launchers, the default elements per
// workgroup and the LDS bytes of a workgroup, lds_fixed + epw * lds_elem.
struct Resolved {
  LaunchFn fn;
  size_t lds;  // at the default epw
  int ns;
  BatchLaunchFn bfn;
  int epw;     // default elements per workgroup
  size_t lds_fixed, lds_elem;
  int max_threads = 512;  // largest workgroup the kernels are built for
  // batched launcher shared by the members of one family of different Nq (one
  // launch per stage for all of them), or null
  BatchLaunchFn mbfn = nullptr;
};

inline Resolved make_resolved(LaunchFn fn, BatchLaunchFn bfn, int ns, int epw, size_t fixed, size_t elem,
                              int max_threads = 512) {
  return Resolved{fn, fixed + (size_t)epw * elem, ns, bfn, epw, fixed, elem, max_threads};
}

template <class L, int NQ, int VF, int SF>
Resolved make_fd(int Nq, int Np) {
  return make_resolved(launch_fd<L, NQ, VF, SF>, launch_fd_batch<L, NQ, VF, SF>, L::Ns, elements_per_workgroup(Np),
                       sizeof(double) * Nq * Nq, lds_elem_bytes<L>(ATM_FORM_FLUX_DIFFERENCING, VF, Np));
}

template <class L, int NQ, int SF>
Resolved make_weak(int Nq, int Np) {
  return make_resolved(launch_weak<L, NQ, SF>, launch_weak_batch<L, NQ, SF>, L::Ns, elements_per_workgroup(Np),
                       sizeof(double) * Nq * Nq, lds_elem_bytes<L>(ATM_FORM_WEAK, 0, Np));
}

// Generic (runtime Nq, runtime surface flux) kernel for a law/dimension.
template <class L>
Resolved resolve_generic(int form, int vflux, int Nq, int Np) {
  Resolved r = make_resolved(nullptr, nullptr, L::Ns, elements_per_workgroup(Np), sizeof(double) * Nq * Nq,
                             lds_elem_bytes<L>(form, vflux, Np));
  if (form == ATM_FORM_WEAK) {
    r.fn = launch_weak<L, 0, -1>;
    r.bfn = launch_weak_batch<L, 0, -1>;
  } else {
    if constexpr (L::has_vflux(ATM_FLUX_CENTRAL))
      if (vflux == ATM_FLUX_CENTRAL) {
        r.fn = launch_fd<L, 0, ATM_FLUX_CENTRAL, -1>;
        r.bfn = launch_fd_batch<L, 0, ATM_FLUX_CENTRAL, -1>;
      }
    if constexpr (L::has_vflux(ATM_FLUX_EC))
      if (vflux == ATM_FLUX_EC) {
        r.fn = launch_fd<L, 0, ATM_FLUX_EC, -1>;
        r.bfn = launch_fd_batch<L, 0, ATM_FLUX_EC, -1>;
      }
    if constexpr (L::has_vflux(ATM_FLUX_KG))
      if (vflux == ATM_FLUX_KG) {
        r.fn = launch_fd<L, 0, ATM_FLUX_KG, -1>;
        r.bfn = launch_fd_batch<L, 0, ATM_FLUX_KG, -1>;
      }
  }
  return r;
}

}  // namespace atm
