// Pair-symmetric flux-differencing kernel (ATM_PRECISION_FAST).
//
// Same operator as the :per_dir kernels of src/dgsem.jl:470-561,
//   dq_i += MJI_i * sum_r sum_n sum_d G_in[d] * F#(q_i, q_n)[d, s],
//   G_in[d] = D[n,i] (MJ g_r)_n[d] - D[i,n] (MJ g_r)_i[d],
// evaluated in a different order: G_ni = -G_in and G_ii = 0, so every
// unordered pair {a, b} of a line is evaluated once (its diagonal term is
// identically zero) and its contribution is given to both nodes.  The
// entropy-conservative flux is symmetric except for the gravity fluctuation
// term -alpha_a (Phi_a - Phi_b) I with alpha_a = b_avg rho_log / (2 b_a)
// (src/balancelaws/euler_gravity.jl:231-233), which is applied per side.
// The logarithmic means use the series of f/atanh(f) and atanh(f)/f in
// u = f^2, f = (a-b)/(a+b), for u < 0.01 (8 terms, truncation < 1e-17) and
// logarithms otherwise, so a pair costs two reciprocals.  Results equal the
// reference to round-off; the summation order (and hence the last bits)
// differs from the reference's.
//
// Work split: thread = node; in direction r the node at line position a
// evaluates the pairs {a, a+m mod Nq} for m = 1..M (tournament pairing; for
// even Nq the pairs at distance Nq/2 are taken by the lower half only) and
// writes the partner's share to an LDS slot that the partner adds after a
// barrier, in a fixed order (deterministic).
#pragma once
#include "dg_kernels.hpp"

namespace atm {

// Series coefficients in u = f^2: LM_COEF[k] of f/atanh(f) = sum_k LM_COEF[k] u^k and
// LM_COEF[8 + k] = 1/(2k+1) of atanh(f)/f = sum_k u^k/(2k+1).  Kept in constant memory:
// the kernel loads them once into scalar registers, where literal 64-bit constants
// would be rebuilt in vector registers at every use inside the pair loop.
static __constant__ double LM_COEF[16] = {1.0, -0.33333333333333331, -0.088888888888888892, -0.046560846560846560,
                                   -0.030194003527336862, -0.021796804019026242, -0.016787551856334924,
                                   -0.013502765051265932, 1.0, 1.0 / 3, 1.0 / 5, 1.0 / 7, 1.0 / 9, 1.0 / 11,
                                   1.0 / 13, 1.0 / 15};
__device__ __forceinline__ double lm_inv_series(double u) {
#pragma clang fp contract(fast)
  double s = LM_COEF[7];
#pragma unroll
  for (int k = 6; k >= 0; --k) s = s * u + LM_COEF[k];
  return s;
}

__device__ __forceinline__ double lm_series(double u) {
#pragma clang fp contract(fast)
  double s = LM_COEF[15];
#pragma unroll
  for (int k = 6; k >= 0; --k) s = s * u + LM_COEF[8 + k];
  return s;
}

constexpr double LM_U0 = 0.01;

// doubles per element of the LDS exchange buffer: pair shares (M = NQ/2 slots of Ns per
// node; for even NQ the last slot, at pair distance NQ/2, receives shares only in the upper
// half of each line and holds Ns per node of that half) and, afterwards, the face-point
// surface results (Ns per face point)
template <class L, int NQ>
constexpr int fast_xbuf() {
  constexpr int D = L::D, Ns = L::Ns;
  constexpr int Np = D == 1 ? NQ : (D == 2 ? NQ * NQ : NQ * NQ * NQ);
  constexpr int a = NQ % 2 == 0 ? (NQ / 2 - 1) * Ns * Np + Ns * (Np / 2) : (NQ / 2) * Ns * Np;
  constexpr int b = 2 * D * (Np / NQ) * Ns;
  return a > b ? a : b;
}

// The lean register and LDS form of the kernel (fd_fast_body) for 3-D elements of N >= 4.
template <class L, int NQ>
constexpr bool fast_lean() {
  return L::D == 3 && NQ >= 5;
}

// Per-node primitives: [rho, u (D), b, u.u, phi, 1/b]
template <int D, bool GRAV>
struct FastEC {
  static constexpr int Ns = D + 2, NP = D + 5;
  ATM_DEV static void prim(const LawConst& c, const double* q, const double* a, double* P) {
#pragma clang fp contract(fast)
    const double rho = q[0];
    const double ir = fast_rcp(rho);
    double u[D];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = q[1 + d] * ir;
    const double phi = GRAV ? a[0] : 0.0;
    double ke = u[0] * q[1];
#pragma unroll
    for (int d = 1; d < D; ++d) ke += u[d] * q[1 + d];
    double p = (c.gamma - 1) * (q[D + 1] - 0.5 * ke);
    if (GRAV) p = p - (c.gamma - 1) * rho * phi;
    const double ib = 2 * p * ir;  // 1/b
    P[0] = rho;
#pragma unroll
    for (int d = 0; d < D; ++d) P[1 + d] = u[d];
    P[D + 1] = fast_rcp(ib);
    double uu = u[0] * u[0];
#pragma unroll
    for (int d = 1; d < D; ++d) uu += u[d] * u[d];
    P[D + 2] = uu;
    P[D + 3] = phi;
    P[D + 4] = ib;
  }

  // The symmetric part of F#(a, b) has the form (u = the pair average of the velocity)
  //   F[d] = [rho_log u_d,  rho_log u_d u + pd e_d,  u_d (rho_log (cc + |u|^2) + pd)],
  // so its contraction with the metric weights G needs only Gu = G.u:
  //   G.F = [rho_log Gu,  u (rho_log Gu) + pd G,  Gu (rho_log (cc + |u|^2) + pd)].
  // pair() returns rho_log, pd, u and the energy coefficient, and the two gravity
  // fluctuation scalars.
  struct Pair {
    double rho_log, pd, ecoef, u[D], fla, flb;
  };
  ATM_DEV static void contract(const Pair& w, const double* G, double* gf) {
#pragma clang fp contract(fast)
    double Gu = G[0] * w.u[0];
#pragma unroll
    for (int d = 1; d < D; ++d) Gu += G[d] * w.u[d];
    gf[0] = w.rho_log * Gu;
#pragma unroll
    for (int e = 0; e < D; ++e) gf[1 + e] = w.u[e] * gf[0] + w.pd * G[e];
    gf[D + 1] = Gu * w.ecoef;
  }
  ATM_DEV static void pair(const LawConst& c, double inv_gm1, const double* Pa, const double* Pb, Pair& w) {
#pragma clang fp contract(fast)
    const double ra = Pa[0], rb = Pb[0], ba = Pa[D + 1], bb = Pb[D + 1];
    const double sr = ra + rb, sb = ba + bb;
    const double isr = fast_rcp(sr), isb = fast_rcp(sb);
    const double fr = (ra - rb) * isr, fb = (ba - bb) * isb;
    const double ur = fr * fr, ub = fb * fb;
    double Gr, Fb;
    if (ur < LM_U0 && ub < LM_U0) {
      Gr = lm_inv_series(ur);
      Fb = lm_series(ub);
    } else {
      Gr = (ur < LM_U0) ? lm_inv_series(ur) : 2 * fr / log(ra / rb);
      Fb = (ub < LM_U0) ? lm_series(ub) : log(ba / bb) / (2 * fb);
    }
    const double rho_log = 0.5 * sr * Gr;
    const double pd = 0.5 * sr * isb;              // rho_avg / (2 b_avg)
    const double inv2bl = Fb * isb * inv_gm1;      // 1 / (2 (gamma-1) b_log)
    double uu = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      w.u[d] = 0.5 * (Pa[1 + d] + Pb[1 + d]);
      uu += w.u[d] * w.u[d];
    }
    double cc = inv2bl - 0.25 * (Pa[D + 2] + Pb[D + 2]);
    if (GRAV) cc += 0.5 * (Pa[D + 3] + Pb[D + 3]);
    w.rho_log = rho_log;
    w.pd = pd;
    w.ecoef = rho_log * (cc + uu) + pd;
    if (GRAV) {
      const double wg = 0.25 * sb * rho_log;  // b_avg rho_log / 2
      const double dphi = Pa[D + 3] - Pb[D + 3];
      w.fla = wg * Pa[D + 4] * dphi;
      w.flb = -wg * Pb[D + 4] * dphi;
    } else {
      w.fla = w.flb = 0.0;
    }
  }
};

// The surface term of the FAST kernels: an element's face points dealt out over its threads
// (point f = t + i Np in round i, faces of the directions in the mask), each writing faceMJ*nf
// to buf[s * NFP + f], f = face*NFPF + fp; gather() then subtracts a node's slots times its MJI
// in direction order, as the node-centric form of the reference-order kernels does (the same
// products, (faceMJ*nf)*MJI, in the same order).  Everything that indexes a face point is
// compile-time.  The element's own state and geopotential are read from LDS, where the
// threads staged them by node; each round's boundary tag and neighbour index are loaded by
// load() before the barrier that precedes run(), so the neighbour's state is the only global
// load on the path to each flux.
template <class L, int SF, int NQ>
struct FastFaces {
  static constexpr int D = L::D, Ns = L::Ns, NA = L::NA;
  static constexpr int Np = D == 1 ? NQ : (D == 2 ? NQ * NQ : NQ * NQ * NQ);
  static constexpr int NFPF = Np / NQ, NFP = 2 * D * NFPF;
  static constexpr int R = (NFP + Np - 1) / Np;  // rounds
  int tag[R];
  int64_t nb[R];
  // a round past the last face point reads the last face point's entries and ignores them
  ATM_DEV void load(const DevDG& p, int64_t e, int t) {
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int f = min(t + i * Np, NFP - 1);
      tag[i] = p.bctag[e * (2 * D) + f / NFPF];
      nb[i] = p.faceix_p[e * p.Nfp + f];
    }
  }
  // sQ: [s][node] the element's state (s < Ns) and, for Euler-gravity, its geopotential (s = Ns)
  ATM_DEV void run(const DevDG& p, int64_t e, int t, const double* q, double* buf, const double* sQ) const {
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int f = t + i * Np;
      if (f >= NFP) break;
      const int face = f / NFPF, fp = f - face * NFPF, r = face >> 1, side = face & 1;
      if (!((p.dirmask >> r) & 1)) continue;
      int id[3] = {0, 0, 0}, rem = fp;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (d == r) continue;
        id[d] = rem % NQ;
        rem /= NQ;
      }
      id[r] = side ? NQ - 1 : 0;
      const int ln = id[0] + NQ * (id[1] + NQ * id[2]);
      const int64_t node = e * Np + ln;
      const int64_t jj = e * p.Nfp + f;
      // a boundary face point reads its own node as the neighbour (every load unconditional)
      // and replaces it by the boundary state below
      const int64_t idp = tag[i] == 0 ? nb[i] : node;
      double qm[Ns], am[NA], n[D], qp[Ns], ap[NA], nf[Ns];
#pragma unroll
      for (int s = 0; s < Ns; ++s) qm[s] = sQ[s * Np + ln];
      load_aux<L>(p, node, am);
      if constexpr (L::LAW == ATM_LAW_EULER_GRAVITY) am[0] = sQ[Ns * Np + ln];
#pragma unroll
      for (int d = 0; d < D; ++d) n[d] = p.normal[d * p.NF + jj];
      const double fMJ = p.faceMJ[jj];
#pragma unroll
      for (int s = 0; s < Ns; ++s) qp[s] = q[s * p.Ntot + idp];
      load_aux<L>(p, idp, ap);
      if (tag[i] != 0) {
        boundarystate<L>(p.c, n, qm, tag[i], qp);
#pragma unroll
        for (int s = 0; s < NA; ++s) ap[s] = am[s];
      }
      if constexpr (has_fast_sflux<L, SF>()) fast_sflux<L, SF>(p.c, n, qm, am, qp, ap, nf);
      else dispatch_sflux<L, SF>(p, n, qm, am, qp, ap, nf);
#pragma unroll
      for (int s = 0; s < Ns; ++s) buf[s * NFP + f] = fMJ * nf[s];
    }
  }
  ATM_DEV static void gather(const DevDG& p, const int idx[3], const double* buf, double* acc, double MJIn) {
#pragma unroll
    for (int r = 0; r < D; ++r) {
      if (!((p.dirmask >> r) & 1)) continue;
      const int ir = idx[r];
      const int side = (ir == 0) ? 0 : ((ir == NQ - 1) ? 1 : -1);
      if (side < 0) continue;
      int fp = 0;
      if (D == 2) fp = (r == 0) ? idx[1] : idx[0];
      if (D == 3) fp = (r == 0) ? idx[1] + NQ * idx[2] : ((r == 1) ? idx[0] + NQ * idx[2] : idx[0] + NQ * idx[1]);
      const int f = (2 * r + side) * NFPF + fp;
#pragma unroll
      for (int s = 0; s < Ns; ++s) acc[s] -= buf[s * NFP + f] * MJIn;
    }
  }
};

template <class L, int NQ, int SF>
ATM_DEV void fd_fast_body(const DevDG& p, double* lds, int64_t blk, const double* __restrict__ q, double* __restrict__ dq, double alpha,
               int assign, double* __restrict__ qout, double bdt, int epw) {
#pragma clang fp contract(fast)
  constexpr bool GRAV = L::LAW == ATM_LAW_EULER_GRAVITY;
  using FE = FastEC<L::D, GRAV>;
  constexpr int D = L::D, Ns = L::Ns, NA = L::NA, NP = FE::NP;
  constexpr int Nq = NQ;
  constexpr int Np = D == 1 ? NQ : (D == 2 ? NQ * NQ : NQ * NQ * NQ);
  constexpr int M = Nq / 2;  // pair slots per node and direction
  double* sD = lds;
  // The lean form (3-D elements of N >= 4): the LDS primitives without u.u (recomputed from
  // u where a partner's are read), each later direction's metric row loaded at that
  // direction, and the stage's dq, MJI and the state for the update loaded at the end, so
  // that no register holds them through the volume term.  It brings the 3-D N = 4
  // Euler-gravity kernel from 150 to 128 VGPRs and 20 KB of LDS per element: 4 waves per
  // SIMD instead of 3.  For smaller elements the workgroup's life is too short to hide the
  // loads at the end (measured slower), and they keep every load at the start.
  constexpr bool LEAN = fast_lean<L, NQ>();
  constexpr int NPL = LEAN ? NP - 1 : NP;  // LDS planes of primitives
  constexpr int stride = (NPL + D) * Np + fast_xbuf<L, NQ>();
  static_assert(NPL + D >= Ns + 1, "the surface term stages the state and geopotential in the primitives' planes");
  const int tid = threadIdx.x;
  const int el = tid / Np;
  const int ijk = tid - el * Np;
  const int64_t le = blk * epw + el;  // position in the processing order
  const bool active = el < epw && le < p.K;
  const int64_t e = (p.sched && active) ? (int64_t)p.sched[le] : le;
  double* sP = lds + Nq * Nq + el * stride;  // [k][node]
  double* slg = sP + NPL * Np;               // [d][node]
  double* sx = slg + D * Np;                 // [m][s][node]
  for (int t = tid; t < Nq * Nq; t += blockDim.x) sD[t] = p.D[t];
  double q1[Ns], a1[NA], P1[NP];
  const int64_t node = e * Np + ijk;
  // Every load of the node's data first and unconditionally (a lane past the last element
  // reads node 0 and discards it): a load on a conditional path would make the compiler
  // wait for it before the next ones, so state, aux, MJ, MJI, the stage's dq and the metric
  // rows of every direction are in flight together (one memory latency); the primitives
  // follow.
  const int64_t nl = active ? node : 0;
#pragma unroll
  for (int s = 0; s < Ns; ++s) q1[s] = q[s * p.Ntot + nl];
  load_aux<L>(p, nl, a1);
  const double MJn = p.MJ[nl];
  double MJIn = 0.0, dq0[Ns];
#pragma unroll
  for (int s = 0; s < Ns; ++s) dq0[s] = 0.0;
  double gall[LEAN ? 1 : D][D];  // metric rows: the first direction's only in the lean form
  if constexpr (!LEAN) {
    MJIn = p.MJI[nl];
    if (!assign) {  // uniform over the launch
#pragma unroll
      for (int s = 0; s < Ns; ++s) dq0[s] = dq[s * p.Ntot + nl];
    }
  }
#pragma unroll
  for (int r = 0; r < (LEAN ? 1 : D); ++r)
#pragma unroll
    for (int d = 0; d < D; ++d) gall[r][d] = p.g[(int64_t)(r + D * d) * p.Ntot + nl];
  FE::prim(p.c, q1, a1, P1);
  if (active) {
#pragma unroll
    for (int k = 0; k < NP; ++k)
      if (!LEAN || k != D + 2) sP[(!LEAN || k < D + 2 ? k : k - 1) * Np + ijk] = P1[k];
  }
  int idx[3];
  node_index<D>(ijk, Nq, idx);
  // the last exchange slot of even Nq (pair distance M): a node of the upper half of its line
  // in direction r, at line position c >= M, by its position in that half
  constexpr bool HALF = Nq % 2 == 0;
  auto half_pos = [&](int r, int c) {
    int pos = 0, st = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      pos += (d == r ? c - M : idx[d]) * st;
      st *= d == r ? M : Nq;
    }
    return pos;
  };
  const double inv_gm1 = 1 / (p.c.gamma - 1);
  double vol[Ns];
#pragma unroll
  for (int s = 0; s < Ns; ++s) vol[s] = 0.0;
  double src[Ns];
#pragma unroll
  for (int s = 0; s < Ns; ++s) src[s] = 0.0;
  if (LEAN && active) {  // every direction's sources first, in direction order
#pragma unroll
    for (int r = 0; r < D; ++r)
      if ((p.dirmask >> r) & 1) L::sources(p.c, src, q1, a1, 1 << r, false);
  }
#pragma unroll
  for (int r = 0; r < D; ++r) {
    if (!((p.dirmask >> r) & 1)) continue;
    // No barrier before overwriting slg: every thread passed the previous
    // direction's post-pair barrier, so all reads of the old slg are done; the
    // slots are rewritten only after the next barrier, which every thread
    // reaches after gathering its previous slots.
    double gr[D];
#pragma unroll
    for (int d = 0; d < D; ++d) gr[d] = (!LEAN || r == 0) ? gall[LEAN ? 0 : r][d] : p.g[(int64_t)(r + D * d) * p.Ntot + nl];
    if (active) {
#pragma unroll
      for (int d = 0; d < D; ++d) slg[d * Np + ijk] = MJn * gr[d];
    }
    __syncthreads();
    const int a = idx[r];
    const int lstride = (r == 0) ? 1 : ((r == 1) ? Nq : Nq * Nq);
    const int lbase = ijk - a * lstride;
    if (active) {
      if constexpr (!LEAN) L::sources(p.c, src, q1, a1, 1 << r, false);
      double lga[D];
#pragma unroll
      for (int d = 0; d < D; ++d) lga[d] = slg[d * Np + ijk];
#pragma unroll
      for (int m = 1; m <= M; ++m) {
        if ((Nq % 2 == 0) && m == M && a >= M) continue;
        const int b = (a + m) % Nq;
        const int ib = lbase + b * lstride;
        double P2[NP];
#pragma unroll
        for (int k = 0; k < NP; ++k)
          if (!LEAN || k != D + 2) P2[k] = sP[(!LEAN || k < D + 2 ? k : k - 1) * Np + ib];
        if constexpr (LEAN) {  // the partner's u.u, the same operations as in prim()
          double uu = P2[1] * P2[1];
#pragma unroll
          for (int d = 1; d < D; ++d) uu += P2[1 + d] * P2[1 + d];
          P2[D + 2] = uu;
        }
        typename FE::Pair w;
        FE::pair(p.c, inv_gm1, P1, P2, w);
        const double Dba = sD[b + Nq * a], Dab = sD[a + Nq * b];
        double G[D];
#pragma unroll
        for (int d = 0; d < D; ++d) G[d] = Dba * slg[d * Np + ib] - Dab * lga[d];
        double gf[Ns];
        FE::contract(w, G, gf);
#pragma unroll
        for (int s = 0; s < Ns; ++s) vol[s] += gf[s];
#pragma unroll
        for (int d = 0; d < D; ++d) vol[1 + d] -= G[d] * w.fla;
        // the partner's share: -G.F, plus its own gravity fluctuation term
        const bool hs = HALF && m == M;
        double* slot = sx + (m - 1) * Ns * Np + (hs ? half_pos(r, b) : ib);
        const int xs = hs ? Np / 2 : Np;
        slot[0] = -gf[0];
#pragma unroll
        for (int d = 0; d < D; ++d) slot[(1 + d) * xs] = G[d] * w.flb - gf[1 + d];
        slot[(D + 1) * xs] = -gf[D + 1];
      }
    }
    __syncthreads();
    if (active) {
#pragma unroll
      for (int m = 1; m <= M; ++m) {
        if ((Nq % 2 == 0) && m == M && a < M) continue;
        const bool hs = HALF && m == M;
        const double* slot = sx + (m - 1) * Ns * Np + (hs ? half_pos(r, a) : ijk);
        const int xs = hs ? Np / 2 : Np;
#pragma unroll
        for (int s = 0; s < Ns; ++s) vol[s] += slot[s * xs];
      }
    }
  }
  // Surface term.  No thread reads a partner's primitives after the last direction's barrier,
  // so the element's state and geopotential are staged by node in the primitives' LDS planes
  // (the lean form reloads the state with one coalesced load); the face points read their own
  // side there instead of with strided global loads, and the update reads the state back.
  using FF = FastFaces<L, SF, NQ>;
  FF ff;
  ff.load(p, active ? e : 0, ijk);
  double* sQ = sP;
  if (active) {
#pragma unroll
    for (int s = 0; s < Ns; ++s) sQ[s * Np + ijk] = LEAN ? q[s * p.Ntot + node] : q1[s];
    if constexpr (GRAV) sQ[Ns * Np + ijk] = P1[D + 3];  // = aux[0]
  }
  __syncthreads();  // every thread has consumed its exchange slots
  if (active) ff.run(p, e, ijk, q, sx, sQ);
  __syncthreads();
  if (!active) return;
  if constexpr (LEAN) {
#pragma unroll
    for (int s = 0; s < Ns; ++s) dq0[s] = assign ? 0.0 : dq[s * p.Ntot + node];
    MJIn = p.MJI[node];
#pragma unroll
    for (int s = 0; s < Ns; ++s) q1[s] = sQ[s * Np + ijk];
  }
  double acc[Ns];
#pragma unroll
  for (int s = 0; s < Ns; ++s) {
    const double v = src[s] + MJIn * vol[s];
    acc[s] = assign ? v : alpha * dq0[s] + v;
  }
  FF::gather(p, idx, sx, acc, MJIn);
  epilogue<L>(p, node, q1, acc, dq, qout, bdt);
}

// Largest workgroup of the kernel.  An element of 5 to 7 waves (343 nodes at
// N = 6) leaves a CU with one such workgroup resident at 3 waves per SIMD;
// two elements in one workgroup of up to 12 waves keep almost twice the waves
// in flight (N = 6 bubble: 74 -> 58 us per stage).  Elsewhere 512 threads (the
// compiler then budgets registers for 2 waves per SIMD).  The variant with the
// runtime surface-flux switch (SF < 0) needs more registers than 3 waves per
// SIMD leave and keeps 512.
template <class L, int NQ, int SF>
constexpr int fast_max_threads() {
  constexpr int Np = L::D == 1 ? NQ : (L::D == 2 ? NQ * NQ : NQ * NQ * NQ);
  constexpr int waves = (Np + 63) / 64;
  return (SF >= 0 && waves > 4 && waves < 8) ? 768 : 512;
}

template <class L, int NQ, int SF>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, fast_max_threads<L, NQ, SF>())))
fd_fast_kernel(const DevDG p, const double* __restrict__ q, double* __restrict__ dq, double alpha, int assign,
               double* __restrict__ qout, double bdt, int epw) {
  extern __shared__ double lds[];
  fd_fast_body<L, NQ, SF>(p, lds, xcd_block(), q, dq, alpha, assign, qout, bdt, epw);
}

template <class L, int NQ, int SF>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, fast_max_threads<L, NQ, SF>())))
fd_fast_kernel_batch(const BatchEntry* __restrict__ tab, int n, int parity, double alpha, int assign, double b,
                     int epw) {
  extern __shared__ double lds[];
  const int64_t lb = xcd_block();
  const int i = batch_find(tab, n, lb);
  const DevDG p = tab[i].p;
  fd_fast_body<L, NQ, SF>(p, lds, lb - tab[i].block_start, tab[i].buf[parity], tab[i].dq, alpha, assign,
                          tab[i].buf[1 - parity], b * tab[i].dt, epw);
}

// One launch per stage for 2-D members of different degrees (N = 3..7, Nq =
// 4..8) with the same law and surface flux: each workgroup runs its member's
// Nq-specialised body with the member's own elements per workgroup and LDS
// share from the table (as weak_multi_kernel_batch).
template <class L, int SF>
__global__ void __launch_bounds__(512)
    fd_fast_multi_kernel_batch(const BatchEntry* __restrict__ tab, int n, int parity, double alpha, int assign,
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
    case 4: fd_fast_body<L, 4, SF>(p, lds, blk, qin, tab[i].dq, alpha, assign, qo, bdt, p.epw); break;
    case 5: fd_fast_body<L, 5, SF>(p, lds, blk, qin, tab[i].dq, alpha, assign, qo, bdt, p.epw); break;
    case 6: fd_fast_body<L, 6, SF>(p, lds, blk, qin, tab[i].dq, alpha, assign, qo, bdt, p.epw); break;
    case 7: fd_fast_body<L, 7, SF>(p, lds, blk, qin, tab[i].dq, alpha, assign, qo, bdt, p.epw); break;
    case 8: fd_fast_body<L, 8, SF>(p, lds, blk, qin, tab[i].dq, alpha, assign, qo, bdt, p.epw); break;
    default: break;
  }
}

// `epw` is unused: every member carries its own in the table.
template <class L, int SF>
hipError_t launch_fd_fast_multi_batch(const BatchEntry* tab, int n, int64_t blocks, int threads, size_t lds, int epw,
                                      hipStream_t st, int parity, double alpha, int assign, double b) {
  (void)epw;
  hipLaunchKernelGGL((fd_fast_multi_kernel_batch<L, SF>), dim3(blocks), dim3(threads), lds, st, tab, n, parity,
                     alpha, assign, b);
  return hipGetLastError();
}

template <class L, int NQ, int SF>
constexpr int fast_epw() {
  constexpr int Np = L::D == 1 ? NQ : (L::D == 2 ? NQ * NQ : NQ * NQ * NQ);
  if (fast_max_threads<L, NQ, SF>() > 512) return 2;
  return Np >= 64 ? 1 : 256 / Np;
}

// LDS bytes per element (primitives, metric weights, exchange slots); the
// workgroup adds NQ*NQ doubles for D
template <class L, int NQ>
constexpr size_t fast_lds_elem_bytes() {
  constexpr int D = L::D, NP = FastEC<L::D, true>::NP;
  constexpr int Np = D == 1 ? NQ : (D == 2 ? NQ * NQ : NQ * NQ * NQ);
  return sizeof(double) * ((size_t)(NP - (fast_lean<L, NQ>() ? 1 : 0) + D) * Np + fast_xbuf<L, NQ>());
}

template <class L, int NQ, int SF>
hipError_t launch_fd_fast(const DevDG& p, hipStream_t st, const double* q, double* dq, double alpha,
                          int assign, double* qout, double bdt) {
  constexpr int D = L::D;
  constexpr int Np = D == 1 ? NQ : (D == 2 ? NQ * NQ : NQ * NQ * NQ);
  const int epw = p.epw;
  const int threads = ((epw * Np + 63) / 64) * 64;
  if (epw < 1 || threads > fast_max_threads<L, NQ, SF>()) return hipErrorInvalidConfiguration;
  const int blocks = (p.K + epw - 1) / epw;
  if (blocks == 0) return hipSuccess;
  const size_t lds = p.lds;
  hipLaunchKernelGGL((fd_fast_kernel<L, NQ, SF>), dim3(blocks), dim3(threads), lds, st, p, q, dq, alpha,
                     assign, qout, bdt, epw);
  return hipGetLastError();
}

template <class L, int NQ, int SF>
hipError_t launch_fd_fast_batch(const BatchEntry* tab, int n, int64_t blocks, int threads, size_t lds, int epw,
                                hipStream_t st, int parity, double alpha, int assign, double b) {
  hipLaunchKernelGGL((fd_fast_kernel_batch<L, NQ, SF>), dim3(blocks), dim3(threads), lds, st, tab, n, parity,
                     alpha, assign, b, epw);
  return hipGetLastError();
}

}  // namespace atm
