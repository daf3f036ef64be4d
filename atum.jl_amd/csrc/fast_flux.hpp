// Surface fluxes of the FAST kernels for the Euler laws: the Roe flux
// (src/balancelaws/euler.jl:91-138, euler_gravity.jl:149-198: the minus-side
// geopotential in both pressures of the dissipation, each side's own aux in
// its physical flux) and the Rusanov flux (src/numericalfluxes.jl:36-42).
//
// The pressures are computed by the law's own pressure function, rounded as in
// the reference: on near-balanced states the pressure jump across a face is
// itself of round-off size, and the Roe dissipation carries it with the factor
// h/c, so a differently rounded pressure would change the result far beyond
// round-off of the flux.  Everything downstream is reformulated: the normal
// physical flux is formed directly (F(q).n = [rho u.n, rho u (u.n) + p n,
// (u.n)(rho e + p)]), reciprocals are shared, FMA contraction is allowed, and the
// reciprocals and square roots outside the pressures are the hardware estimates
// refined by Newton steps (about one ulp) instead of the IEEE sequences.
#pragma once
#include "laws.hpp"

namespace atm {

// 1/x for positive normal x: the hardware reciprocal refined by two Newton
// steps (within about one ulp), about a third of an IEEE division.
__device__ __forceinline__ double fast_rcp(double x) {
#pragma clang fp contract(fast)
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// sqrt(x) for positive normal x: the hardware reciprocal square root y refined by
// two coupled Newton (Goldschmidt) steps on g = x*y ~ sqrt(x) and h = y/2 ~
// 1/(2 sqrt(x)), then one residual correction g + h (x - g^2) (within about one ulp).
__device__ __forceinline__ double fast_sqrt(double x) {
#pragma clang fp contract(fast)
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
  }
  return fma(fma(-g, g, x), h, g);
}

template <class L, int SF>
constexpr bool has_fast_sflux() {
  return (L::LAW == ATM_LAW_EULER || L::LAW == ATM_LAW_EULER_GRAVITY) &&
         (SF == ATM_FLUX_ROE || SF == ATM_FLUX_RUSANOV);
}

template <int D, bool GRAV>
struct FastSide {
  double rho, ir, u[D], un, p, pf, rhoe;
};

// p: pressure of the dissipation (geopotential phi_diss); pf: pressure of the
// physical flux (own geopotential a[0])
template <int D, bool GRAV>
ATM_DEV void fast_side(const LawConst& c, const double* q, const double* a, double phi_diss, const double* n,
                       FastSide<D, GRAV>& S) {
  S.rho = q[0];
  S.rhoe = q[D + 1];
  if constexpr (GRAV) {
    S.p = EulerGravity<D>::pressure(c, S.rho, q + 1, S.rhoe, phi_diss);
    S.pf = (phi_diss == a[0]) ? S.p : EulerGravity<D>::pressure(c, S.rho, q + 1, S.rhoe, a[0]);
  } else {
    S.p = S.pf = Euler<D>::pressure(c, S.rho, q + 1, S.rhoe);
  }
  {
#pragma clang fp contract(fast)
    S.ir = fast_rcp(S.rho);
    S.un = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      S.u[d] = q[1 + d] * S.ir;
      S.un += S.u[d] * n[d];
    }
  }
}

template <int D, bool GRAV>
ATM_DEV void fast_normal_flux(const LawConst& c, const FastSide<D, GRAV>& S, const double* q, const double* a,
                              const double* n, double* fn) {
#pragma clang fp contract(fast)
  double dp = S.pf;
  if (GRAV && c.pde_level_balance) dp -= a[D + 2];
  fn[0] = S.rho * S.un;
#pragma unroll
  for (int e = 0; e < D; ++e) fn[1 + e] = q[1 + e] * S.un + dp * n[e];
  fn[D + 1] = S.un * (S.rhoe + S.pf);
}

template <int D, int Ns, bool GRAV, int SF>
ATM_DEV void fast_flux_combine(const LawConst& c, const double* n, const double* qm, const double* am,
                               const double* qp, const double* ap, const FastSide<D, GRAV>& M,
                               const FastSide<D, GRAV>& P, double phim, double* nf);

template <class L, int SF>
ATM_DEV void fast_sflux(const LawConst& c, const double* n, const double* qm, const double* am, const double* qp,
                        const double* ap, double* nf) {
  static_assert(has_fast_sflux<L, SF>(), "no fast form of this surface flux");
  constexpr int D = L::D, Ns = L::Ns;
  constexpr bool GRAV = L::LAW == ATM_LAW_EULER_GRAVITY;
  FastSide<D, GRAV> M, P;
  const double phim = GRAV ? am[0] : 0.0;
  fast_side<D, GRAV>(c, qm, am, phim, n, M);
  // Roe: minus-side geopotential in both dissipation pressures; Rusanov: each side's own
  fast_side<D, GRAV>(c, qp, ap, SF == ATM_FLUX_ROE ? phim : (GRAV ? ap[0] : 0.0), n, P);
  fast_flux_combine<D, Ns, GRAV, SF>(c, n, qm, am, qp, ap, M, P, phim, nf);
}

template <int D, int Ns, bool GRAV, int SF>
ATM_DEV void fast_flux_combine(const LawConst& c, const double* n, const double* qm, const double* am,
                               const double* qp, const double* ap, const FastSide<D, GRAV>& M,
                               const FastSide<D, GRAV>& P, double phim, double* nf) {
#pragma clang fp contract(fast)
  const double g = c.gamma;
  double fm[Ns], fpn[Ns];
  fast_normal_flux<D, GRAV>(c, M, qm, am, n, fm);
  fast_normal_flux<D, GRAV>(c, P, qp, ap, n, fpn);
  if constexpr (SF == ATM_FLUX_RUSANOV) {
    const double wm = fabs(M.un) + fast_sqrt(g * M.pf * M.ir);
    const double wp = fabs(P.un) + fast_sqrt(g * P.pf * P.ir);
    const double ws = 0.5 * fmax(wm, wp);
#pragma unroll
    for (int s = 0; s < Ns; ++s) nf[s] = 0.5 * (fm[s] + fpn[s]) - ws * (qp[s] - qm[s]);
  } else {
    const double hm = (M.rhoe + M.p) * M.ir, hp = (P.rhoe + P.p) * P.ir;
    const double cm = fast_sqrt(g * M.p * M.ir), cp = fast_sqrt(g * P.p * P.ir);
    const double srm = fast_sqrt(M.rho), srp = fast_sqrt(P.rho);
    const double iden = fast_rcp(srm + srp);
    const double rho = srm * srp;
    double u[D], du[D];
    double un = 0, dun = 0, uu = 0, udu = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      u[d] = (srm * M.u[d] + srp * P.u[d]) * iden;
      du[d] = P.u[d] - M.u[d];
      un += u[d] * n[d];
      dun += du[d] * n[d];
      uu += u[d] * u[d];
      udu += u[d] * du[d];
    }
    const double h = (srm * hm + srp * hp) * iden;
    const double cs = (srm * cm + srp * cp) * iden;
    const double drho = P.rho - M.rho, dp = P.p - M.p;
    const double cm2 = fast_rcp(cs * cs);
    const double w1 = 0.5 * fabs(un - cs) * (dp - rho * cs * dun) * cm2;
    const double w2 = 0.5 * fabs(un + cs) * (dp + rho * cs * dun) * cm2;
    const double w3 = fabs(un) * (drho - dp * cm2);
    const double w4 = fabs(un) * rho;
    const double e3 = GRAV ? w3 * (0.5 * uu + phim) : 0.5 * w3 * uu;
    nf[0] = 0.5 * (fm[0] + fpn[0]) - 0.5 * (w1 + w2 + w3);
#pragma unroll
    for (int d = 0; d < D; ++d)
      nf[1 + d] = 0.5 * (fm[1 + d] + fpn[1 + d]) -
                  0.5 * (w1 * (u[d] - cs * n[d]) + w2 * (u[d] + cs * n[d]) + w3 * u[d] + w4 * (du[d] - dun * n[d]));
    nf[D + 1] = 0.5 * (fm[D + 1] + fpn[D + 1]) -
                0.5 * (w1 * (h - cs * un) + w2 * (h + cs * un) + e3 + w4 * (udu - un * dun));
  }
}

}  // namespace atm
