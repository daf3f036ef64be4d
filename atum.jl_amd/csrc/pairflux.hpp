// Two-point volume fluxes evaluated from per-node primitives.
//
// Every quantity of a two-point flux that depends on one node only (velocity,
// pressure, beta = rho/2p, |u|^2, the physical flux for the central flux, ...)
// is computed once per node and staged in LDS; the pair evaluation then does
// only the genuinely two-point arithmetic.  Each primitive is computed with
// exactly the operations the reference's flux evaluates for it, so the pair
// result is bit-identical to evaluating the flux from the two states
// (src/balancelaws/euler.jl:140-196, euler_gravity.jl:200-269,
// shallow_water.jl:112-136, src/numericalfluxes.jl:31-34).
#pragma once
#include "laws.hpp"

namespace atm {

template <class L, int VF>
struct PairFlux;

// Central: primitive = physical flux F(q), pair = (F1 + F2) / 2
template <class L>
struct PairFlux<L, ATM_FLUX_CENTRAL> {
  static constexpr int D = L::D, Ns = L::Ns, NP = L::D * L::Ns;
  ATM_DEV static void prim(const LawConst& c, const double* q, const double* a, double* P) {
    double F[D][Ns];
    L::flux(c, q, a, F);
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
      for (int s = 0; s < Ns; ++s) P[d * Ns + s] = F[d][s];
  }
  ATM_DEV static void pair(const LawConst&, const double* P1, const double* P2, double F[D][Ns]) {
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
      for (int s = 0; s < Ns; ++s) F[d][s] = (P1[d * Ns + s] + P2[d * Ns + s]) / 2;
  }
};

// Euler-type EC / KG primitives: [rho, u (D), x, y, phi]
//   EC: x = b = rho/(2p), y = u.u ;  KG: x = e = rhoe/rho, y = p
template <int D, bool GRAV, int VF>
struct EulerPair {
  static constexpr int Ns = D + 2, NP = D + 4;
  ATM_DEV static void prim(const LawConst& c, const double* q, const double* a, double* P) {
    const double rho = q[0];
    double u[D];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = q[1 + d] / rho;
    const double phi = GRAV ? a[0] : 0.0;
    const double p = GRAV ? EulerGravity<D>::pressure(c, rho, q + 1, q[D + 1], phi)
                          : Euler<D>::pressure(c, rho, q + 1, q[D + 1]);
    P[0] = rho;
#pragma unroll
    for (int d = 0; d < D; ++d) P[1 + d] = u[d];
    if (VF == ATM_FLUX_EC) {
      P[D + 1] = rho / (2 * p);
      P[D + 2] = dotv<D>(u, u);
    } else {
      P[D + 1] = q[D + 1] / rho;
      P[D + 2] = p;
    }
    P[D + 3] = phi;
  }
  ATM_DEV static void pair(const LawConst& c, const double* P1, const double* P2, double F[D][Ns]) {
    const double rho1 = P1[0], rho2 = P2[0];
    const double* u1 = P1 + 1;
    const double* u2 = P2 + 1;
    const double phi1 = P1[D + 3], phi2 = P2[D + 3];
    const double rho_avg = avg(rho1, rho2);
    double u_avg[D];
#pragma unroll
    for (int d = 0; d < D; ++d) u_avg[d] = avg(u1[d], u2[d]);
    if (VF == ATM_FLUX_EC) {
      const double g = c.gamma;
      const double b1 = P1[D + 1], b2 = P2[D + 1];
      const double b_avg = avg(b1, b2);
      const double u2_avg = avg(P1[D + 2], P2[D + 2]);
      const double rho_log = logavg(rho1, rho2);
      const double b_log = logavg(b1, b2);
      double frho[D];
#pragma unroll
      for (int d = 0; d < D; ++d) frho[d] = u_avg[d] * rho_log;
      const double pd = rho_avg / (2 * b_avg);
      double cc = 1 / (2 * (g - 1) * b_log) - u2_avg / 2;
      if (GRAV) cc = cc + avg(phi1, phi2);
#pragma unroll
      for (int d = 0; d < D; ++d) {
        double mv = 0;
#pragma unroll
        for (int e = 0; e < D; ++e) {
          const double fu = (d == e) ? u_avg[d] * frho[e] + pd : u_avg[d] * frho[e];
          F[d][1 + e] = fu;
          mv = (e == 0) ? fu * u_avg[0] : mv + fu * u_avg[e];
        }
        F[d][0] = frho[d];
        F[d][D + 1] = cc * frho[d] + mv;
      }
      if (GRAV) {
        const double alpha = b_avg * rho_log / (2 * b1);
        const double fl = alpha * (phi1 - phi2);
#pragma unroll
        for (int d = 0; d < D; ++d) F[d][1 + d] -= fl;
      }
    } else {
      const double e_avg = avg(P1[D + 1], P2[D + 1]), p_avg = avg(P1[D + 2], P2[D + 2]);
      const double fl = GRAV ? (rho_avg / 2) * (phi1 - phi2) : 0.0;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        F[d][0] = u_avg[d] * rho_avg;
#pragma unroll
        for (int e = 0; e < D; ++e) {
          const double fre = u_avg[e] * rho_avg;
          double v = (d == e) ? u_avg[d] * fre + p_avg : u_avg[d] * fre;
          if (GRAV && d == e) v -= fl;
          F[d][1 + e] = v;
        }
        F[d][D + 1] = u_avg[d] * (rho_avg * e_avg + p_avg);
      }
    }
  }
};

template <int D>
struct PairFlux<Euler<D>, ATM_FLUX_EC> : EulerPair<D, false, ATM_FLUX_EC> {};
template <int D>
struct PairFlux<Euler<D>, ATM_FLUX_KG> : EulerPair<D, false, ATM_FLUX_KG> {};
template <int D>
struct PairFlux<EulerGravity<D>, ATM_FLUX_EC> : EulerPair<D, true, ATM_FLUX_EC> {};
template <int D>
struct PairFlux<EulerGravity<D>, ATM_FLUX_KG> : EulerPair<D, true, ATM_FLUX_KG> {};

// Shallow-water EC: [rho, u (D), rhou (D), theta, rho^2]
template <int D>
struct PairFlux<ShallowWater<D>, ATM_FLUX_EC> {
  static constexpr int Ns = D + 2, NP = 2 * D + 3;
  ATM_DEV static void prim(const LawConst&, const double* q, const double*, double* P) {
    const double rho = q[0];
    P[0] = rho;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      P[1 + d] = q[1 + d] / rho;
      P[1 + D + d] = q[1 + d];
    }
    P[2 * D + 1] = q[D + 1] / rho;
    P[2 * D + 2] = rho * rho;
  }
  ATM_DEV static void pair(const LawConst& c, const double* P1, const double* P2, double F[D][Ns]) {
    double u_avg[D], ru_avg[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      u_avg[d] = avg(P1[1 + d], P2[1 + d]);
      ru_avg[d] = avg(P1[1 + D + d], P2[1 + D + d]);
    }
    const double th_avg = avg(P1[2 * D + 1], P2[2 * D + 1]);
    const double rho_avg = avg(P1[0], P2[0]);
    const double rho2_avg = avg(P1[2 * D + 2], P2[2 * D + 2]);
    const double pd = c.grav * (rho_avg * rho_avg - rho2_avg / 2);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      F[d][0] = ru_avg[d];
#pragma unroll
      for (int e = 0; e < D; ++e) F[d][1 + e] = (d == e) ? ru_avg[d] * u_avg[e] + pd : ru_avg[d] * u_avg[e];
      F[d][D + 1] = ru_avg[d] * th_avg;
    }
  }
};

}  // namespace atm
