// Device-side balance laws and numerical fluxes for the gfx950 DGSEM kernels.
//
// The arithmetic keeps the reference's evaluation order (Julia evaluates
// `a * b * c` as `(a * b) * c`, `x / 2y` as `x / (2 * y)` and StaticArrays dot
// products as a left fold) so that, built without FMA contraction, a kernel
// that also keeps the reference's summation order reproduces it to round-off
// of the transcendental functions only.
//
//   avg / logavg / roe_avg ........ src/numericalfluxes.jl:8-23
//   Central / Rusanov / default ... src/numericalfluxes.jl:26-42
//   AdvectionLaw .................. src/balancelaws/advection.jl:8-19
//   EulerLaw ...................... src/balancelaws/euler.jl:9-256
//   EulerGravityLaw ............... src/balancelaws/euler_gravity.jl:10-332
//   LinearEulerGravityLaw ......... src/balancelaws/linear_euler_gravity.jl:3-65
//   ShallowWaterLaw ............... src/balancelaws/shallow_water.jl:9-136
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/atum_hip.h"

namespace atm {

#define ATM_DEV __device__ __forceinline__

struct LawConst {
  double gamma, grav, omega;
  double u[3];
  int pde_level_balance;
  int bc, source, n_bc_states;
  const double* bc_states;
};

template <int D>
ATM_DEV double dotv(const double* a, const double* b) {
  double s = a[0] * b[0];
#pragma unroll
  for (int d = 1; d < D; ++d) s += a[d] * b[d];
  return s;
}

ATM_DEV double avg(double a, double b) { return (a + b) / 2; }

ATM_DEV double logavg(double a, double b) {
  double zeta = a / b;
  double f = (zeta - 1) / (zeta + 1);
  double u = f * f;
  double F;
  if (u < 2.220446049250313e-16) {
    F = 1.0 + u * (1.0 / 3 + u * (1.0 / 5 + u * (1.0 / 7 + u * (1.0 / 9))));
  } else {
    F = log(zeta) / (2 * f);
  }
  return (a + b) / (2 * F);
}

ATM_DEV double roe_avg(double rm, double rp, double sm, double sp) {
  return (sqrt(rm) * sm + sqrt(rp) * sp) / (sqrt(rm) + sqrt(rp));
}

// Momentum-reflecting wall: q+ = (ρ, ρu - 2 (n·ρu) n, last)
template <int D, int Ns>
ATM_DEV void reflect_state(const double* n, const double* qm, double* qp) {
  double dn = dotv<D>(n, qm + 1);
  qp[0] = qm[0];
#pragma unroll
  for (int d = 0; d < D; ++d) qp[1 + d] = qm[1 + d] - 2 * dn * n[d];
  qp[Ns - 1] = qm[Ns - 1];
}

// ---------------------------------------------------------------------------
template <int D_>
struct Advection {
  static constexpr int D = D_, Ns = 1, NA = 1, LAW = ATM_LAW_ADVECTION;
  ATM_DEV static void flux(const LawConst& c, const double* q, const double*, double F[D][Ns]) {
#pragma unroll
    for (int d = 0; d < D; ++d) F[d][0] = c.u[d] * q[0];
  }
  ATM_DEV static double wavespeed(const LawConst& c, const double* n, const double*, const double*) {
    return fabs(dotv<D>(n, c.u));
  }
  // advection.jl:18-19
  static constexpr bool HAS_ENTROPY = true;
  ATM_DEV static double entropy(const LawConst&, const double* q, const double*) { return q[0] * q[0] / 2; }
  ATM_DEV static void entropyvariables(const LawConst&, const double* q, const double*, double* v) { v[0] = q[0]; }
  template <int VF>
  ATM_DEV static void twopoint(const LawConst& c, const double* q1, const double* a1,
                               const double* q2, const double* a2, double F[D][Ns]);
  template <int SF>
  ATM_DEV static void special_surface(const LawConst&, const double*, const double*, const double*,
                                      const double*, const double*, double*) {}
  ATM_DEV static void sources(const LawConst&, double*, const double*, const double*, int, bool) {}
  ATM_DEV static void reflect(const double*, const double* qm, double* qp) { qp[0] = qm[0]; }
  static constexpr bool has_vflux(int vf) { return vf == ATM_FLUX_CENTRAL; }
  static constexpr bool has_sflux(int sf) { return sf == ATM_FLUX_CENTRAL || sf == ATM_FLUX_RUSANOV; }
};

// ---------------------------------------------------------------------------
template <int D_>
struct Euler {
  static constexpr int D = D_, Ns = D_ + 2, NA = 1, LAW = ATM_LAW_EULER;
  ATM_DEV static double pressure(const LawConst& c, double rho, const double* rhou, double rhoe) {
    return (c.gamma - 1) * (rhoe - dotv<D>(rhou, rhou) / (2 * rho));
  }
  ATM_DEV static void flux(const LawConst& c, const double* q, const double*, double F[D][Ns]) {
    const double rho = q[0], rhoe = q[D + 1];
    double u[D];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = q[1 + d] / rho;
    const double p = pressure(c, rho, q + 1, rhoe);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      F[d][0] = q[1 + d];
#pragma unroll
      for (int e = 0; e < D; ++e) F[d][1 + e] = (d == e) ? q[1 + d] * u[e] + p : q[1 + d] * u[e];
      F[d][D + 1] = u[d] * (rhoe + p);
    }
  }
  ATM_DEV static double wavespeed(const LawConst& c, const double* n, const double* q, const double*) {
    const double rho = q[0];
    double u[D];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = q[1 + d] / rho;
    const double p = pressure(c, rho, q + 1, q[D + 1]);
    return fabs(dotv<D>(n, u)) + sqrt(c.gamma * p / rho);
  }
  // euler.jl:68-75
  static constexpr bool HAS_ENTROPY = true;
  ATM_DEV static double entropy(const LawConst& c, const double* q, const double*) {
    const double rho = q[0];
    const double p = pressure(c, rho, q + 1, q[D + 1]);
    const double s = log(p / pow(rho, c.gamma));
    return -rho * s / (c.gamma - 1);
  }
  ATM_DEV static void entropyvariables(const LawConst& c, const double* q, const double*, double* v) {
    const double rho = q[0], g = c.gamma;
    const double p = pressure(c, rho, q + 1, q[D + 1]);
    const double s = log(p / pow(rho, g));
    const double b = rho / (2 * p);
    double u[D];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = q[1 + d] / rho;
    v[0] = (g - s) / (g - 1) - dotv<D>(u, u) * b;
#pragma unroll
    for (int d = 0; d < D; ++d) v[1 + d] = (2 * b) * u[d];
    v[D + 1] = -2 * b;
  }
  template <int VF>
  ATM_DEV static void twopoint(const LawConst& c, const double* q1, const double* a1,
                               const double* q2, const double* a2, double F[D][Ns]);
  template <int SF>
  ATM_DEV static void special_surface(const LawConst& c, const double* n, const double* qm,
                                      const double* am, const double* qp, const double* ap,
                                      double* nf);
  ATM_DEV static void sources(const LawConst&, double*, const double*, const double*, int, bool) {}
  ATM_DEV static void reflect(const double* n, const double* qm, double* qp) {
    reflect_state<D, Ns>(n, qm, qp);
  }
  static constexpr bool has_vflux(int vf) {
    return vf == ATM_FLUX_CENTRAL || vf == ATM_FLUX_EC || vf == ATM_FLUX_KG;
  }
  static constexpr bool has_sflux(int) { return true; }
};

// ---------------------------------------------------------------------------
// aux = [Φ, ∇Φ (D), ρ_ref, p_ref]
template <int D_>
struct EulerGravity {
  static constexpr int D = D_, Ns = D_ + 2, NA = D_ + 3, LAW = ATM_LAW_EULER_GRAVITY;
  ATM_DEV static double pressure(const LawConst& c, double rho, const double* rhou, double rhoe,
                                 double phi) {
    return (c.gamma - 1) * (rhoe - dotv<D>(rhou, rhou) / (2 * rho) - rho * phi);
  }
  ATM_DEV static void flux(const LawConst& c, const double* q, const double* aux, double F[D][Ns]) {
    const double rho = q[0], rhoe = q[D + 1], phi = aux[0];
    double u[D];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = q[1 + d] / rho;
    const double p = pressure(c, rho, q + 1, rhoe, phi);
    double dp = p;
    if (c.pde_level_balance) dp -= aux[D + 2];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      F[d][0] = q[1 + d];
#pragma unroll
      for (int e = 0; e < D; ++e) F[d][1 + e] = (d == e) ? q[1 + d] * u[e] + dp : q[1 + d] * u[e];
      F[d][D + 1] = u[d] * (rhoe + p);
    }
  }
  ATM_DEV static double wavespeed(const LawConst& c, const double* n, const double* q, const double* aux) {
    const double rho = q[0];
    double u[D];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = q[1 + d] / rho;
    const double p = pressure(c, rho, q + 1, q[D + 1], aux[0]);
    return fabs(dotv<D>(n, u)) + sqrt(c.gamma * p / rho);
  }
  // euler_gravity.jl:124-132
  static constexpr bool HAS_ENTROPY = true;
  ATM_DEV static double entropy(const LawConst& c, const double* q, const double* aux) {
    const double rho = q[0];
    const double p = pressure(c, rho, q + 1, q[D + 1], aux[0]);
    const double s = log(p / pow(rho, c.gamma));
    return -rho * s / (c.gamma - 1);
  }
  ATM_DEV static void entropyvariables(const LawConst& c, const double* q, const double* aux, double* v) {
    const double rho = q[0], g = c.gamma, phi = aux[0];
    const double p = pressure(c, rho, q + 1, q[D + 1], phi);
    const double s = log(p / pow(rho, g));
    const double b = rho / (2 * p);
    double u[D];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = q[1 + d] / rho;
    v[0] = (g - s) / (g - 1) - (dotv<D>(u, u) - 2 * phi) * b;
#pragma unroll
    for (int d = 0; d < D; ++d) v[1 + d] = (2 * b) * u[d];
    v[D + 1] = -2 * b;
  }
  template <int VF>
  ATM_DEV static void twopoint(const LawConst& c, const double* q1, const double* a1,
                               const double* q2, const double* a2, double F[D][Ns]);
  template <int SF>
  ATM_DEV static void special_surface(const LawConst& c, const double* n, const double* qm,
                                      const double* am, const double* qp, const double* ap,
                                      double* nf);
  // nonconservative_term! (:100-113) when `nonconservative`, problem Coriolis source
  // (experiment/euler_gravity/baroclinicwave.jl:27-34)
  ATM_DEV static void sources(const LawConst& c, double* dq, const double* q, const double* aux,
                              int dirmask, bool nonconservative) {
    const bool vert = dirmask & (1 << (D - 1));
    if (c.source == ATM_SOURCE_CORIOLIS && D == 3 && vert) {
      const double a3 = 2 * c.omega;
      const double c0 = 0.0 * q[3] - a3 * q[2];
      const double c1 = a3 * q[1] - 0.0 * q[3];
      const double c2 = 0.0 * q[2] - 0.0 * q[1];
      dq[1] -= c0;
      dq[2] -= c1;
      dq[D] -= c2;
    }
    if (nonconservative && vert) {
      double rho = q[0];
      if (c.pde_level_balance) rho -= aux[D + 1];
#pragma unroll
      for (int d = 0; d < D; ++d) dq[1 + d] -= rho * aux[1 + d];
    }
  }
  ATM_DEV static void reflect(const double* n, const double* qm, double* qp) {
    reflect_state<D, Ns>(n, qm, qp);
  }
  static constexpr bool has_vflux(int vf) {
    return vf == ATM_FLUX_CENTRAL || vf == ATM_FLUX_EC || vf == ATM_FLUX_KG;
  }
  static constexpr bool has_sflux(int) { return true; }
};

// ---------------------------------------------------------------------------
template <int D_>
struct LinearEulerGravity {
  static constexpr int D = D_, Ns = D_ + 2, NA = D_ + 3, LAW = ATM_LAW_LINEAR_EULER_GRAVITY;
  ATM_DEV static void flux(const LawConst& c, const double* q, const double* aux, double F[D][Ns]) {
    const double rho = q[0], rhoe = q[D + 1], phi = aux[0];
    const double p = (c.gamma - 1) * (rhoe - rho * phi);
    const double ref_rho = aux[D + 1], ref_p = aux[D + 2];
    const double ref_rhoe = ref_p / (c.gamma - 1) + ref_rho * phi;
    const double h = (ref_rhoe + ref_p) / ref_rho;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      F[d][0] = q[1 + d];
#pragma unroll
      for (int e = 0; e < D; ++e) F[d][1 + e] = (d == e) ? p : 0.0;
      F[d][D + 1] = h * q[1 + d];
    }
  }
  ATM_DEV static double wavespeed(const LawConst& c, const double*, const double*, const double* aux) {
    return sqrt(c.gamma * aux[D + 2] / aux[D + 1]);
  }
  // the linearized law defines no entropy (linear_euler_gravity.jl)
  static constexpr bool HAS_ENTROPY = false;
  ATM_DEV static double entropy(const LawConst&, const double*, const double*) { return 0.0; }
  ATM_DEV static void entropyvariables(const LawConst&, const double*, const double*, double* v) {
#pragma unroll
    for (int s = 0; s < Ns; ++s) v[s] = 0.0;
  }
  template <int VF>
  ATM_DEV static void twopoint(const LawConst& c, const double* q1, const double* a1,
                               const double* q2, const double* a2, double F[D][Ns]);
  template <int SF>
  ATM_DEV static void special_surface(const LawConst&, const double*, const double*, const double*,
                                      const double*, const double*, double*) {}
  // source! (:58-65) applies in every volume form
  ATM_DEV static void sources(const LawConst&, double* dq, const double* q, const double* aux,
                              int dirmask, bool) {
    if (dirmask & (1 << (D - 1))) {
      const double rho = q[0];
#pragma unroll
      for (int d = 0; d < D; ++d) dq[1 + d] -= rho * aux[1 + d];
    }
  }
  ATM_DEV static void reflect(const double* n, const double* qm, double* qp) {
    reflect_state<D, Ns>(n, qm, qp);
  }
  static constexpr bool has_vflux(int vf) { return vf == ATM_FLUX_CENTRAL; }
  static constexpr bool has_sflux(int sf) { return sf == ATM_FLUX_CENTRAL || sf == ATM_FLUX_RUSANOV; }
};

// ---------------------------------------------------------------------------
template <int D_>
struct ShallowWater {
  static constexpr int D = D_, Ns = D_ + 2, NA = 1, LAW = ATM_LAW_SHALLOW_WATER;
  ATM_DEV static void flux(const LawConst& c, const double* q, const double*, double F[D][Ns]) {
    const double rho = q[0], rhot = q[D + 1];
    double u[D];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = q[1 + d] / rho;
    const double p = c.grav * (rho * rho) / 2;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      F[d][0] = q[1 + d];
#pragma unroll
      for (int e = 0; e < D; ++e) F[d][1 + e] = (d == e) ? q[1 + d] * u[e] + p : q[1 + d] * u[e];
      F[d][D + 1] = u[d] * rhot;
    }
  }
  ATM_DEV static double wavespeed(const LawConst& c, const double* n, const double* q, const double*) {
    const double rho = q[0];
    double u[D];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = q[1 + d] / rho;
    return fabs(dotv<D>(n, u)) + sqrt(c.grav * rho);
  }
  // shallow_water.jl:52-66
  static constexpr bool HAS_ENTROPY = true;
  ATM_DEV static double entropy(const LawConst& c, const double* q, const double*) {
    const double rho = q[0];
    return dotv<D>(q + 1, q + 1) / (2 * rho) + c.grav * (rho * rho) / 2;
  }
  ATM_DEV static void entropyvariables(const LawConst& c, const double* q, const double*, double* v) {
    const double rho = q[0];
    double mr[D];
#pragma unroll
    for (int d = 0; d < D; ++d) mr[d] = -q[1 + d];
    v[0] = dotv<D>(mr, q + 1) / (2 * (rho * rho)) + c.grav * rho;
#pragma unroll
    for (int d = 0; d < D; ++d) v[1 + d] = q[1 + d] / rho;
    v[D + 1] = 0.0;
  }
  template <int VF>
  ATM_DEV static void twopoint(const LawConst& c, const double* q1, const double* a1,
                               const double* q2, const double* a2, double F[D][Ns]);
  template <int SF>
  ATM_DEV static void special_surface(const LawConst& c, const double* n, const double* qm,
                                      const double* am, const double* qp, const double* ap,
                                      double* nf);
  ATM_DEV static void sources(const LawConst&, double*, const double*, const double*, int, bool) {}
  ATM_DEV static void reflect(const double* n, const double* qm, double* qp) {
    reflect_state<D, Ns>(n, qm, qp);
  }
  static constexpr bool has_vflux(int vf) { return vf == ATM_FLUX_CENTRAL || vf == ATM_FLUX_EC; }
  static constexpr bool has_sflux(int sf) { return sf != ATM_FLUX_KG && sf != ATM_FLUX_MATRIX; }
};

// ---------------------------------------------------------------------------
// Central two-point flux (src/numericalfluxes.jl:31-34)
template <class L>
ATM_DEV void central(const LawConst& c, const double* q1, const double* a1, const double* q2,
                     const double* a2, double F[L::D][L::Ns]) {
  double F2[L::D][L::Ns];
  L::flux(c, q1, a1, F);
  L::flux(c, q2, a2, F2);
#pragma unroll
  for (int d = 0; d < L::D; ++d)
#pragma unroll
    for (int s = 0; s < L::Ns; ++s) F[d][s] = (F[d][s] + F2[d][s]) / 2;
}

template <int D>
template <int VF>
ATM_DEV void Advection<D>::twopoint(const LawConst& c, const double* q1, const double* a1,
                                    const double* q2, const double* a2, double F[D][Ns]) {
  central<Advection<D>>(c, q1, a1, q2, a2, F);
}

template <int D>
template <int VF>
ATM_DEV void LinearEulerGravity<D>::twopoint(const LawConst& c, const double* q1, const double* a1,
                                             const double* q2, const double* a2, double F[D][Ns]) {
  central<LinearEulerGravity<D>>(c, q1, a1, q2, a2, F);
}

// Ismail–Roe / Chandrashekar entropy-conservative and Kennedy–Gruber fluxes for
// the Euler laws; with gravity the fluctuation term -α(Φ1-Φ2)I is added
// (euler.jl:140-196, euler_gravity.jl:200-269).
template <int D, bool GRAV, int VF>
ATM_DEV void euler_twopoint(const LawConst& c, const double* q1, const double* a1,
                            const double* q2, const double* a2, double F[D][D + 2]) {
  const double g = c.gamma;
  const double rho1 = q1[0], rhoe1 = q1[D + 1], rho2 = q2[0], rhoe2 = q2[D + 1];
  const double phi1 = GRAV ? a1[0] : 0.0, phi2 = GRAV ? a2[0] : 0.0;
  double u1[D], u2[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    u1[d] = q1[1 + d] / rho1;
    u2[d] = q2[1 + d] / rho2;
  }
  double p1, p2;
  if (GRAV) {
    p1 = EulerGravity<D>::pressure(c, rho1, q1 + 1, rhoe1, phi1);
    p2 = EulerGravity<D>::pressure(c, rho2, q2 + 1, rhoe2, phi2);
  } else {
    p1 = Euler<D>::pressure(c, rho1, q1 + 1, rhoe1);
    p2 = Euler<D>::pressure(c, rho2, q2 + 1, rhoe2);
  }
  if (VF == ATM_FLUX_EC) {
    const double b1 = rho1 / (2 * p1), b2 = rho2 / (2 * p2);
    const double rho_avg = avg(rho1, rho2);
    double u_avg[D];
#pragma unroll
    for (int d = 0; d < D; ++d) u_avg[d] = avg(u1[d], u2[d]);
    const double b_avg = avg(b1, b2);
    const double u2_avg = avg(dotv<D>(u1, u1), dotv<D>(u2, u2));
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
  } else {  // Kennedy–Gruber
    const double e1 = rhoe1 / rho1, e2 = rhoe2 / rho2;
    const double rho_avg = avg(rho1, rho2);
    double u_avg[D];
#pragma unroll
    for (int d = 0; d < D; ++d) u_avg[d] = avg(u1[d], u2[d]);
    const double e_avg = avg(e1, e2), p_avg = avg(p1, p2);
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

template <int D>
template <int VF>
ATM_DEV void Euler<D>::twopoint(const LawConst& c, const double* q1, const double* a1,
                                const double* q2, const double* a2, double F[D][Ns]) {
  if constexpr (VF == ATM_FLUX_CENTRAL) central<Euler<D>>(c, q1, a1, q2, a2, F);
  else euler_twopoint<D, false, VF>(c, q1, a1, q2, a2, F);
}

template <int D>
template <int VF>
ATM_DEV void EulerGravity<D>::twopoint(const LawConst& c, const double* q1, const double* a1,
                                       const double* q2, const double* a2, double F[D][Ns]) {
  if constexpr (VF == ATM_FLUX_CENTRAL) central<EulerGravity<D>>(c, q1, a1, q2, a2, F);
  else euler_twopoint<D, true, VF>(c, q1, a1, q2, a2, F);
}

// shallow_water.jl:112-136
template <int D>
template <int VF>
ATM_DEV void ShallowWater<D>::twopoint(const LawConst& c, const double* q1, const double* a1,
                                       const double* q2, const double* a2, double F[D][Ns]) {
  if constexpr (VF == ATM_FLUX_CENTRAL) {
    central<ShallowWater<D>>(c, q1, a1, q2, a2, F);
  } else {
    const double rho1 = q1[0], rho2 = q2[0];
    double u_avg[D], ru_avg[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      u_avg[d] = avg(q1[1 + d] / rho1, q2[1 + d] / rho2);
      ru_avg[d] = avg(q1[1 + d], q2[1 + d]);
    }
    const double th_avg = avg(q1[D + 1] / rho1, q2[D + 1] / rho2);
    const double rho_avg = avg(rho1, rho2);
    const double rho2_avg = avg(rho1 * rho1, rho2 * rho2);
    const double pd = c.grav * (rho_avg * rho_avg - rho2_avg / 2);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      F[d][0] = ru_avg[d];
#pragma unroll
      for (int e = 0; e < D; ++e)
        F[d][1 + e] = (d == e) ? ru_avg[d] * u_avg[e] + pd : ru_avg[d] * u_avg[e];
      F[d][D + 1] = ru_avg[d] * th_avg;
    }
  }
}

// Roe flux for the Euler laws (euler.jl:91-138; euler_gravity.jl:149-198, which
// uses the minus-side geopotential for both pressures).
template <int D, bool GRAV>
ATM_DEV void euler_roe(const LawConst& c, const double* n, const double* qm, const double* am,
                       const double* qp, const double* ap, double* nf) {
  constexpr int Ns = D + 2;
  const double phi = GRAV ? am[0] : 0.0;
  double Fm[D][Ns], Fp[D][Ns];
  if (GRAV) {
    EulerGravity<D>::flux(c, qm, am, Fm);
    EulerGravity<D>::flux(c, qp, ap, Fp);
  } else {
    Euler<D>::flux(c, qm, am, Fm);
    Euler<D>::flux(c, qp, ap, Fp);
  }
  const double rm = qm[0], rp = qp[0];
  double um[D], up[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    um[d] = qm[1 + d] / rm;
    up[d] = qp[1 + d] / rp;
  }
  const double em = qm[D + 1] / rm, ep = qp[D + 1] / rp;
  double pm, pp;
  if (GRAV) {
    pm = EulerGravity<D>::pressure(c, rm, qm + 1, qm[D + 1], phi);
    pp = EulerGravity<D>::pressure(c, rp, qp + 1, qp[D + 1], phi);
  } else {
    pm = Euler<D>::pressure(c, rm, qm + 1, qm[D + 1]);
    pp = Euler<D>::pressure(c, rp, qp + 1, qp[D + 1]);
  }
  const double hm = em + pm / rm, hp = ep + pp / rp;
  const double cm = sqrt(c.gamma * pm / rm), cp = sqrt(c.gamma * pp / rp);
  const double rho = sqrt(rm * rp);
  double u[D];
#pragma unroll
  for (int d = 0; d < D; ++d) u[d] = roe_avg(rm, rp, um[d], up[d]);
  const double h = roe_avg(rm, rp, hm, hp);
  const double cs = roe_avg(rm, rp, cm, cp);
  const double un = dotv<D>(u, n);
  const double drho = rp - rm, dp = pp - pm;
  double du[D];
#pragma unroll
  for (int d = 0; d < D; ++d) du[d] = up[d] - um[d];
  const double dun = dotv<D>(du, n);
  const double cm2 = 1 / (cs * cs);
  const double w1 = fabs(un - cs) * (dp - rho * cs * dun) * cm2 / 2;
  const double w2 = fabs(un + cs) * (dp + rho * cs * dun) * cm2 / 2;
  const double w3 = fabs(un) * (drho - dp * cm2);
  const double w4 = fabs(un) * rho;
  double fp[Ns];
  fp[0] = (w1 + w2 + w3) / 2;
#pragma unroll
  for (int d = 0; d < D; ++d)
    fp[1 + d] = (w1 * (u[d] - cs * n[d]) + w2 * (u[d] + cs * n[d]) + w3 * u[d] +
                 w4 * (du[d] - dun * n[d])) / 2;
  double e3;
  if (GRAV) {
    e3 = w3 * (dotv<D>(u, u) / 2 + phi);
  } else {
    double w3u[D];
#pragma unroll
    for (int d = 0; d < D; ++d) w3u[d] = w3 * u[d];
    e3 = dotv<D>(w3u, u) / 2;
  }
  fp[D + 1] = (w1 * (h - cs * un) + w2 * (h + cs * un) + e3 + w4 * (dotv<D>(u, du) - un * dun)) / 2;
#pragma unroll
  for (int s = 0; s < Ns; ++s) {
    double a = (Fm[0][s] + Fp[0][s]) * n[0];
#pragma unroll
    for (int d = 1; d < D; ++d) a += (Fm[d][s] + Fp[d][s]) * n[d];
    nf[s] = a / 2 - fp[s];
  }
}

// Matrix (entropy-stable) flux (euler.jl:198-256; euler_gravity.jl:271-332)
template <int D, bool GRAV>
ATM_DEV void euler_matrix(const LawConst& c, const double* n, const double* qm, const double* am,
                          const double* qp, const double* ap, double* nf) {
  constexpr int Ns = D + 2;
  const double g = c.gamma;
  double Fe[D][Ns];
  euler_twopoint<D, GRAV, ATM_FLUX_EC>(c, qm, am, qp, ap, Fe);
  double ec[Ns];
#pragma unroll
  for (int s = 0; s < Ns; ++s) {
    double a = Fe[0][s] * n[0];
#pragma unroll
    for (int d = 1; d < D; ++d) a += Fe[d][s] * n[d];
    ec[s] = a;
  }
  const double rm = qm[0], rp = qp[0];
  const double phim = GRAV ? am[0] : 0.0, phip = GRAV ? ap[0] : 0.0;
  double um[D], up[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    um[d] = qm[1 + d] / rm;
    up[d] = qp[1 + d] / rp;
  }
  double pm, pp;
  double vm[Ns], vp[Ns];
  if (GRAV) {
    pm = EulerGravity<D>::pressure(c, rm, qm + 1, qm[D + 1], phim);
    pp = EulerGravity<D>::pressure(c, rp, qp + 1, qp[D + 1], phip);
    EulerGravity<D>::entropyvariables(c, qm, am, vm);
    EulerGravity<D>::entropyvariables(c, qp, ap, vp);
  } else {
    pm = Euler<D>::pressure(c, rm, qm + 1, qm[D + 1]);
    pp = Euler<D>::pressure(c, rp, qp + 1, qp[D + 1]);
    Euler<D>::entropyvariables(c, qm, am, vm);
    Euler<D>::entropyvariables(c, qp, ap, vp);
  }
  const double bm = rm / (2 * pm), bp = rp / (2 * pp);
  const double phi_avg = avg(phim, phip);
  const double rho_log = logavg(rm, rp), b_log = logavg(bm, bp);
  double u_avg[D];
#pragma unroll
  for (int d = 0; d < D; ++d) u_avg[d] = avg(um[d], up[d]);
  const double p_avg = avg(rm, rp) / (2 * avg(bm, bp));
  const double u2_bar = 2 * sqrt(dotv<D>(u_avg, u_avg)) - avg(sqrt(dotv<D>(um, um)), sqrt(dotv<D>(up, up)));
  double h_bar = g / (2 * b_log * (g - 1)) + u2_bar / 2;
  if (GRAV) h_bar = h_bar + phi_avg;
  const double c_bar = sqrt(g * p_avg / rho_log);
  double umc[D], upc[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    umc[d] = u_avg[d] - c_bar * n[d];
    upc[d] = u_avg[d] + c_bar * n[d];
  }
  const double uavgn = dotv<D>(u_avg, n);
  double dv[Ns];
#pragma unroll
  for (int s = 0; s < Ns; ++s) dv[s] = vp[s] - vm[s];
  const double l1 = fabs(uavgn - c_bar) * rho_log / (2 * g);
  const double l2 = fabs(uavgn) * rho_log * (g - 1) / g;
  const double l3 = fabs(uavgn + c_bar) * rho_log / (2 * g);
  const double l4 = fabs(uavgn) * p_avg;
  const double* dvu = dv + 1;
  const double dvr = dv[0], dve = dv[D + 1];
  double ut[D];
#pragma unroll
  for (int d = 0; d < D; ++d) ut[d] = u_avg[d] - uavgn * n[d];
  const double w1 = l1 * (dvr + dotv<D>(umc, dvu) + (h_bar - c_bar * uavgn) * dve);
  const double k2 = GRAV ? (u2_bar / 2 + phi_avg) : 0.0;
  const double w2 = GRAV ? l2 * (dvr + dotv<D>(u_avg, dvu) + k2 * dve)
                         : l2 * (dvr + dotv<D>(u_avg, dvu) + u2_bar / 2 * dve);
  const double w3 = l3 * (dvr + dotv<D>(upc, dvu) + (h_bar + c_bar * uavgn) * dve);
  double Dm[Ns];
  Dm[0] = w1 + w2 + w3;
  const double ndv = dotv<D>(n, dvu);
#pragma unroll
  for (int d = 0; d < D; ++d)
    Dm[1 + d] = w1 * umc[d] + w2 * u_avg[d] + w3 * upc[d] + l4 * (dvu[d] - ndv * n[d] + dve * ut[d]);
  const double e2 = GRAV ? w2 * k2 : w2 * u2_bar / 2;
  Dm[D + 1] = w1 * (h_bar - c_bar * uavgn) + e2 + w3 * (h_bar + c_bar * uavgn) +
              l4 * (dotv<D>(ut, dvu) + dve * (dotv<D>(u_avg, u_avg) - uavgn * uavgn));
#pragma unroll
  for (int s = 0; s < Ns; ++s) nf[s] = ec[s] - Dm[s] / 2;
}

template <int D>
template <int SF>
ATM_DEV void Euler<D>::special_surface(const LawConst& c, const double* n, const double* qm,
                                       const double* am, const double* qp, const double* ap,
                                       double* nf) {
  if constexpr (SF == ATM_FLUX_ROE) euler_roe<D, false>(c, n, qm, am, qp, ap, nf);
  else euler_matrix<D, false>(c, n, qm, am, qp, ap, nf);
}

template <int D>
template <int SF>
ATM_DEV void EulerGravity<D>::special_surface(const LawConst& c, const double* n, const double* qm,
                                              const double* am, const double* qp, const double* ap,
                                              double* nf) {
  if constexpr (SF == ATM_FLUX_ROE) euler_roe<D, true>(c, n, qm, am, qp, ap, nf);
  else euler_matrix<D, true>(c, n, qm, am, qp, ap, nf);
}

// shallow_water.jl:68-110
template <int D>
template <int SF>
ATM_DEV void ShallowWater<D>::special_surface(const LawConst& c, const double* n, const double* qm,
                                              const double* am, const double* qp, const double* ap,
                                              double* nf) {
  const double g = c.grav;
  double Fm[D][Ns], Fp[D][Ns];
  flux(c, qm, am, Fm);
  flux(c, qp, ap, Fp);
  const double rm = qm[0], rp = qp[0];
  double um[D], up[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    um[d] = qm[1 + d] / rm;
    up[d] = qp[1 + d] / rp;
  }
  const double thm = qm[D + 1] / rm, thp = qp[D + 1] / rp;
  const double pm = g * (rm * rm) / 2, pp = g * (rp * rp) / 2;
  const double cm = sqrt(g * rm), cp = sqrt(g * rp);
  const double rho = sqrt(rm * rp);
  double u[D];
#pragma unroll
  for (int d = 0; d < D; ++d) u[d] = roe_avg(rm, rp, um[d], up[d]);
  const double th = roe_avg(rm, rp, thm, thp);
  const double cs = roe_avg(rm, rp, cm, cp);
  const double un = dotv<D>(u, n);
  const double drho = rp - rm, dp = pp - pm;
  double du[D];
#pragma unroll
  for (int d = 0; d < D; ++d) du[d] = up[d] - um[d];
  const double drt = qp[D + 1] - qm[D + 1];
  const double dun = dotv<D>(du, n);
  const double cm2 = 1 / (cs * cs);
  const double w1 = fabs(un - cs) * (dp - rho * cs * dun) * cm2 / 2;
  const double w2 = fabs(un + cs) * (dp + rho * cs * dun) * cm2 / 2;
  const double w3 = fabs(un) * (drho - dp * cm2);
  const double w4 = fabs(un) * rho;
  const double w5 = fabs(un) * (drt - th * dp * cm2);
  double fp[Ns];
  fp[0] = (w1 + w2 + w3) / 2;
#pragma unroll
  for (int d = 0; d < D; ++d)
    fp[1 + d] = (w1 * (u[d] - cs * n[d]) + w2 * (u[d] + cs * n[d]) + w3 * u[d] + w4 * (du[d] - dun * n[d])) / 2;
  fp[D + 1] = ((w1 + w2) * th + w5) / 2;
#pragma unroll
  for (int s = 0; s < Ns; ++s) {
    double a = (Fm[0][s] + Fp[0][s]) * n[0];
#pragma unroll
    for (int d = 1; d < D; ++d) a += (Fm[d][s] + Fp[d][s]) * n[d];
    nf[s] = a / 2 - fp[s];
  }
}

// Numerical surface flux nf(n, q-, q+) (src/numericalfluxes.jl:26-42).
template <class L, int SF>
ATM_DEV void surfaceflux(const LawConst& c, const double* n, const double* qm, const double* am,
                         const double* qp, const double* ap, double* nf) {
  constexpr int D = L::D, Ns = L::Ns;
  if constexpr (SF == ATM_FLUX_ROE || SF == ATM_FLUX_MATRIX) {
    L::template special_surface<SF>(c, n, qm, am, qp, ap, nf);
  } else {
    double F[D][Ns];
    constexpr int TP = (SF == ATM_FLUX_RUSANOV) ? ATM_FLUX_CENTRAL : SF;
    L::template twopoint<TP>(c, qm, am, qp, ap, F);
#pragma unroll
    for (int s = 0; s < Ns; ++s) {
      double a = F[0][s] * n[0];
#pragma unroll
      for (int d = 1; d < D; ++d) a += F[d][s] * n[d];
      nf[s] = a;
    }
    if constexpr (SF == ATM_FLUX_RUSANOV) {
      const double ws = fmax(L::wavespeed(c, n, qm, am), L::wavespeed(c, n, qp, ap));
#pragma unroll
      for (int s = 0; s < Ns; ++s) nf[s] = nf[s] - ws * (qp[s] - qm[s]) / 2;
    }
  }
}

// boundarystate plug-ins (enumerated): copy (src/balancelaw.jl:21), reflective
// wall (experiment/euler_gravity/risingbubble.jl:10-15), zero inflow
// (test/advection/imex_wave_2d.jl:19-26), prescribed by tag (experiment/euler/sod.jl:20-23)
template <class L>
ATM_DEV void boundarystate(const LawConst& c, const double* n, const double* qm, int tag, double* qp) {
  constexpr int D = L::D, Ns = L::Ns;
  if (c.bc == ATM_BC_REFLECT) {
    L::reflect(n, qm, qp);
  } else if (c.bc == ATM_BC_ZERO_INFLOW) {
    const double un = dotv<D>(c.u, n);
#pragma unroll
    for (int s = 0; s < Ns; ++s) qp[s] = (un <= 0) ? 0.0 : qm[s];
  } else if (c.bc == ATM_BC_PRESCRIBED) {
    const int t = tag <= c.n_bc_states ? tag : c.n_bc_states;
#pragma unroll
    for (int s = 0; s < Ns; ++s) qp[s] = c.bc_states[(t - 1) * Ns + s];
  } else {
#pragma unroll
    for (int s = 0; s < Ns; ++s) qp[s] = qm[s];
  }
}

}  // namespace atm
