// CPU ORACLE — test infrastructure only.
//
// This file is part of the FP64 CPU restatement of Atum.jl's DGSEM path that
// the tests, the smoke check and the bench's CPU-baseline leg use as the
// checker.  Nothing in the product library includes, links or calls it.
//
// Balance laws and numerical fluxes, restated from the reference text with
// its operation order kept (Julia evaluates `a * b * c` as `(a * b) * c`,
// `x / 2y` as `x / (2 * y)`, StaticArrays dot products as a left fold).
//
//   src/numericalfluxes.jl:8-47            avg, logavg, roe_avg, Central, Rusanov
//   src/balancelaws/advection.jl:8-19      AdvectionLaw
//   src/balancelaws/euler.jl:9-256         EulerLaw (flux, Roe, EC, KG, Matrix)
//   src/balancelaws/euler_gravity.jl:10-332 EulerGravityLaw
//   src/balancelaws/linear_euler_gravity.jl:3-65 LinearEulerGravityLaw
//   src/balancelaws/shallow_water.jl:9-136 ShallowWaterLaw
//   src/balancelaw.jl:17-24                default auxiliary / boundarystate / sources
#pragma once
#include <cfloat>
#include <cmath>
#include <cstdint>

namespace orc {

enum LawId { LAW_ADVECTION = 0, LAW_EULER = 1, LAW_EULER_GRAVITY = 2,
             LAW_LINEAR_EULER_GRAVITY = 3, LAW_SHALLOW_WATER = 4 };
enum FluxId { FLUX_CENTRAL = 0, FLUX_RUSANOV = 1, FLUX_EC = 2, FLUX_KG = 3,
              FLUX_ROE = 4, FLUX_MATRIX = 5 };
enum BcId { BC_COPY = 0, BC_REFLECT = 1, BC_ZERO_INFLOW = 2, BC_PRESCRIBED = 3 };
enum SourceId { SRC_NONE = 0, SRC_CORIOLIS = 1 };

struct Params {
  double gamma = 1.4;
  double grav = 9.81;
  double u[3] = {1, 1, 1};
  int pde_level_balance = 0;
  int bc = BC_COPY;
  int source = SRC_NONE;
  double omega = 0;
  int n_bc_states = 0;
  const double* bc_states = nullptr;
};

// Thrown (inside the oracle only) when a law x flux pair has no method in the
// reference (Julia would raise a MethodError).
struct Unsupported {};

template <int D>
static inline double dotv(const double* a, const double* b) {
  double s = a[0] * b[0];
  for (int d = 1; d < D; ++d) s += a[d] * b[d];
  return s;
}

// src/numericalfluxes.jl:8
static inline double avg(double a, double b) { return (a + b) / 2; }

// src/numericalfluxes.jl:9-22
static inline double logavg(double a, double b) {
  double zeta = a / b;
  double f = (zeta - 1) / (zeta + 1);
  double u = f * f;
  double F;
  if (u < DBL_EPSILON) {
    // @evalpoly(u, 1, 1/3, 1/5, 1/7, 1/9) in Horner form
    F = 1.0 + u * (1.0 / 3 + u * (1.0 / 5 + u * (1.0 / 7 + u * (1.0 / 9))));
  } else {
    F = std::log(zeta) / (2 * f);
  }
  return (a + b) / (2 * F);
}

// src/numericalfluxes.jl:23
static inline double roe_avg(double rm, double rp, double sm, double sp) {
  return (std::sqrt(rm) * sm + std::sqrt(rp) * sp) / (std::sqrt(rm) + std::sqrt(rp));
}

// ---------------------------------------------------------------------------
// Advection: src/balancelaws/advection.jl:8-19
template <int D_>
struct Advection {
  static constexpr int D = D_, Ns = 1;
  static void flux(const Params& P, const double* q, const double*, double F[D][Ns]) {
    for (int d = 0; d < D; ++d) F[d][0] = P.u[d] * q[0];
  }
  static double wavespeed(const Params& P, const double* n, const double*, const double*) {
    return std::fabs(dotv<D>(n, P.u));
  }
  static double entropy(const Params&, const double* q, const double*) { return q[0] * q[0] / 2; }
  static void entropyvariables(const Params&, const double* q, const double*, double* v) { v[0] = q[0]; }
  static void twopoint(int vf, const Params& P, const double* q1, const double* a1,
                       const double* q2, const double* a2, double F[D][Ns]);
  static void surface_special(int, const Params&, const double*, const double*, const double*,
                              const double*, const double*, double*) { throw Unsupported(); }
  static void law_source(const Params&, double*, const double*, const double*, int) {}
  static void nonconservative(const Params&, double*, const double*, const double*, int) {}
  static void problem_source(const Params&, double*, const double*, const double*, int) {}
  static void reflect(const double*, const double* qm, double* qp) { qp[0] = qm[0]; }
};

// ---------------------------------------------------------------------------
// Euler: src/balancelaws/euler.jl
template <int D_>
struct Euler {
  static constexpr int D = D_, Ns = D_ + 2;
  // :32-43 (pressure, soundspeed)
  static double pressure(const Params& P, double rho, const double* rhou, double rhoe) {
    return (P.gamma - 1) * (rhoe - dotv<D>(rhou, rhou) / (2 * rho));
  }
  static double soundspeed(const Params& P, double rho, double p) {
    return std::sqrt(P.gamma * p / rho);
  }
  // :48-59
  static void flux(const Params& P, const double* q, const double*, double F[D][Ns]) {
    double rho = q[0], rhoe = q[D + 1];
    const double* rhou = q + 1;
    double u[D];
    for (int d = 0; d < D; ++d) u[d] = rhou[d] / rho;
    double p = pressure(P, rho, rhou, rhoe);
    for (int d = 0; d < D; ++d) {
      F[d][0] = rhou[d];
      // `ρu⃗ * u⃗' + p * I`: p is added on the diagonal only
      for (int e = 0; e < D; ++e) F[d][1 + e] = (d == e) ? rhou[d] * u[e] + p : rhou[d] * u[e];
      F[d][D + 1] = u[d] * (rhoe + p);
    }
  }
  // :61-66
  static double wavespeed(const Params& P, const double* n, const double* q, const double*) {
    double rho = q[0];
    double u[D];
    for (int d = 0; d < D; ++d) u[d] = q[1 + d] / rho;
    return std::fabs(dotv<D>(n, u)) + soundspeed(P, rho, pressure(P, rho, q + 1, q[D + 1]));
  }
  // :68-75
  static double entropy(const Params& P, const double* q, const double*) {
    double rho = q[0];
    double p = pressure(P, rho, q + 1, q[D + 1]);
    double s = std::log(p / std::pow(rho, P.gamma));
    return -rho * s / (P.gamma - 1);
  }
  // :77-89
  static void entropyvariables(const Params& P, const double* q, const double*, double* v) {
    double rho = q[0];
    double g = P.gamma;
    double p = pressure(P, rho, q + 1, q[D + 1]);
    double s = std::log(p / std::pow(rho, g));
    double b = rho / (2 * p);
    double u[D];
    for (int d = 0; d < D; ++d) u[d] = q[1 + d] / rho;
    v[0] = (g - s) / (g - 1) - dotv<D>(u, u) * b;
    for (int d = 0; d < D; ++d) v[1 + d] = (2 * b) * u[d];
    v[D + 1] = -2 * b;
  }
  static void twopoint(int vf, const Params& P, const double* q1, const double* a1,
                       const double* q2, const double* a2, double F[D][Ns]);
  static void surface_special(int sf, const Params& P, const double* n, const double* qm,
                              const double* am, const double* qp, const double* ap, double* nf);
  static void law_source(const Params&, double*, const double*, const double*, int) {}
  static void nonconservative(const Params&, double*, const double*, const double*, int) {}
  static void problem_source(const Params&, double*, const double*, const double*, int) {}
  static void reflect(const double* n, const double* qm, double* qp) {
    // experiment/euler_gravity/risingbubble.jl:10-15 (same form for every law with momentum)
    double dn = dotv<D>(n, qm + 1);
    qp[0] = qm[0];
    for (int d = 0; d < D; ++d) qp[1 + d] = qm[1 + d] - 2 * dn * n[d];
    qp[D + 1] = qm[D + 1];
  }
};

// ---------------------------------------------------------------------------
// EulerGravity: src/balancelaws/euler_gravity.jl
// aux layout (:42-55): [Phi, gradPhi(D), rho_ref, p_ref (optional)]
template <int D_>
struct EulerGravity {
  static constexpr int D = D_, Ns = D_ + 2;
  static double Phi(const double* aux) { return aux[0]; }
  // :64-67
  static double pressure(const Params& P, double rho, const double* rhou, double rhoe, double phi) {
    return (P.gamma - 1) * (rhoe - dotv<D>(rhou, rhou) / (2 * rho) - rho * phi);
  }
  static double soundspeed(const Params& P, double rho, double p) {
    return std::sqrt(P.gamma * p / rho);
  }
  // :80-98
  static void flux(const Params& P, const double* q, const double* aux, double F[D][Ns]) {
    double rho = q[0], rhoe = q[D + 1];
    const double* rhou = q + 1;
    double phi = Phi(aux);
    double u[D];
    for (int d = 0; d < D; ++d) u[d] = rhou[d] / rho;
    double p = pressure(P, rho, rhou, rhoe, phi);
    double dp = p;
    if (P.pde_level_balance) dp -= aux[D + 2];
    for (int d = 0; d < D; ++d) {
      F[d][0] = rhou[d];
      for (int e = 0; e < D; ++e) F[d][1 + e] = (d == e) ? rhou[d] * u[e] + dp : rhou[d] * u[e];
      F[d][D + 1] = u[d] * (rhoe + p);
    }
  }
  // :100-113
  static void nonconservative(const Params& P, double* dq, const double* q, const double* aux,
                              int dirmask) {
    if (dirmask & (1 << (D - 1))) {
      double rho = q[0];
      if (P.pde_level_balance) rho -= aux[D + 1];
      for (int d = 0; d < D; ++d) dq[1 + d] -= rho * aux[1 + d];
    }
  }
  // :115-122
  static double wavespeed(const Params& P, const double* n, const double* q, const double* aux) {
    double rho = q[0];
    double u[D];
    for (int d = 0; d < D; ++d) u[d] = q[1 + d] / rho;
    return std::fabs(dotv<D>(n, u)) +
           soundspeed(P, rho, pressure(P, rho, q + 1, q[D + 1], Phi(aux)));
  }
  // :124-132
  static double entropy(const Params& P, const double* q, const double* aux) {
    double rho = q[0];
    double p = pressure(P, rho, q + 1, q[D + 1], Phi(aux));
    double s = std::log(p / std::pow(rho, P.gamma));
    return -rho * s / (P.gamma - 1);
  }
  // :134-147
  static void entropyvariables(const Params& P, const double* q, const double* aux, double* v) {
    double rho = q[0], g = P.gamma, phi = Phi(aux);
    double p = pressure(P, rho, q + 1, q[D + 1], phi);
    double s = std::log(p / std::pow(rho, g));
    double b = rho / (2 * p);
    double u[D];
    for (int d = 0; d < D; ++d) u[d] = q[1 + d] / rho;
    v[0] = (g - s) / (g - 1) - (dotv<D>(u, u) - 2 * phi) * b;
    for (int d = 0; d < D; ++d) v[1 + d] = (2 * b) * u[d];
    v[D + 1] = -2 * b;
  }
  static void twopoint(int vf, const Params& P, const double* q1, const double* a1,
                       const double* q2, const double* a2, double F[D][Ns]);
  static void surface_special(int sf, const Params& P, const double* n, const double* qm,
                              const double* am, const double* qp, const double* ap, double* nf);
  static void law_source(const Params&, double*, const double*, const double*, int) {}
  // experiment/euler_gravity/baroclinicwave.jl:27-34 (Coriolis, 3-D only)
  static void problem_source(const Params& P, double* dq, const double* q, const double*, int dirmask) {
    if (P.source == SRC_CORIOLIS && D == 3 && (dirmask & (1 << (D - 1)))) {
      double a3 = 2 * P.omega;
      const double* b = q + 1;
      double c0 = 0.0 * b[2] - a3 * b[1];
      double c1 = a3 * b[0] - 0.0 * b[2];
      double c2 = 0.0 * b[1] - 0.0 * b[0];
      dq[1] -= c0;
      if (D > 1) dq[2] -= c1;
      if (D > 2) dq[3] -= c2;
    }
  }
  static void reflect(const double* n, const double* qm, double* qp) {
    Euler<D>::reflect(n, qm, qp);
  }
};

// ---------------------------------------------------------------------------
// LinearEulerGravity: src/balancelaws/linear_euler_gravity.jl
template <int D_>
struct LinearEulerGravity {
  static constexpr int D = D_, Ns = D_ + 2;
  // :29-32
  static double pressure(const Params& P, double rho, double rhoe, double phi) {
    return (P.gamma - 1) * (rhoe - rho * phi);
  }
  // :37-54
  static void flux(const Params& P, const double* q, const double* aux, double F[D][Ns]) {
    double rho = q[0], rhoe = q[D + 1];
    const double* rhou = q + 1;
    double phi = aux[0];
    double p = pressure(P, rho, rhoe, phi);
    double ref_rho = aux[D + 1], ref_p = aux[D + 2];
    double ref_rhoe = ref_p / (P.gamma - 1) + ref_rho * phi;  // :17-20
    double h = (ref_rhoe + ref_p) / ref_rho;
    for (int d = 0; d < D; ++d) {
      F[d][0] = rhou[d];
      for (int e = 0; e < D; ++e) F[d][1 + e] = (d == e) ? p : 0.0;
      F[d][D + 1] = h * rhou[d];
    }
  }
  // :56 (reference sound speed, state independent)
  static double wavespeed(const Params& P, const double*, const double*, const double* aux) {
    return std::sqrt(P.gamma * aux[D + 2] / aux[D + 1]);
  }
  static double entropy(const Params&, const double*, const double*) { throw Unsupported(); }
  static void entropyvariables(const Params&, const double*, const double*, double*) { throw Unsupported(); }
  static void twopoint(int vf, const Params& P, const double* q1, const double* a1,
                       const double* q2, const double* a2, double F[D][Ns]);
  static void surface_special(int, const Params&, const double*, const double*, const double*,
                              const double*, const double*, double*) { throw Unsupported(); }
  // :58-65
  static void law_source(const Params&, double* dq, const double* q, const double* aux, int dirmask) {
    if (dirmask & (1 << (D - 1))) {
      double rho = q[0];
      for (int d = 0; d < D; ++d) dq[1 + d] -= rho * aux[1 + d];
    }
  }
  static void nonconservative(const Params&, double*, const double*, const double*, int) {}
  static void problem_source(const Params&, double*, const double*, const double*, int) {}
  static void reflect(const double* n, const double* qm, double* qp) {
    Euler<D>::reflect(n, qm, qp);
  }
};

// ---------------------------------------------------------------------------
// ShallowWater: src/balancelaws/shallow_water.jl
template <int D_>
struct ShallowWater {
  static constexpr int D = D_, Ns = D_ + 2;
  // :32-43
  static void flux(const Params& P, const double* q, const double*, double F[D][Ns]) {
    double rho = q[0], rhot = q[D + 1];
    const double* rhou = q + 1;
    double u[D];
    for (int d = 0; d < D; ++d) u[d] = rhou[d] / rho;
    double p = P.grav * (rho * rho) / 2;
    for (int d = 0; d < D; ++d) {
      F[d][0] = rhou[d];
      for (int e = 0; e < D; ++e) F[d][1 + e] = (d == e) ? rhou[d] * u[e] + p : rhou[d] * u[e];
      F[d][D + 1] = u[d] * rhot;
    }
  }
  // :45-50
  static double wavespeed(const Params& P, const double* n, const double* q, const double*) {
    double rho = q[0];
    double u[D];
    for (int d = 0; d < D; ++d) u[d] = q[1 + d] / rho;
    return std::fabs(dotv<D>(n, u)) + std::sqrt(P.grav * rho);
  }
  // :52-56
  static double entropy(const Params& P, const double* q, const double*) {
    double rho = q[0];
    return dotv<D>(q + 1, q + 1) / (2 * rho) + P.grav * (rho * rho) / 2;
  }
  // :58-66
  static void entropyvariables(const Params& P, const double* q, const double*, double* v) {
    double rho = q[0];
    double mr[D];
    for (int d = 0; d < D; ++d) mr[d] = -q[1 + d];
    v[0] = dotv<D>(mr, q + 1) / (2 * (rho * rho)) + P.grav * rho;
    for (int d = 0; d < D; ++d) v[1 + d] = q[1 + d] / rho;
    v[D + 1] = 0.0;
  }
  static void twopoint(int vf, const Params& P, const double* q1, const double* a1,
                       const double* q2, const double* a2, double F[D][Ns]);
  static void surface_special(int sf, const Params& P, const double* n, const double* qm,
                              const double* am, const double* qp, const double* ap, double* nf);
  static void law_source(const Params&, double*, const double*, const double*, int) {}
  static void nonconservative(const Params&, double*, const double*, const double*, int) {}
  static void problem_source(const Params&, double*, const double*, const double*, int) {}
  static void reflect(const double* n, const double* qm, double* qp) {
    Euler<D>::reflect(n, qm, qp);  // experiment/shallow_water/bickleyjet.jl:11-16
  }
};

// ---------------------------------------------------------------------------
// Central two-point flux: src/numericalfluxes.jl:31-34
template <class L>
static void central_twopoint(const Params& P, const double* q1, const double* a1,
                             const double* q2, const double* a2, double F[L::D][L::Ns]) {
  double F1[L::D][L::Ns], F2[L::D][L::Ns];
  L::flux(P, q1, a1, F1);
  L::flux(P, q2, a2, F2);
  for (int d = 0; d < L::D; ++d)
    for (int s = 0; s < L::Ns; ++s) F[d][s] = (F1[d][s] + F2[d][s]) / 2;
}

template <int D>
void Advection<D>::twopoint(int vf, const Params& P, const double* q1, const double* a1,
                            const double* q2, const double* a2, double F[D][Ns]) {
  if (vf == FLUX_CENTRAL) return central_twopoint<Advection<D>>(P, q1, a1, q2, a2, F);
  throw Unsupported();
}

template <int D>
void LinearEulerGravity<D>::twopoint(int vf, const Params& P, const double* q1, const double* a1,
                                     const double* q2, const double* a2, double F[D][Ns]) {
  if (vf == FLUX_CENTRAL) return central_twopoint<LinearEulerGravity<D>>(P, q1, a1, q2, a2, F);
  throw Unsupported();
}

// Euler EC: src/balancelaws/euler.jl:140-169;  KG: :171-196
template <int D>
void Euler<D>::twopoint(int vf, const Params& P, const double* q1, const double* a1,
                        const double* q2, const double* a2, double F[D][Ns]) {
  if (vf == FLUX_CENTRAL) return central_twopoint<Euler<D>>(P, q1, a1, q2, a2, F);
  const double g = P.gamma;
  double rho1 = q1[0], rhoe1 = q1[D + 1], rho2 = q2[0], rhoe2 = q2[D + 1];
  double u1[D], u2[D];
  for (int d = 0; d < D; ++d) { u1[d] = q1[1 + d] / rho1; u2[d] = q2[1 + d] / rho2; }
  double p1 = pressure(P, rho1, q1 + 1, rhoe1);
  double p2 = pressure(P, rho2, q2 + 1, rhoe2);
  if (vf == FLUX_EC) {
    double b1 = rho1 / (2 * p1), b2 = rho2 / (2 * p2);
    double rho_avg = avg(rho1, rho2);
    double u_avg[D];
    for (int d = 0; d < D; ++d) u_avg[d] = avg(u1[d], u2[d]);
    double b_avg = avg(b1, b2);
    double u2_avg = avg(dotv<D>(u1, u1), dotv<D>(u2, u2));
    double rho_log = logavg(rho1, rho2);
    double b_log = logavg(b1, b2);
    double frho[D];
    for (int d = 0; d < D; ++d) frho[d] = u_avg[d] * rho_log;
    double pd = rho_avg / (2 * b_avg);
    double frhou[D][D];
    for (int d = 0; d < D; ++d)
      for (int e = 0; e < D; ++e) frhou[d][e] = (d == e) ? u_avg[d] * frho[e] + pd : u_avg[d] * frho[e];
    double c = 1 / (2 * (g - 1) * b_log) - u2_avg / 2;
    for (int d = 0; d < D; ++d) {
      double mv = frhou[d][0] * u_avg[0];
      for (int e = 1; e < D; ++e) mv += frhou[d][e] * u_avg[e];
      F[d][0] = frho[d];
      for (int e = 0; e < D; ++e) F[d][1 + e] = frhou[d][e];
      F[d][D + 1] = c * frho[d] + mv;
    }
    return;
  }
  if (vf == FLUX_KG) {
    double e1 = rhoe1 / rho1, e2 = rhoe2 / rho2;
    double rho_avg = avg(rho1, rho2);
    double u_avg[D];
    for (int d = 0; d < D; ++d) u_avg[d] = avg(u1[d], u2[d]);
    double e_avg = avg(e1, e2), p_avg = avg(p1, p2);
    for (int d = 0; d < D; ++d) {
      double frho_d = u_avg[d] * rho_avg;
      F[d][0] = frho_d;
    }
    for (int d = 0; d < D; ++d) {
      for (int e = 0; e < D; ++e) {
        double frho_e = u_avg[e] * rho_avg;
        F[d][1 + e] = (d == e) ? u_avg[d] * frho_e + p_avg : u_avg[d] * frho_e;
      }
      F[d][D + 1] = u_avg[d] * (rho_avg * e_avg + p_avg);
    }
    return;
  }
  throw Unsupported();
}

// EulerGravity EC: src/balancelaws/euler_gravity.jl:200-236;  KG: :238-269
template <int D>
void EulerGravity<D>::twopoint(int vf, const Params& P, const double* q1, const double* a1,
                               const double* q2, const double* a2, double F[D][Ns]) {
  if (vf == FLUX_CENTRAL) return central_twopoint<EulerGravity<D>>(P, q1, a1, q2, a2, F);
  const double g = P.gamma;
  double rho1 = q1[0], rhoe1 = q1[D + 1], rho2 = q2[0], rhoe2 = q2[D + 1];
  double phi1 = Phi(a1), phi2 = Phi(a2);
  double u1[D], u2[D];
  for (int d = 0; d < D; ++d) { u1[d] = q1[1 + d] / rho1; u2[d] = q2[1 + d] / rho2; }
  double p1 = pressure(P, rho1, q1 + 1, rhoe1, phi1);
  double p2 = pressure(P, rho2, q2 + 1, rhoe2, phi2);
  if (vf == FLUX_EC) {
    double b1 = rho1 / (2 * p1), b2 = rho2 / (2 * p2);
    double rho_avg = avg(rho1, rho2);
    double u_avg[D];
    for (int d = 0; d < D; ++d) u_avg[d] = avg(u1[d], u2[d]);
    double b_avg = avg(b1, b2);
    double phi_avg = avg(phi1, phi2);
    double u2_avg = avg(dotv<D>(u1, u1), dotv<D>(u2, u2));
    double rho_log = logavg(rho1, rho2);
    double b_log = logavg(b1, b2);
    double frho[D];
    for (int d = 0; d < D; ++d) frho[d] = u_avg[d] * rho_log;
    double pd = rho_avg / (2 * b_avg);
    double frhou[D][D];
    for (int d = 0; d < D; ++d)
      for (int e = 0; e < D; ++e) frhou[d][e] = (d == e) ? u_avg[d] * frho[e] + pd : u_avg[d] * frho[e];
    double c = 1 / (2 * (g - 1) * b_log) - u2_avg / 2 + phi_avg;
    double frhoe[D];
    for (int d = 0; d < D; ++d) {
      double mv = frhou[d][0] * u_avg[0];
      for (int e = 1; e < D; ++e) mv += frhou[d][e] * u_avg[e];
      frhoe[d] = c * frho[d] + mv;
    }
    // fluctuation (:231-233)
    double alpha = b_avg * rho_log / (2 * b1);
    double fl = alpha * (phi1 - phi2);
    for (int d = 0; d < D; ++d) frhou[d][d] -= fl;
    for (int d = 0; d < D; ++d) {
      F[d][0] = frho[d];
      for (int e = 0; e < D; ++e) F[d][1 + e] = frhou[d][e];
      F[d][D + 1] = frhoe[d];
    }
    return;
  }
  if (vf == FLUX_KG) {
    double e1 = rhoe1 / rho1, e2 = rhoe2 / rho2;
    double rho_avg = avg(rho1, rho2);
    double u_avg[D];
    for (int d = 0; d < D; ++d) u_avg[d] = avg(u1[d], u2[d]);
    double e_avg = avg(e1, e2), p_avg = avg(p1, p2);
    double alpha = rho_avg / 2;
    double fl = alpha * (phi1 - phi2);
    for (int d = 0; d < D; ++d) F[d][0] = u_avg[d] * rho_avg;
    for (int d = 0; d < D; ++d) {
      for (int e = 0; e < D; ++e) {
        double frho_e = u_avg[e] * rho_avg;
        double v = (d == e) ? u_avg[d] * frho_e + p_avg : u_avg[d] * frho_e;
        if (d == e) v -= fl;
        F[d][1 + e] = v;
      }
      F[d][D + 1] = u_avg[d] * (rho_avg * e_avg + p_avg);
    }
    return;
  }
  throw Unsupported();
}

// ShallowWater EC: src/balancelaws/shallow_water.jl:112-136
template <int D>
void ShallowWater<D>::twopoint(int vf, const Params& P, const double* q1, const double* a1,
                               const double* q2, const double* a2, double F[D][Ns]) {
  if (vf == FLUX_CENTRAL) return central_twopoint<ShallowWater<D>>(P, q1, a1, q2, a2, F);
  if (vf != FLUX_EC) throw Unsupported();
  double rho1 = q1[0], rho2 = q2[0];
  double u1[D], u2[D];
  for (int d = 0; d < D; ++d) { u1[d] = q1[1 + d] / rho1; u2[d] = q2[1 + d] / rho2; }
  double th1 = q1[D + 1] / rho1, th2 = q2[D + 1] / rho2;
  double rho_avg = avg(rho1, rho2);
  double rho2_avg = avg(rho1 * rho1, rho2 * rho2);
  double u_avg[D], ru_avg[D];
  for (int d = 0; d < D; ++d) { u_avg[d] = avg(u1[d], u2[d]); ru_avg[d] = avg(q1[1 + d], q2[1 + d]); }
  double th_avg = avg(th1, th2);
  double pd = P.grav * (rho_avg * rho_avg - rho2_avg / 2);
  for (int d = 0; d < D; ++d) {
    F[d][0] = ru_avg[d];
    for (int e = 0; e < D; ++e)
      F[d][1 + e] = (d == e) ? ru_avg[d] * u_avg[e] + pd : ru_avg[d] * u_avg[e];
    F[d][D + 1] = ru_avg[d] * th_avg;
  }
}

// ---------------------------------------------------------------------------
// Roe / Matrix surface fluxes

// Euler Roe: src/balancelaws/euler.jl:91-138 ; Matrix: :198-256
template <int D>
void Euler<D>::surface_special(int sf, const Params& P, const double* n, const double* qm,
                               const double* am, const double* qp, const double* ap, double* nf) {
  const double g = P.gamma;
  if (sf == FLUX_ROE) {
    double Fm[D][Ns], Fp[D][Ns];
    flux(P, qm, am, Fm);
    flux(P, qp, ap, Fp);
    double rm = qm[0], rp = qp[0];
    double um[D], up[D];
    for (int d = 0; d < D; ++d) { um[d] = qm[1 + d] / rm; up[d] = qp[1 + d] / rp; }
    double em = qm[D + 1] / rm, ep = qp[D + 1] / rp;
    double pm = pressure(P, rm, qm + 1, qm[D + 1]), pp = pressure(P, rp, qp + 1, qp[D + 1]);
    double hm = em + pm / rm, hp = ep + pp / rp;
    double cm = soundspeed(P, rm, pm), cp = soundspeed(P, rp, pp);
    double rho = std::sqrt(rm * rp);
    double u[D];
    for (int d = 0; d < D; ++d) u[d] = roe_avg(rm, rp, um[d], up[d]);
    double h = roe_avg(rm, rp, hm, hp);
    double c = roe_avg(rm, rp, cm, cp);
    double un = dotv<D>(u, n);
    double drho = rp - rm, dp = pp - pm;
    double du[D];
    for (int d = 0; d < D; ++d) du[d] = up[d] - um[d];
    double dun = dotv<D>(du, n);
    double cm2 = 1 / (c * c);
    double w1 = std::fabs(un - c) * (dp - rho * c * dun) * cm2 / 2;
    double w2 = std::fabs(un + c) * (dp + rho * c * dun) * cm2 / 2;
    double w3 = std::fabs(un) * (drho - dp * cm2);
    double w4 = std::fabs(un) * rho;
    double fp[Ns];
    fp[0] = (w1 + w2 + w3) / 2;
    for (int d = 0; d < D; ++d)
      fp[1 + d] = (w1 * (u[d] - c * n[d]) + w2 * (u[d] + c * n[d]) + w3 * u[d] +
                   w4 * (du[d] - dun * n[d])) / 2;
    double w3u[D];
    for (int d = 0; d < D; ++d) w3u[d] = w3 * u[d];
    fp[D + 1] = (w1 * (h - c * un) + w2 * (h + c * un) + dotv<D>(w3u, u) / 2 +
                 w4 * (dotv<D>(u, du) - un * dun)) / 2;
    for (int s = 0; s < Ns; ++s) {
      double a = (Fm[0][s] + Fp[0][s]) * n[0];
      for (int d = 1; d < D; ++d) a += (Fm[d][s] + Fp[d][s]) * n[d];
      nf[s] = a / 2 - fp[s];
    }
    return;
  }
  if (sf == FLUX_MATRIX) {
    double Fe[D][Ns];
    twopoint(FLUX_EC, P, qm, am, qp, ap, Fe);
    double ec[Ns];
    for (int s = 0; s < Ns; ++s) {
      double a = Fe[0][s] * n[0];
      for (int d = 1; d < D; ++d) a += Fe[d][s] * n[d];
      ec[s] = a;
    }
    double rm = qm[0], rp = qp[0];
    double um[D], up[D];
    for (int d = 0; d < D; ++d) { um[d] = qm[1 + d] / rm; up[d] = qp[1 + d] / rp; }
    double pm = pressure(P, rm, qm + 1, qm[D + 1]), pp = pressure(P, rp, qp + 1, qp[D + 1]);
    double bm = rm / (2 * pm), bp = rp / (2 * pp);
    double rho_log = logavg(rm, rp), b_log = logavg(bm, bp);
    double u_avg[D];
    for (int d = 0; d < D; ++d) u_avg[d] = avg(um[d], up[d]);
    double p_avg = avg(rm, rp) / (2 * avg(bm, bp));
    double u2_bar = 2 * std::sqrt(dotv<D>(u_avg, u_avg)) -
                    avg(std::sqrt(dotv<D>(um, um)), std::sqrt(dotv<D>(up, up)));
    double h_bar = g / (2 * b_log * (g - 1)) + u2_bar / 2;
    double c_bar = std::sqrt(g * p_avg / rho_log);
    double umc[D], upc[D];
    for (int d = 0; d < D; ++d) { umc[d] = u_avg[d] - c_bar * n[d]; upc[d] = u_avg[d] + c_bar * n[d]; }
    double uavgn = dotv<D>(u_avg, n);
    double vm[Ns], vp[Ns], dv[Ns];
    entropyvariables(P, qm, am, vm);
    entropyvariables(P, qp, ap, vp);
    for (int s = 0; s < Ns; ++s) dv[s] = vp[s] - vm[s];
    double l1 = std::fabs(uavgn - c_bar) * rho_log / (2 * g);
    double l2 = std::fabs(uavgn) * rho_log * (g - 1) / g;
    double l3 = std::fabs(uavgn + c_bar) * rho_log / (2 * g);
    double l4 = std::fabs(uavgn) * p_avg;
    const double* dvu = dv + 1;
    double dvr = dv[0], dve = dv[D + 1];
    double ut[D];
    for (int d = 0; d < D; ++d) ut[d] = u_avg[d] - uavgn * n[d];
    double w1 = l1 * (dvr + dotv<D>(umc, dvu) + (h_bar - c_bar * uavgn) * dve);
    double w2 = l2 * (dvr + dotv<D>(u_avg, dvu) + u2_bar / 2 * dve);
    double w3 = l3 * (dvr + dotv<D>(upc, dvu) + (h_bar + c_bar * uavgn) * dve);
    double Dm[Ns];
    Dm[0] = w1 + w2 + w3;
    double ndv = dotv<D>(n, dvu);
    for (int d = 0; d < D; ++d)
      Dm[1 + d] = w1 * umc[d] + w2 * u_avg[d] + w3 * upc[d] +
                  l4 * (dvu[d] - ndv * n[d] + dve * ut[d]);
    Dm[D + 1] = w1 * (h_bar - c_bar * uavgn) + w2 * u2_bar / 2 + w3 * (h_bar + c_bar * uavgn) +
                l4 * (dotv<D>(ut, dvu) + dve * (dotv<D>(u_avg, u_avg) - uavgn * uavgn));
    for (int s = 0; s < Ns; ++s) nf[s] = ec[s] - Dm[s] / 2;
    return;
  }
  throw Unsupported();
}

// EulerGravity Roe: src/balancelaws/euler_gravity.jl:149-198 ; Matrix: :271-332
template <int D>
void EulerGravity<D>::surface_special(int sf, const Params& P, const double* n, const double* qm,
                                      const double* am, const double* qp, const double* ap,
                                      double* nf) {
  const double g = P.gamma;
  if (sf == FLUX_ROE) {
    double phi = Phi(am);  // minus-side geopotential for both pressures (:150)
    double Fm[D][Ns], Fp[D][Ns];
    flux(P, qm, am, Fm);
    flux(P, qp, ap, Fp);
    double rm = qm[0], rp = qp[0];
    double um[D], up[D];
    for (int d = 0; d < D; ++d) { um[d] = qm[1 + d] / rm; up[d] = qp[1 + d] / rp; }
    double em = qm[D + 1] / rm, ep = qp[D + 1] / rp;
    double pm = pressure(P, rm, qm + 1, qm[D + 1], phi), pp = pressure(P, rp, qp + 1, qp[D + 1], phi);
    double hm = em + pm / rm, hp = ep + pp / rp;
    double cm = soundspeed(P, rm, pm), cp = soundspeed(P, rp, pp);
    double rho = std::sqrt(rm * rp);
    double u[D];
    for (int d = 0; d < D; ++d) u[d] = roe_avg(rm, rp, um[d], up[d]);
    double h = roe_avg(rm, rp, hm, hp);
    double c = roe_avg(rm, rp, cm, cp);
    double un = dotv<D>(u, n);
    double drho = rp - rm, dp = pp - pm;
    double du[D];
    for (int d = 0; d < D; ++d) du[d] = up[d] - um[d];
    double dun = dotv<D>(du, n);
    double cm2 = 1 / (c * c);
    double w1 = std::fabs(un - c) * (dp - rho * c * dun) * cm2 / 2;
    double w2 = std::fabs(un + c) * (dp + rho * c * dun) * cm2 / 2;
    double w3 = std::fabs(un) * (drho - dp * cm2);
    double w4 = std::fabs(un) * rho;
    double fp[Ns];
    fp[0] = (w1 + w2 + w3) / 2;
    for (int d = 0; d < D; ++d)
      fp[1 + d] = (w1 * (u[d] - c * n[d]) + w2 * (u[d] + c * n[d]) + w3 * u[d] +
                   w4 * (du[d] - dun * n[d])) / 2;
    fp[D + 1] = (w1 * (h - c * un) + w2 * (h + c * un) + w3 * (dotv<D>(u, u) / 2 + phi) +
                 w4 * (dotv<D>(u, du) - un * dun)) / 2;
    for (int s = 0; s < Ns; ++s) {
      double a = (Fm[0][s] + Fp[0][s]) * n[0];
      for (int d = 1; d < D; ++d) a += (Fm[d][s] + Fp[d][s]) * n[d];
      nf[s] = a / 2 - fp[s];
    }
    return;
  }
  if (sf == FLUX_MATRIX) {
    double Fe[D][Ns];
    twopoint(FLUX_EC, P, qm, am, qp, ap, Fe);
    double ec[Ns];
    for (int s = 0; s < Ns; ++s) {
      double a = Fe[0][s] * n[0];
      for (int d = 1; d < D; ++d) a += Fe[d][s] * n[d];
      ec[s] = a;
    }
    double rm = qm[0], rp = qp[0];
    double phim = Phi(am), phip = Phi(ap);
    double um[D], up[D];
    for (int d = 0; d < D; ++d) { um[d] = qm[1 + d] / rm; up[d] = qp[1 + d] / rp; }
    double pm = pressure(P, rm, qm + 1, qm[D + 1], phim), pp = pressure(P, rp, qp + 1, qp[D + 1], phip);
    double bm = rm / (2 * pm), bp = rp / (2 * pp);
    double phi_avg = avg(phim, phip);
    double rho_log = logavg(rm, rp), b_log = logavg(bm, bp);
    double u_avg[D];
    for (int d = 0; d < D; ++d) u_avg[d] = avg(um[d], up[d]);
    double p_avg = avg(rm, rp) / (2 * avg(bm, bp));
    double u2_bar = 2 * std::sqrt(dotv<D>(u_avg, u_avg)) -
                    avg(std::sqrt(dotv<D>(um, um)), std::sqrt(dotv<D>(up, up)));
    double h_bar = g / (2 * b_log * (g - 1)) + u2_bar / 2 + phi_avg;
    double c_bar = std::sqrt(g * p_avg / rho_log);
    double umc[D], upc[D];
    for (int d = 0; d < D; ++d) { umc[d] = u_avg[d] - c_bar * n[d]; upc[d] = u_avg[d] + c_bar * n[d]; }
    double uavgn = dotv<D>(u_avg, n);
    double vm[Ns], vp[Ns], dv[Ns];
    entropyvariables(P, qm, am, vm);
    entropyvariables(P, qp, ap, vp);
    for (int s = 0; s < Ns; ++s) dv[s] = vp[s] - vm[s];
    double l1 = std::fabs(uavgn - c_bar) * rho_log / (2 * g);
    double l2 = std::fabs(uavgn) * rho_log * (g - 1) / g;
    double l3 = std::fabs(uavgn + c_bar) * rho_log / (2 * g);
    double l4 = std::fabs(uavgn) * p_avg;
    const double* dvu = dv + 1;
    double dvr = dv[0], dve = dv[D + 1];
    double ut[D];
    for (int d = 0; d < D; ++d) ut[d] = u_avg[d] - uavgn * n[d];
    double w1 = l1 * (dvr + dotv<D>(umc, dvu) + (h_bar - c_bar * uavgn) * dve);
    double w2 = l2 * (dvr + dotv<D>(u_avg, dvu) + (u2_bar / 2 + phi_avg) * dve);
    double w3 = l3 * (dvr + dotv<D>(upc, dvu) + (h_bar + c_bar * uavgn) * dve);
    double Dm[Ns];
    Dm[0] = w1 + w2 + w3;
    double ndv = dotv<D>(n, dvu);
    for (int d = 0; d < D; ++d)
      Dm[1 + d] = w1 * umc[d] + w2 * u_avg[d] + w3 * upc[d] +
                  l4 * (dvu[d] - ndv * n[d] + dve * ut[d]);
    Dm[D + 1] = w1 * (h_bar - c_bar * uavgn) + w2 * (u2_bar / 2 + phi_avg) +
                w3 * (h_bar + c_bar * uavgn) +
                l4 * (dotv<D>(ut, dvu) + dve * (dotv<D>(u_avg, u_avg) - uavgn * uavgn));
    for (int s = 0; s < Ns; ++s) nf[s] = ec[s] - Dm[s] / 2;
    return;
  }
  throw Unsupported();
}

// ShallowWater Roe: src/balancelaws/shallow_water.jl:68-110
template <int D>
void ShallowWater<D>::surface_special(int sf, const Params& P, const double* n, const double* qm,
                                      const double* am, const double* qp, const double* ap,
                                      double* nf) {
  if (sf != FLUX_ROE) throw Unsupported();
  const double g = P.grav;
  double Fm[D][Ns], Fp[D][Ns];
  flux(P, qm, am, Fm);
  flux(P, qp, ap, Fp);
  double rm = qm[0], rp = qp[0];
  double um[D], up[D];
  for (int d = 0; d < D; ++d) { um[d] = qm[1 + d] / rm; up[d] = qp[1 + d] / rp; }
  double thm = qm[D + 1] / rm, thp = qp[D + 1] / rp;
  double pm = g * (rm * rm) / 2, pp = g * (rp * rp) / 2;
  double cm = std::sqrt(g * rm), cp = std::sqrt(g * rp);
  double rho = std::sqrt(rm * rp);
  double u[D];
  for (int d = 0; d < D; ++d) u[d] = roe_avg(rm, rp, um[d], up[d]);
  double th = roe_avg(rm, rp, thm, thp);
  double c = roe_avg(rm, rp, cm, cp);
  double un = dotv<D>(u, n);
  double drho = rp - rm, dp = pp - pm;
  double du[D];
  for (int d = 0; d < D; ++d) du[d] = up[d] - um[d];
  double drt = qp[D + 1] - qm[D + 1];
  double dun = dotv<D>(du, n);
  double cm2 = 1 / (c * c);
  double w1 = std::fabs(un - c) * (dp - rho * c * dun) * cm2 / 2;
  double w2 = std::fabs(un + c) * (dp + rho * c * dun) * cm2 / 2;
  double w3 = std::fabs(un) * (drho - dp * cm2);
  double w4 = std::fabs(un) * rho;
  double w5 = std::fabs(un) * (drt - th * dp * cm2);
  double fp[Ns];
  fp[0] = (w1 + w2 + w3) / 2;
  for (int d = 0; d < D; ++d)
    fp[1 + d] = (w1 * (u[d] - c * n[d]) + w2 * (u[d] + c * n[d]) + w3 * u[d] +
                 w4 * (du[d] - dun * n[d])) / 2;
  fp[D + 1] = ((w1 + w2) * th + w5) / 2;
  for (int s = 0; s < Ns; ++s) {
    double a = (Fm[0][s] + Fp[0][s]) * n[0];
    for (int d = 1; d < D; ++d) a += (Fm[d][s] + Fp[d][s]) * n[d];
    nf[s] = a / 2 - fp[s];
  }
}

// ---------------------------------------------------------------------------
// Surface flux dispatch: src/numericalfluxes.jl:26-42 + per-law specialisations
template <class L>
static void surfaceflux(int sf, const Params& P, const double* n, const double* qm,
                        const double* am, const double* qp, const double* ap, double* nf) {
  constexpr int D = L::D, Ns = L::Ns;
  if (sf == FLUX_ROE || sf == FLUX_MATRIX) return L::surface_special(sf, P, n, qm, am, qp, ap, nf);
  if (sf == FLUX_RUSANOV) {
    double F[D][Ns];
    L::twopoint(FLUX_CENTRAL, P, qm, am, qp, ap, F);
    double wm = L::wavespeed(P, n, qm, am), wp = L::wavespeed(P, n, qp, ap);
    double ws = std::fmax(wm, wp);
    for (int s = 0; s < Ns; ++s) {
      double a = F[0][s] * n[0];
      for (int d = 1; d < D; ++d) a += F[d][s] * n[d];
      nf[s] = a - ws * (qp[s] - qm[s]) / 2;
    }
    return;
  }
  // default: twopointflux(flux, ...)' * n
  double F[D][Ns];
  L::twopoint(sf, P, qm, am, qp, ap, F);
  for (int s = 0; s < Ns; ++s) {
    double a = F[0][s] * n[0];
    for (int d = 1; d < D; ++d) a += F[d][s] * n[d];
    nf[s] = a;
  }
}

// boundarystate plug-ins: src/balancelaw.jl:21 (copy); reflective wall
// (experiment/euler_gravity/risingbubble.jl:10-15); zero inflow
// (test/advection/imex_wave_2d.jl:19-26); prescribed by tag (experiment/euler/sod.jl:20-23)
template <class L>
static void boundarystate(const Params& P, const double* n, const double* qm, int tag, double* qp) {
  constexpr int D = L::D, Ns = L::Ns;
  switch (P.bc) {
    case BC_REFLECT: L::reflect(n, qm, qp); return;
    case BC_ZERO_INFLOW: {
      double un = P.u[0] * n[0];
      for (int d = 1; d < D; ++d) un += P.u[d] * n[d];
      for (int s = 0; s < Ns; ++s) qp[s] = (un <= 0) ? 0.0 : qm[s];
      return;
    }
    case BC_PRESCRIBED: {
      int t = tag <= P.n_bc_states ? tag : P.n_bc_states;
      for (int s = 0; s < Ns; ++s) qp[s] = P.bc_states[(t - 1) * Ns + s];
      return;
    }
    default:
      for (int s = 0; s < Ns; ++s) qp[s] = qm[s];
  }
}

}  // namespace orc
