// Pair-symmetric EC flux-differencing kernels for EulerGravityLaw (ATM_PRECISION_FAST).
#include "resolve.hpp"

namespace atm {
Resolved resolve_fast_euler_gravity(int dim, int sflux, int Nq) {
  if (dim == 2) return fast_nq<EulerGravity, 2>(sflux, Nq);
  if (dim == 3) return fast_nq<EulerGravity, 3>(sflux, Nq);
  return Resolved{nullptr, 0, 0, nullptr, 1};
}
}  // namespace atm
