// Pair-symmetric EC flux-differencing kernels for EulerLaw (ATM_PRECISION_FAST).
#include "resolve.hpp"

namespace atm {
Resolved resolve_fast_euler(int dim, int sflux, int Nq) {
  if (dim == 2) return fast_nq<Euler, 2>(sflux, Nq);
  if (dim == 3) return fast_nq<Euler, 3>(sflux, Nq);
  return Resolved{nullptr, 0, 0, nullptr, 1};
}
}  // namespace atm
