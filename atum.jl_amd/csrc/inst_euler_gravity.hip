// Element-kernel instantiations for EulerGravityLaw (src/balancelaws/euler_gravity.jl).
#include "resolve.hpp"

namespace atm {
Resolved resolve_euler_gravity(int dim, int form, int vflux, int sflux, int Nq, int Np) {
  using EG3 = EulerGravity<3>;
  if (dim == 3 && form == ATM_FORM_FLUX_DIFFERENCING && vflux == ATM_FLUX_EC) {
    // rising bubble: N = 4, Roe surface
    if (Nq == 5 && sflux == ATM_FLUX_ROE)
      return make_fd<EG3, 5, ATM_FLUX_EC, ATM_FLUX_ROE>(Nq, Np);
    // IMEX explicit operator: N = 3, Rusanov surface
    if (Nq == 4 && sflux == ATM_FLUX_RUSANOV)
      return make_fd<EG3, 4, ATM_FLUX_EC, ATM_FLUX_RUSANOV>(Nq, Np);
  }
  return resolve_dims<EulerGravity>(dim, form, vflux, Nq, Np);
}
}  // namespace atm
