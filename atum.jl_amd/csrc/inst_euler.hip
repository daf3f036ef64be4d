// Element-kernel instantiations for EulerLaw (src/balancelaws/euler.jl).
#include "resolve.hpp"

namespace atm {
Resolved resolve_euler(int dim, int form, int vflux, int sflux, int Nq, int Np) {
  // isentropic vortex: 2-D, N = 4, EC flux differencing, Rusanov surface
  if (dim == 2 && Nq == 5 && form == ATM_FORM_FLUX_DIFFERENCING && vflux == ATM_FLUX_EC &&
      sflux == ATM_FLUX_RUSANOV)
    return make_fd<Euler<2>, 5, ATM_FLUX_EC, ATM_FLUX_RUSANOV>(Nq, Np);
  return resolve_dims<Euler>(dim, form, vflux, Nq, Np);
}
}  // namespace atm
