// Element-kernel instantiations for LinearEulerGravityLaw (src/balancelaws/linear_euler_gravity.jl).
#include "resolve.hpp"

namespace atm {
Resolved resolve_linear_euler_gravity(int dim, int form, int vflux, int sflux, int Nq, int Np) {
  using LE3 = LinearEulerGravity<3>;
  // IMEX implicit (vertical) operator: N = 3, weak form, Rusanov surface
  if (dim == 3 && Nq == 4 && form == ATM_FORM_WEAK && sflux == ATM_FLUX_RUSANOV)
    return make_weak<LE3, 4, ATM_FLUX_RUSANOV>(Nq, Np);
  return resolve_dims<LinearEulerGravity>(dim, form, vflux, Nq, Np);
}
}  // namespace atm
