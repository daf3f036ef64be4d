// Element-kernel instantiations for ShallowWaterLaw (src/balancelaws/shallow_water.jl).
#include "resolve.hpp"

namespace atm {
Resolved resolve_shallow_water(int dim, int form, int vflux, int sflux, int Nq, int Np) {
  using SW2 = ShallowWater<2>;
  // shallow-water ensemble: 2-D, weak form, Rusanov surface, N = 3..7
  if (dim == 2 && form == ATM_FORM_WEAK && sflux == ATM_FLUX_RUSANOV) {
    switch (Nq) {
      case 4: return make_weak<SW2, 4, ATM_FLUX_RUSANOV>(Nq, Np);
      case 5: return make_weak<SW2, 5, ATM_FLUX_RUSANOV>(Nq, Np);
      case 6: return make_weak<SW2, 6, ATM_FLUX_RUSANOV>(Nq, Np);
      case 7: return make_weak<SW2, 7, ATM_FLUX_RUSANOV>(Nq, Np);
      case 8: return make_weak<SW2, 8, ATM_FLUX_RUSANOV>(Nq, Np);
      default: break;
    }
  }
  return resolve_dims<ShallowWater>(dim, form, vflux, Nq, Np);
}
}  // namespace atm
