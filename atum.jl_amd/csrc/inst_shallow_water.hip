// Element-kernel instantiations for ShallowWaterLaw (src/balancelaws/shallow_water.jl).
#include "resolve.hpp"

namespace atm {
Resolved resolve_shallow_water(int dim, int form, int vflux, int sflux, int Nq, int Np) {
  using SW2 = ShallowWater<2>;
  // shallow-water ensemble: 2-D, weak form, Rusanov surface, N = 3..7; batches
  // of these members take one launch per stage whatever their N
  if (dim == 2 && form == ATM_FORM_WEAK && sflux == ATM_FLUX_RUSANOV) {
    Resolved r{nullptr, 0, 0, nullptr, 1};
    switch (Nq) {
      case 4: r = make_weak<SW2, 4, ATM_FLUX_RUSANOV>(Nq, Np); break;
      case 5: r = make_weak<SW2, 5, ATM_FLUX_RUSANOV>(Nq, Np); break;
      case 6: r = make_weak<SW2, 6, ATM_FLUX_RUSANOV>(Nq, Np); break;
      case 7: r = make_weak<SW2, 7, ATM_FLUX_RUSANOV>(Nq, Np); break;
      case 8: r = make_weak<SW2, 8, ATM_FLUX_RUSANOV>(Nq, Np); break;
      default: break;
    }
    if (r.fn) {
      r.mbfn = launch_weak_multi_batch<SW2, ATM_FLUX_RUSANOV>;
      return r;
    }
  }
  return resolve_dims<ShallowWater>(dim, form, vflux, Nq, Np);
}
}  // namespace atm
