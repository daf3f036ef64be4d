// Element-kernel instantiations for AdvectionLaw (src/balancelaws/advection.jl).
#include "resolve.hpp"

namespace atm {
Resolved resolve_advection(int dim, int form, int vflux, int sflux, int Nq, int Np) {
  (void)sflux;
  return resolve_dims<Advection>(dim, form, vflux, Nq, Np);
}
}  // namespace atm
