// Element-kernel instantiations for EulerGravityLaw (src/balancelaws/euler_gravity.jl).
#include "resolve.hpp"

namespace atm {
Resolved resolve_euler_gravity(int dim, int form, int vflux, int sflux, int Nq, int Np) {
  using EG3 = EulerGravity<3>;
  if (dim == 3 && form == ATM_FORM_FLUX_DIFFERENCING && vflux == ATM_FLUX_EC) {
    // 3-D EC flux differencing with the Roe surface flux (the rising bubble, N = 4 the
    // metric's configuration) and the Rusanov one (the IMEX explicit operator, N = 3), every
    // degree N = 1..7 specialised
    if (sflux == ATM_FLUX_ROE) {
      switch (Nq) {
        case 2: return make_fd<EG3, 2, ATM_FLUX_EC, ATM_FLUX_ROE>(Nq, Np);
        case 3: return make_fd<EG3, 3, ATM_FLUX_EC, ATM_FLUX_ROE>(Nq, Np);
        case 4: return make_fd<EG3, 4, ATM_FLUX_EC, ATM_FLUX_ROE>(Nq, Np);
        case 5: return make_fd<EG3, 5, ATM_FLUX_EC, ATM_FLUX_ROE>(Nq, Np);
        case 6: return make_fd<EG3, 6, ATM_FLUX_EC, ATM_FLUX_ROE>(Nq, Np);
        case 7: return make_fd<EG3, 7, ATM_FLUX_EC, ATM_FLUX_ROE>(Nq, Np);
        case 8: return make_fd<EG3, 8, ATM_FLUX_EC, ATM_FLUX_ROE>(Nq, Np);
        default: break;
      }
    }
    if (sflux == ATM_FLUX_RUSANOV) {
      switch (Nq) {
        case 2: return make_fd<EG3, 2, ATM_FLUX_EC, ATM_FLUX_RUSANOV>(Nq, Np);
        case 3: return make_fd<EG3, 3, ATM_FLUX_EC, ATM_FLUX_RUSANOV>(Nq, Np);
        case 4: return make_fd<EG3, 4, ATM_FLUX_EC, ATM_FLUX_RUSANOV>(Nq, Np);
        case 5: return make_fd<EG3, 5, ATM_FLUX_EC, ATM_FLUX_RUSANOV>(Nq, Np);
        case 6: return make_fd<EG3, 6, ATM_FLUX_EC, ATM_FLUX_RUSANOV>(Nq, Np);
        case 7: return make_fd<EG3, 7, ATM_FLUX_EC, ATM_FLUX_RUSANOV>(Nq, Np);
        case 8: return make_fd<EG3, 8, ATM_FLUX_EC, ATM_FLUX_RUSANOV>(Nq, Np);
        default: break;
      }
    }
  }
  return resolve_dims<EulerGravity>(dim, form, vflux, Nq, Np);
}
}  // namespace atm
