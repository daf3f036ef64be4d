// Kernel selection: one translation unit per balance law instantiates the
// element kernels; each exports a resolver that maps an operator description
// to a launcher.  Hot configurations get kernels with the LGL order and the
// surface flux fixed at compile time; everything else uses the generic
// kernels (runtime Nq, runtime surface-flux switch).
#pragma once
#include "fd_fast.hpp"

namespace atm {

Resolved resolve_advection(int dim, int form, int vflux, int sflux, int Nq, int Np);
Resolved resolve_euler(int dim, int form, int vflux, int sflux, int Nq, int Np);
Resolved resolve_euler_gravity(int dim, int form, int vflux, int sflux, int Nq, int Np);
Resolved resolve_linear_euler_gravity(int dim, int form, int vflux, int sflux, int Nq, int Np);
Resolved resolve_shallow_water(int dim, int form, int vflux, int sflux, int Nq, int Np);
// pair-symmetric EC flux-differencing kernels (ATM_PRECISION_FAST); fn == nullptr if none
Resolved resolve_fast_euler(int dim, int sflux, int Nq);
Resolved resolve_fast_euler_gravity(int dim, int sflux, int Nq);

template <class L, int NQ, int SF>
Resolved fast_variant() {
  Resolved r = make_resolved(launch_fd_fast<L, NQ, SF>, launch_fd_fast_batch<L, NQ, SF>, L::Ns,
                             fast_epw<L, NQ, SF>(), sizeof(double) * NQ * NQ, fast_lds_elem_bytes<L, NQ>(),
                             fast_max_threads<L, NQ, SF>());
  // 2-D and 3-D members of N = 3..7 with a compiled-in surface flux share one batched launch
  if constexpr (L::D >= 2 && SF >= 0 && NQ >= 4 && NQ <= 8) r.mbfn = launch_fd_fast_multi_batch<L, SF>;
  return r;
}

template <class L, int NQ>
Resolved fast_entry(int sflux) {
  if (sflux == ATM_FLUX_ROE) return fast_variant<L, NQ, ATM_FLUX_ROE>();
  if (sflux == ATM_FLUX_RUSANOV) return fast_variant<L, NQ, ATM_FLUX_RUSANOV>();
  return fast_variant<L, NQ, -1>();
}

template <template <int> class LAW, int D>
Resolved fast_nq(int sflux, int Nq) {
  switch (Nq) {
    case 2: return fast_entry<LAW<D>, 2>(sflux);
    case 3: return fast_entry<LAW<D>, 3>(sflux);
    case 4: return fast_entry<LAW<D>, 4>(sflux);
    case 5: return fast_entry<LAW<D>, 5>(sflux);
    case 6: return fast_entry<LAW<D>, 6>(sflux);
    case 7: return fast_entry<LAW<D>, 7>(sflux);
    case 8: return fast_entry<LAW<D>, 8>(sflux);
    default: return Resolved{nullptr, 0, 0, nullptr, 1};
  }
}

template <template <int> class LAW>
Resolved resolve_dims(int dim, int form, int vflux, int Nq, int Np) {
  if (dim == 1) return resolve_generic<LAW<1>>(form, vflux, Nq, Np);
  if (dim == 2) return resolve_generic<LAW<2>>(form, vflux, Nq, Np);
  if (dim == 3) return resolve_generic<LAW<3>>(form, vflux, Nq, Np);
  return Resolved{nullptr, 0, 0, nullptr, 1};
}

template <class L>
bool sflux_supported(int sf) {
  switch (sf) {
    case ATM_FLUX_CENTRAL:
    case ATM_FLUX_RUSANOV:
      return L::has_sflux(sf);
    case ATM_FLUX_EC:
      return L::has_sflux(sf) && L::has_vflux(ATM_FLUX_EC);
    case ATM_FLUX_KG:
      return L::has_sflux(sf) && L::has_vflux(ATM_FLUX_KG);
    case ATM_FLUX_ROE:
      return L::LAW == ATM_LAW_EULER || L::LAW == ATM_LAW_EULER_GRAVITY || L::LAW == ATM_LAW_SHALLOW_WATER;
    case ATM_FLUX_MATRIX:
      return L::LAW == ATM_LAW_EULER || L::LAW == ATM_LAW_EULER_GRAVITY;
    default:
      return false;
  }
}

}  // namespace atm
