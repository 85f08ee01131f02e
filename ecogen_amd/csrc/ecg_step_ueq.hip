// Instantiations of the finite-volume step kernels for the ueq model.
#include "ecg_step_impl.hpp"

namespace ecg {

step_fn get_step_fn_ueq(int nph, int dim) {
  if (nph == 2 && dim == 1) return &launch_step_t<M_UEQ, 2, 1>;
  if (nph == 2 && dim == 2) return &launch_step_t<M_UEQ, 2, 2>;
  if (nph == 2 && dim == 3) return &launch_step_t<M_UEQ, 2, 3>;
  if (nph == 3 && dim == 1) return &launch_step_t<M_UEQ, 3, 1>;
  if (nph == 3 && dim == 2) return &launch_step_t<M_UEQ, 3, 2>;
  if (nph == 3 && dim == 3) return &launch_step_t<M_UEQ, 3, 3>;
  return nullptr;
}

fill_fn get_fill_fn_ueq(int nph, int dim) {
  if (nph == 2 && dim == 1) return &launch_fill_t<M_UEQ, 2, 1>;
  if (nph == 2 && dim == 2) return &launch_fill_t<M_UEQ, 2, 2>;
  if (nph == 2 && dim == 3) return &launch_fill_t<M_UEQ, 2, 3>;
  if (nph == 3 && dim == 1) return &launch_fill_t<M_UEQ, 3, 1>;
  if (nph == 3 && dim == 2) return &launch_fill_t<M_UEQ, 3, 2>;
  if (nph == 3 && dim == 3) return &launch_fill_t<M_UEQ, 3, 3>;
  return nullptr;
}

}  // namespace ecg
