// Instantiations of the step kernels: euler model, float face arithmetic.
#include "ecg_step_impl.hpp"

namespace ecg {

step_fn get_step_fn_euler_f32(int nph, int dim) {
  if (nph == 1 && dim == 1) return &launch_step_t<M_EULER, 1, 1, float>;
  if (nph == 1 && dim == 2) return &launch_step_t<M_EULER, 1, 2, float>;
  if (nph == 1 && dim == 3) return &launch_step_t<M_EULER, 1, 3, float>;
  return nullptr;
}

fill_fn get_fill_fn_euler_f32(int nph, int dim) {
  if (nph == 1 && dim == 1) return &launch_fill_t<M_EULER, 1, 1, float>;
  if (nph == 1 && dim == 2) return &launch_fill_t<M_EULER, 1, 2, float>;
  if (nph == 1 && dim == 3) return &launch_fill_t<M_EULER, 1, 3, float>;
  return nullptr;
}

}  // namespace ecg
