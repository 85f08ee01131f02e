// Instantiations of the step kernels: euler model, double face arithmetic.
#include "ecg_step_impl.hpp"

namespace ecg {

step_fn get_step_fn_euler_f64(int nph, int dim) {
  if (nph == 1 && dim == 1) return &launch_step_t<M_EULER, 1, 1, double>;
  if (nph == 1 && dim == 2) return &launch_step_t<M_EULER, 1, 2, double>;
  if (nph == 1 && dim == 3) return &launch_step_t<M_EULER, 1, 3, double>;
  return nullptr;
}

fill_fn get_fill_fn_euler_f64(int nph, int dim) {
  if (nph == 1 && dim == 1) return &launch_fill_t<M_EULER, 1, 1, double>;
  if (nph == 1 && dim == 2) return &launch_fill_t<M_EULER, 1, 2, double>;
  if (nph == 1 && dim == 3) return &launch_fill_t<M_EULER, 1, 3, double>;
  return nullptr;
}

}  // namespace ecg
