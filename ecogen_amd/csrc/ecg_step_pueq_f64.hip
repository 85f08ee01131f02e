// Instantiations of the step kernels: pueq model, double face arithmetic.
#include "ecg_step_impl.hpp"

namespace ecg {

step_fn get_step_fn_pueq_f64(int nph, int dim) {
  if (nph == 2 && dim == 1) return &launch_step_t<M_PUEQ, 2, 1, double>;
  if (nph == 2 && dim == 2) return &launch_step_t<M_PUEQ, 2, 2, double>;
  if (nph == 2 && dim == 3) return &launch_step_t<M_PUEQ, 2, 3, double>;
  if (nph == 3 && dim == 1) return &launch_step_t<M_PUEQ, 3, 1, double>;
  if (nph == 3 && dim == 2) return &launch_step_t<M_PUEQ, 3, 2, double>;
  if (nph == 3 && dim == 3) return &launch_step_t<M_PUEQ, 3, 3, double>;
  return nullptr;
}

fill_fn get_fill_fn_pueq_f64(int nph, int dim) {
  if (nph == 2 && dim == 1) return &launch_fill_t<M_PUEQ, 2, 1, double>;
  if (nph == 2 && dim == 2) return &launch_fill_t<M_PUEQ, 2, 2, double>;
  if (nph == 2 && dim == 3) return &launch_fill_t<M_PUEQ, 2, 3, double>;
  if (nph == 3 && dim == 1) return &launch_fill_t<M_PUEQ, 3, 1, double>;
  if (nph == 3 && dim == 2) return &launch_fill_t<M_PUEQ, 3, 2, double>;
  if (nph == 3 && dim == 3) return &launch_fill_t<M_PUEQ, 3, 3, double>;
  return nullptr;
}

}  // namespace ecg
