// Kernel entry points of the finite-volume step (gfx950), shared by the
// per-model translation units and the C ABI layer.
#pragma once

#include "ecg_physics.hpp"

namespace ecg {

// Domain description on the device (Cell::fill, GeometricalDomain.cpp:64-83).
struct DomainDev {
  int kind, axis, axis2, dir;
  double origin;
  double center[3];
  double radius;
  double lo[3], len[3];
  double alpha[4], rho[4], p[4];
  double vel[3];
};

struct StepLaunch {
  const double* wn;     // state at time n
  const double* wsrc;   // state the fluxes are evaluated on (== wn for first order / first stage)
  double* wout;         // written state
  double scale_half;    // 1: the update uses 0.5*dt (second-order predictor)
  int stage;            // 0: first order; 1: predictor; 2: corrector
};

// Per (model, phases, dimension) launchers; defined in ecg_step_<model>.hip.
typedef void (*step_fn)(const Params& P, const StepLaunch& L, Clock* clk, Counters* cnt, hipStream_t s,
                        int blocks, int threads);
typedef void (*fill_fn)(const Params& P, const DomainDev* doms, int ndom, double* w, hipStream_t s, int blocks,
                        int threads);

step_fn get_step_fn(int model, int nph, int dim, int order);
fill_fn get_fill_fn(int model, int nph, int dim);

void launch_clock(Clock* clk, double cfl, hipStream_t s);

}  // namespace ecg
