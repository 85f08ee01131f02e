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

// Buffers of one launch; the element type is the state type of the precision form.
struct StepLaunch {
  const void* wn;       // state at time n
  const void* wsrc;     // state the fluxes are evaluated on (== wn for first order / predictor)
  void* wout;           // written state
  int stage;            // 0: first order; 1: predictor; 2: corrector
};

template <typename R>
struct StepLaunchT {
  const R* wn;
  const R* wsrc;
  R* wout;
  int stage;
};

// Per (model, phases, dimension, precision) launchers; defined in ecg_step_<model>_<precision>.hip.
typedef void (*step_fn)(const Params& P, const StepLaunch& L, Clock* clk, Counters* cnt, hipStream_t s);
typedef void (*fill_fn)(const Params& P, const DomainDev* doms, int ndom, void* w, hipStream_t s);

// Tile extents (cells per block along x, y, z) for a problem dimension.
void tile_extent(int dim, int t[3]);

void launch_clock(Clock* clk, double cfl, hipStream_t s);

}  // namespace ecg
