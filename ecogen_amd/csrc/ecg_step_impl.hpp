// Kernel templates of the finite-volume step.  One thread owns one cell.
//
// First order (Run::solveHyperbolic, Run.cpp:696-716, then solveRelaxations,
// Run.cpp:771-804) is a single kernel: the cell gathers its 2*D face fluxes in
// face-creation order (X faces, Y faces, Z faces, then boundary faces X-, X+,
// Y-, Y+, Z-, Z+; MeshCartesian.cpp:283-557), updates, relaxes, corrects and
// writes the new persistent state.
//
// Second order (Run::solveHyperbolicO2, Run.cpp:627-692) is two launches of the
// same kernel: the predictor evaluates MUSCL fluxes on W^n and writes the
// half-step state W* (CellO2::predictionOrdre2, CellO2.cpp:111-135); the
// corrector evaluates them on W* and updates W^n to W^{n+1}.
#pragma once

#include "ecg_kernels.hpp"

namespace ecg {

// Persistent variable layout (see include/ecogen_amd.h).
template <int MODEL, int N>
struct Layout {
  static constexpr int phase_vars = (MODEL == M_EULER) ? 2 : 3 * N;
  template <int D> static constexpr int nvar() { return phase_vars + D + 1; }
};

template <int MODEL, int N, int D>
__device__ __forceinline__ void load_prim(const Params& P, const double* __restrict__ W, int64_t c, Cell<N>& s) {
  const int64_t ld = P.pitch;
  int v = 0;
  if (MODEL == M_EULER) {
    s.a[0] = 1.;
    s.r[0] = W[(v++) * ld + c];
    s.p[0] = W[(v++) * ld + c];
  } else {
#pragma unroll
    for (int k = 0; k < N; k++) {
      s.a[k] = W[(v++) * ld + c];
      s.r[k] = W[(v++) * ld + c];
      s.p[k] = W[(v++) * ld + c];
    }
  }
  s.u[0] = W[(v++) * ld + c];
  s.u[1] = (D >= 2) ? W[(v++) * ld + c] : 0.;
  s.u[2] = (D >= 3) ? W[(v++) * ld + c] : 0.;
  s.E = W[v * ld + c];
}

template <int MODEL, int N, int D>
__device__ __forceinline__ void store_prim(const Params& P, double* __restrict__ W, int64_t c, const Cell<N>& s) {
  const int64_t ld = P.pitch;
  int v = 0;
  if (MODEL == M_EULER) {
    W[(v++) * ld + c] = s.r[0];
    W[(v++) * ld + c] = s.p[0];
  } else {
#pragma unroll
    for (int k = 0; k < N; k++) {
      W[(v++) * ld + c] = s.a[k];
      W[(v++) * ld + c] = s.r[k];
      W[(v++) * ld + c] = s.p[k];
    }
  }
  W[(v++) * ld + c] = s.u[0];
  if (D >= 2) W[(v++) * ld + c] = s.u[1];
  if (D >= 3) W[(v++) * ld + c] = s.u[2];
  W[v * ld + c] = s.E;
}

// alpha of phase 0 at time n (PhaseEuler::getAlpha returns 0, PhaseEuler.h:110)
template <int MODEL>
__device__ __forceinline__ double alpha0(const Params& P, const double* __restrict__ W, int64_t c) {
  return (MODEL == M_EULER) ? 0. : W[c];
}

// Interior face slope variables: phase (alpha, rho, p) per phase, velocity (3).
template <int N>
struct Slope {
  double a[N], r[N], p[N];
  double u[3];
};

template <int MODEL, int N, int D>
__device__ __forceinline__ void face_slope(const Cell<N>& L, const Cell<N>& R, double dist, Slope<N>& s) {
  // CellInterfaceO2Cartesian::computeSlopes (CellInterfaceO2Cartesian.cpp:77-99)
#pragma unroll
  for (int k = 0; k < N; k++) {
    if (MODEL != M_EULER) s.a[k] = (R.a[k] - L.a[k]) / dist;
    s.r[k] = (R.r[k] - L.r[k]) / dist;
    s.p[k] = (R.p[k] - L.p[k]) / dist;
  }
#pragma unroll
  for (int d = 0; d < 3; d++) s.u[d] = (d < D) ? (R.u[d] - L.u[d]) / dist : 0.;
}

// CellO2Cartesian::computeLocalSlopes (CellO2Cartesian.cpp:23-171) on a uniform
// Cartesian cell: slope1 is the face's own slope, slope2 the opposite face's
// slope when that face is interior (boundary faces carry none: slope2 = 0 and
// the opposite alpha stays 0); then limiting and extrapolation by dist with the
// sign sgn (+1 left cell, -1 right cell: changeSign, CellInterfaceO2Cartesian.cpp:151-156).
template <int MODEL, int N, int D>
__device__ __forceinline__ void extrapolate(const Params& P, Cell<N>& b, const Slope<N>& s1, const Slope<N>& s2,
                                            bool has2, double aAfter, double aCell, double aOther, double dist,
                                            bool negate) {
  int lg, lvf;
  if ((aCell >= 1.e-4) && (aCell <= 1. - 1.e-4) && ((aOther - aCell) * (aCell - aAfter) >= 1.e-8)) {
    lg = P.lim[1]; lvf = P.lim[3];
  } else {
    lg = P.lim[0]; lvf = P.lim[2];
  }
#pragma unroll
  for (int k = 0; k < N; k++) {
    double la = 0., lr, lp;
    if (MODEL != M_EULER) la = limit(lvf, s1.a[k], has2 ? s2.a[k] : 0.);
    lr = limit(lg, s1.r[k], has2 ? s2.r[k] : 0.);
    lp = limit(lg, s1.p[k], has2 ? s2.p[k] : 0.);
    if (negate) { la = -la; lr = -lr; lp = -lp; }
    if (MODEL != M_EULER) b.a[k] += la * dist;
    b.r[k] += lr * dist;
    b.p[k] += lp * dist;
    if (MODEL == M_EULER) {
#pragma unroll
      for (int d = 0; d < D; d++) {
        double lu = limit(lg, s1.u[d], has2 ? s2.u[d] : 0.);
        if (negate) lu = lu * -1.;
        b.u[d] = b.u[d] + lu * dist;
      }
    }
    verify_correct<MODEL, N>(P, b, k);
  }
  if (MODEL != M_EULER) {
#pragma unroll
    for (int d = 0; d < D; d++) {
      double lu = limit(lg, s1.u[d], has2 ? s2.u[d] : 0.);
      if (negate) lu = lu * -1.;
      b.u[d] = b.u[d] + lu * dist;
    }
  }
}

// Extrapolated left or right state of the interior face (x | x+1) along an axis
// (CellInterfaceO2Cartesian::solveRiemann, CellInterfaceO2Cartesian.cpp:101-202).
// m: axis index of cell x; n: cells along the axis; st: index stride.
template <int MODEL, int N, int D, bool RIGHT>
__device__ __forceinline__ void muscl_state(const Params& P, const double* __restrict__ wsrc,
                                            const double* __restrict__ wn, int64_t x, int m, int n, int64_t st,
                                            int axis, Cell<N>& b) {
  Cell<N> cL, cR, cO;
  load_prim<MODEL, N, D>(P, wsrc, x, cL);
  load_prim<MODEL, N, D>(P, wsrc, x + st, cR);
  Slope<N> s1, s2;
  face_slope<MODEL, N, D>(cL, cR, P.dcc[axis][m], s1);
  const double aL = (MODEL == M_EULER) ? 0. : cL.a[0];
  const double aR = (MODEL == M_EULER) ? 0. : cR.a[0];
  if (!RIGHT) {
    const bool has2 = (m >= 1);
    double aLL = 0.;
    if (has2) {
      load_prim<MODEL, N, D>(P, wsrc, x - st, cO);
      face_slope<MODEL, N, D>(cO, cL, P.dcc[axis][m - 1], s2);
      aLL = alpha0<MODEL>(P, wn, x - st);
    }
    b = cL;
    extrapolate<MODEL, N, D>(P, b, s1, s2, has2, aLL, aL, aR, P.dfl[axis][m], false);
  } else {
    const bool has2 = (m + 1 <= n - 2);
    double aRR = 0.;
    if (has2) {
      load_prim<MODEL, N, D>(P, wsrc, x + 2 * st, cO);
      face_slope<MODEL, N, D>(cR, cO, P.dcc[axis][m + 1], s2);
      aRR = alpha0<MODEL>(P, wn, x + 2 * st);
    }
    b = cR;
    extrapolate<MODEL, N, D>(P, b, s1, s2, has2, aRR, aR, aL, P.dfr[axis][m], true);
  }
  derive<MODEL, N>(P, b, MODEL == M_EULER);
}

template <int MODEL, int N, int D, bool O2, int FR>
__device__ __forceinline__ void interior_pair(const Params& P, const StepLaunch& L, const Cell<N>& own, int64_t c,
                                              int m, int n, int64_t st, int axis, double coef, const double* an,
                                              const double* pn, double& dtm, Cons<N>& acc) {
  Flux<N> F;
  if (!O2) {
    // First order: the stored cells (derived quantities rebuilt identically).
    if (m > 0) {
      Cell<N> nb;
      load_prim<MODEL, N, D>(P, L.wsrc, c - st, nb);
      derive<MODEL, N>(P, nb, false);
      face_flux<MODEL, N, FR>(P, nb, nb.E, own, own.E, dtm, F);
      add_flux<MODEL, N>(acc, F, coef, an, pn);
    }
    if (m < n - 1) {
      Cell<N> nb;
      load_prim<MODEL, N, D>(P, L.wsrc, c + st, nb);
      derive<MODEL, N>(P, nb, false);
      face_flux<MODEL, N, FR>(P, own, own.E, nb, nb.E, dtm, F);
      sub_flux<MODEL, N>(acc, F, coef, an, pn);
    }
  } else {
    if (m > 0) {
      Cell<N> bl, br;
      muscl_state<MODEL, N, D, false>(P, L.wsrc, L.wn, c - st, m - 1, n, st, axis, bl);
      muscl_state<MODEL, N, D, true>(P, L.wsrc, L.wn, c - st, m - 1, n, st, axis, br);
      face_flux<MODEL, N, FR>(P, bl, bl.E, br, br.E, dtm, F);
      add_flux<MODEL, N>(acc, F, coef, an, pn);
    }
    if (m < n - 1) {
      Cell<N> bl, br;
      muscl_state<MODEL, N, D, false>(P, L.wsrc, L.wn, c, m, n, st, axis, bl);
      muscl_state<MODEL, N, D, true>(P, L.wsrc, L.wn, c, m, n, st, axis, br);
      face_flux<MODEL, N, FR>(P, bl, bl.E, br, br.E, dtm, F);
      sub_flux<MODEL, N>(acc, F, coef, an, pn);
    }
  }
}

template <int MODEL, int N, int FR>
__device__ __forceinline__ void boundary_face(const Params& P, const Cell<N>& b, int bc, double coef,
                                              const double* an, const double* pn, double& dtm, Cons<N>& acc) {
  Flux<N> F;
  const double EB = (MODEL == M_EULER) ? b.e[0] + 0.5 * sqn3(b.u) : b.E;
  boundary_flux<MODEL, N, FR>(P, b, EB, bc, dtm, F);
  sub_flux<MODEL, N>(acc, F, coef, an, pn);
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = dmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int MODEL, int N, int D, bool O2>
__global__ __launch_bounds__(256) void k_step(Params P, StepLaunch L, Clock* __restrict__ clk,
                                              Counters* __restrict__ cnt) {
  __shared__ double s_dt[4];
  __shared__ unsigned long long s_c[3][4];
  if (clk->done) return;
  const double dt = clk->dt;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double dtm = 1.e10;
  unsigned long long n_it = 0, n_fail = 0, n_relax = 0;
  if (c < P.ncell) {
    const int i = (int)(c % P.nx);
    const int j = (D >= 2) ? (int)((c / P.nx) % P.ny) : 0;
    const int k = (D >= 3) ? (int)(c / ((int64_t)P.nx * P.ny)) : 0;
    // Time-n cell: non-conservative terms and U^n (Cell::getPhase default type).
    Cell<N> own;
    load_prim<MODEL, N, D>(P, L.wn, c, own);
    derive<MODEL, N>(P, own, false);
    // Cell the fluxes see: W^n (first order / predictor) or W* (corrector).
    Cell<N> src;
    if (L.stage == 2) {
      load_prim<MODEL, N, D>(P, L.wsrc, c, src);
      derive<MODEL, N>(P, src, false);
    } else {
      src = own;
    }
    double an[N], pn[N];
#pragma unroll
    for (int q = 0; q < N; q++) { an[q] = own.a[q]; pn[q] = own.p[q]; }
    Cons<N> acc;
#pragma unroll
    for (int q = 0; q < N; q++) { acc.alpha[q] = 0.; acc.mass[q] = 0.; acc.energ[q] = 0.; }
    acc.mom[0] = 0.; acc.mom[1] = 0.; acc.mom[2] = 0.; acc.energM = 0.;

    interior_pair<MODEL, N, D, O2, FR_XI>(P, L, src, c, i, P.nx, 1, 0, P.coef[0], an, pn, dtm, acc);
    if (D >= 2) interior_pair<MODEL, N, D, O2, FR_YI>(P, L, src, c, j, P.ny, P.nx, 1, P.coef[1], an, pn, dtm, acc);
    if (D >= 3)
      interior_pair<MODEL, N, D, O2, FR_ZI>(P, L, src, c, k, P.nz, (int64_t)P.nx * P.ny, 2, P.coef[2], an, pn, dtm, acc);
    if (i == 0) boundary_face<MODEL, N, FR_XM>(P, src, P.bc[0], P.coef[0], an, pn, dtm, acc);
    if (i == P.nx - 1) boundary_face<MODEL, N, FR_XP>(P, src, P.bc[1], P.coef[0], an, pn, dtm, acc);
    if (D >= 2) {
      if (j == 0) boundary_face<MODEL, N, FR_YM>(P, src, P.bc[2], P.coef[1], an, pn, dtm, acc);
      if (j == P.ny - 1) boundary_face<MODEL, N, FR_YP>(P, src, P.bc[3], P.coef[1], an, pn, dtm, acc);
    }
    if (D >= 3) {
      if (k == 0) boundary_face<MODEL, N, FR_ZM>(P, src, P.bc[4], P.coef[2], an, pn, dtm, acc);
      if (k == P.nz - 1) boundary_face<MODEL, N, FR_ZP>(P, src, P.bc[5], P.coef[2], an, pn, dtm, acc);
    }

    const double scale = (L.stage == 1) ? 0.5 * dt : dt;
    time_evolution<MODEL, N>(P, acc, own, scale);
    Cell<N> nw;
    build_prim<MODEL, N>(P, acc, nw);
    // Relaxation: after the predictor (CellO2.cpp:128-134) and after the full step
    // (Run::advancingProcedure, Run.cpp:593).
    if (MODEL != M_EULER && P.relax) {
      const int it = relax_newton<N>(P, nw, relax_guess<N>(nw));
      if (it > 0) { n_relax = 1; n_it = (unsigned long long)it; n_fail = (it >= 100) ? 1 : 0; }
      correct_and_derive<MODEL, N>(P, nw);
    } else if (L.stage == 1) {
      correct_and_derive<MODEL, N>(P, nw);
    }
    store_prim<MODEL, N, D>(P, L.wout, c, nw);
  }
  // Block reductions: CFL minimum and relaxation counters.
  dtm = wave_min(dtm);
  n_it = wave_sum(n_it);
  n_fail = wave_sum(n_fail);
  n_relax = wave_sum(n_relax);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) { s_dt[wv] = dtm; s_c[0][wv] = n_it; s_c[1][wv] = n_fail; s_c[2][wv] = n_relax; }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    double m = s_dt[0];
    unsigned long long a = s_c[0][0], b = s_c[1][0], r = s_c[2][0];
    for (int w = 1; w < nw; w++) { m = dmin(m, s_dt[w]); a += s_c[0][w]; b += s_c[1][w]; r += s_c[2][w]; }
    atomicMin(&clk->dtmax_bits, (unsigned long long)__double_as_longlong(m));
    if (r) {
      atomicAdd(&cnt->newton_iters, a);
      atomicAdd(&cnt->newton_fail, b);
      atomicAdd(&cnt->relax_calls, r);
    }
  }
}

// Initial condition: Cell::fill over the domains (later ones override), then
// Cell::completeFulfillState (Cell.cpp:344-353) and E = e + |u|^2/2.
template <int MODEL, int N, int D>
__global__ __launch_bounds__(256) void k_fill(Params P, const DomainDev* __restrict__ doms, int ndom,
                                              double* __restrict__ W) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.ncell) return;
  const int i = (int)(c % P.nx);
  const int j = (int)((c / P.nx) % P.ny);
  const int k = (int)(c / ((int64_t)P.nx * P.ny));
  double x[3];
  x[0] = (i + 0.5) * P.d[0];
  x[1] = (j + 0.5) * P.d[1];
  x[2] = (k + 0.5) * P.d[2];
  Cell<N> s;
#pragma unroll
  for (int q = 0; q < N; q++) { s.a[q] = 0.; s.r[q] = 0.; s.p[q] = 0.; }
  s.u[0] = s.u[1] = s.u[2] = 0.;
  // The last domain containing the cell wins: scan backwards and stop at the first hit.
  for (int q = ndom - 1; q >= 0; q--) {
    const DomainDev& d = doms[q];
    bool in = false;
    if (d.kind == 0) {
      in = true;
    } else if (d.kind == 1) {
      in = (d.dir >= 0) ? (x[d.axis] >= d.origin) : (x[d.axis] <= d.origin);
    } else if (d.kind == 2) {
      const double a = x[d.axis] - d.center[d.axis], b = x[d.axis2] - d.center[d.axis2];
      double sum = 0.;
      sum += a * a;
      sum += b * b;
      in = (sum <= d.radius * d.radius);
    } else if (d.kind == 3) {
      const double a = x[0] - d.center[0], b = x[1] - d.center[1], e = x[2] - d.center[2];
      const double sum = a * a + b * b + e * e;
      in = (sum <= d.radius * d.radius);
    } else if (d.kind == 4) {
      in = !((x[d.axis] - d.lo[d.axis] < 0) || (x[d.axis] - d.lo[d.axis] > d.len[d.axis]) ||
             (x[d.axis2] - d.lo[d.axis2] < 0) || (x[d.axis2] - d.lo[d.axis2] > d.len[d.axis2]));
    } else if (d.kind == 5) {
      in = true;
      for (int a = 0; a < 3; a++)
        if ((x[a] - d.lo[a] < 0) || (x[a] - d.lo[a] > d.len[a])) in = false;
    }
    if (in) {
#pragma unroll
      for (int q2 = 0; q2 < N; q2++) {
        s.a[q2] = (MODEL == M_EULER) ? 1. : d.alpha[q2];
        s.r[q2] = d.rho[q2];
        s.p[q2] = d.p[q2];
      }
      s.u[0] = d.vel[0];
      s.u[1] = (D >= 2) ? d.vel[1] : 0.;
      s.u[2] = (D >= 3) ? d.vel[2] : 0.;
      break;
    }
  }
  derive<MODEL, N>(P, s, MODEL == M_EULER);
  if (MODEL != M_EULER) s.E = s.em + 0.5 * sqn3(s.u);
  store_prim<MODEL, N, D>(P, W, c, s);
}

template <int MODEL, int N, int D>
void launch_step_t(const Params& P, const StepLaunch& L, Clock* clk, Counters* cnt, hipStream_t s, int blocks,
                   int threads) {
  if (L.stage == 0) hipLaunchKernelGGL((k_step<MODEL, N, D, false>), dim3(blocks), dim3(threads), 0, s, P, L, clk, cnt);
  else hipLaunchKernelGGL((k_step<MODEL, N, D, true>), dim3(blocks), dim3(threads), 0, s, P, L, clk, cnt);
}

template <int MODEL, int N, int D>
void launch_fill_t(const Params& P, const DomainDev* doms, int ndom, double* w, hipStream_t s, int blocks,
                   int threads) {
  hipLaunchKernelGGL((k_fill<MODEL, N, D>), dim3(blocks), dim3(threads), 0, s, P, doms, ndom, w);
}

}  // namespace ecg
