// Device physics of the explicit finite-volume step (gfx950).
//
// Per-cell register formulation: a thread owns one cell, gathers the states of
// its face neighbours, evaluates every face flux of the cell in the reference's
// face-creation order and performs the cell update, relaxation and energy
// correction in registers.  Every expression keeps the association order of
// the reference expression it evaluates (built with -ffp-contract=off) so that
// the FP64 path reproduces the reference's CPU arithmetic.
//
// Reference files (paths relative to the reference src/):
//   EOS             Eos/EosSG.cpp:66-257, Eos/EosIG.cpp:64-255
//   HLLC            Models/Euler/ModEuler.cpp:105-221, Models/UEq/ModUEq.cpp:113-268
//   wall            Models/Euler/ModEuler.cpp:368-409, Models/UEq/ModUEq.cpp:452-493
//   cons/prim       Models/Euler/FluxEuler.cpp:113-150, Models/UEq/FluxUEq.cpp:147-207
//   non-conservative Models/UEq/FluxUEq.cpp:224-244, Models/PUEq/FluxPUEq.cpp:44-65
//   scheme corr.    Models/UEq/FluxUEq.cpp:248-269 (VELOCITYEQ only)
//   energy corr.    Models/PUEq/FluxPUEq.cpp:68-116
//   relaxation      Relaxations/RelaxationPInfinite.cpp:43-98, Relaxations/RelaxationP.cpp:43-93
//   limiters        Order2/Limiter*.cpp
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ecg {

enum { M_EULER = 0, M_PUEQ = 1, M_UEQ = 2 };
enum { LIM_NONE = 0, LIM_MINMOD = 1, LIM_VANLEER = 2, LIM_VANALBADA = 3, LIM_SUPERBEE = 4, LIM_MC = 5 };
enum { BC_NONREFLECTING = 0, BC_WALL = 1 };

struct EosDev {
  int sg;          // 1: stiffened gas, 0: ideal gas
  double g;        // gamma
  double pinf;     // SG p_infinity (0 for IG)
  double gm1;      // gamma - 1
  double gpinf;    // gamma * p_infinity
  double eref;
  double mixg;     // gamma*pinf/(gamma-1) (IG: 0)   EosSG.cpp:208-214
  double mixinv;   // 1/(gamma-1)
  double pclamp;   // SG pressure floor -(1-1e-15)*pinf + 1e-15
};

struct Params {
  int nx, ny, nz;
  int64_t ncell;
  int64_t pitch;            // elements between two SoA variables
  double d[3];              // cell sizes
  double coef[3];           // face surface / cell volume per axis
  double lcfl;              // CFL length
  double eps;               // epsilonAlphaNull
  double cfl;
  int alpha_null;
  int relax;                // pressure relaxation active
  int bc[6];
  int lim[4];
  EosDev eos[4];
  const double* dcc[3];     // cell-to-cell distance of face m (cells m | m+1)
  const double* dfl[3];     // face m to its left cell centre
  const double* dfr[3];     // face m to its right cell centre
};

// Device clock: lagged time step kept on the GPU (Run.cpp:385-498).
struct Clock {
  double t, dt, dt_next, dt_max;
  double t_end;
  int64_t iteration, max_iter;
  uint64_t dtmax_bits;      // atomicMin target (non-negative doubles order as integers)
  int done;
  int pad;
};

struct Counters {
  unsigned long long newton_fail, newton_iters, relax_calls;
};

// std::min / std::max tie semantics.
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }

// ---------------------------------------------------------------------------
// EOS
// ---------------------------------------------------------------------------
__device__ __forceinline__ double eos_energy(const EosDev& q, double eps, double rho, double p) {
  return (q.sg ? (p + q.gpinf) : p) / q.gm1 / dmax(rho, eps) + q.eref;
}
__device__ __forceinline__ double eos_pressure(const EosDev& q, double rho, double e) {
  double v = q.gm1 * rho * (e - q.eref);
  return q.sg ? v - q.gpinf : v;
}
__device__ __forceinline__ double eos_sound(const EosDev& q, double eps, double rho, double p) {
  return sqrt(q.g * (q.sg ? (p + q.pinf) : p) / dmax(rho, eps));
}
__device__ __forceinline__ double eos_p_isentropic(const EosDev& q, double eps, double p0, double rho0, double rhof) {
  double base = pow(rhof / dmax(rho0, eps), q.g);
  return q.sg ? (p0 + q.pinf) * base - q.pinf : p0 * base;
}
__device__ __forceinline__ double eos_clamp_p(const EosDev& q, double p) {
  if (q.sg) return (p <= q.pclamp) ? q.pclamp : p;
  return (p < 1.e-8) ? 1.e-8 : p;
}
// Tools::returnNonZeroValue (Tools.cpp:134-145)
__device__ __forceinline__ double nonzero(double a, double eps) {
  if (a > eps || a < -eps) return a;
  return (double)((0. < a) - (a < 0.)) * eps;
}
// computeDensityPfinal (EosSG.cpp:163-170, EosIG.cpp:161-168)
__device__ __forceinline__ double eos_rho_pfinal(const EosDev& q, double eps, double p0, double rho0, double pf,
                                                 double& drhodp) {
  double num = q.sg ? q.g * (pf + q.pinf) : q.g * pf;
  double denom = num + p0 - pf;
  double rhof = rho0 * num / (q.sg ? nonzero(denom, eps) : dmax(denom, eps));
  drhodp = rho0 * q.g * (q.sg ? (p0 + q.pinf) : p0) / dmax(denom * denom, eps);
  return rhof;
}

// ---------------------------------------------------------------------------
// Limiters (Order2/LimiterMinmod.cpp:46-56, LimiterVanLeer.cpp:43-49,
// LimiterVanAlbada.cpp:43-51, LimiterSuperBee.cpp:46-55, LimiterMC.cpp:43-53)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double limit(int kind, double s1, double s2) {
  double prod = s1 * s2, out = 0.;
  if (kind == LIM_MINMOD) {
    if (prod > 1e-9) { out = dmin(fabs(s1), fabs(s2)); if (s1 < 0) out = -out; }
  } else if (kind == LIM_VANLEER) {
    double sum = s1 + s2;
    if ((fabs(sum) > 1e-6) && (prod > 1e-6)) out = 2. * prod / sum;
  } else if (kind == LIM_VANALBADA) {
    double sum = s1 + s2;
    if ((fabs(s1) > 1e-6) && (fabs(s2) > 1e-6) && (fabs(sum) > 1e-6) && (prod > 1e-6))
      out = prod * sum / (s1 * s1 + s2 * s2);
  } else if (kind == LIM_SUPERBEE) {
    if (prod > 1e-6) out = dmax(dmin(2. * fabs(s1), fabs(s2)), dmin(fabs(s1), 2. * fabs(s2)));
    if (s1 < 0.) out = -out;
  } else if (kind == LIM_MC) {
    if (prod > 1e-9) { out = dmin(2. * fabs(s1), 2. * fabs(s2)); out = dmin(out, 0.5 * (fabs(s1) + fabs(s2))); }
    if (s1 < 0.) out = -out;
  }
  return out;
}

// ---------------------------------------------------------------------------
// State
// ---------------------------------------------------------------------------
// Persistent primitives plus derived quantities of one cell.  For Euler only
// index 0 of the phase arrays is used (a[0] unused).
template <int N>
struct Cell {
  double a[N], r[N], p[N];
  double e[N], c[N];
  double u[3];
  double rho, pm, em, cF;
  double E;
};

template <int N>
struct Cons {
  double alpha[N], mass[N], energ[N];
  double mom[3];
  double energM;
};

template <int N>
struct Flux {
  Cons<N> f;
  double sM, uStar;
};

// Face frames: local (n, t, b) components are a signed permutation of the
// global velocity (MeshCartesian.cpp:282, 315, 348, 387, 411, 440, 482, 511, 535).
enum { FR_XI = 0, FR_YI, FR_ZI, FR_XM, FR_XP, FR_YM, FR_YP, FR_ZM, FR_ZP };
template <int FR> struct Frame;
template <> struct Frame<FR_XI> { static constexpr int i0 = 0, i1 = 1, i2 = 2; static constexpr double s0 = 1, s1 = 1, s2 = 1; };
template <> struct Frame<FR_XP> { static constexpr int i0 = 0, i1 = 1, i2 = 2; static constexpr double s0 = 1, s1 = 1, s2 = 1; };
template <> struct Frame<FR_XM> { static constexpr int i0 = 0, i1 = 1, i2 = 2; static constexpr double s0 = -1, s1 = -1, s2 = 1; };
template <> struct Frame<FR_YI> { static constexpr int i0 = 1, i1 = 0, i2 = 2; static constexpr double s0 = 1, s1 = -1, s2 = 1; };
template <> struct Frame<FR_YP> { static constexpr int i0 = 1, i1 = 0, i2 = 2; static constexpr double s0 = 1, s1 = -1, s2 = 1; };
template <> struct Frame<FR_YM> { static constexpr int i0 = 1, i1 = 0, i2 = 2; static constexpr double s0 = -1, s1 = 1, s2 = 1; };
template <> struct Frame<FR_ZI> { static constexpr int i0 = 2, i1 = 0, i2 = 1; static constexpr double s0 = 1, s1 = 1, s2 = 1; };
template <> struct Frame<FR_ZP> { static constexpr int i0 = 2, i1 = 0, i2 = 1; static constexpr double s0 = 1, s1 = 1, s2 = 1; };
template <> struct Frame<FR_ZM> { static constexpr int i0 = 2, i1 = 0, i2 = 1; static constexpr double s0 = -1, s1 = -1, s2 = 1; };

template <int FR>
__device__ __forceinline__ void to_local(const double* u, double* l) {
  using F = Frame<FR>;
  l[0] = F::s0 < 0 ? -u[F::i0] : u[F::i0];
  l[1] = F::s1 < 0 ? -u[F::i1] : u[F::i1];
  l[2] = F::s2 < 0 ? -u[F::i2] : u[F::i2];
}
template <int FR>
__device__ __forceinline__ void to_global(const double* l, double* g) {
  using F = Frame<FR>;
  g[F::i0] = F::s0 < 0 ? -l[0] : l[0];
  g[F::i1] = F::s1 < 0 ? -l[1] : l[1];
  g[F::i2] = F::s2 < 0 ? -l[2] : l[2];
}

__device__ __forceinline__ double sqn3(const double* u) { return u[0] * u[0] + u[1] * u[1] + u[2] * u[2]; }

// PhaseUEq::extendedCalculusPhase + MixUEq::computeMixtureVariables; Euler:
// PhaseEuler::extendedCalculusPhase (also rebuilds E when rebuildE).
template <int MODEL, int N>
__device__ __forceinline__ void derive(const Params& P, Cell<N>& s, bool rebuildE) {
  if (MODEL == M_EULER) {
    s.e[0] = eos_energy(P.eos[0], P.eps, s.r[0], s.p[0]);
    s.c[0] = eos_sound(P.eos[0], P.eps, s.r[0], s.p[0]);
    if (rebuildE) s.E = s.e[0] + 0.5 * sqn3(s.u);
    return;
  }
#pragma unroll
  for (int k = 0; k < N; k++) {
    s.e[k] = eos_energy(P.eos[k], P.eps, s.r[k], s.p[k]);
    s.c[k] = eos_sound(P.eos[k], P.eps, s.r[k], s.p[k]);
  }
  double rho = 0., pm = 0.;
#pragma unroll
  for (int k = 0; k < N; k++) { rho += s.a[k] * s.r[k]; pm += s.a[k] * s.p[k]; }
  double den = dmax(rho, P.eps), em = 0., cf = 0.;
#pragma unroll
  for (int k = 0; k < N; k++) {
    double Y = s.a[k] * s.r[k] / den;
    em += Y * s.e[k];
    cf += Y * s.c[k] * s.c[k];
  }
  s.rho = rho; s.pm = pm; s.em = em; s.cF = sqrt(cf);
}

// verifyAndCorrectPhase (PhaseUEq.cpp:355-370, PhaseEuler.cpp:395-399)
template <int MODEL, int N>
__device__ __forceinline__ void verify_correct(const Params& P, Cell<N>& s, int k) {
  if (MODEL == M_EULER) {
    if (s.r[0] < 1.e-10) s.r[0] = 1.e-10;
    s.p[0] = eos_clamp_p(P.eos[0], s.p[0]);
    return;
  }
  if (P.eps > 1.e-20) {
    if (s.a[k] < 0.) s.a[k] = 0.;
    if (s.a[k] > 1.) s.a[k] = 1.;
    if (s.r[k] <= 1.e-15) s.r[k] = 1.e-15;
  } else {
    if (s.a[k] < 1e-15) s.a[k] = 1e-14;
    if (s.a[k] > 1. - 1e-15) s.a[k] = 1. - 1e-14;
    if (s.r[k] < 1.e-15) s.r[k] = 1.e-14;
  }
  s.p[k] = eos_clamp_p(P.eos[k], s.p[k]);
}

// ---------------------------------------------------------------------------
// Riemann solvers in the face frame; lu = local velocity of each side.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cfl_bound(double& dtm, double s, double dx) {
  if (fabs(s) > 1.e-3) dtm = dmin(dtm, 1. * dx / fabs(s));
}

// ModEuler::solveRiemannIntern
template <int N>
__device__ __forceinline__ void hllc_euler(const Cell<N>& Lc, const double* lu, const double EL, const Cell<N>& Rc,
                                           const double* ru, const double ER, double lcfl, double& dtm,
                                           Flux<N>& F) {
  const double uL = lu[0], vL = lu[1], wL = lu[2], pL = Lc.p[0], rhoL = Lc.r[0], cL = Lc.c[0];
  const double uR = ru[0], vR = ru[1], wR = ru[2], pR = Rc.p[0], rhoR = Rc.r[0], cR = Rc.c[0];
  const double sL = dmin(uL - cL, uR - cR);
  const double sR = dmax(uR + cR, uL + cL);
  cfl_bound(dtm, sL, lcfl);
  cfl_bound(dtm, sR, lcfl);
  const double mL = rhoL * (sL - uL), mR = rhoR * (sR - uR);
  double sM = (pR - pL + mL * uL - mR * uR) / (mL - mR);
  if (fabs(sM) < 1.e-8) sM = 0.;
  double mass, m0, m1, m2, en;
  if (sL > 0.) {
    mass = rhoL * uL; m0 = rhoL * uL * uL + pL; m1 = rhoL * vL * uL; m2 = rhoL * wL * uL; en = (rhoL * EL + pL) * uL;
  } else if (sR < 0.) {
    mass = rhoR * uR; m0 = rhoR * uR * uR + pR; m1 = rhoR * vR * uR; m2 = rhoR * wR * uR; en = (rhoR * ER + pR) * uR;
  } else {
    const bool left = (sM >= 0.);
    const double mS = left ? mL : mR, uS = left ? uL : uR, pS = left ? pL : pR, sS = left ? sL : sR;
    const double ES = left ? EL : ER, vS = left ? vL : vR, wS = left ? wL : wR;
    const double pStar = mS * (sM - uS) + pS;
    const double rhoStar = mS / (sS - sM);
    const double Estar = ES + (sM - uS) * (sM + pS / mS);
    mass = rhoStar * sM; m0 = rhoStar * sM * sM + pStar; m1 = rhoStar * sM * vS; m2 = rhoStar * sM * wS;
    en = (rhoStar * Estar + pStar) * sM;
  }
  F.f.mass[0] = mass; F.f.mom[0] = m0; F.f.mom[1] = m1; F.f.mom[2] = m2; F.f.energM = en;
  F.sM = sM; F.uStar = sM;
}

// ModUEq::solveRiemannIntern
template <int N>
__device__ __forceinline__ void hllc_ueq(const Params& P, const Cell<N>& Lc, const double* lu, const Cell<N>& Rc,
                                         const double* ru, double& dtm, Flux<N>& F) {
  const double uL = lu[0], cL = Lc.cF, pL = Lc.pm, rhoL = Lc.rho;
  const double uR = ru[0], cR = Rc.cF, pR = Rc.pm, rhoR = Rc.rho;
  const double sL = dmin(uL - cL, uR - cR);
  const double sR = dmax(uR + cR, uL + cL);
  cfl_bound(dtm, sL, P.lcfl);
  cfl_bound(dtm, sR, P.lcfl);
  const double mL = rhoL * (sL - uL), mR = rhoR * (sR - uR);
  double sM = (pR - pL + mL * uL - mR * uR) / (mL - mR);
  if (fabs(sM) < 1.e-8) sM = 0.;
  if (sL >= 0. || sR <= 0.) {
    const bool left = (sL >= 0.);
    const Cell<N>& S = left ? Lc : Rc;
    const double* su = left ? lu : ru;
    const double uS = su[0], rhoS = left ? rhoL : rhoR, pS = left ? pL : pR;
    const double totalEnergy = S.em + 0.5 * sqn3(su);
#pragma unroll
    for (int k = 0; k < N; k++) {
      F.f.alpha[k] = S.a[k] * sM;
      F.f.mass[k] = S.a[k] * S.r[k] * uS;
      F.f.energ[k] = S.a[k] * S.r[k] * S.e[k] * uS;
    }
    F.f.mom[0] = rhoS * uS * uS + pS;
    F.f.mom[1] = rhoS * su[1] * uS;
    F.f.mom[2] = rhoS * su[2] * uS;
    F.f.energM = (rhoS * totalEnergy + pS) * uS;
    F.uStar = uS;
  } else {
    const bool left = (sM >= 0.);
    const Cell<N>& S = left ? Lc : Rc;
    const double* su = left ? lu : ru;
    const double uS = su[0], pS = left ? pL : pR, mS = left ? mL : mR, sS = left ? sL : sR;
    const double totalEnergy = S.em + 0.5 * sqn3(su);
    const double rhoStar = mS / (sS - sM);
    const double EStar = totalEnergy + (sM - uS) * (sM + pS / mS);
    const double pStar = mS * (sM - uS) + pS;
#pragma unroll
    for (int k = 0; k < N; k++) {
      const double mk = S.r[k] * (sS - uS);
      const double rks = mk / (sS - sM);
      const double pks = eos_p_isentropic(P.eos[k], P.eps, S.p[k], S.r[k], rks);
      const double eks = eos_energy(P.eos[k], P.eps, rks, pks);
      F.f.alpha[k] = S.a[k] * sM;
      F.f.mass[k] = S.a[k] * rks * sM;
      F.f.energ[k] = S.a[k] * rks * eks * sM;
    }
    F.f.mom[0] = rhoStar * sM * sM + pStar;
    F.f.mom[1] = rhoStar * su[1] * sM;
    F.f.mom[2] = rhoStar * su[2] * sM;
    F.f.energM = (rhoStar * EStar + pStar) * sM;
    F.uStar = sM;
  }
  F.sM = sM;
}

// Face flux between two derived cells (global-frame velocities), returned in
// the global frame.  EL/ER: total energies the Euler solver uses.
template <int MODEL, int N, int FR>
__device__ __forceinline__ void face_flux(const Params& P, const Cell<N>& L, double EL, const Cell<N>& R, double ER,
                                          double& dtm, Flux<N>& F) {
  double lu[3], ru[3], loc[3];
  to_local<FR>(L.u, lu);
  to_local<FR>(R.u, ru);
  if (MODEL == M_EULER) hllc_euler<N>(L, lu, EL, R, ru, ER, P.lcfl, dtm, F);
  else hllc_ueq<N>(P, L, lu, R, ru, dtm, F);
  loc[0] = F.f.mom[0]; loc[1] = F.f.mom[1]; loc[2] = F.f.mom[2];
  to_global<FR>(loc, F.f.mom);
}

// Boundary face (BoundCond::solveRiemann, BoundCond.cpp:78-98): the buffer cell
// is the interior cell with its derived state rebuilt (Euler: E rebuilt).
template <int MODEL, int N, int FR>
__device__ __forceinline__ void boundary_flux(const Params& P, const Cell<N>& B, double EB, int bc, double& dtm,
                                              Flux<N>& F) {
  if (bc == BC_WALL) {
    double lu[3];
    to_local<FR>(B.u, lu);
    const double uL = lu[0];
    const double cL = (MODEL == M_EULER) ? B.c[0] : B.cF;
    const double pL = (MODEL == M_EULER) ? B.p[0] : B.pm;
    const double rhoL = (MODEL == M_EULER) ? B.r[0] : B.rho;
    const double sL = dmin(uL - cL, -uL - cL);
    cfl_bound(dtm, sL, P.lcfl);
    const double pStar = (MODEL == M_EULER) ? rhoL * uL * (uL - sL) + pL : rhoL * (uL - sL) * uL + pL;
#pragma unroll
    for (int k = 0; k < N; k++) { F.f.alpha[k] = 0.; F.f.mass[k] = 0.; F.f.energ[k] = 0.; }
    double loc[3] = {pStar, 0., 0.};
    to_global<FR>(loc, F.f.mom);
    F.f.energM = 0.;
    F.sM = 0.;
    F.uStar = 0.;
  } else {
    face_flux<MODEL, N, FR>(P, B, EB, B, EB, dtm, F);
  }
}

// CellInterface::addFlux / subtractFlux with the non-conservative terms of the
// cell's time-n phases (FluxUEq.cpp:61-84, 224-244; FluxPUEq.cpp:44-65).
template <int MODEL, int N>
__device__ __forceinline__ void add_flux(Cons<N>& A, const Flux<N>& F, double coefA, const double* an,
                                         const double* pn) {
  if (MODEL == M_EULER) {
    A.mass[0] += coefA * F.f.mass[0];
  } else {
#pragma unroll
    for (int k = 0; k < N; k++) {
      A.alpha[k] += coefA * F.f.alpha[k];
      A.mass[k] += coefA * F.f.mass[k];
      A.energ[k] += coefA * F.f.energ[k];
    }
  }
  A.mom[0] += F.f.mom[0] * coefA;
  A.mom[1] += F.f.mom[1] * coefA;
  A.mom[2] += F.f.mom[2] * coefA;
  A.energM += coefA * F.f.energM;
  if (MODEL != M_EULER) {
#pragma unroll
    for (int k = 0; k < N; k++) {
      A.alpha[k] += -coefA * an[k] * F.sM;
      if (MODEL == M_PUEQ) A.energ[k] += coefA * an[k] * pn[k] * F.uStar;
    }
  }
}

template <int MODEL, int N>
__device__ __forceinline__ void sub_flux(Cons<N>& A, const Flux<N>& F, double coefA, const double* an,
                                         const double* pn) {
  if (MODEL == M_EULER) {
    A.mass[0] -= coefA * F.f.mass[0];
  } else {
#pragma unroll
    for (int k = 0; k < N; k++) {
      A.alpha[k] -= coefA * F.f.alpha[k];
      A.mass[k] -= coefA * F.f.mass[k];
      A.energ[k] -= coefA * F.f.energ[k];
    }
  }
  A.mom[0] -= F.f.mom[0] * coefA;
  A.mom[1] -= F.f.mom[1] * coefA;
  A.mom[2] -= F.f.mom[2] * coefA;
  A.energM -= coefA * F.f.energM;
  if (MODEL != M_EULER) {
#pragma unroll
    for (int k = 0; k < N; k++) {
      A.alpha[k] -= -coefA * an[k] * F.sM;
      if (MODEL == M_PUEQ) A.energ[k] -= coefA * an[k] * pn[k] * F.uStar;
    }
  }
}

// Cell::timeEvolution (Cell.cpp:286-299): A <- scale*A (+ scheme correction) + U^n.
template <int MODEL, int N>
__device__ __forceinline__ void time_evolution(const Params& P, Cons<N>& A, const Cell<N>& n, double scale) {
  Cons<N> U;
  if (MODEL == M_EULER) {
    U.mass[0] = n.r[0];
    U.mom[0] = n.u[0] * U.mass[0]; U.mom[1] = n.u[1] * U.mass[0]; U.mom[2] = n.u[2] * U.mass[0];
    U.energM = U.mass[0] * n.E;
    A.mass[0] *= scale;
  } else {
#pragma unroll
    for (int k = 0; k < N; k++) {
      U.alpha[k] = n.a[k];
      U.mass[k] = n.a[k] * n.r[k];
      U.energ[k] = n.a[k] * n.r[k] * eos_energy(P.eos[k], P.eps, n.r[k], n.p[k]);
      A.alpha[k] *= scale; A.mass[k] *= scale; A.energ[k] *= scale;
    }
    U.mom[0] = n.u[0] * n.rho; U.mom[1] = n.u[1] * n.rho; U.mom[2] = n.u[2] * n.rho;
    U.energM = n.rho * n.E;
  }
  A.mom[0] *= scale; A.mom[1] *= scale; A.mom[2] *= scale;
  A.energM *= scale;
  if (MODEL == M_UEQ) {
    double rhoNp1 = 0., sumD = 0., m[3];
#pragma unroll
    for (int k = 0; k < N; k++) { rhoNp1 += A.mass[k] + U.mass[k]; sumD += A.energ[k]; }
    m[0] = A.mom[0] + U.mom[0]; m[1] = A.mom[1] + U.mom[1]; m[2] = A.mom[2] + U.mom[2];
    const double dRU2 = sqn3(m) / rhoNp1 - sqn3(U.mom) / n.rho;
    const double err = (A.energM - 0.5 * dRU2) - sumD;
#pragma unroll
    for (int k = 0; k < N; k++) A.energ[k] += n.a[k] * err;
  }
  if (MODEL == M_EULER) {
    A.mass[0] += 1. * U.mass[0];
  } else {
#pragma unroll
    for (int k = 0; k < N; k++) {
      A.alpha[k] += 1. * U.alpha[k]; A.mass[k] += 1. * U.mass[k]; A.energ[k] += 1. * U.energ[k];
    }
  }
  A.mom[0] += U.mom[0] * 1.; A.mom[1] += U.mom[1] * 1.; A.mom[2] += U.mom[2] * 1.;
  A.energM += 1. * U.energM;
}

// Flux::buildPrim (FluxUEq.cpp:161-207, FluxEuler.cpp:124-150); derived rebuilt.
template <int MODEL, int N>
__device__ __forceinline__ void build_prim(const Params& P, Cons<N>& A, Cell<N>& s) {
  if (MODEL == M_EULER) {
    const double m = A.mass[0];
    s.r[0] = m;
    s.u[0] = A.mom[0] / m; s.u[1] = A.mom[1] / m; s.u[2] = A.mom[2] / m;
    if (fabs(s.u[0]) < 1.e-8) s.u[0] = 0.;
    if (fabs(s.u[1]) < 1.e-8) s.u[1] = 0.;
    if (fabs(s.u[2]) < 1.e-8) s.u[2] = 0.;
    s.E = A.energM / m;
    const double e = s.E - 0.5 * sqn3(s.u);
    s.e[0] = e;
    s.p[0] = eos_pressure(P.eos[0], m, e);
    s.c[0] = eos_sound(P.eos[0], P.eps, m, s.p[0]);
    return;
  }
  double un = 0., rhoMel = 0.;
  if (P.eps > 1.e-20) {
#pragma unroll
    for (int k = 0; k < N; k++) {
      if (A.alpha[k] < 0.) A.alpha[k] = 0.;
      if (A.alpha[k] > 1.) A.alpha[k] = 1.;
      un += A.alpha[k];
    }
  } else {
#pragma unroll
    for (int k = 0; k < N; k++) {
      if (A.alpha[k] <= 1.e-15) A.alpha[k] = 1e-15;
      if (A.alpha[k] >= 1. - 1.e-15) A.alpha[k] = 1.0 - 1e-15;
      un += A.alpha[k];
    }
  }
#pragma unroll
  for (int k = 0; k < N; k++) A.alpha[k] /= un;
#pragma unroll
  for (int k = 0; k < N; k++) {
    rhoMel += A.mass[k];
    s.a[k] = A.alpha[k];
    s.r[k] = A.mass[k] / dmax(A.alpha[k], P.eps);
    const double ek = A.energ[k] / dmax(A.mass[k], P.eps);
    s.p[k] = eos_pressure(P.eos[k], s.r[k], ek);
    verify_correct<MODEL, N>(P, s, k);
  }
  s.u[0] = A.mom[0] / rhoMel; s.u[1] = A.mom[1] / rhoMel; s.u[2] = A.mom[2] / rhoMel;
  if (fabs(s.u[0]) < 1.e-8) s.u[0] = 0.;
  if (fabs(s.u[1]) < 1.e-8) s.u[1] = 0.;
  if (fabs(s.u[2]) < 1.e-8) s.u[2] = 0.;
  derive<MODEL, N>(P, s, false);
  s.E = A.energM / rhoMel;
}

// RelaxationPInfinite::relaxation + RelaxationP::NewtonRaphson from initial
// guess p0 (the reference's guess is sum alpha_k p_k).  Returns iterations
// (0 when gated out).
template <int N>
__device__ __forceinline__ int relax_newton(const Params& P, Cell<N>& s, double pStar) {
  if (P.eps > 1.e-20) {
#pragma unroll
    for (int k = 0; k < N; k++)
      if (s.a[k] > (1. - 1.e-5)) return 0;
  }
  double ak[N], rk[N], pk[N], akS[N], rkS[N];
#pragma unroll
  for (int k = 0; k < N; k++) { ak[k] = s.a[k]; rk[k] = s.r[k]; pk[k] = s.p[k]; }
  double f = 0., df = 1.;
  int it = 0;
  do {
    it++;
    pStar -= f / df;
#pragma unroll
    for (int k = 0; k < N; k++) pStar = eos_clamp_p(P.eos[k], pStar);
    f = -1.;
    df = 0.;
#pragma unroll
    for (int k = 0; k < N; k++) {
      double drhodp;
      rkS[k] = eos_rho_pfinal(P.eos[k], P.eps, pk[k], rk[k], pStar, drhodp);
      akS[k] = ak[k] * rk[k] / rkS[k];
      const double dalpha = ak[k] * rk[k] * drhodp / (rkS[k] * rkS[k]);
      f += akS[k];
      df -= dalpha;
    }
  } while (fabs(f) > 1e-10 && it < 100);
  if (it < 100) {
#pragma unroll
    for (int k = 0; k < N; k++) { s.a[k] = akS[k]; s.r[k] = rkS[k]; s.p[k] = pStar; }
  } else {
    // Interface-pressure fallback (RelaxationPInfinite.cpp:82-97, RelaxationP.cpp:72-93)
    double z[N], pI = 0., sumZk = 0.;
#pragma unroll
    for (int k = 0; k < N; k++) z[k] = rk[k] * s.c[k];
#pragma unroll
    for (int k = 0; k < N; k++) {
      sumZk += z[k];
      double sumZj = 0.;
#pragma unroll
      for (int j = 0; j < N; j++)
        if (j != k) sumZj += z[j];
      pI += sumZj * (pk[k] - 0.);
    }
    pI /= sumZk * (double)(N - 1);
#pragma unroll
    for (int k = 0; k < N; k++) pI = eos_clamp_p(P.eos[k], pI);
#pragma unroll
    for (int k = 0; k < N; k++) {
      double drhodp;
      const double rho_s = eos_rho_pfinal(P.eos[k], P.eps, pk[k], rk[k], pI, drhodp);
      s.a[k] = ak[k] * rk[k] / rho_s;
      s.r[k] = rho_s;
      s.p[k] = pI;
    }
  }
  return it;
}

template <int N>
__device__ __forceinline__ double relax_guess(const Cell<N>& s) {
  double p = 0.;
#pragma unroll
  for (int k = 0; k < N; k++) p += s.a[k] * s.p[k];
  return p;
}

// Cell::correctionEnergy + FluxPUEq::correctionEnergy (Kapila only), then fulfillState.
template <int MODEL, int N>
__device__ __forceinline__ void correct_and_derive(const Params& P, Cell<N>& s) {
  if (MODEL == M_PUEQ) {
    const double em = s.E - 0.5 * sqn3(s.u);
    double pp = s.rho * em, denom = 0.;
#pragma unroll
    for (int k = 0; k < N; k++) {
      pp -= s.a[k] * (P.eos[k].mixg * (1. - s.r[k] * 0.) + s.r[k] * P.eos[k].eref);
      denom += s.a[k] * P.eos[k].mixinv * (1. - s.r[k] * 0.);
    }
    pp /= denom;
#pragma unroll
    for (int k = 0; k < N; k++) {
      s.p[k] = pp;
      verify_correct<MODEL, N>(P, s, k);
    }
  }
  derive<MODEL, N>(P, s, MODEL == M_EULER);
}

}  // namespace ecg
