// Checker driver for the reference-compiled EOS and limiter sources.
//
// Links the reference's own EosSG / EosIG / Limiter* objects (compiled from
// the reference tree by oracle/ref.mk) and evaluates them on the binary input
// records read from stdin, writing the results to stdout.  Used only to pin
// the oracle's restatement of those functions.
//
// Input: int32 op, int32 eos_kind, then 8 doubles (gamma, pinf, cv, eref,
// eps, x0, x1, x2) per record.  Output: 2 doubles per record.
#include <cstdio>
#include <cstdint>
#include <string>
#include <vector>

#include "Eos/EosSG.h"
#include "Eos/EosIG.h"
#include "Order2/LimiterMinmod.h"
#include "Order2/LimiterVanLeer.h"
#include "Order2/LimiterVanAlbada.h"
#include "Order2/LimiterSuperBee.h"
#include "Order2/LimiterMC.h"

int main() {
  // EosSG::computeDensityPfinal calls TB->returnNonZeroValue, which reads no
  // member; the Tools constructor itself writes mesh globals that live in code
  // not built here, so the object is given storage without being constructed.
  TB = static_cast<Tools*>(::operator new(sizeof(Tools)));
  LimiterMinmod lmm; LimiterVanLeer lvl; LimiterVanAlbada lva; LimiterSuperBee lsb; LimiterMC lmc;
  Limiter* lims[6] = {nullptr, &lmm, &lvl, &lva, &lsb, &lmc};
  struct Rec { int32_t op, kind; double v[8]; } r;
  while (std::fread(&r, sizeof(r), 1, stdin) == 1) {
    std::vector<std::string> names;
    int num = 0;
    Eos* eos;
    if (r.kind == 1) {
      eos = new EosSG(names, num);
      eos->assignParametersEos("sg", std::vector<double>{r.v[0], r.v[1], r.v[2], r.v[3], 0.});
    } else {
      eos = new EosIG(names, num);
      eos->assignParametersEos("ig", std::vector<double>{r.v[0], r.v[2], r.v[3], 0.});
    }
    eos->assignEpsilonForAlphaNull(r.v[4] > 0.);
    double out[2] = {0., 0.};
    const double a = r.v[5], b = r.v[6], c = r.v[7];
    switch (r.op) {
      case 0: out[0] = eos->computeEnergy(a, b); break;
      case 1: out[0] = eos->computePressure(a, b); break;
      case 2: out[0] = eos->computeSoundSpeed(a, b); break;
      case 3: out[0] = eos->computeTemperature(a, b); break;
      case 4: out[0] = eos->computePressureIsentropic(a, b, c); break;
      case 5: out[0] = eos->computeDensityPfinal(a, b, c, &out[1]); break;
      case 6: out[0] = eos->computeDensityHugoniot(a, b, c, &out[1]); break;
      case 7: { double p = a; eos->verifyAndModifyPressure(p); out[0] = p; } break;
      case 8: out[0] = lims[r.kind]->limiteSlope(a, b); break;
      default: break;
    }
    std::fwrite(out, sizeof(out), 1, stdout);
    delete eos;
  }
  return 0;
}
