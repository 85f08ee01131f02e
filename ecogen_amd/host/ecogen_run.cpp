// ecogen_run: runs a case written in the reference's XML input format on the
// GPU through libecogen_amd (Run::initialize + Run::solver, Run.cpp:82-498),
// writing the reference's result files under <work>/results/<run>/ at the
// case's print times (and resuming from them when main.xml asks for a restart).
//
//   ecogen_run <case dir> [--eos-dir DIR] [--precision fp64|fp32|fp32_learned]
//              [--max-iter N] [--work DIR] [--no-output] [--out FILE]
//
// --out also writes the final state as a text header line (nx ny nz nvar t
// iterations) followed by nvar arrays of nx*ny*nz little-endian doubles in the
// C ABI's variable order.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "ecogen.hpp"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <case dir> [--eos-dir DIR] [--precision fp64|fp32|fp32_learned] "
                         "[--max-iter N] [--work DIR] [--no-output] [--out FILE]\n", argv[0]);
    return 2;
  }
  std::string caseDir = argv[1], eosDir, out, work = ".";
  bool write = true;
  int precision = ECG_FP64;
  long long maxIter = -1;
  for (int i = 2; i < argc; i++) {
    const std::string a = argv[i];
    if (a == "--eos-dir" && i + 1 < argc) eosDir = argv[++i];
    else if (a == "--out" && i + 1 < argc) out = argv[++i];
    else if (a == "--work" && i + 1 < argc) work = argv[++i];
    else if (a == "--no-output") write = false;
    else if (a == "--max-iter" && i + 1 < argc) maxIter = std::atoll(argv[++i]);
    else if (a == "--precision" && i + 1 < argc) {
      const std::string p = argv[++i];
      if (p == "fp64") precision = ECG_FP64;
      else if (p == "fp32") precision = ECG_FP32_RELAX64;
      else if (p == "fp32_learned") precision = ECG_FP32_LEARNED;
      else { std::fprintf(stderr, "unknown precision %s\n", p.c_str()); return 2; }
    } else {
      std::fprintf(stderr, "unknown argument %s\n", a.c_str());
      return 2;
    }
  }
  try {
    ecg_host::CaseDescription c = ecg_host::Input::readInputXML(caseDir, eosDir, precision);
    if (maxIter > 0) {
      c.control.iterationsMode = true;
      c.control.iterations = maxIter;
      if (c.control.iterFreq <= 0) c.control.iterFreq = (int)maxIter;
    }
    std::printf("case %s: %s, %d phase(s), %dx%dx%d cells, %s, CFL %g, %zu domains\n", c.control.runName.c_str(),
                c.model.whoAmI().c_str(), c.model.numberPhases, c.config.nx, c.config.ny, c.config.nz,
                c.order.c_str(), c.config.cfl, c.domains.size());
    ecg_host::Run run(c, work, write);
    run.initialize();
    const auto t0 = std::chrono::steady_clock::now();
    const ecg_step_stats st = run.solver();
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const double cells = (double)c.config.nx * c.config.ny * c.config.nz;
    std::printf("iterations %lld, t = %.10e, dt_next = %.6e, %.3f s (%.1f Mcell-updates/s), "
                "Newton evaluations %lld over %lld relaxations, fallbacks %lld\n",
                (long long)st.iteration, st.t, st.dt_next, el, cells * st.iteration / el / 1e6,
                (long long)st.newton_iters, (long long)st.relax_calls, (long long)st.newton_fail);
    if (!out.empty()) {
      const auto s = run.state();
      FILE* f = std::fopen(out.c_str(), "wb");
      if (!f) { std::fprintf(stderr, "cannot write %s\n", out.c_str()); return 1; }
      std::fprintf(f, "%d %d %d %zu %.17g %lld\n", c.config.nx, c.config.ny, c.config.nz, s.size(), st.t,
                   (long long)st.iteration);
      for (const auto& v : s) std::fwrite(v.data(), sizeof(double), v.size(), f);
      std::fclose(f);
      std::printf("state written to %s\n", out.c_str());
    }
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
