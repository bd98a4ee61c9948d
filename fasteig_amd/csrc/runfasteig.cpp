// runfasteig_hip -- input-deck driver on the HIP backend (reference:
// examples/runfasteig.cpp:24-113). Usage:
//   runfasteig_hip <deck.input> [key=value;key=value]
//   runfasteig_hip --version | --build-info | --check-runtime
// Writes <logfile>_<solver>.log with the RAW RESULTS / DIAGNOSTICS blocks the
// reference's test harness parses. Exit codes as in the reference: 0 success,
// 2 when the solve ends with a nonzero status, 77 when no GPU is available,
// 1 on errors.
#include <fstream>
#include <iostream>
#include <string>

#include "fasteig_problem.hpp"
#include "fe_hip.h"
#include "parseinput.hpp"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << "usage: runfasteig_hip <input deck> [key=value;...]\n";
    return 1;
  }
  const std::string arg = argv[1];
  if (arg == "--version") {
    std::cout << fe_version() << "\n";
    return 0;
  }
  if (arg == "--build-info") {
    std::cout << fe_version() << "\nbackend: HIP (gfx950)\n";
    return 0;
  }
  if (arg == "--check-runtime") {
    fe_ctx ctx = nullptr;
    if (fe_ctx_create(0, &ctx) != 0) {
      std::cerr << "no usable GPU: " << fe_last_error() << "\n";
      return 77;
    }
    fe_ctx_destroy(ctx);
    return 0;
  }
  try {
    fasteig_params params;
    if (parseinput(argv[1], &params) != 0) {
      std::cerr << "failed to parse " << argv[1] << "\n";
      return 1;
    }
    fasteig_problem prob(params);
    fasteig_options opts;
    opts.configure_from(params);
    if (argc > 2) opts.apply_string(argv[2]);
    {
      fe_ctx probe = nullptr;
      if (fe_ctx_create(opts.device(), &probe) != 0) {
        std::cerr << "no usable GPU: " << fe_last_error() << "\n";
        return 77;
      }
      fe_ctx_destroy(probe);
    }
    const FASTEIG_INT status = prob.solve(opts);
    const std::string path = opts.output_path(prob.solver());
    if (!path.empty()) {
      std::ofstream log(path);
      log << "FASTEIG (HIP backend) results for " << params.amatfname << "\n";
      log << "  n = " << prob.n() << ", nnz = " << prob.nnz() << ", solver = " << prob.solver()
          << ", which = " << opts.which() << ", nev = " << opts.nev() << ", tol = " << opts.tol()
          << ", precision = " << opts.precision() << "\n";
      prob.write_results_text(log);
    }
    if (opts.verbosity() >= 1) prob.write_results_text(std::cout);
    return status == 0 ? 0 : 2;
  } catch (const std::exception& e) {
    std::cerr << "runfasteig_hip: " << e.what() << "\n";
    return 1;
  }
}
