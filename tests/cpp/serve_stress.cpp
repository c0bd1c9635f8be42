// Serving runtime under many workers, for a ThreadSanitizer build of the
// host layer (tests/cpp/Makefile target tsan: every host_*.cpp compiled
// with -fsanitize=thread, the HIP kernels as usual).  Twelve workers serve
// presolve + postsolve jobs on LPs each worker generates on the device
// (C-ABI gp_server_submit_spec), sizes from 300 to 20,000 rows, with the
// duplicate-row and duplicate-column rules on; every job must succeed.
// Usage: serve_stress [jobs] [workers].  Prints one summary line.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gp_presolve_capi.h"

int main(int argc, char** argv) {
  const int jobs = argc > 1 ? std::atoi(argv[1]) : 400;
  const int workers = argc > 2 ? std::atoi(argv[2]) : 12;
  gp_params prm;
  gp_params_default(&prm);
  prm.enable_structure_specific_rules = 0;
  gp_serve_config cfg{workers, 1, 0.0, 0x5eedull};
  gp_server* srv = nullptr;
  if (gp_server_create(&prm, &cfg, &srv) != GP_OK) {
    std::printf("server create failed: %s\n", gp_last_error());
    return 2;
  }
  for (int k = 0; k < jobs; ++k) {
    const int m = 300 + (k * 7919) % 19700;
    const gp_lp_spec spec{0x1000ull + static_cast<uint64_t>(k), m, m + m / 2, 2, 10, (k % 3) ? 6.0 : 0.0, 1};
    if (gp_server_submit_spec(srv, &spec, static_cast<uint64_t>(k), -1.0, nullptr) != GP_OK) {
      std::printf("submit %d failed: %s\n", k, gp_last_error());
      return 2;
    }
  }
  std::vector<gp_job_result> res(64);
  int done = 0, failed = 0;
  while (done < jobs) {
    const int n = gp_server_wait(srv, res.data(), static_cast<int>(res.size()), 5.0);
    if (n < 0) {
      std::printf("wait failed: %s\n", gp_last_error());
      return 2;
    }
    for (int i = 0; i < n; ++i) failed += res[static_cast<size_t>(i)].status != GP_OK;
    done += n;
  }
  gp_server_destroy(srv);
  std::printf("serve_stress: %d jobs on %d workers, %d failed\n", jobs, workers, failed);
  return failed == 0 ? 0 : 1;
}
