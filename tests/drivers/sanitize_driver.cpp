// sanitize_driver.cpp -- a host-side workout of micropp_amd for the
// sanitizer builds (libmicropp_amd_asan.so / libmicropp_amd_tsan.so, host
// code instrumented, GPU code as usual): every host path that allocates,
// copies between host and device, or runs host threads, through the drop-in
// C++ interface.
//   1. a J2 instance with 4 FE Gauss points: the batch of GP slots, 2 load
//      steps with history, update_vars;
//   2. the same with MPPD_GP_BATCH=0: concurrent solvers driven from host
//      threads;
//   3. an all-elastic instance with mixed precision (the multi-RHS solver
//      with its coarse level) and 6 GPs;
//   4. a J2 instance with mixed precision and MPPD_GP_BATCH=0 (the mixed
//      Newton of each GP, with its coarse level);
//   5. write_restart / read_restart and output2 of instance 1 in a
//      temporary directory.
// It prints one checksum line per case and exits 0; a sanitizer report ends
// the process with a non-zero status (halt_on_error).  It leaves with _exit
// after its last case: the HIP runtime's own teardown at exit trips an
// internal check of the ROCm sanitizer runtime's device-memory allocator
// (sanitizer_allocator_device.h, "dev_runtime_unloaded_"), which is outside
// what this driver tests.
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <string>

#include "micropp.hpp"

namespace {

material_base mat(int type, double E, double nu, double Ka = 0, double Sy = 0) {
  material_base m;
  material_set(&m, type, E, nu, Ka, Sy, 0.0);
  return m;
}

micropp_params_t params(int n, int ngp, const int *coupling, bool plastic) {
  micropp_params_t p;
  p.ngp = ngp;
  p.size[0] = p.size[1] = p.size[2] = n;
  p.type = MIC_SPHERE;
  p.geo_params[0] = 0.25;
  p.materials[0] = plastic ? mat(1, 1e6, 0.3, 1e5, 1e3) : mat(0, 1e6, 0.3);
  p.materials[1] = mat(0, 1e7, 0.25);
  p.materials[2] = mat(0, 1e7, 0.25);
  p.coupling = const_cast<int *>(coupling);
  p.lin_stress = false;
  p.cg_rel_tol = 1e-8;
  p.cg_max_its = 2000;
  return p;
}

double run(micropp<3> &m, int ngp, int steps) {
  double sum = 0.0;
  for (int t = 0; t < steps; ++t) {
    for (int g = 0; g < ngp; ++g) {
      double eps[6] = {0, 0, 0, 0, 0, 0};
      eps[0] = 1e-3 * (t + 1) * (1 + 0.2 * g);
      eps[3] = -5e-4 * (t + 1);
      m.set_strain(g, eps);
    }
    m.homogenize();
    for (int g = 0; g < ngp; ++g) {
      double sig[6], ctan[36];
      m.get_stress(g, sig);
      m.get_ctan(g, ctan);
      for (double v : sig) sum += v;
      sum += ctan[0] * 1e-6;
    }
    m.update_vars();
  }
  return sum;
}

}  // namespace

int main() {
  setvbuf(stdout, nullptr, _IOLBF, 0);
  const int fe4[4] = {FE_FULL, FE_ONE_WAY, FE_ONE_WAY, FE_FULL};
  const int one6[6] = {FE_ONE_WAY, FE_ONE_WAY, FE_ONE_WAY, FE_ONE_WAY, FE_ONE_WAY, FE_ONE_WAY};
  {
    micropp<3> m(params(9, 4, fe4, true));
    printf("batch j2: %.12e\n", run(m, 4, 2));
    // restart and output of this instance
    const std::filesystem::path dir = std::filesystem::temp_directory_path() / "mpp_sanitize";
    std::filesystem::create_directories(dir);
    const std::filesystem::path cwd = std::filesystem::current_path();
    std::filesystem::current_path(dir);
    m.write_restart(7);
    m.read_restart(7);
    m.output2(0, 1, 2);
    std::filesystem::current_path(cwd);
    std::filesystem::remove_all(dir);
    printf("restart and output: ok\n");
  }
  setenv("MPPD_GP_BATCH", "0", 1);
  {
    micropp<3> m(params(9, 4, fe4, true));
    printf("concurrent j2: %.12e\n", run(m, 4, 2));
  }
  unsetenv("MPPD_GP_BATCH");
  {
    micropp_options_t o;
    o.mixed_precision = true;
    micropp<3> m(params(11, 6, one6, false), o);
    printf("multi-rhs elastic: %.12e\n", run(m, 6, 2));
  }
  setenv("MPPD_GP_BATCH", "0", 1);
  {
    micropp_options_t o;
    o.mixed_precision = true;
    micropp<3> m(params(9, 2, fe4, true), o);
    printf("mixed newton j2: %.12e\n", run(m, 2, 2));
  }
  unsetenv("MPPD_GP_BATCH");
  printf("done\n");
  fflush(stdout);
  _exit(0);
}
