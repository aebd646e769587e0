// fpc_main.cpp -- command-line driver of the channel solver (the reference's
// main(), main.cu:8-424, as an executable over libfpc).
//
// Same flags (common.hpp:87-127), same console output: the device line
// (common.hpp:129-134), the banner (main.cu:223-245), one line per printed step
// (main.cu:365-368: steps 1, nsteps and every print_every-th) and the timing
// summary (main.cu:380-389), so the logs of the two programs compare line by
// line.  With -write-vtk 1 (the reference default) the cell-centred VTK file is
// written at steps 1, nsteps and every max(1, vtk_every)-th step with the
// "_%06d" step suffix (main.cu:370-378), or once without suffix when
// nsteps <= 0.  The time step itself runs on the GPU through fpc_sim.
//
// Extension flags: -init-profile, -noise-amp, -seed, -mg-precision (fpc_parse_args)
// and -dump-fields PATH (raw dump of u, v, w, p after the last step).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>

#include "fpc_sim.h"
#include "fpc_struct.h"

namespace {

double wall_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

int fail(const char *what, int rc) {
  const char *msg = fpc_GetErrorMessage();
  std::fprintf(stderr, "fpc_main: %s failed with code %d: %s\n", what, rc, msg ? msg : "");
  return rc;
}

std::string step_file(const char *base, int step) {
  char buf[512];
  if (fpc_vtk_filename_with_step(base, step, buf, (int)sizeof(buf))) return std::string(base);
  return std::string(buf);
}

}  // namespace

int main(int argc, char **argv) {
  fpc_params par;
  fpc_default_params(&par);
  fpc_parse_args(argc, argv, &par);
  const char *dump = nullptr;
  for (int a = 1; a + 1 < argc; ++a)
    if (!std::strcmp(argv[a], "-dump-fields")) dump = argv[a + 1];

  if (hipSetDevice(par.device) != hipSuccess) {
    std::fprintf(stderr, "fpc_main: hipSetDevice(%d) failed\n", par.device);
    return 1;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, par.device) != hipSuccess) {
    std::fprintf(stderr, "fpc_main: hipGetDeviceProperties(%d) failed\n", par.device);
    return 1;
  }
  std::printf("GPU                   : %s\n", prop.name);
  int rc = fpc_Initialize();
  if (rc) return fail("fpc_Initialize", rc);

  fpc_sim sim = nullptr;
  rc = fpc_sim_create(&par, &sim);
  if (rc) return fail("fpc_sim_create", rc);
  auto count = [&](const char *n) { return (int)fpc_sim_scalar(sim, n); };

  std::printf("============================================================\n");
  std::printf("3D Schaefer-Turek channel with immersed full-height cylinder\n");
  std::printf("Multi-step reused-object device-solve code (CPU assembly, GPU momentum+pressure dense ops)\n");
  std::printf("Grid                  : %d x %d x %d pressure cells\n", par.nx, par.ny, par.nz);
  std::printf("Domain                : [0,%g] x [0,%g] x [0,%g]\n", par.Lx, par.Ly, par.Lz);
  std::printf("Cylinder center/rad   : (%.6f, %.6f), %.6f\n", par.xcirc, par.ycirc, par.Rcirc);
  std::printf("nu                    : %.6e\n", par.nu);
  std::printf("dt                    : %.6e\n", par.dt);
  std::printf("nsteps                : %d\n", par.nsteps);
  std::printf("Pressure cells        : %d\n", count("Np"));
  std::printf("Active u/v/w          : %d / %d / %d\n", count("Nu"), count("Nv"), count("Nw"));
  std::printf("Projection q faces    : %d / %d / %d\n", count("Nqu"), count("Nqv"), count("Nqw"));
  std::printf("Pruned isolated x-y columns: %d\n", count("n_pruned"));
  std::printf("Momentum solver       : explicit row-major Struct GMRES + Struct Jacobi (DEVICE mode, CPU assembly, "
              "reused objects, GPU momentum dense ops)\n");
  std::printf("Pressure solver       : explicit row-major Struct GMRES + Struct PFMG on sign-flipped exact A_p (DEVICE "
              "mode, CPU assembly, reused objects)\n");
  std::printf("Pressure matrix build : once\n");
  std::printf("VTK writing           : %s", par.write_vtk ? "enabled" : "disabled");
  if (par.write_vtk) std::printf(" (every %d step(s), base=%s)", std::max(1, par.vtk_every), par.vtk_filename);
  std::printf("\n");
  std::printf("============================================================\n");
  std::printf("step  ramp         max|u|         max|v|         max|w|         divL2          its(u,v,w,p)         "
              "       relres(u,v,w,p)\n");
  std::fflush(stdout);

  const double t0 = wall_s();
  for (int step = 1; step <= par.nsteps; ++step) {
    rc = fpc_sim_step(sim, 1);
    if (rc) return fail("fpc_sim_step", rc);
    auto s = [&](const char *n) { return fpc_sim_scalar(sim, n); };
    if (step == 1 || step == par.nsteps || (par.print_every > 0 && step % par.print_every == 0)) {
      std::printf("%4d  %.6f  %.6e  %.6e  %.6e  %.6e  (%d,%d,%d,%d)  (%.3e,%.3e,%.3e,%.3e)\n", step, s("ramp"),
                  s("umax"), s("vmax"), s("wmax"), s("div_l2"), (int)s("its_u"), (int)s("its_v"), (int)s("its_w"),
                  (int)s("its_p"), s("relres_u"), s("relres_v"), s("relres_w"), s("relres_p"));
      std::fflush(stdout);
    }
    if (par.write_vtk) {
      const int every = std::max(1, par.vtk_every);
      if (step == 1 || step == par.nsteps || step % every == 0) {
        rc = fpc_sim_write_vtk(sim, step_file(par.vtk_filename, step).c_str());
        if (rc) return fail("fpc_sim_write_vtk", rc);
      }
    }
  }
  if (par.write_vtk && par.nsteps <= 0) {
    rc = fpc_sim_write_vtk(sim, par.vtk_filename);
    if (rc) return fail("fpc_sim_write_vtk", rc);
  }
  const double total = wall_s() - t0;
  std::printf("============================================================\n");
  std::printf("Timing summary\n");
  std::printf("momentum build time    : %.6e s\n", fpc_sim_timer(sim, "t_build"));
  std::printf("momentum solve time    : %.6e s\n", fpc_sim_timer(sim, "t_mom"));
  std::printf("pressure solve time    : %.6e s\n", fpc_sim_timer(sim, "t_p"));
  std::printf("total wall time        : %.6e s\n", total);
  std::printf("============================================================\n");
  if (dump) {
    rc = fpc_sim_write_fields(sim, dump);
    if (rc) return fail("fpc_sim_write_fields", rc);
  }
  fpc_sim_destroy(sim);
  fpc_Finalize();
  return 0;
}
