// fsx_solver -- command-line driver in the shape of the reference's main()
// (BCH:383-521): reads the problem name from ProblemName.txt (BCH:536-540),
// or a .inp path / "cavity N" / "duct" from the command line, sets up, runs
// the time loop and prints one monitor line per step (BCH:3091-3093).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../../include/fsx_host.h"

static void usage() {
  std::fprintf(stderr,
               "usage: fsx_solver [file.inp | cavity N | duct] [--ngp 8|27] [--steps K] [--mode fp64|mixed|direct]\n"
               "                  [--rtol R] [--re RE] [--dt DT]\n"
               "       without a mesh argument the .inp named in ProblemName.txt is read\n");
}

int main(int argc, char** argv) {
  fsxh_mesh* m = nullptr;
  char err[512] = {0};
  int a = 1;
  int steps = -1, ngp = -1, mode = FSX_SOLVER_FP64_JACOBI;
  double rtol = 1e-8, re = -1, dt = -1;
  if (a < argc && std::strcmp(argv[a], "cavity") == 0 && a + 1 < argc) {
    m = fsxh_mesh_cavity(std::atoi(argv[a + 1]), 0.0, 1.0);
    a += 2;
  } else if (a < argc && std::strcmp(argv[a], "duct") == 0) {
    m = fsxh_mesh_duct(16, 40, 40, 80, 5.0, 2.0, 10.0, 1.0, 12345, 1);
    a += 1;
  } else {
    std::string path;
    if (a < argc && argv[a][0] != '-') path = argv[a++];
    else {
      std::ifstream pn("ProblemName.txt");
      std::string name;
      if (!(pn >> name)) { usage(); return 2; }
      path = name + ".inp";
    }
    if (fsxh_mesh_read_inp(path.c_str(), &m, err, sizeof err) != FSX_OK) {
      std::fprintf(stderr, "cannot read %s: %s\n", path.c_str(), err);
      return 1;
    }
  }
  if (!m) { usage(); return 2; }
  for (; a < argc; a++) {
    std::string s = argv[a];
    auto next = [&]() { return a + 1 < argc ? argv[++a] : (char*)"0"; };
    if (s == "--steps") steps = std::atoi(next());
    else if (s == "--ngp") ngp = std::atoi(next());
    else if (s == "--mode") {
      const char* m = next();
      mode = std::strcmp(m, "mixed") == 0 ? FSX_SOLVER_MIXED_IR
             : std::strcmp(m, "direct") == 0 ? FSX_SOLVER_DIRECT : FSX_SOLVER_FP64_JACOBI;
    }
    else if (s == "--rtol") rtol = std::atof(next());
    else if (s == "--re") re = std::atof(next());
    else if (s == "--dt") dt = std::atof(next());
    else { usage(); return 2; }
  }
  long long info[8];
  double par[10];
  fsxh_mesh_info(m, info);
  fsxh_mesh_params(m, par);
  if (ngp < 0) ngp = (int)info[5];
  double visc = re > 0 ? 1.0 / re : par[6];
  if (dt < 0) dt = par[0];
  fsxh_mesh_set_params(m, ngp, dt, (int)info[6], par[3], par[4], par[5], visc);
  if (steps < 0) {
    // the reference advances while timeT < t_final (BCH:2859)
    steps = 0;
    for (double t = par[1]; t < par[2]; t += dt) steps++;
  }
  fsxd_options o;
  fsxd_default_options(&o);
  o.NGP = ngp; o.maxIter = (int)info[6]; o.tolerance = par[3]; o.converge = par[4];
  o.solver_mode = mode; o.pcg_rtol = rtol;
  fsxd_solver* s = nullptr;
  if (fsxd_create(m, &o, &s, err, sizeof err) != FSX_OK) {
    std::fprintf(stderr, "setup failed: %s\n", err);
    return 1;
  }
  double st[3];
  fsxd_setup_times(s, st);
  long long ti[11];
  fsxh_topo_info(fsxd_topo(s), ti);
  std::printf(" NN = %lld\n NNp = %lld\n sparseM_NNZ = %lld\n sparseG_NNZ = %lld\n NNZ of Z = %lld\n", ti[2], ti[3],
              3 * ti[4], 3 * ti[5], ti[6]);
  std::printf(" setup: topology %.3f s, upload %.3f s, assembly %.3f s\n\n", st[0], st[1], st[2]);
  std::printf("Time step  Iter     Time       u_monitor     v_monitor     w_monitor     p_monitor     TimeSpend      maxAcc   PCG\n");
  std::vector<fsxd_step_stats> stats(1);
  for (int n = 0; n < steps; n++) {
    int done = fsxd_run(s, 1, stats.data());
    if (done < 0) { std::fprintf(stderr, "step failed: %s\n", fsxd_error(s)); return 1; }
    if (done == 0) break;
    const fsxd_step_stats& q = stats[0];
    std::printf("%6d  %6d  %10.5f  %12.5f  %12.5f  %12.5f  %12.5f %12.5f %12.5g %5d\n", q.step, q.inner_iters, q.time,
                q.mon[0], q.mon[1], q.mon[2], q.mon[3], q.wall_s, q.max_acc, q.pcg_iters);
    if (q.steady) break;
  }
  fsxd_free(s);
  fsxh_mesh_free(m);
  return 0;
}
