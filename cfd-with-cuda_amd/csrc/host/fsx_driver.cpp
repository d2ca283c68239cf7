// fsx_driver.cpp -- the fractional-step driver on the host: builds the
// topology, feeds the device context through the fsx_* C-ABI and runs the
// time loop in the order of the reference's timeLoop (BCH:2815-3120).
#include "fsx_host_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace {
double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

extern "C" void fsxd_default_options(fsxd_options* o) {
  o->device = 0;
  o->NGP = 27;
  o->maxIter = 3;
  o->tolerance = 1e-3;
  o->converge = 1e-6;
  o->solver_mode = FSX_SOLVER_FP64_JACOBI;
  o->pcg_rtol = 1e-8;
  o->pcg_maxit = 5000;
  o->timing = 0;
}

#define DRV_CHECK(call)                                                             \
  do {                                                                              \
    int _rc = (call);                                                               \
    if (_rc != FSX_OK) {                                                            \
      s->err = std::string(#call) + ": " + (s->ctx ? fsx_last_error(s->ctx) : ""); \
      return _rc;                                                                   \
    }                                                                               \
  } while (0)

extern "C" int fsxd_create(const fsxh_mesh* m, const fsxd_options* o, fsxd_solver** out, char* err, int errlen) {
  auto s = std::make_unique<fsxd_solver>();
  s->opt = *o;
  auto fail = [&](int rc) {
    if (err && errlen > 0) std::snprintf(err, errlen, "%s", s->err.c_str());
    return rc;
  };
  double t0 = now_s();
  char buf[512] = {0};
  int rc = fsxh_topology(m, &s->topo, buf, sizeof buf);                 // BCH:414-480 setup chain
  if (rc != FSX_OK) { s->err = std::string("topology: ") + buf; return fail(rc); }
  s->t_topo = now_s() - t0;
  const fsxh_topo* t = s->topo;
  t0 = now_s();
  auto body = [&]() -> int {
    DRV_CHECK(fsx_create(o->device, &s->ctx));                             // selectCUDAdevice (BCH:397-399)
    DRV_CHECK(fsx_set_params(s->ctx, o->NGP, m->density, m->viscosity, m->dt));
    DRV_CHECK(fsx_set_solver(s->ctx, o->solver_mode, o->pcg_rtol, o->pcg_maxit));
    std::vector<double> corner(t->coord.begin(), t->coord.begin() + (size_t)t->NNp * 3);
    DRV_CHECK(fsx_upload_mesh(s->ctx, t->NE, t->NN, t->NNp, t->LtoGnode.data(), corner.data(),
                              (int)t->color_ptr.size() - 1, t->color_ptr.data(), t->color_order.data()));
    DRV_CHECK(fsx_upload_csr(s->ctx, FSX_CSR_M, t->NN, t->Mrow.data(), t->Mcol.data()));
    DRV_CHECK(fsx_upload_csr(s->ctx, FSX_CSR_G, t->NN, t->Grow.data(), t->Gcol.data()));
    DRV_CHECK(fsx_upload_csr(s->ctx, FSX_CSR_Z, t->NNp, t->Zrow.data(), t->Zcol.data()));
    DRV_CHECK(fsx_upload_bc(s->ctx, (int)t->BCnodes.size(), t->BCnodes.data(), t->BCvalues.data(), t->zeroP));
    DRV_CHECK(fsx_sync(s->ctx));
    s->t_upload = now_s() - t0;
    t0 = now_s();
    DRV_CHECK(fsx_assemble_static(s->ctx));                                // step0 (BCH:2833)
    DRV_CHECK(fsx_build_Z(s->ctx));                                        // calculateZ_CUSP (BCH:3303)
    DRV_CHECK(fsx_init_state(s->ctx));                                     // initializeAndAllocate (BCH:2825)
    DRV_CHECK(fsx_sync(s->ctx));
    s->t_assemble = now_s() - t0;
    return FSX_OK;
  };
  rc = body();
  if (rc != FSX_OK) return fail(rc);
  s->dt = m->dt;
  s->t = m->t_ini;
  s->t0 = m->t_ini;
  *out = s.release();
  return FSX_OK;
}

extern "C" void fsxd_free(fsxd_solver* s) {
  if (!s) return;
  if (s->ctx) fsx_destroy(s->ctx);
  fsxh_topo_free(s->topo);
  delete s;
}

// fsxd_run reports a failed call as the negative status (its positive
// returns count the steps done)
#define RUN_CHECK(call)                                                             \
  do {                                                                              \
    int _rc = (call);                                                               \
    if (_rc != FSX_OK) {                                                            \
      s->err = std::string(#call) + ": " + (s->ctx ? fsx_last_error(s->ctx) : ""); \
      return -_rc;                                                                  \
    }                                                                               \
  } while (0)

// BCH:2859-3118, one iteration of the while loop per time step.
extern "C" int fsxd_run(fsxd_solver* s, int nsteps, fsxd_step_stats* stats) {
  const fsxd_options& o = s->opt;
  for (int n = 0; n < nsteps; n++) {
    if (s->steady) return n;
    double t0 = now_s();
    s->timeN++;
    s->t += s->dt;
    RUN_CHECK(fsx_begin_step(s->ctx));                                      // BCH:2865-2881
    int iter, pcg_total = 0, it = 0, noconv = 0;
    double rel = 0.0, rel_max = 0.0;
    for (iter = 1; iter <= o.maxIter; iter++) {                            // BCH:2885
      RUN_CHECK(fsx_convection(s->ctx));                                     // step1 (BCH:3702)
      RUN_CHECK(fsx_step1(s->ctx, iter));
      int rc = fsx_step2(s->ctx, iter, &it, &rel);                           // BCH:2898
      if (rc != FSX_OK && rc != FSX_ENOCONV) { s->err = fsx_last_error(s->ctx); return -rc; }
      if (rc == FSX_ENOCONV) noconv++;                                       // counted, not hidden
      pcg_total += it;
      rel_max = std::max(rel_max, rel);
      RUN_CHECK(fsx_step3(s->ctx, iter));                                    // BCH:2911
      int conv = 0;
      double ru, rp;
      RUN_CHECK(fsx_check_convergence(s->ctx, o.tolerance, &conv, &ru, &rp));  // BCH:2925
      if (conv) break;
      RUN_CHECK(fsx_next_iteration(s->ctx));                                 // BCH:2928-2932
      if (iter < o.maxIter) RUN_CHECK(fsx_kacc(s->ctx));                     // BCH:2982 (unused after the last iteration)
    }
    int steady = 0;
    double maxAcc = 0.0;
    RUN_CHECK(fsx_end_step(s->ctx, o.converge, &steady, &maxAcc));         // BCH:3046-3077
    double mon[4] = {0, 0, 0, 0};
    RUN_CHECK(fsx_monitor(s->ctx, s->topo->monPoint, mon));                // BCH:3088
    double wall = now_s() - t0;
    if (stats) {
      fsxd_step_stats& st = stats[n];
      st.step = s->timeN;
      st.inner_iters = iter > o.maxIter ? o.maxIter : iter;
      st.pcg_iters = pcg_total;
      st.steady = steady;
      st.time = s->t;
      st.max_acc = maxAcc;
      std::memcpy(st.mon, mon, sizeof mon);
      st.wall_s = wall;
      st.pcg_relres = rel;
      st.pcg_noconv = noconv;
      st.pad_ = 0;
      st.pcg_relres_max = rel_max;
    }
    if (steady) { s->steady = 1; return n + 1; }
  }
  return nsteps;
}

extern "C" int fsxd_reset(fsxd_solver* s) {
  DRV_CHECK(fsx_init_state(s->ctx));
  DRV_CHECK(fsx_sync(s->ctx));
  s->timeN = 0;
  s->t = s->t0;
  s->steady = 0;
  return FSX_OK;
}

extern "C" fsx_ctx* fsxd_ctx(fsxd_solver* s) { return s->ctx; }
extern "C" const fsxh_topo* fsxd_topo(fsxd_solver* s) { return s->topo; }
extern "C" const char* fsxd_error(fsxd_solver* s) { return s->err.c_str(); }
extern "C" int fsxd_setup_times(fsxd_solver* s, double* out3) {
  out3[0] = s->t_topo; out3[1] = s->t_upload; out3[2] = s->t_assemble;
  return FSX_OK;
}
