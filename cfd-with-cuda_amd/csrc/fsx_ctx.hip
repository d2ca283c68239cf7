// fsx_ctx.hip -- context lifecycle, uploads/downloads and the step-level
// C-ABI of the fractional-step core (include/fsx.h). The kernels live in
// fsx_assemble.hip (element work) and fsx_solve.hip (sparse, solver, steps).
#include "fsx_internal.h"
#include "fsx_ztensor.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <utility>

namespace fsx {

int dev_alloc(fsx_ctx* ctx, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    ctx->err = "hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e);
    (void)hipGetLastError();
    return FSX_ENOMEM;
  }
  ctx->allocs.push_back(*p);
  ctx->bytes += bytes;
  return FSX_OK;
}

void dev_free(fsx_ctx* ctx, void* p) {
  if (!p) return;
  auto it = std::find(ctx->allocs.begin(), ctx->allocs.end(), p);
  if (it != ctx->allocs.end()) {
    ctx->allocs.erase(it);
    (void)hipFree(p);
  }
}

int fail(fsx_ctx* ctx, int code, const std::string& msg) {
  ctx->err = msg;
  return code;
}

void iter_graph_reset(IterGraph& g) {
  for (int i = 0; i < 2; i++) {
    if (g.exec[i]) (void)hipGraphExecDestroy(g.exec[i]);
    if (g.graph[i]) (void)hipGraphDestroy(g.graph[i]);
    g.exec[i] = nullptr;
    g.graph[i] = nullptr;
  }
  g.key = 0;
}

int launch_check(fsx_ctx* ctx, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ctx, FSX_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return FSX_OK;
}

int row_grid(long long nrows, int rows_per_block) {
  long long g = (nrows + rows_per_block - 1) / rows_per_block;
  if (g > kMaxGrid) g = kMaxGrid;
  if (g < 1) g = 1;
  g = (g + 7) / 8 * 8;   // a multiple of the 8 XCDs
  return (int)g;
}

// Parametric position (-1, 0, +1) of each hex27 node (BCH:2312-2343) and the
// 1-D quadratic Lagrange factors used there: (x^2-x)/2, 1-x^2, (x^2+x)/2.
static const int kNodePos[27][3] = {
  {-1,-1,-1},{1,-1,-1},{1,1,-1},{-1,1,-1},{-1,-1,1},{1,-1,1},{1,1,1},{-1,1,1},
  {0,-1,-1},{1,0,-1},{0,1,-1},{-1,0,-1},
  {-1,-1,0},{1,-1,0},{1,1,0},{-1,1,0},
  {0,-1,1},{1,0,1},{0,1,1},{-1,0,1},
  {0,0,-1},{0,-1,0},{1,0,0},{0,1,0},{-1,0,0},{0,0,1},{0,0,0}};
static double lag(int a, double x) { return a < 0 ? 0.5 * (x*x - x) : (a > 0 ? 0.5 * (x*x + x) : (1 - x*x)); }
static double dlag(int a, double x) { return a < 0 ? 0.5 * (2*x - 1) : (a > 0 ? 0.5 * (2*x + 1) : (-2*x)); }

// Gauss rule (BCH:2166-2208; NGP 27 = tensor 3-point rule, xi fastest) and
// shape tables (BCH:2215-2488) in the GP-major layout of ShapeTables.
static void build_tables(int NGP, std::vector<double>& out) {
  std::vector<double> gp(NGP * 3, 0.0), gw(NGP, 0.0);
  if (NGP == 1) {
    gw[0] = 4.0;
  } else if (NGP == 8) {
    const double a = std::sqrt(1. / 3);
    for (int k = 0; k < 8; k++) {
      gp[3*k] = (k & 1) ? a : -a; gp[3*k+1] = (k & 2) ? a : -a; gp[3*k+2] = (k & 4) ? a : -a; gw[k] = 1.0;
    }
  } else {
    const double p[3] = {-std::sqrt(0.6), 0.0, std::sqrt(0.6)};
    const double w[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    for (int kz = 0; kz < 3; kz++) for (int ky = 0; ky < 3; ky++) for (int kx = 0; kx < 3; kx++) {
      int k = kx + 3*ky + 9*kz;
      gp[3*k] = p[kx]; gp[3*k+1] = p[ky]; gp[3*k+2] = p[kz]; gw[k] = w[kx] * w[ky] * w[kz];
    }
  }
  out.assign((size_t)NGP * (27 + 81 + 8 + 24 + 1), 0.0);
  double* Sv = out.data();
  double* dSv = Sv + NGP * 27;
  double* Sp = dSv + NGP * 81;
  double* dSp = Sp + NGP * 8;
  double* w = dSp + NGP * 24;
  for (int k = 0; k < NGP; k++) {
    double x = gp[3*k], y = gp[3*k+1], z = gp[3*k+2];
    for (int i = 0; i < 27; i++) {
      int a = kNodePos[i][0], b = kNodePos[i][1], c = kNodePos[i][2];
      Sv[k*27 + i] = lag(a, x) * lag(b, y) * lag(c, z);
      dSv[(k*27 + i)*3 + 0] = dlag(a, x) * lag(b, y) * lag(c, z);
      dSv[(k*27 + i)*3 + 1] = lag(a, x) * dlag(b, y) * lag(c, z);
      dSv[(k*27 + i)*3 + 2] = lag(a, x) * lag(b, y) * dlag(c, z);
    }
    for (int j = 0; j < 8; j++) {
      double sx = kNodePos[j][0], sy = kNodePos[j][1], sz = kNodePos[j][2];
      Sp[k*8 + j] = 0.125 * (1 + sx*x) * (1 + sy*y) * (1 + sz*z);
      dSp[(k*8 + j)*3 + 0] = 0.125 * sx * (1 + sy*y) * (1 + sz*z);
      dSp[(k*8 + j)*3 + 1] = 0.125 * (1 + sx*x) * sy * (1 + sz*z);
      dSp[(k*8 + j)*3 + 2] = 0.125 * (1 + sx*x) * (1 + sy*y) * sz;
    }
    w[k] = gw[k];
  }
}

// One wavefront copies the solve scalars to host-coherent memory, then
// releases the sequence number the host waits for.
__global__ void k_publish(const SolveScalars* __restrict__ sc, Published* pub, unsigned seq) {
  constexpr int nw = (int)(sizeof(SolveScalars) / sizeof(unsigned));
  const unsigned* src = reinterpret_cast<const unsigned*>(sc);
  volatile unsigned* dst = reinterpret_cast<volatile unsigned*>(&pub->s);
  for (int i = threadIdx.x; i < nw; i += blockDim.x) dst[i] = src[i];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) *reinterpret_cast<volatile unsigned*>(&pub->seq) = seq;
}

// Waits for the scalars with a host spin on the published sequence number
// (a copy plus hipStreamSynchronize costs tens of microseconds of idle GPU
// per solver decision). The stream is queried every few thousand spins so
// that a failed kernel or a lost publication ends the wait with an error.
int wait_published(fsx_ctx* ctx, unsigned seq, const char* what) {
  volatile unsigned* flag = &ctx->h_pub->seq;
  for (long spin = 1; *flag != seq; spin++) {
    if ((spin & 4095) == 0) {
      const hipError_t e = hipStreamQuery(ctx->stream);
      if (e != hipSuccess && e != hipErrorNotReady)
        return fail(ctx, FSX_EHIP, std::string(what) + ": " + hipGetErrorString(e));
      if (e == hipSuccess) {
        std::atomic_thread_fence(std::memory_order_seq_cst);
        if (*flag != seq) return fail(ctx, FSX_EHIP, std::string(what) + ": publication not visible after the stream drained");
      }
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  return FSX_OK;
}

int read_scalars(fsx_ctx* ctx, const SolveScalars* sc) {
  static_assert(sizeof(SolveScalars) % sizeof(unsigned) == 0, "SolveScalars is copied as 32-bit words");
  const unsigned seq = ++ctx->pub_seq;
  hipLaunchKernelGGL(k_publish, dim3(1), dim3(kWave), 0, ctx->stream, sc ? sc : ctx->sc, ctx->d_pub, seq);
  FSX_TRY(launch_check(ctx, "k_publish"));
  FSX_TRY(wait_published(ctx, seq, "solver scalars"));
  std::memcpy(ctx->h_sc, const_cast<const SolveScalars*>(&ctx->h_pub->s), sizeof(SolveScalars));
  return FSX_OK;
}

}  // namespace fsx

using namespace fsx;

extern "C" {

int fsx_create(int device, fsx_ctx** out) {
  if (!out) return FSX_EINVAL;
  *out = nullptr;
  auto* ctx = new fsx_ctx();
  ctx->device = device;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0 || device < 0 || device >= n) {
    delete ctx;
    return FSX_EHIP;
  }
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return FSX_EHIP;
  }
  int rc = alloc(ctx, &ctx->partials, 8 * (size_t)kMaxGrid);
  if (rc == FSX_OK) rc = alloc(ctx, &ctx->sc, 1);
  if (rc == FSX_OK) rc = alloc(ctx, &ctx->st, 1);
  if (rc == FSX_OK) rc = alloc(ctx, &ctx->d_mon, 4);
  if (rc == FSX_OK && hipHostMalloc((void**)&ctx->h_sc, sizeof(SolveScalars)) != hipSuccess) rc = FSX_ENOMEM;
  if (rc == FSX_OK && hipHostMalloc((void**)&ctx->h_pub, sizeof(Published), hipHostMallocMapped | hipHostMallocCoherent) !=
                          hipSuccess)
    rc = FSX_ENOMEM;
  if (rc == FSX_OK) {
    std::memset((void*)ctx->h_pub, 0, sizeof(Published));
    if (hipHostGetDevicePointer((void**)&ctx->d_pub, ctx->h_pub, 0) != hipSuccess) rc = FSX_EHIP;
  }
  if (rc == FSX_OK && hipHostMalloc((void**)&ctx->h_st, sizeof(StepScalars)) != hipSuccess) rc = FSX_ENOMEM;
  if (rc == FSX_OK && hipHostMalloc((void**)&ctx->h_mon, 4 * sizeof(double)) != hipSuccess) rc = FSX_ENOMEM;
  if (rc == FSX_OK && (hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess)) rc = FSX_EHIP;
  if (rc == FSX_OK) rc = scratch_setup(ctx);   // the device is current here
  if (rc == FSX_OK) {
    if (hipMemsetAsync(ctx->sc, 0, sizeof(SolveScalars), ctx->stream) != hipSuccess ||
        hipMemsetAsync(ctx->st, 0, sizeof(StepScalars), ctx->stream) != hipSuccess)
      rc = FSX_EHIP;
  }
  if (rc != FSX_OK) {
    fsx_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return FSX_OK;
}

void fsx_destroy(fsx_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  fsx::unet_free(ctx);
  fsx::mf_free(ctx);
  for (fsx::IterGraph* g : {&ctx->graph_fp64, &ctx->graph_mixed, &ctx->graph_learned}) fsx::iter_graph_reset(*g);
  for (void* p : ctx->allocs) (void)hipFree(p);
  if (ctx->h_sc) (void)hipHostFree(ctx->h_sc);
  if (ctx->h_pub) (void)hipHostFree(ctx->h_pub);
  if (ctx->h_st) (void)hipHostFree(ctx->h_st);
  if (ctx->h_mon) (void)hipHostFree(ctx->h_mon);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* fsx_last_error(const fsx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

size_t fsx_device_bytes(const fsx_ctx* ctx) { return ctx ? ctx->bytes : 0; }

int fsx_set_params(fsx_ctx* ctx, int NGP, double density, double viscosity, double dt) {
  if (!ctx) return FSX_EINVAL;
  if (NGP != 1 && NGP != 8 && NGP != 27) return fail(ctx, FSX_EINVAL, "NGP must be 1, 8 or 27");
  if (!(density > 0) || !(viscosity >= 0) || !(dt > 0)) return fail(ctx, FSX_EINVAL, "density, viscosity, dt");
  if (ctx->assembled && NGP != ctx->NGP) return fail(ctx, FSX_ESTATE, "NGP cannot change after assembly");
  std::vector<double> t;
  build_tables(NGP, t);
  if (ctx->tab_mem) dev_free(ctx, ctx->tab_mem);
  FSX_TRY(alloc(ctx, &ctx->tab_mem, t.size()));
  FSX_HIP(hipMemcpy(ctx->tab_mem, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
  ctx->tab.Sv = ctx->tab_mem;
  ctx->tab.dSv = ctx->tab.Sv + NGP * 27;
  ctx->tab.Sp = ctx->tab.dSv + NGP * 81;
  ctx->tab.dSp = ctx->tab.Sp + NGP * 8;
  ctx->tab.w = ctx->tab.dSp + NGP * 24;
  {
    std::vector<float> tf(t.begin(), t.end());
    if (ctx->tabf_mem) dev_free(ctx, ctx->tabf_mem);
    FSX_TRY(alloc(ctx, &ctx->tabf_mem, tf.size()));
    FSX_HIP(hipMemcpy(ctx->tabf_mem, tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice));
    ctx->tabf.Sv = ctx->tabf_mem;
    ctx->tabf.dSv = ctx->tabf.Sv + NGP * 27;
    ctx->tabf.Sp = ctx->tabf.dSv + NGP * 81;
    ctx->tabf.dSp = ctx->tabf.Sp + NGP * 8;
    ctx->tabf.w = ctx->tabf.dSp + NGP * 24;
  }
  FSX_TRY(elem_tables_setup(ctx, t.data(), NGP));
  ctx->NGP = NGP; ctx->density = density; ctx->viscosity = viscosity; ctx->dt = dt;
  ctx->have_params = true;
  return FSX_OK;
}

static int build_z_forms(fsx_ctx* ctx);

int fsx_set_matrix_precision(fsx_ctx* ctx, int bits) {
  if (!ctx) return FSX_EINVAL;
  if (bits != 32 && bits != 64) return fail(ctx, FSX_EINVAL, "fsx_set_matrix_precision: 32 or 64");
  ctx->mat_bits = bits;
  ctx->gpn_valid = ctx->gpn_next_valid = false;   // G Pn is formed again from the values now in use
  return refresh_matrix_copies(ctx);
}

int fsx_set_matrix_classes(fsx_ctx* ctx, int on) {
  if (!ctx) return FSX_EINVAL;
  ctx->use_g_classes = on != 0;
  ctx->gpn_valid = ctx->gpn_next_valid = false;
  return refresh_matrix_copies(ctx);
}

int fsx_set_z_tensor(fsx_ctx* ctx, int on) {
  if (!ctx) return FSX_EINVAL;
  ctx->use_z_tensor = on != 0;
  if (!ctx->z_built) return FSX_OK;
  return build_z_forms(ctx);   // the tiles again, in the form now asked for
}

int fsx_z_tensor_info(fsx_ctx* ctx, int* in_use, double* err) {
  if (!ctx) return FSX_EINVAL;
  if (in_use) *in_use = ctx->tile.ntiles > 0 && ctx->tile.tensor ? 1 : 0;
  if (err) *err = ctx->z_tensor_err;
  return FSX_OK;
}

int fsx_g_class_info(fsx_ctx* ctx, int* classes, long long* entries, int* t_classes, long long* t_entries) {
  if (!ctx) return FSX_EINVAL;
  const bool on = g_classes_on(ctx);
  if (classes) *classes = on ? ctx->gc_ncls : 0;
  if (entries) *entries = on ? ctx->gc_nent : 0;
  if (t_classes) *t_classes = on ? ctx->gtc_ncls : 0;
  if (t_entries) *t_entries = on ? ctx->gtc_nent : 0;
  return FSX_OK;
}

int fsx_set_solver_x0(fsx_ctx* ctx, int policy) {
  if (!ctx) return FSX_EINVAL;
  if (policy != FSX_X0_ZERO && policy != FSX_X0_PREVIOUS && policy != FSX_X0_PROJECTION)
    return fail(ctx, FSX_EINVAL, "fsx_set_solver_x0: unknown policy");
  ctx->x0_policy = policy;
  return FSX_OK;
}

int fsx_set_convection_precision(fsx_ctx* ctx, int bits) {
  if (!ctx) return FSX_EINVAL;
  if (bits != 32 && bits != 64) return fail(ctx, FSX_EINVAL, "fsx_set_convection_precision: 32 or 64");
  ctx->conv_bits = bits;
  return FSX_OK;
}

int fsx_set_convection_form(fsx_ctx* ctx, int elements) {
  if (!ctx) return FSX_EINVAL;
  ctx->conv_elements = elements != 0;
  return FSX_OK;
}

int fsx_set_step1_window(fsx_ctx* ctx, int enable) {
  if (!ctx) return FSX_EINVAL;
  ctx->use_step1_window = enable != 0;
  return FSX_OK;
}

int fsx_set_step1_elements(fsx_ctx* ctx, int enable) {
  if (!ctx) return FSX_EINVAL;
  FSX_TRY(settle_deferred(ctx));
  ctx->use_step1_elements = enable != 0;
  return FSX_OK;
}

int fsx_set_graphs(fsx_ctx* ctx, int enable) {
  if (!ctx) return FSX_EINVAL;
  ctx->use_graphs = enable != 0;
  return FSX_OK;
}

int fsx_set_solver(fsx_ctx* ctx, int mode, double rtol, int maxit) {
  if (!ctx) return FSX_EINVAL;
  if (mode < FSX_SOLVER_FP64_JACOBI || mode > FSX_SOLVER_DIRECT) return fail(ctx, FSX_EINVAL, "unknown solver mode");
  if (mode == FSX_SOLVER_LEARNED && !ctx->unet) return fail(ctx, FSX_ESTATE, "learned mode needs fsx_unet_setup first");
  if (!(rtol >= 0) || maxit < 1) return fail(ctx, FSX_EINVAL, "rtol >= 0 and maxit >= 1 required");
  ctx->solver_mode = mode; ctx->rtol = rtol; ctx->maxit = maxit;
  return FSX_OK;
}

int fsx_upload_mesh(fsx_ctx* ctx, int NE, int NN, int NNp, const int* LtoGnode, const double* corner_coord,
                    int ncolors, const int* color_ptr, const int* color_order) {
  if (!ctx || NE <= 0 || NN <= 0 || NNp <= 0 || !LtoGnode || !corner_coord || ncolors <= 0 || !color_ptr || !color_order)
    return ctx ? fail(ctx, FSX_EINVAL, "fsx_upload_mesh: bad arguments") : FSX_EINVAL;
  if (ctx->have_mesh) return fail(ctx, FSX_ESTATE, "mesh already uploaded");
  // host-side validation: ids in range, corners are pressure nodes, the
  // colour table is a permutation of the elements
  for (size_t i = 0; i < (size_t)NE * 27; i++) {
    int n = LtoGnode[i];
    if (n < 0 || n >= NN || ((i % 27) < 8 && n >= NNp)) return fail(ctx, FSX_EINVAL, "LtoGnode entry out of range");
  }
  if (color_ptr[0] != 0 || color_ptr[ncolors] != NE) return fail(ctx, FSX_EINVAL, "colour offsets");
  {
    std::vector<char> seen(NE, 0);
    for (int i = 0; i < NE; i++) {
      int e = color_order[i];
      if (e < 0 || e >= NE || seen[e]) return fail(ctx, FSX_EINVAL, "colour order is not a permutation");
      seen[e] = 1;
    }
    // no two elements of one colour share a node
    std::vector<int> owner(NN, -1);
    for (int c = 0; c < ncolors; c++) {
      if (color_ptr[c + 1] < color_ptr[c]) return fail(ctx, FSX_EINVAL, "colour offsets");
      for (int i = color_ptr[c]; i < color_ptr[c + 1]; i++) {
        int e = color_order[i];
        for (int j = 0; j < 27; j++) {
          int n = LtoGnode[(size_t)e*27 + j];
          if (owner[n] == c) return fail(ctx, FSX_EINVAL, "two elements of one colour share a node");
          owner[n] = c;
        }
      }
    }
  }
  ctx->NE = NE; ctx->NN = NN; ctx->NNp = NNp;
  // Structured pressure grid: every element's corner ids differ by 1, nx and
  // nx*ny along its x, y and z edges (local corners 0-1, 0-3, 0-4, BCH:2312-
  // 2343), as the structured generators number them (x fastest). The Z
  // products of the solvers then stage the grid through LDS tiles.
  ctx->pg_nx = ctx->pg_ny = ctx->pg_nz = 0;
  {
    const int nx = LtoGnode[3] - LtoGnode[0], nxy = LtoGnode[4] - LtoGnode[0];
    bool ok = nx > 1 && nxy > nx && nxy % nx == 0 && NNp % nxy == 0;
    for (int e = 0; ok && e < NE; e++) {
      const int* L = LtoGnode + (size_t)e * 27;
      const int c = L[0];
      ok = L[1] - c == 1 && L[2] - c == 1 + nx && L[3] - c == nx && L[4] - c == nxy && L[5] - c == 1 + nxy &&
           L[6] - c == 1 + nx + nxy && L[7] - c == nx + nxy;
    }
    if (ok) { ctx->pg_nx = nx; ctx->pg_ny = nxy / nx; ctx->pg_nz = NNp / nxy; }
  }
  // internal velocity numbering: the nodes in the order of their coordinates
  // (z, y, x), each node placed by the trilinear map of the first element
  // that holds it, coordinates compared on a grid of 1e-9 of the mesh extent.
  // (Ordering each grid line even-ranked x first, so that a wavefront's rows
  // have equal lengths, measured no faster.)
  {
    std::vector<double> X((size_t)NN * 3, 0.0);
    std::vector<char> have(NN, 0);
    for (int e = 0; e < NE; e++) {
      const int* L = LtoGnode + (size_t)e * 27;
      for (int j = 0; j < 27; j++) {
        const int nd = L[j];
        if (have[nd]) continue;
        have[nd] = 1;
        for (int m = 0; m < 8; m++) {
          const double w = 0.125 * (1 + kNodePos[m][0] * kNodePos[j][0]) * (1 + kNodePos[m][1] * kNodePos[j][1]) *
                           (1 + kNodePos[m][2] * kNodePos[j][2]);
          for (int d = 0; d < 3; d++) X[3 * (size_t)nd + d] += w * corner_coord[3 * (size_t)L[m] + d];
        }
      }
    }
    double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (int nd = 0; nd < NN; nd++)
      for (int d = 0; d < 3; d++) { lo[d] = std::min(lo[d], X[3 * (size_t)nd + d]); hi[d] = std::max(hi[d], X[3 * (size_t)nd + d]); }
    const double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    const double h = ext > 0 ? ext * 1e-9 : 1.0;
    struct Key { long long z, y, x; int id; };
    std::vector<Key> key(NN);
    for (int nd = 0; nd < NN; nd++)
      key[nd] = Key{std::llround((X[3 * (size_t)nd + 2] - lo[2]) / h), std::llround((X[3 * (size_t)nd + 1] - lo[1]) / h),
                    std::llround((X[3 * (size_t)nd] - lo[0]) / h), nd};
    std::sort(key.begin(), key.end(), [](const Key& a, const Key& b) {
      if (a.z != b.z) return a.z < b.z;
      if (a.y != b.y) return a.y < b.y;
      if (a.x != b.x) return a.x < b.x;
      return a.id < b.id;
    });
    ctx->viperm.resize(NN);
    ctx->vperm.resize(NN);
    for (int k = 0; k < NN; k++) { ctx->viperm[k] = key[k].id; ctx->vperm[key[k].id] = k; }
  }
  {
    std::vector<int> li((size_t)NE * 27), p8((size_t)NE * 8);
    for (size_t i = 0; i < (size_t)NE * 27; i++) li[i] = ctx->vperm[LtoGnode[i]];
    for (int e = 0; e < NE; e++)
      for (int m = 0; m < 8; m++) p8[(size_t)e * 8 + m] = LtoGnode[(size_t)e * 27 + m];
    FSX_TRY(alloc(ctx, &ctx->L27, (size_t)NE * 27));
    FSX_TRY(alloc(ctx, &ctx->P8, (size_t)NE * 8));
    FSX_HIP(hipMemcpy(ctx->L27, li.data(), li.size() * sizeof(int), hipMemcpyHostToDevice));
    FSX_HIP(hipMemcpy(ctx->P8, p8.data(), p8.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  FSX_TRY(alloc(ctx, &ctx->xc, (size_t)NNp * 3));
  FSX_TRY(alloc(ctx, &ctx->color_order, (size_t)NE));
  FSX_HIP(hipMemcpy(ctx->xc, corner_coord, (size_t)NNp * 3 * sizeof(double), hipMemcpyHostToDevice));
  FSX_HIP(hipMemcpy(ctx->color_order, color_order, (size_t)NE * sizeof(int), hipMemcpyHostToDevice));
  ctx->color_ptr.assign(color_ptr, color_ptr + ncolors + 1);
  // state vectors
  const size_t n3 = (size_t)3 * NN;
  for (double** v : {&ctx->Un, &ctx->Unp1, &ctx->Unp1_prev, &ctx->UnpHalf, &ctx->UnpHalf_prev, &ctx->Acc, &ctx->Acc_prev,
                     &ctx->KAcc, &ctx->R1, &ctx->GPn, &ctx->GPn_next})
    FSX_TRY(alloc(ctx, v, n3));
  for (double** v : {&ctx->Pn, &ctx->Pnp1, &ctx->Pnp1_prev, &ctx->Pdot, &ctx->R2, &ctx->wr, &ctx->wz, &ctx->wp, &ctx->wq,
                     &ctx->wh, &ctx->w0r, &ctx->w0e})
    FSX_TRY(alloc(ctx, v, (size_t)NNp));
  for (float** v : {&ctx->fr, &ctx->fz, &ctx->fp, &ctx->fq, &ctx->fx}) FSX_TRY(alloc(ctx, v, (size_t)NNp));
  FSX_TRY(alloc(ctx, &ctx->MdInv, (size_t)NN));
  FSX_TRY(alloc(ctx, &ctx->MdOrigInv, (size_t)NN));
  FSX_TRY(alloc(ctx, &ctx->Md, (size_t)NN));
  FSX_TRY(alloc(ctx, &ctx->bcflag, (size_t)NN));
  FSX_HIP(hipMemsetAsync(ctx->bcflag, 0, NN, ctx->stream));
  ctx->have_mesh = true;
  return FSX_OK;
}

// Dictionary of the distinct per-row lists of a CSR pattern: each row's
// column offsets (col - row) and, when `vbits` is given, the 32-bit words of
// its values. Structured meshes have a handful of lists (interior rows all
// share one), so the SpMVs can read a pattern id per row instead of a
// column (and value) per entry. Returns false when there are more than
// max_pat distinct lists or they hold more than max_len entries in total.
static bool build_row_dict(int nrows, const int* rowptr, const int* col, const unsigned* vbits, int max_pat,
                           long long max_len, std::vector<int>& row_pat, std::vector<int>& pat_ptr,
                           std::vector<int>& pat_off, std::vector<unsigned>* pat_val) {
  std::unordered_map<unsigned long long, std::vector<int>> by_hash;
  row_pat.assign(nrows, 0);
  pat_ptr.assign(1, 0);
  pat_off.clear();
  if (pat_val) pat_val->clear();
  for (int r = 0; r < nrows; r++) {
    const int rs = rowptr[r], len = rowptr[r + 1] - rs;
    unsigned long long h = 1469598103934665603ull ^ (unsigned long long)len;
    for (int k = 0; k < len; k++) {
      h = (h ^ (unsigned long long)(unsigned)(col[rs + k] - r)) * 1099511628211ull;
      if (vbits) h = (h ^ (unsigned long long)vbits[rs + k]) * 1099511628211ull;
    }
    int id = -1;
    auto& cand = by_hash[h];
    for (int c : cand) {
      const int cs = pat_ptr[c], clen = pat_ptr[c + 1] - cs;
      if (clen != len) continue;
      bool same = true;
      for (int k = 0; k < len && same; k++)
        same = pat_off[cs + k] == col[rs + k] - r && (!vbits || (*pat_val)[cs + k] == vbits[rs + k]);
      if (same) { id = c; break; }
    }
    if (id < 0) {
      id = (int)pat_ptr.size() - 1;
      if (id >= max_pat || (long long)pat_off.size() + len > max_len) return false;
      for (int k = 0; k < len; k++) {
        pat_off.push_back(col[rs + k] - r);
        if (vbits) pat_val->push_back(vbits[rs + k]);
      }
      pat_ptr.push_back((int)pat_off.size());
      cand.push_back(id);
    }
    row_pat[r] = id;
  }
  return true;
}

static bool build_patterns(int nrows, const int* rowptr, const int* col, int max_pat, long long max_len,
                           std::vector<int>& row_pat, std::vector<int>& pat_ptr, std::vector<int>& pat_off) {
  return build_row_dict(nrows, rowptr, col, nullptr, max_pat, max_len, row_pat, pat_ptr, pat_off, nullptr);
}

// The solvers' forms of the built Z (structural zeros dropped): FP64 rows
// with a pattern dictionary, and FP32 row classes when they compress.
// Runs once per fsx_build_Z on host copies of Z; captured solver graphs that
// name the previous forms are dropped.
static int build_z_forms(fsx_ctx* ctx) {
  const int n = ctx->NNp;
  const long long nnz = ctx->nnzZ;
  std::vector<int> rp(n + 1), ci(nnz);
  std::vector<double> v(nnz);
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  FSX_HIP(hipMemcpy(rp.data(), ctx->Zrow, (size_t)(n + 1) * sizeof(int), hipMemcpyDeviceToHost));
  FSX_HIP(hipMemcpy(ci.data(), ctx->Zcol, (size_t)nnz * sizeof(int), hipMemcpyDeviceToHost));
  FSX_HIP(hipMemcpy(v.data(), ctx->Z, (size_t)nnz * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<int> rp2(n + 1, 0), ci2;
  std::vector<double> v2;
  std::vector<unsigned> f2;
  ci2.reserve(nnz); v2.reserve(nnz); f2.reserve(nnz);
  for (int r = 0; r < n; r++) {
    double m = 0.0;
    for (int k = rp[r]; k < rp[r + 1]; k++) m = std::max(m, std::fabs(v[k]));
    const double thr = m * kPruneRel;
    for (int k = rp[r]; k < rp[r + 1]; k++) {
      if (!(std::fabs(v[k]) >= thr) && ci[k] != r) continue;
      ci2.push_back(ci[k]);
      v2.push_back(v[k]);
      const float f = (float)v[k];
      unsigned u;
      std::memcpy(&u, &f, sizeof u);
      f2.push_back(u);
    }
    rp2[r + 1] = (int)ci2.size();
  }
  for (fsx::IterGraph* g : {&ctx->graph_fp64, &ctx->graph_mixed, &ctx->graph_learned}) iter_graph_reset(*g);
  for (void* p : {(void*)ctx->Zr_row, (void*)ctx->Zr_val, (void*)ctx->Zr_valf, (void*)ctx->Zr_col, (void*)ctx->Zr_pat, (void*)ctx->Zr_off,
                  (void*)ctx->Zc_ent, (void*)ctx->Zc_chunk, (void*)ctx->Zc_rows, (void*)ctx->Zc_delta,
                  (void*)ctx->Zc_doff})
    if (p) dev_free(ctx, p);
  ctx->Zr_row = nullptr; ctx->Zr_val = nullptr; ctx->Zr_valf = nullptr; ctx->Zr_col = nullptr; ctx->Zr_pat = nullptr;
  ctx->Zr_off = nullptr;
  ctx->Zc_ent = nullptr; ctx->Zc_chunk = nullptr; ctx->Zc_rows = nullptr; ctx->Zc_delta = nullptr; ctx->Zc_doff = nullptr;
  ctx->Zr_npat = 0; ctx->Zr_patlen = 0; ctx->Zc_n = 0; ctx->Zc_len = 0; ctx->Zc_nchunk = 0;
  const long long nnz2 = rp2[n];
  ctx->nnzZr = nnz2;
  FSX_TRY(alloc(ctx, &ctx->Zr_row, (size_t)n + 1));
  FSX_TRY(alloc(ctx, &ctx->Zr_val, (size_t)std::max(1LL, nnz2)));
  FSX_HIP(hipMemcpy(ctx->Zr_row, rp2.data(), (size_t)(n + 1) * sizeof(int), hipMemcpyHostToDevice));
  FSX_HIP(hipMemcpy(ctx->Zr_val, v2.data(), (size_t)nnz2 * sizeof(double), hipMemcpyHostToDevice));
  FSX_TRY(alloc(ctx, &ctx->Zr_valf, (size_t)std::max(1LL, nnz2)));
  static_assert(sizeof(unsigned) == sizeof(float), "FP32 bits");
  FSX_HIP(hipMemcpy(ctx->Zr_valf, f2.data(), (size_t)nnz2 * sizeof(float), hipMemcpyHostToDevice));
  std::vector<int> rowpat, pp, po;
  if (build_patterns(n, rp2.data(), ci2.data(), 1 << 16, 1 << 22, rowpat, pp, po)) {
    for (int& id : rowpat) id = pp[id];
    FSX_TRY(alloc(ctx, &ctx->Zr_pat, (size_t)n));
    FSX_TRY(alloc(ctx, &ctx->Zr_off, std::max<size_t>(1, po.size())));
    FSX_HIP(hipMemcpy(ctx->Zr_pat, rowpat.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    if (!po.empty()) FSX_HIP(hipMemcpy(ctx->Zr_off, po.data(), po.size() * sizeof(int), hipMemcpyHostToDevice));
    ctx->Zr_npat = (int)pp.size() - 1;
    ctx->Zr_patlen = (long long)po.size();
  } else {
    FSX_TRY(alloc(ctx, &ctx->Zr_col, (size_t)std::max(1LL, nnz2)));
    FSX_HIP(hipMemcpy(ctx->Zr_col, ci2.data(), (size_t)nnz2 * sizeof(int), hipMemcpyHostToDevice));
  }
  std::vector<int> cls, cp, co;
  std::vector<unsigned> cv;
  const bool classes = build_row_dict(n, rp2.data(), ci2.data(), f2.data(), 65535, 1 << 20, cls, cp, co, &cv);
  // LDS tiles of a structured pressure grid where Z's FP32 rows form classes:
  // Z applied as Kronecker products of 1D stencils where the grid is a
  // tensor-product box grid and that form reproduces Z's rows (to 1e-10 of
  // each row's largest), else by the classes. (Without classes, as on the
  // sinh-clustered 64^3 grid, there are no tiles and the learned solver
  // runs its additive form: there the two-level form's Jacobi smoothing
  // took twice the iterations, 51 against 25 a solve.)
  ctx->tile = TilePlan{};
  ctx->z_tensor_err = -1.0;
  if (ctx->zt64) dev_free(ctx, ctx->zt64);
  ctx->zt64 = nullptr;
  if (ctx->pg_nx > 0) {
    std::vector<double> xc(3 * (size_t)n);
    FSX_HIP(hipMemcpy(xc.data(), ctx->xc, xc.size() * sizeof(double), hipMemcpyDeviceToHost));
    ZTensor zt;
    if (z_tensor_tables(ctx->pg_nx, ctx->pg_ny, ctx->pg_nz, xc.data(), ctx->NGP, ctx->density, &zt) == 0) {
      const int held = z_tensor_check(n, rp2.data(), ci2.data(), v2.data(), ctx->zeroP, 1e-10, &zt);
      ctx->z_tensor_err = zt.err;
      if (held == 0 && ctx->use_z_tensor) {
        if (classes) FSX_TRY(build_tile_plan_tensor(ctx, zt));
        // the FP64 tables for the mixed and learned FP64 residuals
        std::vector<double> c64;
        for (int d = 0; d < 3; d++)
          for (int i = 0; i < zt.n[d]; i++) {
            for (int c = 0; c < 5; c++) c64.push_back(zt.k[d][5 * (size_t)i + c]);
            for (int c = 0; c < 5; c++) c64.push_back(zt.m[d][5 * (size_t)i + c]);
          }
        FSX_TRY(alloc(ctx, &ctx->zt64, c64.size()));
        FSX_HIP(hipMemcpy(ctx->zt64, c64.data(), c64.size() * sizeof(double), hipMemcpyHostToDevice));
        for (int d = 0; d < 3; d++) ctx->zt_n[d] = zt.n[d];
        ctx->zt_zp = zt.zero_p;
        ctx->zt_pen = zt.pen;
      }
    }
  }
  if (ctx->tile.ntiles == 0) {
    if (classes) FSX_TRY(build_tile_plan(ctx, cls, cp, co, cv));
    else FSX_TRY(build_tile_plan(ctx, {}, {}, {}, {}));
  }
  if (classes) {
    std::vector<int2> ent(co.size());
    for (size_t k = 0; k < co.size(); k++) ent[k] = make_int2(co[k], (int)cv[k]);
    FSX_TRY(alloc(ctx, &ctx->Zc_ent, std::max<size_t>(1, ent.size())));
    if (!ent.empty()) FSX_HIP(hipMemcpy(ctx->Zc_ent, ent.data(), ent.size() * sizeof(int2), hipMemcpyHostToDevice));
    ctx->Zc_n = (int)cp.size() - 1;
    ctx->Zc_len = (long long)ent.size();
    // launch plan: the rows of each class in ascending order, cut into
    // chunks of <= 64 rows; chunks ordered by their first row
    std::vector<std::vector<int>> members(cp.size() - 1);
    for (int r = 0; r < n; r++) members[cls[r]].push_back(r);
    std::vector<int4> chunks;
    std::vector<int> rows;
    rows.reserve(n);
    for (size_t c = 0; c < members.size(); c++)
      for (size_t s0 = 0; s0 < members[c].size(); s0 += 64) {
        const int len = (int)std::min<size_t>(64, members[c].size() - s0);
        chunks.push_back(make_int4((int)rows.size(), len, cp[c], cp[c + 1] - cp[c]));
        rows.insert(rows.end(), members[c].begin() + s0, members[c].begin() + s0 + len);
      }
    std::stable_sort(chunks.begin(), chunks.end(),
                     [&](const int4& a, const int4& b) { return rows[a.x] < rows[b.x]; });
    FSX_TRY(alloc(ctx, &ctx->Zc_chunk, std::max<size_t>(1, chunks.size())));
    FSX_TRY(alloc(ctx, &ctx->Zc_rows, (size_t)n));
    FSX_HIP(hipMemcpy(ctx->Zc_chunk, chunks.data(), chunks.size() * sizeof(int4), hipMemcpyHostToDevice));
    FSX_HIP(hipMemcpy(ctx->Zc_rows, rows.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    ctx->Zc_nchunk = (int)chunks.size();
    // the FP64 corrections: a class entry k is entry k of each member row
    // (same column offsets in the same order); rows that differ keep no split
    std::vector<long long> doff(chunks.size() + 1, 0);
    for (size_t c = 0; c < chunks.size(); c++) doff[c + 1] = doff[c] + 64LL * chunks[c].w;
    std::vector<float> delta((size_t)doff.back(), 0.0f);
    bool same = true;
    for (size_t c = 0; c < chunks.size() && same; c++) {
      const int4 ch = chunks[c];
      for (int l = 0; l < ch.y && same; l++) {
        const int r = rows[ch.x + l];
        if (rp2[r + 1] - rp2[r] != ch.w) { same = false; break; }
        for (int k = 0; k < ch.w; k++) {
          const int kk = rp2[r] + k;
          if (ci2[kk] - r != co[ch.z + k]) { same = false; break; }
          const double v = v2[kk];
          delta[(size_t)doff[c] + (size_t)k * 64 + l] = (float)(v - (double)(float)v);
        }
      }
    }
    if (same) {
      FSX_TRY(alloc(ctx, &ctx->Zc_delta, std::max<size_t>(1, delta.size())));
      FSX_TRY(alloc(ctx, &ctx->Zc_doff, chunks.size()));
      if (!delta.empty())
        FSX_HIP(hipMemcpy(ctx->Zc_delta, delta.data(), delta.size() * sizeof(float), hipMemcpyHostToDevice));
      FSX_HIP(hipMemcpy(ctx->Zc_doff, doff.data(), chunks.size() * sizeof(long long), hipMemcpyHostToDevice));
    }
  }
  return FSX_OK;
}

int fsx_upload_csr(fsx_ctx* ctx, int kind, int nrows, const int* rowptr, const int* col) {
  if (!ctx || !rowptr || !col) return ctx ? fail(ctx, FSX_EINVAL, "fsx_upload_csr: null") : FSX_EINVAL;
  if (!ctx->have_mesh) return fail(ctx, FSX_ESTATE, "upload the mesh first");
  int ncols;
  if (kind == FSX_CSR_M) { if (nrows != ctx->NN) return fail(ctx, FSX_EINVAL, "M rows != NN"); ncols = ctx->NN; }
  else if (kind == FSX_CSR_G) { if (nrows != ctx->NN) return fail(ctx, FSX_EINVAL, "G rows != NN"); ncols = ctx->NNp; }
  else if (kind == FSX_CSR_Z) { if (nrows != ctx->NNp) return fail(ctx, FSX_EINVAL, "Z rows != NNp"); ncols = ctx->NNp; }
  else return fail(ctx, FSX_EINVAL, "unknown CSR kind");
  if (rowptr[0] != 0) return fail(ctx, FSX_EINVAL, "rowptr[0] != 0");
  int rmax = 0;
  for (int r = 0; r < nrows; r++) {
    if (rowptr[r + 1] < rowptr[r]) return fail(ctx, FSX_EINVAL, "rowptr not monotone");
    rmax = std::max(rmax, rowptr[r + 1] - rowptr[r]);
    for (int k = rowptr[r]; k < rowptr[r + 1]; k++) {
      if (col[k] < 0 || col[k] >= ncols) return fail(ctx, FSX_EINVAL, "column out of range");
      if (k > rowptr[r] && col[k] <= col[k - 1]) return fail(ctx, FSX_EINVAL, "columns not sorted/unique");
    }
  }
  const long long nnz = rowptr[nrows];
  // M (rows and columns) and G (rows) go to the device in the internal
  // velocity numbering, rows in internal order with sorted columns; the
  // reference pattern is kept on the host for value downloads
  std::vector<int> irow, icol;
  if (kind == FSX_CSR_M || kind == FSX_CSR_G) {
    const std::vector<int>& vp = ctx->vperm;
    irow.assign((size_t)nrows + 1, 0);
    for (int r = 0; r < nrows; r++) irow[vp[r] + 1] = rowptr[r + 1] - rowptr[r];
    for (int r = 0; r < nrows; r++) irow[r + 1] += irow[r];
    icol.resize((size_t)nnz);
    for (int r = 0; r < nrows; r++) {
      int* dst = icol.data() + irow[vp[r]];
      const int len = rowptr[r + 1] - rowptr[r];
      for (int k = 0; k < len; k++) dst[k] = kind == FSX_CSR_M ? vp[col[rowptr[r] + k]] : col[rowptr[r] + k];
      if (kind == FSX_CSR_M) std::sort(dst, dst + len);
    }
    if (kind == FSX_CSR_M) {
      ctx->refMrow.assign(rowptr, rowptr + nrows + 1);
      ctx->refMcol.assign(col, col + nnz);
      ctx->posM.clear();
    } else {
      ctx->refGrow.assign(rowptr, rowptr + nrows + 1);
      ctx->refGcol.assign(col, col + nnz);
      ctx->posG.clear();
    }
    rowptr = irow.data();
    col = icol.data();
  }
  int* drow; int* dcol;
  FSX_TRY(alloc(ctx, &drow, (size_t)nrows + 1));
  FSX_TRY(alloc(ctx, &dcol, (size_t)nnz));
  FSX_HIP(hipMemcpy(drow, rowptr, ((size_t)nrows + 1) * sizeof(int), hipMemcpyHostToDevice));
  FSX_HIP(hipMemcpy(dcol, col, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice));
  if (kind == FSX_CSR_M) {
    ctx->Mrow = drow; ctx->Mcol = dcol; ctx->nnzM = nnz;
    FSX_TRY(alloc(ctx, &ctx->K, (size_t)nnz));
    FSX_TRY(alloc(ctx, &ctx->M, (size_t)nnz));
    // K's columns as a dictionary of row patterns: in the coordinate order a
    // structured mesh has a few hundred distinct rows
    std::vector<int> rp, pp, po;
    if (build_patterns(nrows, rowptr, col, 1 << 16, 1 << 22, rp, pp, po)) {
      for (int& id : rp) id = pp[id];
      FSX_TRY(alloc(ctx, &ctx->Mrowpat, rp.size()));
      FSX_TRY(alloc(ctx, &ctx->Mpatoff, std::max<size_t>(1, po.size())));
      FSX_HIP(hipMemcpy(ctx->Mrowpat, rp.data(), rp.size() * sizeof(int), hipMemcpyHostToDevice));
      if (!po.empty()) FSX_HIP(hipMemcpy(ctx->Mpatoff, po.data(), po.size() * sizeof(int), hipMemcpyHostToDevice));
      ctx->Mnpat = (int)pp.size() - 1;
      ctx->Mpatlen = (long long)po.size();
      // the LDS window's slots, when every offset is a (dx, dy, dz) step of
      // at most 2 on the velocity grid of a structured mesh
      if (ctx->pg_nx > 0 && !po.empty()) {
        const long long vnx = 2LL * ctx->pg_nx - 1, vnxy = vnx * (2LL * ctx->pg_ny - 1);
        std::vector<int> slot(po.size());
        bool ok = vnx > 4;
        for (size_t k = 0; ok && k < po.size(); k++) {
          const long long o = po[k];
          const long long dz = (o + 2 * vnxy * 3 + vnxy / 2) / vnxy - 6;   // nearest multiple of vnxy
          const long long rz = o - dz * vnxy;
          const long long dy = (rz + 2 * vnx * 3 + vnx / 2) / vnx - 6;
          const long long dx = rz - dy * vnx;
          ok = dz >= -2 && dz <= 2 && dy >= -2 && dy <= 2 && dx >= -2 && dx <= 2;
          slot[k] = (int)(2 + dx + 36 * ((dz + 2) * 5 + (dy + 2)));
        }
        if (ok) {
          FSX_TRY(alloc(ctx, &ctx->Mpatslot, slot.size()));
          FSX_HIP(hipMemcpy(ctx->Mpatslot, slot.data(), slot.size() * sizeof(int), hipMemcpyHostToDevice));
          ctx->vnx = (int)vnx;
          ctx->vnxy = vnxy;
        }
      }
    }
    ctx->have_M = true;
  } else if (kind == FSX_CSR_G) {
    ctx->Grow = drow; ctx->Gcol = dcol; ctx->nnzG = nnz;
    FSX_TRY(alloc(ctx, &ctx->G, (size_t)nnz * 3));
    // transposed pattern: rows = pressure nodes, entries in ascending velocity row
    const int NNp = ctx->NNp;
    std::vector<int> trow(NNp + 1, 0), tcol(nnz), tpos(nnz);
    for (long long k = 0; k < nnz; k++) trow[col[k] + 1]++;
    for (int p = 0; p < NNp; p++) trow[p + 1] += trow[p];
    std::vector<int> next(trow.begin(), trow.end() - 1);
    for (int r = 0; r < nrows; r++)
      for (int k = rowptr[r]; k < rowptr[r + 1]; k++) { int p = col[k]; tcol[next[p]] = r; tpos[next[p]] = k; next[p]++; }
    FSX_TRY(alloc(ctx, &ctx->GTrow, (size_t)NNp + 1));
    FSX_TRY(alloc(ctx, &ctx->GTcol, (size_t)nnz));
    FSX_TRY(alloc(ctx, &ctx->GTpos, (size_t)nnz));
    FSX_TRY(alloc(ctx, &ctx->GT, (size_t)nnz * 3));
    FSX_HIP(hipMemcpy(ctx->GTrow, trow.data(), ((size_t)NNp + 1) * sizeof(int), hipMemcpyHostToDevice));
    FSX_HIP(hipMemcpy(ctx->GTcol, tcol.data(), (size_t)nnz * sizeof(int), hipMemcpyHostToDevice));
    FSX_HIP(hipMemcpy(ctx->GTpos, tpos.data(), (size_t)nnz * sizeof(int), hipMemcpyHostToDevice));
    ctx->have_G = true;
  } else {
    if (rmax > kZRowMax) return fail(ctx, FSX_EINVAL, "a Z row is longer than the Z-build kernel supports");
    for (int r = 0; r < nrows; r++) {
      bool diag = false;
      for (int k = rowptr[r]; k < rowptr[r + 1]; k++) diag |= (col[k] == r);
      if (!diag) return fail(ctx, FSX_EINVAL, "Z pattern lacks a diagonal entry");
    }
    ctx->Zrow = drow; ctx->Zcol = dcol; ctx->nnzZ = nnz; ctx->zrow_max = rmax;
    FSX_TRY(alloc(ctx, &ctx->Z, (size_t)nnz));
    FSX_TRY(alloc(ctx, &ctx->Zdinv, (size_t)nrows));
    FSX_TRY(alloc(ctx, &ctx->Zdinvf, (size_t)nrows));
    ctx->have_Z = true;
  }
  return FSX_OK;
}

int fsx_upload_bc(fsx_ctx* ctx, int n, const int* nodes, const double* values, int zero_p_node) {
  if (!ctx || n < 0 || (n > 0 && (!nodes || !values))) return ctx ? fail(ctx, FSX_EINVAL, "fsx_upload_bc") : FSX_EINVAL;
  if (!ctx->have_mesh) return fail(ctx, FSX_ESTATE, "upload the mesh first");
  if (zero_p_node >= ctx->NNp) return fail(ctx, FSX_EINVAL, "zero-pressure node out of range");
  std::vector<unsigned char> flag(ctx->NN, 0);
  std::vector<int> inodes(std::max(n, 1));   // internal ids
  for (int i = 0; i < n; i++) {
    if (nodes[i] < 0 || nodes[i] >= ctx->NN) return fail(ctx, FSX_EINVAL, "BC node out of range");
    inodes[i] = ctx->vperm[nodes[i]];
    flag[inodes[i]] = 1;
  }
  if (ctx->bcnodes) { dev_free(ctx, ctx->bcnodes); dev_free(ctx, ctx->bcvals); }
  FSX_TRY(alloc(ctx, &ctx->bcnodes, (size_t)std::max(n, 1)));
  FSX_TRY(alloc(ctx, &ctx->bcvals, (size_t)std::max(n, 1) * 3));
  if (n > 0) {
    FSX_HIP(hipMemcpy(ctx->bcnodes, inodes.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    FSX_HIP(hipMemcpy(ctx->bcvals, values, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice));
  }
  FSX_HIP(hipMemcpy(ctx->bcflag, flag.data(), (size_t)ctx->NN, hipMemcpyHostToDevice));
  ctx->nbc = n;
  ctx->zeroP = zero_p_node;
  ctx->mass_dinv_ready = false;
  ctx->have_bc = true;
  return FSX_OK;
}

int fsx_assemble_static(fsx_ctx* ctx) {
  if (!ctx) return FSX_EINVAL;
  if (!(ctx->have_mesh && ctx->have_M && ctx->have_G && ctx->have_bc && ctx->have_params))
    return fail(ctx, FSX_ESTATE, "fsx_assemble_static needs mesh, M and G patterns, BCs and parameters");
  FSX_TRY(launch_assemble_static(ctx));
  ctx->assembled = true;
  ctx->gc_built = false;   // G's classes are formed again from the new values
  ctx->gpn_valid = ctx->gpn_next_valid = false;
  FSX_TRY(refresh_matrix_copies(ctx));
  ctx->mass_dinv_ready = false;
  ctx->mass_m_uploaded = false;   // M is the elements' sum again
  return FSX_OK;
}

int fsx_build_Z(fsx_ctx* ctx) {
  if (!ctx) return FSX_EINVAL;
  if (!ctx->assembled || !ctx->have_Z) return fail(ctx, FSX_ESTATE, "fsx_build_Z needs assembly and the Z pattern");
  FSX_TRY(launch_build_Z(ctx));
  FSX_TRY(build_z_forms(ctx));
  ctx->z_built = true;
  ctx->band_ready = false;   // a direct solve factors the new Z
  fsx::mf_invalidate(ctx);
  fsx::proj_reset(ctx);      // the kept solutions are of the old Z
  return FSX_OK;
}

int fsx_init_state(fsx_ctx* ctx) {
  if (!ctx) return FSX_EINVAL;
  if (!ctx->have_mesh || !ctx->have_bc) return fail(ctx, FSX_ESTATE, "fsx_init_state needs mesh and BCs");
  const long long n3 = 3LL * ctx->NN;
  ctx->kacc_pending = false;
  ctx->conv_pending = false;
  ctx->gpn_valid = ctx->gpn_next_valid = false;
  drop_records(ctx);
  fsx::proj_reset(ctx);
  for (double* v : {ctx->Un, ctx->Unp1, ctx->Unp1_prev, ctx->UnpHalf, ctx->UnpHalf_prev, ctx->Acc, ctx->Acc_prev,
                    ctx->KAcc, ctx->R1})
    FSX_TRY(launch_fill(ctx, v, 0.0, n3));
  for (double* v : {ctx->Pn, ctx->Pnp1, ctx->Pnp1_prev, ctx->Pdot, ctx->R2}) FSX_TRY(launch_fill(ctx, v, 0.0, ctx->NNp));
  if (ctx->mass_x) FSX_TRY(launch_fill(ctx, ctx->mass_x, 0.0, n3));   // the consistent step1's warm start
  // applyBC_initial (BCH:4062-4107): Un at BC nodes <- BC values
  if (ctx->nbc > 0) {
    std::vector<int> nodes(ctx->nbc);
    std::vector<double> vals((size_t)ctx->nbc * 3), un((size_t)n3, 0.0);
    FSX_HIP(hipMemcpy(nodes.data(), ctx->bcnodes, nodes.size() * sizeof(int), hipMemcpyDeviceToHost));
    FSX_HIP(hipMemcpy(vals.data(), ctx->bcvals, vals.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < ctx->nbc; i++)
      for (int c = 0; c < 3; c++) un[3 * (size_t)nodes[i] + c] = vals[(size_t)i*3 + c];
    FSX_HIP(hipStreamSynchronize(ctx->stream));
    FSX_HIP(hipMemcpy(ctx->Un, un.data(), un.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  ctx->state_ready = true;
  return FSX_OK;
}

static int need_step(fsx_ctx* ctx) {
  if (!ctx) return FSX_EINVAL;
  if (!ctx->z_built || !ctx->state_ready) return fail(ctx, FSX_ESTATE, "assemble, build Z and init the state first");
  return FSX_OK;
}

int fsx_begin_step(fsx_ctx* ctx) {
  FSX_TRY(need_step(ctx));
  FSX_TRY(settle_deferred(ctx));
  FSX_TRY(launch_copy2(ctx, ctx->UnpHalf_prev, ctx->Un, 3LL * ctx->NN, ctx->Pnp1_prev, ctx->Pn,
                       ctx->NNp));    // BCH:2872-2876
  FSX_TRY(launch_fill(ctx, ctx->Acc_prev, 0.0, 3LL * ctx->NN));
  return records_begin_step(ctx);
}

int fsx_convection(fsx_ctx* ctx) { FSX_TRY(need_step(ctx)); return launch_convection(ctx); }
int fsx_step1(fsx_ctx* ctx, int iter) { FSX_TRY(need_step(ctx)); return launch_step1(ctx, iter); }

int fsx_step2(fsx_ctx* ctx, int iter, int* pcg_iters, double* relres) {
  FSX_TRY(need_step(ctx));
  FSX_TRY(settle_deferred(ctx));
  ctx->gpn_next_valid = false;   // Pnp1 changes
  FSX_TRY(launch_step2_rhs(ctx, iter));
  int it = 0;
  double rel = 0.0;
  int rc = pressure_solve(ctx, ctx->R2, ctx->Pdot, &it, &rel);
  if (rc != FSX_OK && rc != FSX_ENOCONV) return rc;
  if (pcg_iters) *pcg_iters = it;
  if (relres) *relres = rel;
  FSX_TRY(launch_pnp1(ctx));
  return rc;
}

int fsx_step3(fsx_ctx* ctx, int iter) {
  FSX_TRY(need_step(ctx));
  FSX_TRY(settle_deferred(ctx));
  return launch_step3(ctx, iter);
}

int fsx_check_convergence(fsx_ctx* ctx, double tol, int* converged, double* rel_u, double* rel_p) {
  FSX_TRY(need_step(ctx));
  FSX_HIP(hipMemcpyAsync(ctx->h_st, ctx->st, sizeof(StepScalars), hipMemcpyDeviceToHost, ctx->stream));
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  double nu = std::sqrt(ctx->h_st->ud) / std::sqrt(ctx->h_st->uu);        // BCH:2936-2958
  double np = std::sqrt(ctx->h_st->pd) / std::sqrt(ctx->h_st->pp);
  if (rel_u) *rel_u = nu;
  if (rel_p) *rel_p = np;
  if (converged) *converged = (nu < tol && np < tol) ? 1 : 0;
  return FSX_OK;
}

int fsx_next_iteration(fsx_ctx* ctx) {
  FSX_TRY(need_step(ctx));
  FSX_TRY(settle_deferred(ctx));
  // BCH:2964-2972. UnpHalf and Acc are rewritten by step1/step3 before they
  // are read again, so their "prev" copies are pointer swaps. Unp1 and Pnp1
  // must keep their values: when this was the last inner iteration the time
  // loop still moves them into Un and Pn (BCH:3070-3077), so they are copied.
  std::swap(ctx->UnpHalf_prev, ctx->UnpHalf);
  std::swap(ctx->Acc_prev, ctx->Acc);
  std::swap(ctx->Urec_prev, ctx->Urec);   // their FP32 shadows go with them
  std::swap(ctx->Arec_prev, ctx->Arec);
  std::swap(ctx->urec_prev_ok, ctx->urec_ok);
  std::swap(ctx->arec_prev_ok, ctx->arec_ok);
  ctx->drec_ok = false;   // step1 of the next inner iteration writes it again
  ctx->ucmp_ok = false;   // UnpHalf_prev is no longer Un
  FSX_TRY(launch_copy2(ctx, ctx->Unp1_prev, ctx->Unp1, 3LL * ctx->NN, ctx->Pnp1_prev, ctx->Pnp1, ctx->NNp));
  return FSX_OK;
}

// KAcc = K Acc_prev: recorded here and formed by the next fsx_step1 in its
// pass over K (k_step1 with KACC); any other entry point that reads KAcc or
// changes Acc_prev first forms it on its own (settle_deferred).
int fsx_kacc(fsx_ctx* ctx) {
  FSX_TRY(need_step(ctx));
  ctx->kacc_pending = true;
  return FSX_OK;
}

int fsx_end_step(fsx_ctx* ctx, double converge, int* steady, double* max_acc) {
  FSX_TRY(need_step(ctx));
  FSX_TRY(settle_deferred(ctx));
  FSX_TRY(launch_maxacc(ctx));
  FSX_HIP(hipMemcpyAsync(ctx->h_st, ctx->st, sizeof(StepScalars), hipMemcpyDeviceToHost, ctx->stream));
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  double m = ctx->h_st->maxacc;
  if (max_acc) *max_acc = m;
  if (steady) *steady = (m > converge) ? 0 : 1;                            // BCH:3059-3061
  std::swap(ctx->Un, ctx->Unp1);                                           // BCH:3070-3077
  std::swap(ctx->Pn, ctx->Pnp1);
  // G Pn changes with Pn: the G Pnp1 that step3 carried becomes it
  if (ctx->gpn_next_valid) std::swap(ctx->GPn, ctx->GPn_next);
  ctx->gpn_valid = ctx->gpn_next_valid;
  ctx->gpn_next_valid = false;
  ctx->ucmp_ok = ctx->ucmp_un = false;                                     // and Un's shadow with Un
  return FSX_OK;
}

int fsx_monitor(fsx_ctx* ctx, int node, double* out4) {
  if (!ctx || !out4) return FSX_EINVAL;
  if (node < 0 || node >= ctx->NNp) return fail(ctx, FSX_EINVAL, "monitor node out of range");
  for (int c = 0; c < 3; c++)
    FSX_HIP(hipMemcpyAsync(ctx->h_mon + c, ctx->Un + 3 * (size_t)ctx->vperm[node] + c, sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
  FSX_HIP(hipMemcpyAsync(ctx->h_mon + 3, ctx->Pn + node, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(out4, ctx->h_mon, 4 * sizeof(double));
  return FSX_OK;
}

static int vec_info(fsx_ctx* ctx, int id, double** p, size_t* n, int* kind) {
  // kind 0: plain, 1: NN vector replicated x3, 2: component c of G (c = id - FSX_VAL_G1),
  // 3: velocity vector, [u|v|w] at the boundary and node-major (u, v, w of a node together) on the device,
  // 4: values on the M pattern; kinds 1-4 are in the internal velocity numbering on the device
  const size_t n3 = (size_t)3 * ctx->NN;
  *kind = 0;
  switch (id) {
    case FSX_VEC_UN: *p = ctx->Un; *n = n3; *kind = 3; break;
    case FSX_VEC_UNP1: *p = ctx->Unp1; *n = n3; *kind = 3; break;
    case FSX_VEC_UNP1_PREV: *p = ctx->Unp1_prev; *n = n3; *kind = 3; break;
    case FSX_VEC_UNPHALF: *p = ctx->UnpHalf; *n = n3; *kind = 3; break;
    case FSX_VEC_UNPHALF_PREV: *p = ctx->UnpHalf_prev; *n = n3; *kind = 3; break;
    case FSX_VEC_ACC: *p = ctx->Acc; *n = n3; *kind = 3; break;
    case FSX_VEC_KACC: *p = ctx->KAcc; *n = n3; *kind = 3; break;
    case FSX_VEC_R1: *p = ctx->R1; *n = n3; *kind = 3; break;
    case FSX_VEC_PN: *p = ctx->Pn; *n = ctx->NNp; break;
    case FSX_VEC_PNP1: *p = ctx->Pnp1; *n = ctx->NNp; break;
    case FSX_VEC_PNP1_PREV: *p = ctx->Pnp1_prev; *n = ctx->NNp; break;
    case FSX_VEC_PDOT: *p = ctx->Pdot; *n = ctx->NNp; break;
    case FSX_VEC_R2: *p = ctx->R2; *n = ctx->NNp; break;
    case FSX_VEC_MDINV: *p = ctx->MdInv; *n = n3; *kind = 1; break;
    case FSX_VEC_MDORIGINV: *p = ctx->MdOrigInv; *n = n3; *kind = 1; break;
    case FSX_VAL_K: *p = ctx->K; *n = ctx->nnzM; *kind = 4; break;
    case FSX_VAL_M: *p = ctx->M; *n = ctx->nnzM; *kind = 4; break;
    case FSX_VAL_G1: case FSX_VAL_G2: case FSX_VAL_G3: *p = ctx->G; *n = ctx->nnzG; *kind = 2; break;
    case FSX_VAL_Z: *p = ctx->Z; *n = ctx->nnzZ; break;
    default: return fail(ctx, FSX_EINVAL, "unknown vector id");
  }
  if (!*p) return fail(ctx, FSX_ESTATE, "vector not allocated yet");
  return FSX_OK;
}

// Position in the device's (internal) CSR of every entry of an uploaded
// reference pattern: kind M maps rows and columns through vperm, kind G only
// the rows (its columns are pressure nodes).
static int build_pos(fsx_ctx* ctx, bool is_m, std::vector<int>& pos) {
  const std::vector<int>& rr = is_m ? ctx->refMrow : ctx->refGrow;
  const std::vector<int>& rc = is_m ? ctx->refMcol : ctx->refGcol;
  const int nrows = ctx->NN;
  const size_t nnz = rc.size();
  std::vector<int> irow((size_t)nrows + 1), icol(nnz);
  FSX_HIP(hipMemcpy(irow.data(), is_m ? ctx->Mrow : ctx->Grow, irow.size() * sizeof(int), hipMemcpyDeviceToHost));
  FSX_HIP(hipMemcpy(icol.data(), is_m ? ctx->Mcol : ctx->Gcol, nnz * sizeof(int), hipMemcpyDeviceToHost));
  pos.resize(nnz);
  for (int r = 0; r < nrows; r++) {
    const int pr = ctx->vperm[r];
    const int* lo = icol.data() + irow[pr];
    const int* hi = icol.data() + irow[pr + 1];
    for (int k = rr[r]; k < rr[r + 1]; k++) {
      const int c = is_m ? ctx->vperm[rc[k]] : rc[k];
      const int* f = std::lower_bound(lo, hi, c);
      if (f == hi || *f != c) return fail(ctx, FSX_ESTATE, "internal pattern lacks an uploaded entry");
      pos[k] = (int)(f - icol.data());
    }
  }
  return FSX_OK;
}

int fsx_download(fsx_ctx* ctx, int id, double* host, size_t count) {
  if (!ctx || !host) return FSX_EINVAL;
  double* p; size_t n; int kind;
  FSX_TRY(vec_info(ctx, id, &p, &n, &kind));
  if (count != n) return fail(ctx, FSX_EINVAL, "fsx_download: count mismatch (expected " + std::to_string(n) + ")");
  FSX_TRY(settle_deferred(ctx));
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  const size_t NN = ctx->NN;
  const std::vector<int>& ip = ctx->viperm;
  if (kind == 0) {
    FSX_HIP(hipMemcpy(host, p, n * sizeof(double), hipMemcpyDeviceToHost));
  } else if (kind == 3) {   // node-major, internal order -> [u|v|w], reference order
    std::vector<double> t(n);
    FSX_HIP(hipMemcpy(t.data(), p, n * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < NN; i++)
      for (int c = 0; c < 3; c++) host[c * NN + ip[i]] = t[3 * i + c];
  } else if (kind == 1) {   // one NN vector in internal order, replicated x3
    std::vector<double> t(NN);
    FSX_HIP(hipMemcpy(t.data(), p, NN * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < NN; i++) host[ip[i]] = t[i];
    std::memcpy(host + NN, host, NN * sizeof(double));
    std::memcpy(host + 2 * NN, host, NN * sizeof(double));
  } else if (kind == 4) {   // values on the M pattern, in the reference CSR order
    if (ctx->posM.empty()) FSX_TRY(build_pos(ctx, true, ctx->posM));
    std::vector<double> t(n);
    FSX_HIP(hipMemcpy(t.data(), p, n * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < n; k++) host[k] = t[ctx->posM[k]];
  } else {                  // component c of the interleaved G values, reference CSR order
    if (ctx->posG.empty()) FSX_TRY(build_pos(ctx, false, ctx->posG));
    std::vector<double> g(n * 3);
    FSX_HIP(hipMemcpy(g.data(), p, g.size() * sizeof(double), hipMemcpyDeviceToHost));
    const int c = id - FSX_VAL_G1;
    for (size_t k = 0; k < n; k++) host[k] = g[3 * (size_t)ctx->posG[k] + c];
  }
  return FSX_OK;
}

int fsx_upload_vec(fsx_ctx* ctx, int id, const double* host, size_t count) {
  if (!ctx || !host) return FSX_EINVAL;
  double* p; size_t n; int kind;
  FSX_TRY(vec_info(ctx, id, &p, &n, &kind));
  if (kind != 0 && kind != 3 && kind != 4)
    return fail(ctx, FSX_EINVAL, "fsx_upload_vec: only state vectors and K/M/Z values");
  FSX_TRY(settle_deferred(ctx));
  if (count != n) return fail(ctx, FSX_EINVAL, "fsx_upload_vec: count mismatch");
  ctx->gpn_valid = ctx->gpn_next_valid = false;
  drop_records(ctx);   // the FP32 shadows may no longer match the vectors
  if (id == FSX_VAL_Z) {
    ctx->band_ready = false;
    mf_invalidate(ctx);
    proj_reset(ctx);
  }
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  if (kind == 3) {
    const size_t NN = ctx->NN;
    std::vector<double> t(n);
    for (size_t i = 0; i < NN; i++)
      for (int c = 0; c < 3; c++) t[3 * i + c] = host[c * NN + ctx->viperm[i]];
    FSX_HIP(hipMemcpy(p, t.data(), n * sizeof(double), hipMemcpyHostToDevice));
  } else if (kind == 4) {
    if (ctx->posM.empty()) FSX_TRY(build_pos(ctx, true, ctx->posM));
    std::vector<double> t(n);
    for (size_t k = 0; k < n; k++) t[ctx->posM[k]] = host[k];
    FSX_HIP(hipMemcpy(p, t.data(), n * sizeof(double), hipMemcpyHostToDevice));
    if (id == FSX_VAL_K) {
      FSX_TRY(refresh_matrix_copies(ctx));
      ctx->k_uploaded = true;   // step1 applies these CSR values until the next fsx_assemble_static
    }
    if (id == FSX_VAL_M) {
      // the lumped masses follow M (Z keeps the lumped mass it was built
      // from until fsx_build_Z runs again)
      FSX_TRY(launch_lump(ctx));
      ctx->mass_dinv_ready = false;
      ctx->mass_m_uploaded = true;   // consistent solves apply these CSR values
    }
  } else {
    FSX_HIP(hipMemcpy(p, host, n * sizeof(double), hipMemcpyHostToDevice));
    if (id == FSX_VAL_Z) {
      // the uploaded values carry the penalty pin; every form the solvers
      // read (Jacobi inverses, the FP32 copy, the rows without structural
      // zeros, the row classes) is rebuilt from them
      FSX_TRY(launch_z_refresh(ctx));
      FSX_TRY(build_z_forms(ctx));
      ctx->z_built = true;
    }
  }
  return FSX_OK;
}

int fsx_sync(fsx_ctx* ctx) {
  if (!ctx) return FSX_EINVAL;
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  return FSX_OK;
}

int fsx_solve_pressure(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres) {
  if (!ctx || !b || !x) return FSX_EINVAL;
  if (!ctx->z_built) return fail(ctx, FSX_ESTATE, "build Z first");
  FSX_TRY(settle_deferred(ctx));
  FSX_HIP(hipMemcpyAsync(ctx->R2, b, (size_t)ctx->NNp * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  int it = 0;
  double rel = 0;
  int rc = pressure_solve(ctx, ctx->R2, ctx->Pdot, &it, &rel);
  if (rc != FSX_OK && rc != FSX_ENOCONV) return rc;
  FSX_HIP(hipMemcpyAsync(x, ctx->Pdot, (size_t)ctx->NNp * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  if (iters) *iters = it;
  if (relres) *relres = rel;
  return rc;
}

int fsx_spmv(fsx_ctx* ctx, int which, const double* x, double* y, double alpha, double beta) {
  if (!ctx || !x || !y) return FSX_EINVAL;
  if (!ctx->assembled) return fail(ctx, FSX_ESTATE, "assemble first");
  if ((which == 4 || which == 8 || which == 9) && !ctx->z_built) return fail(ctx, FSX_ESTATE, "build Z first");
  FSX_TRY(settle_deferred(ctx));
  if ((which == 8 || which == 9) && !(alpha == 1.0 && beta == 0.0))
    return fail(ctx, FSX_EINVAL, "fsx_spmv: the solvers' Z forms take alpha = 1, beta = 0");
  size_t nx, ny;
  if (which == 0) { nx = ny = ctx->NN; }
  else if (which >= 1 && which <= 3) { nx = ctx->NNp; ny = ctx->NN; }
  else if (which == 4 || which == 8 || which == 9) { nx = ny = ctx->NNp; }
  else if (which >= 5 && which <= 7) { nx = ctx->NN; ny = ctx->NNp; }
  else return fail(ctx, FSX_EINVAL, "fsx_spmv: unknown operator");
  double *dx, *dy;
  FSX_TRY(alloc(ctx, &dx, nx));
  FSX_TRY(alloc(ctx, &dy, ny));
  int rc = FSX_OK;
  // velocity-sized operands move between the reference and the internal numbering
  const bool x_vel = which == 0 || (which >= 5 && which <= 7), y_vel = which >= 0 && which <= 3;
  std::vector<double> hx(x, x + nx), hy(y, y + ny);
  if (x_vel) for (size_t i = 0; i < nx; i++) hx[ctx->vperm[i]] = x[i];
  if (y_vel) for (size_t i = 0; i < ny; i++) hy[ctx->vperm[i]] = y[i];
  if (hipMemcpy(dx, hx.data(), nx * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dy, hy.data(), ny * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(ctx, FSX_EHIP, "fsx_spmv copy");
  if (rc == FSX_OK) rc = spmv_host_op(ctx, which, dx, dy, alpha, beta);
  if (rc == FSX_OK) {
    if (hipStreamSynchronize(ctx->stream) != hipSuccess ||
        hipMemcpy(hy.data(), dy, ny * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(ctx, FSX_EHIP, "fsx_spmv copy back");
    else if (y_vel) for (size_t i = 0; i < ny; i++) y[i] = hy[ctx->vperm[i]];
    else std::memcpy(y, hy.data(), ny * sizeof(double));
  }
  dev_free(ctx, dx);
  dev_free(ctx, dy);
  ctx->bytes -= (nx + ny) * sizeof(double);
  return rc;
}

int fsx_timing_reset(fsx_ctx* ctx, int enable) {
  if (!ctx) return FSX_EINVAL;
  ctx->timing = enable != 0;
  ctx->tm = fsx_timing{};
  return FSX_OK;
}

int fsx_timing_get(fsx_ctx* ctx, fsx_timing* out) {
  if (!ctx || !out) return FSX_EINVAL;
  *out = ctx->tm;
  return FSX_OK;
}

}  // extern "C"
