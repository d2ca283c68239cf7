// fsx_mass.hip -- consistent-mass velocity update: the three velocity
// components solved together as one 3-right-hand-side Jacobi-PCG on the
// consistent mass matrix.
//
// No counterpart in the fractional-step reference, which lumps M (row sums,
// BCH:3262-3295) and updates UnpHalf = Un + dt R1 MdInv (BCH:3781-3783) and
// Acc = R3 MdInv (BCH:3960-3962). In the consistent mode those become
// M_bc a = R1, UnpHalf = Un + dt a and M_bc Acc = R3, where M_bc is M with the
// rows and columns of BC nodes replaced by identity rows and columns. R is
// zero at BC nodes, so a and Acc are too, as with the lumped update.
//
// Block form: one pass over M's pattern applies M to all three components of
// the node-major search direction (u, v, w of a node side by side), while
// each component keeps its own CG scalars (alpha_c, beta_c) and its own stop
// test ||r_c|| <= rtol ||b_c||. Iterates are p, q, r, z, x of 3NN doubles;
// BC rows of q are zero and the Jacobi inverse is 0 there, so r, z, p and x
// stay zero at BC nodes and M p equals M_bc p on the free rows.
//
// One iteration is four kernels (SpMV, p.q per component, update, direction)
// with the scalars on the device, like the pressure PCG; the host polls the
// stop flag after every kMassChunk iterations.
#include "fsx_device.h"

#include <algorithm>
#include <cmath>

namespace fsx {

struct MassScalars {
  double rz[2][3];      // r.z per component, by iteration parity
  double bb[3], rr[3];  // |b_c|^2, |r_c|^2
  double rtol2;
  int iters, maxit, stop;   // stop: 1 converged, 2 maxit, 3 NaN
};

constexpr int kMassChunk = 8;

// Jacobi inverse of M's diagonal, 0 at BC nodes
__global__ void k_mass_dinv(int NN, const int* __restrict__ Mrow, const int* __restrict__ Mcol, const double* __restrict__ M,
                            const unsigned char* __restrict__ bc, double* __restrict__ dinv) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < NN; r += gridDim.x * blockDim.x) {
    const int k = find_col(Mcol, Mrow[r], Mrow[r + 1], r);
    dinv[r] = bc[r] ? 0.0 : 1.0 / M[k];
  }
}

// x = 0, r = b, z = dinv r, p = z; partials of r.z (slots 0-2) and b.b (3-5)
__global__ __launch_bounds__(kBlock) void k_mass_init(int NN, const double* __restrict__ b, const double* __restrict__ dinv,
                                                      double* __restrict__ x, double* __restrict__ r,
                                                      double* __restrict__ z, double* __restrict__ p,
                                                      double* __restrict__ partials) {
  double v[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < NN; i += kRedGrid * kBlock) {
    const double di = dinv[i];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const size_t k = 3 * (size_t)i + c;
      const double bi = b[k], zi = di * bi;
      x[k] = 0.0; r[k] = bi; z[k] = zi; p[k] = zi;
      v[c] += bi * zi;
      v[3 + c] += bi * bi;
    }
  }
  write_partials<6>(v, partials);
}

__global__ __launch_bounds__(kBlock) void k_mass_init_finish(const double* __restrict__ partials, MassScalars* sc,
                                                             double rtol, int maxit) {
  double s[6];
  read_partials<6>(partials, s);
  if (threadIdx.x == 0) {
    bool zero = true;
    for (int c = 0; c < 3; c++) {
      sc->rz[0][c] = s[c]; sc->rz[1][c] = 0.0;
      sc->bb[c] = s[3 + c]; sc->rr[c] = s[3 + c];
      zero = zero && s[3 + c] == 0.0;
    }
    sc->rtol2 = rtol * rtol;
    sc->iters = 0; sc->maxit = maxit;
    sc->stop = zero ? 1 : 0;
  }
}

// q = M p for the three components (rows = nodes), q = 0 at BC rows
template <int LPR, int U, bool KPAT>
__global__ __launch_bounds__(kBlock) void k_mass_spmv(int NN, const int* __restrict__ Mrow, ColSrc mcs,
                                                      const double* __restrict__ M, const unsigned char* __restrict__ bc,
                                                      const double* __restrict__ p, double* __restrict__ q,
                                                      const MassScalars* __restrict__ sc) {
  if (sc->stop) return;
  const int l = threadIdx.x % LPR;
  const long long r = (long long)blockIdx.x * (kBlock / LPR) + threadIdx.x / LPR;
  double s0 = 0, s1 = 0, s2 = 0;
  if (r < NN) {
    const int rs = Mrow[r], re = Mrow[r + 1];
    const int cb = col_base<KPAT>(mcs, r, rs);
    for (int base = rs + l; base < re; base += U * LPR) {
      int c[U];
      double a[U];
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int k = base + j * LPR, kk = k < re ? k : re - 1;
        c[j] = col_at<KPAT>(mcs, r, cb, kk);
        a[j] = k < re ? M[kk] : 0.0;
      }
#pragma unroll
      for (int j = 0; j < U; j++) {
        const size_t x = 3 * (size_t)c[j];
        s0 += a[j] * p[x]; s1 += a[j] * p[x + 1]; s2 += a[j] * p[x + 2];
      }
    }
  }
  s0 = group_sum<LPR>(s0); s1 = group_sum<LPR>(s1); s2 = group_sum<LPR>(s2);
  if (r < NN && l < 3) q[3 * (size_t)r + l] = bc[r] ? 0.0 : (l == 0 ? s0 : (l == 1 ? s1 : s2));
}

// p.q per component (slots 6-8)
__global__ __launch_bounds__(kBlock) void k_mass_dot(int NN, const double* __restrict__ p, const double* __restrict__ q,
                                                     double* __restrict__ partials, const MassScalars* __restrict__ sc) {
  if (sc->stop) return;
  double v[3] = {0, 0, 0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < NN; i += kRedGrid * kBlock)
#pragma unroll
    for (int c = 0; c < 3; c++) v[c] += p[3 * (size_t)i + c] * q[3 * (size_t)i + c];
  write_partials<3>(v, partials + 6 * kRedGrid);
}

// alpha_c = r.z_c / p.q_c (0 for a component with no work left);
// x += alpha p, r -= alpha q, z = dinv r; partials of r.r (slots 0-2) and r.z (3-5)
__global__ __launch_bounds__(kBlock) void k_mass_update(int j, int NN, double* __restrict__ x, double* __restrict__ r,
                                                        double* __restrict__ z, const double* __restrict__ p,
                                                        const double* __restrict__ q, const double* __restrict__ dinv,
                                                        double* __restrict__ partials, const MassScalars* __restrict__ sc) {
  if (sc->stop) return;
  double pq[3];
  read_partials<3>(partials + 6 * kRedGrid, pq);
  double a[3];
#pragma unroll
  for (int c = 0; c < 3; c++) a[c] = pq[c] > 0.0 ? sc->rz[j & 1][c] / pq[c] : 0.0;
  double v[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < NN; i += kRedGrid * kBlock) {
    const double di = dinv[i];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const size_t k = 3 * (size_t)i + c;
      x[k] += a[c] * p[k];
      const double ri = r[k] - a[c] * q[k];
      const double zi = di * ri;
      r[k] = ri; z[k] = zi;
      v[c] += ri * ri;
      v[3 + c] += ri * zi;
    }
  }
  write_partials<6>(v, partials);
}

// beta_c = r.z_c(new) / r.z_c(old); p = z + beta p; block 0 records the scalars
// and the stop test: every component at ||r_c|| <= rtol ||b_c||
__global__ __launch_bounds__(kBlock) void k_mass_direction(int j, int NN, const double* __restrict__ z,
                                                           double* __restrict__ p, const double* __restrict__ partials,
                                                           MassScalars* sc) {
  if (sc->stop) return;
  double s[6];
  read_partials<6>(partials, s);
  double be[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const double old = sc->rz[j & 1][c];
    be[c] = old != 0.0 ? s[3 + c] / old : 0.0;
  }
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < NN; i += kRedGrid * kBlock)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const size_t k = 3 * (size_t)i + c;
      p[k] = z[k] + be[c] * p[k];
    }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    bool done = true, bad = false;
    for (int c = 0; c < 3; c++) {
      sc->rz[(j + 1) & 1][c] = s[3 + c];
      sc->rr[c] = s[c];
      bad = bad || !(s[c] == s[c]) || !(s[3 + c] == s[3 + c]);
      done = done && s[c] <= sc->rtol2 * sc->bb[c];
    }
    const int it = sc->iters + 1;
    sc->iters = it;
    if (bad) sc->stop = 3;
    else if (done) sc->stop = 1;
    else if (it >= sc->maxit) sc->stop = 2;
  }
}

// y = u + dt x (node-major 3NN)
__global__ void k_mass_axpy(long long n, const double* __restrict__ u, const double* __restrict__ x, double dt,
                            double* __restrict__ y) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = u[i] + dt * x[i];
}

// Solves M_bc x = b (3NN, node-major) to the context's mass tolerance.
int mass_solve(fsx_ctx* ctx, const double* b, double* x) {
  const int NN = ctx->NN;
  if (!ctx->mass_dinv_ready) {
    hipLaunchKernelGGL(k_mass_dinv, dim3(std::min((NN + kBlock - 1) / kBlock, 4096)), dim3(kBlock), 0, ctx->stream, NN,
                       ctx->Mrow, ctx->Mcol, ctx->M, ctx->bcflag, ctx->mass_dinv);
    FSX_TRY(launch_check(ctx, "k_mass_dinv"));
    ctx->mass_dinv_ready = true;
  }
  MassScalars* sc = reinterpret_cast<MassScalars*>(ctx->mass_sc);
  double *r = ctx->mass_w, *z = r + 3 * (size_t)NN, *p = z + 3 * (size_t)NN, *q = p + 3 * (size_t)NN;
  hipLaunchKernelGGL(k_mass_init, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, NN, b, ctx->mass_dinv, x, r, z, p,
                     ctx->partials);
  hipLaunchKernelGGL(k_mass_init_finish, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->partials, sc, ctx->mass_rtol,
                     ctx->mass_maxit);
  FSX_TRY(launch_check(ctx, "k_mass_init"));
  constexpr int LPR = 16, U = 4;
  const ColSrc mcs{ctx->Mcol, ctx->Mrowpat, ctx->Mpatoff};
  const unsigned grid = (unsigned)std::max(1, (NN + kBlock / LPR - 1) / (kBlock / LPR));
  MassScalars h{};
  int j = 0;
  for (;;) {
    for (int t = 0; t < kMassChunk; t++, j++) {
      if (mcs.row_pat)
        hipLaunchKernelGGL((k_mass_spmv<LPR, U, true>), dim3(grid), dim3(kBlock), 0, ctx->stream, NN, ctx->Mrow, mcs,
                           ctx->M, ctx->bcflag, p, q, sc);
      else
        hipLaunchKernelGGL((k_mass_spmv<LPR, U, false>), dim3(grid), dim3(kBlock), 0, ctx->stream, NN, ctx->Mrow, mcs,
                           ctx->M, ctx->bcflag, p, q, sc);
      hipLaunchKernelGGL(k_mass_dot, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, NN, p, q, ctx->partials, sc);
      hipLaunchKernelGGL(k_mass_update, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, j, NN, x, r, z, p, q,
                         ctx->mass_dinv, ctx->partials, sc);
      hipLaunchKernelGGL(k_mass_direction, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, j, NN, z, p, ctx->partials, sc);
    }
    FSX_TRY(launch_check(ctx, "mass PCG"));
    FSX_HIP(hipMemcpyAsync(&h, sc, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    FSX_HIP(hipStreamSynchronize(ctx->stream));
    if (h.stop) break;
    if (j > ctx->mass_maxit + 2 * kMassChunk) return fail(ctx, FSX_ESTATE, "mass PCG did not terminate");
  }
  double rel = 0.0;
  for (int c = 0; c < 3; c++) rel = std::max(rel, h.bb[c] > 0.0 ? std::sqrt(h.rr[c] / h.bb[c]) : 0.0);
  ctx->mass_last_iters = h.iters;
  ctx->mass_last_relres = rel;
  ctx->mass_total_iters += h.iters;
  if (h.stop == 3) return fail(ctx, FSX_ENUMERIC, "mass PCG met NaN/Inf");
  if (h.stop == 2) { ctx->err = "mass PCG reached maxit"; return FSX_ENOCONV; }
  return FSX_OK;
}

// step1's consistent-mass update: R1 is final, UnpHalf = Un + dt M_bc^-1 R1
int mass_update_step1(fsx_ctx* ctx) {
  const int rc = mass_solve(ctx, ctx->R1, ctx->mass_x);
  if (rc != FSX_OK && rc != FSX_ENOCONV) return rc;
  const long long n = 3LL * ctx->NN;
  hipLaunchKernelGGL(k_mass_axpy, dim3((unsigned)std::min<long long>((n + kBlock - 1) / kBlock, 4096)), dim3(kBlock), 0,
                     ctx->stream, n, ctx->Un, ctx->mass_x, ctx->dt, ctx->UnpHalf);
  FSX_TRY(launch_check(ctx, "k_mass_axpy"));
  return rc;
}

// step3's: R3 is in mass_b, Acc = M_bc^-1 R3, Unp1 = UnpHalf + dt Acc
int mass_update_step3(fsx_ctx* ctx) {
  const int rc = mass_solve(ctx, ctx->mass_b, ctx->Acc);
  if (rc != FSX_OK && rc != FSX_ENOCONV) return rc;
  const long long n = 3LL * ctx->NN;
  hipLaunchKernelGGL(k_mass_axpy, dim3((unsigned)std::min<long long>((n + kBlock - 1) / kBlock, 4096)), dim3(kBlock), 0,
                     ctx->stream, n, ctx->UnpHalf, ctx->Acc, ctx->dt, ctx->Unp1);
  FSX_TRY(launch_check(ctx, "k_mass_axpy"));
  return rc;
}

size_t mass_scalars_bytes() { return sizeof(MassScalars); }

}  // namespace fsx

extern "C" int fsx_set_mass(fsx_ctx* ctx, int mode, double rtol, int maxit) {
  using namespace fsx;
  if (!ctx) return FSX_EINVAL;
  if (mode != FSX_MASS_LUMPED && mode != FSX_MASS_CONSISTENT) return fail(ctx, FSX_EINVAL, "fsx_set_mass: unknown mode");
  if (!(rtol >= 0) || maxit < 1) return fail(ctx, FSX_EINVAL, "fsx_set_mass: rtol >= 0 and maxit >= 1 required");
  if (mode == FSX_MASS_CONSISTENT) {
    if (!ctx->have_mesh) return fail(ctx, FSX_ESTATE, "fsx_set_mass: upload the mesh first");
    const size_t n3 = 3 * (size_t)ctx->NN;
    if (!ctx->mass_w) {
      FSX_TRY(alloc(ctx, &ctx->mass_w, 4 * n3));
      FSX_TRY(alloc(ctx, &ctx->mass_x, n3));
      FSX_TRY(alloc(ctx, &ctx->mass_b, n3));
      FSX_TRY(alloc(ctx, &ctx->mass_dinv, (size_t)ctx->NN));
      FSX_TRY(alloc(ctx, &ctx->mass_sc, (mass_scalars_bytes() + 7) / 8));
    }
  }
  ctx->mass_mode = mode;
  ctx->mass_rtol = rtol;
  ctx->mass_maxit = maxit;
  return FSX_OK;
}

extern "C" int fsx_mass_stats(fsx_ctx* ctx, int* last_iters, double* last_relres, long long* total_iters) {
  if (!ctx) return FSX_EINVAL;
  if (last_iters) *last_iters = ctx->mass_last_iters;
  if (last_relres) *last_relres = ctx->mass_last_relres;
  if (total_iters) *total_iters = ctx->mass_total_iters;
  return FSX_OK;
}
