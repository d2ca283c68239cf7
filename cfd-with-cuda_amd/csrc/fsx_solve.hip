// fsx_solve.hip -- CSR row kernels of the time step (fused SpMV + elementwise
// epilogues), deterministic reductions and the Jacobi-PCG pressure solver.
//
// Every row kernel maps a group of LPR lanes to one CSR row: the group reads
// the row's nonzeros contiguously (coalesced), gathers x through the column
// indices (L2-resident for the local meshes here), and reduces with xor
// shuffles inside the wavefront. Blocks own contiguous row ranges and are
// renumbered so that each XCD walks one contiguous eighth of the rows.
#include "fsx_device.h"

#include <algorithm>
#include <cmath>

namespace fsx {

// ------------------------------------------------------------ elementwise
__global__ void k_fill(double* __restrict__ p, double v, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_copy(double* __restrict__ d, const double* __restrict__ s, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) d[i] = s[i];
}

static int ew_grid(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  return (int)std::max(1LL, std::min(g, 4096LL));
}

int launch_fill(fsx_ctx* ctx, double* p, double v, long long n) {
  hipLaunchKernelGGL(k_fill, dim3(ew_grid(n)), dim3(kBlock), 0, ctx->stream, p, v, n);
  return launch_check(ctx, "k_fill");
}
int launch_copy(fsx_ctx* ctx, double* dst, const double* src, long long n) {
  hipLaunchKernelGGL(k_copy, dim3(ew_grid(n)), dim3(kBlock), 0, ctx->stream, dst, src, n);
  return launch_check(ctx, "k_copy");
}

// ------------------------------------------------------------ step 1
// R1 <- (R1 - K UnpHalf_prev) - G Pn; BC rows 0; UnpHalf = Un + dt R1 MdInv.
// (BCH:3746-3783; the convection part of R1 is already in R1.)
template <int LPR>
__global__ __launch_bounds__(kBlock) void k_step1(int NN, const int* __restrict__ Mrow, const int* __restrict__ Mcol,
                                                  const double* __restrict__ K, const int* __restrict__ Grow,
                                                  const int* __restrict__ Gcol, const double* __restrict__ G,
                                                  const double* __restrict__ Uprev, const double* __restrict__ Pn,
                                                  const double* __restrict__ Un, double* __restrict__ R1,
                                                  const double* __restrict__ MdInv, const unsigned char* __restrict__ bc,
                                                  double* __restrict__ UnpHalf, double dt) {
  constexpr int RPB = kBlock / LPR;
  long long r0, r1;
  block_rows(NN, xcd_block(), gridDim.x, RPB, &r0, &r1);
  const int grp = threadIdx.x / LPR, l = threadIdx.x % LPR;
  for (long long base = r0; base < r1; base += RPB) {
    const long long r = base + grp;
    double k0 = 0, k1 = 0, k2 = 0, g0 = 0, g1 = 0, g2 = 0;
    if (r < r1) {
      const int e = Mrow[r + 1];
      for (int k = Mrow[r] + l; k < e; k += LPR) {
        const int c = Mcol[k];
        const double a = K[k];
        k0 += a * Uprev[c]; k1 += a * Uprev[NN + c]; k2 += a * Uprev[2 * (size_t)NN + c];
      }
      const int eg = Grow[r + 1];
      for (int k = Grow[r] + l; k < eg; k += LPR) {
        const double p = Pn[Gcol[k]];
        g0 += G[3 * (size_t)k] * p; g1 += G[3 * (size_t)k + 1] * p; g2 += G[3 * (size_t)k + 2] * p;
      }
    }
    k0 = group_sum<LPR>(k0); k1 = group_sum<LPR>(k1); k2 = group_sum<LPR>(k2);
    g0 = group_sum<LPR>(g0); g1 = group_sum<LPR>(g1); g2 = group_sum<LPR>(g2);
    if (r < r1 && l < 3) {
      const double kk = l == 0 ? k0 : (l == 1 ? k1 : k2);
      const double gg = l == 0 ? g0 : (l == 1 ? g1 : g2);
      const size_t i = (size_t)l * NN + r;
      double R = (R1[i] - kk) - gg;
      if (bc[r]) R = 0.0;
      R1[i] = R;
      UnpHalf[i] = Un[i] + dt * R * MdInv[r];
    }
  }
}

int launch_step1(fsx_ctx* ctx, int iter) {
  (void)iter;
  constexpr int LPR = 16;
  hipLaunchKernelGGL(k_step1<LPR>, dim3(row_grid(ctx->NN, kBlock / LPR)), dim3(kBlock), 0, ctx->stream, ctx->NN,
                     ctx->Mrow, ctx->Mcol, ctx->K, ctx->Grow, ctx->Gcol, ctx->G, ctx->UnpHalf_prev, ctx->Pn, ctx->Un,
                     ctx->R1, ctx->MdInv, ctx->bcflag, ctx->UnpHalf, ctx->dt);
  return launch_check(ctx, "k_step1");
}

// ------------------------------------------------------------ K * Acc_prev
// KAcc_c = K Acc_prev_c, c = 0..2, one pass over K (BCH:2981-3035)
template <int LPR>
__global__ __launch_bounds__(kBlock) void k_k3(int NN, const int* __restrict__ Mrow, const int* __restrict__ Mcol,
                                               const double* __restrict__ K, const double* __restrict__ x,
                                               double* __restrict__ y, double alpha, double beta) {
  constexpr int RPB = kBlock / LPR;
  long long r0, r1;
  block_rows(NN, xcd_block(), gridDim.x, RPB, &r0, &r1);
  const int grp = threadIdx.x / LPR, l = threadIdx.x % LPR;
  for (long long base = r0; base < r1; base += RPB) {
    const long long r = base + grp;
    double k0 = 0, k1 = 0, k2 = 0;
    if (r < r1) {
      const int e = Mrow[r + 1];
      for (int k = Mrow[r] + l; k < e; k += LPR) {
        const int c = Mcol[k];
        const double a = K[k];
        k0 += a * x[c]; k1 += a * x[NN + c]; k2 += a * x[2 * (size_t)NN + c];
      }
    }
    k0 = group_sum<LPR>(k0); k1 = group_sum<LPR>(k1); k2 = group_sum<LPR>(k2);
    if (r < r1 && l < 3) {
      const double kk = l == 0 ? k0 : (l == 1 ? k1 : k2);
      const size_t i = (size_t)l * NN + r;
      y[i] = (beta == 0.0 ? 0.0 : beta * y[i]) + alpha * kk;
    }
  }
}

int launch_kacc(fsx_ctx* ctx) {
  constexpr int LPR = 16;
  hipLaunchKernelGGL(k_k3<LPR>, dim3(row_grid(ctx->NN, kBlock / LPR)), dim3(kBlock), 0, ctx->stream, ctx->NN, ctx->Mrow,
                     ctx->Mcol, ctx->K, ctx->Acc_prev, ctx->KAcc, 1.0, 0.0);
  return launch_check(ctx, "k_k3");
}

// ------------------------------------------------------------ step 2 RHS
// R2[p] = sum_c G_c^T d_c, d_c = UnpHalf_c / dt^2 - [iter>1] MdOrigInv KAcc_c,
// summed as (c0 + c1) + c2 like the three csrmv calls of BCH:3864-3868;
// R2[p0] = 0 for p0 > 0 (BCH:4180-4184).
template <int LPR>
__global__ __launch_bounds__(kBlock) void k_step2_rhs(int NNp, int NN, const int* __restrict__ GTrow,
                                                      const int* __restrict__ GTcol, const double* __restrict__ GT,
                                                      const double* __restrict__ UnpHalf, const double* __restrict__ MdOrigInv,
                                                      const double* __restrict__ KAcc, int use_kacc, double o2,
                                                      int zeroP, double* __restrict__ R2) {
  constexpr int RPB = kBlock / LPR;
  long long r0, r1;
  block_rows(NNp, xcd_block(), gridDim.x, RPB, &r0, &r1);
  const int grp = threadIdx.x / LPR, l = threadIdx.x % LPR;
  for (long long base = r0; base < r1; base += RPB) {
    const long long p = base + grp;
    double s0 = 0, s1 = 0, s2 = 0;
    if (p < r1) {
      const int e = GTrow[p + 1];
      for (int k = GTrow[p] + l; k < e; k += LPR) {
        const int i = GTcol[k];
        double d0 = o2 * UnpHalf[i], d1 = o2 * UnpHalf[NN + i], d2 = o2 * UnpHalf[2 * (size_t)NN + i];
        if (use_kacc) {
          const double m = MdOrigInv[i];
          d0 = d0 - m * KAcc[i]; d1 = d1 - m * KAcc[NN + i]; d2 = d2 - m * KAcc[2 * (size_t)NN + i];
        }
        s0 += GT[3 * (size_t)k] * d0; s1 += GT[3 * (size_t)k + 1] * d1; s2 += GT[3 * (size_t)k + 2] * d2;
      }
    }
    s0 = group_sum<LPR>(s0); s1 = group_sum<LPR>(s1); s2 = group_sum<LPR>(s2);
    if (p < r1 && l == 0) R2[p] = (zeroP > 0 && p == zeroP) ? 0.0 : (s0 + s1) + s2;
  }
}

int launch_step2_rhs(fsx_ctx* ctx, int iter) {
  constexpr int LPR = 32;
  const double o2 = 1.0000000000000000 / (ctx->dt * ctx->dt);
  hipLaunchKernelGGL(k_step2_rhs<LPR>, dim3(row_grid(ctx->NNp, kBlock / LPR)), dim3(kBlock), 0, ctx->stream, ctx->NNp,
                     ctx->NN, ctx->GTrow, ctx->GTcol, ctx->GT, ctx->UnpHalf, ctx->MdOrigInv, ctx->KAcc, iter != 1 ? 1 : 0,
                     o2, ctx->zeroP, ctx->R2);
  return launch_check(ctx, "k_step2_rhs");
}

// Pnp1 = Pn + dt Pdot (BCH:3891-3893) and the pressure norms of the
// convergence check (BCH:2950-2957).
__global__ __launch_bounds__(kBlock) void k_pnp1(int NNp, const double* __restrict__ Pn, const double* __restrict__ Pdot,
                                                 const double* __restrict__ Pprev, double* __restrict__ Pnp1, double dt,
                                                 double* partials, StepScalars* st) {
  long long r0, r1;
  block_rows(NNp, xcd_block(), gridDim.x, kBlock, &r0, &r1);
  double v[2] = {0.0, 0.0};
  for (long long i = r0 + threadIdx.x; i < r1; i += kBlock) {
    const double p = Pn[i] + dt * Pdot[i];
    Pnp1[i] = p;
    const double d = p - Pprev[i];
    v[0] += p * p; v[1] += d * d;
  }
  double out[2];
  if (grid_sum<2>(v, partials, &st->cnt[0], out) && threadIdx.x == 0) { st->pp = out[0]; st->pd = out[1]; }
}

int launch_pnp1(fsx_ctx* ctx) {
  hipLaunchKernelGGL(k_pnp1, dim3(row_grid(ctx->NNp, kBlock)), dim3(kBlock), 0, ctx->stream, ctx->NNp, ctx->Pn, ctx->Pdot,
                     ctx->Pnp1_prev, ctx->Pnp1, ctx->dt, ctx->partials, ctx->st);
  return launch_check(ctx, "k_pnp1");
}

// ------------------------------------------------------------ step 3
// R3_c = -dt G_c Pdot - [iter>1] dt KAcc_c; BC rows 0; Acc = R3 MdInv;
// Unp1 = UnpHalf + dt Acc (BCH:3912-3974), plus the velocity norms of the
// convergence check (BCH:2936-2943).
template <int LPR>
__global__ __launch_bounds__(kBlock) void k_step3(int NN, const int* __restrict__ Grow, const int* __restrict__ Gcol,
                                                  const double* __restrict__ G, const double* __restrict__ Pdot,
                                                  const double* __restrict__ KAcc, int use_kacc,
                                                  const unsigned char* __restrict__ bc, const double* __restrict__ MdInv,
                                                  const double* __restrict__ UnpHalf, const double* __restrict__ Uprev1,
                                                  double* __restrict__ Acc, double* __restrict__ Unp1, double dt,
                                                  double* partials, StepScalars* st) {
  constexpr int RPB = kBlock / LPR;
  long long r0, r1;
  block_rows(NN, xcd_block(), gridDim.x, RPB, &r0, &r1);
  const int grp = threadIdx.x / LPR, l = threadIdx.x % LPR;
  double v[2] = {0.0, 0.0};
  for (long long base = r0; base < r1; base += RPB) {
    const long long r = base + grp;
    double g0 = 0, g1 = 0, g2 = 0;
    if (r < r1) {
      const int eg = Grow[r + 1];
      for (int k = Grow[r] + l; k < eg; k += LPR) {
        const double p = Pdot[Gcol[k]];
        g0 += G[3 * (size_t)k] * p; g1 += G[3 * (size_t)k + 1] * p; g2 += G[3 * (size_t)k + 2] * p;
      }
    }
    g0 = group_sum<LPR>(g0); g1 = group_sum<LPR>(g1); g2 = group_sum<LPR>(g2);
    if (r < r1 && l < 3) {
      const double gg = l == 0 ? g0 : (l == 1 ? g1 : g2);
      const size_t i = (size_t)l * NN + r;
      double R = -dt * gg;
      if (use_kacc) R = R - dt * KAcc[i];
      if (bc[r]) R = 0.0;
      const double a = R * MdInv[r];
      Acc[i] = a;
      const double u = UnpHalf[i] + dt * a;
      Unp1[i] = u;
      const double d = u - Uprev1[i];
      v[0] += u * u; v[1] += d * d;
    }
  }
  double out[2];
  if (grid_sum<2>(v, partials, &st->cnt[1], out) && threadIdx.x == 0) { st->uu = out[0]; st->ud = out[1]; }
}

int launch_step3(fsx_ctx* ctx, int iter) {
  constexpr int LPR = 4;
  hipLaunchKernelGGL(k_step3<LPR>, dim3(row_grid(ctx->NN, kBlock / LPR)), dim3(kBlock), 0, ctx->stream, ctx->NN, ctx->Grow,
                     ctx->Gcol, ctx->G, ctx->Pdot, ctx->KAcc, iter != 1 ? 1 : 0, ctx->bcflag, ctx->MdInv, ctx->UnpHalf,
                     ctx->Unp1_prev, ctx->Acc, ctx->Unp1, ctx->dt, ctx->partials, ctx->st);
  return launch_check(ctx, "k_step3");
}

// max |Unp1 - Un| / dt over 3NN (BCH:3048-3057)
__global__ __launch_bounds__(kBlock) void k_maxacc(long long n, const double* __restrict__ Unp1, const double* __restrict__ Un,
                                                   double oneOverdt, double* partials, StepScalars* st) {
  long long r0, r1;
  block_rows(n, xcd_block(), gridDim.x, kBlock, &r0, &r1);
  double m = 0.0;
  for (long long i = r0 + threadIdx.x; i < r1; i += kBlock) m = fmax(m, fabs((Unp1[i] - Un[i]) * oneOverdt));
  double out;
  if (grid_max(m, partials, &st->cnt[2], &out) && threadIdx.x == 0) st->maxacc = out;
}

int launch_maxacc(fsx_ctx* ctx) {
  const long long n = 3LL * ctx->NN;
  hipLaunchKernelGGL(k_maxacc, dim3(row_grid(n, kBlock * 4)), dim3(kBlock), 0, ctx->stream, n, ctx->Unp1, ctx->Un,
                     1.0000000000000000 / ctx->dt, ctx->partials, ctx->st);
  return launch_check(ctx, "k_maxacc");
}

// ------------------------------------------------------------ generic SpMVs (tooling)
// y = alpha * A x + beta * y for one CSR value array with value stride `vs`
// (G components are interleaved by 3).
template <int LPR>
__global__ __launch_bounds__(kBlock) void k_csrmv(int nrows, const int* __restrict__ rp, const int* __restrict__ ci,
                                                  const double* __restrict__ v, int vs, const double* __restrict__ x,
                                                  double* __restrict__ y, double alpha, double beta) {
  constexpr int RPB = kBlock / LPR;
  long long r0, r1;
  block_rows(nrows, xcd_block(), gridDim.x, RPB, &r0, &r1);
  const int grp = threadIdx.x / LPR, l = threadIdx.x % LPR;
  for (long long base = r0; base < r1; base += RPB) {
    const long long r = base + grp;
    double s = 0;
    if (r < r1)
      for (int k = rp[r] + l; k < rp[r + 1]; k += LPR) s += v[(size_t)vs * k] * x[ci[k]];
    s = group_sum<LPR>(s);
    if (r < r1 && l == 0) y[r] = (beta == 0.0 ? 0.0 : beta * y[r]) + alpha * s;
  }
}

int spmv_host_op(fsx_ctx* ctx, int which, const double* dx, double* dy, double alpha, double beta) {
  constexpr int LPR = 16;
  if (which == 0) {
    hipLaunchKernelGGL(k_csrmv<LPR>, dim3(row_grid(ctx->NN, kBlock / LPR)), dim3(kBlock), 0, ctx->stream, ctx->NN, ctx->Mrow,
                       ctx->Mcol, ctx->K, 1, dx, dy, alpha, beta);
  } else if (which >= 1 && which <= 3) {
    hipLaunchKernelGGL(k_csrmv<LPR>, dim3(row_grid(ctx->NN, kBlock / LPR)), dim3(kBlock), 0, ctx->stream, ctx->NN, ctx->Grow,
                       ctx->Gcol, ctx->G + (which - 1), 3, dx, dy, alpha, beta);
  } else if (which == 4) {
    hipLaunchKernelGGL(k_csrmv<LPR>, dim3(row_grid(ctx->NNp, kBlock / LPR)), dim3(kBlock), 0, ctx->stream, ctx->NNp,
                       ctx->Zrow, ctx->Zcol, ctx->Z, 1, dx, dy, alpha, beta);
  } else {
    hipLaunchKernelGGL(k_csrmv<LPR>, dim3(row_grid(ctx->NNp, kBlock / LPR)), dim3(kBlock), 0, ctx->stream, ctx->NNp,
                       ctx->GTrow, ctx->GTcol, ctx->GT + (which - 5), 3, dx, dy, alpha, beta);
  }
  return launch_check(ctx, "k_csrmv");
}

// ------------------------------------------------------------ Jacobi-PCG
// Textbook preconditioned CG with x0 = 0 (BCH:3999-4004), Jacobi
// preconditioner (ipar[10] = 1, BCH:4017) and the stop test
// ||r||_2 <= rtol ||b||_2 checked after each update. Three kernels per
// iteration, all scalars stay on the device; launches after convergence
// return at once (done != 0), so the host only polls every few iterations.
// T is the storage/arithmetic type of Z and the vectors (double, or float
// for the inner solve of mixed precision); dot products accumulate in FP64.
template <class T>
__global__ __launch_bounds__(kBlock) void k_pcg_init(int n, const double* __restrict__ b, const T* __restrict__ dinv,
                                                     T* __restrict__ x, T* __restrict__ r, T* __restrict__ p,
                                                     double rtol, int maxit, double* partials, SolveScalars* sc) {
  long long r0, r1;
  block_rows(n, xcd_block(), gridDim.x, kBlock, &r0, &r1);
  double v[2] = {0.0, 0.0};
  for (long long i = r0 + threadIdx.x; i < r1; i += kBlock) {
    const T bi = (T)b[i];
    const T z = dinv[i] * bi;
    x[i] = (T)0; r[i] = bi; p[i] = z;
    v[0] += (double)bi * (double)z;
    v[1] += (double)bi * (double)bi;
  }
  double out[2];
  if (grid_sum<2>(v, partials, &sc->cnt[0], out) && threadIdx.x == 0) {
    sc->rz = out[0]; sc->bb = out[1]; sc->rr = out[1];
    sc->rtol2bb = rtol * rtol * out[1];
    sc->iters = 0; sc->maxit = maxit;
    sc->done = (out[1] == 0.0) ? 1 : 0;
    sc->alpha = 0.0; sc->beta = 0.0;
  }
}

// q = Z p and p.q; alpha = rz / pq
template <class T, int LPR>
__global__ __launch_bounds__(kBlock) void k_pcg_spmv(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                     const T* __restrict__ val, const T* __restrict__ p, T* __restrict__ q,
                                                     double* partials, SolveScalars* sc) {
  if (sc->done) return;
  constexpr int RPB = kBlock / LPR;
  long long r0, r1;
  block_rows(n, xcd_block(), gridDim.x, RPB, &r0, &r1);
  const int grp = threadIdx.x / LPR, l = threadIdx.x % LPR;
  double v[1] = {0.0};
  for (long long base = r0; base < r1; base += RPB) {
    const long long r = base + grp;
    T s = 0;
    if (r < r1) {
      const int e = rp[r + 1];
      for (int k = rp[r] + l; k < e; k += LPR) s += val[k] * p[ci[k]];
    }
    if constexpr (sizeof(T) == 8) s = group_sum<LPR>(s); else s = group_sumf<LPR>(s);
    if (r < r1 && l == 0) {
      q[r] = s;
      v[0] += (double)p[r] * (double)s;
    }
  }
  double out[1];
  if (grid_sum<1>(v, partials, &sc->cnt[1], out) && threadIdx.x == 0) {
    sc->pq = out[0];
    sc->alpha = sc->rz / out[0];
  }
}

// x += alpha p; r -= alpha q; z = dinv r; rr, rz_new; convergence / beta
template <class T>
__global__ __launch_bounds__(kBlock) void k_pcg_update(int n, T* __restrict__ x, T* __restrict__ r,
                                                       const T* __restrict__ p, const T* __restrict__ q,
                                                       const T* __restrict__ dinv, double* partials, SolveScalars* sc) {
  if (sc->done) return;
  long long r0, r1;
  block_rows(n, xcd_block(), gridDim.x, kBlock, &r0, &r1);
  const T a = (T)sc->alpha;
  double v[2] = {0.0, 0.0};
  for (long long i = r0 + threadIdx.x; i < r1; i += kBlock) {
    x[i] += a * p[i];
    const T ri = r[i] - a * q[i];
    r[i] = ri;
    const T z = dinv[i] * ri;
    v[0] += (double)ri * (double)ri;
    v[1] += (double)ri * (double)z;
  }
  double out[2];
  if (grid_sum<2>(v, partials, &sc->cnt[2], out) && threadIdx.x == 0) {
    sc->iters += 1;
    sc->rr = out[0];
    if (!(out[0] == out[0]) || !(out[1] == out[1]) || !(sc->alpha == sc->alpha)) sc->done = 3;   // NaN
    else if (out[0] <= sc->rtol2bb) sc->done = 1;
    else if (sc->iters >= sc->maxit) sc->done = 2;
    else {
      sc->beta = out[1] / sc->rz;
      sc->rz = out[1];
    }
  }
}

// p = dinv r + beta p
template <class T>
__global__ __launch_bounds__(kBlock) void k_pcg_pupdate(int n, const T* __restrict__ r, const T* __restrict__ dinv,
                                                        T* __restrict__ p, const SolveScalars* sc) {
  if (sc->done) return;
  long long r0, r1;
  block_rows(n, xcd_block(), gridDim.x, kBlock, &r0, &r1);
  const T be = (T)sc->beta;
  for (long long i = r0 + threadIdx.x; i < r1; i += kBlock) p[i] = dinv[i] * r[i] + be * p[i];
}

// Runs one Jacobi-PCG solve of Z x = b to completion. b is FP64 (converted to
// T on load); x is the T-typed solution. Returns the iterations and the
// recurrence residual ||r||/||b||.
template <class T>
static int pcg_run(fsx_ctx* ctx, const T* val, const T* dinv, const double* b, T* x, T* r, T* p, T* q, double rtol,
                   int maxit, int* iters_out, double* relres_out, int* status_out) {
  constexpr int LPR = 16;
  const int n = ctx->NNp;
  const int g_rows = row_grid(n, kBlock / LPR), g_vec = row_grid(n, kBlock);
  if (ctx->timing) (void)hipEventRecord(ctx->ev0, ctx->stream);
  hipLaunchKernelGGL(k_pcg_init<T>, dim3(g_vec), dim3(kBlock), 0, ctx->stream, n, b, dinv, x, r, p, rtol, maxit,
                     ctx->partials, ctx->sc);
  FSX_TRY(launch_check(ctx, "k_pcg_init"));
  int issued = 0;
  int chunk = ctx->last_iters > 16 ? (int)(0.9 * ctx->last_iters) : 16;
  long long spmv_launches = 0;
  for (;;) {
    for (int j = 0; j < chunk; j++) {
      hipLaunchKernelGGL((k_pcg_spmv<T, LPR>), dim3(g_rows), dim3(kBlock), 0, ctx->stream, n, ctx->Zrow, ctx->Zcol, val, p,
                         q, ctx->partials, ctx->sc);
      hipLaunchKernelGGL(k_pcg_update<T>, dim3(g_vec), dim3(kBlock), 0, ctx->stream, n, x, r, p, q, dinv, ctx->partials,
                         ctx->sc);
      hipLaunchKernelGGL(k_pcg_pupdate<T>, dim3(g_vec), dim3(kBlock), 0, ctx->stream, n, r, dinv, p, ctx->sc);
    }
    spmv_launches += chunk;
    issued += chunk;
    FSX_TRY(launch_check(ctx, "pcg iteration"));
    FSX_HIP(hipMemcpyAsync(ctx->h_sc, ctx->sc, sizeof(SolveScalars), hipMemcpyDeviceToHost, ctx->stream));
    FSX_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->h_sc->done) break;
    if (issued > maxit + 64) return fail(ctx, FSX_ESTATE, "PCG did not terminate");
    chunk = 16;
  }
  if (ctx->timing) {
    (void)hipEventRecord(ctx->ev1, ctx->stream);
    (void)hipEventSynchronize(ctx->ev1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    ctx->tm.pcg_ms += ms;
    ctx->tm.pcg_spmv_launches += ctx->h_sc->iters;
    ctx->tm.pcg_iters += ctx->h_sc->iters;
  }
  *iters_out = ctx->h_sc->iters;
  *relres_out = ctx->h_sc->bb > 0 ? std::sqrt(ctx->h_sc->rr / ctx->h_sc->bb) : 0.0;
  *status_out = ctx->h_sc->done;
  ctx->last_iters = ctx->h_sc->iters;
  (void)spmv_launches;
  return FSX_OK;
}

// ------------------------------------------------------------ SpMV timing
__global__ void k_arm(SolveScalars* sc) {
  sc->done = 0;
  sc->rz = 1.0;
}

template <class T>
__global__ void k_fill_t(T* __restrict__ p, T v, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

template <class T>
static int bench_spmv_t(fsx_ctx* ctx, const T* val, T* p, T* q, int reps, double* ms) {
  constexpr int LPR = 16;
  const int n = ctx->NNp;
  hipLaunchKernelGGL(k_fill_t<T>, dim3(1024), dim3(kBlock), 0, ctx->stream, p, (T)1, n);
  hipLaunchKernelGGL(k_arm, dim3(1), dim3(1), 0, ctx->stream, ctx->sc);
  const int g_rows = row_grid(n, kBlock / LPR);
  for (int w = 0; w < 3; w++)
    hipLaunchKernelGGL((k_pcg_spmv<T, LPR>), dim3(g_rows), dim3(kBlock), 0, ctx->stream, n, ctx->Zrow, ctx->Zcol, val, p, q,
                       ctx->partials, ctx->sc);
  FSX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL((k_pcg_spmv<T, LPR>), dim3(g_rows), dim3(kBlock), 0, ctx->stream, n, ctx->Zrow, ctx->Zcol, val, p, q,
                       ctx->partials, ctx->sc);
  FSX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  FSX_HIP(hipEventSynchronize(ctx->ev1));
  FSX_TRY(launch_check(ctx, "k_pcg_spmv (timing)"));
  float t = 0;
  FSX_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
  *ms = t / reps;
  return FSX_OK;
}

// ------------------------------------------------------------ mixed precision
__global__ void k_to_float(int n, const double* __restrict__ a, float* __restrict__ b) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = (float)a[i];
}
__global__ void k_add_float(int n, double* __restrict__ x, const float* __restrict__ d) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] += (double)d[i];
}

// r = b - Z x and ||r||^2 (FP64), the refinement residual.
template <int LPR>
__global__ __launch_bounds__(kBlock) void k_residual(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                     const double* __restrict__ val, const double* __restrict__ x,
                                                     const double* __restrict__ b, double* __restrict__ r,
                                                     double* partials, SolveScalars* sc) {
  constexpr int RPB = kBlock / LPR;
  long long r0, r1;
  block_rows(n, xcd_block(), gridDim.x, RPB, &r0, &r1);
  const int grp = threadIdx.x / LPR, l = threadIdx.x % LPR;
  double v[1] = {0.0};
  for (long long base = r0; base < r1; base += RPB) {
    const long long i = base + grp;
    double s = 0;
    if (i < r1)
      for (int k = rp[i] + l; k < rp[i + 1]; k += LPR) s += val[k] * x[ci[k]];
    s = group_sum<LPR>(s);
    if (i < r1 && l == 0) {
      const double ri = b[i] - s;
      r[i] = ri;
      v[0] += ri * ri;
    }
  }
  double out[1];
  if (grid_sum<1>(v, partials, &sc->cnt[3], out) && threadIdx.x == 0) sc->rr = out[0];
}

int pressure_solve(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres) {
  const int n = ctx->NNp;
  int status = 0;
  if (ctx->solver_mode == FSX_SOLVER_FP64_JACOBI) {
    FSX_TRY(pcg_run<double>(ctx, ctx->Z, ctx->Zdinv, b, x, ctx->wr, ctx->wp, ctx->wq, ctx->rtol, ctx->maxit, iters, relres,
                            &status));
    if (status == 3) return fail(ctx, FSX_ENUMERIC, "PCG met NaN/Inf");
    ctx->last_relres = *relres;
    if (status == 2) { ctx->err = "PCG reached maxit"; return FSX_ENOCONV; }
    return FSX_OK;
  }
  // FSX_SOLVER_MIXED_IR: FP32 inner Jacobi-PCG, FP64 true residual, FP64 correction
  constexpr int LPR = 16;
  const double eta = 1e-3;            // inner relative tolerance
  const int g_vec = std::max(1, std::min((n + kBlock - 1) / kBlock, 4096));
  FSX_TRY(launch_fill(ctx, x, 0.0, n));
  FSX_TRY(launch_copy(ctx, ctx->wr, b, n));
  // ||b||^2 via the residual kernel with x = 0
  hipLaunchKernelGGL(k_residual<LPR>, dim3(row_grid(n, kBlock / LPR)), dim3(kBlock), 0, ctx->stream, n, ctx->Zrow, ctx->Zcol,
                     ctx->Z, x, b, ctx->wr, ctx->partials, ctx->sc);
  FSX_HIP(hipMemcpyAsync(ctx->h_sc, ctx->sc, sizeof(SolveScalars), hipMemcpyDeviceToHost, ctx->stream));
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  const double bb = ctx->h_sc->rr;
  double rr = bb;
  int total = 0, outer = 0;
  const int saved_last = ctx->last_iters;
  while (bb > 0.0 && rr > ctx->rtol * ctx->rtol * bb) {
    if (total >= ctx->maxit || outer >= 100) { status = 2; break; }
    hipLaunchKernelGGL(k_to_float, dim3(g_vec), dim3(kBlock), 0, ctx->stream, n, ctx->wr, ctx->fr);
    // inner solve: Zf d = r  (d in fx); the inner b is the FP64 residual
    int it = 0, st = 0;
    double rel = 0;
    ctx->last_iters = saved_last > 0 ? saved_last / 2 : 0;
    FSX_TRY(pcg_run<float>(ctx, ctx->Zf, ctx->Zdinvf, ctx->wr, ctx->fx, ctx->fr, ctx->fp, ctx->fq, eta, ctx->maxit - total,
                           &it, &rel, &st));
    if (st == 3) return fail(ctx, FSX_ENUMERIC, "inner PCG met NaN/Inf");
    total += it;
    hipLaunchKernelGGL(k_add_float, dim3(g_vec), dim3(kBlock), 0, ctx->stream, n, x, ctx->fx);
    hipLaunchKernelGGL(k_residual<LPR>, dim3(row_grid(n, kBlock / LPR)), dim3(kBlock), 0, ctx->stream, n, ctx->Zrow,
                       ctx->Zcol, ctx->Z, x, b, ctx->wr, ctx->partials, ctx->sc);
    FSX_TRY(launch_check(ctx, "mixed refinement"));
    FSX_HIP(hipMemcpyAsync(ctx->h_sc, ctx->sc, sizeof(SolveScalars), hipMemcpyDeviceToHost, ctx->stream));
    FSX_HIP(hipStreamSynchronize(ctx->stream));
    rr = ctx->h_sc->rr;
    if (!(rr == rr)) return fail(ctx, FSX_ENUMERIC, "refinement residual is NaN");
    outer++;
  }
  *iters = total;
  *relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
  ctx->last_iters = total;
  ctx->last_relres = *relres;
  if (status == 2) { ctx->err = "mixed-precision solve reached maxit"; return FSX_ENOCONV; }
  return FSX_OK;
}

}  // namespace fsx

extern "C" int fsx_bench_spmv(fsx_ctx* ctx, int precision, int reps, double* ms_per_launch, double* bytes_per_launch) {
  using namespace fsx;
  if (!ctx || reps < 1 || !ms_per_launch) return FSX_EINVAL;
  if (!ctx->z_built) return fail(ctx, FSX_ESTATE, "build Z first");
  double ms = 0;
  const double nnz = (double)ctx->nnzZ, n = ctx->NNp;
  double bytes;
  if (precision == 0) {
    FSX_TRY(bench_spmv_t<double>(ctx, ctx->Z, ctx->wp, ctx->wq, reps, &ms));
    bytes = 12.0 * nnz + 4.0 * (n + 1) + 16.0 * n;
  } else {
    FSX_TRY(bench_spmv_t<float>(ctx, ctx->Zf, ctx->fp, ctx->fq, reps, &ms));
    bytes = 8.0 * nnz + 4.0 * (n + 1) + 8.0 * n;
  }
  *ms_per_launch = ms;
  if (bytes_per_launch) *bytes_per_launch = bytes;
  return FSX_OK;
}
