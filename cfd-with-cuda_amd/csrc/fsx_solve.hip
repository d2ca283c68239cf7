// fsx_solve.hip -- CSR row kernels of the time step (fused SpMV + elementwise
// epilogues), deterministic reductions and the Jacobi-PCG pressure solvers.
//
// Row kernels map a group of LPR lanes to one CSR row. Each lane walks the
// row in rounds of U entries whose column/value loads are issued together
// (non-temporal: they stream once) before the dependent gathers and FMAs, so
// a wavefront keeps several HBM requests in flight instead of one dependent
// chain. Row pointers of the next pass are prefetched. Blocks own contiguous
// row ranges and are renumbered so that each XCD walks one contiguous eighth
// of the rows (its gathers then stay in its own L2).
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

// Row-group iteration scaffold: calls body(r, rs, re) for every row of the
// block's range (r >= r1 marks an idle group in the last pass, rs = re = 0),
// with the next pass's row pointers loaded ahead.
template <int LPR, class Body>
__device__ __forceinline__ void for_rows(long long nrows, const int* __restrict__ rp, Body&& body) {
  constexpr int RPB = kBlock / LPR;
  long long r0, r1;
  block_rows(nrows, xcd_block(), gridDim.x, RPB, &r0, &r1);
  const int grp = threadIdx.x / LPR;
  long long r = r0 + grp;
  int rs = r < r1 ? rp[r] : 0, re = r < r1 ? rp[r + 1] : 0;
  for (long long base = r0; base < r1; base += RPB) {
    const long long rn = base + RPB + grp;
    const int rsn = rn < r1 ? rp[rn] : 0, ren = rn < r1 ? rp[rn + 1] : 0;
    body(base + grp, r1, rs, re);
    rs = rsn; re = ren;
  }
}

// ------------------------------------------------------------ step 1
// R1 <- (R1 - K UnpHalf_prev) - G Pn; BC rows 0; UnpHalf = Un + dt R1 MdInv.
// (BCH:3746-3783; the convection part of R1 is already in R1.)
template <int LPR, int UK, int UG>
__global__ __launch_bounds__(kBlock) void k_step1(int NN, const int* __restrict__ Mrow, const int* __restrict__ Mcol,
                                                  const double* __restrict__ K, const int* __restrict__ Grow,
                                                  const int* __restrict__ Gcol, const double* __restrict__ G,
                                                  const double* __restrict__ Uprev, const double* __restrict__ Pn,
                                                  const double* __restrict__ Un, double* __restrict__ R1,
                                                  const double* __restrict__ MdInv, const unsigned char* __restrict__ bc,
                                                  double* __restrict__ UnpHalf, double dt) {
  const int l = threadIdx.x % LPR;
  const size_t n1 = NN, n2 = 2 * (size_t)NN;
  for_rows<LPR>(NN, Mrow, [&](long long r, long long r1, int rs, int re) {
    double k0 = 0, k1 = 0, k2 = 0, g0 = 0, g1 = 0, g2 = 0;
    for (int base = rs + l; base < re; base += UK * LPR) {
      int c[UK];
      double a[UK];
#pragma unroll
      for (int j = 0; j < UK; j++) {
        const int k = base + j * LPR, kk = k < re ? k : re - 1;
        c[j] = ld_stream(Mcol + kk);
        a[j] = k < re ? ld_stream(K + kk) : 0.0;
      }
#pragma unroll
      for (int j = 0; j < UK; j++) {
        k0 += a[j] * Uprev[c[j]]; k1 += a[j] * Uprev[n1 + c[j]]; k2 += a[j] * Uprev[n2 + c[j]];
      }
    }
    const int gs = r < r1 ? Grow[r] : 0, ge = r < r1 ? Grow[r + 1] : 0;
    for (int base = gs + l; base < ge; base += UG * LPR) {
      int c[UG];
      double a0[UG], a1[UG], a2[UG];
#pragma unroll
      for (int j = 0; j < UG; j++) {
        const int k = base + j * LPR, kk = k < ge ? k : ge - 1;
        c[j] = ld_stream(Gcol + kk);
        const bool ok = k < ge;
        a0[j] = ok ? ld_stream(G + 3 * (size_t)kk) : 0.0;
        a1[j] = ok ? ld_stream(G + 3 * (size_t)kk + 1) : 0.0;
        a2[j] = ok ? ld_stream(G + 3 * (size_t)kk + 2) : 0.0;
      }
#pragma unroll
      for (int j = 0; j < UG; j++) {
        const double p = Pn[c[j]];
        g0 += a0[j] * p; g1 += a1[j] * p; g2 += a2[j] * p;
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
  });
}

int launch_step1(fsx_ctx* ctx, int iter) {
  (void)iter;
  constexpr int LPR = 16;
  hipLaunchKernelGGL((k_step1<LPR, 4, 2>), dim3(row_grid(ctx->NN, kBlock / LPR)), dim3(kBlock), 0, ctx->stream, ctx->NN,
                     ctx->Mrow, ctx->Mcol, ctx->K, ctx->Grow, ctx->Gcol, ctx->G, ctx->UnpHalf_prev, ctx->Pn, ctx->Un,
                     ctx->R1, ctx->MdInv, ctx->bcflag, ctx->UnpHalf, ctx->dt);
  return launch_check(ctx, "k_step1");
}

// ------------------------------------------------------------ K * Acc_prev
// KAcc_c = K Acc_prev_c, c = 0..2, one pass over K (BCH:2981-3035)
template <int LPR, int U>
__global__ __launch_bounds__(kBlock) void k_k3(int NN, const int* __restrict__ Mrow, const int* __restrict__ Mcol,
                                               const double* __restrict__ K, const double* __restrict__ x,
                                               double* __restrict__ y, double alpha, double beta) {
  const int l = threadIdx.x % LPR;
  const size_t n1 = NN, n2 = 2 * (size_t)NN;
  for_rows<LPR>(NN, Mrow, [&](long long r, long long r1, int rs, int re) {
    double k0 = 0, k1 = 0, k2 = 0;
    for (int base = rs + l; base < re; base += U * LPR) {
      int c[U];
      double a[U];
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int k = base + j * LPR, kk = k < re ? k : re - 1;
        c[j] = ld_stream(Mcol + kk);
        a[j] = k < re ? ld_stream(K + kk) : 0.0;
      }
#pragma unroll
      for (int j = 0; j < U; j++) { k0 += a[j] * x[c[j]]; k1 += a[j] * x[n1 + c[j]]; k2 += a[j] * x[n2 + c[j]]; }
    }
    k0 = group_sum<LPR>(k0); k1 = group_sum<LPR>(k1); k2 = group_sum<LPR>(k2);
    if (r < r1 && l < 3) {
      const double kk = l == 0 ? k0 : (l == 1 ? k1 : k2);
      const size_t i = (size_t)l * NN + r;
      y[i] = (beta == 0.0 ? 0.0 : beta * y[i]) + alpha * kk;
    }
  });
}

int launch_kacc(fsx_ctx* ctx) {
  constexpr int LPR = 16;
  hipLaunchKernelGGL((k_k3<LPR, 4>), dim3(row_grid(ctx->NN, kBlock / LPR)), dim3(kBlock), 0, ctx->stream, ctx->NN,
                     ctx->Mrow, ctx->Mcol, ctx->K, ctx->Acc_prev, ctx->KAcc, 1.0, 0.0);
  return launch_check(ctx, "k_k3");
}

// ------------------------------------------------------------ step 2 RHS
// R2[p] = sum_c G_c^T d_c, d_c = UnpHalf_c / dt^2 - [iter>1] MdOrigInv KAcc_c,
// summed as (c0 + c1) + c2 like the three csrmv calls of BCH:3864-3868;
// R2[p0] = 0 for p0 > 0 (BCH:4180-4184).
template <int LPR, int U>
__global__ __launch_bounds__(kBlock) void k_step2_rhs(int NNp, int NN, const int* __restrict__ GTrow,
                                                      const int* __restrict__ GTcol, const double* __restrict__ GT,
                                                      const double* __restrict__ UnpHalf, const double* __restrict__ MdOrigInv,
                                                      const double* __restrict__ KAcc, int use_kacc, double o2,
                                                      int zeroP, double* __restrict__ R2) {
  const int l = threadIdx.x % LPR;
  const size_t n1 = NN, n2 = 2 * (size_t)NN;
  for_rows<LPR>(NNp, GTrow, [&](long long p, long long r1, int rs, int re) {
    double s0 = 0, s1 = 0, s2 = 0;
    for (int base = rs + l; base < re; base += U * LPR) {
      int c[U];
      double a0[U], a1[U], a2[U];
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int k = base + j * LPR, kk = k < re ? k : re - 1;
        c[j] = ld_stream(GTcol + kk);
        const bool ok = k < re;
        a0[j] = ok ? ld_stream(GT + 3 * (size_t)kk) : 0.0;
        a1[j] = ok ? ld_stream(GT + 3 * (size_t)kk + 1) : 0.0;
        a2[j] = ok ? ld_stream(GT + 3 * (size_t)kk + 2) : 0.0;
      }
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int i = c[j];
        double d0 = o2 * UnpHalf[i], d1 = o2 * UnpHalf[n1 + i], d2 = o2 * UnpHalf[n2 + i];
        if (use_kacc) {
          const double m = MdOrigInv[i];
          d0 = d0 - m * KAcc[i]; d1 = d1 - m * KAcc[n1 + i]; d2 = d2 - m * KAcc[n2 + i];
        }
        s0 += a0[j] * d0; s1 += a1[j] * d1; s2 += a2[j] * d2;
      }
    }
    s0 = group_sum<LPR>(s0); s1 = group_sum<LPR>(s1); s2 = group_sum<LPR>(s2);
    if (p < r1 && l == 0) R2[p] = (zeroP > 0 && p == zeroP) ? 0.0 : (s0 + s1) + s2;
  });
}

int launch_step2_rhs(fsx_ctx* ctx, int iter) {
  constexpr int LPR = 16;
  const double o2 = 1.0000000000000000 / (ctx->dt * ctx->dt);
  hipLaunchKernelGGL((k_step2_rhs<LPR, 8>), dim3(row_grid(ctx->NNp, kBlock / LPR)), dim3(kBlock), 0, ctx->stream, ctx->NNp,
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
template <int LPR, int U>
__global__ __launch_bounds__(kBlock) void k_step3(int NN, const int* __restrict__ Grow, const int* __restrict__ Gcol,
                                                  const double* __restrict__ G, const double* __restrict__ Pdot,
                                                  const double* __restrict__ KAcc, int use_kacc,
                                                  const unsigned char* __restrict__ bc, const double* __restrict__ MdInv,
                                                  const double* __restrict__ UnpHalf, const double* __restrict__ Uprev1,
                                                  double* __restrict__ Acc, double* __restrict__ Unp1, double dt,
                                                  double* partials, StepScalars* st) {
  const int l = threadIdx.x % LPR;
  double v[2] = {0.0, 0.0};
  for_rows<LPR>(NN, Grow, [&](long long r, long long r1, int rs, int re) {
    double g0 = 0, g1 = 0, g2 = 0;
    for (int base = rs + l; base < re; base += U * LPR) {
      int c[U];
      double a0[U], a1[U], a2[U];
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int k = base + j * LPR, kk = k < re ? k : re - 1;
        c[j] = ld_stream(Gcol + kk);
        const bool ok = k < re;
        a0[j] = ok ? ld_stream(G + 3 * (size_t)kk) : 0.0;
        a1[j] = ok ? ld_stream(G + 3 * (size_t)kk + 1) : 0.0;
        a2[j] = ok ? ld_stream(G + 3 * (size_t)kk + 2) : 0.0;
      }
#pragma unroll
      for (int j = 0; j < U; j++) {
        const double p = Pdot[c[j]];
        g0 += a0[j] * p; g1 += a1[j] * p; g2 += a2[j] * p;
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
  });
  double out[2];
  if (grid_sum<2>(v, partials, &st->cnt[1], out) && threadIdx.x == 0) { st->uu = out[0]; st->ud = out[1]; }
}

int launch_step3(fsx_ctx* ctx, int iter) {
  constexpr int LPR = 4;
  hipLaunchKernelGGL((k_step3<LPR, 4>), dim3(row_grid(ctx->NN, kBlock / LPR)), dim3(kBlock), 0, ctx->stream, ctx->NN,
                     ctx->Grow, ctx->Gcol, ctx->G, ctx->Pdot, ctx->KAcc, iter != 1 ? 1 : 0, ctx->bcflag, ctx->MdInv,
                     ctx->UnpHalf, ctx->Unp1_prev, ctx->Acc, ctx->Unp1, ctx->dt, ctx->partials, ctx->st);
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
  const int l = threadIdx.x % LPR;
  for_rows<LPR>(nrows, rp, [&](long long r, long long r1, int rs, int re) {
    double s = 0;
    for (int k = rs + l; k < re; k += LPR) s += v[(size_t)vs * k] * x[ci[k]];
    s = group_sum<LPR>(s);
    if (r < r1 && l == 0) y[r] = (beta == 0.0 ? 0.0 : beta * y[r]) + alpha * s;
  });
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
// ||r||_2 <= rtol ||b||_2 after each update. Two kernels per iteration:
//   A: p_new = z + beta p_old (own row), q = Z p_new with the gathered
//      z + beta p_old, p_new . q  ->  alpha = rz / pq
//   B: x += alpha p_new, r -= alpha q, z = dinv r, r.r, r.z  ->  stop test,
//      beta = rz_new / rz
// p is double-buffered (A reads p_old at neighbours while rows write p_new).
// All scalars stay on the device; launches after the stop test return at
// once (done != 0), so the host only polls every few iterations.
// T is the type of Z and the vectors (double, or float for the inner solve
// of the mixed-precision solver); dot products accumulate in FP64.
constexpr int kPcgLPR = 16;
constexpr int kPcgU = 8;

template <class T>
__global__ __launch_bounds__(kBlock) void k_pcg_init(int n, const double* __restrict__ b, const T* __restrict__ dinv,
                                                     T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                                                     T* __restrict__ p_old, double rtol, int maxit, double* partials,
                                                     SolveScalars* sc) {
  long long r0, r1;
  block_rows(n, xcd_block(), gridDim.x, kBlock, &r0, &r1);
  double v[2] = {0.0, 0.0};
  for (long long i = r0 + threadIdx.x; i < r1; i += kBlock) {
    const T bi = (T)b[i];
    const T zi = dinv[i] * bi;
    x[i] = (T)0; r[i] = bi; z[i] = zi; p_old[i] = (T)0;
    v[0] += (double)bi * (double)zi;
    v[1] += (double)bi * (double)bi;
  }
  double out[2];
  if (grid_sum<2>(v, partials, &sc->cnt[0], out) && threadIdx.x == 0) {
    sc->rz = out[0]; sc->bb = out[1]; sc->rr = out[1]; sc->rr_anchor = out[1];
    sc->rtol2bb = rtol * rtol * out[1];
    sc->iters = 0; sc->maxit = maxit;
    sc->done = (out[1] == 0.0) ? 1 : 0;
    sc->alpha = 0.0; sc->beta = 0.0;
  }
}

template <class T, int LPR, int U>
__global__ __launch_bounds__(kBlock) void k_pcg_a(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                  const T* __restrict__ val, const T* __restrict__ z, T* p0, T* p1,
                                                  T* __restrict__ q, double* partials, SolveScalars* sc) {
  if (sc->done) return;
  const T be = (T)sc->beta;
  // iteration i (0-based, = iterations executed so far) reads p_(i%2), writes the other
  const bool odd = sc->iters & 1;
  const T* __restrict__ p_old = odd ? p1 : p0;
  T* __restrict__ p_new = odd ? p0 : p1;
  const int l = threadIdx.x % LPR;
  double v[1] = {0.0};
  for_rows<LPR>(n, rp, [&](long long r, long long r1, int rs, int re) {
    T s = 0;
    for (int base = rs + l; base < re; base += U * LPR) {
      int c[U];
      T a[U];
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int k = base + j * LPR, kk = k < re ? k : re - 1;
        c[j] = ld_stream(ci + kk);
        a[j] = k < re ? ld_stream(val + kk) : T(0);
      }
#pragma unroll
      for (int j = 0; j < U; j++) s += a[j] * (z[c[j]] + be * p_old[c[j]]);
    }
    if constexpr (sizeof(T) == 8) s = group_sum<LPR>(s); else s = group_sumf<LPR>(s);
    if (r < r1 && l == 0) {
      const T pn = z[r] + be * p_old[r];
      p_new[r] = pn;
      q[r] = s;
      v[0] += (double)pn * (double)s;
    }
  });
  double out[1];
  if (grid_sum<1>(v, partials, &sc->cnt[1], out) && threadIdx.x == 0) {
    sc->pq = out[0];
    sc->alpha = sc->rz / out[0];
  }
}

// kind 0: plain stop test. kind 1 (FP32 inner solve of the mixed solver):
// the recurrence residual reaching the target (done = 1) or dropping by
// delta^2 since the last FP64 residual (done = 5) pauses for a reliable
// update by the host.
template <class T, int KIND>
__global__ __launch_bounds__(kBlock) void k_pcg_b(int n, T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                                                  const T* p0, const T* p1, const T* __restrict__ q,
                                                  const T* __restrict__ dinv, double delta2, double* partials,
                                                  SolveScalars* sc) {
  if (sc->done) return;
  const T* __restrict__ p = (sc->iters & 1) ? p0 : p1;   // written by this iteration's k_pcg_a
  long long r0, r1;
  block_rows(n, xcd_block(), gridDim.x, kBlock, &r0, &r1);
  const T a = (T)sc->alpha;
  double v[2] = {0.0, 0.0};
  for (long long i = r0 + threadIdx.x; i < r1; i += kBlock) {
    x[i] += a * p[i];
    const T ri = r[i] - a * q[i];
    r[i] = ri;
    const T zi = dinv[i] * ri;
    z[i] = zi;
    v[0] += (double)ri * (double)ri;
    v[1] += (double)ri * (double)zi;
  }
  double out[2];
  if (grid_sum<2>(v, partials, &sc->cnt[2], out) && threadIdx.x == 0) {
    sc->iters += 1;
    sc->rr = out[0];
    const bool bad = !(out[0] == out[0]) || !(out[1] == out[1]) || !(sc->alpha == sc->alpha);
    if (!bad) {
      sc->beta = out[1] / sc->rz;
      sc->rz = out[1];
    }
    if (bad) sc->done = 3;
    else if (out[0] <= sc->rtol2bb) sc->done = 1;
    else if (KIND == 1 && out[0] <= delta2 * sc->rr_anchor) sc->done = 5;
    else if (sc->iters >= sc->maxit) sc->done = 2;
  }
}

template <class T>
static void pcg_launch_iteration(fsx_ctx* ctx, const T* val, const T* dinv, T* x, T* r, T* z, T* p0, T* p1, T* q,
                                 int kind, double delta2) {
  const int n = ctx->NNp;
  const int g_rows = row_grid(n, kBlock / kPcgLPR), g_vec = row_grid(n, kBlock);
  hipLaunchKernelGGL((k_pcg_a<T, kPcgLPR, kPcgU>), dim3(g_rows), dim3(kBlock), 0, ctx->stream, n, ctx->Zrow, ctx->Zcol,
                     val, z, p0, p1, q, ctx->partials, ctx->sc);
  if (kind == 0)
    hipLaunchKernelGGL((k_pcg_b<T, 0>), dim3(g_vec), dim3(kBlock), 0, ctx->stream, n, x, r, z, p0, p1, q, dinv, delta2,
                       ctx->partials, ctx->sc);
  else
    hipLaunchKernelGGL((k_pcg_b<T, 1>), dim3(g_vec), dim3(kBlock), 0, ctx->stream, n, x, r, z, p0, p1, q, dinv, delta2,
                       ctx->partials, ctx->sc);
}

// Enqueues iterations in chunks (the first sized from the previous solve)
// and polls the device scalars after each chunk until done != 0. Launches
// after a stop or a pause do nothing, and the p double buffer is chosen on
// the device from the executed iteration count, so no-op launches cannot
// put it out of step.
template <class T>
static int pcg_iterate(fsx_ctx* ctx, const T* val, const T* dinv, T* x, T* r, T* z, T* p0, T* p1, T* q, int kind,
                       double delta2, int first_chunk, int maxit) {
  int issued = 0;
  int chunk = first_chunk > 8 ? first_chunk : 8;
  for (;;) {
    for (int j = 0; j < chunk; j++) pcg_launch_iteration<T>(ctx, val, dinv, x, r, z, p0, p1, q, kind, delta2);
    issued += chunk;
    FSX_TRY(launch_check(ctx, "pcg iteration"));
    FSX_HIP(hipMemcpyAsync(ctx->h_sc, ctx->sc, sizeof(SolveScalars), hipMemcpyDeviceToHost, ctx->stream));
    FSX_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->h_sc->done) break;
    if (issued > maxit + 64) return fail(ctx, FSX_ESTATE, "PCG did not terminate");
    chunk = 8;
  }
  return FSX_OK;
}

template <class T>
static void timing_begin(fsx_ctx* ctx) {
  if (ctx->timing) (void)hipEventRecord(ctx->ev0, ctx->stream);
}
static void timing_end(fsx_ctx* ctx, int iters) {
  if (!ctx->timing) return;
  (void)hipEventRecord(ctx->ev1, ctx->stream);
  (void)hipEventSynchronize(ctx->ev1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  ctx->tm.pcg_ms += ms;
  ctx->tm.pcg_spmv_launches += iters;
  ctx->tm.pcg_iters += iters;
}

static int solve_fp64(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres) {
  const int n = ctx->NNp;
  timing_begin<double>(ctx);
  hipLaunchKernelGGL(k_pcg_init<double>, dim3(row_grid(n, kBlock)), dim3(kBlock), 0, ctx->stream, n, b, ctx->Zdinv, x,
                     ctx->wr, ctx->wz, ctx->wp, ctx->rtol, ctx->maxit, ctx->partials, ctx->sc);
  FSX_TRY(launch_check(ctx, "k_pcg_init"));
  FSX_TRY(pcg_iterate<double>(ctx, ctx->Z, ctx->Zdinv, x, ctx->wr, ctx->wz, ctx->wp, ctx->wq2, ctx->wq, 0, 0.0,
                              (int)(0.85 * ctx->last_iters), ctx->maxit));
  const SolveScalars& s = *ctx->h_sc;
  timing_end(ctx, s.iters);
  *iters = s.iters;
  *relres = s.bb > 0 ? std::sqrt(s.rr / s.bb) : 0.0;
  ctx->last_iters = s.iters;
  if (s.done == 3) return fail(ctx, FSX_ENUMERIC, "PCG met NaN/Inf");
  if (s.done == 2) { ctx->err = "PCG reached maxit"; return FSX_ENOCONV; }
  return FSX_OK;
}

// ------------------------------------------------------------ mixed precision
// FP32 Jacobi-PCG with reliable updates: the CG recurrence (Z, x, r, z, p, q)
// runs in FP32; whenever its residual has dropped by delta since the last
// anchor, or reaches the target, the FP32 iterate is folded into the FP64
// solution x_hi, the residual is recomputed in FP64 as b - Z x_hi (the
// refinement residual), the FP64 stop test ||b - Z x_hi|| <= rtol ||b|| is
// applied, and the FP32 residual is replaced by the rounded FP64 one. The
// search direction is kept, so the Krylov information is not lost as in a
// restarted refinement.
__global__ __launch_bounds__(kBlock) void k_mx_fold(int n, double* __restrict__ xh, float* __restrict__ xl) {
  long long r0, r1;
  block_rows(n, xcd_block(), gridDim.x, kBlock, &r0, &r1);
  for (long long i = r0 + threadIdx.x; i < r1; i += kBlock) {
    xh[i] += (double)xl[i];
    xl[i] = 0.0f;
  }
}

// rh = b - Z xh (FP64), ||rh||^2 -> final stop test or new anchor
template <int LPR, int U>
__global__ __launch_bounds__(kBlock) void k_mx_residual(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                        const double* __restrict__ val, const double* __restrict__ xh,
                                                        const double* __restrict__ b, double* __restrict__ rh,
                                                        double* partials, SolveScalars* sc) {
  const int l = threadIdx.x % LPR;
  double v[1] = {0.0};
  for_rows<LPR>(n, rp, [&](long long i, long long r1, int rs, int re) {
    double s = 0;
    for (int base = rs + l; base < re; base += U * LPR) {
      int c[U];
      double a[U];
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int k = base + j * LPR, kk = k < re ? k : re - 1;
        c[j] = ld_stream(ci + kk);
        a[j] = k < re ? ld_stream(val + kk) : 0.0;
      }
#pragma unroll
      for (int j = 0; j < U; j++) s += a[j] * xh[c[j]];
    }
    s = group_sum<LPR>(s);
    if (i < r1 && l == 0) {
      const double ri = b[i] - s;
      rh[i] = ri;
      v[0] += ri * ri;
    }
  });
  double out[1];
  if (grid_sum<1>(v, partials, &sc->cnt[3], out) && threadIdx.x == 0) {
    sc->rr_hi = out[0];
    if (!(out[0] == out[0])) sc->done = 3;
    else if (out[0] <= sc->rtol2bb) sc->done = 1;
    else { sc->done = 0; sc->rr_anchor = out[0]; }
  }
}

// r = float(rh), z = dinv r, rz recomputed
__global__ __launch_bounds__(kBlock) void k_mx_replace(int n, const double* __restrict__ rh, const float* __restrict__ dinv,
                                                       float* __restrict__ r, float* __restrict__ z, double* partials,
                                                       SolveScalars* sc) {
  long long r0, r1;
  block_rows(n, xcd_block(), gridDim.x, kBlock, &r0, &r1);
  double v[1] = {0.0};
  for (long long i = r0 + threadIdx.x; i < r1; i += kBlock) {
    const float ri = (float)rh[i];
    const float zi = dinv[i] * ri;
    r[i] = ri; z[i] = zi;
    v[0] += (double)ri * (double)zi;
  }
  double out[1];
  if (grid_sum<1>(v, partials, &sc->cnt[0], out) && threadIdx.x == 0) sc->rz = out[0];
}

static int solve_mixed(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres) {
  const int n = ctx->NNp;
  const double delta = 1e-2;
  timing_begin<float>(ctx);
  FSX_TRY(launch_fill(ctx, x, 0.0, n));
  hipLaunchKernelGGL(k_pcg_init<float>, dim3(row_grid(n, kBlock)), dim3(kBlock), 0, ctx->stream, n, b, ctx->Zdinvf, ctx->fx,
                     ctx->fr, ctx->fz, ctx->fp, ctx->rtol, ctx->maxit, ctx->partials, ctx->sc);
  FSX_TRY(launch_check(ctx, "k_pcg_init<float>"));
  int guess = (int)(0.5 * ctx->last_iters);
  int status = 0, updates = 0;
  for (;;) {
    FSX_TRY(pcg_iterate<float>(ctx, ctx->Zf, ctx->Zdinvf, ctx->fx, ctx->fr, ctx->fz, ctx->fp, ctx->fp2, ctx->fq, 1,
                               delta * delta, guess, ctx->maxit));
    const int done = ctx->h_sc->done;
    if (done == 3) return fail(ctx, FSX_ENUMERIC, "mixed PCG met NaN/Inf");
    // fold the FP32 iterate into x and take the FP64 residual
    hipLaunchKernelGGL(k_mx_fold, dim3(row_grid(n, kBlock)), dim3(kBlock), 0, ctx->stream, n, x, ctx->fx);
    hipLaunchKernelGGL((k_mx_residual<kPcgLPR, kPcgU>), dim3(row_grid(n, kBlock / kPcgLPR)), dim3(kBlock), 0, ctx->stream,
                       n, ctx->Zrow, ctx->Zcol, ctx->Z, x, b, ctx->wr, ctx->partials, ctx->sc);
    FSX_TRY(launch_check(ctx, "mixed reliable update"));
    FSX_HIP(hipMemcpyAsync(ctx->h_sc, ctx->sc, sizeof(SolveScalars), hipMemcpyDeviceToHost, ctx->stream));
    FSX_HIP(hipStreamSynchronize(ctx->stream));
    updates++;
    if (ctx->h_sc->done == 3) return fail(ctx, FSX_ENUMERIC, "mixed residual met NaN/Inf");
    if (ctx->h_sc->done == 1) break;
    if (done == 2 || ctx->h_sc->iters >= ctx->maxit) { status = 2; break; }
    hipLaunchKernelGGL(k_mx_replace, dim3(row_grid(n, kBlock)), dim3(kBlock), 0, ctx->stream, n, ctx->wr, ctx->Zdinvf,
                       ctx->fr, ctx->fz, ctx->partials, ctx->sc);
    FSX_TRY(launch_check(ctx, "k_mx_replace"));
    guess = 8;
  }
  const SolveScalars& s = *ctx->h_sc;
  timing_end(ctx, s.iters);
  *iters = s.iters;
  *relres = s.bb > 0 ? std::sqrt(s.rr_hi / s.bb) : 0.0;
  ctx->last_iters = s.iters;
  ctx->last_updates = updates;
  if (status == 2) { ctx->err = "mixed-precision PCG reached maxit"; return FSX_ENOCONV; }
  return FSX_OK;
}

int pressure_solve(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres) {
  int rc = ctx->solver_mode == FSX_SOLVER_MIXED_IR ? solve_mixed(ctx, b, x, iters, relres)
                                                     : solve_fp64(ctx, b, x, iters, relres);
  ctx->last_relres = *relres;
  return rc;
}

// ------------------------------------------------------------ SpMV timing
__global__ void k_arm(SolveScalars* sc) {
  sc->done = 0;
  sc->iters = 0;
  sc->rz = 1.0;
  sc->beta = 0.0;
}

template <class T>
__global__ void k_fill_t(T* __restrict__ p, T v, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

template <class T>
static int bench_spmv_t(fsx_ctx* ctx, const T* val, T* z, T* pa, T* pb, T* q, int reps, double* ms) {
  const int n = ctx->NNp;
  hipLaunchKernelGGL(k_fill_t<T>, dim3(1024), dim3(kBlock), 0, ctx->stream, z, (T)1, n);
  hipLaunchKernelGGL(k_fill_t<T>, dim3(1024), dim3(kBlock), 0, ctx->stream, pa, (T)0, n);
  hipLaunchKernelGGL(k_arm, dim3(1), dim3(1), 0, ctx->stream, ctx->sc);
  const int g_rows = row_grid(n, kBlock / kPcgLPR);
  for (int w = 0; w < 3; w++)
    hipLaunchKernelGGL((k_pcg_a<T, kPcgLPR, kPcgU>), dim3(g_rows), dim3(kBlock), 0, ctx->stream, n, ctx->Zrow, ctx->Zcol,
                       val, z, pa, pb, q, ctx->partials, ctx->sc);
  FSX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL((k_pcg_a<T, kPcgLPR, kPcgU>), dim3(g_rows), dim3(kBlock), 0, ctx->stream, n, ctx->Zrow, ctx->Zcol,
                       val, z, pa, pb, q, ctx->partials, ctx->sc);
  FSX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  FSX_HIP(hipEventSynchronize(ctx->ev1));
  FSX_TRY(launch_check(ctx, "k_pcg_a (timing)"));
  float t = 0;
  FSX_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
  *ms = t / reps;
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
  // algorithmic bytes: values + columns + row pointers, z and p_old read
  // once, p_new and q written once
  if (precision == 0) {
    FSX_TRY(bench_spmv_t<double>(ctx, ctx->Z, ctx->wz, ctx->wp, ctx->wq2, ctx->wq, reps, &ms));
    bytes = 12.0 * nnz + 4.0 * (n + 1) + 32.0 * n;
  } else {
    FSX_TRY(bench_spmv_t<float>(ctx, ctx->Zf, ctx->fz, ctx->fp, ctx->fp2, ctx->fq, reps, &ms));
    bytes = 8.0 * nnz + 4.0 * (n + 1) + 16.0 * n;
  }
  *ms_per_launch = ms;
  if (bytes_per_launch) *bytes_per_launch = bytes;
  return FSX_OK;
}
