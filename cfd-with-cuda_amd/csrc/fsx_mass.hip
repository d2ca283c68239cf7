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
// M is applied element by element (FSX_MASS_CONSISTENT): per element the
// products y_e = M_e p_e = sum_k Sv_k f_k (Sv_k^T p_e) with f_k = detJ_k w_k,
// the quadrature k_assemble sums M with, then per node the sum of its
// elements' y_e rows in a fixed element order (deterministic, no atomics).
// That reads NE x (27 ids + NGP factors + 81 products) instead of M's
// values and columns: on the 16x16x160 bent duct 0.1 GB against 0.26 GB per
// application. FSX_MASS_CONSISTENT_CSR applies the assembled CSR values.
//
// One iteration is four kernels (the two of the element form with p.q
// fused into the second, or the CSR SpMV and p.q; update; direction) with
// the scalars on the device, like the pressure PCG; the host polls the stop
// flag after every kMassChunk iterations.
//
// Mixed precision (fsx_set_mass_precision 32, element form): the same
// iteration in FP32 (vectors, element products, factors and shape table) on
// the correction M_bc d = r, in segments: a segment starts from the FP64
// residual r = b - M_bc x, runs until every component's FP32 residual has
// fallen by 1e-4 (or below the FP64 target), then x += d and the residual is
// formed again in FP64 (the FP64 element form), whose per-component test
// ||r_c|| <= rtol ||b_c|| decides convergence, as in the FP64 solve.
#include "fsx_device.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace fsx {

struct MassScalars {
  double rz[2][3];      // r.z per component, by iteration parity
  double bb[3], rr[3];  // |b_c|^2, |r_c|^2
  double rtol2;
  double rr0[3];        // |r_c|^2 at a mixed segment's start
  double delta2;        // a mixed segment ends at |r_c|^2 <= delta2 rr0_c (0: FP64 solve)
  int iters, maxit, stop;   // stop: 1 converged (or the segment done), 2 maxit, 3 NaN
};

constexpr double kMassSegment = 1e-6;   // residual reduction of a mixed segment
constexpr int kMassSegs = 4;            // segments whose lengths the next solve reuses

constexpr int kMassChunk = 8;

// Jacobi inverse of M's diagonal, 0 at BC nodes
__global__ void k_mass_dinv(int NN, const int* __restrict__ Mrow, const int* __restrict__ Mcol, const double* __restrict__ M,
                            const unsigned char* __restrict__ bc, double* __restrict__ dinv) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < NN; r += gridDim.x * blockDim.x) {
    const int k = find_col(Mcol, Mrow[r], Mrow[r + 1], r);
    dinv[r] = bc[r] ? 0.0 : 1.0 / M[k];
  }
}

// Warm start, first pass: with q = M x0, the partials of |b_c|^2 (slots
// 16-18) and |b_c - q_c|^2 (19-21)
__global__ __launch_bounds__(kBlock) void k_mass_warm_a(int NN, const double* __restrict__ b, const double* __restrict__ q,
                                                        double* __restrict__ partials) {
  double v[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < NN; i += kRedGrid * kBlock)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const size_t k = 3 * (size_t)i + c;
      const double bi = b[k], ri = bi - q[k];
      v[c] += bi * bi;
      v[3 + c] += ri * ri;
    }
  write_partials<6>(v, partials + 16 * kRedGrid);
}

// Second pass: component c keeps x0 when |b_c - M x0_c| < |b_c|, else starts
// from 0; r = b - M x, z = dinv r, p = z; partials of r.z (0-2), b.b (3-5)
// and r.r (6-8)
__global__ __launch_bounds__(kBlock) void k_mass_warm_b(int NN, const double* __restrict__ b, const double* __restrict__ q,
                                                        const double* __restrict__ dinv, double* __restrict__ x,
                                                        double* __restrict__ r, double* __restrict__ z,
                                                        double* __restrict__ p, double* __restrict__ partials) {
  double s[6];
  read_partials<6>(partials + 16 * kRedGrid, s);
  bool warm[3];
#pragma unroll
  for (int c = 0; c < 3; c++) warm[c] = s[3 + c] < s[c];
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll 2   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < NN; i += kRedGrid * kBlock) {
    const double di = dinv[i];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const size_t k = 3 * (size_t)i + c;
      const double bi = b[k];
      const double ri = warm[c] ? bi - q[k] : bi, zi = di * ri;
      if (!warm[c]) x[k] = 0.0;
      r[k] = ri; z[k] = zi; p[k] = zi;
      v[c] += ri * zi;
      v[3 + c] += bi * bi;
      v[6 + c] += ri * ri;
    }
  }
  write_partials<9>(v, partials);
}

__global__ __launch_bounds__(kBlock) void k_mass_warm_finish(const double* __restrict__ partials, MassScalars* sc,
                                                             double rtol, int maxit) {
  double s[9];
  read_partials<9>(partials, s);
  if (threadIdx.x == 0) {
    bool done = true;
    for (int c = 0; c < 3; c++) {
      sc->rz[0][c] = s[c]; sc->rz[1][c] = 0.0;
      sc->bb[c] = s[3 + c]; sc->rr[c] = s[6 + c];
      done = done && s[6 + c] <= rtol * rtol * s[3 + c];
    }
    sc->rtol2 = rtol * rtol;
    sc->delta2 = 0.0;
    for (int c = 0; c < 3; c++) sc->rr0[c] = 0.0;
    sc->iters = 0; sc->maxit = maxit;
    sc->stop = done ? 1 : 0;
  }
}

// q = M p for the three components (rows = nodes), q = 0 at BC rows
template <int LPR, int U, bool KPAT>
__global__ __launch_bounds__(kBlock) void k_mass_spmv(int NN, const int* __restrict__ Mrow, ColSrc mcs,
                                                      const double* __restrict__ M, const unsigned char* __restrict__ bc,
                                                      const double* __restrict__ p, double* __restrict__ q,
                                                      const MassScalars* __restrict__ sc) {
  if (sc && sc->stop) return;
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

// The element form, first pass: y_e = M_e p_e, one thread per element and
// velocity component. The thread holds its element's 27 values of p in
// registers; for each Gauss point k it takes row k of the shape table (the
// same for every thread: scalar loads), u = sum_j Sv[k][j] p_j, then adds
// Sv[k][i] f_k u to y_i for the 27 nodes i. The element arrays are
// transposed so that consecutive threads read consecutive words: ids
// L[j * NE + e], factors fac[k * NE + e], products ye[(i * NE + e) * 3 + c].
// (T = float: the mixed solve's FP32 form, the same operations in FP32)
template <class T>
__global__ __launch_bounds__(kBlock) void k_mass_elem(int NE, int NGP, const int* __restrict__ L,
                                                      const T* __restrict__ fac, const T* __restrict__ Sv,
                                                      const T* __restrict__ p, T* __restrict__ ye,
                                                      const MassScalars* __restrict__ sc) {
  if (sc && sc->stop) return;
  const long long t = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (t >= 3LL * NE) return;
  const int e = (int)(t / 3), c = (int)(t % 3);
  T pe[27], y[27];
#pragma unroll
  for (int j = 0; j < 27; j++) {
    pe[j] = p[3 * (size_t)L[(size_t)j * NE + e] + c];
    y[j] = T(0);
  }
  for (int k = 0; k < NGP; k++) {
    const T* row = Sv + k * 27;
    T u = T(0);
#pragma unroll
    for (int j = 0; j < 27; j++) u += row[j] * pe[j];
    const T v = fac[(size_t)k * NE + e] * u;
#pragma unroll
    for (int i = 0; i < 27; i++) y[i] += row[i] * v;
  }
#pragma unroll
  for (int i = 0; i < 27; i++) ye[((size_t)i * NE + e) * 3 + c] = y[i];
}

// Second pass: q_r = sum of y_e rows of the elements holding node r (0 at
// BC rows); with DOT the per-block p.q partials per component at
// partials[kGatherPq + c kGatherGrid + block] (the grid is kGatherGrid blocks)
constexpr int kGatherGrid = 2048;
constexpr int kGatherPq = 32 * kRedGrid;

template <bool DOT, class T>
__global__ __launch_bounds__(kBlock) void k_mass_gather(int NN, const int* __restrict__ inc_ptr, const int* __restrict__ inc,
                                                        const T* __restrict__ ye, const unsigned char* __restrict__ bc,
                                                        const T* __restrict__ p, T* __restrict__ q,
                                                        double* __restrict__ partials, const MassScalars* __restrict__ sc) {
  if (sc && sc->stop) return;
  double v[3] = {0.0, 0.0, 0.0};
#pragma unroll 2   // independent rows: their loads in flight together
  for (int r = blockIdx.x * kBlock + threadIdx.x; r < NN; r += gridDim.x * kBlock) {
    T s0 = T(0), s1 = T(0), s2 = T(0);
    if (!bc[r]) {
      const int t1 = inc_ptr[r + 1];
      for (int t = inc_ptr[r]; t < t1; t++) {
        const T* y = ye + 3 * (size_t)inc[t];
        s0 += y[0]; s1 += y[1]; s2 += y[2];
      }
    }
    const size_t k = 3 * (size_t)r;
    q[k] = s0; q[k + 1] = s1; q[k + 2] = s2;
    if (DOT) {
      v[0] += (double)p[k] * (double)s0; v[1] += (double)p[k + 1] * (double)s1; v[2] += (double)p[k + 2] * (double)s2;
    }
  }
  if (DOT) {
    __shared__ double red[3 * (kBlock / kWave)];
    block_sum_all<3>(v, red);
    if (threadIdx.x == 0)
#pragma unroll
      for (int c = 0; c < 3; c++) partials[kGatherPq + c * kGatherGrid + blockIdx.x] = v[c];
  }
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

// alpha_c = r.z_c / p.q_c (0 for a component with no work left), p.q_c the
// sum of npq per-block partials at pqp + c npq;
// x += alpha p, r -= alpha q, z = dinv r; partials of r.r (slots 0-2) and r.z (3-5)
template <class T>
__global__ __launch_bounds__(kBlock) void k_mass_update(int j, int NN, T* __restrict__ x, T* __restrict__ r,
                                                        T* __restrict__ z, const T* __restrict__ p,
                                                        const T* __restrict__ q, const T* __restrict__ dinv,
                                                        const double* __restrict__ pqp, int npq,
                                                        double* __restrict__ partials, const MassScalars* __restrict__ sc) {
  if (sc->stop) return;
  double pq[3];
#pragma unroll
  for (int c = 0; c < 3; c++) pq[c] = sum_parts(pqp + c * npq, npq);
  T a[3];
#pragma unroll
  for (int c = 0; c < 3; c++) a[c] = (T)(pq[c] > 0.0 ? sc->rz[j & 1][c] / pq[c] : 0.0);
  double v[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < NN; i += kRedGrid * kBlock) {
    const T di = dinv[i];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const size_t k = 3 * (size_t)i + c;
      x[k] += a[c] * p[k];
      const T ri = r[k] - a[c] * q[k];
      const T zi = di * ri;
      r[k] = ri; z[k] = zi;
      v[c] += (double)ri * (double)ri;
      v[3 + c] += (double)ri * (double)zi;
    }
  }
  write_partials<6>(v, partials);
}

// beta_c = r.z_c(new) / r.z_c(old); p = z + beta p; block 0 records the scalars
// and the stop test: every component at ||r_c|| <= rtol ||b_c||
template <class T>
__global__ __launch_bounds__(kBlock) void k_mass_direction(int j, int NN, const T* __restrict__ z,
                                                           T* __restrict__ p, const double* __restrict__ partials,
                                                           MassScalars* sc) {
  if (sc->stop) return;
  double s[6];
  read_partials<6>(partials, s);
  T be[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const double old = sc->rz[j & 1][c];
    be[c] = (T)(old != 0.0 ? s[3 + c] / old : 0.0);
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
      done = done && s[c] <= fmax(sc->rtol2 * sc->bb[c], sc->delta2 * sc->rr0[c]);
    }
    const int it = sc->iters + 1;
    sc->iters = it;
    if (bad) sc->stop = 3;
    else if (done) sc->stop = 1;
    else if (it >= sc->maxit) sc->stop = 2;
  }
}

// ---- mixed precision: FP32 segments on the correction d, FP64 residuals

// A segment's start from the FP64 residual r64: r = float(r64), d = 0,
// z = dinv r, p = z; partials of r.z (slots 0-2) and r.r (3-5)
__global__ __launch_bounds__(kBlock) void k_mass_seg_start(int NN, const double* __restrict__ r64,
                                                           const float* __restrict__ dinv, float* __restrict__ r,
                                                           float* __restrict__ z, float* __restrict__ p,
                                                           float* __restrict__ d, double* __restrict__ partials) {
  double v[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 4
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < NN; i += kRedGrid * kBlock) {
    const float di = dinv[i];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const size_t k = 3 * (size_t)i + c;
      const float ri = (float)r64[k], zi = di * ri;
      r[k] = ri; z[k] = zi; p[k] = zi; d[k] = 0.0f;
      v[c] += (double)ri * (double)zi;
      v[3 + c] += (double)ri * (double)ri;
    }
  }
  write_partials<6>(v, partials);
}

// the segment's scalars: r.z for iteration j's parity, the start residual
// rr0 and the segment's reduction target; stop at once if every component
// already meets the FP64 target
__global__ __launch_bounds__(kBlock) void k_mass_seg_finish(int j, const double* __restrict__ partials, MassScalars* sc,
                                                            double delta) {
  double s[6];
  read_partials<6>(partials, s);
  if (threadIdx.x == 0) {
    bool done = true;
    for (int c = 0; c < 3; c++) {
      sc->rz[j & 1][c] = s[c];
      sc->rr0[c] = s[3 + c];
      sc->rr[c] = s[3 + c];
      done = done && s[3 + c] <= sc->rtol2 * sc->bb[c];
    }
    sc->delta2 = delta * delta;
    sc->stop = done ? 1 : 0;
  }
}

// b = float(a), rounded to nearest
__global__ void k_mass_round(long long n, const double* __restrict__ a, float* __restrict__ b) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// x += d (the FP64 iterate takes the segment's correction)
__global__ void k_mass_fold(long long n, const float* __restrict__ d, double* __restrict__ x) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] += (double)d[i];
}

// r64 = b - q (q = M_bc x, zero at BC rows, where b is zero too); partials
// of |r_c|^2 (slots 0-2)
__global__ __launch_bounds__(kBlock) void k_mass_res64(int NN, const double* __restrict__ b, const double* __restrict__ q,
                                                       double* __restrict__ r, double* __restrict__ partials) {
  double v[3] = {0, 0, 0};
#pragma unroll 4
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < NN; i += kRedGrid * kBlock)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const size_t k = 3 * (size_t)i + c;
      const double ri = b[k] - q[k];
      r[k] = ri;
      v[c] += ri * ri;
    }
  write_partials<3>(v, partials);
}

// the FP64 stop test on the formed residual: stop 1 when every component
// meets ||r_c|| <= rtol ||b_c||, else 0 (a new segment follows), 3 on NaN,
// 2 at the iteration cap
__global__ __launch_bounds__(kBlock) void k_mass_res_finish(const double* __restrict__ partials, MassScalars* sc) {
  double s[3];
  read_partials<3>(partials, s);
  if (threadIdx.x == 0) {
    bool done = true, bad = false;
    for (int c = 0; c < 3; c++) {
      sc->rr[c] = s[c];
      bad = bad || !(s[c] == s[c]);
      done = done && s[c] <= sc->rtol2 * sc->bb[c];
    }
    sc->stop = bad ? 3 : (done ? 1 : (sc->iters >= sc->maxit ? 2 : 0));
  }
}

// y = u + dt x (node-major 3NN)
__global__ void k_mass_axpy(long long n, const double* __restrict__ u, const double* __restrict__ x, double dt,
                            double* __restrict__ y) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = u[i] + dt * x[i];
}

// The element forms' tables (host-built once): the elements sorted by
// their centre node's internal id, their velocity ids transposed, their
// corners' pressure ids, and the node -> (sorted element, local node)
// incidence, ascending per node.
int elements_setup(fsx_ctx* ctx) {
  const int NE = ctx->NE, NN = ctx->NN;
  for (void* q : {(void*)ctx->el_L27, (void*)ctx->el_P8, (void*)ctx->el_inc_ptr, (void*)ctx->el_inc})
    dev_free(ctx, q);   // what an earlier, failed set-up left
  ctx->el_L27 = nullptr; ctx->el_P8 = nullptr; ctx->el_inc_ptr = nullptr; ctx->el_inc = nullptr;
  std::vector<int> L((size_t)NE * 27), P((size_t)NE * 8);
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  FSX_HIP(hipMemcpy(L.data(), ctx->L27, L.size() * sizeof(int), hipMemcpyDeviceToHost));
  FSX_HIP(hipMemcpy(P.data(), ctx->P8, P.size() * sizeof(int), hipMemcpyDeviceToHost));
  std::vector<int> order(NE);
  for (int e = 0; e < NE; e++) order[e] = e;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return L[(size_t)a * 27 + 26] < L[(size_t)b * 27 + 26]; });
  std::vector<int> Ls((size_t)NE * 27), Ps((size_t)NE * 8);   // Ls[i * NE + k]: node i of sorted element k
  for (int k = 0; k < NE; k++) {
    for (int i = 0; i < 27; i++) Ls[(size_t)i * NE + k] = L[(size_t)order[k] * 27 + i];
    std::copy_n(P.begin() + (size_t)order[k] * 8, 8, Ps.begin() + (size_t)k * 8);
  }
  std::vector<int> ptr((size_t)NN + 1, 0), inc((size_t)NE * 27);
  for (int v : Ls) ptr[(size_t)v + 1]++;
  for (int r = 0; r < NN; r++) ptr[r + 1] += ptr[r];
  std::vector<int> next(ptr.begin(), ptr.end() - 1);
  for (int k = 0; k < NE; k++)   // entries i * NE + k, ascending k per node
    for (int i = 0; i < 27; i++) inc[next[Ls[(size_t)i * NE + k]]++] = i * NE + k;
  FSX_TRY(alloc(ctx, &ctx->el_L27, Ls.size()));
  FSX_TRY(alloc(ctx, &ctx->el_P8, Ps.size()));
  FSX_TRY(alloc(ctx, &ctx->el_inc_ptr, ptr.size()));
  FSX_TRY(alloc(ctx, &ctx->el_inc, inc.size()));
  FSX_HIP(hipMemcpy(ctx->el_L27, Ls.data(), Ls.size() * sizeof(int), hipMemcpyHostToDevice));
  FSX_HIP(hipMemcpy(ctx->el_P8, Ps.data(), Ps.size() * sizeof(int), hipMemcpyHostToDevice));
  FSX_HIP(hipMemcpy(ctx->el_inc_ptr, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice));
  FSX_HIP(hipMemcpy(ctx->el_inc, inc.data(), inc.size() * sizeof(int), hipMemcpyHostToDevice));
  ctx->el_ready = true;
  ctx->ek_xl_ready = false;   // the element kernels' corner table follows the new order
  return FSX_OK;
}

// the consistent mass's factors and products on top of the element tables
static int mass_elements_setup(fsx_ctx* ctx) {
  if (!ctx->el_ready) FSX_TRY(elements_setup(ctx));
  if (!ctx->mass_fac) FSX_TRY(alloc(ctx, &ctx->mass_fac, (size_t)ctx->NE * ctx->NGP));
  if (!ctx->mass_ye) FSX_TRY(alloc(ctx, &ctx->mass_ye, (size_t)ctx->NE * 81));
  return launch_mass_fac(ctx, ctx->el_P8, ctx->mass_fac);
}

// q = M_bc p; with dot the p.q partials (slots 6-8). sc null: always runs.
static void mass_apply(fsx_ctx* ctx, bool elements, const double* p, double* q, bool dot, const MassScalars* sc) {
  const int NN = ctx->NN;
  if (elements) {
    hipLaunchKernelGGL(k_mass_elem<double>, dim3((unsigned)((3LL * ctx->NE + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       ctx->stream, ctx->NE, ctx->NGP, ctx->el_L27, ctx->mass_fac, ctx->tab.Sv, p, ctx->mass_ye, sc);
    if (dot)
      hipLaunchKernelGGL((k_mass_gather<true, double>), dim3(kGatherGrid), dim3(kBlock), 0, ctx->stream, NN,
                         ctx->el_inc_ptr, ctx->el_inc, ctx->mass_ye, ctx->bcflag, p, q, ctx->partials, sc);
    else
      hipLaunchKernelGGL((k_mass_gather<false, double>), dim3((unsigned)std::min((NN + kBlock - 1) / kBlock, 4096)),
                         dim3(kBlock), 0, ctx->stream, NN, ctx->el_inc_ptr, ctx->el_inc, ctx->mass_ye, ctx->bcflag, p, q,
                         ctx->partials, sc);
    return;
  }
  constexpr int LPR = 16, U = 4;
  const ColSrc mcs{ctx->Mcol, ctx->Mrowpat, ctx->Mpatoff};
  const unsigned grid = (unsigned)std::max(1, (NN + kBlock / LPR - 1) / (kBlock / LPR));
  if (mcs.row_pat)
    hipLaunchKernelGGL((k_mass_spmv<LPR, U, true>), dim3(grid), dim3(kBlock), 0, ctx->stream, NN, ctx->Mrow, mcs,
                       ctx->M, ctx->bcflag, p, q, sc);
  else
    hipLaunchKernelGGL((k_mass_spmv<LPR, U, false>), dim3(grid), dim3(kBlock), 0, ctx->stream, NN, ctx->Mrow, mcs,
                       ctx->M, ctx->bcflag, p, q, sc);
  if (dot) hipLaunchKernelGGL(k_mass_dot, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, NN, p, q, ctx->partials, sc);
}

// The statistics and return code of a finished solve from its scalars h
static int mass_finish(fsx_ctx* ctx, const MassScalars& h) {
  double rel = 0.0;
  for (int c = 0; c < 3; c++) rel = std::max(rel, h.bb[c] > 0.0 ? std::sqrt(h.rr[c] / h.bb[c]) : 0.0);
  ctx->mass_last_iters = h.iters;
  ctx->mass_last_relres = rel;
  ctx->mass_total_iters += h.iters;
  if (h.stop == 3) return fail(ctx, FSX_ENUMERIC, "mass PCG met NaN/Inf");
  if (h.stop == 2) { ctx->err = "mass PCG reached maxit"; return FSX_ENOCONV; }
  return FSX_OK;
}

// The mixed form after the FP64 warm start (x, the FP64 residual r64 and the
// scalars set): FP32 segments on M_bc d = r64, each folded into x and
// checked on the FP64 residual of the FP64 element form.
static int mass_solve_mixed(fsx_ctx* ctx, const double* b, double* x, double* r64, double* q64, MassScalars* sc) {
  const int NN = ctx->NN, NE = ctx->NE;
  const size_t n3 = 3 * (size_t)NN;
  if (!ctx->mass_w32) {
    FSX_TRY(alloc(ctx, &ctx->mass_w32, 5 * n3));
    FSX_TRY(alloc(ctx, &ctx->mass_fac32, (size_t)NE * ctx->NGP));
    FSX_TRY(alloc(ctx, &ctx->mass_ye32, (size_t)NE * 81));
    FSX_TRY(alloc(ctx, &ctx->mass_dinv32, (size_t)NN));
    ctx->mass32_ready = false;
  }
  if (!ctx->mass32_ready) {   // the FP32 factors and Jacobi inverse, rounded to nearest
    hipLaunchKernelGGL(k_mass_round, dim3(4096), dim3(kBlock), 0, ctx->stream, (long long)NE * ctx->NGP, ctx->mass_fac,
                       ctx->mass_fac32);
    hipLaunchKernelGGL(k_mass_round, dim3(4096), dim3(kBlock), 0, ctx->stream, (long long)NN, ctx->mass_dinv,
                       ctx->mass_dinv32);
    FSX_TRY(launch_check(ctx, "mass FP32 copies"));
    ctx->mass32_ready = true;
  }
  float *r = ctx->mass_w32, *z = r + n3, *p = z + n3, *q = p + n3, *d = q + n3;
  MassScalars h{};
  int j = 0;
  const unsigned eg = (unsigned)((3LL * NE + kBlock - 1) / kBlock);
  // (a warm start that already meets the target stops the first segment at
  // once: k_mass_seg_finish tests the FP64 residual)
  for (int seg = 0; seg == 0 || !h.stop; seg++) {
    hipLaunchKernelGGL(k_mass_seg_start, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, NN, r64, ctx->mass_dinv32, r, z, p,
                       d, ctx->partials);
    hipLaunchKernelGGL(k_mass_seg_finish, dim3(1), dim3(kBlock), 0, ctx->stream, j, ctx->partials, sc, kMassSegment);
    // the first chunk as long as the same segment of the last solve, then
    // kMassChunk iterations between polls (iterations past the stop are no-ops)
    const int it0 = h.iters;
    int chunk = std::max(kMassChunk, seg < kMassSegs ? ctx->mass_seg_iters[seg] : 0);
    for (;;) {
      for (int t = 0; t < chunk; t++, j++) {
        hipLaunchKernelGGL(k_mass_elem<float>, dim3(eg), dim3(kBlock), 0, ctx->stream, NE, ctx->NGP, ctx->el_L27,
                           ctx->mass_fac32, ctx->tabf.Sv, p, ctx->mass_ye32, sc);
        hipLaunchKernelGGL((k_mass_gather<true, float>), dim3(kGatherGrid), dim3(kBlock), 0, ctx->stream, NN,
                           ctx->el_inc_ptr, ctx->el_inc, ctx->mass_ye32, ctx->bcflag, p, q, ctx->partials, sc);
        hipLaunchKernelGGL(k_mass_update<float>, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, j, NN, d, r, z, p, q,
                           ctx->mass_dinv32, ctx->partials + kGatherPq, kGatherGrid, ctx->partials, sc);
        hipLaunchKernelGGL(k_mass_direction<float>, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, j, NN, z, p,
                           ctx->partials, sc);
      }
      FSX_TRY(launch_check(ctx, "mixed mass PCG"));
      FSX_HIP(hipMemcpyAsync(&h, sc, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
      FSX_HIP(hipStreamSynchronize(ctx->stream));
      if (h.stop) break;
      if (j > 2 * ctx->mass_maxit + 64) return fail(ctx, FSX_ESTATE, "mixed mass PCG did not terminate");
      chunk = kMassChunk;
    }
    if (seg < kMassSegs) ctx->mass_seg_iters[seg] = h.iters - it0;
    if (h.stop == 3) break;
    // x += d, then the FP64 residual and stop test
    hipLaunchKernelGGL(k_mass_fold, dim3((unsigned)std::min<long long>(((long long)n3 + kBlock - 1) / kBlock, 4096)),
                       dim3(kBlock), 0, ctx->stream, (long long)n3, d, x);
    mass_apply(ctx, true, x, q64, false, nullptr);
    hipLaunchKernelGGL(k_mass_res64, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, NN, b, q64, r64, ctx->partials);
    hipLaunchKernelGGL(k_mass_res_finish, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->partials, sc);
    FSX_TRY(launch_check(ctx, "mixed mass residual"));
    FSX_HIP(hipMemcpyAsync(&h, sc, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    FSX_HIP(hipStreamSynchronize(ctx->stream));
  }
  return mass_finish(ctx, h);
}

// Solves M_bc x = b (3NN, node-major) to the context's mass tolerance; x
// holds the previous solution of the same update (the warm start).
int mass_solve(fsx_ctx* ctx, const double* b, double* x) {
  const int NN = ctx->NN;
  if (!ctx->mass_dinv_ready) {
    hipLaunchKernelGGL(k_mass_dinv, dim3(std::min((NN + kBlock - 1) / kBlock, 4096)), dim3(kBlock), 0, ctx->stream, NN,
                       ctx->Mrow, ctx->Mcol, ctx->M, ctx->bcflag, ctx->mass_dinv);
    FSX_TRY(launch_check(ctx, "k_mass_dinv"));
    ctx->mass_dinv_ready = true;
    ctx->mass32_ready = false;   // the FP32 copy follows
  }
  const bool elements = ctx->mass_mode == FSX_MASS_CONSISTENT && !ctx->mass_m_uploaded;
  if (elements && !ctx->mass_fac_ready) FSX_TRY(mass_elements_setup(ctx));
  if (elements) ctx->mass_fac_ready = true;
  MassScalars* sc = reinterpret_cast<MassScalars*>(ctx->mass_sc);
  double *r = ctx->mass_w, *z = r + 3 * (size_t)NN, *p = z + 3 * (size_t)NN, *q = p + 3 * (size_t)NN;
  mass_apply(ctx, elements, x, q, false, nullptr);
  hipLaunchKernelGGL(k_mass_warm_a, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, NN, b, q, ctx->partials);
  hipLaunchKernelGGL(k_mass_warm_b, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, NN, b, q, ctx->mass_dinv, x, r, z, p,
                     ctx->partials);
  hipLaunchKernelGGL(k_mass_warm_finish, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->partials, sc, ctx->mass_rtol,
                     ctx->mass_maxit);
  FSX_TRY(launch_check(ctx, "mass PCG start"));
  MassScalars h{};
  int j = 0;
  if (elements && ctx->mass_bits == 32) return mass_solve_mixed(ctx, b, x, r, q, sc);
  // the first chunk as long as the last solve, then kMassChunk iterations
  // between polls (iterations past the stop are no-ops)
  int chunk = std::max(kMassChunk, ctx->mass_last_iters);
  for (;;) {
    for (int t = 0; t < chunk; t++, j++) {
      mass_apply(ctx, elements, p, q, true, sc);
      hipLaunchKernelGGL(k_mass_update<double>, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, j, NN, x, r, z, p, q,
                         ctx->mass_dinv, ctx->partials + (elements ? kGatherPq : 6 * kRedGrid),
                         elements ? kGatherGrid : kRedGrid, ctx->partials, sc);
      hipLaunchKernelGGL(k_mass_direction<double>, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, j, NN, z, p,
                         ctx->partials, sc);
    }
    FSX_TRY(launch_check(ctx, "mass PCG"));
    FSX_HIP(hipMemcpyAsync(&h, sc, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    FSX_HIP(hipStreamSynchronize(ctx->stream));
    if (h.stop) break;
    if (j > 2 * ctx->mass_maxit + 64) return fail(ctx, FSX_ESTATE, "mass PCG did not terminate");
    chunk = kMassChunk;
  }
  return mass_finish(ctx, h);
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
  if (mode != FSX_MASS_LUMPED && mode != FSX_MASS_CONSISTENT && mode != FSX_MASS_CONSISTENT_CSR)
    return fail(ctx, FSX_EINVAL, "fsx_set_mass: unknown mode");
  if (!(rtol >= 0) || maxit < 1) return fail(ctx, FSX_EINVAL, "fsx_set_mass: rtol >= 0 and maxit >= 1 required");
  if (mode != FSX_MASS_LUMPED) {
    if (!ctx->have_mesh) return fail(ctx, FSX_ESTATE, "fsx_set_mass: upload the mesh first");
    const size_t n3 = 3 * (size_t)ctx->NN;
    if (!ctx->mass_w) {
      FSX_TRY(alloc(ctx, &ctx->mass_w, 4 * n3));
      FSX_TRY(alloc(ctx, &ctx->mass_x, n3));
      FSX_TRY(alloc(ctx, &ctx->mass_b, n3));
      FSX_TRY(alloc(ctx, &ctx->mass_dinv, (size_t)ctx->NN));
      FSX_TRY(alloc(ctx, &ctx->mass_sc, (mass_scalars_bytes() + 7) / 8));
      FSX_TRY(launch_fill(ctx, ctx->mass_x, 0.0, (long long)n3));   // the first warm start is 0
    }
  }
  ctx->mass_mode = mode;
  ctx->mass_rtol = rtol;
  ctx->mass_maxit = maxit;
  return FSX_OK;
}

extern "C" int fsx_set_mass_precision(fsx_ctx* ctx, int bits) {
  if (!ctx) return FSX_EINVAL;
  if (bits != 64 && bits != 32) return fsx::fail(ctx, FSX_EINVAL, "fsx_set_mass_precision: 64 or 32");
  ctx->mass_bits = bits;
  return FSX_OK;
}

extern "C" int fsx_mass_stats(fsx_ctx* ctx, int* last_iters, double* last_relres, long long* total_iters) {
  if (!ctx) return FSX_EINVAL;
  if (last_iters) *last_iters = ctx->mass_last_iters;
  if (last_relres) *last_relres = ctx->mass_last_relres;
  if (total_iters) *total_iters = ctx->mass_total_iters;
  return FSX_OK;
}
