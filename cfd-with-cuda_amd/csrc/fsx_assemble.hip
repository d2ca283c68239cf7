// fsx_assemble.hip -- element-level kernels: static assembly (step0), the
// numeric Z = sum_d G_d^T diag(MdOrigInv) G_d, and the matrix-free
// convection residual.
//
// Element kernels run one launch per colour of a race-free element colouring,
// so the scatter-add into CSR values and into R1 needs no atomics and the
// result is bitwise reproducible run to run.
#include "fsx_device.h"

#include <algorithm>
#include <cmath>

namespace fsx {

// Jacobian of the trilinear (hex8) map at one Gauss point, its cofactor
// inverse and determinant, in the arithmetic of calcJacob (BCH:2553-2590):
// J[i][j] = sum_m dSp[m][i] x_m[j], dSp[m][i] read at dSp_k[(m * 3 + i) * stride].
__device__ __forceinline__ void jacobian(const double (*ec)[3], const double* __restrict__ dSp_k, double iJ[3][3],
                                         double* det, int stride = 1) {
  double J[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double s = 0;
#pragma unroll
      for (int m = 0; m < 8; m++) s = s + dSp_k[(m * 3 + i) * stride] * ec[m][j];
      J[i][j] = s;
    }
  iJ[0][0] =   J[1][1]*J[2][2] - J[2][1]*J[1][2];
  iJ[0][1] = -(J[0][1]*J[2][2] - J[0][2]*J[2][1]);
  iJ[0][2] =   J[0][1]*J[1][2] - J[1][1]*J[0][2];
  iJ[1][0] = -(J[1][0]*J[2][2] - J[1][2]*J[2][0]);
  iJ[1][1] =   J[2][2]*J[0][0] - J[2][0]*J[0][2];
  iJ[1][2] = -(J[1][2]*J[0][0] - J[1][0]*J[0][2]);
  iJ[2][0] =   J[1][0]*J[2][1] - J[2][0]*J[1][1];
  iJ[2][1] = -(J[2][1]*J[0][0] - J[2][0]*J[0][1]);
  iJ[2][2] =   J[1][1]*J[0][0] - J[1][0]*J[0][1];
  const double d = J[0][0]*(J[1][1]*J[2][2] - J[2][1]*J[1][2]) +
                   J[0][1]*(J[1][2]*J[2][0] - J[1][0]*J[2][2]) +
                   J[0][2]*(J[1][0]*J[2][1] - J[1][1]*J[2][0]);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) iJ[i][j] = iJ[i][j] / d;
  *det = d;
}

// ------------------------------------------------------------ step0
// One wavefront per element of the current colour. Per element (BCH:3176-3231):
//   Me_ij = sum_k Sv_ki Sv_kj f_k,  Ke_ij = nu sum_k gradSv_i . gradSv_j f_k,
//   Ge_d,ij = -(1/rho) sum_k Sp_kj d_d Sv_ki f_k,  f_k = detJ_k w_k,
// scattered into the CSR values through a binary search of the sorted row.
constexpr int kMaxNGP = 27;

__global__ __launch_bounds__(kWave) void k_assemble(int ncolor_elems, const int* __restrict__ elems,
                                                    const int* __restrict__ L27, const int* __restrict__ P8,
                                                    const double* __restrict__ xc,
                                                    ShapeTables tab, int NGP, double nu, double inv_rho,
                                                    const int* __restrict__ Mrow, const int* __restrict__ Mcol,
                                                    double* __restrict__ Mv, double* __restrict__ Kv,
                                                    const int* __restrict__ Grow, const int* __restrict__ Gcol,
                                                    double* __restrict__ Gv) {
  __shared__ double gd[kMaxNGP * 81];   // [k][j][d]
  __shared__ double fac[kMaxNGP];
  __shared__ double ec[8][3];
  __shared__ int L[27], P[8];   // velocity ids (internal), pressure ids of the corners
  const int t = threadIdx.x;
  const int e = elems[blockIdx.x];
  if (t < 27) L[t] = L27[(size_t)e * 27 + t];
  if (t < 8) P[t] = P8[(size_t)e * 8 + t];
  __syncthreads();
  if (t < 24) ec[t / 3][t % 3] = xc[3 * (size_t)P[t / 3] + t % 3];
  __syncthreads();
  if (t < NGP) {
    double iJ[3][3], det;
    jacobian(ec, tab.dSp + t * 24, iJ, &det);
    for (int j = 0; j < 27; j++) {
      const double* ds = tab.dSv + (t * 27 + j) * 3;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        double s = 0;
#pragma unroll
        for (int m = 0; m < 3; m++) s = s + iJ[i][m] * ds[m];
        gd[(t * 27 + j) * 3 + i] = s;
      }
    }
    fac[t] = det * tab.w[t];
  }
  __syncthreads();
  for (int idx = t; idx < 729; idx += kWave) {
    const int i = idx / 27, j = idx % 27;
    double me = 0.0, ke = 0.0;
    for (int k = 0; k < NGP; k++) {
      const double f = fac[k];
      const double* gi = gd + (k * 27 + i) * 3;
      const double* gj = gd + (k * 27 + j) * 3;
      me = me + tab.Sv[k * 27 + i] * tab.Sv[k * 27 + j] * f;
      ke = ke + nu * (gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2]) * f;
    }
    const int r = L[i];
    const int pos = find_col(Mcol, Mrow[r], Mrow[r + 1], L[j]);
    Mv[pos] += me;
    Kv[pos] += ke;
  }
  for (int idx = t; idx < 216; idx += kWave) {
    const int i = idx / 8, j = idx % 8;
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int k = 0; k < NGP; k++) {
      const double f = fac[k];
      const double* gi = gd + (k * 27 + i) * 3;
      const double sp = tab.Sp[k * 8 + j];
      g0 = g0 - inv_rho * sp * gi[0] * f;
      g1 = g1 - inv_rho * sp * gi[1] * f;
      g2 = g2 - inv_rho * sp * gi[2] * f;
    }
    const int r = L[i];
    const int pos = find_col(Gcol, Grow[r], Grow[r + 1], P[j]);
    Gv[3 * (size_t)pos] += g0;
    Gv[3 * (size_t)pos + 1] += g1;
    Gv[3 * (size_t)pos + 2] += g2;
  }
}

// detJ_k w_k of every element at every Gauss point (the f_k of k_assemble,
// in the same arithmetic), for the element-by-element mass products:
// fac[k * NE + e]
__global__ void k_mass_fac(int NE, int NGP, const int* __restrict__ P8, const double* __restrict__ xc, ShapeTables tab,
                           double* __restrict__ fac) {
  const long long n = (long long)NE * NGP;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const int e = (int)(t / NGP), k = (int)(t % NGP);
    double ec[8][3];
#pragma unroll
    for (int m = 0; m < 8; m++) {
      const int pn = P8[(size_t)e * 8 + m];
#pragma unroll
      for (int d = 0; d < 3; d++) ec[m][d] = xc[3 * (size_t)pn + d];
    }
    double iJ[3][3], det;
    jacobian(ec, tab.dSp + k * 24, iJ, &det);
    fac[(size_t)k * NE + e] = det * tab.w[k];
  }
}

// for the elements whose corner pressure ids P8 lists (NE x 8, any order)
int launch_mass_fac(fsx_ctx* ctx, const int* P8, double* fac) {
  const long long n = (long long)ctx->NE * ctx->NGP;
  const unsigned g = (unsigned)std::max(1LL, std::min((n + kBlock - 1) / kBlock, 8192LL));
  hipLaunchKernelGGL(k_mass_fac, dim3(g), dim3(kBlock), 0, ctx->stream, ctx->NE, ctx->NGP, P8, ctx->xc, ctx->tab,
                     fac);
  return launch_check(ctx, "k_mass_fac");
}

// Lumped mass: Md = row sums of M (BCH:3262-3266), MdOrigInv = 1/Md
// (BCH:3287-3289), MdInv = 1/Md with BC rows 1 (BCH:3292-3295, 4121-4135).
__global__ void k_lump(int NN, const int* __restrict__ Mrow, const double* __restrict__ Mv,
                       const unsigned char* __restrict__ bc, double* __restrict__ Md, double* __restrict__ MdOrigInv,
                       double* __restrict__ MdInv) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < NN; r += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int k = Mrow[r]; k < Mrow[r + 1]; k++) s += Mv[k];
    Md[r] = s;
    MdOrigInv[r] = 1.0 / s;
    MdInv[r] = bc[r] ? 1.0 : 1.0 / s;
  }
}

// Values of G in the transposed pattern.
__global__ void k_gt_values(long long nnz, const int* __restrict__ GTpos, const double* __restrict__ G,
                            double* __restrict__ GT) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < nnz; t += (long long)gridDim.x * blockDim.x) {
    const size_t k = GTpos[t];
    GT[3 * t] = G[3 * k]; GT[3 * t + 1] = G[3 * k + 1]; GT[3 * t + 2] = G[3 * k + 2];
  }
}

int launch_assemble_static(fsx_ctx* ctx) {
  FSX_TRY(launch_fill(ctx, ctx->M, 0.0, ctx->nnzM));
  FSX_TRY(launch_fill(ctx, ctx->K, 0.0, ctx->nnzM));
  FSX_TRY(launch_fill(ctx, ctx->G, 0.0, 3 * ctx->nnzG));
  const int ncol = (int)ctx->color_ptr.size() - 1;
  for (int c = 0; c < ncol; c++) {
    const int n = ctx->color_ptr[c + 1] - ctx->color_ptr[c];
    if (n == 0) continue;
    hipLaunchKernelGGL(k_assemble, dim3(n), dim3(kWave), 0, ctx->stream, n, ctx->color_order + ctx->color_ptr[c], ctx->L27,
                       ctx->P8, ctx->xc, ctx->tab, ctx->NGP, ctx->viscosity, 1.0 / ctx->density, ctx->Mrow, ctx->Mcol, ctx->M, ctx->K,
                       ctx->Grow, ctx->Gcol, ctx->G);
    FSX_TRY(launch_check(ctx, "k_assemble"));
  }
  ctx->k_viscosity = ctx->viscosity;   // the element form of step1 applies this K
  ctx->k_uploaded = false;
  FSX_TRY(launch_lump(ctx));
  hipLaunchKernelGGL(k_gt_values, dim3(4096), dim3(kBlock), 0, ctx->stream, ctx->nnzG, ctx->GTpos, ctx->G, ctx->GT);
  return launch_check(ctx, "k_gt_values");
}

int launch_lump(fsx_ctx* ctx) {
  const int g = std::max(1, std::min((ctx->NN + kBlock - 1) / kBlock, 4096));
  hipLaunchKernelGGL(k_lump, dim3(g), dim3(kBlock), 0, ctx->stream, ctx->NN, ctx->Mrow, ctx->M, ctx->bcflag, ctx->Md,
                     ctx->MdOrigInv, ctx->MdInv);
  return launch_check(ctx, "k_lump");
}

// ------------------------------------------------------------ Z
// One wavefront per pressure row p. The row's structural columns sit in LDS;
// for each velocity row i of G^T(p) in ascending order the lanes take the
// entries (i, q) of G row i, find q in the Z row and add
// G_c(i,p) * (G_c(i,q) * MdOrigInv[i]) to the c-th partial sum. Rows of G
// have distinct columns, so no two lanes of one step hit the same slot; the
// steps are ordered by a barrier. Z = (Z1 + Z2) + Z3 reproduces the value
// order of three cs_multiply calls and two value-wise adds (BCH:3384-3457).
__global__ __launch_bounds__(kWave) void k_build_z(int NNp, const int* __restrict__ Zrow, const int* __restrict__ Zcol,
                                                   double* __restrict__ Z, const int* __restrict__ GTrow,
                                                   const int* __restrict__ GTcol, const double* __restrict__ GT,
                                                   const int* __restrict__ Grow, const int* __restrict__ Gcol,
                                                   const double* __restrict__ G, const double* __restrict__ MdOrigInv) {
  __shared__ int zc[kZRowMax];
  __shared__ double z1[kZRowMax], z2[kZRowMax], z3[kZRowMax];
  const int p = blockIdx.x;
  const int t = threadIdx.x;
  const int zb = Zrow[p], zn = Zrow[p + 1] - zb;
  for (int s = t; s < zn; s += kWave) { zc[s] = Zcol[zb + s]; z1[s] = 0.0; z2[s] = 0.0; z3[s] = 0.0; }
  __syncthreads();
  for (int k = GTrow[p]; k < GTrow[p + 1]; k++) {
    const int i = GTcol[k];
    const double a0 = GT[3 * (size_t)k], a1 = GT[3 * (size_t)k + 1], a2 = GT[3 * (size_t)k + 2];
    const double m = MdOrigInv[i];
    const int gb = Grow[i], ge = Grow[i + 1];
    for (int g = gb + t; g < ge; g += kWave) {
      const int s = find_col(zc, 0, zn, Gcol[g]);
      z1[s] += a0 * (G[3 * (size_t)g] * m);
      z2[s] += a1 * (G[3 * (size_t)g + 1] * m);
      z3[s] += a2 * (G[3 * (size_t)g + 2] * m);
    }
    __syncthreads();
  }
  for (int s = t; s < zn; s += kWave) Z[zb + s] = (z1[s] + z2[s]) + z3[s];
}

// Penalty pin (BCH:4167-4177, only for node > 0) and the Jacobi inverse
// diagonal in FP64 and FP32 (the solvers' FP32 values of Z are made from
// the pruned rows, fsx_build_Z).
__global__ void k_z_finish(int NNp, const int* __restrict__ Zrow, const int* __restrict__ Zcol, double* __restrict__ Z,
                           double* __restrict__ dinv, float* __restrict__ dinvf, int zeroP) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < NNp; p += gridDim.x * blockDim.x) {
    double d = 0.0;
    for (int k = Zrow[p]; k < Zrow[p + 1]; k++) {
      double v = Z[k];
      if (Zcol[k] == p) {
        if (zeroP > 0 && p == zeroP) { v = v * 1000.0; Z[k] = v; }
        d = v;
      }
    }
    dinv[p] = 1.0 / d;
    dinvf[p] = (float)(1.0 / d);
  }
}

int launch_build_Z(fsx_ctx* ctx) {
  hipLaunchKernelGGL(k_build_z, dim3(ctx->NNp), dim3(kWave), 0, ctx->stream, ctx->NNp, ctx->Zrow, ctx->Zcol, ctx->Z,
                     ctx->GTrow, ctx->GTcol, ctx->GT, ctx->Grow, ctx->Gcol, ctx->G, ctx->MdOrigInv);
  FSX_TRY(launch_check(ctx, "k_build_z"));
  const int g = std::max(1, std::min((ctx->NNp + kBlock - 1) / kBlock, 4096));
  hipLaunchKernelGGL(k_z_finish, dim3(g), dim3(kBlock), 0, ctx->stream, ctx->NNp, ctx->Zrow, ctx->Zcol, ctx->Z,
                     ctx->Zdinv, ctx->Zdinvf, ctx->zeroP);
  return launch_check(ctx, "k_z_finish");
}

// The Jacobi inverses of Z values that already carry the penalty pin (an
// uploaded Z): k_z_finish without pinning again.
int launch_z_refresh(fsx_ctx* ctx) {
  const int g = std::max(1, std::min((ctx->NNp + kBlock - 1) / kBlock, 4096));
  hipLaunchKernelGGL(k_z_finish, dim3(g), dim3(kBlock), 0, ctx->stream, ctx->NNp, ctx->Zrow, ctx->Zcol, ctx->Z,
                     ctx->Zdinv, ctx->Zdinvf, -1);
  return launch_check(ctx, "k_z_finish");
}

// ------------------------------------------------------------ convection
// R1 -= A(Un) UnpHalf_prev element by element, matrix free (BCH:3560-3654):
//   R1e_c,i = sum_k Sv_ki f_k (u0_k . grad uprev_c(k)),
//   u0_k = sum_j Sv_kj Un_j,  grad uprev_c(k) = J_k^-1 sum_j dSv_j(k) uprev_c,j,
// which equals sum_j Ae_ij uprev_c,j of the reference's explicit Ae.
//
// Persistent form: a block stages the shape tables in LDS once, and each
// half-wavefront (32 lanes) walks elements of the colour at a grid stride,
// one element at a time (lane = Gauss point, then lane = node), so two
// elements share a wavefront and 54 of its 64 lanes work. The next
// element's node values and corner coordinates are loaded into registers
// while the current one computes. LDS holds the tables packed per (node,
// Gauss point) as (Sv, dSv/dxi, dSv/deta, dSv/dzeta) and the node values as
// (Un_x, Un_y, Un_z, Uprev_x, Uprev_y, Uprev_z), so one step of the inner
// sum reads five 16-byte words; the arithmetic and its order are those of
// the per-element form. Each table a lane-indexed phase reads has that
// lane's index fastest (Gauss point in the first phase, node in the second),
// so the lanes of a wavefront read consecutive LDS words instead of a few
// banks at a stride of 27 or 24 entries.
constexpr int kConvWaves = 4;      // 3 blocks per CU fit (140 VGPRs, 46 KB LDS); 8-wave blocks at 128 VGPRs spill and run slower
constexpr int kConvHalves = 2 * kConvWaves;

// orders LDS traffic between the lanes of one wavefront
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// packed LDS records of the compute type CT
template <class CT> struct alignas(4 * sizeof(CT)) CT4 { CT x, y, z, w; };
template <class CT> struct alignas(2 * sizeof(CT)) CT2 { CT x, y; };

// jacobian() in the compute type CT
template <class CT>
__device__ __forceinline__ void jacobian_ct(const CT (*ec)[3], const CT* __restrict__ dSp_k, CT iJ[3][3], CT* det,
                                            int stride) {
  CT J[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      CT s = 0;
#pragma unroll
      for (int m = 0; m < 8; m++) s = s + dSp_k[(m * 3 + i) * stride] * ec[m][j];
      J[i][j] = s;
    }
  iJ[0][0] =   J[1][1]*J[2][2] - J[2][1]*J[1][2];
  iJ[0][1] = -(J[0][1]*J[2][2] - J[0][2]*J[2][1]);
  iJ[0][2] =   J[0][1]*J[1][2] - J[1][1]*J[0][2];
  iJ[1][0] = -(J[1][0]*J[2][2] - J[1][2]*J[2][0]);
  iJ[1][1] =   J[2][2]*J[0][0] - J[2][0]*J[0][2];
  iJ[1][2] = -(J[1][2]*J[0][0] - J[1][0]*J[0][2]);
  iJ[2][0] =   J[1][0]*J[2][1] - J[2][0]*J[1][1];
  iJ[2][1] = -(J[2][1]*J[0][0] - J[2][0]*J[0][1]);
  iJ[2][2] =   J[1][1]*J[0][0] - J[1][0]*J[0][1];
  const CT d = J[0][0]*(J[1][1]*J[2][2] - J[2][1]*J[1][2]) +
               J[0][1]*(J[1][2]*J[2][0] - J[1][0]*J[2][2]) +
               J[0][2]*(J[1][0]*J[2][1] - J[1][1]*J[2][0]);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) iJ[i][j] = iJ[i][j] / d;
  *det = d;
}

// CT is the arithmetic type: double (the reference's arithmetic) or float
// (the mixed forms' FP32 assembly: tables, node values, Jacobians and sums in
// FP32, the result added to the FP64 R1).
template <class CT>
__global__ __launch_bounds__(kWave * kConvWaves) void k_convection(int ncolor_elems, const int* __restrict__ elems,
                                                                   const int* __restrict__ L27,
                                                                   const int* __restrict__ P8,
                                                                   const double* __restrict__ xc, ShapeTables tab, int NGP,
                                                                   int NN, const double* __restrict__ Un,
                                                                   const double* __restrict__ Uprev, double* __restrict__ R1) {
  __shared__ CT4<CT> sT[27 * kMaxNGP];                 // (Sv, dSv_0, dSv_1, dSv_2) at [j][k]
  __shared__ CT sSv[kMaxNGP * 27];                     // Sv at [k][i]
  __shared__ CT sdSp[24 * kMaxNGP];                    // dSp at [m * 3 + i][k]
  __shared__ CT sw[kMaxNGP];
  // In FP32 a node's six values sit in two 16-byte records and a Gauss
  // point's three products in one, so that the two loops read 3 and 2 LDS
  // words per step instead of 4 (FP64 keeps 16-byte pairs).
  constexpr bool kPk = sizeof(CT) == 4;
  __shared__ CT2<CT> nod[kPk ? 1 : kConvHalves][27][3];   // (Un_x, Un_y), (Un_z, Up_x), (Up_y, Up_z)
  __shared__ CT4<CT> nod4[kPk ? kConvHalves : 1][27][2];  // (Un_x, Un_y, Un_z, Up_x), (Up_y, Up_z, -, -)
  __shared__ CT conv[kPk ? 1 : kConvHalves][3][kMaxNGP];
  __shared__ CT4<CT> conv4[kPk ? kConvHalves : 1][kMaxNGP];
  __shared__ CT ec[kConvHalves][8][3];
  const int t = threadIdx.x;
  for (int i = t; i < NGP * 27; i += blockDim.x) {   // i = k * 27 + j
    const int k = i / 27, j = i - 27 * k;
    sT[j * NGP + k] = CT4<CT>{(CT)tab.Sv[i], (CT)tab.dSv[3 * i], (CT)tab.dSv[3 * i + 1], (CT)tab.dSv[3 * i + 2]};
    sSv[i] = (CT)tab.Sv[i];
  }
  for (int i = t; i < NGP * 24; i += blockDim.x) {   // i = k * 24 + m * 3 + d
    const int k = i / 24, m3 = i - 24 * k;
    sdSp[m3 * NGP + k] = (CT)tab.dSp[i];
  }
  for (int i = t; i < NGP; i += blockDim.x) sw[i] = (CT)tab.w[i];
  __syncthreads();
  const int lane = t % kWave, l = lane & 31;
  const int hw = (t / kWave) * 2 + (lane >> 5);        // half-wave of this lane in the block
  const int stride = gridDim.x * kConvHalves;
  int slot = blockIdx.x * kConvHalves + hw;
  // registers of the element being fetched: node id and values (l < 27),
  // one corner coordinate (l < 24)
  int nn = 0;
  CT un[3] = {0, 0, 0}, up[3] = {0, 0, 0}, xv = 0;
  auto fetch = [&](int sl) {
    const bool act = sl < ncolor_elems;
    const int e = act ? elems[sl] : 0;
    if (act && l < 27) {
      nn = L27[(size_t)e * 27 + l];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        un[c] = (CT)Un[3 * (size_t)nn + c];
        up[c] = (CT)Uprev[3 * (size_t)nn + c];
      }
    }
    int pn = 0;   // pressure id (row of xc) of corner l
    if (act && l < 8) pn = P8[(size_t)e * 8 + l];
    const int cn = __shfl(pn, (lane & 32) + (l < 24 ? l / 3 : 0), kWave);
    if (act && l < 24) xv = (CT)xc[3 * (size_t)cn + l % 3];
  };
  fetch(slot);
  for (; slot < ncolor_elems; slot += stride) {
    // stage the fetched element, then start fetching the next one
    const int node = nn;
    if (l < 27) {
      if constexpr (kPk) {
        nod4[hw][l][0] = CT4<CT>{un[0], un[1], un[2], up[0]};
        nod4[hw][l][1] = CT4<CT>{up[1], up[2], 0, 0};
      } else {
        nod[hw][l][0] = CT2<CT>{un[0], un[1]};
        nod[hw][l][1] = CT2<CT>{un[2], up[0]};
        nod[hw][l][2] = CT2<CT>{up[1], up[2]};
      }
    }
    if (l < 24) ec[hw][l / 3][l % 3] = xv;
    wave_sync();
    fetch(slot + stride);
    if (l < NGP) {
      const int k = l;
      CT iJ[3][3], det;
      jacobian_ct<CT>(ec[hw], sdSp + k, iJ, &det, NGP);
      const CT f = det * sw[k];
      CT u0[3] = {0, 0, 0}, gr[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int j = 0; j < 27; j++) {
        const CT4<CT> T = sT[j * NGP + k];
        CT u0v[3], upv[3];
        if constexpr (kPk) {
          const CT4<CT> a = nod4[hw][j][0], b = nod4[hw][j][1];
          u0v[0] = a.x; u0v[1] = a.y; u0v[2] = a.z; upv[0] = a.w; upv[1] = b.x; upv[2] = b.y;
        } else {
          const CT2<CT> a = nod[hw][j][0], b = nod[hw][j][1], c2 = nod[hw][j][2];
          u0v[0] = a.x; u0v[1] = a.y; u0v[2] = b.x; upv[0] = b.y; upv[1] = c2.x; upv[2] = c2.y;
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
          u0[c] += T.x * u0v[c];
          gr[c][0] += T.y * upv[c]; gr[c][1] += T.z * upv[c]; gr[c][2] += T.w * upv[c];
        }
      }
      CT cv[3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        CT g[3];
#pragma unroll
        for (int d = 0; d < 3; d++) g[d] = iJ[d][0] * gr[c][0] + iJ[d][1] * gr[c][1] + iJ[d][2] * gr[c][2];
        cv[c] = f * (u0[0] * g[0] + u0[1] * g[1] + u0[2] * g[2]);
      }
      if constexpr (kPk) {
        conv4[hw][k] = CT4<CT>{cv[0], cv[1], cv[2], 0};
      } else {
#pragma unroll
        for (int c = 0; c < 3; c++) conv[hw][c][k] = cv[c];
      }
    }
    wave_sync();
    if (l < 27) {
      const int i = l;
      CT r[3] = {0, 0, 0};
      for (int k = 0; k < NGP; k++) {
        const CT sv = sSv[k * 27 + i];
        if constexpr (kPk) {
          const CT4<CT> cv = conv4[hw][k];
          r[0] += sv * cv.x; r[1] += sv * cv.y; r[2] += sv * cv.z;
        } else {
#pragma unroll
          for (int c = 0; c < 3; c++) r[c] += sv * conv[hw][c][k];
        }
      }
#pragma unroll
      for (int c = 0; c < 3; c++) R1[3 * (size_t)node + c] -= (double)r[c];
    }
    wave_sync();   // the next element's staging overwrites this one's LDS
  }
}

// The FP32 convection element by element, without the colouring: one
// thread per (element, velocity component) of the sorted element order. The
// thread holds its element's Un (three components), UnpHalf_prev (its
// component) and corner coordinates in registers and walks the Gauss points
// with the shape tables as scalar loads (the same row for the whole
// wavefront), in the arithmetic of k_convection<float>:
//   ye[(i * NE + e) * 3 + c] = sum_k Sv_ki f_k (u0_k . grad uprev_c(k)).
// k_conv_gather then sets R1 = -(sum of each node's element rows), the rows
// of one node in ascending element order. (Measured at 64^3: 262 us, and
// 43.5 us for the gather, against 8 colour launches of 40.5 us and the
// zeroing of R1. Sharing u0 between an element's three threads, so that
// each holds one component of Un, took it to 291 us.)
__global__ __launch_bounds__(kBlock) void k_conv_elem(int NE, int NGP, const int* __restrict__ L,
                                                      const int* __restrict__ P8, const double* __restrict__ xc,
                                                      ShapeTablesF tab, const double* __restrict__ Un,
                                                      const double* __restrict__ Uprev, float* __restrict__ ye) {
  const long long t = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (t >= 3LL * NE) return;
  const int e = (int)(t / 3), c = (int)(t % 3);
  float un[27][3], up[27], y[27];
#pragma unroll
  for (int j = 0; j < 27; j++) {
    const size_t n = 3 * (size_t)L[(size_t)j * NE + e];
    un[j][0] = (float)Un[n]; un[j][1] = (float)Un[n + 1]; un[j][2] = (float)Un[n + 2];
    up[j] = (float)Uprev[n + c];
    y[j] = 0.0f;
  }
  float ec[8][3];
#pragma unroll
  for (int m = 0; m < 8; m++) {
    const size_t pn = 3 * (size_t)P8[(size_t)e * 8 + m];
#pragma unroll
    for (int d = 0; d < 3; d++) ec[m][d] = (float)xc[pn + d];
  }
  for (int k = 0; k < NGP; k++) {
    float iJ[3][3], det;
    jacobian_ct<float>(ec, tab.dSp + k * 24, iJ, &det, 1);
    const float f = det * tab.w[k];
    const float* __restrict__ sv = tab.Sv + k * 27;
    const float* __restrict__ ds = tab.dSv + k * 81;
    float u0[3] = {0.0f, 0.0f, 0.0f}, gr[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int j = 0; j < 27; j++) {
#pragma unroll
      for (int d = 0; d < 3; d++) u0[d] += sv[j] * un[j][d];
      gr[0] += ds[3 * j] * up[j]; gr[1] += ds[3 * j + 1] * up[j]; gr[2] += ds[3 * j + 2] * up[j];
    }
    float g[3];
#pragma unroll
    for (int d = 0; d < 3; d++) g[d] = iJ[d][0] * gr[0] + iJ[d][1] * gr[1] + iJ[d][2] * gr[2];
    const float cv = f * (u0[0] * g[0] + u0[1] * g[1] + u0[2] * g[2]);
#pragma unroll
    for (int i = 0; i < 27; i++) y[i] += sv[i] * cv;
  }
#pragma unroll
  for (int i = 0; i < 27; i++) ye[((size_t)i * NE + e) * 3 + c] = y[i];
}

__global__ __launch_bounds__(kBlock) void k_conv_gather(int NN, const int* __restrict__ inc_ptr,
                                                        const int* __restrict__ inc, const float* __restrict__ ye,
                                                        double* __restrict__ R1) {
  for (int r = blockIdx.x * kBlock + threadIdx.x; r < NN; r += gridDim.x * kBlock) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    const int t1 = inc_ptr[r + 1];
    for (int t = inc_ptr[r]; t < t1; t++) {
      const float* y = ye + 3 * (size_t)inc[t];
      s0 += (double)y[0]; s1 += (double)y[1]; s2 += (double)y[2];
    }
    R1[3 * (size_t)r] = -s0; R1[3 * (size_t)r + 1] = -s1; R1[3 * (size_t)r + 2] = -s2;
  }
}

static int launch_convection_elements(fsx_ctx* ctx) {
  if (!ctx->el_ready) FSX_TRY(elements_setup(ctx));
  if (!ctx->conv_ye) FSX_TRY(alloc(ctx, &ctx->conv_ye, (size_t)ctx->NE * 81));
  hipLaunchKernelGGL(k_conv_elem, dim3((unsigned)((3LL * ctx->NE + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream,
                     ctx->NE, ctx->NGP, ctx->el_L27, ctx->el_P8, ctx->xc, ctx->tabf, ctx->Un, ctx->UnpHalf_prev,
                     ctx->conv_ye);
  hipLaunchKernelGGL(k_conv_gather, dim3((unsigned)std::min((ctx->NN + kBlock - 1) / kBlock, 8192)), dim3(kBlock), 0,
                     ctx->stream, ctx->NN, ctx->el_inc_ptr, ctx->el_inc, ctx->conv_ye, ctx->R1);
  return launch_check(ctx, "k_conv_elem");
}

template <class CT>
static int launch_convection_t(fsx_ctx* ctx, int* grid_cache) {
  if (*grid_cache == 0) {
    // as many blocks as fit on the device at once (the kernel is persistent)
    int per_cu = 0, cus = 0;
    FSX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_convection<CT>, kWave * kConvWaves, 0));
    FSX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    *grid_cache = std::max(1, per_cu * cus);
  }
  const int ncol = (int)ctx->color_ptr.size() - 1;
  for (int c = 0; c < ncol; c++) {
    const int n = ctx->color_ptr[c + 1] - ctx->color_ptr[c];
    if (n == 0) continue;
    const int g = std::min(*grid_cache, (n + kConvHalves - 1) / kConvHalves);
    hipLaunchKernelGGL(k_convection<CT>, dim3(g), dim3(kWave * kConvWaves), 0, ctx->stream, n,
                       ctx->color_order + ctx->color_ptr[c], ctx->L27, ctx->P8, ctx->xc, ctx->tab, ctx->NGP, ctx->NN, ctx->Un,
                       ctx->UnpHalf_prev, ctx->R1);
  }
  return launch_check(ctx, "k_convection");
}

// FP64 arithmetic, or FP32 (fsx_set_convection_precision, the mixed forms).
// When the next step1 can run element by element, the FP32 convection is
// only recorded here and formed in that pass (launch_step1).
int launch_convection(fsx_ctx* ctx) {
  if (step1_elements_ok(ctx)) {
    ctx->conv_pending = true;
    return FSX_OK;
  }
  return launch_convection_now(ctx);
}

int launch_convection_now(fsx_ctx* ctx) {
  ctx->conv_pending = false;
  // the FP32 element form writes every entry of R1 (BCH:3541-3545's zeroing included)
  if (ctx->conv_bits == 32 && ctx->conv_elements) return launch_convection_elements(ctx);
  FSX_TRY(launch_fill(ctx, ctx->R1, 0.0, 3LL * ctx->NN));                  // BCH:3541-3545
  if (ctx->conv_bits == 32) return launch_convection_t<float>(ctx, &ctx->conv_grid_f);
  return launch_convection_t<double>(ctx, &ctx->conv_grid);
}

}  // namespace fsx
