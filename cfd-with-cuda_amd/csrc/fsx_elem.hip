// fsx_elem.hip -- step1 of the mixed forms element by element on the matrix
// cores: the FP32 convection, K UnpHalf_prev and K Acc_prev of every element
// in one pass (v_mfma_f32_16x16x4f32), then one pass over the nodes that
// sums each node's element rows and applies step1's update.
//
// The K products need no CSR values: per element (BCH:3176-3231)
//   (K_e u)_i = nu sum_k f_k gradSv_i(k) . grad u(k),   f_k = detJ_k w_k,
// with gradSv_i = iJ dSv_i (the reference gradient dSv_i mapped by the
// inverse Jacobian), so that at each Gauss point
//   gr = sum_j dSv_j u_j          (reference gradient of the field)
//   V  = nu f iJ^T iJ gr          (3 values)
//   (K_e u)_i += sum_m dSv_i[m] V[m].
// The convection adds sum_k Sv_ki f_k (u0_k . iJ gr_k) (BCH:3548-3700) to the
// rows of K UnpHalf_prev, u0 = sum_j Sv_j Un_j.
//
// A wavefront takes 16 elements (the MFMA columns) and one field set: set 0
// the convection and K UnpHalf_prev (fields Un and UnpHalf_prev), set 1
// K Acc_prev. For a tile of 16 Gauss points the interpolations are MFMAs
// with the shape tables as the A operand (rows Gauss points, K the 27 nodes
// padded to 28) and the node values as B; their results leave lane
// (r, q) = (l % 16, l / 16) holding, for element r, the Gauss points
// 16t + 4q + v (v = 0..3). The pointwise work (Jacobian, V, convection) runs
// there in registers, and the products back onto the nodes are MFMAs whose
// K index visits (Gauss point, direction) pairs in that same lane order, so
// the pointwise results are the B operand as they stand: no LDS round trip.
// The A operands of all MFMAs are packed per lane on the host, in the order
// of use, and staged in LDS once per block.
//
// Node values come from the FP32 shadows (Un's compact record, the
// interleaved UnpHalf_prev / Acc_prev records); products and sums are FP32,
// the node sums FP64 in ascending element order (bitwise reproducible).
#include "fsx_device.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace fsx {

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Packed A operands of one Gauss-point tile t (points 16t..16t+15), each
// entry 64 floats (one a lane), r = lane % 16, q = lane / 16:
//   F1 [ks]       Sv[16t + r][4ks + q]                u0 = Sv Un
//   F2 [m][ks]    dSv[16t + r][4ks + q][m]            gr_m = dSv_m u
//   BK [mt][v][m] dSv[16t + 4q + v][16mt + r][m]      y += dSv^T V
//   BS [mt][v]    Sv[16t + 4q + v][16mt + r]          y += Sv^T conv
//   FJ [ks][i]    dSp[16t + r][4ks + q][i]            J_ij = dSp_i x_j
// Entries outside the 27 nodes, the 8 corners or the NGP points are zero.
// Then the NGP weights w (zero-padded to 32).
constexpr int kEkF1 = 0, kEkF2 = 7, kEkBK = 28, kEkBS = 52, kEkFJ = 60, kEkPerTile = 66, kEkMaxTiles = 2;
constexpr int kEkWaves = kBlock / kWave;

int elem_tables_setup(fsx_ctx* ctx, const double* t, int NGP) {
  const double* Sv = t;
  const double* dSv = t + NGP * 27;
  const double* dSp = dSv + NGP * 81 + NGP * 8;
  const double* w = dSp + NGP * 24;
  const int ntiles = (NGP + 15) / 16;
  std::vector<float> a((size_t)ntiles * kEkPerTile * kWave + 32, 0.0f);
  auto sv = [&](int p, int j) { return p < NGP && j < 27 ? (float)Sv[p * 27 + j] : 0.0f; };
  auto dsv = [&](int p, int j, int m) { return p < NGP && j < 27 ? (float)dSv[(p * 27 + j) * 3 + m] : 0.0f; };
  auto dsp = [&](int p, int m, int i) { return p < NGP ? (float)dSp[p * 24 + m * 3 + i] : 0.0f; };
  for (int tt = 0; tt < ntiles; tt++) {
    float* T = a.data() + (size_t)tt * kEkPerTile * kWave;
    for (int l = 0; l < kWave; l++) {
      const int r = l % 16, q = l / 16;
      for (int ks = 0; ks < 7; ks++) T[(kEkF1 + ks) * kWave + l] = sv(16 * tt + r, 4 * ks + q);
      for (int m = 0; m < 3; m++)
        for (int ks = 0; ks < 7; ks++) T[(kEkF2 + m * 7 + ks) * kWave + l] = dsv(16 * tt + r, 4 * ks + q, m);
      for (int mt = 0; mt < 2; mt++)
        for (int v = 0; v < 4; v++) {
          for (int m = 0; m < 3; m++)
            T[(kEkBK + (mt * 4 + v) * 3 + m) * kWave + l] = dsv(16 * tt + 4 * q + v, 16 * mt + r, m);
          T[(kEkBS + mt * 4 + v) * kWave + l] = sv(16 * tt + 4 * q + v, 16 * mt + r);
        }
      for (int ks = 0; ks < 2; ks++)
        for (int i = 0; i < 3; i++) T[(kEkFJ + ks * 3 + i) * kWave + l] = dsp(16 * tt + r, 4 * ks + q, i);
    }
  }
  float* W = a.data() + (size_t)ntiles * kEkPerTile * kWave;
  for (int p = 0; p < NGP; p++) W[p] = (float)w[p];
  if (ctx->ek_tab) dev_free(ctx, ctx->ek_tab);
  ctx->ek_tab = nullptr;
  FSX_TRY(alloc(ctx, &ctx->ek_tab, a.size()));
  FSX_HIP(hipMemcpy(ctx->ek_tab, a.data(), a.size() * sizeof(float), hipMemcpyHostToDevice));
  ctx->ek_ntiles = ntiles;
  return FSX_OK;
}

__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Lane (r, q) of batch b (16 elements from b * 16) holds nodes 4ks + q of
// element r (node 27 is padding) and its centre node 26: their ids in n[0..6]
// and n[7], -1 where there is none.
__device__ __forceinline__ void ek_ids(int NE, int nb, int b, const int* __restrict__ L, int r, int q, int* n) {
  const int e = b * 16 + r;
  const bool ok = b < nb && e < NE;
#pragma unroll
  for (int ks = 0; ks < 7; ks++) n[ks] = ok && 4 * ks + q < 27 ? L[(unsigned)(4 * ks + q) * (unsigned)NE + e] : -1;
  n[7] = ok ? L[26u * (unsigned)NE + e] : -1;
}

// The differentiated field's values at those ids and corners q and 4 + q of
// the element (k_elem_corners). The field enters less its value at the
// centre node: the gradients are the same (sum_j dSv_j = 0), and their FP32
// sums no longer round a constant part that cancels.
__device__ __forceinline__ void ek_vals(int b, const int* n, const float* __restrict__ Xl,
                                        const float4* __restrict__ Fp, int r, int q, float (&fv)[7][3],
                                        float (&xl)[2][3]) {
  const bool ok = n[7] >= 0;
  const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float4 c0 = ok ? Fp[2u * (unsigned)n[7]] : z;
#pragma unroll
  for (int ks = 0; ks < 7; ks++) {
    float4 c = z;
    if (n[ks] >= 0) {
      c = Fp[2u * (unsigned)n[ks]];
      c.x -= c0.x; c.y -= c0.y; c.z -= c0.z;
    }
    fv[ks][0] = c.x; fv[ks][1] = c.y; fv[ks][2] = c.z;
  }
  const unsigned e = (unsigned)b * 16u + r;
#pragma unroll
  for (int ks = 0; ks < 2; ks++)
#pragma unroll
    for (int d = 0; d < 3; d++) xl[ks][d] = ok ? Xl[e * 24u + (4 * ks + q) * 3 + d] : 0.0f;
}

// Un at those ids (set 0)
__device__ __forceinline__ void ek_un(const int* n, const float4* __restrict__ Ucmp, float (&un)[7][3]) {
#pragma unroll
  for (int ks = 0; ks < 7; ks++) {
    const float4 a = n[ks] >= 0 ? Ucmp[n[ks]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    un[ks][0] = a.x; un[ks][1] = a.y; un[ks][2] = a.z;
  }
}

// The interpolations of one Gauss-point tile (A: its packed operands, this
// lane's column): J, the reference gradients gr and (set 0) u0, in that
// order, so that Un's loads have the time of the others' products.
template <int FS>
__device__ __forceinline__ void ek_forward(const float* __restrict__ A, const float (&un)[7][3],
                                           const float (&fv)[7][3], const float (&xl)[2][3], f32x4 (&J)[3][3],
                                           f32x4 (&u0)[3], f32x4 (&gr)[3][3]) {
  const f32x4 z = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  // J[i][j] = sum_m dSp[m][i] x_m[j] at the tile's points (calcJacob, BCH:2553-2590)
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) J[i][j] = z;
#pragma unroll
  for (int ks = 0; ks < 2; ks++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const float a = A[(kEkFJ + ks * 3 + i) * kWave];
#pragma unroll
      for (int j = 0; j < 3; j++) J[i][j] = mfma4(a, xl[ks][j], J[i][j]);
    }
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int m = 0; m < 3; m++) gr[c][m] = z;
#pragma unroll
  for (int m = 0; m < 3; m++)
#pragma unroll
    for (int ks = 0; ks < 7; ks++) {
      const float a = A[(kEkF2 + m * 7 + ks) * kWave];
#pragma unroll
      for (int c = 0; c < 3; c++) gr[c][m] = mfma4(a, fv[ks][c], gr[c][m]);
    }
  if constexpr (FS == 0) {
#pragma unroll
    for (int c = 0; c < 3; c++) u0[c] = z;
#pragma unroll
    for (int ks = 0; ks < 7; ks++) {
      const float a = A[(kEkF1 + ks) * kWave];
#pragma unroll
      for (int c = 0; c < 3; c++) u0[c] = mfma4(a, un[ks][c], u0[c]);
    }
  }
}

// The pointwise work at the tile's points 16t + 4q + v (V = nu f iJ^T iJ gr
// and, set 0, the convection) and the products back onto the nodes, added
// to out[field][node tile].
template <int FS>
__device__ __forceinline__ void ek_back(const float* __restrict__ A, int t, int q, bool ok, int NGP,
                                        const float* __restrict__ sW, float nu, const f32x4 (&J)[3][3],
                                        const f32x4 (&u0)[3], f32x4 (&gr)[3][3], f32x4 (&out)[3][2]) {
  f32x4 cv[3];
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int p = 16 * t + 4 * q + v;
    // iJ = C / det with C the cofactor matrix (the inverse of calcJacob);
    // f = det w, so nu f iJ^T iJ = (nu w / det) C^T C and f iJ^T = w C^T
    float C[3][3];
    const float j00 = J[0][0][v], j01 = J[0][1][v], j02 = J[0][2][v];
    const float j10 = J[1][0][v], j11 = J[1][1][v], j12 = J[1][2][v];
    const float j20 = J[2][0][v], j21 = J[2][1][v], j22 = J[2][2][v];
    C[0][0] = j11 * j22 - j21 * j12;
    C[0][1] = j02 * j21 - j01 * j22;
    C[0][2] = j01 * j12 - j11 * j02;
    C[1][0] = j12 * j20 - j10 * j22;
    C[1][1] = j22 * j00 - j20 * j02;
    C[1][2] = j10 * j02 - j12 * j00;
    C[2][0] = j10 * j21 - j11 * j20;
    C[2][1] = j20 * j01 - j21 * j00;
    C[2][2] = j11 * j00 - j10 * j01;
    const float det = j00 * C[0][0] + j01 * C[1][0] + j02 * C[2][0];
    const bool val = ok && p < NGP;
    const float wp = val ? sW[p] : 0.0f;
    const float sq = val ? nu * wp * __builtin_amdgcn_rcpf(det) : 0.0f;
    float Q[3][3];
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
      for (int n = m; n < 3; n++) {
        const float s2 = C[0][m] * C[0][n] + C[1][m] * C[1][n] + C[2][m] * C[2][n];
        Q[m][n] = Q[n][m] = sq * s2;
      }
    if constexpr (FS == 0) {
      float w3[3];   // f iJ^T u0: conv_c = w3 . gr_c
#pragma unroll
      for (int m = 0; m < 3; m++) w3[m] = wp * (u0[0][v] * C[0][m] + u0[1][v] * C[1][m] + u0[2][v] * C[2][m]);
#pragma unroll
      for (int c = 0; c < 3; c++) cv[c][v] = w3[0] * gr[c][0][v] + w3[1] * gr[c][1][v] + w3[2] * gr[c][2][v];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const float g0 = gr[c][0][v], g1 = gr[c][1][v], g2 = gr[c][2][v];
#pragma unroll
      for (int m = 0; m < 3; m++) gr[c][m][v] = Q[m][0] * g0 + Q[m][1] * g1 + Q[m][2] * g2;
    }
    __builtin_amdgcn_sched_barrier(0);   // one point's temporaries at a time
  }
#pragma unroll
  for (int mt = 0; mt < 2; mt++) {
#pragma unroll
    for (int v = 0; v < 4; v++)
#pragma unroll
      for (int m = 0; m < 3; m++) {
        const float a = A[(kEkBK + (mt * 4 + v) * 3 + m) * kWave];
#pragma unroll
        for (int c = 0; c < 3; c++) out[c][mt] = mfma4(a, gr[c][m][v], out[c][mt]);
      }
    if constexpr (FS == 0) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const float a = A[(kEkBS + mt * 4 + v) * kWave];
#pragma unroll
        for (int c = 0; c < 3; c++) out[c][mt] = mfma4(a, cv[c][v], out[c][mt]);
      }
    }
  }
}

// Batches b, b + bs, ... < nb of field set FS. Writes ye[(i * NE + e) * ys +
// 3 FS + c] for the 27 nodes i of each element e. Software-pipelined: the
// next batch's node ids, differentiated field and corners are loaded once
// the last tile's interpolations have used this batch's (they arrive during
// its pointwise work and products); Un follows at the batch's start.
template <int FS>
__device__ __forceinline__ void elem_run(int NE, int NGP, int ntiles, int b, int bs, int nb,
                                         const int* __restrict__ L, const float* __restrict__ Xl,
                                         const float* __restrict__ sA, const float* __restrict__ sW, float nu,
                                         const float4* __restrict__ Ucmp, const float4* __restrict__ Fp,
                                         float* __restrict__ ye, int ys) {
  const int lane = threadIdx.x & (kWave - 1), r = lane & 15, q = lane >> 4;
  if (b >= nb) return;
  const float* __restrict__ A0 = sA + lane;
  const float* __restrict__ A1 = sA + kEkPerTile * kWave + lane;
  int nid[8];
  float un[7][3], fv[7][3], xl[2][3];
  ek_ids(NE, nb, b, L, r, q, nid);
  ek_vals(b, nid, Xl, Fp, r, q, fv, xl);
  for (; b < nb; b += bs) {
    const int e = b * 16 + r;
    const bool ok = e < NE;
    if constexpr (FS == 0) ek_un(nid, Ucmp, un);
    f32x4 out[3][2];
#pragma unroll
    for (int c = 0; c < 3; c++) out[c][0] = out[c][1] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    f32x4 J[3][3], u0[3], gr[3][3];
    ek_forward<FS>(A0, un, fv, xl, J, u0, gr);
    if (ntiles == 2) {
      ek_back<FS>(A0, 0, q, ok, NGP, sW, nu, J, u0, gr, out);
      // Un again rather than kept over the products (the loads hit the cache)
      if constexpr (FS == 0) ek_un(nid, Ucmp, un);
      ek_forward<FS>(A1, un, fv, xl, J, u0, gr);
    }
    ek_ids(NE, nb, b + bs, L, r, q, nid);
    ek_vals(b + bs, nid, Xl, Fp, r, q, fv, xl);
    ek_back<FS>(ntiles == 2 ? A1 : A0, ntiles - 1, q, ok, NGP, sW, nu, J, u0, gr, out);
    if (ok) {
#pragma unroll
      for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int i = 16 * mt + 4 * q + v;
          if (i < 27) {
            float* y = ye + (((unsigned)i * (unsigned)NE + e) * (unsigned)ys + 3 * FS);
#pragma unroll
            for (int c = 0; c < 3; c++) y[c] = out[c][mt][v];
          }
        }
    }
  }
}

// The corner coordinates of each (sorted) element relative to its corner 0,
// differenced in FP64 and rounded: Xl[e * 24 + 3m + d]. J is the same
// (sum_m dSp_m = 0) and its FP32 entries keep their relative precision,
// which absolute coordinates would lose in proportion to |x| / h.
__global__ void k_elem_corners(int NE, const int* __restrict__ P8, const double* __restrict__ xc,
                               float* __restrict__ Xl) {
  const long long n = 24LL * NE;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const long long e = t / 24;
    const int m = (int)(t % 24) / 3, d = (int)(t % 3);
    Xl[t] = (float)(xc[3 * (size_t)P8[e * 8 + m] + d] - xc[3 * (size_t)P8[e * 8] + d]);
  }
}

// Work items (16 elements, field set): with K Acc_prev the even wavefronts
// of the grid take field set 0 and the odd ones set 1, each walking the
// batches with the stride of its half of the grid (the two sets of a batch
// run on neighbouring wavefronts); without, every wavefront takes set 0.
template <bool ACC>
__global__ __launch_bounds__(kBlock, 2) void k_elem_step1(int NE, int NGP, int ntiles, int nb,
                                                       const int* __restrict__ L, const float* __restrict__ Xl,
                                                       const float* __restrict__ tab, float nu,
                                                       const float4* __restrict__ Ucmp,
                                                       const float4* __restrict__ Urp, const float4* __restrict__ Arp,
                                                       float* __restrict__ ye, int ys) {
  __shared__ float sA[kEkMaxTiles * kEkPerTile * kWave + 32];
  const int na = ntiles * kEkPerTile * kWave + 32;   // the tiles' operands, then the weights
  for (int i = threadIdx.x; i < na; i += kBlock) sA[i] = tab[i];
  __syncthreads();
  const float* sW = sA + ntiles * kEkPerTile * kWave;
  const int w = blockIdx.x * kEkWaves + threadIdx.x / kWave, nw = gridDim.x * kEkWaves;
  if (ACC && (w & 1))
    elem_run<1>(NE, NGP, ntiles, w >> 1, nw >> 1, nb, L, Xl, sA, sW, nu, Ucmp, Arp, ye, ys);
  else
    elem_run<0>(NE, NGP, ntiles, ACC ? w >> 1 : w, ACC ? nw >> 1 : nw, nb, L, Xl, sA, sW, nu, Ucmp, Urp, ye, ys);
}

// step1 at each node from the element rows: s = sum of the node's rows
// (ascending element order, FP64); R = -s_(conv + K UnpHalf_prev) - G Pn, 0
// at BC rows; UnpHalf = Un + dt R MdInv; KAcc = s_(K Acc_prev) with ACC; the
// FP32 shadows as k_step1 writes them. GMODE 1 forms G Pn (FP32 values, FP64
// products, as k_step1) and keeps it in GPn, GMODE 2 reads it back. LPR
// lanes a node (8 for GMODE 1, whose G rows hold about 16 entries, else 1):
// lane l adds entries l, l + LPR, ... of the node's lists, UG at a time, and
// the lane sums are added in a fixed tree; lane c then updates component c.
template <bool ACC, int GMODE>
__global__ __launch_bounds__(kBlock) void k_step1_nodes(int NN, const int* __restrict__ inc_ptr,
                                                        const int* __restrict__ inc, const float* __restrict__ ye,
                                                        int ys, const int* __restrict__ Grow,
                                                        const int* __restrict__ Gcol, const float* __restrict__ G,
                                                        const double* __restrict__ Pn, double* __restrict__ GPn,
                                                        const double* __restrict__ Un, double* __restrict__ R1,
                                                        const double* __restrict__ MdInv,
                                                        const unsigned char* __restrict__ bc,
                                                        double* __restrict__ UnpHalf, double dt,
                                                        double* __restrict__ KAcc, float* __restrict__ Uout,
                                                        const double* __restrict__ MdOrigInv, double o2,
                                                        float* __restrict__ Dout) {
  constexpr int LPR = GMODE == 1 ? 8 : 1, UG = 2, NS = ACC ? 6 : 3;
  const int l = threadIdx.x % LPR;
  const int r = blockIdx.x * (kBlock / LPR) + threadIdx.x / LPR;
  double s[NS], gg[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < NS; c++) s[c] = 0.0;
  if (r < NN) {
    const int t0 = inc_ptr[r], t1 = inc_ptr[r + 1];
    if constexpr (LPR == 1) {
      // a hexahedral node has up to 8 elements: their ids are loaded before
      // the rows, so the row loads are all in flight together
      int id[8];
#pragma unroll
      for (int k = 0; k < 8; k++) id[k] = t0 + k < t1 ? inc[t0 + k] : 0;
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (t0 + k < t1) {
          const float* y = ye + (size_t)ys * id[k];
#pragma unroll
          for (int c = 0; c < NS; c++) s[c] += (double)y[c];
        }
    }
    for (int t = t0 + (LPR == 1 ? 8 : l); t < t1; t += LPR) {
      const float* y = ye + (size_t)ys * inc[t];
#pragma unroll
      for (int c = 0; c < NS; c++) s[c] += (double)y[c];
    }
    if constexpr (GMODE == 1) {
      const int gs = Grow[r], ge = Grow[r + 1];
      for (int base = gs + l; base < ge; base += UG * LPR) {
        int cl[UG];
        double b[UG][3];
#pragma unroll
        for (int j = 0; j < UG; j++) {
          const int k = base + j * LPR, kk = k < ge ? k : ge - 1;
          cl[j] = Gcol[kk];
#pragma unroll
          for (int c = 0; c < 3; c++) b[j][c] = k < ge ? (double)G[3 * (size_t)kk + c] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < UG; j++) {
          const double p = Pn[cl[j]];
#pragma unroll
          for (int c = 0; c < 3; c++) gg[c] += b[j][c] * p;
        }
      }
    }
  }
  if constexpr (LPR > 1) {
#pragma unroll
    for (int c = 0; c < NS; c++) s[c] = group_sum<LPR>(s[c]);
#pragma unroll
    for (int c = 0; c < 3; c++) gg[c] = group_sum<LPR>(gg[c]);
  }
  if (r >= NN) return;
  const bool isbc = bc[r] != 0;
  const double mi = MdInv[r], mo = ACC ? MdOrigInv[r] : 0.0;
  double un[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const bool mine = LPR == 1 || c == l;
    un[c] = mine ? Un[3 * (size_t)r + c] : 0.0;
    if constexpr (GMODE == 2) gg[c] = mine ? GPn[3 * (size_t)r + c] : 0.0;
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if (LPR > 1 && c != l) continue;
    const size_t i = 3 * (size_t)r + c;
    if constexpr (GMODE == 1) GPn[i] = gg[c];
    const double R = isbc ? 0.0 : -s[c] - gg[c];
    R1[i] = R;
    const double uh = un[c] + dt * R * mi;
    UnpHalf[i] = uh;
    Uout[8 * (size_t)r + c] = (float)uh;
    double d = o2 * uh;
    if constexpr (ACC) {
      KAcc[i] = s[3 + c];
      d = d - mo * s[3 + c];
    }
    Dout[4 * (size_t)r + c] = (float)d;
  }
}

// Whether step1 of the coming inner iteration can run element by element
// with the convection: FP32 convection and matrices, lumped mass, the FP32
// shadows of Un, UnpHalf_prev (and Acc_prev when K Acc_prev is pending)
// current, and K the assembled one (an uploaded K has no element form).
bool step1_elements_ok(const fsx_ctx* ctx) {
  return ctx->use_step1_elements && ctx->conv_bits == 32 && ctx->conv_elements && ctx->Kf && ctx->rec_mem &&
         ctx->mass_mode == FSX_MASS_LUMPED && ctx->ek_tab && !ctx->k_uploaded && ctx->ucmp_un &&
         ctx->urec_prev_ok && (!ctx->kacc_pending || ctx->arec_prev_ok) &&
         27LL * 6 * ctx->NE < (1LL << 32) && 2LL * ctx->NN < (1LL << 32);   // the kernel's 32-bit offsets
}

// The element pass: k_elem_step1 writes every element's rows of the
// convection plus K UnpHalf_prev (and K Acc_prev with kacc) into ek_ye.
int launch_elem_pass(fsx_ctx* ctx, bool kacc) {
  const int NE = ctx->NE;
  if (!ctx->el_ready) FSX_TRY(elements_setup(ctx));
  if (!ctx->ek_ye) FSX_TRY(alloc(ctx, &ctx->ek_ye, (size_t)NE * 27 * 6));
  if (!ctx->ek_xl_ready) {
    if (!ctx->ek_xl) FSX_TRY(alloc(ctx, &ctx->ek_xl, (size_t)NE * 24));
    hipLaunchKernelGGL(k_elem_corners, dim3(std::max(1, std::min((24 * NE + kBlock - 1) / kBlock, 4096))), dim3(kBlock), 0,
                       ctx->stream, NE, ctx->el_P8, ctx->xc, ctx->ek_xl);
    FSX_TRY(launch_check(ctx, "k_elem_corners"));
    ctx->ek_xl_ready = true;
  }
  if (ctx->ek_grid == 0) {
    int cus = 0;
    FSX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    ctx->ek_grid = std::max(1, 8 * cus);
  }
  const int nb = (NE + 15) / 16, nitems = kacc ? 2 * nb : nb, ys = kacc ? 6 : 3;
  const unsigned grid = (unsigned)std::max(1, std::min(ctx->ek_grid, (nitems + kEkWaves - 1) / kEkWaves));
  // (kEkWaves is even: with K Acc_prev each block holds both sets' wavefronts)
  const float nu = (float)ctx->k_viscosity;
  const float4* ucmp = reinterpret_cast<const float4*>(ctx->Ucmp);
  const float4* urp = reinterpret_cast<const float4*>(ctx->Urec_prev);
  const float4* arp = reinterpret_cast<const float4*>(ctx->Arec_prev);
  if (kacc)
    hipLaunchKernelGGL(k_elem_step1<true>, dim3(grid), dim3(kBlock), 0, ctx->stream, NE, ctx->NGP, ctx->ek_ntiles,
                       nb, ctx->el_L27, ctx->ek_xl, ctx->ek_tab, nu, ucmp, urp, arp, ctx->ek_ye, ys);
  else
    hipLaunchKernelGGL(k_elem_step1<false>, dim3(grid), dim3(kBlock), 0, ctx->stream, NE, ctx->NGP, ctx->ek_ntiles,
                       nb, ctx->el_L27, ctx->ek_xl, ctx->ek_tab, nu, ucmp, urp, arp, ctx->ek_ye, ys);
  return launch_check(ctx, "k_elem_step1");
}

// The node pass: k_step1_nodes sums each node's element rows from ek_ye and
// applies step1's update (gload: G Pn kept from an earlier pass, else formed).
int launch_nodes_pass(fsx_ctx* ctx, bool kacc, bool gload) {
  const int NN = ctx->NN;
  const int ys = kacc ? 6 : 3;
  const double o2 = 1.0000000000000000 / (ctx->dt * ctx->dt);
  const dim3 ng1((unsigned)((NN + kBlock - 1) / kBlock)), ng8((unsigned)((NN + kBlock / 8 - 1) / (kBlock / 8)));
#define NODES_LAUNCH(KA, GM)                                                                                          \
  hipLaunchKernelGGL((k_step1_nodes<KA, GM>), GM == 1 ? ng8 : ng1, dim3(kBlock), 0, ctx->stream, NN,                 \
                     ctx->el_inc_ptr, ctx->el_inc, ctx->ek_ye, ys, ctx->Grow, ctx->Gcol, ctx->Gf, ctx->Pn, ctx->GPn,   \
                     ctx->Un, ctx->R1, ctx->MdInv, ctx->bcflag, ctx->UnpHalf, ctx->dt, ctx->KAcc, ctx->Urec,           \
                     ctx->MdOrigInv, o2, ctx->Drec)
  if (kacc) {
    if (gload) NODES_LAUNCH(true, 2); else NODES_LAUNCH(true, 1);
  } else {
    if (gload) NODES_LAUNCH(false, 2); else NODES_LAUNCH(false, 1);
  }
#undef NODES_LAUNCH
  return launch_check(ctx, "k_step1_nodes");
}

int launch_step1_elements(fsx_ctx* ctx, bool kacc, bool gload) {
  FSX_TRY(launch_elem_pass(ctx, kacc));
  return launch_nodes_pass(ctx, kacc, gload);
}

}  // namespace fsx
