// fsx_unet.hip -- learned pressure preconditioner: a symmetric multilevel
// 3-D convolutional network with its own BF16-MFMA forward, transpose and
// weight-gradient kernels, Adam training on the GPU, and training pairs
// produced by the solver itself.
//
// Preconditioner z = dinv r + gain J^T J r (linear, so M(cr) = c M(r); the
// training pairs are normalised by rms(r)). J maps the pressure grid gf to
// features on three coarser node-aligned grids (G0 = every other node of gf,
// G1 of G0, G2 of G1; 65^3 -> 33^3 -> 17^3 -> 9^3), one stride-2 3x3x3
// convolution per level:
//   f0 = B0 C0 r           gf -> G0, 1 -> 16 channels       (VALU)
//   f1 = B1 C1 f0          G0 -> G1, 16 -> 32               (MFMA)
//   f2 = B2 C2 f1          G1 -> G2, 32 -> 64               (MFMA)
// and J^T J r = C0^T B0 (f0 + C1^T B1 (f1 + C2^T f2)), the transpose pass
// running the same convolutions with transposed weights (a U-net whose
// decoder is tied to its encoder; the last step is fused with z). B_l zeroes
// the boundary nodes of grid l: the pressure operator penalises boundary
// values (its mass term), so the slow modes that the coarse levels must
// carry vanish there. The correction is symmetric positive semidefinite by
// construction, so M is SPD and the flexible CG keeps descent. With one
// channel, full-weighting weights (C0) and 1/sqrt(2) x full weighting (C1,
// C2) reproduce the additive multilevel (BPX) preconditioner
// sum_l 2^-(l-1) P_l P_l^T; the network starts there and is trained from it.
// Convolutions are linear, without bias; weights are BF16 for the MFMA
// operands with FP32 master copies; activations are FP32 and rounded to BF16
// where they feed an MFMA; accumulation is FP32.
//
// MFMA mapping (v_mfma_f32_16x16x32_bf16, layout checked on gfx950): a conv
// is an implicit GEMM out[voxel][co] = sum_k A[voxel][k] W[k][co] with
// k = tap * Cin + ci. A wavefront owns 32 output voxels x Cout; for each
// 32-wide k chunk every lane gathers 8 contiguous channels (16 bytes) of one
// neighbour voxel, zero outside the grid.
#include "fsx_device.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace fsx {

struct Grid3 {
  int nx, ny, nz;
  __host__ __device__ long long n() const { return (long long)nx * ny * nz; }
};

// (i, j, k) of node v, x fastest. The grids hold far fewer than 2^31 nodes,
// so two 32-bit divisions do (64-bit integer division is a long emulated
// sequence on the GPU).
__device__ __forceinline__ void node_ijk(long long v, const Grid3& g, int& i, int& j, int& k) {
  const unsigned u = (unsigned)v, nx = (unsigned)g.nx, nxy = nx * (unsigned)g.ny;
  const unsigned kk = u / nxy, rem = u - kk * nxy, jj = rem / nx;
  k = (int)kk; j = (int)jj; i = (int)(rem - jj * nx);
}

__device__ __forceinline__ __bf16 to_bf16(float f) {
  // round to nearest even (finite inputs)
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  unsigned short h = (unsigned short)(u >> 16);
  return __builtin_bit_cast(__bf16, h);
}
__device__ __forceinline__ float from_bf16(__bf16 b) {
  return __uint_as_float(((unsigned)__builtin_bit_cast(unsigned short, b)) << 16);
}

// Input coordinate of tap t (t = dx + 3 dy + 9 dz, offsets in {-1,0,1}) for
// output node (i,j,k). MODE 0: o + d; 1: o - d; 2: 2o + d (stride-2 down);
// 3: (o - d) / 2 when even (transpose of MODE 2). Returns -1 outside.
template <int MODE>
__device__ __forceinline__ long long tap_input(int i, int j, int k, int t, const Grid3& gi) {
  const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
  int a, b, c;
  if (MODE == 0) { a = i + dx; b = j + dy; c = k + dz; }
  else if (MODE == 1) { a = i - dx; b = j - dy; c = k - dz; }
  else if (MODE == 2) { a = 2 * i + dx; b = 2 * j + dy; c = 2 * k + dz; }
  else {
    a = i - dx; b = j - dy; c = k - dz;
    if ((a | b | c) & 1) return -1;
    a >>= 1; b >>= 1; c >>= 1;
  }
  if (a < 0 || b < 0 || c < 0 || a >= gi.nx || b >= gi.ny || c >= gi.nz) return -1;
  return (long long)(((unsigned)c * (unsigned)gi.ny + (unsigned)b) * (unsigned)gi.nx + (unsigned)a);
}

__host__ __device__ __forceinline__ bool on_boundary(long long v, const Grid3& g) {
  int i, j, k;
  node_ijk(v, g, i, j, k);
  return i == 0 || j == 0 || k == 0 || i == g.nx - 1 || j == g.ny - 1 || k == g.nz - 1;
}

// ------------------------------------------------------------ MFMA conv
// out32[v][co] = sum_k A[v][k] W[co][k] (+ res[v][co]); zero at the boundary
// nodes of go when mask is set; optional BF16 copy. W is [COUT][KP] BF16,
// KP = 27 * CIN rounded up to 32 (zero padded).
//
// The grids are small (the network's coarsest level is 9^3 for a 65^3
// pressure grid), so the kernel is organised for latency rather than
// throughput: a block owns 16 output voxels (one MFMA row tile) x COUT and
// splits the K loop over its KS wavefronts (wave w takes chunks w, w + KS,
// ...); the KS partial tiles are summed in LDS in wave order, so results do
// not depend on scheduling.
constexpr int kConvVox = 16;   // output voxels per block

template <int CIN>
constexpr int conv_ks() { return ((27 * CIN + 31) / 32) >= 48 ? 16 : ((27 * CIN + 31) / 32) >= 24 ? 8 : 4; }

template <int CIN, int COUT, int MODE>
__global__ __launch_bounds__(kWave * conv_ks<CIN>()) void k_conv_mfma(Grid3 gi, Grid3 go, const __bf16* __restrict__ in,
                                                                      const __bf16* __restrict__ W, float* __restrict__ out32,
                                                                      __bf16* __restrict__ out16, const float* __restrict__ res,
                                                                      int mask, const SolveScalars* __restrict__ sc) {
  if (sc && sc->stop) return;
  constexpr int KS = conv_ks<CIN>();
  constexpr int KP = (27 * CIN + 31) / 32 * 32;
  constexpr int NCH = KP / 32;
  constexpr int NCT = COUT / 16;
  constexpr int RS = COUT + 1;   // LDS row stride (odd: rows of one wave land in different banks)
  __shared__ float red[KS][kConvVox][RS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long nout = go.n();
  const long long vbase = (long long)blockIdx.x * kConvVox;
  const int grp = lane >> 4, col = lane & 15;
  const long long va = vbase + col;   // A-operand row of this lane
  const bool ov = va < nout;
  const long long vv = ov ? va : 0;
  int oi, oj, ok;
  node_ijk(vv, go, oi, oj, ok);
  f32x4 acc[NCT];
#pragma unroll
  for (int ct = 0; ct < NCT; ct++) acc[ct] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16x8 zero8 = {};
#pragma unroll 2
  for (int c = wave; c < NCH; c += KS) {
    const int kk0 = c * 32 + grp * 8;
    const int t = kk0 / CIN, ci0 = kk0 % CIN;
    long long src = -1;
    if (t < 27 && ov) src = tap_input<MODE>(oi, oj, ok, t, gi);
    const bf16x8 a = src >= 0 ? *reinterpret_cast<const bf16x8*>(in + src * CIN + ci0) : zero8;
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(W + (size_t)(ct * 16 + col) * KP + kk0);
      acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[ct], 0, 0, 0);
    }
  }
#pragma unroll
  for (int ct = 0; ct < NCT; ct++)
#pragma unroll
    for (int r = 0; r < 4; r++) red[wave][grp * 4 + r][ct * 16 + col] = acc[ct][r];
  __syncthreads();
  for (int e = threadIdx.x; e < kConvVox * COUT; e += KS * kWave) {
    const int row = e / COUT, co = e % COUT;
    const long long v = vbase + row;
    if (v >= nout) continue;
    float val = 0.f;
#pragma unroll
    for (int w = 0; w < KS; w++) val += red[w][row][co];
    const size_t o = (size_t)v * COUT + co;
    if (res) val += res[o];
    if (mask && on_boundary(v, go)) val = 0.0f;
    out32[o] = val;
    if (out16) out16[o] = to_bf16(val);
  }
}

// Transposed stride-2 convolution (the MODE 3 form of k_conv_mfma), tiled by
// output parity class: output node (i, j, k) receives only the taps with
// d == parity per axis (d = 0 at an even index, d = +-1 at an odd one), so a
// block owns 16 nodes of one class (blockIdx.y) and walks only that class's
// 1 to 8 taps instead of all 27. Same epilogue as k_conv_mfma.
//
// With hout (COUT = 16, the layer that produces g0): also the first layer's
// transpose taps h[v][t] = sum_c W0[c][t] g0[v][c] (t < 27, rows of 32), so
// the final pressure-grid layer only sums the <= 8 taps that reach a node.
template <int CIN, int COUT>
__global__ __launch_bounds__(kWave * 4) void k_conv_mfma_t(Grid3 gi, Grid3 go, const __bf16* __restrict__ in,
                                                           const __bf16* __restrict__ W, float* __restrict__ out32,
                                                           __bf16* __restrict__ out16, const float* __restrict__ res,
                                                           int mask, const float* __restrict__ w0, float* __restrict__ hout,
                                                           const SolveScalars* __restrict__ sc) {
  if (sc && sc->stop) return;
  constexpr int KS = 4;
  constexpr int KP = (27 * CIN + 31) / 32 * 32;
  constexpr int NCT = COUT / 16;
  constexpr int RS = COUT + 1;
  __shared__ float red[KS][kConvVox][RS];
  __shared__ float fin[kConvVox][RS];
  __shared__ float w0s[COUT == 16 ? 16 * 27 : 1];
  const int cls = blockIdx.y;
  const int pi = cls & 1, pj = (cls >> 1) & 1, pk = cls >> 2;
  // nodes of this class: i = 2 ic + pi, ic < cx, ...
  const int cx = (go.nx - pi + 1) >> 1, cy = (go.ny - pj + 1) >> 1, cz = (go.nz - pk + 1) >> 1;
  const long long ncls = (long long)cx * cy * cz;
  const long long vbase = (long long)blockIdx.x * kConvVox;
  if (vbase >= ncls) return;   // uniform over the block
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int grp = lane >> 4, col = lane & 15;
  const int ntx = 1 + pi, nty = 1 + pj, ntz = 1 + pk;
  const int nch = ntx * nty * ntz * CIN / 32;
  // this lane's A-operand row
  const long long va = vbase + col;
  const bool ov = va < ncls;
  int oi = 0, oj = 0, ok = 0;
  if (ov) {
    const unsigned u = (unsigned)va, cxy = (unsigned)cx * (unsigned)cy;
    const unsigned kc = u / cxy, rem = u - kc * cxy, jc = rem / (unsigned)cx, ic = rem - jc * (unsigned)cx;
    oi = 2 * (int)ic + pi; oj = 2 * (int)jc + pj; ok = 2 * (int)kc + pk;
  }
  f32x4 acc[NCT];
#pragma unroll
  for (int ct = 0; ct < NCT; ct++) acc[ct] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16x8 zero8 = {};
#pragma unroll 2
  for (int c = wave; c < nch; c += KS) {
    const int kl = c * 32 + grp * 8;           // k within this class's taps
    const int tl = kl / CIN, ci0 = kl % CIN;   // class-local tap, channel
    const int ax = tl % ntx, ay = (tl / ntx) % nty, az = tl / (ntx * nty);
    const int dx = pi ? (ax ? -1 : 1) : 0, dy = pj ? (ay ? -1 : 1) : 0, dz = pk ? (az ? -1 : 1) : 0;
    const int t = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
    bf16x8 a = zero8;
    if (ov) {
      const int sx = (oi - dx) >> 1, sy = (oj - dy) >> 1, sz = (ok - dz) >> 1;
      if (sx >= 0 && sy >= 0 && sz >= 0 && sx < gi.nx && sy < gi.ny && sz < gi.nz) {
        const unsigned src = ((unsigned)sz * (unsigned)gi.ny + (unsigned)sy) * (unsigned)gi.nx + (unsigned)sx;
        a = *reinterpret_cast<const bf16x8*>(in + (size_t)src * CIN + ci0);
      }
    }
    const int kg = t * CIN + ci0;              // k in the weight layout
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(W + (size_t)(ct * 16 + col) * KP + kg);
      acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[ct], 0, 0, 0);
    }
  }
#pragma unroll
  for (int ct = 0; ct < NCT; ct++)
#pragma unroll
    for (int r = 0; r < 4; r++) red[wave][grp * 4 + r][ct * 16 + col] = acc[ct][r];
  __syncthreads();
  for (int e = threadIdx.x; e < kConvVox * COUT; e += KS * kWave) {
    const int row = e / COUT, co = e % COUT;
    const long long vl = vbase + row;
    if (vl >= ncls) continue;
    const unsigned u = (unsigned)vl, cxy = (unsigned)cx * (unsigned)cy;
    const unsigned kc = u / cxy, rem = u - kc * cxy, jc = rem / (unsigned)cx, ic = rem - jc * (unsigned)cx;
    const long long v = ((long long)(2 * (int)kc + pk) * go.ny + (2 * (int)jc + pj)) * go.nx + (2 * (int)ic + pi);
    float val = 0.f;
#pragma unroll
    for (int w = 0; w < KS; w++) val += red[w][row][co];
    const size_t o = (size_t)v * COUT + co;
    if (res) val += res[o];
    if (mask && on_boundary(v, go)) val = 0.0f;
    out32[o] = val;
    if (out16) out16[o] = to_bf16(val);
    fin[row][co] = val;
  }
  if constexpr (COUT == 16) {
    if (hout) {
      for (int e = threadIdx.x; e < 16 * 27; e += KS * kWave) w0s[e] = w0[(e / 27) * 32 + e % 27];
      __syncthreads();
      for (int e = threadIdx.x; e < kConvVox * 32; e += KS * kWave) {
        const int row = e >> 5, t = e & 31;
        const long long vl = vbase + row;
        if (vl >= ncls || t >= 27) continue;
        const unsigned u = (unsigned)vl, cxy = (unsigned)cx * (unsigned)cy;
        const unsigned kc = u / cxy, rem = u - kc * cxy, jc = rem / (unsigned)cx, ic = rem - jc * (unsigned)cx;
        const long long v = ((long long)(2 * (int)kc + pk) * go.ny + (2 * (int)jc + pj)) * go.nx + (2 * (int)ic + pi);
        float h = 0.f;
#pragma unroll
        for (int c = 0; c < 16; c++) h += w0s[c * 27 + t] * fin[row][c];
        hout[(size_t)v * 32 + t] = h;
      }
    }
  }
}

// ------------------------------------------------------------ pressure-grid layers (VALU)
// f0[o][c] = B0(o) sum_t W0[c][t] x[2o + d_t]: the first layer, a stride-2
// conv from the pressure grid gf (1 channel) to G0 (16 channels), zero at
// the boundary nodes of G0; FP32 and a BF16 copy for the next (MFMA) layer.
// One thread per (node, channel); weights in LDS as [t][c], so the 16 lanes
// of a node read 16 consecutive words.
__global__ __launch_bounds__(kBlock) void k_conv0(Grid3 gf, Grid3 g0, const float* __restrict__ x,
                                                  const float* __restrict__ W0, float* __restrict__ out32,
                                                  __bf16* __restrict__ out16, const SolveScalars* __restrict__ sc) {
  if (sc && sc->stop) return;
  __shared__ float w[27 * 16];
  for (int i = threadIdx.x; i < 27 * 16; i += blockDim.x) w[i] = W0[(i & 15) * 32 + (i >> 4)];
  __syncthreads();
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g0.n() * 16) return;
  const int c = (int)(idx & 15);
  const long long v = idx >> 4;
  int i, j, k;
  node_ijk(v, g0, i, j, k);
  float acc = 0.f;
  if (!(i == 0 || j == 0 || k == 0 || i == g0.nx - 1 || j == g0.ny - 1 || k == g0.nz - 1)) {
    // an interior node's taps 2o + d stay inside gf, so all 27 loads are
    // issued together (no per-tap bounds test)
    const int sx = 1, sy = gf.nx, sz = gf.nx * gf.ny;
    const float* xc = x + ((2 * k) * sz + (2 * j) * sy + 2 * i);
    float xv[27];
#pragma unroll
    for (int t = 0; t < 27; t++) xv[t] = xc[(t / 9 - 1) * sz + ((t / 3) % 3 - 1) * sy + (t % 3 - 1) * sx];
#pragma unroll
    for (int t = 0; t < 27; t++) acc += w[t * 16 + c] * xv[t];
  }
  out32[idx] = acc;
  if (out16) out16[idx] = to_bf16(acc);
}

// The last layer, fused with the solver's use of it: for each pressure node
// f, y = gain sum_{o,t: 2o + d_t = f} h[o][t], with h[o][t] = sum_c W0[c][t]
// g0[o][c] from the previous layer (the transpose of k_conv0; g0 is already
// zero on the boundary of G0). With z set: z = dinv r + y and the r.z, z.q
// partials (slots 3, 4) of the flexible CG; otherwise y is stored. Grid:
// kRedGrid blocks of kWideBlock threads, each block a contiguous range of
// nodes at a stride of the block size (about two nodes per thread on 65^3).
// Measured per application at 64^3 (threads x nodes in flight per thread):
// 512 x 1 7.0 us, 512 x 2 7.5, 256 x 4 7.9, 256 x 2 8.2, 1024 x 2 10.3.
constexpr int kWideBlock = 512;

// With omega (the two-level form) z = omega dinv r + y and no partials.
__global__ __launch_bounds__(kWideBlock) void k_conv0T(Grid3 g0, Grid3 gf, const float* __restrict__ h, float gain,
                                                   float* __restrict__ y, const float* __restrict__ r,
                                                   const float* __restrict__ dinv, const float* __restrict__ q,
                                                   float* __restrict__ z, double* __restrict__ partials,
                                                   const SolveScalars* __restrict__ sc, float omega) {
  if (sc && sc->stop) return;   // sc is null on a start, which always runs
  double acc2[2] = {0.0, 0.0};
  // block b owns nodes [b per, (b + 1) per); a thread takes kNodes of them
  // at a stride of the block size, all loads issued before the sums
  constexpr int kNodes = 1;
  const long long n = gf.n(), per = (n + kRedGrid - 1) / kRedGrid;
  const long long lo = blockIdx.x * per, hi = lo + per < n ? lo + per : n;
  for (long long f0 = lo + threadIdx.x; f0 < hi; f0 += (long long)kNodes * kWideBlock) {
  float hv[kNodes][8];
  float rv[kNodes], dv[kNodes], qv[kNodes];
  unsigned mk[kNodes];
#pragma unroll
  for (int u = 0; u < kNodes; u++) {
    const long long fu = f0 + (long long)u * kWideBlock;
    const long long f = fu < hi ? fu : f0;   // an idle slot repeats the first node, unused
    int i, j, k;
    node_ijk(f, gf, i, j, k);
    // coarse parents o = (f - d) / 2 per axis: one (d = 0) at an even index,
    // two (d = +1, -1) at an odd one. Every node walks 2 x 2 x 2 slots; the
    // slots an even index does not have repeat its parent (a valid address)
    // with weight 0, so the loads carry no branches and issue together.
    const int dx0 = i & 1, dy0 = j & 1, dz0 = k & 1;   // slot 0: d = +1 when odd, else 0
    const unsigned ox[2] = {(unsigned)(i - dx0) >> 1, (unsigned)(i + dx0) >> 1};
    const unsigned oy[2] = {(unsigned)(j - dy0) >> 1, (unsigned)(j + dy0) >> 1};
    const unsigned oz[2] = {(unsigned)(k - dz0) >> 1, (unsigned)(k + dz0) >> 1};
    const int tx[2] = {dx0 + 1, 1 - dx0}, ty[2] = {dy0 + 1, 1 - dy0}, tz[2] = {dz0 + 1, 1 - dz0};
    mk[u] = (unsigned)dx0 | ((unsigned)dy0 << 1) | ((unsigned)dz0 << 2);
#pragma unroll
    for (int s8 = 0; s8 < 8; s8++) {
      const int a = s8 & 1, b = (s8 >> 1) & 1, c = s8 >> 2;
      const unsigned o = (oz[c] * (unsigned)g0.ny + oy[b]) * (unsigned)g0.nx + ox[a];
      hv[u][s8] = h[(size_t)o * 32 + (tx[a] + 3 * ty[b] + 9 * tz[c])];
    }
    if (z) { rv[u] = r[f]; dv[u] = dinv[f]; qv[u] = partials ? q[f] : 0.f; }
  }
#pragma unroll
  for (int u = 0; u < kNodes; u++) {
    const long long f = f0 + (long long)u * kWideBlock;
    if (f >= hi) continue;
    // slot s8 exists when each axis with a second parent (s8 bit set) is odd
    float acc = 0.f;
#pragma unroll
    for (int s8 = 0; s8 < 8; s8++) acc += ((s8 & ~mk[u]) ? 0.f : 1.f) * hv[u][s8];
    const float yv = gain * acc;
    if (z) {
      const float zi = omega * dv[u] * rv[u] + yv;
      z[f] = zi;
      acc2[0] += (double)rv[u] * (double)zi;
      acc2[1] += (double)zi * (double)qv[u];
    } else {
      y[f] = yv;
    }
  }
  }
  if (z && partials) write_partials<2, kWideBlock>(acc2, partials + 3 * kRedGrid);
}

// ------------------------------------------------------------ weight gradients
// part[chunk][co][t*CIN + ci] = sum over the chunk's output voxels v of
//   in[tap_input(v, t)][ci] * gout[v][co]
// (FP32 activations and gradients). grid = (kWChunks, 27).
constexpr int kWChunks = 64;
constexpr int kWBatch = 32;

template <int CIN, int COUT, int MODE>
__global__ __launch_bounds__(kBlock) void k_wgrad(Grid3 gi, Grid3 go, const float* __restrict__ in,
                                                  const float* __restrict__ gout, float* __restrict__ part) {
  constexpr int P = CIN * COUT;
  constexpr int PPT = (P + kBlock - 1) / kBlock;
  __shared__ float sa[kWBatch][CIN];
  __shared__ float sg[kWBatch][COUT];
  const int t = blockIdx.y;
  const long long nout = go.n();
  const long long per = (nout + kWChunks - 1) / kWChunks;
  const long long v0 = blockIdx.x * per, v1 = std::min(nout, v0 + per);
  float acc[PPT];
#pragma unroll
  for (int q = 0; q < PPT; q++) acc[q] = 0.f;
  for (long long vb = v0; vb < v1; vb += kWBatch) {
    __syncthreads();
    for (int e = threadIdx.x; e < kWBatch * CIN; e += kBlock) {
      const int b = e / CIN, ci = e % CIN;
      const long long v = vb + b;
      float val = 0.f;
      if (v < v1) {
        int i, j, k;
  node_ijk(v, go, i, j, k);
        const long long src = tap_input<MODE>(i, j, k, t, gi);
        if (src >= 0) val = in[src * CIN + ci];
      }
      sa[b][ci] = val;
    }
    for (int e = threadIdx.x; e < kWBatch * COUT; e += kBlock) {
      const int b = e / COUT, co = e % COUT;
      const long long v = vb + b;
      sg[b][co] = v < v1 ? gout[v * COUT + co] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PPT; q++) {
      const int pidx = threadIdx.x + q * kBlock;
      if (pidx < P) {
        const int co = pidx / CIN, ci = pidx % CIN;
        float s = 0.f;
#pragma unroll 8
        for (int b = 0; b < kWBatch; b++) s += sa[b][ci] * sg[b][co];
        acc[q] += s;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < PPT; q++) {
    const int pidx = threadIdx.x + q * kBlock;
    if (pidx < P) {
      const int co = pidx / CIN, ci = pidx % CIN;
      part[((size_t)blockIdx.x * COUT + co) * (27 * CIN) + t * CIN + ci] = acc[q];
    }
  }
}

// grad[co][k] (+)= sum_chunk part[chunk][co][k] in chunk order; grad has row stride KP
__global__ void k_wgrad_reduce(int cout, int k27, int kp, const float* __restrict__ part, float* __restrict__ grad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= cout * k27) return;
  const int co = e / k27, k = e % k27;
  float s = 0.f;
  for (int c = 0; c < kWChunks; c++) s += part[((size_t)c * cout + co) * k27 + k];
  grad[(size_t)co * kp + k] += s;
}

// ------------------------------------------------------------ optimizer, weights
// Adam on the FP32 master weights; refreshes the BF16 copy.
__global__ void k_adam(int n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, __bf16* __restrict__ w16, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi; v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (std::sqrt(vi / c2) + eps);
  w[i] = wi;
  if (w16) w16[i] = to_bf16(wi);
}

// BF16 copy and the transposed copy used by backward data:
//   WT[ci][t*COUT + co] = W[co][t*CIN + ci]
__global__ void k_weights_refresh(int cin, int cout, int kp, int kpt, const float* __restrict__ w32,
                                  __bf16* __restrict__ w16, __bf16* __restrict__ wt16, float* __restrict__ wr) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= cout * kp) return;
  const int co = e / kp, k = e % kp;
  const float v = k < 27 * cin ? w32[e] : 0.f;
  const __bf16 b = to_bf16(v);
  w16[e] = b;
  if (wr) wr[e] = from_bf16(b);   // rounded FP32 values for the VALU layers
  if (k < 27 * cin && wt16) {
    const int t = k / cin, ci = k % cin;
    wt16[(size_t)ci * kpt + t * cout + co] = b;
  }
}

__global__ void k_scale_f(long long n, float* __restrict__ a, float s) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) a[i] *= s;
}

// relative squared error of one sample and its gradient:
// L = sum (y - t)^2 / sum t^2 ; gy = 2 (y - t) / sum t^2   (tt = sum t^2 precomputed)
__global__ __launch_bounds__(kBlock) void k_loss(int n, const float* __restrict__ y, const float* __restrict__ tgt, double tt,
                                                 float* __restrict__ gy, double* __restrict__ partials) {
  double v[1] = {0.0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kRedGrid * kBlock) {
    const float d = y[i] - tgt[i];
    gy[i] = (float)(2.0 * d / tt);
    v[0] += (double)d * d;
  }
  write_partials<1>(v, partials);
}

// energy-norm loss of one sample: d = y - t (in gy on entry), q = Z d;
// L = d.q / tt, gy = 2 q / tt (gradient w.r.t. the network output y)
__global__ __launch_bounds__(kBlock) void k_eloss(int n, const float* __restrict__ q, double tt, float* __restrict__ gy,
                                                  double* __restrict__ partials) {
  double v[1] = {0.0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kRedGrid * kBlock) {
    const float d = gy[i];
    v[0] += (double)d * (double)q[i];
    gy[i] = (float)(2.0 * q[i] / tt);
  }
  write_partials<1>(v, partials);
}

// The two-level loss: with d = y - t (the coarse correction's error, y = C r1
// and t = e1 the error after pre-smoothing), the error left by the whole
// preconditioner is f = (I - S Z) d, S = omega diag(dinv). These kernels form
// f = d - S (Z d) in place of d, then f.(Z f) with ws = S (Z f), and last the
// gradient 2 (I - Z S) Z f / tt = 2 (Z f - Z ws) / tt of f.Z f / tt.
__global__ void k_smooth_sub_f(int n, float omega, const double* __restrict__ dinv, const float* __restrict__ zd,
                               float* __restrict__ d) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    d[i] = d[i] - omega * (float)dinv[i] * zd[i];
}

__global__ __launch_bounds__(kBlock) void k_eloss2a(int n, float omega, const double* __restrict__ dinv,
                                                    const float* __restrict__ f, const float* __restrict__ zf,
                                                    float* __restrict__ ws, double* __restrict__ partials) {
  double v[1] = {0.0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kRedGrid * kBlock) {
    const float q = zf[i];
    v[0] += (double)f[i] * (double)q;
    ws[i] = omega * (float)dinv[i] * q;
  }
  write_partials<1>(v, partials);
}

__global__ void k_eloss2b(int n, const float* __restrict__ zf, double tt, float* __restrict__ gy) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    gy[i] = (float)(2.0 * ((double)zf[i] - (double)gy[i]) / tt);
}

__global__ void k_sub_f(int n, const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ c) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) c[i] = a[i] - b[i];
}

__global__ __launch_bounds__(kBlock) void k_dot_f(int n, const float* __restrict__ a, const float* __restrict__ b,
                                                  double* __restrict__ partials) {
  double v[1] = {0.0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kRedGrid * kBlock) v[0] += (double)a[i] * (double)b[i];
  write_partials<1>(v, partials);
}

}  // namespace fsx

// ====================================================================== host side
using namespace fsx;

namespace {

struct Layer {
  int cin, cout, mode;      // mode 0 stride 1, 2 stride-2 down
  Grid3 gi, go;
  int kp, kpt;              // row strides of W [cout][kp] and WT [cin][kpt]
  size_t off;               // offset of this layer's weights in the flat FP32 parameter array
  __bf16* w16 = nullptr;
  __bf16* wt16 = nullptr;
  float* wr = nullptr;      // BF16-rounded FP32 copy (VALU layers)
};

// Activations of one application of J^T J: encoder values f0, f1, f2 and the
// transpose-pass values g1 = B1 (f1 + C2^T f2), g0 = B0 (f0 + C1^T g1). The
// transpose values are the gradients of <J x, G> with respect to the encoder
// outputs for G = (f0, f1, f2), which the weight gradient uses.
struct Acts {
  float *f0 = nullptr, *f1 = nullptr, *f2 = nullptr, *g1 = nullptr, *g0 = nullptr;
  __bf16 *f0b = nullptr, *f1b = nullptr, *f2b = nullptr, *g1b = nullptr;
};

}  // namespace

struct fsx_unet {
  Grid3 gf{};               // pressure grid
  Grid3 g0{}, g1{}, g2{};   // coarse levels
  std::vector<Layer> L;     // C0, C1, C2
  size_t nparam = 0;        // padded parameter count (rows of stride kp)
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  float *x = nullptr, *y = nullptr, *tgt = nullptr, *gy = nullptr, *zq = nullptr, *ws = nullptr;   // pressure grid
  Acts A, B;                // A: the preconditioner / the sample; B: the loss gradient (training)
  float* wpart = nullptr;
  float* h0 = nullptr;      // first layer's transpose taps on G0, 32 per node (27 used)
  // training set (normalised pairs, FP32)
  int ncap = 0, cap_max = 0, cap_limit = 0;   // captured, allocated, asked for
  double *cap_b = nullptr, *cap_x = nullptr;     // captured (R2, Pdot) pairs
  int nsamp = 0, samp_max = 0;
  float *s_in = nullptr, *s_tgt = nullptr;
  std::vector<double> s_tt;   // per-pair loss normaliser (error energy)
  int s_kind = 1;             // loss of the pairs: 1 energy (additive form), 2 two-level energy
  int adam_t = 0;
  float gain = 1.0f;        // fixed output gain of the network
  double* tmp_x = nullptr;  // truncated iterate (or drawn error) while preparing samples
  double* zero_x = nullptr; // zeros: the iterate of a drawn error's pair
};

namespace {

Grid3 coarse_of(Grid3 g) { return Grid3{(g.nx + 1) / 2, (g.ny + 1) / 2, (g.nz + 1) / 2}; }

unsigned blocks_for(long long n, int per) { return (unsigned)std::max(1LL, (n + per - 1) / per); }

template <int CIN, int COUT, int MODE>
void conv_launch(fsx_ctx* ctx, Grid3 gi, Grid3 go, const __bf16* in, const __bf16* W, float* out32, __bf16* out16,
                 const float* res, int mask, const float* w0, float* hout, const SolveScalars* sc) {
  if constexpr (MODE == 3) {
    // class (0,0,0) is the largest: ceil(n/2) nodes per axis
    const long long nmax = (long long)((go.nx + 1) / 2) * ((go.ny + 1) / 2) * ((go.nz + 1) / 2);
    hipLaunchKernelGGL((k_conv_mfma_t<CIN, COUT>), dim3(blocks_for(nmax, kConvVox), 8), dim3(kWave * 4), 0, ctx->stream, gi, go,
                       in, W, out32, out16, res, mask, w0, hout, sc);
  } else {
    hipLaunchKernelGGL((k_conv_mfma<CIN, COUT, MODE>), dim3(blocks_for(go.n(), kConvVox)), dim3(kWave * conv_ks<CIN>()), 0,
                       ctx->stream, gi, go, in, W, out32, out16, res, mask, sc);
  }
}

// dispatch on (cin, cout, mode): the encoder convolutions (mode 2, stride 2)
// and their transposes (mode 3, with WT); hout (the 32 -> 16 transpose only)
// also receives the first layer's transpose taps (k_conv_mfma_t)
int conv(fsx_ctx* ctx, int cin, int cout, int mode, Grid3 gi, Grid3 go, const __bf16* in, const __bf16* W, float* out32,
         __bf16* out16, const float* res, int mask, const SolveScalars* sc, const float* w0 = nullptr,
         float* hout = nullptr) {
  if (hout && !(cin == 32 && cout == 16 && mode == 3)) return fail(ctx, FSX_EINVAL, "conv: taps output on another layer");
#define CONV_CASE(A, B, M) \
  if (cin == A && cout == B && mode == M) { conv_launch<A, B, M>(ctx, gi, go, in, W, out32, out16, res, mask, w0, hout, sc); return FSX_OK; }
  CONV_CASE(16, 32, 2) CONV_CASE(32, 64, 2) CONV_CASE(64, 32, 3) CONV_CASE(32, 16, 3)
#undef CONV_CASE
  return fail(ctx, FSX_EINVAL, "conv shape not instantiated");
}

int wgrad(fsx_ctx* ctx, int cin, int cout, int mode, Grid3 gi, Grid3 go, const float* in, const float* gout, float* part) {
#define WG_CASE(A, B, M) \
  if (cin == A && cout == B && mode == M) { \
    hipLaunchKernelGGL((k_wgrad<A, B, M>), dim3(kWChunks, 27), dim3(kBlock), 0, ctx->stream, gi, go, in, gout, part); \
    return FSX_OK; }
  WG_CASE(1, 16, 2) WG_CASE(16, 32, 2) WG_CASE(32, 64, 2)
#undef WG_CASE
  return fail(ctx, FSX_EINVAL, "wgrad shape not instantiated");
}

int refresh_weights(fsx_ctx* ctx, fsx_unet* u) {
  for (Layer& l : u->L) {
    const int n = l.cout * l.kp;
    hipLaunchKernelGGL(k_weights_refresh, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, l.cin, l.cout, l.kp, l.kpt,
                       u->w + l.off, l.w16, l.wt16, l.wr);
  }
  return launch_check(ctx, "k_weights_refresh");
}

template <class T>
int ualloc(fsx_ctx* ctx, T** p, size_t n) { return alloc(ctx, p, n); }

// One application on activation set S, x FP32 on gf. Output modes:
// kOutY: y = gain J^T J x on gf; kOutZ: the solver's z = dinv x + gain J^T J x
// with its r.z / z.q partials (fused into the last layer); kOutNone: stop
// after the transpose pass (training only needs its values).
enum { kOutNone = 0, kOutY = 1, kOutZ = 2 };

int apply_net(fsx_ctx* ctx, fsx_unet* u, Acts& S, const float* x, float* y, const SolveScalars* sc, int out = kOutY) {
  const Layer* L = u->L.data();
  const long long n0 = u->g0.n();
  hipLaunchKernelGGL(k_conv0, dim3(blocks_for(n0 * 16, kBlock)), dim3(kBlock), 0, ctx->stream, u->gf, u->g0, x, L[0].wr, S.f0,
                     S.f0b, sc);
  FSX_TRY(conv(ctx, 16, 32, 2, u->g0, u->g1, S.f0b, L[1].w16, S.f1, S.f1b, nullptr, 1, sc));
  FSX_TRY(conv(ctx, 32, 64, 2, u->g1, u->g2, S.f1b, L[2].w16, S.f2, S.f2b, nullptr, 1, sc));
  // transpose pass
  FSX_TRY(conv(ctx, 64, 32, 3, u->g2, u->g1, S.f2b, L[2].wt16, S.g1, S.g1b, S.f1, 1, sc));
  FSX_TRY(conv(ctx, 32, 16, 3, u->g1, u->g0, S.g1b, L[1].wt16, S.g0, nullptr, S.f0, 1, sc, L[0].wr,
               out == kOutNone ? nullptr : u->h0));
  if (out == kOutY)
    hipLaunchKernelGGL(k_conv0T, dim3(kRedGrid), dim3(kWideBlock), 0, ctx->stream, u->g0, u->gf, u->h0, u->gain, y, nullptr,
                       nullptr, nullptr, nullptr, nullptr, sc, 1.0f);
  else if (out == kOutZ)
    hipLaunchKernelGGL(k_conv0T, dim3(kRedGrid), dim3(kWideBlock), 0, ctx->stream, u->g0, u->gf, u->h0, u->gain, nullptr,
                       x, ctx->Zdinvf, ctx->fq, ctx->fz, ctx->partials, sc, 1.0f);
  return launch_check(ctx, "network application");
}

int forward(fsx_ctx* ctx, fsx_unet* u, const float* x, float* y) { return apply_net(ctx, u, u->A, x, y, nullptr); }

// The two-level form's coarse correction: the network on the pre-smoothed
// residual r1, its output added to the pre-smoothing of r:
// z2 = omega dinv r + gain J^T J r1.
int apply_net_mult(fsx_ctx* ctx, fsx_unet* u, const float* r1, const float* r, float* z2, const SolveScalars* sc) {
  const Layer* L = u->L.data();
  Acts& S = u->A;
  const long long n0 = u->g0.n();
  hipLaunchKernelGGL(k_conv0, dim3(blocks_for(n0 * 16, kBlock)), dim3(kBlock), 0, ctx->stream, u->gf, u->g0, r1, L[0].wr,
                     S.f0, S.f0b, sc);
  FSX_TRY(conv(ctx, 16, 32, 2, u->g0, u->g1, S.f0b, L[1].w16, S.f1, S.f1b, nullptr, 1, sc));
  FSX_TRY(conv(ctx, 32, 64, 2, u->g1, u->g2, S.f1b, L[2].w16, S.f2, S.f2b, nullptr, 1, sc));
  FSX_TRY(conv(ctx, 64, 32, 3, u->g2, u->g1, S.f2b, L[2].wt16, S.g1, S.g1b, S.f1, 1, sc));
  FSX_TRY(conv(ctx, 32, 16, 3, u->g1, u->g0, S.g1b, L[1].wt16, S.g0, nullptr, S.f0, 1, sc, L[0].wr, u->h0));
  hipLaunchKernelGGL(k_conv0T, dim3(kRedGrid), dim3(kWideBlock), 0, ctx->stream, u->g0, u->gf, u->h0, u->gain, nullptr, r,
                     ctx->Zdinvf, nullptr, z2, nullptr, sc, ctx->learned_omega);
  return launch_check(ctx, "network application (two-level)");
}

// Parameter gradient of <gy, y> after forward() of x on set A (accumulates
// into u->g): with y = gain J^T J x, <gy, y> = gain <J gy, J x>, whose
// gradient is the encoder backpropagation at input x with output gradients
// J (gain gy) plus the one at input gain gy with output gradients J x. Both
// sets of output-side gradients are transpose-pass values: set A holds those
// of x, set B (the network applied to gain gy) those of gain gy.
int backward(fsx_ctx* ctx, fsx_unet* u, const float* x) {
  hipLaunchKernelGGL(k_scale_f, dim3(256), dim3(kBlock), 0, ctx->stream, (long long)u->gf.n(), u->gy, u->gain);
  FSX_TRY(apply_net(ctx, u, u->B, u->gy, nullptr, nullptr, kOutNone));
  const Layer* L = u->L.data();
  auto wg = [&](int li, const float* inA, const float* gB, const float* inB, const float* gA) -> int {
    const Layer& l = L[li];
    const int n = l.cout * 27 * l.cin;
    FSX_TRY(wgrad(ctx, l.cin, l.cout, l.mode, l.gi, l.go, inA, gB, u->wpart));
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, l.cout, 27 * l.cin, l.kp,
                       u->wpart, u->g + l.off);
    FSX_TRY(wgrad(ctx, l.cin, l.cout, l.mode, l.gi, l.go, inB, gA, u->wpart));
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, l.cout, 27 * l.cin, l.kp,
                       u->wpart, u->g + l.off);
    return launch_check(ctx, "wgrad");
  };
  Acts &A = u->A, &B = u->B;
  FSX_TRY(wg(2, A.f1, B.f2, B.f1, A.f2));
  FSX_TRY(wg(1, A.f0, B.g1, B.f0, A.g1));
  FSX_TRY(wg(0, x, B.g0, u->gy, A.g0));
  return FSX_OK;
}

int alloc_acts(fsx_ctx* ctx, fsx_unet* u, Acts& S) {
  const size_t n0 = u->g0.n(), n1 = u->g1.n(), n2 = u->g2.n();
  FSX_TRY(ualloc(ctx, &S.f0, n0 * 16)); FSX_TRY(ualloc(ctx, &S.f0b, n0 * 16)); FSX_TRY(ualloc(ctx, &S.g0, n0 * 16));
  FSX_TRY(ualloc(ctx, &S.f1, n1 * 32)); FSX_TRY(ualloc(ctx, &S.f1b, n1 * 32));
  FSX_TRY(ualloc(ctx, &S.g1, n1 * 32)); FSX_TRY(ualloc(ctx, &S.g1b, n1 * 32));
  FSX_TRY(ualloc(ctx, &S.f2, n2 * 64)); FSX_TRY(ualloc(ctx, &S.f2b, n2 * 64));
  return FSX_OK;
}

// Initial weights: channel 0 of every layer carries the additive multilevel
// preconditioner (C0: full weighting; C1, C2: full weighting scaled by
// 1/sqrt(2)), so J^T J = sum_l 2^-(l-1) P_l P_l^T over the three coarse
// levels; the other entries are small seeded noise (1/10 of a fan-in-scaled
// normal) so that training can use them; gain = median(dinv) / 4.
void init_weights(const std::vector<Layer>& L, size_t nparam, unsigned long long seed, std::vector<float>& w) {
  w.assign(nparam, 0.0f);
  std::mt19937_64 rng(seed);
  for (size_t li = 0; li < L.size(); li++) {
    const Layer& l = L[li];
    std::normal_distribution<float> nd(0.0f, 0.1f * (float)std::sqrt(1.0 / (27.0 * l.cin)));
    for (int co = 0; co < l.cout; co++)
      for (int k = 0; k < 27 * l.cin; k++) w[l.off + (size_t)co * l.kp + k] = nd(rng);
    const float a = li == 0 ? 1.0f : (float)std::sqrt(0.5);
    for (int t = 0; t < 27; t++) {
      const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
      w[l.off + (size_t)t * l.cin] = a * (dx ? 0.5f : 1.0f) * (dy ? 0.5f : 1.0f) * (dz ? 0.5f : 1.0f);   // co = ci = 0
    }
  }
}

int setup_unet(fsx_ctx* ctx, int nx, int ny, int nz, unsigned long long seed, fsx_unet** out) {
  if ((nx - 1) % 8 || (ny - 1) % 8 || (nz - 1) % 8 || nx < 9 || ny < 9 || nz < 9)
    return fail(ctx, FSX_EINVAL, "pressure grid needs 8m+1 nodes (m >= 1) per axis for three node-aligned coarsenings");
  if ((long long)nx * ny * nz != ctx->NNp) return fail(ctx, FSX_EINVAL, "U-net grid size != NNp");
  auto u = std::make_unique<fsx_unet>();
  u->gf = Grid3{nx, ny, nz};
  u->g0 = coarse_of(u->gf);
  u->g1 = coarse_of(u->g0);
  u->g2 = coarse_of(u->g1);
  struct Spec { int cin, cout, mode; Grid3 gi, go; };
  const Spec spec[3] = {{1, 16, 2, u->gf, u->g0}, {16, 32, 2, u->g0, u->g1}, {32, 64, 2, u->g1, u->g2}};
  size_t off = 0;
  for (const Spec& sp : spec) {
    Layer l;
    l.cin = sp.cin; l.cout = sp.cout; l.mode = sp.mode; l.gi = sp.gi; l.go = sp.go;
    l.kp = (27 * sp.cin + 31) / 32 * 32;
    l.kpt = (27 * sp.cout + 31) / 32 * 32;
    l.off = off;
    off += (size_t)l.cout * l.kp;
    u->L.push_back(l);
  }
  u->nparam = off;
  FSX_TRY(ualloc(ctx, &u->w, off)); FSX_TRY(ualloc(ctx, &u->g, off));
  FSX_TRY(ualloc(ctx, &u->m, off)); FSX_TRY(ualloc(ctx, &u->v, off));
  for (Layer& l : u->L) {
    FSX_TRY(ualloc(ctx, &l.w16, (size_t)l.cout * l.kp));
    FSX_TRY(ualloc(ctx, &l.wt16, (size_t)l.cin * l.kpt + 8));
    FSX_TRY(ualloc(ctx, &l.wr, (size_t)l.cout * l.kp));
    FSX_HIP(hipMemsetAsync(l.wt16, 0, ((size_t)l.cin * l.kpt + 8) * sizeof(__bf16), ctx->stream));
  }
  std::vector<float> w;
  init_weights(u->L, off, seed, w);
  FSX_HIP(hipMemcpy(u->w, w.data(), off * sizeof(float), hipMemcpyHostToDevice));
  FSX_HIP(hipMemsetAsync(u->m, 0, off * sizeof(float), ctx->stream));
  FSX_HIP(hipMemsetAsync(u->v, 0, off * sizeof(float), ctx->stream));
  FSX_HIP(hipMemsetAsync(u->g, 0, off * sizeof(float), ctx->stream));
  // gain: a quarter of the median Jacobi scaling
  {
    std::vector<double> dinv(ctx->NNp);
    FSX_HIP(hipMemcpy(dinv.data(), ctx->Zdinv, dinv.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::nth_element(dinv.begin(), dinv.begin() + dinv.size() / 2, dinv.end());
    u->gain = (float)(0.25 * dinv[dinv.size() / 2]);
  }
  const size_t nf = u->gf.n();
  for (float** p : {&u->x, &u->y, &u->tgt, &u->gy, &u->zq, &u->ws}) FSX_TRY(ualloc(ctx, p, nf));
  FSX_TRY(alloc_acts(ctx, u.get(), u->A));
  FSX_TRY(alloc_acts(ctx, u.get(), u->B));
  FSX_TRY(ualloc(ctx, &u->wpart, (size_t)kWChunks * 64 * 27 * 64));
  FSX_TRY(ualloc(ctx, &u->h0, (size_t)u->g0.n() * 32));
  FSX_TRY(refresh_weights(ctx, u.get()));
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  *out = u.release();
  return FSX_OK;
}

}  // namespace

// ====================================================================== learned solve
// Flexible CG in the FCG(1) form (Notay 2000): alpha = r.p / p.Zp, the new
// direction Z-orthogonalised against the previous one, beta = -z.Zp / p.Zp.
// Both stay exact when the preconditioner varies from one application to the
// next (BF16 rounding, a scale taken from the residual) and when a reliable
// update replaces r. p.Zp comes from the SpMV's own partial sums (one per
// row-class chunk); r.p is carried as a scalar: r.p_new = r.z + beta r.p_old
// (bilinear, so exact up to rounding), with r.p_old summed by the update
// kernel and, after a reliable update replaces r, by k_fpcg_load_r.
// Partial-sum slots (kRedGrid doubles each): 1 r.p_old, 2 r.r
// (k_fpcg_update, k_fpcg_load_r); 3 r.z, 4 z.q (k_conv0T).
namespace fsx {

constexpr int kSlotRP = 1;

// alpha = r.p / p.q (0 when p is not a descent direction, which
// k_fpcg_direction then reports); x += alpha p, r -= alpha q; partials of
// r.p (the direction before its update) and r.r
__global__ __launch_bounds__(kBlock) void k_fpcg_update(int j, int n, float* __restrict__ x, float* __restrict__ r,
                                                        const float* __restrict__ p, const float* __restrict__ q,
                                                        double* __restrict__ partials, const double* __restrict__ pq,
                                                        int npq, SolveScalars* sc) {
  const int stop = sc->stop;
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->active_v[j & 1] = stop ? 0 : 1;
  if (stop) return;
  const double pqv = sum_parts(pq, npq), rp = sc->rp_v[j & 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->pq_v[j & 1] = pqv;
  const float a = (pqv > 0.0 && rp > 0.0) ? (float)(rp / pqv) : 0.0f;
  double v[2] = {0.0, 0.0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kRedGrid * kBlock) {
    const float pi = p[i];
    x[i] += a * pi;
    const float ri = r[i] - a * q[i];
    r[i] = ri;
    v[0] += (double)ri * (double)pi;
    v[1] += (double)ri * (double)ri;
  }
  write_partials<2>(v, partials + kSlotRP * kRedGrid);
}

// p = z - (z.q / p.q) p and the next r.p; stop tests (block 0): 3 NaN/Inf,
// 6 loss of descent (p.Zp, r.p or r.z not positive), 1 converged (FP32
// residual; the FP64 check decides), 5 pause for a reliable update, 2
// iteration cap.
__global__ __launch_bounds__(kBlock) void k_fpcg_direction(int j, int n, const float* __restrict__ z, float* __restrict__ p,
                                                           const double* __restrict__ partials, double delta2,
                                                           SolveScalars* sc) {
  if (!sc->active_v[j & 1]) return;
  double s[4];
  read_partials<4>(partials + kSlotRP * kRedGrid, s);
  const double rpo = s[0], rr = s[1], rz = s[2], zq = s[3];
  const double pq = sc->pq_v[j & 1], rp = sc->rp_v[j & 1];
  const bool bad = !(pq == pq) || !(rp == rp) || !(rr == rr) || !(rz == rz) || !(zq == zq) || !(rpo == rpo);
  const bool lost = !(pq > 0.0) || !(rp > 0.0) || !(rz > 0.0);
  const float be = (bad || lost) ? 0.f : (float)(-zq / pq);
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kRedGrid * kBlock) p[i] = z[i] + be * p[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->rp_v[(j + 1) & 1] = rz + (double)be * rpo;
    const int it = sc->iters + 1;
    sc->iters = it;
    sc->rr = rr;
    if (bad) sc->stop = 3;
    else if (lost) sc->stop = 6;
    else if (rr <= sc->rtol2bb) sc->stop = 1;
    else if (rr <= delta2 * sc->rr_anchor) sc->stop = 5;
    else if (it >= sc->maxit) sc->stop = 2;
  }
}

// first direction p = z; r.p = r.z (slot 3) enters iteration 0
__global__ __launch_bounds__(kBlock) void k_fpcg_start(int n, const float* __restrict__ z, float* __restrict__ p,
                                                       const double* __restrict__ partials, SolveScalars* sc) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) p[i] = z[i];
  if (blockIdx.x == 0) {
    double rz[1];
    read_partials<1>(partials + 3 * kRedGrid, rz);
    if (threadIdx.x == 0) sc->rp_v[0] = rz[0];
  }
}

// r.p of a replaced residual (slot 1, from k_fpcg_load_r) enters iteration j
__global__ __launch_bounds__(kBlock) void k_fpcg_set_rp(int j, const double* __restrict__ partials, SolveScalars* sc) {
  double rp[1];
  read_partials<1>(partials + kSlotRP * kRedGrid, rp);
  if (threadIdx.x == 0) sc->rp_v[j & 1] = rp[0];
}

// r = float(rh), x_lo = 0; partials of r.p (the direction kept) and r.r
__global__ __launch_bounds__(kBlock) void k_fpcg_load_r(int n, const double* __restrict__ rh, float* __restrict__ r,
                                                        float* __restrict__ x, const float* __restrict__ p,
                                                        double* __restrict__ partials) {
  double v[2] = {0.0, 0.0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kRedGrid * kBlock) {
    const float ri = (float)rh[i];
    r[i] = ri;
    x[i] = 0.0f;
    v[0] += (double)ri * (double)p[i];
    v[1] += (double)ri * (double)ri;
  }
  write_partials<2>(v, partials + kSlotRP * kRedGrid);
}

__global__ void k_zero_f(int n, float* __restrict__ a) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = 0.0f;
}

__global__ __launch_bounds__(kBlock) void k_sumsq_u(int n, const double* __restrict__ a, double* __restrict__ partials) {
  double v[1] = {0.0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kRedGrid * kBlock) v[0] += a[i] * a[i];
  write_partials<1>(v, partials);
}

__global__ __launch_bounds__(kBlock) void k_learned_init(const double* __restrict__ partials, SolveScalars* sc, double rtol,
                                                         int maxit) {
  double bb[1];
  read_partials<1>(partials, bb);
  if (threadIdx.x == 0) {
    sc->bb = bb[0]; sc->rr = bb[0]; sc->rr_anchor = bb[0]; sc->rr_hi = bb[0];
    sc->rtol2bb = rtol * rtol * bb[0];
    sc->iters = 0; sc->maxit = maxit;
    sc->stop = bb[0] == 0.0 ? 1 : 0;
    sc->active_v[0] = 0; sc->active_v[1] = 0;
  }
}

// training sample from (b, x*, x_k): r = b - Z x_k (in rh, |r|^2 partials in
// slot 0), s = rms(r); in = r / s; t = (x* - x_k - dinv r) / s; slot 1 gets
// the error energy (x* - x_k).r / s^2, the loss normaliser of the sample
__global__ __launch_bounds__(kBlock) void k_make_sample(int n, const double* __restrict__ rh, const double* __restrict__ xs,
                                                        const double* __restrict__ xk, const double* __restrict__ dinv,
                                                        double* __restrict__ partials, float* __restrict__ in,
                                                        float* __restrict__ tgt) {
  double rr[1];
  read_partials<1>(partials, rr);
  const double inv_s = rr[0] > 0.0 ? 1.0 / std::sqrt(rr[0] / n) : 0.0;
  double v[1] = {0.0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kRedGrid * kBlock) {
    const double r = rh[i], e = xs[i] - xk[i];
    in[i] = (float)(r * inv_s);
    tgt[i] = (float)((e - dinv[i] * r) * inv_s);
    v[0] += e * r * inv_s * inv_s;
  }
  write_partials<1>(v, partials + kRedGrid);
}

// the two-level form's sample: with t = omega dinv r (rh) and r1 = r - Z t,
// in = r1 / s, t' = (x* - x_k - t) / s: the network, applied to the
// pre-smoothed residual, is trained towards the error that the smoothing
// leaves; slot 1 gets the error energy (x* - x_k).r / s^2 (slot 0 holds |r|^2)
__global__ __launch_bounds__(kBlock) void k_make_sample2(int n, const double* __restrict__ rh, const double* __restrict__ r1,
                                                         const double* __restrict__ t, const double* __restrict__ xs,
                                                         const double* __restrict__ xk, double* __restrict__ partials,
                                                         float* __restrict__ in, float* __restrict__ tgt) {
  double rr[1];
  read_partials<1>(partials, rr);
  const double inv_s = rr[0] > 0.0 ? 1.0 / std::sqrt(rr[0] / n) : 0.0;
  double v[1] = {0.0};
#pragma unroll 4   // independent elements: their loads in flight together
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kRedGrid * kBlock) {
    const double e = xs[i] - xk[i];
    in[i] = (float)(r1[i] * inv_s);
    tgt[i] = (float)((e - t[i]) * inv_s);
    v[0] += e * rh[i] * inv_s * inv_s;
  }
  write_partials<1>(v, partials + kRedGrid);
}

__global__ void k_scaled_dinv(int n, double w, const double* __restrict__ dinv, const double* __restrict__ r,
                              double* __restrict__ t) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) t[i] = w * dinv[i] * r[i];
}

// e_i ~ N(0, 1), seeded by (seed, draw) and the node index (SplitMix64
// finaliser for two uniforms, Box-Muller), 0 at the pinned pressure node
__device__ inline unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

__global__ void k_randn(int n, unsigned long long seed, int draw, int pin, double* __restrict__ e) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const unsigned long long k = mix64(seed + 0x9e3779b97f4a7c15ULL * (unsigned long long)(draw + 1)) + 2ULL * i;
    const double u1 = ((mix64(k) >> 11) + 1) * 0x1.0p-53, u2 = (mix64(k + 1) >> 11) * 0x1.0p-53;
    e[i] = i == pin ? 0.0 : std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
}

// e <- e - omega dinv (Z e), one damped Jacobi sweep on the error
__global__ void k_jacobi_err(int n, double omega, const double* __restrict__ dinv, const double* __restrict__ ze,
                             double* __restrict__ e) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) e[i] -= omega * dinv[i] * ze[i];
}

__global__ void k_stop_clear(SolveScalars* sc) {
  sc->stop = 0;
  sc->pq_prev = 0.0;
}

}  // namespace fsx

namespace {

// z = dinv r + gain J^T J r for the current FP32 residual r, with the r.z
// and z.q partials. J^T J is linear, so this equals gain s J^T J (r / s) for
// any s; the residual is not rescaled here (BF16 has the FP32 exponent
// range, so a residual reduced by 1e-8 stays far from underflow), while the
// training pairs are normalised by rms(r).
int precondition(fsx_ctx* ctx, fsx_unet* u, const SolveScalars* sc) {
  FSX_TRY(apply_net(ctx, u, u->A, ctx->fr, nullptr, sc, kOutZ));
  return launch_check(ctx, "precondition");
}

}  // namespace

namespace fsx {

int launch_pcg_spmv_f32(fsx_ctx* ctx, const float* p, float* q, const SolveScalars* sc);
int launch_pcg_spmv_pq_f32(fsx_ctx* ctx, const float* p, float* q, const SolveScalars* sc, double* pq, int* npq);
int launch_fold_residual(fsx_ctx* ctx, double* x, const double* b, int* stop_out, double* rr_hi_out);
int solve_mixed_entry(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres);

// Flexible PCG preconditioned by Jacobi plus the U-net correction, in FP32
// with FP64 reliable updates exactly like the mixed solver (b - Z x in FP64
// decides convergence). If the flexible recurrence loses descent, meets the
// iteration cap or NaN, the remaining FP64 residual is handed to the mixed
// Jacobi-PCG, which then also decides the final residual. truncate_at > 0
// instead stops after that many flexible iterations (or at a loss of
// descent) and returns the iterate without the fallback (sample generation).
int solve_learned(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres, int truncate_at) {
  fsx_unet* u = ctx->unet;
  const int n = ctx->NNp;
  const double delta = 1e-2;
  const int max_learned = truncate_at > 0 ? truncate_at : std::min(ctx->maxit, ctx->learned_maxit);
  FSX_TRY(launch_fill(ctx, x, 0.0, n));
  FSX_TRY(launch_copy(ctx, ctx->wh, b, n));
  hipLaunchKernelGGL(k_sumsq_u, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, b, ctx->partials);
  // (b = 0 sets the stop flag here: every iteration is then a no-op and the
  // first FP64 check finds the zero residual; no read-back before the start)
  hipLaunchKernelGGL(k_learned_init, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->partials, ctx->sc, ctx->rtol, max_learned);
  FSX_TRY(ensure_fpq(ctx));
  // start: r = b, z = M r, p = z, r.p = r.z
  hipLaunchKernelGGL(k_fpcg_load_r, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, ctx->wh, ctx->fr, ctx->fx, ctx->fp,
                     ctx->partials);
  hipLaunchKernelGGL(k_zero_f, dim3(256), dim3(kBlock), 0, ctx->stream, n, ctx->fq);
  FSX_TRY(precondition(ctx, u, nullptr));
  hipLaunchKernelGGL(k_fpcg_start, dim3(256), dim3(kBlock), 0, ctx->stream, n, ctx->fz, ctx->fp, ctx->partials, ctx->sc);
  FSX_TRY(launch_check(ctx, "fpcg start"));
  int j = 0;
  const int seg = std::min(max_learned, segment_guess(ctx->last_learned_iters, delta, ctx->rtol, 2));
  int guess = seg;
  int fallback = 0;
  auto issue = [&](int jj) -> int {
    int npq = 0;
    FSX_TRY(launch_pcg_spmv_pq_f32(ctx, ctx->fp, ctx->fq, ctx->sc, ctx->fpq, &npq));
    hipLaunchKernelGGL(k_fpcg_update, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, jj, n, ctx->fx, ctx->fr, ctx->fp,
                       ctx->fq, ctx->partials, ctx->fpq, npq, ctx->sc);
    FSX_TRY(precondition(ctx, u, ctx->sc));
    hipLaunchKernelGGL(k_fpcg_direction, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, jj, n, ctx->fz, ctx->fp,
                       ctx->partials, delta * delta, ctx->sc);
    return FSX_OK;
  };
  // the captured launches bake in the network's output gain
  unsigned gbits = 0;
  std::memcpy(&gbits, &u->gain, sizeof gbits);
  const unsigned long long key = ((unsigned long long)gbits << 1) | 1ull;
  for (;;) {
    int issued = 0;
    for (;;) {
      int done = 0;
      FSX_TRY(run_iterations(ctx, ctx->graph_learned, key, j, guess, &done, issue));
      j += done;
      issued += done;
      FSX_TRY(read_scalars(ctx));
      if (ctx->h_sc->stop) break;
      if (issued > max_learned + 64) return fail(ctx, FSX_ESTATE, "FPCG did not terminate");
      guess = 2;
    }
    const int stop = ctx->h_sc->stop;
    j = ctx->h_sc->iters;
    // reliable update: fold x_lo into x, FP64 residual wh = b - Z x, FP64 stop test
    int st = 0;
    double rr_hi = 0;
    FSX_TRY(launch_fold_residual(ctx, x, b, &st, &rr_hi));
    if (st == 3 || stop == 3) return fail(ctx, FSX_ENUMERIC, "learned solve met NaN/Inf");
    if (st == 1) break;                                     // converged in FP64
    if (stop == 6 || stop == 2 || st == 2) { fallback = stop == 6 ? 6 : 2; break; }
    // replace r by the FP64 residual and keep p: the next alpha uses r.p
    hipLaunchKernelGGL(k_fpcg_load_r, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, ctx->wh, ctx->fr, ctx->fx, ctx->fp,
                       ctx->partials);
    hipLaunchKernelGGL(k_fpcg_set_rp, dim3(1), dim3(kBlock), 0, ctx->stream, j, ctx->partials, ctx->sc);
    guess = seg;
  }
  const int learned_iters = ctx->h_sc->iters;
  const double bb = ctx->h_sc->bb, rh = ctx->h_sc->rr_hi;
  if (truncate_at > 0) {
    *iters = learned_iters;
    *relres = bb > 0 ? std::sqrt(rh / bb) : 0.0;
    return FSX_OK;
  }
  ctx->last_learned_iters = learned_iters;
  ctx->last_fallback = fallback;
  if (fallback) {
    // Jacobi-PCG (mixed) on the remaining FP64 residual wh: x += Z^-1 wh,
    // to the same absolute threshold rtol ||b||
    int it2 = 0;
    double rel2 = 0;
    const double saved_rtol = ctx->rtol;
    ctx->rtol = rh > 0 ? std::min(1.0, saved_rtol * std::sqrt(bb / rh)) : saved_rtol;
    FSX_TRY(launch_copy(ctx, ctx->wr, ctx->wh, n));        // the mixed solver overwrites wh
    const int rc = solve_mixed_entry(ctx, ctx->wr, ctx->wz, &it2, &rel2);
    ctx->rtol = saved_rtol;
    if (rc != FSX_OK && rc != FSX_ENOCONV) return rc;
    FSX_TRY(launch_add(ctx, x, ctx->wz, n));
    int st = 0;
    double rr2 = 0;
    FSX_TRY(launch_residual_only(ctx, x, b, &st, &rr2));
    *iters = learned_iters + it2;
    *relres = bb > 0 ? std::sqrt(rr2 / bb) : 0.0;
    ctx->last_iters = *iters;
    if (st != 1) { ctx->err = "learned solve and Jacobi fallback reached maxit"; return FSX_ENOCONV; }
    return FSX_OK;
  }
  *iters = learned_iters;
  *relres = bb > 0 ? std::sqrt(rh / bb) : 0.0;
  ctx->last_iters = learned_iters;
  return FSX_OK;
}

int solve_fp64_entry(fsx_ctx* ctx, const double* b, double* x, int* iters, double* relres);

// FP32 |b|^2 and |r0|^2 (slots 0 and 1) -> the scalars of a tiled solve: the
// stop threshold rtol^2 |b|^2 is the cold solve's even when the solve starts
// from x0 (r0 = b - Z x0); an r0 already within it stops at once.
__global__ __launch_bounds__(kBlock) void k_tiled_init(const double* __restrict__ partials, SolveScalars* sc, double rtol,
                                                       int maxit) {
  double s[2];
  read_partials<2>(partials, s);
  if (threadIdx.x == 0) {
    sc->bb = s[0];
    sc->rr = s[1]; sc->rr_anchor = s[1]; sc->rr_hi = s[1];
    sc->rtol2bb = rtol * rtol * s[0];
    sc->iters = 0; sc->maxit = maxit;
    sc->stop = s[1] <= sc->rtol2bb ? 1 : (s[1] == s[1] ? 0 : 3);
    sc->active_v[0] = 0; sc->active_v[1] = 0;
    sc->rp = 0.0; sc->pq_prev = 0.0;
  }
}

// Flexible CG with the two-level preconditioner
//   z = S r + C r1 + S (r - Z (S r + C r1)),  r1 = r - Z S r,
// S = omega D^-1 (Jacobi smoothing) and C = gain J^T J (the network as the
// coarse correction); I - M Z = (I - S Z)(I - C Z)(I - S Z), so M is SPD
// whenever the smoothing converges (omega < 2 / lambda_max(D^-1 Z)), for any
// network weights. One iteration is nine launches: the direction with q = Z p
// (k_tile kTileDir), the update with the pre-smoothing (kTileSmooth), the
// network's five convolutions and its last layer forming z2, and the
// post-smoothing with z (kTilePost); the three Z products read a vector
// staged in LDS. The recurrence is FCG(1) in FP32 as in solve_learned, with
// FP64 reliable updates (every 1e-6 of residual reduction here) and the same
// FP64 stop test; a reliable update replaces r and keeps p, q and z, whose
// partial sums with the new r are rescanned (k_tile_rescan), so the next
// direction is Z-orthogonalised against the kept one as usual. With warm, x
// holds the initial guess and r0 = b - Z x is formed on the device.
int solve_learned_tiled(fsx_ctx* ctx, const double* b, double* x, bool warm, int* iters, double* relres,
                        int truncate_at) {
  fsx_unet* u = ctx->unet;
  const int n = ctx->NNp;
  // reliable updates every 1e-6 of FP32 residual reduction: the recurrence
  // keeps about 1e-7 of the residual it started the segment from, so the
  // FP64 check decides convergence before FP32 rounding does (1e-4 took the
  // same flexible iterations and one more FP64 residual a solve: 64^3 cavity
  // solve 72.6 against 69.8 ms over 20 steps)
  const double delta = 1e-6;
  const int max_learned = truncate_at > 0 ? truncate_at : std::min(ctx->maxit, ctx->learned_maxit);
  for (float** v : {&ctx->fr2, &ctx->fp2, &ctx->fr1, &ctx->fz2})
    if (!*v) FSX_TRY(alloc(ctx, v, (size_t)n));
  float* R[2] = {ctx->fr, ctx->fr2};
  float* P[2] = {ctx->fp, ctx->fp2};
  if (warm) {
    FSX_TRY(launch_warm_residual(ctx, x, b, ctx->wh));
  } else {
    FSX_TRY(launch_fill(ctx, x, 0.0, n));
    FSX_TRY(launch_copy(ctx, ctx->wh, b, n));
    hipLaunchKernelGGL(k_sumsq_u, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, b, ctx->partials);
    hipLaunchKernelGGL(k_sumsq_u, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, b, ctx->partials + kRedGrid);
  }
  hipLaunchKernelGGL(k_tiled_init, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->partials, ctx->sc, ctx->rtol, max_learned);
  // the direction before the first (p[1]) and q zero; r_0 = float(wh), x_lo = 0
  hipLaunchKernelGGL(k_zero_f, dim3(256), dim3(kBlock), 0, ctx->stream, n, P[1]);
  hipLaunchKernelGGL(k_zero_f, dim3(256), dim3(kBlock), 0, ctx->stream, n, ctx->fq);
  hipLaunchKernelGGL(k_fpcg_load_r, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, ctx->wh, R[0], ctx->fx, P[1],
                     ctx->partials);
  FSX_TRY(launch_check(ctx, "tiled learned start"));
  TileVecs tv{};
  tv.dinv = ctx->Zdinvf;
  tv.omega = ctx->learned_omega;
  tv.part = ctx->tile_part;
  tv.sc = ctx->sc;
  tv.maxit = max_learned;
  tv.delta2 = delta * delta;
  // z for the residual r[j & 1] (a start or a replaced residual)
  auto precondition_at = [&](int j) -> int {
    TileVecs a = tv;
    a.rin = R[j & 1]; a.pin = P[(j + 1) & 1]; a.r1 = ctx->fr1;
    FSX_TRY(launch_tile(ctx, kTilePre, a));
    FSX_TRY(apply_net_mult(ctx, u, ctx->fr1, R[j & 1], ctx->fz2, ctx->sc));
    TileVecs c = tv;
    c.rin = R[j & 1]; c.z2 = ctx->fz2; c.zout = ctx->fz; c.q = ctx->fq;
    return launch_tile(ctx, kTilePost, c);
  };
  auto issue = [&](int j) -> int {
    TileVecs d = tv;
    d.z = ctx->fz; d.pin = P[(j + 1) & 1]; d.pout = P[j & 1]; d.q = ctx->fq; d.j = j;
    FSX_TRY(launch_tile(ctx, kTileDir, d));
    TileVecs m = tv;
    m.pin = P[j & 1]; m.q = ctx->fq; m.rin = R[j & 1]; m.rout = R[(j + 1) & 1]; m.x = ctx->fx; m.r1 = ctx->fr1;
    FSX_TRY(launch_tile(ctx, kTileSmooth, m));
    FSX_TRY(apply_net_mult(ctx, u, ctx->fr1, R[(j + 1) & 1], ctx->fz2, ctx->sc));
    TileVecs c = tv;
    c.rin = R[(j + 1) & 1]; c.z2 = ctx->fz2; c.zout = ctx->fz; c.q = ctx->fq;
    return launch_tile(ctx, kTilePost, c);
  };
  FSX_TRY(precondition_at(0));
  int j = 0;
  int fallback = 0;
  // each segment enqueues as many iterations as the same segment of the last
  // solve took (4 the first time), then one at a time: iterations past the
  // stop are no-op launches, a read-back idles the device, and consecutive
  // pressure solves converge alike
  const bool keep_segs = truncate_at == 0;
  for (int segi = 0;; segi++) {
    const int j_seg = j;
    const int known = segi < 16 ? ctx->learned_seg[segi] : 0;
    int guess = std::max(1, std::min(known > 0 ? known : 4, max_learned - j + 1));
    int issued = 0;
    for (;;) {
      for (int c = 0; c < guess; c++) FSX_TRY(issue(j + c));
      j += guess;
      issued += guess;
      FSX_TRY(read_scalars(ctx));
      if (ctx->h_sc->stop) break;
      if (issued > max_learned + 64) return fail(ctx, FSX_ESTATE, "tiled FPCG did not terminate");
      guess = 1;
    }
    const int stop = ctx->h_sc->stop;
    j = ctx->h_sc->iters;
    if (keep_segs && segi < 16) ctx->learned_seg[segi] = std::max(1, j - j_seg);
    // reliable update: fold x_lo into x, FP64 residual wh = b - Z x, FP64 stop test
    int st = 0;
    double rr_hi = 0;
    FSX_TRY(launch_fold_residual(ctx, x, b, &st, &rr_hi));
    if (st == 3 || stop == 3) return fail(ctx, FSX_ENUMERIC, "learned solve met NaN/Inf");
    if (st == 1) break;                                     // converged in FP64
    if (stop == 6 || stop == 2 || st == 2) { fallback = stop == 6 ? 6 : 2; break; }
    // replace r by the FP64 residual; keep p, q and z (rescan its partials)
    FSX_TRY(launch_tile_rescan(ctx, ctx->wh, R[j & 1], P[(j + 1) & 1]));
  }
  const int learned_iters = ctx->h_sc->iters;
  const double bb = ctx->h_sc->bb, rh = ctx->h_sc->rr_hi;
  if (truncate_at > 0) {
    *iters = learned_iters;
    *relres = bb > 0 ? std::sqrt(rh / bb) : 0.0;
    return FSX_OK;
  }
  ctx->last_learned_iters = learned_iters;
  ctx->last_fallback = fallback;
  if (fallback) {
    // Jacobi-PCG (mixed) on the remaining FP64 residual wh, to the same
    // absolute threshold rtol ||b||
    int it2 = 0;
    double rel2 = 0;
    const double saved_rtol = ctx->rtol;
    ctx->rtol = rh > 0 ? std::min(1.0, saved_rtol * std::sqrt(bb / rh)) : saved_rtol;
    FSX_TRY(launch_copy(ctx, ctx->wr, ctx->wh, n));        // the mixed solver overwrites wh
    const int rc = solve_mixed_entry(ctx, ctx->wr, ctx->wz, &it2, &rel2);
    ctx->rtol = saved_rtol;
    if (rc != FSX_OK && rc != FSX_ENOCONV) return rc;
    FSX_TRY(launch_add(ctx, x, ctx->wz, n));
    int st = 0;
    double rr2 = 0;
    FSX_TRY(launch_residual_only(ctx, x, b, &st, &rr2));
    *iters = learned_iters + it2;
    *relres = bb > 0 ? std::sqrt(rr2 / bb) : 0.0;
    ctx->last_iters = *iters;
    if (st != 1) { ctx->err = "learned solve and Jacobi fallback reached maxit"; return FSX_ENOCONV; }
    return FSX_OK;
  }
  *iters = learned_iters;
  *relres = bb > 0 ? std::sqrt(rh / bb) : 0.0;
  ctx->last_iters = learned_iters;
  return FSX_OK;
}

void unet_free(fsx_ctx* ctx) {
  delete ctx->unet;   // device buffers belong to the context's allocation list
  ctx->unet = nullptr;
  ctx->capture_hook = nullptr;
}


}  // namespace fsx

namespace {

// stores (R2, Pdot) of every converged pressure solve while capacity lasts
void capture_pair(fsx_ctx* ctx, const double* b, const double* x) {
  fsx_unet* u = ctx->unet;
  if (!u || u->ncap >= u->cap_limit) return;
  const size_t n = ctx->NNp;
  (void)hipMemcpyAsync(u->cap_b + (size_t)u->ncap * n, b, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream);
  (void)hipMemcpyAsync(u->cap_x + (size_t)u->ncap * n, x, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream);
  u->ncap++;
}

int sum_partials_host(fsx_ctx* ctx, double* out, int slot = 0) {
  double h[kRedGrid];
  FSX_HIP(hipMemcpyAsync(h, ctx->partials + (size_t)slot * kRedGrid, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  double s = 0;
  for (int i = 0; i < kRedGrid; i++) s += h[i];
  *out = s;
  return FSX_OK;
}

int need_unet(fsx_ctx* ctx) {
  if (!ctx) return FSX_EINVAL;
  if (!ctx->unet) return fail(ctx, FSX_ESTATE, "call fsx_unet_setup first");
  return FSX_OK;
}

// y = Z x in FP32 (the solvers' FP32 form of Z)
void spmv_f(fsx_ctx* ctx, const float* x, float* y) { (void)launch_pcg_spmv_f32(ctx, x, y, nullptr); }

// One training sample: forward, loss, output gradient, backward (accumulates
// u->g). kind 0: relative L2 loss |y - t|^2 / tt with tt = |t|^2; kind 1:
// energy-norm loss (y - t)' Z (y - t) / tt with tt = t' Z t.
int train_sample(fsx_ctx* ctx, fsx_unet* u, int kind, const float* x, const float* t, double tt, double* loss) {
  const int n = ctx->NNp;
  FSX_TRY(forward(ctx, u, x, u->y));
  if (kind == 0) {
    hipLaunchKernelGGL(k_loss, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, u->y, t, tt, u->gy, ctx->partials);
  } else if (kind == 1) {
    hipLaunchKernelGGL(k_sub_f, dim3(256), dim3(kBlock), 0, ctx->stream, n, u->y, t, u->gy);
    spmv_f(ctx, u->gy, u->zq);
    hipLaunchKernelGGL(k_eloss, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, u->zq, tt, u->gy, ctx->partials);
  } else {
    const float om = ctx->learned_omega;
    hipLaunchKernelGGL(k_sub_f, dim3(256), dim3(kBlock), 0, ctx->stream, n, u->y, t, u->gy);
    spmv_f(ctx, u->gy, u->zq);
    hipLaunchKernelGGL(k_smooth_sub_f, dim3(256), dim3(kBlock), 0, ctx->stream, n, om, ctx->Zdinv, u->zq, u->gy);
    spmv_f(ctx, u->gy, u->zq);
    hipLaunchKernelGGL(k_eloss2a, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, om, ctx->Zdinv, u->gy, u->zq, u->ws,
                       ctx->partials);
    spmv_f(ctx, u->ws, u->gy);
    hipLaunchKernelGGL(k_eloss2b, dim3(256), dim3(kBlock), 0, ctx->stream, n, u->zq, tt, u->gy);
  }
  FSX_TRY(launch_check(ctx, "unet loss"));
  FSX_TRY(backward(ctx, u, x));
  if (loss) {
    double d2 = 0;
    FSX_TRY(sum_partials_host(ctx, &d2));
    *loss = d2 / tt;
  }
  return FSX_OK;
}

// Training pair s from the residual r = b - Z x_k in ctx->wh (|r|^2 partial
// sums in slot 0 of ctx->partials) and the error xs - xk of the iterate, in
// the additive (input r, target e - dinv r) or the two-level form (input
// r1 = r - Z S r, target e - S r); the pair's error energy goes to s_tt[s].
int make_pair(fsx_ctx* ctx, fsx_unet* u, bool two_level, const double* xs, const double* xk, int s) {
  const int n = ctx->NNp;
  if (two_level) {
    // t = omega dinv r, r1 = r - Z t (|r|^2 stays in slot 0)
    hipLaunchKernelGGL(k_scaled_dinv, dim3(256), dim3(kBlock), 0, ctx->stream, n, (double)ctx->learned_omega, ctx->Zdinv,
                       ctx->wh, ctx->w0e);
    FSX_TRY(launch_warm_residual(ctx, ctx->w0e, ctx->wh, ctx->w0r));
    hipLaunchKernelGGL(k_make_sample2, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, ctx->wh, ctx->w0r, ctx->w0e, xs, xk,
                       ctx->partials, u->s_in + (size_t)s * n, u->s_tgt + (size_t)s * n);
  } else {
    hipLaunchKernelGGL(k_make_sample, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, ctx->wh, xs, xk, ctx->Zdinv,
                       ctx->partials, u->s_in + (size_t)s * n, u->s_tgt + (size_t)s * n);
  }
  FSX_TRY(launch_check(ctx, "k_make_sample"));
  FSX_TRY(sum_partials_host(ctx, &u->s_tt[s], 1));
  if (!(u->s_tt[s] > 0)) return fail(ctx, FSX_ENUMERIC, "a training sample has no error energy");
  return FSX_OK;
}

// room for `total` pairs, keeping the first `keep`
int reserve_pairs(fsx_ctx* ctx, fsx_unet* u, int total, int keep) {
  const size_t n = ctx->NNp;
  if (total > u->samp_max) {
    float *ni = nullptr, *nt = nullptr;
    FSX_TRY(alloc(ctx, &ni, (size_t)total * n));
    FSX_TRY(alloc(ctx, &nt, (size_t)total * n));
    if (keep > 0) {
      FSX_HIP(hipMemcpyAsync(ni, u->s_in, (size_t)keep * n * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
      FSX_HIP(hipMemcpyAsync(nt, u->s_tgt, (size_t)keep * n * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    }
    FSX_HIP(hipStreamSynchronize(ctx->stream));
    if (u->s_in) { dev_free(ctx, u->s_in); dev_free(ctx, u->s_tgt); }
    u->s_in = ni;
    u->s_tgt = nt;
    u->samp_max = total;
  }
  u->s_tt.resize(total, 0.0);
  if (!u->tmp_x) FSX_TRY(alloc(ctx, &u->tmp_x, n));
  return FSX_OK;
}

}  // namespace

extern "C" {

int fsx_unet_setup(fsx_ctx* ctx, int nx, int ny, int nz, unsigned long long seed) {
  if (!ctx) return FSX_EINVAL;
  if (!ctx->z_built) return fail(ctx, FSX_ESTATE, "build Z before setting up the U-net");
  if (ctx->unet) return fail(ctx, FSX_ESTATE, "U-net already set up");
  fsx_unet* u = nullptr;
  FSX_TRY(setup_unet(ctx, nx, ny, nz, seed, &u));
  ctx->unet = u;
  return FSX_OK;
}

int fsx_unet_params(fsx_ctx* ctx, long long* n) {
  FSX_TRY(need_unet(ctx));
  *n = (long long)ctx->unet->nparam;
  return FSX_OK;
}

int fsx_unet_get_weights(fsx_ctx* ctx, float* host, size_t n) {
  FSX_TRY(need_unet(ctx));
  if (!host || n != ctx->unet->nparam) return fail(ctx, FSX_EINVAL, "weight count mismatch");
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  FSX_HIP(hipMemcpy(host, ctx->unet->w, n * sizeof(float), hipMemcpyDeviceToHost));
  return FSX_OK;
}

int fsx_unet_set_weights(fsx_ctx* ctx, const float* host, size_t n) {
  FSX_TRY(need_unet(ctx));
  if (!host || n != ctx->unet->nparam) return fail(ctx, FSX_EINVAL, "weight count mismatch");
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  FSX_HIP(hipMemcpy(ctx->unet->w, host, n * sizeof(float), hipMemcpyHostToDevice));
  FSX_TRY(refresh_weights(ctx, ctx->unet));
  return fsx_sync(ctx);
}

int fsx_unet_gain(fsx_ctx* ctx, float set, float* get) {
  FSX_TRY(need_unet(ctx));
  if (set > 0.0f) ctx->unet->gain = set;
  if (get) *get = ctx->unet->gain;
  return FSX_OK;
}

int fsx_unet_apply(fsx_ctx* ctx, const float* x, float* y) {
  FSX_TRY(need_unet(ctx));
  fsx_unet* u = ctx->unet;
  const size_t n = ctx->NNp;
  FSX_HIP(hipMemcpyAsync(u->x, x, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  FSX_TRY(forward(ctx, u, u->x, u->y));
  FSX_HIP(hipMemcpyAsync(y, u->y, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  return FSX_OK;
}

int fsx_unet_loss_grad(fsx_ctx* ctx, int kind, const float* x, const float* t, double* loss, float* grad) {
  FSX_TRY(need_unet(ctx));
  if (kind < 0 || kind > 2) return fail(ctx, FSX_EINVAL, "loss kind must be 0 (L2), 1 (energy) or 2 (two-level energy)");
  fsx_unet* u = ctx->unet;
  const size_t n = ctx->NNp;
  FSX_HIP(hipMemcpyAsync(u->x, x, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  FSX_HIP(hipMemcpyAsync(u->tgt, t, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  double tt = 0;
  if (kind == 0) {
    for (size_t i = 0; i < n; i++) tt += (double)t[i] * t[i];
  } else {
    spmv_f(ctx, u->tgt, u->zq);
    hipLaunchKernelGGL(k_dot_f, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, (int)n, u->tgt, u->zq, ctx->partials);
    FSX_TRY(sum_partials_host(ctx, &tt));
  }
  if (!(tt > 0)) return fail(ctx, FSX_EINVAL, "target has no positive norm");
  FSX_HIP(hipMemsetAsync(u->g, 0, u->nparam * sizeof(float), ctx->stream));
  double l = 0;
  FSX_TRY(train_sample(ctx, u, kind, u->x, u->tgt, tt, &l));
  if (loss) *loss = l;
  if (grad) FSX_HIP(hipMemcpy(grad, u->g, u->nparam * sizeof(float), hipMemcpyDeviceToHost));
  return FSX_OK;
}

int fsx_unet_capture(fsx_ctx* ctx, int capacity) {
  FSX_TRY(need_unet(ctx));
  fsx_unet* u = ctx->unet;
  if (capacity <= 0) { ctx->capture_hook = nullptr; return FSX_OK; }
  if (capacity > u->cap_max) {
    if (u->cap_b) { dev_free(ctx, u->cap_b); dev_free(ctx, u->cap_x); }
    u->cap_b = u->cap_x = nullptr;
    u->cap_max = u->cap_limit = u->ncap = 0;
    FSX_TRY(alloc(ctx, &u->cap_b, (size_t)capacity * ctx->NNp));
    FSX_TRY(alloc(ctx, &u->cap_x, (size_t)capacity * ctx->NNp));
  }
  u->cap_max = std::max(u->cap_max, capacity);
  u->cap_limit = capacity;
  u->ncap = 0;   // a new capture replaces the previous one
  ctx->capture_hook = capture_pair;
  return FSX_OK;
}

int fsx_unet_captured(fsx_ctx* ctx, int* n) {
  FSX_TRY(need_unet(ctx));
  *n = ctx->unet->ncap;
  return FSX_OK;
}

// Training pairs from the captured solves: for each (b, x*) and each k in
// ks, x_k = the k-th iterate from x = 0 of FP64 Jacobi-PCG, or (flag 1) of
// the learned flexible PCG itself, so that the network also sees the
// residuals it leaves; r = b - Z x_k; input r / rms(r), target
// (x* - x_k - dinv r) / rms(r), loss normaliser (x* - x_k).r / rms(r)^2 (the
// error energy). Pairs whose residual is below 1e-5 ||b|| are skipped (x*
// itself is only that accurate). Flag 2 appends to the existing pairs,
// otherwise they are replaced.
int fsx_unet_prepare(fsx_ctx* ctx, const int* ks, int nk, int flags, int* nsamples) {
  FSX_TRY(need_unet(ctx));
  fsx_unet* u = ctx->unet;
  if (!ks || nk <= 0) return fail(ctx, FSX_EINVAL, "no truncation depths");
  if (u->ncap == 0) return fail(ctx, FSX_ESTATE, "no captured solves");
  const bool use_learned = flags & 1, append = (flags & 2) && u->nsamp > 0;
  const bool two_level = ctx->learned_form == 1 && ctx->tile.ntiles > 0;
  const int n = ctx->NNp;
  if (append && u->s_kind != (two_level ? 2 : 1))
    return fail(ctx, FSX_ESTATE, "appended training pairs must be of the same preconditioner form");
  const int start = append ? u->nsamp : 0;
  FSX_TRY(reserve_pairs(ctx, u, start + u->ncap * nk, start));
  const double saved_rtol = ctx->rtol;
  const int saved_maxit = ctx->maxit, saved_last = ctx->last_iters;
  const int saved_li = ctx->last_learned_iters, saved_fb = ctx->last_fallback;
  auto hook = ctx->capture_hook;
  ctx->capture_hook = nullptr;
  auto restore = [&]() {
    ctx->rtol = saved_rtol; ctx->maxit = saved_maxit; ctx->last_iters = saved_last;
    ctx->last_learned_iters = saved_li; ctx->last_fallback = saved_fb; ctx->capture_hook = hook;
  };
  int s = start;
  for (int p = 0; p < u->ncap; p++) {
    const double* b = u->cap_b + (size_t)p * n;
    const double* xs = u->cap_x + (size_t)p * n;
    double bb = 0;
    hipLaunchKernelGGL(k_sumsq_u, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, b, ctx->partials);
    FSX_TRY(sum_partials_host(ctx, &bb));
    if (!(bb > 0)) continue;
    for (int q = 0; q < nk; q++) {
      if (ks[q] <= 0) {
        FSX_TRY(launch_fill(ctx, u->tmp_x, 0.0, n));
      } else {
        ctx->rtol = 0.0;
        ctx->maxit = ks[q];
        ctx->last_iters = ks[q];
        int it = 0;
        double rel = 0;
        const int rc = !use_learned ? solve_fp64_entry(ctx, b, u->tmp_x, &it, &rel)
                       : two_level ? solve_learned_tiled(ctx, b, u->tmp_x, false, &it, &rel, ks[q])
                                   : solve_learned(ctx, b, u->tmp_x, &it, &rel, ks[q]);
        if (rc != FSX_OK && rc != FSX_ENOCONV) { restore(); return rc; }
      }
      int st = 0;
      double rr = 0;
      FSX_TRY(launch_residual_only(ctx, u->tmp_x, b, &st, &rr));
      // an iterate closer than 1e-5 ||b|| to the solution is compared with x*
      // that is itself only good to the solver's tolerance: no information
      if (!(rr >= 1e-10 * bb)) continue;
      const int rc = make_pair(ctx, u, two_level, xs, u->tmp_x, s);
      if (rc != FSX_OK) { restore(); return rc; }
      s++;
    }
  }
  restore();
  if (s == 0) return fail(ctx, FSX_ESTATE, "no informative training pairs (all truncated iterates converged)");
  u->nsamp = s;
  u->s_kind = two_level ? 2 : 1;
  if (nsamples) *nsamples = s;
  return fsx_sync(ctx);
}

// Training pairs without solves: seeded errors e = (I - S Z)^sweeps e0, e0
// i.i.d. normal per node (0 at the pinned node), r = Z e in FP64, made into
// pairs of the current form like fsx_unet_prepare's and appended to them.
int fsx_unet_prepare_random(fsx_ctx* ctx, int count, int sweeps, unsigned long long seed, int* nsamples) {
  FSX_TRY(need_unet(ctx));
  fsx_unet* u = ctx->unet;
  if (count < 1 || sweeps < 0) return fail(ctx, FSX_EINVAL, "count >= 1 and sweeps >= 0 required");
  const bool two_level = ctx->learned_form == 1 && ctx->tile.ntiles > 0;
  if (u->nsamp > 0 && u->s_kind != (two_level ? 2 : 1))
    return fail(ctx, FSX_ESTATE, "appended training pairs must be of the same preconditioner form");
  const int n = ctx->NNp, start = u->nsamp;
  FSX_TRY(reserve_pairs(ctx, u, start + count, start));
  if (!u->zero_x) FSX_TRY(alloc(ctx, &u->zero_x, (size_t)n));
  FSX_TRY(launch_fill(ctx, u->zero_x, 0.0, n));
  const double om = ctx->learned_omega;
  for (int p = 0; p < count; p++) {
    hipLaunchKernelGGL(k_randn, dim3(256), dim3(kBlock), 0, ctx->stream, n, seed, p, ctx->zeroP, u->tmp_x);
    for (int k = 0; k < sweeps; k++) {
      FSX_TRY(spmv_host_op(ctx, 4, u->tmp_x, ctx->wh, 1.0, 0.0));
      hipLaunchKernelGGL(k_jacobi_err, dim3(256), dim3(kBlock), 0, ctx->stream, n, om, ctx->Zdinv, ctx->wh, u->tmp_x);
    }
    FSX_TRY(spmv_host_op(ctx, 4, u->tmp_x, ctx->wh, 1.0, 0.0));
    hipLaunchKernelGGL(k_sumsq_u, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, n, ctx->wh, ctx->partials);
    FSX_TRY(launch_check(ctx, "random pair"));
    FSX_TRY(make_pair(ctx, u, two_level, u->tmp_x, u->zero_x, start + p));
  }
  u->nsamp = start + count;
  u->s_kind = two_level ? 2 : 1;
  if (nsamples) *nsamples = u->nsamp;
  return fsx_sync(ctx);
}

// Adam on mini-batches drawn (seeded) from the prepared samples. Returns the
// mean relative loss of the last min(50, steps) steps.
int fsx_unet_train(fsx_ctx* ctx, int steps, int batch, float lr, unsigned long long seed, double* loss_out) {
  FSX_TRY(need_unet(ctx));
  fsx_unet* u = ctx->unet;
  if (u->nsamp == 0) return fail(ctx, FSX_ESTATE, "no prepared samples");
  if (steps < 1 || batch < 1 || !(lr > 0)) return fail(ctx, FSX_EINVAL, "steps, batch, lr");
  const int n = ctx->NNp;
  std::mt19937_64 rng(seed);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  double tail = 0;
  int tail_n = 0;
  for (int step = 0; step < steps; step++) {
    FSX_HIP(hipMemsetAsync(u->g, 0, u->nparam * sizeof(float), ctx->stream));
    double lsum = 0;
    for (int b = 0; b < batch; b++) {
      const int idx = (int)(rng() % (unsigned long long)u->nsamp);
      double l = 0;
      FSX_TRY(train_sample(ctx, u, u->s_kind, u->s_in + (size_t)idx * n, u->s_tgt + (size_t)idx * n, u->s_tt[idx], &l));
      lsum += l;
    }
    u->adam_t++;
    const float c1 = 1.0f - std::pow(b1, (float)u->adam_t), c2 = 1.0f - std::pow(b2, (float)u->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for((long long)u->nparam, 256)), dim3(256), 0, ctx->stream, (int)u->nparam, u->w,
                       u->g, u->m, u->v, nullptr, lr, b1, b2, eps, c1, c2);
    FSX_TRY(refresh_weights(ctx, u));
    if (step >= steps - 50) { tail += lsum / batch; tail_n++; }
  }
  if (loss_out) *loss_out = tail_n ? tail / tail_n : 0.0;
  return fsx_sync(ctx);
}

int fsx_unet_form(fsx_ctx* ctx, int form, float omega) {
  if (!ctx) return FSX_EINVAL;
  if (form != 0 && form != 1) return fail(ctx, FSX_EINVAL, "form must be 0 (additive) or 1 (two-level)");
  if (form == 1 && !(omega > 0.0f && omega < 1.0f)) return fail(ctx, FSX_EINVAL, "omega must lie in (0, 1)");
  ctx->learned_form = form;
  if (form == 1) ctx->learned_omega = omega;
  return FSX_OK;
}

int fsx_unet_form_get(fsx_ctx* ctx, int* form_in_use, float* omega) {
  if (!ctx) return FSX_EINVAL;
  if (form_in_use) *form_in_use = (ctx->learned_form == 1 && ctx->tile.ntiles > 0) ? 1 : 0;
  if (omega) *omega = ctx->learned_omega;
  return FSX_OK;
}

int fsx_unet_precondition(fsx_ctx* ctx, const float* r, float* z) {
  FSX_TRY(need_unet(ctx));
  if (!r || !z) return FSX_EINVAL;
  fsx_unet* u = ctx->unet;
  const int n = ctx->NNp;
  FSX_HIP(hipMemcpyAsync(ctx->fr, r, (size_t)n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_stop_clear, dim3(1), dim3(1), 0, ctx->stream, ctx->sc);
  hipLaunchKernelGGL(k_zero_f, dim3(256), dim3(kBlock), 0, ctx->stream, n, ctx->fq);
  if (ctx->learned_form == 1 && ctx->tile.ntiles > 0) {
    for (float** v : {&ctx->fr2, &ctx->fp2, &ctx->fr1, &ctx->fz2})
      if (!*v) FSX_TRY(alloc(ctx, v, (size_t)n));
    hipLaunchKernelGGL(k_zero_f, dim3(256), dim3(kBlock), 0, ctx->stream, n, ctx->fp2);
    TileVecs a{};
    a.dinv = ctx->Zdinvf; a.omega = ctx->learned_omega; a.part = ctx->tile_part; a.sc = ctx->sc;
    a.rin = ctx->fr; a.pin = ctx->fp2; a.r1 = ctx->fr1;
    FSX_TRY(launch_tile(ctx, kTilePre, a));
    FSX_TRY(apply_net_mult(ctx, u, ctx->fr1, ctx->fr, ctx->fz2, ctx->sc));
    TileVecs c = a;
    c.rin = ctx->fr; c.z2 = ctx->fz2; c.zout = ctx->fz; c.q = ctx->fq;
    FSX_TRY(launch_tile(ctx, kTilePost, c));
  } else {
    FSX_TRY(precondition(ctx, u, nullptr));
  }
  FSX_HIP(hipMemcpyAsync(z, ctx->fz, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  return FSX_OK;
}

int fsx_unet_limit(fsx_ctx* ctx, int learned_maxit) {
  FSX_TRY(need_unet(ctx));
  if (learned_maxit < 1) return fail(ctx, FSX_EINVAL, "learned_maxit >= 1 required");
  ctx->learned_maxit = learned_maxit;
  return FSX_OK;
}

int fsx_unet_stats(fsx_ctx* ctx, int* last_learned_iters, int* last_fallback) {
  if (!ctx) return FSX_EINVAL;
  if (last_learned_iters) *last_learned_iters = ctx->last_learned_iters;
  if (last_fallback) *last_fallback = ctx->last_fallback;
  return FSX_OK;
}

}  // extern "C"

