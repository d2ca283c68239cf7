// fsx_direct.hip -- direct pressure solve: Cholesky of Z as a band matrix.
//
// The MATLAB prototype of the reference factors the pressure matrix once
// before the time loop (chol with an AMD ordering, ML:1876-1885) and solves
// two triangular systems per pressure step (ML:2021-2030). Here Z is
// reordered to a small half-bandwidth w (the device's pressure numbering or
// reverse Cuthill-McKee, whichever is narrower), stored as its lower band,
// and factored on the device in 64-column panels:
//   k_band_potrf  the 64x64 diagonal block, in LDS (one workgroup);
//   k_band_trsm   the panel's rows below it, L21 = A21 L11^-T (one row per lane);
//   k_band_syrk   the trailing window, A22 -= L21 L21^T, in 64x64 LDS tiles.
// A band factor has no fill outside the band, so the factor overwrites Z's
// band in place. The 1024x1024 diagonal blocks of L are then inverted
// (k_band_inv) and L is also kept transposed as an upper band
// (k_band_transpose), so that both triangular solves read rows. Each solve
// runs in 1024-row steps: the off-diagonal band part as row sums, a
// workgroup a row with its four wavefronts on four slices of the columns
// (k_band_rowpart), then the block's inverse, a wavefront a row
// (k_band_apply_inv). Every sum has a fixed order, so the solve is
// deterministic.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "fsx_internal.h"

namespace fsx {
namespace {

constexpr int kPanel = 64;     // factor panel width
constexpr int kSB = 1024;      // rows per triangular-solve step (= lanes of k_band_inv)
constexpr int kQ = 4;          // slices of a row's off-diagonal columns in a solve step
constexpr int kTileLd = 66;    // LDS row length of the trailing-update tiles (64 + 2)
constexpr double kBandMaxBytes = 64.0 * 1024 * 1024 * 1024;   // band storage limit

__device__ __forceinline__ long long bidx(long long W, int w, long long i, long long j) { return i * W + (j - i + w); }

// band <- the lower band of Z in band order (row r of Z is band row iperm[r])
__global__ void k_band_from_z(int n, const int* __restrict__ Zrow, const int* __restrict__ Zcol,
                              const double* __restrict__ Z, const int* __restrict__ iperm, double* __restrict__ band,
                              long long W, int w) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int i = iperm[r];
  for (int k = Zrow[r]; k < Zrow[r + 1]; k++) {
    const int j = iperm[Zcol[k]];
    if (j <= i && i - j <= w) band[bidx(W, w, i, j)] = Z[k];
  }
}

// the padding rows n..np-1 are identity rows
__global__ void k_band_pad(int n, int np, double* __restrict__ band, long long W, int w) {
  const int i = n + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) band[bidx(W, w, i, i)] = 1.0;
}

// Cholesky of the 64x64 diagonal block at k0 in LDS. A pivot that is not
// positive sets *flag and is replaced by 1 so that the rest stays finite.
__global__ __launch_bounds__(256) void k_band_potrf(double* __restrict__ band, long long W, int w, int k0,
                                                    int* __restrict__ flag) {
  __shared__ double A[kPanel][kPanel + 1];
  const int t = threadIdx.x;
  for (int e = t; e < kPanel * kPanel; e += 256) {
    const int r = e / kPanel, c = e % kPanel;
    A[r][c] = (c <= r && r - c <= w) ? band[bidx(W, w, k0 + r, k0 + c)] : 0.0;
  }
  __syncthreads();
  for (int c = 0; c < kPanel; c++) {
    if (t == 0) {
      double d = A[c][c];
      if (!(d > 0.0)) { flag[0] = 1; d = 1.0; }
      A[c][c] = sqrt(d);
    }
    __syncthreads();
    const double d = A[c][c];
    for (int r = c + 1 + t; r < kPanel; r += 256) A[r][c] = A[r][c] / d;
    __syncthreads();
    const int m = kPanel - 1 - c;
    for (int e = t; e < m * m; e += 256) {
      const int r = c + 1 + e / m, s = c + 1 + e % m;
      if (s <= r) A[r][s] -= A[r][c] * A[s][c];
    }
    __syncthreads();
  }
  for (int e = t; e < kPanel * kPanel; e += 256) {
    const int r = e / kPanel, c = e % kPanel;
    if (c <= r && r - c <= w) band[bidx(W, w, k0 + r, k0 + c)] = A[r][c];
  }
}

// rows i in [k0+64, i_end): their entries in the panel's columns become
// L21 = A21 L11^-T (forward substitution along the row, one row per lane,
// the row's solved entries in LDS, lane fastest)
__global__ __launch_bounds__(64) void k_band_trsm(double* __restrict__ band, long long W, int w, int k0, int i_end) {
  __shared__ double L[kPanel][kPanel + 1];
  __shared__ double xs[kPanel * 64];
  const int t = threadIdx.x;
  for (int e = t; e < kPanel * kPanel; e += 64) {
    const int r = e / kPanel, c = e % kPanel;
    L[r][c] = (c <= r && r - c <= w) ? band[bidx(W, w, k0 + r, k0 + c)] : 0.0;
  }
  __syncthreads();
  const long long i = (long long)k0 + kPanel + (long long)blockIdx.x * 64 + t;
  if (i >= i_end) return;
  for (int c = 0; c < kPanel; c++) {
    const long long j = k0 + c;
    double a = (j >= i - w) ? band[bidx(W, w, i, j)] : 0.0;
    for (int q = 0; q < c; q++) a -= xs[q * 64 + t] * L[c][q];
    a = a / L[c][c];
    xs[c * 64 + t] = a;
    if (j >= i - w) band[bidx(W, w, i, j)] = a;
  }
}

// trailing window rows/cols [T0, T1), T0 = k0 + 64, T1 - T0 <= w:
// A(i, j) -= sum_c L(i, k0+c) L(j, k0+c) for j <= i, one 64x64 tile per
// workgroup (tile t -> row tile I, column tile J <= I), 4x4 outputs a lane
__global__ __launch_bounds__(256) void k_band_syrk(double* __restrict__ band, long long W, int w, int k0, int T1) {
  __shared__ double Pi[kPanel * kTileLd];
  __shared__ double Pj[kPanel * kTileLd];
  const int T0 = k0 + kPanel;
  const long long tt = blockIdx.x;
  int I = (int)((sqrt(8.0 * (double)tt + 1.0) - 1.0) * 0.5);
  while ((long long)(I + 1) * (I + 2) / 2 <= tt) I++;
  while ((long long)I * (I + 1) / 2 > tt) I--;
  const int J = (int)(tt - (long long)I * (I + 1) / 2);
  const int i0 = T0 + 64 * I, j0 = T0 + 64 * J;
  const int t = threadIdx.x;
  for (int e = t; e < 64 * kPanel; e += 256) {
    const int a = e / kPanel, c = e % kPanel;
    const long long col = k0 + c;
    const long long ri = i0 + a, rj = j0 + a;
    Pi[c * kTileLd + a] = (ri < T1 && col >= ri - w) ? band[bidx(W, w, ri, col)] : 0.0;
    Pj[c * kTileLd + a] = (rj < T1 && col >= rj - w) ? band[bidx(W, w, rj, col)] : 0.0;
  }
  __syncthreads();
  const int ta = t / 16, tb = t % 16;
  double acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) acc[x][y] = 0.0;
  for (int c = 0; c < kPanel; c++) {
    double pa[4], pb[4];
#pragma unroll
    for (int q = 0; q < 4; q++) { pa[q] = Pi[c * kTileLd + 4 * ta + q]; pb[q] = Pj[c * kTileLd + 4 * tb + q]; }
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
      for (int y = 0; y < 4; y++) acc[x][y] += pa[x] * pb[y];
  }
#pragma unroll
  for (int x = 0; x < 4; x++) {
    const long long i = i0 + 4 * ta + x;
    if (i >= T1) continue;
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const long long j = j0 + 4 * tb + y;
      if (j <= i) band[bidx(W, w, i, j)] -= acc[x][y];
    }
  }
}

// X = inverse of the kSB x kSB diagonal block k of L (lower triangular):
// lane t forms column t by forward substitution; written row-major
// (inv_rm, X(r, t) at r*kSB + t) and transposed (inv_cm, at t*kSB + r)
__global__ __launch_bounds__(kSB) void k_band_inv(const double* __restrict__ band, long long W, int w,
                                                  double* __restrict__ inv_rm, double* __restrict__ inv_cm) {
  const long long R0 = (long long)blockIdx.x * kSB;
  const int t = threadIdx.x;
  double* X = inv_rm + (long long)blockIdx.x * kSB * kSB;
  double* Xc = inv_cm + (long long)blockIdx.x * kSB * kSB;
  for (int r = 0; r < kSB; r++) {
    double v;
    if (r < t) {
      v = 0.0;
    } else if (r == t) {
      v = 1.0 / band[bidx(W, w, R0 + r, R0 + r)];
    } else {
      double acc = 0.0;
      for (int s = (r - w > t ? r - w : t); s < r; s++) acc += band[bidx(W, w, R0 + r, R0 + s)] * X[(long long)s * kSB + t];
      v = -acc / band[bidx(W, w, R0 + r, R0 + r)];
    }
    X[(long long)r * kSB + t] = v;
    Xc[(long long)t * kSB + r] = v;
  }
}

// U = L^T as an upper band: U(j, i) = L(i, j) at j*W + (i - j), i in [j, j + w]
__global__ void k_band_transpose(const double* __restrict__ band, double* __restrict__ up, long long W, int w, int np) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;   // L entry (i, o), o = j - i + w
  const long long i = e / (w + 1);
  const int o = (int)(e % (w + 1));
  if (i >= np) return;
  const long long j = i - w + o;
  if (j < 0) return;
  up[j * W + (i - j)] = band[i * W + o];
}

// y[k] = b[perm[k]] (0 in the padding), x[perm[k]] = y[k]
__global__ void k_band_gather(int n, int np, const int* __restrict__ perm, const double* __restrict__ b,
                              double* __restrict__ y) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < np) y[k] = k < n ? b[perm[k]] : 0.0;
}
__global__ void k_band_scatter(int n, const int* __restrict__ perm, const double* __restrict__ y, double* __restrict__ x) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[perm[k]] = y[k];
}

// One triangular-solve step's right-hand side for the kSB rows R0 + r:
// t[r] = rhs[R0 + r] - sum_j M(R0 + r, j) v[j] over the row's off-diagonal
// band columns, a workgroup per row, its kQ = 4 wavefronts on four slices of
// the columns (8 loads in flight a lane), summed in LDS in slice order.
// Forward (UPPER false): M = L, columns [max(0, i - w), R0), L(i, j) at
// i*W + (j - i + w). Backward (UPPER true): M = U, columns
// [R0 + kSB, min(np, i + w + 1)), U(i, j) at i*W + (j - i).
template <bool UPPER>
__global__ __launch_bounds__(64 * kQ) void k_band_rowpart(const double* __restrict__ M, long long W, int w, int np,
                                                          int R0, const double* __restrict__ v,
                                                          const double* __restrict__ rhs, double* __restrict__ t) {
  __shared__ double ws[kQ];
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int r = blockIdx.x;
  const long long i = (long long)R0 + r;
  long long lo, hi;
  if (UPPER) { lo = (long long)R0 + kSB; hi = i + w + 1 < np ? i + w + 1 : np; }
  else { lo = i - w > 0 ? i - w : 0; hi = R0; }
  double acc = 0.0;
  if (hi > lo) {
    const long long len = hi - lo, sl = (len + kQ - 1) / kQ;
    const long long a = lo + q * sl, b = a + sl < hi ? a + sl : hi;
    const long long base = UPPER ? i * W - i : i * W + w - i;   // M(i, j) at base + j
    for (long long j0 = a; j0 < b; j0 += 8 * 64) {
      double m[8], x[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const long long j = j0 + u * 64 + lane;
        const bool ok = j < b;
        m[u] = ok ? M[base + j] : 0.0;
        x[u] = ok ? v[j] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; u++) acc += m[u] * x[u];
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if (lane == 0) ws[q] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double p = rhs[i];
#pragma unroll
    for (int k = 0; k < kQ; k++) p -= ws[k];
    t[r] = p;
  }
}

// The step's diagonal block: out[R0 + r] = sum_s X(r, s) t[s], X the block's
// inverse (forward: Linv, nonzero for s <= r) or its transpose (backward:
// nonzero for s >= r), rows contiguous; a wavefront per row
__global__ __launch_bounds__(256) void k_band_apply_inv(const double* __restrict__ t, const double* __restrict__ X,
                                                        int upper, int R0, double* __restrict__ out) {
  __shared__ double ts[kSB];
  for (int s = threadIdx.x; s < kSB; s += 256) ts[s] = t[s];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  const double* xr = X + (long long)r * kSB;
  const int s_lo = upper ? r : 0, s_hi = upper ? kSB : r + 1;
  double acc = 0.0;
  for (int s0 = s_lo; s0 < s_hi; s0 += 4 * 64) {
    double m[4], y[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int s = s0 + u * 64 + lane;
      const bool ok = s < s_hi;
      m[u] = ok ? xr[s] : 0.0;
      y[u] = ok ? ts[s] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) acc += m[u] * y[u];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if (lane == 0) out[R0 + r] = acc;
}

// half-bandwidth of Z's pattern under the order iperm (pressure id -> band row)
int bandwidth_of(int n, const std::vector<int>& rp, const std::vector<int>& ci, const std::vector<int>& iperm) {
  int w = 0;
  for (int r = 0; r < n; r++)
    for (int k = rp[r]; k < rp[r + 1]; k++) w = std::max(w, std::abs(iperm[r] - iperm[ci[k]]));
  return w;
}

// reverse Cuthill-McKee over Z's pattern: breadth-first from a node of a
// last level reached from a minimum-degree node, neighbours by increasing
// degree, per connected component; returns iperm (pressure id -> position)
std::vector<int> rcm_iperm(int n, const std::vector<int>& rp, const std::vector<int>& ci) {
  std::vector<int> deg(n), order;
  order.reserve(n);
  for (int r = 0; r < n; r++) deg[r] = rp[r + 1] - rp[r];
  std::vector<char> seen(n, 0);
  std::vector<int> nb;
  auto bfs = [&](int s, std::vector<int>& out) {
    out.clear();
    std::vector<char> vis(n, 0);
    out.push_back(s);
    vis[s] = 1;
    for (size_t h = 0; h < out.size(); h++) {
      const int u = out[h];
      nb.clear();
      for (int k = rp[u]; k < rp[u + 1]; k++) if (!vis[ci[k]]) { vis[ci[k]] = 1; nb.push_back(ci[k]); }
      std::sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
      out.insert(out.end(), nb.begin(), nb.end());
    }
  };
  std::vector<int> comp;
  for (int start = 0; start < n; start++) {
    if (seen[start]) continue;
    bfs(start, comp);                       // the component of start
    int s = start;
    for (int u : comp) if (deg[u] < deg[s]) s = u;
    bfs(s, comp);                           // from its minimum-degree node
    bfs(comp.back(), comp);                 // from a node of the last level reached

    for (int u : comp) seen[u] = 1;
    order.insert(order.end(), comp.begin(), comp.end());
  }
  std::reverse(order.begin(), order.end());
  std::vector<int> iperm(n);
  for (int k = 0; k < n; k++) iperm[order[k]] = k;
  return iperm;
}

}  // namespace

// Factor Z (the assembled FP64 Z with its penalty pin) once; kept until Z
// changes (fsx_build_Z, a Z value upload). The band order is the narrower of
// the device's pressure numbering and reverse Cuthill-McKee.
int direct_factor(fsx_ctx* ctx) {
  if (ctx->band_ready) return FSX_OK;
  const auto t0 = std::chrono::steady_clock::now();
  const int n = ctx->NNp;
  std::vector<int> rp(n + 1), ci(ctx->nnzZ);
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  FSX_HIP(hipMemcpy(rp.data(), ctx->Zrow, rp.size() * sizeof(int), hipMemcpyDeviceToHost));
  FSX_HIP(hipMemcpy(ci.data(), ctx->Zcol, ci.size() * sizeof(int), hipMemcpyDeviceToHost));
  std::vector<int> iperm(n);
  for (int r = 0; r < n; r++) iperm[r] = r;
  int w = bandwidth_of(n, rp, ci, iperm);
  std::vector<int> iperm2 = rcm_iperm(n, rp, ci);
  const int w2 = bandwidth_of(n, rp, ci, iperm2);
  if (w2 < w) { w = w2; iperm.swap(iperm2); }
  const int np = (n + kSB - 1) / kSB * kSB;
  w = std::min(w, np - 1);
  const long long W = ((long long)w + 1 + 3) / 4 * 4;
  const double band_bytes = 8.0 * (double)np * (double)W, inv_bytes = 2.0 * 8.0 * (double)np * kSB;
  if (2.0 * band_bytes > kBandMaxBytes)
    return fail(ctx, FSX_ENOMEM, "direct solve: the factor of Z (half-bandwidth " + std::to_string(w) + ") and its "
                                 "transpose need " + std::to_string(2.0 * band_bytes / 1e9) +
                                 " GB, above the 64 GiB limit");
  std::vector<int> perm(n);
  for (int r = 0; r < n; r++) perm[iperm[r]] = r;
  // (re)allocate when the shape changed, or after an allocation failed part way
  const bool have_all = ctx->band && ctx->band_up && ctx->band_inv && ctx->band_vec && ctx->band_part &&
                        ctx->band_perm && ctx->band_iperm && ctx->band_flag;
  if (!have_all || ctx->band_n != np || ctx->band_W != W || ctx->band_nn != n) {
    for (void* p : {(void*)ctx->band, (void*)ctx->band_up, (void*)ctx->band_inv, (void*)ctx->band_vec,
                    (void*)ctx->band_part, (void*)ctx->band_perm, (void*)ctx->band_iperm, (void*)ctx->band_flag})
      if (p) dev_free(ctx, p);
    ctx->band = ctx->band_up = ctx->band_inv = ctx->band_vec = ctx->band_part = nullptr;
    ctx->band_perm = ctx->band_iperm = ctx->band_flag = nullptr;
    FSX_TRY(alloc(ctx, &ctx->band, (size_t)np * (size_t)W));
    FSX_TRY(alloc(ctx, &ctx->band_up, (size_t)np * (size_t)W));
    FSX_TRY(alloc(ctx, &ctx->band_inv, (size_t)(inv_bytes / 8.0)));
    FSX_TRY(alloc(ctx, &ctx->band_vec, 3 * (size_t)np));
    FSX_TRY(alloc(ctx, &ctx->band_part, (size_t)kQ * kSB));
    FSX_TRY(alloc(ctx, &ctx->band_perm, (size_t)n));
    FSX_TRY(alloc(ctx, &ctx->band_iperm, (size_t)n));
    FSX_TRY(alloc(ctx, &ctx->band_flag, 1));
  }
  ctx->band_n = np; ctx->band_W = W; ctx->band_w = w; ctx->band_nn = n;   // the shape the buffers are sized for
  FSX_HIP(hipMemcpy(ctx->band_perm, perm.data(), n * sizeof(int), hipMemcpyHostToDevice));
  FSX_HIP(hipMemcpy(ctx->band_iperm, iperm.data(), n * sizeof(int), hipMemcpyHostToDevice));
  const int zero = 0;
  FSX_HIP(hipMemcpy(ctx->band_flag, &zero, sizeof(int), hipMemcpyHostToDevice));
  FSX_TRY(launch_fill(ctx, ctx->band, 0.0, (long long)np * W));
  hipLaunchKernelGGL(k_band_from_z, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, ctx->Zrow, ctx->Zcol, ctx->Z,
                     ctx->band_iperm, ctx->band, W, w);
  if (np > n) hipLaunchKernelGGL(k_band_pad, dim3((np - n + 255) / 256), dim3(256), 0, ctx->stream, n, np, ctx->band, W, w);
  FSX_TRY(launch_check(ctx, "k_band_from_z"));
  for (int k0 = 0; k0 < np; k0 += kPanel) {
    hipLaunchKernelGGL(k_band_potrf, dim3(1), dim3(256), 0, ctx->stream, ctx->band, W, w, k0, ctx->band_flag);
    const int T0 = k0 + kPanel, T1 = (int)std::min<long long>(np, (long long)k0 + kPanel + w);
    if (T1 > T0) {
      hipLaunchKernelGGL(k_band_trsm, dim3((T1 - T0 + 63) / 64), dim3(64), 0, ctx->stream, ctx->band, W, w, k0, T1);
      const long long R = (T1 - T0 + 63) / 64;
      hipLaunchKernelGGL(k_band_syrk, dim3((unsigned)(R * (R + 1) / 2)), dim3(256), 0, ctx->stream, ctx->band, W, w, k0, T1);
    }
  }
  FSX_TRY(launch_check(ctx, "k_band_potrf/trsm/syrk"));
  hipLaunchKernelGGL(k_band_inv, dim3(np / kSB), dim3(kSB), 0, ctx->stream, ctx->band, W, w, ctx->band_inv,
                     ctx->band_inv + (size_t)np * kSB);
  FSX_TRY(launch_check(ctx, "k_band_inv"));
  FSX_TRY(launch_fill(ctx, ctx->band_up, 0.0, (long long)np * W));
  const long long nent = (long long)np * (w + 1);
  hipLaunchKernelGGL(k_band_transpose, dim3((unsigned)((nent + 255) / 256)), dim3(256), 0, ctx->stream, ctx->band,
                     ctx->band_up, W, w, np);
  FSX_TRY(launch_check(ctx, "k_band_transpose"));
  int flag = 0;
  FSX_HIP(hipMemcpyAsync(&flag, ctx->band_flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  if (flag) return fail(ctx, FSX_ENUMERIC, "direct solve: Cholesky of Z met a non-positive pivot");
  ctx->band_ready = true;
  ctx->band_factor_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return FSX_OK;
}

// x = Z^-1 b with the kept factor (ML:2021-2030): gather into band order,
// L y = b and L^T x = y in kSB-row steps (the off-diagonal band part in
// kQ slices a row, k_band_rowpart, then the diagonal block's inverse,
// k_band_apply_inv), scatter back
int direct_apply(fsx_ctx* ctx, const double* b, double* x) {
  const int n = ctx->NNp, np = ctx->band_n, w = ctx->band_w;
  const long long W = ctx->band_W;
  double* bb = ctx->band_vec;
  double* y = bb + np;
  double* xx = bb + 2 * (size_t)np;
  const double* inv_rm = ctx->band_inv;
  const double* inv_cm = ctx->band_inv + (size_t)np * kSB;
  double* t = ctx->band_part;
  const dim3 gpart(kSB), gapply(kSB / 4);
  hipLaunchKernelGGL(k_band_gather, dim3((np + 255) / 256), dim3(256), 0, ctx->stream, n, np, ctx->band_perm, b, bb);
  for (int R0 = 0; R0 < np; R0 += kSB) {
    hipLaunchKernelGGL(k_band_rowpart<false>, gpart, dim3(64 * kQ), 0, ctx->stream, ctx->band, W, w, np, R0, y, bb, t);
    hipLaunchKernelGGL(k_band_apply_inv, gapply, dim3(256), 0, ctx->stream, t, inv_rm + (size_t)R0 * kSB, 0, R0, y);
  }
  for (int R0 = np - kSB; R0 >= 0; R0 -= kSB) {
    hipLaunchKernelGGL(k_band_rowpart<true>, gpart, dim3(64 * kQ), 0, ctx->stream, ctx->band_up, W, w, np, R0, xx, y, t);
    hipLaunchKernelGGL(k_band_apply_inv, gapply, dim3(256), 0, ctx->stream, t, inv_cm + (size_t)R0 * kSB, 1, R0, xx);
  }
  hipLaunchKernelGGL(k_band_scatter, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, ctx->band_perm, xx, x);
  return launch_check(ctx, "direct solve");
}

}  // namespace fsx

extern "C" int fsx_direct_stats(fsx_ctx* ctx, int* half_bandwidth, double* factor_ms, double* band_bytes) {
  if (!ctx) return FSX_EINVAL;
  if (ctx->direct_order == FSX_DIRECT_ND) {
    double lb = 0.0, ms = 0.0;
    fsx::mf_info(ctx, &lb, nullptr, nullptr, nullptr, &ms);
    if (half_bandwidth) *half_bandwidth = -1;
    if (factor_ms) *factor_ms = ms;
    if (band_bytes) *band_bytes = lb;   // L's front columns
    return FSX_OK;
  }
  if (half_bandwidth) *half_bandwidth = ctx->band_ready ? ctx->band_w : -1;
  if (factor_ms) *factor_ms = ctx->band_ready ? ctx->band_factor_ms : 0.0;
  if (band_bytes) *band_bytes = ctx->band_ready ? 8.0 * (double)ctx->band_n * (double)ctx->band_W : 0.0;   // L's band
  return FSX_OK;
}

extern "C" int fsx_set_direct_order(fsx_ctx* ctx, int order) {
  if (!ctx) return FSX_EINVAL;
  if (order != FSX_DIRECT_BAND && order != FSX_DIRECT_ND) return fsx::fail(ctx, FSX_EINVAL, "fsx_set_direct_order: unknown order");
  ctx->direct_order = order;
  return FSX_OK;
}

extern "C" int fsx_direct_info(fsx_ctx* ctx, double* factor_bytes, double* factor_flops, int* fronts, int* depth) {
  if (!ctx) return FSX_EINVAL;
  return fsx::mf_info(ctx, factor_bytes, factor_flops, fronts, depth, nullptr);
}
