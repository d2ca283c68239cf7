// fsx_sparse.hip -- direct pressure solve by a multifrontal Cholesky of Z in
// a nested-dissection order.
//
// The MATLAB prototype of the reference orders the pressure matrix with
// symamd, factors it once before the time loop (ML:1876-1885) and solves two
// triangular systems per pressure step (ML:2021-2030). The order here is the
// nested dissection of fsx_nd.cpp (host): every front's pivots are a
// contiguous range of the new numbering, its rows are the pivots followed by
// its boundary, and all fronts of one tree level are independent.
//
// Factor, level by level from the deepest. Each front is a dense m x m
// column-major matrix in a level workspace: zeroed, Z's lower entries of its
// pivot columns scattered in (k_mf_scatter_z), its children's update
// matrices added (k_mf_extend, one child slot per launch, so no two writes
// meet), then its pivot columns factored in 64-column panels across all
// fronts of the level at once (k_mf_potrf in LDS, k_mf_trsm a row a lane,
// k_mf_syrk 64x64 tiles with 4x4 outputs a lane). That leaves L's columns
// and, in the trailing boundary block, the front's update matrix; the
// columns go to the factor store (k_mf_store), the update matrix waits in
// the workspace for the parent level. The 128x128 diagonal blocks of every
// front's pivot block are inverted once (k_mf_inv).
//
// After the levels, every front's pivot block is inverted explicitly, X =
// inv(L11): its 128x128 diagonal blocks first (k_mf_inv), then block row by
// block row, right-looking, from the identity (k_mf_xrow applies the
// diagonal block's inverse to the block row, k_mf_xupd subtracts its
// product with L's block column from the rows below), all fronts together.
//
// Solve. Forward, from the deepest level: a front's vector is its pivots'
// right-hand side plus its children's update vectors (k_mf_fwd_init, one
// workgroup a front, children in slot order; on levels of few fronts
// k_mf_fwd_base and k_mf_fwd_child per slot, a lane an entry); y = X w_P (k_mf_fwd_x, column
// chunks of 256 in separate workgroups, their partial sums kept), then the
// update vector u = w_B - L21 y (k_mf_fwd_l21 sums y from those partials and
// forms L21's partial sums by column chunk, k_mf_fwd_reduce adds them).
// Backward, from the roots: y_P -= L21^T x_B (k_mf_bwd_bound_lds, a wavefront a
// column, x_B staged in LDS), then x_P = X^T y_P (k_mf_bwd_x, a wavefront a column of X). Each
// level is six launches whatever its fronts' sizes, and every sum has a
// fixed order: the solve is deterministic.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "fsx_internal.h"
#include "fsx_nd.h"

namespace fsx {

namespace {

constexpr int kLeaf = 64;       // nodes of a leaf front
constexpr int kPanel = 64;      // factor panel width
constexpr int kSB = 128;        // rows of a triangular-solve step (inverted diagonal blocks)
constexpr int kTileLd = 66;     // LDS row length of the trailing-update tiles
constexpr int kMaxGridY = 65535;
constexpr size_t kMaxBoundLds = 128 * 1024;   // LDS bytes of a front's staged boundary values
constexpr int kWideLevel = 64;                 // levels of at most this many fronts start their vectors wide

// A front as the kernels read it (offsets in doubles or ints)
struct MfFront {
  long long off;      // its m x m matrix in the level workspace
  long long loff;     // L's columns (m x npiv, column-major) in the factor store
  long long ioff;     // inverses of its 128x128 diagonal blocks
  long long woff;     // its vector in the forward-solve workspace
  long long boff;     // its boundary rows (new numbering) in the index store
  long long coff[2];  // its children's rows in this front (nb of the child each)
  long long xoff;     // X = inv(L11), npiv x npiv column-major, in the inverse store
  long long poff;     // partial sums of X w by 256-column chunk (ceil(npiv / 256) x npiv)
  long long qoff;     // partial sums of L21 y by 256-column chunk (ceil(npiv / 256) x nb)
  int m, npiv, p0, nb;
  int child[2];       // device indices of the children (-1: none)
};

}  // namespace

struct MfLevel {
  int f0 = 0, nf = 0;           // device fronts [f0, f0 + nf)
  int maxpiv = 0, maxm = 0, maxnb = 0;
  int maxnb_child[2] = {0, 0};  // largest boundary of a child in each slot
  long long ws = 0;             // doubles of the level workspace
  long long a0 = 0, na = 0;     // range of the level's Z entries in the assembly list
};

struct MfState {
  std::vector<MfLevel> lev;     // by depth (0: roots)
  std::vector<int> hm, hpiv;    // per device front
  int nfront = 0, depth = 0, n = 0;
  double nnz_l = 0, flops = 0, l_entries = 0, x_entries = 0, factor_ms = 0;
  MfFront* F = nullptr;
  int* cmap = nullptr;
  int* bidx = nullptr;
  int* perm = nullptr;
  long long* adest = nullptr;
  int* asrc = nullptr;
  double* ws[2] = {nullptr, nullptr};
  double* L = nullptr;
  double* invc = nullptr;       // inverses of the 128x128 diagonal blocks, column-major
  int2* itask = nullptr;        // (front, block) of each inverse
  int ninv = 0;
  double* Xs = nullptr;         // every front's X = inv(L11)
  double* P = nullptr;          // partial sums of the forward solve
  int* xlist = nullptr;         // fronts by descending block count (the inversion's steps)
  std::vector<int> xcnt;        // fronts with more than s blocks, s = 0, 1, ...
  std::vector<int> xorder;      // the host copy of xlist
  int maxpiv = 0;
  double* W = nullptr;
  double* Y = nullptr;
  double* X = nullptr;
  double* bp = nullptr;
  int* flag = nullptr;
  bool ready = false;
};

namespace {

// --------------------------------------------------------------- factor kernels

__global__ void k_mf_scatter_z(long long na, const long long* __restrict__ dest, const int* __restrict__ src,
                               const double* __restrict__ Z, double* __restrict__ ws) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < na; e += (long long)gridDim.x * blockDim.x)
    ws[dest[e]] = Z[src[e]];
}

// parent rows += the child's update matrix (its trailing nb x nb block),
// lower triangle, for the children in one slot: block (b, front) adds
// column b of the child; the maps ascend, so lower entries stay lower
__global__ __launch_bounds__(256) void k_mf_extend(const MfFront* __restrict__ F, int f0, int slot,
                                                   const int* __restrict__ cmap, const double* __restrict__ wsb,
                                                   double* __restrict__ ws) {
  const MfFront T = F[f0 + blockIdx.y];
  const int c = T.child[slot];
  if (c < 0) return;
  const MfFront C = F[c];
  const int b = blockIdx.x;
  if (b >= C.nb) return;
  const int* __restrict__ mp = cmap + T.coff[slot];
  const long long col = T.off + (long long)T.m * mp[b];
  const double* __restrict__ src = wsb + C.off + (long long)C.m * (C.npiv + b) + C.npiv;
  for (int a = b + threadIdx.x; a < C.nb; a += blockDim.x) ws[col + mp[a]] += src[a];
}

// Cholesky of the panel's diagonal block (kb x kb, kb <= 64) of every front
// of the level with pivots left at k0, in LDS. A pivot that is not positive
// sets *flag and is replaced by 1 so that the rest stays finite.
__global__ __launch_bounds__(256) void k_mf_potrf(const MfFront* __restrict__ F, int f0, int k0, double* __restrict__ ws,
                                                  int* __restrict__ flag) {
  __shared__ double A[kPanel][kPanel + 1];
  const MfFront T = F[f0 + blockIdx.x];
  if (T.npiv <= k0) return;
  const int kb = min(kPanel, T.npiv - k0);
  const long long base = T.off + (long long)k0 * T.m + k0;
  const int t = threadIdx.x;
  for (int e = t; e < kPanel * kPanel; e += 256) {
    const int c = e / kPanel, r = e % kPanel;
    A[r][c] = (r < kb && c < kb && c <= r) ? ws[base + (long long)c * T.m + r] : (r == c ? 1.0 : 0.0);
  }
  __syncthreads();
  for (int c = 0; c < kb; c++) {
    if (t == 0) {
      double d = A[c][c];
      if (!(d > 0.0)) { flag[0] = 1; d = 1.0; }
      A[c][c] = sqrt(d);
    }
    __syncthreads();
    const double d = A[c][c];
    for (int r = c + 1 + t; r < kb; r += 256) A[r][c] = A[r][c] / d;
    __syncthreads();
    const int m = kb - 1 - c;
    for (int e = t; e < m * m; e += 256) {
      const int r = c + 1 + e / m, s = c + 1 + e % m;
      if (s <= r) A[r][s] -= A[r][c] * A[s][c];
    }
    __syncthreads();
  }
  for (int e = t; e < kPanel * kPanel; e += 256) {
    const int c = e / kPanel, r = e % kPanel;
    if (r < kb && c <= r) ws[base + (long long)c * T.m + r] = A[r][c];
  }
}

// rows [k0 + kb, m) of the panel's columns become L21 = A21 L11^-T: a row a
// lane, forward substitution along the row, solved entries in LDS
__global__ __launch_bounds__(64) void k_mf_trsm(const MfFront* __restrict__ F, int f0, int k0, double* __restrict__ ws) {
  __shared__ double Ld[kPanel][kPanel + 1];
  __shared__ double xs[kPanel * 64];
  const MfFront T = F[f0 + blockIdx.y];
  if (T.npiv <= k0) return;
  const int kb = min(kPanel, T.npiv - k0);
  const long long r0 = (long long)k0 + kb + (long long)blockIdx.x * 64;
  if (r0 >= T.m) return;
  const int t = threadIdx.x;
  const long long base = T.off + (long long)k0 * T.m;
  for (int e = t; e < kPanel * kPanel; e += 64) {
    const int c = e / kPanel, r = e % kPanel;
    Ld[r][c] = (r < kb && c < kb && c <= r) ? ws[base + (long long)c * T.m + k0 + r] : 0.0;
  }
  __syncthreads();
  const long long i = r0 + t;
  if (i >= T.m) return;
  for (int c = 0; c < kb; c++) {
    const long long at = base + (long long)c * T.m + i;
    double a = ws[at];
    for (int q = 0; q < c; q++) a -= xs[q * 64 + t] * Ld[c][q];
    a = a / Ld[c][c];
    xs[c * 64 + t] = a;
    ws[at] = a;
  }
}

// trailing block [T0, m)^2, T0 = k0 + kb: A(i, j) -= sum_c L(i, k0 + c)
// L(j, k0 + c) for j <= i, one 64x64 tile (I >= J) a workgroup, 4x4 outputs
// a lane
__global__ __launch_bounds__(256) void k_mf_syrk(const MfFront* __restrict__ F, int f0, int k0, double* __restrict__ ws) {
  __shared__ double Pi[kPanel * kTileLd];
  __shared__ double Pj[kPanel * kTileLd];
  const MfFront T = F[f0 + blockIdx.y];
  if (T.npiv <= k0) return;
  const int kb = min(kPanel, T.npiv - k0);
  const long long T0 = (long long)k0 + kb;
  const long long nt = (T.m - T0 + 63) / 64;
  const long long tt = blockIdx.x;
  if (tt >= nt * (nt + 1) / 2) return;
  long long I = (long long)((sqrt(8.0 * (double)tt + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= tt) I++;
  while (I * (I + 1) / 2 > tt) I--;
  const long long J = tt - I * (I + 1) / 2;
  const long long i0 = T0 + 64 * I, j0 = T0 + 64 * J;
  const int t = threadIdx.x;
  const long long pbase = T.off + (long long)k0 * T.m;
  for (int e = t; e < 64 * kPanel; e += 256) {
    const int c = e / 64, a = e % 64;
    const long long ri = i0 + a, rj = j0 + a;
    const long long col = pbase + (long long)c * T.m;
    Pi[c * kTileLd + a] = (c < kb && ri < T.m) ? ws[col + ri] : 0.0;
    Pj[c * kTileLd + a] = (c < kb && rj < T.m) ? ws[col + rj] : 0.0;
  }
  __syncthreads();
  const int ta = t / 16, tb = t % 16;
  double acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) acc[x][y] = 0.0;
  for (int c = 0; c < kb; c++) {
    double pa[4], pb[4];
#pragma unroll
    for (int q = 0; q < 4; q++) { pa[q] = Pi[c * kTileLd + 4 * ta + q]; pb[q] = Pj[c * kTileLd + 4 * tb + q]; }
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
      for (int y = 0; y < 4; y++) acc[x][y] += pa[x] * pb[y];
  }
#pragma unroll
  for (int y = 0; y < 4; y++) {
    const long long j = j0 + 4 * tb + y;
    if (j >= T.m) continue;
#pragma unroll
    for (int x = 0; x < 4; x++) {
      const long long i = i0 + 4 * ta + x;
      if (i < T.m && j <= i) ws[T.off + j * T.m + i] -= acc[x][y];
    }
  }
}

// L's columns of the level's fronts into the factor store: block (j, front)
// copies column j
__global__ __launch_bounds__(256) void k_mf_store(const MfFront* __restrict__ F, int f0, const double* __restrict__ ws,
                                                  double* __restrict__ L) {
  const MfFront T = F[f0 + blockIdx.y];
  const int j = blockIdx.x;
  if (j >= T.npiv) return;
  const double* __restrict__ s = ws + T.off + (long long)j * T.m;
  double* __restrict__ d = L + T.loff + (long long)j * T.m;
  for (int r = j + threadIdx.x; r < T.m; r += blockDim.x) d[r] = s[r];
}

// inverse of diagonal block s of a front's pivot block (kb x kb, lower
// triangular; identity past kb): lane c forms column c by forward
// substitution, column-major (invc, X(r, c) at c*128 + r)
__global__ __launch_bounds__(kSB) void k_mf_inv(const MfFront* __restrict__ F, const int2* __restrict__ task,
                                                const double* __restrict__ L, double* __restrict__ invc) {
  const int2 tk = task[blockIdx.x];
  const MfFront T = F[tk.x];
  const int s = tk.y, R0 = s * kSB, kb = min(kSB, T.npiv - R0);
  double* __restrict__ Xc = invc + T.ioff + (long long)s * kSB * kSB;
  const int c = threadIdx.x;
  const double* __restrict__ Lb = L + T.loff + (long long)R0 * T.m + R0;   // L(R0 + r, R0 + q) at q*m + r
  for (int r = 0; r < kSB; r++) {
    double v;
    if (r >= kb || c >= kb) {
      v = r == c ? 1.0 : 0.0;
    } else if (r < c) {
      v = 0.0;
    } else if (r == c) {
      v = 1.0 / Lb[(long long)r * T.m + r];
    } else {
      double acc = 0.0;
      for (int q = c; q < r; q++) acc += Lb[(long long)q * T.m + r] * Xc[c * kSB + q];
      v = -acc / Lb[(long long)r * T.m + r];
    }
    Xc[c * kSB + r] = v;
  }
}

// ones on every front's diagonal of X (the store is zeroed before)
__global__ void k_mf_xdiag(const MfFront* __restrict__ F, double* __restrict__ Xs) {
  const MfFront T = F[blockIdx.x];
  for (int c = threadIdx.x; c < T.npiv; c += blockDim.x) Xs[T.xoff + (long long)c * T.npiv + c] = 1.0;
}

// block row s of X, columns [0, min(npiv, 128(s+1))): X_s = inv(L_ss) X_s
// (the rows hold what the earlier steps left of the identity), a 64-column
// chunk a workgroup: lane (i, h) forms row i for 32 of its columns
__global__ __launch_bounds__(256) void k_mf_xrow(const MfFront* __restrict__ F, const int* __restrict__ list, int s,
                                                 const double* __restrict__ invc, double* __restrict__ Xs) {
  __shared__ double Ws[kSB][65];
  const MfFront T = F[list[blockIdx.y]];
  const int R0 = s * kSB;
  if (T.npiv <= R0) return;
  const int kb = min(kSB, T.npiv - R0);
  const int ncols = min(T.npiv, R0 + kSB);
  const int c0 = 64 * blockIdx.x;
  if (c0 >= ncols) return;
  const int t = threadIdx.x;
  double* __restrict__ Xf = Xs + T.xoff;   // X(r, c) at c*npiv + r
  for (int e = t; e < kSB * 64; e += 256) {
    const int c = e / kSB, i = e % kSB;
    Ws[i][c] = (i < kb && c0 + c < ncols) ? Xf[(long long)(c0 + c) * T.npiv + R0 + i] : 0.0;
  }
  __syncthreads();
  const int i = t & (kSB - 1), h = t >> 7;
  if (i >= kb) return;
  double acc[32];
#pragma unroll
  for (int c = 0; c < 32; c++) acc[c] = 0.0;
  const double* __restrict__ inv = invc + T.ioff + (long long)s * kSB * kSB + i;   // inv(i, j) at j*128 + i
  for (int j = 0; j <= i; j++) {
    const double a = inv[(long long)j * kSB];
#pragma unroll
    for (int c = 0; c < 32; c++) acc[c] += a * Ws[j][32 * h + c];
  }
#pragma unroll
  for (int c = 0; c < 32; c++) {
    const int col = c0 + 32 * h + c;
    if (col < ncols) Xf[(long long)col * T.npiv + R0 + i] = acc[c];
  }
}

// rows [R1, npiv) of X, R1 = 128(s+1), columns [0, R1): X(r, c) -= sum_k
// L(r, R0 + k) X(R0 + k, c) over the 128 columns of block s, one 64x64 tile
// a workgroup, 4x4 outputs a lane
__global__ __launch_bounds__(256) void k_mf_xupd(const MfFront* __restrict__ F, const int* __restrict__ list, int s,
                                                 const double* __restrict__ L, double* __restrict__ Xs) {
  __shared__ double A[kPanel * kTileLd];   // L(r0 + a, R0 + kc + k) at k*kTileLd + a
  __shared__ double B[kPanel * kTileLd];   // X(R0 + kc + k, c0 + c) at k*kTileLd + c
  const MfFront T = F[list[blockIdx.y]];
  const int R0 = s * kSB, R1 = R0 + kSB;
  if (T.npiv <= R1) return;
  const int ct = R1 / 64;
  const long long I = blockIdx.x / ct, J = blockIdx.x % ct;
  const long long r0 = R1 + 64 * I, c0 = 64 * J;
  if (r0 >= T.npiv) return;
  const int t = threadIdx.x, ta = t / 16, tb = t % 16;
  double acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) acc[x][y] = 0.0;
  for (int kc = 0; kc < kSB; kc += kPanel) {
    for (int e = t; e < 64 * kPanel; e += 256) {
      const int k = e / 64, a = e % 64;
      const long long r = r0 + a;
      A[k * kTileLd + a] = r < T.npiv ? L[T.loff + (long long)(R0 + kc + k) * T.m + r] : 0.0;
      B[a * kTileLd + k] = Xs[T.xoff + (c0 + k) * (long long)T.npiv + R0 + kc + a];
    }
    __syncthreads();
    for (int k = 0; k < kPanel; k++) {
      double pa[4], pb[4];
#pragma unroll
      for (int q = 0; q < 4; q++) { pa[q] = A[k * kTileLd + 4 * ta + q]; pb[q] = B[k * kTileLd + 4 * tb + q]; }
#pragma unroll
      for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) acc[x][y] += pa[x] * pb[y];
    }
    __syncthreads();
  }
#pragma unroll
  for (int y = 0; y < 4; y++) {
    const long long c = c0 + 4 * tb + y;
#pragma unroll
    for (int x = 0; x < 4; x++) {
      const long long r = r0 + 4 * ta + x;
      if (r < T.npiv) Xs[T.xoff + c * T.npiv + r] -= acc[x][y];
    }
  }
}

// --------------------------------------------------------------- solve kernels

__global__ void k_mf_permute(int n, const int* __restrict__ perm, const double* __restrict__ b, double* __restrict__ bp) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) bp[k] = b[perm[k]];
}
__global__ void k_mf_unpermute(int n, const int* __restrict__ perm, const double* __restrict__ X, double* __restrict__ x) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[perm[k]] = X[k];
}

// a front's vector: its pivots' right-hand side, zeros on the boundary, plus
// its children's update vectors in slot order
__global__ __launch_bounds__(256) void k_mf_fwd_init(const MfFront* __restrict__ F, int f0, const int* __restrict__ cmap,
                                                     const double* __restrict__ bp, double* __restrict__ W) {
  const MfFront T = F[f0 + blockIdx.x];
  for (int k = threadIdx.x; k < T.m; k += blockDim.x) W[T.woff + k] = k < T.npiv ? bp[T.p0 + k] : 0.0;
  __syncthreads();
#pragma unroll
  for (int slot = 0; slot < 2; slot++) {
    const int c = T.child[slot];
    if (c >= 0) {
      const MfFront C = F[c];
      const int* __restrict__ mp = cmap + T.coff[slot];
      for (int a = threadIdx.x; a < C.nb; a += blockDim.x) W[T.woff + mp[a]] += W[C.woff + C.npiv + a];
    }
    __syncthreads();
  }
}

// The same in three launches for levels of few, large fronts: the vectors'
// start (b on the pivots, zeros on the boundary), then each child slot's
// update vectors added, a 256-entry chunk a workgroup
__global__ __launch_bounds__(256) void k_mf_fwd_base(const MfFront* __restrict__ F, int f0, const double* __restrict__ bp,
                                                     double* __restrict__ W) {
  const MfFront T = F[f0 + blockIdx.y];
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < T.m) W[T.woff + k] = k < T.npiv ? bp[T.p0 + k] : 0.0;
}
__global__ __launch_bounds__(256) void k_mf_fwd_child(const MfFront* __restrict__ F, int f0, int slot,
                                                      const int* __restrict__ cmap, double* __restrict__ W) {
  const MfFront T = F[f0 + blockIdx.y];
  const int c = T.child[slot];
  if (c < 0) return;
  const MfFront C = F[c];
  const int a = blockIdx.x * 256 + threadIdx.x;
  if (a < C.nb) W[T.woff + cmap[T.coff[slot] + a]] += W[C.woff + C.npiv + a];
}

// partial sums of y = X w_P: workgroup (row chunk, column chunk cc, front)
// sums columns [256 cc, 256 cc + 256) of its 64 rows, 64 columns a
// wavefront, the wavefronts' sums added in a fixed order; rows r >= 256 cc
// keep P(cc, r) (X is lower triangular: the later chunks of a row are zero)
__global__ __launch_bounds__(256) void k_mf_fwd_x(const MfFront* __restrict__ F, int f0, const double* __restrict__ W,
                                                  const double* __restrict__ Xs, double* __restrict__ P) {
  __shared__ double red[4][64];
  const MfFront T = F[f0 + blockIdx.z];
  const int r0 = 64 * blockIdx.x, cc = blockIdx.y, j0 = 256 * cc;
  if (r0 >= T.npiv || j0 > r0 + 63) return;
  const int t = threadIdx.x, lane = t & 63, q = t >> 6;
  const int r = r0 + lane;
  const int ja = j0 + 64 * q, jb = min(min(ja + 64, T.npiv), r0 + 64);
  double acc = 0.0;
  if (r < T.npiv) {
    const double* __restrict__ Xc = Xs + T.xoff + r;
    const double* __restrict__ w = W + T.woff;
    for (int j = ja; j < jb; j += 16) {
      double a[16], x[16];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int jj = j + u;
        const bool ok = jj < jb;
        a[u] = ok ? Xc[(long long)jj * T.npiv] : 0.0;
        x[u] = ok ? w[jj] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 16; u++) acc += a[u] * x[u];
    }
  }
  red[q][lane] = acc;
  __syncthreads();
  if (q == 0 && r < T.npiv && r >= j0)
    P[T.poff + (long long)cc * T.npiv + r] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

// y of the chunk's 256 columns from the partial sums (chunks in order; chunk
// row 0 stores them in Y), then partial sums of L21 y: workgroup (row
// chunk, column chunk cc, front) over its 64 boundary rows, kept in Q(cc, r)
__global__ __launch_bounds__(256) void k_mf_fwd_l21(const MfFront* __restrict__ F, int f0, const double* __restrict__ P,
                                                    const double* __restrict__ L, double* __restrict__ Y,
                                                    double* __restrict__ Q) {
  __shared__ double ys[256];
  __shared__ double red[4][64];
  const MfFront T = F[f0 + blockIdx.z];
  const int cc = blockIdx.y, j0 = 256 * cc;
  if (j0 >= T.npiv) return;
  const int r0 = 64 * blockIdx.x;
  if (blockIdx.x > 0 && r0 >= T.nb) return;
  const int t = threadIdx.x;
  {
    const int j = j0 + t;
    double y = 0.0;
    if (j < T.npiv)
      for (int c = 0; c <= j / 256; c++) y += P[T.poff + (long long)c * T.npiv + j];
    ys[t] = y;
    if (blockIdx.x == 0 && j < T.npiv) Y[T.p0 + j] = y;
  }
  __syncthreads();
  const int lane = t & 63, q = t >> 6;
  const int r = r0 + lane;
  const int ja = j0 + 64 * q, jb = min(ja + 64, T.npiv);
  double acc = 0.0;
  if (r < T.nb) {
    const double* __restrict__ Lc = L + T.loff + T.npiv + r;
    for (int j = ja; j < jb; j += 16) {
      double a[16];
#pragma unroll
      for (int u = 0; u < 16; u++) a[u] = j + u < jb ? Lc[(long long)(j + u) * T.m] : 0.0;
#pragma unroll
      for (int u = 0; u < 16; u++) acc += a[u] * ys[j + u - j0];   // (j + u - j0 < 256; a[u] = 0 past jb)
    }
  }
  red[q][lane] = acc;
  __syncthreads();
  if (q == 0 && r < T.nb)
    Q[T.qoff + (long long)cc * T.nb + r] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

// the update vector: w_B -= sum over the column chunks of Q (in order)
__global__ __launch_bounds__(256) void k_mf_fwd_reduce(const MfFront* __restrict__ F, int f0, const double* __restrict__ Q,
                                                       double* __restrict__ W) {
  const MfFront T = F[f0 + blockIdx.y];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= T.nb) return;
  const int ncc = (T.npiv + 255) / 256;
  double acc = 0.0;
  for (int c = 0; c < ncc; c++) acc += Q[T.qoff + (long long)c * T.nb + r];
  W[T.woff + T.npiv + r] -= acc;
}

// y_j -= sum_r L(npiv + r, j) x(B_r) for the pivot columns j, a wavefront a
// column
__global__ __launch_bounds__(256) void k_mf_bwd_bound(const MfFront* __restrict__ F, int f0, const double* __restrict__ L,
                                                      const int* __restrict__ bidx, const double* __restrict__ X,
                                                      double* __restrict__ Y) {
  const MfFront T = F[f0 + blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= T.npiv || T.nb == 0) return;
  const double* __restrict__ Lc = L + T.loff + (long long)j * T.m + T.npiv;
  const int* __restrict__ bi = bidx + T.boff;
  double acc = 0.0;
  for (int r0 = 0; r0 < T.nb; r0 += 4 * 64) {
    double a[4], x[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int r = r0 + u * 64 + lane;
      const bool ok = r < T.nb;
      a[u] = ok ? Lc[r] : 0.0;
      x[u] = ok ? X[bi[r]] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) acc += a[u] * x[u];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if (lane == 0) Y[T.p0 + j] -= acc;
}

// The same with the front's boundary values x(B) staged in LDS once per
// workgroup of 16 wavefronts (16 columns), so that the column loop reads
// L21 and LDS only (the boundary rows' values are gathered once, not once a
// column); dynamic LDS of the level's largest boundary
__global__ __launch_bounds__(1024) void k_mf_bwd_bound_lds(const MfFront* __restrict__ F, int f0,
                                                           const double* __restrict__ L, const int* __restrict__ bidx,
                                                           const double* __restrict__ X, double* __restrict__ Y) {
  extern __shared__ double xb[];
  const MfFront T = F[f0 + blockIdx.y];
  const int j0 = blockIdx.x * 16;
  if (j0 >= T.npiv || T.nb == 0) return;
  const int* __restrict__ bi = bidx + T.boff;
  for (int r = threadIdx.x; r < T.nb; r += 1024) xb[r] = X[bi[r]];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int j = j0 + (threadIdx.x >> 6);
  if (j >= T.npiv) return;
  const double* __restrict__ Lc = L + T.loff + (long long)j * T.m + T.npiv;
  double acc = 0.0;
  for (int r0 = 0; r0 < T.nb; r0 += 4 * 64) {
    double a[4], x[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int r = r0 + u * 64 + lane;
      const bool ok = r < T.nb;
      a[u] = ok ? Lc[r] : 0.0;
      x[u] = ok ? xb[r] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) acc += a[u] * x[u];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if (lane == 0) Y[T.p0 + j] -= acc;
}

// x_j = sum_{i >= j} X(i, j) y_i, a wavefront a column of X
__global__ __launch_bounds__(256) void k_mf_bwd_x(const MfFront* __restrict__ F, int f0, const double* __restrict__ Xs,
                                                  const double* __restrict__ Y, double* __restrict__ Xo) {
  const MfFront T = F[f0 + blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= T.npiv) return;
  const double* __restrict__ Xc = Xs + T.xoff + (long long)j * T.npiv;
  const double* __restrict__ y = Y + T.p0;
  double acc = 0.0;
  for (int i0 = j; i0 < T.npiv; i0 += 4 * 64) {
    double a[4], v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int i = i0 + u * 64 + lane;
      const bool ok = i < T.npiv;
      a[u] = ok ? Xc[i] : 0.0;
      v[u] = ok ? y[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) acc += a[u] * v[u];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if (lane == 0) Xo[T.p0 + j] = acc;
}

template <class Fn>
void over_y(int nf, Fn fn) {
  for (int y0 = 0; y0 < nf; y0 += kMaxGridY) fn(y0, std::min(kMaxGridY, nf - y0));
}

}  // namespace

void mf_free(fsx_ctx* ctx) {
  MfState* S = ctx->mf;
  if (!S) return;
  for (void* p : {(void*)S->F, (void*)S->cmap, (void*)S->bidx, (void*)S->perm, (void*)S->adest, (void*)S->asrc,
                  (void*)S->ws[0], (void*)S->ws[1], (void*)S->L, (void*)S->invc, (void*)S->itask, (void*)S->Xs, (void*)S->P, (void*)S->xlist,
                  (void*)S->W, (void*)S->Y, (void*)S->X, (void*)S->bp, (void*)S->flag})
    if (p) dev_free(ctx, p);
  delete S;
  ctx->mf = nullptr;
}

// The order, the fronts and the assembly lists (host), then the device
// buffers, for the current Z pattern and pressure coordinates.
static int mf_plan(fsx_ctx* ctx) {
  mf_free(ctx);
  const int n = ctx->NNp;
  std::vector<int> rp(n + 1), ci(ctx->nnzZ);
  std::vector<double> xyz(3 * (size_t)n);
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  FSX_HIP(hipMemcpy(rp.data(), ctx->Zrow, rp.size() * sizeof(int), hipMemcpyDeviceToHost));
  FSX_HIP(hipMemcpy(ci.data(), ctx->Zcol, ci.size() * sizeof(int), hipMemcpyDeviceToHost));
  FSX_HIP(hipMemcpy(xyz.data(), ctx->xc, xyz.size() * sizeof(double), hipMemcpyDeviceToHost));
  NdPlan P;
  if (nd_plan(n, rp.data(), ci.data(), xyz.data(), kLeaf, &P) != 0)
    return fail(ctx, FSX_EINVAL, "direct solve: nested dissection of Z's pattern failed");
  MfState* S = new MfState();
  ctx->mf = S;
  S->n = n;
  S->depth = P.depth;
  S->nnz_l = P.nnz_l;
  S->flops = P.flops;
  const int nf = (int)P.fronts.size();
  S->nfront = nf;
  // device order: by level, creation order within a level
  std::vector<int> did(nf);
  S->lev.assign(P.depth + 1, MfLevel{});
  {
    std::vector<int> cnt(P.depth + 1, 0);
    for (const NdFront& f : P.fronts) cnt[f.depth]++;
    int at = 0;
    for (int d = 0; d <= P.depth; d++) { S->lev[d].f0 = at; S->lev[d].nf = cnt[d]; at += cnt[d]; }
    std::vector<int> fill(P.depth + 1, 0);
    for (int t = 0; t < nf; t++) { const int d = P.fronts[t].depth; did[t] = S->lev[d].f0 + fill[d]++; }
  }
  std::vector<MfFront> F(nf);
  std::vector<int> cmap, bidx;
  S->hm.assign(nf, 0);
  S->hpiv.assign(nf, 0);
  long long loff = 0, ioff = 0, woff = 0, xoff = 0, poff = 0;
  std::vector<int2> itask;
  for (int t = 0; t < nf; t++) {
    const NdFront& f = P.fronts[t];
    MfFront& g = F[did[t]];
    g.npiv = f.npiv; g.nb = (int)f.bnd.size(); g.m = g.npiv + g.nb; g.p0 = f.p0;
    g.child[0] = f.child[0] >= 0 ? did[f.child[0]] : -1;
    g.child[1] = f.child[1] >= 0 ? did[f.child[1]] : -1;
    MfLevel& L = S->lev[f.depth];
    g.off = L.ws;
    L.ws += (long long)g.m * g.m;
    L.maxpiv = std::max(L.maxpiv, g.npiv);
    L.maxm = std::max(L.maxm, g.m);
    g.loff = loff; loff += (long long)g.m * g.npiv;
    const int nblk = (g.npiv + kSB - 1) / kSB;
    g.ioff = ioff; ioff += (long long)nblk * kSB * kSB;
    for (int s = 0; s < nblk; s++) itask.push_back(make_int2(did[t], s));
    g.woff = woff; woff += g.m;
    g.xoff = xoff; xoff += (long long)g.npiv * g.npiv;
    const long long ncc = (g.npiv + 255) / 256;
    g.poff = poff; poff += ncc * g.npiv;
    g.qoff = poff; poff += ncc * g.nb;
    L.maxnb = std::max(L.maxnb, g.nb);
    g.boff = (long long)bidx.size();
    bidx.insert(bidx.end(), f.bnd.begin(), f.bnd.end());
    S->hm[did[t]] = g.m;
    S->hpiv[did[t]] = g.npiv;
  }
  // the inversion's steps: fronts by descending block count
  {
    std::vector<int> order(nf);
    for (int t = 0; t < nf; t++) order[t] = t;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return S->hpiv[a] > S->hpiv[b]; });
    S->xcnt.clear();
    const int maxblk = nf ? (S->hpiv[order[0]] + kSB - 1) / kSB : 0;
    for (int st = 0; st < maxblk; st++) {
      int c = 0;
      while (c < nf && S->hpiv[order[c]] > st * kSB) c++;
      S->xcnt.push_back(c);
    }
    S->maxpiv = nf ? S->hpiv[order[0]] : 0;
    S->xorder = order;
  }
  // children's rows in their parents
  for (int t = 0; t < nf; t++) {
    const NdFront& f = P.fronts[t];
    MfFront& g = F[did[t]];
    for (int s = 0; s < 2; s++) {
      g.coff[s] = 0;
      if (f.child[s] < 0) continue;
      const NdFront& c = P.fronts[f.child[s]];
      g.coff[s] = (long long)cmap.size();
      MfLevel& Lc = S->lev[c.depth];
      Lc.maxnb_child[s] = std::max(Lc.maxnb_child[s], (int)c.bnd.size());
      for (int k : c.bnd) {
        int pos;
        if (k < f.p0 + f.npiv) {
          pos = k - f.p0;
        } else {
          const auto it = std::lower_bound(f.bnd.begin(), f.bnd.end(), k);
          pos = (it != f.bnd.end() && *it == k) ? f.npiv + (int)(it - f.bnd.begin()) : -1;
        }
        if (pos < 0) return fail(ctx, FSX_ESTATE, "direct solve: a child's boundary row is not in its parent");
        cmap.push_back(pos);
      }
    }
  }
  // Z's lower entries of every pivot column, level by level (the value from
  // the row of the later node)
  std::vector<long long> adest;
  std::vector<int> asrc;
  adest.reserve((size_t)ctx->nnzZ / 2 + n);
  asrc.reserve((size_t)ctx->nnzZ / 2 + n);
  std::vector<std::vector<int>> by_level(P.depth + 1);
  for (int t = 0; t < nf; t++) by_level[P.fronts[t].depth].push_back(t);
  for (int d = 0; d <= P.depth; d++) {
    S->lev[d].a0 = (long long)adest.size();
    for (int t : by_level[d]) {
      const NdFront& f = P.fronts[t];
      const MfFront& g = F[did[t]];
      for (int lj = 0; lj < f.npiv; lj++) {
        const int c = f.p0 + lj, v = P.perm[c];
        for (int e = rp[v]; e < rp[v + 1]; e++) {
          const int w = ci[e], k = P.iperm[w];
          if (k < c) continue;
          int li;
          if (k < f.p0 + f.npiv) {
            li = k - f.p0;
          } else {
            const auto it = std::lower_bound(f.bnd.begin(), f.bnd.end(), k);
            if (it == f.bnd.end() || *it != k) return fail(ctx, FSX_ESTATE, "direct solve: a row of Z is not in its front");
            li = f.npiv + (int)(it - f.bnd.begin());
          }
          int src = e;
          if (w != v) {
            const int* lo = ci.data() + rp[w];
            const int* hi = ci.data() + rp[w + 1];
            const int* it = std::lower_bound(lo, hi, v);
            if (it != hi && *it == v) src = (int)(it - ci.data());
          }
          adest.push_back(g.off + (long long)li + (long long)g.m * lj);
          asrc.push_back(src);
        }
      }
    }
    S->lev[d].na = (long long)adest.size() - S->lev[d].a0;
  }
  // device memory
  long long wsmax = 1;
  for (const MfLevel& L : S->lev) wsmax = std::max(wsmax, L.ws);
  S->l_entries = (double)loff;
  const double need = 8.0 * (2.0 * wsmax + (double)loff + (double)ioff + (double)xoff + (double)poff + (double)woff + 4.0 * n) +
                      12.0 * (double)adest.size() + 4.0 * (double)(cmap.size() + bidx.size());
  size_t freeb = 0, totalb = 0;
  FSX_HIP(hipMemGetInfo(&freeb, &totalb));
  if (need > 0.9 * (double)freeb)
    return fail(ctx, FSX_ENOMEM, "direct solve: the multifrontal factor needs " + std::to_string(need / 1e9) +
                                     " GB of device memory, " + std::to_string(freeb / 1e9) + " GB are free");
  FSX_TRY(alloc(ctx, &S->F, (size_t)nf));
  FSX_TRY(alloc(ctx, &S->cmap, std::max<size_t>(1, cmap.size())));
  FSX_TRY(alloc(ctx, &S->bidx, std::max<size_t>(1, bidx.size())));
  FSX_TRY(alloc(ctx, &S->perm, (size_t)n));
  FSX_TRY(alloc(ctx, &S->adest, std::max<size_t>(1, adest.size())));
  FSX_TRY(alloc(ctx, &S->asrc, std::max<size_t>(1, asrc.size())));
  FSX_TRY(alloc(ctx, &S->ws[0], (size_t)wsmax));
  FSX_TRY(alloc(ctx, &S->ws[1], (size_t)wsmax));
  FSX_TRY(alloc(ctx, &S->L, std::max<size_t>(1, (size_t)loff)));
  FSX_TRY(alloc(ctx, &S->invc, std::max<size_t>(1, (size_t)ioff)));
  FSX_TRY(alloc(ctx, &S->Xs, std::max<size_t>(1, (size_t)xoff)));
  FSX_TRY(alloc(ctx, &S->P, std::max<size_t>(1, (size_t)poff)));
  FSX_TRY(alloc(ctx, &S->xlist, std::max<size_t>(1, (size_t)nf)));
  S->x_entries = (double)xoff;
  FSX_TRY(alloc(ctx, &S->itask, std::max<size_t>(1, itask.size())));
  FSX_TRY(alloc(ctx, &S->W, std::max<size_t>(1, (size_t)woff)));
  FSX_TRY(alloc(ctx, &S->Y, (size_t)n));
  FSX_TRY(alloc(ctx, &S->X, (size_t)n));
  FSX_TRY(alloc(ctx, &S->bp, (size_t)n));
  FSX_TRY(alloc(ctx, &S->flag, 1));
  S->ninv = (int)itask.size();
  FSX_HIP(hipMemcpy(S->F, F.data(), F.size() * sizeof(MfFront), hipMemcpyHostToDevice));
  if (!cmap.empty()) FSX_HIP(hipMemcpy(S->cmap, cmap.data(), cmap.size() * sizeof(int), hipMemcpyHostToDevice));
  if (!bidx.empty()) FSX_HIP(hipMemcpy(S->bidx, bidx.data(), bidx.size() * sizeof(int), hipMemcpyHostToDevice));
  FSX_HIP(hipMemcpy(S->perm, P.perm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
  if (!adest.empty()) {
    FSX_HIP(hipMemcpy(S->adest, adest.data(), adest.size() * sizeof(long long), hipMemcpyHostToDevice));
    FSX_HIP(hipMemcpy(S->asrc, asrc.data(), asrc.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  if (!itask.empty()) FSX_HIP(hipMemcpy(S->itask, itask.data(), itask.size() * sizeof(int2), hipMemcpyHostToDevice));
  FSX_HIP(hipMemcpy(S->xlist, S->xorder.data(), (size_t)nf * sizeof(int), hipMemcpyHostToDevice));
  return FSX_OK;
}

// Factor Z once per Z (the prototype factors before the time loop,
// ML:1876-1885): plan, then the levels from the deepest.
int mf_factor(fsx_ctx* ctx) {
  if (ctx->mf && ctx->mf->ready) return FSX_OK;
  const auto t0 = std::chrono::steady_clock::now();
  FSX_TRY(mf_plan(ctx));
  MfState* S = ctx->mf;
  const int zero = 0;
  FSX_HIP(hipMemcpy(S->flag, &zero, sizeof(int), hipMemcpyHostToDevice));
  int cur = 0;
  for (int d = S->depth; d >= 0; d--) {
    const MfLevel& L = S->lev[d];
    double* ws = S->ws[cur];
    const double* wsb = S->ws[cur ^ 1];   // the level below, factored
    FSX_TRY(launch_fill(ctx, ws, 0.0, L.ws));
    if (L.na > 0) {
      const unsigned g = (unsigned)std::min<long long>(4096, (L.na + 255) / 256);
      hipLaunchKernelGGL(k_mf_scatter_z, dim3(g), dim3(256), 0, ctx->stream, L.na, S->adest + L.a0, S->asrc + L.a0,
                         ctx->Z, ws);
    }
    if (d < S->depth) {
      const MfLevel& Lb = S->lev[d + 1];
      for (int slot = 0; slot < 2; slot++) {
        if (Lb.maxnb_child[slot] == 0) continue;
        over_y(L.nf, [&](int y0, int ny) {
          hipLaunchKernelGGL(k_mf_extend, dim3(Lb.maxnb_child[slot], ny), dim3(256), 0, ctx->stream, S->F,
                             L.f0 + y0, slot, S->cmap, wsb, ws);
        });
      }
    }
    FSX_TRY(launch_check(ctx, "k_mf_scatter_z/k_mf_extend"));
    for (int k0 = 0; k0 < L.maxpiv; k0 += kPanel) {
      long long chunks = 0, tiles = 0;
      for (int t = L.f0; t < L.f0 + L.nf; t++) {
        if (S->hpiv[t] <= k0) continue;
        const long long kb = std::min(kPanel, S->hpiv[t] - k0), rows = S->hm[t] - k0 - kb;
        const long long nt = (rows + 63) / 64;
        chunks = std::max(chunks, nt);
        tiles = std::max(tiles, nt * (nt + 1) / 2);
      }
      over_y(L.nf, [&](int y0, int ny) {
        hipLaunchKernelGGL(k_mf_potrf, dim3(ny), dim3(256), 0, ctx->stream, S->F, L.f0 + y0, k0, ws, S->flag);
        if (chunks > 0) {
          hipLaunchKernelGGL(k_mf_trsm, dim3((unsigned)chunks, ny), dim3(64), 0, ctx->stream, S->F, L.f0 + y0, k0, ws);
          hipLaunchKernelGGL(k_mf_syrk, dim3((unsigned)tiles, ny), dim3(256), 0, ctx->stream, S->F, L.f0 + y0, k0, ws);
        }
      });
    }
    FSX_TRY(launch_check(ctx, "k_mf_potrf/trsm/syrk"));
    if (L.maxpiv > 0)
      over_y(L.nf, [&](int y0, int ny) {
        hipLaunchKernelGGL(k_mf_store, dim3(L.maxpiv, ny), dim3(256), 0, ctx->stream, S->F, L.f0 + y0, ws, S->L);
      });
    FSX_TRY(launch_check(ctx, "k_mf_store"));
    cur ^= 1;
  }
  // X = inv(L11) of every front: the diagonal blocks' inverses, then block
  // rows from the identity
  if (S->ninv > 0)
    hipLaunchKernelGGL(k_mf_inv, dim3(S->ninv), dim3(kSB), 0, ctx->stream, S->F, S->itask, S->L, S->invc);
  FSX_TRY(launch_check(ctx, "k_mf_inv"));
  FSX_TRY(launch_fill(ctx, S->Xs, 0.0, (long long)S->x_entries));
  if (S->nfront > 0) hipLaunchKernelGGL(k_mf_xdiag, dim3(S->nfront), dim3(256), 0, ctx->stream, S->F, S->Xs);
  for (int st = 0; st < (int)S->xcnt.size(); st++) {
    const int R1 = (st + 1) * kSB;
    hipLaunchKernelGGL(k_mf_xrow, dim3((unsigned)(2 * (st + 1)), S->xcnt[st]), dim3(256), 0, ctx->stream, S->F,
                       S->xlist, st, S->invc, S->Xs);
    if (st + 1 < (int)S->xcnt.size()) {
      const long long rt = (S->maxpiv - R1 + 63) / 64;
      hipLaunchKernelGGL(k_mf_xupd, dim3((unsigned)(rt * (R1 / 64)), S->xcnt[st + 1]), dim3(256), 0, ctx->stream,
                         S->F, S->xlist, st, S->L, S->Xs);
    }
  }
  FSX_TRY(launch_check(ctx, "k_mf_xrow/k_mf_xupd"));
  int flag = 0;
  FSX_HIP(hipMemcpyAsync(&flag, S->flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FSX_HIP(hipStreamSynchronize(ctx->stream));
  if (flag) return fail(ctx, FSX_ENUMERIC, "direct solve: Cholesky of Z met a non-positive pivot");
  S->ready = true;
  S->factor_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return FSX_OK;
}

// x = Z^-1 b with the kept factor (ML:2021-2030)
int mf_apply(fsx_ctx* ctx, const double* b, double* x) {
  MfState* S = ctx->mf;
  if (!S || !S->ready) return fail(ctx, FSX_ESTATE, "direct solve: no factor");
  const int n = S->n;
  hipLaunchKernelGGL(k_mf_permute, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, S->perm, b, S->bp);
  for (int d = S->depth; d >= 0; d--) {
    const MfLevel& L = S->lev[d];
    if (L.nf <= kWideLevel) {   // few, large fronts: every entry its own lane
      hipLaunchKernelGGL(k_mf_fwd_base, dim3((unsigned)((L.maxm + 255) / 256), L.nf), dim3(256), 0, ctx->stream, S->F,
                         L.f0, S->bp, S->W);
      if (d < S->depth)
        for (int slot = 0; slot < 2; slot++) {
          const int nbc = S->lev[d + 1].maxnb_child[slot];
          if (nbc > 0)
            hipLaunchKernelGGL(k_mf_fwd_child, dim3((unsigned)((nbc + 255) / 256), L.nf), dim3(256), 0, ctx->stream,
                               S->F, L.f0, slot, S->cmap, S->W);
        }
    } else {
      over_y(L.nf, [&](int z0, int nz) {
        hipLaunchKernelGGL(k_mf_fwd_init, dim3(nz), dim3(256), 0, ctx->stream, S->F, L.f0 + z0, S->cmap, S->bp, S->W);
      });
    }
    if (L.maxpiv > 0) {
      const unsigned rch = (unsigned)((L.maxpiv + 63) / 64), cch = (unsigned)((L.maxpiv + 255) / 256);
      const unsigned bch = (unsigned)std::max(1, (L.maxnb + 63) / 64);
      over_y(L.nf, [&](int z0, int nz) {
        hipLaunchKernelGGL(k_mf_fwd_x, dim3(rch, cch, nz), dim3(256), 0, ctx->stream, S->F, L.f0 + z0, S->W, S->Xs, S->P);
        hipLaunchKernelGGL(k_mf_fwd_l21, dim3(bch, cch, nz), dim3(256), 0, ctx->stream, S->F, L.f0 + z0, S->P, S->L,
                           S->Y, S->P);
      });
    }
    if (L.maxnb > 0 && L.maxpiv > 0)
      over_y(L.nf, [&](int z0, int nz) {
        hipLaunchKernelGGL(k_mf_fwd_reduce, dim3((unsigned)((L.maxnb + 255) / 256), nz), dim3(256), 0, ctx->stream,
                           S->F, L.f0 + z0, S->P, S->W);
      });
  }
  for (int d = 0; d <= S->depth; d++) {
    const MfLevel& L = S->lev[d];
    if (L.maxpiv == 0) continue;
    over_y(L.nf, [&](int y0, int ny) {
      if (L.maxnb > 0 && (size_t)L.maxnb * sizeof(double) <= kMaxBoundLds)
        hipLaunchKernelGGL(k_mf_bwd_bound_lds, dim3((L.maxpiv + 15) / 16, ny), dim3(1024),
                           (size_t)L.maxnb * sizeof(double), ctx->stream, S->F, L.f0 + y0, S->L, S->bidx, S->X, S->Y);
      else if (L.maxnb > 0)
        hipLaunchKernelGGL(k_mf_bwd_bound, dim3((L.maxpiv + 3) / 4, ny), dim3(256), 0, ctx->stream, S->F, L.f0 + y0,
                           S->L, S->bidx, S->X, S->Y);
      hipLaunchKernelGGL(k_mf_bwd_x, dim3((L.maxpiv + 3) / 4, ny), dim3(256), 0, ctx->stream, S->F, L.f0 + y0, S->Xs,
                         S->Y, S->X);
    });
  }
  hipLaunchKernelGGL(k_mf_unpermute, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, S->perm, S->X, x);
  return launch_check(ctx, "multifrontal solve");
}

int mf_info(const fsx_ctx* ctx, double* l_bytes, double* flops, int* fronts, int* depth, double* factor_ms) {
  const MfState* S = ctx->mf;
  const bool ok = S && S->ready;
  if (l_bytes) *l_bytes = ok ? 8.0 * S->l_entries : 0.0;
  if (flops) *flops = ok ? S->flops : 0.0;
  if (fronts) *fronts = ok ? S->nfront : 0;
  if (depth) *depth = ok ? S->depth : -1;
  if (factor_ms) *factor_ms = ok ? S->factor_ms : 0.0;
  return FSX_OK;
}

void mf_invalidate(fsx_ctx* ctx) {
  if (ctx->mf) ctx->mf->ready = false;
}

}  // namespace fsx
