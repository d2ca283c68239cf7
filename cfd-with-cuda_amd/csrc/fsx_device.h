// Device helpers shared by the kernels: wavefront and block reductions,
// deterministic grid reductions through kernel boundaries, search in a sorted
// CSR row.
#pragma once

#include <hip/hip_runtime.h>

#include "fsx_internal.h"

namespace fsx {

template <int W>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = W / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

template <int W>
__device__ __forceinline__ float group_sumf(float v) {
#pragma unroll
  for (int off = W / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
  return v;
}

// Sum N values over the block in a fixed tree (deterministic); every thread
// receives the totals. `red` holds N * (blockDim.x / 64) doubles of LDS.
template <int N>
__device__ __forceinline__ void block_sum_all(double (&v)[N], double* red) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave, nw = blockDim.x / kWave;
#pragma unroll
  for (int n = 0; n < N; n++) v[n] = group_sum<kWave>(v[n]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int n = 0; n < N; n++) red[n * nw + wid] = v[n];
  __syncthreads();
#pragma unroll
  for (int n = 0; n < N; n++) {
    double s = 0.0;
    for (int w = 0; w < nw; w++) s += red[n * nw + w];
    v[n] = s;
  }
}

// Grid-wide reductions go through a kernel boundary instead of fences and
// atomics (an agent-scope fence writes back the L2 of every XCD and costs
// far more than a launch): a fixed-size grid writes one partial per block,
// and the next kernel reduces the partials redundantly in every block, in a
// fixed order, so all blocks agree bit for bit and runs are reproducible.
constexpr int kRedGrid = 256;

// BS: the block size the kernel is launched with (the LDS scratch is sized
// for BS / 64 wavefronts).
template <int N, int BS = kBlock>
__device__ __forceinline__ void write_partials(double (&v)[N], double* partials) {
  __shared__ double red[N * (BS / kWave)];
  block_sum_all<N>(v, red);
  if (threadIdx.x == 0)
#pragma unroll
    for (int n = 0; n < N; n++) partials[n * kRedGrid + blockIdx.x] = v[n];
}

// n partial sums in a fixed order (thread t adds entries t, t + kBlock, ...,
// then the block tree), for blocks of kBlock threads; every thread receives
// the total. For n = kRedGrid it equals read_partials<1>.
__device__ __forceinline__ double sum_parts(const double* __restrict__ parts, int n) {
  __shared__ double red[kBlock / kWave];
  double a[4] = {0.0, 0.0, 0.0, 0.0};   // four loads in flight a lane
  int i = threadIdx.x;
  for (; i + 3 * kBlock < n; i += 4 * kBlock) {
#pragma unroll
    for (int u = 0; u < 4; u++) a[u] += parts[i + u * kBlock];
  }
  for (int u = 0; i < n; i += kBlock, u++) a[u] += parts[i];
  double v[1] = {(a[0] + a[1]) + (a[2] + a[3])};
  block_sum_all<1>(v, red);
  return v[0];
}

// partials[n * kRedGrid + b], b < kRedGrid, summed in a fixed order by a
// block of BS >= kRedGrid threads (threads past kRedGrid add exact zeros, so
// the sum does not depend on BS).
template <int N, int BS = kBlock>
__device__ __forceinline__ void read_partials(const double* partials, double (&out)[N]) {
  __shared__ double red[N * (BS / kWave)];
#pragma unroll
  for (int n = 0; n < N; n++) out[n] = threadIdx.x < kRedGrid ? partials[n * kRedGrid + threadIdx.x] : 0.0;
  block_sum_all<N>(out, red);
}

__device__ __forceinline__ double block_max_all(double v) {
  __shared__ double red[kBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave, nw = blockDim.x / kWave;
  v = wave_max(v);
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double m = 0.0;
  for (int w = 0; w < nw; w++) m = fmax(m, red[w]);
  return m;
}

// Column source of a CSR matrix: the explicit column array, or the
// dictionary of row patterns (column = row + pat_off[row_pat[row] + k -
// rowptr[row]], row_pat holding where the row's pattern starts in pat_off),
// which replaces a 4-byte column per entry by 4 bytes per row.
struct ColSrc {
  const int* ci;
  const int* row_pat;
  const int* pat_off;
};

// per-row offset that turns an entry index kk into its pattern slot
template <bool PAT>
__device__ __forceinline__ int col_base(const ColSrc& cs, long long r, int rs) {
  if constexpr (PAT) return cs.row_pat[r] - rs;
  else return 0;
}

template <bool PAT>
__device__ __forceinline__ int col_at(const ColSrc& cs, long long r, int base, int kk) {
  if constexpr (PAT) return (int)r + cs.pat_off[base + kk];
  else return cs.ci[kk];
}

// Position of `c` in the sorted columns col[lo, hi); the entry must exist.
__device__ __forceinline__ int find_col(const int* __restrict__ col, int lo, int hi, int c) {
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (col[mid] <= c) lo = mid; else hi = mid;
  }
  return lo;
}

}  // namespace fsx
