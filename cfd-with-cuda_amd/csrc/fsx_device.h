// Device helpers shared by the kernels: XCD-aware block mapping, wavefront
// and block reductions, deterministic last-block grid reductions, search in a
// sorted CSR row.
#pragma once

#include <hip/hip_runtime.h>

#include "fsx_internal.h"

namespace fsx {

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).
// Renumber so that each XCD works on one contiguous 1/8 of the logical blocks
// and keeps its rows' gathers in its own L2. gridDim.x is a multiple of 8.
__device__ __forceinline__ int xcd_block() {
  const int b = blockIdx.x, g = gridDim.x;
  return (b & 7) * (g >> 3) + (b >> 3);
}

// Contiguous row range of the logical block `lb` among `g` blocks.
__device__ __forceinline__ void block_rows(long long nrows, int lb, int g, int granule, long long* r0, long long* r1) {
  long long chunk = (nrows + g - 1) / g;
  chunk = (chunk + granule - 1) / granule * granule;
  long long a = (long long)lb * chunk;
  long long b = a + chunk;
  *r0 = a < nrows ? a : nrows;
  *r1 = b < nrows ? b : nrows;
}

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

// Sum N values over the block (fixed tree, deterministic). Result valid in
// thread 0. `red` must hold N * (blockDim.x / 64) doubles.
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double* red) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave, nw = blockDim.x / kWave;
#pragma unroll
  for (int n = 0; n < N; n++) v[n] = group_sum<kWave>(v[n]);
  if (lane == 0)
#pragma unroll
    for (int n = 0; n < N; n++) red[n * nw + wid] = v[n];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int n = 0; n < N; n++) {
      double s = 0.0;
      for (int w = 0; w < nw; w++) s += red[n * nw + w];
      v[n] = s;
    }
  }
  __syncthreads();
}

// Grid-wide deterministic sum of N values: every block deposits its block sum
// in partials[n * gridDim.x + block]; the last block to arrive (counter) sums
// the partials in a fixed order and returns true with the totals in out[] for
// thread 0. The counter is reset by the last block. Must be called by all
// threads of every block.
template <int N>
__device__ bool grid_sum(double (&v)[N], double* partials, unsigned int* counter, double (&out)[N]) {
  __shared__ double red[N * (kBlock / kWave)];
  __shared__ int am_last;
  block_sum<N>(v, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int n = 0; n < N; n++) partials[n * gridDim.x + blockIdx.x] = v[n];
    __threadfence();
    unsigned int t = atomicAdd(counter, 1u);
    am_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!am_last) return false;
  __threadfence();
  double s[N];
#pragma unroll
  for (int n = 0; n < N; n++) {
    s[n] = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
      s[n] += __builtin_nontemporal_load(&partials[n * gridDim.x + b]);
  }
  block_sum<N>(s, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int n = 0; n < N; n++) out[n] = s[n];
    *counter = 0u;
  }
  return true;
}

// Grid-wide max of one non-negative value (order independent).
__device__ inline bool grid_max(double v, double* partials, unsigned int* counter, double* out) {
  __shared__ double red[kBlock / kWave];
  __shared__ int am_last;
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave, nw = blockDim.x / kWave;
  v = wave_max(v);
  if (lane == 0) red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < nw; w++) m = fmax(m, red[w]);
    partials[blockIdx.x] = m;
    __threadfence();
    unsigned int t = atomicAdd(counter, 1u);
    am_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!am_last) return false;
  __threadfence();
  double m = 0.0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) m = fmax(m, __builtin_nontemporal_load(&partials[b]));
  m = wave_max(m);
  if (lane == 0) red[wid] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int w = 0; w < nw; w++) r = fmax(r, red[w]);
    *out = r;
    *counter = 0u;
  }
  return true;
}

// Streaming loads: CSR values and column indices are read once per launch,
// so they are loaded non-temporally and leave L2 to the gathered vectors.
template <class T>
__device__ __forceinline__ T ld_stream(const T* p) {
  return __builtin_nontemporal_load(p);
}

// Partial dot products of one CSR row for the LPR lanes of a row group: lane
// l takes entries start+l, start+l+LPR, ... in rounds of U entries whose loads
// are all issued before the first FMA (memory-level parallelism instead of a
// dependent load chain). Out-of-row slots are clamped to a valid index and
// their products selected away. `g(c, v, acc)` consumes one (column, value).
template <int LPR, int U, class V, class G>
__device__ __forceinline__ void row_foreach(const int* __restrict__ ci, const V* __restrict__ val, int vstride,
                                            int start, int end, int l, G&& g) {
  for (int base = start + l; base < end; base += U * LPR) {
    int c[U];
    V v[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int k = base + j * LPR;
      const int kk = k < end ? k : end - 1;
      c[j] = ld_stream(ci + kk);
      v[j] = k < end ? ld_stream(val + (size_t)vstride * kk) : V(0);
    }
#pragma unroll
    for (int j = 0; j < U; j++) g(c[j], v[j], (base + j * LPR) < end);
  }
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
