// Device-side helpers: 64-lane wavefront scans and reductions.
#pragma once

#include <hip/hip_runtime.h>

namespace tmd {

constexpr int kWave = 64;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;  // 1024 items per block

__device__ __forceinline__ int wave_inclusive_scan(int v) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        int t = __shfl_up(v, d, kWave);
        if (lane >= d) v += t;
    }
    return v;
}

// Exclusive scan over the block (blockDim.x == kScanThreads); returns the
// exclusive prefix of v and the block total in *total.
__device__ __forceinline__ int block_exclusive_scan(int v, int* total) {
    __shared__ int wsum[kScanThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    int inc = wave_inclusive_scan(v);
    if (lane == kWave - 1) wsum[wid] = inc;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / kWave; ++w) {
        int s = wsum[w];
        if (w < wid) off += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

__device__ __forceinline__ int block_sum_int(int v) {
    int tot;
    (void)block_exclusive_scan(v, &tot);
    return tot;
}

// Deterministic block reduction of doubles (fixed shuffle tree, fixed wave
// order).  Valid in thread 0.
template <int NT>
__device__ __forceinline__ double block_sum_double(double v) {
    __shared__ double wsum[NT / kWave];
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_down(v, d, kWave);
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    if (lane == 0) wsum[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < NT / kWave; ++w) r += wsum[w];
    }
    __syncthreads();
    return r;
}

}  // namespace tmd
