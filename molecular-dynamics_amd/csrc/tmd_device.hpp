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

// Deterministic block sum of doubles for any blockDim.x multiple of 64
// (fixed shuffle tree, fixed wave order).  Valid in thread 0.
__device__ __forceinline__ double block_sum_double_any(double v) {
    __shared__ double wsum[16];  // up to 1024 threads
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_down(v, d, kWave);
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    if (lane == 0) wsum[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        const int nw = (int)blockDim.x / kWave;
        for (int w = 0; w < nw; ++w) r += wsum[w];
    }
    __syncthreads();
    return r;
}

// Counting-sort slot of a binned atom: lanes of one wavefront that fall in
// the same cell take consecutive slots in lane order with one atomic per
// (wavefront, cell).  c < 0: not binned (slot -1).  Every lane of the
// wavefront must call it.
__device__ __forceinline__ int wave_cell_slot(int c, int* __restrict__ count) {
    const int lane = threadIdx.x & (kWave - 1);
    unsigned long long pending = __ballot(c >= 0);
    // group the lanes by cell first (no memory round trip inside the loop),
    // then every group's leader takes its slots with one atomic, all groups
    // at once
    int leader_of = lane;
    unsigned long long grp_of = 0ull;
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int cl = __shfl(c, leader, kWave);
        const unsigned long long grp = __ballot(c == cl) & pending;
        if ((grp >> lane) & 1ull) {
            leader_of = leader;
            grp_of = grp;
        }
        pending &= ~grp;
    }
    int base = 0;
    if (c >= 0 && lane == leader_of) base = atomicAdd(&count[c], __popcll(grp_of));
    base = __shfl(base, leader_of, kWave);
    return c >= 0 ? base + __popcll(grp_of & ((1ull << lane) - 1ull)) : -1;
}

// Sort the cell segment parts[b, e) (distinct atom indices) ascending with
// one wavefront and set inv[value] = its position.  Segments of up to 64 by
// rank (each lane counts the smaller values); longer ones by lane 0
// (insertion sort).  Returns, in every lane, how many values are below lim
// (the locals, which lead the sorted segment).
__device__ __forceinline__ int wave_sort_cell(int b, int e, int lim, int* __restrict__ parts, int* __restrict__ inv) {
    const int lane = threadIdx.x & (kWave - 1);
    const int m = e - b;
    if (m <= kWave) {
        const int v = lane < m ? parts[b + lane] : 0x7fffffff;
        int rank = 0;
        for (int k = 0; k < m; ++k) rank += (__shfl(v, k, kWave) < v) ? 1 : 0;
        if (lane < m) {
            parts[b + rank] = v;
            inv[v] = b + rank;
        }
        return __popcll(__ballot(lane < m && v < lim));
    }
    int nl = 0;
    if (lane == 0) {
        for (int k = b + 1; k < e; ++k) {
            const int v = parts[k];
            int q = k - 1;
            while (q >= b && parts[q] > v) {
                parts[q + 1] = parts[q];
                --q;
            }
            parts[q + 1] = v;
        }
        while (b + nl < e && parts[b + nl] < lim) ++nl;
        for (int k = b; k < e; ++k) inv[parts[k]] = k;
    }
    return __shfl(nl, 0, kWave);
}

// Streaming (non-temporal) load, for data a launch reads once (list groups):
// it does not displace the lines that neighbouring tiles re-read from L2
// (the staged window positions).
typedef int tmd_v4i __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int4 ld_stream(const int4* p) {
    const tmd_v4i t = __builtin_nontemporal_load(reinterpret_cast<const tmd_v4i*>(p));
    return make_int4(t.x, t.y, t.z, t.w);
}

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each
// with a private L2.  Remap so that XCD x processes one contiguous range of
// logical blocks: atoms that are neighbours in cell order (and so share
// neighbour positions) then share an L2.  A bijection on [0, nblocks).
constexpr int kNumXcd = 8;
__device__ __forceinline__ int xcd_block(int b, int nblocks) {
    const int q = nblocks / kNumXcd, r = nblocks % kNumXcd;
    const int x = b % kNumXcd, k = b / kNumXcd;
    return x * q + min(x, r) + k;
}

}  // namespace tmd
