// Device exclusive prefix scan (reduce-then-scan over 1024-item tiles).
// Used by the counting-sort binning (cell offsets) and by the ghost
// compaction.  Integer addition is associative, so the result is exact and
// independent of the launch shape.
#include "tmd_internal.hpp"
#include "tmd_device.hpp"

namespace tmd {

__global__ void __launch_bounds__(kScanThreads) k_tile_sums(const int* __restrict__ in, int n, int* __restrict__ tile_sums) {
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v = 0;
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
        int idx = base + it;
        if (idx < n) v += in[idx];
    }
    int tot = block_sum_int(v);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// One block: exclusive scan of the tile sums in place; grand total to *total.
__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(int* __restrict__ tile_sums, int ntiles, int* __restrict__ total) {
    // up to kPer items per thread: one block-wide scan of per-thread sums
    // (a few microseconds less than 256 items at a time for thousands of tiles)
    constexpr int kPer = 16;
    if (ntiles <= kScanThreads * kPer) {
        const int per = (ntiles + kScanThreads - 1) / kScanThreads;
        const int base = (int)threadIdx.x * per;
        int v[kPer];
        int sum = 0;
#pragma unroll
        for (int it = 0; it < kPer; ++it) {
            const int idx = base + it;
            v[it] = (it < per && idx < ntiles) ? tile_sums[idx] : 0;
            sum += v[it];
        }
        int tot;
        int run = block_exclusive_scan(sum, &tot);
#pragma unroll
        for (int it = 0; it < kPer; ++it) {
            const int idx = base + it;
            if (it < per && idx < ntiles) tile_sums[idx] = run;
            run += v[it];
        }
        if (threadIdx.x == 0 && total) *total = tot;
        return;
    }
    int carry = 0;
    for (int base = 0; base < ntiles; base += kScanThreads) {
        int idx = base + threadIdx.x;
        int v = idx < ntiles ? tile_sums[idx] : 0;
        int tot;
        int ex = block_exclusive_scan(v, &tot);
        if (idx < ntiles) tile_sums[idx] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_finish(const int* __restrict__ in, int n, const int* __restrict__ tile_off,
                                                              int* __restrict__ out) {
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int sum = 0;
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
        int idx = base + it;
        v[it] = idx < n ? in[idx] : 0;
        sum += v[it];
    }
    int tot;
    int run = tile_off[blockIdx.x] + block_exclusive_scan(sum, &tot);
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
        int idx = base + it;
        if (idx < n) out[idx] = run;
        run += v[it];
    }
}

void scan_tiles(Sim& s, int* tiles, int ntiles, int* total_dev) {
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, s.stream, tiles, ntiles, total_dev);
    TMD_HIP(hipGetLastError());
}

void scan_exclusive(const int* in, int* out, int n, int* total_dev, int* tiles, hipStream_t st) {
    const int ntiles = div_up(n, kScanTile);
    hipLaunchKernelGGL(k_tile_sums, dim3(ntiles), dim3(kScanThreads), 0, st, in, n, tiles);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, st, tiles, ntiles, total_dev);
    hipLaunchKernelGGL(k_scan_finish, dim3(ntiles), dim3(kScanThreads), 0, st, in, n, (const int*)tiles, out);
    TMD_HIP(hipGetLastError());
}

void exclusive_scan_int(const int* in, int* out, int n, int* total_dev, Sim& s) {
    if (n <= 0) {
        if (total_dev) TMD_HIP(hipMemsetAsync(total_dev, 0, sizeof(int), s.stream));
        return;
    }
    const int ntiles = div_up(n, kScanTile);
    s.scan_blocks.reserve(sizeof(int) * (size_t)ntiles, false, s.stream);
    int* tiles = s.scan_blocks.as<int>();
    hipLaunchKernelGGL(k_tile_sums, dim3(ntiles), dim3(kScanThreads), 0, s.stream, in, n, tiles);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, s.stream, tiles, ntiles, total_dev);
    hipLaunchKernelGGL(k_scan_finish, dim3(ntiles), dim3(kScanThreads), 0, s.stream, in, n, (const int*)tiles, out);
    TMD_HIP(hipGetLastError());
}

}  // namespace tmd
