// Cell binning as a counting sort: per-cell histogram (integer atomics give
// each atom a provisional slot), device exclusive scan to cell offsets,
// scatter, then an in-place sort of every cell segment by atom index so the
// cell lists are exactly the index-ordered lists of the reference's CPU
// mapping (core/grid.impala:280-321; cell math core/grid.impala:406-426).
#include "tmd_internal.hpp"
#include "tmd_device.hpp"

namespace tmd {

struct BinArgs {
    double xmin, ymin, zmin, spacing;
    int nx, ny, ncells, nall;
};

__global__ void k_bin_count(const double4* __restrict__ pos, BinArgs a, int* __restrict__ cell_of, int* __restrict__ slot,
                            int* __restrict__ count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nall) return;
    const double4 p = pos[i];
    const double fx = floor((p.x - a.xmin) / a.spacing);
    const double fy = floor((p.y - a.ymin) / a.spacing);
    const double fz = floor((p.z - a.zmin) / a.spacing);
    int c = -1;
    // non-finite or absurd coordinates are treated as outside the grid
    if (fabs(fx) < 1e9 && fabs(fy) < 1e9 && fabs(fz) < 1e9) {
        const int ci = (int)fx, cj = (int)fy, ck = (int)fz;
        const long long flat = ((long long)ck * a.ny + cj) * a.nx + ci;
        if (flat >= 0 && flat < a.ncells) c = (int)flat;
    }
    cell_of[i] = c;
    slot[i] = (c >= 0) ? atomicAdd(&count[c], 1) : -1;
}

__global__ void k_bin_scatter(const int* __restrict__ cell_of, const int* __restrict__ slot, const int* __restrict__ start,
                              int nall, int* __restrict__ parts) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nall) return;
    const int c = cell_of[i];
    if (c >= 0) parts[start[c] + slot[i]] = i;
}

// One thread per cell: insertion sort of the (short) cell segment, then the
// number of locals in it (locals have the lowest indices, so they lead every
// sorted segment).
__global__ void k_bin_sort(const int* __restrict__ start, int ncells, int n, int* __restrict__ parts,
                           int* __restrict__ local_count) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const int b = start[c], e = start[c + 1];
    for (int k = b + 1; k < e; ++k) {
        const int v = parts[k];
        int m = k - 1;
        while (m >= b && parts[m] > v) {
            parts[m + 1] = parts[m];
            --m;
        }
        parts[m + 1] = v;
    }
    int nl = 0;
    while (b + nl < e && parts[b + nl] < n) ++nl;
    local_count[c] = nl;
}

// Work slots: the locals in cell order.  Neighbour-list rows and the force
// kernels walk atoms in this order (coalesced list reads, cell-local
// gathers); storage keeps the reference's atom index order.
__global__ void k_work_fill(const int* __restrict__ start, const int* __restrict__ work_start, int ncells,
                            const int* __restrict__ parts, int* __restrict__ work) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const int b = start[c], wb = work_start[c], nl = work_start[c + 1] - wb;
    for (int a = 0; a < nl; ++a) work[wb + a] = parts[b + a];
}

void bin_particles(Sim& s) {
    if (has_error()) return;
    const int nall = s.n + s.ng;
    const Geom& G = s.g;
    s.cell_count.reserve(sizeof(int) * (size_t)G.ncells, false, s.stream);
    s.cell_start.reserve(sizeof(int) * ((size_t)G.ncells + 1), false, s.stream);
    s.work_start.reserve(sizeof(int) * ((size_t)G.ncells + 1), false, s.stream);
    s.work.reserve(sizeof(int) * ((size_t)s.n + 1), false, s.stream);
    if (has_error()) return;
    BinArgs a{G.aabb[0], G.aabb[2], G.aabb[4], G.spacing, G.nx, G.ny, G.ncells, nall};
    const int nt = 256;
    TMD_HIP(hipMemsetAsync(s.cell_count.p, 0, sizeof(int) * (size_t)G.ncells, s.stream));
    if (nall > 0)
        hipLaunchKernelGGL(k_bin_count, dim3(div_up(nall, nt)), dim3(nt), 0, s.stream, (const double4*)s.pos.as<double4>(), a,
                           s.cell_of.as<int>(), s.cell_slot.as<int>(), s.cell_count.as<int>());
    exclusive_scan_int(s.cell_count.as<int>(), s.cell_start.as<int>(), G.ncells, s.cell_start.as<int>() + G.ncells, s);
    if (nall > 0)
        hipLaunchKernelGGL(k_bin_scatter, dim3(div_up(nall, nt)), dim3(nt), 0, s.stream, (const int*)s.cell_of.as<int>(),
                           (const int*)s.cell_slot.as<int>(), (const int*)s.cell_start.as<int>(), nall, s.cell_parts.as<int>());
    hipLaunchKernelGGL(k_bin_sort, dim3(div_up(G.ncells, nt)), dim3(nt), 0, s.stream, (const int*)s.cell_start.as<int>(), G.ncells,
                       s.n, s.cell_parts.as<int>(), s.cell_count.as<int>());
    exclusive_scan_int(s.cell_count.as<int>(), s.work_start.as<int>(), G.ncells, s.work_start.as<int>() + G.ncells, s);
    hipLaunchKernelGGL(k_work_fill, dim3(div_up(G.ncells, nt)), dim3(nt), 0, s.stream, (const int*)s.cell_start.as<int>(),
                       (const int*)s.work_start.as<int>(), G.ncells, (const int*)s.cell_parts.as<int>(), s.work.as<int>());
    TMD_HIP(hipGetLastError());
}

}  // namespace tmd
