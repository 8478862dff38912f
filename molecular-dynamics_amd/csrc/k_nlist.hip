// Verlet neighbour lists (full and half) from the cell grid.
//
// One 64-lane wavefront per cell.  The 27-cell stencil is 9 rows of 3 cells
// that are contiguous in the cell CSR (cells are flattened x-fastest), so the
// candidates of every atom of the cell are 9 contiguous ranges, staged once
// into LDS (index + FP64 position).  The wave then walks the cell's atoms one
// at a time and tests 64 candidates per pass, compacting the accepted ones
// with a ballot so that every list keeps the reference's order: stencil
// (z, y, x) then cell order (core/neighborlist.impala:1-78).  Distances are
// computed in FP64 without contraction, so the sets are bit-exact with the
// reference's `rsq <= r_nb^2` test.
//
// List layout ("tiled column-major"): work slot w (locals in cell order) owns
// column (w & 63) of tile (w >> 6); entry k of w is at
//   tile * cap * 64 + k * 64 + (w & 63)
// so a wavefront of 64 consecutive slots reads 256 contiguous bytes per k.
#include "tmd_internal.hpp"
#include "tmd_device.hpp"

namespace tmd {

constexpr int kNbStage = 768;

struct NbArgs {
    int n, nx, ny, ncells, cap, half, half_mode;
    double cut2;
};

__device__ __forceinline__ size_t nb_index(int w, int k, int cap) {
    return (size_t)(w >> 6) * (size_t)cap * 64 + (size_t)k * 64 + (size_t)(w & 63);
}

__global__ void __launch_bounds__(64) k_nlist(const double4* __restrict__ pos, const int* __restrict__ start,
                                               const int* __restrict__ parts, const int* __restrict__ work_start,
                                               NbArgs a, int* __restrict__ nb_count, int* __restrict__ nb_list,
                                               int* __restrict__ overflow) {
    const int c = blockIdx.x;
    const int lane = threadIdx.x;
    const int wbeg = work_start[c];
    const int nloc = work_start[c + 1] - wbeg;
    if (nloc == 0) return;

    __shared__ int rbeg[9], roff[10];
    __shared__ int lj[kNbStage];
    __shared__ double lx[kNbStage], ly[kNbStage], lz[kNbStage];

    if (lane < 9) {
        const int oz = lane / 3 - 1, oy = lane % 3 - 1;
        const int mid = c + oz * a.nx * a.ny + oy * a.nx;
        const int lo = max(mid - 1, 0);
        const int hi = min(mid + 1, a.ncells - 1);
        int b = 0, len = 0;
        if (lo <= hi) {
            b = start[lo];
            len = start[hi + 1] - b;
        }
        rbeg[lane] = b;
        roff[lane + 1] = len;
    }
    __syncthreads();
    if (lane == 0) {
        roff[0] = 0;
        for (int r = 0; r < 9; ++r) roff[r + 1] += roff[r];
    }
    __syncthreads();
    const int S = roff[9];
    const bool staged = S <= kNbStage;
    if (staged) {
        for (int t = lane; t < S; t += 64) {
            int r = 0;
            while (t >= roff[r + 1]) ++r;
            const int j = parts[rbeg[r] + (t - roff[r])];
            const double4 p = pos[j];
            lj[t] = j;
            lx[t] = p.x;
            ly[t] = p.y;
            lz[t] = p.z;
        }
    }
    __syncthreads();

    const int cbeg = start[c];
    for (int at = 0; at < nloc; ++at) {
        const int w = wbeg + at;
        const int i = parts[cbeg + at];
        const double4 pi = pos[i];
        int cnt = 0;
        for (int base = 0; base < S; base += 64) {
            const int t = base + lane;
            bool ok = false;
            int j = -1;
            if (t < S) {
                double xj, yj, zj;
                if (staged) {
                    j = lj[t];
                    xj = lx[t]; yj = ly[t]; zj = lz[t];
                } else {
                    int r = 0;
                    while (t >= roff[r + 1]) ++r;
                    j = parts[rbeg[r] + (t - roff[r])];
                    const double4 p = pos[j];
                    xj = p.x; yj = p.y; zj = p.z;
                }
                bool keep = (j != i);
                if (keep && a.half) keep = (a.half_mode == HALF_REF_LITERAL) ? (i > j) : (j >= a.n || i > j);
                if (keep) {
                    const double dx = pi.x - xj, dy = pi.y - yj, dz = pi.z - zj;
                    const double rsq = dx * dx + dy * dy + dz * dz;
                    ok = rsq <= a.cut2;
                }
            }
            const unsigned long long m = __ballot(ok);
            if (ok) {
                const int idx = cnt + __popcll(m & ((1ull << lane) - 1ull));
                if (idx < a.cap) nb_list[nb_index(w, idx, a.cap)] = j;
            }
            cnt += __popcll(m);
        }
        if (lane == 0) {
            nb_count[w] = cnt;
            if (cnt > a.cap) atomicMax(overflow, cnt);
        }
    }
}

void build_neighbor_lists(Sim& s, int half, double cutoff) {
    if (has_error()) return;
    const Geom& G = s.g;
    int* counters = s.counters.as<int>();
    // counters[2] = overflow, counters[3] = number of work slots (binned locals)
    for (;;) {
        const int ntile = div_up(s.n > 0 ? s.n : 1, 64);
        s.nb_stride = ntile * 64;
        s.nb_count.reserve(sizeof(int) * (size_t)s.nb_stride, false, s.stream);
        s.nb_list.reserve(sizeof(int) * (size_t)ntile * 64 * (size_t)s.nb_cap, false, s.stream);
        if (has_error()) return;
        TMD_HIP(hipMemsetAsync(counters + 2, 0, sizeof(int), s.stream));
        NbArgs a{s.n, G.nx, G.ny, G.ncells, s.nb_cap, half, s.half_mode, cutoff * cutoff};
        hipLaunchKernelGGL(k_nlist, dim3(G.ncells), dim3(64), 0, s.stream, (const double4*)s.pos.as<double4>(),
                           (const int*)s.cell_start.as<int>(), (const int*)s.cell_parts.as<int>(),
                           (const int*)s.work_start.as<int>(), a, s.nb_count.as<int>(), s.nb_list.as<int>(), counters + 2);
        TMD_HIP(hipGetLastError());
        int host[2] = {0, 0};
        TMD_HIP(hipMemcpyAsync(&host[0], counters + 2, sizeof(int), hipMemcpyDeviceToHost, s.stream));
        TMD_HIP(hipMemcpyAsync(&host[1], s.work_start.as<int>() + G.ncells, sizeof(int), hipMemcpyDeviceToHost, s.stream));
        TMD_HIP(hipStreamSynchronize(s.stream));
        if (has_error()) return;
        if (host[1] != s.n) {
            char msg[160];
            snprintf(msg, sizeof msg, "%d local atoms lie outside the cell grid (non-finite or escaped positions)", s.n - host[1]);
            set_error(__FILE__, __LINE__, msg);
            return;
        }
        if (host[0] > 0) {
            s.nb_cap = host[0] * 2;
            continue;
        }
        break;
    }
    s.nb_half = half;
    s.nb_valid = 1;
}

}  // namespace tmd
