// Cell binning as a counting sort: per-cell histogram (integer atomics give
// each atom a provisional slot), device exclusive scan to cell offsets,
// scatter, then an in-place sort of every cell segment by atom index so the
// cell lists are exactly the index-ordered lists of the reference's CPU
// mapping (core/grid.impala:280-321; cell math core/grid.impala:406-426).
#include "tmd_internal.hpp"
#include "tmd_device.hpp"
#include "tmd_window.hpp"

namespace tmd {

struct BinArgs {
    double xmin, ymin, zmin, spacing;
    int nx, ny, ncells, nall;
};

__global__ void k_bin_count(const double4* __restrict__ pos, BinArgs a, int* __restrict__ cell_of, int* __restrict__ slot,
                            int* __restrict__ count, int* __restrict__ inv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = -1;
    if (i < a.nall) {
        const double4 p = pos[i];
        const double fx = floor((p.x - a.xmin) / a.spacing);
        const double fy = floor((p.y - a.ymin) / a.spacing);
        const double fz = floor((p.z - a.zmin) / a.spacing);
        // non-finite or absurd coordinates are treated as outside the grid
        if (fabs(fx) < 1e9 && fabs(fy) < 1e9 && fabs(fz) < 1e9) {
            const int ci = (int)fx, cj = (int)fy, ck = (int)fz;
            const long long flat = ((long long)ck * a.ny + cj) * a.nx + ci;
            if (flat >= 0 && flat < a.ncells) c = (int)flat;
        }
    }
    const int sl = wave_cell_slot(c, count);  // all lanes take part
    if (i >= a.nall) return;
    cell_of[i] = c;
    slot[i] = sl;
    inv[i] = -1;  // set by k_bin_sort for binned atoms
}

__global__ void k_bin_scatter(const int* __restrict__ cell_of, const int* __restrict__ slot, const int* __restrict__ start,
                              int nall, int* __restrict__ parts) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nall) return;
    const int c = cell_of[i];
    if (c >= 0) parts[start[c] + slot[i]] = i;
}

// One thread per cell (sparse grids: a few atoms per cell): insertion sort
// of the cell segment by atom index, then its locals and the CSR index of
// every atom; the same result as k_bin_sort.
__global__ void k_bin_sort_small(const int* __restrict__ start, int ncells, int n, int* __restrict__ parts,
                                 int* __restrict__ local_count, int* __restrict__ inv) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const int b = start[c], e = start[c + 1];
    for (int k = b + 1; k < e; ++k) {
        const int v = parts[k];
        int q = k - 1;
        while (q >= b && parts[q] > v) {
            parts[q + 1] = parts[q];
            --q;
        }
        parts[q + 1] = v;
    }
    int nl = 0;
    while (b + nl < e && parts[b + nl] < n) ++nl;
    local_count[c] = nl;
    for (int k = b; k < e; ++k) inv[parts[k]] = k;
}

// One wavefront per cell: sort the cell segment by atom index, then the
// number of locals in it (locals have the lowest indices, so they lead every
// sorted segment).
__global__ void __launch_bounds__(256) k_bin_sort(const int* __restrict__ start, int ncells, int n, int* __restrict__ parts,
                                                  int* __restrict__ local_count, int* __restrict__ inv) {
    const int c = blockIdx.x * (blockDim.x / kWave) + (int)threadIdx.x / kWave;
    if (c >= ncells) return;  // whole wavefronts
    const int nl = wave_sort_cell(start[c], start[c + 1], n, parts, inv);
    if ((threadIdx.x & (kWave - 1)) == 0) local_count[c] = nl;
}

// Tile work order (see tmd_window.hpp): one thread per tile counts the
// tile's locals and the size of its window (small tiles).
__global__ void k_tile_stats_small(const int* __restrict__ start, const int* __restrict__ local_count, TileGeom g,
                                   int ntiles, int* __restrict__ tile_count, int* __restrict__ max_window,
                                   int* __restrict__ max_locals) {
    const int tile = blockIdx.x * blockDim.x + threadIdx.x;
    int cnt = 0, S = 0;
    if (tile < ntiles) {
        const int tx_i = tile % g.ntx, ty_i = (tile / g.ntx) % g.nty, tz_i = tile / (g.ntx * g.nty);
        const int cx0 = 1 + tx_i * g.tx, cy0 = 1 + ty_i * g.ty, cz0 = 1 + tz_i * g.tz;
        const int cx1 = min(cx0 + g.tx, g.nx - 1), cy1 = min(cy0 + g.ty, g.ny - 1), cz1 = min(cz0 + g.tz, g.nz - 1);
        for (int z = cz0; z < cz1; ++z)
            for (int y = cy0; y < cy1; ++y)
                for (int x = cx0; x < cx1; ++x) cnt += local_count[(z * g.ny + y) * g.nx + x];
        tile_count[tile] = cnt;
        const int xlo = cx0 - 1, xhi = min(cx0 + g.tx, g.nx - 1);
        for (int z = cz0 - 1; z <= cz0 + g.tz; ++z)
            for (int y = cy0 - 1; y <= cy0 + g.ty; ++y)
                if (y < g.ny && z < g.nz) {
                    const int row = (z * g.ny + y) * g.nx;
                    S += start[row + xhi + 1] - start[row + xlo];
                }
        if (cnt == 0) S = 0;  // empty tiles launch no block
    }
    // one atomic per wavefront for each maximum
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        S = max(S, __shfl_down(S, d, kWave));
        cnt = max(cnt, __shfl_down(cnt, d, kWave));
    }
    if ((threadIdx.x & (kWave - 1)) == 0 && cnt > 0) {
        atomicMax(max_window, S);
        atomicMax(max_locals, cnt);
    }
}

// The same with one wavefront per tile (large tiles: lanes stride over the
// cells and the window rows; fixed-order wavefront sums).
__global__ void __launch_bounds__(kWave) k_tile_stats(const int* __restrict__ start, const int* __restrict__ local_count,
                                                      TileGeom g, int ntiles, int* __restrict__ tile_count,
                                                      int* __restrict__ max_window, int* __restrict__ max_locals) {
    const int tile = blockIdx.x;
    if (tile >= ntiles) return;
    const int lane = threadIdx.x;
    const int tx_i = tile % g.ntx, ty_i = (tile / g.ntx) % g.nty, tz_i = tile / (g.ntx * g.nty);
    const int cx0 = 1 + tx_i * g.tx, cy0 = 1 + ty_i * g.ty, cz0 = 1 + tz_i * g.tz;
    const int cx1 = min(cx0 + g.tx, g.nx - 1), cy1 = min(cy0 + g.ty, g.ny - 1), cz1 = min(cz0 + g.tz, g.nz - 1);
    const int ex = cx1 - cx0, ey = cy1 - cy0, ez = cz1 - cz0;
    int cnt = 0;
    for (int q = lane; q < ex * ey * ez; q += kWave)
        cnt += local_count[((cz0 + q / (ex * ey)) * g.ny + (cy0 + (q / ex) % ey)) * g.nx + (cx0 + q % ex)];
    int S = 0;
    const int xlo = cx0 - 1, xhi = min(cx0 + g.tx, g.nx - 1);
    const int wy = g.ty + 2;
    for (int q = lane; q < wy * (g.tz + 2); q += kWave) {
        const int y = cy0 - 1 + q % wy, z = cz0 - 1 + q / wy;
        if (y < g.ny && z < g.nz) {
            const int row = (z * g.ny + y) * g.nx;
            S += start[row + xhi + 1] - start[row + xlo];
        }
    }
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        cnt += __shfl_down(cnt, d, kWave);
        S += __shfl_down(S, d, kWave);
    }
    if (lane == 0) {
        tile_count[tile] = cnt;
        // read first: only the few tiles that raise a maximum issue the atomic
        // (thousands of same-address atomics from separate wavefronts serialise)
        if (cnt > 0 && S > *(volatile int*)max_window) atomicMax(max_window, S);
        if (cnt > 0 && cnt > *(volatile int*)max_locals) atomicMax(max_locals, cnt);
    }
}

// Work order: one block per tile writes the tile's locals in (z, y, x)
// cell order (lanes of a wavefront are then atoms of adjacent cells, whose
// neighbour windows overlap -- measured faster than grouping lanes by list
// length), with each atom's mirror index and cell.
__global__ void __launch_bounds__(256) k_tile_fill_block(const int* __restrict__ start, const int* __restrict__ local_count,
                                                         const int* __restrict__ tile_start, TileGeom g,
                                                         const int* __restrict__ parts, int* __restrict__ work,
                                                         int* __restrict__ work_k, int* __restrict__ work_c) {
    __shared__ int coff[kMaxTile * kMaxTile * kMaxTile + 1];
    const int tile = blockIdx.x;
    const int w0 = tile_start[tile], nloc = tile_start[tile + 1] - w0;
    if (nloc == 0) return;
    const int tx_i = tile % g.ntx, ty_i = (tile / g.ntx) % g.nty, tz_i = tile / (g.ntx * g.nty);
    const int cx0 = 1 + tx_i * g.tx, cy0 = 1 + ty_i * g.ty, cz0 = 1 + tz_i * g.tz;
    const int cx1 = min(cx0 + g.tx, g.nx - 1), cy1 = min(cy0 + g.ty, g.ny - 1), cz1 = min(cz0 + g.tz, g.nz - 1);
    const int ex = cx1 - cx0, ey = cy1 - cy0, ez = cz1 - cz0;
    const int nc = ex * ey * ez;
    // cell offsets inside the tile: one thread for small tiles, else a
    // block-wide exclusive scan of the cells' local counts, 256 at a time
    if (nc <= 64) {
        if (threadIdx.x == 0) {
            int o = 0;
            for (int q = 0; q < nc; ++q) {
                coff[q] = o;
                o += local_count[((cz0 + q / (ex * ey)) * g.ny + (cy0 + (q / ex) % ey)) * g.nx + (cx0 + q % ex)];
            }
            coff[nc] = o;
        }
    } else {
        int carry = 0;
        for (int q0 = 0; q0 < nc; q0 += kScanThreads) {
            const int q = q0 + (int)threadIdx.x;
            const int v = q < nc ? local_count[((cz0 + q / (ex * ey)) * g.ny + (cy0 + (q / ex) % ey)) * g.nx + (cx0 + q % ex)] : 0;
            int tot;
            const int ex_v = block_exclusive_scan(v, &tot);
            if (q < nc) coff[q] = carry + ex_v;
            carry += tot;
        }
        if (threadIdx.x == 0) coff[nc] = carry;
    }
    __syncthreads();
    // one thread per local (coalesced writes): its cell is the last one whose
    // offset does not exceed its place (empty cells are passed over)
    for (int e = threadIdx.x; e < nloc; e += blockDim.x) {
        int lo = 0, hi = nc - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (coff[mid] <= e) lo = mid;
            else hi = mid - 1;
        }
        const int c = ((cz0 + lo / (ex * ey)) * g.ny + (cy0 + (lo / ex) % ey)) * g.nx + (cx0 + lo % ex);
        const int k = start[c] + (e - coff[lo]);
        work[w0 + e] = parts[k];
        work_k[w0 + e] = k;  // the atom's mirror index
        work_c[w0 + e] = c;
    }
}

// Cells per tile: about 480 locals per tile (one 512-thread block), with
// the window (tile + 1 cell on each side) inside the LDS budget.  The
// density is taken over the domain's volume in cells, not over the grid's
// inner cells: the grid rounds each side up by up to a cell and a half, so
// a small box would look up to twice as sparse as it is.  An explicit choice
// (tmd_set_tile) overrides.
int auto_tile_edge(const Geom& G, int n) {
    double cells = 1.0;
    for (int d = 0; d < 3; ++d) cells *= std::max((G.rank[2 * d + 1] - G.rank[2 * d]) / G.spacing, 1.0);
    const double rho = std::max((double)n / cells, 1e-3);
    int T = (int)std::lround(std::cbrt(480.0 / rho));
    T = std::max(1, std::min(T, kMaxTile));
    while (T > 1 && (T + 2.0) * (T + 2.0) * (T + 2.0) * rho * 28.0 > 0.85 * kLdsBudget) --T;
    return T;
}

TileGeom make_tile_geom(const Geom& G, int tx, int ty, int tz) {
    TileGeom g;
    g.nx = G.nx; g.ny = G.ny; g.nz = G.nz;
    g.tx = std::max(1, std::min(tx, kMaxTile));
    g.ty = std::max(1, std::min(ty, kMaxTile));
    g.tz = std::max(1, std::min(tz, kMaxTile));
    g.ntx = std::max(1, div_up(G.nx - 2, g.tx));
    g.nty = std::max(1, div_up(G.ny - 2, g.ty));
    g.ntz = std::max(1, div_up(G.nz - 2, g.tz));
    return g;
}

static void choose_tiles(Sim& s) {
    int t[3] = {s.tile_req[0], s.tile_req[1], s.tile_req[2]};
    if (t[0] <= 0 || t[1] <= 0 || t[2] <= 0) {
        if (s.tile_auto <= 0) s.tile_auto = auto_tile_edge(s.g, s.n);
        t[0] = t[1] = t[2] = s.tile_auto;
    }
    s.tiles = make_tile_geom(s.g, t[0], t[1], t[2]);
}

// auto_tile_edge sees only the mean density.  Where atoms crowd into part of
// the box (the DEM gas fills half of it), the largest window can outgrow the
// LDS share of the force pass in use (stage_bytes per slot within
// stage_budget per block); the pass then reads such windows from global
// memory.  From the peak window density of the current tiles this picks the
// largest edge whose window fits, for the next binning.  The edge only
// shrinks, and the tiles do not change the results (list order follows the
// cells, not the tiles).
void retune_tile_edge(Sim& s) {
    if (s.tile_req[0] > 0 && s.tile_req[1] > 0 && s.tile_req[2] > 0) return;  // explicit tiles
    const int T = s.tiles.tx;
    if (T <= 1 || (double)(s.max_window + 1) * s.stage_bytes <= (double)s.stage_budget) return;
    const double peak = (s.max_window + 1) / ((T + 2.0) * (T + 2.0) * (T + 2.0));
    int t = T - 1;
    while (t > 1 && (t + 2.0) * (t + 2.0) * (t + 2.0) * peak * s.stage_bytes > (double)s.stage_budget) --t;
    s.tile_auto = std::min(t, s.tile_auto > 0 ? s.tile_auto : t);
}

// Counting sort of atoms 0..nall-1 by cell: cell_count, cell_start (CSR),
// cell_parts (index order inside each cell), cell_of, cell_inv.
static void bin_counting_sort(Sim& s, int nall) {
    const Geom& G = s.g;
    BinArgs a{G.aabb[0], G.aabb[2], G.aabb[4], G.spacing, G.nx, G.ny, G.ncells, nall};
    const int nt = 256;
    TMD_HIP(hipMemsetAsync(s.cell_count.p, 0, sizeof(int) * (size_t)G.ncells, s.stream));
    if (nall > 0)
        hipLaunchKernelGGL(k_bin_count, dim3(div_up(nall, nt)), dim3(nt), 0, s.stream, (const double4*)s.pos.as<double4>(), a,
                           s.cell_of.as<int>(), s.cell_slot.as<int>(), s.cell_count.as<int>(), s.cell_inv.as<int>());
    exclusive_scan_int(s.cell_count.as<int>(), s.cell_start.as<int>(), G.ncells, s.cell_start.as<int>() + G.ncells, s);
    if (nall > 0)
        hipLaunchKernelGGL(k_bin_scatter, dim3(div_up(nall, nt)), dim3(nt), 0, s.stream, (const int*)s.cell_of.as<int>(),
                           (const int*)s.cell_slot.as<int>(), (const int*)s.cell_start.as<int>(), nall, s.cell_parts.as<int>());
    if (nall < 4 * G.ncells)  // sparse grid: a thread per cell beats a wavefront per cell
        hipLaunchKernelGGL(k_bin_sort_small, dim3(div_up(G.ncells, nt)), dim3(nt), 0, s.stream,
                           (const int*)s.cell_start.as<int>(), G.ncells, s.n, s.cell_parts.as<int>(), s.cell_count.as<int>(),
                           s.cell_inv.as<int>());
    else
        hipLaunchKernelGGL(k_bin_sort, dim3(div_up(G.ncells, nt / kWave)), dim3(nt), 0, s.stream,
                           (const int*)s.cell_start.as<int>(), G.ncells, s.n, s.cell_parts.as<int>(), s.cell_count.as<int>(),
                           s.cell_inv.as<int>());
}

// k-th local of the cell order <- local parts[k]: positions, velocities,
// forces and masses into the scratch copies
__global__ void k_permute_locals(int n, const int* __restrict__ parts, const double4* __restrict__ pos,
                                 const double4* __restrict__ vel, const double4* __restrict__ force,
                                 const double* __restrict__ mass, double4* __restrict__ pos2, double4* __restrict__ vel2,
                                 double4* __restrict__ force2, double* __restrict__ mass2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = parts[k];
    pos2[k] = pos[i];
    vel2[k] = vel[i];
    force2[k] = force[i];
    mass2[k] = mass[i];
}

// Cell-order reordering of the locals (tmd_set_sort_particles; the
// reference's sort_particles() hook, core/grid.impala:54, 323-356): at the
// PBC wrap of a rebuild, before the ghosts are built, the locals are
// renumbered in cell order, index order inside a cell (a stable counting
// sort), and their positions, velocities, forces and masses move with them.
// The ghost layer built next refers to the new numbering.  (The reference's
// hook renumbers inside distribute_particles, after borders, and copies only
// positions and velocities, so its ghosts keep the old source indices; it is
// compiled out there.)  A local outside the grid after the wrap is an error,
// as in binning.
void sort_locals(Sim& s) {
    if (has_error() || s.n == 0) return;
    // the fused step loop keeps positions and velocities in the mirrors
    sync_pos_from_mirror(s);
    enforce_pbc(s);
    const Geom& G = s.g;
    s.cell_count.reserve(sizeof(int) * (size_t)G.ncells, false, s.stream);
    s.cell_start.reserve(sizeof(int) * ((size_t)G.ncells + 1), false, s.stream);
    const size_t n = (size_t)s.n;
    s.sort_buf.reserve((sizeof(double4) * 3 + sizeof(double)) * n, false, s.stream);
    if (has_error()) return;
    s.ng = 0;
    bin_counting_sort(s, s.n);
    int binned = 0;
    TMD_HIP(hipMemcpyAsync(&binned, s.cell_start.as<int>() + G.ncells, sizeof(int), hipMemcpyDeviceToHost, s.stream));
    TMD_HIP(hipStreamSynchronize(s.stream));
    if (has_error()) return;
    if (binned != s.n) {
        set_error(__FILE__, __LINE__, "sort: local atoms outside the cell grid after the periodic wrap (non-finite positions?)");
        return;
    }
    double4* p2 = s.sort_buf.as<double4>();
    double4* v2 = p2 + n;
    double4* f2 = v2 + n;
    double* m2 = reinterpret_cast<double*>(f2 + n);
    hipLaunchKernelGGL(k_permute_locals, dim3(div_up(s.n, 256)), dim3(256), 0, s.stream, s.n, (const int*)s.cell_parts.as<int>(),
                       (const double4*)s.pos.as<double4>(), (const double4*)s.vel.as<double4>(),
                       (const double4*)s.force.as<double4>(), (const double*)s.mass.as<double>(), p2, v2, f2, m2);
    TMD_HIP(hipMemcpyAsync(s.pos.p, p2, sizeof(double4) * n, hipMemcpyDeviceToDevice, s.stream));
    TMD_HIP(hipMemcpyAsync(s.vel.p, v2, sizeof(double4) * n, hipMemcpyDeviceToDevice, s.stream));
    TMD_HIP(hipMemcpyAsync(s.force.p, f2, sizeof(double4) * n, hipMemcpyDeviceToDevice, s.stream));
    TMD_HIP(hipMemcpyAsync(s.mass.p, m2, sizeof(double) * n, hipMemcpyDeviceToDevice, s.stream));
    TMD_HIP(hipGetLastError());
    // every atom-indexed table is stale now
    mirror_invalidate(s);
    s.inv_valid = 0;
    s.gsrc_valid = 0;
    s.windows_valid = 0;
    s.nb_valid = 0;
}

void bin_particles(Sim& s) {
    if (has_error()) return;
    const int nall = s.n + s.ng;
    const Geom& G = s.g;
    s.cell_count.reserve(sizeof(int) * (size_t)G.ncells, false, s.stream);
    s.cell_start.reserve(sizeof(int) * ((size_t)G.ncells + 1), false, s.stream);
    s.work.reserve(sizeof(int) * ((size_t)s.n + 1), false, s.stream);
    s.work_k.reserve(sizeof(int) * ((size_t)s.n + 1), false, s.stream);
    s.work_c.reserve(sizeof(int) * ((size_t)s.n + 1), false, s.stream);
    if (has_error()) return;
    mirror_invalidate(s);
    s.inv_valid = 0;
    const int nt = 256;
    bin_counting_sort(s, nall);
    s.inv_valid = 1;
    build_ghost_src(s);
    // tile work order
    choose_tiles(s);
    const TileGeom tg = s.tiles;
    const int ntiles = tg.ntx * tg.nty * tg.ntz;
    s.tile_start.reserve(sizeof(int) * ((size_t)ntiles + 1), false, s.stream);
    s.work_start.reserve(sizeof(int) * ((size_t)ntiles + 1), false, s.stream);  // per-tile counts
    if (has_error()) return;
    int* counters = s.counters.as<int>();
    TMD_HIP(hipMemsetAsync(counters + 6, 0, 2 * sizeof(int), s.stream));
    if (tg.tx * tg.ty * tg.tz <= 64)
        hipLaunchKernelGGL(k_tile_stats_small, dim3(div_up(ntiles, nt)), dim3(nt), 0, s.stream, (const int*)s.cell_start.as<int>(),
                           (const int*)s.cell_count.as<int>(), tg, ntiles, s.work_start.as<int>(), counters + 6, counters + 7);
    else
        hipLaunchKernelGGL(k_tile_stats, dim3(ntiles), dim3(kWave), 0, s.stream, (const int*)s.cell_start.as<int>(),
                           (const int*)s.cell_count.as<int>(), tg, ntiles, s.work_start.as<int>(), counters + 6, counters + 7);
    exclusive_scan_int(s.work_start.as<int>(), s.tile_start.as<int>(), ntiles, s.tile_start.as<int>() + ntiles, s);
    hipLaunchKernelGGL(k_tile_fill_block, dim3(ntiles), dim3(256), 0, s.stream, (const int*)s.cell_start.as<int>(),
                       (const int*)s.cell_count.as<int>(), (const int*)s.tile_start.as<int>(), tg,
                       (const int*)s.cell_parts.as<int>(), s.work.as<int>(), s.work_k.as<int>(), s.work_c.as<int>());
    TMD_HIP(hipGetLastError());
    s.windows_valid = 0;
    build_windows(s);
}

}  // namespace tmd
