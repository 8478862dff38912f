// Batched replicas: many independent periodic LJ boxes of different sizes
// advanced together, every kernel launch covering all of them.
//
// Layout.  The replicas' atoms share one slot space: replica r owns slots
// [a0, a0 + cap), locals first ([a0, a0 + n)), then room for its ghosts.
// Cells, tiles and work slots are concatenated the same way, so one global
// exclusive scan gives every replica's cell CSR and tile work ranges at
// once.  A block finds its replica through a static block -> replica table
// and reads the replica's descriptor (geometry, offsets) with scalar loads.
//
// No host round trip inside a run.  Ghost counts live in device counters;
// every launch is sized by capacity (ghost capacity is set from the volume
// of the ghost shell with headroom), and threads past the live count exit.
// Capacity overflow (ghosts, list entries, 16-bit windows) raises a device
// flag; the host reads the flags once after the run and, if any is set,
// grows the capacities, restores the run's starting state and runs again.
// With no host round trip, the whole run -- reneighbouring included -- is
// captured into one hipGraph and launched once.
//
// Every per-atom and per-pair computation is the single-box code
// (tmd_pair.hpp, and the same integration, ghost and binning arithmetic),
// and every ordering (ghosts, cell lists, neighbour lists) is the
// single-box ordering within each replica, so a replica's trajectory is
// bit-identical to the same box run alone through the md_* path, for any
// tile size and block size.  The cell spacing changes list order, so it
// keeps results identical only to single-box runs with the same spacing.
#include <chrono>
#include <cstring>
#include <cmath>
#include <map>
#include <memory>

#include "tmd_internal.hpp"
#include "tmd_pair.hpp"
#include "../../include/md_abi.h"
#include "../../include/tmd_ext.h"

namespace tmd {

constexpr int kBlk = 256;

struct RepDesc {
    int n, a0, cap, w0;          // locals, first slot, slots (locals + ghost room), first work slot
    int c0, ncells, t0, ntiles;  // first cell / tile, counts
    int sb0, nsb;                // scan blocks (kScanTile slots) over [0, cap)
    int eb0, neb;                // 256-blocks over the locals
    int cb0, ncb;                // 256-blocks over [0, cap)
    int gb0, ngb;                // 256-blocks over the ghost room [n, cap)
    TileGeom g;
    double aabb[6], len[3], spacing;
    double lo_pbc[3], hi_pbc[3];     // world bounds (periodic wrap)
    double hi_thr[3], lo_thr[3];     // ghost selection: aabb max - 2s, aabb min + 2s
    double whi_thr[3], wlo_thr[3];   // image flags: world max - s, world min + s
};

// per-replica device counters
enum { DY_NG = 0, DY_BASE = 1, DY_TOT0 = 4, DY_STRIDE = 8 };
// device flags (max-reduced)
enum { FL_GHOST = 0, FL_LIST = 1, FL_WIN = 2, FL_UNBINNED = 3, FL_MAXWIN = 4, FL_MAXLOC = 5, FL_N = 8 };

__device__ __forceinline__ int b_img_encode(int ix, int iy, int iz) { return (ix + 1) + 3 * (iy + 1) + 9 * (iz + 1); }
__device__ __forceinline__ void b_img_decode(int c, int& ix, int& iy, int& iz) {
    ix = c % 3 - 1;
    iy = (c / 3) % 3 - 1;
    iz = c / 9 - 1;
}
__device__ __forceinline__ double b_comp(const double4& p, int d) { return d == 0 ? p.x : (d == 1 ? p.y : p.z); }


__global__ void k_b_fill(int* __restrict__ p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// integration and periodic wrap (k_integrate.hip arithmetic)
__global__ void k_b_int1(const RepDesc* __restrict__ R, const int* __restrict__ eb_rep, double4* __restrict__ pos,
                         double4* __restrict__ vel, const double4* __restrict__ force, double hdt, double dt,
                         const int* __restrict__ inv, const Mir spos) {
    const int r = eb_rep[blockIdx.x];
    const int i = (blockIdx.x - R[r].eb0) * kBlk + threadIdx.x;
    if (i >= R[r].n) return;
    const int s = R[r].a0 + i;
    double4 v = vel[s];
    const double4 f = force[s];
    double4 x = pos[s];
    v.x = v.x + f.x * hdt;
    v.y = v.y + f.y * hdt;
    v.z = v.z + f.z * hdt;
    x.x = x.x + v.x * dt;
    x.y = x.y + v.y * dt;
    x.z = x.z + v.z * dt;
    vel[s] = v;
    pos[s] = x;
    const int k = inv[s];  // keep the cell-ordered mirror current
    if (k >= 0) spos[k] = x;
}

__global__ void k_b_pbc(const RepDesc* __restrict__ R, const int* __restrict__ eb_rep, double4* __restrict__ pos) {
    const int r = eb_rep[blockIdx.x];
    const RepDesc& D = R[r];
    const int i = (blockIdx.x - D.eb0) * kBlk + threadIdx.x;
    if (i >= D.n) return;
    double4 p = pos[D.a0 + i];
    double q;
    q = p.x; if (p.x < D.lo_pbc[0]) q += D.len[0]; if (p.x > D.hi_pbc[0]) q -= D.len[0]; p.x = q;
    q = p.y; if (p.y < D.lo_pbc[1]) q += D.len[1]; if (p.y > D.hi_pbc[1]) q -= D.len[1]; p.y = q;
    q = p.z; if (p.z < D.lo_pbc[2]) q += D.len[2]; if (p.z > D.hi_pbc[2]) q -= D.len[2]; p.z = q;
    pos[D.a0 + i] = p;
}

// ---------------------------------------------------------------------------
// ghosts (k_ghost.hip order and arithmetic), one dimension per pass
__global__ void k_b_zero_ng(int* __restrict__ dyn, int nrep) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrep) dyn[r * DY_STRIDE + DY_NG] = 0;
}

__global__ void __launch_bounds__(kScanThreads) k_b_border_count(const RepDesc* __restrict__ R, const int* __restrict__ sb_rep,
                                                                 const double4* __restrict__ pos, const int* __restrict__ dyn,
                                                                 int d, int* __restrict__ cnt0, int* __restrict__ cnt1) {
    const int r = sb_rep[blockIdx.x];
    const RepDesc& D = R[r];
    const int scan_end = D.n + dyn[r * DY_STRIDE + DY_NG];
    const int base = (blockIdx.x - D.sb0) * kScanTile + threadIdx.x * kScanItems;
    int c0 = 0, c1 = 0;
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
        const int i = base + it;
        if (i < scan_end) {
            const double p = b_comp(pos[D.a0 + i], d);
            c0 += (p > D.hi_thr[d]) ? 1 : 0;
            c1 += (p < D.lo_thr[d]) ? 1 : 0;
        }
    }
    const int t0 = block_sum_int(c0);
    const int t1 = block_sum_int(c1);
    if (threadIdx.x == 0) {
        cnt0[blockIdx.x] = t0;
        cnt1[blockIdx.x] = t1;
    }
}

// one block per replica: exclusive scans of the replica's block counts,
// ghost bases and the new ghost count (clamped to the capacity, flagged)
__global__ void __launch_bounds__(kScanThreads) k_b_border_scan(const RepDesc* __restrict__ R, int d, int* __restrict__ cnt0,
                                                                int* __restrict__ cnt1, int* __restrict__ dyn,
                                                                int* __restrict__ flags) {
    const int r = blockIdx.x;
    const RepDesc& D = R[r];
    int tot[2];
    for (int side = 0; side < 2; ++side) {
        int* c = (side == 0 ? cnt0 : cnt1) + D.sb0;
        int carry = 0;
        for (int base = 0; base < D.nsb; base += kScanThreads) {
            const int idx = base + threadIdx.x;
            const int v = idx < D.nsb ? c[idx] : 0;
            int t;
            const int ex = block_exclusive_scan(v, &t);
            if (idx < D.nsb) c[idx] = carry + ex;
            carry += t;
        }
        tot[side] = carry;
    }
    if (threadIdx.x == 0) {
        int* dy = dyn + r * DY_STRIDE;
        const int ng = dy[DY_NG];
        dy[DY_BASE + d] = D.n + ng;
        dy[DY_TOT0 + d] = tot[0];
        int ngn = ng + tot[0] + tot[1];
        if (D.n + ngn > D.cap) {
            atomicMax(&flags[FL_GHOST], D.n + ngn);
            ngn = D.cap - D.n;
        }
        dy[DY_NG] = ngn;
    }
}

__device__ __forceinline__ void b_write_ghost(const RepDesc& D, int i, int slot, int side, int d, double4* __restrict__ pos,
                                              int* __restrict__ g_owner, int* __restrict__ g_code) {
    if (slot >= D.cap) return;
    const double4 p = pos[D.a0 + i];
    const double pd = b_comp(p, d);
    const int fl = (side == 0) ? (pd > D.whi_thr[d] ? -1 : 0) : (pd < D.wlo_thr[d] ? 1 : 0);
    int f[3] = {0, 0, 0};
    f[d] = fl;
    double4 q;
    q.x = p.x + D.len[0] * (double)f[0];
    q.y = p.y + D.len[1] * (double)f[1];
    q.z = p.z + D.len[2] * (double)f[2];
    q.w = p.w;
    int owner, ix, iy, iz;
    if (i < D.n) {
        owner = i;
        ix = f[0]; iy = f[1]; iz = f[2];
    } else {
        owner = g_owner[D.a0 + i];
        b_img_decode(g_code[D.a0 + i], ix, iy, iz);
        ix += f[0]; iy += f[1]; iz += f[2];
    }
    pos[D.a0 + slot] = q;
    g_owner[D.a0 + slot] = owner;
    g_code[D.a0 + slot] = b_img_encode(ix, iy, iz);
}

__global__ void __launch_bounds__(kScanThreads) k_b_border_scatter(const RepDesc* __restrict__ R, const int* __restrict__ sb_rep,
                                                                   double4* __restrict__ pos, int* __restrict__ g_owner,
                                                                   int* __restrict__ g_code, const int* __restrict__ off0,
                                                                   const int* __restrict__ off1, const int* __restrict__ dyn,
                                                                   int d) {
    const int r = sb_rep[blockIdx.x];
    const RepDesc& D = R[r];
    const int scan_end = dyn[r * DY_STRIDE + DY_BASE + d];  // n + ghosts before this dimension
    const int total0 = dyn[r * DY_STRIDE + DY_TOT0 + d];
    const int base = (blockIdx.x - D.sb0) * kScanTile + threadIdx.x * kScanItems;
    int f0[kScanItems], f1[kScanItems];
    int c0 = 0, c1 = 0;
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
        const int i = base + it;
        f0[it] = f1[it] = 0;
        if (i < scan_end) {
            const double p = b_comp(pos[D.a0 + i], d);
            f0[it] = (p > D.hi_thr[d]) ? 1 : 0;
            f1[it] = (p < D.lo_thr[d]) ? 1 : 0;
        }
        c0 += f0[it];
        c1 += f1[it];
    }
    int t0, t1;
    int r0 = off0[blockIdx.x] + block_exclusive_scan(c0, &t0);
    int r1 = off1[blockIdx.x] + block_exclusive_scan(c1, &t1);
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
        const int i = base + it;
        if (f0[it]) { b_write_ghost(D, i, scan_end + r0, 0, d, pos, g_owner, g_code); r0++; }
        if (f1[it]) { b_write_ghost(D, i, scan_end + total0 + r1, 1, d, pos, g_owner, g_code); r1++; }
    }
}

// after a force pass that wrote the advanced locals into the mirror only:
// ghost = owner + image * L with the owner read from its mirror entry, into
// pos and the ghost's mirror entry (k_ghost_step's arithmetic)
__global__ void k_b_ghost_step(const RepDesc* __restrict__ R, const int* __restrict__ gb_rep, const int* __restrict__ dyn,
                               double4* __restrict__ pos, const int* __restrict__ g_owner, const int* __restrict__ g_code,
                               const int* __restrict__ inv, const Mir spos) {
    const int r = gb_rep[blockIdx.x];
    const RepDesc& D = R[r];
    const int g = (blockIdx.x - D.gb0) * kBlk + threadIdx.x;
    if (g >= dyn[r * DY_STRIDE + DY_NG]) return;
    const int s = D.a0 + D.n + g;
    int ix, iy, iz;
    b_img_decode(g_code[s], ix, iy, iz);
    const double4 p = spos[inv[D.a0 + g_owner[s]]];
    double4 q;
    q.x = p.x + D.len[0] * (double)ix;
    q.y = p.y + D.len[1] * (double)iy;
    q.z = p.z + D.len[2] * (double)iz;
    q.w = p.w;
    pos[s] = q;
    const int k = inv[s];
    if (k >= 0) spos[k] = q;
}

// mirror indices of each ghost's owner and of the ghost itself (after the
// binning), read in slot order by the step loop's ghost pass
__global__ void k_b_ghost_src(const RepDesc* __restrict__ R, const int* __restrict__ gb_rep, const int* __restrict__ dyn,
                              const int* __restrict__ g_owner, const int* __restrict__ inv, int* __restrict__ g_ko,
                              int* __restrict__ g_kg) {
    const int r = gb_rep[blockIdx.x];
    const RepDesc& D = R[r];
    const int g = (blockIdx.x - D.gb0) * kBlk + threadIdx.x;
    if (g >= dyn[r * DY_STRIDE + DY_NG]) return;
    const int s = D.a0 + D.n + g;
    g_ko[s] = inv[D.a0 + g_owner[s]];
    g_kg[s] = inv[s];
}

// the step loop's ghost pass, mirror only (k_b_ghost_step's arithmetic); the
// ghosts' pos entries are written by the full pass at the end of a run
__global__ void k_b_ghost_step_mirror(const RepDesc* __restrict__ R, const int* __restrict__ gb_rep,
                                      const int* __restrict__ dyn, const int* __restrict__ g_code,
                                      const int* __restrict__ g_ko, const int* __restrict__ g_kg,
                                      const Mir spos) {
    const int r = gb_rep[blockIdx.x];
    const RepDesc& D = R[r];
    const int g = (blockIdx.x - D.gb0) * kBlk + threadIdx.x;
    if (g >= dyn[r * DY_STRIDE + DY_NG]) return;
    const int s = D.a0 + D.n + g;
    const int k = g_kg[s];
    if (k < 0) return;
    int ix, iy, iz;
    b_img_decode(g_code[s], ix, iy, iz);
    const double4 p = spos[g_ko[s]];
    double4 q;
    q.x = p.x + D.len[0] * (double)ix;
    q.y = p.y + D.len[1] * (double)iy;
    q.z = p.z + D.len[2] * (double)iz;
    q.w = p.w;
    spos[k] = q;
}

// pos of every local from its mirror entry (before a rebuild and at the end
// of a run)
__global__ void k_b_unmirror(const RepDesc* __restrict__ R, const int* __restrict__ eb_rep, const int* __restrict__ inv,
                             const CMir spos, double4* __restrict__ pos) {
    const int r = eb_rep[blockIdx.x];
    const int i = (blockIdx.x - R[r].eb0) * kBlk + threadIdx.x;
    if (i >= R[r].n) return;
    const int s = R[r].a0 + i;
    pos[s] = spos[inv[s]];
}

// ---------------------------------------------------------------------------
// binning (k_bin.hip arithmetic): cell indices are replica-local, cell
// counters and the CSR are global
__global__ void k_b_bin_count(const RepDesc* __restrict__ R, const int* __restrict__ cb_rep, const int* __restrict__ dyn,
                              const double4* __restrict__ pos, int* __restrict__ cell_of, int* __restrict__ slot,
                              int* __restrict__ count, int* __restrict__ inv) {
    const int r = cb_rep[blockIdx.x];
    const RepDesc& D = R[r];
    const int i = (blockIdx.x - D.cb0) * kBlk + threadIdx.x;
    const bool live = i < D.n + dyn[r * DY_STRIDE + DY_NG];
    int c = -1;
    if (live) {
        const double4 p = pos[D.a0 + i];
        const double fx = floor((p.x - D.aabb[0]) / D.spacing);
        const double fy = floor((p.y - D.aabb[2]) / D.spacing);
        const double fz = floor((p.z - D.aabb[4]) / D.spacing);
        if (fabs(fx) < 1e9 && fabs(fy) < 1e9 && fabs(fz) < 1e9) {
            const int ci = (int)fx, cj = (int)fy, ck = (int)fz;
            const long long flat = ((long long)ck * D.g.ny + cj) * D.g.nx + ci;
            if (flat >= 0 && flat < D.ncells) c = (int)flat;
        }
    }
    // all lanes take part; a block belongs to one replica, so cells do not mix
    const int sl = wave_cell_slot(c >= 0 ? D.c0 + c : -1, count);
    if (!live) return;
    cell_of[D.a0 + i] = c;
    slot[D.a0 + i] = sl;
    inv[D.a0 + i] = -1;  // set by k_b_bin_sort for binned atoms
}

__global__ void k_b_bin_scatter(const RepDesc* __restrict__ R, const int* __restrict__ cb_rep, const int* __restrict__ dyn,
                                const int* __restrict__ cell_of, const int* __restrict__ slot, const int* __restrict__ start,
                                int* __restrict__ parts) {
    const int r = cb_rep[blockIdx.x];
    const RepDesc& D = R[r];
    const int i = (blockIdx.x - D.cb0) * kBlk + threadIdx.x;
    if (i >= D.n + dyn[r * DY_STRIDE + DY_NG]) return;
    const int c = cell_of[D.a0 + i];
    if (c >= 0) parts[start[D.c0 + c] + slot[D.a0 + i]] = D.a0 + i;
}

__global__ void k_b_bin_sort(const RepDesc* __restrict__ R, const int* __restrict__ cell_rep, const int* __restrict__ start,
                             int ncells_all, int* __restrict__ parts, int* __restrict__ local_count, int* __restrict__ inv) {
    const int c = blockIdx.x * (blockDim.x / kWave) + (int)threadIdx.x / kWave;  // one wavefront per cell
    if (c >= ncells_all) return;
    const int lim = R[cell_rep[c]].a0 + R[cell_rep[c]].n;
    const int nl = wave_sort_cell(start[c], start[c + 1], lim, parts, inv);
    if ((threadIdx.x & (kWave - 1)) == 0) local_count[c] = nl;
}

__global__ void k_b_tile_stats(const RepDesc* __restrict__ R, const int* __restrict__ tile_rep, const int* __restrict__ start,
                               const int* __restrict__ local_count, int ntiles_all, int* __restrict__ tile_count,
                               int* __restrict__ flags) {
    const int tile = blockIdx.x * blockDim.x + threadIdx.x;
    int cnt = 0, S = 0;
    if (tile < ntiles_all) {
        const RepDesc& D = R[tile_rep[tile]];
        const TileGeom g = D.g;
        const int* st = start + D.c0;
        const int* lc = local_count + D.c0;
        const int lt = tile - D.t0;
        const int tx_i = lt % g.ntx, ty_i = (lt / g.ntx) % g.nty, tz_i = lt / (g.ntx * g.nty);
        const int cx0 = 1 + tx_i * g.tx, cy0 = 1 + ty_i * g.ty, cz0 = 1 + tz_i * g.tz;
        const int cx1 = min(cx0 + g.tx, g.nx - 1), cy1 = min(cy0 + g.ty, g.ny - 1), cz1 = min(cz0 + g.tz, g.nz - 1);
        for (int z = cz0; z < cz1; ++z)
            for (int y = cy0; y < cy1; ++y)
                for (int x = cx0; x < cx1; ++x) cnt += lc[(z * g.ny + y) * g.nx + x];
        tile_count[tile] = cnt;
        const int xlo = cx0 - 1, xhi = min(cx0 + g.tx, g.nx - 1);
        for (int z = cz0 - 1; z <= cz0 + g.tz; ++z)
            for (int y = cy0 - 1; y <= cy0 + g.ty; ++y)
                if (y < g.ny && z < g.nz) {
                    const int row = (z * g.ny + y) * g.nx;
                    S += st[row + xhi + 1] - st[row + xlo];
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
        atomicMax(&flags[FL_MAXWIN], S);
        atomicMax(&flags[FL_MAXLOC], cnt);
        if (S > kWinMax) atomicMax(&flags[FL_WIN], S);
    }
}

// one block per tile: the tile's locals in (z, y, x) cell order, and the
// CSR (mirror) index of each (k_tile_fill_block's order)
__global__ void __launch_bounds__(64) k_b_tile_fill(const RepDesc* __restrict__ R, const int* __restrict__ tile_rep,
                                                    const int* __restrict__ start, const int* __restrict__ local_count,
                                                    const int* __restrict__ tile_start, const int* __restrict__ parts,
                                                    int* __restrict__ work, int* __restrict__ work_k) {
    __shared__ int coff[kMaxTile * kMaxTile * kMaxTile + 1];
    const int tile = blockIdx.x;
    const int w0 = tile_start[tile], nloc = tile_start[tile + 1] - w0;
    if (nloc == 0) return;
    const RepDesc& D = R[tile_rep[tile]];
    const TileGeom g = D.g;
    const int* st = start + D.c0;
    const int* lc = local_count + D.c0;
    const int lt = tile - D.t0;
    const int tx_i = lt % g.ntx, ty_i = (lt / g.ntx) % g.nty, tz_i = lt / (g.ntx * g.nty);
    const int cx0 = 1 + tx_i * g.tx, cy0 = 1 + ty_i * g.ty, cz0 = 1 + tz_i * g.tz;
    const int cx1 = min(cx0 + g.tx, g.nx - 1), cy1 = min(cy0 + g.ty, g.ny - 1), cz1 = min(cz0 + g.tz, g.nz - 1);
    const int ex = cx1 - cx0, ey = cy1 - cy0, ez = cz1 - cz0;
    const int nc = ex * ey * ez;
    if (threadIdx.x == 0) {
        int o = 0;
        for (int q = 0; q < nc; ++q) {
            const int c = ((cz0 + q / (ex * ey)) * g.ny + (cy0 + (q / ex) % ey)) * g.nx + (cx0 + q % ex);
            coff[q] = o;
            o += lc[c];
        }
        coff[nc] = o;
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
        const int k = st[c] + (e - coff[lo]);
        work[w0 + e] = parts[k];
        work_k[w0 + e] = k;
    }
}

// every replica's locals must all sit in its tiles
__global__ void k_b_check_work(const RepDesc* __restrict__ R, int nrep, const int* __restrict__ tile_start,
                               int* __restrict__ flags) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrep) return;
    const RepDesc& D = R[r];
    const int got = tile_start[D.t0 + D.ntiles] - tile_start[D.t0];
    if (got != D.n || tile_start[D.t0] != D.w0) atomicMax(&flags[FL_UNBINNED], 1 + r);
}

__global__ void k_b_mirror(const double4* __restrict__ pos, const int* __restrict__ parts, const int* __restrict__ nbinned,
                           const Mir spos, int cap_all) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= min(*nbinned, cap_all)) return;
    spos[k] = pos[parts[k]];
}

// ---------------------------------------------------------------------------
// neighbour lists and forces: one block per tile of any replica
__global__ void __launch_bounds__(kMaxTileThreads) k_b_nlist(const RepDesc* __restrict__ R, const int* __restrict__ tile_rep,
                                                          const CMir spos, const int* __restrict__ work_k,
                                                          const int* __restrict__ cell_of,
                                                          const int* __restrict__ start,
                                                          const int* __restrict__ parts, const int* __restrict__ work,
                                                          const int* __restrict__ tile_start, NbArgs a,
                                                          int* __restrict__ nb_count, nb_t* __restrict__ nb_list,
                                                          int* __restrict__ flags) {
    extern __shared__ double dyn[];
    int* cst = reinterpret_cast<int*>(reinterpret_cast<char*>(dyn) + (size_t)kStageBytes * a.wcap);
    const int tile = blockIdx.x;
    const int w0 = tile_start[tile], nloc = tile_start[tile + 1] - w0;
    if (nloc == 0) return;
    const RepDesc& D = R[tile_rep[tile]];
    NbArgs aa = a;
    aa.n = D.a0 + D.n;  // slots from here on are ghosts
    aa.g = D.g;
    const int* st = start + D.c0;
    __shared__ TileWin W;
    tile_window(tile - D.t0, st, D.g, W);
    stage_row_starts(W, st, D.g, cst);
    const WinStage sp = win_stage(dyn, a.wcap);
    const bool staged = stage_window(W, a.wcap, spos, nullptr, sp, kListPad);
    for (int at = threadIdx.x; at < nloc; at += blockDim.x) {
        const int w = w0 + at;
        const int i = work[w];
        const int k = work_k[w];  // its mirror index (the mirror is current)
        int unused = 0;
        const int cnt = staged ? nlist_atom<false, true>(i, w, cell_of[i], k, spos[k], cst, aa, W, sp, spos, parts, nb_list,
                                                         nullptr, unused)
                               : nlist_atom<false, false>(i, w, cell_of[i], k, spos[k], cst, aa, W, sp, spos, parts, nb_list,
                                                          nullptr, unused);
        nb_count[w] = cnt;
        if (cnt > a.cap) atomicMax(&flags[FL_LIST], cnt);
    }
}

// Positions come from the mirror (own atom at wk[w]).  ADV 1: also the next
// step's first kick and drift, written only into the second mirror, with no
// force store (as k_lj_full<..., 2>); ADV 0 stores the force.
template <int PREC, int SHIFT, int UNIT, int ADV>
__global__ void __launch_bounds__(kMaxTileThreads) k_b_lj_full(const RepDesc* __restrict__ R, const int* __restrict__ tile_rep,
                                                               const CMir spos, const int* __restrict__ start,
                                                               const int* __restrict__ work, const int* __restrict__ wk,
                                                               const int* __restrict__ tile_start, const int* __restrict__ nbc,
                                                               const nb_t* __restrict__ nbl, LJArgs a,
                                                               double4* __restrict__ force, double4* __restrict__ vel, double hdt,
                                                               double dt, const Mir spos_next) {
    extern __shared__ double dyn[];
    const WinStage sp = win_stage(dyn, a.wcap);
    const int tile = xcd_block(blockIdx.x, gridDim.x);
    const int w0 = tile_start[tile], nloc = tile_start[tile + 1] - w0;
    if (nloc == 0) return;
    const RepDesc& D = R[tile_rep[tile]];
    __shared__ TileWin W;
    tile_window(tile - D.t0, start + D.c0, D.g, W);
    const bool staged = stage_window(W, a.wcap, spos, nullptr, sp);
    for (int at = threadIdx.x; at < nloc; at += blockDim.x) {
        const int w = w0 + at;
        const int i = work[w];
        const int kw = wk[w];
        const double4 pi = spos[kw];
        // A captured run goes on after a list build that overflowed (the
        // host sees the flag at the end, restores the start and reruns with
        // a larger capacity), so a count can exceed the capacity here: read
        // at most `cap` entries, which keeps every list load inside the
        // atom's column; that run's results are discarded.
        const int cnt = min(nbc[w], a.cap);
        const int4* lp = reinterpret_cast<const int4*>(nbl + (size_t)(w >> 6) * (size_t)a.cap * 64) + (w & 63);
        const double4 f = staged ? lj_full_atom<PREC, SHIFT, UNIT, true>(pi, cnt, lp, sp, spos, W, a)
                                 : lj_full_atom<PREC, SHIFT, UNIT, false>(pi, cnt, lp, sp, spos, W, a);
        if (!ADV) force[i] = f;
        // velocity-Verlet final kick (k_final_integration's arithmetic)
        double4 v = vel[i];
        v.x = v.x + f.x * hdt;
        v.y = v.y + f.y * hdt;
        v.z = v.z + f.z * hdt;
        if (ADV) {
            // next step's k_b_int1
            v.x = v.x + f.x * hdt;
            v.y = v.y + f.y * hdt;
            v.z = v.z + f.z * hdt;
            double4 x = pi;
            x.x = x.x + v.x * dt;
            x.y = x.y + v.y * dt;
            x.z = x.z + v.z * dt;
            spos_next[kw] = x;
        }
        vel[i] = v;
    }
}

template <int SHIFT>
__global__ void __launch_bounds__(kMaxTileThreads) k_b_energy(const RepDesc* __restrict__ R, const int* __restrict__ tile_rep,
                                                           const double4* __restrict__ pos, const CMir spos,
                                                           const int* __restrict__ start, const int* __restrict__ parts,
                                                           const int* __restrict__ work, const int* __restrict__ tile_start,
                                                           const int* __restrict__ nbc, const nb_t* __restrict__ nbl, LJArgs a,
                                                           double* __restrict__ partials) {
    extern __shared__ double dyn[];
    const int tile = blockIdx.x;
    const int w0 = tile_start[tile], nloc = tile_start[tile + 1] - w0;
    if (nloc == 0) {
        if (threadIdx.x == 0) {
            partials[3 * tile + 0] = 0.0;
            partials[3 * tile + 1] = 0.0;
            partials[3 * tile + 2] = 0.0;
        }
        return;
    }
    const RepDesc& D = R[tile_rep[tile]];
    __shared__ TileWin W;
    tile_window(tile - D.t0, start + D.c0, D.g, W);
    const WinStage sp = win_stage(dyn, a.wcap);
    const bool staged = stage_window(W, a.wcap, spos, parts, sp);
    double pe = 0.0, vir = 0.0, np = 0.0;
    for (int at = threadIdx.x; at < nloc; at += blockDim.x) {
        const int w = w0 + at;
        const int4* lp = reinterpret_cast<const int4*>(nbl + (size_t)(w >> 6) * (size_t)a.cap * 64) + (w & 63);
        const int cnt = min(nbc[w], a.cap);  // as in k_b_lj_full
        if (staged) lj_energy_atom<SHIFT, true>(pos[work[w]], cnt, lp, sp, spos, parts, W, a, 0, D.a0 + D.n, pe, vir, np);
        else lj_energy_atom<SHIFT, false>(pos[work[w]], cnt, lp, sp, spos, parts, W, a, 0, D.a0 + D.n, pe, vir, np);
    }
    const double bpe = block_sum_double_any(pe);
    const double bvir = block_sum_double_any(vir);
    const double bnp = block_sum_double_any(np);
    if (threadIdx.x == 0) {
        partials[3 * tile + 0] = bpe;
        partials[3 * tile + 1] = bvir;
        partials[3 * tile + 2] = bnp;
    }
}

__global__ void __launch_bounds__(kBlk) k_b_kinetic(const RepDesc* __restrict__ R, const int* __restrict__ eb_rep,
                                                    const double4* __restrict__ vel, double* __restrict__ partials) {
    const int r = eb_rep[blockIdx.x];
    const int i = (blockIdx.x - R[r].eb0) * kBlk + threadIdx.x;
    double t = 0.0;
    if (i < R[r].n) {
        const double4 v = vel[R[r].a0 + i];
        t = (v.x * v.x + v.y * v.y + v.z * v.z) * 1.0;
    }
    const double b = block_sum_double<kBlk>(t);
    if (threadIdx.x == 0) partials[blockIdx.x] = b;
}

// one block per replica: out[4 r + {0,1,2,3}] = PE, virial, pairs, sum v^2
__global__ void __launch_bounds__(kBlk) k_b_thermo_reduce(const RepDesc* __restrict__ R, const double* __restrict__ epart,
                                                          const double* __restrict__ kpart, double* __restrict__ out) {
    const RepDesc& D = R[blockIdx.x];
    for (int c = 0; c < 3; ++c) {
        double v = 0.0;
        for (int t = threadIdx.x; t < D.ntiles; t += kBlk) v += epart[3 * (size_t)(D.t0 + t) + c];
        const double s = block_sum_double<kBlk>(v);
        if (threadIdx.x == 0) out[4 * blockIdx.x + c] = s;
    }
    double v = 0.0;
    for (int b = threadIdx.x; b < D.neb; b += kBlk) v += kpart[D.eb0 + b];
    const double s = block_sum_double<kBlk>(v);
    if (threadIdx.x == 0) out[4 * blockIdx.x + 3] = s;
}

// ===========================================================================
// host side
struct BatchCfg {
    int tile = 0;               // cells per tile edge (0: per-replica automatic)
    int threads = 0;            // threads per tile block (multiple of 64, <= 1024); 0: automatic
    double spacing = 1.0;       // cell edge / (cutoff + skin)
    double list_headroom = 1.2; // list capacity / largest count at set-up
    double ghost_headroom = 1.5;// ghost room / expected ghost count
};

struct Batch {
    int nrep = 0;
    std::vector<int> n;
    std::vector<double> aabb;        // 6 per replica
    std::vector<double> h_pos, h_vel;// initial state, 3 per atom, replicas concatenated
    double rc = 2.5, skin = 0.3, eps = 1.0, sigma = 1.0, dt = 0.005;
    int reneigh = 20, precision = PREC_FP32_MIXED, shift = 0;
    BatchCfg cfg;
    double ghost_scale = 1.0;        // grows after a ghost overflow

    std::vector<RepDesc> R;
    int n_sb = 0, n_eb = 0, n_cb = 0, n_gb = 0, n_tiles = 0, n_cells = 0, n_slots = 0, n_work = 0;
    int list_cap = 8, wcap = 0;
    int threads = kTileThreads;      // tile block size in use (cfg.threads, or automatic at set-up)
    DevBuf dR, sb_rep, eb_rep, cb_rep, gb_rep, tile_rep, cell_rep;
    DevBuf pos, vel, force, g_owner, g_code, cell_of, cell_slot, cell_inv, parts, spos, spos2;
    DevBuf cell_count, cell_start, tile_count, tile_start, work, work_k, g_ko, g_kg;
    DevBuf sb_cnt0, sb_cnt1, scan_tmp, nb_count, nb_list;
    DevBuf dyn, flags, epart, kpart, tout;
    DevBuf snap_pos, snap_vel, snap_force;
    hipStream_t stream = nullptr;
    struct CapturedRun {
        hipGraph_t graph;
        hipGraphExec_t exec;
        std::vector<unsigned long long> sig;  // pointers and scalars baked into the graph
    };
    std::map<int, CapturedRun> graphs;
    bool use_graph = true;  // false: a run's kernels are enqueued directly (no capture)
    bool ready = false;

    ~Batch() {
        drop_graphs();
        DevBuf* bufs[] = {&dR, &sb_rep, &eb_rep, &cb_rep, &gb_rep, &tile_rep, &cell_rep, &pos, &vel, &force, &g_owner,
                          &g_code, &cell_of, &cell_slot, &cell_inv, &parts, &spos, &spos2, &cell_count, &cell_start,
                          &tile_count, &tile_start, &work, &work_k, &g_ko, &g_kg, &sb_cnt0, &sb_cnt1, &scan_tmp, &nb_count, &nb_list, &dyn, &flags, &epart,
                          &kpart, &tout, &snap_pos, &snap_vel, &snap_force};
        if (stream) (void)hipStreamSynchronize(stream);
        for (DevBuf* b : bufs) b->release();
        if (stream) (void)hipStreamDestroy(stream);
    }
    void drop_graphs() {
        for (auto& kv : graphs) {
            (void)hipGraphExecDestroy(kv.second.exec);
            (void)hipGraphDestroy(kv.second.graph);
        }
        graphs.clear();
    }
};

static std::map<int, std::unique_ptr<Batch>> g_batches;
static int g_next_batch = 1;

static Batch* find_batch(int h) {
    auto it = g_batches.find(h);
    if (it == g_batches.end()) {
        set_error(__FILE__, __LINE__, "unknown batch handle");
        return nullptr;
    }
    return it->second.get();
}

static int scan_scratch(int n) { return div_up(std::max(n, 1), kScanTile) + 1; }

// Descriptors, tables and buffers for the current configuration.
static void layout(Batch& B) {
    B.R.assign(B.nrep, RepDesc{});
    const double cut = B.rc + B.skin;
    const double spacing = cut * B.cfg.spacing;
    std::vector<int> sb, eb, cb, gb, tr, cr;
    int a0 = 0, w0 = 0, c0 = 0, t0 = 0;
    for (int r = 0; r < B.nrep; ++r) {
        RepDesc& D = B.R[r];
        const double* box = &B.aabb[6 * (size_t)r];
        Geom G;
        make_geom(G, box, box, spacing);
        const int T = B.cfg.tile > 0 ? B.cfg.tile : auto_tile_edge(G, B.n[r]);
        D.g = make_tile_geom(G, T, T, T);
        D.n = B.n[r];
        double vol = 1.0, volx = 1.0;
        for (int d = 0; d < 3; ++d) {
            vol *= G.len[d];
            volx *= G.len[d] + 2.0 * spacing;
        }
        const double rho = D.n / vol;
        const int gcap = (int)std::ceil(B.cfg.ghost_headroom * B.ghost_scale * rho * (volx - vol)) + 256;
        D.cap = D.n + gcap;
        D.a0 = a0;
        D.w0 = w0;
        D.c0 = c0;
        D.ncells = G.ncells;
        D.t0 = t0;
        D.ntiles = D.g.ntx * D.g.nty * D.g.ntz;
        D.sb0 = (int)sb.size(); D.nsb = div_up(D.cap, kScanTile);
        D.eb0 = (int)eb.size(); D.neb = std::max(1, div_up(D.n, kBlk));
        D.cb0 = (int)cb.size(); D.ncb = div_up(D.cap, kBlk);
        D.gb0 = (int)gb.size(); D.ngb = div_up(D.cap - D.n, kBlk);
        for (int k = 0; k < D.nsb; ++k) sb.push_back(r);
        for (int k = 0; k < D.neb; ++k) eb.push_back(r);
        for (int k = 0; k < D.ncb; ++k) cb.push_back(r);
        for (int k = 0; k < D.ngb; ++k) gb.push_back(r);
        for (int k = 0; k < D.ntiles; ++k) tr.push_back(r);
        for (int k = 0; k < D.ncells; ++k) cr.push_back(r);
        for (int d = 0; d < 6; ++d) D.aabb[d] = G.aabb[d];
        D.spacing = G.spacing;
        const double dsp = G.spacing * 2.0;
        for (int d = 0; d < 3; ++d) {
            D.len[d] = G.len[d];
            D.lo_pbc[d] = G.world[2 * d];
            D.hi_pbc[d] = G.world[2 * d + 1];
            D.hi_thr[d] = G.aabb[2 * d + 1] - dsp;
            D.lo_thr[d] = G.aabb[2 * d] + dsp;
            D.whi_thr[d] = G.world[2 * d + 1] - G.spacing;
            D.wlo_thr[d] = G.world[2 * d] + G.spacing;
        }
        a0 += D.cap;
        w0 += D.n;
        c0 += D.ncells;
        t0 += D.ntiles;
    }
    B.n_slots = a0;
    B.n_work = w0;
    B.n_cells = c0;
    B.n_tiles = t0;
    B.n_sb = (int)sb.size();
    B.n_eb = (int)eb.size();
    B.n_cb = (int)cb.size();
    B.n_gb = (int)gb.size();
    hipStream_t st = B.stream;
    auto upload_ints = [&](DevBuf& b, const std::vector<int>& v) {
        b.reserve(sizeof(int) * std::max<size_t>(v.size(), 1), false, st);
        if (!v.empty()) TMD_HIP(hipMemcpyAsync(b.p, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice, st));
    };
    B.dR.reserve(sizeof(RepDesc) * B.nrep, false, st);
    TMD_HIP(hipMemcpyAsync(B.dR.p, B.R.data(), sizeof(RepDesc) * B.nrep, hipMemcpyHostToDevice, st));
    upload_ints(B.sb_rep, sb);
    upload_ints(B.eb_rep, eb);
    upload_ints(B.cb_rep, cb);
    upload_ints(B.gb_rep, gb);
    upload_ints(B.tile_rep, tr);
    upload_ints(B.cell_rep, cr);
    const size_t S = (size_t)B.n_slots;
    B.pos.reserve(sizeof(double4) * S, false, st);
    B.vel.reserve(sizeof(double4) * S, false, st);
    B.force.reserve(sizeof(double4) * S, false, st);
    B.snap_pos.reserve(sizeof(double4) * S, false, st);
    B.snap_vel.reserve(sizeof(double4) * S, false, st);
    B.snap_force.reserve(sizeof(double4) * S, false, st);
    B.g_owner.reserve(sizeof(int) * S, false, st);
    B.g_code.reserve(sizeof(int) * S, false, st);
    B.g_ko.reserve(sizeof(int) * S, false, st);
    B.g_kg.reserve(sizeof(int) * S, false, st);
    B.cell_of.reserve(sizeof(int) * S, false, st);
    B.cell_slot.reserve(sizeof(int) * S, false, st);
    B.cell_inv.reserve(sizeof(int) * S, false, st);
    B.parts.reserve(sizeof(int) * S, false, st);
    B.spos.reserve((size_t)kMirBytes * S, false, st);
    B.spos2.reserve((size_t)kMirBytes * S, false, st);
    B.cell_count.reserve(sizeof(int) * (size_t)B.n_cells, false, st);
    B.cell_start.reserve(sizeof(int) * ((size_t)B.n_cells + 1), false, st);
    B.tile_count.reserve(sizeof(int) * (size_t)B.n_tiles, false, st);
    B.tile_start.reserve(sizeof(int) * ((size_t)B.n_tiles + 1), false, st);
    B.work.reserve(sizeof(int) * ((size_t)B.n_work + 1), false, st);
    B.work_k.reserve(sizeof(int) * ((size_t)B.n_work + 1), false, st);
    B.sb_cnt0.reserve(sizeof(int) * (size_t)B.n_sb, false, st);
    B.sb_cnt1.reserve(sizeof(int) * (size_t)B.n_sb, false, st);
    B.scan_tmp.reserve(sizeof(int) * (size_t)scan_scratch(std::max(B.n_cells, B.n_tiles)), false, st);
    B.nb_count.reserve(sizeof(int) * ((size_t)div_up(B.n_work, 64) * 64), false, st);
    B.dyn.reserve(sizeof(int) * DY_STRIDE * (size_t)B.nrep, false, st);
    B.flags.reserve(sizeof(int) * FL_N, false, st);
    B.epart.reserve(sizeof(double) * 3 * (size_t)B.n_tiles, false, st);
    B.kpart.reserve(sizeof(double) * (size_t)B.n_eb, false, st);
    B.tout.reserve(sizeof(double) * 4 * (size_t)B.nrep, false, st);
    // the per-block tables above were uploaded from local (pageable) vectors
    TMD_HIP(hipStreamSynchronize(st));
}

static void reserve_lists(Batch& B) {
    B.nb_list.reserve(sizeof(nb_t) * (size_t)div_up(B.n_work, 64) * 64 * (size_t)B.list_cap, false, B.stream);
}

static void fill(Batch& B, DevBuf& b, int count, int v) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_b_fill, dim3(div_up(count, kBlk)), dim3(kBlk), 0, B.stream, b.as<int>(), count, v);
}

static LJArgs batch_lj_args(const Batch& B) {
    LJArgs a{};
    const double sr2 = 1.0 / (B.rc * B.rc);
    const double s3 = B.sigma * (B.sigma * B.sigma), s6 = s3 * s3;
    const double sr6 = sr2 * sr2 * sr2 * s6;
    a.rc = B.rc;
    a.rc2 = B.rc * B.rc;
    a.eps = B.eps;
    a.sigma = B.sigma;
    a.sigma6 = s6;
    a.fc = 48.0 * sr6 * (sr6 - 0.5) * sr2 * B.eps;
    a.dfc = 24.0 * B.eps * sr2 * sr2 * sr6 * (4.0 - 14.0 * sr6);
    a.uc = 4.0 * B.eps * sr6 * (sr6 - 1.0);
    a.epsf = (float)B.eps;
    a.sigma6f = (float)s6;
    a.fcf = (float)a.fc;
    a.dfcf = (float)a.dfc;
    a.rc2f = (float)a.rc2;
    a.cap = B.list_cap;
    a.wcap = B.wcap;
    return a;
}

// --- the step pieces (all graph-capturable: kernels only) -----------------
static void enq_rebuild(Batch& B) {
    hipStream_t st = B.stream;
    const RepDesc* R = B.dR.as<RepDesc>();
    hipLaunchKernelGGL(k_b_pbc, dim3(B.n_eb), dim3(kBlk), 0, st, R, (const int*)B.eb_rep.as<int>(), B.pos.as<double4>());
    hipLaunchKernelGGL(k_b_zero_ng, dim3(div_up(B.nrep, kBlk)), dim3(kBlk), 0, st, B.dyn.as<int>(), B.nrep);
    for (int d = 0; d < 3; ++d) {
        hipLaunchKernelGGL(k_b_border_count, dim3(B.n_sb), dim3(kScanThreads), 0, st, R, (const int*)B.sb_rep.as<int>(),
                           (const double4*)B.pos.as<double4>(), (const int*)B.dyn.as<int>(), d, B.sb_cnt0.as<int>(),
                           B.sb_cnt1.as<int>());
        hipLaunchKernelGGL(k_b_border_scan, dim3(B.nrep), dim3(kScanThreads), 0, st, R, d, B.sb_cnt0.as<int>(),
                           B.sb_cnt1.as<int>(), B.dyn.as<int>(), B.flags.as<int>());
        hipLaunchKernelGGL(k_b_border_scatter, dim3(B.n_sb), dim3(kScanThreads), 0, st, R, (const int*)B.sb_rep.as<int>(),
                           B.pos.as<double4>(), B.g_owner.as<int>(), B.g_code.as<int>(), (const int*)B.sb_cnt0.as<int>(),
                           (const int*)B.sb_cnt1.as<int>(), (const int*)B.dyn.as<int>(), d);
    }
    fill(B, B.cell_count, B.n_cells, 0);
    hipLaunchKernelGGL(k_b_bin_count, dim3(B.n_cb), dim3(kBlk), 0, st, R, (const int*)B.cb_rep.as<int>(),
                       (const int*)B.dyn.as<int>(), (const double4*)B.pos.as<double4>(), B.cell_of.as<int>(),
                       B.cell_slot.as<int>(), B.cell_count.as<int>(), B.cell_inv.as<int>());
    scan_exclusive(B.cell_count.as<int>(), B.cell_start.as<int>(), B.n_cells, B.cell_start.as<int>() + B.n_cells,
                   B.scan_tmp.as<int>(), st);
    hipLaunchKernelGGL(k_b_bin_scatter, dim3(B.n_cb), dim3(kBlk), 0, st, R, (const int*)B.cb_rep.as<int>(),
                       (const int*)B.dyn.as<int>(), (const int*)B.cell_of.as<int>(), (const int*)B.cell_slot.as<int>(),
                       (const int*)B.cell_start.as<int>(), B.parts.as<int>());
    hipLaunchKernelGGL(k_b_bin_sort, dim3(div_up(B.n_cells, kBlk / kWave)), dim3(kBlk), 0, st, R, (const int*)B.cell_rep.as<int>(),
                       (const int*)B.cell_start.as<int>(), B.n_cells, B.parts.as<int>(), B.cell_count.as<int>(),
                       B.cell_inv.as<int>());
    hipLaunchKernelGGL(k_b_ghost_src, dim3(B.n_gb), dim3(kBlk), 0, st, R, (const int*)B.gb_rep.as<int>(),
                       (const int*)B.dyn.as<int>(), (const int*)B.g_owner.as<int>(), (const int*)B.cell_inv.as<int>(),
                       B.g_ko.as<int>(), B.g_kg.as<int>());
    hipLaunchKernelGGL(k_b_tile_stats, dim3(div_up(B.n_tiles, kBlk)), dim3(kBlk), 0, st, R, (const int*)B.tile_rep.as<int>(),
                       (const int*)B.cell_start.as<int>(), (const int*)B.cell_count.as<int>(), B.n_tiles,
                       B.tile_count.as<int>(), B.flags.as<int>());
    scan_exclusive(B.tile_count.as<int>(), B.tile_start.as<int>(), B.n_tiles, B.tile_start.as<int>() + B.n_tiles,
                   B.scan_tmp.as<int>(), st);
    hipLaunchKernelGGL(k_b_tile_fill, dim3(B.n_tiles), dim3(64), 0, st, R, (const int*)B.tile_rep.as<int>(),
                       (const int*)B.cell_start.as<int>(), (const int*)B.cell_count.as<int>(),
                       (const int*)B.tile_start.as<int>(), (const int*)B.parts.as<int>(), B.work.as<int>(),
                       B.work_k.as<int>());
    hipLaunchKernelGGL(k_b_check_work, dim3(div_up(B.nrep, kBlk)), dim3(kBlk), 0, st, R, B.nrep,
                       (const int*)B.tile_start.as<int>(), B.flags.as<int>());
    TMD_HIP(hipGetLastError());
}

static void enq_mirror(Batch& B, DevBuf& m) {
    hipLaunchKernelGGL(k_b_mirror, dim3(div_up(B.n_slots, kBlk)), dim3(kBlk), 0, B.stream, (const double4*)B.pos.as<double4>(),
                       (const int*)B.parts.as<int>(), (const int*)(B.cell_start.as<int>() + B.n_cells), Mir{m.as<double>()},
                       B.n_slots);
}

static void enq_unmirror(Batch& B, const DevBuf& m) {
    hipLaunchKernelGGL(k_b_unmirror, dim3(B.n_eb), dim3(kBlk), 0, B.stream, (const RepDesc*)B.dR.as<RepDesc>(),
                       (const int*)B.eb_rep.as<int>(), (const int*)B.cell_inv.as<int>(), CMir{m.as<double>()},
                       B.pos.as<double4>());
}

// LDS layout of the list-build launch: the row starts of the largest tile
// shape among the replicas (rsb bytes) after the staged window, whose atom
// capacity is what the budget leaves.  These values are baked into a
// captured run, so graph_signature includes them.
struct NlistLds {
    int rsb, wcap;
    size_t bytes;
};

static NlistLds nlist_lds(const Batch& B) {
    NlistLds l;
    l.rsb = 0;
    for (const RepDesc& D : B.R) l.rsb = std::max(l.rsb, row_starts_bytes(D.g));
    l.wcap = std::min(B.wcap, (kLdsBudget - l.rsb) / kStageBytes);
    l.bytes = (size_t)l.wcap * kStageBytes + l.rsb;
    return l;
}

static void enq_nlist(Batch& B, const DevBuf& m) {
    NbArgs a{};
    a.cap = B.list_cap;
    a.half = 0;
    a.half_mode = HALF_GHOST_FULL;
    const NlistLds l = nlist_lds(B);
    a.wcap = l.wcap;
    const double cut = B.rc + B.skin;
    a.cut2 = cut * cut;
    const size_t lds = l.bytes;
    hipLaunchKernelGGL(k_b_nlist, dim3(B.n_tiles), dim3(B.threads), lds, B.stream, (const RepDesc*)B.dR.as<RepDesc>(),
                       (const int*)B.tile_rep.as<int>(), CMir{m.as<double>()}, (const int*)B.work_k.as<int>(),
                       (const int*)B.cell_of.as<int>(), (const int*)B.cell_start.as<int>(),
                       (const int*)B.parts.as<int>(),
                       (const int*)B.work.as<int>(), (const int*)B.tile_start.as<int>(), a, B.nb_count.as<int>(),
                       B.nb_list.as<nb_t>(), B.flags.as<int>());
    TMD_HIP(hipGetLastError());
}

static void enq_ghost_step_mirror(Batch& B, DevBuf& m) {
    hipLaunchKernelGGL(k_b_ghost_step_mirror, dim3(B.n_gb), dim3(kBlk), 0, B.stream, (const RepDesc*)B.dR.as<RepDesc>(),
                       (const int*)B.gb_rep.as<int>(), (const int*)B.dyn.as<int>(), (const int*)B.g_code.as<int>(),
                       (const int*)B.g_ko.as<int>(), (const int*)B.g_kg.as<int>(), Mir{m.as<double>()});
}

static void enq_ghost_step(Batch& B, DevBuf& m) {
    hipLaunchKernelGGL(k_b_ghost_step, dim3(B.n_gb), dim3(kBlk), 0, B.stream, (const RepDesc*)B.dR.as<RepDesc>(),
                       (const int*)B.gb_rep.as<int>(), (const int*)B.dyn.as<int>(), B.pos.as<double4>(),
                       (const int*)B.g_owner.as<int>(), (const int*)B.g_code.as<int>(), (const int*)B.cell_inv.as<int>(),
                       Mir{m.as<double>()});
}

// kick: final kick v += F dt/2 (hdt = 0 leaves v unchanged: v + f * 0 == v);
// advance: also the next step's kick and drift, into the mirror `next`
static void enq_force(Batch& B, bool kick, bool advance, const DevBuf& cur, DevBuf& next) {
    const LJArgs a = batch_lj_args(B);
    const size_t lds = (size_t)B.wcap * kStageBytes;
    const bool unit = B.eps == 1.0 && B.sigma == 1.0;
    const double hdt = kick ? 0.5 * B.dt : 0.0;
#define TMD_B_LJ_ADV(P, SH, U, ADV)                                                                                 \
    hipLaunchKernelGGL((k_b_lj_full<P, SH, U, ADV>), dim3(B.n_tiles), dim3(B.threads), lds, B.stream,                \
                       (const RepDesc*)B.dR.as<RepDesc>(), (const int*)B.tile_rep.as<int>(), CMir{cur.as<double>()},          \
                       (const int*)B.cell_start.as<int>(), (const int*)B.work.as<int>(), (const int*)B.work_k.as<int>(), \
                       (const int*)B.tile_start.as<int>(), (const int*)B.nb_count.as<int>(), (const nb_t*)B.nb_list.as<nb_t>(), \
                       a, B.force.as<double4>(), B.vel.as<double4>(), hdt, B.dt, Mir{next.as<double>()})
#define TMD_B_LJ_A(P, SH, U) \
    do { if (advance) TMD_B_LJ_ADV(P, SH, U, 1); else TMD_B_LJ_ADV(P, SH, U, 0); } while (0)
#define TMD_B_LJ_U(P, SH) \
    do { if (unit) TMD_B_LJ_A(P, SH, 1); else TMD_B_LJ_A(P, SH, 0); } while (0)
#define TMD_B_LJ_S(P) \
    do { if (B.shift) TMD_B_LJ_U(P, 1); else TMD_B_LJ_U(P, 0); } while (0)
    if (B.precision == PREC_FP64) TMD_B_LJ_S(PREC_FP64);
    else if (B.precision == PREC_FP32_KAHAN) TMD_B_LJ_S(PREC_FP32_KAHAN);
    else TMD_B_LJ_S(PREC_FP32_MIXED);
#undef TMD_B_LJ_S
#undef TMD_B_LJ_U
#undef TMD_B_LJ_A
#undef TMD_B_LJ_ADV
    TMD_HIP(hipGetLastError());
}

static void enq_int1(Batch& B) {
    hipLaunchKernelGGL(k_b_int1, dim3(B.n_eb), dim3(kBlk), 0, B.stream, (const RepDesc*)B.dR.as<RepDesc>(),
                       (const int*)B.eb_rep.as<int>(), B.pos.as<double4>(), B.vel.as<double4>(),
                       (const double4*)B.force.as<double4>(), 0.5 * B.dt, B.dt, (const int*)B.cell_inv.as<int>(),
                       Mir{B.spos.as<double>()});
}

// the main.cpp step loop over `steps` steps (rebuild when (j + 1) % reneigh
// == 0); each force pass also performs the final kick of its step and the
// first kick + drift of the next, so only the first step integrates alone.
// Positions live in the two mirrors during the run (the force pass reads one
// and writes the other; the enqueue alternates them); pos is restored from
// the mirror before each rebuild and at the end.  The first integration
// writes pos and the mirror, so a run depends only on pos, vel and force.
static void enq_steps(Batch& B, int steps) {
    DevBuf* cur = &B.spos;
    DevBuf* nxt = &B.spos2;
    if (steps > 0) enq_int1(B);
    for (int j = 0; j < steps; ++j) {
        if ((j + 1) % B.reneigh == 0) {
            enq_unmirror(B, *cur);
            enq_rebuild(B);
            enq_mirror(B, *cur);
            enq_nlist(B, *cur);
        } else {
            enq_ghost_step_mirror(B, *cur);
        }
        const bool adv = j + 1 < steps;
        enq_force(B, true, adv, *cur, *nxt);
        if (adv) std::swap(cur, nxt);
    }
    if (steps > 0) {
        enq_ghost_step(B, *cur);  // the ghosts' pos entries (and the same mirror entries)
        enq_unmirror(B, *cur);
    }
    TMD_HIP(hipGetLastError());
}

static void read_flags(Batch& B, int* host) {
    TMD_HIP(hipMemcpyAsync(host, B.flags.p, sizeof(int) * FL_N, hipMemcpyDeviceToHost, B.stream));
    TMD_HIP(hipStreamSynchronize(B.stream));
}

static void upload_initial(Batch& B) {
    std::vector<double4> p((size_t)B.n_slots, make_double4(0, 0, 0, 0)), v((size_t)B.n_slots, make_double4(0, 0, 0, 0));
    size_t off = 0;
    for (int r = 0; r < B.nrep; ++r) {
        for (int i = 0; i < B.n[r]; ++i) {
            const size_t s = (size_t)B.R[r].a0 + i, k = 3 * (off + i);
            p[s] = make_double4(B.h_pos[k], B.h_pos[k + 1], B.h_pos[k + 2], 0.0);
            v[s] = make_double4(B.h_vel[k], B.h_vel[k + 1], B.h_vel[k + 2], 0.0);
        }
        off += B.n[r];
    }
    TMD_HIP(hipMemcpyAsync(B.pos.p, p.data(), sizeof(double4) * p.size(), hipMemcpyHostToDevice, B.stream));
    TMD_HIP(hipMemcpyAsync(B.vel.p, v.data(), sizeof(double4) * v.size(), hipMemcpyHostToDevice, B.stream));
    TMD_HIP(hipStreamSynchronize(B.stream));
}

// Grow every replica's ghost room (a ghost overflow), re-lay out the
// buffers and carry the locals' positions and velocities over to their new
// slots.  Every copy is ordered on the batch stream and waited for: the
// read-back must see the work queued before it, and the re-laid-out
// buffers must hold the state before anything is enqueued behind it.
static void grow_ghost_room(Batch& B) {
    std::vector<double4> p((size_t)B.n_slots), v((size_t)B.n_slots);
    TMD_HIP(hipMemcpyAsync(p.data(), B.pos.p, sizeof(double4) * p.size(), hipMemcpyDeviceToHost, B.stream));
    TMD_HIP(hipMemcpyAsync(v.data(), B.vel.p, sizeof(double4) * v.size(), hipMemcpyDeviceToHost, B.stream));
    TMD_HIP(hipStreamSynchronize(B.stream));
    if (has_error()) return;
    const std::vector<RepDesc> old = B.R;
    B.ghost_scale *= 1.6;
    B.drop_graphs();  // every captured run refers to the old buffers and tables
    layout(B);
    if (has_error()) return;
    std::vector<double4> p2((size_t)B.n_slots, make_double4(0, 0, 0, 0)), v2((size_t)B.n_slots, make_double4(0, 0, 0, 0));
    for (int r = 0; r < B.nrep; ++r)
        for (int i = 0; i < B.n[r]; ++i) {
            p2[(size_t)B.R[r].a0 + i] = p[(size_t)old[r].a0 + i];
            v2[(size_t)B.R[r].a0 + i] = v[(size_t)old[r].a0 + i];
        }
    TMD_HIP(hipMemcpyAsync(B.pos.p, p2.data(), sizeof(double4) * p2.size(), hipMemcpyHostToDevice, B.stream));
    TMD_HIP(hipMemcpyAsync(B.vel.p, v2.data(), sizeof(double4) * v2.size(), hipMemcpyHostToDevice, B.stream));
    TMD_HIP(hipStreamSynchronize(B.stream));
}

// Initial rebuild + force at the current positions, sizing the ghost room,
// the LDS window and the list capacity.  Host synchronisation allowed.
static bool setup_lists(Batch& B) {
    for (int attempt = 0; attempt < 8 && !has_error(); ++attempt) {
        fill(B, B.flags, FL_N, 0);
        enq_rebuild(B);
        int fl[FL_N];
        read_flags(B, fl);
        if (has_error()) return false;
        if (fl[FL_GHOST] > 0) {
            // ghost room too small: grow, re-lay out, restore the state
            grow_ghost_room(B);
            if (has_error()) return false;
            continue;
        }
        if (fl[FL_UNBINNED] > 0) {
            set_error(__FILE__, __LINE__, "batch: a replica has local atoms outside its cell grid (non-finite positions?)");
            return false;
        }
        if (fl[FL_WIN] > 0) {
            set_error(__FILE__, __LINE__, "batch: a tile neighbourhood exceeds the 16-bit window range; use a smaller tile");
            return false;
        }
        const int maxwin = fl[FL_MAXWIN];
        B.threads = tile_block_threads(B.cfg.threads, fl[FL_MAXLOC]);
        // capacities only grow, so a restart keeps a captured run valid
        B.wcap = std::max(B.wcap, std::min(maxwin + maxwin / 4 + 64, kLdsBudget / kStageBytes));
        enq_mirror(B, B.spos);
        for (int k = 0; k < 4 && !has_error(); ++k) {
            reserve_lists(B);
            fill(B, B.flags, FL_N, 0);
            enq_nlist(B, B.spos);
            read_flags(B, fl);
            if (fl[FL_LIST] == 0) break;
            B.list_cap = std::max(B.list_cap, ((int)std::ceil(fl[FL_LIST] * B.cfg.list_headroom) + 7) & ~7);
        }
        if (has_error()) return false;
        if (fl[FL_LIST] > 0) continue;
        enq_force(B, false, false, B.spos, B.spos2);
        TMD_HIP(hipStreamSynchronize(B.stream));
        B.ready = !has_error();
        return B.ready;
    }
    if (!has_error()) set_error(__FILE__, __LINE__, "batch: set-up did not converge");
    return false;
}

// fresh: size the list capacity and the LDS window anew (new configuration);
// otherwise keep the (only growing) capacities of earlier runs
static bool setup_from_initial(Batch& B, bool fresh = true) {
    B.ready = false;
    layout(B);
    if (has_error()) return false;
    if (fresh) {
        B.list_cap = 8;
        B.wcap = 0;
    }
    upload_initial(B);
    if (has_error()) return false;
    return setup_lists(B);
}

// Everything a captured run bakes in: buffer addresses and launch scalars.
static std::vector<unsigned long long> graph_signature(const Batch& B) {
    std::vector<unsigned long long> s;
    const DevBuf* bufs[] = {&B.dR, &B.sb_rep, &B.eb_rep, &B.cb_rep, &B.gb_rep, &B.tile_rep, &B.cell_rep, &B.pos, &B.vel,
                            &B.force, &B.g_owner, &B.g_code, &B.cell_of, &B.cell_slot, &B.cell_inv, &B.parts, &B.spos,
                            &B.spos2, &B.cell_count, &B.cell_start, &B.tile_count, &B.tile_start, &B.work, &B.work_k,
                            &B.g_ko, &B.g_kg,
                            &B.sb_cnt0, &B.sb_cnt1,
                            &B.scan_tmp, &B.nb_count, &B.nb_list, &B.dyn, &B.flags};
    for (const DevBuf* b : bufs) s.push_back((unsigned long long)(uintptr_t)b->p);
    const double dv[] = {B.rc, B.skin, B.eps, B.sigma, B.dt};
    for (double d : dv) {
        unsigned long long u;
        std::memcpy(&u, &d, sizeof u);
        s.push_back(u);
    }
    const NlistLds l = nlist_lds(B);
    const int iv[] = {B.reneigh, B.precision, B.shift, B.list_cap, B.wcap, B.threads, B.n_sb, B.n_eb, B.n_cb, B.n_gb,
                      B.n_tiles, B.n_cells, B.n_slots, B.nrep, l.rsb, l.wcap, (int)l.bytes};
    for (int v : iv) s.push_back((unsigned long long)(unsigned)v);
    // the host-side tile shapes behind the launch sizes above: two layouts
    // with equal totals but different tile edges must not share a graph
    for (const RepDesc& D : B.R) {
        const int gv[] = {D.g.nx, D.g.ny, D.g.nz, D.g.tx, D.g.ty, D.g.tz, D.g.ntx, D.g.nty, D.g.ntz};
        for (int v : gv) s.push_back((unsigned long long)(unsigned)v);
    }
    return s;
}

static hipGraphExec_t get_graph(Batch& B, int steps) {
    const std::vector<unsigned long long> sig = graph_signature(B);
    auto it = B.graphs.find(steps);
    if (it != B.graphs.end()) {
        if (it->second.sig == sig) return it->second.exec;
        (void)hipGraphExecDestroy(it->second.exec);
        (void)hipGraphDestroy(it->second.graph);
        B.graphs.erase(it);
    }
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    TMD_HIP(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
    if (has_error()) return nullptr;
    enq_steps(B, steps);
    TMD_HIP(hipStreamEndCapture(B.stream, &g));
    if (has_error() || !g) return nullptr;
    TMD_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    if (has_error()) {
        (void)hipGraphDestroy(g);
        return nullptr;
    }
    B.graphs[steps] = Batch::CapturedRun{g, ge, sig};
    return ge;
}

static void copy_all(Batch& B, DevBuf& px, DevBuf& pv, DevBuf& pf, const DevBuf& sx, const DevBuf& sv, const DevBuf& sf) {
    const size_t bytes = sizeof(double4) * (size_t)B.n_slots;
    TMD_HIP(hipMemcpyAsync(px.p, sx.p, bytes, hipMemcpyDeviceToDevice, B.stream));
    TMD_HIP(hipMemcpyAsync(pv.p, sv.p, bytes, hipMemcpyDeviceToDevice, B.stream));
    TMD_HIP(hipMemcpyAsync(pf.p, sf.p, bytes, hipMemcpyDeviceToDevice, B.stream));
}

// `steps` steps as one graph launch; on a capacity flag, grow and redo the
// run from its starting state (after a rebuild there).
static bool run_steps(Batch& B, int steps, double* gpu_ms) {
    if (!B.ready) {
        set_error(__FILE__, __LINE__, "batch: not set up");
        return false;
    }
    if (steps <= 0) return true;
    for (int attempt = 0; attempt < 6 && !has_error(); ++attempt) {
        copy_all(B, B.snap_pos, B.snap_vel, B.snap_force, B.pos, B.vel, B.force);
        fill(B, B.flags, FL_N, 0);
        hipGraphExec_t ge = nullptr;
        if (B.use_graph) {
            ge = get_graph(B, steps);
            if (!ge) return false;
        }
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (gpu_ms) {
            TMD_HIP(hipEventCreate(&e0));
            TMD_HIP(hipEventCreate(&e1));
            TMD_HIP(hipEventRecord(e0, B.stream));
        }
        if (ge) TMD_HIP(hipGraphLaunch(ge, B.stream));
        else enq_steps(B, steps);  // the same kernels, launched one by one
        if (gpu_ms) TMD_HIP(hipEventRecord(e1, B.stream));
        int fl[FL_N];
        read_flags(B, fl);
        if (gpu_ms) {
            float ms = 0.f;
            TMD_HIP(hipEventElapsedTime(&ms, e0, e1));
            *gpu_ms = ms;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        if (has_error()) return false;
        if (fl[FL_UNBINNED] > 0) {
            set_error(__FILE__, __LINE__, "batch: a replica has local atoms outside its cell grid (non-finite positions?)");
            return false;
        }
        if (fl[FL_WIN] > 0) {
            set_error(__FILE__, __LINE__, "batch: a tile neighbourhood exceeds the 16-bit window range; use a smaller tile");
            return false;
        }
        if (fl[FL_GHOST] == 0 && fl[FL_LIST] == 0) return true;
        // overflow: restore the starting state, grow, rebuild there, rerun.
        // The restore copies run on the batch stream, which does not
        // synchronise with the null stream: wait for them before the state
        // is read back (grow_ghost_room) or rebuilt.
        copy_all(B, B.pos, B.vel, B.force, B.snap_pos, B.snap_vel, B.snap_force);
        TMD_HIP(hipStreamSynchronize(B.stream));
        if (fl[FL_LIST] > 0) B.list_cap = ((int)std::ceil(fl[FL_LIST] * B.cfg.list_headroom) + 7) & ~7;
        B.drop_graphs();
        if (fl[FL_GHOST] > 0) {
            const int keep_cap = B.list_cap;
            grow_ghost_room(B);
            B.list_cap = keep_cap;
            if (has_error()) return false;
        }
        if (!setup_lists(B)) return false;
    }
    if (!has_error()) set_error(__FILE__, __LINE__, "batch: capacity retries exhausted");
    return false;
}

static void batch_thermo(Batch& B, double* out) {
    const LJArgs a = batch_lj_args(B);
    const size_t lds = (size_t)B.wcap * kStageBytes;
    enq_mirror(B, B.spos);
#define TMD_B_EN(SH)                                                                                                 \
    hipLaunchKernelGGL(k_b_energy<SH>, dim3(B.n_tiles), dim3(B.threads), lds, B.stream,                          \
                       (const RepDesc*)B.dR.as<RepDesc>(), (const int*)B.tile_rep.as<int>(), (const double4*)B.pos.as<double4>(), \
                       CMir{B.spos.as<double>()}, (const int*)B.cell_start.as<int>(), (const int*)B.parts.as<int>(),        \
                       (const int*)B.work.as<int>(), (const int*)B.tile_start.as<int>(), (const int*)B.nb_count.as<int>(), \
                       (const nb_t*)B.nb_list.as<nb_t>(), a, B.epart.as<double>())
    if (B.shift) TMD_B_EN(1); else TMD_B_EN(0);
#undef TMD_B_EN
    hipLaunchKernelGGL(k_b_kinetic, dim3(B.n_eb), dim3(kBlk), 0, B.stream, (const RepDesc*)B.dR.as<RepDesc>(),
                       (const int*)B.eb_rep.as<int>(), (const double4*)B.vel.as<double4>(), B.kpart.as<double>());
    hipLaunchKernelGGL(k_b_thermo_reduce, dim3(B.nrep), dim3(kBlk), 0, B.stream, (const RepDesc*)B.dR.as<RepDesc>(),
                       (const double*)B.epart.as<double>(), (const double*)B.kpart.as<double>(), B.tout.as<double>());
    TMD_HIP(hipGetLastError());
    std::vector<double> t(4 * (size_t)B.nrep);
    TMD_HIP(hipMemcpyAsync(t.data(), B.tout.p, sizeof(double) * t.size(), hipMemcpyDeviceToHost, B.stream));
    TMD_HIP(hipStreamSynchronize(B.stream));
    if (has_error()) return;
    for (int r = 0; r < B.nrep; ++r) {
        const double* q = &t[4 * (size_t)r];
        const double vol = B.R[r].len[0] * B.R[r].len[1] * B.R[r].len[2];
        const int n = B.n[r];
        double* o = out + 6 * (size_t)r;
        o[0] = q[3] * (1.0 / (double)(n * 3 - 3));
        o[1] = 0.5 * q[3];
        o[2] = q[0];
        o[3] = q[1];
        o[4] = (q[3] + q[1]) / (3.0 * vol);
        o[5] = q[2];
    }
}

}  // namespace tmd

using namespace tmd;

extern "C" {

int tmd_generate_replica_set(unsigned long long seed, int nrep, int cells_min, int cells_max, double rho_min, double rho_max,
                             double temp_min, double temp_max, int* cells, double* rho, double* temp) {
    // SplitMix64 finaliser over (seed, replica, draw): a counter-based
    // stream, so replica r's parameters do not depend on nrep
    auto mix = [](unsigned long long z) {
        z += 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto unit = [&](int r, int k) {
        const unsigned long long h = mix(seed ^ mix(3ull * (unsigned long long)r + (unsigned long long)k));
        return (double)(h >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
    };
    if (nrep <= 0 || cells_max < cells_min) return 0;
    for (int r = 0; r < nrep; ++r) {
        const int span = cells_max - cells_min + 1;
        int c = cells_min + (int)std::floor(unit(r, 0) * span);
        cells[r] = std::min(c, cells_max);
        rho[r] = rho_min + unit(r, 1) * (rho_max - rho_min);
        temp[r] = temp_min + unit(r, 2) * (temp_max - temp_min);
    }
    return nrep;
}

int tmd_batch_create(int nrep, const int* counts, const struct Vector3D* pos, const struct Vector3D* vel, const double* aabbs) {
    if (has_error()) return 0;
    if (nrep <= 0) {
        set_error(__FILE__, __LINE__, "batch: nrep must be positive");
        return 0;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error(__FILE__, __LINE__, "no HIP device available");
        return 0;
    }
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, 0) != hipSuccess || lds < kLdsBudget) {
        set_error(__FILE__, __LINE__, "batch: device LDS per block is below the window budget");
        return 0;
    }
    std::unique_ptr<Batch> B(new Batch());
    B->nrep = nrep;
    B->n.assign(counts, counts + nrep);
    B->aabb.assign(aabbs, aabbs + 6 * (size_t)nrep);
    size_t total = 0;
    for (int r = 0; r < nrep; ++r) {
        if (counts[r] < 2) {
            set_error(__FILE__, __LINE__, "batch: every replica needs at least two atoms");
            return 0;
        }
        total += (size_t)counts[r];
    }
    const double* p = reinterpret_cast<const double*>(pos);
    const double* v = reinterpret_cast<const double*>(vel);
    B->h_pos.assign(p, p + 3 * total);
    B->h_vel.assign(v, v + 3 * total);
    TMD_HIP(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    if (has_error()) return 0;
    const int h = g_next_batch++;
    g_batches[h] = std::move(B);
    return h;
}

int tmd_batch_setup(int h, double cutoff, double skin, double epsilon, double sigma, double dt, int reneigh_every,
                    int precision, int lj_shift) {
    Batch* B = find_batch(h);
    if (!B || has_error()) return 0;
    if (reneigh_every <= 0 || !(cutoff > 0) || !(skin >= 0) || !(dt > 0)) {
        set_error(__FILE__, __LINE__, "batch: invalid parameters");
        return 0;
    }
    B->rc = cutoff;
    B->skin = skin;
    B->eps = epsilon;
    B->sigma = sigma;
    B->dt = dt;
    B->reneigh = reneigh_every;
    B->precision = precision == 2 ? PREC_FP32_KAHAN : (precision ? PREC_FP32_MIXED : PREC_FP64);
    B->shift = lj_shift ? 1 : 0;
    return setup_from_initial(*B) ? 1 : 0;
}

int tmd_batch_configure(int h, int tile, int threads, double spacing_factor, double list_headroom, double ghost_headroom) {
    Batch* B = find_batch(h);
    if (!B || has_error()) return 0;
    if ((threads != 0 && (threads < 64 || threads > kMaxTileThreads || threads % 64 != 0)) || tile < 0 || tile > kMaxTile ||
        !(spacing_factor >= 1.0) || !(list_headroom >= 1.0) || !(ghost_headroom > 0.0)) {
        set_error(__FILE__, __LINE__, "batch: invalid configuration");
        return 0;
    }
    B->cfg.tile = tile;
    B->cfg.threads = threads;
    B->cfg.spacing = spacing_factor;
    B->cfg.list_headroom = list_headroom;
    B->cfg.ghost_headroom = ghost_headroom;
    B->ghost_scale = 1.0;
    return setup_from_initial(*B) ? 1 : 0;
}

int tmd_batch_set_graph(int h, int on) {
    Batch* B = find_batch(h);
    if (!B || has_error()) return 0;
    B->use_graph = on != 0;
    if (!B->use_graph) B->drop_graphs();
    return 1;
}

int tmd_batch_run(int h, int steps) {
    Batch* B = find_batch(h);
    if (!B || has_error()) return 0;
    return run_steps(*B, steps, nullptr) ? steps : 0;
}

int tmd_batch_run_timed(int h, int steps, double* gpu_ms) {
    Batch* B = find_batch(h);
    if (!B || has_error()) return 0;
    return run_steps(*B, steps, gpu_ms) ? steps : 0;
}

int tmd_batch_autotune(int h, int trial_steps, double* out) {
    Batch* B = find_batch(h);
    if (!B || has_error()) return 0;
    if (trial_steps <= 0) trial_steps = 2 * B->reneigh;
    const BatchCfg keep = B->cfg;
    double best = 1e300;
    BatchCfg best_cfg = keep;
    // one trial: set up from the initial state, a warm run (graph capture),
    // the same start again, a timed run
    auto trial = [&](int t, int th, double sp, double hr) -> bool {
        B->cfg = keep;
        B->cfg.tile = t;
        B->cfg.threads = th;
        B->cfg.spacing = sp;
        B->cfg.list_headroom = hr;
        B->ghost_scale = 1.0;
        double ms = 0.0;
        if (!setup_from_initial(*B) || !run_steps(*B, trial_steps, nullptr)) return false;
        if (!setup_from_initial(*B) || !run_steps(*B, trial_steps, &ms)) return false;
        if (ms < best) {
            best = ms;
            best_cfg = B->cfg;
        }
        return true;
    };
    // stage 1: the launch shape (cells per tile edge x threads per block) at
    // the reference's cell edge and the smaller list headroom; stage 2: cell
    // edge and list headroom around the best shape (15 trials, not 48)
    const int tiles[] = {0, 2, 3, 4};
    const int threads[] = {0, 256, 512};  // 0: one lane per local of the fullest tile
    for (int t : tiles)
        for (int th : threads)
            if (!trial(t, th, 1.0, 1.1)) return 0;
    const BatchCfg shape = best_cfg;
    const double spacing[] = {1.0, 1.1};
    const double headroom[] = {1.1, 1.3};
    for (double sp : spacing)
        for (double hr : headroom)
            if (!(sp == 1.0 && hr == 1.1) && !trial(shape.tile, shape.threads, sp, hr)) return 0;
    B->cfg = best_cfg;
    B->ghost_scale = 1.0;
    if (!setup_from_initial(*B)) return 0;
    if (out) {
        out[0] = best_cfg.tile;
        out[1] = B->threads;  // the block size in use (automatic when the configuration says 0)
        out[2] = best_cfg.spacing;
        out[3] = best_cfg.list_headroom;
        out[4] = best / trial_steps;
    }
    return 1;
}

int tmd_batch_reset(int h) {
    Batch* B = find_batch(h);
    if (!B || has_error()) return 0;
    return setup_from_initial(*B, false) ? 1 : 0;
}

int tmd_batch_get_state(int h, int rep, struct Vector3D* pos, struct Vector3D* vel, struct Vector3D* force) {
    Batch* B = find_batch(h);
    if (!B || has_error()) return 0;
    if (rep < 0 || rep >= B->nrep) {
        set_error(__FILE__, __LINE__, "batch: replica index out of range");
        return 0;
    }
    const int n = B->n[rep];
    const size_t a0 = (size_t)B->R[rep].a0;
    std::vector<double4> tmp((size_t)n);
    struct Vector3D* outs[3] = {pos, vel, force};
    DevBuf* srcs[3] = {&B->pos, &B->vel, &B->force};
    for (int k = 0; k < 3; ++k) {
        if (!outs[k]) continue;
        TMD_HIP(hipMemcpyAsync(tmp.data(), srcs[k]->as<double4>() + a0, sizeof(double4) * n, hipMemcpyDeviceToHost, B->stream));
        TMD_HIP(hipStreamSynchronize(B->stream));
        if (has_error()) return 0;
        for (int i = 0; i < n; ++i) outs[k][i] = Vector3D{tmp[i].x, tmp[i].y, tmp[i].z};
    }
    return n;
}

int tmd_batch_thermo(int h, double* out) {
    Batch* B = find_batch(h);
    if (!B || has_error()) return 0;
    if (!B->ready) {
        set_error(__FILE__, __LINE__, "batch: not set up");
        return 0;
    }
    batch_thermo(*B, out);
    return has_error() ? 0 : B->nrep;
}

int tmd_batch_info(int h, int* out) {
    Batch* B = find_batch(h);
    if (!B || has_error()) return 0;
    std::vector<int> dy(DY_STRIDE * (size_t)B->nrep, 0);
    if (B->ready) {
        TMD_HIP(hipMemcpyAsync(dy.data(), B->dyn.p, sizeof(int) * dy.size(), hipMemcpyDeviceToHost, B->stream));
        TMD_HIP(hipStreamSynchronize(B->stream));
    }
    for (int r = 0; r < B->nrep; ++r) {
        int* o = out + 8 * (size_t)r;
        const RepDesc& D = B->R[r];
        o[0] = D.n;
        o[1] = dy[r * DY_STRIDE + DY_NG];
        o[2] = D.g.nx;
        o[3] = D.ncells;
        o[4] = D.ntiles;
        o[5] = D.g.tx;
        o[6] = D.cap - D.n;
        o[7] = B->list_cap;
    }
    return B->nrep;
}

void tmd_batch_destroy(int h) {
    auto it = g_batches.find(h);
    if (it != g_batches.end()) g_batches.erase(it);
}

}  // extern "C"
