// Tile windows: LDS staging of the neighbourhood of a tile of cells.
//
// The cells that can hold local atoms (1..n-2 in each dimension) are cut
// into tiles of TX x TY x TZ cells.  The work order
// (local atoms in the order the list and force kernels walk them) is
// tile-major: tiles in (z, y, x) order, the cells of a tile in (z, y, x)
// order, the locals of a cell in index order.  One block handles one tile.
// The window of a tile is the box of cells one wider on every side; each of
// its (TY+2)(TZ+2) x-rows is contiguous in the cell CSR, so the window is the
// concatenation of those CSR ranges, staged into LDS once per block from
// the cell-ordered position mirror (contiguous loads).  A list entry is a
// 16-bit offset into the window.  An atom's 27-cell stencil is 9 rows of 3
// cells, each inside one window row, so lists keep the reference's stencil
// order (core/neighborlist.impala:1-16) and cell order within each cell.
//
// Locals always lie in cells 1..n-2 of each dimension (the grid is the box
// extended by one cell on each side), so their stencils never wrap and
// always lie inside their tile's window.
#pragma once

#include "tmd_device.hpp"

namespace tmd {

constexpr int kTileThreads = 512;  // threads per tile block (default)
constexpr int kMaxTileThreads = 1024;  // launch bound of the single-box tile kernels
constexpr int kAutoTileThreads = 768;   // largest automatic block (two blocks per CU at <= 80 VGPRs)
constexpr int kMaxTile = 8;        // cells per tile dimension, at most
constexpr int kMaxSeg = (kMaxTile + 2) * (kMaxTile + 2);
constexpr int kWinMax = 65535;     // window offsets are 16-bit
// dynamic LDS per block: two blocks per 160 KiB CU, with 1 KiB left for each
// block's static LDS (window rows, reductions)
constexpr int kLdsBudget = 80 * 1024 - 1024;
constexpr int kStageBytes = 28;       // LDS bytes per staged window atom: x, y, z 24, index 4
constexpr int kPosStageBytes = 24;    // the same without the index (kernels that stage positions only)
constexpr int kDemStageBytes = 48;    // DEM step pass: position and velocity, 24 + 24
// LDS share per block of the DEM step pass: three blocks per CU (160 KiB of
// LDS) less the static tile table
constexpr int kDemLdsShare = 160 * 1024 / 3 - 1024;

// The cell-ordered mirrors of the single-box path (positions, velocities)
// hold x, y, z doubles packed at 24 B per atom, a quarter less than double4
// records.  These views read and write them as double4 values (w = 0).
struct MirRef {
    double* q;
    __device__ __forceinline__ operator double4() const { return make_double4(q[0], q[1], q[2], 0.0); }
    __device__ __forceinline__ void operator=(const double4& v) const {
        q[0] = v.x;
        q[1] = v.y;
        q[2] = v.z;
    }
    // element copy (not a rebinding of the view)
    __device__ __forceinline__ void operator=(const MirRef& o) const { *this = (double4)o; }
};
struct Mir {
    double* p;
    __device__ __forceinline__ MirRef operator[](size_t k) const { return MirRef{p + 3 * k}; }
};
struct CMir {
    const double* p;
    __device__ __forceinline__ double4 operator[](size_t k) const {
        const double* q = p + 3 * k;
        return make_double4(q[0], q[1], q[2], 0.0);
    }
};
constexpr int kMirBytes = 24;  // bytes per atom of a mirror

typedef unsigned short nb_t;

// u16 list layout: tiles of 64 work slots, entry k of column l at
// (k / 8) * 512 + l * 8 + (k % 8); one int4 load reads 8 entries of a lane.
__device__ __forceinline__ size_t nb16_slot(int w, int k, int cap) {
    return (size_t)(w >> 6) * (size_t)cap * 64 + (size_t)(k >> 3) * 512 + (size_t)(w & 63) * 8 + (size_t)(k & 7);
}

struct TileGeom {
    int nx, ny, nz;     // cells
    int tx, ty, tz;     // cells per tile
    int ntx, nty, ntz;  // tiles
};

struct TileWin {
    int cx0, cy0, cz0;          // first cell of the tile
    int nseg;                   // (ty + 2) * (tz + 2)
    int seg_beg[kMaxSeg];       // CSR start of each window row
    int seg_off[kMaxSeg + 1];   // window offset of each row (prefix)
};

// Build the tile's window (whole block; barriers inside).  Row s covers
// cells x in [cx0-1, cx0+tx] (clipped) at y = cy0-1 + s % (ty+2),
// z = cz0-1 + s / (ty+2).
__device__ __forceinline__ void tile_window(int tile, const int* __restrict__ start, TileGeom g, TileWin& W) {
    const int t = threadIdx.x;
    const int tx_i = tile % g.ntx, ty_i = (tile / g.ntx) % g.nty, tz_i = tile / (g.ntx * g.nty);
    const int cx0 = 1 + tx_i * g.tx, cy0 = 1 + ty_i * g.ty, cz0 = 1 + tz_i * g.tz;
    const int wy = g.ty + 2;
    const int nseg = wy * (g.tz + 2);
    for (int s = t; s < nseg; s += blockDim.x) {
        const int y = cy0 - 1 + s % wy, z = cz0 - 1 + s / wy;
        const int xlo = max(cx0 - 1, 0), xhi = min(cx0 + g.tx, g.nx - 1);
        int b = 0, len = 0;
        if (y >= 0 && y < g.ny && z >= 0 && z < g.nz) {
            const int row = (z * g.ny + y) * g.nx;
            b = start[row + xlo];
            len = start[row + xhi + 1] - b;
        }
        W.seg_beg[s] = b;
        W.seg_off[s + 1] = len;
    }
    if (t == 0) {
        W.cx0 = cx0;
        W.cy0 = cy0;
        W.cz0 = cz0;
        W.nseg = nseg;
    }
    __syncthreads();
    if (t < kWave) {
        // the first wavefront turns the row lengths into offsets (inclusive
        // scan, 64 rows at a time; blocks have at least one wavefront)
        int carry = 0;
        for (int s0 = 0; s0 < nseg; s0 += kWave) {
            const int s = s0 + t;
            const int inc = wave_inclusive_scan(s < nseg ? W.seg_off[s + 1] : 0);
            if (s < nseg) W.seg_off[s + 1] = carry + inc;
            carry += __shfl(inc, kWave - 1, kWave);
        }
        if (t == 0) W.seg_off[0] = 0;
    }
    __syncthreads();
}

__device__ __forceinline__ int win_size(const TileWin& W) { return W.seg_off[W.nseg]; }

// CSR index (mirror index) of window offset t: binary search over the rows
__device__ __forceinline__ int win_index(int t, const TileWin& W) {
    int lo = 0, hi = W.nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (W.seg_off[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return W.seg_beg[lo] + (t - W.seg_off[lo]);
}

// Window offsets of the cell starts of every window row, for the list
// build: cst[s * (tx + 3) + (x - (cx0 - 1))] for x in [cx0 - 1, xhi + 1].
// A local atom's stencil row then needs no global load.  Whole block.
// cst holds (ty + 2)(tz + 2)(tx + 3) ints (row_starts_bytes).
__host__ __device__ inline int row_starts_bytes(const TileGeom& g) { return (g.ty + 2) * (g.tz + 2) * (g.tx + 3) * 4; }
__device__ __forceinline__ void stage_row_starts(const TileWin& W, const int* __restrict__ start, TileGeom g, int* cst) {
    const int wy = g.ty + 2;
    const int per = g.tx + 3;
    const int xlo = W.cx0 - 1;
    const int xhi = min(W.cx0 + g.tx, g.nx - 1);
    for (int e = threadIdx.x; e < W.nseg * per; e += blockDim.x) {
        const int s = e / per, xo = e % per;
        const int y = W.cy0 - 1 + s % wy, z = W.cz0 - 1 + s / wy;
        const int x = xlo + xo;
        int v = 0;
        if (y >= 0 && y < g.ny && z >= 0 && z < g.nz && x <= xhi + 1) {
            const int row = (z * g.ny + y) * g.nx;
            v = W.seg_off[s] + (start[row + x] - W.seg_beg[s]);
        }
        cst[s * per + xo] = v;
    }
    __syncthreads();
}

// Stage the window's FP64 positions and, when sj != nullptr, atom indices
// into LDS if it fits wcap atoms; otherwise consumers read the mirror.
template <class SP>
__device__ __forceinline__ bool stage_tile(const TileWin& W, int wcap, const SP& spos,
                                           const int* __restrict__ parts, double* sx, double* sy, double* sz, int* sj) {
    const int S = win_size(W);
    const bool staged = S <= wcap;
    if (staged) {
        // one thread group per window row, as in stage_window
        const int nseg = W.nseg;
        const int gsz = max(1, (int)blockDim.x / nseg);
        const int ngrp = (int)blockDim.x / gsz;
        const int g = (int)threadIdx.x / gsz, l = (int)threadIdx.x - g * gsz;
        if (g < ngrp) {
            for (int s = g; s < nseg; s += ngrp) {
                const int off = W.seg_off[s], beg = W.seg_beg[s], len = W.seg_off[s + 1] - off;
                for (int q = l; q < len; q += gsz) {
                    const int t = off + q, k = beg + q;
                    const double4 p = spos[k];
                    sx[t] = p.x;
                    sy[t] = p.y;
                    sz[t] = p.z;
                    if (sj) sj[t] = parts[k];
                }
            }
        }
    }
    __syncthreads();
    return staged;
}

// LDS image of a tile's window: x, y, z interleaved (three 8-byte loads per
// neighbour; measured faster than 16-byte loads for these scattered
// accesses) and optionally the atom index; kStageBytes per atom.
struct WinStage {
    double* xyz;
    int* j;
};

__device__ __forceinline__ WinStage win_stage(double* dyn, int wcap) {
    WinStage st;
    st.xyz = dyn;
    st.j = reinterpret_cast<int*>(dyn + 3 * wcap);
    return st;
}

// Position of the sentinel slot that follows a staged window (offset
// win_size): list groups are padded with that offset, and its distance to
// any atom of the box is far beyond every cutoff, so a padded entry fails
// the cutoff test like any other out-of-range neighbour.
constexpr double kFarPos = 1.0e18;
// sentinel slots of a list-build window: candidates are read four at a time
// without clamping, up to three past the window end
constexpr int kListPad = 3;

// Stage the window (positions, and atom indices when parts != nullptr) and
// `pad` sentinel slots (offsets S .. S + pad - 1) into LDS if they fit wcap
// atoms.  Whole block.  Returns whether it was staged; otherwise consumers
// read the mirror.
template <class SP>
__device__ __forceinline__ bool stage_window(const TileWin& W, int wcap, const SP& spos,
                                             const int* __restrict__ parts, const WinStage& st, int pad = 1) {
    const int S = win_size(W);
    const bool staged = S + pad <= wcap;
    if (staged) {
        // the block is split into one thread group per window row (rows in
        // turn when there are more rows than groups); a group copies its
        // row's contiguous CSR range, so no offset-to-row search is needed
        const int nseg = W.nseg;
        const int gsz = max(1, (int)blockDim.x / nseg);
        const int ngrp = (int)blockDim.x / gsz;
        const int g = (int)threadIdx.x / gsz, l = (int)threadIdx.x - g * gsz;
        if (g < ngrp) {
            for (int s = g; s < nseg; s += ngrp) {
                const int off = W.seg_off[s], beg = W.seg_beg[s], len = W.seg_off[s + 1] - off;
                for (int q = l; q < len; q += gsz) {
                    const int t = off + q, k = beg + q;
                    const double4 p = spos[k];
                    st.xyz[3 * t + 0] = p.x;
                    st.xyz[3 * t + 1] = p.y;
                    st.xyz[3 * t + 2] = p.z;
                    if (parts) st.j[t] = parts[k];
                }
            }
        }
        if ((int)threadIdx.x < pad) {
            const int t = S + (int)threadIdx.x;
            st.xyz[3 * t + 0] = kFarPos;
            st.xyz[3 * t + 1] = kFarPos;
            st.xyz[3 * t + 2] = kFarPos;
            if (parts) st.j[t] = -1;
        }
    }
    __syncthreads();
    return staged;
}

}  // namespace tmd
