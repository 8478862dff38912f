// Tile windows, the cell-ordered position mirror, and the Verlet neighbour
// lists (full and half) as 16-bit offsets into each tile's window
// (tmd_window.hpp).
//
// List build: one block per tile, one lane per local atom of the tile.  The
// block stages its window (atom index + FP64 position) into LDS; each lane
// walks its 27-cell stencil as 9 rows of 3 cells inside the window, in the
// reference's order -- stencil (z, y, x), then cell order
// (core/neighborlist.impala:1-78) -- and keeps a candidate when
// `rsq <= r_nb^2` in FP64 without contraction, so the sets are bit-exact with
// the reference.  Half lists keep j < i among locals (the reference's
// filter) plus, by default, every ghost.
#include "tmd_internal.hpp"
#include "tmd_pair.hpp"

namespace tmd {

// Cell-ordered position mirror: spos[k] = pos[parts[k]] (locals and ghosts);
// with svel, also the locals' velocities in mirror order.
__global__ void k_mirror(const double4* __restrict__ pos, const int* __restrict__ parts, const int* __restrict__ nbinned,
                         const Mir spos, int n, const double4* __restrict__ vel, const Mir svel,
                         int vel_ghosts) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *nbinned) return;
    const int j = parts[k];
    spos[k] = pos[j];
    if (svel.p && (j < n || vel_ghosts)) svel[k] = vel[j];
}

// Per-atom inputs come in work order (mirror index work_k, cell work_c), so
// a wavefront's reads of them are contiguous.
template <bool HALF>
__global__ void __launch_bounds__(kMaxTileThreads) k_nlist(const CMir spos, const int* __restrict__ work_k,
                                                        const int* __restrict__ work_c,
                                                        const int* __restrict__ start, const int* __restrict__ parts,
                                                        const int* __restrict__ work, const int* __restrict__ tile_start,
                                                        NbArgs a, int* __restrict__ nb_count, nb_t* __restrict__ nb_list,
                                                        int* __restrict__ nb_count2, nb_t* __restrict__ nb_list2,
                                                        int* __restrict__ overflow) {
    extern __shared__ double dyn[];
    // half lists stage the atom index too (the i < j test), full lists positions only
    constexpr int SB = HALF ? kStageBytes : kPosStageBytes;
    int* cst = reinterpret_cast<int*>(reinterpret_cast<char*>(dyn) + (size_t)SB * a.wcap);
    const int tile = blockIdx.x;
    const int w0 = tile_start[tile], nloc = tile_start[tile + 1] - w0;
    if (nloc == 0) return;
    __shared__ TileWin W;
    tile_window(tile, start, a.g, W);
    stage_row_starts(W, start, a.g, cst);
    const WinStage sp = win_stage(dyn, a.wcap);
    const bool staged = stage_window(W, a.wcap, spos, HALF ? parts : nullptr, sp, kListPad);
    for (int at = threadIdx.x; at < nloc; at += blockDim.x) {
        const int w = w0 + at;
        const int i = work[w];
        const int k = work_k[w];
        int cnt2 = 0;
        const int cnt = staged ? nlist_atom<HALF, true>(i, w, work_c[w], k, spos[k], cst, a, W, sp, spos, parts, nb_list,
                                                        nb_list2, cnt2)
                               : nlist_atom<HALF, false>(i, w, work_c[w], k, spos[k], cst, a, W, sp, spos, parts, nb_list,
                                                         nb_list2, cnt2);
        nb_count[w] = cnt;
        if (HALF) nb_count2[w] = cnt2;
        const int most = max(cnt, cnt2);
        if (most > a.cap) atomicMax(overflow, most);
    }
}

// ---------------------------------------------------------------------------
// Threads per tile block: one lane per local of the fullest tile (no tile
// takes a second pass), in whole wavefronts, at most kAutoTileThreads so that
// two blocks of the tile kernels (up to 80 VGPRs: 6 waves per SIMD) share a
// CU.  An explicit choice (tmd_set_tile_threads) overrides.
int tile_block_threads(int requested, int max_locals) {
    if (requested > 0) return requested;
    const int t = (std::max(max_locals, 1) + kWave - 1) / kWave * kWave;
    return std::min(std::max(t, kTileThreads / 2), kAutoTileThreads);
}

int window_cap(const Sim& s) { return std::min(s.max_window + 1, kLdsBudget / kStageBytes); }  // + sentinel slot
int window_cap_pos(const Sim& s) { return std::min(s.max_window + 1, kLdsBudget / kPosStageBytes); }

// The window statistics (locals placed, binned atoms, largest window, most
// locals in a tile) decide the LDS size and block size of the tile kernels.
static void apply_window_stats(Sim& s, const int* host) {
    if (host[0] != s.n) {
        char msg[200];
        snprintf(msg, sizeof msg,
                 "%d local atoms lie outside the periodic box or the cell grid (non-finite or unwrapped positions; "
                 "md_enforce_pbc wraps them)",
                 s.n - host[0]);
        set_error(__FILE__, __LINE__, msg);
        return;
    }
    s.nbinned = host[1];
    s.max_window = host[2];
    s.tile_threads = tile_block_threads(s.tile_threads_req, host[3]);
    retune_tile_edge(s);
    s.spos.reserve((size_t)kMirBytes * (size_t)std::max(s.nbinned, 1), false, s.stream);
    s.spos2.reserve((size_t)kMirBytes * (size_t)std::max(s.nbinned, 1), false, s.stream);
    if (has_error()) return;
    if (s.max_window > kWinMax) {
        char msg[200];
        snprintf(msg, sizeof msg, "a tile neighbourhood holds %d atoms, beyond the 16-bit window range (%d)", s.max_window,
                 kWinMax);
        set_error(__FILE__, __LINE__, msg);
        return;
    }
    if (kLdsBudget > s.lds_limit) {
        set_error(__FILE__, __LINE__, "device LDS per block is below the window budget");
        return;
    }
    s.windows_valid = 1;
}

// A rebuild inside the fused step loop (s.defer_window_sync, with statistics
// from an earlier binning at hand) only queues the copy of the statistics
// into pinned host memory: its list build runs with the previous window and
// block sizes (a larger window is then read from the mirror, a fuller tile
// takes a second pass of its block) and reads the new statistics after its
// own synchronisation (finish_windows), saving one host round trip.
void build_windows(Sim& s) {
    if (has_error()) return;
    int* counters = s.counters.as<int>();
    const Geom& G = s.g;
    const int ntiles = s.tiles.ntx * s.tiles.nty * s.tiles.ntz;
    const bool defer = s.defer_window_sync && s.max_window > 0;
    int local[4] = {0, 0, 0, 0};
    int* host = local;
    if (defer) {
        if (!s.h_winstats) TMD_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.h_winstats), 4 * sizeof(int)));
        if (has_error()) return;
        host = s.h_winstats;
    }
    // counters[6] = largest tile window, counters[7] = most locals in a tile (from k_tile_stats)
    TMD_HIP(hipMemcpyAsync(&host[0], s.tile_start.as<int>() + ntiles, sizeof(int), hipMemcpyDeviceToHost, s.stream));
    TMD_HIP(hipMemcpyAsync(&host[1], s.cell_start.as<int>() + G.ncells, sizeof(int), hipMemcpyDeviceToHost, s.stream));
    TMD_HIP(hipMemcpyAsync(&host[2], counters + 6, 2 * sizeof(int), hipMemcpyDeviceToHost, s.stream));
    if (defer) {
        const int bound = std::max(s.n + s.ng, 1);  // binned atoms, at most
        s.spos.reserve((size_t)kMirBytes * (size_t)bound, false, s.stream);
        s.spos2.reserve((size_t)kMirBytes * (size_t)bound, false, s.stream);
        if (has_error()) return;
        s.windows_pending = 1;
        s.windows_valid = 1;
        return;
    }
    TMD_HIP(hipStreamSynchronize(s.stream));
    if (has_error()) return;
    apply_window_stats(s, host);
}

void finish_windows(Sim& s) {
    if (!s.windows_pending) return;
    s.windows_pending = 0;
    TMD_HIP(hipStreamSynchronize(s.stream));  // the copies were queued before
    if (has_error()) return;
    apply_window_stats(s, s.h_winstats);
}

// DEM step loop: positions and velocities of every binned atom, ghosts
// included, into the mirrors (spos, svel); the force pass reads them there.
void dem_mirror(Sim& s) {
    const int bound = s.n + s.ng;
    if (has_error() || bound == 0) return;
    s.svel.reserve((size_t)kMirBytes * (size_t)bound, false, s.stream);
    s.svel2.reserve((size_t)kMirBytes * (size_t)bound, false, s.stream);
    s.spos2.reserve((size_t)kMirBytes * (size_t)bound, false, s.stream);
    if (has_error()) return;
    const int nt = 256;
    hipLaunchKernelGGL(k_mirror, dim3(div_up(bound, nt)), dim3(nt), 0, s.stream, (const double4*)s.pos.as<double4>(),
                       (const int*)s.cell_parts.as<int>(), (const int*)(s.cell_start.as<int>() + s.g.ncells),
                       Mir{s.spos.as<double>()}, s.n, (const double4*)s.vel.as<double4>(), Mir{s.svel.as<double>()}, 1);
    TMD_HIP(hipGetLastError());
    s.mirror_fresh = 1;
    s.dem_mirrored = 1;
}

void refresh_mirror(Sim& s) {
    // the grid covers every atom that can be binned; the kernel stops at the
    // binned count on the device (the host's may still be pending)
    const int bound = s.n + s.ng;
    if (has_error() || bound == 0) return;
    const int nt = 256;
    // inside the fused step loop the velocities follow the new order too
    const bool with_vel = s.vel_mirror_wanted && !s.vel_mirrored && s.n > 0;
    if (with_vel) {
        s.svel.reserve((size_t)kMirBytes * (size_t)bound, false, s.stream);
        if (has_error()) return;
    }
    hipLaunchKernelGGL(k_mirror, dim3(div_up(bound, nt)), dim3(nt), 0, s.stream, (const double4*)s.pos.as<double4>(),
                       (const int*)s.cell_parts.as<int>(), (const int*)(s.cell_start.as<int>() + s.g.ncells),
                       Mir{s.spos.as<double>()}, s.n, (const double4*)s.vel.as<double4>(),
                       Mir{with_vel ? s.svel.as<double>() : nullptr}, 0);
    TMD_HIP(hipGetLastError());
    s.mirror_fresh = 1;
    if (with_vel) s.vel_mirrored = 1;
}

void build_neighbor_lists(Sim& s, int half, double cutoff) {
    if (has_error()) return;
    TMD_REQUIRE(s.windows_valid, "md_assemble_neighborlists: cells not binned (call md_distribute_particles first)");
    if (has_error()) return;
    int* counters = s.counters.as<int>();
    refresh_mirror(s);
    const int ntiles = s.tiles.ntx * s.tiles.nty * s.tiles.ntz;
    // counters[2] = overflow (largest count above the capacity)
    for (;;) {
        s.nb_cap = (s.nb_cap + 7) & ~7;
        const int ntile64 = div_up(s.n > 0 ? s.n : 1, 64);
        s.nb_stride = ntile64 * 64;
        s.nb_count.reserve(sizeof(int) * (size_t)s.nb_stride, false, s.stream);
        s.nb_list.reserve(sizeof(nb_t) * (size_t)ntile64 * 64 * (size_t)s.nb_cap, false, s.stream);
        if (half) {
            // the complementary column of every half list (same layout)
            s.nb_count2.reserve(sizeof(int) * (size_t)s.nb_stride, false, s.stream);
            s.nb_list2.reserve(sizeof(nb_t) * (size_t)ntile64 * 64 * (size_t)s.nb_cap, false, s.stream);
        }
        if (has_error()) return;
        TMD_HIP(hipMemsetAsync(counters + 2, 0, sizeof(int), s.stream));
        const int rsb = row_starts_bytes(s.tiles);
        // window slots (positions, 24 B; half lists also the atom index,
        // 28 B) + sentinel slots, then the row starts
        const int sb = half ? kStageBytes : kPosStageBytes;
        const int wcap = std::min(s.max_window + kListPad, (kLdsBudget - rsb) / sb);
        NbArgs a{s.n, s.nb_cap, half, s.half_mode, wcap, cutoff * cutoff, s.tiles};
        const size_t lds = (size_t)wcap * sb + rsb;
        if (s.n > 0) {
#define TMD_NLIST(H)                                                                                                   \
    hipLaunchKernelGGL(k_nlist<H>, dim3(ntiles), dim3(s.tile_threads), lds, s.stream, CMir{s.spos.as<double>()},           \
                       (const int*)s.work_k.as<int>(), (const int*)s.work_c.as<int>(),                                      \
                       (const int*)s.cell_start.as<int>(), (const int*)s.cell_parts.as<int>(), (const int*)s.work.as<int>(), \
                       (const int*)s.tile_start.as<int>(), a, s.nb_count.as<int>(), s.nb_list.as<nb_t>(), \
                       s.nb_count2.as<int>(), s.nb_list2.as<nb_t>(), counters + 2)
            if (half) TMD_NLIST(true); else TMD_NLIST(false);
#undef TMD_NLIST
        }
        TMD_HIP(hipGetLastError());
        int over = 0;
        TMD_HIP(hipMemcpyAsync(&over, counters + 2, sizeof(int), hipMemcpyDeviceToHost, s.stream));
        TMD_HIP(hipStreamSynchronize(s.stream));
        finish_windows(s);  // statistics of a deferred window build
        if (has_error()) return;
        if (over > 0) {
            s.nb_cap = over * 2;
            continue;
        }
        break;
    }
    s.nb_half = half;
    s.nb_valid = 1;
}

}  // namespace tmd
