// Internal declarations shared by the HIP translation units of libtinymd_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "tmd_window.hpp"

namespace tmd {

// ---------------------------------------------------------------------------
// error handling: every HIP call is checked; the first failure is recorded,
// printed, and reported by tmd_last_error(); later work is skipped.
void set_error(const char* file, int line, const char* what);
bool has_error();
// While an AllocFree guard is alive, a DevBuf that would have to grow sets
// an error instead of calling hipMalloc (the allocation-free entry points).
bool alloc_forbidden();
struct AllocFree {
    AllocFree();
    ~AllocFree();
};

#define TMD_HIP(call)                                                          \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            ::tmd::set_error(__FILE__, __LINE__, hipGetErrorString(e_));       \
        }                                                                      \
    } while (0)

#define TMD_REQUIRE(cond, msg)                                                 \
    do {                                                                       \
        if (!(cond)) {                                                         \
            ::tmd::set_error(__FILE__, __LINE__, msg);                         \
        }                                                                      \
    } while (0)

// ---------------------------------------------------------------------------
// modes
enum Precision { PREC_FP64 = 0, PREC_FP32_MIXED = 1, PREC_FP32_KAHAN = 2 };
enum HalfMode { HALF_GHOST_FULL = 0, HALF_REF_LITERAL = 1 };
enum GhostMode { GHOST_FRESH = 0, GHOST_REF_LAG = 1 };

// Geometry of one periodic box (single rank: rank AABB == world AABB).
struct Geom {
    double world[6];  // xmin xmax ymin ymax zmin zmax
    double aabb[6];   // rank AABB extended by one cell spacing
    double len[3];    // periodic lengths
    double spacing;   // cell spacing
    double rank[6];   // rank AABB as given (aabb is it extended by one spacing)
    int nx, ny, nz, ncells;
};

// A device buffer that grows (optionally preserving its contents).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void reserve(size_t need, bool preserve, hipStream_t s);
    void release();
    template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

// Per-kernel timing with HIP events on the library stream.
enum KernelId {
    K_FORCE = 0, K_INTEGRATE_1, K_INTEGRATE_2, K_GHOST_SYNC, K_NEIGHBOR, K_BIN, K_BORDERS,
    K_THERMO, K_ADJOINT, K_MIRROR, K_NKERNELS
};
struct KernelTimer {
    bool enabled = false;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
    double total_ms[K_NKERNELS] = {0};
    long long count[K_NKERNELS] = {0};
    size_t next = 0;
    void begin(int id, hipStream_t s, hipEvent_t* ev0);
    void end(int id, hipStream_t s, hipEvent_t ev0);
    void collect();
    void reset();
};

struct Sim {
    Geom g;
    int n = 0;          // local atoms
    int ng = 0;         // ghosts
    int cap_all = 0;    // capacity of per-atom arrays (locals + ghosts)
    int nb_cap = 100;   // neighbour-list capacity per atom
    int nb_stride = 0;  // column stride of the list (multiple of 64 >= n)
    int nb_half = 0;    // list built as a half list
    int nb_valid = 0;   // list present
    int nbinned = 0;    // atoms in the cell CSR (locals + ghosts inside the grid)
    int max_window = 0; // largest tile window (atoms)
    TileGeom tiles{};   // tile decomposition of the cell grid
    int tile_req[3] = {0, 0, 0};  // requested cells per tile (0: automatic)
    int tile_auto = 0;            // automatic cells per tile edge in effect (0: chosen at the next binning)
    int stage_bytes = kPosStageBytes;  // LDS bytes per window slot of the force pass in use (LJ: positions)
    int stage_budget = kLdsBudget;  // LDS bytes per block that pass should stay within
    int tile_threads = kTileThreads;  // threads per tile block in use (multiple of 64, <= kMaxTileThreads)
    int tile_threads_req = 0;         // requested block size, 0: automatic (tile_block_threads)
    int windows_valid = 0;
    int inv_valid = 0;      // cell_inv matches the current atom set (binning after the last ghost build)
    int mirror_fresh = 0;   // spos[k] == pos[cell_parts[k]] for every binned atom
    int ghost_stale = 0;    // the last force pass advanced the locals in the mirror; the ghosts (pos and
                            // mirror) still hold the previous step
    int pos_stale = 0;      // local positions in pos lag the mirror (fused step loop; sync_pos_from_mirror)
    int vel_mirrored = 0;   // local velocities live in svel, in mirror order (fused step loop)
    int vel_mirror_wanted = 0;  // fused step loop: refresh_mirror also moves the velocities to svel
    int gsrc_valid = 0;     // g_ko / g_kg match the ghost table and the binning
    int gh_valid = 0;       // gh_start / gh_k / gh_c match the ghost table, binning and work order
    int dem_ghosts_done = 0;  // DEM step loop: the last force pass also wrote the ghosts' mirror entries
    int defer_window_sync = 0;  // rebuild in progress: build_windows may leave its statistics pending
    int windows_pending = 0;    // window statistics queued into h_winstats, not yet applied (finish_windows)
    int* h_winstats = nullptr;  // pinned host copy of the window statistics (4 ints)
    int ghost_pos_stale = 0;  // ghosts in pos lag the mirror (fused step loop; sync_pos_from_mirror)
    int lds_limit = 65536;  // LDS bytes per block (device attribute)

    int precision = PREC_FP64;
    int adj_precision = -1;  // adjoint kernel precision (-1: the forward's)
    int half_mode = HALF_GHOST_FULL;
    int sort_particles = 0;  // renumber the locals in cell order at each wrap (sort_locals)
    int ghost_mode = GHOST_FRESH;
    int sync_calls = 1;
    double dem_radius = 0.0;  // <= 0: contact cutoff / 2
    int lj_shift = 0;

    // host mirrors (filled by md_copy_data_from_accelerator)
    std::vector<double> h_mass;
    std::vector<double> h_pos, h_vel, h_force;  // AoS 3 doubles per atom (locals + ghosts)
    bool host_valid = true;                     // host mirror is the newest copy
    bool dev_valid = false;

    // device state
    DevBuf pos, vel, force, mass;      // double4 x3, double
    DevBuf sort_buf;                   // sort_locals: scratch copies of the four arrays above
    DevBuf g_owner, g_src, g_code;     // ghost table
    DevBuf cell_of, cell_slot, cell_count, cell_start, cell_parts;
    DevBuf cell_inv;                   // CSR index of each atom (-1: outside the grid)
    DevBuf work_start, work;           // per-tile local counts; locals in tile work order (work slots)
    DevBuf tile_start;                 // first work slot of each tile [ntiles + 1]
    DevBuf spos;                       // cell-ordered position mirror (x, y, z packed, 24 B per atom)
    DevBuf spos2;                      // second mirror: the force pass that advances positions writes the
                                       // new local positions here, then the two are swapped
    DevBuf work_k;                     // mirror (CSR) index of each work slot's atom
    DevBuf work_c;                     // cell of each work slot's atom
    DevBuf svel;                       // local velocities in mirror order during the fused step loop
    DevBuf svel2;                      // DEM step loop: the velocity mirror the force pass writes
    int dem_mirrored = 0;              // DEM step loop: positions and velocities (ghosts too) live in the mirrors
    DevBuf face;                       // per atom: border selection bits of the three ghost passes (u8)
    DevBuf g_ko, g_kg;                 // per ghost: mirror index of its owner and its own (-1: outside the grid)
    DevBuf gh_start, gh_k, gh_c;       // ghosts by owner work slot: CSR start [n + 1], mirror index, image code
    DevBuf gh_aux;                     // per work slot: owner's work slot scratch, counts and fill cursors
    DevBuf nb_count, nb_list;
    DevBuf nb_count2, nb_list2;        // half lists: the complementary columns (local j > i)
    DevBuf scan_blocks, scan_aux, counters, red, gbuf;
    // trajectory loss and adjoint: initial state, checkpoints, interval
    // states, adjoints (double4 per local atom), partial sums, accumulators
    DevBuf adj_x0, adj_v0, adj_ckx, adj_ckv, adj_segx, adj_segv;
    DevBuf adj_xbar, adj_vbar, adj_fbar, adj_part, adj_acc;

    hipStream_t stream = nullptr;
    KernelTimer timer;

    // last thermo (device reductions)
    double pe = 0.0, virial = 0.0;
    long long pairs = 0;
    // iteration counters reported by md_report_iterations
    long long force_iters = 0, rsq_iters = 0, integration_iters = 0;
    double last_force_rc = 0.0;  // cutoff of the last pair-force evaluation (0: none yet)
};

// Positions changed outside the fused step loop: the mirror is rebuilt from
// pos before the next pair pass.
inline void mirror_invalidate(Sim& s) {
    s.mirror_fresh = 0;
    s.ghost_stale = 0;
}
Sim& sim();
bool sim_alive();

inline int div_up(long long a, long long b) { return (int)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// launchers (kernels live in the .hip translation units)
// scan / compaction
void exclusive_scan_int(const int* in, int* out, int n, int* total_dev, Sim& s);
void scan_tiles(Sim& s, int* tiles, int ntiles, int* total_dev);  // one-block exclusive scan in place
// exclusive scan of n > 0 ints with caller scratch of div_up(n, kScanTile)
// ints; no allocation, no host synchronisation (graph-capturable)
void scan_exclusive(const int* in, int* out, int n, int* total_dev, int* tiles, hipStream_t st);

// borders / ghosts
void build_ghosts(Sim& s, bool wrap = false);  // wrap: md_enforce_pbc on the locals first (fused)
void refresh_ghosts(Sim& s);
// fresh ghost positions and the cell-ordered mirror of locals and ghosts in
// one pass (after positions were advanced inside the force pass)
void step_refresh(Sim& s);
void build_ghost_src(Sim& s);  // g_ko / g_kg (binning, after the CSR is built)
void build_ghost_index(Sim& s);  // each local's ghosts by work slot (gh_start / gh_k / gh_c)
void sync_pos_from_mirror(Sim& s);
void mirror_velocities(Sim& s, bool into_mirror);

// geometry (core/grid.impala:94-98): world AABB, rank AABB, cell spacing
void make_geom(Geom& G, const double* aabb, const double* rank_aabb, double cell_spacing);

// binning
void bin_particles(Sim& s);
void sort_locals(Sim& s);  // PBC wrap + renumber the locals in cell order
int auto_tile_edge(const Geom& G, int n);  // cells per tile edge for ~480 locals per tile
void retune_tile_edge(Sim& s);  // automatic tiles: a smaller edge at the next binning when windows overflow
TileGeom make_tile_geom(const Geom& G, int tx, int ty, int tz);

// tile windows (after binning), position mirror, neighbour lists
void build_windows(Sim& s);
int window_cap(const Sim& s);      // window slots of kernels that stage positions and atom indices
int window_cap_pos(const Sim& s);  // window slots of kernels that stage positions only
int tile_block_threads(int requested, int max_locals);  // staged window atoms per block (LDS budget, kStageBytes each)
void refresh_mirror(Sim& s);
void finish_windows(Sim& s);  // apply window statistics left pending by build_windows
void build_neighbor_lists(Sim& s, int half, double cutoff);

// forces
// kick_dt > 0: the final velocity-Verlet kick v += F dt/2 is fused into the
// force pass; with advance also the next step's kick and drift (positions
// move; the mirror must be refreshed before the next force evaluation)
void compute_lj(Sim& s, int half, double rc, double eps, double sigma, double kick_dt = 0.0, bool advance = false);
// kick_dt > 0: the final kick of the step fused; advance: also the next
// step's first kick and drift
// DEM step loop with the state in the mirrors: mirror set-up, per-step ghost
// pass, the force pass with the step's final kick and the next kick and
// drift, and the way back to atom order
void dem_mirror(Sim& s);
void dem_ghost_step(Sim& s);
void compute_dem_fused(Sim& s, int half, double rc, double gn, double gt, double k, double mu, double dt, bool advance);
void dem_unmirror(Sim& s);
void compute_dem(Sim& s, int half, double rc, double gn, double gt, double k, double mu, double kick_dt = 0.0,
                 bool advance = false);
// eshift: subtract u(rc) from each in-cutoff pair's energy (plain-cutoff runs)
// buf: 3 * tiles + 8 doubles of device memory for the partial sums (nullptr:
// the simulation's own reduction buffer)
void lj_energy_virial(Sim& s, double rc, double eps, double sigma, double* pe, double* vir, long long* pairs,
                      bool eshift = false, double* buf = nullptr);

void reduce_partials(Sim& s, const double* partials, int nblocks, int width, double* out);
// per-tile (PE, W, pairs) partials of the current full/half list into part[3 * ntiles]
void launch_lj_energy(Sim& s, double rc, double eps, double sigma, double* part, bool eshift = false);

// integration
void initial_integration(Sim& s, double dt);
void final_integration(Sim& s, double dt);
void enforce_pbc(Sim& s);

// Periodic wrap of a local position into the box (md_enforce_pbc): one
// conditional shift per coordinate.
struct PbcArgs {
    double lo[3], hi[3], len[3];
};
PbcArgs pbc_args(const Sim& s);
// image shift (ix, iy, iz) in {-1, 0, 1}^3 as one code 0..26
__device__ __forceinline__ int img_encode(int ix, int iy, int iz) { return (ix + 1) + 3 * (iy + 1) + 9 * (iz + 1); }
__device__ __forceinline__ void img_decode(int c, int& ix, int& iy, int& iz) {
    ix = c % 3 - 1;
    iy = (c / 3) % 3 - 1;
    iz = c / 9 - 1;
}

__host__ __device__ inline void pbc_wrap(double4& p, const PbcArgs& a) {
    double q;
    q = p.x; if (p.x < a.lo[0]) q += a.len[0]; if (p.x > a.hi[0]) q -= a.len[0]; p.x = q;
    q = p.y; if (p.y < a.lo[1]) q += a.len[1]; if (p.y > a.hi[1]) q -= a.len[1]; p.y = q;
    q = p.z; if (p.z < a.lo[2]) q += a.len[2]; if (p.z > a.hi[2]) q -= a.len[2]; p.z = q;
}
// sum m v.v over locals (deterministic); buf: ceil(n / 256) + 8 doubles of
// device memory for the partial sums (nullptr: the simulation's own buffer)
double kinetic_sum(Sim& s, double* buf = nullptr);
int launch_kinetic(Sim& s, double* part);  // per-block partials of sum v.v; returns the block count

// trajectory loss (Ubar, Pbar) and its adjoint (k_adjoint.hip)
bool traj_loss(Sim& s, int steps, int ck, double dt, double rc, double skin, double eps, double sigma, double* ubar,
               double* pbar);
// loss3 = {L, Ubar, Pbar}; grad_theta = {dL/deps, dL/dsigma}; grad_x, grad_v
// (3 n doubles each, may be null) = dL/dx0, dL/dv0
bool traj_adjoint(Sim& s, int steps, int ck, double dt, double rc, double skin, double eps, double sigma, double u_target,
                  double p_target, double* loss3, double* grad_theta, double* grad_x, double* grad_v);

}  // namespace tmd
