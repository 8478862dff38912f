/* Extensions of libtinymd_amd.so beyond the reference ABI (md_abi.h).
 * None of these has a counterpart in the reference (anydsl_includes.h);
 * each comment names the reference behaviour it extends. */
#ifndef TINYMD_AMD_TMD_EXT_H
#define TINYMD_AMD_TMD_EXT_H

#include "md_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- errors ------------------------------------------------------------- */
/* First recorded error (HIP failure, invalid geometry, escaped atoms,
 * 16-bit window overflow), or NULL.  The reference has no error channel. */
const char* tmd_last_error(void);
void tmd_clear_error(void);
/* Device name and architecture (the architecture alone when the runtime
 * reports no name) into buf; returns the number of compute units (0 = no GPU). */
int tmd_device_info(char* buf, int len);
/* Host evaluation (no GPU needed) of the compensated FP32 sum that the
 * FP32-Kahan pair kernels apply to their per-atom force terms: the same
 * inline accumulation and finalisation, returned in double. */
double tmd_kahan_sum(const float* x, int n);

/* ---- modes ---------------------------------------------------------------- */
/* 0 = FP64 (default; the reference's real_t, utils/utilities.impala:1),
 * 1 = FP32-mixed: FP32 pair forces accumulated per atom in FP64 (FP64
 *     positions, distance test, velocities, wrap, ghosts and reductions),
 * 2 = as 1 but accumulated in FP32 with Kahan compensation (full lists). */
void tmd_set_precision(int mode);
/* Precision of the adjoint kernel alone (tmd_traj_adjoint's reverse pass):
 * -1 (default) follows tmd_set_precision, 0 = FP64, 1 = FP32-mixed (its
 * Hessian-vector product's pair scalars in FP32).  Setting it apart from the
 * forward precision separates the adjoint's own rounding from that of the
 * FP32-mixed forward trajectory. */
void tmd_set_adjoint_precision(int mode);
/* 1: renumber the local atoms in cell order (index order inside a cell) at
 * every periodic wrap of a rebuild (md_enforce_pbc, and inside tmd_run_*),
 * before the ghosts are built; positions, velocities, forces and masses move
 * with them, and the arrays written by md_write_grid_data_to_arrays follow
 * the new numbering.  This is the reference's sort_particles() hook
 * (core/grid.impala:54, 323-356; compiled out there), taken before borders
 * so that the ghost layer refers to the new numbering.  0 (default): atoms
 * keep their indices, as in the reference's build. */
void tmd_set_sort_particles(int on);
/* Half-list filter.  0 = local j < i plus every ghost (each periodic pair
 * once; default), 1 = the reference's literal i > j filter, which drops all
 * ghost pairs (core/neighborlist.impala:34). */
void tmd_set_half_mode(int mode);
/* Ghost refresh.  0 = fresh (owner + image * L; default), 1 = the
 * reference's two-phase refresh whose ghost-of-ghost values lag a step
 * (comm/comm.impala:341-387). */
void tmd_set_ghost_mode(int mode);
/* 1 (default): every md_* call returns after the GPU finished it.
 * 0: calls only enqueue; md_barrier / copies synchronise. */
void tmd_set_sync(int on);
/* DEM sphere radius; <= 0 restores the default (contact cutoff / 2).
 * The reference hard-codes 1.0 (core/runtime.impala:210). */
void tmd_set_dem_radius(double radius);
/* 1 = shifted-force LJ, F(r) - F(rc), with energy shifted to match
 * (continuous force at rc; used for the differentiable runs). */
void tmd_set_lj_shift(int on);
/* Cells per tile of the list/force kernels (one block per tile; 0 0 0 =
 * automatic, about 480 atoms per tile).  Takes effect at the next binning. */
void tmd_set_tile(int tx, int ty, int tz);
/* Threads per tile block of the list-build and pair kernels: a multiple of
 * 64 in 64..1024, or 0 (default) for automatic: one lane per local atom of
 * the fullest tile, 256..768, set at each binning. */
void tmd_set_tile_threads(int threads);

/* ---- thermo (deterministic device reductions) ------------------------------
 * out = {T, KE, PE, W, P, pairs}: T = sum v^2 / (3N-3) (core/thermo.impala:1-16),
 * PE and the pair virial W = sum r.F over the current list (the reference
 * computes neither), P = (sum v^2 + W) / (3V), pairs = in-cutoff entries. */
void tmd_thermo(double rc, double epsilon, double sigma, double* out6);
/* As tmd_thermo, with PE taken over the energy-shifted potential U(r) - U(rc)
 * for every in-cutoff pair (continuous at the cutoff; the plain-cutoff force
 * is unchanged), for energy-drift bookkeeping.  Runs with the smooth cutoff
 * (tmd_set_lj_shift) report their own shifted energy either way. */
void tmd_thermo_eshift(double rc, double epsilon, double sigma, double* out6);
/* As tmd_thermo, without allocating: no hipMalloc, hipFree or allocating
 * library call.  No counterpart in the reference; it is for callers that
 * cannot allocate (code under stream capture, tight serving loops,
 * harnesses that forbid allocation).  The reductions' partial sums live in a
 * fixed 1 MiB device array that each GPU holds until the process ends, which
 * limits the call to 3 * tiles + 8 <= 131072 and ceil(nlocal / 256) + 8 <=
 * 131072 (about 43,000 tiles; an error is set beyond that).  The state must
 * already be on the device and its buffers sized (md_copy_data_to_accelerator
 * and a neighbour-list build); a call that would have to grow a buffer sets
 * an error instead.  Calls that overlap on one GPU race on the array.
 * Results are those of tmd_thermo, bit for bit. */
void tmd_thermo_noalloc(double rc, double epsilon, double sigma, double* out6);

/* ---- state I/O -------------------------------------------------------------
 * VTK: the reference's legacy particle file (vtk.h:6-91, written by main.cpp's
 * vtk_write_*_data, main.cpp:67-98, from md_write_grid_*_data_to_arrays) for
 * what = 0 (locals), 1 (ghosts) or 2 (the grid AABB's 8 corners), from the
 * host copy (after md_copy_data_from_accelerator, as main.cpp does).
 * Returns the number of points.
 * Restart (no counterpart in the reference): tmd_save_state writes the
 * locals' mass, position, velocity and force (FP64), the geometry and the
 * modes to a binary file with a checksum; tmd_load_state reads one into a
 * fresh simulation (md_initialize_grid + md_copy_data_to_accelerator, with
 * the saved force).  Ghosts, cells and lists are rebuilt by the next
 * set-up, as after md_initialize_grid.  Both return the number of atoms
 * (0 and an error on failure). */
int tmd_write_vtk(const char* path, int what);
int tmd_save_state(const char* path);
int tmd_load_state(const char* path);

/* ---- introspection for parity tests (host copies) --------------------------- */
/* {nx, ny, nz, ncells, nlocal, nghost, list capacity} */
void tmd_get_grid_info(int* out7);
/* Tile decomposition of the last binning: {cells per tile in x, y, z,
 * tiles, threads per tile block, largest window (atoms), LDS bytes per
 * window slot that automatic tiles are sized for}. */
void tmd_get_tile_info(int* out7);
/* Ghost g (slot nlocal + g): flattened owner, total image (3 ints),
 * immediate source and per-direction flag (3 ints).  Returns nghost. */
int tmd_get_ghost_table(int* owner, int* image3, int* source, int* flag3);
/* Cell CSR: start[ncells + 1], atoms[start[ncells]], cell of each atom
 * [nlocal + nghost] (-1 outside).  Returns the number of binned atoms. */
int tmd_get_cells(int* start, int* atoms, int* atom_cell);
/* Neighbour lists by atom index: counts[nlocal], list[nlocal * stride]
 * (row-major, first min(count, stride) entries).  Returns nlocal. */
int tmd_get_neighbor_lists(int* counts, int* list, int stride);

/* ---- device-resident step loops -------------------------------------------
 * main.cpp's timestep loop (main.cpp:715-776) run inside the library with
 * host synchronisation only at rebuilds.  Return the number of steps run
 * (0 on error). */
int tmd_run_lj(int steps, int reneigh_every, double dt, bool half, double cutoff, double skin, double epsilon, double sigma);
int tmd_run_dem(int steps, int reneigh_every, double dt, bool half, double cutoff, double skin, double damping_n,
                double damping_t, double stiffness, double friction);

/* ---- profiling: HIP events around each launch on the library stream -------- */
/* kernel ids: 0 force, 1 initial integration, 2 final integration,
 * 3 ghost refresh, 4 neighbour list, 5 binning, 6 borders, 7 thermo, 8 adjoint,
 * 9 cell-ordered position mirror */
void tmd_profile_enable(int on);
void tmd_profile_get(int kernel, double* total_ms, long long* count);
void tmd_profile_reset(void);

/* ---- initial conditions (host) ----------------------------------------------
 * FCC lattice exactly as init_rectangular_grid builds it (initialize.h:91-224):
 * returns the number of atoms (arrays filled up to cap). */
int tmd_generate_fcc(int nx, int ny, int nz, double potential_minimum, int half, double* mass, struct Vector3D* pos,
                     struct Vector3D* vel, int cap);
/* Granular gas: FCC sites of an nc^3-cell lattice kept where the lattice
 * indices satisfy i + j < 2 nc (the diagonal half x + y < L), velocities
 * uniform in [-vmax, vmax] from a counter-based hash of (seed, site,
 * component).  Replaces initialize.h:226-295 for the DEM benchmark. */
int tmd_generate_granular_gas(int nc, double potential_minimum, unsigned long long seed, double vmax, double* mass,
                              struct Vector3D* pos, struct Vector3D* vel, int cap);

/* --- differentiable trajectories (no counterpart in the reference) ---------
 * Loss of an LJ velocity-Verlet trajectory of `steps` steps from the current
 * state (x0, v0), reneighbouring every `reneigh_every` steps with full lists
 * as tmd_run_lj does:
 *   Ubar = (1/T) sum_{t=1..T} PE_t / N,
 *   Pbar = (1/T) sum_{t=1..T} (sum_i v_i^2 + W_t) / (3 V),  W = sum_pairs r.F,
 *   L    = (Ubar - u_target)^2 + (Pbar - p_target)^2   (unit masses).
 * The current precision and LJ shift modes apply; with the smooth cutoff
 * (tmd_set_lj_shift(1)) the trajectory is differentiable and the adjoint is
 * its exact gradient.  Both calls leave the state, ghosts, list and forces
 * as they were at x0 (full list).  Return 1 on success, 0 on error. */
int tmd_traj_loss(int steps, int reneigh_every, double dt, double cutoff, double skin, double epsilon, double sigma,
                  double* out2 /* {Ubar, Pbar} */);
/* Reverse-mode gradient of L: state checkpoints every `reneigh_every` steps,
 * each interval recomputed forward and reversed with the hand-written
 * Hessian-vector-product kernel.  out3 = {L, Ubar, Pbar}; grad_theta2 =
 * {dL/depsilon, dL/dsigma}; grad_x0 / grad_v0 (n entries, may be NULL) =
 * dL/dx0, dL/dv0. */
int tmd_traj_adjoint(int steps, int reneigh_every, double dt, double cutoff, double skin, double epsilon, double sigma,
                     double u_target, double p_target, double* out3, double* grad_theta2, struct Vector3D* grad_x0,
                     struct Vector3D* grad_v0);

/* --- batched replicas (no counterpart in the reference) ---------------------
 * Many independent periodic LJ boxes of different sizes advanced together:
 * every kernel launch covers all replicas, and a whole run -- the main.cpp
 * step loop with reneighbouring every `reneigh_every` steps -- is one
 * hipGraph launch.  Full neighbour lists, unit masses, fresh ghosts.  Each
 * replica's trajectory is bit-identical to the same box run alone through
 * the md_* path with the same cell spacing, for any tile and block size.
 * Handles are small positive integers; 0 means failure (tmd_last_error). */

/* Replica parameters from a counter-based generator (SplitMix64 over
 * (seed, replica, draw)): cells per side uniform in [cells_min, cells_max],
 * density and temperature uniform in their ranges.  Host only. */
int tmd_generate_replica_set(unsigned long long seed, int nrep, int cells_min, int cells_max, double rho_min,
                             double rho_max, double temp_min, double temp_max, int* cells, double* rho, double* temp);
/* counts[r] atoms per replica; pos / vel concatenated in replica order;
 * aabbs: 6 per replica (xmin xmax ymin ymax zmin zmax). */
int tmd_batch_create(int nrep, const int* counts, const struct Vector3D* pos, const struct Vector3D* vel,
                     const double* aabbs);
/* Parameters and the initial rebuild + force at the stored initial state.
 * precision 0 FP64, 1 FP32-mixed, 2 FP32 Kahan. */
int tmd_batch_setup(int handle, double cutoff, double skin, double epsilon, double sigma, double dt, int reneigh_every,
                    int precision, int lj_shift);
/* Launch configuration (restarts from the initial state): cells per tile
 * edge (0 automatic per replica), threads per tile block (0: automatic, one
 * lane per local of the fullest tile; else 64..1024, multiple of 64), cell
 * edge / (cutoff + skin) (>= 1), list capacity headroom over the largest
 * count, ghost room headroom over the expected ghost count. */
int tmd_batch_configure(int handle, int tile, int threads, double spacing_factor, double list_headroom,
                        double ghost_headroom);
/* Times trial runs of candidate configurations from the initial state, in
 * two stages: cells per tile edge (automatic, 2, 3, 4) x threads per block
 * (automatic, 256, 512) at cell edge 1.0 and list headroom 1.1, then cell
 * edge (1.0, 1.1) x list headroom (1.1, 1.3) at the best of those; keeps
 * the fastest and restarts from the initial state with it.  out (may be
 * NULL): tile, threads in use, spacing factor, list headroom, ms/step. */
int tmd_batch_autotune(int handle, int trial_steps, double* out5);
/* Restart from the initial state with the current configuration. */
int tmd_batch_reset(int handle);
/* 1 (default): a run is captured once into a hipGraph and launched as one
 * graph; 0: the same kernels are launched one by one on the batch stream
 * (for tools that trace individual launches).  Results are identical. */
int tmd_batch_set_graph(int handle, int on);
/* Advance all replicas `steps` steps (one graph launch); returns steps.
 * The timed variant also returns the GPU time of the launch. */
int tmd_batch_run(int handle, int steps);
int tmd_batch_run_timed(int handle, int steps, double* gpu_ms);
/* Locals of replica `rep` (any pointer may be NULL); returns their number. */
int tmd_batch_get_state(int handle, int rep, struct Vector3D* pos, struct Vector3D* vel, struct Vector3D* force);
/* Per replica 6 values as tmd_thermo: T, KE, PE, virial, P, pairs. */
int tmd_batch_thermo(int handle, double* out);
/* Per replica 8 ints: n, ghosts, cells per x, cells, tiles, tile edge, ghost
 * room, list capacity. */
int tmd_batch_info(int handle, int* out);
void tmd_batch_destroy(int handle);

#ifdef __cplusplus
}
#endif

#endif /* TINYMD_AMD_TMD_EXT_H */
