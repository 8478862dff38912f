/* tinyMD runtime ABI, served by libtinymd_amd.so on one MI355X.
 *
 * These are the 39 extern "C" entry points of the reference's ABI header
 * (anydsl_includes.h:10-48, defined in core/runtime.impala:7-337 and
 * comm/lb.impala:52-85), with identical C signatures, so the reference's
 * host driver (main.cpp, initialize.h) links against this library
 * unchanged.  Each prototype cites the reference definition it replaces.
 *
 * Semantics kept from the reference:
 *  - one global simulation per process, not re-entrant, single host thread;
 *  - every call is synchronous on return (main.cpp's phase timers rely on
 *    it, main.cpp:713-762) unless tmd_set_sync(0) is called;
 *  - input arrays are copied during the call; output arrays are
 *    caller-allocated and filled from the host mirror, which
 *    md_copy_data_from_accelerator refreshes;
 *  - capacity overflow is handled internally by grow-and-retry.
 * Errors (HIP failures, invalid geometry, non-finite positions) are printed
 * once and reported by tmd_last_error() (tmd_ext.h); later calls are no-ops.
 */
#ifndef TINYMD_AMD_MD_ABI_H
#define TINYMD_AMD_MD_ABI_H

#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

#ifndef ANYDSL_INCLUDES_H
/* anydsl_includes.h:5-9 */
struct Vector3D {
    double x;
    double y;
    double z;
};
#endif

/* core/runtime.impala:7-37 -- build the grid from host arrays (m, x, v, n),
 * world AABB, rank AABB, cell spacing, initial cell and list capacities.
 * Returns the number of local atoms. */
int md_initialize_grid(double const *, struct Vector3D const *, struct Vector3D const *, int, double const *, double const *,
                       double, int, int);
/* core/runtime.impala:39-50 */
void md_rescale_grid(double const *);
/* core/runtime.impala:105, core/grid.impala:369-374 */
void md_copy_data_from_accelerator(void);
/* core/runtime.impala:104, core/grid.impala:360-367 */
void md_copy_data_to_accelerator(void);
/* core/runtime.impala:106, core/compute.impala:97-106 */
void md_initial_integration(double);
/* core/runtime.impala:107, core/compute.impala:108-116 */
void md_final_integration(double);
/* core/runtime.impala:108, core/compute.impala:118-122 */
void md_enforce_pbc(void);
/* core/runtime.impala:110-117, core/grid.impala:483-499 */
int md_write_grid_data_to_arrays(double *, struct Vector3D *, struct Vector3D *, struct Vector3D *);
/* core/runtime.impala:119-126, core/grid.impala:501-518 (returns the number
 * of LOCAL atoms, as the reference does) */
int md_write_grid_ghost_data_to_arrays(double *, struct Vector3D *, struct Vector3D *, struct Vector3D *);
/* core/runtime.impala:128-135, core/grid.impala:520-546 */
int md_write_grid_aabb_data_to_arrays(double *, struct Vector3D *, struct Vector3D *, struct Vector3D *);
/* core/runtime.impala:137, core/grid.impala:323-358 */
void md_distribute_particles(void);
/* core/runtime.impala:141-144, core/neighborlist.impala:18-78 */
void md_assemble_neighborlists(bool, double);
/* core/runtime.impala:97-102 */
void md_deallocate_grid(void);
/* core/runtime.impala:146-150 */
void md_print_grid(void);
/* core/runtime.impala:152-156 */
void md_print_ghost(void);
/* core/runtime.impala:158-186 */
void md_create_velocity(double);
/* core/runtime.impala:188-201 (half, cutoff, epsilon, sigma) */
void md_compute_lennard_jones(bool, double, double, double);
/* core/runtime.impala:203-247 (half, cutoff, damping_n, damping_t,
 * stiffness, friction); contact model corrected, radius from
 * tmd_set_dem_radius (default cutoff / 2) */
void md_compute_dem(bool, double, double, double, double, double);
/* core/runtime.impala:250 */
void md_mpi_initialize(void);
/* core/runtime.impala:251 */
void md_mpi_finalize(void);
/* core/runtime.impala:252, comm/comm.impala:334-388 */
void md_synchronize_ghost_layer(void);
/* core/runtime.impala:138, comm/comm.impala:528-642 */
void md_borders(void);
/* core/runtime.impala:139, comm/comm.impala:392-525 (single rank: drops
 * the ghosts) */
void md_exchange_particles(void);
/* core/runtime.impala:258 (load-balancing buffer; single GPU: 0) */
double md_get_send_buffer_value(int);
/* core/runtime.impala:259 (load-balancing buffer; single GPU: no-op) */
void md_set_recv_buffer_value(int, double);
/* core/runtime.impala:260-264 (single GPU: no-op) */
void md_resize_recv_buffer_capacity(int);
/* core/runtime.impala:266-277 (single GPU: nothing leaves, returns 0) */
int md_serialize_particles(double const *);
/* core/runtime.impala:279 (single GPU: no-op) */
void md_deserialize_particles(int);
/* core/runtime.impala:254 */
int md_get_world_size(void);
/* core/runtime.impala:255 */
int md_get_world_rank(void);
/* core/runtime.impala:256 */
int md_get_number_of_particles(void);
/* core/runtime.impala:257 */
int md_get_number_of_ghost_particles(void);
/* core/runtime.impala:52-70, comm/comm.impala:290-331 (single rank: the
 * rank AABB is the world AABB) */
void md_get_node_bounding_box(double const *, double (*)[6]);
/* core/runtime.impala:72-95 (atoms / ghosts inside an AABB) */
void md_compute_boundary_weights(double, double, double, double, double, double, unsigned long int *, unsigned long int *);
/* comm/lb.impala:52-85 (waLBerla neighbourhood; single GPU: no-op) */
void md_update_neighborhood(int, int, int const *, int const *, double const *);
/* core/runtime.impala:280-296 */
void md_report_iterations(void);
/* core/runtime.impala:298-316 */
void md_report_particles(void);
/* core/runtime.impala:318-337 */
void md_report_time(double, double, double, double, double, double);
/* core/runtime.impala:253 */
void md_barrier(void);

#ifdef __cplusplus
}
#endif

#endif /* TINYMD_AMD_MD_ABI_H */
