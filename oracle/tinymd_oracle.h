/* tinyMD CPU oracle -- TEST INFRASTRUCTURE ONLY (tests/, smoke(), bench.py's
 * cpu_baseline leg).  The product library never includes or links this.
 * C API over an opaque simulation handle; arrays are AoS [3*i + d] doubles. */
#ifndef TINYMD_ORACLE_H
#define TINYMD_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif
double orc_pm_random(int* seed);
int orc_lattice(const double* aabb, const double* rank_aabb, int half, double spacing,
                double* mass, double* pos, double* vel, int cap);
void* orc_create(const double* m, const double* pos, const double* vel, int n,
                 const double* aabb, const double* rank_aabb, double cell_spacing, int nb_cap);
void orc_destroy(void* h);
void orc_set_threads(void* h, int t);
void orc_set_ghost_mode(void* h, int mode);
void orc_grid_info(void* h, int* out);
void orc_create_velocity(void* h, double t_request);
void orc_exchange(void* h);
void orc_enforce_pbc(void* h);
void orc_set_sort(void* h, int on);
void orc_borders(void* h);
void orc_distribute(void* h);
void orc_neighbors(void* h, int half, int half_mode, double cutoff);
void orc_sync_ghosts(void* h);
void orc_initial_integration(void* h, double dt);
void orc_final_integration(void* h, double dt);
void orc_lj(void* h, int half, double rc, double eps, double sigma, int shift);
void orc_dem(void* h, int half, double rc, double radius, double gn, double gt, double k, double mu);
void orc_thermo(void* h, double* out);
void orc_get_state(void* h, double* pos, double* vel, double* force);
void orc_set_state(void* h, const double* pos, const double* vel);
void orc_get_ghosts(void* h, int* owner, int* img, int* src, int* flag, double* pos);
void orc_get_cells(void* h, int* start, int* particles, int* pcell);
void orc_get_neighbors(void* h, int* counts, int* list, int stride);
int orc_run_lj(void* h, int steps, int reneigh_every, double dt, int half, int half_mode,
               double rc, double skin, double eps, double sigma, int shift);
int orc_run_dem(void* h, int steps, int reneigh_every, double dt, int half, int half_mode,
                double rc, double skin, double radius, double gn, double gt, double k, double mu);
void orc_setup(void* h, int half, int half_mode, double cutoff);
void orc_traj_thermo(void* h, int steps, int reneigh_every, double dt, double rc, double skin, double eps, double sigma,
                     int shift, double* out);
void orc_traj_adjoint(void* h, int steps, int reneigh_every, double dt, double rc, double skin, double eps,
                      double sigma, int shift, double u_target, double p_target, double* out3, double* gtheta,
                      double* gx, double* gv);
#ifdef __cplusplus
}
#endif
#endif
