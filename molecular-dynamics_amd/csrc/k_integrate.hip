// Velocity Verlet (unit mass), the periodic wrap and the kinetic-energy
// reduction.  core/compute.impala:97-122 (integration, PBC),
// comm/comm.impala:152-162 (single-image wrap), core/thermo.impala:1-16.
#include "tmd_internal.hpp"
#include "tmd_device.hpp"

namespace tmd {

// v += f * (dt/2); x += v * dt; with MIRROR the cell-ordered position
// mirror is kept current too (spos[inv[i]] = x)
template <int MIRROR>
__global__ void k_initial_integration(double4* __restrict__ pos, double4* __restrict__ vel, const double4* __restrict__ force,
                                      int n, double hdt, double dt, const int* __restrict__ inv, const Mir spos) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 v = vel[i];
    const double4 f = force[i];
    double4 x = pos[i];
    v.x = v.x + f.x * hdt;
    v.y = v.y + f.y * hdt;
    v.z = v.z + f.z * hdt;
    x.x = x.x + v.x * dt;
    x.y = x.y + v.y * dt;
    x.z = x.z + v.z * dt;
    vel[i] = v;
    pos[i] = x;
    if (MIRROR) {
        const int k = inv[i];
        if (k >= 0) spos[k] = x;
    }
}

__global__ void k_final_integration(double4* __restrict__ vel, const double4* __restrict__ force, int n, double hdt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 v = vel[i];
    const double4 f = force[i];
    v.x = v.x + f.x * hdt;
    v.y = v.y + f.y * hdt;
    v.z = v.z + f.z * hdt;
    vel[i] = v;
}

__global__ void k_enforce_pbc(double4* __restrict__ pos, int n, PbcArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 p = pos[i];
    pbc_wrap(p, a);
    pos[i] = p;
}

__global__ void __launch_bounds__(256) k_kinetic(const double4* __restrict__ vel, int n, double* __restrict__ partials) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double t = 0.0;
    if (i < n) {
        const double4 v = vel[i];
        t = (v.x * v.x + v.y * v.y + v.z * v.z) * 1.0;
    }
    const double b = block_sum_double<256>(t);
    if (threadIdx.x == 0) partials[blockIdx.x] = b;
}


void initial_integration(Sim& s, double dt) {
    if (has_error() || s.n == 0) return;
    const int nt = 256;
    hipEvent_t ev = nullptr;
    s.timer.begin(K_INTEGRATE_1, s.stream, &ev);
    if (s.inv_valid)
        hipLaunchKernelGGL(k_initial_integration<1>, dim3(div_up(s.n, nt)), dim3(nt), 0, s.stream, s.pos.as<double4>(),
                           s.vel.as<double4>(), (const double4*)s.force.as<double4>(), s.n, 0.5 * dt, dt,
                           (const int*)s.cell_inv.as<int>(), Mir{s.spos.as<double>()});
    else
        hipLaunchKernelGGL(k_initial_integration<0>, dim3(div_up(s.n, nt)), dim3(nt), 0, s.stream, s.pos.as<double4>(),
                           s.vel.as<double4>(), (const double4*)s.force.as<double4>(), s.n, 0.5 * dt, dt, (const int*)nullptr,
                           Mir{nullptr});
    if (!s.inv_valid) mirror_invalidate(s);
    s.timer.end(K_INTEGRATE_1, s.stream, ev);
    TMD_HIP(hipGetLastError());
}

void final_integration(Sim& s, double dt) {
    if (has_error() || s.n == 0) return;
    const int nt = 256;
    hipEvent_t ev = nullptr;
    s.timer.begin(K_INTEGRATE_2, s.stream, &ev);
    hipLaunchKernelGGL(k_final_integration, dim3(div_up(s.n, nt)), dim3(nt), 0, s.stream, s.vel.as<double4>(),
                       (const double4*)s.force.as<double4>(), s.n, 0.5 * dt);
    s.timer.end(K_INTEGRATE_2, s.stream, ev);
    TMD_HIP(hipGetLastError());
}

PbcArgs pbc_args(const Sim& s) {
    PbcArgs a;
    for (int d = 0; d < 3; ++d) {
        a.lo[d] = s.g.world[2 * d];
        a.hi[d] = s.g.world[2 * d + 1];
        a.len[d] = s.g.len[d];
    }
    return a;
}

void enforce_pbc(Sim& s) {
    if (has_error() || s.n == 0) return;
    const PbcArgs a = pbc_args(s);
    const int nt = 256;
    mirror_invalidate(s);
    hipLaunchKernelGGL(k_enforce_pbc, dim3(div_up(s.n, nt)), dim3(nt), 0, s.stream, s.pos.as<double4>(), s.n, a);
    TMD_HIP(hipGetLastError());
}

int launch_kinetic(Sim& s, double* part) {
    const int nt = 256, nb = div_up(s.n, nt);
    hipLaunchKernelGGL(k_kinetic, dim3(nb), dim3(nt), 0, s.stream, (const double4*)s.vel.as<double4>(), s.n, part);
    TMD_HIP(hipGetLastError());
    return nb;
}

double kinetic_sum(Sim& s, double* buf) {
    if (has_error() || s.n == 0) return 0.0;
    const int nb = div_up(s.n, 256);
    if (!buf) s.red.reserve(sizeof(double) * ((size_t)nb + 8), false, s.stream);
    if (has_error()) return 0.0;
    double* part = buf ? buf : s.red.as<double>();
    double* out = part + nb;
    launch_kinetic(s, part);
    reduce_partials(s, part, nb, 1, out);
    double host = 0.0;
    TMD_HIP(hipMemcpyAsync(&host, out, sizeof host, hipMemcpyDeviceToHost, s.stream));
    TMD_HIP(hipStreamSynchronize(s.stream));
    return host;
}

}  // namespace tmd
