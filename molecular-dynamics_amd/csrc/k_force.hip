// Pair-force kernels: Lennard-Jones and spring-dashpot DEM, full and half
// lists, FP64 and FP32-mixed.  One work slot (local atom, cell order) per
// lane; the neighbour indices of a wavefront are read as 256-byte rows of
// the tiled list (see k_nlist.hip).
//
// Full lists accumulate the force of lane i in registers and store it once:
// no atomics (core/compute.impala:92-123, full branch).
//
// Half lists (each pair once) use a deterministic reduction: the force on i
// is summed in registers and every reaction -F on a local j is added to a
// 64-bit fixed-point accumulator (scale 2^40) with integer atomics.  Integer
// addition is associative, so the result does not depend on the order in
// which waves run (the reference's half-list "+=" is a data race on GPUs,
// core/compute.impala:104-111).
//
// FP64 LJ follows the reference closure operation by operation
// (core/runtime.impala:188-201), so with the reference's list order the
// full-list FP64 force is bit-identical to the CPU restatement.
// FP32-mixed: the distance test stays FP64 (identical in/out decisions at the
// cutoff); the LJ algebra is FP32 and each component is accumulated with an
// error-free TwoSum into a (sum, compensation) pair, combined in FP64.
#include "tmd_internal.hpp"
#include "tmd_device.hpp"

namespace tmd {

struct LJArgs {
    double rc2, rc, eps, sigma6, frc, urc;
    float rc2f, epsf, sigma6f, frcf;
    int cap, nwork, n;
};

__device__ __forceinline__ const int* nb_column(const int* nbl, int w, int cap) {
    return nbl + (size_t)(w >> 6) * (size_t)cap * 64 + (size_t)(w & 63);
}

// error-free transformation: s + x = t + e exactly (Knuth TwoSum)
__device__ __forceinline__ void two_sum_acc(float& s, float& c, float x) {
    const float t = s + x;
    const float bp = t - s;
    const float ap = t - bp;
    const float e = (s - ap) + (x - bp);
    s = t;
    c += e;
}

template <int SHIFT>
__device__ __forceinline__ double lj_f64(const LJArgs& a, double rsq) {
    const double sr2 = 1.0 / rsq;
    const double sr6 = sr2 * sr2 * sr2 * a.sigma6;
    double f = 48.0 * sr6 * (sr6 - 0.5) * sr2 * a.eps;
    if (SHIFT) f = f - a.frc / sqrt(rsq);
    return f;
}

template <int SHIFT>
__device__ __forceinline__ float lj_f32(const LJArgs& a, float rsq) {
    const float sr2 = __builtin_amdgcn_rcpf(rsq);
    const float sr6 = sr2 * sr2 * sr2 * a.sigma6f;
    float f = 48.0f * sr6 * (sr6 - 0.5f) * sr2 * a.epsf;
    if (SHIFT) f = f - a.frcf * __builtin_amdgcn_rsqf(rsq);
    return f;
}

template <int SHIFT>
__global__ void __launch_bounds__(256) k_lj_full_f64(const double4* __restrict__ pos, const int* __restrict__ work,
                                                     const int* __restrict__ nbc, const int* __restrict__ nbl, LJArgs a,
                                                     double4* __restrict__ force) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= a.nwork) return;
    const int i = work[w];
    const double4 pi = pos[i];
    const int cnt = nbc[w];
    const int* lp = nb_column(nbl, w, a.cap);
    double fx = 0.0, fy = 0.0, fz = 0.0;
#pragma unroll 4
    for (int k = 0; k < cnt; ++k) {
        const int j = lp[k * 64];
        const double4 pj = pos[j];
        const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const double rsq = dx * dx + dy * dy + dz * dz;
        if (rsq < a.rc2) {
            const double f = lj_f64<SHIFT>(a, rsq);
            fx = fx + dx * f;
            fy = fy + dy * f;
            fz = fz + dz * f;
        }
    }
    force[i] = make_double4(fx, fy, fz, 0.0);
}

template <int SHIFT>
__global__ void __launch_bounds__(256) k_lj_full_f32(const double4* __restrict__ pos, const int* __restrict__ work,
                                                     const int* __restrict__ nbc, const int* __restrict__ nbl, LJArgs a,
                                                     double4* __restrict__ force) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= a.nwork) return;
    const int i = work[w];
    const double4 pi = pos[i];
    const int cnt = nbc[w];
    const int* lp = nb_column(nbl, w, a.cap);
    float sx = 0.f, sy = 0.f, sz = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
#pragma unroll 4
    for (int k = 0; k < cnt; ++k) {
        const int j = lp[k * 64];
        const double4 pj = pos[j];
        const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const double rsq = dx * dx + dy * dy + dz * dz;
        if (rsq < a.rc2) {
            const float f = lj_f32<SHIFT>(a, (float)rsq);
            two_sum_acc(sx, cx, (float)dx * f);
            two_sum_acc(sy, cy, (float)dy * f);
            two_sum_acc(sz, cz, (float)dz * f);
        }
    }
    force[i] = make_double4((double)sx + (double)cx, (double)sy + (double)cy, (double)sz + (double)cz, 0.0);
}

__device__ __forceinline__ unsigned long long to_fix(double v) { return (unsigned long long)__double2ll_rn(v * kFixScale); }

template <int PREC, int SHIFT>
__global__ void __launch_bounds__(256) k_lj_half(const double4* __restrict__ pos, const int* __restrict__ work,
                                                 const int* __restrict__ nbc, const int* __restrict__ nbl, LJArgs a,
                                                 unsigned long long* __restrict__ acc, int* __restrict__ flag) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= a.nwork) return;
    const int i = work[w];
    const double4 pi = pos[i];
    const int cnt = nbc[w];
    const int* lp = nb_column(nbl, w, a.cap);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    bool big = false;
    for (int k = 0; k < cnt; ++k) {
        const int j = lp[k * 64];
        const double4 pj = pos[j];
        const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const double rsq = dx * dx + dy * dy + dz * dz;
        if (rsq < a.rc2) {
            double Fx, Fy, Fz;
            if (PREC == PREC_FP64) {
                const double f = lj_f64<SHIFT>(a, rsq);
                Fx = dx * f; Fy = dy * f; Fz = dz * f;
            } else {
                const float f = lj_f32<SHIFT>(a, (float)rsq);
                Fx = (double)((float)dx * f); Fy = (double)((float)dy * f); Fz = (double)((float)dz * f);
            }
            fx = fx + Fx; fy = fy + Fy; fz = fz + Fz;
            if (j < a.n) {
                big |= fabs(Fx) > kFixMax || fabs(Fy) > kFixMax || fabs(Fz) > kFixMax;
                atomicAdd(&acc[3 * (size_t)j + 0], to_fix(-Fx));
                atomicAdd(&acc[3 * (size_t)j + 1], to_fix(-Fy));
                atomicAdd(&acc[3 * (size_t)j + 2], to_fix(-Fz));
            }
        }
    }
    big |= fabs(fx) > kFixMax || fabs(fy) > kFixMax || fabs(fz) > kFixMax;
    atomicAdd(&acc[3 * (size_t)i + 0], to_fix(fx));
    atomicAdd(&acc[3 * (size_t)i + 1], to_fix(fy));
    atomicAdd(&acc[3 * (size_t)i + 2], to_fix(fz));
    if (big) atomicOr(flag, 1);
}

// fixed point -> FP64 forces; clears the accumulators for the next step
__global__ void k_fix_to_force(unsigned long long* __restrict__ acc, int n, double4* __restrict__ force) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long ax = (long long)acc[3 * (size_t)i + 0];
    const long long ay = (long long)acc[3 * (size_t)i + 1];
    const long long az = (long long)acc[3 * (size_t)i + 2];
    force[i] = make_double4((double)ax * kFixInv, (double)ay * kFixInv, (double)az * kFixInv, 0.0);
    acc[3 * (size_t)i + 0] = 0ull;
    acc[3 * (size_t)i + 1] = 0ull;
    acc[3 * (size_t)i + 2] = 0ull;
}

// ---------------------------------------------------------------------------
// energy and virial (always FP64): U = 4 eps (sr12 - sr6) [+ shift terms],
// W = sum r.F; full lists weight each entry 1/2, half lists weight
// local-local pairs 1 and local-ghost pairs 1/2.
template <int SHIFT>
__global__ void __launch_bounds__(256) k_lj_energy(const double4* __restrict__ pos, const int* __restrict__ work,
                                                   const int* __restrict__ nbc, const int* __restrict__ nbl, LJArgs a, int half,
                                                   double* __restrict__ partials) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    double pe = 0.0, vir = 0.0, np = 0.0;
    if (w < a.nwork) {
        const int i = work[w];
        const double4 pi = pos[i];
        const int cnt = nbc[w];
        const int* lp = nb_column(nbl, w, a.cap);
        for (int k = 0; k < cnt; ++k) {
            const int j = lp[k * 64];
            const double4 pj = pos[j];
            const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const double rsq = dx * dx + dy * dy + dz * dz;
            if (rsq < a.rc2) {
                const double sr2 = 1.0 / rsq;
                const double sr6 = sr2 * sr2 * sr2 * a.sigma6;
                double f = 48.0 * sr6 * (sr6 - 0.5) * sr2 * a.eps;
                double u = 4.0 * a.eps * sr6 * (sr6 - 1.0);
                if (SHIFT) {
                    const double r = sqrt(rsq);
                    f = f - a.frc / r;
                    u = u - a.urc + (r - a.rc) * a.frc;
                }
                const double wt = half ? (j < a.n ? 1.0 : 0.5) : 0.5;
                pe += wt * u;
                vir += wt * (f * rsq);
                np += 1.0;
            }
        }
    }
    const double bpe = block_sum_double<256>(pe);
    const double bvir = block_sum_double<256>(vir);
    const double bnp = block_sum_double<256>(np);
    if (threadIdx.x == 0) {
        partials[3 * blockIdx.x + 0] = bpe;
        partials[3 * blockIdx.x + 1] = bvir;
        partials[3 * blockIdx.x + 2] = bnp;
    }
}

// single block, fixed order: sums `stride`-strided partial tuples of width k
__global__ void __launch_bounds__(256) k_reduce_partials(const double* __restrict__ partials, int nblocks, int width,
                                                         double* __restrict__ out) {
    for (int c = 0; c < width; ++c) {
        double v = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += 256) v += partials[(size_t)width * b + c];
        const double t = block_sum_double<256>(v);
        if (threadIdx.x == 0) out[c] = t;
    }
}

void reduce_partials(Sim& s, const double* partials, int nblocks, int width, double* out) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s.stream, partials, nblocks, width, out);
    TMD_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// DEM spring-dashpot contact (corrected model; see tinymd_amd docs):
// delta = 2R - r > 0, n = del / r, v = v_i - v_j, v_n = v.n, v_t = v - v_n n,
// F = (k delta - g_n v_n) n - min(g_t |v_t|, mu |k delta - g_n v_n|) v_t/|v_t|.
struct DEMArgs {
    double rc2, d2, gn, gt, k, mu;
    int cap, nwork, n;
};

template <typename R>
__device__ __forceinline__ void dem_pair(const DEMArgs& a, R dx, R dy, R dz, R rsq, R vx, R vy, R vz, R& Fx, R& Fy, R& Fz) {
    Fx = Fy = Fz = R(0);
    const R d2 = (R)a.d2;
    if (rsq < d2 * d2) {
        const R r = sqrt(rsq);
        const R delta = d2 - r;
        const R nx = dx / r, ny = dy / r, nz = dz / r;
        const R vn = vx * nx + vy * ny + vz * nz;
        const R tx = vx - vn * nx, ty = vy - vn * ny, tz = vz - vn * nz;
        const R vtl = sqrt(tx * tx + ty * ty + tz * tz);
        const R fn = (R)a.k * delta - (R)a.gn * vn;
        const R ft = fmin((R)a.gt * vtl, (R)a.mu * fabs(fn));
        const R inv = vtl > R(0) ? ft / vtl : R(0);
        Fx = fn * nx - inv * tx;
        Fy = fn * ny - inv * ty;
        Fz = fn * nz - inv * tz;
    }
}

template <int PREC, int HALF>
__global__ void __launch_bounds__(256) k_dem(const double4* __restrict__ pos, const double4* __restrict__ vel,
                                             const int* __restrict__ work, const int* __restrict__ nbc,
                                             const int* __restrict__ nbl, DEMArgs a, double4* __restrict__ force,
                                             unsigned long long* __restrict__ acc, int* __restrict__ flag) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= a.nwork) return;
    const int i = work[w];
    const double4 pi = pos[i];
    const double4 vi = vel[i];
    const int cnt = nbc[w];
    const int* lp = nb_column(nbl, w, a.cap);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    bool big = false;
    for (int k = 0; k < cnt; ++k) {
        const int j = lp[k * 64];
        const double4 pj = pos[j];
        const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const double rsq = dx * dx + dy * dy + dz * dz;
        if (rsq < a.rc2) {
            const double4 vj = vel[j];
            double Fx, Fy, Fz;
            if (PREC == PREC_FP64) {
                dem_pair<double>(a, dx, dy, dz, rsq, vi.x - vj.x, vi.y - vj.y, vi.z - vj.z, Fx, Fy, Fz);
            } else {
                float gx, gy, gz;
                dem_pair<float>(a, (float)dx, (float)dy, (float)dz, (float)rsq, (float)(vi.x - vj.x), (float)(vi.y - vj.y),
                                (float)(vi.z - vj.z), gx, gy, gz);
                Fx = gx; Fy = gy; Fz = gz;
            }
            fx = fx + Fx; fy = fy + Fy; fz = fz + Fz;
            if (HALF && j < a.n) {
                big |= fabs(Fx) > kFixMax || fabs(Fy) > kFixMax || fabs(Fz) > kFixMax;
                atomicAdd(&acc[3 * (size_t)j + 0], to_fix(-Fx));
                atomicAdd(&acc[3 * (size_t)j + 1], to_fix(-Fy));
                atomicAdd(&acc[3 * (size_t)j + 2], to_fix(-Fz));
            }
        }
    }
    if (HALF) {
        big |= fabs(fx) > kFixMax || fabs(fy) > kFixMax || fabs(fz) > kFixMax;
        atomicAdd(&acc[3 * (size_t)i + 0], to_fix(fx));
        atomicAdd(&acc[3 * (size_t)i + 1], to_fix(fy));
        atomicAdd(&acc[3 * (size_t)i + 2], to_fix(fz));
        if (big) atomicOr(flag, 1);
    } else {
        force[i] = make_double4(fx, fy, fz, 0.0);
    }
}

// ---------------------------------------------------------------------------
static LJArgs make_lj_args(const Sim& s, double rc, double eps, double sigma) {
    LJArgs a;
    a.rc = rc;
    a.rc2 = rc * rc;
    a.eps = eps;
    const double s3 = sigma * (sigma * sigma);  // pow(sigma, 6): utils/utilities.impala:27-36
    a.sigma6 = s3 * s3;
    const double sr2 = 1.0 / a.rc2;
    const double sr6 = sr2 * sr2 * sr2 * a.sigma6;
    a.frc = 48.0 * sr6 * (sr6 - 0.5) * sr2 * eps * rc;
    a.urc = 4.0 * eps * sr6 * (sr6 - 1.0);
    a.rc2f = (float)a.rc2;
    a.epsf = (float)eps;
    a.sigma6f = (float)a.sigma6;
    a.frcf = (float)a.frc;
    a.cap = s.nb_cap;
    a.nwork = s.n;
    a.n = s.n;
    return a;
}

static void ensure_acc(Sim& s) {
    const size_t need = sizeof(unsigned long long) * 3 * (size_t)(s.n > 0 ? s.n : 1);
    if (s.acc.bytes < need) {
        s.acc.reserve(need, false, s.stream);
        TMD_HIP(hipMemsetAsync(s.acc.p, 0, s.acc.bytes, s.stream));
    }
}

static void check_fix_flag(Sim& s) {
    int host = 0;
    int* flag = s.counters.as<int>() + 4;
    TMD_HIP(hipMemcpyAsync(&host, flag, sizeof(int), hipMemcpyDeviceToHost, s.stream));
    TMD_HIP(hipStreamSynchronize(s.stream));
    if (host) set_error(__FILE__, __LINE__, "half-list force exceeds the fixed-point accumulator range (|F| > 4e6)");
}

void compute_lj(Sim& s, int half, double rc, double eps, double sigma) {
    if (has_error()) return;
    TMD_REQUIRE(s.nb_valid, "compute forces: no neighbour list (call md_assemble_neighborlists first)");
    TMD_REQUIRE(half == s.nb_half, "compute forces: half flag differs from the one the neighbour list was built with");
    if (has_error() || s.n == 0) return;
    LJArgs a = make_lj_args(s, rc, eps, sigma);
    const int nt = 256, nb = div_up(s.n, nt);
    const int shift = s.lj_shift;
    hipEvent_t ev = nullptr;
    s.timer.begin(K_FORCE, s.stream, &ev);
    if (!half) {
        const double4* P = s.pos.as<double4>();
        const int* W = s.work.as<int>();
        const int* C = s.nb_count.as<int>();
        const int* L = s.nb_list.as<int>();
        double4* F = s.force.as<double4>();
        if (s.precision == PREC_FP64) {
            if (shift) hipLaunchKernelGGL(k_lj_full_f64<1>, dim3(nb), dim3(nt), 0, s.stream, P, W, C, L, a, F);
            else hipLaunchKernelGGL(k_lj_full_f64<0>, dim3(nb), dim3(nt), 0, s.stream, P, W, C, L, a, F);
        } else {
            if (shift) hipLaunchKernelGGL(k_lj_full_f32<1>, dim3(nb), dim3(nt), 0, s.stream, P, W, C, L, a, F);
            else hipLaunchKernelGGL(k_lj_full_f32<0>, dim3(nb), dim3(nt), 0, s.stream, P, W, C, L, a, F);
        }
    } else {
        ensure_acc(s);
        int* flag = s.counters.as<int>() + 4;
        TMD_HIP(hipMemsetAsync(flag, 0, sizeof(int), s.stream));
        const double4* P = s.pos.as<double4>();
        const int* W = s.work.as<int>();
        const int* C = s.nb_count.as<int>();
        const int* L = s.nb_list.as<int>();
        unsigned long long* A = s.acc.as<unsigned long long>();
        if (s.precision == PREC_FP64) {
            if (shift) hipLaunchKernelGGL((k_lj_half<PREC_FP64, 1>), dim3(nb), dim3(nt), 0, s.stream, P, W, C, L, a, A, flag);
            else hipLaunchKernelGGL((k_lj_half<PREC_FP64, 0>), dim3(nb), dim3(nt), 0, s.stream, P, W, C, L, a, A, flag);
        } else {
            if (shift) hipLaunchKernelGGL((k_lj_half<PREC_FP32_MIXED, 1>), dim3(nb), dim3(nt), 0, s.stream, P, W, C, L, a, A, flag);
            else hipLaunchKernelGGL((k_lj_half<PREC_FP32_MIXED, 0>), dim3(nb), dim3(nt), 0, s.stream, P, W, C, L, a, A, flag);
        }
        hipLaunchKernelGGL(k_fix_to_force, dim3(nb), dim3(nt), 0, s.stream, A, s.n, s.force.as<double4>());
    }
    s.timer.end(K_FORCE, s.stream, ev);
    TMD_HIP(hipGetLastError());
    if (half && s.sync_calls) check_fix_flag(s);
}

void compute_dem(Sim& s, int half, double rc, double gn, double gt, double k, double mu) {
    if (has_error()) return;
    TMD_REQUIRE(s.nb_valid, "compute forces: no neighbour list (call md_assemble_neighborlists first)");
    TMD_REQUIRE(half == s.nb_half, "compute forces: half flag differs from the one the neighbour list was built with");
    if (has_error() || s.n == 0) return;
    const double radius = s.dem_radius > 0.0 ? s.dem_radius : 0.5 * rc;
    DEMArgs a{rc * rc, 2.0 * radius, gn, gt, k, mu, s.nb_cap, s.n, s.n};
    const int nt = 256, nb = div_up(s.n, nt);
    ensure_acc(s);
    int* flag = s.counters.as<int>() + 4;
    TMD_HIP(hipMemsetAsync(flag, 0, sizeof(int), s.stream));
    const double4* P = s.pos.as<double4>();
    const double4* V = s.vel.as<double4>();
    const int* W = s.work.as<int>();
    const int* C = s.nb_count.as<int>();
    const int* L = s.nb_list.as<int>();
    double4* F = s.force.as<double4>();
    unsigned long long* A = s.acc.as<unsigned long long>();
    hipEvent_t ev = nullptr;
    s.timer.begin(K_FORCE, s.stream, &ev);
    if (s.precision == PREC_FP64) {
        if (half) hipLaunchKernelGGL((k_dem<PREC_FP64, 1>), dim3(nb), dim3(nt), 0, s.stream, P, V, W, C, L, a, F, A, flag);
        else hipLaunchKernelGGL((k_dem<PREC_FP64, 0>), dim3(nb), dim3(nt), 0, s.stream, P, V, W, C, L, a, F, A, flag);
    } else {
        if (half) hipLaunchKernelGGL((k_dem<PREC_FP32_MIXED, 1>), dim3(nb), dim3(nt), 0, s.stream, P, V, W, C, L, a, F, A, flag);
        else hipLaunchKernelGGL((k_dem<PREC_FP32_MIXED, 0>), dim3(nb), dim3(nt), 0, s.stream, P, V, W, C, L, a, F, A, flag);
    }
    if (half) hipLaunchKernelGGL(k_fix_to_force, dim3(nb), dim3(nt), 0, s.stream, A, s.n, F);
    s.timer.end(K_FORCE, s.stream, ev);
    TMD_HIP(hipGetLastError());
    if (half && s.sync_calls) check_fix_flag(s);
}

void lj_energy_virial(Sim& s, double rc, double eps, double sigma, double* pe, double* vir, long long* pairs) {
    *pe = 0.0; *vir = 0.0; *pairs = 0;
    if (has_error() || !s.nb_valid || s.n == 0) return;
    LJArgs a = make_lj_args(s, rc, eps, sigma);
    const int nt = 256, nb = div_up(s.n, nt);
    s.red.reserve(sizeof(double) * (3 * (size_t)nb + 8), false, s.stream);
    double* part = s.red.as<double>();
    double* out = part + 3 * (size_t)nb;
    if (s.lj_shift)
        hipLaunchKernelGGL(k_lj_energy<1>, dim3(nb), dim3(nt), 0, s.stream, (const double4*)s.pos.as<double4>(), (const int*)s.work.as<int>(),
                           (const int*)s.nb_count.as<int>(), (const int*)s.nb_list.as<int>(), a, s.nb_half, part);
    else
        hipLaunchKernelGGL(k_lj_energy<0>, dim3(nb), dim3(nt), 0, s.stream, (const double4*)s.pos.as<double4>(), (const int*)s.work.as<int>(),
                           (const int*)s.nb_count.as<int>(), (const int*)s.nb_list.as<int>(), a, s.nb_half, part);
    reduce_partials(s, part, nb, 3, out);
    double host[3] = {0, 0, 0};
    TMD_HIP(hipMemcpyAsync(host, out, sizeof host, hipMemcpyDeviceToHost, s.stream));
    TMD_HIP(hipStreamSynchronize(s.stream));
    *pe = host[0];
    *vir = host[1];
    *pairs = (long long)host[2];
}

}  // namespace tmd
