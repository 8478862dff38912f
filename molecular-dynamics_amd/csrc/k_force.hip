// Pair-force kernels: Lennard-Jones and spring-dashpot DEM, full and half
// lists, FP64 and FP32-mixed.
//
// One block per tile of cells, one lane per local atom of the tile.  Each
// block stages the FP64 positions of its window (the tile plus one cell on
// every side, tmd_window.hpp) from the cell-ordered mirror into LDS once; list entries are 16-bit window offsets read 8 at a time
// (one int4 per lane, 1 KiB per wavefront), and the 8 neighbour positions
// are read from LDS before any arithmetic.  The pair loop therefore streams
// 2 bytes of list per entry from HBM and does no global gathers.
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
// (core/runtime.impala:188-201) in list order, so full-list FP64 forces are
// bit-identical to the CPU restatement.  FP32-mixed: the distance test stays
// FP64 (identical in/out decisions at the cutoff); the LJ algebra is FP32 and
// dx * f is accumulated in FP64 (each FP32 term is exact there, and the
// extra mantissa bits absorb the summation rounding).  The Kahan variant
// accumulates in FP32 with a compensation term instead.
#include "tmd_internal.hpp"
#include "tmd_window.hpp"
#include "tmd_lj.hpp"
#include "tmd_pair.hpp"

namespace tmd {



#define TMD_KERNEL_LIST_PARAMS                                                                                   \
    const double4 *__restrict__ pos, const double4 *__restrict__ spos, const int *__restrict__ start,              \
        const int *__restrict__ parts, const int *__restrict__ work, const int *__restrict__ tile_start,           \
        const int *__restrict__ nbc, const nb_t *__restrict__ nbl

// KICK 0: force only.  KICK 1: plus the final velocity-Verlet kick
// v += F dt/2.  KICK 2: plus the next step's first kick and drift,
// v += F dt/2, x += v dt, so the step loop needs no separate integration
// pass.  Positions (the lane's own and its neighbours') come from the
// cell-ordered mirror, at the atom's mirror index wk[w] for its own.  KICK 2
// writes the new position only into the second mirror (spos_next[wk[w]]),
// not into pos, and no force: inside the fused step loop the force of a
// step that is followed by another one is consumed here (the kick) and the
// loop restores pos from the mirror before it rebuilds or returns.  The
// arithmetic is that of the separate integration kernels.
template <int PREC, int SHIFT, int UNIT, int KICK>
__global__ void __launch_bounds__(kMaxTileThreads) k_lj_full(TMD_KERNEL_LIST_PARAMS, LJArgs a, double4* __restrict__ force,
                                                         double4* __restrict__ vel, double hdt, double dt,
                                                         const int* __restrict__ wk, double4* __restrict__ spos_next,
                                                         int vel_by_k) {
    extern __shared__ double dyn[];
    const WinStage sp = win_stage(dyn, a.wcap);
    const int tile = xcd_block(blockIdx.x, gridDim.x);
    const int w0 = tile_start[tile], nloc = tile_start[tile + 1] - w0;
    if (nloc == 0) return;
    __shared__ TileWin W;
    tile_window(tile, start, a.g, W);
    const bool staged = stage_window(W, a.wcap, spos, nullptr, sp);
    for (int at = threadIdx.x; at < nloc; at += blockDim.x) {
        const int w = w0 + at;
        const int i = work[w];
        const int kw = wk[w];
        const double4 pi = spos[kw];
        const int cnt = nbc[w];
        const int4* lp = reinterpret_cast<const int4*>(nbl + (size_t)(w >> 6) * (size_t)a.cap * 64) + (w & 63);
        const double4 f = staged ? lj_full_atom<PREC, SHIFT, UNIT, true>(pi, cnt, lp, sp, spos, W, a)
                                 : lj_full_atom<PREC, SHIFT, UNIT, false>(pi, cnt, lp, sp, spos, W, a);
        if (KICK != 2) force[i] = f;
        if (KICK) {
            // velocity-Verlet final kick (k_final_integration's arithmetic)
            const int iv = vel_by_k ? kw : i;  // velocities in atom or in mirror order
            double4 v = vel[iv];
            v.x = v.x + f.x * hdt;
            v.y = v.y + f.y * hdt;
            v.z = v.z + f.z * hdt;
            if (KICK == 2) {
                // next step's k_initial_integration
                v.x = v.x + f.x * hdt;
                v.y = v.y + f.y * hdt;
                v.z = v.z + f.z * hdt;
                double4 x = pi;
                x.x = x.x + v.x * dt;
                x.y = x.y + v.y * dt;
                x.z = x.z + v.z * dt;
                spos_next[kw] = x;
            }
            vel[iv] = v;
        }
    }
}

__device__ __forceinline__ unsigned long long to_fix(double v) { return (unsigned long long)__double2ll_rn(v * kFixScale); }

// Half-list LJ.  Each pair is evaluated once, by the lane of its higher
// index (local-ghost pairs by the local side; no reaction on ghosts).  The
// force on the lane's own atom is summed in registers in list order and
// stored once (`own`).  Reactions -F on local neighbours are added, as
// 64-bit fixed-point integers (scale 2^40), to LDS accumulators of the
// staged window, and each window atom's accumulated reaction is added to
// its global accumulator with one integer atomic per component when the
// block ends.  Integer addition is associative, so the result does not
// depend on the order in which lanes and blocks run.  A window too large
// for LDS sends reactions straight to the global accumulators.
template <int PREC, int SHIFT>
__global__ void __launch_bounds__(kMaxTileThreads) k_lj_half(TMD_KERNEL_LIST_PARAMS, LJArgs a, int wcap_half,
                                                         double4* __restrict__ own, unsigned long long* __restrict__ acc,
                                                         int* __restrict__ flag) {
    extern __shared__ double dyn[];
    const int tile = xcd_block(blockIdx.x, gridDim.x);
    const int w0 = tile_start[tile], nloc = tile_start[tile + 1] - w0;
    if (nloc == 0) return;
    __shared__ TileWin W;
    tile_window(tile, start, a.g, W);
    const WinStage sp = win_stage(dyn, wcap_half);
    // wcap_half is even, so the accumulators start 8-byte aligned
    unsigned long long* racc = reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(dyn) + (size_t)kStageBytes * wcap_half);
    const bool staged = stage_window(W, wcap_half, spos, parts, sp);
    const int S = win_size(W);
    if (staged) {
        for (int t = threadIdx.x; t < 3 * S; t += blockDim.x) racc[t] = 0ull;
        __syncthreads();
    }
    bool big = false;
    for (int at = threadIdx.x; at < nloc; at += blockDim.x) {
        const int w = w0 + at;
        const int i = work[w];
        const double4 pi = pos[i];
        const int cnt = nbc[w];
        const int4* lp = reinterpret_cast<const int4*>(nbl + (size_t)(w >> 6) * (size_t)a.cap * 64) + (w & 63);
        double fx = 0.0, fy = 0.0, fz = 0.0;
        const int ng = (cnt + 7) >> 3;
        int4 q = lp[0];  // group 0 of the column always exists (capacity >= 8)
        for (int gi = 0; gi < ng; ++gi) {
            const int4 e = q;
            q = lp[(size_t)min(gi + 1, ng - 1) * 64];  // next group (clamped: always a valid load)
            const int k0 = gi * 8;
            double dx[8], dy[8], dz[8], rs[8];
            int jj[8], tt[8];
            bool ok[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const bool valid = k0 + u < cnt;
                const int t = valid ? nb_entry(e, u) : 0;
                double xj, yj, zj;
                if (staged) {
                    xj = sp.xyz[3 * t + 0]; yj = sp.xyz[3 * t + 1]; zj = sp.xyz[3 * t + 2];
                    jj[u] = sp.j[t];
                } else {
                    const int kk = win_index(t, W);
                    const double4 r = spos[kk];
                    xj = r.x; yj = r.y; zj = r.z;
                    jj[u] = parts[kk];
                }
                tt[u] = t;
                dx[u] = pi.x - xj;
                dy[u] = pi.y - yj;
                dz[u] = pi.z - zj;
                const double rsq = dx[u] * dx[u] + dy[u] * dy[u] + dz[u] * dz[u];
                ok[u] = valid && rsq < a.rc2;
                rs[u] = ok[u] ? rsq : 1.0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                double fm;
                if (PREC == PREC_FP64) {
                    const double f = lj_f64u<SHIFT, 0>(a, rs[u]);
                    fm = ok[u] ? f : 0.0;
                } else {
                    const float f = lj_f32u<SHIFT, 0>(a, (float)rs[u]);
                    fm = (double)(ok[u] ? f : 0.0f);
                }
                const double Fx = dx[u] * fm, Fy = dy[u] * fm, Fz = dz[u] * fm;
                fx = fx + Fx;
                fy = fy + Fy;
                fz = fz + Fz;
                if (ok[u] && jj[u] < a.n) {
                    big |= fabs(Fx) > kFixMax || fabs(Fy) > kFixMax || fabs(Fz) > kFixMax;
                    const unsigned long long rx = to_fix(-Fx), ry = to_fix(-Fy), rz = to_fix(-Fz);
                    if (staged) {
                        atomicAdd(&racc[3 * tt[u] + 0], rx);
                        atomicAdd(&racc[3 * tt[u] + 1], ry);
                        atomicAdd(&racc[3 * tt[u] + 2], rz);
                    } else {
                        atomicAdd(&acc[3 * (size_t)jj[u] + 0], rx);
                        atomicAdd(&acc[3 * (size_t)jj[u] + 1], ry);
                        atomicAdd(&acc[3 * (size_t)jj[u] + 2], rz);
                    }
                }
            }
        }
        own[i] = make_double4(fx, fy, fz, 0.0);
    }
    if (big) atomicOr(flag, 1);
    if (staged) {
        __syncthreads();
        for (int t = threadIdx.x; t < S; t += blockDim.x) {
            const int j = sp.j[t];
            if (j < a.n) {
                const unsigned long long rx = racc[3 * t + 0], ry = racc[3 * t + 1], rz = racc[3 * t + 2];
                if ((rx | ry | rz) != 0ull) {
                    atomicAdd(&acc[3 * (size_t)j + 0], rx);
                    atomicAdd(&acc[3 * (size_t)j + 1], ry);
                    atomicAdd(&acc[3 * (size_t)j + 2], rz);
                }
            }
        }
    }
}

// force = own sum + accumulated reactions; clears the accumulators; KICK
// as in k_lj_full
template <int KICK>
__global__ void k_half_finish(unsigned long long* __restrict__ acc, int n, double4* __restrict__ force,
                              double4* __restrict__ vel, double hdt, double dt, double4* __restrict__ pos) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long ax = (long long)acc[3 * (size_t)i + 0];
    const long long ay = (long long)acc[3 * (size_t)i + 1];
    const long long az = (long long)acc[3 * (size_t)i + 2];
    double4 f = force[i];
    f.x = f.x + (double)ax * kFixInv;
    f.y = f.y + (double)ay * kFixInv;
    f.z = f.z + (double)az * kFixInv;
    force[i] = f;
    acc[3 * (size_t)i + 0] = 0ull;
    acc[3 * (size_t)i + 1] = 0ull;
    acc[3 * (size_t)i + 2] = 0ull;
    if (KICK) {
        double4 v = vel[i];
        v.x = v.x + f.x * hdt;
        v.y = v.y + f.y * hdt;
        v.z = v.z + f.z * hdt;
        if (KICK == 2) {
            v.x = v.x + f.x * hdt;
            v.y = v.y + f.y * hdt;
            v.z = v.z + f.z * hdt;
            double4 x = pos[i];
            x.x = x.x + v.x * dt;
            x.y = x.y + v.y * dt;
            x.z = x.z + v.z * dt;
            pos[i] = x;
        }
        vel[i] = v;
    }
}

// fixed-point accumulators -> force (DEM half lists)
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

// End of a DEM step: the force (from the fixed-point reaction accumulators
// with FIX, which are cleared, else as stored by k_dem), the final kick
// v += f dt/2 and, with ADV, the next step's kick and drift, with the
// mirror kept current when inv is given (k_final_integration and
// k_initial_integration arithmetic, in that order).
template <int FIX, int ADV>
__global__ void k_dem_finish(unsigned long long* __restrict__ acc, int n, double4* __restrict__ force,
                             double4* __restrict__ vel, double hdt, double dt, double4* __restrict__ pos,
                             const int* __restrict__ inv, double4* __restrict__ spos) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 f;
    if (FIX) {
        const long long ax = (long long)acc[3 * (size_t)i + 0];
        const long long ay = (long long)acc[3 * (size_t)i + 1];
        const long long az = (long long)acc[3 * (size_t)i + 2];
        f = make_double4((double)ax * kFixInv, (double)ay * kFixInv, (double)az * kFixInv, 0.0);
        force[i] = f;
        acc[3 * (size_t)i + 0] = 0ull;
        acc[3 * (size_t)i + 1] = 0ull;
        acc[3 * (size_t)i + 2] = 0ull;
    } else {
        f = force[i];
    }
    double4 v = vel[i];
    v.x = v.x + f.x * hdt;
    v.y = v.y + f.y * hdt;
    v.z = v.z + f.z * hdt;
    if (ADV) {
        v.x = v.x + f.x * hdt;
        v.y = v.y + f.y * hdt;
        v.z = v.z + f.z * hdt;
        double4 x = pos[i];
        x.x = x.x + v.x * dt;
        x.y = x.y + v.y * dt;
        x.z = x.z + v.z * dt;
        pos[i] = x;
        if (inv) {
            const int k = inv[i];
            if (k >= 0) spos[k] = x;
        }
    }
    vel[i] = v;
}

// ---------------------------------------------------------------------------
// energy and virial (always FP64): U = 4 eps (sr12 - sr6) [+ shift terms],
// W = sum r.F; full lists weight each entry 1/2, half lists weight
// local-local pairs 1 and local-ghost pairs 1/2.  Per-block partial sums in
// a fixed order, reduced by one block in a fixed order: deterministic.
template <int SHIFT>
__global__ void __launch_bounds__(kMaxTileThreads) k_lj_energy(TMD_KERNEL_LIST_PARAMS, LJArgs a, int half,
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
    double pe = 0.0, vir = 0.0, np = 0.0;
    __shared__ TileWin W;
    tile_window(tile, start, a.g, W);
    const WinStage sp = win_stage(dyn, a.wcap);
    const bool staged = stage_window(W, a.wcap, spos, parts, sp);
    for (int at = threadIdx.x; at < nloc; at += blockDim.x) {
        const int w = w0 + at;
        const int4* lp = reinterpret_cast<const int4*>(nbl + (size_t)(w >> 6) * (size_t)a.cap * 64) + (w & 63);
        if (staged) lj_energy_atom<SHIFT, true>(pos[work[w]], nbc[w], lp, sp, spos, parts, W, a, half, a.n, pe, vir, np);
        else lj_energy_atom<SHIFT, false>(pos[work[w]], nbc[w], lp, sp, spos, parts, W, a, half, a.n, pe, vir, np);
    }
    const double bpe = block_sum_double_any(pe);
    const double bvir = block_sum_double_any(vir);
    const double bnp = block_sum_double_any(np);
    if (threadIdx.x == 0) {
        partials[3 * blockIdx.x + 0] = bpe;
        partials[3 * blockIdx.x + 1] = bvir;
        partials[3 * blockIdx.x + 2] = bnp;
    }
}

// single block, fixed order: column sums of `width`-wide partial tuples
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
    int cap, nwork, n, wcap;
    TileGeom g;
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
__global__ void __launch_bounds__(kMaxTileThreads) k_dem(TMD_KERNEL_LIST_PARAMS, const double4* __restrict__ vel, DEMArgs a,
                                                     double4* __restrict__ force, unsigned long long* __restrict__ acc,
                                                     int* __restrict__ flag) {
    extern __shared__ double dyn[];
    const int tile = xcd_block(blockIdx.x, gridDim.x);
    const int w0 = tile_start[tile], nloc = tile_start[tile + 1] - w0;
    if (nloc == 0) return;
    __shared__ TileWin W;
    tile_window(tile, start, a.g, W);
    double* sx = dyn;
    double* sy = dyn + a.wcap;
    double* sz = dyn + 2 * a.wcap;
    int* sj = reinterpret_cast<int*>(dyn + 3 * a.wcap);
    const bool staged = stage_tile(W, a.wcap, spos, parts, sx, sy, sz, sj);
    // STAGED as a compile-time constant: LDS reads, not generic-pointer loads
    auto atoms = [&](auto staged_c) {
    constexpr bool STAGED = decltype(staged_c)::value;
    for (int at = threadIdx.x; at < nloc; at += blockDim.x) {
    const int w = w0 + at;
    const int i = work[w];
    const double4 pi = pos[i];
    const double4 vi = vel[i];
    const int cnt = nbc[w];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    bool big = false;
    for (int k = 0; k < cnt; ++k) {
        const int t = nbl[nb16_slot(w, k, a.cap)];
        int j;
        double xj, yj, zj;
        if (STAGED) {
            j = sj[t];
            xj = sx[t]; yj = sy[t]; zj = sz[t];
        } else {
            const int kk = win_index(t, W);
            j = parts[kk];
            const double4 p = spos[kk];
            xj = p.x; yj = p.y; zj = p.z;
        }
        const double dx = pi.x - xj, dy = pi.y - yj, dz = pi.z - zj;
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
    };
    if (staged) atoms(std::true_type{});
    else atoms(std::false_type{});
}

// ---------------------------------------------------------------------------

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

#define TMD_LIST_ARGS                                                                                              \
    (const double4*)s.pos.as<double4>(), (const double4*)s.spos.as<double4>(), (const int*)s.cell_start.as<int>(), \
        (const int*)s.cell_parts.as<int>(), (const int*)s.work.as<int>(), (const int*)s.tile_start.as<int>(),     \
        (const int*)s.nb_count.as<int>(), (const nb_t*)s.nb_list.as<nb_t>()

static size_t lds_bytes(const Sim& s, bool) { return (size_t)window_cap(s) * kStageBytes; }

void compute_lj(Sim& s, int half, double rc, double eps, double sigma, double kick_dt, bool advance) {
    if (has_error()) return;
    TMD_REQUIRE(s.nb_valid, "compute forces: no neighbour list (call md_assemble_neighborlists first)");
    TMD_REQUIRE(half == s.nb_half, "compute forces: half flag differs from the one the neighbour list was built with");
    if (has_error() || s.n == 0) return;
    LJArgs a = make_lj_args(s, rc, eps, sigma);
    const int nb = s.tiles.ntx * s.tiles.nty * s.tiles.ntz;
    const int shift = s.lj_shift;
    hipEvent_t ev = nullptr;
    s.timer.begin(K_MIRROR, s.stream, &ev);
    if (!s.mirror_fresh) refresh_mirror(s);
    s.timer.end(K_MIRROR, s.stream, ev);
    s.timer.begin(K_FORCE, s.stream, &ev);
    if (!half) {
        double4* F = s.force.as<double4>();
        const size_t lds = lds_bytes(s, false);
        const int unit = (eps == 1.0 && a.sigma6 == 1.0) ? 1 : 0;
        const double hdt = 0.5 * kick_dt;
        // inside the fused step loop the velocities live in mirror order
        double4* V = s.vel_mirrored ? s.svel.as<double4>() : s.vel.as<double4>();
        const int VK = s.vel_mirrored;
        const int* WK = s.work_k.as<int>();
        double4* SN = s.spos2.as<double4>();
#define TMD_LJ_FULL(P, SH, U)                                                                                          \
    do {                                                                                                               \
        if (kick_dt > 0.0 && advance)                                                                                  \
            hipLaunchKernelGGL((k_lj_full<P, SH, U, 2>), dim3(nb), dim3(s.tile_threads), lds, s.stream, TMD_LIST_ARGS, a, F, V, hdt, \
                               kick_dt, WK, SN, VK);                                                                    \
        else if (kick_dt > 0.0)                                                                                        \
            hipLaunchKernelGGL((k_lj_full<P, SH, U, 1>), dim3(nb), dim3(s.tile_threads), lds, s.stream, TMD_LIST_ARGS, a, F, V, hdt, \
                               kick_dt, WK, SN, VK);                                                                    \
        else                                                                                                           \
            hipLaunchKernelGGL((k_lj_full<P, SH, U, 0>), dim3(nb), dim3(s.tile_threads), lds, s.stream, TMD_LIST_ARGS, a, F, V, hdt, \
                               kick_dt, WK, SN, VK);                                                                    \
    } while (0)
        if (s.precision == PREC_FP64) {
            if (shift) { if (unit) TMD_LJ_FULL(PREC_FP64, 1, 1); else TMD_LJ_FULL(PREC_FP64, 1, 0); }
            else { if (unit) TMD_LJ_FULL(PREC_FP64, 0, 1); else TMD_LJ_FULL(PREC_FP64, 0, 0); }
        } else if (s.precision == PREC_FP32_MIXED) {
            if (shift) { if (unit) TMD_LJ_FULL(PREC_FP32_MIXED, 1, 1); else TMD_LJ_FULL(PREC_FP32_MIXED, 1, 0); }
            else { if (unit) TMD_LJ_FULL(PREC_FP32_MIXED, 0, 1); else TMD_LJ_FULL(PREC_FP32_MIXED, 0, 0); }
        } else {
            if (shift) { if (unit) TMD_LJ_FULL(PREC_FP32_KAHAN, 1, 1); else TMD_LJ_FULL(PREC_FP32_KAHAN, 1, 0); }
            else { if (unit) TMD_LJ_FULL(PREC_FP32_KAHAN, 0, 1); else TMD_LJ_FULL(PREC_FP32_KAHAN, 0, 0); }
        }
#undef TMD_LJ_FULL
    } else {
        ensure_acc(s);
        int* flag = s.counters.as<int>() + 4;
        TMD_HIP(hipMemsetAsync(flag, 0, sizeof(int), s.stream));
        unsigned long long* A = s.acc.as<unsigned long long>();
        // window records (32 B) + three 64-bit reaction accumulators per atom
        const int per = kStageBytes + 3 * (int)sizeof(unsigned long long);
        const int wh = std::min(s.max_window + 2, std::min(s.lds_limit, kLdsHalfBudget) / per) & ~1;  // + sentinel slot
        const size_t lds = (size_t)wh * per;
        double4* F = s.force.as<double4>();
#define TMD_LJ_HALF(P, SH) \
    hipLaunchKernelGGL((k_lj_half<P, SH>), dim3(nb), dim3(s.tile_threads), lds, s.stream, TMD_LIST_ARGS, a, wh, F, A, flag)
        if (s.precision == PREC_FP64) {
            if (shift) TMD_LJ_HALF(PREC_FP64, 1); else TMD_LJ_HALF(PREC_FP64, 0);
        } else {
            if (shift) TMD_LJ_HALF(PREC_FP32_MIXED, 1); else TMD_LJ_HALF(PREC_FP32_MIXED, 0);
        }
#undef TMD_LJ_HALF
        double4* V = s.vel.as<double4>();
        double4* X = s.pos.as<double4>();
        const int nb256 = div_up(s.n, 256);
        if (kick_dt > 0.0 && advance)
            hipLaunchKernelGGL(k_half_finish<2>, dim3(nb256), dim3(256), 0, s.stream, A, s.n, F, V, 0.5 * kick_dt, kick_dt, X);
        else if (kick_dt > 0.0)
            hipLaunchKernelGGL(k_half_finish<1>, dim3(nb256), dim3(256), 0, s.stream, A, s.n, F, V, 0.5 * kick_dt, kick_dt, X);
        else
            hipLaunchKernelGGL(k_half_finish<0>, dim3(nb256), dim3(256), 0, s.stream, A, s.n, F, V, 0.0, 0.0, X);
    }
    s.timer.end(K_FORCE, s.stream, ev);
    TMD_HIP(hipGetLastError());
    if (kick_dt > 0.0 && advance) {
        if (!half) {
            // the new local positions are in spos2: it becomes the mirror, whose
            // ghost entries step_refresh rewrites
            std::swap(s.spos, s.spos2);
            s.mirror_fresh = 0;
            s.ghost_stale = 1;
            s.pos_stale = 1;
        } else {
            mirror_invalidate(s);  // local positions advanced
        }
    }
    if (half && s.sync_calls) check_fix_flag(s);
}

void compute_dem(Sim& s, int half, double rc, double gn, double gt, double k, double mu, double kick_dt, bool advance) {
    if (has_error()) return;
    TMD_REQUIRE(s.nb_valid, "compute forces: no neighbour list (call md_assemble_neighborlists first)");
    TMD_REQUIRE(half == s.nb_half, "compute forces: half flag differs from the one the neighbour list was built with");
    if (has_error() || s.n == 0) return;
    const double radius = s.dem_radius > 0.0 ? s.dem_radius : 0.5 * rc;
    DEMArgs a{rc * rc, 2.0 * radius, gn, gt, k, mu, s.nb_cap, s.n, s.n, window_cap(s), s.tiles};
    const int nb = s.tiles.ntx * s.tiles.nty * s.tiles.ntz;
    ensure_acc(s);
    int* flag = s.counters.as<int>() + 4;
    TMD_HIP(hipMemsetAsync(flag, 0, sizeof(int), s.stream));
    const double4* V = s.vel.as<double4>();
    double4* F = s.force.as<double4>();
    unsigned long long* A = s.acc.as<unsigned long long>();
    const size_t lds = lds_bytes(s, true);
    hipEvent_t ev = nullptr;
    s.timer.begin(K_MIRROR, s.stream, &ev);
    if (!s.mirror_fresh) refresh_mirror(s);
    s.timer.end(K_MIRROR, s.stream, ev);
    s.timer.begin(K_FORCE, s.stream, &ev);
    if (s.precision == PREC_FP64) {
        if (half) hipLaunchKernelGGL((k_dem<PREC_FP64, 1>), dim3(nb), dim3(s.tile_threads), lds, s.stream, TMD_LIST_ARGS, V, a, F, A, flag);
        else hipLaunchKernelGGL((k_dem<PREC_FP64, 0>), dim3(nb), dim3(s.tile_threads), lds, s.stream, TMD_LIST_ARGS, V, a, F, A, flag);
    } else {
        if (half) hipLaunchKernelGGL((k_dem<PREC_FP32_MIXED, 1>), dim3(nb), dim3(s.tile_threads), lds, s.stream, TMD_LIST_ARGS, V, a, F, A, flag);
        else hipLaunchKernelGGL((k_dem<PREC_FP32_MIXED, 0>), dim3(nb), dim3(s.tile_threads), lds, s.stream, TMD_LIST_ARGS, V, a, F, A, flag);
    }
    if (kick_dt > 0.0) {
        // final kick (and, advancing, the next kick and drift) fused
        double4* V2 = s.vel.as<double4>();
        double4* X = s.pos.as<double4>();
        const int* I = s.inv_valid ? (const int*)s.cell_inv.as<int>() : nullptr;
        double4* M = s.spos.as<double4>();
        const double hdt = 0.5 * kick_dt;
        const dim3 g(div_up(s.n, 256)), b(256);
        if (half) {
            if (advance) hipLaunchKernelGGL((k_dem_finish<1, 1>), g, b, 0, s.stream, A, s.n, F, V2, hdt, kick_dt, X, I, M);
            else hipLaunchKernelGGL((k_dem_finish<1, 0>), g, b, 0, s.stream, A, s.n, F, V2, hdt, kick_dt, X, I, M);
        } else {
            if (advance) hipLaunchKernelGGL((k_dem_finish<0, 1>), g, b, 0, s.stream, A, s.n, F, V2, hdt, kick_dt, X, I, M);
            else hipLaunchKernelGGL((k_dem_finish<0, 0>), g, b, 0, s.stream, A, s.n, F, V2, hdt, kick_dt, X, I, M);
        }
        if (advance && !s.inv_valid) mirror_invalidate(s);
    } else if (half) {
        hipLaunchKernelGGL(k_fix_to_force, dim3(div_up(s.n, 256)), dim3(256), 0, s.stream, A, s.n, F);
    }
    s.timer.end(K_FORCE, s.stream, ev);
    TMD_HIP(hipGetLastError());
    if (half && s.sync_calls) check_fix_flag(s);
}

void launch_lj_energy(Sim& s, double rc, double eps, double sigma, double* part) {
    LJArgs a = make_lj_args(s, rc, eps, sigma);
    const int nb = s.tiles.ntx * s.tiles.nty * s.tiles.ntz;
    if (!s.mirror_fresh) refresh_mirror(s);
    if (s.lj_shift)
        hipLaunchKernelGGL(k_lj_energy<1>, dim3(nb), dim3(s.tile_threads), lds_bytes(s, true), s.stream, TMD_LIST_ARGS, a, s.nb_half, part);
    else
        hipLaunchKernelGGL(k_lj_energy<0>, dim3(nb), dim3(s.tile_threads), lds_bytes(s, true), s.stream, TMD_LIST_ARGS, a, s.nb_half, part);
    TMD_HIP(hipGetLastError());
}

void lj_energy_virial(Sim& s, double rc, double eps, double sigma, double* pe, double* vir, long long* pairs) {
    *pe = 0.0; *vir = 0.0; *pairs = 0;
    if (has_error() || !s.nb_valid || s.n == 0) return;
    const int nb = s.tiles.ntx * s.tiles.nty * s.tiles.ntz;
    s.red.reserve(sizeof(double) * (3 * (size_t)nb + 8), false, s.stream);
    double* part = s.red.as<double>();
    double* out = part + 3 * (size_t)nb;
    launch_lj_energy(s, rc, eps, sigma, part);
    reduce_partials(s, part, nb, 3, out);
    double host[3] = {0, 0, 0};
    TMD_HIP(hipMemcpyAsync(host, out, sizeof host, hipMemcpyDeviceToHost, s.stream));
    TMD_HIP(hipStreamSynchronize(s.stream));
    *pe = host[0];
    *vir = host[1];
    *pairs = (long long)host[2];
}

}  // namespace tmd
