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
// no atomics (core/compute.impala:14-45, full branch).
//
// Half lists (each pair once in the list) are evaluated by gather as well:
// the list build also writes each atom's complementary column (its local
// neighbours j > i), and the atom sums its half-list terms and its
// complementary terms.  A complementary term is, bit for bit, the negated
// reaction the reference scatters to i with "+=" (a data race on GPUs,
// core/compute.impala:104-111), so the result needs no atomics and does not
// depend on the order in which waves run.
//
// FP64 LJ follows the reference closure operation by operation
// (core/runtime.impala:188-201) in list order, so full-list FP64 forces are
// bit-identical to the CPU restatement.  FP32-mixed: the distance test stays
// FP64, on an rsq formed with FMAs (an ulp or so from the reference's
// rounding, so an in/out decision can differ from the FP64 path's only for a
// pair that close to rc^2); the LJ algebra is FP32 and
// dx * f is accumulated in FP64 (each FP32 term is exact there, and the
// extra mantissa bits absorb the summation rounding).  The Kahan variant
// accumulates in FP32 with a compensation term instead.
#include "tmd_internal.hpp"
#include "tmd_window.hpp"
#include "tmd_lj.hpp"
#include "tmd_pair.hpp"

namespace tmd {



#define TMD_KERNEL_LIST_PARAMS                                                                                   \
    const double4 *__restrict__ pos, const CMir spos, const int *__restrict__ start,                                \
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
//
// PAIR2 (half lists): the atom also gathers its complementary column (local
// neighbours j > i, whose pair terms the half list evaluates from j's side),
// continuing the same running sum.  Each term equals, bit for bit, the
// negated reaction the reference scatters to i (dx and dv change sign
// exactly), so only the summation order differs from the reference's.
template <int PREC, int SHIFT, int UNIT, int KICK, bool PAIR2>
__global__ void __launch_bounds__(kMaxTileThreads) k_lj_full(TMD_KERNEL_LIST_PARAMS, LJArgs a, double4* __restrict__ force,
                                                         double4* __restrict__ vel, const Mir svel, double hdt, double dt,
                                                         const int* __restrict__ wk, const Mir spos_next,
                                                         int vel_by_k, const int* __restrict__ nbc2,
                                                         const nb_t* __restrict__ nbl2) {
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
        // half lists: the complementary column continues the same sum
        const int cnt2 = PAIR2 ? nbc2[w] : 0;
        const int4* lp2 = PAIR2 ? reinterpret_cast<const int4*>(nbl2 + (size_t)(w >> 6) * (size_t)a.cap * 64) + (w & 63)
                                : nullptr;
        const double4 f = staged ? lj_full_atom<PREC, SHIFT, UNIT, true, PAIR2>(pi, cnt, lp, sp, spos, W, a, cnt2, lp2)
                                 : lj_full_atom<PREC, SHIFT, UNIT, false, PAIR2>(pi, cnt, lp, sp, spos, W, a, cnt2, lp2);
        if (KICK != 2) force[i] = f;
        if (KICK) {
            // velocity-Verlet final kick (k_final_integration's arithmetic)
            // velocities in atom order (vel) or in mirror order (svel)
            double4 v = vel_by_k ? (double4)svel[kw] : vel[i];
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
            if (vel_by_k) svel[kw] = v;
            else vel[i] = v;
        }
    }
}

// End of a DEM step: the final kick v += f dt/2 and, with ADV, the next
// step's kick and drift, with the mirror kept current when inv is given
// (k_final_integration and k_initial_integration arithmetic, in that order).
template <int ADV>
__global__ void k_dem_finish(int n, const double4* __restrict__ force, double4* __restrict__ vel, double hdt, double dt,
                             double4* __restrict__ pos, const int* __restrict__ inv, const Mir spos) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 f = force[i];
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
    const int tile = xcd_block(blockIdx.x, gridDim.x);
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
        partials[3 * tile + 0] = bpe;  // by tile: the fixed reduction order does not depend on the block mapping
        partials[3 * tile + 1] = bvir;
        partials[3 * tile + 2] = bnp;
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

// One contact term.  The caller decides contact (rsq < (2R)^2) in FP64 for
// both precisions, so FP32-mixed forces switch on and off at exactly the
// FP64 overlap: the dashpot makes the force jump by g_n v_n at first
// contact, and a contact decided on the rounded FP32 rsq would add or drop
// such a jump.  The overlap is passed as (2R)^2 - rsq, formed in FP64
// before any rounding to R; FP32 turns it into delta = ((2R)^2 - r^2) /
// (2R + r), since 2R - r in FP32 would lose a small overlap's leading
// digits.
//
// FP64 divides and takes square roots as written.  FP32 works from the
// hardware reciprocal square root and reciprocal (v_rsq_f32, v_rcp_f32, about
// 1 ulp each): 1/r once, then products, instead of four IEEE divisions and two
// square roots; the rounding this adds is a few 1e-8 relative, far inside
// the FP32-mixed tolerance, and it halves the contact term's instructions.
template <typename R>
__device__ __forceinline__ void dem_pair(const DEMArgs& a, R dx, R dy, R dz, R rsq, R dd, R vx, R vy, R vz, R& Fx, R& Fy,
                                         R& Fz) {
    const R d2 = (R)a.d2;
    R delta, nx, ny, nz;
    if constexpr (sizeof(R) == sizeof(double)) {
        const R r = sqrt(rsq);
        delta = d2 - r;  // the reference's 2R - r
        nx = dx / r; ny = dy / r; nz = dz / r;
    } else {
        const float ir = __builtin_amdgcn_rsqf(rsq);  // contact: rsq in (0, (2R)^2)
        delta = dd * __builtin_amdgcn_rcpf(d2 + rsq * ir);
        nx = dx * ir; ny = dy * ir; nz = dz * ir;
    }
    const R vn = vx * nx + vy * ny + vz * nz;
    const R tx = vx - vn * nx, ty = vy - vn * ny, tz = vz - vn * nz;
    const R vt2 = tx * tx + ty * ty + tz * tz;
    const R fn = (R)a.k * delta - (R)a.gn * vn;
    R inv;
    if constexpr (sizeof(R) == sizeof(double)) {
        const R vtl = sqrt(vt2);
        const R ft = fmin((R)a.gt * vtl, (R)a.mu * fabs(fn));
        inv = vtl > R(0) ? ft / vtl : R(0);
    } else {
        // |v_t| = vt2 / sqrt(vt2); below 1e-30 (and at 0) the tangential term is 0
        const float ivt = vt2 > 1e-30f ? __builtin_amdgcn_rsqf(vt2) : 0.0f;
        const float ft = fminf((float)a.gt * (vt2 * ivt), (float)a.mu * fabsf(fn));
        inv = ft * ivt;
    }
    Fx = fn * nx - inv * tx;
    Fy = fn * ny - inv * ty;
    Fz = fn * nz - inv * tz;
}

// DEM forces by gather.  PAIR2 (half lists): the atom also gathers its
// complementary column (local neighbours j > i); each such term is, bit for
// bit, the negated reaction the half list would send to i (dx and dv change
// sign exactly), so no scatter is needed.
template <int PREC, bool PAIR2>
__global__ void __launch_bounds__(kMaxTileThreads) k_dem(TMD_KERNEL_LIST_PARAMS, const double4* __restrict__ vel, DEMArgs a,
                                                     double4* __restrict__ force, const int* __restrict__ nbc2,
                                                     const nb_t* __restrict__ nbl2) {
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
            double fx = 0.0, fy = 0.0, fz = 0.0;
            auto entries = [&](const nb_t* __restrict__ L, int cnt) {
                // groups of 8 entries, one 16-byte streaming load each
                const int4* lp = reinterpret_cast<const int4*>(L + (size_t)(w >> 6) * (size_t)a.cap * 64) + (w & 63);
                int4 e = make_int4(0, 0, 0, 0);
                for (int k = 0; k < cnt; ++k) {
                    if ((k & 7) == 0) e = ld_stream(lp + (size_t)(k >> 3) * 64);
                    // entry k & 7 of the group: shift the group's words down
                    // as they are used (no indexed register access)
                    const int t = (k & 1) ? (int)((unsigned)e.x >> 16) : (e.x & 0xffff);
                    if (k & 1) e = make_int4(e.y, e.z, e.w, 0);
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
                    // a contact also needs rsq < (2R)^2, decided here in FP64
                    const double dd = a.d2 * a.d2 - rsq;
                    if (rsq < a.rc2 && dd > 0.0) {
                        const double4 vj = vel[j];
                        double Fx, Fy, Fz;
                        if (PREC == PREC_FP64) {
                            dem_pair<double>(a, dx, dy, dz, rsq, dd, vi.x - vj.x, vi.y - vj.y, vi.z - vj.z, Fx, Fy, Fz);
                        } else {
                            float gx, gy, gz;
                            dem_pair<float>(a, (float)dx, (float)dy, (float)dz, (float)rsq, (float)dd, (float)(vi.x - vj.x),
                                            (float)(vi.y - vj.y), (float)(vi.z - vj.z), gx, gy, gz);
                            Fx = gx; Fy = gy; Fz = gz;
                        }
                        fx = fx + Fx; fy = fy + Fy; fz = fz + Fz;
                    }
                }
            };
            entries(nbl, nbc[w]);
            if (PAIR2) entries(nbl2, nbc2[w]);
            force[i] = make_double4(fx, fy, fz, 0.0);
        }
    };
    if (staged) atoms(std::true_type{});
    else atoms(std::false_type{});
}

// ---------------------------------------------------------------------------
// DEM step loop with the state in the cell-ordered mirrors (spos, svel,
// ghosts included).  The block stages its window's positions and
// velocities into LDS (48 bytes per atom), so a contact's velocity is an LDS
// read rather than a global gather; the lane reads its own atom at its mirror
// index.  With ADV the pass ends the step (final kick v += f dt/2) and runs
// the next step's kick and drift, writing the new position and velocity
// into the other mirrors (spos_next, svel_next; other blocks still read the
// current ones); without it (the last step) it stores the force, the
// velocity and the position in atom order.  The pair arithmetic, the
// contact decisions and the summation order are k_dem's, and the kicks are
// k_dem_finish's, so the trajectory is that of the separate passes.
static int dem_window_cap(const Sim& s) { return std::min(s.max_window + 1, kLdsBudget / kDemStageBytes); }

template <int PREC, bool PAIR2, int ADV>
__global__ void __launch_bounds__(kMaxTileThreads) k_dem_fused(const CMir spos, const CMir svel,
                                                           const int* __restrict__ start, const int* __restrict__ work,
                                                           const int* __restrict__ wk, const int* __restrict__ tile_start,
                                                           const int* __restrict__ nbc, const nb_t* __restrict__ nbl,
                                                           const int* __restrict__ nbc2, const nb_t* __restrict__ nbl2,
                                                           DEMArgs a, double hdt, double dt, const Mir spos_next,
                                                           const Mir svel_next, double4* __restrict__ force,
                                                           double4* __restrict__ vel, double4* __restrict__ pos,
                                                           const int* __restrict__ gh_start, const int* __restrict__ gh_k,
                                                           const int* __restrict__ gh_c, double L0, double L1, double L2) {
    extern __shared__ double dyn[];
    const int tile = xcd_block(blockIdx.x, gridDim.x);
    const int w0 = tile_start[tile], nloc = tile_start[tile + 1] - w0;
    if (nloc == 0) return;
    __shared__ TileWin W;
    tile_window(tile, start, a.g, W);
    double* sx = dyn;
    double* sy = dyn + a.wcap;
    double* sz = dyn + 2 * a.wcap;
    double* su = dyn + 3 * a.wcap;
    double* sv = dyn + 4 * a.wcap;
    double* sw = dyn + 5 * a.wcap;
    const bool staged = win_size(W) <= a.wcap;
    if (staged) {
        // one thread group per window row, as in stage_window
        const int nseg = W.nseg;
        const int gsz = max(1, (int)blockDim.x / nseg);
        const int ngrp = (int)blockDim.x / gsz;
        const int g = (int)threadIdx.x / gsz, l = (int)threadIdx.x - g * gsz;
        if (g < ngrp) {
            for (int sg = g; sg < nseg; sg += ngrp) {
                const int off = W.seg_off[sg], beg = W.seg_beg[sg], len = W.seg_off[sg + 1] - off;
                for (int q = l; q < len; q += gsz) {
                    const int t = off + q, k = beg + q;
                    const double4 p = spos[k];
                    const double4 u = svel[k];
                    sx[t] = p.x; sy[t] = p.y; sz[t] = p.z;
                    su[t] = u.x; sv[t] = u.y; sw[t] = u.z;
                }
            }
        }
    }
    __syncthreads();
    auto atoms = [&](auto staged_c) {
        constexpr bool STAGED = decltype(staged_c)::value;
        for (int at = threadIdx.x; at < nloc; at += blockDim.x) {
            const int w = w0 + at;
            const int kw = wk[w];
            const double4 pi = spos[kw];
            const double4 vi = svel[kw];
            double fx = 0.0, fy = 0.0, fz = 0.0;
            auto entries = [&](const nb_t* __restrict__ L, int cnt) {
                const int4* lp = reinterpret_cast<const int4*>(L + (size_t)(w >> 6) * (size_t)a.cap * 64) + (w & 63);
                int4 e = make_int4(0, 0, 0, 0);
                for (int k = 0; k < cnt; ++k) {
                    if ((k & 7) == 0) e = ld_stream(lp + (size_t)(k >> 3) * 64);
                    const int t = (k & 1) ? (int)((unsigned)e.x >> 16) : (e.x & 0xffff);
                    if (k & 1) e = make_int4(e.y, e.z, e.w, 0);
                    double xj, yj, zj, uj, vj, wj;
                    if (STAGED) {
                        xj = sx[t]; yj = sy[t]; zj = sz[t];
                        uj = su[t]; vj = sv[t]; wj = sw[t];
                    } else {
                        const int kk = win_index(t, W);
                        const double4 p = spos[kk];
                        const double4 u = svel[kk];
                        xj = p.x; yj = p.y; zj = p.z;
                        uj = u.x; vj = u.y; wj = u.z;
                    }
                    const double dx = pi.x - xj, dy = pi.y - yj, dz = pi.z - zj;
                    const double rsq = dx * dx + dy * dy + dz * dz;
                    const double dd = a.d2 * a.d2 - rsq;
                    if (rsq < a.rc2 && dd > 0.0) {
                        double Fx, Fy, Fz;
                        if (PREC == PREC_FP64) {
                            dem_pair<double>(a, dx, dy, dz, rsq, dd, vi.x - uj, vi.y - vj, vi.z - wj, Fx, Fy, Fz);
                        } else {
                            float gx, gy, gz;
                            dem_pair<float>(a, (float)dx, (float)dy, (float)dz, (float)rsq, (float)dd, (float)(vi.x - uj),
                                            (float)(vi.y - vj), (float)(vi.z - wj), gx, gy, gz);
                            Fx = gx; Fy = gy; Fz = gz;
                        }
                        fx = fx + Fx; fy = fy + Fy; fz = fz + Fz;
                    }
                }
            };
            entries(nbl, nbc[w]);
            if (PAIR2) entries(nbl2, nbc2[w]);
            // k_dem_finish's arithmetic
            double4 v = vi;
            v.x = v.x + fx * hdt;
            v.y = v.y + fy * hdt;
            v.z = v.z + fz * hdt;
            if (ADV) {
                v.x = v.x + fx * hdt;
                v.y = v.y + fy * hdt;
                v.z = v.z + fz * hdt;
                double4 x = pi;
                x.x = x.x + v.x * dt;
                x.y = x.y + v.y * dt;
                x.z = x.z + v.z * dt;
                spos_next[kw] = x;
                svel_next[kw] = v;
                // the sphere's ghosts: its new position + image shift and its
                // new velocity (k_ghost_step_mirror_v's arithmetic)
                if (gh_start) {
                    const int e1 = gh_start[w + 1];
                    for (int e = gh_start[w]; e < e1; ++e) {
                        int ix, iy, iz;
                        img_decode(gh_c[e], ix, iy, iz);
                        double4 q;
                        q.x = x.x + L0 * (double)ix;
                        q.y = x.y + L1 * (double)iy;
                        q.z = x.z + L2 * (double)iz;
                        q.w = x.w;
                        const int kg = gh_k[e];
                        spos_next[kg] = q;
                        svel_next[kg] = v;
                    }
                }
            } else {
                const int i = work[w];
                force[i] = make_double4(fx, fy, fz, 0.0);
                vel[i] = v;
                pos[i] = pi;
            }
        }
    };
    if (staged) atoms(std::true_type{});
    else atoms(std::false_type{});
}

void compute_dem_fused(Sim& s, int half, double rc, double gn, double gt, double k, double mu, double dt, bool advance) {
    if (has_error()) return;
    TMD_REQUIRE(s.nb_valid && s.dem_mirrored, "DEM step loop: the state is not in the mirrors");
    TMD_REQUIRE(half == s.nb_half, "compute forces: half flag differs from the one the neighbour list was built with");
    if (has_error() || s.n == 0) return;
    const double radius = s.dem_radius > 0.0 ? s.dem_radius : 0.5 * rc;
    s.last_force_rc = rc;
    DEMArgs a{rc * rc, 2.0 * radius, gn, gt, k, mu, s.nb_cap, s.n, s.n, dem_window_cap(s), s.tiles};
    const int nb = s.tiles.ntx * s.tiles.nty * s.tiles.ntz;
    const size_t lds = (size_t)a.wcap * kDemStageBytes;
    const double hdt = 0.5 * dt;
    // advancing: the pass also writes the ghosts' next mirror entries
    // (from each local's ghost index), so the next step needs no ghost pass
    if (advance && s.ng > 0 && s.ghost_mode == GHOST_FRESH && !(s.gsrc_valid && s.gh_valid)) build_ghost_index(s);
    const bool fuse = advance && s.ng > 0 && s.ghost_mode == GHOST_FRESH && s.gsrc_valid && s.gh_valid && !has_error();
    const int* GS = fuse ? s.gh_start.as<int>() : nullptr;
    hipEvent_t ev = nullptr;
    s.timer.begin(K_FORCE, s.stream, &ev);
#define TMD_DEMF(P, H, A)                                                                                               \
    hipLaunchKernelGGL((k_dem_fused<P, H, A>), dim3(nb), dim3(s.tile_threads), lds, s.stream,                            \
                       CMir{s.spos.as<double>()}, CMir{s.svel.as<double>()},                                           \
                       (const int*)s.cell_start.as<int>(), (const int*)s.work.as<int>(), (const int*)s.work_k.as<int>(), \
                       (const int*)s.tile_start.as<int>(), (const int*)s.nb_count.as<int>(), (const nb_t*)s.nb_list.as<nb_t>(), \
                       (const int*)s.nb_count2.as<int>(), (const nb_t*)s.nb_list2.as<nb_t>(), a, hdt, dt,               \
                       Mir{s.spos2.as<double>()}, Mir{s.svel2.as<double>()}, s.force.as<double4>(), s.vel.as<double4>(),    \
                       s.pos.as<double4>(), GS, (const int*)s.gh_k.as<int>(), (const int*)s.gh_c.as<int>(), s.g.len[0],  \
                       s.g.len[1], s.g.len[2])
#define TMD_DEMF_A(P, H) do { if (advance) TMD_DEMF(P, H, 1); else TMD_DEMF(P, H, 0); } while (0)
#define TMD_DEMF_H(P) do { if (half) TMD_DEMF_A(P, true); else TMD_DEMF_A(P, false); } while (0)
    if (s.precision == PREC_FP64) TMD_DEMF_H(PREC_FP64);
    else TMD_DEMF_H(PREC_FP32_MIXED);
#undef TMD_DEMF_H
#undef TMD_DEMF_A
#undef TMD_DEMF
    s.timer.end(K_FORCE, s.stream, ev);
    TMD_HIP(hipGetLastError());
    if (advance) {
        std::swap(s.spos, s.spos2);
        std::swap(s.svel, s.svel2);
        s.dem_ghosts_done = fuse || s.ng == 0;
    } else {
        // the locals are back in atom order; their ghosts follow from them
        s.dem_mirrored = 0;
        refresh_ghosts(s);
    }
}

#define TMD_LIST_ARGS                                                                                              \
    (const double4*)s.pos.as<double4>(), CMir{s.spos.as<double>()}, (const int*)s.cell_start.as<int>(),           \
        (const int*)s.cell_parts.as<int>(), (const int*)s.work.as<int>(), (const int*)s.tile_start.as<int>(),     \
        (const int*)s.nb_count.as<int>(), (const nb_t*)s.nb_list.as<nb_t>()

static size_t lds_bytes(const Sim& s, bool) { return (size_t)window_cap(s) * kStageBytes; }

void compute_lj(Sim& s, int half, double rc, double eps, double sigma, double kick_dt, bool advance) {
    if (has_error()) return;
    TMD_REQUIRE(s.nb_valid, "compute forces: no neighbour list (call md_assemble_neighborlists first)");
    TMD_REQUIRE(half == s.nb_half, "compute forces: half flag differs from the one the neighbour list was built with");
    if (has_error() || s.n == 0) return;
    LJArgs a = make_lj_args(s, rc, eps, sigma);
    a.wcap = window_cap_pos(s);  // the force pass stages positions only
    const int nb = s.tiles.ntx * s.tiles.nty * s.tiles.ntz;
    const int shift = s.lj_shift;
    s.last_force_rc = rc;
    hipEvent_t ev = nullptr;
    s.timer.begin(K_MIRROR, s.stream, &ev);
    if (!s.mirror_fresh) refresh_mirror(s);
    s.timer.end(K_MIRROR, s.stream, ev);
    s.timer.begin(K_FORCE, s.stream, &ev);
    {
        double4* F = s.force.as<double4>();
        const size_t lds = (size_t)a.wcap * kPosStageBytes;
        const int unit = (eps == 1.0 && a.sigma6 == 1.0) ? 1 : 0;
        const double hdt = 0.5 * kick_dt;
        // inside the fused step loop the velocities live in mirror order
        double4* V = s.vel.as<double4>();
        const Mir SV{s.svel.as<double>()};  // used when the velocities live in the mirror (VK)
        const int VK = s.vel_mirrored;
        const int* WK = s.work_k.as<int>();
        const Mir SN{s.spos2.as<double>()};
        const int* C2 = s.nb_count2.as<int>();
        const nb_t* L2 = s.nb_list2.as<nb_t>();
#define TMD_LJ_K(P, SH, U, K, H)                                                                                       \
    hipLaunchKernelGGL((k_lj_full<P, SH, U, K, H>), dim3(nb), dim3(s.tile_threads), lds, s.stream, TMD_LIST_ARGS, a, F, V, SV, hdt, \
                       kick_dt, WK, SN, VK, C2, L2)
#define TMD_LJ_H(P, SH, U, K)                                                                                          \
    do { if (half) TMD_LJ_K(P, SH, U, K, true); else TMD_LJ_K(P, SH, U, K, false); } while (0)
#define TMD_LJ_FULL(P, SH, U)                                                                                          \
    do {                                                                                                               \
        if (kick_dt > 0.0 && advance) TMD_LJ_H(P, SH, U, 2);                                                           \
        else if (kick_dt > 0.0) TMD_LJ_H(P, SH, U, 1);                                                                 \
        else TMD_LJ_H(P, SH, U, 0);                                                                                    \
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
#undef TMD_LJ_H
#undef TMD_LJ_K
    }
    s.timer.end(K_FORCE, s.stream, ev);
    TMD_HIP(hipGetLastError());
    if (kick_dt > 0.0 && advance) {
        // the new local positions are in spos2: it becomes the mirror, whose
        // ghost entries step_refresh rewrites
        std::swap(s.spos, s.spos2);
        s.mirror_fresh = 0;
        s.ghost_stale = 1;
        s.pos_stale = 1;
    }
}

void compute_dem(Sim& s, int half, double rc, double gn, double gt, double k, double mu, double kick_dt, bool advance) {
    if (has_error()) return;
    TMD_REQUIRE(s.nb_valid, "compute forces: no neighbour list (call md_assemble_neighborlists first)");
    TMD_REQUIRE(half == s.nb_half, "compute forces: half flag differs from the one the neighbour list was built with");
    if (has_error() || s.n == 0) return;
    const double radius = s.dem_radius > 0.0 ? s.dem_radius : 0.5 * rc;
    s.last_force_rc = rc;
    DEMArgs a{rc * rc, 2.0 * radius, gn, gt, k, mu, s.nb_cap, s.n, s.n, window_cap(s), s.tiles};
    const int nb = s.tiles.ntx * s.tiles.nty * s.tiles.ntz;
    const double4* V = s.vel.as<double4>();
    double4* F = s.force.as<double4>();
    const int* C2 = s.nb_count2.as<int>();
    const nb_t* L2 = s.nb_list2.as<nb_t>();
    const size_t lds = lds_bytes(s, true);
    hipEvent_t ev = nullptr;
    s.timer.begin(K_MIRROR, s.stream, &ev);
    if (!s.mirror_fresh) refresh_mirror(s);
    s.timer.end(K_MIRROR, s.stream, ev);
    s.timer.begin(K_FORCE, s.stream, &ev);
#define TMD_DEM(P, H) hipLaunchKernelGGL((k_dem<P, H>), dim3(nb), dim3(s.tile_threads), lds, s.stream, TMD_LIST_ARGS, V, a, F, C2, L2)
    if (s.precision == PREC_FP64) {
        if (half) TMD_DEM(PREC_FP64, true); else TMD_DEM(PREC_FP64, false);
    } else {
        if (half) TMD_DEM(PREC_FP32_MIXED, true); else TMD_DEM(PREC_FP32_MIXED, false);
    }
#undef TMD_DEM
    if (kick_dt > 0.0) {
        // final kick (and, advancing, the next kick and drift) fused
        double4* V2 = s.vel.as<double4>();
        double4* X = s.pos.as<double4>();
        const int* I = s.inv_valid ? (const int*)s.cell_inv.as<int>() : nullptr;
        const Mir M{s.spos.as<double>()};
        const double hdt = 0.5 * kick_dt;
        const dim3 g(div_up(s.n, 256)), b(256);
        if (advance) hipLaunchKernelGGL(k_dem_finish<1>, g, b, 0, s.stream, s.n, (const double4*)F, V2, hdt, kick_dt, X, I, M);
        else hipLaunchKernelGGL(k_dem_finish<0>, g, b, 0, s.stream, s.n, (const double4*)F, V2, hdt, kick_dt, X, I, M);
        if (advance && !s.inv_valid) mirror_invalidate(s);
    }
    s.timer.end(K_FORCE, s.stream, ev);
    TMD_HIP(hipGetLastError());
}

void launch_lj_energy(Sim& s, double rc, double eps, double sigma, double* part, bool eshift) {
    LJArgs a = make_lj_args(s, rc, eps, sigma);
    if (eshift && !s.lj_shift) a.eshift = a.uc;
    const int nb = s.tiles.ntx * s.tiles.nty * s.tiles.ntz;
    if (!s.mirror_fresh) refresh_mirror(s);
    if (s.lj_shift)
        hipLaunchKernelGGL(k_lj_energy<1>, dim3(nb), dim3(s.tile_threads), lds_bytes(s, true), s.stream, TMD_LIST_ARGS, a, s.nb_half, part);
    else
        hipLaunchKernelGGL(k_lj_energy<0>, dim3(nb), dim3(s.tile_threads), lds_bytes(s, true), s.stream, TMD_LIST_ARGS, a, s.nb_half, part);
    TMD_HIP(hipGetLastError());
}

void lj_energy_virial(Sim& s, double rc, double eps, double sigma, double* pe, double* vir, long long* pairs, bool eshift,
                      double* buf) {
    *pe = 0.0; *vir = 0.0; *pairs = 0;
    if (has_error() || !s.nb_valid || s.n == 0) return;
    const int nb = s.tiles.ntx * s.tiles.nty * s.tiles.ntz;
    if (!buf) s.red.reserve(sizeof(double) * (3 * (size_t)nb + 8), false, s.stream);
    if (has_error()) return;
    double* part = buf ? buf : s.red.as<double>();
    double* out = part + 3 * (size_t)nb;
    launch_lj_energy(s, rc, eps, sigma, part, eshift);
    reduce_partials(s, part, nb, 3, out);
    double host[3] = {0, 0, 0};
    TMD_HIP(hipMemcpyAsync(host, out, sizeof host, hipMemcpyDeviceToHost, s.stream));
    TMD_HIP(hipStreamSynchronize(s.stream));
    *pe = host[0];
    *vir = host[1];
    *pairs = (long long)host[2];
}

}  // namespace tmd

// Host evaluation of the compensated sum that the FP32-Kahan pair kernels
// use (kahan_add / kahan_value, the same inline functions), so that the
// finalisation can be checked without a GPU.
extern "C" double tmd_kahan_sum(const float* x, int n) {
    float s = 0.f, c = 0.f;
    for (int k = 0; k < n; ++k) tmd::kahan_add(s, c, x[k]);
    return tmd::kahan_value(s, c);
}
