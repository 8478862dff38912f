// Differentiable LJ trajectories: forward loss and its hand-written adjoint.
//
// Loss of a velocity-Verlet trajectory of T steps from (x0, v0):
//   L = (Ubar - U*)^2 + (Pbar - P*)^2,
//   Ubar = (1/T) sum_{t=1..T} PE(x_t) / N,
//   Pbar = (1/T) sum_{t=1..T} (sum v_t^2 + W(x_t)) / (3V),   W = sum_pairs r.F.
// Gradients with respect to eps, sigma and x0.  The pair potential is the
// smoothly truncated LJ of tmd_lj.hpp (force and its derivative vanish at
// rc), so the discrete trajectory is differentiable in its parameters.
//
// Reverse pass: each forward step is
//   (a) v += h F, (b) x += dt v, (c) F = F(x), (d) v += h F, (e) loss terms,
// h = dt/2, and is undone as
//   (e) vbar += dP/dv, xbar/theta += loss terms at x_t
//   (d) Fbar += h vbar
//   (c) xbar += J_F(x_t)^T Fbar, theta += (dF/dtheta)^T Fbar, Fbar = 0
//   (b) vbar += dt xbar
//   (a) Fbar += h vbar.
// J_F^T Fbar is a Hessian-vector product over the full neighbour list:
//   xbar_i += sum_j [f(s) d + 2 f'(s) (r.d) r],  d = Fbar_i - Fbar_owner(j),
// symmetric in (i, j), so each lane owns its atom (no atomics); ghost
// adjoints are their owners'.  (x, v) is checkpointed every rebuild interval
// and each interval is recomputed forward before it is reversed; within an
// interval the neighbour list and ghost table are fixed, as in the forward.
#include "tmd_internal.hpp"
#include "tmd_lj.hpp"

namespace tmd {

struct AdjCoef {
    double cU;   // dL/dPE_t  = 2 (Ubar - U*) / (T N)
    double cW;   // dL/dW_t   = 2 (Pbar - P*) / (3 V T)
    // sigma-derivatives of the truncation constants
    double dfc_ds, ddfc_ds, duc_ds;
    int loss_terms;  // 0 at x_0 (the loss starts at t = 1)
};

template <int PREC, int SHIFT>
__global__ void __launch_bounds__(kMaxTileThreads) k_lj_adjoint(const double4* __restrict__ pos, const CMir spos,
                                                             const int* __restrict__ start, const int* __restrict__ parts,
                                                             const int* __restrict__ work, const int* __restrict__ tile_start,
                                                             const int* __restrict__ nbc, const nb_t* __restrict__ nbl,
                                                             const int* __restrict__ g_owner, const double4* __restrict__ fbar,
                                                             LJArgs a, AdjCoef c, double4* __restrict__ xbar,
                                                             double* __restrict__ partials) {
    extern __shared__ double dyn[];
    const int tile = xcd_block(blockIdx.x, gridDim.x);
    const int w0 = tile_start[tile], nloc = tile_start[tile + 1] - w0;
    if (nloc == 0) {
        if (threadIdx.x == 0) {
            partials[2 * tile + 0] = 0.0;
            partials[2 * tile + 1] = 0.0;
        }
        return;
    }
    __shared__ TileWin W;
    tile_window(tile, start, a.g, W);
    double* sx = dyn;
    double* sy = dyn + a.wcap;
    double* sz = dyn + 2 * a.wcap;
    int* sj = reinterpret_cast<int*>(dyn + 3 * a.wcap);
    const bool staged = stage_tile(W, a.wcap, spos, parts, sx, sy, sz, sj);
    const double inv_eps = 1.0 / a.eps, inv_sig = 1.0 / a.sigma;
    double g_eps = 0.0, g_sig = 0.0;
    // STAGED as a compile-time constant: LDS reads, not generic-pointer loads
    auto atoms = [&](auto staged_c) {
    constexpr bool STAGED = decltype(staged_c)::value;
    for (int at = threadIdx.x; at < nloc; at += blockDim.x) {
        const int w = w0 + at;
        const int i = work[w];
        const double4 pi = pos[i];
        const double4 fbi = fbar[i];
        const int cnt = nbc[w];
        double ax = 0.0, ay = 0.0, az = 0.0;
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
            const double rx = pi.x - xj, ry = pi.y - yj, rz = pi.z - zj;
            const double s = rx * rx + ry * ry + rz * rz;
            if (!(s < a.rc2)) continue;
            // f(s), f'(s), u(s), d f/d sigma, d u/d sigma in FP64.  FP64 takes
            // 1/s from the IEEE division (the forward kernel's rounding);
            // FP32-mixed from the FP32 reciprocal refined by one FP64 Newton
            // step, whose relative error is about the square of the FP32 one
            // (~1e-14), at the cost of the FP32 path
            double sr2;
            if (PREC == PREC_FP64) {
                sr2 = 1.0 / s;
            } else {
                sr2 = (double)__builtin_amdgcn_rcpf((float)s);
                sr2 = fma(sr2, fma(-s, sr2, 1.0), sr2);
            }
            double f, df, u, dfds, duds;
            {
                const double sr6 = sr2 * sr2 * sr2 * a.sigma6;
                f = 48.0 * a.eps * sr6 * (sr6 - 0.5) * sr2;
                df = 24.0 * a.eps * sr2 * sr2 * sr6 * (4.0 - 14.0 * sr6);
                u = 4.0 * a.eps * sr6 * (sr6 - 1.0);
                dfds = 144.0 * a.eps * inv_sig * sr2 * sr6 * (4.0 * sr6 - 1.0);
                duds = 24.0 * a.eps * inv_sig * sr6 * (2.0 * sr6 - 1.0);
            }
            if (SHIFT) {
                const double ds = s - a.rc2;
                f = f - a.fc - ds * a.dfc;
                df = df - a.dfc;
                u = u - a.uc + ds * (0.5 * a.fc) + ds * ds * (0.25 * a.dfc);
                dfds = dfds - c.dfc_ds - ds * c.ddfc_ds;
                duds = duds - c.duc_ds + ds * (0.5 * c.dfc_ds) + ds * ds * (0.25 * c.ddfc_ds);
            }
            // J_F^T Fbar
            const int o = j < a.n ? j : g_owner[j - a.n];
            const double4 fbj = fbar[o];
            const double dx = fbi.x - fbj.x, dy = fbi.y - fbj.y, dz = fbi.z - fbj.z;
            const double rd = rx * dx + ry * dy + rz * dz;
            double cr = 2.0 * df * rd;  // coefficient of r
            ax += f * dx;
            ay += f * dy;
            az += f * dz;
            // loss: dPE/dx_i = -sum f r ; dW/dx_i = sum 2 (f + s f') r
            if (c.loss_terms) cr += -c.cU * f + c.cW * 2.0 * (f + s * df);
            ax += cr * rx;
            ay += cr * ry;
            az += cr * rz;
            // parameters: Fbar_i . dF_i/dtheta over ordered pairs; loss halves
            const double rf = rx * fbi.x + ry * fbi.y + rz * fbi.z;
            double ge = f * inv_eps * rf;
            double gs = dfds * rf;
            if (c.loss_terms) {
                ge += 0.5 * (c.cU * u * inv_eps + c.cW * s * f * inv_eps);
                gs += 0.5 * (c.cU * duds + c.cW * s * dfds);
            }
            g_eps += ge;
            g_sig += gs;
        }
        double4 xb = xbar[i];
        xb.x += ax;
        xb.y += ay;
        xb.z += az;
        xbar[i] = xb;
    }
    };
    if (staged) atoms(std::true_type{});
    else atoms(std::false_type{});
    const double be = block_sum_double_any(g_eps);
    const double bs = block_sum_double_any(g_sig);
    if (threadIdx.x == 0) {
        partials[2 * tile + 0] = be;
        partials[2 * tile + 1] = bs;
    }
}

// acc[c] += sum_b partials[width * b + c] (one block, fixed order)
__global__ void __launch_bounds__(256) k_accumulate(const double* __restrict__ partials, int nblocks, int width,
                                                    double* __restrict__ acc) {
    for (int c = 0; c < width; ++c) {
        double v = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += 256) v += partials[(size_t)width * b + c];
        const double t = block_sum_double<256>(v);
        if (threadIdx.x == 0) acc[c] += t;
    }
}

// (e) + (d): vbar += cv * v_t ; Fbar += h vbar
__global__ void k_adj_pre(const double4* __restrict__ vel, int n, double cv, double h, double4* __restrict__ vbar,
                          double4* __restrict__ fbar) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 vb = vbar[i];
    const double4 v = vel[i];
    vb.x += cv * v.x; vb.y += cv * v.y; vb.z += cv * v.z;
    vbar[i] = vb;
    double4 fb = fbar[i];
    fb.x += h * vb.x; fb.y += h * vb.y; fb.z += h * vb.z;
    fbar[i] = fb;
}

// (c) Fbar = 0, (b) vbar += dt xbar, (a) Fbar = h vbar
__global__ void k_adj_post(const double4* __restrict__ xbar, int n, double dt, double h, double4* __restrict__ vbar,
                           double4* __restrict__ fbar) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 vb = vbar[i];
    const double4 xb = xbar[i];
    vb.x += dt * xb.x; vb.y += dt * xb.y; vb.z += dt * xb.z;
    vbar[i] = vb;
    fbar[i] = make_double4(h * vb.x, h * vb.y, h * vb.z, 0.0);
}

// ---------------------------------------------------------------------------
static void copy_state(Sim& s, double4* xd, double4* vd, const double4* xs, const double4* vs) {
    const size_t bytes = sizeof(double4) * (size_t)s.n;
    TMD_HIP(hipMemcpyAsync(xd, xs, bytes, hipMemcpyDeviceToDevice, s.stream));
    TMD_HIP(hipMemcpyAsync(vd, vs, bytes, hipMemcpyDeviceToDevice, s.stream));
}

static void snapshot_to(Sim& s, DevBuf& bx, DevBuf& bv, size_t slot) {
    copy_state(s, bx.as<double4>() + slot * s.n, bv.as<double4>() + slot * s.n, s.pos.as<double4>(), s.vel.as<double4>());
}

static void restore_from(Sim& s, const DevBuf& bx, const DevBuf& bv, size_t slot) {
    mirror_invalidate(s);
    copy_state(s, s.pos.as<double4>(), s.vel.as<double4>(), bx.as<double4>() + slot * s.n, bv.as<double4>() + slot * s.n);
}

// ghosts, cells and full list for the current positions (no periodic wrap)
static void relist(Sim& s, double cutoff) {
    s.ng = 0;
    build_ghosts(s);
    bin_particles(s);
    build_neighbor_lists(s, 0, cutoff);
}

static int ntiles_of(const Sim& s) { return s.tiles.ntx * s.tiles.nty * s.tiles.ntz; }

static double* partials(Sim& s, size_t count) {
    s.adj_part.reserve(sizeof(double) * std::max<size_t>(count, 1), false, s.stream);
    return s.adj_part.as<double>();
}

// acc[0..3] += {PE, W, pairs, sum v^2} of the current state
static void accumulate_thermo(Sim& s, double rc, double eps, double sigma, double* acc) {
    const int nt = ntiles_of(s), nkb = div_up(s.n, 256);
    double* part = partials(s, 3 * (size_t)std::max(nt, nkb));
    launch_lj_energy(s, rc, eps, sigma, part);
    hipLaunchKernelGGL(k_accumulate, dim3(1), dim3(256), 0, s.stream, (const double*)part, nt, 3, acc);
    launch_kinetic(s, part);
    hipLaunchKernelGGL(k_accumulate, dim3(1), dim3(256), 0, s.stream, (const double*)part, nkb, 1, acc + 3);
    TMD_HIP(hipGetLastError());
}

static void launch_lj_adjoint(Sim& s, double rc, double eps, double sigma, const AdjCoef& c, double* theta) {
    LJArgs a = make_lj_args(s, rc, eps, sigma);
    if (!s.mirror_fresh) refresh_mirror(s);
    const int ntiles = ntiles_of(s);
    const size_t lds = (size_t)a.wcap * kStageBytes;
    double* part = partials(s, 2 * (size_t)ntiles);
#define TMD_ADJ(P, SH)                                                                                                  \
    hipLaunchKernelGGL((k_lj_adjoint<P, SH>), dim3(ntiles), dim3(s.tile_threads), lds, s.stream,                          \
                       (const double4*)s.pos.as<double4>(), CMir{s.spos.as<double>()},                       \
                       (const int*)s.cell_start.as<int>(), (const int*)s.cell_parts.as<int>(), (const int*)s.work.as<int>(), \
                       (const int*)s.tile_start.as<int>(), (const int*)s.nb_count.as<int>(),                            \
                       (const nb_t*)s.nb_list.as<nb_t>(), (const int*)s.g_owner.as<int>(),                             \
                       (const double4*)s.adj_fbar.as<double4>(), a, c, s.adj_xbar.as<double4>(), part)
    // the adjoint kernel's precision: the forward's unless set apart
    const int prec = s.adj_precision >= 0 ? s.adj_precision : s.precision;
    if (prec == PREC_FP64) {
        if (s.lj_shift) TMD_ADJ(PREC_FP64, 1); else TMD_ADJ(PREC_FP64, 0);
    } else {
        if (s.lj_shift) TMD_ADJ(PREC_FP32_MIXED, 1); else TMD_ADJ(PREC_FP32_MIXED, 0);
    }
#undef TMD_ADJ
    hipLaunchKernelGGL(k_accumulate, dim3(1), dim3(256), 0, s.stream, (const double*)part, ntiles, 2, theta);
    TMD_HIP(hipGetLastError());
}

struct TrajSpec {
    int steps, ck;
    double dt, rc, skin, eps, sigma;
};

// Forward trajectory from (adj_x0, adj_v0), reneighbouring every ck steps
// exactly as the step loop does (wrap, ghosts, cells, list after the
// position update of steps ck, 2ck, ...).  With checkpoints, slot k holds
// the state at t = k ck (after that step's wrap).
static void forward_traj(Sim& s, const TrajSpec& p, bool store_ckpt, double* ubar, double* pbar) {
    const double cut = p.rc + p.skin;
    double* acc = s.adj_acc.as<double>();
    restore_from(s, s.adj_x0, s.adj_v0, 0);
    relist(s, cut);
    compute_lj(s, 0, p.rc, p.eps, p.sigma);
    if (store_ckpt) snapshot_to(s, s.adj_ckx, s.adj_ckv, 0);
    TMD_HIP(hipMemsetAsync(acc, 0, 4 * sizeof(double), s.stream));
    for (int t = 1; t <= p.steps && !has_error(); ++t) {
        initial_integration(s, p.dt);
        if (t % p.ck == 0) {
            s.ng = 0;
            build_ghosts(s, true);  // md_enforce_pbc + borders
            bin_particles(s);
            build_neighbor_lists(s, 0, cut);
        } else {
            refresh_ghosts(s);
        }
        compute_lj(s, 0, p.rc, p.eps, p.sigma, p.dt);  // final kick fused
        accumulate_thermo(s, p.rc, p.eps, p.sigma, acc);
        if (store_ckpt && t % p.ck == 0) snapshot_to(s, s.adj_ckx, s.adj_ckv, (size_t)(t / p.ck));
    }
    double h[4] = {0, 0, 0, 0};
    TMD_HIP(hipMemcpyAsync(h, acc, sizeof h, hipMemcpyDeviceToHost, s.stream));
    TMD_HIP(hipStreamSynchronize(s.stream));
    const double vol = s.g.len[0] * s.g.len[1] * s.g.len[2];
    *ubar = h[0] / (double)s.n / (double)p.steps;
    *pbar = (h[3] + h[1]) / (3.0 * vol) / (double)p.steps;
}

static bool begin_traj(Sim& s, const TrajSpec& p, bool adjoint) {
    if (p.steps <= 0 || p.ck <= 0) {
        set_error(__FILE__, __LINE__, "trajectory: steps and the reneighbouring interval must be positive");
        return false;
    }
    if (s.ghost_mode != GHOST_FRESH) {
        set_error(__FILE__, __LINE__, "trajectory: needs fresh ghost positions (tmd_set_ghost_mode(0))");
        return false;
    }
    if (s.n == 0) {
        set_error(__FILE__, __LINE__, "trajectory: no atoms");
        return false;
    }
    const size_t atom = sizeof(double4) * (size_t)s.n;
    s.adj_x0.reserve(atom, false, s.stream);
    s.adj_v0.reserve(atom, false, s.stream);
    s.adj_acc.reserve(8 * sizeof(double), false, s.stream);
    if (adjoint) {
        const size_t nck = (size_t)(p.steps / p.ck + 1);
        s.adj_ckx.reserve(atom * nck, false, s.stream);
        s.adj_ckv.reserve(atom * nck, false, s.stream);
        const size_t nseg = (size_t)std::min(p.ck, p.steps + 1);  // states per interval
        s.adj_segx.reserve(atom * nseg, false, s.stream);
        s.adj_segv.reserve(atom * nseg, false, s.stream);
        s.adj_xbar.reserve(atom, false, s.stream);
        s.adj_vbar.reserve(atom, false, s.stream);
        s.adj_fbar.reserve(atom, false, s.stream);
    }
    if (has_error()) return false;
    snapshot_to(s, s.adj_x0, s.adj_v0, 0);
    return !has_error();
}

// put the initial state back, with its ghosts, list and forces
static void end_traj(Sim& s, const TrajSpec& p) {
    if (has_error()) return;
    restore_from(s, s.adj_x0, s.adj_v0, 0);
    relist(s, p.rc + p.skin);
    compute_lj(s, 0, p.rc, p.eps, p.sigma);
    TMD_HIP(hipStreamSynchronize(s.stream));
}

bool traj_loss(Sim& s, int steps, int ck, double dt, double rc, double skin, double eps, double sigma, double* ubar,
               double* pbar) {
    const TrajSpec p{steps, ck, dt, rc, skin, eps, sigma};
    if (!begin_traj(s, p, false)) return false;
    forward_traj(s, p, false, ubar, pbar);
    end_traj(s, p);
    return !has_error();
}

bool traj_adjoint(Sim& s, int steps, int ck, double dt, double rc, double skin, double eps, double sigma, double u_target,
                  double p_target, double* loss3, double* grad_theta, double* grad_x, double* grad_v) {
    const TrajSpec p{steps, ck, dt, rc, skin, eps, sigma};
    if (!begin_traj(s, p, true)) return false;
    double ubar = 0.0, pbar = 0.0;
    forward_traj(s, p, true, &ubar, &pbar);
    if (has_error()) return false;
    const int n = s.n;
    const double vol = s.g.len[0] * s.g.len[1] * s.g.len[2];
    loss3[0] = (ubar - u_target) * (ubar - u_target) + (pbar - p_target) * (pbar - p_target);
    loss3[1] = ubar;
    loss3[2] = pbar;

    AdjCoef c;
    c.cU = 2.0 * (ubar - u_target) / ((double)steps * (double)n);
    c.cW = 2.0 * (pbar - p_target) / (3.0 * vol * (double)steps);
    {
        // sigma-derivatives of f(rc^2), f'(rc^2), u(rc^2)
        const double sr2 = 1.0 / (rc * rc);
        const double s3 = sigma * sigma * sigma, sr6 = sr2 * sr2 * sr2 * (s3 * s3);
        c.dfc_ds = 144.0 * eps / sigma * sr2 * sr6 * (4.0 * sr6 - 1.0);
        c.ddfc_ds = 144.0 * eps / sigma * sr2 * sr2 * sr6 * (4.0 - 28.0 * sr6);
        c.duc_ds = 24.0 * eps / sigma * sr6 * (2.0 * sr6 - 1.0);
    }
    c.loss_terms = 0;
    const double h = 0.5 * dt;
    const double cv = 2.0 * c.cW;  // d(sum v^2)/dv = 2 v
    const size_t atom = sizeof(double4) * (size_t)n;
    TMD_HIP(hipMemsetAsync(s.adj_xbar.p, 0, atom, s.stream));
    TMD_HIP(hipMemsetAsync(s.adj_vbar.p, 0, atom, s.stream));
    TMD_HIP(hipMemsetAsync(s.adj_fbar.p, 0, atom, s.stream));
    double* theta = s.adj_acc.as<double>() + 4;
    TMD_HIP(hipMemsetAsync(theta, 0, 2 * sizeof(double), s.stream));
    const int nt = 256, nb = div_up(n, nt);
    const double cut = rc + skin;
    for (int sg = steps / ck; sg >= 0 && !has_error(); --sg) {
        const int t0 = sg * ck, t1 = std::min(t0 + ck - 1, steps);
        // recompute the interval from its checkpoint (its list is fixed)
        restore_from(s, s.adj_ckx, s.adj_ckv, (size_t)sg);
        relist(s, cut);
        compute_lj(s, 0, rc, eps, sigma);
        snapshot_to(s, s.adj_segx, s.adj_segv, 0);
        for (int t = t0 + 1; t <= t1 && !has_error(); ++t) {
            initial_integration(s, dt);
            refresh_ghosts(s);
            compute_lj(s, 0, rc, eps, sigma, dt);  // final kick fused
            snapshot_to(s, s.adj_segx, s.adj_segv, (size_t)(t - t0));
        }
        // reverse it
        for (int t = t1; t >= t0 && !has_error(); --t) {
            if (t < t1) {
                restore_from(s, s.adj_segx, s.adj_segv, (size_t)(t - t0));
                refresh_ghosts(s);
            }
            hipEvent_t ev = nullptr;
            s.timer.begin(K_ADJOINT, s.stream, &ev);
            if (t >= 1)
                hipLaunchKernelGGL(k_adj_pre, dim3(nb), dim3(nt), 0, s.stream, (const double4*)s.vel.as<double4>(), n, cv, h,
                                   s.adj_vbar.as<double4>(), s.adj_fbar.as<double4>());
            c.loss_terms = t >= 1 ? 1 : 0;
            launch_lj_adjoint(s, rc, eps, sigma, c, theta);
            if (t >= 1)
                hipLaunchKernelGGL(k_adj_post, dim3(nb), dim3(nt), 0, s.stream, (const double4*)s.adj_xbar.as<double4>(), n,
                                   dt, h, s.adj_vbar.as<double4>(), s.adj_fbar.as<double4>());
            s.timer.end(K_ADJOINT, s.stream, ev);
            TMD_HIP(hipGetLastError());
        }
    }
    std::vector<double4> hx(grad_x ? n : 0), hv(grad_v ? n : 0);
    double th[2] = {0, 0};
    TMD_HIP(hipMemcpyAsync(th, theta, sizeof th, hipMemcpyDeviceToHost, s.stream));
    if (grad_x) TMD_HIP(hipMemcpyAsync(hx.data(), s.adj_xbar.p, atom, hipMemcpyDeviceToHost, s.stream));
    if (grad_v) TMD_HIP(hipMemcpyAsync(hv.data(), s.adj_vbar.p, atom, hipMemcpyDeviceToHost, s.stream));
    TMD_HIP(hipStreamSynchronize(s.stream));
    if (has_error()) return false;
    grad_theta[0] = th[0];
    grad_theta[1] = th[1];
    for (int i = 0; i < (grad_x ? n : 0); ++i) {
        grad_x[3 * i] = hx[i].x; grad_x[3 * i + 1] = hx[i].y; grad_x[3 * i + 2] = hx[i].z;
    }
    for (int i = 0; i < (grad_v ? n : 0); ++i) {
        grad_v[3 * i] = hv[i].x; grad_v[3 * i + 1] = hv[i].y; grad_v[3 * i + 2] = hv[i].z;
    }
    end_traj(s, p);
    return !has_error();
}

}  // namespace tmd
