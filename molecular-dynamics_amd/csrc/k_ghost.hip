// On-device periodic ghosts: the single-GPU replacement of the MPI border
// exchange.  For each dimension the "upper" (image -1) and "lower" (image +1)
// ghosts are selected over locals plus the ghosts of earlier dimensions by a
// flag -> tile scan -> stable scatter pass, so the ghost order is the index
// order the reference's CPU mapping produces (comm/comm.impala:528-642 with
// the predicates of comm/comm.impala:105-128 and the image flags of
// comm/lb.impala:88-100).  Every ghost records its immediate source and
// per-direction flag (the reference's send_offsets / send_pbc) and its
// flattened owner + total image, so the per-step refresh is one gather from
// the owner (comm/comm.impala:334-388 without the two-phase lag).
#include "tmd_internal.hpp"
#include "tmd_device.hpp"

namespace tmd {

struct BorderArgs {
    int d;          // dimension 0..2
    int scan_end;   // scan locals + ghosts of previous dimensions
    int n;          // number of locals
    int base;       // first new ghost slot (n + ng)
    double hi_thr;  // aabb_max - 2 s   (select for image -1)
    double lo_thr;  // aabb_min + 2 s   (select for image +1)
    double whi_thr; // world_max - s    (flag -1)
    double wlo_thr; // world_min + s    (flag +1)
    double len0, len1, len2;
};


__device__ __forceinline__ double comp(const double4& p, int d) { return d == 0 ? p.x : (d == 1 ? p.y : p.z); }

// Border selection bits of an atom, bit 2d: coordinate d above the upper
// select threshold (the ghost of image -1 side), bit 2d + 1: below the lower
// one.  A ghost made in pass d differs from its source only in coordinate d,
// so it inherits the source's bits of the later passes.  The bits are
// computed once, over the locals, together with the ghost totals of the
// three passes: a local whose family (itself and its ghosts of earlier
// passes) has m members contributes m to each later pass it is selected in.
struct FaceArgs {
    int n, wrap;
    double hi_thr[3], lo_thr[3];
    PbcArgs pbc;
};

// mpos / mvel (the fused step loop's state, in mirror order): the local's
// position and velocity are first copied back from there (k_unmirror's copy,
// through the mirror index inv).
__global__ void __launch_bounds__(kScanThreads) k_face_flags(double4* __restrict__ pos, FaceArgs a, unsigned char* __restrict__ face,
                                                             int* __restrict__ totals, const int* __restrict__ inv,
                                                             const CMir mpos, const CMir mvel,
                                                             double4* __restrict__ vel) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int t[6] = {0, 0, 0, 0, 0, 0};
    if (i < a.n) {
        double4 p;
        if (mpos.p || mvel.p) {
            const int k = inv[i];
            if (mpos.p) p = mpos[k];
            if (mvel.p) vel[i] = mvel[k];
        }
        if (!mpos.p) p = pos[i];
        if (a.wrap) pbc_wrap(p, a.pbc);  // md_enforce_pbc, fused
        if (a.wrap || mpos.p) pos[i] = p;
        const double c[3] = {p.x, p.y, p.z};
        int bits = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            bits |= (c[d] > a.hi_thr[d] ? 1 : 0) << (2 * d);
            bits |= (c[d] < a.lo_thr[d] ? 1 : 0) << (2 * d + 1);
        }
        face[i] = (unsigned char)bits;
        const int fx = 1 + (bits & 1) + ((bits >> 1) & 1);       // the local and its x ghosts
        const int fy = fx * (1 + ((bits >> 2) & 1) + ((bits >> 3) & 1));  // ... and their y ghosts
        t[0] = bits & 1;
        t[1] = (bits >> 1) & 1;
        t[2] = ((bits >> 2) & 1) * fx;
        t[3] = ((bits >> 3) & 1) * fx;
        t[4] = ((bits >> 4) & 1) * fy;
        t[5] = ((bits >> 5) & 1) * fy;
    }
    // block totals of the six counts: wavefront sums by shuffles, then one
    // LDS pass (integer sums: the order does not matter)
    __shared__ int wt[kScanThreads / kWave][6];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        int v = t[k];
#pragma unroll
        for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_down(v, d, kWave);
        if (lane == 0) wt[wid][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        int v = 0;
#pragma unroll
        for (int w = 0; w < kScanThreads / kWave; ++w) v += wt[w][threadIdx.x];
        if (v) atomicAdd(&totals[threadIdx.x], v);
    }
}

__global__ void __launch_bounds__(kScanThreads) k_border_count(const unsigned char* __restrict__ face, BorderArgs a,
                                                               int* __restrict__ tiles0, int* __restrict__ tiles1) {
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int c0 = 0, c1 = 0;
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
        const int i = base + it;
        if (i < a.scan_end) {
            const int b = face[i] >> (2 * a.d);
            c0 += b & 1;
            c1 += (b >> 1) & 1;
        }
    }
    int t0 = block_sum_int(c0);
    int t1 = block_sum_int(c1);
    if (threadIdx.x == 0) {
        tiles0[blockIdx.x] = t0;
        tiles1[blockIdx.x] = t1;
    }
}

__device__ __forceinline__ void write_ghost(int i, int slot, int side, const BorderArgs& a, double4* __restrict__ pos,
                                            double4* __restrict__ vel, double* __restrict__ mass,
                                            int* __restrict__ g_owner, int* __restrict__ g_src, int* __restrict__ g_code,
                                            unsigned char* __restrict__ face) {
    const double4 p = pos[i];
    const double pd = comp(p, a.d);
    int fl = (side == 0) ? (pd > a.whi_thr ? -1 : 0) : (pd < a.wlo_thr ? 1 : 0);
    int f[3] = {0, 0, 0};
    f[a.d] = fl;
    double4 q;
    q.x = p.x + a.len0 * (double)f[0];
    q.y = p.y + a.len1 * (double)f[1];
    q.z = p.z + a.len2 * (double)f[2];
    q.w = p.w;
    int owner, ix, iy, iz;
    if (i < a.n) {
        owner = i;
        ix = f[0]; iy = f[1]; iz = f[2];
    } else {
        const int gi = i - a.n;
        owner = g_owner[gi];
        img_decode(g_code[gi] & 0xff, ix, iy, iz);
        ix += f[0]; iy += f[1]; iz += f[2];
    }
    pos[slot] = q;
    mass[slot] = mass[i];
    vel[slot] = vel[owner];
    const int g = slot - a.n;
    g_owner[g] = owner;
    g_src[g] = i;
    g_code[g] = img_encode(ix, iy, iz) | (img_encode(f[0], f[1], f[2]) << 8);
    face[slot] = face[i];
}

__global__ void __launch_bounds__(kScanThreads) k_border_scatter(double4* __restrict__ pos, double4* __restrict__ vel,
                                                                 double* __restrict__ mass, int* __restrict__ g_owner,
                                                                 int* __restrict__ g_src, int* __restrict__ g_code,
                                                                 const int* __restrict__ off0, const int* __restrict__ off1,
                                                                 const int* __restrict__ totals, BorderArgs a,
                                                                 unsigned char* __restrict__ face) {
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int f0[kScanItems], f1[kScanItems];
    int c0 = 0, c1 = 0;
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
        int i = base + it;
        f0[it] = f1[it] = 0;
        if (i < a.scan_end) {
            const int b = face[i] >> (2 * a.d);
            f0[it] = b & 1;
            f1[it] = (b >> 1) & 1;
        }
        c0 += f0[it];
        c1 += f1[it];
    }
    int t0, t1;
    int r0 = off0[blockIdx.x] + block_exclusive_scan(c0, &t0);
    int r1 = off1[blockIdx.x] + block_exclusive_scan(c1, &t1);
    const int total0 = totals[0];
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
        int i = base + it;
        if (f0[it]) { write_ghost(i, a.base + r0, 0, a, pos, vel, mass, g_owner, g_src, g_code, face); r0++; }
        if (f1[it]) { write_ghost(i, a.base + total0 + r1, 1, a, pos, vel, mass, g_owner, g_src, g_code, face); r1++; }
    }
}

// FRESH refresh: ghost = owner + image * L (bit-identical to the rebuild's
// chained construction because each coordinate receives exactly one shift).
__global__ void k_ghost_refresh(double4* __restrict__ pos, double4* __restrict__ vel, const int* __restrict__ g_owner,
                                const int* __restrict__ g_code, int n, int ng, double L0, double L1, double L2, int with_vel,
                                const int* __restrict__ inv, const Mir spos) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const int o = g_owner[g];
    int ix, iy, iz;
    img_decode(g_code[g] & 0xff, ix, iy, iz);
    const double4 p = pos[o];
    double4 q;
    q.x = p.x + L0 * (double)ix;
    q.y = p.y + L1 * (double)iy;
    q.z = p.z + L2 * (double)iz;
    q.w = p.w;
    pos[n + g] = q;
    if (with_vel) vel[n + g] = vel[o];
    if (inv) {
        const int k = inv[n + g];
        if (k >= 0) spos[k] = q;
    }
}

// One pass over locals and ghosts after the force pass advanced the local
// positions: ghost = owner + image * L, and every binned atom's mirror
// entry (inv >= 0) is rewritten.
__global__ void k_step_refresh(double4* __restrict__ pos, const int* __restrict__ g_owner, const int* __restrict__ g_code,
                               int n, int ng, double L0, double L1, double L2, const int* __restrict__ inv,
                               const Mir spos) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n + ng) return;
    double4 q;
    if (i < n) {
        q = pos[i];
    } else {
        const int g = i - n;
        int ix, iy, iz;
        img_decode(g_code[g] & 0xff, ix, iy, iz);
        const double4 p = pos[g_owner[g]];
        q.x = p.x + L0 * (double)ix;
        q.y = p.y + L1 * (double)iy;
        q.z = p.z + L2 * (double)iz;
        q.w = p.w;
        pos[i] = q;
    }
    const int k = inv[i];
    if (k >= 0) spos[k] = q;
}

// Ghost half of k_step_refresh, after a force pass that wrote the advanced
// local positions into the mirror only: ghost = owner + image * L, with the
// owner read from its mirror entry; written to pos and to the ghost's
// mirror entry (owners are locals, so no entry is both read and written).
__global__ void k_ghost_step(double4* __restrict__ pos, const int* __restrict__ g_owner, const int* __restrict__ g_code, int n,
                             int ng, double L0, double L1, double L2, const int* __restrict__ inv, const Mir spos) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    int ix, iy, iz;
    img_decode(g_code[g] & 0xff, ix, iy, iz);
    const double4 p = spos[inv[g_owner[g]]];
    double4 q;
    q.x = p.x + L0 * (double)ix;
    q.y = p.y + L1 * (double)iy;
    q.z = p.z + L2 * (double)iz;
    q.w = p.w;
    pos[n + g] = q;
    const int k = inv[n + g];
    if (k >= 0) spos[k] = q;
}

// Mirror indices of each ghost's owner and of the ghost itself, so the
// step loop's ghost pass reads them in ghost order instead of through inv.
__global__ void k_ghost_src(const int* __restrict__ g_owner, const int* __restrict__ inv, int n, int ng,
                            int* __restrict__ g_ko, int* __restrict__ g_kg) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    g_ko[g] = inv[g_owner[g]];
    g_kg[g] = inv[n + g];
}

// Each local's ghosts by the local's work slot (a CSR over work slots), so
// that a force pass that advances a local can also write its ghosts' mirror
// entries.  Within one local the order of its ghosts is that of the atomics
// (each ghost's entry is written independently of the others).
__global__ void k_inv_work(const int* __restrict__ work, int n, int* __restrict__ inv_work) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w < n) inv_work[work[w]] = w;
}

__global__ void k_gh_count(const int* __restrict__ g_owner, const int* __restrict__ g_kg, int ng,
                           const int* __restrict__ inv_work, int* __restrict__ cnt) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng || g_kg[g] < 0) return;
    atomicAdd(&cnt[inv_work[g_owner[g]]], 1);
}

__global__ void k_gh_fill(const int* __restrict__ g_owner, const int* __restrict__ g_kg, const int* __restrict__ g_code,
                          int ng, const int* __restrict__ inv_work, const int* __restrict__ start, int* __restrict__ fill,
                          int* __restrict__ gh_k, int* __restrict__ gh_c) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng || g_kg[g] < 0) return;
    const int w = inv_work[g_owner[g]];
    const int e = start[w] + atomicAdd(&fill[w], 1);
    gh_k[e] = g_kg[g];
    gh_c[e] = g_code[g] & 0xff;
}

void build_ghost_index(Sim& s) {
    s.gh_valid = 0;
    if (has_error() || s.n == 0) return;
    if (!s.gsrc_valid) build_ghost_src(s);
    const int n = s.n, ng = s.ng;
    s.gh_start.reserve(sizeof(int) * ((size_t)n + 1), false, s.stream);
    s.gh_aux.reserve(sizeof(int) * 2 * (size_t)n, false, s.stream);
    s.gh_k.reserve(sizeof(int) * (size_t)std::max(ng, 1), false, s.stream);
    s.gh_c.reserve(sizeof(int) * (size_t)std::max(ng, 1), false, s.stream);
    if (has_error()) return;
    int* inv_work = s.gh_aux.as<int>();
    int* cnt = inv_work + n;
    const int nt = 256;
    hipLaunchKernelGGL(k_inv_work, dim3(div_up(n, nt)), dim3(nt), 0, s.stream, (const int*)s.work.as<int>(), n, inv_work);
    TMD_HIP(hipMemsetAsync(cnt, 0, sizeof(int) * (size_t)n, s.stream));
    if (ng > 0)
        hipLaunchKernelGGL(k_gh_count, dim3(div_up(ng, nt)), dim3(nt), 0, s.stream, (const int*)s.g_owner.as<int>(),
                           (const int*)s.g_kg.as<int>(), ng, (const int*)inv_work, cnt);
    exclusive_scan_int(cnt, s.gh_start.as<int>(), n, s.gh_start.as<int>() + n, s);
    TMD_HIP(hipMemsetAsync(cnt, 0, sizeof(int) * (size_t)n, s.stream));
    if (ng > 0)
        hipLaunchKernelGGL(k_gh_fill, dim3(div_up(ng, nt)), dim3(nt), 0, s.stream, (const int*)s.g_owner.as<int>(),
                           (const int*)s.g_kg.as<int>(), (const int*)s.g_code.as<int>(), ng, (const int*)inv_work,
                           (const int*)s.gh_start.as<int>(), cnt, s.gh_k.as<int>(), s.gh_c.as<int>());
    TMD_HIP(hipGetLastError());
    if (!has_error()) s.gh_valid = 1;
}

void build_ghost_src(Sim& s) {
    s.gsrc_valid = 0;
    s.gh_valid = 0;
    if (has_error() || s.ng == 0) return;
    s.g_ko.reserve(sizeof(int) * (size_t)s.cap_all, false, s.stream);
    s.g_kg.reserve(sizeof(int) * (size_t)s.cap_all, false, s.stream);
    if (has_error()) return;
    const int nt = 256;
    hipLaunchKernelGGL(k_ghost_src, dim3(div_up(s.ng, nt)), dim3(nt), 0, s.stream, (const int*)s.g_owner.as<int>(),
                       (const int*)s.cell_inv.as<int>(), s.n, s.ng, s.g_ko.as<int>(), s.g_kg.as<int>());
    TMD_HIP(hipGetLastError());
    s.gsrc_valid = 1;
}

// Ghost refresh inside the fused step loop, mirror only: owner (current in
// the mirror) + image shift, k_ghost_step's arithmetic.  pos is left for
// sync_pos_from_mirror.
__global__ void k_ghost_step_mirror(const int* __restrict__ g_ko, const int* __restrict__ g_kg,
                                    const int* __restrict__ g_code, int ng, double L0, double L1, double L2,
                                    const Mir spos) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const int kg = g_kg[g];
    if (kg < 0) return;
    int ix, iy, iz;
    img_decode(g_code[g] & 0xff, ix, iy, iz);
    const double4 p = spos[g_ko[g]];
    double4 q;
    q.x = p.x + L0 * (double)ix;
    q.y = p.y + L1 * (double)iy;
    q.z = p.z + L2 * (double)iz;
    q.w = p.w;
    spos[kg] = q;
}

// DEM step loop: the ghosts' mirror entries from their owners' (position +
// image shift, velocity copied), k_ghost_refresh's arithmetic.
__global__ void k_ghost_step_mirror_v(const int* __restrict__ g_ko, const int* __restrict__ g_kg,
                                      const int* __restrict__ g_code, int ng, double L0, double L1, double L2,
                                      const Mir spos, const Mir svel) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const int kg = g_kg[g];
    if (kg < 0) return;
    int ix, iy, iz;
    img_decode(g_code[g] & 0xff, ix, iy, iz);
    const int ko = g_ko[g];
    const double4 p = spos[ko];
    double4 q;
    q.x = p.x + L0 * (double)ix;
    q.y = p.y + L1 * (double)iy;
    q.z = p.z + L2 * (double)iz;
    q.w = p.w;
    spos[kg] = q;
    svel[kg] = svel[ko];
}

void dem_ghost_step(Sim& s) {
    if (has_error() || s.ng == 0) return;
    if (!s.gsrc_valid) build_ghost_src(s);
    const int nt = 256;
    hipLaunchKernelGGL(k_ghost_step_mirror_v, dim3(div_up(s.ng, nt)), dim3(nt), 0, s.stream, (const int*)s.g_ko.as<int>(),
                       (const int*)s.g_kg.as<int>(), (const int*)s.g_code.as<int>(), s.ng, s.g.len[0], s.g.len[1], s.g.len[2],
                       Mir{s.spos.as<double>()}, Mir{s.svel.as<double>()});
    TMD_HIP(hipGetLastError());
}

// REF_LAG refresh (diagnostic): the reference's pack from the immediate
// source (which may be a not-yet-refreshed ghost) then unpack.
__global__ void k_ghost_pack_lag(const double4* __restrict__ pos, const int* __restrict__ g_src, const int* __restrict__ g_code,
                                 int ng, double L0, double L1, double L2, double4* __restrict__ buf) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    int fx, fy, fz;
    img_decode((g_code[g] >> 8) & 0xff, fx, fy, fz);
    const double4 p = pos[g_src[g]];
    double4 q;
    q.x = p.x + L0 * (double)fx;
    q.y = p.y + L1 * (double)fy;
    q.z = p.z + L2 * (double)fz;
    q.w = p.w;
    buf[g] = q;
}

__global__ void k_ghost_unpack_lag(double4* __restrict__ pos, double4* __restrict__ vel, const int* __restrict__ g_owner,
                                   const double4* __restrict__ buf, int n, int ng, int with_vel, const int* __restrict__ inv,
                                   const Mir spos) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    pos[n + g] = buf[g];
    if (with_vel) vel[n + g] = vel[g_owner[g]];
    if (inv) {
        const int k = inv[n + g];
        if (k >= 0) spos[k] = buf[g];
    }
}

static void ensure_atom_capacity(Sim& s, int need);

// One pass over the locals (optionally wrapping them into the box first)
// gives every atom's selection bits and the ghost totals of the three passes,
// read back with one synchronisation; the passes then run back to back on
// the stream: per pass a count over the bits, a tile scan and the stable
// scatter, in the reference's order.
void build_ghosts(Sim& s, bool wrap) {
    if (has_error()) return;
    // after the fused step loop the locals' positions and velocities live in
    // the mirror: the selection pass copies them back (binning's inv is still
    // that of the mirror)
    const CMir mpos{s.pos_stale ? s.spos.as<double>() : nullptr};
    const CMir mvel{s.vel_mirrored ? s.svel.as<double>() : nullptr};
    s.pos_stale = 0;
    s.vel_mirrored = 0;
    s.ghost_pos_stale = 0;  // the ghosts are made anew
    s.ng = 0;
    s.inv_valid = 0;
    s.gsrc_valid = 0;
    mirror_invalidate(s);
    if (s.n == 0) return;
    const Geom& G = s.g;
    const double dspacing = G.spacing * 2.0;
    int* counters = s.counters.as<int>();
    s.face.reserve((size_t)std::max(s.cap_all, s.n), true, s.stream);
    if (has_error()) return;
    FaceArgs fa;
    fa.n = s.n;
    fa.wrap = wrap ? 1 : 0;
    for (int d = 0; d < 3; ++d) {
        fa.hi_thr[d] = G.aabb[2 * d + 1] - dspacing;
        fa.lo_thr[d] = G.aabb[2 * d] + dspacing;
    }
    fa.pbc = pbc_args(s);
    TMD_HIP(hipMemsetAsync(counters + 8, 0, 6 * sizeof(int), s.stream));
    hipLaunchKernelGGL(k_face_flags, dim3(div_up(s.n, kScanThreads)), dim3(kScanThreads), 0, s.stream, s.pos.as<double4>(), fa,
                       s.face.as<unsigned char>(), counters + 8, (const int*)s.cell_inv.as<int>(), mpos, mvel,
                       s.vel.as<double4>());
    int tot[6] = {0, 0, 0, 0, 0, 0};
    TMD_HIP(hipMemcpyAsync(tot, counters + 8, sizeof tot, hipMemcpyDeviceToHost, s.stream));
    TMD_HIP(hipStreamSynchronize(s.stream));
    if (has_error()) return;
    const int ng_all = tot[0] + tot[1] + tot[2] + tot[3] + tot[4] + tot[5];
    ensure_atom_capacity(s, s.n + ng_all);
    if (has_error()) return;
    const int max_tiles = div_up(s.n + ng_all, kScanTile);
    s.scan_blocks.reserve(sizeof(int) * 2 * (size_t)max_tiles, false, s.stream);
    if (has_error()) return;
    int ng = 0;
    for (int d = 0; d < 3; ++d) {
        BorderArgs a;
        a.d = d;
        a.n = s.n;
        a.scan_end = s.n + ng;
        a.base = s.n + ng;
        a.hi_thr = fa.hi_thr[d];
        a.lo_thr = fa.lo_thr[d];
        a.whi_thr = G.world[2 * d + 1] - G.spacing;
        a.wlo_thr = G.world[2 * d] + G.spacing;
        a.len0 = G.len[0]; a.len1 = G.len[1]; a.len2 = G.len[2];
        const int add = tot[2 * d] + tot[2 * d + 1];
        if (add > 0) {
            const int ntiles = div_up(a.scan_end, kScanTile);
            int* t0 = s.scan_blocks.as<int>();
            int* t1 = t0 + ntiles;
            hipLaunchKernelGGL(k_border_count, dim3(ntiles), dim3(kScanThreads), 0, s.stream,
                               (const unsigned char*)s.face.as<unsigned char>(), a, t0, t1);
            scan_tiles(s, t0, ntiles, counters + 0);
            scan_tiles(s, t1, ntiles, counters + 1);
            hipLaunchKernelGGL(k_border_scatter, dim3(ntiles), dim3(kScanThreads), 0, s.stream, s.pos.as<double4>(),
                               s.vel.as<double4>(), s.mass.as<double>(), s.g_owner.as<int>(), s.g_src.as<int>(),
                               s.g_code.as<int>(), (const int*)t0, (const int*)t1, (const int*)counters, a,
                               s.face.as<unsigned char>());
            TMD_HIP(hipGetLastError());
        }
        ng += add;
    }
    s.ng = ng;
}

void refresh_ghosts(Sim& s) {
    if (has_error() || s.ng == 0) return;
    const int with_vel = 1;
    const int nt = 256;
    if (s.ghost_mode == GHOST_REF_LAG) {
        s.gbuf.reserve(sizeof(double4) * (size_t)s.ng, false, s.stream);
        hipLaunchKernelGGL(k_ghost_pack_lag, dim3(div_up(s.ng, nt)), dim3(nt), 0, s.stream, (const double4*)s.pos.as<double4>(),
                           (const int*)s.g_src.as<int>(), (const int*)s.g_code.as<int>(), s.ng, s.g.len[0], s.g.len[1], s.g.len[2],
                           s.gbuf.as<double4>());
        hipLaunchKernelGGL(k_ghost_unpack_lag, dim3(div_up(s.ng, nt)), dim3(nt), 0, s.stream, s.pos.as<double4>(), s.vel.as<double4>(),
                           (const int*)s.g_owner.as<int>(), (const double4*)s.gbuf.as<double4>(), s.n, s.ng, with_vel,
                           s.inv_valid ? (const int*)s.cell_inv.as<int>() : nullptr, Mir{s.spos.as<double>()});
    } else {
        hipLaunchKernelGGL(k_ghost_refresh, dim3(div_up(s.ng, nt)), dim3(nt), 0, s.stream, s.pos.as<double4>(), s.vel.as<double4>(),
                           (const int*)s.g_owner.as<int>(), (const int*)s.g_code.as<int>(), s.n, s.ng, s.g.len[0], s.g.len[1],
                           s.g.len[2], with_vel, s.inv_valid ? (const int*)s.cell_inv.as<int>() : nullptr,
                           Mir{s.spos.as<double>()});
    }
    TMD_HIP(hipGetLastError());
}

// pos[parts[k]] = mirror[k] for the locals of the cell CSR (when pos is
// given), and likewise the velocities from their mirror order (when svel is)
__global__ void k_unmirror(const CMir spos, const int* __restrict__ parts, int nbinned, int n,
                           double4* __restrict__ pos, const CMir svel, double4* __restrict__ vel) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nbinned) return;
    const int j = parts[k];
    if (j < n) {
        if (pos) pos[j] = spos[k];
        if (svel.p) vel[j] = svel[k];
    }
}

// Positions (and velocities) of the fused step loop back into atom order.
void sync_pos_from_mirror(Sim& s) {
    if (has_error()) return;
    if (s.ghost_pos_stale) {
        // the ghosts' positions in pos, from their owners in the mirror
        if (s.ng > 0)
            hipLaunchKernelGGL(k_ghost_step, dim3(div_up(s.ng, 256)), dim3(256), 0, s.stream, s.pos.as<double4>(),
                               (const int*)s.g_owner.as<int>(), (const int*)s.g_code.as<int>(), s.n, s.ng, s.g.len[0],
                               s.g.len[1], s.g.len[2], (const int*)s.cell_inv.as<int>(), Mir{s.spos.as<double>()});
        TMD_HIP(hipGetLastError());
        s.ghost_pos_stale = 0;
    }
    if (!s.pos_stale && !s.vel_mirrored) return;
    const int nt = 256;
    if (s.nbinned > 0)
        hipLaunchKernelGGL(k_unmirror, dim3(div_up(s.nbinned, nt)), dim3(nt), 0, s.stream, CMir{s.spos.as<double>()},
                           (const int*)s.cell_parts.as<int>(), s.nbinned, s.n,
                           s.pos_stale ? s.pos.as<double4>() : (double4*)nullptr,
                           CMir{s.vel_mirrored ? s.svel.as<double>() : nullptr}, s.vel.as<double4>());
    TMD_HIP(hipGetLastError());
    s.pos_stale = 0;
    s.vel_mirrored = 0;
}

// velocities of the locals between atom order (vel) and mirror order (svel)
__global__ void k_vel_mirror(const int* __restrict__ work, const int* __restrict__ wk, int n, double4* __restrict__ vel,
                             const Mir svel, int into_mirror) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n) return;
    if (into_mirror) svel[wk[w]] = vel[work[w]];
    else vel[work[w]] = svel[wk[w]];
}

// The fused step loop keeps the locals' velocities in mirror order (the
// force pass then reads and writes them with the same coalesced index as the
// positions); they return to atom order before each rebuild and at loop exit.
void mirror_velocities(Sim& s, bool into_mirror) {
    if (has_error() || s.n == 0) return;
    if (into_mirror == (s.vel_mirrored != 0)) return;
    if (into_mirror) {
        s.svel.reserve((size_t)kMirBytes * (size_t)std::max(s.nbinned, 1), false, s.stream);
        if (has_error()) return;
    }
    const int nt = 256;
    hipLaunchKernelGGL(k_vel_mirror, dim3(div_up(s.n, nt)), dim3(nt), 0, s.stream, (const int*)s.work.as<int>(),
                       (const int*)s.work_k.as<int>(), s.n, s.vel.as<double4>(), Mir{s.svel.as<double>()}, into_mirror ? 1 : 0);
    TMD_HIP(hipGetLastError());
    s.vel_mirrored = into_mirror ? 1 : 0;
}

void step_refresh(Sim& s) {
    if (has_error() || s.n + s.ng == 0) return;
    if (!s.inv_valid || s.ghost_mode != GHOST_FRESH) {
        s.ghost_stale = 0;
        refresh_ghosts(s);
        return;
    }
    const int nt = 256;
    if (s.ghost_stale && s.pos_stale && s.gsrc_valid) {
        // fused step loop: the ghosts in the mirror only (pos at the end)
        if (s.ng > 0)
            hipLaunchKernelGGL(k_ghost_step_mirror, dim3(div_up(s.ng, nt)), dim3(nt), 0, s.stream,
                               (const int*)s.g_ko.as<int>(), (const int*)s.g_kg.as<int>(), (const int*)s.g_code.as<int>(),
                               s.ng, s.g.len[0], s.g.len[1], s.g.len[2], Mir{s.spos.as<double>()});
        s.ghost_pos_stale = 1;
    } else if (s.ghost_stale) {
        // locals are current in the mirror: rewrite the ghosts only
        if (s.ng > 0)
            hipLaunchKernelGGL(k_ghost_step, dim3(div_up(s.ng, nt)), dim3(nt), 0, s.stream, s.pos.as<double4>(),
                               (const int*)s.g_owner.as<int>(), (const int*)s.g_code.as<int>(), s.n, s.ng, s.g.len[0],
                               s.g.len[1], s.g.len[2], (const int*)s.cell_inv.as<int>(), Mir{s.spos.as<double>()});
    } else {
        hipLaunchKernelGGL(k_step_refresh, dim3(div_up(s.n + s.ng, nt)), dim3(nt), 0, s.stream, s.pos.as<double4>(),
                           (const int*)s.g_owner.as<int>(), (const int*)s.g_code.as<int>(), s.n, s.ng, s.g.len[0], s.g.len[1],
                           s.g.len[2], (const int*)s.cell_inv.as<int>(), Mir{s.spos.as<double>()});
    }
    TMD_HIP(hipGetLastError());
    s.ghost_stale = 0;
    s.mirror_fresh = 1;
}

// Grow every per-atom array (locals + ghosts) to hold `need` atoms,
// preserving contents.  Ghost-table arrays are indexed by ghost number.
static void ensure_atom_capacity(Sim& s, int need) {
    if (need <= s.cap_all) return;
    int cap = need + need / 5 + 1024;
    s.pos.reserve(sizeof(double4) * (size_t)cap, true, s.stream);
    s.vel.reserve(sizeof(double4) * (size_t)cap, true, s.stream);
    s.mass.reserve(sizeof(double) * (size_t)cap, true, s.stream);
    s.g_owner.reserve(sizeof(int) * (size_t)cap, true, s.stream);
    s.g_src.reserve(sizeof(int) * (size_t)cap, true, s.stream);
    s.g_code.reserve(sizeof(int) * (size_t)cap, true, s.stream);
    s.cell_of.reserve(sizeof(int) * (size_t)cap, false, s.stream);
    s.cell_slot.reserve(sizeof(int) * (size_t)cap, false, s.stream);
    s.cell_parts.reserve(sizeof(int) * (size_t)cap, false, s.stream);
    s.cell_inv.reserve(sizeof(int) * (size_t)cap, false, s.stream);
    s.face.reserve((size_t)cap, true, s.stream);
    s.inv_valid = 0;
    s.gsrc_valid = 0;
    s.cap_all = cap;
}

// End of the DEM mirror phase: the locals' positions and velocities back to
// atom order, then the ghosts' from their owners.
void dem_unmirror(Sim& s) {
    if (has_error() || !s.dem_mirrored) return;
    s.dem_mirrored = 0;
    const int nt = 256;
    if (s.nbinned > 0)
        hipLaunchKernelGGL(k_unmirror, dim3(div_up(s.nbinned, nt)), dim3(nt), 0, s.stream, CMir{s.spos.as<double>()},
                           (const int*)s.cell_parts.as<int>(), s.nbinned, s.n, s.pos.as<double4>(),
                           CMir{s.svel.as<double>()}, s.vel.as<double4>());
    TMD_HIP(hipGetLastError());
    refresh_ghosts(s);  // also writes the ghosts' mirror entries, so the mirror stays that of pos
}

}  // namespace tmd
