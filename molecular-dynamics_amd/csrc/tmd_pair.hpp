// Per-atom pair-force bodies shared by the single-box kernels (k_force.hip)
// and the batched-replica kernels (k_batch.hip): both evaluate exactly the
// same arithmetic in the same order.
#pragma once

#include "tmd_window.hpp"
#include "tmd_lj.hpp"

namespace tmd {

template <int SHIFT>
__device__ __forceinline__ double lj_f64(const LJArgs& a, double rsq) {
    const double sr2 = 1.0 / rsq;
    const double sr6 = sr2 * sr2 * sr2 * a.sigma6;
    double f = 48.0 * sr6 * (sr6 - 0.5) * sr2 * a.eps;
    if (SHIFT) f = f - a.fc - (rsq - a.rc2) * a.dfc;
    return f;
}

template <int SHIFT>
__device__ __forceinline__ float lj_f32(const LJArgs& a, float rsq) {
    const float sr2 = __builtin_amdgcn_rcpf(rsq);
    const float sr6 = sr2 * sr2 * sr2 * a.sigma6f;
    float f = 48.0f * sr6 * (sr6 - 0.5f) * sr2 * a.epsf;
    if (SHIFT) f = f - a.fcf - (rsq - a.rc2f) * a.dfcf;
    return f;
}

// entry u (0..7) of a packed group of 8 u16 list entries
__device__ __forceinline__ int nb_entry(const int4& q, int u) {
    const int w = (u >> 1) == 0 ? q.x : ((u >> 1) == 1 ? q.y : ((u >> 1) == 2 ? q.z : q.w));
    return (u & 1) ? (int)((unsigned)w >> 16) : (w & 0xffff);
}


// FP64 pair scalar of the reference closure; UNIT skips the exact
// multiplications by eps = 1 and sigma^6 = 1 (x * 1.0 == x), so the result
// is bit-identical either way.
template <int SHIFT, int UNIT>
__device__ __forceinline__ double lj_f64u(const LJArgs& a, double rsq) {
    const double sr2 = 1.0 / rsq;
    const double sr6 = UNIT ? sr2 * sr2 * sr2 : sr2 * sr2 * sr2 * a.sigma6;
    double f = UNIT ? 48.0 * sr6 * (sr6 - 0.5) * sr2 : 48.0 * sr6 * (sr6 - 0.5) * sr2 * a.eps;
    if (SHIFT) f = f - a.fc - (rsq - a.rc2) * a.dfc;
    return f;
}

template <int SHIFT, int UNIT>
__device__ __forceinline__ float lj_f32u(const LJArgs& a, float rsq) {
    const float sr2 = __builtin_amdgcn_rcpf(rsq);
    const float sr6 = UNIT ? sr2 * sr2 * sr2 : sr2 * sr2 * sr2 * a.sigma6f;
    // UNIT: 48 sr6 (sr6 - 1/2) sr2 as (sr6 sr2) (48 sr6 - 24), one FMA
    float f = UNIT ? (sr6 * sr2) * __builtin_fmaf(48.0f, sr6, -24.0f) : 48.0f * sr6 * (sr6 - 0.5f) * sr2 * a.epsf;
    if (SHIFT) f = f - a.fcf - (rsq - a.rc2f) * a.dfcf;
    return f;
}

// Kahan compensated summation (s, c) += x.  c keeps the excess that the
// rounded addition put into s, (t - s) - y, so the compensated value of the
// sum is s - c (kahan_value).  The translation unit is built without FP
// contraction, so the four operations are not fused.
__host__ __device__ __forceinline__ void kahan_add(float& s, float& c, float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

__host__ __device__ __forceinline__ double kahan_value(float s, float c) { return (double)s - (double)c; }

// Neighbour position of window offset t: LDS (x, y, z interleaved) when the
// window is staged, else the cell-ordered mirror.
template <bool STAGED, class SP>
__device__ __forceinline__ void win_pos(int t, const WinStage& st, const SP& spos, const TileWin& W,
                                        double& x, double& y, double& z) {
    if (STAGED) {
        x = st.xyz[3 * t + 0];
        y = st.xyz[3 * t + 1];
        z = st.xyz[3 * t + 2];
    } else {
        const double4 p = spos[win_index(t, W)];
        x = p.x;
        y = p.y;
        z = p.z;
    }
}

// Full-list force on one atom (lane).  Entries are read 8 at a time (one
// 16-byte load, prefetched a group ahead) and evaluated 4 at a time; the
// pair terms are evaluated without branches (a masked entry's scalar is
// replaced by zero) and accumulated in list order.  Adding a masked term
// adds dx * 0.0 = +-0.0, which leaves the running sum unchanged (the sum
// starts at +0.0 and never becomes -0.0), so the FP64 result equals the
// sequential reference sum bit for bit.  From a staged window, the entries
// past the count in the last group are the sentinel offset (beyond the
// cutoff), so only the cutoff test masks; from the mirror, the count does.
// FP32 variants form rsq with FMAs; FP64 keeps the reference's rounding.
//
// PAIR2 (half lists): the lane's groups are column 1's (cnt entries at lp)
// followed by column 2's (cnt2 at lp2), one running sum over both; a
// wavefront then runs as many groups as its busiest lane has in both
// columns together, not the busiest lane of each column in turn.
template <int PREC, int SHIFT, int UNIT, bool STAGED, bool PAIR2 = false, class SP = const double4*>
__device__ __forceinline__ double4 lj_full_atom(const double4 pi, int cnt, const int4* __restrict__ lp, const WinStage& sp,
                                                const SP& spos, const TileWin& W, const LJArgs& a,
                                                int cnt2 = 0, const int4* __restrict__ lp2 = nullptr) {
    double fx = 0.0, fy = 0.0, fz = 0.0;
    float sxa = 0.f, sya = 0.f, sza = 0.f, cxa = 0.f, cya = 0.f, cza = 0.f;
    const int ng1 = (cnt + 7) >> 3;
    const int ng = ng1 + (PAIR2 ? (cnt2 + 7) >> 3 : 0);
    // group g of the lane's entries (group 0 of either column always
    // exists: capacity >= 8)
    auto grp = [&](int g) -> const int4* {
        return (!PAIR2 || g < ng1) ? lp + (size_t)g * 64 : lp2 + (size_t)(g - ng1) * 64;
    };
    // the next group of 8 entries is loaded while the current one is used
    // (streaming loads: each group is read once per pass)
    int4 q = ld_stream(grp(0));
    for (int gi = 0; gi < ng; ++gi) {
        const int4 e = q;
        q = ld_stream(grp(min(gi + 1, ng - 1)));  // next group (clamped: always a valid load)
        const int rem = (!PAIR2 || gi < ng1) ? cnt - gi * 8 : cnt2 - (gi - ng1) * 8;  // entries in this group
        // two passes of 4 entries (fewer live registers than 8 at once),
        // unrolled: the compiler then overlaps one pass's LDS reads with the
        // other's arithmetic
#pragma unroll
        for (int hb = 0; hb < 8; hb += 4) {
        if (hb >= rem) break;  // this lane's entries are done
        double dx[4], dy[4], dz[4], rs[4];
        bool ok[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const bool valid = STAGED || hb + u < rem;
            double xj, yj, zj;
            win_pos<STAGED>(valid ? nb_entry(e, hb + u) : 0, sp, spos, W, xj, yj, zj);
            dx[u] = pi.x - xj;
            dy[u] = pi.y - yj;
            dz[u] = pi.z - zj;
            rs[u] = PREC == PREC_FP64 ? dx[u] * dx[u] + dy[u] * dy[u] + dz[u] * dz[u]
                                      : fma(dx[u], dx[u], fma(dy[u], dy[u], dz[u] * dz[u]));
            ok[u] = valid && rs[u] < a.rc2;
        }
        // The pair scalar is evaluated for every entry and masked afterwards:
        // a masked entry may have rsq = 0 (its own position) or lie beyond
        // the cutoff; its (possibly infinite) scalar is replaced by 0 by the
        // select, so the masked term adds dx * 0 = +-0.
        if (PREC == PREC_FP64) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double f = lj_f64u<SHIFT, UNIT>(a, rs[u]);
                const double fm = ok[u] ? f : 0.0;
                fx = fx + dx[u] * fm;
                fy = fy + dy[u] * fm;
                fz = fz + dz[u] * fm;
            }
        } else if (PREC == PREC_FP32_MIXED) {
            // FP32 pair scalar, FP64 fused multiply-add accumulation of dx * f
            // (one rounding per term; the FP64 mantissa absorbs the sum's
            // rounding)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                double fm;
                if (SHIFT) {
                    const float f = lj_f32u<SHIFT, UNIT>(a, (float)rs[u]);
                    fm = (double)(ok[u] ? f : 0.0f);
                } else {
                    // a masked entry gets rsq = +inf, so sr2 = 1/inf = 0 and
                    // the plain-cutoff scalar is a signed zero (one 32-bit
                    // select instead of a 64-bit one)
                    fm = (double)lj_f32u<SHIFT, UNIT>(a, ok[u] ? (float)rs[u] : __builtin_inff());
                }
                fx = fma(dx[u], fm, fx);
                fy = fma(dy[u], fm, fy);
                fz = fma(dz[u], fm, fz);
            }
        } else {
            // FP32 pair terms with Kahan-compensated FP32 accumulation
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float f = lj_f32u<SHIFT, UNIT>(a, (float)rs[u]);
                const float fm = ok[u] ? f : 0.0f;
                kahan_add(sxa, cxa, (float)dx[u] * fm);
                kahan_add(sya, cya, (float)dy[u] * fm);
                kahan_add(sza, cza, (float)dz[u] * fm);
            }
        }
        }
    }
    if (PREC != PREC_FP32_KAHAN) return make_double4(fx, fy, fz, 0.0);
    return make_double4(kahan_value(sxa, cxa), kahan_value(sya, cya), kahan_value(sza, cza), 0.0);
}


// Energy and virial terms of one atom's list entries (FP64): U = 4 eps
// (sr12 - sr6) [+ shift terms], W = f r^2.  Full lists weight every entry
// 1/2; half lists weight local-local entries 1 and local-ghost entries 1/2
// (an index j >= nloc_end is a ghost).  Entries are read 8 at a time (one
// 16-byte load per group, as in the force pass) and summed in list order.
template <int SHIFT, bool STAGED, class SP>
__device__ __forceinline__ void lj_energy_atom(const double4 pi, int cnt, const int4* __restrict__ lp, const WinStage& sp,
                                               const SP& spos, const int* __restrict__ parts,
                                               const TileWin& W, const LJArgs& a, int half, int nloc_end, double& pe,
                                               double& vir, double& np) {
    const int ng = (cnt + 7) >> 3;
    // the next group of 8 entries is loaded while the current one is used
    int4 q = ld_stream(lp);  // group 0 of the column always exists (capacity >= 8)
    for (int gi = 0; gi < ng; ++gi) {
        const int4 e = q;
        q = ld_stream(lp + (size_t)min(gi + 1, ng - 1) * 64);  // next group (clamped: always a valid load)
        const int kend = cnt - gi * 8;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            // staged, the entries past the count are the sentinel slot (far
            // away, j = -1) and fail the cutoff test; from the mirror they
            // would index past the window
            if (!STAGED && u >= kend) break;
            const int t = nb_entry(e, u);
            int j;
            double xj, yj, zj;
            if (STAGED) {
                j = sp.j[t];
                xj = sp.xyz[3 * t + 0]; yj = sp.xyz[3 * t + 1]; zj = sp.xyz[3 * t + 2];
            } else {
                const int kk = win_index(t, W);
                j = parts[kk];
                const double4 p = spos[kk];
                xj = p.x; yj = p.y; zj = p.z;
            }
            const double dx = pi.x - xj, dy = pi.y - yj, dz = pi.z - zj;
            const double rsq = dx * dx + dy * dy + dz * dz;
            if (rsq < a.rc2) {
                const double sr2 = 1.0 / rsq;
                const double sr6 = sr2 * sr2 * sr2 * a.sigma6;
                double f = 48.0 * sr6 * (sr6 - 0.5) * sr2 * a.eps;
                double uu = 4.0 * a.eps * sr6 * (sr6 - 1.0);
                if (SHIFT) {
                    const double ds = rsq - a.rc2;
                    f = f - a.fc - ds * a.dfc;
                    uu = uu - a.uc + ds * (0.5 * a.fc) + ds * ds * (0.25 * a.dfc);
                } else if (a.eshift != 0.0) {
                    uu = uu - a.eshift;
                }
                const double wt = half ? (j < nloc_end ? 1.0 : 0.5) : 0.5;
                pe += wt * uu;
                vir += wt * (f * rsq);
                np += 1.0;
            }
        }
    }
}

struct NbArgs {
    int n, cap, half, half_mode, wcap;  // n: first ghost index (locals are below it)
    double cut2;
    TileGeom g;
};

// One list column being written: groups of 8 entries assembled in registers
// (q0..q3 hold the last 8 entries, the newest in the high half of q3, and p
// the two before them) and stored whole, one 16-byte store per group; groups
// beyond the capacity are counted, not stored.
struct ListCol {
    int4* col;
    int ngroups;
    unsigned p, q0, q1, q2, q3;
    int cnt;
    __device__ __forceinline__ ListCol(int4* c, int ng)
        : col(c), ngroups(ng), p(0u), q0(0u), q1(0u), q2(0u), q3(0u), cnt(0) {}
    __device__ __forceinline__ void shift_in(unsigned t) {
        q0 = __builtin_amdgcn_alignbit(q1, q0, 16);
        q1 = __builtin_amdgcn_alignbit(q2, q1, 16);
        q2 = __builtin_amdgcn_alignbit(q3, q2, 16);
        q3 = __builtin_amdgcn_alignbit(t, q3, 16);
    }
    __device__ __forceinline__ void store(unsigned d0, unsigned d1, unsigned d2, unsigned d3) {
        const int grp = (cnt >> 3) - 1;
        if (grp < ngroups) col[(size_t)grp * 64] = make_int4((int)d0, (int)d1, (int)d2, (int)d3);
    }
    // one entry (p is not kept up to date here: push2 refills it before any
    // store reads it)
    __device__ __forceinline__ void push(unsigned t) {
        shift_in(t);
        ++cnt;
        if ((cnt & 7) == 0) store(q0, q1, q2, q3);
    }
    // two entries, t1 first: whole dwords move, no funnel shifts.  A group
    // completes either at t2 (the last 8 entries, q0..q3) or at t1 (the 8
    // before t2, one entry lower: the high half of p through the low half
    // of q3).
    __device__ __forceinline__ void push2(unsigned t1, unsigned t2) {
        p = q0;
        q0 = q1;
        q1 = q2;
        q2 = q3;
        q3 = t1 | (t2 << 16);
        cnt += 2;
        if ((cnt & 6) == 0) {
            if (cnt & 1)
                store(__builtin_amdgcn_alignbit(q0, p, 16), __builtin_amdgcn_alignbit(q1, q0, 16),
                      __builtin_amdgcn_alignbit(q2, q1, 16), __builtin_amdgcn_alignbit(q3, q2, 16));
            else
                store(q0, q1, q2, q3);
        }
    }
    // the set bits of m (bit b: entry base + b) in ascending order: an odd
    // one alone, then two per step
    __device__ __forceinline__ void push_mask(unsigned m, unsigned base) {
        if (__builtin_popcount(m) & 1) {
            push(base + (unsigned)__builtin_ctz(m));
            m &= m - 1u;
        }
        while (m != 0u) {
            const unsigned b1 = (unsigned)__builtin_ctz(m);
            m &= m - 1u;
            const unsigned b2 = (unsigned)__builtin_ctz(m);
            m &= m - 1u;
            push2(base + b1, base + b2);
        }
    }
    // the last group's free slots hold the window's sentinel offset
    __device__ __forceinline__ void finish(unsigned pad) {
        if ((cnt & 7) != 0 && (cnt >> 3) < ngroups) {
            for (int p = cnt & 7; p < 8; ++p) shift_in(pad);
            col[(size_t)(cnt >> 3) * 64] = make_int4((int)q0, (int)q1, (int)q2, (int)q3);
        }
    }
};

// Verlet list of local atom i (work slot w, cell c of the tile's grid,
// cell-ordered index k_i) from the staged window: 9 stencil rows of 3 cells
// in the reference's order, `rsq <= cut2` in FP64 without contraction.  Row
// bounds come from the staged row starts (cst).  Each row is tested in
// chunks of 32 candidates that produce a bit mask (bit b: candidate tc + b
// accepted); i itself is removed from the mask by its window offset, so full
// lists read no atom index (half lists read the staged index for the j < i /
// ghost filter).  The set bits are then appended in ascending order, which is
// the reference's order (ListCol, two entries per step).  Only accepted
// candidates pay for the packing.  Returns the count.
//
// HALF also writes the complementary column (nb_list2, count in cnt2): the
// local neighbours j > i that the half list leaves to j's own list.  The
// pair force then needs no reaction scatter: atom i gathers its half-list
// entries and its complementary entries, which are exactly the pairs whose
// reactions the half list would have sent to i.  STAGED is a template
// parameter so that the staged loop addresses LDS directly (ds_read) rather
// than through generic pointers.
template <bool HALF, bool STAGED, class SP>
__device__ __forceinline__ int nlist_atom(int i, int w, int c, int k_i, const double4 pi, const int* cst, const NbArgs& a,
                                          const TileWin& W, const WinStage& sp, const SP& spos,
                                          const int* __restrict__ parts, nb_t* __restrict__ nb_list,
                                          nb_t* __restrict__ nb_list2, int& cnt2) {
    const int cx = c % a.g.nx, cy = (c / a.g.nx) % a.g.ny, cz = c / (a.g.nx * a.g.ny);
    const int wy = a.g.ty + 2;
    const int xo = cx - W.cx0;  // (cx - 1) - (cx0 - 1)
    const int per = a.g.tx + 3;
    const int s_own = (cz - (W.cz0 - 1)) * wy + (cy - (W.cy0 - 1));
    const int t_i = W.seg_off[s_own] + (k_i - W.seg_beg[s_own]);
    const size_t col_off = (size_t)(w >> 6) * (size_t)a.cap * 8 + (w & 63);
    ListCol A(reinterpret_cast<int4*>(nb_list) + col_off, a.cap >> 3);
    ListCol B(HALF ? reinterpret_cast<int4*>(nb_list2) + col_off : nullptr, a.cap >> 3);
    for (int r = 0; r < 9; ++r) {
        const int s = s_own + (r / 3 - 1) * wy + (r % 3 - 1);
        const int t0 = cst[s * per + xo], t1 = cst[s * per + xo + 3];
        for (int tc = t0; tc < t1; tc += 32) {
            const int te = min(tc + 32, t1);
            // the candidates' bits are shifted in at the low end (first
            // candidate highest): half lists a batch of 4 at a time, full
            // lists one sign bit per candidate (one funnel shift)
            unsigned m = 0u, m2 = 0u;
            const int nbat = (te - tc + 3) >> 2;  // batches (the loop test is then one compare)
            for (int q = 0; q < nbat; ++q) {
                const int tb = tc + 4 * q;
                unsigned bq = 0u, bq2 = 0u;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    // staged: up to three candidates past te are read (window
                    // atoms or sentinel slots) and masked below; from the
                    // mirror, t is clamped to stay inside the window
                    const int t = STAGED ? tb + u : min(tb + u, te - 1);
                    double xj, yj, zj;
                    int j = 0;
                    if (STAGED) {
                        xj = sp.xyz[3 * t + 0]; yj = sp.xyz[3 * t + 1]; zj = sp.xyz[3 * t + 2];
                        if (HALF) j = sp.j[t];
                    } else {
                        const int kk = win_index(t, W);
                        const double4 p = spos[kk];
                        xj = p.x; yj = p.y; zj = p.z;
                        if (HALF) j = parts[kk];
                    }
                    const double dx = pi.x - xj, dy = pi.y - yj, dz = pi.z - zj;
                    const double rsq = dx * dx + dy * dy + dz * dz;
                    if (HALF) {
                        bool ok = rsq <= a.cut2;
                        bq2 |= (ok && j < a.n && j > i) ? 8u >> u : 0u;
                        ok = ok && ((a.half_mode == HALF_REF_LITERAL) ? (i > j) : (j >= a.n || i > j));
                        bq |= ok ? 8u >> u : 0u;
                    } else {
                        // the sign of cut2 - rsq is set exactly when rsq > cut2
                        // (the difference of two equal finite doubles is +0);
                        // it is shifted in at the low end
                        const double d = a.cut2 - rsq;
                        m = __builtin_amdgcn_alignbit(m, (unsigned)__double2hiint(d), 31);
                    }
                }
                if (HALF) {
                    m = (m << 4) | bq;
                    m2 = (m2 << 4) | bq2;
                }
            }
            if (!HALF) m = ~m;  // the bits were set for the rejected candidates
            // reversed, bit b is candidate tc + b (nbat >= 1: the chunk is not empty)
            m = __builtin_bitreverse32(m) >> (32 - 4 * nbat);
            if (HALF) m2 = __builtin_bitreverse32(m2) >> (32 - 4 * nbat);
            const unsigned chunk = te - tc == 32 ? 0xffffffffu : (1u << (te - tc)) - 1u;  // candidates of this chunk only
            m &= chunk;
            const unsigned d = (unsigned)(t_i - tc);
            m &= ~(d < 32u ? 1u << d : 0u);  // i is not its own neighbour
            A.push_mask(m, (unsigned)tc);
            if (HALF) B.push_mask(m2 & chunk, (unsigned)tc);  // j > i excludes i itself
        }
    }
    const unsigned pad = (unsigned)win_size(W);
    A.finish(pad);
    if (HALF) B.finish(pad);
    cnt2 = B.cnt;
    return A.cnt;
}

}  // namespace tmd
