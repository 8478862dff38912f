// Lennard-Jones parameters shared by the force, energy and adjoint kernels.
//
// Pair force = f(s) * del with s = |del|^2 and
//   f(s) = 48 eps sr6 (sr6 - 1/2) sr2,  sr2 = 1/s, sr6 = sigma^6 sr2^3
// (core/runtime.impala:188-201).  The smoothly truncated variant used for
// differentiable runs subtracts the value and slope at sc = rc^2:
//   f_s(s) = f(s) - f(sc) - (s - sc) f'(sc)
//   u_s(s) = u(s) - u(sc) + (s - sc) f(sc)/2 + (s - sc)^2 f'(sc)/4
// so force and force derivative vanish at the cutoff (u_s' = -f_s / 2).
// For energy bookkeeping of plain-cutoff runs, eshift = u(sc) can be
// subtracted from each in-cutoff pair's energy (the energy-shifted
// potential, continuous at the cutoff; the force is unchanged).
#pragma once

#include "tmd_internal.hpp"

namespace tmd {

struct LJArgs {
    double rc2, rc, eps, sigma, sigma6, fc, dfc, uc;
    double eshift;  // energy only: subtracted from each in-cutoff pair's U (0: none)
    float epsf, sigma6f, fcf, dfcf, rc2f;
    int cap, nwork, n, wcap;
    TileGeom g;
};

inline LJArgs make_lj_args(const Sim& s, double rc, double eps, double sigma) {
    LJArgs a;
    a.rc = rc;
    a.rc2 = rc * rc;
    a.eps = eps;
    a.sigma = sigma;
    const double s3 = sigma * (sigma * sigma);  // pow(sigma, 6): utils/utilities.impala:27-36
    a.sigma6 = s3 * s3;
    const double sr2 = 1.0 / a.rc2;
    const double sr6 = sr2 * sr2 * sr2 * a.sigma6;
    a.fc = 48.0 * sr6 * (sr6 - 0.5) * sr2 * eps;
    a.dfc = 24.0 * eps * sr2 * sr2 * sr6 * (4.0 - 14.0 * sr6);
    a.uc = 4.0 * eps * sr6 * (sr6 - 1.0);
    a.eshift = 0.0;
    a.epsf = (float)eps;
    a.sigma6f = (float)a.sigma6;
    a.fcf = (float)a.fc;
    a.dfcf = (float)a.dfc;
    a.rc2f = (float)a.rc2;
    a.cap = s.nb_cap;
    a.nwork = s.n;
    a.n = s.n;
    a.wcap = window_cap(s);
    a.g = s.tiles;
    return a;
}

}  // namespace tmd
