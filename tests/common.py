"""Shared set-up for parity tests: the same seeded inputs go to the HIP
library (through the C ABI) and to the CPU oracle (the checker)."""
import itertools
import sys

import numpy as np

import pyoracle as po

POTMIN = 1.6796
RC, SKIN = 2.5, 0.3


def announce(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def perturbed_fcc(nx, ny=None, nz=None, amp=0.1, seed=0, temp=1.44, potmin=POTMIN):
    """FCC lattice with seeded displacements (breaks the lattice symmetry)
    and Park-Miller velocities rescaled to temp.  Returns (aabb, m, x, v)."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    aabb, m, x, v = po.lattice(nx, ny, nz, potmin)
    rng = np.random.default_rng(seed)
    x = x + rng.uniform(-amp, amp, size=x.shape)
    L = aabb[1::2] - aabb[0::2]
    x = np.mod(x, L)  # keep inside [0, L)
    # reference velocity creation on the oracle side fixes v
    s = po.OracleSim(m, x, v, aabb, RC + SKIN)
    s.create_velocity(temp)
    _, v2, _ = s.state()
    s.close()
    return aabb, m, x, v2


def make_pair(aabb, m, x, v, cutoff=RC + SKIN, nb_cap=100, threads=4, **modes):
    """(library Simulation, OracleSim) from identical inputs."""
    import tinymd_amd as tm
    sim = tm.Simulation(m, x, v, aabb, cutoff, nb_cap=nb_cap, **modes)
    sim.copy_to_accelerator()
    orc = po.OracleSim(m, x, v, aabb, cutoff, nb_cap=nb_cap, threads=threads)
    if modes.get("ghost_mode") == "ref_lag":
        orc.set_ghost_mode(po.GHOST_REF_LAG)
    return sim, orc


def force_error(f, fref):
    """(rms error / rms force, max abs error / rms force)."""
    rms = np.sqrt(np.mean(np.sum(fref * fref, axis=1)))
    d = f - fref
    return np.sqrt(np.mean(np.sum(d * d, axis=1))) / rms, np.max(np.abs(d)) / rms


def force_error_atom(f, fref, floor=None):
    """Per-atom relative force error |F_i - F_ref,i| / max(|F_ref,i|, floor),
    returned as (rms over atoms, max over atoms).  The floor defaults to the
    RMS of |F_ref,i| over the atoms, so an atom whose force nearly cancels is
    measured against the typical force rather than against its own
    near-zero magnitude."""
    fn = np.sqrt(np.sum(fref * fref, axis=1))
    if floor is None:
        floor = np.sqrt(np.mean(fn * fn))
    d = np.sqrt(np.sum((f - fref) ** 2, axis=1)) / np.maximum(fn, floor)
    return float(np.sqrt(np.mean(d * d))), float(np.max(d))


# the FP32-mixed bar against FP64: per-atom relative force error (floor: the
# RMS force) at most 1e-5, both as RMS and as maximum over the atoms
FP32_BAR = 1e-5


def boundary_pairs(seed=5, targets=None):
    """Pairs whose squared distance, formed as the reference forms it
    (dx*dx + dy*dy + dz*dz in FP64, no contraction), lands exactly on, one
    ulp below or one ulp above r_nb^2 = (rc + skin)^2 (the list keeps
    rsq <= r_nb^2, core/neighborlist.impala) and rc^2 (the force keeps
    rsq < rc^2, core/compute.impala).  27 anchors 8 apart in a 30^3 box get
    one partner each, placed along a seeded random direction and then moved
    by ulp steps of its coordinates until the target is hit.  No atom lies
    within r_nb of a face, so there are no ghosts and no periodic image in
    range.  Other squared distances to hit can be given as targets, a list
    of (name, rsq).  Returns (aabb, m, x, v, pairs), pairs = [(anchor,
    partner, target name, rsq)]; atoms 0..26 are the anchors, 27..53 the
    partners."""
    if targets is None:
        cut = RC + SKIN
        targets = []
        for name, t in (("cut2", cut * cut), ("rc2", RC * RC)):
            targets += [(name + "_eq", t), (name + "_lo", np.nextafter(t, 0.0)), (name + "_hi", np.nextafter(t, 10.0))]
    rng = np.random.default_rng(seed)
    steps = range(-6, 7)
    anchors, partners, pairs = [], [], []
    for k, (a, b, c) in enumerate(itertools.product(range(3), repeat=3)):
        anchor = np.array([6.0 + 8.0 * a, 6.0 + 8.0 * b, 6.0 + 8.0 * c])
        name, goal = targets[k % len(targets)]
        found = None
        while found is None:
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            start = anchor + u * np.sqrt(goal)
            # each coordinate moved by s ulps, s in -6..6
            cand = []
            for d in range(3):
                row = {0: start[d]}
                up = dn = start[d]
                for s_ in range(1, 7):
                    up = np.nextafter(up, np.inf)
                    dn = np.nextafter(dn, -np.inf)
                    row[s_], row[-s_] = up, dn
                cand.append(row)
            for sx, sy, sz in itertools.product(steps, steps, steps):
                p = np.array([cand[0][sx], cand[1][sy], cand[2][sz]])
                dd = anchor - p
                if dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2] == goal:
                    found = p
                    break
        anchors.append(anchor)
        partners.append(found)
        pairs.append((k, 27 + k, name, float(goal)))
    x = np.vstack(anchors + partners)
    aabb = np.array([0.0, 30.0, 0.0, 30.0, 0.0, 30.0])
    return aabb, np.ones(len(x)), x, np.zeros_like(x), pairs


def fma_rsq(xi, xj):
    """The squared distance as the FP32 force variants form it in FP64,
    fma(dx, dx, fma(dy, dy, dz * dz)), each fused step rounded once: computed
    exactly with rationals and rounded to the nearest double."""
    from fractions import Fraction
    d = xi - xj
    dz2 = float(Fraction(d[2]) * Fraction(d[2]))
    t = float(Fraction(d[1]) * Fraction(d[1]) + Fraction(dz2))
    return float(Fraction(d[0]) * Fraction(d[0]) + Fraction(t))


def brute_force_rsq(x):
    """All-pairs squared distances formed as the reference forms them."""
    d = x[:, None, :] - x[None, :, :]
    return d[:, :, 0] * d[:, :, 0] + d[:, :, 1] * d[:, :, 1] + d[:, :, 2] * d[:, :, 2]
