"""Shared set-up for parity tests: the same seeded inputs go to the HIP
library (through the C ABI) and to the CPU oracle (the checker)."""
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
