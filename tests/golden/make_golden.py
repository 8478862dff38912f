"""Writes the golden vectors in this directory from the reference's own C++
sources (initialize.h, core/force_lj.cpp), compiled from the reference tree
into oracle/_ref by `make -C oracle ref`.  Run in the container that has the
reference tree; the .npz files are committed, the reference is not.

  ref_lattice_5x3x7.npz : generate_rectangular_grid(5x3x7 cells) positions and
                          Park-Miller velocities
  ref_force_lj_6cube.npz: compute_fullneigh forces on a seeded perturbed
                          6x6x6 lattice with the oracle's full list
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import pyoracle as po  # noqa: E402

RC, SKIN = 2.5, 0.3


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def main():
    ref = po.ref_lib()
    if ref is None:
        raise SystemExit("build oracle/_ref first: make -C oracle ref")
    aabb = po.world_aabb(5, 3, 7)
    cap = 4 * 5 * 3 * 7 + 64
    m = np.zeros(cap); x = np.zeros((cap, 3)); v = np.zeros((cap, 3))
    n = ref.ref_generate_rectangular_grid(_dp(aabb), _dp(aabb), 0, 1.6796 * 0.5, _dp(m), _dp(x), _dp(v), cap)
    np.savez_compressed(os.path.join(HERE, "ref_lattice_5x3x7.npz"), pos=x[:n], vel=v[:n])

    aabb, m, x, v = po.lattice(6, 6, 6)
    rng = np.random.default_rng(2024)
    x = np.mod(x + rng.uniform(-0.12, 0.12, x.shape), aabb[1::2] - aabb[0::2])
    s = po.OracleSim(m, x, v, aabb, RC + SKIN)
    s.setup(False, po.HALF_GHOST_FULL, RC + SKIN)
    counts, lst = s.neighbor_lists()
    allx = np.ascontiguousarray(np.vstack([x, s.ghosts()["pos"]]))
    f = np.zeros((len(allx), 3))
    lst = np.ascontiguousarray(lst.astype(np.int32))
    counts = np.ascontiguousarray(counts.astype(np.int32))
    ref.compute_fullneigh(len(x), RC * RC, _dp(allx), _dp(f), _ip(lst), _ip(counts), lst.shape[1])
    np.savez_compressed(os.path.join(HERE, "ref_force_lj_6cube.npz"), pos=x, force=f[:len(x)], counts=counts)


if __name__ == "__main__":
    main()
