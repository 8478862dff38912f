"""Batched replicas: many independent periodic LJ boxes advanced together
in one hipGraph launch per run (tmd_batch_* in libtinymd_amd.so)."""
import ctypes

import numpy as np

from ._lib import check, load
from .sim import PRECISION, _dptr, _iptr, _vptr


def replica_set(seed=2024, nrep=64, cells=(6, 24), rho=(0.70, 0.90), temp=(0.7, 2.0)):
    """Per-replica (cells per side, density, temperature) from the library's
    counter-based generator."""
    L = load()
    c = np.zeros(nrep, dtype=np.int32)
    r = np.zeros(nrep)
    t = np.zeros(nrep)
    L.tmd_generate_replica_set(seed, nrep, cells[0], cells[1], rho[0], rho[1], temp[0], temp[1], _iptr(c), _dptr(r),
                               _dptr(t))
    return c, r, t


def scale_velocities(v, temp):
    """Remove the mean velocity and rescale to T = sum v^2 / (3N - 3), as
    md_create_velocity does (core/runtime.impala:158-186)."""
    v = v - v.mean(axis=0)
    t = np.sum(v * v) / (3 * len(v) - 3)
    return v * np.sqrt(temp / t)


_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix(z):
    """SplitMix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    z = (z + np.uint64(0x9E3779B97F4A7C15)) & _M64
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return z ^ (z >> np.uint64(31))


def fcc_box(nc, a, seed):
    """FCC lattice of nc^3 cubic cells (4 nc^3 sites (i, j, k) a / 2 with
    i + j + k even, k-major order) and velocities uniform in [-0.5, 0.5)
    from a counter-based hash of (seed, site, component)."""
    r = np.arange(2 * nc)
    k, j, i = np.meshgrid(r, r, r, indexing="ij")
    keep = (i + j + k) % 2 == 0
    ijk = np.stack([i[keep], j[keep], k[keep]], axis=1).astype(np.float64)
    x = ijk * (a * 0.5)
    site = np.arange(len(x), dtype=np.uint64)[:, None] * np.uint64(3) + np.arange(3, dtype=np.uint64)[None, :]
    with np.errstate(over="ignore"):
        h = _splitmix(np.uint64(seed) ^ _splitmix(site))
    v = (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0) - 0.5
    return x, v


def replica_boxes(seed=2024, nrep=64, **kw):
    """[(aabb, x, v)] FCC boxes: nc^3 unit cells at lattice constant
    (4 / rho)^(1/3) (4 nc^3 atoms), velocities from fcc_box with seed
    seed * 1000 + r, rescaled to the replica's temperature."""
    cells, rho, temp = replica_set(seed, nrep, **kw)
    boxes = []
    for r, (nc, rh, tt) in enumerate(zip(cells, rho, temp)):
        a = (4.0 / rh) ** (1.0 / 3.0)
        x, v = fcc_box(int(nc), a, seed * 1000 + r)
        aabb = np.array([0.0, nc * a, 0.0, nc * a, 0.0, nc * a])
        boxes.append((aabb, x, scale_velocities(v, tt)))
    return boxes


class Batch:
    def __init__(self, boxes, cutoff=2.5, skin=0.3, eps=1.0, sigma=1.0, dt=0.005, reneigh=20, precision="fp32mixed",
                 lj_shift=False):
        self.L = L = load()
        self.counts = np.array([len(b[1]) for b in boxes], dtype=np.int32)
        x = np.ascontiguousarray(np.concatenate([b[1] for b in boxes]), dtype=np.float64)
        v = np.ascontiguousarray(np.concatenate([b[2] for b in boxes]), dtype=np.float64)
        aabbs = np.ascontiguousarray(np.concatenate([b[0] for b in boxes]), dtype=np.float64)
        self.h = L.tmd_batch_create(len(boxes), _iptr(self.counts), _vptr(x), _vptr(v), _dptr(aabbs))
        check()
        L.tmd_batch_setup(self.h, cutoff, skin, eps, sigma, dt, reneigh, PRECISION[precision], int(bool(lj_shift)))
        check()

    @property
    def nrep(self):
        return len(self.counts)

    def configure(self, tile=0, threads=0, spacing=1.0, list_headroom=1.2, ghost_headroom=1.5):
        self.L.tmd_batch_configure(self.h, tile, threads, spacing, list_headroom, ghost_headroom)
        check()

    def autotune(self, trial_steps=0):
        out = np.zeros(5)
        self.L.tmd_batch_autotune(self.h, trial_steps, _dptr(out))
        check()
        return dict(tile=int(out[0]), threads=int(out[1]), spacing=float(out[2]), list_headroom=float(out[3]),
                    ms_per_step=float(out[4]))

    def reset(self):
        self.L.tmd_batch_reset(self.h)
        check()

    def set_graph(self, on=True):
        """True (default): each run is one hipGraph launch; False: the same
        kernels are launched one by one (for per-launch tracing tools)."""
        self.L.tmd_batch_set_graph(self.h, int(bool(on)))
        check()

    def run(self, steps):
        done = self.L.tmd_batch_run(self.h, steps)
        check()
        return done

    def run_timed(self, steps):
        ms = ctypes.c_double(0.0)
        self.L.tmd_batch_run_timed(self.h, steps, ctypes.byref(ms))
        check()
        return ms.value

    def state(self, rep):
        n = int(self.counts[rep])
        x = np.zeros((n, 3)); v = np.zeros((n, 3)); f = np.zeros((n, 3))
        self.L.tmd_batch_get_state(self.h, rep, _vptr(x), _vptr(v), _vptr(f))
        check()
        return x, v, f

    def thermo(self):
        out = np.zeros((self.nrep, 6))
        self.L.tmd_batch_thermo(self.h, _dptr(out))
        check()
        keys = ["T", "KE", "PE", "virial", "P", "pairs"]
        return [dict(zip(keys, row.tolist())) for row in out]

    def info(self):
        out = np.zeros((self.nrep, 8), dtype=np.int32)
        self.L.tmd_batch_info(self.h, _iptr(out))
        check()
        keys = ["n", "ghosts", "nx", "cells", "tiles", "tile", "ghost_room", "list_cap"]
        return [dict(zip(keys, [int(a) for a in row])) for row in out]

    def close(self):
        if self.h:
            self.L.tmd_batch_destroy(self.h)
            self.h = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
