"""ctypes wrapper over the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker.  The product (molecular-dynamics_amd/) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libtinymd_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libtinymd_ref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_d = ctypes.c_double
_i = ctypes.c_int

GHOST_FRESH, GHOST_REF_LAG = 0, 1
HALF_GHOST_FULL, HALF_REF_LITERAL = 0, 1


def build(force=False):
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "liboracle"], cwd=HERE)


def _ptr(a, t=_dp):
    if a is None:
        return None
    return a.ctypes.data_as(t)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "orc_pm_random": (_d, [_ip]),
            "orc_lattice": (_i, [_dp, _dp, _i, _d, _dp, _dp, _dp, _i]),
            "orc_create": (_vp, [_dp, _dp, _dp, _i, _dp, _dp, _d, _i]),
            "orc_destroy": (None, [_vp]),
            "orc_set_threads": (None, [_vp, _i]),
            "orc_set_ghost_mode": (None, [_vp, _i]),
            "orc_grid_info": (None, [_vp, _ip]),
            "orc_create_velocity": (None, [_vp, _d]),
            "orc_exchange": (None, [_vp]),
            "orc_enforce_pbc": (None, [_vp]),
            "orc_set_sort": (None, [_vp, _i]),
            "orc_borders": (None, [_vp]),
            "orc_distribute": (None, [_vp]),
            "orc_neighbors": (None, [_vp, _i, _i, _d]),
            "orc_sync_ghosts": (None, [_vp]),
            "orc_initial_integration": (None, [_vp, _d]),
            "orc_final_integration": (None, [_vp, _d]),
            "orc_lj": (None, [_vp, _i, _d, _d, _d, _i]),
            "orc_dem": (None, [_vp, _i, _d, _d, _d, _d, _d, _d]),
            "orc_thermo": (None, [_vp, _dp]),
            "orc_get_state": (None, [_vp, _dp, _dp, _dp]),
            "orc_set_state": (None, [_vp, _dp, _dp]),
            "orc_get_ghosts": (None, [_vp, _ip, _ip, _ip, _ip, _dp]),
            "orc_get_cells": (None, [_vp, _ip, _ip, _ip]),
            "orc_get_neighbors": (None, [_vp, _ip, _ip, _i]),
            "orc_run_lj": (_i, [_vp, _i, _i, _d, _i, _i, _d, _d, _d, _d, _i]),
            "orc_run_dem": (_i, [_vp, _i, _i, _d, _i, _i, _d, _d, _d, _d, _d, _d, _d]),
            "orc_setup": (None, [_vp, _i, _i, _d]),
            "orc_traj_thermo": (None, [_vp, _i, _i, _d, _d, _d, _d, _d, _i, _dp]),
            "orc_traj_adjoint": (None, [_vp, _i, _i, _d, _d, _d, _d, _d, _i, _d, _d, _dp, _dp, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ref_lib():
    """The reference's own initialize.h / core/force_lj.cpp, built by
    `make ref` from the reference tree.  None when not built."""
    if not os.path.exists(REF_PATH):
        return None
    L = ctypes.CDLL(REF_PATH)
    L.ref_random.restype = _d
    L.ref_random.argtypes = [_ip]
    L.ref_generate_rectangular_grid.restype = _i
    L.ref_generate_rectangular_grid.argtypes = [_dp, _dp, _i, _d, _dp, _dp, _dp, _i]
    L.compute_fullneigh.restype = None
    L.compute_fullneigh.argtypes = [_i, _d, _dp, _dp, _ip, _ip, _i]
    return L


def world_aabb(nx, ny, nz, potmin=1.6796):
    return np.array([0.0, nx * potmin, 0.0, ny * potmin, 0.0, nz * potmin])


def lattice(nx, ny, nz, potmin=1.6796, half=False):
    """FCC lattice as init_rectangular_grid builds it (initialize.h:203-224)."""
    aabb = world_aabb(nx, ny, nz, potmin)
    cap = 4 * nx * ny * nz + 64
    m = np.zeros(cap)
    x = np.zeros((cap, 3))
    v = np.zeros((cap, 3))
    n = lib().orc_lattice(_ptr(aabb), _ptr(aabb), int(half), potmin * 0.5, _ptr(m), _ptr(x), _ptr(v), cap)
    return aabb, m[:n].copy(), x[:n].copy(), v[:n].copy()


class OracleSim:
    """One oracle simulation; mirrors the md_* call sequence of main.cpp."""

    def __init__(self, m, x, v, aabb, cell_spacing, nb_cap=100, threads=1, rank_aabb=None):
        L = lib()
        self.aabb = np.ascontiguousarray(aabb, dtype=np.float64)
        ra = self.aabb if rank_aabb is None else np.ascontiguousarray(rank_aabb, dtype=np.float64)
        m = np.ascontiguousarray(m, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        self.h = L.orc_create(_ptr(m), _ptr(x), _ptr(v), len(m), _ptr(self.aabb), _ptr(ra), cell_spacing, nb_cap)
        L.orc_set_threads(self.h, threads)
        self.L = L

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        out = np.zeros(7, dtype=np.int32)
        self.L.orc_grid_info(self.h, _ptr(out, _ip))
        keys = ["nx", "ny", "nz", "ncells", "n", "ng", "nb_cap"]
        return dict(zip(keys, [int(a) for a in out]))

    @property
    def n(self):
        return self.info()["n"]

    @property
    def ng(self):
        return self.info()["ng"]

    def set_ghost_mode(self, mode):
        self.L.orc_set_ghost_mode(self.h, mode)

    def set_sort(self, on=True):
        """Renumber the locals in cell order at each wrap (orc_set_sort)."""
        self.L.orc_set_sort(self.h, int(bool(on)))

    def create_velocity(self, t):
        self.L.orc_create_velocity(self.h, t)

    def setup(self, half=False, half_mode=HALF_GHOST_FULL, cutoff=2.8):
        self.L.orc_setup(self.h, int(half), half_mode, cutoff)

    def exchange(self):
        self.L.orc_exchange(self.h)

    def enforce_pbc(self):
        self.L.orc_enforce_pbc(self.h)

    def borders(self):
        self.L.orc_borders(self.h)

    def distribute(self):
        self.L.orc_distribute(self.h)

    def neighbors(self, half=False, half_mode=HALF_GHOST_FULL, cutoff=2.8):
        self.L.orc_neighbors(self.h, int(half), half_mode, cutoff)

    def sync_ghosts(self):
        self.L.orc_sync_ghosts(self.h)

    def initial_integration(self, dt):
        self.L.orc_initial_integration(self.h, dt)

    def final_integration(self, dt):
        self.L.orc_final_integration(self.h, dt)

    def lj(self, half=False, rc=2.5, eps=1.0, sigma=1.0, shift=0):
        self.L.orc_lj(self.h, int(half), rc, eps, sigma, shift)

    def dem(self, half=True, rc=1.0, radius=0.5, gn=0.5, gt=0.2, k=100.0, mu=0.3):
        self.L.orc_dem(self.h, int(half), rc, radius, gn, gt, k, mu)

    def thermo(self):
        out = np.zeros(6)
        self.L.orc_thermo(self.h, _ptr(out))
        return dict(zip(["T", "KE", "PE", "virial", "P", "pairs"], out.tolist()))

    def state(self):
        n = self.n
        x = np.zeros((n, 3)); v = np.zeros((n, 3)); f = np.zeros((n, 3))
        self.L.orc_get_state(self.h, _ptr(x), _ptr(v), _ptr(f))
        return x, v, f

    def set_state(self, x=None, v=None):
        xa = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        va = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        self.L.orc_set_state(self.h, _ptr(xa), _ptr(va))

    def ghosts(self):
        ng = self.ng
        owner = np.zeros(ng, dtype=np.int32); img = np.zeros((ng, 3), dtype=np.int32)
        src = np.zeros(ng, dtype=np.int32); flag = np.zeros((ng, 3), dtype=np.int32)
        pos = np.zeros((ng, 3))
        self.L.orc_get_ghosts(self.h, _ptr(owner, _ip), _ptr(img, _ip), _ptr(src, _ip), _ptr(flag, _ip), _ptr(pos))
        return dict(owner=owner, img=img, src=src, flag=flag, pos=pos)

    def cells(self):
        inf = self.info()
        start = np.zeros(inf["ncells"] + 1, dtype=np.int32)
        pcell = np.zeros(inf["n"] + inf["ng"], dtype=np.int32)
        self.L.orc_get_cells(self.h, _ptr(start, _ip), None, _ptr(pcell, _ip))
        parts = np.zeros(int(start[-1]), dtype=np.int32)
        self.L.orc_get_cells(self.h, None, _ptr(parts, _ip), None)
        return start, parts, pcell

    def neighbor_lists(self):
        inf = self.info()
        n, cap = inf["n"], inf["nb_cap"]
        counts = np.zeros(n, dtype=np.int32)
        lst = np.zeros((n, cap), dtype=np.int32)
        self.L.orc_get_neighbors(self.h, _ptr(counts, _ip), _ptr(lst, _ip), cap)
        return counts, lst

    def run_lj(self, steps, reneigh=20, dt=0.005, half=False, half_mode=HALF_GHOST_FULL,
               rc=2.5, skin=0.3, eps=1.0, sigma=1.0, shift=0):
        return self.L.orc_run_lj(self.h, steps, reneigh, dt, int(half), half_mode, rc, skin, eps, sigma, shift)

    def traj_thermo(self, steps, reneigh=20, dt=0.005, rc=2.5, skin=0.3, eps=1.0, sigma=1.0, shift=1):
        """(Ubar, Pbar): time averages of PE/N and P over `steps` steps."""
        out = np.zeros(2)
        self.L.orc_traj_thermo(self.h, steps, reneigh, dt, rc, skin, eps, sigma, shift, _ptr(out))
        return float(out[0]), float(out[1])

    def traj_adjoint(self, steps, reneigh=20, dt=0.005, rc=2.5, skin=0.3, eps=1.0, sigma=1.0, shift=1,
                     u_target=0.0, p_target=0.0):
        """Loss (Ubar - u_target)^2 + (Pbar - p_target)^2 and its gradient; the
        state is left at x0."""
        n = self.n
        out = np.zeros(3); gt = np.zeros(2); gx = np.zeros((n, 3)); gv = np.zeros((n, 3))
        self.L.orc_traj_adjoint(self.h, steps, reneigh, dt, rc, skin, eps, sigma, shift, u_target, p_target,
                                _ptr(out), _ptr(gt), _ptr(gx), _ptr(gv))
        return dict(loss=float(out[0]), ubar=float(out[1]), pbar=float(out[2]), grad_eps=float(gt[0]),
                    grad_sigma=float(gt[1]), grad_x0=gx, grad_v0=gv)

    def run_dem(self, steps, reneigh=20, dt=0.001, half=True, half_mode=HALF_GHOST_FULL,
                rc=1.0, skin=0.3, radius=0.5, gn=0.5, gt=0.2, k=100.0, mu=0.3):
        return self.L.orc_run_dem(self.h, steps, reneigh, dt, int(half), half_mode, rc, skin, radius, gn, gt, k, mu)


def lj_benchmark(nc, potmin=1.6796, temp=1.44, half=False, half_mode=HALF_GHOST_FULL,
                 rc=2.5, skin=0.3, threads=1, nb_cap=100):
    """main.cpp's default benchmark set-up (main.cpp:664-700) on the oracle."""
    aabb, m, x, v = lattice(nc, nc, nc, potmin)
    s = OracleSim(m, x, v, aabb, rc + skin, nb_cap=nb_cap, threads=threads)
    s.create_velocity(temp)
    s.setup(half, half_mode, rc + skin)
    s.lj(half, rc)
    return s
