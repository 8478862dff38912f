"""Python facade over the md_* runtime ABI, in the call order of the
reference driver (main.cpp:664-784).  The library keeps one global
simulation per process (core/runtime.impala:1-5), so a Simulation object is
a handle to that state, and creating a new one replaces the previous one.
"""
import ctypes
import os

import numpy as np

from ._lib import Vector3D, check, load

PRECISION = {"fp64": 0, "fp32mixed": 1, "fp32": 1, "fp32kahan": 2}
HALF_MODE = {"ghost_full": 0, "ref_literal": 1}
GHOST_MODE = {"fresh": 0, "ref_lag": 1}
KERNELS = {"force": 0, "integrate1": 1, "integrate2": 2, "ghost_sync": 3, "neighbor": 4, "bin": 5,
           "borders": 6, "thermo": 7, "adjoint": 8, "mirror": 9}


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _vptr(a):
    return a.ctypes.data_as(ctypes.POINTER(Vector3D))


def world_aabb(nx, ny, nz, potmin=1.6796):
    """main.cpp:515-518"""
    return np.array([0.0, nx * potmin, 0.0, ny * potmin, 0.0, nz * potmin])


def generate_fcc(nx, ny, nz, potmin=1.6796, half=False):
    """FCC lattice + Park-Miller velocities (initialize.h:91-224) from the library."""
    L = load()
    cap = 4 * nx * ny * nz + 64
    m = np.zeros(cap)
    x = np.zeros((cap, 3))
    v = np.zeros((cap, 3))
    n = L.tmd_generate_fcc(nx, ny, nz, potmin, int(half), _dptr(m), _vptr(x), _vptr(v), cap)
    return m[:n].copy(), x[:n].copy(), v[:n].copy()


def generate_granular_gas(nc, potmin=1.6796, seed=12345, vmax=1.0):
    L = load()
    cap = 4 * nc ** 3 + 64
    m = np.zeros(cap)
    x = np.zeros((cap, 3))
    v = np.zeros((cap, 3))
    n = L.tmd_generate_granular_gas(nc, potmin, seed, vmax, _dptr(m), _vptr(x), _vptr(v), cap)
    return m[:n].copy(), x[:n].copy(), v[:n].copy()


def device_info():
    buf = ctypes.create_string_buffer(256)
    cus = load().tmd_device_info(buf, 256)
    return buf.value.decode(), cus


class Simulation:
    def __init__(self, masses, positions, velocities, aabb, cell_spacing, rank_aabb=None, nb_cap=100,
                 cell_cap=60, precision="fp64", half_mode="ghost_full", ghost_mode="fresh", lj_shift=False,
                 dem_radius=None, sync=True, sort=False):
        self.L = L = load()
        L.tmd_clear_error()
        L.md_deallocate_grid()
        self.set_modes(precision, half_mode, ghost_mode, lj_shift, dem_radius, sync)
        L.tmd_set_sort_particles(int(bool(sort)))
        # tile shape and block size are library-wide settings: a new
        # simulation starts from the automatic ones (set_tile after this)
        L.tmd_set_tile(0, 0, 0)
        L.tmd_set_tile_threads(0)
        m = np.ascontiguousarray(masses, dtype=np.float64)
        x = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(velocities, dtype=np.float64).reshape(-1, 3)
        self.aabb = np.ascontiguousarray(aabb, dtype=np.float64)
        ra = self.aabb if rank_aabb is None else np.ascontiguousarray(rank_aabb, dtype=np.float64)
        self.cell_spacing = cell_spacing
        n = L.md_initialize_grid(_dptr(m), _vptr(x), _vptr(v), len(m), _dptr(self.aabb), _dptr(ra),
                                 cell_spacing, cell_cap, nb_cap)
        check()
        self.n = n

    # ---- modes
    def set_modes(self, precision="fp64", half_mode="ghost_full", ghost_mode="fresh", lj_shift=False,
                  dem_radius=None, sync=True):
        L = self.L
        L.tmd_set_precision(PRECISION[precision])
        L.tmd_set_adjoint_precision(-1)
        L.tmd_set_half_mode(HALF_MODE[half_mode])
        L.tmd_set_ghost_mode(GHOST_MODE[ghost_mode])
        L.tmd_set_lj_shift(int(bool(lj_shift)))
        L.tmd_set_dem_radius(0.0 if dem_radius is None else float(dem_radius))
        L.tmd_set_sync(int(bool(sync)))

    def set_precision(self, precision):
        self.L.tmd_set_precision(PRECISION[precision])

    def set_adjoint_precision(self, precision=None):
        """Precision of the adjoint kernel alone; None follows the forward."""
        self.L.tmd_set_adjoint_precision(-1 if precision is None else PRECISION[precision])

    def set_tile(self, tx=0, ty=None, tz=None, threads=None):
        """Cells per tile (0: automatic) and threads per tile block (0:
        automatic); the tile applies from the next binning."""
        ty = tx if ty is None else ty
        tz = tx if tz is None else tz
        self.L.tmd_set_tile(tx, ty, tz)
        if threads is not None:
            self.L.tmd_set_tile_threads(threads)
        check()

    @classmethod
    def restore(cls, path, sync=True):
        """A simulation read from a restart file written by save()
        (tmd_load_state): the saved atoms, geometry and modes, on the device;
        ghosts and lists are built by the next setup()."""
        self = cls.__new__(cls)
        self.L = L = load()
        L.tmd_clear_error()
        L.md_deallocate_grid()
        L.tmd_set_sync(int(bool(sync)))
        L.tmd_set_adjoint_precision(-1)
        L.tmd_set_sort_particles(0)
        L.tmd_set_tile(0, 0, 0)
        L.tmd_set_tile_threads(0)
        self.n = L.tmd_load_state(os.fsencode(path))
        check()
        self.aabb = None
        self.cell_spacing = None
        return self

    def save(self, path):
        """Restart file of the current state (tmd_save_state)."""
        n = self.L.tmd_save_state(os.fsencode(path))
        check()
        return n

    def write_vtk(self, path, what="locals"):
        """The reference's legacy VTK particle file of the host copy
        (call copy_from_accelerator first): what = locals, ghosts or aabb."""
        n = self.L.tmd_write_vtk(os.fsencode(path), {"locals": 0, "ghosts": 1, "aabb": 2}[what])
        check()
        return n

    @classmethod
    def fcc(cls, nx, ny=None, nz=None, potmin=1.6796, temp=1.44, half_benchmark=False, cutoff=2.5,
            skin=0.3, **kw):
        """init_rectangular_grid + md_create_velocity + upload (main.cpp:670-677)."""
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        m, x, v = generate_fcc(nx, ny, nz, potmin, half_benchmark)
        sim = cls(m, x, v, world_aabb(nx, ny, nz, potmin), cutoff + skin, **kw)
        sim.create_velocity(temp)
        sim.copy_to_accelerator()
        return sim

    # ---- ABI calls
    def create_velocity(self, t):
        self.L.md_create_velocity(t); check()

    def copy_to_accelerator(self):
        self.L.md_copy_data_to_accelerator(); check()

    def copy_from_accelerator(self):
        self.L.md_copy_data_from_accelerator(); check()

    def exchange(self):
        self.L.md_exchange_particles(); check()

    def enforce_pbc(self):
        self.L.md_enforce_pbc(); check()

    def borders(self):
        self.L.md_borders(); check()

    def distribute(self):
        self.L.md_distribute_particles(); check()

    def assemble_neighborlists(self, half=False, cutoff=2.8):
        self.L.md_assemble_neighborlists(bool(half), cutoff); check()

    def lennard_jones(self, half=False, cutoff=2.5, eps=1.0, sigma=1.0):
        self.L.md_compute_lennard_jones(bool(half), cutoff, eps, sigma); check()

    def dem(self, half=True, cutoff=1.0, damping_n=0.5, damping_t=0.2, stiffness=100.0, friction=0.3):
        self.L.md_compute_dem(bool(half), cutoff, damping_n, damping_t, stiffness, friction); check()

    def initial_integration(self, dt):
        self.L.md_initial_integration(dt); check()

    def final_integration(self, dt):
        self.L.md_final_integration(dt); check()

    def synchronize_ghost_layer(self):
        self.L.md_synchronize_ghost_layer(); check()

    def setup(self, half=False, cutoff=2.5, skin=0.3):
        """main.cpp:691-694"""
        self.exchange()
        self.borders()
        self.distribute()
        self.assemble_neighborlists(half, cutoff + skin)

    def run_abi_lj(self, steps, reneigh=20, dt=0.005, half=False, cutoff=2.5, skin=0.3, eps=1.0, sigma=1.0):
        """The reference driver's timestep loop through the md_* calls (main.cpp:715-762)."""
        for j in range(steps):
            self.initial_integration(dt)
            if (j + 1) % reneigh == 0:
                self.exchange()
                self.enforce_pbc()
                self.borders()
                self.distribute()
                self.assemble_neighborlists(half, cutoff + skin)
            else:
                self.synchronize_ghost_layer()
            self.lennard_jones(half, cutoff, eps, sigma)
            self.final_integration(dt)

    def run_lj(self, steps, reneigh=20, dt=0.005, half=False, cutoff=2.5, skin=0.3, eps=1.0, sigma=1.0):
        """Same loop inside the library (tmd_run_lj), synchronising only at rebuilds."""
        done = self.L.tmd_run_lj(steps, reneigh, dt, bool(half), cutoff, skin, eps, sigma)
        check()
        return done

    def run_dem(self, steps, reneigh=20, dt=0.001, half=True, cutoff=1.0, skin=0.3, damping_n=0.5,
                damping_t=0.2, stiffness=100.0, friction=0.3):
        done = self.L.tmd_run_dem(steps, reneigh, dt, bool(half), cutoff, skin, damping_n, damping_t,
                                  stiffness, friction)
        check()
        return done

    # ---- differentiable trajectories
    def traj_loss(self, steps, reneigh=20, dt=0.005, cutoff=2.5, skin=0.3, eps=1.0, sigma=1.0):
        """(Ubar, Pbar) of an LJ trajectory from the current state; the state is left unchanged."""
        out = np.zeros(2)
        self.L.tmd_traj_loss(steps, reneigh, dt, cutoff, skin, eps, sigma, _dptr(out))
        check()
        return float(out[0]), float(out[1])

    def traj_adjoint(self, steps, reneigh=20, dt=0.005, cutoff=2.5, skin=0.3, eps=1.0, sigma=1.0,
                     u_target=0.0, p_target=0.0, grad_state=True):
        """Loss L = (Ubar - u_target)^2 + (Pbar - p_target)^2 and its gradient.

        Returns dict(loss, ubar, pbar, grad_eps, grad_sigma, grad_x0, grad_v0)."""
        out = np.zeros(3)
        gt = np.zeros(2)
        n = self.nlocal
        gx = np.zeros((n, 3)) if grad_state else None
        gv = np.zeros((n, 3)) if grad_state else None
        self.L.tmd_traj_adjoint(steps, reneigh, dt, cutoff, skin, eps, sigma, u_target, p_target, _dptr(out),
                                _dptr(gt), None if gx is None else _vptr(gx), None if gv is None else _vptr(gv))
        check()
        return dict(loss=float(out[0]), ubar=float(out[1]), pbar=float(out[2]), grad_eps=float(gt[0]),
                    grad_sigma=float(gt[1]), grad_x0=gx, grad_v0=gv)

    def barrier(self):
        self.L.md_barrier(); check()

    # ---- state access
    @property
    def nlocal(self):
        return self.L.md_get_number_of_particles()

    @property
    def nghost(self):
        return self.L.md_get_number_of_ghost_particles()

    def state(self):
        """(m, x, v, f) of the locals after a device -> host copy."""
        self.copy_from_accelerator()
        n = self.nlocal
        m = np.zeros(n); x = np.zeros((n, 3)); v = np.zeros((n, 3)); f = np.zeros((n, 3))
        self.L.md_write_grid_data_to_arrays(_dptr(m), _vptr(x), _vptr(v), _vptr(f))
        check()
        return m, x, v, f

    def ghost_state(self):
        self.copy_from_accelerator()
        ng = self.nghost
        m = np.zeros(ng); x = np.zeros((ng, 3)); v = np.zeros((ng, 3)); f = np.zeros((ng, 3))
        self.L.md_write_grid_ghost_data_to_arrays(_dptr(m), _vptr(x), _vptr(v), _vptr(f))
        check()
        return m, x, v, f

    def grid_info(self):
        out = np.zeros(7, dtype=np.int32)
        self.L.tmd_get_grid_info(_iptr(out))
        return dict(zip(["nx", "ny", "nz", "ncells", "n", "ng", "nb_cap"], [int(a) for a in out]))

    def tile_info(self):
        out = np.zeros(7, dtype=np.int32)
        self.L.tmd_get_tile_info(_iptr(out))
        return dict(zip(["tx", "ty", "tz", "ntiles", "threads", "max_window", "stage_bytes"], [int(a) for a in out]))

    def ghost_table(self):
        ng = self.nghost
        owner = np.zeros(ng, dtype=np.int32); img = np.zeros((ng, 3), dtype=np.int32)
        src = np.zeros(ng, dtype=np.int32); flag = np.zeros((ng, 3), dtype=np.int32)
        self.L.tmd_get_ghost_table(_iptr(owner), _iptr(img), _iptr(src), _iptr(flag))
        check()
        return dict(owner=owner, img=img, src=src, flag=flag)

    def cells(self):
        inf = self.grid_info()
        start = np.zeros(inf["ncells"] + 1, dtype=np.int32)
        nb = self.L.tmd_get_cells(_iptr(start), None, None)
        parts = np.zeros(max(nb, 1), dtype=np.int32)
        pcell = np.zeros(inf["n"] + inf["ng"], dtype=np.int32)
        self.L.tmd_get_cells(_iptr(start), _iptr(parts), _iptr(pcell))
        check()
        return start, parts[:nb], pcell

    def neighbor_lists(self, with_lists=True):
        """(counts[n], lists[n, cap]) by atom index; lists is None when not requested."""
        inf = self.grid_info()
        n, cap = inf["n"], inf["nb_cap"]
        counts = np.zeros(n, dtype=np.int32)
        lst = np.zeros((n, cap), dtype=np.int32) if with_lists else None
        self.L.tmd_get_neighbor_lists(_iptr(counts), None if lst is None else _iptr(lst), cap)
        check()
        return counts, lst

    def thermo(self, cutoff=2.5, eps=1.0, sigma=1.0, eshift=False):
        """T, KE, PE, virial, P, pairs; eshift: PE of the energy-shifted
        potential U(r) - U(rc) (for energy-drift bookkeeping)."""
        out = np.zeros(6)
        (self.L.tmd_thermo_eshift if eshift else self.L.tmd_thermo)(cutoff, eps, sigma, _dptr(out))
        check()
        return dict(zip(["T", "KE", "PE", "virial", "P", "pairs"], out.tolist()))

    # ---- profiling
    def profile(self, on=True):
        self.L.tmd_profile_enable(int(on))

    def profile_reset(self):
        self.L.tmd_profile_reset()

    def profile_get(self, kernel):
        ms = ctypes.c_double(0.0)
        cnt = ctypes.c_longlong(0)
        self.L.tmd_profile_get(KERNELS[kernel], ctypes.byref(ms), ctypes.byref(cnt))
        return ms.value, cnt.value

    def close(self):
        self.L.md_deallocate_grid()
