"""GPU parity at the sizes of the five benchmark workloads, against the CPU
oracle (the checker) on the same inputs.

- Plain LJ, FCC 32^3 (131,072 atoms), FP64, full list, 100 steps with a
  rebuild every 20 (the reference driver's default run): positions,
  velocities, forces and neighbour lists bit-exact, thermo to 1e-12.
- Large LJ, FCC 80^3 (2,048,000 atoms) on a thermalised state: FP64
  full-list forces bit-exact; FP32-mixed and FP32-Kahan forces, full and
  half lists, within the per-atom relative bar 1e-5 (floor: RMS force)
  against the oracle's FP64 forces; FP64 half-list forces within 1e-12.
- Differentiable LJ, FCC 16^3 (16,384 atoms), 200 steps, checkpoints every
  20, smooth cutoff: FP64 loss and gradients within 1e-9 of the oracle's
  adjoint; FP32-mixed within 1e-5 of it.
- DEM granular gas: FP32-mixed spring-dashpot forces against the FP64
  oracle on a compressed 8^3 fixture (1e-5 per atom); at the full 64^3 size
  (528,384 spheres) no contact at t = 0 (force exactly zero), momentum
  conserved and mechanical energy (kinetic + spring) non-increasing over
  1,000 steps, and FP32-mixed against FP64 forces on the same state.
- Batched replicas: the 64-box set (seed 2024, 6..24 cells per side,
  1,334,984 atoms) in FP64, 25 steps with a rebuild at step 20, every
  replica bit-exact against the oracle run of the same box.
"""
import os

import numpy as np
import pytest

import pyoracle as po
from common import FP32_BAR, RC, SKIN, announce, force_error, force_error_atom

pytestmark = pytest.mark.gpu

DT, REN = 0.005, 20
THREADS = max(1, min(os.cpu_count() or 1, 16))


@pytest.fixture(scope="module")
def tm():
    import tinymd_amd
    name, cus = tinymd_amd.device_info()
    if cus == 0:
        pytest.fail("no HIP device visible to libtinymd_amd")
    return tinymd_amd


def lists_equal(sim, orc):
    """Per-atom counts equal and the first count entries of every list equal,
    in order.  The two sides' capacities may differ (the library rounds its
    capacity up to a multiple of 8), so only the filled columns compare."""
    cg, lg = sim.neighbor_lists()
    co, lo = orc.neighbor_lists()
    if not np.array_equal(cg, co):
        bad = np.flatnonzero(cg != co)
        announce(f"  list counts differ at {len(bad)} atoms, first {bad[:5]}: {cg[bad[:5]]} vs {co[bad[:5]]}")
        return False
    w = int(cg.max()) if len(cg) else 0
    assert lg.shape[1] >= w and lo.shape[1] >= w
    k = np.arange(w)[None, :] < cg[:, None]
    a, b = np.where(k, lg[:, :w], -1), np.where(k, lo[:, :w], -1)
    if not np.array_equal(a, b):
        bad = np.flatnonzero(np.any(a != b, axis=1))
        announce(f"  lists differ at {len(bad)} atoms, first {bad[:5]}")
        return False
    return True


def test_lj_32cube_fp64_100_steps_bitwise(tm):
    announce("LJ FCC 32^3 FP64 full list, 100 steps, rebuild every 20: bit-exact vs oracle")
    sim = tm.Simulation.fcc(32, precision="fp64")
    sim.setup(False, RC, SKIN)
    sim.lennard_jones(False, RC)
    orc = po.lj_benchmark(32, threads=THREADS)
    _, xg, vg, fg = sim.state()
    xo, vo, fo = orc.state()
    assert np.array_equal(xg, xo) and np.array_equal(vg, vo) and np.array_equal(fg, fo)
    assert sim.run_lj(100, REN, DT, False, RC, SKIN) == 100
    orc.run_lj(100, REN, DT)
    _, xg, vg, fg = sim.state()
    xo, vo, fo = orc.state()
    assert np.array_equal(xg, xo), "positions"
    assert np.array_equal(vg, vo), "velocities"
    assert np.array_equal(fg, fo), "forces"
    assert sim.nghost == orc.ng
    assert lists_equal(sim, orc), "neighbour lists"
    tg = sim.thermo(RC)
    orc.lj(False, RC)
    to = orc.thermo()
    for key in ("T", "KE", "PE", "virial", "P"):
        assert abs(tg[key] - to[key]) <= 1e-12 * max(1.0, abs(to[key])), (key, tg[key], to[key])
    assert tg["pairs"] == to["pairs"]
    sim.close()
    orc.close()


@pytest.fixture(scope="module")
def lj80(tm):
    """80^3 FCC advanced 40 FP64 steps (two rebuilds) on the GPU; the
    oracle's lists and FP64 full-list forces on the resulting positions."""
    announce("LJ FCC 80^3: 40 FP64 steps on the GPU, then the oracle on that state")
    sim = tm.Simulation.fcc(80, precision="fp64")
    sim.setup(False, RC, SKIN)
    sim.lennard_jones(False, RC)
    sim.run_lj(40, REN, DT, False, RC, SKIN)
    m, x, v, _ = sim.state()
    aabb = tm.world_aabb(80, 80, 80)
    orc = po.OracleSim(m, x, v, aabb, RC + SKIN, threads=THREADS)
    orc.setup(False, po.HALF_GHOST_FULL, RC + SKIN)
    orc.lj(False, RC)
    _, _, f64 = orc.state()
    yield sim, orc, f64
    sim.close()
    orc.close()


def test_lj_80cube_fp64_forces_and_lists_bitwise(tm, lj80):
    sim, orc, f64 = lj80
    announce("LJ 80^3 thermalised: FP64 full-list forces and lists vs oracle")
    sim.set_precision("fp64")
    sim.lennard_jones(False, RC)
    _, _, _, fg = sim.state()
    assert np.abs(f64).max() > 10.0  # the state is disordered: forces are not the lattice's zero
    assert np.array_equal(fg, f64)
    assert lists_equal(sim, orc)


@pytest.mark.parametrize("precision,half", [("fp32mixed", False), ("fp32kahan", False), ("fp32mixed", True),
                                            ("fp64", True)])
def test_lj_80cube_reduced_precision_and_half(tm, lj80, precision, half):
    sim, orc, f64 = lj80
    announce(f"LJ 80^3 thermalised: {precision} {'half' if half else 'full'} list vs oracle FP64 full")
    sim.setup(half, RC, SKIN)  # same positions: the list is rebuilt in the requested form
    sim.set_precision(precision)
    sim.lennard_jones(half, RC)
    _, _, _, fg = sim.state()
    sim.set_precision("fp64")
    if precision == "fp64":
        rms, mx = force_error(fg, f64)
        assert rms < 1e-12 and mx < 1e-11, (rms, mx)
    else:
        rms, mx = force_error_atom(fg, f64)
        announce(f"  per-atom relative error rms={rms:.3e} max={mx:.3e}")
        assert rms <= FP32_BAR and mx <= FP32_BAR, (rms, mx)
    if half:
        sim.setup(False, RC, SKIN)


# (loss and parameter gradients, state gradients relative to their largest
# component): FP64 differs from the oracle only in summation order (gather vs
# scatter form of the adjoint); FP32-mixed follows its own FP32-force
# trajectory, whose deviation the 200 chaotic steps amplify in dL/dx0, dL/dv0
@pytest.mark.parametrize("precision,tol,tol_state", [("fp64", 1e-9, 1e-8), ("fp32mixed", 1e-5, 1e-2)])
def test_adjoint_16cube_200_steps_vs_oracle(tm, precision, tol, tol_state):
    announce(f"differentiable LJ 16^3, 200 steps, checkpoints every 20, {precision} adjoint vs oracle FP64")
    aabb, m, x, v = po.lattice(16, 16, 16)
    o0 = po.OracleSim(m, x, v, aabb, RC + SKIN, threads=THREADS)
    o0.create_velocity(1.44)
    _, v, _ = o0.state()
    o0.close()

    def oracle(eps=1.0, sigma=1.0):
        o = po.OracleSim(m, x, v, aabb, RC + SKIN, threads=THREADS)
        o.setup(False, po.HALF_GHOST_FULL, RC + SKIN)
        o.lj(False, RC, eps, sigma, shift=1)
        return o

    # targets from the same initial state at eps = 1.05, sigma = 0.98
    ot = oracle(1.05, 0.98)
    u_t, p_t = ot.traj_thermo(200, REN, DT, RC, SKIN, 1.05, 0.98, shift=1)
    ot.close()
    orc = oracle()
    ro = orc.traj_adjoint(200, REN, DT, RC, SKIN, shift=1, u_target=u_t, p_target=p_t)
    orc.close()
    sim = tm.Simulation(m, x, v, aabb, RC + SKIN, precision=precision, lj_shift=True)
    sim.copy_to_accelerator()
    rg = sim.traj_adjoint(200, REN, DT, RC, SKIN, 1.0, 1.0, u_t, p_t)
    sim.close()
    announce(f"  loss {rg['loss']:.10e} vs {ro['loss']:.10e}; d/deps {rg['grad_eps']:.10e} vs {ro['grad_eps']:.10e}; "
             f"d/dsigma {rg['grad_sigma']:.10e} vs {ro['grad_sigma']:.10e}")
    assert rg["loss"] == pytest.approx(ro["loss"], rel=tol)
    assert rg["grad_eps"] == pytest.approx(ro["grad_eps"], rel=tol)
    assert rg["grad_sigma"] == pytest.approx(ro["grad_sigma"], rel=tol)
    for key in ("grad_x0", "grad_v0"):
        scale = np.max(np.abs(ro[key]))
        err = np.max(np.abs(rg[key] - ro[key]))
        announce(f"  {key}: max |difference| / max |oracle| = {err / scale:.3e}")
        assert err <= tol_state * scale, (key, err, scale)


def dem_fixture(tm):
    m, x, v = tm.generate_granular_gas(8, seed=7, vmax=1.0)
    aabb = np.array([0.0, 8 * 1.6796, 0.0, 8 * 1.6796, 0.0, 8 * 1.6796])
    # compress the lattice so that spheres of diameter 1.0 overlap
    return m, x * 0.55, v, aabb * 0.55


@pytest.mark.parametrize("half", [True, False])
def test_dem_fp32_mixed_vs_oracle(tm, half):
    announce(f"DEM FP32-mixed vs oracle FP64, compressed 8^3 gas, half={half}")
    m, x, v, aabb = dem_fixture(tm)
    sim = tm.Simulation(m, x, v, aabb, 1.3, precision="fp32mixed", dem_radius=0.5)
    sim.copy_to_accelerator()
    orc = po.OracleSim(m, x, v, aabb, 1.3, threads=THREADS)
    sim.setup(half, 1.0, 0.3)
    orc.setup(half, po.HALF_GHOST_FULL, 1.3)
    sim.dem(half, 1.0, 0.5, 0.2, 100.0, 0.3)
    orc.dem(half, 1.0, 0.5, 0.5, 0.2, 100.0, 0.3)
    _, _, _, fg = sim.state()
    _, _, fo = orc.state()
    assert np.abs(fo).max() > 1.0  # contacts exist
    rms, mx = force_error_atom(fg, fo)
    announce(f"  per-atom relative error rms={rms:.3e} max={mx:.3e}")
    assert rms <= FP32_BAR and mx <= FP32_BAR, (rms, mx)
    sim.run_dem(30, 10, 0.001, half)
    orc.run_dem(30, 10, 0.001, half)
    _, xg, vg, _ = sim.state()
    xo, vo, _ = orc.state()
    assert np.max(np.abs(xg - xo)) < 1e-6 and np.max(np.abs(vg - vo)) < 1e-4
    sim.close()
    orc.close()


def dem_energy(x, v, L, k=100.0, diameter=1.0):
    """Kinetic energy plus the spring energy k delta^2 / 2 of every
    overlapping pair (delta = diameter - r), periodic minimum images."""
    from scipy.spatial import cKDTree
    xw = np.mod(x, L)
    xw[xw >= L] = 0.0
    tree = cKDTree(xw, boxsize=L)
    pairs = tree.query_pairs(diameter, output_type="ndarray")
    spring = 0.0
    if len(pairs):
        d = xw[pairs[:, 0]] - xw[pairs[:, 1]]
        d -= L * np.rint(d / L)
        r = np.sqrt(np.sum(d * d, axis=1))
        delta = np.clip(diameter - r, 0.0, None)
        spring = 0.5 * k * float(np.sum(delta * delta))
    return 0.5 * float(np.sum(v * v)), spring, len(pairs)


def test_dem_granular_gas_64cube_invariants(tm):
    """Full-size granular gas (528,384 spheres, half list, FP32-mixed)."""
    announce("DEM granular gas 64^3 (528,384 spheres), FP32-mixed, 1000 steps: invariants")
    m, x, v = tm.generate_granular_gas(64)
    assert len(m) == 528384
    L = 64 * 1.6796
    aabb = np.array([0.0, L, 0.0, L, 0.0, L])
    sim = tm.Simulation(m, x, v, aabb, 1.3, precision="fp32mixed", dem_radius=0.5)
    sim.copy_to_accelerator()
    sim.setup(True, 1.0, 0.3)
    sim.dem(True, 1.0, 0.5, 0.2, 100.0, 0.3)
    _, x0, v0, f0 = sim.state()
    # nearest neighbours are 1.1877 apart: no contact, the force is exactly zero
    assert np.count_nonzero(f0) == 0
    p0 = v0.sum(axis=0)
    ke0, sp0, nc0 = dem_energy(x0, v0, L)
    assert nc0 == 0 and sp0 == 0.0
    e_prev = ke0
    contacts = 0
    for block in range(5):
        sim.run_dem(200, REN, 0.001, True, 1.0, 0.3, 0.5, 0.2, 100.0, 0.3)
        _, xb, vb, _ = sim.state()
        p = vb.sum(axis=0)
        scale = np.sqrt(np.mean(np.sum(v0 * v0, axis=1))) * len(vb)
        assert np.abs(p - p0).max() <= 1e-10 * scale, (block, p - p0)
        ke, sp, nc = dem_energy(xb, vb, L)
        contacts = max(contacts, nc)
        announce(f"  after {200 * (block + 1)} steps: KE {ke:.6f} spring {sp:.6f} contacts {nc}")
        # dissipative contacts: kinetic + spring energy does not grow (the
        # integrator's own fluctuation is far below 1e-6 of the total)
        assert ke + sp <= e_prev * (1.0 + 1e-6), (block, ke + sp, e_prev)
        e_prev = ke + sp
    assert e_prev < ke0 and contacts > 0  # collisions happened and dissipated energy
    # FP32-mixed against FP64 forces on the current (colliding) state
    sim.set_precision("fp64")
    sim.dem(True, 1.0, 0.5, 0.2, 100.0, 0.3)
    f64 = sim.state()[3]
    sim.set_precision("fp32mixed")
    sim.dem(True, 1.0, 0.5, 0.2, 100.0, 0.3)
    f32 = sim.state()[3]
    on = np.any(f64 != 0.0, axis=1)
    assert np.array_equal(on, np.any(f32 != 0.0, axis=1))  # identical contact decisions
    rms, mx = force_error_atom(f32[on], f64[on])
    announce(f"  FP32-mixed vs FP64 on {int(on.sum())} atoms in contact: rms {rms:.3e} max {mx:.3e}")
    assert rms <= FP32_BAR and mx <= FP32_BAR, (rms, mx)
    sim.close()


def test_batched_64_replicas_fp64_vs_oracle(tm):
    announce("batched replicas: 64 boxes (seed 2024), FP64, 25 steps, every replica vs oracle")
    boxes = tm.replica_boxes(seed=2024, nrep=64)
    assert sum(len(b[1]) for b in boxes) == 1334984
    b = tm.Batch(boxes, precision="fp64")
    assert b.run(25) == 25
    for r, (aabb, x, v) in enumerate(boxes):
        orc = po.OracleSim(np.ones(len(x)), x, v, aabb, RC + SKIN, threads=THREADS)
        orc.setup(False, po.HALF_GHOST_FULL, RC + SKIN)
        orc.lj(False, RC)
        orc.run_lj(25, REN, DT)
        xo, vo, fo = orc.state()
        orc.close()
        xb, vb, fb = b.state(r)
        assert np.array_equal(xb, xo), f"replica {r} ({len(x)} atoms) positions"
        assert np.array_equal(vb, vo), f"replica {r} velocities"
        assert np.array_equal(fb, fo), f"replica {r} forces"
    b.close()
