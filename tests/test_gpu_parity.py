"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.  Integer/index work (ghost tables, cell lists,
neighbour lists) must be bit-exact; FP64 full-list forces and trajectories
are bit-exact too (same operation order, no contraction); half lists, the
FP32-mixed path and reductions have stated tolerances."""
import numpy as np
import pytest

import pyoracle as po
from common import (FP32_BAR, RC, SKIN, announce, boundary_pairs, brute_force_rsq, fma_rsq, force_error,
                    force_error_atom, make_pair, perturbed_fcc)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tm():
    import tinymd_amd
    name, cus = tinymd_amd.device_info()
    if cus == 0:
        pytest.fail("no HIP device visible to libtinymd_amd")
    announce(f"device {name} with {cus} CUs")
    return tinymd_amd


def setup_both(sim, orc, half=False, half_mode="ghost_full"):
    hm = po.HALF_GHOST_FULL if half_mode == "ghost_full" else po.HALF_REF_LITERAL
    sim.setup(half, RC, SKIN)
    orc.setup(half, hm, RC + SKIN)


def test_lattice_matches_oracle(tm):
    announce("lattice 8x6x5 + create_velocity")
    m, x, v = tm.generate_fcc(8, 6, 5)
    aabb, m2, x2, v2 = po.lattice(8, 6, 5)
    assert np.array_equal(x, x2) and np.array_equal(v, v2) and np.array_equal(m, m2)
    sim = tm.Simulation(m, x, v, aabb, RC + SKIN)
    sim.create_velocity(1.44)
    sim.copy_to_accelerator()
    orc = po.OracleSim(m2, x2, v2, aabb, RC + SKIN)
    orc.create_velocity(1.44)
    _, _, vg, _ = sim.state()
    _, vo, _ = orc.state()
    assert np.array_equal(vg, vo)


@pytest.mark.parametrize("shape,amp", [((8, 8, 8), 0.0), ((6, 7, 9), 0.15), ((5, 5, 5), 0.3)])
def test_ghosts_cells_exact(tm, shape, amp):
    announce(f"ghosts+cells {shape} amp={amp}")
    aabb, m, x, v = perturbed_fcc(*shape, amp=amp, seed=1)
    sim, orc = make_pair(aabb, m, x, v)
    setup_both(sim, orc)
    assert sim.nghost == orc.ng
    gt, go = sim.ghost_table(), orc.ghosts()
    for key in ("owner", "img", "src", "flag"):
        assert np.array_equal(gt[key], go[key]), key
    _, gx, _, _ = sim.ghost_state()
    assert np.array_equal(gx, go["pos"])
    st, parts, pcell = sim.cells()
    so, po_, pco = orc.cells()
    assert np.array_equal(st, so) and np.array_equal(parts, po_) and np.array_equal(pcell, pco)


@pytest.mark.parametrize("half,half_mode", [(False, "ghost_full"), (True, "ghost_full"), (True, "ref_literal")])
def test_neighbor_lists_exact(tm, half, half_mode):
    announce(f"neighbour lists half={half} mode={half_mode}")
    aabb, m, x, v = perturbed_fcc(7, 8, 6, amp=0.2, seed=2)
    sim, orc = make_pair(aabb, m, x, v, half_mode=half_mode)
    setup_both(sim, orc, half, half_mode)
    cg, lg = sim.neighbor_lists()
    co, lo = orc.neighbor_lists()
    assert np.array_equal(cg, co)
    n = len(cg)
    for i in range(n):
        assert np.array_equal(lg[i, :cg[i]], lo[i, :co[i]]), i
    if half and half_mode == "ref_literal":
        # the reference's literal filter keeps no ghost (index >= n)
        assert all(np.all(lo[i, :co[i]] < n) for i in range(n))


@pytest.mark.parametrize("half", [False, True])
def test_cutoff_boundary_pairs(tm, half):
    """Pairs whose squared distance is exactly r_nb^2, one ulp below and one
    ulp above, and the same around rc^2: the lists equal the oracle's entry
    for entry (the pair at r_nb^2 kept, the one an ulp beyond dropped), and
    the FP64 forces equal the oracle's (bit for bit with full lists), so the
    force of a pair at rc^2 is zero and one an ulp inside is not.  The box
    has no ghosts; with full lists it also runs the fused step loop."""
    announce(f"cutoff boundary pairs, half={half}")
    aabb, m, x, v, pairs = boundary_pairs()
    sim, orc = make_pair(aabb, m, x, v)
    setup_both(sim, orc, half)
    cg, lg = sim.neighbor_lists()
    co, lo = orc.neighbor_lists()
    assert np.array_equal(cg, co)
    for i in range(len(cg)):
        assert np.array_equal(lg[i, :cg[i]], lo[i, :co[i]]), i
    cut2 = (RC + SKIN) * (RC + SKIN)
    for a, p, name, goal in pairs:
        kept = p in lg[a, :cg[a]].tolist() or a in lg[p, :cg[p]].tolist()
        assert kept == (goal <= cut2), name
    sim.lennard_jones(half, RC)
    orc.lj(half, RC)
    _, _, _, fg = sim.state()
    _, _, fo = orc.state()
    if half:
        assert np.allclose(fg, fo, rtol=1e-12, atol=1e-15)
    else:
        assert np.array_equal(fg, fo)
    r2 = brute_force_rsq(x)
    for a, p, name, goal in pairs:
        if any(j not in (a, p) for j in np.nonzero(r2[a] < RC * RC)[0]):
            continue
        assert np.any(fg[a] != 0.0) == (goal < RC * RC), name
    # the FP32 variants decide on their FMA-formed rsq, which can lie an ulp
    # from the reference's rounding: at rc^2 their in/out decisions follow
    # that rule exactly, and where both rules agree the forces meet the bar
    flips = 0
    for precision in ("fp32mixed", "fp32kahan"):
        sim.set_precision(precision)
        sim.lennard_jones(half, RC)
        _, _, _, f32 = sim.state()
        agree = np.ones(len(x), dtype=bool)
        for a, p, name, goal in pairs:
            if any(j not in (a, p) for j in np.nonzero(r2[a] < RC * RC)[0]):
                continue
            inside = fma_rsq(x[a], x[p]) < RC * RC
            assert np.any(f32[a] != 0.0) == inside, (precision, name)
            if inside != (goal < RC * RC):
                agree[[a, p]] = False
                flips += 1
        rms, mx = force_error_atom(f32[agree], fg[agree])
        assert rms < FP32_BAR and mx < FP32_BAR, (precision, rms, mx)
    assert flips > 0  # the construction does hit pairs the two roundings split
    sim.set_precision("fp64")
    sim.lennard_jones(half, RC)  # the FP64 forces the run starts from
    if not half:
        # the fused step loop on a box without ghosts (no atom within r_nb of
        # a face): FP64 trajectory bit for bit, rebuilds included
        assert sim.nghost == 0 and orc.ng == 0
        sim.run_lj(45, reneigh=20, dt=0.005)
        orc.run_lj(45, reneigh=20, dt=0.005)
        _, xg, vg, fg = sim.state()
        xo, vo, fo = orc.state()
        assert np.array_equal(xg, xo) and np.array_equal(vg, vo) and np.array_equal(fg, fo)
    sim.close()
    orc.close()


def test_atoms_exactly_on_box_faces(tm):
    """Atoms moved onto exactly 0 or exactly L in one coordinate (those
    within 0.15 of a face, so the physics is unchanged): the wrap is strict
    (core/compute.impala:118-122), so an atom at L stays at L and is binned
    and given ghosts from there.  Lists, FP64 forces and a 45-step FP64
    trajectory with two rebuilds equal the oracle's bit for bit."""
    announce("atoms exactly on the box faces")
    aabb, m, x, v = perturbed_fcc(5, amp=0.1, seed=9)
    L = aabb[1::2] - aabb[0::2]
    x = x.copy()
    moved = 0
    for i in range(len(x)):
        d = i % 3
        if x[i, d] < 0.15:
            x[i, d] = 0.0 if (i // 3) % 2 else L[d]
            moved += 1
        elif x[i, d] > L[d] - 0.15:
            x[i, d] = L[d] if (i // 3) % 2 else 0.0
            moved += 1
    assert moved > 20 and np.any(x == L[0])
    sim, orc = make_pair(aabb, m, x, v)
    setup_both(sim, orc)
    cg, lg = sim.neighbor_lists()
    co, lo = orc.neighbor_lists()
    assert np.array_equal(cg, co)
    for i in range(len(cg)):
        assert np.array_equal(lg[i, :cg[i]], lo[i, :co[i]]), i
    sim.lennard_jones(False, RC)
    orc.lj(False, RC)
    _, _, _, fg = sim.state()
    _, _, fo = orc.state()
    assert np.array_equal(fg, fo)
    sim.run_lj(45, reneigh=20, dt=0.005)
    orc.run_lj(45, reneigh=20, dt=0.005)
    _, xg, vg, fg = sim.state()
    xo, vo, fo = orc.state()
    assert np.array_equal(xg, xo) and np.array_equal(vg, vo) and np.array_equal(fg, fo)
    sim.close()
    orc.close()


@pytest.mark.parametrize("half", [False, True])
def test_dense_lists_exact(tm, half):
    """A compressed lattice (lattice constant 1.0, about 365 neighbours per
    atom in a 6^3 box whose grid's inner cells also hold ghosts): 32-candidate
    chunks accepted whole or nearly so, lists grown from capacity 100 past
    700, small tiles with windows read from the mirror.  Lists equal the
    oracle's entry for entry; FP64 forces bit for bit with full lists."""
    announce(f"dense lattice lists, half={half}")
    aabb, m, x, v = perturbed_fcc(6, amp=0.05, seed=7, potmin=1.0)
    sim, orc = make_pair(aabb, m, x, v)
    setup_both(sim, orc, half)
    cg, lg = sim.neighbor_lists()
    co, lo = orc.neighbor_lists()
    assert cg.max() > 300  # dense: chunks of 32 candidates mostly accepted
    assert np.array_equal(cg, co)
    for i in range(len(cg)):
        assert np.array_equal(lg[i, :cg[i]], lo[i, :co[i]]), i
    sim.lennard_jones(half, RC)
    orc.lj(half, RC)
    _, _, _, fg = sim.state()
    _, _, fo = orc.state()
    if half:
        rms, mx = force_error(fg, fo)
        assert rms < 1e-12 and mx < 1e-11, (rms, mx)
    else:
        assert np.array_equal(fg, fo)
    sim.close()
    orc.close()


@pytest.mark.parametrize("half", [False, True])
def test_list_capacity_retry(tm, half):
    """Capacity 1 grows and retries; with half lists the complementary
    columns share the capacity, and the forces come out right."""
    announce(f"list capacity 1 -> grow and retry, half={half}")
    aabb, m, x, v = perturbed_fcc(6, amp=0.1, seed=3)
    sim, orc = make_pair(aabb, m, x, v, nb_cap=1)
    setup_both(sim, orc, half)
    cg, lg = sim.neighbor_lists()
    co, lo = orc.neighbor_lists()
    assert np.array_equal(cg, co)
    for i in range(len(cg)):
        assert np.array_equal(lg[i, :cg[i]], lo[i, :co[i]])
    sim.lennard_jones(half, RC)
    orc.lj(half, RC)
    _, _, _, fg = sim.state()
    _, _, fo = orc.state()
    rms, mx = force_error(fg, fo)
    assert rms < 1e-12 and mx < 1e-11, (rms, mx)


def test_lj_full_fp64_bitwise(tm):
    announce("LJ full FP64 forces, bit-exact")
    aabb, m, x, v = perturbed_fcc(8, amp=0.1, seed=4)
    sim, orc = make_pair(aabb, m, x, v)
    setup_both(sim, orc)
    sim.lennard_jones(False, RC)
    orc.lj(False, RC)
    _, _, _, fg = sim.state()
    _, _, fo = orc.state()
    assert np.array_equal(fg, fo)


@pytest.mark.parametrize("half_mode", ["ghost_full", "ref_literal"])
def test_lj_half_fp64(tm, half_mode):
    announce(f"LJ half FP64 forces mode={half_mode}")
    aabb, m, x, v = perturbed_fcc(8, amp=0.1, seed=5)
    sim, orc = make_pair(aabb, m, x, v, half_mode=half_mode)
    hm = po.HALF_GHOST_FULL if half_mode == "ghost_full" else po.HALF_REF_LITERAL
    setup_both(sim, orc, True, half_mode)
    sim.lennard_jones(True, RC)
    orc.lj(True, RC)
    _, _, _, fg = sim.state()
    _, _, fo = orc.state()
    rms, mx = force_error(fg, fo)
    assert rms < 1e-12 and mx < 1e-11, (rms, mx)
    if half_mode == "ghost_full":
        # each periodic pair once: equals the full-list force
        orc2 = po.OracleSim(m, x, v, aabb, RC + SKIN)
        orc2.setup(False, hm, RC + SKIN)
        orc2.lj(False, RC)
        _, _, ff = orc2.state()
        rms, mx = force_error(fg, ff)
        assert rms < 1e-12 and mx < 1e-11, (rms, mx)


@pytest.mark.parametrize("half,precision", [(False, "fp32mixed"), (True, "fp32mixed"), (False, "fp32kahan")])
def test_lj_fp32_mixed_error(tm, half, precision):
    announce(f"LJ {precision} error half={half}")
    aabb, m, x, v = perturbed_fcc(10, amp=0.15, seed=6)
    sim, orc = make_pair(aabb, m, x, v, precision=precision)
    setup_both(sim, orc, half)
    sim.lennard_jones(half, RC)
    orc.lj(half, RC)
    _, _, _, fg = sim.state()
    _, _, fo = orc.state()
    rms, mx = force_error_atom(fg, fo)
    announce(f"  {precision} per-atom relative error rms={rms:.3e} max={mx:.3e}")
    assert rms <= FP32_BAR and mx <= FP32_BAR, (rms, mx)


@pytest.mark.parametrize("half", [False, True])
def test_thermo_matches_oracle(tm, half):
    announce(f"thermo half={half}")
    aabb, m, x, v = perturbed_fcc(8, amp=0.1, seed=7)
    sim, orc = make_pair(aabb, m, x, v)
    setup_both(sim, orc, half)
    sim.lennard_jones(half, RC)
    orc.lj(half, RC)
    tg, to = sim.thermo(RC), orc.thermo()
    for key in ("T", "KE", "PE", "virial", "P"):
        assert abs(tg[key] - to[key]) <= 1e-12 * max(1.0, abs(to[key])), (key, tg[key], to[key])
    assert tg["pairs"] == to["pairs"]


def test_thermo_energy_shift(tm):
    """PE of the energy-shifted potential U(r) - U(rc): with a full list every
    in-cutoff pair is listed twice at weight 1/2, so it is the oracle's PE
    minus U(rc) times half the in-cutoff entries; the half list (each pair
    once, ghost pairs at 1/2 from both owners) gives the same physical sum."""
    announce("thermo with the energy-shifted potential, full and half lists")
    aabb, m, x, v = perturbed_fcc(8, amp=0.1, seed=7)
    sim, orc = make_pair(aabb, m, x, v)
    setup_both(sim, orc, False)
    sim.lennard_jones(False, RC)
    orc.lj(False, RC)
    to = orc.thermo()
    ts = sim.thermo(RC, eshift=True)
    urc = 4.0 * (RC ** -12 - RC ** -6)
    expect = to["PE"] - urc * to["pairs"] / 2
    assert abs(ts["PE"] - expect) <= 1e-12 * abs(expect), (ts["PE"], expect)
    assert ts["pairs"] == to["pairs"] and ts["virial"] == pytest.approx(to["virial"], rel=1e-12)
    sim.setup(True, RC, SKIN)
    th = sim.thermo(RC, eshift=True)
    assert abs(th["PE"] - expect) <= 1e-12 * abs(expect), (th["PE"], expect)
    sim.close()
    orc.close()


def test_trajectory_bitwise_abi_and_runner(tm):
    announce("trajectory 6^3, 60 steps, ABI loop and in-library loop vs oracle")
    aabb, m, x, v = perturbed_fcc(6, amp=0.0, seed=0)
    sim, orc = make_pair(aabb, m, x, v)
    setup_both(sim, orc)
    sim.lennard_jones(False, RC)
    orc.lj(False, RC)
    sim.run_abi_lj(30, reneigh=20, dt=0.005)
    orc.run_lj(30, reneigh=20, dt=0.005)
    _, xg, vg, fg = sim.state()
    xo, vo, fo = orc.state()
    assert np.array_equal(xg, xo) and np.array_equal(vg, vo) and np.array_equal(fg, fo)
    sim.run_lj(30, reneigh=20, dt=0.005)
    orc.run_lj(30, reneigh=20, dt=0.005)
    _, xg, vg, fg = sim.state()
    xo, vo, fo = orc.state()
    assert np.array_equal(xg, xo) and np.array_equal(vg, vo) and np.array_equal(fg, fo)


def test_ref_lag_ghosts_trajectory(tm):
    announce("REF_LAG ghost refresh trajectory vs oracle")
    aabb, m, x, v = perturbed_fcc(6, amp=0.05, seed=8)
    sim, orc = make_pair(aabb, m, x, v, ghost_mode="ref_lag")
    setup_both(sim, orc)
    sim.lennard_jones(False, RC)
    orc.lj(False, RC)
    sim.run_lj(25, reneigh=20, dt=0.005)
    orc.run_lj(25, reneigh=20, dt=0.005)
    _, xg, vg, _ = sim.state()
    xo, vo, _ = orc.state()
    assert np.array_equal(xg, xo) and np.array_equal(vg, vo)


@pytest.mark.parametrize("shape", [(6, 6, 6), (4, 4, 4), (4, 7, 5)])
def test_fused_loop_ghost_positions(tm, shape):
    """Fused step loop (positions advanced in the mirror, ghosts refreshed
    from their owners there): in the small boxes corner owners have the full
    7 images.  The trajectory and the final ghost positions are the oracle's
    bit for bit."""
    announce(f"fused loop ghost positions {shape}, 25 steps")
    aabb, m, x, v = perturbed_fcc(*shape, amp=0.05, seed=31)
    sim, orc = make_pair(aabb, m, x, v)
    setup_both(sim, orc)
    sim.lennard_jones(False, RC)
    orc.lj(False, RC)
    sim.run_lj(25, reneigh=20, dt=0.005)
    orc.run_lj(25, reneigh=20, dt=0.005)
    _, xg, vg, fg = sim.state()
    xo, vo, fo = orc.state()
    assert np.array_equal(xg, xo) and np.array_equal(vg, vo) and np.array_equal(fg, fo)
    assert sim.nghost == orc.ng
    _, gx, _, _ = sim.ghost_state()
    assert np.array_equal(gx, orc.ghosts()["pos"])
    sim.close()


@pytest.mark.parametrize("reneigh,steps", [(1, 6), (3, 13)])
def test_trajectory_bitwise_short_reneighbour(tm, reneigh, steps):
    """Rebuilds every step and every third step of the fused loop.  Each
    rebuild takes the positions and velocities back from the loop's mirrors
    in its ghost selection pass, and its list build runs on the previous
    window statistics.  The FP64 trajectory, the lists and the final ghost
    positions are the oracle's bit for bit."""
    announce(f"trajectory 16^3, {steps} steps, rebuild every {reneigh}")
    aabb, m, x, v = perturbed_fcc(16, amp=0.1, seed=41)
    sim, orc = make_pair(aabb, m, x, v, threads=8)
    setup_both(sim, orc)
    sim.lennard_jones(False, RC)
    orc.lj(False, RC)
    sim.run_lj(steps, reneigh=reneigh, dt=0.005)
    orc.run_lj(steps, reneigh=reneigh, dt=0.005)
    _, xg, vg, fg = sim.state()
    xo, vo, fo = orc.state()
    assert np.array_equal(xg, xo) and np.array_equal(vg, vo) and np.array_equal(fg, fo)
    assert sim.nghost == orc.ng
    _, gx, _, _ = sim.ghost_state()
    assert np.array_equal(gx, orc.ghosts()["pos"])
    cg, lg = sim.neighbor_lists()
    co, lo = orc.neighbor_lists()
    assert np.array_equal(cg, co)
    assert all(np.array_equal(lg[i, :cg[i]], lo[i, :co[i]]) for i in range(0, len(cg), 53))
    sim.close()


def test_half_list_trajectory_close(tm):
    announce("half-list trajectory vs oracle (tolerance)")
    aabb, m, x, v = perturbed_fcc(6, amp=0.05, seed=9)
    sim, orc = make_pair(aabb, m, x, v)
    setup_both(sim, orc, True)
    sim.lennard_jones(True, RC)
    orc.lj(True, RC)
    sim.run_lj(40, reneigh=20, dt=0.005, half=True)
    orc.run_lj(40, reneigh=20, dt=0.005, half=True)
    _, xg, vg, _ = sim.state()
    xo, vo, _ = orc.state()
    assert np.max(np.abs(xg - xo)) < 1e-9 and np.max(np.abs(vg - vo)) < 1e-8


def test_determinism_half_and_full(tm):
    announce("bitwise run-to-run determinism")
    aabb, m, x, v = perturbed_fcc(7, amp=0.1, seed=10)
    outs = []
    for _ in range(2):
        sim, _ = make_pair(aabb, m, x, v)
        sim.setup(True, RC, SKIN)
        sim.lennard_jones(True, RC)
        sim.run_lj(20, reneigh=10, half=True)
        outs.append(sim.state()[1:])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("half", [True, False])
def test_dem_matches_oracle(tm, half):
    announce(f"DEM granular gas 8^3 half={half}")
    m, x, v = tm.generate_granular_gas(8, seed=7, vmax=1.0)
    aabb = np.array([0.0, 8 * 1.6796, 0.0, 8 * 1.6796, 0.0, 8 * 1.6796])
    # compress the lattice so spheres of diameter 1.0 overlap
    x = x * 0.55
    aabb = aabb * 0.55
    sim, orc = make_pair(aabb, m, x, v, cutoff=1.3)
    sim.setup(half, 1.0, 0.3)
    orc.setup(half, po.HALF_GHOST_FULL, 1.3)
    sim.dem(half, 1.0, 0.5, 0.2, 100.0, 0.3)
    orc.dem(half, 1.0, 0.5, 0.5, 0.2, 100.0, 0.3)
    _, _, _, fg = sim.state()
    _, _, fo = orc.state()
    assert np.abs(fo).max() > 1.0  # contacts exist
    rms, mx = force_error(fg, fo)
    assert rms < 1e-12 and mx < 1e-11, (rms, mx)
    sim.run_dem(30, reneigh=10, dt=0.001, half=half)
    orc.run_dem(30, reneigh=10, dt=0.001, half=half)
    _, xg, vg, _ = sim.state()
    xo, vo, _ = orc.state()
    assert np.max(np.abs(xg - xo)) < 1e-10 and np.max(np.abs(vg - vo)) < 1e-9


def test_dem_contact_boundary_pairs(tm):
    """Sphere pairs (diameter 1) at exactly touching distance, one ulp
    inside and one ulp apart, the partner moving: the dashpot term makes the
    force of every pair in contact non-zero even at zero overlap, so the
    contact decision (2R)^2 - rsq > 0 in FP64 is visible.  FP64 forces equal
    the oracle's, zero exactly for the touching and the separated pairs."""
    announce("DEM contact boundary pairs")
    one = 1.0
    targets = [("d2_eq", one), ("d2_lo", np.nextafter(one, 0.0)), ("d2_hi", np.nextafter(one, 2.0))]
    aabb, m, x, v, pairs = boundary_pairs(seed=8, targets=targets)
    v = v.copy()
    v[27:] = [0.3, -0.2, 0.1]
    # cutoff 1.1 > 2R: the contact test alone decides at rsq = (2R)^2
    sim, orc = make_pair(aabb, m, x, v, cutoff=1.4, dem_radius=0.5)
    sim.setup(True, 1.1, 0.3)
    orc.setup(True, po.HALF_GHOST_FULL, 1.4)
    sim.dem(True, 1.1, 0.5, 0.2, 100.0, 0.3)
    orc.dem(True, 1.1, 0.5, 0.5, 0.2, 100.0, 0.3)
    _, _, _, fg = sim.state()
    _, _, fo = orc.state()
    assert np.allclose(fg, fo, rtol=1e-12, atol=0.0)
    for a, p, name, goal in pairs:
        touching = goal < one  # (2R)^2 - rsq > 0
        assert np.any(fg[a] != 0.0) == touching, name
        assert np.any(fo[a] != 0.0) == touching, name
    sim.close()
    orc.close()


def test_full_size_lattice_invariants(tm):
    """80^3 (2,048,000 atoms): size-independent invariants at t = 0."""
    announce("80^3 lattice invariants")
    sim = tm.Simulation.fcc(80)
    sim.setup(False, RC, SKIN)
    assert sim.nghost == 280731
    counts, _ = sim.neighbor_lists(with_lists=False)
    assert np.all(counts == 78)
    sim.lennard_jones(False, RC)
    th = sim.thermo(RC)
    _, _, _, f = sim.state()
    assert np.abs(f).max() < 1e-10  # perfect lattice: F = 0 by symmetry
    assert abs(th["PE"] / sim.nlocal - (-6.7733178)) < 5e-7
    assert abs(th["T"] - 1.44) < 1e-12
    sim.close()


def test_trajectory_bitwise_16cube_staged_windows(tm):
    """16^3 (16,384 atoms): large enough that block windows are staged in
    LDS; FP64 trajectory must stay bit-identical to the oracle."""
    announce("trajectory 16^3, 45 steps, FP64 full, bit-exact")
    aabb, m, x, v = perturbed_fcc(16, amp=0.05, seed=21)
    sim, orc = make_pair(aabb, m, x, v, threads=8)
    setup_both(sim, orc)
    sim.lennard_jones(False, RC)
    orc.lj(False, RC)
    sim.run_lj(45, reneigh=20, dt=0.005)
    orc.run_lj(45, reneigh=20, dt=0.005)
    _, xg, vg, fg = sim.state()
    xo, vo, fo = orc.state()
    assert np.array_equal(xg, xo) and np.array_equal(vg, vo) and np.array_equal(fg, fo)
    cg, lg = sim.neighbor_lists()
    co, lo = orc.neighbor_lists()
    assert np.array_equal(cg, co)
    assert all(np.array_equal(lg[i, :cg[i]], lo[i, :co[i]]) for i in range(0, len(cg), 97))


@pytest.mark.parametrize("tile,threads", [((3, 3, 3), 64), ((2, 2, 2), 1024), ((4, 4, 2), 0), ((8, 8, 8), 256)])
def test_trajectory_bitwise_tile_shapes(tm, tile, threads):
    """Blocks smaller than a tile's local count (several passes per block),
    larger than it, the automatic size, and windows beyond the LDS budget
    (read from the mirror): the FP64 trajectory across a rebuild is the
    oracle's bit for bit."""
    announce(f"trajectory 16^3, 25 steps, tile {tile}, threads {threads}")
    aabb, m, x, v = perturbed_fcc(16, amp=0.05, seed=22)
    sim, orc = make_pair(aabb, m, x, v, threads=8)
    sim.set_tile(*tile, threads=threads)
    setup_both(sim, orc)
    sim.lennard_jones(False, RC)
    orc.lj(False, RC)
    sim.run_lj(25, reneigh=20, dt=0.005)
    orc.run_lj(25, reneigh=20, dt=0.005)
    _, xg, vg, fg = sim.state()
    xo, vo, fo = orc.state()
    assert np.array_equal(xg, xo) and np.array_equal(vg, vo) and np.array_equal(fg, fo)
    sim.close()


@pytest.mark.parametrize("precision,tol_x", [("fp64", 1e-9), ("fp32mixed", 1e-6)])
def test_half_list_trajectory_16cube(tm, precision, tol_x):
    """Half lists at 16^3 (staged windows; the complementary columns carry
    the reactions): a trajectory across a rebuild stays within rounding of
    the oracle's half-list run."""
    announce(f"half-list trajectory 16^3, 25 steps, {precision}")
    aabb, m, x, v = perturbed_fcc(16, amp=0.05, seed=23)
    sim, orc = make_pair(aabb, m, x, v, threads=8, precision=precision)
    setup_both(sim, orc, True)
    sim.lennard_jones(True, RC)
    orc.lj(True, RC)
    sim.run_lj(25, reneigh=20, dt=0.005, half=True)
    orc.run_lj(25, reneigh=20, dt=0.005, half=True)
    _, xg, vg, fg = sim.state()
    xo, vo, fo = orc.state()
    assert np.max(np.abs(xg - xo)) < tol_x and np.max(np.abs(vg - vo)) < 100 * tol_x
    rms, mx = force_error(fg, fo)
    assert rms < (1e-12 if precision == "fp64" else 1e-5), (rms, mx)
    sim.close()


def test_report_lines_match_oracle(tm, capfd):
    """md_report_particles / md_report_iterations print the reference's
    headers and values (core/runtime.impala:280-316): local and ghost counts
    and the temperature of the host mirror; the list entries visited by the
    last force evaluation, those inside the cutoff, and the atoms of the last
    integration.  Values equal the oracle's after the same run."""
    announce("report lines after a 25-step run vs oracle")
    aabb, m, x, v = perturbed_fcc(6, amp=0.05, seed=51)
    sim, orc = make_pair(aabb, m, x, v)
    setup_both(sim, orc)
    sim.lennard_jones(False, RC)
    orc.lj(False, RC)
    sim.run_lj(25, reneigh=20, dt=0.005)
    orc.run_lj(25, reneigh=20, dt=0.005)
    sim.copy_from_accelerator()
    L = tm.load()
    capfd.readouterr()
    L.md_report_particles()
    L.md_report_iterations()
    out = capfd.readouterr().out
    to = orc.thermo()
    counts, _ = orc.neighbor_lists()
    n = len(x)
    expect = ("# Local # Ghost Temp\n%d %d %g \n" % (n, orc.ng, to["T"])
              + "ForceIters RsqIters IntegrationIters\n%d %d %d \n" % (int(counts.sum()), int(to["pairs"]), n))
    assert out == expect, (out, expect)
    sim.close()


def test_aabb_export(tm):
    """md_write_grid_aabb_data_to_arrays: the 8 corners of the grid's AABB
    (the rank AABB extended by one cell spacing) at index k*4 + j*2 + i, zero
    mass, velocity and force; returns the number of local atoms
    (core/grid.impala:520-546)."""
    import ctypes
    from tinymd_amd._lib import Vector3D
    announce("AABB corner export")
    aabb, m, x, v = perturbed_fcc(5, 6, 7, amp=0.05, seed=52)
    sp = RC + SKIN
    sim, _ = make_pair(aabb, m, x, v, cutoff=sp)
    mass = np.full(8, -1.0)
    pos = np.zeros((8, 3)); vel = np.full((8, 3), -1.0); frc = np.full((8, 3), -1.0)
    vp = ctypes.POINTER(Vector3D)
    n = tm.load().md_write_grid_aabb_data_to_arrays(mass.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                     pos.ctypes.data_as(vp), vel.ctypes.data_as(vp),
                                                     frc.ctypes.data_as(vp))
    assert n == len(x)
    lo, hi = aabb[0::2] - sp, aabb[1::2] + sp
    for i in range(2):
        for j in range(2):
            for k in range(2):
                idx = k * 4 + j * 2 + i
                want = [lo[0] if i == 0 else hi[0], lo[1] if j == 0 else hi[1], lo[2] if k == 0 else hi[2]]
                assert np.array_equal(pos[idx], want), (idx, pos[idx], want)
    assert np.all(mass == 0.0) and np.all(vel == 0.0) and np.all(frc == 0.0)
    sim.close()


def parse_vtk(path):
    """Points and the seven scalar fields of a legacy VTK file in the layout
    of the reference's writer (vtk.h:6-91)."""
    lines = open(path).read().split("\n")
    assert lines[:4] == ["# vtk DataFile Version 2.0", "Particle data", "ASCII", "DATASET UNSTRUCTURED_GRID"]
    n = int(lines[4].split()[1])
    assert lines[4] == f"POINTS {n} double"
    pts = np.array([[float(t) for t in ln.split()] for ln in lines[5:5 + n]]).reshape(n, 3)
    k = 5 + n
    assert lines[k:k + 3] == ["", "", f"CELLS {n} {2 * n}"]
    assert lines[k + 3:k + 3 + n] == [f"1 {i}" for i in range(n)]
    fields = {}
    i = lines.index(f"POINT_DATA {n}")
    while i + 1 < len(lines) and lines[i + 1].startswith("SCALARS"):
        name = lines[i + 1].split()[1]
        assert lines[i + 2] == "LOOKUP_TABLE default"
        fields[name] = np.array([float(t) for t in lines[i + 3:i + 3 + n]])
        i = i + 3 + n
    return pts, fields


def test_vtk_files(tm, tmp_path):
    """tmd_write_vtk for locals, ghosts and the grid AABB: the reference's
    layout, and every value equal to the exported arrays at six decimals
    (the writer's std::fixed format)."""
    announce("VTK files of locals, ghosts and AABB")
    aabb, m, x, v = perturbed_fcc(5, amp=0.1, seed=21)
    sim, orc = make_pair(aabb, m, x, v)
    setup_both(sim, orc)
    sim.lennard_jones(False, RC)
    sim.run_lj(7, 5, 0.005, False, RC, SKIN)
    sim.copy_from_accelerator()
    mg, xg, vg, fg = sim.state()
    cases = {"locals": (mg, xg, vg, fg)}
    gm, gx, gv, gf = sim.ghost_state()
    cases["ghosts"] = (gm, gx, gv, gf)
    for what, (mm, xx, vv, ff) in cases.items():
        p = str(tmp_path / f"{what}.vtk")
        assert sim.write_vtk(p, what) == len(mm)
        pts, fields = parse_vtk(p)
        r6 = lambda a: np.array([float(f"{t:.6f}") for t in np.ravel(a)]).reshape(np.shape(a))
        assert np.array_equal(pts, r6(xx)), what
        assert np.array_equal(fields["mass"], r6(mm))
        for c, ax in enumerate("xyz"):
            assert np.array_equal(fields["velocity_" + ax], r6(vv[:, c])), (what, ax)
            assert np.array_equal(fields["force_" + ax], r6(ff[:, c])), (what, ax)
    p = str(tmp_path / "aabb.vtk")
    assert sim.write_vtk(p, "aabb") == 8
    pts, fields = parse_vtk(p)
    lo, hi = aabb[0::2] - (RC + SKIN), aabb[1::2] + (RC + SKIN)
    for idx in range(8):
        i, j, k = idx & 1, (idx >> 1) & 1, idx >> 2
        corner = [lo[0] if i == 0 else hi[0], lo[1] if j == 0 else hi[1], lo[2] if k == 0 else hi[2]]
        assert np.allclose(pts[idx], corner, atol=5e-7), (idx, pts[idx], corner)
    sim.close()
    orc.close()


def test_restart_round_trip(tm, tmp_path):
    """tmd_save_state / tmd_load_state: the restored arrays equal the saved
    ones bit for bit, and a run continued from the file (lists rebuilt at
    the start, as after md_initialize_grid) is bit-identical to the same
    continuation of the original simulation."""
    announce("restart file: save, restore, bit-identical continuation")
    aabb, m, x, v = perturbed_fcc(6, amp=0.1, seed=22)
    sim = tm.Simulation(m, x, v, aabb, RC + SKIN)
    sim.copy_to_accelerator()
    sim.setup(False, RC, SKIN)
    sim.lennard_jones(False, RC)
    sim.run_lj(30, 20, 0.005, False, RC, SKIN)
    p = str(tmp_path / "state.tmdrst")
    assert sim.save(p) == len(m)
    _, x0, v0, f0 = sim.state()
    # continue as the reference's driver starts a run: wrap, ghosts, cells,
    # lists, forces (main.cpp:691-700)
    sim.enforce_pbc()
    sim.setup(False, RC, SKIN)
    sim.lennard_jones(False, RC)
    sim.run_lj(30, 20, 0.005, False, RC, SKIN)
    _, xa, va, fa = sim.state()
    sim.close()
    # a corrupted copy is refused
    raw = bytearray(open(p, "rb").read())
    raw[200] ^= 1
    bad = str(tmp_path / "bad.tmdrst")
    open(bad, "wb").write(bytes(raw))
    with pytest.raises(tm.TinyMDError, match="corrupted"):
        tm.Simulation.restore(bad)
    tm.load().tmd_clear_error()
    sim2 = tm.Simulation.restore(p)
    assert sim2.nlocal == len(m)
    _, x1, v1, f1 = sim2.state()
    assert np.array_equal(x1, x0) and np.array_equal(v1, v0) and np.array_equal(f1, f0)
    sim2.enforce_pbc()
    sim2.setup(False, RC, SKIN)
    sim2.lennard_jones(False, RC)
    sim2.run_lj(30, 20, 0.005, False, RC, SKIN)
    _, xb, vb, fb = sim2.state()
    assert np.array_equal(xa, xb) and np.array_equal(va, vb) and np.array_equal(fa, fb)
    sim2.close()


@pytest.mark.parametrize("loop", ["abi", "runner"])
def test_sort_particles_trajectory_bitwise(tm, loop):
    """Cell-order renumbering of the locals at each rebuild (sort mode), on
    the GPU and in the oracle: FP64 trajectories bit-exact, through the md_*
    driver loop (the sort happens in md_enforce_pbc) and the library's own
    loop; right after a rebuild the exported locals are in cell order."""
    announce(f"sort mode 7^3, 40 + 5 steps, rebuild every 20, loop={loop}")
    aabb, m, x, v = perturbed_fcc(7, amp=0.1, seed=33)
    sim, orc = make_pair(aabb, m, x, v, sort=True)
    orc.set_sort(True)
    setup_both(sim, orc)
    sim.lennard_jones(False, RC)
    orc.lj(False, RC)
    run = sim.run_abi_lj if loop == "abi" else sim.run_lj
    run(40, reneigh=20, dt=0.005)
    orc.run_lj(40, 20, 0.005)
    _, xg, vg, fg = sim.state()
    xo, vo, fo = orc.state()
    assert np.array_equal(xg, xo) and np.array_equal(vg, vo) and np.array_equal(fg, fo)
    # the last step rebuilt (and sorted) after its drift: cell order now
    cut = RC + SKIN
    lo = aabb[0::2] - cut
    inf = orc.info()
    ijk = np.floor((xg - lo) / cut).astype(np.int64)
    key = (ijk[:, 2] * inf["ny"] + ijk[:, 1]) * inf["nx"] + ijk[:, 0]
    assert np.all(np.diff(key) >= 0)
    run(5, reneigh=20, dt=0.005)
    orc.run_lj(5, 20, 0.005)
    _, xg, vg, fg = sim.state()
    xo, vo, fo = orc.state()
    assert np.array_equal(xg, xo) and np.array_equal(vg, vo) and np.array_equal(fg, fo)
    sim.close()
    orc.close()


@pytest.mark.parametrize("half,precision", [(True, "fp64"), (True, "fp32mixed"), (False, "fp64")])
def test_dem_step_loop_matches_driver_loop(tm, half, precision):
    """The library's DEM step loop (state in the cell-ordered mirrors, kicks
    and drift inside the force pass) against the same steps through the
    md_* driver calls (main.cpp:715-762): bit-identical positions,
    velocities and forces, rebuilds included."""
    announce(f"DEM step loop vs md_* driver loop, half={half}, {precision}")
    m, x, v = tm.generate_granular_gas(8, seed=11, vmax=1.0)
    aabb = np.array([0.0, 8 * 1.6796, 0.0, 8 * 1.6796, 0.0, 8 * 1.6796]) * 0.6
    x = x * 0.6  # compressed: spheres of diameter 1.0 in contact

    def fresh():
        s = tm.Simulation(m, x, v, aabb, 1.3, precision=precision, dem_radius=0.5)
        s.copy_to_accelerator()
        s.setup(half, 1.0, 0.3)
        s.dem(half, 1.0, 0.5, 0.2, 100.0, 0.3)
        return s
    steps, ren, dt = 23, 10, 0.001
    a = fresh()
    a.run_dem(steps, ren, dt, half, 1.0, 0.3, 0.5, 0.2, 100.0, 0.3)
    _, xa, va, fa = a.state()
    a.close()
    b = fresh()
    for j in range(steps):
        b.initial_integration(dt)
        if (j + 1) % ren == 0:
            b.exchange()
            b.enforce_pbc()
            b.borders()
            b.distribute()
            b.assemble_neighborlists(half, 1.3)
        else:
            b.synchronize_ghost_layer()
        b.dem(half, 1.0, 0.5, 0.2, 100.0, 0.3)
        b.final_integration(dt)
    _, xb, vb, fb = b.state()
    b.close()
    assert np.abs(fa).max() > 1.0  # contacts
    assert np.array_equal(xa, xb) and np.array_equal(va, vb) and np.array_equal(fa, fb)


def test_dem_tile_edge_retune_keeps_trajectory(tm):
    """A granular gas that fills half of its box: the mean density gives
    8-cell automatic tiles whose densest windows overflow the DEM step
    pass's LDS share (48 B per slot, three blocks per CU); the step loop
    shrinks the edge at its first rebuild.  The trajectory is bit-identical
    to runs with explicit 6- and 8-cell tiles (the lists follow the cells,
    not the tiles)."""
    announce("DEM automatic tile edge retune, 16^3 gas, 45 steps")
    m, x, v = tm.generate_granular_gas(16)
    L = 16 * 1.6796
    aabb = np.array([0.0, L, 0.0, L, 0.0, L])

    def run(tile):
        s = tm.Simulation(m, x, v, aabb, 1.3, precision="fp32mixed", dem_radius=0.5)
        if tile:
            s.set_tile(tile)
        s.copy_to_accelerator()
        s.setup(True, 1.0, 0.3)
        s.dem(True, 1.0, 0.5, 0.2, 100.0, 0.3)
        before = s.tile_info()
        s.run_dem(45, 20, 0.001, True, 1.0, 0.3, 0.5, 0.2, 100.0, 0.3)
        after = s.tile_info()
        _, xs, vs, fs = s.state()
        s.close()
        return before, after, (xs, vs, fs)

    b0, a0, r0 = run(0)
    announce(f"  automatic tiles: {b0['tx']} -> {a0['tx']} cells, largest window {b0['max_window']} -> {a0['max_window']}")
    assert b0["tx"] == 8 and (b0["max_window"] + 1) * 48 > 160 * 1024 // 3 - 1024
    assert a0["tx"] < 8 and (a0["max_window"] + 1) * a0["stage_bytes"] <= 160 * 1024 // 3 - 1024
    for t in (6, 8):
        _, a, r = run(t)
        assert a["tx"] == t
        for p, q in zip(r0, r):
            assert np.array_equal(p, q), f"tile edge {t}"


def test_non_finite_position_fails_loudly(tm):
    """A local atom with a non-finite coordinate falls outside every cell.
    The reference's fill_cells skips such an atom silently
    (core/grid.impala:296-311); the library stops with an error at the
    binning instead of running on without it."""
    announce("non-finite position -> error at binning")
    aabb, m, x, v = perturbed_fcc(4, amp=0.05, seed=3)
    x = x.copy()
    x[5, 1] = np.nan
    sim = tm.Simulation(m, x, v, aabb, RC + SKIN)
    sim.copy_to_accelerator()
    with pytest.raises(tm.TinyMDError, match="outside"):
        sim.setup(False, RC, SKIN)
    sim.close()


def test_dem_step_loop_unstaged_windows(tm):
    """The DEM step pass with a window beyond its LDS budget (one 8-cell tile
    over a compressed gas: every neighbour read from the mirrors) against
    the same loop with staged windows: bit-identical positions, velocities
    and forces, rebuilds included."""
    announce("DEM step loop: unstaged (8-cell tile) vs staged windows")
    m, x, v = tm.generate_granular_gas(8, seed=11, vmax=1.0)
    aabb = np.array([0.0, 8 * 1.6796, 0.0, 8 * 1.6796, 0.0, 8 * 1.6796]) * 0.6
    x = x * 0.6

    def run(tile):
        s = tm.Simulation(m, x, v, aabb, 1.3, precision="fp32mixed", dem_radius=0.5)
        if tile:
            s.set_tile(tile)
        s.copy_to_accelerator()
        s.setup(True, 1.0, 0.3)
        s.dem(True, 1.0, 0.5, 0.2, 100.0, 0.3)
        s.run_dem(23, 10, 0.001, True, 1.0, 0.3, 0.5, 0.2, 100.0, 0.3)
        info = s.tile_info()
        _, xs, vs, fs = s.state()
        s.close()
        return info, (xs, vs, fs)

    i_st, r_st = run(0)
    i_un, r_un = run(8)
    kLdsBudget = 80 * 1024 - 1024
    assert (i_st["max_window"] + 1) * 48 <= kLdsBudget
    assert (i_un["max_window"] + 1) * 48 > kLdsBudget
    assert np.abs(r_st[2]).max() > 1.0  # contacts
    for p, q in zip(r_st, r_un):
        assert np.array_equal(p, q)


def test_window_beyond_16bit_offsets_fails_loudly(tm):
    """List entries are 16-bit offsets into a tile's window: a window of
    more than 65,535 atoms (here 8-cell tiles of 6.0-wide cells over an FCC
    liquid-density box, about 180,000 atoms per window) stops the set-up
    with an error naming the limit instead of building wrong lists."""
    announce("tile window beyond the 16-bit offset range -> error")
    m, x, v = tm.generate_fcc(36, 36, 36)
    sim = tm.Simulation(m, x, v, tm.world_aabb(36, 36, 36), 6.0)
    sim.set_tile(8)
    sim.copy_to_accelerator()
    with pytest.raises(tm.TinyMDError, match="16-bit window range"):
        sim.setup(False, RC, SKIN)
    sim.close()
