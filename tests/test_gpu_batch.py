"""Batched replicas on the GPU against the same boxes run one at a time
through the md_* path: each replica's positions, velocities and forces are
bit-identical (FP64 and FP32-mixed), for any tile and block size, across
capacity retries, and with a non-default cell spacing when the single run
uses the same spacing."""
import numpy as np
import pytest

from common import RC, SKIN, announce

pytestmark = pytest.mark.gpu
DT, REN = 0.005, 20


@pytest.fixture(scope="module")
def tm():
    import tinymd_amd
    name, cus = tinymd_amd.device_info()
    if cus == 0:
        pytest.fail("no HIP device visible to libtinymd_amd")
    return tinymd_amd


@pytest.fixture(scope="module")
def boxes(tm):
    return tm.replica_boxes(seed=7, nrep=5, cells=(6, 9))


def single(tm, box, steps, precision, spacing=1.0, runs=None, ren=REN):
    aabb, x, v = box
    sim = tm.Simulation(np.ones(len(x)), x, v, aabb, (RC + SKIN) * spacing, precision=precision)
    sim.copy_to_accelerator()
    sim.setup(False, RC, SKIN)
    sim.lennard_jones(False, RC)
    for s in (runs or [steps]):
        sim.run_lj(s, ren, DT, False, RC, SKIN)
    _, xs, vs, fs = sim.state()
    th = sim.thermo(RC)
    ng = sim.nghost
    sim.close()
    return xs, vs, fs, th, ng


def assert_bitwise(tm, b, boxes, steps, precision, spacing=1.0, runs=None, ren=REN):
    for r, box in enumerate(boxes):
        xs, vs, fs, _, _ = single(tm, box, steps, precision, spacing, runs, ren)
        xb, vb, fb = b.state(r)
        assert np.array_equal(xb, xs), f"replica {r} positions"
        assert np.array_equal(vb, vs), f"replica {r} velocities"
        assert np.array_equal(fb, fs), f"replica {r} forces"


@pytest.mark.parametrize("precision", ["fp64", "fp32mixed"])
def test_batch_bitwise_vs_single(tm, boxes, precision):
    announce(f"batch 5 replicas {precision} 45 steps vs single runs")
    b = tm.Batch(boxes, precision=precision)
    assert b.run(45) == 45
    assert_bitwise(tm, b, boxes, 45, precision)
    info = b.info()
    for r, box in enumerate(boxes):
        assert info[r]["ghosts"] == single(tm, box, 45, precision)[4]
    b.close()


def test_batch_short_reneighbour(tm, boxes):
    """Rebuilds every third step: each rebuild copies the replicas' state
    back from the step loop's mirrors and rebuilds their ghosts; the
    replicas stay bit-identical to single runs with the same interval."""
    announce("batch 5 replicas fp64, rebuild every 3 steps, 14 steps")
    b = tm.Batch(boxes, precision="fp64", reneigh=3)
    assert b.run(14) == 14
    assert_bitwise(tm, b, boxes, 14, "fp64", ren=3)
    b.close()


def test_batch_thermo_vs_single(tm, boxes):
    announce("batch thermo vs single thermo")
    b = tm.Batch(boxes, precision="fp64")
    b.run(30)
    th = b.thermo()
    for r, box in enumerate(boxes):
        ts = single(tm, box, 30, "fp64")[3]
        for k in ("T", "KE", "PE", "virial", "P"):
            assert th[r][k] == pytest.approx(ts[k], rel=1e-12, abs=1e-12), (r, k)
        assert th[r]["pairs"] == ts["pairs"]
    b.close()


@pytest.mark.parametrize("tile,threads", [(2, 256), (4, 512), (1, 64), (3, 0), (3, 768)])
def test_batch_launch_config_invariance(tm, boxes, tile, threads):
    announce(f"batch config tile={tile} threads={threads}")
    b = tm.Batch(boxes, precision="fp64")
    b.configure(tile=tile, threads=threads)
    b.run(45)
    assert_bitwise(tm, b, boxes, 45, "fp64")
    b.close()


def test_batch_cell_spacing_matches_single_with_same_spacing(tm, boxes):
    announce("batch spacing factor 1.1")
    b = tm.Batch(boxes, precision="fp64")
    b.configure(spacing=1.1)
    b.run(45)
    assert_bitwise(tm, b, boxes, 45, "fp64", spacing=1.1)
    b.close()


def test_batch_ghost_and_list_overflow_retry(tm, boxes):
    """Deliberately small ghost room and list headroom: the set-up grows the
    ghost room and the list capacity; results stay bit-identical."""
    announce("batch capacity retry")
    b = tm.Batch(boxes, precision="fp64")
    b.configure(ghost_headroom=0.25, list_headroom=1.0)
    b.run(45)
    assert_bitwise(tm, b, boxes, 45, "fp64")
    info = b.info()
    assert all(i["ghost_room"] >= i["ghosts"] for i in info)
    b.close()


def test_batch_consecutive_runs(tm, boxes):
    announce("batch run(25) + run(20) vs single run_lj(25) + run_lj(20)")
    b = tm.Batch(boxes, precision="fp64")
    b.run(25)
    b.run(20)
    assert_bitwise(tm, b, boxes, 45, "fp64", runs=[25, 20])
    b.close()


def test_batch_autotune_keeps_results(tm, boxes):
    announce("batch autotune")
    b = tm.Batch(boxes, precision="fp32mixed")
    cfg = b.autotune(trial_steps=20)
    assert cfg["threads"] % 64 == 0 and 64 <= cfg["threads"] <= 1024
    assert cfg["spacing"] in (1.0, 1.1) and cfg["ms_per_step"] > 0
    b.run(45)
    assert_bitwise(tm, b, boxes, 45, "fp32mixed", spacing=cfg["spacing"])
    b.close()


def test_batch_graph_reuse_across_tile_edges(tm, boxes):
    """A run captured under one tile edge is not reused under another: the
    launch shapes and LDS sizes a graph bakes in follow the configuration.
    configure(tile 3) run, configure(tile 4) run: bit-identical to a fresh
    batch configured with tile 4."""
    announce("batch graph reuse: tile 3 run, tile 4 run vs fresh tile 4")
    b = tm.Batch(boxes, precision="fp64")
    b.configure(tile=3)
    b.run(25)
    b.configure(tile=4)
    b.run(25)
    f = tm.Batch(boxes, precision="fp64")
    f.configure(tile=4)
    f.run(25)
    for r in range(len(boxes)):
        for a, c in zip(b.state(r), f.state(r)):
            assert np.array_equal(a, c), r
    b.close()
    f.close()


def expanding_box(n_side=9, spacing=1.25, L=22.4, rate=9.0):
    """A simple-cubic block of atoms in the middle of a periodic box, with
    velocities pointing away from the centre (v = rate * (x - c)): no atom is
    within a cell spacing of a face at t = 0 (no ghosts), but after 20 steps
    the outer layers are, so the ghost room sized at set-up overflows at the
    first rebuild inside the run."""
    g = (np.arange(n_side) - (n_side - 1) / 2.0) * spacing
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    d = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    c = L / 2.0
    return np.array([0.0, L, 0.0, L, 0.0, L]), c + d, rate * d


def test_batch_ghost_overflow_inside_run(tm):
    """Ghost room exhausted at the step-20 rebuild of a run: the run restores
    its starting state, grows the room, rebuilds and runs again.  The result
    is bit-identical to the single-box path, and the room grew."""
    announce("batch ghost overflow at a rebuild inside the run (expanding block)")
    box = expanding_box()
    box2 = expanding_box(n_side=8, spacing=1.3)
    b = tm.Batch([box, box2], precision="fp64")
    b.configure(ghost_headroom=0.25)
    room0 = [i["ghost_room"] for i in b.info()]
    assert all(i["ghosts"] == 0 for i in b.info())
    assert b.run(25) == 25
    info = b.info()
    assert any(i["ghost_room"] > r0 for i, r0 in zip(info, room0)), (room0, info)
    assert all(i["ghost_room"] >= i["ghosts"] for i in info)
    for r, bx in enumerate((box, box2)):
        xs, vs, fs, _, _ = single(tm, bx, 25, "fp64")
        xb, vb, fb = b.state(r)
        assert np.array_equal(xb, xs) and np.array_equal(vb, vs) and np.array_equal(fb, fs), r
    b.close()
