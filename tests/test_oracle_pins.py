"""Pins of the CPU oracle (the checker) -- run on the CPU, no GPU needed.

Every oracle function is checked against something that is not a retyping
of its own formula: the reference's own C++ sources compiled from the
reference tree (oracle/_ref, when present), golden vectors written from them
(tests/golden), textbook values, closed forms, brute force on small inputs,
and the thermo line miniMD printed in the reference's results logs."""
import ctypes
import os

import numpy as np
import pytest

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
RC, SKIN = 2.5, 0.3
A_MINIMD = (4.0 / 0.8442) ** (1.0 / 3.0)


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@pytest.fixture(scope="module")
def ref():
    lib = po.ref_lib()
    if lib is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    return lib


# ---------------------------------------------------------------- RNG + lattice
def test_park_miller_textbook_value():
    # Park & Miller (CACM 1988): seed 1, the 10,000th value is 1043618065
    seed = ctypes.c_int(1)
    for _ in range(10000):
        po.lib().orc_pm_random(ctypes.byref(seed))
    assert seed.value == 1043618065


def test_park_miller_matches_reference(ref):
    for s0 in (1, 7, 12345, 2147483646, 99991):
        a, b = ctypes.c_int(s0), ctypes.c_int(s0)
        for _ in range(200):
            assert po.lib().orc_pm_random(ctypes.byref(a)) == ref.ref_random(ctypes.byref(b))
        assert a.value == b.value


@pytest.mark.parametrize("shape,half", [((4, 4, 4), False), ((5, 3, 7), False), ((6, 6, 6), True), ((16, 16, 16), False)])
def test_lattice_matches_reference_build(ref, shape, half):
    aabb, m, x, v = po.lattice(*shape, half=half)
    cap = 4 * shape[0] * shape[1] * shape[2] + 64
    m2 = np.zeros(cap); x2 = np.zeros((cap, 3)); v2 = np.zeros((cap, 3))
    n = ref.ref_generate_rectangular_grid(_dp(aabb), _dp(aabb), int(half), 1.6796 * 0.5, _dp(m2), _dp(x2), _dp(v2), cap)
    assert n == len(m)
    assert np.array_equal(x, x2[:n]) and np.array_equal(v, v2[:n]) and np.array_equal(m, m2[:n])


def test_lattice_golden_fixture():
    g = np.load(os.path.join(GOLDEN, "ref_lattice_5x3x7.npz"))
    aabb, m, x, v = po.lattice(5, 3, 7)
    assert np.array_equal(x, g["pos"]) and np.array_equal(v, g["vel"])


def test_lattice_closed_form():
    nc = 6
    a = 1.6796
    aabb, m, x, v = po.lattice(nc, nc, nc, a)
    assert len(m) == 4 * nc ** 3
    ijk = np.rint(x / (a / 2)).astype(int)
    assert np.allclose(ijk * (a / 2), x, atol=1e-12)
    assert np.all(ijk.sum(axis=1) % 2 == 0)
    assert len({tuple(r) for r in ijk}) == len(m)
    assert np.all((v >= 0) & (v < 1))


def test_create_velocity_invariants():
    aabb, m, x, v = po.lattice(8, 8, 8)
    s = po.OracleSim(m, x, v, aabb, RC + SKIN)
    s.create_velocity(1.44)
    _, vv, _ = s.state()
    assert np.abs(vv.sum(axis=0)).max() <= 1e-12 * len(m)
    t = np.sum(vv * vv) / (3 * len(m) - 3)
    assert abs(t - 1.44) <= 1e-13


# ---------------------------------------------------------------- LJ force
def _full_list(s):
    counts, lst = s.neighbor_lists()
    return counts, lst


def test_lj_force_matches_reference_force_lj(ref):
    """core/force_lj.cpp (eps = sigma = 1, AoS stride 3, row-major list)
    run on the oracle's own list: forces must agree bit for bit."""
    aabb, m, x, v = po.lattice(6, 6, 6)
    rng = np.random.default_rng(11)
    x = np.mod(x + rng.uniform(-0.12, 0.12, x.shape), aabb[1::2] - aabb[0::2])
    s = po.OracleSim(m, x, v, aabb, RC + SKIN)
    s.setup(False, po.HALF_GHOST_FULL, RC + SKIN)
    s.lj(False, RC)
    _, _, f = s.state()
    counts, lst = _full_list(s)
    n = s.n
    gpos = s.ghosts()["pos"]
    allx = np.ascontiguousarray(np.vstack([x, gpos]))
    fref = np.zeros((len(allx), 3))
    lst = np.ascontiguousarray(lst.astype(np.int32))
    counts = np.ascontiguousarray(counts.astype(np.int32))
    ref.compute_fullneigh(n, RC * RC, _dp(allx), _dp(fref), _ip(lst), _ip(counts), lst.shape[1])
    assert np.array_equal(f, fref[:n])


def test_lj_force_golden_fixture():
    g = np.load(os.path.join(GOLDEN, "ref_force_lj_6cube.npz"))
    aabb, m, x, v = po.lattice(6, 6, 6)
    x = g["pos"]
    s = po.OracleSim(m, x, v, aabb, RC + SKIN)
    s.setup(False, po.HALF_GHOST_FULL, RC + SKIN)
    s.lj(False, RC)
    _, _, f = s.state()
    counts, _ = s.neighbor_lists()
    assert np.array_equal(counts, g["counts"])
    assert np.array_equal(f, g["force"])


def _images_brute(x, L):
    """All 27 images of every atom: positions and (owner, image) labels."""
    imgs = np.array([(a, b, c) for c in (-1, 0, 1) for b in (-1, 0, 1) for a in (-1, 0, 1)])
    pos = (x[None, :, :] + imgs[:, None, :] * L[None, None, :]).reshape(-1, 3)
    owner = np.tile(np.arange(len(x)), len(imgs))
    img = np.repeat(imgs, len(x), axis=0)
    return pos, owner, img


def test_lj_force_brute_force():
    aabb, m, x, v = po.lattice(5, 5, 5)
    rng = np.random.default_rng(12)
    L = aabb[1::2] - aabb[0::2]
    x = np.mod(x + rng.uniform(-0.2, 0.2, x.shape), L)
    s = po.OracleSim(m, x, v, aabb, RC + SKIN)
    s.setup(False, po.HALF_GHOST_FULL, RC + SKIN)
    s.lj(False, RC)
    _, _, f = s.state()
    pos, owner, img = _images_brute(x, L)
    fb = np.zeros_like(x)
    for i in range(len(x)):
        d = x[i] - pos
        r2 = np.sum(d * d, axis=1)
        sel = (r2 < RC * RC) & ~((owner == i) & np.all(img == 0, axis=1))
        sr2 = 1.0 / r2[sel]
        sr6 = sr2 ** 3
        fb[i] = np.sum((48.0 * sr6 * (sr6 - 0.5) * sr2)[:, None] * d[sel], axis=0)
    assert np.max(np.abs(f - fb)) < 1e-10 * max(1.0, np.abs(fb).max())
    assert np.abs(f.sum(axis=0)).max() < 1e-9  # Newton's third law over periodic pairs


# ---------------------------------------------------------------- ghosts, cells, lists
@pytest.mark.parametrize("shape", [(5, 5, 5), (4, 6, 5)])
def test_ghosts_brute_force(shape):
    aabb, m, x, v = po.lattice(*shape)
    rng = np.random.default_rng(13)
    L = aabb[1::2] - aabb[0::2]
    x = np.mod(x + rng.uniform(-0.3, 0.3, x.shape), L)
    s = po.OracleSim(m, x, v, aabb, RC + SKIN)
    s.setup(False, po.HALF_GHOST_FULL, RC + SKIN)
    g = s.ghosts()
    sp = RC + SKIN
    # an image I != 0 of atom o is a ghost iff x_o,d > L_d - s where I_d = -1
    # and x_o,d < s where I_d = +1
    expect = set()
    for o in range(len(x)):
        opts = []
        for d in range(3):
            c = [0]
            if x[o, d] > L[d] - sp:
                c.append(-1)
            if x[o, d] < sp:
                c.append(1)
            opts.append(c)
        for a in opts[0]:
            for b in opts[1]:
                for c in opts[2]:
                    if (a, b, c) != (0, 0, 0):
                        expect.add((o, a, b, c))
    got = {(int(o), *map(int, im)) for o, im in zip(g["owner"], g["img"])}
    assert got == expect and len(g["owner"]) == len(expect)
    assert np.array_equal(g["pos"], x[g["owner"]] + g["img"] * L)


@pytest.mark.parametrize("mode", [po.GHOST_REF_LAG, po.GHOST_FRESH])
def test_ghost_refresh_lag_closed_form(mode):
    """Per-step ghost refresh with every local moving by the same constant
    displacement d per step (no forces; the ghost table fixed at t = 0).

    The reference refreshes the ghost layer in two phases over the ghost
    table built by borders(): first every ghost's value is packed from its
    immediate source (a local, or a ghost of an earlier dimension) as it
    stands, then all ghosts are unpacked (comm/comm.impala:341-387).  A
    ghost whose source is a local is therefore current after refresh t; a
    ghost of a ghost reads its source's value of refresh t - 1, and a ghost
    at chain depth l (1 = its source is a local) lags l - 1 refreshes:
        x_g(t) = x0[owner] + image L + max(0, t - l + 1) d.
    The fresh refresh puts every ghost at owner + image L, so there
    x_g(t) = x0[owner] + image L + t d at any depth."""
    aabb, m, x, v = po.lattice(4, 4, 4)
    rng = np.random.default_rng(17)
    L = aabb[1::2] - aabb[0::2]
    x0 = np.mod(x + rng.uniform(-0.2, 0.2, x.shape), L)
    s = po.OracleSim(m, x0, v, aabb, RC + SKIN)
    s.set_ghost_mode(mode)
    s.setup(False, po.HALF_GHOST_FULL, RC + SKIN)
    g = s.ghosts()
    n = len(x0)
    depth = np.zeros(len(g["src"]), dtype=int)
    for k, src in enumerate(g["src"]):  # sources precede their ghosts
        depth[k] = 1 if src < n else depth[src - n] + 1
    assert depth.max() == 3  # corner ghosts (three chained images) exist in a 4^3 box
    d = np.array([0.0110, -0.0070, 0.0045])
    for t in range(1, 5):
        s.set_state(x=x0 + t * d)
        s.sync_ghosts()
        lag = np.maximum(0, t - depth + 1) if mode == po.GHOST_REF_LAG else np.full(len(depth), t)
        expect = x0[g["owner"]] + g["img"] * L + lag[:, None] * d
        got = s.ghosts()["pos"]
        assert np.abs(got - expect).max() < 1e-12, (t, np.abs(got - expect).max())
    s.close()


def test_ghost_counts_known_sizes():
    for nc, expect in ((16, 13275), (32, 47883)):
        aabb, m, x, v = po.lattice(nc, nc, nc)
        s = po.OracleSim(m, x, v, aabb, RC + SKIN, threads=4)
        s.borders()
        assert s.ng == expect
        s.close()


def test_cells_are_floor_bins_in_index_order():
    aabb, m, x, v = po.lattice(6, 5, 7)
    rng = np.random.default_rng(14)
    x = np.mod(x + rng.uniform(-0.3, 0.3, x.shape), aabb[1::2] - aabb[0::2])
    s = po.OracleSim(m, x, v, aabb, RC + SKIN)
    s.setup(False, po.HALF_GHOST_FULL, RC + SKIN)
    inf = s.info()
    start, parts, pcell = s.cells()
    allx = np.vstack([x, s.ghosts()["pos"]])
    lo = aabb[0::2] - (RC + SKIN)
    ijk = np.floor((allx - lo) / (RC + SKIN)).astype(int)
    flat = (ijk[:, 2] * inf["ny"] + ijk[:, 1]) * inf["nx"] + ijk[:, 0]
    assert np.array_equal(pcell, flat)
    assert start[-1] == len(allx)
    for c in range(inf["ncells"]):
        seg = parts[start[c]:start[c + 1]]
        assert np.all(np.diff(seg) > 0) and np.all(flat[seg] == c)


@pytest.mark.parametrize("half,mode", [(False, po.HALF_GHOST_FULL), (True, po.HALF_GHOST_FULL), (True, po.HALF_REF_LITERAL)])
def test_neighbor_lists_brute_force(half, mode):
    aabb, m, x, v = po.lattice(5, 6, 5)
    rng = np.random.default_rng(15)
    x = np.mod(x + rng.uniform(-0.25, 0.25, x.shape), aabb[1::2] - aabb[0::2])
    s = po.OracleSim(m, x, v, aabb, RC + SKIN)
    s.setup(half, mode, RC + SKIN)
    counts, lst = s.neighbor_lists()
    n = s.n
    allx = np.vstack([x, s.ghosts()["pos"]])
    cut2 = (RC + SKIN) * (RC + SKIN)
    for i in range(n):
        d = allx[i] - allx
        r2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        js = np.nonzero(r2 <= cut2)[0]
        js = js[js != i]
        if half:
            js = js[js < i] if mode == po.HALF_REF_LITERAL else js[(js >= n) | (js < i)]
        assert sorted(lst[i, :counts[i]].tolist()) == js.tolist(), i


@pytest.mark.parametrize("half", [False, True])
def test_list_and_force_cutoff_boundaries(half):
    """Pairs at exactly r_nb^2, one ulp below and one ulp above (the list's
    rsq <= r_nb^2), and the same around rc^2 (the force's rsq < rc^2): the
    oracle's lists equal brute force, the pair at r_nb^2 is kept and the one
    an ulp beyond is not, and the force of a pair at rc^2 (or beyond) is
    exactly zero while one an ulp inside is not."""
    from common import boundary_pairs, brute_force_rsq
    aabb, m, x, v, pairs = boundary_pairs()
    s = po.OracleSim(m, x, v, aabb, RC + SKIN)
    s.setup(half, po.HALF_GHOST_FULL, RC + SKIN)
    assert len(s.ghosts()["pos"]) == 0
    counts, lst = s.neighbor_lists()
    n = s.n
    r2 = brute_force_rsq(x)
    cut2 = (RC + SKIN) * (RC + SKIN)
    for i in range(n):
        js = np.nonzero(r2[i] <= cut2)[0]
        js = js[js != i]
        if half:
            js = js[js < i]
        assert sorted(lst[i, :counts[i]].tolist()) == js.tolist(), i
    for a, p, name, goal in pairs:
        kept = p in lst[a, :counts[a]].tolist() or a in lst[p, :counts[p]].tolist()
        assert kept == (goal <= cut2), name
    assert {name for _, _, name, _ in pairs} >= {"cut2_eq", "cut2_lo", "cut2_hi", "rc2_eq", "rc2_lo", "rc2_hi"}
    s.lj(half, RC)
    _, _, f = s.state()
    rc2 = RC * RC
    checked = set()
    for a, p, name, goal in pairs:
        # anchors whose only neighbour within rc is (or would be) the partner
        others = [j for j in np.nonzero(r2[a] < rc2)[0] if j not in (a, p)]
        if others:
            continue
        checked.add(name)
        if goal < rc2:
            assert np.any(f[a] != 0.0), name
            assert np.array_equal(f[a], -f[p]), name  # one pair: equal and opposite
        else:
            assert np.array_equal(f[a], np.zeros(3)), name
    assert checked >= {"rc2_eq", "rc2_lo", "rc2_hi"}
    s.close()


def test_dem_contact_boundary():
    """Spheres of diameter 1 at exactly touching distance, one ulp inside
    and one ulp apart, the partner moving: only the pair one ulp inside is
    in contact (rsq < (2R)^2); its dashpot force is non-zero and the pair's
    forces are equal and opposite, the others exactly zero."""
    from common import boundary_pairs
    one = 1.0
    targets = [("d2_eq", one), ("d2_lo", np.nextafter(one, 0.0)), ("d2_hi", np.nextafter(one, 2.0))]
    aabb, m, x, v, pairs = boundary_pairs(seed=8, targets=targets)
    v = v.copy()
    v[27:] = [0.3, -0.2, 0.1]
    # cutoff 1.1 > 2R: the contact test alone decides at rsq = (2R)^2
    for half in (False, True):
        s = po.OracleSim(m, x, v, aabb, 1.4)
        s.setup(half, po.HALF_GHOST_FULL, 1.4)
        s.dem(half, 1.1, 0.5, 0.5, 0.2, 100.0, 0.3)
        _, _, f = s.state()
        for a, p, name, goal in pairs:
            if goal < one:
                assert np.any(f[a] != 0.0) and np.array_equal(f[a], -f[p]), name
            else:
                assert np.array_equal(f[a], np.zeros(3)) and np.array_equal(f[p], np.zeros(3)), name
        s.close()


def test_neighbor_counts_perfect_lattice():
    s = po.lj_benchmark(8)
    counts, lst = s.neighbor_lists()
    assert np.all(counts == 78)
    n = s.n
    ll = sum(int(np.sum(lst[i, :counts[i]] < n)) for i in range(n))
    lg = int(counts.sum()) - ll
    # half lists: every local-local pair once, every local-ghost entry kept
    # (GHOST_FULL) or dropped (the reference's literal filter)
    s2 = po.lj_benchmark(8, half=True)
    c2, _ = s2.neighbor_lists()
    assert int(c2.sum()) == ll // 2 + lg
    s3 = po.lj_benchmark(8, half=True, half_mode=po.HALF_REF_LITERAL)
    c3, _ = s3.neighbor_lists()
    assert int(c3.sum()) == ll // 2
    assert s.thermo()["pairs"] == 54 * n


def test_capacity_retry_is_transparent():
    aabb, m, x, v = po.lattice(5, 5, 5)
    a = po.OracleSim(m, x, v, aabb, RC + SKIN, nb_cap=1)
    b = po.OracleSim(m, x, v, aabb, RC + SKIN, nb_cap=200)
    a.setup(False, 0, RC + SKIN)
    b.setup(False, 0, RC + SKIN)
    ca, la = a.neighbor_lists()
    cb, lb = b.neighbor_lists()
    assert np.array_equal(ca, cb)
    for i in range(len(ca)):
        assert np.array_equal(la[i, :ca[i]], lb[i, :cb[i]])


# ---------------------------------------------------------------- thermo known answers
def test_lattice_energy_and_pressure_known_answers():
    s = po.lj_benchmark(8, potmin=A_MINIMD)
    th = s.thermo()
    n = s.n
    # miniMD, step 0, 64x64x32 at density 0.8442: T=1.44 U=-6.773368 P=-5.019672
    # (results/single_gpu_v100.txt:39).  A perfect periodic lattice gives the
    # same per-atom PE and virial at any size, so U and P of the 524,288-atom
    # box follow from this 2,048-atom box exactly.
    assert abs(th["PE"] / n - (-6.773368)) < 5e-7
    N = 64 * 64 * 32 * 4
    rho = 0.8442
    P = rho * (1.44 * (3 * N - 3) / (3 * N) + th["virial"] / n / 3.0)
    assert abs(P - (-5.019672)) < 5e-7
    s2 = po.lj_benchmark(8)
    assert abs(s2.thermo()["PE"] / s2.n - (-6.7733178)) < 5e-8


def test_vv_time_reversibility_and_energy():
    aabb, m, x, v = po.lattice(5, 5, 5)
    s = po.OracleSim(m, x, v, aabb, RC + SKIN)
    s.create_velocity(0.7)
    s.setup(False, 0, RC + SKIN)
    s.lj(False, RC, shift=1)
    x0, v0, _ = s.state()
    th0 = s.thermo()
    s.run_lj(200, reneigh=1000, dt=0.002, shift=1)
    th1 = s.thermo()
    e0, e1 = th0["KE"] + th0["PE"], th1["KE"] + th1["PE"]
    assert abs(e1 - e0) / abs(e0) < 2e-5
    x1, v1, _ = s.state()
    s.set_state(x1, -v1)
    s.run_lj(200, reneigh=1000, dt=0.002, shift=1)
    x2, v2, _ = s.state()
    assert np.max(np.abs(x2 - x0)) < 1e-9 and np.max(np.abs(v2 + v0)) < 1e-9


def test_vv_free_flight():
    # dilute atoms never interact: x(t) = x0 + v t
    L = 40.0
    x = np.array([[5.0, 5.0, 5.0], [20.0, 20.0, 20.0], [33.0, 8.0, 14.0]])
    v = np.array([[0.3, -0.2, 0.1], [-0.1, 0.05, 0.2], [0.0, 0.4, -0.3]])
    aabb = np.array([0.0, L, 0.0, L, 0.0, L])
    s = po.OracleSim(np.ones(3), x, v, aabb, RC + SKIN)
    s.setup(False, 0, RC + SKIN)
    s.lj(False, RC)
    s.run_lj(100, reneigh=1000, dt=0.01)
    x1, v1, _ = s.state()
    assert np.allclose(x1, x + v * 1.0, atol=1e-12) and np.array_equal(v1, v)


def test_minimd_trajectory_soft_pin():
    """miniMD step 100 (dt 0.001, 64x64x32, rebuild every 20):
    T=0.7192753 U=-5.698859 P=0.2176028 (results/single_gpu_v100.txt:40)."""
    aabb, m, x, v = po.lattice(64, 64, 32, A_MINIMD)
    s = po.OracleSim(m, x, v, aabb, RC + SKIN, threads=min(os.cpu_count() or 1, 8))
    s.create_velocity(1.44)
    s.setup(False, 0, RC + SKIN)
    s.lj(False, RC)
    s.run_lj(100, reneigh=20, dt=0.001)
    th = s.thermo()
    assert abs(th["T"] - 0.7192753) < 2e-6
    assert abs(th["PE"] / s.n - (-5.698859)) < 2e-6
    assert abs(th["P"] - 0.2176028) < 2e-5
    s.close()


# ---------------------------------------------------------------- DEM
def test_dem_head_on_restitution():
    """Two unit-mass spheres (R = 0.5) approach head on at relative speed 2;
    the linear spring-dashpot gives e = exp(-pi beta / omega),
    beta = g_n / (2 m*), omega = sqrt(k / m* - beta^2), m* = 1/2."""
    k, gn = 100.0, 0.5
    L = 30.0
    x = np.array([[14.0, 15.0, 15.0], [16.0, 15.0, 15.0]])
    v = np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]])
    aabb = np.array([0.0, L, 0.0, L, 0.0, L])
    s = po.OracleSim(np.ones(2), x, v, aabb, 1.3)
    s.setup(True, 0, 1.3)
    s.dem(True, 1.0, 0.5, gn, 0.0, k, 0.0)
    s.run_dem(20000, reneigh=100, dt=1e-4, half=True, rc=1.0, radius=0.5, gn=gn, gt=0.0, k=k, mu=0.0)
    _, v1, _ = s.state()
    mstar, beta = 0.5, gn / (2 * 0.5)
    omega = np.sqrt(k / mstar - beta ** 2)
    e = np.exp(-np.pi * beta / omega)
    e_num = (v1[1, 0] - v1[0, 0]) / 2.0
    assert abs(e_num - e) < 2e-3, (e_num, e)
    assert abs(v1.sum(axis=0)).max() < 1e-12  # momentum


def test_dem_tangential_coulomb_cap():
    """Overlapping pair sliding tangentially: |F_t| = min(g_t |v_t|, mu |F_n|)."""
    L = 30.0
    x = np.array([[15.0, 15.0, 15.0], [15.9, 15.0, 15.0]])  # overlap 0.1
    v = np.array([[0.0, 0.0, 0.0], [0.0, 2.0, 0.0]])
    aabb = np.array([0.0, L, 0.0, L, 0.0, L])
    for gt, mu in ((0.2, 10.0), (50.0, 0.3)):
        s = po.OracleSim(np.ones(2), x, v, aabb, 1.3)
        s.setup(False, 0, 1.3)
        s.dem(False, 1.0, 0.5, 0.5, gt, 100.0, mu)
        _, _, f = s.state()
        fn = 100.0 * 0.1  # no normal relative velocity
        assert abs(f[0, 0] - (-fn)) < 1e-12
        assert abs(f[0, 1] - min(gt * 2.0, mu * fn)) < 1e-12  # i is dragged along +y
        assert np.allclose(f[0], -f[1], atol=1e-14)


def test_sort_locals_cell_order_and_same_physics():
    """The oracle's cell-order renumbering (sort mode): after a rebuild the
    locals are ordered by cell, index order inside a cell (the stable order
    a counting sort gives, computed here with numpy from the cell formula of
    core/grid.impala:406-426); the positions, velocities and forces are the
    unsorted run's, permuted (forces to 1e-12: the ghost layer is built
    from the new numbering, so summation orders differ)."""
    aabb, m, x, v = po.lattice(6, 5, 7)
    rng = np.random.default_rng(31)
    L = aabb[1::2] - aabb[0::2]
    x = np.mod(x + rng.uniform(-0.1, 0.1, x.shape), L)
    cut = RC + SKIN
    a = po.OracleSim(m, x, v, aabb, cut)
    b = po.OracleSim(m, x, v, aabb, cut)
    b.set_sort(True)
    a.create_velocity(1.44)
    b.create_velocity(1.44)
    for s in (a, b):
        s.enforce_pbc()
        s.setup(False, po.HALF_GHOST_FULL, cut)
        s.lj(False, RC)
    xa, va, fa = a.state()
    xb, vb, fb = b.state()
    lo = aabb[0::2] - cut
    inf = a.info()
    ijk = np.floor((xa - lo) / cut).astype(np.int64)
    key = (ijk[:, 2] * inf["ny"] + ijk[:, 1]) * inf["nx"] + ijk[:, 0]
    order = np.argsort(key, kind="stable")
    assert np.array_equal(xb, xa[order]) and np.array_equal(vb, va[order])
    rms = np.sqrt(np.mean(np.sum(fa * fa, axis=1)))
    assert np.abs(fb - fa[order]).max() <= 1e-12 * rms
    kb = key[order]
    assert np.all(np.diff(kb) >= 0)
    # a short run with rebuilds: still the same trajectory up to the labels
    a.run_lj(25, 10, 0.005)
    b.run_lj(25, 10, 0.005)
    xa, _, _ = a.state()
    xb, _, _ = b.state()
    # both runs wrap at every rebuild; match atoms by their nearest periodic image
    from scipy.spatial import cKDTree
    d, idx = cKDTree(np.mod(xa, L), boxsize=L).query(np.mod(xb, L))
    assert d.max() < 1e-9 and len(set(idx.tolist())) == len(idx)
    a.close()
    b.close()
