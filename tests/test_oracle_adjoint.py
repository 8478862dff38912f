"""Pins of the oracle's trajectory loss and its adjoint (CPU only).

The loss (orc_traj_thermo) is checked against the step loop plus per-step
thermo; the adjoint (orc_traj_adjoint, scatter form of J^T) against central
finite differences of the loss with h = 1e-6 in FP64, for eps, sigma and
random directions in x0 and v0, across a reneighbouring boundary."""
import numpy as np
import pytest

import pyoracle as po
from common import RC, SKIN, perturbed_fcc

STEPS, CK, DT = 40, 20, 0.005
H = 1e-6


def make(aabb, m, x, v, eps=1.0, sigma=1.0):
    o = po.OracleSim(m, x, v, aabb, RC + SKIN)
    o.setup(False, po.HALF_GHOST_FULL, RC + SKIN)
    o.lj(False, RC, eps, sigma, shift=1)
    return o


@pytest.fixture(scope="module")
def system():
    return perturbed_fcc(4, amp=0.05, seed=11)


def loss(system, eps=1.0, sigma=1.0, x=None, v=None, targets=(0.0, 0.0)):
    aabb, m, x0, v0 = system
    o = make(aabb, m, x0 if x is None else x, v0 if v is None else v, eps, sigma)
    u, p = o.traj_thermo(STEPS, CK, DT, RC, SKIN, eps, sigma, shift=1)
    o.close()
    return (u - targets[0]) ** 2 + (p - targets[1]) ** 2


def test_traj_thermo_matches_step_loop(system):
    """Ubar, Pbar over one reneighbouring interval equal the mean of the
    per-step thermo of the plain step loop."""
    aabb, m, x0, v0 = system
    u1, p1 = make(aabb, m, x0, v0).traj_thermo(CK, CK, DT, RC, SKIN, 1.0, 1.0, shift=1)
    o = make(aabb, m, x0, v0)
    us, ps = [], []
    for _ in range(CK - 1):
        o.run_lj(1, reneigh=1_000_000, dt=DT, shift=1)
        th = o.thermo()
        us.append(th["PE"] / o.n)
        ps.append(th["P"])
    o.run_lj(1, reneigh=1, dt=DT, shift=1)  # the interval's last step rebuilds
    th = o.thermo()
    us.append(th["PE"] / o.n)
    ps.append(th["P"])
    assert u1 == pytest.approx(np.mean(us), rel=1e-12)
    assert p1 == pytest.approx(np.mean(ps), rel=1e-10)


def test_adjoint_matches_central_differences(system):
    aabb, m, x0, v0 = system
    base = make(aabb, m, x0, v0).traj_adjoint(STEPS, CK, DT, RC, SKIN, shift=1)
    tg = (base["ubar"] - 0.1, base["pbar"] + 0.2)
    r = make(aabb, m, x0, v0).traj_adjoint(STEPS, CK, DT, RC, SKIN, shift=1, u_target=tg[0], p_target=tg[1])
    assert r["loss"] == pytest.approx(loss(system, targets=tg), rel=1e-12)
    fe = (loss(system, 1 + H, targets=tg) - loss(system, 1 - H, targets=tg)) / (2 * H)
    fs = (loss(system, sigma=1 + H, targets=tg) - loss(system, sigma=1 - H, targets=tg)) / (2 * H)
    assert r["grad_eps"] == pytest.approx(fe, rel=1e-6)
    assert r["grad_sigma"] == pytest.approx(fs, rel=1e-6)
    rng = np.random.default_rng(5)
    for key, which in (("grad_x0", "x"), ("grad_v0", "v")):
        d = rng.normal(size=x0.shape)
        d /= np.linalg.norm(d)
        a0 = x0 if which == "x" else v0
        lp = loss(system, **{which: a0 + H * d}, targets=tg)
        lm = loss(system, **{which: a0 - H * d}, targets=tg)
        fd = (lp - lm) / (2 * H)
        ad = float(np.sum(r[key] * d))
        assert ad == pytest.approx(fd, rel=1e-5, abs=1e-9), key


def test_adjoint_planted_mistake_is_caught(system):
    """A wrong adjoint (sigma derivative of the truncation constants dropped,
    emulated by differentiating with shift=0 on a shifted loss) fails the
    finite-difference check: the check has power."""
    aabb, m, x0, v0 = system
    o = make(aabb, m, x0, v0)
    r_noshift = o.traj_adjoint(STEPS, CK, DT, RC, SKIN, shift=0, u_target=-5.0, p_target=0.5)
    fs = (loss(system, sigma=1 + H, targets=(-5.0, 0.5)) - loss(system, sigma=1 - H, targets=(-5.0, 0.5))) / (2 * H)
    assert abs(r_noshift["grad_sigma"] - fs) > 1e-4 * abs(fs)
