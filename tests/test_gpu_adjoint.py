"""GPU parity of the differentiable trajectory path: the forward loss
(tmd_traj_loss) and the hand-written adjoint (tmd_traj_adjoint) against the
oracle's loss and its scatter-form adjoint, plus central finite differences
(h = 1e-6) of the library's own loss through the torch.autograd.Function."""
import numpy as np
import pytest

import pyoracle as po
from common import RC, SKIN, announce, perturbed_fcc

pytestmark = pytest.mark.gpu

DT = 0.005
H = 1e-6


@pytest.fixture(scope="module")
def tm():
    import tinymd_amd
    name, cus = tinymd_amd.device_info()
    if cus == 0:
        pytest.fail("no HIP device visible to libtinymd_amd")
    return tinymd_amd


def gpu_sim(tm, system, precision="fp64", eps=1.0, sigma=1.0):
    aabb, m, x, v = system
    sim = tm.Simulation(m, x, v, aabb, RC + SKIN, precision=precision, lj_shift=True)
    sim.copy_to_accelerator()
    return sim


def oracle(system, eps=1.0, sigma=1.0):
    aabb, m, x, v = system
    o = po.OracleSim(m, x, v, aabb, RC + SKIN, threads=4)
    o.setup(False, po.HALF_GHOST_FULL, RC + SKIN)
    o.lj(False, RC, eps, sigma, shift=1)
    return o


@pytest.mark.parametrize("shape,steps,ck", [((4, 4, 4), 40, 20), ((6, 5, 7), 50, 20), ((8, 8, 8), 20, 10)])
def test_traj_loss_matches_oracle(tm, shape, steps, ck):
    announce(f"traj loss {shape} steps={steps} ck={ck}")
    system = perturbed_fcc(*shape, amp=0.05, seed=3)
    sim = gpu_sim(tm, system)
    _, x_before, v_before, _ = sim.state()
    ub, pb = sim.traj_loss(steps, ck, DT, RC, SKIN, 1.0, 1.0)
    uo, po_ = oracle(system).traj_thermo(steps, ck, DT, RC, SKIN, 1.0, 1.0, shift=1)
    assert ub == pytest.approx(uo, rel=1e-12)
    assert pb == pytest.approx(po_, rel=1e-10, abs=1e-12)
    # the call leaves the state at x0
    _, x_after, v_after, _ = sim.state()
    assert np.array_equal(x_before, x_after) and np.array_equal(v_before, v_after)


@pytest.mark.parametrize("shape,steps,ck", [((4, 4, 4), 40, 20), ((6, 5, 7), 45, 20)])
def test_adjoint_matches_oracle_fp64(tm, shape, steps, ck):
    announce(f"adjoint fp64 {shape} steps={steps} ck={ck}")
    system = perturbed_fcc(*shape, amp=0.05, seed=4)
    base = oracle(system).traj_adjoint(steps, ck, DT, RC, SKIN, shift=1)
    tg = (base["ubar"] - 0.1, base["pbar"] + 0.2)
    ro = oracle(system).traj_adjoint(steps, ck, DT, RC, SKIN, shift=1, u_target=tg[0], p_target=tg[1])
    rg = gpu_sim(tm, system).traj_adjoint(steps, ck, DT, RC, SKIN, 1.0, 1.0, tg[0], tg[1])
    assert rg["loss"] == pytest.approx(ro["loss"], rel=1e-10)
    assert rg["grad_eps"] == pytest.approx(ro["grad_eps"], rel=1e-9)
    assert rg["grad_sigma"] == pytest.approx(ro["grad_sigma"], rel=1e-9)
    for key in ("grad_x0", "grad_v0"):
        scale = np.max(np.abs(ro[key]))
        assert np.max(np.abs(rg[key] - ro[key])) <= 1e-9 * scale, key


def test_adjoint_matches_central_differences(tm):
    """TrajectoryLoss (autograd.Function over the library) against central
    differences of tmd_traj_loss, h = 1e-6, FP64."""
    import torch
    from tinymd_amd.autograd import TrajectoryLoss, TrajectoryProblem
    announce("adjoint vs central differences 5x5x5, 60 steps, ck=20")
    aabb, m, x0, v0 = perturbed_fcc(5, amp=0.05, seed=7)
    prob = TrajectoryProblem(m, v0, aabb, steps=60, reneigh=20, dt=DT, cutoff=RC, skin=SKIN)
    base = prob.adjoint(x0, 1.0, 1.0)
    prob.u_target, prob.p_target = base["ubar"] - 0.1, base["pbar"] + 0.2
    x = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
    eps = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    sig = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    L = TrajectoryLoss.apply(x, eps, sig, prob)
    L.backward()
    fe = (prob.loss(x0, 1 + H, 1.0) - prob.loss(x0, 1 - H, 1.0)) / (2 * H)
    fs = (prob.loss(x0, 1.0, 1 + H) - prob.loss(x0, 1.0, 1 - H)) / (2 * H)
    assert float(L.detach()) == pytest.approx(prob.loss(x0, 1.0, 1.0), rel=1e-12)
    assert float(eps.grad) == pytest.approx(fe, rel=1e-6)
    assert float(sig.grad) == pytest.approx(fs, rel=1e-6)
    rng = np.random.default_rng(8)
    d = rng.normal(size=x0.shape)
    d /= np.linalg.norm(d)
    fx = (prob.loss(x0 + H * d, 1.0, 1.0) - prob.loss(x0 - H * d, 1.0, 1.0)) / (2 * H)
    assert float((x.grad.numpy() * d).sum()) == pytest.approx(fx, rel=1e-5, abs=1e-9)


def test_adjoint_fp32_mixed_close_to_fp64(tm):
    announce("adjoint fp32mixed vs fp64 6x6x6, 40 steps")
    system = perturbed_fcc(6, amp=0.05, seed=9)
    r64 = gpu_sim(tm, system).traj_adjoint(40, 20, DT, RC, SKIN, 1.0, 1.0, -4.5, 0.5)
    r32 = gpu_sim(tm, system, precision="fp32mixed").traj_adjoint(40, 20, DT, RC, SKIN, 1.0, 1.0, -4.5, 0.5)
    assert r32["loss"] == pytest.approx(r64["loss"], rel=1e-5)
    assert r32["grad_eps"] == pytest.approx(r64["grad_eps"], rel=1e-4)
    assert r32["grad_sigma"] == pytest.approx(r64["grad_sigma"], rel=1e-4)
    scale = np.max(np.abs(r64["grad_x0"]))
    assert np.max(np.abs(r32["grad_x0"] - r64["grad_x0"])) <= 1e-3 * scale


def test_adjoint_checkpoint_interval_invariance(tm):
    """The gradient does not depend on how the run is cut into intervals
    beyond the list rebuilds it implies: with no rebuild inside the run, one
    interval and the oracle agree."""
    announce("adjoint single interval 5x5x5, 30 steps")
    system = perturbed_fcc(5, amp=0.05, seed=10)
    ro = oracle(system).traj_adjoint(30, 1000, DT, RC, SKIN, shift=1, u_target=-4.0, p_target=1.0)
    rg = gpu_sim(tm, system).traj_adjoint(30, 1000, DT, RC, SKIN, 1.0, 1.0, -4.0, 1.0)
    assert rg["grad_sigma"] == pytest.approx(ro["grad_sigma"], rel=1e-9)
    assert np.max(np.abs(rg["grad_v0"] - ro["grad_v0"])) <= 1e-9 * np.max(np.abs(ro["grad_v0"]))


def test_adjoint_16cube_central_differences_by_precision(tm):
    """Differentiable LJ at full size (FCC 16^3, 16,384 atoms, 200 steps,
    checkpoints every 20, targets from the same state at eps = 1.05,
    sigma = 0.98): parameter gradients against central differences (h = 1e-6)
    of the FP64 loss.  The FP64 adjoint and the FP32-mixed adjoint kernel
    run on the FP64 forward trajectory are within 1e-6 relative; with the
    FP32-mixed forward trajectory as well, the gradient belongs to that
    slightly different trajectory, and its distance from the FP64 finite
    differences is reported and bounded at 1e-5."""
    announce("adjoint 16^3, 200 steps, ck 20: gradients vs central differences by precision")
    steps, ck = 200, 20
    sim = tm.Simulation.fcc(16, precision="fp64", lj_shift=True)
    u_t, p_t = sim.traj_loss(steps, ck, DT, RC, SKIN, 1.05, 0.98)

    def loss(e, s):
        u, p = sim.traj_loss(steps, ck, DT, RC, SKIN, e, s)
        return (u - u_t) ** 2 + (p - p_t) ** 2
    fe = (loss(1 + H, 1.0) - loss(1 - H, 1.0)) / (2 * H)
    fs = (loss(1.0, 1 + H) - loss(1.0, 1 - H)) / (2 * H)
    res = {}
    for fwd, adj in (("fp64", "fp64"), ("fp64", "fp32mixed"), ("fp32mixed", "fp32mixed")):
        sim.set_precision(fwd)
        sim.set_adjoint_precision(adj)
        r = sim.traj_adjoint(steps, ck, DT, RC, SKIN, 1.0, 1.0, u_t, p_t, grad_state=False)
        res[(fwd, adj)] = (abs(r["grad_eps"] - fe) / abs(fe), abs(r["grad_sigma"] - fs) / abs(fs))
        announce(f"  forward {fwd}, adjoint {adj}: relative error vs FD eps {res[(fwd, adj)][0]:.3e} "
                 f"sigma {res[(fwd, adj)][1]:.3e}")
    sim.set_adjoint_precision(None)
    sim.set_precision("fp64")
    sim.close()
    for key in (("fp64", "fp64"), ("fp64", "fp32mixed")):
        assert max(res[key]) <= 1e-6, (key, res[key])
    assert max(res[("fp32mixed", "fp32mixed")]) <= 1e-5, res
