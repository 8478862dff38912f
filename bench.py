#!/usr/bin/env python3
"""Benchmark of the tinyMD per-timestep path on one MI355X.

Workload: Lennard-Jones FCC lattice of 80x80x80 unit cells (2,048,000 atoms,
density 0.8442, lattice constant 1.6796), Park-Miller velocities rescaled to
T = 1.44, eps = sigma = 1, cutoff 2.5, skin 0.3, dt 0.005, reneighbour every
20 steps, full neighbour list.  A step is one full timestep (integrate,
ghost refresh or rebuild, pair force, integrate) over all atoms, with all
state resident in HBM.

Printed value: atom-steps/s of the FP32-mixed path (FP32 pair forces with
compensated accumulation, FP64 positions/velocities/ghosts/reductions);
the FP64 path, the half-list variants and the FP32-vs-FP64 force error are
reported beside it.  `roofline` is for the dominant kernel (the pair-force
kernel), timed with HIP events on the library's stream; `cpu_baseline` is
the FP64 CPU oracle on the host cores over a bounded sample of the same
workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "molecular-dynamics_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
XGMI_LINK_GBS = 153.0     # one MI355X xGMI link, per direction (GB/s)
FP64_PEAK_TFLOPS = 78.6   # MI355X vector FP64
FP32_PEAK_TFLOPS = 157.3  # MI355X vector FP32
METRIC = "LJ atom-steps/s, FP32-mixed vs FP64 at ≤1e-5 force err; adjoint/fwd cost"
NC = 80
RC, SKIN, DT, REN = 2.5, 0.3, 0.005, 20


def force_kernel_model(sim, rc=RC):
    """Algorithmic bytes and flops of one full-list pair-force launch at the
    current state, by the compulsory-traffic model of the workload: each
    array read once, each output written once, neighbour gathers assumed
    cache-resident after their first touch.

      bytes = 4 N (list counts) + 4 entries (4-byte list entries, the
              reference's i32 list) + 24 N_all (FP64 positions, locals and
              ghosts, once) + 24 N (one FP64 3-vector out per atom)
      flops = 23 per in-cutoff entry (del 3, rsq 5, 1/rsq 1, sr6 3, scale 5,
              accumulate 6; an FMA counts 2) + 8 per entry in the skin
              (rc < r <= rc + skin: del and rsq)

    Also returned: the same bytes with the 2-byte entries this library
    stores, and a model of the fused step-loop launch (the kick and drift
    read and write the velocity, 48 N, and each work slot reads three 4-byte
    words, 12 N, instead of the count alone)."""
    counts, _ = sim.neighbor_lists(with_lists=False)
    n = sim.nlocal
    nall = n + sim.nghost
    entries = int(counts.astype(np.int64).sum())
    in_cut = int(sim.thermo(rc)["pairs"])
    b = 4 * n + 4 * entries + 24 * nall + 24 * n
    return {"atoms": n, "atoms_with_ghosts": nall, "list_entries": entries, "in_cutoff_entries": in_cut,
            "bytes": b, "bytes_2B_entries": b - 2 * entries,
            "bytes_fused_step_model": 12 * n + 4 * entries + 24 * nall + 24 * n + 48 * n,
            "flops": 23 * in_cut + 8 * (entries - in_cut)}


# the reference's phase timers (main.cpp:716-762): Force includes both
# integration half-steps, Neigh the list build, Comm the ghost borders and
# per-step ghost refresh, Other the PBC wrap, binning and (here) the
# cell-ordered mirror
PHASES = {"force": ("force", "integrate1", "integrate2"), "neigh": ("neighbor",),
          "comm": ("borders", "ghost_sync"), "other": ("bin", "mirror")}


def run_variant(tm, precision, half, steps, warmup, profile=False, nc=None):
    """Timed run of `steps` steps (no per-kernel timers), then, with
    profile, a second run of the same length with HIP-event timers on the
    library stream for per-kernel averages and the reference's phase split.
    Energy drift is taken with the energy-shifted potential U(r) - U(rc)
    (continuous at the cutoff) plus the kinetic energy."""
    sim = tm.Simulation.fcc(nc or NC, precision=precision)
    sim.setup(half, RC, SKIN)
    sim.lennard_jones(half, RC)
    if warmup:
        sim.run_lj(warmup, REN, DT, half, RC, SKIN)
    sim.barrier()
    th0 = sim.thermo(RC, eshift=True)
    t0 = time.perf_counter()
    done = sim.run_lj(steps, REN, DT, half, RC, SKIN)
    sim.barrier()
    t1 = time.perf_counter()
    assert done == steps
    th1 = sim.thermo(RC, eshift=True)
    e0, e1 = th0["KE"] + th0["PE"], th1["KE"] + th1["PE"]
    out = {"seconds": t1 - t0, "atom_steps_per_s": sim.nlocal * steps / (t1 - t0),
           "energy_per_atom_start": e0 / sim.nlocal, "energy_per_atom_end": e1 / sim.nlocal,
           "energy_drift_rel": (e1 - e0) / abs(e0), "energy_form": "KE + sum over in-cutoff pairs of U(r) - U(rc)",
           "T_end": th1["T"], "P_end": th1["P"]}
    if profile:
        sim.profile(True)
        sim.profile_reset()
        sim.run_lj(steps, REN, DT, half, RC, SKIN)
        sim.barrier()
        ms, cnt = sim.profile_get("force")
        out["force_ms_avg"] = ms / max(cnt, 1)
        out["force_launches"] = cnt
        per = {}
        for k in ("force", "integrate1", "integrate2", "ghost_sync", "mirror", "neighbor", "bin", "borders"):
            kms, kc = sim.profile_get(k)
            per[k] = kms
            if k != "force":
                out[k + "_ms_avg"] = kms / max(kc, 1)
        out["phases_ms_per_step"] = {ph: sum(per[k] for k in ks) / steps for ph, ks in PHASES.items()}
        sim.profile(False)
    return sim, out


def reduced_force_error(tm, sim, precision):
    """Per-atom relative force error of a reduced-precision kernel against
    the FP64 kernel on the same (thermalised) positions and neighbour list:
    |F_i - F64_i| / max(|F64_i|, floor) with the floor the RMS of |F64_i|,
    as RMS and maximum over the atoms."""
    sim.set_precision("fp64")
    sim.lennard_jones(False, RC)
    f64 = sim.state()[3]
    sim.set_precision(precision)
    sim.lennard_jones(False, RC)
    f32 = sim.state()[3]
    sim.set_precision("fp32mixed")
    fn = np.sqrt(np.sum(f64 * f64, axis=1))
    floor = np.sqrt(np.mean(fn * fn))
    d = np.sqrt(np.sum((f32 - f64) ** 2, axis=1)) / np.maximum(fn, floor)
    return {"rms_rel": float(np.sqrt(np.mean(d * d))), "max_rel": float(np.max(d)), "floor": float(floor)}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline():
    """FP64 oracle (test infrastructure, the checker, built with -O3
    -fno-fast-math -ffp-contract=off -fopenmp) on the host cores, timed over
    the region main.cpp:712-778 times (after the initial build):
      - FCC 32^3 (131,072 atoms), 100 steps, 1 thread and `threads` threads;
      - FCC 80^3 (2,048,000 atoms), 200 steps on `threads` threads, and a
        20-step sample (one rebuild) on 1 thread, which stands for the
        200-step run (the step cost is flat: one rebuild per 20 steps).
    `value` is the 80^3 rate on `threads` threads (the GPU workload, FP64)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    threads = min(os.cpu_count() or 1, 16)
    runs = {}
    for nc, steps, th in ((32, 100, 1), (32, 100, threads), (80, 200, threads), (80, 20, 1)):
        s = po.lj_benchmark(nc, threads=th)
        t0 = time.perf_counter()
        s.run_lj(steps, REN, DT)
        t1 = time.perf_counter()
        n = s.n
        s.close()
        runs[f"{nc}cube_{steps}steps_{th}thread"] = {"atoms": n, "steps": steps, "threads": th, "seconds": t1 - t0,
                                                       "atom_steps_per_s": n * steps / (t1 - t0)}
        print(f"# cpu oracle {nc}^3 {steps} steps {th} threads: {n * steps / (t1 - t0):.4g} atom-steps/s",
              file=sys.stderr, flush=True)
    head = runs[f"80cube_200steps_{threads}thread"]
    return {"value": head["atom_steps_per_s"], "unit": "atom-steps/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle, FCC 80^3 ({head['atoms']} atoms), 200 steps incl. 10 rebuilds, full list, "
                      f"{threads} OpenMP threads (list build and pair force threaded), {head['seconds']:.1f} s; "
                      f"also 32^3 x 100 steps on 1 and {threads} threads and 80^3 x 20 steps on 1 thread (runs)",
            "runs": runs, "cpu_model": cpu_model(), "nproc": os.cpu_count()}


def c1_leg(tm, nc=32, steps=100, cpu=True):
    """The reference's default run (main.cpp defaults): FCC 32^3 (131,072
    atoms), FP64, full list, 100 steps, on the GPU and on the CPU oracle
    (the checker) from the same initial state; the final thermo of both."""
    sim = tm.Simulation.fcc(nc, precision="fp64")
    sim.setup(False, RC, SKIN)
    sim.lennard_jones(False, RC)
    sim.barrier()
    t0 = time.perf_counter()
    sim.run_lj(steps, REN, DT, False, RC, SKIN)
    sim.barrier()
    t1 = time.perf_counter()
    th = sim.thermo(RC)
    n = sim.nlocal
    sim.close()
    out = {"atoms": n, "steps": steps, "gpu_seconds": t1 - t0, "gpu_atom_steps_per_s": n * steps / (t1 - t0),
           "gpu_T": th["T"], "gpu_PE": th["PE"], "gpu_P": th["P"]}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        threads = min(os.cpu_count() or 1, 16)
        s = po.lj_benchmark(nc, threads=threads)
        s.run_lj(steps, REN, DT)
        s.lj(False, RC)
        tc = s.thermo()
        s.close()
        out.update({"cpu_T": tc["T"], "cpu_PE": tc["PE"], "cpu_P": tc["P"]})
    return out


def adjoint_leg(tm, precision, nc=16, steps=200, ck=20, fd=False):
    """Differentiable LJ: FCC nc^3 (16^3: 16,384 atoms), `steps` velocity-Verlet
    steps with checkpoints every `ck`; loss (Ubar - U*)^2 + (Pbar - P*)^2 with
    targets from the same initial state at eps = 1.05, sigma = 0.98.  Times the
    forward loss (tmd_traj_loss) and the adjoint (tmd_traj_adjoint, which
    includes its own forward pass and the interval recomputation)."""
    sim = tm.Simulation.fcc(nc, precision=precision, lj_shift=True)
    u_t, p_t = sim.traj_loss(steps, ck, DT, RC, SKIN, 1.05, 0.98)
    sim.traj_loss(steps, ck, DT, RC, SKIN, 1.0, 1.0)  # warm-up
    t0 = time.perf_counter()
    ub, pb = sim.traj_loss(steps, ck, DT, RC, SKIN, 1.0, 1.0)
    t1 = time.perf_counter()
    r = sim.traj_adjoint(steps, ck, DT, RC, SKIN, 1.0, 1.0, u_t, p_t)
    t2 = time.perf_counter()
    r = sim.traj_adjoint(steps, ck, DT, RC, SKIN, 1.0, 1.0, u_t, p_t)
    t3 = time.perf_counter()
    out = {"atoms": sim.nlocal, "steps": steps, "checkpoint_every": ck, "forward_s": t1 - t0,
           "adjoint_s": t3 - t2, "adjoint_over_forward": (t3 - t2) / (t1 - t0),
           "loss": r["loss"],
           "grad_eps": r["grad_eps"], "grad_sigma": r["grad_sigma"],
           "grad_x0_norm": float(np.linalg.norm(r["grad_x0"]))}
    if fd:
        # central differences (h = 1e-6) of the FP64 loss: the reference
        # gradient for both precisions
        h = 1e-6
        sim.set_precision("fp64")

        def loss(e, s):
            u, p = sim.traj_loss(steps, ck, DT, RC, SKIN, e, s)
            return (u - u_t) ** 2 + (p - p_t) ** 2
        fe = (loss(1 + h, 1.0) - loss(1 - h, 1.0)) / (2 * h)
        fs = (loss(1.0, 1 + h) - loss(1.0, 1 - h)) / (2 * h)
        out["fd_grad_eps"], out["fd_grad_sigma"] = fe, fs
        out["fd_rel_err_eps"] = abs(r["grad_eps"] - fe) / abs(fe)
        out["fd_rel_err_sigma"] = abs(r["grad_sigma"] - fs) / abs(fs)
        if precision != "fp64":
            # the same adjoint kernel on the FP64 forward trajectory: what is
            # left of the error is the adjoint's own rounding
            sim.set_adjoint_precision(precision)
            r2 = sim.traj_adjoint(steps, ck, DT, RC, SKIN, 1.0, 1.0, u_t, p_t, grad_state=False)
            sim.set_adjoint_precision(None)
            out["fp64_forward_fd_rel_err_eps"] = abs(r2["grad_eps"] - fe) / abs(fe)
            out["fp64_forward_fd_rel_err_sigma"] = abs(r2["grad_sigma"] - fs) / abs(fs)
    sim.close()
    # a plain forward trajectory of the same length (no per-step energy and
    # pressure): the step loop alone, from the same initial state (the
    # library holds one simulation at a time, so after the one above closes)
    sim2 = tm.Simulation.fcc(nc, precision=precision, lj_shift=True)
    sim2.setup(False, RC, SKIN)
    sim2.lennard_jones(False, RC)
    sim2.run_lj(steps, ck, DT, False, RC, SKIN)  # warm-up
    sim2.barrier()
    p0 = time.perf_counter()
    sim2.run_lj(steps, ck, DT, False, RC, SKIN)
    sim2.barrier()
    p1 = time.perf_counter()
    sim2.close()
    out.update({"plain_forward_s": p1 - p0, "adjoint_over_plain_forward": out["adjoint_s"] / (p1 - p0),
                "forward_form": "forward_s: the loss trajectory (step loop plus per-step energy and pressure); "
                                "plain_forward_s: the step loop alone"})
    return out


def batched_leg(tm, steps=500, nrep=64, seed=2024, sequential=True, autotune=True, graph=True):
    """64 independent LJ boxes (cells per side uniform 6..24, density
    U[0.70, 0.90], temperature U[0.7, 2.0]; counter-based generator, seed
    2024), FP32-mixed, cutoff 2.5, skin 0.3, dt 0.005, reneighbour 20:
    autotuned batched run (one hipGraph launch) against the same boxes run
    one after another through the single-box path.  autotune=False keeps the
    default configuration; graph=False launches the run's kernels one by one
    (the same kernels and results) for tools that trace single launches."""
    boxes = tm.replica_boxes(seed=seed, nrep=nrep)
    atoms = int(sum(len(b[1]) for b in boxes))
    b = tm.Batch(boxes, cutoff=RC, skin=SKIN, dt=DT, reneigh=REN, precision="fp32mixed")
    t0 = time.perf_counter()
    cfg = b.autotune(trial_steps=2 * REN) if autotune else None
    t_tune = time.perf_counter() - t0
    if not graph:
        b.set_graph(False)
    b.run(2 * REN)  # warms the kernels
    b.reset()
    b.run(steps)    # with graphs: captures the `steps` graph
    b.reset()
    t0 = time.perf_counter()
    b.run(steps)
    t1 = time.perf_counter()
    b.reset()
    gpu_ms = b.run_timed(steps)
    out = {"replicas": nrep, "atoms": atoms, "steps": steps, "config": cfg, "autotuned": autotune,
           "graph": graph, "autotune_s": t_tune,
           "batched_s": t1 - t0, "batched_atom_steps_per_s": atoms * steps / (t1 - t0),
           "batched_graph_ms": gpu_ms}
    b.close()
    if sequential:
        t_seq = 0.0
        for aabb, x, v in boxes:
            sim = tm.Simulation(np.ones(len(x)), x, v, aabb, RC + SKIN, precision="fp32mixed")
            sim.copy_to_accelerator()
            sim.setup(False, RC, SKIN)
            sim.lennard_jones(False, RC)
            sim.barrier()
            s0 = time.perf_counter()
            sim.run_lj(steps, REN, DT, False, RC, SKIN)
            sim.barrier()
            t_seq += time.perf_counter() - s0
            sim.close()
        out["sequential_s"] = t_seq
        out["sequential_atom_steps_per_s"] = atoms * steps / t_seq
        out["batched_over_sequential"] = t_seq / (t1 - t0)
    return out


def dem_leg(tm, steps=1000, nc=64):
    """Granular gas: FCC 64^3 sites (a = 1.6796) kept where i + j < 2 nc
    (528,384 spheres of diameter 1.0), velocities U[-1, 1], spring-dashpot
    contact k = 100, damping 0.5 / 0.2, friction 0.3, contact cutoff 1.0,
    skin 0.3, dt 0.001, half list, FP32-mixed, reneighbour 20."""
    m, x, v = tm.generate_granular_gas(nc)
    L = nc * 1.6796
    sim = tm.Simulation(m, x, v, np.array([0.0, L, 0.0, L, 0.0, L]), 1.3, precision="fp32mixed", dem_radius=0.5)
    sim.copy_to_accelerator()
    sim.setup(True, 1.0, 0.3)
    sim.dem(True, 1.0, 0.5, 0.2, 100.0, 0.3)
    sim.run_dem(REN, REN, 0.001, True, 1.0, 0.3, 0.5, 0.2, 100.0, 0.3)  # warm-up
    sim.barrier()
    t0 = time.perf_counter()
    sim.run_dem(steps, REN, 0.001, True, 1.0, 0.3, 0.5, 0.2, 100.0, 0.3)
    sim.barrier()
    t1 = time.perf_counter()
    # per-kernel timers on a separate, shorter run (they add host work)
    sim.profile(True)
    sim.profile_reset()
    psteps = 5 * REN
    sim.run_dem(psteps, REN, 0.001, True, 1.0, 0.3, 0.5, 0.2, 100.0, 0.3)
    sim.barrier()
    ms, cnt = sim.profile_get("force")
    per = {k: sim.profile_get(k)[0] for k in ("force", "integrate1", "integrate2", "ghost_sync", "mirror", "neighbor", "bin",
                                             "borders")}
    phases = {ph: sum(per[k] for k in ks) / psteps for ph, ks in PHASES.items()}
    sim.profile(False)
    n = sim.nlocal
    nall = n + sim.nghost
    counts, _ = sim.neighbor_lists(with_lists=False)
    entries = int(counts.astype(np.int64).sum())
    # the step pass in half-list form: counts, 4-byte entries, FP64 positions
    # and velocities of locals and ghosts read once, the advanced positions
    # and velocities of the locals and of their ghosts written once
    nbytes = 4 * n + 4 * entries + 48 * nall + 48 * nall
    ke = sim.thermo(1.0)["KE"]
    sim.close()
    t = ms / max(cnt, 1)
    return {"atoms": n, "steps": steps, "seconds": t1 - t0, "atom_steps_per_s": n * steps / (t1 - t0),
            "force_ms_avg": t, "phases_ms_per_step": phases, "kinetic_energy_end": ke, "half_list_entries": entries,
            "roofline": {"bound": "hbm", "bytes_per_launch": nbytes, "achieved": nbytes / (t * 1e-3) / 1e9,
                         "time": "the force phase by HIP events: one pass computing the contact forces, both kicks and the drift",
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": nbytes / (t * 1e-3) / 1e9 / HBM_PEAK_GBS}}


def shard_leg(tm, steps=200, warmup=20, nc=40):
    """One GPU's share of the 80^3 box split 2 x 2 x 2 over eight GPUs: a
    40^3-cell sub-box (256,000 atoms) with its ghost shell.  A periodic
    40^3 box has the same locals, the same ghost-shell population and the
    same per-step work; only the source of the ghosts' positions differs
    (owners on this GPU here, face neighbours over xGMI there).  Reports the
    step rate, the force pass, the ghost refresh (the pack and unpack of the
    halo) and the bytes a face exchange would move per step."""
    sim, r = run_variant(tm, "fp32mixed", False, steps, warmup, profile=True, nc=nc)
    n, ng = sim.nlocal, sim.nghost
    sim.close()
    halo = 24 * ng  # FP64 positions of the ghost shell, per step
    return {"cells": f"{nc}^3", "atoms": n, "ghosts_last_rebuild": ng, "steps": steps,
            "atom_steps_per_s": r["atom_steps_per_s"], "ms_per_step": r["seconds"] / steps * 1e3,
            "force_ms_avg": r["force_ms_avg"], "ghost_refresh_ms_avg": r["ghost_sync_ms_avg"],
            "phases_ms_per_step": r["phases_ms_per_step"], "halo_bytes_per_step": halo,
            "halo_us_one_xgmi_link": halo / XGMI_LINK_GBS / 1e3,
            "halo_us_three_links": halo / 3 / XGMI_LINK_GBS / 1e3,
            "xgmi_link_GBps": XGMI_LINK_GBS}


def copy_calibration(nbytes=1 << 30, reps=5):
    """A plain device-to-device copy of known size (torch), for scaling the
    FETCH_SIZE / WRITE_SIZE counters of the force kernel."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps}), flush=True)


def hbm_copy_gbs_child():
    """hbm_copy_gbs in a child process of its own (torch sets up the device
    there, apart from the library's use of it in this process)."""
    import subprocess
    try:
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-bw"], capture_output=True, text=True,
                           timeout=300)
        if r.returncode != 0:
            print(f"# copy bandwidth leg failed: {r.stderr[-400:]}", file=sys.stderr, flush=True)
            return None
        return json.loads(r.stdout.strip().splitlines()[-1])["copy_GBps"]
    except (subprocess.TimeoutExpired, ValueError, KeyError, IndexError) as e:
        print(f"# copy bandwidth leg failed: {e}", file=sys.stderr, flush=True)
        return None


def hbm_copy_gbs(nbytes=1 << 30, reps=5):
    """Achievable HBM bandwidth on this device: a plain 1 GiB device-to-device
    copy (torch), timed with events; bytes read + written per second."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2.0 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b
    torch.cuda.empty_cache()
    return gbs


def load_profile_figures():
    """Per-launch figures from the committed rocprofv3 summaries: the LJ
    force kernel's HBM bytes (FETCH_SIZE + WRITE_SIZE, copy-calibrated), the
    FP64 and FP32 VALU flops of its counted instructions and its average
    duration in the kernel trace, and the DEM step pass's HBM bytes."""
    import re
    out = {}
    for name, keys in (("force_lj_fp32mixed_80cube_traffic.md", ("traffic_bytes_per_launch",)),
                       ("force_lj_80cube_valu_flops.md", ("valu_f64_flops_per_launch", "valu_f32_flops_per_launch")),
                       ("lj_80cube_fp32mixed_full_kernel_stats.md", ("trace_avg_us",)),
                       ("dem_528k_step_pass_counters.md", ("dem_traffic_bytes_per_launch",))):
        p = os.path.join(ROOT, "profiles", name)
        if not os.path.exists(p):
            continue
        text = open(p).read()
        for k in keys:
            m = re.search(r"^" + k + r":\s*([0-9.eE+]+)", text, re.M)
            if m:
                out[k] = float(m.group(1))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--quick", action="store_true", help="headline variant only")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-trajectory leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-replica leg")
    ap.add_argument("--no-dem", action="store_true", help="skip the DEM granular-gas leg")
    ap.add_argument("--batch-profile", action="store_true",
                    help="batched leg with the default configuration, kernels launched one by one, 100 steps")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration) and exit")
    ap.add_argument("--copy-bw", action="store_true", help="only measure the device copy bandwidth and exit")
    ap.add_argument("--no-copy-bw", action="store_true", help="skip the device copy bandwidth leg")
    ap.add_argument("--no-shard", action="store_true", help="skip the one-GPU-share (40^3 sub-box) leg")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.copy_bw:
        print(json.dumps({"copy_GBps": hbm_copy_gbs()}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this library runs on one GPU; splitting a box over several GPUs "
              "is not implemented. Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import tinymd_amd as tm
    dev, cus = tm.device_info()
    if cus == 0:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    print(f"# device: {dev}, {cus} CUs", file=sys.stderr, flush=True)

    variants = {}
    sim32, r32 = run_variant(tm, "fp32mixed", False, args.steps, args.warmup, profile=True)
    variants["fp32mixed_full"] = r32
    print(f"# fp32mixed full: {r32}", file=sys.stderr, flush=True)
    model = force_kernel_model(sim32)
    errors = {"fp32mixed": reduced_force_error(tm, sim32, "fp32mixed"),
              "fp32kahan": reduced_force_error(tm, sim32, "fp32kahan")}
    print(f"# force model {model}; errors vs FP64 {errors}", file=sys.stderr, flush=True)
    n = sim32.nlocal
    sim32.close()
    del sim32
    if not args.quick:
        for prec, half in (("fp64", False), ("fp32kahan", False), ("fp32mixed", True), ("fp64", True)):
            sim, r = run_variant(tm, prec, half, args.steps, args.warmup, profile=True)
            variants[f"{prec}_{'half' if half else 'full'}"] = r
            print(f"# {prec} half={half}: {r}", file=sys.stderr, flush=True)
            sim.close()
            del sim

    adjoint = {}
    if not args.no_adjoint:
        for prec in ("fp64", "fp32mixed"):
            adjoint[prec] = adjoint_leg(tm, prec, fd=not args.quick)
            print(f"# adjoint {prec}: {adjoint[prec]}", file=sys.stderr, flush=True)

    c1 = None
    if not args.quick:
        c1 = c1_leg(tm, cpu=not args.no_cpu)
        print(f"# c1 32^3 fp64: {c1}", file=sys.stderr, flush=True)
    batched = None
    if not args.no_batch:
        if args.batch_profile:
            batched = batched_leg(tm, steps=100, sequential=False, autotune=False, graph=False)
        else:
            batched = batched_leg(tm, sequential=not args.quick)
        print(f"# batched: {batched}", file=sys.stderr, flush=True)
    dem = None
    if not args.no_dem:
        dem = dem_leg(tm, steps=200 if args.quick else 1000)
        print(f"# dem: {dem}", file=sys.stderr, flush=True)

    shard = None
    if not args.no_shard:
        shard = shard_leg(tm)
        print(f"# shard 40^3: {shard}", file=sys.stderr, flush=True)

    value = r32["atom_steps_per_s"]
    t_force = r32["force_ms_avg"] * 1e-3
    achieved = model["bytes"] / t_force / 1e9
    prof = load_profile_figures()
    copy_gbs = None if args.no_copy_bw else hbm_copy_gbs_child()
    valu = {"alg_flops_per_launch": model["flops"], "alg_tflops": model["flops"] / t_force / 1e12,
            "frac_fp64_peak": model["flops"] / t_force / 1e12 / FP64_PEAK_TFLOPS,
            "frac_fp32_peak": model["flops"] / t_force / 1e12 / FP32_PEAK_TFLOPS,
            "peaks_tflops": {"fp64": FP64_PEAK_TFLOPS, "fp32": FP32_PEAK_TFLOPS}}
    if "valu_f64_flops_per_launch" in prof:
        f64 = prof["valu_f64_flops_per_launch"] / t_force / 1e12
        f32 = prof["valu_f32_flops_per_launch"] / t_force / 1e12
        valu.update({"counted_fp64_tflops": f64, "counted_fp32_tflops": f32,
                     "counted_frac_fp64_peak": f64 / FP64_PEAK_TFLOPS, "counted_frac_fp32_peak": f32 / FP32_PEAK_TFLOPS,
                     "counted_source": "profiles/force_lj_80cube_valu_flops.md (SQ_INSTS_VALU_* x 64 lanes, FMA x 2)"})
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["seconds"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair forces, f64 accumulation and state",
        "data": "synthetic: FCC lattice + Park-Miller velocities (initialize.h), T=1.44",
        "config": {"workload": "LJ FCC 80x80x80 cells, 2,048,000 atoms, rho 0.8442, rc 2.5, skin 0.3, "
                               "dt 0.005, reneighbour 20, full list, FP32-mixed",
                   "atoms": n, "list_entries": model["list_entries"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": prof.get("traffic_bytes_per_launch"),
                     "kernel": "k_lj_full<FP32-mixed, kick and drift fused> (step-loop force pass)",
                     "bytes_per_launch": model["bytes"],
                     "bytes_model": "4 N + 4 entries + 24 N_all + 24 N (list counts, 4-byte entries, FP64 "
                                    "positions once, FP64 force out)",
                     "avg_launch_ms": r32["force_ms_avg"], "avg_launch_source": "HIP events on the library stream",
                     "trace_avg_ms": prof["trace_avg_us"] * 1e-3 if "trace_avg_us" in prof else None,
                     "frac_at_trace_avg": (model["bytes"] / (prof["trace_avg_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS
                                           if "trace_avg_us" in prof else None),
                     "frac_2B_entries": model["bytes_2B_entries"] / t_force / 1e9 / HBM_PEAK_GBS,
                     "frac_fused_step_model": model["bytes_fused_step_model"] / t_force / 1e9 / HBM_PEAK_GBS,
                     "copy_measured": copy_gbs, "frac_of_copy": achieved / copy_gbs if copy_gbs else None,
                     "valu": valu},
        "force_model": model,
        "reduced_precision_force_err": errors,
        "pair_interactions_per_s": value * model["in_cutoff_entries"] / 2.0 / n,
        "list_entries_per_s": value * model["list_entries"] / n,
        "variants": variants,
    }
    if adjoint:
        line["adjoint"] = adjoint
    if c1:
        line["c1_32cube_fp64"] = c1
    if batched:
        line["batched"] = batched
    if dem:
        # measured HBM bytes of the DEM step pass (copy-calibrated counters)
        dem["roofline"]["traffic"] = prof.get("dem_traffic_bytes_per_launch")
        line["dem"] = dem
    if shard:
        line["shard_40cube_of_80cube_2x2x2"] = shard
    if "fp64_full" in variants:
        line["fp32mixed_over_fp64"] = value / variants["fp64_full"]["atom_steps_per_s"]
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
