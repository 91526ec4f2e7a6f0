#!/usr/bin/env python3
"""Throughput of the fused stochastic-LLG step on one MI355X, with its HBM roofline and a CPU baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload bcc64|tri256|sc16] [--precision fp32|fp64]

A step is one stochastic Heun step (noise + field + predictor, then field + corrector) over the whole
synthetic lattice, inputs resident in HBM.  Default workload: bcc 64^3 conventional cells (524 288
spins), two shells with per-bond random 3x3 tensors, random per-site SIA, T = 300 K, FP32 -- the
HBM-roofline case.  Prints one JSON line.

roofline.achieved = algorithmic bytes per step / measured step time, with the algorithmic bytes of a
Heun step: 6w (read e, write e) + 6w if the SIA is per site + 9w nnb/2 if bond tensors are per bond
(each unordered bond stored once), w = 4 (FP32) or 8 (FP64) bytes.  The step time comes from HIP
events recorded on the stream the kernels run on, around one mx_llg_run call of K steps (the solver's
set-up calls made before it).  Before the W warm-up steps, untimed steps run for SETTLE_S seconds: the
GPU's clocks reach their sustained level only after some milliseconds of continuous work, and a small
W (5 steps of bcc64 are 0.3 ms) would otherwise time part of that ramp (measured: 20 timed steps at
61 us per step after 5 warm-up steps, 58 after 200; 56 in long runs).
cpu_baseline: the C++/OpenMP restatement of magex's fused CPU Heun (oracle/cpu_llg.cpp), FP64, on a
bounded sample of the same workload, on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the CPU baseline runs one OpenMP thread per core, pinned; set before any OpenMP runtime starts
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")
# the CPUs this process may run on, read before OpenMP pins the main thread to one of them
CPUS_USABLE = len(os.sched_getaffinity(0))

METRIC = "spin-steps/sec + % HBM roofline (FP32 vs FP64); batched speedup; grad cost/forward"
HBM_PEAK_GBS = 8000.0
SETTLE_S = 0.05   # untimed work before the warm-up steps (clock settling)


def alg_bytes(system, precision):
    """Compulsory HBM bytes of one Heun step per spin."""
    w = 4 if precision == "fp32" else 8
    b = 6 * w
    A = system.A.reshape(system.ncell, system.uc.nsites, 9)
    if not (A == A[:1]).all():
        b += 6 * w
    if system.bond_tensors is not None:
        nnb = len(system.stencil) / system.uc.nsites
        b += 9 * w * nnb / 2
    return b


def workload(name):
    from magex_amd import configs
    if name == "bcc64":
        s, p = configs.bcc64()
        desc = ("bcc 64^3 conventional (524288 spins), J1=20 J2=10 meV + per-bond random sym/antisym tensors "
                "(sigma 1%), random SIA, mu=2.2, alpha=0.1, T=300 K, dt=0.1 fs")
    elif name == "tri256":
        s, p = configs.tri256()
        desc = "triangular 256^2 (65536 spins), J=(5,-0.5,0.2) meV, DMI 1.5, K=0.3, B=2 T, alpha=0.05, T=20 K, dt=0.5 fs"
    elif name == "sc16":
        s, p = configs.sc16()
        desc = "sc 16^3 (4096 spins), J=1, K=0.1 meV, B=0.5 T, alpha=0.1, T=0, dt=0.1 fs"
    else:
        raise SystemExit(f"unknown workload {name}")
    return s, p, desc


def gpu_rate(system, p, precision, steps, warmup):
    import torch
    from magex_amd.llg import LLGSolver
    sol = LLGSolver(system, p["alpha"], p["dt"], p["temperature"], p["seed"], precision=precision)
    e = torch.as_tensor(system.init_spins("random", seed=0), device="cuda",
                        dtype=torch.float32 if precision == "fp32" else torch.float64).contiguous()
    dm = sol._sync()                     # the solver's set-up calls (dynamics, field, noise grade)
    seed, step, n = sol.noise_seed, 0, 10
    t_settle = time.perf_counter()
    while time.perf_counter() - t_settle < SETTLE_S:
        t = time.perf_counter()
        dm.run(e, n, seed, step)
        step += n
        torch.cuda.synchronize()
        if time.perf_counter() - t < 2e-3:   # chunks of at least 2 ms of work
            n *= 2
    dm.run(e, warmup, seed, step)
    step += warmup
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    dm.run(e, steps, seed, step)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = start.elapsed_time(stop)
    assert torch.isfinite(e).all(), "non-finite spins"
    assert dm.step_info()[1] == 0, "a wait of the step kernels gave up"
    return dict(wall_s=wall, event_ms=ev_ms, step_ms=ev_ms / steps)


def _angle(a, b):
    import torch
    b = b / b.norm(dim=1, keepdim=True)
    t = torch.atan2(torch.linalg.cross(a, b).norm(dim=1), (a * b).sum(1))
    return {"mean_rad": float(t.mean()), "max_rad": float(t.max())}


def fp32_angle_error(system, p, steps=200, start="random", control=False):
    """Mean and max per-spin angle (rad) between the FP32 and FP64 trajectories of the workload after
    `steps` steps from the same start on the same thermal noise (the FP64 kernels match the oracle to
    1e-10).  control: also the angle between FP64 and FP64 started 1e-7 rad away (about FP32's rounding
    of the start), which separates the dynamics' own amplification of small differences from FP32 error."""
    import numpy as np
    import torch
    from magex_amd.llg import LLGSolver
    e0 = system.init_spins("random", seed=0) if start == "random" else system.init_spins("ferro")
    starts = {"fp64": e0, "fp32": e0}
    if control:
        d = np.random.default_rng(7).normal(size=e0.shape)
        d -= (d * e0).sum(1, keepdims=True) * e0
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        ep = e0 + 1e-7 * d
        starts["fp64_perturbed"] = ep / np.linalg.norm(ep, axis=1, keepdims=True)
    out = {}
    for key, x0 in starts.items():
        prec = "fp32" if key == "fp32" else "fp64"
        # the FP64 runs draw the FP32 run's (float-grade) normals: identical noise
        sol = LLGSolver(system, p["alpha"], p["dt"], p["temperature"], p["seed"], precision=prec, noise="float")
        e = torch.as_tensor(x0, device="cuda", dtype=torch.float32 if prec == "fp32" else torch.float64).contiguous()
        sol.step_device(e, steps)
        out[key] = e.double()
    r = {"steps": steps, "start": start, **_angle(out["fp64"], out["fp32"]),
         "note": "FP32 vs FP64 device trajectories, same start and noise"}
    if control:
        r["fp64_started_1e-7_rad_away"] = _angle(out["fp64"], out["fp64_perturbed"])
    return r


def fp32_error_curve(system, p, checkpoints=(50, 200, 1000, 2000, 5000), start="random"):
    """Mean per-spin angle (rad) of FP32 from FP64 along one trajectory, beside the angle of an FP64 run
    started 1e-7 rad away (about FP32's rounding of the start) from FP64: the three runs advance in lock
    step on identical noise (the FP32 run's float-grade normals) and are compared at each checkpoint.
    The control shows how fast the dynamics themselves amplify a difference of FP32's size; the ratio
    FP32 / control is the part FP32 adds beyond one rounding of the start."""
    import numpy as np
    import torch
    from magex_amd.llg import LLGSolver
    e0 = system.init_spins("random", seed=0) if start == "random" else system.init_spins("ferro")
    d = np.random.default_rng(7).normal(size=e0.shape)
    d -= (d * e0).sum(1, keepdims=True) * e0
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    x1 = e0 + 1e-7 * d
    runs = {}
    for key, prec, x0 in (("fp64", "fp64", e0), ("fp32", "fp32", e0),
                          ("control", "fp64", x1 / np.linalg.norm(x1, axis=1, keepdims=True))):
        sol = LLGSolver(system, p["alpha"], p["dt"], p["temperature"], p["seed"], precision=prec, noise="float")
        runs[key] = (sol, torch.as_tensor(x0, device="cuda",
                                          dtype=torch.float32 if prec == "fp32" else torch.float64).contiguous())
    curve, done = [], 0
    for n in checkpoints:
        for sol, e in runs.values():
            sol.step_device(e, n - done)
        done = n
        ref = runs["fp64"][1]
        a32 = _angle(ref, runs["fp32"][1].double())["mean_rad"]
        actl = _angle(ref, runs["control"][1])["mean_rad"]
        curve.append({"steps": n, "fp32_mean_rad": a32, "control_mean_rad": actl, "ratio": a32 / max(actl, 1e-300)})
    return {"start": start, "curve": curve,
            "note": "mean angle to FP64 on identical noise: FP32, and FP64 started 1e-7 rad away (control)"}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_topology():
    """Model, sockets, physical cores and threads per core of this host (lscpu), plus the CPUs this
    process may run on (read at start-up)."""
    import subprocess
    info = {"cpus_usable": CPUS_USABLE}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
    for line in txt.splitlines():
        k, _, v = line.partition(":")
        if k.strip() in keys:
            v = v.strip()
            info[keys[k.strip()]] = int(v) if v.isdigit() else v
    if "sockets" in info and "cores_per_socket" in info:
        info["physical_cores"] = info["sockets"] * info["cores_per_socket"]
    return info


def _best_of_3(h, e, n):
    best = float("inf")
    for _ in range(3):
        t = time.perf_counter()
        h.run(e, n)
        best = min(best, (time.perf_counter() - t) / n)
    return best


def cpu_calibration():
    """The CPU restatement on magex's published set-up (benchmarks/speed_llg.py, sc.json): sc 32^3,
    J = 1 meV, mu = 2, alpha = 1, dt = 1e-4 ps, T = 10 K, ferro start; best of 3 x 50 steps after one
    warm-up step, 1 and 2 threads.  Published: 55-64 and 29 ns/spin/step (docs/benchmarks.md:11-13,
    which include magex's per-step Python overhead)."""
    import numpy as np
    from oracle import cpu
    from oracle import magex_ref as R
    o = R.SpinSystem(R.load_results(R.make_model("sc", a=3.0, J=[1.0], moment=2.0)), (32, 32, 32))
    out = {"published_ns_per_spin_step": {"1": "55-64", "2": "29"}}
    for th in (1, 2):
        cpu.set_threads(th)
        h = cpu.CpuHeun(o, 1.0, 1e-4, 10.0, seed=1)
        e = np.tile([0.0, 0.0, 1.0], (o.N, 1))
        h.run(e, 1)
        out[f"{th}_thread_ns_per_spin_step"] = _best_of_3(h, e, 50) / o.N * 1e9
    cpu.set_threads(cpu.max_threads_available())
    return out


def cpu_rate(system, p, seconds, threads=None, per_bond=False):
    """spin-steps/s of the CPU restatement (FP64) on a bounded sample of the workload."""
    import numpy as np
    from oracle import cpu
    from oracle import magex_ref as R
    if threads:
        cpu.set_threads(threads)
    if per_bond:
        src, dst, J = [], [], []
        for k in range(len(system.stencil)):
            a, b, blk = system._entry_blocks(k)
            ok = b >= 0
            src.append(a[ok])
            dst.append(b[ok])
            J.append(np.asarray(blk)[ok])
        blocks = cpu.per_bond_blocks(system.N, np.concatenate(src), np.concatenate(dst), np.concatenate(J))
    else:
        blocks = R.neighbour_blocks(system.Jsp, system.N)
    h = cpu.CpuHeun(system, p["alpha"], p["dt"], p["temperature"], seed=1, blocks=blocks)
    e = system.init_spins("random", seed=0)
    h.run(e, 1)
    t = time.perf_counter()
    h.run(e, 1)
    one = time.perf_counter() - t
    n = max(1, int(seconds / 3 / max(one, 1e-9)))
    best = _best_of_3(h, e, n)
    th = cpu.max_threads()
    cpu.set_threads(cpu.max_threads_available())
    return system.N / best, n, th


def config_rates():
    """The other benchmark configurations, GPU beside CPU: sc 16^3 (FP64, T = 0) and tri 256^2 (FP32 and
    FP64 on identical noise, with the topological charge after 2000 steps)."""
    import torch
    from magex_amd import configs, texture
    from magex_amd.llg import LLGSolver
    out = {}
    s, p = configs.sc16()
    r = gpu_rate(s, p, "fp64", 2000, 200)
    c, n, th = cpu_rate(s, p, 2.0)
    out["sc16_fp64"] = {"gpu_spin_steps_per_s": s.N / (r["step_ms"] * 1e-3), "gpu_us_per_step": 1e3 * r["step_ms"],
                        "cpu_spin_steps_per_s": c, "cpu_threads": th, "alg_bytes_per_spin_step": alg_bytes(s, "fp64")}
    s, p = configs.tri256()
    c, n, th = cpu_rate(s, p, 3.0)
    q = {}
    for prec in ("fp32", "fp64"):
        r = gpu_rate(s, p, prec, 2000, 200)
        sol = LLGSolver(s, p["alpha"], p["dt"], p["temperature"], p["seed"], precision=prec)
        e = torch.as_tensor(s.init_spins("random", seed=0), device="cuda",
                            dtype=torch.float32 if prec == "fp32" else torch.float64)
        sol.step_device(e, 2000)
        q[prec] = float(texture.topological_charge(s, e))
        out[f"tri256_{prec}"] = {"gpu_spin_steps_per_s": s.N / (r["step_ms"] * 1e-3),
                                 "gpu_us_per_step": 1e3 * r["step_ms"], "Q_after_2000_steps": q[prec],
                                 "roofline_frac": alg_bytes(s, prec) * s.N / (r["step_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS}
    out["tri256_cpu_spin_steps_per_s"] = c
    out["tri256_cpu_threads"] = th
    # random start at T = 20 K forms a skyrmion texture: the dynamics amplify small differences, so the
    # angle error is given at two lengths, the longer with the FP64 control started 1e-7 rad away (the
    # charge Q is the same in both precisions)
    out["tri256_fp32_vs_fp64"] = fp32_error_curve(s, p)
    return out


def cpu_baseline(system, p, seconds):
    """magex's fused Heun restated in C++/OpenMP, FP64, timed on a bounded sample of the workload."""
    import numpy as np
    from oracle import cpu
    src, dst, J = [], [], []
    for k in range(len(system.stencil)):
        a, b, blk = system._entry_blocks(k)
        ok = b >= 0
        src.append(a[ok])
        dst.append(b[ok])
        J.append(np.asarray(blk)[ok])
    blocks = cpu.per_bond_blocks(system.N, np.concatenate(src), np.concatenate(dst), np.concatenate(J))
    h = cpu.CpuHeun(system, p["alpha"], p["dt"], p["temperature"], seed=1, blocks=blocks)
    e = system.init_spins("random", seed=0)
    h.run(e, 1)
    t = time.perf_counter()
    h.run(e, 1)
    one = time.perf_counter() - t
    n = max(1, int(seconds / 3 / max(one, 1e-9)))
    best = float("inf")
    for _ in range(3):
        t = time.perf_counter()
        h.run(e, n)
        best = min(best, (time.perf_counter() - t) / n)
    threads = cpu.max_threads()
    one, n1, _ = cpu_rate(system, p, min(3.0, seconds / 4), threads=1, per_bond=True)
    return dict(value=system.N / best, unit="spin-steps/s", cores=threads, kind="port",
                sample=f"best of 3 x {n} Heun steps of the full {system.N}-spin lattice, FP64, {threads} OpenMP "
                       f"threads (magex's fused two-pass Heun, per-bond 3x3 blocks, ziggurat noise)",
                cpu_model=cpu_model(), topology=cpu_topology(), ns_per_spin_step=1e9 * best / system.N,
                one_thread={"spin_steps_per_s": one, "ns_per_spin_step": 1e9 / one, "steps": n1},
                omp={k: os.environ.get(k) for k in ("OMP_PROC_BIND", "OMP_PLACES", "OMP_NUM_THREADS")},
                calibration=cpu_calibration())


def batch_rates(steps, per_replica_steps):
    """The batched form on 96 triangular replicas with L uniform in {32, 48, ..., 192} (seeded), NN J=1,
    K=0.1 meV, B=0.5 T, a geometric 5-300 K ladder, FP32: one launch pair per step for all replicas
    (autotuned shape) against the same replicas run as separate plain simulations."""
    import numpy as np
    import torch
    from magex_amd import configs
    from magex_amd.batch import ReplicaBatch, per_replica_rate
    systems, p = configs.batch96()
    T, seeds = p["temperatures"], p["seeds"]
    alpha = np.full(len(systems), p["alpha"])
    spins = [s.init_spins("random", seed=seeds[k]) for k, s in enumerate(systems)]
    batch = ReplicaBatch(systems, alpha, p["dt"], T, seeds, "fp32")
    e = batch.pack(spins)
    shape, _, _ = batch.autotune(e, steps=8)
    batch.run(e, 5)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    t0.record()
    batch.run(e, steps)
    t1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    total = int(batch.offsets[-1])
    rate = total * steps / wall
    base, _ = per_replica_rate(systems, alpha, p["dt"], T, seeds, spins, per_replica_steps)
    m = batch.magnetization(e)
    # per-replica <|M|> of FP32 against FP64 (the FP64 kernels match the oracle to 1e-10): |M| sampled
    # every 50 steps over 2000 steps from the same start on the same noise, averaged per replica
    mags = {}
    nsamp = 40
    for prec in ("fp32", "fp64"):
        bp = ReplicaBatch(systems, alpha, p["dt"], T, seeds, prec, noise="float")
        ep = bp.pack(spins)
        acc = np.zeros(len(systems))
        for _ in range(nsamp):
            bp.run(ep, 50)
            acc += bp.magnetization(ep)
        mags[prec] = acc / nsamp
        del bp, ep
    dm = np.abs(mags["fp32"] - mags["fp64"])
    return {"workload": "96 triangular replicas, L in {32..192} (seeded), J=1 K=0.1 meV B=0.5 T, T 5-300 K, "
                        "alpha=0.1, dt=0.1 fs, FP32", "spins": total, "steps": steps,
            "spin_steps_per_s": rate, "ms_per_step": 1e3 * wall / steps,
            "event_ms_per_step": t0.elapsed_time(t1) / steps, "launch": {"block": shape[0],
            "sites_per_thread": shape[1], "graph_steps": shape[2]},
            "per_replica_spin_steps_per_s": base, "per_replica_steps": per_replica_steps,
            "speedup_vs_per_replica": rate / base, "mean_abs_m_range": [float(m.min()), float(m.max())],
            "mean_abs_m_fp32_vs_fp64": {"steps": 50 * nsamp, "samples": nsamp,
                                        "max_abs_diff": float(dm.max()), "mean_abs_diff": float(dm.mean()),
                                        "fp64_range": [float(mags["fp64"].min()), float(mags["fp64"].max())]}}


def adjoint_rates(steps=500, checkpoint_every=50, reps=3):
    """The differentiable form on square 128^2 (16 384 spins): J1, J2, interfacial D, K as parameters at
    (4, -0.4, 1.0, 0.2) meV, 500 damped Heun steps (alpha = 0.3, dt = 1 fs, T = 0) from a seeded random
    start, MSE to the texture of (5, -0.5, 1.5, 0.3) meV, checkpoints every 50 steps, FP32.  forward =
    the recorded forward (mx_adjoint_forward), backward = segment replays + adjoint sweep
    (mx_adjoint_backward; each replay runs beside the previous segment's sweep on a second stream, and
    the backward's events on the caller's stream enclose both); both timed with HIP events on the stream
    they run on, best of reps."""
    import numpy as np
    import torch
    from magex_amd import configs
    from magex_amd.adjoint import DifferentiableLLG
    theta0, theta_t = np.array([4.0, -0.4, 1.0, 0.2]), np.array([5.0, -0.5, 1.5, 0.3])
    sol = DifferentiableLLG(lambda th: configs.square128(*th)[0], theta0, alpha=0.3, dt=1e-3, precision="fp32",
                            checkpoint_every=checkpoint_every)
    e0 = sol.system.init_spins("random", seed=0)
    tgt = sol.forward(theta_t, e0, steps)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    fwd, bwd = [], []
    for r in range(reps + 1):
        torch.cuda.synchronize()
        ev[0].record()
        e = sol.forward(theta0, e0, steps, checked=False)
        ev[1].record()
        g = 2.0 * (e - tgt) / e.numel()
        torch.cuda.synchronize()
        ev[1].synchronize()
        t_f = ev[0].elapsed_time(ev[1])
        ev[1].record()
        gt, _ = sol.backward(g, checked=False)
        ev[2].record()
        torch.cuda.synchronize()
        if r:                                   # the first round warms up
            fwd.append(t_f)
            bwd.append(ev[1].elapsed_time(ev[2]))
    assert sol.dm.step_info()[1] == 0, "a wait of the adjoint kernels gave up"
    f, b = min(fwd), min(bwd)
    return {"workload": "square 128^2 (16384 spins), J1 J2 D K learnable, 500 steps alpha=0.3 dt=1 fs T=0, "
                        "checkpoints every 50, MSE loss, FP32",
            "forward_ms": f, "backward_ms": b, "grad_over_forward": b / f, "forward_plus_grad_over_forward": (f + b) / f,
            "grad_theta": [float(x) for x in gt.cpu().numpy()]}


def adjoint_bcc64_rates(steps=100, checkpoint_every=50, reps=2):
    """The differentiable form on the headline model: bcc 64^3 (524 288 spins) with its random per-bond
    tensors and per-site anisotropy held fixed, J1 and J2 learnable at (20, 10) meV, 100 Heun steps of
    its dynamics (alpha = 0.1, dt = 0.1 fs, T = 300 K, one thermal sample path), checkpoints every 50,
    MSE to the texture reached at (21, 9.5) meV, FP32.  Times the recorded forward, the backward, and the
    backward with dense gradients (dL/dT of all 3.7 M bond tensors and dL/dA of every site), HIP events
    on the stream they run on, best of reps."""
    import numpy as np
    import torch
    from magex_amd import configs
    from magex_amd.adjoint import DifferentiableLLG
    from magex_amd.models import make_model
    from magex_amd.spinmodel import SpinSystem, load_results
    s0, p = configs.bcc64()
    fw = s0.forward_entries()
    Js0 = np.array([J for *_, J in s0.stencil])
    D = s0.bond_tensors - Js0[fw][:, None]
    A = np.asarray(s0.A)

    def build(th):
        s = SpinSystem(load_results(make_model("bcc_conventional", a=2.87, J=[th[0], th[1]], moment=2.2)), s0.L)
        s.set_bond_tensors(np.array([J for *_, J in s.stencil])[fw][:, None] + D)
        s.A = A
        return s

    theta0, theta_t = np.array([20.0, 10.0]), np.array([21.0, 9.5])
    sol = DifferentiableLLG(build, theta0, alpha=p["alpha"], dt=p["dt"], temperature=p["temperature"],
                            seed=p["seed"], precision="fp32", checkpoint_every=checkpoint_every)
    e0 = s0.init_spins("random", seed=0)
    tgt = sol.forward(theta_t, e0, steps)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    fwd, bwd, dns = [], [], []
    for r in range(reps + 1):
        torch.cuda.synchronize()
        ev[0].record()
        e = sol.forward(theta0, e0, steps, checked=False)
        ev[1].record()
        g = 2.0 * (e - tgt) / e.numel()
        torch.cuda.synchronize()
        t_f = ev[0].elapsed_time(ev[1])
        ev[1].record()
        gt, _ = sol.backward(g, want_e0=False, checked=False)
        ev[2].record()
        gt2, _, gb, ga = sol.backward_dense(g, want_e0=False, checked=False)
        ev[3].record()
        torch.cuda.synchronize()
        if r:
            fwd.append(t_f)
            bwd.append(ev[1].elapsed_time(ev[2]))
            dns.append(ev[2].elapsed_time(ev[3]))
    assert sol.dm.step_info()[1] == 0, "a wait of the adjoint kernels gave up"
    f, b, bd = min(fwd), min(bwd), min(dns)
    return {"workload": "bcc 64^3 (524288 spins), per-bond tensors and per-site SIA fixed, J1 J2 learnable, "
                        "100 steps alpha=0.1 dt=0.1 fs T=300 K, checkpoints every 50, MSE loss, FP32",
            "forward_ms": f, "backward_ms": b, "backward_dense_ms": bd, "grad_over_forward": b / f,
            "dense_grad_over_forward": bd / f, "grad_theta": [float(x) for x in gt.cpu().numpy()],
            "dense_equals_plain_theta_grad": bool(torch.equal(gt, gt2))}


def mc_rates(system, p, sweeps=50):
    """Metropolis colour sweeps (gaussian moves, one over-relaxation pass) on the bcc64 workload, FP32:
    site updates per second counting both passes, and the HBM rate by the same algorithmic bytes per
    site visit as a Heun pass (a visit gathers every coupling of the site once)."""
    import torch
    from magex_amd.montecarlo import MetropolisMC
    mc = MetropolisMC(system, p["temperature"], seed=1, sigma=0.3, overrelax=1, move="gaussian", precision="fp32")
    e = torch.as_tensor(system.init_spins("random", seed=0), device="cuda", dtype=torch.float32)
    mc.sweep_device(e, 3)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    mc.sweep_device(e, sweeps)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / sweeps
    visits = 2 * system.N
    B = alg_bytes(system, "fp32") / 2
    return {"workload": "bcc64 Metropolis (gaussian sigma 0.3, 1 over-relaxation pass), T = 300 K, FP32",
            "colours": mc.ncolors, "ms_per_sweep": ms, "site_visits_per_s": visits / (ms * 1e-3),
            "acceptance": mc.acceptance(), "alg_bytes_per_visit": B,
            "roofline_frac": B * visits / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}


def traffic_from_profiles(workload_name, precision):
    """HBM bytes per step measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE), if committed."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload_name}_{precision}.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_step:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="bcc64")
    ap.add_argument("--precision", default="fp32", choices=("fp32", "fp64"))
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=300)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the sc16 / tri256 GPU-beside-CPU table")
    ap.add_argument("--no-mc", action="store_true", help="skip the Monte Carlo sweep leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a plain 256 MiB device copy first (the reference kernel for rocprof FETCH/WRITE_SIZE)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {a.gpus} is not supported (the multi-GPU sharding of this "
              "path is a design, not a measurement)", file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        import torch
        x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
    system, p, desc = workload(a.workload)
    print(f"workload {a.workload}: {system.N} spins; precision {a.precision}", file=sys.stderr, flush=True)
    r = gpu_rate(system, p, a.precision, a.steps, a.warmup)
    other = "fp64" if a.precision == "fp32" else "fp32"
    r2 = None if a.no_fp64 else gpu_rate(system, p, other, max(10, a.steps // 5), max(2, a.warmup // 5))
    N = system.N
    value = N * a.steps / r["wall_s"]
    B = alg_bytes(system, a.precision)
    achieved = B * N / (r["step_ms"] * 1e-3) / 1e9
    traffic = traffic_from_profiles(a.workload, a.precision)
    out = {
        "metric": METRIC, "value": value, "unit": "spin-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": r["wall_s"] * 1e3 / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if a.precision == "fp32" else "f64",
        "data": "synthetic (seeded random couplings and spins)",
        "config": {"workload": f"{a.workload}: {desc}", "spins": N, "precision": a.precision,
                   "alg_bytes_per_spin_step": B},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "Heun step = k_heun pass 1 + pass 2 (event time per step)"},
    }
    if a.precision == "fp32" and not a.no_fp64:
        # at the workload's full 5000 steps: from a ferromagnetic start (thermal dynamics around order)
        # and from the random start, a quench whose dynamics amplify any small difference; the FP64
        # control started 1e-7 rad away shows how much of that is the dynamics rather than FP32
        n = p.get("steps", 5000)
        out["fp32_angle_error"] = {"ferro_start": fp32_angle_error(system, p, steps=n, start="ferro"),
                                   "random_start_curve": fp32_error_curve(system, p, start="random")}
    if r2 is not None:
        out[other] = {"spin_steps_per_s": N / (r2["step_ms"] * 1e-3), "ms_per_step": r2["step_ms"],
                      "roofline_frac": alg_bytes(system, other) * N / (r2["step_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS}
    if not a.no_batch:
        print("batched replicas ...", file=sys.stderr, flush=True)
        out["batched"] = batch_rates(a.batch_steps, max(20, a.batch_steps // 10))
    if not a.no_adjoint:
        print("adjoint ...", file=sys.stderr, flush=True)
        out["adjoint"] = adjoint_rates()
        print("adjoint on bcc64 ...", file=sys.stderr, flush=True)
        out["adjoint"]["bcc64_perbond"] = adjoint_bcc64_rates()
    if not a.no_mc:
        print("monte carlo ...", file=sys.stderr, flush=True)
        out["monte_carlo"] = mc_rates(system, p) if a.workload == "bcc64" else None
    if not a.no_configs:
        print("other configurations ...", file=sys.stderr, flush=True)
        out["configs"] = config_rates()
    if not a.no_cpu_baseline:
        print("cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(system, p, a.cpu_seconds)
        out["vs_cpu"] = value / out["cpu_baseline"]["value"]
        if r2 is not None and other == "fp64":
            # like for like: the FP64 kernels against the FP64 CPU restatement
            out["vs_cpu_fp64"] = out["fp64"]["spin_steps_per_s"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
