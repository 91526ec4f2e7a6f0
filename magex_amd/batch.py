"""Batched replicas: many independent LLG runs of one spin model, advanced together on the GPU.

    from magex_amd.batch import ReplicaBatch
    batch = ReplicaBatch(systems, alpha=0.1, dt=1e-4, temperatures=T_ladder, seeds=seeds, precision="fp32")
    e = batch.pack([s.init_spins("random", seed=k) for k, s in enumerate(systems)])
    batch.autotune(e)                   # picks the launch shape by timing candidates on this batch
    batch.run(e, 10000)
    M = batch.observables(e)            # (nrep, 8): E, sum e, sum s e, max torque per replica

The replicas share one unit cell (couplings, on-site tensor, moments, field) and may differ in supercell,
damping, temperature and seed: a temperature ladder, a set of sampling runs, or replicas of different
sizes.  One launch per Heun pass advances every replica (magex has no batched form: its nearest analogue
is a Python loop over replicas, magex/advanced_mc.py:38-94).  Replica k's thermal noise is the stream a
single-replica LLGSolver with seed seeds[k] draws, so a batched run equals separate runs bit for bit.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _lib
from ._lib import check, hptr
from .device import _stream, _torch

# autotuner results per batch signature (process lifetime)
_TUNED = {}


class ReplicaBatch:
    def __init__(self, systems, alpha=0.1, dt=1e-3, temperatures=0.0, seeds=None, precision="fp64",
                 precession=True, noise=None):
        torch = _torch()
        systems = list(systems)
        if not systems:
            raise ValueError("a batch needs at least one replica")
        proto = systems[0]
        for s in systems[1:]:
            if s.uc is not proto.uc and len(s.stencil) != len(proto.stencil):
                raise ValueError("replicas must share one unit-cell model")
            if not np.array_equal(s.B, proto.B):
                raise ValueError("replicas must share the external field")
        for s in systems:
            if not all(s.pbc):
                raise NotImplementedError("batched replicas need periodic supercells")
        self.systems = systems
        self.n = len(systems)
        self.precision = precision
        self.dtype = torch.float64 if precision == "fp64" else torch.float32
        self.lib = _lib.lib()
        self._proto = proto.device_model(precision)
        # generator grade: the precision's own unless given ("float" makes an FP64 batch draw exactly the
        # noise of the FP32 one)
        self._proto.set_noise(noise or ("double" if precision == "fp64" else "float"))
        L = np.ascontiguousarray([s.L for s in systems], np.int32)
        h = ctypes.c_void_p()
        check(self.lib.mx_batch_create(self._proto.h, self.n, hptr(L), ctypes.byref(h)))
        self.h = h
        off = np.zeros(self.n + 1, np.int64)
        check(self.lib.mx_batch_offsets(h, hptr(off)))
        self.offsets = off
        self.dt = float(dt)
        self.alpha = np.broadcast_to(np.asarray(alpha, float), (self.n,)).copy()
        self.temperatures = np.broadcast_to(np.asarray(temperatures, float), (self.n,)).copy()
        self.precession = bool(precession)
        self._set_dynamics()
        if seeds is None:
            seeds = np.random.SeedSequence().generate_state(self.n, dtype=np.uint64)
        self.seeds = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64).reshape(self.n))
        self.step_count = 0
        self.launch = (256, 1, 0)
        self._obs = torch.empty((self.n, _lib.MX_OBS_LEN), dtype=torch.float64, device="cuda")

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self.lib.mx_batch_destroy(h)
            self.h = None

    def _set_dynamics(self):
        check(self.lib.mx_batch_set_dynamics(self.h, self.dt, hptr(np.ascontiguousarray(self.alpha)),
                                             hptr(np.ascontiguousarray(self.temperatures)), int(self.precession)))

    def set_temperatures(self, T):
        self.temperatures = np.broadcast_to(np.asarray(T, float), (self.n,)).copy()
        self._set_dynamics()

    # ------------------------------------------------------------------ spins
    def pack(self, spins):
        """Per-replica (N_k, 3) arrays or tensors -> one packed device tensor (total, 3)."""
        torch = _torch()
        parts = [torch.as_tensor(np.asarray(e) if not hasattr(e, "device") else e).to("cuda", self.dtype)
                 for e in spins]
        for k, p in enumerate(parts):
            if p.shape != (self.systems[k].N, 3):
                raise ValueError(f"replica {k}: spins must have shape {(self.systems[k].N, 3)}")
        return torch.cat(parts).contiguous()

    def unpack(self, e):
        """Views of replica k's spins in the packed tensor."""
        return [e[self.offsets[k]:self.offsets[k + 1]] for k in range(self.n)]

    def _check(self, e):
        if e.shape != (int(self.offsets[-1]), 3) or e.dtype != self.dtype or not e.is_contiguous():
            raise ValueError(f"packed spins must be a contiguous ({int(self.offsets[-1])}, 3) {self.dtype} tensor")

    # ------------------------------------------------------------------ kernels
    def run(self, e, nsteps):
        """nsteps stochastic Heun steps of every replica, in place."""
        self._check(e)
        check(self.lib.mx_batch_run(self.h, ctypes.c_void_p(e.data_ptr()), int(nsteps), hptr(self.seeds),
                                    int(self.step_count), _stream()))
        self.step_count += int(nsteps)
        return e

    def observables(self, e):
        """(nrep, 8): energy (meV), sum e (3), sum s e (3), max |e x B_eff| (T) per replica."""
        self._check(e)
        check(self.lib.mx_batch_observables(self.h, ctypes.c_void_p(e.data_ptr()),
                                            ctypes.c_void_p(self._obs.data_ptr()), _stream()))
        return self._obs.cpu().numpy().copy()

    def magnetization(self, e):
        """|<e>| per replica."""
        o = self.observables(e)
        N = np.diff(self.offsets).astype(float)
        return np.linalg.norm(o[:, 1:4], axis=1) / N

    # ------------------------------------------------------------------ launch shape
    def set_launch(self, block=256, sites_per_thread=1, graph_steps=0):
        check(self.lib.mx_batch_set_launch(self.h, int(block), int(sites_per_thread), int(graph_steps)))
        self.launch = (int(block), int(sites_per_thread), int(graph_steps))

    def signature(self):
        return (self.precision, tuple(s.L for s in self.systems), self.systems[0].uc.nsites,
                len(self.systems[0].stencil))

    def autotune(self, e, steps=20, blocks=(64, 128, 256, 512), spts=(1, 2, 4), graphs=(0, 16), verbose=False):
        """Time every launch shape for `steps` steps on a copy of e and keep the fastest.

        The copy is advanced from the batch's current step counter and discarded, so the tuning leaves
        neither e nor the noise sequence of the real run changed.  Results are cached per signature.
        """
        torch = _torch()
        sig = self.signature()
        if sig in _TUNED:
            self.set_launch(*_TUNED[sig][0])
            return _TUNED[sig]
        scratch = e.clone()
        results = []
        for g in graphs:
            for b in blocks:
                for spt in spts:
                    self.set_launch(b, spt, g)
                    n = max(steps, 2 * g) if g else steps
                    check(self.lib.mx_batch_run(self.h, ctypes.c_void_p(scratch.data_ptr()), 2, hptr(self.seeds),
                                                int(self.step_count), _stream()))
                    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    t0.record()
                    check(self.lib.mx_batch_run(self.h, ctypes.c_void_p(scratch.data_ptr()), n, hptr(self.seeds),
                                                int(self.step_count), _stream()))
                    t1.record()
                    torch.cuda.synchronize()
                    ms = t0.elapsed_time(t1) / n
                    results.append(((b, spt, g), ms))
                    if verbose:
                        print(f"  block {b:3d} spt {spt} graph {g:2d}: {1e3 * ms:8.1f} us/step", flush=True)
        results.sort(key=lambda r: r[1])
        self.set_launch(*results[0][0])
        _TUNED[sig] = (results[0][0], results[0][1], results)
        del scratch
        return _TUNED[sig]


def per_replica_rate(systems, alpha, dt, temperatures, seeds, spins, steps, precision="fp32"):
    """Spin-steps/s of the same replicas run as separate plain simulations, one after another, each
    advanced `steps` steps by its own LLGSolver (two launches per step per replica)."""
    torch = _torch()
    from .llg import LLGSolver
    sols = [LLGSolver(s, alpha[k], dt, temperatures[k], int(seeds[k]), precision=precision)
            for k, s in enumerate(systems)]
    es = [torch.as_tensor(np.asarray(spins[k]), device="cuda",
                          dtype=torch.float32 if precision == "fp32" else torch.float64).contiguous()
          for k in range(len(systems))]
    for s, e in zip(sols, es):
        s.step_device(e, 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, e in zip(sols, es):
        s.step_device(e, steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total = sum(s.N for s in systems)
    return total * steps / wall, es
