"""Parallel tempering on the device (magex/advanced_mc.py:38-94): M replicas of one lattice at a
temperature ladder, swept together (one launch per colour and pass for all replicas), with
configuration swaps between neighbouring temperatures every `swap_every` sweeps accepted with
probability min(1, exp[(beta_i - beta_j)(E_i - E_j)]).  Same arguments, schedule (cone-width
adaptation every 50 thermalisation sweeps, alternating even/odd swap pairs, sampling every
`sample_every` measurement sweeps) and result rows as magex.  Wang-Landau sampling is not on the device
path.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from ._lib import check, hptr
from .constants import K_B
from .device import _stream, _torch
from .montecarlo import MOVES, MetropolisMC


def _obs(E, m, N, kT):
    return (np.mean(E) / N, (np.mean(E**2) - np.mean(E) ** 2) / (N * kT**2), np.mean(m),
            N * (np.mean(m**2) - np.mean(m) ** 2) / kT, 1 - np.mean(m**4) / (3 * np.mean(m**2) ** 2))


def parallel_tempering(system, temperatures, therm_sweeps=2000, meas_sweeps=4000, sample_every=2, swap_every=5,
                       seed=1, init="reference", init_axis=(0, 0, 1), sigma=0.6, overrelax=1, log=None,
                       precision="fp64", move="gaussian", return_spins=False):
    """Returns (rows like magex.montecarlo._point, swap acceptance per neighbouring pair), and the final
    replica spins (M, N, 3) device tensor, ordered by temperature, when return_spins."""
    torch = _torch()
    Ts = np.array(sorted(float(t) for t in temperatures))
    M = len(Ts)
    rng = np.random.default_rng(seed)
    mc = MetropolisMC(system, Ts[0], seed=seed, sigma=sigma, overrelax=overrelax, move=move, precision=precision)
    dm = mc._sync()
    if getattr(system, "site_T", None) is not None:
        raise NotImplementedError("parallel tempering with per-site temperatures")
    N = system.N
    seeds = np.array([seed + 17 * k for k in range(M)], np.uint64)
    betas = np.where(Ts > 0, 1.0 / (K_B * np.maximum(Ts, 1e-300)), np.inf)
    sig = np.full(M, float(sigma))
    conf = torch.stack([torch.as_tensor(np.asarray(system.init_spins(init, init_axis, seed + k), float),
                                        dtype=dm.dtype) for k in range(M)]).to("cuda").contiguous()
    nfree = sum(len(g) for g in mc.groups)
    tried = np.zeros(M)

    def energies():
        return np.array([float(dm.observables_device(conf[k])[0]) for k in range(M)])

    def read_acc(reset=True):
        out = np.zeros(M, np.uint64)
        check(dm.lib.mx_mc_replica_accepted(dm.h, M, hptr(out), int(reset)))
        return out

    swaps_try = np.zeros(M - 1)
    swaps_acc = np.zeros(M - 1)
    samples = [[] for _ in range(M)]
    t0 = time.time()
    total = therm_sweeps + meas_sweeps

    def adapt(n):
        return n < therm_sweeps and (n + 1) % 50 == 0

    def swap(n):
        return (n + 1) % swap_every == 0

    def sample(n):
        return n >= therm_sweeps and (n - therm_sweeps) % sample_every == 0

    read_acc()                                   # start the counters at zero
    n = 0
    while n < total:
        # sweeps n .. m run in one call; m is the next sweep after which something happens
        m = n
        while m < total - 1 and not (adapt(m) or swap(m) or sample(m)):
            m += 1
        check(dm.lib.mx_mc_sweep_replicas(dm.h, ctypes.c_void_p(conf.data_ptr()), M, m - n + 1, MOVES[mc.move],
                                          overrelax, hptr(np.ascontiguousarray(betas)), hptr(np.ascontiguousarray(sig)),
                                          hptr(seeds), n, _stream()))
        tried += (m - n + 1) * nfree
        if adapt(m):                             # magex/advanced_mc.py:56-58
            a = read_acc() / np.maximum(tried, 1)
            tried[:] = 0
            sig = np.clip(sig * 0.5 / np.maximum(1 - a, 0.05), 0.02, 4.0)
        if swap(m):                              # magex/advanced_mc.py:59-68
            E = energies()
            start = (m // swap_every) % 2
            for k in range(start, M - 1, 2):
                d = (betas[k] - betas[k + 1]) * (E[k] - E[k + 1])
                swaps_try[k] += 1
                if d >= 0 or rng.random() < np.exp(d):
                    conf[[k, k + 1]] = conf[[k + 1, k]]
                    E[k], E[k + 1] = E[k + 1], E[k]
                    swaps_acc[k] += 1
        if sample(m):                            # magex/advanced_mc.py:69-72
            for k in range(M):
                o = dm.observables_device(conf[k]).cpu().numpy()
                samples[k].append((o[0], np.linalg.norm(o[4:7]) / N))
        n = m + 1
    accepted = read_acc()
    rows = []
    for k, T in enumerate(Ts):
        Es = np.array([s[0] for s in samples[k]])
        ms = np.array([s[1] for s in samples[k]])
        kT = K_B * T
        full = _obs(Es, ms, N, kT) if len(Es) else (np.nan,) * 5
        nb = 10
        L = len(Es) // nb
        err = (np.array([_obs(Es[b * L:(b + 1) * L], ms[b * L:(b + 1) * L], N, kT) for b in range(nb)]).std(0, ddof=1)
               / np.sqrt(nb)) if L > 2 else np.zeros(5)
        rows.append(dict(L=list(system.L), N=N, T=float(T), E=full[0], C=full[1], m=full[2], chi=full[3], U4=full[4],
                         dE=err[0], dC=err[1], dm=err[2], dchi=err[3], dU4=err[4],
                         acceptance=float(accepted[k]) / max(tried[k], 1), sigma=float(sig[k]), ncolors=mc.ncolors,
                         mvec=[0, 0, 0], seconds=time.time() - t0))
    swap = (swaps_acc / np.maximum(swaps_try, 1)).tolist()
    if log:
        log(f"  parallel tempering L={system.L}: {M} replicas, swap acceptance "
            f"{np.round(swap, 2).tolist()}, {time.time() - t0:.1f} s")
    return (rows, swap, conf) if return_spins else (rows, swap)
