"""Metropolis Monte Carlo on the device: the colour sweep of magex's fast path (magex/montecarlo.py:71-239,
magex/fast.py:393-503) on libmxllg's stencil gather.

    mc = MetropolisMC(system, temperature, seed=None, sigma=0.6, overrelax=1, move="gaussian")
    e = mc.sweep(e)                # one sweep, in place (numpy or device tensor)
    mc.acceptance()                # accepted / tried since the last reset

A colour is a set of sites with no coupling between them, so a colour moves at once (one launch).  The
colouring is linear over cells when the stencil allows one, colour = (sublattice, a.c mod M), which
gives few, regular colours; otherwise it is magex's greedy colouring of the bond graph
(magex/montecarlo.py:35-49).  Moves: "gaussian" (cone of width sigma), "uniform", "heatbath"; then
`overrelax` over-relaxation passes.  Random numbers come from a counter-based generator keyed by
(seed, sweep, site), so a sweep's outcome does not depend on the launch shape.  "hinzke" moves and
dipolar / four-spin terms are not on the device path (NotImplementedError).
"""
from __future__ import annotations

import ctypes
import weakref

import numpy as np

from ._lib import check, hptr
from .device import _stream, _torch, as_device
from .structure import site_temperature

MOVES = {"gaussian": 0, "uniform": 1, "heatbath": 2}


def greedy_coloring(N, bonds):
    """Colour array (N,) such that bonded sites differ (bonds: (M, 2) src, dst); magex's greedy order
    (magex/montecarlo.py:35-49)."""
    import scipy.sparse as sp
    bonds = np.asarray(bonds, int).reshape(-1, 2)
    if len(bonds) == 0:
        return np.zeros(N, int)
    adj = sp.csr_matrix((np.ones(len(bonds)), (bonds[:, 0], bonds[:, 1])), shape=(N, N))
    adj = ((adj + adj.T) > 0).tocsr()
    color = -np.ones(N, int)
    indptr, indices = adj.indptr, adj.indices
    for v in range(N):
        used = set(color[indices[indptr[v]:indptr[v + 1]]].tolist())
        c = 0
        while c in used:
            c += 1
        color[v] = c
    return color


def linear_coloring(system, max_m=8):
    """(a, M) with colour = sublattice * M + (a . cell mod M) valid for the stencil, or None.

    Valid: no entry couples a site to one of its own colour (a . R != 0 mod M for every same-sublattice
    entry), and the colouring survives the periodic wrap (a_k L_k = 0 mod M on periodic axes)."""
    same = [np.asarray(R, int) for s, t, R, _ in system.stencil if s == t]
    for M in range(2, max_m + 1):
        for a0 in range(M):
            for a1 in range(M):
                for a2 in range(M):
                    a = np.array([a0, a1, a2])
                    if any(p and (a[k] * system.L[k]) % M for k, p in enumerate(system.pbc)):
                        continue
                    if all((a @ R) % M for R in same):
                        return a, M
    return None


def coloring(system):
    """Colour of every site (N,) for the stencil of `system`."""
    lin = linear_coloring(system)
    if lin is not None:
        a, M = lin
        cell = (system.cells @ a) % M
        return system.sublattice * M + cell[system.cell_index]
    bonds = [np.stack(system.neighbour(k), 1) for k in range(len(system.stencil))]
    bonds = np.concatenate(bonds) if bonds else np.zeros((0, 2), int)
    bonds = bonds[bonds[:, 1] >= 0]
    return greedy_coloring(system.N, bonds)


class MetropolisMC:
    """magex's MetropolisMC (magex/montecarlo.py:71-239) on the device; precision "fp64" or "fp32"."""

    def __init__(self, system, temperature, seed=None, sigma=0.6, overrelax=1, move="gaussian", backend="auto",
                 dipolar_exact="auto", precision="fp64"):
        if move not in MOVES:
            if move == "hinzke":
                raise NotImplementedError("hinzke moves are not on the device path (gaussian, uniform, heatbath are)")
            raise ValueError(f"move must be one of {tuple(MOVES)} or 'hinzke'")
        if backend == "numpy":
            raise RuntimeError("magex_amd has no NumPy path: use backend='auto' (HIP kernels)")
        self.move = move
        self.sys = system
        self.T = float(temperature)
        self.rng = np.random.default_rng(seed)
        self.seed = int(seed) if seed is not None else int(self.rng.integers(0, 2**63))
        self.sigma = float(sigma)
        self.overrelax = int(overrelax)
        self.precision = precision
        col = coloring(system)
        # colours that are the parity classes of c0 + c1 + c2 (bcc, sc): the library's checkerboard cell
        # order then keeps every colour contiguous in memory (where the supercell allows that order)
        lin = linear_coloring(system)
        self.layout = None
        if lin is not None and lin[1] == 2 and all(int(a) % 2 == 1 or L == 1 for a, L in zip(lin[0], system.L)):
            self.layout = "checker"
        pinned = system.pinned
        self.groups = []
        for c in range(int(col.max()) + 1 if len(col) else 0):
            idx = np.where(col == c)[0]
            if pinned is not None:
                idx = idx[~pinned[idx]]
            if len(idx):
                self.groups.append(idx)
        self.ncolors = len(self.groups)
        self.accepted = 0
        self.tried = 0
        self.sweeps = 0
        self._dm = None
        self._key = None

    def _sync(self):
        dm = self.sys.device_model(self.precision, layout=self.layout)
        key = (id(dm), self.T, id(getattr(self.sys, "site_T", None)))
        # the device model (colour lists, temperatures, acceptance counter) is shared by every sampler
        # of this system: a sampler that was not the last one to sweep it uploads its own set-up again,
        # and the accepted moves still on the device are credited to the sampler that made them
        owner = getattr(dm, "_mc_owner", None)
        prev = owner() if owner is not None else None
        if prev is not self:
            if prev is not None:
                n = ctypes.c_int64()
                check(dm.lib.mx_mc_accepted(dm.h, ctypes.byref(n), 1))
                prev.accepted += n.value
            self._dm = None
            dm._mc_owner = weakref.ref(self)
        if dm is not self._dm:
            off = np.concatenate([[0], np.cumsum([len(g) for g in self.groups])]).astype(np.int64)
            sites = np.ascontiguousarray(np.concatenate(self.groups) if self.groups else np.zeros(0), np.int32)
            check(dm.lib.mx_mc_set_colors(dm.h, len(self.groups), hptr(off), hptr(sites)))
            self._dm = dm
            self._key = None
        if key != self._key:
            T = np.asarray(site_temperature(self.sys, self.T), float)
            per_site = T.ndim > 0 and T.size > 1
            check(dm.lib.mx_mc_set_temperature(dm.h, int(per_site), hptr(np.ascontiguousarray(np.atleast_1d(T)))))
            self._key = key
        return dm

    def sweep_device(self, e_dev, nsweeps=1):
        """nsweeps sweeps of a device tensor (N, 3) in the solver's precision, in place."""
        dm = self._sync()
        dm._check_spins(e_dev)
        check(dm.lib.mx_mc_sweep(dm.h, ctypes.c_void_p(e_dev.data_ptr()), int(nsweeps), MOVES[self.move],
                                 self.overrelax, self.sigma, self.seed & 0xFFFFFFFFFFFFFFFF, self.sweeps, _stream()))
        self.sweeps += int(nsweeps)
        self.tried += int(nsweeps) * sum(len(g) for g in self.groups)
        return e_dev

    def sweep(self, e):
        """One sweep; numpy spins are updated in place and returned (magex/montecarlo.py:181-219)."""
        torch = _torch()
        if isinstance(e, torch.Tensor) and e.is_cuda:
            return self.sweep_device(e)
        dm = self._sync()
        ed = as_device(e, dm.dtype)
        self.sweep_device(ed)
        e[...] = ed.cpu().numpy()
        return e

    def acceptance(self, reset=True):
        dm = self._sync()
        n = ctypes.c_int64()
        check(dm.lib.mx_mc_accepted(dm.h, ctypes.byref(n), 1))   # the device counter moves into self.accepted
        self.accepted += n.value
        a = self.accepted / max(self.tried, 1)
        if reset:
            self.accepted = self.tried = 0
        return a
