"""Magnon spectra from spin dynamics on the device (magex/spectra.py:49-84 and
magex/observables.py:66-71).

    E, S_perp, S_par, e = structure_factor(system, solver, e, qfrac, nsteps, sample_every=10, nseg=4)

The trajectory runs with `solver.step_device`; every `sample_every` steps the library forms
m_q = (1/sqrt N) sum_i e_i exp(-i q.r_i) for all q on the device (phases computed per site, nothing of
size N x nq stored) into a device array of samples.  Only the samples (nsamp x nq x 3 complex) come back
for the Hann-windowed Welch average, as in magex.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from ._lib import check, hptr
from .device import _stream, _torch, as_device

HBAR_MEV_PS = 0.6582119569     # meV ps (magex/quantum.py:34)


class Sampler:
    """m_q of a SpinSystem's spins for fixed Cartesian wave vectors q (nq, 3) in 1/Angstrom."""

    def __init__(self, system, qcart, precision="fp64"):
        self.dm = system.device_model(precision)
        self.q = np.ascontiguousarray(np.atleast_2d(np.asarray(qcart, float)))
        self.nq = len(self.q)
        self.N = system.N
        h = ctypes.c_void_p()
        cell = np.ascontiguousarray(np.asarray(system.uc.cell, float).reshape(9))
        basis = np.ascontiguousarray(np.asarray(system.uc.positions, float).reshape(-1, 3))
        check(self.dm.lib.mx_sampler_create(self.dm.h, self.nq, hptr(self.q), hptr(cell), hptr(basis),
                                            ctypes.byref(h)))
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self.dm.lib.mx_sampler_destroy(h)
            self.h = None

    def sample(self, e_dev, out=None):
        """m_q (nq, 3) complex as a device float64 tensor view (nq, 6) written into `out` if given."""
        torch = _torch()
        self.dm._check_spins(e_dev)
        if out is None:
            out = torch.empty((self.nq, 6), dtype=torch.float64, device="cuda")
        check(self.dm.lib.mx_sample_mq(self.h, ctypes.c_void_p(e_dev.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                       _stream()))
        return out

    @staticmethod
    def to_complex(mq):
        a = np.asarray(mq.cpu().numpy() if hasattr(mq, "cpu") else mq).reshape(*np.shape(mq)[:-1], 3, 2)
        return a[..., 0] + 1j * a[..., 1]


def _frame(n):
    n = np.asarray(n, float) / np.linalg.norm(n)
    a = np.array([1.0, 0, 0]) if abs(n[0]) < 0.9 else np.array([0, 1.0, 0])
    e1 = np.cross(n, a)
    e1 /= np.linalg.norm(e1)
    return e1, np.cross(n, e1), n


def structure_factor(system, solver, e, qfrac, nsteps, sample_every=10, nseg=4, log=None):
    """Run `nsteps` LLG steps sampling every `sample_every`; returns (E_meV, S_perp, S_par, e)
    (magex/spectra.py:49-84).  qfrac: (nq, 3) fractional coordinates of the unit cell's reciprocal
    lattice.  e: numpy or device tensor; the returned spins have the same type."""
    torch = _torch()
    B = 2 * np.pi * np.linalg.inv(system.uc.cell).T
    qc = np.asarray(qfrac, float) @ B
    samp = Sampler(system, qc, solver.precision)
    solver._sync()
    ed = as_device(e, solver.dm.dtype)
    if ed is e:
        ed = ed.clone()
    nsamp = nsteps // sample_every
    mq = torch.empty((nsamp, len(qc), 6), dtype=torch.float64, device="cuda")
    ref = torch.zeros(3, dtype=torch.float64, device="cuda")
    signs = torch.as_tensor(np.asarray(system.signs, float), device="cuda")[:, None]
    t0 = time.time()
    for s in range(nsamp):
        solver.step_device(ed, sample_every)
        ref += (signs * ed.double()).sum(0)
        samp.sample(ed, mq[s])
    m = Sampler.to_complex(mq)
    e1, e2, n = _frame(ref.cpu().numpy())
    dt = solver.dt * sample_every
    seg = nsamp // nseg
    win = np.hanning(seg)
    Sperp = 0.0
    Spar = 0.0
    for k in range(nseg):
        x = m[k * seg:(k + 1) * seg]
        x = x - x.mean(0, keepdims=True)
        F = np.fft.fft(x * win[:, None, None], axis=0)
        Sperp = Sperp + np.abs(F @ e1) ** 2 + np.abs(F @ e2) ** 2
        Spar = Spar + np.abs(F @ n) ** 2
    f = np.fft.fftfreq(seg, dt)
    E = 2 * np.pi * f * HBAR_MEV_PS
    order = np.argsort(E)
    if log:
        log(f"  S(q,w): {nsamp} samples x {len(qc)} q-points, {nseg} segments, "
            f"dE = {E[1] - E[0]:.3f} meV, E_max = {E.max():.1f} meV, {time.time() - t0:.1f} s")
    out = ed if isinstance(e, torch.Tensor) else ed.cpu().numpy().astype(float)
    return E[order], (Sperp / nseg)[order].T, (Spar / nseg)[order].T, out


def static_structure_factor(system, e, q):
    """S(q) for q (M, 3) in 1/Angstrom: |sum_i e_i exp(i q.r_i)|^2 / N summed over components
    (magex/observables.py:66-71)."""
    samp = Sampler(system, np.atleast_2d(q), "fp64" if not str(getattr(e, "dtype", "")).endswith("32") else "fp32")
    ed = as_device(e, samp.dm.dtype)
    m = Sampler.to_complex(samp.sample(ed))
    return np.sum(np.abs(m) ** 2, axis=1)
