"""Magnon spectra from spin dynamics on the device (magex/spectra.py:49-84 and
magex/observables.py:66-71).

    E, S_perp, S_par, e = structure_factor(system, solver, e, qfrac, nsteps, sample_every=10, nseg=4)

The trajectory runs with `solver.step_device`; every `sample_every` steps the library forms
m_q = (1/sqrt N) sum_i e_i exp(-i q.r_i) for all q on the device (phases computed per site, nothing of
size N x nq stored) into a device array of samples, and the Hann-windowed Welch average of those samples
is taken on the device too (``welch_power``); only the spectra come back.
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


def welch_power(mq, order_dir, dt, nseg):
    """Hann-windowed Welch power spectra of sampled m_q(t), on the device.

    mq: (nsamp, nq, 3) complex device tensor; order_dir: (3,) order direction (normalised here); dt:
    time between samples (ps).  The samples split into nseg segments of nsamp // nseg; each segment has
    its mean removed, is windowed (symmetric Hann) and Fourier transformed in time, and the powers are
    averaged over the segments.  The longitudinal power is |n . F|^2; the transverse power is the total
    minus it (|F|^2 = |n . F|^2 + |F_perp|^2), so no transverse frame is needed.
    Returns (E (seg,) meV ascending, S_perp (nq, seg), S_par (nq, seg)) as NumPy arrays.
    """
    torch = _torch()
    seg = mq.shape[0] // nseg
    x = mq[: seg * nseg].reshape(nseg, seg, *mq.shape[1:])
    x = x - x.mean(dim=1, keepdim=True)
    win = torch.hann_window(seg, periodic=False, dtype=torch.float64, device=mq.device)
    F = torch.fft.fft(x * win[None, :, None, None], dim=1)
    n = torch.as_tensor(np.asarray(order_dir, float), dtype=torch.float64, device=mq.device)
    n = (n / n.norm()).to(F.dtype)
    par = (F * n).sum(-1).abs().square().mean(0)
    perp = F.abs().square().sum(-1).mean(0) - par
    E = 2 * np.pi * HBAR_MEV_PS * torch.fft.fftfreq(seg, dt, dtype=torch.float64)
    E, perp, par = (torch.fft.fftshift(t, dim=0) for t in (E, perp, par))
    return E.cpu().numpy(), perp.T.cpu().numpy(), par.T.cpu().numpy()


def structure_factor(system, solver, e, qfrac, nsteps, sample_every=10, nseg=4, log=None):
    """Run `nsteps` LLG steps sampling every `sample_every`; returns (E_meV, S_perp, S_par, e)
    (magex/spectra.py:49-84).  qfrac: (nq, 3) fractional coordinates of the unit cell's reciprocal
    lattice.  e: numpy or device tensor; the returned spins have the same type.  Samples, the order
    direction and the spectra stay on the device until the end."""
    torch = _torch()
    qc = np.asarray(qfrac, float) @ (2 * np.pi * np.linalg.inv(system.uc.cell).T)
    samp = Sampler(system, qc, solver.precision)
    solver._sync()
    ed = as_device(e, solver.dm.dtype)
    if ed is e:
        ed = ed.clone()
    nsamp = nsteps // sample_every
    raw = torch.empty((nsamp, len(qc), 6), dtype=torch.float64, device="cuda")
    order = torch.zeros(3, dtype=torch.float64, device="cuda")
    signs = torch.as_tensor(np.asarray(system.signs, float), device="cuda")[:, None]
    t0 = time.time()
    for k in range(nsamp):
        solver.step_device(ed, sample_every)
        order += (signs * ed.double()).sum(0)
        samp.sample(ed, raw[k])
    mq = torch.view_as_complex(raw.reshape(nsamp, len(qc), 3, 2))
    E, Sperp, Spar = welch_power(mq, order.cpu().numpy(), solver.dt * sample_every, nseg)
    if log:
        log(f"  S(q,w): {len(qc)} wave vectors, {nsamp} samples in {nseg} windows, energy step "
            f"{E[1] - E[0]:.3f} meV up to {E.max():.1f} meV ({time.time() - t0:.1f} s)")
    out = ed if isinstance(e, torch.Tensor) else ed.cpu().numpy().astype(float)
    return E, Sperp, Spar, out


def static_structure_factor(system, e, q):
    """S(q) for q (M, 3) in 1/Angstrom: |sum_i e_i exp(i q.r_i)|^2 / N summed over components
    (magex/observables.py:66-71)."""
    samp = Sampler(system, np.atleast_2d(q), "fp64" if not str(getattr(e, "dtype", "")).endswith("32") else "fp32")
    ed = as_device(e, samp.dm.dtype)
    m = Sampler.to_complex(samp.sample(ed))
    return np.sum(np.abs(m) ** 2, axis=1)
