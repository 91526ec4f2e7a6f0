"""CPU restatement of the device Monte Carlo colour sweep (test infrastructure only).

TEST INFRASTRUCTURE ONLY.  Nothing in ``magex_amd`` imports this module; only ``tests/`` uses it, as
the checker.  The sweep is magex's numba `_mc_group` (magex/fast.py:393-503) as driven by
`MetropolisMC._fast_sweep` (magex/montecarlo.py:152-179): per colour, H = 2 h + mu mu_B B, a proposal,
dE = dE_onsite - H.(new - old) for the linear moves, acceptance if dE <= 0 or u < exp(-min(beta dE, 700)).
magex draws from xorshift64* streams; this restatement (like the device) draws from Philox4x32-10
keyed by (seed, sweep, site), converted to 53-bit uniforms and Box-Muller normals, so that the two
can be compared step for step.

Pinning (tests/test_oracle.py): single-spin Langevin averages <cos theta> = coth x - 1/x for every
move, the dimer average <e1.e2> = L(2 beta J) (magex's per-ordered-pair convention), a quadrature of
the anisotropic single-spin distribution for the Metropolis-corrected heat bath, and exact energy
conservation of over-relaxation without on-site terms.
"""
from __future__ import annotations

import numpy as np

from . import magex_ref as R
from .philox import philox4x32_10

MC_TAG = 0x4D430000


def uniforms(seed, sweep, sites, tag):
    """Two 53-bit uniforms per site from one Philox call: counter (site, sweep lo, sweep hi, tag)."""
    seed, sweep = int(seed), int(sweep)
    c = np.zeros((len(sites), 4), np.uint32)
    c[:, 0] = np.asarray(sites, np.uint32)
    c[:, 1] = sweep & 0xFFFFFFFF
    c[:, 2] = (sweep >> 32) & 0xFFFFFFFF
    c[:, 3] = tag
    k = np.array([seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF], np.uint32)
    w = philox4x32_10(c, np.broadcast_to(k, (len(sites), 2))).astype(np.float64)
    u0 = (np.floor(w[:, 0] / 32) * 67108864.0 + np.floor(w[:, 1] / 64) + 0.5) * 2.0**-53
    u1 = (np.floor(w[:, 2] / 32) * 67108864.0 + np.floor(w[:, 3] / 64) + 0.5) * 2.0**-53
    return u0, u1


def onsite_energy(system, e, idx):
    """e.A.e plus the terms' local energies of sites idx at directions e (n, 3) (fast.py:66-84)."""
    E = np.einsum("ia,iab,ib->i", e, system.A[idx], e)
    for t in system.terms:
        if isinstance(t, R.CubicAnisotropy):
            q = (e @ t.C.T) ** 2
            E = E + t.Kc[idx] * (q[:, 0] * q[:, 1] + q[:, 1] * q[:, 2] + q[:, 2] * q[:, 0])
        elif isinstance(t, R.UniaxialHigher):
            c = e @ t.u
            E = E - t.K2[idx] * c**4 - t.K3[idx] * c**6
    return E


def _heatbath(H, beta, u0, u1):
    Hn = np.linalg.norm(H, axis=1)
    x = beta * Hn
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        c = 1 + np.log1p(-u0 * (-np.expm1(-2 * x))) / x
    c = np.where((x < 1e-8) | ~np.isfinite(x), np.where(x < 1e-8, 2 * u0 - 1, 1.0), c)
    c = np.clip(c, -1, 1)
    ph = 2 * np.pi * u1
    z = np.where((Hn > 0)[:, None], H / np.where(Hn > 0, Hn, 1)[:, None], np.array([0, 0, 1.0]))
    a = np.where((np.abs(z[:, 0]) < 0.9)[:, None], np.array([1.0, 0, 0]), np.array([0, 1.0, 0]))
    t1 = np.cross(z, a)
    t1 = t1 / np.linalg.norm(t1, axis=1, keepdims=True)
    t2 = np.cross(z, t1)
    sn = np.sqrt(np.maximum(0.0, 1 - c * c))[:, None]
    return c[:, None] * z + sn * (np.cos(ph)[:, None] * t1 + np.sin(ph)[:, None] * t2)


def sweep(system, e, groups, beta, move, overrelax, sigma, seed, sweep_no):
    """One sweep of e (N, 3) in place; beta: scalar or (N,) (inf at T = 0); move 0/1/2.
    Returns the accepted count of the moves (over-relaxation not counted)."""
    N = system.N
    beta = np.broadcast_to(np.asarray(beta, float), (N,))
    if move == 2 and not np.all(np.isfinite(beta)):
        move = 0
    flat = e.reshape(-1)
    accepted = 0
    for pass_ in range(overrelax + 1):
        tag = MC_TAG | (pass_ << 8)
        for idx in groups:
            h = (system.Jsp[np.ravel(3 * idx[:, None] + np.arange(3))] @ flat).reshape(-1, 3)
            H = 2 * h + R.MU_B * system.mu[idx, None] * system.B[None, :]
            o = e[idx].copy()
            bt = beta[idx]
            live = np.ones(len(idx), bool)
            lin = False
            if pass_ > 0:
                HH = np.einsum("ij,ij->i", H, H)
                live = HH >= 1e-28
                p = 2 * np.einsum("ij,ij->i", o, H) / np.where(live, HH, 1.0)
                n = R.normalize(np.where(live[:, None], p[:, None] * H - o, o))
                ua, _ = uniforms(seed, sweep_no, idx, tag)
            elif move == 2:
                u0, u1 = uniforms(seed, sweep_no, idx, tag)
                ua, _ = uniforms(seed, sweep_no, idx, tag | 1)
                n = _heatbath(H, bt, u0, u1)
            else:
                a0, a1 = uniforms(seed, sweep_no, idx, tag)
                b0, b1 = uniforms(seed, sweep_no, idx, tag | 1)
                ua, _ = uniforms(seed, sweep_no, idx, tag | 2)
                r0, r2 = np.sqrt(-2 * np.log(a0)), np.sqrt(-2 * np.log(b0))
                g = np.stack([r0 * np.cos(2 * np.pi * a1), r0 * np.sin(2 * np.pi * a1), r2 * np.cos(2 * np.pi * b1)], 1)
                n = R.normalize(o + sigma * g if move == 0 else g)
                lin = True
            dE = onsite_energy(system, n, idx) - onsite_energy(system, o, idx)
            if lin:
                dE = dE - np.einsum("ij,ij->i", H, n - o)
            with np.errstate(invalid="ignore", over="ignore"):
                ok = live & ((dE <= 0) | (ua < np.exp(-np.minimum(dE * bt, 700.0))))
            e[idx[ok]] = n[ok]
            if pass_ == 0:
                accepted += int(ok.sum())
    return accepted
