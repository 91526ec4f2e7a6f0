"""CPU restatement of magex's geodesic nudged elastic band (test infrastructure only).

TEST INFRASTRUCTURE ONLY.  Nothing in ``magex_amd`` imports this module; only ``tests/`` uses it, as
the checker of the device GNEB.  Restates magex/gneb.py:85-223 (geometry, upwind tangents, GNEB
forces with climbing image, velocity-projection steps with a trust region) over the energy and
gradient of an oracle SpinSystem.

Pinning (tests/test_oracle.py): geodesic distance of a uniform rotation (sqrt(N) theta); the
Stoner-Wohlfarth single spin, whose climbing image must land on the analytic saddle cos(theta*) = -h
with barriers K (1 +- h)^2.
"""
from __future__ import annotations

import numpy as np


def normalize(x):
    return x / np.linalg.norm(x, axis=-1, keepdims=True)


def project(x, v):
    return v - np.einsum("ia,ia->i", v, x)[:, None] * x


def geodesic(a, b):
    c = np.clip(np.einsum("ia,ia->i", a, b), -1.0, 1.0)
    return float(np.sqrt(np.sum(np.arctan2(np.linalg.norm(np.cross(a, b), axis=1), c) ** 2)))


def slerp_path(a, b, nimages):
    """magex's initial chain (gneb.py:97-118): each spin rotated about a x b (Rodrigues) by a fraction of
    its angle; spins with |a x b| < 1e-8 turn about a x u, u = x (y when |a_x| >= 0.9); ends exact."""
    ang = np.arctan2(np.linalg.norm(np.cross(a, b), axis=1), np.clip(np.einsum("ia,ia->i", a, b), -1, 1))
    axis = np.cross(a, b)
    nrm = np.linalg.norm(axis, axis=1)
    bad = nrm < 1e-8
    trial = np.where(np.abs(a[:, 0:1]) < 0.9, [[1.0, 0.0, 0.0]], [[0.0, 1.0, 0.0]])
    axis[bad] = np.cross(a[bad], trial[bad])
    axis = axis / np.linalg.norm(axis, axis=1, keepdims=True)
    out = []
    for t in np.linspace(0, 1, nimages):
        th = (ang * t)[:, None]
        rot = a * np.cos(th) + np.cross(axis, a) * np.sin(th) + axis * np.einsum("ia,ia->i", axis, a)[:, None] * (
            1 - np.cos(th))
        out.append(normalize(rot))
    out[0], out[-1] = a.copy(), b.copy()
    return out


def tangents(x, E):
    t = [None] * len(x)
    for v in range(1, len(x) - 1):
        tp, tm = x[v + 1] - x[v], x[v] - x[v - 1]
        Ep, E0, Em = E[v + 1], E[v], E[v - 1]
        if Ep > E0 > Em:
            tau = tp
        elif Ep < E0 < Em:
            tau = tm
        else:
            hi = max(abs(Ep - E0), abs(Em - E0))
            lo = min(abs(Ep - E0), abs(Em - E0))
            tau = tp * hi + tm * lo if Ep > Em else tp * lo + tm * hi
        tau = project(x[v], tau)
        t[v] = tau / max(np.linalg.norm(tau), 1e-30)
    return t


def forces(system, x, k, climbing=True):
    E = np.array([system.energy(im) for im in x])
    G = [system.gradient(im) for im in x]
    t = tangents(x, E)
    ci = int(np.argmax(E[1:-1])) + 1 if (climbing and len(x) > 2) else -1
    F = [np.zeros_like(im) for im in x]
    for v in range(1, len(x) - 1):
        g = project(x[v], G[v])
        gt = float(np.sum(g * t[v]))
        if v == ci:
            F[v] = -g + 2 * gt * t[v]
        else:
            F[v] = -g + (gt + k * (geodesic(x[v + 1], x[v]) - geodesic(x[v], x[v - 1]))) * t[v]
    return E, F, ci


def spring_constant(system, x, spring):
    g = [float(np.max(np.linalg.norm(project(im, system.gradient(im)), axis=1))) for im in x[1:-1]]
    return float(spring) * max(float(np.median(g)) if g else 1.0, 1e-300)


def run(system, images, spring=1.0, climbing=True, rel_tol=1e-4, tol=0.0, max_iter=20000, max_step=0.02):
    x = [normalize(np.asarray(im, float)) for im in images]
    k = spring_constant(system, x, spring)
    vel = [np.zeros_like(im) for im in x]
    E, F, ci = forces(system, x, k, climbing)
    fmax0 = max(float(np.max(np.linalg.norm(f, axis=1))) for f in F[1:-1])
    scale = 1.0 / max(fmax0, 1e-30)
    it = 0
    for it in range(1, max_iter + 1):
        fmax = max(float(np.max(np.linalg.norm(f, axis=1))) for f in F[1:-1])
        if fmax < tol or fmax < rel_tol * fmax0:
            break
        for v in range(1, len(x) - 1):
            f = F[v] * scale
            fn = np.linalg.norm(f)
            vp = float(np.sum(vel[v] * f))
            vel[v] = ((vp / fn**2) * f if (vp > 0 and fn > 0) else np.zeros_like(f)) + 0.05 * f
            disp = 0.05 * vel[v]
            dmax = float(np.max(np.linalg.norm(disp, axis=1)))
            if dmax > max_step:
                disp *= max_step / dmax
                vel[v] *= max_step / dmax
            x[v] = normalize(x[v] + disp)
            vel[v] = project(x[v], vel[v])
        E, F, ci = forces(system, x, k, climbing)
    return x, E, ci, it
