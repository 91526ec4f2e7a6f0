"""Geodesic nudged elastic band on the device (magex/gneb.py:27-245, Bessarab, Uzdin & Jonsson, Comput.
Phys. Commun. 196, 335 (2015)).

The image chain lives on the device as one (P, N, 3) tensor.  Every iteration evaluates the energies and
gradients of all images in one launch each (mx_observables_images, mx_field_images), and the tangents,
projections, geodesic distances and the velocity-projection step run as whole-chain tensor operations.
Same classes, arguments, schedule and results as magex for the atomistic engine; the micromagnetic
adapter is not on this path.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _lib
from ._lib import check
from .device import _stream, _torch, as_device


# ----------------------------------------------------------------------------- geometry
def _spin_angles(a, b):
    """Angle between corresponding unit spins, 2 atan2(|a - b|, |a + b|): accurate at 0 and at pi."""
    return 2.0 * np.arctan2(np.linalg.norm(a - b, axis=-1), np.linalg.norm(a + b, axis=-1))


def geodesic_distance(a, b):
    """Geodesic distance sqrt(sum_i angle_i^2) between two spin configurations (magex/gneb.py:91-94)."""
    ang = _spin_angles(np.asarray(a, float), np.asarray(b, float))
    return float(np.sqrt(ang @ ang))


def slerp_path(a, b, nimages):
    """nimages configurations from a to b along each spin's great circle (magex/gneb.py:97-118), built on
    the device: Shoemake's slerp (sin((1 - t) th) a + sin(t th) b) / sin th; spins closer than 1e-8 to
    parallel interpolate linearly; spins closer than 1e-8 to antiparallel turn by t th about a x u, with
    u = x unless a is nearly along x (then y) -- magex's choice of the half-turn.  The end images are a
    and b exactly.  Returns a list of (N, 3) NumPy arrays."""
    torch = _torch()
    A = torch.as_tensor(np.asarray(a, float), device="cuda")
    B = torch.as_tensor(np.asarray(b, float), device="cuda")
    th = 2.0 * torch.atan2((A - B).norm(dim=-1), (A + B).norm(dim=-1))[None, :, None]
    t = torch.linspace(0.0, 1.0, nimages, dtype=torch.float64, device="cuda")[:, None, None]
    cross = torch.linalg.cross(A, B, dim=-1).norm(dim=-1)[None, :, None]
    near = cross < 1e-8
    sin_th = torch.sin(th).where(~near, torch.ones_like(th))
    arc = (torch.sin((1.0 - t) * th) * A + torch.sin(t * th) * B) / sin_th
    lin = (1.0 - t) * A + t * B
    u = torch.zeros_like(A)
    along_x = A[:, 0].abs() >= 0.9
    u[:, 0] = (~along_x).to(A.dtype)
    u[:, 1] = along_x.to(A.dtype)
    k = torch.linalg.cross(A, u, dim=-1)
    k = k / k.norm(dim=-1, keepdim=True)
    turn = A * torch.cos(t * th) + torch.linalg.cross(k, A, dim=-1)[None] * torch.sin(t * th)
    antipar = near & ((A * B).sum(-1)[None, :, None] < 0)
    X = torch.where(antipar, turn, torch.where(near, lin, arc))
    X = X / X.norm(dim=-1, keepdim=True)
    out = [x for x in X.cpu().numpy()]
    out[0], out[-1] = np.array(a, float), np.array(b, float)
    return out


# ----------------------------------------------------------------------------- adapter
class AtomisticAdapter:
    """Energy (meV) and gradient dE/de of a SpinSystem, the engine interface of magex's GNEB
    (magex/gneb.py:27-47); every evaluation runs the batched image kernels, one image or many."""
    unit = "meV"
    kB = 8.617333262e-2

    def __init__(self, system, precision="fp64"):
        self.sys = system
        self.precision = precision

    @property
    def dm(self):
        return self.sys.device_model(self.precision)

    def energy_grad(self, x):
        X = as_device(np.asarray(x, float), self.dm.dtype)[None].contiguous()
        E, G = self.images_energy_grad(X)
        return float(E[0]), G[0].double().cpu().numpy()

    def energy(self, x):
        return self.energy_grad(x)[0]

    def grad(self, x):
        return self.energy_grad(x)[1]

    # states are the spin arrays themselves for the atomistic engine
    def to_state(self, x):
        return x

    def from_state(self, e):
        return np.asarray(e, float)

    def images_energy_grad(self, X):
        """(E (P,) float64 device tensor, dE/de (P, N, 3) device tensor) of a contiguous image stack."""
        torch = _torch()
        dm = self.dm
        P = X.shape[0]
        G = torch.empty_like(X)
        obs = torch.empty((P, _lib.MX_OBS_LEN), dtype=torch.float64, device="cuda")
        check(dm.lib.mx_field_images(dm.h, ctypes.c_void_p(X.data_ptr()), P, ctypes.c_void_p(G.data_ptr()),
                                     _lib.MX_FIELD_GRADIENT, _stream()))
        check(dm.lib.mx_observables_images(dm.h, ctypes.c_void_p(X.data_ptr()), P, ctypes.c_void_p(obs.data_ptr()),
                                           _stream()))
        return obs[:, 0], G


class MicromagAdapter:
    def __init__(self, *a, **k):
        raise NotImplementedError("micromagnetic GNEB is not on the device path")


# ----------------------------------------------------------------------------- chain algebra (device)
def _tproject(x, v):
    return v - (v * x).sum(-1, keepdim=True) * x


def _tgeodesic(a, b):
    """Geodesic distance between corresponding images of two stacks (P, N, 3) -> (P,)."""
    torch = _torch()
    c = (a * b).sum(-1).clamp(-1.0, 1.0)
    s = torch.linalg.cross(a, b, dim=-1).norm(dim=-1)
    return torch.atan2(s, c).pow(2).sum(-1).sqrt()


def minimize(adapter, x, tol=1e-6, max_iter=20000):
    """Barzilai-Borwein descent on the spheres; tol on max |grad_perp| (magex/gneb.py:121-138)."""
    dm = adapter.dm
    X = as_device(np.asarray(x, float), dm.dtype)[None]
    X = (X / X.norm(dim=-1, keepdim=True)).contiguous()
    _, G = adapter.images_energy_grad(X)
    g = _tproject(X, G)
    tau = 1e-3 / max(float(g.abs().max()), 1e-30)
    gmax = 0.0
    it = 0
    for it in range(max_iter):
        gmax = float(g.norm(dim=-1).max())
        if gmax < tol:
            break
        Xn = X - tau * g
        Xn = (Xn / Xn.norm(dim=-1, keepdim=True)).contiguous()
        _, Gn = adapter.images_energy_grad(Xn)
        gn = _tproject(Xn, Gn)
        s = (Xn - X).reshape(-1)
        y = (gn - g).reshape(-1)
        sy = float(s @ y)
        tau = (float(s @ s) / sy if it % 2 else sy / float(y @ y)) if sy > 0 else 1e-3 / max(gmax, 1e-30)
        X, g = Xn, gn
    return X[0].double().cpu().numpy(), it, gmax


class GNEB:
    """magex's GNEB (magex/gneb.py:141-237) with the chain held and advanced on the device."""

    def __init__(self, adapter, images, spring=1.0, climbing=True):
        torch = _torch()
        self.ad = adapter
        dm = adapter.dm
        X = torch.as_tensor(np.stack([np.asarray(im, float) for im in images]), dtype=dm.dtype, device="cuda")
        self.X = (X / X.norm(dim=-1, keepdim=True)).contiguous()
        _, G = adapter.images_energy_grad(self.X)
        gp = _tproject(self.X, G)[1:-1]
        gref = gp.norm(dim=-1).amax(dim=-1).double().cpu().numpy() if len(images) > 2 else np.zeros(0)
        self.k = float(spring) * max(float(np.median(gref)) if len(gref) else 1.0, 1e-300)
        self.climbing = climbing
        self.V = torch.zeros_like(self.X)

    @property
    def x(self):
        return [im for im in self.X.double().cpu().numpy()]

    def _tangents(self, E):
        """Upwind tangents of the interior images (magex/gneb.py:152-169), normalised per image."""
        X = self.X
        tp = X[2:] - X[1:-1]
        tm = X[1:-1] - X[:-2]
        Ep, E0, Em = E[2:], E[1:-1], E[:-2]
        wp = np.empty(len(E0))
        wm = np.empty(len(E0))
        for v in range(len(E0)):
            if Ep[v] > E0[v] > Em[v]:
                wp[v], wm[v] = 1.0, 0.0
            elif Ep[v] < E0[v] < Em[v]:
                wp[v], wm[v] = 0.0, 1.0
            else:
                dmax = max(abs(Ep[v] - E0[v]), abs(Em[v] - E0[v]))
                dmin = min(abs(Ep[v] - E0[v]), abs(Em[v] - E0[v]))
                wp[v], wm[v] = (dmax, dmin) if Ep[v] > Em[v] else (dmin, dmax)
        torch = _torch()
        wpt = torch.as_tensor(wp, dtype=X.dtype, device="cuda")[:, None, None]
        wmt = torch.as_tensor(wm, dtype=X.dtype, device="cuda")[:, None, None]
        tau = _tproject(X[1:-1], wpt * tp + wmt * tm)
        n = tau.reshape(len(E0), -1).norm(dim=-1).clamp_min(1e-30)
        return tau / n[:, None, None]

    def forces(self):
        """(E (P,), F (P-2, N, 3) device tensor of the interior images, climbing index)."""
        torch = _torch()
        Et, G = self.ad.images_energy_grad(self.X)
        E = Et.cpu().numpy()
        P = len(E)
        if P <= 2:
            return E, torch.zeros_like(self.X[1:-1]), -1
        t = self._tangents(E)
        ci = int(np.argmax(E[1:-1])) + 1 if self.climbing else -1
        g = _tproject(self.X[1:-1], G[1:-1])
        gt = (g * t).reshape(P - 2, -1).sum(-1)
        Lp = _tgeodesic(self.X[2:], self.X[1:-1])
        Lm = _tgeodesic(self.X[1:-1], self.X[:-2])
        coef = gt + self.k * (Lp - Lm)
        if ci > 0:
            coef[ci - 1] = 2 * gt[ci - 1]
        F = -g + coef.to(g.dtype)[:, None, None] * t
        return E, F, ci

    def run(self, tol=0.0, max_iter=20000, rel_tol=1e-4, log=None, log_every=500, max_step=0.02):
        """Stop when max |F| < tol or < rel_tol x the initial max |F| (magex/gneb.py:189-223)."""
        t0 = time.time()
        E, F, ci = self.forces()
        fmaxes = lambda F: F.norm(dim=-1).amax(dim=-1)                        # noqa: E731
        fmax0 = float(fmaxes(F).max()) if F.shape[0] else 0.0
        scale = 1.0 / max(fmax0, 1e-30)
        it = 0
        for it in range(1, max_iter + 1):
            fmax = float(fmaxes(F).max())
            if fmax < tol or fmax < rel_tol * fmax0:
                break
            f = F * scale
            V = self.V[1:-1]
            fn2 = (f * f).reshape(len(f), -1).sum(-1)
            vp = (V * f).reshape(len(f), -1).sum(-1)
            c = (vp / fn2.clamp_min(1e-300)) * ((vp > 0) & (fn2 > 0))
            V = c[:, None, None] * f + 0.05 * f
            disp = 0.05 * V
            dmax = disp.norm(dim=-1).amax(dim=-1)
            r = (max_step / dmax.clamp_min(1e-300)).clamp(max=1.0)
            r = r.where(dmax > max_step, r.new_ones(()))
            disp = disp * r[:, None, None]
            V = V * r[:, None, None]
            Xn = self.X[1:-1] + disp
            Xn = Xn / Xn.norm(dim=-1, keepdim=True)
            self.X[1:-1] = Xn
            self.V[1:-1] = _tproject(Xn, V)
            E, F, ci = self.forces()
            if log and it % log_every == 0:
                log(f"    chain step {it}: largest image force {fmax:.3e}, current barrier "
                    f"{E.max() - E[0]:.6g} {self.ad.unit}")
        # what result() reports (the attributes magex's GNEB leaves after run)
        self.E, self.climb, self.iterations = E, ci, it
        self.fmax = float(fmaxes(F).max()) if F.shape[0] else 0.0
        self.seconds = time.time() - t0
        if log:
            log(f"    chain relaxed in {it} steps ({self.seconds:.1f} s), largest image force {self.fmax:.2e}")
        return self

    def reaction_coordinate(self):
        L = _tgeodesic(self.X[1:], self.X[:-1]).double().cpu().numpy()
        return np.concatenate([[0.0], np.cumsum(L)])

    def result(self):
        """The path summary in magex's format (magex/gneb.py:231-237): reaction coordinate s, energies
        and energies relative to the first image, barriers seen from either end, saddle and climbing
        images, unit, iterations and the final largest image force."""
        E = np.array(self.E, float)
        top = int(E.argmax())
        out = {"s": self.reaction_coordinate().tolist(), "E": E.tolist(), "E_rel": (E - E[0]).tolist()}
        out.update(barrier_forward=float(E[top] - E[0]), barrier_backward=float(E[top] - E[-1]),
                   saddle_image=top, climbing_image=int(self.climb), unit=self.ad.unit,
                   iterations=int(self.iterations), max_force=float(self.fmax))
        return out


def arrhenius(barrier, unit_kB, T, tau0=1e-9):
    """Arrhenius lifetime tau0 exp(barrier / k_B T) in seconds (magex/gneb.py:240-245); infinite at T = 0
    and where the exponent passes 700."""
    x = barrier / (unit_kB * T) if T > 0 else np.inf
    return float(np.exp(np.log(tau0) + x)) if x < 700 else float("inf")


def skyrmion_on_lattice(system, center=None, radius=10.0, core=-1, helicity_deg=0.0):
    """Skyrmion in a +z background on the atomistic positions (radius in Angstrom; core -1 points down;
    helicity 0 Neel, 90 Bloch), the profile of magex/gneb.py:270-281: tan(theta / 2) = sinh(2) /
    sinh(2 R / radius), written algebraically as cos theta = (s^2 - s0^2) / (s^2 + s0^2) and
    sin theta = 2 s s0 / (s^2 + s0^2) with s = sinh(2 R / radius), s0 = sinh(2)."""
    P = system.positions[:, :2]
    r = P - (P.mean(0) if center is None else np.asarray(center, float)[:2])
    R = np.hypot(r[:, 0], r[:, 1])
    s, s0 = np.sinh(2.0 * R / radius), np.sinh(2.0)
    den = s * s + s0 * s0
    cos_t, sin_t = (s * s - s0 * s0) / den, 2.0 * s * s0 / den
    rhat = np.where(R[:, None] > 0, r / np.where(R > 0, R, 1.0)[:, None], [1.0, 0.0])
    h = np.deg2rad(helicity_deg)
    inplane = np.stack([np.cos(h) * rhat[:, 0] - np.sin(h) * rhat[:, 1],
                        np.sin(h) * rhat[:, 0] + np.cos(h) * rhat[:, 1]], 1)
    m = np.column_stack([sin_t[:, None] * inplane, -core * cos_t])
    return m / np.linalg.norm(m, axis=1, keepdims=True)
