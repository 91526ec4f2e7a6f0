"""Atomistic stochastic Landau-Lifshitz-Gilbert dynamics on the GPU (drop-in for magex.llg).

    de_i/dt = -gamma/(1+alpha^2) [ e_i x B_i + alpha e_i x (e_i x B_i) ]

B_i = B_eff,i + b_i with the Brown thermal field <b b> = 2 alpha k_B T / (gamma mu_i mu_B dt) per
component and step, held over both stages of the stochastic Heun scheme, with |e_i| = 1 restored after
each stage (magex/llg.py:1-9, 216-227).  Units: ps, T, meV, K.

Every step runs in the HIP kernels of libmxllg.so (two fused passes: noise + field + predictor, then
field + corrector).  The noise is not numpy's stream: it is a counter-based generator keyed by (seed,
step, site), so a run is bitwise reproducible and resumable from (spins, step count).
"""
from __future__ import annotations

import time

import numpy as np

from .constants import GAMMA, K_B, MU_B  # noqa: F401
from .spinmodel import normalize  # noqa: F401

INTEGRATORS = ("heun", "rk4", "depondt", "sib", "euler")
DEVICE_INTEGRATORS = ("heun", "depondt")


def llg_rhs(e, B, alpha, precession=True, precision="fp64"):
    """de/dt (1/ps) for unit spins e (N,3) in fields B (N,3) (magex/llg.py:21-28), on the device.

    alpha: scalar.  numpy inputs give numpy output; device tensors stay on the device.
    """
    from .device import DeviceModel, as_device, like_input
    from .models import make_model
    from .spinmodel import SpinSystem, load_results
    if np.ndim(alpha):
        raise NotImplementedError("llg_rhs on the device takes a scalar alpha")
    n = int(np.shape(e)[0])
    key = (n, precision)
    dm = _RHS_MODELS.get(key)
    if dm is None:
        sys_ = SpinSystem(load_results(make_model("chain", J=[0.0])), (n, 1, 1), (False, False, False))
        dm = _RHS_MODELS[key] = DeviceModel(sys_, precision)
    ed, Bd = as_device(e, dm.dtype), as_device(B, dm.dtype)
    return like_input(dm.rhs(ed, Bd, float(alpha), precession), e)


_RHS_MODELS = {}


class LLGSolver:
    """Stochastic Heun LLG of a SpinSystem on the GPU (magex/llg.py:48-283).

    Same constructor as magex's LLGSolver plus ``precision`` ("fp64", magex's arithmetic, or "fp32") and
    ``noise``, the grade of the thermal generator ("double": 53-bit uniforms and double Box-Muller, the
    FP64 default; "float": 24-bit uniforms and float Box-Muller, the FP32 default; an FP64 solver with
    noise="float" draws exactly the FP32 solver's noise).
    backend: "auto", "hip" or "numba" (all mean the device kernels); "numpy" raises: this package
    has no CPU path.  On the device path: the Heun and Depondt integrators, the classical thermostat,
    spin-orbit torque fields (sot), ``field_func(t) -> (3,) or (N, 3)`` evaluated at t in the predictor
    and t + dt in the corrector, and ``extra_field(e) -> (N, 3)`` evaluated at the spins of each stage,
    as magex's fast path does (magex/llg.py:152-166, 197-212).  extra_field receives a NumPy (N, 3) array,
    as in magex; with ``extra_field_device = True`` it receives the device tensor instead and may return a
    tensor.  ValueError for an unknown integrator; NotImplementedError for rk4 / sib / euler, STT and the
    quantum thermostat.
    """

    def __init__(self, system, alpha=0.1, dt=1e-3, temperature=0.0, seed=None, precession=True,
                 thermostat="classical", integrator="heun", stt=None, sot=None, backend="auto", precision="fp64",
                 noise=None):
        self.sys = system
        a = getattr(system, "alpha_site", None)
        self.alpha = np.asarray(a if (a is not None and np.isscalar(alpha)) else alpha, float)
        if self.alpha.ndim and a is not None and np.isscalar(alpha):
            self.alpha = np.where(np.isnan(self.alpha), float(alpha), self.alpha)
        self.dt = float(dt)
        self.T = float(temperature)
        self.rng = np.random.default_rng(seed)
        self.seed = seed
        self.precession = bool(precession)
        self.thermostat = thermostat
        if integrator not in INTEGRATORS:
            raise ValueError(f"integrator must be one of {INTEGRATORS}")
        if integrator not in DEVICE_INTEGRATORS:
            raise NotImplementedError(f"integrator '{integrator}' is not on the device path (heun and depondt are)")
        if thermostat != "classical":
            raise NotImplementedError("only the classical (white-noise) thermostat is on the device path")
        if stt:
            raise NotImplementedError("spin-transfer torque (a lattice-gradient term) is not on the device path")
        self.sot = None
        if sot:                                  # magex/llg.py:83-86
            p = np.asarray(sot.get("p", [0, 1, 0]), float)
            self.sot = (p / np.linalg.norm(p), float(sot.get("B_DL", 0.0)), float(sot.get("B_FL", 0.0)))
        if backend not in ("auto", "hip", "numba"):
            if backend == "numpy":
                raise RuntimeError("magex_amd has no NumPy path: use backend='auto' (HIP kernels)")
            raise ValueError(f"unknown backend '{backend}'")
        self.integrator = integrator
        self.precision = precision
        if noise not in (None, "float", "double"):
            raise ValueError("noise must be 'float' or 'double'")
        self.noise = noise or ("double" if precision == "fp64" else "float")
        self.time = 0.0
        self.field_func = None
        self.extra_field = None
        self.extra_field_device = False
        self.stt = None
        # the noise key: the seed itself, or 64 bits drawn from fresh entropy
        self.noise_seed = int(seed) if seed is not None else int(self.rng.integers(0, 2**63))
        self.step_count = 0
        self.fast = self.dm = system.device_model(precision)

    # ------------------------------------------------------------------ device plumbing
    def _sync(self):
        self.dm = self.sys.device_model(self.precision)
        self.dm.set_noise(self.noise)
        p, bdl, bfl = self.sot if self.sot is not None else ((0.0, 1.0, 0.0), 0.0, 0.0)
        self.dm.set_sot(p, bdl, bfl)
        self.fast = self.dm
        from .structure import site_temperature
        T = site_temperature(self.sys, self.T)
        self.dm.set_dynamics(self.dt, self.alpha, T, self.precession, self.integrator)
        return self.dm

    def _to_device(self, e):
        from .device import as_device
        return as_device(e, self.dm.dtype)

    @property
    def sigma(self):
        """Thermal-field amplitude per site, sqrt(2 alpha k_B T / (gamma mu mu_B dt)) (llg.py:111)."""
        from .structure import site_temperature
        T = site_temperature(self.sys, self.T)
        s = np.sqrt(2.0 * self.alpha * K_B * np.asarray(T, float) / (GAMMA * self.sys.mu * MU_B * self.dt))
        return np.broadcast_to(s, (self.sys.N,))

    def thermal_field(self):
        """The thermal field b (N,3) of the next step (T), as the kernels will draw it."""
        if self.T <= 0 and getattr(self.sys, "site_T", None) is None:
            return 0.0
        dm = self._sync()
        b = self.sigma[:, None] * dm.noise(self.noise_seed, self.step_count).cpu().numpy()
        if self.sys.pinned is not None:
            b[self.sys.pinned] = 0.0
        return b

    def _stage_fields(self, x, t):
        """Uniform field (3,) and per-site extra field (device (N, 3) or None) of the stage evaluated at
        spins x (device) and time t: field_func(t) (uniform or per site) plus extra_field(x)
        (magex/llg.py:152-166)."""
        from .device import as_device
        B = np.asarray(self.sys.B, float).reshape(3).copy()
        X = None
        if self.field_func is not None:
            f = np.asarray(self.field_func(t), float)
            if f.size == 3:
                B += f.reshape(3)
            else:
                X = as_device(np.array(np.broadcast_to(f, (self.sys.N, 3))), self.dm.dtype)
        if self.extra_field is not None:
            arg = x if self.extra_field_device else x.cpu().numpy().astype(float)
            xe = as_device(self.extra_field(arg), self.dm.dtype)
            if tuple(xe.shape) != (self.sys.N, 3):
                raise ValueError(f"extra_field must return an ({self.sys.N}, 3) field")
            X = xe if X is None else X + xe
        return B, (None if X is None else X.contiguous())

    def step_device(self, e_dev, nsteps=1, noise=None):
        """Advance a device tensor in place by nsteps (the CudaLLG.step_device shape,
        magex/cuda_backend.py:362-369).  noise: injected (nsteps, N, 3) standard normals or None."""
        dm = self._sync()
        site_ff = False
        if self.field_func is not None and self.extra_field is None:
            site_ff = np.asarray(self.field_func(self.time), float).size != 3
        if self.extra_field is not None or site_ff:
            # fields that depend on the spins or on the site: evaluated on the host between the stages
            if noise is not None:
                raise NotImplementedError("injected noise together with extra_field or a per-site field_func")
            import torch
            ep = torch.empty_like(e_dev)
            for n in range(int(nsteps)):
                t = self.time + n * self.dt
                step = self.step_count + n
                B1, X1 = self._stage_fields(e_dev, t)
                dm.stage(e_dev, 1, B1, X1, self.noise_seed, step, ep)
                B2, X2 = self._stage_fields(ep, t + self.dt)
                dm.stage(e_dev, 2, B2, X2, self.noise_seed, step)
        elif self.field_func is not None:
            if noise is not None:
                raise NotImplementedError("injected noise together with field_func")
            t = self.time + self.dt * np.arange(int(nsteps))
            B = np.empty((int(nsteps), 2, 3))
            for n, tn in enumerate(t):
                for k, tk in enumerate((tn, tn + self.dt)):
                    B[n, k] = self.sys.B + np.asarray(self.field_func(tk), float).reshape(3)
            dm.run_driven(e_dev, B, self.noise_seed, self.step_count)
        else:
            dm.run(e_dev, nsteps, self.noise_seed, self.step_count, noise)
        self.step_count += int(nsteps)
        self.time += nsteps * self.dt
        return e_dev

    def steps_checked(self, e_dev, nsteps):
        """step_device(e_dev, nsteps), then, if a persistent launch gave up a wait (its spins part-way),
        the same steps again from the same spins one launch per step (DeviceModel.checked): the result
        of an undisturbed run.  Synchronises."""
        dm = self._sync()
        s0, t0, keep = self.step_count, self.time, e_dev.clone()

        def undo():
            e_dev.copy_(keep)
            self.step_count, self.time = s0, t0

        dm.checked(lambda: self.step_device(e_dev, nsteps), undo)
        return e_dev

    def field(self, e, t=None):
        """B_eff + spin-orbit torque fields + field_func(t) + extra_field(e) (T), zero on pinned sites
        (magex/llg.py:124-136); numpy in -> numpy out, device tensors stay on the device.  t defaults to
        the solver's time."""
        from . import _lib
        from .device import as_device, like_input
        dm = self._sync()
        ed = self._to_device(e)
        B = dm.field(ed, _lib.MX_FIELD_BEFF)          # B_eff from the kernels, zero on pinned sites
        extra = self.sot is not None or self.field_func is not None or self.extra_field is not None
        if self.sot is not None:
            p, bdl, bfl = self.sot
            pt = B.new_tensor(p)
            B = B + bdl * ed.cross(pt.expand_as(ed), dim=1) + bfl * pt
        if self.field_func is not None:
            f = np.asarray(self.field_func(self.time if t is None else t), float)
            B = B + B.new_tensor(f.reshape(3) if f.size == 3 else np.array(np.broadcast_to(f, (self.sys.N, 3))))
        if self.extra_field is not None:
            arg = ed if self.extra_field_device else ed.cpu().numpy().astype(float)
            B = B + as_device(self.extra_field(arg), B.dtype)
        if extra and self.sys.pinned is not None:
            B[_pinned_mask(self.sys, B)] = 0.0
        return like_input(B, e)

    def rhs(self, e, b=0.0, t=None):
        """Tangent de/dt (1/ps) in the field field(e, t) + b (magex/llg.py:138-150); b a thermal field
        (N,3) or 0.  Same input/output convention as field()."""
        from .device import as_device, like_input
        dm = self._sync()
        ed = self._to_device(e)
        B = self.field(ed, t)
        if not np.isscalar(b) or b != 0.0:
            B = B + as_device(b, dm.dtype)
        a = self.alpha
        if np.ndim(a) == 0:
            return like_input(dm.rhs(ed, B.contiguous(), float(a), self.precession), e)
        # per-site damping: the same expression with alpha per row
        at = B.new_tensor(np.asarray(a, float))[:, None]
        exB = ed.cross(B, dim=1)
        f = at * ed.cross(exB, dim=1)
        if self.precession:
            f = f + exB
        return like_input(-GAMMA / (1.0 + at * at) * f, e)

    def step(self, e):
        """One integration step; returns new spins (numpy in -> numpy out, the input is not mutated)."""
        from .device import like_input
        self._sync()
        ed = self._to_device(e)
        if ed is e:
            ed = ed.clone()
        self.step_device(ed, 1)
        return like_input(ed, e)

    def max_torque(self, e):
        """max_i |e_i x B_eff,i| (T) (magex/llg.py:250-252)."""
        dm = self._sync()
        return float(dm.observables(self._to_device(e))[7])

    def run(self, e, steps, output_every=100, snapshots=0, callback=None, torque_tol=None, log=None):
        """Integrate steps steps (magex/llg.py:254-283); returns (spins, table, snapshots).

        table columns: t (ps), E/N (meV), <ex>, <ey>, <ez>, order parameter, max torque (T).
        Between output points the steps run back to back on the device without host round trips.
        """
        from .device import like_input
        dm = self._sync()
        ed = self._to_device(e)
        if ed is e:
            ed = ed.clone()
        N = self.sys.N
        rows, snaps = [], []
        snap_every = max(1, steps // snapshots) if snapshots else 0
        t0 = time.time()
        marks = set(range(0, steps + 1, max(1, output_every))) | {steps}
        if snap_every:
            marks |= set(range(0, steps + 1, snap_every))
        if callback:
            marks = set(range(steps + 1))
        order = sorted(marks)
        n = 0
        for m in order:
            if m > n:
                self.steps_checked(ed, m - n)
                n = m
            if n % output_every == 0 or n == steps:
                o = dm.observables(ed)
                mvec = o[1:4] / N
                row = [self.time, o[0] / N, *mvec, float(np.linalg.norm(o[4:7]) / N), float(o[7])]
                rows.append(row)
                if log and (n % (output_every * 10) == 0 or n == steps):
                    log(f"  step {n:8d}  t = {self.time:9.3f} ps  E/N = {row[1]: .6f} meV  "
                        f"m = ({mvec[0]: .3f},{mvec[1]: .3f},{mvec[2]: .3f})  |torque|max = {row[6]:.2e} T")
                if torque_tol is not None and self.T == 0 and row[6] < torque_tol and n > 0:
                    if log:
                        log(f"  converged: max torque {row[6]:.2e} T < {torque_tol:g} T after {n} steps")
                    break
            if snap_every and n % snap_every == 0:
                snaps.append((self.time, ed.cpu().numpy().astype(float)))
            if callback:
                callback(n, like_input(ed, e))
        if log:
            log(f"  LLG wall time {time.time() - t0:.1f} s")
        return like_input(ed, e), np.array(rows), snaps


def _pinned_mask(system, like):
    import torch
    return torch.as_tensor(np.asarray(system.pinned, bool), device=like.device)


def save_spins(path, system, e):
    """spins_*.npz + .csv writer (magex/llg.py:292-296)."""
    e = e.cpu().numpy() if hasattr(e, "cpu") else np.asarray(e)
    np.savez_compressed(path, spins=e, positions=system.positions, mu=system.mu, sublattice=system.sublattice,
                        cell=system.uc.cell, supercell=np.array(system.L))
    np.savetxt(path.replace(".npz", ".csv"), np.column_stack([system.positions, e]), delimiter=",", fmt="%.6f",
               header="x_A,y_A,z_A,ex,ey,ez", comments="")
