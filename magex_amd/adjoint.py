"""Differentiable LLG: gradients of a Heun or Depondt trajectory with respect to the initial spins and
to the constants of the Hamiltonian, through the library's hand-written adjoint.

No counterpart in magex: its solvers (magex/llg.py:87-120) only run forward.  Here a Hamiltonian whose
couplings and on-site tensors are linear in a few constants theta (exchange shells, DMI, anisotropy)
is written J_k(theta) = sum_p theta_p U_p[k] per stencil entry and A_s(theta) = sum_p theta_p Aun_p[s];
the units come from a builder ``build(theta) -> SpinSystem`` (for example ``configs.square128``) by
differencing, and are checked to reproduce the builder's couplings.

    sol = DifferentiableLLG(lambda th: configs.square128(*th)[0], theta0, alpha=0.3, dt=1e-3)
    theta = torch.tensor(theta0, requires_grad=True)
    e = sol.trajectory(theta, e0, 500)           # autograd-aware
    loss = ((e - target) ** 2).mean(); loss.backward()

The backward replays each checkpoint interval (default 50 steps) and runs the adjoint kernels on the
device; see mxllg_adjoint.hip for the equations.  The builder's systems may carry per-bond tensors and
per-site on-site tensors (a disordered model) as long as theta moves them uniformly per stencil entry
and sublattice; the rest is held fixed.  ``backward_dense`` also returns the gradient with
respect to every bond's 3x3 tensor and every site's on-site tensor at the current couplings
(``dmi_gradient`` takes the DMI vectors' part of the former).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import check, hptr
from .device import DeviceModel, _stream, _torch

MAXP = 16   # parameters the library differentiates in one backward (mxllg_adjoint.hip)


def linear_units(build, theta_ref, tol=1e-12):
    """Units of a builder linear in theta: (U (P', nent, 3, 3), Aun (P', ns, 3, 3), offset, system, boff, aoff).

    U_p = J(theta_ref + e_p) - J(theta_ref) entry by entry, Aun_p likewise per sublattice.  A
    theta-independent remainder J(theta_ref) - sum_p theta_ref_p U_p, if any, is appended to U and Aun as
    a fixed unit (offset True).  Per-bond tensors (SpinSystem.set_bond_tensors) and per-site on-site
    tensors are allowed when their dependence on theta is uniform per stencil entry / sublattice: what
    remains is a fixed per-bond part boff (nbonds, ncell, 3, 3) and per-site part aoff (N, 3, 3) (None
    when the model has none).  Raises ValueError when the builder's stencil changes with theta or its
    couplings are not linear in theta with such units."""
    theta_ref = np.asarray(theta_ref, float)
    P = len(theta_ref)
    base = build(theta_ref)
    key = [(s, t, tuple(R)) for s, t, R, _ in base.stencil]
    ns = base.uc.nsites
    fw = base.forward_entries()

    def arrays(sys_):
        """(stencil tensors (nent, 3, 3), per-bond tensors or None, per-site A (ncell, ns, 3, 3))."""
        if [(s, t, tuple(R)) for s, t, R, _ in sys_.stencil] != key:
            raise ValueError("the builder's stencil changes with the parameters (keep every shell nonzero at the "
                             "reference point)")
        Js = np.array([J for *_, J in sys_.stencil], float).reshape(-1, 3, 3)
        Jb = None if sys_.bond_tensors is None else np.asarray(sys_.bond_tensors, float)
        return Js, Jb, np.asarray(sys_.A, float).reshape(sys_.ncell, ns, 3, 3)

    Js0, Jb0, A0 = arrays(base)
    disordered_J = Jb0 is not None
    disordered_A = not np.all(A0 == A0[:1])
    U = np.empty((P,) + Js0.shape)
    Aun = np.empty((P, ns, 3, 3))
    scale = max(np.abs(Js0).max(initial=0.0), np.abs(A0).max(initial=0.0),
                np.abs(Jb0).max(initial=0.0) if disordered_J else 0.0, 1e-300)
    for p in range(P):
        th = theta_ref.copy()
        th[p] += 1.0
        Js1, Jb1, A1 = arrays(build(th))
        if (Jb1 is not None) != disordered_J:
            raise ValueError("the builder sets per-bond tensors at some parameters only")
        U[p] = Js1 - Js0
        dA = A1 - A0
        Aun[p] = dA[0]
        if np.abs(dA - dA[:1]).max(initial=0.0) > tol * scale:
            raise ValueError("the on-site tensors depend on the parameters per site (only per sublattice is "
                             "parameterized)")
        if disordered_J:
            dJ = Jb1 - Jb0
            if np.abs(dJ - U[p][fw][:, None]).max(initial=0.0) > tol * scale:
                raise ValueError("the per-bond tensors must follow the parameters as the stencil does (uniformly "
                                 "per entry)")
    offJ = Js0 - np.tensordot(theta_ref, U, 1)
    offA = A0[0] - np.tensordot(theta_ref, Aun, 1)
    offset = None
    if np.abs(offJ).max(initial=0.0) > tol * scale or np.abs(offA).max(initial=0.0) > tol * scale:
        offset = (offJ, offA)
    Uf = np.concatenate([U, offJ[None]]) if offset else U
    Af = np.concatenate([Aun, offA[None]]) if offset else Aun
    thf = np.append(theta_ref, 1.0) if offset else theta_ref
    boff = Jb0 - np.tensordot(thf, Uf, 1)[fw][:, None] if disordered_J else None
    aoff = (A0 - np.tensordot(thf, Af, 1)[None]).reshape(-1, 3, 3) if disordered_A else None

    # linearity: a second point must be reproduced by the units and fixed parts
    th = theta_ref * 0.5 + 0.25
    thf2 = np.append(th, 1.0) if offset else th
    Js2, Jb2, A2 = arrays(build(th))
    pJ = np.tensordot(thf2, Uf, 1)
    pA = np.tensordot(thf2, Af, 1)[None] + (0.0 if aoff is None else aoff.reshape(A0.shape))
    bad = np.abs(pJ - Js2).max(initial=0.0) > 1e-10 * scale or np.abs(pA - A2).max(initial=0.0) > 1e-10 * scale
    if disordered_J:
        bad = bad or np.abs(boff + pJ[fw][:, None] - Jb2).max(initial=0.0) > 1e-10 * scale
    if bad:
        raise ValueError("the builder's couplings are not linear in the parameters")
    return Uf, Af, offset is not None, base, boff, aoff


class DifferentiableLLG:
    """Deterministic-noise Heun (or Depondt) trajectories of a parameterized SpinSystem with adjoint gradients.

    build(theta) -> SpinSystem defines the Hamiltonian (its field, moments and stencil are taken at
    theta0; couplings and on-site tensors must be linear in theta).  alpha, dt (ps), temperature (K) and
    seed as LLGSolver; the thermal field is a fixed realization (the library's Philox stream), so the
    gradient is that of one sample path.  precision "fp64" or "fp32"; layout as DeviceModel; sot as
    LLGSolver's ({"p": direction, "B_DL": T, "B_FL": T}, magex/llg.py:83-86), held fixed; integrator
    "heun" or "depondt" (magex/llg.py:216-240)."""

    def __init__(self, build, theta0, alpha, dt, temperature=0.0, seed=0, precision="fp64", checkpoint_every=50,
                 precession=True, layout=None, sot=None, integrator="heun"):
        _torch()
        U, Aun, offset, system, boff, aoff = linear_units(build, theta0)
        self.P = len(theta0)
        if len(U) > MAXP:
            raise ValueError(f"at most {MAXP} parameters (including a fixed offset)")
        self.offset = offset
        self.system = system
        self.precision = precision
        self.dtype = torch.float64 if precision == "fp64" else torch.float32
        self.N = system.N
        self.seed = int(seed)
        self.dm = DeviceModel(system, precision, layout=layout)
        self.dm.set_dynamics(dt, alpha, temperature, precession, integrator)
        self.dt = float(dt)
        self._nsteps = 0
        if sot:
            p = np.asarray(sot.get("p", [0, 1, 0]), float)
            self.dm.set_sot(p / np.linalg.norm(p), float(sot.get("B_DL", 0.0)), float(sot.get("B_FL", 0.0)))
        self.checkpoint_every = int(checkpoint_every)
        L = self.dm.lib
        h = ctypes.c_void_p()
        self._U = np.ascontiguousarray(U.reshape(len(U), -1, 9))
        self._Aun = np.ascontiguousarray(Aun.reshape(len(Aun), -1, 9))
        check(L.mx_adjoint_create(self.dm.h, len(U), hptr(self._U), hptr(self._Aun), self.checkpoint_every,
                                  ctypes.byref(h)))
        self.h = h
        if boff is not None or aoff is not None:
            self._boff = None if boff is None else np.ascontiguousarray(boff.reshape(-1))
            self._aoff = None if aoff is None else np.ascontiguousarray(aoff.reshape(-1))
            check(L.mx_adjoint_set_offsets(h, hptr(self._boff) if boff is not None else None,
                                           hptr(self._aoff) if aoff is not None else None))
        self._theta = None
        self._token = 0          # id of the forward the adjoint record belongs to

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self.dm.lib.mx_adjoint_destroy(h)
            self.h = None

    def set_params(self, theta):
        th = np.asarray(theta, float).reshape(-1)
        if th.size != self.P:
            raise ValueError(f"expected {self.P} parameters")
        full = np.append(th, 1.0) if self.offset else th
        key = full.tobytes()
        if key != self._theta:
            check(self.dm.lib.mx_adjoint_set_params(self.h, hptr(np.ascontiguousarray(full))))
            self._theta = key

    def _spins(self, e):
        if not isinstance(e, torch.Tensor):
            e = torch.as_tensor(np.asarray(e, float))
        e = e.detach().to(device="cuda", dtype=self.dtype).contiguous()
        if e.shape != (self.N, 3):
            raise ValueError(f"spins must be ({self.N}, 3)")
        return e.clone()

    def _run(self, work, undo, checked):
        # checked: a persistent launch that gave up a wait is run again one launch per step (synchronises)
        if checked:
            self.dm.checked(work, undo)
        else:
            work()

    def forward(self, theta, e0, nsteps, step0=0, field=None, checked=True):
        """Spins after nsteps steps from e0 (new device tensor); records checkpoints for backward().
        field: the uniform field (T) of the predictor and corrector pass of each step, (nsteps, 2, 3)
        (field_sequence builds it from a field_func), or None for the system's own field.
        checked=False leaves the work queued without the give-up check (and its synchronisation): the
        caller reads self.dm.step_info()[1] itself."""
        self.set_params(theta)
        e = self._spins(e0)
        start = e.clone()
        L = self.dm.lib
        B = None if field is None else np.ascontiguousarray(np.asarray(field, float).reshape(int(nsteps), 2, 3))

        def work():
            if B is None:
                check(L.mx_adjoint_forward(self.h, ctypes.c_void_p(e.data_ptr()), int(nsteps), self.seed,
                                           int(step0), _stream()))
            else:
                check(L.mx_adjoint_forward_driven(self.h, ctypes.c_void_p(e.data_ptr()), int(nsteps), hptr(B),
                                                  self.seed, int(step0), _stream()))

        # a persistent launch that gave up a wait (its spins part-way) is run again one launch per step
        self._run(work, lambda: e.copy_(start), checked)
        self._nsteps = int(nsteps)
        self._token += 1
        return e

    def field_sequence(self, field_func, nsteps, t0=0.0):
        """(nsteps, 2, 3) total fields B + field_func(t) at t = t0 + n dt and t + dt, as LLGSolver drives
        them (magex/llg.py:129-130)."""
        out = np.empty((int(nsteps), 2, 3))
        for n in range(int(nsteps)):
            t = t0 + n * self.dt
            for k, tk in enumerate((t, t + self.dt)):
                out[n, k] = self.system.B + np.asarray(field_func(tk), float).reshape(3)
        return out

    def backward(self, grad_e, want_e0=True, checked=True):
        """(dL/dtheta (P,) float64 device tensor, dL/de0 or None) for the last forward().

        Runs on torch's current stream; on grids the GPU holds twice over at once, each checkpoint segment
        is replayed on a second, internal stream beside the previous segment's sweep, ordered against the
        current stream with events (so the results are the current stream's, as for any other call)."""
        g = grad_e.detach().to(device="cuda", dtype=self.dtype).contiguous()
        gt = torch.empty(MAXP, dtype=torch.float64, device="cuda")
        g0 = torch.empty_like(g) if want_e0 else None
        self._run(lambda: check(self.dm.lib.mx_adjoint_backward(
            self.h, ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(g0.data_ptr() if g0 is not None else 0),
            ctypes.c_void_p(gt.data_ptr()), _stream())), lambda: None, checked)
        return gt[:self.P], g0

    @property
    def nbonds(self):
        """Unordered bonds per cell: the forward stencil entries (SpinSystem.forward_entries)."""
        return int(self.dm.lib.mx_adjoint_num_bonds(self.h))

    def backward_dense(self, grad_e, want_e0=True, checked=True):
        """(dL/dtheta, dL/de0 or None, dL/dT (nbonds, ncell, 3, 3), dL/dA (N, 3, 3)) for the last forward().

        T is each unordered bond's tensor in the layout of SpinSystem.set_bond_tensors (it acts as J_ij = T
        from the forward entry's site and J_ji = T^T back), A each site's on-site tensor; both are
        evaluated at the current couplings.  float64 device tensors."""
        g = grad_e.detach().to(device="cuda", dtype=self.dtype).contiguous()
        gt = torch.empty(MAXP, dtype=torch.float64, device="cuda")
        g0 = torch.empty_like(g) if want_e0 else None
        gb = torch.empty((self.nbonds, self.system.ncell, 3, 3), dtype=torch.float64, device="cuda")
        ga = torch.empty((self.N, 3, 3), dtype=torch.float64, device="cuda")
        self._run(lambda: check(self.dm.lib.mx_adjoint_backward_dense(
            self.h, ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(g0.data_ptr() if g0 is not None else 0),
            ctypes.c_void_p(gt.data_ptr()), ctypes.c_void_p(gb.data_ptr()), ctypes.c_void_p(ga.data_ptr()),
            _stream())), lambda: None, checked)
        return gt[:self.P], g0, gb, ga

    def backward_field(self, grad_e, want_e0=True, checked=True):
        """(dL/dtheta, dL/de0 or None, dL/dB (nsteps, 2, 3)) for the last forward(): dL/dB is the gradient
        with respect to the uniform field of each pass of each step (driven or not), float64 device tensor."""
        g = grad_e.detach().to(device="cuda", dtype=self.dtype).contiguous()
        gt = torch.empty(MAXP, dtype=torch.float64, device="cuda")
        g0 = torch.empty_like(g) if want_e0 else None
        gB = torch.empty((self._nsteps, 2, 3), dtype=torch.float64, device="cuda")
        self._run(lambda: check(self.dm.lib.mx_adjoint_backward_field(
            self.h, ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(g0.data_ptr() if g0 is not None else 0),
            ctypes.c_void_p(gt.data_ptr()), ctypes.c_void_p(gB.data_ptr()), _stream())), lambda: None, checked)
        return gt[:self.P], g0, gB

    def trajectory(self, theta, e0, nsteps, step0=0, field=None):
        """Differentiable e(nsteps) as a function of theta (torch tensor, P), e0 (N, 3) and, when given,
        the driving field (nsteps, 2, 3) of forward()."""
        if not isinstance(theta, torch.Tensor):
            theta = torch.as_tensor(np.asarray(theta, float))
        if not isinstance(e0, torch.Tensor):
            e0 = torch.as_tensor(np.asarray(e0, float))
        if field is not None and not isinstance(field, torch.Tensor):
            field = torch.as_tensor(np.asarray(field, float))
        return _HeunTrajectory.apply(theta, e0, field, self, int(nsteps), int(step0))


class _HeunTrajectory(torch.autograd.Function):
    """e_final = Heun^nsteps(e0; theta, field): forward through mx_adjoint_forward(_driven), backward
    through mx_adjoint_backward(_field) (re-running the forward first if another forward has run since)."""

    @staticmethod
    def forward(ctx, theta, e0, field, sol, nsteps, step0):
        th = theta.detach().cpu().double().numpy()
        fld = None if field is None else field.detach().cpu().double().numpy()
        e = sol.forward(th, e0, nsteps, step0, field=fld)
        ctx.sol, ctx.nsteps, ctx.step0, ctx.theta, ctx.field = sol, nsteps, step0, th, fld
        ctx.token = sol._token
        ctx.e0 = e0.detach().clone()
        ctx.meta = (theta.dtype, theta.device, e0.dtype, e0.device,
                    None if field is None else (field.dtype, field.device))
        return e.to(dtype=e0.dtype, device=e0.device)

    @staticmethod
    def backward(ctx, g):
        sol = ctx.sol
        if sol._token != ctx.token:
            sol.forward(ctx.theta, ctx.e0, ctx.nsteps, ctx.step0, field=ctx.field)
            ctx.token = sol._token
        tdt, tdev, edt, edev, fmeta = ctx.meta
        gB = None
        if fmeta is not None and ctx.needs_input_grad[2]:
            gt, g0, gB = sol.backward_field(g, want_e0=ctx.needs_input_grad[1])
            gB = gB.to(dtype=fmeta[0], device=fmeta[1])
        else:
            gt, g0 = sol.backward(g, want_e0=ctx.needs_input_grad[1])
        gt = gt.to(dtype=tdt, device=tdev)
        if g0 is not None:
            g0 = g0.to(dtype=edt, device=edev)
        return gt, g0, gB, None, None, None


def dmi_gradient(grad_bonds):
    """dL/dD (..., 3) of each bond's DMI vector from dL/dT (..., 3, 3): the DMI part of a bond tensor is
    T_ab = eps_abc D_c (magex/models.py:37), so dL/dD_c = eps_abc dL/dT_ab."""
    G = grad_bonds
    return torch.stack([G[..., 1, 2] - G[..., 2, 1], G[..., 2, 0] - G[..., 0, 2], G[..., 0, 1] - G[..., 1, 0]], -1)
