"""FP64 torch.autograd restatement of magex's Heun trajectory, for gradients (test infrastructure only).

TEST INFRASTRUCTURE ONLY.  Nothing in ``magex_amd`` imports this module; only ``tests/`` uses it, as
the checker of the library's adjoint.  magex has no gradients; this module differentiates its forward
by automatic differentiation: the trajectory of magex_ref.run_heun (magex/llg.py:216-227, the field of
magex/spinmodel.py:454-473) written in torch FP64 on the CPU with the couplings linear in parameters
theta, Jsp(theta) = sum_p theta_p Jsp_p (+ a fixed remainder), A(theta) likewise.

Pinning (tests/test_oracle.py): the forward equals magex_ref.run_heun at the same theta to round-off
(magex_ref is pinned to magex's golden trajectories); the gradient is torch's reverse mode of that
forward, checked against central finite differences of the same forward.  ``run_dense`` is the same
trajectory with the coupling matrix (dense 3N x 3N) and the on-site tensors as free tensors, for
gradients with respect to every bond and site; pinned to ``run`` (same trajectory, and the chain rule
dL/dtheta_p = <Jsp_p, dL/dJ> + <A_p, dL/dA>) and to central differences of single elements.
"""
from __future__ import annotations

import numpy as np
import torch

from . import magex_ref as R


def _coo(M):
    M = M.tocoo()
    idx = torch.as_tensor(np.vstack([M.row, M.col]), dtype=torch.int64)
    return torch.sparse_coo_tensor(idx, torch.as_tensor(M.data, dtype=torch.float64), M.shape).coalesce()


class TorchHeun:
    """build(theta) -> magex_ref.SpinSystem, linear in theta; units by differences at theta_ref."""

    def __init__(self, build, theta_ref, alpha, dt, precession=True, sot=None, integrator="heun"):
        theta_ref = np.asarray(theta_ref, float)
        base = build(theta_ref)
        self.N = base.N
        Jp, Ap = [], []
        for p in range(len(theta_ref)):
            th = theta_ref.copy()
            th[p] += 1.0
            s = build(th)
            Jp.append(s.Jsp - base.Jsp)
            Ap.append(s.A - base.A)
        J0 = base.Jsp - sum(t * M for t, M in zip(theta_ref, Jp))
        A0 = base.A - sum(t * M for t, M in zip(theta_ref, Ap))
        self.Jp = [_coo(M) for M in Jp] + [_coo(J0)]
        self.Ap = torch.as_tensor(np.stack(Ap + [A0]))
        self.mu = torch.as_tensor(base.mu)
        self.B = torch.as_tensor(base.B)
        self.alpha, self.dt, self.precession = float(alpha), float(dt), precession
        self.integrator = integrator
        self.base = base
        # spin-orbit torque fields B_DL (e x p) + B_FL p (magex/llg.py:126-128), sot = (p unit, B_DL, B_FL)
        self.sot = None if sot is None else (torch.as_tensor(np.asarray(sot[0], float)), float(sot[1]), float(sot[2]))
        # pinned spins are carried unchanged (a copy, as the device kernels do; magex renormalises them,
        # which changes nothing but round-off in the value and drops the radial part of the derivative)
        self.pin = None if base.pinned is None else torch.as_tensor(np.asarray(base.pinned, bool))[:, None]

    def terms_gradient(self, e):
        """Gradient of the system's cubic and higher-uniaxial terms (magex_ref.CubicAnisotropy /
        UniaxialHigher, terms.py:35-75), fixed in theta."""
        g = torch.zeros_like(e)
        for t in self.base.terms:
            if isinstance(t, R.CubicAnisotropy):
                C = torch.as_tensor(t.C)
                a = e @ C.T
                q = a * a
                da = 2.0 * a * torch.stack([q[:, 1] + q[:, 2], q[:, 0] + q[:, 2], q[:, 0] + q[:, 1]], 1)
                g = g + torch.as_tensor(np.broadcast_to(t.Kc, (self.N,)).copy())[:, None] * (da @ C)
            elif isinstance(t, R.UniaxialHigher):
                u = torch.as_tensor(t.u)
                c = e @ u
                K2 = torch.as_tensor(np.broadcast_to(t.K2, (self.N,)).copy())
                K3 = torch.as_tensor(np.broadcast_to(t.K3, (self.N,)).copy())
                g = g + (-4.0 * K2 * c**3 - 6.0 * K3 * c**5)[:, None] * u[None, :]
            else:
                raise NotImplementedError(type(t).__name__)
        return g

    def effective_field(self, th, e, B=None):
        """B_eff (spinmodel.py:458-473) with theta th (torch, P) and e (N, 3); B replaces the system's
        uniform field (a driven pass: B + field_func(t), magex/llg.py:129-130)."""
        thf = torch.cat([th, th.new_ones(1)])
        x = e.reshape(-1, 1)
        h = sum(thf[p] * torch.sparse.mm(self.Jp[p], x) for p in range(len(self.Jp))).reshape(-1, 3)
        A = torch.einsum("p,piab->iab", thf, self.Ap)
        Bu = self.B if B is None else B
        g = -2.0 * h - R.MU_B * self.mu[:, None] * Bu[None, :] + 2.0 * torch.einsum("iab,ib->ia", A, e)
        g = g + self.terms_gradient(e)
        return self._sot(-g / (R.MU_B * self.mu[:, None]), e)

    def _sot(self, B, e):
        if self.sot is None:
            return B
        p, bdl, bfl = self.sot
        return B + bdl * torch.cross(e, p.expand_as(e), dim=1) + bfl * p[None, :]

    def rhs(self, e, B):
        a = self.alpha
        pre = -R.GAMMA / (1.0 + a * a)
        exB = torch.cross(e, B, dim=1)
        out = a * torch.cross(e, exB, dim=1)
        if self.precession:
            out = out + exB
        return pre * out

    @staticmethod
    def normalize(u):
        return u / torch.linalg.norm(u, dim=1, keepdim=True)

    def _keep_pinned(self, new, old):
        return new if self.pin is None else torch.where(self.pin, old, new)

    @staticmethod
    def rotate(e, w):
        """Rodrigues rotation of e by the rotation vectors w (magex_ref.rotate, llg.py:31-36)."""
        th = torch.linalg.norm(w, dim=1, keepdim=True)
        big = th > 1e-14
        k = torch.where(big, w / torch.where(big, th, torch.ones_like(th)), torch.zeros_like(w))
        c, s = torch.cos(th), torch.sin(th)
        return e * c + torch.cross(k, e, dim=1) * s + k * (k * e).sum(1, keepdim=True) * (1 - c)

    def _step(self, e, field1, field2, dt):
        """One Heun or Depondt step (magex_ref.heun_step / depondt_step) from the two pass fields."""
        k1 = self.rhs(e, field1(e))
        if self.integrator == "depondt":
            w1 = torch.cross(e, k1, dim=1)
            ep = self._keep_pinned(self.rotate(e, dt * w1), e)
            w2 = torch.cross(ep, self.rhs(ep, field2(ep)), dim=1)
            return self._keep_pinned(self.normalize(self.rotate(e, 0.5 * dt * (w1 + w2))), e)
        ep = self._keep_pinned(self.normalize(e + dt * k1), e)
        k2 = self.rhs(ep, field2(ep))
        return self._keep_pinned(self.normalize(e + 0.5 * dt * (k1 + k2)), e)

    def run(self, th, e, steps, thermal=None, field=None):
        """steps Heun steps; thermal(n) -> (N, 3) thermal field of step n (torch FP64) or None; field
        (steps, 2, 3): the uniform field of each step's predictor and corrector pass, or None."""
        for n in range(steps):
            b = 0.0 if thermal is None else thermal(n)
            B1 = None if field is None else field[n, 0]
            B2 = None if field is None else field[n, 1]
            e = self._step(e, lambda x: self.effective_field(th, x, B1) + b,
                           lambda x: self.effective_field(th, x, B2) + b, self.dt)
        return e

    def dense(self, th):
        """(J (3N, 3N), A (N, 3, 3)) at theta th as dense torch FP64 tensors."""
        thf = torch.cat([torch.as_tensor(th, dtype=torch.float64).reshape(-1), torch.ones(1, dtype=torch.float64)])
        J = sum(thf[p] * self.Jp[p].to_dense() for p in range(len(self.Jp)))
        A = torch.einsum("p,piab->iab", thf, self.Ap)
        return J, A

    def effective_field_dense(self, J, A, e):
        """B_eff (spinmodel.py:458-473) with the coupling matrix J (3N, 3N) and on-site tensors A (N, 3, 3)."""
        h = (J @ e.reshape(-1)).reshape(-1, 3)
        g = -2.0 * h - R.MU_B * self.mu[:, None] * self.B[None, :] + 2.0 * torch.einsum("iab,ib->ia", A, e)
        g = g + self.terms_gradient(e)
        return self._sot(-g / (R.MU_B * self.mu[:, None]), e)

    def run_dense(self, J, A, e, steps, thermal=None):
        """run() with the couplings given as the dense matrix J and on-site tensors A."""
        for n in range(steps):
            b = 0.0 if thermal is None else thermal(n)
            f = lambda x: self.effective_field_dense(J, A, x) + b      # noqa: E731
            e = self._step(e, f, f, self.dt)
        return e
