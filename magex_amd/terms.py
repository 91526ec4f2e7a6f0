"""On-site energy terms that run inside the device kernels (drop-in for magex.terms' local terms).

    CubicAnisotropy   E_i = Kc (a1^2 a2^2 + a2^2 a3^2 + a3^2 a1^2), a_k = e_i . c_k  (magex/terms.py:35-55)
    UniaxialHigher    E_i = -K2 (e_i.u)^4 - K3 (e_i.u)^6                          (magex/terms.py:58-75)

They are folded into the field / Heun kernels as magex folds them into its numba kernels
(magex/fast.py:47-63, 647-652).  The non-local terms of magex.terms (Quadruplets, Dipolar) are not on
this path; constructing them raises NotImplementedError.
"""
from __future__ import annotations

import numpy as np


class _Term:
    long_range = False

    def bonds(self):
        return None


class CubicAnisotropy(_Term):
    def __init__(self, Kc, axes=None):
        self.Kc = np.asarray(Kc, float)
        self.C = np.eye(3) if axes is None else np.asarray(axes, float)


class UniaxialHigher(_Term):
    def __init__(self, K2, K3=0.0, axis=(0, 0, 1)):
        self.K2 = np.asarray(K2, float)
        self.K3 = np.asarray(K3, float) * np.ones_like(self.K2)
        u = np.asarray(axis, float)
        self.u = u / np.linalg.norm(u)


class Quadruplets(_Term):
    def __init__(self, *a, **k):
        raise NotImplementedError("four-spin quadruplets are not on the device LLG path")


class Dipolar(_Term):
    long_range = True

    def __init__(self, *a, **k):
        raise NotImplementedError("dipolar fields are not on the device LLG path")
