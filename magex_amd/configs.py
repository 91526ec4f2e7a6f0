"""Synthetic workloads of the benchmark configurations (random-initialised, seeded; no data files).

    sc16       simple cubic 16^3 (4 096 spins), J = 1 meV (6 NN), K = 0.1 meV along z, B = 0.5 T z,
               mu = 2, alpha = 0.1, T = 0, dt = 0.1 fs, FP64
    tri256     triangular 256^2 (65 536), J = (5, -0.5, 0.2) meV over 3 shells, interfacial DMI 1.5 meV,
               K = 0.3 meV, B = 2 T z, alpha = 0.05, T = 20 K, dt = 0.5 fs
    bcc64      bcc 64^3 conventional cells (524 288), J1 = 20, J2 = 10 meV iso + per-bond random symmetric
               and antisymmetric parts (Gaussian, sigma = 1 % of |J_iso|, J_ji = J_ij^T), random symmetric
               SIA (sigma = 0.05 meV), mu = 2.2, alpha = 0.1, T = 300 K, dt = 0.1 fs
    batch96    96 triangular replicas, L uniform in {32, 48, ..., 192}, sc16's Hamiltonian on NN bonds,
               geometric temperature ladder 5-300 K, alpha = 0.1, dt = 0.1 fs
    square128  square 128^2 (16 384), J1, J2, interfacial D, K as parameters, alpha = 0.3, dt = 1 fs
"""
from __future__ import annotations

import numpy as np

from .models import make_model
from .spinmodel import SpinSystem, load_results


def sc16():
    s = SpinSystem(load_results(make_model("sc", J=[1.0], K=0.1, moment=2.0)), (16, 16, 16), field=(0, 0, 0.5))
    return s, dict(alpha=0.1, dt=1e-4, temperature=0.0, steps=1000, seed=0)


def tri256(L=256):
    s = SpinSystem(load_results(make_model("triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3)), (L, L, 1),
                   field=(0, 0, 2.0))
    return s, dict(alpha=0.05, dt=5e-4, temperature=20.0, steps=20000, seed=3)


def random_bond_tensors(system, rel=0.01, seed=1):
    """Per-bond tensors J_iso I + S + A on the forward stencil entries: S symmetric, A antisymmetric,
    independent entries ~ N(0, (rel |J_iso|)^2) (each unordered bond drawn once)."""
    rng = np.random.default_rng(seed)
    fw = system.forward_entries()
    out = np.empty((len(fw), system.ncell, 3, 3))
    iu = np.triu_indices(3)
    for n, k in enumerate(fw):
        J0 = system.stencil[k][3]
        sig = rel * abs(np.trace(J0) / 3.0)
        S = np.zeros((system.ncell, 3, 3))
        S[:, iu[0], iu[1]] = rng.normal(0.0, sig, (system.ncell, 6))
        S = S + np.triu(S, 1).transpose(0, 2, 1)
        D = rng.normal(0.0, sig, (system.ncell, 3))
        A = np.zeros((system.ncell, 3, 3))
        A[:, 0, 1], A[:, 0, 2], A[:, 1, 2] = D[:, 2], -D[:, 1], D[:, 0]
        A = A - A.transpose(0, 2, 1)
        out[n] = J0[None] + S + A
    return out


def random_sia(N, sigma=0.05, seed=2):
    rng = np.random.default_rng(seed)
    A = rng.normal(0.0, sigma, (N, 3, 3))
    return 0.5 * (A + A.transpose(0, 2, 1))


def bcc64(L=64):
    s = SpinSystem(load_results(make_model("bcc_conventional", a=2.87, J=[20.0, 10.0], moment=2.2)), (L, L, L))
    s.set_bond_tensors(random_bond_tensors(s))
    s.A = random_sia(s.N)
    return s, dict(alpha=0.1, dt=1e-4, temperature=300.0, steps=5000, seed=3)


def batch96(n=96, seed=4):
    rng = np.random.default_rng(seed)
    Ls = rng.choice(np.arange(32, 193, 16), size=n)
    Ts = 5.0 * 60.0 ** (np.arange(n) / (n - 1))
    uc = load_results(make_model("triangular", J=[1.0], K=0.1, moment=2.0))
    systems = [SpinSystem(uc, (int(L), int(L), 1), field=(0, 0, 0.5)) for L in Ls]
    return systems, dict(alpha=0.1, dt=1e-4, temperatures=Ts, steps=10000, seeds=[100 + k for k in range(n)])


def square128(J1=4.0, J2=-0.4, D=1.0, K=0.2, L=128):
    s = SpinSystem(load_results(make_model("square", J=[J1, J2], D=D, K=K)), (L, L, 1))
    return s, dict(alpha=0.3, dt=1e-3, temperature=0.0, steps=500, seed=0)
