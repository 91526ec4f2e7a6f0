"""ctypes wrapper of oracle/cpu_llg.cpp, the C++/OpenMP restatement of magex's fused CPU Heun path.

Test infrastructure and CPU baseline only (``kind: "port"`` in bench.py); never used by magex_amd.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from . import magex_ref as R

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libcpullg.so")
_lib = None


def build(force=False):
    """Compile libcpullg.so in place (make -C oracle)."""
    if force and os.path.exists(_SO):
        os.remove(_SO)
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        P = ctypes.c_void_p
        L.cpu_heun_run.restype = ctypes.c_int
        L.cpu_heun_run.argtypes = [ctypes.c_int64, P, P, P, P, ctypes.c_int, P, P, P, P, P, P, ctypes.c_double,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int64, P, P, ctypes.c_int, P, P, P, P]
        L.cpu_gauss3.argtypes = [ctypes.c_int64, P, P, ctypes.c_int]
        L.cpu_max_threads.restype = ctypes.c_int
        L.cpu_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def seed_streams(seed, n=256):
    """xorshift64* stream states (magex/fast.py:683-688)."""
    st = np.random.SeedSequence(int(seed) if seed is not None else None).generate_state(n, dtype=np.uint64)
    st[st == 0] = np.uint64(0x9E3779B97F4A7C15)
    return st


def gauss3(n, seed):
    out = np.empty((n, 3))
    st = seed_streams(seed)
    lib().cpu_gauss3(n, _p(out), _p(st), len(st))
    return out


class CpuHeun:
    """magex's fused two-pass Heun (fast.py:336-390) over neighbour_blocks of an oracle SpinSystem."""

    def __init__(self, system, alpha, dt, T=0.0, seed=0, precession=True, blocks=None):
        """blocks: optional precomputed (ptr, nbr, typ, table) instead of neighbour_blocks(Jsp)."""
        self.sys = system
        self.N = system.N
        self.ptr, self.nbr, self.typ, self.table = blocks if blocks is not None else \
            R.neighbour_blocks(system.Jsp, system.N)
        self.w = self.table.shape[1]
        A = np.asarray(system.A)
        self.A = np.ascontiguousarray(A.reshape(self.N, 9), float) if np.any(A) else None
        self.mu = np.ascontiguousarray(system.mu, float)
        self.alpha = np.ascontiguousarray(np.broadcast_to(np.asarray(alpha, float), (self.N,)))
        self.dt, self.T, self.prec = float(dt), float(T), int(bool(precession))
        self.sig = np.ascontiguousarray(R.sigma(system, self.alpha, self.T, self.dt))
        self.pinned = None if system.pinned is None else np.ascontiguousarray(system.pinned, np.uint8)
        self.st = seed_streams(seed)
        self.buf = [np.empty((self.N, 3)) for _ in range(3)]


    def run(self, e, nsteps, noise=None):
        """Advance e (N,3) float64 in place by nsteps; noise: injected normals (nsteps, N, 3) or None."""
        assert e.dtype == np.float64 and e.flags.c_contiguous and e.shape == (self.N, 3)
        noisy = int(self.T > 0 or self.sys.site_T is not None)
        if noise is not None:
            noise = np.ascontiguousarray(noise, float)
            assert noise.shape == (nsteps, self.N, 3)
        B = np.ascontiguousarray(self.sys.B, float)
        ep, k1, b = self.buf
        rc = lib().cpu_heun_run(self.N, _p(self.ptr), _p(self.nbr), _p(self.typ), _p(self.table), self.w, _p(self.A),
                                _p(self.mu), _p(B), _p(self.pinned), _p(self.alpha), _p(self.sig), self.dt, self.prec,
                                noisy, int(nsteps), _p(noise), _p(self.st), len(self.st), _p(e), _p(ep), _p(k1), _p(b))
        assert rc == 0
        return e


def set_threads(n):
    max_threads_available()
    lib().cpu_set_threads(int(n))


def max_threads():
    return lib().cpu_max_threads()


_AVAILABLE = None


def max_threads_available():
    """The thread count OpenMP started with (restored after single-thread runs)."""
    global _AVAILABLE
    if _AVAILABLE is None:
        _AVAILABLE = max_threads()
    return _AVAILABLE


def per_bond_blocks(N, src, dst, blocks):
    """(ptr, nbr, typ, table) in magex's per-bond fallback form (fast.py:612-614): every ordered pair's
    3x3 block is its own table row; neighbours of a site sorted by target index (fast.py:594-602)."""
    key = np.asarray(src, np.int64) * N + np.asarray(dst, np.int64)
    order = np.argsort(key, kind="stable")
    src, dst = np.asarray(src)[order], np.asarray(dst)[order]
    table = np.ascontiguousarray(np.asarray(blocks).reshape(-1, 9)[order])
    ptr = np.concatenate([[0], np.cumsum(np.bincount(src, minlength=N))]).astype(np.int64)
    return ptr, dst.astype(np.int32), np.arange(len(dst), dtype=np.int32), table
