"""CPU restatement of the thermal-noise generator used by the HIP kernels (test infrastructure only).

magex draws its thermal field from numpy's PCG64 (magex/llg.py:119) or, on the numba path, from
xorshift64* + ziggurat streams (magex/fast.py:123-200); neither stream is reproduced.  The HIP kernels
use a counter-based generator instead so that a step's noise is a pure function of
(seed, step, site): Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) followed by Box-Muller.
This module restates that generator so tests can pin the device's integer stream bit-exactly and
its normals to float32 round-off.

Two grades: ``normals`` (float grade, the FP32 kernels' default: 24-bit uniforms from one Philox output)
and ``normals_double`` (double grade, the FP64 default: 53-bit uniforms from two outputs).

Pinning: ``philox4x32_10`` against the Random123 known-answer vectors (tests/test_oracle.py);
``normals`` and ``normals_double`` against moments of the standard normal distribution, and the
double grade's uniforms against their defining bit layout (odd multiples of 2^-53, exact).
"""
from __future__ import annotations

import numpy as np

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)

STREAM_THERMAL = 0          # fourth counter word of the thermal-field stream


def philox4x32_10(c, k):
    """c: (..., 4) uint32 counters, k: (..., 2) uint32 keys -> (..., 4) uint32."""
    c = [np.asarray(c[..., i], np.uint64) for i in range(4)]
    k0, k1 = (np.asarray(k[..., i], np.uint64) for i in range(2))
    for r in range(10):
        p0 = _M0 * c[0]
        p1 = _M1 * c[2]
        c = [(p1 >> np.uint64(32)) ^ c[1] ^ k0, p1 & _MASK, (p0 >> np.uint64(32)) ^ c[3] ^ k1, p0 & _MASK]
        k0 = (k0 + _W0) & _MASK
        k1 = (k1 + _W1) & _MASK
    return np.stack(c, -1).astype(np.uint32)


def raw_words(seed, step, n_sites, stream=STREAM_THERMAL):
    """(n_sites, 4) Philox words of one step: counter (site, step_lo, step_hi, stream), key (seed_lo, seed_hi)."""
    seed, step = int(seed), int(step)
    c = np.zeros((n_sites, 4), np.uint32)
    c[:, 0] = np.arange(n_sites, dtype=np.uint32)
    c[:, 1] = step & 0xFFFFFFFF
    c[:, 2] = (step >> 32) & 0xFFFFFFFF
    c[:, 3] = stream
    k = np.array([seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF], np.uint32)
    return philox4x32_10(c, np.broadcast_to(k, (n_sites, 2)))


def normals(seed, step, n_sites, stream=STREAM_THERMAL):
    """(n_sites, 3) standard normals of one step: 23-bit uniforms u = (2 (w >> 9) + 1) 2^-24 in (0,1),
    Box-Muller n0 = r0 cos(2 pi u1), n1 = r0 sin(2 pi u1), n2 = r2 cos(2 pi u3), r = sqrt(-2 ln u)."""
    w = raw_words(seed, step, n_sites, stream)
    u = (2.0 * (w >> np.uint32(9)).astype(np.float64) + 1.0) * 2.0**-24
    r0 = np.sqrt(-2.0 * np.log(u[:, 0]))
    r2 = np.sqrt(-2.0 * np.log(u[:, 2]))
    return np.stack([r0 * np.cos(2 * np.pi * u[:, 1]), r0 * np.sin(2 * np.pi * u[:, 1]),
                     r2 * np.cos(2 * np.pi * u[:, 3])], 1)


def uniforms_double(seed, step, n_sites):
    """(n_sites, 4) double-grade uniforms: two Philox outputs (fourth counter word 0 and 1) give four
    64-bit words (first word high); the top 52 bits k of each give u = (2 k + 1) 2^-53 in (0, 1)."""
    a = raw_words(seed, step, n_sites, 0).astype(np.uint64)
    b = raw_words(seed, step, n_sites, 1).astype(np.uint64)
    words = [(a[:, 0] << np.uint64(32)) | a[:, 1], (a[:, 2] << np.uint64(32)) | a[:, 3],
             (b[:, 0] << np.uint64(32)) | b[:, 1], (b[:, 2] << np.uint64(32)) | b[:, 3]]
    k = np.stack([w >> np.uint64(12) for w in words], 1)
    return (2.0 * k.astype(np.float64) + 1.0) * 2.0**-53


def normals_double(seed, step, n_sites):
    """(n_sites, 3) double-grade standard normals: Box-Muller n0 = r0 cos(2 pi u1), n1 = r0 sin(2 pi u1),
    n2 = r2 cos(2 pi u3), r = sqrt(-2 ln u), in double precision."""
    u = uniforms_double(seed, step, n_sites)
    r0 = np.sqrt(-2.0 * np.log(u[:, 0]))
    r2 = np.sqrt(-2.0 * np.log(u[:, 2]))
    return np.stack([r0 * np.cos(2 * np.pi * u[:, 1]), r0 * np.sin(2 * np.pi * u[:, 1]),
                     r2 * np.cos(2 * np.pi * u[:, 3])], 1)
