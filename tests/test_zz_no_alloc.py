"""mx_llg_run_noalloc (the allocation-free step entry point) against mx_llg_run, through the C ABI.

Its scratch is one device array shared by every call on the GPU and keeping the last call's contents, so
the calls run back to back on models of falling size: a large two-launch model, then a smaller
one-launch model, then a still smaller two-sublattice model (permuted order) on injected noise.  Every
result must equal mx_llg_run's bit for bit; oversized and unconfigured models are refused.
"""
import ctypes
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden
from magex_amd import _lib, configs
from magex_amd import models as M
from magex_amd import spinmodel as S

pytestmark = pytest.mark.gpu


def _stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _noalloc(torch, dm, e, nsteps, seed, step0=0, noise=None):
    L = _lib.lib()
    _lib.check(L.mx_llg_run_noalloc(dm.h, ctypes.c_void_p(e.data_ptr()), int(nsteps), int(seed), int(step0),
                                    None if noise is None else ctypes.c_void_p(noise.data_ptr()), _stream(torch)))


def _compare(torch, s, precision, nsteps, dt, alpha, T, seed, noise=None, mode=1):
    dm = s.device_model(precision)
    dm.set_dynamics(dt, alpha, T, True, "heun")
    dm.set_step_mode(mode)
    e0 = torch.as_tensor(s.init_spins("random", seed=4), device="cuda", dtype=dm.dtype).contiguous()
    nz = None if noise is None else torch.as_tensor(noise, device="cuda", dtype=dm.dtype).contiguous()
    ref = e0.clone()
    dm.run(ref, nsteps, seed, 3, nz)
    got = e0.clone()
    _noalloc(torch, dm, got, nsteps, seed, 3, nz)
    torch.cuda.synchronize()
    assert torch.equal(got, ref), float((got - ref).abs().max())
    return dm


def _bcc_perbond():
    import scipy.sparse as sp
    g = golden("bcc_perbond.npz")
    d = json.load(open(os.path.join(GOLDEN, "bcc_conventional_model.json")))
    s = S.SpinSystem(S.load_results(d), (4, 4, 4))
    src, dst, blocks = g["src"], g["dst"], g["blocks"]
    rows = (3 * src[:, None, None] + np.arange(3)[None, :, None]).repeat(3, 2).ravel()
    cols = (3 * dst[:, None, None] + np.arange(3)[None, None, :]).repeat(3, 1).ravel()
    s.Jsp = sp.csr_matrix((blocks.ravel(), (rows, cols)), shape=(3 * s.N, 3 * s.N))
    s.A = g["A"]
    return s


def test_noalloc_equals_run_on_falling_sizes(gpu):
    print("case: noalloc bcc64 fp32 (two launches per step)", flush=True)
    s, p = configs.bcc64()
    dm = _compare(gpu, s, "fp32", 10, p["dt"], p["alpha"], p["temperature"], 7)
    assert dm.step_info()[0] == 0
    print("case: noalloc tri64 fp64 (one launch per step)", flush=True)
    t = S.SpinSystem(S.load_results(M.make_model("triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3)), (64, 64, 1),
                     field=(0, 0, 2.0))
    dm = _compare(gpu, t, "fp64", 30, 5e-4, 0.05, 20.0, 9)
    assert dm.step_info()[0] == 1
    print("case: noalloc bcc 4^3 per-bond fp32 (two sublattices), injected noise", flush=True)
    b = _bcc_perbond()
    noise = np.random.default_rng(1).normal(size=(12, b.N, 3))
    dm = _compare(gpu, b, "fp32", 12, 1e-4, 0.1, 300.0, 5, noise=noise)
    assert dm.step_info() == (1, 0)
    # the same small call once more, after the scratch held the others' contents
    _compare(gpu, b, "fp32", 5, 1e-4, 0.1, 300.0, 6)
    # persistent launches: their two counter sets and abort word live in the same scratch
    print("case: noalloc bcc 4^3 per-bond fp32, persistent launches", flush=True)
    dm = _compare(gpu, b, "fp32", 7, 1e-4, 0.1, 300.0, 8, mode=2)
    assert dm.step_info() == (2, 0)


def test_noalloc_refuses_oversized_and_unconfigured_models(gpu):
    print("case: noalloc refusals", flush=True)
    big = S.SpinSystem(S.load_results(M.make_model("sc", J=[1.0])), (128, 128, 64))   # 2^20 sites, 72 MiB FP64
    dm = big.device_model("fp64")
    dm.set_dynamics(1e-4, 0.1, 0.0, True, "heun")
    e = gpu.as_tensor(big.init_spins("random", seed=1), device="cuda")
    dm.run(e, 1, 1)
    with pytest.raises(NotImplementedError, match="too large for the allocation-free scratch"):
        _noalloc(gpu, dm, e, 1, 1)
    small = S.SpinSystem(S.load_results(M.make_model("sc", J=[1.0])), (8, 8, 8))
    dm = small.device_model("fp32")
    dm.set_dynamics(1e-4, 0.1, 5.0, True, "heun")          # constants changed, not yet on the device
    e = gpu.as_tensor(small.init_spins("random", seed=1), device="cuda", dtype=gpu.float32)
    with pytest.raises(ValueError, match="constants not on the device"):
        _noalloc(gpu, dm, e, 1, 1)
    ref = e.clone()
    dm.run(ref, 4, 2)
    got = e.clone()
    _noalloc(gpu, dm, got, 4, 2)
    assert gpu.equal(got, ref)
