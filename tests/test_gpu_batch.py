"""Batched replicas (one launch per Heun pass for all replicas) against separate runs and the oracle."""
import numpy as np
import pytest

from oracle import magex_ref as R
from magex_amd import models as M
from magex_amd import spinmodel as S
from magex_amd.batch import ReplicaBatch
from magex_amd.llg import LLGSolver

pytestmark = pytest.mark.gpu


def _replicas(lattice, Ls, **kw):
    uc = S.load_results(M.make_model(lattice, **kw))
    return [S.SpinSystem(uc, L, field=(0, 0, 0.5)) for L in Ls]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("lattice", ["triangular", "honeycomb"])
def test_batch_equals_separate_runs(gpu, precision, lattice):
    """Every launch shape (block, sites per thread, captured loop) gives each replica exactly the
    trajectory of a separate single-replica run with the same seed."""
    print(f"case: batch vs separate {lattice} {precision}", flush=True)
    Ls = [(8, 8, 1), (13, 9, 1), (16, 16, 1), (21, 12, 1), (32, 30, 1)]
    kw = dict(J=[1.0], K=0.1, moment=2.0) if lattice == "triangular" else dict(J=[1.0, 0.1], D=0.2, K=0.1)
    systems = _replicas(lattice, Ls, **kw)
    n = len(systems)
    T = 5.0 * 60.0 ** (np.arange(n) / (n - 1))
    alpha = 0.1 + 0.02 * np.arange(n)
    seeds = [100 + k for k in range(n)]
    dt = gpu.float64 if precision == "fp64" else gpu.float32
    e0 = [s.init_spins("random", seed=100 + k) for k, s in enumerate(systems)]
    ref = []
    for k, s in enumerate(systems):
        ek = gpu.as_tensor(e0[k], device="cuda", dtype=dt)
        LLGSolver(s, alpha[k], 1e-4, T[k], seeds[k], precision=precision).step_device(ek, 40)
        ref.append(ek)
    batch = ReplicaBatch(systems, alpha, 1e-4, T, seeds, precision)
    for shape in ((256, 1, 0), (64, 4, 0), (128, 2, 16), (512, 1, 0)):
        batch.set_launch(*shape)
        batch.step_count = 0
        e = batch.pack(e0)
        batch.run(e, 40)
        for k, part in enumerate(batch.unpack(e)):
            assert gpu.equal(part, ref[k]), (shape, k)
    o = batch.observables(e)
    for k, s in enumerate(systems):
        dm = s.device_model(precision)
        assert np.allclose(o[k], dm.observables(ref[k]), rtol=1e-9, atol=1e-9 * s.N)


def test_batch_magnetisation_tracks_fp64_oracle(gpu):
    """FP32 batch on a temperature ladder vs the FP64 oracle fed the same normals: |M| per replica
    within 2e-4 over 200 steps."""
    print("case: batch fp32 vs fp64 oracle", flush=True)
    Ls = [(12, 12, 1), (16, 20, 1), (24, 24, 1)]
    systems = _replicas("triangular", Ls, J=[1.0], K=0.1, moment=2.0)
    n = len(systems)
    T = np.array([5.0, 60.0, 300.0])
    seeds = [7, 8, 9]
    e0 = [s.init_spins("ferro") for s in systems]
    batch = ReplicaBatch(systems, 0.1, 1e-4, T, seeds, "fp32")
    e = batch.pack(e0)
    batch.run(e, 200)
    m32 = batch.magnetization(e)
    uc = R.load_results(M.make_model("triangular", J=[1.0], K=0.1, moment=2.0))
    for k, s in enumerate(systems):
        o = R.SpinSystem(uc, s.L, field=(0, 0, 0.5))
        dm = s.device_model("fp64")
        dm.set_noise("float")                       # the FP32 batch's normals
        normals = [dm.noise(seeds[k], t).cpu().numpy() for t in range(200)]
        ek = R.run_heun(o, e0[k].astype(float), 200, 0.1, 1e-4, T=T[k], normals=lambda t: normals[t])
        m64 = np.linalg.norm(ek.mean(0))
        print(f"replica {k}: |M| fp32 batch {m32[k]:.6f}  fp64 oracle {m64:.6f}")
        assert abs(m32[k] - m64) < 2e-4


def test_batch_autotuner_picks_a_valid_shape(gpu):
    print("case: autotune", flush=True)
    systems = _replicas("triangular", [(32, 32, 1), (48, 48, 1), (64, 64, 1)], J=[1.0], K=0.1, moment=2.0)
    batch = ReplicaBatch(systems, 0.1, 1e-4, [5.0, 50.0, 300.0], [1, 2, 3], "fp32")
    e = batch.pack([s.init_spins("random", seed=1) for s in systems])
    before = e.clone()
    shape, ms, results = batch.autotune(e, steps=4)
    assert shape in [r[0] for r in results] and ms == min(r[1] for r in results)
    assert gpu.equal(e, before) and batch.step_count == 0
