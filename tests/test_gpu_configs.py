"""The benchmark configurations at their full sizes, through checks that do not depend on size:

- bcc 64^3 (524 288 spins, per-bond random tensors, random SIA): reciprocity of the exchange operator,
  sum_i e1_i . h(e2)_i = sum_i e2_i . h(e1)_i, which holds only if every partner entry applies the
  transposed tensor of its bond (J_ji = J_ij^T); exact norm preservation; the stated FP32 bar against
  FP64 on identical noise (mean angle below 1e-5 rad, after 50 steps from the random start and after
  the workload's full 5000 steps from a ferromagnetic start);
- triangular 256^2 (65 536 spins, three shells + DMI): field and ten Heun steps on injected noise
  against the CPU oracle at the full size (FP64, 1e-10 relative), integer topological charge;
- the 96-replica batch (1.66 M spins, L from 32 to 192): the smallest and the largest replica of the
  batched FP64 run equal separate plain runs bit for bit.
"""
import numpy as np
import pytest

from oracle import magex_ref as R
from magex_amd import _lib, configs, texture
from magex_amd import models as M
from magex_amd.batch import ReplicaBatch
from magex_amd.llg import LLGSolver

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def test_bcc64_exchange_reciprocity_and_norms(gpu):
    print("case: bcc64 reciprocity, norms, fp32 angle", flush=True)
    s, p = configs.bcc64()
    dm = s.device_model("fp64")
    e1 = gpu.as_tensor(s.init_spins("random", seed=11), device="cuda")
    e2 = gpu.as_tensor(s.init_spins("random", seed=12), device="cuda")
    h1 = dm.field(e1, _lib.MX_FIELD_EXCHANGE)
    h2 = dm.field(e2, _lib.MX_FIELD_EXCHANGE)
    a, b = float((e1 * h2).sum()), float((e2 * h1).sum())
    # a transposition slip in the partner entries changes these sums by ~1e-2 relative
    assert abs(a - b) <= 1e-11 * max(abs(a), abs(b)), (a, b)
    e = {}
    for prec, dt in (("fp64", gpu.float64), ("fp32", gpu.float32)):
        sol = LLGSolver(s, p["alpha"], p["dt"], p["temperature"], p["seed"], precision=prec, noise="float")
        x = gpu.as_tensor(s.init_spins("random", seed=0), device="cuda", dtype=dt)
        sol.step_device(x, 50)
        n = x.double().norm(dim=1)
        assert float((n - 1).abs().max()) < (1e-13 if prec == "fp64" else 1e-6)
        e[prec] = x.double()
    b32 = e["fp32"] / e["fp32"].norm(dim=1, keepdim=True)
    ang = gpu.atan2(gpu.linalg.cross(e["fp64"], b32).norm(dim=1), (e["fp64"] * b32).sum(1))
    print(f"bcc64 fp32 vs fp64 after 50 steps: mean {float(ang.mean()):.3e} rad, max {float(ang.max()):.3e}")
    assert float(ang.mean()) < 1e-5


@pytest.mark.parametrize("name", ["bcc64", "tri256"])
def test_fp32_error_grows_like_a_rounding_of_the_start(gpu, name):
    """From the random start (a quench whose dynamics amplify any difference), FP32's mean angle to FP64
    on identical noise stays within 3x that of an FP64 run started 1e-7 rad away, at every checkpoint up
    to 2000 steps (measured 1.6-2.5x on MI355X): FP32 adds about two start roundings' worth of error, and
    the rest is the dynamics' own amplification."""
    print(f"case: fp32 error curve {name}", flush=True)
    import bench
    s, p = getattr(configs, name)()
    c = bench.fp32_error_curve(s, p, checkpoints=(50, 200, 1000, 2000))
    for r in c["curve"]:
        print(f"  {r['steps']:5d} steps: fp32 {r['fp32_mean_rad']:.3e} rad, control {r['control_mean_rad']:.3e} rad, "
              f"ratio {r['ratio']:.2f}", flush=True)
        assert r["ratio"] < 3.0
    assert c["curve"][1]["fp32_mean_rad"] < 1e-5          # 200 steps: FP32 within the bar


def test_bcc64_fp32_bar_over_the_full_run(gpu):
    """The FP32 bar at the workload's full length: 5000 steps from a ferromagnetic start on identical
    noise, mean per-spin angle to FP64 below 1e-5 rad (measured about 7e-7)."""
    print("case: bcc64 fp32 vs fp64, 5000 steps from ferro", flush=True)
    s, p = configs.bcc64()
    out = {}
    for prec, dt in (("fp64", gpu.float64), ("fp32", gpu.float32)):
        sol = LLGSolver(s, p["alpha"], p["dt"], p["temperature"], p["seed"], precision=prec, noise="float")
        x = gpu.as_tensor(s.init_spins("ferro"), device="cuda", dtype=dt)
        sol.step_device(x, 5000)
        out[prec] = x.double()
    b32 = out["fp32"] / out["fp32"].norm(dim=1, keepdim=True)
    ang = gpu.atan2(gpu.linalg.cross(out["fp64"], b32).norm(dim=1), (out["fp64"] * b32).sum(1))
    print(f"mean {float(ang.mean()):.3e} rad, max {float(ang.max()):.3e}")
    assert float(ang.mean()) < 1e-5


def test_tri256_full_size_matches_oracle(gpu):
    print("case: tri256 full size vs oracle", flush=True)
    s, p = configs.tri256()
    o = R.SpinSystem(R.load_results(M.make_model("triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3)), (256, 256, 1),
                     field=(0, 0, 2.0))
    e0 = s.init_spins("random", seed=0)
    assert _rel(s.effective_field(e0), o.effective_field(e0)) < 1e-12
    rng = np.random.default_rng(5)
    noise = rng.standard_normal((10, s.N, 3))
    sol = LLGSolver(s, p["alpha"], p["dt"], p["temperature"], p["seed"], precision="fp64")
    e = gpu.as_tensor(e0, device="cuda")
    sol.step_device(e, 10, noise=gpu.as_tensor(noise, device="cuda"))
    ref = R.run_heun(o, e0.copy(), 10, p["alpha"], p["dt"], T=p["temperature"], normals=lambda n: noise[n])
    assert _rel(e.cpu().numpy(), ref) < 1e-10
    # the charge of any texture on the periodic lattice is an integer
    sol.step_device(e, 400)
    q = texture.topological_charge(s, e)
    assert abs(q - round(q)) < 1e-6, q


def test_batch96_full_size_equals_separate_runs(gpu):
    """The full 96-replica batch (FP64, generated noise): the smallest and the largest replica equal their
    separate runs bit for bit, and the oracle fed its own restatement of each replica's FP64-grade
    noise stream (keyed by the replica's seed, the step and the site) within 1e-10."""
    from oracle import philox
    print("case: batch96 full size vs separate runs and the oracle", flush=True)
    systems, p = configs.batch96()
    T, seeds = p["temperatures"], p["seeds"]
    alpha = np.full(len(systems), p["alpha"])
    spins = [s.init_spins("random", seed=seeds[k]) for k, s in enumerate(systems)]
    batch = ReplicaBatch(systems, alpha, p["dt"], T, seeds, "fp64")
    e = batch.pack(spins)
    batch.run(e, 20)
    parts = batch.unpack(e)
    sizes = [s.N for s in systems]
    for k in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        ek = gpu.as_tensor(spins[k], device="cuda")
        LLGSolver(systems[k], alpha[k], p["dt"], T[k], seeds[k], precision="fp64").step_device(ek, 20)
        assert gpu.equal(parts[k], ek), k
        s = systems[k]
        o = R.SpinSystem(R.load_results(M.make_model("triangular", J=[1.0], K=0.1, moment=2.0)), s.L,
                         field=(0, 0, 0.5))
        ref = R.run_heun(o, spins[k].astype(float), 20, alpha[k], p["dt"], T=T[k],
                         normals=lambda n: philox.normals_double(seeds[k], n, s.N))
        err = _rel(parts[k].cpu().numpy(), ref)
        print(f"replica {k} ({s.L[0]}x{s.L[1]}): |batch - oracle| rel {err:.2e}")
        assert err < 1e-10, (k, err)
