"""Parity of the device Monte Carlo colour sweep with its CPU restatement (oracle/mc_ref.py) on the same
counter-based random numbers, and its sampling against closed forms in FP32.

Bars: FP64 spins within 1e-10 of the oracle after a few sweeps and identical accepted counts (a
round-off difference could flip one acceptance, which would show as an O(1) difference: none occur
in these cases); parallel tempering over 120 sweeps within 1e-6 with identical swaps and cone widths;
FP32 Langevin averages within 0.01 of coth x - 1/x.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden
from oracle import magex_ref as R
from oracle import mc_ref
from magex_amd import models as M
from magex_amd import spinmodel as S
from magex_amd.montecarlo import MetropolisMC

pytestmark = pytest.mark.gpu


def _pair(L, field=(0, 0, 0), pbc=(True, True, True), **kw):
    d = M.make_model(**kw)
    return S.SpinSystem(S.load_results(d), L, pbc, field), R.SpinSystem(R.load_results(d), L, pbc, field)


def _run_both(s, o, mc, e0, nsweeps, beta):
    mv = {"gaussian": 0, "uniform": 1, "heatbath": 2}[mc.move]
    e = e0.copy()
    for _ in range(nsweeps):
        mc.sweep(e)
    ref = e0.copy()
    acc = 0
    for k in range(nsweeps):
        acc += mc_ref.sweep(o, ref, mc.groups, beta, mv, mc.overrelax, mc.sigma, mc.seed, k)
    return e, ref, acc


def test_mc_gaussian_overrelax_matches_oracle(gpu):
    print("case: mc sc gaussian + overrelax", flush=True)
    s, o = _pair((6, 6, 6), (0.1, 0, 0.4), lattice="sc", J=[1.0], D=0.2, K=0.1, moment=2.0)
    mc = MetropolisMC(s, 3.0, seed=11, sigma=0.5, overrelax=1, move="gaussian")
    assert mc.ncolors == 2
    e0 = s.init_spins("random", seed=1)
    e, ref, acc = _run_both(s, o, mc, e0, 5, 1 / (R.K_B * 3.0))
    assert np.abs(e - ref).max() < 1e-10
    a = mc.acceptance()
    assert abs(a - acc / (5 * s.N)) < 1e-15 and 0 < a < 1


def test_mc_heatbath_site_temperature_pinned_matches_oracle(gpu):
    """Greedy colours (triangular 3 shells + DMI), heat bath, per-site temperature, pinned sites."""
    print("case: mc tri heatbath", flush=True)
    s, o = _pair((10, 7, 1), (0, 0, 1.0), lattice="triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3)
    pin = np.zeros(s.N, bool)
    pin[::9] = True
    sT = np.full(s.N, np.nan)
    sT[: s.N // 2] = 40.0
    s.pinned, o.pinned = pin, pin
    s.site_T = sT
    mc = MetropolisMC(s, 15.0, seed=5, overrelax=0, move="heatbath")
    assert all(not pin[g].any() for g in mc.groups)
    T = np.where(np.isnan(sT), 15.0, sT)
    e0 = s.init_spins("random", seed=2)
    e, ref, acc = _run_both(s, o, mc, e0, 4, 1 / (R.K_B * T))
    assert np.array_equal(e[pin], e0[pin])
    assert np.abs(e - ref).max() < 1e-10
    assert mc.acceptance() == acc / (4 * (~pin).sum())


def test_mc_perbond_terms_uniform_matches_oracle(gpu):
    """Per-bond tensors (golden bcc per-bond lattice), per-site SIA, cubic term, uniform moves, two
    over-relaxation passes."""
    print("case: mc bcc per-bond uniform", flush=True)
    import scipy.sparse as sp
    from magex_amd.terms import CubicAnisotropy
    g = golden("bcc_perbond.npz")
    d = json.load(open(os.path.join(GOLDEN, "bcc_conventional_model.json")))
    s = S.SpinSystem(S.load_results(d), (4, 4, 4))
    o = R.SpinSystem(R.load_results(d), (4, 4, 4))
    rows = (3 * g["src"][:, None, None] + np.arange(3)[None, :, None]).repeat(3, 2).ravel()
    cols = (3 * g["dst"][:, None, None] + np.arange(3)[None, None, :]).repeat(3, 1).ravel()
    s.Jsp = sp.csr_matrix((g["blocks"].ravel(), (rows, cols)), shape=(3 * s.N, 3 * s.N))
    o.Jsp = sp.csr_matrix((g["blocks"].ravel(), (rows, cols)), shape=(3 * o.N, 3 * o.N))
    s.A = g["A"]
    o.A = g["A"]
    Kc = np.linspace(0.05, 0.2, s.N)
    s.terms.append(CubicAnisotropy(Kc))
    o.terms.append(R.CubicAnisotropy(Kc))
    mc = MetropolisMC(s, 200.0, seed=9, overrelax=2, move="uniform")
    e0 = g["e0"]
    e, ref, acc = _run_both(s, o, mc, e0, 3, 1 / (R.K_B * 200.0))
    assert np.abs(e - ref).max() < 1e-10
    assert mc.acceptance() == acc / (3 * s.N)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_mc_checker_layout_equals_row_major(gpu, precision):
    """Sweeps on the checkerboard cell layout (each parity colour contiguous) give the same spins and
    acceptance counts, bit for bit, as on the row-major layout."""
    print(f"case: mc checker vs row-major {precision}", flush=True)
    out = []
    for layout in ("checker", None):
        s, _ = _pair((6, 8, 10), (0.1, 0, 0.4), lattice="bcc_conventional", J=[2.0, 1.0], K=0.1, moment=2.2)
        mc = MetropolisMC(s, 40.0, seed=21, sigma=0.4, overrelax=1, move="gaussian", precision=precision)
        assert mc.layout == "checker"
        mc.layout = layout
        dt = gpu.float64 if precision == "fp64" else gpu.float32
        e = gpu.as_tensor(s.init_spins("random", seed=4), device="cuda", dtype=dt)
        for _ in range(4):
            mc.sweep_device(e)
        assert s.device_model(precision, layout=layout).layout() == ((-1, -1, -1) if layout else (0, 0, 0))
        out.append((e.clone(), mc.accepted))
    assert gpu.equal(out[0][0], out[1][0]) and out[0][1] == out[1][1]


def test_cuda_mc_class_shape(gpu):
    """magex's GPU Metropolis class (CudaMC(system, T, seed, sigma, move), sweep_device, acceptance):
    the same sweeps as MetropolisMC without over-relaxation, acceptance resetting per call, and the
    heat-bath move falling back to gaussian moves at T = 0 (where every accepted move lowers E)."""
    print("case: CudaMC", flush=True)
    from magex_amd.hip_backend import CudaMC
    s, _ = _pair((6, 6, 6), (0.1, 0, 0.4), lattice="sc", J=[1.0], D=0.2, K=0.1, moment=2.0)
    e0 = gpu.as_tensor(s.init_spins("random", seed=3), device="cuda")
    a, b = e0.clone(), e0.clone()
    cm = CudaMC(s, 5.0, seed=4, sigma=0.4, move="gaussian")
    mc = MetropolisMC(s, 5.0, seed=4, sigma=0.4, overrelax=0, move="gaussian")
    for _ in range(3):
        cm.sweep_device(a)
        mc.sweep_device(b)
    assert gpu.equal(a, b)
    acc = cm.acceptance()
    assert 0.0 < acc < 1.0 and cm.tried == 0 and acc == mc.acceptance()
    z = CudaMC(s, 0.0, seed=1, move="heatbath")
    assert z.mc.move == "gaussian"
    e = e0.clone()
    E0 = s.energy(e.cpu().numpy())
    for _ in range(5):
        z.sweep_device(e)
    assert s.energy(e.cpu().numpy()) <= E0


def test_two_samplers_on_one_system_do_not_interfere(gpu):
    """Samplers share the system's device model (colour lists, temperatures, acceptance counter): two
    at different temperatures, sweeping alternately, each equal a sampler running alone, and each
    counts only its own accepted moves."""
    print("case: two samplers on one system", flush=True)
    def system():
        return _pair((6, 6, 6), (0.1, 0, 0.4), lattice="sc", J=[1.0], D=0.2, K=0.1, moment=2.0)[0]
    s = system()
    e0 = s.init_spins("random", seed=3)
    a = MetropolisMC(s, 5.0, seed=4, sigma=0.4, overrelax=1)
    b = MetropolisMC(s, 50.0, seed=9, sigma=0.4, overrelax=0, move="heatbath")
    ea, eb = gpu.as_tensor(e0, device="cuda"), gpu.as_tensor(e0, device="cuda")
    for _ in range(3):
        a.sweep_device(ea)
        b.sweep_device(eb)
    solo = []
    for T, seed, orx, mv in ((5.0, 4, 1, "gaussian"), (50.0, 9, 0, "heatbath")):
        m = MetropolisMC(system(), T, seed=seed, sigma=0.4, overrelax=orx, move=mv)
        e = gpu.as_tensor(e0, device="cuda")
        for _ in range(3):
            m.sweep_device(e)
        solo.append((e, m.acceptance()))
    assert gpu.equal(ea, solo[0][0]) and gpu.equal(eb, solo[1][0])
    assert a.acceptance() == solo[0][1] and b.acceptance() == solo[1][1]


@pytest.mark.parametrize("move", ["gaussian", "heatbath"])
def test_mc_fp32_langevin(gpu, move):
    """FP32 sweeps of free spins in a field sample <cos theta> = coth x - 1/x."""
    print(f"case: mc fp32 langevin {move}", flush=True)
    import torch
    s = S.SpinSystem(S.load_results(M.make_model("sc", J=[0.0], moment=2.0)), (32, 32, 32), field=(0, 0, 1.5))
    assert len(s.stencil) == 0
    mc = MetropolisMC(s, 1.0, seed=3, move=move, precision="fp32")
    e = torch.as_tensor(s.init_spins("random", seed=0), device="cuda", dtype=torch.float32)
    mc.sweep_device(e, 20)
    vals = []
    for _ in range(60):
        mc.sweep_device(e, 1)
        vals.append(float(e[:, 2].mean()))
    x = 2.0 * R.MU_B * 1.5 / (R.K_B * 1.0)
    assert abs(np.mean(vals) - (1 / np.tanh(x) - 1 / x)) < 0.01


def _pt_reference(o, groups, Ts, therm, meas, swap_every, sample_every, seed, sigma, overrelax, init):
    """The parallel-tempering schedule of magex/advanced_mc.py:38-94 on the CPU restatement of the sweep."""
    rng = np.random.default_rng(seed)
    M = len(Ts)
    betas = 1 / (R.K_B * np.asarray(Ts))
    conf = [c.copy() for c in init]
    sig = np.full(M, sigma)
    acc, tried = np.zeros(M), np.zeros(M)
    nfree = sum(len(g) for g in groups)
    swaps_acc, swaps_try = np.zeros(M - 1), np.zeros(M - 1)
    for n in range(therm + meas):
        for k in range(M):
            acc[k] += mc_ref.sweep(o, conf[k], groups, betas[k], 0, overrelax, sig[k], seed + 17 * k, n)
            tried[k] += nfree
            if n < therm and (n + 1) % 50 == 0:
                sig[k] = np.clip(sig[k] * 0.5 / max(1 - acc[k] / tried[k], 0.05), 0.02, 4.0)
                acc[k] = tried[k] = 0
        if (n + 1) % swap_every == 0:
            E = np.array([o.energy(c) for c in conf])
            for k in range((n // swap_every) % 2, M - 1, 2):
                d = (betas[k] - betas[k + 1]) * (E[k] - E[k + 1])
                swaps_try[k] += 1
                if d >= 0 or rng.random() < np.exp(d):
                    conf[k], conf[k + 1] = conf[k + 1], conf[k]
                    E[k], E[k + 1] = E[k + 1], E[k]
                    swaps_acc[k] += 1
    return conf, sig, (swaps_acc / np.maximum(swaps_try, 1)).tolist()


def test_parallel_tempering_matches_oracle(gpu):
    """Four replicas swept together on the device, swaps and cone adaptation, against the same schedule
    on the CPU restatement (FP64, identical draws)."""
    print("case: parallel tempering vs oracle", flush=True)
    from magex_amd.advanced_mc import parallel_tempering
    s, o = _pair((4, 4, 4), (0, 0, 0.3), lattice="sc", J=[1.0], K=0.2, moment=2.0)
    Ts = [10.0, 11.0, 12.2, 13.5]
    rows, swap, conf = parallel_tempering(s, Ts, therm_sweeps=100, meas_sweeps=20, sample_every=2, swap_every=5,
                                          seed=3, init="random", sigma=0.6, overrelax=1, return_spins=True)
    mc = MetropolisMC(s, Ts[0], seed=3)
    init = [s.init_spins("random", (0, 0, 1), 3 + k) for k in range(4)]
    ref, sig, swap_ref = _pt_reference(o, mc.groups, Ts, 100, 20, 5, 2, 3, 0.6, 1, init)
    got = conf.cpu().numpy()
    # 120 sweeps carry round-off of two evaluation orders (about 1e-9 here); a single differing accept,
    # reject or swap decision would show as an O(0.1) difference
    assert max(np.abs(got[k] - ref[k]).max() for k in range(4)) < 1e-6
    assert swap == swap_ref and 0 < min(swap)
    assert np.allclose([r["sigma"] for r in rows], sig, rtol=1e-12)


def test_parallel_tempering_two_site_cells_matches_oracle(gpu):
    """Replica sweeps on a two-site cell (bcc conventional, J1 + J2, anisotropy): the replicas are swept in
    the library's internal order and written back; the schedule matches the CPU restatement as above."""
    print("case: parallel tempering vs oracle, bcc 4^3 (two sites per cell)", flush=True)
    from magex_amd.advanced_mc import parallel_tempering
    s, o = _pair((4, 4, 4), (0, 0, 0.5), lattice="bcc_conventional", a=2.87, J=[2.0, 1.0], K=0.1, moment=2.2)
    Ts = [20.0, 24.0, 29.0]
    rows, swap, conf = parallel_tempering(s, Ts, therm_sweeps=60, meas_sweeps=20, sample_every=2, swap_every=5,
                                          seed=5, init="random", sigma=0.6, overrelax=1, return_spins=True)
    mc = MetropolisMC(s, Ts[0], seed=5)
    assert s.uc.nsites == 2
    init = [s.init_spins("random", (0, 0, 1), 5 + k) for k in range(3)]
    ref, sig, swap_ref = _pt_reference(o, mc.groups, Ts, 60, 20, 5, 2, 5, 0.6, 1, init)
    got = conf.cpu().numpy()
    assert max(np.abs(got[k] - ref[k]).max() for k in range(3)) < 1e-6
    assert swap == swap_ref
    assert np.allclose([r["sigma"] for r in rows], sig, rtol=1e-12)


def test_parallel_tempering_free_spins(gpu):
    """Free spins in a field at a ladder of temperatures: E/N = -mu mu_B B L(x) at every rung."""
    print("case: parallel tempering free spins", flush=True)
    from magex_amd.advanced_mc import parallel_tempering
    s = S.SpinSystem(S.load_results(M.make_model("sc", J=[0.0], moment=2.0)), (16, 16, 16), field=(0, 0, 1.5))
    Ts = [0.5, 1.0, 2.0, 4.0]
    rows, swap = parallel_tempering(s, Ts, therm_sweeps=100, meas_sweeps=200, sample_every=2, swap_every=5,
                                    seed=2, init="random", precision="fp32", move="heatbath")
    zf = 2.0 * R.MU_B * 1.5
    for r in rows:
        x = zf / (R.K_B * r["T"])
        assert abs(r["E"] + zf * (1 / np.tanh(x) - 1 / x)) < 0.01 * zf, r
