"""Parity of the device Monte Carlo colour sweep with its CPU restatement (oracle/mc_ref.py) on the same
counter-based random numbers, and its sampling against closed forms in FP32.

Bars: FP64 spins within 1e-10 of the oracle after several sweeps and identical accepted counts (a
round-off difference could flip one acceptance, which would show as an O(1) difference: none occur
in these cases); FP32 Langevin averages within 0.01 of coth x - 1/x.
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
