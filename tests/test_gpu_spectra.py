"""Parity of the on-device spin-wave sampler and the structure factor built on it with the NumPy
restatement of magex/spectra.py (oracle: mq_sample, welch_sqw), and a physics check against the
analytic magnon dispersion of a ferromagnetic chain.

Bars: m_q within 1e-12 (FP64) of ph.T @ e; S(q, w) from the device trajectory within 1e-9 relative of
the restatement fed the same trajectory; magnon peaks within two frequency bins of
hbar gamma (4 J (1 - cos k) + 2 K) / (mu mu_B).
"""
import numpy as np
import pytest

from oracle import magex_ref as R
from magex_amd import models as M
from magex_amd import spinmodel as S
from magex_amd.llg import LLGSolver

pytestmark = pytest.mark.gpu


def _sys(L, field=(0, 0, 0), pbc=(True, True, True), **kw):
    return S.SpinSystem(S.load_results(M.make_model(**kw)), L, pbc, field)


@pytest.mark.parametrize("lattice,L", [("triangular", (12, 10, 1)), ("bcc_conventional", (3, 4, 5))])
def test_sampler_matches_restatement(gpu, lattice, L):
    print(f"case: m_q sampler {lattice}", flush=True)
    import torch
    from magex_amd.spectra import Sampler
    s = _sys(L, lattice=lattice, J=[1.0])
    B = 2 * np.pi * np.linalg.inv(s.uc.cell).T
    qc = np.random.default_rng(0).uniform(-0.5, 0.5, (7, 3)) @ B
    e = s.init_spins("random", seed=1)
    ref = R.mq_sample(s.positions, qc, e)
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-5)):
        smp = Sampler(s, qc, prec)
        got = Sampler.to_complex(smp.sample(torch.as_tensor(e, device="cuda", dtype=smp.dm.dtype)))
        assert np.abs(got - ref).max() < tol * np.sqrt(s.N), prec


def test_structure_factor_matches_restatement(gpu):
    """Device structure_factor against the restatement's Welch average of the same trajectory."""
    print("case: structure factor vs restatement", flush=True)
    from magex_amd.spectra import structure_factor
    s = _sys((16, 1, 1), (0, 0, 0.2), (True, False, False), lattice="chain", J=[1.0], K=0.05, moment=2.0)
    q = np.array([[k / 16, 0, 0] for k in range(9)])
    e0 = s.init_spins("random", seed=3)
    sol = LLGSolver(s, alpha=0.01, dt=2e-3, temperature=2.0, seed=4)
    E, Sp, Sl, _ = structure_factor(s, sol, e0, q, 2048, 8, 2)
    ref_sol = LLGSolver(s, alpha=0.01, dt=2e-3, temperature=2.0, seed=4)
    B = 2 * np.pi * np.linalg.inv(s.uc.cell).T
    e, mq, refv = e0.copy(), [], np.zeros(3)
    for _ in range(2048 // 8):
        e, _, _ = ref_sol.run(e, 8, output_every=8)
        refv += (s.signs[:, None] * e).sum(0)
        mq.append(R.mq_sample(s.positions, q @ B, e))
    E2, Sp2, Sl2 = R.welch_sqw(np.array(mq), refv, 2e-3 * 8, 2)
    assert np.allclose(E, E2, rtol=1e-14, atol=0)
    assert np.abs(Sp - Sp2).max() < 1e-9 * Sp2.max()
    assert np.abs(Sl - Sl2).max() < 1e-9 * max(Sl2.max(), 1e-300)


def test_chain_magnon_dispersion(gpu):
    """magex's own spectrum check (tests/test_micromag.py:118-137 set-up: chain of 64, J = 1, K = 0.05,
    alpha = 0.005, T = 1 K, dt = 2 fs, 16384 steps sampled every 16, 2 segments), with the analytic
    ferromagnetic-chain dispersion in place of the linear spin-wave module."""
    print("case: chain magnon dispersion", flush=True)
    from magex_amd.spectra import structure_factor
    J, K, mu = 1.0, 0.05, 2.0
    s = _sys((64, 1, 1), (0, 0, 0), (True, False, False), lattice="chain", J=[J], K=K, moment=mu)
    q = np.array([[k / 64, 0, 0] for k in range(0, 33, 2)])
    sol = LLGSolver(s, alpha=0.005, dt=2e-3, temperature=1.0, seed=1)
    e = sol.step_device(sol._to_device(s.init_spins("reference")), 2000)
    E, Sp, _, _ = structure_factor(s, sol, e, q, 16384, 16, 2)
    k = 2 * np.pi * q[:, 0]
    band = R.HBAR_MEV_PS * R.GAMMA * (4 * J * (1 - np.cos(k)) + 2 * K) / (mu * R.MU_B)
    pos = E >= 0
    peak = E[pos][np.argmax(Sp[:, pos], axis=1)]
    dE = E[1] - E[0]
    print(f"  peaks {np.round(peak, 2)}\n  bands {np.round(band, 2)}", flush=True)
    assert np.median(np.abs(peak - band)) < 2 * dE
