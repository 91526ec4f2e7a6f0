"""The one-launch Heun step (k_step) against the two-launch step (k_heun pass 1 + pass 2), bit for bit.

Both compute the same arithmetic in the same order, so every case must agree exactly, over many steps,
for every coupling form, geometry and model feature the one-launch kernel accepts; and no wait of the
one-launch step may time out in normal operation.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden
from magex_amd import models as M
from magex_amd import spinmodel as S

pytestmark = pytest.mark.gpu


def _system(L, field=(0, 0, 0), pbc=(True, True, True), **kw):
    return S.SpinSystem(S.load_results(M.make_model(**kw)), L, pbc, field)


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
    return s, g


def _modes_agree(torch, s, precision, steps, dt, alpha=0.1, T=0.0, integrator="heun", seed=5, noise=None,
                 graph=None, sot=None):
    dm = s.device_model(precision)
    if graph is not None:
        dm.set_graph_steps(graph)
    if sot is not None:
        dm.set_sot(*sot)
    dm.set_dynamics(dt, alpha, T, True, integrator)
    e0 = torch.as_tensor(s.init_spins("random", seed=2), device="cuda", dtype=dm.dtype).contiguous()
    nz = None if noise is None else torch.as_tensor(noise, device="cuda", dtype=dm.dtype).contiguous()
    out = {}
    for mode in (0, 1):
        dm.set_step_mode(mode)
        e = e0.clone()
        dm.run(e, steps, seed, 0, nz)
        out[mode] = e
    one, given_up = dm.step_info()
    assert one, "the model should take the one-launch step"
    assert given_up == 0
    assert torch.isfinite(out[1]).all()
    assert torch.equal(out[0], out[1]), float((out[0] - out[1]).abs().max())
    return out[1]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("graph", [0, 8])
def test_triangular_three_shells_dmi(gpu, precision, graph):
    print(f"case: tri 16^2 width-4 table {precision} graph {graph}", flush=True)
    s = _system((16, 16, 1), (0, 0, 2.0), lattice="triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3)
    _modes_agree(gpu, s, precision, 40, 5e-4, 0.05, 20.0, graph=graph)


def test_bcc_per_bond_two_held_batches(gpu):
    """Two sublattices, 14 per-bond tensors per site (both held batches), per-site SIA; generated noise
    (the plain-model kernel) and injected noise."""
    print("case: bcc per-bond fp32", flush=True)
    s, g = _bcc_perbond()
    _modes_agree(gpu, s, "fp32", 30, 1e-4, 0.1, 300.0)
    noise = np.random.default_rng(4).normal(size=(10, s.N, 3))
    _modes_agree(gpu, s, "fp32", 10, 1e-4, 0.1, 300.0, noise=noise)


@pytest.mark.parametrize("integrator", ["heun", "depondt"])
def test_pinned_site_parameters_and_sot(gpu, integrator):
    """The general kernel: pinned sites, per-site damping, moments and temperatures, spin-orbit torque."""
    print(f"case: sc pinned + per-site + sot, {integrator}", flush=True)
    s = _system((6, 6, 4), (0, 0, 1.0), lattice="sc", J=[1.0], K=0.05)
    rng = np.random.default_rng(5)
    s.pinned = rng.random(s.N) < 0.2
    s.mu = 1.5 + rng.random(s.N)
    s.site_T = 50 * rng.random(s.N)
    s.touch()
    alpha = 0.05 + 0.2 * rng.random(s.N)
    e = _modes_agree(gpu, s, "fp64", 40, 5e-4, alpha, 10.0, integrator, sot=((0, 1, 0), 0.3, 0.1))
    e0 = s.init_spins("random", seed=2)
    assert np.array_equal(e.cpu().numpy()[s.pinned], e0[s.pinned])


def test_open_axes_and_tiles_spanning_planes(gpu):
    """Open boundaries (the bounds-checked gather); a chain whose 256-site tiles each cover 256 axis-0
    planes (every tile publishes many plane counters)."""
    print("case: open square + open chain", flush=True)
    s = _system((10, 7, 1), (0, 0, 0.3), (False, True, False), lattice="square", J=[2.0, -0.3], D=0.5, K=0.1)
    _modes_agree(gpu, s, "fp64", 30, 5e-4, 0.2, 15.0)
    c = _system((700, 1, 1), (0.1, 0, 0.2), (False, False, False), lattice="chain", J=[1.0, 0.1], K=0.02)
    _modes_agree(gpu, c, "fp32", 30, 5e-4, 0.2, 15.0)
    p = _system((700, 1, 1), (0.1, 0, 0.2), lattice="chain", J=[1.0, 0.1], K=0.02)
    _modes_agree(gpu, p, "fp32", 30, 5e-4, 0.2, 15.0)


def test_cubic_and_higher_uniaxial(gpu):
    print("case: sc cubic + uniaxial higher", flush=True)
    from scipy.spatial.transform import Rotation
    from magex_amd.terms import CubicAnisotropy, UniaxialHigher
    s = _system((5, 4, 6), (0.1, 0.0, 0.4), lattice="sc", J=[1.0, 0.2], K=0.05, moment=2.0)
    rng = np.random.default_rng(8)
    C = Rotation.from_rotvec([0.2, 0.4, -0.1]).as_matrix()
    s.terms += [CubicAnisotropy(rng.uniform(0.05, 0.3, s.N), C),
                UniaxialHigher(rng.uniform(0.05, 0.2, s.N), rng.uniform(0.0, 0.1, s.N), axis=(0.3, 0, 1))]
    _modes_agree(gpu, s, "fp32", 40, 2e-4, 0.1, 5.0)
