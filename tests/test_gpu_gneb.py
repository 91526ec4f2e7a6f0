"""The device GNEB (image chain held and advanced on the GPU, batched image energies and gradients)
against the CPU restatement of magex's GNEB (oracle/gneb_ref.py) and the Stoner-Wohlfarth closed form.

Bars: FP64 barriers within 1e-6 meV of K (1 +- h)^2; images within 1e-9 of the restatement after 100
iterations and the converged barriers within 1e-6 relative; the batched energies and gradients of an
image stack equal the single-image ones bit for bit.
"""
import numpy as np
import pytest

from oracle import gneb_ref as G
from oracle import magex_ref as R
from magex_amd import models as M
from magex_amd import spinmodel as S

pytestmark = pytest.mark.gpu


def _pair(L, field=(0, 0, 0), pbc=(True, True, True), **kw):
    d = M.make_model(**kw)
    return S.SpinSystem(S.load_results(d), L, pbc, field), R.SpinSystem(R.load_results(d), L, pbc, field)


def test_image_stack_equals_single_images(gpu):
    print("case: image stack field/energy", flush=True)
    import torch
    from magex_amd.gneb import AtomisticAdapter
    for kw, L in ((dict(lattice="triangular", J=[1.0, 0.2], D=0.3, K=0.1), (8, 6, 1)),
                  (dict(lattice="bcc_conventional", J=[20.0, 10.0], moment=2.2), (3, 3, 4))):
        s, _ = _pair(L, (0, 0, 0.5), **kw)
        ad = AtomisticAdapter(s)
        X = torch.as_tensor(np.stack([s.init_spins("random", seed=k) for k in range(5)]), device="cuda")
        E, Gr = ad.images_energy_grad(X.contiguous())
        for k in range(5):
            assert torch.equal(Gr[k], torch.as_tensor(s.gradient(X[k].cpu().numpy()), device="cuda"))
            assert float(E[k]) == s.energy(X[k].cpu().numpy())


def test_slerp_path_matches_restatement(gpu):
    """The device-built initial chain equals magex's Rodrigues construction (restated), including spins
    parallel and antiparallel to their ends, to 1e-12."""
    print("case: slerp path", flush=True)
    from magex_amd.gneb import geodesic_distance, slerp_path
    rng = np.random.default_rng(5)
    a = R.normalize(rng.normal(size=(64, 3)))
    b = R.normalize(rng.normal(size=(64, 3)))
    a[:4] = [[0, 0, 1], [1, 0, 0], [0.95, 0.3, 0], [0.3, 0.2, 0.9]]
    b[:4] = [[0, 0, -1], [-1, 0, 0], [-0.95, -0.3, 0], [0.3, 0.2, 0.9]]
    a, b = R.normalize(a), R.normalize(b)
    mine, ref = slerp_path(a, b, 9), G.slerp_path(a, b, 9)
    assert max(np.abs(m - r).max() for m, r in zip(mine, ref)) < 1e-12
    assert abs(geodesic_distance(a, b) - G.geodesic(a, b)) < 1e-12


def test_gneb_stoner_wohlfarth(gpu):
    print("case: gneb stoner-wohlfarth", flush=True)
    from magex_amd.gneb import GNEB, AtomisticAdapter
    K, h, mu = 0.5, 0.3, 2.0
    B = 2 * K * h / (mu * R.MU_B)
    s, _ = _pair((1, 1, 1), (0, 0, B), (False, False, False), lattice="sc", J=[1.0], K=K, moment=mu)
    th = np.linspace(0, np.pi, 9)
    images = [np.array([[np.sin(t) * np.cos(0.4), np.sin(t) * np.sin(0.4), np.cos(t)]]) for t in th]
    g = GNEB(AtomisticAdapter(s), images).run(rel_tol=1e-8, max_iter=5000)
    r = g.result()
    assert abs(r["barrier_forward"] - K * (1 + h) ** 2) < 1e-6
    assert abs(r["barrier_backward"] - K * (1 - h) ** 2) < 1e-6
    assert abs(g.x[r["climbing_image"]][0, 2] + h) < 1e-4


def test_gneb_chain_matches_restatement(gpu):
    """A ferromagnetic chain of 8 spins reversing between +z and -z (uniaxial K, field): 300 iterations
    of the device chain against the restatement."""
    print("case: gneb chain vs restatement", flush=True)
    from magex_amd.gneb import GNEB, AtomisticAdapter, slerp_path
    s, o = _pair((8, 1, 1), (0, 0, 0.2), (False, False, False), lattice="chain", J=[1.0], K=0.3, moment=2.0)
    a = np.tile([0.0, 0.0, 1.0], (8, 1))
    b = np.tile([0.0, 0.0, -1.0], (8, 1))
    b[:, 0] = 1e-3 * np.arange(8)                      # break the symmetry of the coherent path
    b = R.normalize(b)
    images = slerp_path(a, b, 7)
    # step for step: 100 iterations agree to round-off (the optimiser's velocity-projection and
    # trust-region branches are discontinuous, so later a round-off difference can pick the other
    # branch for one step and the two chains part for a while before they meet again on the path)
    g = GNEB(AtomisticAdapter(s), images).run(rel_tol=0.0, max_iter=100)
    x, E, ci, it = G.run(o, images, rel_tol=0.0, max_iter=100)
    assert g.iterations == it == 100
    assert max(np.abs(g.x[v] - x[v]).max() for v in range(7)) < 1e-9
    assert g.climb == ci
    # converged: the same minimum-energy path and barrier
    g = GNEB(AtomisticAdapter(s), images).run(rel_tol=1e-6, max_iter=20000)
    x, E, ci, it = G.run(o, images, rel_tol=1e-6, max_iter=20000)
    r = g.result()
    assert abs(r["barrier_forward"] - (E[ci] - E[0])) < 1e-6 * abs(E[ci] - E[0])
    assert r["climbing_image"] == ci
    assert max(np.abs(g.x[v] - x[v]).max() for v in range(7)) < 1e-4


def test_minimize_reaches_tolerance(gpu):
    print("case: gneb minimize", flush=True)
    from magex_amd.gneb import AtomisticAdapter, minimize
    s, _ = _pair((6, 6, 1), (0, 0, 0.3), lattice="square", J=[1.0], D=0.2, K=0.1)
    ad = AtomisticAdapter(s)
    x0 = s.init_spins("random", seed=2)
    x, it, gmax = minimize(ad, x0, tol=1e-6, max_iter=20000)
    assert gmax < 1e-6
    assert s.energy(x) < s.energy(x0)
    assert np.abs(np.linalg.norm(x, axis=1) - 1).max() < 1e-12
