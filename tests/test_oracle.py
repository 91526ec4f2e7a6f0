"""Pins the CPU oracle (oracle/) before anything is compared with it.

Against: magex's own outputs (tests/golden, written by tests/golden/make_golden.py from magex's NumPy
path), the Random123 known-answer vectors of Philox4x32-10, and closed forms (Larmor precession,
the damped single-spin solution, finite-difference gradients, energy conservation at alpha = 0).
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import GOLDEN, golden
from oracle import magex_ref as R
from oracle import philox


def _sys(L, field=(0, 0, 0), pbc=(True, True, True), **kw):
    return R.SpinSystem(R.load_results(R.make_model(**kw)), L, pbc, field)


# ------------------------------------------------------------------ magex's own vectors
def test_make_model_matches_magex():
    d = json.load(open(os.path.join(GOLDEN, "make_model_cases.json")))
    for name, kw in d["kwargs"].items():
        mine, ref = R.make_model(**kw), d["models"][name]
        assert len(mine["pairs"]) == len(ref["pairs"]), name
        for p, q in zip(mine["pairs"], ref["pairs"]):
            assert (p["i"], p["j"], p["R"], p["nu"]) == (q["i"], q["j"], q["R"], q["nu"]), name
            assert np.allclose(p["J"], q["J"], rtol=0, atol=1e-15), name
        for k in ("cell", "positions", "moments", "A"):
            assert np.allclose(mine[k], ref[k], rtol=0, atol=1e-15), (name, k)


def test_load_results_matches_magex():
    g = golden("load_results_tri.npz")
    uc = R.load_results(R.make_model(lattice="triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3))
    assert np.array_equal([c[0] for c in uc.couplings], g["i"])
    assert np.array_equal([c[1] for c in uc.couplings], g["j"])
    assert np.array_equal([c[2] for c in uc.couplings], g["R"])
    assert np.allclose([c[3] for c in uc.couplings], g["J"], rtol=0, atol=1e-15)
    assert np.allclose(uc.A, g["A"], rtol=0, atol=1e-15)
    uc2 = R.load_results(R.make_model(lattice="triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3), rcut=3.15, jtol=0.3)
    assert len(uc2.couplings) == int(g["n_filtered"])


def test_c1_trajectory_matches_magex():
    """sc 16^3, T = 0: field, energy and the Heun trajectory after 1, 10 steps (1000 in the next test)."""
    g = golden("c1_sc16.npz")
    s = _sys((16, 16, 16), (0, 0, 0.5), lattice="sc", J=[1.0], K=0.1, moment=2.0)
    e0 = s.init_random(0)
    assert np.array_equal(e0, g["e0"])
    assert np.abs(s.effective_field(e0) - g["B0"]).max() < 1e-13
    assert abs(s.energy(e0) - g["E0"]) < 1e-10 * abs(g["E0"])
    e = R.heun_step(s, e0, 0.1, 1e-4)
    assert np.abs(e - g["e1"]).max() < 1e-14
    e = R.run_heun(s, e, 9, 0.1, 1e-4)
    assert np.abs(e - g["e10"]).max() < 1e-14


def test_c1_1000_steps_c_baseline_matches_magex():
    from oracle import cpu
    g = golden("c1_sc16.npz")
    s = _sys((16, 16, 16), (0, 0, 0.5), lattice="sc", J=[1.0], K=0.1, moment=2.0)
    e = g["e0"].copy()
    cpu.CpuHeun(s, 0.1, 1e-4).run(e, 1000)
    assert np.abs(e - g["e1000"]).max() < 1e-12


def test_tri_noise_injected_matches_magex():
    g = golden("tri12.npz")
    s = _sys((12, 12, 1), (0, 0, 2.0), lattice="triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3)
    e0 = g["e0"]
    assert np.abs(s.exchange_field(e0) - g["H0"]).max() < 1e-13
    assert np.abs(s.gradient(e0) - g["G0"]).max() < 1e-12
    assert abs(s.energy(e0) - g["E0"]) < 1e-11
    e = R.run_heun(s, e0.copy(), 10, 0.05, 5e-4, T=20.0, normals=lambda n: g["noise"][n])
    assert np.abs(e - g["e10"]).max() < 1e-13
    assert abs(s.order_parameter(e) - g["order10"]) < 1e-13


def test_square_matches_magex():
    g = golden("square8.npz")
    s = _sys((8, 8, 1), lattice="square", J=[5.0, -0.5], D=1.5, K=0.3)
    assert np.abs(s.effective_field(g["e0"]) - g["B0"]).max() < 1e-12
    e = R.run_heun(s, g["e0"].copy(), 50, 0.3, 1e-3)
    assert np.abs(e - g["e50"]).max() < 1e-13


def bcc_perbond_system():
    """The golden per-bond bcc system rebuilt in the oracle: 2-site cubic cell 4^3, per-bond tensors
    from the fixture, per-site SIA."""
    g = golden("bcc_perbond.npz")
    d = json.load(open(os.path.join(GOLDEN, "bcc_conventional_model.json")))
    s = R.SpinSystem(R.load_results(d), (4, 4, 4))
    src, dst, blocks = g["src"], g["dst"], g["blocks"]
    rows = (3 * src[:, None, None] + np.arange(3)[None, :, None]).repeat(3, 2).ravel()
    cols = (3 * dst[:, None, None] + np.arange(3)[None, None, :]).repeat(3, 1).ravel()
    s.Jsp = sp.csr_matrix((blocks.ravel(), (rows, cols)), shape=s.Jsp.shape)
    s.A = g["A"]
    return s, g


def test_bcc_perbond_matches_magex():
    s, g = bcc_perbond_system()
    assert np.abs(s.effective_field(g["e0"]) - g["B0"]).max() < 1e-12
    assert abs(s.energy(g["e0"]) - g["E0"]) < 1e-10
    e = R.run_heun(s, g["e0"].copy(), 10, 0.1, 1e-4, T=300.0, normals=lambda n: g["noise"][n])
    assert np.abs(e - g["e10"]).max() < 1e-13


def test_exchange_reciprocity_detects_untransposed_partners():
    """The full-size GPU check sum_i e1_i . h(e2)_i == sum_i e2_i . h(e1)_i has teeth: it holds for
    magex's per-bond couplings (golden bcc per-bond lattice, J_ji = J_ij^T) and fails when the partner
    entries reuse the forward tensor untransposed, a plausible slip in a once-per-bond storage."""
    g = golden("bcc_perbond.npz")
    src, dst, blocks = g["src"], g["dst"], g["blocks"]
    N = int(max(src.max(), dst.max())) + 1
    rng = np.random.default_rng(3)
    e1, e2 = R.normalize(rng.normal(size=(N, 3))), R.normalize(rng.normal(size=(N, 3)))

    def recip(b):
        rows = (3 * src[:, None, None] + np.arange(3)[None, :, None]).repeat(3, 2).ravel()
        cols = (3 * dst[:, None, None] + np.arange(3)[None, None, :]).repeat(3, 1).ravel()
        J = sp.csr_matrix((b.ravel(), (rows, cols)), shape=(3 * N, 3 * N))
        a, c = e1.ravel() @ (J @ e2.ravel()), e2.ravel() @ (J @ e1.ravel())
        return abs(a - c) / max(abs(a), abs(c))

    assert recip(blocks) < 1e-12
    # the slip: for every pair with src > dst, the block of the reverse pair used as it is
    idx = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(src, dst))}
    wrong = blocks.copy()
    for k, (a, b) in enumerate(zip(src, dst)):
        if a > b:
            wrong[k] = blocks[idx[(int(b), int(a))]]
    assert recip(wrong) > 1e-4


def test_charge_matches_magex():
    g = golden("charge.npz")
    uc = R.load_results(R.make_model(lattice="triangular", J=[1.0]))
    s = R.SpinSystem(uc, (30, 30, 1))
    for k in ("tri_sky", "tri_rand"):
        Q = R.topological_charge(s.positions, uc.cell, s.L, s.pbc, g[k + "_e"])
        assert abs(Q - g[k + "_Q"]) < 1e-9
    uc2 = R.load_results(R.make_model(lattice="square", J=[1.0]))
    s2 = R.SpinSystem(uc2, (20, 20, 1))
    assert abs(R.topological_charge(s2.positions, uc2.cell, s2.L, s2.pbc, g["sq_sky_e"]) - g["sq_sky_Q"]) < 1e-9


# ------------------------------------------------------------------ closed forms
def test_larmor_and_damped_spin_closed_form():
    """Single spin in B z: phi(t) = gamma B t / (1 + a^2) (counter-clockwise about B, as de/dt = -g e x B),
    tan(theta/2) = tan(theta0/2) exp(-a gamma B t / (1 + a^2))."""
    s = _sys((1, 1, 1), (0, 0, 2.0), (False, False, False), lattice="sc", J=[0.0])
    for a in (0.0, 0.3):
        th0 = 1.2
        e = np.array([[np.sin(th0), 0, np.cos(th0)]])
        dt, n = 1e-4, 5000
        e = R.run_heun(s, e, n, a, dt)
        t = n * dt
        w = R.GAMMA * 2.0 / (1 + a * a)
        th = 2 * np.arctan(np.tan(th0 / 2) * np.exp(-a * w * t))
        assert abs(np.arccos(e[0, 2]) - th) < 1e-7
        assert abs(np.arctan2(e[0, 1], e[0, 0]) - w * t) < 1e-7


def test_gradient_matches_finite_difference():
    s = _sys((5, 5, 1), (0.1, 0.2, 0.5), lattice="triangular", J=[1.0, 0.2], D=0.3, K=0.1)
    e = s.init_random(1)
    g = s.gradient(e)
    h = 1e-6
    for i, a in [(0, 0), (3, 2), (7, 1), (24, 0)]:
        ep, em = e.copy(), e.copy()
        ep[i, a] += h
        em[i, a] -= h
        assert abs((s.energy(ep) - s.energy(em)) / (2 * h) - g[i, a]) < 1e-7


def test_energy_conserved_without_damping():
    s = _sys((4, 4, 4), lattice="sc", J=[1.0], K=0.05)
    e = s.init_random(3)
    E0 = s.energy(e)
    e = R.run_heun(s, e, 500, 0.0, 2e-4)
    assert abs(s.energy(e) - E0) / abs(E0) < 1e-4


# ------------------------------------------------------------------ generators
def test_philox_known_answers():
    """Random123 kat_vectors for philox4x32-10."""
    cases = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
             ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
             ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
              (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for c, k, out in cases:
        got = philox.philox4x32_10(np.array(c, np.uint32), np.array(k, np.uint32))
        assert tuple(int(x) for x in got) == out


def test_philox_normals_moments():
    x = philox.normals(7, 3, 400000).ravel()
    assert abs(x.mean()) < 0.005 and abs(x.var() - 1) < 0.006
    assert abs((x**4).mean() / x.var() ** 2 - 3) < 0.03
    assert abs((np.abs(x) > 3).mean() / 0.0026998 - 1) < 0.08


def test_philox_double_grade_normals():
    """Double-grade restatement: uniforms are odd multiples of 2^-53 in (0, 1) with uniform moments; the
    normals pass moments, the 3-sigma tail and a Kolmogorov-Smirnov test against the normal CDF."""
    from scipy import stats
    u = philox.uniforms_double(11, 5, 200000)
    k = u * 2.0**53
    assert np.all(k == np.round(k)) and np.all(np.mod(k, 2) == 1) and u.min() > 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.002 and abs(u.var() - 1 / 12) < 0.001
    x = philox.normals_double(7, 3, 400000).ravel()
    assert abs(x.mean()) < 0.005 and abs(x.var() - 1) < 0.006
    assert abs((x**4).mean() / x.var() ** 2 - 3) < 0.03
    assert abs((np.abs(x) > 3).mean() / 0.0026998 - 1) < 0.08
    assert stats.kstest(x, "norm").pvalue > 1e-3
    # a different word order (low word first) is a different stream: the restatement pins the layout
    assert not np.allclose(philox.normals_double(7, 3, 1000), philox.normals(7, 3, 1000), atol=1e-3)


def test_c_ziggurat_moments():
    """The C restatement of magex's ziggurat passes magex's own moment test (test_fast.py:39-46)."""
    from oracle import cpu
    x = cpu.gauss3(300000, 7).ravel()
    assert abs(x.mean()) < 0.006 and abs(x.var() - 1) < 0.006
    assert abs((x**4).mean() / x.var() ** 2 - 3) < 0.03
    assert abs((np.abs(x) > 3).mean() / 0.0026998 - 1) < 0.08


def test_c_baseline_noise_injected_equals_numpy_oracle():
    from oracle import cpu
    s, g = bcc_perbond_system()
    e = g["e0"].copy()
    cpu.CpuHeun(s, 0.1, 1e-4, T=300.0).run(e, 10, noise=g["noise"])
    assert np.abs(e - g["e10"]).max() < 1e-13


@pytest.mark.parametrize("planted", ["dmi_sign", "self_fold", "noise_scale"])
def test_planted_mistakes_are_caught(planted):
    """A plausible oracle mistake must fail the golden comparisons above."""
    g = golden("tri12.npz")
    kw = dict(lattice="triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3)
    if planted == "dmi_sign":
        kw["D"] = -1.5
    s = _sys((12, 12, 1), (0, 0, 2.0), **kw)
    T = 20.0 * (4.0 if planted == "noise_scale" else 1.0)
    if planted == "self_fold":
        s.A = s.A * 0.0
    e = R.run_heun(s, g["e0"].copy(), 10, 0.05, 5e-4, T=T, normals=lambda n: g["noise"][n])
    assert np.abs(e - g["e10"]).max() > 1e-6


# ------------------------------------------------------------------ FP64 autograd restatement
def _square_builder(L):
    return lambda th: R.SpinSystem(R.load_results(R.make_model("square", J=[th[0], th[1]], D=th[2], K=th[3])),
                                   (L, L, 1))


def test_torch_heun_forward_matches_oracle():
    """The autograd restatement's trajectory is magex_ref.run_heun's (T = 0 and with a thermal field)."""
    import torch
    from oracle.torch_heun import TorchHeun
    th = np.array([4.0, -0.4, 1.0, 0.2])
    build = _square_builder(6)
    T = TorchHeun(build, th, alpha=0.3, dt=1e-3)
    o = build(th)
    e0 = o.init_random(5)
    ref = R.run_heun(o, e0.copy(), 25, 0.3, 1e-3)
    got = T.run(torch.as_tensor(th), torch.as_tensor(e0), 25).numpy()
    assert np.abs(got - ref).max() < 1e-13
    # a different theta than the reference point of the units
    th2 = np.array([5.0, -0.5, 1.5, 0.3])
    ref2 = R.run_heun(build(th2), e0.copy(), 25, 0.3, 1e-3)
    got2 = T.run(torch.as_tensor(th2), torch.as_tensor(e0), 25).numpy()
    assert np.abs(got2 - ref2).max() < 1e-13
    # thermal field: magex_ref draws sigma * normals; the restatement takes the field itself
    rng = np.random.default_rng(1)
    nz = [rng.normal(size=(o.N, 3)) for _ in range(10)]
    sg = R.sigma(o, 0.3, 30.0, 1e-3)
    ref3 = R.run_heun(o, e0.copy(), 10, 0.3, 1e-3, T=30.0, normals=lambda n: nz[n])
    got3 = T.run(torch.as_tensor(th), torch.as_tensor(e0), 10,
                 thermal=lambda n: torch.as_tensor(sg[:, None] * nz[n])).numpy()
    assert np.abs(got3 - ref3).max() < 1e-13


def test_torch_heun_dense_form_pinned():
    """run_dense at the dense couplings of theta is run(theta); its gradients obey the chain rule
    dL/dtheta_p = <Jsp_p, dL/dJ> + <A_p, dL/dA> and match central differences of single elements of J
    (an exchange bond, both triangles of the matrix) and of A."""
    import torch
    from oracle.torch_heun import TorchHeun
    th = np.array([4.0, -0.4, 1.0, 0.2])
    T = TorchHeun(_square_builder(5), th, alpha=0.3, dt=1e-3)
    e0 = torch.as_tensor(T.base.init_random(2))
    tgt = torch.as_tensor(T.base.init_random(3))
    J, A = T.dense(th)
    J = J.clone().requires_grad_(True)
    A = A.clone().requires_grad_(True)
    with torch.no_grad():
        assert (T.run_dense(J, A, e0, 20) - T.run(torch.as_tensor(th), e0, 20)).abs().max() < 1e-13
    loss = lambda J_, A_: ((T.run_dense(J_, A_, e0, 20) - tgt) ** 2).mean()      # noqa: E731
    loss(J, A).backward()
    t = torch.tensor(th, requires_grad=True)
    ((T.run(t, e0, 20) - tgt) ** 2).mean().backward()
    chain = [float((T.Jp[p].to_dense() * J.grad).sum() + (T.Ap[p] * A.grad).sum()) for p in range(4)]
    assert np.allclose(chain, t.grad.numpy(), rtol=1e-10, atol=1e-14)
    h = 1e-4
    with torch.no_grad():
        for (r, c) in [(0, 3), (3, 0), (4, 5)]:                    # sites 0 and 1 are nearest neighbours
            d = torch.zeros_like(J)
            d[r, c] = h
            fd = (loss(J + d, A) - loss(J - d, A)) / (2 * h)
            assert abs(fd - J.grad[r, c]) <= 1e-7 * abs(fd) + 1e-12, (r, c)
        d = torch.zeros_like(A)
        d[7, 0, 2] = h
        fd = (loss(J, A + d) - loss(J, A - d)) / (2 * h)
        assert abs(fd - A.grad[7, 0, 2]) <= 1e-7 * abs(fd) + 1e-12


def test_torch_heun_pinned_matches_oracle():
    """With pinned sites the restatement's trajectory is magex_ref.run_heun's (pinned spins held; magex
    renormalises them, which differs by round-off only), and a planted mistake, pinned spins left free,
    is far off."""
    import torch
    from oracle.torch_heun import TorchHeun
    th = np.array([4.0, -0.4, 1.0, 0.2])
    mask = np.zeros(36, bool)
    mask[[0, 7, 14, 20, 33]] = True

    def build(t):
        o = _square_builder(6)(t)
        o.pinned = mask
        return o

    T = TorchHeun(build, th, alpha=0.3, dt=1e-3)
    o = build(th)
    e0 = o.init_random(5)
    ref = R.run_heun(o, e0.copy(), 25, 0.3, 1e-3)
    got = T.run(torch.as_tensor(th), torch.as_tensor(e0), 25).numpy()
    assert np.abs(got - ref).max() < 1e-13
    assert np.abs(got[mask] - e0[mask]).max() < 1e-15
    T.pin = None
    free = T.run(torch.as_tensor(th), torch.as_tensor(e0), 25).numpy()
    assert np.abs(free - ref).max() > 1e-4


def test_torch_heun_terms_match_oracle():
    """Cubic and higher-uniaxial terms in the restatement: its trajectory is magex_ref.run_heun's, and
    its spin gradient matches central differences (the terms' Hessians enter only there)."""
    import torch
    from oracle.torch_heun import TorchHeun
    th = np.array([4.0, -0.4, 1.0, 0.2])
    Kc = np.random.default_rng(3).uniform(0.2, 0.6, 25)
    C = np.linalg.qr(np.random.default_rng(4).normal(size=(3, 3)))[0]

    def build(t):
        o = _square_builder(5)(t)
        o.terms = [R.CubicAnisotropy(Kc, C), R.UniaxialHigher(0.3, 0.1, (0.2, 0.1, 1.0))]
        return o

    T = TorchHeun(build, th, alpha=0.3, dt=1e-3)
    o = build(th)
    e0 = o.init_random(6)
    ref = R.run_heun(o, e0.copy(), 20, 0.3, 1e-3)
    got = T.run(torch.as_tensor(th), torch.as_tensor(e0), 20).numpy()
    assert np.abs(got - ref).max() < 1e-13
    tgt = torch.as_tensor(o.init_random(7))
    e = torch.tensor(e0, requires_grad=True)
    loss = lambda x: ((T.run(torch.as_tensor(th), x, 20) - tgt) ** 2).mean()      # noqa: E731
    loss(e).backward()
    h = 1e-5
    with torch.no_grad():
        for (r, c) in [(3, 0), (11, 2)]:
            d = torch.zeros_like(e)
            d[r, c] = h
            fd = (loss(e + d) - loss(e - d)) / (2 * h)
            assert abs(fd - e.grad[r, c]) <= 1e-6 * abs(fd) + 1e-12, (r, c)


def test_torch_heun_sot_matches_oracle():
    """Spin-orbit torque fields in the restatement: its trajectory is magex_ref.run_heun's with the same
    sot, and differs from the one without (a dropped term shows)."""
    import torch
    from oracle.torch_heun import TorchHeun
    th = np.array([4.0, -0.4, 1.0, 0.2])
    p = np.array([0.3, 1.0, 0.2])
    sot = (p / np.linalg.norm(p), 1.5, 0.7)
    T = TorchHeun(_square_builder(5), th, alpha=0.3, dt=1e-3, sot=sot)
    o = _square_builder(5)(th)
    e0 = o.init_random(8)
    ref = R.run_heun(o, e0.copy(), 20, 0.3, 1e-3, sot=sot)
    got = T.run(torch.as_tensor(th), torch.as_tensor(e0), 20).numpy()
    assert np.abs(got - ref).max() < 1e-13
    assert np.abs(R.run_heun(o, e0.copy(), 20, 0.3, 1e-3) - ref).max() > 1e-5


def test_torch_heun_driven_field_matches_oracle():
    """A time-dependent uniform field: the restatement, given B + field_func at t and t + dt per step,
    follows magex_ref.run_heun with that field_func; its field gradient matches central differences."""
    import torch
    from oracle.torch_heun import TorchHeun
    th = np.array([4.0, -0.4, 1.0, 0.2])
    dt, steps = 1e-3, 20
    ff = lambda t: np.array([3.0 * np.sin(40.0 * t), 0.0, 2.0 * np.cos(25.0 * t)])      # noqa: E731
    T = TorchHeun(_square_builder(5), th, alpha=0.3, dt=dt)
    o = _square_builder(5)(th)
    e0 = o.init_random(9)
    B = np.array([[o.B + ff(n * dt), o.B + ff(n * dt + dt)] for n in range(steps)])
    ref = R.run_heun(o, e0.copy(), steps, 0.3, dt, field_func=ff)
    got = T.run(torch.as_tensor(th), torch.as_tensor(e0), steps, field=torch.as_tensor(B)).numpy()
    assert np.abs(got - ref).max() < 1e-13
    tgt = torch.as_tensor(o.init_random(10))
    Bt = torch.tensor(B, requires_grad=True)
    loss = lambda b: ((T.run(torch.as_tensor(th), torch.as_tensor(e0), steps, field=b) - tgt) ** 2).mean()  # noqa
    loss(Bt).backward()
    h = 1e-5
    with torch.no_grad():
        for idx in [(0, 0, 0), (7, 1, 2), (19, 1, 1)]:
            d = torch.zeros_like(Bt)
            d[idx] = h
            fd = (loss(Bt + d) - loss(Bt - d)) / (2 * h)
            assert abs(fd - Bt.grad[idx]) <= 1e-6 * abs(fd) + 1e-13, idx


def test_torch_depondt_matches_oracle():
    """The Depondt form of the restatement: its trajectory is magex_ref.run_heun(integrator="depondt")'s
    (with a thermal field too), its gradients match central differences, and it differs from Heun's."""
    import torch
    from oracle.torch_heun import TorchHeun
    th = np.array([4.0, -0.4, 1.0, 0.2])
    T = TorchHeun(_square_builder(5), th, alpha=0.3, dt=1e-3, integrator="depondt")
    o = _square_builder(5)(th)
    e0 = o.init_random(11)
    ref = R.run_heun(o, e0.copy(), 20, 0.3, 1e-3, integrator="depondt")
    got = T.run(torch.as_tensor(th), torch.as_tensor(e0), 20).numpy()
    assert np.abs(got - ref).max() < 1e-13
    assert np.abs(R.run_heun(o, e0.copy(), 20, 0.3, 1e-3) - ref).max() > 1e-8
    rng = np.random.default_rng(2)
    nz = [rng.normal(size=(o.N, 3)) for _ in range(10)]
    sg = R.sigma(o, 0.3, 30.0, 1e-3)
    ref3 = R.run_heun(o, e0.copy(), 10, 0.3, 1e-3, T=30.0, normals=lambda n: nz[n], integrator="depondt")
    got3 = T.run(torch.as_tensor(th), torch.as_tensor(e0), 10,
                 thermal=lambda n: torch.as_tensor(sg[:, None] * nz[n])).numpy()
    assert np.abs(got3 - ref3).max() < 1e-13
    tgt = torch.as_tensor(o.init_random(12))
    t = torch.tensor(th, requires_grad=True)
    e = torch.tensor(e0, requires_grad=True)
    loss = lambda a, b: ((T.run(a, b, 20) - tgt) ** 2).mean()                        # noqa: E731
    loss(t, e).backward()
    h = 1e-5
    with torch.no_grad():
        for p in range(4):
            d = torch.zeros(4, dtype=torch.float64)
            d[p] = h
            fd = (loss(t + d, e) - loss(t - d, e)) / (2 * h)
            assert abs(fd - t.grad[p]) <= 1e-6 * abs(fd) + 1e-12, p
        d = torch.zeros_like(e)
        d[6, 2] = h
        fd = (loss(t, e + d) - loss(t, e - d)) / (2 * h)
        assert abs(fd - e.grad[6, 2]) <= 1e-6 * abs(fd) + 1e-12


def test_torch_heun_gradient_matches_central_differences():
    import torch
    from oracle.torch_heun import TorchHeun
    th = np.array([4.0, -0.4, 1.0, 0.2])
    T = TorchHeun(_square_builder(5), th, alpha=0.3, dt=1e-3)
    e0 = T.base.init_random(2)
    tgt = torch.as_tensor(T.base.init_random(3))

    def loss(t, e):
        return ((T.run(t, e, 20) - tgt) ** 2).mean()

    t = torch.tensor(th, requires_grad=True)
    e = torch.tensor(e0, requires_grad=True)
    loss(t, e).backward()
    h = 1e-4
    with torch.no_grad():
        for p in range(4):
            d = torch.zeros(4, dtype=torch.float64)
            d[p] = h
            fd = (loss(t + d, e) - loss(t - d, e)) / (2 * h)
            assert abs(fd - t.grad[p]) <= 1e-7 * abs(fd) + 1e-12, p
        # one spin component
        d = torch.zeros_like(e)
        d[3, 1] = h
        fd = (loss(t, e + d) - loss(t, e - d)) / (2 * h)
        assert abs(fd - e.grad[3, 1]) <= 1e-7 * abs(fd) + 1e-12
    # a planted mistake: a sign error in the DMI unit must show in the gradient check above
    Tbad = TorchHeun(_square_builder(5), th, alpha=0.3, dt=1e-3)
    Tbad.Jp[2] = -Tbad.Jp[2]
    with torch.no_grad():
        a = ((Tbad.run(torch.as_tensor(th), torch.as_tensor(e0), 20) - tgt) ** 2).mean()
        b = loss(torch.as_tensor(th), torch.as_tensor(e0))
    assert abs(a - b) > 1e-9


# ------------------------------------------------------------------ Depondt
def test_rotate_is_rodrigues():
    """oracle rotate against scipy's rotation-vector routine and a quarter turn about z."""
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(0)
    e = R.normalize(rng.normal(size=(50, 3)))
    w = rng.normal(size=(50, 3)) * 0.7
    w[3] = 0.0
    ref = np.array([Rotation.from_rotvec(w[i]).apply(e[i]) for i in range(50)])
    assert np.abs(R.rotate(e, w) - ref).max() < 1e-14
    q = R.rotate(np.array([[1.0, 0, 0]]), np.array([[0, 0, np.pi / 2]]))
    assert np.allclose(q, [[0, 1, 0]], atol=1e-15)


def _flow_error(step_fn, s, e0, dt, t_end):
    """Max deviation from the exact LLG flow (DOP853, tolerance 1e-13) after t_end ps."""
    from scipy.integrate import solve_ivp

    def f(t, y):
        e = y.reshape(-1, 3)
        return R.llg_rhs(e, s.effective_field(e), 0.1).ravel()

    exact = solve_ivp(f, (0, t_end), e0.ravel(), method="DOP853", rtol=1e-13, atol=1e-13).y[:, -1].reshape(-1, 3)
    e = e0.copy()
    for _ in range(int(round(t_end / dt))):
        e = step_fn(s, e, 0.1, dt)
    return np.abs(e - exact).max()


def test_depondt_is_second_order():
    """Halving dt divides the global error by 4 (second order); a corrector that forgets the second
    rotation vector is only first order and fails the same check."""
    s = _sys((3, 3, 3), field=(0.3, 0, 1.0), lattice="sc", J=[1.0], K=0.1, D=0.0)
    e0 = s.init_random(1)
    errs = [_flow_error(R.depondt_step, s, e0, dt, 0.2) for dt in (8e-3, 4e-3, 2e-3)]
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.5 < r1 < 4.5 and 3.5 < r2 < 4.5, errs

    def first_order(sys_, e, alpha, dt):
        w1 = np.cross(e, R.llg_rhs(e, sys_.effective_field(e), alpha))
        return R.normalize(R.rotate(e, dt * w1))

    bad = [_flow_error(first_order, s, e0, dt, 0.2) for dt in (8e-3, 4e-3)]
    assert bad[0] / bad[1] < 2.5
    # Depondt and Heun agree to O(dt^2)
    h = R.run_heun(s, e0.copy(), 50, 0.1, 2e-3)
    d = R.run_heun(s, e0.copy(), 50, 0.1, 2e-3, integrator="depondt")
    assert np.abs(h - d).max() < 1e-4


# ------------------------------------------------------------------ on-site terms
def test_onsite_terms_closed_forms_and_gradients():
    rng = np.random.default_rng(3)
    N = 7
    Kc, K2, K3 = rng.uniform(0.1, 1, N), rng.uniform(0.1, 1, N), rng.uniform(0.1, 1, N)
    from scipy.spatial.transform import Rotation
    C = Rotation.from_rotvec([0.3, -0.2, 0.5]).as_matrix()
    cub, uni = R.CubicAnisotropy(Kc, C), R.UniaxialHigher(K2, K3, axis=(1, 1, 0))
    # symmetry directions: cube axis -> 0; body diagonal -> Kc / 3; along u -> -K2 - K3; normal to u -> 0
    assert abs(cub.energy(np.tile(C[1], (N, 1)))) < 1e-15
    diag = (C.T @ np.ones(3)) / np.sqrt(3)
    assert abs(cub.energy(np.tile(diag, (N, 1))) - Kc.sum() / 3) < 1e-14
    u = np.array([1, 1, 0]) / np.sqrt(2)
    assert abs(uni.energy(np.tile(u, (N, 1))) + (K2 + K3).sum()) < 1e-14
    assert abs(uni.energy(np.tile([0, 0, 1.0], (N, 1)))) < 1e-15
    # gradients by central differences
    e = R.normalize(rng.normal(size=(N, 3)))
    for t in (cub, uni):
        g = t.gradient(e)
        h = 1e-6
        for i, a in ((0, 0), (3, 1), (6, 2)):
            d = np.zeros_like(e)
            d[i, a] = h
            fd = (t.energy(e + d) - t.energy(e - d)) / (2 * h)
            assert abs(fd - g[i, a]) < 1e-8 * max(1.0, abs(fd))


# ------------------------------------------------------------------ solver fields: field_func, SOT
def _single_spin(field=(0, 0, 0)):
    return _sys((1, 1, 1), field=field, pbc=(False, False, False), lattice="sc", J=[1.0], moment=2.0)


def test_field_func_precession_phase():
    """alpha = 0, B(t) = (0, 0, 1 + 0.5 sin(2 pi t / 0.5)) T: the spin keeps its polar angle and turns
    right-handed about z by phi(t) = gamma int B_z dt (de/dt = gamma B x e).  Evaluating field_func at t in the corrector too (instead of t + dt)
    is a first-order error that the same check catches."""
    s = _single_spin()
    w = 2 * np.pi / 0.5
    ff = lambda t: np.array([0.0, 0.0, 1.0 + 0.5 * np.sin(w * t)])          # noqa: E731
    th0 = 0.7
    e0 = np.array([[np.sin(th0), 0.0, np.cos(th0)]])
    dt, n = 1e-4, 4000
    e = R.run_heun(s, e0.copy(), n, 0.0, dt, field_func=ff)
    t = n * dt
    phi = R.GAMMA * (t + 0.5 * (1 - np.cos(w * t)) / w)
    ref = np.array([[np.sin(th0) * np.cos(phi), np.sin(th0) * np.sin(phi), np.cos(th0)]])
    assert np.abs(e - ref).max() < 1e-7
    lagged = e0.copy()
    for k in range(n):
        lagged = R.heun_step(s, lagged, 0.0, dt, field=lambda x, _t: R.solver_field(s, None, ff)(x, k * dt))
    assert np.abs(lagged - ref).max() > 1e-6


def test_sot_damping_like_closed_form():
    """B = B_DL (e x p), alpha = 0: theta(t) to p follows tan(theta/2) = tan(theta0/2) exp(-gamma B_DL t);
    the field-like part B_FL p is a plain field along p (precession about p at fixed theta)."""
    s = _single_spin()
    p = np.array([0.0, 1.0, 0.0])
    th0, bdl = 1.2, 0.8
    e0 = np.array([[np.sin(th0), np.cos(th0), 0.0]])
    dt, n = 1e-4, 5000
    e = R.run_heun(s, e0.copy(), n, 0.0, dt, sot=(p, bdl, 0.0))
    th = 2 * np.arctan(np.tan(th0 / 2) * np.exp(-R.GAMMA * bdl * n * dt))
    assert abs(np.arccos(e[0] @ p) - th) < 1e-7
    e = R.run_heun(s, e0.copy(), n, 0.0, dt, sot=(p, 0.0, 2.0))
    assert abs(e[0] @ p - np.cos(th0)) < 1e-10


@pytest.mark.parametrize("integrator", ["heun", "depondt"])
def test_extra_field_reproducing_an_onsite_tensor(integrator):
    """An extra_field(e) = -2 A e / (mu mu_B) is, field for field, the on-site tensor A: a model without
    anisotropy driven by it follows the model with A (thermal, pinned sites, both integrators).  The same
    extra field evaluated at the step's input spins in the corrector (instead of the predictor spins) is
    a first-order error that the check catches; an (N, 3) field_func with equal rows is the uniform one."""
    rng = np.random.default_rng(3)
    sK = _sys((4, 4, 2), field=(0, 0, 0.3), lattice="sc", J=[1.0, 0.1], K=0.4, moment=2.0)
    s0 = _sys((4, 4, 2), field=(0, 0, 0.3), lattice="sc", J=[1.0, 0.1], K=0.0, moment=2.0)
    pinned = rng.random(sK.N) < 0.15
    sK.pinned = s0.pinned = pinned
    A = sK.A.copy()
    xf = lambda e: -2.0 * np.einsum("iab,ib->ia", A, e) / (R.MU_B * s0.mu[:, None])    # noqa: E731
    e0 = R.normalize(rng.normal(size=(sK.N, 3)))
    normals = rng.normal(size=(25, sK.N, 3))
    dt = 2e-3
    pulse = lambda t: np.array([0.2 * np.sin(40 * t), 0.0, 0.1])                      # noqa: E731
    ref = R.run_heun(sK, e0.copy(), 25, 0.1, dt, T=30.0, normals=lambda n: normals[n], integrator=integrator,
                     field_func=pulse)
    got = R.run_heun(s0, e0.copy(), 25, 0.1, dt, T=30.0, normals=lambda n: normals[n], integrator=integrator,
                     field_func=pulse, extra_field=xf)
    assert np.abs(got - ref).max() < 1e-13
    assert np.abs(got[pinned] - e0[pinned]).max() < 1e-14      # renormalised, as magex's unfused path does
    rows = lambda t: np.tile(pulse(t), (s0.N, 1))                                      # noqa: E731
    got_rows = R.run_heun(s0, e0.copy(), 25, 0.1, dt, T=30.0, normals=lambda n: normals[n],
                          integrator=integrator, field_func=rows, extra_field=xf)
    assert np.abs(got_rows - got).max() < 1e-14
    step = R.heun_step if integrator == "heun" else R.depondt_step
    sg = R.sigma(s0, 0.1, 30.0, dt)
    stale = e0.copy()
    for n in range(25):
        b = sg[:, None] * normals[n]
        b[pinned] = 0.0
        x_in = stale.copy()
        fld = R.solver_field(s0, None, pulse, lambda _e: xf(x_in))
        stale = step(s0, stale, 0.1, dt, b, True, fld, n * dt)
    assert np.abs(stale - ref).max() > 1e-9


# ------------------------------------------------------------------ Monte Carlo restatement
def _langevin(x):
    return 1 / np.tanh(x) - 1 / x


def _free_spins(N, B, mu=2.0):
    s = _sys((N, 1, 1), field=(0, 0, B), pbc=(True, True, True), lattice="chain", J=[1.0], moment=mu)
    s.Jsp = sp.csr_matrix((3 * s.N, 3 * s.N))
    return s


def _mc_mean(s, groups, T, move, f, sweeps=240, burn=40, sigma=0.6, overrelax=0, seed=3):
    from oracle import mc_ref
    rng = np.random.default_rng(seed)
    e = R.normalize(rng.normal(size=(s.N, 3)))
    beta = 1 / (R.K_B * T)
    vals = []
    for k in range(sweeps):
        mc_ref.sweep(s, e, groups, beta, move, overrelax, sigma, seed, k)
        if k >= burn:
            vals.append(f(e))
    return float(np.mean(vals))


@pytest.mark.parametrize("move", [0, 1, 2])
def test_mc_single_spin_langevin(move):
    """Free spins in a field: <cos theta> = coth x - 1/x, x = beta mu mu_B B, for every move."""
    s = _free_spins(4096, 1.5)
    T = 1.0
    x = 2.0 * R.MU_B * 1.5 / (R.K_B * T)
    m = _mc_mean(s, [np.arange(s.N)], T, move, lambda e: e[:, 2].mean())
    assert abs(m - _langevin(x)) < 0.01, (m, _langevin(x))


@pytest.mark.parametrize("move", [0, 2])
def test_mc_dimer_correlation(move):
    """Independent dimers E = -2 J e1.e2 (every ordered pair counted): <e1.e2> = L(2 beta J).  A factor
    of 2 dropped from H = 2h fails it."""
    N = 2048
    s = _free_spins(N, 0.0)
    J = 1.0
    a, b = np.arange(0, N, 2), np.arange(1, N, 2)
    rows, cols = [], []
    for c in range(3):
        rows += [3 * a + c, 3 * b + c]
        cols += [3 * b + c, 3 * a + c]
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    s.Jsp = sp.csr_matrix((np.full(len(rows), J), (rows, cols)), shape=(3 * N, 3 * N))
    T = 1.0 / R.K_B                                          # 2 beta J = 2
    m = _mc_mean(s, [a, b], T, move, lambda e: np.einsum("ij,ij->i", e[a], e[b]).mean())
    assert abs(m - _langevin(2.0)) < 0.01, (m, _langevin(2.0))
    assert abs(m - _langevin(1.0)) > 0.05


def test_mc_heatbath_with_anisotropy_quadrature():
    """Heat bath around H, Metropolis-corrected for an easy axis K: <cos theta> equals the quadrature
    of exp(beta (mu mu_B B cos + K cos^2)) sin."""
    from scipy.integrate import quad
    s = _free_spins(4096, 0.8)
    K, T = 0.3, 2.0
    s.A = np.tile(-K * np.diag([0.0, 0.0, 1.0]), (s.N, 1, 1))
    beta = 1 / (R.K_B * T)
    w = lambda c: np.exp(beta * (2.0 * R.MU_B * 0.8 * c + K * c * c))            # noqa: E731
    ref = quad(lambda c: c * w(c), -1, 1)[0] / quad(w, -1, 1)[0]
    m = _mc_mean(s, [np.arange(s.N)], T, 2, lambda e: e[:, 2].mean())
    assert abs(m - ref) < 0.01, (m, ref)


def test_mc_overrelaxation_conserves_energy():
    from oracle import mc_ref
    s = _sys((4, 4, 4), field=(0.2, 0, 0.5), lattice="sc", J=[1.0], D=0.3)
    s.A[:] = 0.0
    e = s.init_random(1)
    groups = [np.where((s.cells.sum(1) % 2) == c)[0] for c in (0, 1)]  # checkerboard: no NN inside a colour
    E0 = s.energy(e)
    before = e.copy()
    # over-relaxation passes only: move pass uses gaussian with sigma 0 (no change), then two reflections
    mc_ref.sweep(s, e, groups, 1 / (R.K_B * 5.0), 0, 2, 0.0, 7, 0)
    assert abs(s.energy(e) - E0) < 1e-12 * abs(E0)
    assert np.abs(e - before).max() > 0.1


# ------------------------------------------------------------------ spectra restatement
def test_mq_sample_plane_wave_closed_form():
    """A commensurate spin spiral e = (cos k.r, sin k.r, 0): m_k = (sqrt N / 2)(1, -i, 0), zero elsewhere."""
    s = _sys((12, 8, 1), lattice="square", J=[1.0])
    B = 2 * np.pi * np.linalg.inv(s.uc.cell).T
    k = np.array([3 / 12, 2 / 8, 0]) @ B
    ph = s.positions @ k
    e = np.stack([np.cos(ph), np.sin(ph), np.zeros_like(ph)], 1)
    qs = np.array([[3 / 12, 2 / 8, 0], [1 / 12, 0, 0], [5 / 12, 3 / 8, 0]]) @ B
    m = R.mq_sample(s.positions, qs, e)
    rN = np.sqrt(s.N)
    assert np.allclose(m[0], [rN / 2, -1j * rN / 2, 0], atol=1e-12)
    assert np.abs(m[1:]).max() < 1e-12


def test_welch_peak_of_a_precessing_mode():
    """m_q(t) precessing at omega: S_perp peaks at E = +-hbar omega (within one frequency bin), S_par ~ 0."""
    dt, n, w = 0.02, 2048, 3.0
    t = np.arange(n) * dt
    mq = np.zeros((n, 1, 3), complex)
    mq[:, 0, 0] = np.exp(-1j * w * t)
    mq[:, 0, 1] = 1j * np.exp(-1j * w * t)
    E, Sp, Sl = R.welch_sqw(mq, [0, 0, 1.0], dt, 2)
    dE = E[1] - E[0]
    assert abs(abs(E[np.argmax(Sp[0])]) - R.HBAR_MEV_PS * w) <= dE
    assert Sl.max() < 1e-20


# ------------------------------------------------------------------ GNEB restatement
def _sw_spin(K, h, mu=2.0):
    """One spin, easy axis z (A = -K(zz - I/3)), field along +z with h = mu mu_B B / (2K)."""
    B = 2 * K * h / (mu * R.MU_B)
    s = _sys((1, 1, 1), field=(0, 0, B), pbc=(False, False, False), lattice="sc", J=[1.0], K=K, moment=mu)
    return s


def test_slerp_path_closed_forms():
    """The restated initial chain: every spin stays on its great circle at evenly spaced angles from a
    to b; an exactly antiparallel +z -> -z spin passes through +x (the turn about z x x = y), one along
    x -> -x through +y (about x x y = z)."""
    from oracle import gneb_ref as G
    rng = np.random.default_rng(2)
    a = R.normalize(rng.normal(size=(30, 3)))
    b = R.normalize(rng.normal(size=(30, 3)))
    a[0], b[0] = [0, 0, 1], [0, 0, -1]
    a[1], b[1] = [1, 0, 0], [-1, 0, 0]
    path = G.slerp_path(a, b, 5)
    total = np.arctan2(np.linalg.norm(np.cross(a, b), axis=1), np.einsum("ia,ia->i", a, b))
    for k in range(1, 5):
        step = np.arctan2(np.linalg.norm(np.cross(path[k - 1], path[k]), axis=1),
                          np.einsum("ia,ia->i", path[k - 1], path[k]))
        assert np.abs(step - total / 4).max() < 1e-12
        n = np.cross(a[2:], b[2:])
        assert np.abs(np.einsum("ia,ia->i", path[k][2:], n)).max() < 1e-12     # on the great circle
    assert np.allclose(path[2][0], [1, 0, 0], atol=1e-15) and np.allclose(path[2][1], [0, 1, 0], atol=1e-15)
    assert np.array_equal(path[0], a) and np.array_equal(path[-1], b)


def test_geodesic_of_a_uniform_rotation():
    from oracle import gneb_ref as G
    rng = np.random.default_rng(0)
    a = R.normalize(rng.normal(size=(50, 3)))
    th = 0.3
    b = R.rotate(a, np.tile([0, 0, 1.0], (50, 1)) * th)
    perp = np.linalg.norm(a[:, :2], axis=1)
    ang = np.arccos(np.clip((a * b).sum(1), -1, 1))
    assert abs(G.geodesic(a, b) - np.sqrt(np.sum(ang**2))) < 1e-12
    c = np.tile([1.0, 0, 0], (9, 1))
    assert abs(G.geodesic(c, R.rotate(c, np.tile([0, 0, th], (9, 1)))) - 3 * th) < 1e-12
    assert perp.min() >= 0


def test_gneb_stoner_wohlfarth_saddle():
    """Climbing-image GNEB between +z and -z of a uniaxial spin in a field: barriers K (1 +- h)^2 and
    the saddle at cos(theta) = -h."""
    from oracle import gneb_ref as G
    K, h = 0.5, 0.3
    s = _sw_spin(K, h)
    th = np.linspace(0, np.pi, 9)
    images = [np.array([[np.sin(t) * np.cos(0.4), np.sin(t) * np.sin(0.4), np.cos(t)]]) for t in th]
    x, E, ci, it = G.run(s, images, rel_tol=1e-8, max_iter=5000)
    assert abs((E[ci] - E[0]) - K * (1 + h) ** 2) < 1e-6
    assert abs((E[ci] - E[-1]) - K * (1 - h) ** 2) < 1e-6
    assert abs(x[ci][0, 2] + h) < 1e-4
