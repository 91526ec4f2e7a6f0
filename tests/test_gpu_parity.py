"""Parity of the HIP path (libmxllg.so through magex_amd) with magex's outputs and the CPU oracle.

Bars: FP64 within 1e-10 relative of magex's CPU Heun (the golden trajectories, and the oracle on the
same injected noise); the noise generator bit-compatible with its Philox restatement up to float32
Box-Muller round-off; FP32 within a stated per-spin angle error of FP64 on identical noise.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden
from oracle import magex_ref as R
from oracle import philox
from magex_amd import models as M
from magex_amd import spinmodel as S
from magex_amd.llg import LLGSolver

pytestmark = pytest.mark.gpu


def _pair(L, field=(0, 0, 0), pbc=(True, True, True), **kw):
    d = M.make_model(**kw)
    return (S.SpinSystem(S.load_results(d), L, pbc, field), R.SpinSystem(R.load_results(d), L, pbc, field))


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def test_c1_fp64_1000_steps_matches_magex(gpu):
    """Config 1 (sc 16^3, J=1, K=0.1, B_z=0.5, alpha=0.1, dt=0.1 fs, T=0, 1000 steps) vs magex's Heun."""
    print("case: c1 sc16 fp64 1000 steps", flush=True)
    g = golden("c1_sc16.npz")
    s, _ = _pair((16, 16, 16), (0, 0, 0.5), lattice="sc", J=[1.0], K=0.1, moment=2.0)
    assert _rel(s.effective_field(g["e0"]), g["B0"]) < 1e-13
    assert abs(s.energy(g["e0"]) - g["E0"]) <= 1e-12 * abs(g["E0"])
    sol = LLGSolver(s, alpha=0.1, dt=1e-4, seed=0)
    e = sol.step(g["e0"])
    assert _rel(e, g["e1"]) < 1e-13
    e, tab, _ = sol.run(e, 9, output_every=9)
    assert _rel(e, g["e10"]) < 1e-12
    e, tab, _ = sol.run(e, 990, output_every=330)
    assert _rel(e, g["e1000"]) < 1e-10
    assert abs(tab[-1, 1] * s.N - g["E1000"]) < 1e-10 * abs(g["E1000"])


def test_tri_field_energy_and_injected_noise_match_magex(gpu):
    """Triangular 3 shells + DMI (width-4 table): field, gradient, exchange, energy, 10 noisy steps."""
    print("case: tri12 golden", flush=True)
    g = golden("tri12.npz")
    s, o = _pair((12, 12, 1), (0, 0, 2.0), lattice="triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3)
    e0 = g["e0"]
    assert _rel(s.exchange_field(e0), g["H0"]) < 1e-13
    assert _rel(s.gradient(e0), g["G0"]) < 1e-13
    assert _rel(s.effective_field(e0), g["B0"]) < 1e-13
    assert abs(s.energy(e0) - g["E0"]) < 1e-12 * abs(g["E0"])
    sol = LLGSolver(s, alpha=0.05, dt=5e-4, temperature=20.0, seed=1)
    ed = gpu.as_tensor(e0, device="cuda")
    sol.step_device(ed, 10, noise=gpu.as_tensor(g["noise"], device="cuda"))
    e = ed.cpu().numpy()
    assert _rel(e, g["e10"]) < 1e-12
    assert abs(s.order_parameter(e) - g["order10"]) < 1e-12


def test_square_dmi_matches_magex(gpu):
    print("case: square8 golden", flush=True)
    g = golden("square8.npz")
    s, _ = _pair((8, 8, 1), lattice="square", J=[5.0, -0.5], D=1.5, K=0.3)
    assert _rel(s.effective_field(g["e0"]), g["B0"]) < 1e-13
    e, _, _ = LLGSolver(s, alpha=0.3, dt=1e-3).run(g["e0"], 50, output_every=50)
    assert _rel(e, g["e50"]) < 1e-11


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


def test_bcc_perbond_matches_magex(gpu):
    """Two sublattices, per-bond random tensors stored once per bond, per-site symmetric SIA, T=300 K."""
    print("case: bcc per-bond golden", flush=True)
    s, g = _bcc_perbond()
    assert _rel(s.effective_field(g["e0"]), g["B0"]) < 1e-13
    assert abs(s.energy(g["e0"]) - g["E0"]) < 1e-12 * abs(g["E0"])
    sol = LLGSolver(s, alpha=0.1, dt=1e-4, temperature=300.0, seed=3)
    ed = gpu.as_tensor(g["e0"], device="cuda")
    sol.step_device(ed, 10, noise=gpu.as_tensor(g["noise"], device="cuda"))
    assert _rel(ed.cpu().numpy(), g["e10"]) < 1e-12


def test_device_noise_is_the_philox_stream(gpu):
    """Both generator grades against their restatements: double grade (the FP64 default) to 1e-13, float
    grade (the FP32 default) to float32 Box-Muller round-off; an FP64 model set to the float grade draws
    exactly the FP32 model's normals."""
    print("case: noise vs philox", flush=True)
    s, _ = _pair((64, 64, 1), lattice="square", J=[1.0])
    dm = s.device_model("fp64")
    dm.set_noise("double")
    for seed, step in ((0, 0), (12345, 7), (2**40 + 3, 2**33 + 5)):
        dev = dm.noise(seed, step).cpu().numpy()
        assert np.abs(dev - philox.normals_double(seed, step, s.N)).max() < 1e-13
    x = np.concatenate([dm.noise(9, k).cpu().numpy().ravel() for k in range(40)])
    assert abs(x.mean()) < 0.005 and abs(x.var() - 1) < 0.006
    assert abs((x**4).mean() / x.var() ** 2 - 3) < 0.03
    dm.set_noise("float")
    for seed, step in ((0, 0), (12345, 7), (2**40 + 3, 2**33 + 5)):
        dev = dm.noise(seed, step).cpu().numpy()
        assert np.abs(dev - philox.normals(seed, step, s.N)).max() < 2e-5
    f32 = s.device_model("fp32").noise(12345, 7).cpu().numpy()
    assert np.array_equal(f32.astype(np.float64), dm.noise(12345, 7).cpu().numpy())
    dm.set_noise("double")


def test_double_grade_noise_tails(gpu):
    """Tail frequencies of the double-grade normals over 5e7 draws: P(|n| > 4) and P(|n| > 5) within five
    Poisson standard deviations of the normal distribution's (the float grade stops at 5.77 sigma)."""
    print("case: double-grade noise tails", flush=True)
    from scipy.special import erfc
    s, _ = _pair((512, 512, 1), lattice="square", J=[1.0])
    dm = s.device_model("fp64")
    dm.set_noise("double")
    counts = {4.0: 0, 5.0: 0}
    n = 0
    big = 0.0
    for k in range(64):
        x = dm.noise(2024, k).abs()
        n += x.numel()
        big = max(big, float(x.max()))
        for c in counts:
            counts[c] += int((x > c).sum())
    for c, got in counts.items():
        want = n * erfc(c / np.sqrt(2.0))
        print(f"  |n| > {c}: {got} of {n} (expected {want:.1f})", flush=True)
        assert abs(got - want) < 5 * np.sqrt(want)
    print(f"  largest |n| {big:.3f}")
    assert big < 8.7


def test_thermal_run_matches_oracle_on_device_noise(gpu):
    """Internal generator at T > 0: 40 steps on a 32^2 triangular DMI lattice equal the oracle fed the
    same normals (read back with mx_noise)."""
    print("case: tri32 thermal vs oracle", flush=True)
    s, o = _pair((32, 32, 1), (0, 0, 2.0), lattice="triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3)
    e0 = s.init_spins("random", seed=0)
    sol = LLGSolver(s, alpha=0.05, dt=5e-4, temperature=20.0, seed=77)
    e, _, _ = sol.run(e0, 40, output_every=40)
    dm = s.device_model("fp64")
    normals = [dm.noise(77, n).cpu().numpy() for n in range(40)]
    ref = R.run_heun(o, e0.copy(), 40, 0.05, 5e-4, T=20.0, normals=lambda n: normals[n])
    assert _rel(e, ref) < 1e-11


@pytest.mark.parametrize("kw,L,pbc", [
    (dict(lattice="honeycomb", J=[1.0, 0.1], D=0.1, K=0.05), (6, 5, 1), (True, True, True)),
    (dict(lattice="kagome", J=[1.0], K=0.05, K_axis=(1, 0, 1)), (5, 4, 1), (True, True, True)),
    (dict(lattice="sc", J=[-1.0], order="afm"), (3, 4, 3), (True, True, True)),
    (dict(lattice="triangular", J=[1.0, 0.2], D=0.3), (7, 6, 1), (False, True, False)),
    (dict(lattice="sc", J=[1.0, 0.3], K=0.1), (2, 3, 4), (True, True, False)),
    (dict(lattice="bcc_conventional", J=[20.0, 10.0], moment=2.2), (3, 4, 5), (True, True, True)),
    (dict(lattice="sc", J=[1.0], Jani=[[0.1, 0.02, 0], [0.02, -0.05, 0], [0, 0, -0.05]], D=0.1), (4, 4, 4),
     (True, True, True)),
])
def test_lattices_match_oracle(gpu, kw, L, pbc):
    """Several sublattices, open axes, self-image folding, width-9 tables: field, observables, 30 steps."""
    print(f"case: lattice {kw['lattice']} {L} {pbc}", flush=True)
    s, o = _pair(L, (0.1, -0.2, 0.4), pbc, **kw)
    e0 = s.init_spins("random", seed=4)
    assert _rel(s.effective_field(e0), o.effective_field(e0)) < 1e-13
    assert abs(s.energy(e0) - o.energy(e0)) < 1e-12 * max(abs(o.energy(e0)), 1.0)
    assert abs(s.order_parameter(e0) - o.order_parameter(e0)) < 1e-13
    sol = LLGSolver(s, alpha=0.2, dt=1e-3)
    mt = sol.max_torque(e0)
    ref_t = np.max(np.linalg.norm(np.cross(e0, o.effective_field(e0)), axis=1))
    assert abs(mt - ref_t) < 1e-12 * ref_t
    e, _, _ = sol.run(e0, 30, output_every=10)
    assert _rel(e, R.run_heun(o, e0.copy(), 30, 0.2, 1e-3)) < 1e-11


def test_pinned_and_site_dependent_inputs(gpu):
    """Pinned spins stay fixed; per-site alpha, moments and temperatures enter sigma as magex's."""
    print("case: pinned + per-site", flush=True)
    s, o = _pair((6, 6, 2), (0, 0, 1.0), lattice="sc", J=[1.0], K=0.05)
    rng = np.random.default_rng(5)
    pinned = rng.random(s.N) < 0.2
    mu = 1.5 + rng.random(s.N)
    alpha = 0.05 + 0.2 * rng.random(s.N)
    Tsite = 50 * rng.random(s.N)
    for sys_ in (s, o):
        sys_.pinned = pinned
        sys_.mu = mu.copy()
        sys_.site_T = Tsite
    s.touch()
    e0 = s.init_spins("random", seed=6)
    sol = LLGSolver(s, alpha=alpha, dt=5e-4, temperature=10.0, seed=11)
    e, _, _ = sol.run(e0, 20, output_every=20)
    dm = s.device_model("fp64")
    normals = [dm.noise(11, n).cpu().numpy() for n in range(20)]
    ref = R.run_heun(o, e0.copy(), 20, alpha, 5e-4, T=10.0, normals=lambda n: normals[n])
    assert _rel(e, ref) < 1e-11
    assert np.array_equal(e[pinned], e0[pinned])
    assert _rel(s.gradient(e0), o.gradient(e0)) < 1e-13


def test_fp32_angle_error_against_fp64(gpu):
    """FP32 vs FP64 on identical noise (tri 64^2, 3 shells + DMI, T = 20 K, 500 steps): mean per-spin
    angle below 1e-5 rad, max below 1e-3 rad."""
    print("case: fp32 vs fp64 angle", flush=True)
    s, _ = _pair((64, 64, 1), (0, 0, 2.0), lattice="triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3)
    e0 = s.init_spins("random", seed=0)
    a = LLGSolver(s, alpha=0.05, dt=5e-4, temperature=20.0, seed=3, noise="float").run(e0, 500, output_every=500)[0]
    b = LLGSolver(s, alpha=0.05, dt=5e-4, temperature=20.0, seed=3, precision="fp32").run(
        e0.astype(np.float32), 500, output_every=500)[0].astype(np.float64)
    b = R.normalize(b)
    ang = np.arctan2(np.linalg.norm(np.cross(a, b), axis=1), np.einsum("ij,ij->i", a, b))
    print(f"fp32 angle error: mean {ang.mean():.3e} max {ang.max():.3e}")
    assert ang.mean() < 1e-5 and ang.max() < 1e-3


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("lattice", ["triangular", "bcc_conventional"])
def test_captured_step_loop_equals_direct_launches(gpu, precision, lattice):
    """hipGraph replays (device step counter) give bit-identical spins to step-by-step launches, across
    several calls with eager work and synchronisations between the replays."""
    print(f"case: graph vs direct {lattice} {precision}", flush=True)
    kw = dict(lattice=lattice, J=[5.0, -0.5], D=1.5, K=0.3) if lattice == "triangular" else \
        dict(lattice=lattice, J=[20.0, 10.0], moment=2.2)
    L = (48, 40, 1) if lattice == "triangular" else (6, 5, 7)
    s, _ = _pair(L, (0, 0, 1.0), **kw)
    dt = gpu.float64 if precision == "fp64" else gpu.float32
    e0 = gpu.as_tensor(s.init_spins("random", seed=1), device="cuda", dtype=dt)
    direct = LLGSolver(s, alpha=0.1, dt=2e-4, temperature=30.0, seed=5, precision=precision)
    direct.dm.set_graph_steps(0)
    ea = e0.clone()
    direct.step_device(ea, 100)
    graph = LLGSolver(s, alpha=0.1, dt=2e-4, temperature=30.0, seed=5, precision=precision)
    eb = e0.clone()
    for chunk in (37, 45, 18):
        graph.dm.set_graph_steps(16)
        graph.step_device(eb, chunk)
        gpu.cuda.synchronize()
        _ = (eb * 2.0).sum().item()          # eager work between replays
    direct.dm.set_graph_steps(32)
    assert gpu.equal(ea, eb)


@pytest.mark.parametrize("case", ["bcc_perbond", "bcc_columns", "tri_dmi"])
def test_tiled_layout_equals_row_major(gpu, case):
    """The tiled cell layout (tile-major internal order) changes no result bit of the field or of 25
    thermal steps against the row-major layout (FP64 and FP32); the observables, sums taken in the
    internal site order, agree to round-off.  bcc_columns: axis-0 planes of 32 x 64 cells (8 tiles),
    where the row-major step kernels walk per-XCD columns instead of slabs."""
    from magex_amd.configs import random_bond_tensors, random_sia
    from magex_amd.device import DeviceModel
    print(f"case: tiled vs row-major {case}", flush=True)
    if case in ("bcc_perbond", "bcc_columns"):
        L = (8, 8, 16) if case == "bcc_perbond" else (4, 32, 64)
        s, _ = _pair(L, lattice="bcc_conventional", J=[20.0, 10.0], moment=2.2)
        s.set_bond_tensors(random_bond_tensors(s, seed=3))
        s.A = random_sia(s.N, seed=4)
    else:
        s, _ = _pair((32, 64, 1), (0, 0, 2.0), lattice="triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3)
    e0 = s.init_spins("random", seed=2)
    for prec, dt in (("fp64", gpu.float64), ("fp32", gpu.float32)):
        a = DeviceModel(s, prec, tiled=True)
        b = DeviceModel(s, prec, tiled=False)
        assert a.layout() != (0, 0, 0) and b.layout() == (0, 0, 0)
        ed = gpu.as_tensor(e0, device="cuda", dtype=dt)
        assert gpu.equal(a.field(ed), b.field(ed))
        oa, ob = a.observables(ed), b.observables(ed)         # sums in another site order: round-off only
        assert np.allclose(oa, ob, rtol=1e-12, atol=1e-12 * np.abs(ob).max())
        for m in (a, b):
            m.set_dynamics(1e-4, 0.1, 300.0)
        ea, eb = ed.clone(), ed.clone()
        a.run(ea, 25, seed=9)
        b.run(eb, 25, seed=9)
        assert gpu.equal(ea, eb)


@pytest.mark.parametrize("case", ["bcc_perbond", "sc_dmi"])
def test_checker_layout_equals_row_major(gpu, case):
    """The checkerboard cell layout (even c0 + c1 + c2 first) changes no result bit of the field or of
    25 thermal steps against the row-major layout (FP64 and FP32)."""
    from magex_amd.configs import random_bond_tensors, random_sia
    from magex_amd.device import DeviceModel
    print(f"case: checker vs row-major {case}", flush=True)
    if case == "bcc_perbond":
        s, _ = _pair((6, 4, 8), lattice="bcc_conventional", J=[20.0, 10.0], moment=2.2)
        s.set_bond_tensors(random_bond_tensors(s, seed=3))
        s.A = random_sia(s.N, seed=4)
    else:
        s, _ = _pair((5, 6, 10), (0.2, 0, 0.7), lattice="sc", J=[1.0, 0.1], D=0.3, K=0.1, moment=2.0)
    e0 = s.init_spins("random", seed=2)
    for prec, dt in (("fp64", gpu.float64), ("fp32", gpu.float32)):
        a = DeviceModel(s, prec, layout="checker")
        b = DeviceModel(s, prec, layout="rowmajor")
        assert a.layout() == (-1, -1, -1) and b.layout() == (0, 0, 0)
        ed = gpu.as_tensor(e0, device="cuda", dtype=dt)
        assert gpu.equal(a.field(ed), b.field(ed))
        oa, ob = a.observables(ed), b.observables(ed)
        assert np.allclose(oa, ob, rtol=1e-12, atol=1e-12 * np.abs(ob).max())
        for m in (a, b):
            m.set_dynamics(1e-4, 0.1, 300.0)
        ea, eb = ed.clone(), ed.clone()
        a.run(ea, 25, seed=9)
        b.run(eb, 25, seed=9)
        assert gpu.equal(ea, eb)


def test_topological_charge_matches_magex(gpu):
    print("case: topological charge", flush=True)
    from magex_amd.texture import topological_charge
    g = golden("charge.npz")
    s, _ = _pair((30, 30, 1), lattice="triangular", J=[1.0])
    for k in ("tri_sky", "tri_rand"):
        assert abs(topological_charge(s, g[k + "_e"]) - g[k + "_Q"]) < 1e-9
    s2, _ = _pair((20, 20, 1), lattice="square", J=[1.0])
    assert abs(topological_charge(s2, g["sq_sky_e"]) - g["sq_sky_Q"]) < 1e-9
    assert abs(topological_charge(s, np.tile([0, 0, 1.0], (s.N, 1)))) < 1e-12


def test_larmor_on_device(gpu):
    """Single free spin: Larmor phase and damped polar angle against the closed form."""
    print("case: larmor", flush=True)
    s, _ = _pair((1, 1, 1), (0, 0, 2.0), (False, False, False), lattice="sc", J=[0.0])
    th0, dt, n = 1.2, 1e-4, 5000
    for a in (0.0, 0.3):
        e = np.array([[np.sin(th0), 0, np.cos(th0)]])
        e, _, _ = LLGSolver(s, alpha=a, dt=dt).run(e, n, output_every=n)
        w = R.GAMMA * 2.0 / (1 + a * a)
        th = 2 * np.arctan(np.tan(th0 / 2) * np.exp(-a * w * n * dt))
        assert abs(np.arccos(e[0, 2]) - th) < 1e-7
        assert abs(np.arctan2(e[0, 1], e[0, 0]) - w * n * dt) < 1e-7


def test_llg_rhs_on_device(gpu):
    from magex_amd.llg import llg_rhs
    rng = np.random.default_rng(1)
    e = R.normalize(rng.normal(size=(100, 3)))
    B = rng.normal(size=(100, 3))
    for prec in (True, False):
        assert _rel(llg_rhs(e, B, 0.3, prec), R.llg_rhs(e, B, 0.3, prec)) < 1e-14


@pytest.mark.parametrize("precession", [True, False])
def test_depondt_matches_oracle(gpu, precession):
    """LLGSolver(integrator="depondt") against the oracle's Depondt step: T = 0 on a per-bond bcc
    lattice, and the internal generator at T > 0 on a triangular DMI lattice (device normals)."""
    print(f"case: depondt precession={precession}", flush=True)
    s, g = _bcc_perbond()
    d = json.load(open(os.path.join(GOLDEN, "bcc_conventional_model.json")))
    o = R.SpinSystem(R.load_results(d), (4, 4, 4))
    import scipy.sparse as sp
    rows = (3 * g["src"][:, None, None] + np.arange(3)[None, :, None]).repeat(3, 2).ravel()
    cols = (3 * g["dst"][:, None, None] + np.arange(3)[None, None, :]).repeat(3, 1).ravel()
    o.Jsp = sp.csr_matrix((g["blocks"].ravel(), (rows, cols)), shape=(3 * o.N, 3 * o.N))
    o.A = g["A"]
    s.set_field((0.2, 0, 0.5))
    o.set_field((0.2, 0, 0.5))
    e0 = g["e0"]
    sol = LLGSolver(s, alpha=0.2, dt=1e-4, integrator="depondt", precession=precession)
    e, _, _ = sol.run(e0, 60, output_every=60)
    ref = R.run_heun(o, e0.copy(), 60, 0.2, 1e-4, precession=precession, integrator="depondt")
    assert _rel(e, ref) < 1e-11
    s, o = _pair((24, 20, 1), (0, 0, 2.0), lattice="triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3)
    e0 = s.init_spins("random", seed=2)
    sol = LLGSolver(s, alpha=0.05, dt=5e-4, temperature=20.0, seed=5, integrator="depondt", precession=precession)
    e, _, _ = sol.run(e0, 40, output_every=40)
    dm = s.device_model("fp64")
    normals = [dm.noise(5, n).cpu().numpy() for n in range(40)]
    ref = R.run_heun(o, e0.copy(), 40, 0.05, 5e-4, T=20.0, normals=lambda n: normals[n], precession=precession,
                     integrator="depondt")
    assert _rel(e, ref) < 1e-11
    # FP32 Depondt carries no more round-off than FP32 Heun on the same noise: its mean angle from
    # FP64 Depondt is within twice the FP32 Heun angle from FP64 Heun (plus 1e-7 rad)
    def angle(integrator):
        kw = dict(alpha=0.05, dt=5e-4, temperature=20.0, seed=5, integrator=integrator, precession=precession)
        e64, _, _ = LLGSolver(s, noise="float", **kw).run(e0, 40, output_every=40)
        e32, _, _ = LLGSolver(s, precision="fp32", **kw).run(e0, 40, output_every=40)
        return np.arccos(np.clip(np.sum(np.asarray(e32, float) * e64, 1), -1, 1)).mean()

    a_dep, a_heun = angle("depondt"), angle("heun")
    print(f"  FP32 mean angle: depondt {a_dep:.3e}  heun {a_heun:.3e}", flush=True)
    assert a_dep < 2 * a_heun + 1e-7


@pytest.mark.parametrize("integrator", ["heun", "depondt"])
def test_device_class_matches_solver(gpu, integrator):
    """The magex GPU-class shape (CudaLLG(s, 0.2, 1e-3, 0.0, 1, it).run(e0, 20), as magex's
    tests/cuda_sim_check.py drives it) gives LLGSolver's trajectory, and its callback sees the spins."""
    print(f"case: CudaLLG {integrator}", flush=True)
    from magex_amd.hip_backend import CudaLLG
    s, o = _pair((6, 6, 1), (0, 0, 1.0), lattice="square", J=[1.0, 0.2], D=0.3, K=0.05)
    e0 = s.init_spins("random", seed=4)
    seen = []
    eg = CudaLLG(s, 0.2, 1e-3, 0.0, 1, integrator).run(e0, 20, callback=lambda n, e: seen.append((n, e)), every=8)
    ref = R.run_heun(o, e0.copy(), 20, 0.2, 1e-3, integrator=integrator)
    assert _rel(eg, ref) < 1e-12
    assert [n for n, _ in seen] == [8, 16]
    e8, _, _ = LLGSolver(s, alpha=0.2, dt=1e-3, seed=1, integrator=integrator).run(e0, 8, output_every=8)
    assert np.array_equal(seen[0][1], e8)


def test_cubic_and_higher_uniaxial_terms_match_oracle(gpu):
    """CubicAnisotropy (rotated axes, per-site Kc) and UniaxialHigher (per-site K2, K3, tilted axis)
    folded into the kernels: gradient, energy and 60 thermal Heun steps against the oracle's terms."""
    print("case: cubic + higher uniaxial", flush=True)
    from scipy.spatial.transform import Rotation
    from magex_amd.terms import CubicAnisotropy, UniaxialHigher
    s, o = _pair((5, 4, 6), (0.1, 0.0, 0.4), lattice="sc", J=[1.0, 0.2], K=0.05, moment=2.0)
    rng = np.random.default_rng(8)
    C = Rotation.from_rotvec([0.2, 0.4, -0.1]).as_matrix()
    Kc, K2, K3 = rng.uniform(0.05, 0.3, s.N), rng.uniform(0.05, 0.2, s.N), rng.uniform(0.0, 0.1, s.N)
    s.terms += [CubicAnisotropy(Kc, C), UniaxialHigher(K2, K3, axis=(0.3, 0, 1))]
    o.terms += [R.CubicAnisotropy(Kc, C), R.UniaxialHigher(K2, K3, axis=(0.3, 0, 1))]
    e0 = s.init_spins("random", seed=3)
    assert _rel(s.gradient(e0), o.gradient(e0)) < 1e-13
    assert abs(s.energy(e0) - o.energy(e0)) < 1e-12 * abs(o.energy(e0))
    sol = LLGSolver(s, alpha=0.1, dt=2e-4, temperature=5.0, seed=21)
    e, _, _ = sol.run(e0, 60, output_every=60)
    dm = s.device_model("fp64")
    normals = [dm.noise(21, n).cpu().numpy() for n in range(60)]
    ref = R.run_heun(o, e0.copy(), 60, 0.1, 2e-4, T=5.0, normals=lambda n: normals[n])
    assert _rel(e, ref) < 1e-11


@pytest.mark.parametrize("integrator", ["heun", "depondt"])
def test_sot_and_field_func_match_oracle(gpu, integrator):
    """Spin-orbit torque fields and a uniform time-dependent field_func(t) (evaluated at t and t + dt),
    with pinned sites and thermal noise, against the oracle's LLGSolver.field restatement."""
    print(f"case: sot + field_func {integrator}", flush=True)
    s, o = _pair((10, 8, 1), (0, 0, 0.3), lattice="square", J=[2.0, 0.3], D=0.8, K=0.1)
    pin = np.zeros(s.N, bool)
    pin[::7] = True
    s.pinned = pin
    o.pinned = pin
    ff = lambda t: np.array([0.4 * np.cos(40.0 * t), 0.4 * np.sin(40.0 * t), 0.0])      # noqa: E731
    sot = {"p": [0.0, 1.0, 1.0], "B_DL": 0.6, "B_FL": -0.2}
    sol = LLGSolver(s, alpha=0.1, dt=5e-4, temperature=8.0, seed=13, integrator=integrator, sot=sot)
    sol.field_func = ff
    e0 = s.init_spins("random", seed=6)
    e, _, _ = sol.run(e0, 50, output_every=50)
    dm = s.device_model("fp64")
    normals = [dm.noise(13, n).cpu().numpy() for n in range(50)]
    p = np.array(sot["p"]) / np.linalg.norm(sot["p"])
    ref = R.run_heun(o, e0.copy(), 50, 0.1, 5e-4, T=8.0, normals=lambda n: normals[n], integrator=integrator,
                     sot=(p, sot["B_DL"], sot["B_FL"]), field_func=ff)
    assert np.array_equal(e[pin], e0[pin])
    assert _rel(e, ref) < 1e-11
    assert abs(sol.time - 50 * 5e-4) < 1e-15


def _spin_dependent_extra(A, mu):
    """A field that depends on the spins (that of an on-site tensor A, scaled) and a radial pulse."""
    def xf(e):
        return -1.5 * np.einsum("iab,ib->ia", A, np.asarray(e, float)) / (R.MU_B * mu[:, None])
    return xf


@pytest.mark.parametrize("integrator", ["heun", "depondt"])
@pytest.mark.parametrize("layout", ["one-site", "two-site"])
def test_extra_field_and_site_field_func_match_oracle(gpu, integrator, layout):
    """extra_field(e) (evaluated at the step's spins in the predictor and at the predictor spins in the
    corrector) and a site-dependent field_func(t) -> (N, 3), with spin-orbit torque, pinned sites and
    thermal noise, against the oracle within 1e-11 (FP64); LLGSolver.field includes both.  The two-site
    case runs through the internal sublattice-major order (permuted in and out of each stage)."""
    print(f"case: extra_field + (N,3) field_func {integrator} {layout}", flush=True)
    if layout == "one-site":
        s, o = _pair((10, 8, 1), (0, 0, 0.3), lattice="square", J=[2.0, 0.3], D=0.8, K=0.1)
    else:
        s, o = _pair((4, 4, 1), (0, 0, 0.3), lattice="honeycomb", J=[2.0, 0.3], K=0.1)
    rng = np.random.default_rng(9)
    pin = rng.random(s.N) < 0.12
    s.pinned = pin
    o.pinned = pin
    A = rng.normal(0.0, 0.2, (s.N, 3, 3))
    A = A + A.transpose(0, 2, 1)
    xf = _spin_dependent_extra(A, np.asarray(o.mu, float))
    prof = np.exp(-((np.arange(s.N) - s.N / 2) / (0.3 * s.N)) ** 2)
    ff = lambda t: np.outer(prof, [0.5 * np.cos(30.0 * t), 0.0, 0.2])                      # noqa: E731
    sot = {"p": [0.0, 1.0, 1.0], "B_DL": 0.4, "B_FL": -0.1}
    sol = LLGSolver(s, alpha=0.1, dt=5e-4, temperature=8.0, seed=17, integrator=integrator, sot=sot)
    sol.field_func = ff
    sol.extra_field = xf
    e0 = s.init_spins("random", seed=6)
    e, _, _ = sol.run(e0, 40, output_every=40)
    dm = s.device_model("fp64")
    normals = [dm.noise(17, n).cpu().numpy() for n in range(40)]
    p = np.array(sot["p"]) / np.linalg.norm(sot["p"])
    ref = R.run_heun(o, e0.copy(), 40, 0.1, 5e-4, T=8.0, normals=lambda n: normals[n], integrator=integrator,
                     sot=(p, sot["B_DL"], sot["B_FL"]), field_func=ff, extra_field=xf)
    assert np.array_equal(e[pin], e0[pin])
    assert _rel(e, ref) < 1e-11
    ref_field = R.solver_field(o, (p, sot["B_DL"], sot["B_FL"]), ff, xf)
    assert _rel(sol.field(e, 0.01), ref_field(e, 0.01)) < 1e-13
    # the device-tensor form of the callable gives the same trajectory
    sol2 = LLGSolver(s, alpha=0.1, dt=5e-4, temperature=8.0, seed=17, integrator=integrator, sot=sot)
    sol2.field_func = ff
    sol2.extra_field = lambda ed: gpu.as_tensor(xf(ed.cpu().numpy()), device="cuda")
    sol2.extra_field_device = True
    e2, _, _ = sol2.run(e0, 40, output_every=40)
    assert np.array_equal(e2, e)


@pytest.mark.parametrize("per_site_alpha", [False, True])
def test_solver_field_and_rhs_match_oracle(gpu, per_site_alpha):
    """LLGSolver.field(e, t) and .rhs(e, b, t) (magex/llg.py:124-150) with spin-orbit torque fields, a
    time-dependent field_func, pinned sites and a thermal field b, against the oracle's solver_field and
    llg_rhs; numpy in -> numpy out, device tensors stay on the device."""
    print(f"case: solver field/rhs per_site_alpha={per_site_alpha}", flush=True)
    s, o = _pair((9, 7, 1), (0.1, 0, 0.3), lattice="square", J=[2.0, 0.3], D=0.8, K=0.1)
    pin = np.zeros(s.N, bool)
    pin[::5] = True
    s.pinned = pin
    o.pinned = pin
    ff = lambda t: np.array([0.4 * np.cos(40.0 * t), 0.4 * np.sin(40.0 * t), 0.1])      # noqa: E731
    sot = {"p": [0.0, 1.0, 1.0], "B_DL": 0.6, "B_FL": -0.2}
    alpha = np.linspace(0.05, 0.3, s.N) if per_site_alpha else 0.12
    sol = LLGSolver(s, alpha=alpha, dt=5e-4, sot=sot)
    sol.field_func = ff
    p = np.array(sot["p"]) / np.linalg.norm(sot["p"])
    ref_field = R.solver_field(o, (p, sot["B_DL"], sot["B_FL"]), ff)
    e = s.init_spins("random", seed=8)
    b = np.random.default_rng(2).normal(size=(s.N, 3)) * 0.7
    for t in (0.0, 0.013):
        B = sol.field(e, t)
        assert isinstance(B, np.ndarray) and _rel(B, ref_field(e, t)) < 1e-13
        assert np.all(B[pin] == 0.0)
        a = np.asarray(alpha)[:, None] if per_site_alpha else alpha
        f = sol.rhs(e, b, t)
        assert _rel(f, R.llg_rhs(e, ref_field(e, t) + b, a)) < 1e-13
        fd = sol.rhs(gpu.as_tensor(e, device="cuda"), gpu.as_tensor(b, device="cuda"), t)
        assert fd.is_cuda and _rel(fd.cpu().numpy(), f) < 1e-15


def _structured_film():
    from magex_amd import structure
    case = json.load(open(os.path.join(GOLDEN, "structure_cases.json")))["film"]
    d = M.make_model(**case["lattice"])
    s = S.SpinSystem(S.load_results(d), case["L"], case["pbc"], case["field"])
    structure.apply_structure(s, case["cfg"])
    o = R.SpinSystem(R.load_results(d), case["L"], case["pbc"], case["field"])
    return s, o, case


def test_structured_film_matches_magex_and_oracle(gpu):
    """A model set up by apply_structure (two materials in layers, scaled moments and anisotropy, Ku, cubic
    and higher-uniaxial terms, Neel surface anisotropy, interface exchange as per-bond tensors, a pinned
    bottom layer with its direction, per-material damping and temperature under a temperature ramp):
    start vectors, B_eff and energy equal magex's; 10 Heun steps on injected noise within 1e-10 of magex's
    trajectory; 50 steps on the device generator's noise within 1e-11 of the oracle."""
    print("case: structured film", flush=True)
    import scipy.sparse as sp
    g = golden("structure_film.npz")
    s, o, case = _structured_film()
    e0 = s.init_spins("random", seed=3)
    assert np.array_equal(e0, g["e0"])
    assert _rel(s.effective_field(e0), g["B0"]) < 1e-13
    assert abs(s.energy(e0) - g["E0"]) < 1e-12 * abs(g["E0"])
    sol = LLGSolver(s, alpha=0.1, dt=2e-4, temperature=20.0, seed=0)
    ed = gpu.as_tensor(e0, device="cuda")
    sol.step_device(ed, 10, noise=gpu.as_tensor(g["noise"], device="cuda"))
    assert _rel(ed.cpu().numpy(), g["e10"]) < 1e-10
    # the oracle's system carries magex's arrays; the kernels' carry the drop-in's
    o.Jsp = sp.csr_matrix((g["J_val"], (g["J_row"], g["J_col"])), shape=(3 * o.N, 3 * o.N))
    o.A, o.mu, o.pinned, o.site_T = g["A"], g["mu"], g["pinned"], g["site_T"]
    o.terms += [R.CubicAnisotropy(g["Kc"], g["C"]), R.UniaxialHigher(g["K2"], g["K3"], axis=g["u"])]
    alpha = np.where(np.isnan(g["alpha_site"]), 0.1, g["alpha_site"])
    sol = LLGSolver(s, alpha=0.1, dt=2e-4, temperature=20.0, seed=31)
    e, _, _ = sol.run(e0, 50, output_every=50)
    dm = s.device_model("fp64")
    normals = [dm.noise(31, n).cpu().numpy() for n in range(50)]
    ref = R.run_heun(o, e0.copy(), 50, alpha, 2e-4, T=20.0, normals=lambda n: normals[n])
    assert _rel(e, ref) < 1e-11
    assert np.array_equal(e[g["pinned"]], e0[g["pinned"]])


def test_largest_cell_count_and_self_images(gpu):
    """Sizes at the edges of the index arithmetic and of the stencil:
    - 256^3 = 2^24 cells (the largest supercell class, where cell indices leave the float-division
      path): a spin spiral of period 8 gives, site for site, the field and the 5-step T=0 trajectory
      of the 8^3 supercell (translation invariance);
    - a supercell one cell thick on a periodic axis (every bond along it a self-image, folded into the
      on-site tensor) and two cells thick (both directions reach the same neighbour) against the oracle."""
    print("case: 256^3 cells and self images", flush=True)
    import torch
    from magex_amd.device import DeviceModel
    uc = S.load_results(M.make_model(lattice="sc", J=[1.0, 0.1], K=0.05, D=0.2, moment=2.0))
    small = S.SpinSystem(uc, (8, 8, 8), field=(0, 0, 0.4))
    big = S.SpinSystem(uc, (256, 256, 256), field=(0, 0, 0.4))
    c = small.cells.astype(float)
    q = 2 * np.pi / 8 * np.array([1.0, 2.0, 3.0])
    ph = c @ q
    es = np.stack([np.cos(ph) * 0.8, np.sin(ph) * 0.8, np.full_like(ph, 0.6)], 1)
    a = DeviceModel(small, "fp32")
    b = DeviceModel(big, "fp32")
    idx = torch.as_tensor(((big.cells % 8) @ np.array([64, 8, 1])).astype(np.int64), device="cuda")
    xs = torch.as_tensor(es, device="cuda", dtype=torch.float32)
    xb = xs[idx].contiguous()
    assert torch.equal(b.field(xb), a.field(xs)[idx])
    for m in (a, b):
        m.set_dynamics(2e-4, 0.1, 0.0)
    a.run(xs, 5, seed=1)
    b.run(xb, 5, seed=1)
    assert torch.equal(xb, xs[idx])
    del xb, b
    for L in ((1, 3, 4), (2, 3, 4)):
        s, o = _pair(L, (0.1, 0, 0.3), lattice="bcc_conventional", J=[20.0, 10.0], moment=2.2)
        e0 = s.init_spins("random", seed=5)
        assert _rel(s.effective_field(e0), o.effective_field(e0)) < 1e-13
        e, _, _ = LLGSolver(s, alpha=0.1, dt=1e-4).run(e0, 20, output_every=20)
        assert _rel(e, R.run_heun(o, e0.copy(), 20, 0.1, 1e-4)) < 1e-12
