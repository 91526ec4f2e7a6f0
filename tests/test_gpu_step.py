"""The one-launch Heun step (k_step) and the persistent multi-step launch (k_steps) against the
two-launch step (k_heun pass 1 + pass 2), bit for bit.

All three compute the same arithmetic in the same order, so every case must agree exactly, over many
steps, for every coupling form, geometry and model feature the one-launch kernels accept; and no wait
of a one-launch kernel may time out in normal operation.
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
                 graph=None, sot=None, forms=(0, 1, 2), modes=(0, 1, 2)):
    dm = s.device_model(precision)
    if graph is not None:
        dm.set_graph_steps(graph)
    if sot is not None:
        dm.set_sot(*sot)
    dm.set_dynamics(dt, alpha, T, True, integrator)
    e0 = torch.as_tensor(s.init_spins("random", seed=2), device="cuda", dtype=dm.dtype).contiguous()
    nz = None if noise is None else torch.as_tensor(noise, device="cuda", dtype=dm.dtype).contiguous()
    out = {}
    for mode, form_want in zip(modes, forms):
        dm.set_step_mode(mode)
        e = e0.clone()
        dm.run(e, steps, seed, 0, nz)
        out[mode] = e
        form, given_up = dm.step_info()
        assert form == form_want, f"step mode {mode}: the model runs in form {form}"
        assert given_up == 0
    assert torch.isfinite(out[modes[1]]).all()
    for mode in modes[1:]:
        assert torch.equal(out[modes[0]], out[mode]), (mode, float((out[modes[0]] - out[mode]).abs().max()))
    return out[modes[1]]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("graph", [0, 8])
def test_triangular_three_shells_dmi(gpu, precision, graph):
    print(f"case: tri 16^2 width-4 table {precision} graph {graph}", flush=True)
    s = _system((16, 16, 1), (0, 0, 2.0), lattice="triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3)
    _modes_agree(gpu, s, precision, 40, 5e-4, 0.05, 20.0, graph=graph)


@pytest.mark.parametrize("case", ["tri3_periodic", "tri3_open", "tri2_open", "sc2_jani"])
def test_two_batch_tables_fp32(gpu, case):
    """FP32 table stencils of 10 to 18 entries (two batches, both with held record indices in
    persistent launches), width-4 and width-9 tables, periodic and open axes."""
    print(f"case: two-batch table {case}", flush=True)
    if case == "sc2_jani":
        s = _system((6, 5, 4), (0.1, 0, 0.3), lattice="sc", J=[1.0, 0.3],
                    Jani=[[0.1, 0.02, 0], [0.02, -0.05, 0], [0, 0, -0.05]], D=0.1, K=0.05)
    else:
        pbc = (True, True, True) if case == "tri3_periodic" else (False, True, False)
        J = [5.0, -0.5, 0.2] if case.startswith("tri3") else [1.0, 0.2]
        s = _system((23, 19, 1), (0, 0, 2.0), pbc, lattice="triangular", J=J, D=1.5, K=0.3)
    _modes_agree(gpu, s, "fp32", 60, 5e-4, 0.05, 20.0)


def test_bcc_per_bond_two_held_batches(gpu):
    """Two sublattices, 14 per-bond tensors per site (both held batches), per-site SIA; generated noise
    (the plain-model kernel) and injected noise."""
    print("case: bcc per-bond fp32", flush=True)
    s, g = _bcc_perbond()
    _modes_agree(gpu, s, "fp32", 30, 1e-4, 0.1, 300.0)
    noise = np.random.default_rng(4).normal(size=(10, s.N, 3))
    _modes_agree(gpu, s, "fp32", 10, 1e-4, 0.1, 300.0, noise=noise)
    alpha = 0.1 + 0.05 * np.random.default_rng(6).random(s.N)      # the general (non-lean) kernels
    _modes_agree(gpu, s, "fp32", 10, 1e-4, alpha, 300.0)


def test_bcc_per_bond_fp64_persistent(gpu):
    """FP64 per-bond tensors: two launches per step in mode 1 (tensors not held), persistent launches in
    mode 2; per-site damping takes the general kernel."""
    print("case: bcc per-bond fp64", flush=True)
    s, g = _bcc_perbond()
    _modes_agree(gpu, s, "fp64", 30, 1e-4, 0.1, 300.0, forms=(0, 0, 2))
    alpha = 0.1 + 0.05 * np.random.default_rng(6).random(s.N)
    noise = np.random.default_rng(7).normal(size=(12, s.N, 3))
    _modes_agree(gpu, s, "fp64", 12, 1e-4, alpha, 300.0, noise=noise, forms=(0, 0, 2))


def _bcc_large(pbc0=True):
    """bcc 64 x 64 x 32 with random per-bond tensors and SIA (262 144 spins): 1024 tiles of 256 sites,
    twice what the GPU holds at once in the walked step, so every workgroup takes a second ticket; axis-0
    planes of 2048 cells (one tile per XCD per plane and sublattice)."""
    from magex_amd import configs
    s = S.SpinSystem(S.load_results(M.make_model("bcc_conventional", a=2.87, J=[20.0, 10.0], moment=2.2)),
                     (64, 64, 32), (pbc0, True, True))
    s.set_bond_tensors(configs.random_bond_tensors(s, seed=3))
    s.A = configs.random_sia(s.N, seed=4)
    return s


def test_walked_step_on_a_grid_larger_than_the_gpu_holds(gpu):
    """Step mode 3: the walked one-launch step (ticket-ordered tiles, tensors held in registers from the
    predictor to the corrector) equals two launches bit for bit, with generated and injected noise, the
    general kernel (per-site damping, pinned sites, Depondt) and an open axis 0, and no wait gives up."""
    print("case: bcc 64x64x32 per-bond fp32, walked step", flush=True)
    s = _bcc_large()
    _modes_agree(gpu, s, "fp32", 12, 1e-4, 0.1, 300.0, forms=(0, 3), modes=(0, 3))
    noise = np.random.default_rng(9).normal(size=(4, s.N, 3))
    _modes_agree(gpu, s, "fp32", 4, 1e-4, 0.1, 300.0, noise=noise, forms=(0, 3), modes=(0, 3))
    rng = np.random.default_rng(10)
    s.pinned = rng.random(s.N) < 0.05
    s.touch()
    alpha = 0.1 + 0.05 * rng.random(s.N)
    _modes_agree(gpu, s, "fp32", 6, 1e-4, alpha, 300.0, "depondt", forms=(0, 3), modes=(0, 3))
    o = _bcc_large(pbc0=False)
    _modes_agree(gpu, o, "fp32", 6, 1e-4, 0.1, 300.0, forms=(0, 3), modes=(0, 3))


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


def test_persistent_runs_span_launches_and_continue(gpu):
    """A run longer than one persistent launch (1024 steps) and a run split in two (step numbers carry
    on) equal the two-launch steps."""
    print("case: sc 8^3 fp64 persistent 1500 steps, split 700 + 800", flush=True)
    s = _system((8, 8, 8), (0, 0, 0.5), lattice="sc", J=[1.0, 0.1], K=0.05)
    dm = s.device_model("fp64")
    dm.set_dynamics(2e-4, 0.1, 20.0, True, "heun")
    e0 = gpu.as_tensor(s.init_spins("random", seed=3), device="cuda").contiguous()
    dm.set_step_mode(0)
    ref = e0.clone()
    dm.run(ref, 1500, 11, 4)
    dm.set_step_mode(2)
    one = e0.clone()
    dm.run(one, 1500, 11, 4)
    two = e0.clone()
    dm.run(two, 700, 11, 4)
    dm.run(two, 800, 11, 704)
    assert dm.step_info() == (2, 0)
    assert gpu.equal(one, ref), float((one - ref).abs().max())
    assert gpu.equal(two, ref), float((two - ref).abs().max())


def test_launch_forms_follow_the_measured_policy(gpu):
    """Default mode: persistent launches where a round of polls is small (sc 16^3 FP64, triangular 256^2
    FP32), two launches per step for triangular 256^2 FP64 (38 MB of records per poll round) and for
    grids larger than the GPU holds at once (bcc 64^3)."""
    from magex_amd import configs
    print("case: launch forms of the headline configurations", flush=True)
    for name, prec, form in (("sc16", "fp64", 2), ("tri256", "fp32", 2), ("tri256", "fp64", 0),
                             ("bcc64", "fp32", 0)):
        s, p = getattr(configs, name)()
        dm = s.device_model(prec)
        dm.set_dynamics(p["dt"], p["alpha"], p["temperature"], True, "heun")
        assert dm.step_info()[0] == form, (name, prec)
    s, p = configs.tri256()
    dm = s.device_model("fp64")
    dm.set_step_mode(1)
    assert dm.step_info()[0] == 0     # FP64 beyond 16 k sites keeps two launches per step
    s, p = configs.bcc64()
    dm = s.device_model("fp32")
    dm.set_dynamics(p["dt"], p["alpha"], p["temperature"], True, "heun")
    dm.set_step_mode(3)
    assert dm.step_info()[0] == 3     # the walked step only when asked for (slower on this grid)
    dm = s.device_model("fp64")
    dm.set_dynamics(p["dt"], p["alpha"], p["temperature"], True, "heun")
    dm.set_step_mode(3)
    assert dm.step_info()[0] == 0     # FP64 tensors are not held: two launches
