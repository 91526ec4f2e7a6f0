"""Waits that give up, forced from the host with a wait limit of zero polls (every wait then gives up
after its first unsuccessful poll; every kernel still runs to its end).

- One launch per step (step mode 1): the launch's last workgroup finishes the steps whose waits gave up,
  so the spins are those of two launches per step, bit for bit.
- Persistent runs (step mode 2): a launch that gives up stops early and says so in the given-up count;
  LLGSolver.run and the adjoint's forward and backward then run the work again one launch per step
  (DeviceModel.checked) and return the bits of an undisturbed run.
"""
import numpy as np
import pytest
import torch

from magex_amd import models as M
from magex_amd import spinmodel as S

pytestmark = pytest.mark.gpu


def _tri(L):
    return S.SpinSystem(S.load_results(M.make_model("triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3)), (L, L, 1),
                        (True, True, True), (0, 0, 2.0))


def test_one_launch_step_finishes_given_up_waits(gpu):
    """sc 24^3 FP64 (54 workgroups, each waiting on the planes next to its own), 30 steps at T = 10 K."""
    print("case: wait limit 0, one launch per step, sc 24^3 fp64", flush=True)
    s = S.SpinSystem(S.load_results(M.make_model("sc", J=[1.0], K=0.1)), (24, 24, 24), (True, True, True),
                     (0, 0, 0.5))
    dm = s.device_model("fp64")
    dm.set_dynamics(1e-4, 0.1, 10.0, True, "heun")
    e0 = torch.as_tensor(s.init_spins("random", seed=3), device="cuda")
    out = {}
    for mode in (0, 1):
        dm.set_step_mode(mode)
        dm.set_wait_limit(0 if mode == 1 else 1 << 16)
        e = e0.clone()
        dm.run(e, 30, 7, 0)
        out[mode] = e
    form, given_up = dm.step_info()
    assert form == 1
    assert given_up > 0
    assert torch.equal(out[0], out[1])


def test_persistent_run_that_gives_up_is_run_again(gpu):
    """Triangular 48^2 FP32 (three shells + DMI, persistent launches), 300 steps at T = 20 K."""
    print("case: wait limit 0, persistent runs, tri 48^2 fp32", flush=True)
    from magex_amd.llg import LLGSolver
    e0 = _tri(48).init_spins("random", seed=1)
    res = {}
    for limit in (1 << 16, 0):
        s = _tri(48)
        sol = LLGSolver(s, alpha=0.05, dt=5e-4, temperature=20.0, seed=3, precision="fp32")
        dm = sol._sync()
        dm.set_wait_limit(limit)
        assert dm.step_info()[0] == 2
        e, table, _ = sol.run(e0, 300, output_every=100)
        res[limit] = (e, table, dm.step_info()[1])
    assert res[1 << 16][2] == 0
    assert res[0][2] > 0
    assert np.array_equal(res[0][0], res[1 << 16][0])
    assert np.array_equal(res[0][1], res[1 << 16][1])
    # magex's GPU-class shape runs its steps the same way
    from magex_amd.hip_backend import CudaLLG
    ref = CudaLLG(_tri(48), 0.05, 5e-4, 20.0, 3, precision="fp32").run(e0, 150)
    g = CudaLLG(_tri(48), 0.05, 5e-4, 20.0, 3, precision="fp32")
    g.sol._sync().set_wait_limit(0)
    assert np.array_equal(g.run(e0, 150), ref)
    assert g.sol.dm.step_info()[1] > 0
    # the unchecked call reports the stopped launch
    s = _tri(48)
    dm = s.device_model("fp32")
    dm.set_dynamics(5e-4, 0.05, 20.0, True, "heun")
    dm.set_wait_limit(0)
    e = torch.as_tensor(e0, device="cuda", dtype=torch.float32)
    dm.run(e, 200, 3, 0)
    assert dm.step_info()[1] > 0


def test_adjoint_that_gives_up_is_run_again(gpu):
    """Square 48^2 FP64 adjoint (persistent forward, persistent sweeps beside persistent replays), 60
    steps in segments of 10 at T = 30 K: forward spins and gradients equal an undisturbed run's."""
    print("case: wait limit 0, adjoint, square 48^2 fp64", flush=True)
    from magex_amd import configs
    from magex_amd.adjoint import DifferentiableLLG
    th0 = np.array([4.0, -0.4, 1.0, 0.2])
    build = lambda th: configs.square128(*th, L=48)[0]          # noqa: E731
    e0 = build(th0).init_spins("random", seed=40)
    tgt = torch.as_tensor(build(th0).init_spins("random", seed=41), device="cuda")
    res = {}
    for limit in (1 << 16, 0):
        sol = DifferentiableLLG(build, th0, alpha=0.3, dt=1e-3, temperature=30.0, seed=4, precision="fp64",
                                checkpoint_every=10)
        sol.dm.set_wait_limit(limit)
        assert sol.dm.step_info()[0] == 2
        e = sol.forward(th0, e0, 60)
        gt, g0 = sol.backward(2.0 * (e - tgt) / e.numel())
        torch.cuda.synchronize()
        res[limit] = ([x.cpu().numpy() for x in (e, gt, g0)], sol.dm.step_info()[1])
    assert res[1 << 16][1] == 0
    assert res[0][1] > 0
    for a, b in zip(res[0][0], res[1 << 16][0]):
        assert np.array_equal(a, b), float(np.abs(a - b).max())
