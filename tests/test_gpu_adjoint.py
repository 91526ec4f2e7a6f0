"""Parity of the differentiable form: the library's adjoint against the FP64 torch.autograd restatement
of the same trajectory (oracle/torch_heun.py) and against central finite differences of the library's
own forward.

Tolerances: FP64 gradients match autograd to 1e-9 relative (round-off of two different evaluation
orders through a few hundred steps) and central differences with h = 1e-4 meV to 1e-6 relative; the
FP32 gradient of the square 128^2 workload stays within 2e-3 of the FP64 one (its trajectory carries
FP32 round-off through 500 steps).
"""
import numpy as np
import pytest
import torch

from oracle import magex_ref as R
from oracle.torch_heun import TorchHeun

pytestmark = pytest.mark.gpu

THETA0 = np.array([4.0, -0.4, 1.0, 0.2])        # J1, J2, D, K (meV)
THETA_T = np.array([5.0, -0.5, 1.5, 0.3])


def _builders(L):
    from magex_amd import configs
    dev = lambda th: configs.square128(*th, L=L)[0]                                     # noqa: E731
    ref = lambda th: R.SpinSystem(R.load_results(R.make_model("square", J=[th[0], th[1]], D=th[2], K=th[3])),
                                  (L, L, 1))                                           # noqa: E731
    return dev, ref


def _solver(L, precision="fp64", C=50, T=0.0, seed=0, alpha=0.3, dt=1e-3):
    from magex_amd.adjoint import DifferentiableLLG
    dev, _ = _builders(L)
    return DifferentiableLLG(dev, THETA0, alpha=alpha, dt=dt, temperature=T, seed=seed, precision=precision,
                             checkpoint_every=C)


def _mse_grad(e, tgt):
    return 2.0 * (e - tgt) / e.numel()


def _grads(sol, theta, e0, tgt, steps):
    e = sol.forward(theta, e0, steps)
    gt, g0 = sol.backward(_mse_grad(e, tgt))
    torch.cuda.synchronize()
    return e, gt.cpu().numpy(), g0.double().cpu().numpy()


def _relerr(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def test_adjoint_matches_fp64_autograd(gpu):
    print("case: adjoint vs autograd, square 16^2, 120 steps, C=50", flush=True)
    L, steps = 16, 120
    sol = _solver(L, C=50)
    _, ref_build = _builders(L)
    o = ref_build(THETA0)
    e0 = o.init_random(0)
    tgt = torch.as_tensor(o.init_random(1))
    e, gt, g0 = _grads(sol, THETA0, e0, tgt.cuda(), steps)
    Th = TorchHeun(ref_build, THETA0, alpha=0.3, dt=1e-3)
    t = torch.tensor(THETA0, requires_grad=True)
    x0 = torch.tensor(e0, requires_grad=True)
    eT = Th.run(t, x0, steps)
    ((eT - tgt) ** 2).mean().backward()
    assert np.abs(e.cpu().numpy() - eT.detach().numpy()).max() < 1e-12
    assert _relerr(gt, t.grad.numpy()) < 1e-9, (gt, t.grad)
    assert _relerr(g0, x0.grad.numpy()) < 1e-9


def test_adjoint_thermal_matches_autograd(gpu):
    """T > 0: the gradient of one sample path, the thermal field read back from the device generator."""
    print("case: adjoint thermal, square 12^2, 40 steps, C=16", flush=True)
    L, steps, T, seed = 12, 40, 50.0, 7
    sol = _solver(L, C=16, T=T, seed=seed)
    _, ref_build = _builders(L)
    o = ref_build(THETA0)
    e0 = o.init_random(2)
    tgt = torch.as_tensor(o.init_random(3))
    e, gt, g0 = _grads(sol, THETA0, e0, tgt.cuda(), steps)
    sg = R.sigma(o, 0.3, T, 1e-3)
    fields = [torch.as_tensor(sg[:, None] * sol.dm.noise(seed, n).double().cpu().numpy()) for n in range(steps)]
    Th = TorchHeun(ref_build, THETA0, alpha=0.3, dt=1e-3)
    t = torch.tensor(THETA0, requires_grad=True)
    x0 = torch.tensor(e0, requires_grad=True)
    eT = Th.run(t, x0, steps, thermal=lambda n: fields[n])
    ((eT - tgt) ** 2).mean().backward()
    assert np.abs(e.cpu().numpy() - eT.detach().numpy()).max() < 1e-12
    assert _relerr(gt, t.grad.numpy()) < 1e-9, (gt, t.grad)
    assert _relerr(g0, x0.grad.numpy()) < 1e-9


def test_adjoint_full_tensor_gradient(gpu):
    """Ten parameters in one backward: the nearest-neighbour coupling as J1 I plus a symmetric
    anisotropic tensor (six components), DMI, anisotropy and J2.  The FP64 adjoint matches the FP64
    autograd restatement to 1e-9 and central differences of the library's forward to 1e-6."""
    print("case: adjoint with 10 parameters (full NN tensor), square 12^2, 80 steps", flush=True)
    from magex_amd.adjoint import DifferentiableLLG
    from magex_amd import models as Md
    from magex_amd import spinmodel as Sd
    L, steps, h = 12, 80, 1e-4

    def jani(th):
        return [[th[1], th[4], th[5]], [th[4], th[2], th[6]], [th[5], th[6], th[3]]]

    def dev(th):
        return Sd.SpinSystem(Sd.load_results(Md.make_model("square", J=[th[0], th[9]], D=th[7], K=th[8],
                                                           Jani=jani(th))), (L, L, 1))

    def ref(th):
        return R.SpinSystem(R.load_results(R.make_model("square", J=[th[0], th[9]], D=th[7], K=th[8],
                                                        Jani=jani(th))), (L, L, 1))

    th0 = np.array([4.0, 0.3, -0.2, 0.1, 0.15, -0.05, 0.08, 1.0, 0.2, -0.4])
    sol = DifferentiableLLG(dev, th0, alpha=0.3, dt=1e-3, precision="fp64", checkpoint_every=20)
    o = ref(th0)
    e0 = o.init_random(4)
    tgt = torch.as_tensor(o.init_random(5))
    e, gt, g0 = _grads(sol, th0, e0, tgt.cuda(), steps)
    Th = TorchHeun(ref, th0, alpha=0.3, dt=1e-3)
    t = torch.tensor(th0, requires_grad=True)
    x0 = torch.tensor(e0, requires_grad=True)
    ((Th.run(t, x0, steps) - tgt) ** 2).mean().backward()
    assert _relerr(gt, t.grad.numpy()) < 1e-9, (gt, t.grad)
    assert _relerr(g0, x0.grad.numpy()) < 1e-9

    def loss(th):
        x = sol.forward(th, e0, steps)
        return float(((x - tgt.cuda()) ** 2).mean())

    for p in (1, 4, 6):
        d = np.zeros(len(th0))
        d[p] = h
        fd = (loss(th0 + d) - loss(th0 - d)) / (2 * h)
        assert abs(gt[p] - fd) <= 1e-6 * abs(fd), (p, gt[p], fd)


def test_adjoint_checkpoint_interval_is_exact(gpu):
    """Replays from checkpoints recompute the forward bit for bit, so the interval does not change the
    gradient at all (segments of 50, 7 and one segment of the whole run)."""
    print("case: adjoint checkpoint intervals, square 20^2, 100 steps", flush=True)
    L, steps = 20, 100
    _, ref_build = _builders(L)
    o = ref_build(THETA0)
    e0 = o.init_random(4)
    tgt = torch.as_tensor(o.init_random(5)).cuda()
    sols = [_solver(L, C=C) for C in (50, 7, 1000)]
    out = [_grads(sol, THETA0, e0, tgt, steps) for sol in sols]
    for e, gt, g0 in out[1:]:
        assert torch.equal(e, out[0][0])
        assert np.array_equal(gt, out[0][1])
        assert np.array_equal(g0, out[0][2])
    # a second backward of the same forward replays what the first one consumed
    for sol in sols[:2]:
        gt, g0 = sol.backward(_mse_grad(out[0][0], tgt))
        assert np.array_equal(gt.cpu().numpy(), out[0][1])
        assert np.array_equal(g0.double().cpu().numpy(), out[0][2])
    # zero steps: the gradient of the loss itself
    e = sols[0].forward(THETA0, e0, 0)
    gt, g0 = sols[0].backward(_mse_grad(e, tgt))
    assert np.array_equal(gt.cpu().numpy(), np.zeros(4))
    assert torch.equal(g0, _mse_grad(e, tgt))


@pytest.mark.parametrize("integrator", ["heun", "depondt"])
def test_recorded_persistent_replays_equal_single_pass_replays(gpu, integrator):
    """Segments of two or more undriven steps are recorded by the persistent launch (every record in one
    launch) and swept by the persistent adjoint launch; one-step segments by the single-pass kernels.  Both
    must give the same records and sweeps bit for bit, so the gradients with checkpoints every step and
    every 16 steps are equal -- on the general model:
    pinned sites, per-site damping, T = 50 K, the dense gradients too."""
    print(f"case: recorded persistent vs single-pass replays, square 16^2, {integrator}", flush=True)
    from magex_amd.adjoint import DifferentiableLLG
    L, steps, T, seed = 16, 40, 50.0, 9
    dev0, ref0 = _builders(L)
    mask = np.zeros(L * L, bool)
    mask[[3, 40, 77, 200]] = True
    alpha = 0.2 + 0.2 * np.random.default_rng(3).random(L * L)

    def dev(th):
        s = dev0(th)
        s.pinned = mask
        return s

    out = []
    for C in (1, 16):
        sol = DifferentiableLLG(dev, THETA0, alpha=alpha, dt=1e-3, temperature=T, seed=seed, precision="fp64",
                                checkpoint_every=C, integrator=integrator)
        o = ref0(THETA0)
        e0 = o.init_random(21)
        tgt = torch.as_tensor(o.init_random(22)).cuda()
        e = sol.forward(THETA0, e0, steps)
        gt, g0, gb, ga = sol.backward_dense(_mse_grad(e, tgt))
        torch.cuda.synchronize()
        out.append([x.cpu().numpy() for x in (e, gt, g0, gb, ga)])
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b), float(np.abs(a - b).max())


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_persistent_sweep_equals_per_pass_sweep(gpu, precision):
    """On a model that takes the persistent forward launch (step mode 2), every segment of two or more
    steps is swept by one persistent launch; in step mode 1 every step runs three kernels.  Both give the
    same gradients bit for bit: the bcc 4^3 per-bond model (28 stencil entries, per-bond tensors, per-site
    anisotropy, three parameters) with pinned sites, per-site damping and T = 30 K, segments of 16 steps
    and a last one of 9; parameter, spin and dense gradients undriven, field gradients under a drive."""
    print(f"case: persistent vs per-pass adjoint sweep, bcc 4^3 per-bond, {precision}", flush=True)
    from magex_amd.adjoint import DifferentiableLLG
    from magex_amd import models as Md
    from magex_amd import spinmodel as Sd
    L, steps, th0 = 4, 41, np.array([20.0, 10.0, 0.5])

    def dev0(th):
        return Sd.SpinSystem(Sd.load_results(Md.make_model("bcc_conventional", a=2.87, J=[th[0], th[1]], K=th[2],
                                                           moment=2.2)), (L, L, L))

    dev1, _, _ = _disorder(dev0, dev0, th0, seed=23, sym_A=True)
    N = 2 * L ** 3
    mask = np.zeros(N, bool)
    mask[[1, 30, 77]] = True
    alpha = 0.15 + 0.2 * np.random.default_rng(4).random(N)

    def dev(th):
        s = dev1(th)
        s.pinned = mask
        return s

    out = {}
    for mode in (1, 2):
        sol = DifferentiableLLG(dev, th0, alpha=alpha, dt=1e-3, temperature=30.0, seed=5, precision=precision,
                                checkpoint_every=16)
        sol.dm.set_step_mode(mode)
        assert (sol.dm.step_info()[0] == 2) == (mode == 2)
        e0 = sol.system.init_spins("random", seed=12)
        tgt = torch.as_tensor(sol.system.init_spins("random", seed=13), device="cuda",
                              dtype=torch.float64 if precision == "fp64" else torch.float32)
        e = sol.forward(th0, e0, steps)
        res = [e, *sol.backward_dense(_mse_grad(e, tgt))]
        B = sol.field_sequence(lambda t: np.array([2.0 * np.sin(30.0 * t), 0.3, 1.0]), steps)
        e = sol.forward(th0, e0, steps, field=B)
        res += [e, *sol.backward_field(_mse_grad(e, tgt))]
        torch.cuda.synchronize()
        assert sol.dm.step_info()[1] == 0
        out[mode] = [x.cpu().numpy() for x in res]
    for k, (a, b) in enumerate(zip(out[1], out[2])):
        assert np.array_equal(a, b), (k, float(np.abs(a - b).max()))


def test_persistent_sweep_longer_than_one_launch(gpu):
    """A segment longer than one persistent launch (1024 steps): the replay and the sweep each split into
    two launches (lambda and the partial rows carried through memory, a fresh tag epoch each), and the
    parameter terms of all 2200 passes are summed after the sweep.  The gradients equal step mode 1's
    (one kernel per pass) bit for bit, with no poll given up; square 8^2, 1100 steps, one segment."""
    print("case: persistent sweep over 1100 steps in one segment, square 8^2", flush=True)
    L, steps = 8, 1100
    _, ref_build = _builders(L)
    o = ref_build(THETA0)
    e0 = o.init_random(30)
    tgt = torch.as_tensor(o.init_random(31)).cuda()
    out = {}
    for mode in (1, 2):
        sol = _solver(L, C=steps, T=20.0, seed=3)
        sol.dm.set_step_mode(mode)
        e, gt, g0 = _grads(sol, THETA0, e0, tgt, steps)
        assert sol.dm.step_info()[1] == 0
        out[mode] = (e.cpu().numpy(), gt, g0)
    for a, b in zip(out[1], out[2]):
        assert np.array_equal(a, b), float(np.abs(a - b).max())


def test_overlapped_backward_on_a_side_stream(gpu):
    """The overlapped backward (replays on the handle's own stream, ordered with events) issued on a torch
    side stream gives the same bits as on the default stream, twice in a row (square 16^2, 120 steps in
    segments of 10, T = 30 K)."""
    print("case: overlapped backward on a torch side stream, square 16^2", flush=True)
    L, steps = 16, 120
    _, ref_build = _builders(L)
    o = ref_build(THETA0)
    e0 = o.init_random(40)
    tgt = torch.as_tensor(o.init_random(41)).cuda()
    sol = _solver(L, C=10, T=30.0, seed=4)
    ref = _grads(sol, THETA0, e0, tgt, steps)
    side = torch.cuda.Stream()          # default priority
    for _ in range(2):
        with torch.cuda.stream(side):
            e = sol.forward(THETA0, e0, steps)
            gt, g0 = sol.backward(_mse_grad(e, tgt))
        side.synchronize()
        assert torch.equal(e, ref[0])
        assert np.array_equal(gt.cpu().numpy(), ref[1])
        assert np.array_equal(g0.double().cpu().numpy(), ref[2])
    assert sol.dm.step_info()[1] == 0


def test_adjoint_forward_is_the_plain_trajectory(gpu):
    """mx_adjoint_forward at theta is the plain solver on the builder's system at theta."""
    print("case: adjoint forward vs LLGSolver, square 32^2", flush=True)
    from magex_amd.llg import LLGSolver
    L = 32
    dev, _ = _builders(L)
    sol = _solver(L)
    s = dev(THETA_T)
    e0 = s.init_spins("random", seed=6)
    e = sol.forward(THETA_T, e0, 60).cpu().numpy()
    ref = LLGSolver(s, 0.3, 1e-3, 0.0, 0, precision="fp64")
    x = torch.as_tensor(e0, device="cuda")
    ref.step_device(x, 60)
    assert np.abs(e - x.cpu().numpy()).max() < 1e-12


def test_adjoint_square128_matches_central_differences(gpu):
    """The differentiation workload: square 128^2, 500 damped steps, checkpoints every 50, MSE to the
    texture of the target constants from the same start; FP64 gradient vs central differences
    (h = 1e-4 meV) of the library's forward, to 1e-6 relative per parameter."""
    print("case: adjoint square128 vs central differences (FP64)", flush=True)
    from magex_amd import configs
    L, steps, h = 128, 500, 1e-4
    sol = _solver(L, C=50)
    s0 = configs.square128(L=L)[0]
    e0 = s0.init_spins("random", seed=0)
    tgt = sol.forward(THETA_T, e0, steps)

    def loss(th):
        e = sol.forward(th, e0, steps)
        return float(((e - tgt) ** 2).mean())

    e, gt, g0 = _grads(sol, THETA0, e0, tgt, steps)
    for p in range(4):
        d = np.zeros(4)
        d[p] = h
        fd = (loss(THETA0 + d) - loss(THETA0 - d)) / (2 * h)
        print(f"  theta[{p}]: adjoint {gt[p]:.12e}  central difference {fd:.12e}", flush=True)
        assert abs(gt[p] - fd) <= 1e-6 * abs(fd), (p, gt[p], fd)
    # two components of the initial spins (e0 is not renormalized by the solver's first step)
    for site, comp in ((100, 0), (9000, 2)):
        def loss_e(dx):
            x = torch.as_tensor(e0, device="cuda").clone()
            x[site, comp] += dx
            e = sol.forward(THETA0, x, steps)
            return float(((e - tgt) ** 2).mean())
        fd = (loss_e(h) - loss_e(-h)) / (2 * h)
        assert abs(g0[site, comp] - fd) <= 1e-6 * abs(fd) + 1e-14, (site, comp, g0[site, comp], fd)


def test_adjoint_fp32_tracks_fp64(gpu):
    print("case: adjoint square128 FP32 vs FP64", flush=True)
    from magex_amd import configs
    L, steps = 128, 500
    e0 = configs.square128(L=L)[0].init_spins("random", seed=0)
    res = {}
    for prec in ("fp64", "fp32"):
        sol = _solver(L, precision=prec, C=50)
        tgt = sol.forward(THETA_T, e0, steps)
        e = sol.forward(THETA0, e0, steps)
        gt, _ = sol.backward(_mse_grad(e, tgt))
        res[prec] = gt.cpu().numpy()
    print(f"  fp64 {res['fp64']}\n  fp32 {res['fp32']}", flush=True)
    assert _relerr(res["fp32"], res["fp64"]) < 2e-3


def test_autograd_function(gpu):
    """trajectory() as a torch.autograd.Function: FP32 leaf parameters, loss.backward(), and a backward
    that follows another forward (which re-runs its own forward first)."""
    print("case: autograd Function, square 16^2", flush=True)
    L, steps = 16, 80
    sol = _solver(L, C=20)
    _, ref_build = _builders(L)
    o = ref_build(THETA0)
    e0 = torch.as_tensor(o.init_random(8), device="cuda")
    tgt = torch.as_tensor(o.init_random(9), device="cuda")
    theta = torch.tensor(THETA0, dtype=torch.float32, requires_grad=True)
    x0 = e0.clone().requires_grad_(True)
    e = sol.trajectory(theta, x0, steps)
    loss = ((e - tgt) ** 2).mean()
    sol.forward(THETA_T, e0, steps)             # an unrelated forward in between
    loss.backward()
    _, gt, g0 = _grads(sol, theta.detach().double().numpy(), e0, tgt, steps)
    assert theta.grad.dtype == torch.float32
    assert _relerr(theta.grad.double().numpy(), gt) < 1e-6
    assert _relerr(x0.grad.cpu().numpy(), g0) < 1e-12


def _forward_pairs(system):
    """(src, dst) magex site arrays of every forward stencil entry (s < t, or s == t and R > 0), in
    stencil order: the bond's tensor acts at block (src, dst) of the coupling matrix and its transpose
    at (dst, src).  dst = -1 where an open axis cuts the bond."""
    L = np.array(system.L)
    pbc = np.array(system.pbc, bool)
    ns = system.uc.nsites
    cells = np.stack(np.meshgrid(*[np.arange(n) for n in L], indexing="ij"), -1).reshape(-1, 3)
    out = []
    for s, t, Rv, _ in system.stencil:
        if not (s < t or (s == t and tuple(Rv) > (0, 0, 0))):
            continue
        tgt = cells + np.asarray(Rv)
        ok = np.all(pbc | ((tgt >= 0) & (tgt < L)), axis=1)
        tgt = tgt % L
        cn = (tgt[:, 0] * L[1] + tgt[:, 1]) * L[2] + tgt[:, 2]
        out.append((np.arange(len(cells)) * ns + s, np.where(ok, cn * ns + t, -1)))
    return out


def _bond_grads(G, pairs):
    """dL/dT (nbonds, ncell, 3, 3) from the coupling-matrix gradient G[i, j] (N, N, 3, 3); zero on cut bonds."""
    out = []
    for i, j in pairs:
        jj = np.maximum(j, 0)
        g = G[i, jj] + G[jj, i].transpose(0, 2, 1)
        out.append(np.where((j >= 0)[:, None, None], g, 0.0))
    return np.stack(out)


def test_adjoint_dense_gradients_match_autograd(gpu):
    """dL/dT per bond and dL/dA per site (FP64, square 8^2 with J2, DMI and anisotropy, 60 steps over
    three checkpoint segments) against FP64 autograd of the dense coupling matrix and on-site tensors:
    dL/dT_b = G(i, j) + G(j, i)^T from the matrix gradient G, to 1e-9 of the largest element.  The
    parameter gradient of the same call equals mx_adjoint_backward's bit for bit."""
    print("case: dense adjoint gradients vs autograd, square 8^2, 60 steps, C=25", flush=True)
    from magex_amd.adjoint import dmi_gradient
    L, steps = 8, 60
    sol = _solver(L, C=25)
    _, ref_build = _builders(L)
    o = ref_build(THETA0)
    e0 = o.init_random(6)
    tgt = torch.as_tensor(o.init_random(7))
    e = sol.forward(THETA0, e0, steps)
    gt, g0, gb, ga = sol.backward_dense(_mse_grad(e, tgt.cuda()))
    gt_plain, g0_plain = sol.backward(_mse_grad(e, tgt.cuda()))
    torch.cuda.synchronize()
    assert torch.equal(gt, gt_plain) and torch.equal(g0, g0_plain)
    Th = TorchHeun(ref_build, THETA0, alpha=0.3, dt=1e-3)
    J, A = Th.dense(THETA0)
    J = J.clone().requires_grad_(True)
    A = A.clone().requires_grad_(True)
    ((Th.run_dense(J, A, torch.as_tensor(e0), steps) - tgt) ** 2).mean().backward()
    N = o.N
    G = J.grad.reshape(N, 3, N, 3).permute(0, 2, 1, 3).numpy()           # G[i, j] = dL/dJ block (i, j)
    pairs = _forward_pairs(sol.system)
    assert sol.nbonds == len(pairs) == 4
    ref_b = _bond_grads(G, pairs)
    got_b = gb.cpu().numpy()
    assert _relerr(got_b, ref_b) < 1e-9
    assert _relerr(ga.cpu().numpy(), A.grad.numpy()) < 1e-9
    # the DMI vectors' gradient is the antisymmetric part, read with eps_abc
    dD = dmi_gradient(gb).cpu().numpy()
    assert np.allclose(dD[..., 2], ref_b[..., 0, 1] - ref_b[..., 1, 0], rtol=0, atol=1e-12 * np.abs(ref_b).max())


def test_adjoint_dense_gradients_tiled_and_fp32(gpu):
    """square 64^2 (tiled cells in the library): the dense gradients equal the row-major layout's to
    round-off, satisfy the chain rule dL/dtheta_p = sum_b <U_p(b), dL/dT_b> + sum_i <Aun_p(i), dL/dA_i>
    to 1e-10, and FP32 stays within 2e-3 of FP64 (of the largest element) over 80 steps."""
    print("case: dense adjoint gradients, square 64^2 tiled vs row-major, FP64 and FP32", flush=True)
    from magex_amd.adjoint import DifferentiableLLG
    L, steps = 64, 80
    dev, _ = _builders(L)
    s0 = dev(THETA0)
    e0 = s0.init_spins("random", seed=8)
    tgt = torch.as_tensor(s0.init_spins("random", seed=9), device="cuda")
    out = {}
    for layout, prec in (("rowmajor", "fp64"), ("tiled", "fp64"), ("tiled", "fp32")):
        sol = DifferentiableLLG(dev, THETA0, alpha=0.3, dt=1e-3, precision=prec, checkpoint_every=30,
                                layout=layout)
        if layout == "tiled":
            assert tuple(sol.dm.layout()) != (0, 0, 0)
        e = sol.forward(THETA0, e0, steps)
        gt, _, gb, ga = sol.backward_dense(_mse_grad(e.double(), tgt).to(e.dtype), want_e0=False)
        torch.cuda.synchronize()
        out[(layout, prec)] = (gt.cpu().numpy(), gb.cpu().numpy(), ga.cpu().numpy(), sol)
    gt, gb, ga, sol = out[("tiled", "fp64")]
    gt_r, gb_r, ga_r, _ = out[("rowmajor", "fp64")]
    assert _relerr(gb, gb_r) < 1e-12 and _relerr(ga, ga_r) < 1e-12
    fw = sol.system.forward_entries()
    U = sol._U.reshape(len(sol._U), -1, 3, 3)[:, fw]                      # (P, nbonds, 3, 3)
    Aun = sol._Aun.reshape(len(sol._Aun), -1, 3, 3)                       # (P, ns, 3, 3)
    ns = sol.system.uc.nsites
    ga_s = ga.reshape(-1, ns, 3, 3).sum(0)
    chain = np.einsum("pfab,fcab->p", U, gb) + np.einsum("psab,sab->p", Aun, ga_s)
    assert _relerr(chain[:sol.P], gt) < 1e-10, (chain, gt)
    gt32, gb32, ga32, _ = out[("tiled", "fp32")]
    assert _relerr(gb32, gb) < 2e-3 and _relerr(ga32, ga) < 2e-3


def _disorder(dev0, ref0, th0, seed, sym_A):
    """Builders of the same model plus fixed per-bond tensors D_b (on every unordered bond, J_ji = J_ij^T)
    and fixed per-site on-site tensors: the device one through set_bond_tensors / A, the oracle one as
    blocks added to its coupling matrix."""
    import scipy.sparse as sp
    s0 = dev0(th0)
    fw = s0.forward_entries()
    rng = np.random.default_rng(seed)
    D = rng.normal(0.0, 0.3, (len(fw), s0.ncell, 3, 3))
    Af = rng.normal(0.0, 0.05, (s0.N, 3, 3))
    if sym_A:
        Af = 0.5 * (Af + Af.transpose(0, 2, 1))
    pairs = _forward_pairs(s0)
    rows, cols, vals = [], [], []
    for f, (i, j) in enumerate(pairs):
        ok = j >= 0
        i, j = i[ok], j[ok]
        for a in range(3):
            for b in range(3):
                rows += [3 * i + a, 3 * j + b]
                cols += [3 * j + b, 3 * i + a]
                vals += [D[f, ok, a, b], D[f, ok, a, b]]
    rows = np.concatenate([np.broadcast_to(np.asarray(r), (s0.ncell,)) if np.ndim(r) == 0 else r for r in rows])
    cols = np.concatenate([np.broadcast_to(np.asarray(c), (s0.ncell,)) if np.ndim(c) == 0 else c for c in cols])
    Dsp = sp.csr_matrix((np.concatenate(vals), (rows, cols)), shape=(3 * s0.N, 3 * s0.N))
    Dsp.sum_duplicates()

    def dev(th):
        s = dev0(th)
        Js = np.array([J for *_, J in s.stencil], float)
        s.set_bond_tensors(Js[fw][:, None] + D)
        s.A = np.asarray(s.A, float).reshape(s.N, 3, 3) + Af
        return s

    def ref(th):
        o = ref0(th)
        o.Jsp = (o.Jsp + Dsp).tocsr()
        o.A = o.A + Af
        return o

    return dev, ref, pairs


def _check_against_autograd(sol, ref, th0, e0, tgt, steps, pairs):
    e = sol.forward(th0, e0, steps)
    gt, g0, gb, ga = sol.backward_dense(_mse_grad(e, tgt.cuda()))
    torch.cuda.synchronize()
    Th = TorchHeun(ref, th0, alpha=0.3, dt=1e-3)
    t = torch.tensor(th0, requires_grad=True)
    x0 = torch.tensor(e0, requires_grad=True)
    eT = Th.run(t, x0, steps)
    ((eT - tgt) ** 2).mean().backward()
    assert np.abs(e.cpu().numpy() - eT.detach().numpy()).max() < 1e-12
    assert _relerr(gt.cpu().numpy(), t.grad.numpy()) < 1e-9, (gt, t.grad)
    assert _relerr(g0.cpu().numpy(), x0.grad.numpy()) < 1e-9
    J, A = Th.dense(th0)
    J = J.clone().requires_grad_(True)
    A = A.clone().requires_grad_(True)
    ((Th.run_dense(J, A, torch.as_tensor(e0), steps) - tgt) ** 2).mean().backward()
    N = len(e0)
    G = J.grad.reshape(N, 3, N, 3).permute(0, 2, 1, 3).numpy()
    assert _relerr(gb.cpu().numpy(), _bond_grads(G, pairs)) < 1e-9
    assert _relerr(ga.cpu().numpy(), A.grad.numpy()) < 1e-9


def test_adjoint_disordered_square_matches_autograd(gpu):
    """A disordered model: fixed random per-bond tensors on every bond and non-symmetric per-site on-site
    tensors on top of the parameterized square model (FP64, 10^2, 50 steps over three checkpoint
    segments).  Parameter, spin and dense gradients match FP64 autograd to 1e-9."""
    print("case: adjoint, disordered square 10^2 (per-bond + per-site 9), 50 steps", flush=True)
    from magex_amd.adjoint import DifferentiableLLG
    L, steps = 10, 50
    dev0, ref0 = _builders(L)
    dev, ref, pairs = _disorder(dev0, ref0, THETA0, seed=21, sym_A=False)
    sol = DifferentiableLLG(dev, THETA0, alpha=0.3, dt=1e-3, precision="fp64", checkpoint_every=20)
    o = ref(THETA0)
    _check_against_autograd(sol, ref, THETA0, o.init_random(8), torch.as_tensor(o.init_random(9)), steps, pairs)


def test_adjoint_bcc_perbond_model_matches_autograd(gpu):
    """The headline model's form at 4^3 cells: bcc conventional (two sites per cell, 28 stencil entries)
    with random per-bond tensors and symmetric per-site anisotropy tensors; parameters J1, J2, K.
    FP64, 40 steps: parameter, spin and dense gradients match FP64 autograd to 1e-9."""
    print("case: adjoint, bcc 4^3 per-bond + per-site SIA, 40 steps", flush=True)
    from magex_amd.adjoint import DifferentiableLLG
    from magex_amd import models as Md
    from magex_amd import spinmodel as Sd
    L, steps = 4, 40
    th0 = np.array([20.0, 10.0, 0.5])

    def dev0(th):
        return Sd.SpinSystem(Sd.load_results(Md.make_model("bcc_conventional", a=2.87, J=[th[0], th[1]], K=th[2],
                                                           moment=2.2)), (L, L, L))

    def ref0(th):
        # the results.json dict of the two-site cell (data) read by the oracle's own loader
        return R.SpinSystem(R.load_results(Md.make_model("bcc_conventional", a=2.87, J=[th[0], th[1]], K=th[2],
                                                         moment=2.2)), (L, L, L))

    dev, ref, pairs = _disorder(dev0, ref0, th0, seed=22, sym_A=True)
    sol = DifferentiableLLG(dev, th0, alpha=0.3, dt=1e-3, precision="fp64", checkpoint_every=15)
    assert sol.system.uc.nsites == 2 and sol.nbonds == len(pairs)
    o = ref(th0)
    _check_against_autograd(sol, ref, th0, o.init_random(10), torch.as_tensor(o.init_random(11)), steps, pairs)


def test_adjoint_stencil_beyond_the_table(gpu):
    """bcc with three shells: 52 stencil entries, more than the kernels' constant coupling table holds,
    so the adjoint runs its model in the per-bond form (zero fixed part).  FP64, 4^3 cells, 30 steps:
    parameter, spin and dense gradients match FP64 autograd to 1e-9."""
    print("case: adjoint, bcc 4^3 three shells (52 entries), 30 steps", flush=True)
    from magex_amd.adjoint import DifferentiableLLG
    from magex_amd import models as Md
    from magex_amd import spinmodel as Sd
    L, steps = 4, 30
    th0 = np.array([20.0, 10.0, -3.0])

    def dev(th):
        return Sd.SpinSystem(Sd.load_results(Md.make_model("bcc_conventional", a=2.87, J=list(th), moment=2.2)),
                             (L, L, L))

    def ref(th):
        return R.SpinSystem(R.load_results(Md.make_model("bcc_conventional", a=2.87, J=list(th), moment=2.2)),
                            (L, L, L))

    sol = DifferentiableLLG(dev, th0, alpha=0.3, dt=1e-3, precision="fp64", checkpoint_every=12)
    assert len(sol.system.stencil) * 9 > 256
    o = ref(th0)
    _check_against_autograd(sol, ref, th0, o.init_random(12), torch.as_tensor(o.init_random(13)), steps,
                            _forward_pairs(sol.system))


def test_adjoint_pinned_sites_match_autograd(gpu):
    """Pinned sites (a stripe and scattered sites, as structure.apply_structure makes them): they keep
    their spins, feel no field and still act on their neighbours.  FP64, square 12^2, 40 steps at 50 K
    over three checkpoint segments: parameter, spin and dense gradients match FP64 autograd of the
    restatement that holds pinned spins (to 1e-9)."""
    print("case: adjoint with pinned sites, square 12^2, 40 steps, T=50 K", flush=True)
    from magex_amd.adjoint import DifferentiableLLG
    L, steps, T, seed = 12, 40, 50.0, 4
    dev0, ref0 = _builders(L)
    mask = np.zeros(L * L, bool)
    mask[:L] = True                                   # the first row of cells
    mask[[30, 57, 100, 143]] = True

    def dev(th):
        s = dev0(th)
        s.pinned = mask
        return s

    def ref(th):
        o = ref0(th)
        o.pinned = mask
        return o

    sol = DifferentiableLLG(dev, THETA0, alpha=0.3, dt=1e-3, temperature=T, seed=seed, precision="fp64",
                            checkpoint_every=16)
    o = ref(THETA0)
    e0 = o.init_random(14)
    tgt = torch.as_tensor(o.init_random(15))
    e = sol.forward(THETA0, e0, steps)
    gt, g0, gb, ga = sol.backward_dense(_mse_grad(e, tgt.cuda()))
    torch.cuda.synchronize()
    assert np.array_equal(e.cpu().numpy()[mask], e0[mask])
    sg = R.sigma(o, 0.3, T, 1e-3)
    fields = [torch.as_tensor(sg[:, None] * sol.dm.noise(seed, n).double().cpu().numpy()) for n in range(steps)]
    Th = TorchHeun(ref, THETA0, alpha=0.3, dt=1e-3)
    t = torch.tensor(THETA0, requires_grad=True)
    x0 = torch.tensor(e0, requires_grad=True)
    eT = Th.run(t, x0, steps, thermal=lambda n: fields[n])
    ((eT - tgt) ** 2).mean().backward()
    assert np.abs(e.cpu().numpy() - eT.detach().numpy()).max() < 1e-12
    assert _relerr(gt.cpu().numpy(), t.grad.numpy()) < 1e-9, (gt, t.grad)
    assert _relerr(g0.cpu().numpy(), x0.grad.numpy()) < 1e-9
    J, A = Th.dense(THETA0)
    J = J.clone().requires_grad_(True)
    A = A.clone().requires_grad_(True)
    ((Th.run_dense(J, A, torch.as_tensor(e0), steps, thermal=lambda n: fields[n]) - tgt) ** 2).mean().backward()
    N = L * L
    G = J.grad.reshape(N, 3, N, 3).permute(0, 2, 1, 3).numpy()
    assert _relerr(gb.cpu().numpy(), _bond_grads(G, _forward_pairs(sol.system))) < 1e-9
    assert _relerr(ga.cpu().numpy(), A.grad.numpy()) < 1e-9


def test_adjoint_cubic_and_higher_uniaxial_terms(gpu):
    """Fixed cubic (per-site Kc, rotated axes) and higher-uniaxial terms on the parameterized square
    model: their Hessians enter the adjoint's on-site transpose.  FP64, 10^2, 50 steps: parameter,
    spin and dense gradients match FP64 autograd of the restatement with the same terms (to 1e-9)."""
    print("case: adjoint with cubic + K2/K3 terms, square 10^2, 50 steps", flush=True)
    from magex_amd.adjoint import DifferentiableLLG
    from magex_amd.terms import CubicAnisotropy, UniaxialHigher
    L, steps = 10, 50
    dev0, ref0 = _builders(L)
    Kc = np.random.default_rng(5).uniform(0.2, 0.8, L * L)
    C = np.linalg.qr(np.random.default_rng(6).normal(size=(3, 3)))[0]
    K2, K3, u = 0.4, 0.15, (0.3, -0.2, 1.0)

    def dev(th):
        s = dev0(th)
        s.terms.append(CubicAnisotropy(Kc, C))
        s.terms.append(UniaxialHigher(K2, K3, u))
        return s

    def ref(th):
        o = ref0(th)
        o.terms = [R.CubicAnisotropy(Kc, C), R.UniaxialHigher(K2, K3, u)]
        return o

    sol = DifferentiableLLG(dev, THETA0, alpha=0.3, dt=1e-3, precision="fp64", checkpoint_every=20)
    o = ref(THETA0)
    _check_against_autograd(sol, ref, THETA0, o.init_random(16), torch.as_tensor(o.init_random(17)), steps,
                            _forward_pairs(sol.system))


def test_adjoint_open_axis_disordered(gpu):
    """Per-bond tensors on a lattice with an open axis (square 9 x 8, open along axis 0): reverse
    entries whose partner bond leaves the lattice read nothing, and the forward of the library equals
    the restatement's to 1e-12; parameter, spin and dense gradients match FP64 autograd to 1e-9 (zero
    on the cut bonds)."""
    print("case: adjoint, disordered square 9x8 open along axis 0, 40 steps", flush=True)
    from magex_amd.adjoint import DifferentiableLLG
    from magex_amd import models as Md
    from magex_amd import spinmodel as Sd
    L, pbc, steps = (9, 8, 1), (False, True, True), 40

    def dev0(th):
        return Sd.SpinSystem(Sd.load_results(Md.make_model("square", J=[th[0], th[1]], D=th[2], K=th[3])), L, pbc)

    def ref0(th):
        return R.SpinSystem(R.load_results(R.make_model("square", J=[th[0], th[1]], D=th[2], K=th[3])), L, pbc)

    dev, ref, pairs = _disorder(dev0, ref0, THETA0, seed=23, sym_A=True)
    assert any((j < 0).any() for _, j in pairs)
    sol = DifferentiableLLG(dev, THETA0, alpha=0.3, dt=1e-3, precision="fp64", checkpoint_every=15)
    o = ref(THETA0)
    _check_against_autograd(sol, ref, THETA0, o.init_random(18), torch.as_tensor(o.init_random(19)), steps, pairs)


def test_adjoint_spin_orbit_torque_fields(gpu):
    """Spin-orbit torque fields B_DL (e x p) + B_FL p (held fixed): the damping-like term enters the
    adjoint as B_DL (p x dL/dB).  FP64, square 10^2, 50 steps: parameter, spin and dense gradients match
    FP64 autograd of the restatement with the same fields (to 1e-9)."""
    print("case: adjoint with SOT fields, square 10^2, 50 steps", flush=True)
    from magex_amd.adjoint import DifferentiableLLG
    L, steps = 10, 50
    dev, ref = _builders(L)
    p = np.array([0.3, 1.0, 0.2])
    sot = {"p": p, "B_DL": 1.5, "B_FL": 0.7}
    sol = DifferentiableLLG(dev, THETA0, alpha=0.3, dt=1e-3, precision="fp64", checkpoint_every=20, sot=sot)
    o = ref(THETA0)
    e0 = o.init_random(20)
    tgt = torch.as_tensor(o.init_random(21))
    e = sol.forward(THETA0, e0, steps)
    gt, g0, gb, ga = sol.backward_dense(_mse_grad(e, tgt.cuda()))
    torch.cuda.synchronize()
    Th = TorchHeun(ref, THETA0, alpha=0.3, dt=1e-3, sot=(p / np.linalg.norm(p), 1.5, 0.7))
    t = torch.tensor(THETA0, requires_grad=True)
    x0 = torch.tensor(e0, requires_grad=True)
    ((Th.run(t, x0, steps) - tgt) ** 2).mean().backward()
    eT = Th.run(torch.as_tensor(THETA0), torch.as_tensor(e0), steps)
    assert np.abs(e.cpu().numpy() - eT.numpy()).max() < 1e-12
    assert _relerr(gt.cpu().numpy(), t.grad.numpy()) < 1e-9, (gt, t.grad)
    assert _relerr(g0.cpu().numpy(), x0.grad.numpy()) < 1e-9
    J, A = Th.dense(THETA0)
    J = J.clone().requires_grad_(True)
    A = A.clone().requires_grad_(True)
    ((Th.run_dense(J, A, torch.as_tensor(e0), steps) - tgt) ** 2).mean().backward()
    G = J.grad.reshape(L * L, 3, L * L, 3).permute(0, 2, 1, 3).numpy()
    assert _relerr(gb.cpu().numpy(), _bond_grads(G, _forward_pairs(sol.system))) < 1e-9
    assert _relerr(ga.cpu().numpy(), A.grad.numpy()) < 1e-9


def test_adjoint_driven_field_and_field_gradient(gpu):
    """A time-dependent uniform field (field_func at t and t + dt of every step): the driven forward
    equals the restatement's trajectory to 1e-12, and dL/dB of every pass of every step, with dL/dtheta
    and dL/de0, match FP64 autograd to 1e-9 (square 10^2, 45 steps over three checkpoint segments).
    trajectory() with the field as a torch leaf gives the same field gradient through loss.backward()."""
    print("case: adjoint, driven field + field gradient, square 10^2, 45 steps", flush=True)
    L, steps = 10, 45
    sol = _solver(L, C=20)
    _, ref_build = _builders(L)
    o = ref_build(THETA0)
    e0 = o.init_random(22)
    tgt = torch.as_tensor(o.init_random(23))
    ff = lambda t: np.array([3.0 * np.sin(40.0 * t), 0.5, 2.0 * np.cos(25.0 * t)])     # noqa: E731
    B = sol.field_sequence(ff, steps)
    assert np.allclose(B[3, 1], o.B + ff(4 * 1e-3))
    e = sol.forward(THETA0, e0, steps, field=B)
    gt, g0, gB = sol.backward_field(_mse_grad(e, tgt.cuda()))
    torch.cuda.synchronize()
    Th = TorchHeun(ref_build, THETA0, alpha=0.3, dt=1e-3)
    t = torch.tensor(THETA0, requires_grad=True)
    x0 = torch.tensor(e0, requires_grad=True)
    Bt = torch.tensor(B, requires_grad=True)
    eT = Th.run(t, x0, steps, field=Bt)
    ((eT - tgt) ** 2).mean().backward()
    assert np.abs(e.cpu().numpy() - eT.detach().numpy()).max() < 1e-12
    assert _relerr(gt.cpu().numpy(), t.grad.numpy()) < 1e-9
    assert _relerr(g0.cpu().numpy(), x0.grad.numpy()) < 1e-9
    assert _relerr(gB.cpu().numpy(), Bt.grad.numpy()) < 1e-9
    # the autograd Function with the field as a leaf
    th = torch.tensor(THETA0, requires_grad=True)
    fl = torch.tensor(B, requires_grad=True)
    x = sol.trajectory(th, torch.as_tensor(e0), steps, field=fl)
    ((x - tgt.to(x.device)) ** 2).mean().backward()
    assert _relerr(fl.grad.numpy(), Bt.grad.numpy()) < 1e-9
    assert _relerr(th.grad.numpy(), t.grad.numpy()) < 1e-9


def test_adjoint_field_gradient_of_the_static_field(gpu):
    """Without a drive, the field gradient summed over steps and passes is dL/dB of the system's static
    field: it matches a central difference of the library's own forward in the field (h = 1e-4 T)."""
    print("case: adjoint, static-field gradient vs central differences, square 12^2, 60 steps", flush=True)
    L, steps, h = 12, 60, 1e-4
    sol = _solver(L, C=25)
    s0 = sol.system
    e0 = s0.init_spins("random", seed=24)
    tgt = torch.as_tensor(s0.init_spins("random", seed=25), device="cuda")
    e = sol.forward(THETA0, e0, steps)
    _, _, gB = sol.backward_field(_mse_grad(e, tgt), want_e0=False)
    total = gB.sum((0, 1)).cpu().numpy()
    B0 = np.asarray(s0.B, float)
    for c in range(3):
        fd = []
        for sgn in (1, -1):
            Bc = B0.copy()
            Bc[c] += sgn * h
            x = sol.forward(THETA0, e0, steps, field=np.broadcast_to(Bc, (steps, 2, 3)))
            fd.append(float(((x - tgt) ** 2).mean()))
        d = (fd[0] - fd[1]) / (2 * h)
        assert abs(total[c] - d) <= 1e-6 * max(abs(d), 1e-12) + 1e-12, (c, total[c], d)


def test_adjoint_entry_points_reject_bad_arguments(gpu):
    """Host-side validation of the adjoint's entry points: null outputs and inputs, a backward before
    any forward, and a backward after the parameters changed raise ValueError with the library's
    message; nothing is launched for them."""
    print("case: adjoint argument errors", flush=True)
    import ctypes
    from magex_amd._lib import check
    sol = _solver(6, C=5)
    L, h = sol.dm.lib, sol.h
    g = torch.zeros((sol.N, 3), dtype=torch.float64, device="cuda")
    out = torch.zeros(16, dtype=torch.float64, device="cuda")
    null = ctypes.c_void_p(0)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())                                    # noqa: E731
    with pytest.raises(ValueError, match="no forward recorded"):
        check(L.mx_adjoint_backward(h, ptr(g), null, ptr(out), None))
    with pytest.raises(ValueError, match="null field sequence"):
        check(L.mx_adjoint_forward_driven(h, ptr(g), 3, None, 0, 0, None))
    e = sol.forward(THETA0, sol.system.init_spins("random", seed=1), 4)
    with pytest.raises(ValueError, match="null field-gradient output"):
        check(L.mx_adjoint_backward_field(h, ptr(g), null, ptr(out), None, None))
    with pytest.raises(ValueError, match="null argument"):
        check(L.mx_adjoint_backward(h, None, null, ptr(out), None))
    with pytest.raises(ValueError, match="negative step count"):
        check(L.mx_adjoint_forward(h, ptr(e), -1, 0, 0, None))
    sol.set_params(THETA0 + 0.5)                      # the recorded forward belongs to other parameters
    with pytest.raises(ValueError, match="no forward recorded"):
        sol.backward(g)
    assert L.mx_adjoint_num_bonds(h) == 4 and L.mx_adjoint_num_bonds(None) == -1


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_adjoint_depondt_matches_autograd(gpu, precision):
    """The Depondt integrator (Heun on rotation vectors, magex/llg.py:236-240): the adjoint goes through
    both Rodrigues rotations and the final renormalisation.  Square 12^2 at 50 K, 40 steps over three
    checkpoint segments.  FP64: parameter, spin and dense gradients match FP64 autograd of the Depondt
    restatement to 1e-9; FP32: the parameter gradient stays within 2e-3 of it."""
    print(f"case: adjoint Depondt {precision}, square 12^2, 40 steps, T=50 K", flush=True)
    from magex_amd.adjoint import DifferentiableLLG
    L, steps, T, seed = 12, 40, 50.0, 9
    dev, ref = _builders(L)
    sol = DifferentiableLLG(dev, THETA0, alpha=0.3, dt=1e-3, temperature=T, seed=seed, precision=precision,
                            checkpoint_every=16, integrator="depondt")
    o = ref(THETA0)
    e0 = o.init_random(26)
    tgt = torch.as_tensor(o.init_random(27))
    e = sol.forward(THETA0, e0, steps)
    gt, g0, gb, ga = sol.backward_dense(_mse_grad(e.double(), tgt.cuda()).to(e.dtype))
    torch.cuda.synchronize()
    sg = R.sigma(o, 0.3, T, 1e-3)
    fields = [torch.as_tensor(sg[:, None] * sol.dm.noise(seed, n).double().cpu().numpy()) for n in range(steps)]
    Th = TorchHeun(ref, THETA0, alpha=0.3, dt=1e-3, integrator="depondt")
    t = torch.tensor(THETA0, requires_grad=True)
    x0 = torch.tensor(e0, requires_grad=True)
    eT = Th.run(t, x0, steps, thermal=lambda n: fields[n])
    ((eT - tgt) ** 2).mean().backward()
    if precision == "fp32":
        assert _relerr(gt.cpu().numpy(), t.grad.numpy()) < 2e-3, (gt, t.grad)
        return
    assert np.abs(e.cpu().numpy() - eT.detach().numpy()).max() < 1e-12
    assert _relerr(gt.cpu().numpy(), t.grad.numpy()) < 1e-9, (gt, t.grad)
    assert _relerr(g0.cpu().numpy(), x0.grad.numpy()) < 1e-9
    J, A = Th.dense(THETA0)
    J = J.clone().requires_grad_(True)
    A = A.clone().requires_grad_(True)
    ((Th.run_dense(J, A, torch.as_tensor(e0), steps, thermal=lambda n: fields[n]) - tgt) ** 2).mean().backward()
    G = J.grad.reshape(L * L, 3, L * L, 3).permute(0, 2, 1, 3).numpy()
    assert _relerr(gb.cpu().numpy(), _bond_grads(G, _forward_pairs(sol.system))) < 1e-9
    assert _relerr(ga.cpu().numpy(), A.grad.numpy()) < 1e-9
