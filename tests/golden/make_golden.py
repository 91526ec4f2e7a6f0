"""Writes golden input/output vectors of magex's own NumPy LLG path.

Run once, in a container that holds the magex sources (they are imported, not copied):

    python tests/golden/make_golden.py /root/reference

The vectors pin what no closed form pins: the assembly chain
(magex/models.py:40-116 -> magex/spinmodel.py:64-151 -> magex/spinmodel.py:191-252: site order,
image splitting, self-image folding, DMI sign) and the Heun trajectory of
magex/llg.py:216-227 itself.  Every file is plain data (``.npz`` without pickles, ``.json``).
"""
import json
import os
import sys
import tempfile

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REF)

from magex import texture  # noqa: E402
from magex.llg import LLGSolver  # noqa: E402
from magex.models import make_model, write_model  # noqa: E402
from magex.spinmodel import SpinSystem, load_results  # noqa: E402
import scipy.sparse as sp  # noqa: E402


class _InjectedNormals:
    """Stand-in for the solver's numpy Generator: hands out stored standard normals in order."""

    def __init__(self, g):
        self.g, self.k = g, 0

    def normal(self, size=None):
        out = self.g[self.k]
        assert out.shape == tuple(size)
        self.k += 1
        return out.copy()


def _system(lattice_kw, L, field, pbc=(True, True, True)):
    fd, p = tempfile.mkstemp(suffix=".json")
    os.close(fd)
    write_model(p, **lattice_kw)
    s = SpinSystem(load_results(p), L, pbc, field)
    os.remove(p)
    return s


def _save(name, **arrs):
    np.savez_compressed(os.path.join(OUT, name), **arrs)
    print("wrote", name, sum(np.asarray(a).nbytes for a in arrs.values()) // 1024, "KiB raw")


def models_json():
    """make_model outputs for a spread of lattices/options (pins the pair generator)."""
    cases = {
        "sc_J1": dict(lattice="sc", J=[1.0]),
        "tri_3shell_dmi": dict(lattice="triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3),
        "square_2shell_dmi": dict(lattice="square", J=[4.0, -0.4], D=1.0, K=0.2),
        "bcc_2shell": dict(lattice="bcc", J=[20.0, 10.0], moment=2.2),
        "fcc_bulkdmi": dict(lattice="fcc", J=[1.0], D=0.2, dmi="bulk"),
        "honeycomb": dict(lattice="honeycomb", J=[1.0, 0.1], D=0.1),
        "kagome": dict(lattice="kagome", J=[1.0], K=0.05, K_axis=(1, 0, 1)),
        "chain_afm": dict(lattice="chain", J=[-1.0], order="afm"),
        "sc_afm": dict(lattice="sc", J=[-1.0], order="afm"),
        "sc_jani": dict(lattice="sc", J=[1.0], Jani=[[0.1, 0.02, 0], [0.02, -0.05, 0], [0, 0, -0.05]]),
    }
    out = {k: make_model(**v) for k, v in cases.items()}
    with open(os.path.join(OUT, "make_model_cases.json"), "w") as f:
        json.dump(dict(kwargs=cases, models=out), f)
    print("wrote make_model_cases.json")


def couplings_tri():
    """load_results of the triangular 3-shell + DMI model: the coupling list (i, j, R, J)."""
    fd, p = tempfile.mkstemp(suffix=".json")
    os.close(fd)
    write_model(p, lattice="triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3)
    uc = load_results(p)
    uc2 = load_results(p, rcut=1.05 * 3.0, jtol=0.3)           # shell filtering by rcut / jtol
    os.remove(p)
    _save("load_results_tri.npz",
          i=np.array([c[0] for c in uc.couplings]), j=np.array([c[1] for c in uc.couplings]),
          R=np.array([c[2] for c in uc.couplings]), J=np.array([c[3] for c in uc.couplings]),
          A=uc.A, moments=uc.moments, signs=uc.signs, ref_dirs=uc.ref_dirs,
          n_filtered=np.array(len(uc2.couplings)))


def case_c1():
    """sc 16^3, J=1, K=0.1, B=(0,0,0.5), mu=2, alpha=0.1, T=0, dt=1e-4 ps, 1000 steps, e0 from seed 0."""
    s = _system(dict(lattice="sc", J=[1.0], K=0.1, moment=2.0), (16, 16, 16), (0, 0, 0.5))
    e = s.init_spins("random", seed=0)
    e0 = e.copy()
    B0 = s.effective_field(e0)
    E0 = s.energy(e0)
    sol = LLGSolver(s, alpha=0.1, dt=1e-4, temperature=0.0, seed=0)
    assert sol.fast is None
    snaps = {}
    for n in range(1, 1001):
        e = sol.step(e)
        if n in (1, 10, 1000):
            snaps[n] = e.copy()
    _save("c1_sc16.npz", e0=e0, B0=B0, E0=np.array(E0), e1=snaps[1], e10=snaps[10], e1000=snaps[1000],
          E1000=np.array(s.energy(snaps[1000])))


def case_tri():
    """triangular 12x12, 3 shells + interfacial DMI + K + B, T = 20 K, injected noise (10 steps)."""
    s = _system(dict(lattice="triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3), (12, 12, 1), (0, 0, 2.0))
    e0 = s.init_spins("random", seed=5)
    g = np.random.default_rng(11).normal(size=(10, s.N, 3))
    sol = LLGSolver(s, alpha=0.05, dt=5e-4, temperature=20.0, seed=0)
    sol.rng = _InjectedNormals(g)
    e = e0.copy()
    for _ in range(10):
        e = sol.step(e)
    _save("tri12.npz", e0=e0, B0=s.effective_field(e0), G0=s.gradient(e0), H0=s.exchange_field(e0),
          E0=np.array(s.energy(e0)), noise=g, e10=e, E10=np.array(s.energy(e)),
          order10=np.array(s.order_parameter(e)))


def case_square():
    """square 8x8, 2 shells + DMI + K (the differentiation case's Hamiltonian), alpha=0.3, dt=1e-3, T=0."""
    s = _system(dict(lattice="square", J=[5.0, -0.5], D=1.5, K=0.3), (8, 8, 1), (0, 0, 0))
    e0 = s.init_spins("random", seed=0)
    sol = LLGSolver(s, alpha=0.3, dt=1e-3)
    e = e0.copy()
    for _ in range(50):
        e = sol.step(e)
    _save("square8.npz", e0=e0, B0=s.effective_field(e0), E0=np.array(s.energy(e0)), e50=e,
          E50=np.array(s.energy(e)))


def _bcc_conventional(a=2.87, J=(20.0, 10.0), moment=2.2):
    """results.json dict of bcc in its 2-site conventional cubic cell (not a make_model lattice)."""
    import itertools
    cell = np.eye(3) * a
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0.5]]) @ cell
    cand = []
    for i in range(2):
        for j in range(2):
            for R in itertools.product(range(-2, 3), repeat=3):
                if i == j and R == (0, 0, 0):
                    continue
                r = basis[j] + np.array(R) @ cell - basis[i]
                cand.append((round(float(np.linalg.norm(r)), 6), i, j, R, r))
    shells = sorted({c[0] for c in cand})[:len(J)]
    pairs = []
    for d, i, j, R, r in cand:
        if d in shells:
            pairs.append(dict(i=i, j=j, R=list(R), nu=[R[k] % 64 for k in range(3)], dist=d, vec=r.tolist(),
                              J=(J[shells.index(d)] * np.eye(3)).tolist()))
    ref = [0, 0, 1.0]
    return dict(version="model", sites=[0, 1], symbols=["Fe", "Fe"], cell=cell.tolist(), positions=basis.tolist(),
                mesh=[64, 64, 64], moments=[[0, 0, moment]] * 2, refs=[ref], A=[np.zeros((3, 3)).tolist()] * 2,
                mag_pos=[0, 1], pairs=pairs)


def case_bcc_perbond():
    """bcc conventional 4x4x4 (128 sites) with per-bond random tensors (J_ji = J_ij^T) and per-site SIA."""
    d = _bcc_conventional()
    fd, p = tempfile.mkstemp(suffix=".json")
    os.close(fd)
    with open(p, "w") as f:
        json.dump(d, f)
    s = SpinSystem(load_results(p), (4, 4, 4))
    os.remove(p)
    rng = np.random.default_rng(1)
    J = s.Jsp.tocoo()
    bi, bj = J.row // 3, J.col // 3
    keys = np.unique(bi.astype(np.int64) * s.N + bj)
    src, dst = keys // s.N, keys % s.N
    blocks = np.zeros((len(keys), 3, 3))
    np.add.at(blocks, (np.searchsorted(keys, bi.astype(np.int64) * s.N + bj), J.row % 3, J.col % 3), J.data)
    lookup = {(a, b): k for k, (a, b) in enumerate(zip(src, dst))}
    done = np.zeros(len(keys), bool)
    for k in range(len(keys)):
        if done[k]:
            continue
        r = lookup[(dst[k], src[k])]
        iso = blocks[k][0, 0]
        S = rng.normal(0, 0.01 * abs(iso), (3, 3))
        S = np.triu(S) + np.triu(S, 1).T
        Dv = rng.normal(0, 0.01 * abs(iso), 3)
        Aq = np.array([[0, Dv[2], -Dv[1]], [-Dv[2], 0, Dv[0]], [Dv[1], -Dv[0], 0]])
        blocks[k] = blocks[k] + S + Aq
        blocks[r] = blocks[k].T
        done[k] = done[r] = True
    rows = (3 * src[:, None, None] + np.arange(3)[None, :, None]).repeat(3, 2).ravel()
    cols = (3 * dst[:, None, None] + np.arange(3)[None, None, :]).repeat(3, 1).ravel()
    s.Jsp = sp.csr_matrix((blocks.ravel(), (rows, cols)), shape=s.Jsp.shape)
    A = rng.normal(0, 0.05, (s.N, 3, 3))
    s.A = 0.5 * (A + A.transpose(0, 2, 1))
    s.has_A = True
    e0 = s.init_spins("random", seed=0)
    g = np.random.default_rng(3).normal(size=(10, s.N, 3))
    sol = LLGSolver(s, alpha=0.1, dt=1e-4, temperature=300.0, seed=0)
    sol.rng = _InjectedNormals(g)
    e = e0.copy()
    for _ in range(10):
        e = sol.step(e)
    _save("bcc_perbond.npz", src=src, dst=dst, blocks=blocks, A=s.A, mu=s.mu, e0=e0,
          B0=s.effective_field(e0), E0=np.array(s.energy(e0)), noise=g, e10=e, E10=np.array(s.energy(e)))
    with open(os.path.join(OUT, "bcc_conventional_model.json"), "w") as f:
        json.dump(d, f)


def case_charge():
    """Berg-Luscher charge of textures (magex/texture.py:30-63) on triangular and square layers."""
    out = {}
    s = _system(dict(lattice="triangular", J=[1.0]), (30, 30, 1), (0, 0, 0))
    c = s.positions[:, :2].mean(0)
    r = s.positions[:, :2] - c
    R = np.linalg.norm(r, axis=1)
    th = np.pi * np.exp(-(R / 12.0) ** 2)
    ph = np.arctan2(r[:, 1], r[:, 0])
    e = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
    out["tri_sky_e"], out["tri_sky_Q"] = e, np.array(texture.topological_charge(s, e))
    er = s.init_spins("random", seed=4)
    out["tri_rand_e"], out["tri_rand_Q"] = er, np.array(texture.topological_charge(s, er))
    s2 = _system(dict(lattice="square", J=[1.0]), (20, 20, 1), (0, 0, 0))
    c = s2.positions[:, :2].mean(0)
    r = s2.positions[:, :2] - c
    R = np.linalg.norm(r, axis=1)
    th = np.pi * np.exp(-(R / 7.0) ** 2)
    ph = np.arctan2(r[:, 1], r[:, 0])
    e = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
    out["sq_sky_e"], out["sq_sky_Q"] = e, np.array(texture.topological_charge(s2, e))
    _save("charge.npz", **out)


# [structure] configurations of the drop-in's golden case: a layered film (open axis 2) with two
# materials, terms, Neel surface anisotropy, interface exchange, a pinned bottom layer and a temperature
# ramp; and a core-shell alloy on a periodic triangular lattice with a pinned material and a hot spot.
STRUCTURE_FILM = dict(
    lattice=dict(lattice="sc", J=[1.0, 0.2], K=0.1), L=(6, 5, 4), pbc=(True, True, False), field=(0, 0, 0.3),
    cfg={"layers": [{"material": "Co", "thickness": 2}, {"material": "Pt", "thickness": 2}],
         "materials": [{"name": "Co", "moment_scale": 1.2, "anisotropy_scale": 0.5, "Ku": [0.3, [0, 0, 1]],
                        "alpha": 0.05, "ks": 0.04},
                       {"name": "Pt", "moment_scale": 0.6, "Kc": 0.02, "cubic_axes": [[0.8, 0.6, 0], [-0.6, 0.8, 0],
                                                                                     [0, 0, 1]],
                        "K2": 0.01, "K3": 0.005, "K_axis": [0, 1, 1], "T": 40.0}],
         "pinned": [{"region": "bottom", "layers": 1, "direction": [1, 0, 1]}],
         "temperature_profile": {"axis": "x", "T0": 10.0, "T1": 100.0},
         "exchange_scale": [[1, 1, 1], [1, 1, 0.5], [1, 0.5, 1]], "neel_surface": 0.05})
STRUCTURE_ALLOY = dict(
    lattice=dict(lattice="triangular", J=[1.0]), L=(10, 10, 1), pbc=(True, True, True), field=(0, 0, 0.5),
    cfg={"core_shell": {"radius": 6.0, "core": "A", "shell": "B"},
         "alloy": {"host": "B", "guest": "C", "fraction": 0.3, "seed": 3},
         "materials": [{"name": "A", "moment_scale": 1.5, "T": 30.0},
                       {"name": "C", "alpha": 0.2, "Ku": [0.2, [1, 0, 0]]}],
         "pinned": [{"region": "material:A", "direction": [0, 1, 0]}],
         "temperature_profile": {"spot": [10.0, 5.0, 0.0], "sigma": 4.0, "T0": 5.0, "T1": 50.0}})


def case_structure():
    """magex's apply_structure on both configurations: the arrays it leaves on the system; for the film
    also B_eff, the energy and 10 thermal Heun steps (T = 20 K, alpha 0.1 where no material sets it) on
    injected normals."""
    from magex.structure import apply_structure
    from magex import terms
    with open(os.path.join(OUT, "structure_cases.json"), "w") as f:
        json.dump({"film": STRUCTURE_FILM, "alloy": STRUCTURE_ALLOY}, f, indent=1)
    print("wrote structure_cases.json")
    for name, case in (("structure_film.npz", STRUCTURE_FILM), ("structure_alloy.npz", STRUCTURE_ALLOY)):
        s = _system(case["lattice"], case["L"], case["field"], case["pbc"])
        apply_structure(s, case["cfg"])
        out = dict(material=s.material, material_names=np.array(s.material_names), mu=s.mu, A=s.A,
                   pinned=s.pinned, pinned_dir=s.pinned_dir,
                   alpha_site=getattr(s, "alpha_site", None) if getattr(s, "alpha_site", None) is not None
                   else np.full(s.N, np.nan),
                   site_T=s.site_T if s.site_T is not None else np.full(s.N, np.nan))
        for t in s.terms:
            if isinstance(t, terms.CubicAnisotropy):
                out["Kc"], out["C"] = t.Kc, t.C
            elif isinstance(t, terms.UniaxialHigher):
                out["K2"], out["K3"], out["u"] = t.K2, t.K3, t.u
        if hasattr(s, "surface"):
            out["surface"] = s.surface
        J = s.Jsp.tocoo()
        out.update(J_row=J.row, J_col=J.col, J_val=J.data)
        if name == "structure_film.npz":
            e0 = s.init_spins("random", seed=3)
            out.update(e0=e0, B0=s.effective_field(e0), E0=np.array(s.energy(e0)))
            g = np.random.default_rng(8).normal(size=(10, s.N, 3))
            sol = LLGSolver(s, alpha=0.1, dt=2e-4, temperature=20.0, seed=0)
            sol.rng = _InjectedNormals(g)
            e = e0.copy()
            for _ in range(10):
                e = sol.step(e)
            out.update(noise=g, e10=e)
        _save(name, **out)


if __name__ == "__main__":
    case_structure()
    models_json()
    couplings_tri()
    case_c1()
    case_tri()
    case_square()
    case_bcc_perbond()
    case_charge()
