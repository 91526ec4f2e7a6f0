"""Host-side logic of magex_amd (no GPU): model generator, results reader, stencil assembly, per-bond
storage, coupling-table compaction, triangulation, and the C ABI library's exports."""
import ctypes
import json
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import GOLDEN, ROOT, golden
from oracle import magex_ref as R
from magex_amd import models as M
from magex_amd import spinmodel as S


def test_make_model_matches_magex():
    d = json.load(open(os.path.join(GOLDEN, "make_model_cases.json")))
    for name, kw in d["kwargs"].items():
        mine, ref = M.make_model(**kw), d["models"][name]
        assert [(p["i"], p["j"], p["R"], p["nu"]) for p in mine["pairs"]] == \
               [(q["i"], q["j"], q["R"], q["nu"]) for q in ref["pairs"]], name
        assert np.allclose([p["J"] for p in mine["pairs"]], [q["J"] for q in ref["pairs"]], rtol=0, atol=1e-15)
        for k in ("cell", "positions", "moments", "A", "mesh", "refs"):
            assert np.allclose(mine[k], ref[k], rtol=0, atol=1e-15), (name, k)


def test_load_results_matches_magex(tmp_path):
    g = golden("load_results_tri.npz")
    p = str(tmp_path / "m.json")
    M.write_model(p, lattice="triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3)
    uc = S.load_results(p)
    assert np.array_equal([c[0] for c in uc.couplings], g["i"])
    assert np.array_equal([c[2] for c in uc.couplings], g["R"])
    assert np.allclose([c[3] for c in uc.couplings], g["J"], rtol=0, atol=1e-15)
    for k in ("A", "moments", "signs", "ref_dirs"):
        assert np.allclose(getattr(uc, k), g[k])
    assert len(S.load_results(p, rcut=3.15, jtol=0.3).couplings) == int(g["n_filtered"])


CASES = [
    (dict(lattice="sc", J=[1.0], K=0.1), (4, 4, 4), (True,) * 3),
    (dict(lattice="sc", J=[1.0, 0.3]), (2, 3, 1), (True,) * 3),            # self images fold into A
    (dict(lattice="triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3), (6, 5, 1), (True,) * 3),
    (dict(lattice="triangular", J=[1.0, 0.2], D=0.3), (5, 4, 1), (False, True, False)),  # open axis
    (dict(lattice="honeycomb", J=[1.0, 0.1], D=0.1), (4, 3, 1), (True,) * 3),
    (dict(lattice="kagome", J=[1.0], K=0.05, K_axis=(1, 0, 1)), (3, 3, 1), (True,) * 3),
    (dict(lattice="sc", J=[-1.0], order="afm"), (3, 3, 3), (True,) * 3),
    (dict(lattice="bcc_conventional", J=[20.0, 10.0], moment=2.2), (3, 3, 3), (True,) * 3),
    (dict(lattice="sc", J=[1.0], Jani=[[0.1, 0.02, 0], [0.02, -0.05, 0], [0, 0, -0.05]]), (3, 3, 3), (True,) * 3),
]


@pytest.mark.parametrize("kw,L,pbc", CASES)
def test_stencil_assembly_equals_csr(kw, L, pbc):
    """Summing the stencil over cells gives magex's 3N x 3N matrix and on-site tensors exactly."""
    d = M.make_model(**kw)
    s = S.SpinSystem(S.load_results(d), L, pbc)
    o = R.SpinSystem(R.load_results(d), L, pbc)
    assert abs(s.Jsp - o.Jsp).max() < 1e-14
    assert np.abs(s.A - o.A).max() < 1e-15
    assert np.array_equal(s.positions, o.positions) and np.array_equal(s.mu, o.mu)


def test_bond_tensors_roundtrip_through_csr():
    """Golden per-bond bcc: assigning the CSR stores each bond once; reading it back is exact."""
    g = golden("bcc_perbond.npz")
    d = json.load(open(os.path.join(GOLDEN, "bcc_conventional_model.json")))
    s = S.SpinSystem(S.load_results(d), (4, 4, 4))
    src, dst, blocks = g["src"], g["dst"], g["blocks"]
    rows = (3 * src[:, None, None] + np.arange(3)[None, :, None]).repeat(3, 2).ravel()
    cols = (3 * dst[:, None, None] + np.arange(3)[None, None, :]).repeat(3, 1).ravel()
    Jsp = sp.csr_matrix((blocks.ravel(), (rows, cols)), shape=(3 * s.N, 3 * s.N))
    s.Jsp = Jsp
    assert s.bond_tensors.shape == (len(s.forward_entries()), s.ncell, 3, 3)
    s._Jsp = None
    assert abs(s.Jsp - Jsp).max() == 0.0


def test_coupling_table_compaction():
    from magex_amd.device import coupling_table
    iso = [np.eye(3) * 2.0, np.eye(3) * 2.0, np.eye(3) * -1.0]
    t, ty = coupling_table(iso)
    assert t.shape == (2, 1) and list(ty) == [0, 0, 1]
    dm = [np.eye(3) + M.skew([0.1, 0.2, 0.3]), np.eye(3)]
    t, _ = coupling_table(dm)
    assert t.shape == (2, 4) and np.allclose(t[0], [1.0, 0.1, 0.2, 0.3])
    t, _ = coupling_table([np.diag([1.0, 2.0, 3.0])])
    assert t.shape == (1, 9)


@pytest.mark.parametrize("lat,L", [("triangular", (8, 7, 1)), ("square", (6, 6, 1)), ("honeycomb", (4, 4, 1))])
def test_triangulation_matches_oracle_charge(lat, L):
    """The cached CCW triangle list reproduces the oracle's Berg-Luscher sum for random spins."""
    from magex_amd.texture import triangles
    d = M.make_model(lattice=lat, J=[1.0])
    s = S.SpinSystem(S.load_results(d), L)
    e = R.normalize(np.random.default_rng(2).normal(size=(s.N, 3)))
    tri = triangles(s)
    s1, s2, s3 = e[tri[:, 0]], e[tri[:, 1]], e[tri[:, 2]]
    num = np.einsum("ij,ij->i", s1, np.cross(s2, s3))
    den = 1 + np.einsum("ij,ij->i", s1, s2) + np.einsum("ij,ij->i", s2, s3) + np.einsum("ij,ij->i", s3, s1)
    Q = (2 * np.arctan2(num, den)).sum() / (4 * np.pi)
    assert abs(Q - R.topological_charge(s.positions, s.uc.cell, s.L, s.pbc, e)) < 1e-10


# ------------------------------------------------------------------ the C ABI library
def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "mxllg.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(mx_\w+)\(", txt, re.M)))


def _structured(name):
    """The drop-in's apply_structure on one of the golden [structure] configurations."""
    from magex_amd import structure
    case = json.load(open(os.path.join(GOLDEN, "structure_cases.json")))[name]
    s = S.SpinSystem(S.load_results(M.make_model(**case["lattice"])), case["L"], case["pbc"], case["field"])
    return structure.apply_structure(s, case["cfg"])


@pytest.mark.parametrize("case,lattice,L,field,seed", [
    ("c1_sc16.npz", dict(lattice="sc", J=[1.0], K=0.1, moment=2.0), (16, 16, 16), (0, 0, 0.5), 0),
    ("tri12.npz", dict(lattice="triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3), (12, 12, 1), (0, 0, 2.0), 5),
])
def test_init_spins_random_is_magex_start(case, lattice, L, field, seed):
    """SpinSystem.init_spins("random", seed) reproduces magex's start vectors bit for bit (numpy's PCG64
    normals, normalised; magex/spinmodel.py:495-496, 512)."""
    s = S.SpinSystem(S.load_results(M.make_model(**lattice)), L, field=field)
    assert np.array_equal(s.init_spins("random", seed=seed), golden(case)["e0"])


@pytest.mark.parametrize("name", ["film", "alloy"])
def test_apply_structure_matches_magex(name):
    """Materials (layers, core-shell, alloy), moments, on-site tensors (scaling, Ku, Neel surface), cubic
    and higher-uniaxial terms, damping and temperature maps, pinned sites and directions, and interface
    exchange (as per-bond tensors) equal what magex's apply_structure leaves on its system."""
    import scipy.sparse as sp
    from magex_amd.terms import CubicAnisotropy, UniaxialHigher
    g = golden(f"structure_{name}.npz")
    s = _structured(name)
    assert list(s.material_names) == [str(x) for x in g["material_names"]]
    assert np.array_equal(s.material, g["material"])
    assert np.array_equal(s.mu, g["mu"])
    assert np.abs(s.A - g["A"]).max() < 1e-15
    assert np.array_equal(s.pinned, g["pinned"])
    assert np.abs(s.pinned_dir - g["pinned_dir"]).max() < 1e-15
    for key in ("alpha_site", "site_T"):
        mine = getattr(s, key)
        mine = np.full(s.N, np.nan) if mine is None else mine
        assert np.array_equal(np.isnan(mine), np.isnan(g[key])), key
        assert np.allclose(mine[~np.isnan(mine)], g[key][~np.isnan(g[key])], rtol=1e-15, atol=0), key
    cub = [t for t in s.terms if isinstance(t, CubicAnisotropy)]
    uni = [t for t in s.terms if isinstance(t, UniaxialHigher)]
    if "Kc" in g.files:
        assert np.array_equal(cub[0].Kc, g["Kc"]) and np.array_equal(cub[0].C, g["C"])
        assert np.array_equal(uni[0].K2, g["K2"]) and np.array_equal(uni[0].K3, g["K3"])
        assert np.allclose(uni[0].u, g["u"], rtol=0, atol=1e-16)
    else:
        assert not cub and not uni
    if "surface" in g.files:
        assert np.array_equal(s.surface, g["surface"])
    ref = sp.csr_matrix((g["J_val"], (g["J_row"], g["J_col"])), shape=(3 * s.N, 3 * s.N))
    assert abs(s.Jsp - ref).max() < 1e-15


def test_apply_structure_refuses_what_the_device_path_cannot_run():
    from magex_amd import structure
    s = S.SpinSystem(S.load_results(M.make_model(lattice="sc", J=[1.0])), (4, 4, 4))
    for cfg in ({"shape": "sphere", "size": 3.0}, {"vacancies": 0.1}, {"roughness": {"amplitude": 1}},
                {"dipolar": True}, {"quadruplets": {"K": 0.1, "list": [[0, 1, 2, 3]]}}):
        with pytest.raises(NotImplementedError):
            structure.apply_structure(s, cfg)
    s.material_names += ["a", "b"]
    with pytest.raises(NotImplementedError, match="asymmetric"):
        s.scale_exchange([[1, 1, 1], [1, 1, 0.5], [1, 0.7, 1]])


def test_gneb_helpers_closed_forms():
    """The skyrmion profile (magex/gneb.py:270-281): core down at the centre, in-plane at R = radius,
    +z far away, unit length, Neel (radial) at helicity 0 and Bloch (azimuthal) at 90 degrees; the
    geodesic distance of a uniform rotation by th of N spins is sqrt(N) th, also at th = pi; the
    Arrhenius lifetime and its infinite limits."""
    from magex_amd import gneb
    s = S.SpinSystem(S.load_results(M.make_model(lattice="square", J=[1.0])), (41, 41, 1), (False, False, False))
    c = s.positions[:, :2].mean(0)
    a0 = float(np.linalg.norm(s.positions[1] - s.positions[0]))      # lattice spacing
    rad = 4 * a0
    for hel, comp in ((0.0, 0), (90.0, 1)):
        m = gneb.skyrmion_on_lattice(s, radius=rad, helicity_deg=hel)
        r = s.positions[:, :2] - c
        R = np.linalg.norm(r, axis=1)
        assert np.abs(np.linalg.norm(m, axis=1) - 1).max() < 1e-15
        assert np.allclose(m[R == 0], [[0, 0, -1]], atol=1e-15)
        on = np.abs(R - rad) < 1e-9
        assert on.any() and np.abs(m[on, 2]).max() < 1e-14
        far = R > 3 * rad
        assert m[far, 2].min() > 0.99
        ring = on & (np.abs(r[:, 1]) < 1e-9) & (r[:, 0] > 0)           # the point (+radius, 0)
        assert np.allclose(m[ring][0], [1, 0, 0] if comp == 0 else [0, 1, 0], atol=1e-12)
    a = np.random.default_rng(0).normal(size=(50, 3))
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    for th in (0.3, np.pi):
        ax = np.cross(a, [0.0, 0.0, 1.0])
        ax /= np.linalg.norm(ax, axis=1, keepdims=True)
        b = a * np.cos(th) + np.cross(ax, a) * np.sin(th) + ax * np.sum(ax * a, 1, keepdims=True) * (1 - np.cos(th))
        assert abs(gneb.geodesic_distance(a, b) - np.sqrt(50) * th) < 1e-12
    assert gneb.arrhenius(1.0, 1.0, 0.0) == float("inf") and gneb.arrhenius(800.0, 1.0, 1.0) == float("inf")
    assert abs(gneb.arrhenius(10.0, 0.5, 4.0, tau0=2e-9) / (2e-9 * np.exp(5.0)) - 1) < 1e-14


def test_library_exports_every_declared_symbol():
    from magex_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared_symbols()
    assert len(names) >= 18
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures and header disagree"


def test_library_argument_errors_without_gpu():
    """Argument checks that fire before any device call map to ValueError / NotImplementedError."""
    from magex_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    Lv = np.array([4, 4, 4], np.int32)
    pb = np.array([1, 1, 1], np.int32)
    with pytest.raises(ValueError, match="precision"):
        _lib.check(L.mx_model_create(7, 1, _lib.hptr(Lv), _lib.hptr(pb), ctypes.byref(h)))
    with pytest.raises(NotImplementedError, match="sites per cell"):
        _lib.check(L.mx_model_create(0, 99, _lib.hptr(Lv), _lib.hptr(pb), ctypes.byref(h)))
    with pytest.raises(ValueError, match="null model"):
        _lib.check(L.mx_model_set_field(None, _lib.hptr(np.zeros(3))))
    # the 32-bit per-site offsets of the kernels (up to 9 N) cap a model at (2^31 - 1) / 9 sites
    big = np.array([1024, 1024, 228], np.int32)          # 239 075 328 sites: 9 N > 2^31
    with pytest.raises(NotImplementedError, match="238609293 sites"):
        _lib.check(L.mx_model_create(1, 1, _lib.hptr(big), _lib.hptr(pb), ctypes.byref(h)))
    with pytest.raises(NotImplementedError, match="238609293 sites"):
        _lib.check(L.mx_model_create(1, 2, _lib.hptr(np.array([1024, 1024, 114], np.int32)), _lib.hptr(pb),
                                     ctypes.byref(h)))


def test_product_has_no_cpu_path():
    """magex_amd never imports the oracle and refuses the NumPy backend."""
    for dirpath, _, files in os.walk(os.path.join(ROOT, "magex_amd")):
        for f in files:
            if f.endswith(".py"):
                txt = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in re.sub(r"#.*|\"\"\"[\s\S]*?\"\"\"", "", txt), f
    from magex_amd.llg import LLGSolver
    s = S.SpinSystem(S.load_results(M.make_model("sc", J=[1.0])), (2, 2, 2))
    with pytest.raises((RuntimeError, ValueError)):
        LLGSolver(s, backend="numpy")


def test_build_recipe_compiles_every_source_for_gfx950_without_packed_fp32():
    """Every HIP source is one parallel compile unit with the gfx950 target, packed FP32 instructions
    switched off and single-expression FMA contraction, and the link step takes all of their objects."""
    from magex_amd.csrc import build as B
    here = os.path.dirname(B.__file__)
    on_disk = sorted(f for f in os.listdir(here) if f.endswith(".hip"))
    assert sorted(B.SOURCES) == on_disk
    objs = []
    for src in B.SOURCES:
        cmd = B.compile_command(src, src + ".o")
        assert "--offload-arch=gfx950" in cmd and "-c" in cmd and "-ffp-contract=on" in cmd
        k = cmd.index("-target-feature")
        assert cmd[k - 1] == "-Xclang" and cmd[k + 1] == "-Xclang" and cmd[k + 2] == "-packed-fp32-ops"
        assert os.path.join(here, src) in cmd
        objs.append(cmd[cmd.index("-o") + 1])
    link = B.link_command(objs, "out.so")
    assert "-shared" in link and "--offload-arch=gfx950" in link
    assert all(o in link for o in objs) and link[-2:] == ["-o", "out.so"]
