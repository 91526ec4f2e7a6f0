"""Classical atomistic spin model (drop-in for magex.spinmodel), evaluated by the HIP kernels.

Energy (magex convention, meV; unit spins e_i; every ordered pair counted; magex/spinmodel.py:3-11):

    E = sum_i e_i.A_i.e_i  -  sum_{i != j} e_i.J_ij.e_j  -  mu_B sum_i mu_i e_i.B

What changes against magex: the supercell couplings are kept as a stencil over cells, one entry
(s -> t, cell offset R, tensor J) per unit-cell coupling, instead of a 3N x 3N CSR matrix
(magex/spinmodel.py:211-241).  A coupling that reaches a site's own periodic image is folded into the
on-site tensor exactly as magex does (A_i <- A_i - sym(J), spinmodel.py:224-229).  The CSR ``Jsp`` is
still available (built on first access) and may be assigned: assigning it, or calling
``set_bond_tensors``, switches the model to per-bond tensors on the same stencil.

Field, gradient and energy run on the GPU through libmxllg.so; they accept numpy arrays (copied to
and from the device) or torch tensors on the GPU (kept there).
"""
from __future__ import annotations

import itertools
import json
import os
from dataclasses import dataclass, field as dc_field

import numpy as np

from .constants import GAMMA, K_B, MU_B  # noqa: F401  (re-exported, as magex.spinmodel does)


@dataclass
class UnitCellModel:
    """Magnetic sites of one unit cell and their couplings (meV) (magex/spinmodel.py:34-51)."""
    cell: np.ndarray
    positions: np.ndarray
    symbols: list
    moments: np.ndarray
    ref_dirs: np.ndarray
    signs: np.ndarray
    couplings: list                  # [(i, j, R(3,), J(3,3))], every ordered pair
    A: np.ndarray
    info: dict = dc_field(default_factory=dict)
    biquadratic: list = dc_field(default_factory=list)
    four_spin: float = 0.0

    @property
    def nsites(self):
        return len(self.moments)


def _shortest_images(nu, mesh, fi, fj, cell, tol=1e-6):
    """All periodic images R = nu + mesh m (m in -2..2) of a DFT-mesh pair at the minimum distance
    (magex/spinmodel.py:54-61); the pair's coupling is shared equally among them."""
    m = np.array(list(itertools.product(range(-2, 3), repeat=3)))
    R = np.asarray(nu)[None, :] + np.asarray(mesh)[None, :] * m
    d = np.linalg.norm((fj[None, :] + R - fi[None, :]) @ cell, axis=1)
    keep = d - d.min() < tol
    return [tuple(int(x) for x in r) for r in R[keep]], float(d.min())


def load_results(path, J="mft", anisotropic=True, sia=True, rcut=None, jtol=1e-4,
                 moments=None, axis=None, biquadratic=None):
    """results.json -> UnitCellModel (magex/spinmodel.py:64-151).

    J "mft" or a key of the file's "corrections"; rcut (A, default 2 x nearest-neighbour distance) and
    jtol (meV) drop far or weak pairs; moments overrides |M|; axis sets the reference axis.
    biquadratic="lloyd" is not supported (biquadratic terms are not on the device path).
    """
    if isinstance(path, dict):
        d = path
    elif str(path).endswith(".gz"):
        import gzip
        with gzip.open(path, "rt") as f:
            d = json.load(f)
    else:
        with open(path) as f:
            d = json.load(f)
    if biquadratic:
        raise NotImplementedError("biquadratic (Lloyd) couplings are not on the device path")
    mp = list(d.get("mag_pos") or range(len(d["sites"])))
    cell = np.array(d["cell"], float)
    pos_all = np.array(d["positions"], float)
    frac = pos_all @ np.linalg.inv(cell)
    M = np.array(d["moments"], float)
    ax = np.array(axis if axis is not None else d["refs"][-1], float)
    ax = ax / np.linalg.norm(ax)
    sgn = np.sign(M[mp] @ ax)
    sgn[sgn == 0] = 1
    mom = np.linalg.norm(M[mp], axis=1) if moments is None else np.asarray(moments, float)
    corr = None
    if J != "mft":
        if J not in (d.get("corrections") or {}):
            raise ValueError(f"results.json has no '{J}' couplings; available: {list(d.get('corrections') or {})}")
        corr = np.array(d["corrections"][J])
    site = {k: n for n, k in enumerate(mp)}
    raw = []
    for p in d["pairs"]:
        if p["i"] not in site or p["j"] not in site:
            continue
        i, j = site[p["i"]], site[p["j"]]
        Jt = np.array(p["J"], float)
        iso = np.trace(Jt) / 3.0
        if corr is not None:
            Jt = Jt - iso * np.eye(3) + corr[i, j][tuple(p["nu"])] * np.eye(3)
        if not anisotropic:
            Jt = np.trace(Jt) / 3.0 * np.eye(3)
        imgs, dist = _shortest_images(p["nu"], d["mesh"], frac[p["i"]], frac[p["j"]], cell)
        raw.append((i, j, imgs, dist, Jt))
    dmin = min(r[3] for r in raw)
    rcut = 2.0 * dmin if rcut is None else rcut
    couplings, shells = [], set()
    J0_all = J0_kept = 0.0
    for i, j, imgs, dist, Jt in raw:
        w = sgn[i] * sgn[j] * np.trace(Jt) / 3.0
        J0_all += w
        if dist > rcut + 1e-6 or np.max(np.abs(Jt)) < jtol:
            continue
        J0_kept += w
        shells.add(round(dist, 3))
        share = Jt / len(imgs)
        couplings.extend((i, j, np.array(R, int), share) for R in imgs)
    A = np.array(d["A"], float)[mp] if sia else np.zeros((len(mp), 3, 3))
    info = dict(source=os.path.abspath(path) if not isinstance(path, dict) else None, J=J, rcut=float(rcut),
                nn_distance=float(dmin), shells=sorted(float(x) for x in shells),
                ncouplings_per_cell=len(couplings), J0_fraction=float(J0_kept / J0_all) if J0_all else 1.0,
                axis=ax.tolist(), biquadratic=False)
    K4 = float((d.get("four_spin") or {}).get("K", 0.0))
    return UnitCellModel(cell, pos_all[mp], [d["symbols"][k] for k in mp], mom, sgn[:, None] * ax[None, :], sgn,
                         couplings, A, info, [], K4)


def normalize(e):
    """Row-wise unit vectors (magex/spinmodel.py:553-554)."""
    return e / np.linalg.norm(e, axis=1, keepdims=True)


class SpinSystem:
    """Supercell of the magnetic unit cell (magex/spinmodel.py:188-252); couplings as a cell stencil."""

    def __init__(self, uc, supercell=(10, 10, 10), pbc=(True, True, True), field=(0.0, 0.0, 0.0)):
        if uc.four_spin:
            raise NotImplementedError("four-spin terms are not on the device path")
        self.uc = uc
        self.L = tuple(int(x) for x in supercell)
        self.pbc = tuple(bool(x) for x in pbc)
        ns = uc.nsites
        Lv = np.array(self.L)
        # cells row-major with the last axis fastest; site = cell * ns + sublattice (spinmodel.py:197-206)
        cells = np.stack(np.meshgrid(*[np.arange(n) for n in self.L], indexing="ij"), -1).reshape(-1, 3)
        self.cells = cells
        self.ncell = len(cells)
        self.N = self.ncell * ns
        self.sublattice = np.tile(np.arange(ns), self.ncell)
        self.cell_index = np.repeat(np.arange(self.ncell), ns)
        self.positions = cells[self.cell_index] @ uc.cell + uc.positions[self.sublattice]
        self.mu = uc.moments[self.sublattice].astype(float)
        self.signs = uc.signs[self.sublattice].astype(float)
        self.ref = uc.ref_dirs[self.sublattice].copy()
        A_sub = np.array(uc.A, float).copy()
        self.stencil = []            # [(s, t, R(3,), J(3,3))] after self-image folding
        self.self_folded = 0
        for (i, j, R, Jt) in uc.couplings:
            R = np.asarray(R, int)
            if i == j and all((R[a] % Lv[a] == 0) if self.pbc[a] else (R[a] == 0) for a in range(3)):
                A_sub[i] -= 0.5 * (Jt + Jt.T)
                self.self_folded += self.ncell
                continue
            self.stencil.append((int(i), int(j), R.copy(), np.asarray(Jt, float).copy()))
        self._A = A_sub[self.sublattice].copy()
        self.has_A = bool(np.max(np.abs(self._A)) > 0)
        self.terms = []
        self._pinned = None
        self._site_T = None
        self.material = np.zeros(self.N, int)
        self.material_names = ["bulk"]
        self.grid_index = np.arange(self.N)
        self._Jsp = None
        self._bond_J = None          # per-bond tensors on the forward stencil entries, or None
        self._version = 0
        self._dev = {}
        self.set_field(field)

    # ------------------------------------------------------------------ mutable model arrays
    def touch(self):
        """Mark the model as edited (needed after in-place edits of A, mu, pinned, site_T)."""
        self._version += 1
        self._dev.clear()

    @property
    def A(self):
        return self._A

    @A.setter
    def A(self, v):
        self._A = np.asarray(v, float)
        self.touch()

    @property
    def pinned(self):
        return self._pinned

    @pinned.setter
    def pinned(self, v):
        self._pinned = None if v is None else np.asarray(v, bool)
        self.touch()

    @property
    def site_T(self):
        return self._site_T

    @site_T.setter
    def site_T(self, v):
        self._site_T = None if v is None else np.asarray(v, float)
        self.touch()

    def set_field(self, B):
        self.B = np.asarray(B, float).reshape(3)

    # ------------------------------------------------------------------ stencil / bonds
    def forward_entries(self):
        """Indices of the stencil entries that store per-bond tensors: s < t, or s == t and R > 0
        lexicographically (each unordered bond once; its partner entry reads J^T)."""
        return [k for k, (s, t, R, _) in enumerate(self.stencil) if s < t or (s == t and tuple(R) > (0, 0, 0))]

    def neighbour(self, k):
        """(src, dst) site arrays of stencil entry k over all cells (dst = -1 where an open axis cuts it)."""
        s, t, R, _ = self.stencil[k]
        Lv = np.array(self.L)
        tgt = self.cells + R[None, :]
        ok = np.ones(self.ncell, bool)
        for a in range(3):
            if self.pbc[a]:
                tgt[:, a] %= Lv[a]
            else:
                ok &= (tgt[:, a] >= 0) & (tgt[:, a] < Lv[a])
        cn = (tgt[:, 0] * Lv[1] + tgt[:, 1]) * Lv[2] + tgt[:, 2]
        ns = self.uc.nsites
        src = np.arange(self.ncell) * ns + s
        dst = np.where(ok, cn * ns + t, -1)
        return src, dst

    def set_bond_tensors(self, J):
        """Per-bond tensors: J (nfwd, ncell, 3, 3), J[f, c] acts from site (c, s_f) on its neighbour
        (c + R_f, t_f) for forward entry f (``forward_entries``); the reverse bond carries J^T."""
        fw = self.forward_entries()
        J = np.asarray(J, float)
        if J.shape != (len(fw), self.ncell, 3, 3):
            raise ValueError(f"bond tensors must have shape {(len(fw), self.ncell, 3, 3)}")
        keys = {(s, t, tuple(R)) for s, t, R, _ in self.stencil}
        for k in fw:
            s, t, R, _ = self.stencil[k]
            if (t, s, tuple(-R)) not in keys:
                raise ValueError("per-bond tensors need every stencil entry's reverse partner")
        self._bond_J = J
        self._Jsp = None
        self.touch()

    @property
    def bond_tensors(self):
        return self._bond_J

    def scale_exchange(self, f):
        """Multiply every coupling J_ij by f[material_i, material_j] (interface exchange,
        magex/spinmodel.py:433-442).  The couplings stop being translation invariant, so they become
        per-bond tensors; f must be symmetric and partner entries transposed (J_ji = J_ij^T), which
        per-bond storage (each unordered bond once) needs."""
        f = np.asarray(f, float)
        if not np.array_equal(f, f.T):
            raise NotImplementedError("an asymmetric exchange scale breaks J_ji = J_ij^T")
        fw = self.forward_entries()
        base = self._bond_J
        if base is None:
            base = self._table_as_bonds(fw)
        J = np.empty_like(base)
        for n, k in enumerate(fw):
            src, dst = self.neighbour(k)
            w = np.where(dst >= 0, f[self.material[src], self.material[np.maximum(dst, 0)]], 1.0)
            J[n] = base[n] * w[:, None, None]
        self.set_bond_tensors(J)

    def _table_as_bonds(self, fw):
        """The stencil's translation-invariant tensors as per-bond storage on the forward entries."""
        keys = {(s, t, tuple(R)): J for s, t, R, J in self.stencil}
        out = np.empty((len(fw), self.ncell, 3, 3))
        for n, k in enumerate(fw):
            s, t, R, J = self.stencil[k]
            P = keys.get((t, s, tuple(-R)))
            if P is None or not np.allclose(P, J.T, rtol=0, atol=1e-14 * max(np.abs(J).max(), 1e-300)):
                raise NotImplementedError("per-bond storage needs every entry's partner with J_ji = J_ij^T")
            out[n] = J[None]
        return out

    @property
    def Jsp(self):
        """The 3N x 3N CSR coupling matrix of magex (spinmodel.py:230-241), built on first access."""
        if self._Jsp is None:
            self._Jsp = self._build_csr()
        return self._Jsp

    @Jsp.setter
    def Jsp(self, M):
        """Adopt a coupling matrix: its 3x3 blocks must sit on the stencil's bonds (per-bond storage)."""
        self.set_bond_tensors(self._blocks_from_csr(M))
        self._Jsp = M

    def _entry_blocks(self, k):
        """(src, dst, J (ncell, 3, 3)) of entry k as the model currently holds it."""
        src, dst = self.neighbour(k)
        if self._bond_J is None:
            J = np.broadcast_to(self.stencil[k][3], (self.ncell, 3, 3))
        else:
            fw = self.forward_entries()
            if k in fw:
                J = self._bond_J[fw.index(k)]
            else:
                s, t, R, _ = self.stencil[k]
                kp = next(q for q in fw if self.stencil[q][0] == t and self.stencil[q][1] == s
                          and np.array_equal(self.stencil[q][2], -R))
                psrc, pdst = self.neighbour(kp)
                # the partner's tensor sits at the partner's source cell, i.e. this entry's target cell
                pos = np.full(self.N, -1)
                pos[psrc] = np.arange(self.ncell)
                cidx = np.where(dst >= 0, pos[np.maximum(dst, 0)], 0)
                J = self._bond_J[fw.index(kp)][cidx].transpose(0, 2, 1)
        return src, dst, J

    def _build_csr(self):
        import scipy.sparse as sp
        rows, cols, vals = [], [], []
        for k in range(len(self.stencil)):
            src, dst, J = self._entry_blocks(k)
            ok = dst >= 0
            src, dst, J = src[ok], dst[ok], np.asarray(J)[ok]
            for a in range(3):
                for b in range(3):
                    v = J[:, a, b]
                    nz = v != 0.0
                    rows.append(3 * src[nz] + a)
                    cols.append(3 * dst[nz] + b)
                    vals.append(v[nz])
        cat = lambda x, t: np.concatenate(x) if x else np.zeros(0, t)  # noqa: E731
        M = sp.csr_matrix((cat(vals, float), (cat(rows, int), cat(cols, int))), shape=(3 * self.N, 3 * self.N))
        M.sum_duplicates()
        return M

    def _blocks_from_csr(self, M):
        """Per-bond tensors of the forward entries read out of a CSR matrix on the stencil pattern."""
        C = M.tocoo()
        key = (C.row // 3).astype(np.int64) * self.N + C.col // 3
        uniq, inv = np.unique(key, return_inverse=True)
        blocks = np.zeros((len(uniq), 3, 3))
        np.add.at(blocks, (inv.ravel(), C.row % 3, C.col % 3), C.data)
        seen = np.zeros(len(uniq), bool)
        fw = self.forward_entries()
        out = np.zeros((len(fw), self.ncell, 3, 3))
        for k in range(len(self.stencil)):
            src, dst = self.neighbour(k)
            ok = dst >= 0
            q = src.astype(np.int64) * self.N + np.maximum(dst, 0)
            pos = np.searchsorted(uniq, q)
            hit = ok & (pos < len(uniq)) & (uniq[np.minimum(pos, len(uniq) - 1)] == q)
            seen[pos[hit]] = True
            if k in fw:
                out[fw.index(k)][hit] = blocks[pos[hit]]
        if not seen.all():
            raise NotImplementedError("couplings outside the supercell stencil are not supported")
        allkeys = []
        for k in range(len(self.stencil)):
            src, dst = self.neighbour(k)
            allkeys.append(src[dst >= 0].astype(np.int64) * self.N + dst[dst >= 0])
        allkeys = np.concatenate(allkeys) if allkeys else np.zeros(0, np.int64)
        if len(np.unique(allkeys)) != len(allkeys):
            raise NotImplementedError("two stencil entries share a site pair: enlarge the supercell")
        return out

    # ------------------------------------------------------------------ device evaluation
    def device_model(self, precision="fp64", layout=None):
        """Cached DeviceModel of the current arrays (rebuilt after touch() / setters); layout None is the
        library default, "checker" the two-colour cell order of the Monte Carlo sweeps."""
        from .device import DeviceModel
        state = (self._version, len(self.terms), tuple(id(t) for t in self.terms))
        key = (precision, layout) + state
        dm = self._dev.get(key)
        if dm is None:
            for k in [k for k in self._dev if k[2:] != state]:   # models of older arrays
                del self._dev[k]
            dm = self._dev[key] = DeviceModel(self, precision, layout=layout)
        dm.set_field(self.B)
        return dm

    def _eval(self, e, kind):
        from .device import as_device, like_input
        dm = self.device_model("fp64" if not _is_fp32(e) else "fp32")
        ed = as_device(e, dm.dtype)
        return like_input(dm.field(ed, kind), e)

    def exchange_field(self, e):
        """h_i = sum_j J_ij e_j (meV), (N, 3) (magex/spinmodel.py:454-456)."""
        return self._eval(e, 2)

    def gradient(self, e):
        """dE/de_i (meV) (magex/spinmodel.py:458-469)."""
        return self._eval(e, 1)

    def effective_field(self, e):
        """B_eff,i = -(1/(mu_B mu_i)) dE/de_i (T) (magex/spinmodel.py:471-473)."""
        return self._eval(e, 0)

    def energy(self, e, per_spin=False):
        """Total energy (meV) (magex/spinmodel.py:475-484)."""
        from .device import as_device
        dm = self.device_model("fp64" if not _is_fp32(e) else "fp32")
        E = float(dm.observables(as_device(e, dm.dtype))[0])
        return E / self.N if per_spin else E

    # ------------------------------------------------------------------ configurations
    def init_spins(self, kind="reference", axis=(0, 0, 1), seed=None, cone=0.0, path=None, **texture):
        """Initial spins (magex/spinmodel.py:487-515): random (uniform on S^2), ferro, reference,
        reference_axis, file; skyrmion/spiral textures need magex.structure (not provided)."""
        rng = np.random.default_rng(seed)
        if kind == "random":
            e = rng.normal(size=(self.N, 3))
        elif kind == "ferro":
            e = np.tile(np.asarray(axis, float), (self.N, 1))
        elif kind == "reference":
            e = self.ref.copy()
        elif kind == "reference_axis":
            ax = np.asarray(axis, float) / np.linalg.norm(axis)
            e = self.signs[:, None] * ax[None, :]
        elif kind == "file":
            e = np.load(path)["spins"] if str(path).endswith(".npz") else np.loadtxt(path, delimiter=",")[:, -3:]
            if e.shape != (self.N, 3):
                raise ValueError(f"spin file has shape {e.shape}, expected {(self.N, 3)}")
        elif kind in ("spiral", "skyrmion", "skyrmions", "hopfion", "domain", "bubble"):
            raise NotImplementedError(f"texture '{kind}' initial states are outside the LLG hot path")
        else:
            raise ValueError(f"unknown initial state '{kind}'")
        if cone > 0:
            e = e + cone * rng.normal(size=e.shape)
        e = normalize(e)
        if self.pinned is not None and hasattr(self, "pinned_dir"):
            e[self.pinned] = self.pinned_dir[self.pinned]
        return e

    def order_parameter(self, e):
        """|sum_i s_i e_i| / N (magex/spinmodel.py:517-519)."""
        from .device import as_device
        dm = self.device_model("fp64" if not _is_fp32(e) else "fp32")
        o = dm.observables(as_device(e, dm.dtype))
        return float(np.linalg.norm(np.asarray(o[4:7])) / self.N)

    def magnetization(self, e, weighted=False):
        """<e> (or mu-weighted) (magex/spinmodel.py:521-524)."""
        e = _host(e)
        if weighted:
            return (self.mu[:, None] * e).sum(0) / self.mu.sum()
        return e.mean(0)

    def kind(self):
        s, mu = self.uc.signs, self.uc.moments
        if np.all(s == s[0]):
            return "ferromagnet"
        return "antiferromagnet" if abs(np.sum(s * mu)) < 1e-3 * np.sum(mu) else "ferrimagnet"

    def summary(self):
        i = self.uc.info
        return "\n".join([
            f"spin model: {self.uc.nsites} magnetic sites/cell, moments {np.round(self.uc.moments, 3).tolist()} mu_B, "
            f"reference = {self.kind()}",
            f"  couplings: rcut = {i.get('rcut', 0):.3f} A, {len(i.get('shells', []))} shells, "
            f"{len(self.stencil)} stencil entries",
            f"  supercell {self.L}, pbc {self.pbc}: N = {self.N} spins",
            f"  single-ion anisotropy: {'on' if self.has_A else 'off'};  field B = {self.B.tolist()} T"])


def _is_fp32(e):
    return str(getattr(e, "dtype", "")) in ("float32", "torch.float32")


def _host(e):
    if hasattr(e, "detach"):
        return e.detach().cpu().numpy()
    return np.asarray(e)


def build_from_config(cfg, base="."):
    """(UnitCellModel, [spinmodel] table) from a config dict (magex/spinmodel.py:557-573)."""
    sm = dict(cfg.get("spinmodel", {}))
    if "model" in sm:
        from .models import make_model
        src = make_model(**sm["model"])
    else:
        src = sm.get("results", "magex_results/results.json")
        if not os.path.isabs(src):
            src = os.path.join(base, src)
    uc = load_results(src, J=sm.get("J", "mft"), anisotropic=sm.get("anisotropic", True), sia=sm.get("sia", True),
                      rcut=sm.get("rcut"), jtol=sm.get("jtol", 1e-4), moments=sm.get("moments"), axis=sm.get("axis"),
                      biquadratic=sm.get("biquadratic"))
    return uc, sm
