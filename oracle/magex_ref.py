"""CPU oracle: a NumPy/scipy restatement of magex's atomistic LLG hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``magex_amd`` imports this module; only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it, and only as the
checker.  Each function cites the magex file:line it restates (all paths inside the magex tree).

Pinning (what each part is checked against, see tests/test_oracle.py):
  make_model / load_results / SpinSystem assembly  -> golden vectors of magex itself (tests/golden)
  exchange_field / gradient / energy               -> golden vectors + central finite differences
  CubicAnisotropy / UniaxialHigher                 -> closed-form energies on symmetry directions,
                                                      central differences of their energies
  solver_field (spin-orbit torque, field_func,     -> closed-form single-spin solutions: precession
               extra_field)                           phase under B_z(t); damping-like relaxation
                                                      tan(theta/2) = tan(theta0/2) exp(-gamma B_DL t);
                                                      an extra_field equal to an on-site tensor's field
                                                      follows the model with that tensor
  heun_step                                        -> golden trajectories (T=0 and injected noise),
                                                      Larmor closed form, damped-spin closed form
  rotate / depondt_step                            -> Rodrigues closed form (rotation about an axis),
                                                      second-order convergence to the exact LLG flow
                                                      (a fine-step integration), agreement with Heun
  topological_charge                               -> golden charges + integer-Q textures
  neighbour_blocks                                 -> exact equality of its mat-vec with Jsp @ e
"""
from __future__ import annotations

import itertools
import json

import numpy as np
import scipy.sparse as sp

# magex/spinmodel.py:28-30
MU_B = 5.7883818060e-2      # meV / T
K_B = 8.617333262e-2        # meV / K
GAMMA = 0.1760859630        # rad / (ps T)

# magex/models.py:20-32: (cell rows in units of a, fractional basis, number of periodic dims)
LATTICES = {
    "chain": (np.diag([1.0, 10.0, 10.0]), [[0, 0, 0]], 1),
    "square": (np.diag([1.0, 1.0, 10.0]), [[0, 0, 0]], 2),
    "triangular": (np.array([[1, 0, 0], [0.5, np.sqrt(3) / 2, 0], [0, 0, 10.0]]), [[0, 0, 0]], 2),
    "honeycomb": (np.array([[1, 0, 0], [0.5, np.sqrt(3) / 2, 0], [0, 0, 10.0]]) * np.sqrt(3),
                  [[1 / 3, 1 / 3, 0], [2 / 3, 2 / 3, 0]], 2),
    "kagome": (np.array([[2, 0, 0], [1, np.sqrt(3), 0], [0, 0, 10.0]]), [[0, 0, 0], [0.5, 0, 0], [0, 0.5, 0]], 2),
    "sc": (np.eye(3), [[0, 0, 0]], 3),
    "bcc": (np.array([[-0.5, 0.5, 0.5], [0.5, -0.5, 0.5], [0.5, 0.5, -0.5]]), [[0, 0, 0]], 3),
    "fcc": (np.array([[0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]]), [[0, 0, 0]], 3),
}


def skew(D):
    """Antisymmetric 3x3 with (M_yz, M_zx, M_xy) = D (magex/models.py:35-37)."""
    return np.array([[0.0, D[2], -D[1]], [-D[2], 0.0, D[0]], [D[1], -D[0], 0.0]])


def make_model(lattice="sc", a=3.0, J=(1.0,), D=0.0, dmi="interfacial", normal=(0, 0, 1), K=0.0,
               K_axis=(0, 0, 1), moment=2.0, order="ferro", afm_signs=None, Jani=None, mesh=None,
               Bq=None, K4=0.0):
    """results.json-style dict of a textbook lattice (magex/models.py:40-116)."""
    cell0, basis, pdim = LATTICES[lattice]
    cell = np.array(cell0, float) * a
    basis = np.array(basis, float)
    signs = np.ones(len(basis))
    if order == "afm":                                      # models.py:58-74
        if afm_signs is not None:
            signs = np.array(afm_signs, float)
        elif lattice == "chain":
            cell[0] *= 2
            basis = np.array([[0, 0, 0], [0.5, 0, 0]])
            signs = np.array([1.0, -1.0])
        elif lattice == "square":
            cell = np.array([[a, a, 0], [a, -a, 0], [0, 0, 10 * a]])
            basis = np.array([[0, 0, 0], [0.5, 0.5, 0]])
            signs = np.array([1.0, -1.0])
        elif lattice == "sc":
            cell = np.array([[a, a, 0], [a, 0, a], [0, a, a]])
            basis = np.array([[0, 0, 0], [0.5, 0.5, 0.5]])
            signs = np.array([1.0, -1.0])
        else:
            raise ValueError("give afm_signs for this lattice")
    ns = len(basis)
    pos = basis @ cell
    span = range(-4, 5)                                     # models.py:78-87
    cand = []
    for i in range(ns):
        for j in range(ns):
            for R in itertools.product(span, span if pdim >= 2 else [0], span if pdim >= 3 else [0]):
                if i == j and R == (0, 0, 0):
                    continue
                r = pos[j] + np.array(R) @ cell - pos[i]
                cand.append((round(float(np.linalg.norm(r)), 6), i, j, R, r))
    shells = sorted({c[0] for c in cand})[:len(J)]
    nvec = np.asarray(normal, float) / np.linalg.norm(normal)
    mesh = mesh or [64, 64 if pdim >= 2 else 1, 64 if pdim >= 3 else 1]
    pairs = []
    for d, i, j, R, r in cand:                              # models.py:92-107
        if d not in shells:
            continue
        s = shells.index(d)
        Jt = J[s] * np.eye(3)
        if Jani is not None and s == 0:
            Jt = Jt + np.asarray(Jani, float)
        if s == 0 and D != 0.0:
            rh = r / np.linalg.norm(r)
            Jt = Jt + skew(D * (np.cross(nvec, rh) if dmi == "interfacial" else rh))
        pr = dict(i=i, j=j, R=list(R), nu=[int(R[k]) % mesh[k] for k in range(3)], dist=d, vec=r.tolist(),
                  J=Jt.tolist(), resid=0.0, degeneracy=1)
        if Bq is not None and s < len(Bq) and Bq[s]:
            pr["B"] = float(Bq[s])
        pairs.append(pr)
    kax = np.asarray(K_axis, float) / np.linalg.norm(K_axis)
    Aion = -K * (np.outer(kax, kax) - np.eye(3) / 3)        # models.py:108-109
    ref = np.array([0, 0, 1.0])
    return dict(version="model", sites=list(range(ns)), symbols=["X"] * ns, cell=cell.tolist(), positions=pos.tolist(),
                mesh=mesh, moments=(signs[:, None] * moment * ref[None, :]).tolist(), charges=[0.0] * ns,
                refs=[ref.tolist()], full_tensor=True, A=[Aion.tolist()] * ns, mu=0.0, mag_pos=list(range(ns)),
                lig_pos=[], pairs=pairs, corrections={}, lloyd=None, mae=None, symmetry=None,
                four_spin=dict(K=float(K4)) if K4 else None,
                model=dict(lattice=lattice, a=a, J=list(J), D=D, dmi=dmi, K=K, moment=moment, order=order))


def all_images(nu, mesh, fi, fj, cell, tol=1e-6):
    """Shortest periodic images of a DFT-mesh pair vector (magex/spinmodel.py:54-61)."""
    nu, mesh = np.array(nu), np.array(mesh)
    cand = []
    for m in itertools.product((-2, -1, 0, 1, 2), repeat=3):
        R = nu + mesh * np.array(m)
        cand.append((np.linalg.norm((fj + R - fi) @ cell), tuple(int(x) for x in R)))
    dmin = min(c[0] for c in cand)
    return [R for d, R in cand if d - dmin < tol], dmin


class UnitCell:
    """Plain record of what magex's UnitCellModel carries on the hot path (spinmodel.py:34-47)."""

    def __init__(self, cell, positions, moments, ref_dirs, signs, couplings, A):
        self.cell, self.positions, self.moments = cell, positions, moments
        self.ref_dirs, self.signs, self.couplings, self.A = ref_dirs, signs, couplings, A

    @property
    def nsites(self):
        return len(self.moments)


def load_results(src, anisotropic=True, sia=True, rcut=None, jtol=1e-4, moments=None, axis=None):
    """results.json (path or dict) -> UnitCell, J="mft" only (magex/spinmodel.py:64-151)."""
    if isinstance(src, dict):
        d = src
    else:
        with open(src) as f:
            d = json.load(f)
    mp = list(d.get("mag_pos") or range(len(d["sites"])))
    cell = np.array(d["cell"], float)
    pos_all = np.array(d["positions"], float)
    frac = pos_all @ np.linalg.inv(cell)
    M = np.array(d["moments"], float)
    ax = np.array(axis if axis is not None else d["refs"][-1], float)
    ax /= np.linalg.norm(ax)
    s = np.sign(M[mp] @ ax)
    s[s == 0] = 1
    mom = np.linalg.norm(M[mp], axis=1) if moments is None else np.asarray(moments, float)
    raw = []
    for p in d["pairs"]:
        if p["i"] not in mp or p["j"] not in mp:
            continue
        Jt = np.array(p["J"], float)
        if not anisotropic:
            Jt = np.trace(Jt) / 3 * np.eye(3)
        imgs, dist = all_images(p["nu"], d["mesh"], frac[p["i"]], frac[p["j"]], cell)
        raw.append((mp.index(p["i"]), mp.index(p["j"]), imgs, dist, Jt))
    dmin = min(r[3] for r in raw)
    rcut = 2.0 * dmin if rcut is None else rcut
    couplings = []
    for i, j, imgs, dist, Jt in raw:                        # spinmodel.py:133-143
        if dist > rcut + 1e-6 or np.max(np.abs(Jt)) < jtol:
            continue
        for R in imgs:
            couplings.append((i, j, np.array(R, int), Jt / len(imgs)))
    A = np.array(d["A"], float)[mp] if sia else np.zeros((len(mp), 3, 3))
    return UnitCell(cell, pos_all[mp], mom, s[:, None] * ax[None, :], s, couplings, A)


class SpinSystem:
    """Supercell with the 3N x 3N CSR coupling matrix (magex/spinmodel.py:191-252)."""

    def __init__(self, uc, supercell=(10, 10, 10), pbc=(True, True, True), field=(0.0, 0.0, 0.0)):
        self.uc = uc
        self.L = tuple(int(x) for x in supercell)
        self.pbc = tuple(bool(x) for x in pbc)
        ns = uc.nsites
        cells = np.array(list(itertools.product(*[range(n) for n in self.L])), int)
        self.ncell = len(cells)
        self.N = self.ncell * ns
        Lv = np.array(self.L)
        cidx = lambda c: (c[:, 0] * Lv[1] + c[:, 1]) * Lv[2] + c[:, 2]  # noqa: E731
        self.sublattice = np.tile(np.arange(ns), self.ncell)
        self.cell_index = np.repeat(np.arange(self.ncell), ns)
        self.cells = cells
        self.positions = cells[self.cell_index] @ uc.cell + uc.positions[self.sublattice]
        self.mu = uc.moments[self.sublattice].astype(float)
        self.signs = uc.signs[self.sublattice].astype(float)
        A = uc.A[self.sublattice].copy()
        rows, cols, vals = [], [], []
        for (i, j, R, Jt) in uc.couplings:
            tgt = cells + R[None, :]
            ok = np.ones(self.ncell, bool)
            for a in range(3):
                if self.pbc[a]:
                    tgt[:, a] %= Lv[a]
                else:
                    ok &= (tgt[:, a] >= 0) & (tgt[:, a] < Lv[a])
            src = cidx(cells[ok]) * ns + i
            dst = cidx(tgt[ok]) * ns + j
            same = src == dst
            if np.any(same):                                 # spinmodel.py:224-229
                A[src[same]] -= 0.5 * (Jt + Jt.T)
                src, dst = src[~same], dst[~same]
            for a in range(3):
                for b in range(3):
                    if Jt[a, b] != 0.0:
                        rows.append(3 * src + a)
                        cols.append(3 * dst + b)
                        vals.append(np.full(len(src), Jt[a, b]))
        rows = np.concatenate(rows) if rows else np.zeros(0, int)
        cols = np.concatenate(cols) if cols else np.zeros(0, int)
        vals = np.concatenate(vals) if vals else np.zeros(0)
        self.Jsp = sp.csr_matrix((vals, (rows, cols)), shape=(3 * self.N, 3 * self.N))
        self.Jsp.sum_duplicates()
        self.A = A
        self.pinned = None
        self.site_T = None
        self.terms = []
        self.B = np.asarray(field, float).reshape(3)

    def set_field(self, B):
        self.B = np.asarray(B, float).reshape(3)

    @property
    def has_A(self):
        return bool(np.max(np.abs(self.A)) > 0)

    def exchange_field(self, e):                             # spinmodel.py:454-456
        return (self.Jsp @ e.reshape(-1)).reshape(-1, 3)

    def gradient(self, e):                                   # spinmodel.py:458-469
        g = -2.0 * self.exchange_field(e) - MU_B * self.mu[:, None] * self.B[None, :]
        g += 2.0 * np.einsum("iab,ib->ia", self.A, e)
        for t in self.terms:
            g += t.gradient(e)
        if self.pinned is not None:
            g[self.pinned] = 0.0
        return g

    def effective_field(self, e):                            # spinmodel.py:471-473
        return -self.gradient(e) / (MU_B * self.mu[:, None])

    def energy(self, e, per_spin=False):                     # spinmodel.py:475-484
        E = -np.sum(e * self.exchange_field(e)) - MU_B * np.sum(self.mu * (e @ self.B))
        E += np.einsum("ia,iab,ib->", e, self.A, e)
        for t in self.terms:
            E += t.energy(e)
        return E / self.N if per_spin else E

    def init_random(self, seed):                             # spinmodel.py:495-496, 512
        return normalize(np.random.default_rng(seed).normal(size=(self.N, 3)))

    def order_parameter(self, e):                            # spinmodel.py:517-519
        return float(np.linalg.norm((self.signs[:, None] * e).sum(0)) / self.N)


class CubicAnisotropy:
    """E_i = Kc_i (a1^2 a2^2 + a2^2 a3^2 + a3^2 a1^2), a = C e_i (terms.py:35-55)."""

    def __init__(self, Kc, axes=None):
        self.Kc = np.asarray(Kc, float)
        self.C = np.eye(3) if axes is None else np.asarray(axes, float)

    def energy(self, e):
        q = (e @ self.C.T) ** 2
        return float(np.sum(self.Kc * (q[:, 0] * q[:, 1] + q[:, 1] * q[:, 2] + q[:, 2] * q[:, 0])))

    def gradient(self, e):
        a = e @ self.C.T
        q = a**2
        da = 2 * a * np.stack([q[:, 1] + q[:, 2], q[:, 0] + q[:, 2], q[:, 0] + q[:, 1]], 1)
        return self.Kc[:, None] * (da @ self.C)


class UniaxialHigher:
    """E_i = -K2_i (e_i.u)^4 - K3_i (e_i.u)^6 (terms.py:58-75)."""

    def __init__(self, K2, K3=0.0, axis=(0, 0, 1)):
        self.K2 = np.asarray(K2, float)
        self.K3 = np.asarray(K3, float) * np.ones_like(self.K2)
        u = np.asarray(axis, float)
        self.u = u / np.linalg.norm(u)

    def energy(self, e):
        c = e @ self.u
        return float(np.sum(-self.K2 * c**4 - self.K3 * c**6))

    def gradient(self, e):
        c = e @ self.u
        return (-4 * self.K2 * c**3 - 6 * self.K3 * c**5)[:, None] * self.u[None, :]


def normalize(e):                                            # spinmodel.py:553-554
    return e / np.linalg.norm(e, axis=1, keepdims=True)


def llg_rhs(e, B, alpha, precession=True):                   # llg.py:21-28
    pre = -GAMMA / (1.0 + alpha * alpha)
    exB = np.cross(e, B)
    out = alpha * np.cross(e, exB)
    if precession:
        out = out + exB
    return pre * out


def sigma(system, alpha, T, dt):
    """Thermal-field amplitude per site, sqrt(2 alpha k_B T / (gamma mu mu_B dt)) (llg.py:106-122)."""
    Ts = T if system.site_T is None else np.where(np.isnan(system.site_T), T, system.site_T)
    s = np.sqrt(2.0 * np.asarray(alpha, float) * K_B * Ts / (GAMMA * system.mu * MU_B * dt))
    return np.broadcast_to(s, (system.N,)).astype(float)


def solver_field(system, sot=None, field_func=None, extra_field=None):   # llg.py:124-136
    """LLGSolver.field: B_eff + spin-orbit torque fields + field_func(t) ((3,) or (N, 3)) + extra_field(e)
    ((N, 3), a function of the spins), zero on pinned sites when any of the extras is present.
    sot = (p unit vector, B_DL, B_FL)."""
    def field(e, t):
        B = system.effective_field(e)
        if sot is not None:
            p, bdl, bfl = sot
            B = B + bdl * np.cross(e, p[None, :]) + bfl * p[None, :]
        if field_func is not None:
            B = B + field_func(t)
        if extra_field is not None:
            B = B + extra_field(e)
        if system.pinned is not None and (sot is not None or field_func is not None or extra_field is not None):
            B[system.pinned] = 0.0
        return B
    return field


def heun_step(system, e, alpha, dt, b=0.0, precession=True, field=None, t=0.0):
    """One stochastic Heun step with thermal field b held over both stages (llg.py:216-227); field(e, t)
    defaults to the system's B_eff (time t in the predictor, t + dt in the corrector)."""
    field = field or (lambda x, _t: system.effective_field(x))
    a = alpha[:, None] if np.ndim(alpha) else alpha
    k1 = llg_rhs(e, field(e, t) + b, a, precession)
    ep = normalize(e + dt * k1)
    k2 = llg_rhs(ep, field(ep, t + dt) + b, a, precession)
    return normalize(e + 0.5 * dt * (k1 + k2))


def rotate(e, w):                                            # llg.py:31-36
    """Rodrigues rotation of unit vectors e (N,3) by rotation vectors w (N,3) (angle |w|)."""
    th = np.linalg.norm(w, axis=1, keepdims=True)
    k = np.where(th > 1e-14, w / np.maximum(th, 1e-300), 0.0)
    c, s_ = np.cos(th), np.sin(th)
    return e * c + np.cross(k, e) * s_ + k * np.einsum("ia,ia->i", k, e)[:, None] * (1 - c)


def depondt_step(system, e, alpha, dt, b=0.0, precession=True, field=None, t=0.0):
    """One Depondt step (Heun on rotation vectors), thermal field b held over both stages
    (llg.py:236-240; fast.py:284-306)."""
    field = field or (lambda x, _t: system.effective_field(x))
    a = alpha[:, None] if np.ndim(alpha) else alpha
    w1 = np.cross(e, llg_rhs(e, field(e, t) + b, a, precession))
    ep = rotate(e, dt * w1)
    w2 = np.cross(ep, llg_rhs(ep, field(ep, t + dt) + b, a, precession))
    return normalize(rotate(e, 0.5 * dt * (w1 + w2)))


def run_heun(system, e, steps, alpha, dt, T=0.0, normals=None, precession=True, integrator="heun", sot=None,
             field_func=None, extra_field=None):
    """steps Heun (or Depondt) steps from t = 0; normals(n) -> (N,3) standard normals of step n
    (injected noise); sot / field_func / extra_field as solver_field."""
    step = {"heun": heun_step, "depondt": depondt_step}[integrator]
    extras = sot is not None or field_func is not None or extra_field is not None
    field = solver_field(system, sot, field_func, extra_field) if extras else None
    sg = sigma(system, alpha, T, dt) if (T > 0 or system.site_T is not None) else None
    for n in range(steps):
        b = 0.0
        if sg is not None:
            b = sg[:, None] * normals(n)
            if system.pinned is not None:
                b[system.pinned] = 0.0
        e = step(system, e, alpha, dt, b, precession, field, n * dt)
    return e


def topological_charge(positions, cell, L, pbc, e, plane="xy", layer=0, tol=0.05):
    """Berg-Luscher charge over a periodic Delaunay triangulation (magex/texture.py:14-63)."""
    from scipy.spatial import Delaunay
    planes = {"xy": (0, 1, 2), "xz": (0, 2, 1), "yz": (1, 2, 0)}
    a, b, n = planes[plane]
    z = positions[:, n]
    levels = np.unique(np.round(z / tol) * tol)
    idx = np.where(np.abs(z - levels[layer]) < tol * 1.01)[0]
    if len(idx) < 3:
        return 0.0
    P, S = positions[idx][:, [a, b]], e[idx]
    V3 = np.array(L)[:, None] * np.asarray(cell)
    order = np.argsort(np.abs(V3[:, n]))
    V = V3[order[:2]][:, [a, b]]
    per = [bool(pbc[k]) for k in order[:2]]
    shifts = [(i, j) for i in ((-1, 0, 1) if per[0] else (0,)) for j in ((-1, 0, 1) if per[1] else (0,))]
    Pt = np.concatenate([P + i * V[0] + j * V[1] for i, j in shifts])
    St = np.concatenate([S for _ in shifts])
    tri = Delaunay(Pt).simplices
    frac = np.linalg.solve(V.T, Pt[tri].mean(1).T).T
    keep = np.all((frac >= -1e-9) & (frac < 1 - 1e-9), axis=1) if len(shifts) > 1 else np.ones(len(tri), bool)
    tri = tri[keep]
    p1, p2, p3 = Pt[tri[:, 0]], Pt[tri[:, 1]], Pt[tri[:, 2]]
    orient = np.sign((p2[:, 0] - p1[:, 0]) * (p3[:, 1] - p1[:, 1]) - (p2[:, 1] - p1[:, 1]) * (p3[:, 0] - p1[:, 0]))
    s1, s2, s3 = St[tri[:, 0]], St[tri[:, 1]], St[tri[:, 2]]
    num = np.einsum("ij,ij->i", s1, np.cross(s2, s3))
    den = 1 + np.einsum("ij,ij->i", s1, s2) + np.einsum("ij,ij->i", s2, s3) + np.einsum("ij,ij->i", s3, s1)
    hand = 1.0 if plane in ("xy", "yz") else -1.0
    return float(hand * (2 * np.arctan2(num, den) * orient).sum() / (4 * np.pi))


def neighbour_blocks(Jsp, N):
    """CSR -> per-site neighbour list + coupling table of width 1, 4 or 9 (magex/fast.py:587-627).

    Returns ptr (N+1,) int64, nbr int32, typ int32, table (ntypes, w) float64.
    """
    J = Jsp.tocoo()
    i, j = J.row // 3, J.col // 3
    key = i.astype(np.int64) * N + j
    uniq, inv = np.unique(key, return_inverse=True)
    Jb = np.zeros((len(uniq), 9))
    np.add.at(Jb, (inv, 3 * (J.row % 3) + J.col % 3), J.data)
    si = uniq // N
    nbr = (uniq % N).astype(np.int32)
    ptr = np.cumsum(np.concatenate([[0], np.bincount(si, minlength=N)])).astype(np.int64)
    if len(Jb) == 0:
        return ptr, nbr, np.zeros(0, np.int32), np.zeros((1, 1))
    scale = max(float(np.abs(Jb).max()), 1e-300)
    keys = np.round(Jb / scale * 2**40).astype(np.int64)
    tk, typ = np.unique(keys, axis=0, return_inverse=True)
    typ = typ.ravel()
    table = np.zeros((len(tk), 9))
    np.add.at(table, typ, Jb)
    table /= np.bincount(typ, minlength=len(tk))[:, None]
    if len(table) > 0.5 * len(Jb):                          # no translational symmetry (fast.py:612-614)
        table, typ = Jb, np.arange(len(Jb))
    M = table.reshape(-1, 3, 3)
    S = 0.5 * (M + M.transpose(0, 2, 1))
    iso = np.trace(M, axis1=1, axis2=2) / 3
    tol = 1e-14 * max(float(np.abs(M).max()), 1e-300)
    if np.all(np.abs(S - iso[:, None, None] * np.eye(3)) <= tol):
        Aq = 0.5 * (M - M.transpose(0, 2, 1))
        D = np.stack([Aq[:, 1, 2], Aq[:, 2, 0], Aq[:, 0, 1]], 1)
        table = iso[:, None].copy() if np.all(np.abs(D) <= tol) else np.column_stack([iso, D])
    return ptr, nbr, typ.astype(np.int32), np.ascontiguousarray(table)


# ----------------------------------------------------------------------------- spectra
HBAR_MEV_PS = 0.6582119569                                   # quantum.py:34


def mq_sample(system_positions, qc, e):
    """m_q = ph.T @ e with ph = exp(-i r q^T) / sqrt(N) (spectra.py:55-66)."""
    ph = np.exp(-1j * system_positions @ qc.T) / np.sqrt(len(e))
    return ph.T @ e


def welch_sqw(mq, ref, dt, nseg):
    """Hann-windowed Welch average of m_q(t) into (E, S_perp, S_par) (spectra.py:67-84)."""
    n = np.asarray(ref, float) / np.linalg.norm(ref)
    a = np.array([1.0, 0, 0]) if abs(n[0]) < 0.9 else np.array([0, 1.0, 0])
    e1 = np.cross(n, a)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(n, e1)
    seg = len(mq) // nseg
    win = np.hanning(seg)
    Sperp = Spar = 0.0
    for k in range(nseg):
        x = mq[k * seg:(k + 1) * seg]
        x = x - x.mean(0, keepdims=True)
        F = np.fft.fft(x * win[:, None, None], axis=0)
        Sperp = Sperp + np.abs(F @ e1) ** 2 + np.abs(F @ e2) ** 2
        Spar = Spar + np.abs(F @ n) ** 2
    E = 2 * np.pi * np.fft.fftfreq(seg, dt) * HBAR_MEV_PS
    order = np.argsort(E)
    return E[order], (Sperp / nseg)[order].T, (Spar / nseg)[order].T
