"""Textbook spin Hamiltonians in the results.json format (drop-in for magex.models).

    make_model("triangular", J=[5.0, -0.5, 0.2], D=1.5, K=0.3)

Same convention and output as magex/models.py:40-123: couplings per ordered pair,
E = -sum_{i!=j} e_i.J_ij.e_j; interfacial DMI D_ij = D (n x r_ij)/|r_ij| (bulk: D r_ij/|r_ij|) on
the first shell, J^A with (J^A_yz, J^A_zx, J^A_xy) = D; single-ion A = -K (k k^T - I/3).

One lattice is added to magex's table: "bcc_conventional", bcc in its two-site cubic cell, so that a
64 x 64 x 64 supercell holds 524 288 spins on a cube.
"""
from __future__ import annotations

import itertools
import json

import numpy as np

_S3 = np.sqrt(3.0)
_TRI = np.array([[1.0, 0.0, 0.0], [0.5, _S3 / 2, 0.0], [0.0, 0.0, 10.0]])

# name: (cell rows in units of a, fractional basis, periodic dimensions)  (magex/models.py:20-32)
LATTICES = {
    "chain": (np.diag([1.0, 10.0, 10.0]), [[0, 0, 0]], 1),
    "square": (np.diag([1.0, 1.0, 10.0]), [[0, 0, 0]], 2),
    "triangular": (_TRI, [[0, 0, 0]], 2),
    "honeycomb": (_TRI * _S3, [[1 / 3, 1 / 3, 0], [2 / 3, 2 / 3, 0]], 2),
    "kagome": (np.array([[2.0, 0.0, 0.0], [1.0, _S3, 0.0], [0.0, 0.0, 10.0]]), [[0, 0, 0], [0.5, 0, 0], [0, 0.5, 0]], 2),
    "sc": (np.eye(3), [[0, 0, 0]], 3),
    "bcc": (np.array([[-0.5, 0.5, 0.5], [0.5, -0.5, 0.5], [0.5, 0.5, -0.5]]), [[0, 0, 0]], 3),
    "fcc": (np.array([[0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]]), [[0, 0, 0]], 3),
    "bcc_conventional": (np.eye(3), [[0, 0, 0], [0.5, 0.5, 0.5]], 3),
}


def skew(D):
    """Antisymmetric tensor with (M_yz, M_zx, M_xy) = D."""
    dx, dy, dz = D
    return np.array([[0.0, dz, -dy], [-dz, 0.0, dx], [dy, -dx, 0.0]])


def _cell(lattice, a, order, afm_signs):
    """Cell (rows, A), fractional basis, periodic dims and reference signs."""
    rows, basis, pdim = LATTICES[lattice]
    cell = np.asarray(rows, float) * a
    basis = np.asarray(basis, float)
    signs = np.ones(len(basis))
    if order != "afm":
        return cell, basis, pdim, signs
    if afm_signs is not None:
        return cell, basis, pdim, np.asarray(afm_signs, float)
    # checkerboard / rocksalt doubling of one-site bipartite lattices (magex/models.py:58-74)
    doubled = {
        "chain": (np.vstack([cell[0] * 2, cell[1:]]), [[0, 0, 0], [0.5, 0, 0]]),
        "square": (np.array([[a, a, 0], [a, -a, 0], [0, 0, 10 * a]], float), [[0, 0, 0], [0.5, 0.5, 0]]),
        "sc": (np.array([[a, a, 0], [a, 0, a], [0, a, a]], float), [[0, 0, 0], [0.5, 0.5, 0.5]]),
    }
    if lattice not in doubled:
        raise ValueError("give afm_signs for this lattice")
    c, b = doubled[lattice]
    return c, np.asarray(b, float), pdim, np.array([1.0, -1.0])


def make_model(lattice="sc", a=3.0, J=(1.0,), D=0.0, dmi="interfacial", normal=(0, 0, 1), K=0.0,
               K_axis=(0, 0, 1), moment=2.0, order="ferro", afm_signs=None, Jani=None, mesh=None,
               Bq=None, K4=0.0):
    """results.json-style dict (magex/models.py:40-116).

    J    couplings per neighbour shell (meV per ordered pair), J[0] nearest neighbours
    D    nearest-neighbour DMI magnitude (meV), dmi "interfacial" (normal n) or "bulk"
    K    single-ion anisotropy along K_axis (K > 0 easy axis)
    Bq   biquadratic constants per shell; K4 four-spin plaquette constant (both stored, not on the
         device path)
    """
    if lattice not in LATTICES:
        raise ValueError(f"lattice must be one of {list(LATTICES)}")
    cell, basis, pdim, signs = _cell(lattice, a, order, afm_signs)
    ns = len(basis)
    pos = basis @ cell
    span = np.arange(-4, 5)
    axes = [span, span if pdim >= 2 else np.zeros(1, int), span if pdim >= 3 else np.zeros(1, int)]
    Rs = np.array(list(itertools.product(*axes)), int)
    cand = []
    for i, j in itertools.product(range(ns), range(ns)):
        vecs = pos[j] + Rs @ cell - pos[i]
        dist = np.round(np.linalg.norm(vecs, axis=1), 6)
        for R, v, d in zip(Rs, vecs, dist):
            if i == j and not R.any():
                continue
            cand.append((float(d), i, j, R, v))
    shells = sorted({c[0] for c in cand})[:len(J)]
    shell_of = {d: k for k, d in enumerate(shells)}
    n = np.asarray(normal, float)
    n = n / np.linalg.norm(n)
    if mesh is None:
        mesh = [64, 64 if pdim >= 2 else 1, 64 if pdim >= 3 else 1]
    pairs = []
    for d, i, j, R, v in cand:
        s = shell_of.get(d)
        if s is None:
            continue
        Jt = float(J[s]) * np.eye(3)
        if s == 0:
            if Jani is not None:
                Jt = Jt + np.asarray(Jani, float)
            if D != 0.0:
                rhat = v / np.linalg.norm(v)
                Jt = Jt + skew(D * (np.cross(n, rhat) if dmi == "interfacial" else rhat))
        p = dict(i=i, j=j, R=[int(x) for x in R], nu=[int(R[k]) % mesh[k] for k in range(3)], dist=d,
                 vec=v.tolist(), J=Jt.tolist(), resid=0.0, degeneracy=1)
        if Bq is not None and s < len(Bq) and Bq[s]:
            p["B"] = float(Bq[s])
        pairs.append(p)
    k = np.asarray(K_axis, float)
    k = k / np.linalg.norm(k)
    A = -K * (np.outer(k, k) - np.eye(3) / 3.0)
    z = np.array([0.0, 0.0, 1.0])
    return dict(version="model", sites=list(range(ns)), symbols=["X"] * ns, cell=cell.tolist(), positions=pos.tolist(),
                mesh=list(mesh), moments=(signs[:, None] * moment * z[None, :]).tolist(), charges=[0.0] * ns,
                refs=[z.tolist()], full_tensor=True, A=[A.tolist()] * ns, mu=0.0, mag_pos=list(range(ns)),
                lig_pos=[], pairs=pairs, corrections={}, lloyd=None, mae=None, symmetry=None,
                four_spin=dict(K=float(K4)) if K4 else None,
                model=dict(lattice=lattice, a=a, J=list(J), D=D, dmi=dmi, K=K, moment=moment, order=order))


def write_model(path, **kw):
    """make_model(**kw) written to path as JSON (magex/models.py:119-123)."""
    with open(path, "w") as f:
        json.dump(make_model(**kw), f)
    return path
