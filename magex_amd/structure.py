"""The [structure] set-up of magex (magex/structure.py:86-239) on complete supercells.

``apply_structure(system, cfg)`` edits a freshly built SpinSystem in place and maps everything onto the
arrays the LLG kernels take: per-site moments, on-site tensors (per-material anisotropy scaling, extra
uniaxial Ku, Neel surface pair anisotropy), cubic and higher-uniaxial term arrays, per-site damping
(``system.alpha_site``, NaN = the solver's alpha) and temperature (``system.site_T``, NaN = the
solver's T), pinned regions with their directions (``system.pinned``, ``system.pinned_dir``), and
interface exchange scaling between materials (as per-bond tensors).

Materials come from ``layers`` (along axis 2, in cells), ``core_shell`` (by radius about ``center``) and
``alloy`` (random substitution).  What would remove sites (a ``shape`` other than film/bulk/none,
``vacancies``, ``roughness``) and the terms that are not on the device path (``quadruplets``,
``dipolar``) raise NotImplementedError: the kernels run complete supercells (a stencil over cells).
"""
from __future__ import annotations

import numpy as np

_AXIS = {"x": 0, "y": 1, "z": 2}
_REGION_AXIS = {"bottom": 2, "top": 2, "xmin": 0, "xmax": 0, "ymin": 1, "ymax": 1}


def site_temperature(system, T):
    """Per-site temperature array, or the scalar T when no map is set (magex/structure.py:234-239)."""
    sT = getattr(system, "site_T", None)
    if sT is None:
        return T
    return np.where(np.isnan(sT), T, sT)


def _material_id(system, name):
    """Index of material `name`, appended to system.material_names on first use (structure.py:68-71)."""
    names = system.material_names
    if name not in names:
        names.append(name)
    return names.index(name)


def _centre(system, center):
    if center is not None:
        return np.asarray(center, float)
    lo, hi = system.positions.min(0), system.positions.max(0)
    return 0.5 * (lo + hi)


def _assign_materials(system, st):
    """Material of every site from layers, core_shell and alloy, in that order (structure.py:99-118)."""
    cz = system.cells[system.cell_index][:, 2]
    if "layers" in st:
        z = 0
        for layer in st["layers"]:
            t = int(layer["thickness"])
            system.material[(cz >= z) & (cz < z + t)] = _material_id(system, layer["material"])
            z += t
    if "core_shell" in st:
        cs = st["core_shell"]
        dist = np.linalg.norm(system.positions - _centre(system, st.get("center")), axis=1)
        inner = dist <= float(cs["radius"])
        system.material[inner] = _material_id(system, cs.get("core", "core"))
        system.material[~inner] = _material_id(system, cs.get("shell", "shell"))
    if "alloy" in st:
        al = st["alloy"]
        host = system.material == _material_id(system, al.get("host", "bulk"))
        guest = _material_id(system, al["guest"])
        draw = np.random.default_rng(al.get("seed", 1)).random(system.N)
        system.material[host & (draw < float(al["fraction"]))] = guest


def _refuse_site_removal(st):
    shape = st.get("shape")
    if shape not in (None, "none", "film", "bulk"):
        raise NotImplementedError(f"shape '{shape}' removes sites: the device path runs complete supercells")
    for key in ("vacancies", "roughness"):
        if key in st:
            raise NotImplementedError(f"'{key}' removes sites: the device path runs complete supercells")
    for key in ("quadruplets", "dipolar"):
        if st.get(key):
            raise NotImplementedError(f"'{key}' is not on the device path")


def _min_image(system, v):
    """Displacements reduced to the nearest periodic image of the supercell."""
    V = np.array(system.L)[:, None] * system.uc.cell
    f = v @ np.linalg.inv(V)
    for a in range(3):
        if system.pbc[a]:
            f[:, a] -= np.round(f[:, a])
    return f @ V


def neel_surface(system, ks, sel):
    """Neel pair anisotropy E_i = (ks / 2) sum_{j nn} (e_i . u_ij)^2 on the under-coordinated sites of sel
    (fewer nearest neighbours than the most coordinated site), folded into their on-site tensors
    (magex/structure.py:273-290); sets system.surface and returns the number of surface sites."""
    src, dst = [], []
    for k in range(len(system.stencil)):
        a, b = system.neighbour(k)
        ok = b >= 0
        src.append(a[ok])
        dst.append(b[ok])
    src, dst = np.concatenate(src), np.concatenate(dst)
    d = _min_image(system, system.positions[dst] - system.positions[src])
    r = np.linalg.norm(d, axis=1)
    nn = np.abs(r - r.min()) < 1e-3
    count = np.bincount(src[nn], minlength=system.N)
    surf = (count < count.max()) & np.asarray(sel, bool)
    u = d[nn] / r[nn, None]
    T = np.zeros((system.N, 3, 3))
    np.add.at(T, src[nn], 0.5 * ks * u[:, :, None] * u[:, None, :])
    A = system.A.copy()
    A[surf] += T[surf]
    system.A = A
    system.surface = surf
    return int(surf.sum())


def _region(system, spec):
    """Sites of a pinned region: the first / last `layers` cells along an axis, or a material."""
    reg = spec.get("region", "bottom")
    if reg.startswith("material:"):
        return system.material == system.material_names.index(reg.split(":", 1)[1])
    c = system.cells[system.cell_index][:, _REGION_AXIS[reg]]
    n = int(spec.get("layers", 1))
    if reg in ("bottom", "xmin", "ymin"):
        return c < c.min() + n
    return c > c.max() - n


def _temperature_profile(system, tp):
    """Linear ramp T0 -> T1 along an axis, or a Gaussian spot of width sigma about `spot`."""
    T0, T1 = float(tp["T0"]), float(tp["T1"])
    P = system.positions
    if "spot" in tp:
        r2 = ((P - np.asarray(tp["spot"], float)) ** 2).sum(1)
        return T0 + (T1 - T0) * np.exp(-r2 / (2.0 * float(tp["sigma"]) ** 2))
    x = P[:, _AXIS[tp.get("axis", "x")]]
    return T0 + (T1 - T0) * (x - x.min()) / max(np.ptp(x), 1e-9)


def apply_structure(system, cfg, log=None):
    """Apply a [structure] table to a SpinSystem in place (magex/structure.py:86-231); returns the system."""
    from .terms import CubicAnisotropy, UniaxialHigher
    if not cfg:
        return system
    log = log or (lambda *a: None)
    st = dict(cfg)
    _refuse_site_removal(st)
    materials = st.get("materials", [])
    for m in materials:
        _material_id(system, m["name"])
    _assign_materials(system, st)
    # per-material moments, anisotropy, terms, damping and temperature
    N = system.N
    mu, A = system.mu.copy(), system.A.copy()
    Kc, K2, K3 = np.zeros(N), np.zeros(N), np.zeros(N)
    cubic_axes, K_axis = None, (0, 0, 1)
    alpha_site, site_T = getattr(system, "alpha_site", None), system.site_T
    for m in materials:
        sel = system.material == system.material_names.index(m["name"])
        if "moment_scale" in m:
            mu[sel] *= float(m["moment_scale"])
        if "anisotropy_scale" in m:
            A[sel] *= float(m["anisotropy_scale"])
        if "Ku" in m:                                  # extra uniaxial E = -K (e.u)^2
            ax = np.asarray(m["Ku"][1], float)
            ax = ax / np.linalg.norm(ax)
            A[sel] -= float(m["Ku"][0]) * np.outer(ax, ax)
        if "Kc" in m:
            Kc[sel] = float(m["Kc"])
            cubic_axes = m.get("cubic_axes", cubic_axes)
        if "K2" in m or "K3" in m:
            K2[sel] = float(m.get("K2", 0.0))
            K3[sel] = float(m.get("K3", 0.0))
            K_axis = m.get("K_axis", K_axis)
        if "ks" in m:                                  # on the tensors as they stand at this point
            system.A = A
            neel_surface(system, float(m["ks"]), sel)
            A = system.A.copy()
        if "alpha" in m:
            alpha_site = np.full(N, np.nan) if alpha_site is None else alpha_site.copy()
            alpha_site[sel] = float(m["alpha"])
        if "T" in m:
            site_T = np.full(N, np.nan) if site_T is None else site_T.copy()
            site_T[sel] = float(m["T"])
    system.mu = mu
    system.A = A
    if "exchange_scale" in st:
        E = np.asarray(st["exchange_scale"], float)
        f = np.ones((len(system.material_names),) * 2)
        f[: E.shape[0], : E.shape[1]] = E
        system.scale_exchange(f)
        log(f"  structure: interface exchange scaling {E.tolist()} over materials {system.material_names}")
    if "neel_surface" in st:
        n = neel_surface(system, float(st["neel_surface"]), np.ones(N, bool))
        log(f"  structure: Neel surface anisotropy ks = {st['neel_surface']} meV on {n} surface sites")
    system.has_A = bool(np.max(np.abs(system.A)) > 0)
    if np.any(Kc):
        system.terms.append(CubicAnisotropy(Kc, cubic_axes))
    if np.any(K2) or np.any(K3):
        system.terms.append(UniaxialHigher(K2, K3, K_axis))
    system.alpha_site = alpha_site
    # pinned regions, then the temperature profile under the per-material temperatures
    pinned = None if system.pinned is None else system.pinned.copy()
    for spec in st.get("pinned", []):
        sel = _region(system, spec)
        if pinned is None:
            pinned = np.zeros(N, bool)
        if getattr(system, "pinned_dir", None) is None:
            system.pinned_dir = system.ref.copy()
        pinned |= sel
        if "direction" in spec:
            d = np.asarray(spec["direction"], float)
            system.pinned_dir[sel] = d / np.linalg.norm(d)
        log(f"  structure: {int(sel.sum())} spins pinned ({spec.get('region', 'bottom')})")
    system.pinned = pinned
    if st.get("temperature_profile"):
        prof = _temperature_profile(system, st["temperature_profile"])
        site_T = prof if site_T is None else np.where(np.isnan(site_T), prof, site_T)
    if site_T is not None and np.all(np.isnan(site_T)):
        site_T = None
    system.site_T = site_T
    system.touch()
    return system
