"""A SpinSystem's arrays in GPU memory, driven through the C ABI of libmxllg.so.

Plays the role of magex's fast.FastModel (magex/fast.py:630-680) and cuda_backend._Device
(magex/cuda_backend.py:307-335): it is built once from a SpinSystem and must be rebuilt after the
system's couplings, on-site tensors, moments or terms are edited (SpinSystem.device_model does that
from the system's edit counter).  Device memory and streams come from PyTorch; every kernel is the
library's.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from ._lib import check, hptr


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("magex_amd needs a HIP device: none is visible")
    return torch


def as_device(e, dtype):
    """numpy / torch (N, 3) -> contiguous device tensor of dtype (torch.float64 / float32)."""
    torch = _torch()
    if isinstance(e, torch.Tensor):
        return e.to(device="cuda", dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(e), dtype=dtype, device="cuda")


def like_input(out, e):
    """Return device result out as numpy if e was numpy, else as the tensor."""
    if isinstance(e, np.ndarray) or not hasattr(e, "device"):
        return out.cpu().numpy()
    return out


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def coupling_table(tensors):
    """Deduplicated coupling table of width 1 (J I), 4 (J I + DMI) or 9 (full), and per-entry types.

    Width 1 / 4 are used when every tensor's symmetric part is isotropic (and, for 1, its antisymmetric
    part vanishes) to 1e-14 relative, as magex compacts its table (magex/fast.py:615-626).
    """
    uniq, typ = [], []
    for J in tensors:
        for k, U in enumerate(uniq):
            if np.array_equal(U, J):
                typ.append(k)
                break
        else:
            typ.append(len(uniq))
            uniq.append(np.asarray(J, float))
    if not uniq:
        return np.zeros((1, 1)), np.zeros(0, np.int32)
    M = np.array(uniq)
    S = 0.5 * (M + M.transpose(0, 2, 1))
    iso = np.trace(M, axis1=1, axis2=2) / 3.0
    tol = 1e-14 * max(float(np.abs(M).max()), 1e-300)
    if np.all(np.abs(S - iso[:, None, None] * np.eye(3)) <= tol):
        Aq = 0.5 * (M - M.transpose(0, 2, 1))
        D = np.stack([Aq[:, 1, 2], Aq[:, 2, 0], Aq[:, 0, 1]], 1)
        table = iso[:, None] if np.all(np.abs(D) <= tol) else np.column_stack([iso, D])
    else:
        table = M.reshape(len(M), 9)
    return np.ascontiguousarray(table, float), np.asarray(typ, np.int32)


MAXTAB = 256   # coupling-table values the kernels hold in their constant block (mxllg_internal.h)

# Cell layout inside the library: "tiled" (3D tiles of 256 cells where the supercell allows it) or
# "rowmajor".  Results are identical; only the memory access pattern changes.
DEFAULT_LAYOUT = os.environ.get("MAGEX_AMD_LAYOUT", "rowmajor")

# Models of at most this many spins replay a captured hipGraph of GRAPH_STEPS steps by default: their
# step kernels take a few microseconds, and on MI355X the captured loop measured 5-9 % faster there
# (sc 16^3, triangular 256^2); on bcc 64^3 (524 288 spins) direct launches measured faster.
GRAPH_MAX_SPINS = 1 << 18
GRAPH_STEPS = 32


class DeviceModel:
    """libmxllg model of a SpinSystem in precision "fp64" or "fp32"."""

    def __init__(self, system, precision="fp64", tiled=None, layout=None):
        from .terms import CubicAnisotropy, UniaxialHigher
        torch = _torch()
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        L = _lib.lib()
        self.lib = L
        self.precision = precision
        self.dtype = torch.float64 if precision == "fp64" else torch.float32
        self.sys = system
        self.N = system.N
        ns = system.uc.nsites
        self.ns = ns
        h = ctypes.c_void_p()
        check(L.mx_model_create(_lib.MX_FP64 if precision == "fp64" else _lib.MX_FP32, ns,
                                hptr(np.array(system.L, np.int32)), hptr(np.array(system.pbc, np.int32)),
                                ctypes.byref(h)))
        self.h = h
        # layout: "rowmajor", "tiled" or "checker" (tiled=True/False is the older spelling)
        if layout is None:
            layout = ("tiled" if tiled else "rowmajor") if tiled is not None else DEFAULT_LAYOUT
        codes = {"rowmajor": 0, "tiled": 1, "checker": 2}
        if layout not in codes:
            raise ValueError(f"layout must be one of {tuple(codes)}")
        check(L.mx_model_set_layout(h, codes[layout]))
        if system.N <= GRAPH_MAX_SPINS:
            check(L.mx_model_set_graph_steps(h, GRAPH_STEPS))
        # exchange stencil + coupling table (or per-bond tensors)
        st = system.stencil
        table, typ = coupling_table([J for *_, J in st])
        ent = np.array([[s, t, R[0], R[1], R[2], ty] for (s, t, R, _), ty in zip(st, typ)], np.int32).reshape(-1, 6)
        bond = system.bond_tensors
        if bond is None and table.size > MAXTAB:
            bond = self._broadcast_bonds(system)
            table, ent[:, 5] = np.zeros((1, 1)), 0
        check(L.mx_model_set_stencil(h, len(ent), hptr(ent), table.shape[0], table.shape[1], hptr(table)))
        if bond is not None:
            check(L.mx_model_set_bond_tensors(h, hptr(np.ascontiguousarray(bond, float))))
        # on-site tensor: one per sublattice when every cell agrees, else per site
        A = np.asarray(system.A, float).reshape(system.ncell, ns, 9)
        if np.all(A == A[:1]):
            check(L.mx_model_set_onsite(h, 0, hptr(np.ascontiguousarray(A[0]))))
        else:
            check(L.mx_model_set_onsite(h, 1, hptr(np.ascontiguousarray(A.reshape(-1, 9)))))
        # kernel-folded on-site terms (magex/fast.py:647-656); anything else is non-local
        has_c = has_u = False
        for t in system.terms:
            if isinstance(t, CubicAnisotropy) and not has_c:
                Kc = np.ascontiguousarray(np.broadcast_to(np.asarray(t.Kc, float), (self.N,)))
                check(L.mx_model_set_cubic(h, hptr(Kc), hptr(np.ascontiguousarray(t.C, float))))
                has_c = True
            elif isinstance(t, UniaxialHigher) and not has_u:
                K2 = np.ascontiguousarray(np.broadcast_to(np.asarray(t.K2, float), (self.N,)))
                K3 = np.ascontiguousarray(np.broadcast_to(np.asarray(t.K3, float), (self.N,)))
                check(L.mx_model_set_uniaxial_higher(h, hptr(K2), hptr(K3), hptr(np.ascontiguousarray(t.u, float))))
                has_u = True
            else:
                raise NotImplementedError(f"{type(t).__name__}: only cubic and higher uniaxial on-site terms run "
                                          "on the device")
        mu = np.asarray(system.mu, float).reshape(system.ncell, ns)
        sg = np.asarray(system.signs, float).reshape(system.ncell, ns)
        if np.all(mu == mu[:1]) and np.all(sg == sg[:1]):
            check(L.mx_model_set_moments(h, 0, hptr(np.ascontiguousarray(mu[0])), hptr(np.ascontiguousarray(sg[0]))))
        else:
            check(L.mx_model_set_moments(h, 1, hptr(np.ascontiguousarray(system.mu, float)),
                                         hptr(np.ascontiguousarray(system.signs, float))))
        if system.pinned is not None:
            check(L.mx_model_set_pinned(h, hptr(np.ascontiguousarray(system.pinned, np.uint8))))
        self.set_field(system.B)
        self._dyn = None
        self._mode = 2       # the library's default step mode
        self._gu_seen = 0    # given-up waits accounted for by checked()

    @staticmethod
    def _broadcast_bonds(system):
        """Translation-invariant tensors written as per-bond storage (for tables too large for the
        kernel arguments); needs J_ji = J_ij^T between partner entries."""
        fw = system.forward_entries()
        st = system.stencil
        for k in fw:
            s, t, R, J = st[k]
            part = [q for q, (s2, t2, R2, _) in enumerate(st) if s2 == t and t2 == s and np.array_equal(R2, -R)]
            if not part or not np.allclose(st[part[0]][3], J.T, rtol=0, atol=1e-14 * max(np.abs(J).max(), 1e-300)):
                raise NotImplementedError("coupling table too large for the kernels and J_ji != J_ij^T")
        return np.stack([np.broadcast_to(st[k][3], (system.ncell, 3, 3)) for k in fw])

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self.lib.mx_model_destroy(h)
            self.h = None

    # ------------------------------------------------------------------ configuration
    def set_field(self, B):
        check(self.lib.mx_model_set_field(self.h, hptr(np.ascontiguousarray(np.asarray(B, float).reshape(3)))))

    def set_dynamics(self, dt, alpha, T, precession=True, integrator="heun"):
        """alpha, T: scalars or per-site arrays (N,); integrator "heun" or "depondt"."""
        a = np.ascontiguousarray(np.atleast_1d(np.asarray(alpha, float)))
        t = np.ascontiguousarray(np.atleast_1d(np.asarray(T, float)))
        codes = {"heun": _lib.MX_HEUN, "depondt": _lib.MX_DEPONDT}
        if integrator not in codes:
            raise ValueError(f"integrator must be one of {tuple(codes)}")
        key = (float(dt), a.tobytes(), t.tobytes(), bool(precession), integrator)
        if key == self._dyn:
            return
        check(self.lib.mx_model_set_dynamics(self.h, float(dt), int(a.size > 1), hptr(a), int(t.size > 1), hptr(t),
                                             int(bool(precession)), codes[integrator]))
        self._dyn = key

    def layout(self):
        """Tile sides of the library's cell layout ((0, 0, 0): row-major)."""
        t = np.zeros(3, np.int32)
        check(self.lib.mx_model_layout(self.h, hptr(t)))
        return tuple(int(x) for x in t)

    def set_graph_steps(self, n):
        """Steps per captured hipGraph step loop (0: launch every step directly)."""
        check(self.lib.mx_model_set_graph_steps(self.h, int(n)))

    def set_noise(self, grade):
        """Thermal-noise generator grade: "double" (default in FP64) or "float" (default in FP32; an FP64
        model set to "float" draws exactly the FP32 model's noise)."""
        codes = {"float": _lib.MX_NOISE_FLOAT, "double": _lib.MX_NOISE_DOUBLE}
        if grade not in codes:
            raise ValueError("noise grade must be 'float' or 'double'")
        check(self.lib.mx_model_set_noise(self.h, codes[grade]))

    def set_step_mode(self, mode):
        """0: two kernel launches per Heun step; 1: one launch per step where the model allows it; 2
        (default): as 1, and runs of two or more steps in persistent launches where they pay; 3: as 2, and
        one walked launch per step for FP32 per-bond grids larger than the GPU holds at once."""
        check(self.lib.mx_model_set_step_mode(self.h, int(mode)))
        self._mode = int(mode)

    def set_wait_limit(self, polls):
        """Polls after which a wait of the one-launch, persistent and adjoint-sweep kernels gives up
        (default 65536; 0: after the first unsuccessful poll), for tests of the give-up paths."""
        check(self.lib.mx_model_set_wait_limit(self.h, int(polls)))

    def checked(self, work, undo):
        """work() (device calls on the current stream), done again after undo() with step mode 1 if a
        wait gave up meanwhile.  A persistent launch that gives up stops early and leaves its spins
        part-way; one launch per step finishes a given-up step in the launch's last workgroup instead, and
        every step form gives the same bits, so the repeat is the result work() should have given.
        Synchronises the device once (twice on a repeat)."""
        work()
        gu = self.step_info()[1]
        if gu != self._gu_seen and self._mode >= 2:  # modes 0 and 1 never stop part-way
            undo()
            mode = self._mode
            self.set_step_mode(min(mode, 1))
            try:
                work()
            finally:
                self.set_step_mode(mode)
            gu = self.step_info()[1]
        self._gu_seen = gu

    def step_info(self):
        """(launch form: 0 two per step, 1 one per step, 2 persistent runs, 3 one walked launch per step;
        waits given up or persistent launches stopped early since creation -- synchronises)."""
        one, gu = ctypes.c_int32(), ctypes.c_int64()
        check(self.lib.mx_model_step_info(self.h, ctypes.byref(one), ctypes.byref(gu)))
        return int(one.value), int(gu.value)

    # ------------------------------------------------------------------ kernels
    def _check_spins(self, e):
        if e.shape != (self.N, 3) or e.dtype != self.dtype or not e.is_contiguous() or e.device.type != "cuda":
            raise ValueError(f"spins must be a contiguous ({self.N}, 3) {self.dtype} device tensor")

    def run(self, e, nsteps, seed, step0=0, noise=None):
        """nsteps Heun steps on the device tensor e, in place; noise: (nsteps, N, 3) normals or None."""
        self._check_spins(e)
        if noise is not None:
            if noise.shape != (nsteps, self.N, 3) or noise.dtype != self.dtype or not noise.is_contiguous():
                raise ValueError(f"noise must be a contiguous ({nsteps}, {self.N}, 3) {self.dtype} tensor")
        seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        check(self.lib.mx_llg_run(self.h, ctypes.c_void_p(e.data_ptr()), int(nsteps), seed, int(step0),
                                  None if noise is None else ctypes.c_void_p(noise.data_ptr()), _stream()))
        return e

    def run_driven(self, e, B, seed, step0=0):
        """len(B) steps on e with the uniform field B[n, 0] in the predictor and B[n, 1] in the corrector
        of step n (B: (nsteps, 2, 3) host array, in place of the model's field)."""
        self._check_spins(e)
        B = np.ascontiguousarray(np.asarray(B, float).reshape(-1, 2, 3))
        seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        check(self.lib.mx_llg_run_driven(self.h, ctypes.c_void_p(e.data_ptr()), len(B), hptr(B), seed, int(step0),
                                         _stream()))
        return e

    def stage(self, e, stage, B, extra, seed, step, ep_out=None):
        """One stage (1: predictor, 2: corrector) of a step on e with the uniform field B ((3,) or None for
        the model's) and the per-site field extra (device (N, 3) tensor of the model dtype, or None);
        stage 1 writes the predictor spins into ep_out when given."""
        self._check_spins(e)
        for t in (extra, ep_out):
            if t is not None:
                self._check_spins(t)
        Bh = None if B is None else np.ascontiguousarray(np.asarray(B, float).reshape(3))
        check(self.lib.mx_llg_stage(self.h, ctypes.c_void_p(e.data_ptr()), int(stage), hptr(Bh),
                                    None if extra is None else ctypes.c_void_p(extra.data_ptr()),
                                    int(seed) & 0xFFFFFFFFFFFFFFFF, int(step),
                                    None if ep_out is None else ctypes.c_void_p(ep_out.data_ptr()), _stream()))
        return e

    def set_sot(self, p, B_DL, B_FL):
        """Spin-orbit torque fields B_DL (e x p) + B_FL p on the dynamics (p a unit vector)."""
        check(self.lib.mx_model_set_sot(self.h, hptr(np.ascontiguousarray(np.asarray(p, float).reshape(3))),
                                        float(B_DL), float(B_FL)))

    def field(self, e, kind=_lib.MX_FIELD_BEFF):
        self._check_spins(e)
        out = _torch().empty_like(e)
        check(self.lib.mx_field(self.h, ctypes.c_void_p(e.data_ptr()), ctypes.c_void_p(out.data_ptr()), int(kind),
                                _stream()))
        return out

    def observables_device(self, e):
        """[E, sum e (3), sum s e (3), max torque] as a new device float64 tensor (no synchronisation)."""
        self._check_spins(e)
        out = _torch().empty(_lib.MX_OBS_LEN, dtype=_torch().float64, device="cuda")
        check(self.lib.mx_observables(self.h, ctypes.c_void_p(e.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                      _stream()))
        return out

    def observables(self, e):
        return self.observables_device(e).cpu().numpy()

    def noise(self, seed, step):
        out = _torch().empty((self.N, 3), dtype=self.dtype, device="cuda")
        check(self.lib.mx_noise(self.h, ctypes.c_void_p(out.data_ptr()), int(seed) & 0xFFFFFFFFFFFFFFFF, int(step),
                                _stream()))
        return out

    def rhs(self, e, B, alpha, precession=True):
        self._check_spins(e)
        self._check_spins(B)
        out = _torch().empty_like(e)
        check(self.lib.mx_rhs(self.h, ctypes.c_void_p(e.data_ptr()), ctypes.c_void_p(B.data_ptr()), float(alpha),
                              int(bool(precession)), ctypes.c_void_p(out.data_ptr()), _stream()))
        return out

    def topological_charge(self, e, tri, hand=1.0):
        """tri: device int32 (ntri, 3) counter-clockwise triangles."""
        self._check_spins(e)
        out = _torch().empty(1, dtype=_torch().float64, device="cuda")
        check(self.lib.mx_topo_charge(self.h, ctypes.c_void_p(e.data_ptr()), ctypes.c_void_p(tri.data_ptr()),
                                      int(tri.shape[0]), float(hand), ctypes.c_void_p(out.data_ptr()), _stream()))
        return float(out.cpu()[0])
