"""Loader of libmxllg.so (the HIP kernels behind include/mxllg.h) through ctypes.

There is no CPU fallback: if the library is missing, or no HIP device is visible, every entry point
that needs the GPU raises RuntimeError.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmxllg.so")

MX_OK, MX_EINVAL, MX_ENOMEM, MX_EHIP, MX_EUNSUPPORTED = 0, 1, 2, 3, 4
MX_FP64, MX_FP32 = 0, 1
MX_FIELD_BEFF, MX_FIELD_GRADIENT, MX_FIELD_EXCHANGE = 0, 1, 2
MX_HEUN, MX_DEPONDT = 0, 1
MX_OBS_LEN = 8
MX_NOISE_FLOAT, MX_NOISE_DOUBLE = 0, 1

_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64
_D = ctypes.c_double

# name -> (restype, argtypes); every symbol declared in include/mxllg.h
SIGNATURES = {
    "mx_last_error": (ctypes.c_char_p, []),
    "mx_version": (_I, []),
    "mx_model_create": (_I, [_I, _I, _P, _P, _P]),
    "mx_model_destroy": (None, [_P]),
    "mx_model_info": (_I, [_P, _P, _P, _P]),
    "mx_model_set_layout": (_I, [_P, _I]),
    "mx_model_layout": (_I, [_P, _P]),
    "mx_model_set_stencil": (_I, [_P, ctypes.c_int32, _P, ctypes.c_int32, ctypes.c_int32, _P]),
    "mx_model_set_bond_tensors": (_I, [_P, _P]),
    "mx_model_set_onsite": (_I, [_P, _I, _P]),
    "mx_model_set_cubic": (_I, [_P, _P, _P]),
    "mx_model_set_uniaxial_higher": (_I, [_P, _P, _P, _P]),
    "mx_model_set_moments": (_I, [_P, _I, _P, _P]),
    "mx_model_set_pinned": (_I, [_P, _P]),
    "mx_model_set_field": (_I, [_P, _P]),
    "mx_model_set_dynamics": (_I, [_P, _D, _I, _P, _I, _P, _I, _I]),
    "mx_model_set_graph_steps": (_I, [_P, ctypes.c_int32]),
    "mx_model_set_step_mode": (_I, [_P, ctypes.c_int32]),
    "mx_model_set_wait_limit": (_I, [_P, ctypes.c_int32]),
    "mx_model_set_noise": (_I, [_P, ctypes.c_int32]),
    "mx_model_step_info": (_I, [_P, _P, _P]),
    "mx_llg_run_driven": (_I, [_P, _P, _I64, _P, _U64, _U64, _P]),
    "mx_llg_stage": (_I, [_P, _P, ctypes.c_int32, _P, _P, _U64, _U64, _P, _P]),
    "mx_model_set_sot": (_I, [_P, _P, _D, _D]),
    "mx_mc_set_colors": (_I, [_P, ctypes.c_int32, _P, _P]),
    "mx_mc_set_temperature": (_I, [_P, _I, _P]),
    "mx_mc_sweep": (_I, [_P, _P, _I64, _I, _I, _D, _U64, _U64, _P]),
    "mx_mc_accepted": (_I, [_P, _P, _I]),
    "mx_mc_sweep_replicas": (_I, [_P, _P, ctypes.c_int32, _I64, _I, _I, _P, _P, _P, _U64, _P]),
    "mx_mc_replica_accepted": (_I, [_P, ctypes.c_int32, _P, _I]),
    "mx_field_images": (_I, [_P, _P, ctypes.c_int32, _P, _I, _P]),
    "mx_observables_images": (_I, [_P, _P, ctypes.c_int32, _P, _P]),
    "mx_sampler_create": (_I, [_P, ctypes.c_int32, _P, _P, _P, _P]),
    "mx_sampler_destroy": (None, [_P]),
    "mx_sample_mq": (_I, [_P, _P, _P, _P]),
    "mx_llg_run": (_I, [_P, _P, _I64, _U64, _U64, _P, _P]),
    "mx_llg_run_noalloc": (_I, [_P, _P, _I64, _U64, _U64, _P, _P]),
    "mx_noise": (_I, [_P, _P, _U64, _U64, _P]),
    "mx_field": (_I, [_P, _P, _P, _I, _P]),
    "mx_observables": (_I, [_P, _P, _P, _P]),
    "mx_rhs": (_I, [_P, _P, _P, _D, _I, _P, _P]),
    "mx_topo_charge": (_I, [_P, _P, _P, _I64, _D, _P, _P]),
    "mx_batch_create": (_I, [_P, ctypes.c_int32, _P, _P]),
    "mx_batch_destroy": (None, [_P]),
    "mx_batch_offsets": (_I, [_P, _P]),
    "mx_batch_set_dynamics": (_I, [_P, _D, _P, _P, _I]),
    "mx_batch_set_launch": (_I, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "mx_batch_run": (_I, [_P, _P, _I64, _P, _U64, _P]),
    "mx_batch_observables": (_I, [_P, _P, _P, _P]),
    "mx_adjoint_create": (_I, [_P, ctypes.c_int32, _P, _P, ctypes.c_int32, _P]),
    "mx_adjoint_destroy": (None, [_P]),
    "mx_adjoint_set_params": (_I, [_P, _P]),
    "mx_adjoint_set_offsets": (_I, [_P, _P, _P]),
    "mx_adjoint_forward": (_I, [_P, _P, ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint64, _P]),
    "mx_adjoint_backward": (_I, [_P, _P, _P, _P, _P]),
    "mx_adjoint_forward_driven": (_I, [_P, _P, ctypes.c_int64, _P, ctypes.c_uint64, ctypes.c_uint64, _P]),
    "mx_adjoint_backward_field": (_I, [_P, _P, _P, _P, _P, _P]),
    "mx_adjoint_backward_dense": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "mx_adjoint_num_bonds": (_I, [_P]),
}

_lib = None


def lib():
    """The loaded library (raises RuntimeError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m magex_amd.csrc.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    """Map an MX_* status to a Python exception carrying mx_last_error()."""
    if rc == MX_OK:
        return
    msg = lib().mx_last_error().decode(errors="replace")
    if rc == MX_EINVAL:
        raise ValueError(msg)
    if rc == MX_EUNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == MX_ENOMEM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def hptr(a):
    """Host numpy array -> void* (None passes NULL)."""
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)
