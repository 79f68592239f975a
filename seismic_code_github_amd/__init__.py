"""seismic_code_github_amd — MI355X (gfx950) finite-difference wave propagators.

Python mirror of the C ABI in ``include/seisfd.h`` (ctypes over the in-tree
``libseisfd.so``).  The propagators themselves are HIP kernels; this module only
marshals arguments.  It raises if the native library is missing: there is no
CPU fallback.

Reference: stone0312/Seismic_code_github forward/Acoustic_FD_by_ss/src/
Acoustic_1order_FD.cpp — its constant block (:28-45, 137-216) becomes the
arguments of :func:`acoustic_forward`, its time loop (:237-393) the library.
"""
from __future__ import annotations

import ctypes
import os
import weakref
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libseisfd.so")

SFD_OK = 0
SFD_EINVAL = -1
SFD_ENOMEM = -2
SFD_EDEVICE = -3
SFD_ECFL = -4
SFD_ENAN = -5
SFD_ENOTSUP = -6

BOUNDARY_SPONGE = 0
BOUNDARY_CPML = 1
ARITH_FAST = 0
ARITH_REFERENCE = 1
FIELD_P, FIELD_VX, FIELD_VZ = 1, 2, 4
FIELD_SXX, FIELD_SZZ, FIELD_SXZ = 8, 16, 32
ELASTIC_FREE_SURFACE, ELASTIC_ABSORBING_TOP = 0, 1

_STATUS = {SFD_EINVAL: "EINVAL", SFD_ENOMEM: "ENOMEM", SFD_EDEVICE: "EDEVICE",
           SFD_ECFL: "ECFL", SFD_ENAN: "ENAN", SFD_ENOTSUP: "ENOTSUP"}


class SeisfdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_STATUS.get(code, code)}: {msg}")
        self.code = code


_IP = ctypes.POINTER(ctypes.c_int)
_FP = ctypes.POINTER(ctypes.c_float)


class AcousticParamsC(ctypes.Structure):
    """Mirror of ``struct sfd_acoustic_params``."""
    _fields_ = [
        ("nx", ctypes.c_int), ("nz", ctypes.c_int), ("nt", ctypes.c_int),
        ("dx", ctypes.c_float), ("dz", ctypes.c_float), ("dt", ctypes.c_float),
        ("radius", ctypes.c_int), ("coef", _FP),
        ("boundary", ctypes.c_int), ("width", ctypes.c_int),
        ("pml_r", ctypes.c_float), ("pml_f0", ctypes.c_float),
        ("sx", ctypes.c_int), ("sz", ctypes.c_int),
        ("nrec", ctypes.c_int), ("rx", _IP), ("rz", _IP),
        ("nsnap", ctypes.c_int), ("snap_steps", _IP), ("snap_fields", ctypes.c_int),
        ("arith", ctypes.c_int),
    ]


class ViscoParamsC(ctypes.Structure):
    """Mirror of ``struct sfd_visco_params``."""
    _fields_ = [
        ("nx", ctypes.c_int), ("nz", ctypes.c_int), ("nt", ctypes.c_int), ("delay", ctypes.c_int),
        ("dx", ctypes.c_float), ("dz", ctypes.c_float), ("dt", ctypes.c_float), ("f0", ctypes.c_float),
        ("radius", ctypes.c_int), ("coef", _FP), ("nmech", ctypes.c_int), ("q_scale", ctypes.c_float),
        ("f_low", ctypes.c_float), ("f_high", ctypes.c_float),
        ("sponge_xlo", ctypes.c_int), ("sponge_xhi", ctypes.c_int),
        ("sponge_zlo", ctypes.c_int), ("sponge_zhi", ctypes.c_int), ("sponge_r", ctypes.c_float),
        ("sx", ctypes.c_int), ("sz", ctypes.c_int), ("src_sigma", ctypes.c_float),
        ("nrec", ctypes.c_int), ("rx", _IP), ("rz", _IP),
        ("nsnap", ctypes.c_int), ("snap_steps", _IP), ("snap_fields", ctypes.c_int),
        ("qs", _FP),
    ]


class ElasticParamsC(ctypes.Structure):
    """Mirror of ``struct sfd_elastic_params``."""
    _fields_ = [
        ("nx", ctypes.c_int), ("nz", ctypes.c_int), ("nt", ctypes.c_int), ("delay", ctypes.c_int),
        ("dx", ctypes.c_float), ("dz", ctypes.c_float), ("dt", ctypes.c_float), ("f0", ctypes.c_float),
        ("itopo", _IP),
        ("sponge_xlo", ctypes.c_int), ("sponge_xhi", ctypes.c_int), ("sponge_zhi", ctypes.c_int),
        ("sponge_r", ctypes.c_float),
        ("sx", ctypes.c_int), ("sz", ctypes.c_int), ("src_radius", ctypes.c_int),
        ("nrec", ctypes.c_int), ("rx", _IP), ("rz", _IP),
        ("nsnap", ctypes.c_int), ("snap_steps", _IP), ("snap_fields", ctypes.c_int),
        ("mode", ctypes.c_int), ("sponge_zlo", ctypes.c_int),
    ]


class VtiParamsC(ctypes.Structure):
    """Mirror of ``struct sfd_vti_params``."""
    _fields_ = [
        ("nx", ctypes.c_int), ("nz", ctypes.c_int), ("nt", ctypes.c_int), ("delay", ctypes.c_int),
        ("dx", ctypes.c_float), ("dz", ctypes.c_float), ("dt", ctypes.c_float), ("f0", ctypes.c_float),
        ("pml_xlo", ctypes.c_int), ("pml_xhi", ctypes.c_int), ("pml_zlo", ctypes.c_int), ("pml_zhi", ctypes.c_int),
        ("pml_r", ctypes.c_float), ("c2", _FP), ("c1", _FP),
        ("sx", ctypes.c_int), ("sz", ctypes.c_int),
        ("nrec", ctypes.c_int), ("rx", _IP), ("rz", _IP),
        ("allow_eps_lt_delta", ctypes.c_int),
    ]


class SrmParamsC(ctypes.Structure):
    """Mirror of ``struct sfd_srm_params``."""
    _fields_ = [("nshot", ctypes.c_int), ("ntrace", ctypes.c_int), ("lt", ctypes.c_int), ("ncrg", ctypes.c_int)]


class QParC(ctypes.Structure):
    _fields_ = [(n, ctypes.c_char * 256) for n in ("vp_fn", "vs_fn", "rho_fn", "qfp_fn", "qfs_fn",
                                                   "shot_fn1", "shot_fn2", "shot_fn3", "shot_fn4")] + [
        ("nshot", ctypes.c_int),
        *[(n, ctypes.c_float) for n in ("offset_min", "offset_max", "sx0", "sz0", "rz0", "dsx")],
        ("nx", ctypes.c_int), ("nz", ctypes.c_int),
        ("dx", ctypes.c_float), ("dz", ctypes.c_float),
        ("nt", ctypes.c_int),
        *[(n, ctypes.c_float) for n in ("dt", "f0", "ppxn", "ppxp", "ppzn", "ppzp")],
        ("fdo2", ctypes.c_int), ("fdo1", ctypes.c_int)]


class IsoParC(ctypes.Structure):
    _fields_ = [(n, ctypes.c_char * 256) for n in ("mode", "vp_fn", "vs_fn", "rho_fn", "topo_fn", "snapx_fn",
                                                   "snapz_fn", "shot_fn1", "shot_fn2", "shot_fn3",
                                                   "shot_fn4")] + [
        ("nshot", ctypes.c_int),
        *[(n, ctypes.c_float) for n in ("ofsmin", "ofsmax", "sx0", "sz0", "rz0", "dsx")],
        ("nx", ctypes.c_int), ("nz", ctypes.c_int),
        ("dx", ctypes.c_float), ("dz", ctypes.c_float),
        ("nt", ctypes.c_int),
        *[(n, ctypes.c_float) for n in ("dt", "dtsnap", "f0", "decay", "ppxn", "ppxp", "ppzn", "ppzp")],
        ("rgau", ctypes.c_int)]


# Every symbol declared in include/seisfd.h, with its ctypes signature.
_SIGNATURES = {
    "sfd_abi_version": (ctypes.c_int, []),
    "sfd_context_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "sfd_context_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sfd_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "sfd_device_alloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    "sfd_device_copy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int]),
    "sfd_device_free": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "sfd_acoustic_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(AcousticParamsC),
                                            _FP, _FP, _FP, _FP, _FP]),
    "sfd_acoustic_plan_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(AcousticParamsC),
                                                _FP, _FP, _FP, ctypes.POINTER(ctypes.c_void_p)]),
    "sfd_acoustic_plan_run": (ctypes.c_int, [ctypes.c_void_p, _FP]),
    "sfd_acoustic_plan_fetch": (ctypes.c_int, [ctypes.c_void_p, _FP, _FP]),
    "sfd_acoustic_plan_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sfd_acoustic_plan_set_state": (ctypes.c_int, [ctypes.c_void_p, _FP, _FP]),
    "sfd_acoustic_plan_get_state": (ctypes.c_int, [ctypes.c_void_p, _FP, _FP]),
    "sfd_acoustic_plan_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "sfd_acoustic_plan_set_tiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "sfd_acoustic_plan_set_tiled": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "sfd_acoustic_plan_set_fused": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "sfd_acoustic_cfl_limit": (ctypes.c_double, [ctypes.c_int, _FP]),
    "sfd_acoustic_batch_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(AcousticParamsC),
                                                 ctypes.POINTER(_FP), ctypes.POINTER(_FP), ctypes.POINTER(_FP),
                                                 ctypes.POINTER(ctypes.c_void_p)]),
    "sfd_acoustic_batch_set_tiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "sfd_acoustic_batch_set_tiled": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "sfd_acoustic_batch_set_fused": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "sfd_acoustic_plan_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _FP]),
    "sfd_acoustic_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _FP]),
    "sfd_acoustic_plan_get_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                                   ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "sfd_acoustic_batch_get_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                                    ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "sfd_read_q_par": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(QParC), ctypes.c_char_p]),
    "sfd_read_iso_par": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(IsoParC), ctypes.c_char_p]),
    "sfd_write_su_shot": (ctypes.c_int, [ctypes.c_char_p, _FP, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_int]),
    "sfd_merge_su": (ctypes.c_int, [ctypes.POINTER(ctypes.c_char_p), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_char_p, ctypes.c_int]),
    "sfd_elastic_plan_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _FP]),
    "sfd_vti_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(VtiParamsC), ctypes.c_int, _FP, _FP, _FP, _FP,
                                       _FP]),
    "sfd_vti_plan_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(VtiParamsC), ctypes.c_int, _FP, _FP, _FP,
                                           _FP, ctypes.POINTER(ctypes.c_void_p)]),
    "sfd_vti_plan_run": (ctypes.c_int, [ctypes.c_void_p, _FP]),
    "sfd_vti_plan_fetch": (ctypes.c_int, [ctypes.c_void_p, _FP, _FP]),
    "sfd_vti_plan_set_tiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "sfd_vti_plan_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _FP]),
    "sfd_vti_plan_get_shape": (ctypes.c_int, [ctypes.c_void_p, _IP, _IP]),
    "sfd_vti_plan_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sfd_vti_coefficients": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _FP]),
    "sfd_srm_predict": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SrmParamsC), _FP, _FP, _FP]),
    "sfd_srm_crg_traces": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "sfd_srm_plan_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SrmParamsC), _FP, _FP,
                                           ctypes.POINTER(ctypes.c_void_p)]),
    "sfd_srm_plan_run": (ctypes.c_int, [ctypes.c_void_p, _FP]),
    "sfd_srm_plan_fetch": (ctypes.c_int, [ctypes.c_void_p, _FP]),
    "sfd_srm_plan_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sfd_acoustic_forward_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(AcousticParamsC),
                                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "sfd_noalloc_scratch_bytes": (ctypes.c_size_t, []),
    "sfd_bandwidth_probe": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_double)]),
    "sfd_srm_plan_set_method": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "sfd_srm_plan_get_method": (ctypes.c_int, [ctypes.c_void_p]),
    "sfd_vti_wavelet": (ctypes.c_int, [ctypes.c_float, ctypes.c_float, ctypes.c_int, _FP]),
    "sfd_vti_window": (ctypes.c_int, [_FP, _FP, _FP] + [ctypes.c_int] * 9 + [_FP, _FP, _FP]),
    "sfd_elastic_plan_get_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                                  ctypes.POINTER(ctypes.c_int)]),
    "sfd_visco_plan_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _FP]),
    "sfd_visco_plan_get_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                                ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "sfd_context_save_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "sfd_context_load_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "sfd_context_tuning_size": (ctypes.c_int, [ctypes.c_void_p]),
    "sfd_acoustic_batch_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _FP]),
    "sfd_acoustic_batch_fetch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _FP]),
    "sfd_acoustic_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sfd_acoustic_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(AcousticParamsC), _FP, _FP, _FP, _FP,
                                             ctypes.c_int, _FP, _FP]),
    "sfd_acoustic_misfit_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(AcousticParamsC),
                                                    _FP, _FP, ctypes.POINTER(_FP), ctypes.POINTER(_FP), ctypes.c_int,
                                                    _FP, ctypes.POINTER(ctypes.c_double)]),
    "sfd_acoustic_misfit_gradient_ex": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                                       ctypes.POINTER(AcousticParamsC), _FP, _FP,
                                                       ctypes.POINTER(_FP), ctypes.POINTER(_FP), ctypes.c_int,
                                                       ctypes.c_int, _FP, ctypes.POINTER(ctypes.c_double)]),
    "sfd_acoustic_misfit_gradient_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                                        ctypes.POINTER(AcousticParamsC), _FP, _FP,
                                                        ctypes.POINTER(_FP), ctypes.POINTER(_FP), ctypes.c_int,
                                                        ctypes.c_int, _FP, ctypes.POINTER(ctypes.c_double)]),
    "sfd_visco_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ViscoParamsC), _FP, _FP, _FP, _FP, _FP,
                                         _FP, _FP]),
    "sfd_visco_plan_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ViscoParamsC), _FP, _FP, _FP, _FP, _FP,
                                             ctypes.POINTER(ctypes.c_void_p)]),
    "sfd_visco_plan_run": (ctypes.c_int, [ctypes.c_void_p, _FP]),
    "sfd_visco_plan_fetch": (ctypes.c_int, [ctypes.c_void_p, _FP, _FP]),
    "sfd_visco_plan_fetch_xx": (ctypes.c_int, [ctypes.c_void_p, _FP]),
    "sfd_visco_plan_set_tiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "sfd_visco_plan_set_tiled": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "sfd_visco_plan_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sfd_acoustic_batch_set_group_bytes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t]),
    "sfd_acoustic_batch_get_groups": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "sfd_vti_plan_set_group_bytes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t]),
    "sfd_vti_plan_get_groups": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "sfd_visco_batch_set_group_bytes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t]),
    "sfd_visco_batch_get_groups": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "sfd_elastic_batch_set_group_bytes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t]),
    "sfd_elastic_batch_get_groups": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "sfd_visco_batch_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ViscoParamsC),
                                              ctypes.POINTER(_FP), ctypes.POINTER(_FP), ctypes.POINTER(_FP),
                                              ctypes.POINTER(_FP), ctypes.POINTER(_FP),
                                              ctypes.POINTER(ctypes.c_void_p)]),
    "sfd_visco_batch_run": (ctypes.c_int, [ctypes.c_void_p, _FP]),
    "sfd_visco_batch_fetch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _FP, _FP, _FP]),
    "sfd_visco_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sfd_elastic_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ElasticParamsC), _FP, _FP, _FP, _FP,
                                           _FP, _FP, _FP]),
    "sfd_elastic_plan_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ElasticParamsC), _FP, _FP, _FP, _FP,
                                               ctypes.POINTER(ctypes.c_void_p)]),
    "sfd_elastic_plan_run": (ctypes.c_int, [ctypes.c_void_p, _FP]),
    "sfd_elastic_plan_fetch": (ctypes.c_int, [ctypes.c_void_p, _FP, _FP, _FP]),
    "sfd_elastic_plan_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sfd_elastic_plan_set_tiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "sfd_elastic_batch_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ElasticParamsC),
                                                ctypes.POINTER(_FP), ctypes.POINTER(_FP), ctypes.POINTER(_FP),
                                                ctypes.POINTER(_FP), ctypes.POINTER(ctypes.c_void_p)]),
    "sfd_elastic_batch_set_tiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "sfd_elastic_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _FP]),
    "sfd_elastic_batch_get_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                                   ctypes.POINTER(ctypes.c_int)]),
    "sfd_elastic_batch_run": (ctypes.c_int, [ctypes.c_void_p, _FP]),
    "sfd_elastic_batch_fetch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _FP, _FP, _FP]),
    "sfd_elastic_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sfd_window_model": (ctypes.c_int, [_FP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, _FP]),
}

BATCH_GRAPH = 0
BATCH_LAUNCHES = 1
BATCH_PER_SHOT = 2

_lib = None


def lib() -> ctypes.CDLL:
    """Load libseisfd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"native library {LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, t=ctypes.c_float):
    if a is None:
        return ctypes.POINTER(t)()
    return a.ctypes.data_as(ctypes.POINTER(t))


class Context:
    """An ``sfd_context``: one HIP device and stream."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        rc = lib().sfd_context_create(device, ctypes.byref(h))
        if rc != SFD_OK:
            raise SeisfdError(rc, f"sfd_context_create(device={device}) failed")
        self.handle = h
        # Plans and batches made in this context: closed before the context
        # (the library requires it).
        self._children = weakref.WeakSet()

    def adopt(self, obj):
        self._children.add(obj)

    def check(self, rc: int, what: str):
        if rc != SFD_OK:
            msg = lib().sfd_last_error(self.handle) or b""
            raise SeisfdError(rc, f"{what}: {msg.decode()}")

    def save_tuning(self, path: str):
        """Writes (merges) the autotuner cache into a JSON file."""
        self.check(lib().sfd_context_save_tuning(self.handle, os.fsencode(path)), "save_tuning")

    def load_tuning(self, path: str):
        """Adds the entries of a JSON tuning file measured on this GPU's architecture."""
        self.check(lib().sfd_context_load_tuning(self.handle, os.fsencode(path)), "load_tuning")

    def tuning_size(self) -> int:
        return lib().sfd_context_tuning_size(self.handle)

    def close(self):
        if self.handle:
            for c in list(self._children):
                c.close()
            lib().sfd_context_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class DeviceArray:
    """A float32 array in the context's device memory (``sfd_device_alloc``),
    for inputs kept resident across calls, such as observed gathers over FWI
    iterations.  Exposes ``__cuda_array_interface__`` (version 3)."""

    def __init__(self, a, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        a = _f32(a)
        self.shape = tuple(a.shape)
        self.nbytes = a.nbytes
        p = ctypes.c_void_p()
        self.ctx.check(lib().sfd_device_alloc(self.ctx.handle, self.nbytes, ctypes.byref(p)), "sfd_device_alloc")
        self.ptr = p.value or 0
        if self.nbytes:
            self.ctx.check(lib().sfd_device_copy(self.ctx.handle, self.ptr, a.ctypes.data, self.nbytes, 0),
                           "sfd_device_copy")

    @property
    def __cuda_array_interface__(self):
        return {"shape": self.shape, "typestr": "<f4", "data": (self.ptr, False), "version": 3, "strides": None}

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float32)
        if self.nbytes:
            self.ctx.check(lib().sfd_device_copy(self.ctx.handle, out.ctypes.data, self.ptr, self.nbytes, 1),
                           "sfd_device_copy")
        return out

    def close(self):
        if self.ptr and self.ctx.handle:
            lib().sfd_device_free(self.ctx.handle, self.ptr)
        self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class AcousticShot:
    """One acoustic shot: the reference program's constant block as data.

    vp, den: [nx, nz] float32, z fastest.  wavelet: [nt].  src = (sx, sz).
    rec = (rx[], rz[]).  width: sponge/PML width (reference PML)."""
    vp: np.ndarray
    wavelet: np.ndarray
    src: tuple
    rec: tuple
    dx: float = 10.0
    dz: float = 10.0
    dt: float = 5e-4
    den: Optional[np.ndarray] = None
    radius: int = 8
    coef: Optional[np.ndarray] = None
    boundary: str = "sponge"
    width: int = 20
    pml_r: float = 1e-3
    pml_f0: float = 0.0
    snap_steps: Sequence[int] = ()
    snap_fields: int = FIELD_P | FIELD_VX | FIELD_VZ
    arith: str = "fast"
    nt: Optional[int] = None
    _keep: list = field(default_factory=list, repr=False)

    def params(self, keep: Optional[list] = None) -> AcousticParamsC:
        """The C struct.  Its receiver, snapshot and coefficient arrays live as
        long as the list they are appended to: ``keep`` (owned by the caller,
        which must hold it until the library call returns; pass one list for
        every struct made for the same call) or, without one, the shot's own
        list, replaced by the next call."""
        vp = _f32(self.vp)
        nx, nz = vp.shape
        nt = int(self.nt if self.nt is not None else np.asarray(self.wavelet).shape[0])
        rx = np.ascontiguousarray(self.rec[0], dtype=np.int32)
        rz = np.ascontiguousarray(self.rec[1], dtype=np.int32)
        steps = np.ascontiguousarray(list(self.snap_steps), dtype=np.int32)
        coef = None if self.coef is None else _f32(self.coef)
        if keep is None:
            self._keep = keep = []
        keep += [rx, rz, steps, coef]
        return AcousticParamsC(
            nx, nz, nt, self.dx, self.dz, self.dt, self.radius, _ptr(coef),
            {"sponge": BOUNDARY_SPONGE, "cpml": BOUNDARY_CPML}[self.boundary], self.width,
            self.pml_r, self.pml_f0, int(self.src[0]), int(self.src[1]),
            rx.size, _ptr(rx, ctypes.c_int), _ptr(rz, ctypes.c_int),
            steps.size, _ptr(steps, ctypes.c_int), self.snap_fields,
            {"fast": ARITH_FAST, "reference": ARITH_REFERENCE}[self.arith])

    def nfields(self) -> int:
        return bin(self.snap_fields & 7).count("1")


class AcousticPlan:
    """``sfd_acoustic_plan``: model and wavelet resident on the device."""

    def __init__(self, shot: AcousticShot, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.ctx.adopt(self)
        self.shot = shot
        self._pkeep = []
        self.p = shot.params(self._pkeep)
        self.nx, self.nz = np.asarray(shot.vp).shape
        self.nt = self.p.nt
        vp = _f32(shot.vp)
        den = None if shot.den is None else _f32(shot.den)
        w = _f32(shot.wavelet)
        h = ctypes.c_void_p()
        self.ctx.check(lib().sfd_acoustic_plan_create(self.ctx.handle, ctypes.byref(self.p), _ptr(vp),
                                                      _ptr(den), _ptr(w), ctypes.byref(h)),
                       "sfd_acoustic_plan_create")
        self.handle = h

    def set_tiling(self, waves: int = 0, chunk_rows: int = 0):
        self.ctx.check(lib().sfd_acoustic_plan_set_tiling(self.handle, waves, chunk_rows), "set_tiling")

    def set_tiled(self, tile_rows: int = 0, waves: int = 0):
        """Fast arithmetic: the two-launch tiled step at tile_rows x waves
        (0 = the default shape; tile_rows < 0 = the fused march kernel)."""
        self.ctx.check(lib().sfd_acoustic_plan_set_tiled(self.handle, tile_rows, waves), "set_tiled")

    def set_fused(self, rows: int = 0, waves: int = 0):
        """One launch per step: `rows` x 64 cells per workgroup of `waves` (0 = 32 x 8); rows 16 or
        32 select the tiled step, rows >= 64 the streaming step (a chunk marched in row blocks)."""
        self.ctx.check(lib().sfd_acoustic_plan_set_fused(self.handle, rows, waves), "set_fused")

    def autotune(self, trial_steps: int = 20) -> float:
        """Times every launch shape, keeps the fastest (cached in the context);
        returns its device time per step in us.  Resets the state."""
        us = ctypes.c_float()
        self.ctx.check(lib().sfd_acoustic_plan_autotune(self.handle, trial_steps, ctypes.byref(us)), "autotune")
        return us.value

    def shape(self):
        """(kind, a, b): kind 0 = tiled (rows, waves), 1 = march (waves, chunk rows)."""
        k, a, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self.ctx.check(lib().sfd_acoustic_plan_get_shape(self.handle, ctypes.byref(k), ctypes.byref(a),
                                                         ctypes.byref(b)), "get_shape")
        return k.value, a.value, b.value

    def run(self) -> float:
        ms = ctypes.c_float()
        self.ctx.check(lib().sfd_acoustic_plan_run(self.handle, ctypes.byref(ms)), "sfd_acoustic_plan_run")
        return ms.value

    def step(self, k0: int, nsteps: int):
        self.ctx.check(lib().sfd_acoustic_plan_step(self.handle, k0, nsteps), "sfd_acoustic_plan_step")

    def fetch(self):
        nrec = self.p.nrec
        record = np.zeros((nrec, self.nt), dtype=np.float32)
        snaps = np.zeros((self.p.nsnap, self.shot.nfields(), self.nx, self.nz), dtype=np.float32)
        self.ctx.check(lib().sfd_acoustic_plan_fetch(self.handle, _ptr(record) if nrec else None,
                                                     _ptr(snaps) if self.p.nsnap else None),
                       "sfd_acoustic_plan_fetch")
        return record, snaps

    def get_state(self):
        st = np.zeros((3, self.nx, self.nz), dtype=np.float32)
        pml = np.zeros((4, self.nx, self.nz), dtype=np.float32) if self.p.boundary == BOUNDARY_CPML else None
        self.ctx.check(lib().sfd_acoustic_plan_get_state(self.handle, _ptr(st), _ptr(pml)), "get_state")
        return st, pml

    def set_state(self, state, pml=None):
        st = _f32(state)
        pm = None if pml is None else _f32(pml)
        self.ctx.check(lib().sfd_acoustic_plan_set_state(self.handle, _ptr(st), _ptr(pm)), "set_state")

    def close(self):
        if getattr(self, "handle", None):
            lib().sfd_acoustic_plan_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def acoustic_forward(shot: AcousticShot, ctx: Optional[Context] = None):
    """One shot end to end through ``sfd_acoustic_forward``.
    Returns (record [nrec, nt], snaps [nsnap, nfields, nx, nz])."""
    ctx = ctx or default_context()
    pkeep = []
    p = shot.params(pkeep)
    nx, nz = np.asarray(shot.vp).shape
    vp = _f32(shot.vp)
    den = None if shot.den is None else _f32(shot.den)
    w = _f32(shot.wavelet)
    record = np.zeros((p.nrec, p.nt), dtype=np.float32)
    snaps = np.zeros((p.nsnap, shot.nfields(), nx, nz), dtype=np.float32)
    ctx.check(lib().sfd_acoustic_forward(ctx.handle, ctypes.byref(p), _ptr(vp), _ptr(den), _ptr(w),
                                         _ptr(record) if p.nrec else None, _ptr(snaps) if p.nsnap else None),
              "sfd_acoustic_forward")
    return record, snaps


def acoustic_gradient(shot: AcousticShot, seed, ckpt_every: int = 50, ctx: Optional[Context] = None):
    """dJ/dvp for J = sum(seed * record) through ``sfd_acoustic_gradient``
    (discrete adjoint of the fast sponge or C-PML step, checkpoints every
    ``ckpt_every`` steps; C-PML profiles held fixed).  Returns (grad [nx, nz],
    record [nrec, nt])."""
    ctx = ctx or default_context()
    pkeep = []
    p = shot.params(pkeep)
    nx, nz = np.asarray(shot.vp).shape
    vp = _f32(shot.vp)
    den = None if shot.den is None else _f32(shot.den)
    w = _f32(shot.wavelet)
    sd = _f32(seed)
    if sd.shape != (p.nrec, p.nt):
        raise ValueError(f"seed must be [nrec, nt] = {(p.nrec, p.nt)}, got {sd.shape}")
    grad = np.zeros((nx, nz), dtype=np.float32)
    record = np.zeros((p.nrec, p.nt), dtype=np.float32)
    ctx.check(lib().sfd_acoustic_gradient(ctx.handle, ctypes.byref(p), _ptr(vp), _ptr(den), _ptr(w), _ptr(sd),
                                          ckpt_every, _ptr(grad), _ptr(record)), "sfd_acoustic_gradient")
    return grad, record


GRADIENT_METHODS = {"recompute": 0, "boundary": 1}


def misfit_gradient(shots, d_obs, vp=None, den=None, ckpt_every: int = 50, ctx: Optional[Context] = None,
                    method: str = "recompute"):
    """J = sum_s 1/2 |record_s - d_obs[s]|^2 over shots sharing one model and
    its dJ/dvp, through ``sfd_acoustic_misfit_gradient_ex`` (d_obs as host
    arrays) or ``sfd_acoustic_misfit_gradient_dev`` (d_obs as contiguous
    float32 device arrays on the context's GPU -- ``DeviceArray`` or any
    object with ``__cuda_array_interface__`` -- read in place).  The
    model is ``vp``/``den`` (default: the first shot's).  method "recompute"
    (checkpoints + segment recomputation) or "boundary" (checkpoints +
    boundary saving; the gradient inside the box the sponge or C-PML leaves
    alone only).  ``ckpt_every >= nt`` with "recompute" keeps every step's
    divergence on the device (nt fields per shot) and recomputes nothing:
    faster where memory allows, the same result bit for bit.  Every shot
    needs the same boundary kind.
    Returns (J, grad [nx, nz])."""
    ctx = ctx or default_context()
    shots = list(shots)
    n = len(shots)
    vp = _f32(shots[0].vp if vp is None else vp)
    den = shots[0].den if den is None else den
    den = None if den is None else _f32(den)
    params = (AcousticParamsC * n)()
    ws, ds = (_FP * n)(), (_FP * n)()
    keep = []
    on_device = n > 0 and all(hasattr(d, "__cuda_array_interface__") for d in d_obs)
    for i, sh in enumerate(shots):
        params[i] = sh.params(keep)
        w = _f32(sh.wavelet)
        d = d_obs[i]
        if on_device:
            iface = d.__cuda_array_interface__
            shape = tuple(iface["shape"])
            if iface["typestr"] != "<f4" or iface.get("strides") not in (None, (4 * shape[-1], 4) if shape else None):
                raise ValueError(f"d_obs[{i}] must be a contiguous float32 device array")
            ds[i] = ctypes.cast(iface["data"][0], _FP)
        else:
            d = _f32(d)
            shape = d.shape
            ds[i] = _ptr(d)
        if shape != (params[i].nrec, params[i].nt):
            raise ValueError(f"d_obs[{i}] must be [nrec, nt]")
        keep += [w, d, sh]
        ws[i] = _ptr(w)
    grad = np.zeros(vp.shape, dtype=np.float32)
    J = ctypes.c_double()
    if on_device:
        fn, name = lib().sfd_acoustic_misfit_gradient_dev, "sfd_acoustic_misfit_gradient_dev"
    else:
        fn, name = lib().sfd_acoustic_misfit_gradient_ex, "sfd_acoustic_misfit_gradient_ex"
    ctx.check(fn(ctx.handle, n, params, _ptr(vp), _ptr(den), ws, ds, ckpt_every, GRADIENT_METHODS[method],
                 _ptr(grad), ctypes.byref(J)), name)
    return J.value, grad


def cfl_limit(radius: int, coef=None) -> float:
    c = None if coef is None else _f32(coef)
    return lib().sfd_acoustic_cfl_limit(radius, _ptr(c))


def window_model(model, x0: int, nxe: int, pad_x: tuple, pad_z: tuple) -> np.ndarray:
    """Aperture window of one shot (``sfd_window_model``): columns
    clamp(x0 + w) of ``model`` [mnx, mnz] for w < nxe, edge-padded by
    pad_x = (lo, hi) columns and pad_z = (lo, hi) rows."""
    m = _f32(model)
    mnx, mnz = m.shape
    out = np.empty((nxe + pad_x[0] + pad_x[1], mnz + pad_z[0] + pad_z[1]), dtype=np.float32)
    rc = lib().sfd_window_model(_ptr(m), mnx, mnz, x0, nxe, pad_x[0], pad_x[1], pad_z[0], pad_z[1], _ptr(out))
    if rc != SFD_OK:
        raise SeisfdError(rc, "sfd_window_model")
    return out


class AcousticBatch:
    """``sfd_acoustic_batch``: many shots advanced by one launch per step."""

    def __init__(self, shots, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.ctx.adopt(self)
        self.shots = list(shots)
        n = len(self.shots)
        params = (AcousticParamsC * n)()
        self._keep = []
        vps, dens, ws = (_FP * n)(), (_FP * n)(), (_FP * n)()
        for i, sh in enumerate(self.shots):
            params[i] = sh.params(self._keep)
            vp = _f32(sh.vp)
            w = _f32(sh.wavelet)
            den = None if sh.den is None else _f32(sh.den)
            self._keep += [vp, w, den, sh]
            vps[i], ws[i], dens[i] = _ptr(vp), _ptr(w), _ptr(den)
        h = ctypes.c_void_p()
        self.ctx.check(lib().sfd_acoustic_batch_create(self.ctx.handle, n, params, vps, dens, ws, ctypes.byref(h)),
                       "sfd_acoustic_batch_create")
        self.handle = h
        self.nt = params[0].nt
        self.nrec = [params[i].nrec for i in range(n)]

    def set_tiling(self, waves: int = 0, chunk_rows: int = 0):
        self.ctx.check(lib().sfd_acoustic_batch_set_tiling(self.handle, waves, chunk_rows), "batch_set_tiling")

    def set_tiled(self, tile_rows: int = 0, waves: int = 0):
        self.ctx.check(lib().sfd_acoustic_batch_set_tiled(self.handle, tile_rows, waves), "batch_set_tiled")

    def set_group_bytes(self, nbytes: int = 0):
        """Budget of one shot group's resident fields (0 = the default,
        -1 = every shot in one group); returns the number of groups."""
        v = (1 << 64) - 1 if nbytes < 0 else int(nbytes)
        self.ctx.check(lib().sfd_acoustic_batch_set_group_bytes(self.handle, v), "batch set_group_bytes")
        return self.groups()

    def groups(self) -> int:
        """Number of shot groups the batch runs one after another."""
        n = ctypes.c_int()
        self.ctx.check(lib().sfd_acoustic_batch_get_groups(self.handle, ctypes.byref(n)), "batch get_groups")
        return n.value

    def set_fused(self, rows: int = 0, waves: int = 0):
        """One launch per step for every shot (0 = 32-row tiles, 8 waves)."""
        self.ctx.check(lib().sfd_acoustic_batch_set_fused(self.handle, rows, waves), "batch_set_fused")

    def autotune(self, trial_steps: int = 20) -> float:
        """Times every launch shape, keeps the fastest (cached in the context);
        returns its device time per step in us."""
        us = ctypes.c_float()
        self.ctx.check(lib().sfd_acoustic_batch_autotune(self.handle, trial_steps, ctypes.byref(us)),
                       "batch_autotune")
        return us.value

    def shape(self):
        """(kind, a, b): kind 0 = tiled (rows, waves), 1 = march (waves, chunk rows)."""
        k, a, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self.ctx.check(lib().sfd_acoustic_batch_get_shape(self.handle, ctypes.byref(k), ctypes.byref(a),
                                                          ctypes.byref(b)), "batch_get_shape")
        return k.value, a.value, b.value

    def run(self, how: int = BATCH_GRAPH) -> float:
        ms = ctypes.c_float()
        self.ctx.check(lib().sfd_acoustic_batch_run(self.handle, how, ctypes.byref(ms)), "sfd_acoustic_batch_run")
        return ms.value

    def fetch(self, shot: int) -> np.ndarray:
        rec = np.zeros((self.nrec[shot], self.nt), dtype=np.float32)
        self.ctx.check(lib().sfd_acoustic_batch_fetch(self.handle, shot, _ptr(rec)), "sfd_acoustic_batch_fetch")
        return rec

    def close(self):
        if getattr(self, "handle", None):
            lib().sfd_acoustic_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- visco-acoustic

@dataclass
class ViscoShot:
    """One visco-acoustic shot on the reference's padded shot grid
    (forward/Q_forward single_shot_modelling, vs = 0).

    vp, qp, rho: [nx, nz] float32, z fastest (rho None = 2400).  src = (sx, sz);
    rec = (rx[], rz[]).  sponge = (npxn, npxp, npzn, npzp).  nt = recorded
    samples; delay = steps before the first (None = the reference's
    nint(1.2 / (f0 dt))).  wavelet None = the reference's built-in Ricker."""
    vp: np.ndarray
    qp: np.ndarray
    src: tuple
    rec: tuple
    nt: int
    dx: float = 10.0
    dz: float = 10.0
    dt: float = 1e-3
    f0: float = 20.0
    rho: Optional[np.ndarray] = None
    vs: Optional[np.ndarray] = None
    radius: int = 5
    coef: Optional[np.ndarray] = None
    nmech: int = 1
    q_scale: float = 0.9
    f_low: float = 0.0
    f_high: float = 0.0
    sponge: tuple = (40, 40, 40, 40)
    sponge_r: float = 1e-3
    src_sigma: float = 1.5
    delay: Optional[int] = None
    wavelet: Optional[np.ndarray] = None
    snap_steps: Sequence[int] = ()
    snap_fields: int = FIELD_P
    qs: Optional[np.ndarray] = None
    _keep: list = field(default_factory=list, repr=False)

    def params(self, keep: Optional[list] = None) -> ViscoParamsC:
        """The C struct; its arrays live in ``keep`` (as AcousticShot.params)."""
        nx, nz = np.asarray(self.vp).shape
        rx = np.ascontiguousarray(self.rec[0], dtype=np.int32)
        rz = np.ascontiguousarray(self.rec[1], dtype=np.int32)
        steps = np.ascontiguousarray(list(self.snap_steps), dtype=np.int32)
        coef = None if self.coef is None else _f32(self.coef)
        qs = None if self.qs is None else _f32(self.qs)
        if keep is None:
            self._keep = keep = []
        keep += [rx, rz, steps, coef, qs]
        return ViscoParamsC(nx, nz, int(self.nt), -1 if self.delay is None else int(self.delay),
                            self.dx, self.dz, self.dt, self.f0, self.radius, _ptr(coef), self.nmech,
                            self.q_scale, self.f_low, self.f_high, *[int(v) for v in self.sponge], self.sponge_r,
                            int(self.src[0]), int(self.src[1]), self.src_sigma, rx.size, _ptr(rx, ctypes.c_int),
                            _ptr(rz, ctypes.c_int), steps.size, _ptr(steps, ctypes.c_int), self.snap_fields, _ptr(qs))

    def elastic(self) -> bool:
        return self.vs is not None and bool(np.any(np.asarray(self.vs) != 0))

    def nfields(self) -> int:
        return bin(self.snap_fields & (63 if self.elastic() else 7)).count("1")

    def ref_delay(self) -> int:
        """The reference's decay = nint(1.2 / (f0 dt)) (main:279-280), in float."""
        if self.delay is not None:
            return int(self.delay)
        v = np.float32(1.2) / (np.float32(self.f0) * np.float32(self.dt))
        return int(np.floor(float(v) + 0.5))

    def _arrays(self):
        vp = _f32(self.vp)
        qp = _f32(self.qp)
        rho = None if self.rho is None else _f32(self.rho)
        vs = None if self.vs is None else _f32(self.vs)
        w = None if self.wavelet is None else _f32(self.wavelet)
        return vp, vs, rho, qp, w


class ViscoPlan:
    """``sfd_visco_plan``: model resident on the device."""

    def __init__(self, shot: ViscoShot, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.ctx.adopt(self)
        self.shot = shot
        self._pkeep = []
        self.p = shot.params(self._pkeep)
        self.nx, self.nz = np.asarray(shot.vp).shape
        self._arr = shot._arrays()
        vp, vs, rho, qp, w = self._arr
        h = ctypes.c_void_p()
        self.ctx.check(lib().sfd_visco_plan_create(self.ctx.handle, ctypes.byref(self.p), _ptr(vp), _ptr(vs),
                                                   _ptr(rho), _ptr(qp), _ptr(w), ctypes.byref(h)),
                       "sfd_visco_plan_create")
        self.handle = h

    def set_tiling(self, waves: int = 0, chunk_rows: int = 0):
        self.ctx.check(lib().sfd_visco_plan_set_tiling(self.handle, waves, chunk_rows), "visco set_tiling")

    def set_tiled(self, tile_rows: int = 0, waves: int = 0):
        """The tiled two-launch step at tile_rows x 64 columns, `waves` waves
        (8 x 8, the default, or 16 x 8)."""
        self.ctx.check(lib().sfd_visco_plan_set_tiled(self.handle, tile_rows, waves), "visco set_tiled")

    def autotune(self, trial_steps: int = 20) -> float:
        us = ctypes.c_float()
        self.ctx.check(lib().sfd_visco_plan_autotune(self.handle, trial_steps, ctypes.byref(us)), "visco autotune")
        return us.value

    def shape(self):
        """(kind, a, b): kind 0 = tiled (rows, waves), 1 = march (waves, chunk rows)."""
        k, a, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self.ctx.check(lib().sfd_visco_plan_get_shape(self.handle, ctypes.byref(k), ctypes.byref(a),
                                                      ctypes.byref(b)), "visco get_shape")
        return k.value, a.value, b.value

    def run(self) -> float:
        ms = ctypes.c_float()
        self.ctx.check(lib().sfd_visco_plan_run(self.handle, ctypes.byref(ms)), "sfd_visco_plan_run")
        return ms.value

    def fetch(self):
        nrec = self.p.nrec
        record = np.zeros((nrec, self.p.nt), dtype=np.float32)
        snaps = np.zeros((self.p.nsnap, self.shot.nfields(), self.nx, self.nz), dtype=np.float32)
        self.ctx.check(lib().sfd_visco_plan_fetch(self.handle, _ptr(record) if nrec and self.p.nt else None,
                                                  _ptr(snaps) if self.p.nsnap else None), "sfd_visco_plan_fetch")
        return record, snaps

    def fetch_xx(self) -> np.ndarray:
        """strxx at the receivers [nrec, nt] (the reference's record2)."""
        record = np.zeros((self.p.nrec, self.p.nt), dtype=np.float32)
        self.ctx.check(lib().sfd_visco_plan_fetch_xx(self.handle, _ptr(record) if self.p.nrec and self.p.nt else None),
                       "sfd_visco_plan_fetch_xx")
        return record

    def close(self):
        if getattr(self, "handle", None):
            lib().sfd_visco_plan_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ViscoBatch:
    """``sfd_visco_batch``: visco shots (windows of any width) advanced by one
    launch per half-step for all of them; the shots share nt, the delay, the
    radius, L and whether vs > 0."""

    def __init__(self, shots, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.ctx.adopt(self)
        self.shots = list(shots)
        n = len(self.shots)
        params = (ViscoParamsC * n)()
        self._keep = []
        arrs = [(_FP * n)() for _ in range(5)]  # vp, vs, rho, qp, wavelet
        for i, sh in enumerate(self.shots):
            params[i] = sh.params(self._keep)
            vp, vs, rho, qp, w = sh._arrays()
            self._keep += [vp, vs, rho, qp, w, sh]
            for a, v in zip(arrs, (vp, vs, rho, qp, w)):
                a[i] = _ptr(v)
        self.params = params
        h = ctypes.c_void_p()
        self.ctx.check(lib().sfd_visco_batch_create(self.ctx.handle, n, params, arrs[0], arrs[1], arrs[2], arrs[3],
                                                    arrs[4], ctypes.byref(h)), "sfd_visco_batch_create")
        self.handle = h

    def set_group_bytes(self, nbytes: int = 0) -> int:
        """Budget of one shot group's fields (0 = the default, -1 = every shot
        in one group); returns the number of groups."""
        v = (1 << 64) - 1 if nbytes < 0 else int(nbytes)
        self.ctx.check(lib().sfd_visco_batch_set_group_bytes(self.handle, v), "visco_batch set_group_bytes")
        return self.groups()

    def groups(self) -> int:
        n = ctypes.c_int()
        self.ctx.check(lib().sfd_visco_batch_get_groups(self.handle, ctypes.byref(n)), "visco_batch get_groups")
        return n.value

    def run(self) -> float:
        ms = ctypes.c_float()
        self.ctx.check(lib().sfd_visco_batch_run(self.handle, ctypes.byref(ms)), "sfd_visco_batch_run")
        return ms.value

    def fetch(self, shot: int):
        """Shot `shot`'s (record_zz [nrec, nt], record_xx [nrec, nt], snaps [nsnap, nfields, nx, nz])."""
        p, sh = self.params[shot], self.shots[shot]
        nx, nz = np.asarray(sh.vp).shape
        rzz = np.zeros((p.nrec, p.nt), dtype=np.float32)
        rxx = np.zeros((p.nrec, p.nt), dtype=np.float32)
        snaps = np.zeros((p.nsnap, sh.nfields(), nx, nz), dtype=np.float32)
        has = p.nrec and p.nt
        self.ctx.check(lib().sfd_visco_batch_fetch(self.handle, shot, _ptr(rzz) if has else None,
                                                   _ptr(rxx) if has else None, _ptr(snaps) if p.nsnap else None),
                       "sfd_visco_batch_fetch")
        return rzz, rxx, snaps

    def close(self):
        if getattr(self, "handle", None):
            lib().sfd_visco_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def visco_forward(shot: ViscoShot, ctx: Optional[Context] = None):
    """One visco-acoustic shot through ``sfd_visco_forward``.
    Returns (record [nrec, nt], snaps [nsnap, nfields, nx, nz])."""
    ctx = ctx or default_context()
    pkeep = []
    p = shot.params(pkeep)
    nx, nz = np.asarray(shot.vp).shape
    vp, vs, rho, qp, w = shot._arrays()
    record = np.zeros((p.nrec, p.nt), dtype=np.float32)
    snaps = np.zeros((p.nsnap, shot.nfields(), nx, nz), dtype=np.float32)
    ctx.check(lib().sfd_visco_forward(ctx.handle, ctypes.byref(p), _ptr(vp), _ptr(vs), _ptr(rho), _ptr(qp), _ptr(w),
                                      _ptr(record) if p.nrec and p.nt else None, _ptr(snaps) if p.nsnap else None),
              "sfd_visco_forward")
    return record, snaps


# ---------------------------------------------------------------- elastic

@dataclass
class ElasticShot:
    """One elastic free-surface shot on the reference's padded shot grid
    (forward/yang_version_2016_2_2, isoewm2dfsbc in 'fsbc' mode).

    vp, vs, rho: [nx, nz] float32, z fastest.  itopo [nx]: 0-based surface row
    of every column.  src = (sx, sz); rec = (rx[], rz[]).  sponge = (npxn,
    npxp, npzp).  nt = recorded samples; delay = steps before the first (None =
    nint(1.2 / (f0 dt))).  wavelet None = the reference's built-in Ricker.
    snap_fields: mask of FIELD_VX, FIELD_VZ, FIELD_SXX, FIELD_SZZ, FIELD_SXZ.
    mode: ELASTIC_FREE_SURFACE ('fsbc') or ELASTIC_ABSORBING_TOP ('abc', with
    sponge_zlo = the reference's npzn)."""
    vp: np.ndarray
    vs: np.ndarray
    rho: np.ndarray
    itopo: np.ndarray
    src: tuple
    rec: tuple
    nt: int
    dx: float = 5.0
    dz: float = 5.0
    dt: float = 1e-3
    f0: float = 25.0
    sponge: tuple = (40, 40, 40)
    sponge_r: float = 1e-3
    src_radius: int = 50
    delay: Optional[int] = None
    wavelet: Optional[np.ndarray] = None
    snap_steps: Sequence[int] = ()
    snap_fields: int = FIELD_VX | FIELD_VZ
    mode: int = 0
    sponge_zlo: int = 0
    _keep: list = field(default_factory=list, repr=False)

    def params(self, keep: Optional[list] = None) -> ElasticParamsC:
        """The C struct; its arrays live in ``keep`` (as AcousticShot.params)."""
        nx, nz = np.asarray(self.vp).shape
        it = np.ascontiguousarray(self.itopo, dtype=np.int32)
        rx = np.ascontiguousarray(self.rec[0], dtype=np.int32)
        rz = np.ascontiguousarray(self.rec[1], dtype=np.int32)
        steps = np.ascontiguousarray(list(self.snap_steps), dtype=np.int32)
        if keep is None:
            self._keep = keep = []
        keep += [it, rx, rz, steps]
        return ElasticParamsC(nx, nz, int(self.nt), -1 if self.delay is None else int(self.delay),
                              self.dx, self.dz, self.dt, self.f0, _ptr(it, ctypes.c_int),
                              *[int(v) for v in self.sponge], self.sponge_r, int(self.src[0]), int(self.src[1]),
                              int(self.src_radius), rx.size, _ptr(rx, ctypes.c_int), _ptr(rz, ctypes.c_int),
                              steps.size, _ptr(steps, ctypes.c_int), self.snap_fields, int(self.mode),
                              int(self.sponge_zlo))

    def nfields(self) -> int:
        return bin(self.snap_fields & 62).count("1")

    def ref_delay(self) -> int:
        """nint(1.2 / (f0 dt)) in float, the delay used when none is given."""
        if self.delay is not None:
            return int(self.delay)
        v = np.float32(1.2) / (np.float32(self.f0) * np.float32(self.dt))
        return int(np.floor(float(v) + 0.5))

    def _arrays(self):
        w = None if self.wavelet is None else _f32(self.wavelet)
        return _f32(self.vp), _f32(self.vs), _f32(self.rho), w


class ElasticPlan:
    """``sfd_elastic_plan``: model resident on the device."""

    def __init__(self, shot: ElasticShot, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.ctx.adopt(self)
        self.shot = shot
        self._pkeep = []
        self.p = shot.params(self._pkeep)
        self.nx, self.nz = np.asarray(shot.vp).shape
        self._arr = shot._arrays()
        vp, vs, rho, w = self._arr
        h = ctypes.c_void_p()
        self.ctx.check(lib().sfd_elastic_plan_create(self.ctx.handle, ctypes.byref(self.p), _ptr(vp), _ptr(vs),
                                                     _ptr(rho), _ptr(w), ctypes.byref(h)), "sfd_elastic_plan_create")
        self.handle = h

    def set_tiling(self, tile_cols: int = 0, waves: int = 0):
        self.ctx.check(lib().sfd_elastic_plan_set_tiling(self.handle, tile_cols, waves), "elastic set_tiling")

    def autotune(self, trial_steps: int = 20) -> float:
        us = ctypes.c_float()
        self.ctx.check(lib().sfd_elastic_plan_autotune(self.handle, trial_steps, ctypes.byref(us)),
                       "elastic autotune")
        return us.value

    def shape(self):
        """(tile columns, waves)."""
        a, b = ctypes.c_int(), ctypes.c_int()
        self.ctx.check(lib().sfd_elastic_plan_get_shape(self.handle, ctypes.byref(a), ctypes.byref(b)),
                       "elastic get_shape")
        return a.value, b.value

    def run(self) -> float:
        ms = ctypes.c_float()
        self.ctx.check(lib().sfd_elastic_plan_run(self.handle, ctypes.byref(ms)), "sfd_elastic_plan_run")
        return ms.value

    def fetch(self):
        """(record_vx [nrec, nt], record_vz [nrec, nt], snaps [nsnap, nfields, nx, nz])."""
        nrec, nt = self.p.nrec, self.p.nt
        rvx = np.zeros((nrec, nt), dtype=np.float32)
        rvz = np.zeros((nrec, nt), dtype=np.float32)
        snaps = np.zeros((self.p.nsnap, self.shot.nfields(), self.nx, self.nz), dtype=np.float32)
        has = nrec and nt
        self.ctx.check(lib().sfd_elastic_plan_fetch(self.handle, _ptr(rvx) if has else None,
                                                    _ptr(rvz) if has else None,
                                                    _ptr(snaps) if self.p.nsnap else None), "sfd_elastic_plan_fetch")
        return rvx, rvz, snaps

    def close(self):
        if getattr(self, "handle", None):
            lib().sfd_elastic_plan_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ElasticBatch:
    """``sfd_elastic_batch``: several elastic shots advanced by one launch per
    half-step (the shots share the mode, nt and the delay)."""

    def __init__(self, shots, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.ctx.adopt(self)
        self.shots = list(shots)
        n = len(self.shots)
        params = (ElasticParamsC * n)()
        self._keep = []
        vps, vss, rhos, ws = (_FP * n)(), (_FP * n)(), (_FP * n)(), (_FP * n)()
        for i, sh in enumerate(self.shots):
            params[i] = sh.params(self._keep)
            vp, vs, rho, w = sh._arrays()
            self._keep += [vp, vs, rho, w, sh]
            vps[i], vss[i], rhos[i], ws[i] = _ptr(vp), _ptr(vs), _ptr(rho), _ptr(w)
        self.params = params
        h = ctypes.c_void_p()
        self.ctx.check(lib().sfd_elastic_batch_create(self.ctx.handle, n, params, vps, vss, rhos, ws,
                                                      ctypes.byref(h)), "sfd_elastic_batch_create")
        self.handle = h

    def set_tiling(self, tile_cols: int = 0, waves: int = 0):
        self.ctx.check(lib().sfd_elastic_batch_set_tiling(self.handle, tile_cols, waves), "elastic batch set_tiling")

    def autotune(self, trial_steps: int = 20) -> float:
        us = ctypes.c_float()
        self.ctx.check(lib().sfd_elastic_batch_autotune(self.handle, trial_steps, ctypes.byref(us)),
                       "elastic batch autotune")
        return us.value

    def shape(self):
        """(tile columns, waves)."""
        a, b = ctypes.c_int(), ctypes.c_int()
        self.ctx.check(lib().sfd_elastic_batch_get_shape(self.handle, ctypes.byref(a), ctypes.byref(b)),
                       "elastic batch get_shape")
        return a.value, b.value

    def set_group_bytes(self, nbytes: int = 0) -> int:
        """Budget of one shot group's fields (0 = the default, -1 = every shot
        in one group); returns the number of groups."""
        v = (1 << 64) - 1 if nbytes < 0 else int(nbytes)
        self.ctx.check(lib().sfd_elastic_batch_set_group_bytes(self.handle, v), "elastic_batch set_group_bytes")
        return self.groups()

    def groups(self) -> int:
        n = ctypes.c_int()
        self.ctx.check(lib().sfd_elastic_batch_get_groups(self.handle, ctypes.byref(n)), "elastic_batch get_groups")
        return n.value

    def run(self) -> float:
        ms = ctypes.c_float()
        self.ctx.check(lib().sfd_elastic_batch_run(self.handle, ctypes.byref(ms)), "sfd_elastic_batch_run")
        return ms.value

    def fetch(self, shot: int):
        """Shot `shot`'s (record_vx [nrec, nt], record_vz [nrec, nt], snaps [nsnap, nfields, nx, nz])."""
        p, sh = self.params[shot], self.shots[shot]
        nx, nz = np.asarray(sh.vp).shape
        rvx = np.zeros((p.nrec, p.nt), dtype=np.float32)
        rvz = np.zeros((p.nrec, p.nt), dtype=np.float32)
        snaps = np.zeros((p.nsnap, sh.nfields(), nx, nz), dtype=np.float32)
        has = p.nrec and p.nt
        self.ctx.check(lib().sfd_elastic_batch_fetch(self.handle, shot, _ptr(rvx) if has else None,
                                                     _ptr(rvz) if has else None, _ptr(snaps) if p.nsnap else None),
                       "sfd_elastic_batch_fetch")
        return rvx, rvz, snaps

    def close(self):
        if getattr(self, "handle", None):
            lib().sfd_elastic_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def elastic_forward(shot: ElasticShot, ctx: Optional[Context] = None):
    """One elastic shot through ``sfd_elastic_forward``.
    Returns (record_vx, record_vz [nrec, nt], snaps [nsnap, nfields, nx, nz])."""
    ctx = ctx or default_context()
    pkeep = []
    p = shot.params(pkeep)
    nx, nz = np.asarray(shot.vp).shape
    vp, vs, rho, w = shot._arrays()
    rvx = np.zeros((p.nrec, p.nt), dtype=np.float32)
    rvz = np.zeros((p.nrec, p.nt), dtype=np.float32)
    snaps = np.zeros((p.nsnap, shot.nfields(), nx, nz), dtype=np.float32)
    has = p.nrec and p.nt
    ctx.check(lib().sfd_elastic_forward(ctx.handle, ctypes.byref(p), _ptr(vp), _ptr(vs), _ptr(rho), _ptr(w),
                                        _ptr(rvx) if has else None, _ptr(rvz) if has else None,
                                        _ptr(snaps) if p.nsnap else None), "sfd_elastic_forward")
    return rvx, rvz, snaps


# ---------------------------------------------------------------- VTI

def vti_coefficients(kind: int, k: int = 5) -> np.ndarray:
    """coefficient_2nd (kind 2) / coefficient_1st (kind 1), solved as the reference does."""
    out = np.zeros(k, dtype=np.float32)
    rc = lib().sfd_vti_coefficients(kind, k, _ptr(out))
    if rc != 0:
        raise SeisfdError(rc, "sfd_vti_coefficients: bad request")
    return out


def vti_wavelet(dt: float, f0: float) -> np.ndarray:
    """The VTI driver's wavelet, nwt = nint(0.14 / dt) + 1 samples."""
    n = lib().sfd_vti_wavelet(dt, f0, 0, None)
    if n < 0:
        raise SeisfdError(n, "sfd_vti_wavelet: bad dt")
    w = np.zeros(n, dtype=np.float32)
    lib().sfd_vti_wavelet(dt, f0, n, _ptr(w))
    return w


def vti_window(vz, eps, delta, xmin: int, rng: int, pml, nsz0: int):
    """get_currshot_parameters plus the isotropic top rows (host).  Model
    [nxv, nzv] z fastest; pml = (left, right, top, bottom) cells.  Returns the
    padded shot grid's (vz, eps, delta), each [rng + nl + nr - 2, nzv + nu + nd - 2]."""
    vz, eps, delta = _f32(vz), _f32(eps), _f32(delta)
    nxv, nzv = vz.shape
    nl, nr, nu, nd = (int(v) for v in pml)
    shape = (rng + nl + nr - 2, nzv + nu + nd - 2)
    out = [np.zeros(shape, dtype=np.float32) for _ in range(3)]
    rc = lib().sfd_vti_window(_ptr(vz), _ptr(eps), _ptr(delta), nxv, nzv, int(xmin), int(rng), nl, nr, nu, nd,
                              int(nsz0), *(_ptr(o) for o in out))
    if rc != 0:
        raise SeisfdError(rc, "sfd_vti_window: bad window")
    return tuple(out)


class VtiPlan:
    """``sfd_vti_plan``: nshot shots on grids of one shape, models resident on
    the device.  vz, eps, delta [nshot, nx, nz] (or [nx, nz] for one shot);
    pml = (left, right, top, bottom) cells; rec = (rx, rz) cell lists."""

    def __init__(self, vz, eps, delta, *, dx, dz, dt, nt, src, rec, pml=(0, 0, 0, 0), f0=30.0, delay=None,
                 pml_r=0.0, wavelet=None, c2=None, c1=None, allow_eps_lt_delta=False,
                 zero_first_sample=False, ctx: Optional[Context] = None):
        # zero_first_sample: the driver's rule when its receiver row was moved
        # down to row 6 (VTI main:469-476): sample 1 of every trace is never
        # written.  fetch applies it.
        self.zero_first_sample = bool(zero_first_sample)
        vz, eps, delta = (_f32(a) for a in (vz, eps, delta))
        if vz.ndim == 2:
            vz, eps, delta = vz[None], eps[None], delta[None]
        if not (vz.shape == eps.shape == delta.shape):
            raise ValueError("vz, eps and delta must have one shape")
        self.nshot, self.nx, self.nz = vz.shape
        self._rx = np.ascontiguousarray(rec[0], dtype=np.int32)
        self._rz = np.ascontiguousarray(rec[1], dtype=np.int32)
        self._c2 = None if c2 is None else _f32(c2)
        self._c1 = None if c1 is None else _f32(c1)
        for name, c in (("c2", self._c2), ("c1", self._c1)):
            if c is not None and c.size != 5:
                raise ValueError(f"{name} must hold 5 weights")
        self._w = None if wavelet is None else _f32(wavelet)
        if self._w is not None:
            # the library reads delay + nt samples; the reference's delay is
            # nint((nwt - 1) / 2) with nwt the built-in wavelet's length
            d = int(np.floor((vti_wavelet(dt, f0).size - 1) / 2.0 + 0.5)) if delay is None or delay < 0 else delay
            if self._w.size < d + int(nt):
                raise ValueError(f"wavelet needs delay + nt = {d + int(nt)} samples, got {self._w.size}")
        self._arr = (vz, eps, delta)
        self.p = VtiParamsC(self.nx, self.nz, int(nt), -1 if delay is None else int(delay), dx, dz, dt, f0,
                            *(int(v) for v in pml), pml_r, _ptr(self._c2), _ptr(self._c1), int(src[0]), int(src[1]),
                            self._rx.size, _ptr(self._rx, ctypes.c_int), _ptr(self._rz, ctypes.c_int),
                            int(bool(allow_eps_lt_delta)))
        self.ctx = ctx or default_context()
        self.ctx.adopt(self)
        h = ctypes.c_void_p()
        self.ctx.check(lib().sfd_vti_plan_create(self.ctx.handle, ctypes.byref(self.p), self.nshot, _ptr(vz),
                                                 _ptr(eps), _ptr(delta), _ptr(self._w), ctypes.byref(h)),
                       "sfd_vti_plan_create")
        self.handle = h

    def set_tiling(self, tile_cols: int = 0, waves: int = 0):
        self.ctx.check(lib().sfd_vti_plan_set_tiling(self.handle, tile_cols, waves), "vti set_tiling")

    def autotune(self, trial_steps: int = 20) -> float:
        us = ctypes.c_float()
        self.ctx.check(lib().sfd_vti_plan_autotune(self.handle, trial_steps, ctypes.byref(us)), "vti autotune")
        return us.value

    def shape(self):
        """(tile columns, waves)."""
        a, b = ctypes.c_int(), ctypes.c_int()
        self.ctx.check(lib().sfd_vti_plan_get_shape(self.handle, ctypes.byref(a), ctypes.byref(b)), "vti get_shape")
        return a.value, b.value

    def set_group_bytes(self, nbytes: int = 0) -> int:
        """Budget of one shot group's fields (0 = the default, -1 = every shot
        in one group); returns the number of groups."""
        v = (1 << 64) - 1 if nbytes < 0 else int(nbytes)
        self.ctx.check(lib().sfd_vti_plan_set_group_bytes(self.handle, v), "vti set_group_bytes")
        return self.groups()

    def groups(self) -> int:
        n = ctypes.c_int()
        self.ctx.check(lib().sfd_vti_plan_get_groups(self.handle, ctypes.byref(n)), "vti get_groups")
        return n.value

    def run(self) -> float:
        ms = ctypes.c_float()
        self.ctx.check(lib().sfd_vti_plan_run(self.handle, ctypes.byref(ms)), "sfd_vti_plan_run")
        return ms.value

    def fetch(self, state: bool = False):
        """record [nshot, nrec, nt] (and with state=True, p and q [nshot, 2, nx, nz])."""
        nrec, nt = self.p.nrec, self.p.nt
        record = np.zeros((self.nshot, nrec, nt), dtype=np.float32)
        st = np.zeros((self.nshot, 2, self.nx, self.nz), dtype=np.float32) if state else None
        self.ctx.check(lib().sfd_vti_plan_fetch(self.handle, _ptr(record) if nrec and nt else None,
                                                _ptr(st) if state else None), "sfd_vti_plan_fetch")
        if self.zero_first_sample and nt:
            record[:, :, 0] = 0.0
        return (record, st) if state else record

    def close(self):
        if getattr(self, "handle", None):
            lib().sfd_vti_plan_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- SRM

SRM_AUTO, SRM_DIRECT, SRM_GEMM = 0, 1, 2


def srm_crg_traces(nshot: int, ntrace: int) -> int:
    """Traces of the common-receiver array the prediction reads: ntrace (2 nshot - 1)."""
    return lib().sfd_srm_crg_traces(nshot, ntrace)


class SrmPlan:
    """``sfd_srm_plan``: convolution multiple prediction with csg [nshot,
    ntrace, lt] and crg [ncrg, lt] resident on the device."""

    def __init__(self, csg, crg, ctx: Optional[Context] = None):
        csg, crg = _f32(csg), _f32(crg)
        if csg.ndim != 3 or crg.ndim != 2 or crg.shape[1] != csg.shape[2]:
            raise ValueError("csg must be [nshot, ntrace, lt] and crg [ncrg, lt]")
        self.ctx = ctx or default_context()
        self.ctx.adopt(self)
        self.shape = csg.shape
        self.p = SrmParamsC(csg.shape[0], csg.shape[1], csg.shape[2], crg.shape[0])
        h = ctypes.c_void_p()
        self.ctx.check(lib().sfd_srm_plan_create(self.ctx.handle, ctypes.byref(self.p), _ptr(csg), _ptr(crg),
                                                 ctypes.byref(h)), "sfd_srm_plan_create")
        self.handle = h

    def set_method(self, method: int):
        """SRM_AUTO, SRM_DIRECT or SRM_GEMM (ntrace = nshot only)."""
        self.ctx.check(lib().sfd_srm_plan_set_method(self.handle, method), "sfd_srm_plan_set_method")

    def method(self) -> int:
        return lib().sfd_srm_plan_get_method(self.handle)

    def run(self) -> float:
        ms = ctypes.c_float()
        self.ctx.check(lib().sfd_srm_plan_run(self.handle, ctypes.byref(ms)), "sfd_srm_plan_run")
        return ms.value

    def fetch(self) -> np.ndarray:
        out = np.zeros(self.shape, dtype=np.float32)
        self.ctx.check(lib().sfd_srm_plan_fetch(self.handle, _ptr(out)), "sfd_srm_plan_fetch")
        return out

    def close(self):
        if getattr(self, "handle", None):
            lib().sfd_srm_plan_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def srm_predict(csg, crg, ctx: Optional[Context] = None) -> np.ndarray:
    """Predicted multiples [nshot, ntrace, lt] through ``sfd_srm_predict``."""
    csg, crg = _f32(csg), _f32(crg)
    if csg.ndim != 3 or crg.ndim != 2 or crg.shape[1] != csg.shape[2]:
        raise ValueError("csg must be [nshot, ntrace, lt] and crg [ncrg, lt]")
    ctx = ctx or default_context()
    p = SrmParamsC(csg.shape[0], csg.shape[1], csg.shape[2], crg.shape[0])
    out = np.zeros(csg.shape, dtype=np.float32)
    ctx.check(lib().sfd_srm_predict(ctx.handle, ctypes.byref(p), _ptr(csg), _ptr(crg), _ptr(out)), "sfd_srm_predict")
    return out


# ---------------------------------------------------------------- files

def _par_dict(st) -> dict:
    out = {}
    for name, _ in st._fields_:
        v = getattr(st, name)
        out[name] = v.decode() if isinstance(v, bytes) else v
    return out


def _read_par(fn: str, cls, path: str) -> dict:
    par = cls()
    err = ctypes.create_string_buffer(256)
    rc = getattr(lib(), fn)(os.fsencode(path), ctypes.byref(par), err)
    if rc != SFD_OK:
        raise SeisfdError(rc, f"{fn}({path}): {err.value.decode()}")
    return _par_dict(par)


def read_q_par(path: str) -> dict:
    """Q_forward's parameter card (read_par, Q sub:121-234) as a dict."""
    return _read_par("sfd_read_q_par", QParC, path)


def read_iso_par(path: str) -> dict:
    """yang_version's parameter card (read_par, pub:48-145) as a dict."""
    return _read_par("sfd_read_iso_par", IsoParC, path)


def write_su_shot(path: str, record, shot_no: int, src_col: int, first_col: int, dx: float, dt: float,
                  with_cdp: bool = False):
    """One shot's gather [ntrace, nt] as an SU file (write_currshot_disk)."""
    rec = _f32(record)
    rc = lib().sfd_write_su_shot(os.fsencode(path), _ptr(rec), rec.shape[0], rec.shape[1], shot_no, src_col,
                                 first_col, dx, dt, int(with_cdp))
    if rc != SFD_OK:
        raise SeisfdError(rc, f"sfd_write_su_shot({path})")


def merge_su(shot_files, ntrace: int, nt: int, out_path: str, remove_inputs: bool = False):
    """Concatenates per-shot SU files of ntrace traces each (merge_shot_files)."""
    arr = (ctypes.c_char_p * len(shot_files))(*[os.fsencode(f) for f in shot_files])
    rc = lib().sfd_merge_su(arr, len(shot_files), ntrace, nt, os.fsencode(out_path), int(remove_inputs))
    if rc != SFD_OK:
        raise SeisfdError(rc, f"sfd_merge_su({out_path})")


BW_READ = 0
BW_WRITE = 1
BW_COPY = 2


def bandwidth_probe(nbytes: int, kind: int = BW_COPY, reps: int = 5, ctx: Optional[Context] = None) -> float:
    """Streaming rate of a hand-written kernel over an nbytes device buffer
    (BW_READ, BW_WRITE or BW_COPY, the copy counting both buffers), in GB/s:
    the achievable HBM rate the stencil steps are read against."""
    ctx = ctx or default_context()
    g = ctypes.c_double()
    ctx.check(lib().sfd_bandwidth_probe(ctx.handle, int(nbytes), int(kind), int(reps), ctypes.byref(g)),
              "sfd_bandwidth_probe")
    return g.value
