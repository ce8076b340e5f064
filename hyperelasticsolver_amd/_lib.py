"""ctypes binding of libhes.so (the gfx950 kernels behind include/hes.h).

The library is required: importing the package works without it, but every
call that needs the GPU raises `HesLibraryMissing` if it has not been built.
There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
# HES_LIB selects another build of the library (e.g. a variant under test)
LIB_PATH = os.environ.get("HES_LIB") or os.path.join(_PKG, "libhes.so")

HES_OK = 0
HES_ERR_ARG = -1
HES_ERR_HIP = -2
HES_ERR_DOMAIN = -3

FLUX_RUSANOV = 0
FLUX_HLL = 1
FLUX_LXF = 2


class HesError(RuntimeError):
    """A libhes call failed (argument error or HIP runtime error)."""


class DomainError(HesError, ValueError):
    """Non-positive det(rho A) or a non-finite state: the reference raises a
    Julia DomainError from sqrt (HyperelasticityMPh.jl:114)."""


class HesLibraryMissing(HesError, ImportError):
    """libhes.so has not been built."""


_dp = ctypes.POINTER(ctypes.c_double)
_ll = ctypes.c_longlong
_i = ctypes.c_int
_d = ctypes.c_double
_vp = ctypes.c_void_p

# name -> argtypes (all return int status)
_SIGS = {
    "hes_version": [],
    "hes_device_info": [ctypes.c_char_p, _i, ctypes.POINTER(_i), _dp],
    "hes_solver_create": [ctypes.POINTER(_vp), _i, _i, _ll, _dp, _d, _d],
    "hes_solver_destroy": [_vp],
    "hes_solver_set_state": [_vp, _dp, _d],
    "hes_solver_set_state_device": [_vp, _vp, _d],
    "hes_solver_set_piecewise": [_vp, _i, _dp, _dp, _d],
    "hes_solver_get_state": [_vp, _dp],
    "hes_solver_get_speeds": [_vp, _dp],
    "hes_solver_time": [_vp, _dp, ctypes.POINTER(_ll)],
    "hes_solver_next_dt": [_vp, _dp],
    "hes_solver_run": [_vp, _ll, _d, _i, ctypes.POINTER(_ll), _dp,
                       ctypes.POINTER(ctypes.c_float), _dp],
    "hes_solver_time_steps": [_vp, _ll, ctypes.POINTER(ctypes.c_float)],
    "hes_solver_set_relaxation": [_vp, _d],
    "hes_solver_set_pressure_relaxation": [_vp, _d],
    "hes_copy_bandwidth": [_ll, _i, _dp],
    "hes_adjoint_create": [ctypes.POINTER(_vp), _ll, _dp, _d, _d, _ll, _ll],
    "hes_adjoint_destroy": [_vp],
    "hes_adjoint_set_initial": [_vp, _i, _dp, _dp],
    "hes_adjoint_set_target": [_vp, _dp, _dp],
    "hes_adjoint_forward": [_vp, _dp, ctypes.POINTER(ctypes.c_float)],
    "hes_adjoint_backward": [_vp, _dp, _dp, ctypes.POINTER(ctypes.c_float)],
    "hes_adjoint_set_sparse": [_vp, _i],
    "hes_adjoint_set_launch": [_vp, _i],
    "hes_adjoint_get_state": [_vp, _dp],
    "hes_adjoint_get_profiles": [_vp, _dp, _dp],
    "hes_adjoint_info": [_vp, ctypes.POINTER(_ll), ctypes.POINTER(_ll)],
    "hes_badjoint_create": [ctypes.POINTER(_vp), _i, ctypes.POINTER(_ll), _dp, _d, _dp, _ll, _ll],
    "hes_badjoint_destroy": [_vp],
    "hes_badjoint_set_initial": [_vp, ctypes.POINTER(_i), _dp, _dp],
    "hes_badjoint_set_target": [_vp, _dp, _dp],
    "hes_badjoint_forward": [_vp, _dp, _dp, ctypes.POINTER(ctypes.c_float)],
    "hes_badjoint_backward": [_vp, _dp, _dp, ctypes.POINTER(ctypes.c_float)],
    "hes_badjoint_set_sparse": [_vp, _i],
    "hes_badjoint_set_launch": [_vp, _i],
    "hes_badjoint_get_state": [_vp, _dp],
    "hes_badjoint_get_profiles": [_vp, _dp, _dp],
    "hes_badjoint_info": [_vp, ctypes.POINTER(_ll), ctypes.POINTER(_ll), ctypes.POINTER(_ll)],
    "hes_batch_create": [ctypes.POINTER(_vp), _i, _i, _i, ctypes.POINTER(_ll), _dp, _dp, _dp],
    "hes_batch_destroy": [_vp],
    "hes_batch_info": [_vp, ctypes.POINTER(_ll), ctypes.POINTER(_i), ctypes.POINTER(_i),
                       ctypes.POINTER(_i), ctypes.POINTER(_i)],
    "hes_batch_set_state": [_vp, _dp, _dp],
    "hes_batch_set_piecewise": [_vp, ctypes.POINTER(_i), _dp, _dp, _dp],
    "hes_batch_get_state": [_vp, _dp],
    "hes_batch_get_times": [_vp, _dp, ctypes.POINTER(_ll)],
    "hes_batch_set_t_end": [_vp, _dp],
    "hes_batch_set_launch": [_vp, _i, _i, _i],
    "hes_batch_run": [_vp, _ll, _i, ctypes.POINTER(ctypes.c_float)],
    "hes_batch_autotune": [_vp, _i, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i),
                           ctypes.POINTER(ctypes.c_float)],
    "hes_cons2prim": [_i, _dp, _dp, _dp, _ll],
    "hes_prim2cons": [_i, _dp, _dp, _dp, _ll],
    "hes_flux": [_i, _dp, _dp, _dp, _ll],
    "hes_get_eigvals": [_i, _dp, _dp, _dp, _dp, _ll],
    "hes_max_speed": [_i, _dp, _dp, _dp, _ll],
    "hes_max_speed_noalloc": [_i, _dp, _vp, _vp, _dp, _ll],
    "hes_noncons_flux": [_dp, _dp, _dp, _ll],
    "hes_face_flux": [_i, _i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ll],
    "hes_phase_cons2prim": [_dp, _dp, _dp, _ll],
    "hes_phase_flux": [_dp, _dp, _dp, _ll],
    "hes_phase_get_eigvals": [_dp, _dp, _dp, _dp, _ll],
    "hes_sp_cons2prim": [_dp, _dp, _dp, _ll],
    "hes_sp_prim2cons": [_dp, _dp, _dp, _ll],
    "hes_sp_flux": [_dp, _dp, _dp, _ll],
    "hes_energy": [_dp, _dp, _dp, _dp, _ll],
    "hes_entropy": [_dp, _dp, _dp, _dp, _ll],
    "hes_temperature": [_dp, _dp, _dp, _dp, _ll],
    "hes_stress": [_dp, _dp, _dp, _dp, _ll],
    "hes_stress_den": [_dp, _dp, _dp, _dp, _dp, _ll],
    "hes_acoustic": [_dp, _dp, _dp, _dp, _dp, _ll],
    "hes_acoustic_prim": [_dp, _dp, _dp, _dp, _ll],
    "hes_finger": [_dp, _dp, _ll],
    "hes_invariants": [_dp, _dp, _ll],
}

_lib = None


def lib():
    """Load libhes.so once; raise HesLibraryMissing if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HesLibraryMissing(
            f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    L.hes_last_error.argtypes = []
    L.hes_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def exported_symbols():
    """Names of the C-ABI entry points this binding knows about."""
    return sorted(list(_SIGS) + ["hes_last_error"])


def check(rc):
    if rc == HES_OK:
        return
    msg = lib().hes_last_error().decode(errors="replace")
    if rc == HES_ERR_DOMAIN:
        raise DomainError(msg)
    raise HesError(f"libhes error {rc}: {msg}")


def ptr(a):
    """ctypes double* of a contiguous float64 numpy array."""
    assert a.dtype == np.float64 and (a.flags.c_contiguous or a.flags.f_contiguous)
    return a.ctypes.data_as(_dp)


def columns(a, k, name="array"):
    """Return (F-ordered float64 (k, n) array, n, was_vector) for a Julia-style
    k-vector or (k, n) matrix (component fastest in memory)."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        if a.shape[0] != k:
            raise ValueError(f"{name}: expected length {k}, got {a.shape[0]}")
        return np.asfortranarray(a.reshape(k, 1)), 1, True
    if a.ndim != 2 or a.shape[0] != k:
        raise ValueError(f"{name}: expected shape ({k}, n), got {a.shape}")
    return np.asfortranarray(a), a.shape[1], False
