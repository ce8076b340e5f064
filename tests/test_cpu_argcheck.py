"""Argument checks of the C-ABI entry points that create handles.  Each bad
argument is rejected with HES_ERR_ARG and a message before any device call is
made, so these run without a GPU; a valid call would allocate device memory
and is left to the GPU tests."""
import ctypes
import os

import numpy as np
import pytest

from hyperelasticsolver_amd import _lib

ERR_ARG = -1
MAT = np.array([8.93, 4.6, 3.9e-4, 300.0, 2.1, 1.0, 3.0, 2.0])   # Barton2009() defaults
MAT_GEN = np.array([8.93, 6.22, 9.0e-4, 300.0, 3.16, 1.0, 3.577, 2.088])  # main.jl:133


@pytest.fixture(scope="module")
def L():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail(f"{_lib.LIB_PATH} not built")
    return _lib.lib()


def _mat(**kw):
    m = MAT.copy()
    for k, v in kw.items():
        m[["rho0", "c0", "cv", "t0", "b0", "alpha", "beta", "gamma"].index(k)] = v
    return m


def _expect(L, rc, h, words):
    assert rc == ERR_ARG
    assert not h.value
    msg = L.hes_last_error().decode()
    assert words in msg, msg


@pytest.mark.parametrize("nphase,flux,ncells,mat,cfl,dx,words", [
    (3, 0, 100, MAT, 0.6, 0.01, "nphase"),
    (1, 7, 100, MAT, 0.6, 0.01, "unknown flux"),
    (1, 0, 2, MAT, 0.6, 0.01, "at least 3 cells"),
    (1, 0, 100, _mat(rho0=0.0), 0.6, 0.01, "material"),
    (1, 0, 100, _mat(cv=np.nan), 0.6, 0.01, "material"),
    (1, 0, 100, MAT, 0.0, 0.01, "positive"),
    (1, 0, 100, MAT, 0.6, -1.0, "positive"),
])
def test_solver_create_rejects(L, nphase, flux, ncells, mat, cfl, dx, words):
    m = np.ascontiguousarray(np.tile(mat, nphase if nphase in (1, 2) else 1))
    h = ctypes.c_void_p()
    rc = L.hes_solver_create(ctypes.byref(h), nphase, flux, ncells, _lib.ptr(m), cfl, dx)
    _expect(L, rc, h, words)


def _ll(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int64))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))


@pytest.mark.parametrize("nphase,flux,ncells,words", [
    (1, 0, [100, 2, 50], "at least 3 cells"),
    (1, 5, [100], "unknown flux"),
    (4, 0, [100], "nphase"),
])
def test_batch_create_rejects(L, nphase, flux, ncells, words):
    nc, pnc = _ll(ncells)
    m = np.ascontiguousarray(np.tile(MAT, len(ncells) * (nphase if nphase in (1, 2) else 1)))
    h = ctypes.c_void_p()
    rc = L.hes_batch_create(ctypes.byref(h), nphase, flux, len(ncells), pnc, _lib.ptr(m), None, None)
    _expect(L, rc, h, words)


def test_batch_create_rejects_a_bad_material_and_dx(L):
    nc, pnc = _ll([100, 200])
    h = ctypes.c_void_p()
    m = np.ascontiguousarray(np.concatenate([MAT, _mat(t0=-1.0)]))
    _expect(L, L.hes_batch_create(ctypes.byref(h), 1, 0, 2, pnc, _lib.ptr(m), None, None), h,
            "material")
    m = np.ascontiguousarray(np.tile(MAT, 2))
    dx = np.array([0.01, 0.0])
    _expect(L, L.hes_batch_create(ctypes.byref(h), 1, 0, 2, pnc, _lib.ptr(m), None, _lib.ptr(dx)),
            h, "positive")


@pytest.mark.parametrize("ncells,mat,cfl,nsteps,every,words", [
    (2, MAT, 0.6, 10, 5, "at least 3 cells"),
    (100, MAT, 0.6, 0, 5, "at least one step"),
    (100, MAT, 0.6, 10, 0, "interval"),
    (100, MAT_GEN, 0.6, 10, 5, "exponents"),
    (100, MAT, -0.6, 10, 5, "positive"),
    (100, MAT, 0.6, 1 << 40, 5, "too many steps"),
    (1 << 28, MAT, 0.6, 10, 5, "2^28"),
])
def test_adjoint_create_rejects(L, ncells, mat, cfl, nsteps, every, words):
    m = np.ascontiguousarray(mat)
    h = ctypes.c_void_p()
    rc = L.hes_adjoint_create(ctypes.byref(h), ncells, _lib.ptr(m), cfl, 0.01, nsteps, every)
    _expect(L, rc, h, words)


@pytest.mark.parametrize("ncells,mats,dx,nsteps,every,words", [
    ([100, 2], [MAT, MAT], None, 10, 5, "at least 3 cells"),
    ([100, 200], [MAT, MAT_GEN], None, 10, 5, "exponents"),
    ([100, 200], [MAT, _mat(alpha=0.0)], None, 10, 5, "material"),
    ([100, 200], [MAT, MAT], [0.01, -0.01], 10, 5, "dx"),
    ([100, 200], [MAT, MAT], None, 0, 5, "at least one step"),
    ([100, 200], [MAT, MAT], None, 10, 0, "interval"),
    ([100, 200], [MAT, MAT], None, (1 << 31) - 1, 5, "too many steps"),
    ([1 << 27, 1 << 27], [MAT, MAT], None, 10, 5, "2^28"),
])
def test_batched_adjoint_create_rejects(L, ncells, mats, dx, nsteps, every, words):
    nc, pnc = _ll(ncells)
    m = np.ascontiguousarray(np.concatenate(mats))
    d = None if dx is None else np.ascontiguousarray(np.asarray(dx, dtype=np.float64))
    h = ctypes.c_void_p()
    rc = L.hes_badjoint_create(ctypes.byref(h), len(ncells), pnc, _lib.ptr(m), 0.6,
                               None if d is None else _lib.ptr(d), nsteps, every)
    _expect(L, rc, h, words)


def test_null_arguments_are_rejected(L):
    h = ctypes.c_void_p()
    _expect(L, L.hes_solver_create(ctypes.byref(h), 1, 0, 100, None, 0.6, 0.01), h, "null")
    _expect(L, L.hes_batch_create(ctypes.byref(h), 1, 0, 0, None, None, None, None), h, "null")
    _expect(L, L.hes_badjoint_create(ctypes.byref(h), 0, None, None, 0.6, None, 10, 5), h, "null")
    assert L.hes_solver_create(None, 1, 0, 100, None, 0.6, 0.01) == ERR_ARG
