"""hes_max_speed_noalloc (no device allocation; its partial maxima live in a
fixed device array) against hes_max_speed on the same states: the per-state
speeds bit for bit, and the maximum equal to their maximum, on two calls in a
row, the second with fewer states.  The first call's fastest state sits in its
last block, beyond the second call's blocks, so a second call that read the
first call's partial maxima would return the wrong maximum.  One and two
phases; invalid states raise DomainError; more states than the fixed array
allows is an argument error raised before any device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)


class _Hip:
    """The HIP runtime libhes.so is linked against, for the test's own device
    buffers (the entry point under test allocates nothing)."""

    def __init__(self):
        L = ctypes.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L = L

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(int(nbytes), 8)) == 0
        return p

    def free(self, p):
        assert self.L.hipFree(p) == 0

    def upload(self, dst, a):
        assert a.flags.f_contiguous and a.dtype == np.float64
        assert self.L.hipMemcpy(dst, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0

    def download(self, src, n):
        out = np.empty(n)
        assert self.L.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), src, out.nbytes, 2) == 0
        return out


def _mats(nph):
    from test_oracle import MAT, MAT2
    return (MAT if nph == 1 else MAT2).astype(np.float64).copy()


def _states(ora, rng, nph, n, hot_last):
    """n states drawn from 7 seeded ones; with hot_last the last state moves at
    u1 = 20 in every phase (the others have |u1| <= 1/2 and sound speeds of
    about 10, so its speed is about twice theirs)."""
    from test_oracle import _rand_prims
    mats = _mats(nph)
    Qk = np.stack([ora.prim2cons(nph, mats, _rand_prims(rng, nph, mats)) for _ in range(7)], axis=1)
    Q = np.asfortranarray(Qk[:, rng.integers(0, 7, n)])
    if hot_last:
        P = _rand_prims(rng, nph, mats)
        for p in range(nph):
            P[(0 if nph == 1 else 15 * p + 1) + 1] = 20.0
        Q[:, -1] = ora.prim2cons(nph, mats, P)
    return Q


def _max_speed(nph, Q):
    from hyperelasticsolver_amd._lib import check, lib
    lam = np.empty(Q.shape[1])
    check(lib().hes_max_speed(nph, _mats(nph).ctypes.data_as(_dp), Q.ctypes.data_as(_dp),
                              lam.ctypes.data_as(_dp), Q.shape[1]))
    return lam


def _noalloc(hip, nph, Q):
    from hyperelasticsolver_amd._lib import check, lib
    n = Q.shape[1]
    dq, dl = hip.malloc(Q.nbytes), hip.malloc(8 * n)
    try:
        hip.upload(dq, Q)
        lmax = ctypes.c_double(-1.0)
        check(lib().hes_max_speed_noalloc(nph, _mats(nph).ctypes.data_as(_dp), dq, dl,
                                          ctypes.byref(lmax), n))
        lam = hip.download(dl, n)
    finally:
        hip.free(dq)
        hip.free(dl)
    return lam, lmax.value


@pytest.mark.parametrize("nph", [1, 2])
def test_noalloc_max_speed_matches_hes_max_speed(ora, nph):
    hip = _Hip()
    rng = np.random.default_rng(5 + nph)
    maxima = []
    for n, hot in ((300_001, True), (1_000, False)):
        print(f"[noalloc] nph={nph} n={n}", flush=True)
        Q = _states(ora, rng, nph, n, hot)
        lam, lmax = _noalloc(hip, nph, Q)
        ref = _max_speed(nph, Q)
        assert np.array_equal(lam, ref), n
        assert lmax == ref.max(), (n, lmax, ref.max())
        maxima.append(lmax)
    assert maxima[0] > 1.5 * maxima[1]  # the stale maximum would have been visible


def test_noalloc_errors(ora):
    from hyperelasticsolver_amd._lib import DomainError, HesError, check, lib
    print("[noalloc] errors", flush=True)
    hip = _Hip()
    rng = np.random.default_rng(9)
    Q = _states(ora, rng, 1, 500, False)
    Q[5:, 123] = 0.0  # det(rho A) = 0: the reference's DomainError
    dq, dl = hip.malloc(Q.nbytes), hip.malloc(8 * 500)
    try:
        hip.upload(dq, Q)
        lmax = ctypes.c_double()
        with pytest.raises(DomainError):
            check(lib().hes_max_speed_noalloc(1, _mats(1).ctypes.data_as(_dp), dq, dl,
                                              ctypes.byref(lmax), 500))
        # too many states: rejected on the host before any device access
        with pytest.raises(HesError):
            check(lib().hes_max_speed_noalloc(1, _mats(1).ctypes.data_as(_dp), dq, dl,
                                              ctypes.byref(lmax), (1 << 26) + 1))
        # the next valid call is unaffected by the failed ones
        good = _states(ora, rng, 1, 500, False)
        lam, lm = _noalloc(hip, 1, good)
        assert np.array_equal(lam, _max_speed(1, good)) and lm == lam.max()
    finally:
        hip.free(dq)
        hip.free(dl)
