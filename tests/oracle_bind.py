"""ctypes binding of the CPU oracle (oracle/libhes_oracle.so) for the tests.

Test infrastructure only: the oracle checks results and is never the thing measured.
"""
import ctypes
import math
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "libhes_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_d = ctypes.c_double
_i = ctypes.c_int


def P(a):
    assert a.dtype == np.float64 and (a.flags.c_contiguous or a.flags.f_contiguous)
    return a.ctypes.data_as(_dp)


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            import subprocess
            subprocess.run(["make", "-s", "-C", os.path.dirname(path)], check=True)
        L = ctypes.CDLL(path)
        for name in ("ora_energy", "ora_entropy", "ora_det", "ora_det_lu", "ora_temperature"):
            getattr(L, name).restype = _d
        self.L = L

    # --- EoS level, single point
    def energy(self, mat, S, G):
        return self.L.ora_energy(P(mat), _d(S), P(np.asarray(G, float).copy()))

    def entropy(self, mat, e, G):
        return self.L.ora_entropy(P(mat), _d(e), P(np.asarray(G, float).copy()))

    def temperature(self, mat, S, G):
        return self.L.ora_temperature(P(mat), _d(S), P(np.asarray(G, float).copy()))

    def stress(self, mat, ent, F):
        out = np.zeros(9)
        self.L.ora_stress(P(mat), _d(ent), P(np.asarray(F, float).copy()), P(out))
        return out

    def stress_den(self, mat, den, e_int, F):
        out = np.zeros(9)
        self.L.ora_stress_den(P(mat), _d(den), _d(e_int), P(np.asarray(F, float).copy()), P(out))
        return out

    def acoustic(self, mat, ent, F, n=(1.0, 0.0, 0.0)):
        out = np.zeros(9)
        self.L.ora_acoustic(P(mat), _d(ent), P(np.asarray(F, float).copy()),
                            P(np.asarray(n, float).copy()), P(out))
        return out.reshape(3, 3, order="F")

    def acoustic_prim(self, mat, Pv, n=(1.0, 0.0, 0.0)):
        out = np.zeros(9)
        self.L.ora_acoustic_prim(P(mat), P(np.asarray(Pv, float).copy()),
                                 P(np.asarray(n, float).copy()), P(out))
        return out.reshape(3, 3, order="F")

    def eigvals3(self, M):
        w = np.zeros(3)
        self.L.ora_eigvals3(P(np.asfortranarray(M, dtype=float).reshape(-1, order="F").copy()), P(w))
        return w

    def finger(self, A9):
        g = np.zeros(9)
        self.L.ora_finger(P(np.asarray(A9, float).copy()), P(g))
        return g

    def invariants(self, G9):
        i = np.zeros(3)
        self.L.ora_invariants(P(np.asarray(G9, float).copy()), P(i))
        return i

    def inv(self, M9):
        r = np.zeros(9)
        self.L.ora_inv(P(np.asarray(M9, float).copy()), P(r))
        return r

    def det(self, M9):
        return self.L.ora_det(P(np.asarray(M9, float).copy()))

    def det_lu(self, M9):
        return self.L.ora_det_lu(P(np.asarray(M9, float).copy()))

    # --- model level (nph = 1: 14 comps, nph = 2: 30 comps)
    @staticmethod
    def nc(nph):
        return 14 if nph == 1 else 30

    def _map(self, fn, nph, mats, X, out_w=None):
        X = np.asarray(X, float)
        vec = X.ndim == 1
        X2 = X.reshape(X.shape[0], -1, order="F")
        out_w = out_w or X2.shape[0]
        out = np.zeros((out_w, X2.shape[1]), order="F")
        for j in range(X2.shape[1]):
            col = np.ascontiguousarray(X2[:, j])
            o = np.zeros(out_w)
            fn(nph, P(mats), P(col), P(o))
            out[:, j] = o
        return out[:, 0] if vec else out

    def prim2cons(self, nph, mats, Pm):
        return self._map(self.L.ora_prim2cons, nph, mats, Pm)

    def cons2prim(self, nph, mats, Q):
        return self._map(self.L.ora_cons2prim, nph, mats, Q)

    def flux(self, nph, mats, Q):
        return self._map(self.L.ora_flux, nph, mats, Q)

    def get_eigvals(self, nph, mats, Q):
        return self._map(self.L.ora_get_eigvals, nph, mats, Q, out_w=6 * nph)

    def noncons_flux(self, mats, Q30):
        B = np.zeros(900)
        self.L.ora_noncons_flux(P(mats), P(np.ascontiguousarray(Q30, dtype=float)), P(B))
        return B.reshape(30, 30, order="F")

    def lxf(self, nph, mats, Ql, Qr, lam):
        nc = self.nc(nph)
        c, m, p = np.zeros(nc), np.zeros(nc), np.zeros(nc)
        self.L.ora_lxf(nph, P(mats), P(np.ascontiguousarray(Ql, float)),
                       P(np.ascontiguousarray(Qr, float)), _d(lam), P(c), P(m), P(p))
        return c, m, p

    def hll(self, nph, mats, Ql, Qr, eig_l, eig_r):
        nc = self.nc(nph)
        c, m, p = np.zeros(nc), np.zeros(nc), np.zeros(nc)
        self.L.ora_hll(nph, P(mats), P(np.ascontiguousarray(Ql, float)),
                       P(np.ascontiguousarray(Qr, float)), P(np.ascontiguousarray(eig_l, float)),
                       P(np.ascontiguousarray(eig_r, float)), P(c), P(m), P(p))
        return c, m, p

    def quadrature(self, which):
        x, w = np.zeros(6), np.zeros(6)
        self.L.ora_quadrature(which, P(x), P(w))
        return x, w

    def run(self, nph, mats, flux, Q0, cfl, dx, steps=None, t_end=math.inf, threads=0, t0=0.0):
        Q = np.asfortranarray(np.array(Q0, dtype=float))
        nx = Q.shape[1]
        max_steps = steps if steps is not None else 10 ** 9
        hist = np.zeros(max_steps if steps is not None else 200000)
        t = ctypes.c_double(t0)
        s = ctypes.c_int(0)
        self.L.ora_run(nph, P(mats), flux, nx, P(Q), _d(cfl), _d(dx), int(max_steps), _d(t_end),
                       ctypes.byref(t), ctypes.byref(s), P(hist), int(threads))
        return Q, t.value, s.value, hist[: s.value].copy()


    def adjoint_dual(self, mats, prims, x_ifc, nx, steps, cfl, rho_t, u_t, threads=0):
        """Misfit of the single-phase Rusanov run from two pieces (prims: 13 x 2,
        u(3), vec(A), S) and its gradient with respect to both pieces' prims and
        (b0, c0), by forward-mode duals through the whole time loop."""
        pr = np.asfortranarray(np.array(prims, dtype=float))
        rt = np.ascontiguousarray(np.array(rho_t, dtype=float))
        ut = np.ascontiguousarray(np.array(u_t, dtype=float))
        assert pr.shape == (13, 2) and rt.size == nx and ut.size == nx
        loss = ctypes.c_double()
        g = np.zeros(28)
        done = self.L.ora_adjoint_dual(P(np.ascontiguousarray(mats, dtype=float)), P(pr), _d(x_ifc),
                                       int(nx), int(steps), _d(cfl), P(rt), P(ut), int(threads),
                                       ctypes.byref(loss), P(g))
        if done != steps:
            raise RuntimeError(f"dual run stopped after {done} of {steps} steps")
        return loss.value, g


_ORA = None


def load():
    global _ORA
    if _ORA is None:
        _ORA = Oracle()
    return _ORA


def rel_err_per_component(A, B):
    """max_i |A - B| / max_i |B| per conserved component (rows)."""
    A = np.asarray(A)
    B = np.asarray(B)
    scale = np.max(np.abs(B), axis=1)
    scale = np.where(scale > 0, scale, 1.0)
    return np.max(np.abs(A - B), axis=1) / scale
