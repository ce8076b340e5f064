"""Initial states: the reference's Riemann test cases and seeded random
piecewise-constant problems.

Primitive layouts (HyperelasticityMPh.jl:126-130):
  one phase  : [rho, u1, u2, u3, S, vec(A)]            (14)
  two phases : [a, rho, u1, u2, u3, S, vec(A)] x 2     (30)
vec(A) is column-major.  Conversion to conservative variables runs on the GPU
(`prim2cons`); the piecewise state is assembled on the GPU by
`Solver.set_piecewise`.
"""
import numpy as np

from . import _lib
from ._lib import check, columns, lib, ptr
from .solver import materials_array, nphase_of


def prim2cons(eos, P):
    """prim2cons_mph for a tuple of two EoS, or the one-phase block for one."""
    nph = nphase_of(eos)
    nc = 14 if nph == 1 else 30
    p, n, was_vec = columns(P, nc, "P")
    mats = materials_array(eos if nph == 2 else (eos[0] if isinstance(eos, (tuple, list)) else eos))
    out = np.empty((nc, n), order="F")
    check(lib().hes_prim2cons(nph, ptr(mats), ptr(p), ptr(out), n))
    return out[:, 0].copy() if was_vec else out


def random_phase_prims(rng, eos, k, a_scale=0.05, u_half=0.5, entropy=1e-3):
    """k random one-phase primitive states: A = I + a_scale N(0,1) per entry,
    u ~ U(-u_half, u_half)^3, S = entropy, rho = rho0 / det A."""
    P = np.zeros((14, k), order="F")
    for j in range(k):
        A = np.eye(3) + a_scale * rng.standard_normal((3, 3))
        while np.linalg.det(A) <= 0.5:
            A = np.eye(3) + a_scale * rng.standard_normal((3, 3))
        u = rng.uniform(-u_half, u_half, 3)
        P[0, j] = eos.rho0 / np.linalg.det(A)
        P[1:4, j] = u
        P[4, j] = entropy
        P[5:14, j] = A.reshape(-1, order="F")
    return P


def random_two_phase_prims(rng, eos2, k, alpha_lo=0.01, alpha_hi=0.99, **kw):
    """k random two-phase primitive states, a1 ~ U(alpha_lo, alpha_hi), a2 = 1 - a1."""
    P = np.zeros((30, k), order="F")
    for j in range(k):
        a1 = rng.uniform(alpha_lo, alpha_hi)
        for p, a in enumerate((a1, 1.0 - a1)):
            ph = random_phase_prims(rng, eos2[p], 1, **kw)[:, 0]
            P[15 * p, j] = a
            P[15 * p + 1:15 * p + 15, j] = ph
    return P


def random_interfaces(rng, k):
    """k sorted interface positions in (0, 1)."""
    return np.sort(rng.uniform(0.0, 1.0, k))


def piecewise_cells(nx, states, interfaces):
    """Host-side expansion of a piecewise-constant state (ncomp x nx), same rule
    as Solver.set_piecewise: cell i takes state #{j : (i+1/2)/nx > s_j}."""
    st = np.asarray(states)
    x = (np.arange(nx) + 0.5) / nx
    idx = np.searchsorted(np.asarray(interfaces, dtype=np.float64), x, side="left")
    return np.asfortranarray(st[:, idx])


def initial_condition(Ql, Qr, nx):
    """main.jl:99-106: cell i (1-based) is Ql iff (i-1) < nx/2."""
    Ql = np.asarray(Ql, dtype=np.float64)
    Qr = np.asarray(Qr, dtype=np.float64)
    Q = np.empty((Ql.shape[0], nx), order="F")
    for i in range(nx):
        Q[:, i] = Ql if i < nx / 2 else Qr
    return Q


def initial_condition_tanh(eos, Ql, Qr, nx, eps):
    """main.jl:110-123: volume fraction a1 = 0.1 (tanh(4(x-1/2)/eps) + 1) + 0.4."""
    from .HyperelasticityMPh import cons2prim_mph, prim2cons_mph
    Pl = cons2prim_mph(eos, Ql)
    Pr = cons2prim_mph(eos, Qr)
    P = np.empty((30, nx), order="F")
    for i in range(1, nx + 1):
        x = (i - 0.5) / nx
        Pi = (Pl if x < 0.5 else Pr).copy()
        Pi[0] = 0.2 / 2 * (np.tanh(4 * (x - 0.5) / eps) + 1) + 0.4
        Pi[15] = 1 - Pi[0]
        P[:, i - 1] = Pi
    return prim2cons_mph(eos, P)
