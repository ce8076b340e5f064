"""Strains: Finger tensor and its invariants (Strains.jl:26-52), on the GPU."""
import numpy as np

from ._lib import check, columns, lib, ptr


def finger(a):
    """G = vec(inv(A A^T)) for a column-major 9-vector A, or (9, n) array."""
    x, n, was_vec = columns(a, 9, "a")
    out = np.empty((9, n), order="F")
    check(lib().hes_finger(ptr(x), ptr(out), n))
    return out[:, 0].copy() if was_vec else out


def invariants(g):
    """[tr G, (tr(G)^2 - tr(G^2))/2, det G] of a column-major 9-vector G."""
    x, n, was_vec = columns(g, 9, "g")
    out = np.empty((3, n), order="F")
    check(lib().hes_invariants(ptr(x), ptr(out), n))
    return out[:, 0].copy() if was_vec else out


__all__ = ["finger", "invariants"]
