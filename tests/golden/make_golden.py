"""Writes the golden fixtures under tests/golden/ with the oracle.

The reference publishes no vectors of its own (it has no tests), so these are
the oracle's outputs on seeded inputs of the benchmark's configurations:
  config1_rusanov.npz : 1000-cell single-phase Riemann problem, Rusanov,
                        CFL 0.6, t_end 0.06 (interface at x = 1/2, main.jl:103)
  config1_hll.npz     : the same with HLL
  testcase6_hll.npz   : HyperelasticityMPh test case 6 (HyperelasticityMPh.jl:348-367),
                        200 cells, HLL, 30 steps (main.jl's own case at small size)
Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_bind  # noqa: E402

MAT = np.array([8.93, 4.6, 3.9e-4, 300.0, 2.1, 1.0, 3.0, 2.0])


def config1_states(ora, seed=2024):
    """Two single-phase states: A = I + 0.05 N(0,1), u ~ U(-1/2,1/2)^3, S = 1e-3, rho = rho0/det A."""
    rng = np.random.default_rng(seed)
    P = np.zeros((14, 2), order="F")
    for j in range(2):
        A = np.eye(3) + 0.05 * rng.standard_normal((3, 3))
        P[0, j] = MAT[0] / np.linalg.det(A)
        P[1:4, j] = rng.uniform(-0.5, 0.5, 3)
        P[4, j] = 1e-3
        P[5:, j] = A.reshape(-1, order="F")
    return P, ora.prim2cons(1, MAT, P)


def testcase6_prims():
    """Left/right primitives of HyperelasticityMPh test case 6 (den_k = 8.9 / det F)."""
    Fl = np.array([[0.98, 0.0, 0.0], [0.02, 1.0, 0.1], [0.0, 0.0, 1.0]])
    Fr = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.1], [0.0, 0.0, 1.0]])
    def blk(a, F, u, S):
        return [a, 8.9 / np.linalg.det(F), *u, S, *F.reshape(-1, order="F")]
    Pl = np.array(blk(0.1, Fl, (0.0, 0.5, 1.0), 1e-3) + blk(0.9, Fl, (0.0, 0.5, 1.0), 1e-3))
    Pr = np.array(blk(0.9, Fr, (0.0, 0.0, 0.0), 0.0) + blk(0.1, Fr, (0.0, 0.0, 0.0), 0.0))
    return Pl, Pr


def main():
    ora = oracle_bind.load()
    nx = 1000
    P, Qs = config1_states(ora)
    Q0 = np.asfortranarray(np.stack([Qs[:, 0] if i < nx / 2 else Qs[:, 1] for i in range(nx)], axis=1))
    for name, fk in (("config1_rusanov", 0), ("config1_hll", 1)):
        Q, t, s, dt = ora.run(1, MAT, fk, Q0, 0.6, 1.0 / nx, t_end=0.06, threads=8)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), mat=MAT, prims=P, Q0=Q0, Q=Q, t=t,
                            steps=s, dt=dt)
        print(name, "steps", s, "t", t)
    mats2 = np.concatenate([MAT, MAT])  # main.jl:134 eos = (Barton2009(), Barton2009())
    Pl, Pr = testcase6_prims()
    Ql, Qr = ora.prim2cons(2, mats2, Pl), ora.prim2cons(2, mats2, Pr)
    nx2 = 200
    Q02 = np.asfortranarray(np.stack([Ql if i < nx2 / 2 else Qr for i in range(nx2)], axis=1))
    Q, t, s, dt = ora.run(2, mats2, 1, Q02, 0.6, 1.0 / nx2, steps=30, threads=8)
    np.savez_compressed(os.path.join(HERE, "testcase6_hll.npz"), mat=mats2, Pl=Pl, Pr=Pr, Q0=Q02,
                        Q=Q, t=t, steps=s, dt=dt)
    print("testcase6_hll steps", s, "t", t)


if __name__ == "__main__":
    main()
