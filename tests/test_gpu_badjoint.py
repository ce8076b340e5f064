"""Batched adjoint on the GPU: a batch of independent single-phase Rusanov runs
differentiated in one pass.

Checks: every problem of a batch evolves exactly as the one-problem adjoint
evolves it (bit for bit) and gets the same loss and gradients up to the order
of its sums; the checkpoint interval, zero-tile skipping and graph replay
change no bit; problems do not interact; gradients agree with central
differences (h = 1e-6 |theta_k|) of the batched forward model.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHEAR_UP, BULK_DOWN = 1.05, 0.97   # target runs: b0 * 1.05, c0 * 0.97
NCELLS = [3, 62, 63, 200, 517, 1000]
PIECES = [2, 2, 1, 2, 3, 2]


def _eos(rng, P):
    from hyperelasticsolver_amd.EquationsOfState import Barton2009
    out = []
    for _ in range(P):
        s = rng.uniform(0.8, 1.2, 3)
        out.append(Barton2009(_rho0=8.93 * s[0], _c0=4.6 * s[1], _b0=2.1 * s[2]))
    return out


def _batch_problem(ncells, pieces, steps, seed):
    """Per problem: scaled material, seeded pieces and the target profiles of
    a run with b0 * 1.05 and c0 * 0.97."""
    from hyperelasticsolver_amd.EquationsOfState import Barton2009
    from hyperelasticsolver_amd.adjoint import BatchAdjoint, phase_prims_to_adjoint
    from hyperelasticsolver_amd.initial import random_phase_prims
    rng = np.random.default_rng(seed)
    eos = _eos(rng, len(ncells))
    prims, ifcs = [], []
    for e, k in zip(eos, pieces):
        prims.append(phase_prims_to_adjoint(random_phase_prims(rng, e, k)))
        ifcs.append(np.sort(rng.uniform(0.2, 0.8, k - 1)))
    eos_t = [Barton2009(_rho0=e.rho0, _b0=e.b0 * SHEAR_UP, _c0=e.c0 * BULK_DOWN) for e in eos]
    N = int(np.sum(ncells))
    with BatchAdjoint(eos_t, ncells, steps, every=steps) as t:
        t.set_initial(prims, ifcs)
        t.set_target(np.zeros(N), np.zeros(N))
        t.forward()
        rho_t, u_t = t.profiles()
    return eos, prims, ifcs, rho_t, u_t


def _batch_run(eos, ncells, prims, ifcs, rho_t, u_t, steps, every, sparse=True, graph=True,
               grad=True):
    from hyperelasticsolver_amd.adjoint import BatchAdjoint
    with BatchAdjoint(eos, ncells, steps, every=every) as b:
        b.set_sparse(sparse)
        b.set_launch(graph)
        b.set_initial(prims, ifcs)
        b.set_target(rho_t, u_t)
        L = b.forward()
        if not grad:
            return L
        Q = b.final_state()
        return L, Q, b.backward()


def test_batch_matches_one_problem_adjoints():
    from hyperelasticsolver_amd.adjoint import Adjoint
    print("[badjoint] batch vs one-problem adjoints", flush=True)
    steps, every = 60, 13
    eos, prims, ifcs, rho_t, u_t = _batch_problem(NCELLS, PIECES, steps, seed=31)
    L, Q, grads = _batch_run(eos, NCELLS, prims, ifcs, rho_t, u_t, steps, every)
    # a one-piece problem stays uniform, and so does its target run (rho0 and u
    # are the same): its misfit is exactly zero
    assert np.all((L > 0) == (np.array(PIECES) > 1)), L
    off = np.concatenate([[0], np.cumsum(NCELLS)])
    for p, n in enumerate(NCELLS):
        sl = slice(off[p], off[p + 1])
        with Adjoint(eos[p], n, steps, every=every) as a:
            a.set_initial(prims[p], list(ifcs[p]))
            a.set_target(rho_t[sl], u_t[sl])
            La = a.forward()
            Qa = a.final_state()
            ga, gma = a.backward()
        assert np.array_equal(Q[:, sl], Qa), f"problem {p}: forward state"
        assert L[p] == pytest.approx(La, rel=1e-12, abs=1e-300)
        gb, gmb = grads[p]
        scale = max(np.abs(ga).max(), np.abs(gma).max())
        assert np.abs(gb - ga).max() <= 1e-9 * scale, (p, np.abs(gb - ga).max(), scale)
        assert np.abs(gmb - gma).max() <= 1e-9 * scale, (p, gmb, gma)


def test_batch_of_more_tiles_than_reverse_waves():
    """More tiles than the reverse step has persistent waves (4 per CU): the
    waves take several tiles each, every later one prefetched into LDS while
    the previous one is evaluated.  Three problems' gradients agree with the
    one-problem adjoint (which gets one wave per tile here); skipping and
    graph replay change no bit."""
    from hyperelasticsolver_amd.adjoint import Adjoint
    ncells = [2000 + 37 * k for k in range(48)]
    ntiles = sum((n + 61) // 62 for n in ncells)
    assert ntiles > 4 * 256, ntiles
    print(f"[badjoint] {len(ncells)} problems, {ntiles} tiles", flush=True)
    pieces = [2] * len(ncells)
    steps, every = 40, 9
    eos, prims, ifcs, rho_t, u_t = _batch_problem(ncells, pieces, steps, seed=47)
    L, Q, grads = _batch_run(eos, ncells, prims, ifcs, rho_t, u_t, steps, every)
    L2, Q2, grads2 = _batch_run(eos, ncells, prims, ifcs, rho_t, u_t, steps, every,
                                sparse=False, graph=False)
    assert np.array_equal(L, L2) and np.array_equal(Q, Q2)
    for (g, gm), (g2, gm2) in zip(grads, grads2):
        assert np.array_equal(g, g2) and np.array_equal(gm, gm2)
    off = np.concatenate([[0], np.cumsum(ncells)])
    for p in (0, 23, 47):
        sl = slice(off[p], off[p + 1])
        with Adjoint(eos[p], ncells[p], steps, every=every) as a:
            a.set_initial(prims[p], list(ifcs[p]))
            a.set_target(rho_t[sl], u_t[sl])
            La = a.forward()
            ga, gma = a.backward()
        assert L[p] == pytest.approx(La, rel=1e-12, abs=1e-300)
        gb, gmb = grads[p]
        scale = max(np.abs(ga).max(), np.abs(gma).max())
        assert np.abs(gb - ga).max() <= 1e-9 * scale, (p, np.abs(gb - ga).max(), scale)
        assert np.abs(gmb - gma).max() <= 1e-9 * scale, (p, gmb, gma)


def test_batch_with_its_own_dx_per_problem():
    """Per-problem dx (here half and twice 1 / n_p): each problem still evolves
    and differentiates as the one-problem adjoint with that dx."""
    from hyperelasticsolver_amd.adjoint import Adjoint, BatchAdjoint
    print("[badjoint] per-problem dx", flush=True)
    steps, every = 30, 8
    ncells, pieces = [70, 200, 333], [2, 3, 2]
    eos, prims, ifcs, rho_t, u_t = _batch_problem(ncells, pieces, steps, seed=35)
    dx = np.array([0.5 / ncells[0], 2.0 / ncells[1], 1.0 / ncells[2]])
    with BatchAdjoint(eos, ncells, steps, every=every, dx=dx) as b:
        b.set_initial(prims, ifcs)
        b.set_target(rho_t, u_t)
        L = b.forward()
        Q = b.final_state()
        grads = b.backward()
    off = np.concatenate([[0], np.cumsum(ncells)])
    for p, n in enumerate(ncells):
        sl = slice(off[p], off[p + 1])
        with Adjoint(eos[p], n, steps, every=every, dx=dx[p]) as a:
            a.set_initial(prims[p], list(ifcs[p]))
            a.set_target(rho_t[sl], u_t[sl])
            La = a.forward()
            Qa = a.final_state()
            ga, gma = a.backward()
        assert np.array_equal(Q[:, sl], Qa), f"problem {p}: forward state"
        assert L[p] == pytest.approx(La, rel=1e-12, abs=1e-300)
        scale = max(np.abs(ga).max(), np.abs(gma).max())
        assert np.abs(grads[p][0] - ga).max() <= 1e-9 * scale
        assert np.abs(grads[p][1] - gma).max() <= 1e-9 * scale


def test_batch_checkpoints_skipping_and_launch_change_nothing():
    print("[badjoint] checkpoint interval, skipping and launch invariance", flush=True)
    steps = 45
    eos, prims, ifcs, rho_t, u_t = _batch_problem(NCELLS, PIECES, steps, seed=32)
    ref = None
    for every, sparse, graph in ((1, True, True), (7, True, True), (45, False, True),
                                 (45, True, False), (20, False, False)):
        L, Q, grads = _batch_run(eos, NCELLS, prims, ifcs, rho_t, u_t, steps, every, sparse, graph)
        flat = np.concatenate([np.concatenate([g.ravel(order="F"), gm]) for g, gm in grads])
        assert np.all(np.isfinite(flat))
        if ref is None:
            ref = (L, Q, flat)
        else:
            assert np.array_equal(L, ref[0]) and np.array_equal(Q, ref[1])
            assert np.array_equal(flat, ref[2]), (every, sparse, graph)


def test_batch_problems_do_not_interact():
    from hyperelasticsolver_amd.EquationsOfState import Barton2009
    print("[badjoint] problems do not interact", flush=True)
    steps = 40
    eos, prims, ifcs, rho_t, u_t = _batch_problem(NCELLS, PIECES, steps, seed=33)
    L0, Q0, g0 = _batch_run(eos, NCELLS, prims, ifcs, rho_t, u_t, steps, 11)
    eos2 = list(eos)
    eos2[3] = Barton2009(_rho0=eos[3].rho0, _b0=eos[3].b0 * 1.01, _c0=eos[3].c0)
    L1, Q1, g1 = _batch_run(eos2, NCELLS, prims, ifcs, rho_t, u_t, steps, 11)
    assert L1[3] != L0[3]
    for p in range(len(NCELLS)):
        if p == 3:
            continue
        assert L1[p] == L0[p]
        assert np.array_equal(g1[p][0], g0[p][0]) and np.array_equal(g1[p][1], g0[p][1])


def test_batch_gradient_matches_central_differences():
    """b0, c0 and entries of u, A and S of two problems against central
    differences of their own losses in the batched forward model."""
    from hyperelasticsolver_amd.EquationsOfState import Barton2009
    print("[badjoint] gradient vs central differences", flush=True)
    steps = 120
    eos, prims, ifcs, rho_t, u_t = _batch_problem(NCELLS, PIECES, steps, seed=34)
    L, _, grads = _batch_run(eos, NCELLS, prims, ifcs, rho_t, u_t, steps, 25)

    def loss_with(p, b0=None, c0=None, prim=None):
        e2 = list(eos)
        if b0 is not None or c0 is not None:
            e = eos[p]
            e2[p] = Barton2009(_rho0=e.rho0, _b0=e.b0 if b0 is None else b0,
                               _c0=e.c0 if c0 is None else c0)
        pr = list(prims)
        if prim is not None:
            pr[p] = prim
        return _batch_run(e2, NCELLS, pr, ifcs, rho_t, u_t, steps, steps, grad=False)[p]

    for p in (3, 4):
        g, gm = grads[p]
        scale = max(np.abs(g).max(), np.abs(gm).max())
        for j, which in enumerate(("b0", "c0")):
            v = getattr(eos[p], which)
            h = 1e-6 * v
            fd = (loss_with(p, **{which: v + h}) - loss_with(p, **{which: v - h})) / (2 * h)
            assert abs(gm[j] - fd) <= 1e-6 * max(abs(fd), abs(gm[j])) + 1e-9 * scale, (p, which, gm[j], fd)
        for k, piece in ((0, 0), (3, 1), (12, 0)):   # u1, A11, S
            th = prims[p][k, piece]
            h = 1e-6 * abs(th)
            pp, pm = prims[p].copy(), prims[p].copy()
            pp[k, piece] += h
            pm[k, piece] -= h
            fd = (loss_with(p, prim=pp) - loss_with(p, prim=pm)) / (2 * h)
            assert abs(g[k, piece] - fd) <= 1e-6 * max(abs(fd), abs(g[k, piece])) + 1e-9 * scale, \
                (p, k, piece, g[k, piece], fd)
