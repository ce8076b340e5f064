"""Batched solver on the GPU: every problem of a batch must evolve exactly as
it does alone (same kernels, same tiling), and match the oracle to 1e-10."""
import math

import numpy as np
import pytest

from oracle_bind import rel_err_per_component

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _eos_list(rng, P, nph=1):
    from hyperelasticsolver_amd.EquationsOfState import Barton2009
    out = []
    for _ in range(P):
        s = rng.uniform(0.8, 1.2, 3)
        e = Barton2009(_rho0=8.93 * s[0], _c0=4.6 * s[1], _b0=2.1 * s[2])
        if nph == 1:
            out.append(e)
        else:
            s2 = rng.uniform(0.8, 1.2, 3)
            out.append((e, Barton2009(_rho0=2.7 * s2[0], _c0=4.6 * s2[1], _b0=3.1 * s2[2])))
    return out


def _problem_states(ora, rng, eos, nph):
    from hyperelasticsolver_amd.solver import materials_array
    k = int(rng.integers(1, 5))
    mats = materials_array(eos)
    P = np.zeros((14 if nph == 1 else 30, k), order="F")
    for j in range(k):
        for ph in range(nph):
            A = np.eye(3) + 0.05 * rng.standard_normal((3, 3))
            m = mats[8 * ph:8 * ph + 8]
            blk = [m[0] / np.linalg.det(A), *rng.uniform(-0.5, 0.5, 3), 1e-3, *A.reshape(-1, order="F")]
            if nph == 1:
                P[:, j] = blk
            else:
                a = rng.uniform(0.01, 0.99) if ph == 0 else 1.0 - P[0, j]
                P[15 * ph:15 * ph + 15, j] = [a] + blk
    Qs = ora.prim2cons(nph, mats, P)
    ifc = np.sort(rng.uniform(0.05, 0.95, k - 1))
    return Qs, ifc


def _build(ora, ncells, nph=1, flux="rusanov", seed=0):
    from hyperelasticsolver_amd.batched import Batch
    rng = np.random.default_rng(seed)
    eos = _eos_list(rng, len(ncells), nph)
    sts, ifs = [], []
    for e in eos:
        Qs, ifc = _problem_states(ora, rng, e, nph)
        sts.append(Qs)
        ifs.append(ifc)
    b = Batch(eos, ncells, flux=flux)
    b.set_piecewise(sts, ifs)
    return b, eos, sts, ifs


NCELLS = [3, 61, 62, 63, 64, 124, 125, 200, 517]


@pytest.mark.parametrize("flux", ["rusanov", "hll"])
def test_batch_equals_single_problem_runs(ora, flux):
    from hyperelasticsolver_amd.solver import Solver
    print(f"[batch] vs single runs, flux={flux}", flush=True)
    steps = 15
    b, eos, sts, ifs = _build(ora, NCELLS, flux=flux, seed=1)
    Q0 = b.get_state()
    b.run(steps, mode="graph")
    Qb = b.get_state()
    tb, sb = b.times()
    b.close()
    for p, n in enumerate(NCELLS):
        with Solver(eos[p], n, flux=flux, cfl=0.6) as s:
            s.set_state(Q0[:, b.offsets[p]:b.offsets[p + 1]])
            r = s.run(steps=steps, use_graph=False)
            Qs = s.get_state()
        assert sb[p] == steps == r["steps"]
        assert tb[p] == r["t"]
        assert np.array_equal(Qb[:, b.offsets[p]:b.offsets[p + 1]], Qs), f"problem {p}"


@pytest.mark.parametrize("flux,fk", [("rusanov", 0), ("hll", 1)])
def test_batch_over_two_to_the_21_cells(ora, flux, fk):
    """A whole-batch launch over 2^21 cells or more takes the pinned schedule
    (four waves per SIMD); one-problem launches (mode "loop") take the other.
    Both must give the same bits, and the oracle the same states to 1e-10."""
    from hyperelasticsolver_amd.solver import materials_array
    rng = np.random.default_rng(21)
    ncells = [int(n) for n in rng.integers(40000, 52000, 48)]
    assert sum((n + 61) // 62 for n in ncells) * 62 >= 1 << 21
    steps = 8
    print(f"[batch] over 2^21 cells ({sum(ncells)}), flux={flux}", flush=True)
    res = []
    for mode, tpw in (("graph", 1), ("loop", 1), ("graph", 3)):
        b, eos, sts, ifs = _build(ora, ncells, flux=flux, seed=4)
        b.set_launch(256, True, tpw)
        Q0 = b.get_state()
        b.run(steps, mode=mode)
        res.append((b.get_state(), b.times()))
        offs = list(b.offsets)
        b.close()
    for r in res[1:]:
        assert np.array_equal(res[0][0], r[0])
        assert np.array_equal(res[0][1][0], r[1][0])
    for p in (0, 17, 47):
        n = ncells[p]
        Qo, _, so, _ = ora.run(1, materials_array(eos[p]), fk, Q0[:, offs[p]:offs[p + 1]], 0.6,
                               1.0 / n, steps=steps)
        err = rel_err_per_component(res[0][0][:, offs[p]:offs[p + 1]], Qo)
        assert so == steps and np.all(err <= TOL), (p, err.max())


def test_large_problems_in_a_batch_equal_single_problem_runs(ora):
    """Problems above 65536 cells max their CFL speeds into the slot only when
    the slot holds a smaller one; smaller problems always issue the atomic.
    Problems on both sides of that switch must step exactly as they do alone."""
    from hyperelasticsolver_amd.solver import Solver
    ncells = [300001, 70, 65537, 65536, 65535]
    steps = 6
    print(f"[batch] large problems {ncells}", flush=True)
    b, eos, sts, ifs = _build(ora, ncells, seed=12)
    Q0 = b.get_state()
    b.run(steps, mode="graph")
    Qb = b.get_state()
    tb, sb = b.times()
    offs = list(b.offsets)
    b.close()
    for p, n in enumerate(ncells):
        with Solver(eos[p], n, flux="rusanov", cfl=0.6) as s:
            s.set_state(Q0[:, offs[p]:offs[p + 1]])
            r = s.run(steps=steps, use_graph=False)
            Qs = s.get_state()
        assert sb[p] == steps == r["steps"]
        assert tb[p] == r["t"], p
        assert np.array_equal(Qb[:, offs[p]:offs[p + 1]], Qs), f"problem {p}"


def test_batch_matches_oracle(ora):
    from hyperelasticsolver_amd.solver import materials_array
    print("[batch] vs oracle", flush=True)
    ncells = [64, 150, 333]
    steps = 12
    for flux, fk in (("rusanov", 0), ("hll", 1), ("lxf", 2)):
        b, eos, sts, ifs = _build(ora, ncells, flux=flux, seed=2)
        Q0 = b.get_state()
        b.run(steps, mode="eager")
        Qb = b.get_state()
        b.close()
        for p, n in enumerate(ncells):
            Qo, to, so, _ = ora.run(1, materials_array(eos[p]), fk,
                                    Q0[:, b.offsets[p]:b.offsets[p + 1]], 0.6, 1.0 / n, steps=steps)
            err = rel_err_per_component(Qb[:, b.offsets[p]:b.offsets[p + 1]], Qo)
            assert np.all(err <= TOL), (flux, p, err.max())


@pytest.mark.parametrize("flux,fk", [("hll", 1), ("lxf", 2)])
def test_batch_two_phase_matches_oracle(ora, flux, fk):
    from hyperelasticsolver_amd.solver import materials_array
    print(f"[batch] two-phase vs oracle, flux={flux}", flush=True)
    ncells = [40, 70]
    b, eos, sts, ifs = _build(ora, ncells, nph=2, flux=flux, seed=3)
    Q0 = b.get_state()
    b.run(5, mode="graph")
    Qb = b.get_state()
    b.close()
    for p, n in enumerate(ncells):
        Qo, _, _, _ = ora.run(2, materials_array(eos[p]), fk, Q0[:, b.offsets[p]:b.offsets[p + 1]],
                              0.6, 1.0 / n, steps=5)
        err = rel_err_per_component(Qb[:, b.offsets[p]:b.offsets[p + 1]], Qo)
        assert np.all(err <= TOL), (p, err.max())


@pytest.mark.parametrize("flux", ["rusanov", "hll", "lxf"])
def test_two_phase_batch_equals_single_problem_runs(ora, flux):
    """Two-phase problems take the jump-face scan and path-term prepass per
    launch, with each face's own materials.  Every launch mode and shape gives
    the same bits; every problem takes the one-problem solver's steps and
    times exactly, also when the problems stop at different end times, and its
    state equals the solver's: bit for bit for Rusanov and Lax-Friedrichs, to
    rounding (1e-13 relative) for HLL, whose path code the two kernels may
    contract into fused multiply-adds differently."""
    from hyperelasticsolver_amd.solver import Solver
    print(f"[batch] two-phase vs single runs, flux={flux}", flush=True)
    ncells = [3, 61, 62, 63, 64, 130, 300]
    steps = 9
    t_end = [np.inf, np.inf, 2e-4, np.inf, 5e-4, np.inf, 3e-4]
    res = []
    for mode, shape in (("graph", (256, True, 1)), ("eager", (64, False, 3)), ("loop", (64, True, 1))):
        b, eos, sts, ifs = _build(ora, ncells, nph=2, flux=flux, seed=9)
        b.set_launch(*shape)
        b.set_t_end(t_end)
        Q0 = b.get_state()
        b.run(steps, mode=mode)
        res.append((b.get_state(), b.times()))
        offs = list(b.offsets)
        b.close()
    for r in res[1:]:
        assert np.array_equal(r[0], res[0][0]) and np.array_equal(r[1][1], res[0][1][1])
    Qb, (tb, sb) = res[0]
    assert any(s < steps for s in sb) and any(s == steps for s in sb)
    for p, n in enumerate(ncells):
        with Solver(eos[p], n, flux=flux, cfl=0.6) as s:
            s.set_state(Q0[:, offs[p]:offs[p + 1]])
            r = s.run(steps=steps, t_end=t_end[p], use_graph=False)
            Qs = s.get_state()
        assert sb[p] == r["steps"] and tb[p] == r["t"], (p, sb[p], r["steps"])
        err = rel_err_per_component(Qb[:, offs[p]:offs[p + 1]], Qs)
        print(f"  problem {p}: max rel diff to the solver {err.max():.2e}", flush=True)
        if flux == "hll":
            assert np.all(err <= 1e-13), (p, err.max())
        else:
            assert np.array_equal(Qb[:, offs[p]:offs[p + 1]], Qs), f"problem {p}"


def test_batch_modes_agree_and_autotune_restores(ora):
    print("[batch] modes + autotune", flush=True)
    ncells = [70, 5, 300, 1000, 62]
    res = []
    for mode in ("eager", "graph", "loop"):
        b, *_ = _build(ora, ncells, seed=4)
        if mode == "graph":
            tuned = b.autotune(steps=2)
            assert tuned["block"] in (64, 256) and tuned["tiles_per_wave"] in (1, 2, 4)
        b.run(9, mode=mode)
        res.append((b.get_state(), b.times()))
        b.close()
    for Q, (t, s) in res[1:]:
        assert np.array_equal(Q, res[0][0])
        assert np.array_equal(t, res[0][1][0])
        assert np.array_equal(s, res[0][1][1])


def test_every_launch_shape_gives_the_same_bits(ora):
    """Block size, XCD block order and tiles per wave (each wave stepping
    through its tiles in turn) change only the schedule: every shape gives the
    bits of the default one, eager and graph-replayed."""
    print("[batch] launch shapes", flush=True)
    ncells = [70, 5, 300, 1000, 62, 2500, 63, 64]
    ref = None
    for block, xcd, tpw, mode in ((256, True, 1, "eager"), (64, False, 1, "graph"),
                                  (64, True, 2, "eager"), (256, False, 3, "graph"),
                                  (256, True, 4, "graph"), (64, True, 7, "eager"),
                                  (64, False, 64, "graph")):
        b, *_ = _build(ora, ncells, seed=8)
        b.set_launch(block, xcd, tpw)
        assert b.info()["tiles_per_wave"] == tpw
        b.run(11, mode=mode)
        out = (b.get_state(), b.times())
        b.close()
        if ref is None:
            ref = out
        else:
            assert np.array_equal(out[0], ref[0]), (block, xcd, tpw, mode)
            assert np.array_equal(out[1][0], ref[1][0]) and np.array_equal(out[1][1], ref[1][1])


def test_batch_per_problem_t_end(ora):
    print("[batch] per-problem t_end", flush=True)
    ncells = [100, 200]
    b, eos, *_ = _build(ora, ncells, seed=5)
    b.set_t_end([0.002, 0.004])
    b.run(400, mode="graph")
    t, s = b.times()
    b.close()
    assert t[0] >= 0.002 and t[1] >= 0.004
    assert s[0] < s[1] < 400


def test_autotune_on_reused_memory_matches_clean_run(ora):
    """Device memory full of NaNs from an earlier allocation must never leak in:
    the autotuner's save/restore copies are ordered with its kernels.  The
    poisoning goes through this library's own allocations (a second HIP
    runtime in the process, such as torch's, may not see the device once this
    library's runtime holds it)."""
    from hyperelasticsolver_amd.EquationsOfState import Barton2009
    from hyperelasticsolver_amd._lib import DomainError
    from hyperelasticsolver_amd.batched import Batch
    print("[batch] autotune on NaN-filled reused memory", flush=True)
    ncells = [300, 4000, 777, 5000] * 4
    b, *_ = _build(ora, ncells, seed=6)
    b.run(6, mode="eager")
    clean = b.get_state()
    b.close()
    for n in (ncells, [sum(ncells) * 8]):
        p = Batch([Barton2009()] * len(n), n)
        with pytest.raises(DomainError):  # the state is stored, then rejected
            p.set_state(np.full((14, sum(n)), np.nan, order="F"))
        p.close()
    b, *_ = _build(ora, ncells, seed=6)
    b.autotune(steps=3)
    b.set_launch(256, True)
    b.run(6, mode="eager")
    assert np.array_equal(b.get_state(), clean)
    b.close()


def test_batch_general_exponents_match_oracle(ora):
    """A batch with one problem of general exponents (main.jl:133 constants)
    runs the general-exponent kernels for all its problems."""
    from hyperelasticsolver_amd.EquationsOfState import Barton2009
    from hyperelasticsolver_amd.batched import Batch
    from hyperelasticsolver_amd.solver import materials_array
    print("[batch] general exponents vs oracle", flush=True)
    rng = np.random.default_rng(21)
    eos = [Barton2009(), Barton2009(_c0=6.22, _cv=9.0e-4, _b0=3.16, _beta=3.577, _gamma=2.088),
           Barton2009(_rho0=2.7, _b0=3.1)]
    ncells = [90, 130, 75]
    sts, ifs = [], []
    for e in eos:
        Qs, ifc = _problem_states(ora, rng, e, 1)
        sts.append(Qs)
        ifs.append(ifc)
    with Batch(eos, ncells, flux="hll") as b:
        b.set_piecewise(sts, ifs)
        Q0 = b.get_state()
        b.run(10, mode="eager")
        Qb = b.get_state()
        offs = b.offsets
    for p, n in enumerate(ncells):
        Qo, _, _, _ = ora.run(1, materials_array(eos[p]), 1, Q0[:, offs[p]:offs[p + 1]], 0.6,
                              1.0 / n, steps=10)
        err = rel_err_per_component(Qb[:, offs[p]:offs[p + 1]], Qo)
        assert np.all(err <= TOL), (p, err.max())


def _pieces(ora, rng, eos, nph, k):
    """k distinct conservative states (every piece of a problem differs)."""
    from hyperelasticsolver_amd.solver import materials_array
    mats = materials_array(eos)
    P = np.zeros((14 if nph == 1 else 30, k), order="F")
    for j in range(k):
        for ph in range(nph):
            A = np.eye(3) + 0.05 * rng.standard_normal((3, 3))
            m = mats[8 * ph:8 * ph + 8]
            blk = [m[0] / np.linalg.det(A), *rng.uniform(-0.5, 0.5, 3), 1e-3, *A.reshape(-1, order="F")]
            if nph == 1:
                P[:, j] = blk
            else:
                a = rng.uniform(0.01, 0.99) if ph == 0 else 1.0 - P[0, j]
                P[15 * ph:15 * ph + 15, j] = [a] + blk
    return ora.prim2cons(nph, mats, P), np.sort(rng.uniform(0.1, 0.9, k - 1))


@pytest.mark.parametrize("nph,flux,mode", [(1, "rusanov", "graph"), (1, "hll", "eager"),
                                           (1, "rusanov", "loop"), (2, "rusanov", "graph")])
def test_stopped_problems_keep_their_state_and_resume(ora, nph, flux, mode):
    """A problem that reaches its end time inside a run stops stepping while the
    batch goes on.  Problems that stop after 1, 2 and 4 of their own steps in a
    7-step run (an odd number of launches) must come back with the state,
    time and step count of a one-problem run to the same end time, bit for
    bit; with the end times lifted, a further 4 launches must continue them
    as a one-problem run continues (its CFL maximum carried over the stopped
    launches)."""
    from hyperelasticsolver_amd.batched import Batch
    from hyperelasticsolver_amd.solver import Solver
    print(f"[batch] stopped problems resume, nph={nph} flux={flux} mode={mode}", flush=True)
    rng = np.random.default_rng(77)
    ncells = [150, 64, 300, 97]
    stop_after = [2, 1, 4, None]
    eos = _eos_list(rng, len(ncells), nph)
    sts, ifs = zip(*[_pieces(ora, rng, e, nph, 3) for e in eos])
    # the end time of a problem that stops after k steps: between its times
    # after k - 1 and k steps (`while t < T` then takes exactly k steps)
    t_end = []
    for p, n in enumerate(ncells):
        if stop_after[p] is None:
            t_end.append(np.inf)
            continue
        with Solver(eos[p], n, flux=flux, cfl=0.6) as s:
            s.set_piecewise(sts[p], ifs[p])
            dts = s.run(steps=stop_after[p], record_dt=True, use_graph=False)["dt"]
        tk = np.concatenate([[0.0], np.cumsum(dts)])
        t_end.append(0.5 * (tk[-2] + tk[-1]))
    with Batch(eos, ncells, flux=flux) as b:
        b.set_piecewise(sts, ifs)
        Q0 = b.get_state()
        b.set_t_end(t_end)
        b.run(7, mode=mode)
        Q1, (t1, s1) = b.get_state(), b.times()
        b.set_t_end(None)
        b.run(4, mode=mode)
        Q2, (t2, s2) = b.get_state(), b.times()
        offs = list(b.offsets)
    for p, n in enumerate(ncells):
        k = 7 if stop_after[p] is None else stop_after[p]
        with Solver(eos[p], n, flux=flux, cfl=0.6) as s:
            s.set_state(Q0[:, offs[p]:offs[p + 1]])
            r1 = s.run(steps=7, t_end=t_end[p], use_graph=False)
            S1 = s.get_state()
            r2 = s.run(steps=4, use_graph=False)
            S2 = s.get_state()
        assert r1["steps"] == s1[p] == k and t1[p] == r1["t"], (p, r1, s1[p], t1[p])
        assert np.array_equal(Q1[:, offs[p]:offs[p + 1]], S1), f"problem {p} after the stop"
        assert r2["steps"] == 4 and s2[p] == k + 4 and t2[p] == r2["t"], (p, r2, s2[p], t2[p])
        assert np.array_equal(Q2[:, offs[p]:offs[p + 1]], S2), f"problem {p} after resuming"
