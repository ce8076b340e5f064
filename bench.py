#!/usr/bin/env python3
"""Benchmark of the FP64 finite-volume time step on one MI355X.

Workload (the headline configuration): single-phase 1D, 2^22 cells on [0,1],
65 seeded piecewise-constant states (A = I + 0.05 N(0,1), u ~ U(-1/2,1/2)^3,
S = 1e-3, rho = rho0/det A) separated by 64 seeded interfaces, Barton2009
copper constants, CFL 0.6, Rusanov flux.  A "step" is one time step over all
cells (one launch of the fused step kernel).  The state is resident in HBM
before the timed region; HIP events on the solver's stream bracket exactly
the K timed steps (replayed from one graph captured beforehand).

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 cell-updates/s on 1 GPU (% HBM peak); batched vs plain; fwd+adj/fwd"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, 8 TB/s
COUNTERS = os.path.join(ROOT, "profiles", "k_step_2p22_counters.json")
BYTES_PER_CELL_1PH = 2 * 14 * 8  # read + write the 14-component FP64 state


def headline_problem(nx, seed=2):
    from hyperelasticsolver_amd.EquationsOfState import Barton2009
    from hyperelasticsolver_amd.initial import prim2cons, random_interfaces, random_phase_prims
    eos = Barton2009()
    rng = np.random.default_rng(seed)
    P = random_phase_prims(rng, eos, 65)
    Qs = prim2cons(eos, P)
    ifc = random_interfaces(rng, 64)
    return eos, Qs, ifc


def time_solver(eos, Qs, ifc, nx, flux, steps, warmup, sources=None):
    """Milliseconds of `steps` steps; sources = (tau_velocity, tau_pressure)
    switches on the two-phase relaxation sources."""
    from hyperelasticsolver_amd.solver import Solver
    s = Solver(eos, nx, flux=flux, cfl=0.6)
    if sources is not None:
        s.set_relaxation(sources[0])
        s.set_pressure_relaxation(sources[1])
    s.set_piecewise(Qs, ifc)
    if warmup:
        s.time_steps(warmup)
    ms = s.time_steps(steps)
    Q = s.get_state()
    s.close()
    if not np.all(np.isfinite(Q)):
        raise RuntimeError("non-finite state after the timed steps")
    return ms


def host_transfer(eos, Qs, ifc, nx):
    """Upload and download of the headline state between host memory (Julia's
    ncomp x n column-major layout) and the device layout: what a caller that
    hands over host arrays pays once per run, outside `value`."""
    from hyperelasticsolver_amd.solver import Solver
    x = (np.arange(nx) + 0.5) / nx
    Q0 = np.asfortranarray(Qs[:, np.searchsorted(ifc, x, side="left")])
    up, down = [], []
    with Solver(eos, nx, flux="rusanov", cfl=0.6) as s:
        for _ in range(3):
            t0 = time.perf_counter()
            s.set_state(Q0)
            t1 = time.perf_counter()
            Q = s.get_state()
            t2 = time.perf_counter()
            up.append(t1 - t0)
            down.append(t2 - t1)
    if not np.array_equal(Q, Q0):
        raise RuntimeError("state changed in an upload/download round trip")
    nbytes = Q0.nbytes
    return {"state_bytes": nbytes, "upload_ms": min(up) * 1e3, "download_ms": min(down) * 1e3,
            "upload_gbs": nbytes / min(up) / 1e9, "download_gbs": nbytes / min(down) / 1e9}


def cpu_baseline(eos, Qs, ifc, flux_kind, threads):
    """The oracle (C++ restatement of the reference algorithm) on host cores,
    on a bounded sample of the same workload: 2^17 cells, 64 steps (about
    10 s on 16 host cores)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind
    ora = oracle_bind.load()
    nx = 1 << 17
    steps = 64
    x = (np.arange(nx) + 0.5) / nx
    Q0 = np.asfortranarray(Qs[:, np.searchsorted(ifc, x, side="left")])
    t0 = time.perf_counter()
    _, _, s, _ = ora.run(1, eos.params(), flux_kind, Q0, 0.6, 1.0 / nx, steps=steps, threads=threads)
    el = time.perf_counter() - t0
    return {"value": nx * s / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(),
            "sample": f"{nx} cells x {s} steps of the same workload ({'Rusanov' if flux_kind == 0 else 'HLL'}),"
                      f" OpenMP over cells on {threads} threads, {el:.1f} s wall"}


BYTES_PER_CELL_2PH = 2 * 30 * 8  # read + write the 30-component FP64 state


def two_phase_problem(nx, seed=3):
    """Two-phase workload: 4 seeded pieces split by 3 interfaces, a1 ~ U(0.01,
    0.99), a2 = 1 - a1, per-phase A = I + 0.05 N(0,1), u ~ U(-1/2,1/2)^3,
    S = 1e-3, materials Barton2009(_rho0=8.93, _b0=2.1) / (_rho0=2.7, _b0=3.1)."""
    from hyperelasticsolver_amd.EquationsOfState import Barton2009
    from hyperelasticsolver_amd.HyperelasticityMPh import prim2cons_mph
    from hyperelasticsolver_amd.initial import random_interfaces, random_phase_prims
    rng = np.random.default_rng(seed)
    eos = (Barton2009(_rho0=8.93, _b0=2.1), Barton2009(_rho0=2.7, _b0=3.1))
    k = 4
    P = np.zeros((30, k), order="F")
    for j in range(k):
        a = rng.uniform(0.01, 0.99)
        for p, al in enumerate((a, 1.0 - a)):
            P[15 * p, j] = al
            P[15 * p + 1:15 * p + 15, j] = random_phase_prims(rng, eos[p], 1)[:, 0]
    return eos, prim2cons_mph(eos, P), random_interfaces(rng, k - 1)


TAU_STIFF = 1e-10  # relaxation time scale far below dt (about 5e-8 at 2^21 cells)


def time_two_phase(nx=1 << 21, steps=200, warmup=2):
    """2^21 two-phase cells, 200 steps, with the stiff relaxation sources
    (velocity and pressure, tau = 1e-10) and without them."""
    out = {}
    eos, Qs, ifc = two_phase_problem(nx)
    for flux in ("hll", "rusanov"):
        for key, src in ((flux + "_sources", (TAU_STIFF, TAU_STIFF)), (flux, None)):
            ms = time_solver(eos, Qs, ifc, nx, flux, steps, warmup, sources=src)
            ms_step = ms / steps
            out[key] = {"value": nx * steps / (ms / 1e3), "ms_per_step": ms_step,
                        "hbm_frac": BYTES_PER_CELL_2PH * nx / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS}
    out.update({"cells": nx, "interfaces": len(ifc), "steps": steps,
                "sources": {"tau_velocity": TAU_STIFF, "tau_pressure": TAU_STIFF},
                "bytes_per_cell_update": BYTES_PER_CELL_2PH})
    return out


def batched_problem(seed=4):
    """4096 problems, lengths floor(exp(U(ln 256, ln 16384))), (rho0, c0, b0)
    scaled by U(0.8, 1.2), 1-8 seeded interfaces each, single phase."""
    from hyperelasticsolver_amd.EquationsOfState import Barton2009
    from hyperelasticsolver_amd.initial import prim2cons, random_interfaces, random_phase_prims
    rng = np.random.default_rng(seed)
    P = 4096
    ncells = np.floor(np.exp(rng.uniform(np.log(256), np.log(16384), P))).astype(np.int64)
    eos, states, ifcs = [], [], []
    for p in range(P):
        s = rng.uniform(0.8, 1.2, 3)
        e = Barton2009(_rho0=8.93 * s[0], _c0=4.6 * s[1], _b0=2.1 * s[2])
        k = int(rng.integers(1, 9)) + 1
        eos.append(e)
        states.append(prim2cons(e, random_phase_prims(rng, e, k)))
        ifcs.append(random_interfaces(rng, k - 1))
    return eos, ncells, states, ifcs


def time_batched(steps, loop_steps=5, eager_only=False):
    """eager_only: the step kernels launched one by one with the default launch
    shape (no autotuning, no captured graph, no per-problem loop), for kernel
    counters."""
    from hyperelasticsolver_amd.batched import Batch
    eos, ncells, states, ifcs = batched_problem()
    total = int(ncells.sum())
    with Batch(eos, ncells, flux="rusanov") as b:
        b.set_piecewise(states, ifcs)
        if eager_only:
            b.set_launch(64, True, 1)
            b.run(1, mode="eager")
            ms = b.run(steps, mode="eager")
            Q = b.get_state()
            if not np.all(np.isfinite(Q)):
                raise RuntimeError("non-finite batched state")
            return {"problems": len(ncells), "cells": total, "steps": steps, "mode": "eager",
                    "ms_per_step": ms / steps}
        tuned = b.autotune(steps=3)
        b.run(2, mode="graph")                     # warm-up
        ms_graph = b.run(steps, mode="graph")       # whole loop in one hipGraph launch
        ms_loop = b.run(loop_steps, mode="loop")    # one launch per problem per step
        Q = b.get_state()
    if not np.all(np.isfinite(Q)):
        raise RuntimeError("non-finite batched state")
    per_step = ms_graph / steps
    loop_per_step = ms_loop / loop_steps
    return {"value": total * steps / (ms_graph / 1e3), "unit": "cell-updates/s",
            "problems": len(ncells), "cells": total, "steps": steps, "ms_per_step": per_step,
            "hbm_frac": BYTES_PER_CELL_1PH * total / (per_step / 1e3) / 1e9 / HBM_PEAK_GBS,
            "loop_ms_per_step": loop_per_step, "speedup_vs_one_problem_per_launch": loop_per_step / per_step,
            "autotune": tuned}


def time_batched_two_phase(nprob=512, n=4096, steps=20, loop_steps=3, eager_only=False):
    """512 two-phase problems of 4096 cells (2^21 cells in all), each the
    two-phase workload's four seeded pieces, HLL and Rusanov; one launch
    sequence (jump-face scan, path terms, step) per step over the whole batch,
    replayed from a graph, against one sequence per problem per step.
    eager_only: the launch sequences issued one by one, nothing else."""
    from hyperelasticsolver_amd.batched import Batch
    eos, Qs, ifc = two_phase_problem(1 << 21)
    out = {"problems": nprob, "cells": nprob * n, "bytes_per_cell_update": BYTES_PER_CELL_2PH}
    for flux in ("hll", "rusanov"):
        with Batch([eos] * nprob, [n] * nprob, flux=flux) as b:
            b.set_piecewise([Qs] * nprob, [ifc] * nprob)
            if eager_only:
                b.set_launch(64, True, 1)
                b.run(1, mode="eager")
                out[flux] = {"mode": "eager", "ms_per_step": b.run(steps, mode="eager") / steps}
                continue
            tuned = b.autotune(steps=3)
            b.run(2, mode="graph")
            ms = b.run(steps, mode="graph") / steps
            loop_ms = b.run(loop_steps, mode="loop") / loop_steps
            if not np.all(np.isfinite(b.get_state())):
                raise RuntimeError("non-finite batched two-phase state")
        out[flux] = {"ms_per_step": ms, "value": nprob * n / (ms / 1e3),
                     "hbm_frac": BYTES_PER_CELL_2PH * nprob * n / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                     "loop_ms_per_step": loop_ms, "speedup_vs_one_problem_per_launch": loop_ms / ms,
                     "autotune": tuned}
    return out


def adjoint_problem(n, steps, every, seed=2):
    """Two seeded pieces split at x = 1/2 (A = I + 0.05 N(0,1), u ~ U(-1/2,1/2)^3,
    S = 1e-3) and the target profiles of a run with b0 * 1.05 and c0 * 0.97."""
    from hyperelasticsolver_amd.EquationsOfState import Barton2009
    from hyperelasticsolver_amd.adjoint import Adjoint, phase_prims_to_adjoint
    from hyperelasticsolver_amd.initial import random_phase_prims
    rng = np.random.default_rng(seed)
    eos = Barton2009()
    P14 = random_phase_prims(rng, eos, 2)
    prims = phase_prims_to_adjoint(P14)
    eos_t = Barton2009(_b0=eos.b0 * 1.05, _c0=eos.c0 * 0.97)
    with Adjoint(eos_t, n, steps, every=every) as t:
        t.set_initial(prims, [0.5])
        t.set_target(np.zeros(n), np.zeros(n))
        t.forward()
        rho_t, u_t = t.profiles()
    return eos, P14, prims, rho_t, u_t


ADJOINT_MODES = ("graph_sparse", "graph_dense", "eager_sparse", "eager_dense")


def time_adjoint(n=1 << 18, steps=2000, every=50, modes=ADJOINT_MODES):
    """(forward + adjoint) / forward: the plain forward is the solver's step
    loop on the same problem (its steps replayed from one captured graph, as
    the adjoint's graph modes replay theirs); the adjoint run is its
    checkpointed forward pass plus the reverse pass (recompute + reverse
    steps).  Reported for each launch mode, with zero-tile skipping (the
    default) and without it (every tile of every reverse step evaluated)."""
    from hyperelasticsolver_amd.adjoint import Adjoint
    from hyperelasticsolver_amd.initial import prim2cons
    from hyperelasticsolver_amd.solver import Solver
    eos, P14, prims, rho_t, u_t = adjoint_problem(n, steps, every)
    with Solver(eos, n, flux="rusanov", cfl=0.6) as s:
        s.set_piecewise(prim2cons(eos, P14), [0.5])
        s.time_steps(50)
        s.set_piecewise(prim2cons(eos, P14), [0.5])
        fwd_ms = s.time_steps(steps)
    runs = {}
    grads = []
    with Adjoint(eos, n, steps, every=every) as a:
        a.set_initial(prims, [0.5])
        a.set_target(rho_t, u_t)
        info = a.info()
        for mode in modes:
            graph, sparse = mode.startswith("graph"), mode.endswith("sparse")
            a.set_launch(graph)
            a.set_sparse(sparse)
            a.forward()
            a.backward()          # the first calls capture the graphs
            loss = a.forward()
            g, gm = a.backward()
            grads.append(np.concatenate([g.ravel(), gm]))
            runs[mode] = {"forward_ms": a.last_forward_ms, "backward_ms": a.last_backward_ms,
                          "ratio": (a.last_forward_ms + a.last_backward_ms) / fwd_ms}
    if not all(np.array_equal(grads[0], x) for x in grads[1:]):
        raise RuntimeError("launch modes disagree")
    if not np.all(np.isfinite(grads[0])):
        raise RuntimeError("non-finite gradient")
    out = {"cells": n, "steps": steps, "checkpoint_every": every, "loss": loss,
           "forward_ms": fwd_ms, "runs": runs, "device_bytes": info["device_bytes"],
           "checkpoints": info["checkpoints"], "grad_b0_c0": [float(gm[0]), float(gm[1])]}
    for kind in ("sparse", "dense"):
        cand = [(v["ratio"], k) for k, v in runs.items() if k.endswith(kind)]
        if cand:
            r, k = min(cand)
            out["fwd_plus_adj_over_fwd" + ("" if kind == "sparse" else "_dense")] = r
            out["mode" + ("" if kind == "sparse" else "_dense")] = k
    return out


def time_batched_adjoint(nprob=256, lo=1024, hi=8192, steps=2000, every=50, seed=5,
                         modes=("graph_sparse", "graph_dense")):
    """The adjoint workload for a batch: 256 problems of floor(exp(U(ln 1024,
    ln 8192))) cells, (rho0, c0, b0) of each scaled by U(0.8, 1.2), two seeded
    pieces split at x = 1/2, 2000 Rusanov steps, checkpoints every 50, each
    problem's target from a run with its b0 * 1.05 and c0 * 0.97.  The plain
    forward is the batched solver's step loop on the same problems (one
    captured graph); the adjoint run is the batched adjoint's checkpointed
    forward pass plus its reverse pass (graph modes: each replayed from one
    captured graph; eager modes: kernels launched one by one)."""
    from hyperelasticsolver_amd.EquationsOfState import Barton2009
    from hyperelasticsolver_amd.adjoint import BatchAdjoint, phase_prims_to_adjoint
    from hyperelasticsolver_amd.batched import Batch
    from hyperelasticsolver_amd.initial import prim2cons, random_phase_prims
    rng = np.random.default_rng(seed)
    ncells = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), nprob))).astype(np.int64)
    eos, P14s = [], []
    for _ in range(nprob):
        sc = rng.uniform(0.8, 1.2, 3)
        e = Barton2009(_rho0=8.93 * sc[0], _c0=4.6 * sc[1], _b0=2.1 * sc[2])
        eos.append(e)
        P14s.append(random_phase_prims(rng, e, 2))
    prims = [phase_prims_to_adjoint(P) for P in P14s]
    ifcs = [[0.5]] * nprob
    N = int(ncells.sum())
    eos_t = [Barton2009(_rho0=e.rho0, _b0=e.b0 * 1.05, _c0=e.c0 * 0.97) for e in eos]
    with BatchAdjoint(eos_t, ncells, steps, every=every) as t:
        t.set_initial(prims, ifcs)
        t.set_target(np.zeros(N), np.zeros(N))
        t.forward()
        rho_t, u_t = t.profiles()
    b = Batch(eos, ncells, flux="rusanov")
    try:
        sts = [prim2cons(e, P) for e, P in zip(eos, P14s)]
        b.set_piecewise(sts, ifcs)
        b.run(50, mode="graph")
        b.set_piecewise(sts, ifcs)
        fwd_ms = b.run(steps, mode="graph")
    finally:
        b.close()
    runs = {}
    grads = []
    with BatchAdjoint(eos, ncells, steps, every=every) as a:
        a.set_initial(prims, ifcs)
        a.set_target(rho_t, u_t)
        info = a.info()
        for mode in modes:
            a.set_launch(mode.startswith("graph"))
            a.set_sparse(mode.endswith("sparse"))
            a.forward()
            a.backward()          # the first calls capture the graphs
            losses = a.forward()
            g = a.backward()
            grads.append(np.concatenate([np.concatenate([x.ravel(), m]) for x, m in g]))
            runs[mode] = {"forward_ms": a.last_forward_ms, "backward_ms": a.last_backward_ms,
                          "ratio": (a.last_forward_ms + a.last_backward_ms) / fwd_ms}
    if not all(np.array_equal(grads[0], x) for x in grads[1:]):
        raise RuntimeError("batched adjoint: launch modes disagree")
    if not np.all(np.isfinite(grads[0])):
        raise RuntimeError("non-finite gradient")
    out = {"problems": nprob, "cells": N, "steps": steps, "checkpoint_every": every,
           "loss": float(losses.sum()), "forward_ms": fwd_ms, "runs": runs,
           "device_bytes": info["device_bytes"], "checkpoints": info["checkpoints"]}
    for kind in ("sparse", "dense"):
        cand = [(v["ratio"], k) for k, v in runs.items() if k.endswith(kind)]
        if cand:
            r, k = min(cand)
            out["fwd_plus_adj_over_fwd" + ("" if kind == "sparse" else "_dense")] = r
    return out


def measured_traffic(nx, flux):
    """HBM bytes per launch of the step kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs (calibrated on a copy kernel), when they were
    taken on this workload; else None."""
    try:
        with open(COUNTERS) as f:
            c = json.load(f)
    except (OSError, ValueError):
        return None
    rec = c.get(flux)
    if not rec or rec.get("cells") != nx:
        return None
    return rec["traffic_bytes"]


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=1 << 22)
    ap.add_argument("--flux", default="rusanov", choices=["rusanov", "hll"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="only the headline line (no HLL, batched or adjoint lines)")
    ap.add_argument("--parts",
                    default="hll,two_phase,transfer,batched,batched_two_phase,adjoint,adjoint_batched",
                    help="extra lines to add: any of hll, two_phase, transfer, batched, "
                         "batched_two_phase, adjoint, adjoint_batched")
    ap.add_argument("--adjoint-modes", default=",".join(ADJOINT_MODES))
    ap.add_argument("--batched-steps", type=int, default=200)
    ap.add_argument("--batched-adjoint-modes", default="graph_sparse,graph_dense")
    ap.add_argument("--batched-eager", action="store_true",
                    help="batched lines: only eager launches of the step kernels with the "
                         "default launch shape (for kernel counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; the multi-GPU slab "
                                   "decomposition is a design, not measured",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from hyperelasticsolver_amd.solver import device_info
    nx = args.nx
    eos, Qs, ifc = headline_problem(nx)
    ms = time_solver(eos, Qs, ifc, nx, args.flux, args.steps, args.warmup)
    ms_step = ms / args.steps
    value = nx * args.steps / (ms / 1e3)
    achieved = BYTES_PER_CELL_1PH * nx / (ms_step / 1e3) / 1e9
    from hyperelasticsolver_amd.solver import copy_bandwidth
    copy_gbs = copy_bandwidth(1 << 30, 5)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"single-phase 1D Riemann, {nx} cells, 64 seeded interfaces, "
                               f"{args.flux}, CFL 0.6, Barton2009 copper",
                   "cells": nx, "flux": args.flux, "interfaces": 64},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(nx, args.flux),
                     "kernel": f"k_step<1, {int(args.flux == 'hll')}, 1, {int(nx >= 1 << 21)}, 0>",
                     "bytes_per_cell_update": BYTES_PER_CELL_1PH,
                     "copy_gbs": copy_gbs, "frac_of_copy": achieved / copy_gbs},
        "device": device_info(),
    }
    parts = set() if args.no_extra else set(p for p in args.parts.split(",") if p)
    if "hll" in parts:
        other = "hll" if args.flux == "rusanov" else "rusanov"
        ms2 = time_solver(eos, Qs, ifc, nx, other, args.steps, args.warmup)
        v2 = nx * args.steps / (ms2 / 1e3)
        out[other] = {"value": v2, "ms_per_step": ms2 / args.steps,
                      "hbm_frac": BYTES_PER_CELL_1PH * v2 / 1e9 / HBM_PEAK_GBS}
    if "two_phase" in parts:
        out["two_phase"] = time_two_phase()
    if "transfer" in parts:
        out["host_transfer"] = host_transfer(eos, Qs, ifc, nx)
    if "batched" in parts:
        out["batched"] = time_batched(args.batched_steps, eager_only=args.batched_eager)
    if "batched_two_phase" in parts:
        out["batched_two_phase"] = time_batched_two_phase(
            steps=min(20, args.batched_steps), eager_only=args.batched_eager)
    if "adjoint" in parts:
        out["adjoint"] = time_adjoint(modes=tuple(m for m in args.adjoint_modes.split(",") if m))
    if "adjoint_batched" in parts:
        out["adjoint_batched"] = time_batched_adjoint(
            modes=tuple(m for m in args.batched_adjoint_modes.split(",") if m))
    if not args.no_cpu_baseline:
        # every core this process may run on (the OpenMP default would follow
        # OMP_NUM_THREADS, which may be set lower)
        threads = len(os.sched_getaffinity(0))
        out["cpu_baseline"] = cpu_baseline(eos, Qs, ifc, 0 if args.flux == "rusanov" else 1, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
