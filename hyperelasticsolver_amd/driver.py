"""The driver of main.jl: initial conditions, snapshots, restart and the time
loop, on the GPU solver.

- get_filename(step_num)           main.jl:108   "sol_%06i.csv"
- save_data(fname, Q, eos)         main.jl:67-77 TSV of cons2prim_mph per cell, header of main.jl:71
- read_data(fname)                 main.jl:84-92 -> (P (30 x nx), nx)
- initial_condition(Ql, Qr, nx)    main.jl:99-106 (cell i, 1-based, is left iff (i - 1) < nx / 2)
- initial_condition_tanh(eos, Ql, Qr, nx, eps)   main.jl:110-123
- run(...)                         main.jl:130-245

Numbers are written the way Julia's `string(::Float64)` prints them (shortest
round-trip digits; decimal for decimal exponents -4..5, else `d.ddde-x`), so
the files read the same in both.  The reference restarts at
`t = step_num * dt` with the placeholder `dt = 5e-6` of main.jl:144, which is
not the time of the snapshot; here every snapshot also gets a sidecar file
`<name>.t` with its time and step, and a restart uses it (falling back to the
reference's estimate when it is missing).
"""
import logging
import math
import os

import numpy as np

from . import HyperelasticityMPh as H
from .solver import Solver

log = logging.getLogger("hyperelasticsolver_amd.driver")

HEADER = ("a1\tr1\tu11\tu21\tu31\tS1\tF111\tF211\tF311\tF121\tF221\tF321\tF131\tF231\tF331\t"
          "a2\tr2\tu12\tu22\tu32\tS2\tF112\tF212\tF312\tF122\tF222\tF322\tF132\tF232\tF332")


def get_filename(step_num):
    """main.jl:108"""
    return "sol_%06i.csv" % int(step_num)


def julia_float(x):
    """`string(x)` of a Julia Float64: the shortest digits that round-trip,
    in decimal notation when the decimal exponent is in -4..5 (with at least
    one digit after the point), else as `d.ddd` + `e` + exponent."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    r = repr(abs(x))                        # shortest round-trip digits
    mant, _, ex = r.partition("e")
    e = int(ex) if ex else 0
    if "." in mant:
        ip, fp = mant.split(".")
    else:
        ip, fp = mant, ""
    digits = (ip + fp).lstrip("0")
    # decimal exponent of the leading digit
    if ip.strip("0"):
        e10 = e + len(ip.lstrip("0")) - 1
    else:
        e10 = e - (len(fp) - len(fp.lstrip("0"))) - 1
    digits = digits.rstrip("0") or "0"
    sign = "-" if x < 0 else ""
    if -4 <= e10 <= 5:
        if e10 >= 0:
            ip2 = digits[:e10 + 1].ljust(e10 + 1, "0")
            fp2 = digits[e10 + 1:] or "0"
        else:
            ip2 = "0"
            fp2 = "0" * (-e10 - 1) + digits
        return f"{sign}{ip2}.{fp2}"
    return f"{sign}{digits[0]}.{digits[1:] or '0'}e{e10}"


def save_data(fname, Q, eos):
    """main.jl:67-77: the header line, then the 30 primitives of every cell."""
    P = H.cons2prim_mph(eos, np.asfortranarray(Q))
    with open(fname, "w") as f:
        f.write(HEADER + "\n")
        for i in range(P.shape[1]):
            f.write("\t".join(julia_float(v) for v in P[:, i]) + "\n")


def read_data(fname):
    """main.jl:84-92: (P, nx) with P 30 x nx."""
    with open(fname) as f:
        lines = f.read().splitlines()
    rows = [ln.split() for ln in lines[1:] if ln.strip()]
    P = np.asfortranarray(np.array(rows, dtype=np.float64).T)
    return P, P.shape[1]


def initial_condition(Ql, Qr, nx):
    """main.jl:99-106: cell i (1-based) takes Ql iff (i - 1) < nx / 2."""
    i = np.arange(nx)
    left = i < nx / 2
    return np.asfortranarray(np.where(left[None, :], np.asarray(Ql)[:, None], np.asarray(Qr)[:, None]))


def initial_condition_tanh(eos, Ql, Qr, nx, eps):
    """main.jl:110-123: the left/right primitives with a tanh volume-fraction
    profile a1 = 0.1 (tanh(4 (x - 1/2) / eps) + 1) + 0.4, a2 = 1 - a1."""
    Pl = H.cons2prim_mph(eos, Ql)
    Pr = H.cons2prim_mph(eos, Qr)
    x = (np.arange(1, nx + 1) - 0.5) / nx
    P = np.asfortranarray(np.where((x < 0.5)[None, :], Pl[:, None], Pr[:, None]))
    P[0] = 0.2 / 2 * (np.tanh(4 * (x - 0.5) / eps) + 1) + 0.4
    P[15] = 1 - P[0]
    return H.prim2cons_mph(eos, P)


def _write_time(fname, t, step_num):
    with open(fname + ".t", "w") as f:
        f.write(f"{t!r} {int(step_num)}\n")


def _read_time(fname):
    try:
        with open(fname + ".t") as f:
            t, s = f.read().split()
        return float(t), int(s)
    except (OSError, ValueError):
        return None


def run(eos=None, testcase=6, X=1.0, T=0.06, nx=1000, cfl=0.6, log_freq=100,
        dir_name="barton_data", flux="hll", tanh_eps=None, dt_guess=5e-6):
    """main.jl:130-245 on the GPU.  Starts from the test case (or resumes from
    the last snapshot in `dir_name` unless it holds `result.csv` or nothing),
    writes `sol_%06i.csv` every `log_freq` steps and `result.csv` at the end,
    and logs every step as main.jl does.  Returns (Q, t, step_num)."""
    if eos is None:
        from .EquationsOfState import Barton2009
        eos = (Barton2009(), Barton2009())
    dx = X / nx
    os.makedirs(dir_name, exist_ok=True)
    files = sorted(f for f in os.listdir(dir_name) if f.endswith(".csv"))
    t, step_num = 0.0, 0
    if "result.csv" in files or not files:
        for f in os.listdir(dir_name):
            os.remove(os.path.join(dir_name, f))
        log.info("Cleaning data directory: %s", dir_name)
        Ql, Qr = H.initial_states(eos, testcase)
        Q0 = (initial_condition(Ql, Qr, nx) if tanh_eps is None
              else initial_condition_tanh(eos, Ql, Qr, nx, tanh_eps))
    else:
        last = os.path.join(dir_name, files[-1])
        log.info("Found file: %s", last)
        step_num = int(os.path.basename(last).split(".")[0].split("_")[1])
        tt = _read_time(last)
        if tt is not None:
            t, step_num = tt
        else:  # the reference's estimate (main.jl:178)
            t = step_num * dt_guess
            log.warning("no time recorded for %s: resuming at step_num * dt = %g", last, t)
        P0, nx = read_data(last)
        dx = X / nx
        Q0 = H.prim2cons_mph(eos, P0)
    fname = os.path.join(dir_name, get_filename(step_num))
    save_data(fname, Q0, eos)
    _write_time(fname, t, step_num)
    log.info("Initial state saved to: %s", fname)

    with Solver(eos, nx, flux=flux, cfl=cfl, dx=dx) as s:
        s.set_state(Q0, t)
        while t < T:
            chunk = log_freq - step_num % log_freq
            r = s.run(steps=chunk, t_end=T, use_graph=False, record_dt=True)
            for dt in r["dt"]:
                t += dt
                step_num += 1
                log.info("Step = %d,\t t = %.6f / %.6f,\t Δt = %.6f", step_num, t, T, dt)
            t = r["t"]
            if step_num % log_freq == 0 and r["steps"] > 0:
                fname = os.path.join(dir_name, get_filename(step_num))
                Q = s.get_state()
                save_data(fname, Q, eos)
                _write_time(fname, t, step_num)
                log.info("Solution saved to: %s", fname)
            if r["steps"] < chunk:
                break
        Q = s.get_state()
    fname = os.path.join(dir_name, "result.csv")
    save_data(fname, Q, eos)
    _write_time(fname, t, step_num)
    log.info("Result solution saved to: %s", fname)
    return Q, t, step_num


__all__ = ["get_filename", "save_data", "read_data", "initial_condition", "initial_condition_tanh",
           "julia_float", "run", "HEADER"]
