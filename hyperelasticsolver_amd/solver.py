"""Time stepping on the GPU: the loop of main.jl:202-241 behind libhes.so.

`Solver` owns the device state of one 1D problem (structure-of-arrays FP64 in
HBM) and advances it with one fused kernel launch per step.  The state crosses
the host boundary only in `set_state` / `get_state`, in the reference's Julia
layout (ncomp x nx, component fastest).
"""
import ctypes
import math

import numpy as np

from . import _lib
from ._lib import FLUX_HLL, FLUX_LXF, FLUX_RUSANOV, check, columns, lib, ptr

# flux kinds of the time step: the face speed of NumFluxes.lxf as main.jl:46-50
# chooses it (Rusanov: max over the face's two cells; Lax-Friedrichs: dx / dt),
# or NumFluxes.hll (main.jl:56-57)
FLUXES = {"rusanov": FLUX_RUSANOV, "hll": FLUX_HLL, "lxf": FLUX_LXF, "lax_friedrichs": FLUX_LXF}


def materials_array(eos):
    """(rho0, c0, cv, t0, b0, alpha, beta, gamma) per phase, concatenated."""
    if isinstance(eos, (tuple, list)):
        return np.concatenate([e.params() for e in eos]).astype(np.float64)
    return eos.params().astype(np.float64)


def nphase_of(eos):
    if isinstance(eos, (tuple, list)):
        if len(eos) not in (1, 2):
            raise ValueError("one or two phases are supported")
        return len(eos)
    return 1


def ncomp_of(nph):
    return 14 if nph == 1 else 30


class Solver:
    """One 1D finite-volume problem on the GPU.

    eos   : Barton2009 (one phase, 14-component states) or a tuple of two
            (two phases, 30-component states, HyperelasticityMPh layout)
    flux  : "rusanov" (lxf with the local max speed, main.jl:46-48), "hll"
            (main.jl:56-57) or "lxf" / "lax_friedrichs" (lxf with
            lambda = dx / dt on every face, main.jl:30-41, 50)
    """

    def __init__(self, eos, nx, flux="hll", cfl=0.6, X=1.0, dx=None):
        self.eos = eos
        self.nph = nphase_of(eos)
        if self.nph == 1 and isinstance(eos, (tuple, list)):
            eos = eos[0]
        self.nc = ncomp_of(self.nph)
        self.nx = int(nx)
        self.flux = FLUXES[flux] if isinstance(flux, str) else int(flux)
        self.cfl = float(cfl)
        self.dx = float(X / nx if dx is None else dx)  # main.jl:147
        self._mats = materials_array(self.eos if self.nph == 2 else eos)
        h = ctypes.c_void_p()
        check(lib().hes_solver_create(ctypes.byref(h), self.nph, self.flux, self.nx,
                                      ptr(self._mats), self.cfl, self.dx))
        self._h = h

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().hes_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state
    def set_state(self, Q, t=0.0):
        q, n, _ = columns(Q, self.nc, "Q")
        if n != self.nx:
            raise ValueError(f"Q has {n} cells, solver has {self.nx}")
        check(lib().hes_solver_set_state(self._h, ptr(q), float(t)))

    def set_piecewise(self, states, interfaces, t=0.0):
        """Piecewise-constant state built on the GPU from conservative `states`
        (ncomp x k) separated by k-1 sorted interfaces in (0, 1)."""
        st, k, _ = columns(states, self.nc, "states")
        ifc = np.ascontiguousarray(np.asarray(interfaces, dtype=np.float64).reshape(-1))
        if ifc.size != k - 1:
            raise ValueError("need one interface fewer than states")
        if ifc.size == 0:
            ifc = np.zeros(1)
        check(lib().hes_solver_set_piecewise(self._h, k, ptr(st), ptr(ifc), float(t)))

    def set_state_device(self, dev_ptr, t=0.0):
        """State from a device pointer (Julia layout, e.g. a torch cuda tensor's data_ptr())."""
        check(lib().hes_solver_set_state_device(self._h, ctypes.c_void_p(int(dev_ptr)), float(t)))

    def get_state(self):
        out = np.empty((self.nc, self.nx), order="F")
        check(lib().hes_solver_get_state(self._h, ptr(out)))
        return out

    def get_speeds(self):
        out = np.empty((self.nph, self.nx))
        check(lib().hes_solver_get_speeds(self._h, ptr(out)))
        return out

    @property
    def time(self):
        t = ctypes.c_double()
        s = ctypes.c_longlong()
        check(lib().hes_solver_time(self._h, ctypes.byref(t), ctypes.byref(s)))
        return t.value

    @property
    def steps(self):
        t = ctypes.c_double()
        s = ctypes.c_longlong()
        check(lib().hes_solver_time(self._h, ctypes.byref(t), ctypes.byref(s)))
        return s.value

    def next_dt(self):
        d = ctypes.c_double()
        check(lib().hes_solver_next_dt(self._h, ctypes.byref(d)))
        return d.value

    def set_relaxation(self, tau=math.inf):
        """Two phases: interphase velocity relaxation with time scale tau after
        every flux update (an optional source; the reference has none).
        tau = inf switches it off."""
        check(lib().hes_solver_set_relaxation(self._h, float(tau)))

    def set_pressure_relaxation(self, tau=math.inf):
        """Two phases: implicit pressure relaxation with time scale tau after
        every flux update (an optional stiff source; the reference has none):
        the volume fractions move towards equal phase pressures, conserving
        mixture mass, momentum and energy.  tau = inf switches it off."""
        check(lib().hes_solver_set_pressure_relaxation(self._h, float(tau)))

    # -- stepping
    def run(self, steps=None, t_end=math.inf, use_graph=True, record_dt=False):
        """Advance up to `steps` steps while t < t_end.  Returns a dict with the
        executed step count, the time, the elapsed GPU time (ms) and optionally
        the dt of every step."""
        if steps is None:
            if not math.isfinite(t_end):
                raise ValueError("give steps or a finite t_end")
            steps = 1 << 40
        steps = int(steps)
        done = ctypes.c_longlong()
        tnow = ctypes.c_double()
        ms = ctypes.c_float()
        dts = None
        dtp = None
        if record_dt:
            if steps > 50_000_000:
                raise ValueError("record_dt needs a bounded step count")
            dts = np.zeros(steps)
            dtp = ptr(dts)
        check(lib().hes_solver_run(self._h, steps, float(t_end), 1 if use_graph else 0,
                                   ctypes.byref(done), ctypes.byref(tnow), ctypes.byref(ms), dtp))
        res = {"steps": done.value, "t": tnow.value, "ms": ms.value}
        if record_dt:
            res["dt"] = dts[: done.value]
        return res

    def time_steps(self, reps):
        """GPU time (ms) of `reps` back-to-back steps, replayed from one captured graph."""
        ms = ctypes.c_float()
        check(lib().hes_solver_time_steps(self._h, int(reps), ctypes.byref(ms)))
        return ms.value


def device_info():
    name = ctypes.create_string_buffer(64)
    cus = ctypes.c_int()
    mem = ctypes.c_double()
    check(lib().hes_device_info(name, 64, ctypes.byref(cus), ctypes.byref(mem)))
    return {"arch": name.value.decode(), "compute_units": cus.value, "mem_gb": mem.value}


def copy_bandwidth(nbytes=1 << 31, reps=5):
    """Measured device-copy bandwidth in GB/s (read + write bytes per second)."""
    g = ctypes.c_double()
    check(lib().hes_copy_bandwidth(int(nbytes), int(reps), ctypes.byref(g)))
    return g.value
