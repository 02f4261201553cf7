"""Low-storage Runge-Kutta over the fused HIP stage (odesolvers.jl:9-71).

``LSRK54(dg, q, dt; t0)`` / ``LSRK144`` build a solver; ``solve_(q, timeend,
solver)`` is Raven's ``solve!`` (final step shortened to land on timeend) and
runs entirely inside libravenhip (rv_lsrk_run); ``dostep_`` runs one step
stage by stage through rv_lsrk_stage.
"""
import ctypes

import numpy as np

from . import _lib as L

class LSRK:
    def __init__(self, dg, q, dt, t0, scheme):
        self.dg = dg
        self.dt = float(dt)
        self.time = float(t0)
        self.scheme = scheme
        self.dq = dg.GridArray(0.0)

    def coefficients(self):
        """(rka, rkb, rkc) of the scheme as Python floats, read from the
        library (coefficients_lsrk54/144, odesolvers.jl:163-244)."""
        a, b, c = (np.zeros(14) for _ in range(3))
        n = ctypes.c_int(0)
        L.check(L.lib().rv_lsrk_coefficients(self.scheme, L.ptr(a, ctypes.c_double), L.ptr(b, ctypes.c_double),
                                             L.ptr(c, ctypes.c_double), ctypes.byref(n)))
        k = n.value
        return [float(v) for v in a[:k]], [float(v) for v in b[:k]], [float(v) for v in c[:k]]


def LSRK54(dg, q, dt, t0=0.0):
    return LSRK(dg, q, dt, t0, L.RV_LSRK54)


def LSRK144(dg, q, dt, t0=0.0):
    return LSRK(dg, q, dt, t0, L.RV_LSRK144)


def solve_(q, timeend, solver):
    """solve!(q, timeend, solver) with adjust_final = true; returns #steps."""
    t = ctypes.c_double(0.0)
    n = ctypes.c_int64(0)
    dg = solver.dg
    L.check(L.lib().rv_lsrk_run(dg.handle, dg._check_state(q), dg._check_state(solver.dq),
                                solver.scheme, solver.time, solver.dt, float(timeend),
                                ctypes.byref(t), ctypes.byref(n)), dg.handle)
    solver.time = t.value
    return n.value


def dostep_(q, solver, qtmp):
    """dostep!(q, lsrk) (odesolvers.jl:50-61) through the per-stage entry
    point, stage times time + rkc*dt; q updated in place (qtmp is the
    ping-pong buffer; an odd stage count ends with a copy)."""
    a, b, c = solver.coefficients()
    if solver.dg.dtype == "f32":
        # LSRK{Float32}: coefficients and dt in the state type, rkb*dt in FP32
        # (odesolvers.jl:41-48, 57)
        f = np.float32
        a = [float(f(v)) for v in a]
        bdt = [float(f(v) * f(solver.dt)) for v in b]
    else:
        bdt = [v * solver.dt for v in b]
    cur, alt = q, qtmp
    for s in range(len(a)):
        solver.dg.lsrk_stage(alt, cur, solver.dq, a[s], bdt[s],
                             solver.time + c[s] * solver.dt)
        cur, alt = alt, cur
    if cur is not q:
        q.copy_(cur)
    solver.time += solver.dt
