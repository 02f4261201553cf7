"""Low-storage Runge-Kutta over the fused HIP stage (odesolvers.jl:9-71).

``LSRK54(dg, q, dt; t0)`` / ``LSRK144`` build a solver; ``solve_(q, timeend,
solver)`` is Raven's ``solve!`` (final step shortened to land on timeend) and
runs entirely inside libravenhip (rv_lsrk_run); ``dostep_`` runs one step
stage by stage through rv_lsrk_stage.
"""
import ctypes

from . import _lib as L

# odesolvers.jl:163-189
_LSRK54 = (
    (0, (-567301805773, 1357537059087), (-2404267990393, 2016746695238),
     (-3550918686646, 2091501179385), (-1275806237668, 842570457699)),
    ((1432997174477, 9575080441755), (5161836677717, 13612068292357),
     (1720146321549, 2090206949498), (3134564353537, 4481467310338),
     (2277821191437, 14882151754819)),
)


def _frac(v):
    return float(v) if not isinstance(v, tuple) else v[0] / v[1]


class LSRK:
    def __init__(self, dg, q, dt, t0, scheme):
        self.dg = dg
        self.dt = float(dt)
        self.time = float(t0)
        self.scheme = scheme
        self.dq = dg.GridArray(0.0)

    def coefficients(self):
        if self.scheme == L.RV_LSRK54:
            a = [_frac(v) for v in _LSRK54[0]]
            b = [_frac(v) for v in _LSRK54[1]]
            return a, b
        raise NotImplementedError("per-stage stepping is provided for LSRK54")


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
    """One LSRK54 step through the per-stage entry point; q updated in place
    (qtmp is the ping-pong buffer; an odd stage count ends with a copy)."""
    a, b = solver.coefficients()
    cur, alt = q, qtmp
    for s in range(len(a)):
        solver.dg.lsrk_stage(alt, cur, solver.dq, a[s], b[s] * solver.dt, solver.time)
        cur, alt = alt, cur
    if cur is not q:
        q.copy_(cur)
    solver.time += solver.dt
