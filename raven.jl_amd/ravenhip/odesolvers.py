"""Low-storage Runge-Kutta over the fused HIP stage (odesolvers.jl:9-162).

``LSRK54(dg, q, dt; t0)`` / ``LSRK144`` build a solver; ``solve_(q, timeend,
solver)`` is Raven's ``solve!`` (final step shortened to land on timeend) and
runs entirely inside libravenhip (rv_lsrk_solve); ``dostep_`` runs one step
stage by stage through rv_lsrk_stage.  ``RLSRK54`` / ``RLSRK144`` are the
relaxation variants (entropy-conserving or -stable steps, odesolvers.jl:72-162):
``solve_`` steps them one rv_rlsrk_step at a time.
"""
import ctypes

import numpy as np

from . import _lib as L


def _clock(dg):
    """The arithmetic of the solver's clock: FT(v) rounding and time + dt in
    the state type (LSRK{Float32} keeps dt and time in Float32,
    odesolvers.jl:32-49), plain doubles for FP64."""
    if getattr(dg, "dtype", "f64") == "f32":
        f = np.float32
        return (lambda v: float(f(v))), (lambda a, b: float(f(a) + f(b)))
    return float, (lambda a, b: a + b)


class LSRK:
    def __init__(self, dg, q, dt, t0, scheme):
        self.dg = dg
        rnd, _ = _clock(dg)
        self.dt = rnd(dt)
        self.time = rnd(t0)
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


class RLSRK(LSRK):
    """RLSRK (odesolvers.jl:72-115): the LSRK scheme plus the relaxation
    parameter gamma (1 at the start) and the work arrays q0, k."""

    def __init__(self, dg, q, dt, t0, scheme):
        super().__init__(dg, q, dt, t0, scheme)
        self.gamma = 1.0
        self.k = dg.GridArray(0.0)
        self.q0 = dg.GridArray(0.0)
        self.newton_iterations = 0


def RLSRK54(dg, q, dt, t0=0.0):
    return RLSRK(dg, q, dt, t0, L.RV_LSRK54)


def RLSRK144(dg, q, dt, t0=0.0):
    return RLSRK(dg, q, dt, t0, L.RV_LSRK144)


def _rlsrk_dostep(q, solver):
    """dostep!(q, rlsrk) (odesolvers.jl:117-151) through rv_rlsrk_step."""
    dg = solver.dg
    g = np.array([solver.gamma])
    it = ctypes.c_int32(0)
    L.check(L.lib().rv_rlsrk_step(dg.handle, dg._check_state(q), dg._check_state(solver.dq),
                                  dg._check_state(solver.k), dg._check_state(solver.q0),
                                  solver.scheme, solver.time, solver.dt,
                                  L.ptr(g, ctypes.c_double), ctypes.byref(it)), dg.handle)
    solver.gamma = float(g[0])
    solver.newton_iterations = it.value
    solver.time = _clock(dg)[1](solver.time, solver.dt)


def solve_(q, timeend, solver, after_step=None, after_stage=None, adjust_final=True):
    """solve!(q, timeend, solver; after_step, after_stage, adjust_final)
    (odesolvers.jl:9-31); returns #steps.  Steps are taken until time + dt >=
    timeend; with adjust_final the last one is shortened to timeend - time.
    Without callbacks an LSRK solve runs inside the library (rv_lsrk_solve).
    With them the steps are taken from here: after_step(step, time, q) after
    every step (time at the step's start), after_stage(stagetime, q_stage)
    after every LSRK stage, q_stage being the stage's output (q itself after a
    step's last stage).  An RLSRK step is one library call, so RLSRK solvers
    take after_step only.  FP32 solvers keep their clock in FP32, comparing
    with timeend as given (the library loop's arithmetic)."""
    rnd, add = _clock(solver.dg)
    if isinstance(solver, RLSRK):
        if after_stage is not None:
            raise ValueError("RLSRK steps are single library calls: no after_stage")
        step_fn = _rlsrk_dostep
    elif after_step is None and after_stage is None:
        t = ctypes.c_double(0.0)
        n = ctypes.c_int64(0)
        dg = solver.dg
        L.check(L.lib().rv_lsrk_solve(dg.handle, dg._check_state(q), dg._check_state(solver.dq),
                                      solver.scheme, solver.time, solver.dt, float(timeend),
                                      1 if adjust_final else 0, ctypes.byref(t),
                                      ctypes.byref(n)), dg.handle)
        solver.time = t.value
        return n.value
    else:
        # LSRK's dq starts at zero (odesolvers.jl:43-44); rv_lsrk_solve's
        # schemes (first a = 0) give the same without reading it
        solver.dq.zero_()
        qtmp = q.new_empty(q.shape)

        def step_fn(q, solver):
            dostep_(q, solver, qtmp, after_stage=after_stage)
    step = 0
    while True:
        step += 1
        time = solver.time
        final = add(time, solver.dt) >= timeend
        if final and adjust_final:
            solver.dt = rnd(timeend - time)
        step_fn(q, solver)
        if after_step is not None:
            after_step(step, time, q)
        if final:
            return step


def dostep_(q, solver, qtmp=None, after_stage=None):
    """dostep!(q, lsrk) (odesolvers.jl:50-61) through the per-stage entry
    point, stage times time + rkc*dt; q updated in place (qtmp is the
    ping-pong buffer; an odd stage count ends with a copy).  An RLSRK solver
    takes its step through rv_rlsrk_step and needs no qtmp."""
    if isinstance(solver, RLSRK):
        _rlsrk_dostep(q, solver)
        return
    if qtmp is None:
        raise ValueError("dostep_ of an LSRK solver needs the ping-pong buffer qtmp")
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
        if after_stage is not None:
            if s == len(a) - 1 and cur is not q:
                q.copy_(cur)
                cur = q
            after_stage(solver.time + c[s] * solver.dt, cur)
    if cur is not q:
        q.copy_(cur)
    solver.time = _clock(solver.dg)[1](solver.time, solver.dt)
