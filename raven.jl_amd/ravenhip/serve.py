"""Serving many solves on one GPU with continuous batching.

A request is one brick solve (2-D or 3-D, advection or Euler, degree 3-7,
FP32 or FP64) to a final time T.  The server keeps every admitted solve's
state resident in device memory, admits waiting requests in arrival order as
soon as their device footprint fits the budget, and advances all admitted
solves of one dimension together: LSRK54 steps are rv_batch_step(s) calls
over the current mix, with every member's own dt and its final step shortened
to land on its T exactly as ``solve!`` does (odesolvers.jl:9-30).  Requests
that arrive while others run join at the next step; finished ones leave and
free their memory.  Each member's stages are the ones it would get alone, so
its result does not depend on the mix.

``run_plain`` serves the same stream one request at a time through the plain
per-operator path (rv_lsrk_run) for comparison.
"""
import itertools
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .batch import Batch
from .dgsem import DGSEM, AdvectionLaw, EulerLaw, RusanovFlux
from .grids import GridManager, LobattoCell, brick

GAMMA = 1.4
CFL = 0.125


@dataclass
class SolveRequest:
    """One solve: a periodic brick of n^dim unit cells, final time T."""
    rid: int
    arrival: float          # seconds after the stream starts
    dim: int
    law: str                # "advection" or "euler"
    nq: int
    n: int                  # cells per direction
    T: float
    dtype: str              # "f64" or "f32"
    seed: int = 0
    # filled in by the server
    admitted: float = math.nan
    finished: float = math.nan
    steps: int = 0
    result: object = field(default=None, repr=False)

    @property
    def ns(self):
        return 1 if self.law == "advection" else self.dim + 2

    @property
    def cells(self):
        return self.n ** self.dim

    @property
    def dofs(self):
        return self.cells * self.nq ** self.dim * self.ns

    def velocity(self):
        rng = np.random.default_rng(self.seed)
        u = rng.uniform(-1.0, 1.0, self.dim)
        return u / np.linalg.norm(u)

    def speed(self):
        # |u| for advection; |u0| + c0 of the vortex background for Euler
        return 1.0 if self.law == "advection" else 0.5 + math.sqrt(GAMMA)

    def dt(self):
        """CFL rule of Raven's tests, cfl * h / N / speed (h = 1 here)."""
        return CFL * 1.0 / (self.nq - 1) / self.speed()

    def state_bytes(self):
        return self.dofs * (4 if self.dtype == "f32" else 8)

    def footprint(self):
        """Device bytes while resident: q, dq and the ping-pong copy, plus the
        operator's FP64 MJ, face records and per-cell geometry."""
        return (3 * self.state_bytes() + self.cells * self.nq ** self.dim * 8
                + self.cells * (6 * 8 + 4 * 8) + (1 << 16))


def make_operator(req, stream=None):
    cell = LobattoCell(*([req.nq] * req.dim))
    gm = GridManager(cell, _brick(req))
    law = AdvectionLaw(req.dim, list(req.velocity())) if req.law == "advection" \
        else EulerLaw(req.dim, GAMMA)
    return gm, DGSEM.brick(gm, law, RusanovFlux(), dtype=req.dtype, stream=stream)


def _brick(req):
    coords = [np.linspace(0.0, float(req.n), req.n + 1)] * req.dim
    return brick(coords, (True,) * req.dim)


def initial_state(req, gm, dg):
    """A smooth initial state evaluated on the device from the cells' lattice
    positions and the Gauss-Lobatto points: a Gaussian pulse (advection) or an
    isentropic vortex (Euler, extruded in z for 3-D) at a seeded centre."""
    import torch
    desc = gm.brick_desc()
    ijk = np.zeros((dg.K, req.dim), dtype=np.int32)
    L.check(L.lib().rv_brick_cells(L.ctypes.byref(desc), L.ptr(ijk, L.ctypes.c_int32)))
    xi = torch.tensor((gm.cell.points_1d + 1.0) / 2.0, dtype=torch.float64, device="cuda")
    nq, dim = req.nq, req.dim
    base = torch.from_numpy(ijk.astype(np.float64)).cuda()           # (K, dim)
    # node m = i + nq*j (+ nq^2*k): coordinate d uses index (m // nq^d) % nq
    m = torch.arange(nq ** dim, device="cuda")
    X = [base[:, d:d + 1] + xi[(m // nq ** d) % nq][None, :] for d in range(dim)]  # (K, Np)
    rng = np.random.default_rng(req.seed + 1)
    c = rng.uniform(0.3, 0.7, dim) * req.n
    if req.law == "advection":
        r2 = sum((X[d] - c[d]) ** 2 for d in range(dim))
        q = torch.exp(-r2 / (2.0 * (req.n / 8.0) ** 2))[:, None, :]
    else:
        # isentropic vortex (test/balancelaws/euler/isentropicvortex.jl:18-34),
        # radius scaled to n/10, background velocity 0.5 along x
        beta, u0 = 5.0, 0.5
        s = req.n / 10.0
        dx, dy = (X[0] - c[0]) / s, (X[1] - c[1]) / s
        f = 1.0 - (dx * dx + dy * dy)
        ef = torch.exp(f / 2.0)
        u = u0 - beta / (2.0 * math.pi) * ef * dy
        v = beta / (2.0 * math.pi) * ef * dx
        Tt = 1.0 - (GAMMA - 1.0) * beta ** 2 / (8.0 * GAMMA * math.pi ** 2) * torch.exp(f)
        rho = Tt ** (1.0 / (GAMMA - 1.0))
        p = rho ** GAMMA
        comps = [rho, rho * u, rho * v] + ([torch.zeros_like(rho)] if dim == 3 else [])
        E = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v)
        q = torch.stack(comps + [E], dim=1)
    return q.to(dg.torch_dtype).contiguous()


class _Active:
    """A resident solve: operator, states and its own clock (rv_lsrk_run's
    arithmetic, FP32 time for FP32 states)."""

    _serial = itertools.count()

    def __init__(self, req, a, b, stream):
        import torch
        self.key = next(_Active._serial)  # unique for the process (ids are reused)
        self.req = req
        self.gm, self.dg = make_operator(req, stream)
        q = initial_state(req, self.gm, self.dg)
        self.bufs = [q, torch.empty_like(q)]
        self.dq = torch.zeros_like(q)
        self.cur = 0
        f32 = req.dtype == "f32"
        rnd = (lambda v: float(np.float32(v))) if f32 else float
        self.f32 = f32
        self.a = [rnd(v) for v in a]
        self.b = [rnd(v) for v in b]
        self.time = rnd(0.0)
        self.dt = rnd(req.dt())
        self.T = float(req.T)
        self.rnd = rnd
        self.done = False

    def next_step(self):
        """dt of the coming step and whether it is the last (odesolvers.jl:21-25):
        time + dt in the state type, compared with T as given; the last step
        is T - time rounded to the state type."""
        if self.f32:
            f = np.float32
            tdt = float(f(self.time) + f(self.dt))
        else:
            tdt = self.time + self.dt
        if tdt >= self.T:
            return self.rnd(self.T - self.time), True
        return self.dt, False

    def full_steps(self, limit):
        """How many of the coming steps, at most limit, are whole steps of dt
        before the final one (the clock arithmetic of next_step)."""
        t, n = self.time, 0
        while n < limit:
            if self.f32:
                tdt = float(np.float32(t) + np.float32(self.dt))
            else:
                tdt = t + self.dt
            if tdt >= self.T:
                break
            t, n = tdt, n + 1
        return n

    def bdt(self, s, step):
        if self.f32:
            return float(np.float32(self.b[s]) * np.float32(step))
        return self.b[s] * step

    def advance_clock(self, step, last):
        if self.f32:
            self.time = float(np.float32(self.time) + np.float32(step))
        else:
            self.time = self.time + step
        self.done = last


class _Group:
    """The resident solves of one dimension and their batch.  Finished solves
    stay in the batch with step 0 until the next admission rebuilds it."""

    def __init__(self, members):
        import ctypes
        self.members = list(members)
        self.batch = Batch([m.dg for m in self.members])
        n = len(self.members)
        self.cur = np.zeros(n, dtype=np.int32)
        self.dt = np.zeros(n, dtype=np.float64)
        vp = ctypes.c_void_p * n
        self.q0 = vp(*[m.bufs[0].data_ptr() for m in self.members])
        self.q1 = vp(*[m.bufs[1].data_ptr() for m in self.members])
        self.dq = vp(*[m.dq.data_ptr() for m in self.members])

    def step(self, scheme, nstages, max_steps=8):
        """LSRK steps of every unfinished member in one C call: as many whole
        steps (at most max_steps) as come before any member's final, shortened
        step, else that one step.  Returns (steps taken, members that
        finished with them)."""
        plan = []
        run = max_steps
        for k, m in enumerate(self.members):
            if m.done:
                self.dt[k] = 0.0
                plan.append(None)
                continue
            st, last = m.next_step()
            self.dt[k] = st
            self.cur[k] = m.cur
            plan.append((st, last))
            run = min(run, 0 if last else m.full_steps(run))
        if run >= 1:
            # whole steps of every member's dt, no final step among them
            L.check_batch(L.lib().rv_batch_steps(self.batch._b, self.q0, self.q1, self.dq,
                                                 L.ptr(self.cur, L.ctypes.c_int32),
                                                 L.ptr(self.dt, L.ctypes.c_double), scheme,
                                                 run), self.batch._b)
        else:
            run = 1
            L.check_batch(L.lib().rv_batch_step(self.batch._b, self.q0, self.q1, self.dq,
                                                L.ptr(self.cur, L.ctypes.c_int32),
                                                L.ptr(self.dt, L.ctypes.c_double), scheme),
                          self.batch._b)
        finished = []
        for m, p in zip(self.members, plan):
            if p is None:
                continue
            st, last = p
            m.cur = (m.cur + run * nstages) % 2
            for _ in range(run):
                m.advance_clock(st, last)
            m.req.steps += run
            if last:
                finished.append(m)
        return run, finished


class Server:
    """Continuous-batching server over one device.

    hbm_budget  -- device bytes the resident solves may hold together
    admit_slice -- host seconds one pass of the loop may wait for set-ups
                   under way before it steps (0: solves join as soon as
                   their set-up is done)
    max_steps   -- most LSRK steps issued in one call

    Solves step on a stream of the server's own.  Each pass of the loop
    admits arrived requests in arrival order while their footprint fits
    (reserving it), and a worker thread builds their operators and initial
    states on the default stream, which the solves' stream does not wait for,
    so set-up overlaps the device's work and the loop's step issue.  Admitted
    solves join the mix in order as their set-up completes.  The loop then
    issues LSRK steps of every resident solve: one call per dimension present,
    of as many whole steps (at most max_steps) as come before any member's
    final, shortened step, or that one step (one step per pass while requests
    wait or are being set up).  The
    host never waits for the device between steps (the batch's parameter ring
    keeps it at most a few steps ahead); a solve's completion is seen through an
    event recorded after its last step, and its states go back to the
    allocator then.  Operators and batches that are no longer used are
    destroyed when the server is idle or done, since freeing device memory
    waits for the device.
    """

    def __init__(self, hbm_budget, scheme=L.RV_LSRK54, admit_slice=0.0, max_steps=8):
        from .odesolvers import LSRK
        self.hbm_budget = int(hbm_budget)
        self.scheme = scheme
        self.admit_slice = float(admit_slice)
        self.max_steps = int(max_steps)  # steps per call when no request waits

        class _NoDG:
            def GridArray(self, fill):
                return None
        a, b, _ = LSRK(_NoDG(), None, 1.0, 0.0, scheme).coefficients()
        self.a, self.b = a, b
        self.peak_bytes = 0       # largest footprint held at once
        self.peak_active = 0      # most solves resident at once
        self.steps_run = 0        # batched steps (one per dimension present)
        # host seconds spent admitting (operator set-up, initial state, batch
        # rebuilds), issuing steps, waiting for completions, and idle
        self.host_time = {"admit": 0.0, "step": 0.0, "wait": 0.0, "idle": 0.0}

    def run(self, requests, keep_results=False, stream=None):
        """Serve the requests (arrival times relative to the call); returns
        the requests with admitted / finished times filled in.  stream: the
        stream the solves step on (default: a new stream)."""
        import concurrent.futures as cf

        import torch
        compute = stream if stream is not None else torch.cuda.Stream()
        setup = torch.cuda.current_stream()
        handle = L.ctypes.c_void_p(compute.cuda_stream)
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting = []
        # admitted solves whose operator and initial state a worker thread is
        # building (arrival order): set-up overlaps the device's work and the
        # loop's step issue instead of holding the loop
        preparing = []
        pool = cf.ThreadPoolExecutor(max_workers=1, thread_name_prefix="rv-admit")

        def prepare(r):
            with torch.cuda.stream(setup):
                return _Active(r, self.a, self.b, handle)
        groups = {}               # dim -> _Group of the solves still stepping
        stepping = []             # admitted solves whose last step is not issued
        inflight = []             # (solve, event after its last step)
        graveyard = []            # operators and batches to destroy when idle
        used = 0
        t0 = time.perf_counter()
        k = 0
        clock = self.host_time

        def bury():
            # destroying waits for the device: only when it has nothing queued
            for obj in graveyard:
                if isinstance(obj, _Group):
                    obj.batch.close()
                else:
                    obj.dg.close()
            graveyard.clear()

        while k < len(pending) or waiting or preparing or stepping or inflight:
            now = time.perf_counter() - t0
            while k < len(pending) and pending[k].arrival <= now:
                waiting.append(pending[k])
                k += 1
            # completions seen so far
            still = []
            for m, ev in inflight:
                if ev.query():
                    m.req.finished = time.perf_counter() - t0
                    m.req.result = m.bufs[m.cur] if keep_results else None
                    used -= m.req.footprint()
                    m.bufs, m.dq = None, None
                    graveyard.append(m)
                else:
                    still.append((m, ev))
            inflight = still
            # admission in arrival order while the footprint fits: the
            # memory is reserved and the set-up handed to the worker
            ta = time.perf_counter()
            while waiting and (used + waiting[0].footprint() <= self.hbm_budget
                               or not (stepping or inflight or preparing)):
                if not (stepping or inflight or preparing) and graveyard:
                    bury()
                r = waiting.pop(0)
                r.admitted = time.perf_counter() - t0
                used += r.footprint()
                preparing.append((r, pool.submit(prepare, r)))
                self.peak_bytes = max(self.peak_bytes, used)
            # the solves whose set-up is done join, in order (waiting for the
            # first only when nothing else could run)
            admitted = []
            while preparing and (preparing[0][1].done() or not (stepping or inflight)
                                 or time.perf_counter() - ta < self.admit_slice):
                m = preparing.pop(0)[1].result()
                stepping.append(m)
                admitted.append(m)
                self.peak_active = max(self.peak_active, len(stepping) + len(inflight))
            if admitted:
                # the first steps wait for the initial states
                compute.wait_stream(setup)
                for m in admitted:
                    for t in m.bufs + [m.dq]:
                        t.record_stream(compute)
                for dim in {m.req.dim for m in admitted}:
                    if dim in groups:
                        graveyard.append(groups[dim])
                    groups[dim] = _Group([m for m in stepping if m.req.dim == dim])
            tb = time.perf_counter()
            clock["admit"] += tb - ta
            # one step of every solve still stepping
            issued = False
            for dim in sorted(groups):
                g = groups[dim]
                if any(not m.done for m in g.members):
                    # no admission waits on these steps when nothing is
                    # waiting; else one step per pass
                    nrun, done = g.step(self.scheme, len(self.a),
                                        max_steps=1 if waiting or preparing
                                        else self.max_steps)
                    for m in done:
                        ev = torch.cuda.Event()
                        ev.record(compute)
                        inflight.append((m, ev))
                    self.steps_run += nrun
                    issued = True
            stepping = [m for m in stepping if not m.done]
            for dim in [d for d, g in groups.items() if all(m.done for m in g.members)]:
                graveyard.append(groups.pop(dim))
            tc = time.perf_counter()
            clock["step"] += tc - tb
            if issued:
                continue
            if preparing:
                # nothing to issue until a set-up is done
                cf.wait([preparing[0][1]], timeout=0.002)
                clock["wait"] += time.perf_counter() - tc
            elif inflight:
                # nothing to issue: wait for the earliest completion, or
                # until the next arrival could be admitted
                inflight[0][1].synchronize()
                clock["wait"] += time.perf_counter() - tc
            elif k < len(pending) and not waiting:
                bury()
                time.sleep(max(0.0, min(pending[k].arrival - (time.perf_counter() - t0), 0.01)))
                clock["idle"] += time.perf_counter() - tc
        compute.synchronize()
        pool.shutdown(wait=True)
        bury()
        return requests


def run_plain(requests, keep_results=False, stream=None, scheme=L.RV_LSRK54):
    """The same stream served one request at a time (FIFO) by the plain path:
    one operator, rv_lsrk_run, then the next request."""
    import torch
    from .odesolvers import LSRK, solve_
    pending = sorted(requests, key=lambda r: r.arrival)
    t0 = time.perf_counter()
    for r in pending:
        now = time.perf_counter() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        r.admitted = time.perf_counter() - t0
        gm, dg = make_operator(r, stream)
        q = initial_state(r, gm, dg)
        solver = LSRK(dg, q, r.dt(), 0.0, scheme)
        r.steps = solve_(q, r.T, solver)
        torch.cuda.synchronize()
        r.finished = time.perf_counter() - t0
        r.result = q if keep_results else None
        del solver, dg, gm
    return requests


def random_requests(n, rate, seed=0, dims=(2, 3), cells=(1e3, 1e5), degrees=(3, 7),
                    T=(0.1, 1.0)):
    """Poisson arrivals at `rate` per second; each request a 2-D or 3-D
    advection or Euler solve, degree uniform in `degrees`, cell count
    log-uniform in `cells`, T uniform, FP32 or FP64."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        dim = int(rng.choice(dims))
        ncell = math.exp(rng.uniform(math.log(cells[0]), math.log(cells[1])))
        side = max(2, int(round(ncell ** (1.0 / dim))))
        out.append(SolveRequest(rid=i, arrival=t, dim=dim,
                                law=str(rng.choice(["advection", "euler"])),
                                nq=int(rng.integers(degrees[0], degrees[1] + 1)) + 1,
                                n=side, T=float(rng.uniform(*T)),
                                dtype=str(rng.choice(["f32", "f64"])),
                                seed=int(rng.integers(1 << 30))))
    return out


def summarize(requests, wall):
    tts = np.array([r.finished - r.arrival for r in requests])
    dof_updates = sum(r.dofs * r.steps * 5 for r in requests)
    return {"requests": len(requests), "wall_s": wall,
            "throughput_dof_updates_per_s": dof_updates / wall,
            "p50_time_to_solution_s": float(np.percentile(tts, 50)),
            "p99_time_to_solution_s": float(np.percentile(tts, 99)),
            "mean_time_to_solution_s": float(tts.mean())}
