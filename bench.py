"""Benchmark: 3-D compressible Euler, degree 7 (Nq = 8), periodic 32^3 hex brick
over [0,10]^3 (16.8 M nodes, 83.9 M DOF), LSRK54 on one MI355X.

A step is one LSRK54 time step = 5 fused stages (dq = a*dq + L(q); q += b*dt*dq),
each one launch of the fused DG stage kernel.  DOF-updates/s counts
Ns x nodes x stage evaluations per second (inputs resident in HBM).

Prints ONE JSON line.  --gpus N with N > 1 exits with an explanation: this
build runs on one GPU.
"""
import argparse
import json
import math
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "raven.jl_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published
NQ, NCELL, L = 8, 32, 10.0
GAMMA = 1.4
KERNEL_NAME = {"f64": "rv::dg_plane3_kernel<8, Euler, double, affine, Rusanov, streaming>",
               "f32": "rv::dg_plane3_kernel<8, Euler, float, affine, Rusanov, streaming>"}
# rocprofv3 counter summary of this workload's stage kernel (HBM bytes per
# dispatch); read for roofline.traffic, which stays null without it
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "dg_plane3_kernel_euler3d_32cube_nq8_traffic.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(dtype):
    """HBM bytes per stage-kernel launch from the committed counter summary:
    FETCH_SIZE and WRITE_SIZE per dispatch (KiB), each scaled by the factor the
    same counter shows for a plain device copy of known size, averaged over
    the five stages of an LSRK54 step.  None when no summary is present."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            text = f.read()
    except OSError:
        return None
    m = re.search(r"```json\n(.*?)\n```", text, re.S)
    if not m:
        return None
    entry = json.loads(m.group(1)).get(dtype)
    return None if entry is None else float(entry["traffic_bytes_per_launch"])


def c2_points(R):
    """Node coordinates of the 32^3 brick (library generator, host only)."""
    cell = R.LobattoCell(NQ, NQ, NQ)
    gm = R.GridManager(cell, R.brick([np.linspace(0, L, NCELL + 1)] * 3, (True,) * 3))
    desc = gm.brick_desc()
    from ravenhip import _lib as RL
    K = RL.lib().rv_brick_numcells(RL.ctypes.byref(desc))
    pts = np.zeros((K, 3, NQ ** 3))
    RL.check(RL.lib().rv_brick_arrays(RL.ctypes.byref(desc), RL.ptr(pts, RL.ctypes.c_double),
                                      None, None, None, None, None, None, None))
    return gm, pts


def c2_initial_state(pts, seed=1234):
    """Isentropic vortex (test/balancelaws/euler/isentropicvortex.jl:18-34)
    extruded in z, centred at (5, 5), plus seeded uniform density noise 1e-3."""
    import problems as PB
    rng = np.random.default_rng(seed)
    noise = 1e-3 * rng.uniform(-1.0, 1.0, pts[:, 0, :].shape)
    return PB.isentropic_vortex_3d(pts, 0.0, gamma=GAMMA, centre=(5.0, 5.0), noise=noise)


def lsrk54():
    a = [0.0, -567301805773 / 1357537059087, -2404267990393 / 2016746695238,
         -3550918686646 / 2091501179385, -1275806237668 / 842570457699]
    b = [1432997174477 / 9575080441755, 5161836677717 / 13612068292357,
         1720146321549 / 2090206949498, 3134564353537 / 4481467310338,
         2277821191437 / 14882151754819]
    return a, b


def run_gpu(R, torch, gm, q0, dtype, steps, warmup, dt, copy_probe=False):
    law = R.EulerLaw(3, GAMMA)
    dg = R.DGSEM.brick(gm, law, dtype=dtype)
    tdt = dg.torch_dtype
    q = torch.from_numpy(q0).to(tdt).cuda().reshape(dg.K, dg.ns, dg.Np).contiguous()
    alt = torch.empty_like(q)
    dq = torch.zeros_like(q)
    a, b = lsrk54()
    if dtype == "f32":
        a = [float(np.float32(v)) for v in a]
        bdt = [float(np.float32(np.float32(v) * np.float32(dt))) for v in b]
    else:
        bdt = [v * dt for v in b]
    bufs = [q, alt]
    cur = [0]

    def step(evs=None):
        for s in range(5):
            if evs is not None:
                evs[s][0].record()
            dg.lsrk_stage(bufs[1 - cur[0]], bufs[cur[0]], dq, a[s], bdt[s])
            if evs is not None:
                evs[s][1].record()
            cur[0] = 1 - cur[0]

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    # timed region: exactly `steps` steps, events around every launch
    evs = [[[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(5)]
           for _ in range(steps)]
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    for k in range(steps):
        step(evs[k])
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    elapsed = start.elapsed_time(end) / 1e3
    durs = np.array([[evs[k][s][0].elapsed_time(evs[k][s][1]) / 1e3 for s in range(5)]
                     for k in range(steps)])
    info = dg.info
    qf = bufs[cur[0]]
    # single right-hand side dg(dq, q, t; increment = false), 5 launches
    scratch = bufs[1 - cur[0]]
    r0, r1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    r0.record()
    for _ in range(5):
        dg(scratch, qf, 0.0, increment=False)
    r1.record()
    torch.cuda.synchronize()
    rhs_ms = r0.elapsed_time(r1) / 5
    # the measured copy peak: a plain device copy of one state array (read
    # and write state_bytes), best of three; profilers also use this copy to
    # scale HBM traffic counters
    copy_gbs = 0.0
    for _ in range(3 if copy_probe or dtype == "f64" else 0):
        c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        c0.record()
        bufs[1 - cur[0]].copy_(qf)
        c1.record()
        torch.cuda.synchronize()
        copy_gbs = max(copy_gbs, 2 * info.state_bytes / (c0.elapsed_time(c1) * 1e-3) / 1e9)
    ints = dg.integrals(qf)
    res = dict(dg=dg, elapsed=elapsed, wall=wall, durs=durs, info=info, rhs_ms=rhs_ms,
               copy_gbs=copy_gbs,
               q=qf.double().cpu().numpy().reshape(-1), integrals=ints,
               integrals0=None)
    # conservation reference: integrals of the initial state
    q0d = torch.from_numpy(q0).to(tdt).cuda().reshape(dg.K, dg.ns, dg.Np).contiguous()
    res["integrals0"] = dg.integrals(q0d)
    del q0d
    return res


def cpu_baseline(q0, dt, budget_s):
    """The CPU oracle (a C restatement of Raven's CPU path, OpenMP) on the
    same 32^3 Nq=8 grid and state, timed for whole LSRK54 stages."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    os.environ["OMP_NUM_THREADS"] = str(cores)
    import problems as PB
    g = O.BrickGrid(NQ, [PB.julia_range(0, 10, NCELL + 1)] * 3, (True,) * 3)
    _ORACLE_GRID["grid"], _ORACLE_GRID["cores"] = g, cores
    dg = O.DG(g, "euler", gamma=GAMMA)
    dg.set_threads(cores)
    q = q0.copy()
    dq = np.zeros_like(q)
    t0 = time.perf_counter()
    dg.rhs(dq, q, 0.0, increment=True)   # one stage's right-hand side
    rhs_s = time.perf_counter() - t0
    dofs = g.K * g.Np * 5
    stages, t_total = 1, rhs_s
    sample = "one right-hand side of the full 32^3 Nq=8 FP64 workload"
    if 5 * rhs_s * 1.3 < budget_s:
        # whole LSRK54 steps until about a third of the budget is spent
        q = q0.copy()
        dq[:] = 0
        nsteps, t_total = 0, 0.0
        while nsteps == 0 or (t_total < budget_s / 3 and t_total * (nsteps + 1) / nsteps < budget_s):
            t0 = time.perf_counter()
            dg.lsrk_step(q, dq, nsteps * dt, dt)
            t_total += time.perf_counter() - t0
            nsteps += 1
        stages = 5 * nsteps
        sample = (f"{nsteps} LSRK54 step(s) ({stages} stages) of the full 32^3 Nq=8 FP64 "
                  "workload")
    return {"value": dofs * stages / t_total, "unit": "DOF-updates/s", "cores": cores,
            "kind": "port", "sample": sample, "seconds": round(t_total, 3),
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "note": "C restatement of Raven's CPU path (oracle/), not Raven itself"}


def cpu_model():
    """The host CPU's model name (what lscpu reports), from /proc/cpuinfo."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


_ORACLE_GRID = {}


def cpu_baseline_fd(q0, dt):
    """The CPU oracle on the flux-differencing and relaxation workloads of the
    same grid and state (after cpu_baseline, whose oracle grid it reuses): one
    LSRK54 stage of FluxDifferencingForm(EntropyConservativeFlux()) with
    Rusanov faces, and one RLSRK54 step with the entropy-conservative pairing."""
    import oracle as O
    g, cores = _ORACLE_GRID["grid"], _ORACLE_GRID["cores"]
    dofs = g.K * g.Np * 5
    fd = O.DG(g, "euler", gamma=GAMMA, volume_flux="ec")
    fd.set_threads(cores)
    q, dq = q0.copy(), np.zeros_like(q0)
    t0 = time.perf_counter()
    fd.rhs(dq, q, 0.0, increment=True)
    fd_s = time.perf_counter() - t0
    out_fd = {"ms_per_stage": fd_s * 1e3, "dof_updates_per_s": dofs / fd_s, "cores": cores,
              "kind": "port", "sample": "one right-hand side (a stage's work) of the full "
                                       "32^3 Nq=8 FP64 flux-differencing workload"}
    del fd
    rl = O.DG(g, "euler", gamma=GAMMA, volume_flux="ec", surface_flux="ec")
    rl.set_threads(cores)
    q = q0.copy()
    n = q.size
    import ctypes
    gam = ctypes.c_double(1.0)
    t0 = time.perf_counter()
    O.lib().orc_dg_rlsrk_step(rl.h, 54, q, np.zeros(n), np.zeros(n), np.zeros(n), 0.0, dt,
                              ctypes.byref(gam))
    rl_s = time.perf_counter() - t0
    out_rl = {"ms_per_step": rl_s * 1e3, "cores": cores, "kind": "port",
              "sample": "one RLSRK54 step of the full 32^3 Nq=8 FP64 workload, "
                        "entropy-conservative volume and surface fluxes"}
    return out_fd, out_rl


def run_batched(R, torch, n_inst=256, steps=10, seed=2024):
    """256 independent 2-D advection / Euler solves (degree 3-7, n x n periodic
    cells with n uniform in 8..64, FP64 or FP32, seeded velocities and vortex
    centres) advanced together: the launch layout picked by the batch
    autotuner, each LSRK54 step a replayed hipGraph.  Compared with the same
    solves run one after the other through the plain per-operator path
    (rv_lsrk_run)."""
    import ravenhip.serve as S
    rng = np.random.default_rng(seed)
    reqs = [S.SolveRequest(rid=k, arrival=0.0, dim=2,
                           law=str(rng.choice(["advection", "euler"])),
                           nq=int(rng.integers(3, 8)) + 1, n=int(rng.integers(8, 65)),
                           T=1.0, dtype=str(rng.choice(["f64", "f32"])),
                           seed=int(rng.integers(1 << 30))) for k in range(n_inst)]
    ops = [S.make_operator(r) for r in reqs]
    dgs = [op[1] for op in ops]
    q0 = [S.initial_state(r, gm, dg) for r, (gm, dg) in zip(reqs, ops)]
    dts = [r.dt() for r in reqs]
    dofs = sum(r.dofs for r in reqs)
    batch = R.Batch(dgs)
    qs = [q.clone() for q in q0]
    dqs = [torch.empty_like(q) for q in q0]
    # launch layout chosen by timing (each kind's waves per block, the mixed
    # launch's waves per block, the own-launch threshold and the number of
    # concurrent streams)
    tuned = batch.autotune(qs, dts)  # {layout: (median, min, max) ms per step}
    layout = min(tuned, key=lambda k: tuned[k][0])
    batch.lsrk_run(qs, dqs, dts, steps)  # builds and caches the step graphs
    for q, q_init in zip(qs, q0):
        q.copy_(q_init)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    batch.lsrk_run(qs, dqs, dts, steps)
    ev[1].record()
    torch.cuda.synchronize()
    t_batch = ev[0].elapsed_time(ev[1]) / 1e3
    finite = all(bool(torch.isfinite(q).all()) for q in qs)
    # one after the other: solve! per instance, (steps - 1/4)*dt so that each
    # takes exactly `steps` steps (the last one shortened)
    solvers = [R.odesolvers.LSRK(dg, q, dt, 0.0, 54) for dg, q, dt in zip(dgs, q0, dts)]
    qs2 = [q.clone() for q in q0]
    R.solve_(qs2[0], (steps - 0.25) * dts[0], solvers[0])  # warm-up
    qs2[0].copy_(q0[0])
    solvers[0].time = 0.0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nsteps = 0
    for q, dt, sv in zip(qs2, dts, solvers):
        nsteps += R.solve_(q, (steps - 0.25) * dt, sv)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    info = batch.info
    batch.close()
    # every member's batched state against its plain stage-by-stage run
    # (rv_lsrk_stage per stage, `steps` whole steps): bit for bit
    same = 0
    for r, dg, q_init, dt, qb in zip(reqs, dgs, q0, dts, qs):
        qp = q_init.clone()
        sv = R.LSRK54(dg, qp, dt)
        tmp = torch.empty_like(qp)
        for _ in range(steps):
            R.dostep_(qp, sv, tmp)
        torch.cuda.synchronize()
        same += int(torch.equal(qp, qb))
    # algorithmic HBM bytes of one LSRK54 step of every member (stage 1 does
    # not read dq), per second of the batched run
    step_bytes = sum(dg.info.stage_bytes_first + 4 * dg.info.stage_bytes for dg in dgs)
    gbs = step_bytes * steps / t_batch / 1e9
    return {
        "workload": f"{n_inst} independent 2-D periodic brick solves (advection and Euler, "
                    "degree 3-7, 8^2..64^2 cells, FP64 and FP32), LSRK54",
        "instances": n_inst, "steps": steps, "dof": int(dofs),
        "batched_dof_updates_per_s": dofs * 5 * steps / t_batch,
        "batched_ms_per_step": t_batch / steps * 1e3,
        "sequential_dof_updates_per_s": dofs * 5 * steps / t_seq,
        "sequential_ms_per_step": t_seq / steps * 1e3,
        "speedup_vs_sequential": t_seq / t_batch,
        "sequential_steps_match": nsteps == steps * n_inst,
        "states_equal_to_plain_stepping": f"{same}/{n_inst}",
        "finite": finite,
        "algorithmic_bytes_per_step": step_bytes,
        "batched_achieved_GBs": gbs, "batched_frac_of_hbm_peak": gbs / HBM_PEAK_GBS,
        "work_items": info["items"], "persistent_waves": info["waves"],
        "launch": "per stage: one launch per kind large enough to fill the device, one "
                  "persistent work-list launch for the small kinds, these chains spread "
                  "over concurrent streams; one hipGraph per step",
        "autotuned_layout": {"mixed_waves_per_block": layout[0],
                             "own_launch_items": layout[1],
                             "streams": layout[2],
                             "step_ms_by_layout": {f"{w},{t},{s}": round(v[0], 4)
                                                    for (w, t, s), v in tuned.items()},
                             "step_ms_spread_by_layout": {
                                 f"{w},{t},{s}": [round(v[1], 4), round(v[2], 4)]
                                 for (w, t, s), v in tuned.items()},
                             "timed_runs_per_layout": 3},
    }


def run_fluxdiff(R, torch, gm, q0, dt, steps=4):
    """The C2 problem (32^3 Nq=8 Euler) with FluxDifferencingForm and the
    entropy-conservative two-point flux, FP64 and FP32: LSRK54 steps timed
    with events.  The form is compute-bound (3 Nq two-point fluxes per node,
    two logarithms each), so it reports two-point flux evaluations per second
    next to DOF-updates/s."""
    out = {"workload": "32^3 Nq=8 3-D Euler, FluxDifferencingForm(EntropyConservativeFlux()), "
                       "Rusanov surface flux, LSRK54", "steps": steps}
    for dtype in ("f64", "f32"):
        dg = R.DGSEM.brick(gm, R.EulerLaw(3, GAMMA), dtype=dtype,
                           volume_form=R.FluxDifferencingForm(R.EntropyConservativeFlux()))
        q = torch.from_numpy(q0).to(dg.torch_dtype).cuda().reshape(dg.K, dg.ns, dg.Np)
        q = q.contiguous()
        solver = R.LSRK54(dg, q, dt)
        R.solve_(q, dt * 0.75, solver)  # warm-up: one step
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        n = R.solve_(q, solver.time + (steps - 0.25) * dt, solver)
        ev[1].record()
        torch.cuda.synchronize()
        sec = ev[0].elapsed_time(ev[1]) / 1e3
        nodes = dg.K * dg.Np
        stages = 5 * n
        out[dtype] = {"ms_per_stage": sec / stages * 1e3,
                      "dof_updates_per_s": nodes * dg.ns * stages / sec,
                      "two_point_fluxes_per_s": nodes * 3 * NQ * stages / sec,
                      "finite": bool(torch.isfinite(q).all())}
        del dg, q, solver
    out["fp32_speedup"] = out["f32"]["dof_updates_per_s"] / out["f64"]["dof_updates_per_s"]
    return out


def _rlsrk_case(R, torch, dg, q0, dt, steps):
    """RLSRK54 against LSRK54 on one handle from one state (see run_rlsrk)."""
    qi = torch.from_numpy(q0).to(dg.torch_dtype).cuda().reshape(dg.K, dg.ns, dg.Np)
    qi = qi.contiguous()
    eta0 = R.entropyintegral(dg, qi)
    q = qi.clone()
    solver = R.RLSRK54(dg, q, dt)
    R.dostep_(q, solver)  # first step untimed (allocations)
    torch.cuda.synchronize()
    iters = []
    t0 = time.perf_counter()
    for _ in range(steps):
        R.dostep_(q, solver)
        iters.append(solver.newton_iterations)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    d_eta = R.entropyintegral(dg, q) - eta0
    q.copy_(qi)
    plain = R.LSRK54(dg, q, dt)
    R.solve_(q, dt * 0.75, plain)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    n = R.solve_(q, plain.time + (steps - 0.25) * dt, plain)
    torch.cuda.synchronize()
    sec_plain = time.perf_counter() - t1
    d_eta_plain = R.entropyintegral(dg, q) - eta0
    return {"ms_per_step": sec / steps * 1e3,
            "ms_per_lsrk54_step": sec_plain / n * 1e3,
            "relaxation_overhead": (sec / steps) / (sec_plain / n) - 1.0,
            "newton_iterations_per_step": float(np.mean(iters)),
            "gamma": solver.gamma,
            "eta0": eta0,
            "delta_eta": d_eta,
            "delta_eta_lsrk54": d_eta_plain,
            "finite": bool(torch.isfinite(q).all())}


def run_rlsrk(R, torch, gm, q0, dt, steps=4):
    """RLSRK54 (relaxation LSRK, odesolvers.jl:72-162) on the C2 problem, FP64
    and FP32: with the entropy-conservative pairing (EC two-point flux in the
    volume and on the faces), and with the weak form and Rusanov faces (the
    headline operator).  Wall time per step (each step reads back one pair of
    sums per Newton iteration) next to the same handle's plain LSRK54 step,
    Newton iterations per step, and the entropy change eta(t) - eta(0) of both
    runs over the same steps from the same state.  The isentropic vortex has
    p = rho^gamma up to the density noise, so eta(0) itself is a small number;
    the changes are reported as they are, not relative to it."""
    out = {"workload": "32^3 Nq=8 3-D Euler, FluxDifferencingForm(EntropyConservativeFlux()), "
                       "EntropyConservativeFlux surface flux, RLSRK54 (fused relaxation stages)",
           "steps": steps + 1}
    for dtype in ("f64", "f32"):
        ec = R.EntropyConservativeFlux()
        dg = R.DGSEM.brick(gm, R.EulerLaw(3, GAMMA), surface_numericalflux=ec, dtype=dtype,
                           volume_form=R.FluxDifferencingForm(ec))
        out[dtype] = _rlsrk_case(R, torch, dg, q0, dt, steps)
        del dg
    out["weak_form"] = {"workload": "the same brick, WeakForm(), Rusanov surface flux, RLSRK54 "
                                    "(right-hand side, then a separate update pass per stage)"}
    for dtype in ("f64", "f32"):
        dg = R.DGSEM.brick(gm, R.EulerLaw(3, GAMMA), dtype=dtype)
        out["weak_form"][dtype] = _rlsrk_case(R, torch, dg, q0, dt, steps)
        del dg
    return out


def run_adaptive(R, torch, steps=200, nq=6, seed=99):
    """3-D Euler on a 16^3 brick over [0,10]^3 with a seeded smooth warp,
    refined 2:1-balanced to level 2 inside the sphere of radius 2 at the
    centre, degree 5: FP32 state and flux arithmetic (FP64 metrics) against
    the all-FP64 run after `steps` LSRK54 steps, with hanging faces coupled
    through mortars, and Raven's self-mapped hanging faces for comparison."""
    import problems as PB
    warp, rng = PB.c3_warp(seed)
    gm = PB.c3_gridmanager(R, nq=nq)
    t0 = time.perf_counter()
    grid = R.generate(gm, warp=warp)
    t_grid = time.perf_counter() - t0
    P = grid.points
    q0 = PB.isentropic_vortex_3d(P, 0.0, gamma=GAMMA, centre=(5.0, 5.0),
                                 noise=1e-3 * rng.uniform(-1, 1, P[:, 0, :].shape))
    hmin = 10 / 16 / 4
    dt = 0.125 * hmin / (nq - 1) / (1.0 + np.sqrt(GAMMA))
    out = {"workload": "3-D Euler, 16^3 periodic brick on [0,10]^3, seeded smooth warp, refined "
                       "2:1-balanced to level 2 inside r=2 at the centre, degree 5, isentropic "
                       "vortex + 1e-3 density noise, LSRK54",
           "cells": int(grid.K), "levels": np.bincount(grid.levels).tolist(),
           "hanging_faces": int(np.sum(grid.face_kind == 1)), "steps": steps, "dt": dt,
           "grid_setup_s": t_grid}
    for mortar in (True, False):
        res = {}
        for dtype in ("f64", "f32"):
            dg = R.DGSEM(law=R.EulerLaw(3, GAMMA), grid=grid,
                         surface_numericalflux=R.RusanovFlux(), dtype=dtype)
            if mortar:
                dg.enable_mortars(grid)
            q = torch.from_numpy(q0).to(dg.torch_dtype).cuda().reshape(dg.K, dg.ns, dg.Np)
            q = q.contiguous()
            i0 = dg.integrals(q)
            solver = R.LSRK54(dg, q, dt)
            R.solve_(q, dt * 2, solver)  # warm-up (two steps), then restart
            q.copy_(torch.from_numpy(q0).to(dg.torch_dtype).reshape(q.shape))
            solver.time = 0.0
            torch.cuda.synchronize()
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            ev[0].record()
            n = R.solve_(q, (steps - 0.25) * dt, solver)
            ev[1].record()
            torch.cuda.synchronize()
            sec = ev[0].elapsed_time(ev[1]) / 1e3
            i1 = dg.integrals(q)
            info = dg.info
            res[dtype] = dict(q=q.double().cpu().numpy().reshape(-1), sec=sec, steps=n,
                              step_bytes=info.stage_bytes_first + 4 * info.stage_bytes,
                              mass_drift=float((i1[0] - i0[0]) / i0[0]),
                              energy_drift=float((i1[4] - i0[4]) / i0[4]))
            del dg, q, solver
        q64, q32 = res["f64"]["q"], res["f32"]["q"]
        key = "mortar" if mortar else "self_mapped_hanging_faces"
        dofs = grid.K * nq ** 3 * 5
        finite = bool(np.isfinite(q64).all() and np.isfinite(q32).all())

        def num(v):
            return float(v) if np.isfinite(v) else None
        out[key] = {
            "diverged": not finite,
            "fp32_rel_l2_vs_fp64": num(np.linalg.norm(q32 - q64) / np.linalg.norm(q64))
            if finite else None,
            "fp64_dof_updates_per_s": dofs * 5 * res["f64"]["steps"] / res["f64"]["sec"],
            "fp32_dof_updates_per_s": dofs * 5 * res["f32"]["steps"] / res["f32"]["sec"],
            "mass_drift_fp64": num(res["f64"]["mass_drift"]),
            "mass_drift_fp32": num(res["f32"]["mass_drift"]),
            "energy_drift_fp64": num(res["f64"]["energy_drift"]),
            "energy_drift_fp32": num(res["f32"]["energy_drift"]),
            "steps_run": res["f64"]["steps"],
        }
        # algorithmic HBM bytes per step (state, FP64 per-node geometry and
        # per-face-node normals and weights; mortar work not counted) over
        # the measured time of the run
        for dtype in ("f64", "f32"):
            gbs = res[dtype]["step_bytes"] * res[dtype]["steps"] / res[dtype]["sec"] / 1e9
            out[key][f"{dtype}_bytes_per_step"] = res[dtype]["step_bytes"]
            out[key][f"{dtype}_achieved_GBs"] = gbs
            out[key][f"{dtype}_frac_of_hbm_peak"] = gbs / HBM_PEAK_GBS
    return out


def run_served(R, torch, rate, n_req=32, seed=7, budget_gb=64):
    """A Poisson request stream (2-D / 3-D advection / Euler, degree 3-7,
    1e3..1e5 cells log-uniform, T in [0.1, 1], FP32 / FP64) served with
    continuous batching, and the same stream served one request at a time.
    Afterwards (outside both timed runs): each request's served result is
    compared with its one-at-a-time result bit for bit, and each FP32
    request's result with the same request solved in FP64 one at a time."""
    import dataclasses

    import ravenhip.serve as S
    served = S.random_requests(n_req, rate, seed=seed)
    server = S.Server(hbm_budget=budget_gb << 30)
    t0 = time.perf_counter()
    server.run(served, keep_results=True)
    wall = time.perf_counter() - t0
    plain = S.random_requests(n_req, rate, seed=seed)
    t0 = time.perf_counter()
    S.run_plain(plain, keep_results=True)
    wall_plain = time.perf_counter() - t0
    a, b = S.summarize(served, wall), S.summarize(plain, wall_plain)
    equal = sum(int(torch.equal(x.result, y.result)) for x, y in zip(served, plain))
    errs = []
    for x in served:
        if x.dtype != "f32":
            continue
        ref = dataclasses.replace(x, dtype="f64", arrival=0.0, result=None, steps=0)
        S.run_plain([ref], keep_results=True)
        q32, q64 = x.result.double(), ref.result
        errs.append(float(torch.linalg.norm(q32 - q64) / torch.linalg.norm(q64)))
        ref.result = None
    for x in served + plain:
        x.result = None
    # algorithmic HBM bytes of the served work: every step of every request,
    # stage 1 reading q (writing dq, q), stages 2-5 also reading dq
    work_bytes = sum(x.steps * x.state_bytes() * (3 + 4 * 4) for x in served)
    return {"rate_per_s": rate if math.isfinite(rate) else "burst", "requests": n_req,
            "hbm_budget_gb": budget_gb,
            "served": a, "one_at_a_time": b,
            "throughput_ratio": a["throughput_dof_updates_per_s"] /
            b["throughput_dof_updates_per_s"],
            "peak_resident": server.peak_active,
            "server_host_s": {k: round(v, 3) for k, v in server.host_time.items()},
            "server_steps": server.steps_run,
            "steps_match": all(x.steps == y.steps for x, y in zip(served, plain)),
            "results_equal_to_one_at_a_time": f"{equal}/{n_req}",
            "fp32_rel_l2_vs_fp64_solve": {"requests": len(errs),
                                          "max": max(errs) if errs else None,
                                          "median": float(np.median(errs)) if errs else None},
            "served_algorithmic_GBs": work_bytes / wall / 1e9}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one plain state copy (calibrates profiler traffic counters)")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the 256-solve batched workload")
    ap.add_argument("--no-served", action="store_true",
                    help="skip the served request stream")
    ap.add_argument("--no-fluxdiff", action="store_true",
                    help="skip the flux-differencing (entropy-conservative) workload")
    ap.add_argument("--no-rlsrk", action="store_true",
                    help="skip the relaxation RK (RLSRK54) workload")
    ap.add_argument("--no-adaptive", action="store_true",
                    help="skip the adaptive warped degree-5 FP32/FP64 workload")
    ap.add_argument("--adaptive-steps", type=int, default=200,
                    help="LSRK54 steps of the adaptive workload (200: the configuration)")
    ap.add_argument("--serve-rates", default="8,32,64,burst",
                    help="request arrival rates (per second) of the served stream; "
                         "'burst': every request arrives at once")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU path is "
              "not built in this repository", file=sys.stderr)
        sys.exit(2)

    import torch
    import ravenhip as R
    assert torch.cuda.is_available(), "bench needs a HIP device"
    torch.cuda.set_device(0)

    log("bench: building the 32^3 Nq=8 brick and initial state")
    gm, pts = c2_points(R)
    q0 = c2_initial_state(pts)
    h = L / NCELL
    # CFL 1/8 on h/N: with grid-scale density noise the 1/4 rule of Raven's
    # tests is outside LSRK54's stability region at N = 7 (the oracle shows
    # exponential growth of the highest modes); 1/8 is stable
    dt = 0.125 * h / (NQ - 1) / (1.0 + np.sqrt(GAMMA))
    K = pts.shape[0]
    nodes = K * NQ ** 3
    dofs = nodes * 5
    del pts

    log(f"bench: FP64 {args.warmup} warmup + {args.steps} timed LSRK54 steps")
    r64 = run_gpu(R, torch, gm, q0, "f64", args.steps, args.warmup, dt, args.copy_probe)
    stages = 5 * args.steps
    value = dofs * stages / r64["elapsed"]
    info = r64["info"]
    # algorithmic bytes: stage 1 (a = 0) reads q, writes dq and q; the others
    # also read dq; affine geometry is per cell
    bytes_step = info.stage_bytes_first + 4 * info.stage_bytes
    kavg = float(np.mean(r64["durs"]))
    achieved = bytes_step / 5 / kavg / 1e9
    out = {
        "metric": "3D Euler p=7 DOF-updates/s + %HBM peak (FP64 vs FP32); served p50/p99 latency",
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["elapsed"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "3-D Euler (5 states, gamma 1.4), periodic 32^3 hex brick on [0,10]^3, "
                        "degree 7 (Nq=8), isentropic vortex + 1e-3 seeded density noise, "
                        "LSRK54, Rusanov flux, weak form",
            "cells": int(K), "nodes": int(nodes), "dof": int(dofs), "dt": dt,
            "step": "one LSRK54 step = 5 fused stage launches",
            "device": torch.cuda.get_device_name(0),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic("fp64"),
            "kernel": KERNEL_NAME["f64"],
            "avg_kernel_ms": kavg * 1e3,
            "single_rhs_ms": r64["rhs_ms"],
            # dg(dq, q, t; increment = false): read q, write dq
            "single_rhs_GBs": 2 * r64["info"].state_bytes / (r64["rhs_ms"] * 1e-3) / 1e9,
            "single_rhs_frac": 2 * r64["info"].state_bytes / (r64["rhs_ms"] * 1e-3) / 1e9
            / HBM_PEAK_GBS,
            "algorithmic_bytes_per_launch": bytes_step / 5,
            # a device copy of one state measured in the same run: the
            # attainable streaming rate on this box
            "measured_copy_GBs": r64["copy_gbs"],
            "frac_of_measured_copy": achieved / r64["copy_gbs"] if r64["copy_gbs"] else None,
        },
    }
    if not args.no_fp32:
        log("bench: FP32 run")
        r32 = run_gpu(R, torch, gm, q0, "f32", args.steps, args.warmup, dt, args.copy_probe)
        v32 = dofs * stages / r32["elapsed"]
        k32 = float(np.mean(r32["durs"]))
        i32 = r32["info"]
        b32 = (i32.stage_bytes_first + 4 * i32.stage_bytes) / 5
        q64, q32 = r64["q"], r32["q"]
        err = float(np.linalg.norm(q32 - q64) / np.linalg.norm(q64))
        m0, m64, m32 = r64["integrals0"], r64["integrals"], r32["integrals"]
        out["fp32"] = {
            "value": v32, "unit": "DOF-updates/s", "speedup_vs_fp64": v32 / value,
            "ms_per_step": r32["elapsed"] / args.steps * 1e3,
            "single_rhs_ms": r32["rhs_ms"],
            "single_rhs_GBs": 2 * i32.state_bytes / (r32["rhs_ms"] * 1e-3) / 1e9,
            "avg_kernel_ms": k32 * 1e3, "achieved_GBs": b32 / k32 / 1e9,
            "frac": b32 / k32 / 1e9 / HBM_PEAK_GBS,
            "kernel": KERNEL_NAME["f32"], "traffic": measured_traffic("fp32"),
            "rel_l2_vs_fp64_after_steps": err,
            "mass_drift_fp64": float((m64[0] - m0[0]) / m0[0]),
            "mass_drift_fp32": float((m32[0] - m0[0]) / m0[0]),
            "energy_drift_fp64": float((m64[4] - m0[4]) / m0[4]),
            "energy_drift_fp32": float((m32[4] - m0[4]) / m0[4]),
            "steps_compared": args.warmup + args.steps,
        }
        del r32
    del r64
    torch.cuda.empty_cache()
    if not args.no_fluxdiff:
        log("bench: flux differencing (entropy-conservative) on the 32^3 Nq=8 brick")
        out["fluxdiff"] = run_fluxdiff(R, torch, gm, q0, dt)
        out["fluxdiff"]["weak_form_ms_per_stage"] = kavg * 1e3
        torch.cuda.empty_cache()
    if not args.no_rlsrk:
        log("bench: relaxation RK (entropy-conservative pairing) on the 32^3 Nq=8 brick")
        out["rlsrk"] = run_rlsrk(R, torch, gm, q0, dt)
        torch.cuda.empty_cache()
    if not args.no_adaptive:
        log("bench: adaptive warped 16^3 degree-5 workload")
        out["adaptive"] = run_adaptive(R, torch, steps=args.adaptive_steps)
        torch.cuda.empty_cache()
    if not args.no_batched:
        log("bench: 256 batched 2-D solves")
        out["batched"] = run_batched(R, torch)
        torch.cuda.empty_cache()
    if not args.no_served:
        out["served"] = []
        for tok in [x.strip() for x in args.serve_rates.split(",") if x.strip()]:
            rate = float("inf") if tok == "burst" else float(tok)
            log(f"bench: served stream at {tok} requests/s")
            out["served"].append(run_served(R, torch, rate))
            torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (oracle)")
        out["cpu_baseline"] = cpu_baseline(q0, dt, args.cpu_budget)
        if "fluxdiff" in out or "rlsrk" in out:
            log("bench: CPU baseline of the flux-differencing and relaxation legs (oracle)")
            cfd, crl = cpu_baseline_fd(q0, dt)
            if "fluxdiff" in out:
                out["fluxdiff"]["cpu_baseline"] = cfd
            if "rlsrk" in out:
                out["rlsrk"]["cpu_baseline"] = crl
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
