"""Batched solves (rv_batch_*): one persistent launch per stage over many
operators of different degree, size, law and precision.  Every member's
result must be bit-identical to what its own operator computes alone, and the
FP64 members must agree with the oracle."""
import numpy as np
import pytest

import oracle as O
import problems as PB

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


@pytest.fixture(scope="module")
def R(torch):
    import ravenhip
    return ravenhip


def say(*a):
    print(*a, flush=True)


def member(R, torch, dim, nq, n, law, dtype, seed):
    """A periodic brick solve on [0, n/4]^dim with a smooth state plus noise."""
    cell = R.LobattoCell(*([nq] * dim))
    coords = [PB.julia_range(0, n / 4, n + 1)] * dim
    grid = R.generate(R.GridManager(cell, R.brick(coords, (True,) * dim)))
    rng = np.random.default_rng(seed)
    u = list(rng.uniform(0.5, 1.5, dim))
    lawobj = R.AdvectionLaw(dim, u) if law == "advection" else R.EulerLaw(dim)
    dg = R.DGSEM(law=lawobj, grid=grid, surface_numericalflux=R.RusanovFlux(), dtype=dtype)
    og = O.BrickGrid(nq, coords, (True,) * dim)
    P = og.point_fields()
    if law == "advection":
        q = PB.advection_wave(P, dim, 0.0, u) + 0.01 * rng.standard_normal(og.K * og.Np)
    elif dim == 2:
        q = PB.euler_wave(P, 2, 0.0) + 1e-3 * rng.standard_normal(og.K * og.Np * 4)
    else:
        c = [0.5 * coords[0][-1]] * 2
        q = PB.isentropic_vortex_3d(P, 0.0, centre=c,
                                    noise=1e-3 * rng.uniform(-1, 1, P[:, 0, :].shape))
    qd = torch.from_numpy(np.ascontiguousarray(q)).to(dg.torch_dtype).cuda()
    qd = qd.reshape(dg.K, dg.ns, dg.Np).contiguous()
    odg = O.DG(og, law, u=u if law == "advection" else None)
    return dg, qd, q, odg


MIX2 = [(3, 6, "advection", "f64"), (4, 5, "euler", "f64"), (5, 7, "advection", "f32"),
        (6, 4, "euler", "f32"), (7, 3, "euler", "f64"), (8, 2, "advection", "f64"),
        (2, 9, "euler", "f32"), (5, 8, "euler", "f64")]
MIX3 = [(4, 3, "euler", "f64"), (5, 2, "advection", "f32"), (6, 2, "euler", "f32"),
        (8, 2, "euler", "f64"), (7, 3, "advection", "f64")]
# kinds with enough items to launch on their own (two members of one of them)
# beside small kinds that share the mixed launch (1024 work-list items launch
# on their own by default).  A work-list item of the 2-D node kernel is one
# wave: one cell at Nq = 8 in FP64 Euler (96^2 and 92^2 cells: 9216 and 8464
# items), 16 cells at Nq = 4 in FP32 advection (130^2: 1057 items), one cell
# at Nq = 6 in FP64 Euler (34^2: 1156 items), whose own launch packs 7 cells
# into a workgroup of 4 waves instead: configuring every kind into the
# work-list launch then lengthens the list.
MIX2H = [(8, 96, "euler", "f64"), (3, 6, "advection", "f64"), (8, 92, "euler", "f64"),
         (4, 130, "advection", "f32"), (5, 7, "advection", "f32"), (6, 34, "euler", "f64")]
# several members of one 3-D kind
MIX3H = [(4, 3, "euler", "f64"), (5, 2, "advection", "f32"), (4, 4, "euler", "f64"),
         (5, 3, "advection", "f32")]


@pytest.mark.parametrize("dim,mix", [(2, MIX2), (3, MIX3), (2, MIX2H), (3, MIX3H)])
def test_batch_stage_equals_members_alone(torch, R, dim, mix):
    say(f"batch stage dim={dim} members={len(mix)}")
    ms = [member(R, torch, dim, nq, n, law, dt, seed) for seed, (nq, n, law, dt) in enumerate(mix)]
    dgs = [m[0] for m in ms]
    batch = R.Batch(dgs)
    say(f"   {batch.info}")
    rng = np.random.default_rng(7)
    q_in = [m[1] for m in ms]
    dq0 = [torch.from_numpy(1e-3 * rng.standard_normal(q.numel())).to(q.dtype).cuda()
           .reshape(q.shape) for q in q_in]
    a = [-0.5 - 0.1 * k for k in range(len(ms))]
    bdt = [1e-3 * (k + 1) for k in range(len(ms))]
    # alone
    want = []
    for dg, q, d0, ak, bk in zip(dgs, q_in, dq0, a, bdt):
        qo, dq = torch.empty_like(q), d0.clone()
        dg.lsrk_stage(qo, q, dq, ak, bk)
        want.append((qo, dq))
    # together, twice (the parameter ring is reused)
    for rep in range(2):
        q_out = [torch.empty_like(q) for q in q_in]
        dq = [d.clone() for d in dq0]
        batch.stage(q_out, q_in, dq, a, bdt)
        torch.cuda.synchronize()
        for k, (qo, d) in enumerate(zip(q_out, dq)):
            assert torch.equal(qo, want[k][0]), f"member {k} q_out differs"
            assert torch.equal(d, want[k][1]), f"member {k} dq differs"
    batch.close()


def test_batch_rhs_stage_matches_oracle(torch, R):
    """A stage with a = 0, b_dt = 0 leaves q and returns L(q): FP64 members
    against the oracle at the FP64 bar."""
    mix = [(4, 5, "euler", "f64"), (7, 3, "advection", "f64"), (5, 6, "euler", "f64")]
    ms = [member(R, torch, 2, nq, n, law, dt, seed) for seed, (nq, n, law, dt) in enumerate(mix)]
    batch = R.Batch([m[0] for m in ms])
    q_in = [m[1] for m in ms]
    q_out = [torch.empty_like(q) for q in q_in]
    dq = [torch.zeros_like(q) for q in q_in]
    batch.stage(q_out, q_in, dq, [0.0] * 3, [0.0] * 3)
    torch.cuda.synchronize()
    for k, (dg, qd, q, odg) in enumerate(ms):
        want = np.zeros(q.size)
        odg.rhs(want, q.copy(), 0.0, increment=False)
        got = dq[k].cpu().numpy().reshape(-1)
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        say(f"   member {k}: rel L2 {err:.2e}")
        assert err <= 1e-12


@pytest.mark.parametrize("dim,mix", [(2, MIX2[:5]), (3, MIX3[:3]), (2, MIX2H), (3, MIX3H)])
def test_batch_graph_run_equals_stage_loop(torch, R, dim, mix):
    """rv_batch_lsrk_run (a captured hipGraph per step, replayed) against each
    member stepping alone through its own stages; odd step counts exercise the
    final copy back from the ping-pong buffer.  dq enters as NaN: LSRK54's
    first stage (a = 0) must not read it."""
    ms = [member(R, torch, dim, nq, n, law, dt, 10 + s) for s, (nq, n, law, dt) in enumerate(mix)]
    dgs = [m[0] for m in ms]
    dts = [2e-3 * (1 + 0.25 * k) for k in range(len(ms))]
    for nsteps in (3, 2):
        want = []
        for dg, q, dt in zip(dgs, [m[1] for m in ms], dts):
            qa = q.clone()
            solver = R.LSRK54(dg, qa, dt)
            tmp = torch.empty_like(qa)
            for _ in range(nsteps):
                R.dostep_(qa, solver, tmp)
            want.append(qa)
        batch = R.Batch(dgs)
        qs = [m[1].clone() for m in ms]
        dqs = [torch.full_like(q, float("nan")) for q in qs]
        batch.lsrk_run(qs, dqs, dts, nsteps)
        # a replay of the cached graphs from the same buffers
        qs2 = qs
        for k in range(len(ms)):
            qs2[k].copy_(ms[k][1])
            dqs[k].fill_(float("nan"))
        batch.lsrk_run(qs2, dqs, dts, nsteps)
        torch.cuda.synchronize()
        for k in range(len(ms)):
            assert torch.equal(qs2[k], want[k]), f"member {k}, {nsteps} steps"
        batch.close()


@pytest.mark.parametrize("dim,mix", [(2, MIX2H), (3, MIX3H)])
def test_batch_step_skips_resting_members(torch, R, dim, mix):
    """rv_batch_step with some members at dt = 0: those keep their states
    bit for bit (their kernels are not launched, or return at once), the
    others match two steps taken alone."""
    ms = [member(R, torch, dim, nq, n, law, dt, 20 + s) for s, (nq, n, law, dt) in enumerate(mix)]
    dgs = [m[0] for m in ms]
    dts = [1.5e-3 * (1 + 0.5 * k) for k in range(len(ms))]
    rest = {1, 2}  # one small and one large member (or two of one 3-D kind)
    want = []
    for k, (dg, q) in enumerate(zip(dgs, [m[1] for m in ms])):
        qa = q.clone()
        if k not in rest:
            solver = R.LSRK54(dg, qa, dts[k])
            tmp = torch.empty_like(qa)
            for _ in range(2):
                R.dostep_(qa, solver, tmp)
        want.append(qa)
    batch = R.Batch(dgs)
    q0 = [m[1].clone() for m in ms]
    q1 = [torch.empty_like(q) for q in q0]
    dq = [torch.zeros_like(q) for q in q0]
    cur = [0] * len(ms)
    for _ in range(2):
        batch.step(q0, q1, dq, cur, [0.0 if k in rest else dts[k] for k in range(len(ms))])
        cur = [c if k in rest else (c + 5) % 2 for k, c in enumerate(cur)]
    torch.cuda.synchronize()
    for k in range(len(ms)):
        got = (q0, q1)[cur[k]][k]
        assert torch.equal(got, want[k]), f"member {k} ({'resting' if k in rest else 'stepping'})"
    batch.close()


def test_batch_layout_does_not_change_results(torch, R):
    """Every launch layout rv_batch_configure offers (all kinds on their own,
    all in the mixed launch, each mixed waves-per-block), each number of
    concurrent streams (rv_batch_set_streams) and the autotuner's choice give
    bit-identical stages."""
    ms = [member(R, torch, 2, nq, n, law, dt, 30 + s) for s, (nq, n, law, dt) in enumerate(MIX2H)]
    dgs = [m[0] for m in ms]
    batch = R.Batch(dgs)
    q_in = [m[1] for m in ms]
    rng = np.random.default_rng(5)
    dq0 = [torch.from_numpy(1e-3 * rng.standard_normal(q.numel())).to(q.dtype).cuda()
           .reshape(q.shape) for q in q_in]
    a = [-0.3 - 0.05 * k for k in range(len(ms))]
    bdt = [2e-3 * (k + 1) for k in range(len(ms))]

    def run():
        q_out = [torch.empty_like(q) for q in q_in]
        dq = [d.clone() for d in dq0]
        batch.stage(q_out, q_in, dq, a, bdt)
        torch.cuda.synchronize()
        return q_out, dq

    want = run()
    for w, t, s in [(1, 1, 1), (4, 1, 4), (1, 1 << 40, 2), (2, 1 << 40, 1), (4, 1 << 40, 3),
                    (2, 1024, 1), (2, 1024, 4), (2, 64, 3)]:
        batch.configure(w, t, s)
        got = run()
        for k in range(len(ms)):
            assert torch.equal(got[0][k], want[0][k]) and torch.equal(got[1][k], want[1][k]), \
                f"layout ({w}, {t}, {s} streams): member {k} differs"
    with pytest.raises(R.RavenHIPError):
        batch.configure(2, 1024, 5)
    times = batch.autotune(q_in, [1e-3] * len(ms))
    say(f"   autotune: {times}")
    got = run()
    for k in range(len(ms)):
        assert torch.equal(got[0][k], want[0][k]) and torch.equal(got[1][k], want[1][k])
    batch.close()


def test_batch_rejects_unsupported_members(torch, R):
    cell = R.LobattoCell(5, 5)
    grid = R.generate(R.GridManager(cell, R.brick([PB.julia_range(0, 1, 4)] * 2, (True, True))))
    dg = R.DGSEM(law=R.EulerLaw(2), grid=grid, surface_numericalflux=R.CentralFlux())
    with pytest.raises(R.RavenHIPError):
        R.Batch([dg])
    cell3 = R.LobattoCell(3, 3, 3)
    g3 = R.generate(R.GridManager(cell3, R.brick([PB.julia_range(0, 1, 3)] * 3, (True,) * 3)))
    dg3 = R.DGSEM(law=R.EulerLaw(3), grid=g3, surface_numericalflux=R.RusanovFlux())
    with pytest.raises(R.RavenHIPError):
        R.Batch([dg3])  # 3-D batches start at Nq = 4
