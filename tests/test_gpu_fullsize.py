"""Parity at the benchmarked sizes (needs an MI355X).

The 32^3 Nq = 8 Euler brick is large enough (3 states > 256 MB) that its
stages take the streaming kernel variant, which loads dq and stores dq and
q_out past L2; the smaller grids of the other parity tests take the cached
variant.  These tests run the variant the benchmark times, and the full
adaptive warped forest, against the oracle: FP64 within 1e-12 relative L2,
FP32 against the FP64 oracle within 1e-4 (FP32 rounding of a flux
divergence).
"""
import sys

import numpy as np
import pytest

import oracle as O
import problems as PB

pytestmark = pytest.mark.gpu

RTOL_F64 = 1e-12
RTOL_F32 = 1e-4


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
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def to_dev(torch, dg, x):
    t = torch.from_numpy(np.ascontiguousarray(x)).to(dg.torch_dtype).cuda()
    return t.reshape(dg.K, dg.ns, dg.Np).contiguous()


def host(t):
    return t.double().cpu().numpy().reshape(-1)


def lsrk54_coefficients(R):
    """rka, rkb of LSRK54 (odesolvers.jl:163-189) as the library holds them."""
    c = R._lib.ctypes
    a, b = np.zeros(14), np.zeros(14)
    n = c.c_int(0)
    R._lib.check(R.lib().rv_lsrk_coefficients(54, R._lib.ptr(a, c.c_double),
                                              R._lib.ptr(b, c.c_double), None, c.byref(n)))
    return [float(v) for v in a[:n.value]], [float(v) for v in b[:n.value]]


@pytest.fixture(scope="module")
def c2(R):
    """The 32^3 periodic brick on [0,10]^3, Nq = 8, in the oracle, with the
    benchmark's initial state (isentropic vortex + 1e-3 density noise)."""
    coords = [PB.julia_range(0, 10, 33)] * 3
    og = O.BrickGrid(8, coords, (True,) * 3)
    odg = O.DG(og, "euler")
    P = og.point_fields()
    rng = np.random.default_rng(1234)
    q = PB.isentropic_vortex_3d(P, 0.0, centre=(5.0, 5.0),
                                noise=1e-3 * rng.uniform(-1, 1, P[:, 0, :].shape))
    gm = R.GridManager(R.LobattoCell(8, 8, 8), R.brick(coords, (True,) * 3))
    dt = 0.125 * (10 / 32) / 7 / (1.0 + np.sqrt(1.4))
    return gm, og, odg, q, dt


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_full_size_c2_stages_match_oracle(torch, R, c2, dtype):
    """The five stages of an LSRK54 step on the full 32^3 Nq = 8 brick, in the
    streaming variant, one rv_lsrk_stage at a time (stage 1 with a = 0,
    stages 2-5 reading dq, every stage writing q_out).  Each stage's
    right-hand side L(q_s) = dq_out - a_s dq_in is compared with the oracle's
    right-hand side of the same stage input q_s (dgsem.jl:84-126): the
    per-stage bar, 1e-12 (FP64); in FP32, 5e-4 against the FP64 oracle (FP32
    rounding of a flux divergence: on this smooth state the right-hand side
    is about three orders smaller than the flux terms it is the divergence
    of, so eps32 ~ 6e-8 becomes ~1e-4 relative to it; the state itself, q_out,
    stays within 1e-7); and q_out with q_s + b_s dt dq_out."""
    gm, og, odg, q, dt = c2
    say(f"full-size C2 stages {dtype}")
    dg = R.DGSEM.brick(gm, R.EulerLaw(3), dtype=dtype)
    assert dg.info.streaming == 1 and dg.info.geometry == 1
    a, b = lsrk54_coefficients(R)
    if dtype == "f32":
        f = np.float32
        a = [float(f(v)) for v in a]
        bdt = [float(f(v) * f(dt)) for v in b]
    else:
        bdt = [v * dt for v in b]
    tol = RTOL_F64 if dtype == "f64" else 5e-4
    qd = to_dev(torch, dg, q)
    dqd = torch.zeros_like(qd)
    q_s, dq_s = host(qd), np.zeros(q.size)
    for s in range(5):
        qo = torch.empty_like(qd)
        dg.lsrk_stage(qo, qd, dqd, a[s], bdt[s])
        dg.synchronize()
        dq_new, q_new = host(dqd), host(qo)
        L_gpu = dq_new - a[s] * dq_s
        L_o = np.zeros(q.size)
        odg.rhs(L_o, q_s.copy(), 0.0, increment=False)
        e_L = rel_l2(L_gpu, L_o)
        e_q = rel_l2(q_new, q_s + bdt[s] * dq_new)
        say(f"   stage {s + 1}: L(q_s) {e_L:.2e}, q_out {e_q:.2e}")
        assert e_L <= tol
        assert e_q <= (1e-15 if dtype == "f64" else 1e-7)
        q_s, dq_s, qd = q_new, dq_new, qo
    del dg, qd, dqd
    torch.cuda.empty_cache()


def test_full_size_c2_step_matches_oracle(torch, R, c2):
    """One whole LSRK54 step through rv_lsrk_run (the streaming variant) on
    the full brick against the oracle's step (odesolvers.jl:51-61), FP64: q
    and the step's increment q1 - q0 within 1e-12; the LSRK register dq is a
    combination of the stages' right-hand sides with cancellation, held to
    1e-11 as in the small-grid step test."""
    gm, og, odg, q, dt = c2
    dg = R.DGSEM.brick(gm, R.EulerLaw(3))
    qo, dqo = q.copy(), np.zeros(q.size)
    odg.lsrk_step(qo, dqo, 0.0, dt)
    qd = to_dev(torch, dg, q)
    solver = R.LSRK54(dg, qd, dt)
    n = R.solve_(qd, dt, solver)
    dg.synchronize()
    assert n == 1
    e_dq = rel_l2(host(solver.dq), dqo)
    e_inc = rel_l2(host(qd) - q, qo - q)
    e_q = rel_l2(host(qd), qo)
    say(f"   whole step: dq {e_dq:.2e}, q1 - q0 {e_inc:.2e}, q1 {e_q:.2e}")
    assert e_q <= RTOL_F64 and e_inc <= RTOL_F64 and e_dq <= 1e-11
    del dg, qd, solver
    torch.cuda.empty_cache()


def test_full_size_c2_rhs_increment_matches_oracle(torch, R, c2):
    """dg(dq, q, t; increment = true) at full size, FP64: the streaming
    variant reads dq and adds the right-hand side (dgsem.jl:342-346)."""
    gm, og, odg, q, dt = c2
    dg = R.DGSEM.brick(gm, R.EulerLaw(3))
    assert dg.info.streaming == 1
    rng = np.random.default_rng(5)
    dq0 = rng.standard_normal(q.size)
    want = dq0.copy()
    odg.rhs(want, q.copy(), 0.0, increment=True)
    qd, dqd = to_dev(torch, dg, q), to_dev(torch, dg, dq0)
    dg(dqd, qd, 0.0, increment=True)
    dg.synchronize()
    err = rel_l2(host(dqd), want)
    say(f"   full-size rhs increment: rel L2 {err:.2e}")
    assert err <= RTOL_F64
    del dg, qd, dqd
    torch.cuda.empty_cache()


@pytest.fixture(scope="module")
def c3(R):
    """The adaptive workload's grid: 16^3 periodic brick on [0,10]^3 with the
    seeded warp, refined 2:1-balanced to level 2 inside r = 2, Nq = 6."""
    warp, rng = PB.c3_warp(99)
    grid = R.generate(PB.c3_gridmanager(R, nq=6), warp=warp)
    P = grid.points
    q = PB.isentropic_vortex_3d(P, 0.0, centre=(5.0, 5.0),
                                noise=1e-3 * rng.uniform(-1, 1, P[:, 0, :].shape))
    return grid, q


class _ArrayGrid:
    """The arrays the oracle's operator reads, taken from a generated grid."""

    def __init__(self, grid, surf=None):
        self.dim, self.nq, self.K, self.Np = grid.dim, grid.nq, grid.K, grid.Np
        self.x1d, self.w1d, self.D = O.gll(grid.nq)
        self.metrics = np.ascontiguousarray(grid.metrics).reshape(-1)
        self.surf = np.ascontiguousarray(grid.surface if surf is None else surf).reshape(-1)
        self.vmapM = np.ascontiguousarray(grid.vmapM).reshape(-1)
        self.vmapP = np.ascontiguousarray(grid.vmapP).reshape(-1)
        self.bc = np.ascontiguousarray(grid.boundarycodes).reshape(-1)
        self.points = grid.points


@pytest.mark.parametrize("mortar", [False, True])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_full_c3_forest_stage_matches_oracle(torch, R, c3, mortar, dtype):
    """One fused LSRK stage (dq = a dq + L(q), q_out = q + b dt dq) on the
    whole adaptive warped forest (13224 cells, Nq = 6), with Raven's
    self-mapped hanging faces or with mortars, against the oracle fed the
    same arrays (plus its mortar restatement)."""
    grid, q = c3
    nq = grid.nq
    say(f"full C3 forest stage: K={grid.K} mortar={mortar} {dtype}")
    assert grid.K == 13224
    dg = R.DGSEM(law=R.EulerLaw(3), grid=grid, surface_numericalflux=R.RusanovFlux(),
                 dtype=dtype)
    assert dg.info.geometry == 2
    surf_full = np.ascontiguousarray(grid.surface).reshape(-1).copy()
    if mortar:
        dg.enable_mortars(grid)
        surf = np.array(grid.surface).reshape(grid.K, 6, nq * nq, 5)
        surf[grid.face_kind != 0, :, 4] = 0.0
        ag = _ArrayGrid(grid, surf)
    else:
        ag = _ArrayGrid(grid)
    odg = O.DG(ag, "euler")
    rng = np.random.default_rng(21)
    dq0 = 1e-3 * rng.standard_normal(q.size)
    a, bdt = -0.6, 2e-3
    L_o = np.zeros(q.size)
    odg.rhs(L_o, q.copy(), 0.0, increment=False)
    if mortar:
        lower, upper = np.zeros(nq * nq), np.zeros(nq * nq)
        R._lib.check(R.lib().rv_lobatto_halves(nq, R._lib.ptr(lower, R._lib.ctypes.c_double),
                                               R._lib.ptr(upper, R._lib.ctypes.c_double)))
        odg.mortar_add(grid.face_kind, grid.mortar_partner, lower, upper, surf_full, q, L_o)
    want_dq = a * dq0 + L_o
    want_q = q + bdt * want_dq
    qd, dqd = to_dev(torch, dg, q), to_dev(torch, dg, dq0)
    qo = torch.empty_like(qd)
    dg.lsrk_stage(qo, qd, dqd, a, bdt)
    dg.synchronize()
    tol = RTOL_F64 if dtype == "f64" else RTOL_F32
    e1 = rel_l2(host(dqd), want_dq)
    e2 = rel_l2(host(qo), want_q)
    say(f"   dq {e1:.2e}, q_out {e2:.2e}")
    assert e1 <= tol and e2 <= tol
    if dtype == "f64":
        e3 = rel_l2(host(qo) - q, want_q - q)
        say(f"   q_out - q {e3:.2e}")
        assert e3 <= tol


def test_self_mapped_forest_diverges_with_the_oracle(torch, R, c3):
    """Raven leaves hanging faces self-mapped (mapP is the identity there,
    lobattocells.jl:2072-2074), so the coarse and fine sides of a hanging face
    never exchange fluxes.  On the adaptive workload (dt = h_min / 8 / N /
    (1 + sqrt(gamma)), as the benchmark runs it) the oracle's LSRK54 run is
    unstable: finite after 32 steps, not after 33.  The library, stepping the
    same grid and state in FP64, follows it step for step (relative L2 within
    1e-9 while both are finite) into the same divergence."""
    grid, q = c3
    nq = grid.nq
    dt = 0.125 * (10 / 16 / 4) / (nq - 1) / (1.0 + np.sqrt(1.4))
    dg = R.DGSEM(law=R.EulerLaw(3), grid=grid, surface_numericalflux=R.RusanovFlux())
    odg = O.DG(_ArrayGrid(grid), "euler")
    qo = np.ascontiguousarray(q).reshape(-1).copy()
    dqo = np.zeros_like(qo)
    qd = to_dev(torch, dg, q)
    qtmp = torch.empty_like(qd)
    solver = R.LSRK54(dg, qd, dt)
    diverged = None
    worst = 0.0
    for n in range(1, 41):
        odg.lsrk_step(qo, dqo, solver.time, dt)
        R.dostep_(qd, solver, qtmp)
        got = host(qd)
        fo, fg = bool(np.isfinite(qo).all()), bool(np.isfinite(got).all())
        if fo and fg:
            err = rel_l2(got, qo)
            worst = max(worst, err)
            assert err <= 1e-9, f"step {n}: rel L2 {err:.2e}"
            continue
        assert fo == fg, f"step {n}: oracle finite {fo}, library finite {fg}"
        diverged = n
        break
    say(f"   self-mapped forest: both diverge at step {diverged}; worst rel L2 before "
        f"{worst:.2e}")
    assert diverged == 33


def test_lsrk144_run_matches_oracle(torch, R):
    """solve!(q, T, LSRK144(dg, q, dt)) through rv_lsrk_run (14 stages, the
    last step shortened to land on T) against the oracle's solve, FP64."""
    coords = [PB.julia_range(0, 10, 5)] * 3
    cell = R.LobattoCell(6, 6, 6)
    grid = R.generate(R.GridManager(cell, R.brick(coords, (True,) * 3)))
    dg = R.DGSEM(law=R.EulerLaw(3), grid=grid, surface_numericalflux=R.RusanovFlux())
    og = O.BrickGrid(6, coords, (True,) * 3)
    odg = O.DG(og, "euler")
    P = og.point_fields()
    rng = np.random.default_rng(3)
    q = PB.isentropic_vortex_3d(P, 0.0, noise=1e-3 * rng.uniform(-1, 1, P[:, 0, :].shape))
    dt, T = 0.01, 0.037
    qo = q.copy()
    n_o, t_o = odg.solve(qo, dt, T, scheme=144)
    qd = to_dev(torch, dg, q)
    solver = R.LSRK144(dg, qd, dt)
    n = R.solve_(qd, T, solver)
    dg.synchronize()
    err = rel_l2(host(qd), qo)
    say(f"   LSRK144 run: {n} steps (oracle {n_o}), rel L2 {err:.2e}")
    assert n == n_o == 4 and abs(t_o - T) < 1e-15
    assert err <= RTOL_F64


@pytest.mark.parametrize("vary", [0, 2])
def test_directions_bit_order(torch, R, vary):
    """`directions` names directions by Raven's 1-based index: for a state
    that varies along one coordinate only, the right-hand side restricted to
    that direction is the whole right-hand side and the other directions'
    part vanishes to rounding (the face fluxes of uniform neighbours cancel
    the weak-form volume term).  A reversed bit order fails both."""
    nq, dim = 5, 3
    coords = [PB.julia_range(0, 2, 4)] * 3
    og = O.BrickGrid(nq, coords, (True,) * 3)
    P = og.point_fields()
    x = P[:, vary, :]
    rho = 2 + 0.3 * np.sin(np.pi * x)
    mom = [rho * 0.4, rho * 0.2, rho * -0.1]
    q = np.ascontiguousarray(np.stack([rho] + mom + [PB.euler_energy(rho, mom, 1.0, 1.4)],
                                      axis=1)).reshape(-1)
    cell = R.LobattoCell(nq, nq, nq)
    grid = R.generate(R.GridManager(cell, R.brick(coords, (True,) * 3)))
    others = tuple(d for d in (1, 2, 3) if d != vary + 1)
    out = {}
    for dirs in [(1, 2, 3), (vary + 1,), others]:
        dg = R.DGSEM(law=R.EulerLaw(3), grid=grid, surface_numericalflux=R.RusanovFlux(),
                     directions=dirs)
        qd = to_dev(torch, dg, q)
        dqd = dg.GridArray(0.0)
        dg(dqd, qd, 0.0, increment=False)
        dg.synchronize()
        out[dirs] = host(dqd)
    full = out[(1, 2, 3)]
    scale = np.linalg.norm(full)
    say(f"   vary {vary}: |L(d)-L| {np.linalg.norm(out[(vary + 1,)] - full) / scale:.2e}, "
        f"|L(others)| {np.linalg.norm(out[others]) / scale:.2e}")
    assert np.linalg.norm(out[(vary + 1,)] - full) <= 1e-12 * scale
    assert np.linalg.norm(out[others]) <= 1e-12 * scale
    assert scale > 1.0


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_gatherscatter_on_the_adaptive_forest(torch, R, c3, dtype):
    """GatherScatter (gatherscatter.jl:1-103) with the forest's continuous
    numbering, hanging nodes included (rv_forest_dtoc): gather and scatter
    bit-exact with the oracle's restatement on the same dtoc (FP64), and
    S(G(x) / deg) = x for a continuous x (mpitest_n3_gatherscatter.jl:12-47)
    in both precisions."""
    grid, _ = c3
    say(f"gather/scatter on the adapted forest: K={grid.K}, nc={grid.ncontinuous} {dtype}")
    dg = R.DGSEM(law=R.EulerLaw(3), grid=grid, surface_numericalflux=R.RusanovFlux(),
                 dtype=dtype)
    Np, K, nc, nf = grid.Np, grid.K, grid.ncontinuous, 5
    dtoc = np.ascontiguousarray(grid.dtoc, dtype=np.int64).reshape(-1)
    tdt = dg.torch_dtype
    rng = np.random.default_rng(8)
    if dtype == "f64":
        src = rng.standard_normal(Np * nf * K)
        want = np.zeros(nc * nf)
        O.lib().orc_gather(Np, nf, K, nc, dtoc, src, want)
        d = torch.zeros(nc * nf, dtype=tdt, device="cuda")
        dg.gather(d, torch.from_numpy(src).cuda(), nf)
        cvec = rng.standard_normal(nc * nf)
        want_s = np.zeros(Np * nf * K)
        O.lib().orc_scatter(Np, nf, K, nc, dtoc, cvec, want_s)
        out = torch.zeros(Np * nf * K, dtype=tdt, device="cuda")
        dg.scatter(out, torch.from_numpy(cvec).cuda(), nf)
        dg.synchronize()
        assert np.array_equal(d.cpu().numpy(), want)
        assert np.array_equal(out.cpu().numpy(), want_s)
    # round trip: x = S(c) is continuous; G(x) = deg * c; S(G(x) / deg) = x
    cvec = torch.from_numpy(rng.uniform(1, 2, nc * nf)).to(tdt).cuda()
    x = torch.zeros(Np * nf * K, dtype=tdt, device="cuda")
    dg.scatter(x, cvec, nf)
    ones = torch.ones(Np * nf * K, dtype=tdt, device="cuda")
    deg = torch.zeros(nc * nf, dtype=tdt, device="cuda")
    dg.gather(deg, ones, nf)
    g = torch.zeros(nc * nf, dtype=tdt, device="cuda")
    dg.gather(g, x, nf)
    back = torch.zeros_like(x)
    dg.scatter(back, (g / deg).contiguous(), nf)
    dg.synchronize()
    eps = torch.finfo(tdt).eps
    err = float(torch.max(torch.abs(back - x)).item())
    say(f"   max |S(G(x)/deg) - x| = {err:.2e}, max degree {int(deg.max().item())}")
    assert err <= 10 * eps * 2


def test_mortars_from_ravens_tables_match(torch, R, c3):
    """rv_enable_mortars fed from Raven's own facemaps tables (vmapNC,
    nctoface through rv_mortars_from_nctables) gives the stage of the
    forest's own face kinds and partners bit for bit."""
    grid, q = c3
    outs = []
    for from_tables in (False, True):
        dg = R.DGSEM(law=R.EulerLaw(3), grid=grid, surface_numericalflux=R.RusanovFlux())
        dg.enable_mortars(grid, from_nctables=from_tables)
        qd = to_dev(torch, dg, q)
        dq = dg.GridArray(0.0)
        qo = torch.empty_like(qd)
        dg.lsrk_stage(qo, qd, dq, 0.0, 1e-3)
        dg.synchronize()
        outs.append((dq, qo))
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("warped", [False, True])
def test_forest_grid_generated_on_the_device(torch, R, warped):
    """rv_create_forest builds the adaptive workload's grid on the device
    (points unless given, curved metrics, geometry): its face maps
    (rv_get_maps) equal the host generator's bit for bit, and its stage and
    GatherScatter equal those of the operator made from the host arrays
    (rv_create, general geometry) bit for bit."""
    gm = PB.c3_gridmanager(R, nq=6)
    warp, rng = PB.c3_warp(99)
    grid = R.generate(gm, warp=warp if warped else None)
    say(f"device-generated forest grid: K={grid.K}, warped={warped}")
    law = R.EulerLaw(3)
    ref = R.DGSEM(law=law, grid=grid, surface_numericalflux=R.RusanovFlux(), geometry="general")
    pts = torch.from_numpy(np.ascontiguousarray(grid.points)).cuda() if warped else None
    dg = R.DGSEM.forest(gm, law, points=pts)
    assert dg.K == grid.K and dg.info.geometry == 2
    vM, vP, _, _ = dg.maps()
    assert np.array_equal(vM, grid.vmapM.reshape(-1)) and np.array_equal(vP, grid.vmapP.reshape(-1))
    P = grid.points
    q = PB.isentropic_vortex_3d(P, 0.0, centre=(5.0, 5.0),
                                noise=1e-3 * rng.uniform(-1, 1, P[:, 0, :].shape))
    outs = []
    for op in (ref, dg):
        qd = to_dev(torch, op, q)
        dq = to_dev(torch, op, 1e-3 * np.ones(q.size))
        qo = torch.empty_like(qd)
        op.lsrk_stage(qo, qd, dq, -0.5, 2e-3)
        g = torch.zeros(grid.ncontinuous * 5, dtype=torch.float64, device="cuda")
        op.gather(g, qd.reshape(-1), 5)
        op.synchronize()
        outs.append((dq, qo, g))
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])
    assert torch.equal(outs[0][2], outs[1][2])
    np.testing.assert_allclose(dg.integrals(outs[1][1]), ref.integrals(outs[0][1]), rtol=0,
                               atol=0)
