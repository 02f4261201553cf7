"""Pin the CPU oracle before trusting it (runs on CPU).

Checks come from outside the oracle: Raven's own golden errors and face-map
vectors, 256-bit mpmath fixtures, closed forms and invariants chosen so that a
dropped term, a wrong sign or a transposed operand fails one of them.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
import problems as PB

GOLD = os.path.join(os.path.dirname(__file__), "golden")
EPS = np.finfo(np.float64).eps


def load_json(name):
    with open(os.path.join(GOLD, name)) as fh:
        return json.load(fh)


# --------------------------------------------------------------------------
# GLL operators (lobattocells.jl:1-47; test/cells.jl:13-21,53-62,91-98)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("nq", range(2, 17))
def test_gll_matches_256bit_fixture(nq):
    cell = load_json("gll.json")["cells"][str(nq)]
    x, w, D = O.gll(nq)
    fx = np.array([float.fromhex(v) for v in cell["points_f64"]])
    fw = np.array([float.fromhex(v) for v in cell["weights_f64"]])
    fD = np.array([[float.fromhex(v) for v in row] for row in cell["D_f64"]])
    assert np.array_equal(x, fx)
    assert np.array_equal(w, fw)
    # fixture rows are D[a][b]; the oracle stores column-major D[a + nq*b]
    assert np.array_equal(D.reshape(nq, nq).T, fD)


@pytest.mark.parametrize("nq", range(2, 17))
def test_gll_closed_form_properties(nq):
    cell = load_json("gll.json")["cells"][str(nq)]
    x = np.array([float.fromhex(v) for v in cell["points_f64"]])
    w = np.array([float.fromhex(v) for v in cell["weights_f64"]])
    D = np.array([[float.fromhex(v) for v in row] for row in cell["D_f64"]])
    N = nq - 1
    assert abs(w.sum() - 2) < 20 * EPS           # sum(mass) = 2 (test/cells.jl:13)
    assert np.array_equal(x, -x[::-1])
    # D differentiates polynomials of degree <= N exactly: D x^k = k x^(k-1)
    for k in range(0, N + 1):
        got = D @ x**k
        want = k * x ** max(k - 1, 0) if k > 0 else np.zeros(nq)
        assert np.max(np.abs(got - want)) < 1e3 * nq * nq * EPS * max(1, k)
    # Gauss-Lobatto quadrature is exact to degree 2N-1: int x^k = (1-(-1)^(k+1))/(k+1)
    for k in range(0, 2 * N):
        exact = (1 - (-1) ** (k + 1)) / (k + 1)
        assert abs(w @ x**k - exact) < 50 * EPS
    # summation by parts: W D + (W D)^T = diag(-1, 0, ..., 0, 1)
    Q = np.diag(w) @ D
    B = np.zeros((nq, nq))
    B[0, 0], B[-1, -1] = -1, 1
    assert np.max(np.abs(Q + Q.T - B)) < 1e3 * nq * EPS


def test_tohalves_fixture_interpolates():
    cell = load_json("gll.json")["cells"]["6"]
    x = np.array([float.fromhex(v) for v in cell["points_f64"]])
    lo = np.array([[float.fromhex(v) for v in r] for r in cell["tolowerhalf_f64"]])
    hi = np.array([[float.fromhex(v) for v in r] for r in cell["toupperhalf_f64"]])
    f = lambda s: 3 * s**5 - s**3 + 0.5 * s - 2
    assert np.allclose(lo @ f(x), f((x - 1) / 2), atol=1e-13)
    assert np.allclose(hi @ f(x), f((x + 1) / 2), atol=1e-13)


# --------------------------------------------------------------------------
# LSRK coefficients (odesolvers.jl:163-244)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("scheme,order", [(54, 4), (144, 4)])
def test_lsrk_order_on_linear_ode(scheme, order):
    a, b, c = O.lsrk_coefficients(scheme)

    def step(y, lam, dt):
        dq = 0.0
        for s in range(len(a)):
            dq = a[s] * dq + lam * y
            y = y + b[s] * dt * dq
        return y

    lam = -1.0 + 0.5j
    errs = []
    for dt in (0.2, 0.1):
        errs.append(abs(step(1.0 + 0j, lam, dt) - np.exp(lam * dt)))
    rate = np.log2(errs[0] / errs[1])
    assert rate > order + 0.6  # local error O(dt^(order+1))


@pytest.mark.parametrize("scheme", [54, 144])
def test_lsrk_stage_times_match_quadrature(scheme):
    # for y' = t^k the 2N scheme integrates exactly up to k = 3 when the stage
    # times c are right: a wrong c_s fails here
    a, b, c = O.lsrk_coefficients(scheme)
    for k in range(4):
        dq, y, dt = 0.0, 0.0, 0.7
        for s in range(len(a)):
            dq = a[s] * dq + (c[s] * dt) ** k
            y = y + b[s] * dt * dq
        assert abs(y - dt ** (k + 1) / (k + 1)) < 1e-13


def test_lsrk54_rationals_correctly_rounded():
    from fractions import Fraction
    a, b, c = O.lsrk_coefficients(54)
    assert a[1] == float(Fraction(-567301805773, 1357537059087))
    assert b[4] == float(Fraction(2277821191437, 14882151754819))
    assert c[4] == float(Fraction(2802321613138, 2924317926251))


# --------------------------------------------------------------------------
# Brick grids: order, face maps, numbering, metrics
# --------------------------------------------------------------------------
def test_vmapP_golden_single_cell_periodic():
    g = load_json("reference_goldens.json")["vmapP_periodic_1cell"]
    hexg = O.BrickGrid(2, [[0.0, 1.0]] * 3, [True] * 3)
    assert hexg.vmapP.tolist() == g["hex_nq2"]
    quad = O.BrickGrid(2, [[0.0, 1.0]] * 2, [True] * 2)
    assert quad.vmapP.tolist() == g["quad_nq2"]


def test_brick_tree_order_is_morton():
    g = O.BrickGrid(2, [PB.julia_range(0, 4, 5)] * 2, [False, False])
    xy = g.tree.reshape(-1, 2)
    want = [(0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (3, 0), (2, 1), (3, 1)]
    assert [tuple(v) for v in xy[:8]] == want
    # refined single tree: quadrants in Morton order inside the tree
    g = O.BrickGrid(2, [[0.0, 1.0]] * 3, [False] * 3, level=1)
    assert [tuple(v) for v in g.quad.reshape(-1, 3)][:4] == [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0)]


@pytest.mark.parametrize("dim,per", [(2, (False, False)), (2, (True, False)), (3, (True, True, False)),
                                     (3, (False, False, False)), (1, (True,))])
def test_facemap_invariants(dim, per):
    nq = 4
    coords = [PB.julia_range(-1, 1, 4), PB.julia_range(0, 3, 3), PB.julia_range(2, 3, 3)][:dim]
    g = O.BrickGrid(nq, coords, per)
    nfpt = 2 * dim * g.nfp
    P = g.points.reshape(g.K, dim, g.Np)

    def pt(v):
        v = v - 1
        e, n = v // g.Np, v % g.Np
        return P[e, :, n]

    L = [c[-1] - c[0] for c in coords]
    mapP = g.mapP - 1
    assert np.array_equal(mapP[mapP], np.arange(nfpt * g.K))          # mapP∘mapP = id
    assert np.array_equal(g.vmapP, g.vmapM[mapP])
    xm, xp = pt(g.vmapM), pt(g.vmapP)
    for d in range(dim):
        diff = np.abs(xm[:, d] - xp[:, d])
        if per[d]:
            diff = np.minimum(diff, np.abs(diff - L[d]))
        assert diff.max() < 1e-12
    # boundary codes: 1 exactly on non-periodic outer faces, which self-map
    bc = g.bc.reshape(g.K, 2 * dim)
    for e in range(g.K):
        for f in range(2 * dim):
            slots = np.arange(f * g.nfp, (f + 1) * g.nfp) + nfpt * e
            on_edge = np.all(np.isclose(P[e, f // 2, g.vmapM[slots] - 1 - g.Np * e],
                                        coords[f // 2][-1 if f % 2 else 0]))
            assert bc[e, f] == (1 if (on_edge and not per[f // 2]) else 0)
            if bc[e, f]:
                assert np.array_equal(g.vmapP[slots], g.vmapM[slots])


@pytest.mark.parametrize("dim,per", [(2, (False, False)), (2, (True, True)), (3, (True, False, True))])
def test_dtoc_counts_and_geometry(dim, per):
    nq = 4
    n = [3, 2, 2][:dim]
    coords = [PB.julia_range(0, m, m + 1) for m in n]
    g = O.BrickGrid(nq, coords, per)
    want = 1
    for m, p in zip(n, per):
        want *= m * (nq - 1) + (0 if p else 1)
    assert g.nc == want
    # first-appearance order: ids of element 0 start at 1 and never skip
    seen = 0
    for v in g.dtoc:
        assert v <= seen + 1
        seen = max(seen, v)
    # nodes sharing an id sit at the same place (mod periods)
    P = g.points.reshape(g.K, dim, g.Np).transpose(0, 2, 1).reshape(-1, dim)
    order = np.argsort(g.dtoc, kind="stable")
    ids = g.dtoc[order]
    first = np.concatenate([[True], ids[1:] != ids[:-1]])
    rep = np.maximum.accumulate(np.where(first, np.arange(len(ids)), 0))
    d = np.abs(P[order] - P[order][rep])
    for k in range(dim):
        if per[k]:
            d[:, k] = np.minimum(d[:, k], np.abs(d[:, k] - n[k]))
    assert d.max() < 1e-12


def test_gatherscatter_round_trip():
    nq = 5
    g = O.BrickGrid(nq, [PB.julia_range(0, 3, 4), PB.julia_range(0, 2, 3)], [True, False])
    rng = np.random.default_rng(3)
    nf = 2
    c = rng.standard_normal(g.nc * nf)
    d = np.zeros(g.Np * nf * g.K)
    O.lib().orc_scatter(g.Np, nf, g.K, g.nc, g.dtoc, c, d)
    ones = np.ones(g.Np * nf * g.K)
    deg = np.zeros(g.nc * nf)
    O.lib().orc_gather(g.Np, nf, g.K, g.nc, g.dtoc, ones, deg)
    back = np.zeros(g.nc * nf)
    O.lib().orc_gather(g.Np, nf, g.K, g.nc, g.dtoc, d, back)
    assert np.allclose(back / deg, c, rtol=10 * EPS, atol=0)
    # multiplicities: interior corner of a 2-D brick is shared by 4 elements
    assert deg.max() == 4


def test_brick_metrics_exact_3d():
    """test/grids.jl:650-739 restated for an isotropic cell."""
    nq = 5
    xr, yr, zr = PB.julia_range(-1000, 1000, 21), PB.julia_range(-2000, 2000, 11), PB.julia_range(-3000, 3000, 6)
    sx, sy, sz = 100.0, 400.0, 1200.0
    g = O.BrickGrid(nq, [xr, yr, zr], [False] * 3)
    M = g.metrics.reshape(g.K, 11, g.Np)
    w = g.w1d
    W3 = (w[:, None, None] * w[None, :, None] * w[None, None, :]).transpose(2, 1, 0).reshape(-1)
    assert np.all(M[:, 9, :] == sx * sy * sz / 8)
    assert np.all(M[:, 10, :] == (sx * sy * sz / 8) * W3)
    assert np.all(M[:, 0, :] == 2 / sx) and np.all(M[:, 4, :] == 2 / sy) and np.all(M[:, 8, :] == 2 / sz)
    for f in (1, 2, 3, 5, 6, 7):
        assert np.all(M[:, f, :] == 0)
    S = g.surf.reshape(g.K, 6, g.nfp, 5)
    for f in range(6):
        nvec = np.zeros(3)
        nvec[f // 2] = -1 if f % 2 == 0 else 1
        assert np.all(S[:, f, :, :3] == nvec)
    ww = np.outer(w, w).T.reshape(-1)  # face node m = a + nq*b -> w[a] w[b]
    assert np.allclose(S[:, 0, :, 4], sy * sz / 4 * ww, rtol=10 * EPS, atol=0)
    assert np.allclose(S[:, 2, :, 4], sx * sz / 4 * ww, rtol=10 * EPS, atol=0)
    assert np.allclose(S[:, 5, :, 3], sx * sy / 4, rtol=10 * EPS, atol=0)


def smooth_warp_3d(x, y, z):
    return (x + 0.1 * np.sin(np.pi * x) * np.sin(np.pi * y),
            y + 0.05 * np.sin(2 * np.pi * z) * np.cos(np.pi * x),
            z + 0.1 * np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z))


def multilinear_warp_3d(x, y, z):
    return (x + 0.1 * y * z, y + 0.2 * x * z, z + 0.15 * x * y)


def test_curved_metrics_match_analytic_jacobian():
    """Analytic Jacobian of a smooth map (test/grids.jl:272-360 style).  The map
    is multilinear, so X x dX is a polynomial of degree <= 2 per variable and the
    curl-invariant metrics are exact up to rounding for N = 4."""
    nq = 5
    g = O.BrickGrid(nq, [PB.julia_range(-1, 1, 3)] * 3, [False] * 3, warp=multilinear_warp_3d)
    gb = O.BrickGrid(nq, [PB.julia_range(-1, 1, 3)] * 3, [False] * 3)
    X = gb.points.reshape(gb.K, 3, gb.Np)
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    J = np.zeros(x.shape + (3, 3))
    J[..., 0, 0], J[..., 0, 1], J[..., 0, 2] = 1, 0.1 * z, 0.1 * y
    J[..., 1, 0], J[..., 1, 1], J[..., 1, 2] = 0.2 * z, 1, 0.2 * x
    J[..., 2, 0], J[..., 2, 1], J[..., 2, 2] = 0.15 * y, 0.15 * x, 1
    J = J * 0.5  # element size 1: dx/dxi = 1/2
    M = g.metrics.reshape(g.K, 11, g.Np)
    assert np.max(np.abs(M[:, 9, :] - np.linalg.det(J))) < 1e-14
    inv = np.linalg.inv(J)  # dRdX = (dX/dR)^{-1}, stored column-major per node
    got = M[:, 0:9, :].transpose(0, 2, 1).reshape(g.K, g.Np, 3, 3, order="F")
    assert np.max(np.abs(got - inv)) < 1e-13
    W3 = np.einsum("i,j,k->kji", g.w1d, g.w1d, g.w1d).reshape(-1)
    assert np.max(np.abs(M[:, 10, :] - W3 * M[:, 9, :])) < 1e-15


def test_curved_metrics_free_stream():
    """Sum_d D^T (J dRdX)[d,:] = 0 at every node (test/grids.jl:392-425)."""
    nq = 6
    g = O.BrickGrid(nq, [PB.julia_range(-1, 1, 3)] * 3, [True] * 3, warp=smooth_warp_3d)
    M = g.metrics.reshape(g.K, 11, g.Np)
    D = g.D.reshape(nq, nq, order="F")
    J = M[:, 9, :]
    for c in range(3):
        tot = np.zeros((g.K, nq, nq, nq))
        for d in range(3):
            a = (J * M[:, d + 3 * c, :]).reshape(g.K, nq, nq, nq)  # [e, k, j, i]
            ax = 3 - d
            tot += np.moveaxis(np.tensordot(a, D, axes=([ax], [1])), -1, ax)
        assert np.max(np.abs(tot)) < 200 * EPS  # the reference's own bound


# --------------------------------------------------------------------------
# RHS invariants and Raven's golden errors
# --------------------------------------------------------------------------
def _euler_grid(dim=3, nq=5, warp=None, per=True):
    coords = [PB.julia_range(-1, 1, 4)] * dim
    g = O.BrickGrid(nq, coords, [per] * dim, warp=warp)
    return g, O.DG(g, "euler")


def test_rhs_free_stream_and_conservation():
    g, dg = _euler_grid(warp=smooth_warp_3d)
    q = np.zeros(dg.ns * g.Np * g.K).reshape(g.K, dg.ns, g.Np)
    q[:, 0], q[:, 1], q[:, 2], q[:, 3], q[:, 4] = 1.2, 0.3, -0.2, 0.1, 3.0
    dq = np.zeros(q.size)
    dg.rhs(dq, q.reshape(-1).copy(), 0.0, increment=False)
    assert np.max(np.abs(dq)) < 1e-11
    # random smooth state: sum_e sum_n wJ dq = 0 (periodic, conservative)
    P = g.point_fields()
    q = PB.isentropic_vortex_3d(P, 0.0, centre=(0.1, -0.2))
    dq = np.zeros(q.size)
    dg.rhs(dq, q, 0.0, increment=False)
    MJ = g.MJ
    dq = dq.reshape(g.K, dg.ns, g.Np)
    for s in range(dg.ns):
        assert abs(np.sum(MJ * dq[:, s, :])) < 1e-12 * np.sum(MJ * np.abs(dq[:, s, :]))


def test_rhs_increment_and_linearity_advection():
    g = O.BrickGrid(4, [PB.julia_range(0, 1, 4)] * 2, [True, True])
    dg = O.DG(g, "advection", u=[1.0, 0.5])
    rng = np.random.default_rng(0)
    q1, q2 = rng.standard_normal(g.Np * g.K), rng.standard_normal(g.Np * g.K)
    r1, r2, r12 = (np.zeros(g.Np * g.K) for _ in range(3))
    dg.rhs(r1, q1, increment=False)
    dg.rhs(r2, q2, increment=False)
    dg.rhs(r12, 2 * q1 - q2, increment=False)
    assert np.allclose(r12, 2 * r1 - r2, atol=1e-12)
    base = rng.standard_normal(g.Np * g.K)
    inc = base.copy()
    dg.rhs(inc, q1, increment=True)
    assert np.allclose(inc - base, r1, atol=1e-13)
    # a constant is a steady state: D has zero row sums and the flux balances
    r = np.zeros(g.Np * g.K)
    dg.rhs(r, np.full(g.Np * g.K, 3.25), increment=False)
    assert np.max(np.abs(r)) < 1e-12


def _golden_cases():
    out = []
    for entry in load_json("reference_goldens.json")["errors"]:
        for K, want in zip(entry["K"], entry["expected"]):
            if entry["case"] in ("advection_3d",) and K > 5:
                continue  # 10^3 elements x 98 steps: covered on the GPU
            if entry["case"] == "euler_vortex_2d" and K > 20:
                continue
            out.append((entry["case"], K, want, 10.0 if "1d" in entry["case"] else 1.0))
    return out


def _fd_golden_cases():
    out = []
    for entry in load_json("reference_goldens.json")["errors_fluxdifferencing"]:
        for K, want in zip(entry["K"], entry["expected"]):
            if entry["case"] in ("advection_3d", "euler_wave_3d_onestep") and K > 5:
                continue
            if entry["case"] in ("euler_vortex_2d", "euler_wave_2d", "advection_2d") and K > 10:
                continue
            if "1d" in entry["case"] and K > 20:
                continue
            out.append((entry["case"], entry["volume_flux"], K, want,
                        10.0 if "1d" in entry["case"] else 1.0))
    return out


@pytest.mark.parametrize("case,vflux,K,want,tolscale", _fd_golden_cases())
def test_oracle_fluxdifferencing_reproduces_raven_golden_errors(case, vflux, K, want, tolscale):
    """FluxDifferencingForm (esvolumeterm_dir!, dgsem.jl:514-608) with
    CentralFlux (advection) or EntropyConservativeFlux (Euler, euler.jl:150-184),
    Rusanov surface flux: Raven's own expected errors."""
    s = PB.golden_setup(case, K)
    g = O.BrickGrid(5, s["coords"], [True] * s["dim"])
    dg = O.DG(g, s["law"], u=s["u"], volume_flux=vflux)
    P = g.point_fields()
    q = s["ic"](P, 0.0)
    _, t = dg.solve(q, s["dt"], s["T"])
    err, _ = dg.weightednorm(q - s["ic"](P, t))
    assert err == pytest.approx(want, rel=tolscale * np.sqrt(EPS))


def test_entropy_conservative_flux_is_consistent_and_symmetric():
    """twopointflux(EC) restated: f(q, q) equals Euler's flux F(q) (consistency)
    and f(q1, q2) = f(q2, q1); the log mean switches to its series near a = b.
    Checked through the flux-differencing RHS of a one-cell grid with a constant
    state (zero) and against the central-flux operator on a smooth state."""
    s = PB.golden_setup("euler_wave_2d", 3)
    g = O.BrickGrid(5, s["coords"], [True, True])
    P = g.point_fields()
    q = s["ic"](P, 0.0)
    for vf in ("ec", "central"):
        dg = O.DG(g, "euler", volume_flux=vf)
        r = np.zeros(q.size)
        dg.rhs(r, np.tile(np.repeat([1.2, 0.3, -0.1, 2.6], g.Np), g.K).copy(), increment=False)
        assert np.max(np.abs(r)) < 1e-12
    # EC and central two-point fluxes agree to second order for a smooth field:
    # the difference is far below the RHS itself but not zero
    dge, dgc = O.DG(g, "euler", volume_flux="ec"), O.DG(g, "euler", volume_flux="central")
    re, rc = np.zeros(q.size), np.zeros(q.size)
    dge.rhs(re, q.copy(), increment=False)
    dgc.rhs(rc, q.copy(), increment=False)
    rel = np.linalg.norm(re - rc) / np.linalg.norm(rc)
    assert 1e-12 < rel < 1e-1


@pytest.mark.parametrize("case,K,want,tolscale", _golden_cases())
def test_oracle_reproduces_raven_golden_errors(case, K, want, tolscale):
    s = PB.golden_setup(case, K)
    g = O.BrickGrid(5, s["coords"], [True] * s["dim"])
    dg = O.DG(g, s["law"], u=s["u"])
    P = g.point_fields()
    q = s["ic"](P, 0.0)
    _, t = dg.solve(q, s["dt"], s["T"])
    err, _ = dg.weightednorm(q - s["ic"](P, t))
    assert err == pytest.approx(want, rel=tolscale * np.sqrt(EPS))


# --------------------------------------------------------------------------
# Entropy, two-point surface fluxes and relaxation RK (odesolvers.jl:72-162,
# measures.jl:23-33, numericalfluxes.jl:27-37)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("scheme", [54, 144])
def test_rlsrk_weights_are_the_butcher_weights(scheme):
    """bfull (odesolvers.jl:96-101) are the Butcher b of the 2N scheme: with the
    stage times c they satisfy the quadrature conditions sum b c^(k-1) = 1/k of
    a fourth-order method, and agree with the weights read off the scheme's
    action on y' = f(t) (dq enters y with weight sum_{j>=s} B_j prod A)."""
    a, b, c = (np.asarray(v) for v in O.lsrk_coefficients(scheme))
    bf = O.rlsrk_weights(scheme)
    for k in range(1, 5):
        assert abs(np.sum(bf * c ** (k - 1)) - 1.0 / k) < 1e-13
    # brute force: the weight of stage s's right-hand side in y after all stages
    n = len(a)
    for s in range(n):
        w, carry = 0.0, 1.0
        for j in range(s, n):
            if j > s:
                carry *= a[j]
            w += b[j] * carry
        assert abs(w - bf[s]) < 1e-14


def _entropy_grid(law, dim, vflux, sflux="rusanov", K=3, nq=5):
    coords = [PB.julia_range(-1, 1, K + 1)] * dim
    g = O.BrickGrid(nq, coords, [True] * dim)
    u = [1.0, 0.5, 0.25][:dim] if law == "advection" else None
    return g, O.DG(g, law, u=u, volume_flux=vflux, surface_flux=sflux)


def _smooth_state(g, law, dim, seed=0):
    P = g.point_fields()
    rng = np.random.default_rng(seed)
    if law == "advection":
        return PB.advection_wave(P, dim, 0.0) + 0.05 * rng.standard_normal(g.K * g.Np)
    return PB.euler_wave(P, dim, 0.0) + 1e-2 * rng.standard_normal(g.K * g.Np * (dim + 2))


def test_entropy_integral_closed_form():
    """A uniform Euler state: eta = -rho log(p / rho^gamma) / (gamma - 1) times
    the domain's area; advection: q^2 / 2 times the area."""
    g, dg = _entropy_grid("euler", 2, None)
    rho, m, n, p, gam = 1.3, 0.4, -0.2, 0.9, 1.4
    rhoe = p / (gam - 1) + (m * m + n * n) / (2 * rho)
    q = np.tile(np.repeat([rho, m, n, rhoe], g.Np), g.K).copy()
    want = -rho * np.log(p / rho ** gam) / (gam - 1) * 4.0
    assert dg.entropy_integral(q) == pytest.approx(want, rel=1e-14)
    g, dg = _entropy_grid("advection", 3, None)
    assert dg.entropy_integral(np.full(g.K * g.Np, 1.5)) == pytest.approx(1.5 ** 2 / 2 * 8.0,
                                                                          rel=1e-14)


@pytest.mark.parametrize("law,dim", [("euler", 1), ("euler", 2), ("euler", 3), ("advection", 2)])
def test_entropy_product_is_the_derivative_of_the_integral(law, dim):
    """v(q) = d eta / dq: entropyproduct(q, k) is the directional derivative of
    entropyintegral at q along k (central differences, step 1e-5)."""
    g, dg = _entropy_grid(law, dim, None, K=2, nq=4)
    q = _smooth_state(g, law, dim)
    k = np.random.default_rng(7).standard_normal(q.size)
    h = 1e-5
    fd = (dg.entropy_integral(q + h * k) - dg.entropy_integral(q - h * k)) / (2 * h)
    assert dg.entropy_product(q, k) == pytest.approx(fd, rel=1e-8)


@pytest.mark.parametrize("law,dim,flux", [("euler", 1, "ec"), ("euler", 2, "ec"), ("euler", 3, "ec"),
                                          ("advection", 1, "central"),
                                          ("advection", 2, "central")])
def test_two_point_surface_flux_conserves_entropy_semidiscretely(law, dim, flux):
    """With the entropy-conservative (Euler) or central (advection) two-point
    flux in the volume and on the faces, the periodic semi-discretisation
    conserves entropy: sum MJ v(q) . L(q) vanishes to rounding, while the
    Rusanov faces dissipate it (a strictly negative sum)."""
    g, dg = _entropy_grid(law, dim, flux, sflux=flux)
    q = _smooth_state(g, law, dim, seed=3)
    L = np.zeros(q.size)
    dg.rhs(L, q.copy(), increment=False)
    scale = dg.entropy_product(q, np.abs(L)) if law == "advection" else np.linalg.norm(L)
    prod = dg.entropy_product(q, L)
    assert abs(prod) <= 1e-12 * max(abs(scale), 1.0)
    _, dgr = _entropy_grid(law, dim, flux, sflux="rusanov")
    Lr = np.zeros(q.size)
    dgr.rhs(Lr, q.copy(), increment=False)
    assert dgr.entropy_product(q, Lr) < -1e-6


@pytest.mark.parametrize("law", ["euler", "advection"])
def test_oracle_rlsrk_conserves_entropy_as_ravens_test(law):
    """test/balancelaws/{euler,advection}/entropy_conservation_1d.jl: N = 4,
    K = 5, square pulse, flux differencing with the matching two-point surface
    flux, RLSRK54 to T = 5: |eta(T) - eta(0)| / |eta(0)| <= 40 eps.  The plain
    LSRK54 run of the same set-up misses it by orders of magnitude, so the
    relaxation is what the bound checks."""
    s = PB.entropy_square_setup(law)
    g = O.BrickGrid(s["nq"], s["coords"], [True])
    dg = O.DG(g, law, u=s["u"], volume_flux=s["flux"], surface_flux=s["flux"])
    q = s["ic"](g.point_fields())
    eta0 = dg.entropy_integral(q)
    gammas = []
    n, t, _ = dg.rlsrk_solve(q, s["dt"], s["T"], after_step=lambda *a: gammas.append(a[3]))
    assert t == s["T"] and n == len(gammas)
    d_eta = (dg.entropy_integral(q) - eta0) / abs(eta0)
    assert abs(d_eta) <= 40 * EPS
    assert all(abs(gm - 1) < 1e-2 for gm in gammas)   # relaxation is a small correction
    q = s["ic"](g.point_fields())
    dg.solve(q, s["dt"], s["T"])
    assert abs((dg.entropy_integral(q) - eta0) / abs(eta0)) > 1e-6


# --------------------------------------------------------------------------
# min_node_distance (grids.jl:94-145)
# --------------------------------------------------------------------------
def min_node_dist_warp(dim):
    """test/grids.jl:3-17: stretch by 2 past 1/2 (horizontal) and 3/2 (last)."""
    def stretch(x, at):
        return np.where(x >= at, at + 2 * (x - at), x)

    def warp(*X):
        return tuple(stretch(X[d], 0.5 if d < dim - 1 else 1.5) for d in range(dim))
    return warp


@pytest.mark.parametrize("dim,nq", [(2, 5), (3, 5), (2, 4), (3, 3)])
def test_min_node_distance_closed_form(dim, nq):
    """test/grids.jl:753-790: Kh = 10 cells across [0, 1] horizontally and
    Kv = 4 across [1, 2] vertically, warped; the minimum spacing is the first
    GLL gap over the cell width: Delta xi / (2 Kh) horizontally, Delta xi /
    (2 Kv) vertically (the unstretched cells are the narrow ones)."""
    Kh, Kv = 10, 4
    coords = [PB.julia_range(0, 1, Kh + 1)] * (dim - 1) + [PB.julia_range(1, 2, Kv + 1)]
    g = O.BrickGrid(nq, coords, [True] * dim, warp=min_node_dist_warp(dim))
    dxi = g.x1d[1] - g.x1d[0]
    hmnd, vmnd = dxi / (2 * Kh), dxi / (2 * Kv)
    assert g.min_node_distance() == pytest.approx(hmnd, rel=1e-12)
    assert g.min_node_distance(dims=(dim,)) == pytest.approx(vmnd, rel=1e-12)
    assert g.min_node_distance(dims=range(1, dim)) == pytest.approx(hmnd, rel=1e-12)


# --------------------------------------------------------------------------
# directions (dgsem.jl:32-36, 298-340, 756)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("law,dim,vflux", [("euler", 3, None), ("euler", 2, None),
                                           ("advection", 3, None), ("euler", 3, "ec"),
                                           ("euler", 2, "central")])
def test_directions_split_the_right_hand_side(law, dim, vflux):
    """Volume and surface terms are sums over directions, so the right-hand
    sides of a partition of the directions add up to the full one (to
    rounding), and each part differs from the full one: a mask ignored, or a
    volume or surface term taken from the wrong direction, fails this."""
    coords = [PB.julia_range(-1, 1, 4)] * dim
    g = O.BrickGrid(4, coords, [True] * dim, warp=None)
    u = [1.0, 0.5, 0.25][:dim] if law == "advection" else None
    P = g.point_fields()
    rng = np.random.default_rng(5)
    if law == "advection":
        q = PB.advection_wave(P, dim, 0.0) + 0.05 * rng.standard_normal(g.K * g.Np)
    else:
        q = PB.euler_wave(P, dim, 0.0) + 1e-2 * rng.standard_normal(g.K * g.Np * (dim + 2))
    full = np.zeros(q.size)
    O.DG(g, law, u=u, volume_flux=vflux).rhs(full, q.copy(), increment=False)
    parts = [tuple(range(1, dim)), (dim,)]
    tot = np.zeros(q.size)
    for dirs in parts:
        r = np.zeros(q.size)
        O.DG(g, law, u=u, volume_flux=vflux, directions=dirs).rhs(r, q.copy(), increment=False)
        assert np.linalg.norm(r - full) > 1e-3 * np.linalg.norm(full)
        tot += r
    assert np.linalg.norm(tot - full) <= 1e-12 * np.linalg.norm(full)


@pytest.mark.parametrize("vary", [0, 1, 2])
def test_directions_name_the_varying_coordinate(vary):
    """`directions` counts Raven's directions from 1 = x (dgsem.jl:32-36): for
    an Euler state that varies along one coordinate only, the right-hand side
    restricted to that direction is the whole right-hand side, and the other
    directions' part vanishes to rounding (uniform neighbours' face fluxes
    cancel the weak-form volume term).  A reversed or rotated bit order fails
    both checks, which the split identity alone would not notice."""
    nq = 5
    coords = [PB.julia_range(0, 2, 4)] * 3
    og = O.BrickGrid(nq, coords, (True,) * 3)
    x = og.point_fields()[:, vary, :]
    rho = 2 + 0.3 * np.sin(np.pi * x)
    mom = [rho * 0.4, rho * 0.2, rho * -0.1]
    q = np.ascontiguousarray(np.stack([rho] + mom + [PB.euler_energy(rho, mom, 1.0, 1.4)],
                                      axis=1)).reshape(-1)
    others = tuple(d for d in (1, 2, 3) if d != vary + 1)
    out = {}
    for dirs in [None, (vary + 1,), others]:
        r = np.zeros(q.size)
        O.DG(og, "euler", directions=dirs).rhs(r, q.copy(), increment=False)
        out[dirs] = r
    full = out[None]
    scale = np.linalg.norm(full)
    assert scale > 1.0
    assert np.linalg.norm(out[(vary + 1,)] - full) <= 1e-12 * scale
    assert np.linalg.norm(out[others]) <= 1e-12 * scale
