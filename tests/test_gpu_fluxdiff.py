"""FluxDifferencingForm on the GPU (needs an MI355X).

The volume term from a two-point flux (dgsem.jl:514-608): CentralFlux for
advection and Euler, EntropyConservativeFlux for Euler (euler.jl:150-184),
with the Rusanov surface flux.  Bar: FP64 right-hand sides and LSRK stages
within 1e-12 relative L2 of the oracle (itself pinned to Raven's golden
errors for this form, tests/test_oracle.py), FP32 stages within 1e-4 (dq)
and 1e-6 (q_out), and Raven's
golden errors reproduced through the GPU path at the reference tests' own
tolerances.
"""
import json
import os
import sys

import numpy as np
import pytest

import oracle as O
import problems as PB

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
EPS = np.finfo(np.float64).eps
RTOL_F64 = 1e-12


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


def volume_form(R, vflux):
    return R.FluxDifferencingForm(R.EntropyConservativeFlux() if vflux == "ec"
                                  else R.CentralFlux())


def make_pair(R, dim, nq, coords, per, law, vflux, u=None, dtype="f64", geometry="auto"):
    cell = R.LobattoCell(*([nq] * dim))
    grid = R.generate(R.GridManager(cell, R.brick(coords, per)))
    lawobj = R.AdvectionLaw(dim, u) if law == "advection" else R.EulerLaw(dim)
    dg = R.DGSEM(law=lawobj, grid=grid, surface_numericalflux=R.RusanovFlux(), dtype=dtype,
                 geometry=geometry, volume_form=volume_form(R, vflux))
    og = O.BrickGrid(nq, coords, per)
    odg = O.DG(og, law, u=u, volume_flux=vflux)
    return grid, dg, og, odg


def state(og, law, dim, seed=0):
    P = og.point_fields()
    rng = np.random.default_rng(seed)
    if law == "advection":
        return PB.advection_wave(P, dim, 0.0) + 0.01 * rng.standard_normal(og.K * og.Np)
    if dim == 3:
        c = [0.5 * (og.coords[0][0] + og.coords[0][-1]), 0.5 * (og.coords[1][0] + og.coords[1][-1])]
        return PB.isentropic_vortex_3d(P, 0.0, centre=c,
                                       noise=1e-3 * rng.uniform(-1, 1, P[:, 0, :].shape))
    return PB.euler_wave(P, dim, 0.0) + 1e-3 * rng.standard_normal(og.K * og.Np * (dim + 2))


CASES = [
    # dim, nq, law, volume flux, coords, periodic, geometry
    (1, 5, "euler", "ec", [PB.julia_range(-1, 1, 9)], (True,), "auto"),
    (1, 7, "advection", "central", [PB.julia_range(-1, 1, 9)], (False,), "general"),
    (2, 5, "euler", "ec", [PB.julia_range(-1, 1, 6)] * 2, (True, True), "auto"),
    (2, 8, "euler", "central", [PB.julia_range(0, 2, 5), PB.julia_range(0, 1, 3)],
     (True, False), "general"),
    (2, 4, "advection", "central", [PB.julia_range(0, 1, 7)] * 2, (True, True), "auto"),
    (3, 4, "euler", "ec", [PB.julia_range(0, 10, 5)] * 3, (True, True, True), "auto"),
    (3, 8, "euler", "ec", [PB.julia_range(0, 10, 4)] * 3, (True, True, True), "auto"),
    (3, 6, "euler", "ec", [PB.julia_range(0, 10, 4)] * 3, (True, True, False), "general"),
    (3, 5, "advection", "central", [PB.julia_range(-1, 1, 4)] * 3, (True, False, True), "auto"),
    (3, 8, "euler", "central", [PB.julia_range(0, 10, 3)] * 3, (True, True, True), "general"),
]


@pytest.mark.parametrize("dim,nq,law,vflux,coords,per,geometry", CASES)
@pytest.mark.parametrize("increment", [False, True])
def test_fd_rhs_matches_oracle(torch, R, dim, nq, law, vflux, coords, per, geometry, increment):
    say(f"fd rhs dim={dim} nq={nq} {law}/{vflux} geometry={geometry} increment={increment}")
    u = [1.0, 0.5, 0.25][:dim] if law == "advection" else None
    grid, dg, og, odg = make_pair(R, dim, nq, coords, per, law, vflux, u=u, geometry=geometry)
    q = state(og, law, dim)
    rng = np.random.default_rng(1)
    dq0 = rng.standard_normal(q.size) if increment else np.zeros(q.size)
    want = dq0.copy()
    odg.rhs(want, q.copy(), 0.0, increment=increment)
    qd, dqd = to_dev(torch, dg, q), to_dev(torch, dg, dq0)
    dg(dqd, qd, 0.0, increment=increment)
    dg.synchronize()
    err = rel_l2(dqd.cpu().numpy().reshape(-1), want)
    say(f"   rel L2 = {err:.3e}")
    assert err <= RTOL_F64


@pytest.mark.parametrize("dim,nq,vflux,geometry", [(2, 5, "ec", "auto"), (3, 8, "ec", "auto"),
                                                 (3, 6, "central", "auto"),
                                                 (3, 5, "ec", "general")])
def test_fd_stage_matches_oracle_fp64_and_fp32(torch, R, dim, nq, vflux, geometry):
    """One fused LSRK stage: dq = a dq + L(q), q_out = q + b dt dq."""
    coords = [PB.julia_range(0, 10, 4)] * dim
    # FP32: dq is a flux-difference sum with cancellation, rounded at the FP32
    # level as the weak form's (1e-4 there too); q_out = q + b dt dq is tight
    for dtype, tol_dq, tol_q in (("f64", RTOL_F64, RTOL_F64), ("f32", 1e-4, 1e-6)):
        grid, dg, og, odg = make_pair(R, dim, nq, coords, (True,) * dim, "euler", vflux,
                                      dtype=dtype, geometry=geometry)
        q = state(og, "euler", dim, seed=3)
        rng = np.random.default_rng(4)
        dq0 = 1e-2 * rng.standard_normal(q.size)
        a, bdt = -0.7, 3e-3
        L = np.zeros(q.size)
        odg.rhs(L, q.copy(), 0.0, increment=False)
        want_dq = a * dq0 + L
        want_q = q + bdt * want_dq
        qd, dqd = to_dev(torch, dg, q), to_dev(torch, dg, dq0)
        qo = torch.empty_like(qd)
        dg.lsrk_stage(qo, qd, dqd, a, bdt)
        dg.synchronize()
        e1 = rel_l2(dqd.double().cpu().numpy().reshape(-1), want_dq)
        e2 = rel_l2(qo.double().cpu().numpy().reshape(-1), want_q)
        say(f"   stage dim={dim} nq={nq} {vflux} {dtype}: dq {e1:.2e}, q_out {e2:.2e}")
        assert e1 <= tol_dq and e2 <= tol_q


def test_fd_on_brick_handle_matches_oracle(torch, R):
    """The library-generated brick (rv_create_brick) with the EC volume form."""
    nq, coords = 6, [PB.julia_range(0, 10, 5)] * 3
    gm = R.GridManager(R.LobattoCell(nq, nq, nq), R.brick(coords, (True,) * 3))
    dg = R.DGSEM.brick(gm, R.EulerLaw(3), volume_form=volume_form(R, "ec"))
    og = O.BrickGrid(nq, coords, (True,) * 3)
    odg = O.DG(og, "euler", volume_flux="ec")
    q = state(og, "euler", 3, seed=8)
    want = np.zeros(q.size)
    odg.rhs(want, q.copy(), 0.0, increment=False)
    qd, dqd = to_dev(torch, dg, q), dg.GridArray(0.0)
    dg(dqd, qd, 0.0, increment=False)
    dg.synchronize()
    err = rel_l2(dqd.cpu().numpy().reshape(-1), want)
    say(f"   brick handle: rel L2 = {err:.3e}")
    assert err <= RTOL_F64


def test_fd_entropy_conservative_volume_term_conserves(torch, R):
    """On a periodic grid the flux-differencing RHS is conservative: sum MJ dq
    vanishes for every state (a discrete identity of the SBP form)."""
    nq, coords = 5, [PB.julia_range(0, 10, 5)] * 3
    _, dg, og, _ = make_pair(R, 3, nq, coords, (True,) * 3, "euler", "ec")
    q = state(og, "euler", 3, seed=5)
    qd, dqd = to_dev(torch, dg, q), dg.GridArray(0.0)
    dg(dqd, qd, 0.0, increment=False)
    tot = dg.integrals(dqd)
    scale = dg.integrals(torch.abs(dqd).contiguous())
    say(f"   sum MJ dq = {tot}")
    # measured against the largest component: the z momentum of this vortex
    # (extruded in z, w = 0) has dq at rounding level everywhere
    assert np.all(np.abs(tot) <= 1e-12 * scale.max())


def test_fd_rejects_unsupported_combinations(torch, R):
    cell = R.LobattoCell(4, 4)
    grid = R.generate(R.GridManager(cell, R.brick([PB.julia_range(0, 1, 3)] * 2, (True, True))))
    with pytest.raises(R.RavenHIPError):   # EC is defined for Euler only
        R.DGSEM(law=R.AdvectionLaw(2), grid=grid, surface_numericalflux=R.RusanovFlux(),
                volume_form=volume_form(R, "ec"))
    with pytest.raises(R.RavenHIPError):   # faces: Rusanov or the volume flux itself
        R.DGSEM(law=R.EulerLaw(2), grid=grid, surface_numericalflux=R.CentralFlux(),
                volume_form=volume_form(R, "ec"))
    with pytest.raises(R.RavenHIPError):   # the weak form has no EC surface flux
        R.DGSEM(law=R.EulerLaw(2), grid=grid, surface_numericalflux=R.EntropyConservativeFlux())
    dg = R.DGSEM(law=R.EulerLaw(2), grid=grid, surface_numericalflux=R.RusanovFlux(),
                 volume_form=volume_form(R, "ec"))
    with pytest.raises(R.RavenHIPError):   # batches run the weak form only
        R.Batch([dg])


def _fd_golden_cases():
    with open(os.path.join(GOLD, "reference_goldens.json")) as fh:
        entries = json.load(fh)["errors_fluxdifferencing"]
    out = []
    for e in entries:
        for K, want in zip(e["K"], e["expected"]):
            out.append((e["case"], e["volume_flux"], K, want, 10.0 if "1d" in e["case"] else 1.0))
    return out


@pytest.mark.parametrize("case,vflux,K,want,tolscale", _fd_golden_cases())
def test_fd_gpu_reproduces_raven_golden_errors(torch, R, case, vflux, K, want, tolscale):
    say(f"fd golden {case} K={K}")
    s = PB.golden_setup(case, K)
    dim = s["dim"]
    cell = R.LobattoCell(*([5] * dim))
    grid = R.generate(R.GridManager(cell, R.brick(s["coords"], (True,) * dim)))
    law = R.AdvectionLaw(dim, s["u"]) if s["law"] == "advection" else R.EulerLaw(dim)
    dg = R.DGSEM(law=law, grid=grid, surface_numericalflux=R.RusanovFlux(),
                 volume_form=volume_form(R, vflux))
    q = to_dev(torch, dg, s["ic"](grid.points, 0.0))
    solver = R.LSRK54(dg, q, s["dt"])
    R.solve_(q, s["T"], solver)
    qe = to_dev(torch, dg, s["ic"](grid.points, solver.time))
    err = R.weightednorm(dg, q - qe)
    say(f"   err = {err!r} want {want!r}")
    assert err == pytest.approx(want, rel=tolscale * np.sqrt(EPS))
