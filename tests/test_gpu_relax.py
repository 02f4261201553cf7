"""Entropy diagnostics, two-point surface fluxes and relaxation RK on the GPU
(needs an MI355X).

- entropyintegral / entropyproduct (measures.jl:23-33) against the oracle,
  FP64 to 1e-13 relative, FP32 states (FP64 arithmetic) to 1e-13 as well,
  since the device evaluates the FP32 values in FP64 as the oracle does.
- The flux-differencing right-hand side with its own two-point flux on the
  faces (twopointflux(q-, q+)' n, numericalfluxes.jl:27-37) against the
  oracle to 1e-12 relative L2.
- rv_rlsrk_step (dostep! of RLSRK, odesolvers.jl:117-151) step by step
  against the oracle: state to 1e-10 relative L2 and gamma to 1e-10 (the
  root is fixed only to |residual| <= 10 eps, which the two Newton runs reach
  from sums rounded in different orders).
- Raven's entropy-conservation tests (test/balancelaws/{euler,advection}/
  entropy_conservation_1d.jl) through the GPU path: |delta eta| <= 40 eps.
"""
import sys

import numpy as np
import pytest

import oracle as O
import problems as PB

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


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


FLUX = {"ec": "EntropyConservativeFlux", "central": "CentralFlux", "rusanov": "RusanovFlux"}


def make_pair(R, dim, nq, coords, law, vflux, sflux, u=None, dtype="f64", geometry="auto",
              per=None):
    per = per or (True,) * dim
    cell = R.LobattoCell(*([nq] * dim))
    grid = R.generate(R.GridManager(cell, R.brick(coords, per)))
    lawobj = R.AdvectionLaw(dim, u) if law == "advection" else R.EulerLaw(dim)
    vf = None if vflux is None else R.FluxDifferencingForm(getattr(R, FLUX[vflux])())
    dg = R.DGSEM(law=lawobj, grid=grid, surface_numericalflux=getattr(R, FLUX[sflux])(),
                 dtype=dtype, geometry=geometry, volume_form=vf)
    og = O.BrickGrid(nq, coords, per)
    odg = O.DG(og, law, u=u, volume_flux=vflux, surface_flux=sflux)
    return grid, dg, og, odg


def state(og, law, dim, seed=0):
    P = og.point_fields()
    rng = np.random.default_rng(seed)
    if law == "advection":
        return PB.advection_wave(P, dim, 0.0) + 0.05 * rng.standard_normal(og.K * og.Np)
    return PB.euler_wave(P, dim, 0.0) + 1e-2 * rng.standard_normal(og.K * og.Np * (dim + 2))


@pytest.mark.parametrize("law,dim,dtype", [("euler", 1, "f64"), ("euler", 2, "f64"),
                                           ("euler", 3, "f64"), ("advection", 2, "f64"),
                                           ("euler", 3, "f32"), ("advection", 3, "f32")])
def test_entropy_integral_and_product_match_oracle(torch, R, law, dim, dtype):
    say(f"entropy sums {law} dim={dim} {dtype}")
    coords = [PB.julia_range(-1, 1, 4)] * dim
    u = [1.0, 0.5, 0.25][:dim] if law == "advection" else None
    _, dg, og, odg = make_pair(R, dim, 5, coords, law, None, "rusanov", u=u, dtype=dtype)
    q = state(og, law, dim)
    k = np.random.default_rng(2).standard_normal(q.size)
    qd, kd = to_dev(torch, dg, q), to_dev(torch, dg, k)
    # the oracle sees exactly the values the device holds
    qh = qd.double().cpu().numpy().reshape(-1).copy()
    kh = kd.double().cpu().numpy().reshape(-1).copy()
    eta, want = R.entropyintegral(dg, qd), odg.entropy_integral(qh)
    prod, wantp = R.entropyproduct(dg, qd, kd), odg.entropy_product(qh, kh)
    say(f"   eta {eta!r} vs {want!r}; product {prod!r} vs {wantp!r}")
    assert eta == pytest.approx(want, rel=1e-13, abs=1e-13)
    scale = abs(odg.entropy_product(qh, np.abs(kh))) + abs(wantp)
    assert abs(prod - wantp) <= 1e-13 * scale


SURF_CASES = [
    # dim, nq, law, flux (volume and surface), geometry, periodic
    (1, 5, "euler", "ec", "auto", (True,)),
    (1, 6, "advection", "central", "general", (False,)),
    (2, 5, "euler", "ec", "auto", (True, True)),
    (2, 4, "euler", "ec", "general", (True, False)),
    (2, 6, "euler", "central", "auto", (True, True)),
    (2, 5, "advection", "central", "auto", (True, True)),
    (3, 4, "euler", "ec", "auto", (True, True, True)),
    (3, 5, "euler", "ec", "general", (True, False, True)),
    (3, 4, "advection", "central", "general", (True, True, True)),
]


@pytest.mark.parametrize("dim,nq,law,flux,geometry,per", SURF_CASES)
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_two_point_surface_flux_rhs_matches_oracle(torch, R, dim, nq, law, flux, geometry, per,
                                                   dtype):
    say(f"two-point faces dim={dim} nq={nq} {law}/{flux} {geometry} {dtype}")
    coords = [PB.julia_range(-1, 1, 4)] * dim
    u = [1.0, 0.5, 0.25][:dim] if law == "advection" else None
    _, dg, og, odg = make_pair(R, dim, nq, coords, law, flux, flux, u=u, dtype=dtype,
                               geometry=geometry, per=per)
    q = state(og, law, dim, seed=1)
    want = np.zeros(q.size)
    odg.rhs(want, q.copy(), 0.0, increment=False)
    qd, dqd = to_dev(torch, dg, q), dg.GridArray(0.0)
    dg(dqd, qd, 0.0, increment=False)
    dg.synchronize()
    err = rel_l2(dqd.double().cpu().numpy().reshape(-1), want)
    say(f"   rel L2 = {err:.3e}")
    assert err <= (1e-12 if dtype == "f64" else 1e-4)


# fused (the stage kernel forms the entropy production): flux differencing
# with its own flux on the faces, and the 3-D weak form (affine and general
# cells); separate update passes: every other operator
RLSRK_CASES = [
    # law, dim, volume flux, surface flux, scheme, geometry
    ("euler", 2, "ec", "ec", 54, "auto"),
    ("euler", 1, "ec", "ec", 144, "auto"),
    ("advection", 2, "central", "central", 54, "auto"),
    ("euler", 3, "ec", "ec", 54, "auto"),
    ("euler", 2, "ec", "rusanov", 54, "auto"),
    ("euler", 3, None, "rusanov", 54, "auto"),
    ("euler", 3, None, "central", 144, "general"),
    ("advection", 3, None, "rusanov", 54, "auto"),
    ("advection", 1, None, "rusanov", 144, "auto"),
]


@pytest.mark.parametrize("law,dim,flux,sflux,scheme,geometry", RLSRK_CASES)
def test_rlsrk_steps_match_oracle(torch, R, law, dim, flux, sflux, scheme, geometry):
    say(f"rlsrk steps {law} dim={dim} volume {flux} surface {sflux} LSRK{scheme} {geometry}")
    coords = [PB.julia_range(-1, 1, 4)] * dim
    u = [1.0, 0.5, 0.25][:dim] if law == "advection" else None
    _, dg, og, odg = make_pair(R, dim, 5, coords, law, flux, sflux, u=u, geometry=geometry)
    q = state(og, law, dim, seed=4)
    dt = 0.02
    qd = to_dev(torch, dg, q)
    solver = (R.RLSRK54 if scheme == 54 else R.RLSRK144)(dg, qd, dt)
    n = q.size
    dq, k, q0 = np.zeros(n), np.zeros(n), np.zeros(n)
    import ctypes
    g = ctypes.c_double(1.0)
    lib = O.lib()
    for step in range(6):
        lib.orc_dg_rlsrk_step(odg.h, scheme, q, dq, k, q0, step * dt, dt, ctypes.byref(g))
        R.dostep_(qd, solver)
        err = rel_l2(qd.cpu().numpy().reshape(-1), q)
        say(f"   step {step}: gamma {solver.gamma!r} vs {g.value!r}, "
            f"{solver.newton_iterations} Newton iterations, q rel L2 {err:.2e}")
        assert abs(solver.gamma - g.value) <= 1e-10
        assert err <= 1e-10
        assert solver.gamma != 1.0   # the relaxation acted


@pytest.mark.parametrize("law", ["euler", "advection"])
def test_gpu_rlsrk_conserves_entropy_as_ravens_test(torch, R, law):
    """entropy_conservation_1d.jl (euler and advection) through the GPU:
    N = 4, K = 5, RLSRK54 to T = 5, |eta(T) - eta(0)| / |eta(0)| <= 40 eps,
    and the final state against the oracle's run."""
    s = PB.entropy_square_setup(law)
    say(f"entropy conservation 1-D {law}")
    _, dg, og, odg = make_pair(R, 1, s["nq"], s["coords"], law, s["flux"], s["flux"], u=s["u"])
    q = s["ic"](og.point_fields())
    qd = to_dev(torch, dg, q)
    eta0 = R.entropyintegral(dg, qd)
    solver = R.RLSRK54(dg, qd, s["dt"])
    steps = R.solve_(qd, s["T"], solver)
    d_eta = (R.entropyintegral(dg, qd) - eta0) / abs(eta0)
    say(f"   {steps} steps to t = {solver.time!r}: delta eta = {d_eta:.3e} "
        f"({d_eta / EPS:.1f} eps)")
    assert solver.time == s["T"]
    assert abs(d_eta) <= 40 * EPS
    nref, tref, _ = odg.rlsrk_solve(q, s["dt"], s["T"])
    assert nref == steps
    err = rel_l2(qd.cpu().numpy().reshape(-1), q)
    say(f"   final state vs oracle: rel L2 {err:.2e}")
    assert err <= 1e-9


def test_gpu_rlsrk_fp32_entropy_drift(torch, R):
    """The FP32 Euler run of the same set-up: the relaxation holds the entropy
    to the FP32 level (|delta eta| <= 40 eps(Float32)), where plain FP32 LSRK54
    drifts by about 1e-3."""
    s = PB.entropy_square_setup("euler")
    _, dg, og, _ = make_pair(R, 1, s["nq"], s["coords"], "euler", "ec", "ec", dtype="f32")
    qd = to_dev(torch, dg, s["ic"](og.point_fields()))
    eta0 = R.entropyintegral(dg, qd)
    solver = R.RLSRK54(dg, qd, s["dt"])
    R.solve_(qd, s["T"], solver)
    d_eta = (R.entropyintegral(dg, qd) - eta0) / abs(eta0)
    e32 = float(np.finfo(np.float32).eps)
    say(f"   FP32 delta eta = {d_eta:.3e} ({d_eta / e32:.1f} eps32)")
    assert abs(d_eta) <= 40 * e32


def test_rlsrk_rejects_aliased_and_unknown(torch, R):
    s = PB.entropy_square_setup("advection")
    _, dg, og, _ = make_pair(R, 1, s["nq"], s["coords"], "advection", "central", "central",
                             u=s["u"])
    qd = to_dev(torch, dg, s["ic"](og.point_fields()))
    solver = R.RLSRK54(dg, qd, s["dt"])
    solver.k = solver.dq
    with pytest.raises(R.RavenHIPError):
        R.dostep_(qd, solver)
    solver = R.RLSRK54(dg, qd, s["dt"])
    solver.scheme = 77
    with pytest.raises(R.RavenHIPError):
        R.dostep_(qd, solver)
