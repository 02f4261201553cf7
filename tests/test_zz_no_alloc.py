"""rv_lsrk_run_noalloc: rv_lsrk_run with the ping-pong state in the library's
file-scope device scratch instead of a device allocation (needs an MI355X).

Called through the C ABI (ctypes) on two operators in a row, the second with
a smaller state, each against rv_lsrk_run on the same input: the states, the
final time and the step count must be identical bit for bit, and the handle's
device memory must not grow.  A state beyond the scratch is refused.
"""
import ctypes
import sys

import numpy as np
import pytest

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
    sys.stdout.flush()


def operator(R, dim, nq, n, law, dtype):
    cell = R.LobattoCell(*([nq] * dim))
    coords = [PB.julia_range(0, n / 4, n + 1)] * dim
    grid = R.generate(R.GridManager(cell, R.brick(coords, (True,) * dim)))
    lawobj = R.AdvectionLaw(dim, [0.7, 0.4, 0.2][:dim]) if law == "advection" \
        else R.EulerLaw(dim)
    dg = R.DGSEM(law=lawobj, grid=grid, surface_numericalflux=R.RusanovFlux(), dtype=dtype)
    P = grid.points
    rng = np.random.default_rng(17 + dim + nq)
    if law == "advection":
        q = PB.advection_wave(P, dim, 0.0) + 0.01 * rng.standard_normal(dg.K * dg.Np)
    else:
        c = [n / 8.0, n / 8.0]
        q = PB.isentropic_vortex_3d(P, 0.0, centre=c,
                                    noise=1e-3 * rng.uniform(-1, 1, P[:, 0, :].shape))
    return dg, q


def run(R, torch, dg, q, fn, dt, T):
    qd = torch.from_numpy(q).to(dg.torch_dtype).cuda().reshape(dg.K, dg.ns, dg.Np).contiguous()
    dq = torch.empty_like(qd)
    t = ctypes.c_double(0.0)
    n = ctypes.c_int64(0)
    rc = fn(dg.handle, ctypes.c_void_p(qd.data_ptr()), ctypes.c_void_p(dq.data_ptr()),
            R._lib.RV_LSRK54, 0.0, dt, T, ctypes.byref(t), ctypes.byref(n))
    assert rc == 0, R.lib().rv_last_error(dg.handle)
    dg.synchronize()
    return qd, dq, t.value, n.value


def device_bytes(R, dg):
    info = R._lib.RvInfo()
    assert R.lib().rv_get_info(dg.handle, ctypes.byref(info)) == 0
    return info.device_bytes


def test_noalloc_run_matches_lsrk_run_twice(torch, R):
    lib = R.lib()
    cases = [(3, 6, 4, "euler", "f64", 0.004, 0.0131),   # 1.1 MB state
             (2, 5, 6, "advection", "f32", 0.01, 0.047)]  # smaller, FP32
    for dim, nq, n, law, dtype, dt, T in cases:
        say(f"noalloc: dim={dim} nq={nq} n={n} {law} {dtype}")
        dg, q = operator(R, dim, nq, n, law, dtype)
        before = device_bytes(R, dg)
        qa, dqa, ta, na = run(R, torch, dg, q, lib.rv_lsrk_run_noalloc, dt, T)
        assert device_bytes(R, dg) == before  # no ping-pong buffer was allocated
        qb, dqb, tb, nb = run(R, torch, dg, q, lib.rv_lsrk_run, dt, T)
        assert device_bytes(R, dg) > before   # rv_lsrk_run allocates one
        say(f"   steps {na} / {nb}, t {ta!r} / {tb!r}")
        assert na == nb and na >= 4 and ta == tb
        assert torch.equal(qa, qb) and torch.equal(dqa, dqb)
        assert bool(torch.isfinite(qa).all())


def test_noalloc_refuses_states_beyond_the_scratch(torch, R):
    """A 16 x 16 x 32 brick at Nq = 8, Euler FP64 (168 MB per state), exceeds
    the 32 MiB scratch: RV_E_ARG with a message, and nothing runs."""
    cell = R.LobattoCell(8, 8, 8)
    gm = R.GridManager(cell, R.brick([PB.julia_range(0, 1, 17), PB.julia_range(0, 1, 17),
                                      PB.julia_range(0, 2, 33)], (True,) * 3))
    dg = R.DGSEM.brick(gm, R.EulerLaw(3))
    assert dg.info.state_bytes > R._lib.RV_NOALLOC_SCRATCH_BYTES
    q = dg.GridArray(1.0)
    dq = dg.GridArray(0.0)
    rc = R.lib().rv_lsrk_run_noalloc(dg.handle, ctypes.c_void_p(q.data_ptr()),
                                     ctypes.c_void_p(dq.data_ptr()), R._lib.RV_LSRK54, 0.0,
                                     1e-3, 1e-2, None, None)
    assert rc == R._lib.RV_E_ARG
    assert b"scratch" in R.lib().rv_last_error(dg.handle)
    dg.synchronize()
    assert bool((q == 1.0).all())
