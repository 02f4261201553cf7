"""Mortar coupling of hanging faces, host only: the oracle's restatement of the
library's mortar (Raven has none in its right-hand side; its hanging faces are
self-mapped, lobattocells.jl:2072-2074) is pinned by what the mathematics fixes:
tohalves_1d against 256-bit values, exact discrete conservation on a periodic
adapted grid (which self-mapped faces break), and exactness for a linear field,
which any wrong sub-face, half or transposition would spoil."""
import json
import os

import numpy as np
import pytest

import oracle as O
import problems as PB
import ravenhip as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("nq", range(2, 9))
def test_tohalves_match_256bit_fixture(nq):
    with open(os.path.join(ROOT, "tests", "golden", "gll.json")) as fh:
        cell = json.load(fh)["cells"][str(nq)]
    lo, hi = np.zeros(nq * nq), np.zeros(nq * nq)
    c = R._lib.ctypes
    R._lib.check(R.lib().rv_lobatto_halves(nq, R._lib.ptr(lo, c.c_double),
                                           R._lib.ptr(hi, c.c_double)))
    want_lo = np.array([[float.fromhex(v) for v in row] for row in cell["tolowerhalf_f64"]])
    want_hi = np.array([[float.fromhex(v) for v in row] for row in cell["toupperhalf_f64"]])
    assert np.array_equal(lo.reshape(nq, nq), want_lo)
    assert np.array_equal(hi.reshape(nq, nq), want_hi)


class ForestGrid:
    """An adapted brick forest with the oracle's curved metrics (host only)."""

    def __init__(self, dim, nq, n, rounds, radius, periodic, warp=None):
        cell = R.LobattoCell(*([nq] * dim))
        coords = [PB.julia_range(0, n, n + 1)] * dim
        gm = R.GridManager(cell, R.brick(coords, (periodic,) * dim))
        centre = np.full(dim, n / 2.0)
        for r in range(rounds):
            cen = gm.cell_centres()
            flags = np.where((np.linalg.norm(cen - centre, axis=1) < radius) &
                             (gm.cells()[1] <= r), R.AdaptRefine, R.AdaptNone)
            gm.adapt(flags)
        self.gm = gm
        self.dim, self.nq = dim, nq
        K = self.K = len(gm)
        Np = self.Np = nq ** dim
        nfp = nq ** (dim - 1)
        nfpt = 2 * dim * nfp
        self.x1d, self.w1d, self.D = O.gll(nq)
        pts = np.zeros((K, dim, Np))
        self.vmapM = np.zeros(K * nfpt, np.int64)
        self.vmapP = np.zeros(K * nfpt, np.int64)
        self.bc = np.zeros(K * 2 * dim, np.int64)
        self.kind = np.zeros(K * 2 * dim, np.int32)
        self.partner = np.zeros(K * 2 * dim * 4, np.int64)
        c = R._lib.ctypes
        f = gm._forest
        R._lib.check(R.lib().rv_forest_arrays(f, nq, R._lib.ptr(pts, c.c_double),
                                              R._lib.ptr(self.vmapM, c.c_int64),
                                              R._lib.ptr(self.vmapP, c.c_int64),
                                              R._lib.ptr(self.bc, c.c_int64),
                                              R._lib.ptr(self.kind, c.c_int32)))
        R._lib.check(R.lib().rv_forest_mortars(f, R._lib.ptr(self.partner, c.c_int64)))
        if warp is not None:
            pts = np.ascontiguousarray(np.stack(warp(*[pts[:, d, :] for d in range(dim)]), axis=1))
        self.points = pts.reshape(-1)
        L = O.lib()
        self.metrics = np.zeros(Np * (dim * dim + 2) * K)
        L.orc_volume_metrics(dim, nq, K, 0, self.points, self.D, self.w1d, self.metrics)
        self.surf_full = np.zeros((dim + 2) * nfpt * K)
        L.orc_surface_metrics(dim, nq, K, self.metrics, self.vmapM, self.w1d, self.surf_full)
        # the operator's own surface metrics: lift weight 0 on non-conforming faces
        self.surf = self.surf_full.copy().reshape(K, 2 * dim, nfp, dim + 2)
        self.surf[self.kind.reshape(K, 2 * dim) != 0, :, dim + 1] = 0.0
        self.surf = self.surf.reshape(-1)
        self.lower, self.upper = np.zeros(nq * nq), np.zeros(nq * nq)
        R._lib.check(R.lib().rv_lobatto_halves(nq, R._lib.ptr(self.lower, c.c_double),
                                               R._lib.ptr(self.upper, c.c_double)))

    def point_fields(self):
        return self.points.reshape(self.K, self.dim, self.Np)

    @property
    def MJ(self):
        L = self.dim * self.dim + 2
        return self.metrics.reshape(self.K, L, self.Np)[:, L - 1, :]

    def rhs(self, law, q, mortar=True, u=None):
        """dq of the coupled operator (mortar=True) or of Raven's self-mapped
        hanging faces (mortar=False)."""
        if mortar:
            dg = O.DG(self, law, u=u)
            dq = np.zeros(q.size)
            dg.rhs(dq, q.copy(), 0.0, increment=False)
            dg.mortar_add(self.kind, self.partner, self.lower, self.upper, self.surf_full, q, dq)
            return dq
        full = _Full(self)
        dg = O.DG(full, law, u=u)
        dq = np.zeros(q.size)
        dg.rhs(dq, q.copy(), 0.0, increment=False)
        return dq


class _Full:
    def __init__(self, g):
        self.__dict__.update(g.__dict__)
        self.surf = g.surf_full


def _warp(n):
    k = 2 * np.pi / n

    def w(*X):
        if len(X) == 2:
            x, y = X
            return (x + 0.03 * n * np.sin(k * y), y + 0.02 * n * np.sin(k * x))
        x, y, z = X
        return (x + 0.03 * n * np.sin(k * y) * np.sin(k * z),
                y + 0.02 * n * np.sin(k * x) * np.cos(k * z),
                z + 0.02 * n * np.sin(k * x) * np.sin(k * y))
    return w


@pytest.mark.parametrize("dim,nq,law", [(3, 4, "euler"), (3, 3, "advection"), (2, 5, "euler"),
                                        (2, 4, "advection")])
def test_mortar_conserves_on_periodic_adapted_grid(dim, nq, law):
    n = 4
    g = ForestGrid(dim, nq, n, 2, 1.2, True, warp=_warp(n))
    assert np.any(g.kind == 1)
    rng = np.random.default_rng(3)
    P = g.point_fields()
    u = [1.0, 0.6, 0.3][:dim]
    if law == "advection":
        q = PB.advection_wave(P / n, dim, 0.0, u) + 0.01 * rng.standard_normal(g.K * g.Np)
    else:
        q = PB.euler_wave(P / n, dim, 0.0) + 1e-3 * rng.standard_normal(g.K * g.Np * (dim + 2))
    ns = 1 if law == "advection" else dim + 2
    for mortar in (True, False):
        dq = g.rhs(law, q, mortar=mortar, u=u if law == "advection" else None)
        d = dq.reshape(g.K, ns, g.Np)
        tot = np.einsum("kn,ksn->s", g.MJ, d)
        scale = np.einsum("kn,ksn->s", g.MJ, np.abs(d))
        if mortar:
            assert np.all(np.abs(tot) <= 1e-12 * scale), (tot, scale)
        else:
            # self-mapped hanging faces are not conservative
            assert np.max(np.abs(tot) / scale) > 1e-8


@pytest.mark.parametrize("dim,nq", [(3, 4), (2, 5), (3, 3)])
def test_mortar_is_exact_for_a_linear_field(dim, nq):
    """Advection of q = 1 + c.x on an unwarped, non-periodic adapted brick:
    the traces on both sides of every mortar agree, so dq = -u.c everywhere."""
    g = ForestGrid(dim, nq, 4, 2, 1.2, False)
    P = g.point_fields()
    cvec = np.array([0.7, -0.4, 0.25][:dim])
    u = [1.0, 0.5, -0.3][:dim]
    q = (1.0 + np.einsum("d,kdn->kn", cvec, P)).reshape(-1)
    dq = g.rhs("advection", q, mortar=True, u=u)
    want = -float(np.dot(u, cvec))
    # rounding of the fine cells' large MJI stays far below what a wrong sub-face or
    # half (an O(h) jump in the traces) would give
    assert np.max(np.abs(dq - want)) <= 1e-9
