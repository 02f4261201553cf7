"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the reference results are checked against
and as the CPU baseline.  The product
(raven.jl_amd) never imports it.

Arrays are numpy arrays in Raven's layout (Fortran-ordered views of Julia's
column-major arrays are stored here as flat C arrays with the same offsets).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.orc_gll.argtypes = [ctypes.c_int, f64p, f64p, f64p]
        L.orc_brick_numcells.argtypes = [ctypes.c_int, i32p, ctypes.c_int]
        L.orc_brick_numcells.restype = ctypes.c_int64
        L.orc_brick_quadrants.argtypes = [ctypes.c_int, i32p, ctypes.c_int, i32p, i32p]
        L.orc_brick_points.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                                       i32p, i32p, ctypes.POINTER(ctypes.c_void_p), f64p, f64p]
        L.orc_volume_metrics.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                                         f64p, f64p, f64p, f64p]
        L.orc_surface_metrics.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, f64p, i64p,
                                          f64p, f64p]
        L.orc_brick_facemaps.argtypes = [ctypes.c_int, ctypes.c_int, i32p, i32p, ctypes.c_int,
                                         ctypes.c_int64, i32p, i32p, i64p, i64p, i64p, i64p, i64p]
        L.orc_brick_dtoc.argtypes = [ctypes.c_int, ctypes.c_int, i32p, i32p, ctypes.c_int,
                                     ctypes.c_int64, i32p, i32p, i64p]
        L.orc_brick_dtoc.restype = ctypes.c_int64
        L.orc_gather.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                 i64p, f64p, f64p]
        L.orc_scatter.argtypes = L.orc_gather.argtypes
        L.orc_lsrk_coefficients.argtypes = [ctypes.c_int, f64p, f64p, f64p]
        L.orc_lsrk_coefficients.restype = ctypes.c_int
        L.orc_dg_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                                 ctypes.c_double, f64p, f64p, f64p, f64p, i64p, i64p, i64p]
        L.orc_dg_new.restype = ctypes.c_void_p
        L.orc_dg_free.argtypes = [ctypes.c_void_p]
        L.orc_dg_ns.argtypes = [ctypes.c_void_p]
        L.orc_dg_ns.restype = ctypes.c_int
        L.orc_dg_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_dg_set_volume_form.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.orc_dg_mortar_add.argtypes = [ctypes.c_void_p, i32p, i64p, f64p, f64p, f64p, f64p,
                                         f64p]
        L.orc_dg_rhs.argtypes = [ctypes.c_void_p, f64p, f64p, ctypes.c_double, ctypes.c_int]
        L.orc_dg_lsrk_step.argtypes = [ctypes.c_void_p, ctypes.c_int, f64p, f64p,
                                       ctypes.c_double, ctypes.c_double]
        L.orc_dg_solve.argtypes = [ctypes.c_void_p, ctypes.c_int, f64p, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double,
                                   ctypes.POINTER(ctypes.c_double)]
        L.orc_dg_solve.restype = ctypes.c_int64
        L.orc_dg_weightednorm.argtypes = [ctypes.c_void_p, f64p, f64p]
        L.orc_dg_set_surface_flux.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_dg_set_directions.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_dg_entropy_integral.argtypes = [ctypes.c_void_p, f64p]
        L.orc_dg_entropy_integral.restype = ctypes.c_double
        L.orc_dg_entropy_product.argtypes = [ctypes.c_void_p, f64p, f64p]
        L.orc_dg_entropy_product.restype = ctypes.c_double
        L.orc_dg_rlsrk_step.argtypes = [ctypes.c_void_p, ctypes.c_int, f64p, f64p, f64p, f64p,
                                        ctypes.c_double, ctypes.c_double,
                                        ctypes.POINTER(ctypes.c_double)]
        L.orc_dg_rlsrk_step.restype = ctypes.c_int
        L.orc_min_node_distance.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, f64p,
                                            ctypes.c_int]
        L.orc_min_node_distance.restype = ctypes.c_double
        L.orc_rlsrk_weights.argtypes = [ctypes.c_int, f64p]
        L.orc_rlsrk_weights.restype = ctypes.c_int
        _LIB = L
    return _LIB


def rlsrk_weights(scheme=54):
    b = np.zeros(14)
    n = lib().orc_rlsrk_weights(scheme, b)
    return b[:n].copy()


def gll(nq):
    x = np.zeros(nq)
    w = np.zeros(nq)
    D = np.zeros(nq * nq)
    lib().orc_gll(nq, x, w, D)
    return x, w, D  # D column-major: D[a + nq*b]


def lsrk_coefficients(scheme):
    a, b, c = np.zeros(14), np.zeros(14), np.zeros(14)
    n = lib().orc_lsrk_coefficients(scheme, a, b, c)
    return a[:n], b[:n], c[:n]


class BrickGrid:
    """Oracle restatement of brick(...) + GridManager + generate(warp, gm)."""

    def __init__(self, nq, coords, periodic, level=0, warp=None):
        self.dim = dim = len(coords)
        self.nq = nq
        self.coords = [np.ascontiguousarray(c, dtype=np.float64) for c in coords]
        self.n = np.array([len(c) - 1 for c in coords], dtype=np.int32)
        self.periodic = np.array([1 if p else 0 for p in periodic], dtype=np.int32)
        self.level = level
        L = lib()
        K = L.orc_brick_numcells(dim, self.n, level)
        self.K = K
        self.tree = np.zeros(K * dim, dtype=np.int32)
        self.quad = np.zeros(K * dim, dtype=np.int32)
        L.orc_brick_quadrants(dim, self.n, level, self.tree, self.quad)
        self.Np = nq ** dim
        self.nfp = nq ** (dim - 1)
        self.x1d, self.w1d, self.D = gll(nq)
        pts = np.zeros(self.Np * dim * K)
        ptrs = (ctypes.c_void_p * dim)(*[c.ctypes.data for c in self.coords])
        L.orc_brick_points(dim, nq, K, level, self.tree, self.quad, ptrs, self.x1d, pts)
        if warp is not None:
            P = pts.reshape(K, dim, self.Np)
            X = warp(*[P[:, d, :] for d in range(dim)])
            pts = np.ascontiguousarray(np.stack(X, axis=1)).reshape(-1)
        self.points = pts
        nfpt = 2 * dim * self.nfp
        self.vmapM = np.zeros(nfpt * K, dtype=np.int64)
        self.vmapP = np.zeros(nfpt * K, dtype=np.int64)
        self.mapM = np.zeros(nfpt * K, dtype=np.int64)
        self.mapP = np.zeros(nfpt * K, dtype=np.int64)
        self.bc = np.zeros(2 * dim * K, dtype=np.int64)
        L.orc_brick_facemaps(dim, nq, self.n, self.periodic, level, K, self.tree, self.quad,
                             self.vmapM, self.vmapP, self.mapM, self.mapP, self.bc)
        self.metrics = np.zeros(self.Np * (dim * dim + 2) * K)
        L.orc_volume_metrics(dim, nq, K, 1 if warp is None else 0, pts, self.D, self.w1d,
                             self.metrics)
        self.surf = np.zeros((dim + 2) * nfpt * K)
        L.orc_surface_metrics(dim, nq, K, self.metrics, self.vmapM, self.w1d, self.surf)
        self.dtoc = np.zeros(self.Np * K, dtype=np.int64)
        self.nc = L.orc_brick_dtoc(dim, nq, self.n, self.periodic, level, K, self.tree,
                                   self.quad, self.dtoc)

    def min_node_distance(self, dims=None):
        """min_node_distance(grid; dims) (grids.jl:94-145); dims 1-based."""
        dims = range(1, self.dim + 1) if dims is None else dims
        mask = sum(1 << (d - 1) for d in dims)
        return lib().orc_min_node_distance(self.dim, self.nq, self.K, self.points, mask)

    def point_fields(self):
        """(K, dim, Np) view of the node coordinates."""
        return self.points.reshape(self.K, self.dim, self.Np)

    @property
    def MJ(self):
        L = self.dim * self.dim + 2
        return self.metrics.reshape(self.K, L, self.Np)[:, L - 1, :]


FLUXES = {"rusanov": 1, "central": 2, "ec": 3}


class DG:
    """Oracle DGSEM operator for law 'advection' or 'euler'."""

    def __init__(self, grid, law, gamma=1.4, u=None, volume_flux=None, surface_flux="rusanov",
                 directions=None):
        """volume_flux None: WeakForm; "central" or "ec": FluxDifferencingForm
        with CentralFlux or EntropyConservativeFlux (dgsem.jl:1-17).
        surface_flux "rusanov", "central" or "ec" (numericalfluxes.jl:27-37).
        directions: DGSEM's `directions` (1-based), all by default."""
        self.grid = grid
        self.law = {"advection": 1, "euler": 2}[law]
        u = np.zeros(3) if u is None else np.asarray(u, dtype=np.float64)
        self._u = np.ascontiguousarray(np.resize(u, 3), dtype=np.float64)
        self.gamma = gamma
        self.h = lib().orc_dg_new(grid.dim, grid.nq, grid.K, self.law, gamma, self._u,
                                  grid.D, grid.metrics, grid.surf, grid.vmapM, grid.vmapP,
                                  grid.bc)
        self.ns = lib().orc_dg_ns(self.h)
        if volume_flux is not None:
            lib().orc_dg_set_volume_form(self.h, 1, {"central": 2, "ec": 3}[volume_flux])
        lib().orc_dg_set_surface_flux(self.h, FLUXES[surface_flux])
        if directions is not None:
            lib().orc_dg_set_directions(self.h, sum(1 << (d - 1) for d in directions))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_dg_free(self.h)
            self.h = None

    def set_threads(self, n):
        lib().orc_dg_set_threads(self.h, n)

    def new_state(self):
        return np.zeros(self.grid.Np * self.ns * self.grid.K)

    def rhs(self, dq, q, t=0.0, increment=True):
        lib().orc_dg_rhs(self.h, dq, q, t, 1 if increment else 0)

    def lsrk_step(self, q, dq, time, dt, scheme=54):
        lib().orc_dg_lsrk_step(self.h, scheme, q, dq, time, dt)

    def solve(self, q, dt, timeend, t0=0.0, scheme=54):
        t = ctypes.c_double(0)
        n = lib().orc_dg_solve(self.h, scheme, q, dt, t0, timeend, ctypes.byref(t))
        return n, t.value

    def mortar_add(self, kind, partner, lower, upper, surf_full, q, dq):
        """Add the mortar fluxes of hanging faces to dq (orc_dg_mortar_add)."""
        self._mortar_keep = [np.ascontiguousarray(kind, dtype=np.int32).reshape(-1),
                             np.ascontiguousarray(partner, dtype=np.int64).reshape(-1),
                             np.ascontiguousarray(lower, dtype=np.float64).reshape(-1),
                             np.ascontiguousarray(upper, dtype=np.float64).reshape(-1),
                             np.ascontiguousarray(surf_full, dtype=np.float64).reshape(-1)]
        lib().orc_dg_mortar_add(self.h, *self._mortar_keep, q, dq)

    def entropy_integral(self, q):
        return lib().orc_dg_entropy_integral(self.h, q)

    def entropy_product(self, p, q):
        return lib().orc_dg_entropy_product(self.h, p, q)

    def rlsrk_solve(self, q, dt, timeend, t0=0.0, scheme=54, after_step=None):
        """solve!(q, timeend, RLSRK) (odesolvers.jl:9-31, 117-151): the final
        step is shortened to end at timeend; after_step(step, time, q, gamma)
        after every step.  Returns (steps, end time, last gamma)."""
        n = q.size
        dq, k, q0 = np.zeros(n), np.zeros(n), np.zeros(n)
        gamma = ctypes.c_double(1.0)
        time, step = t0, 0
        while True:
            step += 1
            final = time + dt >= timeend
            if final:
                dt = timeend - time
            it = lib().orc_dg_rlsrk_step(self.h, scheme, q, dq, k, q0, time, dt,
                                         ctypes.byref(gamma))
            if it < 0:
                raise FloatingPointError("relaxation Newton iteration broke down")
            if after_step is not None:
                after_step(step, time, q, gamma.value)
            time += dt
            if final:
                return step, time, gamma.value

    def weightednorm(self, q):
        comp = np.zeros(self.ns)
        lib().orc_dg_weightednorm(self.h, q, comp)
        return float(np.sqrt(np.sum(comp * comp))), comp
