"""Reference cell and brick grids, mirroring Raven.jl's host API.

``LobattoCell`` (src/lobattocells.jl:170-240), ``brick`` (src/coarsegrids.jl:
233-405), ``GridManager`` (src/gridmanager.jl:30-91) and ``generate``
(src/gridmanager.jl:1084-1208) with the same names and argument meaning.  The
arrays come from the native generator in libravenhip (rv_brick_arrays) and use
Raven's memory layout; numpy views are shaped C-order, i.e. Julia's
``(Np, L, K)`` parent becomes ``(K, L, Np)`` here.
"""
import numpy as np

from . import _lib as L


class LobattoCell:
    """LobattoCell{T}(Nq, ..., Nq): tensor-product Gauss-Lobatto cell."""

    def __init__(self, *sizes):
        if len(sizes) < 1 or len(sizes) > 3 or len(set(sizes)) != 1:
            raise ValueError("LobattoCell needs 1-3 equal sizes (DGSEM asserts equal Nq)")
        self.size = tuple(int(s) for s in sizes)
        self.nq = self.size[0]
        self.dim = len(sizes)
        nq = self.nq
        self.points_1d = np.zeros(nq)
        self.weights_1d = np.zeros(nq)
        self._D = np.zeros(nq * nq)
        L.check(L.lib().rv_lobatto(nq, L.ptr(self.points_1d, L.ctypes.c_double),
                                   L.ptr(self.weights_1d, L.ctypes.c_double),
                                   L.ptr(self._D, L.ctypes.c_double)))

    def derivatives_1d(self):
        """D[a, b] = l_b'(x_a) as an (Nq, Nq) array."""
        return self._D.reshape(self.nq, self.nq, order="F")

    @property
    def D_colmajor(self):
        return self._D

    def __len__(self):
        return self.nq ** self.dim


class BrickGrid:
    """brick(coordinates, periodic): uniform coarse brick."""

    def __init__(self, coordinates, periodic):
        self.coordinates = [np.asarray(c, dtype=np.float64) for c in coordinates]
        self.dim = len(self.coordinates)
        self.periodic = tuple(bool(p) for p in periodic)
        if len(self.periodic) != self.dim:
            raise ValueError("periodic must have one flag per dimension")
        self.n = [len(c) - 1 for c in self.coordinates]
        self.lo = [float(c[0]) for c in self.coordinates]
        self.hi = [float(c[-1]) for c in self.coordinates]


def brick(coordinates, periodic=None):
    coordinates = tuple(coordinates)
    if periodic is None:
        periodic = (False,) * len(coordinates)
    return BrickGrid(coordinates, periodic)


# adapt! flags (gridmanager.jl:24-27)
AdaptCoarsen, AdaptNone, AdaptRefine, AdaptTouched = 0x00, 0x01, 0x10, 0x11


class GridManager:
    def __init__(self, cell, coarsegrid, min_level=0):
        if cell.dim != coarsegrid.dim:
            raise ValueError("cell and coarse grid dimensions differ")
        self.cell = cell
        self.coarsegrid = coarsegrid
        self.min_level = int(min_level)
        self._forest = None

    def __del__(self):
        try:
            if getattr(self, "_forest", None) is not None:
                L.lib().rv_forest_destroy(self._forest)
                self._forest = None
        except Exception:
            pass  # interpreter shutdown: the library may already be unloaded

    def _ensure_forest(self):
        if self._forest is None:
            desc = self.brick_desc()
            _check_uniform(self.coarsegrid, desc)
            f = L.H()
            L.check(L.lib().rv_forest_create(L.ctypes.byref(desc), L.ctypes.byref(f)))
            self._forest = f
        return self._forest

    def __len__(self):
        if self._forest is None:
            return int(L.lib().rv_brick_numcells(L.ctypes.byref(self.brick_desc())))
        return int(L.lib().rv_forest_numcells(self._forest))

    @property
    def adapted(self):
        return self._forest is not None

    def adapt(self, flags):
        """adapt!(gm, flags) (gridmanager.jl:160-172): coarsen every complete
        family of leaves flagged AdaptCoarsen, refine the other leaves flagged
        AdaptRefine by one level, then balance the forest 2:1.  As adapt!, the
        flags are overwritten with what happened to each old leaf (AdaptRefine,
        AdaptCoarsen or AdaptNone): in place when `flags` is a contiguous uint8
        array, and returned in any case."""
        f = self._ensure_forest()
        inplace = (isinstance(flags, np.ndarray) and flags.dtype == np.uint8
                   and flags.flags.c_contiguous and flags.flags.writeable)
        fl = flags if inplace else np.array(flags, dtype=np.uint8).reshape(-1)
        if fl.size != len(self):
            raise ValueError("one flag per cell")
        L.check(L.lib().rv_forest_adapt(f, fl.ctypes.data_as(L.ctypes.POINTER(L.ctypes.c_uint8))))
        return fl

    def cells(self):
        """(tree, level, quad): tree lattice coordinates (K, dim), levels (K)
        and each leaf's coordinates inside its tree at its level (K, dim)."""
        dim = self.cell.dim
        K = len(self)
        tree = np.zeros((K, dim), dtype=np.int32)
        level = np.zeros(K, dtype=np.int32)
        quad = np.zeros((K, dim), dtype=np.int32)
        if self._forest is None:
            desc = self.brick_desc()
            ijk = np.zeros((K, dim), dtype=np.int32)
            L.check(L.lib().rv_brick_cells(L.ctypes.byref(desc), L.ptr(ijk, L.ctypes.c_int32)))
            tree[:] = ijk >> self.min_level
            quad[:] = ijk - (tree << self.min_level)
            level[:] = self.min_level
        else:
            L.check(L.lib().rv_forest_cells(self._forest, L.ptr(tree, L.ctypes.c_int32),
                                            L.ptr(level, L.ctypes.c_int32),
                                            L.ptr(quad, L.ctypes.c_int32)))
        return tree, level, quad

    def cell_centres(self):
        """Centres of the leaves in brick coordinates (before any warp)."""
        tree, level, quad = self.cells()
        cg = self.coarsegrid
        out = np.zeros(tree.shape)
        for d in range(cg.dim):
            v = cg.coordinates[d]
            frac = (quad[:, d] + 0.5) / (1 << level).astype(np.float64)
            out[:, d] = v[tree[:, d]] + (v[tree[:, d] + 1] - v[tree[:, d]]) * frac
        return out

    def brick_desc(self, law=L.RV_LAW_ADVECTION, dtype=L.RV_F64, flux=L.RV_FLUX_RUSANOV,
                   gamma=1.4, velocity=(0.0, 0.0, 0.0), stream=None,
                   volume_form=L.RV_FORM_WEAK, volume_flux=0):
        cg = self.coarsegrid
        d = L.RvBrickDesc()
        d.abi_version = L.ABI_VERSION
        d.dim, d.nq, d.law, d.flux, d.dtype = cg.dim, self.cell.nq, law, flux, dtype
        for k in range(3):
            d.n[k] = cg.n[k] if k < cg.dim else 1
            d.periodic[k] = int(cg.periodic[k]) if k < cg.dim else 0
            d.lo[k] = cg.lo[k] if k < cg.dim else 0.0
            d.hi[k] = cg.hi[k] if k < cg.dim else 1.0
            d.velocity[k] = velocity[k] if k < len(velocity) else 0.0
        d.level = self.min_level
        d.gamma = gamma
        d.stream = stream
        d.volume_form, d.volume_flux = volume_form, volume_flux
        return d


class Grid:
    """Arrays of a generated grid in Raven's layout (C-order views)."""

    def __init__(self, gm, points, metrics, surface, vmapM, vmapP, bc, dtoc, nc):
        self.gm = gm
        self.referencecell = gm.cell
        self.dim = gm.cell.dim
        self.nq = gm.cell.nq
        self.K = points.shape[0]
        self.Np = self.nq ** self.dim
        self.nfp = self.nq ** (self.dim - 1)
        self.points = points          # (K, dim, Np)
        self.metrics = metrics        # (K, dim*dim+2, Np): dRdX (column-major), J, wJ
        self.surface = surface        # (K, 2dim*Nfp, dim+2): n, sJ, wsJ
        self.vmapM = vmapM            # (K, 2dim*Nfp) int64, 1-based
        self.vmapP = vmapP
        self.boundarycodes = bc       # (K, 2dim)
        self.dtoc = dtoc              # (K, Np) int64, 1-based
        self.ncontinuous = nc

    def __len__(self):
        return self.K

    @property
    def MJ(self):
        return self.metrics[:, -1, :]


def _check_uniform(cg, desc):
    for d in range(cg.dim):
        n = cg.n[d]
        got = cg.coordinates[d]
        # the generator's lattice lo + (hi-lo)*i/n, correctly rounded
        from fractions import Fraction
        lo, hi = Fraction(cg.lo[d]), Fraction(cg.hi[d])
        want = np.array([float(lo + (hi - lo) * i / n) for i in range(n + 1)])
        if not np.array_equal(got, want):
            raise NotImplementedError("only uniformly spaced brick coordinates are supported")


def generate(gm, warp=None, with_metrics=True):
    """generate(gm) / generate(warp, gm) (gridmanager.jl:1084-1208).

    warp(x, y[, z]) maps arrays of node coordinates to a tuple of arrays (the
    vectorised form of Raven's point-wise warp); warped grids take the curved
    metric path (gridmanager.jl:1160-1166), which runs on the device
    (rv_curved_metrics), as do the refined cells of an adapted forest
    (gm.adapt).  An adapted forest's grid also carries Raven's non-conforming
    tables (grid.vmapNC, .nctoface, .nctypes, .ncids, one entry per face
    group) and its continuous numbering with hanging nodes (rv_forest_dtoc).
    with_metrics=False skips the device-side metrics of curved or adapted grids
    (grid.metrics and grid.surface are then None): the integer arrays and
    points alone, without a device."""
    if gm.adapted:
        return _generate_forest(gm, warp, with_metrics)
    desc = gm.brick_desc()
    _check_uniform(gm.coarsegrid, desc)
    lib = L.lib()
    K = lib.rv_brick_numcells(L.ctypes.byref(desc))
    if K < 0:
        L.check(int(K))
    dim, nq = gm.cell.dim, gm.cell.nq
    Np, nfp = nq ** dim, nq ** (dim - 1)
    nfpt = 2 * dim * nfp
    points = np.zeros((K, dim, Np))
    metrics = np.zeros((K, dim * dim + 2, Np))
    surface = np.zeros((K, nfpt, dim + 2))
    vmapM = np.zeros((K, nfpt), dtype=np.int64)
    vmapP = np.zeros((K, nfpt), dtype=np.int64)
    bc = np.zeros((K, 2 * dim), dtype=np.int64)
    dtoc = np.zeros((K, Np), dtype=np.int64)
    nc = L.ctypes.c_int64(0)
    c = L.ctypes
    L.check(lib.rv_brick_arrays(c.byref(desc), L.ptr(points, c.c_double), L.ptr(metrics, c.c_double),
                                L.ptr(surface, c.c_double), L.ptr(vmapM, c.c_int64),
                                L.ptr(vmapP, c.c_int64), L.ptr(bc, c.c_int64),
                                L.ptr(dtoc, c.c_int64), c.byref(nc)))
    if warp is not None:
        X = warp(*[points[:, d, :] for d in range(dim)])
        points = np.ascontiguousarray(
            np.stack([np.broadcast_to(np.asarray(x, dtype=np.float64), (K, Np)) for x in X],
                     axis=1))
        metrics, surface = curved_metrics(gm.cell, points) if with_metrics else (None, None)
    return Grid(gm, points, metrics, surface, vmapM, vmapP, bc, dtoc, nc.value)


def _generate_forest(gm, warp, with_metrics=True):
    f = gm._forest
    dim, nq = gm.cell.dim, gm.cell.nq
    K = len(gm)
    Np, nfp = nq ** dim, nq ** (dim - 1)
    nfpt = 2 * dim * nfp
    points = np.zeros((K, dim, Np))
    vmapM = np.zeros((K, nfpt), dtype=np.int64)
    vmapP = np.zeros((K, nfpt), dtype=np.int64)
    bc = np.zeros((K, 2 * dim), dtype=np.int64)
    kind = np.zeros((K, 2 * dim), dtype=np.int32)
    c = L.ctypes
    L.check(L.lib().rv_forest_arrays(f, nq, L.ptr(points, c.c_double), L.ptr(vmapM, c.c_int64),
                                     L.ptr(vmapP, c.c_int64), L.ptr(bc, c.c_int64),
                                     L.ptr(kind, c.c_int32)))
    if warp is not None:
        X = warp(*[points[:, d, :] for d in range(dim)])
        points = np.ascontiguousarray(
            np.stack([np.broadcast_to(np.asarray(x, dtype=np.float64), (K, Np)) for x in X],
                     axis=1))
    metrics, surface = curved_metrics(gm.cell, points) if with_metrics else (None, None)
    partner = np.zeros((K, 2 * dim, 4), dtype=np.int64)
    L.check(L.lib().rv_forest_mortars(f, L.ptr(partner, c.c_int64)))
    dtoc = np.zeros((K, Np), dtype=np.int64)
    nc = c.c_int64(0)
    L.check(L.lib().rv_forest_dtoc(f, nq, L.ptr(dtoc, c.c_int64), c.byref(nc)))
    grid = Grid(gm, points, metrics, surface, vmapM, vmapP, bc, dtoc, nc.value)
    grid.face_kind = kind
    grid.mortar_partner = partner
    grid.levels = gm.cells()[1]
    (grid.vmapNC, grid.nctoface, grid.nctypes, grid.ncids) = forest_nctables(f, dim, nq)
    return grid


def forest_nctables(forest, dim, nq):
    """Raven's facemaps(grid).vmapNC, .nctoface, .nctypes, .ncids
    (lobattocells.jl:2121-2276) of a forest: tuples over the face groups of
    arrays in Julia's column-major layout stored C-order reversed, i.e.
    vmapNC[g] has shape (nunique, 1 + 2^(dim-1), Nfp), nctoface[g] (nunique, 2)."""
    c = L.ctypes
    nun = np.zeros(dim, dtype=np.int64)
    nfc = np.zeros(dim, dtype=np.int64)
    L.check(L.lib().rv_forest_nc_counts(forest, L.ptr(nun, c.c_int64), L.ptr(nfc, c.c_int64)))
    nfp, nch = nq ** (dim - 1), 1 << (dim - 1)
    vmapNC = tuple(np.zeros((int(n), 1 + nch, nfp), dtype=np.int64) for n in nun)
    nctoface = tuple(np.zeros((int(n), 2), dtype=np.int8) for n in nun)
    nctypes = tuple(np.zeros(int(n), dtype=np.int8) for n in nfc)
    ncids = tuple(np.zeros(int(n), dtype=np.int64) for n in nfc)
    P64, P8 = c.POINTER(c.c_int64), c.POINTER(c.c_int8)
    arr = lambda T, xs: (T * dim)(*[x.ctypes.data_as(T) for x in xs])  # noqa: E731
    L.check(L.lib().rv_forest_nctables(forest, nq, arr(P64, vmapNC), arr(P8, nctoface),
                                       arr(P8, nctypes), arr(P64, ncids)))
    return vmapNC, nctoface, nctypes, ncids


def mortars_from_nctables(dim, nq, K, vmapNC, nctoface):
    """(face_kind (K, 2dim), partner (K, 2dim, 4)) for DGSEM.enable_mortars
    from Raven's vmapNC / nctoface tuples (rv_mortars_from_nctables)."""
    c = L.ctypes
    vm = [np.ascontiguousarray(v, dtype=np.int64) for v in vmapNC]
    nf = [np.ascontiguousarray(v, dtype=np.int8) for v in nctoface]
    nun = np.array([v.shape[0] for v in vm], dtype=np.int64)
    P64, P8 = c.POINTER(c.c_int64), c.POINTER(c.c_int8)
    kind = np.zeros((K, 2 * dim), dtype=np.int32)
    partner = np.zeros((K, 2 * dim, 4), dtype=np.int64)
    L.check(L.lib().rv_mortars_from_nctables(
        dim, nq, K, L.ptr(nun, c.c_int64), (P64 * dim)(*[v.ctypes.data_as(P64) for v in vm]),
        (P8 * dim)(*[v.ctypes.data_as(P8) for v in nf]), L.ptr(kind, c.c_int32),
        L.ptr(partner, c.c_int64)))
    return kind, partner


def min_node_distance(grid, dims=None):
    """min_node_distance(grid; dims = 1:ndims(cell)) (grids.jl:94-145): the
    least distance between neighbouring nodes of a cell along the given
    directions (1-based), reduced on the device (rv_min_node_distance)."""
    import torch
    dims = range(1, grid.dim + 1) if dims is None else tuple(dims)
    if not dims or max(dims) > grid.dim or min(dims) < 1:
        raise ValueError("dims are not valid")
    mask = sum(1 << (d - 1) for d in dims)
    pd = torch.from_numpy(np.ascontiguousarray(grid.points)).cuda()
    out = np.zeros(1)
    c = L.ctypes
    L.check(L.lib().rv_min_node_distance(grid.dim, grid.nq, grid.K, c.c_void_p(pd.data_ptr()),
                                         mask, L.ptr(out, c.c_double),
                                         c.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return float(out[0])


def curved_metrics(cell, points):
    """Volume and surface metrics of curved cells (rv_curved_metrics) for
    host node coordinates (K, dim, Np); returns host arrays in Raven's
    layout."""
    import torch
    K, dim, Np = points.shape
    nfpt = 2 * dim * cell.nq ** (dim - 1)
    pd = torch.from_numpy(np.ascontiguousarray(points)).cuda()
    md = torch.empty((K, dim * dim + 2, Np), dtype=torch.float64, device="cuda")
    sd = torch.empty((K, nfpt, dim + 2), dtype=torch.float64, device="cuda")
    c = L.ctypes
    D = np.ascontiguousarray(cell.D_colmajor)
    L.check(L.lib().rv_curved_metrics(dim, cell.nq, K, c.c_void_p(pd.data_ptr()),
                                      L.ptr(D, c.c_double), L.ptr(cell.weights_1d, c.c_double),
                                      c.c_void_p(md.data_ptr()), c.c_void_p(sd.data_ptr()),
                                      c.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return md.cpu().numpy(), sd.cpu().numpy()
