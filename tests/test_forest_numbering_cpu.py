"""Continuous numbering and non-conforming face tables of adapted forests
(rv_forest_dtoc, rv_forest_nctables, rv_mortars_from_nctables), host only.

Pins:
  * uniform refinement through adapt gives the brick generator's numbering
    bit for bit (which is pinned to Raven's golden numbering);
  * the numbering the reference's own test draws for a 2 x 2 brick with one
    refined tree (test/gridnumbering.jl:101-111): hanging nodes repeat the
    coarse edge's ids by position;
  * on 3-D forests: every node of a hanging face carries the id of the coarse
    face's node at the same face position, interior nodes of conforming faces
    are shared, and non-hanging nodes with one id coincide;
  * the non-conforming tables: children's points are the parent face's points
    interpolated with tohalves (the check of the reference's
    mpitest_n2_gridnumbering.jl:339-366), every interface is listed once per
    parent and child, and Raven's tables convert back to the forest's own
    face kinds and partners.
"""
import itertools

import numpy as np
import pytest

import problems as PB
import ravenhip as R
from ravenhip import grids as G

c = R._lib.ctypes


def forest_gm(dim, nq, n, rounds, radius, periodic=True):
    cell = R.LobattoCell(*([nq] * dim))
    gm = R.GridManager(cell, R.brick([PB.julia_range(0, n, n + 1)] * dim, (periodic,) * dim))
    centre = np.full(dim, n / 2.0)
    for r in range(rounds):
        cc = gm.cell_centres()
        lev = gm.cells()[1]
        flags = np.where((np.linalg.norm(cc - centre, axis=1) < radius) & (lev <= r),
                         R.AdaptRefine, R.AdaptNone)
        gm.adapt(flags)
    return gm


def first_appearance(ids):
    out, seen = np.zeros_like(ids), {}
    for i, v in enumerate(ids):
        out[i] = seen.setdefault(int(v), len(seen) + 1)
    return out


@pytest.mark.parametrize("dim,nq,n,level,periodic", [(2, 4, 3, 1, True), (2, 3, 2, 2, False),
                                                     (3, 3, 2, 1, True), (3, 4, 3, 1, False),
                                                     (3, 5, 1, 1, True)])
def test_uniform_refinement_numbering_is_the_bricks(dim, nq, n, level, periodic):
    coords = [PB.julia_range(0, 1, n + 1)] * dim
    per = (periodic,) * dim
    ref = R.generate(R.GridManager(R.LobattoCell(*([nq] * dim)), R.brick(coords, per),
                                   min_level=level))
    gm = R.GridManager(R.LobattoCell(*([nq] * dim)), R.brick(coords, per))
    for _ in range(level):
        gm.adapt(np.full(len(gm), R.AdaptRefine, dtype=np.uint8))
    grid = R.generate(gm, with_metrics=False)
    assert grid.ncontinuous == ref.ncontinuous
    assert np.array_equal(grid.dtoc, ref.dtoc)


def test_numbering_matches_the_reference_test_drawing():
    """gridnumbering.jl:101-111: four trees on [0,4]^2 (non-periodic), tree 4
    (top right) refined once, three nodes per direction.  The drawing gives
    each cell's ids row by row from the bottom, in physical orientation; the
    children's edges along the coarse trees repeat the coarse edge's ids."""
    drawn = [
        [[1, 2, 3], [4, 5, 6], [7, 8, 9]],             # tree 1 (0,0)
        [[3, 15, 14], [6, 13, 12], [9, 11, 10]],       # tree 2 (1,0)
        [[7, 8, 9], [16, 17, 18], [19, 20, 21]],       # tree 3 (0,1)
        [[9, 11, 10], [18, 22, 23], [21, 24, 25]],     # tree 4, child (0,0)
        [[9, 11, 10], [23, 26, 27], [25, 28, 29]],     # child (1,0)
        [[9, 24, 25], [18, 30, 31], [21, 32, 33]],     # child (0,1)
        [[25, 28, 29], [31, 34, 35], [33, 36, 37]],    # child (1,1)
    ]
    gm = R.GridManager(R.LobattoCell(3, 3), R.brick([PB.julia_range(0, 4, 3)] * 2, (False, False)))
    tree = gm.cells()[0]
    flags = np.where((tree[:, 0] == 1) & (tree[:, 1] == 1), R.AdaptRefine, R.AdaptNone)
    gm.adapt(flags.astype(np.uint8))
    tree, level, quad = gm.cells()
    order = [(0, 0, 0, 0, 0), (1, 0, 0, 0, 0), (0, 1, 0, 0, 0),
             (1, 1, 1, 0, 0), (1, 1, 1, 1, 0), (1, 1, 1, 0, 1), (1, 1, 1, 1, 1)]
    assert [tuple(tree[e]) + (level[e],) + tuple(quad[e]) for e in range(len(gm))] == order
    grid = R.generate(gm, with_metrics=False)
    want = first_appearance(np.array(drawn).reshape(-1))
    assert grid.ncontinuous == 37
    assert np.array_equal(grid.dtoc.reshape(-1), want)


def hanging_face_targets(grid, e, fc):
    """Face node m of face fc of cell e -> volume index (cell, node) of the
    node of the coarse cell's opposite face at the same face position."""
    nq, dim = grid.nq, grid.dim
    Np, nfp, nfpt = grid.Np, grid.nfp, 2 * dim * grid.nfp
    cpar = int(grid.mortar_partner[e, fc, 0])
    mine = grid.vmapM[e, fc * nfp:(fc + 1) * nfp] - 1
    g = fc ^ 1
    theirs = grid.vmapM[cpar, g * nfp:(g + 1) * nfp] - 1
    assert nfpt == grid.vmapM.shape[1]
    return mine, theirs, Np


@pytest.mark.parametrize("dim,nq,n,rounds,radius", [(3, 3, 4, 2, 1.2), (3, 4, 4, 1, 1.5),
                                                    (2, 5, 6, 2, 2.0)])
def test_forest_numbering_invariants(dim, nq, n, rounds, radius):
    gm = forest_gm(dim, nq, n, rounds, radius)
    grid = R.generate(gm, with_metrics=False)
    K, Np, nfp = grid.K, grid.Np, grid.nfp
    d = grid.dtoc.reshape(-1)
    # ids 1..n, numbered by first appearance
    assert d.max() == grid.ncontinuous
    firsts = np.unique(d, return_index=True)[1]
    assert np.all(np.diff(d[np.sort(firsts)]) == 1) and d[0] == 1
    kind = grid.face_kind
    assert np.any(kind == 1)
    hanging = np.zeros(K * Np, dtype=bool)
    for e in range(K):
        for fc in range(2 * dim):
            if kind[e, fc] != 1:
                continue
            mine, theirs, _ = hanging_face_targets(grid, e, fc)
            hanging[mine] = True
            assert np.array_equal(d[mine], d[theirs])  # by position on the coarse face
    # conforming faces share their interior nodes
    inner = np.ones((nq,) * (dim - 1), dtype=bool)
    if dim == 3:
        inner[0, :] = inner[-1, :] = inner[:, 0] = inner[:, -1] = False
    else:
        inner[0] = inner[-1] = False
    inner = inner.T.reshape(-1)
    vM, vP = grid.vmapM.reshape(K, 2 * dim, nfp) - 1, grid.vmapP.reshape(K, 2 * dim, nfp) - 1
    for e in range(K):
        for fc in range(2 * dim):
            if kind[e, fc] == 0:
                assert np.array_equal(d[vM[e, fc][inner]], d[vP[e, fc][inner]])
    # in 3-D a node is also hanging on an edge whose diagonal neighbour is
    # coarser while neither face at the edge hangs: find those edges from the
    # leaves' lattice positions
    tree, level, quad = gm.cells()
    finest = int(level.max())
    N = n << finest
    owner = {}
    for e in range(K):
        s = 1 << (finest - level[e])
        lo = [(int(tree[e][k]) << finest) + (int(quad[e][k]) << (finest - level[e]))
              for k in range(dim)]
        for off in itertools.product(range(s), repeat=dim):
            owner[tuple(lo[k] + off[k] for k in range(dim))] = e
    nedge = 0
    if dim == 3:
        for e in range(K):
            s = 1 << (finest - level[e])
            lo = [(int(tree[e][k]) << finest) + (int(quad[e][k]) << (finest - level[e]))
                  for k in range(3)]
            for a in range(3):
                bx, cx = [k for k in range(3) if k != a]
                for sb, sc in itertools.product((0, 1), repeat=2):
                    p = [0, 0, 0]
                    p[a] = lo[a]
                    p[bx] = lo[bx] + s if sb else lo[bx] - 1
                    p[cx] = lo[cx] + s if sc else lo[cx] - 1
                    o = owner[tuple(v % N for v in p)]
                    if level[o] >= level[e] or kind[e, 2 * bx + sb] == 1 or \
                            kind[e, 2 * cx + sc] == 1:
                        continue
                    ii = np.arange(Np)
                    idx = [(ii // nq ** k) % nq for k in range(3)]
                    on = (idx[bx] == (nq - 1 if sb else 0)) & (idx[cx] == (nq - 1 if sc else 0))
                    hanging[e * Np + ii[on]] = True
                    # by position on the coarse leaf's edge (the opposite sides)
                    onc = (idx[bx] == (0 if sb else nq - 1)) & (idx[cx] == (0 if sc else nq - 1))
                    assert np.array_equal(d[e * Np + ii[on]], d[o * Np + ii[onc]])
                    nedge += 1
    assert dim == 2 or rounds < 2 or nedge > 0  # the two-level forest has edge-only hanging
    # non-hanging nodes sharing an id coincide (up to the period)
    P = grid.points.transpose(0, 2, 1).reshape(K * Np, dim)
    L = float(n)
    free = ~hanging
    order = np.argsort(d[free], kind="stable")
    ids, pts = d[free][order], P[free][order]
    starts = np.flatnonzero(np.r_[True, ids[1:] != ids[:-1]])
    first = np.repeat(pts[starts], np.diff(np.r_[starts, ids.size]), axis=0)
    wrapped = np.remainder(pts - first + L / 2, L) - L / 2
    assert np.allclose(wrapped, 0, atol=1e-12)
    # and co-located non-hanging nodes share an id
    key = np.round(np.remainder(P[free], L) * 1e6).astype(np.int64)
    _, inv = np.unique(key, axis=0, return_inverse=True)
    for grp in np.unique(inv)[:5000]:
        assert np.unique(d[free][inv.reshape(-1) == grp]).size == 1


def test_forest_nctables_points_are_tohalves_of_the_parent():
    """mpitest_n2_gridnumbering.jl:339-366 on a 3-D forest: for every
    interface, child c's face points are the parent's face points
    interpolated with tohalves along the two tangential directions, the
    halves of c - 1 = s1 + 2 s2."""
    nq = 4
    gm = forest_gm(3, nq, 4, 2, 1.2)
    grid = R.generate(gm, with_metrics=False)
    lower, upper = np.zeros(nq * nq), np.zeros(nq * nq)
    R._lib.check(R.lib().rv_lobatto_halves(nq, R._lib.ptr(lower, c.c_double),
                                           R._lib.ptr(upper, c.c_double)))
    H = [lower.reshape(nq, nq), upper.reshape(nq, nq)]
    P = grid.points.transpose(0, 2, 1).reshape(grid.K * grid.Np, 3)
    nun = 0
    for g in range(3):
        vm = grid.vmapNC[g]
        nun += vm.shape[0]
        for i in range(vm.shape[0]):
            pp = P[vm[i, 0] - 1].reshape(nq, nq, 3)  # [b, a, xyz]: a (first tangential) fastest
            for ch in range(4):
                s1, s2 = ch & 1, ch >> 1
                cp = P[vm[i, 1 + ch] - 1].reshape(nq, nq, 3)
                want = np.einsum("ia,jb,bax->jix", H[s1], H[s2], pp)
                assert np.allclose(cp, want, atol=1e-12)
        # one parent entry and four child entries per interface
        types, ids = grid.nctypes[g], grid.ncids[g]
        for i in range(vm.shape[0]):
            assert sorted(types[ids == i + 1].tolist()) == [1, 2, 3, 4, 5]
        assert set(grid.nctoface[g][:, 0].tolist()) <= {2 * g + 1, 2 * g + 2}
        assert np.all(grid.nctoface[g][:, 1] == (grid.nctoface[g][:, 0] - 1 ^ 1) + 1)
    assert nun == int(np.sum(grid.face_kind == 2))


@pytest.mark.parametrize("dim,nq", [(3, 3), (2, 5)])
def test_nctables_convert_to_the_forest_partners(dim, nq):
    gm = forest_gm(dim, nq, 6 if dim == 2 else 4, 2, 2.0 if dim == 2 else 1.2)
    grid = R.generate(gm, with_metrics=False)
    kind, partner = G.mortars_from_nctables(dim, nq, grid.K, grid.vmapNC, grid.nctoface)
    assert np.array_equal(kind, grid.face_kind)
    assert np.array_equal(partner, grid.mortar_partner)
    # the tables are listed in cell-then-face order, ids by first encounter
    for g in range(dim):
        ids = grid.ncids[g]
        firsts = np.unique(ids, return_index=True)[1]
        assert np.all(np.diff(ids[np.sort(firsts)]) == 1) and ids[0] == 1


def test_nctables_with_rotated_nodes_are_refused():
    gm = forest_gm(3, 3, 4, 1, 1.5)
    grid = R.generate(gm, with_metrics=False)
    vm = [v.copy() for v in grid.vmapNC]
    g = next(i for i in range(3) if vm[i].shape[0])
    vm[g][0, 0, [1, 3]] = vm[g][0, 0, [3, 1]]  # transpose two nodes of a parent face
    with pytest.raises(R.RavenHIPError):
        G.mortars_from_nctables(3, 3, grid.K, vm, grid.nctoface)
