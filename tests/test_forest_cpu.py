"""The adaptive brick forest (rv_forest_*), host only.

p4est, which orders and balances Raven's forests, is not available, so the
adapted forests are checked through what must hold for any correct forest:
uniform refinement reproduces the brick generator (which is pinned to Raven's
golden face maps and numbering), leaves tile every tree exactly, leaves are in
tree-then-Morton order, every pair of touching leaves differs by at most one
level (2:1 balance across faces, edges and corners), conforming face maps are
involutions whose paired nodes coincide, and non-conforming faces are
self-mapped as in Raven (lobattocells.jl:2072-2074)."""
import itertools

import numpy as np
import pytest

import problems as PB
import ravenhip as R


def sphere_forest(dim, nq, n, rounds, radius, periodic=True):
    cell = R.LobattoCell(*([nq] * dim))
    gm = R.GridManager(cell, R.brick([PB.julia_range(0, n, n + 1)] * dim, (periodic,) * dim))
    centre = np.full(dim, n / 2.0)
    for r in range(rounds):
        c = gm.cell_centres()
        lev = gm.cells()[1]
        flags = np.where((np.linalg.norm(c - centre, axis=1) < radius) & (lev <= r),
                         R.AdaptRefine, R.AdaptNone)
        gm.adapt(flags)
    return gm


def morton(c, dim, bits=19):
    out = 0
    for b in range(bits):
        for d in range(dim):
            out |= ((int(c[d]) >> b) & 1) << (b * dim + d)
    return out


@pytest.mark.parametrize("dim,nq,n,level", [(2, 4, 5, 1), (3, 3, 3, 1), (2, 3, 4, 2)])
def test_uniform_refinement_reproduces_the_brick_generator(dim, nq, n, level):
    coords = [PB.julia_range(0, 1, n + 1)] * dim
    per = (True,) + (False,) * (dim - 1)
    ref = R.generate(R.GridManager(R.LobattoCell(*([nq] * dim)), R.brick(coords, per),
                                   min_level=level))
    gm = R.GridManager(R.LobattoCell(*([nq] * dim)), R.brick(coords, per))
    for _ in range(level):
        gm.adapt(np.full(len(gm), R.AdaptRefine, dtype=np.uint8))
    f = gm._forest
    K = len(gm)
    assert K == ref.K
    Np, nfpt = nq ** dim, 2 * dim * nq ** (dim - 1)
    pts = np.zeros((K, dim, Np))
    vM, vP = np.zeros((K, nfpt), np.int64), np.zeros((K, nfpt), np.int64)
    bc = np.zeros((K, 2 * dim), np.int64)
    c = R._lib.ctypes
    R._lib.check(R.lib().rv_forest_arrays(f, nq, R._lib.ptr(pts, c.c_double),
                                          R._lib.ptr(vM, c.c_int64), R._lib.ptr(vP, c.c_int64),
                                          R._lib.ptr(bc, c.c_int64), None))
    assert np.array_equal(pts, ref.points)
    assert np.array_equal(vM, ref.vmapM) and np.array_equal(vP, ref.vmapP)
    assert np.array_equal(bc, ref.boundarycodes)


@pytest.mark.parametrize("dim,nq,n,rounds,radius", [(2, 4, 8, 2, 2.5), (3, 3, 6, 2, 1.6),
                                                    (3, 4, 4, 3, 1.1)])
def test_adapted_forest_invariants(dim, nq, n, rounds, radius):
    gm = sphere_forest(dim, nq, n, rounds, radius)
    tree, level, quad = gm.cells()
    K = len(gm)
    assert level.max() == rounds  # the sphere reaches the finest requested level
    # leaves tile every tree exactly
    vol = np.zeros(n ** dim)
    lin = sum(tree[:, d] * n ** d for d in range(dim))
    np.add.at(vol, lin, 0.5 ** (dim * level.astype(np.float64)))
    assert np.allclose(vol, 1.0, rtol=0, atol=1e-15)
    # tree order (the brick's, i.e. level-0 cell order), then Morton inside a tree
    base = R.GridManager(gm.cell, gm.coarsegrid).cells()[0]
    rank = {tuple(t): i for i, t in enumerate(base)}
    keys = [(rank[tuple(tree[e])], morton(quad[e] << (19 - level[e]), dim)) for e in range(K)]
    assert keys == sorted(keys) and len(set(keys)) == K
    # 2:1 balance across faces, edges and corners: look up the leaf holding the
    # region next to every leaf at its own size
    finest = rounds
    owner = {}
    for e in range(K):
        s = 1 << (finest - level[e])
        lo = [(tree[e][d] << finest) + (quad[e][d] << (finest - level[e])) for d in range(dim)]
        for off in itertools.product(range(s), repeat=dim):
            owner[tuple(lo[d] + off[d] for d in range(dim))] = e
    N = n << finest
    for e in range(K):
        s = 1 << (finest - level[e])
        lo = [(tree[e][d] << finest) + (quad[e][d] << (finest - level[e])) for d in range(dim)]
        for delta in itertools.product((-1, 0, 1), repeat=dim):
            if not any(delta):
                continue
            p = tuple((lo[d] + (s if delta[d] > 0 else -1 if delta[d] < 0 else 0)) % N
                      for d in range(dim))
            assert abs(int(level[owner[p]]) - int(level[e])) <= 1
    # face maps: conforming faces pair up, their nodes coincide; the rest are
    # self-mapped with boundary code 0 on this periodic forest
    f = gm._forest
    Np, nfp = nq ** dim, nq ** (dim - 1)
    nfpt = 2 * dim * nfp
    pts = np.zeros((K, dim, Np))
    vM, vP = np.zeros((K, nfpt), np.int64), np.zeros((K, nfpt), np.int64)
    bc = np.zeros((K, 2 * dim), np.int64)
    kind = np.zeros((K, 2 * dim), np.int32)
    c = R._lib.ctypes
    R._lib.check(R.lib().rv_forest_arrays(f, nq, R._lib.ptr(pts, c.c_double),
                                          R._lib.ptr(vM, c.c_int64), R._lib.ptr(vP, c.c_int64),
                                          R._lib.ptr(bc, c.c_int64), R._lib.ptr(kind, c.c_int32)))
    assert np.all(bc == 0)
    flatM, flatP = vM.reshape(-1) - 1, vP.reshape(-1) - 1
    P = pts.transpose(0, 2, 1).reshape(K * Np, dim)  # node coordinates by volume index
    L = n  # periodic period
    for e in range(K):
        for fc in range(2 * dim):
            sl = slice(fc * nfp + e * nfpt, (fc + 1) * nfp + e * nfpt)
            if kind[e, fc] != 0:
                assert np.array_equal(flatP[sl], flatM[sl])
                continue
            # paired nodes coincide (up to the period) and the map is an involution
            d = P[flatP[sl]] - P[flatM[sl]]
            assert np.allclose(np.remainder(d + L / 2, L) - L / 2, 0, atol=1e-12)
            # the neighbour's opposite face maps back onto this face, node by node
            en = int(flatP[sl][0]) // Np
            assert kind[en, fc ^ 1] == 0
            back = slice((fc ^ 1) * nfp + en * nfpt, ((fc ^ 1) + 1) * nfp + en * nfpt)
            assert np.array_equal(flatP[back], flatM[sl])
    # hanging faces sit on a leaf one level finer than the neighbour covering it
    assert np.any(kind == 1) and np.any(kind == 2)


def leaf_set(gm):
    tree, level, quad = gm.cells()
    return [(tuple(t), int(lv), tuple(q)) for t, lv, q in zip(tree, level, quad)]


def assert_balanced(gm, n):
    """Every pair of leaves touching across a face, edge or corner differs by
    at most one level (periodic brick of n cells per direction)."""
    tree, level, quad = gm.cells()
    dim, K = tree.shape[1], len(gm)
    finest = int(level.max())
    owner = {}
    for e in range(K):
        s = 1 << (finest - level[e])
        lo = [(tree[e][d] << finest) + (quad[e][d] << (finest - level[e])) for d in range(dim)]
        for off in itertools.product(range(s), repeat=dim):
            owner[tuple(lo[d] + off[d] for d in range(dim))] = e
    N = n << finest
    for e in range(K):
        s = 1 << (finest - level[e])
        lo = [(tree[e][d] << finest) + (quad[e][d] << (finest - level[e])) for d in range(dim)]
        for delta in itertools.product((-1, 0, 1), repeat=dim):
            if any(delta):
                p = tuple((lo[d] + (s if delta[d] > 0 else -1 if delta[d] < 0 else 0)) % N
                          for d in range(dim))
                assert abs(int(level[owner[p]]) - int(level[e])) <= 1


def test_adapt_takes_ravens_flag_values():
    """AdaptNone (0x01) and AdaptTouched (0x11) keep a leaf, AdaptRefine (0x10)
    splits it (gridmanager.jl:24-27, refine_quads :133-136), and the flags
    come back as fill_adapt_flags writes them (:117-129)."""
    assert (R.AdaptCoarsen, R.AdaptNone, R.AdaptRefine, R.AdaptTouched) == (0, 1, 0x10, 0x11)
    gm = R.GridManager(R.LobattoCell(3, 3), R.brick([PB.julia_range(0, 1, 5)] * 2, (True, True)))
    before = leaf_set(gm)
    flags = np.full(len(gm), R.AdaptNone, dtype=np.uint8)
    flags[3] = R.AdaptTouched
    gm.adapt(flags)
    assert leaf_set(gm) == before
    assert np.all(flags == R.AdaptNone)
    # coarsening flags on level-0 leaves (trees) have no family: kept
    flags = np.full(len(gm), R.AdaptCoarsen, dtype=np.uint8)
    gm.adapt(flags)
    assert leaf_set(gm) == before and np.all(flags == R.AdaptNone)
    flags = np.full(len(gm), R.AdaptNone, dtype=np.uint8)
    flags[5] = R.AdaptRefine
    out = gm.adapt(list(flags))  # a list: the result is returned
    assert len(gm) == 16 + 3
    assert out[5] == R.AdaptRefine and np.sum(out == R.AdaptNone) == 15


@pytest.mark.parametrize("dim", [2, 3])
def test_refine_then_coarsen_restores_the_forest(dim):
    """Coarsening every complete family of a uniformly refined forest gives the
    coarse forest back leaf for leaf; the flags report AdaptCoarsen."""
    gm = R.GridManager(R.LobattoCell(*([3] * dim)),
                       R.brick([PB.julia_range(0, 1, 4)] * dim, (True,) * dim))
    before = leaf_set(gm)
    gm.adapt(np.full(len(gm), R.AdaptRefine, dtype=np.uint8))
    assert len(gm) == len(before) * 2 ** dim
    flags = np.full(len(gm), R.AdaptCoarsen, dtype=np.uint8)
    gm.adapt(flags)
    assert leaf_set(gm) == before
    assert np.all(flags == R.AdaptCoarsen)


def test_coarsening_needs_the_whole_family():
    """A family with one member not flagged AdaptCoarsen stays (coarsen_quads
    requires every child, gridmanager.jl:138-146); coarsening is not recursive
    (one level per call)."""
    gm = R.GridManager(R.LobattoCell(3, 3), R.brick([PB.julia_range(0, 1, 3)] * 2, (True, True)),
                       min_level=2)
    K = len(gm)
    flags = np.full(K, R.AdaptCoarsen, dtype=np.uint8)
    flags[0] = R.AdaptNone  # breaks the first family only
    gm.adapt(flags)
    tree, level, _ = gm.cells()
    assert len(gm) == K - 3 * (K // 4 - 1)
    assert np.sum(level == 2) == 4 and np.sum(level == 1) == K // 4 - 1
    assert flags[0] == R.AdaptNone and np.sum(flags == R.AdaptCoarsen) == K - 4
    assert np.all(flags[1:4] == R.AdaptNone)


def test_coarsening_is_rebalanced():
    """A coarsened family next to a two-level-finer region is refined again by
    the 2:1 balance (coarsen!, refine!, balance! in that order,
    gridmanager.jl:160-172), and every other family coarsens."""
    n = 4
    gm = sphere_forest(2, 3, n, 2, 1.2)
    tree, level, quad = gm.cells()
    assert level.max() == 2
    flags = np.where(level == 1, R.AdaptCoarsen, R.AdaptNone).astype(np.uint8)
    gm.adapt(flags)
    assert_balanced(gm, n)
    _, level2, _ = gm.cells()
    assert np.sum(level2 == 2) == np.sum(level == 2)  # the finest leaves are kept
    # families next to level-2 leaves come back at level 1 (AdaptNone); the
    # others report AdaptCoarsen
    assert np.any(flags == R.AdaptCoarsen) and np.any((level == 1) & (flags == R.AdaptNone))
    # the balanced forest is the coarsest one: coarsening again changes nothing
    # where balance forbids it
    before = leaf_set(gm)
    tree, lv, quad = gm.cells()
    gm.adapt(np.where(lv == 1, R.AdaptCoarsen, R.AdaptNone).astype(np.uint8))
    assert leaf_set(gm) == before
