// Adaptive brick forest: the leaves Raven's GridManager holds after
// adapt!(gm, flags) (gridmanager.jl:160-172), kept without p4est.
//
//  * the coarse brick's trees come in the brick's tree order (the same order
//    brick_quadrants uses); inside a tree the leaves are in Morton order of
//    their lower corners (x fastest), children of a refined leaf in the order
//    (x, y, z) bits of the child index;
//  * adapt follows adapt! (gridmanager.jl:160-172): first every complete
//    family of sibling leaves whose flags are all AdaptCoarsen (0x00) is
//    replaced by its parent, once (p4est's non-recursive coarsen, the callback
//    coarsen_quads :138-146); then every remaining leaf flagged AdaptRefine
//    (0x10) is split into its children, once (refine_quads :133-136); any
//    other flag value keeps the leaf.  Then the forest is balanced 2:1 across
//    faces, edges and corners: a leaf whose same-size neighbour region lies
//    inside a leaf more than one level coarser forces that coarse leaf to
//    refine, repeated until no such pair is left (each refinement is forced,
//    so this is the coarsest balanced refinement).  Last, the flags are
//    overwritten with what happened to each old leaf, as fill_adapt_flags
//    (:117-129) does: AdaptRefine if its successors are finer, AdaptCoarsen if
//    coarser, AdaptNone (0x01) otherwise, written at the index of the old
//    leaf each new leaf descends from (a coarsened family's first child);
//  * faces are classified against the same-size neighbour region: a leaf of the
//    same level there is a conforming neighbour (brick trees are aligned, so
//    the neighbour face is the opposite one with the identity orientation); a
//    coarser leaf (this face hangs) or finer leaves (this face is covered by
//    several) make the face non-conforming, which Raven's face maps leave
//    self-mapped (lobattocells.jl:2072-2074); outside a non-periodic brick the
//    face is a boundary face (code 1).
//
// Whether p4est's 2:1 balance and numbering produce exactly these leaves for
// every forest is not pinned (p4est is not available here); the leaves satisfy
// the invariants the tests check.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <array>
#include <numeric>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace rv {

namespace {

constexpr int MAXL = 19;  // finest level (octant coordinates fit 3 x 19 bits)

struct Leaf {
    int32_t t[3];   // tree lattice coordinates
    int32_t level;
    int32_t c[3];   // coordinates inside the tree at the leaf's own level
    // during adapt: the index before the call of the leaf this one descends
    // from, and that leaf's level (QuadData's old_id / old_level)
    int64_t origin;
    int32_t old_level;
};

constexpr uint8_t ADAPT_COARSEN = 0x00, ADAPT_NONE = 0x01, ADAPT_REFINE = 0x10;

uint64_t spread_bits(uint64_t v, int dim)
{
    uint64_t out = 0;
    for (int b = 0; b < MAXL; ++b)
        out |= ((v >> b) & 1ull) << (b * dim);
    return out;
}

// volume index of node m of face f (faces -x,+x,-y,+y,-z,+z; face nodes over
// the remaining directions, first fastest: lobattocells.jl:2020-2050)
int64_t face_node_of(int dim, int nq, int f, int64_t m)
{
    const int d = f / 2;
    const int64_t fixed = (f & 1) ? nq - 1 : 0;
    if (dim == 2)
        return d == 0 ? fixed + nq * m : m + nq * fixed;
    const int64_t a = m % nq, c = m / nq;
    if (d == 0)
        return fixed + nq * a + (int64_t)nq * nq * c;
    if (d == 1)
        return a + nq * fixed + (int64_t)nq * nq * c;
    return a + nq * c + (int64_t)nq * nq * fixed;
}

struct KeyHash {
    size_t operator()(const std::pair<uint64_t, uint64_t> &k) const
    {
        return std::hash<uint64_t>()(k.first * 0x9E3779B97F4A7C15ull ^ k.second);
    }
};

}  // namespace

struct Forest {
    BrickSpec b;
    std::vector<int64_t> tree_rank;  // lattice tree index -> position in tree order
    std::vector<Leaf> leaves;
    std::unordered_map<std::pair<uint64_t, uint64_t>, int64_t, KeyHash> index;

    int64_t tree_lin(const int32_t *t) const
    {
        int64_t lin = 0, stride = 1;
        for (int d = 0; d < b.dim; ++d) {
            lin += t[d] * stride;
            stride *= b.n[d];
        }
        return lin;
    }
    std::pair<uint64_t, uint64_t> key(const int32_t *t, int level, const int32_t *c) const
    {
        uint64_t m = 0;
        for (int d = 0; d < b.dim; ++d)
            m |= spread_bits((uint64_t)c[d], b.dim) << d;
        return {((uint64_t)tree_lin(t) << 5) | (uint64_t)level, m};
    }
    // order key: tree position, then the Morton code of the lower corner at MAXL
    std::pair<int64_t, uint64_t> order(const Leaf &l) const
    {
        uint64_t m = 0;
        for (int d = 0; d < b.dim; ++d)
            m |= spread_bits((uint64_t)l.c[d] << (MAXL - l.level), b.dim) << d;
        return {tree_rank[tree_lin(l.t)], m};
    }
    void reindex()
    {
        std::sort(leaves.begin(), leaves.end(),
                  [&](const Leaf &x, const Leaf &y) { return order(x) < order(y); });
        index.clear();
        index.reserve(leaves.size() * 2);
        for (int64_t i = 0; i < (int64_t)leaves.size(); ++i)
            index[key(leaves[i].t, leaves[i].level, leaves[i].c)] = i;
    }
    // The same-size octant next to leaf l in direction delta (per-axis -1, 0,
    // +1), moved into the neighbouring tree when it leaves l's tree; false when
    // it lies outside a non-periodic brick.
    bool shifted(const Leaf &l, const int *delta, Leaf *out) const
    {
        *out = l;
        const int32_t side = 1 << l.level;
        for (int d = 0; d < b.dim; ++d) {
            int32_t c = l.c[d] + delta[d], t = l.t[d];
            if (c < 0) {
                c += side;
                t -= 1;
            } else if (c >= side) {
                c -= side;
                t += 1;
            }
            if (t < 0 || t >= b.n[d]) {
                if (!b.periodic[d])
                    return false;
                t = (t + b.n[d]) % b.n[d];
            }
            out->c[d] = c;
            out->t[d] = t;
        }
        return true;
    }
    // The leaf containing the octant o (its own level or coarser): index, or
    // -1 when o is covered by finer leaves.
    int64_t containing(const Leaf &o, int32_t *level_out) const
    {
        for (int lev = o.level; lev >= 0; --lev) {
            int32_t c[3] = {0, 0, 0};
            for (int d = 0; d < b.dim; ++d)
                c[d] = o.c[d] >> (o.level - lev);
            auto it = index.find(key(o.t, lev, c));
            if (it != index.end()) {
                *level_out = lev;
                return it->second;
            }
        }
        return -1;
    }
    void refine(const std::vector<char> &mark)
    {
        std::vector<Leaf> next;
        next.reserve(leaves.size() + 8 * std::count(mark.begin(), mark.end(), 1));
        const int nchild = 1 << b.dim;
        for (size_t i = 0; i < leaves.size(); ++i) {
            if (!mark[i]) {
                next.push_back(leaves[i]);
                continue;
            }
            for (int ch = 0; ch < nchild; ++ch) {
                Leaf k = leaves[i];
                k.level += 1;
                for (int d = 0; d < b.dim; ++d)
                    k.c[d] = 2 * leaves[i].c[d] + ((ch >> d) & 1);
                next.push_back(k);
            }
        }
        leaves.swap(next);
        reindex();
    }
    void balance()
    {
        const int ndir = b.dim == 3 ? 27 : 9;
        for (;;) {
            std::vector<char> mark(leaves.size(), 0);
            bool any = false;
            for (const Leaf &l : leaves) {
                if (l.level < 2)
                    continue;
                for (int code = 0; code < ndir; ++code) {
                    int delta[3] = {code % 3 - 1, (code / 3) % 3 - 1, b.dim == 3 ? code / 9 - 1 : 0};
                    if (b.dim == 2)
                        delta[2] = 0;
                    if (!delta[0] && !delta[1] && !delta[2])
                        continue;
                    Leaf o;
                    if (!shifted(l, delta, &o))
                        continue;
                    int32_t lev = 0;
                    const int64_t c = containing(o, &lev);
                    if (c >= 0 && lev < l.level - 1 && !mark[c]) {
                        mark[c] = 1;
                        any = true;
                    }
                }
            }
            if (!any)
                return;
            refine(mark);
        }
    }
};

Forest *forest_new(const BrickSpec &b)
{
    Forest *f = new Forest();
    f->b = b;
    const int64_t K = brick_numcells(b);
    std::vector<int32_t> tree(K * b.dim), quad(K * b.dim);
    brick_quadrants(b, tree.data(), quad.data());
    int64_t ntrees = 1;
    for (int d = 0; d < b.dim; ++d)
        ntrees *= b.n[d];
    f->tree_rank.assign(ntrees, 0);
    const int64_t per_tree = K / ntrees;
    f->leaves.resize(K);
    for (int64_t e = 0; e < K; ++e) {
        Leaf &l = f->leaves[e];
        l.level = b.level;
        for (int d = 0; d < 3; ++d) {
            l.t[d] = d < b.dim ? tree[e * b.dim + d] : 0;
            l.c[d] = d < b.dim ? quad[e * b.dim + d] : 0;
        }
        if (e % per_tree == 0)
            f->tree_rank[f->tree_lin(l.t)] = e / per_tree;
    }
    f->reindex();
    return f;
}

void forest_free(Forest *f) { delete f; }

Forest *forest_clone(const Forest *f) { return new Forest(*f); }

BrickSpec forest_spec(const Forest *f) { return f->b; }

void forest_faces(const Forest *f, int64_t *nbr, int32_t *kind, int64_t *bc)
{
    const int dim = f->b.dim, nf = 2 * dim;
    const int64_t K = forest_size(f);
    for (int64_t e = 0; e < K; ++e) {
        const Leaf &l = f->leaves[e];
        for (int fc = 0; fc < nf; ++fc) {
            int delta[3] = {0, 0, 0};
            delta[fc / 2] = (fc & 1) ? 1 : -1;
            Leaf o;
            int64_t nb = -1;
            int code = 0, k = 0;
            if (!f->shifted(l, delta, &o)) {
                code = 1;
            } else {
                int32_t lev = 0;
                const int64_t c = f->containing(o, &lev);
                if (c < 0)
                    k = 2;
                else if (lev < l.level)
                    k = 1;
                else
                    nb = c;
            }
            if (nbr)
                nbr[e * nf + fc] = nb;
            if (kind)
                kind[e * nf + fc] = k;
            if (bc)
                bc[e * nf + fc] = code;
        }
    }
}

int64_t forest_size(const Forest *f) { return (int64_t)f->leaves.size(); }

int forest_adapt(Forest *f, uint8_t *flags)
{
    std::vector<Leaf> &lv = f->leaves;
    const int64_t K0 = (int64_t)lv.size();
    const int dim = f->b.dim, nchild = 1 << dim;
    for (int64_t i = 0; i < K0; ++i) {
        if (flags[i] == ADAPT_REFINE && lv[i].level >= MAXL)
            return -1;
        lv[i].origin = i;
        lv[i].old_level = lv[i].level;
    }
    // coarsen: complete families (consecutive in Morton order, child j at
    // position j) all flagged AdaptCoarsen; the parent carries the first
    // child's origin and is not refined in this call
    std::vector<Leaf> next;
    std::vector<char> touched;
    next.reserve(K0);
    touched.reserve(K0);
    for (int64_t i = 0; i < K0;) {
        const Leaf &l = lv[i];
        bool family = l.level > 0 && i + nchild <= K0;
        for (int j = 0; family && j < nchild; ++j) {
            const Leaf &s = lv[i + j];
            if (s.level != l.level || flags[i + j] != ADAPT_COARSEN)
                family = false;
            for (int d = 0; family && d < dim; ++d)
                if (s.t[d] != l.t[d] || (s.c[d] >> 1) != (l.c[d] >> 1) ||
                    (s.c[d] & 1) != ((j >> d) & 1))
                    family = false;
        }
        if (family) {
            Leaf p = l;
            p.level -= 1;
            for (int d = 0; d < dim; ++d)
                p.c[d] = l.c[d] >> 1;
            next.push_back(p);
            touched.push_back(1);
            i += nchild;
        } else {
            next.push_back(l);
            touched.push_back(0);
            i += 1;
        }
    }
    std::vector<char> mark(next.size(), 0);
    for (size_t i = 0; i < next.size(); ++i)
        mark[i] = !touched[i] && flags[next[i].origin] == ADAPT_REFINE;
    lv.swap(next);
    f->refine(mark);
    f->balance();
    for (const Leaf &l : lv)
        flags[l.origin] = l.level > l.old_level   ? ADAPT_REFINE
                          : l.level < l.old_level ? ADAPT_COARSEN
                                                  : ADAPT_NONE;
    return 0;
}

void forest_cells(const Forest *f, int32_t *tree, int32_t *level, int32_t *quad)
{
    const int dim = f->b.dim;
    for (size_t i = 0; i < f->leaves.size(); ++i) {
        const Leaf &l = f->leaves[i];
        for (int d = 0; d < dim; ++d) {
            if (tree)
                tree[i * dim + d] = l.t[d];
            if (quad)
                quad[i * dim + d] = l.c[d];
        }
        if (level)
            level[i] = l.level;
    }
}

void forest_arrays(const Forest *f, int nq, double *points, int64_t *vmapM, int64_t *vmapP,
                   int64_t *bc, int32_t *kind)
{
    const BrickSpec &b = f->b;
    const int dim = b.dim, nf = 2 * dim;
    const int64_t K = forest_size(f);
    int64_t Np = 1, nfp = 1;
    for (int d = 0; d < dim; ++d)
        Np *= nq;
    nfp = Np / nq;
    const int64_t nfpt = nf * nfp;
    if (points) {
        std::vector<double> r(nq), w(nq), D(nq * nq), verts[3];
        lobatto_operators(nq, r.data(), w.data(), D.data());
        for (int d = 0; d < dim; ++d) {
            verts[d].resize(b.n[d] + 1);
            brick_vertices(b, d, verts[d].data());
        }
        for (int64_t e = 0; e < K; ++e) {
            const Leaf &l = f->leaves[e];
            const double h = 1.0 / (double)(1 << (l.level + 1));
            for (int64_t node = 0; node < Np; ++node) {
                int64_t rem = node;
                for (int d = 0; d < dim; ++d) {
                    const double c1 = verts[d][l.t[d]];
                    const double dx = verts[d][l.t[d] + 1] - verts[d][l.t[d]];
                    const double cr =
                        (double)((int64_t)l.c[d] << (30 - l.level)) / (double)(1 << 30);
                    points[node + Np * (d + (int64_t)dim * e)] =
                        std::fma(dx, std::fma(h, r[rem % nq] + 1, cr), c1);
                    rem /= nq;
                }
            }
        }
    }
    for (int64_t e = 0; e < K; ++e) {
        const Leaf &l = f->leaves[e];
        for (int fc = 0; fc < nf; ++fc) {
            int delta[3] = {0, 0, 0};
            delta[fc / 2] = (fc & 1) ? 1 : -1;
            Leaf o;
            int64_t nb = -1;
            int code = 0, k = 0;  // k: 0 conforming or boundary, 1 hanging, 2 covered
            if (!f->shifted(l, delta, &o)) {
                code = 1;
            } else {
                int32_t lev = 0;
                const int64_t c = f->containing(o, &lev);
                if (c < 0)
                    k = 2;
                else if (lev < l.level)
                    k = 1;
                else
                    nb = c;
            }
            if (bc)
                bc[fc + nf * e] = code;
            if (kind)
                kind[fc + nf * e] = k;
            for (int64_t m = 0; m < nfp; ++m) {
                const int64_t idx = fc * nfp + m + nfpt * e;
                const int64_t self = face_node_of(dim, nq, fc, m) + Np * e + 1;
                if (vmapM)
                    vmapM[idx] = self;
                if (vmapP)
                    vmapP[idx] = nb < 0 ? self : face_node_of(dim, nq, fc ^ 1, m) + Np * nb + 1;
            }
        }
    }
}

void forest_mortars(const Forest *f, int64_t *partner)
{
    const BrickSpec &b = f->b;
    const int dim = b.dim, nf = 2 * dim;
    const int64_t K = forest_size(f);
    for (int64_t e = 0; e < K; ++e) {
        const Leaf &l = f->leaves[e];
        for (int fc = 0; fc < nf; ++fc) {
            int64_t *out = partner + 4 * (fc + nf * e);
            for (int k = 0; k < 4; ++k)
                out[k] = -1;
            const int d = fc / 2;
            int tang[2] = {0, 0}, nt = 0;
            for (int c = 0; c < dim; ++c)
                if (c != d)
                    tang[nt++] = c;
            int delta[3] = {0, 0, 0};
            delta[d] = (fc & 1) ? 1 : -1;
            Leaf o;
            if (!f->shifted(l, delta, &o))
                continue;
            int32_t lev = 0;
            const int64_t c = f->containing(o, &lev);
            if (c >= 0 && lev < l.level) {
                // hanging: the sub-face of the coarse face this leaf covers
                const Leaf &cl = f->leaves[c];
                int sub = 0;
                for (int k = 0; k < nt; ++k)
                    sub |= (l.c[tang[k]] - 2 * cl.c[tang[k]]) << k;
                out[0] = c;
                out[1] = sub;
            } else if (c < 0) {
                // covered: the finer leaves touching this face, sub-face order
                for (int sub = 0; sub < (1 << nt); ++sub) {
                    Leaf ch = o;
                    ch.level = o.level + 1;
                    for (int cc = 0; cc < dim; ++cc)
                        ch.c[cc] = 2 * o.c[cc];
                    ch.c[d] += (fc & 1) ? 0 : 1;
                    for (int k = 0; k < nt; ++k)
                        ch.c[tang[k]] += (sub >> k) & 1;
                    int32_t lv = 0;
                    const int64_t fine = f->containing(ch, &lv);
                    out[sub] = (fine >= 0 && lv == ch.level) ? fine : -2;
                }
            }
        }
    }
}

namespace {

// Non-conforming classification of face fc of leaf e: kind 1 (hanging: the
// coarser leaf *c and the sub-face s1 + 2 s2 this leaf covers), 2 (covered:
// the finer leaves in sub-face order in cells[]), 0 otherwise.
int nc_face(const Forest *f, int64_t e, int fc, int64_t *c_out, int *sub_out, int64_t *cells)
{
    const BrickSpec &b = f->b;
    const int dim = b.dim, d = fc / 2;
    const Leaf &l = f->leaves[e];
    int tang[2] = {0, 0}, nt = 0;
    for (int c = 0; c < dim; ++c)
        if (c != d)
            tang[nt++] = c;
    int delta[3] = {0, 0, 0};
    delta[d] = (fc & 1) ? 1 : -1;
    Leaf o;
    if (!f->shifted(l, delta, &o))
        return 0;
    int32_t lev = 0;
    const int64_t c = f->containing(o, &lev);
    if (c >= 0 && lev < l.level) {
        const Leaf &cl = f->leaves[c];
        int sub = 0;
        for (int k = 0; k < nt; ++k)
            sub |= (l.c[tang[k]] - 2 * cl.c[tang[k]]) << k;
        *c_out = c;
        *sub_out = sub;
        return 1;
    }
    if (c >= 0)
        return 0;
    for (int sub = 0; sub < (1 << nt); ++sub) {
        Leaf ch = o;
        ch.level = o.level + 1;
        for (int cc = 0; cc < dim; ++cc)
            ch.c[cc] = 2 * o.c[cc];
        ch.c[d] += (fc & 1) ? 0 : 1;
        for (int k = 0; k < nt; ++k)
            ch.c[tang[k]] += (sub >> k) & 1;
        int32_t lv = 0;
        const int64_t fine = f->containing(ch, &lv);
        cells[sub] = (fine >= 0 && lv == ch.level) ? fine : -1;
    }
    return 2;
}

struct NcScan {
    // per group: interface key (parent cell, parent face) -> 0-based id
    std::vector<std::unordered_map<int64_t, int64_t>> ids;
    std::vector<int64_t> nfaces;
};

// The first loop of materializefacemaps (lobattocells.jl:2084-2110): the
// interfaces of each face group numbered in order of first encounter over
// cells, then faces.
NcScan nc_scan(const Forest *f)
{
    const int dim = f->b.dim, nf = 2 * dim;
    NcScan sc;
    sc.ids.resize(dim);
    sc.nfaces.assign(dim, 0);
    const int64_t K = forest_size(f);
    int64_t cells[4];
    for (int64_t e = 0; e < K; ++e)
        for (int fc = 0; fc < nf; ++fc) {
            int64_t c = -1;
            int sub = 0;
            const int k = nc_face(f, e, fc, &c, &sub, cells);
            if (!k)
                continue;
            const int g = fc / 2;
            const int64_t key = k == 2 ? e * nf + fc : c * nf + (fc ^ 1);
            sc.ids[g].emplace(key, (int64_t)sc.ids[g].size());
            sc.nfaces[g]++;
        }
    return sc;
}

}  // namespace

void forest_nc_counts(const Forest *f, int64_t *nunique, int64_t *nfaces)
{
    const NcScan sc = nc_scan(f);
    for (int g = 0; g < f->b.dim; ++g) {
        if (nunique)
            nunique[g] = (int64_t)sc.ids[g].size();
        if (nfaces)
            nfaces[g] = sc.nfaces[g];
    }
}

// The second loop of materializefacemaps (lobattocells.jl:2155-2272) on an
// aligned brick forest: every face orientation is the identity (the four
// face-interior degree-3 ids of a face, and of the hanging faces that take a
// coarse face's ids, ascend in face order), so vmapNC lists the parent face's
// nodes and then each child's, in face node order; the children are in
// global-id order, which is the sub-face order s1 + 2 s2 (the order the
// reference's mpitest_n2_gridnumbering.jl:339-366 checks with tohalves).
void forest_nctables(const Forest *f, int nq, int64_t *const *vmapNC, int8_t *const *nctoface,
                     int8_t *const *nctypes, int64_t *const *ncids)
{
    const int dim = f->b.dim, nf = 2 * dim, nchild = 1 << (dim - 1);
    int64_t Np = 1;
    for (int d = 0; d < dim; ++d)
        Np *= nq;
    const int64_t nfp = Np / nq;
    const NcScan sc = nc_scan(f);
    std::vector<int64_t> k(dim, 0);
    std::vector<std::vector<char>> filled(dim);
    for (int g = 0; g < dim; ++g)
        filled[g].assign(sc.ids[g].size(), 0);
    const int64_t K = forest_size(f);
    int64_t cells[4];
    for (int64_t e = 0; e < K; ++e)
        for (int fc = 0; fc < nf; ++fc) {
            int64_t c = -1;
            int sub = 0;
            const int kind = nc_face(f, e, fc, &c, &sub, cells);
            if (!kind)
                continue;
            const int g = fc / 2;
            const int64_t parent = kind == 2 ? e : c;
            const int pface = kind == 2 ? fc : (fc ^ 1);
            const int64_t id = sc.ids[g].at(parent * nf + pface);
            if (ncids && ncids[g])
                ncids[g][k[g]] = id + 1;
            if (nctypes && nctypes[g])
                nctypes[g][k[g]] = (int8_t)(kind == 2 ? 1 : sub + 2);
            k[g]++;
            if (filled[g][id])
                continue;
            filled[g][id] = 1;
            const int cface = pface ^ 1;
            if (nctoface && nctoface[g]) {
                nctoface[g][2 * id] = (int8_t)(pface + 1);
                nctoface[g][2 * id + 1] = (int8_t)(cface + 1);
            }
            if (!(vmapNC && vmapNC[g]))
                continue;
            int64_t kids[4];
            if (kind == 2) {
                std::copy(cells, cells + nchild, kids);
            } else {
                int sb = 0;
                int64_t cc = -1;
                nc_face(f, parent, pface, &cc, &sb, kids);
            }
            int64_t *out = vmapNC[g] + (size_t)nfp * (1 + nchild) * id;
            for (int64_t m = 0; m < nfp; ++m) {
                out[m] = face_node_of(dim, nq, pface, m) + Np * parent + 1;
                for (int ch = 0; ch < nchild; ++ch)
                    out[m + nfp * (1 + ch)] = face_node_of(dim, nq, cface, m) + Np * kids[ch] + 1;
            }
        }
}

int mortars_from_nctables(int dim, int nq, int64_t K, const int64_t *nunique,
                          const int64_t *const *vmapNC, const int8_t *const *nctoface,
                          int32_t *kind, int64_t *partner)
{
    const int nf = 2 * dim, nchild = 1 << (dim - 1);
    int64_t Np = 1;
    for (int d = 0; d < dim; ++d)
        Np *= nq;
    const int64_t nfp = Np / nq;
    for (int64_t i = 0; i < nf * K; ++i) {
        kind[i] = 0;
        for (int j = 0; j < 4; ++j)
            partner[4 * i + j] = -1;
    }
    // node v (1-based) of the tables: its cell, checked against the face
    // node it must be in an aligned pairing
    auto cell_of = [&](int64_t v, int face, int64_t m, int64_t *cell) {
        if (v < 1 || v > Np * K)
            return -1;
        *cell = (v - 1) / Np;
        return (v - 1) % Np == face_node_of(dim, nq, face, m) ? 0 : -2;
    };
    for (int g = 0; g < dim; ++g) {
        if (nunique[g] == 0)
            continue;
        if (!vmapNC[g] || !nctoface[g])
            return -1;
        for (int64_t id = 0; id < nunique[g]; ++id) {
            const int pface = nctoface[g][2 * id] - 1, cface = nctoface[g][2 * id + 1] - 1;
            if (pface < 0 || pface >= nf || pface / 2 != g || cface != (pface ^ 1))
                return -2;
            const int64_t *tab = vmapNC[g] + (size_t)nfp * (1 + nchild) * id;
            int64_t parent = -1, kids[4] = {-1, -1, -1, -1};
            for (int64_t m = 0; m < nfp; ++m) {
                int64_t cell = -1;
                int rc = cell_of(tab[m], pface, m, &cell);
                if (rc)
                    return rc;
                if (m > 0 && cell != parent)
                    return -2;
                parent = cell;
                for (int ch = 0; ch < nchild; ++ch) {
                    rc = cell_of(tab[m + nfp * (1 + ch)], cface, m, &cell);
                    if (rc)
                        return rc;
                    if (m > 0 && cell != kids[ch])
                        return -2;
                    kids[ch] = cell;
                }
            }
            kind[pface + nf * parent] = 2;
            for (int ch = 0; ch < nchild; ++ch) {
                partner[4 * (pface + nf * parent) + ch] = kids[ch];
                kind[cface + nf * kids[ch]] = 1;
                partner[4 * (cface + nf * kids[ch])] = parent;
                partner[4 * (cface + nf * kids[ch]) + 1] = ch;
            }
        }
    }
    return 0;
}

namespace {

struct Dsu {
    std::vector<int64_t> p;
    explicit Dsu(int64_t n) : p(n) { std::iota(p.begin(), p.end(), 0); }
    int64_t find(int64_t x)
    {
        while (p[x] != x) {
            p[x] = p[p[x]];
            x = p[x];
        }
        return x;
    }
    void unite(int64_t a, int64_t b)
    {
        a = find(a);
        b = find(b);
        if (a != b)
            p[std::max(a, b)] = std::min(a, b);
    }
};

struct Key3Hash {
    size_t operator()(const std::array<int64_t, 3> &k) const
    {
        uint64_t h = 0x9E3779B97F4A7C15ull;
        for (int64_t v : k)
            h = (h ^ (uint64_t)v) * 0x100000001B3ull + (h >> 29);
        return (size_t)h;
    }
};

}  // namespace

// Continuous numbering of an adapted forest, as Raven builds it from p4est's
// degree-3 lnodes (gridmanager.jl:349-379, 479-481), materializedtoc
// (lobattocells.jl:1718-1870) and renumberdtocdtop! (:1872-1936):
//  * a node of a cell lying on a hanging face (the neighbour across it is
//    coarser) or, in 3-D, on a hanging edge (a coarser leaf diagonally across
//    the edge, neither face at the edge hanging) is not a degree of freedom
//    of its own: lnodes gives it the node of
//    the coarse face or edge at the same position of that face or edge, as
//    if the fine face were the coarse one (the reference test's numbering in
//    gridnumbering.jl:101-111 shows this: the children's nodes along a hanging
//    edge repeat the coarse edge's ids 9 11 10 and 9 18 21 by position);
//  * every other node is shared by the cells it coincides with (faces and
//    edges between leaves of one level, corners between any leaves);
//  * ids are then numbered by first appearance over cells and their nodes in
//    column-major order, which makes them independent of the order lnodes
//    numbered them in.
// Hanging nodes are joined to their coarse node, the rest by position key, in
// one union-find over (cell, node).
int64_t forest_dtoc(const Forest *f, int nq, int64_t *dtoc)
{
    const BrickSpec &b = f->b;
    const int dim = b.dim, nf = 2 * dim;
    const int64_t K = forest_size(f);
    int64_t Np = 1;
    for (int d = 0; d < dim; ++d)
        Np *= nq;
    // coarser leaf across each face, and (3-D) across each edge only
    // (edge a*4 + sb + 2 sc: along axis a, at sides sb, sc of the other two)
    std::vector<int64_t> fcoarse(K * nf, -1), ecoarse(dim == 3 ? K * 12 : 0, -1);
    for (int64_t e = 0; e < K; ++e) {
        const Leaf &l = f->leaves[e];
        for (int fc = 0; fc < nf; ++fc) {
            int delta[3] = {0, 0, 0};
            delta[fc / 2] = (fc & 1) ? 1 : -1;
            Leaf o;
            if (!f->shifted(l, delta, &o))
                continue;
            int32_t lev = 0;
            const int64_t c = f->containing(o, &lev);
            if (c >= 0 && lev < l.level)
                fcoarse[e * nf + fc] = c;
        }
        if (dim != 3)
            continue;
        for (int a = 0; a < 3; ++a) {
            const int ax = a == 0 ? 1 : 0, bx = a == 2 ? 1 : 2;
            for (int sd = 0; sd < 4; ++sd) {
                int delta[3] = {0, 0, 0};
                delta[ax] = (sd & 1) ? 1 : -1;
                delta[bx] = (sd & 2) ? 1 : -1;
                Leaf o;
                if (!f->shifted(l, delta, &o))
                    continue;
                int32_t lev = 0;
                const int64_t c = f->containing(o, &lev);
                if (c >= 0 && lev < l.level)
                    ecoarse[e * 12 + a * 4 + sd] = c;
            }
        }
    }
    Dsu dsu(K * Np);
    std::unordered_map<std::array<int64_t, 3>, int64_t, Key3Hash> shared;
    shared.reserve((size_t)(K * Np / 2));
    int64_t period[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d)
        period[d] = (int64_t)b.n[d] << MAXL;
    for (int64_t e = 0; e < K; ++e) {
        const Leaf &l = f->leaves[e];
        const int64_t size = (int64_t)1 << (MAXL - l.level);
        for (int64_t node = 0; node < Np; ++node) {
            int ii[3] = {0, 0, 0}, side[3] = {1, 1, 1};  // side 0 min, 1 interior, 2 max
            int64_t rem = node;
            for (int d = 0; d < dim; ++d) {
                ii[d] = (int)(rem % nq);
                rem /= nq;
                side[d] = ii[d] == 0 ? 0 : ii[d] == nq - 1 ? 2 : 1;
            }
            const int64_t me = e * Np + node;
            bool hanging = false;
            auto target = [&](int64_t c, const int *jj) {
                int64_t idx = 0, stride = 1;
                for (int d = 0; d < dim; ++d) {
                    idx += jj[d] * stride;
                    stride *= nq;
                }
                dsu.unite(me, c * Np + idx);
                hanging = true;
            };
            for (int d = 0; d < dim; ++d) {
                if (side[d] == 1)
                    continue;
                const int64_t c = fcoarse[e * nf + 2 * d + (side[d] == 2)];
                if (c < 0)
                    continue;
                int jj[3] = {ii[0], ii[1], ii[2]};
                jj[d] = side[d] == 2 ? 0 : nq - 1;
                target(c, jj);
            }
            if (dim == 3)
                for (int a = 0; a < 3; ++a) {
                    const int ax = a == 0 ? 1 : 0, bx = a == 2 ? 1 : 2;
                    if (side[ax] == 1 || side[bx] == 1)
                        continue;
                    // an edge on a hanging face is that face's (the diagonal
                    // coarser leaf may be the face's own coarse neighbour,
                    // whose face the edge then crosses)
                    if (fcoarse[e * nf + 2 * ax + (side[ax] == 2)] >= 0 ||
                        fcoarse[e * nf + 2 * bx + (side[bx] == 2)] >= 0)
                        continue;
                    const int sd = (side[ax] == 2) + 2 * (side[bx] == 2);
                    const int64_t c = ecoarse[e * 12 + a * 4 + sd];
                    if (c < 0)
                        continue;
                    int jj[3] = {ii[0], ii[1], ii[2]};
                    jj[ax] = side[ax] == 2 ? 0 : nq - 1;
                    jj[bx] = side[bx] == 2 ? 0 : nq - 1;
                    target(c, jj);
                }
            if (hanging)
                continue;
            // shared by position: boundary planes on the finest lattice
            // (wrapped on periodic axes); along an axis where the node is
            // inside the cell, the cell's extent and the node's index
            std::array<int64_t, 3> key{0, 0, 0};
            for (int d = 0; d < dim; ++d) {
                const int64_t lo = ((int64_t)l.t[d] << MAXL) + ((int64_t)l.c[d] << (MAXL - l.level));
                if (side[d] == 1) {
                    key[d] = (((lo << 5) | l.level) << 8 | ii[d]) << 1 | 1;
                } else {
                    int64_t x = side[d] == 2 ? lo + size : lo;
                    if (b.periodic[d])
                        x %= period[d];
                    key[d] = x << 1;
                }
            }
            auto it = shared.emplace(key, me);
            if (!it.second)
                dsu.unite(me, it.first->second);
        }
    }
    std::vector<int64_t> id(K * Np, 0);
    int64_t n = 0;
    for (int64_t i = 0; i < K * Np; ++i) {
        const int64_t r = dsu.find(i);
        if (!id[r])
            id[r] = ++n;
        dtoc[i] = id[r];
    }
    return n;
}

}  // namespace rv
