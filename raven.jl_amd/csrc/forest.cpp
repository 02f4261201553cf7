// Adaptive brick forest: the leaves Raven's GridManager holds after
// adapt!(gm, flags) (gridmanager.jl:160-172), kept without p4est.
//
//  * the coarse brick's trees come in the brick's tree order (the same order
//    brick_quadrants uses); inside a tree the leaves are in Morton order of
//    their lower corners (x fastest), children of a refined leaf in the order
//    (x, y, z) bits of the child index;
//  * adapt refines every flagged leaf by one level (AdaptRefine = 0x10,
//    gridmanager.jl:27; coarsening is not supported) and then balances the
//    forest 2:1 across faces, edges and corners: a leaf whose same-size
//    neighbour region lies inside a leaf more than one level coarser forces
//    that coarse leaf to refine, repeated until no such pair is left;
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
};

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

int64_t forest_size(const Forest *f) { return (int64_t)f->leaves.size(); }

int forest_adapt(Forest *f, const uint8_t *flags)
{
    std::vector<char> mark(f->leaves.size(), 0);
    for (size_t i = 0; i < f->leaves.size(); ++i) {
        if (flags[i] == 0x10) {
            if (f->leaves[i].level >= MAXL)
                return -1;
            mark[i] = 1;
        } else if (flags[i] != 0) {
            return -1;
        }
    }
    f->refine(mark);
    f->balance();
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

}  // namespace rv
