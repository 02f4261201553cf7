// Host-side grid generation for uniform bricks and the LobattoCell 1-D
// operators, compiled with g++ (uses __float128 / libquadmath).
//
// Restates what Raven's GridManager(cell, brick(coords, periodic); min_level)
// followed by generate(gm) produces on one rank:
//   tree order        p4est brick connectivity (coarsegrids.jl:297), quadrants
//                     in Morton order inside a tree (uniform min_level)
//   points            hexbrickpoints!/quadbrickpoints!/linebrickpoints!
//                     (lobattocells.jl:695-760, 801-872, 1590-1674)
//   volume metrics    *volumebrickmetrics! (lobattocells.jl:2590-2628,
//                     2815-2860, 3182-3251)
//   surface metrics   *surfacemetrics! (lobattocells.jl:2630-2665, 2862-2900,
//                     3253-3295)
//   face maps         materializefacemaps (lobattocells.jl:1992-2295): brick
//                     neighbours are aligned, so node m of face f meets node m
//                     of the opposite face; boundary faces map to themselves
//   numbering         renumberdtocdtop! (lobattocells.jl:1872-1936): continuous
//                     ids in order of first appearance over (cell, node)
#include <quadmath.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace rv {

namespace {

int ceil_log2(int m)
{
    int l = 0;
    while ((1 << l) < m)
        ++l;
    return l;
}

int64_t ipow(int64_t b, int e)
{
    int64_t r = 1;
    while (e-- > 0)
        r *= b;
    return r;
}

// Tree coordinates of linear index `lin` in p4est's brick numbering: the bits
// of the coordinates are interleaved (x lowest) over the levels all active
// dimensions still have; a dimension drops out of the interleaving once its
// bits are used up, the remaining ones continue, and the dimension with the
// most bits takes the leftover high bits.
void brick_tree_coords(int dim, const int *logs, const int *order, int64_t lin, int64_t *t)
{
    for (int d = 0; d < dim; ++d)
        t[d] = 0;
    int done = 0;  // bit levels consumed per active dimension
    for (int stage = 0; stage + 1 < dim; ++stage) {
        const int active = dim - stage;
        const int levels = logs[order[stage]] - done;
        int slot[3] = {-1, -1, -1};
        for (int k = stage; k < dim; ++k)
            slot[order[k]] = 0;
        int next = 0;
        for (int d = 0; d < dim; ++d)
            if (slot[d] == 0)
                slot[d] = next++;
        for (int l = 0; l < levels; ++l)
            for (int d = 0; d < dim; ++d)
                if (slot[d] >= 0)
                    t[d] |= ((lin >> (active * l + slot[d])) & 1) << (done + l);
        done += levels;
        lin >>= active * levels;
    }
    t[order[dim - 1]] += lin << done;
}

__float128 legendre(int n, __float128 x, __float128 *deriv)
{
    __float128 pm = 1, p = x;
    if (n == 0) {
        *deriv = 0;
        return 1;
    }
    for (int k = 2; k <= n; ++k) {
        const __float128 pn = ((2 * k - 1) * x * p - (k - 1) * pm) / k;
        pm = p;
        p = pn;
    }
    *deriv = n * (x * p - pm) / (x * x - 1);
    return p;
}

struct Lattice {
    int dim;
    int64_t N[3];
    std::vector<int64_t> pos_to_cell;
    std::vector<int64_t> g;  // K x dim finest-level coordinates
};

Lattice make_lattice(const BrickSpec &b)
{
    Lattice L;
    L.dim = b.dim;
    const int64_t K = brick_numcells(b);
    std::vector<int32_t> tree(K * b.dim), quad(K * b.dim);
    brick_quadrants(b, tree.data(), quad.data());
    int64_t tot = 1;
    for (int d = 0; d < 3; ++d)
        L.N[d] = 1;
    for (int d = 0; d < b.dim; ++d) {
        L.N[d] = (int64_t)b.n[d] << b.level;
        tot *= L.N[d];
    }
    L.pos_to_cell.assign(tot, -1);
    L.g.resize(K * b.dim);
    for (int64_t e = 0; e < K; ++e) {
        int64_t pos = 0, stride = 1;
        for (int d = 0; d < b.dim; ++d) {
            const int64_t g = ((int64_t)tree[e * b.dim + d] << b.level) + quad[e * b.dim + d];
            L.g[e * b.dim + d] = g;
            pos += g * stride;
            stride *= L.N[d];
        }
        L.pos_to_cell[pos] = e;
    }
    return L;
}

// node coordinate of cell e along d at reference index i (hexbrickpoints!)
struct PointEval {
    const BrickSpec &b;
    std::vector<double> verts[3];
    std::vector<double> r;
    std::vector<int32_t> tree, quad;
    explicit PointEval(const BrickSpec &bs) : b(bs)
    {
        const int64_t K = brick_numcells(b);
        tree.resize(K * b.dim);
        quad.resize(K * b.dim);
        brick_quadrants(b, tree.data(), quad.data());
        for (int d = 0; d < b.dim; ++d) {
            verts[d].resize(b.n[d] + 1);
            brick_vertices(b, d, verts[d].data());
        }
        std::vector<double> w(b.nq), D(b.nq * b.nq);
        r.resize(b.nq);
        lobatto_operators(b.nq, r.data(), w.data(), D.data());
    }
    double operator()(int64_t e, int d, double xi) const
    {
        const int32_t t = tree[e * b.dim + d];
        const double c1 = verts[d][t];
        const double dx = verts[d][t + 1] - verts[d][t];
        const double h = 1.0 / (double)(1 << (b.level + 1));
        const double cr = (double)((int64_t)quad[e * b.dim + d] << (30 - b.level)) /
                          (double)(1 << 30);
        return std::fma(dx, std::fma(h, xi + 1, cr), c1);
    }
};

int64_t face_volume_node(int dim, int nq, int f, int64_t m)
{
    const int d = f / 2;
    const int64_t fixed = (f & 1) ? nq - 1 : 0;
    if (dim == 1)
        return fixed;
    if (dim == 2)
        return d == 0 ? fixed + nq * m : m + nq * fixed;
    const int64_t a = m % nq, c = m / nq;
    if (d == 0)
        return fixed + nq * a + (int64_t)nq * nq * c;
    if (d == 1)
        return a + nq * fixed + (int64_t)nq * nq * c;
    return a + nq * c + (int64_t)nq * nq * fixed;
}

}  // namespace

int64_t brick_numcells(const BrickSpec &b)
{
    int64_t K = 1;
    for (int d = 0; d < b.dim; ++d)
        K *= b.n[d];
    return K << (b.dim * b.level);
}

void brick_quadrants(const BrickSpec &b, int32_t *tree, int32_t *quad)
{
    const int dim = b.dim;
    int logs[3] = {0, 0, 0}, order[3] = {0, 1, 2};
    for (int d = 0; d < dim; ++d)
        logs[d] = ceil_log2(b.n[d]);
    // dimensions ordered by bit count, ties by index
    for (int i = 1; i < dim; ++i)
        for (int j = i; j > 0 && logs[order[j]] < logs[order[j - 1]]; --j) {
            const int tmp = order[j];
            order[j] = order[j - 1];
            order[j - 1] = tmp;
        }
    int64_t span = 1;
    for (int d = 0; d < dim; ++d)
        span <<= logs[d];
    const int64_t nsub = (int64_t)1 << (dim * b.level);
    int64_t e = 0;
    for (int64_t lin = 0; lin < span; ++lin) {
        int64_t t[3] = {0, 0, 0};
        if (dim == 1)
            t[0] = lin;
        else
            brick_tree_coords(dim, logs, order, lin, t);
        bool inside = true;
        for (int d = 0; d < dim; ++d)
            inside = inside && t[d] < b.n[d];
        if (!inside)
            continue;
        for (int64_t m = 0; m < nsub; ++m) {
            for (int d = 0; d < dim; ++d) {
                int32_t c = 0;
                for (int l = 0; l < b.level; ++l)
                    c |= (int32_t)(((m >> (dim * l + d)) & 1) << l);
                tree[e * dim + d] = (int32_t)t[d];
                quad[e * dim + d] = c;
            }
            ++e;
        }
    }
}

void brick_vertices(const BrickSpec &b, int d, double *v)
{
    const int n = b.n[d];
    const __float128 lo = b.lo[d], hi = b.hi[d];
    for (int i = 0; i <= n; ++i)
        v[i] = (i == 0) ? b.lo[d] : (i == n) ? b.hi[d] : (double)((lo * (n - i) + hi * i) / n);
}

// Gauss-Lobatto nodes in quad precision: Newton on P_n'(x) from the
// Chebyshev-Lobatto nodes, then exact symmetry (and an exact zero centre).
void gll_nodes_quad(int nq, std::vector<__float128> &x)
{
    const int n = nq - 1;
    x.assign(nq, 0);
    x[0] = -1;
    x[n] = 1;
    for (int i = 1; i < n; ++i) {
        __float128 xi = -cosq(M_PIq * i / n);
        for (int it = 0; it < 100; ++it) {
            __float128 dp;
            const __float128 p = legendre(n, xi, &dp);
            const __float128 step = dp / ((2 * xi * dp - n * (n + 1) * p) / (1 - xi * xi));
            xi -= step;
            if (fabsq(step) < 1e-33Q)
                break;
        }
        x[i] = xi;
    }
    for (int i = 0; i < nq / 2; ++i)
        x[n - i] = -x[i];
    if (nq % 2)
        x[n / 2] = 0;
}

void lobatto_operators(int nq, double *xo, double *wo, double *Do)
{
    const int n = nq - 1;
    std::vector<__float128> x, pn(nq);
    gll_nodes_quad(nq, x);
    for (int i = 0; i < nq; ++i) {
        __float128 dp;
        pn[i] = legendre(n, x[i], &dp);
        xo[i] = (double)x[i];
        wo[i] = (double)(2 / ((__float128)n * (n + 1) * pn[i] * pn[i]));
    }
    for (int a = 0; a < nq; ++a)
        for (int c = 0; c < nq; ++c) {
            __float128 v = 0;
            if (a != c)
                v = pn[a] / (pn[c] * (x[a] - x[c]));
            else if (a == 0)
                v = -(__float128)n * (n + 1) / 4;
            else if (a == n)
                v = (__float128)n * (n + 1) / 4;
            Do[a + nq * c] = (double)v;
        }
}

void lobatto_halves(int nq, double *lower, double *upper)
{
    std::vector<__float128> x;
    gll_nodes_quad(nq, x);
    // I[a][b] = l_b(y_a), y_a = (x_a -+ 1) / 2 (spectralinterpolation, tohalves)
    for (int h = 0; h < 2; ++h) {
        double *out = h ? upper : lower;
        for (int a = 0; a < nq; ++a) {
            const __float128 y = (x[a] + (h ? 1 : -1)) / 2;
            for (int b = 0; b < nq; ++b) {
                __float128 v = 1;
                for (int c = 0; c < nq; ++c)
                    if (c != b)
                        v *= (y - x[c]) / (x[b] - x[c]);
                out[a * nq + b] = (double)v;
            }
        }
    }
}

void brick_affine(const BrickSpec &b, double *drdx, double *J)
{
    const PointEval X(b);
    const int64_t K = brick_numcells(b);
    // the 3-D brick metrics kernel reads a cell's extents from cell p, the
    // local index of its workgroup's third dimension (Q cells per group,
    // lobattocells.jl:3202); restated so that the arrays match bit for bit
    const int64_t Q = std::max<int64_t>(512 / (b.nq * b.nq), 1);
    for (int64_t e = 0; e < K; ++e) {
        const int64_t src = b.dim == 3 ? e % Q : e;
        double Je = 0;
        for (int d = 0; d < b.dim; ++d) {
            const double ext = X(src, d, 1.0) - X(src, d, -1.0);
            drdx[e * b.dim + d] = 2.0 / ext;
            Je = d == 0 ? ext / 2.0 : Je * (ext / 2.0);
        }
        if (J)
            J[e] = Je;
    }
}

void brick_neighbours(const BrickSpec &b, int64_t *nbr)
{
    const Lattice L = make_lattice(b);
    const int64_t K = brick_numcells(b);
    const int nf = 2 * b.dim;
    for (int64_t e = 0; e < K; ++e)
        for (int f = 0; f < nf; ++f) {
            const int d = f / 2;
            int64_t g[3] = {0, 0, 0};
            for (int c = 0; c < b.dim; ++c)
                g[c] = L.g[e * b.dim + c];
            g[d] += (f & 1) ? 1 : -1;
            int64_t out = -1;
            bool ok = true;
            if (g[d] < 0 || g[d] >= L.N[d]) {
                if (b.periodic[d])
                    g[d] = (g[d] + L.N[d]) % L.N[d];
                else
                    ok = false;
            }
            if (ok) {
                int64_t pos = 0, stride = 1;
                for (int c = 0; c < b.dim; ++c) {
                    pos += g[c] * stride;
                    stride *= L.N[c];
                }
                out = L.pos_to_cell[pos];
            }
            nbr[e * nf + f] = out;
        }
}

void brick_arrays(const BrickSpec &b, double *points, double *metrics, double *surf,
                  int64_t *vmapM, int64_t *vmapP, int64_t *bc)
{
    const int dim = b.dim, nq = b.nq, nf = 2 * dim;
    const int64_t K = brick_numcells(b);
    const int64_t Np = ipow(nq, dim), nfp = ipow(nq, dim - 1), nfpt = nf * nfp;
    const int L = dim * dim + 2;
    std::vector<double> r(nq), w(nq), D(nq * nq);
    lobatto_operators(nq, r.data(), w.data(), D.data());
    if (points) {
        const PointEval X(b);
        for (int64_t e = 0; e < K; ++e)
            for (int64_t node = 0; node < Np; ++node) {
                int64_t rem = node;
                for (int d = 0; d < dim; ++d) {
                    points[node + Np * (d + (int64_t)dim * e)] = X(e, d, r[rem % nq]);
                    rem /= nq;
                }
            }
    }
    std::vector<double> drdx(K * dim), J(K);
    if (metrics || surf)
        brick_affine(b, drdx.data(), J.data());
    if (metrics) {
        for (int64_t e = 0; e < K; ++e)
            for (int64_t node = 0; node < Np; ++node) {
                double *m = metrics + node + Np * (int64_t)L * e;
                int64_t rem = node;
                double wJ = 0;
                for (int d = 0; d < dim; ++d) {
                    const double wd = w[rem % nq];
                    wJ = d == 0 ? wd : wJ * wd;
                    rem /= nq;
                }
                for (int c = 0; c < dim; ++c)
                    for (int a = 0; a < dim; ++a)
                        m[Np * (a + dim * c)] = a == c ? drdx[e * dim + a] : -0.0;
                m[Np * (dim * dim)] = J[e];
                m[Np * (dim * dim + 1)] = wJ * J[e];
            }
    }
    if (surf) {
        for (int64_t e = 0; e < K; ++e)
            for (int f = 0; f < nf; ++f) {
                const int fg = f / 2;
                const double snJ = ((f & 1) ? 1.0 : -1.0) * J[e];
                for (int64_t m = 0; m < nfp; ++m) {
                    double *s = surf + (dim + 2) * (f * nfp + m + nfpt * e);
                    double nv[3] = {0.0, 0.0, 0.0};
                    for (int c = 0; c < dim; ++c)
                        nv[c] = snJ * (c == fg ? drdx[e * dim + fg] : -0.0);
                    double ss = nv[0] * nv[0];
                    for (int c = 1; c < dim; ++c)
                        ss = ss + nv[c] * nv[c];
                    const double sJ = std::sqrt(ss);
                    for (int c = 0; c < dim; ++c)
                        s[c] = nv[c] / sJ;
                    s[dim] = sJ;
                    double wsJ = sJ;
                    if (dim == 2)
                        wsJ = sJ * w[m];
                    else if (dim == 3)
                        wsJ = sJ * w[m % nq] * w[m / nq];
                    s[dim + 1] = wsJ;
                }
            }
    }
    if (vmapM || vmapP || bc) {
        std::vector<int64_t> nbr(K * nf);
        brick_neighbours(b, nbr.data());
        for (int64_t e = 0; e < K; ++e)
            for (int f = 0; f < nf; ++f) {
                const int64_t en = nbr[e * nf + f];
                if (bc)
                    bc[f + nf * e] = en < 0 ? 1 : 0;
                for (int64_t m = 0; m < nfp; ++m) {
                    const int64_t idx = f * nfp + m + nfpt * e;
                    if (vmapM)
                        vmapM[idx] = face_volume_node(dim, nq, f, m) + Np * e + 1;
                    if (vmapP)
                        vmapP[idx] = en < 0 ? face_volume_node(dim, nq, f, m) + Np * e + 1
                                            : face_volume_node(dim, nq, f ^ 1, m) + Np * en + 1;
                }
            }
    }
}

int64_t brick_dtoc(const BrickSpec &b, int64_t *dtoc)
{
    const Lattice L = make_lattice(b);
    const int dim = b.dim, nq = b.nq;
    const int64_t K = brick_numcells(b), Np = ipow(nq, dim);
    int64_t M[3] = {1, 1, 1}, tot = 1;
    for (int d = 0; d < dim; ++d) {
        M[d] = L.N[d] * (nq - 1) + (b.periodic[d] ? 0 : 1);
        tot *= M[d];
    }
    std::vector<int64_t> id(tot, 0);
    int64_t count = 0;
    for (int64_t e = 0; e < K; ++e)
        for (int64_t node = 0; node < Np; ++node) {
            int64_t rem = node, pos = 0, stride = 1;
            for (int d = 0; d < dim; ++d) {
                int64_t x = L.g[e * dim + d] * (nq - 1) + rem % nq;
                rem /= nq;
                if (b.periodic[d])
                    x %= M[d];
                pos += x * stride;
                stride *= M[d];
            }
            if (id[pos] == 0)
                id[pos] = ++count;
            dtoc[node + Np * e] = id[pos];
        }
    return count;
}

}  // namespace rv
