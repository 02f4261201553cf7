/*
 * oracle.c -- CPU restatement of Raven.jl's DGSEM hot path (see oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: the reference the HIP library is tested against,
 * never part of it.
 *
 * Build with -ffp-contract=off: Julia evaluates a*b + c with two roundings
 * unless muladd/fma is written, and its muladd is restated as fma() here.
 * Loops follow the reference's accumulation order so that the FP64 results
 * differ from Raven's only where Raven's own order is set by a dependency
 * (StaticArrays det/norm, AcceleratedKernels sum), which is noted in place.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Gauss-Lobatto-Legendre operators (lobattocells.jl:1-47).                  */
/* Raven computes them in 256-bit BigFloat; __float128 (113 bits) rounds to  */
/* the same doubles except within 2^-113 of a tie (checked against the       */
/* 256-bit fixtures in tests/golden/gll.json).                               */
/* ------------------------------------------------------------------------ */
static void legendre_q(int n, __float128 x, __float128 *p, __float128 *dp)
{
    __float128 p0 = 1, p1 = x;
    if (n == 0) {
        *p = 1;
        *dp = 0;
        return;
    }
    for (int k = 2; k <= n; ++k) {
        __float128 p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    *p = p1;
    *dp = n * (x * p1 - p0) / (x * x - 1);
}

void orc_gll(int nq, double *xo, double *wo, double *Do)
{
    int n = nq - 1;
    __float128 x[64], pn[64];
    x[0] = -1;
    x[n] = 1;
    for (int i = 1; i < n; ++i) {
        __float128 xi = -cosq(M_PIq * i / n);
        for (int it = 0; it < 100; ++it) {
            __float128 p, dp;
            legendre_q(n, xi, &p, &dp);
            __float128 ddp = (2 * xi * dp - n * (n + 1) * p) / (1 - xi * xi);
            __float128 dx = dp / ddp;
            xi -= dx;
            if (fabsq(dx) < 1e-33Q)
                break;
        }
        x[i] = xi;
    }
    for (int i = 0; i < nq / 2; ++i)
        x[n - i] = -x[i];
    if (nq % 2 == 1)
        x[n / 2] = 0;
    for (int i = 0; i < nq; ++i) {
        __float128 dp;
        legendre_q(n, x[i], &pn[i], &dp);
    }
    for (int i = 0; i < nq; ++i) {
        xo[i] = (double)x[i];
        wo[i] = (double)(2 / ((__float128)n * (n + 1) * pn[i] * pn[i]));
    }
    for (int a = 0; a < nq; ++a)
        for (int b = 0; b < nq; ++b) {
            __float128 v = 0;
            if (a != b)
                v = pn[a] / (pn[b] * (x[a] - x[b]));
            else if (a == 0)
                v = -(__float128)n * (n + 1) / 4;
            else if (a == n)
                v = (__float128)n * (n + 1) / 4;
            Do[a + nq * b] = (double)v;
        }
}

/* ------------------------------------------------------------------------ */
/* Brick forest: tree order and quadrant order.                              */
/* ------------------------------------------------------------------------ */
static int ceil_log2(int32_t m)
{
    int l = 0;
    while ((1 << l) < m)
        ++l;
    return l;
}

/* Decode a linear brick index into tree coordinates the way p4est's brick
 * connectivity numbers trees: bits are interleaved over all dimensions
 * (x lowest) for as many levels as the dimension with the fewest bits has,
 * then over the remaining dimensions, and the last dimension's leftover bits
 * are appended.  Indices outside the brick are skipped by the caller. */
static void brick_decode(int dim, const int *logs, const int *rank, int64_t ti, int64_t *t)
{
    int lastlog = 0;
    for (int d = 0; d < dim; ++d)
        t[d] = 0;
    for (int st = 0; st < dim - 1; ++st) {
        int nact = dim - st;
        int logi = logs[rank[st]] - lastlog;
        int id[3] = {-1, -1, -1}, c = 0;
        for (int k = 0; k < nact; ++k)
            id[rank[st + k]] = 0;
        for (int k = 0; k < dim; ++k)
            if (id[k] == 0)
                id[k] = c++;
        for (int j = 0; j < logi; ++j)
            for (int k = 0; k < dim; ++k)
                if (id[k] >= 0)
                    t[k] |= ((ti >> (nact * j + id[k])) & 1) << (lastlog + j);
        lastlog += logi;
        ti >>= nact * logi;
    }
    t[rank[dim - 1]] += ti << lastlog;
}

int64_t orc_brick_numcells(int dim, const int32_t *n, int level)
{
    int64_t K = 1;
    for (int d = 0; d < dim; ++d)
        K *= n[d];
    return K << (dim * level);
}

void orc_brick_quadrants(int dim, const int32_t *n, int level, int32_t *tree_xyz,
                         int32_t *quad_xyz)
{
    int logs[3] = {0, 0, 0}, rank[3] = {0, 1, 2};
    for (int d = 0; d < dim; ++d)
        logs[d] = ceil_log2(n[d]);
    /* stable ascending sort of the dimensions by bit count */
    for (int a = 0; a < dim; ++a)
        for (int b = a + 1; b < dim; ++b)
            if (logs[rank[b]] < logs[rank[a]]) {
                int tmp = rank[a];
                rank[a] = rank[b];
                rank[b] = tmp;
            }
    int64_t total = 1;
    for (int d = 0; d < dim; ++d)
        total <<= logs[d];
    int64_t nsub = (int64_t)1 << (dim * level);
    int64_t e = 0;
    for (int64_t ti = 0; ti < total; ++ti) {
        int64_t t[3];
        if (dim == 1) {
            t[0] = ti; /* 1-D forests order cells along the line */
        } else {
            brick_decode(dim, logs, rank, ti, t);
        }
        int inside = 1;
        for (int d = 0; d < dim; ++d)
            if (t[d] >= n[d])
                inside = 0;
        if (!inside)
            continue;
        for (int64_t m = 0; m < nsub; ++m) {
            int32_t qc[3] = {0, 0, 0};
            for (int l = 0; l < level; ++l)
                for (int d = 0; d < dim; ++d)
                    qc[d] |= (int32_t)(((m >> (dim * l + d)) & 1) << l);
            for (int d = 0; d < dim; ++d) {
                tree_xyz[e * dim + d] = (int32_t)t[d];
                quad_xyz[e * dim + d] = qc[d];
            }
            ++e;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Brick points: linebrickpoints! / quadbrickpoints! / hexbrickpoints!       */
/* r = muladd(h, xi+1, ix/2^30), x = muladd(dx, r, c1).                      */
/* ------------------------------------------------------------------------ */
void orc_brick_points(int dim, int nq, int64_t K, int level, const int32_t *tree_xyz,
                      const int32_t *quad_xyz, const double *const *coords,
                      const double *r, double *points)
{
    int64_t Np = 1;
    for (int d = 0; d < dim; ++d)
        Np *= nq;
    const double root_len = (double)(1 << 30);
    const double h = 1.0 / (double)(1 << (level + 1));
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < K; ++e) {
        double c1[3], dx[3], cr[3];
        for (int d = 0; d < dim; ++d) {
            int32_t t = tree_xyz[e * dim + d];
            c1[d] = coords[d][t];
            dx[d] = coords[d][t + 1] - coords[d][t];
            int64_t ix = (int64_t)quad_xyz[e * dim + d] << (30 - level);
            cr[d] = (double)ix / root_len;
        }
        for (int64_t node = 0; node < Np; ++node) {
            int64_t rem = node;
            for (int d = 0; d < dim; ++d) {
                int id = (int)(rem % nq);
                rem /= nq;
                double rr = fma(h, r[id] + 1, cr[d]);
                points[node + Np * (d + (int64_t)dim * e)] = fma(dx[d], rr, c1[d]);
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Volume metrics.                                                           */
/* ------------------------------------------------------------------------ */
static inline double P(const double *points, int64_t Np, int dim, int64_t e, int64_t node, int d)
{
    return points[node + Np * (d + (int64_t)dim * e)];
}

static void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

void orc_volume_metrics(int dim, int nq, int64_t K, int brick, const double *points,
                        const double *D, const double *w, double *metrics)
{
    const int L = dim * dim + 2;
    int64_t Np = 1;
    for (int d = 0; d < dim; ++d)
        Np *= nq;
#define M(node, f, e) metrics[(node) + Np * ((f) + (int64_t)L * (e))]
    if (brick) {
        /* hexvolumebrickmetrics! reads the element's extents from element p,
         * the local index of the workgroup's third dimension (Q elements per
         * group), instead of element q (lobattocells.jl:3202).  Restated. */
        const int64_t Q = (512 / (nq * nq)) > 1 ? (512 / (nq * nq)) : 1;
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < K; ++e) {
            int64_t src = (dim == 3) ? (e % Q) : e;
            double ext[3] = {1.0, 1.0, 1.0};
            for (int d = 0; d < dim; ++d) {
                int64_t stride = 1;
                for (int dd = 0; dd < d; ++dd)
                    stride *= nq;
                int64_t last = (int64_t)(nq - 1) * stride;
                ext[d] = P(points, Np, dim, src, last, d) - P(points, Np, dim, src, 0, d);
            }
            double drdx[3], J = 1.0;
            for (int d = 0; d < dim; ++d)
                drdx[d] = 2.0 / ext[d];
            /* J = (dx / dr) * (dy / ds) * (dz / dt) */
            J = ext[0] / 2.0;
            for (int d = 1; d < dim; ++d)
                J = J * (ext[d] / 2.0);
            for (int64_t node = 0; node < Np; ++node) {
                int id[3] = {0, 0, 0};
                int64_t rem = node;
                for (int d = 0; d < dim; ++d) {
                    id[d] = (int)(rem % nq);
                    rem /= nq;
                }
                for (int a = 0; a < dim; ++a)
                    for (int b = 0; b < dim; ++b)
                        M(node, a + dim * b, e) = (a == b) ? drdx[a] : -0.0;
                M(node, dim * dim, e) = J;
                double wJ = w[id[0]];
                for (int d = 1; d < dim; ++d)
                    wJ = wJ * w[id[d]];
                M(node, dim * dim + 1, e) = wJ * J;
            }
        }
        return;
    }
    if (dim == 1) {
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < K; ++e)
            for (int i = 0; i < nq; ++i) {
                double dxdr = -0.0;
                for (int m = 0; m < nq; ++m)
                    dxdr += D[i + nq * m] * P(points, Np, 1, e, m, 0);
                double G = dxdr * dxdr;
                double invG = 1.0 / G;
                M(i, 0, e) = invG * dxdr;
                double J = fabs(dxdr);
                M(i, 1, e) = J;
                M(i, 2, e) = w[i] * J;
            }
        return;
    }
    if (dim == 2) {
        /* quadvolumemetrics! (lobattocells.jl:2756-2813) */
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < K; ++e)
            for (int j = 0; j < nq; ++j)
                for (int i = 0; i < nq; ++i) {
                    double xr[2] = {-0.0, -0.0}, xs[2] = {-0.0, -0.0};
                    for (int m = 0; m < nq; ++m)
                        for (int d = 0; d < 2; ++d)
                            xr[d] += D[i + nq * m] * P(points, Np, 2, e, m + nq * j, d);
                    for (int m = 0; m < nq; ++m)
                        for (int d = 0; d < 2; ++d)
                            xs[d] += D[j + nq * m] * P(points, Np, 2, e, i + nq * m, d);
                    double g11 = xr[0] * xr[0] + xr[1] * xr[1];
                    double g12 = xr[0] * xs[0] + xr[1] * xs[1];
                    double g22 = xs[0] * xs[0] + xs[1] * xs[1];
                    /* inv of a 2x2 SMatrix: adjugate / det */
                    double det = g11 * g22 - g12 * g12;
                    double i11 = g22 / det, i12 = -g12 / det, i22 = g11 / det;
                    /* dRdX = invG * [xr xs]' : row a = invG[a,:] * [xr'; xs'] */
                    double dRdX[2][2];
                    for (int b = 0; b < 2; ++b) {
                        dRdX[0][b] = i11 * xr[b] + i12 * xs[b];
                        dRdX[1][b] = i12 * xr[b] + i22 * xs[b];
                    }
                    double J = fabs(xr[0] * xs[1] - xr[1] * xs[0]);
                    int64_t node = i + nq * j;
                    for (int a = 0; a < 2; ++a)
                        for (int b = 0; b < 2; ++b)
                            M(node, a + 2 * b, e) = dRdX[a][b];
                    M(node, 4, e) = J;
                    M(node, 5, e) = w[i] * w[j] * J;
                }
        return;
    }
    /* hexvolumemetrics! (lobattocells.jl:3038-3180), curl-invariant form of
     * Kopriva (2006) eq. (37). Per element, with the reference's loop order. */
#pragma omp parallel
    {
        const int64_t n2 = (int64_t)nq * nq;
        double *X = malloc(sizeof(double) * 3 * Np);
        double *Xr = malloc(sizeof(double) * 3 * Np);
        double *Xs = malloc(sizeof(double) * 3 * Np);
        double *Xt = malloc(sizeof(double) * 3 * Np);
        double *cs = malloc(sizeof(double) * 3 * Np); /* cross(X, Xs) */
        double *ct = malloc(sizeof(double) * 3 * Np); /* cross(X, Xt) */
        double *cr = malloc(sizeof(double) * 3 * Np); /* cross(X, Xr) */
#pragma omp for schedule(static)
        for (int64_t e = 0; e < K; ++e) {
            for (int64_t node = 0; node < Np; ++node)
                for (int d = 0; d < 3; ++d) {
                    X[3 * node + d] = P(points, Np, 3, e, node, d);
                    Xr[3 * node + d] = -0.0;
                    Xs[3 * node + d] = -0.0;
                    Xt[3 * node + d] = -0.0;
                }
            for (int k = 0; k < nq; ++k)
                for (int j = 0; j < nq; ++j)
                    for (int i = 0; i < nq; ++i) {
                        int64_t node = i + nq * j + n2 * k;
                        for (int m = 0; m < nq; ++m)
                            for (int d = 0; d < 3; ++d)
                                Xr[3 * node + d] += D[i + nq * m] * X[3 * (m + nq * j + n2 * k) + d];
                        for (int m = 0; m < nq; ++m)
                            for (int d = 0; d < 3; ++d)
                                Xs[3 * node + d] += D[j + nq * m] * X[3 * (i + nq * m + n2 * k) + d];
                    }
            /* dXdt[o] += Dt[o,k] * X[k], accumulated over k in order */
            for (int j = 0; j < nq; ++j)
                for (int i = 0; i < nq; ++i)
                    for (int k = 0; k < nq; ++k)
                        for (int o = 0; o < nq; ++o)
                            for (int d = 0; d < 3; ++d)
                                Xt[3 * (i + nq * j + n2 * o) + d] +=
                                    D[o + nq * k] * X[3 * (i + nq * j + n2 * k) + d];
            for (int64_t node = 0; node < Np; ++node) {
                cross3(&X[3 * node], &Xs[3 * node], &cs[3 * node]);
                cross3(&X[3 * node], &Xt[3 * node], &ct[3 * node]);
                cross3(&X[3 * node], &Xr[3 * node], &cr[3 * node]);
            }
            for (int k = 0; k < nq; ++k)
                for (int j = 0; j < nq; ++j)
                    for (int i = 0; i < nq; ++i) {
                        int64_t node = i + nq * j + n2 * k;
                        double a1[3] = {-0.0, -0.0, -0.0}, a2[3] = {-0.0, -0.0, -0.0},
                               a3[3] = {-0.0, -0.0, -0.0};
                        for (int m = 0; m < nq; ++m)
                            for (int d = 0; d < 3; ++d)
                                a3[d] += D[i + nq * m] * cs[3 * (m + nq * j + n2 * k) + d];
                        for (int m = 0; m < nq; ++m)
                            for (int d = 0; d < 3; ++d)
                                a2[d] -= D[i + nq * m] * ct[3 * (m + nq * j + n2 * k) + d];
                        for (int m = 0; m < nq; ++m)
                            for (int d = 0; d < 3; ++d)
                                a1[d] += D[j + nq * m] * ct[3 * (i + nq * m + n2 * k) + d];
                        for (int m = 0; m < nq; ++m)
                            for (int d = 0; d < 3; ++d)
                                a3[d] -= D[j + nq * m] * cr[3 * (i + nq * m + n2 * k) + d];
                        for (int o = 0; o < nq; ++o)
                            for (int d = 0; d < 3; ++d)
                                a2[d] += D[k + nq * o] * cr[3 * (i + nq * j + n2 * o) + d];
                        for (int o = 0; o < nq; ++o)
                            for (int d = 0; d < 3; ++d)
                                a1[d] -= D[k + nq * o] * cs[3 * (i + nq * j + n2 * o) + d];
                        /* J = [Xr Xs Xt]; detJ = det(J) (StaticArrays 3x3 det;
                         * its operation order is not pinned bit-for-bit) */
                        const double *c0 = &Xr[3 * node], *c1 = &Xs[3 * node], *c2 = &Xt[3 * node];
                        double detJ = c0[0] * (c1[1] * c2[2] - c2[1] * c1[2]) -
                                      c1[0] * (c0[1] * c2[2] - c2[1] * c0[2]) +
                                      c2[0] * (c0[1] * c1[2] - c1[1] * c0[2]);
                        double twodet = 2 * detJ;
                        /* invJ = [a1 a2 a3]' ./ 2detJ : row r = a_r / 2detJ */
                        const double *arow[3] = {a1, a2, a3};
                        for (int a = 0; a < 3; ++a)
                            for (int b = 0; b < 3; ++b)
                                M(node, a + 3 * b, e) = arow[a][b] / twodet;
                        M(node, 9, e) = detJ;
                        M(node, 10, e) = w[i] * w[j] * w[k] * detJ;
                    }
        }
        free(X);
        free(Xr);
        free(Xs);
        free(Xt);
        free(cs);
        free(ct);
        free(cr);
    }
#undef M
}

/* ------------------------------------------------------------------------ */
/* Face node bookkeeping: face f = 2*(d) + side, node m inside the face runs */
/* over the other dimensions in order, first fastest (facedims, lobattocells */
/* .jl:1980-1990; cellfaceindices :2020-2050).                               */
/* ------------------------------------------------------------------------ */
static int64_t face_volume_node(int dim, int nq, int f, int64_t m)
{
    int d = f / 2, side = f % 2;
    int fixed = side ? nq - 1 : 0;
    if (dim == 1)
        return fixed;
    if (dim == 2) {
        if (d == 0)
            return fixed + (int64_t)nq * m;
        return m + (int64_t)nq * fixed;
    }
    int64_t a = m % nq, b = m / nq, n2 = (int64_t)nq * nq;
    if (d == 0)
        return fixed + nq * a + n2 * b;
    if (d == 1)
        return a + nq * fixed + n2 * b;
    return a + nq * b + n2 * fixed;
}

static int64_t face_nodes(int dim, int nq)
{
    int64_t nfp = 1;
    for (int d = 1; d < dim; ++d)
        nfp *= nq;
    return nfp;
}

/* lattice position of each element (finest-level units) and inverse map */
typedef struct {
    int64_t N[3];
    int64_t *pos2e;
} lattice;

static void lattice_build(int dim, const int32_t *n, int level, int64_t K,
                          const int32_t *tree_xyz, const int32_t *quad_xyz, lattice *lt)
{
    int64_t tot = 1;
    for (int d = 0; d < 3; ++d)
        lt->N[d] = 1;
    for (int d = 0; d < dim; ++d) {
        lt->N[d] = (int64_t)n[d] << level;
        tot *= lt->N[d];
    }
    lt->pos2e = malloc(sizeof(int64_t) * tot);
    for (int64_t e = 0; e < K; ++e) {
        int64_t p = 0, stride = 1;
        for (int d = 0; d < dim; ++d) {
            int64_t g = ((int64_t)tree_xyz[e * dim + d] << level) + quad_xyz[e * dim + d];
            p += g * stride;
            stride *= lt->N[d];
        }
        lt->pos2e[p] = e;
    }
}

void orc_brick_facemaps(int dim, int nq, const int32_t *n, const int32_t *periodic,
                        int level, int64_t K, const int32_t *tree_xyz,
                        const int32_t *quad_xyz, int64_t *vmapM, int64_t *vmapP,
                        int64_t *mapM, int64_t *mapP, int64_t *bc)
{
    const int nfaces = 2 * dim;
    const int64_t nfp = face_nodes(dim, nq);
    const int64_t nfpt = nfaces * nfp;
    int64_t Np = 1;
    for (int d = 0; d < dim; ++d)
        Np *= nq;
    lattice lt;
    lattice_build(dim, n, level, K, tree_xyz, quad_xyz, &lt);
    for (int64_t e = 0; e < K; ++e) {
        int64_t g[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d)
            g[d] = ((int64_t)tree_xyz[e * dim + d] << level) + quad_xyz[e * dim + d];
        for (int f = 0; f < nfaces; ++f) {
            int d = f / 2, side = f % 2;
            int64_t gn[3] = {g[0], g[1], g[2]};
            gn[d] += side ? 1 : -1;
            int has = 1;
            if (gn[d] < 0 || gn[d] >= lt.N[d]) {
                if (periodic[d])
                    gn[d] = (gn[d] + lt.N[d]) % lt.N[d];
                else
                    has = 0;
            }
            int64_t en = -1;
            if (has) {
                int64_t p = 0, stride = 1;
                for (int dd = 0; dd < dim; ++dd) {
                    p += gn[dd] * stride;
                    stride *= lt.N[dd];
                }
                en = lt.pos2e[p];
            }
            bc[f + (int64_t)nfaces * e] = has ? 0 : 1;
            for (int64_t m = 0; m < nfp; ++m) {
                int64_t slot = f * nfp + m;
                int64_t idx = slot + nfpt * e;
                vmapM[idx] = face_volume_node(dim, nq, f, m) + Np * e + 1;
                mapM[idx] = idx + 1;
                /* boundary faces keep the identity map (lobattocells.jl:2072-2074);
                 * conforming brick neighbours are aligned, so node m of face f
                 * meets node m of the opposite face f^1 */
                mapP[idx] = has ? (((int64_t)(f ^ 1) * nfp + m) + nfpt * en + 1) : idx + 1;
            }
        }
    }
    for (int64_t idx = 0; idx < nfpt * K; ++idx) {
        int64_t t = mapP[idx] - 1;
        int64_t et = t / nfpt, st = t % nfpt;
        vmapP[idx] = face_volume_node(dim, nq, (int)(st / nfp), st % nfp) + Np * et + 1;
    }
    free(lt.pos2e);
}

int64_t orc_brick_dtoc(int dim, int nq, const int32_t *n, const int32_t *periodic,
                       int level, int64_t K, const int32_t *tree_xyz,
                       const int32_t *quad_xyz, int64_t *dtoc)
{
    int64_t Np = 1;
    for (int d = 0; d < dim; ++d)
        Np *= nq;
    int64_t M[3] = {1, 1, 1}, tot = 1;
    for (int d = 0; d < dim; ++d) {
        int64_t cells = (int64_t)n[d] << level;
        M[d] = cells * (nq - 1) + (periodic[d] ? 0 : 1);
        tot *= M[d];
    }
    int64_t *id = malloc(sizeof(int64_t) * tot);
    for (int64_t i = 0; i < tot; ++i)
        id[i] = 0;
    int64_t count = 0;
    for (int64_t e = 0; e < K; ++e)
        for (int64_t node = 0; node < Np; ++node) {
            int64_t rem = node, p = 0, stride = 1;
            for (int d = 0; d < dim; ++d) {
                int64_t i = rem % nq;
                rem /= nq;
                int64_t g = ((((int64_t)tree_xyz[e * dim + d] << level) + quad_xyz[e * dim + d]) *
                             (nq - 1)) + i;
                if (periodic[d])
                    g %= M[d];
                p += g * stride;
                stride *= M[d];
            }
            if (id[p] == 0)
                id[p] = ++count;
            dtoc[node + Np * e] = id[p];
        }
    free(id);
    return count;
}

/* ------------------------------------------------------------------------ */
/* Surface metrics.                                                          */
/* ------------------------------------------------------------------------ */
void orc_surface_metrics(int dim, int nq, int64_t K, const double *metrics,
                         const int64_t *vmapM, const double *w, double *surf)
{
    const int L = dim * dim + 2;
    const int nfaces = 2 * dim;
    const int64_t nfp = face_nodes(dim, nq);
    const int64_t nfpt = nfaces * nfp;
    int64_t Np = 1;
    for (int d = 0; d < dim; ++d)
        Np *= nq;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < K; ++e)
        for (int f = 0; f < nfaces; ++f) {
            int fg = f / 2;
            double sn = (f % 2 == 0) ? -1.0 : 1.0;
            for (int64_t m = 0; m < nfp; ++m) {
                int64_t slot = f * nfp + m + nfpt * e;
                int64_t k = vmapM[slot] - 1;
                int64_t ek = k / Np, nk = k % Np;
                const double *mb = metrics + nk + Np * (int64_t)L * ek;
                double J = mb[Np * (dim * dim)];
                double nv[3] = {0.0, 0.0, 0.0};
                double snJ = sn * J;
                for (int c = 0; c < dim; ++c)
                    nv[c] = snJ * mb[Np * (fg + dim * c)]; /* dRdX[fg, c] */
                double ss = nv[0] * nv[0];
                for (int c = 1; c < dim; ++c)
                    ss += nv[c] * nv[c];
                double sJ = sqrt(ss);
                double *out = surf + (dim + 2) * slot;
                for (int c = 0; c < dim; ++c)
                    out[c] = nv[c] / sJ;
                out[dim] = sJ;
                double wsJ = sJ;
                if (dim == 2) {
                    wsJ = sJ * w[m];
                } else if (dim == 3) {
                    wsJ = sJ * w[m % nq] * w[m / nq];
                }
                out[dim + 1] = wsJ;
            }
        }
}

/* ------------------------------------------------------------------------ */
/* GatherScatter.                                                            */
/* ------------------------------------------------------------------------ */
void orc_gather(int64_t np, int nf, int64_t K, int64_t nc, const int64_t *dtoc,
                const double *src, double *dest)
{
    /* ctod = sparse(1:n, dtoc, true): column c holds the discontinuous
     * indices i with dtoc[i] == c in ascending order; a counting sort over i
     * reproduces that order. */
    int64_t n = np * K;
    int64_t *ptr = calloc(nc + 1, sizeof(int64_t));
    int64_t *rows = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i)
        ptr[dtoc[i]] += 1;
    for (int64_t c = 0; c < nc; ++c)
        ptr[c + 1] += ptr[c];
    int64_t *fill = malloc(sizeof(int64_t) * (nc + 1));
    memcpy(fill, ptr, sizeof(int64_t) * (nc + 1));
    for (int64_t i = 0; i < n; ++i) {
        int64_t c = dtoc[i] - 1;
        rows[fill[c]++] = i;
    }
    for (int s = 0; s < nf; ++s)
        for (int64_t c = 0; c < nc; ++c) {
            double v = 0.0;
            for (int64_t r = ptr[c]; r < ptr[c + 1]; ++r) {
                int64_t i = rows[r];
                int64_t e = i / np, node = i % np;
                v = v + src[node + np * (s + (int64_t)nf * e)];
            }
            dest[c + nc * s] = v;
        }
    free(ptr);
    free(rows);
    free(fill);
}

void orc_scatter(int64_t np, int nf, int64_t K, int64_t nc, const int64_t *dtoc,
                 const double *src, double *dest)
{
    for (int64_t e = 0; e < K; ++e)
        for (int s = 0; s < nf; ++s)
            for (int64_t node = 0; node < np; ++node)
                dest[node + np * (s + (int64_t)nf * e)] = src[(dtoc[node + np * e] - 1) + nc * s];
}

/* ------------------------------------------------------------------------ */
/* LSRK coefficients (odesolvers.jl:163-244).  Rationals become fl(num/den) */
/* as Julia's Float64(::Rational{Int64}) does for |num|, den < 2^53.        */
/* ------------------------------------------------------------------------ */
int orc_lsrk_coefficients(int scheme, double *a, double *b, double *c)
{
    if (scheme == 54) {
        const double A[5] = {0.0, -567301805773.0 / 1357537059087.0,
                             -2404267990393.0 / 2016746695238.0,
                             -3550918686646.0 / 2091501179385.0,
                             -1275806237668.0 / 842570457699.0};
        const double B[5] = {1432997174477.0 / 9575080441755.0,
                             5161836677717.0 / 13612068292357.0,
                             1720146321549.0 / 2090206949498.0,
                             3134564353537.0 / 4481467310338.0,
                             2277821191437.0 / 14882151754819.0};
        const double C[5] = {0.0, 1432997174477.0 / 9575080441755.0,
                             2526269341429.0 / 6820363962896.0,
                             2006345519317.0 / 3224310063776.0,
                             2802321613138.0 / 2924317926251.0};
        for (int s = 0; s < 5; ++s) {
            a[s] = A[s];
            b[s] = B[s];
            c[s] = C[s];
        }
        return 5;
    }
    if (scheme == 144) {
        const double A[14] = {0,
                              -0.7188012108672410,
                              -0.7785331173421570,
                              -0.0053282796654044,
                              -0.8552979934029281,
                              -3.9564138245774565,
                              -1.5780575380587385,
                              -2.0837094552574054,
                              -0.7483334182761610,
                              -0.7032861106563359,
                              0.0013917096117681,
                              -0.0932075369637460,
                              -0.9514200470875948,
                              -7.1151571693922548};
        const double B[14] = {0.0367762454319673, 0.3136296607553959, 0.1531848691869027,
                              0.0030097086818182, 0.3326293790646110, 0.2440251405350864,
                              0.3718879239592277, 0.6204126221582444, 0.1524043173028741,
                              0.0760894927419266, 0.0077604214040978, 0.0024647284755382,
                              0.0780348340049386, 5.5059777270269628};
        const double C[14] = {0,
                              0.0367762454319673,
                              0.1249685262725025,
                              0.2446177702277698,
                              0.2476149531070420,
                              0.2969311120382472,
                              0.3978149645802642,
                              0.5270854589440328,
                              0.6981269994175695,
                              0.8190890835352128,
                              0.8527059887098624,
                              0.8604711817462826,
                              0.8627060376969976,
                              0.8734213127600976};
        for (int s = 0; s < 14; ++s) {
            a[s] = A[s];
            b[s] = B[s];
            c[s] = C[s];
        }
        return 14;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* DGSEM operator.                                                           */
/* ------------------------------------------------------------------------ */
struct orc_dg {
    int dim, nq, ns, law, nthreads;
    int volume_form, volume_flux;  /* 0 weak / 1 flux differencing; 2 central / 3 EC */
    int surface_flux;              /* 1 Rusanov (default), 2 central, 3 EC */
    int dir_mask;                  /* directions (bit d: direction d + 1), all by default */
    int64_t K, Np, nfp;
    double gamma, u[3];
    const double *D, *metrics, *surf;
    const int64_t *vmapM, *vmapP, *bc;
    double *MJI;
};

orc_dg *orc_dg_new(int dim, int nq, int64_t K, int law, double gamma, const double *u,
                   const double *D, const double *metrics, const double *surf,
                   const int64_t *vmapM, const int64_t *vmapP, const int64_t *bc)
{
    orc_dg *dg = calloc(1, sizeof(orc_dg));
    dg->dim = dim;
    dg->nq = nq;
    dg->K = K;
    dg->law = law;
    dg->ns = (law == 1) ? 1 : dim + 2;
    dg->gamma = gamma;
    for (int d = 0; d < 3; ++d)
        dg->u[d] = (u && d < dim) ? u[d] : 0.0;
    dg->D = D;
    dg->metrics = metrics;
    dg->surf = surf;
    dg->vmapM = vmapM;
    dg->vmapP = vmapP;
    dg->bc = bc;
    dg->Np = 1;
    for (int d = 0; d < dim; ++d)
        dg->Np *= nq;
    dg->nfp = face_nodes(dim, nq);
    dg->nthreads = 0;
    dg->surface_flux = 1;
    dg->dir_mask = (1 << dim) - 1;
    /* MJI = 1 ./ MJ with MJ = wJ (dgsem.jl:55-56) */
    const int L = dim * dim + 2;
    dg->MJI = malloc(sizeof(double) * dg->Np * K);
    for (int64_t e = 0; e < K; ++e)
        for (int64_t n = 0; n < dg->Np; ++n)
            dg->MJI[n + dg->Np * e] = 1.0 / metrics[n + dg->Np * ((L - 1) + (int64_t)L * e)];
    return dg;
}

void orc_dg_free(orc_dg *dg)
{
    if (!dg)
        return;
    free(dg->MJI);
    free(dg);
}

int orc_dg_ns(const orc_dg *dg) { return dg->ns; }

void orc_dg_set_volume_form(orc_dg *dg, int form, int volume_flux)
{
    dg->volume_form = form;
    dg->volume_flux = volume_flux;
}

void orc_dg_set_surface_flux(orc_dg *dg, int flux) { dg->surface_flux = flux; }

/* DGSEM(...; directions) (dgsem.jl:32-36, 46-81): the volume term keeps the
 * derivatives along these directions (volumeterm! :298-340, the per-direction
 * loop of esvolumeterm_dir! :207-224) and the surface term their faces
 * (surfaceterm! :756) */
void orc_dg_set_directions(orc_dg *dg, int mask) { dg->dir_mask = mask; }
void orc_dg_set_threads(orc_dg *dg, int nthreads) { dg->nthreads = nthreads; }

#ifdef _OPENMP
/* threads of a parallel loop over the operator: the caller's choice, else
 * about one per 32768 state values up to the available ones (small grids run
 * thousands of steps, where thread start-up would dominate); every loop is
 * over independent cells or values, so the results do not depend on it */
static int dg_threads(const orc_dg *dg)
{
    if (dg->nthreads > 0)
        return dg->nthreads;
    const int64_t n = 1 + dg->K * dg->Np * dg->ns / 32768;
    const int mx = omp_get_max_threads();
    return n < mx ? (int)n : mx;
}
#endif

/* flux(law, q) as a dim x Ns matrix F[dd + dim*s]
 * Euler: euler.jl:50-61; advection: advection.jl:18 */
static void law_flux(const orc_dg *dg, const double *q, double *F)
{
    const int dim = dg->dim;
    if (dg->law == 1) {
        for (int dd = 0; dd < dim; ++dd)
            F[dd] = dg->u[dd] * q[0];
        return;
    }
    const double gm1 = dg->gamma - 1;
    double rho = q[0], rhoe = q[dim + 1];
    double uu[3], dot = 0;
    for (int d = 0; d < dim; ++d)
        uu[d] = q[1 + d] / rho;
    dot = q[1] * q[1];
    for (int d = 1; d < dim; ++d)
        dot = dot + q[1 + d] * q[1 + d];
    double p = gm1 * (rhoe - dot / (2 * rho));
    for (int dd = 0; dd < dim; ++dd) {
        F[dd + dim * 0] = q[1 + dd];
        for (int b = 0; b < dim; ++b) {
            double v = q[1 + dd] * uu[b];
            if (dd == b)
                v = v + p;
            F[dd + dim * (1 + b)] = v;
        }
        F[dd + dim * (dim + 1)] = uu[dd] * (rhoe + p);
    }
}

/* wavespeed (euler.jl:63-68, advection.jl:19) */
static double law_wavespeed(const orc_dg *dg, const double *n, const double *q)
{
    const int dim = dg->dim;
    if (dg->law == 1) {
        double s = n[0] * dg->u[0];
        for (int d = 1; d < dim; ++d)
            s = s + n[d] * dg->u[d];
        return fabs(s);
    }
    const double gm1 = dg->gamma - 1;
    double rho = q[0], rhoe = q[dim + 1];
    double uu[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < dim; ++d)
        uu[d] = q[1 + d] / rho;
    double nu = n[0] * uu[0];
    for (int d = 1; d < dim; ++d)
        nu = nu + n[d] * uu[d];
    double dot = q[1] * q[1];
    for (int d = 1; d < dim; ++d)
        dot = dot + q[1 + d] * q[1 + d];
    double p = gm1 * (rhoe - dot / (2 * rho));
    return fabs(nu) + sqrt(dg->gamma * p / rho);
}

/* surfaceflux(::RusanovFlux, ...) (numericalfluxes.jl:27-50) */
static void rusanov(const orc_dg *dg, const double *n, const double *qm, const double *qp,
                    double *nf)
{
    const int dim = dg->dim, ns = dg->ns;
    double Fm[15], Fp[15];
    law_flux(dg, qm, Fm);
    law_flux(dg, qp, Fp);
    double wm = law_wavespeed(dg, n, qm), wp = law_wavespeed(dg, n, qp);
    double lam = wm > wp ? wm : wp;
    for (int s = 0; s < ns; ++s) {
        double fc = ((Fm[0 + dim * s] + Fp[0 + dim * s]) / 2) * n[0];
        for (int d = 1; d < dim; ++d)
            fc = fc + ((Fm[d + dim * s] + Fp[d + dim * s]) / 2) * n[d];
        nf[s] = fc - (lam * (qp[s] - qm[s])) / 2;
    }
}

static void gather_state(const orc_dg *dg, const double *q, int64_t id, double *out)
{
    int64_t e = id / dg->Np, node = id % dg->Np;
    for (int s = 0; s < dg->ns; ++s)
        out[s] = q[node + dg->Np * (s + (int64_t)dg->ns * e)];
}

/* logavg (numericalfluxes.jl:9-22): the logarithmic mean with the series of
 * log(zeta)/(2f) in u = f^2 below eps */
static double logavg(double a, double b)
{
    double zeta = a / b;
    double f = (zeta - 1) / (zeta + 1);
    double u = f * f;
    double F;
    if (u < DBL_EPSILON)
        F = 1.0 + u * (1.0 / 3 + u * (1.0 / 5 + u * (1.0 / 7 + u * (1.0 / 9))));
    else
        F = log(zeta) / (2 * f);
    return (a + b) / (2 * F);
}

/* twopointflux as a dim x Ns matrix F[d + dim*s]:
 * CentralFlux (numericalfluxes.jl:39-42): (F(q1) + F(q2)) / 2;
 * EntropyConservativeFlux for Euler (euler.jl:150-184) */
static void twopoint(const orc_dg *dg, int flux, const double *q1, const double *q2, double *F)
{
    const int dim = dg->dim, ns = dg->ns;
    if (flux == 2 || dg->law == 1) {
        double F1[15], F2[15];
        law_flux(dg, q1, F1);
        law_flux(dg, q2, F2);
        for (int k = 0; k < dim * ns; ++k)
            F[k] = (F1[k] + F2[k]) / 2;
        return;
    }
    const double gm1 = dg->gamma - 1;
    double u1[3], u2[3], uavg[3], fr[3];
    for (int d = 0; d < dim; ++d) {
        u1[d] = q1[1 + d] / q1[0];
        u2[d] = q2[1 + d] / q2[0];
    }
    double dot1 = q1[1] * q1[1], dot2 = q2[1] * q2[1];
    for (int d = 1; d < dim; ++d) {
        dot1 = dot1 + q1[1 + d] * q1[1 + d];
        dot2 = dot2 + q2[1 + d] * q2[1 + d];
    }
    double p1 = gm1 * (q1[dim + 1] - dot1 / (2 * q1[0]));
    double p2 = gm1 * (q2[dim + 1] - dot2 / (2 * q2[0]));
    double b1 = q1[0] / (2 * p1), b2 = q2[0] / (2 * p2);
    double ravg = (q1[0] + q2[0]) / 2;
    for (int d = 0; d < dim; ++d)
        uavg[d] = (u1[d] + u2[d]) / 2;
    double bavg = (b1 + b2) / 2;
    double uu1 = u1[0] * u1[0], uu2 = u2[0] * u2[0];
    for (int d = 1; d < dim; ++d) {
        uu1 = uu1 + u1[d] * u1[d];
        uu2 = uu2 + u2[d] * u2[d];
    }
    double u2avg = (uu1 + uu2) / 2;
    double rlog = logavg(q1[0], q2[0]);
    double blog = logavg(b1, b2);
    for (int d = 0; d < dim; ++d)
        fr[d] = uavg[d] * rlog;
    double pav = ravg / (2 * bavg);
    for (int d = 0; d < dim; ++d) {
        F[d + dim * 0] = fr[d];
        for (int b = 0; b < dim; ++b)
            F[d + dim * (1 + b)] = uavg[d] * fr[b] + (d == b ? pav : 0.0);
    }
    double c = 1 / (2 * gm1 * blog) - u2avg / 2;
    for (int d = 0; d < dim; ++d) {
        double v = F[d + dim * 1] * uavg[0];
        for (int b = 1; b < dim; ++b)
            v = v + F[d + dim * (1 + b)] * uavg[b];
        F[d + dim * (dim + 1)] = c * fr[d] + v;
    }
}

/* FluxDifferencingForm volume term, one direction at a time as
 * esvolumeterm_dir! (dgsem.jl:514-608): for every node and line partner n,
 *   dq -= MJI D[id,n] l_g(ijk) . f  and  dq += f . l_g(n) MJI D[n,id],
 * l_g = MJ dRdX[dir, :]; the first of the operator's directions sets (or
 * increments) dq, the others add to it. */
static void fd_volume(orc_dg *dg, double *dq, const double *q, int increment, int nth)
{
    const int dim = dg->dim, nq = dg->nq, ns = dg->ns;
    const int64_t Np = dg->Np, K = dg->K;
    const int L = dim * dim + 2;
    const double *D = dg->D;
    (void)nth;
    int first = 1;
    for (int dir = 0; dir < dim; ++dir) {
        if (!((dg->dir_mask >> dir) & 1))
            continue;
        const int inc = first ? increment : 1;
        first = 0;
#pragma omp parallel for schedule(static) num_threads(nth)
        for (int64_t e = 0; e < K; ++e) {
            for (int64_t node = 0; node < Np; ++node) {
                int i = (int)(node % nq);
                int j = dim > 1 ? (int)((node / nq) % nq) : 0;
                int k = dim > 2 ? (int)(node / ((int64_t)nq * nq)) : 0;
                const int id = dir == 0 ? i : dir == 1 ? j : k;
                double q1[5], q2[5], F[15], acc[5], lg1[3], lg2[3];
                for (int s = 0; s < ns; ++s) {
                    q1[s] = q[node + Np * (s + (int64_t)ns * e)];
                    acc[s] = -0.0;
                }
                const double MJ1 = dg->metrics[node + Np * ((L - 1) + (int64_t)L * e)];
                for (int d = 0; d < dim; ++d)
                    lg1[d] = MJ1 * dg->metrics[node + Np * ((dir + dim * d) + (int64_t)L * e)];
                const double MJI = dg->MJI[node + Np * e];
                for (int n = 0; n < nq; ++n) {
                    int64_t ild = dir == 0 ? n + nq * (j + (int64_t)nq * k)
                                 : dir == 1 ? i + nq * (n + (int64_t)nq * k)
                                            : i + nq * (j + (int64_t)nq * n);
                    for (int s = 0; s < ns; ++s)
                        q2[s] = q[ild + Np * (s + (int64_t)ns * e)];
                    const double MJ2 = dg->metrics[ild + Np * ((L - 1) + (int64_t)L * e)];
                    for (int d = 0; d < dim; ++d)
                        lg2[d] = MJ2 * dg->metrics[ild + Np * ((dir + dim * d) + (int64_t)L * e)];
                    twopoint(dg, dg->volume_flux, q1, q2, F);
                    for (int s = 0; s < ns; ++s) {
                        double Ddn = MJI * D[id + nq * n];
                        double Dnd = MJI * D[n + nq * id];
                        for (int d = 0; d < dim; ++d) {
                            acc[s] = acc[s] - (Ddn * lg1[d]) * F[d + dim * s];
                            acc[s] = acc[s] + (F[d + dim * s] * lg2[d]) * Dnd;
                        }
                    }
                }
                for (int s = 0; s < ns; ++s) {
                    double *dst = &dq[node + Np * (s + (int64_t)ns * e)];
                    *dst = inc ? (*dst + acc[s]) : acc[s];
                }
            }
        }
    }
}

void orc_dg_rhs(orc_dg *dg, double *dq, const double *q, double t, int increment)
{
    (void)t;
    const int dim = dg->dim, nq = dg->nq, ns = dg->ns;
    const int64_t Np = dg->Np, K = dg->K, nfp = dg->nfp;
    const int L = dim * dim + 2;
    const int nfaces = 2 * dim;
    const int64_t nfpt = nfaces * nfp;
    const double *D = dg->D;
#ifdef _OPENMP
    int nth = dg_threads(dg);
#endif
#ifndef _OPENMP
    int nth = 1;
#endif
    if (dg->volume_form == 1) {
        fd_volume(dg, dq, q, increment, nth);
        goto surface;
    }
    /* volumeterm! (dgsem.jl:257-348) */
#pragma omp parallel num_threads(nth)
    {
        double *lF = malloc(sizeof(double) * dim * ns * Np);
#pragma omp for schedule(static)
        for (int64_t e = 0; e < K; ++e) {
            for (int64_t node = 0; node < Np; ++node) {
                double qv[5], F[15];
                for (int s = 0; s < ns; ++s)
                    qv[s] = q[node + Np * (s + (int64_t)ns * e)];
                law_flux(dg, qv, F);
                double MJ = dg->metrics[node + Np * ((L - 1) + (int64_t)L * e)];
                for (int s = 0; s < ns; ++s)
                    for (int d = 0; d < dim; ++d) {
                        double v = 0;
                        for (int dd = 0; dd < dim; ++dd) {
                            double g = dg->metrics[node + Np * ((d + dim * dd) + (int64_t)L * e)];
                            v += g * F[dd + dim * s];
                        }
                        v *= MJ;
                        lF[node + Np * (d + dim * s)] = v;
                    }
            }
            for (int64_t node = 0; node < Np; ++node) {
                int i = (int)(node % nq);
                int j = dim > 1 ? (int)((node / nq) % nq) : 0;
                int k = dim > 2 ? (int)(node / ((int64_t)nq * nq)) : 0;
                double MJI = dg->MJI[node + Np * e];
                double acc[5];
                for (int s = 0; s < ns; ++s)
                    acc[s] = -0.0;
                const int m = dg->dir_mask;
                for (int n = 0; n < nq; ++n) {
                    double Dni = D[n + nq * i] * MJI;
                    double Dnj = D[n + nq * j] * MJI;
                    double Dnk = D[n + nq * k] * MJI;
                    for (int s = 0; s < ns; ++s) {
                        if (m & 1)
                            acc[s] += Dni * lF[(n + nq * j + (int64_t)nq * nq * k) + Np * (0 + dim * s)];
                        if (dim > 1 && (m & 2))
                            acc[s] += Dnj * lF[(i + nq * n + (int64_t)nq * nq * k) + Np * (1 + dim * s)];
                        if (dim > 2 && (m & 4))
                            acc[s] += Dnk * lF[(i + nq * j + (int64_t)nq * nq * n) + Np * (2 + dim * s)];
                    }
                }
                for (int s = 0; s < ns; ++s) {
                    double *dst = &dq[node + Np * (s + (int64_t)ns * e)];
                    *dst = increment ? (*dst + acc[s]) : acc[s];
                }
            }
        }
        free(lF);
    }
surface:;
    /* surfaceterm! (dgsem.jl:733-784): each element updates only its own
     * nodes, faces in the order d = 1..dim, f = 1, 2 */
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int64_t e = 0; e < K; ++e)
        for (int face = 0; face < nfaces; ++face)
            for (int64_t m = 0; m < nfp; ++m) {
                if (!((dg->dir_mask >> (face / 2)) & 1))
                    continue;
                int64_t slot = face * nfp + m + nfpt * e;
                int64_t idm = dg->vmapM[slot] - 1;
                const double *nv = dg->surf + (dim + 2) * slot;
                double fMJ = nv[dim + 1];
                double qm[5], qp[5], nf[5];
                gather_state(dg, q, idm, qm);
                if (dg->bc[face + (int64_t)nfaces * e] == 0) {
                    gather_state(dg, q, dg->vmapP[slot] - 1, qp);
                } else {
                    /* boundarystate defaults to q⁻ (balancelaw.jl:22) */
                    for (int s = 0; s < ns; ++s)
                        qp[s] = qm[s];
                }
                if (dg->surface_flux == 1) {
                    rusanov(dg, nv, qm, qp, nf);
                } else {
                    /* surfaceflux default: twopointflux(q-, q+)' n (numericalfluxes.jl:27-37) */
                    double F[15];
                    twopoint(dg, dg->surface_flux, qm, qp, F);
                    for (int s = 0; s < ns; ++s) {
                        double v = F[0 + dim * s] * nv[0];
                        for (int d = 1; d < dim; ++d)
                            v = v + F[d + dim * s] * nv[d];
                        nf[s] = v;
                    }
                }
                int64_t ee = idm / Np, nn = idm % Np;
                double MJI = dg->MJI[idm];
                for (int s = 0; s < ns; ++s) {
                    double *dst = &dq[nn + Np * (s + (int64_t)ns * ee)];
                    *dst = *dst - (fMJ * nf[s]) * MJI;
                }
            }
}

/* Mortar coupling of hanging faces: restatement of the library's opt-in
 * extension (Raven itself leaves these faces self-mapped and has no mortar in
 * its right-hand side, lobattocells.jl:2072-2074).  Adds to dq, for every cell
 * and face in order: on a hanging face (kind 1) -MJI wsJ_f nf_s, with nf_s the
 * Rusanov flux between the fine state and the coarse trace interpolated with
 * tohalves_1d (lower/upper, row-major [a][b]); on a covered face (kind 2) the
 * transpose projection of the fine faces' wsJ_f nf_s onto the coarse nodes,
 * times MJI.  `surf_full` holds the surface metrics with the lift weights of
 * these faces (the operator's own `surf` has them zeroed).  partner as
 * rv_forest_mortars: (4, 2dim, K). */
void orc_dg_mortar_add(const orc_dg *dg, const int32_t *kind, const int64_t *partner,
                       const double *lower, const double *upper, const double *surf_full,
                       const double *q, double *dq)
{
    const int dim = dg->dim, nq = dg->nq, ns = dg->ns, nf = 2 * dim;
    const int64_t Np = dg->Np, nfp = dg->nfp, nfpt = nf * nfp;
    const int nsub = 1 << (dim - 1);
    const double *H[2] = {lower, upper};
    double tr[5 * 64], g[4 * 5 * 64];
    for (int64_t e = 0; e < dg->K; ++e)
        for (int f = 0; f < nf; ++f) {
            const int k = kind[f + nf * e];
            if (k == 0)
                continue;
            const int64_t *pt = partner + 4 * (f + nf * e);
            /* the trace to interpolate: coarse neighbour's face f^1 or own face f */
            const int64_t src = k == 1 ? pt[0] : e;
            const int fsrc = k == 1 ? (f ^ 1) : f;
            for (int64_t m = 0; m < nfp; ++m) {
                int64_t node = face_volume_node(dim, nq, fsrc, m);
                for (int s = 0; s < ns; ++s)
                    tr[s * nfp + m] = q[node + Np * (s + (int64_t)ns * src)];
            }
            const int first = k == 1 ? (int)pt[1] : 0, last = k == 1 ? (int)pt[1] : nsub - 1;
            for (int sub = first; sub <= last; ++sub) {
                const int s1 = sub & 1, s2 = (sub >> 1) & 1;
                const int64_t fine = k == 1 ? e : pt[sub];
                const int ffine = k == 1 ? f : (f ^ 1);
                for (int64_t m = 0; m < nfp; ++m) {
                    const int a = (int)(m % nq), b = (int)(m / nq);
                    double qi[5], qf[5], nv[3], wsJ, nflx[5];
                    for (int s = 0; s < ns; ++s) {
                        double v = 0.0;
                        if (dim == 2) {
                            for (int i = 0; i < nq; ++i)
                                v = v + H[s1][a * nq + i] * tr[s * nfp + i];
                        } else {
                            for (int jj = 0; jj < nq; ++jj) {
                                double inner = 0.0;
                                for (int i = 0; i < nq; ++i)
                                    inner = inner + H[s1][a * nq + i] * tr[s * nfp + i + nq * jj];
                                v = v + H[s2][b * nq + jj] * inner;
                            }
                        }
                        qi[s] = v;
                    }
                    const int64_t fnode = face_volume_node(dim, nq, ffine, m);
                    for (int s = 0; s < ns; ++s)
                        qf[s] = q[fnode + Np * (s + (int64_t)ns * fine)];
                    const double *sg = surf_full + (dim + 2) * (ffine * nfp + m + nfpt * fine);
                    for (int c = 0; c < dim; ++c)
                        nv[c] = sg[c];
                    wsJ = sg[dim + 1];
                    rusanov(dg, nv, qf, qi, nflx);
                    for (int s = 0; s < ns; ++s) {
                        const double w = wsJ * nflx[s];
                        if (k == 1) {
                            const double mji = dg->MJI[fnode + Np * e];
                            dq[fnode + Np * (s + (int64_t)ns * e)] += -(mji * w);
                        } else {
                            g[(sub * ns + s) * nfp + m] = w;
                        }
                    }
                }
            }
            if (k == 2)
                for (int64_t m = 0; m < nfp; ++m) {
                    const int i = (int)(m % nq), j = (int)(m / nq);
                    const int64_t node = face_volume_node(dim, nq, f, m);
                    for (int s = 0; s < ns; ++s) {
                        double v = 0.0;
                        for (int sub = 0; sub < nsub; ++sub) {
                            const int s1 = sub & 1, s2 = (sub >> 1) & 1;
                            if (dim == 2) {
                                for (int a = 0; a < nq; ++a)
                                    v = v + H[s1][a * nq + i] * g[(sub * ns + s) * nfp + a];
                            } else {
                                for (int b = 0; b < nq; ++b) {
                                    double inner = 0.0;
                                    for (int a = 0; a < nq; ++a)
                                        inner = inner +
                                                H[s1][a * nq + i] * g[(sub * ns + s) * nfp + a + nq * b];
                                    v = v + H[s2][b * nq + j] * inner;
                                }
                            }
                        }
                        dq[node + Np * (s + (int64_t)ns * e)] += dg->MJI[node + Np * e] * v;
                    }
                }
        }
}

void orc_dg_lsrk_step(orc_dg *dg, int scheme, double *q, double *dq, double time,
                      double dt)
{
    double a[14], b[14], c[14];
    int nst = orc_lsrk_coefficients(scheme, a, b, c);
    const int64_t n = dg->Np * dg->ns * dg->K;
#ifdef _OPENMP
    int nth = dg_threads(dg);
#endif
    for (int s = 0; s < nst; ++s) {
        double stagetime = time + c[s] * dt;
        double as = a[s];
#pragma omp parallel for schedule(static) num_threads(nth)
        for (int64_t i = 0; i < n; ++i)
            dq[i] = dq[i] * as;
        orc_dg_rhs(dg, dq, q, stagetime, 1);
        double bdt = b[s] * dt;
#pragma omp parallel for schedule(static) num_threads(nth)
        for (int64_t i = 0; i < n; ++i)
            q[i] = q[i] + bdt * dq[i];
    }
}

int64_t orc_dg_solve(orc_dg *dg, int scheme, double *q, double dt, double t0,
                     double timeend, double *t_out)
{
    const int64_t n = dg->Np * dg->ns * dg->K;
    double *dq = calloc(n, sizeof(double));
    double time = t0;
    int64_t step = 0;
    for (;;) {
        ++step;
        int finalstep = 0;
        if (time + dt >= timeend) {
            dt = timeend - time;
            finalstep = 1;
        }
        orc_dg_lsrk_step(dg, scheme, q, dq, time, dt);
        time += dt;
        if (finalstep)
            break;
    }
    free(dq);
    if (t_out)
        *t_out = time;
    return step;
}

/* entropy and entropy variables (euler.jl:70-91, advection.jl:20-21) */
static double law_entropy(const orc_dg *dg, const double *q)
{
    const int dim = dg->dim;
    if (dg->law == 1)
        return q[0] * q[0] / 2;
    const double g = dg->gamma;
    double dot = q[1] * q[1];
    for (int d = 1; d < dim; ++d)
        dot = dot + q[1 + d] * q[1 + d];
    double p = (g - 1) * (q[dim + 1] - dot / (2 * q[0]));
    double s = log(p / pow(q[0], g));
    return -q[0] * s / (g - 1);
}

static void law_entropy_variables(const orc_dg *dg, const double *q, double *v)
{
    const int dim = dg->dim;
    if (dg->law == 1) {
        v[0] = q[0];
        return;
    }
    const double g = dg->gamma;
    double dot = q[1] * q[1];
    for (int d = 1; d < dim; ++d)
        dot = dot + q[1 + d] * q[1 + d];
    double p = (g - 1) * (q[dim + 1] - dot / (2 * q[0]));
    double s = log(p / pow(q[0], g));
    double b = q[0] / (2 * p);
    double u[3] = {0, 0, 0}, uu;
    for (int d = 0; d < dim; ++d)
        u[d] = q[1 + d] / q[0];
    uu = u[0] * u[0];
    for (int d = 1; d < dim; ++d)
        uu = uu + u[d] * u[d];
    v[0] = (g - s) / (g - 1) - uu * b;
    for (int d = 0; d < dim; ++d)
        v[1 + d] = 2 * b * u[d];
    v[dim + 1] = -2 * b;
}

/* entropyintegral(dg, q) = sum MJ eta(q) (measures.jl:23-27) */
double orc_dg_entropy_integral(const orc_dg *dg, const double *q)
{
    const int dim = dg->dim, ns = dg->ns, L = dim * dim + 2;
    const int64_t Np = dg->Np;
    double sum = 0;
#ifdef _OPENMP
    const int nth = dg_threads(dg);
#endif
#pragma omp parallel for schedule(static) reduction(+ : sum) num_threads(nth)
    for (int64_t e = 0; e < dg->K; ++e)
        for (int64_t n = 0; n < Np; ++n) {
            double qv[5];
            for (int s = 0; s < ns; ++s)
                qv[s] = q[n + Np * (s + (int64_t)ns * e)];
            sum += dg->metrics[n + Np * ((L - 1) + (int64_t)L * e)] * law_entropy(dg, qv);
        }
    return sum;
}

/* entropyproduct(dg, p, q) = sum MJ v(p) . q (measures.jl:28-33) */
double orc_dg_entropy_product(const orc_dg *dg, const double *p, const double *q)
{
    const int dim = dg->dim, ns = dg->ns, L = dim * dim + 2;
    const int64_t Np = dg->Np;
    double sum = 0;
#ifdef _OPENMP
    const int nth = dg_threads(dg);
#endif
#pragma omp parallel for schedule(static) reduction(+ : sum) num_threads(nth)
    for (int64_t e = 0; e < dg->K; ++e)
        for (int64_t n = 0; n < Np; ++n) {
            double pv[5], v[5], dot;
            for (int s = 0; s < ns; ++s)
                pv[s] = p[n + Np * (s + (int64_t)ns * e)];
            law_entropy_variables(dg, pv, v);
            dot = v[0] * q[n + Np * (0 + (int64_t)ns * e)];
            for (int s = 1; s < ns; ++s)
                dot = dot + v[s] * q[n + Np * (s + (int64_t)ns * e)];
            sum += dg->metrics[n + Np * ((L - 1) + (int64_t)L * e)] * dot;
        }
    return sum;
}

/* min_node_distance(grid; dims) (grids.jl:94-145): per node, the distance to
 * the nodes at linear index ijk -+ stride(d) inside [0, Np) for every d in the
 * mask (min_neighbour_distance_kernel, :94-119), minimised over all nodes */
double orc_min_node_distance(int dim, int nq, int64_t K, const double *points, int mask)
{
    const int np = dim == 3 ? nq * nq * nq : dim == 2 ? nq * nq : nq;
    double md = INFINITY;
    for (int64_t e = 0; e < K; ++e)
        for (int ijk = 0; ijk < np; ++ijk) {
            const double *x = points + (int64_t)dim * np * e;
            int stride = 1;
            for (int d = 0; d < dim; ++d, stride *= nq) {
                if (!((mask >> d) & 1))
                    continue;
                for (int m = -1; m <= 1; m += 2) {
                    const int nb = ijk + stride * m;
                    if (nb < 0 || nb >= np)
                        continue;
                    double ss = 0;
                    for (int c = 0; c < dim; ++c) {
                        const double dx = x[ijk + np * c] - x[nb + np * c];
                        ss += dx * dx;
                    }
                    const double dist = sqrt(ss);
                    if (dist < md)
                        md = dist;
                }
            }
        }
    return md;
}

int orc_rlsrk_weights(int scheme, double *bfull);

/* One relaxation LSRK step, dostep!(q, ::RLSRK) (odesolvers.jl:117-151): the
 * LSRK stages with the entropy production dEta = sum_s bfull_s dt v(q).k, then
 * gamma solving eta(q0 + gamma k) - eta0 - gamma dEta = 0 (k = q - q0) from the
 * previous gamma.  Raven calls NLsolve (trust region, ftol = 10 eps); this is
 * Newton's method on the same residual and derivative r'(gamma) = v(q).k - dEta,
 * stopped at the same |residual| <= 10 eps, or when the residual no longer
 * halves (its rounding floor, which for a large integral sits above the
 * absolute 10 eps; the better of the last two roots is kept), or when the
 * Newton update is below 2 eps |gamma|, at most 50 iterations.  bfull_s are the
 * standard RK weights of the low-storage scheme (odesolvers.jl:96-101).
 * Returns the number of Newton iterations, -1 when the update is not finite. */
int orc_dg_rlsrk_step(orc_dg *dg, int scheme, double *q, double *dq, double *k, double *q0,
                      double time, double dt, double *gamma)
{
    double A[14], B[14], C[14];
    const int nst = orc_lsrk_coefficients(scheme, A, B, C);
    const int64_t n = dg->Np * dg->ns * dg->K;
    double bfull[14];
    orc_rlsrk_weights(scheme, bfull);
    memcpy(q0, q, sizeof(double) * n);
    const double eta0 = orc_dg_entropy_integral(dg, q0);
    double deta = -0.0;
    for (int s = 0; s < nst; ++s) {
        for (int64_t i = 0; i < n; ++i)
            dq[i] *= A[s];
        orc_dg_rhs(dg, k, q, time + C[s] * dt, 0);
        for (int64_t i = 0; i < n; ++i)
            dq[i] += k[i];
        deta += bfull[s] * dt * orc_dg_entropy_product(dg, q, k);
        for (int64_t i = 0; i < n; ++i)
            q[i] += B[s] * dt * dq[i];
    }
    for (int64_t i = 0; i < n; ++i)
        k[i] = q[i] - q0[i];
    double g = *gamma, gprev = g, Fprev = INFINITY;
    int it = 0;
    for (; it < 50; ++it) {
        for (int64_t i = 0; i < n; ++i)
            q[i] = q0[i] + g * k[i];
        const double F = orc_dg_entropy_integral(dg, q) - eta0 - g * deta;
        if (fabs(F) <= 10 * DBL_EPSILON)
            break;
        if (fabs(F) > 0.5 * fabs(Fprev)) {
            /* no longer contracting: the residual is at its rounding floor;
             * keep the better of the last two roots */
            if (fabs(F) > fabs(Fprev))
                g = gprev;
            break;
        }
        const double step = F / (orc_dg_entropy_product(dg, q, k) - deta);
        if (!isfinite(step))
            return -1;
        if (!(fabs(step) > 2 * DBL_EPSILON * fabs(g)))
            break;
        gprev = g;
        Fprev = F;
        g = g - step;
    }
    *gamma = g;
    for (int64_t i = 0; i < n; ++i)
        q[i] = q0[i] + g * k[i];
    return it;
}

/* bfull of dostep!(q, ::RLSRK) (odesolvers.jl:96-101): the Butcher weights
 * b_i = rka_{i+1} b_{i+1} + rkb_i, b_last = rkb_last */
int orc_rlsrk_weights(int scheme, double *bfull)
{
    double A[14], B[14], C[14];
    const int nst = orc_lsrk_coefficients(scheme, A, B, C);
    bfull[nst - 1] = B[nst - 1];
    for (int i = nst - 2; i >= 0; --i)
        bfull[i] = A[i + 1] * bfull[i + 1] + B[i];
    return nst;
}

void orc_dg_weightednorm(const orc_dg *dg, const double *q, double *comp)
{
    const int L = dg->dim * dg->dim + 2;
    for (int s = 0; s < dg->ns; ++s) {
        /* AcceleratedKernels' summation order is not restated; a plain sum */
        double acc = 0;
        for (int64_t e = 0; e < dg->K; ++e)
            for (int64_t node = 0; node < dg->Np; ++node) {
                double MJ = dg->metrics[node + dg->Np * ((L - 1) + (int64_t)L * e)];
                double v = fabs(q[node + dg->Np * (s + (int64_t)dg->ns * e)]);
                acc += MJ * (v * v);
            }
        comp[s] = sqrt(acc);
    }
}
