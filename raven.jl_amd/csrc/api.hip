// C ABI of the MI355X DGSEM library (include/ravenhip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ravenhip.h"
#include "common.hpp"

namespace rv {
StageLauncher find_d1_f64_adv(const KernelKey &);
StageLauncher find_d1_f64_euler(const KernelKey &);
StageLauncher find_d1_f32_adv(const KernelKey &);
StageLauncher find_d1_f32_euler(const KernelKey &);
StageLauncher find_d2_f64_adv(const KernelKey &);
StageLauncher find_d2_f64_euler(const KernelKey &);
StageLauncher find_d2_f32_adv(const KernelKey &);
StageLauncher find_d2_f32_euler(const KernelKey &);
StageLauncher find_d3_f64_adv(const KernelKey &);
StageLauncher find_d3_f64_euler(const KernelKey &);
StageLauncher find_d3_f32_adv(const KernelKey &);
StageLauncher find_d3_f32_euler(const KernelKey &);

StageLauncher find_fd_d1_f64(const KernelKey &);
StageLauncher find_fd_d1_f32(const KernelKey &);
StageLauncher find_fd_d2_f64(const KernelKey &);
StageLauncher find_fd_d2_f32(const KernelKey &);
StageLauncher find_fd_d3_f64(const KernelKey &);
StageLauncher find_fd_d3_f32(const KernelKey &);

StageLauncher find_stage_launcher(const KernelKey &k)
{
    using Finder = StageLauncher (*)(const KernelKey &);
    if (k.vflux != 0) {
        // flux differencing form (fluxdiff.hip)
        static const Finder fd[] = {find_fd_d1_f64, find_fd_d1_f32, find_fd_d2_f64,
                                    find_fd_d2_f32, find_fd_d3_f64, find_fd_d3_f32};
        for (Finder f : fd)
            if (StageLauncher l = f(k))
                return l;
        return nullptr;
    }
    static const Finder finders[] = {
        find_d1_f64_adv, find_d1_f64_euler, find_d1_f32_adv, find_d1_f32_euler,
        find_d2_f64_adv, find_d2_f64_euler, find_d2_f32_adv, find_d2_f32_euler,
        find_d3_f64_adv, find_d3_f64_euler, find_d3_f32_adv, find_d3_f32_euler};
    for (Finder f : finders)
        if (StageLauncher l = f(k))
            return l;
    return nullptr;
}
}  // namespace rv

using namespace rv;

struct rv_handle {
    std::string err;
    int dim = 0, nq = 0, ns = 0, law = 0, flux = 0, dtype = 0, geo = 0, dir_mask = 0;
    int vflux = 0;  // two-point flux of the flux-differencing volume form, 0: weak form
    int64_t K = 0, Np = 0, nfp = 0;
    double gamma = 1.4, u[3] = {0, 0, 0};
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int wpb = 1;      // waves per workgroup of a fused stage
    int wpb_rhs = 1;  // ... of a right-hand side alone (no q_out written)
    StageLauncher launcher = nullptr;
    StageLauncher relax_launcher = nullptr;  // fused relaxation stage, when built
    // device buffers
    FaceRec *faces = nullptr;
    int64_t *explicit_idx = nullptr;
    double *affine = nullptr, *vgeo = nullptr, *mji = nullptr, *sgeo = nullptr;
    double *Dt = nullptr, *Dw = nullptr, *winv = nullptr;
    double *MJ = nullptr;
    double *cellJ = nullptr;     // brick handles: per-cell J that MJ is filled from
    double *partials = nullptr;  // reduction scratch
    double *relax_partials = nullptr;  // entropy passes: RELAX_SLOTS slots of 4*RELAX_BLOCKS
    double *relax_sums = nullptr;       // device sums of the partials (relax_read)
    double *relax_sums_host = nullptr;  // pinned copy of them
    double *eprod = nullptr;           // fused relaxation stages: 14 slots of relax_partials_cap
    void *qtmp = nullptr;        // ping-pong state of rv_lsrk_run
    int64_t *dtoc = nullptr;     // 0-based
    int64_t *ctod_ptr = nullptr;
    int32_t *ctod_rows = nullptr;  // (cell, node) pairs
    int64_t nc = 0;
    int64_t device_bytes = 0;
    int faces_self = 0, faces_explicit = 0;
    bool is_brick = false;
    BrickSpec brick{};
    Forest *forest = nullptr;  // forest handles: the leaves, for the continuous numbering
    // mortar coupling of hanging faces (rv_enable_mortars)
    MortarLauncher mortar = nullptr;
    int64_t mortar_cells = 0;
    int64_t *m_cells = nullptr, *m_partner = nullptr;
    int64_t mortar_pairs = 0;        // non-conforming (cell, face) pairs,
    MortarPair *m_pairs = nullptr;  // covered sides first
    int32_t *m_kind = nullptr, *m_slot = nullptr;
    double *m_sgeo = nullptr, *m_halves = nullptr;
    // FP32 handles with 3-D general cells: vgeo, mji, sgeo rounded to FP32
    float *vgeo32 = nullptr, *mji32 = nullptr, *sgeo32 = nullptr, *mj32 = nullptr;
    unsigned char *m_delta = nullptr;  // per (mortar cell, face) updates of one stage
};

namespace {

std::string g_last_error;

int fail(rv_handle *h, int code, const std::string &msg)
{
    if (h)
        h->err = msg;
    g_last_error = msg;
    return code;
}

#define HIP_TRY(h, expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail((h), RV_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_));    \
    } while (0)

template <typename T>
int dalloc(rv_handle *h, T **p, size_t count)
{
    *p = nullptr;
    if (count == 0)
        count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(h, RV_E_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                       " bytes failed: " + hipGetErrorString(e));
    }
    h->device_bytes += (int64_t)(count * sizeof(T));
    return RV_OK;
}

template <typename T>
int upload(rv_handle *h, T **p, const T *src, size_t count)
{
    int rc = dalloc(h, p, count);
    if (rc)
        return rc;
    if (count)
        HIP_TRY(h, hipMemcpy(*p, src, count * sizeof(T), hipMemcpyHostToDevice));
    return RV_OK;
}

int64_t ipow(int64_t b, int e)
{
    int64_t r = 1;
    while (e-- > 0)
        r *= b;
    return r;
}

__host__ __device__ int64_t face_node_rt(int dim, int nq, int f, int64_t m)
{
    const int d = f >> 1;
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

__host__ __device__ int64_t orient_rt(int dim, int nq, int o, int64_t m)
{
    if (dim == 1)
        return 0;
    if (dim == 2)
        return (o & 1) ? nq - 1 - m : m;
    int64_t a = m % nq, c = m / nq;
    if (o & 4) {
        const int64_t t = a;
        a = c;
        c = t;
    }
    if (o & 1)
        a = nq - 1 - a;
    if (o & 2)
        c = nq - 1 - c;
    return a + nq * c;
}

void free_handle(rv_handle *h)
{
    if (!h)
        return;
    void *bufs[] = {h->faces,  h->explicit_idx, h->affine,   h->vgeo,     h->mji,
                    h->sgeo,   h->Dt,           h->Dw,       h->winv,     h->MJ,
                    h->partials, h->qtmp,       h->dtoc,     h->ctod_ptr, h->ctod_rows,
                    h->m_cells, h->m_partner,   h->m_kind,   h->m_sgeo,   h->m_halves,
                    h->m_delta, h->m_slot, h->cellJ,  h->relax_partials, h->eprod,
                    h->relax_sums, h->m_pairs, h->vgeo32, h->mji32, h->sgeo32, h->mj32};
    for (void *b : bufs)
        if (b)
            (void)hipFree(b);
    if (h->relax_sums_host)
        (void)hipHostFree(h->relax_sums_host);
    if (h->own_stream && h->stream)
        (void)hipStreamDestroy(h->stream);
    if (h->forest)
        forest_free(h->forest);
    delete h;
}

// --------------------------------------------------------------------------
// Auxiliary kernels.
// --------------------------------------------------------------------------

// Neumaier two-sum accumulation, fixed order per thread and fixed tree.
struct Comp {
    double s, c;
};

__device__ inline Comp comp_add(Comp a, double x)
{
    const double t = a.s + x;
    if (fabs(a.s) >= fabs(x))
        a.c += (a.s - t) + x;
    else
        a.c += (x - t) + a.s;
    a.s = t;
    return a;
}

__device__ inline Comp comp_merge(Comp a, Comp b)
{
    Comp r = comp_add(a, b.s);
    r.c += b.c;
    return r;
}

constexpr int RED_BLOCKS = 512;
constexpr int RED_THREADS = 256;

// max that returns NaN when either argument is NaN, as Julia's maximum does
// (fmax would drop it and hide a diverged state)
__host__ __device__ inline double max_nan(double a, double b)
{
    return a != a ? a : b != b ? b : (a > b ? a : b);
}

// partial[s * RED_BLOCKS + block] = compensated sum over the block's share of
// MJ * f(q_s); f = square (power 2), identity (power 1) or |q|^p (power 3);
// power 4: the block's max |q_s| (in the sum slot, compensation 0)
template <typename R>
__global__ void __launch_bounds__(RED_THREADS)
    reduce_kernel(const R *q, const double *MJ, int64_t Np, int ns, int64_t K, int power,
                  double pexp, double *partial)
{
    __shared__ Comp sh[RED_THREADS];
    const int64_t n = Np * K;
    const int s = blockIdx.y;
    Comp acc{0.0, 0.0};
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * chunk;
    const int64_t hi = (lo + chunk < n) ? lo + chunk : n;
    for (int64_t i = lo + threadIdx.x; i < hi; i += RED_THREADS) {
        const int64_t e = i / Np, node = i - e * Np;
        const double v = (double)q[node + Np * (s + (int64_t)ns * e)];
        if (power == 4) {
            acc.s = max_nan(acc.s, fabs(v));
            continue;
        }
        const double x = power == 2   ? MJ[i] * (v * v)
                         : power == 3 ? MJ[i] * pow(fabs(v), pexp)
                                      : MJ[i] * v;
        acc = comp_add(acc, x);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            if (power == 4)
                sh[threadIdx.x].s = max_nan(sh[threadIdx.x].s, sh[threadIdx.x + w].s);
            else
                sh[threadIdx.x] = comp_merge(sh[threadIdx.x], sh[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * (s * RED_BLOCKS + blockIdx.x)] = sh[0].s;
        partial[2 * (s * RED_BLOCKS + blockIdx.x) + 1] = sh[0].c;
    }
}

// GatherScatter: dest[c + nc*s] = 0.0 + src[rows...] in ascending row order
template <typename R>
__global__ void gather_kernel(R *dest, const R *src, const int64_t *ptr, const int32_t *rows,
                              int64_t nc, int64_t Np, int nf)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc)
        return;
    const int64_t lo = ptr[c], hi = ptr[c + 1];
    for (int s = 0; s < nf; ++s) {
        R v = R(0);
        for (int64_t r = lo; r < hi; ++r) {
            const int64_t e = rows[2 * r], node = rows[2 * r + 1];
            v = v + src[node + Np * (s + (int64_t)nf * e)];
        }
        dest[c + nc * s] = v;
    }
}

template <typename R>
__global__ void scatter_kernel(R *dest, const R *src, const int64_t *dtoc, int64_t nc,
                               int64_t Np, int64_t K, int nf)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Np * K)
        return;
    const int64_t e = i / Np, node = i - e * Np;
    const int64_t c = dtoc[i];
    for (int s = 0; s < nf; ++s)
        dest[node + Np * (s + (int64_t)nf * e)] = src[c + nc * s];
}

// Rebuild 1-based vmapM/vmapP/mapM/mapP from the face records.
__global__ void maps_kernel(const FaceRec *faces, const int64_t *explicit_idx, int dim, int nq,
                            int64_t Np, int64_t nfp, int ns, int64_t K, int64_t *vmapM,
                            int64_t *vmapP, int64_t *mapM, int64_t *mapP)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nf = 2 * dim;
    const int64_t nfpt = nf * nfp;
    if (idx >= nfpt * K)
        return;
    const int64_t e = idx / nfpt, slot = idx - e * nfpt;
    const int f = (int)(slot / nfp);
    const int64_t m = slot - f * nfp;
    const int64_t vm = face_node_rt(dim, nq, f, m) + Np * e + 1;
    const FaceRec rec = faces[e * nf + f];
    int64_t vp = vm, mp = idx + 1;
    if (rec.x >= 0) {
        const int fn = rec.y & 7, o = (rec.y >> 3) & 7;
        const int64_t mm = orient_rt(dim, nq, o, m);
        vp = face_node_rt(dim, nq, fn, mm) + Np * rec.x + 1;
        mp = fn * nfp + mm + nfpt * rec.x + 1;
    } else if (rec.x == FACE_EXPLICIT) {
        const int64_t off = explicit_idx[(int64_t)rec.y + m];
        const int64_t en = off / (ns * Np), node = off - en * ns * Np;
        vp = node + Np * en + 1;
        mp = 0;
    }
    vmapM[idx] = vm;
    vmapP[idx] = vp;
    mapM[idx] = idx + 1;
    mapP[idx] = mp;
}

// Volume metrics of curved cells from node coordinates, one block of Np
// threads per cell (thread = node).  3-D: hexvolumemetrics!
// (lobattocells.jl:3038-3180), the curl-invariant form of Kopriva (2006)
// eq. (37): a_r are reference derivatives of X x X_s, X x X_t, X x X_r, and
// dRdX[r, :] = a_r / (2 det J).  2-D: quadvolumemetrics! (:2756-2813),
// dRdX = inv(G) [X_r X_s]' with G the metric tensor.  The sums run in the
// reference's loop order and no multiply-add is fused, so every entry is the
// correctly rounded result of the same operation sequence as the host
// restatement.  Output: (Np, dim*dim + 2, K) = (dRdX column-major, J, wJ).
__global__ void __launch_bounds__(512)
curved_volume_metrics_kernel(int dim, int nq, const double *points, const double *D,
                             const double *w, double *metrics)
{
#pragma clang fp contract(off)
    extern __shared__ double mlds[];
    const int64_t e = blockIdx.x;
    const int np = dim == 3 ? nq * nq * nq : nq * nq;
    const int n2 = nq * nq;
    const int L = dim * dim + 2;
    const int node = threadIdx.x;
    const bool on = node < np;
    double *X = mlds;             // [3][np]
    double *cs = mlds + 3 * np;   // X x X_s
    double *ct = mlds + 6 * np;   // X x X_t
    double *cr = mlds + 9 * np;   // X x X_r
    const double *pe = points + (int64_t)dim * np * e;
    if (on)
        for (int d = 0; d < dim; ++d)
            X[d * np + node] = pe[d * np + node];
    __syncthreads();
    const int i = node % nq, j = (node / nq) % nq, k = dim == 3 ? node / n2 : 0;
    double *M = metrics + (int64_t)L * np * e;
    if (dim == 2) {
        if (!on)
            return;
        double xr[2] = {-0.0, -0.0}, xs[2] = {-0.0, -0.0};
        for (int m = 0; m < nq; ++m)
            for (int d = 0; d < 2; ++d)
                xr[d] += D[i + nq * m] * X[d * np + m + nq * j];
        for (int m = 0; m < nq; ++m)
            for (int d = 0; d < 2; ++d)
                xs[d] += D[j + nq * m] * X[d * np + i + nq * m];
        const double g11 = xr[0] * xr[0] + xr[1] * xr[1];
        const double g12 = xr[0] * xs[0] + xr[1] * xs[1];
        const double g22 = xs[0] * xs[0] + xs[1] * xs[1];
        const double det = g11 * g22 - g12 * g12;
        const double i11 = g22 / det, i12 = -g12 / det, i22 = g11 / det;
        for (int b = 0; b < 2; ++b) {
            M[(0 + 2 * b) * np + node] = i11 * xr[b] + i12 * xs[b];
            M[(1 + 2 * b) * np + node] = i12 * xr[b] + i22 * xs[b];
        }
        const double J = fabs(xr[0] * xs[1] - xr[1] * xs[0]);
        M[4 * np + node] = J;
        M[5 * np + node] = w[i] * w[j] * J;
        return;
    }
    double xr[3] = {-0.0, -0.0, -0.0}, xs[3] = {-0.0, -0.0, -0.0}, xt[3] = {-0.0, -0.0, -0.0};
    if (on) {
        for (int m = 0; m < nq; ++m)
            for (int d = 0; d < 3; ++d)
                xr[d] += D[i + nq * m] * X[d * np + m + nq * j + n2 * k];
        for (int m = 0; m < nq; ++m)
            for (int d = 0; d < 3; ++d)
                xs[d] += D[j + nq * m] * X[d * np + i + nq * m + n2 * k];
        for (int m = 0; m < nq; ++m)
            for (int d = 0; d < 3; ++d)
                xt[d] += D[k + nq * m] * X[d * np + i + nq * j + n2 * m];
        const double x[3] = {X[node], X[np + node], X[2 * np + node]};
        const double *v[3] = {xs, xt, xr};
        double *c[3] = {cs, ct, cr};
        for (int q = 0; q < 3; ++q) {
            c[q][0 * np + node] = x[1] * v[q][2] - x[2] * v[q][1];
            c[q][1 * np + node] = x[2] * v[q][0] - x[0] * v[q][2];
            c[q][2 * np + node] = x[0] * v[q][1] - x[1] * v[q][0];
        }
    }
    __syncthreads();
    if (!on)
        return;
    double a1[3] = {-0.0, -0.0, -0.0}, a2[3] = {-0.0, -0.0, -0.0}, a3[3] = {-0.0, -0.0, -0.0};
    for (int m = 0; m < nq; ++m)
        for (int d = 0; d < 3; ++d)
            a3[d] += D[i + nq * m] * cs[d * np + m + nq * j + n2 * k];
    for (int m = 0; m < nq; ++m)
        for (int d = 0; d < 3; ++d)
            a2[d] -= D[i + nq * m] * ct[d * np + m + nq * j + n2 * k];
    for (int m = 0; m < nq; ++m)
        for (int d = 0; d < 3; ++d)
            a1[d] += D[j + nq * m] * ct[d * np + i + nq * m + n2 * k];
    for (int m = 0; m < nq; ++m)
        for (int d = 0; d < 3; ++d)
            a3[d] -= D[j + nq * m] * cr[d * np + i + nq * m + n2 * k];
    for (int o = 0; o < nq; ++o)
        for (int d = 0; d < 3; ++d)
            a2[d] += D[k + nq * o] * cr[d * np + i + nq * j + n2 * o];
    for (int o = 0; o < nq; ++o)
        for (int d = 0; d < 3; ++d)
            a1[d] -= D[k + nq * o] * cs[d * np + i + nq * j + n2 * o];
    const double detJ = xr[0] * (xs[1] * xt[2] - xt[1] * xs[2]) -
                        xs[0] * (xr[1] * xt[2] - xt[1] * xr[2]) +
                        xt[0] * (xr[1] * xs[2] - xs[1] * xr[2]);
    const double twodet = 2 * detJ;
    const double *arow[3] = {a1, a2, a3};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            M[(a + 3 * b) * np + node] = arow[a][b] / twodet;
    M[9 * np + node] = detJ;
    M[10 * np + node] = w[i] * w[j] * w[k] * detJ;
}

// Surface metrics from the volume metrics at the face nodes
// (hexsurfacemetrics! lobattocells.jl:3253-3295, quadsurfacemetrics!
// :2862-2900): n = sn J dRdX[f/2, :], sJ = |n|, n /= sJ, wsJ = sJ * (face
// weights).  Output (dim+2, 2dim*Nfp, K): n, sJ, wsJ per face-node slot.
__global__ void curved_surface_metrics_kernel(int dim, int nq, int64_t K, const double *metrics,
                                              const double *w, double *surf)
{
#pragma clang fp contract(off)
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t np = dim == 3 ? nq * nq * nq : nq * nq;
    const int64_t nfp = dim == 3 ? nq * nq : nq;
    const int64_t nfpt = 2 * dim * nfp;
    if (idx >= nfpt * K)
        return;
    const int L = dim * dim + 2;
    const int64_t e = idx / nfpt, slot = idx - e * nfpt;
    const int f = (int)(slot / nfp);
    const int64_t m = slot - f * nfp;
    const int fg = f / 2;
    const double sn = (f % 2 == 0) ? -1.0 : 1.0;
    const int64_t nk = face_node_rt(dim, nq, f, m);
    const double *mb = metrics + nk + np * L * e;
    const double J = mb[np * (dim * dim)];
    const double snJ = sn * J;
    double nv[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < dim; ++c)
        nv[c] = snJ * mb[np * (fg + dim * c)];
    double ss = nv[0] * nv[0];
    for (int c = 1; c < dim; ++c)
        ss += nv[c] * nv[c];
    const double sJ = sqrt(ss);
    double *out = surf + (dim + 2) * idx;
    for (int c = 0; c < dim; ++c)
        out[c] = nv[c] / sJ;
    out[dim] = sJ;
    double wsJ = sJ;
    if (dim == 2)
        wsJ = sJ * w[m];
    else
        wsJ = sJ * w[m % nq] * w[m / nq];
    out[dim + 1] = wsJ;
}

// MJ of affine bricks: w_i w_j (w_k) J_e at every node.
struct WeightsArg {
    double w[32];
};
__global__ void affine_mj_kernel(double *MJ, const double *J, WeightsArg wa, int dim, int nq,
                                 int64_t K)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t np = nq;
    for (int d = 1; d < dim; ++d)
        np *= nq;
    if (idx >= np * K)
        return;
    const int64_t e = idx / np;
    int64_t rem = idx - e * np;
    double wp = 0;
    for (int c = 0; c < dim; ++c) {
        wp = c == 0 ? wa.w[rem % nq] : wp * wa.w[rem % nq];
        rem /= nq;
    }
    MJ[idx] = wp * J[e];
}

__global__ void round_to_f32_kernel(const double *__restrict__ src, float *__restrict__ dst,
                                    int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}

// The FP32 copies of the general geometry (and MJ) that the 3-D column kernel
// of an FP32 handle reads (StageParams::vgeo32 ...), made again after every
// change of the FP64 arrays.
int refresh_geo32(rv_handle *h)
{
    if (!(h->dtype == RV_F32 && h->geo == GEO_GENERAL && h->dim == 3) || h->K <= 0)
        return RV_OK;
    const int64_t nv = h->K * 9 * h->Np, nm = h->K * h->Np, ns = h->K * 6 * 4 * h->nfp;
    int rc;
    if (!h->vgeo32 && ((rc = dalloc(h, &h->vgeo32, nv)) || (rc = dalloc(h, &h->mji32, nm)) ||
                       (rc = dalloc(h, &h->sgeo32, ns)) || (rc = dalloc(h, &h->mj32, nm))))
        return rc;
    const struct {
        const double *src;
        float *dst;
        int64_t n;
    } parts[] = {{h->vgeo, h->vgeo32, nv}, {h->mji, h->mji32, nm}, {h->sgeo, h->sgeo32, ns},
                 {h->MJ, h->mj32, nm}};
    for (const auto &c : parts)
        hipLaunchKernelGGL(round_to_f32_kernel,
                           dim3((unsigned)std::min<int64_t>((c.n + 255) / 256, 4096)), dim3(256),
                           0, h->stream, c.src, c.dst, c.n);
    HIP_TRY(h, hipGetLastError());
    return RV_OK;
}

// Lift weight 0 on the non-conforming faces: the stage kernels then leave
// them out and the mortar kernel supplies their flux.
__global__ void zero_lift_kernel(double *sgeo, const int32_t *kind, int64_t K, int nf, int dim,
                                 int64_t nfp)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= K * nf * nfp)
        return;
    const int64_t ef = idx / nfp, m = idx - ef * nfp;
    if (kind[ef] != 0)
        sgeo[(ef * (dim + 1) + dim) * nfp + m] = 0.0;
}

// ---------------------------------------------------------------------------

int resolve_stream(rv_handle *h, void *stream)
{
    HIP_TRY(h, hipGetDevice(&h->device));
    // NULL selects the default (null) stream, which orders with the caller's
    // other default-stream work (torch's default stream is that stream)
    h->stream = static_cast<hipStream_t>(stream);
    h->own_stream = false;
    return RV_OK;
}

int finish_setup(rv_handle *h, const double *D, const double *w)
{
    const int nq = h->nq;
    std::vector<double> Dt(nq * nq), Dw(nq * nq), winv(nq);
    for (int i = 0; i < nq; ++i) {
        winv[i] = 1.0 / w[i];
        for (int n = 0; n < nq; ++n) {
            Dt[i * nq + n] = D[n + nq * i];
            Dw[i * nq + n] = D[n + nq * i] * w[n] / w[i];
        }
    }
    int rc;
    if ((rc = upload(h, &h->Dt, Dt.data(), Dt.size())) ||
        (rc = upload(h, &h->Dw, Dw.data(), Dw.size())) ||
        (rc = upload(h, &h->winv, winv.data(), winv.size())) ||
        (rc = dalloc(h, &h->partials, (size_t)2 * RED_BLOCKS * h->ns)) ||
        (rc = refresh_geo32(h)))
        return rc;
    KernelKey key{h->dim, nq, h->law, h->dtype == RV_F64 ? 8 : 4, h->geo, h->flux};
    key.vflux = h->vflux;
    h->launcher = find_stage_launcher(key);
    {
        // the fused relaxation stage: flux-differencing kernels, and the 3-D
        // column kernel of the weak form
        KernelKey rk = key;
        rk.relax = 1;
        h->relax_launcher = find_stage_launcher(rk);
    }
    // waves per workgroup measured on MI355X for the 3-D column kernel on the
    // 32^3 Nq = 8 Euler brick: a fused stage is fastest with four for FP64
    // (0.50 ms against 0.52 ms with one or two) and two for FP32 (one, two
    // and four within 2 %); a right-hand side alone with one for FP64
    // (0.36 ms against 0.37 and 0.40 ms) and two for FP32
    h->wpb = h->dim == 3 ? (h->dtype == RV_F32 ? 2 : 4) : 1;
    h->wpb_rhs = h->dim == 3 && h->dtype == RV_F32 ? 2 : 1;
    if (!h->launcher)
        return fail(h, RV_E_UNSUPPORTED,
                    "no kernel for dim=" + std::to_string(h->dim) + " nq=" + std::to_string(nq) +
                        " law=" + std::to_string(h->law) + " dtype=" + std::to_string(h->dtype) +
                        (h->vflux ? " (flux differencing: Nq^dim <= 512, Rusanov surface flux or"
                                    " the volume flux)"
                                  : " (weak form: Rusanov or central surface flux)"));
    return RV_OK;
}

int build_gatherscatter(rv_handle *h, const int64_t *dtoc, int64_t nc)
{
    if (!dtoc)
        return RV_OK;
    const int64_t n = h->Np * h->K;
    std::vector<int64_t> d0(n), ptr(nc + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
        if (dtoc[i] < 1 || dtoc[i] > nc)
            return fail(h, RV_E_ARG, "dtoc entry out of range at " + std::to_string(i));
        d0[i] = dtoc[i] - 1;
        ptr[d0[i] + 1] += 1;
    }
    for (int64_t c = 0; c < nc; ++c)
        ptr[c + 1] += ptr[c];
    // ctod = sparse(1:n, dtoc, true): rows of a column ascend
    std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
    std::vector<int32_t> rows(2 * std::max<int64_t>(n, 1));
    for (int64_t i = 0; i < n; ++i) {
        const int64_t r = fill[d0[i]]++;
        rows[2 * r] = (int32_t)(i / h->Np);
        rows[2 * r + 1] = (int32_t)(i % h->Np);
    }
    h->nc = nc;
    int rc;
    if ((rc = upload(h, &h->dtoc, d0.data(), d0.size())) ||
        (rc = upload(h, &h->ctod_ptr, ptr.data(), ptr.size())) ||
        (rc = upload(h, &h->ctod_rows, rows.data(), rows.size())))
        return rc;
    return RV_OK;
}

bool close_rel(double a, double b, double rtol)
{
    return std::fabs(a - b) <= rtol * std::max(std::fabs(a), std::fabs(b));
}

int create_from_desc(const rv_desc *d, rv_handle *h)
{
    const int dim = d->dim, nq = d->nq;
    const int64_t K = d->ncells;
    h->dim = dim;
    h->nq = nq;
    h->law = d->law;
    h->flux = d->flux ? d->flux : RV_FLUX_RUSANOV;
    h->dtype = d->dtype;
    h->K = K;
    h->ns = d->law == RV_LAW_ADVECTION ? 1 : dim + 2;
    h->Np = ipow(nq, dim);
    h->nfp = ipow(nq, dim - 1);
    h->gamma = d->gamma;
    h->dir_mask = d->directions ? d->directions : (1 << dim) - 1;
    for (int c = 0; c < 3; ++c)
        h->u[c] = d->velocity[c];
    const int64_t Np = h->Np, nfp = h->nfp, nf = 2 * dim, nfpt = nf * nfp;
    const int L = dim * dim + 2;
    if (K > 0 && (!d->D || !d->weights || !d->metrics || !d->surface_metrics || !d->vmapM ||
                  !d->vmapP))
        return fail(h, RV_E_ARG, "descriptor misses D, weights, metrics, surface metrics or maps");
    if ((int64_t)K * h->ns * Np >= ((int64_t)1 << 31) * 8)
        return fail(h, RV_E_ARG, "state too large");
    int rc = resolve_stream(h, d->stream);
    if (rc)
        return rc;
    const double *M = d->metrics, *S = d->surface_metrics, *w = d->weights;

    // MJ = wJ, MJI = 1 ./ MJ (dgsem.jl:55-56)
    std::vector<double> MJ(Np * K);
    for (int64_t e = 0; e < K; ++e)
        for (int64_t n = 0; n < Np; ++n)
            MJ[n + Np * e] = M[n + Np * ((L - 1) + (int64_t)L * e)];

    // ---- geometry: affine detection
    bool affine = d->geometry != RV_GEOM_GENERAL;
    std::vector<double> aff(K * dim);
    for (int64_t e = 0; e < K && affine; ++e) {
        const double *me = M + Np * (int64_t)L * e;
        const double J = me[Np * (dim * dim)];
        for (int a = 0; a < dim; ++a)
            aff[e * dim + a] = me[Np * (a + dim * a)];
        if (!(J > 0))
            affine = false;
        for (int a = 0; a < dim && affine; ++a)
            if (!(aff[e * dim + a] > 0))
                affine = false;
        for (int64_t n = 0; n < Np && affine; ++n) {
            for (int c = 0; c < dim && affine; ++c)
                for (int a = 0; a < dim && affine; ++a) {
                    const double g = me[n + Np * (a + dim * c)];
                    if (a == c ? g != aff[e * dim + a] : g != 0.0)
                        affine = false;
                }
            if (me[n + Np * (dim * dim)] != J)
                affine = false;
            int64_t rem = n;
            double wprod = 0;
            for (int c = 0; c < dim; ++c) {
                wprod = c == 0 ? w[rem % nq] : wprod * w[rem % nq];
                rem /= nq;
            }
            if (!close_rel(me[n + Np * (dim * dim + 1)], wprod * J, 1e-14))
                affine = false;
        }
        // surface metrics consistent with the volume ones: n = +-e_d, wsJ
        for (int f = 0; f < nf && affine; ++f) {
            const int fg = f / 2;
            const double sn = (f & 1) ? 1.0 : -1.0;
            for (int64_t m = 0; m < nfp && affine; ++m) {
                const double *s = S + (dim + 2) * (f * nfp + m + nfpt * e);
                for (int c = 0; c < dim; ++c)
                    if (s[c] != (c == fg ? sn : 0.0))
                        affine = false;
                double wf = 1.0;
                if (dim == 2)
                    wf = w[m];
                else if (dim == 3)
                    wf = w[m % nq] * w[m / nq];
                if (!close_rel(s[dim + 1], J * aff[e * dim + fg] * wf, 1e-14))
                    affine = false;
            }
        }
    }
    if (d->geometry == RV_GEOM_AFFINE && !affine)
        return fail(h, RV_E_ARG, "geometry forced affine but the metrics are not affine");
    h->geo = affine ? GEO_AFFINE : GEO_GENERAL;
    if (affine) {
        if ((rc = upload(h, &h->affine, aff.data(), aff.size())))
            return rc;
    } else {
        std::vector<double> vg((size_t)K * dim * dim * Np), mji(K * Np),
            sg((size_t)K * nf * (dim + 1) * nfp);
        for (int64_t e = 0; e < K; ++e) {
            const double *me = M + Np * (int64_t)L * e;
            for (int64_t n = 0; n < Np; ++n) {
                const double wJ = me[n + Np * (dim * dim + 1)];
                mji[e * Np + n] = 1.0 / wJ;
                for (int a = 0; a < dim; ++a)
                    for (int c = 0; c < dim; ++c)
                        vg[((e * dim + a) * dim + c) * Np + n] = wJ * me[n + Np * (a + dim * c)];
            }
            for (int f = 0; f < nf; ++f)
                for (int64_t m = 0; m < nfp; ++m) {
                    const double *s = S + (dim + 2) * (f * nfp + m + nfpt * e);
                    for (int c = 0; c < dim; ++c)
                        sg[((e * nf + f) * (dim + 1) + c) * nfp + m] = s[c];
                    sg[((e * nf + f) * (dim + 1) + dim) * nfp + m] = s[dim + 1];
                }
        }
        if ((rc = upload(h, &h->vgeo, vg.data(), vg.size())) ||
            (rc = upload(h, &h->mji, mji.data(), mji.size())) ||
            (rc = upload(h, &h->sgeo, sg.data(), sg.size())))
            return rc;
    }
    if ((rc = upload(h, &h->MJ, MJ.data(), MJ.size())))
        return rc;

    // ---- face records from vmapM / vmapP / boundary codes
    std::vector<FaceRec> recs(K * nf);
    std::vector<int64_t> expl;
    for (int64_t e = 0; e < K; ++e)
        for (int f = 0; f < nf; ++f) {
            const int64_t *vm = d->vmapM + f * nfp + nfpt * e;
            const int64_t *vp = d->vmapP + f * nfp + nfpt * e;
            for (int64_t m = 0; m < nfp; ++m)
                if (vm[m] != face_node_rt(dim, nq, f, m) + Np * e + 1)
                    return fail(h, RV_E_MAPS, "vmapM differs from the cell-face layout at cell " +
                                                  std::to_string(e) + " face " + std::to_string(f));
            const int64_t code = d->boundary_codes ? d->boundary_codes[f + nf * e] : 0;
            bool self = true;
            for (int64_t m = 0; m < nfp; ++m)
                self = self && vp[m] == vm[m];
            if (code != 0 || self) {
                // boundarystate is the identity for the laws here (balancelaw.jl:22)
                recs[e * nf + f] = {FACE_SELF, code != 0 ? 1 : 0};
                h->faces_self++;
                continue;
            }
            const int64_t en = (vp[0] - 1) / Np;
            bool found = false;
            if (en >= 0 && en < K) {
                for (int cand = 0; cand < nf && !found; ++cand) {
                    const int fn = (cand == 0) ? (f ^ 1) : (cand <= (f ^ 1) ? cand - 1 : cand);
                    for (int o = 0; o < (dim == 3 ? 8 : dim == 2 ? 2 : 1) && !found; ++o) {
                        bool ok = true;
                        for (int64_t m = 0; m < nfp && ok; ++m)
                            ok = vp[m] - 1 == face_node_rt(dim, nq, fn, orient_rt(dim, nq, o, m)) + Np * en;
                        if (ok) {
                            recs[e * nf + f] = {(int32_t)en, fn | (o << 3)};
                            found = true;
                        }
                    }
                }
            }
            if (!found) {
                recs[e * nf + f] = {FACE_EXPLICIT, (int32_t)expl.size()};
                for (int64_t m = 0; m < nfp; ++m) {
                    const int64_t v = vp[m] - 1;
                    if (v < 0 || v >= Np * K)
                        return fail(h, RV_E_MAPS, "vmapP entry out of range");
                    expl.push_back((v / Np) * h->ns * Np + v % Np);
                }
                h->faces_explicit++;
            }
        }
    if ((rc = upload(h, &h->faces, recs.data(), recs.size())))
        return rc;
    if (!expl.empty() && (rc = upload(h, &h->explicit_idx, expl.data(), expl.size())))
        return rc;
    if ((rc = build_gatherscatter(h, d->dtoc, d->ncontinuous)))
        return rc;
    return finish_setup(h, d->D, d->weights);
}

constexpr int64_t STREAM_BYTES = (int64_t)256 << 20;

StageParams stage_params(rv_handle *h)
{
    StageParams P{};
    P.K = h->K;
    P.dir_mask = h->dir_mask;
    P.faces = as_global(h->faces);
    P.explicit_idx = as_global(h->explicit_idx);
    P.affine = as_global(h->affine);
    P.vgeo = as_global(h->vgeo);
    P.mji = as_global(h->mji);
    P.mj = as_global(h->MJ);
    P.sgeo = as_global(h->sgeo);
    P.vgeo32 = as_global(h->vgeo32);
    P.mji32 = as_global(h->mji32);
    P.sgeo32 = as_global(h->sgeo32);
    P.mj32 = as_global(h->mj32);
    P.Dt = as_global(h->Dt);
    P.Dw = as_global(h->Dw);
    P.winv = as_global(h->winv);
    P.gamma = h->gamma;
    for (int c = 0; c < 3; ++c)
        P.u[c] = h->u[c];
    // a stage touches q_in, dq and q_out; beyond the 256 MB Infinity Cache of
    // MI355X nothing of dq and q_out survives to the next stage, so they are
    // streamed past L2 (which then keeps the q lines that face reads share)
    const int64_t state = h->Np * h->ns * h->K * (h->dtype == RV_F32 ? 4 : 8);
    P.streaming = 3 * state > STREAM_BYTES ? 1 : 0;
    return P;
}

int launch(rv_handle *h, const StageParams &P0, StageLauncher fn = nullptr)
{
    StageParams P = P0;
    MortarParams M{};
    // the 3-D weak-form column kernel adds the mortar updates after its own
    // stores; the other stage kernels leave that to mortar_apply_kernel
    const bool fused = h->mortar && !fn && h->dim == 3 && h->vflux == 0;
    if (h->mortar) {
        M.stage = P;
        M.ncells = h->mortar_cells;
        M.cells = as_global(h->m_cells);
        M.npairs = h->mortar_pairs;
        M.pairs = as_global(h->m_pairs);
        M.kind = as_global(h->m_kind);
        M.partner = as_global(h->m_partner);
        M.sgeo = as_global(h->m_sgeo);
        M.halves = as_global(h->m_halves);
        M.delta = as_global(static_cast<void *>(h->m_delta));
        if (h->mortar(M, h->stream, 0) != 0) {
            hipError_t e = hipGetLastError();
            return fail(h, RV_E_HIP, std::string("mortar launch failed: ") + hipGetErrorString(e));
        }
        if (fused) {
            P.mslot = as_global(h->m_slot);
            P.mkind = as_global(h->m_kind);
            P.mdelta = M.delta;
        }
    }
    if ((fn ? fn : h->launcher)(P, h->stream, P.write_q ? h->wpb : h->wpb_rhs) != 0) {
        hipError_t e = hipGetLastError();
        return fail(h, RV_E_HIP, std::string("stage launch failed: ") + hipGetErrorString(e));
    }
    if (h->mortar && !fused && h->mortar(M, h->stream, 1) != 0) {
        hipError_t e = hipGetLastError();
        return fail(h, RV_E_HIP, std::string("mortar launch failed: ") + hipGetErrorString(e));
    }
    return RV_OK;
}

int reduce(rv_handle *h, const void *q, int power, double *out, double pexp = 2.0)
{
    if (!q)
        return fail(h, RV_E_ARG, "null state");
    dim3 grid(RED_BLOCKS, h->ns);
    if (h->dtype == RV_F64)
        hipLaunchKernelGGL(reduce_kernel<double>, grid, dim3(RED_THREADS), 0, h->stream,
                           static_cast<const double *>(q), h->MJ, h->Np, h->ns, h->K, power,
                           pexp, h->partials);
    else
        hipLaunchKernelGGL(reduce_kernel<float>, grid, dim3(RED_THREADS), 0, h->stream,
                           static_cast<const float *>(q), h->MJ, h->Np, h->ns, h->K, power,
                           pexp, h->partials);
    HIP_TRY(h, hipGetLastError());
    std::vector<double> part((size_t)2 * RED_BLOCKS * h->ns);
    HIP_TRY(h, hipMemcpyAsync(part.data(), h->partials, part.size() * sizeof(double),
                              hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    for (int s = 0; s < h->ns; ++s) {
        if (power == 4) {
            double m = 0.0;
            for (int b = 0; b < RED_BLOCKS; ++b)
                m = max_nan(m, part[2 * (s * RED_BLOCKS + b)]);
            out[s] = m;
            continue;
        }
        // Neumaier over the block partials in block order
        double sum = 0, comp = 0;
        for (int b = 0; b < RED_BLOCKS; ++b) {
            for (int k = 0; k < 2; ++k) {
                const double x = part[2 * (s * RED_BLOCKS + b) + k];
                const double t = sum + x;
                comp += std::fabs(sum) >= std::fabs(x) ? (sum - t) + x : (x - t) + sum;
                sum = t;
            }
        }
        out[s] = sum + comp;
    }
    return RV_OK;
}

// Entropy passes (entropy.hip): slot 0 holds Newton evaluations and the
// diagnostics, slot 1 the entropy integral of q0, slots 2.. the stages
constexpr int RELAX_SLOTS = 16;
// relax_sums: 2 per slot, then one per fused relaxation stage (at most 14)
constexpr int RELAX_SUMS = 2 * RELAX_SLOTS + 16;
// Newton iterations of the relaxation root per step at most
constexpr int RELAX_MAX_NEWTON = 50;

int ensure_relax(rv_handle *h)
{
    if (h->law != RV_LAW_ADVECTION && h->law != RV_LAW_EULER)
        return fail(h, RV_E_UNSUPPORTED, "no entropy for this law");
    if (h->relax_partials)
        return RV_OK;
    if (int rc = dalloc(h, &h->relax_sums, (size_t)RELAX_SUMS))
        return rc;
    HIP_TRY(h, hipHostMalloc((void **)&h->relax_sums_host, RELAX_SUMS * sizeof(double)));
    return dalloc(h, &h->relax_partials, (size_t)4 * RELAX_BLOCKS * RELAX_SLOTS);
}

// partial sums one fused relaxation stage leaves: one per workgroup of the
// flux-differencing launch
// Product slots a fused relaxation stage writes (one per workgroup of the
// flux-differencing kernel, one per wave of the 3-D column kernel) ...
int64_t relax_partials_count(const rv_handle *h)
{
    if (h->vflux == 0)
        return plane3_relax_slots(h->nq, h->K, h->wpb);
    const int cpb = fd_cells_per_workgroup((int)h->Np);
    return (h->K + cpb - 1) / cpb;
}

// ... and the most it can write whatever the waves per workgroup (the stride
// of the per-stage slots in eprod, fixed when eprod is allocated)
int64_t relax_partials_cap(const rv_handle *h)
{
    if (h->vflux == 0)
        return plane3_relax_slots(h->nq, h->K, 4);
    return relax_partials_count(h);
}

RelaxParams relax_params(rv_handle *h, int mode, int slot)
{
    RelaxParams P{};
    P.mode = mode;
    P.law = h->law == RV_LAW_ADVECTION ? LAW_ADVECTION : LAW_EULER;
    P.dim = h->dim;
    P.real_bytes = h->dtype == RV_F64 ? 8 : 4;
    P.Np = h->Np;
    P.K = h->K;
    P.gas_gamma = h->gamma;
    P.MJ = h->MJ;
    P.partial = h->relax_partials + (size_t)4 * RELAX_BLOCKS * slot;
    return P;
}

int relax(rv_handle *h, const RelaxParams &P)
{
    if (launch_relax(P, h->stream) != 0) {
        hipError_t e = hipGetLastError();
        return fail(h, RV_E_HIP, std::string("entropy pass launch failed: ") + hipGetErrorString(e));
    }
    return RV_OK;
}

// Compensated sums of slots [0, nslots), formed on the device in a fixed
// order: out[2*slot] the entropy integral, out[2*slot + 1] the entropy
// product.  With nstage > 0 also the fused stages' product sums (launched
// into relax_sums + 2 RELAX_SLOTS by the caller) into stage_out.  One pinned
// read-back, one synchronisation.
int relax_read(rv_handle *h, int nslots, double *out, int nstage = 0, double *stage_out = nullptr)
{
    if (launch_slot_sums(h->relax_partials, 2 * RELAX_BLOCKS, nslots, 2,
                         (int64_t)4 * RELAX_BLOCKS, 2, 2, 4, h->relax_sums, h->stream) != 0) {
        hipError_t e = hipGetLastError();
        return fail(h, RV_E_HIP, std::string("entropy sum launch failed: ") + hipGetErrorString(e));
    }
    HIP_TRY(h, hipMemcpyAsync(h->relax_sums_host, h->relax_sums, 2 * nslots * sizeof(double),
                              hipMemcpyDeviceToHost, h->stream));
    if (nstage > 0)
        HIP_TRY(h, hipMemcpyAsync(h->relax_sums_host + 2 * RELAX_SLOTS,
                                  h->relax_sums + 2 * RELAX_SLOTS, nstage * sizeof(double),
                                  hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    for (int i = 0; i < 2 * nslots; ++i)
        out[i] = h->relax_sums_host[i];
    for (int i = 0; i < nstage; ++i)
        stage_out[i] = h->relax_sums_host[2 * RELAX_SLOTS + i];
    return RV_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

// The volume form of a descriptor as the kernel key's two-point flux (0: weak
// form); the entropy-conservative flux is defined for Euler only
static int volume_flux_of(int form, int flux, int law, int *vflux)
{
    *vflux = 0;
    if (form == RV_FORM_WEAK)
        return RV_OK;
    if (form != RV_FORM_FLUX_DIFFERENCING)
        return RV_E_ARG;
    if (flux == RV_FLUX_CENTRAL || (flux == RV_FLUX_EC && law == RV_LAW_EULER)) {
        *vflux = flux;
        return RV_OK;
    }
    return RV_E_ARG;
}

int rv_create(const rv_desc *desc, rv_handle **out)
{
    if (!out)
        return fail(nullptr, RV_E_ARG, "out is null");
    *out = nullptr;
    if (!desc)
        return fail(nullptr, RV_E_ARG, "descriptor is null");
    if (desc->abi_version != 1 && desc->abi_version != RAVENHIP_ABI_VERSION)
        return fail(nullptr, RV_E_ARG, "ABI version mismatch");
    int vflux = 0;
    if (desc->abi_version >= 2 &&
        volume_flux_of(desc->volume_form, desc->volume_flux, desc->law, &vflux) != RV_OK)
        return fail(nullptr, RV_E_ARG, "invalid volume form or volume flux");
    if (desc->dim < 1 || desc->dim > 3 || desc->nq < 2 || desc->nq > 16 || desc->ncells < 0 ||
        (desc->law != RV_LAW_ADVECTION && desc->law != RV_LAW_EULER) ||
        (desc->dtype != RV_F64 && desc->dtype != RV_F32) ||
        (desc->flux != 0 && desc->flux != RV_FLUX_RUSANOV && desc->flux != RV_FLUX_CENTRAL &&
         !(desc->flux == RV_FLUX_EC && desc->law == RV_LAW_EULER)))
        return fail(nullptr, RV_E_ARG, "invalid descriptor field (dim, nq, ncells, law, dtype or flux)");
    rv_handle *h = new rv_handle();
    h->vflux = vflux;
    int rc = create_from_desc(desc, h);
    if (rc) {
        g_last_error = h->err;
        free_handle(h);
        return rc;
    }
    *out = h;
    return RV_OK;
}

static int brick_spec(const rv_brick_desc *bd, BrickSpec *b)
{
    if (!bd || (bd->abi_version != 1 && bd->abi_version != RAVENHIP_ABI_VERSION))
        return fail(nullptr, RV_E_ARG, "null descriptor or ABI version mismatch");
    if (bd->dim < 1 || bd->dim > 3 || bd->nq < 2 || bd->nq > 16 || bd->level < 0 ||
        bd->level > 10 || (bd->law != RV_LAW_ADVECTION && bd->law != RV_LAW_EULER) ||
        (bd->dtype != RV_F64 && bd->dtype != RV_F32))
        return fail(nullptr, RV_E_ARG, "invalid brick descriptor");
    *b = BrickSpec{};
    b->dim = bd->dim;
    b->nq = bd->nq;
    b->level = bd->level;
    int64_t cells = 1;
    for (int d = 0; d < 3; ++d) {
        b->n[d] = d < bd->dim ? bd->n[d] : 1;
        b->periodic[d] = d < bd->dim ? bd->periodic[d] : 0;
        b->lo[d] = bd->lo[d];
        b->hi[d] = bd->hi[d];
        if (d < bd->dim && (b->n[d] < 1 || b->n[d] > (1 << 20) || !(b->hi[d] > b->lo[d])))
            return fail(nullptr, RV_E_ARG, "invalid brick extents");
        cells *= b->n[d];
    }
    if ((cells << (bd->dim * bd->level)) > ((int64_t)1 << 31))
        return fail(nullptr, RV_E_ARG, "brick has too many cells");
    return RV_OK;
}

int64_t rv_brick_numcells(const rv_brick_desc *bd)
{
    BrickSpec b;
    const int rc = brick_spec(bd, &b);
    return rc ? rc : brick_numcells(b);
}

int rv_create_brick(const rv_brick_desc *bd, rv_handle **out)
{
    if (!out)
        return fail(nullptr, RV_E_ARG, "out is null");
    *out = nullptr;
    BrickSpec b;
    if (int rc0 = brick_spec(bd, &b))
        return rc0;
    rv_handle *h = new rv_handle();
    h->is_brick = true;
    h->brick = b;
    const int dim = b.dim, nq = b.nq;
    h->dim = dim;
    h->nq = nq;
    h->law = bd->law;
    h->flux = bd->flux ? bd->flux : RV_FLUX_RUSANOV;
    if (bd->abi_version >= 2 &&
        volume_flux_of(bd->volume_form, bd->volume_flux, bd->law, &h->vflux) != RV_OK) {
        free_handle(h);
        return fail(nullptr, RV_E_ARG, "invalid volume form or volume flux");
    }
    h->dtype = bd->dtype;
    h->K = brick_numcells(b);
    h->ns = bd->law == RV_LAW_ADVECTION ? 1 : dim + 2;
    h->Np = ipow(nq, dim);
    h->nfp = ipow(nq, dim - 1);
    h->gamma = bd->gamma;
    h->dir_mask = (1 << dim) - 1;
    for (int c = 0; c < 3; ++c)
        h->u[c] = bd->velocity[c];
    h->geo = GEO_AFFINE;
    int rc = resolve_stream(h, bd->stream);
    const int64_t K = h->K, Np = h->Np, nf = 2 * dim;
    std::vector<double> x(nq), w(nq), D(nq * nq), drdx(K * dim), J(K);
    std::vector<int64_t> nbr(K * nf);
    std::vector<FaceRec> recs(K * nf);
    if (!rc) {
        lobatto_operators(nq, x.data(), w.data(), D.data());
        brick_affine(b, drdx.data(), J.data());
        brick_neighbours(b, nbr.data());
        for (int64_t e = 0; e < K; ++e) {
            for (int f = 0; f < nf; ++f) {
                const int64_t en = nbr[e * nf + f];
                recs[e * nf + f] = en < 0 ? FaceRec{FACE_SELF, 1} : FaceRec{(int32_t)en, f ^ 1};
                if (en < 0)
                    h->faces_self++;
            }
        }
        // MJ = w_i w_j (w_k) J per node, filled on the device from the
        // per-cell J by a kernel queued on the handle's stream (ordered before
        // every later use of MJ, which runs on that stream too); set-up neither
        // waits for the stream nor frees memory, so it does not stall work
        // already queued on the device.  The continuous numbering is built on
        // the first rv_gather / rv_scatter (solving never needs it).
        if (!(rc = upload(h, &h->affine, drdx.data(), drdx.size())) &&
            !(rc = upload(h, &h->cellJ, J.data(), J.size())) &&
            !(rc = dalloc(h, &h->MJ, K * Np)) &&
            !(rc = upload(h, &h->faces, recs.data(), recs.size()))) {
            WeightsArg wa{};
            for (int i = 0; i < nq; ++i)
                wa.w[i] = w[i];
            const int64_t n = K * Np;
            if (n > 0)
                hipLaunchKernelGGL(affine_mj_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256),
                                   0, h->stream, h->MJ, h->cellJ, wa, dim, nq, K);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess)
                rc = fail(h, RV_E_HIP, std::string("MJ fill: ") + hipGetErrorString(e));
            if (!rc)
                rc = finish_setup(h, D.data(), w.data());
        }
    }
    if (rc) {
        g_last_error = h->err;
        free_handle(h);
        return rc;
    }
    *out = h;
    return RV_OK;
}

void rv_destroy(rv_handle *h)
{
    if (h && h->stream)
        (void)hipStreamSynchronize(h->stream);
    free_handle(h);
}

const char *rv_last_error(const rv_handle *h)
{
    return h ? h->err.c_str() : g_last_error.c_str();
}

int rv_get_info(const rv_handle *h, rv_info *info)
{
    if (!h || !info)
        return fail(nullptr, RV_E_ARG, "null argument");
    std::memset(info, 0, sizeof(*info));
    info->dim = h->dim;
    info->nq = h->nq;
    info->ns = h->ns;
    info->dtype = h->dtype;
    info->geometry = h->geo == GEO_AFFINE ? RV_GEOM_AFFINE : RV_GEOM_GENERAL;
    info->faces_self = h->faces_self;
    info->faces_explicit = h->faces_explicit;
    info->streaming = stage_params(const_cast<rv_handle *>(h)).streaming;
    info->ncells = h->K;
    const int64_t b = h->dtype == RV_F64 ? 8 : 4;
    const int64_t nodes = h->Np * h->K;
    info->state_bytes = nodes * h->ns * b;
    info->device_bytes = h->device_bytes;
    // compulsory traffic per node: read q, [read dq], write dq, write q; plus
    // per-node geometry (wJ*dRdX, 1/wJ) and per face-node (n, wsJ) when the
    // cells are not affine, FP64 except for the 3-D weak form of FP32
    // handles, whose stage reads the FP32 copies; affine geometry is per cell
    // (~0 per node)
    int64_t geo = 0;
    if (h->geo == GEO_GENERAL) {
        const int gb = h->dtype == RV_F32 && h->dim == 3 && h->vflux == 0 ? 4 : 8;
        geo = nodes * (h->dim * h->dim + 1) * gb + h->K * 2 * h->dim * h->nfp * (h->dim + 1) * gb;
    }
    info->stage_bytes = 4 * info->state_bytes + geo;
    info->stage_bytes_first = 3 * info->state_bytes + geo;
    info->rhs_bytes = 2 * info->state_bytes + geo;
    return RV_OK;
}

int rv_synchronize(rv_handle *h)
{
    if (!h)
        return fail(nullptr, RV_E_ARG, "null handle");
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return RV_OK;
}

void *rv_stream(const rv_handle *h) { return h ? (void *)h->stream : nullptr; }

int rv_state_alloc(rv_handle *h, void **state)
{
    if (!h || !state)
        return fail(h, RV_E_ARG, "null argument");
    const size_t bytes = (size_t)h->Np * h->ns * h->K * (h->dtype == RV_F32 ? 4 : 8);
    *state = nullptr;
    hipError_t e = hipMalloc(state, bytes ? bytes : 1);
    if (e != hipSuccess) {
        *state = nullptr;
        return fail(h, RV_E_NOMEM, "hipMalloc of a state failed");
    }
    return RV_OK;
}

int rv_state_free(rv_handle *h, void *state)
{
    if (!h)
        return fail(nullptr, RV_E_ARG, "null handle");
    if (!state)
        return RV_OK;
    // work queued on the handle's stream may still read or write the state
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    HIP_TRY(h, hipFree(state));
    return RV_OK;
}

int rv_state_upload(rv_handle *h, void *state, const void *host)
{
    if (!h || !state || !host)
        return fail(h, RV_E_ARG, "null argument");
    const size_t bytes = (size_t)h->Np * h->ns * h->K * (h->dtype == RV_F32 ? 4 : 8);
    HIP_TRY(h, hipMemcpyAsync(state, host, bytes, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return RV_OK;
}

int rv_state_download(rv_handle *h, void *host, const void *state)
{
    if (!h || !state || !host)
        return fail(h, RV_E_ARG, "null argument");
    const size_t bytes = (size_t)h->Np * h->ns * h->K * (h->dtype == RV_F32 ? 4 : 8);
    HIP_TRY(h, hipMemcpyAsync(host, state, bytes, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return RV_OK;
}

int rv_set_waves_per_block(rv_handle *h, int waves_per_block)
{
    if (!h)
        return fail(nullptr, RV_E_ARG, "null handle");
    if (waves_per_block != 1 && waves_per_block != 2 && waves_per_block != 4)
        return fail(h, RV_E_ARG, "waves per block must be 1, 2 or 4");
    h->wpb = waves_per_block;
    h->wpb_rhs = waves_per_block;
    return RV_OK;
}

int rv_get_waves_per_block(const rv_handle *h) { return h ? h->wpb : 0; }

int rv_enable_mortars(rv_handle *h, const int32_t *face_kind, const int64_t *partner)
{
    if (!h || !face_kind || !partner)
        return fail(h, RV_E_ARG, "null argument");
    if (h->geo != GEO_GENERAL)
        return fail(h, RV_E_UNSUPPORTED, "mortars need per-node (general) geometry");
    if (h->mortar)
        return fail(h, RV_E_STATE, "mortars are already enabled");
    const int nf = 2 * h->dim, nsub = 1 << (h->dim - 1);
    std::vector<int64_t> cells;
    for (int64_t e = 0; e < h->K; ++e) {
        bool any = false;
        for (int f = 0; f < nf; ++f) {
            const int k = face_kind[e * nf + f];
            const int64_t *p = partner + 4 * (e * nf + f);
            if (k < 0 || k > 2)
                return fail(h, RV_E_ARG, "face kind must be 0, 1 or 2");
            if (k == 1 && (p[0] < 0 || p[0] >= h->K || p[1] < 0 || p[1] >= nsub))
                return fail(h, RV_E_ARG, "hanging face without a valid coarse partner");
            if (k == 2)
                for (int s = 0; s < nsub; ++s)
                    if (p[s] < 0 || p[s] >= h->K)
                        return fail(h, RV_E_ARG, "covered face without its finer partners");
            any = any || k != 0;
        }
        if (any)
            cells.push_back(e);
    }
    MortarLauncher ml = find_mortar_launcher(h->dim, h->nq, h->law, h->dtype == RV_F32 ? 4 : 8,
                                             h->flux);
    if (!ml)
        return fail(h, RV_E_UNSUPPORTED,
                    "no mortar kernel for this (dim, Nq, surface flux): 2-D or 3-D, Nq <= 8,"
                    " Rusanov or central");
    std::vector<double> halves(2 * h->nq * h->nq);
    lobatto_halves(h->nq, halves.data(), halves.data() + h->nq * h->nq);
    std::vector<int32_t> slot(h->K, -1);
    // the face kernel's workgroups: covered sides (the longer ones) first,
    // then hanging sides, each in (cell, face) order
    std::vector<MortarPair> pairs, hanging;
    for (size_t li = 0; li < cells.size(); ++li) {
        const int64_t e = cells[li];
        slot[e] = (int32_t)li;
        for (int f = 0; f < nf; ++f) {
            const int k = face_kind[e * nf + f];
            if (k == 0)
                continue;
            MortarPair pr{};
            pr.e = e;
            pr.li = (int64_t)li;
            for (int t = 0; t < 4; ++t)
                pr.part[t] = partner[4 * (e * nf + f) + t];
            pr.f = f;
            pr.kind = k;
            (k == 2 ? pairs : hanging).push_back(pr);
        }
    }
    pairs.insert(pairs.end(), hanging.begin(), hanging.end());
    if (pairs.size() > (size_t)UINT32_MAX)
        return fail(h, RV_E_UNSUPPORTED, "too many non-conforming faces for one launch");
    const int64_t npairs = (int64_t)pairs.size();
    const size_t nsg = (size_t)h->K * nf * (h->dim + 1) * h->nfp;
    const int64_t ncells = (int64_t)cells.size();
    // (a grid without non-conforming faces keeps one-entry lists)
    if (cells.empty())
        cells.push_back(0);
    if (pairs.empty())
        pairs.push_back(MortarPair{});
    int rc;
    if ((rc = upload(h, &h->m_cells, cells.data(), cells.size())) ||
        (rc = upload(h, &h->m_kind, face_kind, (size_t)h->K * nf)) ||
        (rc = upload(h, &h->m_partner, partner, (size_t)h->K * nf * 4)) ||
        (rc = upload(h, &h->m_halves, halves.data(), halves.size())) ||
        (rc = upload(h, &h->m_slot, slot.data(), slot.size())) ||
        (rc = upload(h, &h->m_pairs, pairs.data(), pairs.size())) ||
        (rc = dalloc(h, &h->m_sgeo, nsg)) ||
        (rc = dalloc(h, &h->m_delta, cells.size() * nf * h->ns * h->nfp *
                                         (h->dtype == RV_F32 ? 4 : 8))))
        return rc;
    HIP_TRY(h, hipMemcpyAsync(h->m_sgeo, h->sgeo, nsg * sizeof(double), hipMemcpyDeviceToDevice,
                              h->stream));
    const int64_t n = h->K * nf * h->nfp;
    hipLaunchKernelGGL(zero_lift_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       h->stream, h->sgeo, h->m_kind, h->K, nf, h->dim, h->nfp);
    HIP_TRY(h, hipGetLastError());
    if ((rc = refresh_geo32(h)))
        return rc;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    h->mortar_cells = ncells;
    h->mortar_pairs = npairs;
    h->mortar = ml;
    return RV_OK;
}

int rv_rhs(rv_handle *h, void *dq, const void *q, double t, int increment)
{
    (void)t;  // no time-dependent sources for the laws here (balancelaw.jl:23-26)
    if (!h || !dq || !q)
        return fail(h, RV_E_ARG, "null argument");
    StageParams P = stage_params(h);
    P.q_in = as_global(q);
    P.q_out = nullptr;
    P.dq = as_global(dq);
    P.a = 1.0;
    P.b_dt = 0.0;
    P.read_dq = increment ? 1 : 0;
    P.write_q = 0;
    return launch(h, P);
}

int rv_lsrk_stage(rv_handle *h, void *q_out, const void *q_in, void *dq, double a, double b_dt,
                  double t)
{
    (void)t;
    if (!h || !q_out || !q_in || !dq)
        return fail(h, RV_E_ARG, "null argument");
    if (q_out == q_in)
        return fail(h, RV_E_ARG, "q_out must not alias q_in");
    StageParams P = stage_params(h);
    P.q_in = as_global(q_in);
    P.q_out = as_global(q_out);
    P.dq = as_global(dq);
    P.a = a;
    P.b_dt = b_dt;
    P.read_dq = a != 0.0 ? 1 : 0;
    P.write_q = 1;
    return launch(h, P);
}

// LSRK54 (Carpenter & Kennedy) and LSRK144 (Niegemann, Diehl & Busch)
// coefficients, odesolvers.jl:163-244; the rationals are rounded to double
// like Float64(::Rational)
namespace {
static const double a54[5] = {0.0, -567301805773.0 / 1357537059087.0,
                              -2404267990393.0 / 2016746695238.0,
                              -3550918686646.0 / 2091501179385.0,
                              -1275806237668.0 / 842570457699.0};
static const double b54[5] = {1432997174477.0 / 9575080441755.0,
                              5161836677717.0 / 13612068292357.0,
                              1720146321549.0 / 2090206949498.0,
                              3134564353537.0 / 4481467310338.0,
                              2277821191437.0 / 14882151754819.0};
static const double c54[5] = {0.0, 1432997174477.0 / 9575080441755.0,
                              2526269341429.0 / 6820363962896.0,
                              2006345519317.0 / 3224310063776.0,
                              2802321613138.0 / 2924317926251.0};
static const double a144[14] = {0, -0.7188012108672410, -0.7785331173421570,
                                -0.0053282796654044, -0.8552979934029281,
                                -3.9564138245774565, -1.5780575380587385,
                                -2.0837094552574054, -0.7483334182761610,
                                -0.7032861106563359, 0.0013917096117681,
                                -0.0932075369637460, -0.9514200470875948,
                                -7.1151571693922548};
static const double b144[14] = {0.0367762454319673, 0.3136296607553959, 0.1531848691869027,
                                0.0030097086818182, 0.3326293790646110, 0.2440251405350864,
                                0.3718879239592277, 0.6204126221582444, 0.1524043173028741,
                                0.0760894927419266, 0.0077604214040978, 0.0024647284755382,
                                0.0780348340049386, 5.5059777270269628};
static const double c144[14] = {0, 0.0367762454319673, 0.1249685262725025,
                                0.2446177702277698, 0.2476149531070420,
                                0.2969311120382472, 0.3978149645802642,
                                0.5270854589440328, 0.6981269994175695,
                                0.8190890835352128, 0.8527059887098624,
                                0.8604711817462826, 0.8627060376969976,
                                0.8734213127600976};

// number of stages (0 for an unknown scheme) and the three tables
int lsrk_tables(int scheme, const double **A, const double **B, const double **C)
{
    if (scheme == RV_LSRK54) {
        *A = a54, *B = b54, *C = c54;
        return 5;
    }
    if (scheme == RV_LSRK144) {
        *A = a144, *B = b144, *C = c144;
        return 14;
    }
    return 0;
}
}  // namespace

int rv_lsrk_coefficients(int scheme, double *a, double *b, double *c, int *nstages)
{
    const double *A, *B, *C;
    const int n = lsrk_tables(scheme, &A, &B, &C);
    if (n == 0)
        return RV_E_ARG;
    if (nstages)
        *nstages = n;
    for (int s = 0; s < n; ++s) {
        if (a)
            a[s] = A[s];
        if (b)
            b[s] = B[s];
        if (c)
            c[s] = C[s];
    }
    return RV_OK;
}

namespace {

// solve!(q, timeend, LSRK(dg, q, dt; t0); adjust_final) (odesolvers.jl:9-61)
// with the ping-pong state `alt` (distinct from q and dq, one state's bytes)
int lsrk_run_impl(rv_handle *h, void *q, void *dq, int scheme, double t0, double dt,
                  double timeend, int adjust_final, void *alt, double *t_out,
                  int64_t *nsteps_out)
{
    const double *A, *B, *C;
    const int nst = lsrk_tables(scheme, &A, &B, &C);
    if (nst == 0)
        return fail(h, RV_E_ARG, "unknown LSRK scheme");
    (void)C;  // the laws here do not depend on the stage time
    const bool f32 = h->dtype == RV_F32;
    const size_t bytes = (size_t)h->Np * h->ns * h->K * (f32 ? 4 : 8);
    // dq starts at zero (odesolvers.jl:43-44); a first stage with a = 0 does
    // not read it (dq = L(q), as dq .*= 0 plus the increment gives from zero)
    if (A[0] != 0.0)
        HIP_TRY(h, hipMemsetAsync(dq, 0, bytes, h->stream));
    // LSRK(rhs!, ...) converts dt, t0 and the coefficients to the state type
    // (odesolvers.jl:41-48); in FP32 the clock is FP32 too.  solve! compares
    // time + dt with timeend as given and assigns timeend - time to the FP32
    // dt field (rounding the difference), as Julia does for a Float64 timeend.
    auto rnd = [f32](double v) { return f32 ? (double)(float)v : v; };
    double time = rnd(t0), step_dt = rnd(dt);
    void *cur = q;
    int64_t steps = 0;
    StageParams P = stage_params(h);
    for (;;) {
        ++steps;
        bool final_step = false;
        const double tdt = f32 ? (double)((float)time + (float)step_dt) : time + step_dt;
        if (tdt >= timeend) {
            if (adjust_final)
                step_dt = rnd(timeend - time);
            final_step = true;
        }
        for (int s = 0; s < nst; ++s) {
            P.q_in = as_global(cur);
            P.q_out = as_global(alt);
            P.dq = as_global(dq);
            P.a = rnd(A[s]);
            P.b_dt = f32 ? (double)((float)rnd(B[s]) * (float)step_dt) : B[s] * step_dt;
            P.read_dq = P.a != 0.0 ? 1 : 0;
            P.write_q = 1;
            int rc = launch(h, P);
            if (rc)
                return rc;
            std::swap(cur, alt);
        }
        time = f32 ? (double)((float)time + (float)step_dt) : time + step_dt;
        if (final_step)
            break;
        if (steps > (int64_t)1 << 40)
            return fail(h, RV_E_ARG, "time step too small");
    }
    if (cur != q)
        HIP_TRY(h, hipMemcpyAsync(q, cur, bytes, hipMemcpyDeviceToDevice, h->stream));
    if (t_out)
        *t_out = time;
    if (nsteps_out)
        *nsteps_out = steps;
    return RV_OK;
}

}  // namespace

int rv_lsrk_solve(rv_handle *h, void *q, void *dq, int scheme, double t0, double dt,
                  double timeend, int adjust_final, double *t_out, int64_t *nsteps_out)
{
    if (!h || !q || !dq)
        return fail(h, RV_E_ARG, "null argument");
    if (!(dt > 0))
        return fail(h, RV_E_ARG, "dt must be positive");
    const size_t bytes = (size_t)h->Np * h->ns * h->K * (h->dtype == RV_F32 ? 4 : 8);
    if (!h->qtmp) {
        hipError_t e = hipMalloc(&h->qtmp, bytes ? bytes : 1);
        if (e != hipSuccess) {
            h->qtmp = nullptr;
            return fail(h, RV_E_NOMEM, "hipMalloc of the ping-pong state failed");
        }
        h->device_bytes += (int64_t)bytes;
    }
    return lsrk_run_impl(h, q, dq, scheme, t0, dt, timeend, adjust_final, h->qtmp, t_out,
                         nsteps_out);
}

int rv_lsrk_run(rv_handle *h, void *q, void *dq, int scheme, double t0, double dt,
                double timeend, double *t_out, int64_t *nsteps_out)
{
    return rv_lsrk_solve(h, q, dq, scheme, t0, dt, timeend, 1, t_out, nsteps_out);
}

// The ping-pong state of rv_lsrk_run_noalloc: a file-scope device array of
// RV_NOALLOC_SCRATCH_BYTES.  Every GPU the library runs on holds its own copy
// at its own address until the process ends.  In host code `scratch` names a
// host-side stand-in, never the device memory: kernels and copies get the
// address hipGetSymbolAddress returns for the device.
static __device__ __align__(256) unsigned char scratch[RV_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// Device address of `scratch` on the handle's device, looked up on the first
// call on that device with the device current, then kept.
int scratch_address(rv_handle *h, void **out)
{
    if (h->device < 0 || h->device >= SCRATCH_MAX_DEVICES)
        return fail(h, RV_E_UNSUPPORTED, "device index beyond the no-alloc scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[h->device]) {
        int cur = 0;
        HIP_TRY(h, hipGetDevice(&cur));
        if (cur != h->device)
            HIP_TRY(h, hipSetDevice(h->device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != h->device)
            HIP_TRY(h, hipSetDevice(cur));
        if (e != hipSuccess || !p)
            return fail(h, RV_E_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                         hipGetErrorString(e));
        g_scratch_ptr[h->device] = p;
    }
    *out = g_scratch_ptr[h->device];
    return RV_OK;
}

}  // namespace

int rv_lsrk_run_noalloc(rv_handle *h, void *q, void *dq, int scheme, double t0, double dt,
                        double timeend, double *t_out, int64_t *nsteps_out)
{
    if (!h || !q || !dq)
        return fail(h, RV_E_ARG, "null argument");
    if (!(dt > 0))
        return fail(h, RV_E_ARG, "dt must be positive");
    const size_t bytes = (size_t)h->Np * h->ns * h->K * (h->dtype == RV_F32 ? 4 : 8);
    if (bytes > (size_t)RV_NOALLOC_SCRATCH_BYTES)
        return fail(h, RV_E_ARG,
                    "rv_lsrk_run_noalloc: a state of " + std::to_string(bytes) +
                        " bytes exceeds the " + std::to_string(RV_NOALLOC_SCRATCH_BYTES) +
                        "-byte scratch; use rv_lsrk_run");
    void *alt = nullptr;
    if (int rc = scratch_address(h, &alt))
        return rc;
    // the scratch is written in full by the first stage before it is read,
    // so nothing in it has to start at zero
    return lsrk_run_impl(h, q, dq, scheme, t0, dt, timeend, 1, alt, t_out, nsteps_out);
}

int rv_entropy_integral(rv_handle *h, const void *q, double *out)
{
    if (!h || !q || !out)
        return fail(h, RV_E_ARG, "null argument");
    if (int rc = ensure_relax(h))
        return rc;
    RelaxParams P = relax_params(h, RELAX_ETA, 0);
    P.q = const_cast<void *>(q);
    if (int rc = relax(h, P))
        return rc;
    double v[2];
    if (int rc = relax_read(h, 1, v))
        return rc;
    *out = v[0];
    return RV_OK;
}

int rv_entropy_product(rv_handle *h, const void *p, const void *q, double *out)
{
    if (!h || !p || !q || !out)
        return fail(h, RV_E_ARG, "null argument");
    if (int rc = ensure_relax(h))
        return rc;
    RelaxParams P = relax_params(h, RELAX_PRODUCT, 0);
    P.q = const_cast<void *>(p);
    P.k = const_cast<void *>(q);
    if (int rc = relax(h, P))
        return rc;
    double v[2];
    if (int rc = relax_read(h, 1, v))
        return rc;
    *out = v[1];
    return RV_OK;
}

int rv_rlsrk_step(rv_handle *h, void *q, void *dq, void *k, void *q0, int scheme, double time,
                  double dt, double *gamma, int32_t *iterations)
{
    (void)time;  // the laws here do not depend on the stage time
    if (!h || !q || !dq || !k || !q0 || !gamma)
        return fail(h, RV_E_ARG, "null argument");
    if (q == dq || q == k || q == q0 || dq == k || dq == q0 || k == q0)
        return fail(h, RV_E_ARG, "q, dq, k and q0 must be distinct arrays");
    const double *A, *B, *C;
    const int nst = lsrk_tables(scheme, &A, &B, &C);
    if (nst == 0)
        return fail(h, RV_E_ARG, "unknown LSRK scheme");
    (void)C;
    if (int rc = ensure_relax(h))
        return rc;
    const bool f32 = h->dtype == RV_F32;
    auto rnd = [f32](double v) { return f32 ? (double)(float)v : v; };
    // standard RK weights of the low-storage scheme (odesolvers.jl:96-101)
    double bfull[14];
    bfull[nst - 1] = rnd(B[nst - 1]);
    for (int i = nst - 2; i >= 0; --i)
        bfull[i] = rnd(rnd(A[i + 1]) * bfull[i + 1] + rnd(B[i]));
    const double step_dt = rnd(dt);

    // q0 .= q; eta0 = entropyintegral(q0), one pass
    RelaxParams E = relax_params(h, RELAX_ETA, 1);
    E.q = q;
    E.q0 = q0;
    if (int rc = relax(h, E))
        return rc;
    double stage_prod[14];
    double eta0;
    const void *qfinal = q;
    if (h->relax_launcher && !h->mortar) {
        // fused: each flux-differencing stage forms k = L(q) in registers,
        // dq = a dq + k and q_out = q + b dt dq, and leaves its workgroups'
        // sums of MJ v(q) . k; the stages ping-pong between q and k (k is
        // needed as an array only after the stages)
        const int64_t nwg = relax_partials_count(h);
        const int64_t cap = relax_partials_cap(h);
        if (!h->eprod)
            if (int rc = dalloc(h, &h->eprod, (size_t)14 * cap))
                return rc;
        void *cur = q, *alt = k;
        for (int st = 0; st < nst; ++st) {
            StageParams S = stage_params(h);
            S.q_in = as_global(cur);
            S.q_out = as_global(alt);
            S.dq = as_global(dq);
            S.a = rnd(A[st]);
            S.read_dq = S.a != 0.0 ? 1 : 0;
            S.b_dt = f32 ? (double)((float)rnd(B[st]) * (float)step_dt) : B[st] * step_dt;
            S.write_q = 1;
            S.eprod = as_global(h->eprod + (size_t)st * cap);
            if (int rc = launch(h, S, h->relax_launcher))
                return rc;
            std::swap(cur, alt);
        }
        qfinal = cur;
        // each stage's product: its slots summed on the device
        if (launch_slot_sums(h->eprod, nwg, nst, 1, cap, 0, 1, 1, h->relax_sums + 2 * RELAX_SLOTS,
                             h->stream) != 0) {
            hipError_t e = hipGetLastError();
            return fail(h, RV_E_HIP, std::string("entropy sum launch failed: ") +
                                         hipGetErrorString(e));
        }
        double e2[4];
        if (int rc = relax_read(h, 2, e2, nst, stage_prod))  // synchronises the stream
            return rc;
        eta0 = e2[2];
    } else {
        // stages: k = L(q); dq = a dq + k; product v(q).k into slot 2 + s; q += b dt dq
        StageParams S = stage_params(h);
        S.q_in = as_global(q);
        S.q_out = nullptr;
        S.dq = as_global(k);
        S.a = 1.0;
        S.b_dt = 0.0;
        S.read_dq = 0;
        S.write_q = 0;
        for (int st = 0; st < nst; ++st) {
            if (int rc = launch(h, S))
                return rc;
            RelaxParams U = relax_params(h, RELAX_STAGE, 2 + st);
            U.q = q;
            U.k = k;
            U.dq = dq;
            U.a = rnd(A[st]);
            U.read_dq = U.a != 0.0 ? 1 : 0;
            U.b_dt = f32 ? (double)((float)rnd(B[st]) * (float)step_dt) : B[st] * step_dt;
            if (int rc = relax(h, U))
                return rc;
        }
        double sums[2 * RELAX_SLOTS];
        if (int rc = relax_read(h, 2 + nst, sums))
            return rc;
        eta0 = sums[2];
        for (int st = 0; st < nst; ++st)
            stage_prod[st] = sums[2 * (2 + st) + 1];
    }
    double deta = -0.0;
    for (int st = 0; st < nst; ++st)
        deta += bfull[st] * step_dt * stage_prod[st];

    // gamma: root of eta(q0 + gamma k) - eta0 - gamma deta near the previous one
    const double eps = f32 ? (double)FLT_EPSILON : DBL_EPSILON;
    // (stops as the oracle's: |residual| <= 10 eps; the residual no longer
    // halves, keeping the better of the last two roots; a negligible update)
    double g = rnd(*gamma), gprev = g, Fprev = INFINITY, g_in_q = NAN;
    int it = 0;
    for (; it < RELAX_MAX_NEWTON; ++it) {
        RelaxParams N = relax_params(h, RELAX_NEWTON, 0);
        N.q = q;
        N.k = k;
        N.q0 = q0;
        N.first = it == 0;
        N.qf = qfinal;
        N.relax = g;
        if (int rc = relax(h, N))
            return rc;
        double v[2];
        if (int rc = relax_read(h, 1, v))
            return rc;
        g_in_q = g;
        const double F = v[0] - eta0 - g * deta;
        if (std::fabs(F) <= 10 * eps)
            break;
        if (std::fabs(F) > 0.5 * std::fabs(Fprev)) {
            if (std::fabs(F) > std::fabs(Fprev))
                g = gprev;
            break;
        }
        const double step = F / (v[1] - deta);
        if (!std::isfinite(step))
            return fail(h, RV_E_STATE, "relaxation Newton iteration broke down");
        if (!(std::fabs(step) > 2 * eps * std::fabs(g)))
            break;
        gprev = g;
        Fprev = F;
        g = rnd(g - step);
    }
    if (!(g == g_in_q)) {
        RelaxParams X = relax_params(h, RELAX_AXPY, 0);
        X.q = q;
        X.k = k;
        X.q0 = q0;
        X.relax = g;
        if (int rc = relax(h, X))
            return rc;
    }
    *gamma = g;
    if (iterations)
        *iterations = it;
    return RV_OK;
}

int rv_weightednorm(rv_handle *h, const void *q, double *out)
{
    if (!h || !out)
        return fail(h, RV_E_ARG, "null argument");
    int rc = reduce(h, q, 2, out);
    if (rc)
        return rc;
    for (int s = 0; s < h->ns; ++s)
        out[s] = std::sqrt(out[s]);
    return RV_OK;
}

int rv_weightednorm_p(rv_handle *h, const void *q, double p, double *out)
{
    if (!h || !out)
        return fail(h, RV_E_ARG, "null argument");
    if (std::isinf(p) && p > 0)
        return reduce(h, q, 4, out);
    if (!(p > 0) || !std::isfinite(p))
        return fail(h, RV_E_ARG, "weightednorm: p must be positive");
    int rc = reduce(h, q, 3, out, p);
    if (rc)
        return rc;
    for (int s = 0; s < h->ns; ++s)
        out[s] = std::pow(out[s], 1.0 / p);
    return RV_OK;
}

int rv_integrals(rv_handle *h, const void *q, double *out)
{
    if (!h || !out)
        return fail(h, RV_E_ARG, "null argument");
    return reduce(h, q, 1, out);
}

namespace {
// GatherScatter tables: given to rv_create, or built from the brick on first use
int ensure_gatherscatter(rv_handle *h)
{
    if (h->dtoc)
        return RV_OK;
    if (!h->is_brick && !h->forest)
        return fail(h, RV_E_STATE, "no dtoc was given to rv_create");
    std::vector<int64_t> dtoc(h->K * h->Np);
    const int64_t nc =
        h->forest ? forest_dtoc(h->forest, h->nq, dtoc.data()) : brick_dtoc(h->brick, dtoc.data());
    return build_gatherscatter(h, dtoc.data(), nc);
}
}  // namespace

int rv_gather(rv_handle *h, void *dest, const void *src, int nf)
{
    if (!h || !dest || !src || nf < 1)
        return fail(h, RV_E_ARG, "invalid argument");
    if (int rc = ensure_gatherscatter(h))
        return rc;
    const int64_t blocks = (h->nc + 255) / 256;
    if (h->dtype == RV_F64)
        hipLaunchKernelGGL(gather_kernel<double>, dim3((unsigned)std::max<int64_t>(blocks, 1)),
                           dim3(256), 0, h->stream, static_cast<double *>(dest),
                           static_cast<const double *>(src), h->ctod_ptr, h->ctod_rows, h->nc,
                           h->Np, nf);
    else
        hipLaunchKernelGGL(gather_kernel<float>, dim3((unsigned)std::max<int64_t>(blocks, 1)),
                           dim3(256), 0, h->stream, static_cast<float *>(dest),
                           static_cast<const float *>(src), h->ctod_ptr, h->ctod_rows, h->nc,
                           h->Np, nf);
    HIP_TRY(h, hipGetLastError());
    return RV_OK;
}

int rv_scatter(rv_handle *h, void *dest, const void *src, int nf)
{
    if (!h || !dest || !src || nf < 1)
        return fail(h, RV_E_ARG, "invalid argument");
    if (int rc = ensure_gatherscatter(h))
        return rc;
    const int64_t n = h->Np * h->K;
    const int64_t blocks = std::max<int64_t>((n + 255) / 256, 1);
    if (h->dtype == RV_F64)
        hipLaunchKernelGGL(scatter_kernel<double>, dim3((unsigned)blocks), dim3(256), 0, h->stream,
                           static_cast<double *>(dest), static_cast<const double *>(src), h->dtoc,
                           h->nc, h->Np, h->K, nf);
    else
        hipLaunchKernelGGL(scatter_kernel<float>, dim3((unsigned)blocks), dim3(256), 0, h->stream,
                           static_cast<float *>(dest), static_cast<const float *>(src), h->dtoc,
                           h->nc, h->Np, h->K, nf);
    HIP_TRY(h, hipGetLastError());
    return RV_OK;
}

int rv_get_maps(rv_handle *h, int64_t *vmapM, int64_t *vmapP, int64_t *mapM, int64_t *mapP)
{
    if (!h)
        return fail(nullptr, RV_E_ARG, "null handle");
    const int64_t n = 2 * h->dim * h->nfp * h->K;
    int64_t *buf = nullptr;
    HIP_TRY(h, hipMalloc(&buf, std::max<int64_t>(4 * n, 1) * sizeof(int64_t)));
    const int64_t blocks = std::max<int64_t>((n + 255) / 256, 1);
    hipLaunchKernelGGL(maps_kernel, dim3((unsigned)blocks), dim3(256), 0, h->stream, h->faces,
                       h->explicit_idx, h->dim, h->nq, h->Np, h->nfp, h->ns, h->K, buf, buf + n,
                       buf + 2 * n, buf + 3 * n);
    hipError_t e = hipGetLastError();
    std::vector<int64_t> host(4 * n);
    if (e == hipSuccess && n)
        e = hipMemcpyAsync(host.data(), buf, 4 * n * sizeof(int64_t), hipMemcpyDeviceToHost,
                           h->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(h->stream);
    (void)hipFree(buf);
    if (e != hipSuccess)
        return fail(h, RV_E_HIP, std::string("rv_get_maps: ") + hipGetErrorString(e));
    int64_t *dst[4] = {vmapM, vmapP, mapM, mapP};
    for (int k = 0; k < 4; ++k)
        if (dst[k] && n)
            std::memcpy(dst[k], host.data() + k * n, n * sizeof(int64_t));
    return RV_OK;
}

int rv_brick_arrays(const rv_brick_desc *bd, double *points, double *metrics,
                    double *surface_metrics, int64_t *vmapM, int64_t *vmapP, int64_t *bc,
                    int64_t *dtoc, int64_t *ncontinuous)
{
    BrickSpec b;
    if (int rc = brick_spec(bd, &b))
        return rc;
    brick_arrays(b, points, metrics, surface_metrics, vmapM, vmapP, bc);
    if (dtoc || ncontinuous) {
        std::vector<int64_t> tmp;
        int64_t *dst = dtoc;
        if (!dst) {
            tmp.resize(ipow(b.nq, b.dim) * brick_numcells(b));
            dst = tmp.data();
        }
        const int64_t nc = brick_dtoc(b, dst);
        if (ncontinuous)
            *ncontinuous = nc;
    }
    return RV_OK;
}

namespace {
constexpr int MND_BLOCKS = 1024;
constexpr int MND_THREADS = 256;

// min_neighbour_distance_kernel (grids.jl:94-119): per node, the distance to
// the nodes at linear index ijk -+ stride(d) inside the cell for every d in
// the mask (the linear index may leave the node's line, as there); one
// minimum per workgroup
__global__ void __launch_bounds__(MND_THREADS)
    min_node_distance_kernel(int dim, int nq, int64_t ncells, const double *points, int mask,
                             double *partial)
{
    __shared__ double sh[MND_THREADS];
    const int np = dim == 3 ? nq * nq * nq : dim == 2 ? nq * nq : nq;
    const int64_t n = (int64_t)np * ncells;
    double md = INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * MND_THREADS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * MND_THREADS) {
        const int64_t e = i / np;
        const int ijk = (int)(i - e * np);
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
                md = fmin(md, sqrt(ss));
            }
        }
    }
    sh[threadIdx.x] = md;
    __syncthreads();
    for (int w = MND_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partial[blockIdx.x] = sh[0];
}
}  // namespace

int rv_min_node_distance(int dim, int nq, int64_t ncells, const double *points,
                         int32_t dims_mask, double *out, void *stream)
{
    if (dim < 1 || dim > 3 || nq < 2 || nq > 16 || ncells < 0 || !points || !out ||
        dims_mask <= 0 || dims_mask >= (1 << dim))
        return fail(nullptr, RV_E_ARG, "rv_min_node_distance: invalid argument (dims not valid)");
    hipStream_t st = static_cast<hipStream_t>(stream);
    double *part = nullptr;
    hipError_t e = hipMallocAsync((void **)&part, sizeof(double) * MND_BLOCKS, st);
    std::vector<double> host(MND_BLOCKS);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(min_node_distance_kernel, dim3(MND_BLOCKS), dim3(MND_THREADS), 0, st,
                           dim, nq, ncells, points, (int)dims_mask, part);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(host.data(), part, sizeof(double) * MND_BLOCKS, hipMemcpyDeviceToHost,
                           st);
    if (part)
        (void)hipFreeAsync(part, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return fail(nullptr, RV_E_HIP, std::string("rv_min_node_distance: ") +
                                           hipGetErrorString(e));
    double md = INFINITY;
    for (double v : host)
        md = std::fmin(md, v);
    *out = md;
    return RV_OK;
}

int rv_curved_metrics(int dim, int nq, int64_t ncells, const double *points, const double *D,
                      const double *weights, double *metrics, double *surface_metrics,
                      void *stream)
{
    if ((dim != 2 && dim != 3) || nq < 2 || nq > 8 || ncells < 0 || !points || !D ||
        !weights || !metrics || !surface_metrics)
        return fail(nullptr, RV_E_ARG, "rv_curved_metrics: invalid argument");
    if (ncells == 0)
        return RV_OK;
    const int np = dim == 3 ? nq * nq * nq : nq * nq;
    const int nfp = dim == 3 ? nq * nq : nq;
    // D and the weights are small host arrays: stage them on the device
    double *dDw = nullptr;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e = hipMallocAsync((void **)&dDw, sizeof(double) * (nq * nq + nq), st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dDw, D, sizeof(double) * nq * nq, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dDw + nq * nq, weights, sizeof(double) * nq, hipMemcpyHostToDevice,
                           st);
    if (e == hipSuccess) {
        const int threads = (np + 63) / 64 * 64;
        hipLaunchKernelGGL(curved_volume_metrics_kernel, dim3((unsigned)ncells), dim3(threads),
                           sizeof(double) * 12 * np, st, dim, nq, points, dDw, dDw + nq * nq,
                           metrics);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        const int64_t n = (int64_t)2 * dim * nfp * ncells;
        hipLaunchKernelGGL(curved_surface_metrics_kernel, dim3((unsigned)((n + 255) / 256)),
                           dim3(256), 0, st, dim, nq, (int64_t)ncells, metrics,
                           dDw + nq * nq, surface_metrics);
        e = hipGetLastError();
    }
    if (dDw)
        (void)hipFreeAsync(dDw, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return fail(nullptr, RV_E_HIP, std::string("rv_curved_metrics: ") + hipGetErrorString(e));
    return RV_OK;
}

int rv_brick_cells(const rv_brick_desc *bd, int32_t *ijk)
{
    BrickSpec b;
    if (int rc = brick_spec(bd, &b))
        return rc;
    if (!ijk)
        return fail(nullptr, RV_E_ARG, "null output");
    const int64_t K = brick_numcells(b);
    std::vector<int32_t> tree((size_t)K * b.dim), quad((size_t)K * b.dim);
    brick_quadrants(b, tree.data(), quad.data());
    for (int64_t k = 0; k < K * b.dim; ++k)
        ijk[k] = (tree[k] << b.level) + quad[k];
    return RV_OK;
}

}  // extern "C"

namespace {

// Node coordinates of forest leaves (Np, dim, K): the brick formula of
// hexbrickpoints! / quadbrickpoints! (lobattocells.jl:1590-1674, 801-872) at
// each leaf's level, x = fma(dx, fma(h, r + 1, c / 2^30), x1) with
// h = 2^-(level+1), as the host generator computes it.
struct NodesArg {
    double r[32];
};

__global__ void forest_points_kernel(int dim, int nq, int64_t K, const int32_t *tree,
                                     const int32_t *level, const int32_t *quad,
                                     const double *verts, const int32_t *vert_off, NodesArg na,
                                     double *points)
{
    int64_t Np = nq;
    for (int d = 1; d < dim; ++d)
        Np *= nq;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Np * K)
        return;
    const int64_t e = i / Np, node = i - e * Np;
    const int lv = level[e];
    const double h = 1.0 / (double)(1 << (lv + 1));
    int64_t rem = node;
    for (int d = 0; d < dim; ++d) {
        const int t = tree[e * dim + d];
        const double *v = verts + vert_off[d];
        const double c1 = v[t], dx = v[t + 1] - v[t];
        const double cr = (double)((int64_t)quad[e * dim + d] << (30 - lv)) / (double)(1 << 30);
        points[node + Np * (d + (int64_t)dim * e)] = fma(dx, fma(h, na.r[rem % nq] + 1, cr), c1);
        rem /= nq;
    }
}

// Raven's metric arrays (Np, L, K) and surface metrics (dim+2, 2dim Nfp, K)
// into the stage kernels' general geometry: wJ dRdX (K, dim, dim, Np),
// 1 / wJ and wJ (K, Np), normals and wsJ (K, 2dim, dim+1, Nfp).
__global__ void general_geometry_kernel(int dim, int nq, int64_t K, const double *M,
                                        const double *S, double *vgeo, double *mji, double *mj,
                                        double *sgeo)
{
    int64_t Np = nq;
    for (int d = 1; d < dim; ++d)
        Np *= nq;
    const int64_t nfp = Np / nq, nf = 2 * dim, nfpt = nf * nfp;
    const int L = dim * dim + 2;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < Np * K) {
        const int64_t e = i / Np, n = i - e * Np;
        const double *me = M + Np * (int64_t)L * e;
        const double wJ = me[n + Np * (dim * dim + 1)];
        mji[i] = 1.0 / wJ;
        mj[i] = wJ;
        for (int a = 0; a < dim; ++a)
            for (int c = 0; c < dim; ++c)
                vgeo[((e * dim + a) * dim + c) * Np + n] = wJ * me[n + Np * (a + dim * c)];
    }
    if (i < nfpt * K) {
        const int64_t e = i / nfpt, r = i - e * nfpt, f = r / nfp, m = r - f * nfp;
        const double *sp = S + (dim + 2) * (f * nfp + m + nfpt * e);
        for (int c = 0; c < dim; ++c)
            sgeo[((e * nf + f) * (dim + 1) + c) * nfp + m] = sp[c];
        sgeo[((e * nf + f) * (dim + 1) + dim) * nfp + m] = sp[dim + 1];
    }
}

}  // namespace

struct rv_forest {
    Forest *f = nullptr;
    int dim = 0;
};

extern "C" {

int rv_forest_create(const rv_brick_desc *bd, rv_forest **out)
{
    if (!out)
        return fail(nullptr, RV_E_ARG, "null output");
    *out = nullptr;
    BrickSpec b;
    if (int rc = brick_spec(bd, &b))
        return rc;
    if (b.dim < 2)
        return fail(nullptr, RV_E_UNSUPPORTED, "adaptive forests need dim 2 or 3");
    rv_forest *f = new rv_forest();
    f->f = forest_new(b);
    f->dim = b.dim;
    *out = f;
    return RV_OK;
}

void rv_forest_destroy(rv_forest *f)
{
    if (f) {
        forest_free(f->f);
        delete f;
    }
}

int64_t rv_forest_numcells(const rv_forest *f)
{
    return f ? forest_size(f->f) : fail(nullptr, RV_E_ARG, "null forest");
}

int rv_forest_adapt(rv_forest *f, uint8_t *flags)
{
    if (!f || !flags)
        return fail(nullptr, RV_E_ARG, "null argument");
    if (forest_adapt(f->f, flags) != 0)
        return fail(nullptr, RV_E_UNSUPPORTED, "a leaf flagged AdaptRefine is at the finest level");
    return RV_OK;
}

int rv_forest_cells(const rv_forest *f, int32_t *tree, int32_t *level, int32_t *quad)
{
    if (!f)
        return fail(nullptr, RV_E_ARG, "null forest");
    forest_cells(f->f, tree, level, quad);
    return RV_OK;
}

int rv_forest_mortars(const rv_forest *f, int64_t *partner)
{
    if (!f || !partner)
        return fail(nullptr, RV_E_ARG, "null argument");
    forest_mortars(f->f, partner);
    return RV_OK;
}

int rv_forest_nc_counts(const rv_forest *f, int64_t *nunique, int64_t *nfaces)
{
    if (!f)
        return fail(nullptr, RV_E_ARG, "null forest");
    forest_nc_counts(f->f, nunique, nfaces);
    return RV_OK;
}

int rv_forest_nctables(const rv_forest *f, int32_t nq, int64_t *const *vmapNC,
                       int8_t *const *nctoface, int8_t *const *nctypes, int64_t *const *ncids)
{
    if (!f || nq < 2 || nq > 32)
        return fail(nullptr, RV_E_ARG, "invalid argument");
    forest_nctables(f->f, nq, vmapNC, nctoface, nctypes, ncids);
    return RV_OK;
}

int rv_mortars_from_nctables(int32_t dim, int32_t nq, int64_t ncells, const int64_t *nunique,
                             const int64_t *const *vmapNC, const int8_t *const *nctoface,
                             int32_t *face_kind, int64_t *partner)
{
    if (dim < 2 || dim > 3 || nq < 2 || nq > 32 || ncells < 0 || !nunique || !vmapNC ||
        !nctoface || !face_kind || !partner)
        return fail(nullptr, RV_E_ARG, "rv_mortars_from_nctables: invalid argument");
    const int rc = mortars_from_nctables(dim, nq, ncells, nunique, vmapNC, nctoface, face_kind,
                                         partner);
    if (rc == -1)
        return fail(nullptr, RV_E_ARG, "rv_mortars_from_nctables: a node index is out of range");
    if (rc == -2)
        return fail(nullptr, RV_E_UNSUPPORTED,
                    "rv_mortars_from_nctables: a table is not an aligned face pairing "
                    "(rotated faces are not supported)");
    return RV_OK;
}

int rv_create_forest(const rv_forest *f, const rv_brick_desc *bd, const double *points,
                     rv_handle **out)
{
    if (!out)
        return fail(nullptr, RV_E_ARG, "out is null");
    *out = nullptr;
    if (!f || !bd)
        return fail(nullptr, RV_E_ARG, "null argument");
    const BrickSpec b = forest_spec(f->f);
    const int dim = b.dim, nq = bd->nq;
    if (bd->dim != dim || nq < 2 || nq > 8)
        return fail(nullptr, RV_E_ARG, "rv_create_forest: dim must be the forest's, 2 <= Nq <= 8");
    if (bd->law != RV_LAW_ADVECTION && bd->law != RV_LAW_EULER)
        return fail(nullptr, RV_E_ARG, "unknown law");
    if (bd->dtype != RV_F64 && bd->dtype != RV_F32)
        return fail(nullptr, RV_E_ARG, "unknown dtype");
    rv_handle *h = new rv_handle();
    h->dim = dim;
    h->nq = nq;
    h->law = bd->law;
    h->flux = bd->flux ? bd->flux : RV_FLUX_RUSANOV;
    if (bd->abi_version >= 2 &&
        volume_flux_of(bd->volume_form, bd->volume_flux, bd->law, &h->vflux) != RV_OK) {
        free_handle(h);
        return fail(nullptr, RV_E_ARG, "invalid volume form or volume flux");
    }
    h->dtype = bd->dtype;
    h->K = forest_size(f->f);
    h->ns = bd->law == RV_LAW_ADVECTION ? 1 : dim + 2;
    h->Np = ipow(nq, dim);
    h->nfp = ipow(nq, dim - 1);
    h->gamma = bd->gamma;
    h->dir_mask = (1 << dim) - 1;
    for (int c = 0; c < 3; ++c)
        h->u[c] = bd->velocity[c];
    h->geo = GEO_GENERAL;
    const int64_t K = h->K, Np = h->Np, nfp = h->nfp, nf = 2 * dim, nfpt = nf * nfp;
    const int L = dim * dim + 2;
    int rc = resolve_stream(h, bd->stream);
    // host: the leaves and one record per face (conforming neighbour, or the
    // face itself: boundary faces, and hanging and covered faces as Raven's
    // self-mapped vmapP, lobattocells.jl:2072-2074)
    std::vector<int32_t> tree(K * dim), level(K), quad(K * dim);
    std::vector<int64_t> nbr(K * nf), bc(K * nf);
    std::vector<FaceRec> recs(K * nf);
    std::vector<double> x(nq), w(nq), D(nq * nq), verts;
    std::vector<int32_t> voff(3, 0);
    if (!rc) {
        forest_cells(f->f, tree.data(), level.data(), quad.data());
        forest_faces(f->f, nbr.data(), nullptr, bc.data());
        for (int64_t e = 0; e < K; ++e)
            for (int fc = 0; fc < nf; ++fc) {
                const int64_t en = nbr[e * nf + fc];
                recs[e * nf + fc] = en < 0 ? FaceRec{FACE_SELF, (int32_t)bc[e * nf + fc]}
                                           : FaceRec{(int32_t)en, fc ^ 1};
                if (en < 0)
                    h->faces_self++;
            }
        lobatto_operators(nq, x.data(), w.data(), D.data());
        for (int d = 0; d < dim; ++d) {
            voff[d] = (int32_t)verts.size();
            std::vector<double> v(b.n[d] + 1);
            brick_vertices(b, d, v.data());
            verts.insert(verts.end(), v.begin(), v.end());
        }
    }
    // device: points (unless given), curved metrics, the stage kernels'
    // geometry; temporaries are stream-ordered and freed on the stream
    hipStream_t st = h->stream;
    double *dpts = nullptr, *dM = nullptr, *dS = nullptr, *dDw = nullptr, *dverts = nullptr;
    int32_t *dcells = nullptr;
    hipError_t e = hipSuccess;
    auto amalloc = [&](void **p, size_t bytes) {
        if (e == hipSuccess)
            e = hipMallocAsync(p, std::max<size_t>(bytes, 1), st);
    };
    if (!rc) {
        amalloc((void **)&dM, sizeof(double) * Np * L * K);
        amalloc((void **)&dS, sizeof(double) * (dim + 2) * nfpt * K);
        amalloc((void **)&dDw, sizeof(double) * (nq * nq + nq));
        if (e == hipSuccess)
            e = hipMemcpyAsync(dDw, D.data(), sizeof(double) * nq * nq, hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dDw + nq * nq, w.data(), sizeof(double) * nq,
                               hipMemcpyHostToDevice, st);
        if (!points) {
            amalloc((void **)&dpts, sizeof(double) * Np * dim * K);
            amalloc((void **)&dverts, sizeof(double) * verts.size());
            amalloc((void **)&dcells, sizeof(int32_t) * (K * (2 * dim + 1) + 3));
            if (e == hipSuccess)
                e = hipMemcpyAsync(dverts, verts.data(), sizeof(double) * verts.size(),
                                   hipMemcpyHostToDevice, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(dcells, tree.data(), sizeof(int32_t) * K * dim,
                                   hipMemcpyHostToDevice, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(dcells + K * dim, quad.data(), sizeof(int32_t) * K * dim,
                                   hipMemcpyHostToDevice, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(dcells + 2 * K * dim, level.data(), sizeof(int32_t) * K,
                                   hipMemcpyHostToDevice, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(dcells + K * (2 * dim + 1), voff.data(), sizeof(int32_t) * 3,
                                   hipMemcpyHostToDevice, st);
            if (e == hipSuccess && K > 0) {
                NodesArg na{};
                for (int i = 0; i < nq; ++i)
                    na.r[i] = x[i];
                const int64_t n = Np * K;
                hipLaunchKernelGGL(forest_points_kernel, dim3((unsigned)((n + 255) / 256)),
                                   dim3(256), 0, st, dim, nq, K, dcells, dcells + 2 * K * dim,
                                   dcells + K * dim, dverts, dcells + K * (2 * dim + 1), na, dpts);
                e = hipGetLastError();
            }
        }
        const double *P = points ? points : dpts;
        if (e == hipSuccess && K > 0) {
            const int threads = (int)((Np + 63) / 64 * 64);
            hipLaunchKernelGGL(curved_volume_metrics_kernel, dim3((unsigned)K), dim3(threads),
                               sizeof(double) * 12 * Np, st, dim, nq, P, dDw, dDw + nq * nq, dM);
            e = hipGetLastError();
        }
        if (e == hipSuccess && K > 0) {
            const int64_t n = nfpt * K;
            hipLaunchKernelGGL(curved_surface_metrics_kernel, dim3((unsigned)((n + 255) / 256)),
                               dim3(256), 0, st, dim, nq, K, dM, dDw + nq * nq, dS);
            e = hipGetLastError();
        }
        if (e != hipSuccess)
            rc = fail(h, RV_E_HIP, std::string("rv_create_forest: ") + hipGetErrorString(e));
        if (!rc && !(rc = dalloc(h, &h->vgeo, (size_t)K * dim * dim * Np)) &&
            !(rc = dalloc(h, &h->mji, K * Np)) && !(rc = dalloc(h, &h->MJ, K * Np)) &&
            !(rc = dalloc(h, &h->sgeo, (size_t)K * nf * (dim + 1) * nfp)) &&
            !(rc = upload(h, &h->faces, recs.data(), recs.size())) && K > 0) {
            const int64_t n = std::max(Np, nfpt) * K;
            hipLaunchKernelGGL(general_geometry_kernel, dim3((unsigned)((n + 255) / 256)),
                               dim3(256), 0, st, dim, nq, K, dM, dS, h->vgeo, h->mji, h->MJ,
                               h->sgeo);
            e = hipGetLastError();
            if (e != hipSuccess)
                rc = fail(h, RV_E_HIP, std::string("rv_create_forest: ") + hipGetErrorString(e));
        }
    }
    for (void *p : {(void *)dpts, (void *)dM, (void *)dS, (void *)dDw, (void *)dverts,
                    (void *)dcells})
        if (p)
            (void)hipFreeAsync(p, st);
    if (!rc) {
        h->forest = forest_clone(f->f);
        rc = finish_setup(h, D.data(), w.data());
    }
    if (rc) {
        g_last_error = h->err;
        free_handle(h);
        return rc;
    }
    *out = h;
    return RV_OK;
}

int rv_forest_dtoc(const rv_forest *f, int32_t nq, int64_t *dtoc, int64_t *ncontinuous)
{
    if (!f || nq < 2 || nq > 32 || !dtoc)
        return fail(nullptr, RV_E_ARG, "invalid argument");
    const int64_t n = forest_dtoc(f->f, nq, dtoc);
    if (ncontinuous)
        *ncontinuous = n;
    return RV_OK;
}

int rv_lobatto_halves(int nq, double *lower, double *upper)
{
    if (nq < 2 || nq > 32 || !lower || !upper)
        return fail(nullptr, RV_E_ARG, "invalid argument");
    lobatto_halves(nq, lower, upper);
    return RV_OK;
}

int rv_forest_arrays(const rv_forest *f, int32_t nq, double *points, int64_t *vmapM,
                     int64_t *vmapP, int64_t *bc, int32_t *face_kind)
{
    if (!f || nq < 2 || nq > 32)
        return fail(nullptr, RV_E_ARG, "invalid argument");
    forest_arrays(f->f, nq, points, vmapM, vmapP, bc, face_kind);
    return RV_OK;
}

int rv_lobatto(int nq, double *points, double *weights, double *D)
{
    if (nq < 2 || nq > 32 || !points || !weights || !D)
        return fail(nullptr, RV_E_ARG, "invalid argument");
    lobatto_operators(nq, points, weights, D);
    return RV_OK;
}

}  // extern "C"

// ===========================================================================
// Batches: many operators, one persistent launch per stage (batch.hip).
// ===========================================================================
struct rv_batch {
    static constexpr int RING = 4;   // parameter tables in flight (rv_batch_stage / _step)
    static constexpr int MAXST = 14;  // stages of the longest scheme (LSRK144)
    // a kind with at least this many items (waves) fills the device on its
    // own: 256 CUs x 4 waves
    static constexpr int64_t OWN_LAUNCH_ITEMS = 1024;
    int64_t own_items = OWN_LAUNCH_ITEMS;  // rv_batch_configure may change it
    std::string err;
    int dim = 0, wpb = 2, slab_bytes = 0;
    int64_t nitems = 0, nwaves = 0;
    hipStream_t stream = nullptr;          // the first member's stream
    hipStream_t capture = nullptr;         // private stream graphs are captured on
    // Launch chains (a kind's stages, or the work-list launch's) touch
    // disjoint members and are spread over `lanes` streams: lane 0 is the
    // calling stream, lanes 1.. private streams forked from it and joined back
    // with events (captured step graphs get parallel branches the same way).
    static constexpr int MAXLANES = 4;
    int lanes = 2;
    hipStream_t lane[MAXLANES] = {};       // [0] unused
    hipEvent_t fork_ev = nullptr;
    hipEvent_t join_ev[MAXLANES] = {};
    std::vector<rv_handle *> members;
    BatchItem *items = nullptr;            // device work list
    StageParams *host_ring = nullptr;      // pinned, RING * n
    StageParams *dev_ring = nullptr;       // device, RING * n
    hipEvent_t ring_ev[RING] = {};
    int ring = 0;
    // rv_batch_lsrk_run: ping-pong states, stage tables and step graphs
    std::vector<void *> qtmp;
    StageParams *dev_tables = nullptr;
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    std::vector<const void *> graph_key;   // pointers, scheme and dt the graphs were built for
    int64_t device_bytes = 0;
    // Kinds with enough items to fill the device on their own (every 3-D
    // kind: one kernel holding all 3-D kinds would run out of registers)
    // launch once per kind, with the kind's own slab and the single
    // operator's launch shape.  The remaining small kinds share one mixed
    // persistent launch over the tail of the work list [tail_first, nitems).
    // Members are sorted by (small kind, kind), so a kind's members follow
    // each other and the items of the members still stepping form a few
    // contiguous runs.
    struct KindRange {
        int kind, slab_bytes, wpb;
        int mbegin, mend;                  // positions in `order`
    };
    std::vector<KindRange> kind_ranges;
    int64_t tail_first = 0;                // first item of the mixed launch
    int tail_mbegin = 0;                   // position in `order` of its first member
    std::vector<int> order;                // members sorted by (small kind, kind)
    std::vector<int64_t> mfirst, mcount;   // each member's items
    // each member's kind, and cells and LDS bytes per item in the work-list
    // launch (epw, slab) and when its kind launches on its own (epw_own, slab_own)
    std::vector<int> kind, epw, slab, epw_own, slab_own;
};

namespace {

int bfail(rv_batch *b, int code, const std::string &msg)
{
    if (b)
        b->err = msg;
    g_last_error = msg;
    return code;
}

void drop_graphs(rv_batch *b)
{
    for (auto &e : b->exec)
        if (e) {
            (void)hipGraphExecDestroy(e);
            e = nullptr;
        }
    b->graph_key.clear();
}

void free_batch(rv_batch *b)
{
    if (!b)
        return;
    if (b->stream)
        (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    for (void *p : b->qtmp)
        if (p)
            (void)hipFree(p);
    void *bufs[] = {b->items, b->dev_ring, b->dev_tables};
    for (void *p : bufs)
        if (p)
            (void)hipFree(p);
    if (b->host_ring)
        (void)hipHostFree(b->host_ring);
    for (auto &ev : b->ring_ev)
        if (ev)
            (void)hipEventDestroy(ev);
    if (b->capture)
        (void)hipStreamDestroy(b->capture);
    for (int l = 1; l < rv_batch::MAXLANES; ++l) {
        if (b->lane[l]) {
            (void)hipStreamSynchronize(b->lane[l]);
            (void)hipStreamDestroy(b->lane[l]);
        }
        if (b->join_ev[l])
            (void)hipEventDestroy(b->join_ev[l]);
    }
    if (b->fork_ev)
        (void)hipEventDestroy(b->fork_ev);
    delete b;
}

#define BHIP_TRY(b, expr)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return bfail(b, RV_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_));    \
    } while (0)

// The work list and launch layout of a batch from its members' kinds: members
// sorted by (small kind, kind), each member's cells in order; kinds with at
// least own_items items launch on their own, the rest form the mixed tail.
void batch_layout(rv_batch *b, std::vector<BatchItem> &items)
{
    const int n = (int)b->members.size();
    const std::vector<int> &kind = b->kind, &slab = b->slab;
    std::map<int, int64_t> kind_items;
    for (int m = 0; m < n; ++m)
        kind_items[kind[m]] += (b->members[m]->K + b->epw[m] - 1) / b->epw[m];
    std::vector<char> small(n);
    for (int m = 0; m < n; ++m)
        small[m] = b->dim != 3 && kind_items[kind[m]] < b->own_items;
    std::vector<int> &order = b->order;
    order.resize(n);
    for (int m = 0; m < n; ++m)
        order[m] = m;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return small[x] != small[y] ? small[x] < small[y] : kind[x] < kind[y];
    });
    items.clear();
    b->mfirst.assign(n, 0);
    b->mcount.assign(n, 0);
    b->tail_mbegin = n;
    b->tail_first = -1;
    b->slab_bytes = 16;
    b->kind_ranges.clear();
    for (int p = 0; p < n; ++p) {
        const int m = order[p];
        if (small[m] && b->tail_first < 0) {
            b->tail_first = (int64_t)items.size();
            b->tail_mbegin = p;
        }
        if (small[m])
            b->slab_bytes = std::max(b->slab_bytes, slab[m]);
        b->mfirst[m] = (int64_t)items.size();
        // a kind launching on its own takes its own item size (packed
        // workgroups for 2-D Nq = 5, 6, 7)
        const int ep = small[m] ? b->epw[m] : b->epw_own[m];
        for (int64_t c = 0; c < b->members[m]->K; c += ep)
            items.push_back(BatchItem{m, kind[m], c});
        b->mcount[m] = (int64_t)items.size() - b->mfirst[m];
    }
    b->nitems = (int64_t)items.size();
    if (b->tail_first < 0)
        b->tail_first = b->nitems;
    // a kind launching on its own takes its first member's launch shape (the
    // single operator's, autotuned or default)
    for (int p = 0; p < b->tail_mbegin; ++p) {
        const int m = order[p];
        if (p == 0 || kind[m] != kind[order[p - 1]])
            b->kind_ranges.push_back({kind[m], b->slab_own[m], b->members[m]->wpb, p, p + 1});
        else
            b->kind_ranges.back().mend = p + 1;
    }
}

// persistent waves of the mixed launch: what the device holds at once
int batch_waves(rv_batch *b)
{
    int64_t resident = 0;
    if (batch_resident_waves(b->dim, b->wpb, b->slab_bytes, &resident) != 0 || resident <= 0)
        return RV_E_HIP;
    b->nwaves = std::min<int64_t>(resident, std::max<int64_t>(b->nitems - b->tail_first, 1));
    return RV_OK;
}

// Stages 0..nst-1 of the batch, stage s's table at table + s * n.  active
// (optional) marks the members that step.  Kind by kind, all stages of the
// kind's stepping members run back to back (a member's state stays in the
// caches from one stage to the next, as when it runs alone); the work-list
// launch runs each stage once, or not at all when none of its members steps
// (the others' items return at once, their K being 0).  These chains touch
// disjoint members: they go to the batch's lanes, heaviest first, each to
// the lane with the fewest bytes so far, so that one kind's launches overlap
// another's ramp and tail.
int batch_launch(rv_batch *b, const StageParams *table, int nst, hipStream_t st,
                 const std::vector<char> *active = nullptr)
{
    const size_t n = b->members.size();
    struct Chain {
        int range;                                      // kind_ranges index, -1: work list
        std::vector<std::pair<int64_t, int64_t>> runs;  // [first item, count)
        double bytes;
    };
    auto member_bytes = [&](int m) {
        const rv_handle *h = b->members[m];
        return (double)h->Np * h->ns * h->K * (h->dtype == RV_F32 ? 4 : 8);
    };
    std::vector<Chain> chains;
    for (size_t ri = 0; ri < b->kind_ranges.size(); ++ri) {
        const auto &r = b->kind_ranges[ri];
        Chain c{(int)ri, {}, 0.0};
        // runs of consecutive stepping members
        int64_t first = -1, count = 0;
        for (int p = r.mbegin; p <= r.mend; ++p) {
            const int m = p < r.mend ? b->order[p] : -1;
            if (m < 0 || (active && !(*active)[m])) {
                if (count > 0)
                    c.runs.push_back({first, count});
                first = -1, count = 0;
                continue;
            }
            if (first < 0)
                first = b->mfirst[m];
            count += b->mcount[m];
            c.bytes += member_bytes(m);
        }
        if (!c.runs.empty())
            chains.push_back(std::move(c));
    }
    const int64_t ntail = b->nitems - b->tail_first;
    if (ntail > 0) {
        Chain c{-1, {}, 0.0};
        bool any = false;
        for (int p = b->tail_mbegin; p < (int)n; ++p) {
            const int m = b->order[p];
            if (!active || (*active)[m]) {
                any = true;
                c.bytes += member_bytes(m);
            }
        }
        if (any)
            chains.push_back(std::move(c));
    }
    std::stable_sort(chains.begin(), chains.end(),
                     [](const Chain &x, const Chain &y) { return x.bytes > y.bytes; });
    const int nl = std::max(1, std::min<int>(b->lanes, (int)chains.size()));
    std::vector<double> load(nl, 0.0);
    std::vector<int> lane_of(chains.size());
    for (size_t c = 0; c < chains.size(); ++c) {
        const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        lane_of[c] = l;
        load[l] += chains[c].bytes;
    }
    if (nl > 1) {
        BHIP_TRY(b, hipEventRecord(b->fork_ev, st));
        for (int l = 1; l < nl; ++l)
            BHIP_TRY(b, hipStreamWaitEvent(b->lane[l], b->fork_ev, 0));
    }
    int rc = 0;
    for (size_t c = 0; c < chains.size() && rc == 0; ++c) {
        hipStream_t ls = lane_of[c] == 0 ? st : b->lane[lane_of[c]];
        if (chains[c].range >= 0) {
            const auto &r = b->kind_ranges[chains[c].range];
            for (int s = 0; s < nst && rc == 0; ++s)
                for (const auto &run : chains[c].runs)
                    if ((rc = launch_batch_kind(b->dim, r.kind, table + s * n, b->items,
                                                run.first, run.second, r.slab_bytes, r.wpb, ls)))
                        break;
        } else {
            for (int s = 0; s < nst && rc == 0; ++s)
                rc = launch_batch(b->dim, table + s * n, b->items + b->tail_first, ntail,
                                  b->nwaves, b->slab_bytes, b->wpb, ls);
        }
    }
    const hipError_t launch_err = rc ? hipGetLastError() : hipSuccess;
    // the lanes join back also after a failed launch (a capture must end joined)
    for (int l = 1; l < nl; ++l) {
        BHIP_TRY(b, hipEventRecord(b->join_ev[l], b->lane[l]));
        BHIP_TRY(b, hipStreamWaitEvent(st, b->join_ev[l], 0));
    }
    if (rc)
        return bfail(b, RV_E_HIP, std::string("batch launch failed: ") +
                                      hipGetErrorString(launch_err));
    return RV_OK;
}

}  // namespace

extern "C" {

int rv_batch_create(rv_handle *const *members, int32_t n, rv_batch **out)
{
    if (!out)
        return bfail(nullptr, RV_E_ARG, "null output");
    *out = nullptr;
    if (!members || n <= 0)
        return bfail(nullptr, RV_E_ARG, "a batch needs at least one member");
    rv_batch *b = new rv_batch();
    b->members.assign(members, members + n);
    std::vector<int> kind(n), epw(n), slab(n), epw_own(n), slab_own(n);
    for (int m = 0; m < n; ++m) {
        rv_handle *h = members[m];
        if (!h) {
            free_batch(b);
            return bfail(nullptr, RV_E_ARG, "null member");
        }
        if (m == 0) {
            b->dim = h->dim;
            b->stream = h->stream;
        }
        if (h->dim != b->dim || h->stream != b->stream) {
            free_batch(b);
            return bfail(nullptr, RV_E_ARG, "batch members must share dim and stream");
        }
        if (h->geo != GEO_AFFINE || h->flux != RV_FLUX_RUSANOV || h->vflux != 0 ||
            h->dir_mask != (1 << h->dim) - 1) {
            free_batch(b);
            return bfail(nullptr, RV_E_UNSUPPORTED,
                         "batch members need affine geometry, the weak form, the Rusanov flux "
                         "and all directions");
        }
        int ep[2] = {0, 0}, sb[2] = {0, 0};
        if (!batch_kind(h->dim, h->nq, h->law, h->dtype == RV_F32 ? 4 : 8, &kind[m], ep, sb)) {
            free_batch(b);
            return bfail(nullptr, RV_E_UNSUPPORTED, "no batch kernel for this (dim, Nq)");
        }
        epw[m] = ep[0];
        slab[m] = (sb[0] + 15) / 16 * 16;
        epw_own[m] = ep[1];
        slab_own[m] = (sb[1] + 15) / 16 * 16;
    }
    b->kind = kind;
    b->epw = epw;
    b->epw_own = epw_own;
    b->slab_own = slab_own;
    b->slab = slab;
    std::vector<BatchItem> items;
    batch_layout(b, items);
    auto bail = [&](const char *what, hipError_t e) {
        free_batch(b);
        return bfail(nullptr, RV_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    // room for the longest work list any layout gives: every member in
    // work-list items (the smallest items)
    size_t cap = 1;
    for (int m = 0; m < n; ++m)
        cap += (size_t)((members[m]->K + epw[m] - 1) / epw[m]);
    if ((e = hipMalloc(&b->items, cap * sizeof(BatchItem))))
        return bail("hipMalloc", e);
    if ((e = hipMemcpy(b->items, items.data(), items.size() * sizeof(BatchItem),
                       hipMemcpyHostToDevice)))
        return bail("hipMemcpy", e);
    const size_t ring_entries = (size_t)n * rv_batch::MAXST * rv_batch::RING;
    if ((e = hipHostMalloc((void **)&b->host_ring, sizeof(StageParams) * ring_entries)))
        return bail("hipHostMalloc", e);
    if ((e = hipMalloc(&b->dev_ring, sizeof(StageParams) * ring_entries)))
        return bail("hipMalloc", e);
    for (auto &ev : b->ring_ev)
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)))
            return bail("hipEventCreate", e);
    if ((e = hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming)))
        return bail("hipEventCreate", e);
    for (int l = 1; l < rv_batch::MAXLANES; ++l) {
        if ((e = hipStreamCreateWithFlags(&b->lane[l], hipStreamNonBlocking)))
            return bail("hipStreamCreate", e);
        if ((e = hipEventCreateWithFlags(&b->join_ev[l], hipEventDisableTiming)))
            return bail("hipEventCreate", e);
    }
    b->device_bytes = (int64_t)(cap * sizeof(BatchItem) + sizeof(StageParams) * ring_entries);
    if (batch_waves(b) != RV_OK) {
        free_batch(b);
        return bfail(nullptr, RV_E_HIP, "occupancy query of the batch kernel failed");
    }
    *out = b;
    return RV_OK;
}

void rv_batch_destroy(rv_batch *b) { free_batch(b); }

const char *rv_batch_last_error(const rv_batch *b)
{
    return b ? b->err.c_str() : g_last_error.c_str();
}

int rv_batch_info(const rv_batch *b, int64_t *nitems, int64_t *nwaves, int32_t *slab_bytes)
{
    if (!b)
        return bfail(nullptr, RV_E_ARG, "null batch");
    if (nitems)
        *nitems = b->nitems;
    if (nwaves)
        *nwaves = b->nwaves;
    if (slab_bytes)
        *slab_bytes = b->slab_bytes;
    return RV_OK;
}

int rv_batch_configure(rv_batch *b, int mixed_waves_per_block, int64_t own_launch_items)
{
    if (!b)
        return bfail(nullptr, RV_E_ARG, "null batch");
    if (mixed_waves_per_block != 1 && mixed_waves_per_block != 2 && mixed_waves_per_block != 4)
        return bfail(b, RV_E_ARG, "waves per block must be 1, 2 or 4");
    if (own_launch_items < 1)
        return bfail(b, RV_E_ARG, "own_launch_items must be positive");
    // kernels in flight read the work list and the cached graphs hold the old
    // launches: wait for them, then lay the batch out again
    BHIP_TRY(b, hipStreamSynchronize(b->stream));
    drop_graphs(b);
    b->wpb = mixed_waves_per_block;
    b->own_items = own_launch_items;
    std::vector<BatchItem> items;
    batch_layout(b, items);
    BHIP_TRY(b, hipMemcpy(b->items, items.data(), items.size() * sizeof(BatchItem),
                          hipMemcpyHostToDevice));
    if (batch_waves(b) != RV_OK)
        return bfail(b, RV_E_HIP, "occupancy query of the batch kernel failed");
    return RV_OK;
}

int rv_batch_set_streams(rv_batch *b, int32_t nstreams)
{
    if (!b)
        return bfail(nullptr, RV_E_ARG, "null batch");
    if (nstreams < 1 || nstreams > rv_batch::MAXLANES)
        return bfail(b, RV_E_ARG, "streams must be 1 to 4");
    BHIP_TRY(b, hipStreamSynchronize(b->stream));
    drop_graphs(b);
    b->lanes = nstreams;
    return RV_OK;
}

int rv_batch_stage(rv_batch *b, void *const *q_out, const void *const *q_in, void *const *dq,
                   const double *a, const double *b_dt)
{
    if (!b || !q_out || !q_in || !dq || !a || !b_dt)
        return bfail(b, RV_E_ARG, "null argument");
    const int n = (int)b->members.size();
    const int slot = b->ring;
    StageParams *host = b->host_ring + (size_t)slot * n * rv_batch::MAXST;
    StageParams *dev = b->dev_ring + (size_t)slot * n * rv_batch::MAXST;
    // the pinned slot is rewritten only after its previous upload completed
    BHIP_TRY(b, hipEventSynchronize(b->ring_ev[slot]));
    for (int m = 0; m < n; ++m) {
        if (!q_out[m] || !q_in[m] || !dq[m])
            return bfail(b, RV_E_ARG, "null state pointer");
        if (q_out[m] == q_in[m])
            return bfail(b, RV_E_ARG, "q_out must not alias q_in");
        StageParams P = stage_params(b->members[m]);
        P.q_in = as_global(q_in[m]);
        P.q_out = as_global(q_out[m]);
        P.dq = as_global(dq[m]);
        P.a = a[m];
        P.b_dt = b_dt[m];
        P.read_dq = a[m] != 0.0 ? 1 : 0;
        P.write_q = 1;
        host[m] = P;
    }
    BHIP_TRY(b, hipMemcpyAsync(dev, host, sizeof(StageParams) * n, hipMemcpyHostToDevice,
                               b->stream));
    BHIP_TRY(b, hipEventRecord(b->ring_ev[slot], b->stream));
    b->ring = (slot + 1) % rv_batch::RING;
    return batch_launch(b, dev, 1, b->stream);
}

int rv_batch_step(rv_batch *b, void *const *q0, void *const *q1, void *const *dq,
                  const int32_t *cur, const double *dt, int scheme)
{
    if (!b || !q0 || !q1 || !dq || !cur || !dt)
        return bfail(b, RV_E_ARG, "null argument");
    const double *A, *B, *C;
    const int nst = lsrk_tables(scheme, &A, &B, &C);
    if (nst == 0 || nst > rv_batch::MAXST)
        return bfail(b, RV_E_ARG, "unknown LSRK scheme");
    (void)C;
    const int n = (int)b->members.size();
    const int slot = b->ring;
    StageParams *host = b->host_ring + (size_t)slot * n * rv_batch::MAXST;
    StageParams *dev = b->dev_ring + (size_t)slot * n * rv_batch::MAXST;
    BHIP_TRY(b, hipEventSynchronize(b->ring_ev[slot]));
    std::vector<char> active(n);
    for (int m = 0; m < n; ++m) {
        rv_handle *h = b->members[m];
        const bool f32 = h->dtype == RV_F32;
        const bool on = dt[m] != 0.0;
        active[m] = on;
        if (on && (!q0[m] || !q1[m] || !dq[m] || q0[m] == q1[m] || (cur[m] & ~1)))
            return bfail(b, RV_E_ARG, "invalid state pointers or parity");
        StageParams base = stage_params(h);
        if (!on)
            base.K = 0;  // the member's items do nothing this step
        for (int s = 0; s < nst; ++s) {
            StageParams P = base;
            void *bufs[2] = {q0[m], q1[m]};
            const int c = (cur[m] + s) & 1;
            P.q_in = as_global(bufs[c]);
            P.q_out = as_global(bufs[1 - c]);
            P.dq = as_global(dq[m]);
            // LSRK{FT}: coefficients and dt in the state type (odesolvers.jl:41-48)
            P.a = f32 ? (double)(float)A[s] : A[s];
            P.b_dt = f32 ? (double)((float)B[s] * (float)dt[m]) : B[s] * dt[m];
            P.read_dq = P.a != 0.0 ? 1 : 0;
            P.write_q = 1;
            host[(size_t)s * n + m] = P;
        }
    }
    BHIP_TRY(b, hipMemcpyAsync(dev, host, sizeof(StageParams) * n * nst, hipMemcpyHostToDevice,
                               b->stream));
    BHIP_TRY(b, hipEventRecord(b->ring_ev[slot], b->stream));
    b->ring = (slot + 1) % rv_batch::RING;
    return batch_launch(b, dev, nst, b->stream, &active);
}

int rv_batch_steps(rv_batch *b, void *const *q0, void *const *q1, void *const *dq,
                   const int32_t *cur, const double *dt, int scheme, int64_t nsteps)
{
    if (!b || !cur || nsteps < 0)
        return bfail(b, RV_E_ARG, "invalid argument");
    const double *A, *B, *C;
    const int nst = lsrk_tables(scheme, &A, &B, &C);
    if (nst == 0)
        return bfail(b, RV_E_ARG, "unknown LSRK scheme");
    std::vector<int32_t> c(cur, cur + b->members.size());
    for (int64_t k = 0; k < nsteps; ++k) {
        if (int rc = rv_batch_step(b, q0, q1, dq, c.data(), dt, scheme))
            return rc;
        for (auto &x : c)
            x = (x + nst) & 1;
    }
    return RV_OK;
}

int rv_batch_lsrk_run(rv_batch *b, void *const *q, void *const *dq, int scheme, const double *dt,
                      int64_t nsteps)
{
    if (!b || !q || !dq || !dt || nsteps < 0)
        return bfail(b, RV_E_ARG, "invalid argument");
    const double *A, *B, *C;
    const int nst = lsrk_tables(scheme, &A, &B, &C);
    if (nst == 0)
        return bfail(b, RV_E_ARG, "unknown LSRK scheme");
    (void)C;
    const int n = (int)b->members.size();
    for (int m = 0; m < n; ++m)
        if (!q[m] || !dq[m])
            return bfail(b, RV_E_ARG, "null state pointer");
    if (b->qtmp.empty()) {
        b->qtmp.assign(n, nullptr);
        for (int m = 0; m < n; ++m) {
            rv_handle *h = b->members[m];
            const size_t bytes = (size_t)h->Np * h->ns * h->K * (h->dtype == RV_F32 ? 4 : 8);
            BHIP_TRY(b, hipMalloc(&b->qtmp[m], bytes ? bytes : 1));
            b->device_bytes += (int64_t)bytes;
        }
    }
    // the graphs bake pointers and coefficients: rebuild when any changed
    std::vector<const void *> key;
    key.push_back(reinterpret_cast<const void *>((intptr_t)scheme));
    for (int m = 0; m < n; ++m) {
        key.push_back(q[m]);
        key.push_back(dq[m]);
        int64_t bits;
        std::memcpy(&bits, &dt[m], sizeof bits);
        key.push_back(reinterpret_cast<const void *>((intptr_t)bits));
    }
    const int parities = (nst % 2) ? 2 : 1;
    if (!b->capture)
        BHIP_TRY(b, hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking));
    if (key != b->graph_key) {
        drop_graphs(b);
        if (b->dev_tables) {
            (void)hipFree(b->dev_tables);
            b->dev_tables = nullptr;
        }
        std::vector<StageParams> tab((size_t)parities * nst * n);
        for (int p = 0; p < parities; ++p)
            for (int s = 0; s < nst; ++s)
                for (int m = 0; m < n; ++m) {
                    rv_handle *h = b->members[m];
                    const bool f32 = h->dtype == RV_F32;
                    void *bufs[2] = {q[m], b->qtmp[m]};
                    const int cur = (p + s) % 2;
                    StageParams P = stage_params(h);
                    P.q_in = as_global(bufs[cur]);
                    P.q_out = as_global(bufs[1 - cur]);
                    P.dq = as_global(dq[m]);
                    // coefficients rounded to the state type as in rv_lsrk_run
                    P.a = f32 ? (double)(float)A[s] : A[s];
                    P.b_dt = f32 ? (double)((float)B[s] * (float)dt[m]) : B[s] * dt[m];
                    P.read_dq = P.a != 0.0 ? 1 : 0;
                    P.write_q = 1;
                    tab[((size_t)p * nst + s) * n + m] = P;
                }
        BHIP_TRY(b, hipMalloc(&b->dev_tables, tab.size() * sizeof(StageParams)));
        BHIP_TRY(b, hipMemcpy(b->dev_tables, tab.data(), tab.size() * sizeof(StageParams),
                              hipMemcpyHostToDevice));
        for (int p = 0; p < parities; ++p) {
            hipGraph_t g = nullptr;
            BHIP_TRY(b, hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal));
            int rc = batch_launch(b, b->dev_tables + (size_t)p * nst * n, nst, b->capture);
            hipError_t ec = hipStreamEndCapture(b->capture, &g);
            if (rc != RV_OK) {
                if (g)
                    (void)hipGraphDestroy(g);
                return rc;
            }
            if (ec != hipSuccess)
                return bfail(b, RV_E_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
            hipError_t ei = hipGraphInstantiate(&b->exec[p], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess)
                return bfail(b, RV_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
        }
        b->graph_key = key;
    }
    // LSRK's dq starts at zero (odesolvers.jl:43-44).  A first stage with
    // a = 0 (LSRK54, LSRK144) writes dq = L(q) without reading it, which is
    // what dq .*= 0 followed by the increment gives from that zero, so the
    // clear is needed only when a scheme's first a is not 0.  It stays outside
    // the graphs.
    for (int m = 0; m < n && A[0] != 0.0; ++m) {
        rv_handle *h = b->members[m];
        const size_t bytes = (size_t)h->Np * h->ns * h->K * (h->dtype == RV_F32 ? 4 : 8);
        BHIP_TRY(b, hipMemsetAsync(dq[m], 0, bytes, b->stream));
    }
    int cur = 0;  // 0: state in q, 1: in qtmp
    for (int64_t step = 0; step < nsteps; ++step) {
        BHIP_TRY(b, hipGraphLaunch(b->exec[parities == 2 ? cur : 0], b->stream));
        cur = (cur + nst) % 2;
    }
    if (cur == 1)
        for (int m = 0; m < n; ++m) {
            rv_handle *h = b->members[m];
            const size_t bytes = (size_t)h->Np * h->ns * h->K * (h->dtype == RV_F32 ? 4 : 8);
            BHIP_TRY(b, hipMemcpyAsync(q[m], b->qtmp[m], bytes, hipMemcpyDeviceToDevice,
                                       b->stream));
        }
    return RV_OK;
}

}  // extern "C"
