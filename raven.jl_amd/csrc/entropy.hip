// Entropy passes of the relaxation low-storage Runge-Kutta step
// (RLSRK, odesolvers.jl:72-151) and the entropy diagnostics
// entropyintegral / entropyproduct (measures.jl:23-33).
//
// Every pass is one sweep over the state: HBM-bound elementwise work with a
// Neumaier-compensated sum per workgroup, in a fixed order (contiguous node
// range per workgroup, fixed thread stride, fixed tree), so the host's sum of
// the workgroup partials is reproducible run to run.  The per-node entropy
// arithmetic runs in FP64 for both state precisions.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "entropy.hpp"

namespace rv {

namespace {

struct CSum {
    double s, c;
};

__device__ inline CSum cadd(CSum a, double x)
{
    const double t = a.s + x;
    if (fabs(a.s) >= fabs(x))
        a.c += (a.s - t) + x;
    else
        a.c += (x - t) + a.s;
    a.s = t;
    return a;
}

__device__ inline CSum cmerge(CSum a, CSum b)
{
    CSum r = cadd(a, b.s);
    r.c += b.c;
    return r;
}

template <typename R, int LAW, int DIM, int MODE>
__global__ void __launch_bounds__(RELAX_THREADS) relax_kernel(const RelaxParams P)
{
    constexpr int NS = LAW == LAW_ADVECTION ? 1 : DIM + 2;
    const int64_t Np = P.Np, n = P.Np * P.K;
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * chunk;
    const int64_t hi = lo + chunk < n ? lo + chunk : n;
    R *q = static_cast<R *>(P.q);
    R *k = static_cast<R *>(P.k);
    R *dq = static_cast<R *>(P.dq);
    R *q0 = static_cast<R *>(P.q0);
    const R *qf = static_cast<const R *>(P.qf);
    const R a = R(P.a), bdt = R(P.b_dt), gam = R(P.relax);
    const double g = P.gas_gamma;
    CSum eta{0.0, 0.0}, prod{0.0, 0.0};
    for (int64_t i = lo + threadIdx.x; i < hi; i += RELAX_THREADS) {
        const int64_t e = i / Np, node = i - e * Np;
        const int64_t base = node + Np * NS * e;
        double x[NS], kk[NS];
        if constexpr (MODE == RELAX_ETA) {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const R v = q[base + Np * s];
                if (q0)
                    q0[base + Np * s] = v;
                x[s] = (double)v;
            }
            eta = cadd(eta, P.MJ[i] * law_entropy<LAW, DIM>(x, g));
        } else if constexpr (MODE == RELAX_PRODUCT) {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                x[s] = (double)q[base + Np * s];
                kk[s] = (double)k[base + Np * s];
            }
            prod = cadd(prod, P.MJ[i] * law_entropy_dot<LAW, DIM>(x, kk, g));
        } else if constexpr (MODE == RELAX_STAGE) {
            // dq .*= rka; rhs!(k, q); dq .+= k; deta term at q; q += rkb dt dq
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const int64_t o = base + Np * s;
                const R qs = q[o], ks = k[o];
                const R d = P.read_dq ? a * dq[o] + ks : ks;
                dq[o] = d;
                q[o] = qs + bdt * d;
                x[s] = (double)qs;
                kk[s] = (double)ks;
            }
            prod = cadd(prod, P.MJ[i] * law_entropy_dot<LAW, DIM>(x, kk, g));
        } else if constexpr (MODE == RELAX_NEWTON) {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const int64_t o = base + Np * s;
                const R z = q0[o];
                R ks;
                if (P.first) {
                    ks = qf[o] - z;
                    k[o] = ks;
                } else {
                    ks = k[o];
                }
                const R xs = z + gam * ks;
                q[o] = xs;
                x[s] = (double)xs;
                kk[s] = (double)ks;
            }
            const double mj = P.MJ[i];
            double et, dt;
            law_entropy_and_dot<LAW, DIM>(x, kk, g, et, dt);
            eta = cadd(eta, mj * et);
            prod = cadd(prod, mj * dt);
        } else {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const int64_t o = base + Np * s;
                q[o] = q0[o] + gam * k[o];
            }
        }
    }
    if constexpr (MODE != RELAX_AXPY) {
        __shared__ CSum sh[2][RELAX_THREADS];
        sh[0][threadIdx.x] = eta;
        sh[1][threadIdx.x] = prod;
        __syncthreads();
        for (int w = RELAX_THREADS / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) {
                sh[0][threadIdx.x] = cmerge(sh[0][threadIdx.x], sh[0][threadIdx.x + w]);
                sh[1][threadIdx.x] = cmerge(sh[1][threadIdx.x], sh[1][threadIdx.x + w]);
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            double *out = P.partial + 4 * (int64_t)blockIdx.x;
            out[0] = sh[0][0].s;
            out[1] = sh[0][0].c;
            out[2] = sh[1][0].s;
            out[3] = sh[1][0].c;
        }
    }
}

template <typename R, int LAW, int DIM>
int launch_mode(const RelaxParams &P, hipStream_t st)
{
    const dim3 grid(RELAX_BLOCKS), block(RELAX_THREADS);
    switch (P.mode) {
    case RELAX_ETA:
        hipLaunchKernelGGL((relax_kernel<R, LAW, DIM, RELAX_ETA>), grid, block, 0, st, P);
        break;
    case RELAX_PRODUCT:
        hipLaunchKernelGGL((relax_kernel<R, LAW, DIM, RELAX_PRODUCT>), grid, block, 0, st, P);
        break;
    case RELAX_STAGE:
        hipLaunchKernelGGL((relax_kernel<R, LAW, DIM, RELAX_STAGE>), grid, block, 0, st, P);
        break;
    case RELAX_NEWTON:
        hipLaunchKernelGGL((relax_kernel<R, LAW, DIM, RELAX_NEWTON>), grid, block, 0, st, P);
        break;
    case RELAX_AXPY:
        hipLaunchKernelGGL((relax_kernel<R, LAW, DIM, RELAX_AXPY>), grid, block, 0, st, P);
        break;
    default:
        return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename R>
int launch_real(const RelaxParams &P, hipStream_t st)
{
    if (P.law == LAW_ADVECTION) {
        switch (P.dim) {
        case 1: return launch_mode<R, LAW_ADVECTION, 1>(P, st);
        case 2: return launch_mode<R, LAW_ADVECTION, 2>(P, st);
        case 3: return launch_mode<R, LAW_ADVECTION, 3>(P, st);
        }
    } else if (P.law == LAW_EULER) {
        switch (P.dim) {
        case 1: return launch_mode<R, LAW_EULER, 1>(P, st);
        case 2: return launch_mode<R, LAW_EULER, 2>(P, st);
        case 3: return launch_mode<R, LAW_EULER, 3>(P, st);
        }
    }
    return -1;
}

}  // namespace

// One workgroup per (slot, which): each thread sums its strided share in
// index order, then a fixed tree over the threads; (sum, compensation) pairs
// combined by TwoSum throughout.  Element i sits at
// (i / group) * group_stride + i % group; G is group when it is 1 or 2 (no
// division), else 0.  Sixteen elements are requested before they are added,
// in order.
template <int G>
__global__ void __launch_bounds__(256) slot_sums_kernel(const double *__restrict__ src, int64_t n,
                                                        int64_t slot_stride, int64_t which_stride,
                                                        int group, int64_t group_stride,
                                                        double *__restrict__ out)
{
    __shared__ double red[2 * 256];
    const int t = threadIdx.x;
    const double *base = src + blockIdx.x * slot_stride + blockIdx.y * which_stride;
    auto at = [&](int64_t i) {
        if constexpr (G == 1)
            return base[i * group_stride];
        else if constexpr (G == 2)
            return base[(i >> 1) * group_stride + (i & 1)];
        else
            return base[(i / group) * group_stride + i % group];
    };
    constexpr int U = 16;
    double s = 0.0, c = 0.0;
    int64_t i = t;
    for (; i + (U - 1) * 256 < n; i += U * 256) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            v[u] = at(i + u * 256);
#pragma unroll
        for (int u = 0; u < U; ++u)
            comp_add(s, c, v[u], 0.0);
    }
    for (; i < n; i += 256)
        comp_add(s, c, at(i), 0.0);
    red[2 * t] = s;
    red[2 * t + 1] = c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            double a = red[2 * t], ac = red[2 * t + 1];
            comp_add(a, ac, red[2 * (t + w)], red[2 * (t + w) + 1]);
            red[2 * t] = a;
            red[2 * t + 1] = ac;
        }
        __syncthreads();
    }
    if (t == 0)
        out[blockIdx.x * gridDim.y + blockIdx.y] = red[0] + red[1];
}

int launch_slot_sums(const double *src, int64_t n, int nslots, int nwhich, int64_t slot_stride,
                     int64_t which_stride, int group, int64_t group_stride, double *out,
                     void *stream)
{
    if (nslots <= 0 || nwhich <= 0)
        return 0;
    const dim3 grid((unsigned)nslots, (unsigned)nwhich);
    hipStream_t st = static_cast<hipStream_t>(stream);
    group = group < 1 ? 1 : group;
    if (group == 1)
        hipLaunchKernelGGL(slot_sums_kernel<1>, grid, dim3(256), 0, st, src, n, slot_stride,
                           which_stride, group, group_stride, out);
    else if (group == 2)
        hipLaunchKernelGGL(slot_sums_kernel<2>, grid, dim3(256), 0, st, src, n, slot_stride,
                           which_stride, group, group_stride, out);
    else
        hipLaunchKernelGGL(slot_sums_kernel<0>, grid, dim3(256), 0, st, src, n, slot_stride,
                           which_stride, group, group_stride, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_relax(const RelaxParams &P, void *stream)
{
    if (P.K <= 0 && P.mode == RELAX_AXPY)
        return 0;
    hipStream_t st = static_cast<hipStream_t>(stream);
    return P.real_bytes == 8 ? launch_real<double>(P, st) : launch_real<float>(P, st);
}

}  // namespace rv
