// Mortar coupling of hanging faces (opt-in; Raven leaves these faces
// self-mapped, lobattocells.jl:2072-2074, and never reads its vmapNC /
// tohalves structures in the right-hand side).
//
// A coarse face covered by 2^(dim-1) finer faces is coupled through the fine
// faces (the mortar).  On sub-face s = s1 + 2 s2 the coarse trace is
// interpolated to the fine GLL nodes with tohalves_1d (lobattocells.jl:15-16):
//     q_c->s(a, b) = sum_j H[s2][b][j] sum_i H[s1][a][i] q_c(i, j),
// and the numerical flux nf_s = nflux(n_f; q_f, q_c->s) is evaluated with the
// fine face's normal.  The fine cell adds -MJI wsJ_f nf_s at its face nodes, as
// a conforming face would.  The coarse cell computes the same sub-face fluxes
// (same inputs, same operations) and adds their projection
//     +MJI(i, j) sum_s sum_b H[s2][b][j] sum_a H[s1][a][i] wsJ_f nf_s(a, b),
// the transpose of the interpolation, so the fluxes through the two sides of
// the mortar cancel exactly up to rounding (discrete conservation).  The stage
// kernel leaves these faces out (their lift weight is zero on the device).
//  * mortar_face_kernel, one wave per (cell with non-conforming faces, face),
//    runs before the stage kernel on the same stream: the face's lift L at
//    its face nodes (the update is delta = -MJI L), into a per-face buffer
//    (the faces only read q_in, so they are independent of each other and of
//    the stage): L = wsJ_f nf_s on the fine side, minus the projection on the
//    coarse side;
//  * the 3-D column kernel takes L as the lift of these faces, as it takes
//    lift x flux on a conforming face; the other stage kernels leave them to
//    mortar_apply_kernel (a thread per node of those cells), which adds
//    dq += delta and, in stage mode, q_out += b_dt * delta after the stage,
//    faces in ascending order, so every update is deterministic (and the
//    same as applying the faces one after the other).
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "dg_kernels.hpp"

namespace rv {

namespace {

template <int DIM, int NQ, int LAW, typename R, int FLUX>
__global__ void __launch_bounds__(64) mortar_face_kernel(const MortarParams M)
{
    using S = Stage<DIM, NQ, LAW, R, GEO_GENERAL, FLUX>;
    constexpr int NS = S::NS, NP = S::NP, NFP = S::NP2, NF = 2 * DIM;
    constexpr int NSUB = 1 << (DIM - 1);
    static_assert(NFP <= 64, "one face node per lane");
    // tohalves in the state type ([half][a][b]), a face trace, the first pass
    // of a separable 2-D interpolation, and the coarse side's weighted
    // sub-face fluxes with the first pass of their projection
    __shared__ R H[2][NQ][NQ];
    __shared__ R tr[NS][NFP];
    __shared__ R t1[NS][NFP];
    __shared__ R g[NSUB][NS][NFP];
    __shared__ R u[NSUB][NS][NFP];
    const StageParams &P = M.stage;
    const Law<DIM, LAW, R> law(P);
    // one workgroup per non-conforming (cell, face) pair, its cell, face,
    // kind and partners in one record
    const MortarPair &pr = M.pairs[blockIdx.x];
    const int64_t e = pr.e, li = pr.li;
    const int f = pr.f, kind = pr.kind;
    const int m = threadIdx.x;
    const bool on = m < NFP;
    const int a = m % NQ, b = m / NQ;  // face node m = a + NQ*b (b = 0 in 2-D)
    const R *qin = static_cast<const R *>(P.q_in);
    R *out = static_cast<R *>(M.delta) + (li * NF + f) * (NS * NFP);
    for (int t = m; t < 2 * NQ * NQ; t += 64)
        (&H[0][0][0])[t] = R(M.halves[t]);
    __syncthreads();

    // the trace in tr interpolated to node m of sub-face sub:
    //   sum_jj H[s2][b][jj] (sum_i H[s1][a][i] tr(i, jj)),
    // the inner sums computed once per (a, jj) by lane a + NQ*jj
    auto interp = [&](int sub, R *out) {
        const int s1 = sub & 1, s2 = (sub >> 1) & 1;
        if constexpr (DIM == 2) {
            if (on) {
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    R v = R(0);
#pragma unroll
                    for (int i = 0; i < NQ; ++i)
                        v = v + H[s1][a][i] * tr[s][i];
                    out[s] = v;
                }
            }
        } else {
            if (on) {
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    R inner = R(0);
#pragma unroll
                    for (int i = 0; i < NQ; ++i)
                        inner = inner + H[s1][a][i] * tr[s][i + NQ * b];
                    t1[s][m] = inner;
                }
            }
            __syncthreads();
            if (on) {
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    R v = R(0);
#pragma unroll
                    for (int jj = 0; jj < NQ; ++jj)
                        v = v + H[s2][b][jj] * t1[s][a + NQ * jj];
                    out[s] = v;
                }
            }
            __syncthreads();  // t1 is reused by the next interpolation
        }
    };
    // what a sub-face flux reads of the fine side, as the fine cell sees it:
    // q- = the fine state at its face node, the fine face's normal and weight
    struct Fine {
        R q[NS], n[DIM], wsJ;
    };
    auto load_fine = [&](int64_t fine, int ffine, Fine &F) {
        const int node = face_node<DIM, NQ>(ffine, m);
#pragma unroll
        for (int s = 0; s < NS; ++s)
            F.q[s] = qin[fine * NS * NP + s * NP + node];
        const double *sg = M.sgeo + ((fine * NF + ffine) * (DIM + 1)) * NFP + m;
#pragma unroll
        for (int c = 0; c < DIM; ++c)
            F.n[c] = R(sg[c * NFP]);
        F.wsJ = R(sg[DIM * NFP]);
    };
    // the sub-face flux with q+ = the interpolated coarse trace qi
    auto subflux = [&](const Fine &F, const R *qi, R *wnf) {
        R qf[NS], nf[NS], n[DIM], qp[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            qf[s] = F.q[s];
            qp[s] = qi[s];
        }
#pragma unroll
        for (int c = 0; c < DIM; ++c)
            n[c] = F.n[c];
        S::nflux_general(law, n, qf, qp, nf);
#pragma unroll
        for (int s = 0; s < NS; ++s)
            wnf[s] = F.wsJ * nf[s];
    };
    // the face's lift at face node m, to the per-face buffer
    auto store_lift = [&](const R *lift) {
#pragma unroll
        for (int s = 0; s < NS; ++s)
            out[s * NFP + m] = lift[s];
    };

    int64_t part[NSUB];
#pragma unroll
    for (int t = 0; t < NSUB; ++t)
        part[t] = pr.part[t];
    const int node = face_node<DIM, NQ>(f, m);
    if (kind == 1) {
        // this cell is the fine side: the coarse trace on its face f^1
        const int64_t c = part[0];
        const int sub = (int)part[1];
        Fine F;
        if (on) {
            const int cn = face_node<DIM, NQ>(f ^ 1, m);
            R trv[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s)
                trv[s] = qin[c * NS * NP + s * NP + cn];
            load_fine(e, f, F);
#pragma unroll
            for (int s = 0; s < NS; ++s)
                tr[s][m] = trv[s];
        }
        __syncthreads();
        R qi[NS];
        interp(sub, qi);
        if (on) {
            R wnf[NS];
            subflux(F, qi, wnf);
            store_lift(wnf);
        }
    } else {
        // this cell is the coarse side: its own trace, the fine fluxes of
        // every sub-face, then their projection onto its face nodes
        Fine F[NSUB];
        if (on) {
            R trv[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s)
                trv[s] = qin[e * NS * NP + s * NP + node];
#pragma unroll
            for (int sub = 0; sub < NSUB; ++sub)
                load_fine(part[sub], f ^ 1, F[sub]);
#pragma unroll
            for (int s = 0; s < NS; ++s)
                tr[s][m] = trv[s];
        }
        __syncthreads();
#pragma unroll
        for (int sub = 0; sub < NSUB; ++sub) {
            R qi[NS];
            interp(sub, qi);
            if (on) {
                R wnf[NS];
                subflux(F[sub], qi, wnf);
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    g[sub][s][m] = wnf[s];
            }
        }
        __syncthreads();
        // projection: sum_sub sum_b H[s2][b][j] (sum_a H[s1][a][i] g(a, b)),
        // the inner sums once per (i, b) by lane i + NQ*b
        if constexpr (DIM == 3) {
            if (on) {
                const int i = a, bb = b;
                for (int sub = 0; sub < NSUB; ++sub) {
                    const int s1 = sub & 1;
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        R inner = R(0);
#pragma unroll
                        for (int aa = 0; aa < NQ; ++aa)
                            inner = inner + H[s1][aa][i] * g[sub][s][aa + NQ * bb];
                        u[sub][s][m] = inner;
                    }
                }
            }
            __syncthreads();
        }
        if (on) {
            const int i = a, j = b;
            R lift[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                R v = R(0);
                for (int sub = 0; sub < NSUB; ++sub) {
                    const int s1 = sub & 1, s2 = (sub >> 1) & 1;
                    if constexpr (DIM == 2) {
#pragma unroll
                        for (int aa = 0; aa < NQ; ++aa)
                            v = v + H[s1][aa][i] * g[sub][s][aa];
                    } else {
#pragma unroll
                        for (int bb = 0; bb < NQ; ++bb)
                            v = v + H[s2][bb][j] * u[sub][s][i + NQ * bb];
                    }
                }
                lift[s] = -v;
            }
            store_lift(lift);
        }
    }
}

// Index of node `node` on face f (the inverse of face_node), or -1 when the
// node is not on that face.
template <int DIM, int NQ>
__device__ inline int node_on_face(int f, int node)
{
    const int x = node % NQ, y = (node / NQ) % NQ, z = DIM == 3 ? node / (NQ * NQ) : 0;
    const int d = f >> 1;
    const int fixed = (f & 1) ? NQ - 1 : 0;
    if constexpr (DIM == 2) {
        if (d == 0)
            return x == fixed ? y : -1;
        return y == fixed ? x : -1;
    } else {
        if (d == 0)
            return x == fixed ? y + NQ * z : -1;
        if (d == 1)
            return y == fixed ? x + NQ * z : -1;
        return z == fixed ? x + NQ * y : -1;
    }
}

// dq += delta (and q_out += b_dt delta) at node `node` of mortar cell li, for
// its non-conforming faces in ascending order; delta = -MJI L.
template <int DIM, int NQ, int LAW, typename R>
__global__ void __launch_bounds__(256) mortar_apply_kernel(const MortarParams M)
{
    constexpr int NS = LawInfo<DIM, LAW>::NS, NP = IPow<NQ, DIM>::v, NFP = NP / NQ;
    constexpr int NF = 2 * DIM;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t li = t / NP;
    if (li >= M.ncells)
        return;
    const int node = (int)(t - li * NP);
    const int64_t e = M.cells[li];
    int mf[NF];
    bool any = false;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        mf[f] = node_on_face<DIM, NQ>(f, node);
        if (mf[f] >= 0 && M.kind[e * NF + f] == 0)
            mf[f] = -1;
        any = any || mf[f] >= 0;
    }
    if (!any)
        return;
    const StageParams &P = M.stage;
    R *dq = static_cast<R *>(P.dq);
    R *qo = static_cast<R *>(P.q_out);
    const R *lift = static_cast<const R *>(M.delta) + li * NF * (NS * NFP);
    const R bdt = R(P.b_dt);
    const R mji = R(P.mji[e * NP + node]);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int64_t idx = e * NS * NP + s * NP + node;
        R v = dq[idx];
        R w = P.write_q ? qo[idx] : R(0);
#pragma unroll
        for (int f = 0; f < NF; ++f)
            if (mf[f] >= 0) {
                const R d = -(mji * lift[(f * NS + s) * NFP + mf[f]]);
                v = v + d;
                w = w + bdt * d;
            }
        dq[idx] = v;
        if (P.write_q)
            qo[idx] = w;
    }
}

template <int DIM, int NQ, int LAW, typename R, int FLUX>
int launch_mortar(const MortarParams &M, void *stream, int phase)
{
    constexpr int NP = IPow<NQ, DIM>::v;
    if (M.ncells <= 0)
        return 0;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (phase == 0) {
        if (M.npairs <= 0)
            return 0;
        hipLaunchKernelGGL((mortar_face_kernel<DIM, NQ, LAW, R, FLUX>),
                           dim3((unsigned)M.npairs), dim3(64), 0, st, M);
    } else {
        const int64_t nodes = M.ncells * NP;
        hipLaunchKernelGGL((mortar_apply_kernel<DIM, NQ, LAW, R>),
                           dim3((unsigned)((nodes + 255) / 256)), dim3(256), 0, st, M);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int DIM, int NQ, typename R>
MortarLauncher pick_law(int law, int flux)
{
    if (law == LAW_EULER)
        return flux == FLUX_RUSANOV ? &launch_mortar<DIM, NQ, LAW_EULER, R, FLUX_RUSANOV>
                                    : &launch_mortar<DIM, NQ, LAW_EULER, R, FLUX_CENTRAL>;
    return flux == FLUX_RUSANOV ? &launch_mortar<DIM, NQ, LAW_ADVECTION, R, FLUX_RUSANOV>
                                : &launch_mortar<DIM, NQ, LAW_ADVECTION, R, FLUX_CENTRAL>;
}

template <int DIM, int NQ>
MortarLauncher pick_nq(int nq, int law, int real_bytes, int flux)
{
    if constexpr (NQ > 8) {
        return nullptr;
    } else {
        if (nq != NQ)
            return pick_nq<DIM, NQ + 1>(nq, law, real_bytes, flux);
        return real_bytes == 4 ? pick_law<DIM, NQ, float>(law, flux)
                               : pick_law<DIM, NQ, double>(law, flux);
    }
}

}  // namespace

MortarLauncher find_mortar_launcher(int dim, int nq, int law, int real_bytes, int flux)
{
    if (flux != FLUX_RUSANOV && flux != FLUX_CENTRAL)
        return nullptr;
    if (dim == 2)
        return pick_nq<2, 2>(nq, law, real_bytes, flux);
    if (dim == 3)
        return pick_nq<3, 2>(nq, law, real_bytes, flux);
    return nullptr;
}

}  // namespace rv
