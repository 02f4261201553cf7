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
// kernel leaves these faces out (their lift weight is zero on the device), and
// this kernel runs after it on the same stream: one wave per cell with
// non-conforming faces, the faces in order, so every update is deterministic.
// dq += delta and, in stage mode, q_out += b_dt * delta.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "dg_kernels.hpp"

namespace rv {

namespace {

template <int DIM, int NQ, int LAW, typename R, int FLUX>
__global__ void __launch_bounds__(64) mortar_kernel(const MortarParams M)
{
    using S = Stage<DIM, NQ, LAW, R, GEO_GENERAL, FLUX>;
    constexpr int NS = S::NS, NP = S::NP, NFP = S::NP2, NF = 2 * DIM;
    constexpr int NSUB = 1 << (DIM - 1);
    __shared__ R tr[NS][NFP];
    __shared__ R g[NSUB][NS][NFP];
    const StageParams &P = M.stage;
    const Law<DIM, LAW, R> law(P);
    const int64_t e = M.cells[blockIdx.x];
    const int m = threadIdx.x;
    const bool on = m < NFP;
    const R *qin = static_cast<const R *>(P.q_in);
    R *dq = static_cast<R *>(P.dq);
    R *qo = static_cast<R *>(P.q_out);
    const double *H = M.halves;  // [half][a][b]

    // interpolation of the trace in tr to node m of sub-face sub
    auto interp = [&](int sub, R *out) {
        const int s1 = sub & 1, s2 = (sub >> 1) & 1;
        const int a = m % NQ, b = m / NQ;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            R v = R(0);
            if constexpr (DIM == 2) {
                for (int i = 0; i < NQ; ++i)
                    v = v + R(H[(s1 * NQ + a) * NQ + i]) * tr[s][i];
            } else {
                for (int jj = 0; jj < NQ; ++jj) {
                    R inner = R(0);
                    for (int i = 0; i < NQ; ++i)
                        inner = inner + R(H[(s1 * NQ + a) * NQ + i]) * tr[s][i + NQ * jj];
                    v = v + R(H[(s2 * NQ + b) * NQ + jj]) * inner;
                }
            }
            out[s] = v;
        }
    };
    // sub-face flux as the fine cell sees it: q- = fine state, q+ = the
    // interpolated coarse trace, the fine face's normal and weight
    auto subflux = [&](int64_t fine, int ffine, int sub, R *wnf) {
        R qi[NS], qf[NS], nf[NS], n[DIM];
        interp(sub, qi);
        const int node = face_node<DIM, NQ>(ffine, m);
#pragma unroll
        for (int s = 0; s < NS; ++s)
            qf[s] = qin[fine * NS * NP + s * NP + node];
        const double *sg = M.sgeo + ((fine * NF + ffine) * (DIM + 1)) * NFP + m;
#pragma unroll
        for (int c = 0; c < DIM; ++c)
            n[c] = R(sg[c * NFP]);
        const R wsJ = R(sg[DIM * NFP]);
        S::nflux_general(law, n, qf, qi, nf);
#pragma unroll
        for (int s = 0; s < NS; ++s)
            wnf[s] = wsJ * nf[s];
    };
    auto apply = [&](int node, const R *delta) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int64_t idx = e * NS * NP + s * NP + node;
            dq[idx] = dq[idx] + delta[s];
            if (P.write_q)
                qo[idx] = qo[idx] + R(P.b_dt) * delta[s];
        }
    };

    for (int f = 0; f < NF; ++f) {
        const int kind = M.kind[e * NF + f];
        if (kind == 0)
            continue;
        const int64_t *pt = M.partner + 4 * (e * NF + f);
        if (kind == 1) {
            // this cell is the fine side: the coarse trace on its face f^1
            const int64_t c = pt[0];
            const int sub = (int)pt[1];
            if (on) {
                const int cn = face_node<DIM, NQ>(f ^ 1, m);
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    tr[s][m] = qin[c * NS * NP + s * NP + cn];
            }
            __syncthreads();
            if (on) {
                R wnf[NS], delta[NS];
                subflux(e, f, sub, wnf);
                const int node = face_node<DIM, NQ>(f, m);
                const R mji = R(M.stage.mji[e * NP + node]);
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    delta[s] = -(mji * wnf[s]);
                apply(node, delta);
            }
            __syncthreads();
        } else {
            // this cell is the coarse side: its own trace, the fine fluxes of
            // every sub-face, then their projection onto its face nodes
            const int node = face_node<DIM, NQ>(f, m);
            if (on) {
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    tr[s][m] = qin[e * NS * NP + s * NP + node];
            }
            __syncthreads();
            if (on) {
                for (int sub = 0; sub < NSUB; ++sub) {
                    R wnf[NS];
                    subflux(pt[sub], f ^ 1, sub, wnf);
#pragma unroll
                    for (int s = 0; s < NS; ++s)
                        g[sub][s][m] = wnf[s];
                }
            }
            __syncthreads();
            if (on) {
                const int i = m % NQ, j = m / NQ;
                R delta[NS];
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    R v = R(0);
                    for (int sub = 0; sub < NSUB; ++sub) {
                        const int s1 = sub & 1, s2 = (sub >> 1) & 1;
                        if constexpr (DIM == 2) {
                            for (int a = 0; a < NQ; ++a)
                                v = v + R(H[(s1 * NQ + a) * NQ + i]) * g[sub][s][a];
                        } else {
                            for (int b = 0; b < NQ; ++b) {
                                R inner = R(0);
                                for (int a = 0; a < NQ; ++a)
                                    inner = inner + R(H[(s1 * NQ + a) * NQ + i]) *
                                                        g[sub][s][a + NQ * b];
                                v = v + R(H[(s2 * NQ + b) * NQ + j]) * inner;
                            }
                        }
                    }
                    delta[s] = R(M.stage.mji[e * NP + node]) * v;
                }
                apply(node, delta);
            }
            __syncthreads();
        }
    }
}

template <int DIM, int NQ, int LAW, typename R, int FLUX>
int launch_mortar(const MortarParams &M, void *stream)
{
    if (M.ncells <= 0)
        return 0;
    hipLaunchKernelGGL((mortar_kernel<DIM, NQ, LAW, R, FLUX>), dim3((unsigned)M.ncells), dim3(64),
                       0, static_cast<hipStream_t>(stream), M);
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
    if (dim == 2)
        return pick_nq<2, 2>(nq, law, real_bytes, flux);
    if (dim == 3)
        return pick_nq<3, 2>(nq, law, real_bytes, flux);
    return nullptr;
}

}  // namespace rv
