// Fused 2-D DGSEM stage kernel on affine cells for gfx950: a lane per node.
//
// Same operator and the same arithmetic, operation for operation, as the
// pencil kernel of dg_kernels.hpp (Raven's weak-form volume term, Rusanov /
// central surface term, LSRK update: dgsem.jl:257-348, 733-784,
// odesolvers.jl:55-57), so the two give bit-identical results; what changes
// is the memory access pattern:
//
//  * lane (i, j) of a cell owns node i + Nq j: q and dq_old are read, and dq
//    and q_out written, with lane-contiguous addresses (the pencil kernel's
//    x-pencils are strided by Nq), and q stays in registers for the update;
//  * each lane evaluates both flux rows at its node and publishes them in LDS,
//    x-fluxes in node order and y-fluxes transposed, so that the two
//    derivative sums a lane needs each read one contiguous row;
//  * the 4 Nq face nodes of a cell are spread over the lanes (one face node
//    per lane): the face record, the own and the neighbour state are requested
//    with the node loads, and the numerical flux is published in LDS, where
//    the lanes on the cell's boundary pick it up (interior lanes read a zero
//    spare, which leaves their sums unchanged);
//  * Nq^2 = 25, 36, 49 leave lanes of a wave idle, so G = 2, 4, 4 waves
//    share one workgroup's cells (5, 7, 5 cells over 128, 256, 256 lanes) and
//    synchronise with workgroup barriers (so do batches launching a kind on
//    its own); the batches' work-list kernel uses G = 1;
//  * light cells (advection, FP32 Euler) give a lane V = 4 or 2 nodes, one in
//    each of V sets of cells, all requested before the first wait: a wave
//    then has V times the bytes in flight.
#pragma once

#include "dg_kernels.hpp"

namespace rv {

template <int NQ, int LAW, typename R, int FLUX, int GW, int VW>
struct Quad2 {
    static constexpr int NS = LawInfo<2, LAW>::NS;
    static constexpr int NP = NQ * NQ;
    static constexpr int G = GW;                 // waves sharing the cells
    static constexpr int V = VW;                 // nodes per lane (sets of cells)
    static constexpr int LANES = 64 * G;
    static constexpr int CPS = LANES / NP;       // cells per set
    static constexpr int CELLS = V * CPS;        // cells per group
    static constexpr int NFU = 4 * NQ;           // face nodes per cell
    static constexpr int FPASS = (CPS * NFU + LANES - 1) / LANES;
    // LDS of a group, in elements of R.  The spare sits first, at the same
    // place for every (Nq, law), so that kinds sharing a slab in the work-list
    // kernel never overwrite it; then one part per set of cells.
    static constexpr int SPARE = 0;                        // SPARE_LEN zeros: 32 bytes in
    static constexpr int SPARE_LEN = 32 / (int)sizeof(R);  // either precision
    static constexpr int FXO = 0;                          // [c][s][j][n] = f_x(n, j)
    static constexpr int FYO = CPS * NS * NP;              // [c][s][i][n] = f_y(i, n)
    static constexpr int NFO = 2 * CPS * NS * NP;          // [c][f][p][s] numerical flux
    static constexpr int PART = NFO + CPS * NFU * NS;
    static constexpr int SLAB = SPARE_LEN + V * PART;
    static_assert(CPS >= 1, "Nq^2 must not exceed the group's lanes");
    static_assert(NS <= SPARE_LEN, "the spare must cover every state");
};

// Waves per group of the single-operator kernel: pack cells across waves
// where Nq^2 leaves a wave's lanes idle.
template <int NQ>
constexpr int quad2_group_waves()
{
    return NQ == 5 ? 2 : (NQ == 6 || NQ == 7) ? 4 : 1;
}

// Nodes per lane: as many as the registers allow for the law and precision.
template <int LAW, typename R>
constexpr int quad2_nodes_per_lane()
{
    return LAW == LAW_ADVECTION ? 4 : sizeof(R) == 4 ? 2 : 1;
}

// The cells [cell0, cell0 + CELLS) of one group: set v holds the cells
// [cell0 + v CPS, cell0 + (v + 1) CPS).  gl is the lane's index in the group
// (0 .. 64 G - 1) and glds the group's slab, whose spare must hold zeros
// (quad2_clear_spare) when the group starts.
template <int NQ, int LAW, typename R, int FLUX, int ST, int G, int V>
__device__ __attribute__((always_inline)) inline void quad2_group(const StageParams &P,
                                                                  int64_t cell0, R *glds, int gl)
{
    using Q = Quad2<NQ, LAW, R, FLUX, G, V>;
    using S = Stage<2, NQ, LAW, R, GEO_AFFINE, FLUX>;
    constexpr int NS = Q::NS, NP = Q::NP;
    const Law<2, LAW, R> law(P);
    const R *qin = static_cast<const R *>(P.q_in);
    R *dqp = static_cast<R *>(P.dq);
    R *qo = static_cast<R *>(P.q_out);
    const bool onx = (P.dir_mask & 1) != 0, ony = (P.dir_mask & 2) != 0;

    const int c = gl / NP;
    const int node = gl - c * NP;
    const int i = node % NQ, j = node / NQ;
    bool active[V];
    int64_t off[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int64_t e = cell0 + v * Q::CPS + c;
        active[v] = c < Q::CPS && e < P.K;
        off[v] = e * (NS * NP) + node;
    }

    // the node's state and dq_old, the cell scales, and the operator rows and
    // lift weights of the node (requested before the first wait)
    R q[V][NS], dqo[V][NS], sx[V], sy[V], mx[NQ], my[NQ];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        sx[v] = sy[v] = R(0);
        if (active[v]) {
#pragma unroll
            for (int s = 0; s < NS; ++s)
                q[v][s] = qin[off[v] + s * NP];
            if (P.read_dq) {
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    dqo[v][s] = load_dq<ST>(dqp + off[v] + s * NP);
            }
            const int64_t e = cell0 + v * Q::CPS + c;
            sx[v] = R(P.affine[e * 2 + 0]);
            sy[v] = R(P.affine[e * 2 + 1]);
        }
    }
    R wx = R(0), wy = R(0);
    if (active[0]) {
        const double *M = P.Dw;
#pragma unroll
        for (int n = 0; n < NQ; ++n) {
            mx[n] = R(M[i * NQ + n]);
            my[n] = R(M[j * NQ + n]);
        }
        wx = R(P.winv[i == 0 ? 0 : NQ - 1]);
        wy = R(P.winv[j == 0 ? 0 : NQ - 1]);
    }
    // face nodes: own and neighbour state (face record first)
    R qm[V][Q::FPASS][NS], qp[V][Q::FPASS][NS];
#pragma unroll
    for (int v = 0; v < V; ++v)
#pragma unroll
        for (int ps = 0; ps < Q::FPASS; ++ps) {
            const int u = gl + ps * Q::LANES;
            const int fc = u / Q::NFU;
            const int r = u - fc * Q::NFU;
            const int f = r / NQ, p = r - f * NQ;
            const int64_t fe = cell0 + v * Q::CPS + fc;
            const bool d_on = (f >> 1) == 0 ? onx : ony;
            if (fc < Q::CPS && fe < P.K && d_on) {
                const R *qe = qin + fe * (NS * NP) + face_node<2, NQ>(f, p);
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    qm[v][ps][s] = qe[s * NP];
                S::neighbour_state(P, qin, fe, f, p, qm[v][ps], qp[v][ps]);
            }
        }
    // flux rows at the node, published for the derivative sums
#pragma unroll
    for (int v = 0; v < V; ++v)
        if (active[v]) {
            R fx[NS], fy[NS];
            law.flux_row(0, q[v], fx);
            law.flux_row(1, q[v], fy);
            R *cl = glds + Q::SPARE_LEN + v * Q::PART + c * (NS * NP);
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                cl[Q::FXO + s * NP + node] = fx[s];
                cl[Q::FYO + s * NP + i * NQ + j] = fy[s];
            }
        }
    // numerical fluxes of the face nodes
#pragma unroll
    for (int v = 0; v < V; ++v)
#pragma unroll
        for (int ps = 0; ps < Q::FPASS; ++ps) {
            const int u = gl + ps * Q::LANES;
            const int fc = u / Q::NFU;
            const int r = u - fc * Q::NFU;
            const int f = r / NQ;
            const int64_t fe = cell0 + v * Q::CPS + fc;
            const int d = f >> 1;
            const bool d_on = d == 0 ? onx : ony;
            if (fc < Q::CPS && fe < P.K && d_on) {
                // the direction as a constant: a run-time one would index the
                // law's private arrays (scratch memory)
                R fm[NS], nf[NS];
                const R sgn = (f & 1) ? R(1) : R(-1);
                if (d == 0) {
                    law.flux_row(0, qm[v][ps], fm);
                    S::nflux_axis(law, 0, sgn, fm, qm[v][ps], qp[v][ps], nf);
                } else {
                    law.flux_row(1, qm[v][ps], fm);
                    S::nflux_axis(law, 1, sgn, fm, qm[v][ps], qp[v][ps], nf);
                }
                R *nb = glds + Q::SPARE_LEN + v * Q::PART + Q::NFO;
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    nb[u * NS + s] = nf[s];
            }
        }
    if constexpr (G == 1)
        wave_lds_sync();
    else
        __syncthreads();

    const R a = R(P.a), bdt = R(P.b_dt);
#pragma unroll
    for (int v = 0; v < V; ++v) {
        if (!active[v])
            continue;
        // numerical flux of the face the node lies on (the spare's zeros for
        // interior nodes)
        const int fcell = Q::SPARE_LEN + v * Q::PART + Q::NFO + c * Q::NFU * NS;
        const int bx = i == 0 ? fcell + j * NS
                              : (i == NQ - 1 ? fcell + (NQ + j) * NS : Q::SPARE);
        const int by = j == 0 ? fcell + (2 * NQ + i) * NS
                              : (j == NQ - 1 ? fcell + (3 * NQ + i) * NS : Q::SPARE);
        const R *cl = glds + Q::SPARE_LEN + v * Q::PART + c * (NS * NP);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            R vx = R(0), vy = R(0);
            if (onx) {
                const R *row = cl + Q::FXO + s * NP + j * NQ;
                vx = mx[0] * row[0];
#pragma unroll
                for (int n = 1; n < NQ; ++n)
                    vx = vx + mx[n] * row[n];
                vx = vx - wx * glds[bx + s];
                vx = sx[v] * vx;
            }
            if (ony) {
                const R *row = cl + Q::FYO + s * NP + i * NQ;
                vy = my[0] * row[0];
#pragma unroll
                for (int n = 1; n < NQ; ++n)
                    vy = vy + my[n] * row[n];
                vy = vy - wy * glds[by + s];
                vy = sy[v] * vy;
            }
            const R tot = vx + vy;
            R val = tot;
            if (P.read_dq)
                val = a * dqo[v][s] + tot;
            store_out<ST>(val, dqp + off[v] + s * NP);
            if (P.write_q)
                store_out<ST>(R(q[v][s] + bdt * val), qo + off[v] + s * NP);
        }
    }
}

// Zero the spare (the first 32 bytes) of a group's slab, once per slab;
// nothing else writes it, whatever the precision of the cells using the slab.
template <typename R>
__device__ inline void quad2_clear_spare(R *glds, int gl)
{
    if (gl < 32 / (int)sizeof(R))
        glds[gl] = R(0);
}

// Single operator: a workgroup of wpb waves holds wpb / G groups (G > 1:
// wpb = G, one group).
template <int NQ, int LAW, typename R, int FLUX, int ST, int G, int V>
__global__ void __launch_bounds__(256) dg_quad2_kernel(const StageParams P)
{
    using Q = Quad2<NQ, LAW, R, FLUX, G, V>;
    extern __shared__ __attribute__((aligned(16))) unsigned char dg_lds[];
    const int grp = (int)(threadIdx.x / Q::LANES);
    const int gl = (int)(threadIdx.x - grp * Q::LANES);
    const int gpb = (int)(blockDim.x / Q::LANES);
    const int64_t blk = xcd_remap(blockIdx.x, gridDim.x);
    R *glds = reinterpret_cast<R *>(dg_lds) + grp * Q::SLAB;
    quad2_clear_spare(glds, gl);
    quad2_group<NQ, LAW, R, FLUX, ST, G, V>(P, (blk * gpb + grp) * Q::CELLS, glds, gl);
}

template <int NQ, int LAW, typename R, int FLUX>
int launch_quad2(const StageParams &P, void *stream, int wpb)
{
    constexpr int G = quad2_group_waves<NQ>();
    constexpr int V = quad2_nodes_per_lane<LAW, R>();
    using Q = Quad2<NQ, LAW, R, FLUX, G, V>;
    if (P.K <= 0)
        return 0;
    if (G > 1)
        wpb = G;
    else if (wpb != 1 && wpb != 2 && wpb != 4)
        wpb = 1;
    const int gpb = wpb / G;
    const int64_t groups = (P.K + Q::CELLS - 1) / Q::CELLS;
    const int64_t blocks = (groups + gpb - 1) / gpb;
    const size_t lds = (size_t)gpb * Q::SLAB * sizeof(R);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (P.streaming)
        hipLaunchKernelGGL((dg_quad2_kernel<NQ, LAW, R, FLUX, 1, G, V>), dim3((unsigned)blocks),
                           dim3(64 * wpb), lds, st, P);
    else
        hipLaunchKernelGGL((dg_quad2_kernel<NQ, LAW, R, FLUX, 0, G, V>), dim3((unsigned)blocks),
                           dim3(64 * wpb), lds, st, P);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace rv
