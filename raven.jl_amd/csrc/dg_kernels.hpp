// Fused DGSEM stage kernel for gfx950 (CDNA4, wave64).
//
// One fused launch computes, for every cell, Raven's weak-form volume term
// (dgsem.jl:257-348), its surface term with the Rusanov/central flux
// (dgsem.jl:733-784, numericalfluxes.jl:27-50) and the low-storage RK update
// (odesolvers.jl:55-57):
//     dq = a*dq + L(q_in);   q_out = q_in + b_dt*dq
//
// Work decomposition ("pencil sweeps").  A wave owns EPW = 64 / Nq^(dim-1)
// cells; a lane owns one pencil (a line of Nq nodes) of one cell.  For each
// direction d the lane loads its pencil along d, evaluates the flux component
// needed by d, applies the 1-D derivative in registers (D is wave-uniform and
// comes through scalar loads), and adds the lift of the two faces the pencil
// ends on (faces 2d and 2d+1: their face node index equals the pencil index).
// The directions' contributions meet in LDS (one slab per cell); the last
// sweep adds them, applies the RK update and stores with lane-contiguous
// addresses.  Neighbour states are read through per-face records (neighbour
// cell, face, orientation), so a face costs 8 bytes of index traffic instead of
// Nq^(dim-1) 8-byte node indices.
//
// Owner-computes: each cell evaluates the numerical flux of all of its faces;
// no atomics, deterministic results.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace rv {

template <int B, int E>
struct IPow {
    static constexpr int v = B * IPow<B, E - 1>::v;
};
template <int B>
struct IPow<B, 0> {
    static constexpr int v = 1;
};

// Blocks are dealt to the 8 XCDs round-robin; renumber them so that each XCD
// walks a contiguous range of cells (neighbours then share that XCD's L2).
__device__ inline int64_t xcd_remap(int64_t b, int64_t nb)
{
    constexpr int64_t NX = 8;
    const int64_t q = nb / NX, r = nb % NX;
    const int64_t x = b % NX, idx = b / NX;
    return x * q + (x < r ? x : r) + idx;
}

// Each wave works in its own LDS slab and a wavefront's LDS operations
// complete in issue order, so ordering the compiler's memory operations is all
// the synchronisation a slab needs (no s_barrier tying a block's waves).
__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A value every lane of the wave holds, moved to scalar registers.
__device__ inline int64_t wave_uniform(int64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Lane l's value of v, broadcast to the whole wave (two v_readlane_b32).
__device__ inline double readlane_f64(double v, int l)
{
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Volume index of node m of face f (faces ordered -x,+x,-y,+y,-z,+z; face
// nodes over the remaining directions, first fastest: lobattocells.jl:2020-2050).
template <int DIM, int NQ>
__device__ __host__ inline int face_node(int f, int m)
{
    const int d = f >> 1;
    const int fixed = (f & 1) ? NQ - 1 : 0;
    if constexpr (DIM == 1) {
        return fixed;
    } else if constexpr (DIM == 2) {
        return d == 0 ? fixed + NQ * m : m + NQ * fixed;
    } else {
        const int a = m % NQ, b = m / NQ;
        if (d == 0)
            return fixed + NQ * a + NQ * NQ * b;
        if (d == 1)
            return a + NQ * fixed + NQ * NQ * b;
        return a + NQ * b + NQ * NQ * fixed;
    }
}

// Face-node permutation of a conforming neighbour: bit 2 swaps the two face
// directions, then bit 0 / bit 1 reverse the first / second one.
template <int DIM, int NQ>
__device__ __host__ inline int orient_node(int o, int m)
{
    if constexpr (DIM == 1) {
        return 0;
    } else if constexpr (DIM == 2) {
        return (o & 1) ? NQ - 1 - m : m;
    } else {
        int a = m % NQ, b = m / NQ;
        if (o & 4) {
            const int t = a;
            a = b;
            b = t;
        }
        if (o & 1)
            a = NQ - 1 - a;
        if (o & 2)
            b = NQ - 1 - b;
        return a + NQ * b;
    }
}

template <int DIM, int NQ, int D>
__device__ inline int pencil_base(int pen)
{
    if constexpr (DIM == 1) {
        return 0;
    } else if constexpr (DIM == 2) {
        return D == 0 ? NQ * pen : pen;
    } else {
        if constexpr (D == 0)
            return NQ * pen;
        else if constexpr (D == 1)
            return (pen % NQ) + NQ * NQ * (pen / NQ);
        else
            return pen;
    }
}

// --------------------------------------------------------------------------
// Balance laws (euler.jl:29-68, advection.jl:18-19).
// --------------------------------------------------------------------------
template <int DIM, int LAW>
struct LawInfo {
    static constexpr int NS = LAW == LAW_ADVECTION ? 1 : DIM + 2;
};

// Primitive helpers of the Euler law: velocity and pressure.
template <int DIM, typename R>
__device__ inline void euler_prim(const R *q, R gm1, R *u, R &p, R &rinv)
{
    rinv = R(1) / q[0];
#pragma unroll
    for (int d = 0; d < DIM; ++d)
        u[d] = q[1 + d] * rinv;
    R dot = q[1] * q[1];
#pragma unroll
    for (int d = 1; d < DIM; ++d)
        dot = dot + q[1 + d] * q[1 + d];
    p = gm1 * (q[DIM + 1] - dot * (R(0.5) * rinv));
}

// Row c of Raven's Euler flux matrix F (dim x Ns): F[c, s] for all s.
template <int DIM, typename R>
__device__ inline void euler_flux_row(int c, const R *q, const R *u, R p, R *f)
{
    f[0] = q[1 + c];
#pragma unroll
    for (int b = 0; b < DIM; ++b) {
        R v = q[1 + c] * u[b];
        if (b == c)
            v = v + p;
        f[1 + b] = v;
    }
    f[DIM + 1] = u[c] * (q[DIM + 1] + p);
}

template <int DIM, int LAW, typename R>
struct Law {
    static constexpr int NS = LawInfo<DIM, LAW>::NS;
    R gm1, gamma, u[3];

    __device__ inline Law(const StageParams &P)
    {
        gamma = R(P.gamma);
        gm1 = R(P.gamma - 1.0);
        u[0] = R(P.u[0]);
        u[1] = R(P.u[1]);
        u[2] = R(P.u[2]);
    }

    // flux row c of state q
    __device__ inline void flux_row(int c, const R *q, R *f) const
    {
        if constexpr (LAW == LAW_ADVECTION) {
            f[0] = u[c] * q[0];
        } else {
            R uu[DIM], p, rinv;
            euler_prim<DIM, R>(q, gm1, uu, p, rinv);
            euler_flux_row<DIM, R>(c, q, uu, p, f);
        }
    }

    // all flux rows F[c][s]
    __device__ inline void flux_all(const R *q, R (*F)[NS]) const
    {
        if constexpr (LAW == LAW_ADVECTION) {
#pragma unroll
            for (int c = 0; c < DIM; ++c)
                F[c][0] = u[c] * q[0];
        } else {
            R uu[DIM], p, rinv;
            euler_prim<DIM, R>(q, gm1, uu, p, rinv);
#pragma unroll
            for (int c = 0; c < DIM; ++c)
                euler_flux_row<DIM, R>(c, q, uu, p, F[c]);
        }
    }

    // wavespeed along unit normal n (|n.u| + c for Euler, |n.u| for advection)
    __device__ inline R wavespeed(const R *n, const R *q) const
    {
        if constexpr (LAW == LAW_ADVECTION) {
            R s = n[0] * u[0];
#pragma unroll
            for (int d = 1; d < DIM; ++d)
                s = s + n[d] * u[d];
            return fabs(s);
        } else {
            R uu[DIM], p, rinv;
            euler_prim<DIM, R>(q, gm1, uu, p, rinv);
            R nu = n[0] * uu[0];
#pragma unroll
            for (int d = 1; d < DIM; ++d)
                nu = nu + n[d] * uu[d];
            return fabs(nu) + sqrt(gamma * p * rinv);
        }
    }

    // wavespeed along +-e_c: |u_c| + c
    __device__ inline R wavespeed_axis(int c, const R *q) const
    {
        if constexpr (LAW == LAW_ADVECTION) {
            return fabs(u[c]);
        } else {
            R uu[DIM], p, rinv;
            euler_prim<DIM, R>(q, gm1, uu, p, rinv);
            return fabs(uu[c]) + sqrt(gamma * p * rinv);
        }
    }
};

// --------------------------------------------------------------------------
// The kernel.
// --------------------------------------------------------------------------

// Loads and stores of a cell's own dq and q_out: plain (ST = 0), or streaming
// (ST = 1: non-temporal, no reuse expected in L2).  The choice is a kernel
// variant picked by the launcher from StageParams::streaming: behind a
// run-time branch the compiler merges the two accesses and drops the hint.
template <int ST, typename T>
__device__ __attribute__((always_inline)) inline T load_dq(const T *p)
{
    if constexpr (ST)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

template <int ST, typename T>
__device__ __attribute__((always_inline)) inline void store_out(T v, T *p)
{
    if constexpr (ST)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}
template <int DIM, int NQ, int LAW, typename R, int GEO, int FLUX>
struct Stage {
    static constexpr int NS = LawInfo<DIM, LAW>::NS;
    static constexpr int NP = IPow<NQ, DIM>::v;
    static constexpr int NP2 = NP / NQ;   // pencils per cell = face nodes per face
    static constexpr int EPW = 64 / NP2;  // cells per wave
    static constexpr int NF = 2 * DIM;
    static_assert(EPW >= 1, "Nq^(dim-1) must not exceed the wave size");

    // Neighbour state of face f at face node pen.
    __device__ static inline void neighbour_state(const StageParams &P, const R *qbase,
                                                  int64_t e, int f, int pen, const R *qm,
                                                  R *qp)
    {
        const FaceRec rec = P.faces[e * NF + f];
        if (rec.x >= 0) {
            const int fn = rec.y & 7, o = (rec.y >> 3) & 7;
            const int node = face_node<DIM, NQ>(fn, orient_node<DIM, NQ>(o, pen));
            const R *qn = qbase + (int64_t)rec.x * (NS * NP) + node;
#pragma unroll
            for (int s = 0; s < NS; ++s)
                qp[s] = qn[s * NP];
        } else if (rec.x == FACE_SELF) {
#pragma unroll
            for (int s = 0; s < NS; ++s)
                qp[s] = qm[s];
        } else {
            const int64_t off = P.explicit_idx[(int64_t)rec.y + pen];
#pragma unroll
            for (int s = 0; s < NS; ++s)
                qp[s] = qbase[off + s * NP];
        }
    }

    // Numerical flux for an axis-aligned unit normal sgn*e_c.
    __device__ static inline void nflux_axis(const Law<DIM, LAW, R> &law, int c, R sgn,
                                             const R *fm, const R *qm, const R *qp, R *nf)
    {
        R fp[NS];
        law.flux_row(c, qp, fp);
        if constexpr (FLUX == FLUX_RUSANOV) {
            const R wm = law.wavespeed_axis(c, qm), wp = law.wavespeed_axis(c, qp);
            const R lam = wm > wp ? wm : wp;
#pragma unroll
            for (int s = 0; s < NS; ++s)
                nf[s] = ((fm[s] + fp[s]) * R(0.5)) * sgn - (lam * (qp[s] - qm[s])) * R(0.5);
        } else {
#pragma unroll
            for (int s = 0; s < NS; ++s)
                nf[s] = ((fm[s] + fp[s]) * R(0.5)) * sgn;
        }
    }

    // Numerical flux for a general unit normal n.
    __device__ static inline void nflux_general(const Law<DIM, LAW, R> &law, const R *n,
                                                const R *qm, const R *qp, R *nf)
    {
        R Fm[DIM][NS], Fp[DIM][NS];
        law.flux_all(qm, Fm);
        law.flux_all(qp, Fp);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            R fc = ((Fm[0][s] + Fp[0][s]) * R(0.5)) * n[0];
#pragma unroll
            for (int c = 1; c < DIM; ++c)
                fc = fc + ((Fm[c][s] + Fp[c][s]) * R(0.5)) * n[c];
            nf[s] = fc;
        }
        if constexpr (FLUX == FLUX_RUSANOV) {
            const R wm = law.wavespeed(n, qm), wp = law.wavespeed(n, qp);
            const R lam = wm > wp ? wm : wp;
#pragma unroll
            for (int s = 0; s < NS; ++s)
                nf[s] = nf[s] - (lam * (qp[s] - qm[s])) * R(0.5);
        }
    }

    // One direction sweep; directions meet in the cell's LDS slab.  ST
    // selects streaming accesses for dq and q_out (load_dq / store_out).
    template <int D, int ST>
    __device__ static inline void sweep(const StageParams &P, const Law<DIM, LAW, R> &law,
                                        bool active, int64_t e, int pen, R *slab)
    {
        constexpr int STRIDE = IPow<NQ, D>::v;
        constexpr bool FIRST = (D == 0);
        constexpr bool LAST = (D == DIM - 1);
        const int base = pencil_base<DIM, NQ, D>(pen);
        const R *qbase = static_cast<const R *>(P.q_in);
        const bool on = active && ((P.dir_mask >> D) & 1);

        R F[NS][NQ];     // flux entering the derivative (contravariant for GENERAL)
        R nf[2][NS];     // numerical flux at the two face nodes
        R liftw[2];      // lift weight of the two face nodes
        R scale = R(1);  // AFFINE: dRdX[D,D] of the cell

        if (on) {
            const R *qe = qbase + e * (NS * NP) + base;
            R qv[NS][NQ];
#pragma unroll
            for (int n = 0; n < NQ; ++n)
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    qv[s][n] = qe[s * NP + n * STRIDE];
            // neighbour states first: their latency hides behind the flux work
            R qp[2][NS];
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                R qm[NS];
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    qm[s] = qv[s][side ? NQ - 1 : 0];
                neighbour_state(P, qbase, e, 2 * D + side, pen, qm, qp[side]);
            }
#pragma unroll
            for (int n = 0; n < NQ; ++n) {
                R qn[NS];
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    qn[s] = qv[s][n];
                if constexpr (GEO == GEO_AFFINE) {
                    R f[NS];
                    law.flux_row(D, qn, f);
#pragma unroll
                    for (int s = 0; s < NS; ++s)
                        F[s][n] = f[s];
                } else {
                    R Fa[DIM][NS];
                    law.flux_all(qn, Fa);
                    const double *g = P.vgeo + ((e * DIM + D) * DIM) * NP + base + n * STRIDE;
                    R gc[DIM];
#pragma unroll
                    for (int c = 0; c < DIM; ++c)
                        gc[c] = R(g[c * NP]);
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        R v = gc[0] * Fa[0][s];
#pragma unroll
                        for (int c = 1; c < DIM; ++c)
                            v = v + gc[c] * Fa[c][s];
                        F[s][n] = v;
                    }
                }
            }
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const int iend = side ? NQ - 1 : 0;
                R qm[NS];
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    qm[s] = qv[s][iend];
                if constexpr (GEO == GEO_AFFINE) {
                    R fm[NS];
#pragma unroll
                    for (int s = 0; s < NS; ++s)
                        fm[s] = F[s][iend];
                    nflux_axis(law, D, side ? R(1) : R(-1), fm, qm, qp[side], nf[side]);
                    liftw[side] = R(P.winv[iend]);
                } else {
                    const double *sg = P.sgeo + ((e * NF + 2 * D + side) * (DIM + 1)) * NP2 + pen;
                    R nv[DIM];
#pragma unroll
                    for (int c = 0; c < DIM; ++c)
                        nv[c] = R(sg[c * NP2]);
                    liftw[side] = R(sg[DIM * NP2]);
                    nflux_general(law, nv, qm, qp[side], nf[side]);
                }
            }
            if constexpr (GEO == GEO_AFFINE)
                scale = R(P.affine[e * DIM + D]);
        }

        // state by state: derivative, lift, combine
        const double *M = (GEO == GEO_AFFINE) ? P.Dw : P.Dt;
        R *dq = static_cast<R *>(P.dq) + e * (NS * NP) + base;
        const R *qe = qbase + e * (NS * NP) + base;
        R *qo = static_cast<R *>(P.q_out) + e * (NS * NP) + base;
        const R a = R(P.a), bdt = R(P.b_dt);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            R acc[NQ];
#pragma unroll
            for (int i = 0; i < NQ; ++i)
                acc[i] = R(0);
            if (on) {
#pragma unroll
                for (int i = 0; i < NQ; ++i) {
                    R v = R(M[i * NQ]) * F[s][0];
#pragma unroll
                    for (int n = 1; n < NQ; ++n)
                        v = v + R(M[i * NQ + n]) * F[s][n];
                    acc[i] = v;
                }
                acc[0] = acc[0] - liftw[0] * nf[0][s];
                acc[NQ - 1] = acc[NQ - 1] - liftw[1] * nf[1][s];
                if constexpr (GEO == GEO_AFFINE) {
#pragma unroll
                    for (int i = 0; i < NQ; ++i)
                        acc[i] = scale * acc[i];
                }
            }
            if (!active)
                continue;
            if constexpr (!LAST) {
#pragma unroll
                for (int n = 0; n < NQ; ++n) {
                    R *dst = slab + s * NP + base + n * STRIDE;
                    if constexpr (FIRST)
                        *dst = acc[n];
                    else
                        *dst = *dst + acc[n];
                }
            } else {
#pragma unroll
                for (int n = 0; n < NQ; ++n) {
                    R tot = acc[n];
                    if constexpr (DIM > 1)
                        tot = slab[s * NP + base + n * STRIDE] + tot;
                    if constexpr (GEO == GEO_GENERAL)
                        tot = R(P.mji[e * NP + base + n * STRIDE]) * tot;
                    // dq and q_out belong to this cell alone: on operators
                    // larger than the caches they are streamed, which keeps
                    // L2 for the shared q lines
                    R v = tot;
                    if (P.read_dq)
                        v = a * load_dq<ST>(dq + s * NP + n * STRIDE) + tot;
                    store_out<ST>(v, dq + s * NP + n * STRIDE);
                    if (P.write_q)
                        store_out<ST>(R(qe[s * NP + n * STRIDE] + bdt * v), qo + s * NP + n * STRIDE);
                }
            }
        }
    }
};

// The cells [cell0, cell0 + EPW) of one wave (cell0 is the same on every
// lane): the direction sweeps meet in the wave's own LDS slabs `wlds`
// (EPW * Ns * Np values).
template <int DIM, int NQ, int LAW, typename R, int GEO, int FLUX, int ST = 0>
__device__ __attribute__((always_inline)) inline void stage_wave(const StageParams &P, int64_t cell0, R *wlds)
{
    using S = Stage<DIM, NQ, LAW, R, GEO, FLUX>;
    constexpr int SLAB = S::NS * S::NP;
    const int lane = threadIdx.x & 63;
    const int slot = lane / S::NP2;
    const int pen = lane - slot * S::NP2;
    const int64_t e = cell0 + slot;
    const bool active = (slot < S::EPW) && (e < P.K);
    R *slab = wlds + (slot < S::EPW ? slot : 0) * SLAB;
    const Law<DIM, LAW, R> law(P);

    S::template sweep<0, ST>(P, law, active, e, pen, slab);
    if constexpr (DIM > 1) {
        wave_lds_sync();
        S::template sweep<1, ST>(P, law, active, e, pen, slab);
    }
    if constexpr (DIM > 2) {
        wave_lds_sync();
        S::template sweep<2, ST>(P, law, active, e, pen, slab);
    }
}

// Waves per block (cells of one block share nothing; the knob trades LDS
// granularity against scheduling) come from blockDim.x; the LDS slabs are
// dynamic shared memory sized by the launcher.
template <int DIM, int NQ, int LAW, typename R, int GEO, int FLUX, int ST>
__global__ void __launch_bounds__(256) dg_stage_kernel(const StageParams P)
{
    using S = Stage<DIM, NQ, LAW, R, GEO, FLUX>;
    extern __shared__ __attribute__((aligned(16))) unsigned char dg_lds[];
    const int wave = threadIdx.x >> 6;
    const int wpb = blockDim.x >> 6;
    const int64_t blk = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t cell0 = (blk * wpb + wave) * S::EPW;
    R *wlds = reinterpret_cast<R *>(dg_lds) + wave * S::EPW * S::NS * S::NP;
    stage_wave<DIM, NQ, LAW, R, GEO, FLUX, ST>(P, cell0, wlds);
}

template <int DIM, int NQ, int LAW, typename R, int GEO, int FLUX>
int launch_stage(const StageParams &P, void *stream, int wpb)
{
    using S = Stage<DIM, NQ, LAW, R, GEO, FLUX>;
    if (P.K <= 0)
        return 0;
    if (wpb != 1 && wpb != 2 && wpb != 4)
        wpb = 1;
    const int64_t waves = (P.K + S::EPW - 1) / S::EPW;
    const int64_t blocks = (waves + wpb - 1) / wpb;
    const size_t lds = DIM > 1 ? (size_t)wpb * S::EPW * S::NS * S::NP * sizeof(R) : 16;
    if (P.streaming)
        hipLaunchKernelGGL((dg_stage_kernel<DIM, NQ, LAW, R, GEO, FLUX, 1>), dim3((unsigned)blocks),
                           dim3(64 * wpb), lds, static_cast<hipStream_t>(stream), P);
    else
        hipLaunchKernelGGL((dg_stage_kernel<DIM, NQ, LAW, R, GEO, FLUX, 0>), dim3((unsigned)blocks),
                           dim3(64 * wpb), lds, static_cast<hipStream_t>(stream), P);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace rv
