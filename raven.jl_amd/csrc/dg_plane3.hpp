// Fused 3-D DGSEM stage kernel for gfx950: "column" decomposition.
//
// Same operator as dg_kernels.hpp (Raven's weak-form volume term, Rusanov /
// central surface term, LSRK update: dgsem.jl:257-348, 733-784,
// odesolvers.jl:55-57), organised so that every state value crosses HBM once:
//
//  * a wave owns EPW = 64 / Nq^2 cells; lane (i, j) of a cell holds the column
//    q(i, j, 0..Nq-1) of all states in registers (loaded with lane-contiguous
//    addresses) together with its accumulators;
//  * the k loop visits the cell's planes: the lane evaluates the flux at (i,j,k),
//    scatters the z-derivative contribution into its own accumulators
//    (acc[kk] += Dz[k][kk] f3), and publishes f1, f2 in a double-buffered LDS
//    plane from which every lane gathers its x- and y-derivatives;
//  * faces: the z faces sit in each lane's own column; the four x / y faces
//    (4 Nq^2 face nodes) are spread over the lanes, one node of each face per
//    lane, with q- and the lift exchanged through LDS;
//  * the accumulators start from a*dq_old (affine cells) so dq_old and q load
//    together at the top; dq and q_out are stored at the end; dq_old, dq and
//    q_out are touched by this cell alone, so on operators larger than the
//    caches (P.streaming: the kernel variant ST = 1) they go through L2 as
//    streaming (non-temporal)
//    accesses and leave it to the q lines that the neighbours' face reads hit
//    (measured on the 32^3 Nq = 8 Euler brick: 7% faster, 17% fewer bytes
//    fetched); smaller operators keep them cached for the next stage;
//  * every load of a cell is issued before its first wait; the face records
//    come in with the state (scalar loads when a wave owns one cell); the
//    x-face neighbour states are read with the state and parked in LDS, the
//    y- and z-face ones are requested together after the plane loop.
#pragma once

#include "dg_kernels.hpp"
#include "entropy.hpp"

namespace rv {

template <int NQ, int LAW, typename R, int GEO, int FLUX>
struct Plane3 {
    static constexpr int DIM = 3;
    static constexpr int NS = LawInfo<3, LAW>::NS;
    static constexpr int NQ2 = NQ * NQ;
    static constexpr int NP = NQ2 * NQ;
    static constexpr int EPW = 64 / NQ2;
    // plane rows are read 16 bytes at a time when they can be: VEC values per
    // read, rows padded to a multiple of 16 bytes whose 32-bit bank offsets
    // differ between the Nq rows a wave reads at once
    static constexpr int VEC = (16 / (int)sizeof(R));
    static constexpr bool VECREAD = (NQ % VEC) == 0;
    static constexpr bool rows_disjoint(int pad)
    {
        // 16-byte blocks of the NQ rows (32 banks of 4 bytes) must not overlap
        const int w = (int)sizeof(R) / 4;
        for (int a = 0; a < NQ; ++a)
            for (int b = a + 1; b < NQ; ++b) {
                const int d = ((pad * w * (b - a)) % 32 + 32) % 32;
                if (d < 4 || d > 28)
                    return false;
            }
        return true;
    }
    static constexpr int pick_pad()
    {
        if (!VECREAD)
            return NQ + 1;
        for (int p = NQ; p < NQ + 64; p += VEC)
            if (rows_disjoint(p))
                return p;
        return NQ + VEC;
    }
    static constexpr int PAD = pick_pad();
    static constexpr int PLANE = NS * NQ * PAD;      // one flux component of one plane
    static constexpr int PBUF = 2 * PLANE;           // f1 and f2 of one plane
    // q- / lift of the four x,y faces, then a spare face (lanes off the x or
    // y faces store to it and subtract it, zeroed, without branches)
    static constexpr int FACEBUF = 5 * NS * NQ2;
    static constexpr int SLAB0 = (2 * PBUF > FACEBUF) ? 2 * PBUF : FACEBUF;
    static constexpr int XFB = 2 * NS * NQ2;         // x-face neighbour states, read early
    // FP64 general cells keep their x / y derivative weights in LDS (the
    // column, the accumulators and two planes of metric terms fill the
    // registers); the others in registers
    static constexpr bool WLDS = GEO == GEO_GENERAL;
    static constexpr int CTAB = WLDS ? 2 * NQ2 : 0;
    static constexpr int SLAB = SLAB0 + XFB + CTAB;
    static_assert(EPW >= 1, "Nq^2 must not exceed the wave size");
    // Waves that share their cells: Nq^2 of 25, 36 or 49 columns leave 14 to
    // 44 % of a wave's lanes idle when cells do not straddle waves, so a
    // workgroup of G waves packs CPG cells across its 64 G lanes instead
    // (125 of 128, 252 of 256, 245 of 256 lanes busy) and synchronises its
    // LDS slabs with workgroup barriers.
    static constexpr int G = (NQ == 5) ? 2 : (NQ == 6 || NQ == 7) ? 4 : 1;
    static constexpr int CPG = 64 * G / NQ2;
    using S = Stage<3, NQ, LAW, R, GEO, FLUX>;
};

// q+ at node `pen` of a face from its neighbour record (conforming neighbour or
// explicit offset); self-mapped faces (FACE_SELF) are filled from q- later.
template <int NQ, int NS, typename R>
__device__ inline void load_neighbour(const StageParams &P, const R *qbase, const FaceRec rec,
                                      int pen, R *qp)
{
    constexpr int NP = NQ * NQ * NQ;
    if (rec.x >= 0) {
        const int fn = rec.y & 7, o = (rec.y >> 3) & 7;
        const int node = face_node<3, NQ>(fn, orient_node<3, NQ>(o, pen));
        const R *qn = qbase + (int64_t)rec.x * (NS * NP) + node;
#pragma unroll
        for (int s = 0; s < NS; ++s)
            qp[s] = qn[s * NP];
    } else if (rec.x == FACE_EXPLICIT) {
        const int64_t off = P.explicit_idx[(int64_t)rec.y + pen];
#pragma unroll
        for (int s = 0; s < NS; ++s)
            qp[s] = qbase[off + s * NP];
    }
}

// The cells [cell0, cell0 + EPW) of one wave (cell0 is the same on every
// lane), working in the wave's own LDS slabs `wlds` (EPW * SLAB values); with
// GW > 1, the cells [cell0, cell0 + CPG) of a workgroup of GW waves (the
// workgroup's first 64 GW threads) and their CPG slabs.
// RX = 1: the fused relaxation stage (rv_rlsrk_step).  The accumulators form
// k = L(q_in) alone; the store loop then makes dq = a dq_old + k and
// q_out = q_in + b_dt dq as usual, and the wave's sum of MJ v(q_in) . k
// (TwoSum-compensated, v the entropy variables) goes to P.eprod[eslot].
template <int NQ, int LAW, typename R, int GEO, int FLUX, int ST = 0, int GW = 1, int RX = 0>
__device__ __attribute__((always_inline)) inline void plane3_wave(const StageParams &P, int64_t cell0, R *wlds,
                                                                  int64_t eslot = 0)
{
    using T = Plane3<NQ, LAW, R, GEO, FLUX>;
    using S = typename T::S;
    constexpr int NS = T::NS, NQ2 = T::NQ2, NP = T::NP, PAD = T::PAD;
    constexpr int PLANE = T::PLANE, PBUF = T::PBUF;

    constexpr int CELLS = GW == 1 ? T::EPW : 64 * GW / NQ2;
    const int lane = threadIdx.x & 63;
    const int gl = GW == 1 ? lane : (int)threadIdx.x;  // position in the wave / workgroup
    const int slot = gl / NQ2;
    const int l = gl - slot * NQ2;  // column index = i + NQ*j
    const int i = l % NQ, j = l / NQ;
    int64_t e = cell0 + slot;
    if constexpr (T::EPW == 1 && GW == 1)
        e = wave_uniform(e);  // one cell per wave: per-cell data comes in scalar loads
    const bool active = (slot < CELLS) && (e < P.K);
    R *slab = wlds + (slot < CELLS ? slot : 0) * T::SLAB;
    // slabs shared by the waves of a workgroup need its barrier
    auto lds_sync = [] {
        if constexpr (GW == 1)
            wave_lds_sync();
        else
            __syncthreads();
    };
    const Law<3, LAW, R> law(P);
    const R *qbase = static_cast<const R *>(P.q_in);
    // general cells' geometry: FP32 state reads the FP32 copies (every use
    // rounds to R, so the values are the ones the FP64 arrays give)
    constexpr bool G32 = GEO == GEO_GENERAL && sizeof(R) == 4;
    using GT = typename std::conditional<G32, float, double>::type;
    const GT RV_GLOBAL *vgeo_p, *sgeo_p, *mji_p;
    const GT RV_GLOBAL *mj_p;  // MJ for the a*MJ factor of dq
    if constexpr (G32) {
        vgeo_p = P.vgeo32, sgeo_p = P.sgeo32, mji_p = P.mji32, mj_p = P.mj32;
    } else {
        vgeo_p = P.vgeo, sgeo_p = P.sgeo, mji_p = P.mji, mj_p = P.mj;
    }
    const bool dx_on = P.dir_mask & 1, dy_on = (P.dir_mask >> 1) & 1,
               dz_on = (P.dir_mask >> 2) & 1;

    R q[NS][NQ], acc[NS][NQ];
    R xq[2][NS];       // x-face neighbour states until they are parked in LDS
    R cx[T::WLDS ? 1 : NQ], cy[T::WLDS ? 1 : NQ];  // this lane's x / y derivative weights
    R *ctab = slab + T::SLAB0 + T::XFB;  // with WLDS: [output node][n], x rows then y rows
    R sz = R(1);
    FaceRec rec[6];
    // cells with non-conforming faces (mortar_face_kernel ran before this
    // stage): their slot in the mortar lift buffer and which faces those are
    int mli = -1;
    unsigned mmask = 0;
    double prod = 0.0, pc = 0.0;  // RX: the lane's MJ v(q) . k and its compensation
    const R a = R(P.a), bdt = R(P.b_dt);
    // a*dq_old is formed at the first z scatter so that the dq loads need not
    // complete before the first plane's flux; general cells accumulate
    // MJ * dq (the final update multiplies by MJI), so their factor is a*MJ
    // the relaxation stage accumulates k alone and reads dq_old at the store
    const bool rdq = RX ? false : P.read_dq != 0;
    const bool late_scale = rdq && dz_on;
    R zscale[NQ];  // general cells: a * MJ of the column's nodes

    // the z-derivative weights M[kk*Nq + k], one per lane, broadcast with
    // v_readlane inside the plane loop (no scalar loads or waits there)
    const double ztab = ((GEO == GEO_AFFINE) ? P.Dw : P.Dt)[lane % NQ2];
    if (active) {
        // every load of the cell is issued before the first wait: derivative
        // weights, geometry factors, face records, then q and dq
        const double *M = (GEO == GEO_AFFINE) ? P.Dw : P.Dt;
        double mi[T::WLDS ? 1 : NQ], mj[T::WLDS ? 1 : NQ];
        if constexpr (!T::WLDS) {
#pragma unroll
            for (int n = 0; n < NQ; ++n) {
                mi[n] = M[i * NQ + n];
                mj[n] = M[j * NQ + n];
            }
        }
        R sx = R(1), sy = R(1);
        if constexpr (GEO == GEO_AFFINE) {
            sx = R(P.affine[e * 3 + 0]);
            sy = R(P.affine[e * 3 + 1]);
            sz = R(P.affine[e * 3 + 2]);
        }
#pragma unroll
        for (int f = 0; f < 6; ++f)
            rec[f] = P.faces[e * 6 + f];
        if constexpr (GEO == GEO_GENERAL) {
            if (P.mslot) {
                mli = P.mslot[e];
                unsigned mk = 0;
#pragma unroll
                for (int f = 0; f < 6; ++f)
                    mk |= (P.mkind[e * 6 + f] != 0 ? 1u : 0u) << f;
                mmask = mli >= 0 ? mk : 0u;
            }
        }
        const R *qe = qbase + e * (NS * NP) + l;
#pragma unroll
        for (int k = 0; k < NQ; ++k)
#pragma unroll
            for (int s = 0; s < NS; ++s)
                q[s][k] = qe[s * NP + k * NQ2];
        const R *dqe = static_cast<const R *>(P.dq) + e * (NS * NP) + l;
        if (rdq) {
#pragma unroll
            for (int k = 0; k < NQ; ++k)
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    acc[s][k] = load_dq<ST>(dqe + s * NP + k * NQ2);
#pragma unroll
            for (int k = 0; k < NQ; ++k)
                zscale[k] = GEO == GEO_AFFINE ? a : a * R(mj_p[e * NP + l + k * NQ2]);
            if (!late_scale) {
#pragma unroll
                for (int k = 0; k < NQ; ++k)
#pragma unroll
                    for (int s = 0; s < NS; ++s)
                        acc[s][k] = zscale[k] * acc[s][k];
            }
        } else {
#pragma unroll
            for (int k = 0; k < NQ; ++k)
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    acc[s][k] = R(0);
        }
        // the x neighbours are processed at about the same time (they are
        // next to this cell in the forest order), so their face nodes are
        // read now, while those lines are still in L2; an x face touches every
        // cache line of the neighbour's state
        if (dx_on) {
            load_neighbour<NQ, NS, R>(P, qbase, rec[0], l, xq[0]);
            load_neighbour<NQ, NS, R>(P, qbase, rec[1], l, xq[1]);
        }
        const R fx = dx_on ? sx : R(0), fy = dy_on ? sy : R(0);
        if constexpr (T::WLDS) {
            // column l stores entry l of the operator (ztab is entry
            // lane % Nq^2, another one where cells straddle waves)
            const double ml = M[l];
            ctab[l] = fx * R(ml);
            ctab[NQ2 + l] = fy * R(ml);
        } else {
#pragma unroll
            for (int n = 0; n < NQ; ++n) {
                cx[n] = fx * R(mi[n]);
                cy[n] = fy * R(mj[n]);
            }
        }
    }

    // ---- plane loop: volume term, software-pipelined: the flux of plane
    // k+1 is evaluated and published while plane k's gather reads LDS
    struct alignas(16) Vec {
        R v[T::VEC];
    };
    // general cells: wJ*dRdX of plane k in gbuf[k & 1], requested two planes
    // ahead (one plane's work is shorter than a load's latency)
    GT gbuf[2][3][3];
    auto load_geo = [&](const int k) {
        if constexpr (GEO == GEO_GENERAL) {
            const int node = l + k * NQ2;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    gbuf[k & 1][d][c] = vgeo_p[((e * 3 + d) * 3 + c) * NP + node];
        }
    };
    auto plane_flux = [&](const int k) {
        R *buf = slab + (k & 1) * PBUF;
        R qn[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s)
            qn[s] = q[s][k];
        R f1[NS], f2[NS], f3[NS];
        R Fa[3][NS];
        law.flux_all(qn, Fa);
        if constexpr (GEO == GEO_AFFINE) {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                f1[s] = Fa[0][s];
                f2[s] = Fa[1][s];
                f3[s] = sz * Fa[2][s];
            }
        } else {
            R g[3][3];
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    g[d][c] = R(gbuf[k & 1][d][c]);
            if (k + 2 < NQ)
                load_geo(k + 2);
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                f1[s] = g[0][0] * Fa[0][s] + g[0][1] * Fa[1][s] + g[0][2] * Fa[2][s];
                f2[s] = g[1][0] * Fa[0][s] + g[1][1] * Fa[1][s] + g[1][2] * Fa[2][s];
                f3[s] = g[2][0] * Fa[0][s] + g[2][1] * Fa[1][s] + g[2][2] * Fa[2][s];
            }
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            buf[s * NQ * PAD + j * PAD + i] = f1[s];          // row j, entry i
            buf[PLANE + s * NQ * PAD + i * PAD + j] = f2[s];  // row i, entry j
        }
        if (dz_on) {
#pragma unroll
            for (int kk = 0; kk < NQ; ++kk) {
                // contribution of plane k to output kk
                const R w = R(readlane_f64(ztab, kk * NQ + k));
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    R base = acc[s][kk];
                    if (k == 0 && late_scale)
                        base = zscale[kk] * base;
                    acc[s][kk] = base + w * f3[s];
                }
            }
        }
    };
    auto gather = [&](const int k) {
        const R *buf = slab + (k & 1) * PBUF;
        if constexpr (T::WLDS) {
            // weights from LDS, read once for the five states; per state the
            // same order of terms as below (x then y terms of each VEC block)
            const R *cxr = ctab + i * NQ, *cyr = ctab + NQ2 + j * NQ;
            if constexpr (T::VECREAD) {
#pragma unroll
                for (int t = 0; t < NQ / T::VEC; ++t) {
                    const Vec c1 = reinterpret_cast<const Vec *>(cxr)[t];
                    const Vec c2 = reinterpret_cast<const Vec *>(cyr)[t];
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        const Vec x1 = reinterpret_cast<const Vec *>(buf + s * NQ * PAD + j * PAD)[t];
                        const Vec x2 =
                            reinterpret_cast<const Vec *>(buf + PLANE + s * NQ * PAD + i * PAD)[t];
                        R v = acc[s][k];
#pragma unroll
                        for (int c = 0; c < T::VEC; ++c)
                            v = v + c1.v[c] * x1.v[c];
#pragma unroll
                        for (int c = 0; c < T::VEC; ++c)
                            v = v + c2.v[c] * x2.v[c];
                        acc[s][k] = v;
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            } else {
#pragma unroll
                for (int n = 0; n < NQ; ++n) {
                    const R c1 = cxr[n];
#pragma unroll
                    for (int s = 0; s < NS; ++s)
                        acc[s][k] = acc[s][k] + c1 * buf[s * NQ * PAD + j * PAD + n];
                }
#pragma unroll
                for (int n = 0; n < NQ; ++n) {
                    const R c2 = cyr[n];
#pragma unroll
                    for (int s = 0; s < NS; ++s)
                        acc[s][k] = acc[s][k] + c2 * buf[PLANE + s * NQ * PAD + i * PAD + n];
                }
            }
            return;
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const R *r1 = buf + s * NQ * PAD + j * PAD;
            const R *r2 = buf + PLANE + s * NQ * PAD + i * PAD;
            R v = acc[s][k];
            if constexpr (T::VECREAD) {
#pragma unroll
                for (int t = 0; t < NQ / T::VEC; ++t) {
                    const Vec x1 = reinterpret_cast<const Vec *>(r1)[t];
                    const Vec x2 = reinterpret_cast<const Vec *>(r2)[t];
#pragma unroll
                    for (int c = 0; c < T::VEC; ++c)
                        v = v + cx[t * T::VEC + c] * x1.v[c];
#pragma unroll
                    for (int c = 0; c < T::VEC; ++c)
                        v = v + cy[t * T::VEC + c] * x2.v[c];
                }
            } else {
#pragma unroll
                for (int n = 0; n < NQ; ++n)
                    v = v + cx[n] * r1[n];
#pragma unroll
                for (int n = 0; n < NQ; ++n)
                    v = v + cy[n] * r2[n];
            }
            acc[s][k] = v;
        }
    };
    if (active) {
        load_geo(0);
        // FP64 general cells request plane 1 once plane 0 is done (with the
        // column, the register and the x-face states all in flight at the
        // top, a second plane of metric terms would not fit)
        if (NQ > 1 && !T::WLDS)
            load_geo(1);
        plane_flux(0);
        if (NQ > 1 && T::WLDS)
            load_geo(1);
    }
    if (active && dx_on) {
        R *xf = slab + T::SLAB0;
#pragma unroll
        for (int f = 0; f < 2; ++f)
#pragma unroll
            for (int s = 0; s < NS; ++s)
                xf[(f * NS + s) * NQ2 + l] = xq[f][s];
    }
    // ---- neighbour states of the y and z faces, issued together: node l of
    // each y face (one node per lane) and the ends of the lane's own column
    // for the two z faces (the x faces come from LDS).  General cells request
    // them before the last plane's gather, in the registers the metric terms
    // no longer need, together with the z faces' normals and weights; affine
    // cells after the plane loop
    R qp[6][NS];
    GT zsg[2][4];  // general cells: n and wsJ of this lane's z-face nodes
    auto face_loads = [&] {
#pragma unroll
        for (int f = 2; f < 6; ++f)
            if ((P.dir_mask >> (f >> 1)) & 1)
                load_neighbour<NQ, NS, R>(P, qbase, rec[f], l, qp[f]);
        if constexpr (GEO == GEO_GENERAL) {
#pragma unroll
            for (int side = 0; side < 2; ++side)
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    zsg[side][c] = sgeo_p[((e * 6 + 4 + side) * 4 + c) * NQ2 + l];
        }
    };
    lds_sync();
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        if (active) {
            if (k + 1 < NQ)
                plane_flux(k + 1);
            if (GEO == GEO_GENERAL && k == NQ - 1)
                face_loads();
            gather(k);
        }
        lds_sync();
    }
    if (GEO == GEO_AFFINE && active)
        face_loads();

    // ---- x and y faces: owners publish q-, every lane evaluates one node of
    // each of the four faces, owners collect the lift.  A lane's x slot is
    // face 0 or 1 when it sits on one (i == 0, i == Nq-1) and the spare face
    // otherwise, likewise its y slot: every lane stores and later subtracts
    // both slots without branches, the spare face being zeroed before the
    // lifts are subtracted
    R *fb = slab;  // [face 0..3, spare][s][face node]
    constexpr int SPARE = 4 * NS * NQ2;
    const int xslot = (dx_on && i == 0) ? 0 : (dx_on && i == NQ - 1) ? NS * NQ2 : SPARE;
    const int yslot = (dy_on && j == 0) ? 2 * NS * NQ2 : (dy_on && j == NQ - 1) ? 3 * NS * NQ2
                                                                               : SPARE;
    R *xs = fb + xslot + j, *ys = fb + yslot + i;
    // FP32 general cells: n and wsJ of this lane's node on each x / y face,
    // requested before the z faces are evaluated (FP64 has no registers to
    // spare for them and reads them where the faces are evaluated)
    constexpr bool XYPRE = GEO == GEO_GENERAL && sizeof(R) == 4;
    GT xysg[XYPRE ? 4 : 1][4];
    if (active) {
        if constexpr (XYPRE) {
#pragma unroll
            for (int f = 0; f < 4; ++f)
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    xysg[f][c] = sgeo_p[((e * 6 + f) * 4 + c) * NQ2 + l];
        }
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                xs[s * NQ2 + NQ * k] = q[s][k];
                ys[s * NQ2 + NQ * k] = q[s][k];
            }
        }
        // z faces, in the lane's own column
        if (dz_on) {
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const int kend = side ? NQ - 1 : 0;
                R qm[NS], nf[NS];
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    qm[s] = q[s][kend];
                if (rec[4 + side].x == FACE_SELF) {
#pragma unroll
                    for (int s = 0; s < NS; ++s)
                        qp[4 + side][s] = qm[s];
                }
                R lift;
                if constexpr (GEO == GEO_AFFINE) {
                    R fm[NS];
                    law.flux_row(2, qm, fm);
                    S::nflux_axis(law, 2, side ? R(1) : R(-1), fm, qm, qp[4 + side], nf);
                    lift = sz * R(P.winv[kend]);
                } else {
                    R nv[3] = {R(zsg[side][0]), R(zsg[side][1]), R(zsg[side][2])};
                    lift = R(zsg[side][3]);
                    S::nflux_general(law, nv, qm, qp[4 + side], nf);
                }
                R lv[NS];
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    lv[s] = lift * nf[s];
                if constexpr (GEO == GEO_GENERAL) {
                    // a non-conforming face: its mortar lift instead
                    if ((mmask >> (4 + side)) & 1u) {
                        const R *L = static_cast<const R *>(P.mdelta) +
                                     ((int64_t)mli * 6 + 4 + side) * (NS * NQ2) + l;
#pragma unroll
                        for (int s = 0; s < NS; ++s)
                            lv[s] = L[s * NQ2];
                    }
                }
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    acc[s][kend] = acc[s][kend] - lv[s];
            }
        }
    }
    lds_sync();
    if (active) {
        // the spare face reads as zero when the lifts are subtracted
#pragma unroll
        for (int s = 0; s < NS; ++s)
            fb[SPARE + s * NQ2 + l] = R(0);
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const int d = f >> 1;
            if (!((P.dir_mask >> d) & 1))
                continue;
            R qm[NS], nf[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s)
                qm[s] = fb[(f * NS + s) * NQ2 + l];
            if (f < 2) {
                // x-face neighbour states parked in LDS at the top
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    qp[f][s] = slab[T::SLAB0 + (f * NS + s) * NQ2 + l];
            }
            if (rec[f].x == FACE_SELF) {
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    qp[f][s] = qm[s];
            }
            R lift;
            if constexpr (GEO == GEO_AFFINE) {
                R fm[NS];
                law.flux_row(d, qm, fm);
                S::nflux_axis(law, d, (f & 1) ? R(1) : R(-1), fm, qm, qp[f], nf);
                lift = R(P.affine[e * 3 + d]) * R(P.winv[(f & 1) ? NQ - 1 : 0]);
            } else if constexpr (XYPRE) {
                R nv[3] = {R(xysg[f][0]), R(xysg[f][1]), R(xysg[f][2])};
                lift = R(xysg[f][3]);
                S::nflux_general(law, nv, qm, qp[f], nf);
            } else {
                const GT *sg = sgeo_p + ((e * 6 + f) * 4) * NQ2 + l;
                R nv[3] = {R(sg[0]), R(sg[NQ2]), R(sg[2 * NQ2])};
                lift = R(sg[3 * NQ2]);
                S::nflux_general(law, nv, qm, qp[f], nf);
            }
            R lv[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s)
                lv[s] = lift * nf[s];
            if constexpr (GEO == GEO_GENERAL) {
                // a non-conforming face: its mortar lift instead
                if ((mmask >> f) & 1u) {
                    const R *L = static_cast<const R *>(P.mdelta) +
                                 ((int64_t)mli * 6 + f) * (NS * NQ2) + l;
#pragma unroll
                    for (int s = 0; s < NS; ++s)
                        lv[s] = L[s * NQ2];
                }
            }
#pragma unroll
            for (int s = 0; s < NS; ++s)
                fb[(f * NS + s) * NQ2 + l] = lv[s];
        }
    }
    lds_sync();
    if (active) {
        // ---- lifts, update and store, one plane at a time (the scheduler
        // may not hoist later planes' LDS reads past the barrier: with the
        // column, the accumulators and the reads of every plane in flight
        // at once a wave would run out of registers)
        R *dqo = static_cast<R *>(P.dq) + e * (NS * NP) + l;
        R *qo = static_cast<R *>(P.q_out) + e * (NS * NP) + l;
        // general cells: 1 / (w J) of the whole column, requested at once
        R mjiv[NQ];
#pragma unroll
        for (int k = 0; k < NQ; ++k)
            mjiv[k] = GEO == GEO_GENERAL ? R(mji_p[e * NP + l + k * NQ2]) : R(1);
        if constexpr (RX) {
            const R *dqe = static_cast<const R *>(P.dq) + e * (NS * NP) + l;
            const double *mje = P.mj + e * NP + l;
            // dq_old and MJ of a plane are requested before its sums; affine
            // cells request them one plane ahead (general cells have no
            // registers to spare for that)
            constexpr bool AHEAD = GEO == GEO_AFFINE;
            R dnext[NS];
            double mjnext = 0.0;
            if constexpr (AHEAD) {
                mjnext = mje[0];
                if (P.read_dq) {
#pragma unroll
                    for (int s = 0; s < NS; ++s)
                        dnext[s] = load_dq<ST>(dqe + s * NP);
                }
            }
#pragma unroll
            for (int k = 0; k < NQ; ++k) {
                const R mji = mjiv[k];
                R dcur[NS];
                double mjk;
                if constexpr (AHEAD) {
#pragma unroll
                    for (int s = 0; s < NS; ++s)
                        dcur[s] = dnext[s];
                    mjk = mjnext;
                    if (k + 1 < NQ) {
                        if (P.read_dq) {
#pragma unroll
                            for (int s = 0; s < NS; ++s)
                                dnext[s] = load_dq<ST>(dqe + s * NP + (k + 1) * NQ2);
                        }
                        mjnext = mje[(k + 1) * NQ2];
                    }
                } else {
                    if (P.read_dq) {
#pragma unroll
                        for (int s = 0; s < NS; ++s)
                            dcur[s] = load_dq<ST>(dqe + s * NP + k * NQ2);
                    }
                    mjk = mje[k * NQ2];
                }
                R kv[NS];
                double x[NS], kk[NS];
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    R v = acc[s][k];
                    v = v - xs[s * NQ2 + NQ * k];
                    v = v - ys[s * NQ2 + NQ * k];
                    if constexpr (GEO == GEO_GENERAL)
                        v = mji * v;
                    kv[s] = v;
                    x[s] = (double)q[s][k];
                    kk[s] = (double)v;
                }
                comp_add(prod, pc, mjk * law_entropy_dot_3d<LAW>(x, kk, P.gamma), 0.0);
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    R v = kv[s];
                    if (P.read_dq)
                        v = a * dcur[s] + v;
                    store_out<ST>(v, dqo + s * NP + k * NQ2);
                    store_out<ST>(R(q[s][k] + bdt * v), qo + s * NP + k * NQ2);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        } else {
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            const R mji = mjiv[k];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                // v - 0 is v, so lanes off a face keep their sum bit for bit
                R v = acc[s][k];
                v = v - xs[s * NQ2 + NQ * k];
                v = v - ys[s * NQ2 + NQ * k];
                if constexpr (GEO == GEO_GENERAL)
                    v = mji * v;
                store_out<ST>(v, dqo + s * NP + k * NQ2);
                if (P.write_q)
                    store_out<ST>(R(q[s][k] + bdt * v), qo + s * NP + k * NQ2);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        }
    }
    if constexpr (RX) {
        // the wave's sum: a butterfly of (sum, compensation) pairs, every
        // lane forming the same total
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double os = __shfl_xor(prod, off), oc = __shfl_xor(pc, off);
            comp_add(prod, pc, os, oc);
        }
        if ((threadIdx.x & 63) == 0)
            P.eprod[eslot] = prod + pc;
    }
}

template <int NQ, int LAW, typename R, int GEO, int FLUX, int ST, int RX = 0>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
dg_plane3_kernel(const StageParams P)
{
    using T = Plane3<NQ, LAW, R, GEO, FLUX>;
    extern __shared__ __attribute__((aligned(16))) unsigned char dg_lds[];
    const int64_t blk = xcd_remap(blockIdx.x, gridDim.x);
    const int wave = threadIdx.x >> 6;
    const int wpb = blockDim.x >> 6;
    // RX: one product slot per wave of the launch
    const int64_t eslot = blk * wpb + wave;
    if constexpr (T::G > 1) {
        // one workgroup of G waves packs CPG cells
        plane3_wave<NQ, LAW, R, GEO, FLUX, ST, T::G, RX>(P, blk * T::CPG,
                                                         reinterpret_cast<R *>(dg_lds), eslot);
    } else {
        R *wlds = reinterpret_cast<R *>(dg_lds) + wave * T::EPW * T::SLAB;
        plane3_wave<NQ, LAW, R, GEO, FLUX, ST, 1, RX>(P, (blk * wpb + wave) * T::EPW, wlds,
                                                      eslot);
    }
}

template <int NQ, int LAW, typename R, int GEO, int FLUX, int RX>
int launch_plane3_rx(const StageParams &P, void *stream, int wpb)
{
    using T = Plane3<NQ, LAW, R, GEO, FLUX>;
    if (P.K <= 0)
        return 0;
    if (wpb != 1 && wpb != 2 && wpb != 4)
        wpb = 1;
    int64_t blocks;
    size_t lds;
    if constexpr (T::G > 1) {
        // packed cells: the workgroup is the G waves that share them
        wpb = T::G;
        blocks = (P.K + T::CPG - 1) / T::CPG;
        lds = (size_t)T::CPG * T::SLAB * sizeof(R);
    } else {
        const int64_t waves = (P.K + T::EPW - 1) / T::EPW;
        blocks = (waves + wpb - 1) / wpb;
        lds = (size_t)wpb * T::EPW * T::SLAB * sizeof(R);
    }
    if (P.streaming)
        hipLaunchKernelGGL((dg_plane3_kernel<NQ, LAW, R, GEO, FLUX, 1, RX>), dim3((unsigned)blocks),
                           dim3(64 * wpb), lds, static_cast<hipStream_t>(stream), P);
    else
        hipLaunchKernelGGL((dg_plane3_kernel<NQ, LAW, R, GEO, FLUX, 0, RX>), dim3((unsigned)blocks),
                           dim3(64 * wpb), lds, static_cast<hipStream_t>(stream), P);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int NQ, int LAW, typename R, int GEO, int FLUX>
int launch_plane3(const StageParams &P, void *stream, int wpb)
{
    return launch_plane3_rx<NQ, LAW, R, GEO, FLUX, 0>(P, stream, wpb);
}

template <int NQ, int LAW, typename R, int GEO, int FLUX>
int launch_plane3_relax(const StageParams &P, void *stream, int wpb)
{
    return launch_plane3_rx<NQ, LAW, R, GEO, FLUX, 1>(P, stream, wpb);
}


}  // namespace rv
