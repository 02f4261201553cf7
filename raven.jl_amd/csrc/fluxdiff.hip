// FluxDifferencingForm stage kernel (dgsem.jl:1-17, 163-255, 514-608):
// the volume term from a symmetric two-point flux instead of the weak form,
//   dq(ijk) += sum_dir sum_n [ -MJI D[id,n] l_g(ijk) . f#(q_ijk, q_n)
//                              + f#(q_ijk, q_n) . l_g(n) MJI D[n,id] ],
// l_g = MJ dRdX[dir, :], n running over the line through ijk along dir, with
// f# = CentralFlux (numericalfluxes.jl:39-42) or, for Euler,
// EntropyConservativeFlux (euler.jl:150-184, logavg numericalfluxes.jl:9-22),
// fused with the surface term (dgsem.jl:733-784) and the LSRK update
// dq = a dq + L(q), q_out = q + b dt dq (odesolvers.jl:55-57).  The surface
// flux is Rusanov or the volume form's own two-point flux f#(q-, q+)' n
// (surfaceflux, numericalfluxes.jl:27-37), the pairing of Raven's
// entropy-conservation tests.
//
// The form is compute-bound (Nq (Nq - 1) / 2 two-point fluxes per line and
// direction, two logarithms each for the entropy-conservative flux).  Whole
// cells fill a workgroup with their states staged in LDS; each symmetric pair
// is evaluated once by threads packed over the workgroup's pairs, and a thread
// per node gathers the pair terms.  Affine cells evaluate only the row dir of
// the two-point flux (their l_g has no other component).
//
// A relaxation stage (RX, the RLSRK step of odesolvers.jl:117-151) keeps
// k = L(q) apart from a dq, forms dq = a dq + k at the end and leaves each
// workgroup's sum of MJ v(q) . k (the stage's entropy production) in
// P.eprod[workgroup].
//
// Compiled once per (dim, precision) with -DFD_DIM and -DFD_REAL; api.hip asks
// every unit for a launcher.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "dg_kernels.hpp"
#include "entropy.hpp"

#ifndef FD_DIM
#error "FD_DIM must be defined"
#endif

namespace rv {

namespace {

template <int DIM, int NS, int LAW, int VF, typename R>
struct TwoPoint {
    // row `row` of the dim x Ns two-point flux matrix between q1 and q2
    __device__ static inline void rows(const Law<DIM, LAW, R> &law, const R *q1, const R *q2,
                                       int row0, int nrows, R (*F)[NS])
    {
        if constexpr (VF == FLUX_CENTRAL || LAW == LAW_ADVECTION) {
            for (int r = 0; r < nrows; ++r) {
                R f1[NS], f2[NS];
                law.flux_row(row0 + r, q1, f1);
                law.flux_row(row0 + r, q2, f2);
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    F[r][s] = (f1[s] + f2[s]) / R(2);
            }
        } else {
            ec_rows(law, q1, prim(law, q1), q2, prim(law, q2), row0, nrows, F);
        }
    }

    // Per-node quantities of the entropy-conservative flux: velocity u = rho u / rho,
    // b = rho / 2p with p = (gamma - 1)(rho e - |rho u|^2 / 2 rho), |u|^2; computed
    // with the operations of euler.jl:163-175, so staging them per node gives the
    // values a per-pair evaluation would
    struct Prim {
        R u[DIM], b, uu;
    };
    __device__ static inline Prim prim(const Law<DIM, LAW, R> &law, const R *q)
    {
        Prim P;
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            P.u[d] = q[1 + d] / q[0];
        R dot = q[1] * q[1];
#pragma unroll
        for (int d = 1; d < DIM; ++d)
            dot = dot + q[1 + d] * q[1 + d];
        const R p = law.gm1 * (q[DIM + 1] - dot / (R(2) * q[0]));
        P.b = q[0] / (R(2) * p);
        P.uu = P.u[0] * P.u[0];
#pragma unroll
        for (int d = 1; d < DIM; ++d)
            P.uu = P.uu + P.u[d] * P.u[d];
        return P;
    }

    // EntropyConservativeFlux (euler.jl:150-184), rows row0 .. row0 + nrows - 1
    __device__ static inline void ec_rows(const Law<DIM, LAW, R> &law, const R *q1,
                                          const Prim &P1, const R *q2, const Prim &P2, int row0,
                                          int nrows, R (*F)[NS])
    {
        R uavg[DIM], fr[DIM];
        const R ravg = (q1[0] + q2[0]) / R(2);
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            uavg[d] = (P1.u[d] + P2.u[d]) / R(2);
        const R bavg = (P1.b + P2.b) / R(2);
        const R u2avg = (P1.uu + P2.uu) / R(2);
        // the logarithmic means as (a + b) / 2F: rho's taken whole, b's only
        // through 1 / (2 gm1 blog) = F_b / (2 bavg gm1), which shares
        // 1 / (2 bavg) with the pressure average
        const R rlog = ravg / logavg_factor(q1[0], q2[0]);
        const R Fb = logavg_factor(P1.b, P2.b);
        const R inv2b = R(1) / (R(2) * bavg);
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            fr[d] = uavg[d] * rlog;
        const R pav = ravg * inv2b;
        const R c = Fb * inv2b * (R(1) / law.gm1) - u2avg / R(2);
        for (int r = 0; r < nrows; ++r) {
            const int d = row0 + r;
            F[r][0] = fr[d];
#pragma unroll
            for (int b = 0; b < DIM; ++b)
                F[r][1 + b] = uavg[d] * fr[b] + (d == b ? pav : R(0));
            R v = F[r][1] * uavg[0];
#pragma unroll
            for (int b = 1; b < DIM; ++b)
                v = v + F[r][1 + b] * uavg[b];
            F[r][DIM + 1] = c * fr[d] + v;
        }
    }

    // logavg (numericalfluxes.jl:9-22): the logarithmic mean is (a + b) / (2F),
    // F = log(zeta) / (2f) = atanh(f) / f with zeta = a / b and
    // f = (zeta - 1) / (zeta + 1) = (a - b) / (a + b).  Raven takes the series
    // of F in u = f^2 below the type's epsilon and the logarithm above; here
    // the series runs up to u < 1e-2, to the term that leaves its remainder
    // below the type's rounding (u^9 / 19 in FP64, u^5 / 11 in FP32), and
    // f comes from a - b: as accurate as Raven's form (within 2.2 eps of the
    // exact mean on 20,000 random pairs, against 2.4 eps), and neighbouring
    // nodes of a resolved state never reach the logarithm, two divisions
    // and a log fewer per mean; returns F
    __device__ static inline R logavg_factor(R a, R b)
    {
        const R f = (a - b) / (a + b);
        const R u = f * f;
        R F;
        if (u < R(1e-2)) {
            if constexpr (sizeof(R) == 8) {
                F = R(1) / R(17);
#pragma unroll
                for (int k = 7; k >= 0; --k)
                    F = F * u + R(1) / R(2 * k + 1);
            } else {
                F = R(1) / R(9);
#pragma unroll
                for (int k = 3; k >= 0; --k)
                    F = F * u + R(1) / R(2 * k + 1);
            }
        } else {
            F = log(a / b) / (R(2) * f);
        }
        return F;
    }
};

template <int DIM, int NQ, int LAW, typename R, int GEO, int VF, int SF>
struct FD {
    static constexpr int NS = LawInfo<DIM, LAW>::NS;
    static constexpr int NP = IPow<NQ, DIM>::v;
    static constexpr int NFP = IPow<NQ, DIM - 1>::v;
    static constexpr int NF = 2 * DIM;
    static constexpr int CPB = fd_cells_per_workgroup(NP);
    // volume pairs: the Nq (Nq - 1) / 2 symmetric pairs of each line, NL lines
    // per cell and direction; HL lines per pass through the LDS pair buffer of
    // at most 36 KB (two 512-thread FP64 workgroups per CU, as the register
    // bound allows; a smaller FP32 buffer for more workgroups per CU measured
    // slower: 2.03 against 1.80 ms per stage at Nq = 8)
    static constexpr int NL = NP / NQ;
    static constexpr int NPAIR = NQ * (NQ - 1) / 2;
    static constexpr int BL = CPB * NL;
    static constexpr int HL_FIT = 36864 / (NPAIR * NS * (int)sizeof(R));
    static constexpr int HL = HL_FIT < 1 ? 1 : (HL_FIT < BL ? HL_FIT : BL);
    // the pair buffer is stored line-fastest in rows of HLP values, HLP odd, so
    // that lanes reading one pair of consecutive lines, or different pairs of
    // one line, fall in different LDS banks
    static constexpr int HLP = HL % 2 == 0 ? HL + 1 : HL;
    static constexpr int THREADS = CPB * NP;
    using S = Stage<DIM, NQ, LAW, R, GEO, FLUX_RUSANOV>;
};

template <int DIM, int NQ, int LAW, typename R, int GEO, int VF, int SF, bool RX>
__device__ __forceinline__ void fd_stage(const StageParams &P)
{
    using T = FD<DIM, NQ, LAW, R, GEO, VF, SF>;
    using S = typename T::S;
    constexpr int NS = T::NS, NP = T::NP, NFP = T::NFP, NF = T::NF;
    __shared__ R qs[T::CPB][NS][NP];
    // entropy-conservative flux: per-node primitives of the line partners
    constexpr bool EC = VF == FLUX_EC && LAW == LAW_EULER;
    using TP = TwoPoint<DIM, NS, LAW, VF, R>;
    using PrimT = typename TP::Prim;
    __shared__ PrimT ps[EC ? T::CPB : 1][EC ? NP : 1];
    const int t = threadIdx.x;
    const int c = t / NP, node = t - c * NP;
    const int64_t e0 = (int64_t)blockIdx.x * T::CPB + c;
    const bool active = e0 < P.K;
    // every thread takes part in the pair passes (and a relaxation stage's
    // workgroup sum): a thread past the last cell repeats the last cell's work
    // and stores nothing
    const int64_t e = active ? e0 : P.K - 1;
    const Law<DIM, LAW, R> law(P);
    const R *qin = static_cast<const R *>(P.q_in);
    R acc[NS];
    {
        R q[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            q[s] = qin[(e * NS + s) * NP + node];
            qs[c][s][node] = q[s];
        }
        if constexpr (EC)
            ps[c][node] = TP::prim(law, q);
    }
    __syncthreads();
    const int ijk[3] = {node % NQ, DIM > 1 ? (node / NQ) % NQ : 0,
                        DIM > 2 ? node / (NQ * NQ) : 0};
    const double MJ1 = P.mj[e * NP + node];
    const double MJI = GEO == GEO_GENERAL ? P.mji[e * NP + node] : 1.0 / MJ1;
    const R a = R(P.a);
    R *dq = static_cast<R *>(P.dq);
#pragma unroll
    for (int s = 0; s < NS; ++s)
        acc[s] = P.read_dq && !RX ? a * dq[(e * NS + s) * NP + node] : R(0);

    // ---- volume: one direction after another (esvolumeterm_dir!).  Node i's
    // term from line partner n is MJI_i (D[n,i] G_n - D[i,n] G_i) with
    // G_x = l_g(x) . f#(q_i, q_n): with f# symmetric, the pair's
    // H = D[n,i] G_n - D[i,n] G_i serves both ends (+H at i, -H at n), so each
    // pair is evaluated once (as the reference's :per_dir_symmetric kernel),
    // by threads packed densely over the workgroup's pairs, and every node
    // then adds its partners' +-MJI H in partner order from LDS.  The
    // diagonal term is zero.
    // the pair buffer; a relaxation stage's workgroup sum reuses it at the end
    // (and the surface pass's face terms after it)
    constexpr int HS_BYTES = T::NPAIR * NS * T::HLP * (int)sizeof(R);
    constexpr int FS_BYTES = T::CPB * NF * NFP * NS * (int)sizeof(R);
    constexpr int RED_BYTES = RX ? 2 * T::THREADS * (int)sizeof(double) : 0;
    constexpr int POOL0 = HS_BYTES > FS_BYTES ? HS_BYTES : FS_BYTES;
    __shared__ __align__(16) char pool[POOL0 > RED_BYTES ? POOL0 : RED_BYTES];
    R *hs = reinterpret_cast<R *>(pool);
#pragma unroll
    for (int dir = 0; dir < DIM; ++dir) {
        if (!((P.dir_mask >> dir) & 1))
            continue;
        const int id = ijk[dir];
        int stride = 1;
#pragma unroll
        for (int d = 0; d < dir; ++d)
            stride *= NQ;
        const int myline = c * T::NL + node % stride + (node / (stride * NQ)) * stride;
        const R mji = R(MJI);
        // pair order over the lanes: along x, the pairs of one line side by
        // side (a line's nodes are contiguous); along y and z, one pair of
        // consecutive lines side by side (lines are contiguous there), so
        // that lanes read different LDS banks
        const bool line_major = dir == 0;
        for (int l0 = 0; l0 < T::BL; l0 += T::HL) {
            const int nl = T::BL - l0 < T::HL ? T::BL - l0 : T::HL;
            for (int item = t; item < T::HL * T::NPAIR; item += T::THREADS) {
                int cl, pr;
                if (line_major) {
                    cl = item / T::NPAIR;
                    pr = item - cl * T::NPAIR;
                } else {
                    pr = item / T::HL;
                    cl = item - pr * T::HL;
                }
                if (cl >= nl)
                    continue;
                int rem = pr, i = 0;
                while (rem >= NQ - 1 - i) {
                    rem -= NQ - 1 - i;
                    ++i;
                }
                const int n = i + 1 + rem;
                const int gl = l0 + cl, pc = gl / T::NL, l = gl - pc * T::NL;
                const int base = l % stride + (l / stride) * stride * NQ;
                const int na = base + i * stride, nb = base + n * stride;
                int64_t pe = (int64_t)blockIdx.x * T::CPB + pc;
                pe = pe < P.K ? pe : P.K - 1;
                R qa[NS], qb[NS];
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    qa[s] = qs[pc][s][na];
                    qb[s] = qs[pc][s][nb];
                }
                const R Din = R(P.Dt[n * NQ + i]);  // D[i, n]
                const R Dni = R(P.Dt[i * NQ + n]);  // D[n, i]
                R h[NS];
                if constexpr (GEO == GEO_GENERAL) {
                    R F[DIM][NS], la[DIM], lb[DIM];
                    if constexpr (EC)
                        TP::ec_rows(law, qa, ps[pc][na], qb, ps[pc][nb], 0, DIM, F);
                    else
                        TP::rows(law, qa, qb, 0, DIM, F);
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        la[d] = R(P.vgeo[((pe * DIM + dir) * DIM + d) * NP + na]);
                        lb[d] = R(P.vgeo[((pe * DIM + dir) * DIM + d) * NP + nb]);
                    }
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        R ga = la[0] * F[0][s], gb = F[0][s] * lb[0];
#pragma unroll
                        for (int d = 1; d < DIM; ++d) {
                            ga = ga + la[d] * F[d][s];
                            gb = gb + F[d][s] * lb[d];
                        }
                        h[s] = Dni * gb - Din * ga;
                    }
                } else {
                    R F[1][NS];
                    if constexpr (EC)
                        TP::ec_rows(law, qa, ps[pc][na], qb, ps[pc][nb], dir, 1, F);
                    else
                        TP::rows(law, qa, qb, dir, 1, F);
                    const double aff = P.affine[pe * DIM + dir];
                    const R la = R(P.mj[pe * NP + na] * aff), lb = R(P.mj[pe * NP + nb] * aff);
#pragma unroll
                    for (int s = 0; s < NS; ++s)
                        h[s] = Dni * (F[0][s] * lb) - Din * (la * F[0][s]);
                }
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    hs[(pr * NS + s) * T::HLP + cl] = h[s];
            }
            __syncthreads();
            const int ml = myline - l0;
            if (ml >= 0 && ml < nl) {
                for (int n = 0; n < NQ; ++n) {
                    if (n == id)
                        continue;
                    const int lo = id < n ? id : n, hi = id < n ? n : id;
                    const int pr = lo * (2 * NQ - lo - 1) / 2 + (hi - lo - 1);
                    const R w = id < n ? mji : -mji;
#pragma unroll
                    for (int s = 0; s < NS; ++s)
                        acc[s] = acc[s] + w * hs[(pr * NS + s) * T::HLP + ml];
                }
            }
            __syncthreads();
        }
    }

    // ---- surface (surfaceterm!): each face node's flux is evaluated once, by
    // threads packed densely over the workgroup's 2 dim Nq^(dim-1) face nodes
    // per cell (a thread per volume node would run the flux at the width of
    // the wave for the eight nodes of a cell edge); lift * nf goes to LDS,
    // m fastest, and every node subtracts its faces' terms in face order
    // d = 1..dim, f = 1, 2
    R *fs = reinterpret_cast<R *>(pool);
    for (int item = t; item < T::CPB * NF * NFP; item += T::THREADS) {
        const int fc = item / (NF * NFP), r = item - fc * (NF * NFP);
        const int f = r / NFP, m = r - f * NFP;
        const int d = f >> 1;
        if (!((P.dir_mask >> d) & 1))
            continue;
        int64_t fe = (int64_t)blockIdx.x * T::CPB + fc;
        fe = fe < P.K ? fe : P.K - 1;
        const int side = (f & 1) ? NQ - 1 : 0;
        int nm;  // the face node's volume node
        if constexpr (DIM == 1)
            nm = side;
        else if constexpr (DIM == 2)
            nm = d == 0 ? side + NQ * m : m + NQ * side;
        else
            nm = d == 0   ? side + NQ * m
                 : d == 1 ? m % NQ + NQ * side + NQ * NQ * (m / NQ)
                          : m + NQ * NQ * side;
        R qm[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s)
            qm[s] = qs[fc][s][nm];
        const FaceRec rec = P.faces[fe * NF + f];
        R qp[NS], nf[NS];
        if (rec.x >= 0) {
            const int fn = rec.y & 7, o = (rec.y >> 3) & 7;
            const int nn = face_node<DIM, NQ>(fn, orient_node<DIM, NQ>(o, m));
#pragma unroll
            for (int s = 0; s < NS; ++s)
                qp[s] = qin[((int64_t)rec.x * NS + s) * NP + nn];
        } else if (rec.x == FACE_EXPLICIT) {
            const int64_t off = P.explicit_idx[(int64_t)rec.y + m];
#pragma unroll
            for (int s = 0; s < NS; ++s)
                qp[s] = qin[off + s * NP];
        } else {
#pragma unroll
            for (int s = 0; s < NS; ++s)
                qp[s] = qm[s];
        }
        R lift;
        if constexpr (SF != FLUX_RUSANOV) {
            // twopointflux(q-, q+)' n: on an affine face n = +-e_d, so the
            // product is row d of the flux with the face's sign
            using TS = TwoPoint<DIM, NS, LAW, SF, R>;
            if constexpr (GEO == GEO_AFFINE) {
                R F[1][NS];
                if constexpr (SF == FLUX_EC && LAW == LAW_EULER)
                    TS::ec_rows(law, qm, ps[fc][nm], qp, TS::prim(law, qp), d, 1, F);
                else
                    TS::rows(law, qm, qp, d, 1, F);
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    nf[s] = (f & 1) ? F[0][s] : -F[0][s];
                lift = R(P.affine[fe * DIM + d] * P.winv[side]);
            } else {
                const double *sg = P.sgeo + ((fe * NF + f) * (DIM + 1)) * NFP + m;
                R F[DIM][NS];
                if constexpr (SF == FLUX_EC && LAW == LAW_EULER)
                    TS::ec_rows(law, qm, ps[fc][nm], qp, TS::prim(law, qp), 0, DIM, F);
                else
                    TS::rows(law, qm, qp, 0, DIM, F);
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    R v = F[0][s] * R(sg[0]);
#pragma unroll
                    for (int cc = 1; cc < DIM; ++cc)
                        v = v + F[cc][s] * R(sg[cc * NFP]);
                    nf[s] = v;
                }
                lift = R(sg[DIM * NFP] * P.mji[fe * NP + nm]);
            }
        } else if constexpr (GEO == GEO_AFFINE) {
            R fm[NS];
            law.flux_row(d, qm, fm);
            S::nflux_axis(law, d, (f & 1) ? R(1) : R(-1), fm, qm, qp, nf);
            lift = R(P.affine[fe * DIM + d] * P.winv[side]);
        } else {
            const double *sg = P.sgeo + ((fe * NF + f) * (DIM + 1)) * NFP + m;
            R nv[DIM];
#pragma unroll
            for (int cc = 0; cc < DIM; ++cc)
                nv[cc] = R(sg[cc * NFP]);
            S::nflux_general(law, nv, qm, qp, nf);
            lift = R(sg[DIM * NFP] * P.mji[fe * NP + nm]);
        }
#pragma unroll
        for (int s = 0; s < NS; ++s)
            fs[((fc * NF + f) * NS + s) * NFP + m] = lift * nf[s];
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const int d = f >> 1;
        if (!((P.dir_mask >> d) & 1))
            continue;
        if (ijk[d] != ((f & 1) ? NQ - 1 : 0))
            continue;
        int m = 0;
        if constexpr (DIM == 2)
            m = d == 0 ? ijk[1] : ijk[0];
        if constexpr (DIM == 3)
            m = d == 0 ? ijk[1] + NQ * ijk[2] : d == 1 ? ijk[0] + NQ * ijk[2] : ijk[0] + NQ * ijk[1];
#pragma unroll
        for (int s = 0; s < NS; ++s)
            acc[s] = acc[s] - fs[((c * NF + f) * NS + s) * NFP + m];
    }
    // the node's own state back from LDS (not held in registers through the
    // pair and face passes)
    R q1[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s)
        q1[s] = qs[c][s][node];

    // ---- update
    if constexpr (RX) {
        // dq .*= rka; dq .+= k; the dEta term v(q) . k (odesolvers.jl:130-134)
        double x[NS], kk[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            x[s] = (double)q1[s];
            kk[s] = (double)acc[s];
            if (P.read_dq)
                acc[s] = a * dq[(e * NS + s) * NP + node] + acc[s];
        }
        double prod = 0.0;
        if constexpr (DIM == 3)
            prod = active ? MJ1 * law_entropy_dot_3d<LAW>(x, kk, P.gamma) : 0.0;
        else
            prod = active ? MJ1 * law_entropy_dot<LAW, DIM>(x, kk, P.gamma) : 0.0;
        double pc = 0.0;  // compensation of prod
        // the workgroup's sum, compensated: a fixed tree over the threads of
        // (sum, compensation) pairs, each addition's rounding error (TwoSum,
        // exact and symmetric in its operands) carried in the compensation;
        // in the pair buffer, once every thread has read its face terms
        double *red = reinterpret_cast<double *>(pool);
        __syncthreads();
        if constexpr (T::THREADS % 64 == 0) {
            // whole waves: a butterfly inside each wave (both lanes of a pair
            // form the same sum), then the wave sums in wave order
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double os = __shfl_xor(prod, off), oc = __shfl_xor(pc, off);
                comp_add(prod, pc, os, oc);
            }
            if ((t & 63) == 0) {
                red[2 * (t >> 6)] = prod;
                red[2 * (t >> 6) + 1] = pc;
            }
            __syncthreads();
            if (t == 0) {
                double sum = red[0], c = red[1];
                for (int w = 1; w < T::THREADS / 64; ++w)
                    comp_add(sum, c, red[2 * w], red[2 * w + 1]);
                P.eprod[blockIdx.x] = sum + c;
            }
        } else {
            red[2 * t] = prod;
            red[2 * t + 1] = pc;
            __syncthreads();
            for (int w = 256; w > 0; w >>= 1) {
                if (t < w && t + w < T::THREADS) {
                    double sum = red[2 * t], c = red[2 * t + 1];
                    comp_add(sum, c, red[2 * (t + w)], red[2 * (t + w) + 1]);
                    red[2 * t] = sum;
                    red[2 * t + 1] = c;
                }
                __syncthreads();
            }
            if (t == 0)
                P.eprod[blockIdx.x] = red[0] + red[1];
        }
    }
    if (!active)
        return;
    const R bdt = R(P.b_dt);
    R *qo = static_cast<R *>(P.q_out);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        dq[(e * NS + s) * NP + node] = acc[s];
        if (P.write_q)
            qo[(e * NS + s) * NP + node] = q1[s] + bdt * acc[s];
    }
}

// A 512-thread workgroup (two waves per SIMD) is held to 128 VGPRs so that two
// workgroups share a CU: measured on MI355X at Nq = 8 (3-D Euler, entropy
// conservative, FP64) 3.6 ms per stage against 5.0 ms at the compiler's own
// 140 VGPRs (one workgroup per CU), a few spilled registers included
template <int DIM, int NQ, int LAW, typename R, int GEO, int VF, int SF>
__global__ void __launch_bounds__(512)
    __attribute__((amdgpu_waves_per_eu(FD<DIM, NQ, LAW, R, GEO, VF, SF>::THREADS == 512 ? 4 : 1)))
        dg_fd_kernel(const StageParams P)
{
    fd_stage<DIM, NQ, LAW, R, GEO, VF, SF, false>(P);
}

// the relaxation stage, under the same register bound
template <int DIM, int NQ, int LAW, typename R, int GEO, int VF, int SF>
__global__ void __launch_bounds__(512)
    __attribute__((amdgpu_waves_per_eu(FD<DIM, NQ, LAW, R, GEO, VF, SF>::THREADS == 512 ? 4 : 1)))
        dg_fd_relax_kernel(const StageParams P)
{
    fd_stage<DIM, NQ, LAW, R, GEO, VF, SF, true>(P);
}

template <int DIM, int NQ, int LAW, typename R, int GEO, int VF, int SF, bool RX>
int launch_fd(const StageParams &P, void *stream, int)
{
    using T = FD<DIM, NQ, LAW, R, GEO, VF, SF>;
    if (P.K <= 0)
        return 0;
    const int64_t blocks = (P.K + T::CPB - 1) / T::CPB;
    if constexpr (RX)
        hipLaunchKernelGGL((dg_fd_relax_kernel<DIM, NQ, LAW, R, GEO, VF, SF>),
                           dim3((unsigned)blocks), dim3(T::THREADS), 0,
                           static_cast<hipStream_t>(stream), P);
    else
        hipLaunchKernelGGL((dg_fd_kernel<DIM, NQ, LAW, R, GEO, VF, SF>), dim3((unsigned)blocks),
                           dim3(T::THREADS), 0, static_cast<hipStream_t>(stream), P);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int NQ, int LAW, int VF, int SF, bool RX>
StageLauncher pick_geo(const KernelKey &k)
{
    using R = FD_REAL;
    return k.geo == GEO_GENERAL ? &launch_fd<FD_DIM, NQ, LAW, R, GEO_GENERAL, VF, SF, RX>
                                : &launch_fd<FD_DIM, NQ, LAW, R, GEO_AFFINE, VF, SF, RX>;
}

// surface flux: Rusanov, or the volume form's own two-point flux; the fused
// relaxation stage is built for the latter (the entropy-conservative pairing)
template <int NQ, int LAW, int VF>
StageLauncher pick_surface(const KernelKey &k)
{
    if (k.flux == FLUX_RUSANOV)
        return k.relax ? nullptr : pick_geo<NQ, LAW, VF, FLUX_RUSANOV, false>(k);
    if (k.flux == VF)
        return k.relax ? pick_geo<NQ, LAW, VF, VF, true>(k) : pick_geo<NQ, LAW, VF, VF, false>(k);
    return nullptr;
}

template <int NQ>
StageLauncher pick_fd(const KernelKey &k)
{
    if constexpr (IPow<NQ, FD_DIM>::v > 512) {
        return nullptr;
    } else {
        if (k.law == LAW_ADVECTION)
            return k.vflux == FLUX_CENTRAL ? pick_surface<NQ, LAW_ADVECTION, FLUX_CENTRAL>(k)
                                           : nullptr;
        if (k.vflux == FLUX_CENTRAL)
            return pick_surface<NQ, LAW_EULER, FLUX_CENTRAL>(k);
        if (k.vflux == FLUX_EC)
            return pick_surface<NQ, LAW_EULER, FLUX_EC>(k);
        return nullptr;
    }
}

}  // namespace

StageLauncher FD_FN(const KernelKey &k)
{
    if (k.dim != FD_DIM || k.real_bytes != (int)sizeof(FD_REAL))
        return nullptr;
    switch (k.nq) {
    case 2: return pick_fd<2>(k);
    case 3: return pick_fd<3>(k);
    case 4: return pick_fd<4>(k);
    case 5: return pick_fd<5>(k);
    case 6: return pick_fd<6>(k);
    case 7: return pick_fd<7>(k);
    case 8: return pick_fd<8>(k);
    default: return nullptr;
    }
}

}  // namespace rv
