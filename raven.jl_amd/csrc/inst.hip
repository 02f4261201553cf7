// Explicit instantiations of the fused stage kernels for one (dim, precision,
// law) triple; the Makefile compiles this file once per triple with
//   -DINST_DIM=<1|2|3> -DINST_REAL=<double|float> -DINST_LAW=<1|2> -DINST_FN=<name>
// and api.hip asks every unit for a launcher.  3-D uses the column kernel
// (dg_plane3.hpp), 2-D affine cells the node kernel (dg_quad2.hpp), 1-D and
// 2-D general cells the pencil kernel (dg_kernels.hpp).
#include "dg_plane3.hpp"
#include "dg_quad2.hpp"

#ifndef INST_DIM
#error "INST_DIM must be defined"
#endif

namespace rv {

namespace {

template <int NQ, int GEO, int FLUX>
StageLauncher pick_one(bool relax)
{
    constexpr int DIM = INST_DIM, LAW = INST_LAW;
    using R = INST_REAL;
    if constexpr (DIM == 3) {
        return relax ? &launch_plane3_relax<NQ, LAW, R, GEO, FLUX>
                     : &launch_plane3<NQ, LAW, R, GEO, FLUX>;
    } else {
        if (relax)
            return nullptr;  // the fused relaxation stage is built for 3-D
        if constexpr (DIM == 2 && GEO == GEO_AFFINE)
            return &launch_quad2<NQ, LAW, R, FLUX>;
        else
            return &launch_stage<DIM, NQ, LAW, R, GEO, FLUX>;
    }
}

template <int NQ>
StageLauncher pick(const KernelKey &k)
{
    if constexpr (IPow<NQ, INST_DIM - 1>::v > 64) {
        return nullptr;
    } else {
        if (k.flux != FLUX_RUSANOV && k.flux != FLUX_CENTRAL)
            return nullptr;
        const bool rx = k.relax != 0;
        if (k.geo == GEO_AFFINE)
            return k.flux == FLUX_RUSANOV ? pick_one<NQ, GEO_AFFINE, FLUX_RUSANOV>(rx)
                                          : pick_one<NQ, GEO_AFFINE, FLUX_CENTRAL>(rx);
        return k.flux == FLUX_RUSANOV ? pick_one<NQ, GEO_GENERAL, FLUX_RUSANOV>(rx)
                                      : pick_one<NQ, GEO_GENERAL, FLUX_CENTRAL>(rx);
    }
}

}  // namespace

StageLauncher INST_FN(const KernelKey &k)
{
    if (k.dim != INST_DIM || k.law != INST_LAW || k.real_bytes != (int)sizeof(INST_REAL) ||
        k.vflux != 0)
        return nullptr;
    switch (k.nq) {
    case 2: return pick<2>(k);
    case 3: return pick<3>(k);
    case 4: return pick<4>(k);
    case 5: return pick<5>(k);
    case 6: return pick<6>(k);
    case 7: return pick<7>(k);
    case 8: return pick<8>(k);
    default: return nullptr;
    }
}

}  // namespace rv
