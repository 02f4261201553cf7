// Batched fused LSRK stages: many independent operators (different Nq, law,
// precision and cell count) advanced by one persistent launch per stage (2-D)
// or one launch per kind present, each with that kind's own LDS slab and the
// single operator's launch shape (3-D).
//
// The work list holds one item per wave's worth of cells of one member
// (member, kind, first cell); items are grouped by kind so that the waves
// running at the same time mostly execute the same code.  Each persistent wave
// walks the list with a stride of the launched wave count and hands every item
// to the same per-wave routine the single-operator kernels use
// (stage_wave / plane3_wave), with that member's StageParams read from a
// device table.  Members are restricted to affine geometry and the Rusanov
// flux (the brick solves batched and served here); 2-D members take Nq 2..8,
// 3-D members Nq 4..8.  2-D items run the node kernel's group routine
// (quad2_group) with one wave per group: the same arithmetic as the single
// operator's packed groups, so results are bit-identical.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "dg_plane3.hpp"
#include "dg_quad2.hpp"

namespace rv {

namespace {

constexpr int KINDS_PER_NQ = 4;  // (advection, Euler) x (double, float)

template <int DIM>
struct BatchRange {
    static constexpr int NQ_MIN = DIM == 3 ? 4 : 2;
    static constexpr int NQ_MAX = 8;
};

// Waves of one item of a kind launching on its own: 2-D Nq = 5, 6, 7 pack
// their cells into a workgroup of G waves, as the single-operator kernel does;
// the work-list kernel (and every other kind) runs one wave per item.
template <int DIM, int NQ, int LAW, typename R>
constexpr int kind_group_waves()
{
    return DIM == 2 ? quad2_group_waves<NQ>() : 1;
}

// LDS of one wave for a kind: the single-operator kernel's slab, in elements
template <int DIM, int NQ, int LAW, typename R>
constexpr int kind_slab_elems()
{
    if constexpr (DIM == 3) {
        using T = Plane3<NQ, LAW, R, GEO_AFFINE, FLUX_RUSANOV>;
        return T::EPW * T::SLAB;
    } else {
        return Quad2<NQ, LAW, R, FLUX_RUSANOV, 1, quad2_nodes_per_lane<LAW, R>()>::SLAB;
    }
}

// ... and in bytes, rounded up to 16 (the mixed kernel's slab stride)
template <int DIM, int NQ, int LAW, typename R>
constexpr int kind_slab_bytes()
{
    return (int)((kind_slab_elems<DIM, NQ, LAW, R>() * sizeof(R) + 15) / 16 * 16);
}

// A packed 2-D kind's group: cells and LDS bytes of one workgroup
template <int NQ, int LAW, typename R>
using PackedQuad2 = Quad2<NQ, LAW, R, FLUX_RUSANOV, quad2_group_waves<NQ>(),
                          quad2_nodes_per_lane<LAW, R>()>;

template <int DIM, int NQ, int LAW, typename R>
__device__ __attribute__((always_inline)) inline void run_item(const StageParams &P, int64_t cell0, unsigned char *wl)
{
    if constexpr (DIM == 3)
        plane3_wave<NQ, LAW, R, GEO_AFFINE, FLUX_RUSANOV>(P, cell0, reinterpret_cast<R *>(wl));
    else {
        // the lane index is made opaque here, so that what each kind derives
        // from it stays inside that kind's code: hoisted out of the work-list
        // loop for every kind at once, it would stay live across the loop
        int lane = (int)(threadIdx.x & 63);
        asm volatile("" : "+v"(lane));
        quad2_group<NQ, LAW, R, FLUX_RUSANOV, 0, 1, quad2_nodes_per_lane<LAW, R>()>(
            P, cell0, reinterpret_cast<R *>(wl), lane);
    }
}

template <int DIM, int NQ>
__device__ __attribute__((always_inline)) inline void run_nq(int sub, const StageParams &P, int64_t cell0, unsigned char *wl)
{
    switch (sub) {
    case 0: run_item<DIM, NQ, LAW_ADVECTION, double>(P, cell0, wl); break;
    case 1: run_item<DIM, NQ, LAW_ADVECTION, float>(P, cell0, wl); break;
    case 2: run_item<DIM, NQ, LAW_EULER, double>(P, cell0, wl); break;
    default: run_item<DIM, NQ, LAW_EULER, float>(P, cell0, wl); break;
    }
}

template <int DIM, int NQ>
__device__ __attribute__((always_inline)) inline void run_kind(int kind, const StageParams &P, int64_t cell0, unsigned char *wl)
{
    if constexpr (NQ > BatchRange<DIM>::NQ_MAX) {
        (void)kind, (void)P, (void)cell0, (void)wl;
    } else {
        if (kind / KINDS_PER_NQ == NQ - BatchRange<DIM>::NQ_MIN)
            run_nq<DIM, NQ>(kind % KINDS_PER_NQ, P, cell0, wl);
        else
            run_kind<DIM, NQ + 1>(kind, P, cell0, wl);
    }
}

// Every wave exits after its last item: the loop bound is the list length.
template <int DIM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DIM == 2 ? 4 : 2)))
dg_batch_kernel(const StageParams *__restrict__ members, const BatchItem *__restrict__ items,
                int64_t nitems, int64_t nwaves, int slab_bytes)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char dg_lds[];
    const int wave = threadIdx.x >> 6;
    const int wpb = blockDim.x >> 6;
    // slab_bytes is a multiple of 16: 16-byte LDS accesses stay available
    unsigned char *wl = static_cast<unsigned char *>(
        __builtin_assume_aligned(dg_lds + (size_t)wave * slab_bytes, 16));
    // consecutive waves take consecutive items, so the waves running at
    // the same time share kinds and neighbouring cells, and every XCD sees
    // the whole mix of kinds (cheap and costly items alike)
    const int64_t first = wave_uniform(xcd_remap(blockIdx.x, gridDim.x) * wpb + wave);
    if constexpr (DIM == 2)
        quad2_clear_spare(reinterpret_cast<double *>(wl), (int)(threadIdx.x & 63));
    for (int64_t it = first; it < nitems; it += nwaves) {
        const BatchItem w = items[it];
        run_kind<DIM, BatchRange<DIM>::NQ_MIN>(w.kind, members[w.member], w.cell0, wl);
        // the next item reuses the slab
        wave_lds_sync();
    }
}

// One kind only: a wave per item of the range [first, first + count), the
// same code and register allocation as the single-operator kernel.  3-D
// batches launch one of these per kind present.
template <int DIM, int NQ, int LAW, typename R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
dg_batch_kind_kernel(const StageParams *__restrict__ members, const BatchItem *__restrict__ items,
                     int64_t first, int64_t count, int slab_bytes)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char dg_lds[];
    if constexpr (kind_group_waves<DIM, NQ, LAW, R>() > 1) {
        // one packed item per workgroup (the whole workgroup returns or runs it)
        constexpr int G = kind_group_waves<DIM, NQ, LAW, R>();
        const int64_t k = xcd_remap(blockIdx.x, gridDim.x);
        (void)slab_bytes;
        if (k >= count)
            return;
        const BatchItem w = items[first + k];
        R *gl = reinterpret_cast<R *>(dg_lds);
        quad2_clear_spare(gl, (int)threadIdx.x);
        quad2_group<NQ, LAW, R, FLUX_RUSANOV, 0, G, quad2_nodes_per_lane<LAW, R>()>(
            members[w.member], w.cell0, gl, (int)threadIdx.x);
        return;
    }
    const int wave = threadIdx.x >> 6;
    const int wpb = blockDim.x >> 6;
    const int64_t k = wave_uniform(xcd_remap(blockIdx.x, gridDim.x) * wpb + wave);
    if (k >= count)
        return;
    const BatchItem w = items[first + k];
    // the kind's own slab at the single-operator kernel's offset: the same
    // known alignment gives the same LDS access widths (16-byte reads where
    // that kernel has them; for FP32 those measured slower)
    (void)slab_bytes;
    R *wl = reinterpret_cast<R *>(dg_lds) + wave * kind_slab_elems<DIM, NQ, LAW, R>();
    if constexpr (DIM == 2)
        quad2_clear_spare(wl, (int)(threadIdx.x & 63));
    run_item<DIM, NQ, LAW, R>(members[w.member], w.cell0, reinterpret_cast<unsigned char *>(wl));
}

// slab_bytes: a wave's slab, or a packed kind's whole workgroup's
template <int DIM, int NQ, int LAW, typename R>
void launch_kind_one(const StageParams *members, const BatchItem *items, int64_t first,
                     int64_t count, int slab_bytes, int wpb, hipStream_t st)
{
    constexpr int G = kind_group_waves<DIM, NQ, LAW, R>();
    if (G > 1)
        wpb = G;
    const int64_t blocks = G > 1 ? count : (count + wpb - 1) / wpb;
    const size_t lds = G > 1 ? (size_t)slab_bytes : (size_t)wpb * slab_bytes;
    hipLaunchKernelGGL((dg_batch_kind_kernel<DIM, NQ, LAW, R>), dim3((unsigned)blocks),
                       dim3(64 * wpb), lds, st, members, items, first, count, slab_bytes);
}

template <int DIM, int NQ>
int launch_kind_nq(int sub, const StageParams *members, const BatchItem *items, int64_t first,
                   int64_t count, int slab_bytes, int wpb, hipStream_t st)
{
    switch (sub) {
    case 0:
        launch_kind_one<DIM, NQ, LAW_ADVECTION, double>(members, items, first, count, slab_bytes,
                                                        wpb, st);
        break;
    case 1:
        launch_kind_one<DIM, NQ, LAW_ADVECTION, float>(members, items, first, count, slab_bytes,
                                                       wpb, st);
        break;
    case 2:
        launch_kind_one<DIM, NQ, LAW_EULER, double>(members, items, first, count, slab_bytes, wpb,
                                                    st);
        break;
    default:
        launch_kind_one<DIM, NQ, LAW_EULER, float>(members, items, first, count, slab_bytes, wpb,
                                                   st);
        break;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int DIM, int NQ>
int launch_kind(int kind, const StageParams *members, const BatchItem *items, int64_t first,
                int64_t count, int slab_bytes, int wpb, hipStream_t st)
{
    if constexpr (NQ > BatchRange<DIM>::NQ_MAX) {
        return -1;
    } else {
        if (kind / KINDS_PER_NQ == NQ - BatchRange<DIM>::NQ_MIN)
            return launch_kind_nq<DIM, NQ>(kind % KINDS_PER_NQ, members, items, first, count,
                                           slab_bytes, wpb, st);
        return launch_kind<DIM, NQ + 1>(kind, members, items, first, count, slab_bytes, wpb, st);
    }
}

// cells per item and LDS bytes per item: ep[0], sb[0] in the work-list
// kernel, ep[1], sb[1] in a launch of the kind alone
template <int DIM, int NQ, int LAW, typename R>
void layout_of(int *ep, int *sb)
{
    if constexpr (DIM == 3)
        ep[0] = Plane3<NQ, LAW, R, GEO_AFFINE, FLUX_RUSANOV>::EPW;
    else
        ep[0] = Quad2<NQ, LAW, R, FLUX_RUSANOV, 1, quad2_nodes_per_lane<LAW, R>()>::CELLS;
    sb[0] = kind_slab_bytes<DIM, NQ, LAW, R>();
    ep[1] = ep[0];
    sb[1] = sb[0];
    if constexpr (DIM == 2 && kind_group_waves<DIM, NQ, LAW, R>() > 1) {
        using Q = PackedQuad2<NQ, LAW, R>;
        ep[1] = Q::CELLS;
        sb[1] = (int)((Q::SLAB * sizeof(R) + 15) / 16 * 16);
    }
}

template <int DIM, int NQ>
bool layout_nq(int nq, int sub, int *ep, int *sb)
{
    if constexpr (NQ > BatchRange<DIM>::NQ_MAX) {
        return false;
    } else {
        if (nq != NQ)
            return layout_nq<DIM, NQ + 1>(nq, sub, ep, sb);
        switch (sub) {
        case 0: layout_of<DIM, NQ, LAW_ADVECTION, double>(ep, sb); break;
        case 1: layout_of<DIM, NQ, LAW_ADVECTION, float>(ep, sb); break;
        case 2: layout_of<DIM, NQ, LAW_EULER, double>(ep, sb); break;
        default: layout_of<DIM, NQ, LAW_EULER, float>(ep, sb); break;
        }
        return true;
    }
}

template <int DIM>
const void *kernel_of()
{
    return reinterpret_cast<const void *>(&dg_batch_kernel<DIM>);
}

}  // namespace

bool batch_kind(int dim, int nq, int law, int real_bytes, int *kind, int *cells_per_item,
                int *slab_bytes)
{
    const int sub = (law == LAW_EULER ? 2 : 0) + (real_bytes == 4 ? 1 : 0);
    if (law != LAW_EULER && law != LAW_ADVECTION)
        return false;
    if (real_bytes != 4 && real_bytes != 8)
        return false;
    bool ok = false;
    if (dim == 2) {
        ok = layout_nq<2, BatchRange<2>::NQ_MIN>(nq, sub, cells_per_item, slab_bytes);
        *kind = (nq - BatchRange<2>::NQ_MIN) * KINDS_PER_NQ + sub;
    } else if (dim == 3) {
        ok = layout_nq<3, BatchRange<3>::NQ_MIN>(nq, sub, cells_per_item, slab_bytes);
        *kind = (nq - BatchRange<3>::NQ_MIN) * KINDS_PER_NQ + sub;
    }
    return ok;
}

int batch_resident_waves(int dim, int wpb, int slab_bytes, int64_t *waves)
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return -1;
    const void *k = dim == 3 ? kernel_of<3>() : kernel_of<2>();
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 64 * wpb,
                                                     (size_t)wpb * slab_bytes) != hipSuccess)
        return -1;
    *waves = (int64_t)cus * (per_cu > 0 ? per_cu : 1) * wpb;
    return 0;
}

int launch_batch(int dim, const StageParams *members, const BatchItem *items, int64_t nitems,
                 int64_t nwaves, int slab_bytes, int wpb, void *stream)
{
    if (nitems <= 0)
        return 0;
    const int64_t blocks = (nwaves + wpb - 1) / wpb;
    const int64_t launched = blocks * wpb;
    const size_t lds = (size_t)wpb * slab_bytes;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (dim == 3)
        hipLaunchKernelGGL(dg_batch_kernel<3>, dim3((unsigned)blocks), dim3(64 * wpb), lds, st,
                           members, items, nitems, launched, slab_bytes);
    else
        hipLaunchKernelGGL(dg_batch_kernel<2>, dim3((unsigned)blocks), dim3(64 * wpb), lds, st,
                           members, items, nitems, launched, slab_bytes);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_batch_kind(int dim, int kind, const StageParams *members, const BatchItem *items,
                      int64_t first, int64_t count, int slab_bytes, int wpb, void *stream)
{
    if (count <= 0)
        return 0;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (dim == 3)
        return launch_kind<3, BatchRange<3>::NQ_MIN>(kind, members, items, first, count,
                                                      slab_bytes, wpb, st);
    return launch_kind<2, BatchRange<2>::NQ_MIN>(kind, members, items, first, count, slab_bytes,
                                                  wpb, st);
}

}  // namespace rv
