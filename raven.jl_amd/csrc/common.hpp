// Internal definitions shared by the host API and the kernel instantiation units.
#pragma once

#include <cstdint>
#include <string>

namespace rv {

enum LawId : int { LAW_ADVECTION = 1, LAW_EULER = 2 };
enum GeoId : int { GEO_AFFINE = 1, GEO_GENERAL = 2 };
enum FluxId : int { FLUX_RUSANOV = 1, FLUX_CENTRAL = 2, FLUX_EC = 3 };

// Per-face neighbour record, one per (face, cell):
//   x >= 0 : neighbour cell; y = neighbour face | orientation << 3
//   x == -1: q+ = q- (boundary code != 0 with the identity boundary state,
//            or a self-mapped face: Raven's non-conforming / ghost faces)
//   x == -2: explicit per-node offsets; y = index of the face's first entry
//            in the explicit table (cell offsets e*Ns*Np + node, int64)
constexpr int FACE_SELF = -1;
constexpr int FACE_EXPLICIT = -2;

struct FaceRec {
    int32_t x;
    int32_t y;
};

// Device code reads the arrays below through global-memory (address space
// 1) pointers, also when a StageParams comes out of a device table (batches):
// a pointer loaded from memory is otherwise a generic one, compiled to flat
// loads that also count against the LDS wait counter.  The host pass sees
// plain pointers; the layout is the same.
#if defined(__HIP_DEVICE_COMPILE__)
#define RV_GLOBAL __attribute__((address_space(1)))
#else
#define RV_GLOBAL
#endif

// A host-side device pointer as the parameter structs hold it.
template <typename T>
inline T RV_GLOBAL *as_global(T *p)
{
    return (T RV_GLOBAL *)p;
}

// Arguments of one fused DG stage launch.  Pointers are device pointers.
struct StageParams {
    int64_t K;            // cells
    const void RV_GLOBAL *q_in;  // state read (R)
    void RV_GLOBAL *q_out;       // state written (R), nullptr in RHS mode
    void RV_GLOBAL *dq;          // register (R)
    double a;             // dq = a*dq + L
    double b_dt;          // q_out = q_in + b_dt*dq
    int read_dq;          // load dq (a != 0 or increment)
    int write_q;          // stage mode
    int dir_mask;         // active directions (bit d)
    int streaming;        // dq and q_out through L2 as streaming (non-temporal) accesses
    const FaceRec RV_GLOBAL *faces;        // (2dim, K)
    const int64_t RV_GLOBAL *explicit_idx; // explicit face-node offsets
    // affine geometry: per cell, per direction: dRdX[d,d] (FP64)
    const double RV_GLOBAL *affine;        // (K, dim)
    // general geometry (FP64): wJ*dRdX rows, MJI, face normals and wsJ
    const double RV_GLOBAL *vgeo;  // (K, dim, dim, Np): [e][d][c][node] = wJ*dRdX[d,c]
    const double RV_GLOBAL *mji;   // (K, Np)
    const double RV_GLOBAL *mj;    // (K, Np) = wJ
    const double RV_GLOBAL *sgeo;  // (K, 2dim, dim+1, Nfp): n_c ..., wsJ
    // FP32 handles, 3-D general cells: vgeo, mji and sgeo rounded to FP32 once
    // (the 3-D column kernel rounds every use of them to the state type, so
    // reading these gives the same results with half the geometry bytes)
    const float RV_GLOBAL *vgeo32;
    const float RV_GLOBAL *mji32;
    const float RV_GLOBAL *sgeo32;
    const float RV_GLOBAL *mj32;
    // operators (FP64, device)
    const double RV_GLOBAL *Dt;    // [i*Nq + n] = D[n,i]            (general)
    const double RV_GLOBAL *Dw;    // [i*Nq + n] = D[n,i] w_n / w_i  (affine)
    const double RV_GLOBAL *winv;  // [i] = 1/w_i
    double gamma;
    double u[3];
    // relaxation stage (the flux-differencing kernels built for it): dq = a dq + k
    // with k = L(q_in) accumulated alone, and each workgroup stores
    // sum MJ v(q_in) . k over its nodes at eprod[workgroup]
    double RV_GLOBAL *eprod;
    // mortar updates the 3-D column kernel adds after its stores (nullptr:
    // none): per cell its slot in mdelta or -1, per (cell, face) the face
    // kind, and the per-(slot, face) updates of mortar_face_kernel
    const int32_t RV_GLOBAL *mslot;
    const int32_t RV_GLOBAL *mkind;
    const void RV_GLOBAL *mdelta;
};

// Product slots (StageParams::eprod) the 3-D column kernel's relaxation stage
// writes, one per wave launched: Nq = 5, 6, 7 pack 64 G / Nq^2 cells into a
// workgroup of G = 2, 4, 4 waves, the others give a wave 64 / Nq^2 cells and a
// workgroup wpb waves.
inline int64_t plane3_relax_slots(int nq, int64_t K, int wpb)
{
    const int nq2 = nq * nq;
    const int g = nq == 5 ? 2 : (nq == 6 || nq == 7) ? 4 : 1;
    if (g > 1) {
        const int cpg = 64 * g / nq2;
        return (K + cpg - 1) / cpg * g;
    }
    if (wpb != 1 && wpb != 2 && wpb != 4)
        wpb = 1;
    const int epw = 64 / nq2;
    const int64_t waves = (K + epw - 1) / epw;
    return (waves + wpb - 1) / wpb * wpb;
}

// Launch the fused stage for one kernel configuration; returns false when no
// kernel is instantiated for it.
using StageLauncher = int (*)(const StageParams &, void *stream, int waves_per_block);

struct KernelKey {
    int dim, nq, law, real_bytes, geo, flux;
    int vflux = 0;  // 0: weak form; else the flux-differencing form's two-point flux
    int relax = 0;  // the fused relaxation stage (StageParams::eprod) of a flux-differencing kernel
};

// Mortar coupling of hanging faces (mortar.hip), launched after a stage.
// One non-conforming face of a mortar cell: its cell e (and slot li in the
// list of mortar cells), face f, kind (1 hanging, 2 covered) and partners
// (hanging: coarse cell, sub-face; covered: the 2^(dim-1) fine cells)
struct MortarPair {
    int64_t e, li;
    int64_t part[4];
    int32_t f, kind;
};

struct MortarParams {
    StageParams stage;       // the stage just launched (states, law, mji)
    int64_t ncells;          // cells with non-conforming faces
    const int64_t RV_GLOBAL *cells;
    // their non-conforming (cell, face) pairs: covered (coarse) sides first,
    // which take longer, then hanging (fine) sides
    int64_t npairs;
    const MortarPair RV_GLOBAL *pairs;
    const int32_t RV_GLOBAL *kind;     // (K, 2dim): 0 conforming, 1 hanging, 2 covered
    const int64_t RV_GLOBAL *partner;  // (K, 2dim, 4), see forest_mortars
    const double RV_GLOBAL *sgeo;      // surface geometry with the lift weights kept
    const double RV_GLOBAL *halves;    // tohalves_1d, (2, Nq, Nq) row-major
    void RV_GLOBAL *delta;             // (ncells, 2dim, Ns, Nfp) per-face updates (R)
};
// phase 0: the per-face updates (before the stage kernel); phase 1: adding
// them to dq and q_out (after a stage kernel that does not add them itself)
using MortarLauncher = int (*)(const MortarParams &, void *stream, int phase);
MortarLauncher find_mortar_launcher(int dim, int nq, int law, int real_bytes, int flux);

// Flux-differencing kernels (fluxdiff.hip): whole cells per workgroup, a
// thread per node, at least 256 threads where a cell has fewer nodes
constexpr int fd_cells_per_workgroup(int nodes) { return nodes >= 256 ? 1 : 256 / nodes; }

// Entropy passes of the relaxation LSRK step (entropy.hip).  Each pass runs
// RELAX_BLOCKS workgroups over contiguous node ranges and leaves per workgroup
// two Neumaier sums (value, compensation) in partial[4*block ..]: the entropy
// integral sum MJ eta(x) at +0 and the entropy product sum MJ v(x).k at +2.
enum RelaxMode : int {
    RELAX_ETA = 0,      // x = q; eta sum; q0 = q when q0 is given (the step's copy)
    RELAX_PRODUCT = 1,  // x = q (p of entropyproduct); product sum with k
    RELAX_STAGE = 2,    // x = q; product with k; dq = a dq + k; q += b dt dq
    RELAX_NEWTON = 3,   // (first: k = qf - q0); q = x = q0 + gamma k; both sums
    RELAX_AXPY = 4      // q = q0 + gamma k; no sums
};
constexpr int RELAX_BLOCKS = 2048;  // 8 workgroups of 4 waves per CU
constexpr int RELAX_THREADS = 256;
struct RelaxParams {
    int mode, law, dim, real_bytes;
    int64_t Np, K;
    double gas_gamma;     // the Euler law's ratio of specific heats
    const double *MJ;     // (K, Np)
    void *q, *k, *dq;     // (K, Ns, Np) states of the handle's precision
    void *q0;             // read by RELAX_NEWTON / RELAX_AXPY, written by RELAX_ETA
    double a, b_dt;       // RELAX_STAGE
    int read_dq;          // RELAX_STAGE: a != 0
    int first;            // RELAX_NEWTON: form k = qf - q0 first
    const void *qf;       // RELAX_NEWTON first: the state after the stages (may be q or k)
    double relax;         // RELAX_NEWTON / RELAX_AXPY: gamma
    double *partial;      // 4 * RELAX_BLOCKS doubles
};
int launch_relax(const RelaxParams &P, void *stream);

// Compensated sums of per-workgroup partials on the device, in a fixed order
// (deterministic): for slot a in [0, nslots) and b in [0, nwhich),
//   out[a * nwhich + b] = sum over i in [0, n) of
//       src[a * slot_stride + b * which_stride + (i / group) * group_stride + i % group].
int launch_slot_sums(const double *src, int64_t n, int nslots, int nwhich, int64_t slot_stride,
                     int64_t which_stride, int group, int64_t group_stride, double *out,
                     void *stream);

// Batched stages (batch.hip): one work-list item per wave's worth of cells of
// one member; `kind` selects the member's (Nq, law, precision) routine.
struct BatchItem {
    int32_t member;
    int32_t kind;
    int64_t cell0;
};
// kind id, and cells and LDS bytes per item of a batch member: [0] in the
// work-list kernel (a wave per item), [1] when its kind launches on its own
// (2-D Nq = 5, 6, 7: a packed workgroup per item); false when the batch
// kernels do not cover (dim, nq, law, precision)
bool batch_kind(int dim, int nq, int law, int real_bytes, int *kind, int cells_per_item[2],
                int slab_bytes[2]);
// waves of the batch kernel resident on the whole device at once
int batch_resident_waves(int dim, int wpb, int slab_bytes, int64_t *waves);
// one persistent launch over the work list (members: device table)
int launch_batch(int dim, const StageParams *members, const BatchItem *items, int64_t nitems,
                 int64_t nwaves, int slab_bytes, int wpb, void *stream);
// one launch for the items [first, first + count) of a single kind
int launch_batch_kind(int dim, int kind, const StageParams *members, const BatchItem *items,
                      int64_t first, int64_t count, int slab_bytes, int wpb, void *stream);

// Host-side grid generation (host_grid.cpp): Raven's brick GridManager +
// generate restated for uniform bricks, and the LobattoCell 1-D operators.
struct BrickSpec {
    int dim, nq, level;
    int n[3], periodic[3];
    double lo[3], hi[3];
};
int64_t brick_numcells(const BrickSpec &b);
// tree and quadrant coordinates (K x dim), p4est brick order
void brick_quadrants(const BrickSpec &b, int32_t *tree, int32_t *quad);
// vertex coordinates of the coarse brick along d (n[d]+1 values, correctly
// rounded lattice lo + (hi-lo)*i/n like Julia's range(lo, stop=hi, length=n+1))
void brick_vertices(const BrickSpec &b, int d, double *v);
void lobatto_operators(int nq, double *x, double *w, double *D);
// tohalves_1d (lobattocells.jl:15-16): interpolation from the GLL points to
// the GLL points of the lower / upper half, row-major [a][b] = l_b(y_a)
void lobatto_halves(int nq, double *lower, double *upper);
// full Raven-layout arrays (any pointer may be null)
void brick_arrays(const BrickSpec &b, double *points, double *metrics, double *surf,
                  int64_t *vmapM, int64_t *vmapP, int64_t *bc);
int64_t brick_dtoc(const BrickSpec &b, int64_t *dtoc);
// per-cell diagonal of dRdX (K x dim) and neighbour cell per face (K x 2dim,
// -1 on non-periodic boundaries) without per-node arrays
void brick_affine(const BrickSpec &b, double *drdx, double *J);
void brick_neighbours(const BrickSpec &b, int64_t *nbr);

// Adaptive brick forest (forest.cpp): leaves after adapt! with 2:1 balance.
struct Forest;
Forest *forest_new(const BrickSpec &b);
void forest_free(Forest *f);
Forest *forest_clone(const Forest *f);
// per (cell, face), e * 2dim + f: the conforming neighbour (-1 when none),
// the face kind (0 conforming or boundary, 1 hanging, 2 covered) and the
// boundary code; any pointer may be null
void forest_faces(const Forest *f, int64_t *nbr, int32_t *kind, int64_t *bc);
BrickSpec forest_spec(const Forest *f);
int64_t forest_size(const Forest *f);
// adapt! with Raven's flags (0x00 coarsen, 0x10 refine, anything else keeps),
// flags overwritten with the outcome; -1 when a refined leaf would pass the
// finest level
int forest_adapt(Forest *f, uint8_t *flags);
// tree lattice coordinates (K x dim), levels (K), coordinates inside the tree
// at each leaf's level (K x dim); any pointer may be null
void forest_cells(const Forest *f, int32_t *tree, int32_t *level, int32_t *quad);
// points (Np, dim, K), vmapM / vmapP (2dim*Nfp, K) 1-based, boundary codes
// (2dim, K) and face kinds (2dim, K: 0 conforming or boundary, 1 hanging on a
// coarser neighbour, 2 covered by finer neighbours); any pointer may be null
void forest_arrays(const Forest *f, int nq, double *points, int64_t *vmapM, int64_t *vmapP,
                   int64_t *bc, int32_t *kind);
// partners of the non-conforming faces, (K, 2dim, 4): a hanging face holds
// (coarse cell, sub-face index s1 + 2*s2, -1, -1); a covered face holds the
// 2^(dim-1) finer cells in sub-face order; conforming faces hold -1
void forest_mortars(const Forest *f, int64_t *partner);
// Raven's non-conforming face tables per face group g < dim: the number of
// distinct non-conforming interfaces and of non-conforming (cell, face) pairs
void forest_nc_counts(const Forest *f, int64_t *nunique, int64_t *nfaces);
// vmapNC[g] (Nfp, 1 + 2^(dim-1), nunique[g]) 1-based, nctoface[g] (2,
// nunique[g]), nctypes[g] and ncids[g] (nfaces[g], ncids 1-based); any
// group's pointers may be null
void forest_nctables(const Forest *f, int nq, int64_t *const *vmapNC, int8_t *const *nctoface,
                     int8_t *const *nctypes, int64_t *const *ncids);
// face kinds and partners (as forest_arrays / forest_mortars give them) from
// Raven's vmapNC / nctoface; 0, -1 (an index out of range) or -2 (a table that
// is not an aligned face pairing)
int mortars_from_nctables(int dim, int nq, int64_t K, const int64_t *nunique,
                          const int64_t *const *vmapNC, const int8_t *const *nctoface,
                          int32_t *kind, int64_t *partner);
// continuous numbering of the forest's nodes, hanging nodes included, as
// Raven's dtoc (1-based, first appearance); returns the number of continuous
// nodes
int64_t forest_dtoc(const Forest *f, int nq, int64_t *dtoc);

}  // namespace rv
