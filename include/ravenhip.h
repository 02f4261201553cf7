/*
 * ravenhip.h -- C ABI of the MI355X (gfx950) DGSEM right-hand side and
 * low-storage Runge-Kutta library.
 *
 * Drop-in boundary for Raven.jl's `dg(dq, q, t; increment)` call
 * (src/balancelaws/dgsem.jl:84-126) and the LSRK stage loop around it
 * (src/balancelaws/odesolvers.jl:9-61).  Julia reaches it through `ccall`;
 * every argument is a plain pointer or scalar.
 *
 * Conventions
 *  - Arrays use Raven's memory layout (Julia column-major):
 *      state q, dq           (Nq^dim, Ns, K)            parent(GridArray), field axis dim+1
 *      volume metrics        (Nq^dim, dim*dim+2, K)      (dRdX column-major, J, wJ)
 *      surface metrics       (dim+2, 2dim*Nq^(dim-1), K) field first: (n, sJ, wsJ)
 *      vmapM, vmapP          (2dim*Nq^(dim-1), K) int64, 1-based
 *      boundary codes        (2dim, K) int64
 *      D                     (Nq, Nq) column-major, D[a + Nq*b] = l_b'(x_a)
 *  - Host arrays passed to rv_create are read during the call only.
 *  - State pointers passed to compute calls are DEVICE pointers (HIP device
 *    memory of the current device) holding FP64 or FP32 values as the
 *    descriptor's dtype says.  Results are complete when the handle's stream
 *    has been synchronised (rv_synchronize) or ordered after by the caller.
 *  - Every function returns RV_OK (0) or a negative RV_E* code; the message of
 *    the last failure is in rv_last_error(h) (or rv_last_error(NULL) when no
 *    handle exists).  The library never aborts the process.
 *  - A handle is not reentrant; distinct handles may be used from different
 *    host threads.
 */
#ifndef RAVENHIP_H
#define RAVENHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RAVENHIP_ABI_VERSION 2

enum {
    RV_OK = 0,
    RV_E_ARG = -1,       /* invalid argument or descriptor */
    RV_E_HIP = -2,       /* HIP runtime error */
    RV_E_NOMEM = -3,     /* device allocation failed */
    RV_E_UNSUPPORTED = -4, /* no kernel built for this (dim, Nq, law, dtype) */
    RV_E_MAPS = -5,      /* face maps could not be encoded */
    RV_E_STATE = -6      /* call not valid in the handle's state */
};

/* law ids (src/balancelaws/balancelaws/{advection,euler}.jl) */
enum { RV_LAW_ADVECTION = 1, RV_LAW_EULER = 2 };
/* numerical fluxes (src/balancelaws/numericalfluxes.jl:39-50, euler.jl:150-184):
 * surface fluxes RUSANOV and CENTRAL with either volume form, and with the
 * flux differencing form also its own volume flux as twopointflux(q-, q+)' n
 * (numericalfluxes.jl:27-37; EC on faces is Raven's Euler entropy-conservation
 * pairing); volume (two-point) fluxes of the flux differencing form CENTRAL
 * and EC (EntropyConservativeFlux, Euler only) */
enum { RV_FLUX_RUSANOV = 1, RV_FLUX_CENTRAL = 2, RV_FLUX_EC = 3 };
/* volume form (dgsem.jl:1-17): WeakForm, or FluxDifferencingForm(volume flux)
 * evaluated direction by direction as esvolumeterm_dir! (dgsem.jl:514-608) */
enum { RV_FORM_WEAK = 0, RV_FORM_FLUX_DIFFERENCING = 1 };
/* state precision: FP64, or FP32 state/flux arithmetic with FP64 geometry
 * (the 3-D weak-form stage reads FP32 copies of the general-cell geometry,
 * rounded once, which give the values it would round each use to) */
enum { RV_F64 = 1, RV_F32 = 2 };
/* geometry handling: AUTO detects affine elements (per-element scale factors)
 * and otherwise keeps per-node FP64 metrics; GENERAL forces the latter. */
enum { RV_GEOM_AUTO = 0, RV_GEOM_AFFINE = 1, RV_GEOM_GENERAL = 2 };
/* LSRK schemes (odesolvers.jl:63-71) */
enum { RV_LSRK54 = 54, RV_LSRK144 = 144 };

typedef struct rv_handle rv_handle;

/* Descriptor of one DGSEM operator: the fields of Raven's `DGSEM` struct
 * (dgsem.jl:19-30) plus the grid arrays it reads. */
typedef struct rv_desc {
    int32_t abi_version;   /* RAVENHIP_ABI_VERSION */
    int32_t dim;           /* 1, 2 or 3 */
    int32_t nq;            /* Nq (equal in every direction, dgsem.jl:92) */
    int32_t law;           /* RV_LAW_* */
    int32_t flux;          /* RV_FLUX_* */
    int32_t dtype;         /* RV_F64 / RV_F32 */
    int32_t geometry;      /* RV_GEOM_* */
    int32_t directions;    /* bit mask of active directions (Raven's `directions`), 0 = all */
    int64_t ncells;        /* K, local cells (no ghosts on one device) */
    double gamma;          /* Euler gamma (EulerLaw: FT(7//5)) */
    double velocity[3];    /* advection u (AdvectionLaw) */
    const double *D;       /* derivatives_1d(cell)[1] */
    const double *weights; /* weights_1d(cell)[1] */
    const double *metrics; /* first(volumemetrics(grid)) parent */
    const double *surface_metrics; /* first(surfacemetrics(grid)) parent */
    const int64_t *vmapM;  /* facemaps(grid).vmapM */
    const int64_t *vmapP;  /* facemaps(grid).vmapP */
    const int64_t *boundary_codes; /* boundarycodes(grid), may be NULL (all 0) */
    const int64_t *dtoc;   /* discontinuoustocontinuous(grid), may be NULL */
    int64_t ncontinuous;   /* numcnodes(grid) when dtoc is given */
    void *stream;          /* hipStream_t every call runs on; NULL: the default stream */
    /* from ABI version 2 (version 1 descriptors end above: weak form) */
    int32_t volume_form;   /* RV_FORM_* */
    int32_t volume_flux;   /* RV_FLUX_CENTRAL or RV_FLUX_EC with RV_FORM_FLUX_DIFFERENCING */
} rv_desc;

/* Brick descriptor: the grid a GridManager(LobattoCell(Nq,...), brick(coords,
 * periodic); min_level) + generate() pair builds, generated inside the library
 * (coarsegrids.jl:233-405, gridmanager.jl:1084-1208) with uniform spacing. */
typedef struct rv_brick_desc {
    int32_t abi_version;
    int32_t dim, nq, law, flux, dtype;
    int32_t n[3];          /* cells per direction of the coarse brick */
    int32_t periodic[3];
    int32_t level;         /* uniform refinement level (min_level) */
    double lo[3], hi[3];   /* domain extents */
    double gamma;
    double velocity[3];
    void *stream;
    /* from ABI version 2, as in rv_desc */
    int32_t volume_form, volume_flux;
} rv_brick_desc;

typedef struct rv_info {
    int32_t dim, nq, ns, dtype;
    int32_t geometry;           /* resolved RV_GEOM_AFFINE or RV_GEOM_GENERAL */
    int32_t faces_self;         /* faces with q+ = q- (boundary or self-mapped) */
    int32_t faces_explicit;     /* faces needing per-node neighbour indices */
    int32_t streaming;          /* 1: stages stream dq and q_out past L2 (3 states > 256 MB) */
    int64_t ncells;
    int64_t state_bytes;        /* bytes of one state array */
    int64_t device_bytes;       /* device memory held by the handle */
    int64_t stage_bytes;        /* algorithmic HBM bytes of one fused LSRK stage (a != 0) */
    int64_t stage_bytes_first;  /* ... of a stage with a = 0 */
    int64_t rhs_bytes;          /* ... of rv_rhs with increment = 0 */
} rv_info;

/* --- lifetime --------------------------------------------------------- */

/* DGSEM(; law, grid, surface_numericalflux) (dgsem.jl:46-81): uploads the
 * operator, metrics and face maps, encodes per-face neighbour records. */
int rv_create(const rv_desc *desc, rv_handle **out);

/* GridManager + generate for a uniform brick (see rv_brick_desc), then as
 * rv_create.  Grid arrays can be read back with rv_brick_arrays. */
int rv_create_brick(const rv_brick_desc *desc, rv_handle **out);

void rv_destroy(rv_handle *h);
const char *rv_last_error(const rv_handle *h);
int rv_get_info(const rv_handle *h, rv_info *info);
int rv_synchronize(rv_handle *h);
/* the hipStream_t every call of this handle is ordered on */
void *rv_stream(const rv_handle *h);

/* Resident device states for callers without a device allocator of their own
 * (a C driver; Julia without AMDGPU.jl): one state array of the handle's shape
 * (Nq^dim, Ns, K) and precision, Raven's GridArray parent layout (a11).
 * Upload and download copy one whole state between host memory and such an
 * array on the handle's stream and return when the copy is complete; free
 * waits for the handle's stream first.  The reference's GridArray needs no
 * counterpart (gridarrays.jl:29-108): these are the allocation side of the
 * boundary's device-pointer convention. */
int rv_state_alloc(rv_handle *h, void **state);
int rv_state_free(rv_handle *h, void *state);
int rv_state_upload(rv_handle *h, void *state, const void *host);
int rv_state_download(rv_handle *h, void *host, const void *state);

/* Launch shape of the stage kernels: waves per workgroup (1, 2 or 4; each wave
 * keeps its own cells and LDS slab, so results do not depend on it).  The
 * library picks a default per configuration, for fused stages and for
 * right-hand sides alone; setting it sets both, and get returns the stage
 * value.  3-D cells of Nq 5 to 7, which the kernel packs across the waves of
 * a workgroup, keep their own workgroup.  An autotuner that times the
 * choices on the caller's grid may override it.  No counterpart in Raven
 * (KernelAbstractions picks its workgroup size at launch). */
int rv_set_waves_per_block(rv_handle *h, int waves_per_block);
int rv_get_waves_per_block(const rv_handle *h);

/* Couple the hanging faces through mortars instead of Raven's self-mapping
 * (lobattocells.jl:2072-2074): face kinds (2dim, K) as rv_forest_arrays gives
 * them and partners as from rv_forest_mortars.  On each fine sub-face the coarse trace
 * is interpolated with tohalves_1d (lobattocells.jl:15-16), the numerical flux
 * is evaluated with the fine face's normal and weight, and the coarse cell
 * receives the transpose projection, so the scheme stays conservative.  Every
 * rv_rhs / rv_lsrk_stage / rv_lsrk_run then runs one more kernel before the
 * stage kernel (the non-conforming faces' lifts, into a device buffer of
 * (cells with such faces) x 2dim x Ns x Nq^(dim-1) values this call allocates)
 * and, except with the 3-D weak form, whose stage kernel takes those lifts
 * itself, one after it that adds them.  Needs general geometry.  No
 * counterpart in Raven's RHS, which builds these structures but does not read
 * them. */
int rv_enable_mortars(rv_handle *h, const int32_t *face_kind, const int64_t *partner);

/* tohalves_1d (lobattocells.jl:15-16): row-major [a][b] = l_b((x_a -+ 1)/2),
 * computed in quad precision and rounded. */
int rv_lobatto_halves(int nq, double *lower, double *upper);

/* --- the hot path ------------------------------------------------------ */

/* dg(dq, q, t; increment) (dgsem.jl:84-126):
 * dq = (increment ? dq : 0) + volume term + surface term. */
int rv_rhs(rv_handle *h, void *dq, const void *q, double t, int increment);

/* One fused LSRK stage (odesolvers.jl:53-58):
 *   dq = a*dq + L(q_in, t);  q_out = q_in + b_dt*dq
 * dq is not read when a == 0.  q_out must not alias q_in (neighbouring cells
 * read q_in's face values while q_out is written). */
int rv_lsrk_stage(rv_handle *h, void *q_out, const void *q_in, void *dq, double a,
                  double b_dt, double t);

/* solve!(q, timeend, LSRK(dg, q, dt; t0); adjust_final) (odesolvers.jl:9-61):
 * dq starts at zero (odesolvers.jl:43-44): its entry values are not read by
 * schemes whose first stage has a = 0 (LSRK54, LSRK144; that stage writes
 * dq = L(q)) and are zeroed first otherwise; steps are taken until
 * time + dt >= timeend, and with adjust_final the last one is shortened to
 * timeend - time (odesolvers.jl:21-25), else it keeps dt and ends past
 * timeend.  FP32 handles keep t0, dt and the clock in FP32 as LSRK{Float32}
 * does, comparing with timeend as given.  q is updated in place (the library
 * ping-pongs through a state-sized buffer it allocates on the first call).
 * dt must be positive.  t_out / nsteps_out may be NULL. */
int rv_lsrk_solve(rv_handle *h, void *q, void *dq, int scheme, double t0, double dt,
                  double timeend, int adjust_final, double *t_out, int64_t *nsteps_out);
/* rv_lsrk_solve with adjust_final = 1, solve!'s default. */
int rv_lsrk_run(rv_handle *h, void *q, void *dq, int scheme, double t0, double dt,
                double timeend, double *t_out, int64_t *nsteps_out);

/* rv_lsrk_run without any device allocation.  No counterpart in Raven: for
 * callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid hipMalloc).  The ping-pong state is a library-owned device
 * array of RV_NOALLOC_SCRATCH_BYTES (32 MiB), one per GPU, held until the
 * process ends; states larger than that are rejected with RV_E_ARG.  The
 * caller passes q and dq as for rv_lsrk_run.  Every call on one GPU shares
 * that GPU's scratch: calls that overlap on one GPU (other streams or host
 * threads) race and must be ordered by the caller. */
#define RV_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
int rv_lsrk_run_noalloc(rv_handle *h, void *q, void *dq, int scheme, double t0, double dt,
                        double timeend, double *t_out, int64_t *nsteps_out);

/* The stage coefficients rka, rkb, rkc of a scheme (coefficients_lsrk54 /
 * coefficients_lsrk144, odesolvers.jl:163-244) in double precision, for
 * callers that step stage by stage through rv_lsrk_stage.  Any array pointer
 * may be NULL; *nstages receives 5 or 14.  No handle, no device work. */
int rv_lsrk_coefficients(int scheme, double *a, double *b, double *c, int *nstages);

/* One relaxation LSRK step, dostep!(q, ::RLSRK) (odesolvers.jl:117-151), the
 * solver of RLSRK54 / RLSRK144 (odesolvers.jl:154-162):
 *   q0 = q;  eta0 = entropyintegral(q0);  dEta = -0
 *   per stage s: k = L(q);  dq = rka_s dq + k;  dEta += bfull_s dt v(q).k;
 *                q += rkb_s dt dq
 *   k = q - q0;  gamma: eta(q0 + gamma k) - eta0 - gamma dEta = 0;  q = q0 + gamma k
 * bfull are the Butcher weights of the scheme (odesolvers.jl:96-101).  The
 * root is found by Newton's method from *gamma (the previous step's root, 1 at
 * the start), stopping at |residual| <= 10 eps of the state type (NLsolve's
 * ftol there), when the residual no longer halves (its rounding floor; the
 * better of the last two roots is kept), or when the update falls below
 * 2 eps |gamma|, after at most 50 iterations; *gamma receives the root and *iterations (may be NULL) the
 * Newton iterations taken.  dq carries between steps like rv_lsrk_stage's
 * (not read at a stage with rka = 0); k and q0 are work arrays.  All four
 * arrays have the state's shape and must be distinct.  Blocks on the handle's
 * stream (each Newton iteration reads back its two sums).  A Newton update
 * that is not finite (a vanishing derivative) fails with RV_E_STATE and
 * leaves q at the last iterate q0 + gamma k, q0 holding the step's start. */
int rv_rlsrk_step(rv_handle *h, void *q, void *dq, void *k, void *q0, int scheme, double time,
                  double dt, double *gamma, int32_t *iterations);

/* --- adaptive forests ---------------------------------------------------- */

/* The leaves of GridManager(cell, brick(...); min_level) after adapt!(gm,
 * flags) calls (gridmanager.jl:30-91, 160-172), without p4est: the brick's
 * trees in tree order, Morton order inside a tree.  rv_forest_adapt takes
 * Raven's flags (gridmanager.jl:24-27), one per leaf: every complete family of
 * siblings all flagged AdaptCoarsen (0x00) becomes its parent, then every
 * other leaf flagged AdaptRefine (0x10) is split once; any other value
 * (AdaptNone 0x01, AdaptTouched 0x11, ...) keeps the leaf.  The forest is then
 * balanced 2:1 across faces, edges and corners, and `flags` is overwritten as
 * adapt! overwrites it (fill_adapt_flags, :117-129): AdaptRefine, AdaptCoarsen
 * or AdaptNone by what happened to each old leaf.  Host only. */
typedef struct rv_forest rv_forest;
int rv_forest_create(const rv_brick_desc *desc, rv_forest **out);
void rv_forest_destroy(rv_forest *f);
int64_t rv_forest_numcells(const rv_forest *f);
int rv_forest_adapt(rv_forest *f, uint8_t *flags);
/* tree lattice coordinates (K x dim), levels (K) and coordinates inside the
 * tree at each leaf's level (K x dim); any pointer may be NULL */
int rv_forest_cells(const rv_forest *f, int32_t *tree, int32_t *level, int32_t *quad);
/* Grid arrays generate() builds for the forest (gridmanager.jl:1084-1208):
 * brick points (Np, dim, K) from the leaves' levels and coordinates
 * (lobattocells.jl:1590-1674), vmapM / vmapP (1-based), boundary codes, and
 * per (face, cell) kinds: 0 conforming or boundary, 1 hanging on a coarser
 * neighbour, 2 covered by finer neighbours.  Non-conforming faces are
 * self-mapped in vmapP as in Raven (lobattocells.jl:2072-2074).  Curved or
 * refined cells take their metrics from rv_curved_metrics.  Any output
 * pointer may be NULL. */
int rv_forest_arrays(const rv_forest *f, int32_t nq, double *points, int64_t *vmapM,
                     int64_t *vmapP, int64_t *bc, int32_t *face_kind);
/* Partners of the non-conforming faces, (4, 2dim, K) int64 (column-major as
 * the other face arrays: entry k of face f of cell e at [k + 4 (f + 2dim e)]),
 * 0-based cells: a hanging face holds (coarse cell, sub-face s1 + 2 s2, -1, -1)
 * where s1, s2 are the halves along the face's first and second tangential
 * direction; a covered face holds its 2^(dim-1) finer cells in sub-face
 * order; conforming faces hold -1.  These are the facts Raven keeps in
 * vmapNC / nctoface / nctypes (lobattocells.jl:2208-2275). */
int rv_forest_mortars(const rv_forest *f, int64_t *partner);

/* Raven's non-conforming face tables (materializefacemaps,
 * lobattocells.jl:2121-2276) for the forest, one set per face group g < dim
 * (faces normal to direction g), as Raven's tuples hold them:
 *   vmapNC[g]   (Nfp, 1 + 2^(dim-1), nunique[g]) int64, 1-based volume
 *               indices: the parent (coarse) face's nodes, then each finer
 *               face's, in face node order; the finer faces in sub-face order
 *               s1 + 2 s2 (halves along the first and second tangential
 *               direction)
 *   nctoface[g] (2, nunique[g]) int8: parent face, child face (1-based)
 *   nctypes[g]  (nfaces[g]) int8, per non-conforming (cell, face) pair in
 *               cell-then-face order: 1 for the parent, 2 + sub-face for a child
 *   ncids[g]    (nfaces[g]) int64: the pair's interface (1-based column of
 *               vmapNC), interfaces numbered in order of first encounter.
 * rv_forest_nc_counts gives nunique and nfaces (dim entries each; either may
 * be NULL); in rv_forest_nctables each argument is an array of dim pointers,
 * and any of them, or any entry, may be NULL.  Brick trees are aligned, so
 * every face orientation is the identity.  Host only. */
int rv_forest_nc_counts(const rv_forest *f, int64_t *nunique, int64_t *nfaces);
int rv_forest_nctables(const rv_forest *f, int32_t nq, int64_t *const *vmapNC,
                       int8_t *const *nctoface, int8_t *const *nctypes, int64_t *const *ncids);
/* The face kinds (2dim, K) and partners (4, 2dim, K) rv_enable_mortars takes,
 * from Raven's own vmapNC and nctoface (the tables above, as facemaps(grid)
 * holds them), so a Raven grid can couple its hanging faces: a parent face
 * becomes kind 2 with its children as partners, each child face kind 1 with
 * (parent, sub-face).  Tables whose node lists are not the aligned face
 * pairing (rotated trees) are refused with RV_E_UNSUPPORTED.  Host only. */
int rv_mortars_from_nctables(int32_t dim, int32_t nq, int64_t ncells, const int64_t *nunique,
                             const int64_t *const *vmapNC, const int8_t *const *nctoface,
                             int32_t *face_kind, int64_t *partner);
/* discontinuoustocontinuous(grid) of the forest's grid (Np, K), 1-based, and
 * numcnodes(grid): nodes on hanging faces and hanging edges take the ids of
 * the coarse face's or edge's nodes at the same position (p4est lnodes,
 * gridmanager.jl:349-379), the others are shared where they coincide, and ids
 * are numbered by first appearance (lobattocells.jl:1718-1936).  Host only. */
int rv_forest_dtoc(const rv_forest *f, int32_t nq, int64_t *dtoc, int64_t *ncontinuous);

/* DGSEM(; law, grid = generate(warp, gm)) (dgsem.jl:46-81,
 * gridmanager.jl:1084-1208) for an adapted forest, with the grid generated on
 * the device: the node coordinates (the brick formula at each leaf's level,
 * lobattocells.jl:1590-1674) unless `points` (a DEVICE array (Np, dim, K),
 * e.g. the caller's warp of them) is given, the curved volume and surface
 * metrics (as rv_curved_metrics) and the stage kernels' geometry, all on the
 * descriptor's stream; the host builds one record per face from the leaves
 * (hanging and covered faces self-mapped as Raven's vmapP).  No per-node array
 * crosses PCIe.  `params` gives law, flux, dtype, Nq, gamma, velocity, stream
 * and volume form (its dim must be the forest's; the brick fields are the
 * forest's own).  The handle keeps a copy of the leaves for the continuous
 * numbering, built on the first rv_gather / rv_scatter; rv_get_maps gives the
 * face maps.  Nq <= 8. */
int rv_create_forest(const rv_forest *f, const rv_brick_desc *params, const double *points,
                     rv_handle **out);

/* --- batches of independent solves -------------------------------------- */

/* Many operators advanced together over a work list of (member, first cell)
 * items grouped by (Nq, law, precision): kinds with enough items to fill the
 * device (every 3-D kind) launch once per stage with the single operator's
 * code, and the small 2-D kinds share ONE persistent launch, so solves of
 * different degree, size and precision share each stage's launches.  Members
 * are handles from rv_create / rv_create_brick with the same dim and stream,
 * affine geometry, the Rusanov flux and all directions; 2-D members may have
 * Nq 2..8, 3-D members Nq 4..8.  The handles must outlive the batch.  Each
 * member's result is the one its own rv_lsrk_stage gives.  No counterpart in
 * Raven, which advances one solver per `solve!` call (odesolvers.jl:9-30). */
typedef struct rv_batch rv_batch;
int rv_batch_create(rv_handle *const *members, int32_t nmembers, rv_batch **out);
void rv_batch_destroy(rv_batch *b);
const char *rv_batch_last_error(const rv_batch *b);
/* work-list length, persistent waves of the mixed launch, its LDS bytes per wave */
int rv_batch_info(const rv_batch *b, int64_t *nitems, int64_t *nwaves, int32_t *slab_bytes);
/* Launch layout (the autotuner's knobs): waves per block of the mixed launch
 * (1, 2 or 4; default 2) and the item count from which a 2-D kind launches on
 * its own (default 1024).  Kinds launching on their own take their first
 * member's waves per block (rv_set_waves_per_block) as it is at this call.
 * Waits for the batch's stream; cached step graphs are rebuilt on next use.
 * Results do not depend on the layout. */
int rv_batch_configure(rv_batch *b, int mixed_waves_per_block, int64_t own_launch_items);
/* Concurrent streams (1 to 4, default 2) the launches of one call are spread
 * over.  The launches of one kind, and those of the work-list launch, form a
 * chain that touches no other member; chains go, heaviest first, to the
 * batch's stream and up to three private streams forked from it and joined
 * back with events, so a call stays ordered on the batch's stream and
 * captured step graphs get parallel branches.  Waits for the batch's stream;
 * cached step graphs are rebuilt on next use.  Results do not depend on it.
 * No counterpart in the reference, which advances one solver per call
 * (odesolvers.jl:9-30). */
int rv_batch_set_streams(rv_batch *b, int32_t nstreams);
/* One fused stage for every member m (arrays of nmembers entries):
 *   dq[m] = a[m]*dq[m] + L_m(q_in[m]);  q_out[m] = q_in[m] + b_dt[m]*dq[m]. */
int rv_batch_stage(rv_batch *b, void *const *q_out, const void *const *q_in, void *const *dq,
                   const double *a, const double *b_dt);
/* One whole LSRK step of every member m with its own step size dt[m] (its
 * final, shortened step included), ping-ponging between q0[m] and q1[m]: the
 * state is read from q1[m] when cur[m] is 1, else from q0[m], and after an odd
 * number of stages it ends in the other buffer.  dq[m] carries over between
 * steps (zero it before the first).  Members with dt[m] == 0 are left
 * untouched and launch nothing, so finished solves need not leave the batch.
 * The stage tables of the step go to the device in one copy, then each stage
 * launches once (2-D) or once per kind of the members stepping (3-D). */
int rv_batch_step(rv_batch *b, void *const *q0, void *const *q1, void *const *dq,
                  const int32_t *cur, const double *dt, int scheme);
/* nsteps calls of rv_batch_step with the same dt[m], the parities advancing
 * by the scheme's stage count after each step (cur itself is not written):
 * a serving loop issues the steps before the next final, shortened step of
 * any member in one call. */
int rv_batch_steps(rv_batch *b, void *const *q0, void *const *q1, void *const *dq,
                   const int32_t *cur, const double *dt, int scheme, int64_t nsteps);
/* nsteps whole LSRK steps of every member with its own dt[m]: dq[m] starts
 * at zero as in rv_lsrk_run (its entry values unread when the first a is 0),
 * each step replays a captured hipGraph of the scheme's stages, and q[m] holds
 * the result (the library ping-pongs through buffers of its own).  The graphs
 * are kept while q, dq, scheme and dt stay the same. */
int rv_batch_lsrk_run(rv_batch *b, void *const *q, void *const *dq, int scheme, const double *dt,
                      int64_t nsteps);

/* --- diagnostics ------------------------------------------------------- */

/* weightednorm(dg, q; componentwise = true) (measures.jl:6-14), p = 2, with
 * compensated (Neumaier) FP64 sums in a fixed order: out[s], s < Ns. */
int rv_weightednorm(rv_handle *h, const void *q, double *out);
/* weightednorm(dg, q, p; componentwise = true) (measures.jl:6-21) for any
 * p > 0: (sum MJ |q_s|^p)^(1/p), compensated as above; p = +infinity gives
 * max |q_s| (no weights, as Raven's Val{Inf} method).  out[s], s < Ns. */
int rv_weightednorm_p(rv_handle *h, const void *q, double p, double *out);
/* Compensated conservation integrals sum(MJ .* q_s): out[s], s < Ns. */
int rv_integrals(rv_handle *h, const void *q, double *out);
/* entropyintegral(dg, q) = sum(MJ .* entropy.(q)) (measures.jl:23-27) and
 * entropyproduct(dg, p, q) = sum(MJ .* entropyvariables.(p) . q)
 * (measures.jl:28-33), entropy and entropy variables of euler.jl:70-91 /
 * advection.jl:20-21, in FP64 with compensated sums in a fixed order. */
int rv_entropy_integral(rv_handle *h, const void *q, double *out);
int rv_entropy_product(rv_handle *h, const void *p, const void *q, double *out);

/* --- GatherScatter (gatherscatter.jl:1-103) ----------------------------- */

/* gather!: dest (ncontinuous, nf) = fold(+, ctod column, ascending rows)
 * of src (Nq^dim, nf, K).  Bit-exact with the reference order. */
int rv_gather(rv_handle *h, void *dest, const void *src, int nf);
/* scatter!: dest (Nq^dim, nf, K)[i] = src[dtoc[i]] */
int rv_scatter(rv_handle *h, void *dest, const void *src, int nf);

/* --- integer maps ------------------------------------------------------- */

/* Rebuild vmapM, vmapP, mapM, mapP ((2dim*Nfp, K) int64, 1-based, host
 * arrays) from the device-side face records, for bit-exact checks against
 * facemaps(grid) (lobattocells.jl:2277-2291).  Any pointer may be NULL. */
int rv_get_maps(rv_handle *h, int64_t *vmapM, int64_t *vmapP, int64_t *mapM, int64_t *mapP);

/* Number of cells of a brick descriptor (prod(n) * 2^(dim*level)), or a
 * negative RV_E_* code. */
int64_t rv_brick_numcells(const rv_brick_desc *desc);

/* Host arrays of the grid generate(GridManager(cell, brick(...))) builds for
 * the brick (gridmanager.jl:1084-1208): points (Np, dim, K), volume metrics,
 * surface metrics, vmapM, vmapP, boundary codes, dtoc (layouts as in
 * rv_desc); any pointer may be NULL.  Host only; no device is touched. */
int rv_brick_arrays(const rv_brick_desc *desc, double *points, double *metrics,
                    double *surface_metrics, int64_t *vmapM, int64_t *vmapP, int64_t *bc,
                    int64_t *dtoc, int64_t *ncontinuous);

/* Volume and surface metrics of curved cells from their node coordinates, on
 * the device: hexvolumemetrics! / quadvolumemetrics! (lobattocells.jl:
 * 3038-3180, 2756-2813; Kopriva's curl-invariant form in 3-D) and
 * hexsurfacemetrics! / quadsurfacemetrics! (:3253-3295, :2862-2900), in the
 * layouts of rv_desc.  points (Np, dim, K), metrics and surface_metrics are
 * DEVICE pointers; D (Nq x Nq column-major) and weights are host arrays.
 * Runs on `stream` and returns when the results are complete.  This is the
 * setup step generate(warp, gm) performs for warped grids
 * (gridmanager.jl:1111-1166); dim 2 or 3, Nq <= 8. */
int rv_curved_metrics(int dim, int nq, int64_t ncells, const double *points, const double *D,
                      const double *weights, double *metrics, double *surface_metrics,
                      void *stream);

/* min_node_distance(grid; dims) (grids.jl:94-145): the least distance between
 * a node and its neighbours at linear index ijk -+ stride(d) of the same cell,
 * over every node and every direction d whose bit is set in dims_mask (bit 0
 * is the first direction; the reference's default is all of them).  points
 * (Np, dim, K) is a DEVICE pointer; runs on `stream` and returns the minimum
 * in *out.  A mask naming no direction or one beyond dim is an invalid
 * argument (the reference's ArgumentError). */
int rv_min_node_distance(int dim, int nq, int64_t ncells, const double *points,
                         int32_t dims_mask, double *out, void *stream);

/* The integer lattice position of every cell of the brick at the finest level,
 * ijk[e*dim + d] (cell e spans [lo + h*ijk, lo + h*(ijk+1)) along d): the
 * quadrant coordinates generate() reads from the forest
 * (gridmanager.jl:476-639), without per-node arrays. */
int rv_brick_cells(const rv_brick_desc *desc, int32_t *ijk);

/* LobattoCell 1-D operators (lobattocells.jl:1-47): nodes, weights and D
 * (column-major) for Nq points, rounded from 256-bit closed forms. */
int rv_lobatto(int nq, double *points, double *weights, double *D);

#ifdef __cplusplus
}
#endif
#endif
