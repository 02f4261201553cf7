/*
 * oracle.h -- CPU restatement of Raven.jl's DGSEM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is what results are checked
 * against: tests/, the smoke check in __graft_entry__.py and the cpu_baseline
 * leg of bench.py may load it; the HIP library (raven.jl_amd/csrc) never
 * links, loads or calls it.
 *
 * Every function names the Raven.jl source it restates (file:line, paths
 * relative to the Raven.jl repository root).  Arrays use Raven's memory
 * layout: Julia column-major, 1-based integer tables stored as int64.
 *
 *   state / volume fields  (Np, L, K)       offset ijk + Np*(s + L*e)
 *   surface metrics        (dim+2, Nfp*2dim, K) field-first (lobattocells.jl:3392-3400)
 *   vmapM / vmapP / mapM / mapP  (Nfp*2dim, K) int64, 1-based
 *   boundary codes         (2dim, K) int64
 *   D                      (Nq, Nq) column-major, D[a + Nq*b] = l_b'(x_a)
 */
#ifndef RAVEN_ORACLE_H
#define RAVEN_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Gauss-Lobatto nodes, weights and derivative matrix in __float128, rounded
 * to double (lobattocells.jl:1-47). */
void orc_gll(int nq, double *x, double *w, double *D);

/* Number of quadrants of a uniformly refined brick forest. */
int64_t orc_brick_numcells(int dim, const int32_t *n, int level);

/* Quadrant list of a brick forest refined uniformly to `level`:
 * tree order of p4est's brick connectivity (Morton over tree coordinates,
 * coarsegrids.jl:297 -> P4estTypes.brick), quadrants in Morton order inside a
 * tree.  tree_xyz[e*dim+d] = tree coordinate, quad_xyz[e*dim+d] = quadrant
 * coordinate inside the tree in units of 2^-level. */
void orc_brick_quadrants(int dim, const int32_t *n, int level, int32_t *tree_xyz,
                         int32_t *quad_xyz);

/* Brick node coordinates (lobattocells.jl:695-760, 801-872, 1590-1674).
 * coords[d] = vertex coordinates along d (n[d]+1 values).
 * points: (Np, dim, K). */
void orc_brick_points(int dim, int nq, int64_t K, int level, const int32_t *tree_xyz,
                      const int32_t *quad_xyz, const double *const *coords,
                      const double *r, double *points);

/* First-order volume metrics (dRdX, J, wJ): (Np, dim*dim+2, K).
 * brick != 0: *volumebrickmetrics! (lobattocells.jl:2590-2628, 2815-2860,
 * 3182-3251, including the 3-D element-index quirk at :3202);
 * brick == 0: *volumemetrics! (:2545-2588, 2756-2813, curl-invariant 3038-3180). */
void orc_volume_metrics(int dim, int nq, int64_t K, int brick, const double *points,
                        const double *D, const double *w, double *metrics);

/* Surface metrics (n, sJ, wsJ): (dim+2, 2dim*Nfp, K) (lobattocells.jl:2630-2665,
 * 2862-2900, 3253-3295). */
void orc_surface_metrics(int dim, int nq, int64_t K, const double *metrics,
                         const int64_t *vmapM, const double *w, double *surf);

/* Face maps and boundary codes of a conforming brick (lobattocells.jl:1992-2295,
 * quadranttoboundary :2076-2100). */
void orc_brick_facemaps(int dim, int nq, const int32_t *n, const int32_t *periodic,
                        int level, int64_t K, const int32_t *tree_xyz,
                        const int32_t *quad_xyz, int64_t *vmapM, int64_t *vmapP,
                        int64_t *mapM, int64_t *mapP, int64_t *bc);

/* Discontinuous-to-continuous numbering of a conforming brick: continuous ids
 * in first-appearance order (renumberdtocdtop!, lobattocells.jl:1872-1936).
 * Returns the number of continuous nodes. dtoc: (Np, K) int64 1-based. */
int64_t orc_brick_dtoc(int dim, int nq, const int32_t *n, const int32_t *periodic,
                       int level, int64_t K, const int32_t *tree_xyz,
                       const int32_t *quad_xyz, int64_t *dtoc);

/* GatherScatter (gatherscatter.jl:1-103): ctod columns folded with + from +0.0
 * in ascending row order; scatter dest[i] = src[dtoc[i]].  nf fields,
 * discontinuous layout (Np, nf, K), continuous layout (nc, nf). */
void orc_gather(int64_t np, int nf, int64_t K, int64_t nc, const int64_t *dtoc,
                const double *src, double *dest);
void orc_scatter(int64_t np, int nf, int64_t K, int64_t nc, const int64_t *dtoc,
                 const double *src, double *dest);

/* LSRK coefficients (odesolvers.jl:163-244). scheme 54 or 144. Returns stages. */
int orc_lsrk_coefficients(int scheme, double *a, double *b, double *c);

/* DGSEM operator (dgsem.jl:19-81): keeps pointers to caller arrays (the
 * caller keeps them alive) and owns MJI = 1 ./ MJ. law: 1 advection, 2 Euler. */
typedef struct orc_dg orc_dg;
orc_dg *orc_dg_new(int dim, int nq, int64_t K, int law, double gamma, const double *u,
                   const double *D, const double *metrics, const double *surf,
                   const int64_t *vmapM, const int64_t *vmapP, const int64_t *bc);
void orc_dg_free(orc_dg *dg);
int orc_dg_ns(const orc_dg *dg);
void orc_dg_set_threads(orc_dg *dg, int nthreads);

/* dg(dq, q, t; increment) (dgsem.jl:84-126): volume then surface term. */
void orc_dg_rhs(orc_dg *dg, double *dq, const double *q, double t, int increment);

/* One LSRK step (dostep!, odesolvers.jl:51-61) on q with register dq. */
void orc_dg_lsrk_step(orc_dg *dg, int scheme, double *q, double *dq, double time,
                      double dt);

/* solve!(q, timeend, LSRK(dg, q, dt; t0)) (odesolvers.jl:9-30,41-48): dq zeroed
 * internally. Returns the number of steps; *t_out = final solver time. */
int64_t orc_dg_solve(orc_dg *dg, int scheme, double *q, double dt, double t0,
                     double timeend, double *t_out);

/* Mortar coupling of hanging faces (the library's opt-in extension; Raven has
 * none): adds the mortar fluxes to dq.  kind (2dim, K), partner (4, 2dim, K)
 * as rv_forest_mortars, tohalves_1d lower/upper row-major, surf_full the
 * surface metrics with the non-conforming faces' lift weights kept. */
void orc_dg_mortar_add(const orc_dg *dg, const int32_t *kind, const int64_t *partner,
                       const double *lower, const double *upper, const double *surf_full,
                       const double *q, double *dq);

/* Volume form (dgsem.jl:1-17): 0 = WeakForm, 1 = FluxDifferencingForm with
 * volume_flux 2 = CentralFlux or 3 = EntropyConservativeFlux (Euler),
 * evaluated direction by direction as esvolumeterm_dir! (dgsem.jl:514-608). */
void orc_dg_set_volume_form(orc_dg *dg, int form, int volume_flux);
/* surface numerical flux: 1 Rusanov, 2 central, 3 entropy conservative
 * (the last two as twopointflux(q-, q+)' n, numericalfluxes.jl:27-37) */
void orc_dg_set_surface_flux(orc_dg *dg, int flux);
/* DGSEM(...; directions) as a bit mask (bit d: direction d + 1) */
void orc_dg_set_directions(orc_dg *dg, int mask);
/* sum MJ eta(q) and sum MJ v(p).q (measures.jl:23-33) */
double orc_dg_entropy_integral(const orc_dg *dg, const double *q);
double orc_dg_entropy_product(const orc_dg *dg, const double *p, const double *q);
/* one RLSRK step (odesolvers.jl:117-151); gamma in: previous root, out: this one;
 * dq carries between steps, k and q0 are work arrays of the state's size */
int orc_dg_rlsrk_step(orc_dg *dg, int scheme, double *q, double *dq, double *k, double *q0,
                      double time, double dt, double *gamma);
/* min_node_distance(grid; dims) (grids.jl:94-145), dims as a bit mask */
double orc_min_node_distance(int dim, int nq, int64_t K, const double *points, int mask);
/* Butcher weights of the low-storage scheme (odesolvers.jl:96-101); returns #stages */
int orc_rlsrk_weights(int scheme, double *bfull);

/* weightednorm(dg, q) component norms (measures.jl:6-14), p = 2. */
void orc_dg_weightednorm(const orc_dg *dg, const double *q, double *comp);

#ifdef __cplusplus
}
#endif
#endif
