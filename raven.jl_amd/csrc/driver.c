/*
 * rv_driver: a C caller of libravenhip that goes through the C ABI alone (no
 * Python, no device allocator of its own), the way a Julia or C host would.
 *
 * It reproduces Raven's time loop for the 2-D advection case of
 * test/balancelaws/advection/wave_2d.jl's set-up family (16x16 periodic brick
 * on [0,1]^2, Nq = 5, u = (1, 0.5), a Gaussian pulse), in one of two modes:
 *   run    solve!(q, T, LSRK54(dg, q, dt)) in one call (rv_lsrk_run,
 *          odesolvers.jl:9-71);
 *   stage  dostep! stage by stage (odesolvers.jl:50-61): dq .*= rka;
 *          dg(dq, q, t + rkc dt); q .+= rkb dt dq, through rv_lsrk_stage with
 *          the coefficients of rv_lsrk_coefficients, final step clipped to T
 *          as solve! does (odesolvers.jl:20-25).
 * The final state is written as raw FP64 to the output file, and one JSON
 * line reports the step count, end time and weightednorm (measures.jl:6-14).
 *
 * usage: rv_driver run|stage <steps-worth of T in units of dt> <out.bin>
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ravenhip.h"

static int check(int rc, rv_handle *h, const char *what)
{
    if (rc != RV_OK) {
        fprintf(stderr, "rv_driver: %s failed (%d): %s\n", what, rc, rv_last_error(h));
        exit(1);
    }
    return rc;
}

int main(int argc, char **argv)
{
    if (argc != 4 || (strcmp(argv[1], "run") && strcmp(argv[1], "stage"))) {
        fprintf(stderr, "usage: rv_driver run|stage <steps> <out.bin>\n");
        return 2;
    }
    const int stage_mode = strcmp(argv[1], "stage") == 0;
    const double nsteps = atof(argv[2]);

    rv_brick_desc bd;
    memset(&bd, 0, sizeof bd);
    bd.abi_version = RAVENHIP_ABI_VERSION;
    bd.dim = 2;
    bd.nq = 5;
    bd.law = RV_LAW_ADVECTION;
    bd.flux = RV_FLUX_RUSANOV;
    bd.dtype = RV_F64;
    bd.n[0] = bd.n[1] = 16;
    bd.periodic[0] = bd.periodic[1] = 1;
    bd.hi[0] = bd.hi[1] = 1.0;
    bd.gamma = 1.4;
    bd.velocity[0] = 1.0;
    bd.velocity[1] = 0.5;
    bd.volume_form = RV_FORM_WEAK;

    const int64_t K = rv_brick_numcells(&bd);
    if (K <= 0)
        return check((int)K, NULL, "rv_brick_numcells");
    const int64_t Np = 25, n = Np * K;
    double *points = malloc(sizeof(double) * Np * 2 * K);
    double *q = malloc(sizeof(double) * n);
    check(rv_brick_arrays(&bd, points, NULL, NULL, NULL, NULL, NULL, NULL, NULL), NULL,
          "rv_brick_arrays");
    for (int64_t e = 0; e < K; ++e)
        for (int64_t i = 0; i < Np; ++i) {
            const double x = points[i + Np * (0 + 2 * e)], y = points[i + Np * (1 + 2 * e)];
            q[i + Np * e] = exp(-((x - 0.5) * (x - 0.5) + (y - 0.5) * (y - 0.5)) / 0.01);
        }

    rv_handle *h = NULL;
    check(rv_create_brick(&bd, &h), NULL, "rv_create_brick");
    void *qd = NULL, *dqd = NULL, *qtmp = NULL;
    check(rv_state_alloc(h, &qd), h, "rv_state_alloc");
    check(rv_state_alloc(h, &dqd), h, "rv_state_alloc");
    check(rv_state_alloc(h, &qtmp), h, "rv_state_alloc");
    check(rv_state_upload(h, qd, q), h, "rv_state_upload");

    /* dt = cfl * h / N / |u| with cfl 1/2 (wave_2d.jl:38) */
    const double dt = 0.5 * (1.0 / 16) / 4 / sqrt(1.0 * 1.0 + 0.5 * 0.5);
    const double T = nsteps * dt;
    double t = 0.0;
    int64_t steps = 0;
    if (!stage_mode) {
        check(rv_lsrk_run(h, qd, dqd, RV_LSRK54, 0.0, dt, T, &t, &steps), h, "rv_lsrk_run");
    } else {
        double a[14], b[14], c[14];
        int nst = 0;
        check(rv_lsrk_coefficients(RV_LSRK54, a, b, c, &nst), h, "rv_lsrk_coefficients");
        /* LSRK's dq starts at zero (odesolvers.jl:43-44) */
        memset(q, 0, sizeof(double) * n);
        check(rv_state_upload(h, dqd, q), h, "rv_state_upload");
        double step_dt = dt;
        void *cur = qd, *alt = qtmp;
        for (;;) {
            ++steps;
            int final_step = 0;
            if (t + step_dt >= T) {
                step_dt = T - t;
                final_step = 1;
            }
            for (int s = 0; s < nst; ++s) {
                check(rv_lsrk_stage(h, alt, cur, dqd, a[s], b[s] * step_dt, t + c[s] * step_dt),
                      h, "rv_lsrk_stage");
                void *tmp = cur;
                cur = alt;
                alt = tmp;
            }
            t += step_dt;
            if (final_step)
                break;
        }
        qd = cur;
        qtmp = alt;
    }
    double comp[1];
    check(rv_weightednorm(h, qd, comp), h, "rv_weightednorm");
    check(rv_state_download(h, q, qd), h, "rv_state_download");
    FILE *f = fopen(argv[3], "wb");
    if (!f || fwrite(q, sizeof(double), (size_t)n, f) != (size_t)n) {
        fprintf(stderr, "rv_driver: cannot write %s\n", argv[3]);
        return 1;
    }
    fclose(f);
    printf("{\"mode\": \"%s\", \"steps\": %lld, \"time\": %.17g, \"dt\": %.17g, "
           "\"weightednorm\": %.17g}\n",
           argv[1], (long long)steps, t, dt, comp[0]);
    check(rv_state_free(h, qd), h, "rv_state_free");
    check(rv_state_free(h, dqd), h, "rv_state_free");
    check(rv_state_free(h, qtmp), h, "rv_state_free");
    rv_destroy(h);
    free(points);
    free(q);
    return 0;
}
