/*
 * gpair_cg.c -- the MEM-CG reconstruction loop as C host code over the C ABI.
 *
 * Control flow of the reference's GPU path, gpair1/trunk/gpair.c:682-915:
 *   J(I) = chi2(I) - alpha S(I); gradient; beta by method (SD/FR/PR/HS) with
 *   the |g.g_old|/|g.g| > 0.5 restart; d = beta d - g; bracket phase from
 *   s = 1 growing x10 up to s_max with the Armijo / strong-Wolfe tests, zoom
 *   (gpu_linesearch_zoom) when bracketed; I = max(I + s d, minvalue).
 * Scalars are FP64.  Additions of this build: the chi2/ndof stop rule, the
 * per-line-search evaluation cap, status codes instead of exit().
 */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "../../include/gpair_hip.h"
#include "gpair_io_internal.h"

void gpair_cg_default_params(gpair_cg_params *p)
{
    memset(p, 0, sizeof(*p));
    p->alpha = 1000.0;        /* gpair1/trunk/gpair.c:85 */
    p->method = 2;            /* Polak-Ribiere */
    p->max_iter = 10;         /* gpair1/trunk/gpair.c:255 */
    p->chi2r_target = 1.0;
    p->ls_max_evals = 64;
    p->minvalue = 1e-8;       /* gpair1/trunk/gpair.c:265 */
    p->wolfe1 = 1e-4;         /* gpair1/trunk/gpair.c:268 */
    p->wolfe2 = 0.1;
    p->step_max = 100.0;      /* gpair1/trunk/gpair.c:803 */
    p->verbose = 0;
}

#define TRY(x)                    \
    do {                          \
        int rc_ = (x);            \
        if (rc_ != GPAIR_OK) {    \
            res.status = rc_;     \
            goto done;            \
        }                         \
    } while (0)

int gpair_cg_run(gpair_ctx *ctx, const gpair_cg_params *prm, gpair_cg_result *result, double *trace,
                 int trace_rows)
{
    gpair_cg_result res;
    memset(&res, 0, sizeof(res));
    if (!ctx || !prm)
        return GPAIR_E_ARG;
    int W = 0, nuv = 0, npow = 0, nbis = 0;
    gpair_sizes(ctx, &W, &nuv, &npow, &nbis);
    const int n = W * W, ndof = npow + 2 * nbis;
    const double alpha = prm->alpha, c1 = prm->wolfe1, c2 = prm->wolfe2, minv = prm->minvalue;
    gpair_buf fgn = gpu_getp_fgn(ctx), fg = gpu_getp_fg(ctx), dd = gpu_getp_dd(ctx), tg = gpu_getp_tg(ctx);
    double beta = 0.0, sel = 0.0, chi2 = 0.0, S = 0.0;

    for (int it = 0; it < prm->max_iter; it++) {
        TRY(gpu_get_chi2_curr(ctx, nuv, npow, nbis, ndof, nuv, &chi2));
        TRY(gpu_get_entropy_curr(ctx, W, &S));
        const double J = chi2 - alpha * S;
        res.criterion_evals++;
        if (!isfinite(J)) {
            res.status = GPAIR_E_NUMERIC;
            goto done;
        }
        res.iterations = it + 1;
        res.J = J;
        res.chi2 = chi2;
        res.entropy = S;
        res.chi2r = chi2 / ndof;
        if (trace && it < trace_rows) {
            double *r = trace + 8 * it;
            r[0] = J; r[1] = chi2; r[2] = S; r[3] = sel; r[4] = beta;
            r[5] = res.criterion_evals; r[6] = res.grad_evals; r[7] = 0.0;
        }
        if (prm->verbose)
            printf("it %d J %.10e chi2r %.6f S %.6e step %.3e beta %.3e evals %d/%d\n", it, J, chi2 / ndof, S, sel,
                   beta, res.criterion_evals, res.grad_evals);
        if (gpair_snapshot_due(ctx, it))   /* gpair1/trunk/gpair.c:698-702 */
            TRY(gpair_snapshot_write(ctx));
        if (prm->chi2r_target > 0.0 && chi2 / ndof <= prm->chi2r_target) {
            res.converged = 1;
            break;
        }
        TRY(gpu_compute_data_gradient_curr(ctx, nuv, npow, nbis, W));
        TRY(gpu_compute_entropy_gradient_curr(ctx, W));
        TRY(gpu_compute_criterion_gradient_curr(ctx, W, alpha));
        res.grad_evals++;

        /* beta, gpair1/trunk/gpair.c:727-758 */
        if (it == 0 || prm->method == 0) {
            beta = 0.0;
        } else {
            double gg, go, oo;
            TRY(gpu_get_scalprod(ctx, W, W, fgn, fgn, &gg));
            TRY(gpu_get_scalprod(ctx, W, W, fgn, fg, &go));
            TRY(gpu_get_scalprod(ctx, W, W, fg, fg, &oo));
            if (prm->method == 1)
                beta = gg / oo;
            if (prm->method == 2) {
                beta = (gg - go) / oo;
                if (beta < 0.0) beta = 0.0;
            }
            if (prm->method == 3) {
                double dg, dgo;
                TRY(gpu_get_scalprod(ctx, W, W, dd, fgn, &dg));
                TRY(gpu_get_scalprod(ctx, W, W, dd, fg, &dgo));
                beta = (gg - go) / (dg - dgo);
                if (beta < 0.0) beta = 0.0;
            }
            if (fabs(go / gg) > 0.5)
                beta = 0.0;
        }
        TRY(gpu_compute_descent_dir(ctx, W, beta));
        double w1;
        TRY(gpu_get_scalprod(ctx, W, W, dd, fgn, &w1));

        /* bracket phase, gpair1/trunk/gpair.c:791-906 */
        double s = 1.0, s_old = 0.0, J_init = J, J_old = J;
        int lsit = 1, evals = 0;
        for (;;) {
            if (prm->ls_max_evals > 0 && evals >= prm->ls_max_evals) {
                sel = s_old;
                break;
            }
            double chi2t, St;
            TRY(gpu_update_tempimage(ctx, W, s, minv, dd));
            TRY(gpu_get_chi2_temp(ctx, nuv, npow, nbis, ndof, nuv, &chi2t));
            TRY(gpu_get_entropy_temp(ctx, W, &St));
            const double Jt = chi2t - alpha * St;
            res.criterion_evals++;
            evals++;
            if (Jt > J_init + c1 * s * w1 || (Jt >= J_old && lsit > 1)) {
                const int cap = prm->ls_max_evals > 0 ? prm->ls_max_evals - evals : -1;
                TRY(gpu_linesearch_zoom(ctx, nuv, npow, nbis, ndof, nuv, W, s_old, s, J_old, w1, J_init,
                                        &res.criterion_evals, &res.grad_evals, dd, tg, alpha, cap, &sel));
                break;
            }
            double w2;
            TRY(gpu_compute_data_gradient_temp(ctx, nuv, npow, nbis, W));
            TRY(gpu_compute_entropy_gradient_temp(ctx, W));
            TRY(gpu_compute_criterion_gradient_temp(ctx, W, alpha));
            res.grad_evals++;
            TRY(gpu_get_scalprod(ctx, W, W, dd, tg, &w2));
            if (fabs(w2) <= -c2 * w1) {
                sel = s;
                break;
            }
            if (w2 >= 0.0) {
                const int cap = prm->ls_max_evals > 0 ? prm->ls_max_evals - evals : -1;
                TRY(gpu_linesearch_zoom(ctx, nuv, npow, nbis, ndof, nuv, W, s, s_old, Jt, w1, J_init,
                                        &res.criterion_evals, &res.grad_evals, dd, tg, alpha, cap, &sel));
                break;
            }
            s_old = s;
            J_old = Jt;
            s = 10.0 * s;
            if (s > prm->step_max)
                s = prm->step_max;
            lsit++;
        }
        TRY(gpu_update_image(ctx, W, sel, minv, dd));
        TRY(gpu_backup_gradient(ctx, n, fgn, fg));
    }
done:
    if (result)
        *result = res;
    return res.status;
}
