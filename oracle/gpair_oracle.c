/*
 * gpair_oracle.c -- FP64 CPU restatement of the GPAIR MEM-CG reconstruction
 * path.  TEST INFRASTRUCTURE ONLY: tests, the smoke check and the CPU leg of
 * the benchmark load it as the checker.  The product (libgpair_hip.so) never
 * links, loads or calls it.
 *
 * Every function restates one piece of the reference; the citation is in the
 * comment above it (paths relative to the reference repository root).  All
 * arithmetic is double precision (the reference mixes float and double).
 *
 * Conventions shared with the product:
 *   image[x + W*y] holds I[y][x]                 (gpair1/trunk/cpu.c:469)
 *   au[p] = RPMAS*pixellation*u[p]  (turns per pixel along x)
 *   av[p] = RPMAS*pixellation*v[p]  (turns per pixel along y)
 *   X[p][x] = exp(+2 pi i au[p] x), Y[p][y] = exp(-2 pi i av[p] y)
 *                                                 (gpair1/trunk/gpair.c:242-243)
 *   data/err/mock: [V2 (npow) | (Re, Im) per T3 (2*nbis)]
 *   phasor: complex per T3, uvpnt/sign: 3 per T3 (ab, bc, ca)
 *
 * Parity status of each function is listed in oracle/README.md.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int orc_verbose = 0;
void orc_set_verbose(int v) { orc_verbose = v; }

/* Build-side safeguard, absent from the reference: at most this many J
 * evaluations per line search (bracket + zoom).  The reference zoom can loop
 * without bound when quadratic interpolation keeps landing next to an end of
 * [lo, hi] and the bracket loop never ends on a NaN objective.  When the cap
 * is hit the line search returns its current low end (a step that satisfied
 * the sufficient-decrease test, or 0). */
static int orc_ls_cap = 64;
void orc_set_ls_cap(int cap) { orc_ls_cap = cap; }

#define ORC_TWO_PI 6.283185307179586476925286766559

/* phase of exp(2 pi i t) reduced to one turn in FP64 before the cosine/sine */
static inline double complex turn_exp(double t)
{
    double f = t - nearbyint(t);
    return cos(ORC_TWO_PI * f) + I * sin(ORC_TWO_PI * f);
}

/* DFT tables, gpair1/trunk/gpair.c:231-245 (there: float complex [nuv][W],
 * phase evaluated with cexp of a double argument). Here complex128 [nuv][W]
 * stored as interleaved (re, im) doubles. */
void orc_tables(int W, int nuv, const double *au, const double *av, double *X, double *Y)
{
#pragma omp parallel for schedule(static)
    for (int p = 0; p < nuv; p++) {
        for (int i = 0; i < W; i++) {
            double complex x = turn_exp(au[p] * (double)i);
            double complex y = turn_exp(-av[p] * (double)i);
            X[2 * ((size_t)p * W + i)] = creal(x);
            X[2 * ((size_t)p * W + i) + 1] = cimag(x);
            Y[2 * ((size_t)p * W + i)] = creal(y);
            Y[2 * ((size_t)p * W + i) + 1] = cimag(y);
        }
    }
}

/* Total flux F = sum I: gpair1/trunk/cpu.c:270-277 (compute_flux), used by
 * gpu_compute_flux, gpair1/trunk/gpu.c:860-884. */
double orc_flux(int n, const double *img)
{
    double f = 0.0;
    for (int i = 0; i < n; i++)
        f += img[i];
    return f;
}

/* image -> complex visibilities, gpair1/trunk/cpu.c:449-475 (image2vis) and
 * the visi kernel gpair1/trunk/kernel_visi.cl:26-66:
 *   V_p = (1/F) sum_y sum_x I[y][x] X[p][x] Y[p][y].
 * Returns F. vis: 2*nuv doubles. */
double orc_image2vis(int W, int nuv, const double *au, const double *av, const double *img,
                     double *vis)
{
    double F = orc_flux(W * W, img);
    double inv = (F > 0.0) ? 1.0 / F : 1.0;
#pragma omp parallel
    {
        double complex *xr = malloc(sizeof(double complex) * (size_t)W);
#pragma omp for schedule(static)
        for (int p = 0; p < nuv; p++) {
            for (int i = 0; i < W; i++)
                xr[i] = turn_exp(au[p] * (double)i);
            double complex acc = 0.0;
            for (int y = 0; y < W; y++) {
                double complex t = 0.0;
                const double *row = img + (size_t)y * W;
                for (int x = 0; x < W; x++)
                    t += row[x] * xr[x];
                acc += t * turn_exp(-av[p] * (double)y);
            }
            acc *= inv;
            vis[2 * p] = creal(acc);
            vis[2 * p + 1] = cimag(acc);
        }
        free(xr);
    }
    return F;
}

static inline double complex leg(const double *vis, int idx, int sign)
{
    double complex v = vis[2 * idx] + I * vis[2 * idx + 1];
    return sign < 0 ? conj(v) : v;
}

/* visibilities -> mock data, gpair1/trunk/cpu.c:758-802 (vis2data) and the
 * kernels gpair1/trunk/kernel_pow_spec.cl:1-11, kernel_bispec.cl:52-81:
 *   mock[k] = |V_k|^2 (k < npow)
 *   T_t = V'_ab V'_bc V'_ca rho_t, ' = conj when sign < 0
 *   mock[npow+2t] = Re T_t, mock[npow+2t+1] = Im T_t */
void orc_vis2data(int npow, int nbis, const double *vis, const int *uvpnt, const int *sign,
                  const double *phasor, double *mock)
{
    for (int k = 0; k < npow; k++)
        mock[k] = vis[2 * k] * vis[2 * k] + vis[2 * k + 1] * vis[2 * k + 1];
    for (int t = 0; t < nbis; t++) {
        double complex T = leg(vis, uvpnt[3 * t], sign[3 * t]) *
                           leg(vis, uvpnt[3 * t + 1], sign[3 * t + 1]) *
                           leg(vis, uvpnt[3 * t + 2], sign[3 * t + 2]) *
                           (phasor[2 * t] + I * phasor[2 * t + 1]);
        mock[npow + 2 * t] = creal(T);
        mock[npow + 2 * t + 1] = cimag(T);
    }
}

/* chi2 = sum ((mock - data) * err)^2, gpair1/trunk/cpu.c:371-383 (data2chi2),
 * gpair1/trunk/kernel_chi2.cl:1-11 plus the reduction of gpu.c:1454-1515. */
double orc_data2chi2(int n, const double *data, const double *err, const double *mock)
{
    double c = 0.0;
    for (int i = 0; i < n; i++) {
        double r = (mock[i] - data[i]) * err[i];
        c += r * r;
    }
    return c;
}

/* Gull-Skilling entropy, gpair1/trunk/cpu.c:385-400 and
 * gpair1/trunk/kernel_entropy_gs.cl:3-26. */
double orc_entropy(int n, const double *img, const double *model)
{
    double S = 0.0;
    for (int i = 0; i < n; i++) {
        if (img[i] > 0.0 && model[i] > 0.0)
            S += img[i] - model[i] - img[i] * log(img[i] / model[i]);
        else
            S += -model[i];
    }
    return S;
}

/* entropy gradient, gpair1/trunk/cpu.c:402-412 and
 * gpair1/trunk/kernel_entropy_gs_grad.cl:3-24: grad = -ln(I/M), and the old
 * value is KEPT (grad is in/out) where I <= 0 or M <= 0. */
void orc_entropy_grad(int n, const double *img, const double *model, double *grad)
{
    for (int i = 0; i < n; i++)
        if (img[i] > 0.0 && model[i] > 0.0)
            grad[i] = -log(img[i] / model[i]);
}

/* chi2 gradient by the reference's per-pixel formula, gpair1/trunk/cpu.c:279-369
 * (compute_data_gradient) == kernel_grad_pow.cl:58-103 + kernel_grad_bis.cl:53-137.
 * O(W^2 (npow + nbis)): small problems only. vis/mock must belong to img. */
void orc_grad_direct(int W, int nuv, int npow, int nbis, const double *au, const double *av,
                     const double *img, const double *vis, const double *mock,
                     const double *data, const double *err, const int *uvpnt, const int *sign,
                     const double *phasor, double *grad)
{
    (void)nuv;
    double F = orc_flux(W * W, img);
    double invflux = 1.0 / F;
#pragma omp parallel for schedule(static)
    for (int jj = 0; jj < W; jj++) {
        for (int ii = 0; ii < W; ii++) {
            double g = 0.0;
            for (int k = 0; k < npow; k++) {
                double complex V = vis[2 * k] + I * vis[2 * k + 1];
                double complex E = turn_exp(au[k] * ii) * turn_exp(-av[k] * jj);
                g += 4.0 * err[k] * err[k] * invflux * (mock[k] - data[k]) * creal(conj(V) * (E - V));
            }
            for (int t = 0; t < nbis; t++) {
                int a = uvpnt[3 * t], b = uvpnt[3 * t + 1], c = uvpnt[3 * t + 2];
                double complex vab = leg(vis, a, sign[3 * t]);
                double complex vbc = leg(vis, b, sign[3 * t + 1]);
                double complex vca = leg(vis, c, sign[3 * t + 2]);
                double complex eab = turn_exp(au[a] * ii) * turn_exp(-av[a] * jj);
                double complex ebc = turn_exp(au[b] * ii) * turn_exp(-av[b] * jj);
                double complex eca = turn_exp(au[c] * ii) * turn_exp(-av[c] * jj);
                if (sign[3 * t] < 0) eab = conj(eab);
                if (sign[3 * t + 1] < 0) ebc = conj(ebc);
                if (sign[3 * t + 2] < 0) eca = conj(eca);
                double complex d = ((eab - vab) * vbc * vca + vab * (ebc - vbc) * vca + vab * vbc * (eca - vca)) *
                                   (phasor[2 * t] + I * phasor[2 * t + 1]) * invflux;
                int r = npow + 2 * t;
                g += 2.0 * err[r] * err[r] * (mock[r] - data[r]) * creal(d);
                g += 2.0 * err[r + 1] * err[r + 1] * mock[r + 1] * cimag(d);
            }
            grad[ii + jj * W] = g;
        }
    }
}

/* Per-uv adjoint coefficients of the separable gradient form (derived from
 * gpair1/trunk/cpu.c:310-365):
 *   V2:  a_k += 4 e_k^2 (m_k - d_k) conj(V_k)
 *   T3:  c_t = 2 e_r^2 (m_r - d_r) - i 2 e_i^2 m_i,
 *        b = c_t rho_t (product of the other two legs'), a_p += b (sign>0) or conj(b)
 * coef: 2*nuv doubles (zeroed here). Returns c0 = Re sum_p a_p V_p. */
double orc_grad_coeffs(int nuv, int npow, int nbis, const double *vis, const double *mock,
                       const double *data, const double *err, const int *uvpnt, const int *sign,
                       const double *phasor, double *coef)
{
    double complex *a = calloc((size_t)nuv, sizeof(double complex));
    for (int k = 0; k < npow; k++) {
        double complex V = vis[2 * k] + I * vis[2 * k + 1];
        a[k] += 4.0 * err[k] * err[k] * (mock[k] - data[k]) * conj(V);
    }
    for (int t = 0; t < nbis; t++) {
        int r = npow + 2 * t;
        double complex ct = 2.0 * err[r] * err[r] * (mock[r] - data[r]) - I * 2.0 * err[r + 1] * err[r + 1] * mock[r + 1];
        double complex rho = phasor[2 * t] + I * phasor[2 * t + 1];
        double complex l[3];
        for (int j = 0; j < 3; j++)
            l[j] = leg(vis, uvpnt[3 * t + j], sign[3 * t + j]);
        for (int j = 0; j < 3; j++) {
            double complex b = ct * rho * l[(j + 1) % 3] * l[(j + 2) % 3];
            a[uvpnt[3 * t + j]] += (sign[3 * t + j] < 0) ? conj(b) : b;
        }
    }
    double c0 = 0.0;
    for (int p = 0; p < nuv; p++) {
        coef[2 * p] = creal(a[p]);
        coef[2 * p + 1] = cimag(a[p]);
        c0 += creal(a[p] * (vis[2 * p] + I * vis[2 * p + 1]));
    }
    free(a);
    return c0;
}

/* Separable adjoint: grad[y][x] = (Re sum_p a_p X[p][x] Y[p][y] - c0) / F. */
void orc_grad_separable(int W, int nuv, const double *au, const double *av, const double *coef,
                        double c0, double F, double *grad)
{
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    size_t n = (size_t)W * W;
    double *part = calloc(n * (size_t)nthr, sizeof(double));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *g = part + n * (size_t)tid;
        double complex *xr = malloc(sizeof(double complex) * (size_t)W);
        double complex *by = malloc(sizeof(double complex) * (size_t)W);
#pragma omp for schedule(static)
        for (int p = 0; p < nuv; p++) {
            double complex a = coef[2 * p] + I * coef[2 * p + 1];
            for (int i = 0; i < W; i++) {
                xr[i] = turn_exp(au[p] * (double)i);
                by[i] = a * turn_exp(-av[p] * (double)i);
            }
            for (int y = 0; y < W; y++) {
                double br = creal(by[y]), bi = cimag(by[y]);
                double *row = g + (size_t)y * W;
                for (int x = 0; x < W; x++)
                    row[x] += br * creal(xr[x]) - bi * cimag(xr[x]);
            }
        }
        free(xr);
        free(by);
    }
    double inv = 1.0 / F;
    for (size_t i = 0; i < n; i++) {
        double s = 0.0;
        for (int t = 0; t < nthr; t++)
            s += part[(size_t)t * n + i];
        grad[i] = (s - c0) * inv;
    }
    free(part);
}

/* ------------------------------------------------------------------------
 * The MEM-CG loop with the GPU-path control flow of the reference:
 * gpair1/trunk/gpair.c:682-915 (outer loop, beta, bracket phase) and
 * gpair1/trunk/gpu.c:1947-2065 (gpu_linesearch_zoom).  Scalars are FP64 here.
 * Build-side addition: stop when chi2/ndof <= chi2r_target after evaluating J
 * on the current image (the reference runs a fixed iteration count).
 * ---------------------------------------------------------------------- */
typedef struct {
    int W, nuv, npow, nbis;
    const double *au, *av, *data, *err, *phasor, *model;
    const int *uvpnt, *sign;
    double alpha;
    /* work */
    double *vis, *mock, *coef, *gdata, *gent;
    int nJ, nG;
} orc_prob;

static double eval_J(orc_prob *P, const double *img, double *chi2, double *S)
{
    int n = P->W * P->W;
    orc_image2vis(P->W, P->nuv, P->au, P->av, img, P->vis);
    orc_vis2data(P->npow, P->nbis, P->vis, P->uvpnt, P->sign, P->phasor, P->mock);
    *chi2 = orc_data2chi2(P->npow + 2 * P->nbis, P->data, P->err, P->mock);
    *S = orc_entropy(n, img, P->model);
    P->nJ++;
    return *chi2 - P->alpha * *S;
}

/* gradient of J at img; requires vis/mock from the last eval_J on img
 * (the reference's implicit contract, gpair1/trunk/gpu.c:1598-1675). */
static void eval_grad(orc_prob *P, const double *img, double *g)
{
    int n = P->W * P->W;
    double F = orc_flux(n, img);
    double c0 = orc_grad_coeffs(P->nuv, P->npow, P->nbis, P->vis, P->mock, P->data, P->err, P->uvpnt,
                                P->sign, P->phasor, P->coef);
    orc_grad_separable(P->W, P->nuv, P->au, P->av, P->coef, c0, F, P->gdata);
    orc_entropy_grad(n, img, P->model, P->gent);
    for (int i = 0; i < n; i++)
        g[i] = P->gdata[i] - P->alpha * P->gent[i];
    P->nG++;
}

static double dot(int n, const double *a, const double *b)
{
    double s = 0.0;
    for (int i = 0; i < n; i++)
        s += a[i] * b[i];
    return s;
}

static void step_image(int n, const double *cur, const double *d, double s, double minv, double *out)
{
    for (int i = 0; i < n; i++) {
        double t = cur[i] + s * d[i];
        out[i] = t < minv ? minv : t;
    }
}

/* ------------------------------------------------------------------------
 * The reference's line search on the 1-D objective phi(s) = J(max(I + s d, floor))
 * with slope w1 = d.g at s = 0: the bracket phase of gpair1/trunk/gpair.c:791-906
 * and the GPU path's zoom, gpair1/trunk/gpu.c:1947-2065, with their constants
 * (c1 = 1e-4, c2 = 0.1, growth x10, s_max = 100, interval width 1e-14) and
 * quirks: the zoom's curvature test is the weak one (w2 >= c2 w1, gpu.c:2045)
 * where the bracket's is the strong one (|w2| <= -c2 w1, gpair.c:865); the
 * zoom never updates phi(lo) (gpu.c:2020 compares with the J_lo it was given);
 * its step is the minimiser of the quadratic through (0, J0) with slope w1 and
 * (s_old, J_old) (gpu.c:1976-1978) when that lies inside [lo, hi], else the
 * midpoint; and it accepts after 10 interpolations whatever the curvature
 * (counter > 10, gpu.c:2045).  Scalars are FP64 (the reference's are float).
 *
 * phi(user, s) evaluates J at step s; dphi(user, s) the slope d.grad J at the
 * image of the last phi call (the reference's gradient uses the visibilities of
 * its last chi2 call, gpu.c:1598-1675), and is only called right after phi(s).
 *
 * info (may be NULL), 8 doubles about the returned step:
 *   [0] exit: ORC_LS_BRACKET (strong Wolfe in the bracket), ORC_LS_ZOOM_CURV
 *       (Armijo and weak curvature in the zoom), ORC_LS_ZOOM_COUNT (Armijo,
 *       counter > 10), ORC_LS_ZOOM_WIDTH (|hi - lo| < 1e-14: returns lo),
 *       ORC_LS_CAP (evaluation cap: returns the bracket's s_old or the zoom's lo)
 *   [1] the returned step, [2] J there (the last evaluation for the first three
 *   exits, the J the search holds for lo / s_old otherwise), [3] its slope w2
 *   (NaN when not evaluated), [4] J evaluations, [5] slope evaluations,
 *   [6] the zoom's counter (-1 without zoom), [7] the zoom's last |hi - lo|.
 *
 * Planted mistakes (orc_set_plant, for the tests that show the closed-form
 * cases catch them; 0 in every other use): bit 0 the zoom takes the strong
 * curvature test, bit 1 the zoom updates J_lo when lo moves (the textbook
 * rule), bit 2 the zoom bisects only.
 * ---------------------------------------------------------------------- */
enum { ORC_LS_BRACKET = 1, ORC_LS_ZOOM_CURV = 2, ORC_LS_ZOOM_COUNT = 3, ORC_LS_ZOOM_WIDTH = 4, ORC_LS_CAP = 5 };
typedef double (*orc_phi_fn)(void *user, double s);

static int orc_plant = 0;
void orc_set_plant(int mask) { orc_plant = mask; }

#define LS_C1 1e-4
#define LS_C2 0.1
#define LS_SMAX 100.0

static void ls_info(double *info, int exitc, double s, double J, double w2, int nJ, int nG, int counter,
                    double width)
{
    if (!info)
        return;
    info[0] = exitc; info[1] = s; info[2] = J; info[3] = w2;
    info[4] = nJ; info[5] = nG; info[6] = counter; info[7] = width;
}

/* gpair1/trunk/gpu.c:1947-2065; *nJ / *nG count on from their values (the
 * bracket's), cap bounds *nJ */
static double ls_zoom(orc_phi_fn phi, orc_phi_fn dphi, void *user, double lo, double hi, double J_lo,
                      double w1, double J_init, int cap, int *nJ, int *nG, double *info)
{
    double s = 0.0, s_old = 0.0, J_old = 0.0, J_at_lo = J_lo;
    int counter = 0;
    for (;;) {
        if (!(orc_plant & 4) && counter > 0 && (J_old - J_init - w1 * s_old) != 0.0)
            s = fabs(w1 * s_old * s_old / (2.0 * (J_old - J_init - w1 * s_old)));
        if (counter == 0 || (orc_plant & 4) || s < lo || s > hi)
            s = (hi - lo) / 2.0 + lo;
        if (fabs(hi - lo) < 1e-14) {
            ls_info(info, ORC_LS_ZOOM_WIDTH, lo, J_at_lo, NAN, *nJ, *nG, counter, fabs(hi - lo));
            return lo;
        }
        if (*nJ >= cap) {
            ls_info(info, ORC_LS_CAP, lo, J_at_lo, NAN, *nJ, *nG, counter, fabs(hi - lo));
            return lo;
        }
        const double J = phi(user, s);
        (*nJ)++;
        if (orc_verbose)
            fprintf(stderr, "zoom c=%d s=%.6e lo=%.6e hi=%.6e J=%.10e Jlo=%.10e Jinit=%.10e w1=%.4e\n", counter, s,
                    lo, hi, J, J_lo, J_init, w1);
        if (J > J_init + LS_C1 * s * w1 || J >= J_lo) {
            hi = s;
        } else {
            const double w2 = dphi(user, s);
            (*nG)++;
            const int curv = (orc_plant & 1) ? fabs(w2) <= -LS_C2 * w1 : w2 >= LS_C2 * w1;
            if (curv || counter > 10) {
                ls_info(info, curv ? ORC_LS_ZOOM_CURV : ORC_LS_ZOOM_COUNT, s, J, w2, *nJ, *nG, counter,
                        fabs(hi - lo));
                return s;
            }
            if (w2 * (hi - lo) >= 0.0)
                hi = lo;
            lo = s;
            J_at_lo = J;
            if (orc_plant & 2)
                J_lo = J;
        }
        s_old = s;
        J_old = J;
        counter++;
    }
}

/* the zoom alone, as gpu_linesearch_zoom is called (tests) */
double orc_ls_zoom(orc_phi_fn phi, orc_phi_fn dphi, void *user, double lo, double hi, double J_lo, double w1,
                   double J_init, int cap, double *info)
{
    int nJ = 0, nG = 0;
    return ls_zoom(phi, dphi, user, lo, hi, J_lo, w1, J_init, cap, &nJ, &nG, info);
}

/* bracket phase, gpair1/trunk/gpair.c:791-906, with the zoom above; J_init is
 * phi(0), w1 its slope */
double orc_linesearch(orc_phi_fn phi, orc_phi_fn dphi, void *user, double J_init, double w1, int cap,
                      double *info)
{
    double s = 1.0, s_old = 0.0, J_old = J_init;
    int lsit = 1, nJ = 0, nG = 0;
    for (;;) {
        if (nJ >= cap) {
            ls_info(info, ORC_LS_CAP, s_old, J_old, NAN, nJ, nG, -1, 0.0);
            return s_old;
        }
        const double Jt = phi(user, s);
        nJ++;
        if (orc_verbose)
            fprintf(stderr, "bracket s=%.6e J=%.10e Jold=%.10e w1=%.4e\n", s, Jt, J_old, w1);
        if (Jt > J_init + LS_C1 * s * w1 || (Jt >= J_old && lsit > 1))
            return ls_zoom(phi, dphi, user, s_old, s, J_old, w1, J_init, cap, &nJ, &nG, info);
        const double w2 = dphi(user, s);
        nG++;
        if (fabs(w2) <= -LS_C2 * w1) {
            ls_info(info, ORC_LS_BRACKET, s, Jt, w2, nJ, nG, -1, 0.0);
            return s;
        }
        if (w2 >= 0.0)
            return ls_zoom(phi, dphi, user, s, s_old, Jt, w1, J_init, cap, &nJ, &nG, info);
        s_old = s;
        J_old = Jt;
        s = 10.0 * s;
        if (s > LS_SMAX)
            s = LS_SMAX;
        lsit++;
    }
}

/* phi and its slope for the MEM-CG objective along d from cur */
typedef struct {
    orc_prob *P;
    const double *cur, *d;
    double *tmp, *gt, minv;
} cg_line;

static double cg_phi(void *u, double s)
{
    cg_line *L = u;
    double chi2, S;
    step_image(L->P->W * L->P->W, L->cur, L->d, s, L->minv, L->tmp);
    return eval_J(L->P, L->tmp, &chi2, &S);
}

static double cg_dphi(void *u, double s)
{
    cg_line *L = u;
    (void)s;   /* the image of the last cg_phi call */
    eval_grad(L->P, L->tmp, L->gt);
    return dot(L->P->W * L->P->W, L->d, L->gt);
}

/* optional per-iteration log of orc_cg's line searches: info rows of 8 doubles
 * (above) followed by w1 and J0 of the search (10 doubles per row) */
static double *orc_ls_log = NULL;
static int orc_ls_log_rows = 0;
void orc_set_ls_log(double *buf, int rows)
{
    orc_ls_log = buf;
    orc_ls_log_rows = buf ? rows : 0;
}

/* Runs at most max_iter outer iterations of MEM-CG on img (in/out).
 * method: 0 steepest descent, 1 Fletcher-Reeves, 2 Polak-Ribiere, 3 Hestenes-Stiefel.
 * trace (may be NULL): 8 doubles per outer iteration:
 *   J, chi2, S, selected step, beta, nJ (cumulative), nG (cumulative), flux.
 * Returns the number of outer iterations whose J was evaluated (trace rows). */
int orc_cg(int W, int nuv, int npow, int nbis, const double *au, const double *av, const double *data,
           const double *err, const int *uvpnt, const int *sign, const double *phasor, const double *model,
           double alpha, int method, int max_iter, double chi2r_target, double *img, double *trace)
{
    const double minv = 1e-8;
    int n = W * W, ndof = npow + 2 * nbis;
    orc_prob P = {W, nuv, npow, nbis, au, av, data, err, phasor, model, uvpnt, sign, alpha,
                  NULL, NULL, NULL, NULL, NULL, 0, 0};
    P.vis = malloc(sizeof(double) * 2 * (size_t)nuv);
    P.coef = malloc(sizeof(double) * 2 * (size_t)nuv);
    P.mock = malloc(sizeof(double) * (size_t)ndof);
    P.gdata = malloc(sizeof(double) * (size_t)n);
    P.gent = calloc((size_t)n, sizeof(double));
    double *g = calloc((size_t)n, sizeof(double)), *g_old = calloc((size_t)n, sizeof(double));
    double *d = calloc((size_t)n, sizeof(double)), *tmp = malloc(sizeof(double) * (size_t)n);
    double *gt = calloc((size_t)n, sizeof(double));
    double beta = 0.0, sel = 0.0, chi2, S;
    int rows = 0;
    for (int it = 0; it < max_iter; it++) {
        double J = eval_J(&P, img, &chi2, &S);
        if (trace) {
            double *r = trace + 8 * rows;
            r[0] = J; r[1] = chi2; r[2] = S; r[3] = sel; r[4] = beta;
            r[5] = P.nJ; r[6] = P.nG; r[7] = orc_flux(n, img);
        }
        rows++;
        if (chi2 / ndof <= chi2r_target)
            break;
        eval_grad(&P, img, g);
        /* beta, gpair1/trunk/gpair.c:727-758 */
        if (it == 0 || method == 0) {
            beta = 0.0;
        } else {
            double gg = dot(n, g, g), go = dot(n, g, g_old), oo = dot(n, g_old, g_old);
            if (method == 1)
                beta = gg / oo;
            if (method == 2) {
                beta = (gg - go) / oo;
                if (beta < 0.0) beta = 0.0;
            }
            if (method == 3) {
                beta = (gg - go) / (dot(n, d, g) - dot(n, d, g_old));
                if (beta < 0.0) beta = 0.0;
            }
            if (fabs(go / gg) > 0.5)
                beta = 0.0;
        }
        for (int i = 0; i < n; i++)
            d[i] = beta * d[i] - g[i];
        double w1 = dot(n, d, g);
        /* bracket phase and zoom, gpair1/trunk/gpair.c:791-906, gpu.c:1947-2065 */
        cg_line L = {&P, img, d, tmp, gt, minv};
        double info[8];
        sel = orc_linesearch(cg_phi, cg_dphi, &L, J, w1, orc_ls_cap, info);
        if (orc_ls_log && it < orc_ls_log_rows) {
            memcpy(orc_ls_log + 10 * it, info, sizeof(info));
            orc_ls_log[10 * it + 8] = w1;
            orc_ls_log[10 * it + 9] = J;
        }
        step_image(n, img, d, sel, minv, img);
        memcpy(g_old, g, sizeof(double) * (size_t)n);
    }
    free(P.vis); free(P.coef); free(P.mock); free(P.gdata); free(P.gent);
    free(g); free(g_old); free(d); free(tmp); free(gt);
    return rows;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------
 * Accelerated loop (a build-side method; the reference has no counterpart):
 * preconditioned Polak-Ribiere CG with a batched line search.
 *   z = P o g (P = 1, or the MEM image metric P = I, optionally times a
 *   per-pixel factor), beta = max(0, z.(g - g_old) / z_old.g_old), d = -z + beta d,
 *   restart d = -z when d.g >= 0.
 * Line search: the DFT is linear, so with U_I = A I, U_d = A d (A = unnormalised
 * DFT) the visibilities of I + s d are (U_I + s U_d) / (F_I + s F_d).  All
 * LS_K candidate steps (s_0 = 0, s_k = s_ref 2^((k - LS_TOP)/2)) are scored at
 * once from that (the pixel floor is applied to the entropy term; in V it is
 * ignored); a step is taken when some s_k > 0 scores below s_0, otherwise the
 * window moves down 16 octaves (4 rounds), then the direction restarts as -z,
 * a parabola through the best candidate and its neighbours gives s*, and s* is
 * committed with an exact evaluation that must pass the Armijo test
 * (J <= J0 + c1 s w1), falling back to the best grid step and then halving.
 * ---------------------------------------------------------------------- */
#define LS_K 32
#define LS_TOP 26

void orc_dft_raw(int W, int nuv, const double *au, const double *av, const double *img, double *out)
{
#pragma omp parallel
    {
        double complex *xr = malloc(sizeof(double complex) * (size_t)W);
#pragma omp for schedule(static)
        for (int p = 0; p < nuv; p++) {
            for (int i = 0; i < W; i++)
                xr[i] = turn_exp(au[p] * (double)i);
            double complex acc = 0.0;
            for (int y = 0; y < W; y++) {
                double complex t = 0.0;
                const double *row = img + (size_t)y * W;
                for (int x = 0; x < W; x++)
                    t += row[x] * xr[x];
                acc += t * turn_exp(-av[p] * (double)y);
            }
            out[2 * p] = creal(acc);
            out[2 * p + 1] = cimag(acc);
        }
        free(xr);
    }
}

/* chi2 of the visibilities (U_I + s U_d) / (F_I + s F_d) */
static double chi2_along(orc_prob *P, const double *UI, const double *Ud, double FI, double Fd, double s,
                         double *vis_tmp, double *mock_tmp)
{
    double inv = 1.0 / (FI + s * Fd);
    for (int p = 0; p < P->nuv; p++) {
        vis_tmp[2 * p] = (UI[2 * p] + s * Ud[2 * p]) * inv;
        vis_tmp[2 * p + 1] = (UI[2 * p + 1] + s * Ud[2 * p + 1]) * inv;
    }
    orc_vis2data(P->npow, P->nbis, vis_tmp, P->uvpnt, P->sign, P->phasor, mock_tmp);
    return orc_data2chi2(P->npow + 2 * P->nbis, P->data, P->err, mock_tmp);
}

static double entropy_along(orc_prob *P, const double *img, const double *d, double s, double minv)
{
    int n = P->W * P->W;
    double S = 0.0;
    for (int i = 0; i < n; i++) {
        double t = img[i] + s * d[i];
        if (t < minv) t = minv;
        double M = P->model[i];
        S += (t > 0.0 && M > 0.0) ? (t - M - t * log(t / M)) : -M;
    }
    return S;
}

/* parabola through (a, Ja), (b, Jb), (c, Jc) with a < b < c: vertex if it is a
 * minimum inside (a, c), else b */
double orc_parabola_step(double a, double b, double c, double Ja, double Jb, double Jc)
{
    double p = (b - a) * (Jb - Jc), q = (b - c) * (Jb - Ja);
    double den = p - q;
    if (!(den < 0.0 || den > 0.0))
        return b;
    double s = b - 0.5 * ((b - a) * p - (b - c) * q) / den;
    double curv = (Ja - Jb) / (a - b) - (Jb - Jc) / (b - c);   /* >0 for a minimum (a<b<c) */
    if (!(curv < 0.0) || !(s > a && s < c))
        return b;
    return s;
}

/* trace: 8 doubles per outer iteration: J, chi2, S, step, beta, forward DFTs
 * (cumulative), adjoint DFTs (cumulative), line-search batches (cumulative). */
int orc_pcg(int W, int nuv, int npow, int nbis, const double *au, const double *av, const double *data,
            const double *err, const int *uvpnt, const int *sign, const double *phasor, const double *model,
            double alpha, int precond, const double *pfactor, int max_iter, double chi2r_target, double *img,
            double *trace)
{
    const double c1 = 1e-4, minv = 1e-8;
    int n = W * W, ndof = npow + 2 * nbis;
    orc_prob P = {W, nuv, npow, nbis, au, av, data, err, phasor, model, uvpnt, sign, alpha,
                  NULL, NULL, NULL, NULL, NULL, 0, 0};
    P.vis = malloc(sizeof(double) * 2 * (size_t)nuv);
    P.coef = malloc(sizeof(double) * 2 * (size_t)nuv);
    P.mock = malloc(sizeof(double) * (size_t)ndof);
    P.gdata = malloc(sizeof(double) * (size_t)n);
    P.gent = calloc((size_t)n, sizeof(double));
    double *g = calloc((size_t)n, sizeof(double)), *g_old = calloc((size_t)n, sizeof(double));
    double *z = calloc((size_t)n, sizeof(double)), *z_old = calloc((size_t)n, sizeof(double));
    double *d = calloc((size_t)n, sizeof(double)), *tmp = malloc(sizeof(double) * (size_t)n);
    double *UI = malloc(sizeof(double) * 2 * (size_t)nuv), *Ud = malloc(sizeof(double) * 2 * (size_t)nuv);
    double *vt = malloc(sizeof(double) * 2 * (size_t)nuv), *mt = malloc(sizeof(double) * (size_t)ndof);
    int nfwd = 0, nadj = 0, nbatch = 0, rows = 0;
    double chi2, S, beta = 0.0, sel = 0.0, s_ref = 0.0, zg_old = 0.0;
    double J = eval_J(&P, img, &chi2, &S);
    nfwd++;
    for (int it = 0; it < max_iter; it++) {
        if (trace) {
            double *r = trace + 8 * rows;
            r[0] = J; r[1] = chi2; r[2] = S; r[3] = sel; r[4] = beta; r[5] = nfwd; r[6] = nadj; r[7] = nbatch;
        }
        rows++;
        if (chi2 / ndof <= chi2r_target)
            break;
        eval_grad(&P, img, g);
        nadj++;
        for (int i = 0; i < n; i++) {
            double pf = precond ? img[i] : 1.0;
            if (pfactor)
                pf *= pfactor[i];
            z[i] = pf * g[i];
        }
        const double zg = dot(n, z, g), zgo = dot(n, z, g_old);
        const double FI = orc_flux(n, img);
        for (int p = 0; p < nuv; p++) {
            UI[2 * p] = P.vis[2 * p] * FI;
            UI[2 * p + 1] = P.vis[2 * p + 1] * FI;
        }
        double Jk[LS_K], sk[LS_K], w1 = 0.0;
        int kbest = -1;
        const double s_ref_it = s_ref;   /* a restart along -z searches from here again */
        /* attempt 0: conjugate direction; attempt 1 (only when attempt 0 used
         * beta > 0 and found no decrease): restart along -z */
        for (int attempt = 0; attempt < 2 && kbest < 0; attempt++) {
            double b0 = 0.0;
            if (attempt == 0 && it > 0) {
                b0 = (zg - zgo) / zg_old;
                if (!(b0 > 0.0)) b0 = 0.0;
            }
            if (attempt == 1 && beta == 0.0)
                break;
            if (attempt == 1)
                s_ref = s_ref_it;   /* not the 2^-64 the failed rounds shrank it to */
            beta = b0;
            /* pixels on the floor do not move further down: the direction is
             * projected onto the feasible side there, so d.g is the slope of
             * the clamped path */
            for (int pass = 0; pass < 2; pass++) {
                for (int i = 0; i < n; i++) {
                    d[i] = -z[i] + beta * d[i];
                    if (img[i] <= minv && d[i] < 0.0)
                        d[i] = 0.0;
                }
                w1 = dot(n, d, g);
                if (w1 < 0.0 || beta == 0.0)
                    break;
                /* the product scores a direction before it reads d.g back: a
                 * conjugate direction that is not a descent one costs it the DFT
                 * of d, one batch and the commit launched behind the batch
                 * (counted here so the counters agree) */
                nfwd += 2;
                nbatch++;
                beta = 0.0;
            }
            if (!(w1 < 0.0))
                break;
            double Fd = 0.0, dabs = 0.0;
            for (int i = 0; i < n; i++) {
                Fd += d[i];
                dabs += fabs(d[i]);
            }
            if (s_ref == 0.0)
                s_ref = FI / dabs;
            orc_dft_raw(W, nuv, au, av, d, Ud);
            nfwd++;
            /* candidate 0 is s = 0: every candidate is compared with the same
             * linearised evaluation of the current image */
            for (int round = 0; round < 4 && kbest < 0; round++) {
                /* the product launches the commit at the chosen step right behind
                 * each batch, before the host sees the batch: one DFT per batch,
                 * which is the first commit try when the batch gives a step */
                nbatch++;
                nfwd++;
                for (int k = 0; k < LS_K; k++) {
                    sk[k] = k == 0 ? 0.0 : s_ref * pow(2.0, 0.5 * (k - LS_TOP));
                    double c2 = chi2_along(&P, UI, Ud, FI, Fd, sk[k], vt, mt);
                    Jk[k] = c2 - alpha * entropy_along(&P, img, d, sk[k], minv);
                }
                int kb = 0;
                for (int k = 1; k < LS_K; k++)
                    if (Jk[k] < Jk[kb]) kb = k;
                if (kb > 0)
                    kbest = kb;
                else
                    s_ref *= pow(2.0, -16.0);
            }
        }
        if (kbest < 0) {
            if (orc_verbose)
                fprintf(stderr, "pcg it %d: no decrease (J %.10e, w1 %.3e, s_ref %.3e)\n", it, J, w1, s_ref);
            break;   /* no decrease along d at any scale: stalled */
        }
        double s = sk[kbest];
        if (kbest > 0 && kbest < LS_K - 1)
            s = orc_parabola_step(sk[kbest - 1], sk[kbest], sk[kbest + 1], Jk[kbest - 1], Jk[kbest], Jk[kbest + 1]);
        /* commit with an exact evaluation and the Armijo test */
        double Jn = 0.0, chi2n = 0.0, Sn = 0.0;
        int ok = 0;
        for (int tries = 0; tries < 32; tries++) {
            step_image(n, img, d, s, minv, tmp);
            Jn = eval_J(&P, tmp, &chi2n, &Sn);
            if (tries > 0)
                nfwd++;
            if (Jn <= J + c1 * s * w1) {
                ok = 1;
                break;
            }
            s = (tries == 0 && s != sk[kbest]) ? sk[kbest] : 0.5 * s;
        }
        if (!ok) {
            if (orc_verbose)
                fprintf(stderr, "pcg it %d: commit failed (J %.10e, Jn %.10e, s %.3e, w1 %.3e)\n", it, J, Jn, s, w1);
            break;
        }
        memcpy(img, tmp, sizeof(double) * (size_t)n);
        J = Jn;
        chi2 = chi2n;
        S = Sn;
        sel = s;
        s_ref = s;
        memcpy(g_old, g, sizeof(double) * (size_t)n);
        memcpy(z_old, z, sizeof(double) * (size_t)n);
        zg_old = zg;
    }
    free(P.vis); free(P.coef); free(P.mock); free(P.gdata); free(P.gent);
    free(g); free(g_old); free(z); free(z_old); free(d); free(tmp); free(UI); free(Ud); free(vt); free(mt);
    return rows;
}

/* ---------------------------------------------------------------------------
 * Single-pixel moves: update_vis_fluxchange (gpair1/trunk/cpu.c:477-486) and
 * update_vis_positionchange (gpair1/trunk/cpu.c:488-497), restated through the
 * identity they rest on.  With U = F V the unnormalised visibilities of an image
 * of total flux F, changing pixel (x, y) by d gives
 *   V'_p = (F V_p + d X[p][x] Y[p][y]) / (F + d),
 * the reference's V' = r V + (1 - r) X Y with r = F / (F + d): the ratio of the
 * TOTAL fluxes before and after (cpu.c:480-481 leave the total to the caller).
 * A position change moves flux f between two pixels: F is unchanged and
 * V' = V + (f / F) (X_new Y_new - X_old Y_old).  Table entries are evaluated
 * from au, av here (turn_exp), not read from tables.
 * ------------------------------------------------------------------------- */
/* npx pixel changes d[j] at pix[j] = x + W*y; vis, vis_new: 2*nuv doubles;
 * returns the new flux */
double orc_px_update(int W, int nuv, const double *au, const double *av, const double *vis, double F, int npx,
                     const int *pix, const double *d, double *vis_new)
{
    double Fn = F;
    for (int j = 0; j < npx; j++)
        Fn += d[j];
#pragma omp parallel for schedule(static)
    for (int p = 0; p < nuv; p++) {
        double complex acc = F * (vis[2 * p] + I * vis[2 * p + 1]);
        for (int j = 0; j < npx; j++) {
            const int x = pix[j] % W, y = pix[j] / W;
            acc += d[j] * turn_exp(au[p] * (double)x) * turn_exp(-av[p] * (double)y);
        }
        acc /= Fn;
        vis_new[2 * p] = creal(acc);
        vis_new[2 * p + 1] = cimag(acc);
    }
    return Fn;
}

static double px_pix_entropy(double t, double M)
{
    return (t > 0.0 && M > 0.0) ? (t - M - t * log(t / M)) : -M;
}

/* Scores `count` independent moves against an image (img, its vis and flux F):
 * move k sets pixel pix_a[k] to flux_a[k] and, when pix_b[k] >= 0, pixel
 * pix_b[k] to flux_b[k].  dchi2[k] = chi2 after - chi2 before (vis2data +
 * data2chi2 of the updated visibilities), dS[k] = entropy after - before. */
void orc_px_propose(int W, int nuv, int npow, int nbis, const double *au, const double *av, const double *data,
                    const double *err, const int *uvpnt, const int *sign, const double *phasor,
                    const double *model, const double *img, const double *vis, double F, int count,
                    const int *pix_a, const int *pix_b, const double *flux_a, const double *flux_b, double *dchi2,
                    double *dS)
{
    const int nd = npow + 2 * nbis;
    double *mock0 = malloc(sizeof(double) * (size_t)nd);
    orc_vis2data(npow, nbis, vis, uvpnt, sign, phasor, mock0);
    const double chi0 = orc_data2chi2(nd, data, err, mock0);
    free(mock0);
#pragma omp parallel
    {
        double *vn = malloc(sizeof(double) * 2 * (size_t)nuv);
        double *mk = malloc(sizeof(double) * (size_t)nd);
        double complex *ca = malloc(sizeof(double complex) * (size_t)nuv);
#pragma omp for schedule(dynamic)
        for (int k = 0; k < count; k++) {
            int pix[2] = {pix_a[k], pix_b[k]};
            double d[2] = {flux_a[k] - img[pix_a[k]], pix_b[k] >= 0 ? flux_b[k] - img[pix_b[k]] : 0.0};
            const int npx = pix_b[k] >= 0 ? 2 : 1;
            double Fn = F + d[0] + (npx > 1 ? d[1] : 0.0);
            for (int p = 0; p < nuv; p++) {
                double complex acc = F * (vis[2 * p] + I * vis[2 * p + 1]);
                for (int j = 0; j < npx; j++)
                    acc += d[j] * turn_exp(au[p] * (double)(pix[j] % W)) * turn_exp(-av[p] * (double)(pix[j] / W));
                ca[p] = acc / Fn;
                vn[2 * p] = creal(ca[p]);
                vn[2 * p + 1] = cimag(ca[p]);
            }
            orc_vis2data(npow, nbis, vn, uvpnt, sign, phasor, mk);
            dchi2[k] = orc_data2chi2(nd, data, err, mk) - chi0;
            double s = 0.0;
            for (int j = 0; j < npx; j++)
                s += px_pix_entropy(j == 0 ? flux_a[k] : flux_b[k], model[pix[j]]) -
                     px_pix_entropy(img[pix[j]], model[pix[j]]);
            dS[k] = s;
        }
        free(vn);
        free(mk);
        free(ca);
    }
}
