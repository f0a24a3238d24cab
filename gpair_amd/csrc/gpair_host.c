/*
 * gpair_host.c -- host-side data preparation of the reconstruction path:
 * prior images (set_model, gpair1/trunk/cpu.c:645-749) and the packing of
 * OIFITS-style arrays into data / inverse errors / phasors (init_data,
 * gpair1/trunk/gpair.c:997-1066), including the triple-amplitude
 * extrapolation from V2.  Plain C, no device calls: usable without a GPU.
 * Arithmetic follows the reference's float / double mix so that results
 * match it bit for bit where the reference is deterministic.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../../include/gpair_hip.h"

/* gpair1/trunk/getoifits.h:87 */
#define REF_PI 3.14159265358979323

static float sq(float x) { return x * x; }

int gpair_set_model(float *model, int W, float pixellation, int type, float width, float flux, int centered)
{
    if (!model || W <= 0 || !(pixellation > 0.f))
        return GPAIR_E_ARG;
    const int n = W * W;
    float total = 0.0f;
    switch (type) {
    case 0: { /* flat, cpu.c:653-662 */
        float norm = (float)W * (float)W;
        for (int i = 0; i < n; i++)
            model[i] = flux / norm;
        return GPAIR_OK;
    }
    case 1: /* centred Dirac, cpu.c:664-673 */
        for (int i = 0; i < n; i++)
            model[i] = 1e-8f;
        model[(W * (W + 1)) / 2] = flux;
        return GPAIR_OK;
    case 2: /* uniform disk of radius width (mas), cpu.c:675-698 */
        for (int i = 0; i < W; i++)
            for (int j = 0; j < W; j++) {
                float rsq = sq((float)(W / 2 - i)) + sq((float)(W / 2 - j));
                model[j * W + i] = (sqrt(rsq) <= (width / pixellation)) ? 1.0f : 1e-8f;
                total += model[j * W + i];
            }
        break;
    case 3: { /* Gaussian of FWHM width (mas), cpu.c:700-722; the reference places
               * it 80 px off centre in y (cpu.c:714-715) unless centered != 0 */
        float sigma = width / (2. * sqrt(2. * log(2.)));
        float yoff = centered ? 0.0f : 80.0f;
        for (int i = 0; i < W; i++)
            for (int j = 0; j < W; j++) {
                model[j * W + i] = exp(-(sq((float)W / 2 - i) + sq((float)W / 2 - j - yoff)) /
                                       (2. * sq(sigma / pixellation)));
                total += model[j * W + i];
            }
        break;
    }
    case 4: { /* Lorentzian, cpu.c:724-745 */
        float sigma = width;
        for (int i = 0; i < W; i++)
            for (int j = 0; j < W; j++) {
                model[j * W + i] = sigma / (sq(sigma) + sq((float)W / 2 - i) + sq((float)W / 2 - j));
                total += model[j * W + i];
            }
        break;
    }
    default:
        return GPAIR_E_ARG;
    }
    for (int i = 0; i < n; i++)
        model[i] *= flux / total;
    return GPAIR_OK;
}

int gpair_init_data(int npow, int nbis, const float *pow, const float *powerr, float *bisamp, float *bisamperr,
                    const float *bisphs, const float *bisphserr, const int64_t *uvpnt, int do_extrapolation,
                    float *data, float *data_err, float *phasor, int *ndof)
{
    if (npow < 0 || nbis < 0 || (npow > 0 && (!pow || !powerr)) ||
        (nbis > 0 && (!bisamp || !bisamperr || !bisphs || !bisphserr || !uvpnt || !phasor)) || !data || !data_err)
        return GPAIR_E_ARG;
    int nd = npow + 2 * nbis;
    for (int k = 0; k < npow; k++) {
        data[k] = pow[k];
        data_err[k] = 1. / fabs(powerr[k]);
    }
    for (int t = 0; t < nbis; t++) {
        if (do_extrapolation == 1 && (bisamperr[t] <= 0. || bisamperr[t] > 1e3)) {
            const int64_t a = uvpnt[3 * t], b = uvpnt[3 * t + 1], c = uvpnt[3 * t + 2];
            if (a < npow && b < npow && c < npow && a >= 0 && b >= 0 && c >= 0) {
                /* optimal |V|^2 estimates and variances from the three V2 points,
                 * gpair1/trunk/gpair.c:1017-1038 */
                float p1 = pow[a], e1 = powerr[a], p2 = pow[b], e2 = powerr[b], p3 = pow[c], e3 = powerr[c];
                float s1 = (p1 + sqrt(sq(p1) + 2.0 * sq(e1))) / 2.;
                float v1 = 1. / (1. / s1 + 2. * (3. * s1 - p1) / sq(e1));
                float s2 = (p2 + sqrt(sq(p2) + 2.0 * sq(e2))) / 2.;
                float v2 = 1. / (1. / s2 + 2. * (3. * s2 - p2) / sq(e2));
                float s3 = (p3 + sqrt(sq(p3) + 2.0 * sq(e3))) / 2.;
                float v3 = 1. / (1. / s3 + 2. * (3. * s3 - p3) / sq(e3));
                bisamp[t] = sqrt(s1 * s2 * s3);
                bisamperr[t] = bisamp[t] * sqrt(v1 / s1 + v2 / s2 + v3 / s3);
            } else {
                /* no V2 for a leg: the amplitude carries no weight, gpair.c:1045-1050 */
                bisamp[t] = 1.0f;
                bisamperr[t] = 1e99;   /* float: +inf, so the weight below is 0 */
                nd--;
            }
        }
        data[npow + 2 * t] = fabs(bisamp[t]);
        data[npow + 2 * t + 1] = 0.;
        data_err[npow + 2 * t] = 1. / bisamperr[t];
        data_err[npow + 2 * t + 1] = 1. / fabs(bisamp[t] * bisphserr[t] * REF_PI / 180.);
        float complex ph = cexp(-I * bisphs[t] * REF_PI / 180.);
        phasor[2 * t] = crealf(ph);
        phasor[2 * t + 1] = cimagf(ph);
    }
    if (ndof)
        *ndof = nd;
    return GPAIR_OK;
}
