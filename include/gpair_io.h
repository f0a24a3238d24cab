/*
 * gpair_io.h -- host-side data I/O of the GPAIR path: OIFITS loading and
 * writing (replacing gpair1/trunk/getoifits.c + read_fits.c, which sit on
 * cfitsio and John Young's exchange routines) and FITS image reading and
 * writing (replacing read_fits_image / writefits, gpair1/trunk/gpair.c:954-995,
 * 1164-1194).  Plain C, no device and no cfitsio needed.
 *
 * Errors: every call returns GPAIR_OK (0) or a GPAIR_E_* code of gpair_hip.h;
 * gpair_io_last_error() gives the message of the calling thread's last failure.
 */
#ifndef GPAIR_IO_H
#define GPAIR_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char *gpair_io_last_error(void);

/* Data read from an OIFITS file for one target and one waveband, in the
 * reference's oi_data layout (gpair1/trunk/getoifits.h:61-76):
 *   uv points: the npow V2 points first (one per V2 datum, in file order), then
 *   every T3 leg that matched no earlier point; all in the +u half plane.
 *   uvpnt / sign: 3 legs (ab, bc, ca) per T3; sign -1 when the leg is the
 *   conjugate of the stored point. */
typedef struct {
    int npow, nbis, nuv;
    float *pow, *powerr;                             /* [npow] */
    float *bisamp, *bisamperr, *bisphs, *bisphserr;  /* [nbis] (degrees for phases) */
    float *u, *v, *wavelength, *bandwidth;           /* [nuv] (u, v in wavelengths; metres) */
    int64_t *uvpnt;                                  /* [nbis][3] */
    int16_t *sign;                                   /* [nbis][3] */
    int target_id;                                   /* the selection used */
    float minband, maxband;                          /* nm */
    long long uv_compares;                           /* uv comparisons the deduplication made */
} gpair_oidata;

/* Targets of an OIFITS file (OI_TARGET): up to max ids and names (17 chars
 * each, NUL-terminated); *count receives the number in the file. */
int gpair_oifits_targets(const char *path, int *ids, char *names, int max, int *count);

/* get_oi_fits_selection + get_oi_fits_data, gpair1/trunk/getoifits.c:65-357,
 * 359-733, without the interactive prompts: target_id < 0 selects the only
 * target (an error when there are several, as the reference would prompt);
 * minband < 0 or maxband <= minband selects 1 .. 50000 nm (the reference's
 * automatic full-band choice, getoifits.c:608-613).  A V2 / T3 datum is taken
 * when its table's first record has the target id, its channel's EFF_WAVE*1e9
 * lies strictly between minband and maxband and it is not flagged; a NaN T3
 * amplitude gives amplitude 1 with error (float)1e99, as the reference.  T3
 * legs are matched against the stored uv points with the reference's relative
 * threshold 5e-5 and equal wavelength, first match in index order, through a
 * hash on (wavelength, log|u|, log|v|) cells instead of the reference's linear
 * scan; the result is the same.  Release the arrays with gpair_oidata_free. */
int gpair_oifits_read(const char *path, int target_id, float minband, float maxband, gpair_oidata *out);
void gpair_oidata_free(gpair_oidata *d);

/* No reference counterpart (the reference only reads OIFITS): writes an OIFITS
 * v1 file with one target (id 1), one OI_ARRAY / OI_WAVELENGTH (nwave channels,
 * metres), an OI_VIS2 table of nvis2 records and an OI_T3 table of nt3 records.
 * Per record: coordinates in metres; data arrays [record][nwave]; flags may be
 * NULL (nothing flagged); t3 phases in degrees. */
int gpair_oifits_write(const char *path, const char *target, int nwave, const float *eff_wave,
                       const float *eff_band, int nvis2, const double *v2_u, const double *v2_v,
                       const double *vis2data, const double *vis2err, const unsigned char *v2_flag, int nt3,
                       const double *t3_u1, const double *t3_v1, const double *t3_u2, const double *t3_v2,
                       const double *t3amp, const double *t3amperr, const double *t3phi, const double *t3phierr,
                       const unsigned char *t3_flag);

/* read_fits_image, gpair1/trunk/gpair.c:1164-1194: the primary image (any
 * BITPIX, BSCALE / BZERO applied) as float, x fastest.  With out == NULL only
 * the sizes are returned; otherwise capacity must hold width * height floats. */
int gpair_fits_read_image(const char *path, float *out, long capacity, int *width, int *height);
/* writefits, gpair1/trunk/gpair.c:954-990: a primary image of width x height
 * pixels (BITPIX -64 as the reference, or -32 when bitpix == -32) with the
 * reference's DATAFILE, TARGET, PIXSIZE (mas) and WIDTH keywords; datafile and
 * target may be NULL ("dummy"). */
int gpair_fits_write_image(const char *path, const float *image, int width, int height, int bitpix,
                           float pixsize, const char *datafile, const char *target);

#ifdef __cplusplus
}
#endif

#endif /* GPAIR_IO_H */
