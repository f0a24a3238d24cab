/*
 * gpair_io_internal.h -- FITS helpers shared by the library's own files (not
 * part of the public ABI): multi-HDU image files for loop checkpoints, and the
 * snapshot hooks of the CG loops.
 */
#ifndef GPAIR_IO_INTERNAL_H
#define GPAIR_IO_INTERNAL_H

#include "../../include/gpair_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A file of images: the first one added is the primary HDU, the others IMAGE
 * extensions, each with an EXTNAME.  bitpix -32 (float), -64 (double) or 32
 * (int32); data in host byte order, n1 fastest. */
typedef struct gpair_fitsw gpair_fitsw;
gpair_fitsw *gpair_fitsw_new(void);
int gpair_fitsw_add(gpair_fitsw *w, const char *extname, int bitpix, const void *data, long n1, long n2);
int gpair_fitsw_save(gpair_fitsw *w, const char *path);
void gpair_fitsw_free(gpair_fitsw *w);

/* Reads the image called extname (its type and sizes must be as given). */
typedef struct gpair_fitsr gpair_fitsr;
int gpair_fitsr_open(const char *path, gpair_fitsr **r);
int gpair_fitsr_get(gpair_fitsr *r, const char *extname, int bitpix, void *out, long n1, long n2);
int gpair_fitsr_has(gpair_fitsr *r, const char *extname);
void gpair_fitsr_close(gpair_fitsr *r);

/* snapshot hooks (gpair_set_snapshot): due at objective evaluation `it`, write */
int gpair_snapshot_due(const gpair_ctx *ctx, int it);
int gpair_snapshot_write(gpair_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* GPAIR_IO_INTERNAL_H */
