/*
 * Driver of the reference's own OIFITS loader (TEST INFRASTRUCTURE ONLY).
 *
 * Linked by `make -C oracle ref` with gpair1/trunk/getoifits.c and
 * gpair1/trunk/read_fits.c compiled from where they lie under the reference
 * tree, against the installed cfitsio.  It runs the reference's selection and
 * extraction, get_oi_fits_selection + get_oi_fits_data
 * (gpair1/trunk/getoifits.c:359-714, 65-357), with the band given here (the
 * reference prompts only when it is not set) and, for files with several
 * targets, the target id read by the reference's own prompt from stdin; then it
 * writes what the reference extracted, in the oi_data layout
 * (gpair1/trunk/getoifits.h:61-75), to a binary file:
 *   int32 npow, nbis, nuv
 *   float32 pow[npow], powerr[npow], bisamp[nbis], bisamperr[nbis],
 *           bisphs[nbis], bisphserr[nbis], u[nuv], v[nuv], wavelength[nuv],
 *           bandwidth[nuv]
 *   int64 uvpnt[nbis][3] (ab, bc, ca), int16 sign[nbis][3]
 *
 *   usage: ref_oifits <file.oifits> <out.bin> <minband nm> <maxband nm>
 */
#include "getoifits.h"

#include <stdint.h>

static void put(FILE *f, const void *p, size_t size, size_t n)
{
    if (n && fwrite(p, size, n, f) != n) {
        fprintf(stderr, "ref_oifits: short write\n");
        exit(2);
    }
}

int main(int argc, char **argv)
{
    if (argc != 5) {
        fprintf(stderr, "usage: %s file.oifits out.bin minband maxband\n", argv[0]);
        return 2;
    }
    static oi_usersel usersel;   /* zero-initialised, as the reference's global is */
    static oi_data data;
    int status = 0;
    snprintf(usersel.file, sizeof(usersel.file), "%s", argv[1]);
    usersel.minband = (float)atof(argv[3]);
    usersel.maxband = (float)atof(argv[4]);
    get_oi_fits_selection(&usersel, &status);
    get_oi_fits_data(&usersel, &data, &status);
    if (status) {
        fprintf(stderr, "ref_oifits: cfitsio status %d\n", status);
        return 1;
    }
    FILE *f = fopen(argv[2], "wb");
    if (!f)
        return 2;
    const int32_t hdr[3] = {data.npow, data.nbis, data.nuv};
    put(f, hdr, sizeof(int32_t), 3);
    put(f, data.pow, sizeof(float), (size_t)data.npow);
    put(f, data.powerr, sizeof(float), (size_t)data.npow);
    put(f, data.bisamp, sizeof(float), (size_t)data.nbis);
    put(f, data.bisamperr, sizeof(float), (size_t)data.nbis);
    put(f, data.bisphs, sizeof(float), (size_t)data.nbis);
    put(f, data.bisphserr, sizeof(float), (size_t)data.nbis);
    for (int k = 0; k < 4; k++)
        for (int i = 0; i < data.nuv; i++) {
            const oi_uv *p = &data.uv[i];
            const float x = k == 0 ? p->u : k == 1 ? p->v : k == 2 ? p->wavelength : p->bandwidth;
            put(f, &x, sizeof(float), 1);
        }
    for (int i = 0; i < data.nbis; i++) {
        const oi_uvpnt *leg[3] = {&data.bsref[i].ab, &data.bsref[i].bc, &data.bsref[i].ca};
        for (int j = 0; j < 3; j++) {
            const int64_t v = leg[j]->uvpnt;
            put(f, &v, sizeof(int64_t), 1);
        }
    }
    for (int i = 0; i < data.nbis; i++) {
        const oi_uvpnt *leg[3] = {&data.bsref[i].ab, &data.bsref[i].bc, &data.bsref[i].ca};
        for (int j = 0; j < 3; j++) {
            const int16_t s = leg[j]->sign;
            put(f, &s, sizeof(int16_t), 1);
        }
    }
    fclose(f);
    return 0;
}
