/*
 * CPU harness for the FITS / OIFITS readers of gpair_io.c (test infrastructure).
 * Built by tests/test_fits_fuzz.py with gcc -fsanitize=address,undefined: every
 * file named on the command line is handed to each reader; return codes are
 * ignored (crafted and mutated files are expected to be refused), only memory
 * and undefined-behaviour errors, which the sanitizers turn into a non-zero
 * exit, fail the run.
 */
#include "../include/gpair_io.h"
#include "../gpair_amd/csrc/gpair_io_internal.h"

#include <stdio.h>
#include <stdlib.h>

int main(int argc, char **argv)
{
    for (int i = 1; i < argc; i++) {
        const char *path = argv[i];
        int ids[8], count = 0, w = 0, h = 0;
        char names[8 * 17];
        (void)gpair_oifits_targets(path, ids, names, 8, &count);
        gpair_oidata d;
        if (gpair_oifits_read(path, -1, -1.0f, 0.0f, &d) == 0)
            gpair_oidata_free(&d);
        if (gpair_oifits_read(path, 1, 1.0f, 50000.0f, &d) == 0)
            gpair_oidata_free(&d);
        if (gpair_fits_read_image(path, NULL, 0, &w, &h) == 0 && (long)w * h <= (1L << 24)) {
            float *img = malloc(sizeof(float) * ((size_t)w * h + 1));
            if (img) {
                (void)gpair_fits_read_image(path, img, (long)w * h, &w, &h);
                free(img);
            }
        }
        gpair_fitsr *r = NULL;
        if (gpair_fitsr_open(path, &r) == 0) {
            double buf[64 * 64];
            if (gpair_fitsr_has(r, "STATE"))
                (void)gpair_fitsr_get(r, "STATE", -64, buf, 25, 1);
            (void)gpair_fitsr_get(r, "IMAGE", -32, buf, 64, 64);
            (void)gpair_fitsr_get(r, "IMAGE", -64, buf, 64, 32);
            gpair_fitsr_close(r);
        }
    }
    printf("fits_fuzz: %d files\n", argc - 1);
    return 0;
}
