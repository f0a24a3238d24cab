/*
 * gpair_io.c -- FITS / OIFITS reading and writing for the GPAIR data path
 * (declarations and semantics in include/gpair_io.h).  Plain C99; no cfitsio.
 *
 * The file is read whole; its HDUs are indexed once (header cards, data
 * offsets, binary-table column layouts), and table cells are decoded from the
 * big-endian bytes on demand.  OIFITS extraction follows
 * gpair1/trunk/getoifits.c:65-357 value for value (same float / double
 * arithmetic); the T3-leg deduplication replaces that file's linear scan over
 * all stored uv points by a hash over (wavelength, log|u|, log|v|) cells.
 */
#include "../../include/gpair_io.h"
#include "../../include/gpair_hip.h"
#include "gpair_io_internal.h"

#include <limits.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#define FITS_BLOCK 2880
#define FITS_CARD 80

static __thread char io_msg[512];

const char *gpair_io_last_error(void) { return io_msg; }

static int io_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(io_msg, sizeof(io_msg), fmt, ap);
    va_end(ap);
    return code;
}

/* ------------------------------------------------------------------------
 * FITS structure
 * ---------------------------------------------------------------------- */
typedef struct {
    char key[9];
    char val[72]; /* columns 11-80 of a card with a value indicator */
} fcard;

typedef struct {
    char name[72];
    char code;    /* TFORM type letter */
    long repeat;
    long offset;  /* byte offset in the row */
    int size;     /* bytes per element */
    double scale, zero;
} fcol;

typedef struct {
    long data, datalen;
    int ncard;
    fcard *cards;
    int bintable;
    long naxis1, naxis2;
    int ncol;
    fcol *cols;
} fhdu;

typedef struct {
    unsigned char *buf;
    long len;
    int nhdu;
    fhdu *h;
} ffile;

static void ff_close(ffile *f)
{
    for (int i = 0; i < f->nhdu; i++) {
        free(f->h[i].cards);
        free(f->h[i].cols);
    }
    free(f->h);
    free(f->buf);
    memset(f, 0, sizeof(*f));
}

static const fcard *hdu_card(const fhdu *h, const char *key)
{
    for (int i = 0; i < h->ncard; i++)
        if (strcmp(h->cards[i].key, key) == 0)
            return &h->cards[i];
    return NULL;
}

static void rstrip(char *s)
{
    size_t n = strlen(s);
    while (n > 0 && s[n - 1] == ' ')
        s[--n] = '\0';
}

/* string value: quoted, '' stands for ', trailing blanks are not significant */
static int card_string(const fcard *c, char *out, size_t cap)
{
    const char *p = c->val;
    while (*p == ' ')
        p++;
    if (*p != '\'')
        return -1;
    p++;
    size_t n = 0;
    while (*p) {
        if (*p == '\'') {
            if (p[1] == '\'') {
                p += 2;
                if (n + 1 < cap)
                    out[n++] = '\'';
                continue;
            }
            break;
        }
        if (n + 1 < cap)
            out[n++] = *p;
        p++;
    }
    out[n] = '\0';
    rstrip(out);
    return 0;
}

static int card_double(const fcard *c, double *out)
{
    char tmp[72];
    size_t n = 0;
    for (const char *p = c->val; *p && *p != '/' && n + 1 < sizeof(tmp); p++)
        tmp[n++] = (*p == 'D' || *p == 'd') ? 'E' : *p;
    tmp[n] = '\0';
    char *end;
    *out = strtod(tmp, &end);
    return end == tmp ? -1 : 0;
}

/* an integer keyword; values outside +-2^53 (or NaN) are refused, so the
 * conversion to long is always defined */
static int hdu_long(const fhdu *h, const char *key, long *out)
{
    const fcard *c = hdu_card(h, key);
    double v;
    if (!c || card_double(c, &v) || !(v >= -9007199254740992.0 && v <= 9007199254740992.0))
        return -1;
    *out = (long)v;
    return 0;
}

static int hdu_string(const fhdu *h, const char *key, char *out, size_t cap)
{
    const fcard *c = hdu_card(h, key);
    return c ? card_string(c, out, cap) : -1;
}

static int tform_size(char code, long repeat, int *size)
{
    switch (code) {
    case 'L': case 'B': case 'A': *size = 1; return 0;
    case 'I': *size = 2; return 0;
    case 'J': case 'E': *size = 4; return 0;
    case 'K': case 'D': case 'C': *size = 8; return 0;
    case 'M': *size = 16; return 0;
    case 'X': *size = 1; (void)repeat; return 0;
    default: return -1;
    }
}

static int parse_columns(fhdu *h)
{
    long tfields;
    if (hdu_long(h, "TFIELDS", &tfields) || tfields < 0 || tfields > 999)
        return io_fail(GPAIR_E_ARG, "binary table without a valid TFIELDS");
    h->ncol = (int)tfields;
    h->cols = calloc(tfields > 0 ? (size_t)tfields : 1, sizeof(fcol));
    if (!h->cols)
        return io_fail(GPAIR_E_NOMEM, "out of host memory");
    long off = 0;
    for (int i = 0; i < h->ncol; i++) {
        fcol *c = &h->cols[i];
        char key[16], form[72];
        snprintf(key, sizeof(key), "TTYPE%d", i + 1);
        if (hdu_string(h, key, c->name, sizeof(c->name)))
            c->name[0] = '\0';
        snprintf(key, sizeof(key), "TFORM%d", i + 1);
        if (hdu_string(h, key, form, sizeof(form)))
            return io_fail(GPAIR_E_ARG, "column %d without TFORM", i + 1);
        const char *p = form;
        while (*p == ' ')
            p++;
        long rep = 1;
        if (*p >= '0' && *p <= '9')
            rep = strtol(p, (char **)&p, 10);
        if (rep < 0 || rep > (1L << 30))
            return io_fail(GPAIR_E_ARG, "column %d: repeat count in TFORM '%s' out of range", i + 1, form);
        c->code = *p;
        c->repeat = rep;
        if (tform_size(c->code, rep, &c->size))
            return io_fail(GPAIR_E_ARG, "column %s: TFORM '%s' not supported", c->name, form);
        c->offset = off;
        off += c->code == 'X' ? (rep + 7) / 8 : rep * c->size;
        c->scale = 1.0;
        c->zero = 0.0;
        const fcard *sc;
        snprintf(key, sizeof(key), "TSCAL%d", i + 1);
        if ((sc = hdu_card(h, key)))
            card_double(sc, &c->scale);
        snprintf(key, sizeof(key), "TZERO%d", i + 1);
        if ((sc = hdu_card(h, key)))
            card_double(sc, &c->zero);
    }
    if (off != h->naxis1)
        return io_fail(GPAIR_E_ARG, "binary table row of %ld bytes, columns take %ld", h->naxis1, off);
    long rows_bytes;
    if (h->naxis1 == 0 && h->naxis2 > 0)
        return io_fail(GPAIR_E_ARG, "binary table of %ld rows of zero bytes", h->naxis2);
    if (__builtin_mul_overflow(h->naxis1, h->naxis2, &rows_bytes) || rows_bytes > h->datalen)
        return io_fail(GPAIR_E_ARG, "binary table of %ld rows of %ld bytes does not fit its %ld data bytes",
                       h->naxis2, h->naxis1, h->datalen);
    return GPAIR_OK;
}

static int ff_open(const char *path, ffile *f)
{
    memset(f, 0, sizeof(*f));
    FILE *fp = fopen(path, "rb");
    if (!fp)
        return io_fail(GPAIR_E_ARG, "cannot open %s", path);
    if (fseek(fp, 0, SEEK_END) != 0) {
        fclose(fp);
        return io_fail(GPAIR_E_ARG, "cannot seek %s", path);
    }
    long len = ftell(fp);
    rewind(fp);
    f->buf = malloc(len > 0 ? (size_t)len : 1);
    if (!f->buf) {
        fclose(fp);
        return io_fail(GPAIR_E_NOMEM, "out of host memory (%ld bytes)", len);
    }
    if (len > 0 && fread(f->buf, 1, (size_t)len, fp) != (size_t)len) {
        fclose(fp);
        ff_close(f);
        return io_fail(GPAIR_E_ARG, "short read of %s", path);
    }
    fclose(fp);
    f->len = len;
    long pos = 0;
    int cap = 0;
    while (pos + FITS_BLOCK <= len) {
        if (f->nhdu == cap) {
            cap = cap ? 2 * cap : 8;
            fhdu *nh = realloc(f->h, sizeof(fhdu) * (size_t)cap);
            if (!nh) {
                ff_close(f);
                return io_fail(GPAIR_E_NOMEM, "out of host memory");
            }
            f->h = nh;
        }
        fhdu *h = &f->h[f->nhdu];
        memset(h, 0, sizeof(*h));
        f->nhdu++;
        int ccap = 0, end = 0;
        while (!end) {
            if (pos + FITS_CARD > len) {
                ff_close(f);
                return io_fail(GPAIR_E_ARG, "%s: header without END", path);
            }
            const char *card = (const char *)f->buf + pos;
            pos += FITS_CARD;
            if (strncmp(card, "END", 3) == 0 && (card[3] == ' ' || FITS_CARD == 3)) {
                end = 1;
                break;
            }
            if (card[8] != '=' || card[9] != ' ')
                continue; /* COMMENT, HISTORY, blank */
            if (h->ncard == ccap) {
                ccap = ccap ? 2 * ccap : 64;
                fcard *nc = realloc(h->cards, sizeof(fcard) * (size_t)ccap);
                if (!nc) {
                    ff_close(f);
                    return io_fail(GPAIR_E_NOMEM, "out of host memory");
                }
                h->cards = nc;
            }
            fcard *c = &h->cards[h->ncard++];
            memcpy(c->key, card, 8);
            c->key[8] = '\0';
            rstrip(c->key);
            memcpy(c->val, card + 10, 70);
            c->val[70] = '\0';
        }
        pos = (pos + FITS_BLOCK - 1) / FITS_BLOCK * FITS_BLOCK;
        long bitpix = 0, naxis = 0, pcount = 0, gcount = 1;
        if (hdu_long(h, "BITPIX", &bitpix) || hdu_long(h, "NAXIS", &naxis)) {
            ff_close(f);
            return io_fail(GPAIR_E_ARG, "%s: HDU %d without BITPIX / NAXIS", path, f->nhdu);
        }
        hdu_long(h, "PCOUNT", &pcount);
        hdu_long(h, "GCOUNT", &gcount);
        if ((bitpix != 8 && bitpix != 16 && bitpix != 32 && bitpix != 64 && bitpix != -32 && bitpix != -64) ||
            naxis < 0 || naxis > 999 || pcount < 0 || gcount < 1) {
            ff_close(f);
            return io_fail(GPAIR_E_ARG, "%s: HDU %d with BITPIX %ld, NAXIS %ld, PCOUNT %ld, GCOUNT %ld", path,
                           f->nhdu, bitpix, naxis, pcount, gcount);
        }
        long n = naxis > 0 ? 1 : 0;
        int ovf = 0;
        for (int a = 1; a <= naxis; a++) {
            char key[16];
            long na;
            snprintf(key, sizeof(key), "NAXIS%d", a);
            if (hdu_long(h, key, &na) || na < 0) {
                ff_close(f);
                return io_fail(GPAIR_E_ARG, "%s: HDU %d without %s", path, f->nhdu, key);
            }
            ovf |= __builtin_mul_overflow(n, na, &n);
            if (a == 1)
                h->naxis1 = na;
            if (a == 2)
                h->naxis2 = na;
        }
        h->data = pos;
        /* data bytes = |BITPIX|/8 * GCOUNT * (PCOUNT + NAXIS1 * ... * NAXISn), every
         * step checked: the readers below index the buffer with these sizes */
        long words;
        ovf |= __builtin_add_overflow(pcount, n, &words);
        ovf |= __builtin_mul_overflow(words, gcount, &words);
        ovf |= __builtin_mul_overflow(words, labs(bitpix) / 8, &h->datalen);
        if (ovf || h->datalen > len - h->data) {
            ff_close(f);
            return io_fail(GPAIR_E_ARG, "%s: HDU %d data runs past the end of the file", path, f->nhdu);
        }
        char xt[72];
        if (f->nhdu > 1 && hdu_string(h, "XTENSION", xt, sizeof(xt)) == 0 && strcmp(xt, "BINTABLE") == 0) {
            h->bintable = 1;
            int rc = parse_columns(h);
            if (rc) {
                ff_close(f);
                return rc;
            }
        }
        pos += (h->datalen + FITS_BLOCK - 1) / FITS_BLOCK * FITS_BLOCK;
    }
    if (f->nhdu == 0) {
        ff_close(f);
        return io_fail(GPAIR_E_ARG, "%s: not a FITS file", path);
    }
    return GPAIR_OK;
}

static const fcol *find_col(const fhdu *h, const char *name)
{
    for (int i = 0; i < h->ncol; i++)
        if (strcasecmp(h->cols[i].name, name) == 0)
            return &h->cols[i];
    return NULL;
}

static uint64_t be_bits(const unsigned char *p, int n)
{
    uint64_t v = 0;
    for (int i = 0; i < n; i++)
        v = (v << 8) | p[i];
    return v;
}

/* element el of a numeric / logical cell as double */
static double cell_double(const ffile *f, const fhdu *h, const fcol *c, long row, long el)
{
    if (row < 0 || row >= h->naxis2 || el < 0 || el >= c->repeat)
        return NAN;   /* outside the table: parse_columns bounds rows x row bytes by the data */
    const unsigned char *p = f->buf + h->data + row * h->naxis1 + c->offset + el * c->size;
    double v;
    switch (c->code) {
    case 'L': return p[0] == 'T' ? 1.0 : 0.0;
    case 'B': v = (double)p[0]; break;
    case 'I': v = (double)(int16_t)be_bits(p, 2); break;
    case 'J': v = (double)(int32_t)be_bits(p, 4); break;
    case 'K': v = (double)(int64_t)be_bits(p, 8); break;
    case 'E': {
        uint32_t b = (uint32_t)be_bits(p, 4);
        float x;
        memcpy(&x, &b, 4);
        v = (double)x;
        break;
    }
    case 'D': {
        uint64_t b = be_bits(p, 8);
        memcpy(&v, &b, 8);
        break;
    }
    default: return NAN;
    }
    if (c->scale != 1.0 || c->zero != 0.0)
        v = c->scale * v + c->zero;
    return v;
}

static void cell_string(const ffile *f, const fhdu *h, const fcol *c, long row, char *out, size_t cap)
{
    size_t n = 0;
    if (row < 0 || row >= h->naxis2) {
        out[0] = '\0';
        return;
    }
    const unsigned char *p = f->buf + h->data + row * h->naxis1 + c->offset;
    while (n < (size_t)c->repeat && n + 1 < cap && p[n])
        out[n] = (char)p[n], n++;
    out[n] = '\0';
    rstrip(out);
}

static int is_ext(const fhdu *h, const char *extname)
{
    char name[72];
    return h->bintable && hdu_string(h, "EXTNAME", name, sizeof(name)) == 0 && strcmp(name, extname) == 0;
}

/* the columns a table must have; 0 when all are there */
static int need_cols(const fhdu *h, const char *ext, const char *const *names, int n, const fcol **out)
{
    for (int i = 0; i < n; i++) {
        out[i] = find_col(h, names[i]);
        if (!out[i])
            return io_fail(GPAIR_E_ARG, "%s table without column %s", ext, names[i]);
    }
    return GPAIR_OK;
}

/* ------------------------------------------------------------------------
 * growable arrays
 * ---------------------------------------------------------------------- */
static int grow(void **p, size_t elem, long *cap, long need)
{
    if (need <= *cap)
        return 0;
    long nc = *cap ? *cap : 256;
    while (nc < need)
        nc *= 2;
    void *np = realloc(*p, elem * (size_t)nc);
    if (!np)
        return -1;
    *p = np;
    *cap = nc;
    return 0;
}

/* ------------------------------------------------------------------------
 * uv matching (compare_uv, gpair1/trunk/getoifits.c:17-62, same arithmetic:
 * float differences and sums, the ratio formed in double, stored as float)
 * ---------------------------------------------------------------------- */
#define UV_THRESHOLD 5.0e-5f

static int uv_match(float u, float v, float su, float sv)
{
    const float t2 = UV_THRESHOLD * UV_THRESHOLD;
    float pu = (float)(2.0 * (double)(u - su) * (double)(u - su) / (double)(u * u + su * su));
    float pv = (float)(2.0 * (double)(v - sv) * (double)(v - sv) / (double)(v * v + sv * sv));
    float mu = (float)(2.0 * (double)(u + su) * (double)(u + su) / (double)(u * u + su * su));
    float mv = (float)(2.0 * (double)(v + sv) * (double)(v + sv) / (double)(v * v + sv * sv));
    if (u - su == 0.0f)
        pu = 0.0f;
    if (v - sv == 0.0f)
        pv = 0.0f;
    if (u + su == 0.0f)
        mu = 0.0f;
    if (v + sv == 0.0f)
        mv = 0.0f;
    if (pu < t2 && pv < t2)
        return 1;
    if (mu < t2 && mv < t2)
        return -1;
    return 0;
}

/* Both tests bound ||x| - |x'|| by UV_THRESHOLD * max(|x|, |x'|) for x = u and
 * x = v, so a match lies in a cell of log2|x| (4096 per octave) within
 * |x| (1 -+ 1e-4) of the query.  Zero matches only zero; non-finite values
 * match nothing. */
#define CELL_ZERO (-2147483647 - 1)
#define CELL_NONFINITE (-2147483647)
#define CELLS_PER_OCTAVE 4096.0

static int32_t uv_cell(double a)
{
    if (a == 0.0)
        return CELL_ZERO;
    if (!isfinite(a))
        return CELL_NONFINITE;
    return (int32_t)floor(log2(a) * CELLS_PER_OCTAVE);
}

static void cell_range(float x, int32_t *lo, int32_t *hi)
{
    const double a = fabs((double)x);
    if (a == 0.0 || !isfinite(a)) {
        *lo = *hi = uv_cell(a);
        return;
    }
    *lo = uv_cell(a * (1.0 - 1e-4));
    *hi = uv_cell(a * (1.0 + 1e-4));
}

typedef struct {
    uint32_t wl;
    int32_t cu, cv;
} uvkey;

typedef struct {
    int *head;    /* [cap] most recent point of each slot, -1 when empty */
    int *next;    /* [points] chain */
    uvkey *key;   /* [points] */
    long cap, npt, ptcap;
} uvindex;

static uint32_t key_hash(uint32_t wl, int32_t cu, int32_t cv)
{
    uint64_t h = (uint64_t)wl * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)(uint32_t)cu * 0xC2B2AE3D27D4EB4Full + (h >> 29);
    h ^= (uint64_t)(uint32_t)cv * 0x165667B19E3779F9ull + (h >> 31);
    h ^= h >> 33;
    return (uint32_t)h;
}

static uint32_t float_bits(float x)
{
    uint32_t b;
    memcpy(&b, &x, 4);
    return b;
}

static void uvi_free(uvindex *x)
{
    free(x->head);
    free(x->next);
    free(x->key);
    memset(x, 0, sizeof(*x));
}

static int uvi_rehash(uvindex *x, long cap)
{
    int *h = malloc(sizeof(int) * (size_t)cap);
    if (!h)
        return -1;
    for (long i = 0; i < cap; i++)
        h[i] = -1;
    for (long p = 0; p < x->npt; p++) {
        const uint32_t s = key_hash(x->key[p].wl, x->key[p].cu, x->key[p].cv) & (uint32_t)(cap - 1);
        x->next[p] = h[s];
        h[s] = (int)p;
    }
    free(x->head);
    x->head = h;
    x->cap = cap;
    return 0;
}

/* point p (== x->npt) with stored coordinates (u, v) and wavelength wl */
static int uvi_add(uvindex *x, float u, float v, float wl)
{
    if (x->npt + 1 > x->cap / 2 && uvi_rehash(x, x->cap ? 2 * x->cap : 1024))
        return -1;
    if (x->npt + 1 > x->ptcap) {
        long kc = x->ptcap;
        if (grow((void **)&x->next, sizeof(int), &x->ptcap, x->npt + 1) ||
            grow((void **)&x->key, sizeof(uvkey), &kc, x->npt + 1))
            return -1;
    }
    const long p = x->npt++;
    x->key[p].wl = float_bits(wl);
    x->key[p].cu = uv_cell(fabs((double)u));
    x->key[p].cv = uv_cell(fabs((double)v));
    const uint32_t s = key_hash(x->key[p].wl, x->key[p].cu, x->key[p].cv) & (uint32_t)(x->cap - 1);
    x->next[p] = x->head[s];
    x->head[s] = (int)p;
    return 0;
}

/* first stored point (lowest index) matching (u, v) at wavelength wl, with its
 * sign in *sgn; -1 when none */
static long uvi_find(const uvindex *x, const float *su, const float *sv, const float *swl, float u, float v, float wl,
                     int *sgn, long long *compares)
{
    if (x->npt == 0)
        return -1;
    int32_t ulo, uhi, vlo, vhi;
    cell_range(u, &ulo, &uhi);
    cell_range(v, &vlo, &vhi);
    const uint32_t w = float_bits(wl);
    long best = -1;
    int bs = 0;
    for (int32_t cu = ulo; cu <= uhi; cu++) {
        for (int32_t cv = vlo; cv <= vhi; cv++) {
            const uint32_t s = key_hash(w, cu, cv) & (uint32_t)(x->cap - 1);
            for (int p = x->head[s]; p >= 0; p = x->next[p]) {
                if (x->key[p].wl != w || x->key[p].cu != cu || x->key[p].cv != cv)
                    continue;
                if (best >= 0 && p > best)
                    continue;
                (*compares)++;
                const int m = uv_match(u, v, su[p], sv[p]);
                if (m != 0 && wl == swl[p]) {
                    best = p;
                    bs = m;
                }
            }
            if (cv == CELL_ZERO || cv == CELL_NONFINITE)
                break;
        }
        if (cu == CELL_ZERO || cu == CELL_NONFINITE)
            break;
    }
    *sgn = bs;
    return best;
}

/* ------------------------------------------------------------------------
 * OIFITS
 * ---------------------------------------------------------------------- */
typedef struct {
    long cap_pow, cap_bis, cap_uv;
    uvindex idx;
} oi_build;

void gpair_oidata_free(gpair_oidata *d)
{
    if (!d)
        return;
    free(d->pow); free(d->powerr);
    free(d->bisamp); free(d->bisamperr); free(d->bisphs); free(d->bisphserr);
    free(d->u); free(d->v); free(d->wavelength); free(d->bandwidth);
    free(d->uvpnt); free(d->sign);
    memset(d, 0, sizeof(*d));
}

static int read_targets(const ffile *f, int *ids, char (*names)[17], int max, int *count)
{
    for (int i = 1; i < f->nhdu; i++) {
        const fhdu *h = &f->h[i];
        if (!is_ext(h, "OI_TARGET"))
            continue;
        const fcol *cid = find_col(h, "TARGET_ID"), *cname = find_col(h, "TARGET");
        if (!cid || !cname)
            return io_fail(GPAIR_E_ARG, "OI_TARGET without TARGET_ID / TARGET");
        *count = h->naxis2 > INT_MAX ? INT_MAX : (int)h->naxis2;
        for (long r = 0; r < h->naxis2 && r < max; r++) {
            if (ids)
                ids[r] = (int)cell_double(f, h, cid, r, 0);
            if (names)
                cell_string(f, h, cname, r, names[r], 17);
        }
        return GPAIR_OK;
    }
    return io_fail(GPAIR_E_ARG, "no OI_TARGET table");
}

int gpair_oifits_targets(const char *path, int *ids, char *names, int max, int *count)
{
    if (!path || !count || max < 0)
        return io_fail(GPAIR_E_ARG, "gpair_oifits_targets: NULL argument");
    ffile f;
    int rc = ff_open(path, &f);
    if (rc)
        return rc;
    rc = read_targets(&f, ids, (char(*)[17])names, max, count);
    ff_close(&f);
    return rc;
}

/* OI_WAVELENGTH table of an instrument: the first extension with that INSNAME
 * (read_oi_wavelength, gpair1/trunk/read_fits.c:228-283: values as float) */
static int find_wave(const ffile *f, const char *insname, float **wave, float **band, long *nwave)
{
    for (int i = 1; i < f->nhdu; i++) {
        const fhdu *h = &f->h[i];
        char ins[72];
        if (!is_ext(h, "OI_WAVELENGTH") || hdu_string(h, "INSNAME", ins, sizeof(ins)) || strcmp(ins, insname) != 0)
            continue;
        const fcol *cw = find_col(h, "EFF_WAVE"), *cb = find_col(h, "EFF_BAND");
        if (!cw || !cb)
            return io_fail(GPAIR_E_ARG, "OI_WAVELENGTH %s without EFF_WAVE / EFF_BAND", insname);
        *nwave = h->naxis2;
        *wave = malloc(sizeof(float) * (size_t)(h->naxis2 > 0 ? h->naxis2 : 1));
        *band = malloc(sizeof(float) * (size_t)(h->naxis2 > 0 ? h->naxis2 : 1));
        if (!*wave || !*band) {
            free(*wave);
            free(*band);
            return io_fail(GPAIR_E_NOMEM, "out of host memory");
        }
        for (long r = 0; r < h->naxis2; r++) {
            (*wave)[r] = (float)cell_double(f, h, cw, r, 0);
            (*band)[r] = (float)cell_double(f, h, cb, r, 0);
        }
        return GPAIR_OK;
    }
    return io_fail(GPAIR_E_ARG, "no OI_WAVELENGTH table for INSNAME '%s'", insname);
}

static int in_band(float wave, float minband, float maxband)
{
    const double nm = (double)wave * 1.0e9;
    return nm > (double)minband && nm < (double)maxband;
}

/* a new uv point; flipped into the +u half plane with sign -1 when u < 0 */
static int add_uv(gpair_oidata *d, oi_build *b, float u, float v, float wl, float bw, int *sgn)
{
    if (d->nuv + 1 > b->cap_uv) {
        long c1 = b->cap_uv, c2 = b->cap_uv, c3 = b->cap_uv;
        if (grow((void **)&d->u, sizeof(float), &b->cap_uv, d->nuv + 1) ||
            grow((void **)&d->v, sizeof(float), &c1, d->nuv + 1) ||
            grow((void **)&d->wavelength, sizeof(float), &c2, d->nuv + 1) ||
            grow((void **)&d->bandwidth, sizeof(float), &c3, d->nuv + 1))
            return -1;
    }
    *sgn = 1;
    if (u < 0.0f) {
        u = -u;
        v = -v;
        *sgn = -1;
    }
    d->u[d->nuv] = u;
    d->v[d->nuv] = v;
    d->wavelength[d->nuv] = wl;
    d->bandwidth[d->nuv] = bw;
    if (uvi_add(&b->idx, u, v, wl))
        return -1;
    d->nuv++;
    return 0;
}

static int read_vis2_tables(const ffile *f, gpair_oidata *d, oi_build *b)
{
    static const char *const cn[] = {"TARGET_ID", "VIS2DATA", "VIS2ERR", "UCOORD", "VCOORD", "FLAG"};
    for (int i = 1; i < f->nhdu; i++) {
        const fhdu *h = &f->h[i];
        if (!is_ext(h, "OI_VIS2") || h->naxis2 < 1)
            continue;
        const fcol *c[6];
        char ins[72];
        int rc;
        if ((rc = need_cols(h, "OI_VIS2", cn, 6, c)))
            return rc;
        if (hdu_string(h, "INSNAME", ins, sizeof(ins)))
            return io_fail(GPAIR_E_ARG, "OI_VIS2 table (HDU %d) without INSNAME", i + 1);
        if ((int)cell_double(f, h, c[0], 0, 0) != d->target_id)
            continue; /* the table goes by its first record (getoifits.c:129) */
        float *wave, *band;
        long nwave;
        if ((rc = find_wave(f, ins, &wave, &band, &nwave)))
            return rc;
        const long nw = c[1]->repeat;
        if (nw > nwave || c[2]->repeat < nw || c[5]->repeat < nw) {
            free(wave);
            free(band);
            return io_fail(GPAIR_E_ARG, "OI_VIS2 (HDU %d): %ld channels, OI_WAVELENGTH %s has %ld", i + 1, nw, ins,
                           nwave);
        }
        for (long r = 0; r < h->naxis2; r++) {
            const double uc = cell_double(f, h, c[3], r, 0), vc = cell_double(f, h, c[4], r, 0);
            for (long k = 0; k < nw; k++) {
                if (!in_band(wave[k], d->minband, d->maxband) || cell_double(f, h, c[5], r, k) != 0.0)
                    continue;
                long cap2 = b->cap_pow;
                if (grow((void **)&d->pow, sizeof(float), &b->cap_pow, d->npow + 1) ||
                    grow((void **)&d->powerr, sizeof(float), &cap2, d->npow + 1)) {
                    free(wave);
                    free(band);
                    return io_fail(GPAIR_E_NOMEM, "out of host memory");
                }
                d->pow[d->npow] = (float)cell_double(f, h, c[1], r, k);
                d->powerr[d->npow] = (float)cell_double(f, h, c[2], r, k);
                int sgn;
                if (add_uv(d, b, (float)(uc / (double)wave[k]), (float)(vc / (double)wave[k]), wave[k], band[k],
                           &sgn)) {
                    free(wave);
                    free(band);
                    return io_fail(GPAIR_E_NOMEM, "out of host memory");
                }
                d->npow++;
            }
        }
        free(wave);
        free(band);
    }
    return GPAIR_OK;
}

static int read_t3_tables(const ffile *f, gpair_oidata *d, oi_build *b)
{
    static const char *const cn[] = {"TARGET_ID", "T3AMP", "T3AMPERR", "T3PHI", "T3PHIERR",
                                     "U1COORD",   "V1COORD", "U2COORD", "V2COORD", "FLAG"};
    for (int i = 1; i < f->nhdu; i++) {
        const fhdu *h = &f->h[i];
        if (!is_ext(h, "OI_T3") || h->naxis2 < 1)
            continue;
        const fcol *c[10];
        char ins[72];
        int rc;
        if ((rc = need_cols(h, "OI_T3", cn, 10, c)))
            return rc;
        if (hdu_string(h, "INSNAME", ins, sizeof(ins)))
            return io_fail(GPAIR_E_ARG, "OI_T3 table (HDU %d) without INSNAME", i + 1);
        if ((int)cell_double(f, h, c[0], 0, 0) != d->target_id)
            continue;
        float *wave, *band;
        long nwave;
        if ((rc = find_wave(f, ins, &wave, &band, &nwave)))
            return rc;
        const long nw = c[1]->repeat;
        if (nw > nwave || c[2]->repeat < nw || c[3]->repeat < nw || c[4]->repeat < nw || c[9]->repeat < nw) {
            free(wave);
            free(band);
            return io_fail(GPAIR_E_ARG, "OI_T3 (HDU %d): %ld channels, OI_WAVELENGTH %s has %ld", i + 1, nw, ins,
                           nwave);
        }
        for (long r = 0; r < h->naxis2; r++) {
            const double u1 = cell_double(f, h, c[5], r, 0), v1 = cell_double(f, h, c[6], r, 0);
            const double u2 = cell_double(f, h, c[7], r, 0), v2 = cell_double(f, h, c[8], r, 0);
            for (long k = 0; k < nw; k++) {
                if (!in_band(wave[k], d->minband, d->maxband) || cell_double(f, h, c[9], r, k) != 0.0)
                    continue;
                const long n = d->nbis + 1;
                long cc[6] = {b->cap_bis, b->cap_bis, b->cap_bis, b->cap_bis, b->cap_bis, b->cap_bis};
                if (grow((void **)&d->bisamp, sizeof(float), &cc[0], n) ||
                    grow((void **)&d->bisamperr, sizeof(float), &cc[1], n) ||
                    grow((void **)&d->bisphs, sizeof(float), &cc[2], n) ||
                    grow((void **)&d->bisphserr, sizeof(float), &cc[3], n) ||
                    grow((void **)&d->uvpnt, 3 * sizeof(int64_t), &cc[4], n) ||
                    grow((void **)&d->sign, 3 * sizeof(int16_t), &cc[5], n)) {
                    free(wave);
                    free(band);
                    return io_fail(GPAIR_E_NOMEM, "out of host memory");
                }
                b->cap_bis = cc[0];
                const double amp = cell_double(f, h, c[1], r, k);
                const long t = d->nbis;
                if (isnan(amp)) {
                    d->bisamp[t] = 1.0f;
                    d->bisamperr[t] = (float)1e99;
                } else {
                    d->bisamp[t] = (float)amp;
                    d->bisamperr[t] = (float)cell_double(f, h, c[2], r, k);
                }
                d->bisphs[t] = (float)cell_double(f, h, c[3], r, k);
                d->bisphserr[t] = (float)cell_double(f, h, c[4], r, k);
                float lu[3], lv[3];
                lu[0] = (float)(u1 / (double)wave[k]);
                lv[0] = (float)(v1 / (double)wave[k]);
                lu[1] = (float)(u2 / (double)wave[k]);
                lv[1] = (float)(v2 / (double)wave[k]);
                lu[2] = -(lu[0] + lu[1]);
                lv[2] = -(lv[0] + lv[1]);
                for (int j = 0; j < 3; j++) {
                    int sgn;
                    long p = uvi_find(&b->idx, d->u, d->v, d->wavelength, lu[j], lv[j], wave[k], &sgn,
                                      &d->uv_compares);
                    if (p < 0) {
                        p = d->nuv;
                        if (add_uv(d, b, lu[j], lv[j], wave[k], band[k], &sgn)) {
                            free(wave);
                            free(band);
                            return io_fail(GPAIR_E_NOMEM, "out of host memory");
                        }
                    }
                    d->uvpnt[3 * t + j] = p;
                    d->sign[3 * t + j] = (int16_t)sgn;
                }
                d->nbis++;
            }
        }
        free(wave);
        free(band);
    }
    return GPAIR_OK;
}

int gpair_oifits_read(const char *path, int target_id, float minband, float maxband, gpair_oidata *out)
{
    if (!path || !out)
        return io_fail(GPAIR_E_ARG, "gpair_oifits_read: NULL argument");
    memset(out, 0, sizeof(*out));
    ffile f;
    int rc = ff_open(path, &f);
    if (rc)
        return rc;
    int ids[64], count = 0;
    char names[64][17];
    if ((rc = read_targets(&f, ids, names, 64, &count))) {
        ff_close(&f);
        return rc;
    }
    if (target_id < 0) {
        if (count != 1) {
            char list[256] = "";
            for (int i = 0; i < count && i < 64; i++) {
                char one[40];
                snprintf(one, sizeof(one), "%s%d (%s)", i ? ", " : "", ids[i], names[i]);
                strncat(list, one, sizeof(list) - strlen(list) - 1);
            }
            ff_close(&f);
            return io_fail(GPAIR_E_ARG, "%s has %d targets, choose one: %s", path, count, list);
        }
        target_id = ids[0];
    } else {
        int known = 0;
        for (int i = 0; i < count && i < 64; i++)
            known |= ids[i] == target_id;
        if (!known) {
            ff_close(&f);
            return io_fail(GPAIR_E_ARG, "%s: target id %d is not in OI_TARGET", path, target_id);
        }
    }
    if (minband < 0.0f || maxband <= minband) {
        minband = 1.0f;
        maxband = 50000.0f;
    }
    out->target_id = target_id;
    out->minband = minband;
    out->maxband = maxband;
    oi_build b;
    memset(&b, 0, sizeof(b));
    rc = read_vis2_tables(&f, out, &b);
    if (!rc)
        rc = read_t3_tables(&f, out, &b);
    uvi_free(&b.idx);
    ff_close(&f);
    if (!rc && out->npow == 0 && out->nbis == 0)
        rc = io_fail(GPAIR_E_ARG, "%s: no data of target %d between %g and %g nm", path, target_id, (double)minband,
                     (double)maxband);
    if (rc)
        gpair_oidata_free(out);
    return rc;
}

/* ------------------------------------------------------------------------
 * writing
 * ---------------------------------------------------------------------- */
typedef struct {
    unsigned char *p;
    long n, cap;
} obuf;

static int ob_put(obuf *b, const void *src, long n)
{
    if (grow((void **)&b->p, 1, &b->cap, b->n + n))
        return -1;
    memcpy(b->p + b->n, src, (size_t)n);
    b->n += n;
    return 0;
}

static int ob_pad(obuf *b, unsigned char fill)
{
    long pad = (FITS_BLOCK - b->n % FITS_BLOCK) % FITS_BLOCK;
    if (grow((void **)&b->p, 1, &b->cap, b->n + pad))
        return -1;
    memset(b->p + b->n, fill, (size_t)pad);
    b->n += pad;
    return 0;
}

static int ob_card(obuf *b, const char *text)
{
    char c[FITS_CARD];
    size_t n = strlen(text);
    memset(c, ' ', FITS_CARD);
    memcpy(c, text, n < FITS_CARD ? n : FITS_CARD);
    return ob_put(b, c, FITS_CARD);
}

static int card_int(obuf *b, const char *key, long v, const char *comment)
{
    char t[96];
    snprintf(t, sizeof(t), "%-8.8s= %20ld%s%s", key, v, comment ? " / " : "", comment ? comment : "");
    return ob_card(b, t);
}

static int card_log(obuf *b, const char *key, int v, const char *comment)
{
    char t[96];
    snprintf(t, sizeof(t), "%-8.8s= %20s%s%s", key, v ? "T" : "F", comment ? " / " : "", comment ? comment : "");
    return ob_card(b, t);
}

static int card_dbl(obuf *b, const char *key, double v, const char *comment)
{
    char t[96];
    snprintf(t, sizeof(t), "%-8.8s= %20.12E%s%s", key, v, comment ? " / " : "", comment ? comment : "");
    return ob_card(b, t);
}

static int card_str(obuf *b, const char *key, const char *v, const char *comment)
{
    char q[72];
    size_t n = 0;
    for (const char *p = v; *p && n + 2 < 68; p++) {
        q[n++] = *p;
        if (*p == '\'')
            q[n++] = '\'';
    }
    while (n < 8)
        q[n++] = ' ';
    q[n] = '\0';
    char t[128];
    snprintf(t, sizeof(t), "%-8.8s= '%s'%s%s", key, q, comment ? " / " : "", comment ? comment : "");
    return ob_card(b, t);
}

static void put_be(unsigned char *dst, uint64_t v, int n)
{
    for (int i = n - 1; i >= 0; i--) {
        dst[i] = (unsigned char)(v & 0xff);
        v >>= 8;
    }
}

static void put_d(unsigned char **p, double x)
{
    uint64_t b;
    memcpy(&b, &x, 8);
    put_be(*p, b, 8);
    *p += 8;
}

static void put_e(unsigned char **p, float x)
{
    put_be(*p, float_bits(x), 4);
    *p += 4;
}

static void put_i(unsigned char **p, int v)
{
    put_be(*p, (uint64_t)(uint16_t)(int16_t)v, 2);
    *p += 2;
}

static void put_a(unsigned char **p, const char *s, int n)
{
    int k = 0;
    for (; k < n && s[k]; k++)
        (*p)[k] = (unsigned char)s[k];
    for (; k < n; k++)
        (*p)[k] = ' ';
    *p += n;
}

typedef struct {
    const char *name, *form, *unit;
} colspec;

static int table_header(obuf *b, const char *ext, long rowbytes, long nrows, const colspec *cols, int ncol,
                        const char *insname, int dated)
{
    int rc = card_str(b, "XTENSION", "BINTABLE", "binary table extension") || card_int(b, "BITPIX", 8, NULL) ||
             card_int(b, "NAXIS", 2, NULL) || card_int(b, "NAXIS1", rowbytes, "bytes per row") ||
             card_int(b, "NAXIS2", nrows, "rows") || card_int(b, "PCOUNT", 0, NULL) ||
             card_int(b, "GCOUNT", 1, NULL) || card_int(b, "TFIELDS", ncol, NULL);
    for (int i = 0; i < ncol && !rc; i++) {
        char k[16];
        snprintf(k, sizeof(k), "TTYPE%d", i + 1);
        rc = card_str(b, k, cols[i].name, NULL);
        snprintf(k, sizeof(k), "TFORM%d", i + 1);
        rc = rc || card_str(b, k, cols[i].form, NULL);
        if (!rc && cols[i].unit) {
            snprintf(k, sizeof(k), "TUNIT%d", i + 1);
            rc = card_str(b, k, cols[i].unit, NULL);
        }
    }
    rc = rc || card_str(b, "EXTNAME", ext, NULL) || card_int(b, "OI_REVN", 1, "OIFITS revision");
    if (!rc && dated)
        rc = card_str(b, "DATE-OBS", "2000-01-01", NULL);
    if (!rc && insname)
        rc = card_str(b, "INSNAME", insname, NULL);
    rc = rc || ob_card(b, "END") || ob_pad(b, ' ');
    return rc;
}

int gpair_oifits_write(const char *path, const char *target, int nwave, const float *eff_wave,
                       const float *eff_band, int nvis2, const double *v2_u, const double *v2_v,
                       const double *vis2data, const double *vis2err, const unsigned char *v2_flag, int nt3,
                       const double *t3_u1, const double *t3_v1, const double *t3_u2, const double *t3_v2,
                       const double *t3amp, const double *t3amperr, const double *t3phi, const double *t3phierr,
                       const unsigned char *t3_flag)
{
    if (!path || nwave < 1 || nwave > 9999 || !eff_wave || !eff_band || nvis2 < 0 || nt3 < 0 ||
        (nvis2 > 0 && (!v2_u || !v2_v || !vis2data || !vis2err)) ||
        (nt3 > 0 && (!t3_u1 || !t3_v1 || !t3_u2 || !t3_v2 || !t3amp || !t3amperr || !t3phi || !t3phierr)))
        return io_fail(GPAIR_E_ARG, "gpair_oifits_write: bad arguments");
    const char *ins = "GPAIR_SYNTH";
    obuf b = {0};
    int rc = 0;
    /* primary HDU */
    rc = card_log(&b, "SIMPLE", 1, "FITS standard") || card_int(&b, "BITPIX", 8, NULL) ||
         card_int(&b, "NAXIS", 0, NULL) || card_log(&b, "EXTEND", 1, NULL) || ob_card(&b, "END") || ob_pad(&b, ' ');
    /* OI_TARGET, one row (134 bytes) */
    static const colspec tcols[] = {
        {"TARGET_ID", "1I", NULL}, {"TARGET", "16A", NULL},   {"RAEP0", "1D", "deg"},    {"DECEP0", "1D", "deg"},
        {"EQUINOX", "1E", "year"}, {"RA_ERR", "1D", "deg"},   {"DEC_ERR", "1D", "deg"},  {"SYSVEL", "1D", "m/s"},
        {"VELTYP", "8A", NULL},    {"VELDEF", "8A", NULL},    {"PMRA", "1D", "deg/year"}, {"PMDEC", "1D", "deg/year"},
        {"PMRA_ERR", "1D", "deg/year"}, {"PMDEC_ERR", "1D", "deg/year"}, {"PARALLAX", "1E", "deg"},
        {"PARA_ERR", "1E", "deg"}, {"SPECTYP", "16A", NULL}};
    rc = rc || table_header(&b, "OI_TARGET", 134, 1, tcols, 17, NULL, 0);
    if (!rc) {
        unsigned char row[134], *p = row;
        put_i(&p, 1);
        put_a(&p, target ? target : "SYNTH", 16);
        put_d(&p, 0.0); put_d(&p, 0.0); put_e(&p, 2000.0f); put_d(&p, 0.0); put_d(&p, 0.0); put_d(&p, 0.0);
        put_a(&p, "UNKNOWN", 8); put_a(&p, "OPTICAL", 8);
        put_d(&p, 0.0); put_d(&p, 0.0); put_d(&p, 0.0); put_d(&p, 0.0); put_e(&p, 0.0f); put_e(&p, 0.0f);
        put_a(&p, "UNKNOWN", 16);
        rc = ob_put(&b, row, 134) || ob_pad(&b, 0);
    }
    /* OI_WAVELENGTH */
    static const colspec wcols[] = {{"EFF_WAVE", "1E", "m"}, {"EFF_BAND", "1E", "m"}};
    rc = rc || table_header(&b, "OI_WAVELENGTH", 8, nwave, wcols, 2, ins, 0);
    for (int k = 0; k < nwave && !rc; k++) {
        unsigned char row[8], *p = row;
        put_e(&p, eff_wave[k]);
        put_e(&p, eff_band[k]);
        rc = ob_put(&b, row, 8);
    }
    rc = rc || ob_pad(&b, 0);
    char fd[16], fl[16];
    snprintf(fd, sizeof(fd), "%dD", nwave);
    snprintf(fl, sizeof(fl), "%dL", nwave);
    unsigned char *row = NULL;
    if (!rc && nvis2 > 0) {
        const colspec vcols[] = {{"TARGET_ID", "1I", NULL}, {"TIME", "1D", "s"}, {"MJD", "1D", "day"},
                                 {"INT_TIME", "1D", "s"},   {"VIS2DATA", fd, NULL}, {"VIS2ERR", fd, NULL},
                                 {"UCOORD", "1D", "m"},     {"VCOORD", "1D", "m"},  {"STA_INDEX", "2I", NULL},
                                 {"FLAG", fl, NULL}};
        const long rb = 2 + 8 * 3 + 16L * nwave + 16 + 4 + nwave;
        rc = table_header(&b, "OI_VIS2", rb, nvis2, vcols, 10, ins, 1);
        row = malloc((size_t)rb);
        if (!row)
            rc = 1;
        for (int r = 0; r < nvis2 && !rc; r++) {
            unsigned char *p = row;
            put_i(&p, 1);
            put_d(&p, 0.0); put_d(&p, 51544.0); put_d(&p, 1.0);
            for (int k = 0; k < nwave; k++)
                put_d(&p, vis2data[(size_t)r * nwave + k]);
            for (int k = 0; k < nwave; k++)
                put_d(&p, vis2err[(size_t)r * nwave + k]);
            put_d(&p, v2_u[r]);
            put_d(&p, v2_v[r]);
            put_i(&p, 1);
            put_i(&p, 2);
            for (int k = 0; k < nwave; k++)
                *p++ = (v2_flag && v2_flag[(size_t)r * nwave + k]) ? 'T' : 'F';
            rc = ob_put(&b, row, rb);
        }
        free(row);
        row = NULL;
        rc = rc || ob_pad(&b, 0);
    }
    if (!rc && nt3 > 0) {
        const colspec ccols[] = {{"TARGET_ID", "1I", NULL}, {"TIME", "1D", "s"},    {"MJD", "1D", "day"},
                                 {"INT_TIME", "1D", "s"},   {"T3AMP", fd, NULL},    {"T3AMPERR", fd, NULL},
                                 {"T3PHI", fd, "deg"},      {"T3PHIERR", fd, "deg"}, {"U1COORD", "1D", "m"},
                                 {"V1COORD", "1D", "m"},    {"U2COORD", "1D", "m"}, {"V2COORD", "1D", "m"},
                                 {"STA_INDEX", "3I", NULL}, {"FLAG", fl, NULL}};
        const long rb = 2 + 8 * 3 + 32L * nwave + 32 + 6 + nwave;
        rc = table_header(&b, "OI_T3", rb, nt3, ccols, 14, ins, 1);
        row = malloc((size_t)rb);
        if (!row)
            rc = 1;
        for (int r = 0; r < nt3 && !rc; r++) {
            unsigned char *p = row;
            put_i(&p, 1);
            put_d(&p, 0.0); put_d(&p, 51544.0); put_d(&p, 1.0);
            const double *arr[4] = {t3amp, t3amperr, t3phi, t3phierr};
            for (int a = 0; a < 4; a++)
                for (int k = 0; k < nwave; k++)
                    put_d(&p, arr[a][(size_t)r * nwave + k]);
            put_d(&p, t3_u1[r]); put_d(&p, t3_v1[r]); put_d(&p, t3_u2[r]); put_d(&p, t3_v2[r]);
            put_i(&p, 1); put_i(&p, 2); put_i(&p, 3);
            for (int k = 0; k < nwave; k++)
                *p++ = (t3_flag && t3_flag[(size_t)r * nwave + k]) ? 'T' : 'F';
            rc = ob_put(&b, row, rb);
        }
        free(row);
        rc = rc || ob_pad(&b, 0);
    }
    if (rc) {
        free(b.p);
        return io_fail(GPAIR_E_NOMEM, "gpair_oifits_write: out of host memory");
    }
    FILE *fp = fopen(path, "wb");
    if (!fp) {
        free(b.p);
        return io_fail(GPAIR_E_ARG, "cannot create %s", path);
    }
    const int ok = fwrite(b.p, 1, (size_t)b.n, fp) == (size_t)b.n;
    free(b.p);
    if (fclose(fp) != 0 || !ok)
        return io_fail(GPAIR_E_ARG, "short write of %s", path);
    return GPAIR_OK;
}

/* ------------------------------------------------------------------------
 * images
 * ---------------------------------------------------------------------- */
int gpair_fits_read_image(const char *path, float *out, long capacity, int *width, int *height)
{
    if (!path || !width || !height)
        return io_fail(GPAIR_E_ARG, "gpair_fits_read_image: NULL argument");
    ffile f;
    int rc = ff_open(path, &f);
    if (rc)
        return rc;
    const fhdu *h = &f.h[0];
    long bitpix = 0, naxis = 0;
    hdu_long(h, "BITPIX", &bitpix);
    hdu_long(h, "NAXIS", &naxis);
    if (naxis < 2 || h->naxis1 < 1 || h->naxis2 < 1) {
        ff_close(&f);
        return io_fail(GPAIR_E_ARG, "%s: the primary HDU holds no 2-D image", path);
    }
    const long n = h->naxis1 * h->naxis2;   /* no overflow: bounded by the data size (ff_open) */
    if (h->naxis1 > INT_MAX || h->naxis2 > INT_MAX || n > h->datalen / (labs(bitpix) / 8)) {
        const long n1 = h->naxis1, n2 = h->naxis2;
        ff_close(&f);
        return io_fail(GPAIR_E_ARG, "%s: %ld x %ld pixels do not fit the primary HDU's data", path, n1, n2);
    }
    *width = (int)h->naxis1;
    *height = (int)h->naxis2;
    if (out) {
        if (capacity < n) {
            ff_close(&f);
            return io_fail(GPAIR_E_ARG, "%s: %d x %d pixels, room for %ld", path, *width, *height, capacity);
        }
        double bscale = 1.0, bzero = 0.0;
        const fcard *c;
        if ((c = hdu_card(h, "BSCALE")))
            card_double(c, &bscale);
        if ((c = hdu_card(h, "BZERO")))
            card_double(c, &bzero);
        const int nb = (int)(labs(bitpix) / 8);
        const unsigned char *p = f.buf + h->data;
        for (long i = 0; i < n; i++, p += nb) {
            double v;
            const uint64_t bits = be_bits(p, nb);
            switch (bitpix) {
            case 8: v = (double)p[0]; break;
            case 16: v = (double)(int16_t)bits; break;
            case 32: v = (double)(int32_t)bits; break;
            case 64: v = (double)(int64_t)bits; break;
            case -32: {
                float x;
                uint32_t b32 = (uint32_t)bits;
                memcpy(&x, &b32, 4);
                v = (double)x;
                break;
            }
            case -64: memcpy(&v, &bits, 8); break;
            default:
                ff_close(&f);
                return io_fail(GPAIR_E_ARG, "%s: BITPIX %ld not supported", path, bitpix);
            }
            if (bscale != 1.0 || bzero != 0.0)
                v = bscale * v + bzero;
            out[i] = (float)v;
        }
    }
    ff_close(&f);
    return GPAIR_OK;
}

int gpair_fits_write_image(const char *path, const float *image, int width, int height, int bitpix,
                           float pixsize, const char *datafile, const char *target)
{
    if (!path || !image || width < 1 || height < 1)
        return io_fail(GPAIR_E_ARG, "gpair_fits_write_image: bad arguments");
    if (bitpix != -32)
        bitpix = -64;
    obuf b = {0};
    int rc = card_log(&b, "SIMPLE", 1, "file does conform to FITS standard") ||
             card_int(&b, "BITPIX", bitpix, "number of bits per data pixel") ||
             card_int(&b, "NAXIS", 2, "number of data axes") || card_int(&b, "NAXIS1", width, "length of data axis 1") ||
             card_int(&b, "NAXIS2", height, "length of data axis 2") ||
             card_log(&b, "EXTEND", 1, "FITS dataset may contain extensions") ||
             card_str(&b, "DATAFILE", datafile ? datafile : "dummy", "Data File Name") ||
             card_str(&b, "TARGET", target ? target : "dummy", "Target Name") ||
             card_dbl(&b, "PIXSIZE", (double)pixsize, "Pixelation (mas)") ||
             card_int(&b, "WIDTH", width, "Size (pixels)") || ob_card(&b, "END") || ob_pad(&b, ' ');
    const long n = (long)width * height;
    const int nb = bitpix == -32 ? 4 : 8;
    if (!rc && grow((void **)&b.p, 1, &b.cap, b.n + n * nb))
        rc = 1;
    if (!rc) {
        unsigned char *p = b.p + b.n;
        for (long i = 0; i < n; i++) {
            if (nb == 4)
                put_e(&p, image[i]);
            else
                put_d(&p, (double)image[i]);
        }
        b.n += n * nb;
        rc = ob_pad(&b, 0);
    }
    if (rc) {
        free(b.p);
        return io_fail(GPAIR_E_NOMEM, "gpair_fits_write_image: out of host memory");
    }
    FILE *fp = fopen(path, "wb");
    if (!fp) {
        free(b.p);
        return io_fail(GPAIR_E_ARG, "cannot create %s", path);
    }
    const int ok = fwrite(b.p, 1, (size_t)b.n, fp) == (size_t)b.n;
    free(b.p);
    if (fclose(fp) != 0 || !ok)
        return io_fail(GPAIR_E_ARG, "short write of %s", path);
    return GPAIR_OK;
}

/* ------------------------------------------------------------------------
 * multi-HDU image files (gpair_io_internal.h)
 * ---------------------------------------------------------------------- */

struct gpair_fitsw {
    obuf b;
    int nhdu;
};

gpair_fitsw *gpair_fitsw_new(void) { return calloc(1, sizeof(gpair_fitsw)); }

void gpair_fitsw_free(gpair_fitsw *w)
{
    if (w) {
        free(w->b.p);
        free(w);
    }
}

int gpair_fitsw_add(gpair_fitsw *w, const char *extname, int bitpix, const void *data, long n1, long n2)
{
    if (!w || !extname || !data || n1 < 1 || n2 < 1 || (bitpix != -32 && bitpix != -64 && bitpix != 32))
        return io_fail(GPAIR_E_ARG, "gpair_fitsw_add: bad arguments");
    obuf *b = &w->b;
    int rc = w->nhdu == 0 ? card_log(b, "SIMPLE", 1, NULL) : card_str(b, "XTENSION", "IMAGE", NULL);
    rc = rc || card_int(b, "BITPIX", bitpix, NULL) || card_int(b, "NAXIS", 2, NULL) ||
         card_int(b, "NAXIS1", n1, NULL) || card_int(b, "NAXIS2", n2, NULL);
    if (w->nhdu == 0)
        rc = rc || card_log(b, "EXTEND", 1, NULL);
    else
        rc = rc || card_int(b, "PCOUNT", 0, NULL) || card_int(b, "GCOUNT", 1, NULL);
    rc = rc || card_str(b, "EXTNAME", extname, NULL) || ob_card(b, "END") || ob_pad(b, ' ');
    const int nb = bitpix == -64 ? 8 : 4;
    const long n = n1 * n2;
    if (!rc && grow((void **)&b->p, 1, &b->cap, b->n + n * nb))
        rc = 1;
    if (!rc) {
        unsigned char *p = b->p + b->n;
        for (long i = 0; i < n; i++) {
            uint64_t bits = 0;
            if (nb == 8)
                memcpy(&bits, (const double *)data + i, 8);
            else {
                uint32_t b32;
                memcpy(&b32, (const uint32_t *)data + i, 4);
                bits = b32;
            }
            put_be(p, bits, nb);
            p += nb;
        }
        b->n += n * nb;
        rc = ob_pad(b, 0);
    }
    if (rc)
        return io_fail(GPAIR_E_NOMEM, "gpair_fitsw_add: out of host memory");
    w->nhdu++;
    return GPAIR_OK;
}

int gpair_fitsw_save(gpair_fitsw *w, const char *path)
{
    if (!w || !path || w->nhdu == 0)
        return io_fail(GPAIR_E_ARG, "gpair_fitsw_save: nothing to write");
    FILE *fp = fopen(path, "wb");
    if (!fp)
        return io_fail(GPAIR_E_ARG, "cannot create %s", path);
    const int ok = fwrite(w->b.p, 1, (size_t)w->b.n, fp) == (size_t)w->b.n;
    if (fclose(fp) != 0 || !ok)
        return io_fail(GPAIR_E_ARG, "short write of %s", path);
    return GPAIR_OK;
}

struct gpair_fitsr {
    ffile f;
};

int gpair_fitsr_open(const char *path, gpair_fitsr **r)
{
    if (!path || !r)
        return io_fail(GPAIR_E_ARG, "gpair_fitsr_open: NULL argument");
    *r = calloc(1, sizeof(gpair_fitsr));
    if (!*r)
        return io_fail(GPAIR_E_NOMEM, "out of host memory");
    const int rc = ff_open(path, &(*r)->f);
    if (rc) {
        free(*r);
        *r = NULL;
    }
    return rc;
}

void gpair_fitsr_close(gpair_fitsr *r)
{
    if (r) {
        ff_close(&r->f);
        free(r);
    }
}

static const fhdu *named_hdu(const ffile *f, const char *extname)
{
    for (int i = 0; i < f->nhdu; i++) {
        char name[72];
        if (hdu_string(&f->h[i], "EXTNAME", name, sizeof(name)) == 0 && strcmp(name, extname) == 0)
            return &f->h[i];
    }
    return NULL;
}

int gpair_fitsr_has(gpair_fitsr *r, const char *extname) { return r && named_hdu(&r->f, extname) != NULL; }

int gpair_fitsr_get(gpair_fitsr *r, const char *extname, int bitpix, void *out, long n1, long n2)
{
    if (!r || !extname || !out)
        return io_fail(GPAIR_E_ARG, "gpair_fitsr_get: NULL argument");
    if (bitpix != -32 && bitpix != -64)
        return io_fail(GPAIR_E_ARG, "gpair_fitsr_get: BITPIX %d (-32 or -64 only)", bitpix);
    const fhdu *h = named_hdu(&r->f, extname);
    long bp = 0;
    if (!h)
        return io_fail(GPAIR_E_ARG, "no %s image in the file", extname);
    hdu_long(h, "BITPIX", &bp);
    if (bp != bitpix || h->naxis1 != n1 || h->naxis2 != n2)
        return io_fail(GPAIR_E_ARG, "%s: BITPIX %ld, %ld x %ld (expected %d, %ld x %ld)", extname, bp, h->naxis1,
                       h->naxis2, bitpix, n1, n2);
    const int nb = bitpix == -64 ? 8 : 4;
    if (n1 < 0 || n2 < 0 || (n2 > 0 && n1 > h->datalen / nb / n2))
        return io_fail(GPAIR_E_ARG, "%s: %ld x %ld values do not fit its data", extname, n1, n2);
    const unsigned char *p = r->f.buf + h->data;
    for (long i = 0; i < n1 * n2; i++, p += nb) {
        const uint64_t bits = be_bits(p, nb);
        if (nb == 8)
            memcpy((double *)out + i, &bits, 8);
        else {
            const uint32_t b32 = (uint32_t)bits;
            memcpy((uint32_t *)out + i, &b32, 4);
        }
    }
    return GPAIR_OK;
}
