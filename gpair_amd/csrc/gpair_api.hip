// C ABI of the MEM-CG path (include/gpair_hip.h): context, buffers, launches.
// Each entry point keeps the meaning of the reference's gpu_* function of the
// same name (gpair1/trunk/gpu.c); the citation sits in the header.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "../../include/gpair_hip.h"
#include "../../include/gpair_io.h"
#include "gpair_io_internal.h"
#include "gpair_kernels.hip"

// gpair1/trunk/getoifits.h:87-88
#define REF_PI 3.14159265358979323
#define REF_RPMAS ((REF_PI / 180.0) / 3600000.0)

struct gpair_buf_s {
    float *p;
    int n;
};

struct gpair_ctx {
    int device;
    hipStream_t stream;
    int W, n, nuv, npow, nbis, ndof;
    int have_data, have_tables, built;
    int precision;
    // data
    float *data, *err;
    float2 *phasor;
    int4 *tri;
    int *csr_ptr, *csr_ent;
    float2 *t3c;   // per T3 leg: its term of the adjoint coefficient of its uv point
    float2 *X, *Y;
    float2 *FX, *FY;   // factored tables (coarse then fine factors per point), from gpair_set_uv
    int have_ft;
    // images
    gpair_buf_s img, tmp, model, gdata, gent, gnew, gold, gtmp, dir;
    float2 *vis, *coef;
    float *mock;
    // reductions
    double *scal;
    float *invflux;
    double *part;
    double *zpart;   // z.g, z.g_old partials of k_grad_finish_crit, 2 per block
    unsigned *ticket;
    double *host_scal;   // pinned
    // adjoint split-K partials
    float *adj_part;
    int adj_nsplit, adj_pper, adj_cap;
    // split-precision (BF16x3) operands
    unsigned short *Xh, *Xl, *Ih, *Il;
    unsigned short *Xth, *Xtl;   // x-major split X for the adjoint, points padded to nuvp
    float2 *fwdp;                // partial sums of the forward's blocks split between two workgroups
    int nuvp;
    int split_valid;
    int adj_b3_nsplit, adj_b3_pper;
    // accelerated loop
    gpair_buf_s z, zold, delta;
    float2 *Ud, *Udel, *vis2;   // DFTs of the direction and of a clamped change
    double2 *Vd, *Vn;
    float *mock2;
    float *pfac;
    double *ls_out, *host_ls;   // 2*LS_K line-search sums, right after scal / host_scal
    // launch shapes chosen by gpair_autotune (0: the built-in default)
    int ls_blocks_chi2, ls_blocks_ent, coef_blocks;
    int fwd_nbk;   // image rows per BF16x3 forward group / 32 (0: min(W, 256) / 32)
    int graphs;    // replay the accelerated loop's gradient phase from a hipGraph
    // sparse commits: the clamped pixels of a committed step, per k_commit_step block
    // (clist_cper slots each), and the most pixels for which the change's DFT is taken
    // as s U_d plus a correction over them (-1: W^2 / 16, 0: always the full DFT)
    int2 *clist;
    int *ccount;
    int *dense;   // set by the sparse kernel: 1 when the commit takes the full DFT
    int clist_cper;
    int commit_cap;
    int adj_tr, adj_tc;   // output tile of the factored split-precision adjoints (rows x columns)
    int adj_groups;       // their workgroups in total (tiles x uv splits), about
    // single-pixel moves: state image, FP64 unnormalised visibilities, x-major
    // tables, trial visibilities and mock data, per-move buffers
    float *px_img;
    double2 *px_U;
    float2 *px_XT, *px_YT, *px_vis, *px_Ud;
    float *px_mock;
    double *px_dF, *px_out, *px_part, *px_dpx;
    unsigned *px_ticket;
    gpair_pixel_move *px_moves;
    double *px_host;                // pinned, 2 * px_cap
    gpair_pixel_move *px_staged;    // pinned, px_cap
    hipEvent_t px_ev;               // last upload from px_staged
    int px_cap, px_valid, px_trial, px_tables, px_n, px_nuv, px_ndof;
    // snapshots of the current image written by the loops
    char snap_path[1024];
    int snap_every;
    float snap_pix;
    // optional per-launch timing of the DFT kernels (HIP events on the context stream)
    int timing;                          // 0: off, N: events around every N-th launch of each kind
    std::vector<hipEvent_t> ev[GPAIR_TK_COUNT];
    size_t ev_used[GPAIR_TK_COUNT];
    unsigned tk_seen[GPAIR_TK_COUNT];    // launches of each kind since timing was enabled
    char msg[512];
};

static int fail(gpair_ctx *c, int code, const char *fmt, ...)
{
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->msg, sizeof(c->msg), fmt, ap);
        va_end(ap);
    }
    return code;
}

#define HIPCHK(ctx, call)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail((ctx), GPAIR_E_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                                 \
    } while (0)

#define LAUNCHCHK(ctx) HIPCHK(ctx, hipGetLastError())

extern "C" const char *gpair_strerror(int code)
{
    switch (code) {
    case GPAIR_OK: return "ok";
    case GPAIR_E_ARG: return "invalid argument";
    case GPAIR_E_HIP: return "HIP runtime error";
    case GPAIR_E_STATE: return "call out of order";
    case GPAIR_E_NOMEM: return "device allocation failed";
    case GPAIR_E_NUMERIC: return "non-finite value";
    default: return "unknown error";
    }
}

extern "C" const char *gpair_last_error(const gpair_ctx *ctx) { return ctx ? ctx->msg : "no context"; }

template <typename T>
static int dalloc(gpair_ctx *c, T **p, size_t count)
{
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    if (count == 0)
        count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, GPAIR_E_NOMEM, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    }
    return GPAIR_OK;
}

template <typename T>
static void dfree(T **p)
{
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
}

// one element per thread, at most 128 workgroups: every workgroup of a reduction
// takes the same ticket with an atomic add, and past ~128 of them that serialised
// atomic costs more than the extra elements per thread
static int red_blocks(int n)
{
    int b = (n + RED_THREADS - 1) / RED_THREADS;
    if (b < 1) b = 1;
    if (b > RED_GRID_MAX) b = RED_GRID_MAX;
    return b;
}

static inline int ew_blocks(int n) { return (n + 255) / 256; }

static inline int tuned_blocks(int v, int dflt) { return v > 0 ? v : dflt; }

// device address of the no-alloc scratch on this device (gpair_noalloc.inc)
static int noalloc_setup(int device);

// ---------------------------------------------------------------------------
// setup / teardown
// ---------------------------------------------------------------------------
extern "C" int gpu_init(gpair_ctx **out, int device)
{
    if (!out)
        return GPAIR_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return GPAIR_E_HIP;
    if (device < 0 || device >= ndev)
        return GPAIR_E_ARG;
    gpair_ctx *c = new (std::nothrow) gpair_ctx();
    if (!c)
        return GPAIR_E_NOMEM;
    c->device = device;
    c->precision = GPAIR_PREC_FP32;
    c->commit_cap = -1;
    c->adj_tr = ADJ_TILE_R;
    c->adj_tc = ADJ_TILE_C;
    c->adj_groups = ADJ_WIDE_GROUPS;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return GPAIR_E_HIP;
    }
    int rc;
    if ((rc = noalloc_setup(device))) {
        gpu_cleanup(c);
        return rc;
    }
    // the line-search sums and the commit partials follow the scalars, so that one
    // copy reads them back (RB_* offsets)
    if ((rc = dalloc(c, &c->scal, RB_COUNT)) || (rc = dalloc(c, &c->invflux, 4)) ||
        (rc = dalloc(c, &c->part, (size_t)RED_MAX_BLOCKS * RED_MAX_NV)) || (rc = dalloc(c, &c->ticket, GPAIR_NTICKET))) {
        gpu_cleanup(c);
        return rc;
    }
    c->ls_out = c->scal + RB_LS;
    if (hipHostMalloc((void **)&c->host_scal, sizeof(double) * RB_COUNT, hipHostMallocDefault) != hipSuccess) {
        gpu_cleanup(c);
        return GPAIR_E_NOMEM;
    }
    c->host_ls = c->host_scal + RB_LS;
    if (hipMemsetAsync(c->ticket, 0, GPAIR_NTICKET * sizeof(unsigned), c->stream) != hipSuccess ||
        hipMemsetAsync(c->scal, 0, RB_COUNT * sizeof(double), c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        gpu_cleanup(c);
        return GPAIR_E_HIP;
    }
    *out = c;
    return GPAIR_OK;
}

static void px_free(gpair_ctx *c);
static inline void px_invalidate(gpair_ctx *c);

static void free_images(gpair_ctx *c)
{
    gpair_buf_s *bufs[] = {&c->img,  &c->tmp,  &c->model, &c->gdata, &c->gent, &c->gnew,
                           &c->gold, &c->gtmp, &c->dir,   &c->z,     &c->zold, &c->delta};
    for (auto *b : bufs) {
        dfree(&b->p);
        b->n = 0;
    }
    dfree(&c->zpart);
    dfree(&c->clist);
    dfree(&c->ccount);
    dfree(&c->dense);
}

extern "C" int gpu_cleanup(gpair_ctx *c)
{
    if (!c)
        return GPAIR_E_ARG;
    (void)hipSetDevice(c->device);
    if (c->stream)
        (void)hipStreamSynchronize(c->stream);
    free_images(c);
    px_free(c);
    dfree(&c->data); dfree(&c->err); dfree(&c->phasor); dfree(&c->tri); dfree(&c->csr_ptr); dfree(&c->csr_ent);
    dfree(&c->t3c);
    dfree(&c->X); dfree(&c->Y); dfree(&c->FX); dfree(&c->FY); dfree(&c->vis); dfree(&c->coef); dfree(&c->mock);
    dfree(&c->scal); dfree(&c->invflux); dfree(&c->part); dfree(&c->ticket); dfree(&c->adj_part);
    dfree(&c->Ud); dfree(&c->Udel); dfree(&c->pfac);
    c->ls_out = nullptr;   // inside scal
    dfree(&c->vis2); dfree(&c->Vd); dfree(&c->Vn); dfree(&c->mock2);
    dfree(&c->Xh); dfree(&c->Xl); dfree(&c->Ih); dfree(&c->Il); dfree(&c->Xth); dfree(&c->Xtl); dfree(&c->fwdp);
    if (c->host_scal)
        (void)hipHostFree(c->host_scal);
    c->host_ls = nullptr;   // inside host_scal
    for (int k = 0; k < GPAIR_TK_COUNT; k++)
        for (hipEvent_t e : c->ev[k])
            (void)hipEventDestroy(e);
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
    return GPAIR_OK;
}

extern "C" int gpu_shutdown(gpair_ctx *c) { return gpu_cleanup(c); }

extern "C" int gpair_sizes(const gpair_ctx *c, int *W, int *nuv, int *npow, int *nbis)
{
    if (!c)
        return GPAIR_E_ARG;
    if (W) *W = c->W;
    if (nuv) *nuv = c->nuv;
    if (npow) *npow = c->npow;
    if (nbis) *nbis = c->nbis;
    return GPAIR_OK;
}

extern "C" int gpair_sync(gpair_ctx *c)
{
    if (!c)
        return GPAIR_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GPAIR_OK;
}

static int check_width(gpair_ctx *c, int W)
{
    if (W < 64 || W > 1024 || W % 64 != 0)
        return fail(c, GPAIR_E_ARG, "image width %d: must be a multiple of 64 in [64, 1024]", W);
    int nb = W / 64;
    if (nb != 1 && nb != 2 && nb != 3 && nb != 4 && nb != 8)
        return fail(c, GPAIR_E_ARG, "image width %d: supported widths are 64, 128, 192, 256, 512", W);
    return GPAIR_OK;
}

static int alloc_images(gpair_ctx *c, int W)
{
    int n = W * W, rc;
    if (c->W == W && c->img.p)
        return GPAIR_OK;
    free_images(c);
    dfree(&c->pfac);   // sized for the previous width
    gpair_buf_s *bufs[] = {&c->img,  &c->tmp,  &c->model, &c->gdata, &c->gent, &c->gnew,
                           &c->gold, &c->gtmp, &c->dir,   &c->z,     &c->zold, &c->delta};
    for (auto *b : bufs) {
        if ((rc = dalloc(c, &b->p, n)))
            return rc;
        b->n = n;
        HIPCHK(c, hipMemsetAsync(b->p, 0, sizeof(float) * n, c->stream));
    }
    if ((rc = dalloc(c, &c->zpart, 2 * (size_t)((n + GF_PIX - 1) / GF_PIX))))
        return rc;
    // k_commit_step's blocks take RED_THREADS pixels per pass, grid-stride
    const int nb = red_blocks(n);
    c->clist_cper = (n + nb * RED_THREADS - 1) / (nb * RED_THREADS) * RED_THREADS;
    if ((rc = dalloc(c, &c->clist, (size_t)nb * c->clist_cper)) || (rc = dalloc(c, &c->ccount, (size_t)nb)) ||
        (rc = dalloc(c, &c->dense, 1)))
        return rc;
    c->W = W;
    c->n = n;
    return GPAIR_OK;
}

extern "C" int gpu_copy_data(gpair_ctx *c, const float *data, const float *data_err, int data_size,
                             int data_size_uv, const float *phasor, int phasor_size, int pow_size,
                             const int64_t *uvpnt, const int16_t *sign, int bsref_size, const float *default_model,
                             int image_size, int image_width)
{
    if (!c || !data || !data_err || !default_model || (bsref_size > 0 && (!phasor || !uvpnt || !sign)))
        return fail(c, GPAIR_E_ARG, "gpu_copy_data: NULL pointer");
    int rc;
    if ((rc = check_width(c, image_width)))
        return rc;
    const int nuv = data_size_uv, npow = pow_size, nbis = bsref_size;
    if (nuv <= 0 || npow < 0 || npow > nuv || nbis < 0 || phasor_size != nbis || data_size != npow + 2 * nbis ||
        image_size != image_width * image_width)
        return fail(c, GPAIR_E_ARG,
                    "gpu_copy_data: inconsistent sizes (data %d, nuv %d, npow %d, nbis %d, phasor %d, image %d)",
                    data_size, nuv, npow, nbis, phasor_size, image_size);
    if (npow + nbis == 0)   // chi2 / ndof and the step scales need at least one datum
        return fail(c, GPAIR_E_ARG, "gpu_copy_data: no V2 and no T3 data");
    // pack T3 references: int4 (ab, bc, ca, conj bits) and the uv -> leg CSR
    std::vector<int4> tri(nbis > 0 ? nbis : 1);
    std::vector<int> cnt(nuv + 1, 0);
    for (int t = 0; t < nbis; t++) {
        int idx[3], bits = 0;
        for (int j = 0; j < 3; j++) {
            int64_t p = uvpnt[4 * t + j];
            int s = sign[4 * t + j];
            if (p < 0 || p >= nuv || (s != 1 && s != -1))
                return fail(c, GPAIR_E_ARG, "gpu_copy_data: T3 %d leg %d: uv index %lld sign %d invalid", t, j,
                            (long long)p, s);
            idx[j] = (int)p;
            if (s < 0)
                bits |= 1 << j;
            cnt[p + 1]++;
        }
        tri[t] = make_int4(idx[0], idx[1], idx[2], bits);
    }
    for (int p = 0; p < nuv; p++)
        cnt[p + 1] += cnt[p];
    std::vector<int> ent(3 * (size_t)nbis + 1), fill(cnt.begin(), cnt.end() - 1);
    for (int t = 0; t < nbis; t++)
        for (int j = 0; j < 3; j++)
            ent[fill[uvpnt[4 * t + j]]++] = 3 * t + j;
    HIPCHK(c, hipSetDevice(c->device));
    px_invalidate(c);   // sizes may change: the pixel state is rebuilt on its next use
    const int ndof = npow + 2 * nbis;
    if ((rc = dalloc(c, &c->data, ndof)) || (rc = dalloc(c, &c->err, ndof)) || (rc = dalloc(c, &c->mock, ndof)) ||
        (rc = dalloc(c, &c->phasor, nbis)) || (rc = dalloc(c, &c->tri, nbis)) ||
        (rc = dalloc(c, &c->csr_ptr, nuv + 1)) || (rc = dalloc(c, &c->csr_ent, 3 * (size_t)nbis)) ||
        (rc = dalloc(c, &c->t3c, 3 * (size_t)nbis)) ||
        (rc = dalloc(c, &c->vis, nuv)) || (rc = dalloc(c, &c->coef, nuv)) || (rc = dalloc(c, &c->Ud, nuv)) ||
        (rc = dalloc(c, &c->Udel, nuv)) || (rc = dalloc(c, &c->vis2, nuv)) || (rc = dalloc(c, &c->Vd, nuv)) || (rc = dalloc(c, &c->Vn, nuv)) ||
        (rc = dalloc(c, &c->mock2, ndof)))
        return rc;
    if ((rc = alloc_images(c, image_width)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->data, data, sizeof(float) * ndof, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->err, data_err, sizeof(float) * ndof, hipMemcpyHostToDevice, c->stream));
    if (nbis > 0) {
        HIPCHK(c, hipMemcpyAsync(c->phasor, phasor, sizeof(float2) * nbis, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->tri, tri.data(), sizeof(int4) * nbis, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->csr_ent, ent.data(), sizeof(int) * 3 * (size_t)nbis, hipMemcpyHostToDevice,
                                 c->stream));
    }
    HIPCHK(c, hipMemcpyAsync(c->csr_ptr, cnt.data(), sizeof(int) * (nuv + 1), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->model.p, default_model, sizeof(float) * image_size, hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipMemsetAsync(c->mock, 0, sizeof(float) * ndof, c->stream));
    HIPCHK(c, hipMemsetAsync(c->vis, 0, sizeof(float2) * nuv, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));   // host vectors go out of scope
    if (c->nuv != nuv) {
        c->have_tables = 0;
        c->have_ft = 0;
    }
    c->nuv = nuv;
    c->npow = npow;
    c->nbis = nbis;
    c->ndof = ndof;
    c->have_data = 1;
    c->built = 0;
    return GPAIR_OK;
}

extern "C" int gpu_copy_image(gpair_ctx *c, const float *image, int x_size, int y_size)
{
    if (!c || !image)
        return fail(c, GPAIR_E_ARG, "gpu_copy_image: NULL pointer");
    if (!c->have_data)
        return fail(c, GPAIR_E_STATE, "gpu_copy_image before gpu_copy_data");
    if (x_size != c->W || y_size != c->W)
        return fail(c, GPAIR_E_ARG, "gpu_copy_image: %dx%d, expected %dx%d", x_size, y_size, c->W, c->W);
    px_invalidate(c);
    HIPCHK(c, hipMemcpyAsync(c->img.p, image, sizeof(float) * c->n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GPAIR_OK;
}

extern "C" int gpair_set_uv(gpair_ctx *c, const float *u, const float *v, int nuv, float pixellation)
{
    if (!c || !u || !v)
        return fail(c, GPAIR_E_ARG, "gpair_set_uv: NULL pointer");
    if (!c->have_data)
        return fail(c, GPAIR_E_STATE, "gpair_set_uv before gpu_copy_data");
    if (nuv != c->nuv)
        return fail(c, GPAIR_E_ARG, "gpair_set_uv: nuv %d, data has %d", nuv, c->nuv);
    if (!(pixellation > 0.f) || !isfinite(pixellation))
        return fail(c, GPAIR_E_ARG, "gpair_set_uv: pixellation %g", (double)pixellation);
    int rc;
    const size_t tsz = (size_t)nuv * c->W, fsz = (size_t)nuv * ft_len(c->W);
    c->have_ft = 0;
    if ((rc = dalloc(c, &c->X, tsz)) || (rc = dalloc(c, &c->Y, tsz)) || (rc = dalloc(c, &c->FX, fsz)) ||
        (rc = dalloc(c, &c->FY, fsz)))
        return rc;
    float *du = nullptr, *dv = nullptr;
    if ((rc = dalloc(c, &du, nuv)) || (rc = dalloc(c, &dv, nuv))) {
        dfree(&du);
        return rc;
    }
    int err = GPAIR_OK;
    if (hipMemcpyAsync(du, u, sizeof(float) * nuv, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(dv, v, sizeof(float) * nuv, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        err = fail(c, GPAIR_E_HIP, "gpair_set_uv: upload failed");
    if (!err) {
        const double a = REF_RPMAS * (double)pixellation;
        k_tables<<<(unsigned)((tsz + 255) / 256), 256, 0, c->stream>>>(du, dv, a, nuv, c->W, c->X, c->Y);
        k_ftables<<<(unsigned)((fsz + 255) / 256), 256, 0, c->stream>>>(du, dv, a, nuv, c->W, c->FX, c->FY);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
            err = fail(c, GPAIR_E_HIP, "gpair_set_uv: table kernel failed");
    }
    dfree(&du);
    dfree(&dv);
    if (!err) {
        c->have_tables = 1;
        c->have_ft = 1;
    }
    c->split_valid = 0;
    px_invalidate(c);
    c->px_tables = 0;
    return err;
}

extern "C" int gpu_copy_dft(gpair_ctx *c, const float *dft_x, const float *dft_y, int dft_size)
{
    if (!c || !dft_x || !dft_y)
        return fail(c, GPAIR_E_ARG, "gpu_copy_dft: NULL pointer");
    if (!c->have_data)
        return fail(c, GPAIR_E_STATE, "gpu_copy_dft before gpu_copy_data");
    const size_t tsz = (size_t)c->nuv * c->W;
    if ((size_t)dft_size < tsz)
        return fail(c, GPAIR_E_ARG, "gpu_copy_dft: %d entries, need nuv*W = %zu", dft_size, tsz);
    int rc;
    c->have_ft = 0;   // arbitrary tables: the BF16x3 DFTs read their full split copies
    if ((rc = dalloc(c, &c->X, tsz)) || (rc = dalloc(c, &c->Y, tsz)))
        return rc;
    float2 *tmp = nullptr;
    if ((rc = dalloc(c, &tmp, tsz)))
        return rc;
    const float *src[2] = {dft_x, dft_y};
    float2 *dst[2] = {c->X, c->Y};
    int err = GPAIR_OK;
    for (int k = 0; k < 2 && !err; k++) {
        if (hipMemcpyAsync(tmp, src[k], sizeof(float2) * tsz, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
            err = fail(c, GPAIR_E_HIP, "gpu_copy_dft: upload failed");
            break;
        }
        k_transpose_tables<<<(unsigned)((tsz + 255) / 256), 256, 0, c->stream>>>(tmp, c->nuv, c->W, dst[k]);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
            err = fail(c, GPAIR_E_HIP, "gpu_copy_dft: transpose failed");
    }
    dfree(&tmp);
    if (!err)
        c->have_tables = 1;
    c->split_valid = 0;
    px_invalidate(c);
    c->px_tables = 0;
    return err;
}

// split of the adjoint over uv for about `groups` workgroups in total (b3: the
// split-precision adjoint's 64 x 128 tiles, else the FP32 adjoint's 64 x 64 tiles);
// grows the shared partial buffer when needed
// the split-precision adjoint takes 256 x 256 quadrants (k_adj_b3w) when the width allows
static inline int adj_wide(const gpair_ctx *c) { return ADJ_WIDE && c->W % AW_T == 0; }

static inline int adj_factored(const gpair_ctx *c);

// output tile of the split-precision adjoints on factored tables: the configured one
// when it divides the width, else 128 x 128 when that does, else 64 x 64
static inline void adj_tile(const gpair_ctx *c, int *tr, int *tc)
{
    const int fit = c->W % c->adj_tr == 0 && c->W % c->adj_tc == 0;
    const int t = c->W % 128 == 0 ? 128 : 64;
    *tr = fit ? c->adj_tr : t;
    *tc = fit ? c->adj_tc : t;
}

static inline int adj_tiles(const gpair_ctx *c)
{
    int tr, tc;
    adj_tile(c, &tr, &tc);
    return ((c->W + tr - 1) / tr) * ((c->W + tc - 1) / tc);
}

static int set_adj_split(gpair_ctx *c, int b3, int groups)
{
    const int tiles = b3 ? (adj_factored(c) ? adj_tiles(c)
                            : adj_wide(c)   ? (c->W / AW_T) * (c->W / AW_T)
                                            : (c->W / AB3_YT) * (c->W / AB3_XT))
                         : (c->W / 64) * (c->W / 64);
    // split-precision splits take at least ADJ_MIN_PTS points (four rounds of 16):
    // below that a workgroup's build and store of its tile outweigh its products,
    // and the partial images to add up grow with the splits
    const int gran = b3 ? AB3_PR : 8, minpts = b3 ? ADJ_MIN_PTS : 8;
    int nsplit = groups / tiles;
    if (nsplit < 1) nsplit = 1;
    const int maxsplit = (c->nuv + minpts - 1) / minpts;
    if (nsplit > maxsplit) nsplit = maxsplit;
    int pper = (c->nuv + nsplit - 1) / nsplit;
    pper = (pper + gran - 1) / gran * gran;
    nsplit = (c->nuv + pper - 1) / pper;
    if (nsplit > c->adj_cap) {   // one partial buffer serves both adjoints
        int rc;
        if ((rc = dalloc(c, &c->adj_part, (size_t)nsplit * c->n)))
            return rc;
        c->adj_cap = nsplit;
    }
    if (b3) {
        c->adj_b3_nsplit = nsplit;
        c->adj_b3_pper = pper;
    } else {
        c->adj_nsplit = nsplit;
        c->adj_pper = pper;
    }
    return GPAIR_OK;
}

extern "C" int gpu_build_kernels(gpair_ctx *c, int data_size, int image_width, int image_size)
{
    if (!c)
        return GPAIR_E_ARG;
    if (!c->have_data || !c->have_tables)
        return fail(c, GPAIR_E_STATE, "gpu_build_kernels needs data and DFT tables first");
    if (data_size != c->ndof || image_width != c->W || image_size != c->n)
        return fail(c, GPAIR_E_ARG, "gpu_build_kernels: sizes (%d, %d, %d) differ from the data (%d, %d, %d)",
                    data_size, image_width, image_size, c->ndof, c->W, c->n);
    // adjoint split over uv: about 1024 workgroups in total
    int rc;
    c->adj_cap = 0;
    c->adj_b3_nsplit = 0;
    if ((rc = set_adj_split(c, 0, 1024)))
        return rc;
    c->split_valid = 0;   // the split-precision adjoint sizes its partials on first use
    c->built = 1;
    return GPAIR_OK;
}

extern "C" int gpair_set_precision(gpair_ctx *c, int precision)
{
    if (!c)
        return GPAIR_E_ARG;
    if (precision != GPAIR_PREC_FP32 && precision != GPAIR_PREC_BF16X3 && precision != GPAIR_PREC_FP16S)
        return fail(c, GPAIR_E_ARG, "precision %d not available", precision);
    c->precision = precision;
    return GPAIR_OK;
}

// BF16x3 operands: the split image buffers always; the split copies of the X table
// only where a kernel reads them (tables copied in by gpu_copy_dft, or an adjoint
// width without the factored quadrant kernel)
static inline int adj_factored(const gpair_ctx *c) { return c->have_ft && c->W % 64 == 0; }

static int ensure_split_tables(gpair_ctx *c)
{
    if (c->split_valid)
        return GPAIR_OK;
    if (c->have_ft && ft_len(c->W) > FYS_MAX)   // the factored forwards stage a block's FY rows in LDS
        return fail(c, GPAIR_E_ARG, "width %d above the factored DFTs' maximum of 512", c->W);
    const size_t tsz = (size_t)c->nuv * c->W;
    int rc;
    if ((rc = dalloc(c, &c->Ih, (size_t)c->n)) || (rc = dalloc(c, &c->Il, (size_t)c->n)) ||
        (rc = dalloc(c, &c->fwdp, 2 * (size_t)c->nuv)))
        return rc;
    if (!c->have_ft) {
        if ((rc = dalloc(c, &c->Xh, 2 * tsz)) || (rc = dalloc(c, &c->Xl, 2 * tsz)))
            return rc;
        k_split_tables<<<(unsigned)((tsz + 255) / 256), 256, 0, c->stream>>>(c->X, c->nuv, c->W, c->Xh, c->Xl);
        LAUNCHCHK(c);
    } else {
        dfree(&c->Xh);
        dfree(&c->Xl);
    }
    if (adj_factored(c)) {   // tiles of the factored tables: any multiple of 64
        dfree(&c->Xth);
        dfree(&c->Xtl);
        if (!c->adj_b3_nsplit && (rc = set_adj_split(c, 1, c->adj_groups)))
            return rc;
    } else if (c->W % AB3_XT == 0) {   // uploaded tables: 64 x 128 or 256 x 256 tiles of split copies
        c->nuvp = (c->nuv + AB3_PR - 1) / AB3_PR * AB3_PR;
        const size_t tt = (size_t)c->nuvp * c->W;
        if ((rc = dalloc(c, &c->Xth, 2 * tt)) || (rc = dalloc(c, &c->Xtl, 2 * tt)))
            return rc;
        k_split_tables_t<<<(unsigned)((tt + 255) / 256), 256, 0, c->stream>>>(c->X, c->nuv, c->nuvp, c->W,
                                                                             c->Xth, c->Xtl);
        LAUNCHCHK(c);
        if (!c->adj_b3_nsplit && (rc = set_adj_split(c, 1, adj_wide(c) ? ADJ_WIDE_GROUPS : 512)))
            return rc;
    }
    c->split_valid = 1;
    return GPAIR_OK;
}

// FP16 with scaling: needs the factored tables; the FP16 image shares the BF16x3 hi
// buffer, the split forward blocks and the quadrant adjoint's uv split are the same
static int ensure_f16(gpair_ctx *c)
{
    if (!c->have_ft)
        return fail(c, GPAIR_E_STATE, "FP16 DFTs need tables set from uv (gpair_set_uv), not uploaded ones");
    return ensure_split_tables(c);
}

// ---------------------------------------------------------------------------
// buffer handles
// ---------------------------------------------------------------------------
extern "C" gpair_buf gpu_getp_ci(gpair_ctx *c) { return c ? &c->img : nullptr; }
extern "C" gpair_buf gpu_getp_ti(gpair_ctx *c) { return c ? &c->tmp : nullptr; }
extern "C" gpair_buf gpu_getp_fgn(gpair_ctx *c) { return c ? &c->gnew : nullptr; }
extern "C" gpair_buf gpu_getp_fg(gpair_ctx *c) { return c ? &c->gold : nullptr; }
extern "C" gpair_buf gpu_getp_dd(gpair_ctx *c) { return c ? &c->dir : nullptr; }
extern "C" gpair_buf gpu_getp_tg(gpair_ctx *c) { return c ? &c->gtmp : nullptr; }
extern "C" gpair_buf gpu_getp_eg(gpair_ctx *c) { return c ? &c->gent : nullptr; }
extern "C" gpair_buf gpair_getp_dg(gpair_ctx *c) { return c ? &c->gdata : nullptr; }

static int ready(gpair_ctx *c)
{
    if (!c)
        return GPAIR_E_ARG;
    if (!c->built)
        return fail(c, GPAIR_E_STATE, "call gpu_copy_data, gpair_set_uv/gpu_copy_dft and gpu_build_kernels first");
    return GPAIR_OK;
}

static int own_buf(gpair_ctx *c, gpair_buf b)
{
    gpair_buf_s *bufs[] = {&c->img,  &c->tmp,  &c->model, &c->gdata, &c->gent, &c->gnew,
                           &c->gold, &c->gtmp, &c->dir,   &c->z,     &c->zold, &c->delta};
    for (auto *x : bufs)
        if (x == b && b->p)
            return 1;
    return 0;
}

static int read_scalar(gpair_ctx *c, int slot, double *out)
{
    HIPCHK(c, hipMemcpyAsync(c->host_scal + slot, c->scal + slot, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = c->host_scal[slot];
    return GPAIR_OK;
}

// ---------------------------------------------------------------------------
// kernel timing: event pairs around the DFT launches while enabled
// ---------------------------------------------------------------------------
#define GPAIR_TIMING_MAX_EVENTS 400000

static hipEvent_t timing_event(gpair_ctx *c, int k)
{
    if (c->ev_used[k] >= GPAIR_TIMING_MAX_EVENTS)
        return nullptr;
    if (c->ev_used[k] == c->ev[k].size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess)
            return nullptr;
        c->ev[k].push_back(e);
    }
    return c->ev[k][c->ev_used[k]++];
}

// records a start/stop pair around one launch; pairs are kept even-aligned
#define TIMED_LAUNCH(c, k, launch)                                                   \
    do {                                                                             \
        hipEvent_t a_ = nullptr, b_ = nullptr;                                       \
        if ((c)->timing && (c)->tk_seen[k]++ % (unsigned)(c)->timing == 0 &&         \
            (c)->ev_used[k] + 2 <= GPAIR_TIMING_MAX_EVENTS) {                         \
            a_ = timing_event(c, k);                                                 \
            b_ = a_ ? timing_event(c, k) : nullptr;                                  \
            if (a_ && !b_)                                                           \
                (c)->ev_used[k]--;                                                   \
        }                                                                            \
        if (a_ && b_)                                                                \
            (void)hipEventRecord(a_, (c)->stream);                                   \
        launch;                                                                      \
        if (a_ && b_)                                                                \
            (void)hipEventRecord(b_, (c)->stream);                                   \
    } while (0)

extern "C" int gpair_timing_enable(gpair_ctx *c, int on)
{
    if (!c)
        return GPAIR_E_ARG;
    c->timing = on > 0 ? on : 0;
    for (int k = 0; k < GPAIR_TK_COUNT; k++) {
        c->ev_used[k] = 0;
        c->tk_seen[k] = 0;
    }
    return GPAIR_OK;
}

extern "C" int gpair_timing_read(gpair_ctx *c, int kernel, int *count, double *total_ms)
{
    if (!c || kernel < 0 || kernel >= GPAIR_TK_COUNT || !count || !total_ms)
        return fail(c, GPAIR_E_ARG, "gpair_timing_read: bad arguments");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double tot = 0.0;
    int n = 0;
    for (size_t i = 0; i + 1 < c->ev_used[kernel]; i += 2) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev[kernel][i], c->ev[kernel][i + 1]));
        tot += ms;
        n++;
    }
    *count = n;
    *total_ms = tot;
    return GPAIR_OK;
}

// ---------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------
static int launch_flux(gpair_ctx *c, const float *img)
{
    k_flux<<<red_blocks(c->n), RED_THREADS, 0, c->stream>>>(img, c->n, c->part, c->ticket, c->scal, c->invflux);
    LAUNCHCHK(c);
    return GPAIR_OK;
}

// vis_out = scale ? (*scale) A img : A img  (A = unnormalised DFT)
// presplit: the BF16 split of img is already in Ih / Il (written by its producer)
// run_if (FP32 and factored BF16x3 only): a device flag; the launches do nothing when
// it is 0 (a commit the sparse correction gave); such launches are not timed
static int launch_forward_to(gpair_ctx *c, const float *img, const float *scale, float2 *vis_out, int presplit = 0,
                             const int *run_if = nullptr)
{
    if (c->precision == GPAIR_PREC_FP16S) {
        int rc;
        if ((rc = ensure_f16(c)))
            return rc;
        // scale: max |img| (deterministic max reduction), then the FP16 image
        k_absmax<<<red_blocks(c->n), RED_THREADS, 0, c->stream>>>(img, c->n, c->part, c->ticket, c->scal + SC_AMAXI);
        k_split_image_h<<<ew_blocks(c->n), 256, 0, c->stream>>>(img, c->n, c->scal + SC_AMAXI, c->Ih);
        LAUNCHCHK(c);
        const int nblk = (c->nuv + B3_PW - 1) / B3_PW, nchunk = c->W / B3_KC;
        const int nwg = nblk < B3_WG ? nblk : B3_WG;
        const int upw = (nblk * nchunk + nwg - 1) / nwg;
        const int nbk = c->fwd_nbk ? c->fwd_nbk : (c->W < B3_ROWS ? c->W : B3_ROWS) / 32;
        if (c->W % (32 * nbk) != 0)
            return fail(c, GPAIR_E_ARG, "FP16 forward: width %d is not a multiple of %d rows", c->W, 32 * nbk);
        switch (nbk) {
#define H1_CASE(N)                                                                                            \
    case N:                                                                                                   \
        TIMED_LAUNCH(c, GPAIR_TK_FORWARD,                                                                     \
                     (k_vis_fwd_h1f<N><<<nwg, 256, 0, c->stream>>>(c->Ih, c->FX, c->FY, c->nuv, c->W, upw, scale, \
                                                                   c->scal + SC_AMAXI, vis_out, c->fwdp),       \
                      k_vis_fwd_b3_combine<<<nwg > 1 ? nwg - 1 : 1, B3_PW, 0, c->stream>>>(c->fwdp, c->nuv, c->W, \
                                                                                         upw, scale, vis_out, \
                                                                                         nullptr)));          \
        break;
            H1_CASE(2)
            H1_CASE(4)
            H1_CASE(6)
            H1_CASE(8)
#undef H1_CASE
        default: return fail(c, GPAIR_E_ARG, "unsupported width %d", c->W);
        }
        LAUNCHCHK(c);
        return GPAIR_OK;
    }
    if (c->precision == GPAIR_PREC_BF16X3) {
        int rc;
        if ((rc = ensure_split_tables(c)))
            return rc;
        if (!presplit) {
            k_split_image<<<ew_blocks(c->n), 256, 0, c->stream>>>(img, c->n, c->Ih, c->Il);
            LAUNCHCHK(c);
        }
        // stream-K balance: at most B3_WG workgroups (one per CU at one wave per SIMD),
        // each taking an equal run of (64-point block, 32-column chunk) units
        const int nblk = (c->nuv + B3_PW - 1) / B3_PW, nchunk = c->W / B3_KC;
        const int nwg = nblk < B3_WG ? nblk : B3_WG;
        const int upw = (nblk * nchunk + nwg - 1) / nwg;
        const int nbk = c->fwd_nbk ? c->fwd_nbk : (c->W < B3_ROWS ? c->W : B3_ROWS) / 32;
        if (c->W % (32 * nbk) != 0)
            return fail(c, GPAIR_E_ARG, "BF16x3 forward: width %d is not a multiple of %d rows", c->W, 32 * nbk);
        switch (nbk) {
#define B3_CASE(N)                                                                                             \
    case N:                                                                                                    \
        if (c->have_ft && run_if) {                                                                            \
            k_vis_fwd_b3f<N><<<nwg, 256, 0, c->stream>>>(c->Ih, c->Il, c->FX, c->FY, c->nuv, c->W, upw, scale,      \
                                                          vis_out, c->fwdp, run_if);                               \
            k_vis_fwd_b3_combine<<<nwg > 1 ? nwg - 1 : 1, B3_PW, 0, c->stream>>>(c->fwdp, c->nuv, c->W, upw,       \
                                                                                 scale, vis_out, run_if);          \
        } else if (c->have_ft)                                                                                 \
            TIMED_LAUNCH(c, GPAIR_TK_FORWARD,                                                                  \
                         (k_vis_fwd_b3f<N><<<nwg, 256, 0, c->stream>>>(c->Ih, c->Il, c->FX, c->FY, c->nuv, c->W, \
                                                                       upw, scale, vis_out, c->fwdp, nullptr), \
                          k_vis_fwd_b3_combine<<<nwg > 1 ? nwg - 1 : 1, B3_PW, 0, c->stream>>>(             \
                              c->fwdp, c->nuv, c->W, upw, scale, vis_out, nullptr)));                          \
        else                                                                                                   \
            TIMED_LAUNCH(c, GPAIR_TK_FORWARD,                                                                  \
                         (k_vis_fwd_b3<N><<<nwg, 256, 0, c->stream>>>(c->Ih, c->Il, c->Xh, c->Xl, c->Y, c->nuv,  \
                                                                      c->W, upw, scale, vis_out, c->fwdp),     \
                          k_vis_fwd_b3_combine<<<nwg > 1 ? nwg - 1 : 1, B3_PW, 0, c->stream>>>(             \
                              c->fwdp, c->nuv, c->W, upw, scale, vis_out, nullptr)));                          \
        break;
            B3_CASE(2)
            B3_CASE(4)
            B3_CASE(6)
            B3_CASE(8)
#undef B3_CASE
        default: return fail(c, GPAIR_E_ARG, "unsupported width %d", c->W);
        }
        LAUNCHCHK(c);
        return GPAIR_OK;
    }
    const unsigned grid = (unsigned)((c->nuv + FWD_PT - 1) / FWD_PT);
    switch (c->W / 64) {
#define FWD_CASE(NB)                                                                                       \
    case NB:                                                                                               \
        if (run_if)                                                                                        \
            k_vis_fwd_f32<NB><<<grid, 256, 0, c->stream>>>(img, c->X, c->Y, c->nuv, scale, vis_out, run_if); \
        else                                                                                               \
            TIMED_LAUNCH(c, GPAIR_TK_FORWARD, (k_vis_fwd_f32<NB><<<grid, 256, 0, c->stream>>>(               \
                                                  img, c->X, c->Y, c->nuv, scale, vis_out, nullptr)));     \
        break;
    FWD_CASE(1)
    FWD_CASE(2)
    FWD_CASE(3)
    FWD_CASE(4)
    FWD_CASE(8)
#undef FWD_CASE
    default: return fail(c, GPAIR_E_ARG, "unsupported width %d", c->W);
    }
    LAUNCHCHK(c);
    return GPAIR_OK;
}

static int launch_forward(gpair_ctx *c, const float *img) { return launch_forward_to(c, img, c->invflux, c->vis); }

static int launch_chi2(gpair_ctx *c, const float *img)
{
    int rc;
    if ((rc = launch_flux(c, img)) || (rc = launch_forward(c, img)))
        return rc;
    const int items = c->npow + c->nbis;
    k_model_chi2<<<red_blocks(items), RED_THREADS, 0, c->stream>>>(c->vis, c->data, c->err, c->phasor, c->tri,
                                                                    c->npow, c->nbis, c->mock, c->part,
                                                                    c->ticket, c->scal);
    LAUNCHCHK(c);
    return GPAIR_OK;
}

static int launch_adjoint(gpair_ctx *c);

static int launch_adjoint_dft(gpair_ctx *c, int *nsplit_out);

// per-uv adjoint coefficients from the visibilities and mock data of the last chi2
// (k_coef) and the flux of img (as launch_flux) in one launch, with the same results
static int launch_t3_contrib(gpair_ctx *c)
{
    if (c->nbis > 0) {
        k_t3_contrib<<<ew_blocks(c->nbis), 256, 0, c->stream>>>(c->vis, c->mock, c->data, c->err, c->phasor, c->tri,
                                                               c->npow, c->nbis, c->t3c);
        LAUNCHCHK(c);
    }
    return GPAIR_OK;
}

static int launch_coef_flux(gpair_ctx *c, const float *img)
{
    int rc;
    if ((rc = launch_t3_contrib(c)))
        return rc;
    const int nbc = tuned_blocks(c->coef_blocks, red_nparts(c->nuv, COEF_SPAN)), nbf = red_blocks(c->n);
    k_coef_flux<<<nbc + nbf, RED_THREADS, 0, c->stream>>>(c->vis, c->mock, c->data, c->err, c->t3c, c->csr_ptr,
                                                           c->csr_ent, c->npow, c->nuv, c->coef, nbc, img, c->n,
                                                           c->part, c->ticket, c->scal, c->invflux);
    LAUNCHCHK(c);
    return GPAIR_OK;
}

static int launch_data_gradient(gpair_ctx *c, const float *img)
{
    int rc;
    if ((rc = launch_coef_flux(c, img)))
        return rc;
    return launch_adjoint(c);
}

// adjoint DFT of the coefficients in c->coef into the split partials; *nsplit_out = their count
static int launch_adjoint_dft(gpair_ctx *c, int *nsplit_out)
{
    int rc;
    int nsplit = c->adj_nsplit;
    if (c->precision == GPAIR_PREC_FP16S && adj_factored(c)) {
        if ((rc = ensure_f16(c)))
            return rc;
        nsplit = c->adj_b3_nsplit;
        k_absmax<<<red_blocks(2 * c->nuv), RED_THREADS, 0, c->stream>>>(reinterpret_cast<const float *>(c->coef),
                                                                         2 * c->nuv, c->part, c->ticket,
                                                                         c->scal + SC_AMAXC);
        LAUNCHCHK(c);
        int tr, tc;
        adj_tile(c, &tr, &tc);
        const unsigned grid = (unsigned)(adj_tiles(c) * nsplit);
#define ADJ_H1(TR, TC)                                                                                              \
    TIMED_LAUNCH(c, GPAIR_TK_ADJOINT,                                                                               \
                 (k_adj_h1wf<TR, TC><<<grid, ADJ_NTH(TR), 0, c->stream>>>(c->coef, c->FX, c->FY, c->nuv, c->W,      \
                                                                          c->adj_b3_pper, nsplit,                   \
                                                                          c->scal + SC_AMAXC, c->adj_part)))
        if (tr == 256 && tc == 256)
            ADJ_H1(256, 256);
        else if (tr == 128 && tc == 256)
            ADJ_H1(128, 256);
        else if (tr == 128)
            ADJ_H1(128, 128);
        else
            ADJ_H1(64, 64);
#undef ADJ_H1
    } else if (c->precision == GPAIR_PREC_BF16X3 && (adj_factored(c) || c->W % AB3_XT == 0)) {
        if ((rc = ensure_split_tables(c)))
            return rc;
        nsplit = c->adj_b3_nsplit;
        if (adj_factored(c)) {
            int tr, tc;
            adj_tile(c, &tr, &tc);
            const unsigned grid = (unsigned)(adj_tiles(c) * nsplit);
#define ADJ_B3(TR, TC)                                                                                              \
    TIMED_LAUNCH(c, GPAIR_TK_ADJOINT,                                                                               \
                 (k_adj_b3wf<TR, TC><<<grid, ADJ_NTH(TR), 0, c->stream>>>(c->coef, c->FX, c->FY, c->nuv, c->W,      \
                                                                          c->adj_b3_pper, nsplit, c->adj_part)))
            if (tr == 256 && tc == 256)
                ADJ_B3(256, 256);
            else if (tr == 128 && tc == 256)
                ADJ_B3(128, 256);
            else if (tr == 128)
                ADJ_B3(128, 128);
            else
                ADJ_B3(64, 64);
#undef ADJ_B3
        } else if (adj_wide(c)) {
            dim3 grid((c->W / AW_T) * (c->W / AW_T), nsplit);
            TIMED_LAUNCH(c, GPAIR_TK_ADJOINT,
                         (k_adj_b3w<<<grid, 512, 0, c->stream>>>(c->coef, c->Y, c->Xth, c->Xtl, c->nuv, c->nuvp,
                                                                 c->W, c->adj_b3_pper, c->adj_part)));
        } else {
            dim3 grid((c->W / AB3_YT) * (c->W / AB3_XT), nsplit);
            TIMED_LAUNCH(c, GPAIR_TK_ADJOINT,
                         (k_adj_b3<<<grid, 256, 0, c->stream>>>(c->coef, c->Y, c->Xth, c->Xtl, c->nuv, c->nuvp,
                                                                c->W, c->adj_b3_pper, c->adj_part)));
        }
    } else {
        dim3 grid((c->W / 64) * (c->W / 64), nsplit);
        TIMED_LAUNCH(c, GPAIR_TK_ADJOINT,
                     (k_adj_f32<<<grid, 256, 0, c->stream>>>(c->coef, c->X, c->Y, c->nuv, c->W, c->adj_pper,
                                                             c->adj_part)));
    }
    LAUNCHCHK(c);
    *nsplit_out = nsplit;
    return GPAIR_OK;
}

// adjoint DFT of the coefficients in c->coef and the fixed-order sum of its partials
static int launch_adjoint(gpair_ctx *c)
{
    int rc, nsplit;
    if ((rc = launch_adjoint_dft(c, &nsplit)))
        return rc;
    k_grad_finish<<<(c->n + GF_PIX - 1) / GF_PIX, 256, 0, c->stream>>>(c->adj_part, nsplit, c->n, c->scal,
                                                                          c->gdata.p);
    LAUNCHCHK(c);
    return GPAIR_OK;
}

static int check_counts(gpair_ctx *c, int nuv, int npow, int nbis)
{
    if (nuv != c->nuv || npow != c->npow || nbis != c->nbis)
        return fail(c, GPAIR_E_ARG, "sizes (nuv %d, npow %d, nbis %d) differ from the data (%d, %d, %d)", nuv, npow,
                    nbis, c->nuv, c->npow, c->nbis);
    return GPAIR_OK;
}

static int get_chi2(gpair_ctx *c, const float *img, int nuv, int npow, int nbis, double *chi2)
{
    int rc;
    if ((rc = ready(c)) || (rc = check_counts(c, nuv, npow, nbis)))
        return rc;
    if (!chi2)
        return fail(c, GPAIR_E_ARG, "chi2 output is NULL");
    if ((rc = launch_chi2(c, img)))
        return rc;
    return read_scalar(c, SC_CHI2, chi2);
}

extern "C" int gpu_get_chi2_curr(gpair_ctx *c, int nuv, int npow, int nbis, int data_alloc, int data_alloc_uv,
                                 double *chi2)
{
    (void)data_alloc; (void)data_alloc_uv;
    return get_chi2(c, c ? c->img.p : nullptr, nuv, npow, nbis, chi2);
}

extern "C" int gpu_get_chi2_temp(gpair_ctx *c, int nuv, int npow, int nbis, int data_alloc, int data_alloc_uv,
                                 double *chi2)
{
    (void)data_alloc; (void)data_alloc_uv;
    return get_chi2(c, c ? c->tmp.p : nullptr, nuv, npow, nbis, chi2);
}

static int get_entropy(gpair_ctx *c, const float *img, int W, double *S)
{
    int rc;
    if ((rc = ready(c)))
        return rc;
    if (W != c->W || !S)
        return fail(c, GPAIR_E_ARG, "entropy: width %d (expected %d) or NULL output", W, c->W);
    k_entropy<<<red_blocks(c->n), RED_THREADS, 0, c->stream>>>(img, c->model.p, c->n, c->part, c->ticket, c->scal);
    LAUNCHCHK(c);
    return read_scalar(c, SC_ENT, S);
}

extern "C" int gpu_get_entropy_curr(gpair_ctx *c, int W, double *S) { return get_entropy(c, c ? c->img.p : nullptr, W, S); }
extern "C" int gpu_get_entropy_temp(gpair_ctx *c, int W, double *S) { return get_entropy(c, c ? c->tmp.p : nullptr, W, S); }

static int data_gradient(gpair_ctx *c, const float *img, int nuv, int npow, int nbis, int W)
{
    int rc;
    if ((rc = ready(c)) || (rc = check_counts(c, nuv, npow, nbis)))
        return rc;
    if (W != c->W)
        return fail(c, GPAIR_E_ARG, "data gradient: width %d, expected %d", W, c->W);
    return launch_data_gradient(c, img);
}

extern "C" int gpu_compute_data_gradient_curr(gpair_ctx *c, int nuv, int npow, int nbis, int W)
{
    return data_gradient(c, c ? c->img.p : nullptr, nuv, npow, nbis, W);
}

extern "C" int gpu_compute_data_gradient_temp(gpair_ctx *c, int nuv, int npow, int nbis, int W)
{
    return data_gradient(c, c ? c->tmp.p : nullptr, nuv, npow, nbis, W);
}

static int entropy_gradient(gpair_ctx *c, const float *img, int W)
{
    int rc;
    if ((rc = ready(c)))
        return rc;
    if (W != c->W)
        return fail(c, GPAIR_E_ARG, "entropy gradient: width %d, expected %d", W, c->W);
    k_entropy_grad<<<ew_blocks(c->n), 256, 0, c->stream>>>(img, c->model.p, c->n, c->gent.p);
    LAUNCHCHK(c);
    return GPAIR_OK;
}

extern "C" int gpu_compute_entropy_gradient_curr(gpair_ctx *c, int W) { return entropy_gradient(c, c ? c->img.p : nullptr, W); }
extern "C" int gpu_compute_entropy_gradient_temp(gpair_ctx *c, int W) { return entropy_gradient(c, c ? c->tmp.p : nullptr, W); }

static int criterion_gradient(gpair_ctx *c, float *out, int W, double alpha)
{
    int rc;
    if ((rc = ready(c)))
        return rc;
    if (W != c->W)
        return fail(c, GPAIR_E_ARG, "criterion gradient: width %d, expected %d", W, c->W);
    k_crit_grad<<<ew_blocks(c->n), 256, 0, c->stream>>>(c->gdata.p, c->gent.p, (float)alpha, c->n, out);
    LAUNCHCHK(c);
    return GPAIR_OK;
}

extern "C" int gpu_compute_criterion_gradient_curr(gpair_ctx *c, int W, double alpha)
{
    return criterion_gradient(c, c ? c->gnew.p : nullptr, W, alpha);
}

extern "C" int gpu_compute_criterion_gradient_temp(gpair_ctx *c, int W, double alpha)
{
    return criterion_gradient(c, c ? c->gtmp.p : nullptr, W, alpha);
}

extern "C" int gpu_get_scalprod(gpair_ctx *c, int w, int h, gpair_buf a, gpair_buf b, double *out)
{
    int rc;
    if ((rc = ready(c)))
        return rc;
    if (!own_buf(c, a) || !own_buf(c, b) || !out || w * h != c->n)
        return fail(c, GPAIR_E_ARG, "gpu_get_scalprod: bad buffer, size %dx%d or NULL output", w, h);
    k_dot<<<red_blocks(c->n), RED_THREADS, 0, c->stream>>>(a->p, b->p, c->n, c->part, c->ticket, c->scal + SC_DOT);
    LAUNCHCHK(c);
    return read_scalar(c, SC_DOT, out);
}

extern "C" int gpu_compute_descent_dir(gpair_ctx *c, int W, double beta)
{
    int rc;
    if ((rc = ready(c)))
        return rc;
    if (W != c->W)
        return fail(c, GPAIR_E_ARG, "descent dir: width %d, expected %d", W, c->W);
    k_descent<<<ew_blocks(c->n), 256, 0, c->stream>>>(c->dir.p, c->gnew.p, (float)beta, c->n);
    LAUNCHCHK(c);
    return GPAIR_OK;
}

static int step(gpair_ctx *c, float *out, int W, double s, double minval, gpair_buf d)
{
    int rc;
    if ((rc = ready(c)))
        return rc;
    if (W != c->W || !own_buf(c, d))
        return fail(c, GPAIR_E_ARG, "image update: width %d or direction buffer invalid", W);
    if (!isfinite(s))
        return fail(c, GPAIR_E_NUMERIC, "image update: non-finite step %g", s);
    k_step<<<ew_blocks(c->n), 256, 0, c->stream>>>(c->img.p, d->p, (float)s, (float)minval, c->n, out);
    LAUNCHCHK(c);
    return GPAIR_OK;
}

extern "C" int gpu_update_tempimage(gpair_ctx *c, int W, double s, double minval, gpair_buf d)
{
    return step(c, c ? c->tmp.p : nullptr, W, s, minval, d);
}

extern "C" int gpu_update_image(gpair_ctx *c, int W, double s, double minval, gpair_buf d)
{
    px_invalidate(c);
    return step(c, c ? c->img.p : nullptr, W, s, minval, d);
}

extern "C" int gpu_backup_gradient(gpair_ctx *c, int n, gpair_buf in, gpair_buf out)
{
    if (!c)
        return GPAIR_E_ARG;
    if (!own_buf(c, in) || !own_buf(c, out) || n < 0 || n > in->n || n > out->n)
        return fail(c, GPAIR_E_ARG, "gpu_backup_gradient: bad buffers or size %d", n);
    if (out == &c->img)
        px_invalidate(c);
    HIPCHK(c, hipMemcpyAsync(out->p, in->p, sizeof(float) * n, hipMemcpyDeviceToDevice, c->stream));
    return GPAIR_OK;
}

extern "C" int gpu_get_image(gpair_ctx *c, int size, float *host, gpair_buf which)
{
    if (!c || !host)
        return fail(c, GPAIR_E_ARG, "gpu_get_image: NULL pointer");
    if (!own_buf(c, which) || size < 0 || size > which->n)
        return fail(c, GPAIR_E_ARG, "gpu_get_image: bad buffer or size %d", size);
    HIPCHK(c, hipMemcpyAsync(host, which->p, sizeof(float) * size, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GPAIR_OK;
}

extern "C" int gpair_put_image(gpair_ctx *c, gpair_buf which, const float *host, int size)
{
    if (!c || !host)
        return fail(c, GPAIR_E_ARG, "gpair_put_image: NULL pointer");
    if (!own_buf(c, which) || size < 0 || size > which->n)
        return fail(c, GPAIR_E_ARG, "gpair_put_image: bad buffer or size %d", size);
    if (which == &c->img)
        px_invalidate(c);
    HIPCHK(c, hipMemcpyAsync(which->p, host, sizeof(float) * size, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GPAIR_OK;
}

extern "C" int gpair_get_vis(gpair_ctx *c, float *vis, int nuv)
{
    if (!c || !vis || nuv != c->nuv)
        return fail(c, GPAIR_E_ARG, "gpair_get_vis: NULL pointer or nuv mismatch");
    HIPCHK(c, hipMemcpyAsync(vis, c->vis, sizeof(float2) * nuv, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GPAIR_OK;
}

extern "C" int gpair_get_mock(gpair_ctx *c, float *mock, int ndata)
{
    if (!c || !mock || ndata != c->ndof)
        return fail(c, GPAIR_E_ARG, "gpair_get_mock: NULL pointer or size mismatch");
    HIPCHK(c, hipMemcpyAsync(mock, c->mock, sizeof(float) * ndata, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GPAIR_OK;
}

// ---------------------------------------------------------------------------
// gpu_linesearch_zoom, gpair1/trunk/gpu.c:1947-2065, with FP64 scalars
// ---------------------------------------------------------------------------
extern "C" int gpu_linesearch_zoom(gpair_ctx *c, int nuv, int npow, int nbis, int data_alloc, int data_alloc_uv,
                                   int W, double lo, double hi, double J_lo, double w1, double J_init,
                                   int *criterion_evals, int *grad_evals, gpair_buf d, gpair_buf tg, double alpha,
                                   int max_evals, double *selected)
{
    const double minv = 1e-8, c1 = 1e-4, c2 = 0.1;
    int rc;
    if ((rc = ready(c)))
        return rc;
    if (!selected || !own_buf(c, d) || !own_buf(c, tg))
        return fail(c, GPAIR_E_ARG, "gpu_linesearch_zoom: bad buffers or NULL output");
    double s = 0.0, s_old = 0.0, J_old = 0.0;
    int counter = 0, evals = 0;
    for (;;) {
        if (counter > 0 && (J_old - J_init - w1 * s_old) != 0.0)
            s = fabs(w1 * s_old * s_old / (2.0 * (J_old - J_init - w1 * s_old)));
        if (counter == 0 || s < lo || s > hi)
            s = (hi - lo) / 2.0 + lo;
        if (fabs(hi - lo) < 1e-14 || (max_evals >= 0 && evals >= max_evals)) {
            *selected = lo;
            return GPAIR_OK;
        }
        double chi2, S;
        if ((rc = gpu_update_tempimage(c, W, s, minv, d)) ||
            (rc = gpu_get_chi2_temp(c, nuv, npow, nbis, data_alloc, data_alloc_uv, &chi2)) ||
            (rc = gpu_get_entropy_temp(c, W, &S)))
            return rc;
        const double J = chi2 - alpha * S;
        evals++;
        if (criterion_evals)
            (*criterion_evals)++;
        if (J > J_init + c1 * s * w1 || J >= J_lo) {
            hi = s;
        } else {
            double w2;
            if ((rc = gpu_compute_data_gradient_temp(c, nuv, npow, nbis, W)) ||
                (rc = gpu_compute_entropy_gradient_temp(c, W)) ||
                (rc = gpu_compute_criterion_gradient_temp(c, W, alpha)) ||
                (rc = gpu_get_scalprod(c, W, W, d, tg, &w2)))
                return rc;
            if (grad_evals)
                (*grad_evals)++;
            if (w2 >= c2 * w1 || counter > 10) {
                *selected = s;
                return GPAIR_OK;
            }
            if (w2 * (hi - lo) >= 0.0)
                hi = lo;
            lo = s;
        }
        s_old = s;
        J_old = J;
        counter++;
    }
}

// ---------------------------------------------------------------------------
// accelerated loop: preconditioned PR-CG with the batched line search
// (declared in include/gpair_hip.h; restated in FP64 by the oracle's orc_pcg)
// ---------------------------------------------------------------------------
extern "C" void gpair_pcg_default_params(gpair_pcg_params *p)
{
    memset(p, 0, sizeof(*p));
    p->alpha = 1.0;
    p->precond = 1;
    p->max_iter = 1000;
    p->chi2r_target = 1.0;
    p->minvalue = 1e-8;
    p->wolfe1 = 1e-4;
    p->verbose = 0;
}

extern "C" int gpair_set_sparse_commit(gpair_ctx *c, int max_pixels)
{
    if (!c)
        return GPAIR_E_ARG;
    if (max_pixels < -1)
        return fail(c, GPAIR_E_ARG, "gpair_set_sparse_commit: max_pixels %d (-1, 0 or a count)", max_pixels);
    c->commit_cap = max_pixels;
    return GPAIR_OK;
}

extern "C" int gpair_set_adjoint_tile(gpair_ctx *c, int rows, int cols, int workgroups)
{
    if (!c)
        return GPAIR_E_ARG;
    if (!((rows == 256 && cols == 256) || (rows == 128 && cols == 256) || (rows == 128 && cols == 128)))
        return fail(c, GPAIR_E_ARG, "gpair_set_adjoint_tile: %d x %d (256 x 256, 128 x 256 or 128 x 128)", rows,
                    cols);
    if (workgroups < 0 || workgroups > 65536)
        return fail(c, GPAIR_E_ARG, "gpair_set_adjoint_tile: %d workgroups (0: the default, up to 65536)",
                    workgroups);
    c->adj_tr = rows;
    c->adj_tc = cols;
    c->adj_groups = workgroups ? workgroups : ADJ_WIDE_GROUPS;
    c->adj_b3_nsplit = 0;   // the split follows the tile: set again on the next split-precision adjoint
    c->split_valid = 0;
    return GPAIR_OK;
}

extern "C" int gpair_set_pfactor(gpair_ctx *c, const float *pf, int size)
{
    if (!c)
        return GPAIR_E_ARG;
    if (!c->have_data)
        return fail(c, GPAIR_E_STATE, "gpair_set_pfactor before gpu_copy_data");
    if (!pf) {
        dfree(&c->pfac);
        return GPAIR_OK;
    }
    if (size != c->n)
        return fail(c, GPAIR_E_ARG, "gpair_set_pfactor: %d values, image has %d pixels", size, c->n);
    int rc;
    if ((rc = dalloc(c, &c->pfac, c->n)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->pfac, pf, sizeof(float) * c->n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GPAIR_OK;
}

static int read_scalars(gpair_ctx *c, int slot, int count, double *out)
{
    HIPCHK(c, hipMemcpyAsync(c->host_scal + slot, c->scal + slot, sizeof(double) * count, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < count; i++)
        out[i] = c->host_scal[slot + i];
    return GPAIR_OK;
}

// chi2 and entropy of an image (leaves vis / mock of that image)
static int exact_eval(gpair_ctx *c, const float *img, double *chi2, double *S)
{
    int rc;
    if ((rc = launch_chi2(c, img)))
        return rc;
    k_entropy<<<red_blocks(c->n), RED_THREADS, 0, c->stream>>>(img, c->model.p, c->n, c->part, c->ticket, c->scal);
    LAUNCHCHK(c);
    double v[2];
    if ((rc = read_scalars(c, SC_CHI2, 2, v)))
        return rc;
    *chi2 = v[0];
    *S = v[1];
    return GPAIR_OK;
}

template <typename T>
static inline void swap_ptr(T *&a, T *&b)
{
    T *t = a;
    a = b;
    b = t;
}
// ---------------------------------------------------------------------------
// snapshots of the current image (the reference's periodic reconst.fits)
// ---------------------------------------------------------------------------
extern "C" int gpair_set_snapshot(gpair_ctx *c, const char *path, int every, float pixsize)
{
    if (!c)
        return GPAIR_E_ARG;
    if (!path || every <= 0) {
        c->snap_every = 0;
        c->snap_path[0] = '\0';
        return GPAIR_OK;
    }
    if (strlen(path) >= sizeof(c->snap_path))
        return fail(c, GPAIR_E_ARG, "gpair_set_snapshot: path longer than %zu", sizeof(c->snap_path) - 1);
    strcpy(c->snap_path, path);
    c->snap_every = every;
    c->snap_pix = pixsize;
    return GPAIR_OK;
}

extern "C" int gpair_snapshot_due(const gpair_ctx *c, int it) { return c && c->snap_every > 0 && it % c->snap_every == 0; }

extern "C" int gpair_snapshot_write(gpair_ctx *c)
{
    if (!c || !c->img.p)
        return GPAIR_E_ARG;
    std::vector<float> host(c->n);
    HIPCHK(c, hipMemcpyAsync(host.data(), c->img.p, sizeof(float) * c->n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int rc = gpair_fits_write_image(c->snap_path, host.data(), c->W, c->W, -64, c->snap_pix, nullptr, nullptr);
    return rc ? fail(c, rc, "snapshot: %s", gpair_io_last_error()) : GPAIR_OK;
}

#include "gpair_pcg.inc"
#include "gpair_group.inc"
#include "gpair_tune.inc"
#include "gpair_pixel.inc"
#include "gpair_noalloc.inc"
