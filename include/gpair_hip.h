/*
 * gpair_hip.h -- C ABI of the MI355X (gfx950) MEM-CG reconstruction path.
 *
 * The entry points keep the names and meaning of GPAIR's OpenCL host layer,
 * gpair1/trunk/gpu.h:1-120 / gpair1/trunk/gpu.c, so that the CG loop of
 * gpair1/trunk/gpair.c:567-922 drops in.  Differences, all deliberate:
 *   - every call takes an explicit gpair_ctx* (the reference keeps file-static
 *     globals, gpair1/trunk/gpu.c:22-115); contexts are independent;
 *   - every call returns an int status (GPAIR_OK == 0) instead of exit(0)
 *     on error (gpair1/trunk/gpu.c:186-202); scalars come back through double*;
 *   - scalars are FP64 (the reference returns float);
 *   - device buffers are opaque gpair_buf handles instead of cl_mem*;
 *   - sizes are exact counts (npow + 2*nbis, nbis, nuv), not the pow2-padded
 *     allocation sizes the OpenCL kernels needed (gpair1/trunk/gpair.c:166-172);
 *   - the image width must be a multiple of 64 (64..1024).
 * Threading: one HIP stream per context, calls on one context are serialised
 * by the caller (as in the reference); different contexts may be used from
 * different host threads.
 */
#ifndef GPAIR_HIP_H
#define GPAIR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gpair_ctx gpair_ctx;
typedef struct gpair_buf_s *gpair_buf;

enum {
    GPAIR_OK = 0,
    GPAIR_E_ARG = 1,      /* invalid argument (sizes, indices, NULL) */
    GPAIR_E_HIP = 2,      /* HIP runtime error */
    GPAIR_E_STATE = 3,    /* call out of order (e.g. chi2 before data) */
    GPAIR_E_NOMEM = 4,    /* device allocation failed */
    GPAIR_E_NUMERIC = 5   /* non-finite objective or step */
};

/* precision of the forward / adjoint DFT contraction */
enum {
    GPAIR_PREC_FP32 = 0,   /* FP32 MFMA, FP32 accumulation */
    GPAIR_PREC_BF16X3 = 1, /* split BF16 hi/lo, 3 MFMA products, FP32 accumulation */
    GPAIR_PREC_FP16S = 2   /* FP16 with power-of-two scaling, 1 MFMA product, FP32 accumulation;
                              needs DFT tables set from uv (gpair_set_uv) */
};

const char *gpair_strerror(int code);
const char *gpair_last_error(const gpair_ctx *ctx);

/* ---- setup / teardown ------------------------------------------------- */
/* gpu_init, gpair1/trunk/gpu.c:1693-1732 (first GPU, in-order queue) */
int gpu_init(gpair_ctx **ctx, int device);
/* gpu_copy_data, gpair1/trunk/gpu.c:1175-1415: data/err [npow + 2 nbis] packed as
 * init_data does (gpair1/trunk/gpair.c:997-1066); phasor = (re, im) float pairs
 * [nbis]; uvpnt = 4 x int64 per T3 (cl_long4, gpair1/trunk/gpair.c:585-592);
 * sign = 4 x int16 per T3 (cl_short4); default_model [image_width^2].
 * data_size = npow + 2*bsref_size, data_size_uv = nuv, phasor_size = bsref_size = nbis.
 * Either of npow and nbis may be 0 (V2 only, T3 only), not both (GPAIR_E_ARG). */
int gpu_copy_data(gpair_ctx *ctx, const float *data, const float *data_err, int data_size, int data_size_uv,
                  const float *phasor, int phasor_size, int pow_size, const int64_t *uvpnt, const int16_t *sign,
                  int bsref_size, const float *default_model, int image_size, int image_width);
/* gpu_copy_image, gpair1/trunk/gpu.c:1441-1451: the current image, [y][x] */
int gpu_copy_image(gpair_ctx *ctx, const float *image, int x_size, int y_size);
/* gpu_copy_dft, gpair1/trunk/gpair.c:653: tables in the reference's GPU layout
 * float2 [W][nuv] (gpair1/trunk/gpair.c:610-636); dft_size >= W*nuv. */
int gpu_copy_dft(gpair_ctx *ctx, const float *dft_x, const float *dft_y, int dft_size);
/* No reference counterpart (the reference builds tables on the host,
 * gpair1/trunk/gpair.c:231-245): builds X[p][x] = exp(+2 pi i a u_p x) and
 * Y[p][y] = exp(-2 pi i a v_p y), a = RPMAS*pixellation, on the device with the
 * phase range-reduced in FP64.  u, v in wavelengths, [nuv]. */
int gpair_set_uv(gpair_ctx *ctx, const float *u, const float *v, int nuv, float pixellation);
/* gpu_build_kernels, gpair1/trunk/gpu.c:306-411: kernels are compiled ahead of
 * time here; this validates the sizes and allocates the work buffers. */
int gpu_build_kernels(gpair_ctx *ctx, int data_size, int image_width, int image_size);
/* gpu_cleanup / gpu_shutdown, gpair1/trunk/gpu.c:887-960: free everything */
int gpu_cleanup(gpair_ctx *ctx);
int gpu_shutdown(gpair_ctx *ctx);

/* select the DFT precision (default GPAIR_PREC_FP32); no reference counterpart */
int gpair_set_precision(gpair_ctx *ctx, int precision);

/* ---- buffer handles, gpair1/trunk/gpu.c:1829-1867 -------------------- */
gpair_buf gpu_getp_ci(gpair_ctx *ctx);   /* current image */
gpair_buf gpu_getp_ti(gpair_ctx *ctx);   /* temporary (trial) image */
gpair_buf gpu_getp_fgn(gpair_ctx *ctx);  /* full gradient, new */
gpair_buf gpu_getp_fg(gpair_ctx *ctx);   /* full gradient, previous */
gpair_buf gpu_getp_dd(gpair_ctx *ctx);   /* descent direction */
gpair_buf gpu_getp_tg(gpair_ctx *ctx);   /* gradient at the trial image */
gpair_buf gpu_getp_eg(gpair_ctx *ctx);   /* entropy gradient */
gpair_buf gpair_getp_dg(gpair_ctx *ctx); /* chi2 (data) gradient; no reference getter */

/* ---- objective -------------------------------------------------------- */
/* gpu_get_chi2_curr|temp, gpair1/trunk/gpu.c:1734-1763: flux, DFT, V2/T3 model,
 * chi2.  Leaves visibilities and mock data for the next gradient call. */
int gpu_get_chi2_curr(gpair_ctx *ctx, int nuv, int npow, int nbis, int data_alloc, int data_alloc_uv,
                      double *chi2);
int gpu_get_chi2_temp(gpair_ctx *ctx, int nuv, int npow, int nbis, int data_alloc, int data_alloc_uv,
                      double *chi2);
/* gpu_get_entropy_curr|temp, gpair1/trunk/gpu.c:1767-1802 */
int gpu_get_entropy_curr(gpair_ctx *ctx, int image_width, double *entropy);
int gpu_get_entropy_temp(gpair_ctx *ctx, int image_width, double *entropy);

/* ---- gradients -------------------------------------------------------- */
/* gpu_compute_data_gradient_curr|temp, gpair1/trunk/gpu.c:1598-1691 (adjoint DFT);
 * uses the visibilities / mock data of the last chi2 call, as the reference does. */
int gpu_compute_data_gradient_curr(gpair_ctx *ctx, int nuv, int npow, int nbis, int image_width);
int gpu_compute_data_gradient_temp(gpair_ctx *ctx, int nuv, int npow, int nbis, int image_width);
/* gpu_compute_entropy_gradient_curr|temp, gpair1/trunk/gpu.c:796-857 */
int gpu_compute_entropy_gradient_curr(gpair_ctx *ctx, int image_width);
int gpu_compute_entropy_gradient_temp(gpair_ctx *ctx, int image_width);
/* gpu_compute_criterion_gradient_curr|temp, gpair1/trunk/gpu.c:615-666:
 * fgn (curr) or tg (temp) = data gradient - alpha * entropy gradient */
int gpu_compute_criterion_gradient_curr(gpair_ctx *ctx, int image_width, double alpha);
int gpu_compute_criterion_gradient_temp(gpair_ctx *ctx, int image_width, double alpha);

/* ---- CG --------------------------------------------------------------- */
/* gpu_get_scalprod, gpair1/trunk/gpu.c:2114-2129 (FP64 accumulation) */
int gpu_get_scalprod(gpair_ctx *ctx, int data_width, int data_height, gpair_buf a, gpair_buf b, double *out);
/* gpu_compute_descent_dir, gpair1/trunk/gpu.c:705-739: dd = beta*dd - fgn */
int gpu_compute_descent_dir(gpair_ctx *ctx, int image_width, double beta);
/* gpu_update_tempimage, gpair1/trunk/gpu.c:2300-2339: ti = max(ci + s*d, minval) */
int gpu_update_tempimage(gpair_ctx *ctx, int image_width, double steplength, double minval, gpair_buf d);
/* gpu_update_image, gpair1/trunk/gpu.c:2260-2298: ci = max(ci + s*d, minval) */
int gpu_update_image(gpair_ctx *ctx, int image_width, double steplength, double minval, gpair_buf d);
/* gpu_backup_gradient, gpair1/trunk/gpu.c:257-268: out <- in (data_size floats) */
int gpu_backup_gradient(gpair_ctx *ctx, int data_size, gpair_buf in, gpair_buf out);
/* gpu_linesearch_zoom, gpair1/trunk/gpu.c:1947-2065.  The evaluation counters
 * are incremented (the reference increments the pointers instead,
 * gpair1/trunk/gpu.c:2004,2033).  max_evals bounds the J evaluations (a
 * build-side safeguard; < 0 means unbounded, as in the reference); when it is
 * reached the current low end of the interval is returned. */
int gpu_linesearch_zoom(gpair_ctx *ctx, int nuv, int npow, int nbis, int data_alloc, int data_alloc_uv,
                        int image_width, double steplength_low, double steplength_high,
                        double criterion_steplength_low, double wolfe_product1, double criterion_init,
                        int *criterion_evals, int *grad_evals, gpair_buf descent_direction,
                        gpair_buf temp_gradient, double hyperparameter_entropy, int max_evals,
                        double *selected_steplength);

/* ---- output ----------------------------------------------------------- */
/* gpu_get_image, gpair1/trunk/gpu.c:1815-1827: blocking device -> host copy */
int gpu_get_image(gpair_ctx *ctx, int size, float *cpu_buffer, gpair_buf which);
/* copies of the last visibilities (float2 [nuv]) / mock data ([npow+2nbis]);
 * no reference getter (the reference compares them in gpu_check_data) */
int gpair_get_vis(gpair_ctx *ctx, float *vis, int nuv);
int gpair_get_mock(gpair_ctx *ctx, float *mock, int ndata);
/* host -> device copy into any image-sized buffer (tests, warm starts) */
int gpair_put_image(gpair_ctx *ctx, gpair_buf which, const float *host, int size);

/* ---- the MEM-CG loop (gpair1/trunk/gpair.c:682-915), C host code -------- */
typedef struct {
    double alpha;          /* entropy hyperparameter (gpair.c -a) */
    int method;            /* 0 SD, 1 FR, 2 PR, 3 HS (gpair1/trunk/gpair.c:727-758) */
    int max_iter;          /* outer iterations */
    double chi2r_target;   /* stop when chi2/ndof <= target (0 disables) */
    int ls_max_evals;      /* J evaluations per line search (<= 0: unbounded) */
    double minvalue;       /* pixel floor, 1e-8 in the reference */
    double wolfe1, wolfe2; /* 1e-4, 0.1 */
    double step_max;       /* 100 */
    int verbose;
} gpair_cg_params;

typedef struct {
    int iterations;        /* outer iterations whose J was evaluated */
    int criterion_evals;
    int grad_evals;
    double J, chi2, entropy, chi2r;
    int converged;         /* chi2r <= target reached */
    int status;
} gpair_cg_result;

void gpair_cg_default_params(gpair_cg_params *p);
/* Runs the loop on the current image of ctx.  trace (may be NULL) receives 8
 * doubles per outer iteration: J, chi2, S, selected step, beta, J evals,
 * gradient evals, 0. */
int gpair_cg_run(gpair_ctx *ctx, const gpair_cg_params *params, gpair_cg_result *result, double *trace,
                 int trace_rows);

/* ---- measurement (no reference counterpart; the reference queue has no
 * profiling, gpair1/trunk/gpu.c:1724) ------------------------------------ */
enum {
    GPAIR_TK_FORWARD = 0,
    GPAIR_TK_ADJOINT = 1,
    GPAIR_TK_PX_DELTA = 2,   /* per-move changes of U (gpair_pixel_propose) */
    GPAIR_TK_PX_DCHI2 = 3,   /* per-move chi2 changes (gpair_pixel_propose) */
    GPAIR_TK_COUNT = 4
};
/* start (on > 0) or stop (on <= 0) recording HIP events around the launches of
 * the kernels above on the context stream: every launch for on == 1, every on-th
 * launch of each kernel otherwise (each recorded pair adds work to the stream);
 * enabling clears what was recorded */
int gpair_timing_enable(gpair_ctx *ctx, int on);
/* number of timed launches of a kernel and their summed duration (ms) */
int gpair_timing_read(gpair_ctx *ctx, int kernel, int *count, double *total_ms);

/* ---- accelerated loop (no reference counterpart) ---------------------------
 * Preconditioned Polak-Ribiere CG: z = P o g with P = image (MEM metric, when
 * precond != 0) times an optional per-pixel factor (gpair_set_pfactor), the
 * direction projected to zero on floor pixels moving down, and a batched line
 * search: with U = A I and A d (A = unnormalised DFT, one extra forward DFT of d)
 * the visibilities of I + s d are (U_I + s U_d) / (F_I + s F_d), so 32 steps
 * s_ref 2^((k-26)/2) are scored in one launch; a parabola through the best three
 * gives s, committed with an exact evaluation that must pass the Armijo test.
 * Per iteration: 2 forward DFTs, 1 adjoint DFT, about 4 scalar read-backs. */
typedef struct {
    double alpha;
    int precond;           /* 0: P = 1, 1: P = image */
    int max_iter;
    double chi2r_target;
    double minvalue;       /* 1e-8 */
    double wolfe1;         /* Armijo constant, 1e-4 */
    int verbose;
    double J_target;       /* > 0: the target also needs J <= J_target (a run judged
                              against another run's objective); 0: chi2/ndof only */
} gpair_pcg_params;

typedef struct {
    int iterations;        /* trace rows (objective evaluations of accepted images) */
    int forward_dfts, adjoint_dfts, ls_batches;
    double J, chi2, entropy, chi2r;
    int converged;
    int stalled;           /* no decrease found along the direction */
    int status;
    int resyncs;           /* exact re-evaluations of the tracked objective (their DFTs are not in forward_dfts) */
    int sparse_commits;    /* commits whose change DFT was s U_d plus the correction over the clamped pixels
                              (counted in forward_dfts: full forward DFTs = forward_dfts - sparse_commits) */
} gpair_pcg_result;

void gpair_pcg_default_params(gpair_pcg_params *p);
/* Runs the accelerated loop on the current image.  trace: 8 doubles per row:
 * J, chi2, S, step, beta, forward DFTs, adjoint DFTs, line-search batches
 * (the last three cumulative). */
int gpair_pcg_run(gpair_ctx *ctx, const gpair_pcg_params *params, gpair_pcg_result *result, double *trace,
                  int trace_rows);
/* No reference counterpart.  A commit's trial image is max(I + s d, floor), so its
 * change is s d plus a correction c_j at the pixels the floor clamps; its DFT is then
 * taken as s U_d (the direction's DFT, at hand) plus sum_j c_j X[p][x_j] Y[p][y_j]
 * when at most max_pixels pixels clamp (tables set from uv only), else by a full
 * forward DFT (the choice is made on the device).  -1 (default): W*W/40, about
 * where the correction costs as much as the full DFT; 0: always the full DFT. */
int gpair_set_sparse_commit(gpair_ctx *ctx, int max_pixels);
/* No reference counterpart.  Output tile of the split-precision adjoint on factored
 * tables: rows x cols = 256 x 256, 128 x 256 or 128 x 128 per workgroup where the
 * tile divides W (otherwise 128 x 128 when that divides W, else 64 x 64), with the
 * uv points split (at least 64 points per split) so that about `workgroups` workgroups run
 * (0: 256).  Smaller tiles mean fewer uv splits for as many workgroups, and so
 * fewer partial images to add up, at more operand building per product.  The
 * split fixes the summation order: results are reproducible for a given setting,
 * and differ between settings by rounding. */
int gpair_set_adjoint_tile(gpair_ctx *ctx, int rows, int cols, int workgroups);
/* per-pixel preconditioner factor [W*W] multiplying P (NULL clears it) */
int gpair_set_pfactor(gpair_ctx *ctx, const float *pfactor, int size);
/* on != 0: gpair_pcg_run replays each iteration's gradient phase (flux, adjoint
 * coefficients, adjoint DFT, entropy and criterion gradients, preconditioner)
 * from a hipGraph captured at the start of the run; results are bit-identical
 * to the launched path.  Off by default; ignored while DFT timing is on. */
int gpair_set_graphs(gpair_ctx *ctx, int on);
/* Batch (no reference counterpart; it replaces the one-trial-step-at-a-time loop
 * of gpair1/trunk/gpair.c:808-906 run once per problem): runs the accelerated
 * loop on `count` independent contexts with params[i] and results[i] per
 * context.  When every context is BF16x3 on tables set from uv (gpair_set_uv),
 * of a width that is a multiple of 64 up to 512, without snapshots or DFT
 * timing, and count <= 1024, one launch sequence per iteration serves all
 * problems: each kernel takes the blocks of every problem in one launch, the
 * loop's decisions (step choice, Armijo test, restarts, re-evaluations) run on
 * the device, and the sequence is replayed from one hipGraph.  Otherwise each
 * context runs on its own stream (gpair_batch_pcg_run_streams).  Either way every
 * problem follows exactly the steps gpair_pcg_run would take (bit-identical
 * results) and the contexts end as gpair_pcg_run leaves them.  Returns the first
 * non-zero status, GPAIR_OK when all succeed. */
int gpair_batch_pcg_run(gpair_ctx *const *ctxs, int count, const gpair_pcg_params *params,
                        gpair_pcg_result *results);
/* The same with one stream per context: the host resumes whichever problem has
 * finished its read-back, so the GPU works on the others meanwhile. */
int gpair_batch_pcg_run_streams(gpair_ctx *const *ctxs, int count, const gpair_pcg_params *params,
                                gpair_pcg_result *results);

/* No reference counterpart (the reference restarts from scratch): the
 * accelerated loop with a checkpoint.  resume_from (NULL: start from the current
 * image as gpair_pcg_run) is a file written by save_to of an earlier run on a
 * context with the same data, sizes, precision, alpha and preconditioner; the
 * run then continues at the saved iteration, up to params->max_iter iterations
 * in total, exactly as the uninterrupted run would (bit for bit).  save_to (NULL:
 * none) receives the state when the run ends at an iteration boundary (iteration
 * cap or target reached): a FITS file whose primary image is the current image,
 * with the direction, previous gradient, FP64 and float visibilities, mock
 * data, preconditioner factor and loop scalars as IMAGE extensions.  Trace rows
 * are indexed by iteration number. */
int gpair_pcg_run_checkpointed(gpair_ctx *ctx, const gpair_pcg_params *params, gpair_pcg_result *result,
                               double *trace, int trace_rows, const char *resume_from, const char *save_to);
/* The reference writes the current image to reconst.fits every 5 iterations
 * (gpair1/trunk/gpair.c:698-702, writefits at gpair.c:954-990).  With path set
 * and every > 0, gpair_cg_run and the accelerated loops write the current image
 * (FITS, BITPIX -64, PIXSIZE = pixsize mas) at every objective evaluation of an
 * accepted image whose index is a multiple of `every`; NULL or every <= 0
 * turns it off.  Each snapshot reads the image back and writes the file. */
int gpair_set_snapshot(gpair_ctx *ctx, const char *path, int every, float pixsize);

/* Autotuner (no reference counterpart: the reference fixes its work-group sizes
 * when it builds the kernels, gpair1/trunk/gpu.c:413-498).  Times candidate
 * launch shapes of the size-dependent reductions on the context's data and
 * current image -- the grids of the batched line-search kernels and of the
 * adjoint coefficients, the row grouping of the BF16x3 forward DFT -- and keeps
 * the fastest for the context (reps launches per candidate, <= 0: 5).  No
 * summation order depends on these shapes, so tuning changes no result.
 * chosen, if not NULL, receives up to 4 values: line-search chi2 grid,
 * line-search entropy grid, coefficient grid, forward rows per group / 32.
 * Call it after gpu_build_kernels. */
int gpair_autotune(gpair_ctx *ctx, int reps, int *chosen, int nchosen);

/* ---- single-pixel moves: incremental visibility updates ------------------
 * For MCMC / annealing-style samplers.  With U = F V the unnormalised
 * visibilities of an image of total flux F, changing pixel (x, y) by d gives
 *   V'_p = (U_p + d X[p][x] Y[p][y]) / (F + d),
 * the identity behind update_vis_fluxchange (gpair1/trunk/cpu.c:477-486), whose
 * V' = r V + (1 - r) X Y holds with r = F / (F + d), the ratio of the TOTAL
 * fluxes before and after.  The pixel state is an image (a copy of the current
 * image at gpair_pixel_begin, plus the moves applied since), its U in FP64 and
 * its flux in FP64; proposals are scored against it exactly, in FP64
 * differences, without a DFT.  Pixels are indexed pix = x + W*y. */
typedef struct {
    int pix_a;             /* first pixel, x + W*y */
    int pix_b;             /* second pixel, or -1 for a one-pixel move (must differ from pix_a) */
    float flux_a, flux_b;  /* their new fluxes */
} gpair_pixel_move;

/* no reference counterpart: (re)builds the pixel state from the current image
 * (one forward DFT); any later change of the current image by other calls is not
 * seen until the next gpair_pixel_begin.  The calls below build it when absent;
 * gpu_copy_image, gpu_update_image, the CG loops and new tables drop it. */
int gpair_pixel_begin(gpair_ctx *ctx);
/* gpu_update_vis_fluxchange, gpair1/trunk/gpu.c:2137-2182 (kernel
 * gpair1/branches/cpu_dev/kernel_vis_update_fluxchange.cl:1-40): the trial
 * visibilities (the reference's visi1) of the pixel state with pixel (x, y) set
 * to new_pixel_flux; the state is not changed.  Differs from the reference as
 * the identity above requires: r is the ratio of total fluxes (the reference
 * divides pixel fluxes), the y table is indexed by y (the reference kernel uses
 * x_pos for both), and the pixel is read on the device (the reference reads it
 * back with a byte offset equal to the pixel index).  data_alloc_uv is unused. */
int gpu_update_vis_fluxchange(gpair_ctx *ctx, int x, int y, float new_pixel_flux, int image_width, int nuv,
                              int data_alloc_uv);
/* update_vis_positionchange, gpair1/trunk/cpu.c:488-497: trial visibilities with
 * flux moved from (x_old, y_old) to (x_new, y_new); the total flux is unchanged,
 * so V' = V + (flux / F) (X_new Y_new - X_old Y_old) */
int gpair_update_vis_positionchange(gpair_ctx *ctx, int x_old, int y_old, int x_new, int y_new, float flux,
                                    int image_width, int nuv);
/* gpu_new_chi2, gpair1/trunk/gpu.c:2067-2071: chi2 of the trial visibilities
 * (vis2data + data2chi2 on visi1); its mock data stay apart from the current
 * image's.  data_alloc is unused. */
int gpu_new_chi2(gpair_ctx *ctx, int nuv, int npow, int nbis, int data_alloc, double *chi2);
/* copy of the trial visibilities, float2 [nuv] */
int gpair_get_trial_vis(gpair_ctx *ctx, float *vis, int nuv);
/* no reference counterpart: scores `count` (<= 4096) independent moves against
 * the pixel state in a few launches: dchi2[k] = chi2(state after move k) -
 * chi2(state), dentropy[k] = S(after) - S(before), both FP64 (either may be NULL).
 * A sampler's dJ is dchi2 - alpha dentropy. */
int gpair_pixel_propose(gpair_ctx *ctx, const gpair_pixel_move *moves, int count, double *dchi2,
                        double *dentropy);
/* no reference counterpart: applies `count` (<= 4096) moves in order to the
 * pixel state and writes their pixels into the current image too */
int gpair_pixel_apply(gpair_ctx *ctx, const gpair_pixel_move *moves, int count);
/* no reference counterpart: chi2 (from the FP64 state visibilities), entropy and
 * flux of the pixel state */
int gpair_pixel_state(gpair_ctx *ctx, double *chi2, double *entropy, double *flux);

/* No reference counterpart: for callers that cannot allocate device memory
 * (harnesses that forbid hipMalloc, tight serving loops).  chi2 of `image` (an
 * image buffer of the context, e.g. gpu_getp_ci) as gpu_get_chi2_curr computes it
 * -- flux, forward DFT (FP32, or BF16x3 on tables set with gpair_set_uv), V2/T3
 * model, chi2, the same kernels and launch shapes, so the same value -- but every
 * intermediate (visibilities, mock data, split image, forward partials, reduction
 * partials, tickets, scalars) lives in the library's static device scratch
 * (11.3 MiB per device, held until the process ends) instead of context buffers:
 * the call makes no hipMalloc / hipFree.  Limits: nuv <= 262144, npow + 2 nbis <=
 * 1048576, W <= 512; larger problems and FP16 DFTs get GPAIR_E_ARG.  Unlike
 * gpu_get_chi2_curr it leaves the context's visibilities and mock data as they
 * were.  The scratch is shared by every call on one device and holds the last
 * call's contents: calls that overlap on one device (other contexts, other host
 * threads) race. */
int gpair_chi2_noalloc(gpair_ctx *ctx, gpair_buf image, double *chi2);

/* ---- host-side data preparation (no device needed) ----------------------- */
/* set_model, gpair1/trunk/cpu.c:645-749: prior image [W*W] of total flux `flux`.
 * type 0 flat, 1 centred Dirac, 2 uniform disk (radius width mas), 3 Gaussian
 * (FWHM width mas), 4 Lorentzian.  The reference draws the Gaussian 80 px off
 * centre in y (cpu.c:714-715); centered != 0 removes that offset. */
int gpair_set_model(float *model, int W, float pixellation, int type, float width, float flux, int centered);
/* init_data, gpair1/trunk/gpair.c:997-1066: packs data [npow + 2 nbis],
 * inverse errors and phasors (re, im per T3) from V2 / T3 arrays; uvpnt holds 3
 * uv indices per T3.  With do_extrapolation == 1, triple amplitudes whose error
 * is <= 0 or > 1e3 are rebuilt from the three V2 points (bisamp / bisamperr are
 * updated in place, as the reference updates its oi_data); *ndof is decremented
 * for each T3 that cannot be rebuilt. */
int gpair_init_data(int npow, int nbis, const float *pow, const float *powerr, float *bisamp, float *bisamperr,
                    const float *bisphs, const float *bisphserr, const int64_t *uvpnt, int do_extrapolation,
                    float *data, float *data_err, float *phasor, int *ndof);

/* problem sizes of a context (0 when not set) */
int gpair_sizes(const gpair_ctx *ctx, int *W, int *nuv, int *npow, int *nbis);
/* synchronise the context's stream */
int gpair_sync(gpair_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* GPAIR_HIP_H */
