/*
 * gpair_nn.h -- learned start for the MEM-CG loop: a U-Net that maps the
 * dirty image and the prior to a start image and a per-pixel preconditioner,
 * trained and run on the GPU by the library's own convolution kernels.
 *
 * The reference has no counterpart to anything in this header: its loop always
 * starts from the prior (gpair1/trunk/gpair.c:217-225) with an unpreconditioned
 * gradient (gpair1/trunk/gpair.c:760-772).  The learned start feeds the
 * existing entry points: gpu_copy_image (start image) and gpair_set_pfactor
 * (preconditioner), declared in gpair_hip.h.
 *
 * Conventions: tensors are FP32 [N][C][H][W]; "bs" arguments are batch strides
 * in floats (>= C*H*W), so a tensor may sit at a channel offset inside a wider
 * one.  Single-layer calls take device pointers from gpair_nn_alloc and finish
 * before they return.  Status codes are those of gpair_hip.h.
 */
#ifndef GPAIR_NN_H
#define GPAIR_NN_H

#include "gpair_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* device memory for the single-layer calls (zero-filled); NULL on failure */
void *gpair_nn_alloc(long nfloat);
int gpair_nn_free(void *p);
int gpair_nn_upload(void *dst, const float *src, long nfloat);
int gpair_nn_download(float *dst, const void *src, long nfloat);

/* out = conv3x3(in; w [Co][Ci][3][3], b [Co]), zero padding 1; relu != 0 applies ReLU */
int gpair_nn_conv3_fwd(const float *in, long in_bs, int Ci, const float *w, const float *b, float *out,
                       long out_bs, int Co, int N, int H, int W, int relu);
/* gradients of a conv3 layer for the output gradient dout: din (may be NULL), dw, db (overwritten) */
int gpair_nn_conv3_bwd(const float *in, long in_bs, int Ci, const float *w, const float *dout, long dout_bs,
                       int Co, int N, int H, int W, float *din, long din_bs, float *dw, float *db);
/* 2x2 average pooling (backward != 0: in is the [C][H/2][W/2] gradient, out the [C][H][W] one) */
int gpair_nn_pool2(const float *in, long in_bs, float *out, long out_bs, int N, int C, int H, int W,
                   int backward);
/* nearest 2x upsampling of [C][H][W] (backward != 0: in is the [C][2H][2W] gradient) */
int gpair_nn_up2(const float *in, long in_bs, float *out, long out_bs, int N, int C, int H, int W, int backward);

/* U-Net with `levels` resolution levels of base*2^l channels (two 3x3 conv + ReLU
 * each, average pooling down, nearest upsampling and concatenation up, 3x3 head).
 * Parameter order: encoder a/b per level, decoder a/b from the deepest skip up,
 * head; each layer weights [Co][Ci][3][3] then bias [Co].  He-normal from seed. */
typedef struct gpair_unet gpair_unet;
int gpair_unet_create(gpair_unet **u, int in_ch, int out_ch, int base, int levels, int width, int max_batch,
                      unsigned long long seed);
int gpair_unet_destroy(gpair_unet *u);
long gpair_unet_num_params(const gpair_unet *u);
int gpair_unet_get_params(gpair_unet *u, float *host, long n);
int gpair_unet_set_params(gpair_unet *u, const float *host, long n);   /* also resets Adam */
int gpair_unet_get_grads(gpair_unet *u, float *host, long n);
/* x [N][in_ch][W][W] -> y [N][out_ch][W][W] (y may be NULL); device or host pointers */
int gpair_unet_forward(gpair_unet *u, const float *x, int N, float *y);
/* parameter gradients for the output gradient dy of the last forward */
int gpair_unet_backward(gpair_unet *u, const float *dy, int N);
int gpair_unet_adam(gpair_unet *u, float lr);
/* forward, mean-squared error against t, backward, Adam; *loss = the batch MSE before the step */
int gpair_unet_train_step(gpair_unet *u, const float *x, const float *t, int N, float lr, double *loss);

#ifdef __cplusplus
}
#endif

#endif /* GPAIR_NN_H */
