// Learned start: the convolutional network (a U-Net) that maps the dirty image
// and the prior to a start image and a per-pixel preconditioner for the MEM-CG
// loop.  The reference has no counterpart; declared in include/gpair_nn.h.
//
// Tensors are FP32 [N][C][H][W] with an explicit batch stride, so a channel
// concatenation is a channel offset inside one larger buffer (the encoder
// writes its skip output straight into the decoder's concatenation buffer).
// 3x3 convolutions (zero padding 1, cross-correlation) are direct: one thread
// per output pixel and a block of output channels, the weights uniform per
// block.  Data gradients are the transposed convolution; weight gradients are
// a two-stage, fixed-order reduction over the batch and the pixels, so
// training is deterministic.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../../include/gpair_nn.h"

#define NN_COB 8          // output channels per thread (forward / data gradient)
#define NN_WG_CHUNK 2048  // pixels per block of the weight-gradient reduction
#define NN_MSE_BLOCKS 256

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// out[n][co][p] = b[co] + sum_ci sum_k w[co][ci][k] in[n][ci][p + off_k]; optional ReLU
__global__ __launch_bounds__(256) void k_conv3_fwd(const float *__restrict__ in, long in_bs, int Ci,
                                                   const float *__restrict__ w, const float *__restrict__ b,
                                                   float *__restrict__ out, long out_bs, int Co, int H, int W,
                                                   int relu)
{
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    const int co0 = blockIdx.y * NN_COB, n = blockIdx.z;
    if (pix >= H * W)
        return;
    const int y = pix / W, x = pix % W;
    const bool up = y > 0, dn = y < H - 1, lf = x > 0, rt = x < W - 1;
    float acc[NN_COB];
#pragma unroll
    for (int j = 0; j < NN_COB; j++)
        acc[j] = (co0 + j < Co) ? b[co0 + j] : 0.f;
    const long HW = (long)H * W;
    const float *src = in + n * in_bs + pix;
    for (int ci = 0; ci < Ci; ci++) {
        const float *p = src + ci * HW;
        float v[9];
        v[0] = (up && lf) ? p[-W - 1] : 0.f;
        v[1] = up ? p[-W] : 0.f;
        v[2] = (up && rt) ? p[-W + 1] : 0.f;
        v[3] = lf ? p[-1] : 0.f;
        v[4] = p[0];
        v[5] = rt ? p[1] : 0.f;
        v[6] = (dn && lf) ? p[W - 1] : 0.f;
        v[7] = dn ? p[W] : 0.f;
        v[8] = (dn && rt) ? p[W + 1] : 0.f;
#pragma unroll
        for (int j = 0; j < NN_COB; j++) {
            if (co0 + j < Co) {
                const float *wk = w + ((long)(co0 + j) * Ci + ci) * 9;
#pragma unroll
                for (int k = 0; k < 9; k++)
                    acc[j] = fmaf(wk[k], v[k], acc[j]);
            }
        }
    }
    float *dst = out + n * out_bs + pix;
#pragma unroll
    for (int j = 0; j < NN_COB; j++)
        if (co0 + j < Co)
            dst[(co0 + j) * HW] = (relu && acc[j] < 0.f) ? 0.f : acc[j];
}

// din[n][ci][p] = sum_co sum_k w[co][ci][k] dout[n][co][p - off_k]  (overwrites din)
__global__ __launch_bounds__(256) void k_conv3_bwd_data(const float *__restrict__ dout, long dout_bs, int Co,
                                                        const float *__restrict__ w, float *__restrict__ din,
                                                        long din_bs, int Ci, int H, int W)
{
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    const int ci0 = blockIdx.y * NN_COB, n = blockIdx.z;
    if (pix >= H * W)
        return;
    const int y = pix / W, x = pix % W;
    const bool up = y > 0, dn = y < H - 1, lf = x > 0, rt = x < W - 1;
    float acc[NN_COB];
#pragma unroll
    for (int j = 0; j < NN_COB; j++)
        acc[j] = 0.f;
    const long HW = (long)H * W;
    const float *src = dout + n * dout_bs + pix;
    for (int co = 0; co < Co; co++) {
        const float *p = src + co * HW;
        // tap k = (ky, kx) takes dout at (y - ky + 1, x - kx + 1)
        float v[9];
        v[0] = (dn && rt) ? p[W + 1] : 0.f;
        v[1] = dn ? p[W] : 0.f;
        v[2] = (dn && lf) ? p[W - 1] : 0.f;
        v[3] = rt ? p[1] : 0.f;
        v[4] = p[0];
        v[5] = lf ? p[-1] : 0.f;
        v[6] = (up && rt) ? p[-W + 1] : 0.f;
        v[7] = up ? p[-W] : 0.f;
        v[8] = (up && lf) ? p[-W - 1] : 0.f;
#pragma unroll
        for (int j = 0; j < NN_COB; j++) {
            if (ci0 + j < Ci) {
                const float *wk = w + ((long)co * Ci + ci0 + j) * 9;
#pragma unroll
                for (int k = 0; k < 9; k++)
                    acc[j] = fmaf(wk[k], v[k], acc[j]);
            }
        }
    }
    float *dst = din + n * din_bs + pix;
#pragma unroll
    for (int j = 0; j < NN_COB; j++)
        if (ci0 + j < Ci)
            dst[(ci0 + j) * HW] = acc[j];
}

// stage 1 of dw / db: block (ci, co, s) sums over the pixels [s*chunk, (s+1)*chunk)
// of the flattened (n, p) range; part[((co*Ci + ci)*S + s)*10 + k], k = 9 the bias term
__global__ __launch_bounds__(256) void k_conv3_wgrad_part(const float *__restrict__ in, long in_bs, int Ci,
                                                          const float *__restrict__ dout, long dout_bs, int Co,
                                                          int N, int H, int W, int chunk, float *__restrict__ part)
{
    const int ci = blockIdx.x, co = blockIdx.y, s = blockIdx.z, S = gridDim.z;
    const int HW = H * W;
    const long total = (long)N * HW;
    const long beg = (long)s * chunk, end = beg + chunk < total ? beg + chunk : total;
    float acc[10];
#pragma unroll
    for (int k = 0; k < 10; k++)
        acc[k] = 0.f;
    for (long idx = beg + threadIdx.x; idx < end; idx += blockDim.x) {
        const int n = (int)(idx / HW), pix = (int)(idx % HW);
        const int y = pix / W, x = pix % W;
        const bool up = y > 0, dn = y < H - 1, lf = x > 0, rt = x < W - 1;
        const float g = dout[n * dout_bs + (long)co * HW + pix];
        const float *p = in + n * in_bs + (long)ci * HW + pix;
        acc[0] = fmaf(g, (up && lf) ? p[-W - 1] : 0.f, acc[0]);
        acc[1] = fmaf(g, up ? p[-W] : 0.f, acc[1]);
        acc[2] = fmaf(g, (up && rt) ? p[-W + 1] : 0.f, acc[2]);
        acc[3] = fmaf(g, lf ? p[-1] : 0.f, acc[3]);
        acc[4] = fmaf(g, p[0], acc[4]);
        acc[5] = fmaf(g, rt ? p[1] : 0.f, acc[5]);
        acc[6] = fmaf(g, (dn && lf) ? p[W - 1] : 0.f, acc[6]);
        acc[7] = fmaf(g, dn ? p[W] : 0.f, acc[7]);
        acc[8] = fmaf(g, (dn && rt) ? p[W + 1] : 0.f, acc[8]);
        acc[9] += g;
    }
    __shared__ float sh[10][256];
#pragma unroll
    for (int k = 0; k < 10; k++)
        sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
#pragma unroll
            for (int k = 0; k < 10; k++)
                sh[k][threadIdx.x] += sh[k][threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x < 10)
        part[(((long)co * Ci + ci) * S + s) * 10 + threadIdx.x] = sh[threadIdx.x][0];
}

// stage 2: dw[co][ci][k] = sum_s part, db[co] = sum_s part[ci = 0][9]; fixed order
__global__ void k_conv3_wgrad_finish(const float *__restrict__ part, int Ci, int Co, int S, float *__restrict__ dw,
                                     float *__restrict__ db)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nw = (long)Co * Ci * 9;
    const float *p;
    if (i < nw) {
        const long co = i / (Ci * 9), r = i % (Ci * 9), ci = r / 9, k = r % 9;
        p = part + ((co * Ci + ci) * S) * 10 + k;
    } else if (i < nw + Co) {
        p = part + (((i - nw) * Ci) * S) * 10 + 9;
    } else {
        return;
    }
    float s = 0.f;
    for (int j = 0; j < S; j++)
        s += p[j * 10];
    if (i < nw)
        dw[i] = s;
    else
        db[i - nw] = s;
}

// dy[n][i] = 0 where y[n][i] <= 0 (ReLU backward through its output), i < len
__global__ void k_relu_bwd(float *__restrict__ dy, long dy_bs, const float *__restrict__ y, long y_bs, long len)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (i < len && !(y[n * y_bs + i] > 0.f))
        dy[n * dy_bs + i] = 0.f;
}

// y[n][i] += x[n][i], i < len
__global__ void k_add(float *__restrict__ y, long y_bs, const float *__restrict__ x, long x_bs, long len)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (i < len)
        y[n * y_bs + i] += x[n * x_bs + i];
}

// 2x2 average pooling of [C][H][W] -> [C][H/2][W/2]
__global__ void k_pool2_fwd(const float *__restrict__ in, long in_bs, float *__restrict__ out, long out_bs, int C,
                            int H, int W)
{
    const int Ho = H / 2, Wo = W / 2;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (i >= (long)C * Ho * Wo)
        return;
    const long c = i / (Ho * Wo);
    const int r = (int)(i % (Ho * Wo)), y = r / Wo, x = r % Wo;
    const float *p = in + n * in_bs + c * H * W + (2 * y) * W + 2 * x;
    out[n * out_bs + i] = 0.25f * ((p[0] + p[1]) + (p[W] + p[W + 1]));
}

// din[c][y][x] = dout[c][y/2][x/2] / 4  (din [C][H][W])
__global__ void k_pool2_bwd(const float *__restrict__ dout, long dout_bs, float *__restrict__ din, long din_bs,
                            int C, int H, int W)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (i >= (long)C * H * W)
        return;
    const long c = i / (H * W);
    const int r = (int)(i % (H * W)), y = r / W, x = r % W;
    din[n * din_bs + i] = 0.25f * dout[n * dout_bs + c * (H / 2) * (W / 2) + (y / 2) * (W / 2) + x / 2];
}

// nearest 2x upsampling [C][H][W] -> [C][2H][2W]
__global__ void k_up2_fwd(const float *__restrict__ in, long in_bs, float *__restrict__ out, long out_bs, int C,
                          int H, int W)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    const int Ho = 2 * H, Wo = 2 * W;
    if (i >= (long)C * Ho * Wo)
        return;
    const long c = i / (Ho * Wo);
    const int r = (int)(i % (Ho * Wo)), y = r / Wo, x = r % Wo;
    out[n * out_bs + i] = in[n * in_bs + c * H * W + (y / 2) * W + x / 2];
}

// din[c][y][x] = sum of the 2x2 block of dout (dout [C][2H][2W])
__global__ void k_up2_bwd(const float *__restrict__ dout, long dout_bs, float *__restrict__ din, long din_bs, int C,
                          int H, int W)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (i >= (long)C * H * W)
        return;
    const long c = i / (H * W);
    const int r = (int)(i % (H * W)), y = r / W, x = r % W;
    const float *p = dout + n * dout_bs + c * 4 * H * W + (2 * y) * (2 * W) + 2 * x;
    din[n * din_bs + i] = (p[0] + p[1]) + (p[2 * W] + p[2 * W + 1]);
}

// mean squared error: dpred = 2 (pred - tgt) / n, FP64 block partials of sum d^2
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ pred, const float *__restrict__ tgt, long n,
                                             float *__restrict__ dpred, double *__restrict__ partial)
{
    __shared__ double sh[256];
    double s = 0.0;
    const float sc = 2.0f / (float)n;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float d = pred[i] - tgt[i];
        dpred[i] = sc * d;
        s += (double)d * d;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k)
            sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partial[blockIdx.x] = sh[0];
}

// Adam with bias corrections c1 = 1 - b1^t, c2 = 1 - b2^t
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, long n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------------------
// launch helpers (one stream)
// ---------------------------------------------------------------------------
static inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

static void conv_fwd(hipStream_t st, const float *in, long in_bs, int Ci, const float *w, const float *b, float *out,
                     long out_bs, int Co, int N, int H, int W, int relu)
{
    dim3 g(nblk((long)H * W), (Co + NN_COB - 1) / NN_COB, N);
    k_conv3_fwd<<<g, 256, 0, st>>>(in, in_bs, Ci, w, b, out, out_bs, Co, H, W, relu);
}

static int wgrad_splits(int Ci, int Co, int N, int H, int W)
{
    const long total = (long)N * H * W;
    long S = (total + NN_WG_CHUNK - 1) / NN_WG_CHUNK;
    const long want = (4096 + (long)Ci * Co - 1) / ((long)Ci * Co);   // aim at >= 4096 blocks
    if (S > want)
        S = want;
    return S < 1 ? 1 : (int)S;
}

static long wgrad_part_size(int Ci, int Co, int N, int H, int W)
{
    return (long)Ci * Co * wgrad_splits(Ci, Co, N, H, W) * 10;
}

// part holds wgrad_part_size(...) floats
static void conv_bwd(hipStream_t st, const float *in, long in_bs, int Ci, const float *w, const float *dout,
                     long dout_bs, int Co, int N, int H, int W, float *din, long din_bs, float *dw, float *db,
                     float *part)
{
    if (din) {
        dim3 g(nblk((long)H * W), (Ci + NN_COB - 1) / NN_COB, N);
        k_conv3_bwd_data<<<g, 256, 0, st>>>(dout, dout_bs, Co, w, din, din_bs, Ci, H, W);
    }
    const int S = wgrad_splits(Ci, Co, N, H, W);
    const long total = (long)N * H * W;
    const int chunk = (int)((total + S - 1) / S);
    k_conv3_wgrad_part<<<dim3(Ci, Co, S), 256, 0, st>>>(in, in_bs, Ci, dout, dout_bs, Co, N, H, W, chunk, part);
    k_conv3_wgrad_finish<<<nblk((long)Co * Ci * 9 + Co), 256, 0, st>>>(part, Ci, Co, S, dw, db);
}

// ---------------------------------------------------------------------------
// single-layer entry points (device pointers from gpair_nn_alloc; synchronous)
// ---------------------------------------------------------------------------
static int nn_status(void)
{
    if (hipGetLastError() != hipSuccess)
        return GPAIR_E_HIP;
    return hipDeviceSynchronize() == hipSuccess ? GPAIR_OK : GPAIR_E_HIP;
}

extern "C" void *gpair_nn_alloc(long nfloat)
{
    void *p = nullptr;
    const size_t bytes = sizeof(float) * (size_t)(nfloat > 0 ? nfloat : 1);
    if (hipMalloc(&p, bytes) != hipSuccess)
        return nullptr;
    if (hipMemset(p, 0, bytes) != hipSuccess) {
        (void)hipFree(p);
        return nullptr;
    }
    return p;
}

extern "C" int gpair_nn_free(void *p) { return hipFree(p) == hipSuccess ? GPAIR_OK : GPAIR_E_HIP; }

extern "C" int gpair_nn_upload(void *dst, const float *src, long nfloat)
{
    if (!dst || !src || nfloat < 0)
        return GPAIR_E_ARG;
    return hipMemcpy(dst, src, sizeof(float) * nfloat, hipMemcpyHostToDevice) == hipSuccess ? GPAIR_OK
                                                                                          : GPAIR_E_HIP;
}

extern "C" int gpair_nn_download(float *dst, const void *src, long nfloat)
{
    if (!dst || !src || nfloat < 0)
        return GPAIR_E_ARG;
    return hipMemcpy(dst, src, sizeof(float) * nfloat, hipMemcpyDeviceToHost) == hipSuccess ? GPAIR_OK
                                                                                          : GPAIR_E_HIP;
}

extern "C" int gpair_nn_conv3_fwd(const float *in, long in_bs, int Ci, const float *w, const float *b, float *out,
                                  long out_bs, int Co, int N, int H, int W, int relu)
{
    if (!in || !w || !b || !out || Ci <= 0 || Co <= 0 || N <= 0 || H <= 0 || W <= 0 || N > 65535 ||
        in_bs < (long)Ci * H * W || out_bs < (long)Co * H * W)
        return GPAIR_E_ARG;
    conv_fwd(0, in, in_bs, Ci, w, b, out, out_bs, Co, N, H, W, relu);
    return nn_status();
}

extern "C" int gpair_nn_conv3_bwd(const float *in, long in_bs, int Ci, const float *w, const float *dout,
                                  long dout_bs, int Co, int N, int H, int W, float *din, long din_bs, float *dw,
                                  float *db)
{
    if (!in || !w || !dout || !dw || !db || Ci <= 0 || Co <= 0 || N <= 0 || H <= 0 || W <= 0 || N > 65535 ||
        Co > 65535 || in_bs < (long)Ci * H * W || dout_bs < (long)Co * H * W || (din && din_bs < (long)Ci * H * W))
        return GPAIR_E_ARG;
    float *part = nullptr;
    if (hipMalloc((void **)&part, sizeof(float) * wgrad_part_size(Ci, Co, N, H, W)) != hipSuccess)
        return GPAIR_E_NOMEM;
    conv_bwd(0, in, in_bs, Ci, w, dout, dout_bs, Co, N, H, W, din, din_bs, dw, db, part);
    const int rc = nn_status();
    (void)hipFree(part);
    return rc;
}

extern "C" int gpair_nn_pool2(const float *in, long in_bs, float *out, long out_bs, int N, int C, int H, int W,
                              int backward)
{
    if (!in || !out || N <= 0 || N > 65535 || C <= 0 || H < 2 || W < 2 || ((H | W) & 1))
        return GPAIR_E_ARG;
    if (backward)   // in = dout [C][H/2][W/2], out = din [C][H][W]
        k_pool2_bwd<<<dim3(nblk((long)C * H * W), N), 256>>>(in, in_bs, out, out_bs, C, H, W);
    else
        k_pool2_fwd<<<dim3(nblk((long)C * (H / 2) * (W / 2)), N), 256>>>(in, in_bs, out, out_bs, C, H, W);
    return nn_status();
}

extern "C" int gpair_nn_up2(const float *in, long in_bs, float *out, long out_bs, int N, int C, int H, int W,
                            int backward)
{
    if (!in || !out || N <= 0 || N > 65535 || C <= 0 || H <= 0 || W <= 0)
        return GPAIR_E_ARG;
    if (backward)   // in = dout [C][2H][2W], out = din [C][H][W]
        k_up2_bwd<<<dim3(nblk((long)C * H * W), N), 256>>>(in, in_bs, out, out_bs, C, H, W);
    else
        k_up2_fwd<<<dim3(nblk((long)C * 4 * H * W), N), 256>>>(in, in_bs, out, out_bs, C, H, W);
    return nn_status();
}

// ---------------------------------------------------------------------------
// U-Net: levels l = 0..L-1 with C_l = base * 2^l channels at (W >> l)^2 pixels.
//   encoder l:  x_l -> h_l = relu(conv a) -> e_l = relu(conv b);  x_0 the input,
//               x_l = pool(e_{l-1});  e_{L-1} is the bottleneck (= d_{L-1})
//   decoder l = L-2..0:  cat_l = [up(d_{l+1}), e_l] -> hd_l = relu(conv a)
//               -> d_l = relu(conv b);  head: y = conv(d_0), no ReLU
// ---------------------------------------------------------------------------
struct nn_layer {
    int ci, co;
    long w, b;   // offsets into the flat parameter array
};

struct gpair_unet {
    int in_ch, out_ch, base, L, W, maxN;
    std::vector<nn_layer> layers;
    long nparam;
    hipStream_t st;
    float *prm, *grd, *am, *av;
    int adam_t;
    float *x0;                                   // [maxN][in_ch][HW0]
    std::vector<float *> xl, h, cat, hd, d;      // activations per level
    float *e_top, *out;
    std::vector<float *> dcat, s1, s2;           // gradient buffers per level
    float *dout, *part;
    double *mse_part, *mse_host;
};

static int lvl_ch(const gpair_unet *u, int l) { return u->base << l; }
static int lvl_w(const gpair_unet *u, int l) { return u->W >> l; }
static long lvl_hw(const gpair_unet *u, int l) { return (long)lvl_w(u, l) * lvl_w(u, l); }
// channels of the concatenation buffer at level l < L-1
static int cat_ch(const gpair_unet *u, int l) { return lvl_ch(u, l + 1) + lvl_ch(u, l); }

static float *nn_dalloc(long n, int *fail)
{
    void *p = nullptr;
    const size_t bytes = sizeof(float) * (size_t)(n > 0 ? n : 1);
    if (hipMalloc(&p, bytes) != hipSuccess) {
        *fail = 1;
        return nullptr;
    }
    if (hipMemset(p, 0, bytes) != hipSuccess)
        *fail = 1;
    return (float *)p;
}

static void nn_dfree(float *p)
{
    if (p)
        (void)hipFree(p);
}

// splitmix64 + Box-Muller: seeded He-normal initialisation
static uint64_t nn_next(uint64_t *s)
{
    uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static double nn_normal(uint64_t *s)
{
    const double u1 = ((double)(nn_next(s) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
    const double u2 = (double)(nn_next(s) >> 11) * (1.0 / 9007199254740992.0);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// level at which parameter layer k runs
static int layer_level(const gpair_unet *u, int k)
{
    const int L = u->L, nl = (int)u->layers.size();
    if (k < 2 * L)
        return k / 2;
    if (k < nl - 1)
        return L - 2 - (k - 2 * L) / 2;
    return 0;
}

extern "C" int gpair_unet_destroy(gpair_unet *u)
{
    if (!u)
        return GPAIR_OK;
    nn_dfree(u->prm); nn_dfree(u->grd); nn_dfree(u->am); nn_dfree(u->av);
    nn_dfree(u->x0); nn_dfree(u->e_top); nn_dfree(u->out); nn_dfree(u->dout); nn_dfree(u->part);
    for (std::vector<float *> *v : {&u->xl, &u->h, &u->cat, &u->hd, &u->d, &u->dcat, &u->s1, &u->s2})
        for (float *p : *v)
            nn_dfree(p);
    if (u->mse_part)
        (void)hipFree(u->mse_part);
    if (u->mse_host)
        (void)hipHostFree(u->mse_host);
    if (u->st)
        (void)hipStreamDestroy(u->st);
    delete u;
    return GPAIR_OK;
}

extern "C" int gpair_unet_create(gpair_unet **out, int in_ch, int out_ch, int base, int levels, int width,
                                 int max_batch, unsigned long long seed)
{
    if (!out)
        return GPAIR_E_ARG;
    *out = nullptr;
    if (in_ch <= 0 || out_ch <= 0 || base <= 0 || levels < 2 || levels > 8 || max_batch <= 0 ||
        max_batch > 65535 || width <= 0 || (width % (1 << (levels - 1))) != 0 || (base << (levels - 1)) > 4096)
        return GPAIR_E_ARG;
    gpair_unet *u = new gpair_unet();
    u->in_ch = in_ch; u->out_ch = out_ch; u->base = base; u->L = levels; u->W = width; u->maxN = max_batch;
    long off = 0;
    auto add = [&](int ci, int co) {
        nn_layer ly{ci, co, off, off + (long)co * ci * 9};
        off += (long)co * ci * 9 + co;
        u->layers.push_back(ly);
    };
    int cin = in_ch;
    for (int l = 0; l < levels; l++) {
        add(cin, lvl_ch(u, l));
        add(lvl_ch(u, l), lvl_ch(u, l));
        cin = lvl_ch(u, l);
    }
    for (int l = levels - 2; l >= 0; l--) {
        add(cat_ch(u, l), lvl_ch(u, l));
        add(lvl_ch(u, l), lvl_ch(u, l));
    }
    add(lvl_ch(u, 0), out_ch);
    u->nparam = off;
    if (hipStreamCreate(&u->st) != hipSuccess) {
        u->st = nullptr;
        gpair_unet_destroy(u);
        return GPAIR_E_HIP;
    }
    int fail = 0;
    const long N = max_batch;
    u->prm = nn_dalloc(off, &fail);
    u->grd = nn_dalloc(off, &fail);
    u->am = nn_dalloc(off, &fail);
    u->av = nn_dalloc(off, &fail);
    u->x0 = nn_dalloc(N * in_ch * lvl_hw(u, 0), &fail);
    for (int l = 0; l < levels; l++) {
        const long hw = lvl_hw(u, l), c = lvl_ch(u, l);
        const long cmax = l < levels - 1 ? cat_ch(u, l) : c;
        u->xl.push_back(l > 0 ? nn_dalloc(N * lvl_ch(u, l - 1) * hw, &fail) : nullptr);
        u->h.push_back(nn_dalloc(N * c * hw, &fail));
        u->s1.push_back(nn_dalloc(N * cmax * hw, &fail));
        u->s2.push_back(nn_dalloc(N * cmax * hw, &fail));
        if (l < levels - 1) {
            u->cat.push_back(nn_dalloc(N * cat_ch(u, l) * hw, &fail));
            u->dcat.push_back(nn_dalloc(N * cat_ch(u, l) * hw, &fail));
            u->hd.push_back(nn_dalloc(N * c * hw, &fail));
            u->d.push_back(nn_dalloc(N * c * hw, &fail));
        }
    }
    u->e_top = nn_dalloc(N * lvl_ch(u, levels - 1) * lvl_hw(u, levels - 1), &fail);
    u->out = nn_dalloc(N * out_ch * lvl_hw(u, 0), &fail);
    u->dout = nn_dalloc(N * out_ch * lvl_hw(u, 0), &fail);
    long part = 0;
    for (int k = 0; k < (int)u->layers.size(); k++) {
        const int w = lvl_w(u, layer_level(u, k));
        const long p = wgrad_part_size(u->layers[k].ci, u->layers[k].co, max_batch, w, w);
        if (p > part)
            part = p;
    }
    u->part = nn_dalloc(part, &fail);
    if (hipMalloc((void **)&u->mse_part, sizeof(double) * NN_MSE_BLOCKS) != hipSuccess) {
        u->mse_part = nullptr;
        fail = 1;
    }
    if (hipHostMalloc((void **)&u->mse_host, sizeof(double) * NN_MSE_BLOCKS, hipHostMallocDefault) != hipSuccess) {
        u->mse_host = nullptr;
        fail = 1;
    }
    if (fail) {
        gpair_unet_destroy(u);
        return GPAIR_E_NOMEM;
    }
    // He-normal weights (std sqrt(2 / (9 Ci))), zero biases
    std::vector<float> host(off, 0.f);
    uint64_t s = seed ? seed : 0x2545f4914f6cdd1dull;
    for (const nn_layer &ly : u->layers) {
        const double sd = sqrt(2.0 / (9.0 * ly.ci));
        for (long i = 0; i < (long)ly.co * ly.ci * 9; i++)
            host[ly.w + i] = (float)(sd * nn_normal(&s));
    }
    if (hipMemcpy(u->prm, host.data(), sizeof(float) * off, hipMemcpyHostToDevice) != hipSuccess) {
        gpair_unet_destroy(u);
        return GPAIR_E_HIP;
    }
    *out = u;
    return GPAIR_OK;
}

extern "C" long gpair_unet_num_params(const gpair_unet *u) { return u ? u->nparam : -1; }

extern "C" int gpair_unet_get_params(gpair_unet *u, float *host, long n)
{
    if (!u || !host || n != u->nparam)
        return GPAIR_E_ARG;
    return hipMemcpy(host, u->prm, sizeof(float) * n, hipMemcpyDeviceToHost) == hipSuccess ? GPAIR_OK : GPAIR_E_HIP;
}

extern "C" int gpair_unet_set_params(gpair_unet *u, const float *host, long n)
{
    if (!u || !host || n != u->nparam)
        return GPAIR_E_ARG;
    if (hipMemcpy(u->prm, host, sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess)
        return GPAIR_E_HIP;
    // fresh optimiser state for the new parameters
    if (hipMemset(u->am, 0, sizeof(float) * n) != hipSuccess || hipMemset(u->av, 0, sizeof(float) * n) != hipSuccess)
        return GPAIR_E_HIP;
    u->adam_t = 0;
    return GPAIR_OK;
}

extern "C" int gpair_unet_get_grads(gpair_unet *u, float *host, long n)
{
    if (!u || !host || n != u->nparam)
        return GPAIR_E_ARG;
    return hipMemcpy(host, u->grd, sizeof(float) * n, hipMemcpyDeviceToHost) == hipSuccess ? GPAIR_OK : GPAIR_E_HIP;
}

static int unet_forward(gpair_unet *u, const float *x, int N)
{
    hipStream_t st = u->st;
    const int L = u->L;
    if (hipMemcpyAsync(u->x0, x, sizeof(float) * N * u->in_ch * lvl_hw(u, 0), hipMemcpyDefault, st) != hipSuccess)
        return GPAIR_E_HIP;
    int k = 0;
    for (int l = 0; l < L; l++) {
        const int w = lvl_w(u, l), c = lvl_ch(u, l);
        const long hw = lvl_hw(u, l);
        const float *in;
        long in_bs;
        int cin;
        if (l == 0) {
            in = u->x0; in_bs = (long)u->in_ch * hw; cin = u->in_ch;
        } else {
            // x_l = pool(e_{l-1}); e_{l-1} sits in cat_{l-1} at channel C_l
            const int cp = lvl_ch(u, l - 1);
            const long hwp = lvl_hw(u, l - 1);
            const float *e = u->cat[l - 1] + (long)c * hwp;
            k_pool2_fwd<<<dim3(nblk((long)cp * hw), N), 256, 0, st>>>(e, (long)cat_ch(u, l - 1) * hwp, u->xl[l],
                                                                     (long)cp * hw, cp, 2 * w, 2 * w);
            in = u->xl[l]; in_bs = (long)cp * hw; cin = cp;
        }
        const nn_layer &a = u->layers[k++], &b = u->layers[k++];
        conv_fwd(st, in, in_bs, cin, u->prm + a.w, u->prm + a.b, u->h[l], (long)c * hw, c, N, w, w, 1);
        float *e = (l < L - 1) ? u->cat[l] + (long)lvl_ch(u, l + 1) * hw : u->e_top;
        const long e_bs = (l < L - 1) ? (long)cat_ch(u, l) * hw : (long)c * hw;
        conv_fwd(st, u->h[l], (long)c * hw, c, u->prm + b.w, u->prm + b.b, e, e_bs, c, N, w, w, 1);
    }
    const float *dprev = u->e_top;
    for (int l = L - 2; l >= 0; l--) {
        const int w = lvl_w(u, l), c = lvl_ch(u, l), cu = lvl_ch(u, l + 1), cc = cat_ch(u, l);
        const long hw = lvl_hw(u, l), hwu = lvl_hw(u, l + 1);
        k_up2_fwd<<<dim3(nblk((long)cu * hw), N), 256, 0, st>>>(dprev, (long)cu * hwu, u->cat[l], (long)cc * hw, cu,
                                                                w / 2, w / 2);
        const nn_layer &a = u->layers[k++], &b = u->layers[k++];
        conv_fwd(st, u->cat[l], (long)cc * hw, cc, u->prm + a.w, u->prm + a.b, u->hd[l], (long)c * hw, c, N, w, w,
                 1);
        conv_fwd(st, u->hd[l], (long)c * hw, c, u->prm + b.w, u->prm + b.b, u->d[l], (long)c * hw, c, N, w, w, 1);
        dprev = u->d[l];
    }
    const nn_layer &hd = u->layers[k];
    conv_fwd(st, dprev, (long)lvl_ch(u, 0) * lvl_hw(u, 0), lvl_ch(u, 0), u->prm + hd.w, u->prm + hd.b, u->out,
             (long)u->out_ch * lvl_hw(u, 0), u->out_ch, N, u->W, u->W, 0);
    return hipGetLastError() == hipSuccess ? GPAIR_OK : GPAIR_E_HIP;
}

// parameter gradients into u->grd for the output gradient in u->dout (last forward)
static int unet_backward(gpair_unet *u, int N)
{
    hipStream_t st = u->st;
    const int L = u->L;
    const int nl = (int)u->layers.size();
    const long hw0 = lvl_hw(u, 0);
    {   // head: input d_0, gradient of d_0 -> s1[0]
        const nn_layer &hd = u->layers[nl - 1];
        const int c0 = lvl_ch(u, 0);
        conv_bwd(st, u->d[0], (long)c0 * hw0, c0, u->prm + hd.w, u->dout, (long)u->out_ch * hw0, u->out_ch, N, u->W,
                 u->W, u->s1[0], (long)c0 * hw0, u->grd + hd.w, u->grd + hd.b, u->part);
    }
    // decoder, shallowest level first; s1[l] holds the gradient of d_l
    for (int l = 0; l <= L - 2; l++) {
        const int w = lvl_w(u, l), c = lvl_ch(u, l), cu = lvl_ch(u, l + 1), cc = cat_ch(u, l);
        const long hw = lvl_hw(u, l), hwu = lvl_hw(u, l + 1);
        const int ka = 2 * L + 2 * (L - 2 - l), kb = ka + 1;
        const nn_layer &a = u->layers[ka], &b = u->layers[kb];
        k_relu_bwd<<<dim3(nblk((long)c * hw), N), 256, 0, st>>>(u->s1[l], (long)c * hw, u->d[l], (long)c * hw,
                                                                (long)c * hw);
        conv_bwd(st, u->hd[l], (long)c * hw, c, u->prm + b.w, u->s1[l], (long)c * hw, c, N, w, w, u->s2[l],
                 (long)c * hw, u->grd + b.w, u->grd + b.b, u->part);
        k_relu_bwd<<<dim3(nblk((long)c * hw), N), 256, 0, st>>>(u->s2[l], (long)c * hw, u->hd[l], (long)c * hw,
                                                                (long)c * hw);
        conv_bwd(st, u->cat[l], (long)cc * hw, cc, u->prm + a.w, u->s2[l], (long)c * hw, c, N, w, w, u->dcat[l],
                 (long)cc * hw, u->grd + a.w, u->grd + a.b, u->part);
        // gradient of d_{l+1} (the bottleneck at l = L-2) from the upsampled channels -> s1[l+1]
        k_up2_bwd<<<dim3(nblk((long)cu * hwu), N), 256, 0, st>>>(u->dcat[l], (long)cc * hw, u->s1[l + 1],
                                                                 (long)cu * hwu, cu, w / 2, w / 2);
    }
    // encoder, deepest level first
    for (int l = L - 1; l >= 0; l--) {
        const int w = lvl_w(u, l), c = lvl_ch(u, l);
        const long hw = lvl_hw(u, l);
        const nn_layer &a = u->layers[2 * l], &b = u->layers[2 * l + 1];
        float *ge;   // gradient of e_l
        long ge_bs;
        const float *e;
        if (l == L - 1) {
            ge = u->s1[l]; ge_bs = (long)c * hw; e = u->e_top;
        } else {
            // skip path (dcat_l at channel C_{l+1}) + pooling path from level l+1 (s2[l])
            ge_bs = (long)cat_ch(u, l) * hw;
            ge = u->dcat[l] + (long)lvl_ch(u, l + 1) * hw;
            e = u->cat[l] + (long)lvl_ch(u, l + 1) * hw;
            k_add<<<dim3(nblk((long)c * hw), N), 256, 0, st>>>(ge, ge_bs, u->s2[l], (long)c * hw, (long)c * hw);
        }
        k_relu_bwd<<<dim3(nblk((long)c * hw), N), 256, 0, st>>>(ge, ge_bs, e, ge_bs, (long)c * hw);
        float *gh = (l == L - 1) ? u->s2[l] : u->s1[l];   // gradient of h_l
        conv_bwd(st, u->h[l], (long)c * hw, c, u->prm + b.w, ge, ge_bs, c, N, w, w, gh, (long)c * hw, u->grd + b.w,
                 u->grd + b.b, u->part);
        k_relu_bwd<<<dim3(nblk((long)c * hw), N), 256, 0, st>>>(gh, (long)c * hw, u->h[l], (long)c * hw,
                                                                (long)c * hw);
        if (l == 0) {
            conv_bwd(st, u->x0, (long)u->in_ch * hw, u->in_ch, u->prm + a.w, gh, (long)c * hw, c, N, w, w, nullptr,
                     0, u->grd + a.w, u->grd + a.b, u->part);
        } else {
            // gradient of x_l = pool(e_{l-1}), then the pooling backward -> s2[l-1]
            const int cp = lvl_ch(u, l - 1);
            float *gx = (l == L - 1) ? u->s1[l] : u->s2[l];
            conv_bwd(st, u->xl[l], (long)cp * hw, cp, u->prm + a.w, gh, (long)c * hw, c, N, w, w, gx, (long)cp * hw,
                     u->grd + a.w, u->grd + a.b, u->part);
            k_pool2_bwd<<<dim3(nblk((long)cp * 4 * hw), N), 256, 0, st>>>(gx, (long)cp * hw, u->s2[l - 1],
                                                                          (long)cp * 4 * hw, cp, 2 * w, 2 * w);
        }
    }
    return hipGetLastError() == hipSuccess ? GPAIR_OK : GPAIR_E_HIP;
}

extern "C" int gpair_unet_forward(gpair_unet *u, const float *x, int N, float *y)
{
    if (!u || !x || N <= 0 || N > u->maxN)
        return GPAIR_E_ARG;
    int rc = unet_forward(u, x, N);
    if (rc)
        return rc;
    if (y && hipMemcpyAsync(y, u->out, sizeof(float) * N * u->out_ch * lvl_hw(u, 0), hipMemcpyDefault, u->st) !=
                 hipSuccess)
        return GPAIR_E_HIP;
    return hipStreamSynchronize(u->st) == hipSuccess ? GPAIR_OK : GPAIR_E_HIP;
}

extern "C" int gpair_unet_backward(gpair_unet *u, const float *dy, int N)
{
    if (!u || !dy || N <= 0 || N > u->maxN)
        return GPAIR_E_ARG;
    if (hipMemcpyAsync(u->dout, dy, sizeof(float) * N * u->out_ch * lvl_hw(u, 0), hipMemcpyDefault, u->st) !=
        hipSuccess)
        return GPAIR_E_HIP;
    int rc = unet_backward(u, N);
    if (rc)
        return rc;
    return hipStreamSynchronize(u->st) == hipSuccess ? GPAIR_OK : GPAIR_E_HIP;
}

static int unet_adam_async(gpair_unet *u, float lr)
{
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    u->adam_t++;
    const float c1 = (float)(1.0 - pow((double)b1, u->adam_t)), c2 = (float)(1.0 - pow((double)b2, u->adam_t));
    k_adam<<<nblk(u->nparam), 256, 0, u->st>>>(u->prm, u->grd, u->am, u->av, u->nparam, lr, b1, b2, eps, c1, c2);
    return hipGetLastError() == hipSuccess ? GPAIR_OK : GPAIR_E_HIP;
}

extern "C" int gpair_unet_adam(gpair_unet *u, float lr)
{
    if (!u || !(lr > 0.f))
        return GPAIR_E_ARG;
    int rc = unet_adam_async(u, lr);
    if (rc)
        return rc;
    return hipStreamSynchronize(u->st) == hipSuccess ? GPAIR_OK : GPAIR_E_HIP;
}

extern "C" int gpair_unet_train_step(gpair_unet *u, const float *x, const float *t, int N, float lr, double *loss)
{
    if (!u || !x || !t || N <= 0 || N > u->maxN || !(lr > 0.f))
        return GPAIR_E_ARG;
    int rc = unet_forward(u, x, N);
    if (rc)
        return rc;
    const long n = (long)N * u->out_ch * lvl_hw(u, 0);
    k_mse<<<NN_MSE_BLOCKS, 256, 0, u->st>>>(u->out, t, n, u->dout, u->mse_part);
    if ((rc = unet_backward(u, N)))
        return rc;
    if (hipMemcpyAsync(u->mse_host, u->mse_part, sizeof(double) * NN_MSE_BLOCKS, hipMemcpyDeviceToHost, u->st) !=
        hipSuccess)
        return GPAIR_E_HIP;
    if ((rc = unet_adam_async(u, lr)))
        return rc;
    if (hipStreamSynchronize(u->st) != hipSuccess)
        return GPAIR_E_HIP;
    if (loss) {
        double s = 0.0;
        for (int i = 0; i < NN_MSE_BLOCKS; i++)
            s += u->mse_host[i];
        *loss = s / (double)n;
    }
    return GPAIR_OK;
}
