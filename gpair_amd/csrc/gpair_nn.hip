// Learned start: the convolutional network (a U-Net) that maps the dirty image
// and the prior to a start image and a per-pixel preconditioner for the MEM-CG
// loop.  The reference has no counterpart; declared in include/gpair_nn.h.
//
// Tensors are FP32 [N][C][H][W] with an explicit batch stride, so a channel
// concatenation is a channel offset inside one larger buffer (the encoder
// writes its skip output straight into the decoder's concatenation buffer).
// 3x3 convolutions (zero padding 1, cross-correlation) are implicit GEMMs on
// v_mfma_f32_32x32x16_bf16: BF16 weights and activations (each operand rounded
// to nearest even from its FP32 value as it is loaded), FP32 accumulation, FP32
// bias.  The forward and the data gradient (the transposed convolution) share
// one kernel; weight and bias gradients are a GEMM over the pixels of the batch,
// split in fixed pixel ranges whose partials a second kernel adds in order, so
// training is deterministic.  The master weights and Adam's moments stay FP32.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../../include/gpair_nn.h"
#include "gpair_kernels.h"

#define NN_MSE_BLOCKS 256
#define NN_PT 128          // pixels (columns) per block of the convolution GEMMs: 4 waves x 32
#define NN_WG_TARGET 2048  // blocks the weight-gradient split aims at

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// MFMA operand layout (32x32x16 BF16): lane l holds A[l % 32][8 (l / 32) .. + 7] and
// B[8 (l / 32) .. + 7][l % 32]; accumulator register r holds D[8 (r / 4) + 4 (l / 32) +
// r % 4][l % 32].
__device__ inline bf16x8 nn_pack8(const float (&v)[8])
{
    uint4 u;
    u.x = bf16_bits_rne(v[0]) | ((unsigned)bf16_bits_rne(v[1]) << 16);
    u.y = bf16_bits_rne(v[2]) | ((unsigned)bf16_bits_rne(v[3]) << 16);
    u.z = bf16_bits_rne(v[4]) | ((unsigned)bf16_bits_rne(v[5]) << 16);
    u.w = bf16_bits_rne(v[6]) | ((unsigned)bf16_bits_rne(v[7]) << 16);
    return __builtin_bit_cast(bf16x8, u);
}

// Convolution as a GEMM with rows = output channels r, columns = pixels q and
// K = (c, k) over the Csrc source channels and the 9 taps k = 3 (dy + 1) + dx + 1:
//   BWD = 0 (forward):       dst[r][q] = bias[r] + sum w[r][c][k] src[c][q + (dy, dx)], optional ReLU
//   BWD = 1 (data gradient): dst[r][q] = sum w[c][r][k] src[c][q - (dy, dx)]
// (src = dout, dst = din, the transposed convolution).  Taps outside the image
// read zero.  Block: 4 waves, RT x 32 rows by 128 pixels of one image (blockIdx.z);
// wave v takes pixels 32 v .. + 31 of the block for all RT row tiles.
template <int BWD, int RT>
__global__ __launch_bounds__(256) void k_conv3_gemm(const float *__restrict__ src, long src_bs, int Csrc,
                                                    const float *__restrict__ w, const float *__restrict__ bias,
                                                    float *__restrict__ dst, long dst_bs, int Cdst, int H, int W,
                                                    int relu)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane & 31, lh = lane >> 5;
    const int HW = H * W, K = Csrc * 9, nks = (K + 15) / 16;
    const int q = blockIdx.x * NN_PT + wave * 32 + lr;   // this lane's B column
    const bool qv = q < HW;
    const int y = qv ? q / W : 0, x = qv ? q - (q / W) * W : 0;
    const int row0 = blockIdx.y * 32 * RT;
    const float *s = src + (long)blockIdx.z * src_bs;
    v16f acc[RT];
#pragma unroll
    for (int t = 0; t < RT; t++)
#pragma unroll
        for (int r = 0; r < 16; r++)
            acc[t][r] = 0.f;
    for (int ks = 0; ks < nks; ks++) {
        const int kb = ks * 16 + 8 * lh;
        float bv[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int kk = kb + j;
            float v = 0.f;
            if (qv && kk < K) {
                const int c = kk / 9, k = kk - 9 * c;
                int dy = k / 3 - 1, dx = k - 3 * (k / 3) - 1;
                if (BWD) {
                    dy = -dy;
                    dx = -dx;
                }
                const int yy = y + dy, xx = x + dx;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                    v = s[(long)c * HW + yy * W + xx];
            }
            bv[j] = v;
        }
        const bf16x8 b = nn_pack8(bv);
#pragma unroll
        for (int t = 0; t < RT; t++) {
            const int r = row0 + 32 * t + lr;
            float av[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int kk = kb + j;
                float v = 0.f;
                if (r < Cdst && kk < K) {
                    const int c = kk / 9, k = kk - 9 * c;
                    v = BWD ? w[((long)c * Cdst + r) * 9 + k] : w[(long)r * K + kk];
                }
                av[j] = v;
            }
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nn_pack8(av), b, acc[t], 0, 0, 0);
        }
    }
    if (!qv)
        return;
    float *d = dst + (long)blockIdx.z * dst_bs + q;
#pragma unroll
    for (int t = 0; t < RT; t++)
#pragma unroll
        for (int rr = 0; rr < 16; rr++) {
            const int r = row0 + 32 * t + 8 * (rr >> 2) + 4 * lh + (rr & 3);
            if (r < Cdst) {
                float v = acc[t][rr];
                if (!BWD)
                    v += bias[r];
                if (relu && v < 0.f)
                    v = 0.f;
                d[(long)r * HW] = v;
            }
        }
}

// Weight and bias gradients as a GEMM with rows = output channels co, columns
// = (ci, k) for the Ci * 9 weights of a row plus one column of ones (the bias),
// and K = the pixels of the batch (n, p) flattened:
//   part[s][co][ci 9 + k] = sum over (n, p) in [s chunk, (s + 1) chunk) of
//                           dout[n][co][p] in[n][ci][p + (dy, dx)]
//   part[s][co][Ci 9]     = the same sum of dout[n][co][p]
// Block: 4 waves, 32 rows by 128 columns, split s = blockIdx.z (chunk a multiple of 16).
__global__ __launch_bounds__(256) void k_conv3_wgrad(const float *__restrict__ in, long in_bs, int Ci,
                                                     const float *__restrict__ dout, long dout_bs, int Co, int N,
                                                     int H, int W, int chunk, float *__restrict__ part)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane & 31, lh = lane >> 5;
    const int HW = H * W, K = Ci * 9, ncol = K + 1;
    const int col = blockIdx.x * NN_PT + wave * 32 + lr;   // this lane's B column
    const int co = blockIdx.y * 32 + lr;                   // this lane's A row
    const int cc = col < K ? col / 9 : 0, k = col - 9 * cc;
    const int dy = k / 3 - 1, dx = k - 3 * (k / 3) - 1;
    const long total = (long)N * HW;
    const long beg = (long)blockIdx.z * chunk, end = beg + chunk < total ? beg + chunk : total;
    v16f acc;
#pragma unroll
    for (int r = 0; r < 16; r++)
        acc[r] = 0.f;
    for (long i0 = beg; i0 < end; i0 += 16) {
        float av[8], bv[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const long idx = i0 + 8 * lh + j;
            float a = 0.f, b = 0.f;
            if (idx < end) {
                const int n = (int)(idx / HW), p = (int)(idx - (long)n * HW);
                if (co < Co)
                    a = dout[n * dout_bs + (long)co * HW + p];
                if (col == K) {
                    b = 1.f;
                } else if (col < K) {
                    const int py = p / W, px = p - (p / W) * W;
                    const int yy = py + dy, xx = px + dx;
                    if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                        b = in[n * in_bs + (long)cc * HW + yy * W + xx];
                }
            }
            av[j] = a;
            bv[j] = b;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nn_pack8(av), nn_pack8(bv), acc, 0, 0, 0);
    }
    const int oc = blockIdx.x * NN_PT + wave * 32 + lr;
    if (oc >= ncol)
        return;
    float *pp = part + (long)blockIdx.z * Co * ncol + oc;
#pragma unroll
    for (int rr = 0; rr < 16; rr++) {
        const int r = blockIdx.y * 32 + 8 * (rr >> 2) + 4 * lh + (rr & 3);
        if (r < Co)
            pp[(long)r * ncol] = acc[rr];
    }
}

// dw[co][ci][k] = sum over s of part[s][co][ci 9 + k], db[co] = sum of part[s][co][Ci 9]; s in order
__global__ void k_conv3_wgrad_finish(const float *__restrict__ part, int Ci, int Co, int S, float *__restrict__ dw,
                                     float *__restrict__ db)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ncol = Ci * 9 + 1;
    if (i >= (long)Co * ncol)
        return;
    const long co = i / ncol, c = i - co * ncol;
    float s = 0.f;
    for (int j = 0; j < S; j++)
        s += part[(long)j * Co * ncol + i];
    if (c < ncol - 1)
        dw[co * (ncol - 1) + c] = s;
    else
        db[co] = s;
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
    const unsigned gx = (unsigned)((H * W + NN_PT - 1) / NN_PT);
    if (Co > 32)
        k_conv3_gemm<0, 2><<<dim3(gx, (Co + 63) / 64, N), 256, 0, st>>>(in, in_bs, Ci, w, b, out, out_bs, Co, H, W,
                                                                        relu);
    else
        k_conv3_gemm<0, 1><<<dim3(gx, 1, N), 256, 0, st>>>(in, in_bs, Ci, w, b, out, out_bs, Co, H, W, relu);
}

// splits of the weight-gradient GEMM over the pixels of the batch, and their length
// (a multiple of 16 pixels)
static int wgrad_chunk(int Ci, int Co, int N, int H, int W, int *S_out)
{
    const long total = (long)N * H * W;
    const long blocks = (long)((Ci * 9 + 1 + NN_PT - 1) / NN_PT) * ((Co + 31) / 32);
    long S = (NN_WG_TARGET + blocks - 1) / blocks;
    const long smax = (total + 255) / 256;   // at least 256 pixels per split
    if (S > smax)
        S = smax;
    if (S < 1)
        S = 1;
    long chunk = (total + S - 1) / S;
    chunk = (chunk + 15) / 16 * 16;
    *S_out = (int)((total + chunk - 1) / chunk);
    return (int)chunk;
}

static long wgrad_part_size(int Ci, int Co, int N, int H, int W)
{
    int S;
    (void)wgrad_chunk(Ci, Co, N, H, W, &S);
    return (long)S * Co * (Ci * 9 + 1);
}

// part holds wgrad_part_size(...) floats
static void conv_bwd(hipStream_t st, const float *in, long in_bs, int Ci, const float *w, const float *dout,
                     long dout_bs, int Co, int N, int H, int W, float *din, long din_bs, float *dw, float *db,
                     float *part)
{
    if (din) {
        const unsigned gx = (unsigned)((H * W + NN_PT - 1) / NN_PT);
        if (Ci > 32)
            k_conv3_gemm<1, 2><<<dim3(gx, (Ci + 63) / 64, N), 256, 0, st>>>(dout, dout_bs, Co, w, nullptr, din,
                                                                            din_bs, Ci, H, W, 0);
        else
            k_conv3_gemm<1, 1><<<dim3(gx, 1, N), 256, 0, st>>>(dout, dout_bs, Co, w, nullptr, din, din_bs, Ci, H, W,
                                                               0);
    }
    int S;
    const int chunk = wgrad_chunk(Ci, Co, N, H, W, &S);
    const int ncol = Ci * 9 + 1;
    k_conv3_wgrad<<<dim3((ncol + NN_PT - 1) / NN_PT, (Co + 31) / 32, S), 256, 0, st>>>(in, in_bs, Ci, dout, dout_bs,
                                                                                        Co, N, H, W, chunk, part);
    k_conv3_wgrad_finish<<<nblk((long)Co * ncol), 256, 0, st>>>(part, Ci, Co, S, dw, db);
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
