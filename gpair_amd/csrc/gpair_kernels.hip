// HIP kernels of the MEM-CG path for gfx950.  Host wrappers live in gpair_api.hip.
//
// Forward DFT (image -> visibilities) and adjoint DFT (coefficients -> gradient)
// use the separable kernel exp(2 pi i (a u x - a v y)) = X[p][x] Y[p][y]:
//   forward:  T[p][y] = sum_x X[p][x] I[y][x]   (complex x real GEMM on MFMA)
//             V_p     = (1/F) sum_y Y[p][y] T[p][y]   (fused row reduction)
//   adjoint:  G[y][x] = sum_p Re(a_p Y[p][y]) Re X[p][x] - Im(a_p Y[p][y]) Im X[p][x]
//             (real GEMM with K = 2*nuv on MFMA, split over uv chunks)
// Everything else is HBM-bound elementwise work with FP64 reductions.
#include "gpair_kernels.h"

// ---------------------------------------------------------------------------
// tables
// ---------------------------------------------------------------------------
__global__ void k_tables(const float *__restrict__ u, const float *__restrict__ v, double a, int nuv, int W,
                         float2 *__restrict__ X, float2 *__restrict__ Y)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nuv * W)
        return;
    int p = (int)(i / W), x = (int)(i % W);
    double au = a * (double)u[p], av = a * (double)v[p];
    X[i] = turn_exp(au * (double)x);
    Y[i] = turn_exp(-av * (double)x);
}

// reference GPU layout float2 [W][nuv] (gpair1/trunk/gpair.c:610-636) -> [nuv][W]
__global__ void k_transpose_tables(const float2 *__restrict__ in, int nuv, int W, float2 *__restrict__ out)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nuv * W)
        return;
    int p = (int)(i / W), x = (int)(i % W);
    out[i] = in[(size_t)x * nuv + p];
}

// ---------------------------------------------------------------------------
// image reductions
// ---------------------------------------------------------------------------
// flux F = sum I (gpair1/trunk/gpu.c:860-884), also 1/F as float
__global__ __launch_bounds__(RED_THREADS) void k_flux(const float *__restrict__ img, int n, double *part,
                                                      unsigned *ticket, double *scal, float *invflux)
{
    double v[1] = {0.0};
    for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += gridDim.x * RED_THREADS)
        v[0] += (double)img[i];
    reduce_finish<1>(v, part, ticket, scal + SC_FLUX, invflux);
}

// Gull-Skilling entropy (gpair1/trunk/kernel_entropy_gs.cl:3-26), per-pixel in FP64
__global__ __launch_bounds__(RED_THREADS) void k_entropy(const float *__restrict__ img,
                                                         const float *__restrict__ model, int n, double *part,
                                                         unsigned *ticket, double *scal)
{
    double v[1] = {0.0};
    for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += gridDim.x * RED_THREADS) {
        double I = img[i], M = model[i];
        v[0] += (I > 0.0 && M > 0.0) ? (I - M - I * log(I / M)) : -M;
    }
    reduce_finish<1>(v, part, ticket, scal + SC_ENT, nullptr);
}

// sum a*b (gpair1/trunk/kernel_scalarprod.cl:2-9 + reduction), FP64 accumulation
__global__ __launch_bounds__(RED_THREADS) void k_dot(const float *__restrict__ a, const float *__restrict__ b,
                                                     int n, double *part, unsigned *ticket, double *out)
{
    double v[1] = {0.0};
    for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += gridDim.x * RED_THREADS)
        v[0] += (double)a[i] * (double)b[i];
    reduce_finish<1>(v, part, ticket, out, nullptr);
}

// ---------------------------------------------------------------------------
// forward DFT: one workgroup = 32 uv points x all W rows, K = x in chunks of 32
// MFMA v_mfma_f32_16x16x4_f32: lane l holds A[l%16][l/16], B[l/16][l%16],
// D[4*(l/16)+r][l%16].
// ---------------------------------------------------------------------------
#define FWD_PT 32
#define FWD_KC 32
#define FWD_SA (2 * FWD_PT + 16)

template <int NB>
__global__ __launch_bounds__(256) void k_vis_fwd_f32(const float *__restrict__ img, const float2 *__restrict__ X,
                                                     const float2 *__restrict__ Y, int nuv,
                                                     const float *__restrict__ invflux, float2 *__restrict__ vis)
{
    constexpr int W = 64 * NB;
    constexpr int SB = W + 16;
    __shared__ float sA[FWD_KC * FWD_SA];
    __shared__ float sB[FWD_KC * SB];
    __shared__ float2 red[4][FWD_PT];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int p0 = blockIdx.x * FWD_PT;
    v4f acc[4][NB];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int n = 0; n < NB; n++)
            acc[m][n] = v4f{0.f, 0.f, 0.f, 0.f};

    for (int x0 = 0; x0 < W; x0 += FWD_KC) {
        __syncthreads();
        {   // A tile: rows 0..31 = Re X[p0+pp][x], rows 32..63 = Im
            const int pp = tid >> 3, kq = (tid & 7) * 4, p = p0 + pp;
            float4 q0 = make_float4(0.f, 0.f, 0.f, 0.f), q1 = q0;
            if (p < nuv) {
                const float4 *src = reinterpret_cast<const float4 *>(X + (size_t)p * W + x0 + kq);
                q0 = src[0];
                q1 = src[1];
            }
            sA[(kq + 0) * FWD_SA + pp] = q0.x;
            sA[(kq + 0) * FWD_SA + FWD_PT + pp] = q0.y;
            sA[(kq + 1) * FWD_SA + pp] = q0.z;
            sA[(kq + 1) * FWD_SA + FWD_PT + pp] = q0.w;
            sA[(kq + 2) * FWD_SA + pp] = q1.x;
            sA[(kq + 2) * FWD_SA + FWD_PT + pp] = q1.y;
            sA[(kq + 3) * FWD_SA + pp] = q1.z;
            sA[(kq + 3) * FWD_SA + FWD_PT + pp] = q1.w;
        }
        for (int y = tid; y < W; y += 256) {   // B tile: sB[k][y] = I[y][x0+k]
            const float4 *src = reinterpret_cast<const float4 *>(img + (size_t)y * W + x0);
#pragma unroll
            for (int q = 0; q < FWD_KC / 4; q++) {
                float4 t = src[q];
                sB[(4 * q + 0) * SB + y] = t.x;
                sB[(4 * q + 1) * SB + y] = t.y;
                sB[(4 * q + 2) * SB + y] = t.z;
                sB[(4 * q + 3) * SB + y] = t.w;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k4 = 0; k4 < FWD_KC / 4; k4++) {
            const int kr = k4 * 4 + (lane >> 4);
            float a[4], b[NB];
#pragma unroll
            for (int m = 0; m < 4; m++)
                a[m] = sA[kr * FWD_SA + m * 16 + (lane & 15)];
#pragma unroll
            for (int n = 0; n < NB; n++)
                b[n] = sB[kr * SB + wave * 16 * NB + n * 16 + (lane & 15)];
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int n = 0; n < NB; n++)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[m], b[n], acc[m][n], 0, 0, 0);
        }
    }

    // epilogue: V_p = sum_y Y[p][y] (Re T + i Im T); block rows m=0,1 are Re, m=2,3 Im
    float vr[2][4], vi[2][4];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int pl = m * 16 + 4 * (lane >> 4) + r, p = p0 + pl;
            float sr = 0.f, si = 0.f;
            if (p < nuv) {
#pragma unroll
                for (int n = 0; n < NB; n++) {
                    const int y = wave * 16 * NB + n * 16 + (lane & 15);
                    const float2 yv = Y[(size_t)p * W + y];
                    const float tr = acc[m][n][r], ti = acc[m + 2][n][r];
                    sr += tr * yv.x - ti * yv.y;
                    si += tr * yv.y + ti * yv.x;
                }
            }
            vr[m][r] = sr;
            vi[m][r] = si;
        }
#pragma unroll
    for (int o = 1; o < 16; o <<= 1)
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                vr[m][r] += __shfl_xor(vr[m][r], o, 64);
                vi[m][r] += __shfl_xor(vi[m][r], o, 64);
            }
    if ((lane & 15) == 0) {
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                red[wave][m * 16 + 4 * (lane >> 4) + r] = make_float2(vr[m][r], vi[m][r]);
    }
    __syncthreads();
    if (tid < FWD_PT && p0 + tid < nuv) {
        float2 s = red[0][tid];
#pragma unroll
        for (int w = 1; w < 4; w++) {
            s.x += red[w][tid].x;
            s.y += red[w][tid].y;
        }
        const float f = invflux ? *invflux : 1.0f;   // no scale: unnormalised A I
        vis[p0 + tid] = make_float2(s.x * f, s.y * f);
    }
}

// ---------------------------------------------------------------------------
// model data + chi2 (gpair1/trunk/kernel_pow_spec.cl, kernel_bispec.cl, kernel_chi2.cl)
// ---------------------------------------------------------------------------
__device__ inline float2 leg_vis(const float2 *vis, int idx, int conj_bit)
{
    float2 v = vis[idx];
    return conj_bit ? cconj(v) : v;
}

__global__ __launch_bounds__(RED_THREADS) void k_model_chi2(const float2 *__restrict__ vis,
                                                            const float *__restrict__ data,
                                                            const float *__restrict__ err,
                                                            const float2 *__restrict__ phasor,
                                                            const int4 *__restrict__ tri, int npow, int nbis,
                                                            float *__restrict__ mock, double *part,
                                                            unsigned *ticket, double *scal)
{
    double v[1] = {0.0};
    const int total = npow + nbis;
    for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < total; i += gridDim.x * RED_THREADS) {
        if (i < npow) {
            const float2 V = vis[i];
            const float m = V.x * V.x + V.y * V.y;
            mock[i] = m;
            const double r = ((double)m - (double)data[i]) * (double)err[i];
            v[0] += r * r;
        } else {
            const int t = i - npow;
            const int4 tr = tri[t];
            float2 T = cmul(leg_vis(vis, tr.x, tr.w & 1), leg_vis(vis, tr.y, tr.w & 2));
            T = cmul(T, leg_vis(vis, tr.z, tr.w & 4));
            T = cmul(T, phasor[t]);
            const int k = npow + 2 * t;
            mock[k] = T.x;
            mock[k + 1] = T.y;
            const double r0 = ((double)T.x - (double)data[k]) * (double)err[k];
            const double r1 = ((double)T.y - (double)data[k + 1]) * (double)err[k + 1];
            v[0] += r0 * r0 + r1 * r1;
        }
    }
    reduce_finish<1>(v, part, ticket, scal + SC_CHI2, nullptr);
}

// ---------------------------------------------------------------------------
// adjoint coefficients a_p (one thread per uv point, T3 legs gathered via CSR)
//   V2: a_k = 4 e_k^2 (m_k - d_k) conj(V_k)
//   T3 leg j of t: b = c_t rho_t V'_{j+1} V'_{j+2}, c_t = 2e_r^2(m_r-d_r) - i 2e_i^2 m_i,
//                  a_p += b, or conj(b) for a conjugated leg
// c0 = Re sum_p a_p V_p
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(RED_THREADS) void k_coef(const float2 *__restrict__ vis,
                                                      const float *__restrict__ mock,
                                                      const float *__restrict__ data,
                                                      const float *__restrict__ err,
                                                      const float2 *__restrict__ phasor,
                                                      const int4 *__restrict__ tri,
                                                      const int *__restrict__ csr_ptr,
                                                      const int *__restrict__ csr_ent, int npow, int nuv,
                                                      float2 *__restrict__ coef, double *part, unsigned *ticket,
                                                      double *scal)
{
    double v[1] = {0.0};
    for (int p = blockIdx.x * RED_THREADS + threadIdx.x; p < nuv; p += gridDim.x * RED_THREADS) {
        const float2 Vp = vis[p];
        float2 a = make_float2(0.f, 0.f);
        if (p < npow) {
            const float e = err[p];
            const float w = 4.0f * e * e * (mock[p] - data[p]);
            a = make_float2(w * Vp.x, -w * Vp.y);
        }
        for (int q = csr_ptr[p]; q < csr_ptr[p + 1]; q++) {
            const int ent = csr_ent[q];
            const int t = ent / 3, j = ent - 3 * t;
            const int4 tr = tri[t];
            const int idx[3] = {tr.x, tr.y, tr.z};
            const int k = npow + 2 * t;
            const float er = err[k], ei = err[k + 1];
            const float2 ct = make_float2(2.0f * er * er * (mock[k] - data[k]), -2.0f * ei * ei * mock[k + 1]);
            const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            float2 b = cmul(ct, phasor[t]);
            b = cmul(b, leg_vis(vis, idx[j1], (tr.w >> j1) & 1));
            b = cmul(b, leg_vis(vis, idx[j2], (tr.w >> j2) & 1));
            if ((tr.w >> j) & 1)
                b = cconj(b);
            a.x += b.x;
            a.y += b.y;
        }
        coef[p] = a;
        v[0] += (double)a.x * (double)Vp.x - (double)a.y * (double)Vp.y;
    }
    reduce_finish<1>(v, part, ticket, scal + SC_C0, nullptr);
}

// ---------------------------------------------------------------------------
// adjoint DFT partials: one workgroup = one 64x64 (y, x) tile of one uv split;
// its 4 waves take interleaved uv pairs and are summed through LDS in order.
// K index of the MFMA: k = 2*(pair slot) + c, c = 0 (Re) / 1 (Im):
//   A[y][k] = c ? -Im(a_p Y[p][y]) : Re(a_p Y[p][y]),   B[k][x] = c ? Im X[p][x] : Re X[p][x]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adj_f32(const float2 *__restrict__ coef, const float2 *__restrict__ X,
                                                 const float2 *__restrict__ Y, int nuv, int W, int p_per_split,
                                                 float *__restrict__ part)
{
    __shared__ float red[3][64 * 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tilesx = W / 64;
    const int y0 = (blockIdx.x / tilesx) * 64, x0 = (blockIdx.x % tilesx) * 64;
    const int pstart = blockIdx.y * p_per_split;
    const int pend = min(pstart + p_per_split, nuv);
    const int kk = lane >> 4, c = kk & 1, li = lane & 15;
    v4f acc[4][4];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int n = 0; n < 4; n++)
            acc[m][n] = v4f{0.f, 0.f, 0.f, 0.f};
    const int npairs = (pend - pstart + 1) / 2;
    for (int q = wave; q < npairs; q += 4) {
        const int p = pstart + 2 * q + (kk >> 1);
        float av[4], bv[4];
        if (p < pend) {
            const float2 a = coef[p];
            const float2 *yr = Y + (size_t)p * W + y0 + li;
            const float2 *xr = X + (size_t)p * W + x0 + li;
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const float2 ay = cmul(a, yr[m * 16]);
                av[m] = c ? -ay.y : ay.x;
            }
#pragma unroll
            for (int n = 0; n < 4; n++) {
                const float2 xv = xr[n * 16];
                bv[n] = c ? xv.y : xv.x;
            }
        } else {
#pragma unroll
            for (int m = 0; m < 4; m++) {
                av[m] = 0.f;
                bv[m] = 0.f;
            }
        }
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int n = 0; n < 4; n++)
                acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[m], bv[n], acc[m][n], 0, 0, 0);
    }
    if (wave > 0) {
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int n = 0; n < 4; n++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    red[wave - 1][(m * 16 + 4 * kk + r) * 64 + n * 16 + li] = acc[m][n][r];
    }
    __syncthreads();
    if (wave == 0) {
        float *out = part + (size_t)blockIdx.y * W * W;
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int n = 0; n < 4; n++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int o = (m * 16 + 4 * kk + r) * 64 + n * 16 + li;
                    float s = acc[m][n][r];
                    s += red[0][o];
                    s += red[1][o];
                    s += red[2][o];
                    out[(size_t)(y0 + m * 16 + 4 * kk + r) * W + x0 + n * 16 + li] = s;
                }
    }
}

// data gradient = (sum over splits - c0) / F, splits summed in order in FP64
__global__ void k_grad_finish(const float *__restrict__ part, int nsplit, int n, const double *__restrict__ scal,
                              float *__restrict__ gdata)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    double s = 0.0;
    for (int k = 0; k < nsplit; k++)
        s += (double)part[(size_t)k * n + i];
    gdata[i] = (float)((s - scal[SC_C0]) / scal[SC_FLUX]);
}

// ---------------------------------------------------------------------------
// elementwise image kernels
// ---------------------------------------------------------------------------
// entropy gradient, stale value kept where I <= 0 or M <= 0 (kernel_entropy_gs_grad.cl:3-24)
__global__ void k_entropy_grad(const float *__restrict__ img, const float *__restrict__ model, int n,
                               float *__restrict__ g)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float I = img[i], M = model[i];
    if (I > 0.f && M > 0.f)
        g[i] = -logf(I / M);
}

// out = gdata - alpha * gent (kernel_criterion_grad.cl:2-15)
__global__ void k_crit_grad(const float *__restrict__ gd, const float *__restrict__ ge, float alpha, int n,
                            float *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = gd[i] - alpha * ge[i];
}

// d = beta*d - g (kernel_descent_dir.cl:2-14)
__global__ void k_descent(float *__restrict__ d, const float *__restrict__ g, float beta, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        d[i] = beta * d[i] - g[i];
}

// out = max(img + s*d, minval) (kernel_update_tempimage.cl / kernel_update_image.cl);
// out may alias img
__global__ void k_step(const float *img, const float *__restrict__ d, float s, float minval, int n, float *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float t = img[i] + s * d[i];
        out[i] = t < minval ? minval : t;
    }
}

// ---------------------------------------------------------------------------
// accelerated loop: preconditioned direction and batched line search
// ---------------------------------------------------------------------------
// z = P g with P = image (MEM metric) or 1, times an optional per-pixel factor;
// out[0] = z.g, out[1] = z.g_old
__global__ __launch_bounds__(RED_THREADS) void k_pcg_precond(const float *__restrict__ img,
                                                             const float *__restrict__ g,
                                                             const float *__restrict__ g_old,
                                                             const float *__restrict__ pfac, int precond, int n,
                                                             float *__restrict__ z, double *part, unsigned *ticket,
                                                             double *out)
{
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += gridDim.x * RED_THREADS) {
        float pf = precond ? img[i] : 1.0f;
        if (pfac)
            pf *= pfac[i];
        const float zi = pf * g[i];
        z[i] = zi;
        v[0] += (double)zi * (double)g[i];
        v[1] += (double)zi * (double)g_old[i];
    }
    reduce_finish<2>(v, part, ticket, out, nullptr);
}

// d = -z + beta d, projected to zero where the pixel sits on the floor and d < 0;
// out[0] = d.g, out[1] = sum d, out[2] = sum |d|
__global__ __launch_bounds__(RED_THREADS) void k_pcg_dir(const float *__restrict__ img, const float *__restrict__ z,
                                                         const float *__restrict__ g, float beta, float minval, int n,
                                                         float *__restrict__ d, double *part, unsigned *ticket,
                                                         double *out)
{
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += gridDim.x * RED_THREADS) {
        float di = -z[i] + beta * d[i];
        if (img[i] <= minval && di < 0.f)
            di = 0.f;
        d[i] = di;
        v[0] += (double)di * (double)g[i];
        v[1] += (double)di;
        v[2] += fabs((double)di);
    }
    reduce_finish<3>(v, part, ticket, out, nullptr);
}

// candidate 0 is s = 0 (the linearised current image, the common reference)
__device__ inline double ls_step(double s_ref, int k)
{
    return k == 0 ? 0.0 : s_ref * exp2(0.5 * (double)(k - LS_TOP));
}

// chi2 of the visibilities (U_I + s_k U_d) / (F_I + s_k F_d) for all LS_K candidates,
// U_I = F_I V_I; out[k] = chi2_k
__global__ __launch_bounds__(RED_THREADS) void k_ls_chi2(const float2 *__restrict__ vis,
                                                         const float2 *__restrict__ Ud,
                                                         const double *__restrict__ scal, double s_ref,
                                                         const float *__restrict__ data,
                                                         const float *__restrict__ err,
                                                         const float2 *__restrict__ phasor,
                                                         const int4 *__restrict__ tri, int npow, int nbis,
                                                         double *part, unsigned *ticket, double *out)
{
    double v[LS_K];
#pragma unroll
    for (int k = 0; k < LS_K; k++)
        v[k] = 0.0;
    const double FI = scal[SC_FLUX], Fd = scal[SC_FD];
    float sk[LS_K], ik[LS_K];
#pragma unroll
    for (int k = 0; k < LS_K; k++) {
        const double s = ls_step(s_ref, k);
        sk[k] = (float)s;
        ik[k] = (float)(1.0 / (FI + s * Fd));
    }
    const float fI = (float)FI;
    const int total = npow + nbis;
    for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < total; i += gridDim.x * RED_THREADS) {
        if (i < npow) {
            const float2 a = vis[i], b = Ud[i];
            const float ax = a.x * fI, ay = a.y * fI, d0 = data[i], e = err[i];
#pragma unroll
            for (int k = 0; k < LS_K; k++) {
                const float x = (ax + sk[k] * b.x) * ik[k], y = (ay + sk[k] * b.y) * ik[k];
                const double r = ((double)(x * x + y * y) - (double)d0) * (double)e;
                v[k] += r * r;
            }
        } else {
            const int t = i - npow;
            const int4 tr = tri[t];
            const float2 a0 = vis[tr.x], a1 = vis[tr.y], a2 = vis[tr.z];
            const float2 b0 = Ud[tr.x], b1 = Ud[tr.y], b2 = Ud[tr.z];
            const float2 rho = phasor[t];
            const int kk = npow + 2 * t;
            const float dr = data[kk], di = data[kk + 1], er = err[kk], ei = err[kk + 1];
            const float c0 = (tr.w & 1) ? -1.f : 1.f, c1 = (tr.w & 2) ? -1.f : 1.f, c2 = (tr.w & 4) ? -1.f : 1.f;
#pragma unroll
            for (int k = 0; k < LS_K; k++) {
                const float s = sk[k], iv = ik[k];
                const float2 l0 = make_float2((a0.x * fI + s * b0.x) * iv, c0 * (a0.y * fI + s * b0.y) * iv);
                const float2 l1 = make_float2((a1.x * fI + s * b1.x) * iv, c1 * (a1.y * fI + s * b1.y) * iv);
                const float2 l2 = make_float2((a2.x * fI + s * b2.x) * iv, c2 * (a2.y * fI + s * b2.y) * iv);
                const float2 T = cmul(cmul(cmul(l0, l1), l2), rho);
                const double r0 = ((double)T.x - (double)dr) * (double)er;
                const double r1 = ((double)T.y - (double)di) * (double)ei;
                v[k] += r0 * r0 + r1 * r1;
            }
        }
    }
    reduce_finish<LS_K>(v, part, ticket, out, nullptr);
}

// entropy of max(I + s_k d, minval) for all candidates; out[k] = S_k
__global__ __launch_bounds__(RED_THREADS) void k_ls_entropy(const float *__restrict__ img,
                                                            const float *__restrict__ d,
                                                            const float *__restrict__ model, double s_ref,
                                                            double minval, int n, double *part, unsigned *ticket,
                                                            double *out)
{
    double v[LS_K], sk[LS_K];
#pragma unroll
    for (int k = 0; k < LS_K; k++) {
        v[k] = 0.0;
        sk[k] = ls_step(s_ref, k);
    }
    for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += gridDim.x * RED_THREADS) {
        const double I = img[i], dd = d[i], M = model[i];
#pragma unroll
        for (int k = 0; k < LS_K; k++) {
            double t = I + sk[k] * dd;
            if (t < minval)
                t = minval;
            v[k] += (t > 0.0 && M > 0.0) ? (t - M - t * log(t / M)) : -M;
        }
    }
    reduce_finish<LS_K>(v, part, ticket, out, nullptr);
}

// explicit instantiations for the supported widths (W = 64 * NB)
template __global__ void k_vis_fwd_f32<1>(const float *, const float2 *, const float2 *, int, const float *, float2 *);
template __global__ void k_vis_fwd_f32<2>(const float *, const float2 *, const float2 *, int, const float *, float2 *);
template __global__ void k_vis_fwd_f32<3>(const float *, const float2 *, const float2 *, int, const float *, float2 *);
template __global__ void k_vis_fwd_f32<4>(const float *, const float2 *, const float2 *, int, const float *, float2 *);
template __global__ void k_vis_fwd_f32<8>(const float *, const float2 *, const float2 *, int, const float *, float2 *);
