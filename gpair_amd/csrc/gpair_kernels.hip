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
// flux as a partitioned reduction: the same bits for any grid, alone or fused with k_coef
__device__ __forceinline__ void flux_part(const float *__restrict__ img, int n, int bx, int nbx, double *part,
                                          unsigned *ticket, double *scal, float *invflux)
{
    auto item = [&](int i, double (&v)[1]) { v[0] += (double)img[i]; };
    partitioned_reduce_at<1>(n, 2 * RED_THREADS, item, part, ticket, scal + SC_FLUX, bx, nbx, invflux);
}

__global__ __launch_bounds__(RED_THREADS) void k_flux(const float *__restrict__ img, int n, double *part,
                                                      unsigned *ticket, double *scal, float *invflux)
{
    flux_part(img, n, (int)blockIdx.x, (int)gridDim.x, part, ticket, scal, invflux);
}

// Gull-Skilling entropy (gpair1/trunk/kernel_entropy_gs.cl:3-26), per-pixel in FP64
__device__ __forceinline__ void k_entropy_body(const int BX, const int GX, const float *__restrict__ img,
                                                         const float *__restrict__ model, int n, double *part,
                                                         unsigned *ticket, double *scal){
    double v[1] = {0.0};
    for (int i = BX * RED_THREADS + threadIdx.x; i < n; i += GX * RED_THREADS) {
        double I = img[i], M = model[i];
        v[0] += (I > 0.0 && M > 0.0) ? (I - M - I * log(I / M)) : -M;
    }
    reduce_finish_at<1>(v, part, ticket, scal + SC_ENT, nullptr, BX, GX);
}

__global__ __launch_bounds__(RED_THREADS) void k_entropy(const float *__restrict__ img,
                                                         const float *__restrict__ model, int n, double *part,
                                                         unsigned *ticket, double *scal)
{
    k_entropy_body((int)blockIdx.x, (int)gridDim.x, img, model, n, part, ticket, scal);
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
                                                     const float *__restrict__ invflux, float2 *__restrict__ vis,
                                                     const int *__restrict__ run_if)
{
    if (run_if && *run_if == 0)   // a commit whose change DFT the sparse correction already gave
        return;
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

__device__ __forceinline__ void k_model_chi2_body(const int BX, const int GX, const float2 *__restrict__ vis,
                                                            const float *__restrict__ data,
                                                            const float *__restrict__ err,
                                                            const float2 *__restrict__ phasor,
                                                            const int4 *__restrict__ tri, int npow, int nbis,
                                                            float *__restrict__ mock, double *part,
                                                            unsigned *ticket, double *scal){
    double v[1] = {0.0};
    const int total = npow + nbis;
    for (int i = BX * RED_THREADS + threadIdx.x; i < total; i += GX * RED_THREADS) {
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
    reduce_finish_at<1>(v, part, ticket, scal + SC_CHI2, nullptr, BX, GX);
}

__global__ __launch_bounds__(RED_THREADS) void k_model_chi2(const float2 *__restrict__ vis,
                                                            const float *__restrict__ data,
                                                            const float *__restrict__ err,
                                                            const float2 *__restrict__ phasor,
                                                            const int4 *__restrict__ tri, int npow, int nbis,
                                                            float *__restrict__ mock, double *part,
                                                            unsigned *ticket, double *scal)
{
    k_model_chi2_body((int)blockIdx.x, (int)gridDim.x, vis, data, err, phasor, tri, npow, nbis, mock, part, ticket, scal);
}

// ---------------------------------------------------------------------------
// adjoint coefficients a_p (one thread per uv point, T3 legs gathered via CSR)
//   V2: a_k = 4 e_k^2 (m_k - d_k) conj(V_k)
//   T3 leg j of t: b = c_t rho_t V'_{j+1} V'_{j+2}, c_t = 2e_r^2(m_r-d_r) - i 2e_i^2 m_i,
//                  a_p += b, or conj(b) for a conjugated leg
// c0 = Re sum_p a_p V_p
// ---------------------------------------------------------------------------
// Two launches: k_t3_contrib, one thread per T3, writes the three legs' terms b
// (contrib[3 t + j]), each gathering its three visibilities once; then coef_part,
// one uv point per thread and partition, adds its V2 term and its legs' b in CSR
// order.  The loads of a point's entries are independent of each other, where the
// single gather pass waited on csr -> tri -> vis for every leg in turn.
#define COEF_SPAN RED_THREADS

__device__ __forceinline__ void k_t3_contrib_body(const int BX, const int GX, const float2 *__restrict__ vis,
                                                  const float *__restrict__ mock, const float *__restrict__ data,
                                                  const float *__restrict__ err, const float2 *__restrict__ phasor,
                                                  const int4 *__restrict__ tri, int npow, int nbis,
                                                  float2 *__restrict__ contrib)
{
    for (int t = BX * 256 + (int)threadIdx.x; t < nbis; t += GX * 256) {
        const int4 tr = tri[t];
        const float2 V[3] = {leg_vis(vis, tr.x, tr.w & 1), leg_vis(vis, tr.y, tr.w & 2), leg_vis(vis, tr.z, tr.w & 4)};
        const int k = npow + 2 * t;
        const float er = err[k], ei = err[k + 1];
        const float2 ct = make_float2(2.0f * er * er * (mock[k] - data[k]), -2.0f * ei * ei * mock[k + 1]);
        const float2 cr = cmul(ct, phasor[t]);
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float2 b = cmul(cr, V[(j + 1) % 3]);
            b = cmul(b, V[(j + 2) % 3]);
            if ((tr.w >> j) & 1)
                b = cconj(b);
            contrib[3 * t + j] = b;
        }
    }
}

__global__ __launch_bounds__(256) void k_t3_contrib(const float2 *__restrict__ vis, const float *__restrict__ mock,
                                                    const float *__restrict__ data, const float *__restrict__ err,
                                                    const float2 *__restrict__ phasor, const int4 *__restrict__ tri,
                                                    int npow, int nbis, float2 *__restrict__ contrib)
{
    k_t3_contrib_body((int)blockIdx.x, (int)gridDim.x, vis, mock, data, err, phasor, tri, npow, nbis, contrib);
}

__device__ __forceinline__ void coef_part(const float2 *__restrict__ vis, const float *__restrict__ mock,
                                          const float *__restrict__ data, const float *__restrict__ err,
                                          const float2 *__restrict__ contrib, const int *__restrict__ csr_ptr,
                                          const int *__restrict__ csr_ent, int npow, int nuv,
                                          float2 *__restrict__ coef, double *part, unsigned *ticket, double *scal,
                                          int bx, int nbx)
{
    auto item = [&](int p, double (&v)[1]) {
        const float2 Vp = vis[p];
        float2 a = make_float2(0.f, 0.f);
        if (p < npow) {
            const float e = err[p];
            const float w = 4.0f * e * e * (mock[p] - data[p]);
            a = make_float2(w * Vp.x, -w * Vp.y);
        }
        const int q0 = csr_ptr[p], q1 = csr_ptr[p + 1];
#pragma unroll 4
        for (int q = q0; q < q1; q++) {   // loads in flight four at a time, added in order
            const float2 b = contrib[csr_ent[q]];
            a.x += b.x;
            a.y += b.y;
        }
        coef[p] = a;
        v[0] += (double)a.x * (double)Vp.x - (double)a.y * (double)Vp.y;
    };
    partitioned_reduce_at<1>(nuv, COEF_SPAN, item, part, ticket, scal + SC_C0, bx, nbx);
}

__global__ __launch_bounds__(RED_THREADS) void k_coef(const float2 *__restrict__ vis,
                                                      const float *__restrict__ mock,
                                                      const float *__restrict__ data,
                                                      const float *__restrict__ err,
                                                      const float2 *__restrict__ contrib,
                                                      const int *__restrict__ csr_ptr,
                                                      const int *__restrict__ csr_ent, int npow, int nuv,
                                                      float2 *__restrict__ coef, double *part, unsigned *ticket,
                                                      double *scal)
{
    coef_part(vis, mock, data, err, contrib, csr_ptr, csr_ent, npow, nuv, coef, part, ticket, scal,
              (int)blockIdx.x, (int)gridDim.x);
}

// k_coef and k_flux in one launch: block columns [0, nb_coef) take the
// coefficients, the rest the flux of img (its own partials at part + RED_PARTS and
// ticket + 1); each result is that of the kernel alone
__global__ __launch_bounds__(RED_THREADS) void k_coef_flux(const float2 *__restrict__ vis,
                                                           const float *__restrict__ mock,
                                                           const float *__restrict__ data,
                                                           const float *__restrict__ err,
                                                           const float2 *__restrict__ contrib,
                                                           const int *__restrict__ csr_ptr,
                                                           const int *__restrict__ csr_ent, int npow, int nuv,
                                                           float2 *__restrict__ coef, int nb_coef,
                                                           const float *__restrict__ img, int n, double *part,
                                                           unsigned *ticket, double *scal, float *invflux)
{
    const int bx = (int)blockIdx.x;
    if (bx < nb_coef)
        coef_part(vis, mock, data, err, contrib, csr_ptr, csr_ent, npow, nuv, coef, part, ticket, scal, bx,
                  nb_coef);
    else
        flux_part(img, n, bx - nb_coef, (int)gridDim.x - nb_coef, part + RED_PARTS, ticket + 1, scal, invflux);
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

// data gradient = (sum over splits - c0) / F in FP64.  A block takes 64 pixels; its
// four waves sum a quarter of the splits each, in order, and the quarters are added
// in order: a fixed grouping, so the result is the same on every run.
#define GF_PIX 64
__global__ __launch_bounds__(256) void k_grad_finish(const float *__restrict__ part, int nsplit, int n,
                                                     const double *__restrict__ scal, float *__restrict__ gdata)
{
    __shared__ double sh[4][GF_PIX];
    const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * GF_PIX + lane;
    const int per = (nsplit + 3) / 4;
    const int k0 = q * per, k1 = min(nsplit, k0 + per);
    double s = 0.0;
    if (i < n) {
        int k = k0;
        for (; k + 8 <= k1; k += 8) {
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; j++)
                v[j] = part[(size_t)(k + j) * n + i];
#pragma unroll
            for (int j = 0; j < 8; j++)
                s += (double)v[j];
        }
        for (; k < k1; k++)
            s += (double)part[(size_t)k * n + i];
    }
    sh[q][lane] = s;
    __syncthreads();
    if (q == 0 && i < n) {
        const double t = ((sh[0][lane] + sh[1][lane]) + sh[2][lane]) + sh[3][lane];
        gdata[i] = (float)((t - scal[SC_C0]) / scal[SC_FLUX]);
    }
}

// k_grad_finish, then per pixel the entropy gradient (stale value kept where I <= 0
// or M <= 0, as k_entropy_grad), the criterion gradient g = gdata - alpha gent (as
// k_crit_grad) and the preconditioned gradient z = P g (P = image or 1, times pfac
// when set): one launch instead of four in the accelerated loop.  Block b stores its
// pixels' z.g and z.g_old in zpart[2b], zpart[2b + 1] (plain stores, no ticket);
// k_pcg_dir adds them up.
__device__ __forceinline__ void k_grad_finish_crit_body(const int BX, const int GX, const float *__restrict__ part, int nsplit, int n,
                                                          const double *__restrict__ scal,
                                                          const float *__restrict__ img,
                                                          const float *__restrict__ model, float alpha,
                                                          float *__restrict__ gdata, float *__restrict__ gent,
                                                          float *__restrict__ g, const float *__restrict__ g_old,
                                                          const float *__restrict__ pfac, int precond,
                                                          float *__restrict__ z, double *__restrict__ zpart){
    __shared__ double sh[4][GF_PIX];
    const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = BX * GF_PIX + lane;
    const int per = (nsplit + 3) / 4;
    const int k0 = q * per, k1 = min(nsplit, k0 + per);
    double s = 0.0;
    if (i < n) {
        int k = k0;
        for (; k + 8 <= k1; k += 8) {
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; j++)
                v[j] = part[(size_t)(k + j) * n + i];
#pragma unroll
            for (int j = 0; j < 8; j++)
                s += (double)v[j];
        }
        for (; k < k1; k++)
            s += (double)part[(size_t)k * n + i];
    }
    sh[q][lane] = s;
    __syncthreads();
    if (q != 0)
        return;
    double zg = 0.0, zgo = 0.0;
    if (i < n) {
        const double t = ((sh[0][lane] + sh[1][lane]) + sh[2][lane]) + sh[3][lane];
        const float gd = (float)((t - scal[SC_C0]) / scal[SC_FLUX]);
        gdata[i] = gd;
        const float I = img[i], M = model[i];
        float ge = gent[i];
        if (I > 0.f && M > 0.f) {
            ge = -logf(I / M);
            gent[i] = ge;
        }
        const float gi = gd - alpha * ge;
        g[i] = gi;
        float pf = precond ? I : 1.0f;
        if (pfac)
            pf *= pfac[i];
        const float zi = pf * gi;
        z[i] = zi;
        zg = (double)zi * (double)gi;
        zgo = (double)zi * (double)g_old[i];
    }
    zg = wave_sum(zg);
    zgo = wave_sum(zgo);
    if (lane == 0) {
        zpart[2 * BX] = zg;
        zpart[2 * BX + 1] = zgo;
    }
}

__global__ __launch_bounds__(256) void k_grad_finish_crit(const float *__restrict__ part, int nsplit, int n,
                                                          const double *__restrict__ scal,
                                                          const float *__restrict__ img,
                                                          const float *__restrict__ model, float alpha,
                                                          float *__restrict__ gdata, float *__restrict__ gent,
                                                          float *__restrict__ g, const float *__restrict__ g_old,
                                                          const float *__restrict__ pfac, int precond,
                                                          float *__restrict__ z, double *__restrict__ zpart)
{
    k_grad_finish_crit_body((int)blockIdx.x, (int)gridDim.x, part, nsplit, n, scal, img, model, alpha, gdata, gent, g, g_old, pfac, precond, z, zpart);
}

// ---------------------------------------------------------------------------
// adjoint DFT in split precision (BF16x3) on v_mfma_f32_32x32x16_bf16.
// G[y][x] = sum_k A[y][k] B[k][x], k = 2 p + c:
//   A[y][2p] = Re(a_p Y[p][y]), A[y][2p+1] = -Im(a_p Y[p][y])   (built per call in LDS)
//   B[2p][x] = Re X[p][x],      B[2p+1][x] = Im X[p][x]          (static split table [x][p][2])
// One workgroup = a 64 (y) x 128 (x) tile of one uv split; wave w owns columns
// 32w..32w+31 and both 32-row blocks.  Every round covers 16 uv points (two
// MFMA k-steps): the 256 threads build the 64 x 32 A tile (hi and lo planes)
// in LDS, then each wave runs 2 x 2 x 3 MFMAs.
// ---------------------------------------------------------------------------
#define AB3_YT 64
#define AB3_XT 128
#define AB3_PR 16                 // uv points per round
#define AB3_STRIDE (2 * AB3_PR + 8) // bf16 per A row in LDS: 32 k + 8 padding

// X float2 [nuv][W] -> Xt hi/lo bf16 [W][nuvp][2] (x-major, (Re, Im) per point),
// points nuv..nuvp-1 zero
__global__ void k_split_tables_t(const float2 *__restrict__ X, int nuv, int nuvp, int W,
                                 unsigned short *__restrict__ Xth, unsigned short *__restrict__ Xtl)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nuvp * W)
        return;
    const int x = (int)(i / nuvp), p = (int)(i % nuvp);
    float2 v = make_float2(0.f, 0.f);
    if (p < nuv)
        v = X[(size_t)p * W + x];
    unsigned short h, l;
    split_bf16(v.x, h, l);
    Xth[2 * i] = h;
    Xtl[2 * i] = l;
    split_bf16(v.y, h, l);
    Xth[2 * i + 1] = h;
    Xtl[2 * i + 1] = l;
}

__global__ __launch_bounds__(256) void k_adj_b3(const float2 *__restrict__ coef, const float2 *__restrict__ Y,
                                                const unsigned short *__restrict__ Xth,
                                                const unsigned short *__restrict__ Xtl, int nuv, int nuvp, int W,
                                                int p_per_split, float *__restrict__ part)
{
    __shared__ __attribute__((aligned(16))) unsigned short sAh[AB3_YT * AB3_STRIDE];
    __shared__ __attribute__((aligned(16))) unsigned short sAl[AB3_YT * AB3_STRIDE];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 31, lh = lane >> 5;
    const int tilesx = W / AB3_XT;
    const int y0 = (blockIdx.x / tilesx) * AB3_YT, x0 = (blockIdx.x % tilesx) * AB3_XT;
    const int xw = x0 + 32 * wave + lr;   // this lane's B column
    const int pstart = blockIdx.y * p_per_split;
    const int pend = min(pstart + p_per_split, nuv);
    v16f acc[2];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int r = 0; r < 16; r++)
            acc[m][r] = 0.f;
    // A-tile builder: thread -> point pl = tid / 16 of the round, rows yy = 4 (tid % 16) .. +3
    const int bpl = tid >> 4, byy = (tid & 15) * 4;
    for (int pr = pstart; pr < pend; pr += AB3_PR) {
        {
            const int p = pr + bpl;
            float2 a = make_float2(0.f, 0.f);
            float4 y01 = make_float4(0.f, 0.f, 0.f, 0.f), y23 = y01;
            if (p < pend) {
                a = coef[p];
                const float4 *yr = reinterpret_cast<const float4 *>(Y + (size_t)p * W + y0 + byy);
                y01 = yr[0];
                y23 = yr[1];
            }
            const float2 yv[4] = {make_float2(y01.x, y01.y), make_float2(y01.z, y01.w), make_float2(y23.x, y23.y),
                                  make_float2(y23.z, y23.w)};
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float2 ay = cmul(a, yv[j]);
                unsigned short h0, l0, h1, l1;
                split_bf16(ay.x, h0, l0);
                split_bf16(-ay.y, h1, l1);
                const int o = (byy + j) * AB3_STRIDE + 2 * bpl;
                sAh[o] = h0;
                sAh[o + 1] = h1;
                sAl[o] = l0;
                sAl[o + 1] = l1;
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            // B: points pr + 8 ks + 4 lh .. +3, (Re, Im) interleaved = 8 bf16 at column xw
            const size_t bo = ((size_t)xw * nuvp + pr + 8 * ks + 4 * lh) * 2;
            const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(Xth + bo);
            const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(Xtl + bo);
            bf16x8 ah[2], al[2];
#pragma unroll
            for (int m = 0; m < 2; m++) {
                const int o = (32 * m + lr) * AB3_STRIDE + 16 * ks + 8 * lh;
                ah[m] = *reinterpret_cast<const bf16x8 *>(&sAh[o]);
                al[m] = *reinterpret_cast<const bf16x8 *>(&sAl[o]);
            }
            // product-major over the two accumulators (same per-tile order hh, hl, lh)
#pragma unroll
            for (int m = 0; m < 2; m++)
                acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[m], bh, acc[m], 0, 0, 0);
#pragma unroll
            for (int m = 0; m < 2; m++)
                acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[m], bl, acc[m], 0, 0, 0);
#pragma unroll
            for (int m = 0; m < 2; m++)
                acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[m], bh, acc[m], 0, 0, 0);
        }
        __syncthreads();
    }
    float *out = part + (size_t)blockIdx.y * W * W;
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int y = y0 + 32 * m + 8 * (r >> 2) + 4 * lh + (r & 3);
            out[(size_t)y * W + xw] = acc[m][r];
        }
}

// Wide variant for W a multiple of 256: one workgroup (8 waves) = a 256 x 256
// quadrant of the image for one uv split, so every table entry of the split is read
// once per quadrant (the 64 x 128 tiles above read Y twice and X four times at
// 256^2).  Per round of 16 uv points the A tile (a Y, 256 rows) and the B tile
// (split X, 256 columns) are staged in LDS, double-buffered, from registers loaded
// one round ahead; wave w owns rows 64 (w / 2) .. +63 and columns 128 (w % 2) ..
// +127 of the quadrant (2 x 4 accumulator tiles).
#define AW_T 256
#define AW_STRIDE (2 * AB3_PR + 8)   // bf16 per LDS row: 32 k + 8 padding
#ifndef ADJ_WIDE
#define ADJ_WIDE 1                   // use k_adj_b3w where the width allows
#endif
#ifndef ADJ_WIDE_GROUPS
#define ADJ_WIDE_GROUPS 256          // workgroups of k_adj_b3w (quadrants x uv splits)
#endif
// threads of the factored adjoints: 8 waves, 4 for the 64 x 64 tile
#define ADJ_NTH(TR) ((TR) == 64 ? 256 : 512)
#ifndef ADJ_MIN_PTS
#define ADJ_MIN_PTS 64               // fewest uv points per split of the split-precision adjoints
#endif
#ifndef ADJ_TILE_R
#define ADJ_TILE_R 256               // output tile of k_adj_b3wf / k_adj_h1wf (rows x columns)
#define ADJ_TILE_C 256
#endif

__global__ __launch_bounds__(512) void k_adj_b3w(const float2 *__restrict__ coef, const float2 *__restrict__ Y,
                                                 const unsigned short *__restrict__ Xth,
                                                 const unsigned short *__restrict__ Xtl, int nuv, int nuvp, int W,
                                                 int p_per_split, float *__restrict__ part)
{
    __shared__ __attribute__((aligned(16))) unsigned short sA[2][2][AW_T * AW_STRIDE];   // [buf][hi/lo]
    __shared__ __attribute__((aligned(16))) unsigned short sB[2][2][AW_T * AW_STRIDE];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 31, lh = lane >> 5;
    const int qx = W / AW_T;
    const int y0 = (blockIdx.x / qx) * AW_T, x0 = (blockIdx.x % qx) * AW_T;
    const int pstart = blockIdx.y * p_per_split;
    const int pend = min(pstart + p_per_split, nuv);
    // fill roles: A: point ap = tid / 32 of the round, rows ay .. ay + 7;
    //             B: column bx, points bk .. bk + 7 of the round (k = 2 bk .. 2 bk + 15)
    const int ap = tid >> 5, ay = (tid & 31) * 8;
    const int bx = tid >> 1, bk = (tid & 1) * 8;
    float4 ya0, ya1, ya2, ya3;
    float2 ca;
    uint4 xh0, xh1, xl0, xl1;
    auto load = [&](int pr) {
        // points past the split get a = 0 (their Y row, read from the split's last point,
        // and their X contribute nothing)
        const int p = pr + ap;
        const int pv = p < pend ? p : pend - 1;
        const float2 cv = coef[pv];
        ca = p < pend ? cv : make_float2(0.f, 0.f);
        const float4 *yr = reinterpret_cast<const float4 *>(Y + (size_t)pv * W + y0 + ay);
        ya0 = yr[0];
        ya1 = yr[1];
        ya2 = yr[2];
        ya3 = yr[3];
        // rounds start at multiples of 16 below nuv, so pr + 16 <= nuvp: in the padded table
        const size_t bo = ((size_t)(x0 + bx) * nuvp + pr + bk) * 2;
        xh0 = *reinterpret_cast<const uint4 *>(Xth + bo);
        xh1 = *reinterpret_cast<const uint4 *>(Xth + bo + 8);
        xl0 = *reinterpret_cast<const uint4 *>(Xtl + bo);
        xl1 = *reinterpret_cast<const uint4 *>(Xtl + bo + 8);
    };
    auto store = [&](int buf) {
        auto put = [&](int j, float yr, float yi) {
            const float2 v = cmul(ca, make_float2(yr, yi));
            unsigned short h0, l0, h1, l1;
            split_bf16(v.x, h0, l0);
            split_bf16(-v.y, h1, l1);
            const int o = (ay + j) * AW_STRIDE + 2 * ap;
            *reinterpret_cast<unsigned *>(&sA[buf][0][o]) = (unsigned)h0 | ((unsigned)h1 << 16);
            *reinterpret_cast<unsigned *>(&sA[buf][1][o]) = (unsigned)l0 | ((unsigned)l1 << 16);
        };
        put(0, ya0.x, ya0.y);
        put(1, ya0.z, ya0.w);
        put(2, ya1.x, ya1.y);
        put(3, ya1.z, ya1.w);
        put(4, ya2.x, ya2.y);
        put(5, ya2.z, ya2.w);
        put(6, ya3.x, ya3.y);
        put(7, ya3.z, ya3.w);
        const int o = bx * AW_STRIDE + 2 * bk;
        *reinterpret_cast<uint4 *>(&sB[buf][0][o]) = xh0;
        *reinterpret_cast<uint4 *>(&sB[buf][0][o + 8]) = xh1;
        *reinterpret_cast<uint4 *>(&sB[buf][1][o]) = xl0;
        *reinterpret_cast<uint4 *>(&sB[buf][1][o + 8]) = xl1;
    };
    const int wy = (wave >> 1) * 64, wx = (wave & 1) * 128;
    v16f acc[2][4];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int r = 0; r < 16; r++)
                acc[m][n][r] = 0.f;
    int buf = 0;
    if (pstart < pend) {
        load(pstart);
        store(0);
    }
    __syncthreads();
    for (int pr = pstart; pr < pend; pr += AB3_PR) {
        const bool more = pr + AB3_PR < pend;
        if (more)
            load(pr + AB3_PR);
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            bf16x8 ah[2], al[2], bh[4], bl[4];
#pragma unroll
            for (int m = 0; m < 2; m++) {
                const int o = (wy + 32 * m + lr) * AW_STRIDE + 16 * ks + 8 * lh;
                ah[m] = *reinterpret_cast<const bf16x8 *>(&sA[buf][0][o]);
                al[m] = *reinterpret_cast<const bf16x8 *>(&sA[buf][1][o]);
            }
#pragma unroll
            for (int n = 0; n < 4; n++) {
                const int o = (wx + 32 * n + lr) * AW_STRIDE + 16 * ks + 8 * lh;
                bh[n] = *reinterpret_cast<const bf16x8 *>(&sB[buf][0][o]);
                bl[n] = *reinterpret_cast<const bf16x8 *>(&sB[buf][1][o]);
            }
            // product-major over the eight accumulators (same per-tile order hh, hl, lh)
#pragma unroll
            for (int m = 0; m < 2; m++)
#pragma unroll
                for (int n = 0; n < 4; n++)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[m], bh[n], acc[m][n], 0, 0, 0);
#pragma unroll
            for (int m = 0; m < 2; m++)
#pragma unroll
                for (int n = 0; n < 4; n++)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[m], bl[n], acc[m][n], 0, 0, 0);
#pragma unroll
            for (int m = 0; m < 2; m++)
#pragma unroll
                for (int n = 0; n < 4; n++)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[m], bh[n], acc[m][n], 0, 0, 0);
        }
        // the other buffer was last read in the previous round, before its barrier
        if (more)
            store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    float *out = part + (size_t)blockIdx.y * W * W;
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int y = y0 + wy + 32 * m + 8 * (r >> 2) + 4 * lh + (r & 3);
                out[(size_t)y * W + x0 + wx + 32 * n + lr] = acc[m][n][r];
            }
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
// d = -z + beta d, projected to zero where the pixel sits on the floor and d < 0;
// out[0] = d.g, out[1] = sum d, out[2] = sum |d|.  With zpart set, every block
// first adds up the nzp pairs of z.g and z.g_old partials k_grad_finish_crit left
// (the same fixed order in every block; block 0 keeps them in scal[SC_ZG],
// scal[SC_ZGO]), else it takes those two slots as they are.  conj != 0:
// Polak-Ribiere+ beta = max(0, (z.g - z.g_old) / zg_old) (kept in scal[SC_BETA]);
// conj == 0: beta = 0 (-z).  With dh set, also the BF16 hi / lo split of d for the
// forward DFT that follows.
__device__ __forceinline__ void k_pcg_dir_body(const int BX, const int GX, const float *__restrict__ img, const float *__restrict__ z,
                                                         const float *__restrict__ g, int conj, double zg_old,
                                                         float minval, int n, float *__restrict__ d, double *scal,
                                                         const double *__restrict__ zpart, int nzp,
                                                         double *part, unsigned *ticket, double *out,
                                                         unsigned short *__restrict__ dh, unsigned short *__restrict__ dl){
    __shared__ double shz[2][RED_THREADS / 64];
    __shared__ double zsum[2];
    if (zpart) {
        double zv[2] = {0.0, 0.0};
        for (int b = threadIdx.x; b < nzp; b += RED_THREADS) {
            zv[0] += zpart[2 * b];
            zv[1] += zpart[2 * b + 1];
        }
        block_sum<2>(zv, shz);
        if (threadIdx.x == 0) {
            zsum[0] = zv[0];
            zsum[1] = zv[1];
            if (BX == 0) {
                scal[SC_ZG] = zv[0];
                scal[SC_ZGO] = zv[1];
            }
        }
        __syncthreads();
    } else if (threadIdx.x == 0) {
        zsum[0] = scal[SC_ZG];
        zsum[1] = scal[SC_ZGO];
    }
    __syncthreads();
    double b = 0.0;
    if (conj) {
        b = (zsum[0] - zsum[1]) / zg_old;
        if (!(b > 0.0))
            b = 0.0;
    }
    if (BX == 0 && threadIdx.x == 0)
        scal[SC_BETA] = b;
    const float beta = (float)b;
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = BX * RED_THREADS + threadIdx.x; i < n; i += GX * RED_THREADS) {
        float di = -z[i] + beta * d[i];
        if (img[i] <= minval && di < 0.f)
            di = 0.f;
        d[i] = di;
        if (dh) {
            unsigned short h, l;
            split_bf16(di, h, l);
            dh[i] = h;
            dl[i] = l;
        }
        v[0] += (double)di * (double)g[i];
        v[1] += (double)di;
        v[2] += fabs((double)di);
    }
    reduce_finish_at<3>(v, part, ticket, out, nullptr, BX, GX);
}

__global__ __launch_bounds__(RED_THREADS) void k_pcg_dir(const float *__restrict__ img, const float *__restrict__ z,
                                                         const float *__restrict__ g, int conj, double zg_old,
                                                         float minval, int n, float *__restrict__ d, double *scal,
                                                         const double *__restrict__ zpart, int nzp,
                                                         double *part, unsigned *ticket, double *out,
                                                         unsigned short *__restrict__ dh, unsigned short *__restrict__ dl)
{
    k_pcg_dir_body((int)blockIdx.x, (int)gridDim.x, img, z, g, conj, zg_old, minval, n, d, scal, zpart, nzp, part, ticket, out, dh, dl);
}

// candidate 0 is s = 0 (the linearised current image, the common reference)

// The line search and the commit work on CHANGES of the objective in FP64: the
// current visibilities are kept in FP64 (Vd) and only DFTs of the direction d
// and of the image change are taken in FP32 / BF16x3, so the decision
// variables keep FP64 resolution however small the improvement is.
__device__ inline double2 dcmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ inline double2 dleg(double2 v, int conj_bit) { return conj_bit ? make_double2(v.x, -v.y) : v; }

__device__ inline double pix_entropy(double t, double M)
{
    return (t > 0.0 && M > 0.0) ? (t - M - t * log(t / M)) : -M;
}

// Vd[p] = vis[p] (FP64 copy of freshly evaluated visibilities)
__device__ __forceinline__ void k_vis_to_double_body(const int BX, const int GX, const float2 *__restrict__ vis, int nuv, double2 *__restrict__ Vd){
    const int p = BX * blockDim.x + threadIdx.x;
    if (p < nuv)
        Vd[p] = make_double2(vis[p].x, vis[p].y);
}

__global__ void k_vis_to_double(const float2 *__restrict__ vis, int nuv, double2 *__restrict__ Vd)
{
    k_vis_to_double_body((int)blockIdx.x, (int)gridDim.x, vis, nuv, Vd);
}

// Batched line search, one launch per candidate batch: blocks [0, nb_chi2) take the
// chi2 change of every candidate, one partition at a time; then LS_K / LS_G rows of
// nb_ent blocks the entropy change, row y candidates LS_G y .. LS_G y + LS_G - 1
// (fewer live registers per lane).  Both halves store the sums of fixed partitions
// of LS_SPAN items (part: [2][LS_K / LS_G][RED_PARTS][LS_G], chi2 half first), so
// that neither depends on nb_chi2 / nb_ent; k_ls_choose adds them up in a fixed
// order.
#define LS_SPAN (2 * RED_THREADS)
#define LS_HALF ((LS_K / LS_G) * RED_PARTS * LS_G)   // doubles of partials per half

// chi2(s_k) - chi2(0) from V0 = U / F in FP64 and the direction's DFT U_d.  With
// V(s) = (U + s U_d) / (F + s F_d) = V0 + sigma D, sigma = s / (F + s F_d) and
// D = U_d - F_d V0, every model term moves by a polynomial in sigma without a
// constant term: e (|V(s)|^2 - |V0|^2) = sigma (a1 + sigma a2) for a V2 point, and
// the change of a T3 (the product of three legs V0_j + sigma D_j, conjugated where
// the triangle says, times the phasor) is sigma (c1 + sigma (c2 + sigma c3)).  A
// residual r0 then changes by delta, and r^2 - r0^2 = delta (delta + 2 r0): the
// change itself is evaluated, not a difference of two squares (whose rounding would
// swamp the small changes of the short steps).  One block per partition and all 32
// candidates: the block builds each item's coefficients once (weights folded in)
// into LDS, LS_SPAN items at a time, then wave w evaluates candidates LS_G w ..
// LS_G w + LS_G - 1 over them by Horner's rule, its lanes taking every 64th item
// in order; the candidates' sigma_k are wave-uniform (lane k computes candidate k's,
// the others read it from scalar registers).  Each wave's sums are its group's
// partials: no block reduction.
__device__ __forceinline__ void ls_chi2_part(const double2 *__restrict__ Vd, const float2 *__restrict__ Ud,
                                             const double *__restrict__ scal, double s_ref,
                                             const float *__restrict__ data, const float *__restrict__ err,
                                             const float2 *__restrict__ phasor, const int4 *__restrict__ tri,
                                             int npow, int nbis, int bx, int nbx, double *part)
{
    static_assert(RED_THREADS / 64 * LS_G == LS_K, "one candidate group per wave");
    // per item (c1x, c1y, c2x, c2y) in cf0, (c3x, c3y, 2 r0x, 2 r0y) in cf1; a V2 point
    // (a1, a2, 2 r0, 0) in cf0.  Two arrays, not one of pairs: 32-byte lane strides keep
    // the 16-byte LDS accesses of eight lanes on distinct banks
    __shared__ double4 cf0[LS_SPAN], cf1[LS_SPAN];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const double FI = scal[SC_FLUX], Fd = scal[SC_FD];
    const double s_l = ls_step(s_ref, wave * LS_G + (lane & (LS_G - 1)));
    const double sg_l = s_l / (FI + s_l * Fd);
    double sg[LS_G];
#pragma unroll
    for (int k = 0; k < LS_G; k++)
        sg[k] = readlane_d(sg_l, k);
    const int n = npow + nbis;
    const int nparts = red_nparts(n, LS_SPAN);
    const int span = (n + nparts - 1) / nparts;
    for (int q = bx; q < nparts; q += nbx) {
        double v[LS_G];
#pragma unroll
        for (int k = 0; k < LS_G; k++)
            v[k] = 0.0;
        const int pend = (q + 1) * span < n ? (q + 1) * span : n;
        for (int beg = q * span; beg < pend; beg += LS_SPAN) {
            const int end = beg + LS_SPAN < pend ? beg + LS_SPAN : pend;
            for (int i = beg + tid; i < end; i += RED_THREADS) {
                double4 x0, x1 = make_double4(0.0, 0.0, 0.0, 0.0);
                if (i < npow) {
                    const double2 a = Vd[i];
                    const float2 u = Ud[i];
                    const double d0 = data[i], e = err[i];
                    const double dx = (double)u.x - Fd * a.x, dy = (double)u.y - Fd * a.y;
                    x0 = make_double4(2.0 * e * (a.x * dx + a.y * dy), e * (dx * dx + dy * dy),
                                      2.0 * (a.x * a.x + a.y * a.y - d0) * e, 0.0);
                } else {
                    const int t = i - npow;
                    const int4 tr = tri[t];
                    const int idx[3] = {tr.x, tr.y, tr.z};
                    double2 A[3], B[3];   // legs A_j + sigma B_j
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        const double2 a = Vd[idx[j]];
                        const float2 u = Ud[idx[j]];
                        const double dx = (double)u.x - Fd * a.x, dy = (double)u.y - Fd * a.y;
                        const bool cj = (tr.w >> j) & 1;
                        A[j] = make_double2(a.x, cj ? -a.y : a.y);
                        B[j] = make_double2(dx, cj ? -dy : dy);
                    }
                    const double2 rho = make_double2(phasor[t].x, phasor[t].y);
                    A[2] = dcmul(A[2], rho);
                    B[2] = dcmul(B[2], rho);
                    // (A0 + sigma B0)(A1 + sigma B1) = p0 + sigma p1 + sigma^2 p2, times (A2 + sigma B2)
                    const double2 p0 = dcmul(A[0], A[1]), p2 = dcmul(B[0], B[1]);
                    const double2 p1a = dcmul(A[0], B[1]), p1b = dcmul(B[0], A[1]);
                    const double2 p1 = make_double2(p1a.x + p1b.x, p1a.y + p1b.y);
                    const double2 c0 = dcmul(p0, A[2]), c3 = dcmul(p2, B[2]);
                    const double2 c1a = dcmul(p0, B[2]), c1b = dcmul(p1, A[2]);
                    const double2 c2a = dcmul(p1, B[2]), c2b = dcmul(p2, A[2]);
                    const int kk = npow + 2 * t;
                    const double er = err[kk], ei = err[kk + 1];
                    x0 = make_double4((c1a.x + c1b.x) * er, (c1a.y + c1b.y) * ei, (c2a.x + c2b.x) * er,
                                      (c2a.y + c2b.y) * ei);
                    x1 = make_double4(c3.x * er, c3.y * ei, 2.0 * (c0.x - (double)data[kk]) * er,
                                      2.0 * (c0.y - (double)data[kk + 1]) * ei);
                }
                cf0[i - beg] = x0;
                cf1[i - beg] = x1;
            }
            __syncthreads();
            const int cnt = end - beg, nv = npow - beg;   // items [0, nv) of the chunk are V2 points
            for (int j = lane; j < cnt; j += 64) {
                const double4 x0 = cf0[j];
                if (j < nv) {
#pragma unroll
                    for (int k = 0; k < LS_G; k++) {
                        const double dl = sg[k] * fma(x0.y, sg[k], x0.x);
                        v[k] = fma(dl, dl + x0.z, v[k]);
                    }
                } else {
                    const double4 x1 = cf1[j];
#pragma unroll
                    for (int k = 0; k < LS_G; k++) {
                        const double s = sg[k];
                        const double hx = fma(fma(x1.x, s, x0.z), s, x0.x) * s;
                        const double hy = fma(fma(x1.y, s, x0.w), s, x0.y) * s;
                        v[k] = fma(hx, hx + x1.z, v[k]);
                        v[k] = fma(hy, hy + x1.w, v[k]);
                    }
                }
            }
            __syncthreads();   // the next chunk overwrites cf0, cf1
        }
#pragma unroll
        for (int k = 0; k < LS_G; k++) {
            const double t = wave_sum(v[k]);
            if (lane == 0)
                part[((size_t)wave * RED_PARTS + q) * LS_G + k] = t;
        }
    }
}

// S(max(I + s_k d, minval)) - S(I), summed over the pixels
__device__ __forceinline__ void ls_entropy_part(const float *__restrict__ img, const float *__restrict__ d,
                                                const float *__restrict__ model, double s_ref, double minval, int n,
                                                int grp, int bx, int nbx, double *part)
{
    const int g0 = grp * LS_G;
    float sk[LS_G];
#pragma unroll
    for (int k = 0; k < LS_G; k++)
        sk[k] = (float)ls_step(s_ref, g0 + k);
    auto item = [&](int i, double (&v)[LS_G]) {
        const float I = img[i], dd = d[i], M = model[i];
        if (!(I > 0.f && M > 0.f))
            return;   // the floor keeps I > 0; S = -M there whatever the step
        if (dd == 0.f && I >= (float)minval)
            return;   // t = I for every step: each term below is an exact zero
        // S(t) - S(I) = dl (1 - ln(I/M)) - t log1p(dl / I), dl = t - I: every term is
        // proportional to the change, so FP32 resolves it relative to the change itself
        const float L = logf(I / M), invI = 1.0f / I;
#pragma unroll
        for (int k = 0; k < LS_G; k++) {
            float t = I + sk[k] * dd;
            if (t < (float)minval)
                t = (float)minval;
            const float dl = t - I, q = dl * invI;
            // ln(t/I) = log1p(q).  While the change is small, with u = 1 + q rounded,
            // log(u) q / (u - 1): u - 1 is exact, and the quotient takes the rounding of
            // u out to first order (within 4 ulp of log1p, at a tenth of its
            // instructions).  Once t is far below I (q <= -0.5, where dl rounds to -I)
            // the log of the ratio itself.
            const float u = 1.0f + q;
            const bool near = q > -0.5f;
            const float lx = __logf(near ? u : t * invI);
            const float lr = !near ? lx : (u == 1.0f ? q : lx * (q * __builtin_amdgcn_rcpf(u - 1.0f)));
            v[k] += (double)(dl * (1.0f - L) - t * lr);
        }
    };
    partitioned_partials<LS_G>(n, LS_SPAN, item, part + (size_t)grp * RED_PARTS * LS_G, bx, nbx);
}

__global__ __launch_bounds__(RED_THREADS) void k_ls_batch(const double2 *__restrict__ Vd,
                                                          const float2 *__restrict__ Ud,
                                                          const double *__restrict__ scal, double s_ref,
                                                          const float *__restrict__ data,
                                                          const float *__restrict__ err,
                                                          const float2 *__restrict__ phasor,
                                                          const int4 *__restrict__ tri, int npow, int nbis,
                                                          const float *__restrict__ img, const float *__restrict__ d,
                                                          const float *__restrict__ model, double minval, int n,
                                                          int nb_chi2, int nb_ent, double *part)
{
    const int b = (int)blockIdx.x;
    if (b < nb_chi2) {
        ls_chi2_part(Vd, Ud, scal, s_ref, data, err, phasor, tri, npow, nbis, b, nb_chi2, part);
    } else {
        const int e = b - nb_chi2, grp = e / nb_ent;
        ls_entropy_part(img, d, model, s_ref, minval, n, grp, e - grp * nb_ent, nb_ent, part + LS_HALF);
    }
}

// blocks of a line-search batch launch: nb_chi2 for the chi2 change, then LS_K / LS_G
// rows of nb_ent for the entropy change
__host__ __device__ inline int ls_grid(int nb_chi2, int nb_ent) { return nb_chi2 + (LS_K / LS_G) * nb_ent; }

// block columns of each half of k_ls_batch: one per partition unless a smaller count is given
__host__ inline int ls_nparts_chi2(int npow, int nbis) { return red_nparts(npow + nbis, LS_SPAN); }
__host__ inline int ls_nparts_ent(int n) { return red_nparts(n, LS_SPAN); }

// Sums of a line-search batch and the step they give, on the device so that the
// commit can follow the batch without a read-back.  First ls[m] (m < LS_K: chi2
// change of candidate m, else the entropy change of candidate m - LS_K): LS_FIN_SUB
// lanes per sum take every LS_FIN_SUB-th partition in order, then a fixed butterfly
// adds them.  Then the step: when d is a descent direction (d.g < 0, sum |d| > 0)
// and some candidate k > 0 lowers the objective, the best one (first minimum of
// dJ_k = dchi2_k - alpha dS_k), refined by the parabola through its neighbours;
// otherwise step 0 and index -1.  The host reads both back with the commit's
// results.  One wave: lane k takes candidate k, a shuffle reduction finds the first
// minimum, lane 0 evaluates the parabola.
#define LS_FIN_SUB 16
#define LS_FIN_THREADS (2 * LS_K * LS_FIN_SUB)
__device__ __forceinline__ void k_ls_choose_body(const int BX, const int GX, const double *__restrict__ part, int np_chi2,
                                                              int np_ent, double alpha, double s_ref,
                                                              double *scal, double *__restrict__ ls){
#pragma clang fp contract(off)
    __shared__ double sh[2 * LS_K];
    {
        const int m = threadIdx.x / LS_FIN_SUB, sub = threadIdx.x % LS_FIN_SUB;
        const int half = m / LS_K, c = m % LS_K;
        const double *p = part + (size_t)half * LS_HALF + (size_t)(c / LS_G) * RED_PARTS * LS_G + c % LS_G;
        const int np = half ? np_ent : np_chi2;
        double acc = 0.0;
#pragma unroll 4
        for (int q = sub; q < np; q += LS_FIN_SUB)   // loads in flight four at a time
            acc += p[(size_t)q * LS_G];
#pragma unroll
        for (int o = LS_FIN_SUB / 2; o > 0; o >>= 1)
            acc += __shfl_xor(acc, o, 64);
        if (sub == 0) {
            sh[m] = acc;
            ls[m] = acc;
        }
    }
    __syncthreads();
    if (threadIdx.x >= 64)
        return;
    const int k = threadIdx.x;
    const double dj = k < LS_K ? sh[k] - alpha * sh[LS_K + k] : 0.0;
    double best = k < LS_K && dj == dj ? dj : __builtin_huge_val();   // NaN never the minimum
    int kb = k;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o, 64);
        const int ok = __shfl_xor(kb, o, 64);
        if (ob < best || (ob == best && ok < kb)) {   // first minimum: ties to the smaller index
            best = ob;
            kb = ok;
        }
    }
    const double djm = __shfl(dj, kb > 0 ? kb - 1 : 0, 64), djp = __shfl(dj, kb < LS_K - 1 ? kb + 1 : 0, 64);
    if (k != 0)
        return;
    double s = 0.0, kbo = -1.0;
    if (scal[SC_W1] < 0.0 && scal[SC_DABS] > 0.0 && kb > 0) {
        kbo = kb;
        s = ls_step(s_ref, kb);
        if (kb < LS_K - 1)
            s = parabola_step(ls_step(s_ref, kb - 1), s, ls_step(s_ref, kb + 1), djm, best, djp);
    }
    scal[SC_STEP] = s;
    scal[SC_KBEST] = kbo;
}

__global__ __launch_bounds__(LS_FIN_THREADS) void k_ls_choose(const double *__restrict__ part, int np_chi2,
                                                              int np_ent, double alpha, double s_ref,
                                                              double *scal, double *__restrict__ ls)
{
    k_ls_choose_body((int)blockIdx.x, (int)gridDim.x, part, np_chi2, np_ent, alpha, s_ref, scal, ls);
}

// trial image I_t = max(I + s d, minval), the change delta = I_t - I, and per
// block b: cs_part[3b] = sum delta, cs_part[3b + 1] = S(I_t) - S(I), cs_part[3b + 2]
// = number of pixels where the floor replaced I + s d (plain stores, no ticket:
// k_commit_dchi2 and the host add them up in block order); with th set, also the
// BF16 hi / lo split of delta for the forward DFT that follows
__device__ __forceinline__ void k_commit_step_body(const int BX, const int GX, const float *__restrict__ img,
                                                             const float *__restrict__ d,
                                                             const float *__restrict__ model, float s_arg,
                                                             const double *__restrict__ s_dev, float minval,
                                                             int n, float *__restrict__ trial,
                                                             float *__restrict__ delta, double *__restrict__ cs_part,
                                                             unsigned short *__restrict__ th,
                                                             unsigned short *__restrict__ tl,
                                                             int2 *__restrict__ clist, int *__restrict__ ccount,
                                                             int cper, int W){
    __shared__ double sh[3][RED_THREADS / 64];
    __shared__ int wcnt[RED_THREADS / 64];
    const float s = s_dev ? (float)*s_dev : s_arg;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double v[3] = {0.0, 0.0, 0.0};
    int base = 0;   // clamped pixels listed by this block so far
    for (int i0 = BX * RED_THREADS; i0 < n; i0 += GX * RED_THREADS) {
        const int i = i0 + (int)threadIdx.x;
        bool clamped = false;
        float corr = 0.f;
        if (i < n) {
            const float I = img[i];
            float t = I + s * d[i];
            if (t < minval) {
                t = minval;
                v[2] += 1.0;
                clamped = true;
            }
            trial[i] = t;
            const float dl = t - I;
            delta[i] = dl;
            corr = dl - s * d[i];   // the change beyond s d at a clamped pixel
            if (th) {
                unsigned short h, l;
                split_bf16(dl, h, l);
                th[i] = h;
                tl[i] = l;
            }
            const double M = model[i];
            v[0] += (double)dl;
            v[1] += pix_entropy((double)t, M) - pix_entropy((double)I, M);
        }
        if (clist) {
            // clamped pixels in index order: wave ballots, then the waves in order
            const unsigned long long m = __ballot(clamped);
            if (lane == 0)
                wcnt[wave] = __popcll(m);
            __syncthreads();
            int off = base;
            for (int w = 0; w < wave; w++)
                off += wcnt[w];
            if (clamped) {
                const int pos = off + __popcll(m & ((1ull << lane) - 1ull));
                const int y = i / W, x = i - y * W;   // W <= 512: 16 bits each
                clist[(size_t)BX * cper + pos] = make_int2(x | (y << 16), __float_as_int(corr));
            }
            for (int w = 0; w < RED_THREADS / 64; w++)
                base += wcnt[w];
            __syncthreads();   // wcnt is rewritten by the next pass
        }
    }
    block_sum<3>(v, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < 3; j++)
            cs_part[3 * BX + j] = v[j];
        if (ccount)
            ccount[BX] = base;
    }
}

__global__ __launch_bounds__(RED_THREADS) void k_commit_step(const float *__restrict__ img,
                                                             const float *__restrict__ d,
                                                             const float *__restrict__ model, float s_arg,
                                                             const double *__restrict__ s_dev, float minval,
                                                             int n, float *__restrict__ trial,
                                                             float *__restrict__ delta, double *__restrict__ cs_part,
                                                             unsigned short *__restrict__ th,
                                                             unsigned short *__restrict__ tl,
                                                             int2 *__restrict__ clist, int *__restrict__ ccount,
                                                             int cper, int W)
{
    k_commit_step_body((int)blockIdx.x, (int)gridDim.x, img, d, model, s_arg, s_dev, minval, n, trial, delta, cs_part, th, tl, clist, ccount, cper, W);
}

// visibilities after the committed change: (F V + U_delta) / (F + F_delta)
__device__ __forceinline__ double2 vis_after(const double2 *__restrict__ Vd, const float2 *__restrict__ Udelta,
                                             double FI, double iv, int p)
{
    const double2 a = Vd[p];
    const float2 u = Udelta[p];
    return make_double2((FI * a.x + u.x) * iv, (FI * a.y + u.y) * iv);
}

// exact chi2 change between Vd and the visibilities after the committed change,
// mock data of the latter; dc_part[b] = block b's sum of r(Vn)^2 - r(Vd)^2 (plain
// stores, added up by the host in block order).  The change's flux is the sum of
// k_commit_step's ncs partials, taken in the same order by every block.  The same
// launch writes Vn (FP64) and vis_f (float) for every uv point; the T3 terms
// rebuild their legs with the same expression.
__device__ __forceinline__ void k_commit_dchi2_body(const int BX, const int GX, const double2 *__restrict__ Vd,
                                                              const float2 *__restrict__ Udelta,
                                                              const double *__restrict__ scal,
                                                              const double *__restrict__ cs_part, int ncs, int nuv,
                                                              double2 *__restrict__ Vn, float2 *__restrict__ vis_f,
                                                              const float *__restrict__ data,
                                                              const float *__restrict__ err,
                                                              const float2 *__restrict__ phasor,
                                                              const int4 *__restrict__ tri, int npow, int nbis,
                                                              float *__restrict__ mock_new,
                                                              double *__restrict__ dc_part){
    __shared__ double sh[1][RED_THREADS / 64];
    const double FI = scal[SC_FLUX], iv = 1.0 / (FI + block_fixed_sum(cs_part, 3, ncs));
    for (int p = BX * RED_THREADS + threadIdx.x; p < nuv; p += GX * RED_THREADS) {
        const double2 r = vis_after(Vd, Udelta, FI, iv, p);
        Vn[p] = r;
        vis_f[p] = make_float2((float)r.x, (float)r.y);
    }
    double v[1] = {0.0};
    const int total = npow + nbis;
    for (int i = BX * RED_THREADS + threadIdx.x; i < total; i += GX * RED_THREADS) {
        if (i < npow) {
            const double2 a = Vd[i], b = vis_after(Vd, Udelta, FI, iv, i);
            const double d0 = data[i], e = err[i];
            const double m = b.x * b.x + b.y * b.y;
            const double r0 = (a.x * a.x + a.y * a.y - d0) * e, r1 = (m - d0) * e;
            mock_new[i] = (float)m;
            v[0] += r1 * r1 - r0 * r0;
        } else {
            const int t = i - npow;
            const int4 tr = tri[t];
            const double2 rho = make_double2(phasor[t].x, phasor[t].y);
            const int kk = npow + 2 * t;
            const double dr = data[kk], di = data[kk + 1], er = err[kk], ei = err[kk + 1];
            const double2 T0 =
                dcmul(dcmul(dcmul(dleg(Vd[tr.x], tr.w & 1), dleg(Vd[tr.y], tr.w & 2)), dleg(Vd[tr.z], tr.w & 4)), rho);
            const double2 T1 = dcmul(dcmul(dcmul(dleg(vis_after(Vd, Udelta, FI, iv, tr.x), tr.w & 1),
                                                 dleg(vis_after(Vd, Udelta, FI, iv, tr.y), tr.w & 2)),
                                           dleg(vis_after(Vd, Udelta, FI, iv, tr.z), tr.w & 4)),
                                     rho);
            mock_new[kk] = (float)T1.x;
            mock_new[kk + 1] = (float)T1.y;
            const double a0 = (T0.x - dr) * er, a1 = (T0.y - di) * ei;
            const double b0 = (T1.x - dr) * er, b1 = (T1.y - di) * ei;
            v[0] += (b0 * b0 + b1 * b1) - (a0 * a0 + a1 * a1);
        }
    }
    block_sum<1>(v, sh);
    if (threadIdx.x == 0)
        dc_part[BX] = v[0];
}

__global__ __launch_bounds__(RED_THREADS) void k_commit_dchi2(const double2 *__restrict__ Vd,
                                                              const float2 *__restrict__ Udelta,
                                                              const double *__restrict__ scal,
                                                              const double *__restrict__ cs_part, int ncs, int nuv,
                                                              double2 *__restrict__ Vn, float2 *__restrict__ vis_f,
                                                              const float *__restrict__ data,
                                                              const float *__restrict__ err,
                                                              const float2 *__restrict__ phasor,
                                                              const int4 *__restrict__ tri, int npow, int nbis,
                                                              float *__restrict__ mock_new,
                                                              double *__restrict__ dc_part)
{
    k_commit_dchi2_body((int)blockIdx.x, (int)gridDim.x, Vd, Udelta, scal, cs_part, ncs, nuv, Vn, vis_f, data, err, phasor, tri, npow, nbis, mock_new, dc_part);
}

// ---------------------------------------------------------------------------
// split precision BF16x3 forward DFT (split helpers in gpair_kernels.h)
// ---------------------------------------------------------------------------
// factored tables (below): W/16 coarse then FT_FINE fine factors per uv point
#define FT_FINE 16

__host__ __device__ inline int ft_len(int W) { return W / 16 + FT_FINE; }

// X float2 [nuv][W] -> Xh, Xl bf16 [nuv][2][W] (re plane, im plane per uv point)
__global__ void k_split_tables(const float2 *__restrict__ X, int nuv, int W, unsigned short *__restrict__ Xh,
                               unsigned short *__restrict__ Xl)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nuv * W)
        return;
    const size_t p = i / W, x = i % W;
    const float2 v = X[i];
    unsigned short h, l;
    split_bf16(v.x, h, l);
    Xh[(2 * p) * W + x] = h;
    Xl[(2 * p) * W + x] = l;
    split_bf16(v.y, h, l);
    Xh[(2 * p + 1) * W + x] = h;
    Xl[(2 * p + 1) * W + x] = l;
}

// image float [W*W] -> Ih, Il bf16 [W*W], same [y][x] layout
__device__ __forceinline__ void k_split_image_body(const int BX, const int GX, const float *__restrict__ img, int n, unsigned short *__restrict__ Ih,
                              unsigned short *__restrict__ Il){
    const int i = BX * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    unsigned short h, l;
    split_bf16(img[i], h, l);
    Ih[i] = h;
    Il[i] = l;
}

__global__ void k_split_image(const float *__restrict__ img, int n, unsigned short *__restrict__ Ih,
                              unsigned short *__restrict__ Il)
{
    k_split_image_body((int)blockIdx.x, (int)gridDim.x, img, n, Ih, Il);
}

// Forward DFT on v_mfma_f32_32x32x16_bf16 (gfx950).  Operand layouts: lane l
// holds A[l%32][8(l/32)..+7] and B[8(l/32)..+7][l%32]; accumulator register r
// holds D[8(r/4) + 4(l/32) + r%4][l%32].
// One workgroup = 64 uv points; wave w owns 16 of them as one 32-row block
// (rows 0-15 Re X, rows 16-31 Im X of the same points, so Re and Im of a point
// meet in one lane as accumulator registers r and r+8) and sweeps the image
// rows in groups of up to 256 (NBK blocks of 32).  K = x runs through LDS in
// chunks of 32 columns of the split image [y][x], double-buffered: the global
// loads of chunk c+1 are in flight while chunk c feeds the MFMAs.  The A operand
// (split X rows) comes straight from global memory, one k-step ahead.
#define B3_PW 64
#define B3_KC 32
#define B3_STRIDE 40   // bf16 per LDS row: 32 + 8 padding, conflict-free 16-byte reads
#define B3_ROWS 256
#define B3_WG 256      // stream-K workgroups of the forward (at most)


// cooperative copy of one 32-column chunk of the split image rows into registers,
// then into LDS (row stride B3_STRIDE); up to 4 pieces of 16 bytes per thread, kept
// in named registers (an indexed array would be demoted to scratch memory)
template <int Q, int CH>
__device__ __forceinline__ void b3_fetch1(uint4 &h, uint4 &l, const unsigned short *__restrict__ Ih,
                                          const unsigned short *__restrict__ Il, int tid, int y0, int W, int cc)
{
    const int e = tid + 256 * Q;
    if (CH % 256 == 0 || e < CH) {
        const size_t g = (size_t)(y0 + e / (B3_KC / 8)) * W + cc * B3_KC + 8 * (e % (B3_KC / 8));
        h = *reinterpret_cast<const uint4 *>(Ih + g);
        l = *reinterpret_cast<const uint4 *>(Il + g);
    }
}

template <int Q, int CH>
__device__ __forceinline__ void b3_store1(const uint4 &h, const uint4 &l, unsigned short *sh, unsigned short *sl,
                                          int tid)
{
    const int e = tid + 256 * Q;
    if (CH % 256 == 0 || e < CH) {
        const int o = (e / (B3_KC / 8)) * B3_STRIDE + 8 * (e % (B3_KC / 8));
        *reinterpret_cast<uint4 *>(sh + o) = h;
        *reinterpret_cast<uint4 *>(sl + o) = l;
    }
}

template <int PER, int CH>
__device__ __forceinline__ void b3_fetch(uint4 &h0, uint4 &h1, uint4 &h2, uint4 &h3, uint4 &l0, uint4 &l1,
                                         uint4 &l2, uint4 &l3, const unsigned short *__restrict__ Ih,
                                         const unsigned short *__restrict__ Il, int tid, int y0, int W, int cc)
{
    if constexpr (PER > 0) b3_fetch1<0, CH>(h0, l0, Ih, Il, tid, y0, W, cc);
    if constexpr (PER > 1) b3_fetch1<1, CH>(h1, l1, Ih, Il, tid, y0, W, cc);
    if constexpr (PER > 2) b3_fetch1<2, CH>(h2, l2, Ih, Il, tid, y0, W, cc);
    if constexpr (PER > 3) b3_fetch1<3, CH>(h3, l3, Ih, Il, tid, y0, W, cc);
}

template <int PER, int CH>
__device__ __forceinline__ void b3_store(const uint4 &h0, const uint4 &h1, const uint4 &h2, const uint4 &h3,
                                         const uint4 &l0, const uint4 &l1, const uint4 &l2, const uint4 &l3,
                                         unsigned short *sh, unsigned short *sl, int tid)
{
    if constexpr (PER > 0) b3_store1<0, CH>(h0, l0, sh, sl, tid);
    if constexpr (PER > 1) b3_store1<1, CH>(h1, l1, sh, sl, tid);
    if constexpr (PER > 2) b3_store1<2, CH>(h2, l2, sh, sl, tid);
    if constexpr (PER > 3) b3_store1<3, CH>(h3, l3, sh, sl, tid);
}

// Work is balanced over the workgroups stream-K style: the units are (block of 64
// points, 32-column chunk of x), numbered block-major; workgroup w takes the units
// [w upw, (w+1) upw).  A block whose chunks are split between two workgroups (upw is
// at least the chunks per block, so never more than two) gets two partial sums,
// written to vpart[0] (the part starting at chunk 0) and vpart[1] and added by
// k_vis_fwd_b3_combine; whole blocks are written to vis directly.  The split points
// depend on the sizes only, not on NBK, so the rows-per-group choice changes no bit.
template <int NBK>
__global__ __launch_bounds__(256) void k_vis_fwd_b3(const unsigned short *__restrict__ Ih,
                                                       const unsigned short *__restrict__ Il,
                                                       const unsigned short *__restrict__ Xh,
                                                       const unsigned short *__restrict__ Xl,
                                                       const float2 *__restrict__ Y, int nuv, int W, int upw,
                                                       const float *__restrict__ scale, float2 *__restrict__ vis,
                                                       float2 *__restrict__ vpart)
{
    constexpr int ROWS = NBK * 32;              // image rows per group
    constexpr int CH = ROWS * (B3_KC / 8);      // 16-byte pieces per array per chunk
    constexpr int PER = (CH + 255) / 256;       // pieces per thread
    __shared__ __attribute__((aligned(16))) unsigned short sB[2][2][ROWS * B3_STRIDE];   // [buf][hi/lo]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 31, lh = lane >> 5;
    const int nchunk = W / B3_KC;
    const int nblk = (nuv + B3_PW - 1) / B3_PW;
    const int total = nblk * nchunk;
    const int u_end = min(total, (int)(blockIdx.x + 1) * upw);
    for (int u = (int)blockIdx.x * upw; u < u_end;) {
        const int blk = u / nchunk, c0 = u % nchunk;
        const int c1 = min(nchunk, c0 + (u_end - u));
        u += c1 - c0;
        const int p0 = blk * B3_PW + wave * 16;
        const int pa = p0 + (lr & 15);           // A row of this lane: point pa, component lr >> 4
        const bool aval = pa < nuv;
        const size_t arow = aval ? (size_t)(2 * pa + (lr >> 4)) * W : 0;
        float vr[8], vi[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            vr[r] = 0.f;
            vi[r] = 0.f;
        }
        for (int y0 = 0; y0 < W; y0 += ROWS) {
            v16f acc[NBK];
#pragma unroll
            for (int b = 0; b < NBK; b++)
#pragma unroll
                for (int r = 0; r < 16; r++)
                    acc[b][r] = 0.f;
            static_assert(PER <= 4, "staging holds up to 4 pieces per thread");
            uint4 h0 = make_uint4(0, 0, 0, 0), h1 = h0, h2 = h0, h3 = h0, l0 = h0, l1 = h0, l2 = h0, l3 = h0;
            b3_fetch<PER, CH>(h0, h1, h2, h3, l0, l1, l2, l3, Ih, Il, tid, y0, W, c0);
            __syncthreads();   // the previous readers are done with sB
            b3_store<PER, CH>(h0, h1, h2, h3, l0, l1, l2, l3, sB[c0 & 1][0], sB[c0 & 1][1], tid);
            __syncthreads();
            // A operands (this lane's X row) one whole chunk ahead: the table streams from
            // HBM, and with one wave per SIMD only the wave's own MFMA work hides the latency
            static_assert(B3_KC == 32, "two k-steps per chunk");
            uint4 a0h = make_uint4(0, 0, 0, 0), a0l = a0h, a1h = a0h, a1l = a0h;
            if (aval) {
                const int k0 = c0 * B3_KC;
                a0h = *reinterpret_cast<const uint4 *>(Xh + arow + k0 + 8 * lh);
                a0l = *reinterpret_cast<const uint4 *>(Xl + arow + k0 + 8 * lh);
                a1h = *reinterpret_cast<const uint4 *>(Xh + arow + k0 + 16 + 8 * lh);
                a1l = *reinterpret_cast<const uint4 *>(Xl + arow + k0 + 16 + 8 * lh);
            }
            // One chunk: fetch the next chunk's image rows (L2-resident) and then its X
            // rows (the HBM stream) into the idle register set, run the MFMAs on the
            // current set, store the image rows into the idle LDS buffer.  The last chunk
            // of the range re-fetches itself so that every chunk has the same body; issuing
            // the image loads before the table loads lets the table loads stay in flight
            // across the barrier (in-order load counters).  Chunks go in pairs, so the two
            // register sets swap roles without copies.
            uint4 b0h = a0h, b0l = a0l, b1h = a1h, b1l = a1l;
            auto chunk = [&](int c, const uint4 &c0h, const uint4 &c0l, const uint4 &c1h, const uint4 &c1l,
                             uint4 &d0h, uint4 &d0l, uint4 &d1h, uint4 &d1l) {
                const int buf = c & 1;
                const int cn = c + 1 < c1 ? c + 1 : c;
                b3_fetch<PER, CH>(h0, h1, h2, h3, l0, l1, l2, l3, Ih, Il, tid, y0, W, cn);
                __builtin_amdgcn_sched_barrier(0);   // keep the image loads ahead of the table loads
                {   // unconditional: lanes past nuv read row 0 (arow = 0) and are never stored;
                    // a branch here would make the load counters wait for these loads too
                    const int kn = cn * B3_KC;
                    d0h = *reinterpret_cast<const uint4 *>(Xh + arow + kn + 8 * lh);
                    d0l = *reinterpret_cast<const uint4 *>(Xl + arow + kn + 8 * lh);
                    d1h = *reinterpret_cast<const uint4 *>(Xh + arow + kn + 16 + 8 * lh);
                    d1l = *reinterpret_cast<const uint4 *>(Xl + arow + kn + 16 + 8 * lh);
                }
                __builtin_amdgcn_sched_barrier(0);   // and both ahead of this chunk's MFMAs
#pragma unroll
                for (int ks = 0; ks < B3_KC / 16; ks++) {
                    const bf16x8 ah = *reinterpret_cast<const bf16x8 *>(ks ? &c1h : &c0h);
                    const bf16x8 al = *reinterpret_cast<const bf16x8 *>(ks ? &c1l : &c0l);
                    // every B operand of the k-step read from LDS before the first MFMA: the
                    // reads overlap each other instead of one LDS latency per tile
                    bf16x8 bh[NBK], bl[NBK];
#pragma unroll
                    for (int b = 0; b < NBK; b++) {
                        const int o = (b * 32 + lr) * B3_STRIDE + ks * 16 + 8 * lh;
                        bh[b] = *reinterpret_cast<const bf16x8 *>(&sB[buf][0][o]);
                        bl[b] = *reinterpret_cast<const bf16x8 *>(&sB[buf][1][o]);
                    }
                    // product-major: consecutive MFMAs update different accumulators, so
                    // none waits for the previous one's result (same per-tile order hh, hl, lh)
#pragma unroll
                    for (int b = 0; b < NBK; b++)
                        acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh[b], acc[b], 0, 0, 0);
#pragma unroll
                    for (int b = 0; b < NBK; b++)
                        acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl[b], acc[b], 0, 0, 0);
#pragma unroll
                    for (int b = 0; b < NBK; b++)
                        acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh[b], acc[b], 0, 0, 0);
                }
                // the other buffer was last read in chunk c-1, before the previous barrier
                b3_store<PER, CH>(h0, h1, h2, h3, l0, l1, l2, l3, sB[buf ^ 1][0], sB[buf ^ 1][1], tid);
                __syncthreads();
            };
            int c = c0;
            for (; c + 1 < c1; c += 2) {
                chunk(c, a0h, a0l, a1h, a1l, b0h, b0l, b1h, b1l);
                chunk(c + 1, b0h, b0l, b1h, b1l, a0h, a0l, a1h, a1l);
            }
            if (c < c1)
                chunk(c, a0h, a0l, a1h, a1l, b0h, b0l, b1h, b1l);
            // epilogue of this row group: point 8(r/4) + 4 lh + r%4 (r < 8) has Re T in
            // register r and Im T in register r+8, column y = y0 + 32 b + lr
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const int p = p0 + 8 * (r >> 2) + 4 * lh + (r & 3);
                if (p < nuv) {
                    const float2 *yrow = Y + (size_t)p * W + y0 + lr;
#pragma unroll
                    for (int b = 0; b < NBK; b++) {
                        const float2 yv = yrow[b * 32];
                        const float tr = acc[b][r], ti = acc[b][r + 8];
                        vr[r] += tr * yv.x - ti * yv.y;
                        vi[r] += tr * yv.y + ti * yv.x;
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < 8; r++) {
#pragma unroll
            for (int o = 1; o < 32; o <<= 1) {
                vr[r] += __shfl_xor(vr[r], o, 64);
                vi[r] += __shfl_xor(vi[r], o, 64);
            }
        }
        if (lr == 0) {
            const bool whole = c0 == 0 && c1 == nchunk;
            const float f = whole && scale ? *scale : 1.0f;
            float2 *dst = whole ? vis : vpart + (c0 == 0 ? 0 : nuv);
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const int p = p0 + 8 * (r >> 2) + 4 * lh + (r & 3);
                if (p < nuv)
                    dst[p] = make_float2(vr[r] * f, vi[r] * f);
            }
        }
    }
}

// the blocks split between two workgroups of k_vis_fwd_b3: one 64-thread block per
// workgroup boundary, vis = (part 0 + part 1) * scale
__device__ __forceinline__ void k_vis_fwd_b3_combine_body(const int BX, const int GX, const float2 *__restrict__ vpart, int nuv, int W, int upw,
                                     const float *__restrict__ scale, float2 *__restrict__ vis,
                                     const int *__restrict__ run_if){
    if (run_if && *run_if == 0)
        return;
    const int nchunk = W / B3_KC;
    const int total = (nuv + B3_PW - 1) / B3_PW * nchunk;
    const int ub = (int)(BX + 1) * upw;
    if (ub >= total || ub % nchunk == 0)
        return;
    const int p = ub / nchunk * B3_PW + (int)threadIdx.x;
    if (p >= nuv)
        return;
    const float f = scale ? *scale : 1.0f;
    const float2 a = vpart[p], b = vpart[nuv + p];
    vis[p] = make_float2((a.x + b.x) * f, (a.y + b.y) * f);
}

__global__ void k_vis_fwd_b3_combine(const float2 *__restrict__ vpart, int nuv, int W, int upw,
                                     const float *__restrict__ scale, float2 *__restrict__ vis,
                                     const int *__restrict__ run_if)
{
    k_vis_fwd_b3_combine_body((int)blockIdx.x, (int)gridDim.x, vpart, nuv, W, upw, scale, vis, run_if);
}

template __global__ void k_vis_fwd_b3<2>(const unsigned short *, const unsigned short *, const unsigned short *,
                                         const unsigned short *, const float2 *, int, int, int, const float *, float2 *,
                                         float2 *);
template __global__ void k_vis_fwd_b3<4>(const unsigned short *, const unsigned short *, const unsigned short *,
                                         const unsigned short *, const float2 *, int, int, int, const float *, float2 *,
                                         float2 *);
template __global__ void k_vis_fwd_b3<6>(const unsigned short *, const unsigned short *, const unsigned short *,
                                         const unsigned short *, const float2 *, int, int, int, const float *, float2 *,
                                         float2 *);
template __global__ void k_vis_fwd_b3<8>(const unsigned short *, const unsigned short *, const unsigned short *,
                                         const unsigned short *, const float2 *, int, int, int, const float *, float2 *,
                                         float2 *);

// explicit instantiations for the supported widths (W = 64 * NB)
template __global__ void k_vis_fwd_f32<1>(const float *, const float2 *, const float2 *, int, const float *, float2 *,
                                          const int *);
template __global__ void k_vis_fwd_f32<2>(const float *, const float2 *, const float2 *, int, const float *, float2 *,
                                          const int *);
template __global__ void k_vis_fwd_f32<3>(const float *, const float2 *, const float2 *, int, const float *, float2 *,
                                          const int *);
template __global__ void k_vis_fwd_f32<4>(const float *, const float2 *, const float2 *, int, const float *, float2 *,
                                          const int *);
template __global__ void k_vis_fwd_f32<8>(const float *, const float2 *, const float2 *, int, const float *, float2 *,
                                          const int *);

// ---------------------------------------------------------------------------
// BF16x3 DFTs on factored tables
// ---------------------------------------------------------------------------
// The table entry at column x = 16 i + j factors as exp(2 pi i s a w x) =
// exp(2 pi i s a w 16 i) exp(2 pi i s a w j).  Per uv point the factored table
// holds the W/16 coarse factors followed by the 16 fine ones (both from FP64
// range-reduced phases, like k_tables): 8 (W/16 + 16) bytes per point instead of
// 8 W.  The kernels below rebuild the entries they need as products of two
// factors (about 2 ulp against the directly rounded entry, far below the 2^-17
// of the BF16 split) and split them in registers, so the DFTs read a few MB of
// tables instead of streaming the full ones.

// FX[p][k] for X (w = u, s = +1), FY[p][k] for Y (w = v, s = -1)
__global__ void k_ftables(const float *__restrict__ u, const float *__restrict__ v, double a, int nuv, int W,
                          float2 *__restrict__ FX, float2 *__restrict__ FY)
{
    const int ft = ft_len(W), nco = W / 16;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nuv * ft)
        return;
    const int p = (int)(i / ft), k = (int)(i % ft);
    const double m = k < nco ? 16.0 * k : (double)(k - nco);
    const double au = a * (double)u[p], av = a * (double)v[p];
    FX[i] = turn_exp(au * m);
    FY[i] = turn_exp(-av * m);
}

// 8 consecutive entries of one component of a table row, split: entry e is
// Re(cc g[e]) where g is the fine factor, pre-rotated for the imaginary component
__device__ __forceinline__ void ft_gen8(float cr, float ci, const float2 (&g)[8], uint4 &h, uint4 &l)
{
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; e++)
        v[e] = cr * g[e].x - ci * g[e].y;
    split2_bf16(v[0], v[1], h.x, l.x);
    split2_bf16(v[2], v[3], h.y, l.y);
    split2_bf16(v[4], v[5], h.z, l.z);
    split2_bf16(v[6], v[7], h.w, l.w);
}

// Forward DFT as k_vis_fwd_b3 (same MFMA tiling, stream-K units, LDS staging of
// the split image and summation order), with the A operand generated: lane
// (lr, lh) owns row lr of its wave's block, i.e. point p0 + lr % 16 and component
// lr / 16, and columns 8 lh .. 8 lh + 7 of each 16-column half of a chunk, so its
// 8 fine factors are fixed for the whole block and each chunk needs the two
// coarse factors of its halves (one 16-byte load).  The image rows and coarse
// factors of a chunk are loaded two chunks ahead into two alternating register
// sets, so each load has two chunks of MFMA work to land before its LDS store.
// The y-reduction rebuilds Y[p][y] from FY the same way.
struct B3Stage {
    uint4 h0, h1, h2, h3, l0, l1, l2, l3;
};

// The y-reduction of a 64-point block reads, per lane and point, the fine factor of
// its column and the coarse factors of its 8 column blocks: staged once per block in
// LDS (the block's rows of FY are contiguous), so the reduction reads LDS instead of
// waiting on 72 dependent L2 loads per lane.  FYS_MAX: rows of the widest image (512).
#define FYS_MAX (512 / 16 + FT_FINE)
#define FYS_PER ((B3_PW * FYS_MAX + 255) / 256)

// The committed change's unnormalised DFT without a full DFT when few pixels clamp:
// the change is s d plus, at each pixel the floor clamped, a correction c_j, so
//   U_delta[p] = s U_d[p] + sum_j c_j X[p][x_j] Y[p][y_j]
// with U_d the direction's DFT (already on the device).  k_commit_step lists the
// clamped pixels per block in index order (ncb blocks of cper slots, ccount[b] in
// use); a block of this kernel owns 64 uv points (lane = point) with their factor
// rows staged in LDS, and each of its CS_WAVES waves walks a contiguous share of the list
// (one pixel at a time, the same pixel across the wave; entries hold x | y << 16 and
// the correction; the list is read 64 entries per load, the next 64 in flight while
// these are summed, and broadcast lane by lane).  X and Y are coarse x fine factors:
// a run of pixels of one row under one coarse X factor sums its corrections times
// the fine factors and multiplies by the coarse factor once, a row by Y once.  The
// waves' sums are added in wave order: the result depends on the list only.  With
// more than cap pixels the kernel writes nothing but *dense = 1, and the full DFT of
// the change runs instead: in a batch of problems the one launched behind it (it
// exits at once when *dense = 0), for a single problem the one the host launches when
// the read-back shows the count.
#define CS_PTS 64
// soff[b] = sum of ccount[0 .. b-1] for b <= ncb (<= 128), by one wave: an inclusive
// scan of two counts per lane (a serial loop would wait on 128 dependent loads)
__device__ __forceinline__ void cs_offsets(const int *__restrict__ ccount, int ncb, int *soff, int lane)
{
    int a = lane < ncb ? ccount[lane] : 0;
    int b = lane + 64 < ncb ? ccount[lane + 64] : 0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int ta = __shfl_up(a, o, 64), tb = __shfl_up(b, o, 64);
        if (lane >= o) {
            a += ta;
            b += tb;
        }
    }
    const int ha = __shfl(a, 63, 64);
    if (lane == 0)
        soff[0] = 0;
    if (lane < ncb)
        soff[lane + 1] = a;
    if (lane + 64 < ncb)
        soff[lane + 65] = ha + b;
}
#define CS_WAVES 8
// dynamic LDS: the 64 points' X and Y factor rows, padded (ft + 1 entries each)
__host__ __device__ inline size_t cs_lds_bytes(int W) { return 2 * sizeof(float2) * CS_PTS * (size_t)(ft_len(W) + 1); }

__device__ __forceinline__ void k_commit_sparse_body(const int BX, const int GX, const float2 *__restrict__ Ud,
                                                                const double *__restrict__ s_dev, float s_arg,
                                                                const int2 *__restrict__ clist,
                                                                const int *__restrict__ ccount, int ncb, int cper,
                                                                const float2 *__restrict__ FX,
                                                                const float2 *__restrict__ FY, int nuv, int W,
                                                                int cap, float2 *__restrict__ Udel,
                                                                int *__restrict__ dense){
    extern __shared__ float2 cs_rows[];
    __shared__ int soff[RED_GRID_MAX + 1];
    __shared__ float2 sacc[CS_WAVES - 1][CS_PTS];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // the same in every lane of a wave
    if (wave == 0)
        cs_offsets(ccount, ncb, soff, lane);
    __syncthreads();
    const int total = soff[ncb];
    if (BX == 0 && tid == 0)
        *dense = total > cap;   // read by the full DFT launched behind this kernel
    if (total > cap)
        return;
    const int ft = ft_len(W), nco = W / 16, st = ft + 1;   // padded rows: lanes hit different banks
    float2 *sX = cs_rows, *sY = cs_rows + CS_PTS * st;
    const int p0 = BX * CS_PTS;
    {
        // CS_WAVES threads per row of 64, consecutive factors per thread group
        const int q = tid / CS_WAVES, e0 = tid % CS_WAVES, p = p0 + q;
        const size_t g = (size_t)(p < nuv ? p : 0) * ft;   // points past nuv: any valid row
        for (int e = e0; e < ft; e += CS_WAVES) {
            sX[q * st + e] = FX[g + e];
            sY[q * st + e] = FY[g + e];
        }
    }
    __syncthreads();
    const float2 *rx = sX + lane * st, *ry = sY + lane * st;
    const int j0 = (int)((long long)total * wave / CS_WAVES), j1 = (int)((long long)total * (wave + 1) / CS_WAVES);
    float2 acc = make_float2(0.f, 0.f), row = make_float2(0.f, 0.f), yv = make_float2(0.f, 0.f);
    int prev_y = -1;
    // the list entry at position j (this lane's): its block by a search over the offsets
    auto fetch = [&](int j) -> int2 {
        int2 en = make_int2(0, 0);
        if (j < j1) {
            int lo = 0, hi = ncb;   // soff[lo] <= j < soff[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (soff[mid] <= j)
                    lo = mid;
                else
                    hi = mid;
            }
            en = clist[(size_t)lo * cper + (j - soff[lo])];
        }
        return en;
    };
    // consecutive entries of one row that share a coarse X factor form a run: the run
    // sums c_j times the fine factor and takes the coarse factor once at its end;
    // rows take their Y factor once
    float2 bacc = make_float2(0.f, 0.f), xc = make_float2(0.f, 0.f);
    int prev_key = -1;   // y << 8 | coarse block of the current run
    int2 en = fetch(j0 + lane);
    for (int jb = j0; jb < j1; jb += 64) {
        const int2 en_next = fetch(jb + 64 + lane);   // in flight while this chunk is summed
        const int cnt = min(64, j1 - jb);
        for (int k = 0; k < cnt; k++) {
            const int xy = __builtin_amdgcn_readlane(en.x, k);
            const float cj = __int_as_float(__builtin_amdgcn_readlane(en.y, k));
            const int y = xy >> 16, x = xy & 0xffff, key = (y << 8) | (x >> 4);
            if (key != prev_key) {
                row.x += bacc.x * xc.x - bacc.y * xc.y;
                row.y += bacc.x * xc.y + bacc.y * xc.x;
                bacc = make_float2(0.f, 0.f);
                if (y != prev_y) {
                    if (prev_y >= 0) {
                        acc.x += row.x * yv.x - row.y * yv.y;
                        acc.y += row.x * yv.y + row.y * yv.x;
                    }
                    yv = cmul(ry[y >> 4], ry[nco + (y & 15)]);
                    row = make_float2(0.f, 0.f);
                    prev_y = y;
                }
                xc = rx[x >> 4];
                prev_key = key;
            }
            const float2 xf = rx[nco + (x & 15)];
            bacc.x += cj * xf.x;
            bacc.y += cj * xf.y;
        }
        en = en_next;
    }
    row.x += bacc.x * xc.x - bacc.y * xc.y;
    row.y += bacc.x * xc.y + bacc.y * xc.x;
    if (prev_y >= 0) {
        acc.x += row.x * yv.x - row.y * yv.y;
        acc.y += row.x * yv.y + row.y * yv.x;
    }
    if (wave > 0)
        sacc[wave - 1][lane] = acc;
    __syncthreads();
    const int p = p0 + lane;
    if (wave == 0 && p < nuv) {
        const float s = s_dev ? (float)*s_dev : s_arg;
        const float2 u = Ud[p];
        float sx = acc.x, sy = acc.y;
#pragma unroll
        for (int w = 0; w < CS_WAVES - 1; w++) {
            sx += sacc[w][lane].x;
            sy += sacc[w][lane].y;
        }
        Udel[p] = make_float2(s * u.x + sx, s * u.y + sy);
    }
}

__global__ __launch_bounds__(64 * CS_WAVES) void k_commit_sparse(const float2 *__restrict__ Ud,
                                                                const double *__restrict__ s_dev, float s_arg,
                                                                const int2 *__restrict__ clist,
                                                                const int *__restrict__ ccount, int ncb, int cper,
                                                                const float2 *__restrict__ FX,
                                                                const float2 *__restrict__ FY, int nuv, int W,
                                                                int cap, float2 *__restrict__ Udel,
                                                                int *__restrict__ dense)
{
    k_commit_sparse_body((int)blockIdx.x, (int)gridDim.x, Ud, s_dev, s_arg, clist, ccount, ncb, cper, FX, FY, nuv, W, cap, Udel, dense);
}



struct FyStage {
    float2 v[FYS_PER];
};

__device__ __forceinline__ void fy_fetch(FyStage &S, const float2 *__restrict__ FY, int blk, int nuv, int ft, int tid)
{
    const int nval = B3_PW * ft;
    const size_t base = (size_t)blk * B3_PW * ft, lim = (size_t)nuv * ft;
#pragma unroll
    for (int q = 0; q < FYS_PER; q++) {
        const int i = tid + 256 * q;
        const size_t gi = base + i;
        if (i < nval)
            S.v[q] = FY[gi < lim ? gi : 0];   // points past nuv: any valid row (never stored)
    }
}

__device__ __forceinline__ void fy_store(const FyStage &S, float2 *sY, int ft, int tid)
{
    const int nval = B3_PW * ft;
#pragma unroll
    for (int q = 0; q < FYS_PER; q++) {
        const int i = tid + 256 * q;
        if (i < nval)
            sY[i] = S.v[q];
    }
}

template <int NBK>
__device__ __forceinline__ void k_vis_fwd_b3f_body(const int BX, const int GX, const unsigned short *__restrict__ Ih,
                                                        const unsigned short *__restrict__ Il,
                                                        const float2 *__restrict__ FX,
                                                        const float2 *__restrict__ FY, int nuv, int W, int upw,
                                                        const float *__restrict__ scale, float2 *__restrict__ vis,
                                                        float2 *__restrict__ vpart, const int *__restrict__ run_if){
    if (run_if && *run_if == 0)   // a commit whose change DFT the sparse correction already gave
        return;
    constexpr int ROWS = NBK * 32;
    constexpr int CH = ROWS * (B3_KC / 8);
    constexpr int PER = (CH + 255) / 256;
    static_assert(PER <= 4, "staging holds up to 4 pieces per thread");
    static_assert(B3_KC == 32, "two k-steps per chunk");
    __shared__ __attribute__((aligned(16))) unsigned short sB[2][2][ROWS * B3_STRIDE];
    __shared__ float2 sY[B3_PW * FYS_MAX];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 31, lh = lane >> 5;
    const int ft = ft_len(W), nco = W / 16;
    const int nchunk = W / B3_KC;
    const int nblk = (nuv + B3_PW - 1) / B3_PW;
    const int total = nblk * nchunk;
    const int u_end = min(total, (int)(BX + 1) * upw);
    for (int u = (int)BX * upw; u < u_end;) {
        const int blk = u / nchunk, c0 = u % nchunk;
        const int c1 = min(nchunk, c0 + (u_end - u));
        u += c1 - c0;
        const int p0 = blk * B3_PW + wave * 16;
        const int pa = p0 + (lr & 15);
        // lanes past nuv build row 0's entries: their output rows are never stored
        const float2 *fx = FX + (size_t)(pa < nuv ? pa : 0) * ft;
        float2 g[8];
        {
            const float4 *fp = reinterpret_cast<const float4 *>(fx + nco + 8 * lh);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float4 f = fp[q];
                g[2 * q] = make_float2(f.x, f.y);
                g[2 * q + 1] = make_float2(f.z, f.w);
            }
            if (lr >= 16) {   // Im(c f) = Re(c (Im f, -Re f))
#pragma unroll
                for (int e = 0; e < 8; e++)
                    g[e] = make_float2(g[e].y, -g[e].x);
            }
        }
        float vr[8], vi[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            vr[r] = 0.f;
            vi[r] = 0.f;
        }
        const int n = c1 - c0;
        // chunk of position i in the range; positions past the end reload the last chunk
        auto cidx = [&](int i) { return c0 + (i < n ? i : n - 1); };
        for (int y0 = 0; y0 < W; y0 += ROWS) {
            v16f acc[NBK];
#pragma unroll
            for (int b = 0; b < NBK; b++)
#pragma unroll
                for (int r = 0; r < 16; r++)
                    acc[b][r] = 0.f;
            auto fetch = [&](B3Stage &S, int c) {
                b3_fetch<PER, CH>(S.h0, S.h1, S.h2, S.h3, S.l0, S.l1, S.l2, S.l3, Ih, Il, tid, y0, W, c);
            };
            auto stage = [&](const B3Stage &S, int buf) {
                b3_store<PER, CH>(S.h0, S.h1, S.h2, S.h3, S.l0, S.l1, S.l2, S.l3, sB[buf][0], sB[buf][1], tid);
            };
            B3Stage P, Q;
            P.h0 = P.h1 = P.h2 = P.h3 = P.l0 = P.l1 = P.l2 = P.l3 = make_uint4(0, 0, 0, 0);
            Q = P;
            FyStage FS;
            fy_fetch(FS, FY, blk, nuv, ft, tid);
            fetch(P, cidx(0));
            fetch(Q, cidx(1));
            float4 cxa = *reinterpret_cast<const float4 *>(fx + 2 * cidx(0));
            float4 cxb = *reinterpret_cast<const float4 *>(fx + 2 * cidx(1));
            uint4 a0h, a0l, a1h, a1l, b0h, b0l, b1h, b1l;
            ft_gen8(cxa.x, cxa.y, g, a0h, a0l);
            ft_gen8(cxa.z, cxa.w, g, a1h, a1l);
            __syncthreads();   // the previous readers are done with sB and sY
            stage(P, 0);
            fy_store(FS, sY, ft, tid);
            __syncthreads();
            // position i: MFMAs on LDS buffer i % 2 with A set `cur`; the A set of i + 1
            // from its coarse factors `cxn`; chunk i + 1's rows (`Sn`) into the other
            // buffer; chunk i + 2's rows and factors loaded into `Sf` and `cxf`
            auto step = [&](int i, int buf, const uint4 &c0h, const uint4 &c0l, const uint4 &c1h, const uint4 &c1l,
                            uint4 &d0h, uint4 &d0l, uint4 &d1h, uint4 &d1l, const B3Stage &Sn, B3Stage &Sf,
                            const float4 &cxn, float4 &cxf) {
                cxf = *reinterpret_cast<const float4 *>(fx + 2 * cidx(i + 2));
                fetch(Sf, cidx(i + 2));
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int ks = 0; ks < B3_KC / 16; ks++) {
                    const bf16x8 ah = *reinterpret_cast<const bf16x8 *>(ks ? &c1h : &c0h);
                    const bf16x8 al = *reinterpret_cast<const bf16x8 *>(ks ? &c1l : &c0l);
                    bf16x8 bh[NBK], bl[NBK];
#pragma unroll
                    for (int b = 0; b < NBK; b++) {
                        const int o = (b * 32 + lr) * B3_STRIDE + ks * 16 + 8 * lh;
                        bh[b] = *reinterpret_cast<const bf16x8 *>(&sB[buf][0][o]);
                        bl[b] = *reinterpret_cast<const bf16x8 *>(&sB[buf][1][o]);
                    }
#pragma unroll
                    for (int b = 0; b < NBK; b++)
                        acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh[b], acc[b], 0, 0, 0);
#pragma unroll
                    for (int b = 0; b < NBK; b++)
                        acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl[b], acc[b], 0, 0, 0);
#pragma unroll
                    for (int b = 0; b < NBK; b++)
                        acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh[b], acc[b], 0, 0, 0);
                }
                ft_gen8(cxn.x, cxn.y, g, d0h, d0l);
                ft_gen8(cxn.z, cxn.w, g, d1h, d1l);
                // the other buffer was last read at position i - 1, before the previous barrier
                stage(Sn, buf ^ 1);
                __syncthreads();
            };
            int i = 0;
            for (; i + 1 < n; i += 2) {
                step(i, 0, a0h, a0l, a1h, a1l, b0h, b0l, b1h, b1l, Q, P, cxb, cxa);
                step(i + 1, 1, b0h, b0l, b1h, b1l, a0h, a0l, a1h, a1l, P, Q, cxa, cxb);
            }
            if (i < n)
                step(i, 0, a0h, a0l, a1h, a1l, b0h, b0l, b1h, b1l, Q, P, cxb, cxa);
            // y-reduction: point 8(r/4) + 4 lh + r%4 has Re T in register r and Im T in
            // r+8 at column y = y0 + 32 b + lr = 16 (y0/16 + 2 b + lr/16) + lr % 16;
            // its factors from the block's rows staged in sY
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const float2 *fy = sY + (wave * 16 + 8 * (r >> 2) + 4 * lh + (r & 3)) * ft;
                const float2 fj = fy[nco + (lr & 15)];
                const float2 *cy = fy + (y0 >> 4) + (lr >> 4);
#pragma unroll
                for (int b = 0; b < NBK; b++) {
                    const float2 yv = cmul(cy[2 * b], fj);
                    const float tr = acc[b][r], ti = acc[b][r + 8];
                    vr[r] += tr * yv.x - ti * yv.y;
                    vi[r] += tr * yv.y + ti * yv.x;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < 8; r++) {
#pragma unroll
            for (int o = 1; o < 32; o <<= 1) {
                vr[r] += __shfl_xor(vr[r], o, 64);
                vi[r] += __shfl_xor(vi[r], o, 64);
            }
        }
        if (lr == 0) {
            const bool whole = c0 == 0 && c1 == nchunk;
            const float f = whole && scale ? *scale : 1.0f;
            float2 *dst = whole ? vis : vpart + (c0 == 0 ? 0 : nuv);
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const int p = p0 + 8 * (r >> 2) + 4 * lh + (r & 3);
                if (p < nuv)
                    dst[p] = make_float2(vr[r] * f, vi[r] * f);
            }
        }
    }
}

template <int NBK>
__global__ __launch_bounds__(256) void k_vis_fwd_b3f(const unsigned short *__restrict__ Ih,
                                                        const unsigned short *__restrict__ Il,
                                                        const float2 *__restrict__ FX,
                                                        const float2 *__restrict__ FY, int nuv, int W, int upw,
                                                        const float *__restrict__ scale, float2 *__restrict__ vis,
                                                        float2 *__restrict__ vpart, const int *__restrict__ run_if)
{
    k_vis_fwd_b3f_body<NBK>((int)blockIdx.x, (int)gridDim.x, Ih, Il, FX, FY, nuv, W, upw, scale, vis, vpart, run_if);
}

template __global__ void k_vis_fwd_b3f<2>(const unsigned short *, const unsigned short *, const float2 *,
                                          const float2 *, int, int, int, const float *, float2 *, float2 *,
                                          const int *);
template __global__ void k_vis_fwd_b3f<4>(const unsigned short *, const unsigned short *, const float2 *,
                                          const float2 *, int, int, int, const float *, float2 *, float2 *,
                                          const int *);
template __global__ void k_vis_fwd_b3f<6>(const unsigned short *, const unsigned short *, const float2 *,
                                          const float2 *, int, int, int, const float *, float2 *, float2 *,
                                          const int *);
template __global__ void k_vis_fwd_b3f<8>(const unsigned short *, const unsigned short *, const float2 *,
                                          const float2 *, int, int, int, const float *, float2 *, float2 *,
                                          const int *);

// Adjoint as k_adj_b3w with both tiles built from the factored tables, for output
// tiles of TR x TC pixels (TR, TC in {128, 256}) per workgroup and uv split: 8 waves
// of TR / 128 x TC / 64 accumulator tiles (wave w owns rows TR / 4 (w / 2) .. and
// columns TC / 2 (w % 2) ..), LDS double-buffered per round of 16 points.  Smaller
// tiles need fewer uv splits to fill the chip (the partials the finish kernel adds
// up shrink with the splits) at more builder work per MFMA.  The grid is 1-D,
// block b = split b % nsplit of tile b / nsplit, so that the tiles of one split
// run on one XCD (workgroups go to the XCDs in turn) and share its L2 for the
// split's factors.  Builder role of thread t: point q = t % 16 of the round and
// the row (A) and column (B) runs 4 (t / 16) + 128 h .. + 3, h < TR / 128 (A),
// h < TC / 128 (B): the four entries of a run share one coarse factor and the runs
// the same four fine factors.  In a wave the 16 points are consecutive 32-bit
// words and the four runs start 80 words apart (bank offsets 0, 16, 32, 48), so
// the builders' LDS writes are free of bank conflicts.  W is any multiple of 64:
// rows and columns of an edge tile past W are built from the entries that follow
// in the point's factor row (coarse index < (W + TR) / 16 <= the row's W / 16 +
// FT_FINE entries, as TR, TC <= 256) and are not stored.
template <int TR, int TC>
__device__ __forceinline__ void k_adj_b3wf_body(const int BX, const int GX, const float2 *__restrict__ coef, const float2 *__restrict__ FX,
                                                  const float2 *__restrict__ FY, int nuv, int W, int p_per_split,
                                                  int nsplit, float *__restrict__ part){
    static_assert(((TR == 128 || TR == 256) && (TC == 128 || TC == 256)) || (TR == 64 && TC == 64), "tile");
    // ADJ_NTH(TR) threads; builder runs cover RUN rows (columns) per pass; waves in two
    // columns, each wave TR * 4 / NTH row and TC / 64 column tiles of 32
    constexpr int NTH = ADJ_NTH(TR), RUN = NTH / 4;
    constexpr int HA = TR / RUN, HB = TC / RUN, MT = TR * 4 / NTH, NT = TC / 64;
    __shared__ __attribute__((aligned(16))) unsigned short sA[2][2][TR * AW_STRIDE];   // [buf][hi/lo]
    __shared__ __attribute__((aligned(16))) unsigned short sB[2][2][TC * AW_STRIDE];
    constexpr int WROW = AW_STRIDE / 2;   // 32-bit words per LDS row
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 31, lh = lane >> 5;
    const int ft = ft_len(W), nco = W / 16;
    const int split = BX % nsplit, tile = BX / nsplit;
    const int qx = (W + TC - 1) / TC;   // edge tiles of a width that is not a multiple of TC are partial
    const int y0 = (tile / qx) * TR, x0 = (tile % qx) * TC;
    const int pstart = split * p_per_split;
    const int pend = min(pstart + p_per_split, nuv);
    const int q = tid & 15, o4 = 4 * (tid >> 4);
    float2 ca, cy[HA], cx[HB];
    float4 fy01, fy23, fx01, fx23;
    auto load = [&](int pr) {
        // points past the split get a = 0 (their entries, built from the split's last
        // point, then contribute nothing)
        const int p = pr + q;
        const int pv = min(p, pend - 1);
        const float2 cv = coef[pv];
        const float keep = p < pend ? 1.f : 0.f;   // a multiply, not a branch around the load
        ca = make_float2(cv.x * keep, cv.y * keep);
        const float2 *fy = FY + (size_t)pv * ft, *fx = FX + (size_t)pv * ft;
        const float4 *fyf = reinterpret_cast<const float4 *>(fy + nco + (o4 & 15));
        const float4 *fxf = reinterpret_cast<const float4 *>(fx + nco + (o4 & 15));
        fy01 = fyf[0];
        fy23 = fyf[1];
        fx01 = fxf[0];
        fx23 = fxf[1];
#pragma unroll
        for (int h = 0; h < HA; h++)
            cy[h] = fy[(y0 + o4 + RUN * h) >> 4];
#pragma unroll
        for (int h = 0; h < HB; h++)
            cx[h] = fx[(x0 + o4 + RUN * h) >> 4];
    };
    auto store = [&](int buf) {
        unsigned *ah = reinterpret_cast<unsigned *>(sA[buf][0]), *al = reinterpret_cast<unsigned *>(sA[buf][1]);
        unsigned *bh = reinterpret_cast<unsigned *>(sB[buf][0]), *bl = reinterpret_cast<unsigned *>(sB[buf][1]);
        const float2 fyv[4] = {make_float2(fy01.x, fy01.y), make_float2(fy01.z, fy01.w), make_float2(fy23.x, fy23.y),
                               make_float2(fy23.z, fy23.w)};
        const float2 fxv[4] = {make_float2(fx01.x, fx01.y), make_float2(fx01.z, fx01.w), make_float2(fx23.x, fx23.y),
                               make_float2(fx23.z, fx23.w)};
#pragma unroll
        for (int h = 0; h < HA; h++) {
            // A[y][2q] = Re(a Y[p][y]), A[y][2q + 1] = -Im(a Y[p][y])
            const float2 acy = cmul(ca, cy[h]);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float2 w = cmul(acy, fyv[j]);
                unsigned hh, ll;
                split2_bf16(w.x, -w.y, hh, ll);
                const int o = (o4 + RUN * h + j) * WROW + q;
                ah[o] = hh;
                al[o] = ll;
            }
        }
#pragma unroll
        for (int h = 0; h < HB; h++) {
            // B[2q][x] = Re X[p][x], B[2q + 1][x] = Im X[p][x], stored by column x
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float2 w = cmul(cx[h], fxv[j]);
                unsigned hh, ll;
                split2_bf16(w.x, w.y, hh, ll);
                const int o = (o4 + RUN * h + j) * WROW + q;
                bh[o] = hh;
                bl[o] = ll;
            }
        }
    };
    const int wy = (wave >> 1) * (TR * 128 / NTH), wx = (wave & 1) * (TC / 2);
    v16f acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int n = 0; n < NT; n++)
#pragma unroll
            for (int r = 0; r < 16; r++)
                acc[m][n][r] = 0.f;
    int buf = 0;
    if (pstart < pend) {
        load(pstart);
        store(0);
    }
    __syncthreads();
    for (int pr = pstart; pr < pend; pr += AB3_PR) {
        const bool more = pr + AB3_PR < pend;
        if (more)
            load(pr + AB3_PR);
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            bf16x8 ah[MT], al[MT], bh[NT], bl[NT];
#pragma unroll
            for (int m = 0; m < MT; m++) {
                const int o = (wy + 32 * m + lr) * AW_STRIDE + 16 * ks + 8 * lh;
                ah[m] = *reinterpret_cast<const bf16x8 *>(&sA[buf][0][o]);
                al[m] = *reinterpret_cast<const bf16x8 *>(&sA[buf][1][o]);
            }
#pragma unroll
            for (int n = 0; n < NT; n++) {
                const int o = (wx + 32 * n + lr) * AW_STRIDE + 16 * ks + 8 * lh;
                bh[n] = *reinterpret_cast<const bf16x8 *>(&sB[buf][0][o]);
                bl[n] = *reinterpret_cast<const bf16x8 *>(&sB[buf][1][o]);
            }
            // product-major over the accumulators (same per-tile order hh, hl, lh)
#pragma unroll
            for (int m = 0; m < MT; m++)
#pragma unroll
                for (int n = 0; n < NT; n++)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[m], bh[n], acc[m][n], 0, 0, 0);
#pragma unroll
            for (int m = 0; m < MT; m++)
#pragma unroll
                for (int n = 0; n < NT; n++)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[m], bl[n], acc[m][n], 0, 0, 0);
#pragma unroll
            for (int m = 0; m < MT; m++)
#pragma unroll
                for (int n = 0; n < NT; n++)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[m], bh[n], acc[m][n], 0, 0, 0);
        }
        // the other buffer was last read in the previous round, before its barrier
        if (more)
            store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    float *out = part + (size_t)split * W * W;
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int n = 0; n < NT; n++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int y = y0 + wy + 32 * m + 8 * (r >> 2) + 4 * lh + (r & 3), x = x0 + wx + 32 * n + lr;
                if (y < W && x < W)
                    out[(size_t)y * W + x] = acc[m][n][r];
            }
}

template <int TR, int TC>
__global__ __launch_bounds__(ADJ_NTH(TR)) void k_adj_b3wf(const float2 *__restrict__ coef, const float2 *__restrict__ FX,
                                                  const float2 *__restrict__ FY, int nuv, int W, int p_per_split,
                                                  int nsplit, float *__restrict__ part)
{
    k_adj_b3wf_body<TR, TC>((int)blockIdx.x, (int)gridDim.x, coef, FX, FY, nuv, W, p_per_split, nsplit, part);
}

#define ADJ_TILE_INST(K, TR, TC)                                                                                   \
    template __global__ void K<TR, TC>(const float2 *, const float2 *, const float2 *, int, int, int, int, float *);
ADJ_TILE_INST(k_adj_b3wf, 256, 256)
ADJ_TILE_INST(k_adj_b3wf, 128, 256)
ADJ_TILE_INST(k_adj_b3wf, 128, 128)
ADJ_TILE_INST(k_adj_b3wf, 64, 64)
#undef ADJ_TILE_INST

// ---------------------------------------------------------------------------
// FP16 with scaling: one FP16 MFMA product per multiply-add
// ---------------------------------------------------------------------------
// Table entries (|X|, |Y| <= 1) go to FP16 as they are; the image and the adjoint
// coefficients are scaled by the power of two f16_scale(max |x|) first (FP16 tops
// out at 65504: a direction or a coefficient vector can exceed it), and the result
// is scaled back exactly.  A product carries the FP16 rounding of both operands,
// about 2^-11 each: about 500 times the BF16x3 error, at a third of its MFMAs.

// max |x_i| over n floats into *out (FP64 slot); order-free, so deterministic
__global__ __launch_bounds__(RED_THREADS) void k_absmax(const float *__restrict__ x, int n, double *part,
                                                        unsigned *ticket, double *out)
{
    __shared__ float sh[RED_THREADS / 64];
    __shared__ int is_last;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    float m = 0.f;
    for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += gridDim.x * RED_THREADS)
        m = fmaxf(m, fabsf(x[i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        m = fmaxf(m, __shfl_xor(m, o, 64));
    if (lane == 0)
        sh[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float b = sh[0];
        for (int w = 1; w < RED_THREADS / 64; w++)
            b = fmaxf(b, sh[w]);
        __hip_atomic_store(&part[blockIdx.x], (double)b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        is_last = t == gridDim.x - 1;
    }
    __syncthreads();
    if (!is_last)
        return;
    float mm = 0.f;
    for (unsigned b = threadIdx.x; b < gridDim.x; b += RED_THREADS)
        mm = fmaxf(mm, (float)__hip_atomic_load(&part[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        mm = fmaxf(mm, __shfl_xor(mm, o, 64));
    __syncthreads();
    if (lane == 0)
        sh[wid] = mm;
    __syncthreads();
    if (threadIdx.x == 0) {
        float b = sh[0];
        for (int w = 1; w < RED_THREADS / 64; w++)
            b = fmaxf(b, sh[w]);
        *out = (double)b;
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// image float [W*W] -> FP16 bits of img * f16_scale(*amax), same [y][x] layout
__global__ void k_split_image_h(const float *__restrict__ img, int n, const double *__restrict__ amax,
                                unsigned short *__restrict__ Ih)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float sc = f16_scale(*amax);
    Ih[i] = (unsigned short)(pack_f16(img[i] * sc, 0.f) & 0xffffu);
}

// 8 consecutive entries of one component of a table row, as FP16 (see ft_gen8)
__device__ __forceinline__ void ft_gen8_h(float cr, float ci, const float2 (&g)[8], uint4 &h)
{
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; e++)
        v[e] = cr * g[e].x - ci * g[e].y;
    h.x = pack_f16(v[0], v[1]);
    h.y = pack_f16(v[2], v[3]);
    h.z = pack_f16(v[4], v[5]);
    h.w = pack_f16(v[6], v[7]);
}

// Forward DFT in FP16: the tiling, stream-K units and summation order of
// k_vis_fwd_b3f, with one operand plane in LDS and one MFMA per k-step and block
template <int NBK>
__global__ __launch_bounds__(256) void k_vis_fwd_h1f(const unsigned short *__restrict__ Ih,
                                                        const float2 *__restrict__ FX,
                                                        const float2 *__restrict__ FY, int nuv, int W, int upw,
                                                        const float *__restrict__ scale,
                                                        const double *__restrict__ amax, float2 *__restrict__ vis,
                                                        float2 *__restrict__ vpart)
{
    constexpr int ROWS = NBK * 32;
    constexpr int CH = ROWS * (B3_KC / 8);
    constexpr int PER = (CH + 255) / 256;
    static_assert(PER <= 4, "staging holds up to 4 pieces per thread");
    static_assert(B3_KC == 32, "two k-steps per chunk");
    __shared__ __attribute__((aligned(16))) unsigned short sB[2][ROWS * B3_STRIDE];
    __shared__ float2 sY[B3_PW * FYS_MAX];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 31, lh = lane >> 5;
    const int ft = ft_len(W), nco = W / 16;
    const int nchunk = W / B3_KC;
    const int nblk = (nuv + B3_PW - 1) / B3_PW;
    const int total = nblk * nchunk;
    const int u_end = min(total, (int)(blockIdx.x + 1) * upw);
    const float unscale = 1.0f / f16_scale(*amax);   // a power of two: exact
    for (int u = (int)blockIdx.x * upw; u < u_end;) {
        const int blk = u / nchunk, c0 = u % nchunk;
        const int c1 = min(nchunk, c0 + (u_end - u));
        u += c1 - c0;
        const int p0 = blk * B3_PW + wave * 16;
        const int pa = p0 + (lr & 15);
        const float2 *fx = FX + (size_t)(pa < nuv ? pa : 0) * ft;
        float2 g[8];
        {
            const float4 *fp = reinterpret_cast<const float4 *>(fx + nco + 8 * lh);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float4 f = fp[q];
                g[2 * q] = make_float2(f.x, f.y);
                g[2 * q + 1] = make_float2(f.z, f.w);
            }
            if (lr >= 16) {
#pragma unroll
                for (int e = 0; e < 8; e++)
                    g[e] = make_float2(g[e].y, -g[e].x);
            }
        }
        float vr[8], vi[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            vr[r] = 0.f;
            vi[r] = 0.f;
        }
        for (int y0 = 0; y0 < W; y0 += ROWS) {
            v16f acc[NBK];
#pragma unroll
            for (int b = 0; b < NBK; b++)
#pragma unroll
                for (int r = 0; r < 16; r++)
                    acc[b][r] = 0.f;
            uint4 h0 = make_uint4(0, 0, 0, 0), h1 = h0, h2 = h0, h3 = h0;
            auto fetch = [&](int c) {   // one plane of b3_fetch
                const int cc = c;
                auto one = [&](uint4 &h, int q) {
                    const int e = tid + 256 * q;
                    if (CH % 256 == 0 || e < CH)
                        h = *reinterpret_cast<const uint4 *>(Ih + (size_t)(y0 + e / (B3_KC / 8)) * W + cc * B3_KC +
                                                             8 * (e % (B3_KC / 8)));
                };
                if constexpr (PER > 0) one(h0, 0);
                if constexpr (PER > 1) one(h1, 1);
                if constexpr (PER > 2) one(h2, 2);
                if constexpr (PER > 3) one(h3, 3);
            };
            auto store = [&](int buf) {
                auto one = [&](const uint4 &h, int q) {
                    const int e = tid + 256 * q;
                    if (CH % 256 == 0 || e < CH)
                        *reinterpret_cast<uint4 *>(&sB[buf][(e / (B3_KC / 8)) * B3_STRIDE + 8 * (e % (B3_KC / 8))]) = h;
                };
                if constexpr (PER > 0) one(h0, 0);
                if constexpr (PER > 1) one(h1, 1);
                if constexpr (PER > 2) one(h2, 2);
                if constexpr (PER > 3) one(h3, 3);
            };
            FyStage FS;
            fy_fetch(FS, FY, blk, nuv, ft, tid);
            fetch(c0);
            uint4 a0, a1, b0, b1;
            {
                const float4 cx = *reinterpret_cast<const float4 *>(fx + 2 * c0);
                ft_gen8_h(cx.x, cx.y, g, a0);
                ft_gen8_h(cx.z, cx.w, g, a1);
            }
            __syncthreads();   // the previous readers are done with sB and sY
            store(c0 & 1);
            fy_store(FS, sY, ft, tid);
            __syncthreads();
            auto chunk = [&](int c, const uint4 &ca0, const uint4 &ca1, uint4 &d0, uint4 &d1) {
                const int buf = c & 1;
                const int cn = c + 1 < c1 ? c + 1 : c;
                const float4 cxn = *reinterpret_cast<const float4 *>(fx + 2 * cn);
                fetch(cn);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int ks = 0; ks < B3_KC / 16; ks++) {
                    const f16x8 a = __builtin_bit_cast(f16x8, ks ? ca1 : ca0);
#pragma unroll
                    for (int b = 0; b < NBK; b++) {
                        const int o = (b * 32 + lr) * B3_STRIDE + ks * 16 + 8 * lh;
                        const f16x8 bv = *reinterpret_cast<const f16x8 *>(&sB[buf][o]);
                        acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, bv, acc[b], 0, 0, 0);
                    }
                }
                ft_gen8_h(cxn.x, cxn.y, g, d0);
                ft_gen8_h(cxn.z, cxn.w, g, d1);
                store(buf ^ 1);
                __syncthreads();
            };
            int c = c0;
            for (; c + 1 < c1; c += 2) {
                chunk(c, a0, a1, b0, b1);
                chunk(c + 1, b0, b1, a0, a1);
            }
            if (c < c1)
                chunk(c, a0, a1, b0, b1);
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const float2 *fy = sY + (wave * 16 + 8 * (r >> 2) + 4 * lh + (r & 3)) * ft;
                const float2 fj = fy[nco + (lr & 15)];
                const float2 *cy = fy + (y0 >> 4) + (lr >> 4);
#pragma unroll
                for (int b = 0; b < NBK; b++) {
                    const float2 yv = cmul(cy[2 * b], fj);
                    const float tr = acc[b][r], ti = acc[b][r + 8];
                    vr[r] += tr * yv.x - ti * yv.y;
                    vi[r] += tr * yv.y + ti * yv.x;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < 8; r++) {
#pragma unroll
            for (int o = 1; o < 32; o <<= 1) {
                vr[r] += __shfl_xor(vr[r], o, 64);
                vi[r] += __shfl_xor(vi[r], o, 64);
            }
        }
        if (lr == 0) {
            const bool whole = c0 == 0 && c1 == nchunk;
            const float f = (whole && scale ? *scale : 1.0f) * unscale;
            float2 *dst = whole ? vis : vpart + (c0 == 0 ? 0 : nuv);
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const int p = p0 + 8 * (r >> 2) + 4 * lh + (r & 3);
                if (p < nuv)
                    dst[p] = make_float2(vr[r] * f, vi[r] * f);
            }
        }
    }
}

template __global__ void k_vis_fwd_h1f<2>(const unsigned short *, const float2 *, const float2 *, int, int, int,
                                          const float *, const double *, float2 *, float2 *);
template __global__ void k_vis_fwd_h1f<4>(const unsigned short *, const float2 *, const float2 *, int, int, int,
                                          const float *, const double *, float2 *, float2 *);
template __global__ void k_vis_fwd_h1f<6>(const unsigned short *, const float2 *, const float2 *, int, int, int,
                                          const float *, const double *, float2 *, float2 *);
template __global__ void k_vis_fwd_h1f<8>(const unsigned short *, const float2 *, const float2 *, int, int, int,
                                          const float *, const double *, float2 *, float2 *);

// Adjoint in FP16: k_adj_b3wf's tiles, grid, rounds and builder roles with one plane
// per tile and one MFMA per k-step and tile; the coefficients scaled by
// f16_scale(*amax) (max over their Re and Im parts) and the partials scaled back
template <int TR, int TC>
__global__ __launch_bounds__(ADJ_NTH(TR)) void k_adj_h1wf(const float2 *__restrict__ coef, const float2 *__restrict__ FX,
                                                  const float2 *__restrict__ FY, int nuv, int W, int p_per_split,
                                                  int nsplit, const double *__restrict__ amax,
                                                  float *__restrict__ part)
{
    static_assert(((TR == 128 || TR == 256) && (TC == 128 || TC == 256)) || (TR == 64 && TC == 64), "tile");
    // ADJ_NTH(TR) threads; builder runs cover RUN rows (columns) per pass; waves in two
    // columns, each wave TR * 4 / NTH row and TC / 64 column tiles of 32
    constexpr int NTH = ADJ_NTH(TR), RUN = NTH / 4;
    constexpr int HA = TR / RUN, HB = TC / RUN, MT = TR * 4 / NTH, NT = TC / 64;
    __shared__ __attribute__((aligned(16))) unsigned short sA[2][TR * AW_STRIDE];
    __shared__ __attribute__((aligned(16))) unsigned short sB[2][TC * AW_STRIDE];
    constexpr int WROW = AW_STRIDE / 2;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 31, lh = lane >> 5;
    const int ft = ft_len(W), nco = W / 16;
    const int split = blockIdx.x % nsplit, tile = blockIdx.x / nsplit;
    const int qx = (W + TC - 1) / TC;   // edge tiles of a width that is not a multiple of TC are partial
    const int y0 = (tile / qx) * TR, x0 = (tile % qx) * TC;
    const int pstart = split * p_per_split;
    const int pend = min(pstart + p_per_split, nuv);
    const int q = tid & 15, o4 = 4 * (tid >> 4);
    const float sc = f16_scale(*amax), unscale = 1.0f / sc;
    float2 ca, cy[HA], cx[HB];
    float4 fy01, fy23, fx01, fx23;
    auto load = [&](int pr) {
        const int p = pr + q;
        const int pv = min(p, pend - 1);
        const float2 cv = coef[pv];
        const float keep = p < pend ? sc : 0.f;
        ca = make_float2(cv.x * keep, cv.y * keep);
        const float2 *fy = FY + (size_t)pv * ft, *fx = FX + (size_t)pv * ft;
        const float4 *fyf = reinterpret_cast<const float4 *>(fy + nco + (o4 & 15));
        const float4 *fxf = reinterpret_cast<const float4 *>(fx + nco + (o4 & 15));
        fy01 = fyf[0];
        fy23 = fyf[1];
        fx01 = fxf[0];
        fx23 = fxf[1];
#pragma unroll
        for (int h = 0; h < HA; h++)
            cy[h] = fy[(y0 + o4 + RUN * h) >> 4];
#pragma unroll
        for (int h = 0; h < HB; h++)
            cx[h] = fx[(x0 + o4 + RUN * h) >> 4];
    };
    auto store = [&](int buf) {
        unsigned *aw = reinterpret_cast<unsigned *>(sA[buf]), *bw = reinterpret_cast<unsigned *>(sB[buf]);
        const float2 fyv[4] = {make_float2(fy01.x, fy01.y), make_float2(fy01.z, fy01.w), make_float2(fy23.x, fy23.y),
                               make_float2(fy23.z, fy23.w)};
        const float2 fxv[4] = {make_float2(fx01.x, fx01.y), make_float2(fx01.z, fx01.w), make_float2(fx23.x, fx23.y),
                               make_float2(fx23.z, fx23.w)};
#pragma unroll
        for (int h = 0; h < HA; h++) {
            const float2 acy = cmul(ca, cy[h]);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float2 w = cmul(acy, fyv[j]);
                aw[(o4 + RUN * h + j) * WROW + q] = pack_f16(w.x, -w.y);
            }
        }
#pragma unroll
        for (int h = 0; h < HB; h++) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float2 w = cmul(cx[h], fxv[j]);
                bw[(o4 + RUN * h + j) * WROW + q] = pack_f16(w.x, w.y);
            }
        }
    };
    const int wy = (wave >> 1) * (TR * 128 / NTH), wx = (wave & 1) * (TC / 2);
    v16f acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int n = 0; n < NT; n++)
#pragma unroll
            for (int r = 0; r < 16; r++)
                acc[m][n][r] = 0.f;
    int buf = 0;
    if (pstart < pend) {
        load(pstart);
        store(0);
    }
    __syncthreads();
    for (int pr = pstart; pr < pend; pr += AB3_PR) {
        const bool more = pr + AB3_PR < pend;
        if (more)
            load(pr + AB3_PR);
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            f16x8 a[MT], b[NT];
#pragma unroll
            for (int m = 0; m < MT; m++)
                a[m] = *reinterpret_cast<const f16x8 *>(&sA[buf][(wy + 32 * m + lr) * AW_STRIDE + 16 * ks + 8 * lh]);
#pragma unroll
            for (int n = 0; n < NT; n++)
                b[n] = *reinterpret_cast<const f16x8 *>(&sB[buf][(wx + 32 * n + lr) * AW_STRIDE + 16 * ks + 8 * lh]);
#pragma unroll
            for (int m = 0; m < MT; m++)
#pragma unroll
                for (int n = 0; n < NT; n++)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[m], b[n], acc[m][n], 0, 0, 0);
        }
        if (more)
            store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    float *out = part + (size_t)split * W * W;
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int n = 0; n < NT; n++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int y = y0 + wy + 32 * m + 8 * (r >> 2) + 4 * lh + (r & 3), x = x0 + wx + 32 * n + lr;
                if (y < W && x < W)
                    out[(size_t)y * W + x] = acc[m][n][r] * unscale;
            }
}

template __global__ void k_adj_h1wf<256, 256>(const float2 *, const float2 *, const float2 *, int, int, int, int,
                                              const double *, float *);
template __global__ void k_adj_h1wf<128, 256>(const float2 *, const float2 *, const float2 *, int, int, int, int,
                                              const double *, float *);
template __global__ void k_adj_h1wf<128, 128>(const float2 *, const float2 *, const float2 *, int, int, int, int,
                                              const double *, float *);
template __global__ void k_adj_h1wf<64, 64>(const float2 *, const float2 *, const float2 *, int, int, int, int,
                                              const double *, float *);
