// Device kernels of the MEM-CG path (gfx950 / CDNA4, wave64).
//
// Layouts in HBM (one reconstruction):
//   image-sized float buffers   [W*W], pixel (x, y) at x + W*y
//   X, Y tables                 float2 [nuv][W]  (X = exp(+2 pi i a u x), Y = exp(-2 pi i a v y))
//   vis, coef                   float2 [nuv]
//   data, err, mock             float  [npow + 2*nbis]  ([V2 | (Re, Im) per T3])
//   tri                         int4   [nbis]  (ab, bc, ca uv index, 3 sign bits: bit j set = conjugate leg j)
//   csr_ptr/csr_ent             uv -> T3 legs touching it (entry = 3*t + leg), ascending t
//   scal                        double [GPAIR_NSCAL] reduction results, invflux float alongside
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

enum {
    SC_FLUX = 0,
    SC_CHI2 = 1,
    SC_ENT = 2,
    SC_DOT = 3,
    SC_C0 = 4,
    SC_ZG = 5,     // z.g, z.g_old          (preconditioned CG)
    SC_ZGO = 6,
    SC_W1 = 7,     // d.g, sum d, sum |d|
    SC_FD = 8,
    SC_DABS = 9,
    SC_FDELTA = 10,  // host copy only, added up from the commit partials (RB_CSTEP):
    SC_DSENT = 11,   // sum of the committed image change, its entropy change, and the
    SC_NCLAMP = 12,  // number of pixels the floor clamped (a diagnostic)
    SC_BETA = 13,    // beta the direction kernel used
    SC_PXF = 14,     // flux of the pixel-move state
    SC_PXC = 15,     // chi2 of the pixel-move state
    SC_STEP = 16,    // step chosen on the device from a line-search batch (0: none)
    SC_KBEST = 17,   // its candidate index (-1: no decrease, or not a descent direction)
    SC_AMAXI = 18,   // max |x| of the image of the last FP16 forward DFT
    SC_AMAXC = 19,   // max |Re|, |Im| of the adjoint coefficients of the last FP16 adjoint
    SC_DCHI2 = 20,   // host copy only: chi2 change of the committed image (from RB_DCHI2)
    GPAIR_NSCAL = 21
};

#define RED_THREADS 256
#define RED_MAX_BLOCKS 2048
#define RED_GRID_MAX 128   // blocks of the grid-stride reductions (red_blocks)
#define RED_MAX_NV 32
#define GPAIR_NTICKET 16   // reduction tickets per context (the line-search batch uses 8)

// batched line search: LS_K candidate steps, s_0 = 0 and s_k = s_ref * 2^((k - LS_TOP) / 2)
#define LS_K 32
#define LS_TOP 26
#define LS_G 8   // line-search candidates per wave of the chi2 half and per block row of the entropy half

// The accelerated loop's read-back region, one device allocation read back in one
// copy: the scalars, the line-search sums, then the commit kernels' per-block
// partials (k_commit_step: sum delta, entropy change, clamped pixels per block;
// k_commit_dchi2: chi2 change per block), which the host adds up in block order.
#define RB_LS GPAIR_NSCAL
#define RB_CSTEP (RB_LS + 2 * LS_K)
#define RB_DCHI2 (RB_CSTEP + 3 * RED_GRID_MAX)
#define RB_COUNT (RB_DCHI2 + RED_GRID_MAX)

// candidate k of a batch: s_0 = 0, s_k = s_ref 2^((k - LS_TOP) / 2), built exactly
// (a power of two times 1 or the rounded sqrt 2) so host and device agree bit for bit
__host__ __device__ inline double ls_step(double s_ref, int k)
{
    if (k == 0)
        return 0.0;
    const int e = k - LS_TOP;
    const int h = e >= 0 ? e / 2 : -((1 - e) / 2);   // floor(e / 2)
    return s_ref * ldexp((e & 1) ? 1.4142135623730951 : 1.0, h);
}

// vertex of the parabola through (a, Ja), (b, Jb), (c, Jc), a < b < c, when it is a
// minimum inside (a, c); otherwise b.  No contraction: host and device round alike.
__host__ __device__ inline double parabola_step(double a, double b, double cc, double Ja, double Jb, double Jc)
{
#pragma clang fp contract(off)
    const double p = (b - a) * (Jb - Jc), q = (b - cc) * (Jb - Ja);
    const double den = p - q;
    if (!(den < 0.0 || den > 0.0))
        return b;
    const double s = b - 0.5 * ((b - a) * p - (b - cc) * q) / den;
    const double curv = (Ja - Jb) / (a - b) - (Jb - Jc) / (b - cc);
    if (!(curv < 0.0) || !(s > a && s < cc))
        return b;
    return s;
}

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// split precision BF16x3: x = hi + lo with hi = bf16(x), lo = bf16(x - hi), both
// rounded to nearest even; a*b ~ ah*bh + ah*bl + al*bh (al*bl, ~2^-16 |ab|, dropped),
// every product exact in the FP32 MFMA accumulator
__device__ inline unsigned short bf16_bits_rne(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ inline void split_bf16(float x, unsigned short &hi, unsigned short &lo)
{
    hi = bf16_bits_rne(x);
    lo = bf16_bits_rne(x - __uint_as_float((unsigned)hi << 16));
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// the same split for two values at once on gfx950's packed converter
// (v_cvt_pk_bf16_f32, round to nearest even like bf16_bits_rne for finite
// values): h and l hold (x, y) as bf16 pairs, x in the low half
__device__ inline void split2_bf16(float x, float y, unsigned &h, unsigned &l)
{
    const f32x2 v = {x, y};
    h = __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2));
    const f32x2 r = {x - __uint_as_float(h << 16), y - __uint_as_float(h & 0xffff0000u)};
    l = __builtin_bit_cast(unsigned, __builtin_convertvector(r, bf16x2));
}

typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// two floats as a packed FP16 pair (v_cvt_pk_f16_f32, round to nearest even), x low
__device__ inline unsigned pack_f16(float x, float y)
{
    const f32x2 v = {x, y};
    return __builtin_bit_cast(unsigned, __builtin_convertvector(v, f16x2));
}

// FP16 with scaling: a power of two that maps the largest magnitude m into
// [2^13, 2^14), leaving headroom below FP16's 65504; 1 when m is 0 or not finite
__device__ inline float f16_scale(double m)
{
    if (!(m > 0.0) || !(m < 1e300))
        return 1.0f;
    int e;
    (void)frexp(m, &e);   // m = f 2^e, f in [0.5, 1)
    return ldexpf(1.0f, 14 - e);
}

// ---------------------------------------------------------------------------
// deterministic block reduction + last-block finish (fixed summation order)
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

template <int NV>
__device__ inline void block_sum(double (&v)[NV], double (*sh)[RED_THREADS / 64])
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NV; j++) {
        double s = wave_sum(v[j]);
        if (lane == 0)
            sh[j][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NV; j++) {
            double s = 0.0;
            for (int w = 0; w < RED_THREADS / 64; w++)
                s += sh[j][w];
            v[j] = s;
        }
    }
    __syncthreads();
}

// sum of p[b * stride] over b < count in a fixed order, the same bits in every
// block, returned to every thread (blockDim.x == RED_THREADS)
__device__ inline double block_fixed_sum(const double *p, int stride, int count)
{
    __shared__ double sh[1][RED_THREADS / 64];
    __shared__ double total;
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < count; b += RED_THREADS)
        v[0] += p[(size_t)b * stride];
    block_sum<1>(v, sh);
    if (threadIdx.x == 0)
        total = v[0];
    __syncthreads();
    return total;
}

// red_sum_parts: out[j] = sum over b < count of part[b*NV + j], thread t taking b =
// t, t + RED_THREADS, ... in order, then block_sum (and, for the flux slot, the
// float inverse): the finish of the ticketed reductions below, and of their
// ticket-free form in a kernel of its own (same order, same bits).
template <int NV>
__device__ inline void red_sum_parts(const double *part, int count, double *out, float *invflux_out)
{
    __shared__ double shp[NV][RED_THREADS / 64];
    double acc[NV];
#pragma unroll
    for (int j = 0; j < NV; j++)
        acc[j] = 0.0;
    for (int b = threadIdx.x; b < count; b += RED_THREADS) {
#pragma unroll
        for (int j = 0; j < NV; j++)
            acc[j] += __hip_atomic_load(&part[b * NV + j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    block_sum<NV>(acc, shp);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NV; j++)
            out[j] = acc[j];
        if (invflux_out)
            *invflux_out = (float)(1.0 / acc[0]);
    }
}

// Every block calls this once with its thread-partial values.  Block sums go to
// part[bx*NV + j]; the last block to arrive sums them in block order and writes
// out[j] (and, for the flux slot, the float inverse).  The ticket is reset by the
// last block so no memset is needed between launches.  bx of nbx blocks is given
// explicitly for remapped launches (reduce_finish: blockIdx.x of gridDim.x).  With
// no ticket every block only stores its partial (red_sum_parts adds them up later).
template <int NV>
__device__ inline void reduce_finish_at(double (&v)[NV], double *part, unsigned *ticket, double *out,
                                        float *invflux_out, int bx, int nbx)
{
    __shared__ double sh[NV][RED_THREADS / 64];
    __shared__ int is_last;
    block_sum<NV>(v, sh);
    const unsigned nblk = (unsigned)nbx;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NV; j++)
            __hip_atomic_store(&part[bx * NV + j], v[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (ticket) {
            unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            is_last = (t == nblk - 1);
        }
    }
    if (!ticket)
        return;
    __syncthreads();
    if (!is_last)
        return;
    red_sum_parts<NV>(part, nbx, out, invflux_out);
    if (threadIdx.x == 0)
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int NV>
__device__ inline void reduce_finish(double (&v)[NV], double *part, unsigned *ticket, double *out,
                                     float *invflux_out)
{
    reduce_finish_at<NV>(v, part, ticket, out, invflux_out, (int)blockIdx.x, (int)gridDim.x);
}

// Launch-shape independent reduction over items [0, n): the range is cut into
// nparts = min(RED_PARTS, ceil(n / (2 RED_THREADS))) partitions; partition q is
// summed by one block, thread t taking items q*span + t, + RED_THREADS, ... in
// order, then block_sum, into part[q*NV + j].  Block b takes partitions b,
// b + gridDim.x, ...; the last block to arrive sums the partials in a fixed order.
// The result does not depend on the number of blocks, so a tuned grid changes no
// bit of it.  item(i, v) adds item i into v[0..NV-1].
#define RED_PARTS 512

__host__ __device__ inline int red_nparts(int n, int span)
{
    int nparts = (n + span - 1) / span;
    return nparts < 1 ? 1 : (nparts > RED_PARTS ? RED_PARTS : nparts);
}

// partitioned_reduce_at: the same with partitions of at least `span` items, and
// this block's index bx among the nbx blocks that share the ticket given
// explicitly (for remapped launches).
template <int NV, typename F>
__device__ inline void partitioned_reduce_at(int n, int span_min, F item, double *part, unsigned *ticket,
                                             double *out, int bx, int nbx, float *invflux_out = nullptr)
{
    __shared__ double sh[NV][RED_THREADS / 64];
    __shared__ int is_last;
    const int nparts = red_nparts(n, span_min);
    const int span = (n + nparts - 1) / nparts;
    for (int q = bx; q < nparts; q += nbx) {
        double v[NV];
#pragma unroll
        for (int j = 0; j < NV; j++)
            v[j] = 0.0;
        const int beg = q * span, end = beg + span < n ? beg + span : n;
        for (int i = beg + threadIdx.x; i < end; i += RED_THREADS)
            item(i, v);
        block_sum<NV>(v, sh);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int j = 0; j < NV; j++)
                __hip_atomic_store(&part[q * NV + j], v[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if (!ticket)   // ticket-free: the partials only (red_sum_parts adds them up later)
        return;
    if (threadIdx.x == 0) {
        const unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        is_last = (t == (unsigned)nbx - 1);
    }
    __syncthreads();
    if (!is_last)
        return;
    red_sum_parts<NV>(part, nparts, out, invflux_out);
    if (threadIdx.x == 0)
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// partitioned_partials: the partition sums of partitioned_reduce_at alone, stored
// to part[q*NV + j] with plain stores, no ticket and no finish: a later launch sums
// them in a fixed order.  On MI355X a per-block same-address ticket (an agent-scope
// RMW plus release / acquire) costs more than the line-search arithmetic it guards
// once there are a few hundred blocks; a kernel boundary publishes the partials.
template <int NV, typename F>
__device__ inline void partitioned_partials(int n, int span_min, F item, double *part, int bx, int nbx)
{
    __shared__ double sh[NV][RED_THREADS / 64];
    const int nparts = red_nparts(n, span_min);
    const int span = (n + nparts - 1) / nparts;
    for (int q = bx; q < nparts; q += nbx) {
        double v[NV];
#pragma unroll
        for (int j = 0; j < NV; j++)
            v[j] = 0.0;
        const int beg = q * span, end = beg + span < n ? beg + span : n;
        for (int i = beg + threadIdx.x; i < end; i += RED_THREADS)
            item(i, v);
        block_sum<NV>(v, sh);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int j = 0; j < NV; j++)
                part[q * NV + j] = v[j];
        }
    }
}

// lane k's value of x (wave-uniform result, kept in scalar registers)
__device__ __forceinline__ double readlane_d(double x, int k)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <int NV, typename F>
__device__ inline void partitioned_reduce(int n, F item, double *part, unsigned *ticket, double *out,
                                         float *invflux_out = nullptr)
{
    partitioned_reduce_at<NV>(n, 2 * RED_THREADS, item, part, ticket, out, (int)blockIdx.x, (int)gridDim.x,
                              invflux_out);
}

// ---------------------------------------------------------------------------
// phase helpers: exp(2 pi i t) with t range-reduced in FP64
// ---------------------------------------------------------------------------
__device__ inline float2 turn_exp(double t)
{
    double f = t - rint(t);
    float s, c;
    sincospif((float)(2.0 * f), &s, &c);
    return make_float2(c, s);
}

__device__ inline float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
