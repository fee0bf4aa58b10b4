/*
 * tp_mlp.hip -- the learned field surrogate (no counterpart in the
 * reference): bfloat16 matrix-core kernels for its training (forward,
 * backward, weight gradients, Adam) and for its evaluation inside the FP32
 * Boris step, where the exact mode sum is the per-particle fallback.  C ABI:
 * tp_gpu_mlp_*, tp_gpu_train_surrogate, tp_gpu_set_field_mlp
 * (include/tp_gpu.h).
 *
 * Network: 256 Fourier features [sin 2 pi (q_j . x), cos 2 pi (q_j . x)],
 * three 128-wide gelu layers (bfloat16 weights and activations, FP32
 * accumulation on the matrix cores), an FP32 128 -> 3 output layer.
 *
 * Matrix-core layout of v_mfma_f32_16x16x32_bf16 (one wave of 64 lanes,
 * g = lane / 16, i = lane % 16):
 *   A operand (16 x 32): lane holds A[i][8g .. 8g+7]
 *   B operand (32 x 16): lane holds B[8g .. 8g+7][i]
 *   C / D     (16 x 16): lane holds C[4g .. 4g+3][i]
 * A wave evaluates 64 samples, sample m in lane m.  Layers are computed
 * transposed, Z^T = W^T X^T: the W^T fragments are the A operand, stored in
 * global memory in fragment order (16 contiguous bytes per lane, one
 * kilobyte per fragment, read coalesced and L2-resident); the layer input X
 * is the wave's row-major bfloat16 tile in LDS (B operand, 16 bytes per
 * lane); the result lands as 4 consecutive output features of one sample
 * per lane, which the epilogue (bias, gelu) writes back into the tile as
 * the next layer's input.  The feature layer is done in two halves (sin,
 * then cos) so that a wave's tile is 64 x 128 (17 KB, 8 waves per CU).
 *
 * Backward: dZ_{l-1} = (dZ_l W_l^T) * gelu'(Z_{l-1}) with W_l fragments in
 * a second order (rows = layer input); weight gradients dW_l = X_{l-1}^T dZ_l
 * per chunk of 256 samples, with both operands transposed through LDS so the
 * reduction runs along the samples; chunk partials are summed in a fixed
 * order by the optimiser kernel (deterministic).
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <new>

#include "../../include/tp_gpu.h"
#include "tp_device.h"
#include "tp_ctx.h"
#include "tp_track.h"

#define NF 128
#define NIN 256
#define NH 128
#define P_W1 0
#define P_B1 (P_W1 + NIN * NH)
#define P_W2 (P_B1 + NH)
#define P_B2 (P_W2 + NH * NH)
#define P_W3 (P_B2 + NH)
#define P_B3 (P_W3 + NH * NH)
#define P_W4 (P_B3 + NH)
#define P_B4 (P_W4 + NH * 3)
#define P_TOTAL (P_B4 + 3)
static_assert(P_TOTAL == TP_MLP_NPARAMS, "parameter count");

/* LDS tile of a wave: 64 rows of 128 bf16, row stride 136 (272 bytes: the
 * 16 rows of one fragment read start on 16 distinct 4-bank groups) */
#define TS 136
#define TILE (64 * TS)

/* fragment arrays, in bf16x8 units: W1^T (64 fragments), W2^T, W3^T (32
 * each), W2, W3 (32 each, backward order) */
#define FR_W1F 0
#define FR_W2F 4096
#define FR_W3F 6144
#define FR_W2B 8192
#define FR_W3B 10240
#define FR_TOTAL 12288

/* element index of W[k][n] (layer with K inputs) among the forward
 * fragments: fragment (nb = n/16, kb = k/32), lane (k%32)/8 * 16 + n%16,
 * element k%8 */
__host__ __device__ inline int frag_fwd(int k, int n, int KB)
{
    return (((n >> 4) * KB + (k >> 5)) * 64 + ((k >> 3) & 3) * 16 + (n & 15)) * 8 + (k & 7);
}

/* element index of W[k][n] (128 outputs) among the backward fragments:
 * fragment (ib = k/16, kb = n/32), lane (n%32)/8 * 16 + k%16, element n%8 */
__host__ __device__ inline int frag_bwd(int k, int n)
{
    return (((k >> 4) * 4 + (n >> 5)) * 64 + ((n >> 3) & 3) * 16 + (k & 15)) * 8 + (n & 7);
}

struct TrainBufs {
    __bf16 *h0, *h1, *h2;        /* layer inputs, row-major [rows][256 | 128] */
    float *g1, *g2, *g3, *h3;    /* gelu'(z1..z3) and h3, [rows][128] */
};

enum { MLP_EVAL = 0, MLP_TRAIN = 1 };

#define GELU_C1 1.5957691216057308f   /* 2 sqrt(2/pi) */
#define GELU_C2 0.044715f
#define LOG2E 1.4426950408889634f

/* gelu(z) = z sigma(a), a = 2 sqrt(2/pi) (z + 0.044715 z^3) (the tanh form
 * of Hendrycks & Gimpel); optionally gelu'(z) = sigma + z sigma (1 - sigma)
 * da/dz.  exp2 and the reciprocal are the hardware ones (v_exp_f32,
 * v_rcp_f32); for a -> -inf, sigma -> 0 and both results go to 0 without
 * NaN. */
template <bool DERIV>
__device__ __forceinline__ float gelu_f(float z, float &gd)
{
    float z2 = z * z;
    /* exp(-a) = 2^(-a log2 e) */
    float e = __builtin_amdgcn_exp2f((-GELU_C1 * LOG2E) * z * fmaf(GELU_C2, z2, 1.0f));
    float s = __builtin_amdgcn_rcpf(1.0f + e);
    if (DERIV) gd = fmaf(z * s * (1.0f - s) * GELU_C1, fmaf(3.0f * GELU_C2, z2, 1.0f), s);
    return z * s;
}

/* The network's activation: gelu(z) = max(z, 0) - G(|z|), with the tail
 * G(u) = u - gelu(u) (= u Phi(-u) for the tanh-form Phi) replaced by a
 * degree-8 polynomial in s = min(u, 4) / 2 - 1, evaluated on pairs of
 * values in packed FP16: the FP32 accumulators are converted with round
 * toward zero (v_cvt_pkrtz), then min, FMA, Horner, max and the difference
 * are v_pk_* instructions -- no transcendental.  Within 2^-10 |z| + 1.1e-3
 * of the tanh form for |z| < 65504 (6.3e-4 for z <= 0; 2^-10 |z| is the
 * FP16 rounding of the input), below the bfloat16 rounding the hidden
 * activations get anyway; the coefficients (least squares on the
 * Chebyshev basis of [0, 4], then rounded to FP16) are exact FP16 values. */
typedef _Float16 h16x2 __attribute__((ext_vector_type(2)));
#define H2(v) ((h16x2){(_Float16)(v), (_Float16)(v)})
#define GELU_T0 0.045379638671875f
#define GELU_T1 -0.1727294921875f
#define GELU_T2 0.21728515625f
#define GELU_T3 0.0158538818359375f
#define GELU_T4 -0.3017578125f
#define GELU_T5 0.2413330078125f
#define GELU_T6 0.018310546875f
#define GELU_T7 -0.08465576171875f
#define GELU_T8 0.020843505859375f

__device__ __forceinline__ void gelu_pk(float za, float zb, float &ga, float &gb)
{
    const h16x2 z = __builtin_bit_cast(h16x2, __builtin_amdgcn_cvt_pkrtz(za, zb));
    const h16x2 uc = __builtin_elementwise_min(__builtin_elementwise_abs(z), H2(4.0f));
    const h16x2 sv = __builtin_elementwise_fma(uc, H2(0.5f), H2(-1.0f));
    h16x2 p = H2(GELU_T8);
    p = __builtin_elementwise_fma(p, sv, H2(GELU_T7));
    p = __builtin_elementwise_fma(p, sv, H2(GELU_T6));
    p = __builtin_elementwise_fma(p, sv, H2(GELU_T5));
    p = __builtin_elementwise_fma(p, sv, H2(GELU_T4));
    p = __builtin_elementwise_fma(p, sv, H2(GELU_T3));
    p = __builtin_elementwise_fma(p, sv, H2(GELU_T2));
    p = __builtin_elementwise_fma(p, sv, H2(GELU_T1));
    p = __builtin_elementwise_fma(p, sv, H2(GELU_T0));
    const h16x2 g = __builtin_elementwise_max(z, H2(0.0f)) - p;
    ga = (float)g.x;
    gb = (float)g.y;
}

/* gelu'(z) of the tanh form in FP32 (the hardware exp2 and reciprocal),
 * for the backward pass */
__device__ __forceinline__ float gelu_deriv(float z)
{
    float d = 0.f;
    (void)gelu_f<true>(z, d);
    return d;
}

__device__ __forceinline__ f32x4 mfma(const bf16x8 &a, const bf16x8 &b, const f32x4 &c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

/* hidden-layer epilogue: bias, gelu; layers 1-2 write the bf16 result into
 * the tile (and in training, h and gelu' to the batch buffers); layer 3
 * keeps FP32 h in acc for the output layer. */
template <int MODE, int LAYER>
__device__ __forceinline__ void hidden_epilogue(f32x4 (&acc)[8][4], __bf16 *tile,
                                                const TrainBufs &tb, long row0)
{
    const int lane = threadIdx.x & 63, g = lane >> 4, i16 = lane & 15;
#pragma unroll
    for (int nb = 0; nb < 8; nb++) {
        const int n0 = nb * 16 + 4 * g;
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            const int m = mb * 16 + i16;
            f32x4 h, gd;
            float h0, h1, h2, h3;
            gelu_pk(acc[nb][mb][0], acc[nb][mb][1], h0, h1);
            gelu_pk(acc[nb][mb][2], acc[nb][mb][3], h2, h3);
            h[0] = h0; h[1] = h1; h[2] = h2; h[3] = h3;
            if (MODE == MLP_TRAIN) {
#pragma unroll
                for (int r = 0; r < 4; r++) gd[r] = gelu_deriv(acc[nb][mb][r]);
            }
            if (LAYER < 3) {
                bf16x4 hb;
                hb[0] = (__bf16)h[0]; hb[1] = (__bf16)h[1];
                hb[2] = (__bf16)h[2]; hb[3] = (__bf16)h[3];
                *(bf16x4 *)(tile + m * TS + n0) = hb;
                if (MODE == MLP_TRAIN) {
                    __bf16 *ho = LAYER == 1 ? tb.h1 : tb.h2;
                    float *go = LAYER == 1 ? tb.g1 : tb.g2;
                    *(bf16x4 *)(ho + (row0 + m) * NH + n0) = hb;
                    *(f32x4 *)(go + (row0 + m) * NH + n0) = gd;
                }
            } else {
                acc[nb][mb] = h;
                if (MODE == MLP_TRAIN) {
                    *(f32x4 *)(tb.h3 + (row0 + m) * NH + n0) = h;
                    *(f32x4 *)(tb.g3 + (row0 + m) * NH + n0) = gd;
                }
            }
        }
    }
}

/* accumulators start at the layer's bias (the 4 outputs of this lane) */
__device__ __forceinline__ void acc_bias(f32x4 (&acc)[8][4], const float *bias)
{
    const int g = (threadIdx.x & 63) >> 4;
    /* not hoisted into the previous layer's epilogue, where acc is live */
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int nb = 0; nb < 8; nb++) {
        const f32x4 b = *(const f32x4 *)(bias + nb * 16 + 4 * g);
#pragma unroll
        for (int mb = 0; mb < 4; mb++) acc[nb][mb] = b;
    }
}

/* acc += W^T X^T over k-blocks kb0 .. kb0 + nkb - 1 of the layer (tile
 * column kb * 32 holds k-block kb0 + kb of the input); the first half of
 * the weight fragments of k-block kb + 1 is loaded while kb is multiplied,
 * the second half at the start of its own k-block, behind the MFMAs of the
 * first half (register budget of two waves per SIMD) */
__device__ __forceinline__ void mfma_kblocks(f32x4 (&acc)[8][4], const bf16x8 *wf, int KB,
                                             int kb0, int nkb, const __bf16 *tile)
{
    const int lane = threadIdx.x & 63, g = lane >> 4, i16 = lane & 15;
    bf16x8 an[4];
#pragma unroll
    for (int nb = 0; nb < 4; nb++) an[nb] = wf[(nb * KB + kb0) * 64 + lane];
#pragma unroll 1
    for (int kb = 0; kb < nkb; kb++) {
        bf16x8 a[8];
#pragma unroll
        for (int nb = 0; nb < 4; nb++) a[nb] = an[nb];
#pragma unroll
        for (int nb = 4; nb < 8; nb++) a[nb] = wf[(nb * KB + kb0 + kb) * 64 + lane];
        if (kb + 1 < nkb) {
#pragma unroll
            for (int nb = 0; nb < 4; nb++) an[nb] = wf[(nb * KB + kb0 + kb + 1) * 64 + lane];
        }
        bf16x8 bx[4];
#pragma unroll
        for (int mb = 0; mb < 4; mb++)
            bx[mb] = *(const bf16x8 *)(tile + (mb * 16 + i16) * TS + kb * 32 + 8 * g);
#pragma unroll
        for (int nb = 0; nb < 8; nb++)
#pragma unroll
            for (int mb = 0; mb < 4; mb++) acc[nb][mb] = mfma(a[nb], bx[mb], acc[nb][mb]);
    }
}

/* one 128 -> 128 layer: acc = b + W^T X^T from the wave's tile */
__device__ __forceinline__ void layer128(f32x4 (&acc)[8][4], const bf16x8 *wf,
                                         const float *bias, const __bf16 *tile)
{
    acc_bias(acc, bias);
    mfma_kblocks(acc, wf, 4, 0, 4, tile);
}

/* The network at the wave's 64 points (x, y, z of this lane's sample);
 * every lane of the wave must call it.  out = h3 W4 + b4 of this lane's
 * sample.  In training, row0 is the batch row of lane 0 and the
 * activations are stored for the backward pass. */
template <int MODE>
__device__ __forceinline__ void mlp_wave(const MlpNet &net, float x, float y, float z,
                                         __bf16 *tile, float out[3], const TrainBufs &tb,
                                         long row0)
{
    const int lane = threadIdx.x & 63, g = lane >> 4, i16 = lane & 15;
    typedef const float __attribute__((address_space(4))) *cf;
    const cf q = (cf)net.q;
    f32x4 acc[8][4];
    acc_bias(acc, net.b1);

    /* layer 1 over the sin half, then the cos half of the features */
#pragma unroll
    for (int h = 0; h < 2; h++) {
#pragma unroll 1
        for (int j8 = 0; j8 < NF / 8; j8++) {
            bf16x8 v;
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const int j = j8 * 8 + t;
                /* phase in revolutions, reduced to [0, 1) (the hardware sine
                 * and cosine take revolutions) */
                float u = fmaf(q[3 * j + 2], z, fmaf(q[3 * j + 1], y, q[3 * j] * x));
                float r = __builtin_amdgcn_fractf(u);
                v[t] = (__bf16)(h ? __builtin_amdgcn_cosf(r) : __builtin_amdgcn_sinf(r));
            }
            *(bf16x8 *)(tile + lane * TS + j8 * 8) = v;
            if (MODE == MLP_TRAIN)
                *(bf16x8 *)(tb.h0 + (row0 + lane) * NIN + h * NF + j8 * 8) = v;
        }
        mfma_kblocks(acc, net.w1f, 8, h * 4, 4, tile);
    }
    hidden_epilogue<MODE, 1>(acc, tile, tb, row0);
    layer128(acc, net.w2f, net.b2, tile);
    hidden_epilogue<MODE, 2>(acc, tile, tb, row0);
    layer128(acc, net.w3f, net.b3, tile);
    hidden_epilogue<MODE, 3>(acc, tile, tb, row0);

    /* output layer in FP32: partial sums over this lane's 32 hidden units
     * for its 4 samples, completed across the 4 lane groups */
    float o[4][3];
#pragma unroll
    for (int mb = 0; mb < 4; mb++) o[mb][0] = o[mb][1] = o[mb][2] = 0.f;
#pragma unroll
    for (int nb = 0; nb < 8; nb++) {
        const f32x4 *wp = (const f32x4 *)(net.w4 + 3 * (nb * 16 + 4 * g));
        const f32x4 w0 = wp[0], w1 = wp[1], w2 = wp[2];
        const float w[12] = {w0[0], w0[1], w0[2], w0[3], w1[0], w1[1],
                             w1[2], w1[3], w2[0], w2[1], w2[2], w2[3]};
#pragma unroll
        for (int mb = 0; mb < 4; mb++)
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int c = 0; c < 3; c++) o[mb][c] = fmaf(acc[nb][mb][r], w[3 * r + c], o[mb][c]);
    }
#pragma unroll
    for (int mb = 0; mb < 4; mb++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            o[mb][c] += __shfl_xor(o[mb][c], 16, 64);
            o[mb][c] += __shfl_xor(o[mb][c], 32, 64);
        }
    /* lane m = 16 g + i holds the sums of sample m in o[g] */
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float v = g == 0 ? o[0][c] : g == 1 ? o[1][c] : g == 2 ? o[2][c] : o[3][c];
        out[c] = v + net.b4[c];
    }
}

/* ------------------------------------------------------------------ */
/* kernels                                                             */
/* ------------------------------------------------------------------ */

/* network outputs at n points (double x, y, z) */
__global__ void __launch_bounds__(256, 2) k_mlp_eval(MlpNet net, int n, const double *x,
                                                    double *out)
{
    __shared__ __attribute__((aligned(16))) __bf16 tiles[4 * TILE];
    const int i = blockIdx.x * 256 + threadIdx.x;
    float px = 0.f, py = 0.f, pz = 0.f;
    if (i < n) {
        px = (float)x[3 * i]; py = (float)x[3 * i + 1]; pz = (float)x[3 * i + 2];
    }
    float o[3];
    TrainBufs tb = {};
    mlp_wave<MLP_EVAL>(net, px, py, pz, tiles + (threadIdx.x >> 6) * TILE, o, tb, 0);
    if (i < n) {
        out[3 * i] = o[0]; out[3 * i + 1] = o[1]; out[3 * i + 2] = o[2];
    }
}

/* training forward over batch rows [0, rows): sample s0 + row of the set
 * (x, targets y); rows >= nvalid are padding (zero gradient).  Writes the
 * activations, dL/dout and one loss partial (sum of squared errors) per
 * block. */
__global__ void __launch_bounds__(256, 2) k_mlp_train_fwd(MlpNet net, const float *dsx,
        const float *dsy, long s0, int nvalid, float inv3n, TrainBufs tb, float *dout,
        double *lossp)
{
    __shared__ __attribute__((aligned(16))) __bf16 tiles[4 * TILE];
    __shared__ double wl[4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long row0 = (long)blockIdx.x * 256 + w * 64;
    const long row = row0 + lane;
    const bool valid = row < nvalid;
    float px = 0.f, py = 0.f, pz = 0.f;
    if (valid) {
        const long s = s0 + row;
        px = dsx[3 * s]; py = dsx[3 * s + 1]; pz = dsx[3 * s + 2];
    }
    float o[3];
    mlp_wave<MLP_TRAIN>(net, px, py, pz, tiles + w * TILE, o, tb, row0);
    double e2 = 0.0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float d = 0.f;
        if (valid) {
            float e = o[c] - dsy[3 * (s0 + row) + c];
            e2 += (double)e * (double)e;
            d = 2.0f * e * inv3n;
        }
        dout[3 * row + c] = d;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) e2 += __shfl_xor(e2, off, 64);
    if (lane == 0) wl[w] = e2;
    __syncthreads();
    if (threadIdx.x == 0) lossp[blockIdx.x] = ((wl[0] + wl[1]) + wl[2]) + wl[3];
}

/* backward through layers 3 and 2: dZ3 from dL/dout, W4 and gelu'(z3);
 * dZ2 = (dZ3 W3^T) gelu'(z2); dZ1 = (dZ2 W2^T) gelu'(z1).  All three are
 * stored row-major in bfloat16 for the weight gradients. */
__global__ void __launch_bounds__(256, 2) k_mlp_train_bwd(MlpNet net, TrainBufs tb,
        const float *dout, __bf16 *dz3, __bf16 *dz2, __bf16 *dz1)
{
    __shared__ __attribute__((aligned(16))) __bf16 tiles[4 * TILE];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, i16 = lane & 15;
    __bf16 *tile = tiles + w * TILE;
    const long row0 = (long)blockIdx.x * 256 + w * 64;
    f32x4 acc[8][4];
#pragma unroll
    for (int ib = 0; ib < 8; ib++)
#pragma unroll
        for (int mb = 0; mb < 4; mb++) acc[ib][mb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
    for (int kb = 0; kb < 4; kb++) {
        bf16x8 bz[4];
        const int n0 = kb * 32 + 8 * g;
        const f32x4 *wp = (const f32x4 *)(net.w4 + 3 * n0);
        f32x4 wv[6];
#pragma unroll
        for (int k = 0; k < 6; k++) wv[k] = wp[k];
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            const long r = row0 + mb * 16 + i16;
            const float d0 = dout[3 * r], d1 = dout[3 * r + 1], d2 = dout[3 * r + 2];
            const f32x4 ga = *(const f32x4 *)(tb.g3 + r * NH + n0);
            const f32x4 gb = *(const f32x4 *)(tb.g3 + r * NH + n0 + 4);
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const int e = 3 * t;
                const float w0 = wv[e / 4][e % 4], w1 = wv[(e + 1) / 4][(e + 1) % 4],
                            w2 = wv[(e + 2) / 4][(e + 2) % 4];
                float dh = fmaf(d2, w2, fmaf(d1, w1, d0 * w0));
                bz[mb][t] = (__bf16)(dh * (t < 4 ? ga[t] : gb[t - 4]));
            }
            *(bf16x8 *)(dz3 + r * NH + n0) = bz[mb];
        }
#pragma unroll
        for (int ib = 0; ib < 8; ib++) {
            const bf16x8 a = net.w3b[(ib * 4 + kb) * 64 + lane];
#pragma unroll
            for (int mb = 0; mb < 4; mb++) acc[ib][mb] = mfma(a, bz[mb], acc[ib][mb]);
        }
    }
    /* dZ2 */
#pragma unroll
    for (int ib = 0; ib < 8; ib++)
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            const int m = mb * 16 + i16;
            const int k0 = ib * 16 + 4 * g;
            const f32x4 gg = *(const f32x4 *)(tb.g2 + (row0 + m) * NH + k0);
            bf16x4 v;
#pragma unroll
            for (int r = 0; r < 4; r++) v[r] = (__bf16)(acc[ib][mb][r] * gg[r]);
            *(bf16x4 *)(tile + m * TS + k0) = v;
            *(bf16x4 *)(dz2 + (row0 + m) * NH + k0) = v;
        }
    /* dZ1 */
#pragma unroll
    for (int ib = 0; ib < 8; ib++)
#pragma unroll
        for (int mb = 0; mb < 4; mb++) acc[ib][mb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
    for (int kb = 0; kb < 4; kb++) {
        bf16x8 bx[4];
#pragma unroll
        for (int mb = 0; mb < 4; mb++)
            bx[mb] = *(const bf16x8 *)(tile + (mb * 16 + i16) * TS + kb * 32 + 8 * g);
#pragma unroll
        for (int ib = 0; ib < 8; ib++) {
            const bf16x8 a = net.w2b[(ib * 4 + kb) * 64 + lane];
#pragma unroll
            for (int mb = 0; mb < 4; mb++) acc[ib][mb] = mfma(a, bx[mb], acc[ib][mb]);
        }
    }
#pragma unroll
    for (int ib = 0; ib < 8; ib++)
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            const int m = mb * 16 + i16;
            const int k0 = ib * 16 + 4 * g;
            const f32x4 gg = *(const f32x4 *)(tb.g1 + (row0 + m) * NH + k0);
            bf16x4 v;
#pragma unroll
            for (int r = 0; r < 4; r++) v[r] = (__bf16)(acc[ib][mb][r] * gg[r]);
            *(bf16x4 *)(dz1 + (row0 + m) * NH + k0) = v;
        }
}

/* LDS tiles of the weight-gradient kernel: 32 samples, transposed
 * (row = feature, 40 bf16 per row: 80 bytes, 16-byte aligned reads) */
#define XS 40

/* dW = X^T D over samples [row0, row0 + 256): X [rows][KIN], D [rows][128]
 * bf16; writes pw[k * 128 + n] and the bias gradient pb[n] = sum_s D[s][n] */
template <int KIN>
__device__ __forceinline__ void wgrad_layer(const __bf16 *X, const __bf16 *D, long row0,
                                            float *pw, float *pb, __bf16 *xt, __bf16 *dt)
{
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, g = lane >> 4, i16 = lane & 15;
    constexpr int IBW = KIN / 64;
    constexpr int KC = KIN / 8;
    f32x4 acc[IBW][8];
#pragma unroll
    for (int ii = 0; ii < IBW; ii++)
#pragma unroll
        for (int nb = 0; nb < 8; nb++) acc[ii][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;
    for (int s0 = 0; s0 < 256; s0 += 32) {
        for (int c = tid; c < 32 * KC; c += 256) {
            const int s = c / KC, k8 = c % KC;
            const bf16x8 v = *(const bf16x8 *)(X + (row0 + s0 + s) * KIN + k8 * 8);
#pragma unroll
            for (int t = 0; t < 8; t++) xt[(k8 * 8 + t) * XS + s] = v[t];
        }
        for (int c = tid; c < 32 * 16; c += 256) {
            const int s = c / 16, n8 = c % 16;
            const bf16x8 v = *(const bf16x8 *)(D + (row0 + s0 + s) * NH + n8 * 8);
#pragma unroll
            for (int t = 0; t < 8; t++) dt[(n8 * 8 + t) * XS + s] = v[t];
        }
        __syncthreads();
        bf16x8 bfr[8];
#pragma unroll
        for (int nb = 0; nb < 8; nb++)
            bfr[nb] = *(const bf16x8 *)(dt + (nb * 16 + i16) * XS + 8 * g);
#pragma unroll
        for (int ii = 0; ii < IBW; ii++) {
            const int ib = w * IBW + ii;
            const bf16x8 a = *(const bf16x8 *)(xt + (ib * 16 + i16) * XS + 8 * g);
#pragma unroll
            for (int nb = 0; nb < 8; nb++) acc[ii][nb] = mfma(a, bfr[nb], acc[ii][nb]);
        }
        if (tid < NH)
            for (int s = 0; s < 32; s++) bsum += (float)dt[tid * XS + s];
        __syncthreads();
    }
#pragma unroll
    for (int ii = 0; ii < IBW; ii++)
#pragma unroll
        for (int nb = 0; nb < 8; nb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int k = (w * IBW + ii) * 16 + 4 * g + r;
                const int n = nb * 16 + i16;
                pw[k * NH + n] = acc[ii][nb][r];
            }
    if (tid < NH) pb[tid] = bsum;
}

/* per-chunk parameter gradients: blockIdx.x = chunk of 256 samples,
 * blockIdx.y = layer (0: W1 b1, 1: W2 b2, 2: W3 b3, 3: W4 b4).  part has
 * one row of P_TOTAL floats per chunk. */
__global__ void __launch_bounds__(256) k_mlp_wgrad(TrainBufs tb, const __bf16 *dz1,
        const __bf16 *dz2, const __bf16 *dz3, const float *dout, float *part)
{
    __shared__ __attribute__((aligned(16))) __bf16 xt[NIN * XS];
    __shared__ __attribute__((aligned(16))) __bf16 dt[NH * XS];
    const long row0 = (long)blockIdx.x * 256;
    float *pr = part + (size_t)blockIdx.x * P_TOTAL;
    const int layer = blockIdx.y;
    if (layer == 0) {
        wgrad_layer<NIN>(tb.h0, dz1, row0, pr + P_W1, pr + P_B1, xt, dt);
    } else if (layer == 1) {
        wgrad_layer<NH>(tb.h1, dz2, row0, pr + P_W2, pr + P_B2, xt, dt);
    } else if (layer == 2) {
        wgrad_layer<NH>(tb.h2, dz3, row0, pr + P_W3, pr + P_B3, xt, dt);
    } else {
        /* output layer, FP32: dW4[n][c] = sum_s h3[s][n] dout[s][c] */
        const int tid = threadIdx.x, n = tid & 127, half = tid >> 7;
        float a0 = 0.f, a1 = 0.f;
        for (int s = 0; s < 256; s++) {
            const long r = row0 + s;
            const float h = tb.h3[r * NH + n];
            if (half == 0) {
                a0 = fmaf(h, dout[3 * r], a0);
                a1 = fmaf(h, dout[3 * r + 1], a1);
            } else {
                a0 = fmaf(h, dout[3 * r + 2], a0);
            }
        }
        if (half == 0) {
            pr[P_W4 + 3 * n] = a0;
            pr[P_W4 + 3 * n + 1] = a1;
        } else {
            pr[P_W4 + 3 * n + 2] = a0;
        }
        if (tid < 3) {
            float b = 0.f;
            for (int s = 0; s < 256; s++) b += dout[3 * (row0 + s) + tid];
            pr[P_B4 + tid] = b;
        }
    }
}

enum { ADAM_STEP = 0, ADAM_GRAD_ONLY = 1, ADAM_FRAG_ONLY = 2 };

/* sum of the chunk partials in chunk order (deterministic), then the
 * mode's action: an Adam step (Kingma & Ba 2015, bias-corrected), the
 * gradient only, or nothing; finally the bf16 fragments of the weight. */
__global__ void __launch_bounds__(256) k_mlp_adam(float *params, float *am, float *av,
        const float *part, int nchunks, int mode, float lr, float b1, float b2, float eps,
        float bc1, float bc2, float *gout, bf16x8 *frag)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= P_TOTAL) return;
    float w = params[p];
    if (mode != ADAM_FRAG_ONLY) {
        float gs = 0.f;
        for (int c = 0; c < nchunks; c++) gs += part[(size_t)c * P_TOTAL + p];
        if (mode == ADAM_GRAD_ONLY) {
            gout[p] = gs;
            return;
        }
        const float m = fmaf(b1, am[p], (1.0f - b1) * gs);
        const float v = fmaf(b2, av[p], (1.0f - b2) * gs * gs);
        am[p] = m;
        av[p] = v;
        w -= lr * (m / bc1) / (sqrtf(v / bc2) + eps);
        params[p] = w;
    }
    __bf16 *fr = (__bf16 *)frag;
    const __bf16 wb = (__bf16)w;
    if (p < P_B1) {
        fr[FR_W1F * 8 + frag_fwd(p / NH, p % NH, NIN / 32)] = wb;
    } else if (p >= P_W2 && p < P_B2) {
        const int e = p - P_W2;
        fr[FR_W2F * 8 + frag_fwd(e / NH, e % NH, NH / 32)] = wb;
        fr[FR_W2B * 8 + frag_bwd(e / NH, e % NH)] = wb;
    } else if (p >= P_W3 && p < P_B3) {
        const int e = p - P_W3;
        fr[FR_W3F * 8 + frag_fwd(e / NH, e % NH, NH / 32)] = wb;
        fr[FR_W3B * 8 + frag_bwd(e / NH, e % NH)] = wb;
    }
}

/* sum of the loss partials in block order, added to *acc */
__global__ void k_mlp_loss_acc(const double *lossp, int nblocks, double *acc)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < nblocks; b++) s += lossp[b];
        *acc += s;
    }
}

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

/* training set: x uniform in [0, L)^3 (splitmix64 of seed + index times the
 * golden-ratio increment, 53-bit uniforms, rounded to float), targets the
 * exact FP64 mode sum at the float position, without the guide field,
 * divided by the output scale */
__global__ void __launch_bounds__(256) k_mlp_sample(FieldDev f, long n, double L,
        unsigned long long seed, double inv_scale, float *dsx, float *dsy)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float xs[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const unsigned long long r =
            splitmix64(seed + (unsigned long long)(3 * i + c) * 0x9e3779b97f4a7c15ull);
        xs[c] = (float)((double)(r >> 11) * 0x1.0p-53 * L);
    }
    f.b0z = 0.0;
    Emf e;
    get_emf64<FK_MODES>(f, (double)xs[0], (double)xs[1], (double)xs[2], 0.0, e);
    dsx[3 * i] = xs[0]; dsx[3 * i + 1] = xs[1]; dsx[3 * i + 2] = xs[2];
    dsy[3 * i] = (float)(e.bx * inv_scale);
    dsy[3 * i + 1] = (float)(e.by * inv_scale);
    dsy[3 * i + 2] = (float)(e.bz * inv_scale);
}

/* ------------------------------------------------------------------ */
/* step kernel with the surrogate field                                */
/* ------------------------------------------------------------------ */
/* Residual check of every particle that still uses the network: network
 * and exact FP32 mode sum at its position; a relative residual above tol
 * switches it to the exact field for the rest of the run.  first: the
 * run's first launch (all flags cleared, all particles checked). */
__global__ void __launch_bounds__(256, 2) k_mlp_check(TrackArgs a, int first)
{
    __shared__ __attribute__((aligned(16))) __bf16 tiles[4 * TILE];
    const MlpNet &net = a.f.mlp;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < a.n;
    int fb = 1;
    bool moving = false;
    float px = 0.f, py = 0.f, pz = 0.f;
    if (live) {
        fb = first ? 0 : net.fb[i];
        moving = first || a.esc[i] == 0 || a.bc == 0;
        px = (float)a.x[i]; py = (float)a.y[i]; pz = (float)a.z[i];
    }
    const bool chk = live && moving && fb == 0;
    float o[3] = {0.f, 0.f, 0.f};
    TrainBufs tb = {};
    if (__any(chk))
        mlp_wave<MLP_EVAL>(net, px, py, pz, tiles + (threadIdx.x >> 6) * TILE, o, tb, 0);
    if (chk) {
        const float b0z = (float)a.f.b0z;
        float bx, by, bz;
        modes_sum32(a.f.modes32, a.f.nmodes, px, py, pz, bx, by, bz);
        bz += b0z;
        const float dx = net.scale * o[0] - bx, dy = net.scale * o[1] - by;
        const float dz = fmaf(net.scale, o[2], b0z) - bz;
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        const float bb = fmaf(bx, bx, fmaf(by, by, bz * bz));
        if (!(r2 <= net.tol * net.tol * bb)) {
            fb = 1;
            atomicAdd(net.nfb, 1u);
        }
    }
    if (live) net.fb[i] = fb;
}

/* Splits the particles into the network list (fb == 0) and the exact-field
 * list (fb == 1): ballot prefix sums within a wave, wave offsets within the
 * block, one atomic per block and list for the block's base (the order of
 * blocks within a list is arbitrary; every particle appears exactly once). */
__global__ void __launch_bounds__(256) k_mlp_split(int n, const int *fb, int *idx_net,
                                                   int *idx_ex, unsigned *cnt)
{
    __shared__ unsigned wn[4], we[4], base[2];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool live = i < n;
    const int f = live ? fb[i] : 0;
    const bool isn = live && f == 0, ise = live && f != 0;
    const unsigned long long bn = __ballot(isn), be = __ballot(ise);
    const unsigned long long below = (1ull << lane) - 1ull;
    const unsigned pn = (unsigned)__popcll(bn & below), pe = (unsigned)__popcll(be & below);
    if (lane == 0) {
        wn[w] = (unsigned)__popcll(bn);
        we[w] = (unsigned)__popcll(be);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        base[0] = atomicAdd(&cnt[0], wn[0] + wn[1] + wn[2] + wn[3]);
        base[1] = atomicAdd(&cnt[1], we[0] + we[1] + we[2] + we[3]);
    }
    __syncthreads();
    unsigned on = base[0], oe = base[1];
    for (int k = 0; k < w; k++) {
        on += wn[k];
        oe += we[k];
    }
    if (isn) idx_net[on + pn] = i;
    if (ise) idx_ex[oe + pe] = i;
}

/* The FP32 Boris step of k_track (tp_gpu.hip) for the particles of the
 * network list, the field from the network: every wave evaluates it at its
 * 64 particles each step.  Particles that fell back run in the exact-field
 * kernel instead, so no wave mixes the two. */
__global__ void __launch_bounds__(256, 2) k_track_mlp(TrackArgs a)
{
    extern __shared__ unsigned hist[];
    __shared__ __attribute__((aligned(16))) __bf16 tiles[4 * TILE];
    const int tid = threadIdx.x;
    const int gi = blockIdx.x * blockDim.x + tid;
    const bool live = gi < a.n;
    const int i = live ? a.idx[gi] : 0;
    const MlpNet &net = a.f.mlp;
    __bf16 *tile = tiles + (tid >> 6) * TILE;
    for (int b = tid; b < a.nbins; b += blockDim.x) hist[b] = 0u;

    Ptl p;
    double xs = 0.0, ys = 0.0, zs = 0.0;
    int esc = 1, nst = 0;
    if (live) {
        p.x = a.x[i]; p.y = a.y[i]; p.z = a.z[i];
        p.ux = a.ux[i]; p.uy = a.uy[i]; p.uz = a.uz[i]; p.t = a.t[i];
        xs = a.xs[i]; ys = a.ys[i]; zs = a.zs[i];
        esc = a.esc[i];
        nst = a.nsteps[i];
        if (a.first) {
            esc = 0;
            nst = 1;
            particle_bc(a, p, xs, ys, zs, esc);
            if (esc && 1 < a.nd) atomicAdd(&a.rdec[1], 1);
        }
    } else {
        p.x = p.y = p.z = p.ux = p.uy = p.uz = p.t = 0.0;
    }
    float fux = (float)p.ux, fuy = (float)p.uy, fuz = (float)p.uz;
    const float dtf = (float)a.dt;
    const float hf = (float)(a.qm * (0.5 * a.dt));
    const float b0z = (float)a.f.b0z;
    TrainBufs tb = {};
    __syncthreads();

    int ntp = a.ntp0;
    for (int j = a.j0; j <= a.j1; j++) {
        const bool moving = live && (esc == 0 || a.bc == 0);
        float o[3] = {0.f, 0.f, 0.f};
        if (__any(moving))
            mlp_wave<MLP_EVAL>(net, (float)p.x, (float)p.y, (float)p.z, tile, o, tb, 0);
        EmfF e;
        e.ex = 0.f; e.ey = 0.f; e.ez = 0.f;
        e.bx = net.scale * o[0];
        e.by = net.scale * o[1];
        e.bz = fmaf(net.scale, o[2], b0z);
        int esc_after = esc;
        if (moving) {
            boris_push32(e, p, fux, fuy, fuz, a.dt, dtf, hf);
            p.ux = fux; p.uy = fuy; p.uz = fuz;
            particle_bc(a, p, xs, ys, zs, esc_after);
            nst++;
        }
        if (live && esc == 0 && esc_after == 1 && a.bc == 1 && ntp < a.nt_out)
            atomicAdd(&a.espect_esc[ntp * a.nbins +
                      energy_bin(p.ux, p.uy, p.uz, a.nbins, a.logemin, a.logde)], 1u);
        if (j % a.einterval == 0) {
            if (live && esc_after == 0 && ntp < a.nt_out)
                atomicAdd(&hist[energy_bin(p.ux, p.uy, p.uz, a.nbins, a.logemin, a.logde)], 1u);
            __syncthreads();
            if (ntp < a.nt_out) {
                for (int b = tid; b < a.nbins; b += blockDim.x) {
                    unsigned h = hist[b];
                    if (h) {
                        atomicAdd(&a.espect[ntp * a.nbins + b], h);
                        hist[b] = 0u;
                    }
                }
            }
            __syncthreads();
            ntp++;
        }
        esc = esc_after;
    }
    if (live) {
        a.x[i] = p.x; a.y[i] = p.y; a.z[i] = p.z;
        a.ux[i] = p.ux; a.uy[i] = p.uy; a.uz[i] = p.uz; a.t[i] = p.t;
        a.xs[i] = xs; a.ys[i] = ys; a.zs[i] = zs;
        a.esc[i] = esc;
        a.nsteps[i] = nst;
    }
}

/* ------------------------------------------------------------------ */
/* host side                                                           */
/* ------------------------------------------------------------------ */
struct MlpState {
    float *params = nullptr, *am = nullptr, *av = nullptr, *q = nullptr;
    bf16x8 *frag = nullptr;
    float scale = 1.0f;
    bool set = false;
    double tol = 1e-3;
    int check = 0;
    int *fb = nullptr;
    unsigned *nfb = nullptr;
    int fb_cap = 0;
    int *idx = nullptr;          /* network list, then exact-field list (2 x fb_cap) */
    unsigned *cnt = nullptr;     /* list lengths */
    int n_net = 0, n_ex = 0;
    /* batch buffers */
    long rows_cap = 0;
    __bf16 *h0 = nullptr, *h1 = nullptr, *h2 = nullptr;
    __bf16 *dz1 = nullptr, *dz2 = nullptr, *dz3 = nullptr;
    float *g1 = nullptr, *g2 = nullptr, *g3 = nullptr, *h3 = nullptr, *dout = nullptr;
    double *lossp = nullptr, *lacc = nullptr;
    float *part = nullptr, *grad = nullptr;
    /* training set (or the points of a gradient call) */
    long ds_cap = 0;
    float *dsx = nullptr, *dsy = nullptr;
};

void tp_mlp_release(tp_ctx *c)
{
    MlpState *m = c->mlp;
    if (!m) return;
    void *ptrs[] = {m->params, m->am, m->av, m->q, m->frag, m->fb, m->nfb, m->idx, m->cnt,
                    m->h0, m->h1,
                    m->h2, m->dz1, m->dz2, m->dz3, m->g1, m->g2, m->g3, m->h3, m->dout,
                    m->lossp, m->lacc, m->part, m->grad, m->dsx, m->dsy};
    for (void *p : ptrs) (void)hipFree(p);
    delete m;
    c->mlp = nullptr;
}

static MlpNet net_of(const MlpState *m)
{
    MlpNet n;
    memset(&n, 0, sizeof(n));
    n.w1f = m->frag + FR_W1F;
    n.w2f = m->frag + FR_W2F;
    n.w3f = m->frag + FR_W3F;
    n.w2b = m->frag + FR_W2B;
    n.w3b = m->frag + FR_W3B;
    n.b1 = m->params + P_B1;
    n.b2 = m->params + P_B2;
    n.b3 = m->params + P_B3;
    n.w4 = m->params + P_W4;
    n.b4 = m->params + P_B4;
    n.q = m->q;
    n.scale = m->scale;
    n.tol = (float)m->tol;
    n.check = m->check;
    n.fb = m->fb;
    n.nfb = m->nfb;
    return n;
}

static int dalloc(void **p, size_t bytes)
{
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        return TP_E_NOMEM;
    }
    return TP_OK;
}

static int ensure_rows(MlpState *m, long rows)
{
    if (rows <= m->rows_cap) return TP_OK;
    void *old[] = {m->h0, m->h1, m->h2, m->dz1, m->dz2, m->dz3, m->g1, m->g2, m->g3,
                   m->h3, m->dout, m->lossp, m->part};
    for (void *p : old) (void)hipFree(p);
    m->rows_cap = 0;
    const size_t r = (size_t)rows;
    int rc = TP_OK;
    rc |= dalloc((void **)&m->h0, r * NIN * 2);
    rc |= dalloc((void **)&m->h1, r * NH * 2);
    rc |= dalloc((void **)&m->h2, r * NH * 2);
    rc |= dalloc((void **)&m->dz1, r * NH * 2);
    rc |= dalloc((void **)&m->dz2, r * NH * 2);
    rc |= dalloc((void **)&m->dz3, r * NH * 2);
    rc |= dalloc((void **)&m->g1, r * NH * 4);
    rc |= dalloc((void **)&m->g2, r * NH * 4);
    rc |= dalloc((void **)&m->g3, r * NH * 4);
    rc |= dalloc((void **)&m->h3, r * NH * 4);
    rc |= dalloc((void **)&m->dout, r * 3 * 4);
    rc |= dalloc((void **)&m->lossp, (r / 256) * sizeof(double));
    rc |= dalloc((void **)&m->part, (r / 256) * (size_t)P_TOTAL * 4);
    if (rc) return TP_E_NOMEM;
    m->rows_cap = rows;
    return TP_OK;
}

static int ensure_ds(MlpState *m, long n)
{
    if (n <= m->ds_cap) return TP_OK;
    (void)hipFree(m->dsx); (void)hipFree(m->dsy);
    m->ds_cap = 0;
    if (dalloc((void **)&m->dsx, (size_t)n * 12) || dalloc((void **)&m->dsy, (size_t)n * 12))
        return TP_E_NOMEM;
    m->ds_cap = n;
    return TP_OK;
}

static TrainBufs bufs_of(const MlpState *m)
{
    TrainBufs t;
    t.h0 = m->h0; t.h1 = m->h1; t.h2 = m->h2;
    t.g1 = m->g1; t.g2 = m->g2; t.g3 = m->g3; t.h3 = m->h3;
    return t;
}

/* forward + backward + chunk gradients over rows [0, rows) of batch
 * samples s0.. of the set; then the optimiser kernel in the given mode */
static int train_pass(tp_ctx *c, long s0, int nvalid, long rows, int mode, float lr, float b1,
                      float b2, float eps, float bc1, float bc2, float *gout)
{
    MlpState *m = c->mlp;
    const MlpNet net = net_of(m);
    const TrainBufs tb = bufs_of(m);
    const int nblk = (int)(rows / 256);
    const float inv3n = (float)(1.0 / (3.0 * (double)nvalid));
    hipLaunchKernelGGL(k_mlp_train_fwd, dim3(nblk), dim3(256), 0, c->stream, net, m->dsx,
                       m->dsy, s0, nvalid, inv3n, tb, m->dout, m->lossp);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_train_bwd, dim3(nblk), dim3(256), 0, c->stream, net, tb,
                       (const float *)m->dout, m->dz3, m->dz2, m->dz1);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(nblk, 4), dim3(256), 0, c->stream, tb,
                       (const __bf16 *)m->dz1, (const __bf16 *)m->dz2, (const __bf16 *)m->dz3,
                       (const float *)m->dout, m->part);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_adam, dim3((P_TOTAL + 255) / 256), dim3(256), 0, c->stream,
                       m->params, m->am, m->av, (const float *)m->part, nblk, mode, lr, b1, b2,
                       eps, bc1, bc2, gout, m->frag);
    HIPCHK(hipGetLastError());
    return TP_OK;
}

extern "C" {

int tp_gpu_mlp_set(tp_ctx *c, const float *q, const float *params, double scale)
{
    if (!c || !q || !params || !(scale > 0.0) || !isfinite(scale)) return TP_E_ARG;
    for (int k = 0; k < 3 * NF; k++)
        if (!isfinite(q[k])) return TP_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    if (!c->mlp) {
        c->mlp = new (std::nothrow) MlpState();
        if (!c->mlp) return TP_E_NOMEM;
    }
    MlpState *m = c->mlp;
    if (!m->params) {
        int rc = TP_OK;
        rc |= dalloc((void **)&m->params, P_TOTAL * 4);
        rc |= dalloc((void **)&m->am, P_TOTAL * 4);
        rc |= dalloc((void **)&m->av, P_TOTAL * 4);
        rc |= dalloc((void **)&m->q, 3 * NF * 4);
        rc |= dalloc((void **)&m->frag, FR_TOTAL * sizeof(bf16x8));
        rc |= dalloc((void **)&m->nfb, sizeof(unsigned));
        rc |= dalloc((void **)&m->cnt, 2 * sizeof(unsigned));
        rc |= dalloc((void **)&m->lacc, sizeof(double));
        rc |= dalloc((void **)&m->grad, P_TOTAL * 4);
        if (rc) {
            tp_mlp_release(c);
            return TP_E_NOMEM;
        }
    }
    HIPCHK(hipMemcpy(m->params, params, P_TOTAL * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m->q, q, 3 * NF * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(m->am, 0, P_TOTAL * 4));
    HIPCHK(hipMemset(m->av, 0, P_TOTAL * 4));
    HIPCHK(hipMemset(m->frag, 0, FR_TOTAL * sizeof(bf16x8)));
    m->scale = (float)scale;
    hipLaunchKernelGGL(k_mlp_adam, dim3((P_TOTAL + 255) / 256), dim3(256), 0, c->stream,
                       m->params, m->am, m->av, (const float *)nullptr, 0, (int)ADAM_FRAG_ONLY,
                       0.f, 0.f, 0.f, 0.f, 1.f, 1.f, (float *)nullptr, m->frag);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    m->set = true;
    if (c->f.kind == FK_MLP) c->f.mlp = net_of(m);
    return TP_OK;
}

int tp_gpu_mlp_get(tp_ctx *c, float *params)
{
    if (!c || !params) return TP_E_ARG;
    if (!c->mlp || !c->mlp->set) return TP_E_STATE;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(params, c->mlp->params, P_TOTAL * 4, hipMemcpyDeviceToHost));
    return TP_OK;
}

int tp_gpu_mlp_forward(tp_ctx *c, int n, const double *x, double *out)
{
    if (!c || n < 0 || (n > 0 && (!x || !out))) return TP_E_ARG;
    if (!c->mlp || !c->mlp->set) return TP_E_STATE;
    if (n == 0) return TP_OK;
    HIPCHK(hipSetDevice(c->device));
    double *dx = nullptr, *dout = nullptr;
    if (dalloc((void **)&dx, (size_t)n * 24) || dalloc((void **)&dout, (size_t)n * 24)) {
        (void)hipFree(dx);
        return TP_E_NOMEM;
    }
    int rc = TP_OK;
    const MlpNet net = net_of(c->mlp);
    if (hipMemcpyAsync(dx, x, (size_t)n * 24, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = TP_E_HIP;
    if (!rc) {
        hipLaunchKernelGGL(k_mlp_eval, dim3((n + 255) / 256), dim3(256), 0, c->stream, net, n,
                           (const double *)dx, dout);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(out, dout, (size_t)n * 24, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = TP_E_HIP;
    }
    (void)hipFree(dx);
    (void)hipFree(dout);
    return rc;
}

int tp_gpu_mlp_loss_grad(tp_ctx *c, int n, const float *x, const float *y, float *grad,
                         double *loss)
{
    if (!c || n <= 0 || !x || !y || !grad || !loss) return TP_E_ARG;
    if (!c->mlp || !c->mlp->set) return TP_E_STATE;
    HIPCHK(hipSetDevice(c->device));
    MlpState *m = c->mlp;
    const long rows = ((long)n + 255) / 256 * 256;
    int rc = ensure_rows(m, rows);
    if (!rc) rc = ensure_ds(m, rows);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(m->dsx, x, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(m->dsy, y, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
    rc = train_pass(c, 0, n, rows, ADAM_GRAD_ONLY, 0.f, 0.f, 0.f, 0.f, 1.f, 1.f, m->grad);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(m->lacc, 0, sizeof(double), c->stream));
    hipLaunchKernelGGL(k_mlp_loss_acc, dim3(1), dim3(64), 0, c->stream,
                       (const double *)m->lossp, (int)(rows / 256), m->lacc);
    HIPCHK(hipGetLastError());
    double l = 0.0;
    HIPCHK(hipMemcpyAsync(grad, m->grad, P_TOTAL * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&l, m->lacc, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *loss = l / (3.0 * (double)n);
    return TP_OK;
}

int tp_gpu_train_surrogate(tp_ctx *c, const tp_mlp_train_params *p, tp_mlp_train_stats *st)
{
    if (!c || !p) return TP_E_ARG;
    if (!c->mlp || !c->mlp->set) return TP_E_STATE;
    if (!c->have_field || (c->f.kind != FK_MODES && c->f.kind != FK_MLP) || c->f.nmodes <= 0)
        return TP_E_STATE;
    if (p->batch <= 0 || p->batch % 256 || p->nsamples < p->batch ||
        p->nsamples % p->batch || p->epochs < 1 || !(p->period > 0.0) || !(p->lr > 0.0) ||
        p->lr_min < 0.0 || !(p->beta1 >= 0.0 && p->beta1 < 1.0) ||
        !(p->beta2 >= 0.0 && p->beta2 < 1.0) || !(p->eps > 0.0))
        return TP_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    MlpState *m = c->mlp;
    int rc = ensure_rows(m, p->batch);
    if (!rc) rc = ensure_ds(m, p->nsamples);
    if (rc) return rc;
    hipEvent_t e0 = c->ev0, e1 = c->ev1;
    HIPCHK(hipEventRecord(e0, c->stream));
    const long ns = (long)p->nsamples;
    hipLaunchKernelGGL(k_mlp_sample, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0,
                       c->stream, c->f, ns, p->period, p->seed, 1.0 / (double)m->scale,
                       m->dsx, m->dsy);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e1, c->stream));
    HIPCHK(hipEventSynchronize(e1));
    float sample_ms = 0.f;
    HIPCHK(hipEventElapsedTime(&sample_ms, e0, e1));

    const long spe = ns / p->batch;
    const long total = spe * p->epochs;
    const int nblk = p->batch / 256;
    double first = 0.0;
    HIPCHK(hipEventRecord(e0, c->stream));
    double pb1 = 1.0, pb2 = 1.0;
    for (long t = 0; t < total; t++) {
        const double lr = p->lr_min + 0.5 * (p->lr - p->lr_min) *
                                          (1.0 + cos(TP_PI * (double)t / (double)total));
        pb1 *= p->beta1;
        pb2 *= p->beta2;
        const long s0 = (t % spe) * p->batch;
        rc = train_pass(c, s0, p->batch, p->batch, ADAM_STEP, (float)lr, (float)p->beta1,
                        (float)p->beta2, (float)p->eps, (float)(1.0 - pb1), (float)(1.0 - pb2),
                        nullptr);
        if (rc) return rc;
        if (t == 0) {
            HIPCHK(hipMemsetAsync(m->lacc, 0, sizeof(double), c->stream));
            hipLaunchKernelGGL(k_mlp_loss_acc, dim3(1), dim3(64), 0, c->stream,
                               (const double *)m->lossp, nblk, m->lacc);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(&first, m->lacc, sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            first /= 3.0 * (double)p->batch;
        }
        /* the last epoch's mean loss */
        if (t == total - spe) HIPCHK(hipMemsetAsync(m->lacc, 0, sizeof(double), c->stream));
        if (t >= total - spe) {
            hipLaunchKernelGGL(k_mlp_loss_acc, dim3(1), dim3(64), 0, c->stream,
                               (const double *)m->lossp, nblk, m->lacc);
            HIPCHK(hipGetLastError());
        }
    }
    HIPCHK(hipEventRecord(e1, c->stream));
    double last = 0.0;
    HIPCHK(hipMemcpyAsync(&last, m->lacc, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float train_ms = 0.f;
    HIPCHK(hipEventElapsedTime(&train_ms, e0, e1));
    if (st) {
        st->first_loss = first;
        st->last_loss = last / (3.0 * (double)p->batch * (double)spe);
        st->steps = total;
        st->sample_ms = sample_ms;
        st->train_ms = train_ms;
    }
    if (c->f.kind == FK_MLP) c->f.mlp = net_of(m);
    return TP_OK;
}

int tp_gpu_set_field_mlp(tp_ctx *c, double tol, int check_interval)
{
    if (!c || !(tol >= 0.0) || check_interval < 0) return TP_E_ARG;
    if (!c->mlp || !c->mlp->set) return TP_E_STATE;
    if (!c->have_field || (c->f.kind != FK_MODES && c->f.kind != FK_MLP) || !c->f.modes32)
        return TP_E_STATE;
    c->mlp->tol = tol;
    c->mlp->check = check_interval;
    c->f.kind = FK_MLP;
    c->f.mlp = net_of(c->mlp);
    return TP_OK;
}

int tp_gpu_mlp_fallback_count(tp_ctx *c, long long *count)
{
    if (!c || !count) return TP_E_ARG;
    if (!c->mlp || !c->mlp->nfb) return TP_E_STATE;
    HIPCHK(hipSetDevice(c->device));
    unsigned v = 0;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(&v, c->mlp->nfb, sizeof(unsigned), hipMemcpyDeviceToHost));
    *count = (long long)v;
    return TP_OK;
}

int tp_gpu_mlp_fallback_flags(tp_ctx *c, int n, int *flags)
{
    if (!c || n < 0 || (n > 0 && !flags)) return TP_E_ARG;
    if (!c->mlp || !c->mlp->fb) return TP_E_STATE;
    if (n != c->s.n || n > c->mlp->fb_cap) return TP_E_ARG;
    if (n == 0) return TP_OK;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(flags, c->mlp->fb, sizeof(int) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) flags[i] = flags[i] != 0;
    return TP_OK;
}

}  /* extern "C" */

/* run_fixed's hooks for the surrogate field: per-particle flags and index
 * lists for n particles, the fallback counter cleared, the network
 * pointers refreshed */
int tp_mlp_prepare_run(tp_ctx *c, int n)
{
    MlpState *m = c->mlp;
    if (!m || !m->set) return TP_E_STATE;
    if (n > m->fb_cap) {
        (void)hipFree(m->fb);
        (void)hipFree(m->idx);
        m->fb = nullptr;
        m->idx = nullptr;
        m->fb_cap = 0;
        const size_t k = (size_t)(n > 0 ? n : 1);
        if (dalloc((void **)&m->fb, k * sizeof(int)) ||
            dalloc((void **)&m->idx, 2 * k * sizeof(int)))
            return TP_E_NOMEM;
        m->fb_cap = n;
    }
    HIPCHK(hipMemsetAsync(m->nfb, 0, sizeof(unsigned), c->stream));
    m->n_net = m->n_ex = 0;
    c->f.mlp = net_of(m);
    return TP_OK;
}

/* one launch of the surrogate-field run: optional residual check and
 * split (the list lengths are read back: one host synchronisation per
 * check), then the network kernel over the network list and the
 * exact-field kernel over the other */
int tp_mlp_launch(tp_ctx *c, const TrackArgs &a, size_t shmem, int check, int block,
                  track_fn exact_fn)
{
    MlpState *m = c->mlp;
    const int n = a.n;
    if (check) {
        const dim3 g((n + 255) / 256);
        hipLaunchKernelGGL(k_mlp_check, g, dim3(256), 0, c->stream, a, a.first);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemsetAsync(m->cnt, 0, 2 * sizeof(unsigned), c->stream));
        hipLaunchKernelGGL(k_mlp_split, g, dim3(256), 0, c->stream, n, (const int *)m->fb,
                           m->idx, m->idx + m->fb_cap, m->cnt);
        HIPCHK(hipGetLastError());
        unsigned h[2] = {0u, 0u};
        HIPCHK(hipMemcpyAsync(h, m->cnt, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if ((long)h[0] + (long)h[1] != (long)n) return TP_E_HIP;
        m->n_net = (int)h[0];
        m->n_ex = (int)h[1];
    }
    TrackArgs b = a;
    if (m->n_net > 0) {
        b.idx = m->idx;
        b.n = m->n_net;
        hipLaunchKernelGGL(k_track_mlp, dim3((m->n_net + 255) / 256), dim3(256), shmem,
                           c->stream, b);
        HIPCHK(hipGetLastError());
    }
    if (m->n_ex > 0) {
        b.idx = m->idx + m->fb_cap;
        b.n = m->n_ex;
        exact_fn(dim3((m->n_ex + block - 1) / block), dim3(block), shmem, c->stream, b);
        HIPCHK(hipGetLastError());
    }
    return TP_OK;
}
