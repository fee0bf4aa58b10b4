/*
 * mlpref.c -- CPU oracle of the learned field surrogate (TEST INFRASTRUCTURE
 * ONLY: loaded by tests/ as the checker of the GPU training and inference
 * kernels; never linked into the product library).
 *
 * The surrogate has no counterpart in the reference.  It stands in for the
 * multi-mode turbulent field sum: a coordinate network with Fourier-feature
 * inputs,
 *
 *   u_j  = q_j . x                   (q_j = n_j / L, n_j integer: periodic in L)
 *   r_j  = u_j - rint(u_j)           (revolutions, in [-1/2, 1/2])
 *   f    = [sin 2 pi r_0 .. sin 2 pi r_127, cos 2 pi r_0 .. cos 2 pi r_127]
 *   h1   = gelu(f  W1 + b1)          256 -> 128
 *   h2   = gelu(h1 W2 + b2)          128 -> 128
 *   h3   = gelu(h2 W3 + b3)          128 -> 128
 *   out  = h3 W4 + b4                128 -> 3
 *   B    = b0z z-hat + scale * out
 *
 * gelu is the tanh form of Hendrycks & Gimpel (2016),
 *   gelu(z) = z / 2 (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)))
 *           = z sigma(2 sqrt(2/pi) (z + 0.044715 z^3)),
 * and the loss is the mean over samples and components of (out - y)^2.
 *
 * Parameters are one flat array (PARAM_* offsets below), row index = input.
 *
 * Two modes:
 *   emulate = 0: everything in FP64 (the differentiable reference that the
 *                finite-difference tests pin);
 *   emulate = 1: the GPU kernels' rounding points -- positions and phases in
 *                float, features, h1, h2, W1..W3 and the back-propagated
 *                dz1..dz3 rounded to bfloat16 (round to nearest even), h3 and
 *                layer 4 in float -- with FP64 accumulation in between, and
 *                the kernels' activation (gelu_gpu: the FP16 polynomial form
 *                of gelu, FP16 rounding emulated step by step; gelu' stays
 *                the tanh form's), so that the GPU result differs only by
 *                its FP32 summation order and its hardware sin/cos/exp.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>

#define NF 128
#define NIN 256
#define NH 128
#define NOUT 3
#define P_W1 0
#define P_B1 (P_W1 + NIN * NH)
#define P_W2 (P_B1 + NH)
#define P_B2 (P_W2 + NH * NH)
#define P_W3 (P_B2 + NH)
#define P_B3 (P_W3 + NH * NH)
#define P_W4 (P_B3 + NH)
#define P_B4 (P_W4 + NH * NOUT)
#define P_TOTAL (P_B4 + NOUT)

int tpref_mlp_nparams(void) { return P_TOTAL; }

/* float -> bfloat16 -> float, round to nearest even (finite inputs) */
static float bf16r(float v)
{
    uint32_t u;
    memcpy(&u, &v, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&v, &u, 4);
    return v;
}

double tpref_bf16_round(double v) { return (double)bf16r((float)v); }

#define GELU_C1 1.5957691216057308   /* 2 sqrt(2/pi) */
#define GELU_C2 0.044715

static double gelu(double z)
{
    double a = GELU_C1 * (z + GELU_C2 * z * z * z);
    return z / (1.0 + exp(-a));
}

static double gelu_d(double z)
{
    double a = GELU_C1 * (z + GELU_C2 * z * z * z);
    double s = 1.0 / (1.0 + exp(-a));
    return s + z * s * (1.0 - s) * GELU_C1 * (1.0 + 3.0 * GELU_C2 * z * z);
}

/* The activation of the GPU kernels (tp_mlp.hip, gelu_pk): max(z, 0) minus
 * a degree-8 polynomial for the tail G(u) = u - gelu(u) in
 * s = min(|z|, 4) / 2 - 1, evaluated in FP16 on the matrix-core side:
 * z converted from float with round toward zero, every FMA and the final
 * difference rounded to nearest even in FP16 (the packed FP16 instructions
 * of the kernels).  The coefficients are FP16 values. */
static const double gelu_tail[9] = {
    0.045379638671875, -0.1727294921875, 0.21728515625, 0.0158538818359375,
    -0.3017578125, 0.2413330078125, 0.018310546875, -0.08465576171875,
    0.020843505859375};

/* v rounded to FP16 (11 significant bits, subnormals down to 2^-24, largest
 * finite 65504): to nearest even, or toward zero (which saturates at 65504) */
static double f16r(double v, int rtz)
{
    if (v == 0.0 || isnan(v)) return v;
    if (isinf(v)) return rtz ? copysign(65504.0, v) : v;
    int e;
    const double a = fabs(v);
    frexp(a, &e);                 /* a = m 2^e, m in [0.5, 1) */
    int qe = e - 11;
    if (qe < -24) qe = -24;
    const double q = ldexp(1.0, qe);
    double r = a / q;             /* exact: q is a power of two */
    r = rtz ? trunc(r) : nearbyint(r);
    double w = r * q;
    if (w > 65504.0) w = rtz ? 65504.0 : INFINITY;
    return copysign(w, v);
}

static double gelu_gpu(double z)
{
    const double zh = f16r((double)(float)z, 1);
    const double uc = fmin(fabs(zh), 4.0);
    const double s = f16r(fma(uc, 0.5, -1.0), 0);
    double p = gelu_tail[8];
    for (int k = 7; k >= 0; k--) p = f16r(fma(p, s, gelu_tail[k]), 0);
    return f16r(fmax(zh, 0.0) - p, 0);
}

void tpref_gelu_gpu(int n, const double *z, double *g)
{
    for (int i = 0; i < n; i++) g[i] = gelu_gpu(z[i]);
}

void tpref_gelu(int n, const double *z, double *g, double *gd)
{
    for (int i = 0; i < n; i++) {
        if (g) g[i] = gelu(z[i]);
        if (gd) gd[i] = gelu_d(z[i]);
    }
}

/* per-sample activations */
typedef struct {
    double f[NIN];
    double z1[NH], h1[NH], z2[NH], h2[NH], z3[NH], h3[NH];
    double out[NOUT];
} Act;

static void features(const double *q, const double *x, int em, double *f)
{
    for (int j = 0; j < NF; j++) {
        double r;
        if (em) {
            float xf = (float)x[0], yf = (float)x[1], zf = (float)x[2];
            float u = fmaf((float)q[3 * j + 2], zf,
                           fmaf((float)q[3 * j + 1], yf, (float)q[3 * j] * xf));
            r = (double)(u - rintf(u));
        } else {
            double u = q[3 * j] * x[0] + q[3 * j + 1] * x[1] + q[3 * j + 2] * x[2];
            r = u - rint(u);
        }
        double s = sin(2.0 * M_PI * r), c = cos(2.0 * M_PI * r);
        if (em) {
            s = bf16r((float)s);
            c = bf16r((float)c);
        }
        f[j] = s;
        f[NF + j] = c;
    }
}

/* W (as the kernels use it: bf16 for layers 1-3 when emulating) */
static double wv(const double *p, int i, int em, int layer)
{
    return (em && layer < 4) ? (double)bf16r((float)p[i]) : p[i];
}

static void dense(const double *p, int wo, int bo, int nin, int nout, const double *in,
                  double *z, int em, int layer)
{
    for (int n = 0; n < nout; n++) {
        double s = 0.0;
        for (int k = 0; k < nin; k++) s += in[k] * wv(p, wo + k * nout + n, em, layer);
        if (em) s = (double)((float)s + (float)p[bo + n]);
        else s += p[bo + n];
        z[n] = s;
    }
}

static void forward(const double *p, const double *q, const double *x, int em, Act *a)
{
    features(q, x, em, a->f);
    dense(p, P_W1, P_B1, NIN, NH, a->f, a->z1, em, 1);
    for (int n = 0; n < NH; n++) a->h1[n] = em ? bf16r((float)gelu_gpu(a->z1[n])) : gelu(a->z1[n]);
    dense(p, P_W2, P_B2, NH, NH, a->h1, a->z2, em, 2);
    for (int n = 0; n < NH; n++) a->h2[n] = em ? bf16r((float)gelu_gpu(a->z2[n])) : gelu(a->z2[n]);
    dense(p, P_W3, P_B3, NH, NH, a->h2, a->z3, em, 3);
    for (int n = 0; n < NH; n++) a->h3[n] = em ? gelu_gpu(a->z3[n]) : gelu(a->z3[n]);
    dense(p, P_W4, P_B4, NH, NOUT, a->h3, a->out, em, 4);
}

/* network outputs (before scale and guide field) at n points x[n][3] */
void tpref_mlp_forward(const double *params, const double *q, int n, const double *x,
                       double *out, int emulate)
{
    Act *a = (Act *)malloc(sizeof(Act));
    for (int i = 0; i < n; i++) {
        forward(params, q, x + 3 * i, emulate, a);
        for (int c = 0; c < NOUT; c++) out[3 * i + c] = a->out[c];
    }
    free(a);
}

/* mean-squared-error loss over n samples and its gradient (P_TOTAL) */
void tpref_mlp_loss_grad(const double *params, const double *q, int n, const double *x,
                         const double *y, double *grad, double *loss, int emulate)
{
    const int em = emulate;
    Act *a = (Act *)malloc(sizeof(Act));
    double dz3[NH], dz2[NH], dz1[NH];
    memset(grad, 0, sizeof(double) * P_TOTAL);
    double L = 0.0;
    const double inv = 1.0 / (3.0 * (double)n);
    for (int i = 0; i < n; i++) {
        forward(params, q, x + 3 * i, em, a);
        double dout[NOUT];
        for (int c = 0; c < NOUT; c++) {
            double e = a->out[c] - y[3 * i + c];
            L += e * e;
            dout[c] = 2.0 * e * inv;
            if (em) dout[c] = (double)(float)dout[c];
        }
        /* layer 4 (float on the GPU) */
        for (int k = 0; k < NH; k++)
            for (int c = 0; c < NOUT; c++) grad[P_W4 + k * NOUT + c] += a->h3[k] * dout[c];
        for (int c = 0; c < NOUT; c++) grad[P_B4 + c] += dout[c];
        for (int k = 0; k < NH; k++) {
            double s = 0.0;
            for (int c = 0; c < NOUT; c++) s += dout[c] * params[P_W4 + k * NOUT + c];
            double d = s * gelu_d(a->z3[k]);
            dz3[k] = em ? bf16r((float)d) : d;
        }
        /* layer 3 */
        for (int k = 0; k < NH; k++)
            for (int m = 0; m < NH; m++) grad[P_W3 + k * NH + m] += a->h2[k] * dz3[m];
        for (int m = 0; m < NH; m++) grad[P_B3 + m] += dz3[m];
        for (int k = 0; k < NH; k++) {
            double s = 0.0;
            for (int m = 0; m < NH; m++) s += wv(params, P_W3 + k * NH + m, em, 3) * dz3[m];
            double d = s * gelu_d(a->z2[k]);
            dz2[k] = em ? bf16r((float)d) : d;
        }
        /* layer 2 */
        for (int k = 0; k < NH; k++)
            for (int m = 0; m < NH; m++) grad[P_W2 + k * NH + m] += a->h1[k] * dz2[m];
        for (int m = 0; m < NH; m++) grad[P_B2 + m] += dz2[m];
        for (int k = 0; k < NH; k++) {
            double s = 0.0;
            for (int m = 0; m < NH; m++) s += wv(params, P_W2 + k * NH + m, em, 2) * dz2[m];
            double d = s * gelu_d(a->z1[k]);
            dz1[k] = em ? bf16r((float)d) : d;
        }
        /* layer 1 */
        for (int k = 0; k < NIN; k++)
            for (int m = 0; m < NH; m++) grad[P_W1 + k * NH + m] += a->f[k] * dz1[m];
        for (int m = 0; m < NH; m++) grad[P_B1 + m] += dz1[m];
    }
    *loss = L * inv;
    free(a);
}
