/*
 * tp_math.h -- project-written FP64 sine/cosine shared by the GPU kernels
 * and the CPU oracle.
 *
 * The reference tracer takes sin/cos from the system libm
 * (src/force_free.c:143-145, src/tracking.c:816-817, :853-854).  libm's
 * results are not reproducible on a GPU, and in the chaotic ABC field a
 * one-ulp difference grows over thousands of steps.  Both sides of the
 * parity test therefore evaluate this one routine, with the same IEEE
 * operations in the same order (fma where written, no other contraction),
 * which makes FP64 trajectories bitwise reproducible between host and
 * device.  Accuracy against libm is tested separately (within 2 ulp).
 *
 * Method:
 *   - Cody & Waite (1980) additive argument reduction by pi/2 with a
 *     three-part constant.  x - k*P1 is formed by one fma and is exact for
 *     |x| < 2^52 (the difference is a multiple of 2^-52 below 2 in
 *     magnitude); the P2/P3 terms are carried as a double-double (r, rlo)
 *     with Knuth's TwoSum and fma product errors.
 *   - Taylor series of sin (to r^17) and cos (to r^18) on |r| <= pi/4,
 *     evaluated by Horner's rule with fma; the series truncation error is
 *     below 2^-60 relative there.
 *   - cos uses 1 - r^2/2 split with its rounding error recovered, as in
 *     the classical kernel of Sun's fdlibm.
 * The constants are pi/2 split into doubles and 1/n!.
 */
#ifndef TP_MATH_H
#define TP_MATH_H

#include <math.h>

#if defined(__HIPCC__)
#define TP_HD __host__ __device__ __forceinline__
#else
#define TP_HD static inline
#endif

#define TP_2_OVER_PI 0x1.45f306dc9c883p-1
#define TP_PIO2_1    0x1.921fb54442d18p+0
#define TP_PIO2_2    0x1.1a62633145c07p-54
#define TP_PIO2_3   -0x1.f1976b7ed8fbcp-110

/* sin(x) and cos(x) together; NaN/Inf in give NaN out. */
TP_HD void tp_sincos(double x, double *sp, double *cp)
{
    double k = rint(x * TP_2_OVER_PI);
    double r1 = fma(-k, TP_PIO2_1, x);          /* exact */
    double w = k * TP_PIO2_2;
    double werr = fma(k, TP_PIO2_2, -w);        /* k*P2 = w + werr exactly */
    /* TwoSum(r1, -w) */
    double r = r1 - w;
    double bb = r - r1;
    double rlo = (r1 - (r - bb)) + (-w - bb);
    rlo = rlo - werr;
    rlo = fma(-k, TP_PIO2_3, rlo);

    double z = r * r;
    /* sin(r) = r + r z Ps(z) */
    double ps = 0x1.952c77030ad4ap-49;                      /*  1/17! */
    ps = fma(ps, z, -0x1.ae7f3e733b81fp-41);                /* -1/15! */
    ps = fma(ps, z, 0x1.6124613a86d09p-33);                 /*  1/13! */
    ps = fma(ps, z, -0x1.ae64567f544e4p-26);                /* -1/11! */
    ps = fma(ps, z, 0x1.71de3a556c734p-19);                 /*  1/9!  */
    ps = fma(ps, z, -0x1.a01a01a01a01ap-13);                /* -1/7!  */
    ps = fma(ps, z, 0x1.1111111111111p-7);                  /*  1/5!  */
    ps = fma(ps, z, -0x1.5555555555555p-3);                 /* -1/3!  */
    double s = r + fma(r * z, ps, rlo);
    /* cos(r) = 1 - z/2 + z^2 Pc(z) */
    double pc = 0x1.6827863b97d97p-53;                      /*  1/18! */
    pc = fma(pc, z, -0x1.ae7f3e733b81fp-45);                /* -1/16! */
    pc = fma(pc, z, 0x1.93974a8c07c9dp-37);                 /*  1/14! */
    pc = fma(pc, z, -0x1.1eed8eff8d898p-29);                /* -1/12! */
    pc = fma(pc, z, 0x1.27e4fb7789f5cp-22);                 /*  1/10! */
    pc = fma(pc, z, -0x1.a01a01a01a01ap-16);                /* -1/8!  */
    pc = fma(pc, z, 0x1.6c16c16c16c17p-10);                 /*  1/6!  */
    pc = fma(pc, z, -0x1.5555555555555p-5);                 /* -1/4!  */
    pc = -pc;                                               /* z^2 * (1/4! - ...) */
    double hz = 0.5 * z;
    double wc = 1.0 - hz;
    double c = wc + (((1.0 - wc) - hz) + fma(z * z, pc, -r * rlo));

    /* quadrant q = k mod 4, kept in floating point (no int conversion of
     * huge or non-finite k) */
    double q = k - 4.0 * floor(k * 0.25);
    double so, co;
    if (q == 0.0) {
        so = s; co = c;
    } else if (q == 1.0) {
        so = c; co = -s;
    } else if (q == 2.0) {
        so = -s; co = -c;
    } else {
        so = -c; co = s;
    }
    *sp = so;
    *cp = co;
}

TP_HD double tp_sin(double x)
{
    double s, c;
    tp_sincos(x, &s, &c);
    return s;
}

TP_HD double tp_cos(double x)
{
    double s, c;
    tp_sincos(x, &s, &c);
    return c;
}

/*
 * cos(x) alone, for sums over many modes: reduction by pi instead of pi/2
 * (k = rint(x/pi), r = x - k pi in [-pi/2, pi/2], three-part constant, the
 * first product exact by the same fma argument as above), then one even
 * Taylor polynomial to r^22 (truncation < 4e-20 on that interval) and the
 * sign (-1)^k.  About 40 % fewer operations than tp_sincos.
 */
TP_HD double tp_cos_pi(double x)
{
    double k = rint(x * 0x1.45f306dc9c883p-2);             /* 1/pi */
    double r = fma(-k, 0x1.921fb54442d18p+1, x);            /* exact */
    r = fma(-k, 0x1.1a62633145c07p-53, r);
    r = fma(-k, -0x1.f1976b7ed8fbcp-109, r);
    double z = r * r;
    double p = 0x1.0ce396db7f853p-70;                       /*  1/22! */
    p = fma(p, z, -0x1.e542ba4020225p-62);                  /* -1/20! */
    p = fma(p, z, 0x1.6827863b97d97p-53);                   /*  1/18! */
    p = fma(p, z, -0x1.ae7f3e733b81fp-45);                  /* -1/16! */
    p = fma(p, z, 0x1.93974a8c07c9dp-37);                   /*  1/14! */
    p = fma(p, z, -0x1.1eed8eff8d898p-29);                  /* -1/12! */
    p = fma(p, z, 0x1.27e4fb7789f5cp-22);                   /*  1/10! */
    p = fma(p, z, -0x1.a01a01a01a01ap-16);                  /* -1/8!  */
    p = fma(p, z, 0x1.6c16c16c16c17p-10);                   /*  1/6!  */
    p = fma(p, z, -0x1.5555555555555p-5);                   /* -1/4!  */
    p = fma(p, z, 0.5);
    double c = fma(-z, p, 1.0);                             /* 1 - z/2 + z^2/24 - ... */
    double half = k * 0.5;
    return (half == floor(half)) ? c : -c;
}

/*
 * Deterministic log2 / exp2 / pow for the adaptive stepper's controller
 * (step-size factors err^(1/(2k+1))).  Only reproducibility between host and
 * device matters there; accuracy is ~1e-15 relative.
 *   log2: frexp into [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s) with
 *         s = (m-1)/(m+1), odd series to s^23;
 *   exp2: 2^y = 2^n e^(f ln 2) with n = rint(y), Taylor series to r^14.
 * Arguments: x > 0 finite (log2, pow); |y| < 1000 (exp2).
 */
TP_HD double tp_log2(double x)
{
    int e;
    double m = frexp(x, &e);
    if (m < 0x1.6a09e667f3bcdp-1) {   /* sqrt(1/2) */
        m = 2.0 * m;
        e = e - 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    double p = 0x1.642c8590b2164p-5;              /* 1/23 */
    p = fma(p, s2, 0x1.8618618618618p-5);         /* 1/21 */
    p = fma(p, s2, 0x1.af286bca1af28p-5);         /* 1/19 */
    p = fma(p, s2, 0x1.e1e1e1e1e1e1ep-5);         /* 1/17 */
    p = fma(p, s2, 0x1.1111111111111p-4);         /* 1/15 */
    p = fma(p, s2, 0x1.3b13b13b13b14p-4);         /* 1/13 */
    p = fma(p, s2, 0x1.745d1745d1746p-4);         /* 1/11 */
    p = fma(p, s2, 0x1.c71c71c71c71cp-4);         /* 1/9 */
    p = fma(p, s2, 0x1.2492492492492p-3);         /* 1/7 */
    p = fma(p, s2, 0x1.999999999999ap-3);         /* 1/5 */
    p = fma(p, s2, 0x1.5555555555555p-2);         /* 1/3 */
    double lnm = 2.0 * s + 2.0 * s * s2 * p;
    return (double)e + lnm * 0x1.71547652b82fep+0; /* 1/ln 2 */
}

TP_HD double tp_exp2(double y)
{
    double n = rint(y);
    double r = (y - n) * 0x1.62e42fefa39efp-1;    /* ln 2 */
    double p = 0x1.93974a8c07c9dp-37;              /* 1/14! */
    p = fma(p, r, 0x1.6124613a86d09p-33);          /* 1/13! */
    p = fma(p, r, 0x1.1eed8eff8d898p-29);          /* 1/12! */
    p = fma(p, r, 0x1.ae64567f544e4p-26);          /* 1/11! */
    p = fma(p, r, 0x1.27e4fb7789f5cp-22);          /* 1/10! */
    p = fma(p, r, 0x1.71de3a556c734p-19);          /* 1/9! */
    p = fma(p, r, 0x1.a01a01a01a01ap-16);          /* 1/8! */
    p = fma(p, r, 0x1.a01a01a01a01ap-13);          /* 1/7! */
    p = fma(p, r, 0x1.6c16c16c16c17p-10);          /* 1/6! */
    p = fma(p, r, 0x1.1111111111111p-7);           /* 1/5! */
    p = fma(p, r, 0x1.5555555555555p-5);           /* 1/4! */
    p = fma(p, r, 0x1.5555555555555p-3);           /* 1/3! */
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)n);
}

TP_HD double tp_pow(double x, double y)
{
    return tp_exp2(y * tp_log2(x));
}

/* x^e for an integer e >= 0 by repeated multiplication, left to right */
TP_HD double tp_powi(double x, int e)
{
    double r = 1.0;
    for (int i = 0; i < e; i++) r = r * x;
    return r;
}

#endif /* TP_MATH_H */
