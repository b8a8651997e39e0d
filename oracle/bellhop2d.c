/*
 * ORACLE -- test infrastructure only. Not part of the product.
 *
 * Plain-C restatement of the reference's 2D ray/beam tracing hot path
 * (bellhopcxx, CPU build), used by tests/ and bench.py's cpu_baseline as the
 * checker for the HIP path. It reads the raw environment (bhc_env2d) and
 * performs its own preprocessing, so it shares no code with the product.
 *
 * Compiled twice: OR_REAL=double (FP64) and OR_REAL=float (FP32, the
 * reference's BHC_USE_FLOATS build). Built with gcc -O2 -ffp-contract=off and
 * no -march, like the reference CPU build (config/SetupCommon.cmake:26-38), so
 * every a*b+c is two roundings. Complex arithmetic uses C99 complex, which gcc
 * lowers exactly like libstdc++'s std::complex (inline multiply with the
 * __muldc3 NaN fallback, __divdc3 for division, glibc csqrt/cexp/cabs).
 *
 * Per-function reference citations are given as (file:line) into the reference.
 * Parity status of each part is documented in the module docstring of
 * tests/test_oracle.py (what pins it).
 */
#include <complex.h>
#include <math.h>
#include <float.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <stdatomic.h>
#include <pthread.h>
#include <unistd.h>

#include "../include/bhc_hip.h"

#ifndef OR_REAL
#define OR_REAL double
#endif

typedef OR_REAL real;

#if OR_IS_FLOAT
typedef float complex cpx;
#define RL(a) (a##f)
#define REAL_MAX FLT_MAX
#define REAL_EPSILON FLT_EPSILON
#define REAL_MINPOS FLT_MIN
#define REAL_PI ((float)M_PI)
#define INFINITESIMAL_STEP_SIZE (1e-3f)
#define SFX(name) name##_f
#define RSQRT sqrtf
#define RABS fabsf
#define RCOS cosf
#define RATAN atanf
#define RSIN sinf
#define RATAN2 atan2f
#define REXP expf
#define RROUND roundf
#define RNEXT nextafterf
#define RPOW powf
#define RFMA fmaf
#define CSQRT csqrtf
#define CCONJ conjf
#define CEXP cexpf
#define CABS cabsf
#else
typedef double complex cpx;
#define RL(a) a
#define REAL_MAX DBL_MAX
#define REAL_EPSILON DBL_EPSILON
#define REAL_MINPOS DBL_MIN
#define REAL_PI M_PI
#define INFINITESIMAL_STEP_SIZE (1e-6)
#define SFX(name) name##_d
#define RSQRT sqrt
#define RABS fabs
#define RCOS cos
#define RATAN atan
#define RSIN sin
#define RATAN2 atan2
#define REXP exp
#define RROUND round
#define RNEXT nextafter
#define RPOW pow
#define RFMA fma
#define CSQRT csqrt
#define CCONJ conj
#define CEXP cexp
#define CABS cabs
#endif
#define FL(a) (a##f)
#define MKC(a, b) __builtin_complex((real)(a), (real)(b))
#define SQ(a) ((a) * (a))
#define CUBE(a) ((a) * (a) * (a))

static const real RadDeg = RL(180.0) / REAL_PI; /* src/common.hpp:203-204 */
static const real DegRad = REAL_PI / RL(180.0);

#define MaxN 100000 /* include/bhc/structs.hpp:66 */

typedef struct { real x, y; } v2;
static inline v2 V2(real x, real y) { v2 r = {x, y}; return r; }
static inline v2 vadd(v2 a, v2 b) { return V2(a.x + b.x, a.y + b.y); }
static inline v2 vsub(v2 a, v2 b) { return V2(a.x - b.x, a.y - b.y); }
static inline v2 vscl(real s, v2 a) { return V2(s * a.x, s * a.y); }
static inline v2 vdivs(v2 a, real s) { return V2(a.x / s, a.y / s); }
static inline real vdot(v2 a, v2 b) { return a.x * b.x + a.y * b.y; } /* glm: tmp=a*b; tmp.x+tmp.y */
static inline real vlen(v2 a) { return RSQRT(vdot(a, a)); }
static inline real rmin(real a, real b) { return (b < a) ? b : a; } /* std::min */
static inline real rmax(real a, real b) { return (a < b) ? b : a; } /* std::max */
static inline int imin(int a, int b) { return (b < a) ? b : a; }
static inline int imax(int a, int b) { return (a < b) ? b : a; }
static inline real spacing(real v) { return RABS(RNEXT(v, (real)0.0f) - v); } /* src/common.hpp:216-220 */

/* ------------------------------------------------------------------------- */
/* Preprocessed problem                                                        */
/* ------------------------------------------------------------------------- */

typedef struct { cpx cP, cS; real rho, depth; char bc; } HS;

typedef struct {
    v2 x, t, n, Noden, Nodet;
    real Len, kappa, Dx, Dxx, Dss;
    cpx cP, cS;
    real rho;
} BdryPt;

typedef struct {
    char type[2];
    int npts;
    BdryPt *bd;
} BdryInfo;

typedef struct { real theta, r, phi; } ReflCoef;

typedef struct {
    char RunType[7];
    char Type[4];
    char ssp_type;
    real freq0;
    /* SSP */
    int npts, Nr;
    real *z, *rho;
    cpx *c, *cz, *n2, *n2z, *coef[4], *work[4];
    real *seg_r, *cMat, *czMat;
    /* boundary */
    HS top, bot;
    BdryInfo ati, bty;
    int ntrc, nbrc;
    ReflCoef *trc, *brc;
    /* positions */
    int NSz, NRz, NRr, NRz_per_range;
    float *Sz, *Rz;
    real *Rr;
    real Delta_r;
    /* angles */
    int Nalpha, iSingle;
    real *alpha, Dalpha;
    /* beam */
    real deltas, box_r, box_z, epsMultiplier, rLoop;
    int iBeamWindow;
    int nsbp;
    real *sbp;
} Prob;

/* Francois-Garrison (src/module/atten.cpp:38-72) */
static real franc_garr(real f, const bhc_env2d *env)
{
    real T = (real)env->fg_T, S = (real)env->fg_S, pH = (real)env->fg_pH, zb = (real)env->fg_z_bar;
    real c, a1, a2, a3, p1, p2, p3, f1, f2;
    c = FL(1412.0) + FL(3.21) * T + FL(1.19) * S + FL(0.0167) * zb;
    a1 = FL(8.86) / c * RPOW(FL(10.0), FL(0.78) * pH - FL(5.0));
    p1 = FL(1.0);
    f1 = FL(2.8) * RSQRT(S / 35) * RPOW(FL(1.0), FL(4.0) - FL(1245.0) / (T + FL(273.0)));
    a2 = FL(21.44) * S / c * (FL(1.0) + FL(0.025) * T);
    p2 = FL(1.0) - RL(1.37e-4) * zb + RL(6.2e-9) * SQ(zb);
    f2 = FL(8.17) * RPOW(FL(10.0), FL(8.0) - FL(1990.0) / (T + 273)) / (FL(1.0) + FL(0.0018) * (S - FL(35.0)));
    p3 = FL(1.0) - RL(3.83e-5) * zb + RL(4.9e-10) * SQ(zb);
    if(T < FL(20.0)) {
        a3 = RL(4.937e-4) - RL(2.59e-5) * T + RL(9.11e-7) * SQ(T) - RL(1.5e-8) * CUBE(T);
    } else {
        a3 = RL(3.964e-4) - RL(1.146e-5) * T + RL(1.45e-7) * SQ(T) - RL(6.5e-10) * CUBE(T);
    }
    return a1 * p1 * (f1 * SQ(f)) / (SQ(f1) + SQ(f)) + a2 * p2 * (f2 * SQ(f)) / (SQ(f2) + SQ(f))
        + a3 * p3 * SQ(f);
}

/* crci, src/module/atten.cpp:99-194 */
static cpx crci(const bhc_env2d *env, real freq0, real z, real c, real alpha, const char *AttenUnit)
{
    real freq = freq0;
    real omega = FL(2.0) * REAL_PI * freq;
    real alphaT = FL(0.0);
    switch(AttenUnit[0]) {
    case 'N': alphaT = alpha; break;
    case 'M': alphaT = alpha / RL(8.6858896); break;
    case 'm':
        /* power law in f / freq0; the run frequency is freq0, where both of the
         * reference's branches (below and above fT) give the factor 1 */
        alphaT = alpha / RL(8.6858896);
        alphaT *= RPOW(freq / freq0, FL(1.0));
        break;
    case 'F': alphaT = alpha * freq / RL(8685.8896); break;
    case 'W':
        if(RABS(c) >= REAL_EPSILON) alphaT = alpha * freq / (RL(8.6858896) * c);
        break;
    case 'Q':
        if(RABS(c * alpha) >= REAL_EPSILON) alphaT = omega / (FL(2.0) * c * alpha);
        break;
    case 'L':
        if(RABS(c) >= REAL_EPSILON) alphaT = alpha * omega / c;
        break;
    }
    switch(AttenUnit[1]) {
    case 'T': {
        real f2 = SQ(freq / FL(1000.0));
        real Thorp = FL(3.3e-3) + FL(0.11) * f2 / (FL(1.0) + f2) + FL(44.0) * f2 / (FL(4100.0) + f2)
            + FL(3e-4) * f2;
        Thorp /= FL(8685.8896);
        alphaT += Thorp;
    } break;
    case 'F': {
        real fg = franc_garr(freq / FL(1000.0), env);
        fg /= FL(8685.8896);
        alphaT += fg;
    } break;
    case 'B': /* biological layers (Diachok); 1.0 is a double literal in the reference */
        for(int ib = 0; ib < env->NBioLayers; ++ib) {
            real z1 = (real)env->bio[5 * ib], z2 = (real)env->bio[5 * ib + 1], f0 = (real)env->bio[5 * ib + 2];
            real q = (real)env->bio[5 * ib + 3], a0 = (real)env->bio[5 * ib + 4];
            if(z >= z1 && z <= z2) {
                real a = a0 / (SQ(1.0 - SQ(f0) / SQ(freq)) + 1.0 / SQ(q)); /* dB/km */
                a /= FL(8685.8896);
                alphaT += a;
            }
        }
        break;
    }
    alphaT *= SQ(c) / omega;
    return MKC(c, alphaT);
}

/* cSpline, src/curves.hpp:76-188 (de Boor) */
static void cspline(const real *tau, cpx *c1, cpx *c2, cpx *c3, cpx *c4, int n, int ibcbeg, int ibcend)
{
    cpx g = 0, dtau, divdf1, divdf3;
    int l = n - 1;
    for(int m = 1; m < n; ++m) {
        c3[m] = (cpx)(tau[m] - tau[m - 1]);
        c4[m] = (c1[m] - c1[m - 1]) / c3[m];
    }
    if(ibcbeg == 0) {
        if(n > 2) {
            c4[0] = c3[2];
            c3[0] = c3[1] + c3[2];
            c2[0] = ((c3[1] + FL(2.0) * c3[0]) * c4[1] * c3[2] + SQ(c3[1]) * c4[2]) / c3[0];
        } else {
            c4[0] = (cpx)FL(1.0);
            c3[0] = (cpx)FL(1.0);
            c2[0] = FL(2.0) * c4[1];
        }
    } else if(ibcbeg == 1) {
        c4[0] = (cpx)FL(1.0);
        c3[0] = (cpx)FL(0.0);
    } else if(ibcbeg == 2) {
        c4[0] = (cpx)FL(2.0);
        c3[0] = (cpx)FL(1.0);
        c2[0] = FL(3.0) * c4[1] - c2[0] * c3[1] * FL(0.5);
    }
    for(int m = 1; m < l; ++m) {
        g = -c3[m + 1] / c4[m - 1];
        c2[m] = g * c2[m - 1] + FL(3.0) * (c3[m] * c4[m + 1] + c3[m + 1] * c4[m]);
        c4[m] = g * c3[m - 1] + FL(2.0) * (c3[m] + c3[m + 1]);
    }
    if(ibcend != 1) {
        if(ibcend == 0) {
            if(n == 2 && ibcbeg == 0) {
                c2[n - 1] = c4[n - 1];
            } else if((n == 3 && ibcbeg == 0) || n == 2) {
                c2[n - 1] = FL(2.0) * c4[n - 1];
                c4[n - 1] = (cpx)FL(1.0);
                g = FL(1.0) / -c4[n - 2];
            } else {
                g = c3[n - 2] + c3[n - 1];
                c2[n - 1] = ((c3[n - 1] + FL(2.0) * g) * c4[n - 1] * c3[n - 2]
                             + SQ(c3[n - 1]) * (c1[n - 2] - c1[n - 3]) / c3[n - 2])
                    / g;
                g = -g / c4[n - 2];
                c4[n - 1] = c3[n - 2];
            }
        } else if(ibcend == 2) {
            c2[n - 1] = FL(3.0) * c4[n - 1] + c2[n - 1] * c3[n - 1] * FL(0.5);
            c4[n - 1] = (cpx)FL(2.0);
            g = FL(1.0) / -c4[n - 2];
        }
        if(ibcbeg > 0 || n > 2) {
            c4[n - 1] += g * c3[n - 2];
            c2[n - 1] = (g * c2[n - 2] + c2[n - 1]) / c4[n - 1];
        }
    }
    for(int j = l - 1; j >= 0; --j) c2[j] = (c2[j] - c3[j] * c2[j + 1]) / c4[j];
    for(int i = 1; i < n; ++i) {
        dtau = c3[i];
        divdf1 = (c1[i] - c1[i - 1]) / dtau;
        divdf3 = c2[i - 1] + c2[i] - FL(2.0) * divdf1;
        c3[i - 1] = FL(2.0) * (divdf1 - c2[i - 1] - divdf3) / dtau;
        c4[i - 1] = (divdf3 / dtau) * (FL(6.0) / dtau);
    }
    c3[n - 1] = c3[l - 1] + (tau[n - 1] - tau[l - 1]) * c4[l - 1];
    c4[n - 1] = (cpx)FL(0.0);
    for(int i = 0; i < l; ++i) {
        real dt = tau[i + 1] - tau[i];
        c4[n - 1] += dt * (c1[i] + dt * (c2[i] * FL(0.5) + dt * (c3[i] / FL(6.0) + dt * c4[i] / FL(24.0))));
    }
    c4[n - 1] /= tau[n - 1] - tau[0];
}

/* PCHIP helpers, src/curves.hpp:225-297 */
static real fprime_interior(real del1, real del2, real fprime)
{
    if(del1 * del2 > FL(0.0)) {
        if(del1 > FL(0.0)) return rmin(rmax(fprime, RL(0.0)), RL(3.0) * rmin(del1, del2));
        return rmax(rmin(fprime, RL(0.0)), RL(3.0) * rmax(del1, del2));
    }
    return RL(0.0);
}
static real fprime_left_end(real del1, real del2, real fprime)
{
    if(del1 * fprime <= RL(0.0)) return FL(0.0);
    if((del1 * del2 <= RL(0.0)) && (RABS(fprime) > RABS(RL(3.0) * del1))) return RL(3.0) * del1;
    return fprime;
}
static cpx fpi_c(cpx d1, cpx d2, cpx f)
{
    return fprime_interior(creal(d1), creal(d2), creal(f)) + fprime_interior(cimag(d1), cimag(d2), cimag(f)) * I;
}
static cpx fpl_c(cpx d1, cpx d2, cpx f)
{
    return fprime_left_end(creal(d1), creal(d2), creal(f)) + fprime_left_end(cimag(d1), cimag(d2), cimag(f)) * I;
}
static cpx fpr_c(cpx d1, cpx d2, cpx f) { return fpl_c(d2, d1, f); }

static void h_del(const real *x, const cpx *y, int ix, real *h1, real *h2, cpx *del1, cpx *del2)
{
    *h1 = x[ix] - x[ix - 1];
    *h2 = x[ix + 1] - x[ix];
    *del1 = (y[ix] - y[ix - 1]) / *h1;
    *del2 = (y[ix + 1] - y[ix]) / *h2;
}

/* pchip, src/curves.hpp:299-372 */
static void pchip(const real *x, const cpx *y, int n, cpx *P1, cpx *P2, cpx *P3, cpx *P4, cpx *W1, cpx *W2,
                  cpx *W3, cpx *W4)
{
    real h1, h2;
    cpx del1, del2, f1, f2, f1p, f2p, fpT;
    if(n == 2) {
        P1[0] = y[0];
        P2[0] = (y[1] - y[0]) / (x[1] - x[0]);
        P3[0] = RL(0.0);
        P4[0] = RL(0.0);
        return;
    }
    for(int i = 0; i < n; ++i) P1[i] = y[i];
    h_del(x, y, 1, &h1, &h2, &del1, &del2);
    fpT = ((RL(2.0) * h1 + h2) * del1 - h1 * del2) / (h1 + h2);
    P2[0] = fpl_c(del1, del2, fpT);
    h_del(x, y, n - 2, &h1, &h2, &del1, &del2);
    fpT = (-h2 * del1 + (h1 + RL(2.0) * h2) * del2) / (h1 + h2);
    P2[n - 1] = fpr_c(del1, del2, fpT);
    for(int i = 0; i < n; ++i) W1[i] = P1[i];
    W2[0] = P2[0];
    W2[n - 1] = P2[n - 1];
    cspline(x, W1, W2, W3, W4, n, 1, 1);
    for(int ix = 1; ix < n - 1; ++ix) {
        h_del(x, y, ix, &h1, &h2, &del1, &del2);
        P2[ix] = fpi_c(del1, del2, W2[ix]);
    }
    for(int ix = 0; ix < n - 1; ++ix) {
        real h = x[ix + 1] - x[ix];
        f1 = P1[ix];
        f2 = P1[ix + 1];
        f1p = P2[ix];
        f2p = P2[ix + 1];
        P3[ix] = (RL(3.0) * (f2 - f1) - h * (RL(2.0) * f1p + f2p)) / SQ(h);
        P4[ix] = (h * (f1p + f2p) - RL(2.0) * (f2 - f1)) / CUBE(h);
    }
}

static void *xcalloc(size_t n, size_t s)
{
    void *p = calloc(n ? n : 1, s);
    if(!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}

static void hs_from_env(const bhc_env2d *env, real freq0, const bhc_halfspace_in *in, HS *hs, const char *au)
{
    hs->bc = in->bc;
    hs->depth = (real)in->depth;
    hs->rho = (real)in->rho;
    hs->cP = 0;
    hs->cS = 0;
    if(in->bc == 'A') { /* src/module/boundarycond.hpp:166-176 */
        hs->cP = crci(env, freq0, (real)in->zTemp, (real)in->alphaR, (real)in->alphaI, au);
        hs->cS = crci(env, freq0, (real)in->zTemp, (real)in->betaR, (real)in->betaI, au);
    } else if(in->bc == 'G') { /* grain size, src/module/boundarycond.hpp:177-222 */
        real Mz = (real)in->Mz, vr, alpha2_f, rho;
        if(Mz >= FL(-1.0) && Mz < FL(1.0)) {
            vr = FL(0.002709) * SQ(Mz) - FL(0.056452) * Mz + FL(1.2778);
            rho = FL(0.007797) * SQ(Mz) - FL(0.17057) * Mz + FL(2.3139);
        } else if(Mz >= FL(1.0) && Mz < FL(5.3)) {
            vr = FL(-0.0014881) * CUBE(Mz) + FL(0.0213937) * SQ(Mz) - FL(0.1382798) * Mz + FL(1.3425);
            rho = FL(-0.0165406) * CUBE(Mz) + FL(0.2290201) * SQ(Mz) - FL(1.1069031) * Mz + FL(3.0455);
        } else {
            vr = FL(-0.0024324) * Mz + FL(1.0019);
            rho = FL(-0.0012973) * Mz + FL(1.1565);
        }
        if(Mz >= FL(-1.0) && Mz < FL(0.0)) alpha2_f = FL(0.4556);
        else if(Mz >= FL(0.0) && Mz < FL(2.6)) alpha2_f = FL(0.4556) + FL(0.0245) * Mz;
        else if(Mz >= FL(2.6) && Mz < FL(4.5)) alpha2_f = FL(0.1978) + FL(0.1245) * Mz;
        else if(Mz >= FL(4.5) && Mz < FL(6.0)) alpha2_f = FL(8.0399) - FL(2.5228) * Mz + FL(0.20098) * SQ(Mz);
        else if(Mz >= FL(6.0) && Mz < FL(9.5)) alpha2_f = FL(0.9431) - FL(0.2041) * Mz + FL(0.0117) * SQ(Mz);
        else alpha2_f = FL(0.0601);
        hs->rho = rho;
        real aR = vr * FL(1500.0);
        real aI = alpha2_f * (vr / FL(1000.0)) * FL(1500.0) * (real)logf(10.0f) / (FL(40.0) * REAL_PI);
        hs->cP = crci(env, freq0, (real)in->zTemp, aR, aI, "L ");
        hs->cS = FL(0.0);
    }
}

/* Boundary preprocessing, src/module/boundary.hpp:56-83, 203-226, 499-544, 600-890 (2D) */
static void bdry_from_env(const bhc_env2d *env, real freq0, const bhc_bdry_in *in, const HS *hs, int isTop,
                          BdryInfo *out)
{
    const real NegTop = isTop ? RL(-1.0) : RL(1.0);
    const real BDRYBIG = RSQRT(REAL_MAX) / RL(1.0e5);
    if(in->npts <= 0) {
        out->type[0] = 'L';
        out->type[1] = 'S';
        out->npts = 2;
        out->bd = xcalloc(2, sizeof(BdryPt));
        out->bd[0].x = V2(-BDRYBIG, hs->depth);
        out->bd[1].x = V2(BDRYBIG, hs->depth);
    } else {
        out->type[0] = in->type[0];
        out->type[1] = in->type[1];
        int n = in->npts + 2;
        out->npts = n;
        out->bd = xcalloc(n, sizeof(BdryPt));
        for(int i = 1; i < n - 1; ++i) {
            out->bd[i].x = V2((real)in->r_km[i - 1], (real)in->z[i - 1]);
            if(in->type[1] == 'L') {
                out->bd[i].cP = crci(env, freq0, RL(1.0e20), (real)in->alphaR[i - 1], (real)in->alphaI[i - 1], "W ");
                out->bd[i].cS = crci(env, freq0, RL(1.0e20), (real)in->betaR[i - 1], (real)in->betaI[i - 1], "W ");
                out->bd[i].rho = (real)in->rho[i - 1];
            }
        }
        out->bd[0].x = V2(-BDRYBIG * RL(0.001), out->bd[1].x.y);
        out->bd[n - 1].x = V2(BDRYBIG * RL(0.001), out->bd[n - 2].x.y);
        out->bd[0].cP = out->bd[1].cP;
        out->bd[0].cS = out->bd[1].cS;
        out->bd[0].rho = out->bd[1].rho;
        out->bd[n - 1].cP = out->bd[n - 2].cP;
        out->bd[n - 1].cS = out->bd[n - 2].cS;
        out->bd[n - 1].rho = out->bd[n - 2].rho;
    }
    int n = out->npts;
    BdryPt *bd = out->bd;
    for(int i = 0; i < n; ++i) bd[i].x.x *= RL(1000.0);
    for(int ii = 0; ii < n - 1; ++ii) {
        bd[ii].t = vsub(bd[ii + 1].x, bd[ii].x);
        bd[ii].Dx = bd[ii].t.y / bd[ii].t.x;
        bd[ii].Len = vlen(bd[ii].t);
        bd[ii].t = vdivs(bd[ii].t, bd[ii].Len);
        bd[ii].n.x = -NegTop * bd[ii].t.y;
        bd[ii].n.y = NegTop * bd[ii].t.x;
    }
    if(out->type[0] == 'C') {
        for(int ii = 1; ii < n - 1; ++ii) {
            real sss = FL(0.5);
            bd[ii].Nodet = vadd(vscl(FL(1.0) - sss, bd[ii - 1].t), vscl(sss, bd[ii].t));
        }
        bd[0].Nodet = V2(FL(1.0), FL(0.0));
        bd[n - 1].Nodet = V2(FL(1.0), FL(0.0));
        for(int ii = 0; ii < n; ++ii) {
            bd[ii].Noden.x = -NegTop * bd[ii].Nodet.y;
            bd[ii].Noden.y = NegTop * bd[ii].Nodet.x;
        }
        real curphi = RATAN2(bd[0].Nodet.y, bd[0].Nodet.x);
        for(int ii = 0; ii < n - 1; ++ii) {
            real nextphi = RATAN2(bd[ii + 1].Nodet.y, bd[ii + 1].Nodet.x);
            bd[ii].kappa = (nextphi - curphi) / bd[ii].Len;
            bd[ii].Dxx = (bd[ii + 1].Dx - bd[ii].Dx) / (bd[ii + 1].x.x - bd[ii].x.x);
            bd[ii].Dss = bd[ii].Dxx * CUBE(bd[ii].t.x);
            bd[ii].kappa = bd[ii].Dss;
            curphi = nextphi;
        }
    } else {
        for(int i = 0; i < n; ++i) bd[i].kappa = FL(0.0);
    }
}

/* SubTab-free vectors are passed already expanded by the caller. */
static int prob_from_env(const bhc_env2d *env, Prob *P)
{
    memset(P, 0, sizeof(*P));
    memcpy(P->RunType, env->RunType, 7);
    /* MoveBeamType, src/module/beaminfo.hpp:241-246 */
    P->Type[0] = env->RunType[1];
    P->Type[1] = env->BeamType[0];
    P->Type[2] = env->BeamType[1];
    P->Type[3] = env->RunType[6];
    P->ssp_type = env->ssp_type;
    P->freq0 = (real)env->freq0;
    /* SSP, src/module/ssp.hpp:453-545 */
    int N = env->ssp_npts;
    P->npts = N;
    P->z = xcalloc(N, sizeof(real));
    P->rho = xcalloc(N, sizeof(real));
    P->c = xcalloc(N, sizeof(cpx));
    P->cz = xcalloc(N, sizeof(cpx));
    P->n2 = xcalloc(N, sizeof(cpx));
    P->n2z = xcalloc(N, sizeof(cpx));
    for(int k = 0; k < 4; ++k) {
        P->coef[k] = xcalloc(N, sizeof(cpx));
        P->work[k] = xcalloc(N, sizeof(cpx));
    }
    for(int i = 0; i < N; ++i) {
        P->z[i] = (real)env->ssp_z[i];
        P->rho[i] = (real)env->ssp_rho[i];
    }
    for(int iz = 0; iz < N; ++iz) {
        P->c[iz] = crci(env, P->freq0, P->z[iz], (real)env->ssp_alphaR[iz], (real)env->ssp_alphaI[iz],
                        env->atten_unit);
        if(iz > 0) P->cz[iz - 1] = (P->c[iz] - P->c[iz - 1]) / (P->z[iz] - P->z[iz - 1]);
    }
    P->cz[N - 1] = NAN + NAN * I;
    switch(env->ssp_type) {
    case 'N':
        for(int i = 0; i < N; ++i) P->n2[i] = (real)FL(1.0) / SQ(P->c[i]);
        for(int iz = 1; iz < N; ++iz) P->n2z[iz - 1] = (P->n2[iz] - P->n2[iz - 1]) / (P->z[iz] - P->z[iz - 1]);
        break;
    case 'S':
        for(int i = 0; i < N; ++i) P->coef[0][i] = P->c[i];
        cspline(P->z, P->coef[0], P->coef[1], P->coef[2], P->coef[3], N, 0, 0);
        break;
    case 'P':
        pchip(P->z, P->c, N, P->coef[0], P->coef[1], P->coef[2], P->coef[3], P->work[0], P->work[1], P->work[2],
              P->work[3]);
        break;
    case 'Q': {
        int Nr = env->ssp_nr;
        P->Nr = Nr;
        P->seg_r = xcalloc(Nr, sizeof(real));
        P->cMat = xcalloc((size_t)N * Nr, sizeof(real));
        P->czMat = xcalloc((size_t)N * Nr, sizeof(real));
        for(int i = 0; i < Nr; ++i) P->seg_r[i] = (real)env->ssp_r_km[i] * FL(1000.0);
        for(int i = 0; i < N * Nr; ++i) P->cMat[i] = (real)env->ssp_cmat[i];
        for(int s = 0; s < Nr; ++s)
            for(int iz2 = 1; iz2 < N; ++iz2) {
                real dz = P->z[iz2] - P->z[iz2 - 1];
                P->czMat[(iz2 - 1) * Nr + s] = (P->cMat[iz2 * Nr + s] - P->cMat[(iz2 - 1) * Nr + s]) / dz;
            }
    } break;
    default: break;
    }
    /* boundary conditions */
    hs_from_env(env, P->freq0, &env->top, &P->top, env->atten_unit);
    hs_from_env(env, P->freq0, &env->bot, &P->bot, env->atten_unit);
    bdry_from_env(env, P->freq0, &env->ati, &P->top, 1, &P->ati);
    bdry_from_env(env, P->freq0, &env->bty, &P->bot, 0, &P->bty);
    /* reflection tables, src/module/reflcoef.hpp:156-167 */
    P->ntrc = env->trc.npts;
    P->nbrc = env->brc.npts;
    P->trc = xcalloc(P->ntrc, sizeof(ReflCoef));
    P->brc = xcalloc(P->nbrc, sizeof(ReflCoef));
    for(int i = 0; i < P->ntrc; ++i) {
        P->trc[i].theta = (real)env->trc.theta[i];
        P->trc[i].r = (real)env->trc.r[i];
        P->trc[i].phi = (real)env->trc.phi[i] * DegRad;
    }
    for(int i = 0; i < P->nbrc; ++i) {
        P->brc[i].theta = (real)env->brc.theta[i];
        P->brc[i].r = (real)env->brc.r[i];
        P->brc[i].phi = (real)env->brc.phi[i] * DegRad;
    }
    /* positions, src/module/rcvrranges.hpp:74-85, szrz.hpp:138-142 */
    P->NSz = env->NSz;
    P->NRz = env->NRz;
    P->NRr = env->NRr;
    P->Sz = xcalloc(P->NSz, sizeof(float));
    P->Rz = xcalloc(P->NRz, sizeof(float));
    P->Rr = xcalloc(P->NRr, sizeof(real));
    memcpy(P->Sz, env->Sz, sizeof(float) * P->NSz);
    memcpy(P->Rz, env->Rz, sizeof(float) * P->NRz);
    for(int i = 0; i < P->NRr; ++i) P->Rr[i] = (real)env->Rr_km[i] * FL(1000.0);
    P->Delta_r = FL(0.0);
    if(P->NRr >= 2) P->Delta_r = P->Rr[P->NRr - 1] - P->Rr[P->NRr - 2];
    P->NRz_per_range = (P->RunType[4] == 'I') ? 1 : P->NRz;
    /* angles, src/module/rayangles.hpp:157-190 */
    P->Nalpha = env->Nalpha;
    P->iSingle = env->iSingle_alpha;
    P->alpha = xcalloc(P->Nalpha, sizeof(real));
    for(int i = 0; i < P->Nalpha; ++i) P->alpha[i] = (real)env->alpha_deg[i] * DegRad;
    P->Dalpha = FL(0.0);
    if(P->Nalpha != 1) P->Dalpha = (P->alpha[P->Nalpha - 1] - P->alpha[0]) / (P->Nalpha - 1);
    /* beam, src/module/beaminfo.hpp:221-238 */
    P->box_r = (real)env->box_r_km * FL(1000.0);
    P->box_z = (real)env->box_z;
    P->deltas = (real)env->deltas;
    if(RABS(P->deltas) < REAL_EPSILON) P->deltas = (P->bot.depth - P->top.depth) / FL(10.0);
    P->epsMultiplier = (real)env->epsMultiplier;
    P->rLoop = (real)env->rLoop;
    P->iBeamWindow = env->iBeamWindow;
    /* source beam pattern, src/module/sbp.hpp:43-53, 143-153 */
    if(env->NSBPPts <= 0) {
        P->nsbp = 2;
        P->sbp = xcalloc(4, sizeof(real));
        P->sbp[0] = FL(-180.0);
        P->sbp[1] = (real)pow(10.0, 0.0 / 20.0);
        P->sbp[2] = FL(180.0);
        P->sbp[3] = P->sbp[1];
    } else {
        P->nsbp = env->NSBPPts;
        P->sbp = xcalloc(2 * P->nsbp, sizeof(real));
        for(int i = 0; i < P->nsbp; ++i) {
            P->sbp[2 * i] = (real)env->sbp[2 * i];
            P->sbp[2 * i + 1] = (real)RPOW(FL(10.0), (real)env->sbp[2 * i + 1] / FL(20.0));
        }
    }
    return 0;
}

static void prob_free(Prob *P)
{
    free(P->z); free(P->rho); free(P->c); free(P->cz); free(P->n2); free(P->n2z);
    for(int k = 0; k < 4; ++k) { free(P->coef[k]); free(P->work[k]); }
    free(P->seg_r); free(P->cMat); free(P->czMat);
    free(P->ati.bd); free(P->bty.bd); free(P->trc); free(P->brc);
    free(P->Sz); free(P->Rz); free(P->Rr); free(P->alpha); free(P->sbp);
}

/* ------------------------------------------------------------------------- */
/* Error state, src/util/errors.hpp:29-51                                      */
/* ------------------------------------------------------------------------- */
typedef struct { _Atomic uint32_t error, warning, errCount, warnCount; } ErrState;
static void RunError(ErrState *e, uint32_t code)
{
    atomic_fetch_add_explicit(&e->errCount, 1u, memory_order_relaxed);
    atomic_fetch_or_explicit(&e->error, 1u << code, memory_order_release);
}
static void RunWarning(ErrState *e, uint32_t code)
{
    atomic_fetch_add_explicit(&e->warnCount, 1u, memory_order_relaxed);
    atomic_fetch_or_explicit(&e->warning, 1u << code, memory_order_relaxed);
}
static int HasErrored(ErrState *e) { return atomic_load_explicit(&e->error, memory_order_acquire) != 0u; }

/* ------------------------------------------------------------------------- */
/* SSP evaluation, src/ssp.hpp                                                 */
/* ------------------------------------------------------------------------- */
typedef struct { v2 gradc; cpx ccpx; real rho, czz, crr, crz; } SSPOut;
typedef struct { int r, z; } SegState;

static void UpdateSSPSegment(real x, real t, const real *arr, int n, int *iSeg) /* ssp.hpp:36-51 */
{
    if(t >= RL(0.0)) {
        while(*iSeg > 0 && x < arr[*iSeg]) --*iSeg;
        while(*iSeg < n - 2 && x >= arr[*iSeg + 1]) ++*iSeg;
    } else {
        while(*iSeg < n - 2 && x > arr[*iSeg + 1]) ++*iSeg;
        while(*iSeg > 0 && x <= arr[*iSeg]) --*iSeg;
    }
}
static real LinInterpDensity(real z, const Prob *P, int iz, real *rho) /* ssp.hpp:53-59 */
{
    real w = (z - P->z[iz]) / (P->z[iz + 1] - P->z[iz]);
    *rho = (RL(1.0) - w) * P->rho[iz] + w * P->rho[iz + 1];
    return w;
}

static void EvaluateSSP(v2 x, v2 t, SSPOut *o, const Prob *P, SegState *iSeg, ErrState *err)
{
    o->crr = o->crz = o->czz = RL(0.0);
    switch(P->ssp_type) {
    case 'N': { /* n2Linear ssp.hpp:64-76 */
        UpdateSSPSegment(x.y, t.y, P->z, P->npts, &iSeg->z);
        real w = LinInterpDensity(x.y, P, iSeg->z, &o->rho);
        o->ccpx = RL(1.0) / CSQRT((RL(1.0) - w) * P->n2[iSeg->z] + w * P->n2[iSeg->z + 1]);
        real c = creal(o->ccpx);
        o->gradc = V2(RL(0.0), RL(-0.5) * CUBE(c) * creal(P->n2z[iSeg->z]));
        o->czz = RL(3.0) * o->gradc.y * o->gradc.y / c;
    } break;
    case 'C': { /* cLinear ssp.hpp:81-89 */
        UpdateSSPSegment(x.y, t.y, P->z, P->npts, &iSeg->z);
        LinInterpDensity(x.y, P, iSeg->z, &o->rho);
        o->ccpx = P->c[iSeg->z] + (x.y - P->z[iSeg->z]) * P->cz[iSeg->z];
        o->gradc = V2(RL(0.0), creal(P->cz[iSeg->z]));
    } break;
    case 'P': { /* cPCHIP ssp.hpp:95-128 */
        UpdateSSPSegment(x.y, t.y, P->z, P->npts, &iSeg->z);
        LinInterpDensity(x.y, P, iSeg->z, &o->rho);
        int iz = iSeg->z;
        real xt = x.y - P->z[iz];
        if(RABS(xt) > RL(1.0e10)) RunWarning(err, BHC_WARN_CPCHIP_INVALIDXT);
        for(int i = 0; i < 4; ++i)
            if(CABS(P->coef[i][iz]) > RL(1.0e10)) RunWarning(err, BHC_WARN_CPCHIP_INVALIDCCOEF);
        o->ccpx = P->coef[0][iz] + (P->coef[1][iz] + (P->coef[2][iz] + P->coef[3][iz] * xt) * xt) * xt;
        o->gradc = V2(RL(0.0), creal(P->coef[1][iz] + (RL(2.0) * P->coef[2][iz] + RL(3.0) * P->coef[3][iz] * xt) * xt));
        o->czz = creal(RL(2.0) * P->coef[2][iz] + RL(6.0) * P->coef[3][iz] * xt);
    } break;
    case 'S': { /* cCubic ssp.hpp:133-152 + SplineALL curves.hpp:190-200 */
        UpdateSSPSegment(x.y, t.y, P->z, P->npts, &iSeg->z);
        LinInterpDensity(x.y, P, iSeg->z, &o->rho);
        int iz = iSeg->z;
        real h = x.y - P->z[iz];
        const float half = FL(0.5), sixth = FL(1.0) / FL(6.0);
        cpx c1 = P->coef[0][iz], c2 = P->coef[1][iz], c3 = P->coef[2][iz], c4 = P->coef[3][iz];
        cpx f = c1 + h * (c2 + h * ((real)half * c3 + ((real)sixth * h) * c4));
        cpx fx = c2 + h * (c3 + (h * (real)half) * c4);
        cpx fxx = c3 + h * c4;
        o->ccpx = f;
        o->gradc = V2(RL(0.0), creal(fx));
        o->czz = creal(fxx);
    } break;
    case 'Q': { /* Quad ssp.hpp:157-214 */
        if(x.x < P->seg_r[0] || x.x > P->seg_r[P->Nr - 1]) RunError(err, BHC_ERR_OUTSIDE_SSP);
        UpdateSSPSegment(x.y, t.y, P->z, P->npts, &iSeg->z);
        UpdateSSPSegment(x.x, t.x, P->seg_r, P->Nr, &iSeg->r);
        LinInterpDensity(x.y, P, iSeg->z, &o->rho);
        if(iSeg->z >= P->npts - 1 || iSeg->r >= P->Nr - 1) {
            RunError(err, BHC_ERR_QUAD_ISEG);
            iSeg->z = iSeg->r = 0;
        }
        int Nr = P->Nr, iz = iSeg->z, ir = iSeg->r;
        real cz1 = P->czMat[iz * Nr + ir], cz2 = P->czMat[iz * Nr + ir + 1];
        real s2 = x.y - P->z[iz];
        real delta_z = P->z[iz + 1] - P->z[iz];
        real c1 = P->cMat[iz * Nr + ir] + s2 * cz1;
        real c2 = P->cMat[iz * Nr + ir + 1] + s2 * cz2;
        real delta_r = P->seg_r[ir + 1] - P->seg_r[ir];
        real s1 = (x.x - P->seg_r[ir]) / delta_r;
        s1 = rmin(s1, RL(1.0));
        s1 = rmax(s1, RL(0.0));
        real c = (RL(1.0) - s1) * c1 + s1 * c2;
        s2 = s2 / delta_z;
        real cim = cimag((RL(1.0) - s2) * P->c[iz] + s2 * P->c[iz + 1]);
        o->ccpx = MKC(c, cim);
        real cz = (RL(1.0) - s1) * cz1 + s1 * cz2;
        real cr = (c2 - c1) / delta_r;
        o->crz = (cz2 - cz1) / delta_r;
        o->gradc = V2(cr, cz);
    } break;
    case 'A': { /* Analytic 2D ssp.hpp:338-371 */
        iSeg->z = 0;
        real c0 = FL(1500.0);
        o->rho = FL(1.0);
        const float unk1 = FL(1300.0), unk2 = FL(0.00737);
        real xt = FL(2.0) * (x.y - unk1) / unk1;
        real emxt = REXP(-xt);
        real DxtDz = FL(2.0) / unk1;
        o->ccpx = c0 * (FL(1.0) + unk2 * (xt - FL(1.0) + emxt));
        real cz = c0 * unk2 * (FL(1.0) - emxt) * DxtDz;
        o->czz = c0 * unk2 * emxt * SQ(DxtDz);
        o->gradc = V2(FL(0.0), cz);
    } break;
    default: RunError(err, BHC_ERR_INVALID_SSP_TYPE); break;
    }
}

/* ------------------------------------------------------------------------- */
/* Boundaries, src/boundary.hpp:66-227 (2D branch)                             */
/* ------------------------------------------------------------------------- */
typedef struct { int Iseg; real lmin, lmax; v2 x, n; } BdryState;
typedef struct { HS top, bot; } Bdry;

static void GetBdrySeg(v2 x, v2 t, BdryState *bds, const BdryInfo *bi, HS *hs, int isTop, int isInit, ErrState *err)
{
    if(!isInit
       && !(x.x < bds->lmin || (x.x == bds->lmin && t.x < FL(0.0)) || x.x > bds->lmax
            || (x.x == bds->lmax && t.x >= FL(0.0))))
        return;
    int n = bi->npts;
    bds->Iseg = imin(imax(bds->Iseg, 0), n - 2);
    if(t.x >= FL(0.0)) {
        while(bds->Iseg >= 0 && bi->bd[bds->Iseg].x.x > x.x) --bds->Iseg;
        while(bds->Iseg >= 0 && bds->Iseg < n - 1 && bi->bd[bds->Iseg + 1].x.x <= x.x) ++bds->Iseg;
    } else {
        while(bds->Iseg < n - 1 && bi->bd[bds->Iseg + 1].x.x < x.x) ++bds->Iseg;
        while(bds->Iseg >= 0 && bds->Iseg < n - 1 && bi->bd[bds->Iseg].x.x >= x.x) --bds->Iseg;
    }
    if(bds->Iseg < 0 || bds->Iseg >= n - 1) {
        RunError(err, isTop ? BHC_ERR_OUTSIDE_ALTIMETRY : BHC_ERR_OUTSIDE_BATHYMETRY);
        bds->Iseg = 0;
    }
    bds->lmin = bi->bd[bds->Iseg].x.x;
    bds->lmax = bi->bd[bds->Iseg + 1].x.x;
    bds->x = bi->bd[bds->Iseg].x;
    bds->n = bi->bd[bds->Iseg].n;
    if(bi->type[1] == 'L') {
        hs->cP = bi->bd[bds->Iseg].cP;
        hs->cS = bi->bd[bds->Iseg].cS;
        hs->rho = bi->bd[bds->Iseg].rho;
    }
}

/* ------------------------------------------------------------------------- */
/* Ray point and step, src/step.hpp                                            */
/* ------------------------------------------------------------------------- */
typedef struct {
    v2 p, q;
    int NumTopBnc, NumBotBnc;
    v2 x, t;
    real c, Amp, Phase;
    cpx tau;
} RayPt;

static int outside_box(v2 x, int dim, const Prob *P) /* common_run.hpp:41-47 */
{
    return dim == 0 ? RABS(x.x - RL(0.0)) >= P->box_r : RABS(x.y - RL(0.0)) >= P->box_z;
}

static void DepthInterfaceCrossing(int stepTo, real *hInt, v2 *x, v2 x0, v2 urayt, const SegState *iSeg0, const Prob *P)
{ /* step.hpp:41-79 */
    if(!stepTo) *hInt = REAL_MAX;
    if(RABS(urayt.y) > REAL_EPSILON) {
        if(P->z[iSeg0->z] > x->y) {
            *hInt = (P->z[iSeg0->z] - x0.y) / urayt.y;
            if(stepTo) {
                *x = vadd(x0, vscl(*hInt, urayt));
                x->y = P->z[iSeg0->z];
            }
        } else if(P->z[iSeg0->z + 1] < x->y) {
            *hInt = (P->z[iSeg0->z + 1] - x0.y) / urayt.y;
            if(stepTo) {
                *x = vadd(x0, vscl(*hInt, urayt));
                x->y = P->z[iSeg0->z + 1];
            }
        }
    }
}

static void BeamBoxCrossing(int stepTo, int dim, real *hBox, v2 *x, v2 x0, v2 urayt, const Prob *P)
{ /* step.hpp:84-108; box centre is (0,0) in 2D */
    if(!stepTo) *hBox = REAL_MAX;
    if(outside_box(*x, dim, P)) {
        real box = dim == 0 ? P->box_r : P->box_z;
        real d0 = (dim == 0 ? x0.x : x0.y) - RL(0.0);
        real u = dim == 0 ? urayt.x : urayt.y;
        *hBox = (box - RABS(d0)) / RABS(u);
        if(stepTo) {
            *x = vadd(x0, vscl(*hBox, urayt));
            real s = RL(0.0) + copysign(box, d0);
            if(dim == 0) x->x = s; else x->y = s;
        }
    }
}

static void TopBotCrossing(int stepTo, real *h, const BdryState *bd, v2 *x, v2 x0, v2 urayt, int *refl)
{ /* step.hpp:113-147 */
    if(!stepTo) *h = REAL_MAX;
    v2 d = vsub(*x, bd->x);
    v2 d0 = vsub(x0, bd->x);
    real w = vdot(bd->n, d);
    if(stepTo ? (w > -INFINITESIMAL_STEP_SIZE) : (w >= RL(0.0))) {
        *h = -vdot(d0, bd->n) / vdot(urayt, bd->n);
        if(stepTo) {
            *x = vadd(x0, vscl(*h, urayt));
            if(RABS(bd->n.x) < REAL_EPSILON) x->y = bd->x.y;
        }
        *refl = 1;
    } else {
        *refl = 0;
    }
}

static void TopBotSegCrossing(int stepTo, real *hSeg, const BdryState *top, const BdryState *bot, int iSegr, v2 *x,
                              v2 x0, v2 urayt, const Prob *P, int *topRefl, int *botRefl)
{ /* step.hpp:153-205 (range) */
    real smin = rmax(top->lmin, bot->lmin);
    real smax = rmin(top->lmax, bot->lmax);
    if(P->ssp_type == 'Q') {
        smin = rmax(smin, P->seg_r[iSegr]);
        smax = rmin(smax, P->seg_r[iSegr + 1]);
    }
    if(!stepTo) *hSeg = REAL_MAX;
    if(RABS(urayt.x) > REAL_EPSILON) {
        if(x->x < smin) {
            *hSeg = -(x0.x - smin) / urayt.x;
            if(stepTo) {
                *x = vadd(x0, vscl(*hSeg, urayt));
                x->x = smin;
                *topRefl = *botRefl = 0;
            }
        } else if(x->x > smax) {
            *hSeg = -(x0.x - smax) / urayt.x;
            if(stepTo) {
                *x = vadd(x0, vscl(*hSeg, urayt));
                x->x = smax;
                *topRefl = *botRefl = 0;
            }
        }
    }
}

static void ReduceStep(v2 x0, v2 urayt, const SegState *iSeg0, BdryState *top, BdryState *bot, const Prob *P,
                       ErrState *err, real *h, int *iSmallStepCtr)
{ /* step.hpp:271-342 */
    real hInt, hBoxxr, hBoxyz, hBoxz_, hTop, hBot, hxSeg, hySeg, hTopDiag, hBotDiag;
    int dummy;
    v2 x = vadd(x0, vscl(*h, urayt));
    DepthInterfaceCrossing(0, &hInt, &x, x0, urayt, iSeg0, P);
    BeamBoxCrossing(0, 0, &hBoxxr, &x, x0, urayt, P);
    BeamBoxCrossing(0, 1, &hBoxyz, &x, x0, urayt, P);
    hBoxz_ = REAL_MAX;
    TopBotCrossing(0, &hTop, top, &x, x0, urayt, &dummy);
    TopBotCrossing(0, &hBot, bot, &x, x0, urayt, &dummy);
    TopBotSegCrossing(0, &hxSeg, top, bot, iSeg0->r, &x, x0, urayt, P, &dummy, &dummy);
    hySeg = hTopDiag = hBotDiag = REAL_MAX;
    *h = rmin(rmin(rmin(*h, hInt), rmin(rmin(rmin(hBoxxr, hBoxyz), hBoxz_), rmin(hTop, hBot))),
              rmin(rmin(hxSeg, hySeg), rmin(hTopDiag, hBotDiag)));
    if(*h < RL(-1e-4)) RunWarning(err, BHC_WARN_STEP_NEGATIVE_H);
    if(*h < INFINITESIMAL_STEP_SIZE * P->deltas) {
        *h = INFINITESIMAL_STEP_SIZE * P->deltas;
        ++*iSmallStepCtr;
    } else {
        *iSmallStepCtr = 0;
    }
}

static void StepToBdry(v2 x0, v2 *x2, v2 urayt, real *h, int *topRefl, int *botRefl, const SegState *iSeg0,
                       BdryState *top, BdryState *bot, const Prob *P)
{ /* step.hpp:347-424 */
    *x2 = vadd(x0, vscl(*h, urayt));
    DepthInterfaceCrossing(1, h, x2, x0, urayt, iSeg0, P);
    BeamBoxCrossing(1, 0, h, x2, x0, urayt, P);
    BeamBoxCrossing(1, 1, h, x2, x0, urayt, P);
    TopBotCrossing(1, h, top, x2, x0, urayt, topRefl);
    TopBotCrossing(1, h, bot, x2, x0, urayt, botRefl);
    if(*botRefl) *topRefl = 0;
    TopBotSegCrossing(1, h, top, bot, iSeg0->r, x2, x0, urayt, P, topRefl, botRefl);
    if(*h < INFINITESIMAL_STEP_SIZE * P->deltas) {
        *h = INFINITESIMAL_STEP_SIZE * P->deltas;
        *x2 = vadd(x0, vscl(*h, urayt));
        v2 d = vsub(*x2, top->x);
        if(vdot(top->n, d) > REAL_EPSILON) *topRefl = 1; else *topRefl = 0;
        d = vsub(*x2, bot->x);
        if(vdot(bot->n, d) > REAL_EPSILON) {
            *botRefl = 1;
            *topRefl = 0;
        } else {
            *botRefl = 0;
        }
    }
}

static void Step(RayPt ray0, RayPt *ray2, BdryState *top, BdryState *bot, const Prob *P, SegState *iSeg,
                 ErrState *err, int *iSmallStepCtr, int *topRefl, int *botRefl)
{ /* step.hpp:654-811 (2D) */
    RayPt ray1;
    SSPOut o0, o1, o2;
    EvaluateSSP(ray0.x, ray0.t, &o0, P, iSeg, err);
    real c0 = creal(o0.ccpx);
    real cnn0 = o0.crr * SQ(ray0.t.y) - FL(2.0) * o0.crz * ray0.t.x * ray0.t.y + o0.czz * SQ(ray0.t.x);
    v2 pq0p = vscl(-cnn0, ray0.q), pq0q = vscl(c0, ray0.p);
    SegState iSeg0 = *iSeg;
    real csq0 = SQ(c0);
    v2 urayt0 = vscl(c0, ray0.t);
    real h = P->deltas;
    ReduceStep(ray0.x, urayt0, &iSeg0, top, bot, P, err, &h, iSmallStepCtr);
    real halfh = FL(0.5) * h;
    ray1.x = vadd(ray0.x, vscl(halfh, urayt0));
    ray1.t = vsub(ray0.t, vdivs(vscl(halfh, o0.gradc), csq0));
    ray1.p = vadd(ray0.p, vscl(halfh, pq0p));
    ray1.q = vadd(ray0.q, vscl(halfh, pq0q));

    EvaluateSSP(ray1.x, ray1.t, &o1, P, iSeg, err);
    real c1 = creal(o1.ccpx);
    real cnn1 = o1.crr * SQ(ray1.t.y) - FL(2.0) * o1.crz * ray1.t.x * ray1.t.y + o1.czz * SQ(ray1.t.x);
    v2 pq1p = vscl(-cnn1, ray1.q), pq1q = vscl(c1, ray1.p);
    real csq1 = SQ(c1);
    v2 urayt1 = vscl(c1, ray1.t);
    ReduceStep(ray0.x, urayt1, &iSeg0, top, bot, P, err, &h, iSmallStepCtr);

    real w1 = h / (RL(2.0) * halfh);
    real w0 = RL(1.0) - w1;
    v2 urayt2 = vadd(vscl(w0, urayt0), vscl(w1, urayt1));
    v2 x2;
    StepToBdry(ray0.x, &x2, urayt2, &h, topRefl, botRefl, &iSeg0, top, bot, P);
    ray2->x = x2;
    real hw0 = h * w0, hw1 = h * w1;
    ray2->t = vsub(vsub(ray0.t, vdivs(vscl(hw0, o0.gradc), csq0)), vdivs(vscl(hw1, o1.gradc), csq1));
    ray2->tau = ray0.tau + hw0 / o0.ccpx + hw1 / o1.ccpx;
    ray2->p = vadd(ray0.p, vscl(hw0, pq0p));
    ray2->q = vadd(ray0.q, vscl(hw0, pq0q));
    ray2->p = vadd(ray2->p, vscl(hw1, pq1p));
    ray2->q = vadd(ray2->q, vscl(hw1, pq1q));
    ray2->Amp = ray0.Amp;
    ray2->Phase = ray0.Phase;
    ray2->NumTopBnc = ray0.NumTopBnc;
    ray2->NumBotBnc = ray0.NumBotBnc;

    EvaluateSSP(ray2->x, ray2->t, &o2, P, iSeg, err);
    ray2->c = creal(o2.ccpx);
    if(iSeg->z != iSeg0.z || iSeg->r != iSeg0.r) { /* CurvatureCorrection 2D, step.hpp:626-648 */
        v2 gj = vsub(o2.gradc, o0.gradc);
        v2 ray2n = V2(-ray2->t.y, ray2->t.x);
        real cnjump = vdot(gj, ray2n);
        real csjump = vdot(gj, ray2->t);
        real rm, rn;
        if(iSeg->z != iSeg0.z) rm = ray2->t.x / ray2->t.y;
        else rm = -ray2->t.y / ray2->t.x;
        rn = rm * (FL(2.0) * cnjump - rm * csjump) / ray2->c;
        ray2->p = vsub(ray2->p, vscl(rn, ray2->q));
    }
}

/* ------------------------------------------------------------------------- */
/* Reflection, src/reflect.hpp (2D)                                            */
/* ------------------------------------------------------------------------- */
static void InterpolateReflectionCoefficient(ReflCoef *RInt, const ReflCoef *tab, int n, ErrState *err)
{ /* reflect.hpp:39-92 */
    int iLeft = 0, iRight = n - 1;
    real th = RInt->theta;
    if(th < tab[iLeft].theta) {
        RInt->r = FL(0.0);
        RInt->phi = FL(0.0);
        RunWarning(err, BHC_WARN_OUTSIDE_REFLCOEF);
    } else if(th > tab[iRight].theta) {
        RInt->r = FL(0.0);
        RInt->phi = FL(0.0);
    } else {
        while(iLeft != iRight - 1) {
            int iMid = (iLeft + iRight) / 2;
            if(tab[iMid].theta > th) iRight = iMid; else iLeft = iMid;
        }
        real alpha = (RInt->theta - tab[iLeft].theta) / (tab[iRight].theta - tab[iLeft].theta);
        RInt->r = (FL(1.0) - alpha) * tab[iLeft].r + alpha * tab[iRight].r;
        RInt->phi = (FL(1.0) - alpha) * tab[iLeft].phi + alpha * tab[iRight].phi;
    }
}

static void Reflect(const RayPt *oldp, RayPt *newp, const HS *hs, int isTop, v2 tBdry, v2 nBdry, real kappa,
                    const Prob *P, SegState *iSeg, ErrState *err)
{ /* reflect.hpp:125-452 (2D) */
    real Th = vdot(oldp->t, nBdry);
    real Tg = vdot(oldp->t, tBdry);
    newp->NumTopBnc = oldp->NumTopBnc + isTop;
    newp->NumBotBnc = oldp->NumBotBnc + (1 - isTop);
    newp->x = oldp->x;
    newp->t = vsub(oldp->t, vscl(FL(2.0) * Th, nBdry));
    SSPOut o;
    EvaluateSSP(oldp->x, oldp->t, &o, P, iSeg, err);
    real c = creal(o.ccpx);
    newp->c = c;
    newp->tau = oldp->tau;
    v2 rayt = vscl(c, oldp->t);
    v2 rayn = V2(-rayt.y, rayt.x);
    v2 rayt_t = vscl(c, newp->t);
    v2 rayn_t = V2(-(-rayt_t.y), -(rayt_t.x));
    real rn = FL(2.0) * kappa / SQ(c) / Th;
    real cnjump = -vdot(o.gradc, vsub(rayn_t, rayn));
    real csjump = -vdot(o.gradc, vsub(rayt_t, rayt));
    if(isTop) {
        cnjump = -cnjump;
        rn = -rn;
    }
    real rm = Tg / Th;
    rn += rm * (FL(2.0) * cnjump - rm * csjump) / SQ(c);
    if(P->Type[2] == 'D') rn *= FL(2.0);
    else if(P->Type[2] == 'Z') rn = FL(0.0);
    newp->p = vadd(oldp->p, vscl(rn, oldp->q));
    newp->q = oldp->q;

    if(hs->bc == 'R') {
        newp->Amp = oldp->Amp;
        newp->Phase = oldp->Phase;
    } else if(hs->bc == 'V') {
        newp->Amp = oldp->Amp;
        newp->Phase = oldp->Phase + REAL_PI;
    } else if(hs->bc == 'F') {
        ReflCoef RInt;
        RInt.theta = RadDeg * RABS(RATAN2(Th, Tg));
        if(RInt.theta > FL(90.0)) RInt.theta = FL(180.0) - RInt.theta;
        if(isTop) InterpolateReflectionCoefficient(&RInt, P->trc, P->ntrc, err);
        else InterpolateReflectionCoefficient(&RInt, P->brc, P->nbrc, err);
        newp->Amp = oldp->Amp * RInt.r;
        newp->Phase = oldp->Phase + RInt.phi;
    } else if(hs->bc == 'A' || hs->bc == 'G') {
        real omega = FL(2.0) * REAL_PI * P->freq0;
        cpx kx = omega * Tg;
        cpx kz = omega * Th;
        cpx kx2 = kx * kx;
        cpx f, g;
        if(creal(hs->cS) > FL(0.0)) {
            cpx kzS2 = kx2 - SQ(omega / hs->cS);
            cpx kzP2 = kx2 - SQ(omega / hs->cP);
            cpx kzS = CSQRT(kzS2);
            cpx kzP = CSQRT(kzP2);
            cpx mu = hs->rho * SQ(hs->cS);
            cpx y2 = (SQ(kzS2 + kx2) - RL(4.0) * kzS * kzP * kx2) * mu;
            cpx y4 = kzP * (kx2 - kzS2);
            f = SQ(omega) * y4;
            g = y2;
        } else {
            cpx kzP = CSQRT(kx2 - SQ(omega / hs->cP));
            if(RABS(creal(kzP)) < REAL_EPSILON && cimag(kzP) < RL(0.0)) kzP = -kzP;
            f = kzP;
            g = hs->rho;
        }
        const cpx J = I;
        cpx Refl = -(o.rho * f - J * kz * g) / (o.rho * f + J * kz * g);
        real aR = CABS(Refl);
        if(aR < FL(1.0e-5)) {
            newp->Amp = FL(0.0);
            newp->Phase = oldp->Phase;
        } else {
            newp->Amp = aR * oldp->Amp;
            newp->Phase = oldp->Phase + RATAN2(cimag(Refl), creal(Refl));
            if(P->Type[3] == 'S') {
                /* beam displacement and width change, reflect.hpp:383-442 (2D) */
                cpx ch = oldp->c / CCONJ(hs->cP);
                real co = oldp->t.x * oldp->c;
                real si = oldp->t.y * oldp->c;
                real ck = omega / oldp->c;
                cpx a = FL(2.0) * hs->rho * (FL(1.0) - SQ(ch));
                cpx b = SQ(co) - SQ(ch);
                cpx d = SQ(hs->rho) * SQ(si) + b;
                cpx sb = CSQRT(b);
                real cco = SQ(co);
                real ssi = SQ(si);
                cpx delta;
                if(RABS(si) >= REAL_EPSILON) delta = a * co / si / (ck * sb * d);
                else delta = FL(0.0);
                real pdelta = creal(delta) / (oldp->c / co);
                cpx cksbd = ck * sb * d;
                cpx ddelta = a * (cco / (cksbd * b) - (RL(1.0) + (cco / ssi)) / cksbd
                                  - FL(2.0) * SQ(co) * (SQ(hs->rho) - RL(1.0)) / (cksbd * d));
                real rddelta = -creal(ddelta);
                real sddelta = rddelta / RABS(rddelta);
                real theta_bot = RATAN(tBdry.y / tBdry.x);
                newp->x.x = newp->x.x + creal(delta) * RCOS(theta_bot);
                newp->x.y = newp->x.y + creal(delta) * RSIN(theta_bot);
                newp->tau = newp->tau + pdelta;
                newp->q = vadd(newp->q, vscl(sddelta * rddelta * si * creal(o.ccpx), oldp->p));
            }
        }
    } else {
        RunError(err, BHC_ERR_BOUNDARY_CONDITION_TYPE);
        newp->Amp = NAN;
        newp->Phase = NAN;
    }
}

/* ------------------------------------------------------------------------- */
/* Ray init / update / terminate, src/trace.hpp                                */
/* ------------------------------------------------------------------------- */
typedef struct {
    int isz, ialpha;
    real alpha, SrcDeclAngle;
    /* Nx2D rays (nx = 1): isz above is the field's source slot
     * ((isz3 * NSx + isx) * NSy + isy) * Ntheta + ibeta, GetFieldAddr
     * (common.hpp:393-405) with itheta = ibeta (influence.hpp:439) */
    int nx, isx, isy, isz3, ibeta;
    real SrcAzimAngle;
} RayInitInfo;

typedef struct {
    SegState iSeg;
    BdryState top, bot;
    Bdry bdry;
    v2 xs;
    real DistBegTop, DistBegBot, DistEndTop, DistEndBot;
    int iSmallStepCtr;
} TraceState;

static int BinarySearchLEQ(const real *arr, int n, int stride, int offset, real target)
{ /* common_run.hpp:61-80 */
    int low = 0, hi = n;
    while(low < hi) {
        int t = (low + hi) / 2;
        if(arr[t * stride + offset] > target) hi = t;
        else if(t >= n - 1) return n - 1;
        else if(arr[(t + 1) * stride + offset] > target) return t;
        else low = t + 1;
    }
    return low;
}
static int BinarySearchGT(const real *arr, int n, real target)
{ /* common_run.hpp:119-138 */
    int low = -1, hi = n - 1;
    while(low < hi) {
        int t = (low + hi + 1) / 2;
        if(arr[t] <= target) low = t;
        else if(t <= 0) return 0;
        else if(arr[t - 1] <= target) return t;
        else hi = t - 1;
    }
    return hi;
}

static void Distances(v2 rayx, v2 Topx, v2 Botx, v2 Topn, v2 Botn, real *dt, real *db)
{ /* trace.hpp:37-45 */
    *dt = -vdot(Topn, vsub(rayx, Topx));
    *db = -vdot(Botn, vsub(rayx, Botx));
}

static int RayInit(RayInitInfo *ri, RayPt *p0, TraceState *S, const Prob *P, ErrState *err)
{ /* trace.hpp:76-266 (2D) */
    if(ri->isz < 0 || ri->isz >= P->NSz || ri->ialpha < 0 || ri->ialpha >= P->Nalpha) {
        RunError(err, BHC_ERR_RAYINIT);
        return 0;
    }
    ri->alpha = P->alpha[ri->ialpha];
    ri->SrcDeclAngle = RadDeg * ri->alpha;
    S->iSeg.r = S->iSeg.z = 0;
    S->xs = V2(FL(0.0), (real)P->Sz[ri->isz]);
    v2 tinit = V2(RCOS(ri->alpha), RSIN(ri->alpha));
    SSPOut o;
    EvaluateSSP(S->xs, tinit, &o, P, &S->iSeg, err);
    real c = creal(o.ccpx);
    { /* too few beams */
        real DalphaOpt = RSQRT(c / (FL(6.0) * P->freq0 * P->Rr[P->NRr - 1]));
        int NalphaOpt = 2 + (int)((P->alpha[P->Nalpha - 1] - P->alpha[0]) / DalphaOpt);
        if(P->RunType[0] == 'C' && P->Nalpha < NalphaOpt && ri->ialpha == 0) RunWarning(err, BHC_WARN_TOO_FEW_BEAMS);
    }
    int ibp = BinarySearchLEQ(P->sbp, P->nsbp, 2, 0, ri->SrcDeclAngle);
    ibp = imin(ibp, P->nsbp - 2);
    real s = (ri->SrcDeclAngle - P->sbp[2 * ibp]) / (P->sbp[2 * (ibp + 1)] - P->sbp[2 * ibp]);
    float Amp0 = (float)((FL(1.0) - s) * P->sbp[2 * ibp + 1] + s * P->sbp[2 * (ibp + 1) + 1]);
    if(P->RunType[0] == 'S') {
        float omega = (float)(FL(2.0) * REAL_PI * P->freq0);
        Amp0 = (float)(Amp0 * (sqrtf(FL(2.0)) * RABS(RSIN(omega / c * S->xs.y * RSIN(ri->alpha)))));
    }
    p0->x = S->xs;
    p0->c = c;
    p0->t = vdivs(tinit, c);
    p0->tau = 0;
    p0->Amp = Amp0;
    p0->Phase = FL(0.0);
    p0->NumTopBnc = p0->NumBotBnc = 0;
    p0->p = V2(FL(1.0), FL(0.0));
    p0->q = V2(FL(0.0), FL(1.0));
    if(P->RunType[1] == 'G') p0->q = V2(FL(0.0), FL(0.0));
    S->bdry.top = P->top;
    S->bdry.bot = P->bot;
    S->top.Iseg = S->bot.Iseg = 0;
    GetBdrySeg(S->xs, p0->t, &S->top, &P->ati, &S->bdry.top, 1, 1, err);
    GetBdrySeg(S->xs, p0->t, &S->bot, &P->bty, &S->bdry.bot, 0, 1, err);
    Distances(S->xs, S->top.x, S->bot.x, S->top.n, S->bot.n, &S->DistBegTop, &S->DistBegBot);
    if(S->DistBegTop <= FL(0.0) || S->DistBegBot <= FL(0.0)) {
        RunWarning(err, BHC_WARN_SOURCE_OUTSIDE_BOUNDARIES);
        return 0;
    }
    S->iSmallStepCtr = 0;
    return 1;
}

static int RayUpdate(const RayPt *p0, RayPt *p1, RayPt *p2, TraceState *S, const Prob *P, ErrState *err)
{ /* trace.hpp:274-381 (2D) */
    int topRefl = 0, botRefl = 0;
    Step(*p0, p1, &S->top, &S->bot, P, &S->iSeg, err, &S->iSmallStepCtr, &topRefl, &botRefl);
    GetBdrySeg(p1->x, p1->t, &S->top, &P->ati, &S->bdry.top, 1, 0, err);
    GetBdrySeg(p1->x, p1->t, &S->bot, &P->bty, &S->bdry.bot, 0, 0, err);
    Distances(p1->x, S->top.x, S->bot.x, S->top.n, S->bot.n, &S->DistEndTop, &S->DistEndBot);
    if(topRefl || botRefl) {
        const BdryInfo *bdi = topRefl ? &P->ati : &P->bty;
        const BdryState *bs = topRefl ? &S->top : &S->bot;
        const HS *hs = topRefl ? &S->bdry.top : &S->bdry.bot;
        const BdryPt *bd0 = &bdi->bd[bs->Iseg];
        const BdryPt *bd1 = bd0 + 1;
        v2 nInt, tInt;
        if(bdi->type[0] == 'C') {
            real sss = vdot(vsub(p1->x, bs->x), bd0->t) / bd0->Len;
            nInt = vadd(vscl(FL(1.0) - sss, bd0->Noden), vscl(sss, bd1->Noden));
            tInt = vadd(vscl(FL(1.0) - sss, bd0->Nodet), vscl(sss, bd1->Nodet));
        } else {
            nInt = bd0->n;
            tInt = bd0->t;
        }
        Reflect(p1, p2, hs, topRefl, tInt, nInt, bd0->kappa, P, &S->iSeg, err);
        Distances(p2->x, S->top.x, S->bot.x, S->top.n, S->bot.n, &S->DistEndTop, &S->DistEndBot);
        return 1;
    }
    return 0;
}

static int RayTerminate(const RayPt *pt, int *Nsteps, int is, TraceState *S, int MaxPointsPerRay, const Prob *P,
                        ErrState *err)
{ /* trace.hpp:391-477 (2D) */
    int leftbox = outside_box(pt->x, 0, P) || outside_box(pt->x, 1, P);
    int escaped = (S->DistBegTop < FL(0.0) && S->DistEndTop < FL(0.0))
        || (S->DistBegBot < FL(0.0) && S->DistEndBot < FL(0.0));
    int lostenergy = pt->Amp < FL(0.005);
    if(leftbox || lostenergy || escaped) {
        *Nsteps = is + 1;
        return 1;
    } else if(is >= MaxPointsPerRay - 3) {
        RunWarning(err, BHC_WARN_ONERAY_OUTOFMEMORY);
        *Nsteps = is;
        return 1;
    }
    S->DistBegTop = S->DistEndTop;
    S->DistBegBot = S->DistEndBot;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Influence, src/influence.hpp (geometric, 2D)                                */
/* ------------------------------------------------------------------------- */
typedef struct {
    RayInitInfo init;
    real c0, freq0, omega, RadMax, BeamWindow, Ratio1, rcp_q0, Dalpha;
    real phase, qOld;
    v2 rayn1, x;
    int lastValid;
    int ir;
} InflRay;

typedef struct {
    const Prob *P;
    float *field;
    bhc_arrival *arr;
    _Atomic int32_t *narr;
    int max_narr, allow_merging;
    bhc_eigen_hit *hits;
    _Atomic int32_t *neigen;
    int eigen_memsize;
    _Atomic int64_t contrib;
    int isGaussian, isRayCen;
} Outs;

static void atomic_add_float(float *addr, float v)
{ /* CAS float add, like src/util/atomics.hpp:65-100 */
    _Atomic uint32_t *a = (_Atomic uint32_t *)addr;
    uint32_t old = atomic_load_explicit(a, memory_order_relaxed), nw;
    do {
        float f;
        memcpy(&f, &old, 4);
        f += v;
        memcpy(&nw, &f, 4);
    } while(!atomic_compare_exchange_weak_explicit(a, &old, nw, memory_order_relaxed, memory_order_relaxed));
}

static int IsAtCaustic(const InflRay *ir, real q)
{
    return (q <= RL(0.0) && ir->qOld > RL(0.0)) || (q >= RL(0.0) && ir->qOld < RL(0.0));
}

static void AddArrAt(bhc_arrival *baseArr, _Atomic int32_t *baseNArr, int max_narr, int allow_merging, real Amp,
                     real omega, real Phase, cpx delay, real SrcDeclAngle, real SrcAzimAngle, real RcvrDeclAngle,
                     real RcvrAzimAngle, int NumTopBnc, int NumBotBnc)
{ /* arrivals.hpp:45-130 on one receiver's slots */
    if(allow_merging) {
        int Nt = atomic_load(baseNArr);
        const float PhaseTol = FL(0.05);
        int second = Nt >= 1
            && omega * CABS(delay - (cpx)(baseArr[Nt - 1].delay[0] + baseArr[Nt - 1].delay[1] * I)) < PhaseTol
            && RABS(baseArr[Nt - 1].Phase - Phase) < PhaseTol;
        if(!second) {
            int iArr;
            if(Nt >= max_narr) {
                iArr = -1;
                real weakest = Amp;
                for(Nt = 0; Nt < max_narr; ++Nt) {
                    if(baseArr[Nt].a < weakest) {
                        weakest = baseArr[Nt].a;
                        iArr = Nt;
                    }
                }
                if(iArr < 0) return;
            } else {
                iArr = Nt;
                atomic_store(baseNArr, Nt + 1);
            }
            baseArr[iArr].a = (float)Amp;
            baseArr[iArr].Phase = (float)Phase;
            baseArr[iArr].delay[0] = (float)creal(delay);
            baseArr[iArr].delay[1] = (float)cimag(delay);
            baseArr[iArr].SrcDeclAngle = (float)SrcDeclAngle;
            baseArr[iArr].SrcAzimAngle = (float)SrcAzimAngle;
            baseArr[iArr].RcvrDeclAngle = (float)RcvrDeclAngle;
            baseArr[iArr].RcvrAzimAngle = (float)RcvrAzimAngle;
            baseArr[iArr].NTopBnc = NumTopBnc;
            baseArr[iArr].NBotBnc = NumBotBnc;
        } else {
            float AmpTot = baseArr[Nt - 1].a + (float)Amp;
            float w1 = baseArr[Nt - 1].a / AmpTot;
            float w2 = (float)Amp / AmpTot;
            float dr = (float)creal(delay), di = (float)cimag(delay);
            baseArr[Nt - 1].delay[0] = w1 * baseArr[Nt - 1].delay[0] + w2 * dr;
            baseArr[Nt - 1].delay[1] = w1 * baseArr[Nt - 1].delay[1] + w2 * di;
            baseArr[Nt - 1].a = AmpTot;
            baseArr[Nt - 1].SrcDeclAngle = w1 * baseArr[Nt - 1].SrcDeclAngle + w2 * (float)SrcDeclAngle;
            baseArr[Nt - 1].SrcAzimAngle = w1 * baseArr[Nt - 1].SrcAzimAngle + w2 * (float)SrcAzimAngle;
            baseArr[Nt - 1].RcvrDeclAngle = w1 * baseArr[Nt - 1].RcvrDeclAngle + w2 * (float)RcvrDeclAngle;
            baseArr[Nt - 1].RcvrAzimAngle = w1 * baseArr[Nt - 1].RcvrAzimAngle + w2 * (float)RcvrAzimAngle;
        }
    } else {
        int Nt = atomic_fetch_add(baseNArr, 1);
        if(Nt >= max_narr) return;
        baseArr[Nt].a = (float)Amp;
        baseArr[Nt].Phase = (float)Phase;
        baseArr[Nt].delay[0] = (float)creal(delay);
        baseArr[Nt].delay[1] = (float)cimag(delay);
        baseArr[Nt].SrcDeclAngle = (float)SrcDeclAngle;
        baseArr[Nt].SrcAzimAngle = (float)SrcAzimAngle;
        baseArr[Nt].RcvrDeclAngle = (float)RcvrDeclAngle;
        baseArr[Nt].RcvrAzimAngle = (float)RcvrAzimAngle;
        baseArr[Nt].NTopBnc = NumTopBnc;
        baseArr[Nt].NBotBnc = NumBotBnc;
    }
}

/* Checker hook: a log of the raw AddArr calls of a single-threaded run, in
 * call order, 10 doubles each (receiver index, Amp, Phase, delay re / im,
 * omega, source and receiver declination, top / bottom bounces), so that a
 * test can replay the merge rules on the unrounded values. */
static double *g_arrlog;
static int64_t g_arrlog_cap, g_arrlog_n;

int SFX(oracle_arrlog)(double *buf, int64_t cap)
{
    g_arrlog = buf;
    g_arrlog_cap = cap;
    g_arrlog_n = 0;
    return 0;
}

int64_t SFX(oracle_arrlog_count)(void) { return g_arrlog_n; }

static void AddArr(int id, int ir, real Amp, real omega, real Phase, cpx delay, const RayInitInfo *ri,
                   real RcvrDeclAngle, int NumTopBnc, int NumBotBnc, Outs *O)
{ /* arrivals.hpp:45-130, 2D: no azimuths */
    const Prob *P = O->P;
    size_t base = ((size_t)ri->isz * P->NRz_per_range + id) * P->NRr + ir;
    if(g_arrlog) {
        if(g_arrlog_n < g_arrlog_cap) {
            double *e = g_arrlog + 10 * g_arrlog_n;
            e[0] = (double)base;
            e[1] = (double)Amp;
            e[2] = (double)Phase;
            e[3] = (double)creal(delay);
            e[4] = (double)cimag(delay);
            e[5] = (double)omega;
            e[6] = (double)ri->SrcDeclAngle;
            e[7] = (double)RcvrDeclAngle;
            e[8] = (double)NumTopBnc;
            e[9] = (double)NumBotBnc;
        }
        ++g_arrlog_n;
    }
    /* Nx2D: RcvrAzimAngle = SrcAzimAngle (influence.hpp:122-135) */
    real az = ri->nx ? ri->SrcAzimAngle : (real)NAN;
    AddArrAt(&O->arr[base * O->max_narr], &O->narr[base], O->max_narr, O->allow_merging, Amp, omega, Phase, delay,
             ri->SrcDeclAngle, az, RcvrDeclAngle, az, NumTopBnc, NumBotBnc);
}

static void RecordEigenHit(int ir, int iz, int is, const RayInitInfo *ri, Outs *O)
{ /* eigenrays.hpp:24-41 */
    int mi = atomic_fetch_add(O->neigen, 1);
    if(mi >= O->eigen_memsize) return;
    bhc_eigen_hit *h = &O->hits[mi];
    h->is = is; h->iz = iz; h->ir = ir; h->itheta = ri->nx ? ri->ibeta : 0;
    h->isx = ri->isx; h->isy = ri->isy; h->isz = ri->nx ? ri->isz3 : ri->isz; h->ialpha = ri->ialpha;
    h->ibeta = ri->ibeta;
}

static void ApplyContribution(real cnst, real w, real omega, cpx delay, real phaseInt, real RcvrDeclAngle, int ir,
                              int iz, int is, const InflRay *inf, const RayPt *p1, Outs *O)
{ /* influence.hpp:432-473 */
    const Prob *P = O->P;
    char rt = P->RunType[0];
    atomic_fetch_add_explicit(&O->contrib, 1, memory_order_relaxed);
    if(rt == 'E') {
        RecordEigenHit(ir, iz, is, &inf->init, O);
    } else if(rt == 'A' || rt == 'a' || rt == 'V' || rt == 'v') {
        AddArr(iz, ir, cnst * w, omega, phaseInt, delay, &inf->init, RcvrDeclAngle, p1->NumTopBnc, p1->NumBotBnc, O);
        if(rt == 'V' || rt == 'v') RecordEigenHit(ir, iz, is, &inf->init, O);
    } else {
        float fr, fi;
        if(rt == 'C') {
            cpx e = CEXP(-I * (omega * delay - phaseInt));
            cpx d = (cnst * w) * e;
            fr = (float)creal(d);
            fi = (float)cimag(d);
        } else {
            real v = cnst * REXP(cimag(omega * delay));
            v = SQ(v) * w;
            if(O->isGaussian) v *= RSQRT(FL(2.0) * REAL_PI);
            fr = (float)v;
            fi = 0.0f;
        }
        size_t base = (((size_t)inf->init.isz) * P->NRz_per_range + iz) * P->NRr + ir;
        atomic_add_float(&O->field[2 * base], fr);
        atomic_add_float(&O->field[2 * base + 1], fi);
    }
}

static void InfluenceGeoCore(real s, real n1, v2 dq, cpx dtau, int ir, int iz, int is, const RayPt *p0,
                             const RayPt *p1, real RcvrDeclAngle, const InflRay *inf, Outs *O)
{ /* influence.hpp:897-1024 (2D) */
    v2 qInterp = vadd(p0->q, vscl(s, dq));
    real qFinal = qInterp.x;
    real sigma, sigma_orig;
    sigma = sigma_orig = RABS(qFinal * inf->rcp_q0);
    if(O->isGaussian) {
        real lambda = p0->c / inf->freq0;
        sigma = rmax(sigma, rmin(FL(0.2) * inf->freq0 * creal(p1->tau), REAL_PI * lambda));
    }
    real bcd = n1;
    if(bcd > inf->BeamWindow * sigma || bcd == inf->BeamWindow * sigma) return;
    cpx delay = p0->tau + s * dtau;
    real cfactor = RSQRT(p1->c);
    real cnst = inf->Ratio1 * cfactor * p1->Amp / RSQRT(RABS(qFinal));
    real w;
    if(O->isGaussian) w = REXP(FL(-0.5) * SQ(n1 / sigma)) * (sigma_orig / sigma);
    else w = (sigma - n1) / sigma;
    const RayPt *pp = O->isGaussian ? p1 : p0;
    real phaseInt = pp->Phase + inf->phase;
    if(IsAtCaustic(inf, qFinal)) phaseInt = inf->phase + REAL_PI / FL(2.0);
    ApplyContribution(cnst, w, inf->omega, delay, phaseInt, RcvrDeclAngle, ir, iz, is, inf, p1, O);
}

static int RToIR(real r, const Prob *P)
{ /* influence.hpp:82-89 */
    real temp = (r - P->Rr[0]) / P->Delta_r;
    if(RABS(temp - RROUND(temp)) < RL(1e-6)) temp = RROUND(temp);
    return imax(imin((int)temp, P->NRr - 1), 0);
}

static void Step_InfluenceGeoRayCen(const RayPt *p0, const RayPt *p1, InflRay *inf, int is, Outs *O)
{ /* influence.hpp:1030-1194 (2D) */
    const Prob *P = O->P;
    real phaseq = p0->q.x;
    if(IsAtCaustic(inf, phaseq)) inf->phase += REAL_PI / FL(2.0);
    inf->qOld = phaseq;
    v2 dq = vsub(p1->q, p0->q);
    cpx dtau = p1->tau - p0->tau;
    v2 rayn1 = vscl(p1->c, V2(p1->t.y, -p1->t.x));
    if(RABS(rayn1.y) < RL(1e-10)) return;
    if(RABS(p1->x.x - p0->x.x) < RL(1.0e3) * spacing(p1->x.x)) {
        inf->lastValid = 1;
        inf->x = p1->x;
        inf->rayn1 = rayn1;
        return;
    }
    real RcvrDeclAngle = RadDeg * RATAN2(p1->t.y, p1->t.x);
    for(int iz = 0; iz < P->NRz_per_range; ++iz) {
        real zR = (real)P->Rz[iz];
        real nA, rA, nB, rB;
        int irA, irB;
        if(!inf->lastValid) {
            nA = RL(1e10);
            rA = RL(1e10);
            irA = 0;
        } else {
            nA = (zR - inf->x.y) / inf->rayn1.y;
            rA = inf->x.x + nA * inf->rayn1.x;
            irA = RToIR(rA, P);
        }
        nB = (zR - p1->x.y) / rayn1.y;
        rB = p1->x.x + nB * rayn1.x;
        irB = RToIR(rB, P);
        if(irA == irB) continue;
        int notii = (irB <= irA) ? 0 : 1;
        for(int ir = irA + notii; ir != irB + notii; ir += (notii << 1) - 1) {
            real s = (P->Rr[ir] - rA) / (rB - rA);
            real n1 = RABS(nA + s * (nB - nA));
            InfluenceGeoCore(s, n1, dq, dtau, ir, iz, is, p0, p1, RcvrDeclAngle, inf, O);
        }
    }
    inf->lastValid = 1;
    inf->x = p1->x;
    inf->rayn1 = rayn1;
}

static void Step_InfluenceGeoCart(const RayPt *p0, const RayPt *p1, InflRay *inf, int is, Outs *O)
{ /* influence.hpp:1201-1406 (2D) */
    const Prob *P = O->P;
    real rA = inf->x.x, rB = p1->x.x;
    v2 x_ray = p0->x;
    v2 rayt = vsub(p1->x, p0->x);
    real rlen = vlen(rayt);
    if(rlen < RL(1.0e3) * spacing(p1->x.x)) return;
    rayt = vdivs(rayt, rlen);
    v2 rayn1 = V2(-rayt.y, rayt.x);
    real RcvrDeclAngle = RadDeg * RATAN2(rayt.y, rayt.x);
    v2 dq = vsub(p1->q, p0->q);
    cpx dtau = p1->tau - p0->tau;
    real phaseq = p0->q.x;
    if(IsAtCaustic(inf, phaseq)) inf->phase += REAL_PI / FL(2.0);
    inf->qOld = phaseq;
    real sigma = rmax(RABS(p0->q.x), RABS(p1->q.x)) * inf->rcp_q0 / RABS(rayt.x);
    if(O->isGaussian) {
        real lambda = p0->c / inf->freq0;
        sigma = rmax(sigma, rmin(FL(0.2) * inf->freq0 * creal(p1->tau), REAL_PI * lambda));
    }
    real RadiusMax = inf->BeamWindow * sigma;
    real zmin, zmax;
    if(RABS(rayt.x) > FL(0.50001)) {
        zmin = rmin(p0->x.y, p1->x.y) - RadiusMax;
        zmax = rmax(p0->x.y, p1->x.y) + RadiusMax;
    } else {
        zmin = -REAL_MAX;
        zmax = REAL_MAX;
    }
    for(;;) {
        if(P->Rr[inf->ir] >= rmin(rA, rB) && P->Rr[inf->ir] < rmax(rA, rB)) {
            v2 xr;
            xr.x = P->Rr[inf->ir];
            for(int iz = 0; iz < P->NRz_per_range; ++iz) {
                int tempiz = (P->RunType[4] == 'I') ? inf->ir : iz;
                xr.y = (real)P->Rz[tempiz];
                if(xr.y < zmin || xr.y > zmax) continue;
                v2 xrr = vsub(xr, x_ray);
                real s = vdot(xrr, rayt) / rlen;
                real n1 = RABS(vdot(xrr, rayn1));
                InfluenceGeoCore(s, n1, dq, dtau, inf->ir, iz, is, p0, p1, RcvrDeclAngle, inf, O);
            }
        }
        int irTT;
        if(rB > P->Rr[inf->ir]) {
            if(inf->ir >= P->NRr - 1) break;
            irTT = inf->ir + 1;
            if(P->Rr[irTT] >= rB) break;
        } else {
            if(inf->ir <= 0) break;
            irTT = inf->ir - 1;
            if(P->Rr[irTT] <= rB) break;
        }
        inf->ir = irTT;
    }
    inf->x = p1->x;
}

static void Init_Influence(InflRay *inf, const RayPt *p0, const RayInitInfo *ri, const Prob *P, int isGaussian,
                           int isRayCen)
{ /* influence.hpp:525-644 (geometric 2D) */
    inf->init = *ri;
    inf->freq0 = P->freq0;
    inf->omega = FL(2.0) * REAL_PI * inf->freq0;
    inf->c0 = p0->c;
    inf->RadMax = FL(50.0) * p0->c / P->freq0;
    inf->Dalpha = P->Dalpha;
    inf->BeamWindow = isGaussian ? RL(4.0) : RL(1.0);
    inf->Ratio1 = RSQRT(RABS(RCOS(ri->alpha)));
    if(P->RunType[3] == 'X') inf->Ratio1 = RL(1.0);
    if(isGaussian) inf->Ratio1 /= RSQRT(FL(2.0) * REAL_PI);
    inf->rcp_q0 = inf->Dalpha / p0->c;
    inf->phase = FL(0.0);
    inf->qOld = p0->q.x;
    if(isRayCen) {
        inf->rayn1 = vscl(p0->c, V2(p0->t.y, -p0->t.x));
        inf->x = p0->x;
        inf->lastValid = RABS(inf->rayn1.y) >= RL(1e-6);
    } else {
        inf->x = p0->x;
        inf->ir = BinarySearchGT(P->Rr, P->NRr, p0->x.x);
        if(p0->t.x < RL(0.0) && inf->ir > 0) --inf->ir;
    }
}

/* ------------------------------------------------------------------------- */
/* Per-ray drivers                                                             */
/* ------------------------------------------------------------------------- */
typedef struct {
    int32_t nsteps, ntop, nbot;
    double xend[2];
} RayStat;

static void MainFieldModes(RayInitInfo *ri, Outs *O, ErrState *err, RayStat *st)
{ /* trace.hpp:525-589 */
    const Prob *P = O->P;
    TraceState S;
    RayPt p0, p1, p2;
    InflRay inf;
    memset(&p2, 0, sizeof(p2));
    if(st) { st->nsteps = 0; st->ntop = st->nbot = 0; st->xend[0] = st->xend[1] = 0; }
    if(!RayInit(ri, &p0, &S, P, err)) return;
    Init_Influence(&inf, &p0, ri, P, O->isGaussian, O->isRayCen);
    int is = 0, Nsteps = 0;
    for(;;) {
        if(HasErrored(err)) break;
        int two = RayUpdate(&p0, &p1, &p2, &S, P, err);
        if(O->isRayCen) Step_InfluenceGeoRayCen(&p0, &p1, &inf, is, O);
        else Step_InfluenceGeoCart(&p0, &p1, &inf, is, O);
        ++is;
        if(two) {
            if(O->isRayCen) Step_InfluenceGeoRayCen(&p1, &p2, &inf, is, O);
            else Step_InfluenceGeoCart(&p1, &p2, &inf, is, O);
            p0 = p2;
            ++is;
        } else {
            p0 = p1;
        }
        if(RayTerminate(&p0, &Nsteps, is, &S, MaxN, P, err)) break;
    }
    if(st) {
        st->nsteps = is;
        st->ntop = p0.NumTopBnc;
        st->nbot = p0.NumBotBnc;
        st->xend[0] = (double)p0.x.x;
        st->xend[1] = (double)p0.x.y;
    }
}

static void store_pt(bhc_ray_pt2d *d, const RayPt *s)
{
    d->p[0] = s->p.x; d->p[1] = s->p.y; d->q[0] = s->q.x; d->q[1] = s->q.y;
    d->NumTopBnc = s->NumTopBnc; d->NumBotBnc = s->NumBotBnc;
    d->x[0] = s->x.x; d->x[1] = s->x.y; d->t[0] = s->t.x; d->t[1] = s->t.y;
    d->c = s->c; d->Amp = s->Amp; d->Phase = s->Phase;
    d->tau[0] = creal(s->tau); d->tau[1] = cimag(s->tau);
}

static void MainRayMode(RayInitInfo *ri, RayPt *ray, int *Nsteps, int MaxPointsPerRay, const Prob *P, ErrState *err)
{ /* trace.hpp:482-520 */
    TraceState S;
    if(!RayInit(ri, &ray[0], &S, P, err)) {
        *Nsteps = 1;
        return;
    }
    int is = 0;
    for(;;) {
        if(HasErrored(err)) break;
        int two = RayUpdate(&ray[is], &ray[is + 1], &ray[is + 2], &S, P, err);
        if(*Nsteps >= 0 && is >= *Nsteps) {
            *Nsteps = is + 2;
            break;
        }
        is += two ? 2 : 1;
        if(RayTerminate(&ray[is], Nsteps, is, &S, MaxPointsPerRay, P, err)) break;
    }
}

/* ------------------------------------------------------------------------- */
/* Thread pool (src/mode/fieldimpl.cpp.in:28-77): atomic job counter          */
/* ------------------------------------------------------------------------- */
typedef struct {
    Outs *O;
    ErrState *err;
    _Atomic int32_t next;
    int32_t njobs;
    int32_t job_begin, job_end;
    bhc_field_out *out;
    RayStat *stats;
} FieldJob;

static int num_jobs(const Prob *P) /* common.hpp:347-359 (2D) */
{
    int ret = 1;
    if(P->iSingle == 0) ret *= P->Nalpha;
    ret *= P->NSz;
    return ret;
}
static int job_indices(RayInitInfo *ri, int job, const Prob *P) /* common.hpp:365-391 (2D) */
{
    memset(ri, 0, sizeof(*ri));
    if(P->iSingle >= 1) {
        ri->ialpha = P->iSingle - 1;
    } else {
        ri->ialpha = job % P->Nalpha;
        job /= P->Nalpha;
    }
    ri->isz = job;
    return ri->isz < P->NSz;
}

static void *field_worker(void *arg)
{
    FieldJob *J = (FieldJob *)arg;
    for(;;) {
        int32_t job = atomic_fetch_add(&J->next, 1);
        if(job >= J->job_end) break;
        RayInitInfo ri;
        if(!job_indices(&ri, job, J->O->P)) break;
        MainFieldModes(&ri, J->O, J->err, J->stats ? &J->stats[job - J->job_begin] : NULL);
    }
    return NULL;
}

static int resolve_threads(int n)
{
    if(n > 0) return n;
    long c = sysconf(_SC_NPROCESSORS_ONLN);
    return c > 0 ? (int)c : 1;
}

/* Exported: full field run (TL / arrivals / eigen) on a contiguous job range
 * [job_begin, job_end) (job_end < 0 means all). Outputs accumulate into `out`. */
int SFX(oracle_field_run2d)(const bhc_env2d *env, bhc_field_out *out, int nthreads, int job_begin, int job_end)
{
    Prob P;
    if(prob_from_env(env, &P)) return -1;
    Outs O;
    memset(&O, 0, sizeof(O));
    O.P = &P;
    O.field = out->field;
    O.arr = out->arr;
    O.narr = (_Atomic int32_t *)out->narr;
    O.max_narr = out->max_narr;
    O.allow_merging = out->allow_merging;
    O.hits = out->hits;
    O.neigen = (_Atomic int32_t *)out->neigen;
    O.eigen_memsize = out->eigen_memsize;
    O.isGaussian = (P.Type[0] == 'B' || P.Type[0] == 'b');
    O.isRayCen = (P.Type[0] == 'g' || P.Type[0] == 'b');
    atomic_init(&O.contrib, 0);
    ErrState err;
    atomic_init(&err.error, 0u);
    atomic_init(&err.warning, 0u);
    atomic_init(&err.errCount, 0u);
    atomic_init(&err.warnCount, 0u);
    int nj = num_jobs(&P);
    if(job_end < 0 || job_end > nj) job_end = nj;
    if(job_begin < 0) job_begin = 0;
    int nrays = job_end - job_begin;
    RayStat *stats = NULL;
    if(out->ray_nsteps) stats = xcalloc(nrays > 0 ? nrays : 1, sizeof(RayStat));
    FieldJob J;
    J.O = &O;
    J.err = &err;
    atomic_init(&J.next, job_begin);
    J.job_begin = job_begin;
    J.job_end = job_end;
    J.stats = stats;
    if(O.allow_merging) nthreads = 1; /* src/mode/arr.hpp:79 */
    nthreads = resolve_threads(nthreads);
    pthread_t *th = xcalloc(nthreads, sizeof(pthread_t));
    for(int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, field_worker, &J);
    for(int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th);
    if(stats) {
        for(int i = 0; i < nrays; ++i) {
            out->ray_nsteps[i] = stats[i].nsteps;
            if(out->ray_ntop) out->ray_ntop[i] = stats[i].ntop;
            if(out->ray_nbot) out->ray_nbot[i] = stats[i].nbot;
            if(out->ray_xend) { out->ray_xend[2 * i] = stats[i].xend[0]; out->ray_xend[2 * i + 1] = stats[i].xend[1]; }
        }
        int64_t tot = 0;
        for(int i = 0; i < nrays; ++i) tot += stats[i].nsteps;
        out->total_steps = tot;
        free(stats);
    }
    out->total_contrib = atomic_load(&O.contrib);
    out->err = atomic_load(&err.error);
    out->warn = atomic_load(&err.warning);
    out->err_count = atomic_load(&err.errCount);
    out->warn_count = atomic_load(&err.warnCount);
    prob_free(&P);
    return 0;
}

/* Exported: ray mode (and eigenray re-trace when nsteps_in != NULL). */
int SFX(oracle_ray_run2d)(const bhc_env2d *env, bhc_ray_out *out, const int32_t *ialpha_list,
                          const int32_t *isz_list, const int32_t *nsteps_in, int nrays)
{
    Prob P;
    if(prob_from_env(env, &P)) return -1;
    ErrState err;
    atomic_init(&err.error, 0u);
    atomic_init(&err.warning, 0u);
    atomic_init(&err.errCount, 0u);
    atomic_init(&err.warnCount, 0u);
    int M = out->max_points_per_ray;
    RayPt *buf = xcalloc((size_t)M + 3, sizeof(RayPt));
    for(int k = 0; k < nrays; ++k) {
        RayInitInfo ri;
        memset(&ri, 0, sizeof(ri));
        if(ialpha_list) {
            ri.ialpha = ialpha_list[k];
            ri.isz = isz_list ? isz_list[k] : 0;
        } else if(!job_indices(&ri, k, &P)) {
            break;
        }
        int Nsteps = nsteps_in ? nsteps_in[k] : -1;
        memset(buf, 0, sizeof(RayPt) * ((size_t)M + 3));
        MainRayMode(&ri, buf, &Nsteps, M, &P, &err);
        out->nsteps[k] = Nsteps;
        int n = Nsteps < M ? Nsteps : M;
        for(int i = 0; i < n; ++i) store_pt(&out->points[(size_t)k * M + i], &buf[i]);
    }
    free(buf);
    out->err = atomic_load(&err.error);
    out->warn = atomic_load(&err.warning);
    out->err_count = atomic_load(&err.errCount);
    out->warn_count = atomic_load(&err.warnCount);
    prob_free(&P);
    return 0;
}

/* Exported: ScalePressure for 2D, src/influence.hpp:1556-1627 + src/mode/tl.cpp:106-180. */
int SFX(oracle_scale_pressure2d)(const bhc_env2d *env, float *field)
{
    Prob P;
    if(prob_from_env(env, &P)) return -1;
    ErrState err;
    atomic_init(&err.error, 0u);
    atomic_init(&err.warning, 0u);
    atomic_init(&err.errCount, 0u);
    atomic_init(&err.warnCount, 0u);
    int coherent = P.RunType[0] == 'C';
    int line = P.RunType[3] == 'X';
    for(int isz = 0; isz < P.NSz; ++isz) {
        SegState iSeg = {0, 0};
        SSPOut o;
        v2 xs = V2(FL(0.0), (real)P.Sz[isz]);
        v2 tinit = V2(RCOS(FL(0.0)), RSIN(FL(0.0)));
        EvaluateSSP(xs, tinit, &o, &P, &iSeg, &err);
        real c = creal(o.ccpx);
        (void)c;
        real cnst = FL(-1.0);
        float *u = &field[2 * (size_t)isz * P.NRz_per_range * P.NRr];
        if(!coherent) {
            for(size_t k = 0; k < (size_t)P.NRz_per_range * P.NRr; ++k) {
                u[2 * k] = sqrtf(u[2 * k]);
                u[2 * k + 1] = 0.0f;
            }
        }
        for(int ir = 0; ir < P.NRr; ++ir) {
            real factor;
            if(line) {
                const float local_pi = 3.14159265f;
                factor = FL(-4.0) * sqrtf(local_pi) * cnst;
            } else if(RABS(P.Rr[ir]) < REAL_EPSILON) {
                factor = RL(0.0);
            } else {
                factor = cnst / (real)RSQRT(RABS(P.Rr[ir]));
            }
            for(int irz = 0; irz < P.NRz_per_range; ++irz) {
                size_t a = (size_t)irz * P.NRr + ir;
                float fr = u[2 * a], fi = u[2 * a + 1];
                /* cpxf *= float: (fr*f, fi*f) */
                u[2 * a] = fr * (float)factor;
                u[2 * a + 1] = fi * (float)factor;
            }
        }
    }
    prob_free(&P);
    return 0;
}

/* Exported: dump of the oracle's preprocessed SSP / boundary / receiver arrays,
 * so tests can pin the product's preprocessing against it. Layout:
 * c[2N], cz[2N], coef[4][2N], n2[2N], n2z[2N], rho[N], Rr[NRr], alpha[Nalpha],
 * bty nodes [npts][13]: x,t,n,Noden,Nodet (10), Len, kappa, (pad 0),
 * scalars: deltas, box_r, Dalpha, Delta_r, bot cP re/im, bot cS re/im. */
int SFX(oracle_preprocess_dump2d)(const bhc_env2d *env, double *dst, int64_t cap, int64_t *used)
{
    Prob P;
    if(prob_from_env(env, &P)) return -1;
    int64_t k = 0;
#define PUT(v) do { if(k < cap) dst[k] = (double)(v); ++k; } while(0)
    int N = P.npts;
    for(int i = 0; i < N; ++i) { PUT(creal(P.c[i])); PUT(cimag(P.c[i])); }
    for(int i = 0; i < N; ++i) { PUT(creal(P.cz[i])); PUT(cimag(P.cz[i])); }
    for(int q = 0; q < 4; ++q)
        for(int i = 0; i < N; ++i) { PUT(creal(P.coef[q][i])); PUT(cimag(P.coef[q][i])); }
    for(int i = 0; i < N; ++i) { PUT(creal(P.n2[i])); PUT(cimag(P.n2[i])); }
    for(int i = 0; i < N; ++i) { PUT(creal(P.n2z[i])); PUT(cimag(P.n2z[i])); }
    for(int i = 0; i < N; ++i) PUT(P.rho[i]);
    for(int i = 0; i < P.NRr; ++i) PUT(P.Rr[i]);
    for(int i = 0; i < P.Nalpha; ++i) PUT(P.alpha[i]);
    for(int i = 0; i < P.bty.npts; ++i) {
        BdryPt *b = &P.bty.bd[i];
        PUT(b->x.x); PUT(b->x.y); PUT(b->t.x); PUT(b->t.y); PUT(b->n.x); PUT(b->n.y);
        PUT(b->Noden.x); PUT(b->Noden.y); PUT(b->Nodet.x); PUT(b->Nodet.y);
        PUT(b->Len); PUT(b->kappa); PUT(0.0);
    }
    PUT(P.deltas); PUT(P.box_r); PUT(P.Dalpha); PUT(P.Delta_r);
    PUT(creal(P.bot.cP)); PUT(cimag(P.bot.cP)); PUT(creal(P.bot.cS)); PUT(cimag(P.bot.cS));
#undef PUT
    *used = k;
    prob_free(&P);
    return 0;
}
