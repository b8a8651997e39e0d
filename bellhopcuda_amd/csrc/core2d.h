// core2d.h -- per-ray physics of the 2D tracer, written once for the HIP
// kernels (device) and for the CPU self-check harness (host).
//
// The functions follow the reference's 2D hot path operation-for-operation
// (order of floating-point operations, float-vs-real literal types, tie
// breaks), because trajectories are compared bit-for-bit against the oracle.
// The build uses -ffp-contract=off so that no a*b+c is fused.
//
// Data layout: a problem is one contiguous "blob" of reals (SSP knots and
// coefficients, boundary nodes, receiver ranges, reflection tables) addressed
// through the offsets in Prob2D. The kernel stages the SSP and boundary part
// of the blob into LDS and points the same accessors at the LDS copy.
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>
#include <cmath>
#include <cstddef>
#include <cstdint>

namespace bhc2d {

// ---------------------------------------------------------------------------
// Literals and constants (reference src/common.hpp:144-179)
// ---------------------------------------------------------------------------
template<typename T> struct RC;
template<> struct RC<double> {
    static constexpr double EPS = DBL_EPSILON, MAXV = DBL_MAX, MINPOS = DBL_MIN;
    static constexpr double PI = M_PI, INF_STEP = 1e-6;
};
template<> struct RC<float> {
    static constexpr float EPS = FLT_EPSILON, MAXV = FLT_MAX, MINPOS = FLT_MIN;
    static constexpr float PI = (float)M_PI, INF_STEP = 1e-3f;
};
template<typename T> __host__ __device__ constexpr T rlit(double d, float f);
template<> __host__ __device__ constexpr double rlit<double>(double d, float) { return d; }
template<> __host__ __device__ constexpr float rlit<float>(double, float f) { return f; }
#define RL(a) (::bhc2d::rlit<real>(a, a##f))
#define FL(a) (a##f)

template<typename real> __host__ __device__ inline real rmin(real a, real b) { return (b < a) ? b : a; }
template<typename real> __host__ __device__ inline real rmax(real a, real b) { return (a < b) ? b : a; }
__host__ __device__ inline int imin(int a, int b) { return (b < a) ? b : a; }
__host__ __device__ inline int imax(int a, int b) { return (a < b) ? b : a; }

// (int)x with x86-64 cvttsd2si semantics: out-of-range and NaN give INT_MIN.
template<typename real> __host__ __device__ inline int trunc_int_x86(real x)
{
    if(!(x > (real)-2147483649.0 && x < (real)2147483648.0)) return (int)0x80000000;
    return (int)x;
}

template<typename real> __host__ __device__ inline real spacing(real v)
{
    return std::fabs(std::nextafter(v, (real)0) - v);
}

// ---------------------------------------------------------------------------
// Complex arithmetic as gcc lowers std::complex / C99 complex (no contraction)
// ---------------------------------------------------------------------------
template<typename real> struct cpx {
    real re, im;
};
template<typename real> __host__ __device__ inline cpx<real> C(real a, real b) { return cpx<real>{a, b}; }
template<typename real> __host__ __device__ inline cpx<real> operator+(cpx<real> a, cpx<real> b)
{
    return C(a.re + b.re, a.im + b.im);
}
template<typename real> __host__ __device__ inline cpx<real> operator-(cpx<real> a, cpx<real> b)
{
    return C(a.re - b.re, a.im - b.im);
}
template<typename real> __host__ __device__ inline cpx<real> operator-(cpx<real> a) { return C(-a.re, -a.im); }
template<typename real> __host__ __device__ inline cpx<real> cmul(cpx<real> a, cpx<real> b)
{
    return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template<typename real> __host__ __device__ inline cpx<real> rmul(real r, cpx<real> b) { return C(r * b.re, r * b.im); }
template<typename real> __host__ __device__ inline cpx<real> rdivc(cpx<real> a, real r) { return C(a.re / r, a.im / r); }
// libgcc __divdc3 (Smith's method), finite inputs.
template<typename real> __host__ __device__ inline cpx<real> cdiv(cpx<real> n, cpx<real> dd)
{
    real a = n.re, b = n.im, c = dd.re, d = dd.im, ratio, denom, x, y;
    if(std::fabs(c) < std::fabs(d)) {
        ratio = c / d;
        denom = (c * ratio) + d;
        x = ((a * ratio) + b) / denom;
        y = ((b * ratio) - a) / denom;
    } else {
        ratio = d / c;
        denom = (d * ratio) + c;
        x = ((b * ratio) + a) / denom;
        y = (b - (a * ratio)) / denom;
    }
    return C(x, y);
}
template<typename real> __host__ __device__ inline cpx<real> realdiv(real r, cpx<real> dd)
{
    return cdiv(C(r, (real)0), dd);
}
template<typename real> __host__ __device__ inline real cabs_(cpx<real> a) { return std::hypot(a.re, a.im); }
// glibc csqrt, finite-input branches.
template<typename real> __host__ __device__ inline cpx<real> csqrt_(cpx<real> x)
{
    real re = x.re, im = x.im;
    if(im == (real)0) {
        if(re < (real)0) return C((real)0, std::copysign(std::sqrt(-re), im));
        return C(std::fabs(std::sqrt(re)), std::copysign((real)0, im));
    }
    if(re == (real)0) {
        real r;
        if(std::fabs(im) >= (real)2 * RC<real>::MINPOS) r = std::sqrt((real)M_SQRT1_2 * std::fabs(im));
        else r = (real)0.5 * std::sqrt((real)2 * std::fabs(im));
        return C(r, std::copysign(r, im));
    }
    real d = std::hypot(re, im), r, s;
    if(re > (real)0) {
        r = std::sqrt((real)0.5 * (d + re));
        s = (real)0.5 * (im / r);
    } else {
        s = std::sqrt((real)0.5 * (d - re));
        r = std::fabs((real)0.5 * (im / s));
    }
    return C(r, std::copysign(s, im));
}
// glibc cexp, finite-input branch.
template<typename real> __host__ __device__ inline cpx<real> cexp_(cpx<real> x)
{
    real s, c;
    if(std::fabs(x.im) > RC<real>::MINPOS) {
        s = std::sin(x.im);
        c = std::cos(x.im);
    } else {
        s = x.im;
        c = (real)1;
    }
    real e = std::exp(x.re);
    return C(e * c, e * s);
}

template<typename real> struct v2 {
    real x, y;
};
template<typename real> __host__ __device__ inline v2<real> V(real x, real y) { return v2<real>{x, y}; }
template<typename real> __host__ __device__ inline v2<real> operator+(v2<real> a, v2<real> b) { return V(a.x + b.x, a.y + b.y); }
template<typename real> __host__ __device__ inline v2<real> operator-(v2<real> a, v2<real> b) { return V(a.x - b.x, a.y - b.y); }
template<typename real> __host__ __device__ inline v2<real> sc(real s, v2<real> a) { return V(s * a.x, s * a.y); }
template<typename real> __host__ __device__ inline v2<real> dv(v2<real> a, real s) { return V(a.x / s, a.y / s); }
template<typename real> __host__ __device__ inline real dot(v2<real> a, v2<real> b) { return a.x * b.x + a.y * b.y; }

// ---------------------------------------------------------------------------
// Problem descriptor (device view). All offsets index `base` (reals).
// ---------------------------------------------------------------------------
enum : int { RT_TL = 0, RT_ARR = 1, RT_EIG = 2, RT_RAY = 3 };

// RunType[0] of a TL run: coherent, semi-coherent or incoherent
__host__ __device__ inline bool tl_run(char run0) { return run0 == 'C' || run0 == 'S' || run0 == 'I'; }

// Boundary node record in the blob (BdryPtFull<false> subset). Straight-segment
// boundaries with one set of geoacoustics use the first NODE_SHORT reals of a
// node; curvilinear ('C') or per-node geoacoustics ('L') use all NODE_FULL.
enum : int {
    BN_XX = 0, BN_XY, BN_TX, BN_TY, BN_NX, BN_NY, BN_LEN, BN_KAPPA, NODE_SHORT,
    BN_NDNX = NODE_SHORT, BN_NDNY, BN_NDTX, BN_NDTY, BN_CPR, BN_CPI, BN_CSR, BN_CSI, BN_RHO, BN_PAD, NODE_FULL
};

template<typename real> struct HSd {
    cpx<real> cP, cS;
    real rho;
    char bc;
};

template<typename real> struct Prob2D {
    const real *base;      // blob: staged part first (SSP, boundaries), then the rest
    int32_t staged_reals;  // length of the prefix that may be copied into LDS
    // SSP
    char ssp_type;
    char run0;             // RunType[0]
    char run1;             // RunType[1]
    char run3, run4;       // RunType[3] ('X' line source), RunType[4] ('I' irregular)
    char type2;            // Beam->Type[2] curvature option
    char type3;            // Beam->Type[3]: 'S' beam shift on halfspace reflection
    char ati_curv, bty_curv; // 'C' curvilinear
    char ati_long, bty_long; // 'L' long format (per-node geoacoustics)
    int32_t npts, Nr;
    int32_t o_z, o_rho, o_c, o_cz, o_n2, o_n2z, o_coef, o_segr, o_cmat, o_czmat;
    int32_t ati_n, bty_n, o_ati, o_bty, node;  // node: reals per boundary node
    int32_t o_rzs;         // receiver depths as reals, at the end of the staged prefix
    // rest of blob (global)
    int32_t o_Rr, o_trc, o_brc, ntrc, nbrc, o_sbp, nsbp;
    int32_t o_alpha, o_acos, o_asin;
    const float *Rz;
    const float *Sz;
    int32_t NSz, NRz, NRr, NRz_per_range, Nalpha, iSingle;
    real Delta_r, Dalpha, freq0, deltas, box_r, box_z;
    HSd<real> top, bot;
    int32_t isGaussian, isRayCen;
    int32_t rz_sorted;     // receiver depths strictly increasing (enables depth windows)
    double rz_rdz;         // (NRz - 1) / (Rz[NRz-1] - Rz[0]) when sorted, else 0 (depth-window index guess)
    float *field;          // this problem's complex field (TL runs)
    int32_t njobs;         // rays of this problem
    int32_t pad;
    int32_t *ray_stats;    // batched runs: per-ray steps, top / bottom bounces (or null)
    double *ray_xend;      // batched runs: per-ray final (r, z) (or null)
    // Nx2D runs (nx_ntheta > 0): the source slot isz of a segment is the field
    // slot ((isz3 * NSx + isx) * NSy + isy) * Ntheta + itheta of its bearing
    // plane; o_beta: the bearings (rad) for the arrivals' azimuths
    int32_t nx_ntheta, nx_nsx, nx_nsy, o_beta;
};

// ---------------------------------------------------------------------------
// Ray state
// ---------------------------------------------------------------------------
template<typename real> struct RayPt {
    v2<real> p, q, x, t;
    real c, Amp, Phase;
    cpx<real> tau;
    int32_t NumTopBnc, NumBotBnc;
};

template<typename real> struct BdryState {
    int32_t Iseg;
    real lmin, lmax;
    v2<real> x, n;
};

template<typename real> struct SSPOut {
    v2<real> gradc;
    cpx<real> c;
    real rho, czz, crr, crz;
};

struct Err {
    uint32_t err, warn, nerr, nwarn;
    __host__ __device__ void error(int code)
    {
        err |= 1u << code;
        ++nerr;
    }
    __host__ __device__ void warning(int code)
    {
        warn |= 1u << code;
        ++nwarn;
    }
};

template<typename real> struct Trace {
    int32_t iz, ir;             // SSP segment state
    BdryState<real> top, bot;
    HSd<real> htop, hbot;       // per-segment geoacoustics (type[1]=='L')
    real xs_y;
    real DistBegTop, DistBegBot, DistEndTop, DistEndBot;
};

template<typename real> struct Infl {
    real alpha, SrcDeclAngle, omega, BeamWindow, Ratio1, rcp_q0;
    real phase, qOld;
    v2<real> rayn1, x;
    v2<real> rbr;               // Cartesian: receiver ranges Rr[ir], Rr[ir + 1] (the walk's bracket)
    int32_t lastValid, ir;
};

// The walk's bracket (Rr[ir], Rr[ir + 1]) after ir is set; Rr[ir] twice at the last range.
template<typename real> __host__ __device__ inline v2<real> range_bracket(const real *Rr, int n, int ir)
{
    return V(Rr[ir], ir < n - 1 ? Rr[ir + 1] : Rr[ir]);
}

// ---------------------------------------------------------------------------
// SSP evaluation (reference src/ssp.hpp)
// ---------------------------------------------------------------------------
template<typename real>
__host__ __device__ inline void UpdateSSPSegment(real x, real t, const real *arr, int n, int &iSeg)
{ // ssp.hpp:36-51
    if(t >= RL(0.0)) {
        while(iSeg > 0 && x < arr[iSeg]) --iSeg;
        while(iSeg < n - 2 && x >= arr[iSeg + 1]) ++iSeg;
    } else {
        while(iSeg < n - 2 && x > arr[iSeg + 1]) ++iSeg;
        while(iSeg > 0 && x <= arr[iSeg]) --iSeg;
    }
}

template<typename real> __host__ __device__ inline cpx<real> ldc(const real *p, int i) { return C(p[2 * i], p[2 * i + 1]); }

// ST: SSP type fixed at compile time ('C','S','P','N','Q','A'), or 0 to read it
// from the problem at run time.
template<typename real, char ST = 0>
__host__ __device__ inline void EvaluateSSP(v2<real> x, v2<real> t, SSPOut<real> &o, const Prob2D<real> &P,
                                            const real *B, int &iz, int &ir, Err &e)
{
    const char sst = ST ? ST : P.ssp_type;
    const real *z = B + P.o_z;
    o.crr = o.crz = o.czz = RL(0.0);
    if(sst != 'A' && sst != 'Q') {
        UpdateSSPSegment(x.y, t.y, z, P.npts, iz);
        real w = (x.y - z[iz]) / (z[iz + 1] - z[iz]); // LinInterpDensity ssp.hpp:53-59
        const real *rho = B + P.o_rho;
        o.rho = (RL(1.0) - w) * rho[iz] + w * rho[iz + 1];
        if(sst == 'C') { // cLinear ssp.hpp:81-89
            cpx<real> c = ldc(B + P.o_c, iz), cz = ldc(B + P.o_cz, iz);
            o.c = c + rmul(x.y - z[iz], cz);
            o.gradc = V(RL(0.0), cz.re);
        } else if(sst == 'S') { // cCubic ssp.hpp:133-152, SplineALL curves.hpp:190-200
            const real *cf = B + P.o_coef;
            int n2 = 2 * P.npts;
            cpx<real> c1 = ldc(cf, iz), c2 = ldc(cf + n2, iz), c3 = ldc(cf + 2 * n2, iz), c4 = ldc(cf + 3 * n2, iz);
            real h = x.y - z[iz];
            const float half = FL(0.5), sixth = FL(1.0) / FL(6.0);
            cpx<real> f = c1 + rmul(h, c2 + rmul(h, rmul((real)half, c3) + rmul((real)sixth * h, c4)));
            cpx<real> fx = c2 + rmul(h, c3 + rmul(h * (real)half, c4));
            cpx<real> fxx = c3 + rmul(h, c4);
            o.c = f;
            o.gradc = V(RL(0.0), fx.re);
            o.czz = fxx.re;
        } else if(sst == 'P') { // cPCHIP ssp.hpp:95-128
            const real *cf = B + P.o_coef;
            int n2 = 2 * P.npts;
            cpx<real> c0 = ldc(cf, iz), c1 = ldc(cf + n2, iz), c2 = ldc(cf + 2 * n2, iz), c3 = ldc(cf + 3 * n2, iz);
            real xt = x.y - z[iz];
            if(std::fabs(xt) > RL(1.0e10)) e.warning(16);
            if(cabs_(c0) > RL(1.0e10) || cabs_(c1) > RL(1.0e10) || cabs_(c2) > RL(1.0e10) || cabs_(c3) > RL(1.0e10))
                e.warning(17);
            o.c = c0 + rmul(xt, c1 + rmul(xt, c2 + rmul(xt, c3)));
            o.gradc = V(RL(0.0), (c1 + rmul(xt, rmul(RL(2.0), c2) + rmul(xt, rmul(RL(3.0), c3)))).re);
            o.czz = (rmul(RL(2.0), c2) + rmul(xt, rmul(RL(6.0), c3))).re;
        } else if(sst == 'N') { // n2Linear ssp.hpp:64-76
            cpx<real> a = ldc(B + P.o_n2, iz), b = ldc(B + P.o_n2, iz + 1);
            o.c = realdiv(RL(1.0), csqrt_(rmul(RL(1.0) - w, a) + rmul(w, b)));
            real c = o.c.re;
            o.gradc = V(RL(0.0), RL(-0.5) * (c * c * c) * ldc(B + P.o_n2z, iz).re);
            o.czz = RL(3.0) * o.gradc.y * o.gradc.y / c;
        }
    } else if(sst == 'Q') { // Quad ssp.hpp:157-214
        const real *segr = B + P.o_segr;
        int Nr = P.Nr;
        if(x.x < segr[0] || x.x > segr[Nr - 1]) e.error(8);
        UpdateSSPSegment(x.y, t.y, z, P.npts, iz);
        UpdateSSPSegment(x.x, t.x, segr, Nr, ir);
        real w = (x.y - z[iz]) / (z[iz + 1] - z[iz]);
        const real *rho = B + P.o_rho;
        o.rho = (RL(1.0) - w) * rho[iz] + w * rho[iz + 1];
        if(iz >= P.npts - 1 || ir >= Nr - 1) {
            e.error(9);
            iz = ir = 0;
        }
        const real *cm = B + P.o_cmat, *czm = B + P.o_czmat;
        real cz1 = czm[iz * Nr + ir], cz2 = czm[iz * Nr + ir + 1];
        real s2 = x.y - z[iz];
        real delta_z = z[iz + 1] - z[iz];
        real c1 = cm[iz * Nr + ir] + s2 * cz1;
        real c2 = cm[iz * Nr + ir + 1] + s2 * cz2;
        real delta_r = segr[ir + 1] - segr[ir];
        real s1 = (x.x - segr[ir]) / delta_r;
        s1 = rmin(s1, RL(1.0));
        s1 = rmax(s1, RL(0.0));
        real c = (RL(1.0) - s1) * c1 + s1 * c2;
        s2 = s2 / delta_z;
        real cim = (rmul(RL(1.0) - s2, ldc(B + P.o_c, iz)) + rmul(s2, ldc(B + P.o_c, iz + 1))).im;
        o.c = C(c, cim);
        real cz = (RL(1.0) - s1) * cz1 + s1 * cz2;
        real cr = (c2 - c1) / delta_r;
        o.crz = (cz2 - cz1) / delta_r;
        o.gradc = V(cr, cz);
    } else { // Analytic 2D ssp.hpp:338-371
        iz = 0;
        real c0 = FL(1500.0);
        o.rho = FL(1.0);
        const float unk1 = FL(1300.0), unk2 = FL(0.00737);
        real xt = FL(2.0) * (x.y - unk1) / unk1;
        real emxt = std::exp(-xt);
        real DxtDz = FL(2.0) / unk1;
        o.c = C(c0 * (FL(1.0) + unk2 * (xt - FL(1.0) + emxt)), (real)0);
        real cz = c0 * unk2 * (FL(1.0) - emxt) * DxtDz;
        o.czz = c0 * unk2 * emxt * (DxtDz * DxtDz);
        o.gradc = V((real)FL(0.0), cz);
    }
}

// ---------------------------------------------------------------------------
// Boundary segment lookup (reference src/boundary.hpp:66-227, 2D)
// ---------------------------------------------------------------------------
template<typename real>
__host__ __device__ inline void GetBdrySeg(v2<real> x, v2<real> t, BdryState<real> &bds, const real *bd, int n,
                                           int NODE, bool longfmt, HSd<real> &hs, bool isTop, bool isInit, Err &e)
{
    if(!isInit
       && !(x.x < bds.lmin || (x.x == bds.lmin && t.x < FL(0.0)) || x.x > bds.lmax
            || (x.x == bds.lmax && t.x >= FL(0.0))))
        return;
    int s = imin(imax(bds.Iseg, 0), n - 2);
    if(t.x >= FL(0.0)) {
        while(s >= 0 && bd[s * NODE + BN_XX] > x.x) --s;
        while(s >= 0 && s < n - 1 && bd[(s + 1) * NODE + BN_XX] <= x.x) ++s;
    } else {
        while(s < n - 1 && bd[(s + 1) * NODE + BN_XX] < x.x) ++s;
        while(s >= 0 && s < n - 1 && bd[s * NODE + BN_XX] >= x.x) --s;
    }
    if(s < 0 || s >= n - 1) {
        e.error(isTop ? 6 : 7);
        s = 0;
    }
    bds.Iseg = s;
    const real *b0 = bd + s * NODE;
    bds.lmin = b0[BN_XX];
    bds.lmax = b0[NODE + BN_XX];
    bds.x = V(b0[BN_XX], b0[BN_XY]);
    bds.n = V(b0[BN_NX], b0[BN_NY]);
    if(longfmt) {
        hs.cP = C(b0[BN_CPR], b0[BN_CPI]);
        hs.cS = C(b0[BN_CSR], b0[BN_CSI]);
        hs.rho = b0[BN_RHO];
    }
}

// ---------------------------------------------------------------------------
// Step (reference src/step.hpp, 2D)
// ---------------------------------------------------------------------------
template<typename real> __host__ __device__ inline bool outside_box(v2<real> x, int dim, const Prob2D<real> &P)
{
    return dim == 0 ? std::fabs(x.x - RL(0.0)) >= P.box_r : std::fabs(x.y - RL(0.0)) >= P.box_z;
}

template<typename real>
__host__ __device__ inline void DepthInterfaceCrossing(bool stepTo, real &hInt, v2<real> &x, v2<real> x0,
                                                       v2<real> u, int iz0, const real *z)
{ // step.hpp:41-79
    if(!stepTo) hInt = RC<real>::MAXV;
    if(std::fabs(u.y) > RC<real>::EPS) {
        if(z[iz0] > x.y) {
            hInt = (z[iz0] - x0.y) / u.y;
            if(stepTo) {
                x = x0 + sc(hInt, u);
                x.y = z[iz0];
            }
        } else if(z[iz0 + 1] < x.y) {
            hInt = (z[iz0 + 1] - x0.y) / u.y;
            if(stepTo) {
                x = x0 + sc(hInt, u);
                x.y = z[iz0 + 1];
            }
        }
    }
}

template<typename real>
__host__ __device__ inline void BeamBoxCrossing(bool stepTo, int dim, real &hBox, v2<real> &x, v2<real> x0, v2<real> u,
                                                const Prob2D<real> &P)
{ // step.hpp:84-108
    if(!stepTo) hBox = RC<real>::MAXV;
    if(outside_box(x, dim, P)) {
        real box = dim == 0 ? P.box_r : P.box_z;
        real d0 = (dim == 0 ? x0.x : x0.y) - RL(0.0);
        real uu = dim == 0 ? u.x : u.y;
        hBox = (box - std::fabs(d0)) / std::fabs(uu);
        if(stepTo) {
            x = x0 + sc(hBox, u);
            real s = RL(0.0) + std::copysign(box, d0);
            if(dim == 0) x.x = s;
            else x.y = s;
        }
    }
}

template<typename real>
__host__ __device__ inline void TopBotCrossing(bool stepTo, real &h, const BdryState<real> &bd, v2<real> &x, v2<real> x0,
                                               v2<real> u, bool &refl)
{ // step.hpp:113-147
    if(!stepTo) h = RC<real>::MAXV;
    v2<real> d = x - bd.x;
    v2<real> d0 = x0 - bd.x;
    real w = dot(bd.n, d);
    if(stepTo ? (w > -RC<real>::INF_STEP) : (w >= RL(0.0))) {
        h = -dot(d0, bd.n) / dot(u, bd.n);
        if(stepTo) {
            x = x0 + sc(h, u);
            if(std::fabs(bd.n.x) < RC<real>::EPS) x.y = bd.x.y;
        }
        refl = true;
    } else {
        refl = false;
    }
}

template<typename real>
__host__ __device__ inline void TopBotSegCrossing(bool stepTo, real &hSeg, const BdryState<real> &top,
                                                  const BdryState<real> &bot, int ir0, v2<real> &x, v2<real> x0,
                                                  v2<real> u, const Prob2D<real> &P, const real *B, bool &topRefl,
                                                  bool &botRefl)
{ // step.hpp:153-205
    real smin = rmax(top.lmin, bot.lmin);
    real smax = rmin(top.lmax, bot.lmax);
    if(P.ssp_type == 'Q') {
        const real *segr = B + P.o_segr;
        smin = rmax(smin, segr[ir0]);
        smax = rmin(smax, segr[ir0 + 1]);
    }
    if(!stepTo) hSeg = RC<real>::MAXV;
    if(std::fabs(u.x) > RC<real>::EPS) {
        if(x.x < smin) {
            hSeg = -(x0.x - smin) / u.x;
            if(stepTo) {
                x = x0 + sc(hSeg, u);
                x.x = smin;
                topRefl = botRefl = false;
            }
        } else if(x.x > smax) {
            hSeg = -(x0.x - smax) / u.x;
            if(stepTo) {
                x = x0 + sc(hSeg, u);
                x.x = smax;
                topRefl = botRefl = false;
            }
        }
    }
}

template<typename real>
__host__ __device__ inline void ReduceStep(v2<real> x0, v2<real> u, int iz0, int ir0, const Trace<real> &S,
                                           const Prob2D<real> &P, const real *B, Err &e, real &h)
{ // step.hpp:271-342
    real hInt, hBoxxr, hBoxyz, hTop, hBot, hxSeg;
    bool dummy;
    v2<real> x = x0 + sc(h, u);
    DepthInterfaceCrossing(false, hInt, x, x0, u, iz0, B + P.o_z);
    BeamBoxCrossing(false, 0, hBoxxr, x, x0, u, P);
    BeamBoxCrossing(false, 1, hBoxyz, x, x0, u, P);
    TopBotCrossing(false, hTop, S.top, x, x0, u, dummy);
    TopBotCrossing(false, hBot, S.bot, x, x0, u, dummy);
    TopBotSegCrossing(false, hxSeg, S.top, S.bot, ir0, x, x0, u, P, B, dummy, dummy);
    const real BIG = RC<real>::MAXV;
    h = rmin(rmin(rmin(h, hInt), rmin(rmin(rmin(hBoxxr, hBoxyz), BIG), rmin(hTop, hBot))),
             rmin(rmin(hxSeg, BIG), rmin(BIG, BIG)));
    if(h < RL(-1e-4)) e.warning(8);
    if(h < RC<real>::INF_STEP * P.deltas) h = RC<real>::INF_STEP * P.deltas;
}

template<typename real>
__host__ __device__ inline void StepToBdry(v2<real> x0, v2<real> &x2, v2<real> u, real &h, bool &topRefl, bool &botRefl,
                                           int iz0, int ir0, const Trace<real> &S, const Prob2D<real> &P,
                                           const real *B)
{ // step.hpp:347-424
    x2 = x0 + sc(h, u);
    DepthInterfaceCrossing(true, h, x2, x0, u, iz0, B + P.o_z);
    BeamBoxCrossing(true, 0, h, x2, x0, u, P);
    BeamBoxCrossing(true, 1, h, x2, x0, u, P);
    TopBotCrossing(true, h, S.top, x2, x0, u, topRefl);
    TopBotCrossing(true, h, S.bot, x2, x0, u, botRefl);
    if(botRefl) topRefl = false;
    TopBotSegCrossing(true, h, S.top, S.bot, ir0, x2, x0, u, P, B, topRefl, botRefl);
    if(h < RC<real>::INF_STEP * P.deltas) {
        h = RC<real>::INF_STEP * P.deltas;
        x2 = x0 + sc(h, u);
        v2<real> d = x2 - S.top.x;
        topRefl = dot(S.top.n, d) > RC<real>::EPS;
        d = x2 - S.bot.x;
        if(dot(S.bot.n, d) > RC<real>::EPS) {
            botRefl = true;
            topRefl = false;
        } else {
            botRefl = false;
        }
    }
}

template<typename real, char ST = 0>
__host__ __device__ inline void Step(const RayPt<real> &ray0, RayPt<real> &ray2, Trace<real> &S, const Prob2D<real> &P,
                                     const real *B, Err &e, bool &topRefl, bool &botRefl)
{ // step.hpp:654-811 (2D)
    SSPOut<real> o0, o1, o2;
    EvaluateSSP<real, ST>(ray0.x, ray0.t, o0, P, B, S.iz, S.ir, e);
    real c0 = o0.c.re;
    real cnn0 = o0.crr * (ray0.t.y * ray0.t.y) - FL(2.0) * o0.crz * ray0.t.x * ray0.t.y + o0.czz * (ray0.t.x * ray0.t.x);
    v2<real> pq0p = sc(-cnn0, ray0.q), pq0q = sc(c0, ray0.p);
    int iz0 = S.iz, ir0 = S.ir;
    real csq0 = c0 * c0;
    v2<real> u0 = sc(c0, ray0.t);
    real h = P.deltas;
    ReduceStep(ray0.x, u0, iz0, ir0, S, P, B, e, h);
    real halfh = FL(0.5) * h;
    RayPt<real> ray1;
    ray1.x = ray0.x + sc(halfh, u0);
    ray1.t = ray0.t - dv(sc(halfh, o0.gradc), csq0);
    ray1.p = ray0.p + sc(halfh, pq0p);
    ray1.q = ray0.q + sc(halfh, pq0q);

    EvaluateSSP<real, ST>(ray1.x, ray1.t, o1, P, B, S.iz, S.ir, e);
    real c1 = o1.c.re;
    real cnn1 = o1.crr * (ray1.t.y * ray1.t.y) - FL(2.0) * o1.crz * ray1.t.x * ray1.t.y + o1.czz * (ray1.t.x * ray1.t.x);
    v2<real> pq1p = sc(-cnn1, ray1.q), pq1q = sc(c1, ray1.p);
    real csq1 = c1 * c1;
    v2<real> u1 = sc(c1, ray1.t);
    ReduceStep(ray0.x, u1, iz0, ir0, S, P, B, e, h);

    real w1 = h / (RL(2.0) * halfh);
    real w0 = RL(1.0) - w1;
    v2<real> u2 = sc(w0, u0) + sc(w1, u1);
    v2<real> x2;
    StepToBdry(ray0.x, x2, u2, h, topRefl, botRefl, iz0, ir0, S, P, B);
    ray2.x = x2;
    real hw0 = h * w0, hw1 = h * w1;
    ray2.t = (ray0.t - dv(sc(hw0, o0.gradc), csq0)) - dv(sc(hw1, o1.gradc), csq1);
    ray2.tau = (ray0.tau + realdiv(hw0, o0.c)) + realdiv(hw1, o1.c);
    ray2.p = ray0.p + sc(hw0, pq0p);
    ray2.q = ray0.q + sc(hw0, pq0q);
    ray2.p = ray2.p + sc(hw1, pq1p);
    ray2.q = ray2.q + sc(hw1, pq1q);
    ray2.Amp = ray0.Amp;
    ray2.Phase = ray0.Phase;
    ray2.NumTopBnc = ray0.NumTopBnc;
    ray2.NumBotBnc = ray0.NumBotBnc;

    EvaluateSSP<real, ST>(ray2.x, ray2.t, o2, P, B, S.iz, S.ir, e);
    ray2.c = o2.c.re;
    if(S.iz != iz0 || S.ir != ir0) { // CurvatureCorrection 2D, step.hpp:626-648
        v2<real> gj = o2.gradc - o0.gradc;
        v2<real> ray2n = V(-ray2.t.y, ray2.t.x);
        real cnjump = dot(gj, ray2n);
        real csjump = dot(gj, ray2.t);
        real rm = (S.iz != iz0) ? ray2.t.x / ray2.t.y : -ray2.t.y / ray2.t.x;
        real rn = rm * (FL(2.0) * cnjump - rm * csjump) / ray2.c;
        ray2.p = ray2.p - sc(rn, ray2.q);
    }
}

// ---------------------------------------------------------------------------
// Reflection (reference src/reflect.hpp, 2D)
// ---------------------------------------------------------------------------

// Beam displacement and width change on a halfspace reflection (Tindle's
// displacement as in src/reflect.hpp:383-442, 2D): the reflected point moves
// along the boundary by Re(delta), the delay grows by Re(delta) cos(theta)/c
// and q by |d delta / d theta|-scaled p. Complex operations in the
// reference's order (real / complex as a full complex division).
template<typename real>
__host__ __device__ inline void BeamShift(const RayPt<real> &oldp, RayPt<real> &newp, const HSd<real> &hs, v2<real> tB,
                                          const SSPOut<real> &o, real omega)
{
    cpx<real> ch = realdiv(oldp.c, C(hs.cP.re, -hs.cP.im));
    real co = oldp.t.x * oldp.c;
    real si = oldp.t.y * oldp.c;
    real ck = omega / oldp.c;
    cpx<real> chsq = cmul(ch, ch);
    cpx<real> a = rmul(FL(2.0) * hs.rho, C(FL(1.0) - chsq.re, -chsq.im));
    cpx<real> b = C(co * co - chsq.re, -chsq.im);
    cpx<real> d = C((hs.rho * hs.rho) * (si * si) + b.re, b.im);
    cpx<real> sb = csqrt_(b);
    real cco = co * co;
    real ssi = si * si;
    cpx<real> cksbd = cmul(rmul(ck, sb), d);
    cpx<real> delta = C((real)0, (real)0);
    if(std::fabs(si) >= RC<real>::EPS) delta = cdiv(rdivc(C(a.re * co, a.im * co), si), cksbd);
    real pdelta = delta.re / (oldp.c / co);
    cpx<real> t1 = realdiv(cco, cmul(cksbd, b));
    cpx<real> t2 = realdiv(RL(1.0) + (cco / ssi), cksbd);
    cpx<real> t3 = realdiv(FL(2.0) * (co * co) * ((hs.rho * hs.rho) - RL(1.0)), cmul(cksbd, d));
    cpx<real> ddelta = cmul(a, (t1 - t2) - t3);
    real rddelta = -ddelta.re;
    real sddelta = rddelta / std::fabs(rddelta);
    real theta_bot = std::atan(tB.y / tB.x);
    newp.x.x = newp.x.x + delta.re * std::cos(theta_bot);
    newp.x.y = newp.x.y + delta.re * std::sin(theta_bot);
    newp.tau = C(newp.tau.re + pdelta, newp.tau.im);
    newp.q = newp.q + sc(((sddelta * rddelta) * si) * o.c.re, oldp.p);
}

template<typename real, char ST = 0>
__host__ __device__ inline void Reflect(const RayPt<real> &oldp, RayPt<real> &newp, const HSd<real> &hs, bool isTop,
                                        v2<real> tB, v2<real> nB, real kappa, const Prob2D<real> &P, const real *B,
                                        const real *G, int &iz, int &ir, Err &e)
{
    real Th = dot(oldp.t, nB);
    real Tg = dot(oldp.t, tB);
    newp.NumTopBnc = oldp.NumTopBnc + (isTop ? 1 : 0);
    newp.NumBotBnc = oldp.NumBotBnc + (isTop ? 0 : 1);
    newp.x = oldp.x;
    newp.t = oldp.t - sc(FL(2.0) * Th, nB);
    SSPOut<real> o;
    EvaluateSSP<real, ST>(oldp.x, oldp.t, o, P, B, iz, ir, e);
    real c = o.c.re;
    newp.c = c;
    newp.tau = oldp.tau;
    v2<real> rayt = sc(c, oldp.t);
    v2<real> rayn = V(-rayt.y, rayt.x);
    v2<real> rayt_t = sc(c, newp.t);
    v2<real> rayn_t = V(-(-rayt_t.y), -(rayt_t.x));
    real rn = FL(2.0) * kappa / (c * c) / Th;
    real cnjump = -dot(o.gradc, rayn_t - rayn);
    real csjump = -dot(o.gradc, rayt_t - rayt);
    if(isTop) {
        cnjump = -cnjump;
        rn = -rn;
    }
    real rm = Tg / Th;
    rn += rm * (FL(2.0) * cnjump - rm * csjump) / (c * c);
    if(P.type2 == 'D') rn *= FL(2.0);
    else if(P.type2 == 'Z') rn = FL(0.0);
    newp.p = oldp.p + sc(rn, oldp.q);
    newp.q = oldp.q;

    if(hs.bc == 'R') {
        newp.Amp = oldp.Amp;
        newp.Phase = oldp.Phase;
    } else if(hs.bc == 'V') {
        newp.Amp = oldp.Amp;
        newp.Phase = oldp.Phase + RC<real>::PI;
    } else if(hs.bc == 'F') { // InterpolateReflectionCoefficient reflect.hpp:39-92
        const real *tab = B + (isTop ? P.o_trc : P.o_brc); // staged with the boundaries (LDS when it fits)
        int n = isTop ? P.ntrc : P.nbrc;
        real theta = (RL(180.0) / RC<real>::PI) * std::fabs(std::atan2(Th, Tg));
        if(theta > FL(90.0)) theta = FL(180.0) - theta;
        real r, phi;
        int iL = 0, iR = n - 1;
        if(theta < tab[0]) {
            r = phi = FL(0.0);
            e.warning(5);
        } else if(theta > tab[3 * iR]) {
            r = phi = FL(0.0);
        } else {
            while(iL != iR - 1) {
                int iM = (iL + iR) / 2;
                if(tab[3 * iM] > theta) iR = iM;
                else iL = iM;
            }
            real a = (theta - tab[3 * iL]) / (tab[3 * iR] - tab[3 * iL]);
            r = (FL(1.0) - a) * tab[3 * iL + 1] + a * tab[3 * iR + 1];
            phi = (FL(1.0) - a) * tab[3 * iL + 2] + a * tab[3 * iR + 2];
        }
        newp.Amp = oldp.Amp * r;
        newp.Phase = oldp.Phase + phi;
    } else if(hs.bc == 'A' || hs.bc == 'G') {
        real omega = FL(2.0) * RC<real>::PI * P.freq0;
        cpx<real> kx = C(omega * Tg, (real)0), kz = C(omega * Th, (real)0);
        cpx<real> kx2 = cmul(kx, kx);
        cpx<real> f, g;
        if(hs.cS.re > FL(0.0)) {
            cpx<real> wS = realdiv(omega, hs.cS), wP = realdiv(omega, hs.cP);
            cpx<real> kzS2 = kx2 - cmul(wS, wS);
            cpx<real> kzP2 = kx2 - cmul(wP, wP);
            cpx<real> kzS = csqrt_(kzS2), kzP = csqrt_(kzP2);
            cpx<real> mu = rmul(hs.rho, cmul(hs.cS, hs.cS));
            cpx<real> s = kzS2 + kx2;
            cpx<real> y2 = cmul(cmul(s, s) - cmul(cmul(rmul(RL(4.0), kzS), kzP), kx2), mu);
            cpx<real> y4 = cmul(kzP, kx2 - kzS2);
            f = rmul(omega * omega, y4);
            g = y2;
        } else {
            cpx<real> wP = realdiv(omega, hs.cP);
            cpx<real> kzP = csqrt_(kx2 - cmul(wP, wP));
            if(std::fabs(kzP.re) < RC<real>::EPS && kzP.im < RL(0.0)) kzP = -kzP;
            f = kzP;
            g = C(hs.rho, (real)0);
        }
        cpx<real> J = C((real)0, (real)1);
        cpx<real> a = rmul(o.rho, f), b = cmul(cmul(J, kz), g);
        cpx<real> Refl = cdiv(-(a - b), a + b);
        real aR = cabs_(Refl);
        if(aR < FL(1.0e-5)) {
            newp.Amp = FL(0.0);
            newp.Phase = oldp.Phase;
        } else {
            newp.Amp = aR * oldp.Amp;
            newp.Phase = oldp.Phase + std::atan2(Refl.im, Refl.re);
            if(P.type3 == 'S') BeamShift(oldp, newp, hs, tB, o, omega);
        }
    } else {
        e.error(4);
        newp.Amp = NAN;
        newp.Phase = NAN;
    }
}

template<typename real>
__host__ __device__ inline void Distances(v2<real> x, const Trace<real> &S, real &dt, real &db)
{ // trace.hpp:37-45
    dt = -dot(S.top.n, x - S.top.x);
    db = -dot(S.bot.n, x - S.bot.x);
}

// ---------------------------------------------------------------------------
// Ray init / update / terminate (reference src/trace.hpp, 2D)
// ---------------------------------------------------------------------------
template<typename real>
__host__ __device__ inline int BinarySearchLEQ(const real *arr, int n, int stride, int offset, real target)
{ // common_run.hpp:61-80
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
template<typename real> __host__ __device__ inline int BinarySearchGT(const real *arr, int n, real target)
{ // common_run.hpp:119-138
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

// Returns false when the ray is not traced (source outside the boundaries).
template<typename real, char ST = 0>
__host__ __device__ inline bool RayInit(int isz, int ialpha, RayPt<real> &p0, Trace<real> &S, Infl<real> &I,
                                        const Prob2D<real> &P, const real *B, const real *G, Err &e)
{ // trace.hpp:76-266 (2D) + Init_Influence influence.hpp:525-644
    real alpha = G[P.o_alpha + ialpha];
    real RadDeg = RL(180.0) / RC<real>::PI;
    I.alpha = alpha;
    I.SrcDeclAngle = RadDeg * alpha;
    S.iz = S.ir = 0;
    v2<real> xs = V((real)FL(0.0), (real)P.Sz[isz]);
    S.xs_y = xs.y;
    v2<real> tinit = V(G[P.o_acos + ialpha], G[P.o_asin + ialpha]);
    SSPOut<real> o;
    EvaluateSSP<real, ST>(xs, tinit, o, P, B, S.iz, S.ir, e);
    real c = o.c.re;
    {
        real DalphaOpt = std::sqrt(c / (FL(6.0) * P.freq0 * G[P.o_Rr + P.NRr - 1]));
        int NalphaOpt = 2 + (int)((G[P.o_alpha + P.Nalpha - 1] - G[P.o_alpha]) / DalphaOpt);
        if(P.run0 == 'C' && P.Nalpha < NalphaOpt && ialpha == 0) e.warning(3);
    }
    const real *sbp = G + P.o_sbp;
    int ibp = BinarySearchLEQ(sbp, P.nsbp, 2, 0, I.SrcDeclAngle);
    ibp = imin(ibp, P.nsbp - 2);
    real s = (I.SrcDeclAngle - sbp[2 * ibp]) / (sbp[2 * (ibp + 1)] - sbp[2 * ibp]);
    float Amp0 = (float)((FL(1.0) - s) * sbp[2 * ibp + 1] + s * sbp[2 * (ibp + 1) + 1]);
    if(P.run0 == 'S') {
        float omega = (float)(FL(2.0) * RC<real>::PI * P.freq0);
        Amp0 = (float)(Amp0 * (sqrtf(FL(2.0)) * std::fabs(std::sin(omega / c * xs.y * G[P.o_asin + ialpha]))));
    }
    p0.x = xs;
    p0.c = c;
    p0.t = dv(tinit, c);
    p0.tau = C((real)0, (real)0);
    p0.Amp = Amp0;
    p0.Phase = FL(0.0);
    p0.NumTopBnc = p0.NumBotBnc = 0;
    p0.p = V((real)FL(1.0), (real)FL(0.0));
    p0.q = V((real)FL(0.0), (real)FL(1.0));
    if(P.run1 == 'G') p0.q = V((real)FL(0.0), (real)FL(0.0));
    S.htop = P.top;
    S.hbot = P.bot;
    S.top.Iseg = S.bot.Iseg = 0;
    GetBdrySeg(xs, p0.t, S.top, B + P.o_ati, P.ati_n, P.node, P.ati_long == 'L', S.htop, true, true, e);
    GetBdrySeg(xs, p0.t, S.bot, B + P.o_bty, P.bty_n, P.node, P.bty_long == 'L', S.hbot, false, true, e);
    Distances(xs, S, S.DistBegTop, S.DistBegBot);
    if(S.DistBegTop <= FL(0.0) || S.DistBegBot <= FL(0.0)) {
        e.warning(4);
        return false;
    }
    // Init_Influence
    I.omega = FL(2.0) * RC<real>::PI * P.freq0;
    I.BeamWindow = P.isGaussian ? RL(4.0) : RL(1.0);
    I.Ratio1 = std::sqrt(std::fabs(G[P.o_acos + ialpha]));
    if(P.run3 == 'X') I.Ratio1 = RL(1.0);
    if(P.isGaussian) I.Ratio1 /= std::sqrt(FL(2.0) * RC<real>::PI);
    I.rcp_q0 = P.Dalpha / p0.c;
    I.phase = FL(0.0);
    I.qOld = p0.q.x;
    I.x = p0.x;
    if(P.isRayCen) {
        I.rayn1 = sc(p0.c, V(p0.t.y, -p0.t.x));
        I.lastValid = std::fabs(I.rayn1.y) >= RL(1e-6);
        I.ir = 0;
    } else {
        I.rayn1 = V((real)0, (real)0);
        I.lastValid = 0;
        I.ir = BinarySearchGT(G + P.o_Rr, P.NRr, p0.x.x);
        if(p0.t.x < RL(0.0) && I.ir > 0) --I.ir;
        I.rbr = range_bracket(G + P.o_Rr, P.NRr, I.ir);
    }
    return true;
}

template<typename real, char ST = 0>
__host__ __device__ inline bool RayUpdate(const RayPt<real> &p0, RayPt<real> &p1, RayPt<real> &p2, Trace<real> &S,
                                          const Prob2D<real> &P, const real *B, const real *G, Err &e)
{ // trace.hpp:274-381 (2D)
    bool topRefl = false, botRefl = false;
    Step<real, ST>(p0, p1, S, P, B, e, topRefl, botRefl);
    GetBdrySeg(p1.x, p1.t, S.top, B + P.o_ati, P.ati_n, P.node, P.ati_long == 'L', S.htop, true, false, e);
    GetBdrySeg(p1.x, p1.t, S.bot, B + P.o_bty, P.bty_n, P.node, P.bty_long == 'L', S.hbot, false, false, e);
    Distances(p1.x, S, S.DistEndTop, S.DistEndBot);
    if(topRefl || botRefl) {
        const real *bd = B + (topRefl ? P.o_ati : P.o_bty);
        const BdryState<real> &bs = topRefl ? S.top : S.bot;
        const real *b0 = bd + bs.Iseg * P.node;
        const real *b1 = b0 + P.node;
        v2<real> nI, tI;
        if((topRefl ? P.ati_curv : P.bty_curv) == 'C') {
            real sss = dot(p1.x - bs.x, V(b0[BN_TX], b0[BN_TY])) / b0[BN_LEN];
            nI = sc(FL(1.0) - sss, V(b0[BN_NDNX], b0[BN_NDNY])) + sc(sss, V(b1[BN_NDNX], b1[BN_NDNY]));
            tI = sc(FL(1.0) - sss, V(b0[BN_NDTX], b0[BN_NDTY])) + sc(sss, V(b1[BN_NDTX], b1[BN_NDTY]));
        } else {
            nI = V(b0[BN_NX], b0[BN_NY]);
            tI = V(b0[BN_TX], b0[BN_TY]);
        }
        Reflect<real, ST>(p1, p2, topRefl ? S.htop : S.hbot, topRefl, tI, nI, b0[BN_KAPPA], P, B, G, S.iz, S.ir, e);
        Distances(p2.x, S, S.DistEndTop, S.DistEndBot);
        return true;
    }
    return false;
}

// Returns 0 to continue, 1 for a normal stop (Nsteps = is + 1), 2 when the
// per-ray storage is exhausted (Nsteps = is).
template<typename real>
__host__ __device__ inline int RayTerminate(const RayPt<real> &pt, int is, Trace<real> &S, int MaxPoints,
                                            const Prob2D<real> &P, Err &e)
{ // trace.hpp:391-477 (2D)
    bool leftbox = outside_box(pt.x, 0, P) || outside_box(pt.x, 1, P);
    bool escaped = (S.DistBegTop < FL(0.0) && S.DistEndTop < FL(0.0))
        || (S.DistBegBot < FL(0.0) && S.DistEndBot < FL(0.0));
    bool lost = pt.Amp < FL(0.005);
    if(leftbox || lost || escaped) return 1;
    if(is >= MaxPoints - 3) {
        e.warning(1);
        return 2;
    }
    S.DistBegTop = S.DistEndTop;
    S.DistBegBot = S.DistEndBot;
    return 0;
}

// ---------------------------------------------------------------------------
// Influence (reference src/influence.hpp, geometric 2D)
//
// The reference walks the receivers of each step segment inside the ray's own
// thread. Here a segment is first reduced to a descriptor (SegDesc) by the lane
// that owns the ray -- this applies every update of the ray's influence state
// in the reference's order -- and the receiver work items of the descriptor are
// then evaluated by any lane (the whole wavefront, in the kernel). Every
// contribution is computed from the same operands with the same operations as
// in the reference, so only the order of the field additions changes.
// ---------------------------------------------------------------------------
template<typename real>
__host__ __device__ inline bool IsAtCaustic(real qOld, real q)
{
    return (q <= RL(0.0) && qOld > RL(0.0)) || (q >= RL(0.0) && qOld < RL(0.0));
}


// First index i with Rz[i] >= z (GE) or Rz[i] > z (!GE) of strictly increasing
// depths: an index guessed from the mean spacing (rdz = (n - 1) / (Rz[n-1] -
// Rz[0]), 0 for none) and verified against its neighbours, moved by at most
// four places; a binary search when the guess does not settle. Either way the
// result is the exact bound.
template<bool GE, typename real, typename ZT>
__host__ __device__ inline int rz_bound(const ZT *Rz, int n, double rdz, real z)
{
    if(rdz > 0.0) {
        const double g = ((double)z - (double)(real)Rz[0]) * rdz;
        int i = !(g > 0.0) ? 0 : (g >= (double)n ? n : (int)g + 1);
        for(int t = 0; t < 4; ++t) {
            if(i > 0 && (GE ? (real)Rz[i - 1] >= z : (real)Rz[i - 1] > z)) {
                --i;
                continue;
            }
            if(i < n && (GE ? (real)Rz[i] < z : (real)Rz[i] <= z)) {
                ++i;
                continue;
            }
            return i;
        }
    }
    int a = 0, b = n;
    while(a < b) {
        int m = (a + b) >> 1;
        if(GE ? (real)Rz[m] < z : (real)Rz[m] <= z) a = m + 1;
        else b = m;
    }
    return a;
}

// Receiver-depth index window [i0, i1] (inclusive, clamped, possibly empty with
// i1 < i0) that contains every depth Rz[i] with zlo <= Rz[i] <= zhi, widened by
// one index on each side; requires strictly increasing Rz.
template<typename real, typename ZT>
__host__ __device__ inline void depth_window(const ZT *Rz, int n, double rdz, real zlo, real zhi, int &i0, int &i1)
{
    i0 = imax(rz_bound<true>(Rz, n, rdz, zlo) - 1, 0);
    i1 = imin(rz_bound<false>(Rz, n, rdz, zhi), n - 1);
}

template<typename real> struct SegDesc {
    // core -- every field a Cartesian TL scatter reads, and all that is stored
    // for a Cartesian TL segment. Geometry: Cartesian, the step's start and
    // unit tangent (its normal is (-t_z, t_r)); ray-centred, A = the previous
    // normal point / normal.
    v2<real> xa, ta;
    real rlen, zmin, zmax, lo, hi;
    // beam quantities of InfluenceGeoCore
    real q0x, dqx, phasePt, phase, qOld, rcp_q0;
    real A0;                    // Ratio1 * sqrt(c1) * Amp1: numerator of the contribution amplitude
    real sigfloor;              // Gaussian beams: lower bound of the beam width (AdjustSigma)
    cpx<real> tau0, dtau;
    int32_t ir0, ir1, dir;      // Cartesian: visited receiver ranges ir0..ir1 (step dir)
    int32_t iz0, nz;            // depth window: receiver depths iz0 .. iz0+nz-1
    int32_t nitems;             // work items (0 = nothing to scatter)
    int32_t isz;
    int32_t prob;               // problem index (batched runs)
    // rest -- ray-centred geometry (B = this step's normal point / normal),
    // arrivals and eigen hits
    v2<real> xb, tb;
    real RcvrDecl, SrcDecl;
    real wmax;                  // ray-centred: upper bound of BeamWindow * sigma over the segment
    int32_t lastValid;          // ray-centred
    int32_t ntop, nbot, ialpha, is;
};

// Bytes of a descriptor's core (its leading part up to the rest).
template<typename real> constexpr size_t seg_core_bytes() { return offsetof(SegDesc<real>, xb); }

// Whether a segment needs the descriptor's rest: ray-centred geometry, or a
// run that records arrivals or eigen hits.
template<typename real> __host__ __device__ inline bool seg_full(const Prob2D<real> &P)
{
    return P.isRayCen || !tl_run(P.run0);
}

template<typename real>
__host__ __device__ inline void desc_common(SegDesc<real> &d, const RayPt<real> &p0, const RayPt<real> &p1,
                                            const Infl<real> &I, const Prob2D<real> &P, int is, int isz, int ialpha)
{
    d.q0x = p0.q.x;
    d.dqx = p1.q.x - p0.q.x;
    d.tau0 = p0.tau;
    d.dtau = p1.tau - p0.tau;
    // per-segment factors of InfluenceGeoCore (influence.hpp:970-992), in the
    // reference's operation order so that every contribution is bit-identical
    d.A0 = I.Ratio1 * std::sqrt(p1.c) * p1.Amp;
    d.sigfloor = P.isGaussian ? rmin(FL(0.2) * P.freq0 * p1.tau.re, RC<real>::PI * (p0.c / P.freq0)) : RL(0.0);
    d.phasePt = P.isGaussian ? p1.Phase : p0.Phase;
    d.phase = I.phase;
    d.qOld = I.qOld;
    d.rcp_q0 = I.rcp_q0;
    d.SrcDecl = I.SrcDeclAngle;
    d.ntop = p1.NumTopBnc;
    d.nbot = p1.NumBotBnc;
    d.isz = isz;
    d.ialpha = ialpha;
    d.is = is;
}

// Cartesian prelude: influence.hpp:1201-1297 and the receiver walk :1388-1402.
// WINDOW = false leaves the segment's depth window to the scatter (every
// depth in the descriptor; the scatter's per-column band and the zmin / zmax
// test of each receiver select the same contributions): a batched trace,
// latency-bound at one wavefront per SIMD, hands that search to the scatter.
template<typename real, bool WINDOW = true>
__host__ __device__ inline void prelude_cart(SegDesc<real> &d, const RayPt<real> &p0, const RayPt<real> &p1,
                                             Infl<real> &I, int is, const Prob2D<real> &P, const real *G,
                                             const real *RzT, int isz, int ialpha)
{
    const real *Rr = G + P.o_Rr;
    d.nitems = 0;
    real rA = I.x.x, rB = p1.x.x;
    v2<real> rayt = p1.x - p0.x;
    real rlen = std::sqrt(dot(rayt, rayt));
    if(rlen < RL(1.0e3) * spacing(p1.x.x)) return;
    real phaseq = p0.q.x;
    if(IsAtCaustic(I.qOld, phaseq)) I.phase += RC<real>::PI / FL(2.0);
    I.qOld = phaseq;
    d.lo = rmin(rA, rB);
    d.hi = rmax(rA, rB);
    // walk: the set of visited receiver ranges is contiguous. The bracket
    // (Rr[ir], Rr[ir + 1]) travels with the ray, so a step that crosses no
    // receiver range reads no range from memory; the comparisons are the walk's.
    int ir = I.ir, first = I.ir;
    real rlo = I.rbr.x, rhi = I.rbr.y;
    const real rfirst = rlo;    // Rr[first]
    for(;;) {
        int irTT;
        if(rB > rlo) {
            if(ir >= P.NRr - 1) break;
            irTT = ir + 1;
            if(rhi >= rB) break;
            rlo = rhi;
            if(irTT < P.NRr - 1) rhi = Rr[irTT + 1];
        } else {
            if(ir <= 0) break;
            irTT = ir - 1;
            const real r = Rr[irTT];
            if(r <= rB) break;
            rhi = rlo;
            rlo = r;
        }
        ir = irTT;
    }
    I.ir = ir;
    I.rbr = V(rlo, rhi);
    I.x = p1.x;
    // A step that crosses no receiver range has the one column Rr[first]:
    // outside [lo, hi) the scatter's column test accepts none of its
    // receivers, so the segment is not handed to the scatter at all.
    if(ir == first && !(rfirst >= d.lo && rfirst < d.hi)) return;
    // the beam's reach: only for steps handed to the scatter
    rayt = dv(rayt, rlen);
    real sigma = rmax(std::fabs(p0.q.x), std::fabs(p1.q.x)) * I.rcp_q0 / std::fabs(rayt.x);
    if(P.isGaussian) {
        real lambda = p0.c / P.freq0;
        sigma = rmax(sigma, rmin(FL(0.2) * P.freq0 * p1.tau.re, RC<real>::PI * lambda));
    }
    real RadiusMax = I.BeamWindow * sigma;
    if(std::fabs(rayt.x) > FL(0.50001)) {
        d.zmin = rmin(p0.x.y, p1.x.y) - RadiusMax;
        d.zmax = rmax(p0.x.y, p1.x.y) + RadiusMax;
    } else {
        d.zmin = -RC<real>::MAXV;
        d.zmax = RC<real>::MAXV;
    }
    d.ir0 = first;
    d.ir1 = ir;
    d.dir = (ir >= first) ? 1 : -1;
    int nr = (ir >= first ? ir - first : first - ir) + 1;
    d.iz0 = 0;
    d.nz = P.NRz_per_range;
    if(WINDOW && P.rz_sorted && P.run4 != 'I' && d.zmin > -RC<real>::MAXV) {
        int i0, i1;
        depth_window(RzT, P.NRz_per_range, P.rz_rdz, d.zmin, d.zmax, i0, i1);
        d.iz0 = i0;
        d.nz = imax(0, i1 - i0 + 1);
    }
    d.nitems = nr * d.nz;
    d.xa = p0.x;
    d.ta = rayt;
    d.rlen = rlen;
    // the receiver angle enters arrivals only (TL runs skip the atan2)
    d.RcvrDecl = tl_run(P.run0) ? RL(0.0) : (RL(180.0) / RC<real>::PI) * std::atan2(rayt.y, rayt.x);
    desc_common(d, p0, p1, I, P, is, isz, ialpha);
}

// Ray-centred prelude: influence.hpp:1030-1071.
template<typename real>
__host__ __device__ inline void prelude_raycen(SegDesc<real> &d, const RayPt<real> &p0, const RayPt<real> &p1,
                                               Infl<real> &I, int is, const Prob2D<real> &P, const real *RzT, int isz,
                                               int ialpha)
{
    d.nitems = 0;
    real phaseq = p0.q.x;
    if(IsAtCaustic(I.qOld, phaseq)) I.phase += RC<real>::PI / FL(2.0);
    I.qOld = phaseq;
    v2<real> rayn1 = sc(p1.c, V(p1.t.y, -p1.t.x));
    if(std::fabs(rayn1.y) < RL(1e-10)) return;
    if(std::fabs(p1.x.x - p0.x.x) < RL(1.0e3) * spacing(p1.x.x)) {
        I.lastValid = 1;
        I.x = p1.x;
        I.rayn1 = rayn1;
        return;
    }
    d.lastValid = I.lastValid;
    d.xa = I.x;
    d.ta = I.rayn1;
    d.xb = p1.x;
    d.tb = rayn1;
    d.RcvrDecl = tl_run(P.run0) ? RL(0.0) : (RL(180.0) / RC<real>::PI) * std::atan2(p1.t.y, p1.t.x);
    desc_common(d, p0, p1, I, P, is, isz, ialpha);
    d.nitems = P.NRz_per_range;
    {
        // sigma(s) <= max(|q0|, |q1|) * rcp_q0 (q is linear in s), raised to the
        // AdjustSigma floor; 1e-4 relative slack covers the rounding of s and n1.
        double q0 = std::fabs((double)p0.q.x), q1 = std::fabs((double)p1.q.x);
        double qmax = (q0 > q1 ? q0 : q1) + 1e-3 * std::fabs((double)p1.q.x - (double)p0.q.x);
        double sig = qmax * std::fabs((double)I.rcp_q0);
        if(P.isGaussian) {
            double adj = (double)rmin(FL(0.2) * P.freq0 * p1.tau.re, RC<real>::PI * (p0.c / P.freq0));
            sig = sig > adj ? sig : adj;
        }
        d.wmax = (real)((P.isGaussian ? 4.0 : 1.0) * sig * (1.0 + 1e-4));
    }
    d.iz0 = 0;
    d.nz = P.NRz_per_range;
    if(P.rz_sorted && d.lastValid) {
        // n at depth z is linear between the normals at A and B: a depth can only
        // contribute when that range of n meets [-wmax, wmax] (|normal| = 1), i.e.
        // within wmax (plus slack) of the depths of A and B.
        double ya = (double)d.xa.y, yb = (double)d.xb.y, W = (double)d.wmax;
        double slack = 1e-4 * W + 1e-6 * (std::fabs(ya) + std::fabs(yb)) + 1e-3;
        double zlo = (ya < yb ? ya : yb) - W - slack, zhi = (ya < yb ? yb : ya) + W + slack;
        int i0, i1;
        depth_window(RzT, P.NRz_per_range, P.rz_rdz, (real)zlo, (real)zhi, i0, i1);
        d.iz0 = i0;
        d.nz = imax(0, i1 - i0 + 1);
    }
    d.nitems = d.nz;
    I.lastValid = 1;
    I.x = p1.x;
    I.rayn1 = rayn1;
}

// cos and sin of an FP64 phase at FP32 accuracy: Cody-Waite reduction by 2*pi
// in FP64 (two-part constant, exact to ~1e-11 rad for |a| < 1e6), then the FP32
// sincos of the remainder in [-pi, pi].
__host__ __device__ inline void sincos_phase(double a, float &s, float &c)
{
    const double INV_2PI = 0.15915494309189535;
    const double TWO_PI_HI = 6.283185307179586;       // 2*pi rounded to double
    const double TWO_PI_LO = 2.4492935982947064e-16;  // 2*pi - TWO_PI_HI
    double k = std::rint(a * INV_2PI);
    double r = (a - k * TWO_PI_HI) - k * TWO_PI_LO;
    sincosf((float)r, &s, &c);
}

// InfluenceGeoCore (influence.hpp:897-1024, 2D) + ApplyContribution (:432-473).
// FAST = false follows the reference's arithmetic to the last bit. FAST = true
// evaluates the final TL term -- a unit phasor and Gaussian weight that are
// then rounded to the FP32 field -- at FP32 accuracy: the phase is reduced in
// FP64 first, the geometry, delay and amplitude stay in `real`.
template<typename real, int RT, typename Sink, bool FAST = false>
__host__ __device__ inline void contrib(const SegDesc<real> &d, real s, real n1, int ir, int iz, const Prob2D<real> &P,
                                        Sink &sink)
{
    sink.mark(); // op-count hook (no-op in the kernels)
    real qFinal = d.q0x + s * d.dqx;
    real sigma, sigma_orig;
    sigma = sigma_orig = std::fabs(qFinal * d.rcp_q0);
    if(P.isGaussian) sigma = rmax(sigma, d.sigfloor);
    real BW = P.isGaussian ? RL(4.0) : RL(1.0);
    if(n1 > BW * sigma || n1 == BW * sigma) return;
    cpx<real> delay = d.tau0 + rmul(s, d.dtau);
    real cnst = d.A0 / std::sqrt(std::fabs(qFinal));
    real w;
    if(P.isGaussian) {
        real t = n1 / sigma;
        if constexpr(FAST) w = (real)expf((float)(FL(-0.5) * (t * t))) * (sigma_orig / sigma);
        else w = std::exp(FL(-0.5) * (t * t)) * (sigma_orig / sigma);
    } else {
        w = (sigma - n1) / sigma;
    }
    real phaseInt = d.phasePt + d.phase;
    if(IsAtCaustic(d.qOld, qFinal)) phaseInt = d.phase + RC<real>::PI / FL(2.0);
    size_t idx = ((size_t)d.isz * P.NRz_per_range + iz) * P.NRr + ir;
    real omega = FL(2.0) * RC<real>::PI * P.freq0;
    if constexpr(RT == RT_TL) {
        float fr, fi;
        if(P.run0 == 'C') {
            cpx<real> arg = rmul(omega, delay);
            arg.re = arg.re - phaseInt;
            if constexpr(FAST) {
                // e^{-i arg} = e^{arg.im} (cos arg.re - i sin arg.re)
                float sn, cs;
                sincos_phase((double)arg.re, sn, cs);
                float m = (float)(cnst * w) * expf((float)arg.im);
                fr = m * cs;
                fi = -(m * sn);
            } else {
                cpx<real> ex = cexp_(C(arg.im, -arg.re)); // std::exp(-J * arg)
                cpx<real> dd = rmul(cnst * w, ex);
                fr = (float)dd.re;
                fi = (float)dd.im;
            }
        } else {
            real v;
            if constexpr(FAST) v = cnst * (real)expf((float)(omega * delay.im));
            else v = cnst * std::exp(omega * delay.im);
            v = (v * v) * w;
            if(P.isGaussian) v *= std::sqrt(FL(2.0) * RC<real>::PI);
            fr = (float)v;
            fi = 0.0f;
        }
        sink.add_field(P, idx, fr, fi);
    } else if constexpr(RT == RT_ARR) {
        sink.add_arrival(idx, cnst * w, phaseInt, delay, d.SrcDecl, d.RcvrDecl, d.ntop, d.nbot, d.ialpha, d.isz,
                         d.is);
        if(P.run0 == 'V' || P.run0 == 'v') sink.add_eigen(d.is, ir, iz, d.isz, d.ialpha);
    } else if constexpr(RT == RT_EIG) {
        sink.add_eigen(d.is, ir, iz, d.isz, d.ialpha);
    }
    sink.count();
}

template<typename real> __host__ __device__ inline int RToIR(real r, const Prob2D<real> &P, const real *Rr)
{ // influence.hpp:82-89
    real temp = (r - Rr[0]) / P.Delta_r;
    if(std::fabs(temp - std::round(temp)) < RL(1e-6)) temp = std::round(temp);
    return imax(imin(trunc_int_x86(temp), P.NRr - 1), 0);
}

// Cartesian receiver (ir, iz) of a segment: influence.hpp:1298-1386 (2D).
// Receiver (ir, iz) at range rr = Rr[ir], already inside the segment's [lo, hi).
template<typename real, int RT, typename Sink, bool FAST = false>
__host__ __device__ inline void item_cart_rr(const SegDesc<real> &d, int ir, real rr, int iz, const Prob2D<real> &P,
                                             const float *Rz, Sink &sink)
{
    int tiz = (P.run4 == 'I') ? ir : iz;
    real zr = (real)Rz[tiz];
    if(zr < d.zmin || zr > d.zmax) return;
    v2<real> xrr = V(rr, zr) - d.xa;
    real s = dot(xrr, d.ta) / d.rlen;
    real n1 = std::fabs(dot(xrr, V(-d.ta.y, d.ta.x))); // the unit normal (rayn1)
    contrib<real, RT, Sink, FAST>(d, s, n1, ir, iz, P, sink);
}

template<typename real, int RT, typename Sink, bool FAST = false>
__host__ __device__ inline void item_cart_at(const SegDesc<real> &d, int ir, int iz, const Prob2D<real> &P,
                                             const real *G, const float *Rz, Sink &sink)
{
    real rr = G[P.o_Rr + ir];
    if(!(rr >= d.lo && rr < d.hi)) return;
    int tiz = (P.run4 == 'I') ? ir : iz;
    real zr = (real)Rz[tiz];
    if(zr < d.zmin || zr > d.zmax) return;
    v2<real> xrr = V(rr, zr) - d.xa;
    real s = dot(xrr, d.ta) / d.rlen;
    real n1 = std::fabs(dot(xrr, V(-d.ta.y, d.ta.x))); // the unit normal (rayn1)
    contrib<real, RT, Sink, FAST>(d, s, n1, ir, iz, P, sink);
}

// Cartesian work item k -> (ir, iz) of the segment's bounding box.
template<typename real, int RT, typename Sink, bool FAST = false>
__host__ __device__ inline void item_cart(const SegDesc<real> &d, int k, const Prob2D<real> &P, const real *G,
                                          const float *Rz, Sink &sink)
{
    int j = k / d.nz, iz = d.iz0 + (k - j * d.nz);
    item_cart_at<real, RT, Sink, FAST>(d, d.ir0 + j * d.dir, iz, P, G, Rz, sink);
}

// Cartesian column j (receiver range ir = ir0 + j dir) of a segment: the depth
// indices [i0, i1] that can pass the beam-window test. At range r the segment
// line is at depth z_l, the normal distance of depth z is n = |t_r| |z - z_l|
// (unit tangent t) and the receiver's projection is s = s_l + (z - z_l) t_z / L
// (L the segment length). The beam width obeys
// sigma <= |1/q0| (|q0| + |dq| |s|) (+ the Gaussian floor), so n < BW sigma
// requires |z - z_l| (|t_r| - BW |1/q0| |dq| |t_z| / L) < BW (|1/q0| (|q0| +
// |dq| |s_l|) + floor). Depths outside that band, widened for rounding, are
// skipped; the others go through every test of item_cart_at. Irregular grids,
// unsorted depths and steep segments (non-positive coefficient) keep the
// whole window.
template<typename real>
__host__ __device__ inline void cart_column(const SegDesc<real> &d, int j, const Prob2D<real> &P, const real *G,
                                            int &ir, int &i0, int &i1)
{
    ir = d.ir0 + j * d.dir;
    i0 = d.iz0;
    i1 = d.iz0 + d.nz - 1;
    if(!P.rz_sorted || P.run4 == 'I' || d.nz <= 2) return;
    const double tx = (double)d.ta.x, tz = (double)d.ta.y, atx = std::fabs(tx);
    const double BW = P.isGaussian ? 4.0 : 1.0;
    const double rcp = std::fabs((double)d.rcp_q0), dq = std::fabs((double)d.dqx), L = (double)d.rlen;
    const double coef = atx - BW * rcp * dq * std::fabs(tz) / L;
    if(!(atx > 1e-9) || !(coef > 1e-3 * atx)) return;
    const double rr = (double)G[P.o_Rr + ir], ra = (double)d.xa.x, za = (double)d.xa.y;
    const double zl = za + (rr - ra) * tz / tx;
    const double sl = ((rr - ra) * tx + (zl - za) * tz) / L;
    const double rhs =
        BW * (rcp * (std::fabs((double)d.q0x) + dq * std::fabs(sl)) + (P.isGaussian ? std::fmax((double)d.sigfloor, 0.0) : 0.0));
    const double dz = rhs / coef * (1.0 + 1e-4) + 1e-4 * (std::fabs(zl) + std::fabs(rr) + std::fabs(za)) + 1e-2;
    if(!(dz == dz) || !(zl == zl) || !(dz < 1e30)) return;
    int a, b;
    depth_window(P.Rz, P.NRz_per_range, P.rz_rdz, zl - dz, zl + dz, a, b);
    i0 = imax(i0, a);
    i1 = imin(i1, b);
}

// Ray-centred receiver depth iz (influence.hpp:1075-1187, 2D): the normals at
// A and B cross the depth at ranges rA, rB (normal distances nA, nB); the
// candidate receivers are ir = blo..bhi. Returns false when there are none.
template<typename real>
__host__ __device__ inline bool raycen_range(const SegDesc<real> &d, int iz, const Prob2D<real> &P, const real *Rr,
                                             const float *Rz, real &nA, real &rA, real &nB, real &rB, int &blo,
                                             int &bhi)
{
    real zR = (real)Rz[iz];
    int irA, irB;
    if(!d.lastValid) {
        nA = RL(1e10);
        rA = RL(1e10);
        irA = 0;
    } else {
        nA = (zR - d.xa.y) / d.ta.y;
        rA = d.xa.x + nA * d.ta.x;
        irA = RToIR(rA, P, Rr);
    }
    nB = (zR - d.xb.y) / d.tb.y;
    rB = d.xb.x + nB * d.tb.x;
    irB = RToIR(rB, P, Rr);
    if(irA == irB) return false;
    int notii = (irB <= irA) ? 0 : 1;
    int step = (notii << 1) - 1;
    // The reference visits ir = irA+notii ... irB+notii-step and rejects most of
    // them in the beam-window test. Receivers outside a conservative window
    // (|n| <= wmax, with slack) cannot pass that test and are not visited.
    blo = irA + notii;
    bhi = irB + notii - step;
    if(blo > bhi) {
        int t = blo;
        blo = bhi;
        bhi = t;
    }
    if(bhi - blo >= 4) {
        double dn = (double)nB - (double)nA;
        double W = (double)d.wmax + 1e-4 * (std::fabs((double)nA) + std::fabs((double)nB)) + 1e-30;
        if(dn == 0.0) {
            if(std::fabs((double)nA) > W) return false;
        } else {
            double s1 = (-W - (double)nA) / dn, s2 = (W - (double)nA) / dn;
            double slo = s1 < s2 ? s1 : s2, shi = s1 < s2 ? s2 : s1;
            double dr = (double)rB - (double)rA;
            double ra = (double)rA + slo * dr, rb = (double)rA + shi * dr;
            double rlo = ra < rb ? ra : rb, rhi = ra < rb ? rb : ra;
            double t0 = (rlo - (double)Rr[0]) / (double)P.Delta_r, t1 = (rhi - (double)Rr[0]) / (double)P.Delta_r;
            double lo_lim = (double)blo - 2.0, hi_lim = (double)bhi + 2.0;
            t0 = t0 < lo_lim ? lo_lim : (t0 > hi_lim ? hi_lim : t0);
            t1 = t1 < lo_lim ? lo_lim : (t1 > hi_lim ? hi_lim : t1);
            if(!(t0 == t0) || !(t1 == t1)) { // NaN: no pruning
                t0 = lo_lim;
                t1 = hi_lim;
            }
            int c0 = (int)std::floor(t0) - 1, c1 = (int)std::ceil(t1) + 1;
            blo = imax(blo, c0);
            bhi = imin(bhi, c1);
            if(blo > bhi) return false;
        }
    }
    return true;
}

// One ray-centred receiver (ir on depth iz with the crossing data above).
template<typename real, int RT, typename Sink, bool FAST = false>
__host__ __device__ inline void raycen_one(const SegDesc<real> &d, int ir, int iz, real nA, real rA, real nB, real rB,
                                           const Prob2D<real> &P, const real *Rr, Sink &sink)
{
    real s = (Rr[ir] - rA) / (rB - rA);
    real n1 = std::fabs(nA + s * (nB - nA));
    contrib<real, RT, Sink, FAST>(d, s, n1, ir, iz, P, sink);
}

// Ray-centred work item = one receiver depth.
template<typename real, int RT, typename Sink, bool FAST = false>
__host__ __device__ inline void item_raycen(const SegDesc<real> &d, int iz, const Prob2D<real> &P, const real *G,
                                            const float *Rz, Sink &sink)
{
    const real *Rr = G + P.o_Rr;
    real nA, rA, nB, rB;
    int blo, bhi;
    if(!raycen_range(d, iz, P, Rr, Rz, nA, rA, nB, rB, blo, bhi)) return;
    for(int ir = blo; ir <= bhi; ++ir) raycen_one<real, RT, Sink, FAST>(d, ir, iz, nA, rA, nB, rB, P, Rr, sink);
}

// RzT: the receiver depths as reals (the staged copy when the problem's
// tables sit in LDS).
template<typename real, bool WINDOW = true>
__host__ __device__ inline void prelude(SegDesc<real> &d, const RayPt<real> &p0, const RayPt<real> &p1, Infl<real> &I,
                                        int is, const Prob2D<real> &P, const real *G, const real *RzT, int isz,
                                        int ialpha)
{
    if(P.isRayCen) prelude_raycen(d, p0, p1, I, is, P, RzT, isz, ialpha);
    else prelude_cart<real, WINDOW>(d, p0, p1, I, is, P, G, RzT, isz, ialpha);
}

template<typename real, int RT, typename Sink, bool FAST = false>
__host__ __device__ inline void item(const SegDesc<real> &d, int k, const Prob2D<real> &P, const real *G,
                                     const float *Rz, Sink &sink)
{
    if(P.isRayCen) item_raycen<real, RT, Sink, FAST>(d, d.iz0 + k, P, G, Rz, sink);
    else item_cart<real, RT, Sink, FAST>(d, k, P, G, Rz, sink);
}

// Every receiver of one segment, in one thread: the Cartesian columns with
// their depth bands, or the ray-centred depths with their range windows.
template<typename real, int RT, typename Sink, bool FAST = false>
__host__ __device__ inline void seg_items(const SegDesc<real> &d, const Prob2D<real> &P, const real *G,
                                          const float *Rz, Sink &sink)
{
    if(d.nitems <= 0) return;
    if(P.isRayCen) {
        for(int k = 0; k < d.nz; ++k) item_raycen<real, RT, Sink, FAST>(d, d.iz0 + k, P, G, Rz, sink);
        return;
    }
    const int nr = (d.ir1 - d.ir0) * d.dir + 1;
    for(int j = 0; j < nr; ++j) {
        int ir, i0, i1;
        cart_column(d, j, P, G, ir, i0, i1);
        // the range test depends on the column only: once per column
        const real rr = G[P.o_Rr + ir];
        if(!(rr >= d.lo && rr < d.hi)) continue;
        for(int iz = i0; iz <= i1; ++iz) item_cart_rr<real, RT, Sink, FAST>(d, ir, rr, iz, P, Rz, sink);
    }
}

// Job index -> (isz, ialpha), reference src/common.hpp:365-391 (2D).
template<typename real> __host__ __device__ inline void JobIndices(int job, const Prob2D<real> &P, int &isz, int &ia)
{
    if(P.iSingle >= 1) {
        ia = P.iSingle - 1;
    } else {
        ia = job % P.Nalpha;
        job /= P.Nalpha;
    }
    isz = job;
}

} // namespace bhc2d
