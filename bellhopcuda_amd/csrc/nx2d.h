// nx2d.h -- Nx2D ray tracing (bellhopcxx Nx2D: O3D = true, R3D = false): 2D
// rays (range along a bearing, depth) in the vertical plane of each bearing of
// a 3D environment. The ray state, its integration and the influence are the
// 2D ones (core2d.h); the sound speed, the boundary cells and every step limit
// come from the 3D environment (core3d.h) at the ray's ocean point.
//
// Reference: the O3D && !R3D branches of src/trace.hpp:76-266 (RayInit),
// :274-381 (RayUpdate), :391-477 (RayTerminate); src/step.hpp:654-811 (Step,
// with the step end mapped back by OceanToRayX); src/ssp.hpp:449-462 (the
// projected SSP derivatives); src/reflect.hpp:94-140 + 252-442 (Reflect on the
// projected boundary normal); src/common_run.hpp:221-317 (Origin transforms).
#pragma once
#include "core3d.h"

namespace bhc3d {

// Origin<true, false> (include/bhc/structs.hpp:233-237): source position and
// the bearing's unit vector (cos beta, sin beta).
template<typename real> struct OrgNx {
    v3<real> xs;
    v2<real> tr;
};

template<typename real> __host__ __device__ inline v3<real> RayToOceanX(v2<real> x, const OrgNx<real> &o)
{ // common_run.hpp:221-231
    return V3(o.xs.x + x.x * o.tr.x, o.xs.y + x.x * o.tr.y, x.y);
}
template<typename real> __host__ __device__ inline v3<real> RayToOceanT(v2<real> t, const OrgNx<real> &o)
{ // common_run.hpp:233-242
    return V3(t.x * o.tr.x, t.x * o.tr.y, t.y);
}
template<typename real> __host__ __device__ inline v2<real> OceanToRayT(v3<real> t, const OrgNx<real> &o)
{ // common_run.hpp:308-317
    return V(dot(V(t.x, t.y), o.tr), t.z);
}

// The next representable value(s) by stepping the bit pattern (common.hpp:316-341).
__host__ __device__ inline double bits_add(double r, int i)
{
    return __builtin_bit_cast(double, __builtin_bit_cast(long long, r) + (long long)i);
}
__host__ __device__ inline float bits_add(float r, int i)
{
    return __builtin_bit_cast(float, __builtin_bit_cast(int, r) + i);
}

// OceanToRayX (common_run.hpp:253-306): range from the larger bearing
// component; when a step ended exactly on an x or y plane (snapDim 0 / 1, or
// -2 for a cell diagonal), up to four one-ulp moves along the ray until the
// round trip lands on the plane or just past it in the ray's direction.
template<typename real>
__host__ __device__ inline v2<real> OceanToRayX(v3<real> x, const OrgNx<real> &o, v2<real> t, int snapDim, Err &e)
{
    v2<real> ret;
    ret.y = x.z;
    if(std::fabs(o.tr.x) >= std::fabs(o.tr.y)) ret.x = (x.x - o.xs.x) / o.tr.x;
    else ret.x = (x.y - o.xs.y) / o.tr.y;
    if(snapDim < -2 || snapDim == -1 || snapDim >= 2) return ret;
    for(int i = 0; i < 4; ++i) {
        v3<real> xb = RayToOceanX(ret, o);
        v2<real> want = sc(t.x, o.tr);
        v2<real> err = V(xb.x - x.x, xb.y - x.y);
        bool okx = (want.x * err.x) >= RL(0.0);
        bool oky = (want.y * err.y) >= RL(0.0);
        if((snapDim == 0 && okx) || (snapDim == 1 && oky) || (okx && oky)) return ret;
        ret.x = bits_add(ret.x, (t.x > RL(0.0)) ? 1 : -1);
    }
    e.warning(18); // BHC_WARN_OCEANTORAYX_GAVEUP
    return ret;
}

// EvaluateSSP with O3D && !R3D (ssp.hpp:377-465): the 3D profile at the ocean
// point, its gradient and second derivatives projected on the bearing.
template<typename real>
__host__ __device__ inline void EvaluateSSPNx(v2<real> x, v2<real> t, SSPOut<real> &o, const Prob3D<real> &N,
                                              Trace3<real> &S, const OrgNx<real> &org, Err &e)
{
    SSPOut3<real> p;
    EvaluateSSP3(RayToOceanX(x, org), RayToOceanT(t, org), p, N, S, e);
    o.gradc.x = dot(org.tr, V(p.gradc.x, p.gradc.y));
    o.gradc.y = p.gradc.z;
    o.crz = org.tr.x * p.cxz + org.tr.y * p.cyz;
    o.crr = p.cxx * (org.tr.x * org.tr.x) + FL(2.0) * p.cxy * org.tr.x * org.tr.y + p.cyy * (org.tr.y * org.tr.y);
    o.c = p.c;
    o.rho = p.rho;
    o.czz = p.czz;
}

// Step (step.hpp:654-811, O3D && !R3D): the 2D polygon integration in ray
// coordinates; ReduceStep / StepToBdry run on the 3D ocean point and tangent
// (boundary cells, diagonals, hexahedral cells, beam box), and the step end
// comes back through OceanToRayX. Only depth-layer crossings get the
// curvature correction (step.hpp:801-806).
template<typename real>
__host__ __device__ inline void StepNx(const RayPt<real> &ray0, RayPt<real> &ray2, Trace3<real> &S,
                                       const OrgNx<real> &org, const Prob3D<real> &N, Err &e, bool &topRefl,
                                       bool &botRefl)
{
    SSPOut<real> o0, o1, o2;
    EvaluateSSPNx(ray0.x, ray0.t, o0, N, S, org, e);
    real c0 = o0.c.re;
    real cnn0 = o0.crr * (ray0.t.y * ray0.t.y) - FL(2.0) * o0.crz * ray0.t.x * ray0.t.y + o0.czz * (ray0.t.x * ray0.t.x);
    v2<real> pq0p = sc(-cnn0, ray0.q), pq0q = sc(c0, ray0.p);
    const int iSeg0[3] = {S.sx, S.sy, S.sz};
    real csq0 = c0 * c0;
    v2<real> u0 = sc(c0, ray0.t);
    real h = N.deltas;
    const v3<real> x_o = RayToOceanX(ray0.x, org);
    ReduceStep3(x_o, RayToOceanT(u0, org), iSeg0, S, N, e, h);
    real halfh = FL(0.5) * h;
    RayPt<real> ray1;
    ray1.x = ray0.x + sc(halfh, u0);
    ray1.t = ray0.t - dv(sc(halfh, o0.gradc), csq0);
    ray1.p = ray0.p + sc(halfh, pq0p);
    ray1.q = ray0.q + sc(halfh, pq0q);

    EvaluateSSPNx(ray1.x, ray1.t, o1, N, S, org, e);
    real c1 = o1.c.re;
    real cnn1 = o1.crr * (ray1.t.y * ray1.t.y) - FL(2.0) * o1.crz * ray1.t.x * ray1.t.y + o1.czz * (ray1.t.x * ray1.t.x);
    v2<real> pq1p = sc(-cnn1, ray1.q), pq1q = sc(c1, ray1.p);
    real csq1 = c1 * c1;
    v2<real> u1 = sc(c1, ray1.t);
    ReduceStep3(x_o, RayToOceanT(u1, org), iSeg0, S, N, e, h);

    real w1 = h / (RL(2.0) * halfh);
    real w0 = RL(1.0) - w1;
    v2<real> u2 = sc(w0, u0) + sc(w1, u1);
    v3<real> x2_o;
    int snap = -1;
    StepToBdry3(x_o, x2_o, RayToOceanT(u2, org), h, topRefl, botRefl, iSeg0, S, N, e, &snap);
    ray2.x = OceanToRayX(x2_o, org, u2, snap, e);
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

    EvaluateSSPNx(ray2.x, ray2.t, o2, N, S, org, e);
    ray2.c = o2.c.re;
    if(S.sz != iSeg0[2]) { // CurvatureCorrection, step.hpp:624-648 (crossing in depth)
        v2<real> gj = o2.gradc - o0.gradc;
        v2<real> ray2n = V(-ray2.t.y, ray2.t.x);
        real cnjump = dot(gj, ray2n);
        real csjump = dot(gj, ray2.t);
        real rm = ray2.t.x / ray2.t.y;
        real rn = rm * (FL(2.0) * cnjump - rm * csjump) / ray2.c;
        ray2.p = ray2.p - sc(rn, ray2.q);
    }
}

// Reflect with O3D && !R3D: the ocean normal projected on the bearing plane
// and renormalised, the boundary tangent rebuilt from the incident tangent
// (reflect.hpp:130-142), the 2D curvature update with the cell's curvature
// projected on the bearing (OceanToRayCurvature, reflect.hpp:94-114), the O3D
// halfspace coefficient and the 2D beam shift.
template<typename real>
__host__ __device__ inline void ReflectNx(const RayPt<real> &oldp, RayPt<real> &newp, const HSd<real> &hs, bool isTop,
                                          v3<real> nBdry, real kxx, real kxy, real kyy, const OrgNx<real> &org,
                                          const Prob3D<real> &N, const Prob2D<real> &P, Trace3<real> &S, Err &e,
                                          const real *tabs = nullptr)
{
    v2<real> nB = OceanToRayT(nBdry, org);
    nB = sc(RL(1.0) / std::sqrt(dot(nB, nB)), nB);
    real Th = dot(oldp.t, nB);
    v2<real> tB = oldp.t - sc(Th, nB);
    tB = sc(RL(1.0) / std::sqrt(dot(tB, tB)), tB);
    real Tg = dot(oldp.t, tB);
    real kappa = kxx * (org.tr.x * org.tr.x) + FL(2.0) * kxy * org.tr.x * org.tr.y + kyy * (org.tr.y * org.tr.y);
    if(isTop) kappa = -kappa;
    newp.NumTopBnc = oldp.NumTopBnc + (isTop ? 1 : 0);
    newp.NumBotBnc = oldp.NumBotBnc + (isTop ? 0 : 1);
    newp.x = oldp.x;
    newp.t = oldp.t - sc(FL(2.0) * Th, nB);
    SSPOut<real> o;
    EvaluateSSPNx(oldp.x, newp.t, o, N, S, org, e);
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
    } else if(hs.bc == 'F') {
        real r, phi;
        ReflTableO3D(refl_tab(N, isTop, tabs), isTop ? N.ntrc : N.nbrc, Th, Tg, r, phi, e);
        newp.Amp = oldp.Amp * r;
        newp.Phase = oldp.Phase + phi;
    } else if(hs.bc == 'A' || hs.bc == 'G') {
        real omega = FL(2.0) * RC<real>::PI * N.freq0;
        cpx<real> Refl = HalfspaceReflO3D(omega, Tg, c, hs, o.rho);
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

// Job index -> (ialpha, ibeta, isy, isx, isz) as in 3D (common.hpp:365-391),
// and the field slot of the ray's bearing plane, GetFieldAddr's
// ((isz * NSx + isx) * NSy + isy) * Ntheta + itheta with itheta = ibeta
// (influence.hpp:439).
template<typename real>
__host__ __device__ inline int JobIndicesNx(int job, const Prob3D<real> &N, int &ia, int &ib, int &isx, int &isy,
                                            int &isz)
{
    JobIndices3(job, N, ia, ib, isy, isx, isz);
    return ((isz * N.NSx + isx) * N.NSy + isy) * N.Ntheta + ib;
}

// RayInit (trace.hpp:51-67, 76-266, O3D && !R3D) and the 2D Init_Influence
// (influence.hpp:525-644): the nominal sound speed from the 3D profile at the
// source, the ray starting at range 0 on its bearing plane with p = (1, 0),
// q = (0, 1), and the boundary cells above and below the source.
template<typename real>
__host__ __device__ inline bool RayInitNx(int ia, int ib, int isx, int isy, int isz, RayPt<real> &p0,
                                          Trace3<real> &S, OrgNx<real> &org, Infl<real> &I, const Prob2D<real> &P,
                                          const Prob3D<real> &N, Err &e)
{
    const real *G = N.base;
    real alpha = G[N.o_alpha + ia];
    I.alpha = alpha;
    I.SrcDeclAngle = (RL(180.0) / RC<real>::PI) * alpha;
    S.sx = S.sy = S.sz = S.sr = 0;
    S.xs = V3(G[N.o_Sx + isx], G[N.o_Sy + isy], (real)N.Sz[isz]);
    SSPOut3<real> o;
    EvaluateSSP3(S.xs, V3((real)FL(0.0), (real)FL(0.0), (real)FL(1.0)), o, N, S, e);
    org.xs = S.xs;
    org.tr = V(G[N.o_bcos + ib], G[N.o_bsin + ib]);
    real c = o.c.re;
    const real *sbp = G + N.o_sbp;
    int ibp = BinarySearchLEQ(sbp, N.nsbp, 2, 0, I.SrcDeclAngle);
    ibp = imin(ibp, N.nsbp - 2);
    real s = (I.SrcDeclAngle - sbp[2 * ibp]) / (sbp[2 * (ibp + 1)] - sbp[2 * ibp]);
    float Amp0 = (float)((FL(1.0) - s) * sbp[2 * ibp + 1] + s * sbp[2 * (ibp + 1) + 1]);
    const real ca = G[N.o_acos + ia], sa = G[N.o_asin + ia];
    if(N.run0 == 'S') {
        float omega = (float)(FL(2.0) * RC<real>::PI * N.freq0);
        Amp0 = (float)(Amp0 * (sqrtf(FL(2.0)) * std::fabs(std::sin(omega / c * S.xs.z * sa))));
    }
    p0.x = V((real)RL(0.0), S.xs.z);
    p0.c = c;
    p0.t = dv(V(ca, sa), c);
    p0.tau = C((real)0, (real)0);
    p0.Amp = Amp0;
    p0.Phase = FL(0.0);
    p0.NumTopBnc = p0.NumBotBnc = 0;
    p0.p = V((real)FL(1.0), (real)FL(0.0));
    p0.q = V((real)FL(0.0), (real)FL(1.0));
    S.top = BState3<real>{};
    S.bot = BState3<real>{};
    const v3<real> t_o = RayToOceanT(p0.t, org);
    GetBdrySeg3(S.xs, t_o, S.top, G + N.o_ati, N.ati_nx, N.ati_ny, true, true, e, prov_of(N, true));
    GetBdrySeg3(S.xs, t_o, S.bot, G + N.o_bty, N.bty_nx, N.bty_ny, false, true, e, prov_of(N, false));
    real dt = -dot3(S.top.n, S.xs - S.top.x), db = -dot3(S.bot.n, S.xs - S.bot.x);
    if(dt <= FL(0.0) || db <= FL(0.0)) {
        e.warning(4);
        return false;
    }
    S.iSmall = 0;
    // Init_Influence, 2D geometric beams
    I.omega = FL(2.0) * RC<real>::PI * P.freq0;
    I.BeamWindow = P.isGaussian ? RL(4.0) : RL(1.0);
    I.Ratio1 = std::sqrt(std::fabs(ca));
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
        I.ir = BinarySearchGT(P.base + P.o_Rr, P.NRr, p0.x.x);
        if(p0.t.x < RL(0.0) && I.ir > 0) --I.ir;
        I.rbr = range_bracket(P.base + P.o_Rr, P.NRr, I.ir);
    }
    return true;
}

// RayUpdate (trace.hpp:274-381, O3D && !R3D).
template<typename real>
__host__ __device__ inline bool RayUpdateNx(const RayPt<real> &p0, RayPt<real> &p1, RayPt<real> &p2, Trace3<real> &S,
                                            const OrgNx<real> &org, const Prob3D<real> &N, const Prob2D<real> &P,
                                            Err &e, const real *tabs = nullptr)
{
    bool topRefl = false, botRefl = false;
    StepNx(p0, p1, S, org, N, e, topRefl, botRefl);
    const v3<real> x_o = RayToOceanX(p1.x, org), t_o = RayToOceanT(p1.t, org);
    GetBdrySeg3(x_o, t_o, S.top, N.base + N.o_ati, N.ati_nx, N.ati_ny, true, false, e, prov_of(N, true));
    GetBdrySeg3(x_o, t_o, S.bot, N.base + N.o_bty, N.bty_nx, N.bty_ny, false, false, e, prov_of(N, false));
    if(topRefl || botRefl) {
        ReflB3<real> r;
        refl_bdry3(N, topRefl, topRefl ? S.top : S.bot, x_o, r);
        ReflectNx(p1, p2, r.hs, topRefl, r.n, r.kxx, r.kxy, r.kyy, org, N, P, S, e, tabs);
        return true;
    }
    return false;
}

// RayTerminate (trace.hpp:391-477, O3D && !R3D): the 3D tests on the ocean
// point and tangent. Returns 0 to continue, 1 for a normal stop (Nsteps =
// is + 1), 2 when the per-ray storage is exhausted (Nsteps = is).
template<typename real>
__host__ __device__ inline int RayTerminateNx(const RayPt<real> &pt, int is, const Trace3<real> &S,
                                              const OrgNx<real> &org, const Prob3D<real> &N, Err &e, int maxPoints)
{
    Ray3<real> q;
    q.x = RayToOceanX(pt.x, org);
    q.t = RayToOceanT(pt.t, org);
    q.Amp = pt.Amp;
    int nsteps = 0;
    if(!RayTerminate3(q, is, S, N, e, maxPoints, nsteps)) return 0;
    return nsteps == is + 1 ? 1 : 2;
}

} // namespace bhc3d
