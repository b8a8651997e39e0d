// core3d.h -- per-ray physics of the full-3D tracer (3D rays, triangulated
// boundary grids, depth-only or hexahedral sound speed, Cartesian geometric
// beams with receiver bearings), written once for the device and the host.
//
// Trajectories (positions, tangents, step and bounce counts) follow the
// reference's 3D path operation for operation: they use only + - * / sqrt and
// the host-computed launch cos / sin, so they are compared bit for bit. The
// beam quantities (p, q, torsion phi) also pass through cos / sin / acos of
// the torsion, evaluated with the device's libm; they reach only the field.
//
// Vector / matrix conventions are those of glm, which the reference uses:
// column-major m[col][row], dot(a, b) = (a.x*b.x + a.y*b.y) + a.z*b.z,
// (A*B)[c][r] = A[0][r]*B[c][0] + A[1][r]*B[c][1].
#pragma once
#include "core2d.h"
#include "../../include/bhc_hip.h"

namespace bhc3d {

using namespace bhc2d;

template<typename real> struct v3 {
    real x, y, z;
};
template<typename real> __host__ __device__ inline v3<real> V3(real x, real y, real z) { return v3<real>{x, y, z}; }
template<typename real> __host__ __device__ inline v3<real> operator+(v3<real> a, v3<real> b)
{
    return V3(a.x + b.x, a.y + b.y, a.z + b.z);
}
template<typename real> __host__ __device__ inline v3<real> operator-(v3<real> a, v3<real> b)
{
    return V3(a.x - b.x, a.y - b.y, a.z - b.z);
}
template<typename real> __host__ __device__ inline v3<real> operator-(v3<real> a) { return V3(-a.x, -a.y, -a.z); }
template<typename real> __host__ __device__ inline v3<real> sc3(real s, v3<real> a)
{
    return V3(s * a.x, s * a.y, s * a.z);
}
template<typename real> __host__ __device__ inline v3<real> dv3(v3<real> a, real s)
{
    return V3(a.x / s, a.y / s, a.z / s);
}
template<typename real> __host__ __device__ inline real dot3(v3<real> a, v3<real> b)
{
    return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
template<typename real> __host__ __device__ inline real len3(v3<real> a) { return std::sqrt(dot3(a, a)); }
template<typename real> __host__ __device__ inline v3<real> cross3(v3<real> a, v3<real> b)
{
    return V3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}

template<typename real> struct m2 {
    real a00, a01, a10, a11; // a[col][row]
};
template<typename real> __host__ __device__ inline m2<real> M2(real a00, real a01, real a10, real a11)
{
    return m2<real>{a00, a01, a10, a11};
}
template<typename real> __host__ __device__ inline m2<real> mmul(m2<real> a, m2<real> b)
{
    return M2(a.a00 * b.a00 + a.a10 * b.a01, a.a01 * b.a00 + a.a11 * b.a01, a.a00 * b.a10 + a.a10 * b.a11,
              a.a01 * b.a10 + a.a11 * b.a11);
}
template<typename real> __host__ __device__ inline m2<real> madd(m2<real> a, m2<real> b)
{
    return M2(a.a00 + b.a00, a.a01 + b.a01, a.a10 + b.a10, a.a11 + b.a11);
}
template<typename real> __host__ __device__ inline m2<real> msub(m2<real> a, m2<real> b)
{
    return M2(a.a00 - b.a00, a.a01 - b.a01, a.a10 - b.a10, a.a11 - b.a11);
}
template<typename real> __host__ __device__ inline m2<real> mscl(real s, m2<real> a)
{
    return M2(s * a.a00, s * a.a01, s * a.a10, s * a.a11);
}
template<typename real> __host__ __device__ inline m2<real> mtr(m2<real> a) { return M2(a.a00, a.a10, a.a01, a.a11); }
template<typename real> __host__ __device__ inline real mdet(m2<real> a) { return a.a00 * a.a11 - a.a10 * a.a01; }

template<typename real> struct TD3;
template<> struct TD3<double> {
    static constexpr double THRESH = 3e-6; // TRIDIAG_THRESH, src/boundary.hpp:36-40
};
template<> struct TD3<float> {
    static constexpr float THRESH = 3e-3f;
};

// Boundary grid node in the blob: position, then the unit normals of the two
// triangles of the cell whose lower corner it is.
enum : int { N3_X = 0, N3_N1 = 3, N3_N2 = 6, NODE3 = 9 };
// Per-node extras of curvilinear ('C') and province ('L') grids: node normal,
// the cell's z and kappa curvatures, the cell's province (0-based, as a real).
enum : int { NX_N = 0, NX_ZXX = 3, NX_ZXY = 4, NX_ZYY = 5, NX_KXX = 6, NX_KXY = 7, NX_KYY = 8, NX_PROV = 9, NODEX = 10 };

// Device view of a 3D problem. Offsets index `base` (reals).
template<typename real> struct Prob3D {
    Prob2D<real> s;        // depth-only profile: base, ssp_type, npts, o_z, o_rho, o_c, o_cz, o_n2, o_n2z, o_coef
    const real *base;
    char run0, run1, ssp3; // RunType[0], RunType[1], SSP type ('H' hexahedral)
    char pad;
    int32_t nz3, o_z3, o_rho3;          // depth table and densities (after SegZToZ for 'H')
    int32_t hx, hy, o_hx, o_hy, o_hc, o_hcz;
    int32_t ati_nx, ati_ny, bty_nx, bty_ny, o_ati, o_bty;
    char ati_t0, ati_t1, bty_t0, bty_t1;    // grid types ('R' / 'C', 'S' / 'L')
    int32_t o_atix, o_btyx, o_atip, o_btyp; // per-node extras [NODEX], province tables [5] (cP, cS, rho)
    int32_t o_Rr, o_trc, o_brc, ntrc, nbrc, o_sbp, nsbp;
    int32_t o_alpha, o_acos, o_asin, o_beta, o_bcos, o_bsin, o_Sx, o_Sy, o_trcv;
    const float *Rz, *Sz;
    int32_t NSx, NSy, NSz, NRz, NRr, Ntheta, Nalpha, iSingleA, Nbeta, iSingleB;
    real Dalpha, Dbeta, freq0, deltas, box_x, box_y, box_z;
    real rz_absmax;        // largest |receiver depth|
    HSd<real> top, bot;
    int32_t isGaussian, rz_sorted;
    int32_t isRayCen, pad_rc;  // ray-centred beams ('g', 'b')
    real Delta_r;          // receiver range spacing (RToIR)
    double rz_rdz;         // depth-window index guess, as in the 2D problem
    float *field;
    int32_t njobs, stage_tables; // stage_tables: the scatter copies Rr and the bearing table to LDS
};

template<typename real> struct Ray3 {
    m2<real> p, q;
    real phi;
    v3<real> x, t;
    real c, Amp, Phase;
    cpx<real> tau;
    int32_t NumTopBnc, NumBotBnc;
};

template<typename real> struct BState3 {
    int32_t ix, iy;
    real xmin, xmax, ymin, ymax;
    v3<real> x, xmid, n;
    bool onEdge, side, justSteppedTo, outgoingSide;
    int32_t prov;           // 'L' grids: the current cell's province (CopyHSInfo, boundary.hpp:134-140)
};

template<typename real> struct SSPOut3 {
    v3<real> gradc;
    cpx<real> c;
    real rho, cxx, cyy, czz, cxy, cxz, cyz;
};

template<typename real> struct Trace3 {
    int32_t sx, sy, sz, sr;     // SSP segment state
    BState3<real> top, bot;
    v3<real> xs;
    int32_t iSmall;
};

// ---------------------------------------------------------------------------
// SSP: Hexahedral (src/ssp.hpp:220-307) and the 3D EvaluateSSP (:377-465)
// ---------------------------------------------------------------------------
template<typename real>
__host__ __device__ inline void EvaluateSSP3(v3<real> x, v3<real> t, SSPOut3<real> &o, const Prob3D<real> &P,
                                             Trace3<real> &S, Err &e)
{
    const real *B = P.base;
    if(P.ssp3 == 'H') {
        const real *sx = B + P.o_hx, *sy = B + P.o_hy, *z = B + P.o_z3;
        const int nx = P.hx, ny = P.hy, nz = P.nz3;
        if(x.x < sx[0] || x.x > sx[nx - 1] || x.y < sy[0] || x.y > sy[ny - 1]) e.error(8);
        UpdateSSPSegment(x.x, t.x, sx, nx, S.sx);
        UpdateSSPSegment(x.y, t.y, sy, ny, S.sy);
        UpdateSSPSegment(x.z, t.z, z, nz, S.sz);
        const int ix = S.sx, iy = S.sy, iz = S.sz;
        const real *cm = B + P.o_hc, *czm = B + P.o_hcz;
        size_t b11 = (size_t)ix * ny + iy, b21 = (size_t)(ix + 1) * ny + iy;
        size_t b12 = (size_t)ix * ny + iy + 1, b22 = (size_t)(ix + 1) * ny + iy + 1;
        real cz11 = czm[b11 * (nz - 1) + iz], cz12 = czm[b21 * (nz - 1) + iz];
        real cz21 = czm[b12 * (nz - 1) + iz], cz22 = czm[b22 * (nz - 1) + iz];
        real s3 = x.z - z[iz];
        real c11 = cm[b11 * nz + iz] + s3 * cz11;
        real c21 = cm[b21 * nz + iz] + s3 * cz12;
        real c12 = cm[b12 * nz + iz] + s3 * cz21;
        real c22 = cm[b22 * nz + iz] + s3 * cz22;
        real s1 = (x.x - sx[ix]) / (sx[ix + 1] - sx[ix]);
        s1 = rmax(rmin(s1, RL(1.0)), RL(0.0));
        real s2 = (x.y - sy[iy]) / (sy[iy + 1] - sy[iy]);
        s2 = rmax(rmin(s2, RL(1.0)), RL(0.0));
        real c1 = c11 + s1 * (c21 - c11);
        real c2 = c12 + s1 * (c22 - c12);
        real cy = (c2 - c1) / (sy[iy + 1] - sy[iy]);
        c1 = c11 + s2 * (c12 - c11);
        c2 = c21 + s2 * (c22 - c21);
        real c = c1 + s1 * (c2 - c1);
        // the attenuation term reads a profile that SegZToZ
        // (src/module/ssp.hpp:568-573) set to real NaNs: its imaginary part is +0
        o.c = C(c, (real)0);
        real cx = (c2 - c1) / (sx[ix + 1] - sx[ix]);
        real cz1 = cz11 + s2 * (cz21 - cz11);
        real cz2 = cz12 + s2 * (cz22 - cz12);
        real cz = cz1 + s1 * (cz2 - cz1);
        o.gradc = V3(cx, cy, cz);
        o.cxx = o.cyy = o.czz = o.cxy = o.cxz = o.cyz = RL(0.0);
        const real *rho = B + P.o_rho3;
        real w = (x.z - z[iz]) / (z[iz + 1] - z[iz]);
        o.rho = (RL(1.0) - w) * rho[iz] + w * rho[iz + 1];
        return;
    }
    SSPOut<real> o2;
    EvaluateSSP<real, 0>(V((real)RL(0.0), x.z), V((real)RL(0.0), t.z), o2, P.s, B, S.sz, S.sr, e);
    o.gradc = V3((real)RL(0.0), (real)RL(0.0), o2.gradc.y);
    o.cxx = o.cyy = o.cxy = o.cxz = o.cyz = RL(0.0);
    o.czz = o2.czz;
    o.c = o2.c;
    o.rho = o2.rho;
}

// ---------------------------------------------------------------------------
// Boundary grids (src/boundary.hpp:66-168, 3D)
// ---------------------------------------------------------------------------
template<typename real> __host__ __device__ inline v3<real> ld3(const real *p) { return V3(p[0], p[1], p[2]); }

template<typename real>
__host__ __device__ inline void GetBdrySeg3(v3<real> x, v3<real> t, BState3<real> &b, const real *bd, int nx, int ny,
                                            bool isTop, bool isInit, Err &e, const real *prov = nullptr)
{ // prov: the grid's per-node extras when it has provinces, else null
#define BX(i) bd[(size_t)(i) * NODE3 + 0]
#define BY(i) bd[(size_t)(i) * NODE3 + 1]
    b.ix = imin(imax(b.ix, 0), nx - 2);
    b.iy = imin(imax(b.iy, 0), ny - 2);
    if(t.x >= FL(0.0)) {
        while(b.ix >= 0 && BX(b.ix * ny) > x.x) --b.ix;
        while(b.ix >= 0 && b.ix < nx - 1 && BX((b.ix + 1) * ny) <= x.x) ++b.ix;
    } else {
        while(b.ix < nx - 1 && BX((b.ix + 1) * ny) < x.x) ++b.ix;
        while(b.ix >= 0 && b.ix < nx - 1 && BX(b.ix * ny) >= x.x) --b.ix;
    }
    if(t.y >= FL(0.0)) {
        while(b.iy >= 0 && BY(b.iy) > x.y) --b.iy;
        while(b.iy >= 0 && b.iy < ny - 1 && BY(b.iy + 1) <= x.y) ++b.iy;
    } else {
        while(b.iy < ny - 1 && BY(b.iy + 1) < x.y) ++b.iy;
        while(b.iy >= 0 && b.iy < ny - 1 && BY(b.iy) >= x.y) --b.iy;
    }
    if(b.ix == -1 && BX(0) == x.x) b.ix = 0;
    if(b.ix == nx - 1 && BX((nx - 1) * ny) == x.x) b.ix = nx - 2;
    if(b.iy == -1 && BY(0) == x.y) b.iy = 0;
    if(b.iy == ny - 1 && BY(ny - 1) == x.y) b.iy = ny - 2;
    if(b.ix < 0 || b.ix >= nx - 1 || b.iy < 0 || b.iy >= ny - 1) {
        e.warning(isTop ? 6 : 7);
        b.ix = imin(imax(b.ix, 0), nx - 2);
        b.iy = imin(imax(b.iy, 0), ny - 2);
    }
    b.xmin = BX(b.ix * ny);
    b.xmax = BX((b.ix + 1) * ny);
    b.ymin = BY(b.iy);
    b.ymax = BY(b.iy + 1);
#undef BX
#undef BY
    const real *n0 = bd + ((size_t)b.ix * ny + b.iy) * NODE3;
    const real *n11 = bd + ((size_t)(b.ix + 1) * ny + (b.iy + 1)) * NODE3;
    b.x = ld3(n0);
    if(prov) b.prov = (int32_t)prov[((size_t)b.ix * ny + b.iy) * NODEX + NX_PROV];
    v3<real> s = b.x + ld3(n11);
    b.xmid = V3(s.x * RL(0.5), s.y * RL(0.5), s.z * RL(0.5));
    v2<real> tri_n = V(-(b.ymax - b.ymin), b.xmax - b.xmin);
    tri_n = dv(tri_n, std::sqrt(dot(tri_n, tri_n)));
    v2<real> temp = V(x.x - b.xmid.x, x.y - b.xmid.y);
    real over = dot(temp, tri_n);
    b.onEdge = std::fabs(over) < TD3<real>::THRESH;
    if(!isInit && b.justSteppedTo) b.side = b.outgoingSide;
    else if(!isInit && b.onEdge) b.side = dot(V(t.x, t.y), tri_n) >= RL(0.0);
    else b.side = over >= RL(0.0);
    b.justSteppedTo = false;
    b.n = ld3(n0 + (b.side ? N3_N2 : N3_N1));
    if(!std::isfinite(b.x.z) || !std::isfinite(b.n.x) || !std::isfinite(b.n.y) || !std::isfinite(b.n.z)) e.error(3);
}

// ---------------------------------------------------------------------------
// Step (src/step.hpp, 3D instantiation)
// ---------------------------------------------------------------------------
template<typename real> __host__ __device__ inline real cmp3(v3<real> a, int d) { return d == 0 ? a.x : (d == 1 ? a.y : a.z); }
template<typename real> __host__ __device__ inline void set3(v3<real> &a, int d, real v)
{
    if(d == 0) a.x = v;
    else if(d == 1) a.y = v;
    else a.z = v;
}

template<typename real>
__host__ __device__ inline void DepthInterfaceCrossing3(bool stepTo, real &hInt, v3<real> &x, v3<real> x0,
                                                        v3<real> u, int iz, const Prob3D<real> &P,
                                                        int *snap = nullptr)
{ // step.hpp:41-80 (3D: the top and bottom layers extrapolate); snap: the
  // snapDim of OceanToRayX (Nx2D), set where the step is cut
    if(!stepTo) hInt = RC<real>::MAXV;
    if(std::fabs(u.z) > RC<real>::EPS) {
        const real *z = P.base + P.o_z3;
        if(z[iz] > x.z && iz > 0) {
            hInt = (z[iz] - x0.z) / u.z;
            if(stepTo) {
                x = x0 + sc3(hInt, u);
                x.z = z[iz];
                if(snap) *snap = 2;
            }
        } else if(z[iz + 1] < x.z && iz + 1 < P.nz3 - 1) {
            hInt = (z[iz + 1] - x0.z) / u.z;
            if(stepTo) {
                x = x0 + sc3(hInt, u);
                x.z = z[iz + 1];
                if(snap) *snap = 2;
            }
        }
    }
}

template<typename real>
__host__ __device__ inline void BeamBoxCrossing3(bool stepTo, int dim, real &hBox, v3<real> &x, v3<real> x0,
                                                 v3<real> u, const Prob3D<real> &P, v3<real> xs, int *snap = nullptr)
{ // step.hpp:85-111, common_run.hpp:33-47
    if(!stepTo) hBox = RC<real>::MAXV;
    v3<real> ctr = V3(xs.x, xs.y, (real)RL(0.0));
    real box = dim == 0 ? P.box_x : (dim == 1 ? P.box_y : P.box_z);
    if(std::fabs(cmp3(x, dim) - cmp3(ctr, dim)) >= box) {
        real d0 = cmp3(x0, dim) - cmp3(ctr, dim);
        hBox = (box - std::fabs(d0)) / std::fabs(cmp3(u, dim));
        if(stepTo) {
            x = x0 + sc3(hBox, u);
            set3(x, dim, cmp3(ctr, dim) + std::copysign(box, d0));
            if(snap) *snap = dim;
        }
    }
}

template<typename real>
__host__ __device__ inline void TopBotCrossing3(bool stepTo, real &h, const BState3<real> &bd, v3<real> &x,
                                                v3<real> x0, v3<real> u, bool &refl, int *snap = nullptr)
{ // step.hpp:116-151
    if(!stepTo) h = RC<real>::MAXV;
    v3<real> d = x - bd.x, d0 = x0 - bd.x;
    real w = dot3(bd.n, d);
    if(stepTo ? (w > -RC<real>::INF_STEP) : (w >= RL(0.0))) {
        h = -dot3(d0, bd.n) / dot3(u, bd.n);
        if(stepTo) {
            x = x0 + sc3(h, u);
            if(std::fabs(bd.n.x) < RC<real>::EPS && std::fabs(bd.n.y) < RC<real>::EPS) x.z = bd.x.z;
            if(snap) *snap = 2;
        }
        refl = true;
    } else {
        refl = false;
    }
}

template<typename real>
__host__ __device__ inline void TopBotSegCrossing3(bool stepTo, bool isY, real &hSeg, const Trace3<real> &S,
                                                   int iSeg, v3<real> &x, v3<real> x0, v3<real> u,
                                                   const Prob3D<real> &P, bool &topRefl, bool &botRefl,
                                                   int *snap = nullptr)
{ // step.hpp:157-207
    real lmin = rmax(isY ? S.top.ymin : S.top.xmin, isY ? S.bot.ymin : S.bot.xmin);
    real lmax = rmin(isY ? S.top.ymax : S.top.xmax, isY ? S.bot.ymax : S.bot.xmax);
    if(P.ssp3 == 'H') {
        const real *sw = P.base + (isY ? P.o_hy : P.o_hx);
        lmin = rmax(lmin, sw[iSeg]);
        lmax = rmin(lmax, sw[iSeg + 1]);
    }
    real xw = isY ? x.y : x.x, x0w = isY ? x0.y : x0.x, uw = isY ? u.y : u.x;
    if(!stepTo) hSeg = RC<real>::MAXV;
    if(std::fabs(uw) > RC<real>::EPS) {
        if(xw < lmin) {
            hSeg = -(x0w - lmin) / uw;
            if(stepTo) {
                x = x0 + sc3(hSeg, u);
                if(isY) x.y = lmin;
                else x.x = lmin;
                if(snap) *snap = isY ? 1 : 0;
                topRefl = botRefl = false;
            }
        } else if(xw > lmax) {
            hSeg = -(x0w - lmax) / uw;
            if(stepTo) {
                x = x0 + sc3(hSeg, u);
                if(isY) x.y = lmax;
                else x.x = lmax;
                if(snap) *snap = isY ? 1 : 0;
                topRefl = botRefl = false;
            }
        }
    }
}

template<typename real>
__host__ __device__ inline void TriDiagCrossing(bool stepTo, real &h, BState3<real> &bd, v3<real> &x, v3<real> x0,
                                                v3<real> u, bool &topRefl, bool &botRefl, Err &e,
                                                int *snap = nullptr)
{ // step.hpp:212-260
    if(!stepTo) h = RC<real>::MAXV;
    if(bd.onEdge) return;
    v3<real> d = x - bd.xmid, d0 = x0 - bd.xmid;
    v3<real> tri_n = V3(-(bd.ymax - bd.ymin), bd.xmax - bd.xmin, (real)RL(0.0));
    tri_n = dv3(tri_n, len3(tri_n));
    real over = dot3(d, tri_n);
    bool newSide = over >= RL(0.0);
    if(newSide == bd.side) return;
    if(std::fabs(over) < TD3<real>::THRESH) return;
    real hnew = -dot3(d0, tri_n) / dot3(u, tri_n);
    if(hnew < RL(0.0)) {
        e.warning(9);
        h = RL(0.0);
    } else {
        if(stepTo && hnew >= h) e.warning(10);
        h = hnew;
    }
    if(stepTo) {
        x = x0 + sc3(h, u);
        bd.justSteppedTo = true;
        bd.outgoingSide = newSide;
        if(snap) *snap = -2;
        topRefl = botRefl = false;
    }
}

template<typename real>
__host__ __device__ inline void ReduceStep3(v3<real> x0, v3<real> u, const int iSeg0[3], Trace3<real> &S,
                                            const Prob3D<real> &P, Err &e, real &h)
{ // step.hpp:271-337
    real hInt, hBx, hBy, hBz, hTop, hBot, hxS, hyS, hTD, hBD;
    bool dummy = false;
    v3<real> x = x0 + sc3(h, u);
    DepthInterfaceCrossing3(false, hInt, x, x0, u, iSeg0[2], P);
    BeamBoxCrossing3(false, 0, hBx, x, x0, u, P, S.xs);
    BeamBoxCrossing3(false, 1, hBy, x, x0, u, P, S.xs);
    BeamBoxCrossing3(false, 2, hBz, x, x0, u, P, S.xs);
    TopBotCrossing3(false, hTop, S.top, x, x0, u, dummy);
    TopBotCrossing3(false, hBot, S.bot, x, x0, u, dummy);
    TopBotSegCrossing3(false, false, hxS, S, iSeg0[0], x, x0, u, P, dummy, dummy);
    TopBotSegCrossing3(false, true, hyS, S, iSeg0[1], x, x0, u, P, dummy, dummy);
    TriDiagCrossing(false, hTD, S.top, x, x0, u, dummy, dummy, e);
    TriDiagCrossing(false, hBD, S.bot, x, x0, u, dummy, dummy, e);
    h = rmin(rmin(rmin(h, hInt), rmin(rmin(rmin(hBx, hBy), hBz), rmin(hTop, hBot))), rmin(rmin(hxS, hyS), rmin(hTD, hBD)));
    if(h < RL(-1e-4)) e.warning(8);
    if(h < RC<real>::INF_STEP * P.deltas) {
        h = RC<real>::INF_STEP * P.deltas;
        ++S.iSmall;
    } else {
        S.iSmall = 0;
    }
}

template<typename real>
__host__ __device__ inline void StepToBdry3(v3<real> x0, v3<real> &x2, v3<real> u, real &h, bool &topRefl,
                                            bool &botRefl, const int iSeg0[3], Trace3<real> &S, const Prob3D<real> &P,
                                            Err &e, int *snapDim = nullptr)
{ // step.hpp:342-420; snapDim (Nx2D): which ocean coordinate the step end
  // was put on exactly (common_run.hpp:243-252)
    int snap = -1;
    int *sp = snapDim ? &snap : nullptr;
    x2 = x0 + sc3(h, u);
    DepthInterfaceCrossing3(true, h, x2, x0, u, iSeg0[2], P, sp);
    BeamBoxCrossing3(true, 0, h, x2, x0, u, P, S.xs, sp);
    BeamBoxCrossing3(true, 1, h, x2, x0, u, P, S.xs, sp);
    BeamBoxCrossing3(true, 2, h, x2, x0, u, P, S.xs, sp);
    TopBotCrossing3(true, h, S.top, x2, x0, u, topRefl, sp);
    TopBotCrossing3(true, h, S.bot, x2, x0, u, botRefl, sp);
    if(botRefl) topRefl = false;
    TopBotSegCrossing3(true, false, h, S, iSeg0[0], x2, x0, u, P, topRefl, botRefl, sp);
    TopBotSegCrossing3(true, true, h, S, iSeg0[1], x2, x0, u, P, topRefl, botRefl, sp);
    TriDiagCrossing(true, h, S.top, x2, x0, u, topRefl, botRefl, e, sp);
    TriDiagCrossing(true, h, S.bot, x2, x0, u, topRefl, botRefl, e, sp);
    if(h < RC<real>::INF_STEP * P.deltas) {
        h = RC<real>::INF_STEP * P.deltas;
        x2 = x0 + sc3(h, u);
        snap = -1;
        if(dot3(S.top.n, x2 - S.top.x) > RC<real>::EPS) {
            topRefl = true;
            S.top.justSteppedTo = false;
            S.bot.justSteppedTo = false;
            snap = 2;
        } else {
            topRefl = false;
        }
        if(dot3(S.bot.n, x2 - S.bot.x) > RC<real>::EPS) {
            botRefl = true;
            topRefl = false;
            S.top.justSteppedTo = false;
            S.bot.justSteppedTo = false;
            snap = 2;
        } else {
            botRefl = false;
        }
    }
    if(snapDim) *snapDim = snap;
}

// The torsion's trigonometry (RayNormal, Reflect). A CPU build of the core
// that defines BHC_TORSION_ULP_TEST moves every single-precision result by
// one ulp (up or down, chosen from the argument's bits) to bound, on the
// host, what one-ulp differences between libm implementations do to FP32
// fields; the product never defines it.
template<typename real> __host__ __device__ inline real torsion_ulp(real v, real arg)
{
#ifdef BHC_TORSION_ULP_TEST
    if constexpr(sizeof(real) == 4) {
        uint32_t b;
        __builtin_memcpy(&b, &arg, 4);
        b ^= b >> 13;
        b *= 0x5bd1e995u;
        b ^= b >> 15;
        return std::nextafter(v, (b & 1u) ? (real)1.0e30 : (real)-1.0e30);
    }
#endif
    (void)arg;
    return v;
}
template<typename real> __host__ __device__ inline real torsion_cos(real a) { return torsion_ulp(std::cos(a), a); }
template<typename real> __host__ __device__ inline real torsion_sin(real a) { return torsion_ulp(std::sin(a), a); }
template<typename real> __host__ __device__ inline real torsion_acos(real a) { return torsion_ulp(std::acos(a), a); }

// RayNormal / RayNormal_unit, src/common_run.hpp:144-215
template<typename real>
__host__ __device__ inline void RayNormal3(v3<real> t, real phi, bool ignorephi0, real c, v3<real> &e1, v3<real> &e2)
{
    real rl = std::sqrt(t.x * t.x + t.y * t.y);
    if(std::fabs(phi) >= RC<real>::EPS || ignorephi0) {
        real cosphi = torsion_cos(phi), sinphi = torsion_sin(phi);
        e1.x = (c * t.x * t.z * cosphi + t.y * sinphi) / rl;
        e1.y = (c * t.y * t.z * cosphi - t.x * sinphi) / rl;
        e1.z = -c * rl * cosphi;
        e2.x = (c * t.x * t.z * sinphi - t.y * cosphi) / rl;
        e2.y = (c * t.y * t.z * sinphi + t.x * cosphi) / rl;
        e2.z = -c * rl * sinphi;
    } else {
        e1 = V3(c * t.x * t.z / rl, c * t.y * t.z / rl, -c * rl);
        e2 = V3(-t.y / rl, t.x / rl, (real)RL(0.0));
    }
}

template<typename real> struct PQ3 {
    real phi;
    m2<real> p, q;
};

template<typename real>
__host__ __device__ inline PQ3<real> DeltaPQ3(const Ray3<real> &ray, const SSPOut3<real> &o)
{ // Get_c_partials + ComputeDeltaPQ, step.hpp:425-473 (3D)
    const real c = o.c.re;
    v3<real> e1, e2;
    RayNormal3(ray.t, ray.phi, false, c, e1, e2);
    real cnn = o.cxx * (e1.x * e1.x) + o.cyy * (e1.y * e1.y) + o.czz * (e1.z * e1.z) + FL(2.0) * o.cxy * e1.x * e1.y
        + FL(2.0) * o.cxz * e1.x * e1.z + FL(2.0) * o.cyz * e1.y * e1.z;
    real cmn = o.cxx * e1.x * e2.x + o.cyy * e1.y * e2.y + o.czz * e1.z * e2.z + o.cxy * (e1.x * e2.y + e2.x * e1.y)
        + o.cxz * (e1.x * e2.z + e2.x * e1.z) + o.cyz * (e1.y * e2.z + e2.y * e1.z);
    real cmm = o.cxx * (e2.x * e2.x) + o.cyy * (e2.y * e2.y) + o.czz * (e2.z * e2.z) + FL(2.0) * o.cxy * e2.x * e2.y
        + FL(2.0) * o.cxz * e2.x * e2.z + FL(2.0) * o.cyz * e2.y * e2.z;
    PQ3<real> pq;
    pq.phi = (FL(1.0) / c) * ray.t.z * (ray.t.y * o.gradc.x - ray.t.x * o.gradc.y)
        / (ray.t.x * ray.t.x + ray.t.y * ray.t.y);
    m2<real> cm = mscl(-RL(1.0) / (c * c), M2(cnn, cmn, cmn, cmm));
    pq.p = mmul(ray.q, mtr(cm));
    pq.q = mscl(c, ray.p);
    return pq;
}

template<typename real>
__host__ __device__ inline void UpdatePQ3(Ray3<real> &r1, const Ray3<real> &r0, real h, const PQ3<real> &pq)
{ // step.hpp:475-481
    real phi = r0.phi + h * pq.phi;
    m2<real> p = madd(r0.p, mscl(h, pq.p));
    m2<real> q = madd(r0.q, mscl(h, pq.q));
    r1.phi = phi;
    r1.p = p;
    r1.q = q;
}

template<typename real>
__host__ __device__ inline void CurvCorr3D(Ray3<real> &ray, m2<real> D, real Tg, real Th, real cn1, real cn2, real cs,
                                           v3<real> rayn1, v3<real> rayn2, v3<real> e1, v3<real> e2, bool reflect)
{ // step.hpp:486-572
    real rm = Tg / Th;
    real csq = ray.c * ray.c;
    m2<real> r;
    r.a00 = FL(2.0) / csq * D.a00 / Th + rm * (FL(2.0) * cn1 - rm * cs) / csq;
    r.a01 = FL(2.0) / ray.c * D.a10 * std::copysign(RL(1.0), -Th) + rm * cn2 / csq;
    r.a10 = r.a01;
    r.a11 = FL(2.0) * D.a11 * Th;
    if(reflect) r.a10 = -r.a10;
    else r = mscl(RL(-1.0), r);
    bool noChange = false;
    if(reflect)
        noChange = std::fabs(r.a00) < RC<real>::EPS && std::fabs(r.a01) < RC<real>::EPS
            && std::fabs(r.a11) < RC<real>::EPS;
    if(!noChange) {
        m2<real> R;
        R.a00 = dot3(rayn1, e1);
        R.a10 = dot3(rayn1, e2);
        R.a01 = -R.a10;
        R.a11 = dot3(rayn2, e2);
        m2<real> p_in = mmul(R, ray.p);
        m2<real> q_in = mmul(R, ray.q);
        m2<real> p_out = madd(p_in, mmul(r, q_in));
        ray.p = mmul(mtr(R), p_out);
    }
    if(reflect) {
        real d = dot3(rayn1, e1);
        ray.phi += FL(2.0) * torsion_acos(rmax(rmin(d, RL(1.0)), RL(-1.0)));
    }
}

template<typename real>
__host__ __device__ inline void TangentNormals3(const Ray3<real> &ray, real c, v3<real> nB, v3<real> &rayt,
                                                v3<real> &rayn1, v3<real> &rayn2, real sgn)
{ // CalcTangent_Normals, step.hpp:574-584
    rayt = sc3(c, ray.t);
    rayn2 = sc3(sgn, cross3(rayt, nB));
    rayn2 = dv3(rayn2, len3(rayn2));
    rayn1 = -cross3(rayt, rayn2);
}

template<typename real>
__host__ __device__ inline void CurvCorrSSP3(Ray3<real> &ray, v3<real> gj, bool dz, bool dx)
{ // CurvatureCorrection, step.hpp:589-627 (3D)
    v3<real> nB = V3((real)RL(0.0), (real)RL(0.0), (real)RL(0.0));
    if(dz) nB.z = -std::copysign(RL(1.0), ray.t.z);
    else if(dx) nB.x = -std::copysign(RL(1.0), ray.t.x);
    else nB.y = -std::copysign(RL(1.0), ray.t.y);
    real Th = dot3(ray.t, nB);
    v3<real> tB = ray.t - sc3(Th, nB);
    tB = dv3(tB, len3(tB));
    real Tg = dot3(ray.t, tB);
    v3<real> rayt, rayn1, rayn2, e1, e2;
    TangentNormals3(ray, ray.c, nB, rayt, rayn1, rayn2, RL(1.0));
    RayNormal3(rayt, ray.phi, true, (real)RL(1.0), e1, e2);
    CurvCorr3D(ray, M2((real)RL(0.0), (real)RL(0.0), (real)RL(0.0), (real)RL(0.0)), Tg, Th, dot3(gj, rayn1),
               dot3(gj, rayn2), dot3(gj, rayt), rayn1, rayn2, e1, e2, false);
}

template<typename real>
__host__ __device__ inline void Step3(const Ray3<real> &ray0, Ray3<real> &ray2, Trace3<real> &S,
                                      const Prob3D<real> &P, Err &e, bool &topRefl, bool &botRefl)
{ // step.hpp:654-811 (3D)
    Ray3<real> ray1;
    SSPOut3<real> o0, o1, o2;
    EvaluateSSP3(ray0.x, ray0.t, o0, P, S, e);
    PQ3<real> pq0 = DeltaPQ3(ray0, o0);
    const int iSeg0[3] = {S.sx, S.sy, S.sz};
    real c0 = o0.c.re, csq0 = c0 * c0;
    v3<real> u0 = sc3(c0, ray0.t);
    real h = P.deltas;
    ReduceStep3(ray0.x, u0, iSeg0, S, P, e, h);
    real halfh = FL(0.5) * h;
    ray1.x = ray0.x + sc3(halfh, u0);
    ray1.t = ray0.t - dv3(sc3(halfh, o0.gradc), csq0);
    UpdatePQ3(ray1, ray0, halfh, pq0);
    EvaluateSSP3(ray1.x, ray1.t, o1, P, S, e);
    PQ3<real> pq1 = DeltaPQ3(ray1, o1);
    real c1 = o1.c.re, csq1 = c1 * c1;
    v3<real> u1 = sc3(c1, ray1.t);
    ReduceStep3(ray0.x, u1, iSeg0, S, P, e, h);
    real w1 = h / (RL(2.0) * halfh);
    real w0 = RL(1.0) - w1;
    v3<real> u2 = sc3(w0, u0) + sc3(w1, u1);
    v3<real> x2;
    StepToBdry3(ray0.x, x2, u2, h, topRefl, botRefl, iSeg0, S, P, e);
    ray2.x = x2;
    real hw0 = h * w0, hw1 = h * w1;
    ray2.t = (ray0.t - dv3(sc3(hw0, o0.gradc), csq0)) - dv3(sc3(hw1, o1.gradc), csq1);
    ray2.tau = ray0.tau + realdiv(hw0, o0.c) + realdiv(hw1, o1.c);
    UpdatePQ3(ray2, ray0, hw0, pq0);
    UpdatePQ3(ray2, ray2, hw1, pq1);
    ray2.Amp = ray0.Amp;
    ray2.Phase = ray0.Phase;
    ray2.NumTopBnc = ray0.NumTopBnc;
    ray2.NumBotBnc = ray0.NumBotBnc;
    EvaluateSSP3(ray2.x, ray2.t, o2, P, S, e);
    ray2.c = o2.c.re;
    if(S.sz != iSeg0[2] || S.sx != iSeg0[0] || S.sy != iSeg0[1])
        CurvCorrSSP3(ray2, o2.gradc - o0.gradc, S.sz != iSeg0[2], S.sx != iSeg0[0]);
}

// ---------------------------------------------------------------------------
// Reflection (src/reflect.hpp:125-452, 3D, rectilinear boundaries)
// ---------------------------------------------------------------------------
// The grids' extras as GetBdrySeg3 takes them (null unless the grid has provinces).
template<typename real> __host__ __device__ inline const real *prov_of(const Prob3D<real> &P, bool isTop)
{
    return (isTop ? P.ati_t1 : P.bty_t1) == 'L' ? P.base + (isTop ? P.o_atix : P.o_btyx) : nullptr;
}

// The reflecting boundary's data at the ray's ocean point x (trace.hpp:309-360,
// O3D): the halfspace, its cP, cS, rho replaced by the cell's province on 'L'
// grids; on 'C' grids the node normals interpolated bilinearly over the cell
// (not renormalised) and the cell's curvatures, else the triangle's normal and
// zero curvature.
template<typename real> struct ReflB3 {
    HSd<real> hs;
    v3<real> n;
    real zxx, zxy, zyy, kxx, kxy, kyy;
};
template<typename real>
__host__ __device__ inline void refl_bdry3(const Prob3D<real> &P, bool isTop, const BState3<real> &b, v3<real> x,
                                           ReflB3<real> &r)
{
    r.hs = isTop ? P.top : P.bot;
    if((isTop ? P.ati_t1 : P.bty_t1) == 'L') {
        const real *pv = P.base + (isTop ? P.o_atip : P.o_btyp) + 5 * b.prov;
        r.hs.cP = C(pv[0], pv[1]);
        r.hs.cS = C(pv[2], pv[3]);
        r.hs.rho = pv[4];
    }
    r.n = b.n;
    r.zxx = r.zxy = r.zyy = r.kxx = r.kxy = r.kyy = FL(0.0);
    if((isTop ? P.ati_t0 : P.bty_t0) == 'C') {
        const int ny = isTop ? P.ati_ny : P.bty_ny;
        const real *X = P.base + (isTop ? P.o_atix : P.o_btyx);
        const size_t k00 = (size_t)b.ix * ny + b.iy, k01 = k00 + 1, k10 = k00 + ny, k11 = k10 + 1;
        real s1 = (x.x - b.x.x) / (b.xmax - b.xmin);
        real s2 = (x.y - b.x.y) / (b.ymax - b.ymin);
        real m1 = FL(1.0) - s1, m2 = FL(1.0) - s2;
        r.n = ((sc3(m2, sc3(m1, ld3(X + k00 * NODEX))) + sc3(m2, sc3(s1, ld3(X + k10 * NODEX))))
               + sc3(s2, sc3(s1, ld3(X + k11 * NODEX))))
            + sc3(s2, sc3(m1, ld3(X + k01 * NODEX)));
        const real *c = X + k00 * NODEX;
        r.zxx = c[NX_ZXX];
        r.zxy = c[NX_ZXY];
        r.zyy = c[NX_ZYY];
        r.kxx = c[NX_KXX];
        r.kxy = c[NX_KXY];
        r.kyy = c[NX_KYY];
    }
}

// InterpolateReflectionCoefficient (reflect.hpp:39-92) at the incidence angle
// |atan2(Th, Tg)| folded about 90 degrees; tab holds (theta, |R|, phase) rows.
// The reflection-coefficient table of a boundary: in `tabs`, a staged copy
// of the top then the bottom table (stage_refl_tables), when a kernel made
// one, else in the problem's blob.
template<typename real>
__host__ __device__ inline const real *refl_tab(const Prob3D<real> &P, bool isTop, const real *tabs)
{
    if(tabs) return tabs + (isTop ? 0 : 3 * P.ntrc);
    return P.base + (isTop ? P.o_trc : P.o_brc);
}

// Copy the top and bottom reflection tables into `lds` (cap reals) when they
// fit, for Reflect to read them from LDS; returns the copy, or null. Called
// by every thread of the workgroup (it synchronises it).
template<typename real> __device__ inline const real *stage_refl_tables(const Prob3D<real> &P, real *lds, int cap)
{
    const int nt = 3 * P.ntrc, nb = 3 * P.nbrc;
    if(nt + nb <= 0 || nt + nb > cap) return nullptr;
    for(int i = threadIdx.x; i < nt; i += blockDim.x) lds[i] = P.base[P.o_trc + i];
    for(int i = threadIdx.x; i < nb; i += blockDim.x) lds[nt + i] = P.base[P.o_brc + i];
    __syncthreads();
    return lds;
}

template<typename real>
__host__ __device__ inline void ReflTableO3D(const real *tab, int n, real Th, real Tg, real &r, real &phi, Err &e)
{
    real theta = (RL(180.0) / RC<real>::PI) * std::fabs(std::atan2(Th, Tg));
    if(theta > FL(90.0)) theta = FL(180.0) - theta;
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
}

// Fluid-halfspace reflection coefficient of the O3D branch (reflect.hpp:279-299),
// used by 3D and Nx2D rays: Tg the tangential slowness, c and rho the water's.
template<typename real>
__host__ __device__ inline cpx<real> HalfspaceReflO3D(real omega, real Tg, real c, const HSd<real> &hs, real rho)
{
    cpx<real> gk = C(omega * Tg, (real)0);
    cpx<real> gk2 = cmul(gk, gk);
    real w1 = omega / c;
    cpx<real> eps = C((real)0, (real)RC<real>::MINPOS); // J * REAL_MINPOS
    cpx<real> g1sq = (C(w1 * w1, (real)0) - gk2) - eps;
    cpx<real> wP = realdiv(omega, hs.cP);
    cpx<real> g2sq = (cmul(wP, wP) - gk2) - eps;
    cpx<real> g1 = csqrt_(-g1sq), g2 = csqrt_(-g2sq);
    cpx<real> a = rmul(hs.rho, g1), b = rmul(rho, g2);
    return cdiv(a - b, a + b);
}

template<typename real>
__host__ __device__ inline void Reflect3(const Ray3<real> &oldp, Ray3<real> &newp, const HSd<real> &hs, bool isTop,
                                         v3<real> nB, real zxx, real zxy, real zyy, const Prob3D<real> &P,
                                         Trace3<real> &S, Err &e, const real *tabs = nullptr)
{
    real Th = dot3(oldp.t, nB);
    v3<real> tB = oldp.t - sc3(Th, nB);
    tB = sc3(RL(1.0) / len3(tB), tB);
    real Tg = dot3(oldp.t, tB);
    newp.NumTopBnc = oldp.NumTopBnc + (isTop ? 1 : 0);
    newp.NumBotBnc = oldp.NumBotBnc + (isTop ? 0 : 1);
    newp.x = oldp.x;
    newp.t = oldp.t - sc3(FL(2.0) * Th, nB);
    SSPOut3<real> o;
    EvaluateSSP3(oldp.x, newp.t, o, P, S, e);
    real c = o.c.re;
    newp.c = c;
    newp.tau = oldp.tau;
    v3<real> rayt, rayn1, rayn2, rayt_t, rayn1_t, rayn2_t;
    TangentNormals3(oldp, c, nB, rayt, rayn1, rayn2, RL(-1.0));
    TangentNormals3(newp, c, nB, rayt_t, rayn1_t, rayn2_t, RL(-1.0));
    v2<real> t_rot = V(rayt.x, rayt.y);
    t_rot = sc(RL(1.0) / std::sqrt(dot(t_rot, t_rot)), t_rot);
    v2<real> n_rot = V(rayn2.x, rayn2.y);
    n_rot = sc(RL(1.0) / std::sqrt(dot(n_rot, n_rot)), n_rot);
    m2<real> R = M2(t_rot.x, t_rot.y, n_rot.x, n_rot.y);
    // kappaMat from the boundary's z curvatures (zero off curvilinear grids)
    const m2<real> K = M2(zxx / FL(2.0), zxy / FL(2.0), zxy / FL(2.0), zyy / FL(2.0));
    m2<real> DMat = mmul(mtr(R), mmul(mscl((real)FL(2.0), K), R));
    real cn1 = dot3(o.gradc, -rayn1_t - rayn1);
    real cn2 = dot3(o.gradc, -rayn2_t - rayn2);
    real cs = -dot3(o.gradc, rayt_t - rayt);
    v3<real> e1, e2;
    RayNormal3(oldp.t, oldp.phi, false, oldp.c, e1, e2);
    newp.p = oldp.p;
    newp.q = oldp.q;
    newp.phi = oldp.phi;
    CurvCorr3D(newp, DMat, Tg, Th, cn1, cn2, cs, rayn1, rayn2, e1, e2, true);

    if(hs.bc == 'R') {
        newp.Amp = oldp.Amp;
        newp.Phase = oldp.Phase;
    } else if(hs.bc == 'V') {
        newp.Amp = oldp.Amp;
        newp.Phase = oldp.Phase + RC<real>::PI;
    } else if(hs.bc == 'F') {
        real r, phi;
        ReflTableO3D(refl_tab(P, isTop, tabs), isTop ? P.ntrc : P.nbrc, Th, Tg, r, phi, e);
        newp.Amp = oldp.Amp * r;
        newp.Phase = oldp.Phase + phi;
    } else if(hs.bc == 'A' || hs.bc == 'G') {
        cpx<real> Refl = HalfspaceReflO3D(FL(2.0) * RC<real>::PI * P.freq0, Tg, c, hs, o.rho);
        real aR = cabs_(Refl);
        if(aR < FL(1.0e-5)) {
            newp.Amp = FL(0.0);
            newp.Phase = oldp.Phase;
        } else {
            newp.Amp = aR * oldp.Amp;
            newp.Phase = oldp.Phase + std::atan2(Refl.im, Refl.re);
        }
    } else {
        e.error(4);
        newp.Amp = NAN;
        newp.Phase = NAN;
    }
}

// ---------------------------------------------------------------------------
// Ray init / update / terminate (src/trace.hpp, 3D)
// ---------------------------------------------------------------------------
template<typename real> struct Infl3 {
    real Ratio1, rcp_q0, rcp_qhat0, omega, BeamWindow;
    real phase, qOld;
    v3<real> xs;
    v3<real> rn1, rn2;     // ray-centred beams: the previous point's ray normals (c-scaled)
    v2<real> rbr;          // Cartesian: receiver ranges Rr[ir], Rr[ir + 1] (the walk's bracket)
    int32_t ir;
};

template<typename real>
__host__ __device__ inline void JobIndices3(int job, const Prob3D<real> &P, int &ia, int &ib, int &isy, int &isx,
                                            int &isz)
{ // common.hpp:365-391
    if(P.iSingleA >= 1) ia = P.iSingleA - 1;
    else {
        ia = job % P.Nalpha;
        job /= P.Nalpha;
    }
    if(P.iSingleB >= 1) ib = P.iSingleB - 1;
    else {
        ib = job % P.Nbeta;
        job /= P.Nbeta;
    }
    isy = job % P.NSy;
    job /= P.NSy;
    isx = job % P.NSx;
    job /= P.NSx;
    isz = job;
}

template<typename real>
__host__ __device__ inline bool RayInit3(int ia, int ib, int isx, int isy, int isz, Ray3<real> &p0, Trace3<real> &S,
                                         Infl3<real> &I, const Prob3D<real> &P, Err &e)
{ // trace.hpp:51-67, 76-266 (3D) + Init_Influence influence.hpp:525-644 (geometric Cartesian)
    const real *G = P.base;
    real alpha = G[P.o_alpha + ia];
    real SrcDecl = (RL(180.0) / RC<real>::PI) * alpha;
    S.sx = S.sy = S.sz = S.sr = 0;
    S.xs = V3(G[P.o_Sx + isx], G[P.o_Sy + isy], (real)P.Sz[isz]);
    SSPOut3<real> o;
    EvaluateSSP3(S.xs, V3((real)FL(0.0), (real)FL(0.0), (real)FL(1.0)), o, P, S, e);
    real c = o.c.re;
    const real *sbp = G + P.o_sbp;
    int ibp = BinarySearchLEQ(sbp, P.nsbp, 2, 0, SrcDecl);
    ibp = imin(ibp, P.nsbp - 2);
    real s = (SrcDecl - sbp[2 * ibp]) / (sbp[2 * (ibp + 1)] - sbp[2 * ibp]);
    float Amp0 = (float)((FL(1.0) - s) * sbp[2 * ibp + 1] + s * sbp[2 * (ibp + 1) + 1]);
    const real ca = G[P.o_acos + ia], sa = G[P.o_asin + ia];
    if(P.run0 == 'S') {
        float omega = (float)(FL(2.0) * RC<real>::PI * P.freq0);
        Amp0 = (float)(Amp0 * (sqrtf(FL(2.0)) * std::fabs(std::sin(omega / c * S.xs.z * sa))));
    }
    v3<real> tinit = V3(ca * G[P.o_bcos + ib], ca * G[P.o_bsin + ib], sa);
    p0.x = S.xs;
    p0.c = c;
    p0.t = dv3(tinit, c);
    p0.tau = C((real)0, (real)0);
    p0.Amp = Amp0;
    p0.Phase = FL(0.0);
    p0.NumTopBnc = p0.NumBotBnc = 0;
    p0.p = M2((real)FL(1.0), (real)FL(0.0), (real)FL(0.0), (real)FL(1.0));
    p0.q = M2((real)FL(0.0), (real)FL(0.0), (real)FL(0.0), (real)FL(0.0));
    p0.phi = FL(0.0);
    S.top = BState3<real>{};
    S.bot = BState3<real>{};
    GetBdrySeg3(S.xs, p0.t, S.top, G + P.o_ati, P.ati_nx, P.ati_ny, true, true, e, prov_of(P, true));
    GetBdrySeg3(S.xs, p0.t, S.bot, G + P.o_bty, P.bty_nx, P.bty_ny, false, true, e, prov_of(P, false));
    real dt = -dot3(S.top.n, S.xs - S.top.x), db = -dot3(S.bot.n, S.xs - S.bot.x);
    if(dt <= FL(0.0) || db <= FL(0.0)) {
        e.warning(4);
        return false;
    }
    S.iSmall = 0;
    // Init_Influence
    I.omega = FL(2.0) * RC<real>::PI * P.freq0;
    I.BeamWindow = P.isGaussian ? RL(4.0) : RL(1.0);
    I.Ratio1 = std::sqrt(std::fabs(ca));
    I.Ratio1 *= std::sqrt(P.Dalpha * P.Dbeta) / p0.c;
    if(P.isGaussian) I.Ratio1 /= FL(2.0) * RC<real>::PI;
    I.rcp_q0 = P.Dalpha / p0.c;
    I.rcp_qhat0 = std::fabs(ca) * P.Dbeta / p0.c;
    I.phase = FL(0.0);
    I.qOld = mdet(p0.q);
    I.xs = p0.x;
    I.ir = BinarySearchGT(G + P.o_Rr, P.NRr, (real)RL(0.0));
    I.rbr = range_bracket(G + P.o_Rr, P.NRr, I.ir);
    I.rn1 = I.rn2 = V3((real)RL(0.0), (real)RL(0.0), (real)RL(0.0));
    if(P.isRayCen) RayNormal3(p0.t, p0.phi, false, p0.c, I.rn1, I.rn2); // RayNormalRayCen
    return true;
}

template<typename real>
__host__ __device__ inline bool RayUpdate3(const Ray3<real> &p0, Ray3<real> &p1, Ray3<real> &p2, Trace3<real> &S,
                                           const Prob3D<real> &P, Err &e, const real *tabs = nullptr)
{ // trace.hpp:274-381 (3D, rectilinear boundaries)
    bool topRefl = false, botRefl = false;
    Step3(p0, p1, S, P, e, topRefl, botRefl);
    GetBdrySeg3(p1.x, p1.t, S.top, P.base + P.o_ati, P.ati_nx, P.ati_ny, true, false, e, prov_of(P, true));
    GetBdrySeg3(p1.x, p1.t, S.bot, P.base + P.o_bty, P.bty_nx, P.bty_ny, false, false, e, prov_of(P, false));
    if(topRefl || botRefl) {
        ReflB3<real> r;
        refl_bdry3(P, topRefl, topRefl ? S.top : S.bot, p1.x, r);
        Reflect3(p1, p2, r.hs, topRefl, r.n, r.zxx, r.zxy, r.zyy, P, S, e, tabs);
        return true;
    }
    return false;
}

// Returns true when the ray ends; nsteps is set as the reference sets Nsteps
// (is + 1 when the ray leaves / dies, is at the storage limit).
template<typename real>
__host__ __device__ inline bool RayTerminate3(const Ray3<real> &pt, int is, const Trace3<real> &S,
                                              const Prob3D<real> &P, Err &e, int maxPoints, int &nsteps)
{ // trace.hpp:391-477 (3D)
    v3<real> x = pt.x, t = pt.t, xs = S.xs;
    bool leftbox = std::fabs(x.x - xs.x) >= P.box_x || std::fabs(x.y - xs.y) >= P.box_y
        || std::fabs(x.z - RL(0.0)) >= P.box_z;
    const real *bb = P.base + P.o_bty, *ab = P.base + P.o_ati;
    real minx = rmax(bb[0], ab[0]);
    real miny = rmax(bb[1], ab[1]);
    real maxx = rmin(bb[(size_t)(P.bty_nx - 1) * P.bty_ny * NODE3], ab[(size_t)(P.ati_nx - 1) * P.ati_ny * NODE3]);
    real maxy = rmin(bb[(size_t)(P.bty_ny - 1) * NODE3 + 1], ab[(size_t)(P.ati_ny - 1) * NODE3 + 1]);
    bool e0 = x.x < minx || x.y < miny || (x.x == minx && t.x < RL(0.0)) || (x.y == miny && t.y < RL(0.0));
    bool eN = x.x > maxx || x.y > maxy || (x.x == maxx && t.x > RL(0.0)) || (x.y == maxy && t.y > RL(0.0));
    bool small = S.iSmall > 50;
    bool lost = pt.Amp < FL(0.005);
    if(leftbox || lost || e0 || eN || small) {
        nsteps = is + 1;
        return true;
    }
    if(is >= maxPoints - 3) {
        e.warning(1);
        nsteps = is;
        return true;
    }
    return false;
}

template<typename real>
__host__ __device__ inline bool RayTerminate3(const Ray3<real> &pt, int is, const Trace3<real> &S,
                                              const Prob3D<real> &P, Err &e)
{
    int n;
    return RayTerminate3(pt, is, S, P, e, 100000, n);
}

template<typename real> __host__ __device__ inline void store_pt3(bhc_ray_pt3d &d, const Ray3<real> &r)
{
    d.p[0] = (double)r.p.a00;
    d.p[1] = (double)r.p.a01;
    d.p[2] = (double)r.p.a10;
    d.p[3] = (double)r.p.a11;
    d.q[0] = (double)r.q.a00;
    d.q[1] = (double)r.q.a01;
    d.q[2] = (double)r.q.a10;
    d.q[3] = (double)r.q.a11;
    d.phi = (double)r.phi;
    d.dummy = 0.0;
    d.NumTopBnc = r.NumTopBnc;
    d.NumBotBnc = r.NumBotBnc;
    d.x[0] = (double)r.x.x;
    d.x[1] = (double)r.x.y;
    d.x[2] = (double)r.x.z;
    d.t[0] = (double)r.t.x;
    d.t[1] = (double)r.t.y;
    d.t[2] = (double)r.t.z;
    d.c = (double)r.c;
    d.Amp = (double)r.Amp;
    d.Phase = (double)r.Phase;
    d.tau[0] = (double)r.tau.re;
    d.tau[1] = (double)r.tau.im;
}

// ---------------------------------------------------------------------------
// Influence: geometric Cartesian beams (src/influence.hpp:897-1024, 1201-1406, 3D)
// ---------------------------------------------------------------------------
// Step segment descriptor: the receiver work of one step, scattered later.
template<typename real> struct Seg3D {
    v3<real> xa, ta, e1, e2;    // segment start, unit tangent, ray normals
    real rlen, lo, hi;          // length; receiver ranges in [lo, hi)
    real zlo, zhi, Ldiag;       // depth extent of the segment, beam half-diagonal
    m2<real> q0, dq;
    cpx<real> tau0, dtau;
    real A0;                    // Ratio1 * c1 * Amp1
    real phase0, qOld;          // Phase0 + accumulated caustic phase; det q at the segment start
    real rcp_q0, rcp_qhat0;
    real xs_x, xs_y;
    real F0, Fd;                // Frobenius norms of the scaled q at the segment start and of its change
    int32_t ir0, ir1, dir;      // receiver ranges visited by the walk
    int32_t nitems;             // > 0 when the walk brackets at least one range
    int32_t isrc;               // source (isz * NSx + isx) * NSy + isy
    int32_t job, is, ntop, nbot; // ray (job) number, step index, end point's bounce counts (arrivals, eigen hits)
};

template<typename real>
__host__ __device__ inline void prelude3(Seg3D<real> &d, const Ray3<real> &p0, const Ray3<real> &p1, Infl3<real> &I,
                                         int is, const Prob3D<real> &P, int isz, int isx, int isy)
{
    d.nitems = 0;
    const real *Rr = P.base + P.o_Rr;
    real dax = p0.x.x - I.xs.x, day = p0.x.y - I.xs.y, dbx = p1.x.x - I.xs.x, dby = p1.x.y - I.xs.y;
    real rA = std::sqrt(dax * dax + day * day);
    real rB = std::sqrt(dbx * dbx + dby * dby);
    if(std::fabs(rA - rB) <= RL(1.0e3) * spacing(rA)) return; // IsDuplicatePoint
    if(is == 0) {
        I.ir = (rB > rA) ? 0 : P.NRr - 1;
        I.rbr = range_bracket(Rr, P.NRr, I.ir);
    }
    v3<real> rayt = p1.x - p0.x;
    real rlen = len3(rayt);
    if(rlen <= RL(1.0e3) * spacing(p1.x.x)) return;
    real phaseq = mdet(p0.q);
    if((phaseq <= RL(0.0) && I.qOld > RL(0.0)) || (phaseq >= RL(0.0) && I.qOld < RL(0.0)))
        I.phase += RC<real>::PI / FL(2.0);
    I.qOld = phaseq;
    // receiver walk: the visited ranges are contiguous; the bracket
    // (Rr[ir], Rr[ir + 1]) travels with the ray, so a step that crosses no
    // receiver range reads none (the comparisons are the walk's)
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
    d.lo = rmin(rA, rB);
    d.hi = rmax(rA, rB);
    // a step that crosses no receiver range has the one range Rr[first]:
    // outside [lo, hi) the scatter's row test accepts none of its receivers
    if(ir == first && !(rfirst >= d.lo && rfirst < d.hi)) return;
    // the beam's frame and reach: only for steps handed to the scatter
    rayt = dv3(rayt, rlen);
    v3<real> e1, e2;
    RayNormal3(rayt, p1.phi, true, (real)RL(1.0), e1, e2);
    real a0 = p0.q.a00 * I.rcp_q0, b0 = p0.q.a10 * I.rcp_q0, a1 = p1.q.a00 * I.rcp_q0, b1 = p1.q.a10 * I.rcp_q0;
    real l1 = rmax(std::sqrt(a0 * a0 + b0 * b0), std::sqrt(a1 * a1 + b1 * b1));
    a0 = p0.q.a01 * I.rcp_qhat0;
    b0 = p0.q.a11 * I.rcp_qhat0;
    a1 = p1.q.a01 * I.rcp_qhat0;
    b1 = p1.q.a11 * I.rcp_qhat0;
    real l2 = rmax(std::sqrt(a0 * a0 + b0 * b0), std::sqrt(a1 * a1 + b1 * b1));
    d.Ldiag = std::sqrt(l1 * l1 + l2 * l2);
    d.ir0 = first;
    d.ir1 = ir;
    d.dir = (ir >= first) ? 1 : -1;
    d.xa = p0.x;
    d.ta = rayt;
    d.e1 = e1;
    d.e2 = e2;
    d.rlen = rlen;
    d.zlo = rmin(p0.x.z, p1.x.z);
    d.zhi = rmax(p0.x.z, p1.x.z);
    d.q0 = p0.q;
    d.dq = msub(p1.q, p0.q);
    d.tau0 = p0.tau;
    d.dtau = p1.tau - p0.tau;
    d.A0 = I.Ratio1 * p1.c * p1.Amp;
    d.phase0 = p0.Phase + I.phase;
    d.qOld = I.qOld;
    d.rcp_q0 = I.rcp_q0;
    d.rcp_qhat0 = I.rcp_qhat0;
    {
        real u00 = p0.q.a00 * I.rcp_q0, u10 = p0.q.a10 * I.rcp_q0;
        real u01 = p0.q.a01 * I.rcp_qhat0, u11 = p0.q.a11 * I.rcp_qhat0;
        real v00 = d.dq.a00 * I.rcp_q0, v10 = d.dq.a10 * I.rcp_q0;
        real v01 = d.dq.a01 * I.rcp_qhat0, v11 = d.dq.a11 * I.rcp_qhat0;
        d.F0 = std::sqrt(u00 * u00 + u10 * u10 + u01 * u01 + u11 * u11);
        d.Fd = std::sqrt(v00 * v00 + v10 * v10 + v01 * v01 + v11 * v11);
    }
    d.xs_x = I.xs.x;
    d.xs_y = I.xs.y;
    d.isrc = (isz * P.NSx + isx) * P.NSy + isy;
    d.is = is;
    d.ntop = p1.NumTopBnc;
    d.nbot = p1.NumBotBnc;
    d.nitems = ((ir >= first ? ir - first : first - ir) + 1) * P.Ntheta;
}

// One (range, bearing) row of a segment: the row tests of the reference and
// the depth window [zmin, zmax]. Returns false when the row has no receivers.
// Rr and trcv: the receiver ranges and bearing (cos, sin) table (in the blob or a staged copy).
template<typename real>
__host__ __device__ inline bool row3(const Seg3D<real> &d, int ir, int itheta, const Prob3D<real> &P,
                                     const real *Rr, const real *trcv, real &xr, real &yr, real &zmin, real &zmax)
{
    real rr = Rr[ir];
    if(!(rr >= d.lo && rr < d.hi)) return false;
    const real BW = P.isGaussian ? RL(4.0) : RL(1.0);
    real tx = trcv[2 * itheta], ty = trcv[2 * itheta + 1];
    xr = d.xs_x + rr * tx;
    yr = d.xs_y + rr * ty;
    real mprime = std::fabs((xr - d.xa.x) * (-d.ta.y) + (yr - d.xa.y) * d.ta.x);
    if(!P.isGaussian && mprime > BW * d.Ldiag) return false;
    real sdot = (xr - d.xs_x) * (d.xa.x - d.xs_x) + (yr - d.xs_y) * (d.xa.y - d.xs_y);
    if(sdot < RL(0.0)) return false;
    real nrz = d.ta.x * tx - d.ta.y * (-ty);
    if(std::fabs(nrz) < RL(1e-9)) return false;
    real Lz = BW * d.Ldiag / std::fabs(nrz);
    zmin = d.zlo - Lz;
    zmax = d.zhi + Lz;
    return true;
}

// Depth band of a row that can pass the beam-window test. With u = (n1, n2)
// the receiver's offset from the segment's line (|u| = d, its perpendicular
// distance) and Q the row-scaled q at the receiver's projection s, the test
// needs |Q^-1 u| <= K (K = 4 for Gaussian beams, whose test bounds
// n1' + n2' >= |Q^-1 u|; sqrt 2 for hat beams, max(n1', n2') >= |Q^-1 u| /
// sqrt 2), hence d <= K |Q|_F <= K (F0 + |s| Fd). Along the receiver column,
// d^2 = a^2 (z - zc)^2 + dmin^2 (a = horizontal length of the unit tangent,
// zc the depth of closest approach) and |s| <= |s(zc)| + |z - zc| |tz| / L:
// the receivers are limited to |z - zc| <= K (F0 + Fd |s(zc)|) / (a - K Fd |tz| / L),
// and the row is empty when dmin exceeds the bound at the band's edge. The
// band is widened for rounding; receivers inside it go through every test of
// the reference. Returns false for an empty row; (wlo, whi) is the band, or
// the full window [zmin, zmax] when the bound does not apply.
template<typename real>
__host__ __device__ inline bool band3_circle(const Seg3D<real> &d, real xr, real yr, real zmin, real zmax,
                                             const Prob3D<real> &P, real &wlo, real &whi)
{
    wlo = zmin;
    whi = zmax;
    const double K = P.isGaussian ? 4.0 : 1.4142135623730951;
    const double tx = (double)d.ta.x, ty = (double)d.ta.y, tz = (double)d.ta.z;
    const double a2 = tx * tx + ty * ty, a = std::sqrt(a2);
    const double L = (double)d.rlen, F0 = (double)d.F0, Fd = (double)d.Fd;
    const double coef = a - K * Fd * std::fabs(tz) / L;
    if(!(a > 1e-6) || !(coef > 1e-3 * a)) return true;
    const double hx = (double)xr - (double)d.xa.x, hy = (double)yr - (double)d.xa.y;
    const double hta = hx * tx + hy * ty;
    const double zeta = hta * tz / a2;                 // zc - za
    const double m = hx * ty - hy * tx;                // horizontal distance to the ray's vertical plane x a
    const double dmin = std::fabs(m) / a;
    const double sc = (hta + zeta * tz) / L;           // s at the closest approach
    const double R0 = K * (F0 + Fd * std::fabs(sc));
    const double mag = std::fabs(hx) + std::fabs(hy) + std::fabs((double)d.xa.z) + std::fabs(zeta) + 1.0;
    const double slack = 1e-4 * mag + 1e-2;
    const double dz = R0 / coef * (1.0 + 1e-4) + slack;
    if(!(dz == dz) || !(dz < 1e30) || !(dmin == dmin)) return true;
    const double dlim = K * (F0 + Fd * (std::fabs(sc) + dz * std::fabs(tz) / L)) * (1.0 + 1e-4) + slack;
    if(dmin > dlim) return false;
    const double zc = (double)d.xa.z + zeta;
    wlo = rmax(zmin, (real)(zc - dz));
    whi = rmin(zmax, (real)(zc + dz));
    return whi >= wlo;
}

// Narrows a row's depth band by each beam-plane coordinate separately. With
// w = (n1, n2) the receiver's offsets along e1, e2 and A the row-scaled q at
// its s (rows (a00, a01), (a10, a11) as the beam-window test uses them), the
// test's beam coordinates are m = A^-1 w, so w1 = a00 m1 + a01 m2 and
// w2 = a10 m1 + a11 m2: Gaussian beams (|m1| + |m2| <= 4) need |w1| <= 4
// max(|a00|, |a01|), hat beams (max |mi| <= 1) |w1| <= |a00| + |a01|, and
// likewise for w2. |a_ij| is bounded over the band by |A0_ij| + smax |dA_ij|.
// Along the column w_i is affine in z, so each bound is an interval of z.
// For a flat beam (widths 50:1 on the shelf-canyon field) the coordinate
// along the narrow width limits the band where the circle of radius
// K |A|_F does not. The bounds are widened for the rounding of the test's own
// arithmetic (which grows with A's condition number); ill-conditioned or
// degenerate segments keep the circle band.
// a z <= b applied to the interval [L, H], with ra = 1 / a (a == 0 and b < 0
// empties it; a NaN bound leaves it unchanged).
__host__ __device__ inline void clip_le(double a, double ra, double b, double &L, double &H)
{
    if(a > 0.0) H = rmin(H, b * ra);
    else if(a < 0.0) L = rmax(L, b * ra);
    else if(b < 0.0) H = -1e300;
}

// Bounds |c + g z| <= alpha + beta |s(z)|, s(z) = p + sig z, of one beam-plane
// direction, with the reciprocals of its two slopes g -/+ beta sig.
struct AbsAffine {
    double c, g, alpha, beta, rA, rB;
};

// [L, H] cut to the z where bound f holds, all z of the interval on the side
// sgn of s = 0 (sgn = 1: s >= 0, -1: s <= 0).
__host__ __device__ inline void clip_side(const AbsAffine &f, double p, double sig, int sgn, double &L, double &H)
{
    const double base = f.alpha + sgn * f.beta * p;
    // s >= 0: (g - beta sig) z <= base - c and (-g - beta sig) z <= base + c
    // s <= 0: -(-g - beta sig) z <= base - c and -(g - beta sig) z <= base + c
    const double a1 = f.g - f.beta * sig, a2 = -f.g - f.beta * sig;
    if(sgn > 0) {
        clip_le(a1, f.rA, base - f.c, L, H);
        clip_le(a2, f.rB, base + f.c, L, H);
    } else {
        clip_le(-a2, -f.rB, base - f.c, L, H);
        clip_le(-a1, -f.rA, base + f.c, L, H);
    }
}

// The hull of [zA, zB]'s solutions of bounds f0, f1 and (use2) f2, the
// sides s >= 0 and s <= 0 solved separately (an empty result sets lo > hi).
__host__ __device__ inline void abs_affine_hull(const AbsAffine &f0, const AbsAffine &f1, const AbsAffine &f2,
                                                bool use2, double p, double sig, double zA, double zB, double &lo,
                                                double &hi)
{
    double L1 = zA, H1 = zB, L2 = zA, H2 = zB;
    // the side of s = 0 in z (s = p + sig z)
    if(sig > 0.0) {
        const double z0 = -p / sig;
        L1 = rmax(L1, z0);
        H2 = rmin(H2, z0);
    } else if(sig < 0.0) {
        const double z0 = -p / sig;
        H1 = rmin(H1, z0);
        L2 = rmax(L2, z0);
    } else if(p >= 0.0) {
        L2 = 1.0;
        H2 = -1.0;
    } else {
        L1 = 1.0;
        H1 = -1.0;
    }
    clip_side(f0, p, sig, 1, L1, H1);
    clip_side(f1, p, sig, 1, L1, H1);
    if(use2) clip_side(f2, p, sig, 1, L1, H1);
    clip_side(f0, p, sig, -1, L2, H2);
    clip_side(f1, p, sig, -1, L2, H2);
    if(use2) clip_side(f2, p, sig, -1, L2, H2);
    const bool e1 = !(L1 <= H1), e2 = !(L2 <= H2);
    if(e1 && e2) {
        lo = 1.0;
        hi = -1.0;
    } else if(e1) {
        lo = L2;
        hi = H2;
    } else if(e2) {
        lo = L1;
        hi = H1;
    } else {
        lo = rmin(L1, L2);
        hi = rmax(H1, H2);
    }
}

// Roots of c2 s^2 + c1 s + c0 = 0 in r0, r1 (NaN where there is none).
__host__ __device__ inline void quad_roots(double c2, double c1, double c0, double &r0, double &r1)
{
    r0 = r1 = __builtin_nan("");
    if(std::fabs(c2) <= 1e-300 * (std::fabs(c1) + std::fabs(c0)) || c2 == 0.0) {
        if(c1 != 0.0) r0 = -c0 / c1;
        return;
    }
    const double disc = c1 * c1 - 4.0 * c2 * c0;
    if(!(disc >= 0.0)) return;
    const double sq = std::sqrt(disc);
    const double qq = -0.5 * (c1 + (c1 >= 0.0 ? sq : -sq));
    r0 = qq / c2;
    if(qq != 0.0) r1 = c0 / qq;
}

__host__ __device__ inline void sort2(double &a, double &b)
{
    const double lo = rmin(a, b), hi = rmax(a, b);
    a = lo;
    b = hi;
}

// Narrows a row's depth band by the beam-plane coordinates separately. With
// w = (n1, n2) the receiver's offsets along e1, e2 and A the row-scaled q at
// its s (rows (a00, a01), (a10, a11) as the beam-window test uses them), the
// test's beam coordinates are m = A^-1 w, so w1 = a00 m1 + a01 m2 and
// w2 = a10 m1 + a11 m2: Gaussian beams (|m1| + |m2| <= 4) need |w1| <= 4
// max(|a00|, |a01|), hat beams (max |mi| <= 1) |w1| <= |a00| + |a01|, and
// likewise for w2 and for the column's own direction in the (e1, e2) plane.
// |a_ij(s)| <= |A0_ij| + |s| |dA_ij|, and along the column w_i and s are
// affine in z, so each bound is a condition |c + g z| <= alpha + beta |s(z)|
// solved exactly for z. For a flat beam (widths 50:1 on the shelf-canyon
// field) the coordinate along the narrow width limits the band where the
// circle of radius K |A|_F does not.
// The test's own arithmetic rounds with an error that grows with A's
// condition number (|a00 a11| + |a01 a10|) / |det A|, so the column is cut
// where |det A(s)| = tau N (N bounding the numerator over the band): pieces
// with |det A| >= tau N get the bounds widened by 256 eps / tau, the others
// (around the zeros of det A(s)) are kept whole. The result is the hull of
// the pieces' bands.
template<typename real>
__host__ __device__ inline bool band3_axes(const Seg3D<real> &d, real xr, real yr, const Prob3D<real> &P, real &wlo,
                                           real &whi)
{
    const double zl = (double)wlo, zh = (double)whi;
    if(!(zl == zl) || !(zh == zh) || !(std::fabs(zl) < 1e30) || !(std::fabs(zh) < 1e30)) return true;
    const double L = (double)d.rlen, za = (double)d.xa.z;
    const double hx = (double)xr - (double)d.xa.x, hy = (double)yr - (double)d.xa.y;
    const double sig = (double)d.ta.z / L;
    const double p = (hx * (double)d.ta.x + hy * (double)d.ta.y) / L - za * sig;   // s(z) = p + sig z
    const double smax = rmax(std::fabs(p + sig * zl), std::fabs(p + sig * zh));
    const double rq = (double)d.rcp_q0, rh = (double)d.rcp_qhat0;
    const double sa00 = (double)d.q0.a00 * rq, sa01 = (double)d.q0.a01 * rh;
    const double sa10 = (double)d.q0.a10 * rq, sa11 = (double)d.q0.a11 * rh;
    const double sb00 = (double)d.dq.a00 * rq, sb01 = (double)d.dq.a01 * rh;
    const double sb10 = (double)d.dq.a10 * rq, sb11 = (double)d.dq.a11 * rh;
    const double a00 = std::fabs(sa00), a01 = std::fabs(sa01), a10 = std::fabs(sa10), a11 = std::fabs(sa11);
    const double b00 = std::fabs(sb00), b01 = std::fabs(sb01), b10 = std::fabs(sb10), b11 = std::fabs(sb11);
    // det A(s) = det0 + lin s + ddet s^2
    const double det0 = sa00 * sa11 - sa10 * sa01;
    const double lin = sa00 * sb11 + sb00 * sa11 - sa10 * sb01 - sb10 * sa01;
    const double ddet = sb00 * sb11 - sb10 * sb01;
    const double Nmax = (a00 + smax * b00) * (a11 + smax * b11) + (a01 + smax * b01) * (a10 + smax * b10);
    if(!(Nmax > 0.0) || !(Nmax < 1e300)) return true;
    const bool gauss = P.isGaussian;
    const double u = (double)RC<real>::EPS;
    const double tau = sizeof(real) == 8 ? 1e-9 : 1e-3;
    const double T = tau * Nmax;
    const double rel = rmax(1e-4, 256.0 * u / tau);
    const double K = gauss ? 4.0 : 1.4142135623730951;
    const double mag = std::fabs(hx) + std::fabs(hy) + rmax(std::fabs(zl - za), std::fabs(zh - za));
    // alpha + beta |s| bounds the accepted |w| along a direction whose row
    // weights are (x, y): x |a_0j| + y |a_1j| for columns j = 0, 1
    const double slack0 = rel * K * (double)d.F0 + 1e-4 * mag + 1e-3, slack1 = rel * K * (double)d.Fd;
    auto coef = [&](double x, double y, double &alpha, double &beta) {
        const double c0 = x * a00 + y * a10, c1 = x * a01 + y * a11;
        const double e0 = x * b00 + y * b10, e1 = x * b01 + y * b11;
        alpha = (gauss ? 4.0 * rmax(c0, c1) : c0 + c1) * (1.0 + rel) + slack0;
        beta = (gauss ? 4.0 * rmax(e0, e1) : e0 + e1) * (1.0 + rel) + slack1;
    };
    const double g1 = (double)d.e1.z, g2 = (double)d.e2.z;
    const double c1 = hx * (double)d.e1.x + hy * (double)d.e1.y - za * g1;   // w1 = c1 + z g1
    const double c2 = hx * (double)d.e2.x + hy * (double)d.e2.y - za * g2;   // w2 = c2 + z g2
    const double gn = std::sqrt(g1 * g1 + g2 * g2);
    AbsAffine A0, A1, A2;
    A0.c = c1;
    A0.g = g1;
    coef(1.0, 0.0, A0.alpha, A0.beta);
    A1.c = c2;
    A1.g = g2;
    coef(0.0, 1.0, A1.alpha, A1.beta);
    // the column's direction n = (g1, g2) / |g|: |n . w| <= n' . (|w1|, |w2|), n' = |n|
    const bool use2 = gn > 1e-9;
    A2.c = use2 ? (g1 * c1 + g2 * c2) / gn : 0.0;
    A2.g = gn;
    coef(use2 ? std::fabs(g1) / gn : 0.0, use2 ? std::fabs(g2) / gn : 0.0, A2.alpha, A2.beta);
    if(!(A0.alpha < 1e30) || !(A0.beta < 1e30) || !(A1.alpha < 1e30) || !(A1.beta < 1e30) || !(A2.alpha < 1e30) ||
       !(A2.beta < 1e30))
        return true;
    A0.rA = 1.0 / (A0.g - A0.beta * sig);
    A0.rB = 1.0 / (-A0.g - A0.beta * sig);
    A1.rA = 1.0 / (A1.g - A1.beta * sig);
    A1.rB = 1.0 / (-A1.g - A1.beta * sig);
    A2.rA = 1.0 / (A2.g - A2.beta * sig);
    A2.rB = 1.0 / (-A2.g - A2.beta * sig);
    double lo = 1e300, hi = -1e300;
    if(std::fabs(det0) - smax * std::fabs(lin) - smax * smax * std::fabs(ddet) >= T) {
        // |det A(s)| >= T over the whole band: one piece
        abs_affine_hull(A0, A1, A2, use2, p, sig, zl, zh, lo, hi);
    } else {
        // pieces of [zl, zh] between the crossings of |det A(s)| = T (the
        // crossings outside the band, or missing, collapse onto zh)
        double r0, r1, r2, r3;
        quad_roots(ddet, lin, det0 - T, r0, r1);
        quad_roots(ddet, lin, det0 + T, r2, r3);
        auto toz = [&](double r) {
            const double z = (r - p) / sig;
            return (sig != 0.0 && z > zl && z < zh) ? z : zh;
        };
        double z1 = toz(r0), z2 = toz(r1), z3 = toz(r2), z4 = toz(r3);
        sort2(z1, z2);
        sort2(z3, z4);
        sort2(z1, z3);
        sort2(z2, z4);
        sort2(z2, z3);
        const double zb[6] = {zl, z1, z2, z3, z4, zh};
#pragma unroll
        for(int k = 0; k < 5; ++k) {
            const double zA = zb[k], zB = zb[k + 1];
            const double sm = p + sig * (0.5 * (zA + zB));
            const double fm = det0 + lin * sm + ddet * sm * sm;
            double l = zA, h = zB;
            if(std::fabs(fm) >= T) abs_affine_hull(A0, A1, A2, use2, p, sig, zA, zB, l, h);
            if(l <= h) {
                lo = rmin(lo, l);
                hi = rmax(hi, h);
            }
        }
    }
    if(!(lo <= hi)) return false;
    // round outwards to real
    real rlo = (real)lo, rhi = (real)hi;
    if((double)rlo > lo) rlo = (real)(lo - std::fabs(lo) * 4.0 * u - 1e-30);
    if((double)rhi < hi) rhi = (real)(hi + std::fabs(hi) * 4.0 * u + 1e-30);
    wlo = rmax(wlo, rlo);
    whi = rmin(whi, rhi);
    return whi >= wlo;
}

template<typename real>
__host__ __device__ inline bool band3(const Seg3D<real> &d, real xr, real yr, real zmin, real zmax,
                                      const Prob3D<real> &P, real &wlo, real &whi)
{
    if(!band3_circle(d, xr, yr, zmin, zmax, P, wlo, whi)) return false;
    return band3_axes(d, xr, yr, P, wlo, whi);
}

// Bearing prefilter of a segment: can any receiver on bearing (ct, st) with a
// range in [rlo, rhi] pass the beam-window test? An accepted receiver lies
// within d <= K (F0 + Fd |s|) of the segment's line, with |s| <= |X - xa| / L
// <= (rhi + rA + zspan) / L (rA the segment start's horizontal distance from
// the source, zspan the largest depth offset of a receiver); and d is at least
// the horizontal distance of the receiver column from the line's vertical
// plane, |m0 + r (ct ty - st tx)| / a. The bearing is dropped when that
// distance exceeds the bound for every r in [rlo, rhi] (with rounding slack).
template<typename real>
__host__ __device__ inline bool bearing3(const Seg3D<real> &d, real ct, real st, real rlo, real rhi, real zspan,
                                         const Prob3D<real> &P)
{
    const double K = P.isGaussian ? 4.0 : 1.4142135623730951;
    const double tx = (double)d.ta.x, ty = (double)d.ta.y;
    const double a = std::sqrt(tx * tx + ty * ty);
    if(!(a > 1e-6)) return true;
    const double ex = (double)d.xa.x - (double)d.xs_x, ey = (double)d.xa.y - (double)d.xs_y;
    const double rA = std::sqrt(ex * ex + ey * ey);
    const double smax = ((double)rhi + rA + (double)zspan) / (double)d.rlen;
    const double mag = (double)rhi + rA + (double)zspan + std::fabs((double)d.xa.x) + std::fabs((double)d.xa.y) + 1.0;
    const double dmax = K * ((double)d.F0 + (double)d.Fd * smax) * (1.0 + 1e-4) + 1e-4 * mag + 1e-2;
    const double m0 = -ex * ty + ey * tx;              // (xs - xa) x t
    const double sg = (double)ct * ty - (double)st * tx;
    const double f0 = m0 + (double)rlo * sg, f1 = m0 + (double)rhi * sg;
    const double fmin = (f0 * f1 <= 0.0) ? 0.0 : std::fmin(std::fabs(f0), std::fabs(f1));
    if(!(fmin == fmin) || !(dmax == dmax)) return true;
    return fmin / a <= dmax;
}

template<typename real, typename D, typename Sink, bool FAST, int RT>
__host__ __device__ inline void core3(const D &d, real s, real n1, real n2, int itheta, int ir, int iz,
                                      const v3<real> &rayt, const Prob3D<real> &P, Sink &sink);

// RT selects ApplyContribution's branch (influence.hpp:432-473): TL field
// term, AddArr candidate (arrivals) or RecordEigenHit (eigen hits).
template<typename real, typename Sink, bool FAST = false, int RT = RT_TL>
__host__ __device__ inline void item3(const Seg3D<real> &d, int ir, int itheta, int iz, real xr, real yr, real zmin,
                                      real zmax, const Prob3D<real> &P, const float *Rz, Sink &sink)
{
    real zr = (real)Rz[iz];
    if(zr < zmin || zr > zmax) return;
    v3<real> xrr = V3(xr, yr, zr) - d.xa;
    real s = dot3(xrr, d.ta) / d.rlen;
    real n1 = std::fabs(dot3(xrr, d.e1));
    real n2 = std::fabs(dot3(xrr, d.e2));
    core3<real, Seg3D<real>, Sink, FAST, RT>(d, s, n1, n2, itheta, ir, iz, d.ta, P, sink);
}

template<typename real, typename D>
__host__ __device__ inline bool window3(const D &d, real s, real n1, real n2, const Prob3D<real> &P, m2<real> &qI,
                                        real &qF, real &n1p, real &n2p);

// Whether item3 contributes for this receiver: its depth test and the beam
// tests, with item3's arithmetic.
template<typename real>
__host__ __device__ inline bool item3_accepts(const Seg3D<real> &d, real xr, real yr, int iz, real zmin, real zmax,
                                              const Prob3D<real> &P, const float *Rz)
{
    real zr = (real)Rz[iz];
    if(zr < zmin || zr > zmax) return false;
    v3<real> xrr = V3(xr, yr, zr) - d.xa;
    real s = dot3(xrr, d.ta) / d.rlen;
    real n1 = std::fabs(dot3(xrr, d.e1));
    real n2 = std::fabs(dot3(xrr, d.e2));
    m2<real> qI;
    real qF, n1p, n2p;
    return window3(d, s, n1, n2, P, qI, qF, n1p, n2p);
}

// The tests of InfluenceGeoCore (influence.hpp:897-1024, 3D) for a receiver
// at beam coordinates (s, n1, n2) of segment d: true when it contributes, with
// the row-scaled q at s, its determinant and the beam coordinates n1', n2'.
template<typename real, typename D>
__host__ __device__ inline bool window3(const D &d, real s, real n1, real n2, const Prob3D<real> &P, m2<real> &qI,
                                        real &qF, real &n1p, real &n2p)
{
    qI = madd(d.q0, mscl(s, d.dq));
    qI.a00 *= d.rcp_q0;
    qI.a10 *= d.rcp_q0;
    qI.a01 *= d.rcp_qhat0;
    qI.a11 *= d.rcp_qhat0;
    qF = mdet(qI);
    // numerators of the beam coordinates n1' = |num1 / qF|, n2' = |num2 / qF|
    const real num1 = -qI.a01 * n2 + qI.a11 * n1;
    const real num2 = qI.a00 * n2 - qI.a10 * n1;
    if(!(P.isRayCen && P.isGaussian)) {
        // receivers the beam-window test below rejects, found before its
        // square roots and divisions: its bcd is at least (|num1| + |num2|)
        // / |qF| (their max for hat beams) less two roundings, so a margin of
        // 1e-6 makes this reject only what the test rejects (as do the
        // row-length and |qF| tests it skips)
        const double N1 = std::fabs((double)num1), N2 = std::fabs((double)num2), Q = std::fabs((double)qF);
        const double lhs = P.isGaussian ? N1 + N2 : rmax(N1, N2);
        if(lhs > (P.isGaussian ? 4.0 : 1.0) * Q * (1.0 + 1e-6)) return false;
    }
    if(!P.isRayCen) { // the row-length test is Cartesian only
        real l1 = std::sqrt(qI.a00 * qI.a00 + qI.a10 * qI.a10);
        real l2 = std::sqrt(qI.a01 * qI.a01 + qI.a11 * qI.a11);
        if(std::fabs(l1) < RC<real>::EPS || std::fabs(l2) < RC<real>::EPS) return false;
    }
    if(std::fabs(qF) < RC<real>::EPS) return false;
    n1p = std::fabs(num1 / qF);
    n2p = std::fabs(num2 / qF);
    real bcd = P.isGaussian ? n1p + n2p : rmax(n1p, n2p);
    real BW = P.isGaussian ? RL(4.0) : RL(1.0);
    // the window test is not applied to ray-centred Gaussian beams in 3D
    return !(bcd > BW * RL(1.0) && !(P.isRayCen && P.isGaussian));
}

// InfluenceGeoCore (influence.hpp:897-1024, 3D) + ApplyContribution (:432-473)
// for a receiver at beam coordinates (s, n1, n2) of segment d (Cartesian
// Seg3D or ray-centred Seg3R); rayt: the direction ReceiverAngles uses.
template<typename real, typename D, typename Sink, bool FAST, int RT>
__host__ __device__ inline void core3(const D &d, real s, real n1, real n2, int itheta, int ir, int iz,
                                      const v3<real> &rayt, const Prob3D<real> &P, Sink &sink)
{
    sink.mark();
    m2<real> qI;
    real qF, n1p, n2p;
    if(!window3(d, s, n1, n2, P, qI, qF, n1p, n2p)) return;
    cpx<real> delay = d.tau0 + rmul(s, d.dtau);
    real cnst = d.A0 / std::sqrt(std::fabs(qF));
    real w;
    if(P.isGaussian) {
        if constexpr(FAST) w = (real)expf((float)(FL(-0.5) * (n1p * n1p + n2p * n2p)));
        else w = std::exp(FL(-0.5) * (n1p * n1p + n2p * n2p));
    } else {
        w = (FL(1.0) - n1p) * (FL(1.0) - n2p);
    }
    real phaseInt = d.phase0;
    if((qF <= RL(0.0) && d.qOld > RL(0.0)) || (qF >= RL(0.0) && d.qOld < RL(0.0)))
        phaseInt = phaseInt + RC<real>::PI / FL(2.0);
    size_t idx = (((size_t)d.isrc * P.Ntheta + itheta) * P.NRz + iz) * P.NRr + ir;
    if constexpr(RT == RT_ARR) {
        // ReceiverAngles, influence.hpp:123-136 (3D)
        const real RadDeg = RL(180.0) / RC<real>::PI;
        real azim = RadDeg * std::atan2(rayt.y, rayt.x);
        real decl = RadDeg * std::atan2(rayt.z, std::sqrt(rayt.x * rayt.x + rayt.y * rayt.y));
        sink.arr3(idx, cnst * w, phaseInt, delay, decl, azim, d.job, d.is, d.ntop, d.nbot, P);
        if(P.run0 == 'V' || P.run0 == 'v') sink.eig3(itheta, ir, iz, d.job, d.is, P); // IsAlsoEigenraysRun
        sink.count();
        return;
    } else if constexpr(RT == RT_EIG) {
        sink.eig3(itheta, ir, iz, d.job, d.is, P);
        sink.count();
        return;
    }
    real omega = FL(2.0) * RC<real>::PI * P.freq0;
    float fr, fi;
    if(P.run0 == 'C') {
        cpx<real> arg = rmul(omega, delay);
        arg.re = arg.re - phaseInt;
        if constexpr(FAST) {
            float sn, cs;
            sincos_phase((double)arg.re, sn, cs);
            float m = (float)(cnst * w) * expf((float)arg.im);
            fr = m * cs;
            fi = -(m * sn);
        } else {
            cpx<real> ex = cexp_(C(arg.im, -arg.re));
            cpx<real> dd = rmul(cnst * w, ex);
            fr = (float)dd.re;
            fi = (float)dd.im;
        }
    } else {
        real v;
        if constexpr(FAST) v = cnst * (real)expf((float)(omega * delay.im));
        else v = cnst * std::exp(omega * delay.im);
        v = (v * v) * w;
        if(P.isGaussian) v *= FL(2.0) * RC<real>::PI;
        fr = (float)v;
        fi = 0.0f;
    }
    sink.add3(idx, fr, fi);
    sink.count();
}

// ---------------------------------------------------------------------------
// Ray-centred beams in 3D: Step_InfluenceGeoRayCen (influence.hpp:1030-1194)
// ---------------------------------------------------------------------------

// Step descriptor of a ray-centred segment: the two points with their ray
// normals (c-scaled, RayNormalRayCen) and c t (e1 x e2), and the beam state.
template<typename real> struct Seg3R {
    v3<real> xa, xb;             // points A (previous) and B
    v3<real> n1a, n2a, n1b, n2b; // ray normals at A and B
    v3<real> ca, cb;             // c t at A and B
    m2<real> q0, dq;
    cpx<real> tau0, dtau;
    real A0, phase0, qOld, rcp_q0, rcp_qhat0, xs_x, xs_y;
    real rmax_m, rmax_n;         // Gaussian: BeamWindow x the larger |q| (scaled) at A and B, per normal
    int32_t nitems;              // > 0: rows (depth, bearing) to visit
    int32_t isrc, job, is, ntop, nbot;
};

template<typename real>
__host__ __device__ inline void prelude3rc(Seg3R<real> &d, const Ray3<real> &p0, const Ray3<real> &p1,
                                           Infl3<real> &I, int is, const Prob3D<real> &P, int isz, int isx, int isy)
{
    d.nitems = 0;
    real phaseq = mdet(p0.q);
    if((phaseq <= RL(0.0) && I.qOld > RL(0.0)) || (phaseq >= RL(0.0) && I.qOld < RL(0.0)))
        I.phase += RC<real>::PI / FL(2.0);
    I.qOld = phaseq;
    v3<real> rn1, rn2;
    RayNormal3(p1.t, p1.phi, false, p1.c, rn1, rn2);
    // IsDuplicatePoint: hat beams compare with a constant 1e-4 m threshold
    v3<real> dx = p1.x - p0.x;
    real thresh = P.isGaussian ? RL(1.0e3) * spacing(p1.x.x) : (real)FL(1.0e-4);
    bool dup = len3(dx) <= thresh;
    if(!dup) {
        d.xa = p0.x;
        d.xb = p1.x;
        d.n1a = I.rn1;
        d.n2a = I.rn2;
        d.n1b = rn1;
        d.n2b = rn2;
        d.ca = V3(p0.c * p0.t.x, p0.c * p0.t.y, p0.c * p0.t.z);
        d.cb = V3(p1.c * p1.t.x, p1.c * p1.t.y, p1.c * p1.t.z);
        d.q0 = p0.q;
        d.dq = msub(p1.q, p0.q);
        d.tau0 = p0.tau;
        d.dtau = p1.tau - p0.tau;
        d.A0 = I.Ratio1 * p1.c * p1.Amp;
        d.phase0 = p0.Phase + I.phase;
        d.qOld = I.qOld;
        d.rcp_q0 = I.rcp_q0;
        d.rcp_qhat0 = I.rcp_qhat0;
        d.xs_x = I.xs.x;
        d.xs_y = I.xs.y;
        const real BW = P.isGaussian ? RL(4.0) : RL(1.0);
        d.rmax_m = BW * rmax(std::fabs(p0.q.a11 * I.rcp_qhat0), std::fabs(p1.q.a11 * I.rcp_qhat0));
        d.rmax_n = BW * rmax(std::fabs(p0.q.a00 * I.rcp_q0), std::fabs(p1.q.a00 * I.rcp_q0));
        d.isrc = (isz * P.NSx + isx) * P.NSy + isy;
        d.is = is;
        d.ntop = p1.NumTopBnc;
        d.nbot = p1.NumBotBnc;
        d.nitems = P.NRz * P.Ntheta;
    }
    I.rn1 = rn1;
    I.rn2 = rn2;
}

// RToIR, influence.hpp:82-89 (uniform receiver ranges)
template<typename real> __host__ __device__ inline int rtoir3(real r, const real *Rr, const Prob3D<real> &P)
{
    real temp = (r - Rr[0]) / P.Delta_r;
    if(std::fabs(temp - std::round(temp)) < RL(1e-6)) temp = std::round(temp);
    return imax(imin(trunc_int_x86(temp), P.NRr - 1), 0);
}

// Compute_RayCenCross + Compute_M_N_R_IR (influence.hpp:154-189) at one end
// of the segment for receiver depth zR and bearing (ct, st); false when the
// ray normal is parallel to the receiver line.
template<typename real>
__host__ __device__ inline bool mnr3(const v3<real> &x, const v3<real> &n1, const v3<real> &n2, const v3<real> &e12,
                                     real xsx, real xsy, real zR, real ct, real st, const real *Rr,
                                     const Prob3D<real> &P, real &m, real &n, real &r, int &ir)
{
    v3<real> xt = x - V3(xsx, xsy, zR);
    v3<real> c1 = cross3(xt, n1), c2 = cross3(xt, n2);
    real delta = -(ct * e12.x + st * e12.y);
    if(std::fabs(delta) < FL(1e-3)) return false;
    m = (ct * c1.x + st * c1.y) / delta;
    n = -(ct * c2.x + st * c2.y) / delta;
    r = -dot3(xt, e12) / delta;
    ir = rtoir3(r, Rr, P);
    return true;
}

// One (depth, bearing) row of a ray-centred segment: the bracketed receiver
// ranges irA + notii ... irB + notii - step and the interpolation ends.
// Returns the number of receivers (0: the row is skipped).
template<typename real>
__host__ __device__ inline int row3rc(const Seg3R<real> &d, int iz, int itheta, const Prob3D<real> &P, const real *Rr,
                                      const real *trcv, const float *Rz, real &mA, real &nA, real &rA, real &mB,
                                      real &nB, real &rB, int &first, int &step)
{
    const real zR = (real)Rz[iz];
    const real ct = trcv[2 * itheta], st = trcv[2 * itheta + 1];
    int irA, irB;
    if(!mnr3(d.xa, d.n1a, d.n2a, d.ca, d.xs_x, d.xs_y, zR, ct, st, Rr, P, mA, nA, rA, irA)) return 0;
    if(!mnr3(d.xb, d.n1b, d.n2b, d.cb, d.xs_x, d.xs_y, zR, ct, st, Rr, P, mB, nB, rB, irB)) return 0;
    if(P.isGaussian) {
        if(d.rmax_m <= rmin(std::fabs(mA), std::fabs(mB)) && mA * mB >= RL(0.0)) return 0;
        if(d.rmax_n <= rmin(std::fabs(nA), std::fabs(nB)) && nA * nB >= RL(0.0)) return 0;
    }
    if(irA == irB) return 0;
    const int notii = (irB <= irA) ? 0 : 1;
    step = (notii << 1) - 1;
    first = irA + notii;
    return (irB - irA) * step;
}

// Receiver k of a ray-centred row.
template<typename real, typename Sink, bool FAST, int RT>
__host__ __device__ inline void item3rc(const Seg3R<real> &d, int iz, int itheta, int ir, real mA, real nA, real rA,
                                        real mB, real nB, real rB, const Prob3D<real> &P, const real *Rr, Sink &sink)
{
    real sw = (Rr[ir] - rA) / (rB - rA);
    real n1 = std::fabs(nA + sw * (nB - nA));
    real n2 = std::fabs(mA + sw * (mB - mA));
    core3<real, Seg3R<real>, Sink, FAST, RT>(d, sw, n1, n2, itheta, ir, iz, d.xb - d.xa, P, sink);
}

} // namespace bhc3d
