#ifndef _GNU_SOURCE
#define _GNU_SOURCE
#endif
// host2d.cpp -- host front end of the 2D hot path: preprocessing of a raw
// environment into the problem the launch layer consumes, ScalePressure, and
// the .shd / .arr / .ray writers.
//
// Preprocessing follows the reference's params modules (Preprocess() of
// src/module/{ssp,boundary,boundarycond,reflcoef,beaminfo,rayangles,
// rcvrranges,szrz,sbp}.hpp and crci, src/module/atten.cpp:99-194), in the run's
// real type, with std::complex arithmetic exactly as the reference uses it,
// because the tracer's trajectories depend on these values bit for bit.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bhc_hip.h"

namespace {

thread_local std::string g_host_error;
int hfail(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_host_error = buf;
    return -1;
}

template<typename T> struct Lit;
template<> struct Lit<double> {
    static double pick(double d, float) { return d; }
    static constexpr double EPS = 2.220446049250313e-16;
    static constexpr double MAXV = 1.7976931348623157e308;
    static constexpr double PI = M_PI;
};
template<> struct Lit<float> {
    static float pick(double, float f) { return f; }
    static constexpr float EPS = 1.1920929e-07f;
    static constexpr float MAXV = 3.4028235e38f;
    static constexpr float PI = (float)M_PI;
};
#define RLH(a) (Lit<real>::pick(a, a##f))
#define FLH(a) (a##f)

template<typename real> using cx = std::complex<real>;

inline void host_sincos(double a, double *s, double *c) { ::sincos(a, s, c); }
inline void host_sincos(float a, float *s, float *c) { ::sincosf(a, s, c); }

template<typename real> struct Pre {
    const bhc_env2d *env;
    real freq0;

    // crci, src/module/atten.cpp:99-194 (Thorp 'T', Francois-Garrison 'F' and
    // biological-layer 'B' volume terms); z selects the biological layers
    cx<real> crci(real z, real c, real alpha, const char *au) const
    {
        real freq = freq0;
        real omega = FLH(2.0) * Lit<real>::PI * freq;
        real alphaT = FLH(0.0);
        switch(au[0]) {
        case 'N': alphaT = alpha; break;
        case 'M': alphaT = alpha / RLH(8.6858896); break;
        case 'm':
            // power law in f / freq0: the run frequency is freq0, where both
            // branches of the reference (below and above fT) give the factor 1
            alphaT = alpha / RLH(8.6858896);
            alphaT *= std::pow(freq / freq0, FLH(1.0));
            break;
        case 'F': alphaT = alpha * freq / RLH(8685.8896); break;
        case 'W':
            if(std::abs(c) >= Lit<real>::EPS) alphaT = alpha * freq / (RLH(8.6858896) * c);
            break;
        case 'Q':
            if(std::abs(c * alpha) >= Lit<real>::EPS) alphaT = omega / (FLH(2.0) * c * alpha);
            break;
        case 'L':
            if(std::abs(c) >= Lit<real>::EPS) alphaT = alpha * omega / c;
            break;
        default: break;
        }
        if(au[1] == 'T') {
            real f2 = (freq / FLH(1000.0)) * (freq / FLH(1000.0));
            real Thorp = FLH(3.3e-3) + FLH(0.11) * f2 / (FLH(1.0) + f2) + FLH(44.0) * f2 / (FLH(4100.0) + f2)
                + FLH(3e-4) * f2;
            Thorp /= FLH(8685.8896);
            alphaT += Thorp;
        } else if(au[1] == 'F') {
            real fg = franc_garr(freq / FLH(1000.0));
            fg /= FLH(8685.8896);
            alphaT += fg;
        } else if(au[1] == 'B') {
            for(int ib = 0; ib < env->NBioLayers; ++ib) {
                const double *L = env->bio + 5 * ib;
                real z1 = (real)L[0], z2 = (real)L[1], f0 = (real)L[2], q = (real)L[3], a0 = (real)L[4];
                if(z >= z1 && z <= z2) {
                    // the reference's 1.0 literals are double: the sum is
                    // formed in double precision in both builds (dB/km)
                    double u = 1.0 - (double)((f0 * f0) / (freq * freq));
                    real a = (real)((double)a0 / (u * u + 1.0 / (double)(q * q)));
                    a /= FLH(8685.8896);
                    alphaT += a;
                }
            }
        }
        alphaT *= (c * c) / omega;
        return cx<real>(c, alphaT);
    }

    real franc_garr(real f) const
    { // src/module/atten.cpp:38-72
        real T = (real)env->fg_T, S = (real)env->fg_S, pH = (real)env->fg_pH, zb = (real)env->fg_z_bar;
        real c = FLH(1412.0) + FLH(3.21) * T + FLH(1.19) * S + FLH(0.0167) * zb;
        real a1 = FLH(8.86) / c * std::pow(FLH(10.0), FLH(0.78) * pH - FLH(5.0));
        real p1 = FLH(1.0);
        real f1 = FLH(2.8) * std::sqrt(S / 35) * std::pow(FLH(1.0), FLH(4.0) - FLH(1245.0) / (T + FLH(273.0)));
        real a2 = FLH(21.44) * S / c * (FLH(1.0) + FLH(0.025) * T);
        real p2 = FLH(1.0) - RLH(1.37e-4) * zb + RLH(6.2e-9) * (zb * zb);
        real f2 = FLH(8.17) * std::pow(FLH(10.0), FLH(8.0) - FLH(1990.0) / (T + 273))
            / (FLH(1.0) + FLH(0.0018) * (S - FLH(35.0)));
        real p3 = FLH(1.0) - RLH(3.83e-5) * zb + RLH(4.9e-10) * (zb * zb);
        real a3;
        if(T < FLH(20.0))
            a3 = RLH(4.937e-4) - RLH(2.59e-5) * T + RLH(9.11e-7) * (T * T) - RLH(1.5e-8) * (T * T * T);
        else
            a3 = RLH(3.964e-4) - RLH(1.146e-5) * T + RLH(1.45e-7) * (T * T) - RLH(6.5e-10) * (T * T * T);
        return a1 * p1 * (f1 * (f * f)) / ((f1 * f1) + (f * f)) + a2 * p2 * (f2 * (f * f)) / ((f2 * f2) + (f * f))
            + a3 * p3 * (f * f);
    }
};

// cSpline (de Boor), src/curves.hpp:76-188.
template<typename real>
void cspline(const real *tau, cx<real> *c1, cx<real> *c2, cx<real> *c3, cx<real> *c4, int n, int ibcbeg, int ibcend)
{
    cx<real> g, dtau, divdf1, divdf3;
    int l = n - 1;
    for(int m = 1; m < n; ++m) {
        c3[m] = cx<real>(tau[m] - tau[m - 1], FLH(0.0));
        c4[m] = (c1[m] - c1[m - 1]) / c3[m];
    }
    const cx<real> two(FLH(2.0)), three(FLH(3.0)), half(FLH(0.5)), six(FLH(6.0)), one(FLH(1.0));
    if(ibcbeg == 0) {
        if(n > 2) {
            c4[0] = c3[2];
            c3[0] = c3[1] + c3[2];
            c2[0] = ((c3[1] + (real)FLH(2.0) * c3[0]) * c4[1] * c3[2] + (c3[1] * c3[1]) * c4[2]) / c3[0];
        } else {
            c4[0] = cx<real>(FLH(1.0), FLH(0.0));
            c3[0] = cx<real>(FLH(1.0), FLH(0.0));
            c2[0] = (real)FLH(2.0) * c4[1];
        }
    } else if(ibcbeg == 1) {
        c4[0] = cx<real>(FLH(1.0), FLH(0.0));
        c3[0] = cx<real>(FLH(0.0), FLH(0.0));
    } else if(ibcbeg == 2) {
        c4[0] = cx<real>(FLH(2.0), FLH(0.0));
        c3[0] = cx<real>(FLH(1.0), FLH(0.0));
        c2[0] = (real)FLH(3.0) * c4[1] - c2[0] * c3[1] * (real)FLH(0.5);
    }
    for(int m = 1; m < l; ++m) {
        g = -c3[m + 1] / c4[m - 1];
        c2[m] = g * c2[m - 1] + (real)FLH(3.0) * (c3[m] * c4[m + 1] + c3[m + 1] * c4[m]);
        c4[m] = g * c3[m - 1] + (real)FLH(2.0) * (c3[m] + c3[m + 1]);
    }
    if(ibcend != 1) {
        if(ibcend == 0) {
            if(n == 2 && ibcbeg == 0) {
                c2[n - 1] = c4[n - 1];
            } else if((n == 3 && ibcbeg == 0) || n == 2) {
                c2[n - 1] = (real)FLH(2.0) * c4[n - 1];
                c4[n - 1] = cx<real>(FLH(1.0), FLH(0.0));
                g = (real)FLH(1.0) / -c4[n - 2];
            } else {
                g = c3[n - 2] + c3[n - 1];
                c2[n - 1] = ((c3[n - 1] + (real)FLH(2.0) * g) * c4[n - 1] * c3[n - 2]
                             + (c3[n - 1] * c3[n - 1]) * (c1[n - 2] - c1[n - 3]) / c3[n - 2])
                    / g;
                g = -g / c4[n - 2];
                c4[n - 1] = c3[n - 2];
            }
        } else if(ibcend == 2) {
            c2[n - 1] = (real)FLH(3.0) * c4[n - 1] + c2[n - 1] * c3[n - 1] * (real)FLH(0.5);
            c4[n - 1] = cx<real>(FLH(2.0), FLH(0.0));
            g = (real)FLH(1.0) / -c4[n - 2];
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
        divdf3 = c2[i - 1] + c2[i] - (real)FLH(2.0) * divdf1;
        c3[i - 1] = (real)FLH(2.0) * (divdf1 - c2[i - 1] - divdf3) / dtau;
        c4[i - 1] = (divdf3 / dtau) * ((real)FLH(6.0) / dtau);
    }
    c3[n - 1] = c3[l - 1] + (tau[n - 1] - tau[l - 1]) * c4[l - 1];
    c4[n - 1] = cx<real>(FLH(0.0), FLH(0.0));
    for(int i = 0; i < l; ++i) {
        dtau = tau[i + 1] - tau[i];
        c4[n - 1] += dtau
            * (c1[i] + dtau * (c2[i] * (real)FLH(0.5) + dtau * (c3[i] / (real)FLH(6.0) + dtau * c4[i] / (real)FLH(24.0))));
    }
    c4[n - 1] /= tau[n - 1] - tau[0];
    (void)two; (void)three; (void)half; (void)six; (void)one;
}

// PCHIP, src/curves.hpp:225-372.
template<typename real> real fpi(real d1, real d2, real fp)
{
    if(d1 * d2 > FLH(0.0)) {
        if(d1 > FLH(0.0)) return std::min(std::max(fp, RLH(0.0)), RLH(3.0) * std::min(d1, d2));
        return std::max(std::min(fp, RLH(0.0)), RLH(3.0) * std::max(d1, d2));
    }
    return RLH(0.0);
}
template<typename real> real fpl(real d1, real d2, real fp)
{
    if(d1 * fp <= RLH(0.0)) return FLH(0.0);
    if((d1 * d2 <= RLH(0.0)) && (std::abs(fp) > std::abs(RLH(3.0) * d1))) return RLH(3.0) * d1;
    return fp;
}

template<typename real> void pchip(const real *x, const cx<real> *y, int n, cx<real> *P1, cx<real> *P2, cx<real> *P3,
                                   cx<real> *P4)
{
    std::vector<cx<real>> W1(n), W2(n), W3(n), W4(n);
    if(n == 2) {
        P1[0] = y[0];
        P2[0] = (y[1] - y[0]) / (x[1] - x[0]);
        P3[0] = RLH(0.0);
        P4[0] = RLH(0.0);
        return;
    }
    auto hdel = [&](int ix, real &h1, real &h2, cx<real> &d1, cx<real> &d2) {
        h1 = x[ix] - x[ix - 1];
        h2 = x[ix + 1] - x[ix];
        d1 = (y[ix] - y[ix - 1]) / h1;
        d2 = (y[ix + 1] - y[ix]) / h2;
    };
    real h1, h2;
    cx<real> d1, d2, fpT;
    for(int i = 0; i < n; ++i) P1[i] = y[i];
    hdel(1, h1, h2, d1, d2);
    fpT = ((RLH(2.0) * h1 + h2) * d1 - h1 * d2) / (h1 + h2);
    P2[0] = cx<real>(fpl(d1.real(), d2.real(), fpT.real()), fpl(d1.imag(), d2.imag(), fpT.imag()));
    hdel(n - 2, h1, h2, d1, d2);
    fpT = (-h2 * d1 + (h1 + RLH(2.0) * h2) * d2) / (h1 + h2);
    P2[n - 1] = cx<real>(fpl(d2.real(), d1.real(), fpT.real()), fpl(d2.imag(), d1.imag(), fpT.imag()));
    for(int i = 0; i < n; ++i) W1[i] = P1[i];
    W2[0] = P2[0];
    W2[n - 1] = P2[n - 1];
    cspline<real>(x, W1.data(), W2.data(), W3.data(), W4.data(), n, 1, 1);
    for(int ix = 1; ix < n - 1; ++ix) {
        hdel(ix, h1, h2, d1, d2);
        P2[ix] = cx<real>(fpi(d1.real(), d2.real(), W2[ix].real()), fpi(d1.imag(), d2.imag(), W2[ix].imag()));
    }
    for(int ix = 0; ix < n - 1; ++ix) {
        real h = x[ix + 1] - x[ix];
        cx<real> f1 = P1[ix], f2 = P1[ix + 1], f1p = P2[ix], f2p = P2[ix + 1];
        P3[ix] = (RLH(3.0) * (f2 - f1) - h * (RLH(2.0) * f1p + f2p)) / (h * h);
        P4[ix] = (h * (f1p + f2p) - RLH(2.0) * (f2 - f1)) / (h * h * h);
    }
}

} // namespace

// Owner of a preprocessed problem's arrays (all stored as double).
struct bhc_problem_owner {
    bhc_problem2d v;
    std::vector<double> ssp_z, ssp_c, ssp_cz, ssp_n2, ssp_n2z, ssp_coef, ssp_rho, ssp_segr, ssp_cmat, ssp_czmat;
    std::vector<bhc_bdry_node2d> ati, bty;
    std::vector<double> trc, brc, Rr, alpha, acos_, asin_, sbp;
    std::vector<float> Sz, Rz;
};

namespace {

template<typename real> void put_c(std::vector<double> &dst, const cx<real> &c)
{
    dst.push_back((double)c.real());
    dst.push_back((double)c.imag());
}

template<typename real>
int bdry(const Pre<real> &pre, const bhc_bdry_in &in, const bhc_halfspace_in &hs, bool isTop,
         std::vector<bhc_bdry_node2d> &out, char type[2])
{ // src/module/boundary.hpp:56-83, 203-226, 499-544, 600-890 (2D)
    const real NegTop = isTop ? RLH(-1.0) : RLH(1.0);
    const real BDRYBIG = std::sqrt(Lit<real>::MAXV) / RLH(1.0e5);
    struct N {
        real x[2], t[2], n[2], Noden[2], Nodet[2], Len, kappa, Dx, Dxx, Dss;
        cx<real> cP, cS;
        real rho;
    };
    std::vector<N> bd;
    if(in.npts <= 0) {
        type[0] = 'L';
        type[1] = 'S';
        bd.assign(2, N{});
        bd[0].x[0] = -BDRYBIG;
        bd[0].x[1] = (real)hs.depth;
        bd[1].x[0] = BDRYBIG;
        bd[1].x[1] = (real)hs.depth;
    } else {
        type[0] = in.type[0];
        type[1] = in.type[1];
        if(type[0] != 'L' && type[0] != 'C') return hfail("boundary type '%c' not supported", type[0]);
        int n = in.npts + 2;
        bd.assign(n, N{});
        for(int i = 1; i < n - 1; ++i) {
            bd[i].x[0] = (real)in.r_km[i - 1];
            bd[i].x[1] = (real)in.z[i - 1];
            if(type[1] == 'L') {
                if(!in.alphaR || !in.alphaI || !in.betaR || !in.betaI || !in.rho)
                    return hfail("long-format boundary needs per-point geoacoustics");
                bd[i].cP = pre.crci(RLH(1.0e20), (real)in.alphaR[i - 1], (real)in.alphaI[i - 1], "W ");
                bd[i].cS = pre.crci(RLH(1.0e20), (real)in.betaR[i - 1], (real)in.betaI[i - 1], "W ");
                bd[i].rho = (real)in.rho[i - 1];
            }
            if(i > 1 && !(bd[i].x[0] > bd[i - 1].x[0])) return hfail("boundary ranges must increase");
        }
        bd[0].x[0] = -BDRYBIG * RLH(0.001);
        bd[0].x[1] = bd[1].x[1];
        bd[n - 1].x[0] = BDRYBIG * RLH(0.001);
        bd[n - 1].x[1] = bd[n - 2].x[1];
        bd[0].cP = bd[1].cP;
        bd[0].cS = bd[1].cS;
        bd[0].rho = bd[1].rho;
        bd[n - 1].cP = bd[n - 2].cP;
        bd[n - 1].cS = bd[n - 2].cS;
        bd[n - 1].rho = bd[n - 2].rho;
    }
    int n = (int)bd.size();
    for(int i = 0; i < n; ++i) bd[i].x[0] *= RLH(1000.0);
    for(int ii = 0; ii < n - 1; ++ii) {
        N &b = bd[ii];
        b.t[0] = bd[ii + 1].x[0] - b.x[0];
        b.t[1] = bd[ii + 1].x[1] - b.x[1];
        b.Dx = b.t[1] / b.t[0];
        b.Len = std::sqrt(b.t[0] * b.t[0] + b.t[1] * b.t[1]);
        b.t[0] /= b.Len;
        b.t[1] /= b.Len;
        b.n[0] = -NegTop * b.t[1];
        b.n[1] = NegTop * b.t[0];
    }
    if(type[0] == 'C') {
        for(int ii = 1; ii < n - 1; ++ii) {
            real sss = FLH(0.5);
            bd[ii].Nodet[0] = (FLH(1.0) - sss) * bd[ii - 1].t[0] + sss * bd[ii].t[0];
            bd[ii].Nodet[1] = (FLH(1.0) - sss) * bd[ii - 1].t[1] + sss * bd[ii].t[1];
        }
        bd[0].Nodet[0] = FLH(1.0);
        bd[0].Nodet[1] = FLH(0.0);
        bd[n - 1].Nodet[0] = FLH(1.0);
        bd[n - 1].Nodet[1] = FLH(0.0);
        for(int ii = 0; ii < n; ++ii) {
            bd[ii].Noden[0] = -NegTop * bd[ii].Nodet[1];
            bd[ii].Noden[1] = NegTop * bd[ii].Nodet[0];
        }
        real curphi = std::atan2(bd[0].Nodet[1], bd[0].Nodet[0]);
        for(int ii = 0; ii < n - 1; ++ii) {
            real nextphi = std::atan2(bd[ii + 1].Nodet[1], bd[ii + 1].Nodet[0]);
            bd[ii].kappa = (nextphi - curphi) / bd[ii].Len;
            bd[ii].Dxx = (bd[ii + 1].Dx - bd[ii].Dx) / (bd[ii + 1].x[0] - bd[ii].x[0]);
            bd[ii].Dss = bd[ii].Dxx * (bd[ii].t[0] * bd[ii].t[0] * bd[ii].t[0]);
            bd[ii].kappa = bd[ii].Dss;
            curphi = nextphi;
        }
    } else {
        for(int i = 0; i < n; ++i) bd[i].kappa = FLH(0.0);
    }
    out.resize(n);
    for(int i = 0; i < n; ++i) {
        bhc_bdry_node2d &o = out[i];
        const N &b = bd[i];
        for(int k = 0; k < 2; ++k) {
            o.x[k] = b.x[k];
            o.t[k] = b.t[k];
            o.n[k] = b.n[k];
            o.Noden[k] = b.Noden[k];
            o.Nodet[k] = b.Nodet[k];
        }
        o.Len = b.Len;
        o.kappa = b.kappa;
        o.cP[0] = b.cP.real();
        o.cP[1] = b.cP.imag();
        o.cS[0] = b.cS.real();
        o.cS[1] = b.cS.imag();
        o.rho = b.rho;
    }
    return 0;
}

template<typename real> int halfspace(const Pre<real> &pre, const bhc_halfspace_in &in, bhc_hs2d &out)
{ // src/module/boundarycond.hpp:166-224
    memset(&out, 0, sizeof(out));
    out.bc = in.bc;
    out.depth = in.depth;
    real rho = (real)in.rho;
    cx<real> cP(0), cS(0);
    const char *au = pre.env->atten_unit;
    if(in.bc == 'A') {
        cP = pre.crci((real)in.zTemp, (real)in.alphaR, (real)in.alphaI, au);
        cS = pre.crci((real)in.zTemp, (real)in.betaR, (real)in.betaI, au);
    } else if(in.bc == 'G') {
        real Mz = (real)in.Mz, vr, a2f;
        if(Mz >= FLH(-1.0) && Mz < FLH(1.0)) {
            vr = FLH(0.002709) * (Mz * Mz) - FLH(0.056452) * Mz + FLH(1.2778);
            rho = FLH(0.007797) * (Mz * Mz) - FLH(0.17057) * Mz + FLH(2.3139);
        } else if(Mz >= FLH(1.0) && Mz < FLH(5.3)) {
            vr = FLH(-0.0014881) * (Mz * Mz * Mz) + FLH(0.0213937) * (Mz * Mz) - FLH(0.1382798) * Mz + FLH(1.3425);
            rho = FLH(-0.0165406) * (Mz * Mz * Mz) + FLH(0.2290201) * (Mz * Mz) - FLH(1.1069031) * Mz + FLH(3.0455);
        } else {
            vr = FLH(-0.0024324) * Mz + FLH(1.0019);
            rho = FLH(-0.0012973) * Mz + FLH(1.1565);
        }
        if(Mz >= FLH(-1.0) && Mz < FLH(0.0)) a2f = FLH(0.4556);
        else if(Mz >= FLH(0.0) && Mz < FLH(2.6)) a2f = FLH(0.4556) + FLH(0.0245) * Mz;
        else if(Mz >= FLH(2.6) && Mz < FLH(4.5)) a2f = FLH(0.1978) + FLH(0.1245) * Mz;
        else if(Mz >= FLH(4.5) && Mz < FLH(6.0)) a2f = FLH(8.0399) - FLH(2.5228) * Mz + FLH(0.20098) * (Mz * Mz);
        else if(Mz >= FLH(6.0) && Mz < FLH(9.5)) a2f = FLH(0.9431) - FLH(0.2041) * Mz + FLH(0.0117) * (Mz * Mz);
        else a2f = FLH(0.0601);
        real aR = vr * FLH(1500.0);
        real aI = a2f * (vr / FLH(1000.0)) * FLH(1500.0) * std::log(FLH(10.0)) / (FLH(40.0) * Lit<real>::PI);
        cP = pre.crci((real)in.zTemp, aR, aI, "L ");
        cS = FLH(0.0);
    } else if(in.bc != 'V' && in.bc != 'R' && in.bc != 'F') {
        return hfail("boundary condition '%c' not supported", in.bc);
    }
    out.cP[0] = cP.real();
    out.cP[1] = cP.imag();
    out.cS[0] = cS.real();
    out.cS[1] = cS.imag();
    out.rho = rho;
    return 0;
}

template<typename real> int preprocess(const bhc_env2d *env, bhc_problem_owner *o)
{
    Pre<real> pre{env, (real)env->freq0};
    bhc_problem2d &v = o->v;
    memset(&v, 0, sizeof(v));
    v.precision = env->precision;
    memcpy(v.RunType, env->RunType, 7);
    // MoveBeamType, src/module/beaminfo.hpp:241-246
    v.BeamType[0] = env->RunType[1];
    v.BeamType[1] = env->BeamType[0];
    v.BeamType[2] = env->BeamType[1];
    v.BeamType[3] = env->RunType[6];
    v.ssp_type = env->ssp_type;
    v.freq0 = env->freq0;
    // SSP, src/module/ssp.hpp:453-545
    int N = env->ssp_npts;
    if(env->ssp_type != 'A' && N < 2) return hfail("SSP needs at least 2 points");
    std::vector<real> z(std::max(N, 0));
    std::vector<cx<real>> c(N), cz(N), n2(N), n2z(N), cf[4];
    for(int i = 0; i < N; ++i) z[i] = (real)env->ssp_z[i];
    for(int i = 1; i < N; ++i)
        if(!(z[i] > z[i - 1])) return hfail("SSP depths must increase");
    for(int iz = 0; iz < N; ++iz) {
        c[iz] = pre.crci(z[iz], (real)env->ssp_alphaR[iz], (real)env->ssp_alphaI[iz], env->atten_unit);
        if(iz > 0) cz[iz - 1] = (c[iz] - c[iz - 1]) / (z[iz] - z[iz - 1]);
    }
    if(N > 0) cz[N - 1] = cx<real>(NAN, NAN);
    for(auto &k : cf) k.assign(N, cx<real>(0));
    switch(env->ssp_type) {
    case 'N':
        for(int i = 0; i < N; ++i) n2[i] = (real)FLH(1.0) / (c[i] * c[i]);
        for(int iz = 1; iz < N; ++iz) n2z[iz - 1] = (n2[iz] - n2[iz - 1]) / (z[iz] - z[iz - 1]);
        break;
    case 'S':
        for(int i = 0; i < N; ++i) cf[0][i] = c[i];
        cspline<real>(z.data(), cf[0].data(), cf[1].data(), cf[2].data(), cf[3].data(), N, 0, 0);
        break;
    case 'P': pchip<real>(z.data(), c.data(), N, cf[0].data(), cf[1].data(), cf[2].data(), cf[3].data()); break;
    case 'C':
    case 'A': break;
    case 'Q': {
        int Nr = env->ssp_nr;
        if(Nr < 2 || !env->ssp_r_km || !env->ssp_cmat) return hfail("quad SSP needs >= 2 profiles");
        v.ssp_nr = Nr;
        for(int i = 0; i < Nr; ++i) o->ssp_segr.push_back((double)((real)env->ssp_r_km[i] * FLH(1000.0)));
        std::vector<real> cm((size_t)N * Nr), czm((size_t)N * Nr, (real)0);
        for(size_t i = 0; i < cm.size(); ++i) cm[i] = (real)env->ssp_cmat[i];
        for(int s = 0; s < Nr; ++s)
            for(int iz2 = 1; iz2 < N; ++iz2) {
                real dz = z[iz2] - z[iz2 - 1];
                czm[(iz2 - 1) * Nr + s] = (cm[iz2 * Nr + s] - cm[(iz2 - 1) * Nr + s]) / dz;
            }
        for(size_t i = 0; i < cm.size(); ++i) {
            o->ssp_cmat.push_back((double)cm[i]);
            o->ssp_czmat.push_back((double)czm[i]);
        }
    } break;
    default: return hfail("SSP type '%c' not supported in 2D", env->ssp_type);
    }
    for(int i = 0; i < N; ++i) {
        o->ssp_z.push_back((double)z[i]);
        o->ssp_rho.push_back((double)(real)env->ssp_rho[i]);
        put_c(o->ssp_c, c[i]);
        put_c(o->ssp_cz, cz[i]);
        put_c(o->ssp_n2, n2[i]);
        put_c(o->ssp_n2z, n2z[i]);
    }
    for(int k = 0; k < 4; ++k)
        for(int i = 0; i < N; ++i) put_c(o->ssp_coef, cf[k][i]);
    v.ssp_npts = N;
    // boundary conditions and boundaries
    if(halfspace<real>(pre, env->top, v.top_hs) || halfspace<real>(pre, env->bot, v.bot_hs)) return -1;
    if(bdry<real>(pre, env->ati, env->top, true, o->ati, v.ati_type)
       || bdry<real>(pre, env->bty, env->bot, false, o->bty, v.bty_type))
        return -1;
    v.ati_npts = (int32_t)o->ati.size();
    v.bty_npts = (int32_t)o->bty.size();
    // reflection tables (angles stay in degrees; phase to radians), src/module/reflcoef.hpp:156-167
    const real DegRad = Lit<real>::PI / RLH(180.0);
    auto refl = [&](const bhc_refl_in &in, std::vector<double> &dst) {
        for(int i = 0; i < in.npts; ++i) {
            dst.push_back((double)(real)in.theta[i]);
            dst.push_back((double)(real)in.r[i]);
            dst.push_back((double)((real)in.phi[i] * DegRad));
        }
        return in.npts;
    };
    v.trc_npts = refl(env->trc, o->trc);
    v.brc_npts = refl(env->brc, o->brc);
    if((env->top.bc == 'F' && v.trc_npts < 2) || (env->bot.bc == 'F' && v.brc_npts < 2))
        return hfail("reflection table ('F') needs at least 2 points");
    // positions, src/module/rcvrranges.hpp:74-85, szrz.hpp:138-142
    v.NSz = env->NSz;
    v.NRz = env->NRz;
    v.NRr = env->NRr;
    o->Sz.assign(env->Sz, env->Sz + env->NSz);
    o->Rz.assign(env->Rz, env->Rz + env->NRz);
    std::vector<real> Rr(env->NRr);
    for(int i = 0; i < env->NRr; ++i) Rr[i] = (real)env->Rr_km[i] * FLH(1000.0);
    for(int i = 0; i < env->NRr; ++i) o->Rr.push_back((double)Rr[i]);
    real Delta_r = FLH(0.0);
    if(env->NRr >= 2) Delta_r = Rr[env->NRr - 1] - Rr[env->NRr - 2];
    v.Delta_r = (double)Delta_r;
    v.NRz_per_range = (env->RunType[4] == 'I') ? 1 : env->NRz;
    // angles, src/module/rayangles.hpp:157-190
    v.Nalpha = env->Nalpha;
    v.iSingle_alpha = env->iSingle_alpha;
    std::vector<real> al(env->Nalpha);
    for(int i = 0; i < env->Nalpha; ++i) al[i] = (real)env->alpha_deg[i] * DegRad;
    for(int i = 1; i < env->Nalpha; ++i)
        if(!(al[i] > al[i - 1])) return hfail("launch angles must increase");
    real Dalpha = FLH(0.0);
    if(env->Nalpha != 1) Dalpha = (al[env->Nalpha - 1] - al[0]) / (real)(env->Nalpha - 1);
    v.Dalpha = (double)Dalpha;
    for(int i = 0; i < env->Nalpha; ++i) {
        o->alpha.push_back((double)al[i]);
        // launch tangent (cos, sin), src/trace.hpp:61, with the host libm. The
        // reference evaluates cos(alpha) and sin(alpha) in one expression, which
        // g++ -O3 turns into one sincos call; glibc's sincos can differ from sin
        // in the last bit, so it is called here the same way.
        real sn, cs;
        host_sincos(al[i], &sn, &cs);
        o->acos_.push_back((double)cs);
        o->asin_.push_back((double)sn);
    }
    // beam, src/module/beaminfo.hpp:221-238
    v.box_r = (double)((real)env->box_r_km * FLH(1000.0));
    v.box_z = (double)(real)env->box_z;
    real deltas = (real)env->deltas;
    if(std::abs(deltas) < Lit<real>::EPS) deltas = ((real)env->bot.depth - (real)env->top.depth) / FLH(10.0);
    v.deltas = (double)deltas;
    v.epsMultiplier = env->epsMultiplier;
    v.rLoop = env->rLoop;
    v.iBeamWindow = env->iBeamWindow;
    v.Nimage = env->Nimage;
    // source beam pattern, src/module/sbp.hpp:43-53, 143-153
    if(env->NSBPPts <= 0) {
        real lvl = std::pow(FLH(10.0), (real)0 / FLH(20.0));
        o->sbp = {-180.0, (double)lvl, 180.0, (double)lvl};
        v.NSBPPts = 2;
    } else {
        for(int i = 0; i < env->NSBPPts; ++i) {
            o->sbp.push_back((double)(real)env->sbp[2 * i]);
            o->sbp.push_back((double)std::pow(FLH(10.0), (real)env->sbp[2 * i + 1] / FLH(20.0)));
        }
        v.NSBPPts = env->NSBPPts;
    }
    return 0;
}

void bind(bhc_problem_owner *o)
{
    bhc_problem2d &v = o->v;
    v.ssp_z = o->ssp_z.data();
    v.ssp_c = o->ssp_c.data();
    v.ssp_cz = o->ssp_cz.data();
    v.ssp_n2 = o->ssp_n2.data();
    v.ssp_n2z = o->ssp_n2z.data();
    v.ssp_coef = o->ssp_coef.data();
    v.ssp_rho = o->ssp_rho.data();
    v.ssp_seg_r = o->ssp_segr.empty() ? nullptr : o->ssp_segr.data();
    v.ssp_cmat = o->ssp_cmat.empty() ? nullptr : o->ssp_cmat.data();
    v.ssp_czmat = o->ssp_czmat.empty() ? nullptr : o->ssp_czmat.data();
    v.ati = o->ati.data();
    v.bty = o->bty.data();
    v.trc = o->trc.empty() ? nullptr : o->trc.data();
    v.brc = o->brc.empty() ? nullptr : o->brc.data();
    v.Sz = o->Sz.data();
    v.Rz = o->Rz.data();
    v.Rr = o->Rr.data();
    v.alpha = o->alpha.data();
    v.alpha_cos = o->acos_.data();
    v.alpha_sin = o->asin_.data();
    v.sbp = o->sbp.data();
}

} // namespace

extern "C" const char *bhc_host_last_error(void) { return g_host_error.c_str(); }

// SubTab (src/common_setup.hpp:282-316): x[i] = fma(i, (x1 - x0) / (n - 1), x0)
// in the vector's own precision.
extern "C" int bhc_subtab(double x0, double x1, int32_t n, int32_t precision, void *out)
{
    if(!out || n < 0 || (precision != 64 && precision != 32)) return hfail("bhc_subtab: bad arguments");
    if(precision == 64) {
        double *x = static_cast<double *>(out);
        if(n == 1) x[0] = x0;
        if(n < 2) return 0;
        double d = (x1 - x0) / (double)(n - 1);
        for(int32_t i = 0; i < n; ++i) x[i] = std::fma((double)i, d, x0);
    } else {
        float *x = static_cast<float *>(out);
        float a = (float)x0, b = (float)x1;
        if(n == 1) x[0] = a;
        if(n < 2) return 0;
        float d = (b - a) / (float)(n - 1);
        for(int32_t i = 0; i < n; ++i) x[i] = std::fmaf((float)i, d, a);
    }
    return 0;
}

extern "C" int bhc_preprocess2d(const bhc_env2d *env, bhc_problem_owner **out)
{
    *out = nullptr;
    if(!env) return hfail("null environment");
    if(env->precision != 64 && env->precision != 32) return hfail("precision must be 32 or 64");
    if(env->NSz < 1 || env->NRz < 1 || env->NRr < 1 || env->Nalpha < 1)
        return hfail("sources, receivers and launch angles must be non-empty");
    // Run types of the reference (src/module/runtype.hpp:55-68): 'C'oherent,
    // 'S'emicoherent, 'I'ncoherent TL, 'R'ay, 'E'igenray, 'A'/'a' arrivals,
    // 'V'/'v' arrivals and eigenrays together.
    if(!std::strchr("CSIREAaVv", env->RunType[0]) || env->RunType[0] == '\0')
        return hfail("run type '%c' is not supported", env->RunType[0]);
    // Influence types on this path are the geometric ones (hat / Gaussian,
    // Cartesian / ray-centred); Cerveny ('R', 'C') and simple Gaussian ('S')
    // beams are not.
    if(!std::strchr("GgBb^ ", env->RunType[1]) || env->RunType[1] == '\0')
        return hfail("beam type '%c' is not a geometric type (Cerveny / simple Gaussian beams are not on this path)",
                     env->RunType[1]);
    bhc_problem_owner *o = new bhc_problem_owner();
    int rc = env->precision == 64 ? preprocess<double>(env, o) : preprocess<float>(env, o);
    if(rc) {
        delete o;
        return -1;
    }
    bind(o);
    *out = o;
    return 0;
}

extern "C" const bhc_problem2d *bhc_problem_view(const bhc_problem_owner *p) { return p ? &p->v : nullptr; }

extern "C" void bhc_problem_free(bhc_problem_owner *p) { delete p; }

// ScalePressure (src/influence.hpp:1556-1627) for geometric 2D beams, applied to
// every source of the field as PostProcessTL (src/mode/tl.cpp:106-180) does.
extern "C" int bhc_scale_pressure2d(const bhc_problem2d *p, float *field)
{
    if(!p || !field) return hfail("null argument");
    bool coherent = p->RunType[0] == 'C';
    bool line = p->RunType[3] == 'X';
    size_t plane = (size_t)p->NRz_per_range * p->NRr;
    for(int isz = 0; isz < p->NSz; ++isz) {
        float *u = field + 2 * (size_t)isz * plane;
        if(!coherent)
            for(size_t k = 0; k < plane; ++k) {
                u[2 * k] = std::sqrt(u[2 * k]);
                u[2 * k + 1] = 0.0f;
            }
        for(int ir = 0; ir < p->NRr; ++ir) {
            double factor;
            if(p->precision == 64) {
                double cnst = -1.0;
                if(line) factor = -4.0f * std::sqrt(3.14159265f) * cnst;
                else if(std::abs(p->Rr[ir]) < 2.220446049250313e-16) factor = 0.0;
                else factor = cnst / std::sqrt(std::abs(p->Rr[ir]));
            } else {
                float cnst = -1.0f, r = (float)p->Rr[ir];
                if(line) factor = -4.0f * std::sqrt(3.14159265f) * cnst;
                else if(std::abs(r) < 1.1920929e-07f) factor = 0.0f;
                else factor = cnst / std::sqrt(std::abs(r));
            }
            float f = (float)factor;
            for(int irz = 0; irz < p->NRz_per_range; ++irz) {
                size_t a = (size_t)irz * p->NRr + ir;
                u[2 * a] *= f;
                u[2 * a + 1] *= f;
            }
        }
    }
    return 0;
}

// Internal to the library (used by the 3D preprocessing, not part of the C
// ABI): crci in dB/wavelength with no volume term, for geoacoustics given per
// boundary cell (src/module/boundary.hpp:527-537), in the run's real type.
void bhc_crci_w(int precision, double freq0, double c, double alpha, double out[2])
{
    bhc_env2d e;
    memset(&e, 0, sizeof(e));
    if(precision == 64) {
        Pre<double> pre{&e, freq0};
        cx<double> r = pre.crci(1.0e20, c, alpha, "W ");
        out[0] = r.real();
        out[1] = r.imag();
    } else {
        Pre<float> pre{&e, (float)freq0};
        cx<float> r = pre.crci(1.0e20f, (float)c, (float)alpha, "W ");
        out[0] = (double)r.real();
        out[1] = (double)r.imag();
    }
}
