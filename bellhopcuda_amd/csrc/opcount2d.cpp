// opcount2d.cpp -- algorithmic floating-point operation counts of the 2D hot
// path, for the roofline numerator.
//
// The tracer core (core2d.h) is instantiated on the host with a counting
// scalar type: every +, -, *, / and sqrt / exp / sin / cos / atan2 / hypot on a
// real counts as one operation (complex operations expand into their real
// operations, as in the core). A sample of rays of a problem is traced with it
// and the counts are split into
//   f_I: operations per ray for RayInit (ray initialisation),
//   f_S: operations per step for RayUpdate (Step, boundary lookup, reflection)
//        plus the per-step influence prelude,
//   f_C: operations per accepted contribution (InfluenceGeoCore + the field term),
// so that a launch performs F = f_I R + f_S S + f_C C with R, S, C counted by the
// kernels. Receiver bracketing and beam-window rejections are not counted.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bhc_hip.h"

namespace opc {
thread_local long long n = 0;
}

struct CR {
    double v;
    CR() = default;
    constexpr CR(double x) : v(x) {}
    explicit operator double() const { return v; }
    explicit operator float() const { return (float)v; }
    explicit operator int() const { return (int)v; }
};
inline CR operator+(CR a, CR b) { ++opc::n; return CR(a.v + b.v); }
inline CR operator-(CR a, CR b) { ++opc::n; return CR(a.v - b.v); }
inline CR operator*(CR a, CR b) { ++opc::n; return CR(a.v * b.v); }
inline CR operator/(CR a, CR b) { ++opc::n; return CR(a.v / b.v); }
inline CR operator-(CR a) { return CR(-a.v); }
inline CR &operator+=(CR &a, CR b) { a = a + b; return a; }
inline CR &operator-=(CR &a, CR b) { a = a - b; return a; }
inline CR &operator*=(CR &a, CR b) { a = a * b; return a; }
inline CR &operator/=(CR &a, CR b) { a = a / b; return a; }
inline bool operator<(CR a, CR b) { return a.v < b.v; }
inline bool operator>(CR a, CR b) { return a.v > b.v; }
inline bool operator<=(CR a, CR b) { return a.v <= b.v; }
inline bool operator>=(CR a, CR b) { return a.v >= b.v; }
inline bool operator==(CR a, CR b) { return a.v == b.v; }
inline bool operator!=(CR a, CR b) { return a.v != b.v; }
#define CR_MIXED(T) \
    inline CR operator+(CR a, T b) { return a + CR((double)b); } \
    inline CR operator+(T a, CR b) { return CR((double)a) + b; } \
    inline CR operator-(CR a, T b) { return a - CR((double)b); } \
    inline CR operator-(T a, CR b) { return CR((double)a) - b; } \
    inline CR operator*(CR a, T b) { return a * CR((double)b); } \
    inline CR operator*(T a, CR b) { return CR((double)a) * b; } \
    inline CR operator/(CR a, T b) { return a / CR((double)b); } \
    inline CR operator/(T a, CR b) { return CR((double)a) / b; } \
    inline bool operator<(CR a, T b) { return a.v < (double)b; } \
    inline bool operator<(T a, CR b) { return (double)a < b.v; } \
    inline bool operator>(CR a, T b) { return a.v > (double)b; } \
    inline bool operator>(T a, CR b) { return (double)a > b.v; } \
    inline bool operator<=(CR a, T b) { return a.v <= (double)b; } \
    inline bool operator<=(T a, CR b) { return (double)a <= b.v; } \
    inline bool operator>=(CR a, T b) { return a.v >= (double)b; } \
    inline bool operator>=(T a, CR b) { return (double)a >= b.v; } \
    inline bool operator==(CR a, T b) { return a.v == (double)b; } \
    inline bool operator!=(CR a, T b) { return a.v != (double)b; }
CR_MIXED(double)
CR_MIXED(float)
CR_MIXED(int)

namespace std {
inline CR sqrt(CR a) { ++opc::n; return CR(std::sqrt(a.v)); }
inline CR exp(CR a) { ++opc::n; return CR(std::exp(a.v)); }
inline CR sin(CR a) { ++opc::n; return CR(std::sin(a.v)); }
inline CR cos(CR a) { ++opc::n; return CR(std::cos(a.v)); }
inline CR atan2(CR a, CR b) { ++opc::n; return CR(std::atan2(a.v, b.v)); }
inline CR atan(CR a) { ++opc::n; return CR(std::atan(a.v)); }
inline CR hypot(CR a, CR b) { ++opc::n; return CR(std::hypot(a.v, b.v)); }
inline CR fabs(CR a) { return CR(std::fabs(a.v)); }
inline CR nextafter(CR a, CR b) { return CR(std::nextafter(a.v, b.v)); }
inline CR round(CR a) { return CR(std::round(a.v)); }
inline CR floor(CR a) { return CR(std::floor(a.v)); }
inline CR ceil(CR a) { return CR(std::ceil(a.v)); }
inline CR copysign(CR a, CR b) { return CR(std::copysign(a.v, b.v)); }
inline CR acos(CR a) { ++opc::n; return CR(std::acos(a.v)); }
inline bool isfinite(CR a) { return std::isfinite(a.v); }
} // namespace std

#include "core2d.h"

namespace bhc2d {
template<> struct RC<CR> {
    static constexpr double EPS = RC<double>::EPS, MAXV = RC<double>::MAXV, MINPOS = RC<double>::MINPOS;
    static constexpr double PI = RC<double>::PI, INF_STEP = RC<double>::INF_STEP;
};
template<> __host__ __device__ constexpr CR rlit<CR>(double d, float) { return CR(d); }
} // namespace bhc2d

#include "pack2d.h"
#include "core3d.h"

namespace bhc3d {
template<> struct TD3<CR> {
    static constexpr double THRESH = TD3<double>::THRESH;
};
} // namespace bhc3d

#include "pack3d.h"

using namespace bhc2d;

namespace {
struct CountSink {
    long long contrib_ops = 0, ncontrib = 0, m = 0;
    template<typename P> void add_field(const P &, size_t, float, float) {}
    template<typename R> void add_arrival(size_t, R, R, cpx<R>, R, R, int, int, int, int, int) {}
    void add_eigen(int, int, int, int, int) {}
    void mark() { m = opc::n; }
    void count()
    {
        ++ncontrib;
        contrib_ops += opc::n - m;
    }
};
} // namespace

// Counts over rays [stride*k] of the problem (k < nsample). out: f_I, f_S, f_C,
// and the sample's rays, steps, contributions.
extern "C" int bhc_opcount2d(const bhc_problem2d *prob, int nsample, double *out)
{
    std::vector<CR> blob;
    Prob2D<CR> P;
    std::string err;
    if(pack<CR>(prob, blob, P, err)) return -1;
    P.base = blob.data();
    P.Rz = prob->Rz;
    P.Sz = prob->Sz;
    const CR *B = blob.data(), *G = blob.data();
    int njobs = (prob->iSingle_alpha >= 1 ? 1 : prob->Nalpha) * prob->NSz;
    if(nsample < 1) nsample = 1;
    if(nsample > njobs) nsample = njobs;
    long long ops_init = 0, ops_step = 0, ops_contrib = 0, rays = 0, steps = 0, contrib = 0;
    CountSink sink;
    Err e = {0, 0, 0, 0};
    for(int k = 0; k < nsample; ++k) {
        int job = (int)((long long)k * njobs / nsample);
        int isz, ia;
        JobIndices(job, P, isz, ia);
        RayPt<CR> p0, p1, p2;
        Trace<CR> S;
        Infl<CR> I;
        opc::n = 0;
        bool ok = RayInit(isz, ia, p0, S, I, P, B, G, e);
        ops_init += opc::n;
        ++rays;
        if(!ok) continue;
        int is = 0;
        for(;;) {
            opc::n = 0;
            bool two = RayUpdate(p0, p1, p2, S, P, B, G, e);
            SegDesc<CR> d;
            prelude(d, p0, p1, I, is, P, G, G + P.o_rzs, isz, ia);
            ops_step += opc::n;
            ++is;
            for(int pass = 0; pass < (two ? 2 : 1); ++pass) {
                if(pass == 1) {
                    opc::n = 0;
                    prelude(d, p1, p2, I, is, P, G, G + P.o_rzs, isz, ia);
                    ops_step += opc::n;
                    ++is;
                }
                seg_items<CR, RT_TL>(d, P, G, prob->Rz, sink);
            }
            p0 = two ? p2 : p1;
            if(RayTerminate(p0, is, S, 100000, P, e)) break;
        }
        steps += is;
    }
    out[0] = rays ? (double)ops_init / rays : 0.0;
    out[1] = steps ? (double)ops_step / steps : 0.0;
    contrib = sink.ncontrib;
    ops_contrib = sink.contrib_ops;
    out[2] = contrib ? (double)ops_contrib / contrib : 0.0;
    out[3] = (double)rays;
    out[4] = (double)steps;
    out[5] = (double)contrib;
    return 0;
}

namespace {
struct CountSink3 {
    long long contrib_ops = 0, ncontrib = 0, m = 0;
    void add3(size_t, float, float) {}
    void mark() { m = opc::n; }
    void count()
    {
        ++ncontrib;
        contrib_ops += opc::n - m;
    }
};
} // namespace

// The 3D path: f_I per ray (RayInit3 + Init_Influence), f_S per step
// (RayUpdate3: Step3, boundary lookup, reflection, plus the segment prelude),
// f_C per accepted contribution (InfluenceGeoCore + the TL term), over rays
// [njobs * k / nsample]. Row / depth-band enumeration and rejected receivers
// are not counted.
extern "C" int bhc_opcount3d(const bhc_problem3d *prob, int nsample, double *out)
{
    using namespace bhc3d;
    std::vector<CR> blob;
    Prob3D<CR> P;
    std::string err;
    if(pack3<CR>(prob, blob, P, err)) return -1;
    P.base = blob.data();
    P.s.base = blob.data();
    P.Rz = prob->Rz;
    P.Sz = prob->Sz;
    long long njobs = (long long)(prob->iSingle_alpha >= 1 ? 1 : prob->Nalpha) * (prob->iSingle_beta >= 1 ? 1 : prob->Nbeta)
        * prob->NSx * prob->NSy * prob->NSz;
    if(nsample < 1) nsample = 1;
    if(nsample > njobs) nsample = (int)njobs;
    long long ops_init = 0, ops_step = 0, rays = 0, steps = 0;
    CountSink3 sink;
    Err e = {0, 0, 0, 0};
    auto scatter = [&](const Seg3D<CR> &d) {
        if(d.nitems <= 0) return;
        int nr = (d.ir1 - d.ir0) * d.dir + 1;
        for(int j = 0; j < nr; ++j)
            for(int it = 0; it < P.Ntheta; ++it) {
                long long save = opc::n;
                int ir = d.ir0 + j * d.dir, i0 = 0, i1 = P.NRz - 1;
                CR xr, yr, zmin, zmax, wlo, whi;
                bool ok = row3(d, ir, it, P, P.base + P.o_Rr, P.base + P.o_trcv, xr, yr, zmin, zmax) && band3(d, xr, yr, zmin, zmax, P, wlo, whi);
                opc::n = save;
                if(!ok) continue;
                if(P.rz_sorted) depth_window(P.Rz, P.NRz, P.rz_rdz, wlo.v, whi.v, i0, i1);
                for(int iz = i0; iz <= i1; ++iz) {
                    save = opc::n;
                    item3<CR, CountSink3, false>(d, ir, it, iz, xr, yr, zmin, zmax, P, P.Rz, sink);
                    opc::n = save;
                }
            }
    };
    for(int k = 0; k < nsample; ++k) {
        int job = (int)((long long)k * njobs / nsample);
        int ia, ib, isx, isy, isz;
        JobIndices3(job, P, ia, ib, isy, isx, isz);
        Ray3<CR> p0, p1, p2;
        Trace3<CR> S;
        Infl3<CR> I;
        opc::n = 0;
        bool ok = RayInit3(ia, ib, isx, isy, isz, p0, S, I, P, e);
        ops_init += opc::n;
        ++rays;
        if(!ok) continue;
        int is = 0;
        for(;;) {
            opc::n = 0;
            bool two = RayUpdate3(p0, p1, p2, S, P, e);
            Seg3D<CR> d;
            prelude3(d, p0, p1, I, is, P, isz, isx, isy);
            ops_step += opc::n;
            ++is;
            scatter(d);
            if(two) {
                opc::n = 0;
                prelude3(d, p1, p2, I, is, P, isz, isx, isy);
                ops_step += opc::n;
                ++is;
                scatter(d);
            }
            p0 = two ? p2 : p1;
            if(RayTerminate3(p0, is, S, P, e)) break;
        }
        steps += is;
    }
    out[0] = rays ? (double)ops_init / rays : 0.0;
    out[1] = steps ? (double)ops_step / steps : 0.0;
    out[2] = sink.ncontrib ? (double)sink.contrib_ops / sink.ncontrib : 0.0;
    out[3] = (double)rays;
    out[4] = (double)steps;
    out[5] = (double)sink.ncontrib;
    return 0;
}
