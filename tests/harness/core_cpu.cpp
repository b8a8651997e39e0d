// Test harness only: runs the product's per-ray core (core2d.h + pack2d.h) on
// the CPU, one ray after another, with the wavefront scatter replaced by a
// plain loop over each segment's receivers. It exercises exactly the
// functions the HIP kernel calls, so a disagreement with the oracle found here
// is a bug in the shared core, not in the kernel's launch plumbing.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bhc_hip.h"
#include "../../bellhopcuda_amd/csrc/core2d.h"
#include "../../bellhopcuda_amd/csrc/pack2d.h"

using namespace bhc2d;

namespace {
struct HostSink {
    float *field;
    long long nc = 0;
    template<typename P> void add_field(const P &, size_t idx, float re, float im)
    {
        field[2 * idx] += re;
        field[2 * idx + 1] += im;
    }
    template<typename real>
    void add_arrival(size_t, real, real, cpx<real>, real, real, int, int, int, int, int)
    {
    }
    void add_eigen(int, int, int, int, int) {}
    void count() { ++nc; }
    void mark() {}
};

template<typename real, bool FAST> int run(const bhc_problem2d *prob, bhc_field_out *out)
{
    std::vector<real> blob;
    Prob2D<real> P;
    std::string err;
    if(pack<real>(prob, blob, P, err)) {
        fprintf(stderr, "%s\n", err.c_str());
        return -1;
    }
    P.base = blob.data();
    P.Rz = prob->Rz;
    P.Sz = prob->Sz;
    const real *B = blob.data(), *G = blob.data();
    HostSink sink{out->field};
    Err e = {0, 0, 0, 0};
    int njobs = (prob->iSingle_alpha >= 1 ? 1 : prob->Nalpha) * prob->NSz;
    long long steps = 0;
    for(int job = 0; job < njobs; ++job) {
        int isz, ia;
        JobIndices(job, P, isz, ia);
        RayPt<real> p0, p1, p2;
        Trace<real> S;
        Infl<real> I;
        if(!RayInit(isz, ia, p0, S, I, P, B, G, e)) {
            if(out->ray_nsteps) {
                out->ray_nsteps[job] = 0;
                out->ray_ntop[job] = out->ray_nbot[job] = 0;
                out->ray_xend[2 * job] = out->ray_xend[2 * job + 1] = 0;
            }
            continue;
        }
        int is = 0;
        for(;;) {
            if(e.err) break;
            bool two = RayUpdate(p0, p1, p2, S, P, B, G, e);
            SegDesc<real> d;
            prelude(d, p0, p1, I, is, P, G, G + P.o_rzs, isz, ia);
            seg_items<real, RT_TL, HostSink, FAST>(d, P, G, prob->Rz, sink);
            ++is;
            if(two) {
                prelude(d, p1, p2, I, is, P, G, G + P.o_rzs, isz, ia);
                seg_items<real, RT_TL, HostSink, FAST>(d, P, G, prob->Rz, sink);
                ++is;
            }
            p0 = two ? p2 : p1;
            if(RayTerminate(p0, is, S, 100000, P, e)) break;
        }
        steps += is;
        if(out->ray_nsteps) {
            out->ray_nsteps[job] = is;
            out->ray_ntop[job] = p0.NumTopBnc;
            out->ray_nbot[job] = p0.NumBotBnc;
            out->ray_xend[2 * job] = (double)p0.x.x;
            out->ray_xend[2 * job + 1] = (double)p0.x.y;
        }
    }
    out->total_steps = steps;
    out->total_contrib = sink.nc;
    out->err = e.err;
    out->warn = e.warn;
    out->err_count = e.nerr;
    out->warn_count = e.nwarn;
    return 0;
}
} // namespace

extern "C" int core_field_run2d(const bhc_problem2d *prob, bhc_field_out *out)
{
    return prob->precision == 64 ? run<double, false>(prob, out) : run<float, false>(prob, out);
}

// The same with the FP32-accuracy TL phasor the GPU scatter uses by default.
extern "C" int core_field_run2d_fast(const bhc_problem2d *prob, bhc_field_out *out)
{
    return prob->precision == 64 ? run<double, true>(prob, out) : run<float, true>(prob, out);
}

namespace {
template<typename real> int run_ray(const bhc_problem2d *prob, int isz, int ia, int maxp, bhc_ray_pt2d *pts, int *nsteps)
{
    std::vector<real> blob;
    Prob2D<real> P;
    std::string err;
    if(pack<real>(prob, blob, P, err)) return -1;
    P.base = blob.data();
    P.Rz = prob->Rz;
    P.Sz = prob->Sz;
    const real *B = blob.data(), *G = blob.data();
    Err e = {0, 0, 0, 0};
    RayPt<real> p0, p1, p2;
    Trace<real> S;
    Infl<real> I;
    auto st = [&](int k, const RayPt<real> &r) {
        bhc_ray_pt2d &d = pts[k];
        d.x[0] = r.x.x; d.x[1] = r.x.y; d.t[0] = r.t.x; d.t[1] = r.t.y; d.p[0] = r.p.x; d.p[1] = r.p.y;
        d.q[0] = r.q.x; d.q[1] = r.q.y; d.c = r.c; d.Amp = r.Amp; d.Phase = r.Phase; d.tau[0] = r.tau.re;
        d.tau[1] = r.tau.im; d.NumTopBnc = r.NumTopBnc; d.NumBotBnc = r.NumBotBnc;
    };
    bool ok = RayInit(isz, ia, p0, S, I, P, B, G, e);
    st(0, p0);
    if(!ok) { *nsteps = 1; return 0; }
    int is = 0, N = -1;
    for(;;) {
        bool two = RayUpdate(p0, p1, p2, S, P, B, G, e);
        st(is + 1, p1);
        if(two) st(is + 2, p2);
        is += two ? 2 : 1;
        p0 = two ? p2 : p1;
        int t = RayTerminate(p0, is, S, maxp, P, e);
        if(t == 1) { N = is + 1; break; }
        if(t == 2) { N = is; break; }
    }
    *nsteps = N;
    return 0;
}
} // namespace

extern "C" int core_ray_run2d(const bhc_problem2d *prob, int isz, int ia, int maxp, bhc_ray_pt2d *pts, int *nsteps)
{
    return prob->precision == 64 ? run_ray<double>(prob, isz, ia, maxp, pts, nsteps)
                                  : run_ray<float>(prob, isz, ia, maxp, pts, nsteps);
}

// ---------------------------------------------------------------------------
// 3D: the product's 3D core (core3d.h + pack3d.h) run serially on the CPU,
// with each segment's rows and depth windows enumerated as the scatter kernel
// enumerates them.
// ---------------------------------------------------------------------------
#include "../../bellhopcuda_amd/csrc/core3d.h"
#include "../../bellhopcuda_amd/csrc/pack3d.h"

namespace {
struct HostSink3 {
    float *field;
    long long nc = 0;
    void add3(size_t idx, float re, float im)
    {
        field[2 * idx] += re;
        field[2 * idx + 1] += im;
    }
    void count() { ++nc; }
    void mark() {}
};

template<typename real> int run3(const bhc_problem3d *prob, bhc_field_out *out)
{
    using namespace bhc3d;
    std::vector<real> blob;
    Prob3D<real> P;
    std::string err;
    if(pack3<real>(prob, blob, P, err)) {
        fprintf(stderr, "%s\n", err.c_str());
        return -1;
    }
    P.base = blob.data();
    P.s.base = blob.data();
    P.Rz = prob->Rz;
    P.Sz = prob->Sz;
    HostSink3 sink{out->field};
    Err e = {0, 0, 0, 0};
    int njobs = (prob->iSingle_alpha >= 1 ? 1 : prob->Nalpha) * (prob->iSingle_beta >= 1 ? 1 : prob->Nbeta)
        * prob->NSx * prob->NSy * prob->NSz;
    long long steps = 0;
    auto scatter = [&](const Seg3D<real> &d) {
        if(d.nitems <= 0) return;
        int nr = (d.ir1 - d.ir0) * d.dir + 1;
        const real *Rr = P.base + P.o_Rr, *tr = P.base + P.o_trcv;
        real rlo = rmin(Rr[d.ir0], Rr[d.ir1]), rhi = rmax(Rr[d.ir0], Rr[d.ir1]);
        real zs = P.rz_absmax + std::fabs(d.xa.z);
        for(int j = 0; j < nr; ++j)
            for(int it = 0; it < P.Ntheta; ++it) {
                if(P.Ntheta <= 64 && !bearing3(d, tr[2 * it], tr[2 * it + 1], rlo, rhi, zs, P)) continue;
                int ir = d.ir0 + j * d.dir, i0 = 0, i1 = -1;
                real xr, yr, zmin, zmax;
                real wlo, whi;
                if(!row3(d, ir, it, P, P.base + P.o_Rr, P.base + P.o_trcv, xr, yr, zmin, zmax)) continue;
                if(!band3(d, xr, yr, zmin, zmax, P, wlo, whi)) continue;
                if(P.rz_sorted && wlo == wlo && whi == whi) depth_window(P.Rz, P.NRz, P.rz_rdz, wlo, whi, i0, i1);
                else {
                    i0 = 0;
                    i1 = P.NRz - 1;
                }
                for(int iz = i0; iz <= i1; ++iz) item3<real, HostSink3, false>(d, ir, it, iz, xr, yr, zmin, zmax, P, P.Rz, sink);
            }
    };
    // ray-centred: every (depth, bearing) row's bracketed ranges, in the
    // reference's order
    auto scatter_rc = [&](const Seg3R<real> &d) {
        if(d.nitems <= 0) return;
        const real *Rr = P.base + P.o_Rr, *tr = P.base + P.o_trcv;
        for(int iz = 0; iz < P.NRz; ++iz)
            for(int it = 0; it < P.Ntheta; ++it) {
                real mA, nA, rA, mB, nB, rB;
                int first, step;
                int cnt = row3rc(d, iz, it, P, Rr, tr, P.Rz, mA, nA, rA, mB, nB, rB, first, step);
                for(int k = 0; k < cnt; ++k)
                    item3rc<real, HostSink3, false, RT_TL>(d, iz, it, first + k * step, mA, nA, rA, mB, nB, rB, P, Rr,
                                                           sink);
            }
    };
    for(int job = 0; job < njobs; ++job) {
        int ia, ib, isx, isy, isz;
        JobIndices3(job, P, ia, ib, isy, isx, isz);
        Ray3<real> p0, p1, p2;
        Trace3<real> S;
        Infl3<real> I;
        if(!RayInit3(ia, ib, isx, isy, isz, p0, S, I, P, e)) {
            if(out->ray_nsteps) {
                out->ray_nsteps[job] = out->ray_ntop[job] = out->ray_nbot[job] = 0;
                for(int k = 0; k < 3; ++k) out->ray_xend[3 * job + k] = 0;
            }
            continue;
        }
        int is = 0;
        for(;;) {
            if(e.err) break;
            bool two = RayUpdate3(p0, p1, p2, S, P, e);
            if(P.isRayCen) {
                Seg3R<real> d;
                prelude3rc(d, p0, p1, I, is, P, isz, isx, isy);
                scatter_rc(d);
                ++is;
                if(two) {
                    prelude3rc(d, p1, p2, I, is, P, isz, isx, isy);
                    scatter_rc(d);
                    ++is;
                }
            } else {
                Seg3D<real> d;
                prelude3(d, p0, p1, I, is, P, isz, isx, isy);
                scatter(d);
                ++is;
                if(two) {
                    prelude3(d, p1, p2, I, is, P, isz, isx, isy);
                    scatter(d);
                    ++is;
                }
            }
            p0 = two ? p2 : p1;
            if(RayTerminate3(p0, is, S, P, e)) break;
        }
        steps += is;
        if(out->ray_nsteps) {
            out->ray_nsteps[job] = is;
            out->ray_ntop[job] = p0.NumTopBnc;
            out->ray_nbot[job] = p0.NumBotBnc;
            out->ray_xend[3 * job] = (double)p0.x.x;
            out->ray_xend[3 * job + 1] = (double)p0.x.y;
            out->ray_xend[3 * job + 2] = (double)p0.x.z;
        }
    }
    out->total_steps = steps;
    out->total_contrib = sink.nc;
    out->err = e.err;
    out->warn = e.warn;
    out->err_count = e.nerr;
    out->warn_count = e.nwarn;
    return 0;
}
} // namespace

extern "C" int core_field_run3d(const bhc_problem3d *prob, bhc_field_out *out)
{
    return prob->precision == 64 ? run3<double>(prob, out) : run3<float>(prob, out);
}

namespace {
template<typename real> int ray3(const bhc_problem3d *prob, bhc_ray_pt3d *pts, int32_t *nsteps, int M)
{
    using namespace bhc3d;
    std::vector<real> blob;
    Prob3D<real> P;
    std::string err;
    if(pack3<real>(prob, blob, P, err)) return -1;
    P.base = blob.data();
    P.s.base = blob.data();
    P.Rz = prob->Rz;
    P.Sz = prob->Sz;
    Err e = {0, 0, 0, 0};
    int njobs = (prob->iSingle_alpha >= 1 ? 1 : prob->Nalpha) * (prob->iSingle_beta >= 1 ? 1 : prob->Nbeta)
        * prob->NSx * prob->NSy * prob->NSz;
    for(int job = 0; job < njobs; ++job) {
        int ia, ib, isx, isy, isz;
        JobIndices3(job, P, ia, ib, isy, isx, isz);
        bhc_ray_pt3d *out = pts + (size_t)job * M;
        Ray3<real> p0, p1, p2;
        Trace3<real> S;
        Infl3<real> I;
        int N = -1;
        bool ok = RayInit3(ia, ib, isx, isy, isz, p0, S, I, P, e);
        store_pt3(out[0], p0);
        if(!ok) {
            nsteps[job] = 1;
            continue;
        }
        int is = 0;
        for(;;) {
            if(e.err) break;
            bool two = RayUpdate3(p0, p1, p2, S, P, e);
            store_pt3(out[is + 1], p1);
            if(two) store_pt3(out[is + 2], p2);
            is += two ? 2 : 1;
            p0 = two ? p2 : p1;
            if(RayTerminate3(p0, is, S, P, e, M, N)) break;
        }
        nsteps[job] = N < 0 ? is + 1 : N;
    }
    return 0;
}
} // namespace

extern "C" int core_ray_run3d(const bhc_problem3d *prob, bhc_ray_pt3d *pts, int32_t *nsteps, int M)
{
    return prob->precision == 64 ? ray3<double>(prob, pts, nsteps, M) : ray3<float>(prob, pts, nsteps, M);
}

// Brute-force check of the 3D Cartesian scatter's pruning: every step of
// every ray (one in `stride`) is tested against every (range, bearing, depth)
// receiver the reference's row test admits, with no prefilter or band; each
// receiver the beam-window test accepts must have passed the bearing
// prefilter and lie inside the row's band. out: [0] accepted receivers,
// [1] accepted but pruned by the bearing prefilter, [2] accepted but outside
// the band, [3] receivers in the rows' full depth windows, [4] receivers in
// the bands.
namespace {
struct HitSink3 {
    bool hit = false;
    void add3(size_t, float, float) {}
    void count() { hit = true; }
    void mark() {}
};

template<typename real> int band_check3(const bhc_problem3d *prob, int stride, long long *out)
{
    using namespace bhc3d;
    std::vector<real> blob;
    Prob3D<real> P;
    std::string err;
    if(pack3<real>(prob, blob, P, err)) return -1;
    P.base = blob.data();
    P.s.base = blob.data();
    P.Rz = prob->Rz;
    P.Sz = prob->Sz;
    Err e = {0, 0, 0, 0};
    const int njobs = prob->Nalpha * prob->Nbeta * prob->NSx * prob->NSy * prob->NSz;
    const real *Rr = P.base + P.o_Rr, *tr = P.base + P.o_trcv;
    auto check = [&](const Seg3D<real> &d) {
        if(d.nitems <= 0) return;
        const int nr = (d.ir1 - d.ir0) * d.dir + 1;
        real rlo = rmin(Rr[d.ir0], Rr[d.ir1]), rhi = rmax(Rr[d.ir0], Rr[d.ir1]);
        real zs = P.rz_absmax + std::fabs(d.xa.z);
        for(int it = 0; it < P.Ntheta; ++it) {
            const bool bpass = bearing3(d, tr[2 * it], tr[2 * it + 1], rlo, rhi, zs, P);
            for(int j = 0; j < nr; ++j) {
                const int ir = d.ir0 + j * d.dir;
                real xr, yr, zmin, zmax, wlo = 0, whi = -1;
                if(!row3(d, ir, it, P, Rr, tr, xr, yr, zmin, zmax)) continue;
                const bool in_band = band3(d, xr, yr, zmin, zmax, P, wlo, whi);
                for(int iz = 0; iz < P.NRz; ++iz) {
                    const real zr = (real)P.Rz[iz];
                    if(zr < zmin || zr > zmax) continue;
                    ++out[3];
                    const bool inside = in_band && zr >= wlo && zr <= whi;
                    if(inside) ++out[4];
                    HitSink3 s;
                    item3<real, HitSink3, false>(d, ir, it, iz, xr, yr, zmin, zmax, P, P.Rz, s);
                    if(!s.hit) continue;
                    ++out[0];
                    if(!bpass) ++out[1];
                    if(!inside) ++out[2];
                }
            }
        }
    };
    for(int job = 0; job < njobs; job += stride) {
        int ia, ib, isx, isy, isz;
        JobIndices3(job, P, ia, ib, isy, isx, isz);
        Ray3<real> p0, p1, p2;
        Trace3<real> S;
        Infl3<real> I;
        if(!RayInit3(ia, ib, isx, isy, isz, p0, S, I, P, e)) continue;
        int is = 0;
        for(;;) {
            if(e.err) break;
            const bool two = RayUpdate3(p0, p1, p2, S, P, e);
            Seg3D<real> d;
            prelude3(d, p0, p1, I, is, P, isz, isx, isy);
            check(d);
            ++is;
            if(two) {
                prelude3(d, p1, p2, I, is, P, isz, isx, isy);
                check(d);
                ++is;
            }
            p0 = two ? p2 : p1;
            if(RayTerminate3(p0, is, S, P, e)) break;
        }
    }
    return 0;
}
} // namespace

extern "C" int core_band_check3d(const bhc_problem3d *prob, int stride, long long *out)
{
    for(int k = 0; k < 5; ++k) out[k] = 0;
    return prob->precision == 64 ? band_check3<double>(prob, stride, out) : band_check3<float>(prob, stride, out);
}

// The receiver-depth window of the scatter kernels (guessed index + exact
// verification), for its test against a plain search.
extern "C" void core_depth_window(const float *Rz, int n, double rdz, double zlo, double zhi, int *i0, int *i1)
{
    depth_window(Rz, n, rdz, zlo, zhi, *i0, *i1);
}

// ---------------------------------------------------------------------------
// Nx2D: the product's Nx2D core (nx2d.h) ray by ray, with the 2D influence
// descriptors and the 2D receiver loop of every bearing plane.
// ---------------------------------------------------------------------------
#include "../../bellhopcuda_amd/csrc/nx2d.h"

namespace {
template<typename real> int runnx(const bhc_problem3d *prob, bhc_field_out *out)
{
    using namespace bhc3d;
    std::vector<real> blob;
    Prob3D<real> N;
    Prob2D<real> P;
    std::string err;
    if(pack_nx<real>(prob, blob, N, P, err)) {
        fprintf(stderr, "%s\n", err.c_str());
        return -1;
    }
    P.base = N.base = N.s.base = blob.data();
    P.Rz = N.Rz = prob->Rz;
    P.Sz = N.Sz = prob->Sz;
    const real *G = blob.data();
    HostSink sink{out->field};
    Err e = {0, 0, 0, 0};
    int njobs = (prob->iSingle_alpha >= 1 ? 1 : prob->Nalpha) * (prob->iSingle_beta >= 1 ? 1 : prob->Nbeta)
        * prob->NSx * prob->NSy * prob->NSz;
    long long steps = 0;
    for(int job = 0; job < njobs; ++job) {
        int ia, ib, isx, isy, isz;
        const int isf = JobIndicesNx(job, N, ia, ib, isx, isy, isz);
        RayPt<real> p0, p1, p2;
        Trace3<real> S;
        OrgNx<real> org;
        Infl<real> I;
        if(!RayInitNx(ia, ib, isx, isy, isz, p0, S, org, I, P, N, e)) {
            if(out->ray_nsteps) {
                out->ray_nsteps[job] = out->ray_ntop[job] = out->ray_nbot[job] = 0;
                for(int k = 0; k < 3; ++k) out->ray_xend[3 * job + k] = 0;
            }
            continue;
        }
        int is = 0;
        for(;;) {
            if(e.err) break;
            bool two = RayUpdateNx(p0, p1, p2, S, org, N, P, e);
            SegDesc<real> d;
            prelude(d, p0, p1, I, is, P, G, G + P.o_rzs, isf, ia);
            seg_items<real, RT_TL, HostSink, false>(d, P, G, prob->Rz, sink);
            ++is;
            if(two) {
                prelude(d, p1, p2, I, is, P, G, G + P.o_rzs, isf, ia);
                seg_items<real, RT_TL, HostSink, false>(d, P, G, prob->Rz, sink);
                ++is;
            }
            p0 = two ? p2 : p1;
            if(RayTerminateNx(p0, is, S, org, N, e, 100000)) break;
        }
        steps += is;
        if(out->ray_nsteps) {
            const v3<real> xe = RayToOceanX(p0.x, org);
            out->ray_nsteps[job] = is;
            out->ray_ntop[job] = p0.NumTopBnc;
            out->ray_nbot[job] = p0.NumBotBnc;
            out->ray_xend[3 * job] = (double)xe.x;
            out->ray_xend[3 * job + 1] = (double)xe.y;
            out->ray_xend[3 * job + 2] = (double)xe.z;
        }
    }
    out->total_steps = steps;
    out->total_contrib = sink.nc;
    out->err = e.err;
    out->warn = e.warn;
    out->err_count = e.nerr;
    out->warn_count = e.nwarn;
    return 0;
}
} // namespace

extern "C" int core_field_runnx2d(const bhc_problem3d *prob, bhc_field_out *out)
{
    return prob->precision == 64 ? runnx<double>(prob, out) : runnx<float>(prob, out);
}
