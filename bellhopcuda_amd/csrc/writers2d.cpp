// writers2d.cpp -- output files of the 2D hot path in the reference's layouts.
//
// .shd: Fortran direct-access "shade" file, WriteHeader + WriteOutTL
//       (src/mode/tl.cpp:38-101, 211-248) through DirectOFile
//       (src/util/directio.hpp): fixed-length records, record length in
//       bytes = max over the header vectors and NRr complex floats, the
//       length in 4-byte words stored first; reals are 8 bytes (4 in 32-bit
//       mode); the file is padded to a whole last record.
// .arr: arrivals, PostProcessArrivals + WriteOutArrivals (src/mode/arr.cpp:
//       24-73, 88-238): list-directed ASCII ('A'/'V') or Fortran sequential
//       unformatted binary ('a'/'v', 4-byte record markers on both sides).
// .ray: ray paths, OpenRAYFile + WriteRay (src/mode/ray.hpp:168-244), ASCII.
//
// The list-directed ASCII items follow LDOFile's FORTRAN_OUTPUT rules
// (src/util/ldio.hpp:409-584): integers right-aligned in 12 columns, floats in
// 15 and doubles in 24, fixed or scientific by magnitude; the stream's format
// flags carry over from one item to the next, and the byte layout depends on
// that, so the writer drives a std::ostream through the same flag changes.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "../../include/bhc_hip.h"

namespace {

struct Direct {
    FILE *f = nullptr;
    size_t recl = 0, high = 0, highBytes = 0, cur = 0, curBytes = 0;
    bool ok = true;
    void rec(size_t r)
    {
        if(r >= high) {
            high = r;
            highBytes = 0;
        }
        cur = r;
        curBytes = 0;
        if(fseek(f, (long)(r * recl), SEEK_SET) != 0) ok = false;
    }
    void put(const void *d, size_t n)
    {
        if(curBytes + n > recl) ok = false;
        curBytes += n;
        if(cur == high) highBytes = std::max(highBytes, curBytes);
        if(fwrite(d, 1, n, f) != n) ok = false;
    }
    void str(const char *s, size_t n)
    {
        std::string t(s ? s : "");
        t.resize(n, ' ');
        put(t.data(), n);
    }
    template<typename T> void v(T x) { put(&x, sizeof(T)); }
    void close()
    {
        if(highBytes < recl) {
            fseek(f, (long)((high + 1) * recl - 1), SEEK_SET);
            fputc(0, f);
        }
        fclose(f);
    }
};

template<typename R> int write_shd(const char *path, const bhc_problem2d *p, const char *title, const float *field)
{
    const int Nfreq = 1, Ntheta = 1, NSx = 1, NSy = 1;
    int32_t LRecl = 84;
    LRecl = std::max<int32_t>(LRecl, 2 * Nfreq * (int32_t)sizeof(R));
    LRecl = std::max<int32_t>(LRecl, Ntheta * (int32_t)sizeof(R));
    LRecl = std::max<int32_t>(LRecl, NSx * (int32_t)sizeof(R));
    LRecl = std::max<int32_t>(LRecl, NSy * (int32_t)sizeof(R));
    LRecl = std::max<int32_t>(LRecl, p->NSz * (int32_t)sizeof(float));
    LRecl = std::max<int32_t>(LRecl, p->NRz * (int32_t)sizeof(float));
    LRecl = std::max<int32_t>(LRecl, p->NRr * (int32_t)(2 * sizeof(float)));
    Direct d;
    d.f = fopen(path, "wb");
    if(!d.f) return -1;
    d.recl = (size_t)LRecl;
    int32_t words = LRecl / 4;
    d.rec(0);
    d.v(words);
    d.str(title, 80);
    d.rec(1);
    d.str(p->RunType[4] == 'I' ? "irregular " : "rectilin  ", 10);
    d.rec(2);
    d.v((int32_t)Nfreq);
    d.v((int32_t)Ntheta);
    d.v((int32_t)NSx);
    d.v((int32_t)NSy);
    d.v((int32_t)p->NSz);
    d.v((int32_t)p->NRz);
    d.v((int32_t)p->NRr);
    d.v((R)p->freq0);
    d.v((R)0);
    d.rec(3);
    d.v((R)p->freq0);
    d.rec(4);
    d.v((R)0);
    d.rec(5);
    d.v((R)0);
    d.rec(6);
    d.v((R)0);
    d.rec(7);
    d.put(p->Sz, sizeof(float) * p->NSz);
    d.rec(8);
    d.put(p->Rz, sizeof(float) * p->NRz);
    d.rec(9);
    for(int i = 0; i < p->NRr; ++i) d.v((R)p->Rr[i]);
    for(int isz = 0; isz < p->NSz; ++isz)
        for(int irz = 0; irz < p->NRz_per_range; ++irz) {
            d.rec(10 + (size_t)isz * p->NRz_per_range + irz);
            d.put(field + 2 * (((size_t)isz * p->NRz_per_range + irz) * p->NRr), 2 * sizeof(float) * p->NRr);
        }
    bool ok = d.ok;
    d.close();
    return ok ? 0 : -1;
}

class LDWriter {
public:
    bool open(const char *path)
    {
        os.open(path);
        os << std::setfill(' ');
        return os.good();
    }
    bool close()
    {
        bool ok = os.good();
        os.close();
        return ok;
    }
    void ch(char c) { os << c; }
    void str(const char *s) { os << "'" << (s ? s : "") << "'"; }
    void i32(int32_t v) { os << std::setiosflags(std::ios_base::right) << std::setw(12) << v; }
    void f32(float v) { item((double)v, 15, false); }
    void f64(double v) { item(v, 24, true); }

private:
    std::ofstream os;

    // One real item: two blanks, the sign column, then either a scientific
    // field (|r| < 0.1 or >= 1e6) or a general field with trailing blanks.
    void item(double r, int width, bool exp3)
    {
        os << "  ";
        if(!std::isfinite(r)) {
            os << std::setw(width) << std::left << r;
            return;
        }
        const bool sci = r != 0.0 && (std::fabs(r) < 0.1 || std::fabs(r) >= 1.0e6);
        int w = width;
        if(r < 0.0) {
            os << "-";
            r = -r;
        } else if(sci || r >= 1.0 || r == 0.0) {
            os << " ";
        }
        --w;
        if(sci) --w;
        os << std::setprecision(exp3 ? (w - 6) : (w - 5));
        if(sci) {
            os << std::setiosflags(std::ios_base::uppercase | std::ios_base::scientific)
               << std::setw(exp3 ? (w + 1) : w) << std::left << r;
        } else {
            os << std::setiosflags(std::ios_base::showpoint) << std::defaultfloat << std::setw(w - 5) << r;
            os << (exp3 ? "     " : "    ");
        }
    }
};

// Fortran sequential unformatted output: each record is framed by its byte
// length as a 4-byte integer before and after.
class UnfWriter {
public:
    bool open(const char *path)
    {
        f = fopen(path, "wb");
        return f != nullptr;
    }
    void rec()
    {
        flush();
        open_rec = true;
    }
    template<typename T> void v(T x)
    {
        const unsigned char *p = reinterpret_cast<const unsigned char *>(&x);
        buf.insert(buf.end(), p, p + sizeof(T));
    }
    void bytes(const void *d, size_t n)
    {
        const unsigned char *p = static_cast<const unsigned char *>(d);
        buf.insert(buf.end(), p, p + n);
    }
    bool close()
    {
        flush();
        bool ok = ok_;
        if(f) fclose(f);
        f = nullptr;
        return ok;
    }

private:
    FILE *f = nullptr;
    std::vector<unsigned char> buf;
    bool open_rec = false, ok_ = true;
    void flush()
    {
        if(!open_rec) return;
        int32_t n = (int32_t)buf.size();
        if(fwrite(&n, 4, 1, f) != 1 || (n && fwrite(buf.data(), 1, buf.size(), f) != buf.size())
           || fwrite(&n, 4, 1, f) != 1)
            ok_ = false;
        buf.clear();
        open_rec = false;
    }
};

size_t rcv_index(const bhc_problem2d *p, int isz, int iz, int ir)
{
    return ((size_t)isz * p->NRz_per_range + iz) * p->NRr + ir;
}

template<typename R> int write_arr(const char *path, const bhc_problem2d *p, const bhc_arrival *arr,
                                   const int32_t *narr, int32_t max_narr, const int32_t *maxn, bool binary)
{
    const R RadDeg = (R)180 / (R)M_PI; // RadDeg in the run's real type
    if(!binary) {
        LDWriter o;
        if(!o.open(path)) return -1;
        o.str("2D");
        o.ch('\n');
        if(sizeof(R) == 8) o.f64(p->freq0);
        else o.f32((float)p->freq0);
        o.ch('\n');
        o.i32(p->NSz);
        for(int i = 0; i < p->NSz; ++i) o.f32(p->Sz[i]);
        o.ch('\n');
        o.i32(p->NRz);
        for(int i = 0; i < p->NRz; ++i) o.f32(p->Rz[i]);
        o.ch('\n');
        o.i32(p->NRr);
        for(int i = 0; i < p->NRr; ++i) {
            if(sizeof(R) == 8) o.f64(p->Rr[i]);
            else o.f32((float)p->Rr[i]);
        }
        o.ch('\n');
        for(int isz = 0; isz < p->NSz; ++isz) {
            o.i32(maxn[isz]);
            o.ch('\n');
            for(int iz = 0; iz < p->NRz_per_range; ++iz)
                for(int ir = 0; ir < p->NRr; ++ir) {
                    size_t b = rcv_index(p, isz, iz, ir);
                    o.i32(narr[b]);
                    o.ch('\n');
                    for(int k = 0; k < narr[b]; ++k) {
                        const bhc_arrival &a = arr[b * max_narr + k];
                        o.f32(a.a);
                        o.f32((float)RadDeg * a.Phase);
                        o.f32(a.delay[0]);
                        o.f32(a.delay[1]);
                        o.f32(a.SrcDeclAngle);
                        o.f32(a.RcvrDeclAngle);
                        o.i32(a.NTopBnc);
                        o.i32(a.NBotBnc);
                        o.ch('\n');
                    }
                }
        }
        return o.close() ? 0 : -1;
    }
    UnfWriter o;
    if(!o.open(path)) return -1;
    o.rec();
    o.bytes("'2D'", 4);
    o.rec();
    o.v((float)p->freq0);
    o.rec();
    o.v((int32_t)p->NSz);
    o.bytes(p->Sz, sizeof(float) * p->NSz);
    o.rec();
    o.v((int32_t)p->NRz);
    o.bytes(p->Rz, sizeof(float) * p->NRz);
    o.rec();
    o.v((int32_t)p->NRr);
    for(int i = 0; i < p->NRr; ++i) o.v((R)p->Rr[i]);
    for(int isz = 0; isz < p->NSz; ++isz) {
        o.rec();
        o.v((int32_t)maxn[isz]);
        for(int iz = 0; iz < p->NRz_per_range; ++iz)
            for(int ir = 0; ir < p->NRr; ++ir) {
                size_t b = rcv_index(p, isz, iz, ir);
                o.rec();
                o.v((int32_t)narr[b]);
                for(int k = 0; k < narr[b]; ++k) {
                    const bhc_arrival &a = arr[b * max_narr + k];
                    o.rec();
                    o.v(a.a);
                    o.v((float)(RadDeg * (R)a.Phase));
                    o.v(a.delay[0]);
                    o.v(a.delay[1]);
                    o.v(a.SrcDeclAngle);
                    o.v(a.RcvrDeclAngle);
                    o.v((float)a.NTopBnc);
                    o.v((float)a.NBotBnc);
                }
            }
    }
    return o.close() ? 0 : -1;
}

template<typename R> int write_ray(const char *path, const bhc_problem2d *p, const char *title,
                                   const bhc_ray_pt2d *pts, const int32_t *nsteps, int32_t max_points,
                                   const double *decl_deg, int32_t nrays)
{
    LDWriter o;
    if(!o.open(path)) return -1;
    auto real_item = [&](double v) {
        if(sizeof(R) == 8) o.f64(v);
        else o.f32((float)v);
    };
    o.str(title);
    o.ch('\n');
    real_item(p->freq0);
    o.ch('\n');
    o.i32(1);
    o.i32(1);
    o.i32(p->NSz);
    o.ch('\n');
    o.i32(p->Nalpha);
    o.i32(1);
    o.ch('\n');
    real_item(p->top_hs.depth);
    o.ch('\n');
    real_item(p->bot_hs.depth);
    o.ch('\n');
    o.str("rz");
    o.ch('\n');
    for(int r = 0; r < nrays; ++r) {
        int n = nsteps[r];
        if(n <= 0) continue;
        if(n > max_points) n = max_points;
        const bhc_ray_pt2d *ray = pts + (size_t)r * max_points;
        real_item(decl_deg[r]);
        o.ch('\n');
        o.i32(n);
        o.i32(ray[n - 1].NumTopBnc);
        o.i32(ray[n - 1].NumBotBnc);
        o.ch('\n');
        for(int k = 0; k < n; ++k) {
            real_item(ray[k].x[0]);
            real_item(ray[k].x[1]);
            o.ch('\n');
        }
    }
    return o.close() ? 0 : -1;
}

} // namespace

extern "C" int bhc_write_shd2d(const char *path, const bhc_problem2d *prob, const char *title, const float *field)
{
    if(!path || !prob || !field) return -1;
    return prob->precision == 64 ? write_shd<double>(path, prob, title, field)
                                 : write_shd<float>(path, prob, title, field);
}

// PostProcessArrivals, 2D (src/mode/arr.cpp:24-73): counts clamped to
// max_narr, amplitudes scaled by the cylindrical spreading 1/sqrt(r) (1e5 at
// r = 0, 4 sqrt(pi) for a line source), maximum count per source into maxn.
extern "C" int bhc_postprocess_arrivals2d(const bhc_problem2d *p, bhc_arrival *arr, int32_t *narr, int32_t max_narr,
                                          int32_t *maxn)
{
    if(!p || !arr || !narr || !maxn || max_narr < 1) return -1;
    const bool fp64 = p->precision == 64;
    for(int isz = 0; isz < p->NSz; ++isz) {
        int32_t m = 0;
        for(int iz = 0; iz < p->NRz_per_range; ++iz)
            for(int ir = 0; ir < p->NRr; ++ir) {
                size_t b = rcv_index(p, isz, iz, ir);
                if(narr[b] > max_narr) narr[b] = max_narr;
                m = std::max(m, narr[b]);
                float factor;
                double r = p->Rr[ir];
                if(p->RunType[3] == 'X') {
                    factor = fp64 ? (float)(4.0 * std::sqrt(M_PI)) : 4.0f * std::sqrt((float)M_PI);
                } else if(fp64 ? (std::fabs(r) < 2.220446049250313e-16) : (std::fabs((float)r) < 1.1920929e-07f)) {
                    factor = 1e5f;
                } else {
                    factor = fp64 ? (float)(1.0 / std::sqrt(r)) : 1.0f / std::sqrt((float)r);
                }
                for(int k = 0; k < narr[b]; ++k) arr[b * max_narr + k].a *= factor;
            }
        maxn[isz] = m;
    }
    return 0;
}

// .arr file; binary != 0 writes the unformatted layout (run types 'a' / 'v').
extern "C" int bhc_write_arr2d(const char *path, const bhc_problem2d *prob, const bhc_arrival *arr,
                               const int32_t *narr, int32_t max_narr, const int32_t *maxn, int binary)
{
    if(!path || !prob || !arr || !narr || !maxn) return -1;
    return prob->precision == 64 ? write_arr<double>(path, prob, arr, narr, max_narr, maxn, binary != 0)
                                 : write_arr<float>(path, prob, arr, narr, max_narr, maxn, binary != 0);
}

// .ray file: nrays rays of max_points points each (first nsteps[r] valid),
// with the launch angle of each ray in degrees.
extern "C" int bhc_write_ray2d(const char *path, const bhc_problem2d *prob, const char *title,
                               const bhc_ray_pt2d *pts, const int32_t *nsteps, int32_t max_points,
                               const double *decl_deg, int32_t nrays)
{
    if(!path || !prob || !pts || !nsteps || !decl_deg || max_points < 1) return -1;
    return prob->precision == 64 ? write_ray<double>(path, prob, title, pts, nsteps, max_points, decl_deg, nrays)
                                 : write_ray<float>(path, prob, title, pts, nsteps, max_points, decl_deg, nrays);
}

// 3D .ray file: header with NSx NSy NSz, Nalpha Nbeta and "xyz"; per ray the
// launch declination in radians (SrcDeclAngle * DegRad, src/mode/ray.hpp:226-230),
// the point count, the bounce counts and every point (x, y, z).
template<typename R> int write_ray3(const char *path, const bhc_problem3d *p, const char *title,
                                    const bhc_ray_pt3d *pts, const int32_t *nsteps, int32_t max_points, int32_t nrays)
{
    LDWriter o;
    if(!o.open(path)) return -1;
    auto real_item = [&](double v) {
        if(sizeof(R) == 8) o.f64(v);
        else o.f32((float)v);
    };
    o.str(title);
    o.ch('\n');
    real_item(p->freq0);
    o.ch('\n');
    o.i32(p->NSx);
    o.i32(p->NSy);
    o.i32(p->NSz);
    o.ch('\n');
    o.i32(p->Nalpha);
    o.i32(p->Nbeta);
    o.ch('\n');
    real_item(p->top_hs.depth);
    o.ch('\n');
    real_item(p->bot_hs.depth);
    o.ch('\n');
    o.str("xyz");
    o.ch('\n');
    const R RadDeg = (R)180.0f / (R)M_PI, DegRad = (R)M_PI / (R)180.0f;
    for(int r = 0; r < nrays; ++r) {
        int n = nsteps[r];
        if(n <= 0) continue;
        if(n > max_points) n = max_points;
        int ia = p->iSingle_alpha >= 1 ? p->iSingle_alpha - 1 : r % p->Nalpha;
        R decl = RadDeg * (R)p->alpha[ia];
        const bhc_ray_pt3d *ray = pts + (size_t)r * max_points;
        real_item((double)(decl * DegRad));
        o.ch('\n');
        o.i32(n);
        o.i32(ray[n - 1].NumTopBnc);
        o.i32(ray[n - 1].NumBotBnc);
        o.ch('\n');
        for(int k = 0; k < n; ++k) {
            real_item(ray[k].x[0]);
            real_item(ray[k].x[1]);
            real_item(ray[k].x[2]);
            o.ch('\n');
        }
    }
    return o.close() ? 0 : -1;
}

extern "C" int bhc_write_ray3d(const char *path, const bhc_problem3d *prob, const char *title,
                               const bhc_ray_pt3d *pts, const int32_t *nsteps, int32_t max_points, int32_t nrays)
{
    if(!path || !prob || !pts || !nsteps || max_points < 1) return -1;
    return prob->precision == 64 ? write_ray3<double>(path, prob, title, pts, nsteps, max_points, nrays)
                                 : write_ray3<float>(path, prob, title, pts, nsteps, max_points, nrays);
}

// Receiver index of the 3D field / arrival arrays (GetFieldAddr, src/common.hpp:393-405).
static size_t rcv_index3(const bhc_problem3d *p, int isz, int isx, int isy, int itheta, int iz, int ir)
{
    return (((((size_t)isz * p->NSx + isx) * p->NSy + isy) * p->Ntheta + itheta) * p->NRz + iz) * p->NRr + ir;
}

extern "C" int bhc_postprocess_arrivals3d(const bhc_problem3d *p, bhc_arrival *arr, int32_t *narr, int32_t max_narr,
                                          int32_t *maxn)
{
    (void)arr; // 3D arrivals keep their amplitudes (factor 1)
    if(!p || !narr || !maxn || max_narr < 1) return -1;
    for(int isz = 0; isz < p->NSz; ++isz)
        for(int isx = 0; isx < p->NSx; ++isx)
            for(int isy = 0; isy < p->NSy; ++isy) {
                int32_t m = 0;
                for(int it = 0; it < p->Ntheta; ++it)
                    for(int iz = 0; iz < p->NRz; ++iz)
                        for(int ir = 0; ir < p->NRr; ++ir) {
                            size_t b = rcv_index3(p, isz, isx, isy, it, iz, ir);
                            if(narr[b] > max_narr) narr[b] = max_narr;
                            m = std::max(m, narr[b]);
                        }
                maxn[((size_t)isz * p->NSx + isx) * p->NSy + isy] = m;
            }
    return 0;
}

// 3D .arr: the source and receiver coordinate vectors keep their types
// (Sx / Sy / Rr / theta real, Sz / Rz float); the ASCII phase is written in the
// real type (RadDeg * Phase), the binary one as float.
template<typename R> static int write_arr3(const char *path, const bhc_problem3d *p, const bhc_arrival *arr,
                                           const int32_t *narr, int32_t max_narr, const int32_t *maxn, bool binary)
{
    const R RadDeg = (R)180 / (R)M_PI;
    if(!binary) {
        LDWriter o;
        if(!o.open(path)) return -1;
        auto real_item = [&](double v) {
            if(sizeof(R) == 8) o.f64(v);
            else o.f32((float)v);
        };
        o.str("3D");
        o.ch('\n');
        real_item(p->freq0);
        o.ch('\n');
        o.i32(p->NSx);
        for(int i = 0; i < p->NSx; ++i) real_item(p->Sx[i]);
        o.ch('\n');
        o.i32(p->NSy);
        for(int i = 0; i < p->NSy; ++i) real_item(p->Sy[i]);
        o.ch('\n');
        o.i32(p->NSz);
        for(int i = 0; i < p->NSz; ++i) o.f32(p->Sz[i]);
        o.ch('\n');
        o.i32(p->NRz);
        for(int i = 0; i < p->NRz; ++i) o.f32(p->Rz[i]);
        o.ch('\n');
        o.i32(p->NRr);
        for(int i = 0; i < p->NRr; ++i) real_item(p->Rr[i]);
        o.ch('\n');
        o.i32(p->Ntheta);
        for(int i = 0; i < p->Ntheta; ++i) real_item(p->theta[i]);
        o.ch('\n');
        for(int isz = 0; isz < p->NSz; ++isz)
            for(int isx = 0; isx < p->NSx; ++isx)
                for(int isy = 0; isy < p->NSy; ++isy) {
                    o.i32(maxn[((size_t)isz * p->NSx + isx) * p->NSy + isy]);
                    o.ch('\n');
                    for(int it = 0; it < p->Ntheta; ++it)
                        for(int iz = 0; iz < p->NRz; ++iz)
                            for(int ir = 0; ir < p->NRr; ++ir) {
                                size_t b = rcv_index3(p, isz, isx, isy, it, iz, ir);
                                o.i32(narr[b]);
                                o.ch('\n');
                                for(int k = 0; k < narr[b]; ++k) {
                                    const bhc_arrival &a = arr[b * max_narr + k];
                                    o.f32(a.a);
                                    real_item((double)(RadDeg * (R)a.Phase));
                                    o.f32(a.delay[0]);
                                    o.f32(a.delay[1]);
                                    o.f32(a.SrcDeclAngle);
                                    o.f32(a.SrcAzimAngle);
                                    o.f32(a.RcvrDeclAngle);
                                    o.f32(a.RcvrAzimAngle);
                                    o.i32(a.NTopBnc);
                                    o.i32(a.NBotBnc);
                                    o.ch('\n');
                                }
                            }
                }
        return o.close() ? 0 : -1;
    }
    UnfWriter o;
    if(!o.open(path)) return -1;
    o.rec();
    o.bytes("'3D'", 4);
    o.rec();
    o.v((float)p->freq0);
    o.rec();
    o.v((int32_t)p->NSx);
    for(int i = 0; i < p->NSx; ++i) o.v((R)p->Sx[i]);
    o.rec();
    o.v((int32_t)p->NSy);
    for(int i = 0; i < p->NSy; ++i) o.v((R)p->Sy[i]);
    o.rec();
    o.v((int32_t)p->NSz);
    o.bytes(p->Sz, sizeof(float) * p->NSz);
    o.rec();
    o.v((int32_t)p->NRz);
    o.bytes(p->Rz, sizeof(float) * p->NRz);
    o.rec();
    o.v((int32_t)p->NRr);
    for(int i = 0; i < p->NRr; ++i) o.v((R)p->Rr[i]);
    o.rec();
    o.v((int32_t)p->Ntheta);
    for(int i = 0; i < p->Ntheta; ++i) o.v((R)p->theta[i]);
    for(int isz = 0; isz < p->NSz; ++isz)
        for(int isx = 0; isx < p->NSx; ++isx)
            for(int isy = 0; isy < p->NSy; ++isy) {
                o.rec();
                o.v((int32_t)maxn[((size_t)isz * p->NSx + isx) * p->NSy + isy]);
                for(int it = 0; it < p->Ntheta; ++it)
                    for(int iz = 0; iz < p->NRz; ++iz)
                        for(int ir = 0; ir < p->NRr; ++ir) {
                            size_t b = rcv_index3(p, isz, isx, isy, it, iz, ir);
                            o.rec();
                            o.v((int32_t)narr[b]);
                            for(int k = 0; k < narr[b]; ++k) {
                                const bhc_arrival &a = arr[b * max_narr + k];
                                o.rec();
                                o.v(a.a);
                                o.v((float)(RadDeg * (R)a.Phase));
                                o.v(a.delay[0]);
                                o.v(a.delay[1]);
                                o.v(a.SrcDeclAngle);
                                o.v(a.SrcAzimAngle);
                                o.v(a.RcvrDeclAngle);
                                o.v(a.RcvrAzimAngle);
                                o.v((float)a.NTopBnc);
                                o.v((float)a.NBotBnc);
                            }
                        }
            }
    return o.close() ? 0 : -1;
}

extern "C" int bhc_write_arr3d(const char *path, const bhc_problem3d *prob, const bhc_arrival *arr,
                               const int32_t *narr, int32_t max_narr, const int32_t *maxn, int binary)
{
    if(!path || !prob || !arr || !narr || !maxn || max_narr < 1) return -1;
    return prob->precision == 64 ? write_arr3<double>(path, prob, arr, narr, max_narr, maxn, binary != 0)
                                 : write_arr3<float>(path, prob, arr, narr, max_narr, maxn, binary != 0);
}

// The list-directed writer as a stream, for companion files in the same layout.
struct bhc_ldfile {
    LDWriter w;
};

extern "C" bhc_ldfile *bhc_ld_open(const char *path)
{
    bhc_ldfile *f = new bhc_ldfile();
    if(!f->w.open(path)) {
        delete f;
        return nullptr;
    }
    return f;
}
extern "C" void bhc_ld_i32(bhc_ldfile *f, int32_t v) { f->w.i32(v); }
extern "C" void bhc_ld_f32(bhc_ldfile *f, float v) { f->w.f32(v); }
extern "C" void bhc_ld_f64(bhc_ldfile *f, double v) { f->w.f64(v); }
extern "C" void bhc_ld_str(bhc_ldfile *f, const char *s) { f->w.str(s); }
extern "C" void bhc_ld_char(bhc_ldfile *f, char c) { f->w.ch(c); }
extern "C" int bhc_ld_close(bhc_ldfile *f)
{
    bool ok = f->w.close();
    delete f;
    return ok ? 0 : -1;
}
