#ifndef _GNU_SOURCE
#define _GNU_SOURCE
#endif
// host3d.cpp -- host front end of the 3D path: preprocessing of a raw 3D
// environment into the problem the 3D launch layer consumes, and ScalePressure.
//
// The depth-only part (sound speed profile, halfspaces, reflection tables,
// sources in depth, receivers, declinations, step size, beam pattern) is the
// 2D preprocessing applied to the same values (bhc_preprocess2d): the
// reference's modules compute them identically for O3D. The 3D additions
// follow src/module/{ssp.hpp:144-186 + 453-489 + 560-581 (hexahedral grid),
// boundary.hpp:107-150 + 499-512 + 600-650 (triangulated grids),
// sxsy.hpp:105-112, rcvrbearings.hpp:98-110, rayangles.hpp:181-189,
// beaminfo.hpp:226-230}, in the run's real type.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bhc_hip.h"

struct bhc_problem3d_owner {
    bhc_problem3d v;
    bhc_problem_owner *p2 = nullptr;
    std::vector<double> z3, rho3, hx, hy, hc, hcz, Sx, Sy, trcv, theta, beta, bcos, bsin, atip, btyp;
    std::vector<bhc_bdry_node3d> ati, bty;
    ~bhc_problem3d_owner() { bhc_problem_free(p2); }
};

void bhc_crci_w(int precision, double freq0, double c, double alpha, double out[2]); // host2d.cpp

namespace {

thread_local std::string g_err3;
int fail3(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err3 = buf;
    return -1;
}

template<typename real> struct V3 {
    real x, y, z;
};
template<typename real> V3<real> sub(V3<real> a, V3<real> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template<typename real> V3<real> cross(V3<real> a, V3<real> b)
{
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template<typename real> real len(V3<real> a) { return std::sqrt((a.x * a.x + a.y * a.y) + a.z * a.z); }

inline double atan2_(double y, double x) { return std::atan2(y, x); }
inline float atan2_(float y, float x) { return ::atan2f(y, x); }

// Node normals and, on curvilinear grids, the cells' curvatures:
// ComputeBdryTangentNormal 3D (src/module/boundary.hpp:650-830). The node
// angle phi_xx exists only at the cells' lower corners; a cell in the last
// row or column reads its neighbour's, which the reference leaves unset: 0.
template<typename real>
void curvature(const std::vector<V3<real>> &x, int nx, int ny, bool curvilinear, std::vector<bhc_bdry_node3d> &out)
{
    std::vector<V3<real>> nu(x.size());
    std::vector<real> phi(x.size(), (real)0.0f);
    auto X = [&](int i, int j) -> const V3<real> & { return x[(size_t)i * ny + j]; };
    for(int ix = 0; ix < nx; ++ix)
        for(int iy = 0; iy < ny; ++iy) {
            real mx, my;
            if(ix == 0) mx = (X(ix + 1, iy).z - X(ix, iy).z) / (X(ix + 1, iy).x - X(ix, iy).x);
            else if(ix == nx - 1) mx = (X(ix, iy).z - X(ix - 1, iy).z) / (X(ix, iy).x - X(ix - 1, iy).x);
            else mx = (X(ix + 1, iy).z - X(ix - 1, iy).z) / (X(ix + 1, iy).x - X(ix - 1, iy).x);
            if(iy == 0) my = (X(ix, iy + 1).z - X(ix, iy).z) / (X(ix, iy + 1).y - X(ix, iy).y);
            else if(iy == ny - 1) my = (X(ix, iy).z - X(ix, iy - 1).z) / (X(ix, iy).y - X(ix, iy - 1).y);
            else my = (X(ix, iy + 1).z - X(ix, iy - 1).z) / (X(ix, iy + 1).y - X(ix, iy - 1).y);
            V3<real> n = {-mx, -my, (real)1.0f};
            size_t k = (size_t)ix * ny + iy;
            if(ix < nx - 1 && iy < ny - 1) phi[k] = atan2_(n.z, n.x);
            nu[k] = n;
            real l = len(n);
            out[k].Noden[0] = (double)(n.x / l);
            out[k].Noden[1] = (double)(n.y / l);
            out[k].Noden[2] = (double)(n.z / l);
        }
    if(!curvilinear) return;
    for(int ix = 0; ix < nx - 1; ++ix)
        for(int iy = 0; iy < ny - 1; ++iy) {
            size_t k = (size_t)ix * ny + iy, kx = (size_t)(ix + 1) * ny + iy, ky = k + 1;
            real zxx = -(nu[kx].x - nu[k].x) / (x[kx].x - x[k].x);
            V3<real> tv = sub(x[kx], x[k]);
            real L = std::sqrt(tv.x * tv.x + tv.z * tv.z);
            real kxx = (phi[kx] - phi[k]) / L;
            real zxy = -(nu[ky].x - nu[k].x) / (x[ky].y - x[k].y);
            tv = sub(x[ky], x[k]);
            L = std::sqrt(tv.y * tv.y + tv.z * tv.z);
            real kxy = (phi[ky] - phi[k]) / L;
            real zyy = -(nu[ky].y - nu[k].y) / (x[ky].y - x[k].y);
            real kyy = (phi[ky] - phi[k]) / L;
            L = len(nu[k]);
            out[k].z_xx = (double)(zxx / L);
            out[k].z_xy = (double)(zxy / L);
            out[k].z_yy = (double)(zyy / L);
            out[k].kappa_xx = (double)kxx;
            out[k].kappa_xy = (double)kxy;
            out[k].kappa_yy = (double)kyy;
        }
}

inline void sincos_(double a, double *s, double *c) { ::sincos(a, s, c); }
inline void sincos_(float a, float *s, float *c) { ::sincosf(a, s, c); }

// Triangulated grid (or the flat default) -> nodes with the two cell normals.
template<typename real>
int grid(const bhc_bdry3d_in &in, double depth, bool isTop, std::vector<bhc_bdry_node3d> &out, int32_t &nx,
         int32_t &ny, char type[2])
{
    type[0] = 'R';
    type[1] = 'S';
    if(in.nx <= 0) { // Default, src/module/boundary.hpp:56-78
        const real big = (real)1.0e25;
        const real d = (real)depth, nt = isTop ? (real)-1.0 : (real)1.0;
        nx = ny = 2;
        out.assign(4, bhc_bdry_node3d{});
        const real xs[4] = {-big, -big, big, big}, ys[4] = {-big, big, -big, big};
        for(int i = 0; i < 4; ++i) {
            out[i].x[0] = (double)xs[i];
            out[i].x[1] = (double)ys[i];
            out[i].x[2] = (double)d;
            out[i].n1[2] = out[i].n2[2] = (double)nt;
            out[i].province = -1;
        }
        return 0;
    }
    if((in.type[0] != 'R' && in.type[0] != 'C') || (in.type[1] != 'S' && in.type[1] != 'L'))
        return fail3("3D boundary type '%c%c' is not supported ('R' or 'C', then 'S' or 'L')", in.type[0],
                     in.type[1]);
    type[0] = in.type[0];
    type[1] = in.type[1];
    if(in.nx < 2 || in.ny < 2) return fail3("3D boundary grid needs at least 2 x 2 nodes");
    nx = in.nx;
    ny = in.ny;
    std::vector<V3<real>> x((size_t)nx * ny);
    for(int ix = 0; ix < nx; ++ix)
        for(int iy = 0; iy < ny; ++iy)
            x[(size_t)ix * ny + iy] = {(real)in.x_km[ix] * (real)1000.0f, (real)in.y_km[iy] * (real)1000.0f,
                                       (real)in.z[(size_t)ix * ny + iy]};
    for(int ix = 1; ix < nx; ++ix)
        if(!(x[(size_t)ix * ny].x > x[(size_t)(ix - 1) * ny].x)) return fail3("3D boundary x must increase");
    for(int iy = 1; iy < ny; ++iy)
        if(!(x[iy].y > x[iy - 1].y)) return fail3("3D boundary y must increase");
    out.assign((size_t)nx * ny, bhc_bdry_node3d{});
    for(size_t i = 0; i < x.size(); ++i) {
        out[i].x[0] = (double)x[i].x;
        out[i].x[1] = (double)x[i].y;
        out[i].x[2] = (double)x[i].z;
    }
    // ComputeBdryTangentNormal, src/module/boundary.hpp:600-650
    for(int ix = 0; ix < nx - 1; ++ix)
        for(int iy = 0; iy < ny - 1; ++iy) {
            V3<real> p1 = x[(size_t)ix * ny + iy], p2 = x[(size_t)(ix + 1) * ny + iy];
            V3<real> p3 = x[(size_t)(ix + 1) * ny + iy + 1], p4 = x[(size_t)ix * ny + iy + 1];
            V3<real> n1 = cross(sub(p2, p1), sub(p3, p1));
            V3<real> n2 = cross(sub(p3, p1), sub(p4, p1));
            if(isTop) {
                n1 = {-n1.x, -n1.y, -n1.z};
                n2 = {-n2.x, -n2.y, -n2.z};
            }
            real l1 = len(n1), l2 = len(n2);
            bhc_bdry_node3d &b = out[(size_t)ix * ny + iy];
            b.n1[0] = (double)(n1.x / l1);
            b.n1[1] = (double)(n1.y / l1);
            b.n1[2] = (double)(n1.z / l1);
            b.n2[0] = (double)(n2.x / l2);
            b.n2[1] = (double)(n2.y / l2);
            b.n2[2] = (double)(n2.z / l2);
        }
    curvature<real>(x, nx, ny, in.type[0] == 'C', out);
    for(size_t i = 0; i < out.size(); ++i) out[i].province = -1;
    if(in.type[1] == 'L') {
        if(!in.province || !in.prov || in.nprov < 1) return fail3("province grid needs provinces");
        for(size_t i = 0; i < out.size(); ++i) {
            int pv = in.province[i];
            if(pv < 1 || pv > in.nprov) return fail3("province number %d outside 1..%d", pv, in.nprov);
            out[i].province = pv - 1;
        }
    }
    return 0;
}

template<typename real> int preprocess3d(const bhc_env3d *e, bhc_problem3d_owner *o)
{
    const bhc_problem2d &p = *bhc_problem_view(o->p2);
    bhc_problem3d &v = o->v;
    memset(&v, 0, sizeof(v));
    v.precision = e->precision;
    memcpy(v.RunType, p.RunType, 7);
    v.RunType[5] = e->RunType[5] == '2' ? '2' : '3';   // Nx2D or full 3D
    memcpy(v.BeamType, p.BeamType, 4);
    v.ssp_type = e->ssp_type;
    v.freq0 = p.freq0;
    v.ssp_npts = p.ssp_npts;
    v.ssp_z = p.ssp_z;
    v.ssp_c = p.ssp_c;
    v.ssp_cz = p.ssp_cz;
    v.ssp_n2 = p.ssp_n2;
    v.ssp_n2z = p.ssp_n2z;
    v.ssp_coef = p.ssp_coef;
    v.ssp_rho = p.ssp_rho;
    if(e->ssp_type == 'H') {
        int nx = e->hex_nx, ny = e->hex_ny, nz = e->hex_nz;
        if(nx < 2 || ny < 2 || nz < 2) return fail3("hexahedral SSP needs at least 2 points in x, y and z");
        for(int i = 0; i < nx; ++i) o->hx.push_back((double)((real)e->hex_x_km[i] * (real)1000.0f));
        for(int i = 0; i < ny; ++i) o->hy.push_back((double)((real)e->hex_y_km[i] * (real)1000.0f));
        // SegZToZ (src/module/ssp.hpp:560-581): the depth table becomes the
        // grid's, densities beyond the environment's profile are undefined
        for(int i = 0; i < nz; ++i) {
            o->z3.push_back((double)(real)e->hex_z[i]);
            o->rho3.push_back(i < p.ssp_npts ? p.ssp_rho[i] : (double)NAN);
        }
        for(int i = 1; i < nx; ++i)
            if(!(o->hx[i] > o->hx[i - 1])) return fail3("hexahedral SSP x must increase");
        for(int i = 1; i < ny; ++i)
            if(!(o->hy[i] > o->hy[i - 1])) return fail3("hexahedral SSP y must increase");
        for(int i = 1; i < nz; ++i)
            if(!(o->z3[i] > o->z3[i - 1])) return fail3("hexahedral SSP z must increase");
        size_t n = (size_t)nx * ny * nz;
        o->hc.resize(n);
        for(size_t i = 0; i < n; ++i) o->hc[i] = (double)(real)e->hex_c[i];
        o->hcz.resize((size_t)nx * ny * (nz - 1));
        for(int ix = 0; ix < nx; ++ix)
            for(int iy = 0; iy < ny; ++iy)
                for(int iz = 1; iz < nz; ++iz) {
                    size_t b = (size_t)ix * ny + iy;
                    real c1 = (real)o->hc[b * nz + iz], c0 = (real)o->hc[b * nz + iz - 1];
                    o->hcz[b * (nz - 1) + iz - 1] = (double)((c1 - c0) / ((real)o->z3[iz] - (real)o->z3[iz - 1]));
                }
        v.hex_nx = nx;
        v.hex_ny = ny;
        v.hex_nz = nz;
        v.hex_x = o->hx.data();
        v.hex_y = o->hy.data();
        v.hex_z = o->z3.data();
        v.hex_c = o->hc.data();
        v.hex_cz = o->hcz.data();
        v.ssp_npts = nz;
        v.ssp_z = o->z3.data();
        v.ssp_rho = o->rho3.data();
    }
    v.top_hs = p.top_hs;
    v.bot_hs = p.bot_hs;
    if(grid<real>(e->ati, p.top_hs.depth, true, o->ati, v.ati_nx, v.ati_ny, v.ati_type)) return -1;
    if(grid<real>(e->bty, p.bot_hs.depth, false, o->bty, v.bty_nx, v.bty_ny, v.bty_type)) return -1;
    v.ati = o->ati.data();
    v.bty = o->bty.data();
    // provinces: cP, cS from dB/wavelength at z = 1e20 (boundary.hpp:527-537), rho
    auto provs = [&](const bhc_bdry3d_in &in, const char *type, std::vector<double> &dst, int32_t &n,
                     const double *&ptr) {
        n = 0;
        ptr = nullptr;
        if(type[1] != 'L') return;
        for(int i = 0; i < in.nprov; ++i) {
            const double *r = in.prov + 5 * (size_t)i;
            double cp[2], cs[2];
            bhc_crci_w(e->precision, e->freq0, r[0], r[3], cp);
            bhc_crci_w(e->precision, e->freq0, r[1], r[4], cs);
            dst.insert(dst.end(), {cp[0], cp[1], cs[0], cs[1], (double)(real)r[2]});
        }
        n = in.nprov;
        ptr = dst.data();
    };
    provs(e->ati, v.ati_type, o->atip, v.ati_nprov, v.ati_prov);
    provs(e->bty, v.bty_type, o->btyp, v.bty_nprov, v.bty_prov);
    v.trc_npts = p.trc_npts;
    v.brc_npts = p.brc_npts;
    v.trc = p.trc;
    v.brc = p.brc;
    v.NSx = e->NSx;
    v.NSy = e->NSy;
    v.NSz = p.NSz;
    v.NRz = p.NRz;
    v.NRr = p.NRr;
    v.Ntheta = e->Ntheta;
    for(int i = 0; i < e->NSx; ++i) o->Sx.push_back((double)((real)e->Sx_km[i] * (real)1000.0f));
    for(int i = 0; i < e->NSy; ++i) o->Sy.push_back((double)((real)e->Sy_km[i] * (real)1000.0f));
    v.Sx = o->Sx.data();
    v.Sy = o->Sy.data();
    v.Sz = p.Sz;
    v.Rz = p.Rz;
    v.Rr = p.Rr;
    v.Delta_r = p.Delta_r;
    const real DegRad = (real)M_PI / (real)180.0f;
    for(int i = 0; i < e->Ntheta; ++i) {
        real th = DegRad * (real)e->theta_deg[i], s, c;
        sincos_(th, &s, &c);
        o->trcv.push_back((double)c);
        o->trcv.push_back((double)s);
    }
    v.t_rcvr = o->trcv.data();
    for(int i = 0; i < e->Ntheta; ++i) o->theta.push_back((double)(real)e->theta_deg[i]);
    v.theta = o->theta.data();
    v.Nalpha = p.Nalpha;
    v.iSingle_alpha = p.iSingle_alpha;
    v.alpha = p.alpha;
    v.alpha_cos = p.alpha_cos;
    v.alpha_sin = p.alpha_sin;
    v.Dalpha = p.Dalpha;
    // an Nx2D run that is not a ray run traces its beams on the receiver
    // bearings (rayangles.hpp:163-178)
    const bool on_bearings = e->RunType[5] == '2' && e->RunType[0] != 'R';
    const int nb = on_bearings ? e->Ntheta : e->Nbeta;
    const double *bdeg = on_bearings ? e->theta_deg : e->beta_deg;
    v.Nbeta = nb;
    v.iSingle_beta = e->iSingle_beta;
    if(v.iSingle_beta > nb) return fail3("iSingle_beta out of range");
    std::vector<real> b(nb);
    for(int i = 0; i < nb; ++i) b[i] = (real)bdeg[i] * DegRad;
    for(int i = 1; i < nb; ++i)
        if(!(b[i] > b[i - 1])) return fail3("bearing launch angles must increase");
    real Db = (real)0.0f;
    if(nb != 1) Db = (b[nb - 1] - b[0]) / (real)(nb - 1);
    v.Dbeta = (double)Db;
    for(int i = 0; i < nb; ++i) {
        real s, c;
        sincos_(b[i], &s, &c);
        o->beta.push_back((double)b[i]);
        o->bcos.push_back((double)c);
        o->bsin.push_back((double)s);
    }
    v.beta = o->beta.data();
    v.beta_cos = o->bcos.data();
    v.beta_sin = o->bsin.data();
    v.deltas = p.deltas;
    v.box_x = (double)((real)e->box_x_km * (real)1000.0f);
    v.box_y = (double)((real)e->box_y_km * (real)1000.0f);
    v.box_z = (double)(real)e->box_z;
    v.NSBPPts = p.NSBPPts;
    v.sbp = p.sbp;
    return 0;
}

} // namespace

extern "C" int bhc_preprocess3d(const bhc_env3d *e, bhc_problem3d_owner **out)
{
    *out = nullptr;
    if(!e) return fail3("null environment");
    if(e->precision != 64 && e->precision != 32) return fail3("precision must be 32 or 64");
    if(e->NSx < 1 || e->NSy < 1 || e->NSz < 1 || e->NRz < 1 || e->NRr < 1 || e->Ntheta < 1 || e->Nalpha < 1
       || e->Nbeta < 1)
        return fail3("sources, receivers and launch angles must be non-empty");
    // TL, arrivals ('A' ASCII / 'a' binary file), eigenray, arrivals + eigenray
    // ('V', 'v') and ray run types with geometric beams, Cartesian
    // (src/module/runtype.hpp maps '^' and ' ' to 'G') or ray-centred ('g', 'b').
    const char r0 = e->RunType[0];
    if(r0 != 'C' && r0 != 'S' && r0 != 'I' && r0 != 'R' && r0 != 'A' && r0 != 'a' && r0 != 'E' && r0 != 'V'
       && r0 != 'v')
        return fail3("3D run type '%c' is not supported (C / S / I TL, A / a arrivals, E eigenrays, V / v both, "
                     "R rays)", r0);
    const char r1 = e->RunType[1];
    if(r1 != 'G' && r1 != '^' && r1 != ' ' && r1 != 'B' && r1 != 'g' && r1 != 'b')
        return fail3("3D beam type '%c' is not supported (geometric beams: G / B Cartesian, g / b ray-centred)", r1);
    if(e->ssp_type != 'C' && e->ssp_type != 'N' && e->ssp_type != 'S' && e->ssp_type != 'P' && e->ssp_type != 'H')
        return fail3("3D SSP type '%c' is not supported", e->ssp_type);
    bhc_env2d e2;
    memset(&e2, 0, sizeof(e2));
    e2.precision = e->precision;
    e2.freq0 = e->freq0;
    e2.ssp_type = e->ssp_type == 'H' ? 'C' : e->ssp_type;
    e2.atten_unit[0] = e->atten_unit[0];
    e2.atten_unit[1] = e->atten_unit[1];
    e2.ssp_npts = e->ssp_npts;
    e2.ssp_z = e->ssp_z;
    e2.ssp_alphaR = e->ssp_alphaR;
    e2.ssp_alphaI = e->ssp_alphaI;
    e2.ssp_rho = e->ssp_rho;
    e2.ssp_betaR = e->ssp_betaR;
    e2.ssp_betaI = e->ssp_betaI;
    e2.fg_T = e->fg_T;
    e2.fg_S = e->fg_S;
    e2.fg_pH = e->fg_pH;
    e2.fg_z_bar = e->fg_z_bar;
    e2.top = e->top;
    e2.bot = e->bot;
    e2.trc = e->trc;
    e2.brc = e->brc;
    e2.NSz = e->NSz;
    e2.NRz = e->NRz;
    e2.NRr = e->NRr;
    e2.Sz = e->Sz;
    e2.Rz = e->Rz;
    e2.Rr_km = e->Rr_km;
    e2.Nalpha = e->Nalpha;
    e2.iSingle_alpha = e->iSingle_alpha;
    e2.alpha_deg = e->alpha_deg;
    memcpy(e2.RunType, e->RunType, 7);
    e2.RunType[5] = ' ';
    memcpy(e2.BeamType, e->BeamType, 3);
    e2.deltas = e->deltas;
    e2.box_r_km = e->box_x_km;
    e2.box_z = e->box_z;
    e2.NSBPPts = e->NSBPPts;
    e2.sbp = e->sbp;
    e2.NBioLayers = e->NBioLayers;
    e2.bio = e->bio;
    bhc_problem3d_owner *o = new bhc_problem3d_owner();
    if(bhc_preprocess2d(&e2, &o->p2)) {
        fail3("%s", bhc_host_last_error());
        delete o;
        return -1;
    }
    int rc = e->precision == 64 ? preprocess3d<double>(e, o) : preprocess3d<float>(e, o);
    if(rc) {
        delete o;
        return -1;
    }
    *out = o;
    return 0;
}

extern "C" const bhc_problem3d *bhc_problem3d_view(const bhc_problem3d_owner *p) { return p ? &p->v : nullptr; }

extern "C" void bhc_problem3d_free(bhc_problem3d_owner *p) { delete p; }

extern "C" const char *bhc_host3d_last_error(void) { return g_err3.c_str(); }

// ScalePressure (src/influence.hpp:1556-1627) for 3D geometric beams: no
// spreading factor; incoherent and semicoherent intensities become pressures.
// Nx2D runs (R3D = false) then take the 2D cylindrical spreading per receiver
// range on every bearing plane, as bhc_scale_pressure2d.
extern "C" int bhc_scale_pressure3d(const bhc_problem3d *p, float *field)
{
    if(!p || !field) return fail3("null argument");
    size_t n = (size_t)p->NSz * p->NSx * p->NSy * p->Ntheta * p->NRz * p->NRr;
    if(p->RunType[0] != 'C')
        for(size_t k = 0; k < n; ++k) {
            field[2 * k] = std::sqrt(field[2 * k]);
            field[2 * k + 1] = 0.0f;
        }
    if(p->RunType[5] != '2') return 0;
    const bool line = p->RunType[3] == 'X';
    const size_t rows = n / (size_t)p->NRr;
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
        const float f = (float)factor;
        for(size_t row = 0; row < rows; ++row) {
            size_t a = row * p->NRr + ir;
            field[2 * a] *= f;
            field[2 * a + 1] *= f;
        }
    }
    return 0;
}
