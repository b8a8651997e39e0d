// pack3d.h -- packs a preprocessed 3D problem (bhc_problem3d, doubles) into
// the 3D tracer's blob in the run's real type and fills the device view.
#pragma once
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bhc_hip.h"
#include "core3d.h"

namespace bhc3d {

// Blob layout: [depth table z, rho][depth-only profile c, cz, n2, n2z, coef, or
// the hexahedral x, y, c, cz][bty nodes][ati nodes][Rr][trc][brc][sbp]
// [alpha, cos, sin][beta, cos, sin][Sx][Sy][bearing cos / sin].
template<typename real>
int pack3(const bhc_problem3d *p, std::vector<real> &blob, Prob3D<real> &P, std::string &err)
{
    auto fail = [&](const char *m) {
        err = m;
        return -1;
    };
    memset(&P, 0, sizeof(P));
    auto put = [&](const double *src, size_t n) {
        int32_t off = (int32_t)blob.size();
        for(size_t i = 0; i < n; ++i) blob.push_back(src ? (real)src[i] : (real)0);
        return off;
    };
    const bool hex = p->ssp_type == 'H';
    int N = p->ssp_npts;
    if(N < 2) return fail("3D problem: SSP needs at least 2 depths");
    P.ssp3 = p->ssp_type;
    P.nz3 = N;
    P.o_z3 = put(p->ssp_z, N);
    P.o_rho3 = put(p->ssp_rho, N);
    Prob2D<real> &s = P.s;
    s.ssp_type = hex ? 'C' : p->ssp_type;
    s.npts = N;
    s.o_z = P.o_z3;
    s.o_rho = P.o_rho3;
    if(!hex) {
        s.o_c = put(p->ssp_c, 2 * (size_t)N);
        s.o_cz = put(p->ssp_cz, 2 * (size_t)N);
        if(p->ssp_type == 'N') {
            s.o_n2 = put(p->ssp_n2, 2 * (size_t)N);
            s.o_n2z = put(p->ssp_n2z, 2 * (size_t)N);
        }
        if(p->ssp_type == 'S' || p->ssp_type == 'P') s.o_coef = put(p->ssp_coef, 8 * (size_t)N);
    } else {
        P.hx = p->hex_nx;
        P.hy = p->hex_ny;
        if(p->hex_nz != N) return fail("3D problem: hexahedral depth count mismatch");
        P.o_hx = put(p->hex_x, P.hx);
        P.o_hy = put(p->hex_y, P.hy);
        P.o_hc = put(p->hex_c, (size_t)P.hx * P.hy * N);
        P.o_hcz = put(p->hex_cz, (size_t)P.hx * P.hy * (N - 1));
    }
    auto put_nodes = [&](const bhc_bdry_node3d *nd, int n) {
        int32_t off = (int32_t)blob.size();
        for(int i = 0; i < n; ++i) { // NODE3 reals per node: x, n1, n2
            for(int k = 0; k < 3; ++k) blob.push_back((real)nd[i].x[k]);
            for(int k = 0; k < 3; ++k) blob.push_back((real)nd[i].n1[k]);
            for(int k = 0; k < 3; ++k) blob.push_back((real)nd[i].n2[k]);
        }
        return off;
    };
    if(p->ati_nx < 2 || p->ati_ny < 2 || p->bty_nx < 2 || p->bty_ny < 2) return fail("3D problem: boundary grids need 2 x 2 nodes");
    P.bty_nx = p->bty_nx;
    P.bty_ny = p->bty_ny;
    P.ati_nx = p->ati_nx;
    P.ati_ny = p->ati_ny;
    P.o_bty = put_nodes(p->bty, p->bty_nx * p->bty_ny);
    P.o_ati = put_nodes(p->ati, p->ati_nx * p->ati_ny);
    // curvilinear / province grids: per-node extras [NODEX] and province tables
    auto put_extras = [&](const bhc_bdry_node3d *nd, int n) {
        int32_t off = (int32_t)blob.size();
        for(int i = 0; i < n; ++i) {
            const real rec[NODEX] = {(real)nd[i].Noden[0], (real)nd[i].Noden[1], (real)nd[i].Noden[2],
                                     (real)nd[i].z_xx,     (real)nd[i].z_xy,     (real)nd[i].z_yy,
                                     (real)nd[i].kappa_xx, (real)nd[i].kappa_xy, (real)nd[i].kappa_yy,
                                     (real)nd[i].province};
            for(int k = 0; k < NODEX; ++k) blob.push_back(rec[k]);
        }
        return off;
    };
    P.ati_t0 = p->ati_type[0] ? p->ati_type[0] : 'R';
    P.ati_t1 = p->ati_type[1] ? p->ati_type[1] : 'S';
    P.bty_t0 = p->bty_type[0] ? p->bty_type[0] : 'R';
    P.bty_t1 = p->bty_type[1] ? p->bty_type[1] : 'S';
    if(P.ati_t0 == 'C' || P.ati_t1 == 'L') P.o_atix = put_extras(p->ati, p->ati_nx * p->ati_ny);
    if(P.bty_t0 == 'C' || P.bty_t1 == 'L') P.o_btyx = put_extras(p->bty, p->bty_nx * p->bty_ny);
    if(P.ati_t1 == 'L') {
        if(p->ati_nprov < 1 || !p->ati_prov) return fail("3D problem: province grid without provinces");
        P.o_atip = put(p->ati_prov, 5 * (size_t)p->ati_nprov);
    }
    if(P.bty_t1 == 'L') {
        if(p->bty_nprov < 1 || !p->bty_prov) return fail("3D problem: province grid without provinces");
        P.o_btyp = put(p->bty_prov, 5 * (size_t)p->bty_nprov);
    }
    P.o_Rr = put(p->Rr, p->NRr);
    P.o_trc = put(p->trc, 3 * (size_t)p->trc_npts);
    P.o_brc = put(p->brc, 3 * (size_t)p->brc_npts);
    P.ntrc = p->trc_npts;
    P.nbrc = p->brc_npts;
    P.o_sbp = put(p->sbp, 2 * (size_t)p->NSBPPts);
    P.nsbp = p->NSBPPts;
    P.o_alpha = put(p->alpha, p->Nalpha);
    P.o_acos = put(p->alpha_cos, p->Nalpha);
    P.o_asin = put(p->alpha_sin, p->Nalpha);
    P.o_beta = put(p->beta, p->Nbeta);
    P.o_bcos = put(p->beta_cos, p->Nbeta);
    P.o_bsin = put(p->beta_sin, p->Nbeta);
    P.o_Sx = put(p->Sx, p->NSx);
    P.o_Sy = put(p->Sy, p->NSy);
    P.o_trcv = put(p->t_rcvr, 2 * (size_t)p->Ntheta);
    if(blob.size() >= (size_t)0x7FFFFFFF) return fail("3D problem too large for 32-bit offsets");
    P.run0 = p->RunType[0];
    P.run1 = p->RunType[1];
    P.NSx = p->NSx;
    P.NSy = p->NSy;
    P.NSz = p->NSz;
    P.NRz = p->NRz;
    P.NRr = p->NRr;
    P.Ntheta = p->Ntheta;
    P.Nalpha = p->Nalpha;
    P.iSingleA = p->iSingle_alpha;
    P.Nbeta = p->Nbeta;
    P.iSingleB = p->iSingle_beta;
    P.Dalpha = (real)p->Dalpha;
    P.Dbeta = (real)p->Dbeta;
    P.freq0 = (real)p->freq0;
    P.deltas = (real)p->deltas;
    P.box_x = (real)p->box_x;
    P.box_y = (real)p->box_y;
    P.box_z = (real)p->box_z;
    auto hs = [](const bhc_hs2d &h) {
        HSd<real> r;
        r.cP = C((real)h.cP[0], (real)h.cP[1]);
        r.cS = C((real)h.cS[0], (real)h.cS[1]);
        r.rho = (real)h.rho;
        r.bc = h.bc;
        return r;
    };
    P.top = hs(p->top_hs);
    P.bot = hs(p->bot_hs);
    P.rz_absmax = 0;
    for(int i = 0; i < p->NRz; ++i) P.rz_absmax = rmax(P.rz_absmax, (real)std::fabs((double)p->Rz[i]));
    P.rz_sorted = 1;
    for(int i = 1; i < p->NRz; ++i)
        if(!(p->Rz[i] > p->Rz[i - 1])) P.rz_sorted = 0;
    P.rz_rdz = (P.rz_sorted && p->NRz > 1) ? (double)(p->NRz - 1) / ((double)p->Rz[p->NRz - 1] - (double)p->Rz[0]) : 0.0;
    if(!(P.rz_rdz > 0.0 && P.rz_rdz < 1e300)) P.rz_rdz = 0.0;
    P.isGaussian = p->RunType[1] == 'B' || p->RunType[1] == 'b';
    P.isRayCen = p->RunType[1] == 'g' || p->RunType[1] == 'b';
    P.pad_rc = 0;
    P.Delta_r = (real)p->Delta_r;
    return 0;
}

// Nx2D: the 3D environment's blob (pack3) plus the receiver depths as reals,
// and the 2D view of it that the 2D influence and scatter read -- receivers,
// beam type, the field's source slots (NSz * NSx * NSy * Ntheta bearing
// planes) and the decomposition of a slot for arrivals and eigen hits.
template<typename real>
int pack_nx(const bhc_problem3d *p, std::vector<real> &blob, Prob3D<real> &N, Prob2D<real> &S, std::string &err)
{
    if(pack3<real>(p, blob, N, err)) return -1;
    memset(&S, 0, sizeof(S));
    S.o_rzs = (int32_t)blob.size();
    for(int i = 0; i < p->NRz; ++i) blob.push_back((real)p->Rz[i]);
    if(blob.size() >= (size_t)0x7FFFFFFF) {
        err = "Nx2D problem too large for 32-bit offsets";
        return -1;
    }
    S.staged_reals = 0;
    S.ssp_type = 'C';
    S.run0 = p->RunType[0];
    S.run1 = p->RunType[1];
    S.run3 = p->RunType[3];
    S.run4 = p->RunType[4];
    S.type2 = p->BeamType[2];
    S.type3 = p->BeamType[3];
    S.o_Rr = N.o_Rr;
    S.o_trc = N.o_trc;
    S.o_brc = N.o_brc;
    S.ntrc = N.ntrc;
    S.nbrc = N.nbrc;
    S.o_sbp = N.o_sbp;
    S.nsbp = N.nsbp;
    S.o_alpha = N.o_alpha;
    S.o_acos = N.o_acos;
    S.o_asin = N.o_asin;
    S.NSz = p->NSz * p->NSx * p->NSy * p->Ntheta;
    S.NRz = p->NRz;
    S.NRr = p->NRr;
    S.NRz_per_range = p->NRz;
    S.Nalpha = p->Nalpha;
    S.iSingle = p->iSingle_alpha;
    S.Delta_r = (real)p->Delta_r;
    S.Dalpha = (real)p->Dalpha;
    S.freq0 = (real)p->freq0;
    S.deltas = (real)p->deltas;
    S.box_r = (real)p->box_x;
    S.box_z = (real)p->box_z;
    S.top = N.top;
    S.bot = N.bot;
    S.rz_sorted = N.rz_sorted;
    S.rz_rdz = N.rz_rdz;
    S.isGaussian = N.isGaussian;
    S.isRayCen = N.isRayCen;
    S.nx_ntheta = p->Ntheta;
    S.nx_nsx = p->NSx;
    S.nx_nsy = p->NSy;
    S.o_beta = N.o_beta;
    return 0;
}

} // namespace bhc3d
