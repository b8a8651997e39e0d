// pack2d.h -- packs a preprocessed 2D problem (bhc_problem2d, doubles) into the
// tracer's blob in the run's real type and fills the device descriptor.
#pragma once
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bhc_hip.h"
#include "core2d.h"

namespace bhc2d {

// Blob layout: staged prefix [z, rho, c, cz, n2, n2z, coef, segr, cmat, czmat, ati, bty, trc, brc, Rz]
// then [Rr, sbp, alpha, cos alpha, sin alpha].
template<typename real> int pack(const bhc_problem2d *p, std::vector<real> &blob, Prob2D<real> &P, std::string &err)
{
    auto fail = [&](const char *m, char c = 0) {
        char b[256];
        snprintf(b, sizeof(b), m, c);
        err = b;
        return -1;
    };
    memset(&P, 0, sizeof(P));
    auto put = [&](const double *src, size_t n) {
        int32_t off = (int32_t)blob.size();
        for(size_t i = 0; i < n; ++i) blob.push_back(src ? (real)src[i] : (real)0);
        return off;
    };
    int N = p->ssp_npts;
    if(N < 2 && p->ssp_type != 'A') return fail("pack: SSP needs at least 2 points");
    P.ssp_type = p->ssp_type;
    P.npts = N;
    P.Nr = p->ssp_nr;
    P.o_z = put(p->ssp_z, N);
    P.o_rho = put(p->ssp_rho, N);
    P.o_c = put(p->ssp_c, 2 * (size_t)N);
    P.o_cz = put(p->ssp_cz, 2 * (size_t)N);
    if(p->ssp_type == 'N') {
        P.o_n2 = put(p->ssp_n2, 2 * (size_t)N);
        P.o_n2z = put(p->ssp_n2z, 2 * (size_t)N);
    }
    if(p->ssp_type == 'S' || p->ssp_type == 'P') P.o_coef = put(p->ssp_coef, 8 * (size_t)N);
    if(p->ssp_type == 'Q') {
        P.o_segr = put(p->ssp_seg_r, p->ssp_nr);
        P.o_cmat = put(p->ssp_cmat, (size_t)N * p->ssp_nr);
        P.o_czmat = put(p->ssp_czmat, (size_t)N * p->ssp_nr);
    }
    // node record: compact unless a boundary is curvilinear or long format
    bool full = p->ati_type[0] == 'C' || p->bty_type[0] == 'C' || p->ati_type[1] == 'L' || p->bty_type[1] == 'L';
    P.node = full ? NODE_FULL : NODE_SHORT;
    auto put_bdry = [&](const bhc_bdry_node2d *nodes, int n) {
        int32_t off = (int32_t)blob.size();
        for(int i = 0; i < n; ++i) {
            const bhc_bdry_node2d &b = nodes[i];
            real rec[NODE_FULL] = {(real)b.x[0],     (real)b.x[1],     (real)b.t[0],     (real)b.t[1],
                                   (real)b.n[0],     (real)b.n[1],     (real)b.Len,      (real)b.kappa,
                                   (real)b.Noden[0], (real)b.Noden[1], (real)b.Nodet[0], (real)b.Nodet[1],
                                   (real)b.cP[0],    (real)b.cP[1],    (real)b.cS[0],    (real)b.cS[1],
                                   (real)b.rho,      (real)0};
            for(int k = 0; k < P.node; ++k) blob.push_back(rec[k]);
        }
        return off;
    };
    if(p->ati_npts < 2 || p->bty_npts < 2) return fail("pack: boundaries need at least 2 nodes");
    P.o_ati = put_bdry(p->ati, p->ati_npts);
    P.o_bty = put_bdry(p->bty, p->bty_npts);
    P.ati_n = p->ati_npts;
    P.bty_n = p->bty_npts;
    // reflection-coefficient tables (read at every 'F' reflection)
    P.o_trc = put(p->trc, 3 * (size_t)p->trc_npts);
    P.o_brc = put(p->brc, 3 * (size_t)p->brc_npts);
    // receiver depths (exact as reals) for the per-step depth windows
    P.o_rzs = (int32_t)blob.size();
    for(int i = 0; i < p->NRz; ++i) blob.push_back((real)p->Rz[i]);
    P.staged_reals = (int32_t)blob.size();
    P.o_Rr = put(p->Rr, p->NRr);
    P.ntrc = p->trc_npts;
    P.nbrc = p->brc_npts;
    P.o_sbp = put(p->sbp, 2 * (size_t)p->NSBPPts);
    P.nsbp = p->NSBPPts;
    P.o_alpha = put(p->alpha, p->Nalpha);
    P.o_acos = put(p->alpha_cos, p->Nalpha);
    P.o_asin = put(p->alpha_sin, p->Nalpha);
    P.run0 = p->RunType[0];
    P.run1 = p->RunType[1];
    P.run3 = p->RunType[3];
    P.run4 = p->RunType[4];
    P.type2 = p->BeamType[2];
    P.type3 = p->BeamType[3];
    P.ati_curv = p->ati_type[0];
    P.bty_curv = p->bty_type[0];
    P.ati_long = p->ati_type[1];
    P.bty_long = p->bty_type[1];
    P.NSz = p->NSz;
    P.NRz = p->NRz;
    P.NRr = p->NRr;
    P.NRz_per_range = p->NRz_per_range;
    P.Nalpha = p->Nalpha;
    P.iSingle = p->iSingle_alpha;
    P.Delta_r = (real)p->Delta_r;
    P.Dalpha = (real)p->Dalpha;
    P.freq0 = (real)p->freq0;
    P.deltas = (real)p->deltas;
    P.box_r = (real)p->box_r;
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
    P.rz_sorted = 1;
    for(int i = 1; i < p->NRz; ++i)
        if(!(p->Rz[i] > p->Rz[i - 1])) P.rz_sorted = 0;
    P.rz_rdz = (P.rz_sorted && p->NRz > 1) ? (double)(p->NRz - 1) / ((double)p->Rz[p->NRz - 1] - (double)p->Rz[0]) : 0.0;
    if(!(P.rz_rdz > 0.0 && P.rz_rdz < 1e300)) P.rz_rdz = 0.0;
    char t0 = p->BeamType[0];
    P.isGaussian = (t0 == 'B' || t0 == 'b');
    P.isRayCen = (t0 == 'g' || t0 == 'b');
    if(!(t0 == 'G' || t0 == 'g' || t0 == 'B' || t0 == 'b' || t0 == '^' || t0 == ' '))
        return fail("pack: beam type '%c' is not a geometric type (Cerveny/SGB are not on this path)", t0);
    return 0;
}


} // namespace bhc2d
