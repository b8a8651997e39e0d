/*
 * ORACLE -- test infrastructure only. Not part of the product.
 *
 * Plain-C restatement of the reference's full-3D ray/beam tracing path
 * (bellhopcxx 3D, CFG = geometric beams, O3D = R3D = true): 3D ray steps with
 * 2x2 p / q matrices and torsion, triangulated boundary grids, depth-only and
 * hexahedral sound speed, Cartesian and ray-centred hat / Gaussian influence
 * with receiver bearings; TL, arrivals (AddArr with azimuths), eigen hits,
 * the ray run type and the eigenray re-trace. This file includes the 2D
 * oracle so that the depth-only SSP evaluation, the crci / spline / PCHIP
 * preprocessing, the halfspace setup, the reflection-table lookup and the
 * AddArr merge rule are the 2D oracle's own.
 *
 * What pins it (tests/test_oracle3d.py): free-space spherical spreading
 * (Cartesian and ray-centred beams), the 2D oracle's field on a depth-only
 * environment, flat and laterally uniform grids, straight rays in ray mode,
 * and the reference's own 3D arrivals fixture
 * (test/individual/analysis_3d/example_analysis_3d.arr.txt), reproduced end
 * to end. Parity unpinned: the torsion update and the p rotation at
 * reflections (second order in those checks).
 *
 * The reference computes vectors and 2x2 matrices with glm (an un-vendored
 * submodule, absent here). Its published semantics are restated below:
 * matrices are column-major m[col][row]; dot(a, b) = (a.x*b.x + a.y*b.y) + a.z*b.z;
 * length(v) = sqrt(dot(v, v)); cross(a, b) = (a.y*b.z - b.y*a.z,
 * a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y); (A*B)[c][r] = A[0][r]*B[c][0] +
 * A[1][r]*B[c][1]; determinant = m00*m11 - m10*m01; v / s divides each
 * component; v *= s multiplies each component.
 *
 * Per-function reference citations are given as (file:line).
 */
#include "bellhop2d.c"

typedef struct { real x, y, z; } v3;
static inline v3 V3(real x, real y, real z) { v3 r = {x, y, z}; return r; }
static inline v3 v3add(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 v3sub(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 v3neg(v3 a) { return V3(-a.x, -a.y, -a.z); }
static inline v3 v3scl(real s, v3 a) { return V3(s * a.x, s * a.y, s * a.z); }
static inline v3 v3divs(v3 a, real s) { return V3(a.x / s, a.y / s, a.z / s); }
static inline real v3dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline real v3len(v3 a) { return RSQRT(v3dot(a, a)); }
static inline v3 v3cross(v3 a, v3 b)
{
    return V3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}

typedef struct { real m[2][2]; } m2; /* m[col][row] */
static inline m2 M2(real a00, real a01, real a10, real a11)
{ /* glm mat2x2(x0, y0, x1, y1): column 0 = (x0, y0) */
    m2 r;
    r.m[0][0] = a00; r.m[0][1] = a01; r.m[1][0] = a10; r.m[1][1] = a11;
    return r;
}
static inline m2 m2mul(m2 a, m2 b)
{
    return M2(a.m[0][0] * b.m[0][0] + a.m[1][0] * b.m[0][1], a.m[0][1] * b.m[0][0] + a.m[1][1] * b.m[0][1],
              a.m[0][0] * b.m[1][0] + a.m[1][0] * b.m[1][1], a.m[0][1] * b.m[1][0] + a.m[1][1] * b.m[1][1]);
}
static inline m2 m2add(m2 a, m2 b)
{
    return M2(a.m[0][0] + b.m[0][0], a.m[0][1] + b.m[0][1], a.m[1][0] + b.m[1][0], a.m[1][1] + b.m[1][1]);
}
static inline m2 m2sub(m2 a, m2 b)
{
    return M2(a.m[0][0] - b.m[0][0], a.m[0][1] - b.m[0][1], a.m[1][0] - b.m[1][0], a.m[1][1] - b.m[1][1]);
}
static inline m2 m2scl(real s, m2 a) { return M2(s * a.m[0][0], s * a.m[0][1], s * a.m[1][0], s * a.m[1][1]); }
static inline m2 m2tr(m2 a) { return M2(a.m[0][0], a.m[1][0], a.m[0][1], a.m[1][1]); }
static inline real m2det(m2 a) { return a.m[0][0] * a.m[1][1] - a.m[1][0] * a.m[0][1]; }

#if OR_IS_FLOAT
#define TRIDIAG_THRESH (3e-3f)
#define RACOS acosf
#define RCOPYSIGN copysignf
#define RFMOD fmodf
#else
#define TRIDIAG_THRESH (3e-6)
#define RACOS acos
#define RCOPYSIGN copysign
#define RFMOD fmod
#endif

/* ------------------------------------------------------------------------- */
/* Preprocessed 3D problem                                                     */
/* ------------------------------------------------------------------------- */
typedef struct {
    v3 x, n1, n2;
    v3 Noden, Nu;                           /* node normal, unit and as (-z_x, -z_y, 1) */
    real phi_xx;                            /* node angle (cells' lower corners; 0 elsewhere) */
    real z_xx, z_xy, z_yy, kappa_xx, kappa_xy, kappa_yy; /* 'C': the cell's curvatures */
    int prov;                               /* 'L': the cell's province (0-based) */
} BNode3;
typedef struct { int nx, ny; char type[2]; BNode3 *bd; int nprov; HS *prov; } Bdry3;

typedef struct {
    Prob P2;            /* depth-only SSP, halfspaces, reflection tables, deltas, sbp, alpha, Sz, Rz, Rr */
    char ssp_type;
    /* 'H': hexahedral grid; z3 / rho3 / nz3: the depth table after SegZToZ */
    int hx, hy, hz;
    real *segx, *segy, *cmat, *czmat;
    int nz3;
    real *z3, *rho3;
    Bdry3 ati, bty;
    int NSx, NSy, Ntheta;
    real *Sx, *Sy;
    v2 *t_rcvr;
    int Nbeta, iSingleB;
    real *beta, Dbeta;
    real box_x, box_y, box_z;
} Prob3;

/* Triangle normals of a rectilinear grid, ComputeBdryTangentNormal (src/module/boundary.hpp:600-650). */
static void bdry3_normals(Bdry3 *B, int isTop)
{
    int ny = B->ny;
    for(int ix = 0; ix < B->nx - 1; ++ix)
        for(int iy = 0; iy < ny - 1; ++iy) {
            v3 p1 = B->bd[ix * ny + iy].x, p2 = B->bd[(ix + 1) * ny + iy].x;
            v3 p3 = B->bd[(ix + 1) * ny + iy + 1].x, p4 = B->bd[ix * ny + iy + 1].x;
            v3 n1 = v3cross(v3sub(p2, p1), v3sub(p3, p1));
            if(isTop) n1 = v3neg(n1);
            B->bd[ix * ny + iy].n1 = v3divs(n1, v3len(n1));
            v3 n2 = v3cross(v3sub(p3, p1), v3sub(p4, p1));
            if(isTop) n2 = v3neg(n2);
            B->bd[ix * ny + iy].n2 = v3divs(n2, v3len(n2));
        }
}

/* Node normals and, for curvilinear grids, the cells' curvatures:
 * ComputeBdryTangentNormal 3D (src/module/boundary.hpp:650-830). The node
 * angles phi_xx exist only at the cells' lower corners; the curvature of a
 * cell in the last row or column reads its neighbour's, which the reference
 * never sets (an uninitialised allocation): here it is 0. */
static void bdry3_curvature(Bdry3 *B)
{
    int nx = B->nx, ny = B->ny;
    BNode3 *b = B->bd;
#define BD(i, j) b[(i) * ny + (j)]
    for(int ix = 0; ix < nx; ++ix)
        for(int iy = 0; iy < ny; ++iy) {
            real mx, my;
            if(ix == 0) mx = (BD(ix + 1, iy).x.z - BD(ix, iy).x.z) / (BD(ix + 1, iy).x.x - BD(ix, iy).x.x);
            else if(ix == nx - 1) mx = (BD(ix, iy).x.z - BD(ix - 1, iy).x.z) / (BD(ix, iy).x.x - BD(ix - 1, iy).x.x);
            else mx = (BD(ix + 1, iy).x.z - BD(ix - 1, iy).x.z) / (BD(ix + 1, iy).x.x - BD(ix - 1, iy).x.x);
            if(iy == 0) my = (BD(ix, iy + 1).x.z - BD(ix, iy).x.z) / (BD(ix, iy + 1).x.y - BD(ix, iy).x.y);
            else if(iy == ny - 1) my = (BD(ix, iy).x.z - BD(ix, iy - 1).x.z) / (BD(ix, iy).x.y - BD(ix, iy - 1).x.y);
            else my = (BD(ix, iy + 1).x.z - BD(ix, iy - 1).x.z) / (BD(ix, iy + 1).x.y - BD(ix, iy - 1).x.y);
            v3 n = V3(-mx, -my, RL(1.0));
            BD(ix, iy).phi_xx = (ix < nx - 1 && iy < ny - 1) ? RATAN2(n.z, n.x) : RL(0.0);
            BD(ix, iy).Nu = n;
            BD(ix, iy).Noden = v3divs(n, v3len(n));
        }
    if(B->type[0] != 'C') return;
    for(int ix = 0; ix < nx - 1; ++ix)
        for(int iy = 0; iy < ny - 1; ++iy) {
            BNode3 *c = &BD(ix, iy);
            c->z_xx = -(BD(ix + 1, iy).Nu.x - c->Nu.x) / (BD(ix + 1, iy).x.x - c->x.x);
            v3 tv = v3sub(BD(ix + 1, iy).x, c->x);
            real Len = RSQRT(SQ(tv.x) + SQ(tv.z));
            c->kappa_xx = (BD(ix + 1, iy).phi_xx - c->phi_xx) / Len;
            c->z_xy = -(BD(ix, iy + 1).Nu.x - c->Nu.x) / (BD(ix, iy + 1).x.y - c->x.y);
            tv = v3sub(BD(ix, iy + 1).x, c->x);
            Len = RSQRT(SQ(tv.y) + SQ(tv.z));
            c->kappa_xy = (BD(ix, iy + 1).phi_xx - c->phi_xx) / Len;
            c->z_yy = -(BD(ix, iy + 1).Nu.y - c->Nu.y) / (BD(ix, iy + 1).x.y - c->x.y);
            c->kappa_yy = (BD(ix, iy + 1).phi_xx - c->phi_xx) / Len;
            Len = v3len(c->Nu);
            c->z_xx /= Len;
            c->z_xy /= Len;
            c->z_yy /= Len;
        }
#undef BD
}

static int bdry3_from_env(const bhc_env2d *e2, const bhc_bdry3d_in *in, real depth, int isTop, Bdry3 *B)
{
    B->type[0] = 'R';
    B->type[1] = 'S';
    B->nprov = 0;
    B->prov = NULL;
    if(in->nx <= 0) { /* Default, src/module/boundary.hpp:56-78 */
        const real big = RL(1.0e25);
        real negtop = isTop ? RL(-1.0) : RL(1.0);
        B->nx = B->ny = 2;
        B->bd = xcalloc(4, sizeof(BNode3));
        B->bd[0].x = V3(-big, -big, depth);
        B->bd[1].x = V3(-big, big, depth);
        B->bd[2].x = V3(big, -big, depth);
        B->bd[3].x = V3(big, big, depth);
        for(int i = 0; i < 4; ++i) B->bd[i].n1 = B->bd[i].n2 = V3(FL(0.0), FL(0.0), negtop);
        return 0;
    }
    if((in->type[0] != 'R' && in->type[0] != 'C') || (in->type[1] != 'S' && in->type[1] != 'L')) return -1;
    B->type[0] = in->type[0];
    B->type[1] = in->type[1];
    B->nx = in->nx;
    B->ny = in->ny;
    B->bd = xcalloc((size_t)B->nx * B->ny, sizeof(BNode3));
    for(int ix = 0; ix < B->nx; ++ix)
        for(int iy = 0; iy < B->ny; ++iy) {
            BNode3 *b = &B->bd[ix * B->ny + iy];
            /* km -> m as Preprocess does (boundary.hpp:503-512) */
            b->x = V3((real)in->x_km[ix] * RL(1000.0), (real)in->y_km[iy] * RL(1000.0), (real)in->z[ix * B->ny + iy]);
        }
    bdry3_normals(B, isTop);
    bdry3_curvature(B);
    if(B->type[1] == 'L') {
        /* provinces (boundary.hpp:152-197): cell -> province, geoacoustics in
         * dB/wavelength converted by crci at z = 1e20 (boundary.hpp:527-537) */
        if(!in->province || !in->prov || in->nprov < 1) return -1;
        B->nprov = in->nprov;
        B->prov = xcalloc(B->nprov, sizeof(HS));
        for(int i = 0; i < B->nprov; ++i) {
            const double *r = in->prov + 5 * i;
            B->prov[i].cP = crci(e2, e2->freq0, RL(1.0e20), (real)r[0], (real)r[3], "W ");
            B->prov[i].cS = crci(e2, e2->freq0, RL(1.0e20), (real)r[1], (real)r[4], "W ");
            B->prov[i].rho = (real)r[2];
        }
        for(int i = 0; i < B->nx * B->ny; ++i) {
            int pv = in->province[i];
            if(pv < 1 || pv > B->nprov) return -1;
            B->bd[i].prov = pv - 1;
        }
    }
    return 0;
}

static void prob3_free(Prob3 *Q)
{
    prob_free(&Q->P2);
    free(Q->segx); free(Q->segy); free(Q->cmat); free(Q->czmat);
    free(Q->z3); free(Q->rho3);
    free(Q->ati.bd); free(Q->bty.bd); free(Q->ati.prov); free(Q->bty.prov);
    free(Q->Sx); free(Q->Sy); free(Q->t_rcvr); free(Q->beta);
}

static int prob3_from_env(const bhc_env3d *e, Prob3 *Q)
{
    memset(Q, 0, sizeof(*Q));
    if(e->ssp_type == 'A' || e->ssp_type == 'Q') return -1;
    /* the depth-only part goes through the 2D preprocessing */
    bhc_env2d e2;
    memset(&e2, 0, sizeof(e2));
    e2.precision = e->precision;
    e2.freq0 = e->freq0;
    e2.ssp_type = (e->ssp_type == 'H') ? 'C' : e->ssp_type;
    e2.atten_unit[0] = e->atten_unit[0];
    e2.atten_unit[1] = e->atten_unit[1];
    e2.ssp_npts = e->ssp_npts;
    e2.ssp_z = e->ssp_z; e2.ssp_alphaR = e->ssp_alphaR; e2.ssp_alphaI = e->ssp_alphaI;
    e2.ssp_rho = e->ssp_rho; e2.ssp_betaR = e->ssp_betaR; e2.ssp_betaI = e->ssp_betaI;
    e2.fg_T = e->fg_T; e2.fg_S = e->fg_S; e2.fg_pH = e->fg_pH; e2.fg_z_bar = e->fg_z_bar;
    e2.top = e->top;
    e2.bot = e->bot;
    e2.trc = e->trc;
    e2.brc = e->brc;
    e2.NSz = e->NSz; e2.NRz = e->NRz; e2.NRr = e->NRr;
    e2.Sz = e->Sz; e2.Rz = e->Rz; e2.Rr_km = e->Rr_km;
    e2.Nalpha = e->Nalpha; e2.iSingle_alpha = e->iSingle_alpha; e2.alpha_deg = e->alpha_deg;
    memcpy(e2.RunType, e->RunType, 7);
    memcpy(e2.BeamType, e->BeamType, 3);
    e2.deltas = e->deltas;
    e2.box_r_km = e->box_x_km;
    e2.box_z = e->box_z;
    e2.NSBPPts = e->NSBPPts;
    e2.sbp = e->sbp;
    e2.NBioLayers = e->NBioLayers;
    e2.bio = e->bio;
    if(prob_from_env(&e2, &Q->P2)) return -1;
    Prob *P = &Q->P2;
    Q->ssp_type = e->ssp_type;
    if(e->ssp_type == 'H') {
        /* src/module/ssp.hpp:144-186, 453-489, 560-581 */
        Q->hx = e->hex_nx; Q->hy = e->hex_ny; Q->hz = e->hex_nz;
        Q->segx = xcalloc(Q->hx, sizeof(real));
        Q->segy = xcalloc(Q->hy, sizeof(real));
        Q->z3 = xcalloc(Q->hz, sizeof(real));
        Q->rho3 = xcalloc(Q->hz, sizeof(real));
        for(int i = 0; i < Q->hx; ++i) Q->segx[i] = (real)e->hex_x_km[i] * FL(1000.0);
        for(int i = 0; i < Q->hy; ++i) Q->segy[i] = (real)e->hex_y_km[i] * FL(1000.0);
        for(int i = 0; i < Q->hz; ++i) {
            Q->z3[i] = (real)e->hex_z[i];
            Q->rho3[i] = (i < P->npts) ? P->rho[i] : (real)NAN;
        }
        size_t n = (size_t)Q->hx * Q->hy * Q->hz;
        Q->cmat = xcalloc(n, sizeof(real));
        Q->czmat = xcalloc((size_t)Q->hx * Q->hy * (Q->hz - 1), sizeof(real));
        for(size_t i = 0; i < n; ++i) Q->cmat[i] = (real)e->hex_c[i];
        for(int ix = 0; ix < Q->hx; ++ix)
            for(int iy = 0; iy < Q->hy; ++iy)
                for(int iz = 1; iz < Q->hz; ++iz) {
                    size_t b = (size_t)ix * Q->hy + iy;
                    Q->czmat[b * (Q->hz - 1) + iz - 1] =
                        (Q->cmat[b * Q->hz + iz] - Q->cmat[b * Q->hz + iz - 1]) / (Q->z3[iz] - Q->z3[iz - 1]);
                }
        Q->nz3 = Q->hz;
    } else {
        Q->nz3 = P->npts;
        Q->z3 = xcalloc(P->npts, sizeof(real));
        Q->rho3 = xcalloc(P->npts, sizeof(real));
        memcpy(Q->z3, P->z, sizeof(real) * P->npts);
        memcpy(Q->rho3, P->rho, sizeof(real) * P->npts);
    }
    if(bdry3_from_env(&e2, &e->ati, P->top.depth, 1, &Q->ati) || bdry3_from_env(&e2, &e->bty, P->bot.depth, 0, &Q->bty)) {
        prob3_free(Q);
        return -1;
    }
    /* sources: sxsy.hpp:105-112 */
    Q->NSx = e->NSx;
    Q->NSy = e->NSy;
    Q->Sx = xcalloc(Q->NSx, sizeof(real));
    Q->Sy = xcalloc(Q->NSy, sizeof(real));
    for(int i = 0; i < Q->NSx; ++i) Q->Sx[i] = (real)e->Sx_km[i] * FL(1000.0);
    for(int i = 0; i < Q->NSy; ++i) Q->Sy[i] = (real)e->Sy_km[i] * FL(1000.0);
    /* bearings: rcvrbearings.hpp:98-110 */
    Q->Ntheta = e->Ntheta;
    Q->t_rcvr = xcalloc(Q->Ntheta, sizeof(v2));
    for(int i = 0; i < Q->Ntheta; ++i) {
        real th = DegRad * (real)e->theta_deg[i];
        Q->t_rcvr[i] = V2(RCOS(th), RSIN(th));
    }
    /* beta: rayangles.hpp:157-189; an Nx2D run that is not a ray run traces
     * its beams on the receiver bearings (rayangles.hpp:163-178) */
    const double *bdeg = e->beta_deg;
    Q->Nbeta = e->Nbeta;
    if(e->RunType[5] == '2' && e->RunType[0] != 'R') {
        bdeg = e->theta_deg;
        Q->Nbeta = e->Ntheta;
    }
    Q->iSingleB = e->iSingle_beta;
    Q->beta = xcalloc(Q->Nbeta, sizeof(real));
    for(int i = 0; i < Q->Nbeta; ++i) Q->beta[i] = (real)bdeg[i] * DegRad;
    Q->Dbeta = FL(0.0);
    if(Q->Nbeta != 1) Q->Dbeta = (Q->beta[Q->Nbeta - 1] - Q->beta[0]) / (Q->Nbeta - 1);
    /* beam box: beaminfo.hpp:226-230 */
    Q->box_x = (real)e->box_x_km * FL(1000.0);
    Q->box_y = (real)e->box_y_km * FL(1000.0);
    Q->box_z = (real)e->box_z;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* SSP, src/ssp.hpp:220-307 (Hexahedral), 377-465 (EvaluateSSP, 3D)             */
/* ------------------------------------------------------------------------- */
typedef struct { v3 gradc; cpx ccpx; real rho, cxx, cyy, czz, cxy, cxz, cyz; } SSPOut3;
typedef struct { int x, y, z, r; } Seg3;

static void Hexahedral(v3 x, v3 t, SSPOut3 *o, const Prob3 *Q, Seg3 *iSeg, ErrState *err)
{
    int nx = Q->hx, ny = Q->hy, nz = Q->hz;
    if(x.x < Q->segx[0] || x.x > Q->segx[nx - 1] || x.y < Q->segy[0] || x.y > Q->segy[ny - 1])
        RunError(err, BHC_ERR_OUTSIDE_SSP);
    UpdateSSPSegment(x.x, t.x, Q->segx, nx, &iSeg->x);
    UpdateSSPSegment(x.y, t.y, Q->segy, ny, &iSeg->y);
    UpdateSSPSegment(x.z, t.z, Q->z3, nz, &iSeg->z);
    int ix = iSeg->x, iy = iSeg->y, iz = iSeg->z;
    real cz11 = Q->czmat[((size_t)ix * ny + iy) * (nz - 1) + iz];
    real cz12 = Q->czmat[((size_t)(ix + 1) * ny + iy) * (nz - 1) + iz];
    real cz21 = Q->czmat[((size_t)ix * ny + iy + 1) * (nz - 1) + iz];
    real cz22 = Q->czmat[((size_t)(ix + 1) * ny + iy + 1) * (nz - 1) + iz];
    real s3 = x.z - Q->z3[iz];
    real c11 = Q->cmat[((size_t)ix * ny + iy) * nz + iz] + s3 * cz11;
    real c21 = Q->cmat[((size_t)(ix + 1) * ny + iy) * nz + iz] + s3 * cz12;
    real c12 = Q->cmat[((size_t)ix * ny + iy + 1) * nz + iz] + s3 * cz21;
    real c22 = Q->cmat[((size_t)(ix + 1) * ny + iy + 1) * nz + iz] + s3 * cz22;
    real s1 = (x.x - Q->segx[ix]) / (Q->segx[ix + 1] - Q->segx[ix]);
    s1 = rmax(rmin(s1, RL(1.0)), RL(0.0));
    real s2 = (x.y - Q->segy[iy]) / (Q->segy[iy + 1] - Q->segy[iy]);
    s2 = rmax(rmin(s2, RL(1.0)), RL(0.0));
    real c1 = c11 + s1 * (c21 - c11);
    real c2 = c12 + s1 * (c22 - c12);
    real cy = (c2 - c1) / (Q->segy[iy + 1] - Q->segy[iy]);
    c1 = c11 + s2 * (c12 - c11);
    c2 = c21 + s2 * (c22 - c21);
    real c = c1 + s1 * (c2 - c1);
    /* The volume-attenuation term interpolates the depth profile that
     * SegZToZ (src/module/ssp.hpp:568-573) has overwritten with real NaNs
     * (imaginary part 0), so its imaginary part is +0. */
    o->ccpx = MKC(c, RL(0.0));
    real cx = (c2 - c1) / (Q->segx[ix + 1] - Q->segx[ix]);
    real cz1 = cz11 + s2 * (cz21 - cz11);
    real cz2 = cz12 + s2 * (cz22 - cz12);
    real cz = cz1 + s1 * (cz2 - cz1);
    o->gradc = V3(cx, cy, cz);
    o->cxx = o->cyy = o->czz = o->cxy = o->cxz = o->cyz = RL(0.0);
    real w = (x.z - Q->z3[iz]) / (Q->z3[iz + 1] - Q->z3[iz]); /* LinInterpDensity */
    o->rho = (RL(1.0) - w) * Q->rho3[iz] + w * Q->rho3[iz + 1];
}

static void EvaluateSSP3(v3 x, v3 t, SSPOut3 *o, const Prob3 *Q, Seg3 *iSeg, ErrState *err)
{
    if(Q->ssp_type == 'H') {
        Hexahedral(x, t, o, Q, iSeg, err);
        return;
    }
    SSPOut o2;
    SegState s2;
    s2.r = iSeg->r;
    s2.z = iSeg->z;
    EvaluateSSP(V2(RL(0.0), x.z), V2(RL(0.0), t.z), &o2, &Q->P2, &s2, err);
    iSeg->z = s2.z;
    o->gradc = V3(RL(0.0), RL(0.0), o2.gradc.y);
    o->cxx = o->cyy = o->cxy = o->cxz = o->cyz = RL(0.0);
    o->czz = o2.czz;
    o->ccpx = o2.ccpx;
    o->rho = o2.rho;
}

/* ------------------------------------------------------------------------- */
/* Boundaries, src/boundary.hpp:66-168 (3D branch)                             */
/* ------------------------------------------------------------------------- */
typedef struct {
    int ix, iy;
    real xmin, xmax, ymin, ymax;
    v3 x, xmid, n;
    int onEdge, side, justSteppedTo, outgoingSide;
    int prov; /* 'L' grids: the province of the current cell (CopyHSInfo, boundary.hpp:134-140) */
} BState3;

static void GetBdrySeg3(v3 x, v3 t, BState3 *b, const Bdry3 *B, int isTop, int isInit, ErrState *err)
{
    int nx = B->nx, ny = B->ny;
    const BNode3 *bd = B->bd;
    b->ix = imin(imax(b->ix, 0), nx - 2);
    b->iy = imin(imax(b->iy, 0), ny - 2);
    if(t.x >= FL(0.0)) {
        while(b->ix >= 0 && bd[b->ix * ny].x.x > x.x) --b->ix;
        while(b->ix >= 0 && b->ix < nx - 1 && bd[(b->ix + 1) * ny].x.x <= x.x) ++b->ix;
    } else {
        while(b->ix < nx - 1 && bd[(b->ix + 1) * ny].x.x < x.x) ++b->ix;
        while(b->ix >= 0 && b->ix < nx - 1 && bd[b->ix * ny].x.x >= x.x) --b->ix;
    }
    if(t.y >= FL(0.0)) {
        while(b->iy >= 0 && bd[b->iy].x.y > x.y) --b->iy;
        while(b->iy >= 0 && b->iy < ny - 1 && bd[b->iy + 1].x.y <= x.y) ++b->iy;
    } else {
        while(b->iy < ny - 1 && bd[b->iy + 1].x.y < x.y) ++b->iy;
        while(b->iy >= 0 && b->iy < ny - 1 && bd[b->iy].x.y >= x.y) --b->iy;
    }
    if(b->ix == -1 && bd[0].x.x == x.x) b->ix = 0;
    if(b->ix == nx - 1 && bd[(nx - 1) * ny].x.x == x.x) b->ix = nx - 2;
    if(b->iy == -1 && bd[0].x.y == x.y) b->iy = 0;
    if(b->iy == ny - 1 && bd[ny - 1].x.y == x.y) b->iy = ny - 2;
    if(b->ix < 0 || b->ix >= nx - 1 || b->iy < 0 || b->iy >= ny - 1) {
        RunWarning(err, isTop ? BHC_WARN_OUTSIDE_ALTIMETRY : BHC_WARN_OUTSIDE_BATHYMETRY);
        b->ix = imin(imax(b->ix, 0), nx - 2);
        b->iy = imin(imax(b->iy, 0), ny - 2);
    }
    b->xmin = bd[b->ix * ny].x.x;
    b->xmax = bd[(b->ix + 1) * ny].x.x;
    b->ymin = bd[b->iy].x.y;
    b->ymax = bd[b->iy + 1].x.y;
    b->x = bd[b->ix * ny + b->iy].x;
    b->xmid = v3scl(RL(0.5), v3add(b->x, bd[(b->ix + 1) * ny + (b->iy + 1)].x));
    if(B->type[1] == 'L') b->prov = bd[b->ix * ny + b->iy].prov;
    /* (x + x11) * 0.5 is componentwise */
    v2 tri_n = V2(-(b->ymax - b->ymin), b->xmax - b->xmin);
    tri_n = vdivs(tri_n, vlen(tri_n));
    v2 temp = V2(x.x - b->xmid.x, x.y - b->xmid.y);
    real over = vdot(temp, tri_n);
    b->onEdge = RABS(over) < TRIDIAG_THRESH;
    if(!isInit && b->justSteppedTo) b->side = b->outgoingSide;
    else if(!isInit && b->onEdge) b->side = vdot(V2(t.x, t.y), tri_n) >= RL(0.0);
    else b->side = over >= RL(0.0);
    b->justSteppedTo = 0;
    b->n = b->side ? bd[b->ix * ny + b->iy].n2 : bd[b->ix * ny + b->iy].n1;
    if(!isfinite(b->x.z) || !isfinite(b->n.x) || !isfinite(b->n.y) || !isfinite(b->n.z))
        RunError(err, BHC_ERR_BOUNDARY_SEG_CONTAINS_NAN);
}

/* ------------------------------------------------------------------------- */
/* Step, src/step.hpp (3D instantiation)                                       */
/* ------------------------------------------------------------------------- */
static real v3get(v3 a, int d) { return d == 0 ? a.x : (d == 1 ? a.y : a.z); }
static void v3set(v3 *a, int d, real v) { if(d == 0) a->x = v; else if(d == 1) a->y = v; else a->z = v; }

typedef struct {
    const Prob3 *Q;
    v3 xs;
} Ctx3;

static void DepthInterfaceCrossing3(int stepTo, real *hInt, v3 *x, v3 x0, v3 urayt, const Seg3 *iSeg0, const Prob3 *Q,
                                    int *snap)
{ /* step.hpp:41-80; snap: the snapDim of OceanToRayX (null when not stepping) */
    if(!stepTo) *hInt = REAL_MAX;
    if(RABS(urayt.z) > REAL_EPSILON) {
        const real *z = Q->z3;
        int iz = iSeg0->z;
        if(z[iz] > x->z && iz > 0) {
            *hInt = (z[iz] - x0.z) / urayt.z;
            if(stepTo) {
                *x = v3add(x0, v3scl(*hInt, urayt));
                x->z = z[iz];
                if(snap) *snap = 2;
            }
        } else if(z[iz + 1] < x->z && iz + 1 < Q->nz3 - 1) {
            *hInt = (z[iz + 1] - x0.z) / urayt.z;
            if(stepTo) {
                *x = v3add(x0, v3scl(*hInt, urayt));
                x->z = z[iz + 1];
                if(snap) *snap = 2;
            }
        }
    }
}

static void BeamBoxCrossing3(int stepTo, int dim, real *hBox, v3 *x, v3 x0, v3 urayt, const Prob3 *Q, v3 xs,
                             int *snap)
{ /* step.hpp:85-111, common_run.hpp:33-47 */
    if(!stepTo) *hBox = REAL_MAX;
    v3 ctr = V3(xs.x, xs.y, RL(0.0));
    real box = dim == 0 ? Q->box_x : (dim == 1 ? Q->box_y : Q->box_z);
    if(RABS(v3get(*x, dim) - v3get(ctr, dim)) >= box) {
        real d0 = v3get(x0, dim) - v3get(ctr, dim);
        *hBox = (box - RABS(d0)) / RABS(v3get(urayt, dim));
        if(stepTo) {
            *x = v3add(x0, v3scl(*hBox, urayt));
            v3set(x, dim, v3get(ctr, dim) + RCOPYSIGN(box, d0));
            if(snap) *snap = dim;
        }
    }
}

static void TopBotCrossing3(int stepTo, real *h, const BState3 *bd, v3 *x, v3 x0, v3 urayt, int *refl, int *snap)
{ /* step.hpp:116-151 */
    if(!stepTo) *h = REAL_MAX;
    v3 d = v3sub(*x, bd->x), d0 = v3sub(x0, bd->x);
    real w = v3dot(bd->n, d);
    if(stepTo ? (w > -INFINITESIMAL_STEP_SIZE) : (w >= RL(0.0))) {
        *h = -v3dot(d0, bd->n) / v3dot(urayt, bd->n);
        if(stepTo) {
            *x = v3add(x0, v3scl(*h, urayt));
            if(RABS(bd->n.x) < REAL_EPSILON && RABS(bd->n.y) < REAL_EPSILON) x->z = bd->x.z;
            if(snap) *snap = 2;
        }
        *refl = 1;
    } else {
        *refl = 0;
    }
}

static void TopBotSegCrossing3(int stepTo, int isY, real *hSeg, const BState3 *top, const BState3 *bot, int iSeg,
                               v3 *x, v3 x0, v3 urayt, const Prob3 *Q, int *topRefl, int *botRefl, int *snap)
{ /* step.hpp:157-207 */
    real lmin = rmax(isY ? top->ymin : top->xmin, isY ? bot->ymin : bot->xmin);
    real lmax = rmin(isY ? top->ymax : top->xmax, isY ? bot->ymax : bot->xmax);
    if(Q->ssp_type == 'H') {
        const real *sw = isY ? Q->segy : Q->segx;
        lmin = rmax(lmin, sw[iSeg]);
        lmax = rmin(lmax, sw[iSeg + 1]);
    }
    real xw = isY ? x->y : x->x, x0w = isY ? x0.y : x0.x, uw = isY ? urayt.y : urayt.x;
    if(!stepTo) *hSeg = REAL_MAX;
    if(RABS(uw) > REAL_EPSILON) {
        if(xw < lmin) {
            *hSeg = -(x0w - lmin) / uw;
            if(stepTo) {
                *x = v3add(x0, v3scl(*hSeg, urayt));
                if(isY) x->y = lmin; else x->x = lmin;
                if(snap) *snap = isY;
                *topRefl = *botRefl = 0;
            }
        } else if(xw > lmax) {
            *hSeg = -(x0w - lmax) / uw;
            if(stepTo) {
                *x = v3add(x0, v3scl(*hSeg, urayt));
                if(isY) x->y = lmax; else x->x = lmax;
                if(snap) *snap = isY;
                *topRefl = *botRefl = 0;
            }
        }
    }
}

static void TriDiagCrossing(int stepTo, real *h, BState3 *bd, v3 *x, v3 x0, v3 urayt, int *topRefl, int *botRefl,
                            ErrState *err, int *snap)
{ /* step.hpp:212-260 */
    if(!stepTo) *h = REAL_MAX;
    if(bd->onEdge) return;
    v3 d = v3sub(*x, bd->xmid), d0 = v3sub(x0, bd->xmid);
    v3 tri_n = V3(-(bd->ymax - bd->ymin), bd->xmax - bd->xmin, RL(0.0));
    tri_n = v3divs(tri_n, v3len(tri_n));
    real over = v3dot(d, tri_n);
    int newSide = over >= RL(0.0);
    if(newSide == bd->side) return;
    if(RABS(over) < TRIDIAG_THRESH) return;
    real hnew = -v3dot(d0, tri_n) / v3dot(urayt, tri_n);
    if(hnew < RL(0.0)) {
        RunWarning(err, BHC_WARN_TRIDIAG_H_NEGATIVE);
        *h = RL(0.0);
    } else {
        if(stepTo && hnew >= *h) RunWarning(err, BHC_WARN_TRIDIAG_H_GROWING);
        *h = hnew;
    }
    if(stepTo) {
        *x = v3add(x0, v3scl(*h, urayt));
        bd->justSteppedTo = 1;
        bd->outgoingSide = newSide;
        if(snap) *snap = -2;
        *topRefl = *botRefl = 0;
    }
}

static void ReduceStep3(v3 x0, v3 urayt, const Seg3 *iSeg0, BState3 *top, BState3 *bot, const Prob3 *Q, v3 xs,
                        ErrState *err, real *h, int *iSmallStepCtr)
{ /* step.hpp:271-337 */
    real hInt, hBx, hBy, hBz, hTop, hBot, hxSeg, hySeg, hTD, hBD;
    int dummy;
    v3 x = v3add(x0, v3scl(*h, urayt));
    DepthInterfaceCrossing3(0, &hInt, &x, x0, urayt, iSeg0, Q, NULL);
    BeamBoxCrossing3(0, 0, &hBx, &x, x0, urayt, Q, xs, NULL);
    BeamBoxCrossing3(0, 1, &hBy, &x, x0, urayt, Q, xs, NULL);
    BeamBoxCrossing3(0, 2, &hBz, &x, x0, urayt, Q, xs, NULL);
    TopBotCrossing3(0, &hTop, top, &x, x0, urayt, &dummy, NULL);
    TopBotCrossing3(0, &hBot, bot, &x, x0, urayt, &dummy, NULL);
    TopBotSegCrossing3(0, 0, &hxSeg, top, bot, iSeg0->x, &x, x0, urayt, Q, &dummy, &dummy, NULL);
    TopBotSegCrossing3(0, 1, &hySeg, top, bot, iSeg0->y, &x, x0, urayt, Q, &dummy, &dummy, NULL);
    TriDiagCrossing(0, &hTD, top, &x, x0, urayt, &dummy, &dummy, err, NULL);
    TriDiagCrossing(0, &hBD, bot, &x, x0, urayt, &dummy, &dummy, err, NULL);
    *h = rmin(rmin(rmin(*h, hInt), rmin(rmin(rmin(hBx, hBy), hBz), rmin(hTop, hBot))),
              rmin(rmin(hxSeg, hySeg), rmin(hTD, hBD)));
    if(*h < RL(-1e-4)) RunWarning(err, BHC_WARN_STEP_NEGATIVE_H);
    if(*h < INFINITESIMAL_STEP_SIZE * Q->P2.deltas) {
        *h = INFINITESIMAL_STEP_SIZE * Q->P2.deltas;
        ++*iSmallStepCtr;
    } else {
        *iSmallStepCtr = 0;
    }
}

static void StepToBdry3(v3 x0, v3 *x2, v3 urayt, real *h, int *topRefl, int *botRefl, const Seg3 *iSeg0,
                        BState3 *top, BState3 *bot, const Prob3 *Q, v3 xs, ErrState *err, int *snapDim)
{ /* step.hpp:342-420; snapDim: OceanToRayX's snap choice (common_run.hpp:243-252) */
    int snap = -1;
    *x2 = v3add(x0, v3scl(*h, urayt));
    DepthInterfaceCrossing3(1, h, x2, x0, urayt, iSeg0, Q, &snap);
    BeamBoxCrossing3(1, 0, h, x2, x0, urayt, Q, xs, &snap);
    BeamBoxCrossing3(1, 1, h, x2, x0, urayt, Q, xs, &snap);
    BeamBoxCrossing3(1, 2, h, x2, x0, urayt, Q, xs, &snap);
    TopBotCrossing3(1, h, top, x2, x0, urayt, topRefl, &snap);
    TopBotCrossing3(1, h, bot, x2, x0, urayt, botRefl, &snap);
    if(*botRefl) *topRefl = 0;
    TopBotSegCrossing3(1, 0, h, top, bot, iSeg0->x, x2, x0, urayt, Q, topRefl, botRefl, &snap);
    TopBotSegCrossing3(1, 1, h, top, bot, iSeg0->y, x2, x0, urayt, Q, topRefl, botRefl, &snap);
    TriDiagCrossing(1, h, top, x2, x0, urayt, topRefl, botRefl, err, &snap);
    TriDiagCrossing(1, h, bot, x2, x0, urayt, topRefl, botRefl, err, &snap);
    if(*h < INFINITESIMAL_STEP_SIZE * Q->P2.deltas) {
        *h = INFINITESIMAL_STEP_SIZE * Q->P2.deltas;
        *x2 = v3add(x0, v3scl(*h, urayt));
        snap = -1;
        v3 d = v3sub(*x2, top->x);
        if(v3dot(top->n, d) > REAL_EPSILON) {
            *topRefl = 1;
            top->justSteppedTo = 0;
            bot->justSteppedTo = 0;
            snap = 2;
        } else {
            *topRefl = 0;
        }
        d = v3sub(*x2, bot->x);
        if(v3dot(bot->n, d) > REAL_EPSILON) {
            *botRefl = 1;
            *topRefl = 0;
            top->justSteppedTo = 0;
            bot->justSteppedTo = 0;
            snap = 2;
        } else {
            *botRefl = 0;
        }
    }
    if(snapDim) *snapDim = snap;
}

/* Ray state, rayPt<true> (include/bhc/structs.hpp:397-424). */
typedef struct {
    m2 p, q;
    real phi;
    int NumTopBnc, NumBotBnc;
    v3 x, t;
    real c, Amp, Phase;
    cpx tau;
} RayPt3;

static void RayNormalImpl(v3 t, real phi, int ignorephi0, real c, v3 *e1, v3 *e2)
{ /* common_run.hpp:144-215 */
    real rl = RSQRT(t.x * t.x + t.y * t.y);
    if(RABS(phi) >= REAL_EPSILON || ignorephi0) {
        real cosphi = RCOS(phi), sinphi = RSIN(phi);
        e1->x = (c * t.x * t.z * cosphi + t.y * sinphi) / rl;
        e1->y = (c * t.y * t.z * cosphi - t.x * sinphi) / rl;
        e1->z = -c * rl * cosphi;
        e2->x = (c * t.x * t.z * sinphi - t.y * cosphi) / rl;
        e2->y = (c * t.y * t.z * sinphi + t.x * cosphi) / rl;
        e2->z = -c * rl * sinphi;
    } else {
        *e1 = V3(c * t.x * t.z / rl, c * t.y * t.z / rl, -c * rl);
        *e2 = V3(-t.y / rl, t.x / rl, RL(0.0));
    }
}

typedef struct { real cnn, cmn, cmm; } Part3;
typedef struct { real phi; m2 p, q; } PQ3;

static Part3 Get_c_partials3(const RayPt3 *ray, const SSPOut3 *o)
{ /* step.hpp:425-448 */
    v3 e1, e2;
    RayNormalImpl(ray->t, ray->phi, 0, creal(o->ccpx), &e1, &e2);
    Part3 P;
    P.cnn = o->cxx * SQ(e1.x) + o->cyy * SQ(e1.y) + o->czz * SQ(e1.z) + FL(2.0) * o->cxy * e1.x * e1.y
        + FL(2.0) * o->cxz * e1.x * e1.z + FL(2.0) * o->cyz * e1.y * e1.z;
    P.cmn = o->cxx * e1.x * e2.x + o->cyy * e1.y * e2.y + o->czz * e1.z * e2.z + o->cxy * (e1.x * e2.y + e2.x * e1.y)
        + o->cxz * (e1.x * e2.z + e2.x * e1.z) + o->cyz * (e1.y * e2.z + e2.y * e1.z);
    P.cmm = o->cxx * SQ(e2.x) + o->cyy * SQ(e2.y) + o->czz * SQ(e2.z) + FL(2.0) * o->cxy * e2.x * e2.y
        + FL(2.0) * o->cxz * e2.x * e2.z + FL(2.0) * o->cyz * e2.y * e2.z;
    return P;
}

static PQ3 ComputeDeltaPQ3(const RayPt3 *ray, const SSPOut3 *o, Part3 part)
{ /* step.hpp:450-473 */
    PQ3 pq;
    real c = creal(o->ccpx);
    pq.phi = (FL(1.0) / c) * ray->t.z * (ray->t.y * o->gradc.x - ray->t.x * o->gradc.y)
        / (SQ(ray->t.x) + SQ(ray->t.y));
    m2 cm = M2(part.cnn, part.cmn, part.cmn, part.cmm);
    cm = m2scl(-RL(1.0) / SQ(c), cm);
    pq.p = m2mul(ray->q, m2tr(cm));
    pq.q = m2scl(c, ray->p);
    return pq;
}

static void UpdateRayPQ3(RayPt3 *r1, const RayPt3 *r0, real h, const PQ3 *pq)
{ /* step.hpp:475-481 */
    real phi = r0->phi + h * pq->phi;
    m2 p = m2add(r0->p, m2scl(h, pq->p));
    m2 q = m2add(r0->q, m2scl(h, pq->q));
    r1->phi = phi;
    r1->p = p;
    r1->q = q;
}

static void CurvatureCorrection3D(RayPt3 *ray, m2 DMat, real Tg, real Th, real cn1jump, real cn2jump, real csjump,
                                  v3 rayn1, v3 rayn2, v3 e1, v3 e2, int reflect)
{ /* step.hpp:486-572 */
    real rm = Tg / Th;
    real csq = SQ(ray->c);
    m2 r;
    r.m[0][0] = FL(2.0) / csq * DMat.m[0][0] / Th + rm * (FL(2.0) * cn1jump - rm * csjump) / csq;
    r.m[0][1] = FL(2.0) / ray->c * DMat.m[1][0] * RCOPYSIGN(RL(1.0), -Th) + rm * cn2jump / csq;
    r.m[1][0] = r.m[0][1];
    r.m[1][1] = FL(2.0) * DMat.m[1][1] * Th;
    if(reflect) r.m[1][0] = -r.m[1][0];
    else r = m2scl(RL(-1.0), r);
    int noChange = 0;
    if(reflect)
        noChange = RABS(r.m[0][0]) < REAL_EPSILON && RABS(r.m[0][1]) < REAL_EPSILON && RABS(r.m[1][1]) < REAL_EPSILON;
    if(!noChange) {
        m2 R;
        R.m[0][0] = v3dot(rayn1, e1);
        R.m[1][0] = v3dot(rayn1, e2);
        R.m[0][1] = -R.m[1][0];
        R.m[1][1] = v3dot(rayn2, e2);
        m2 p_in = m2mul(R, ray->p);
        m2 q_in = m2mul(R, ray->q);
        m2 p_out = m2add(p_in, m2mul(r, q_in));
        ray->p = m2mul(m2tr(R), p_out);
    }
    if(reflect) {
        real d = v3dot(rayn1, e1);
        ray->phi += FL(2.0) * RACOS(rmax(rmin(d, RL(1.0)), RL(-1.0)));
    }
}

static void CalcTangent_Normals(const RayPt3 *ray, real c, v3 nBdry, v3 *rayt, v3 *rayn1, v3 *rayn2, real sgn)
{ /* step.hpp:574-584 */
    *rayt = v3scl(c, ray->t);
    *rayn2 = v3scl(sgn, v3cross(*rayt, nBdry));
    *rayn2 = v3divs(*rayn2, v3len(*rayn2));
    *rayn1 = v3neg(v3cross(*rayt, *rayn2));
}

static void CurvatureCorrection3(RayPt3 *ray, v3 gradcjump, const Seg3 *iSeg, const Seg3 *iSeg0)
{ /* step.hpp:589-627 (3D) */
    v3 nB = V3(RL(0.0), RL(0.0), RL(0.0));
    if(iSeg->z != iSeg0->z) nB.z = -RCOPYSIGN(RL(1.0), ray->t.z);
    else if(iSeg->x != iSeg0->x) nB.x = -RCOPYSIGN(RL(1.0), ray->t.x);
    else nB.y = -RCOPYSIGN(RL(1.0), ray->t.y);
    real Th = v3dot(ray->t, nB);
    v3 tB = v3sub(ray->t, v3scl(Th, nB));
    tB = v3divs(tB, v3len(tB));
    real Tg = v3dot(ray->t, tB);
    v3 rayt, rayn1, rayn2, e1, e2;
    CalcTangent_Normals(ray, ray->c, nB, &rayt, &rayn1, &rayn2, RL(1.0));
    RayNormalImpl(rayt, ray->phi, 1, RL(1.0), &e1, &e2);
    real cn1 = v3dot(gradcjump, rayn1), cn2 = v3dot(gradcjump, rayn2), cs = v3dot(gradcjump, rayt);
    CurvatureCorrection3D(ray, M2(RL(0.0), RL(0.0), RL(0.0), RL(0.0)), Tg, Th, cn1, cn2, cs, rayn1, rayn2, e1, e2, 0);
}

typedef struct {
    Seg3 iSeg;
    BState3 top, bot;
    v3 xs;
    int iSmallStepCtr;
} TraceState3;

static void Step3(RayPt3 ray0, RayPt3 *ray2, TraceState3 *S, const Prob3 *Q, ErrState *err, int *topRefl,
                  int *botRefl)
{ /* step.hpp:654-811 (3D) */
    RayPt3 ray1;
    SSPOut3 o0, o1, o2;
    EvaluateSSP3(ray0.x, ray0.t, &o0, Q, &S->iSeg, err);
    Part3 part0 = Get_c_partials3(&ray0, &o0);
    PQ3 pq0 = ComputeDeltaPQ3(&ray0, &o0, part0);
    Seg3 iSeg0 = S->iSeg;
    real c0 = creal(o0.ccpx);
    real csq0 = SQ(c0);
    v3 urayt0 = v3scl(c0, ray0.t);
    real h = Q->P2.deltas;
    ReduceStep3(ray0.x, urayt0, &iSeg0, &S->top, &S->bot, Q, S->xs, err, &h, &S->iSmallStepCtr);
    real halfh = FL(0.5) * h;
    ray1.x = v3add(ray0.x, v3scl(halfh, urayt0));
    ray1.t = v3sub(ray0.t, v3divs(v3scl(halfh, o0.gradc), csq0));
    UpdateRayPQ3(&ray1, &ray0, halfh, &pq0);
    EvaluateSSP3(ray1.x, ray1.t, &o1, Q, &S->iSeg, err);
    Part3 part1 = Get_c_partials3(&ray1, &o1);
    PQ3 pq1 = ComputeDeltaPQ3(&ray1, &o1, part1);
    real c1 = creal(o1.ccpx);
    real csq1 = SQ(c1);
    v3 urayt1 = v3scl(c1, ray1.t);
    ReduceStep3(ray0.x, urayt1, &iSeg0, &S->top, &S->bot, Q, S->xs, err, &h, &S->iSmallStepCtr);
    real w1 = h / (RL(2.0) * halfh);
    real w0 = RL(1.0) - w1;
    v3 urayt2 = v3add(v3scl(w0, urayt0), v3scl(w1, urayt1));
    v3 x2;
    StepToBdry3(ray0.x, &x2, urayt2, &h, topRefl, botRefl, &iSeg0, &S->top, &S->bot, Q, S->xs, err, NULL);
    ray2->x = x2;
    real hw0 = h * w0, hw1 = h * w1;
    ray2->t = v3sub(v3sub(ray0.t, v3divs(v3scl(hw0, o0.gradc), csq0)), v3divs(v3scl(hw1, o1.gradc), csq1));
    ray2->tau = ray0.tau + hw0 / o0.ccpx + hw1 / o1.ccpx;
    UpdateRayPQ3(ray2, &ray0, hw0, &pq0);
    UpdateRayPQ3(ray2, ray2, hw1, &pq1);
    ray2->Amp = ray0.Amp;
    ray2->Phase = ray0.Phase;
    ray2->NumTopBnc = ray0.NumTopBnc;
    ray2->NumBotBnc = ray0.NumBotBnc;
    EvaluateSSP3(ray2->x, ray2->t, &o2, Q, &S->iSeg, err);
    ray2->c = creal(o2.ccpx);
    if(S->iSeg.z != iSeg0.z || S->iSeg.x != iSeg0.x || S->iSeg.y != iSeg0.y)
        CurvatureCorrection3(ray2, v3sub(o2.gradc, o0.gradc), &S->iSeg, &iSeg0);
}

/* ------------------------------------------------------------------------- */
/* Reflect, src/reflect.hpp:125-452 (3D)                                       */
/* ------------------------------------------------------------------------- */
/* The reflecting boundary's data at the ray's ocean point x (trace.hpp:309-360,
 * O3D): the halfspace, its geoacoustics replaced by the cell's province on 'L'
 * grids; on 'C' grids the node normals interpolated bilinearly over the cell
 * (not renormalised) and the cell's curvatures, else the triangle's normal and
 * zero curvature. */
typedef struct { HS hs; v3 n; real z_xx, z_xy, z_yy, kappa_xx, kappa_xy, kappa_yy; } ReflBdry3;

static ReflBdry3 refl_bdry3(const Bdry3 *B, const BState3 *b, const HS *hs0, v3 x)
{
    ReflBdry3 r;
    r.hs = *hs0;
    if(B->type[1] == 'L') {
        const HS *pv = &B->prov[b->prov];
        r.hs.cP = pv->cP;
        r.hs.cS = pv->cS;
        r.hs.rho = pv->rho;
    }
    r.n = b->n;
    r.z_xx = r.z_xy = r.z_yy = r.kappa_xx = r.kappa_xy = r.kappa_yy = FL(0.0);
    if(B->type[0] == 'C') {
        int ny = B->ny;
        const BNode3 *bd00 = &B->bd[b->ix * ny + b->iy], *bd01 = &B->bd[b->ix * ny + b->iy + 1];
        const BNode3 *bd10 = &B->bd[(b->ix + 1) * ny + b->iy], *bd11 = &B->bd[(b->ix + 1) * ny + b->iy + 1];
        real s1 = (x.x - b->x.x) / (b->xmax - b->xmin);
        real s2 = (x.y - b->x.y) / (b->ymax - b->ymin);
        real m1 = FL(1.0) - s1, m2 = FL(1.0) - s2;
        r.n = v3add(v3add(v3add(v3scl(m2, v3scl(m1, bd00->Noden)), v3scl(m2, v3scl(s1, bd10->Noden))),
                          v3scl(s2, v3scl(s1, bd11->Noden))),
                    v3scl(s2, v3scl(m1, bd01->Noden)));
        r.z_xx = bd00->z_xx;
        r.z_xy = bd00->z_xy;
        r.z_yy = bd00->z_yy;
        r.kappa_xx = bd00->kappa_xx;
        r.kappa_xy = bd00->kappa_xy;
        r.kappa_yy = bd00->kappa_yy;
    }
    return r;
}

static void Reflect3(const RayPt3 *oldp, RayPt3 *newp, const HS *hs, int isTop, v3 nB, real z_xx, real z_xy,
                     real z_yy, const Prob3 *Q, Seg3 *iSeg, ErrState *err)
{
    real Th = v3dot(oldp->t, nB);
    v3 tB = v3sub(oldp->t, v3scl(Th, nB));
    tB = v3scl(RL(1.0) / v3len(tB), tB);
    real Tg = v3dot(oldp->t, tB);
    newp->NumTopBnc = oldp->NumTopBnc + isTop;
    newp->NumBotBnc = oldp->NumBotBnc + (1 - isTop);
    newp->x = oldp->x;
    newp->t = v3sub(oldp->t, v3scl(FL(2.0) * Th, nB));
    SSPOut3 o;
    EvaluateSSP3(oldp->x, newp->t, &o, Q, iSeg, err);
    real c = creal(o.ccpx);
    newp->c = c;
    newp->tau = oldp->tau;
    v3 rayt, rayn1, rayn2, rayt_t, rayn1_t, rayn2_t;
    CalcTangent_Normals(oldp, c, nB, &rayt, &rayn1, &rayn2, RL(-1.0));
    CalcTangent_Normals(newp, c, nB, &rayt_t, &rayn1_t, &rayn2_t, RL(-1.0));
    v2 t_rot = V2(rayt.x, rayt.y);
    t_rot = vscl(RL(1.0) / vlen(t_rot), t_rot);
    v2 n_rot = V2(rayn2.x, rayn2.y);
    n_rot = vscl(RL(1.0) / vlen(n_rot), n_rot);
    m2 R = M2(t_rot.x, t_rot.y, n_rot.x, n_rot.y);
    /* kappaMat from the boundary's z curvatures (zero off curvilinear grids) */
    m2 kappa = M2(z_xx / FL(2.0), z_xy / FL(2.0), z_xy / FL(2.0), z_yy / FL(2.0));
    m2 DMatTemp = m2mul(m2scl(FL(2.0), kappa), R);
    m2 DMat = m2mul(m2tr(R), DMatTemp);
    real cn1jump = v3dot(o.gradc, v3sub(v3neg(rayn1_t), rayn1));
    real cn2jump = v3dot(o.gradc, v3sub(v3neg(rayn2_t), rayn2));
    real csjump = -v3dot(o.gradc, v3sub(rayt_t, rayt));
    v3 e1, e2;
    RayNormalImpl(oldp->t, oldp->phi, 0, oldp->c, &e1, &e2);
    newp->p = oldp->p;
    newp->q = oldp->q;
    newp->phi = oldp->phi;
    CurvatureCorrection3D(newp, DMat, Tg, Th, cn1jump, cn2jump, csjump, rayn1, rayn2, e1, e2, 1);

    if(hs->bc == 'R') {
        newp->Amp = oldp->Amp;
        newp->Phase = oldp->Phase;
    } else if(hs->bc == 'V') {
        newp->Amp = oldp->Amp;
        newp->Phase = oldp->Phase + REAL_PI;
    } else if(hs->bc == 'F') {
        ReflCoef RInt;
        RInt.theta = RadDeg * RABS(RATAN2(Th, Tg));
        if(RInt.theta > FL(90.0)) RInt.theta = FL(180.0) - RInt.theta;
        if(isTop) InterpolateReflectionCoefficient(&RInt, Q->P2.trc, Q->P2.ntrc, err);
        else InterpolateReflectionCoefficient(&RInt, Q->P2.brc, Q->P2.nbrc, err);
        newp->Amp = oldp->Amp * RInt.r;
        newp->Phase = oldp->Phase + RInt.phi;
    } else if(hs->bc == 'A' || hs->bc == 'G') {
        real omega = FL(2.0) * REAL_PI * Q->P2.freq0;
        cpx gk = omega * Tg;
        const cpx J = I;
        cpx g1sq = SQ(omega / c) - SQ(gk) - J * REAL_MINPOS;
        cpx g2sq = SQ(omega / hs->cP) - SQ(gk) - J * REAL_MINPOS;
        cpx g1 = CSQRT(-g1sq);
        cpx g2 = CSQRT(-g2sq);
        cpx Refl = (hs->rho * g1 - o.rho * g2) / (hs->rho * g1 + o.rho * g2);
        real aR = CABS(Refl);
        if(aR < FL(1.0e-5)) {
            newp->Amp = FL(0.0);
            newp->Phase = oldp->Phase;
        } else {
            newp->Amp = aR * oldp->Amp;
            newp->Phase = oldp->Phase + RATAN2(cimag(Refl), creal(Refl));
        }
    } else {
        RunError(err, BHC_ERR_BOUNDARY_CONDITION_TYPE);
        newp->Amp = NAN;
        newp->Phase = NAN;
    }
}

/* ------------------------------------------------------------------------- */
/* Ray init / update / terminate, src/trace.hpp (3D)                           */
/* ------------------------------------------------------------------------- */
typedef struct {
    int isx, isy, isz, ialpha, ibeta;
    real alpha, beta, SrcDeclAngle, SrcAzimAngle;
} RayInit3Info;

static void Distances3(v3 x, const BState3 *top, const BState3 *bot, real *dt, real *db)
{ /* trace.hpp:37-45 */
    *dt = -v3dot(top->n, v3sub(x, top->x));
    *db = -v3dot(bot->n, v3sub(x, bot->x));
}

static int RayInit3(RayInit3Info *ri, RayPt3 *p0, TraceState3 *S, const Prob3 *Q, ErrState *err)
{ /* trace.hpp:51-67, 76-266 (3D) */
    const Prob *P = &Q->P2;
    if(ri->isz < 0 || ri->isz >= P->NSz || ri->ialpha < 0 || ri->ialpha >= P->Nalpha || ri->isx < 0
       || ri->isx >= Q->NSx || ri->isy < 0 || ri->isy >= Q->NSy || ri->ibeta < 0 || ri->ibeta >= Q->Nbeta) {
        RunError(err, BHC_ERR_RAYINIT);
        return 0;
    }
    ri->alpha = P->alpha[ri->ialpha];
    ri->SrcDeclAngle = RadDeg * ri->alpha;
    ri->beta = Q->beta[ri->ibeta];
    ri->SrcAzimAngle = RadDeg * ri->beta;
    S->iSeg.x = S->iSeg.y = S->iSeg.z = S->iSeg.r = 0;
    S->xs = V3(Q->Sx[ri->isx], Q->Sy[ri->isy], (real)P->Sz[ri->isz]);
    SSPOut3 o;
    EvaluateSSP3(S->xs, V3(FL(0.0), FL(0.0), FL(1.0)), &o, Q, &S->iSeg, err);
    real c = creal(o.ccpx);
    int ibp = BinarySearchLEQ(P->sbp, P->nsbp, 2, 0, ri->SrcDeclAngle);
    ibp = imin(ibp, P->nsbp - 2);
    real s = (ri->SrcDeclAngle - P->sbp[2 * ibp]) / (P->sbp[2 * (ibp + 1)] - P->sbp[2 * ibp]);
    float Amp0 = (float)((FL(1.0) - s) * P->sbp[2 * ibp + 1] + s * P->sbp[2 * (ibp + 1) + 1]);
    if(P->RunType[0] == 'S') {
        float omega = (float)(FL(2.0) * REAL_PI * P->freq0);
        Amp0 = (float)(Amp0 * (sqrtf(FL(2.0)) * RABS(RSIN(omega / c * S->xs.z * RSIN(ri->alpha)))));
    }
    v3 tinit2 = V3(RCOS(ri->alpha) * RCOS(ri->beta), RCOS(ri->alpha) * RSIN(ri->beta), RSIN(ri->alpha));
    p0->x = S->xs;
    p0->c = c;
    p0->t = v3divs(tinit2, c);
    p0->tau = 0;
    p0->Amp = Amp0;
    p0->Phase = FL(0.0);
    p0->NumTopBnc = p0->NumBotBnc = 0;
    p0->p = M2(FL(1.0), FL(0.0), FL(0.0), FL(1.0));
    p0->q = M2(FL(0.0), FL(0.0), FL(0.0), FL(0.0));
    p0->phi = FL(0.0);
    memset(&S->top, 0, sizeof(S->top));
    memset(&S->bot, 0, sizeof(S->bot));
    GetBdrySeg3(S->xs, p0->t, &S->top, &Q->ati, 1, 1, err);
    GetBdrySeg3(S->xs, p0->t, &S->bot, &Q->bty, 0, 1, err);
    real dt, db;
    Distances3(S->xs, &S->top, &S->bot, &dt, &db);
    if(dt <= FL(0.0) || db <= FL(0.0)) {
        RunWarning(err, BHC_WARN_SOURCE_OUTSIDE_BOUNDARIES);
        return 0;
    }
    S->iSmallStepCtr = 0;
    return 1;
}

static int RayUpdate3(const RayPt3 *p0, RayPt3 *p1, RayPt3 *p2, TraceState3 *S, const Prob3 *Q, ErrState *err)
{ /* trace.hpp:274-381 (3D, rectilinear boundaries) */
    int topRefl = 0, botRefl = 0;
    Step3(*p0, p1, S, Q, err, &topRefl, &botRefl);
    GetBdrySeg3(p1->x, p1->t, &S->top, &Q->ati, 1, 0, err);
    GetBdrySeg3(p1->x, p1->t, &S->bot, &Q->bty, 0, 0, err);
    if(topRefl || botRefl) {
        ReflBdry3 r = refl_bdry3(topRefl ? &Q->ati : &Q->bty, topRefl ? &S->top : &S->bot,
                                 topRefl ? &Q->P2.top : &Q->P2.bot, p1->x);
        Reflect3(p1, p2, &r.hs, topRefl, r.n, r.z_xx, r.z_xy, r.z_yy, Q, &S->iSeg, err);
        return 1;
    }
    return 0;
}

static int RayTerminate3(const RayPt3 *pt, int *Nsteps, int is, const TraceState3 *S, int MaxPointsPerRay,
                         const Prob3 *Q, ErrState *err)
{ /* trace.hpp:391-477 (3D) */
    v3 x = pt->x, t = pt->t, xs = S->xs;
    int leftbox = RABS(x.x - xs.x) >= Q->box_x || RABS(x.y - xs.y) >= Q->box_y || RABS(x.z - RL(0.0)) >= Q->box_z;
    const Bdry3 *bt = &Q->bty, *at = &Q->ati;
    real minx = rmax(bt->bd[0].x.x, at->bd[0].x.x);
    real miny = rmax(bt->bd[0].x.y, at->bd[0].x.y);
    real maxx = rmin(bt->bd[(bt->nx - 1) * bt->ny].x.x, at->bd[(at->nx - 1) * at->ny].x.x);
    real maxy = rmin(bt->bd[bt->ny - 1].x.y, at->bd[at->ny - 1].x.y);
    int e0 = x.x < minx || x.y < miny;
    int eN = x.x > maxx || x.y > maxy;
    e0 = e0 || (x.x == minx && t.x < RL(0.0)) || (x.y == miny && t.y < RL(0.0));
    eN = eN || (x.x == maxx && t.x > RL(0.0)) || (x.y == maxy && t.y > RL(0.0));
    int small = S->iSmallStepCtr > 50;
    int lost = pt->Amp < FL(0.005);
    if(leftbox || lost || e0 || eN || small) {
        *Nsteps = is + 1;
        return 1;
    } else if(is >= MaxPointsPerRay - 3) {
        RunWarning(err, BHC_WARN_ONERAY_OUTOFMEMORY);
        *Nsteps = is;
        return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Influence, src/influence.hpp (geometric Cartesian, 3D)                      */
/* ------------------------------------------------------------------------- */
typedef struct {
    RayInit3Info init;
    real freq0, omega, BeamWindow, Ratio1, rcp_q0, rcp_qhat0, phase, qOld;
    v3 xs, x;
    v3 rayn1, rayn2; /* ray-centred beams: the previous point's ray normals */
    int ir;
} Infl3;

typedef struct {
    const Prob3 *Q;
    float *field;
    bhc_arrival *arr;
    _Atomic int32_t *narr;
    int max_narr, allow_merging;
    bhc_eigen_hit *hits;
    _Atomic int32_t *neigen;
    int eigen_memsize;
    int isGaussian, isRayCen;
    _Atomic long long contrib;
} Outs3;

static size_t field_addr3(const Prob3 *Q, const RayInit3Info *ri, int itheta, int iz, int ir)
{ /* GetFieldAddr, common.hpp:393-405 */
    return ((((((size_t)ri->isz * Q->NSx + ri->isx) * Q->NSy + ri->isy) * Q->Ntheta + itheta)
             * Q->P2.NRz_per_range + iz) * Q->P2.NRr) + ir;
}

static void RecordEigenHit3(int itheta, int ir, int iz, int is, const RayInit3Info *ri, Outs3 *O)
{ /* eigenrays.hpp:24-41 */
    int mi = atomic_fetch_add(O->neigen, 1);
    if(mi >= O->eigen_memsize) return;
    bhc_eigen_hit *h = &O->hits[mi];
    h->is = is; h->iz = iz; h->ir = ir; h->itheta = itheta;
    h->isx = ri->isx; h->isy = ri->isy; h->isz = ri->isz; h->ialpha = ri->ialpha; h->ibeta = ri->ibeta;
}

static void InfluenceGeoCore3(real s, real n1, real n2, m2 dq, cpx dtau, int itheta, int ir, int iz, int is,
                              real RcvrDeclAngle, real RcvrAzimAngle, const RayPt3 *p0, const RayPt3 *p1,
                              const Infl3 *inf, Outs3 *O)
{ /* influence.hpp:897-1024 (3D Cartesian) + ApplyContribution :432-473 */
    const Prob *P = &O->Q->P2;
    m2 qI = m2add(p0->q, m2scl(s, dq));
    qI.m[0][0] *= inf->rcp_q0;
    qI.m[1][0] *= inf->rcp_q0;
    qI.m[0][1] *= inf->rcp_qhat0;
    qI.m[1][1] *= inf->rcp_qhat0;
    real qFinal = m2det(qI);
    if(!O->isRayCen) { /* the row-length test is Cartesian only */
        real l1 = RSQRT(qI.m[0][0] * qI.m[0][0] + qI.m[1][0] * qI.m[1][0]);
        real l2 = RSQRT(qI.m[0][1] * qI.m[0][1] + qI.m[1][1] * qI.m[1][1]);
        if(RABS(l1) < REAL_EPSILON || RABS(l2) < REAL_EPSILON) return;
    }
    if(RABS(qFinal) < REAL_EPSILON) return;
    real n1p = RABS((-qI.m[0][1] * n2 + qI.m[1][1] * n1) / qFinal);
    real n2p = RABS((qI.m[0][0] * n2 - qI.m[1][0] * n1) / qFinal);
    real bcd = O->isGaussian ? n1p + n2p : rmax(n1p, n2p);
    /* the window test is not applied to ray-centred Gaussian beams in 3D */
    if(bcd > inf->BeamWindow * RL(1.0) && !(O->isRayCen && O->isGaussian)) return;
    cpx delay = p0->tau + s * dtau;
    real cnst = inf->Ratio1 * p1->c * p1->Amp / RSQRT(RABS(qFinal));
    real w;
    if(O->isGaussian) w = REXP(FL(-0.5) * (SQ(n1p) + SQ(n2p)));
    else w = (FL(1.0) - n1p) * (FL(1.0) - n2p);
    real phaseInt = p0->Phase + inf->phase;
    if((qFinal <= RL(0.0) && inf->qOld > RL(0.0)) || (qFinal >= RL(0.0) && inf->qOld < RL(0.0)))
        phaseInt = phaseInt + REAL_PI / FL(2.0);
    atomic_fetch_add_explicit(&O->contrib, 1, memory_order_relaxed);
    const char rt = P->RunType[0];
    if(rt == 'E') {
        RecordEigenHit3(itheta, ir, iz, is, &inf->init, O);
        return;
    }
    if(rt == 'A' || rt == 'a' || rt == 'V' || rt == 'v') {
        size_t b = field_addr3(O->Q, &inf->init, itheta, iz, ir);
        AddArrAt(&O->arr[b * O->max_narr], &O->narr[b], O->max_narr, O->allow_merging, cnst * w, inf->omega,
                 phaseInt, delay, inf->init.SrcDeclAngle, inf->init.SrcAzimAngle, RcvrDeclAngle, RcvrAzimAngle,
                 p1->NumTopBnc, p1->NumBotBnc);
        if(rt == 'V' || rt == 'v') RecordEigenHit3(itheta, ir, iz, is, &inf->init, O); /* IsAlsoEigenraysRun */
        return;
    }
    float fr, fi;
    if(rt == 'C') {
        cpx e = CEXP(-I * (inf->omega * delay - phaseInt));
        cpx d = (cnst * w) * e;
        fr = (float)creal(d);
        fi = (float)cimag(d);
    } else {
        real v = cnst * REXP(cimag(inf->omega * delay));
        v = SQ(v) * w;
        if(O->isGaussian) v *= FL(2.0) * REAL_PI;
        fr = (float)v;
        fi = 0.0f;
    }
    size_t base = field_addr3(O->Q, &inf->init, itheta, iz, ir);
    atomic_add_float(&O->field[2 * base], fr);
    atomic_add_float(&O->field[2 * base + 1], fi);
}

static void Step_InfluenceGeoCart3(const RayPt3 *p0, const RayPt3 *p1, Infl3 *inf, int is, Outs3 *O)
{ /* influence.hpp:1201-1406 (3D) */
    const Prob3 *Q = O->Q;
    const Prob *P = &Q->P2;
    v2 xs2 = V2(inf->xs.x, inf->xs.y);
    real rA = vlen(vsub(V2(p0->x.x, p0->x.y), xs2));
    real rB = vlen(vsub(V2(p1->x.x, p1->x.y), xs2));
    if(RABS(rA - rB) <= RL(1.0e3) * spacing(rA)) return;
    if(is == 0) inf->ir = (rB > rA) ? 0 : P->NRr - 1;
    v3 x_ray = p0->x;
    v3 rayt = v3sub(p1->x, p0->x);
    real rlen = v3len(rayt);
    if(rlen <= RL(1.0e3) * spacing(p1->x.x)) return;
    rayt = v3divs(rayt, rlen);
    v3 rayn1, rayn2;
    RayNormalImpl(rayt, p1->phi, 1, RL(1.0), &rayn1, &rayn2);
    /* ReceiverAngles, influence.hpp:123-136 (3D) */
    real RcvrAzimAngle = RadDeg * RATAN2(rayt.y, rayt.x);
    real RcvrDeclAngle = RadDeg * RATAN2(rayt.z, RSQRT(rayt.x * rayt.x + rayt.y * rayt.y));
    m2 dq = m2sub(p1->q, p0->q);
    cpx dtau = p1->tau - p0->tau;
    real phaseq = m2det(p0->q);
    if((phaseq <= RL(0.0) && inf->qOld > RL(0.0)) || (phaseq >= RL(0.0) && inf->qOld < RL(0.0)))
        inf->phase += REAL_PI / FL(2.0);
    inf->qOld = phaseq;
    real a0 = p0->q.m[0][0] * inf->rcp_q0, b0 = p0->q.m[1][0] * inf->rcp_q0;
    real a1 = p1->q.m[0][0] * inf->rcp_q0, b1 = p1->q.m[1][0] * inf->rcp_q0;
    real l1 = rmax(RSQRT(a0 * a0 + b0 * b0), RSQRT(a1 * a1 + b1 * b1));
    a0 = p0->q.m[0][1] * inf->rcp_qhat0; b0 = p0->q.m[1][1] * inf->rcp_qhat0;
    a1 = p1->q.m[0][1] * inf->rcp_qhat0; b1 = p1->q.m[1][1] * inf->rcp_qhat0;
    real l2 = rmax(RSQRT(a0 * a0 + b0 * b0), RSQRT(a1 * a1 + b1 * b1));
    real L_diag = RSQRT(SQ(l1) + SQ(l2));
    for(;;) {
        real rr = P->Rr[inf->ir];
        if(rr >= rmin(rA, rB) && rr < rmax(rA, rB)) {
            for(int itheta = 0; itheta < Q->Ntheta; ++itheta) {
                v2 tr = Q->t_rcvr[itheta];
                v2 xr2 = vadd(xs2, vscl(rr, tr));
                real m_prime = RABS(vdot(vsub(xr2, V2(x_ray.x, x_ray.y)), V2(-rayt.y, rayt.x)));
                if(!O->isGaussian && m_prime > inf->BeamWindow * L_diag) continue;
                real sdot = vdot(vsub(xr2, xs2), vsub(V2(x_ray.x, x_ray.y), xs2));
                if(sdot < RL(0.0)) continue;
                v2 e_theta = V2(-tr.y, tr.x);
                real n_ray_z = rayt.x * e_theta.y - rayt.y * e_theta.x;
                if(RABS(n_ray_z) < RL(1e-9)) continue;
                real L_z = inf->BeamWindow * L_diag / RABS(n_ray_z);
                real zmin = rmin(p0->x.z, p1->x.z) - L_z;
                real zmax = rmax(p0->x.z, p1->x.z) + L_z;
                for(int iz = 0; iz < P->NRz_per_range; ++iz) {
                    real zr = (real)P->Rz[iz];
                    if(zr < zmin || zr > zmax) continue;
                    v3 xrr = v3sub(V3(xr2.x, xr2.y, zr), x_ray);
                    real s = v3dot(xrr, rayt) / rlen;
                    real n1 = RABS(v3dot(xrr, rayn1));
                    real n2 = RABS(v3dot(xrr, rayn2));
                    InfluenceGeoCore3(s, n1, n2, dq, dtau, itheta, inf->ir, iz, is, RcvrDeclAngle, RcvrAzimAngle, p0,
                                      p1, inf, O);
                }
            }
        }
        int irTT;
        if(rB > P->Rr[inf->ir]) {
            if(inf->ir >= P->NRr - 1) break;
            irTT = inf->ir + 1;
            if(P->Rr[irTT] >= rB) break;
        } else {
            if(inf->ir <= 0) break;
            irTT = inf->ir - 1;
            if(P->Rr[irTT] <= rB) break;
        }
        inf->ir = irTT;
    }
    inf->x = p1->x;
}

static void RayCenCross3(v3 x, v3 rn1, v3 rn2, real zR, const Infl3 *inf, v3 *xt, v3 *xtxe1, v3 *xtxe2)
{ /* Compute_RayCenCross, influence.hpp:154-163 */
    *xt = v3sub(x, V3(inf->xs.x, inf->xs.y, zR));
    *xtxe1 = v3cross(*xt, rn1);
    *xtxe2 = v3cross(*xt, rn2);
}

static int M_N_R_IR3(real *m, real *n, real *r, int *ir, v3 e1xe2, v3 xtxe1, v3 xtxe2, v3 xt, v2 tr, const Prob *P)
{ /* Compute_M_N_R_IR, influence.hpp:178-189; 1: the ray normal is parallel to the receiver line */
    real delta = -vdot(tr, V2(e1xe2.x, e1xe2.y));
    if(RABS(delta) < FL(1e-3)) return 1;
    *m = vdot(tr, V2(xtxe1.x, xtxe1.y)) / delta;
    *n = -vdot(tr, V2(xtxe2.x, xtxe2.y)) / delta;
    *r = -v3dot(xt, e1xe2) / delta;
    *ir = RToIR(*r, P);
    return 0;
}

static void Step_InfluenceGeoRayCen3(const RayPt3 *p0, const RayPt3 *p1, Infl3 *inf, int is, Outs3 *O)
{ /* influence.hpp:1030-1194 (3D) */
    const Prob3 *Q = O->Q;
    const Prob *P = &Q->P2;
    real phaseq = m2det(p0->q);
    if((phaseq <= RL(0.0) && inf->qOld > RL(0.0)) || (phaseq >= RL(0.0) && inf->qOld < RL(0.0)))
        inf->phase += REAL_PI / FL(2.0);
    inf->qOld = phaseq;
    m2 dq = m2sub(p1->q, p0->q);
    cpx dtau = p1->tau - p0->tau;
    v3 e1xe2A = v3scl(p0->c, p0->t), e1xe2B = v3scl(p1->c, p1->t);
    v3 rayn1, rayn2;
    RayNormalImpl(p1->t, p1->phi, 0, p1->c, &rayn1, &rayn2);
    /* IsDuplicatePoint: hat beams compare with a constant 1e-4 m threshold */
    real thresh = O->isGaussian ? RL(1.0e3) * spacing(p1->x.x) : FL(1.0e-4);
    if(v3len(v3sub(p1->x, p0->x)) <= thresh) {
        inf->x = p1->x;
        inf->rayn1 = rayn1;
        inf->rayn2 = rayn2;
        return;
    }
    v3 d = v3sub(p1->x, p0->x);
    real RcvrAzimAngle = RadDeg * RATAN2(d.y, d.x);
    real RcvrDeclAngle = RadDeg * RATAN2(d.z, RSQRT(d.x * d.x + d.y * d.y));
    for(int iz = 0; iz < P->NRz_per_range; ++iz) {
        real zR = (real)P->Rz[iz];
        v3 xtA, xtxe1A, xtxe2A, xtB, xtxe1B, xtxe2B;
        RayCenCross3(p0->x, inf->rayn1, inf->rayn2, zR, inf, &xtA, &xtxe1A, &xtxe2A);
        RayCenCross3(p1->x, rayn1, rayn2, zR, inf, &xtB, &xtxe1B, &xtxe2B);
        for(int itheta = 0; itheta < Q->Ntheta; ++itheta) {
            real mA, nA, rA, mB, nB, rB;
            int irA, irB;
            if(M_N_R_IR3(&mA, &nA, &rA, &irA, e1xe2A, xtxe1A, xtxe2A, xtA, Q->t_rcvr[itheta], P)) continue;
            if(M_N_R_IR3(&mB, &nB, &rB, &irB, e1xe2B, xtxe1B, xtxe2B, xtB, Q->t_rcvr[itheta], P)) continue;
            if(O->isGaussian) {
                real MaxRadius_m = inf->BeamWindow
                    * rmax(RABS(p0->q.m[1][1] * inf->rcp_qhat0), RABS(p1->q.m[1][1] * inf->rcp_qhat0));
                real MaxRadius_n = inf->BeamWindow
                    * rmax(RABS(p0->q.m[0][0] * inf->rcp_q0), RABS(p1->q.m[0][0] * inf->rcp_q0));
                if(MaxRadius_m <= rmin(RABS(mA), RABS(mB)) && mA * mB >= RL(0.0)) continue;
                if(MaxRadius_n <= rmin(RABS(nA), RABS(nB)) && nA * nB >= RL(0.0)) continue;
            }
            if(irA == irB) continue;
            int notii = (irB <= irA) ? 0 : 1;
            for(int ir = irA + notii; ir != irB + notii; ir += (notii << 1) - 1) {
                real sw = (P->Rr[ir] - rA) / (rB - rA);
                real n1 = RABS(nA + sw * (nB - nA));
                real n2 = RABS(mA + sw * (mB - mA));
                InfluenceGeoCore3(sw, n1, n2, dq, dtau, itheta, ir, iz, is, RcvrDeclAngle, RcvrAzimAngle, p0, p1,
                                  inf, O);
            }
        }
    }
    inf->x = p1->x;
    inf->rayn1 = rayn1;
    inf->rayn2 = rayn2;
}

static void Init_Influence3(Infl3 *inf, const RayPt3 *p0, const RayInit3Info *ri, const Prob3 *Q, int isGaussian,
                            int isRayCen)
{ /* influence.hpp:525-644 (geometric, 3D) */
    const Prob *P = &Q->P2;
    inf->init = *ri;
    inf->freq0 = P->freq0;
    inf->omega = FL(2.0) * REAL_PI * inf->freq0;
    inf->BeamWindow = isGaussian ? RL(4.0) : RL(1.0);
    inf->Ratio1 = RSQRT(RABS(RCOS(ri->alpha)));
    inf->Ratio1 *= RSQRT(P->Dalpha * Q->Dbeta) / p0->c;
    if(isGaussian) inf->Ratio1 /= FL(2.0) * REAL_PI;
    inf->rcp_q0 = P->Dalpha / p0->c;
    inf->rcp_qhat0 = RABS(RCOS(ri->alpha)) * Q->Dbeta / p0->c;
    inf->phase = FL(0.0);
    inf->qOld = m2det(p0->q);
    inf->xs = p0->x;
    inf->x = p0->x;
    inf->ir = BinarySearchGT(P->Rr, P->NRr, RL(0.0));
    inf->rayn1 = inf->rayn2 = V3(RL(0.0), RL(0.0), RL(0.0));
    if(isRayCen) RayNormalImpl(p0->t, p0->phi, 0, p0->c, &inf->rayn1, &inf->rayn2); /* RayNormalRayCen */
}

typedef struct {
    int32_t nsteps, ntop, nbot;
    double xend[3];
} RayStat3;

static void MainFieldModes3(RayInit3Info *ri, Outs3 *O, ErrState *err, RayStat3 *st)
{ /* trace.hpp:525-589 */
    const Prob3 *Q = O->Q;
    TraceState3 S;
    RayPt3 p0, p1, p2;
    Infl3 inf;
    memset(&p2, 0, sizeof(p2));
    if(st) { memset(st, 0, sizeof(*st)); }
    if(!RayInit3(ri, &p0, &S, Q, err)) return;
    Init_Influence3(&inf, &p0, ri, Q, O->isGaussian, O->isRayCen);
    int is = 0, Nsteps = 0;
    for(;;) {
        if(HasErrored(err)) break;
        int two = RayUpdate3(&p0, &p1, &p2, &S, Q, err);
        if(O->isRayCen) Step_InfluenceGeoRayCen3(&p0, &p1, &inf, is, O);
        else Step_InfluenceGeoCart3(&p0, &p1, &inf, is, O);
        ++is;
        if(two) {
            if(O->isRayCen) Step_InfluenceGeoRayCen3(&p1, &p2, &inf, is, O);
            else Step_InfluenceGeoCart3(&p1, &p2, &inf, is, O);
            p0 = p2;
            ++is;
        } else {
            p0 = p1;
        }
        if(RayTerminate3(&p0, &Nsteps, is, &S, MaxN, Q, err)) break;
    }
    if(st) {
        st->nsteps = is;
        st->ntop = p0.NumTopBnc;
        st->nbot = p0.NumBotBnc;
        st->xend[0] = (double)p0.x.x;
        st->xend[1] = (double)p0.x.y;
        st->xend[2] = (double)p0.x.z;
    }
}

static int num_jobs3(const Prob3 *Q)
{ /* common.hpp:347-359 */
    int ret = 1;
    if(Q->P2.iSingle == 0) ret *= Q->P2.Nalpha;
    if(Q->iSingleB == 0) ret *= Q->Nbeta;
    ret *= Q->NSy;
    ret *= Q->NSx;
    ret *= Q->P2.NSz;
    return ret;
}
static int job_indices3(RayInit3Info *ri, int job, const Prob3 *Q)
{ /* common.hpp:365-391 */
    if(Q->P2.iSingle >= 1) ri->ialpha = Q->P2.iSingle - 1;
    else { ri->ialpha = job % Q->P2.Nalpha; job /= Q->P2.Nalpha; }
    if(Q->iSingleB >= 1) ri->ibeta = Q->iSingleB - 1;
    else { ri->ibeta = job % Q->Nbeta; job /= Q->Nbeta; }
    ri->isy = job % Q->NSy;
    job /= Q->NSy;
    ri->isx = job % Q->NSx;
    job /= Q->NSx;
    ri->isz = job;
    return ri->isz < Q->P2.NSz;
}

typedef struct {
    Outs3 *O;
    ErrState *err;
    _Atomic int32_t next;
    int32_t job_begin, job_end;
    RayStat3 *stats;
} FieldJob3;

static void *field_worker3(void *arg)
{
    FieldJob3 *J = (FieldJob3 *)arg;
    for(;;) {
        int32_t job = atomic_fetch_add(&J->next, 1);
        if(job >= J->job_end) break;
        RayInit3Info ri;
        if(!job_indices3(&ri, job, J->O->Q)) break;
        MainFieldModes3(&ri, J->O, J->err, J->stats ? &J->stats[job - J->job_begin] : NULL);
    }
    return NULL;
}

/* Exported: 3D TL field run on jobs [job_begin, job_end) (job_end < 0: all).
 * out->field is [NSz][NSx][NSy][Ntheta][NRz][NRr] complex floats; ray_xend is [rays][3]. */
int SFX(oracle_field_run3d)(const bhc_env3d *env, bhc_field_out *out, int nthreads, int job_begin, int job_end)
{
    char rt0 = env->RunType[0], rt1 = env->RunType[1];
    if(rt0 != 'C' && rt0 != 'S' && rt0 != 'I' && rt0 != 'A' && rt0 != 'a' && rt0 != 'E' && rt0 != 'V' && rt0 != 'v')
        return -1;
    if(rt1 != 'G' && rt1 != '^' && rt1 != ' ' && rt1 != 'B' && rt1 != 'g' && rt1 != 'b') return -1;
    Prob3 Q;
    if(prob3_from_env(env, &Q)) return -1;
    Outs3 O;
    memset(&O, 0, sizeof(O));
    O.Q = &Q;
    O.field = out->field;
    O.arr = out->arr;
    O.narr = (_Atomic int32_t *)out->narr;
    O.max_narr = out->max_narr;
    O.allow_merging = out->allow_merging;
    O.hits = out->hits;
    O.neigen = (_Atomic int32_t *)out->neigen;
    O.eigen_memsize = out->eigen_memsize;
    int arr_run = rt0 == 'A' || rt0 == 'a' || rt0 == 'V' || rt0 == 'v';
    int eig_run = rt0 == 'E' || rt0 == 'V' || rt0 == 'v';
    if(arr_run && (!O.arr || !O.narr || O.max_narr < 1)) { prob3_free(&Q); return -1; }
    if(eig_run && (!O.hits || !O.neigen)) { prob3_free(&Q); return -1; }
    O.isGaussian = (rt1 == 'B' || rt1 == 'b');
    O.isRayCen = (rt1 == 'g' || rt1 == 'b');
    atomic_init(&O.contrib, 0);
    ErrState err;
    atomic_init(&err.error, 0u);
    atomic_init(&err.warning, 0u);
    atomic_init(&err.errCount, 0u);
    atomic_init(&err.warnCount, 0u);
    int nj = num_jobs3(&Q);
    if(job_end < 0 || job_end > nj) job_end = nj;
    if(job_begin < 0) job_begin = 0;
    int nrays = job_end - job_begin;
    RayStat3 *stats = NULL;
    if(out->ray_nsteps) stats = xcalloc(nrays > 0 ? nrays : 1, sizeof(RayStat3));
    FieldJob3 J;
    J.O = &O;
    J.err = &err;
    atomic_init(&J.next, job_begin);
    J.job_begin = job_begin;
    J.job_end = job_end;
    J.stats = stats;
    nthreads = resolve_threads(nthreads);
    if(arr_run && O.allow_merging) nthreads = 1; /* src/mode/arr.hpp:79 */
    pthread_t *th = xcalloc(nthreads, sizeof(pthread_t));
    for(int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, field_worker3, &J);
    for(int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th);
    if(stats) {
        int64_t tot = 0;
        for(int i = 0; i < nrays; ++i) {
            out->ray_nsteps[i] = stats[i].nsteps;
            if(out->ray_ntop) out->ray_ntop[i] = stats[i].ntop;
            if(out->ray_nbot) out->ray_nbot[i] = stats[i].nbot;
            if(out->ray_xend)
                for(int k = 0; k < 3; ++k) out->ray_xend[3 * i + k] = stats[i].xend[k];
            tot += stats[i].nsteps;
        }
        out->total_steps = tot;
        free(stats);
    }
    out->total_contrib = atomic_load(&O.contrib);
    out->err = atomic_load(&err.error);
    out->warn = atomic_load(&err.warning);
    out->err_count = atomic_load(&err.errCount);
    out->warn_count = atomic_load(&err.warnCount);
    prob3_free(&Q);
    return 0;
}

/* Exported: the preprocessed 3D tables, for the preprocessing parity test.
 * Order: boundary nodes (bty then ati: x, n1, n2), hexahedral cz (if 'H'),
 * receiver bearing cos / sin, beta (rad), Sx, Sy, scalars deltas, box x / y / z, Dbeta. */
/* ------------------------------------------------------------------------- */
/* Nx2D: 2D rays in the vertical plane of each bearing of the 3D environment   */
/* (bellhopcxx Nx2D, O3D = true, R3D = false)                                  */
/* ------------------------------------------------------------------------- */
typedef struct { v3 xs; v2 tradial; } OriginNx; /* Origin<true, false>, include/bhc/structs.hpp:233-237 */

static v3 RayToOceanX(v2 x, const OriginNx *o)
{ /* common_run.hpp:221-231 */
    return V3(o->xs.x + x.x * o->tradial.x, o->xs.y + x.x * o->tradial.y, x.y);
}
static v3 RayToOceanT(v2 t, const OriginNx *o)
{ /* common_run.hpp:233-242 */
    return V3(t.x * o->tradial.x, t.x * o->tradial.y, t.y);
}
static v2 OceanToRayT(v3 t, const OriginNx *o)
{ /* common_run.hpp:308-317 */
    return V2(vdot(V2(t.x, t.y), o->tradial), t.z);
}
static real RealBitsAddInt(real r, int i)
{ /* common.hpp:316-341: step through the representation */
#if OR_IS_FLOAT
    int32_t k;
    memcpy(&k, &r, 4);
    k += i;
#else
    int64_t k;
    memcpy(&k, &r, 8);
    k += (int64_t)i;
#endif
    real x;
    memcpy(&x, &k, sizeof(real));
    return x;
}
static v2 OceanToRayX(v3 x, const OriginNx *o, v2 t, int snapDim, ErrState *err)
{ /* common_run.hpp:253-306: range from the larger bearing component, then up
   * to 4 one-ulp moves along the ray until the round trip lands on (or just
   * past, in the ray's direction) the snapped ocean coordinate */
    v2 ret;
    ret.y = x.z;
    if(RABS(o->tradial.x) >= RABS(o->tradial.y)) ret.x = (x.x - o->xs.x) / o->tradial.x;
    else ret.x = (x.y - o->xs.y) / o->tradial.y;
    if(snapDim < -2 || snapDim == -1 || snapDim >= 2) return ret;
    for(int i = 0; i < 4; ++i) {
        v3 xb = RayToOceanX(ret, o);
        v2 want = vscl(t.x, o->tradial);
        v2 errd = V2(xb.x - x.x, xb.y - x.y);
        int okx = (want.x * errd.x) >= RL(0.0);
        int oky = (want.y * errd.y) >= RL(0.0);
        if((snapDim == 0 && okx) || (snapDim == 1 && oky) || (okx && oky)) return ret;
        ret.x = RealBitsAddInt(ret.x, (t.x > RL(0.0)) ? 1 : -1);
    }
    RunWarning(err, BHC_WARN_OCEANTORAYX_GAVEUP);
    return ret;
}

static void EvaluateSSPNx(v2 x, v2 t, SSPOut *o, const OriginNx *org, const Prob3 *Q, Seg3 *iSeg, ErrState *err)
{ /* ssp.hpp:377-465 with O3D && !R3D: the 3D profile evaluated at the ocean
   * point, gradient and second derivatives projected on the bearing */
    SSPOut3 p;
    EvaluateSSP3(RayToOceanX(x, org), RayToOceanT(t, org), &p, Q, iSeg, err);
    o->gradc.x = vdot(org->tradial, V2(p.gradc.x, p.gradc.y));
    o->gradc.y = p.gradc.z;
    o->crz = org->tradial.x * p.cxz + org->tradial.y * p.cyz;
    o->crr = p.cxx * SQ(org->tradial.x) + FL(2.0) * p.cxy * org->tradial.x * org->tradial.y
        + p.cyy * SQ(org->tradial.y);
    o->ccpx = p.ccpx;
    o->rho = p.rho;
    o->czz = p.czz;
}

typedef struct {
    Seg3 iSeg;
    BState3 top, bot;
    v3 xs;
    OriginNx org;
    int iSmallStepCtr;
} TraceStateNx;

static void StepNx(RayPt ray0, RayPt *ray2, TraceStateNx *S, const Prob3 *Q, ErrState *err, int *topRefl,
                   int *botRefl)
{ /* step.hpp:654-811 with O3D && !R3D: 2D integration in ray coordinates,
   * step limits from the 3D boundaries and SSP cells in ocean coordinates */
    const OriginNx *org = &S->org;
    RayPt ray1;
    SSPOut o0, o1, o2;
    EvaluateSSPNx(ray0.x, ray0.t, &o0, org, Q, &S->iSeg, err);
    real c0 = creal(o0.ccpx);
    real cnn0 = o0.crr * SQ(ray0.t.y) - FL(2.0) * o0.crz * ray0.t.x * ray0.t.y + o0.czz * SQ(ray0.t.x);
    v2 pq0p = vscl(-cnn0, ray0.q), pq0q = vscl(c0, ray0.p);
    Seg3 iSeg0 = S->iSeg;
    real csq0 = SQ(c0);
    v2 urayt0 = vscl(c0, ray0.t);
    real h = Q->P2.deltas;
    v3 x_o = RayToOceanX(ray0.x, org);
    v3 t_o = RayToOceanT(urayt0, org);
    ReduceStep3(x_o, t_o, &iSeg0, &S->top, &S->bot, Q, S->xs, err, &h, &S->iSmallStepCtr);
    real halfh = FL(0.5) * h;
    ray1.x = vadd(ray0.x, vscl(halfh, urayt0));
    ray1.t = vsub(ray0.t, vdivs(vscl(halfh, o0.gradc), csq0));
    ray1.p = vadd(ray0.p, vscl(halfh, pq0p));
    ray1.q = vadd(ray0.q, vscl(halfh, pq0q));

    EvaluateSSPNx(ray1.x, ray1.t, &o1, org, Q, &S->iSeg, err);
    real c1 = creal(o1.ccpx);
    real cnn1 = o1.crr * SQ(ray1.t.y) - FL(2.0) * o1.crz * ray1.t.x * ray1.t.y + o1.czz * SQ(ray1.t.x);
    v2 pq1p = vscl(-cnn1, ray1.q), pq1q = vscl(c1, ray1.p);
    real csq1 = SQ(c1);
    v2 urayt1 = vscl(c1, ray1.t);
    t_o = RayToOceanT(urayt1, org);
    ReduceStep3(x_o, t_o, &iSeg0, &S->top, &S->bot, Q, S->xs, err, &h, &S->iSmallStepCtr);

    real w1 = h / (RL(2.0) * halfh);
    real w0 = RL(1.0) - w1;
    v2 urayt2 = vadd(vscl(w0, urayt0), vscl(w1, urayt1));
    t_o = RayToOceanT(urayt2, org);
    v3 x2_o;
    int snapDim = -1;
    StepToBdry3(x_o, &x2_o, t_o, &h, topRefl, botRefl, &iSeg0, &S->top, &S->bot, Q, S->xs, err, &snapDim);
    ray2->x = OceanToRayX(x2_o, org, urayt2, snapDim, err);
    real hw0 = h * w0, hw1 = h * w1;
    ray2->t = vsub(vsub(ray0.t, vdivs(vscl(hw0, o0.gradc), csq0)), vdivs(vscl(hw1, o1.gradc), csq1));
    ray2->tau = ray0.tau + hw0 / o0.ccpx + hw1 / o1.ccpx;
    ray2->p = vadd(ray0.p, vscl(hw0, pq0p));
    ray2->q = vadd(ray0.q, vscl(hw0, pq0q));
    ray2->p = vadd(ray2->p, vscl(hw1, pq1p));
    ray2->q = vadd(ray2->q, vscl(hw1, pq1q));
    ray2->Amp = ray0.Amp;
    ray2->Phase = ray0.Phase;
    ray2->NumTopBnc = ray0.NumTopBnc;
    ray2->NumBotBnc = ray0.NumBotBnc;

    EvaluateSSPNx(ray2->x, ray2->t, &o2, org, Q, &S->iSeg, err);
    ray2->c = creal(o2.ccpx);
    if(S->iSeg.z != iSeg0.z) { /* CurvatureCorrection, step.hpp:624-648 (depth crossings only) */
        v2 gj = vsub(o2.gradc, o0.gradc);
        v2 ray2n = V2(-ray2->t.y, ray2->t.x);
        real cnjump = vdot(gj, ray2n);
        real csjump = vdot(gj, ray2->t);
        real rm = ray2->t.x / ray2->t.y;
        real rn = rm * (FL(2.0) * cnjump - rm * csjump) / ray2->c;
        ray2->p = vsub(ray2->p, vscl(rn, ray2->q));
    }
}

static void ReflectNx(const RayPt *oldp, RayPt *newp, const HS *hs, int isTop, v3 nBdry, real kxx, real kxy,
                      real kyy, const OriginNx *org, const Prob3 *Q, Seg3 *iSeg, ErrState *err)
{ /* reflect.hpp:94-452 with O3D && !R3D */
    const Prob *P = &Q->P2;
    v2 nB = OceanToRayT(nBdry, org);
    nB = vscl(RL(1.0) / vlen(nB), nB);
    real Th = vdot(oldp->t, nB);
    v2 tB = vsub(oldp->t, vscl(Th, nB));
    tB = vscl(RL(1.0) / vlen(tB), tB);
    real Tg = vdot(oldp->t, tB);
    /* OceanToRayCurvature (reflect.hpp:94-114) */
    real kappa = kxx * SQ(org->tradial.x) + FL(2.0) * kxy * org->tradial.x * org->tradial.y + kyy * SQ(org->tradial.y);
    if(isTop) kappa = -kappa;
    newp->NumTopBnc = oldp->NumTopBnc + isTop;
    newp->NumBotBnc = oldp->NumBotBnc + (1 - isTop);
    newp->x = oldp->x;
    newp->t = vsub(oldp->t, vscl(FL(2.0) * Th, nB));
    SSPOut o;
    EvaluateSSPNx(oldp->x, newp->t, &o, org, Q, iSeg, err);
    real c = creal(o.ccpx);
    newp->c = c;
    newp->tau = oldp->tau;
    v2 rayt = vscl(c, oldp->t);
    v2 rayn = V2(-rayt.y, rayt.x);
    v2 rayt_t = vscl(c, newp->t);
    v2 rayn_t = V2(-(-rayt_t.y), -(rayt_t.x));
    real rn = FL(2.0) * kappa / SQ(c) / Th;
    real cnjump = -vdot(o.gradc, vsub(rayn_t, rayn));
    real csjump = -vdot(o.gradc, vsub(rayt_t, rayt));
    if(isTop) {
        cnjump = -cnjump;
        rn = -rn;
    }
    real rm = Tg / Th;
    rn += rm * (FL(2.0) * cnjump - rm * csjump) / SQ(c);
    if(P->Type[2] == 'D') rn *= FL(2.0);
    else if(P->Type[2] == 'Z') rn = FL(0.0);
    newp->p = vadd(oldp->p, vscl(rn, oldp->q));
    newp->q = oldp->q;

    if(hs->bc == 'R') {
        newp->Amp = oldp->Amp;
        newp->Phase = oldp->Phase;
    } else if(hs->bc == 'V') {
        newp->Amp = oldp->Amp;
        newp->Phase = oldp->Phase + REAL_PI;
    } else if(hs->bc == 'F') {
        ReflCoef RInt;
        RInt.theta = RadDeg * RABS(RATAN2(Th, Tg));
        if(RInt.theta > FL(90.0)) RInt.theta = FL(180.0) - RInt.theta;
        if(isTop) InterpolateReflectionCoefficient(&RInt, P->trc, P->ntrc, err);
        else InterpolateReflectionCoefficient(&RInt, P->brc, P->nbrc, err);
        newp->Amp = oldp->Amp * RInt.r;
        newp->Phase = oldp->Phase + RInt.phi;
    } else if(hs->bc == 'A' || hs->bc == 'G') {
        /* the O3D halfspace formula (reflect.hpp:279-299) */
        real omega = FL(2.0) * REAL_PI * P->freq0;
        cpx gk = omega * Tg;
        const cpx J = I;
        cpx g1sq = SQ(omega / c) - SQ(gk) - J * REAL_MINPOS;
        cpx g2sq = SQ(omega / hs->cP) - SQ(gk) - J * REAL_MINPOS;
        cpx g1 = CSQRT(-g1sq);
        cpx g2 = CSQRT(-g2sq);
        cpx Refl = (hs->rho * g1 - o.rho * g2) / (hs->rho * g1 + o.rho * g2);
        real aR = CABS(Refl);
        if(aR < FL(1.0e-5)) {
            newp->Amp = FL(0.0);
            newp->Phase = oldp->Phase;
        } else {
            newp->Amp = aR * oldp->Amp;
            newp->Phase = oldp->Phase + RATAN2(cimag(Refl), creal(Refl));
            if(P->Type[3] == 'S') {
                /* beam displacement and width change, reflect.hpp:383-442 (!R3D) */
                cpx ch = oldp->c / CCONJ(hs->cP);
                real co = oldp->t.x * oldp->c;
                real si = oldp->t.y * oldp->c;
                real ck = omega / oldp->c;
                cpx a = FL(2.0) * hs->rho * (FL(1.0) - SQ(ch));
                cpx b = SQ(co) - SQ(ch);
                cpx d = SQ(hs->rho) * SQ(si) + b;
                cpx sb = CSQRT(b);
                real cco = SQ(co);
                real ssi = SQ(si);
                cpx delta;
                if(RABS(si) >= REAL_EPSILON) delta = a * co / si / (ck * sb * d);
                else delta = FL(0.0);
                real pdelta = creal(delta) / (oldp->c / co);
                cpx cksbd = ck * sb * d;
                cpx ddelta = a * (cco / (cksbd * b) - (RL(1.0) + (cco / ssi)) / cksbd
                                  - FL(2.0) * SQ(co) * (SQ(hs->rho) - RL(1.0)) / (cksbd * d));
                real rddelta = -creal(ddelta);
                real sddelta = rddelta / RABS(rddelta);
                real theta_bot = RATAN(tB.y / tB.x);
                newp->x.x = newp->x.x + creal(delta) * RCOS(theta_bot);
                newp->x.y = newp->x.y + creal(delta) * RSIN(theta_bot);
                newp->tau = newp->tau + pdelta;
                newp->q = vadd(newp->q, vscl(sddelta * rddelta * si * creal(o.ccpx), oldp->p));
            }
        }
    } else {
        RunError(err, BHC_ERR_BOUNDARY_CONDITION_TYPE);
        newp->Amp = NAN;
        newp->Phase = NAN;
    }
}

static int RayInitNx(RayInit3Info *ri, RayPt *p0, TraceStateNx *S, const Prob3 *Q, ErrState *err)
{ /* trace.hpp:51-67, 76-266 with O3D && !R3D */
    const Prob *P = &Q->P2;
    if(ri->isz < 0 || ri->isz >= P->NSz || ri->ialpha < 0 || ri->ialpha >= P->Nalpha || ri->isx < 0
       || ri->isx >= Q->NSx || ri->isy < 0 || ri->isy >= Q->NSy || ri->ibeta < 0 || ri->ibeta >= Q->Nbeta) {
        RunError(err, BHC_ERR_RAYINIT);
        return 0;
    }
    ri->alpha = P->alpha[ri->ialpha];
    ri->SrcDeclAngle = RadDeg * ri->alpha;
    ri->beta = Q->beta[ri->ibeta];
    ri->SrcAzimAngle = RadDeg * ri->beta;
    S->iSeg.x = S->iSeg.y = S->iSeg.z = S->iSeg.r = 0;
    S->xs = V3(Q->Sx[ri->isx], Q->Sy[ri->isy], (real)P->Sz[ri->isz]);
    SSPOut3 o;
    EvaluateSSP3(S->xs, V3(FL(0.0), FL(0.0), FL(1.0)), &o, Q, &S->iSeg, err);
    S->org.xs = S->xs;
    S->org.tradial = V2(RCOS(ri->beta), RSIN(ri->beta));
    real c = creal(o.ccpx);
    int ibp = BinarySearchLEQ(P->sbp, P->nsbp, 2, 0, ri->SrcDeclAngle);
    ibp = imin(ibp, P->nsbp - 2);
    real s = (ri->SrcDeclAngle - P->sbp[2 * ibp]) / (P->sbp[2 * (ibp + 1)] - P->sbp[2 * ibp]);
    float Amp0 = (float)((FL(1.0) - s) * P->sbp[2 * ibp + 1] + s * P->sbp[2 * (ibp + 1) + 1]);
    if(P->RunType[0] == 'S') {
        float omega = (float)(FL(2.0) * REAL_PI * P->freq0);
        Amp0 = (float)(Amp0 * (sqrtf(FL(2.0)) * RABS(RSIN(omega / c * S->xs.z * RSIN(ri->alpha)))));
    }
    v2 tinit2 = V2(RCOS(ri->alpha), RSIN(ri->alpha));
    p0->x = V2(RL(0.0), S->xs.z);
    p0->c = c;
    p0->t = vdivs(tinit2, c);
    p0->tau = 0;
    p0->Amp = Amp0;
    p0->Phase = FL(0.0);
    p0->NumTopBnc = p0->NumBotBnc = 0;
    p0->p = V2(FL(1.0), FL(0.0));
    p0->q = V2(FL(0.0), FL(1.0));
    memset(&S->top, 0, sizeof(S->top));
    memset(&S->bot, 0, sizeof(S->bot));
    v3 t_o = RayToOceanT(p0->t, &S->org);
    GetBdrySeg3(S->xs, t_o, &S->top, &Q->ati, 1, 1, err);
    GetBdrySeg3(S->xs, t_o, &S->bot, &Q->bty, 0, 1, err);
    real dt, db;
    Distances3(S->xs, &S->top, &S->bot, &dt, &db);
    if(dt <= FL(0.0) || db <= FL(0.0)) {
        RunWarning(err, BHC_WARN_SOURCE_OUTSIDE_BOUNDARIES);
        return 0;
    }
    S->iSmallStepCtr = 0;
    return 1;
}

static int RayUpdateNx(const RayPt *p0, RayPt *p1, RayPt *p2, TraceStateNx *S, const Prob3 *Q, ErrState *err)
{ /* trace.hpp:274-381 with O3D && !R3D, rectilinear boundaries */
    int topRefl = 0, botRefl = 0;
    StepNx(*p0, p1, S, Q, err, &topRefl, &botRefl);
    v3 x_o = RayToOceanX(p1->x, &S->org), t_o = RayToOceanT(p1->t, &S->org);
    GetBdrySeg3(x_o, t_o, &S->top, &Q->ati, 1, 0, err);
    GetBdrySeg3(x_o, t_o, &S->bot, &Q->bty, 0, 0, err);
    if(topRefl || botRefl) {
        ReflBdry3 r = refl_bdry3(topRefl ? &Q->ati : &Q->bty, topRefl ? &S->top : &S->bot,
                                 topRefl ? &Q->P2.top : &Q->P2.bot, x_o);
        ReflectNx(p1, p2, &r.hs, topRefl, r.n, r.kappa_xx, r.kappa_xy, r.kappa_yy, &S->org, Q, &S->iSeg, err);
        return 1;
    }
    return 0;
}

static int RayTerminateNx(const RayPt *pt, int *Nsteps, int is, const TraceStateNx *S, int MaxPointsPerRay,
                          const Prob3 *Q, ErrState *err)
{ /* trace.hpp:391-477 with O3D && !R3D: the 3D box and grid extents on the ocean point */
    v3 x = RayToOceanX(pt->x, &S->org), t = RayToOceanT(pt->t, &S->org), xs = S->xs;
    int leftbox = RABS(x.x - xs.x) >= Q->box_x || RABS(x.y - xs.y) >= Q->box_y || RABS(x.z - RL(0.0)) >= Q->box_z;
    const Bdry3 *bt = &Q->bty, *at = &Q->ati;
    real minx = rmax(bt->bd[0].x.x, at->bd[0].x.x);
    real miny = rmax(bt->bd[0].x.y, at->bd[0].x.y);
    real maxx = rmin(bt->bd[(bt->nx - 1) * bt->ny].x.x, at->bd[(at->nx - 1) * at->ny].x.x);
    real maxy = rmin(bt->bd[bt->ny - 1].x.y, at->bd[at->ny - 1].x.y);
    int e0 = x.x < minx || x.y < miny;
    int eN = x.x > maxx || x.y > maxy;
    e0 = e0 || (x.x == minx && t.x < RL(0.0)) || (x.y == miny && t.y < RL(0.0));
    eN = eN || (x.x == maxx && t.x > RL(0.0)) || (x.y == maxy && t.y > RL(0.0));
    int small = S->iSmallStepCtr > 50;
    int lost = pt->Amp < FL(0.005);
    if(leftbox || lost || e0 || eN || small) {
        *Nsteps = is + 1;
        return 1;
    } else if(is >= MaxPointsPerRay - 3) {
        RunWarning(err, BHC_WARN_ONERAY_OUTOFMEMORY);
        *Nsteps = is;
        return 1;
    }
    return 0;
}

typedef struct {
    Outs *O;
    const Prob3 *Q;
    ErrState *err;
    _Atomic int32_t next;
    int32_t job_begin, job_end;
    RayStat3 *stats;
} FieldJobNx;

static void MainFieldModesNx(RayInit3Info *ri3, Outs *O, const Prob3 *Q, ErrState *err, RayStat3 *st)
{ /* trace.hpp:525-589 with O3D && !R3D: the 2D influence functions on the
   * bearing's receiver plane, field slot itheta = ibeta */
    const Prob *P = &Q->P2;
    TraceStateNx S;
    RayPt p0, p1, p2;
    InflRay inf;
    memset(&p2, 0, sizeof(p2));
    if(st) { st->nsteps = 0; st->ntop = st->nbot = 0; st->xend[0] = st->xend[1] = st->xend[2] = 0; }
    if(!RayInitNx(ri3, &p0, &S, Q, err)) return;
    RayInitInfo ri;
    memset(&ri, 0, sizeof(ri));
    ri.isz = ((ri3->isz * Q->NSx + ri3->isx) * Q->NSy + ri3->isy) * Q->Ntheta + ri3->ibeta;
    ri.ialpha = ri3->ialpha;
    ri.alpha = ri3->alpha;
    ri.SrcDeclAngle = ri3->SrcDeclAngle;
    ri.nx = 1;
    ri.isx = ri3->isx;
    ri.isy = ri3->isy;
    ri.isz3 = ri3->isz;
    ri.ibeta = ri3->ibeta;
    ri.SrcAzimAngle = ri3->SrcAzimAngle;
    Init_Influence(&inf, &p0, &ri, P, O->isGaussian, O->isRayCen);
    int is = 0, Nsteps = 0;
    for(;;) {
        if(HasErrored(err)) break;
        int two = RayUpdateNx(&p0, &p1, &p2, &S, Q, err);
        if(O->isRayCen) Step_InfluenceGeoRayCen(&p0, &p1, &inf, is, O);
        else Step_InfluenceGeoCart(&p0, &p1, &inf, is, O);
        ++is;
        if(two) {
            if(O->isRayCen) Step_InfluenceGeoRayCen(&p1, &p2, &inf, is, O);
            else Step_InfluenceGeoCart(&p1, &p2, &inf, is, O);
            p0 = p2;
            ++is;
        } else {
            p0 = p1;
        }
        if(RayTerminateNx(&p0, &Nsteps, is, &S, MaxN, Q, err)) break;
    }
    if(st) {
        v3 xe = RayToOceanX(p0.x, &S.org);
        st->nsteps = is;
        st->ntop = p0.NumTopBnc;
        st->nbot = p0.NumBotBnc;
        st->xend[0] = (double)xe.x;
        st->xend[1] = (double)xe.y;
        st->xend[2] = (double)xe.z;
    }
}

static void *field_workerNx(void *arg)
{
    FieldJobNx *J = (FieldJobNx *)arg;
    for(;;) {
        int32_t job = atomic_fetch_add(&J->next, 1);
        if(job >= J->job_end) break;
        RayInit3Info ri;
        if(!job_indices3(&ri, job, J->Q)) break;
        MainFieldModesNx(&ri, J->O, J->Q, J->err, J->stats ? &J->stats[job - J->job_begin] : NULL);
    }
    return NULL;
}

/* Exported: Nx2D field run (TL / arrivals / eigen hits) on jobs [job_begin,
 * job_end) (job_end < 0: all); env->RunType[5] must be '2'. out->field is
 * [NSz][NSx][NSy][Ntheta][NRz][NRr] complex floats; ray_xend is [rays][3]
 * (ocean coordinates of the last point). */
int SFX(oracle_field_runnx2d)(const bhc_env3d *env, bhc_field_out *out, int nthreads, int job_begin, int job_end)
{
    char rt0 = env->RunType[0], rt1 = env->RunType[1];
    if(env->RunType[5] != '2') return -1;
    if(rt0 != 'C' && rt0 != 'S' && rt0 != 'I' && rt0 != 'A' && rt0 != 'a' && rt0 != 'E' && rt0 != 'V' && rt0 != 'v')
        return -1;
    if(rt1 != 'G' && rt1 != '^' && rt1 != ' ' && rt1 != 'B' && rt1 != 'g' && rt1 != 'b') return -1;
    Prob3 Q;
    if(prob3_from_env(env, &Q)) return -1;
    Outs O;
    memset(&O, 0, sizeof(O));
    O.P = &Q.P2;
    O.field = out->field;
    O.arr = out->arr;
    O.narr = (_Atomic int32_t *)out->narr;
    O.max_narr = out->max_narr;
    O.allow_merging = out->allow_merging;
    O.hits = out->hits;
    O.neigen = (_Atomic int32_t *)out->neigen;
    O.eigen_memsize = out->eigen_memsize;
    int arr_run = rt0 == 'A' || rt0 == 'a' || rt0 == 'V' || rt0 == 'v';
    int eig_run = rt0 == 'E' || rt0 == 'V' || rt0 == 'v';
    if(arr_run && (!O.arr || !O.narr || O.max_narr < 1)) { prob3_free(&Q); return -1; }
    if(eig_run && (!O.hits || !O.neigen)) { prob3_free(&Q); return -1; }
    O.isGaussian = (rt1 == 'B' || rt1 == 'b');
    O.isRayCen = (rt1 == 'g' || rt1 == 'b');
    atomic_init(&O.contrib, 0);
    ErrState err;
    atomic_init(&err.error, 0u);
    atomic_init(&err.warning, 0u);
    atomic_init(&err.errCount, 0u);
    atomic_init(&err.warnCount, 0u);
    int nj = num_jobs3(&Q);
    if(job_end < 0 || job_end > nj) job_end = nj;
    if(job_begin < 0) job_begin = 0;
    int nrays = job_end - job_begin;
    RayStat3 *stats = NULL;
    if(out->ray_nsteps) stats = xcalloc(nrays > 0 ? nrays : 1, sizeof(RayStat3));
    FieldJobNx J;
    J.O = &O;
    J.Q = &Q;
    J.err = &err;
    atomic_init(&J.next, job_begin);
    J.job_begin = job_begin;
    J.job_end = job_end;
    J.stats = stats;
    nthreads = resolve_threads(nthreads);
    if(arr_run && O.allow_merging) nthreads = 1; /* src/mode/arr.hpp:79 */
    pthread_t *th = xcalloc(nthreads, sizeof(pthread_t));
    for(int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, field_workerNx, &J);
    for(int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th);
    if(stats) {
        int64_t tot = 0;
        for(int i = 0; i < nrays; ++i) {
            out->ray_nsteps[i] = stats[i].nsteps;
            if(out->ray_ntop) out->ray_ntop[i] = stats[i].ntop;
            if(out->ray_nbot) out->ray_nbot[i] = stats[i].nbot;
            if(out->ray_xend)
                for(int k = 0; k < 3; ++k) out->ray_xend[3 * i + k] = stats[i].xend[k];
            tot += stats[i].nsteps;
        }
        out->total_steps = tot;
        free(stats);
    }
    out->total_contrib = atomic_load(&O.contrib);
    out->err = atomic_load(&err.error);
    out->warn = atomic_load(&err.warning);
    out->err_count = atomic_load(&err.errCount);
    out->warn_count = atomic_load(&err.warnCount);
    prob3_free(&Q);
    return 0;
}

static void MainRayModeNx(RayInit3Info *ri, RayPt *ray, int *Nsteps, int MaxPointsPerRay, const Prob3 *Q,
                          ErrState *err, OriginNx *org)
{ /* trace.hpp:482-520 with O3D && !R3D */
    TraceStateNx S;
    if(!RayInitNx(ri, &ray[0], &S, Q, err)) {
        *Nsteps = 1;
        org->xs = V3(Q->Sx[ri->isx], Q->Sy[ri->isy], (real)Q->P2.Sz[ri->isz]);
        org->tradial = V2(RCOS(Q->beta[ri->ibeta]), RSIN(Q->beta[ri->ibeta]));
        return;
    }
    *org = S.org;
    int is = 0;
    for(;;) {
        if(HasErrored(err)) break;
        int two = RayUpdateNx(&ray[is], &ray[is + 1], &ray[is + 2], &S, Q, err);
        if(*Nsteps >= 0 && is >= *Nsteps) {
            *Nsteps = is + 2;
            break;
        }
        is += two ? 2 : 1;
        if(RayTerminateNx(&ray[is], Nsteps, is, &S, MaxPointsPerRay, Q, err)) break;
    }
}

/* Exported: the Nx2D ray run, every ray in job order (beams on the beta fan),
 * points in ray coordinates (range along the bearing, depth) and per ray the
 * origin (xs.x, xs.y, cos beta, sin beta) that maps them to the ocean. */
int SFX(oracle_ray_runnx2d)(const bhc_env3d *env, bhc_ray_pt2d *points, int32_t *nsteps, double *origin,
                            int max_points)
{
    if(env->RunType[5] != '2') return -1;
    Prob3 Q;
    if(prob3_from_env(env, &Q)) return -1;
    ErrState err;
    atomic_init(&err.error, 0u);
    atomic_init(&err.warning, 0u);
    atomic_init(&err.errCount, 0u);
    atomic_init(&err.warnCount, 0u);
    int nj = num_jobs3(&Q);
    RayPt *buf = xcalloc((size_t)max_points + 3, sizeof(RayPt));
    for(int k = 0; k < nj; ++k) {
        RayInit3Info ri;
        if(!job_indices3(&ri, k, &Q)) break;
        int N = -1;
        OriginNx org;
        memset(buf, 0, sizeof(RayPt) * ((size_t)max_points + 3));
        MainRayModeNx(&ri, buf, &N, max_points, &Q, &err, &org);
        nsteps[k] = N;
        if(origin) {
            origin[4 * k] = (double)org.xs.x;
            origin[4 * k + 1] = (double)org.xs.y;
            origin[4 * k + 2] = (double)org.tradial.x;
            origin[4 * k + 3] = (double)org.tradial.y;
        }
        int n = N < max_points ? N : max_points;
        for(int i = 0; i < n; ++i) store_pt(&points[(size_t)k * max_points + i], &buf[i]);
    }
    free(buf);
    prob3_free(&Q);
    return 0;
}

int SFX(oracle_preprocess_dump3d)(const bhc_env3d *env, double *dst, int64_t cap, int64_t *used)
{
    Prob3 Q;
    if(prob3_from_env(env, &Q)) return -1;
    int64_t k = 0;
#define PUT(v) do { if(k < cap) dst[k] = (double)(v); ++k; } while(0)
    const Bdry3 *Bs[2] = {&Q.bty, &Q.ati};
    for(int b = 0; b < 2; ++b)
        for(int i = 0; i < Bs[b]->nx * Bs[b]->ny; ++i) {
            const BNode3 *n = &Bs[b]->bd[i];
            PUT(n->x.x); PUT(n->x.y); PUT(n->x.z);
            PUT(n->n1.x); PUT(n->n1.y); PUT(n->n1.z);
            PUT(n->n2.x); PUT(n->n2.y); PUT(n->n2.z);
        }
    /* curvilinear / province extras of non-default grids */
    for(int b = 0; b < 2; ++b) {
        const Bdry3 *B = Bs[b];
        if(B->type[0] != 'C' && B->type[1] != 'L') continue;
        for(int i = 0; i < B->nx * B->ny; ++i) {
            const BNode3 *n = &B->bd[i];
            PUT(n->Noden.x); PUT(n->Noden.y); PUT(n->Noden.z);
            PUT(n->z_xx); PUT(n->z_xy); PUT(n->z_yy);
            PUT(n->kappa_xx); PUT(n->kappa_xy); PUT(n->kappa_yy);
            PUT(B->type[1] == 'L' ? n->prov : -1);
        }
        for(int i = 0; i < B->nprov; ++i) {
            PUT(creal(B->prov[i].cP)); PUT(cimag(B->prov[i].cP));
            PUT(creal(B->prov[i].cS)); PUT(cimag(B->prov[i].cS));
            PUT(B->prov[i].rho);
        }
    }
    if(Q.ssp_type == 'H')
        for(size_t i = 0; i < (size_t)Q.hx * Q.hy * (Q.hz - 1); ++i) PUT(Q.czmat[i]);
    for(int i = 0; i < Q.Ntheta; ++i) { PUT(Q.t_rcvr[i].x); PUT(Q.t_rcvr[i].y); }
    for(int i = 0; i < Q.Nbeta; ++i) PUT(Q.beta[i]);
    for(int i = 0; i < Q.NSx; ++i) PUT(Q.Sx[i]);
    for(int i = 0; i < Q.NSy; ++i) PUT(Q.Sy[i]);
    PUT(Q.P2.deltas); PUT(Q.box_x); PUT(Q.box_y); PUT(Q.box_z); PUT(Q.Dbeta);
#undef PUT
    *used = k;
    prob3_free(&Q);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Ray run type, src/trace.hpp:482-520 (3D)                                    */
/* ------------------------------------------------------------------------- */
static void store_pt3(bhc_ray_pt3d *d, const RayPt3 *s)
{
    d->p[0] = s->p.m[0][0]; d->p[1] = s->p.m[0][1]; d->p[2] = s->p.m[1][0]; d->p[3] = s->p.m[1][1];
    d->q[0] = s->q.m[0][0]; d->q[1] = s->q.m[0][1]; d->q[2] = s->q.m[1][0]; d->q[3] = s->q.m[1][1];
    d->phi = s->phi;
    d->dummy = 0.0;
    d->NumTopBnc = s->NumTopBnc;
    d->NumBotBnc = s->NumBotBnc;
    d->x[0] = s->x.x; d->x[1] = s->x.y; d->x[2] = s->x.z;
    d->t[0] = s->t.x; d->t[1] = s->t.y; d->t[2] = s->t.z;
    d->c = s->c; d->Amp = s->Amp; d->Phase = s->Phase;
    d->tau[0] = creal(s->tau); d->tau[1] = cimag(s->tau);
}

static void MainRayMode3(RayInit3Info *ri, RayPt3 *ray, int *Nsteps, int MaxPointsPerRay, const Prob3 *Q,
                         ErrState *err)
{
    TraceState3 S;
    if(!RayInit3(ri, &ray[0], &S, Q, err)) {
        *Nsteps = 1;
        return;
    }
    int is = 0;
    for(;;) {
        if(HasErrored(err)) break;
        int two = RayUpdate3(&ray[is], &ray[is + 1], &ray[is + 2], &S, Q, err);
        if(*Nsteps >= 0 && is >= *Nsteps) {
            *Nsteps = is + 2;
            break;
        }
        is += two ? 2 : 1;
        if(RayTerminate3(&ray[is], Nsteps, is, &S, MaxPointsPerRay, Q, err)) break;
    }
}

/* Exported: the 3D ray run, every ray in job order. */
int SFX(oracle_ray_run3d)(const bhc_env3d *env, bhc_ray_pt3d *points, int32_t *nsteps, int max_points)
{
    Prob3 Q;
    if(prob3_from_env(env, &Q)) return -1;
    ErrState err;
    atomic_init(&err.error, 0u);
    atomic_init(&err.warning, 0u);
    atomic_init(&err.errCount, 0u);
    atomic_init(&err.warnCount, 0u);
    int nj = num_jobs3(&Q);
    RayPt3 *buf = xcalloc((size_t)max_points + 3, sizeof(RayPt3));
    for(int k = 0; k < nj; ++k) {
        RayInit3Info ri;
        if(!job_indices3(&ri, k, &Q)) break;
        int N = -1;
        memset(buf, 0, sizeof(RayPt3) * ((size_t)max_points + 3));
        MainRayMode3(&ri, buf, &N, max_points, &Q, &err);
        nsteps[k] = N;
        int n = N < max_points ? N : max_points;
        for(int i = 0; i < n; ++i) store_pt3(&points[(size_t)k * max_points + i], &buf[i]);
    }
    free(buf);
    prob3_free(&Q);
    return 0;
}

/* Exported: the eigenray re-trace in 3D (EigenModePostWorker,
 * src/mode/eigen.cpp:24-48): ray k starts from hit k's source and angles and
 * stops after the hit's step (Nsteps = hit.is on entry). */
int SFX(oracle_ray_run3d_list)(const bhc_env3d *env, const bhc_eigen_hit *hits, int32_t nhits, bhc_ray_pt3d *points,
                               int32_t *nsteps, int max_points)
{
    Prob3 Q;
    if(prob3_from_env(env, &Q)) return -1;
    ErrState err;
    atomic_init(&err.error, 0u);
    atomic_init(&err.warning, 0u);
    atomic_init(&err.errCount, 0u);
    atomic_init(&err.warnCount, 0u);
    RayPt3 *buf = xcalloc((size_t)max_points + 3, sizeof(RayPt3));
    for(int k = 0; k < nhits; ++k) {
        RayInit3Info ri;
        memset(&ri, 0, sizeof(ri));
        ri.isx = hits[k].isx;
        ri.isy = hits[k].isy;
        ri.isz = hits[k].isz;
        ri.ialpha = hits[k].ialpha;
        ri.ibeta = hits[k].ibeta;
        int N = hits[k].is;
        memset(buf, 0, sizeof(RayPt3) * ((size_t)max_points + 3));
        MainRayMode3(&ri, buf, &N, max_points, &Q, &err);
        nsteps[k] = N;
        int n = N < max_points ? N : max_points;
        for(int i = 0; i < n; ++i) store_pt3(&points[(size_t)k * max_points + i], &buf[i]);
    }
    free(buf);
    prob3_free(&Q);
    return 0;
}
