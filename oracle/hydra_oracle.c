/*
 * hydra_oracle.c -- serial FP64 CPU restatement of the HydraPhase per-step
 * finite-volume hot path.  TEST INFRASTRUCTURE ONLY (see hydra_oracle.h).
 *
 * Every routine cites the reference file:line it restates; paths are
 * relative to HydraPhase_60_FPC_5mill/src/.  Expressions keep the
 * reference's left-to-right evaluation order, its single-precision literals
 * (1.0E-8, 1.0E-6, 0.1 and C23 = 2.00/3.00 are REAL*4 constants promoted to
 * double) and its known defects (the missing /wt4 in cTp2,
 * sr_derivative.F90:225; the -wg(ng) of the viscous BC53,
 * sr_viscousgp.F90:408).  Build with -O2 -ffp-contract=off.
 */
#include "hydra_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NEQ 10
#define F1E8 ((double)1.0e-8f)
#define F1E6 ((double)1.0e-6f)
#define F0P1 ((double)0.1f)

struct hyo_case {
    int nodes, neles, nghosts, ntot;
    double *x, *y, *z;
    int *nod;    /* 4 x neles */
    int *nc;     /* 4 x ntot, ghosts point to themselves */
    int *iblank; /* neles */
    int *gparent, *gtype, *gside;
    hyo_params p;
    /* geometry */
    double *sc;  /* 12 x neles, slot 3*f+d */
    double *xc, *yc, *zc, *vol, *dd; /* ntot */
    double *dl, *epssq;              /* neles */
    double xmax, ymax, zmax, ymin, zmin;
    /* state */
    double *cn, *cu, *cn1, *d2q; /* 10 x ntot */
    double *rhs, *d24;           /* 10 x neles */
    double *anu;                 /* 2 x ntot */
    double *q;                   /* HQ_N x ntot */
    double *grad;                /* 39 x ntot */
    double *phi;                 /* ntot */
    double *vis;                 /* 8 x neles */
    double dt, time;
    int total;
    /* overset donor lists (no reference counterpart, see hyo_set_donors) */
    int ndon;
    int *drecv, *dids; /* ndon, 4 x ndon (row r: dids[4r..4r+3]) */
    double *dw;        /* 4 x ndon */
};

#define CN(i, k) c->cn[(size_t)(k) * c->ntot + (i)]
#define CU(i, k) c->cu[(size_t)(k) * c->ntot + (i)]
#define CN1(i, k) c->cn1[(size_t)(k) * c->ntot + (i)]
#define D2Q(i, k) c->d2q[(size_t)(k) * c->ntot + (i)]
#define RHS(i, k) c->rhs[(size_t)(k) * c->neles + (i)]
#define D24(i, k) c->d24[(size_t)(k) * c->neles + (i)]
#define ANU(i, k) c->anu[(size_t)(k) * c->ntot + (i)]
#define Q(s, i) c->q[(size_t)(s) * c->ntot + (i)]
#define GR(s, d, i) c->grad[(size_t)(3 * (s) + (d)) * c->ntot + (i)]
#define VIS(s, i) c->vis[(size_t)(s) * c->neles + (i)]
#define NOD(i, k) c->nod[(size_t)(k) * c->neles + (i)]
#define NC(i, k) c->nc[(size_t)(k) * c->ntot + (i)]
#define SC(i, f, d) c->sc[(size_t)(3 * (f) + (d)) * c->neles + (i)]

#define UG(i) Q(HQ_UG, i)
#define VG(i) Q(HQ_VG, i)
#define WG(i) Q(HQ_WG, i)
#define PG(i) Q(HQ_PG, i)
#define PP(i) Q(HQ_PP, i)
#define TG(i) Q(HQ_TG, i)
#define PHIG(i) Q(HQ_PHIG, i)
#define UP(i) Q(HQ_UP, i)
#define VP(i) Q(HQ_VP, i)
#define WP(i) Q(HQ_WP, i)
#define PHIP(i) Q(HQ_PHIP, i)
#define TP(i) Q(HQ_TP, i)
#define ROP(i) Q(HQ_ROP, i)
#define ROG(i) Q(HQ_ROG, i)
#define EGT(i) Q(HQ_EGT, i)
#define EPT(i) Q(HQ_EPT, i)

static double dmin2(double a, double b) { return a < b ? a : b; }
static double dmax2(double a, double b) { return a > b ? a : b; }

/* mod_MatrixOps.F90:4-9 (arguments named column-wise as in the reference) */
static double det3(double a11, double a21, double a31, double a12, double a22,
                   double a32, double a13, double a23, double a33)
{
    return a11 * (a22 * a33 - a23 * a32) - a21 * (a12 * a33 - a13 * a32) +
           a31 * (a12 * a23 - a13 * a22);
}

/* mod_MatrixOps.F90:12-18 */
static double fvenkat(double delp, double delm, double epssq)
{
    return (((delp * delp + epssq) * delm + 2.0 * delm * delm * delp) /
            (delp * delp + 2.0 * delm * delm + delp * delm + epssq)) /
           delm;
}

/* main.F90:336-349 */
void hyo_derive_params(hyo_params *p)
{
    p->co2 = (double)(1.00f / 16.00f);
    p->co4 = (double)(1.00f / 1024.00f);
    p->c23 = (double)(2.00f / 3.00f);
    p->cmug = (double)11.848e-8f * sqrt(p->amolg);
    p->rrg = 287.0;
    p->gravity = 0.0;
    p->prp = 10.0;
    p->foursigmabyd = 4.0 * p->surfacetension / p->diap;
}

static void *xcalloc(size_t n, size_t s)
{
    void *r = calloc(n ? n : 1, s);
    if (!r) abort();
    return r;
}

/* face k (1..4) node triples of the geometry routine, sr_geometry.F90:149-158 */
static void geom_face_nodes(const hyo_case *c, int i, int f, int *n1, int *n2,
                            int *n3)
{
    switch (f) {
    case 1: *n1 = NOD(i, 1); *n2 = NOD(i, 2); *n3 = NOD(i, 3); break;
    case 2: *n1 = NOD(i, 3); *n2 = NOD(i, 2); *n3 = NOD(i, 0); break;
    case 3: *n1 = NOD(i, 1); *n2 = NOD(i, 3); *n3 = NOD(i, 0); break;
    default: *n1 = NOD(i, 2); *n2 = NOD(i, 1); *n3 = NOD(i, 0); break;
    }
}

/* sr_geometry.F90:112-347 (the MPI ghost-geometry exchange :349-436 is a
 * no-op on one rank) */
static void geometry(hyo_case *c)
{
    const double *x = c->x, *y = c->y, *z = c->z;
    c->xmax = -1.0e30; c->ymax = -1.0e30; c->zmax = -1.0e30;
    c->ymin = 1.0e30; c->zmin = 1.0e30;
    for (int n = 0; n < c->nodes; n++) {
        c->xmax = dmax2(c->xmax, x[n]);
        c->ymax = dmax2(c->ymax, y[n]);
        c->zmax = dmax2(c->zmax, z[n]);
        c->ymin = dmin2(c->ymin, y[n]);
        c->zmin = dmin2(c->zmin, z[n]);
    }
    for (int i = 0; i < c->neles; i++) {
        for (int f = 1; f <= 4; f++) {
            int n1, n2, n3;
            geom_face_nodes(c, i, f, &n1, &n2, &n3);
            double dxyA = (x[n1] - x[n2]) * (y[n1] + y[n2]);
            double dxyB = (x[n2] - x[n3]) * (y[n2] + y[n3]);
            double dxyC = (x[n3] - x[n1]) * (y[n3] + y[n1]);
            double dyzA = (y[n1] - y[n2]) * (z[n1] + z[n2]);
            double dyzB = (y[n2] - y[n3]) * (z[n2] + z[n3]);
            double dyzC = (y[n3] - y[n1]) * (z[n3] + z[n1]);
            double dzxA = (z[n1] - z[n2]) * (x[n1] + x[n2]);
            double dzxB = (z[n2] - z[n3]) * (x[n2] + x[n3]);
            double dzxC = (z[n3] - z[n1]) * (x[n3] + x[n1]);
            SC(i, f - 1, 0) = 0.5 * (dyzA + dyzB + dyzC);
            SC(i, f - 1, 1) = 0.5 * (dzxA + dzxB + dzxC);
            SC(i, f - 1, 2) = 0.5 * (dxyA + dxyB + dxyC);
        }
        int n1 = NOD(i, 0), n2 = NOD(i, 1), n3 = NOD(i, 2), n4 = NOD(i, 3);
        double x1 = x[n1], x2 = x[n2], x3 = x[n3], x4 = x[n4];
        double y1 = y[n1], y2 = y[n2], y3 = y[n3], y4 = y[n4];
        double z1 = z[n1], z2 = z[n2], z3 = z[n3], z4 = z[n4];
        c->xc[i] = 0.25 * (x1 + x2 + x3 + x4);
        c->yc[i] = 0.25 * (y1 + y2 + y3 + y4);
        c->zc[i] = 0.25 * (z1 + z2 + z3 + z4);
        double dl12 = sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2) + (z1 - z2) * (z1 - z2));
        double dl13 = sqrt((x1 - x3) * (x1 - x3) + (y1 - y3) * (y1 - y3) + (z1 - z3) * (z1 - z3));
        double dl14 = sqrt((x1 - x4) * (x1 - x4) + (y1 - y4) * (y1 - y4) + (z1 - z4) * (z1 - z4));
        double dl23 = sqrt((x2 - x3) * (x2 - x3) + (y2 - y3) * (y2 - y3) + (z2 - z3) * (z2 - z3));
        double dl42 = sqrt((x4 - x2) * (x4 - x2) + (y4 - y2) * (y4 - y2) + (z4 - z2) * (z4 - z2));
        double dl43 = sqrt((x4 - x3) * (x4 - x3) + (y4 - y3) * (y4 - y3) + (z4 - z3) * (z4 - z3));
        double d = dl12;
        d = dmin2(d, dl13); d = dmin2(d, dl14); d = dmin2(d, dl23);
        d = dmin2(d, dl42); d = dmin2(d, dl43);
        c->dl[i] = d;
        double v6 = det3(x2, x3, x4, y2, y3, y4, z2, z3, z4) -
                    det3(x1, x3, x4, y1, y3, y4, z1, z3, z4) +
                    det3(x1, x2, x4, y1, y2, y4, z1, z2, z4) -
                    det3(x1, x2, x3, y1, y2, y3, z1, z2, z3);
        c->vol[i] = fabs(v6) / 6.0;
        /* VK epssq=(0.3*dl)**3.0d0, sr_VK.F90:34; the reference build
         * evaluates the constant real power as (t*t)*t, no libm pow.  dl
         * never changes, so it is formed once here. */
        double t = 0.3 * c->dl[i];
        c->epssq[i] = t * t * t;
    }
    for (int g = 0; g < c->nghosts; g++) {
        int np = c->gparent[g], ng = c->neles + g, ns = c->gside[g];
        int n1, n2, n3;
        geom_face_nodes(c, np, ns, &n1, &n2, &n3);
        double xcf = (x[n1] + x[n2] + x[n3]) / 3.0;
        double ycf = (y[n1] + y[n2] + y[n3]) / 3.0;
        double zcf = (z[n1] + z[n2] + z[n3]) / 3.0;
        c->xc[ng] = 2.0 * xcf - c->xc[np];
        c->yc[ng] = 2.0 * ycf - c->yc[np];
        c->zc[ng] = 2.0 * zcf - c->zc[np];
        c->vol[ng] = c->vol[np];
        for (int k = 0; k < 4; k++) NC(ng, k) = ng;
    }
    /* sr_geometry.F90:295-347: dd(np) is overwritten in ghost-list order */
    for (int g = 0; g < c->nghosts; g++) {
        int np = c->gparent[g], nt = c->gtype[g];
        int n1 = NOD(np, 0), n2 = NOD(np, 1), n3 = NOD(np, 2), n4 = NOD(np, 3);
        if (nt == 3) {
            double xcen = (x[n1] + x[n2] + x[n3] + x[n4]) * 0.25;
            c->dd[np] = fabs(c->xmax - xcen);
        } else if (nt == 51) {
            double ycen = (y[n1] + y[n2] + y[n3] + y[n4]) * 0.25;
            c->dd[np] = ycen > 0.0 ? fabs(c->ymax - ycen) : fabs(c->ymin - ycen);
        } else if (nt == 53) {
            double zcen = (z[n1] + z[n2] + z[n3] + z[n4]) * 0.25;
            c->dd[np] = zcen > F0P1 ? fabs(c->zmax - zcen) : fabs(c->zmin - zcen);
        }
    }
}

hyo_case *hyo_create(int nodes, int neles, int nghosts, const double *x,
                     const double *y, const double *z, const int *nod,
                     const int *nc, const int *iblank, const int *gparent,
                     const int *gtype, const int *gside, const hyo_params *p)
{
    hyo_case *c = (hyo_case *)xcalloc(1, sizeof(*c));
    c->nodes = nodes; c->neles = neles; c->nghosts = nghosts;
    c->ntot = neles + nghosts;
    size_t nt = (size_t)c->ntot, ne = (size_t)neles;
    c->x = xcalloc(nodes, sizeof(double));
    c->y = xcalloc(nodes, sizeof(double));
    c->z = xcalloc(nodes, sizeof(double));
    memcpy(c->x, x, sizeof(double) * nodes);
    memcpy(c->y, y, sizeof(double) * nodes);
    memcpy(c->z, z, sizeof(double) * nodes);
    c->nod = xcalloc(4 * ne, sizeof(int));
    memcpy(c->nod, nod, sizeof(int) * 4 * ne);
    c->nc = xcalloc(4 * nt, sizeof(int));
    for (int k = 0; k < 4; k++)
        memcpy(c->nc + k * nt, nc + k * ne, sizeof(int) * ne);
    c->iblank = xcalloc(ne, sizeof(int));
    memcpy(c->iblank, iblank, sizeof(int) * ne);
    c->gparent = xcalloc(nghosts, sizeof(int));
    c->gtype = xcalloc(nghosts, sizeof(int));
    c->gside = xcalloc(nghosts, sizeof(int));
    memcpy(c->gparent, gparent, sizeof(int) * nghosts);
    memcpy(c->gtype, gtype, sizeof(int) * nghosts);
    memcpy(c->gside, gside, sizeof(int) * nghosts);
    c->p = *p;
    hyo_derive_params(&c->p);
    c->sc = xcalloc(12 * ne, sizeof(double));
    c->xc = xcalloc(nt, sizeof(double));
    c->yc = xcalloc(nt, sizeof(double));
    c->zc = xcalloc(nt, sizeof(double));
    c->vol = xcalloc(nt, sizeof(double));
    c->dd = xcalloc(nt, sizeof(double));
    c->dl = xcalloc(ne, sizeof(double));
    c->epssq = xcalloc(ne, sizeof(double));
    c->cn = xcalloc(NEQ * nt, sizeof(double));
    c->cu = xcalloc(NEQ * nt, sizeof(double));
    c->cn1 = xcalloc(NEQ * nt, sizeof(double));
    c->d2q = xcalloc(NEQ * nt, sizeof(double));
    c->rhs = xcalloc(NEQ * ne, sizeof(double));
    c->d24 = xcalloc(NEQ * ne, sizeof(double));
    c->anu = xcalloc(2 * nt, sizeof(double));
    c->q = xcalloc((size_t)HQ_N * nt, sizeof(double));
    c->grad = xcalloc(39 * nt, sizeof(double));
    c->phi = xcalloc(nt, sizeof(double));
    c->vis = xcalloc(HV_N * ne, sizeof(double));
    geometry(c);
    return c;
}

void hyo_destroy(hyo_case *c)
{
    if (!c) return;
    free(c->x); free(c->y); free(c->z); free(c->nod); free(c->nc);
    free(c->iblank); free(c->gparent); free(c->gtype); free(c->gside);
    free(c->sc); free(c->xc); free(c->yc); free(c->zc); free(c->vol);
    free(c->dd); free(c->dl); free(c->epssq); free(c->cn); free(c->cu);
    free(c->cn1); free(c->d2q); free(c->rhs); free(c->d24); free(c->anu);
    free(c->q); free(c->grad); free(c->phi); free(c->vis);
    free(c->drecv); free(c->dids); free(c->dw);
    free(c);
}

/* main.F90:427-482: restart read, then cn = cu */
void hyo_set_state(hyo_case *c, const double *cu, const double *gp,
                   double time, int total)
{
    for (int k = 0; k < NEQ; k++)
        for (int i = 0; i < c->neles; i++)
            CU(i, k) = cu[(size_t)k * c->neles + i];
    static const int slot[6] = {HQ_UG, HQ_VG, HQ_WG, HQ_PG, HQ_TG, HQ_ROG};
    for (int s = 0; s < 6; s++)
        for (int g = 0; g < c->nghosts; g++)
            Q(slot[s], c->neles + g) = gp[(size_t)s * c->nghosts + g];
    memset(c->vis, 0, sizeof(double) * HV_N * c->neles);
    memset(c->d24, 0, sizeof(double) * NEQ * c->neles);
    for (int k = 0; k < NEQ; k++)
        for (int i = 0; i < c->ntot; i++) CN(i, k) = CU(i, k);
    c->time = time;
    c->total = total;
}

/* main.F90:812-824: restart write */
void hyo_get_state(const hyo_case *c, double *cu, double *gp, double *time,
                   int *total)
{
    if (cu)
        for (int k = 0; k < NEQ; k++)
            for (int i = 0; i < c->neles; i++)
                cu[(size_t)k * c->neles + i] = CU(i, k);
    static const int slot[6] = {HQ_UG, HQ_VG, HQ_WG, HQ_PG, HQ_TG, HQ_ROG};
    if (gp)
        for (int s = 0; s < 6; s++)
            for (int g = 0; g < c->nghosts; g++)
                gp[(size_t)s * c->nghosts + g] = Q(slot[s], c->neles + g);
    if (time) *time = c->time;
    if (total) *total = c->total;
}

/* Interior EOS inversion + phase clipping/blending + conserved rewrite:
 * sr_primitive.F90:23-161 and sr_primitiveplusbc.F90:36-172 (the latter
 * also stores ag, ap before clipping, :109-110). */
static void eos_cell(hyo_case *c, int i, int with_sound)
{
    const hyo_params *P = &c->p;
    double ug = CN(i, 1) / CN(i, 0), vg = CN(i, 2) / CN(i, 0), wg = CN(i, 3) / CN(i, 0);
    double up = CN(i, 6) / CN(i, 5), vp = CN(i, 7) / CN(i, 5), wp = CN(i, 8) / CN(i, 5);
    double eg = CN(i, 4) / CN(i, 0) - 0.5 * (ug * ug + vg * vg + wg * wg);
    double ep = CN(i, 9) / CN(i, 5) - 0.5 * (up * up + vp * vp + wp * wp);
    double bbb = (P->foursigmabyd + P->pinf * P->gamap) - (CN(i, 0) * (P->gamag - 1.0) * eg) -
                 (CN(i, 5) * (P->gamap - 1.0) * ep);
    double ccc = (CN(i, 0) * (P->gamag - 1.0) * eg * (P->foursigmabyd + P->pinf * P->gamap));
    double pg = (-bbb + sqrt(bbb * bbb + 4.0 * ccc)) * 0.5;
    double pp = pg + P->foursigmabyd;
    double phig = (CN(i, 0) * (P->gamag - 1.0) * eg) / pg;
    double phip = 1.0 - phig;
    double rog = CN(i, 0) / phig;
    double rop = CN(i, 5) / phip;
    double tg = pg / (rog * P->rrg);
    double tp = ((pp + P->pinf) * P->gamap) / ((P->gamap - 1.0) * rop * P->cpp);
    if (with_sound) {
        Q(HQ_AG, i) = sqrt((P->gamag * pg) / rog);
        Q(HQ_AP, i) = sqrt((P->gamap * (pp + P->pinf)) / rop);
    }
    if (phig < P->epslnmin) {
        phig = P->epslnmin; phip = 1.0 - phig;
        ug = up; vg = vp; wg = wp; tg = tp; pg = pp;
    }
    if (phip < P->epslnmin) {
        phip = P->epslnmin; phig = 1.0 - phip;
        up = ug; vp = vg; wp = wg; tp = tg; pp = pg;
    }
    if (phig >= P->epslnmin && phig <= P->epslnmax) {
        double psig = (phig - P->epslnmin) / (P->epslnmax - P->epslnmin);
        double fg = -psig * psig * (2.0 * psig - 3.0);
        ug = fg * ug + (1.0 - fg) * up;
        vg = fg * vg + (1.0 - fg) * vp;
        wg = fg * wg + (1.0 - fg) * wp;
        tg = fg * tg + (1.0 - fg) * tp;
        pg = fg * pg + (1.0 - fg) * pp;
    }
    if (phip >= P->epslnmin && phip <= P->epslnmax) {
        double psig = (phip - P->epslnmin) / (P->epslnmax - P->epslnmin);
        double fg = -psig * psig * (2.0 * psig - 3.0);
        up = fg * up + (1.0 - fg) * ug;
        vp = fg * vp + (1.0 - fg) * vg;
        wp = fg * wp + (1.0 - fg) * wg;
        tp = fg * tp + (1.0 - fg) * tg;
        pp = fg * pp + (1.0 - fg) * pg;
    }
    rog = pg / (P->rrg * tg);
    rop = (pp + P->pinf) * P->gamap / ((P->gamap - 1.0) * P->cpp * tp);
    eg = pg / ((P->gamag - 1.0) * rog);
    ep = (pp + P->gamap * P->pinf) / ((P->gamap - 1.0) * rop);
    double egt = eg + 0.5 * (ug * ug + vg * vg + wg * wg);
    double ept = ep + 0.5 * (up * up + vp * vp + wp * wp);
    UG(i) = ug; VG(i) = vg; WG(i) = wg; PG(i) = pg; PP(i) = pp; TG(i) = tg;
    PHIG(i) = phig; UP(i) = up; VP(i) = vp; WP(i) = wp; PHIP(i) = phip;
    TP(i) = tp; ROP(i) = rop; ROG(i) = rog; EGT(i) = egt; EPT(i) = ept;
    CN(i, 0) = phig * rog;
    CN(i, 1) = phig * rog * ug;
    CN(i, 2) = phig * rog * vg;
    CN(i, 3) = phig * rog * wg;
    CN(i, 4) = phig * rog * egt;
    CN(i, 5) = phip * rop;
    CN(i, 6) = phip * rop * up;
    CN(i, 7) = phip * rop * vp;
    CN(i, 8) = phip * rop * wp;
    CN(i, 9) = phip * rop * ept;
}

/* sr_primitive.F90:18-178 (partition ghosts do not exist on one rank) */
void hyo_primitive(hyo_case *c)
{
    for (int i = 0; i < c->neles; i++) eos_cell(c, i, 0);
}

/* Boundary-ghost gas/liquid state of sr_primitiveplusbc.F90:198-592 */
void hyo_bc(hyo_case *c)
{
    const hyo_params *P = &c->p;
    double dt = c->dt;
    /* type 2: subsonic inflow, :204-265 */
    for (int g = 0; g < c->nghosts; g++) {
        if (c->gtype[g] != 2) continue;
        int np = c->gparent[g], ng = c->neles + g;
        PHIP(ng) = P->phipzero;
        PHIG(ng) = 1.0 - PHIP(ng);
        UG(ng) = P->ugmin; VG(ng) = P->vgmin; WG(ng) = P->wgmin;
        TG(ng) = P->tgmin;
        PG(ng) = PG(np);
        ROG(ng) = PG(ng) / (P->rrg * TG(ng));
        UP(ng) = 0.99 * P->ugmin; VP(ng) = 0.99 * P->vgmin; WP(ng) = 0.99 * P->wgmin;
        TP(ng) = P->tgmin;
        PP(ng) = PP(np);
        ROP(ng) = P->gamap * (PG(ng) + P->pinf) / ((P->gamap - 1.0) * P->cpp * TP(ng));
        double eg = PG(ng) / (ROG(ng) * (P->gamag - 1.0));
        double ep = (PP(ng) + P->gamap * P->pinf) / (ROP(ng) * (P->gamap - 1.0));
        EGT(ng) = eg + 0.5 * (UG(ng) * UG(ng) + VG(ng) * VG(ng) + WG(ng) * WG(ng));
        EPT(ng) = ep + 0.5 * (UP(ng) * UP(ng) + VP(ng) * VP(ng) + WP(ng) * WP(ng));
    }
    /* type 3: relaxing subsonic outflow, :278-351 */
    for (int g = 0; g < c->nghosts; g++) {
        if (c->gtype[g] != 3) continue;
        int np = c->gparent[g], ng = c->neles + g;
        PHIG(ng) = PHIG(np);
        PHIP(ng) = PHIP(np);
        double wT = UG(np);
        double wU = UG(np) + sqrt(P->gamag * PG(np) / ROG(np));
        double alphaT = wT * dt / c->dd[np];
        double alphaU = wU * dt / c->dd[np];
        double fracT = 1.0 / (1.0 + alphaT);
        double fracU = 1.0 / (1.0 + alphaU);
        if (c->total == 1) {
            UG(ng) = UG(np); VG(ng) = VG(np); WG(ng) = WG(np); TG(ng) = TG(np);
            PG(ng) = P->pmin;
        }
        if (UG(np) > UG(ng)) {
            UG(ng) = UG(ng) * fracU + (1.0 - fracU) * (2.0 * UG(np) - UG(ng));
            VG(ng) = VG(ng) * fracT + (1.0 - fracT) * (2.0 * VG(np) - VG(ng));
            WG(ng) = WG(ng) * fracT + (1.0 - fracT) * (2.0 * WG(np) - WG(ng));
            TG(ng) = TG(ng) * fracT + (1.0 - fracT) * (2.0 * TG(np) - TG(ng));
        } else {
            UG(ng) = UG(np); VG(ng) = VG(np); WG(ng) = WG(np); TG(ng) = TG(np);
        }
        PG(ng) = P->pmin;
        ROG(ng) = PG(ng) / (P->rrg * TG(ng));
        VP(ng) = VP(np); UP(ng) = UP(np); WP(ng) = WP(np);
        PP(ng) = P->pmin;
        TP(ng) = TP(np);
        ROP(ng) = P->gamap * (PP(ng) + P->pinf) / ((P->gamap - 1.0) * P->cpp * TP(ng));
        double eg = P->rrg * TG(ng) / (P->gamag - 1.0);
        double ep = P->cpp * TP(ng) / P->gamap + P->pinf / ROP(ng);
        EGT(ng) = eg + 0.5 * (UG(ng) * UG(ng) + VG(ng) * VG(ng) + WG(ng) * WG(ng));
        EPT(ng) = ep + 0.5 * (UP(ng) * UP(ng) + VP(ng) * VP(ng) + WP(ng) * WP(ng));
    }
    /* type 5: slip wall, :364-412 */
    for (int g = 0; g < c->nghosts; g++) {
        if (c->gtype[g] != 5) continue;
        int np = c->gparent[g], ng = c->neles + g;
        PHIG(ng) = PHIG(np);
        UG(ng) = -UG(np); VG(ng) = -VG(np); WG(ng) = -WG(np);
        TG(ng) = TG(np);
        PHIP(ng) = PHIP(np);
        UP(ng) = -UP(np); VP(ng) = -VP(np); WP(ng) = -WP(np);
        TP(ng) = TP(np);
        PG(ng) = PG(np); PP(ng) = PP(np);
        ROG(ng) = PG(ng) / (P->rrg * TG(ng));
        ROP(ng) = P->gamap * (PP(ng) + P->pinf) / ((P->gamap - 1.0) * P->cpp * TP(ng));
        double eg = P->rrg * TG(ng) / (P->gamag - 1.0);
        double ep = P->cpp * TP(ng) / P->gamap + P->pinf / ROP(ng);
        EGT(ng) = eg + 0.5 * (UG(ng) * UG(ng) + VG(ng) * VG(ng) + WG(ng) * WG(ng));
        EPT(ng) = ep + 0.5 * (UP(ng) * UP(ng) + VP(ng) * VP(ng) + WP(ng) * WP(ng));
    }
    /* type 51: y faces, :425-508 */
    for (int g = 0; g < c->nghosts; g++) {
        if (c->gtype[g] != 51) continue;
        int np = c->gparent[g], ng = c->neles + g;
        double wT = VG(np);
        double wU = VG(np) + sqrt(P->gamag * PG(np) / ROG(np));
        double alphaT = wT * dt / c->dd[np];
        double alphaU = wU * dt / c->dd[np];
        double fracT = 1.0 / (1.0 + alphaT);
        (void)alphaU;
        if (VG(np) > VG(ng)) {
            VG(ng) = -VG(np);
            UG(ng) = UG(ng) * fracT + (1.0 - fracT) * (2.0 * UG(np) - UG(ng));
            WG(ng) = WG(ng) * fracT + (1.0 - fracT) * (2.0 * WG(np) - WG(ng));
            TG(ng) = TG(ng) * fracT + (1.0 - fracT) * (2.0 * TG(np) - TG(ng));
        } else {
            UG(ng) = UG(np); VG(ng) = -VG(np); WG(ng) = WG(np); TG(ng) = TG(np);
        }
        PG(ng) = PG(np);
        PHIG(ng) = PHIG(np);
        PHIP(ng) = PHIP(np);
        UP(ng) = UP(np); VP(ng) = -VP(np); WP(ng) = WP(np);
        TP(ng) = TP(np);
        PG(ng) = PG(np); PP(ng) = PP(np);
        ROG(ng) = PG(ng) / (P->rrg * TG(ng));
        ROP(ng) = P->gamap * (PP(ng) + P->pinf) / ((P->gamap - 1.0) * P->cpp * TP(ng));
        double eg = P->rrg * TG(ng) / (P->gamag - 1.0);
        double ep = P->cpp * TP(ng) / P->gamap + P->pinf / ROP(ng);
        EGT(ng) = eg + 0.5 * (UG(ng) * UG(ng) + VG(ng) * VG(ng) + WG(ng) * WG(ng));
        EPT(ng) = ep + 0.5 * (UP(ng) * UP(ng) + VP(ng) * VP(ng) + WP(ng) * WP(ng));
    }
    /* type 53: z faces, :520-591 */
    for (int g = 0; g < c->nghosts; g++) {
        if (c->gtype[g] != 53) continue;
        int np = c->gparent[g], ng = c->neles + g;
        double wT = WG(np);
        double alphaT = wT * dt / c->dd[np];
        double fracT = 1.0 / (1.0 + alphaT);
        if (WG(np) > WG(ng)) {
            WG(ng) = -WG(np);
            UG(ng) = UG(ng) * fracT + (1.0 - fracT) * (2.0 * UG(np) - UG(ng));
            VG(ng) = VG(ng) * fracT + (1.0 - fracT) * (2.0 * VG(np) - VG(ng));
            TG(ng) = TG(ng) * fracT + (1.0 - fracT) * (2.0 * TG(np) - TG(ng));
        } else {
            UG(ng) = UG(np); VG(ng) = VG(np); WG(ng) = -WG(np); TG(ng) = TG(np);
        }
        PG(ng) = PG(np);
        PHIG(ng) = PHIG(np);
        PHIP(ng) = PHIP(np);
        UP(ng) = UP(np); VP(ng) = VP(np); WP(ng) = -WP(np);
        TP(ng) = TP(np);
        PG(ng) = PG(np); PP(ng) = PP(np);
        ROG(ng) = PG(ng) / (P->rrg * TG(ng));
        ROP(ng) = P->gamap * (PP(ng) + P->pinf) / ((P->gamap - 1.0) * P->cpp * TP(ng));
        double eg = P->rrg * TG(ng) / (P->gamag - 1.0);
        double ep = P->cpp * TP(ng) / P->gamap + P->pinf / ROP(ng);
        EGT(ng) = eg + 0.5 * (UG(ng) * UG(ng) + VG(ng) * VG(ng) + WG(ng) * WG(ng));
        EPT(ng) = ep + 0.5 * (UP(ng) * UP(ng) + VP(ng) * VP(ng) + WP(ng) * WP(ng));
    }
    /* ghost conserved variables, :602-631 */
    for (int g = 0; g < c->nghosts; g++) {
        int ng = c->neles + g;
        CN(ng, 0) = PHIG(ng) * ROG(ng);
        CN(ng, 1) = PHIG(ng) * ROG(ng) * UG(ng);
        CN(ng, 2) = PHIG(ng) * ROG(ng) * VG(ng);
        CN(ng, 3) = PHIG(ng) * ROG(ng) * WG(ng);
        CN(ng, 4) = PHIG(ng) * ROG(ng) * EGT(ng);
        CN(ng, 5) = PHIP(ng) * ROP(ng);
        CN(ng, 6) = PHIP(ng) * ROP(ng) * UP(ng);
        CN(ng, 7) = PHIP(ng) * ROP(ng) * VP(ng);
        CN(ng, 8) = PHIP(ng) * ROP(ng) * WP(ng);
        CN(ng, 9) = PHIP(ng) * ROP(ng) * EPT(ng);
    }
}

/* sr_primitiveplusbc.F90:1-635 */
void hyo_primitiveplusbc(hyo_case *c)
{
    for (int i = 0; i < c->neles; i++) eos_cell(c, i, 1);
    hyo_bc(c);
}

/* main.F90:520-548 */
double hyo_dtmin(hyo_case *c)
{
    const hyo_params *P = &c->p;
    double dtmin = 1.0e8;
    for (int i = 0; i < c->neles; i++) {
        double ag = sqrt((P->gamag * PG(i)) / ROG(i));
        double ap = sqrt((P->gamap * (PP(i) + P->pinf)) / ROP(i));
        Q(HQ_AG, i) = ag;
        Q(HQ_AP, i) = ap;
        double qgi = sqrt(UG(i) * UG(i) + VG(i) * VG(i) + WG(i) * WG(i));
        double qpi = sqrt(UP(i) * UP(i) + VP(i) * VP(i) + WP(i) * WP(i));
        dtmin = dmin2(dtmin, c->dl[i] / (dmax2(ag, ap) + dmax2(qgi, qpi)));
    }
    return dtmin;
}

void hyo_set_dt(hyo_case *c, double dt, double time, int total)
{
    c->dt = dt;
    c->time = time;
    c->total = total;
}

double hyo_get_dt(const hyo_case *c) { return c->dt; }

/* sr_smooth.F90:19-374 */
void hyo_smooth(hyo_case *c)
{
    const hyo_params *P = &c->p;
    const int ne = c->neles;
    for (int g = 0; g < c->nghosts; g++) {
        int ng = ne + g, np = c->gparent[g];
        PG(ng) = PG(np);
        PP(ng) = PP(np);
        for (int k = 0; k < NEQ; k++) CN(ng, k) = CN(np, k);
    }
    for (int i = 0; i < ne; i++) {
        int a = NC(i, 0), b = NC(i, 1), d = NC(i, 2), e = NC(i, 3);
        double anun, anud, anupg, anupp;
        anun = fabs(PG(a) - PG(i)) + fabs(PG(b) - PG(i)) + fabs(PG(d) - PG(i)) + fabs(PG(e) - PG(i));
        anud = fabs(PG(a) + PG(i)) + fabs(PG(b) + PG(i)) + fabs(PG(d) + PG(i)) + fabs(PG(e) + PG(i));
        anupg = anun / anud;
        anun = fabs(CN(a, 4) - CN(i, 4)) + fabs(CN(b, 4) - CN(i, 4)) + fabs(CN(d, 4) - CN(i, 4)) + fabs(CN(e, 4) - CN(i, 4));
        anud = fabs(CN(a, 4) + CN(i, 4)) + fabs(CN(b, 4) + CN(i, 4)) + fabs(CN(d, 4) + CN(i, 4)) + fabs(CN(e, 4) + CN(i, 4));
        ANU(i, 0) = anupg + anun / (anud + F1E6);
        anun = fabs(CN(a, 0) - CN(i, 0)) + fabs(CN(b, 0) - CN(i, 0)) + fabs(CN(d, 0) - CN(i, 0)) + fabs(CN(e, 0) - CN(i, 0));
        anud = fabs(CN(a, 0) + CN(i, 0)) + fabs(CN(b, 0) + CN(i, 0)) + fabs(CN(d, 0) + CN(i, 0)) + fabs(CN(e, 0) + CN(i, 0));
        ANU(i, 0) = ANU(i, 0) + anun / (anud + F1E6);
        anun = fabs(PP(a) - PP(i)) + fabs(PP(b) - PP(i)) + fabs(PP(d) - PP(i)) + fabs(PP(e) - PP(i));
        anud = fabs(PP(a) + PP(i)) + fabs(PP(b) + PP(i)) + fabs(PP(d) + PP(i)) + fabs(PP(e) + PP(i));
        anupp = anun / anud;
        anun = fabs(CN(a, 5) - CN(i, 5)) + fabs(CN(b, 5) - CN(i, 5)) + fabs(CN(d, 5) - CN(i, 5)) + fabs(CN(e, 5) - CN(i, 5));
        anud = fabs(CN(a, 5) + CN(i, 5)) + fabs(CN(b, 5) + CN(i, 5)) + fabs(CN(d, 5) + CN(i, 5)) + fabs(CN(e, 5) + CN(i, 5));
        ANU(i, 1) = anupp + anun / (anud + F1E6);
        anun = fabs(CN(a, 9) - CN(i, 9)) + fabs(CN(b, 9) - CN(i, 9)) + fabs(CN(d, 9) - CN(i, 9)) + fabs(CN(e, 9) - CN(i, 9));
        anud = fabs(CN(a, 9) + CN(i, 9)) + fabs(CN(b, 9) + CN(i, 9)) + fabs(CN(d, 9) + CN(i, 9)) + fabs(CN(e, 9) + CN(i, 9));
        ANU(i, 1) = ANU(i, 1) + anun / (anud + F1E6);
        for (int k = 0; k < NEQ; k++)
            D2Q(i, k) = CN(a, k) - CN(i, k) + CN(b, k) - CN(i, k) + CN(d, k) - CN(i, k) + CN(e, k) - CN(i, k);
    }
    for (int g = 0; g < c->nghosts; g++) {
        int ng = ne + g, np = c->gparent[g];
        for (int k = 0; k < NEQ; k++) D2Q(ng, k) = D2Q(np, k);
        ANU(ng, 0) = ANU(np, 0);
        ANU(ng, 1) = ANU(np, 1);
    }
    const double dt = c->dt;
    for (int i = 0; i < ne; i++) {
        int nb[4] = {NC(i, 0), NC(i, 1), NC(i, 2), NC(i, 3)};
        for (int set = 0; set < 2; set++) {
            double d2c[4], d4c[4];
            for (int f = 0; f < 4; f++) {
                double e2 = P->co2 * dmax2(ANU(i, set), ANU(nb[f], set));
                double e4 = dmax2(0.0, P->co4 - e2);
                double d24n = 0.5 * (c->vol[i] / dt + c->vol[nb[f]] / dt);
                d2c[f] = e2 * d24n;
                d4c[f] = e4 * d24n;
            }
            for (int k = 5 * set; k < 5 * set + 5; k++) {
                double d2 = d2c[0] * (CN(nb[0], k) - CN(i, k)) + d2c[1] * (CN(nb[1], k) - CN(i, k)) +
                            d2c[2] * (CN(nb[2], k) - CN(i, k)) + d2c[3] * (CN(nb[3], k) - CN(i, k));
                double d4 = d4c[0] * (D2Q(nb[0], k) - D2Q(i, k)) + d4c[1] * (D2Q(nb[1], k) - D2Q(i, k)) +
                            d4c[2] * (D2Q(nb[2], k) - D2Q(i, k)) + d4c[3] * (D2Q(nb[3], k) - D2Q(i, k));
                D24(i, k) = d2 - d4;
            }
        }
    }
}

/* Ghost state re-applied at the top of sr_viscousgp.F90:71-446.  Unlike the
 * primitiveplusbc copy it neither forms egtotal/eptotal nor rebuilds ghost
 * conserved variables, and BC53 flips the ghost's own wg (:408). */
void hyo_viscous_bc(hyo_case *c)
{
    const hyo_params *P = &c->p;
    double dt = c->dt;
    for (int g = 0; g < c->nghosts; g++) {
        if (c->gtype[g] != 2) continue;
        int np = c->gparent[g], ng = c->neles + g;
        PHIP(ng) = P->phipzero;
        PHIG(ng) = 1.0 - PHIP(ng);
        UG(ng) = P->ugmin; VG(ng) = P->vgmin; WG(ng) = P->wgmin;
        TG(ng) = P->tgmin;
        PG(ng) = PG(np);
        ROG(ng) = PG(ng) / (P->rrg * TG(ng));
        UP(ng) = 0.99 * P->ugmin; VP(ng) = 0.99 * P->vgmin; WP(ng) = 0.99 * P->wgmin;
        TP(ng) = P->tgmin;
        PP(ng) = PP(np);
        ROP(ng) = P->gamap * (PG(ng) + P->pinf) / ((P->gamap - 1.0) * P->cpp * TP(ng));
    }
    for (int g = 0; g < c->nghosts; g++) {
        if (c->gtype[g] != 3) continue;
        int np = c->gparent[g], ng = c->neles + g;
        PHIG(ng) = PHIG(np);
        PHIP(ng) = PHIP(np);
        double wT = UG(np);
        double wU = UG(np) + sqrt(P->gamag * PG(np) / ROG(np));
        double alphaT = wT * dt / c->dd[np];
        double alphaU = wU * dt / c->dd[np];
        double fracT = 1.0 / (1.0 + alphaT);
        double fracU = 1.0 / (1.0 + alphaU);
        if (c->total == 1) {
            UG(ng) = UG(np); VG(ng) = VG(np); WG(ng) = WG(np); TG(ng) = TG(np);
            PG(ng) = P->pmin;
        }
        if (UG(np) > UG(ng)) {
            UG(ng) = UG(ng) * fracU + (1.0 - fracU) * (2.0 * UG(np) - UG(ng));
            VG(ng) = VG(ng) * fracT + (1.0 - fracT) * (2.0 * VG(np) - VG(ng));
            WG(ng) = WG(ng) * fracT + (1.0 - fracT) * (2.0 * WG(np) - WG(ng));
            TG(ng) = TG(ng) * fracT + (1.0 - fracT) * (2.0 * TG(np) - TG(ng));
        } else {
            UG(ng) = UG(np); VG(ng) = VG(np); WG(ng) = WG(np); TG(ng) = TG(np);
        }
        PG(ng) = P->pmin;
        ROG(ng) = PG(ng) / (P->rrg * TG(ng));
        VP(ng) = VP(np); UP(ng) = UP(np); WP(ng) = WP(np);
        PP(ng) = P->pmin;
        TP(ng) = TP(np);
        ROP(ng) = P->gamap * (PP(ng) + P->pinf) / ((P->gamap - 1.0) * P->cpp * TP(ng));
    }
    for (int g = 0; g < c->nghosts; g++) {
        if (c->gtype[g] != 5) continue;
        int np = c->gparent[g], ng = c->neles + g;
        PHIG(ng) = PHIG(np);
        UG(ng) = -UG(np); VG(ng) = -VG(np); WG(ng) = -WG(np);
        TG(ng) = TG(np);
        PHIP(ng) = PHIP(np);
        UP(ng) = -UP(np); VP(ng) = -VP(np); WP(ng) = -WP(np);
        TP(ng) = TP(np);
        PG(ng) = PG(np); PP(ng) = PP(np);
        ROG(ng) = PG(ng) / (P->rrg * TG(ng));
        ROP(ng) = P->gamap * (PP(ng) + P->pinf) / ((P->gamap - 1.0) * P->cpp * TP(ng));
    }
    for (int g = 0; g < c->nghosts; g++) {
        if (c->gtype[g] != 51) continue;
        int np = c->gparent[g], ng = c->neles + g;
        double wT = VG(np);
        double alphaT = wT * dt / c->dd[np];
        double fracT = 1.0 / (1.0 + alphaT);
        if (VG(np) > VG(ng)) {
            VG(ng) = -VG(np);
            UG(ng) = UG(ng) * fracT + (1.0 - fracT) * (2.0 * UG(np) - UG(ng));
            WG(ng) = WG(ng) * fracT + (1.0 - fracT) * (2.0 * WG(np) - WG(ng));
            TG(ng) = TG(ng) * fracT + (1.0 - fracT) * (2.0 * TG(np) - TG(ng));
        } else {
            UG(ng) = UG(np); VG(ng) = -VG(np); WG(ng) = WG(np); TG(ng) = TG(np);
        }
        PG(ng) = PG(np);
        PHIG(ng) = PHIG(np);
        PHIP(ng) = PHIP(np);
        UP(ng) = UP(np); VP(ng) = -VP(np); WP(ng) = WP(np);
        TP(ng) = TP(np);
        PG(ng) = PG(np); PP(ng) = PP(np);
        ROG(ng) = PG(ng) / (P->rrg * TG(ng));
        ROP(ng) = P->gamap * (PP(ng) + P->pinf) / ((P->gamap - 1.0) * P->cpp * TP(ng));
    }
    for (int g = 0; g < c->nghosts; g++) {
        if (c->gtype[g] != 53) continue;
        int np = c->gparent[g], ng = c->neles + g;
        double wT = WG(np);
        double alphaT = wT * dt / c->dd[np];
        double fracT = 1.0 / (1.0 + alphaT);
        if (WG(np) > WG(ng)) {
            WG(ng) = -WG(ng);
            UG(ng) = UG(ng) * fracT + (1.0 - fracT) * (2.0 * UG(np) - UG(ng));
            VG(ng) = VG(ng) * fracT + (1.0 - fracT) * (2.0 * VG(np) - VG(ng));
            TG(ng) = TG(ng) * fracT + (1.0 - fracT) * (2.0 * TG(np) - TG(ng));
        } else {
            UG(ng) = UG(np); VG(ng) = VG(np); WG(ng) = -WG(np); TG(ng) = TG(np);
        }
        PG(ng) = PG(np);
        PHIG(ng) = PHIG(np);
        PHIP(ng) = PHIP(np);
        UP(ng) = UP(np); VP(ng) = VP(np); WP(ng) = -WP(np);
        TP(ng) = TP(np);
        PG(ng) = PG(np); PP(ng) = PP(np);
        ROG(ng) = PG(ng) / (P->rrg * TG(ng));
        ROP(ng) = P->gamap * (PP(ng) + P->pinf) / ((P->gamap - 1.0) * P->cpp * TP(ng));
    }
}

/* the 13 primitives the face LSQ fits: ug vg wg tg up vp wp tp */
#define NVLSQ 8
static const int vlsq_slot[NVLSQ] = {HQ_UG, HQ_VG, HQ_WG, HQ_TG, HQ_UP, HQ_VP, HQ_WP, HQ_TP};

/* sr_viscousgp.F90:37-61 and :450-1139 (AMUP_CONSTANT build, the Makefile
 * default, Makefile:18,67-69) */
static void viscous_faces(hyo_case *c)
{
    const hyo_params *P = &c->p;
    const int ne = c->neles;
    for (int i = 0; i < ne; i++) {
        if (c->iblank[i] != 1) continue;
        for (int s = 0; s < HV_N; s++) VIS(s, i) = 0.0;
    }
    for (int i = 0; i < ne; i++) {
        if (c->iblank[i] != 1) continue;
        int n1 = NOD(i, 0), n2 = NOD(i, 1), n3 = NOD(i, 2), n4 = NOD(i, 3);
        for (int f = 0; f < 4; f++) {
            double scx = SC(i, f, 0), scy = SC(i, f, 1), scz = SC(i, f, 2);
            int a, b, d;
            switch (f) {
            case 0: a = n2; b = n3; d = n4; break;
            case 1: a = n1; b = n3; d = n4; break;
            case 2: a = n1; b = n2; d = n4; break;
            default: a = n1; b = n2; d = n3; break;
            }
            double xf = (c->x[a] + c->x[b] + c->x[d]) / 3.0;
            double yf = (c->y[a] + c->y[b] + c->y[d]) / 3.0;
            double zf = (c->z[a] + c->z[b] + c->z[d]) / 3.0;
            int neigh = NC(i, f);
            double rogf = (ROG(i) + ROG(neigh)) * 0.5;
            double ugf = (UG(i) + UG(neigh)) * 0.5;
            double vgf = (VG(i) + VG(neigh)) * 0.5;
            double wgf = (WG(i) + WG(neigh)) * 0.5;
            double tgf = (TG(i) + TG(neigh)) * 0.5;
            double phigf = (PHIG(i) + PHIG(neigh)) * 0.5;
            double amug = P->as * sqrt(tgf) / (1.0 + P->ts / tgf);
            double akg = amug * P->cpg / P->prg;
            double ropf = (ROP(i) + ROP(neigh)) * 0.5;
            double upf = (UP(i) + UP(neigh)) * 0.5;
            double vpf = (VP(i) + VP(neigh)) * 0.5;
            double wpf = (WP(i) + WP(neigh)) * 0.5;
            double tpf = (TP(i) + TP(neigh)) * 0.5;
            double phipf = (PHIP(i) + PHIP(neigh)) * 0.5;
            double amup = 0.001;
            double akp = amup * P->cpp / P->prp;
            (void)rogf; (void)ropf;
            double fval[NVLSQ] = {ugf, vgf, wgf, tgf, upf, vpf, wpf, tpf};
            int m[8], nm;
            if (neigh >= ne) { /* :577-822, 5-point stencil */
                m[0] = NC(i, 0); m[1] = NC(i, 1); m[2] = NC(i, 2); m[3] = NC(i, 3); m[4] = i;
                nm = 5;
            } else { /* :823-1136, 8-point stencil */
                m[0] = NC(i, 0); m[1] = NC(i, 1); m[2] = NC(i, 2); m[3] = NC(i, 3);
                m[4] = NC(neigh, 0); m[5] = NC(neigh, 1); m[6] = NC(neigh, 2); m[7] = NC(neigh, 3);
                nm = 8;
            }
            double dx[8], dy[8], dz[8];
            for (int k = 0; k < nm; k++) {
                dx[k] = c->xc[m[k]] - xf;
                dy[k] = c->yc[m[k]] - yf;
                dz[k] = c->zc[m[k]] - zf;
            }
            double a11 = 0, a22 = 0, a33 = 0, a12 = 0, a13 = 0, a23 = 0;
            /* left-to-right sums; the first term starts the sum exactly */
            a11 = dx[0] * dx[0]; a22 = dy[0] * dy[0]; a33 = dz[0] * dz[0];
            a12 = dx[0] * dy[0]; a13 = dx[0] * dz[0]; a23 = dy[0] * dz[0];
            for (int k = 1; k < nm; k++) {
                a11 = a11 + dx[k] * dx[k];
                a22 = a22 + dy[k] * dy[k];
                a33 = a33 + dz[k] * dz[k];
                a12 = a12 + dx[k] * dy[k];
                a13 = a13 + dx[k] * dz[k];
                a23 = a23 + dy[k] * dz[k];
            }
            double gr[NVLSQ][3];
            double dnr = det3(a11, a12, a13, a12, a22, a23, a13, a23, a33);
            for (int v = 0; v < NVLSQ; v++) {
                double b1, b2, b3;
                double dv = Q(vlsq_slot[v], m[0]) - fval[v];
                b1 = dv * dx[0]; b2 = dv * dy[0]; b3 = dv * dz[0];
                for (int k = 1; k < nm; k++) {
                    dv = Q(vlsq_slot[v], m[k]) - fval[v];
                    b1 = b1 + dv * dx[k];
                    b2 = b2 + dv * dy[k];
                    b3 = b3 + dv * dz[k];
                }
                gr[v][0] = det3(b1, b2, b3, a12, a22, a23, a13, a23, a33) / dnr;
                gr[v][1] = det3(a11, a12, a13, b1, b2, b3, a13, a23, a33) / dnr;
                gr[v][2] = det3(a11, a12, a13, a12, a22, a23, b1, b2, b3) / dnr;
            }
            double dugx = gr[0][0], dugy = gr[0][1], dugz = gr[0][2];
            double dvgx = gr[1][0], dvgy = gr[1][1], dvgz = gr[1][2];
            double dwgx = gr[2][0], dwgy = gr[2][1], dwgz = gr[2][2];
            double dtgx = gr[3][0], dtgy = gr[3][1], dtgz = gr[3][2];
            double dupx = gr[4][0], dupy = gr[4][1], dupz = gr[4][2];
            double dvpx = gr[5][0], dvpy = gr[5][1], dvpz = gr[5][2];
            double dwpx = gr[6][0], dwpy = gr[6][1], dwpz = gr[6][2];
            double dtpx = gr[7][0], dtpy = gr[7][1], dtpz = gr[7][2];
            double delvg = dugx + dvgy + dwgz;
            double txxg = amug * (2.0 * dugx - P->c23 * delvg);
            double tyyg = amug * (2.0 * dvgy - P->c23 * delvg);
            double tzzg = amug * (2.0 * dwgz - P->c23 * delvg);
            double txyg = amug * (dvgx + dugy);
            double txzg = amug * (dwgx + dugz);
            double tyzg = amug * (dvgz + dwgy);
            double qxxg = ugf * txxg + vgf * txyg + wgf * txzg + akg * dtgx;
            double qyyg = ugf * txyg + vgf * tyyg + wgf * tyzg + akg * dtgy;
            double qzzg = ugf * txzg + vgf * tyzg + wgf * tzzg + akg * dtgz;
            double delvp = dupx + dvpy + dwpz;
            double txxp = amup * (2.0 * dupx - P->c23 * delvp);
            double tyyp = amup * (2.0 * dvpy - P->c23 * delvp);
            double tzzp = amup * (2.0 * dwpz - P->c23 * delvp);
            double txyp = amup * (dvpx + dupy);
            double txzp = amup * (dwpx + dupz);
            double tyzp = amup * (dvpz + dwpy);
            double qxxp = upf * txxp + vpf * txyp + wpf * txzp + akp * dtpx;
            double qyyp = upf * txyp + vpf * tyyp + wpf * tyzp + akp * dtpy;
            double qzzp = upf * txzp + vpf * tyzp + wpf * tzzp + akp * dtpz;
            VIS(HV_UG, i) = VIS(HV_UG, i) + phigf * (txxg * scx + txyg * scy + txzg * scz);
            VIS(HV_VG, i) = VIS(HV_VG, i) + phigf * (txyg * scx + tyyg * scy + tyzg * scz);
            VIS(HV_WG, i) = VIS(HV_WG, i) + phigf * (txzg * scx + tyzg * scy + tzzg * scz);
            VIS(HV_TG, i) = VIS(HV_TG, i) + phigf * (qxxg * scx + qyyg * scy + qzzg * scz);
            VIS(HV_UP, i) = VIS(HV_UP, i) + phipf * (txxp * scx + txyp * scy + txzp * scz);
            VIS(HV_VP, i) = VIS(HV_VP, i) + phipf * (txyp * scx + tyyp * scy + tyzp * scz);
            VIS(HV_WP, i) = VIS(HV_WP, i) + phipf * (txzp * scx + tyzp * scy + tzzp * scz);
            VIS(HV_TP, i) = VIS(HV_TP, i) + phipf * (qxxp * scx + qyyp * scy + qzzp * scz);
        }
    }
}

void hyo_viscousgp(hyo_case *c)
{
    /* :37-61 zeroing happens inside viscous_faces, before the BCs in the
     * reference; the two do not touch the same arrays */
    hyo_viscous_bc(c);
    viscous_faces(c);
}

/* sr_derivative.F90:45-322; scalar order as the reference's Cramer solves */
static const int deriv_slot[HG_N] = {HQ_UG, HQ_VG, HQ_WG, HQ_TG, HQ_UP, HQ_VP, HQ_WP,
                                     HQ_TP, HQ_PG, HQ_PP, HQ_ROG, HQ_ROP, HQ_PHIG};

void hyo_derivative(hyo_case *c)
{
    for (int i = 0; i < c->neles; i++) {
        if (c->iblank[i] != 1) continue;
        int m[4] = {NC(i, 0), NC(i, 1), NC(i, 2), NC(i, 3)};
        double dx[4], dy[4], dz[4], wt[4];
        for (int k = 0; k < 4; k++) {
            dx[k] = c->xc[m[k]] - c->xc[i];
            dy[k] = c->yc[m[k]] - c->yc[i];
            dz[k] = c->zc[m[k]] - c->zc[i];
            wt[k] = dx[k] * dx[k] + dy[k] * dy[k] + dz[k] * dz[k];
        }
        double a1 = dx[0] * dx[0] / wt[0] + dx[1] * dx[1] / wt[1] + dx[2] * dx[2] / wt[2] + dx[3] * dx[3] / wt[3];
        double a2 = dy[0] * dy[0] / wt[0] + dy[1] * dy[1] / wt[1] + dy[2] * dy[2] / wt[2] + dy[3] * dy[3] / wt[3];
        double a3 = dz[0] * dz[0] / wt[0] + dz[1] * dz[1] / wt[1] + dz[2] * dz[2] / wt[2] + dz[3] * dz[3] / wt[3];
        double b1 = dx[0] * dy[0] / wt[0] + dx[1] * dy[1] / wt[1] + dx[2] * dy[2] / wt[2] + dx[3] * dy[3] / wt[3];
        double b2 = dx[0] * dz[0] / wt[0] + dx[1] * dz[1] / wt[1] + dx[2] * dz[2] / wt[2] + dx[3] * dz[3] / wt[3];
        double b3 = dz[0] * dy[0] / wt[0] + dz[1] * dy[1] / wt[1] + dz[2] * dy[2] / wt[2] + dz[3] * dy[3] / wt[3];
        double dnr = det3(a1, b1, b2, b1, a2, b3, b2, b3, a3);
        for (int s = 0; s < HG_N; s++) {
            double qi = Q(deriv_slot[s], i);
            double dq[4];
            for (int k = 0; k < 4; k++) dq[k] = Q(deriv_slot[s], m[k]) - qi;
            double c1 = dq[0] * dx[0] / wt[0] + dq[1] * dx[1] / wt[1] + dq[2] * dx[2] / wt[2] + dq[3] * dx[3] / wt[3];
            double c2;
            if (s == HG_TP) /* sr_derivative.F90:225: last term lacks /wt4 */
                c2 = dq[0] * dy[0] / wt[0] + dq[1] * dy[1] / wt[1] + dq[2] * dy[2] / wt[2] + dq[3] * dy[3];
            else
                c2 = dq[0] * dy[0] / wt[0] + dq[1] * dy[1] / wt[1] + dq[2] * dy[2] / wt[2] + dq[3] * dy[3] / wt[3];
            double c3 = dq[0] * dz[0] / wt[0] + dq[1] * dz[1] / wt[1] + dq[2] * dz[2] / wt[2] + dq[3] * dz[3] / wt[3];
            GR(s, 0, i) = det3(c1, c2, c3, b1, a2, b3, b2, b3, a3) / dnr;
            GR(s, 1, i) = det3(a1, b1, b2, c1, c2, c3, b2, b3, a3) / dnr;
            GR(s, 2, i) = det3(a1, b1, b2, b1, a2, b3, c1, c2, c3) / dnr;
        }
    }
}

/* face-centroid node triples of the flux and limiter loops
 * (sr_hllcbatten.F90:79-115, sr_VK.F90:52-63), summed in this order and
 * scaled by 0.3333D0 */
static void flux_face_nodes(const hyo_case *c, int i, int f, int *n1, int *n2, int *n3)
{
    switch (f) {
    case 0: *n1 = NOD(i, 1); *n2 = NOD(i, 3); *n3 = NOD(i, 2); break;
    case 1: *n1 = NOD(i, 2); *n2 = NOD(i, 3); *n3 = NOD(i, 0); break;
    case 2: *n1 = NOD(i, 3); *n2 = NOD(i, 1); *n3 = NOD(i, 0); break;
    default: *n1 = NOD(i, 1); *n2 = NOD(i, 2); *n3 = NOD(i, 0); break;
    }
}

/* sr_VK.F90:32-301 */
void hyo_venkat(hyo_case *c)
{
    static const int vq[4] = {HQ_PG, HQ_PP, HQ_ROG, HQ_ROP};
    static const int vgs[4] = {HG_PG, HG_PP, HG_ROG, HG_ROP};
    for (int i = 0; i < c->neles; i++) {
        if (c->iblank[i] != 1) continue;
        double epssq = c->epssq[i];
        double x0 = c->xc[i], y0 = c->yc[i], z0 = c->zc[i];
        int ne[4] = {NC(i, 0), NC(i, 1), NC(i, 2), NC(i, 3)};
        double xf0[4], yf0[4], zf0[4];
        for (int f = 0; f < 4; f++) {
            int a, b, d;
            flux_face_nodes(c, i, f, &a, &b, &d);
            xf0[f] = (c->x[a] + c->x[b] + c->x[d]) * 0.3333 - x0;
            yf0[f] = (c->y[a] + c->y[b] + c->y[d]) * 0.3333 - y0;
            zf0[f] = (c->z[a] + c->z[b] + c->z[d]) * 0.3333 - z0;
        }
        double phv[4][4];
        for (int v = 0; v < 4; v++) {
            int s = vq[v], gs = vgs[v];
            double qi = Q(s, i);
            double vmax = qi, vmin = qi;
            for (int k = 0; k < 4; k++) {
                vmax = dmax2(vmax, Q(s, ne[k]));
                vmin = dmin2(vmin, Q(s, ne[k]));
            }
            for (int f = 0; f < 4; f++) {
                double delf = xf0[f] * GR(gs, 0, i) + yf0[f] * GR(gs, 1, i) + zf0[f] * GR(gs, 2, i);
                if (delf > 0.0)
                    phv[v][f] = fvenkat(vmax - qi, delf, epssq);
                else if (delf < 0.0)
                    phv[v][f] = fvenkat(vmin - qi, delf, epssq);
                else
                    phv[v][f] = 1.0;
            }
        }
        /* :269-272 */
        double ph = dmin2(dmin2(dmin2(phv[0][0], phv[0][1]), phv[0][2]), phv[0][3]);
        for (int v = 1; v < 4; v++)
            ph = dmin2(dmin2(dmin2(dmin2(ph, phv[v][0]), phv[v][1]), phv[v][2]), phv[v][3]);
        c->phi[i] = ph;
    }
    for (int g = 0; g < c->nghosts; g++) c->phi[c->neles + g] = c->phi[c->gparent[g]];
}

/* sr_source.F90:42-178 */
void hyo_source(hyo_case *c)
{
    const hyo_params *P = &c->p;
    for (int i = 0; i < c->neles; i++) {
        if (c->iblank[i] != 1) continue;
        double voli = c->vol[i];
        double amug = P->as * sqrt(TG(i)) / (1.0 + P->ts / TG(i));
        double dug = UG(i) - UP(i), dvg = VG(i) - VP(i), dwg = WG(i) - WP(i);
        double drag = 0.44 * PHIG(i) * (PHIG(i) - 1.0) * ROG(i) * sqrt(dug * dug + dvg * dvg + dwg * dwg) * c->dl[i];
        double fdgx = drag * (UG(i) - UP(i));
        double fdgy = drag * (VG(i) - VP(i));
        double fdgz = drag * (WG(i) - WP(i));
        double enu = 2.0;
        double qsource = 6.0 * PHIP(i) * amug * P->cpg * enu * (TP(i) - TG(i)) / (P->prg * (c->dl[i] * c->dl[i]));
        double delpstar = 2.0 * (PHIG(i) * ROG(i) * PHIP(i) * ROP(i)) / (PHIG(i) * ROP(i) + PHIP(i) * ROG(i)) *
                          (dug * dug + dvg * dvg + dwg * dwg);
        double pint = PG(i) - dmin2(delpstar, 0.01 * PG(i));
        double den = PHIG(i) * ROG(i) + (1.0 - PHIG(i)) * ROP(i);
        double uint = (PHIG(i) * ROG(i) * UG(i) + (1.0 - PHIG(i)) * ROP(i) * UP(i)) / den;
        double vint = (PHIG(i) * ROG(i) * VG(i) + (1.0 - PHIG(i)) * ROP(i) * VP(i)) / den;
        double wint = (PHIG(i) * ROG(i) * WG(i) + (1.0 - PHIG(i)) * ROP(i) * WP(i)) / den;
        int m[4] = {NC(i, 0), NC(i, 1), NC(i, 2), NC(i, 3)};
        double dx[4], dy[4], dz[4], wt[4], dph[4];
        for (int k = 0; k < 4; k++) {
            dx[k] = c->xc[m[k]] - c->xc[i];
            dy[k] = c->yc[m[k]] - c->yc[i];
            dz[k] = c->zc[m[k]] - c->zc[i];
            dph[k] = PHIG(m[k]) - PHIG(i);
            wt[k] = sqrt(dx[k] * dx[k] + dy[k] * dy[k] + dz[k] * dz[k]);
        }
        double a1 = dx[0] * dx[0] / wt[0] + dx[1] * dx[1] / wt[1] + dx[2] * dx[2] / wt[2] + dx[3] * dx[3] / wt[3];
        double a2 = dy[0] * dy[0] / wt[0] + dy[1] * dy[1] / wt[1] + dy[2] * dy[2] / wt[2] + dy[3] * dy[3] / wt[3];
        double a3 = dz[0] * dz[0] / wt[0] + dz[1] * dz[1] / wt[1] + dz[2] * dz[2] / wt[2] + dz[3] * dz[3] / wt[3];
        double b1 = dx[0] * dy[0] / wt[0] + dx[1] * dy[1] / wt[1] + dx[2] * dy[2] / wt[2] + dx[3] * dy[3] / wt[3];
        double b2 = dx[0] * dz[0] / wt[0] + dx[1] * dz[1] / wt[1] + dx[2] * dz[2] / wt[2] + dx[3] * dz[3] / wt[3];
        double b3 = dy[0] * dz[0] / wt[0] + dy[1] * dz[1] / wt[1] + dy[2] * dz[2] / wt[2] + dy[3] * dz[3] / wt[3];
        double c1 = dph[0] * dx[0] / wt[0] + dph[1] * dx[1] / wt[1] + dph[2] * dx[2] / wt[2] + dph[3] * dx[3] / wt[3];
        double c2 = dph[0] * dy[0] / wt[0] + dph[1] * dy[1] / wt[1] + dph[2] * dy[2] / wt[2] + dph[3] * dy[3] / wt[3];
        double c3 = dph[0] * dz[0] / wt[0] + dph[1] * dz[1] / wt[1] + dph[2] * dz[2] / wt[2] + dph[3] * dz[3] / wt[3];
        double dnr = det3(a1, b1, b2, b1, a2, b3, b2, b3, a3);
        double dphx = det3(c1, c2, c3, b1, a2, b3, b2, b3, a3) / dnr;
        double dphy = det3(a1, b1, b2, c1, c2, c3, b2, b3, a3) / dnr;
        double dphz = det3(a1, b1, b2, b1, a2, b3, c1, c2, c3) / dnr;
        double work = pint * (uint * dphx + vint * dphy + wint * dphz) + uint * fdgx + vint * fdgy + wint * fdgz + qsource;
        RHS(i, 1) = RHS(i, 1) - (fdgx + pint * dphx + PHIG(i) * ROG(i) * P->gravity) * voli - VIS(HV_UG, i);
        RHS(i, 2) = RHS(i, 2) - (fdgy + pint * dphy) * voli - VIS(HV_VG, i);
        RHS(i, 3) = RHS(i, 3) - (fdgz + pint * dphz) * voli - VIS(HV_WG, i);
        RHS(i, 4) = RHS(i, 4) - work * voli - VIS(HV_TG, i);
        RHS(i, 6) = RHS(i, 6) + (fdgx + pint * dphx - PHIP(i) * ROP(i) * P->gravity) * voli - VIS(HV_UP, i);
        RHS(i, 7) = RHS(i, 7) + (fdgy + pint * dphy) * voli - VIS(HV_VP, i);
        RHS(i, 8) = RHS(i, 8) + (fdgz + pint * dphz) * voli - VIS(HV_WP, i);
        RHS(i, 9) = RHS(i, 9) + work * voli - VIS(HV_TP, i);
    }
}

/* Left/right face states after MUSCL reconstruction and phase
 * clipping/blending (sr_hllcbatten.F90:135-318, sr_hllczien.F90:138-309) */
typedef struct {
    double rog, ug, vg, wg, tg, phig, rop, up, vp, wp, tp, pg, pp, phip;
} face_state;

static void clip_blend(const hyo_params *P, face_state *s)
{
    if (s->phig < P->epslnmin) {
        s->phig = P->epslnmin; s->phip = 1.0 - s->phig;
        s->ug = s->up; s->vg = s->vp; s->wg = s->wp; s->tg = s->tp; s->pg = s->pp;
    }
    if (s->phip < P->epslnmin) {
        s->phip = P->epslnmin; s->phig = 1.0 - s->phip;
        s->up = s->ug; s->vp = s->vg; s->wp = s->wg; s->tp = s->tg; s->pp = s->pg;
    }
    if (s->phig >= P->epslnmin && s->phig <= P->epslnmax) {
        double psig = (s->phig - P->epslnmin) / (P->epslnmax - P->epslnmin);
        double fg = -psig * psig * (2.0 * psig - 3.0);
        s->ug = fg * s->ug + (1.0 - fg) * s->up;
        s->vg = fg * s->vg + (1.0 - fg) * s->vp;
        s->wg = fg * s->wg + (1.0 - fg) * s->wp;
        s->tg = fg * s->tg + (1.0 - fg) * s->tp;
        s->pg = fg * s->pg + (1.0 - fg) * s->pp;
    }
    if (s->phip >= P->epslnmin && s->phip <= P->epslnmax) {
        double psig = (s->phip - P->epslnmin) / (P->epslnmax - P->epslnmin);
        double fg = -psig * psig * (2.0 * psig - 3.0);
        s->up = fg * s->up + (1.0 - fg) * s->ug;
        s->vp = fg * s->vp + (1.0 - fg) * s->vg;
        s->wp = fg * s->wp + (1.0 - fg) * s->wg;
        s->tp = fg * s->tp + (1.0 - fg) * s->tg;
        s->pp = fg * s->pp + (1.0 - fg) * s->pg;
    }
    s->rog = s->pg / (P->rrg * s->tg);
    s->rop = P->gamap * (s->pp + P->pinf) / ((P->gamap - 1.0) * P->cpp * s->tp);
}

static void load_state(const hyo_case *c, int i, face_state *s)
{
    s->rog = ROG(i); s->ug = UG(i); s->vg = VG(i); s->wg = WG(i); s->tg = TG(i);
    s->phig = PHIG(i); s->rop = ROP(i); s->up = UP(i); s->vp = VP(i); s->wp = WP(i);
    s->tp = TP(i); s->pg = PG(i); s->pp = PP(i);
}

#define GRD(s, i, x0, y0, z0) ((x0) * GR(s, 0, i) + (y0) * GR(s, 1, i) + (z0) * GR(s, 2, i))

/* HLLC-Batten flux with per-phase wave speeds, sr_hllcbatten.F90:338-529 */
static void riemann_batten(const hyo_params *P, const face_state *L, const face_state *R,
                           double enx, double eny, double enz, double area, double F[10])
{
    const double rrg = P->rrg, gamag = P->gamag, gamap = P->gamap, cpp = P->cpp, pinf = P->pinf;
    double EGL = rrg * L->tg / (gamag - 1.0) + 0.5 * (L->ug * L->ug + L->vg * L->vg + L->wg * L->wg);
    double HGL = EGL + L->pg / L->rog;
    double QGL = L->ug * enx + L->vg * eny + L->wg * enz;
    double AGL = sqrt(gamag * L->pg / L->rog);
    double EGR = rrg * R->tg / (gamag - 1.0) + 0.5 * (R->ug * R->ug + R->vg * R->vg + R->wg * R->wg);
    double HGR = EGR + R->pg / R->rog;
    double QGR = R->ug * enx + R->vg * eny + R->wg * enz;
    double AGR = sqrt(gamag * R->pg / R->rog);
    double EPL = (cpp * L->tp) / gamap + pinf / L->rop + 0.5 * (L->up * L->up + L->vp * L->vp + L->wp * L->wp);
    double HPL = EPL + L->pp / L->rop;
    double QPL = L->up * enx + L->vp * eny + L->wp * enz;
    double APL = sqrt(gamap * (L->pp + pinf) / L->rop);
    double EPR = (cpp * R->tp) / gamap + pinf / R->rop + 0.5 * (R->up * R->up + R->vp * R->vp + R->wp * R->wp);
    double HPR = EPR + R->pp / R->rop;
    double QPR = R->up * enx + R->vp * eny + R->wp * enz;
    double APR = sqrt(gamap * (R->pp + pinf) / R->rop);
    double SGL = dmin2(dmin2(QGL - AGL, QGR - AGR), F1E8);
    double SPL = dmin2(dmin2(QPL - APL, QPR - APR), F1E8);
    double SGR = dmax2(dmax2(QGL + AGL, QGR + AGR), F1E8);
    double SPR = dmax2(dmax2(QPL + APL, QPR + APR), F1E8);
    double SGSTARNR = R->pg - L->pg + L->rog * QGL * (SGL - QGL) - R->rog * QGR * (SGR - QGR);
    double SGSTARDR = L->rog * (SGL - QGL) - R->rog * (SGR - QGR);
    double SGSTAR = SGSTARNR / (SGSTARDR + F1E8);
    double SPSTARNR = R->pp - L->pp + L->rop * QPL * (SPL - QPL) - R->rop * QPR * (SPR - QPR);
    double SPSTARDR = L->rop * (SPL - QPL) - R->rop * (SPR - QPR);
    double SPSTAR = SPSTARNR / (SPSTARDR + F1E8);
    double PGSTAR = dmax2(L->pg + L->rog * (QGL - SGL) * (QGL - SGSTAR), F1E6 * dmin2(L->pg, R->pg));
    double PPSTAR = dmax2(L->pp + L->rop * (QPL - SPL) * (QPL - SPSTAR), F1E6 * dmin2(L->pp, R->pp));
    double CNL1 = L->phig * L->rog, CNL2 = L->phig * L->rog * L->ug, CNL3 = L->phig * L->rog * L->vg;
    double CNL4 = L->phig * L->rog * L->wg, CNL5 = L->phig * L->rog * EGL;
    double CNL6 = L->phip * L->rop, CNL7 = L->phip * L->rop * L->up, CNL8 = L->phip * L->rop * L->vp;
    double CNL9 = L->phip * L->rop * L->wp, CNL10 = L->phip * L->rop * EPL;
    double CNR1 = R->phig * R->rog, CNR2 = R->phig * R->rog * R->ug, CNR3 = R->phig * R->rog * R->vg;
    double CNR4 = R->phig * R->rog * R->wg, CNR5 = R->phig * R->rog * EGR;
    double CNR6 = R->phip * R->rop, CNR7 = R->phip * R->rop * R->up, CNR8 = R->phip * R->rop * R->vp;
    double CNR9 = R->phip * R->rop * R->wp, CNR10 = R->phip * R->rop * EPR;
    double DNRG = (SGSTAR - SGL) + F1E6;
    double CSL1 = (QGL - SGL) * CNL1 / DNRG;
    double CSL2 = ((QGL - SGL) * CNL2 + L->phig * (L->pg * enx - PGSTAR * enx)) / DNRG;
    double CSL3 = ((QGL - SGL) * CNL3 + L->phig * (L->pg * eny - PGSTAR * eny)) / DNRG;
    double CSL4 = ((QGL - SGL) * CNL4 + L->phig * (L->pg * enz - PGSTAR * enz)) / DNRG;
    double CSL5 = ((QGL - SGL) * CNL5 + L->phig * (L->pg * QGL - PGSTAR * SGSTAR)) / DNRG;
    double DNRP = (SPSTAR - SPL) + F1E6;
    double CSL6 = (QPL - SPL) * CNL6 / DNRP;
    double CSL7 = ((QPL - SPL) * CNL7 + L->phip * (L->pp * enx - PPSTAR * enx)) / DNRP;
    double CSL8 = ((QPL - SPL) * CNL8 + L->phip * (L->pp * eny - PPSTAR * eny)) / DNRP;
    double CSL9 = ((QPL - SPL) * CNL9 + L->phip * (L->pp * enz - PPSTAR * enz)) / DNRP;
    double CSL10 = ((QPL - SPL) * CNL10 + L->phip * (L->pp * QPL - PPSTAR * SPSTAR)) / DNRP;
    DNRG = (SGSTAR - SGR) + F1E6;
    double CSR1 = (QGR - SGR) * CNR1 / DNRG;
    double CSR2 = ((QGR - SGR) * CNR2 + R->phig * (R->pg * enx - PGSTAR * enx)) / DNRG;
    double CSR3 = ((QGR - SGR) * CNR3 + R->phig * (R->pg * eny - PGSTAR * eny)) / DNRG;
    double CSR4 = ((QGR - SGR) * CNR4 + R->phig * (R->pg * enz - PGSTAR * enz)) / DNRG;
    double CSR5 = ((QGR - SGR) * CNR5 + R->phig * (R->pg * QGR - PGSTAR * SGSTAR)) / DNRG;
    DNRP = (SPSTAR - SPR) + F1E6;
    double CSR6 = (QPR - SPR) * CNR6 / DNRP;
    double CSR7 = ((QPR - SPR) * CNR7 + R->phip * (R->pp * enx - PPSTAR * enx)) / DNRP;
    double CSR8 = ((QPR - SPR) * CNR8 + R->phip * (R->pp * eny - PPSTAR * eny)) / DNRP;
    double CSR9 = ((QPR - SPR) * CNR9 + R->phip * (R->pp * enz - PPSTAR * enz)) / DNRP;
    double CSR10 = ((QPR - SPR) * CNR10 + R->phip * (R->pp * QPR - PPSTAR * SPSTAR)) / DNRP;
    double FL1 = L->phig * L->rog * QGL;
    double FL2 = L->phig * L->rog * QGL * L->ug + L->phig * L->pg * enx;
    double FL3 = L->phig * L->rog * QGL * L->vg + L->phig * L->pg * eny;
    double FL4 = L->phig * L->rog * QGL * L->wg + L->phig * L->pg * enz;
    double FL5 = L->phig * L->rog * QGL * HGL;
    double FL6 = L->phip * L->rop * QPL;
    double FL7 = L->phip * L->rop * QPL * L->up + L->phip * L->pp * enx;
    double FL8 = L->phip * L->rop * QPL * L->vp + L->phip * L->pp * eny;
    double FL9 = L->phip * L->rop * QPL * L->wp + L->phip * L->pp * enz;
    double FL10 = L->phip * L->rop * QPL * HPL;
    double FR1 = R->phig * R->rog * QGR;
    double FR2 = R->phig * R->rog * QGR * R->ug + R->phig * R->pg * enx;
    double FR3 = R->phig * R->rog * QGR * R->vg + R->phig * R->pg * eny;
    double FR4 = R->phig * R->rog * QGR * R->wg + R->phig * R->pg * enz;
    double FR5 = R->phig * R->rog * QGR * HGR;
    double FR6 = R->phip * R->rop * QPR;
    double FR7 = R->phip * R->rop * QPR * R->up + R->phip * R->pp * enx;
    double FR8 = R->phip * R->rop * QPR * R->vp + R->phip * R->pp * eny;
    double FR9 = R->phip * R->rop * QPR * R->wp + R->phip * R->pp * enz;
    double FR10 = R->phip * R->rop * QPR * HPR;
    if (SGL > 0.0) {
        F[0] = FL1 * area; F[1] = FL2 * area; F[2] = FL3 * area; F[3] = FL4 * area; F[4] = FL5 * area;
    } else if (SGL <= 0.0 && SGSTAR > 0.0) {
        F[0] = (FL1 + SGL * (CSL1 - CNL1)) * area;
        F[1] = (FL2 + SGL * (CSL2 - CNL2)) * area;
        F[2] = (FL3 + SGL * (CSL3 - CNL3)) * area;
        F[3] = (FL4 + SGL * (CSL4 - CNL4)) * area;
        F[4] = (FL5 + SGL * (CSL5 - CNL5)) * area;
    } else if (SGSTAR <= 0.0 && SGR >= 0.0) {
        F[0] = (FR1 + SGR * (CSR1 - CNR1)) * area;
        F[1] = (FR2 + SGR * (CSR2 - CNR2)) * area;
        F[2] = (FR3 + SGR * (CSR3 - CNR3)) * area;
        F[3] = (FR4 + SGR * (CSR4 - CNR4)) * area;
        F[4] = (FR5 + SGR * (CSR5 - CNR5)) * area;
    } else if (SGR < 0.0) {
        F[0] = FR1 * area; F[1] = FR2 * area; F[2] = FR3 * area; F[3] = FR4 * area; F[4] = FR5 * area;
    }
    if (SPL > 0.0) {
        F[5] = FL6 * area; F[6] = FL7 * area; F[7] = FL8 * area; F[8] = FL9 * area; F[9] = FL10 * area;
    } else if (SPL <= 0.0 && SPSTAR > 0.0) {
        F[5] = (FL6 + SPL * (CSL6 - CNL6)) * area;
        F[6] = (FL7 + SPL * (CSL7 - CNL7)) * area;
        F[7] = (FL8 + SPL * (CSL8 - CNL8)) * area;
        F[8] = (FL9 + SPL * (CSL9 - CNL9)) * area;
        F[9] = (FL10 + SPL * (CSL10 - CNL10)) * area;
    } else if (SPSTAR <= 0.0 && SPR >= 0.0) {
        F[5] = (FR6 + SPR * (CSR6 - CNR6)) * area;
        F[6] = (FR7 + SPR * (CSR7 - CNR7)) * area;
        F[7] = (FR8 + SPR * (CSR8 - CNR8)) * area;
        F[8] = (FR9 + SPR * (CSR9 - CNR9)) * area;
        F[9] = (FR10 + SPR * (CSR10 - CNR10)) * area;
    } else if (SPR < 0.0) {
        F[5] = FR6 * area; F[6] = FR7 * area; F[7] = FR8 * area; F[8] = FR9 * area; F[9] = FR10 * area;
    }
}

/* HLLC-Zien flux with mixture wave speeds, sr_hllczien.F90:327-487 */
static void riemann_zien(const hyo_params *P, const face_state *L, const face_state *R,
                         double enx, double eny, double enz, double area, double F[10])
{
    const double rrg = P->rrg, gamag = P->gamag, gamap = P->gamap, cpp = P->cpp, pinf = P->pinf;
    double EGL = rrg * L->tg / (gamag - 1.0) + 0.5 * (L->ug * L->ug + L->vg * L->vg + L->wg * L->wg);
    double HGL = EGL + L->pg / L->rog;
    double QGL = L->ug * enx + L->vg * eny + L->wg * enz;
    double AGL = sqrt(gamag * L->pg / L->rog);
    double EGR = rrg * R->tg / (gamag - 1.0) + 0.5 * (R->ug * R->ug + R->vg * R->vg + R->wg * R->wg);
    double HGR = EGR + R->pg / R->rog;
    double QGR = R->ug * enx + R->vg * eny + R->wg * enz;
    double AGR = sqrt(gamag * R->pg / R->rog);
    double EPL = (cpp * L->tp) / gamap + pinf / L->rop + 0.5 * (L->up * L->up + L->vp * L->vp + L->wp * L->wp);
    double HPL = EPL + L->pp / L->rop;
    double QPL = L->up * enx + L->vp * eny + L->wp * enz;
    double APL = sqrt(gamap * (L->pp + pinf) / L->rop);
    double EPR = (cpp * R->tp) / gamap + pinf / R->rop + 0.5 * (R->up * R->up + R->vp * R->vp + R->wp * R->wp);
    double HPR = EPR + R->pp / R->rop;
    double QPR = R->up * enx + R->vp * eny + R->wp * enz;
    double APR = sqrt(gamap * (R->pp + pinf) / R->rop);
    double SL = dmin2(dmin2(dmin2(dmin2(QGL - AGL, QPL - APL), QGR - AGR), QPR - APR), F1E8);
    double SR = dmax2(dmax2(dmax2(dmax2(QGL + AGL, QPL + APL), QGR + AGR), QPR + APR), F1E8);
    double ROL = L->phig * L->rog + L->phip * L->rop;
    double ROQL = L->phig * L->rog * QGL + L->phip * L->rop * QPL;
    double QL = ROQL / ROL;
    double PRL = L->phig * L->pg + L->phip * L->pp;
    double ROR = R->phig * R->rog + R->phip * R->rop;
    double ROQR = R->phig * R->rog * QGR + R->phip * R->rop * QPR;
    double QR = ROQR / ROR;
    double PRR = R->phig * R->pg + R->phip * R->pp;
    double SSTARNR = PRR - PRL + ROQL * (SL - QL) - ROQR * (SR - QR);
    double SSTARDR = ROL * (SL - QL) - ROR * (SR - QR);
    double SSTAR = SSTARNR / (SSTARDR + F1E8);
    double PGSTAR = dmax2(L->pg + L->rog * (QGL - SL) * (QGL - SSTAR), F1E6 * dmin2(L->pg, R->pg));
    double PPSTAR = dmax2(L->pp + L->rop * (QPL - SL) * (QPL - SSTAR), F1E6 * dmin2(L->pp, R->pp));
    double CNL1 = L->phig * L->rog, CNL2 = L->phig * L->rog * L->ug, CNL3 = L->phig * L->rog * L->vg;
    double CNL4 = L->phig * L->rog * L->wg, CNL5 = L->phig * L->rog * EGL;
    double CNL6 = L->phip * L->rop, CNL7 = L->phip * L->rop * L->up, CNL8 = L->phip * L->rop * L->vp;
    double CNL9 = L->phip * L->rop * L->wp, CNL10 = L->phip * L->rop * EPL;
    double CNR1 = R->phig * R->rog, CNR2 = R->phig * R->rog * R->ug, CNR3 = R->phig * R->rog * R->vg;
    double CNR4 = R->phig * R->rog * R->wg, CNR5 = R->phig * R->rog * EGR;
    double CNR6 = R->phip * R->rop, CNR7 = R->phip * R->rop * R->up, CNR8 = R->phip * R->rop * R->vp;
    double CNR9 = R->phip * R->rop * R->wp, CNR10 = R->phip * R->rop * EPR;
    double DNR = (SSTAR - SL) + F1E6;
    double CSL1 = (QGL - SL) * CNL1 / DNR;
    double CSL2 = ((QGL - SL) * CNL2 + L->phig * (L->pg * enx - PGSTAR * enx)) / DNR;
    double CSL3 = ((QGL - SL) * CNL3 + L->phig * (L->pg * eny - PGSTAR * eny)) / DNR;
    double CSL4 = ((QGL - SL) * CNL4 + L->phig * (L->pg * enz - PGSTAR * enz)) / DNR;
    double CSL5 = ((QGL - SL) * CNL5 + L->phig * (L->pg * QGL - PGSTAR * SSTAR)) / DNR;
    double CSL6 = (QPL - SL) * CNL6 / DNR;
    double CSL7 = ((QPL - SL) * CNL7 + L->phip * (L->pp * enx - PPSTAR * enx)) / DNR;
    double CSL8 = ((QPL - SL) * CNL8 + L->phip * (L->pp * eny - PPSTAR * eny)) / DNR;
    double CSL9 = ((QPL - SL) * CNL9 + L->phip * (L->pp * enz - PPSTAR * enz)) / DNR;
    double CSL10 = ((QPL - SL) * CNL10 + L->phip * (L->pp * QPL - PPSTAR * SSTAR)) / DNR;
    DNR = (SSTAR - SR) + F1E6;
    double CSR1 = (QGR - SR) * CNR1 / DNR;
    double CSR2 = ((QGR - SR) * CNR2 + R->phig * (R->pg * enx - PGSTAR * enx)) / DNR;
    double CSR3 = ((QGR - SR) * CNR3 + R->phig * (R->pg * eny - PGSTAR * eny)) / DNR;
    double CSR4 = ((QGR - SR) * CNR4 + R->phig * (R->pg * enz - PGSTAR * enz)) / DNR;
    double CSR5 = ((QGR - SR) * CNR5 + R->phig * (R->pg * QGR - PGSTAR * SSTAR)) / DNR;
    double CSR6 = (QPR - SR) * CNR6 / DNR;
    double CSR7 = ((QPR - SR) * CNR7 + R->phip * (R->pp * enx - PPSTAR * enx)) / DNR;
    double CSR8 = ((QPR - SR) * CNR8 + R->phip * (R->pp * eny - PPSTAR * eny)) / DNR;
    double CSR9 = ((QPR - SR) * CNR9 + R->phip * (R->pp * enz - PPSTAR * enz)) / DNR;
    double CSR10 = ((QPR - SR) * CNR10 + R->phip * (R->pp * QPR - PPSTAR * SSTAR)) / DNR;
    double FL[10], FR[10], CL[10], CR[10], SLs[10], SRs[10];
    FL[0] = L->phig * L->rog * QGL;
    FL[1] = L->phig * L->rog * QGL * L->ug + L->phig * L->pg * enx;
    FL[2] = L->phig * L->rog * QGL * L->vg + L->phig * L->pg * eny;
    FL[3] = L->phig * L->rog * QGL * L->wg + L->phig * L->pg * enz;
    FL[4] = L->phig * L->rog * QGL * HGL;
    FL[5] = L->phip * L->rop * QPL;
    FL[6] = L->phip * L->rop * QPL * L->up + L->phip * L->pp * enx;
    FL[7] = L->phip * L->rop * QPL * L->vp + L->phip * L->pp * eny;
    FL[8] = L->phip * L->rop * QPL * L->wp + L->phip * L->pp * enz;
    FL[9] = L->phip * L->rop * QPL * HPL;
    FR[0] = R->phig * R->rog * QGR;
    FR[1] = R->phig * R->rog * QGR * R->ug + R->phig * R->pg * enx;
    FR[2] = R->phig * R->rog * QGR * R->vg + R->phig * R->pg * eny;
    FR[3] = R->phig * R->rog * QGR * R->wg + R->phig * R->pg * enz;
    FR[4] = R->phig * R->rog * QGR * HGR;
    FR[5] = R->phip * R->rop * QPR;
    FR[6] = R->phip * R->rop * QPR * R->up + R->phip * R->pp * enx;
    FR[7] = R->phip * R->rop * QPR * R->vp + R->phip * R->pp * eny;
    FR[8] = R->phip * R->rop * QPR * R->wp + R->phip * R->pp * enz;
    FR[9] = R->phip * R->rop * QPR * HPR;
    CL[0] = CNL1; CL[1] = CNL2; CL[2] = CNL3; CL[3] = CNL4; CL[4] = CNL5;
    CL[5] = CNL6; CL[6] = CNL7; CL[7] = CNL8; CL[8] = CNL9; CL[9] = CNL10;
    CR[0] = CNR1; CR[1] = CNR2; CR[2] = CNR3; CR[3] = CNR4; CR[4] = CNR5;
    CR[5] = CNR6; CR[6] = CNR7; CR[7] = CNR8; CR[8] = CNR9; CR[9] = CNR10;
    SLs[0] = CSL1; SLs[1] = CSL2; SLs[2] = CSL3; SLs[3] = CSL4; SLs[4] = CSL5;
    SLs[5] = CSL6; SLs[6] = CSL7; SLs[7] = CSL8; SLs[8] = CSL9; SLs[9] = CSL10;
    SRs[0] = CSR1; SRs[1] = CSR2; SRs[2] = CSR3; SRs[3] = CSR4; SRs[4] = CSR5;
    SRs[5] = CSR6; SRs[6] = CSR7; SRs[7] = CSR8; SRs[8] = CSR9; SRs[9] = CSR10;
    (void)QL; (void)QR;
    if (SL >= 0.0) {
        for (int k = 0; k < 10; k++) F[k] = FL[k] * area;
    } else if (SL <= 0.0 && SSTAR > 0.0) {
        for (int k = 0; k < 10; k++) F[k] = (FL[k] + SL * (SLs[k] - CL[k])) * area;
    } else if (SSTAR <= 0.0 && SR >= 0.0) {
        for (int k = 0; k < 10; k++) F[k] = (FR[k] + SR * (SRs[k] - CR[k])) * area;
    } else if (SR <= 0.0) {
        for (int k = 0; k < 10; k++) F[k] = FR[k] * area;
    }
}

/* face loop of sr_hllcbatten.F90:59-556 (myflux=2) or sr_hllczien.F90:62-502
 * (myflux=1); the two differ in the Riemann solver and in whether partition
 * ghosts are reconstructed (none exist on one rank). */
/* One face's left/right states after the second-order reconstruction and
 * before the phase clipping (sr_hllcbatten.F90:135-224 with the
 * reconstruction condition of :169,193; sr_hllczien.F90:138-217, which
 * reconstructs across interior faces only).  Returns the neighbour. */
static int face_states(const hyo_case *c, int i, int f, double *xf_out, face_state *L, face_state *R)
{
    const hyo_params *P = &c->p;
    const int ne = c->neles, ntot_bc = c->neles + c->nghosts;
    const int zien = (P->myflux == 1);
    int a, b, d;
    flux_face_nodes(c, i, f, &a, &b, &d);
    int neigh = NC(i, f);
    double xf = (c->x[a] + c->x[b] + c->x[d]) * 0.3333;
    double yf = (c->y[a] + c->y[b] + c->y[d]) * 0.3333;
    double zf = (c->z[a] + c->z[b] + c->z[d]) * 0.3333;
    xf_out[0] = xf; xf_out[1] = yf; xf_out[2] = zf;
    load_state(c, i, L);
    load_state(c, neigh, R);
    int recon = zien ? (neigh < ne) : (neigh < ne || neigh >= ntot_bc);
    if (P->myorder == 2 && recon) {
        double phimin = dmin2(c->phi[i], c->phi[neigh]);
        double xf0 = xf - c->xc[i], yf0 = yf - c->yc[i], zf0 = zf - c->zc[i];
        double xfn = xf - c->xc[neigh], yfn = yf - c->yc[neigh], zfn = zf - c->zc[neigh];
        L->phig = L->phig + phimin * GRD(HG_PHIG, i, xf0, yf0, zf0);
        L->ug = L->ug + phimin * GRD(HG_UG, i, xf0, yf0, zf0);
        L->vg = L->vg + phimin * GRD(HG_VG, i, xf0, yf0, zf0);
        L->wg = L->wg + phimin * GRD(HG_WG, i, xf0, yf0, zf0);
        L->tg = L->tg + phimin * GRD(HG_TG, i, xf0, yf0, zf0);
        L->pg = L->pg + phimin * GRD(HG_PG, i, xf0, yf0, zf0);
        R->phig = R->phig + phimin * GRD(HG_PHIG, neigh, xfn, yfn, zfn);
        R->ug = R->ug + phimin * GRD(HG_UG, neigh, xfn, yfn, zfn);
        R->vg = R->vg + phimin * GRD(HG_VG, neigh, xfn, yfn, zfn);
        R->wg = R->wg + phimin * GRD(HG_WG, neigh, xfn, yfn, zfn);
        R->tg = R->tg + phimin * GRD(HG_TG, neigh, xfn, yfn, zfn);
        R->pg = R->pg + phimin * GRD(HG_PG, neigh, xfn, yfn, zfn);
        L->up = L->up + phimin * GRD(HG_UP, i, xf0, yf0, zf0);
        L->vp = L->vp + phimin * GRD(HG_VP, i, xf0, yf0, zf0);
        L->wp = L->wp + phimin * GRD(HG_WP, i, xf0, yf0, zf0);
        L->tp = L->tp + phimin * GRD(HG_TP, i, xf0, yf0, zf0);
        L->pp = L->pp + phimin * GRD(HG_PP, i, xf0, yf0, zf0);
        R->up = R->up + phimin * GRD(HG_UP, neigh, xfn, yfn, zfn);
        R->vp = R->vp + phimin * GRD(HG_VP, neigh, xfn, yfn, zfn);
        R->wp = R->wp + phimin * GRD(HG_WP, neigh, xfn, yfn, zfn);
        R->tp = R->tp + phimin * GRD(HG_TP, neigh, xfn, yfn, zfn);
        R->pp = R->pp + phimin * GRD(HG_PP, neigh, xfn, yfn, zfn);
    }
    L->phip = 1.0 - L->phig;
    R->phip = 1.0 - R->phig;
    return neigh;
}

void hyo_face_flux(hyo_case *c)
{
    const hyo_params *P = &c->p;
    const int ne = c->neles, ntot_bc = c->neles + c->nghosts;
    const int zien = (P->myflux == 1);
    for (int i = 0; i < ne; i++) {
        if (c->iblank[i] != 1) continue;
        for (int k = 0; k < NEQ; k++) RHS(i, k) = 0.0;
        double F[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int f = 0; f < 4; f++) {
            double xf[3];
            face_state L, R;
            int neigh = face_states(c, i, f, xf, &L, &R);
            double scx = SC(i, f, 0), scy = SC(i, f, 1), scz = SC(i, f, 2);
            double area = sqrt(scx * scx + scy * scy + scz * scz);
            double enx = scx / area, eny = scy / area, enz = scz / area;
            int nt = 0;
            if (neigh >= ne && neigh < ntot_bc) nt = c->gtype[neigh - ne];
            clip_blend(P, &L);
            clip_blend(P, &R);
            if (nt == 5) {
                RHS(i, 1) = RHS(i, 1) + L.phig * L.pg * enx * area;
                RHS(i, 2) = RHS(i, 2) + L.phig * L.pg * eny * area;
                RHS(i, 3) = RHS(i, 3) + L.phig * L.pg * enz * area;
                RHS(i, 6) = RHS(i, 6) + L.phip * L.pp * enx * area;
                RHS(i, 7) = RHS(i, 7) + L.phip * L.pp * eny * area;
                RHS(i, 8) = RHS(i, 8) + L.phip * L.pp * enz * area;
            } else {
                if (zien)
                    riemann_zien(P, &L, &R, enx, eny, enz, area, F);
                else
                    riemann_batten(P, &L, &R, enx, eny, enz, area, F);
                for (int k = 0; k < NEQ; k++) RHS(i, k) = RHS(i, k) + F[k];
            }
        }
    }
}

/* The face states hyo_face_flux works from, for tests: for every interior
 * cell i and face f, out[((i*4+f)*2+side)*HF_N + field] with side 0 = left
 * (cell i), 1 = right (its neighbour), fields in face_state order
 * (rog ug vg wg tg phig rop up vp wp tp pg pp phip), then the face centroid
 * (0.3333 x node sum).  Reconstructed, not yet clipped or blended. */
_Static_assert(sizeof(face_state) == 14 * sizeof(double), "face_state is 14 doubles");

void hyo_face_states(hyo_case *c, double *out)
{
    for (int i = 0; i < c->neles; i++)
        for (int f = 0; f < 4; f++) {
            double xf[3];
            face_state LR[2];
            face_states(c, i, f, xf, &LR[0], &LR[1]);
            for (int side = 0; side < 2; side++) {
                double *o = out + ((size_t)(i * 4 + f) * 2 + side) * HF_N;
                memcpy(o, &LR[side], sizeof(face_state));
                o[14] = xf[0]; o[15] = xf[1]; o[16] = xf[2];
            }
        }
}

/* mod_Fluxmodule.F90:7-21 -> hllcbatten/hllczien: primitiveplusbc,
 * derivative + limiter (myorder=2), face loop, source */
void hyo_flux_solver(hyo_case *c)
{
    hyo_primitiveplusbc(c);
    if (c->p.myorder == 2) {
        hyo_derivative(c);
        hyo_venkat(c);
    }
    hyo_face_flux(c);
    hyo_source(c);
}

/* stage updates of sr_RK.F90:29-54, 68-90, 102-125, 138-161 */
void hyo_rk_stage(hyo_case *c, int s)
{
    for (int i = 0; i < c->neles; i++) {
        if (c->iblank[i] != 1) continue;
        double dtbyvol = c->dt / c->vol[i];
        for (int j = 0; j < NEQ; j++) {
            double r = RHS(i, j) - D24(i, j);
            if (s == 1) {
                CN(i, j) = CU(i, j) - (11.0 / 20.0) * dtbyvol * r;
                CN1(i, j) = CN(i, j);
            } else if (s == 2) {
                CN(i, j) = (3.0 / 8.0) * CU(i, j) + (5.0 / 8.0) * (CN(i, j) - (11.0 / 20.0) * dtbyvol * r);
            } else if (s == 3) {
                CN(i, j) = (4.0 / 9.0) * CU(i, j) + (5.0 / 9.0) * (CN(i, j) - (11.0 / 20.0) * dtbyvol * r);
            } else {
                CN(i, j) = (111.0 / 1331.0) * CU(i, j) + (260.0 / 1331.0) * CN1(i, j) +
                           (960.0 / 1331.0) * (CN(i, j) - (11.0 / 20.0) * dtbyvol * r);
            }
        }
    }
    hyo_donor_gather(c);  /* overset receptors (none unless hyo_set_donors) */
}

/* Overset donor gather.  The reference has none (it only freezes iblank
 * cells); this restates the rule of the fixed donor lists the HIP library
 * applies after each stage update: receptor r takes, per conserved
 * variable, w1 cn(d1) + w2 cn(d2) + w3 cn(d3) + w4 cn(d4) in that order.
 * Parity against the reference: none possible (no reference function). */
void hyo_set_donors(hyo_case *c, int n, const int *recv, const int *ids, const double *w)
{
    free(c->drecv); free(c->dids); free(c->dw);
    c->drecv = NULL; c->dids = NULL; c->dw = NULL;
    c->ndon = n;
    if (n <= 0) { c->ndon = 0; return; }
    c->drecv = malloc(sizeof(int) * n);
    c->dids = malloc(sizeof(int) * 4 * (size_t)n);
    c->dw = malloc(sizeof(double) * 4 * (size_t)n);
    memcpy(c->drecv, recv, sizeof(int) * n);
    memcpy(c->dids, ids, sizeof(int) * 4 * (size_t)n);
    memcpy(c->dw, w, sizeof(double) * 4 * (size_t)n);
}

void hyo_donor_gather(hyo_case *c)
{
    for (int r = 0; r < c->ndon; r++) {
        const int *d = c->dids + 4 * (size_t)r;
        const double *w = c->dw + 4 * (size_t)r;
        for (int j = 0; j < NEQ; j++)
            CN(c->drecv[r], j) = w[0] * CN(d[0], j) + w[1] * CN(d[1], j) + w[2] * CN(d[2], j) + w[3] * CN(d[3], j);
    }
}

/* sr_RK.F90:5-168 (halo exchanges are no-ops on one rank) */
void hyo_rk(hyo_case *c)
{
    hyo_viscousgp(c);
    for (int s = 1; s <= 4; s++) {
        hyo_flux_solver(c);
        hyo_rk_stage(c, s);
    }
}

/* main.F90:515-642 */
void hyo_step(hyo_case *c, double *diffs)
{
    hyo_primitive(c);
    c->total = c->total + 1;
    double dtmin = hyo_dtmin(c);
    c->dt = dtmin * c->p.cfl;
    c->time = c->time + c->dt;
    hyo_smooth(c);
    hyo_rk(c);
    hyo_primitive(c);
    double dp = 0, du = 0, dv = 0, dw = 0;
    for (int i = 0; i < c->neles; i++) {
        dp = dmax2(dp, fabs(CN(i, 0) - CU(i, 0)));
        du = dmax2(du, fabs(CN(i, 1) - CU(i, 1)));
        dv = dmax2(dv, fabs(CN(i, 2) - CU(i, 2)));
        dw = dmax2(dw, fabs(CN(i, 3) - CU(i, 3)));
    }
    if (diffs) { diffs[0] = du; diffs[1] = dv; diffs[2] = dw; diffs[3] = dp; }
    memcpy(c->cu, c->cn, sizeof(double) * NEQ * (size_t)c->ntot);
}

void hyo_run(hyo_case *c, int nsteps)
{
    for (int s = 0; s < nsteps; s++) hyo_step(c, NULL);
}

double *hyo_array(hyo_case *c, int which)
{
    switch (which) {
    case HA_CN: return c->cn;
    case HA_CU: return c->cu;
    case HA_CN1: return c->cn1;
    case HA_RHS: return c->rhs;
    case HA_D24: return c->d24;
    case HA_D2Q: return c->d2q;
    case HA_ANU: return c->anu;
    case HA_Q: return c->q;
    case HA_GRAD: return c->grad;
    case HA_PHI: return c->phi;
    case HA_VIS: return c->vis;
    case HA_SC: return c->sc;
    case HA_XC: return c->xc;
    case HA_YC: return c->yc;
    case HA_ZC: return c->zc;
    case HA_VOL: return c->vol;
    case HA_DL: return c->dl;
    case HA_DD: return c->dd;
    case HA_EPSSQ: return c->epssq;
    default: return NULL;
    }
}
