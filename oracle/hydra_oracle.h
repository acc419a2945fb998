/*
 * hydra_oracle.h -- CPU restatement of the HydraPhase per-step finite-volume
 * update (two-phase, six-equation model on tetrahedra), FP64, serial.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (libhydra_hip.so) never links or calls it.
 *
 * Indices are 0-based.  Cell arrays of length ntot hold the interior cells
 * 0..neles-1 followed by the boundary ghosts neles..neles+nghosts-1 (ghost g
 * of the boundary list is cell neles+g, as in the reference's bc files).
 * Multi-variable arrays are variable-major ("column-major" like the Fortran
 * cn(ntot,neq)): element (i,k) lives at [k*ld + i].
 */
#ifndef HYDRA_ORACLE_H
#define HYDRA_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* Flow parameters in the order of flow.in (sr_Read_flow.F90:19-268) plus the
 * constants main.F90:336-349 derives.  Identical field order to hp_params in
 * include/hydra_hip.h so one host-side record fills both. */
typedef struct {
    double pmax, pmin, tgmax, tgmin;
    double ugmax, ugmin, vgmax, vgmin, wgmax, wgmin;
    double upmax, upmin, vpmax, vpmin, wpmax, wpmin;
    double tpmax, tpmin;
    double pinf, gamap, gamag, cpp, cpg, amolg, prg;
    double epslnmax, epslnmin, epslninf, phipzero;
    double cfl, diap, surfacetension, as, ts;
    int niter, itersub, istart, myflux, myorder, pad_;
    /* derived (main.F90:336-349); filled by hyo_derive_params */
    double co2, co4, c23, cmug, rrg, gravity, prp, foursigmabyd;
} hyo_params;

/* primitive slots in the q array (each of length ntot) */
enum {
    HQ_UG, HQ_VG, HQ_WG, HQ_PG, HQ_PP, HQ_TG, HQ_PHIG,
    HQ_UP, HQ_VP, HQ_WP, HQ_PHIP, HQ_TP, HQ_ROP, HQ_ROG,
    HQ_EGT, HQ_EPT, HQ_AG, HQ_AP, HQ_N
};

/* gradient slots (each x,y,z => 3 entries), sr_derivative.F90:259-309 */
enum {
    HG_UG, HG_VG, HG_WG, HG_TG, HG_UP, HG_VP, HG_WP, HG_TP,
    HG_PG, HG_PP, HG_ROG, HG_ROP, HG_PHIG, HG_N
};

/* viscous flux slots, sr_viscousgp.F90:797-804 */
enum { HV_UG, HV_VG, HV_WG, HV_TG, HV_UP, HV_VP, HV_WP, HV_TP, HV_N };

/* internal arrays exposed to tests through hyo_array() */
enum {
    HA_CN, HA_CU, HA_CN1, HA_RHS, HA_D24, HA_D2Q, HA_ANU, HA_Q, HA_GRAD,
    HA_PHI, HA_VIS, HA_SC, HA_XC, HA_YC, HA_ZC, HA_VOL, HA_DL, HA_DD,
    HA_EPSSQ, HA_N
};

typedef struct hyo_case hyo_case;

void hyo_derive_params(hyo_params *p);

/* nod, nc: 4 x neles, variable-major, 0-based; nc entries >= neles name
 * ghosts.  gparent/gtype/gside: per boundary ghost (gside 1..4). */
hyo_case *hyo_create(int nodes, int neles, int nghosts,
                     const double *x, const double *y, const double *z,
                     const int *nod, const int *nc, const int *iblank,
                     const int *gparent, const int *gtype, const int *gside,
                     const hyo_params *p);
void hyo_destroy(hyo_case *c);

/* restart state: cu (10 x neles) and ghost gas primitives
 * gp (6 x nghosts: ug vg wg pg tg rog), main.F90:427-441. */
void hyo_set_state(hyo_case *c, const double *cu, const double *gp,
                   double time, int total);
void hyo_get_state(const hyo_case *c, double *cu, double *gp,
                   double *time, int *total);

/* one reference routine each, for per-kernel checks */
void hyo_primitive(hyo_case *c);        /* sr_primitive.F90 */
void hyo_primitiveplusbc(hyo_case *c);  /* sr_primitiveplusbc.F90 */
double hyo_dtmin(hyo_case *c);          /* main.F90:520-548 */
void hyo_set_dt(hyo_case *c, double dt, double time, int total);
void hyo_smooth(hyo_case *c);           /* sr_smooth.F90 */
void hyo_viscousgp(hyo_case *c);        /* sr_viscousgp.F90 */
void hyo_viscous_bc(hyo_case *c);       /* sr_viscousgp.F90:71-446 */
void hyo_bc(hyo_case *c);               /* sr_primitiveplusbc.F90:198-632 */
void hyo_derivative(hyo_case *c);       /* sr_derivative.F90 */
void hyo_venkat(hyo_case *c);           /* sr_VK.F90 */
void hyo_face_flux(hyo_case *c);        /* face loop of hllcbatten/hllczien */
/* reconstructed face states of every (cell, face, side) before clipping:
 * out holds neles*4*2*HF_N doubles (14 fields, then the face centroid) */
enum { HF_N = 17 };
void hyo_face_states(hyo_case *c, double *out);
void hyo_source(hyo_case *c);           /* sr_source.F90 */
void hyo_flux_solver(hyo_case *c);      /* mod_Fluxmodule.F90:7-21 */
void hyo_rk_stage(hyo_case *c, int s);  /* sr_RK.F90 stage s (1..4) update */
void hyo_rk(hyo_case *c);               /* sr_RK.F90 */
/* overset donor lists (0-based; ids row-major n x 4, weights n x 4) and
 * their gather; hyo_rk_stage applies the lists after its update.  No
 * reference counterpart. */
void hyo_set_donors(hyo_case *c, int n, const int *recv, const int *ids, const double *w);
void hyo_donor_gather(hyo_case *c);
/* one time step, main.F90:515-642; diffs (4) = max|cn-cu| vars 2,3,4,1 */
void hyo_step(hyo_case *c, double *diffs);
void hyo_run(hyo_case *c, int nsteps);

double *hyo_array(hyo_case *c, int which);
double hyo_get_dt(const hyo_case *c);

#ifdef __cplusplus
}
#endif
#endif
