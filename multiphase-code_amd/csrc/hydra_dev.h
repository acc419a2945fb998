// hydra_dev.h -- device-side views shared by the kernels and the host
// library of libhydra_hip.so (not part of the public C ABI).
#ifndef HYDRA_DEV_H
#define HYDRA_DEV_H

#include <hip/hip_runtime.h>

#include "hydra_hip.h"

namespace hydra {

typedef hp_params Params;

// primitive slots of q (each ntot long: interior cells, then ghosts)
enum { Q_UG, Q_VG, Q_WG, Q_PG, Q_PP, Q_TG, Q_PHIG, Q_UP, Q_VP, Q_WP, Q_PHIP, Q_TP, Q_ROP, Q_ROG, Q_N };
// stored gradient slots (x, y, z each)
enum { G_UG, G_VG, G_WG, G_TG, G_UP, G_VP, G_WP, G_TP, G_PG, G_PP, G_PHIG, G_N };
// device scalars
enum { SC_DT, SC_TIME, SC_DTMIN, SC_N };

// mesh and geometry, all device pointers, 0-based ids
struct Mesh {
    int ne, ng, nodes;
    const double *x, *y, *z;
    const int *nod;     // 4 x ne
    const int *nc;      // 4 x ne, ghost g -> ne+g
    const int *ncp;     // 4 x ne, ghost -> its parent
    const int *iblank;  // ne
    const int *gpar, *gtype, *gside, *gdd_win;  // ng
    double *sc;         // 12 x ne area vectors, slot 3*face+dim
    double *xc, *yc, *zc, *vol;  // ntot
    double *dl, *epssq, *dd;     // ne
    double *fc0;        // 12 x ne face centroids, flux/limiter variant
    double *fc3;        // 12 x ne face centroids, viscous variant
};

// per-case state, all device pointers
struct State {
    int ne, nt;
    double *cn, *cu, *cn1;  // 10 x ne
    double *q;              // Q_N x nt
    double *grad;           // 3*G_N x ne
    double *phi;            // ne
    float *g32, *phi32;     // FP32-bulk variant: gradients and limiter (else null)
    double *vis;            // 8 x ne
    double *d24, *d2q;      // 10 x ne
    double *anu;            // 2 x ne
    double *rhs;            // 10 x ne
    const double *dl;       // ne (same as Mesh::dl)
    double *scal;           // SC_N
    int *itot;              // total (step counter used by the outflow BC)
};

void launch_geometry(const Mesh &m, double xmax, double ymax, double zmax, double ymin, double zmin,
                     hipStream_t st);
void launch_eos(const Params &P, const State &s, int passes, double *dtpart, hipStream_t st);
void launch_dt_final(const Params &P, const State &s, const double *dtpart, int nparts, hipStream_t st);
void launch_bc(const Params &P, const State &s, const Mesh &m, int viscous, hipStream_t st);
void launch_smooth(const Params &P, const State &s, const Mesh &m, hipStream_t st);
void launch_viscous(const Params &P, const State &s, const Mesh &m, hipStream_t st);
void launch_grad(const Params &P, const State &s, const Mesh &m, hipStream_t st);
void launch_flux(const Params &P, const State &s, const Mesh &m, int stage, hipStream_t st);
void launch_rk_update(const State &s, const Mesh &m, int stage, hipStream_t st);
void launch_end_step(const State &s, double *part, hipStream_t st);
void launch_dt_from_prims(const Params &P, const State &s, double *dtpart, hipStream_t st);

}  // namespace hydra

#endif
