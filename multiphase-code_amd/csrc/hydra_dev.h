// hydra_dev.h -- device-side views shared by the kernels and the host
// library of libhydra_hip.so (not part of the public C ABI).
#ifndef HYDRA_DEV_H
#define HYDRA_DEV_H

#include <hip/hip_runtime.h>

#include <string>

#include "hydra_hip.h"

namespace hydra {

typedef hp_params Params;

// primitive slots of q (each ntot long: interior cells, then ghosts)
enum { Q_UG, Q_VG, Q_WG, Q_PG, Q_PP, Q_TG, Q_PHIG, Q_UP, Q_VP, Q_WP, Q_PHIP, Q_TP, Q_ROP, Q_ROG, Q_N };
// interphase source terms x cell volume, per stage (sr_source.F90:46-155):
// gas x/y/z momentum and energy (the liquid takes y/z momentum and energy
// with the opposite sign) and liquid x momentum, whose gravity term differs
enum { SRC_MX, SRC_MY, SRC_MZ, SRC_E, SRC_MXP, SRC_N };
// stored gradient slots (x, y, z each)
enum { G_UG, G_VG, G_WG, G_TG, G_UP, G_VP, G_WP, G_TP, G_PG, G_PP, G_PHIG, G_N };
// device scalars
enum { SC_DT, SC_TIME, SC_DTMIN, SC_N };
// The FP32-bulk record of a cell (or ghost), 48 floats = 192 bytes read as
// twelve 16-byte loads: the face state's 13 primitives (rog ug vg wg tg phig
// rop up vp wp tp pg pp, rounded from FP64), the 11 stored gradients x
// (x, y, z) in G_* order, the limiter, one pad.  A ghost's record holds its
// primitives only; a cell that is never updated keeps zero gradients and
// limiter, as the reference keeps its derivative arrays.
enum { R32_PRIM = 0, R32_GRAD = 13, R32_PHI = 46, R32_N = 48 };
// floats per face entry of State::fflux: the 10 flux components, 2 pad
enum { FACE32_N = 12 };
// floats per listed face of State::fgeom
enum { FGEO32_N = 9 };
// the FP32-bulk HLLC-Batten form's interphase source, per cell, as the five
// FP32 terms the stage update widens and multiplies by the volume (variable
// 1, variables 2 and 7, 3 and 8, 4 and 9, variable 6)
enum { SRC32_N = 5 };
// planes of State::cgeo32
enum { CGEO32_N = 24 };

// CFL bound of one cell from its primitives (main.F90:525-546): dl over the
// larger phase sound speed plus the larger phase speed.  q is variable-major
// with leading dimension nt.
__device__ __forceinline__ double cfl_bound(const double *q, size_t nt, const double *dl, int i,
                                            double gamag, double gamap, double pinf)
{
    const double ag = sqrt((gamag * q[(size_t)Q_PG * nt + i]) / q[(size_t)Q_ROG * nt + i]);
    const double ap = sqrt((gamap * (q[(size_t)Q_PP * nt + i] + pinf)) / q[(size_t)Q_ROP * nt + i]);
    const double ug = q[(size_t)Q_UG * nt + i], vg = q[(size_t)Q_VG * nt + i], wg = q[(size_t)Q_WG * nt + i];
    const double up = q[(size_t)Q_UP * nt + i], vp = q[(size_t)Q_VP * nt + i], wp = q[(size_t)Q_WP * nt + i];
    const double qgi = sqrt(ug * ug + vg * vg + wg * wg);
    const double qpi = sqrt(up * up + vp * vp + wp * wp);
    const double a = ag > ap ? ag : ap;
    const double v = qgi > qpi ? qgi : qpi;
    return dl[i] / (a + v);
}

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
    double *q2;             // Q_N x nt: the fused step's other primitive buffer
    double *grad;           // 3*G_N x ne
    double *phi;            // ne
    double *src;            // SRC_N x ne interphase source terms x volume (FP64 form; else null)
    float *rec32;           // FP32-bulk variant with second-order faces: per cell and ghost (ntot rows,
                            // cell-major) R32_N floats, the record the face loop reads (else null)
    // FP32-bulk HLLC-Batten (else null): every face's flux computed once for
    // the two cells that share it.  The faces a cell computes (a boundary
    // face, or one towards a higher-numbered cell) are listed cell by cell:
    // flist[k] = 4*cell + face, cell i's from fbase[i] to fbase[i+1] (ne+1
    // entries); fflux holds FACE32_N floats per listed face (10 flux
    // components, 2 pad); fslot (4 x ne, face-major) gives each cell
    // face its entry as (k << 1) | 1 when the cell takes the listed values
    // with the opposite sign; fgeom holds each listed face's FP32 geometry,
    // FGEO32_N floats (the area vector, the face centroid's offsets from both
    // cells' centroids), computed at set-up
    const int *fslot, *fbase, *flist;
    float *fflux, *fgeom;
    float *src32;           // FP32-bulk HLLC-Batten: SRC32_N x ne source terms, written by the gradient
                            // pass (else null)
    float *cgeo32;          // FP32-bulk second order: CGEO32_N x ne, each cell's neighbour centroid offsets
                            // (x of faces 0..3, then y, z) and face-centroid offsets (x, y, z), rounded
                            // from FP64 once at set-up (else null)
    double *vis;            // 8 x ne (FP64 form; else null)
    double *d24, *d2q;      // 10 x ne (d24: FP64 form; else null)
    float *visf, *d24f;     // FP32-bulk form: the viscous fluxes (8 x ne) and the JST dissipation
                            // (10 x ne), fluxes the form keeps in FP32 (else null)
    double *anu;            // 2 x ne
    double *rhs;            // 10 x ne
    const double *dl;       // ne (same as Mesh::dl)
    double *scal;           // SC_N
    int *itot;              // total (step counter used by the outflow BC)
    const unsigned char *dflag;  // ne: bit 0 overset donor, bit 1 overset receptor (null without donor lists)
    double *cnd;            // 10 x ne: the donors' stage update for the fused step's gather (with donor lists)
    double *mon;            // 4: the last fused step's max|cn - cu| of variables 1..4 (main.F90:603-612)
};

// Work of one launch.  Cells (and boundary ghosts) of every case are cut
// into tiles of TILE consecutive ids; a tile belongs to one case slot, so a
// block reads that case's parameters, state and mesh through block-uniform
// (scalar) loads.  Tile counts live in device memory: a launch (or a
// captured graph) with a fixed grid serves whatever set of cases the tables
// describe when it runs.
constexpr int TILE = 64;
struct Tile {
    int cs;     // case slot
    int first;  // first cell (or ghost) of the case in this tile, 0-based
    int count;  // 1..TILE
    int pad;
};
struct Batch {
    const Params *P;        // per slot
    const State *S;         // per slot
    const Mesh *M;          // per slot
    const Tile *ct, *gt;    // cell tiles, ghost tiles
    const int *cnt;         // [0] cell tiles, [1] ghost tiles
    const int *slot_t0;     // per slot: its cell tiles are [slot_t0, slot_t1)
    const int *slot_t1;
    double *part;           // 4 per cell tile: block partials (dt minimum, diffs)
    // a context on its own (Grid::single): its case, carried in the launch
    // arguments, and its cell and ghost counts
    Params P1;
    State S1;
    Mesh M1;
    int n1, n1g;
    // fused step: this launch reads the q2 buffers as q (and writes q); the
    // two primitive buffers trade places from one stage to the next
    int swap;
};
// grid of each launch family (blocks; the tile loops take multiples of 8)
// and the kernel variants shared by every case of the batch
struct Grid {
    int cell;   // per-cell kernels, TILE threads per block
    int quad;   // four lanes per cell, 4*TILE threads per block
    int ghost;  // ghost kernels, TILE threads per block
    int slots;  // k_dt_final, one block per slot
    int persist;  // bits 1 cell, 2 quad, 4 ghost, 8 cellw: the grid may hold fewer blocks than tiles
    int single;   // one case, taken from Batch::P1/S1/M1
    int precision, myflux, myorder;
    int tpb;      // tiles per block (1, 2 or 4) of the streaming per-cell kernels (TILE_LOOP_W)
    int cellw;    // their blocks
};

void launch_geometry(const Mesh &m, double xmax, double ymax, double zmax, double ymin, double zmin,
                     hipStream_t st);
void launch_face32_geometry(const Mesh &m, const State &s, int nf, hipStream_t st);
void launch_grad32_geometry(const Mesh &m, const State &s, hipStream_t st);
void launch_eos(const Batch &B, const Grid &G, int want_dt, hipStream_t st);
void launch_dt_final(const Batch &B, const Grid &G, hipStream_t st);
void launch_bc(const Batch &B, const Grid &G, int viscous, hipStream_t st);
void launch_smooth(const Batch &B, const Grid &G, hipStream_t st);
void launch_viscous(const Batch &B, const Grid &G, hipStream_t st);
void launch_grad(const Batch &B, const Grid &G, hipStream_t st);
// feos: 0 none; 1 fused step, the updated cells' primitive recovery into q2;
// 2 as 1 plus cu = cn (last stage)
void launch_flux(const Batch &B, const Grid &G, int stage, int feos, hipStream_t st);
void launch_rk_update(const Batch &B, const Grid &G, int stage, hipStream_t st);
void launch_end_step(const Batch &B, const Grid &G, int want_diffs, hipStream_t st);
void launch_health(const Batch &B, const Grid &G, int *flags, hipStream_t st);
void launch_check_division(const double *a, const double *b, size_t n, unsigned long long *bad, hipStream_t st);
void launch_dt_from_prims(const Batch &B, const Grid &G, hipStream_t st);

// overset donor lists of one case: per receptor four donor ids and weights
struct Donors {
    int n;
    const int *recv;      // n
    const int4 *ids;      // n, 0-based
    const double2 *w;     // 2n: (w1, w2), (w3, w4)
    double *part;         // 4 per block of 256 receptors: the receptors' monitor maxima at the step's end
};
constexpr int DONOR_BLOCK = 256;
// the monitor maxima of each slot's fused step (main.F90:603-612) from the
// last flux launch's tile partials (and, for a lone case, the last donor
// gather's): into out[0..3] when out is not null, else into the slot's
// State::mon
void launch_monitor_final(const Batch &B, const Grid &G, const Donors *D, double *out, hipStream_t st);
// eos: 0 the gather from cn alone; fused step, 1 the gather from cnd, then
// the receptors' primitive recovery into s.q, 2 as 1 plus cu = cn (last stage)
void launch_donor_gather(const Params &P, const State &s, const Donors &D, int eos, hipStream_t st);

// the no-allocation entry points' scratch array on the current device:
// resolves its address once per device (hydra_noalloc.hip); 0 or an hp_status
int noalloc_setup(std::string *err);
// the error hp_last_error(NULL) reports; returns code
int set_global_error(int code, const std::string &msg);

}  // namespace hydra

#endif
