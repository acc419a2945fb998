// hydra_capi.cpp -- host side of libhydra_hip.so: the extern "C" entry points
// declared in include/hydra_hip.h.  It owns the device buffers and the HIP
// stream of each context and sequences the kernels of hydra_kernels.hip the
// way the reference's step loop sequences its subroutines
// (main.F90:511-657, sr_RK.F90, mod_Fluxmodule.F90).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "hydra_dev.h"
#include "hydra_faces.h"
#include "hydra_hip.h"

using namespace hydra;

// Device tables that let one launch cover a set of cases (hydra_dev.h
// Batch): per-slot copies of the Params/State/Mesh views, the cell and
// ghost tile tables, and the per-tile partials.  A context owns one with a
// single slot; an hp_batch owns one with many.
struct TileSet {
    Params *dP = nullptr;
    State *dS = nullptr;
    Mesh *dM = nullptr;
    Tile *ct = nullptr, *gt = nullptr;
    int *cnt = nullptr, *t0 = nullptr, *t1 = nullptr;
    double *part = nullptr;
    int slots = 0, ct_cap = 0, gt_cap = 0;
    int nct = 0, ngt = 0;
    std::vector<Params> hP;
    std::vector<State> hS;
    std::vector<Mesh> hM;
    std::vector<char> used;
    // host staging of the last upload (kept alive until the next one)
    std::vector<Tile> hct, hgt;
    std::vector<int> h0, h1, hcnt;

    Batch view() const
    {
        Batch B{};
        B.P = dP; B.S = dS; B.M = dM; B.ct = ct; B.gt = gt; B.cnt = cnt;
        B.slot_t0 = t0; B.slot_t1 = t1; B.part = part;
        if (slots == 1 && used[0]) {  // a context's own set: its case rides in the launch
            B.P1 = hP[0];
            B.S1 = hS[0];
            B.M1 = hM[0];
            B.n1 = hS[0].ne;
            B.n1g = hM[0].ng;
        }
        return B;
    }
};

struct hp_batch {
    hipStream_t st = nullptr;
    TileSet ts;
    Grid G{};
    std::vector<void *> bufs;
    std::vector<hp_ctx *> members;  // per slot, null when free
    int ncases = 0;
    int grid_cell = 0, grid_quad = 0, grid_ghost = 0;  // 0: one block per tile at capacity
    int tiles_per_block = 1;                            // streaming per-cell kernels (hp_batch_set_block)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int *health = nullptr;          // per slot, device
    std::string err;
};

struct hp_ctx {
    hipStream_t st = nullptr;
    Params P{};
    Mesh m{};
    State s{};
    std::vector<void *> bufs;
    TileSet ts;
    Grid G{};
    hp_batch *batch = nullptr;  // the batch this case belongs to, if any
    Donors D{};                 // overset donor lists (n = 0: none)
    std::vector<void *> dbufs;  // their device buffers
    int *health = nullptr;      // device flag word of hp_health
    std::string err;
    // timing of the last hp_step
    std::vector<hipEvent_t> fev;  // start/stop pairs around flux launches (hp_set_flux_timing)
    hipEvent_t t0 = nullptr, t1 = nullptr;
    int nflux_timed = 0;
    bool flux_timing = false;
    // per-step monitor maxima of hp_step_diffs (4 per step, device)
    double *mon_log = nullptr;
    int mon_cap = 0;
    // one captured time step for hp_step (rebuilt when the donor lists change)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

thread_local std::string g_create_err;

int fail(hp_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    else g_create_err = msg;
    return code;
}

}  // namespace

int hydra::set_global_error(int code, const std::string &msg) { return fail(nullptr, code, msg); }

// forget the captured time step of hp_step (its launch arguments changed)
static int drop_step_graph(hp_ctx *c)
{
    hipError_t e = hipSuccess;
    if (c->exec) e = hipGraphExecDestroy(c->exec);
    if (c->graph && e == hipSuccess) e = hipGraphDestroy(c->graph);
    c->exec = nullptr;
    c->graph = nullptr;
    if (e != hipSuccess) return fail(c, HP_ERR_HIP, std::string("graph release: ") + hipGetErrorString(e));
    return HP_OK;
}

namespace {

#define HIPCHK(c, call)                                                                           \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail((c), HP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

template <typename T>
int dalloc(hp_ctx *c, T **p, size_t n)
{
    void *d = nullptr;
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    hipError_t e = hipMalloc(&d, bytes);
    if (e != hipSuccess) return fail(c, HP_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    e = hipMemsetAsync(d, 0, bytes, c->st);
    if (e != hipSuccess) return fail(c, HP_ERR_HIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    c->bufs.push_back(d);
    *p = static_cast<T *>(d);
    return HP_OK;
}

// main.F90:336-349 (REAL*4 literals as the Fortran evaluates them)
void derive_params(Params &P)
{
    P.co2 = (double)(1.0f / 16.0f);
    P.co4 = (double)(1.0f / 1024.0f);
    P.c23 = (double)(2.0f / 3.0f);
    P.cmug = (double)11.848e-8f * std::sqrt(P.amolg);
    P.rrg = 287.0;
    P.gravity = 0.0;
    P.prp = 10.0;
    P.foursigmabyd = 4.0 * P.surfacetension / P.diap;
}

int tiles_of(int n) { return (n + TILE - 1) / TILE; }
int round8(int n) { return ((n < 1 ? 1 : n) + 7) & ~7; }

// device tables for `slots` slots and the given tile capacities; alloc(ptr,
// count) is the owner's allocator (its buffers are freed with the owner)
template <class Alloc>
int ts_alloc(TileSet &t, int slots, int ct_cap, int gt_cap, Alloc &&alloc)
{
    t.slots = slots;
    t.ct_cap = ct_cap;
    t.gt_cap = gt_cap;
    t.hP.assign(slots, Params{});
    t.hS.assign(slots, State{});
    t.hM.assign(slots, Mesh{});
    t.used.assign(slots, 0);
    int rc;
    if ((rc = alloc(&t.dP, slots)) || (rc = alloc(&t.dS, slots)) || (rc = alloc(&t.dM, slots)) ||
        (rc = alloc(&t.ct, ct_cap)) || (rc = alloc(&t.gt, gt_cap)) || (rc = alloc(&t.cnt, 2)) ||
        (rc = alloc(&t.t0, slots)) || (rc = alloc(&t.t1, slots)) || (rc = alloc(&t.part, 4 * (size_t)ct_cap)))
        return rc;
    return HP_OK;
}

// rebuild the tile tables from the slots in use and copy them to the device
int ts_upload(TileSet &t, hipStream_t st, std::string &err)
{
    t.hct.clear();
    t.hgt.clear();
    t.h0.assign(t.slots, 0);
    t.h1.assign(t.slots, 0);
    for (int c = 0; c < t.slots; c++) {
        if (!t.used[c]) continue;
        const int ne = t.hS[c].ne, ng = t.hM[c].ng;
        t.h0[c] = (int)t.hct.size();
        for (int f = 0; f < ne; f += TILE) t.hct.push_back(Tile{c, f, std::min(TILE, ne - f), 0});
        t.h1[c] = (int)t.hct.size();
        for (int f = 0; f < ng; f += TILE) t.hgt.push_back(Tile{c, f, std::min(TILE, ng - f), 0});
    }
    if ((int)t.hct.size() > t.ct_cap || (int)t.hgt.size() > t.gt_cap) {
        err = "tile tables over capacity";
        return HP_ERR_ARG;
    }
    t.nct = (int)t.hct.size();
    t.ngt = (int)t.hgt.size();
    t.hcnt = {t.nct, t.ngt};
    auto cp = [&](void *d, const void *h, size_t bytes) -> hipError_t {
        return bytes ? hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st) : hipSuccess;
    };
    hipError_t e;
    if ((e = cp(t.dP, t.hP.data(), sizeof(Params) * t.slots)) != hipSuccess ||
        (e = cp(t.dS, t.hS.data(), sizeof(State) * t.slots)) != hipSuccess ||
        (e = cp(t.dM, t.hM.data(), sizeof(Mesh) * t.slots)) != hipSuccess ||
        (e = cp(t.ct, t.hct.data(), sizeof(Tile) * t.hct.size())) != hipSuccess ||
        (e = cp(t.gt, t.hgt.data(), sizeof(Tile) * t.hgt.size())) != hipSuccess ||
        (e = cp(t.cnt, t.hcnt.data(), sizeof(int) * 2)) != hipSuccess ||
        (e = cp(t.t0, t.h0.data(), sizeof(int) * t.slots)) != hipSuccess ||
        (e = cp(t.t1, t.h1.data(), sizeof(int) * t.slots)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess) {
        err = std::string("tile tables: ") + hipGetErrorString(e);
        return HP_ERR_HIP;
    }
    return HP_OK;
}

// one time step (main.F90:515-625) of every case in B; fev (8 events, may
// be null) brackets the four flux launches; the step's monitor maxima go to
// mon_out (4 doubles) when it is not null, else to each slot's State::mon
void enqueue_step(const Batch &B, const Grid &G, hipStream_t st, hipEvent_t *fev, const Donors *D = nullptr,
                  double *mon_out = nullptr)
{
    launch_eos(B, G, 1, st);       // primitive + CFL bound per tile
    launch_dt_final(B, G, st);     // dt, time, total per case
    launch_smooth(B, G, st);       // :563
    launch_bc(B, G, 1, st);        // sr_RK.F90:19, viscousgp's BCs
    launch_viscous(B, G, st);
    launch_eos(B, G, 0, st);       // stage 1's primitiveplusbc
    // Each stage's flux launch ends with the primitive recovery the next
    // stage starts with (the step's last one, main.F90:595, with cu = cn,
    // :625, and the monitor differences, :603-612), written to the other
    // primitive buffer:
    // stages 1 and 3 read q, stages 2 and 4 read q2 (Bs), and the step ends
    // in q, where it started.
    Batch Bs = B;
    Bs.swap = 1;
    for (int k = 1; k <= 4; k++) {
        const Batch &Bk = (k & 1) ? B : Bs;
        launch_bc(Bk, G, 0, st);
        launch_grad(Bk, G, st);
        if (fev) (void)hipEventRecord(fev[2 * (k - 1)], st);
        launch_flux(Bk, G, k, k == 4 ? 2 : 1, st);  // face flux + source + stage k update + primitives
        if (fev) (void)hipEventRecord(fev[2 * (k - 1) + 1], st);
        if (D) {
            // overset receptors of a lone case, in the buffer just written
            State s = B.S1;
            if (k & 1) std::swap(s.q, s.q2);
            launch_donor_gather(B.P1, s, *D, k == 4 ? 2 : 1, st);
        }
    }
    launch_monitor_final(B, G, D, mon_out, st);  // max|cn - cu| of variables 1..4
}

}  // namespace

extern "C" {

int hp_create(hp_ctx **out, int nodes, int neles, int nghosts, const double *x, const double *y,
              const double *z, const int *nod, int ld_nod, const int *nc1, const int *nc2, const int *nc3,
              const int *nc4, const int *iblank, const int *nparent, const int *nghost, const int *ntype,
              const int *nside, const hp_params *params)
{
    if (!out) return fail(nullptr, HP_ERR_ARG, "hp_create: null ctx pointer");
    *out = nullptr;
    if (nodes <= 0 || neles <= 0 || nghosts < 0 || ld_nod < neles)
        return fail(nullptr, HP_ERR_ARG, "hp_create: bad sizes");
    if (!x || !y || !z || !nod || !nc1 || !nc2 || !nc3 || !nc4 || !iblank || !params ||
        (nghosts > 0 && (!nparent || !nghost || !ntype || !nside)))
        return fail(nullptr, HP_ERR_ARG, "hp_create: null array");
    if (params->myflux != 1 && params->myflux != 2)
        return fail(nullptr, HP_ERR_FLUX, "Invalid flux method! (myflux must be 1 or 2)");
    if (params->myorder != 1 && params->myorder != 2)
        return fail(nullptr, HP_ERR_ARG, "hp_create: myorder must be 1 or 2");
    if (params->precision != 0 && params->precision != 1)
        return fail(nullptr, HP_ERR_ARG, "hp_create: precision must be 0 (FP64) or 1 (FP32 bulk)");

    const int ne = neles, ng = nghosts, ntot = ne + ng;
    // host-side tables, 0-based
    std::vector<int> hnod(4 * (size_t)ne), hnc(4 * (size_t)ne), hncp(4 * (size_t)ne), hib(ne);
    std::vector<int> gpar(ng), gtype(ng), gside(ng), gwin(ng, 0);
    const int *ncs[4] = {nc1, nc2, nc3, nc4};
    for (int k = 0; k < 4; k++)
        for (int i = 0; i < ne; i++) {
            const int a = nod[(size_t)k * ld_nod + i];
            if (a < 1 || a > nodes) return fail(nullptr, HP_ERR_MESH, "hp_create: node id out of range");
            hnod[(size_t)k * ne + i] = a - 1;
            const int b = ncs[k][i];
            if (b < 1 || b > ntot) return fail(nullptr, HP_ERR_MESH, "hp_create: neighbour id out of range");
            hnc[(size_t)k * ne + i] = b - 1;
        }
    for (int i = 0; i < ne; i++) hib[i] = iblank[i];
    for (int g = 0; g < ng; g++) {
        if (nghost[g] != ne + g + 1)
            return fail(nullptr, HP_ERR_MESH, "hp_create: ghost ids must be neles+1..neles+nghosts in list order");
        if (nparent[g] < 1 || nparent[g] > ne || nside[g] < 1 || nside[g] > 4)
            return fail(nullptr, HP_ERR_MESH, "hp_create: bad ghost parent/side");
        gpar[g] = nparent[g] - 1;
        gtype[g] = ntype[g];
        gside[g] = nside[g];
        if (hnc[(size_t)(gside[g] - 1) * ne + gpar[g]] != ne + g)
            return fail(nullptr, HP_ERR_MESH, "hp_create: ghost is not the neighbour of its parent's side");
    }
    // dd(np) keeps the last type 3/51/53 ghost of its parent in list order
    // (sr_geometry.F90:295-347)
    {
        std::vector<int> last(ne, -1);
        for (int g = 0; g < ng; g++)
            if (gtype[g] == 3 || gtype[g] == 51 || gtype[g] == 53) last[gpar[g]] = g;
        for (int i = 0; i < ne; i++)
            if (last[i] >= 0) gwin[last[i]] = 1;
    }
    for (size_t k = 0; k < 4 * (size_t)ne; k++) {
        const int b = hnc[k];
        hncp[k] = b < ne ? b : gpar[b - ne];
    }
    // FP32-bulk HLLC-Batten: the face tables of State::fslot/fbase/flist
    // (each face computed once, hydra_faces.h)
    std::vector<int> hfslot, hfbase, hflist;
    if (params->precision == 1 && params->myflux == 2) {
        if (ne >= (1 << 28)) return fail(nullptr, HP_ERR_ARG, "hp_create: FP32-bulk face tables take < 2^28 cells");
        build_face_tables(ne, hnc.data(), hfslot, hfbase, hflist);
    }
    double xmax = -1.0e30, ymax = -1.0e30, zmax = -1.0e30, ymin = 1.0e30, zmin = 1.0e30;
    for (int n = 0; n < nodes; n++) {
        xmax = std::max(xmax, x[n]);
        ymax = std::max(ymax, y[n]);
        zmax = std::max(zmax, z[n]);
        ymin = std::min(ymin, y[n]);
        zmin = std::min(zmin, z[n]);
    }

    {
        // set-up of the no-allocation entry points on this device
        std::string err;
        if (int rc0 = hydra::noalloc_setup(&err)) return fail(nullptr, rc0, err);
    }
    hp_ctx *c = new hp_ctx();
    c->P = *params;
    if (c->P.rrg == 0.0) derive_params(c->P);
    hipError_t e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    if (e != hipSuccess) {
        std::string msg = std::string("hipStreamCreate: ") + hipGetErrorString(e);
        delete c;
        return fail(nullptr, HP_ERR_HIP, msg);
    }
    int rc = HP_OK;
    Mesh &m = c->m;
    State &s = c->s;
    m.ne = ne; m.ng = ng; m.nodes = nodes;
    s.ne = ne; s.nt = ntot;
    double *dx, *dy, *dz;
    int *dnod, *dnc, *dncp, *dib, *dgpar, *dgtype, *dgside, *dgwin;
#define ALLOC(p, n)                                  \
    if ((rc = dalloc(c, &(p), (n))) != HP_OK) {      \
        g_create_err = c->err;                       \
        hp_destroy(c);                               \
        return rc;                                   \
    }
    ALLOC(dx, nodes) ALLOC(dy, nodes) ALLOC(dz, nodes)
    ALLOC(dnod, 4 * (size_t)ne) ALLOC(dnc, 4 * (size_t)ne) ALLOC(dncp, 4 * (size_t)ne) ALLOC(dib, ne)
    ALLOC(dgpar, ng) ALLOC(dgtype, ng) ALLOC(dgside, ng) ALLOC(dgwin, ng)
    ALLOC(m.sc, 12 * (size_t)ne) ALLOC(m.xc, ntot) ALLOC(m.yc, ntot) ALLOC(m.zc, ntot) ALLOC(m.vol, ntot)
    ALLOC(m.dl, ne) ALLOC(m.epssq, ne) ALLOC(m.dd, ne) ALLOC(m.fc0, 12 * (size_t)ne) ALLOC(m.fc3, 12 * (size_t)ne)
    ALLOC(s.cn, 10 * (size_t)ne) ALLOC(s.cu, 10 * (size_t)ne) ALLOC(s.cn1, 10 * (size_t)ne)
    ALLOC(s.q, (size_t)Q_N * ntot) ALLOC(s.q2, (size_t)Q_N * ntot) ALLOC(s.grad, (size_t)3 * G_N * ne) ALLOC(s.phi, ne)
    if (params->precision == 1) {
        // the FP32-bulk form keeps its viscous fluxes and JST dissipation in FP32
        ALLOC(s.visf, 8 * (size_t)ne) ALLOC(s.d24f, 10 * (size_t)ne)
    } else {
        ALLOC(s.vis, 8 * (size_t)ne) ALLOC(s.d24, 10 * (size_t)ne)
    }
    ALLOC(s.d2q, 10 * (size_t)ne)
    ALLOC(s.anu, 2 * (size_t)ne) ALLOC(s.rhs, 10 * (size_t)ne) ALLOC(s.scal, SC_N) ALLOC(s.itot, 1) ALLOC(s.mon, 4)
    ALLOC(c->health, 1)
    int *dfslot = nullptr, *dfbase = nullptr, *dflist = nullptr;
    if (params->precision == 1) {
        if (params->myorder == 2) {
            ALLOC(s.rec32, (size_t)R32_N * ntot)
            ALLOC(s.cgeo32, (size_t)CGEO32_N * ne)
        }
        if (!hflist.empty()) {
            ALLOC(dfslot, hfslot.size()) ALLOC(dfbase, hfbase.size()) ALLOC(dflist, hflist.size())
            ALLOC(s.fflux, (size_t)FACE32_N * hflist.size())
            ALLOC(s.fgeom, (size_t)FGEO32_N * hflist.size())
            ALLOC(s.src32, (size_t)SRC32_N * ne)
        }
    } else {
        ALLOC(s.src, (size_t)SRC_N * ne)
    }
#undef ALLOC
    m.x = dx; m.y = dy; m.z = dz; m.nod = dnod; m.nc = dnc; m.ncp = dncp; m.iblank = dib;
    m.gpar = dgpar; m.gtype = dgtype; m.gside = dgside; m.gdd_win = dgwin;
    s.dl = m.dl;
    s.fslot = dfslot;
    s.fbase = dfbase;
    s.flist = dflist;
    auto h2d = [&](void *d, const void *h, size_t bytes) -> int {
        hipError_t ee = hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->st);
        if (ee != hipSuccess) return fail(c, HP_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(ee));
        return HP_OK;
    };
    if ((rc = h2d(dx, x, sizeof(double) * nodes)) || (rc = h2d(dy, y, sizeof(double) * nodes)) ||
        (rc = h2d(dz, z, sizeof(double) * nodes)) || (rc = h2d(dnod, hnod.data(), sizeof(int) * hnod.size())) ||
        (rc = h2d(dnc, hnc.data(), sizeof(int) * hnc.size())) ||
        (rc = h2d(dncp, hncp.data(), sizeof(int) * hncp.size())) ||
        (rc = h2d(dib, hib.data(), sizeof(int) * hib.size())) ||
        (ng && ((rc = h2d(dgpar, gpar.data(), sizeof(int) * ng)) || (rc = h2d(dgtype, gtype.data(), sizeof(int) * ng)) ||
                (rc = h2d(dgside, gside.data(), sizeof(int) * ng)) || (rc = h2d(dgwin, gwin.data(), sizeof(int) * ng)))) ||
        (dflist && ((rc = h2d(dfslot, hfslot.data(), sizeof(int) * hfslot.size())) ||
                    (rc = h2d(dfbase, hfbase.data(), sizeof(int) * hfbase.size())) ||
                    (rc = h2d(dflist, hflist.data(), sizeof(int) * hflist.size()))))) {
        g_create_err = c->err;
        hp_destroy(c);
        return rc;
    }
    launch_geometry(m, xmax, ymax, zmax, ymin, zmin, c->st);
    launch_face32_geometry(m, s, (int)hflist.size(), c->st);
    launch_grad32_geometry(m, s, c->st);
    e = hipStreamSynchronize(c->st);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
        g_create_err = std::string("geometry: ") + hipGetErrorString(e);
        hp_destroy(c);
        return HP_ERR_HIP;
    }
    // the context is a batch of one: slot 0, one block per tile
    auto alloc = [c](auto **p, size_t n) { return dalloc(c, p, n); };
    if ((rc = ts_alloc(c->ts, 1, tiles_of(ne), std::max(1, tiles_of(ng)), alloc)) != HP_OK) {
        g_create_err = c->err;
        hp_destroy(c);
        return rc;
    }
    c->ts.hP[0] = c->P;
    c->ts.hS[0] = s;
    c->ts.hM[0] = m;
    c->ts.used[0] = 1;
    if ((rc = ts_upload(c->ts, c->st, c->err)) != HP_OK) {
        g_create_err = c->err;
        hp_destroy(c);
        return rc;
    }
    c->G.cell = round8(c->ts.nct);
    c->G.quad = round8(c->ts.nct);
    c->G.ghost = round8(c->ts.ngt);
    c->G.tpb = 1;
    c->G.cellw = c->G.cell;
    c->G.slots = 1;
    c->G.persist = 0;
    c->G.single = 1;
    c->G.precision = c->P.precision;
    c->G.myflux = c->P.myflux;
    c->G.myorder = c->P.myorder;
    *out = c;
    return HP_OK;
}

int hp_destroy(hp_ctx *c)
{
    if (!c) return HP_OK;
    if (c->batch) (void)hp_batch_remove(c->batch, c);
    if (c->st) (void)hipStreamSynchronize(c->st);
    for (void *p : c->dbufs) (void)hipFree(p);
    if (c->st) (void)hipStreamSynchronize(c->st);
    for (void *p : c->bufs) (void)hipFree(p);
    for (hipEvent_t e : c->fev) (void)hipEventDestroy(e);
    if (c->exec) (void)hipGraphExecDestroy(c->exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->mon_log) (void)hipFree(c->mon_log);
    if (c->t0) (void)hipEventDestroy(c->t0);
    if (c->t1) (void)hipEventDestroy(c->t1);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return HP_OK;
}

const char *hp_last_error(const hp_ctx *c) { return c ? c->err.c_str() : g_create_err.c_str(); }

int hp_synchronize(hp_ctx *c)
{
    if (!c) return fail(nullptr, HP_ERR_ARG, "null context");
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipGetLastError());
    return HP_OK;
}

int hp_upload_state(hp_ctx *c, const double *cu, int ld, const double *gp, double time, int total)
{
    if (!c || !cu || ld < c->m.ne || (c->m.ng > 0 && !gp)) return fail(c, HP_ERR_ARG, "hp_upload_state: bad argument");
    const int ne = c->m.ne, ng = c->m.ng;
    for (int k = 0; k < 10; k++) {
        HIPCHK(c, hipMemcpyAsync(c->s.cu + (size_t)k * ne, cu + (size_t)k * ld, sizeof(double) * ne,
                                 hipMemcpyHostToDevice, c->st));
    }
    HIPCHK(c, hipMemcpyAsync(c->s.cn, c->s.cu, sizeof(double) * 10 * ne, hipMemcpyDeviceToDevice, c->st));
    // the ghost state goes to both primitive buffers (the fused step's BCs
    // relax from the ghost's previous value in whichever buffer they read)
    const int slot[6] = {Q_UG, Q_VG, Q_WG, Q_PG, Q_TG, Q_ROG};
    for (int k = 0; k < 6 && ng > 0; k++)
        for (double *qb : {c->s.q, c->s.q2})
            HIPCHK(c, hipMemcpyAsync(qb + (size_t)slot[k] * c->s.nt + ne, gp + (size_t)k * ng, sizeof(double) * ng,
                                     hipMemcpyHostToDevice, c->st));
    // main.F90:456-476: viscous fluxes and dissipation start at zero
    if (c->s.visf) {
        HIPCHK(c, hipMemsetAsync(c->s.visf, 0, sizeof(float) * 8 * ne, c->st));
        HIPCHK(c, hipMemsetAsync(c->s.d24f, 0, sizeof(float) * 10 * ne, c->st));
    } else {
        HIPCHK(c, hipMemsetAsync(c->s.vis, 0, sizeof(double) * 8 * ne, c->st));
        HIPCHK(c, hipMemsetAsync(c->s.d24, 0, sizeof(double) * 10 * ne, c->st));
    }
    double sc[SC_N] = {0.0, time, 0.0};
    HIPCHK(c, hipMemcpyAsync(c->s.scal, sc, sizeof(sc), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->s.itot, &total, sizeof(int), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return HP_OK;
}

int hp_download_state(hp_ctx *c, double *cu, int ld, double *gp, double *time, int *total)
{
    if (!c || (cu && ld < c->m.ne)) return fail(c, HP_ERR_ARG, "hp_download_state: bad argument");
    const int ne = c->m.ne, ng = c->m.ng;
    if (cu)
        for (int k = 0; k < 10; k++)
            HIPCHK(c, hipMemcpyAsync(cu + (size_t)k * ld, c->s.cu + (size_t)k * ne, sizeof(double) * ne,
                                     hipMemcpyDeviceToHost, c->st));
    const int slot[6] = {Q_UG, Q_VG, Q_WG, Q_PG, Q_TG, Q_ROG};
    if (gp)
        for (int k = 0; k < 6 && ng > 0; k++)
            HIPCHK(c, hipMemcpyAsync(gp + (size_t)k * ng, c->s.q + (size_t)slot[k] * c->s.nt + ne, sizeof(double) * ng,
                                     hipMemcpyDeviceToHost, c->st));
    double sc[SC_N];
    int tot = 0;
    HIPCHK(c, hipMemcpyAsync(sc, c->s.scal, sizeof(sc), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(&tot, c->s.itot, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (time) *time = sc[SC_TIME];
    if (total) *total = tot;
    return HP_OK;
}

static int check_launch(hp_ctx *c, const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, HP_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return HP_OK;
}

// A case that belongs to a batch is also advanced on the batch's stream:
// its own routines first wait for the batch and finish before returning.
static int sync_owner(hp_ctx *c);
static int finish(hp_ctx *c, const char *what)
{
    if (int rc = check_launch(c, what)) return rc;
    if (c->batch) HIPCHK(c, hipStreamSynchronize(c->st));
    return HP_OK;
}

int hp_primitive(hp_ctx *c)
{
    if (!c) return fail(nullptr, HP_ERR_ARG, "null context");
    if (int rc = sync_owner(c)) return rc;
    launch_eos(c->ts.view(), c->G, 0, c->st);
    return finish(c, "primitive");
}

int hp_dt_min(hp_ctx *c, double *dtmin)
{
    if (!c || !dtmin) return fail(c, HP_ERR_ARG, "hp_dt_min: bad argument");
    if (int rc = sync_owner(c)) return rc;
    launch_dt_from_prims(c->ts.view(), c->G, c->st);
    if (int rc = check_launch(c, "dt")) return rc;
    std::vector<double> part(4 * (size_t)c->ts.nct);
    HIPCHK(c, hipMemcpyAsync(part.data(), c->ts.part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    // the tiles folded in cell order as the reference's MIN loop (a tile
    // holding a NaN bound restarts the fold: its part is independent of
    // what came before)
    double r = 1.0e8;
    for (int t = 0; t < c->ts.nct; t++) {
        const double m = part[(size_t)4 * t];
        r = part[(size_t)4 * t + 1] != 0.0 ? m : (r < m ? r : m);
    }
    *dtmin = r;
    return HP_OK;
}

int hp_set_dt(hp_ctx *c, double dt, double time, int total)
{
    if (!c) return fail(nullptr, HP_ERR_ARG, "null context");
    double sc[SC_N] = {dt, time, 0.0};
    HIPCHK(c, hipMemcpyAsync(c->s.scal, sc, sizeof(double) * 2, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->s.itot, &total, sizeof(int), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return HP_OK;
}

int hp_smooth(hp_ctx *c)
{
    if (!c) return fail(nullptr, HP_ERR_ARG, "null context");
    if (int rc = sync_owner(c)) return rc;
    launch_smooth(c->ts.view(), c->G, c->st);
    return finish(c, "smooth");
}

int hp_viscous(hp_ctx *c)
{
    if (!c) return fail(nullptr, HP_ERR_ARG, "null context");
    if (int rc = sync_owner(c)) return rc;
    launch_bc(c->ts.view(), c->G, 1, c->st);
    launch_viscous(c->ts.view(), c->G, c->st);
    return finish(c, "viscous");
}

static void flux_solver(hp_ctx *c, int stage)
{
    const Batch B = c->ts.view();
    launch_eos(B, c->G, 0, c->st);  // primitiveplusbc, interior
    launch_bc(B, c->G, 0, c->st);   // primitiveplusbc, ghosts
    launch_grad(B, c->G, c->st);    // myorder 2 only
    launch_flux(B, c->G, stage, 0, c->st);
}

int hp_flux(hp_ctx *c)
{
    if (!c) return fail(nullptr, HP_ERR_ARG, "null context");
    if (int rc = sync_owner(c)) return rc;
    flux_solver(c, 0);
    return finish(c, "flux");
}

int hp_rk_stage(hp_ctx *c, int stage)
{
    if (!c || stage < 1 || stage > 4) return fail(c, HP_ERR_ARG, "hp_rk_stage: stage must be 1..4");
    if (int rc = sync_owner(c)) return rc;
    launch_rk_update(c->ts.view(), c->G, stage, c->st);
    launch_donor_gather(c->P, c->s, c->D, 0, c->st);
    return finish(c, "rk_stage");
}

int hp_rk(hp_ctx *c)
{
    if (!c) return fail(nullptr, HP_ERR_ARG, "null context");
    if (int rc = sync_owner(c)) return rc;
    launch_bc(c->ts.view(), c->G, 1, c->st);
    launch_viscous(c->ts.view(), c->G, c->st);
    for (int st = 1; st <= 4; st++) {
        flux_solver(c, 0);
        launch_rk_update(c->ts.view(), c->G, st, c->st);
        launch_donor_gather(c->P, c->s, c->D, 0, c->st);
    }
    return finish(c, "rk");
}

int hp_end_step(hp_ctx *c, double *diffs)
{
    if (!c) return fail(nullptr, HP_ERR_ARG, "null context");
    if (int rc = sync_owner(c)) return rc;
    launch_eos(c->ts.view(), c->G, 0, c->st);
    launch_end_step(c->ts.view(), c->G, diffs ? 1 : 0, c->st);
    if (int rc = check_launch(c, "end_step")) return rc;
    if (diffs) {
        std::vector<double> part(4 * (size_t)c->ts.nct);
        HIPCHK(c, hipMemcpyAsync(part.data(), c->ts.part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        double d[4] = {0, 0, 0, 0};
        for (int b = 0; b < c->ts.nct; b++)
            for (int k = 0; k < 4; k++) d[k] = std::max(d[k], part[(size_t)4 * b + k]);
        // main.F90:653 prints diffu, diffv, diffw, diffp = variables 2,3,4,1
        diffs[0] = d[1]; diffs[1] = d[2]; diffs[2] = d[3]; diffs[3] = d[0];
    }
    return HP_OK;
}

int hp_step(hp_ctx *c, int nsteps) { return hp_step_diffs(c, nsteps, nullptr); }

int hp_step_diffs(hp_ctx *c, int nsteps, double *diffs)
{
    if (!c || nsteps < 0) return fail(c, HP_ERR_ARG, "hp_step: bad argument");
    if (c->flux_timing) {
        const int need = 8 * nsteps;
        while ((int)c->fev.size() < need) {
            hipEvent_t e;
            HIPCHK(c, hipEventCreate(&e));
            c->fev.push_back(e);
        }
    }
    if (diffs && nsteps > c->mon_cap) {
        HIPCHK(c, hipStreamSynchronize(c->st));
        if (c->mon_log) HIPCHK(c, hipFree(c->mon_log));
        c->mon_log = nullptr;
        c->mon_cap = 0;
        HIPCHK(c, hipMalloc(&c->mon_log, sizeof(double) * 4 * (size_t)nsteps));
        c->mon_cap = nsteps;
    }
    if (!c->t0) {
        HIPCHK(c, hipEventCreate(&c->t0));
        HIPCHK(c, hipEventCreate(&c->t1));
    }
    if (int rc = sync_owner(c)) return rc;
    const Donors *D = c->D.n ? &c->D : nullptr;
    // The plain step (no per-flux events, monitor into State::mon) replays
    // one captured graph: kernel launches only, every argument fixed for
    // the context's lifetime except the donor lists, whose change drops it.
    const bool graph = !c->flux_timing && !diffs && nsteps > 0;
    if (graph && !c->exec) {
        HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
        enqueue_step(c->ts.view(), c->G, c->st, nullptr, D, nullptr);
        hipGraph_t g = nullptr;
        HIPCHK(c, hipStreamEndCapture(c->st, &g));
        c->graph = g;
        HIPCHK(c, hipGraphInstantiate(&c->exec, c->graph, nullptr, nullptr, 0));
    }
    HIPCHK(c, hipEventRecord(c->t0, c->st));
    int ev = 0;
    for (int n = 0; n < nsteps; n++) {
        if (graph) {
            HIPCHK(c, hipGraphLaunch(c->exec, c->st));
            continue;
        }
        enqueue_step(c->ts.view(), c->G, c->st, c->flux_timing ? &c->fev[ev] : nullptr, D,
                     diffs ? c->mon_log + 4 * (size_t)n : nullptr);
        if (c->flux_timing) ev += 8;
    }
    HIPCHK(c, hipEventRecord(c->t1, c->st));
    c->nflux_timed = ev / 2;
    if (diffs && nsteps > 0) {
        std::vector<double> h(4 * (size_t)nsteps);
        HIPCHK(c, hipMemcpyAsync(h.data(), c->mon_log, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        // main.F90:653 prints diffu, diffv, diffw, diffp = variables 2,3,4,1
        for (int n = 0; n < nsteps; n++) {
            const double *d = &h[4 * (size_t)n];
            diffs[4 * (size_t)n] = d[1]; diffs[4 * (size_t)n + 1] = d[2];
            diffs[4 * (size_t)n + 2] = d[3]; diffs[4 * (size_t)n + 3] = d[0];
        }
    }
    return finish(c, "step");
}

int hp_last_diffs(hp_ctx *c, double *diffs)
{
    if (!c || !diffs) return fail(c, HP_ERR_ARG, "hp_last_diffs: bad argument");
    if (int rc = sync_owner(c)) return rc;
    double d[4];
    HIPCHK(c, hipMemcpyAsync(d, c->s.mon, sizeof(d), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    diffs[0] = d[1]; diffs[1] = d[2]; diffs[2] = d[3]; diffs[3] = d[0];
    return HP_OK;
}

int hp_set_flux_timing(hp_ctx *c, int on)
{
    if (!c) return fail(nullptr, HP_ERR_ARG, "null context");
    c->flux_timing = on != 0;
    return HP_OK;
}

int hp_last_timing(hp_ctx *c, double *total_ms, double *flux_ms, int *flux_launches)
{
    if (!c) return fail(nullptr, HP_ERR_ARG, "null context");
    HIPCHK(c, hipStreamSynchronize(c->st));
    float t = 0.0f;
    if (c->t0) HIPCHK(c, hipEventElapsedTime(&t, c->t0, c->t1));
    double fl = 0.0;
    for (int k = 0; k < c->nflux_timed; k++) {
        float d = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&d, c->fev[2 * k], c->fev[2 * k + 1]));
        fl += d;
    }
    if (total_ms) *total_ms = t;
    if (flux_ms) *flux_ms = fl;
    if (flux_launches) *flux_launches = c->nflux_timed;
    return HP_OK;
}

int hp_health(hp_ctx *c, int *status)
{
    if (!c || !status) return fail(c, HP_ERR_ARG, "hp_health: bad argument");
    if (int rc = sync_owner(c)) return rc;
    HIPCHK(c, hipMemsetAsync(c->health, 0, sizeof(int), c->st));
    launch_health(c->ts.view(), c->G, c->health, c->st);
    if (int rc = check_launch(c, "health")) return rc;
    HIPCHK(c, hipMemcpyAsync(status, c->health, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return HP_OK;
}

int hp_probe_cell(hp_ctx *c, int cell, double *prims)
{
    if (!c || !prims || cell < 1 || cell > c->s.nt) return fail(c, HP_ERR_ARG, "hp_probe_cell: bad argument");
    for (int k = 0; k < Q_N; k++)
        HIPCHK(c, hipMemcpyAsync(prims + k, c->s.q + (size_t)k * c->s.nt + (cell - 1), sizeof(double),
                                 hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return HP_OK;
}

// the FP32-bulk form's FP32 arrays behind array ids 4 (d24) and 10 (vis);
// null in the FP64 form or for other ids
static float *array_view_f32(hp_ctx *c, int which, size_t *n)
{
    const size_t ne = c->m.ne;
    if (which == 4 && c->s.d24f) { *n = 10 * ne; return c->s.d24f; }
    if (which == 10 && c->s.visf) { *n = 8 * ne; return c->s.visf; }
    return nullptr;
}

static int array_view(hp_ctx *c, int which, double **ptr, size_t *n)
{
    const size_t ne = c->m.ne, nt = c->s.nt;
    switch (which) {
    case 0: *ptr = c->s.cn; *n = 10 * ne; return HP_OK;
    case 1: *ptr = c->s.cu; *n = 10 * ne; return HP_OK;
    case 2: *ptr = c->s.cn1; *n = 10 * ne; return HP_OK;
    case 3: *ptr = c->s.rhs; *n = 10 * ne; return HP_OK;
    case 4: *ptr = c->s.d24; *n = 10 * ne; return HP_OK;
    case 5: *ptr = c->s.d2q; *n = 10 * ne; return HP_OK;
    case 6: *ptr = c->s.anu; *n = 2 * ne; return HP_OK;
    case 7: *ptr = c->s.q; *n = (size_t)Q_N * nt; return HP_OK;
    case 8: *ptr = c->s.grad; *n = (size_t)3 * G_N * ne; return HP_OK;
    case 9: *ptr = c->s.phi; *n = ne; return HP_OK;
    case 10: *ptr = c->s.vis; *n = 8 * ne; return HP_OK;
    case 12: *ptr = c->m.dl; *n = ne; return HP_OK;
    default: return fail(c, HP_ERR_ARG, "unknown array id");
    }
}

int hp_get_array(hp_ctx *c, int which, double *host, size_t count)
{
    if (!c || !host) return fail(c, HP_ERR_ARG, "hp_get_array: bad argument");
    if (which == 11) {
        if (count < 3) return fail(c, HP_ERR_ARG, "hp_get_array: need 3 doubles");
        double sc[SC_N];
        int tot = 0;
        HIPCHK(c, hipMemcpyAsync(sc, c->s.scal, sizeof(sc), hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipMemcpyAsync(&tot, c->s.itot, sizeof(int), hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        host[0] = sc[SC_DT]; host[1] = sc[SC_TIME]; host[2] = tot;
        return HP_OK;
    }
    if (which == 13 || which == 14) {
        // the FP32-bulk form's gradients and limiter, from the cells' records
        const size_t ne = c->m.ne, n = which == 13 ? (size_t)3 * G_N * ne : ne;
        if (!c->s.rec32) return fail(c, HP_ERR_ARG, "hp_get_array: arrays 13/14 exist in the FP32-bulk second-order form only");
        if (count != n) return fail(c, HP_ERR_ARG, "hp_get_array: count mismatch");
        std::vector<float> tmp((size_t)R32_N * ne);
        HIPCHK(c, hipMemcpyAsync(tmp.data(), c->s.rec32, sizeof(float) * tmp.size(), hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        for (size_t i = 0; i < ne; i++) {
            const float *r = &tmp[(size_t)R32_N * i];
            if (which == 14) host[i] = r[R32_PHI];
            else
                for (int k = 0; k < 3 * G_N; k++) host[(size_t)k * ne + i] = r[R32_GRAD + k];
        }
        return HP_OK;
    }
    double *p;
    size_t n;
    if (float *pf = array_view_f32(c, which, &n)) {
        // widened to FP64 on the host
        if (count != n) return fail(c, HP_ERR_ARG, "hp_get_array: count mismatch");
        std::vector<float> tmp(n);
        HIPCHK(c, hipMemcpyAsync(tmp.data(), pf, sizeof(float) * n, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        for (size_t k = 0; k < n; k++) host[k] = tmp[k];
        return HP_OK;
    }
    if (int rc = array_view(c, which, &p, &n)) return rc;
    if (count != n) return fail(c, HP_ERR_ARG, "hp_get_array: count mismatch");
    HIPCHK(c, hipMemcpyAsync(host, p, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return HP_OK;
}

int hp_set_array(hp_ctx *c, int which, const double *host, size_t count)
{
    if (!c || !host) return fail(c, HP_ERR_ARG, "hp_set_array: bad argument");
    if (which >= 12) return fail(c, HP_ERR_ARG, "hp_set_array: arrays 12-14 are read-only");
    double *p;
    size_t n;
    if (float *pf = array_view_f32(c, which, &n)) {
        // rounded to FP32 on the host
        if (count != n) return fail(c, HP_ERR_ARG, "hp_set_array: count mismatch");
        std::vector<float> tmp(host, host + n);
        HIPCHK(c, hipMemcpyAsync(pf, tmp.data(), sizeof(float) * n, hipMemcpyHostToDevice, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        return HP_OK;
    }
    if (int rc = array_view(c, which, &p, &n)) return rc;
    if (count != n) return fail(c, HP_ERR_ARG, "hp_set_array: count mismatch");
    HIPCHK(c, hipMemcpyAsync(p, host, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return HP_OK;
}

int hp_set_donors(hp_ctx *c, int n, const int *recv, const int *donors, int ld, const double *weights)
{
    if (!c || n < 0 || (n > 0 && (!recv || !donors || !weights || ld < n)))
        return fail(c, HP_ERR_ARG, "hp_set_donors: bad argument");
    if (c->batch) return fail(c, HP_ERR_ARG, "hp_set_donors: the case belongs to a batch");
    const int ne = c->m.ne;
    std::vector<int> hr(n);
    std::vector<int4> hid(n);
    std::vector<double2> hw(2 * (size_t)n);
    std::vector<char> is_recv(ne, 0);
    for (int r = 0; r < n; r++) {
        if (recv[r] < 1 || recv[r] > ne) return fail(c, HP_ERR_ARG, "hp_set_donors: receptor id out of range");
        if (is_recv[recv[r] - 1]) return fail(c, HP_ERR_ARG, "hp_set_donors: receptor listed twice");
        is_recv[recv[r] - 1] = 1;
        hr[r] = recv[r] - 1;
    }
    for (int r = 0; r < n; r++) {
        int d[4];
        for (int k = 0; k < 4; k++) {
            d[k] = donors[(size_t)k * ld + r] - 1;
            if (d[k] < 0 || d[k] >= ne) return fail(c, HP_ERR_ARG, "hp_set_donors: donor id out of range");
            if (is_recv[d[k]]) return fail(c, HP_ERR_ARG, "hp_set_donors: a donor is itself a receptor");
        }
        hid[r] = make_int4(d[0], d[1], d[2], d[3]);
        hw[2 * (size_t)r] = make_double2(weights[r], weights[(size_t)ld + r]);
        hw[2 * (size_t)r + 1] = make_double2(weights[(size_t)2 * ld + r], weights[(size_t)3 * ld + r]);
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (int rc = drop_step_graph(c)) return rc;
    for (void *p : c->dbufs) HIPCHK(c, hipFree(p));
    c->dbufs.clear();
    c->D = Donors{};
    c->s.dflag = nullptr;
    c->s.cnd = nullptr;
    c->ts.hS[0].dflag = nullptr;
    c->ts.hS[0].cnd = nullptr;
    if (n == 0) return ts_upload(c->ts, c->st, c->err);
    // a per-cell flag (bit 0 donor, bit 1 receptor), and where the fused
    // step's flux keeps the donors' stage update for the gather
    std::vector<unsigned char> hflag(ne, 0);
    for (int r = 0; r < n; r++) {
        hflag[hid[r].x] |= 1; hflag[hid[r].y] |= 1; hflag[hid[r].z] |= 1; hflag[hid[r].w] |= 1;
        hflag[hr[r]] |= 2;
    }
    void *pf = nullptr, *pd = nullptr;
    HIPCHK(c, hipMalloc(&pf, ne));
    c->dbufs.push_back(pf);
    HIPCHK(c, hipMalloc(&pd, sizeof(double) * 10 * (size_t)ne));
    c->dbufs.push_back(pd);
    HIPCHK(c, hipMemcpyAsync(pf, hflag.data(), ne, hipMemcpyHostToDevice, c->st));
    void *pr = nullptr, *pi = nullptr, *pw = nullptr;
    HIPCHK(c, hipMalloc(&pr, sizeof(int) * n));
    c->dbufs.push_back(pr);
    HIPCHK(c, hipMalloc(&pi, sizeof(int4) * n));
    c->dbufs.push_back(pi);
    HIPCHK(c, hipMalloc(&pw, sizeof(double2) * 2 * (size_t)n));
    c->dbufs.push_back(pw);
    void *pp = nullptr;  // the gather's monitor partials, 4 per block of receptors
    HIPCHK(c, hipMalloc(&pp, sizeof(double) * 4 * (size_t)((n + DONOR_BLOCK - 1) / DONOR_BLOCK)));
    c->dbufs.push_back(pp);
    HIPCHK(c, hipMemcpyAsync(pr, hr.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(pi, hid.data(), sizeof(int4) * n, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(pw, hw.data(), sizeof(double2) * 2 * (size_t)n, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->D.n = n;
    c->D.recv = static_cast<const int *>(pr);
    c->D.ids = static_cast<const int4 *>(pi);
    c->D.w = static_cast<const double2 *>(pw);
    c->D.part = static_cast<double *>(pp);
    c->s.dflag = static_cast<const unsigned char *>(pf);
    c->s.cnd = static_cast<double *>(pd);
    c->ts.hS[0].dflag = c->s.dflag;
    c->ts.hS[0].cnd = c->s.cnd;
    if (int rc = ts_upload(c->ts, c->st, c->err)) return rc;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return HP_OK;
}

int hp_apply_donors(hp_ctx *c)
{
    if (!c) return fail(nullptr, HP_ERR_ARG, "null context");
    launch_donor_gather(c->P, c->s, c->D, 0, c->st);
    return finish(c, "apply_donors");
}

static int sync_owner(hp_ctx *c)
{
    if (c->batch) HIPCHK(c, hipStreamSynchronize(c->batch->st));
    return HP_OK;
}

// ---------------------------------------------------------------------------
// batches: many cases advanced by the same launches
// ---------------------------------------------------------------------------
static int bfail(hp_batch *b, int code, const std::string &msg)
{
    if (b) b->err = msg;
    else g_create_err = msg;
    return code;
}

#define BHIPCHK(b, call)                                                                          \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return bfail((b), HP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

// blocks of G.tpb tiles for the streaming per-cell kernels: as many as the
// cell family's one-tile blocks hold tiles, rounded up to multiples of 8;
// they loop over tiles (persist bit 8) when they hold fewer than ntiles
static void cellw_grid(Grid &G, int ntiles)
{
    G.cellw = round8((G.cell + G.tpb - 1) / G.tpb);
    G.persist = (G.persist & ~8) | ((long long)G.cellw * G.tpb < ntiles ? 8 : 0);
}

static void batch_grids(hp_batch *b)
{
    b->G.cell = b->grid_cell ? round8(b->grid_cell) : round8(b->ts.ct_cap);
    b->G.quad = b->grid_quad ? round8(b->grid_quad) : round8(b->ts.ct_cap);
    b->G.ghost = b->grid_ghost ? round8(b->grid_ghost) : round8(b->ts.gt_cap);
    b->G.slots = b->ts.slots;
    b->G.single = 0;
    b->G.tpb = b->tiles_per_block;
    // a grid with fewer blocks than the tables can hold tiles loops
    b->G.persist = (b->G.cell < b->ts.ct_cap ? 1 : 0) | (b->G.quad < b->ts.ct_cap ? 2 : 0) |
                   (b->G.ghost < b->ts.gt_cap ? 4 : 0);
    cellw_grid(b->G, b->ts.ct_cap);
}

// forget the captured step; a launch of it still in flight finishes first
static void batch_drop_graph(hp_batch *b)
{
    if (b->exec && b->st) (void)hipStreamSynchronize(b->st);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int hp_batch_create(hp_batch **out, int max_cases, int max_cells, int max_ghosts)
{
    if (!out) return bfail(nullptr, HP_ERR_ARG, "hp_batch_create: null pointer");
    *out = nullptr;
    if (max_cases < 1 || max_cells < 1 || max_ghosts < 0)
        return bfail(nullptr, HP_ERR_ARG, "hp_batch_create: bad capacity");
    hp_batch *b = new hp_batch();
    hipError_t e = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->e0);
    if (e == hipSuccess) e = hipEventCreate(&b->e1);
    if (e != hipSuccess) {
        std::string msg = std::string("hp_batch_create: ") + hipGetErrorString(e);
        hp_batch_destroy(b);
        return bfail(nullptr, HP_ERR_HIP, msg);
    }
    auto alloc = [b](auto **p, size_t n) -> int {
        void *d = nullptr;
        const size_t bytes = std::max<size_t>(n, 1) * sizeof(**p);
        if (hipMalloc(&d, bytes) != hipSuccess) return bfail(b, HP_ERR_NOMEM, "hp_batch_create: hipMalloc");
        b->bufs.push_back(d);
        if (hipMemsetAsync(d, 0, bytes, b->st) != hipSuccess) return bfail(b, HP_ERR_HIP, "hp_batch_create: memset");
        *p = static_cast<std::remove_reference_t<decltype(*p)>>(d);
        return HP_OK;
    };
    // every case can add one partial tile of cells and one of ghosts
    const int ct_cap = tiles_of(max_cells) + max_cases, gt_cap = tiles_of(max_ghosts) + max_cases;
    if (int rc = ts_alloc(b->ts, max_cases, ct_cap, gt_cap, alloc)) {
        std::string msg = b->err;
        hp_batch_destroy(b);
        return bfail(nullptr, rc, msg);
    }
    if (int rc = ts_upload(b->ts, b->st, b->err)) {
        std::string msg = b->err;
        hp_batch_destroy(b);
        return bfail(nullptr, rc, msg);
    }
    if (int rc = alloc(&b->health, (size_t)max_cases)) {
        std::string msg = b->err;
        hp_batch_destroy(b);
        return bfail(nullptr, rc, msg);
    }
    b->members.assign(max_cases, nullptr);
    batch_grids(b);
    *out = b;
    return HP_OK;
}

int hp_batch_destroy(hp_batch *b)
{
    if (!b) return HP_OK;
    if (b->st) (void)hipStreamSynchronize(b->st);
    for (hp_ctx *c : b->members)
        if (c) c->batch = nullptr;
    batch_drop_graph(b);
    for (void *p : b->bufs) (void)hipFree(p);
    if (b->e0) (void)hipEventDestroy(b->e0);
    if (b->e1) (void)hipEventDestroy(b->e1);
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
    return HP_OK;
}

int hp_batch_add(hp_batch *b, hp_ctx *c, int *slot)
{
    if (!b || !c) return bfail(b, HP_ERR_ARG, "hp_batch_add: null argument");
    if (c->D.n) return bfail(b, HP_ERR_ARG, "hp_batch_add: a case with overset donor lists runs on its own");
    if (c->batch) return bfail(b, HP_ERR_ARG, "hp_batch_add: the case already belongs to a batch");
    if (b->ncases > 0 && (c->P.precision != b->G.precision || c->P.myflux != b->G.myflux ||
                          c->P.myorder != b->G.myorder))
        return bfail(b, HP_ERR_ARG, "hp_batch_add: precision, myflux and myorder must match the batch");
    int k = 0;
    while (k < b->ts.slots && b->members[k]) k++;
    if (k == b->ts.slots) return bfail(b, HP_ERR_NOMEM, "hp_batch_add: no free case slot");
    if (b->ts.nct + tiles_of(c->m.ne) > b->ts.ct_cap || b->ts.ngt + tiles_of(c->m.ng) > b->ts.gt_cap)
        return bfail(b, HP_ERR_NOMEM, "hp_batch_add: batch cell or ghost capacity exceeded");
    BHIPCHK(b, hipStreamSynchronize(c->st));
    BHIPCHK(b, hipStreamSynchronize(b->st));
    if (b->ncases == 0) {
        if (b->G.precision != c->P.precision || b->G.myflux != c->P.myflux || b->G.myorder != c->P.myorder)
            batch_drop_graph(b);
        b->G.precision = c->P.precision;
        b->G.myflux = c->P.myflux;
        b->G.myorder = c->P.myorder;
    }
    b->ts.hP[k] = c->P;
    b->ts.hS[k] = c->s;
    b->ts.hS[k].dflag = nullptr;  // batch steps run no donor gather
    b->ts.hS[k].cnd = nullptr;
    b->ts.hM[k] = c->m;
    b->ts.used[k] = 1;
    if (int rc = ts_upload(b->ts, b->st, b->err)) {
        b->ts.used[k] = 0;
        return rc;
    }
    b->members[k] = c;
    b->ncases++;
    c->batch = b;
    if (slot) *slot = k;
    return HP_OK;
}

int hp_batch_remove(hp_batch *b, hp_ctx *c)
{
    if (!b || !c) return bfail(b, HP_ERR_ARG, "hp_batch_remove: null argument");
    int k = 0;
    while (k < b->ts.slots && b->members[k] != c) k++;
    if (k == b->ts.slots) return bfail(b, HP_ERR_ARG, "hp_batch_remove: the case is not in this batch");
    BHIPCHK(b, hipStreamSynchronize(b->st));
    b->ts.used[k] = 0;
    b->members[k] = nullptr;
    b->ncases--;
    c->batch = nullptr;
    return ts_upload(b->ts, b->st, b->err);
}

int hp_batch_set_grid(hp_batch *b, int cell_blocks, int quad_blocks, int ghost_blocks)
{
    if (!b || cell_blocks < 0 || quad_blocks < 0 || ghost_blocks < 0)
        return bfail(b, HP_ERR_ARG, "hp_batch_set_grid: bad argument");
    b->grid_cell = cell_blocks;
    b->grid_quad = quad_blocks;
    b->grid_ghost = ghost_blocks;
    batch_grids(b);
    batch_drop_graph(b);
    return HP_OK;
}

int hp_batch_set_block(hp_batch *b, int tiles_per_block)
{
    if (!b || (tiles_per_block != 1 && tiles_per_block != 2 && tiles_per_block != 4))
        return bfail(b, HP_ERR_ARG, "hp_batch_set_block: tiles_per_block must be 1, 2 or 4");
    b->tiles_per_block = tiles_per_block;
    batch_grids(b);
    batch_drop_graph(b);
    return HP_OK;
}

int hp_batch_step(hp_batch *b, int nsteps, int use_graph)
{
    if (!b || nsteps < 0) return bfail(b, HP_ERR_ARG, "hp_batch_step: bad argument");
    BHIPCHK(b, hipEventRecord(b->e0, b->st));
    if (b->ncases > 0 && nsteps > 0) {
        const Batch B = b->ts.view();
        if (use_graph) {
            if (!b->exec) {
                BHIPCHK(b, hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
                enqueue_step(B, b->G, b->st, nullptr);
                hipGraph_t g = nullptr;
                BHIPCHK(b, hipStreamEndCapture(b->st, &g));
                b->graph = g;
                BHIPCHK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            }
            for (int n = 0; n < nsteps; n++) BHIPCHK(b, hipGraphLaunch(b->exec, b->st));
        } else {
            // eager launches know the current tile counts: untuned families
            // get exactly a block per tile
            Grid G = b->G;
            if (!b->grid_cell) G.cell = round8(b->ts.nct);
            if (!b->grid_quad) G.quad = round8(b->ts.nct);
            if (!b->grid_ghost) G.ghost = round8(b->ts.ngt);
            G.persist = (G.cell < b->ts.nct ? 1 : 0) | (G.quad < b->ts.nct ? 2 : 0) | (G.ghost < b->ts.ngt ? 4 : 0);
            cellw_grid(G, b->ts.nct);
            for (int n = 0; n < nsteps; n++) enqueue_step(B, G, b->st, nullptr);
        }
    }
    BHIPCHK(b, hipEventRecord(b->e1, b->st));
    BHIPCHK(b, hipGetLastError());
    return HP_OK;
}

int hp_batch_health(hp_batch *b, int *status, int n)
{
    const int slots = b ? (int)b->members.size() : 0;
    if (!b || !status || n < slots) return bfail(b, HP_ERR_ARG, "hp_batch_health: need one int per case slot");
    BHIPCHK(b, hipMemsetAsync(b->health, 0, sizeof(int) * slots, b->st));
    if (b->ncases > 0) {
        Grid G = b->G;
        if (!b->grid_cell) G.cell = round8(b->ts.nct);
        G.persist = G.cell < b->ts.nct ? 1 : 0;
        launch_health(b->ts.view(), G, b->health, b->st);
        BHIPCHK(b, hipGetLastError());
    }
    BHIPCHK(b, hipMemcpyAsync(status, b->health, sizeof(int) * slots, hipMemcpyDeviceToHost, b->st));
    BHIPCHK(b, hipStreamSynchronize(b->st));
    return HP_OK;
}

int hp_batch_synchronize(hp_batch *b)
{
    if (!b) return bfail(nullptr, HP_ERR_ARG, "null batch");
    BHIPCHK(b, hipStreamSynchronize(b->st));
    BHIPCHK(b, hipGetLastError());
    return HP_OK;
}

int hp_batch_last_ms(hp_batch *b, double *ms)
{
    if (!b || !ms) return bfail(b, HP_ERR_ARG, "hp_batch_last_ms: bad argument");
    BHIPCHK(b, hipStreamSynchronize(b->st));
    float t = 0.0f;
    BHIPCHK(b, hipEventElapsedTime(&t, b->e0, b->e1));
    *ms = t;
    return HP_OK;
}

const char *hp_batch_last_error(const hp_batch *b) { return b ? b->err.c_str() : g_create_err.c_str(); }

int hp_check_division(const double *a, const double *b, size_t n, size_t *mismatches)
{
    if ((n && (!a || !b)) || !mismatches) return fail(nullptr, HP_ERR_ARG, "hp_check_division: bad argument");
    *mismatches = 0;
    if (!n) return HP_OK;
    double *da = nullptr, *db = nullptr;
    unsigned long long *dbad = nullptr, bad = 0;
    hipError_t e = hipMalloc(&da, n * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&db, n * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dbad, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemcpy(da, a, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, b, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(dbad, 0, sizeof(unsigned long long));
    if (e == hipSuccess) {
        launch_check_division(da, db, n, dbad, nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(&bad, dbad, sizeof(bad), hipMemcpyDeviceToHost);
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dbad);
    if (e != hipSuccess) return fail(nullptr, HP_ERR_HIP, std::string("hp_check_division: ") + hipGetErrorString(e));
    *mismatches = (size_t)bad;
    return HP_OK;
}

size_t hp_case_bytes(int nodes, int neles, int nghosts, int precision)
{
    if (nodes < 0 || neles < 0 || nghosts < 0) return 0;
    const size_t ne = neles, ng = nghosts, nt = ne + ng, D = sizeof(double), I = sizeof(int);
    size_t b = 3 * (size_t)nodes * D;                        // coordinates
    b += (13 * ne + 4 * ng) * I;                             // nod, nc, ncp, iblank, ghost lists
    b += (12 + 3 + 12 + 12) * ne * D + 4 * nt * D;         // sc, dl/epssq/dd, fc0, fc3, xc..vol
    b += (30 + 3 * G_N + 1 + 10 + 2 + 10) * ne * D + 2 * (size_t)Q_N * nt * D + (SC_N + 4) * D + I;
    b += (8 + 10) * ne * (precision == 1 ? sizeof(float) : D);  // viscous fluxes, JST dissipation
    if (precision == 1) {
        b += (size_t)R32_N * nt * sizeof(float) + (size_t)CGEO32_N * ne * sizeof(float);  // FP32-bulk records, geometry
        b += (9 * ne + 1) * I + 4 * ne * (FACE32_N + FGEO32_N) * sizeof(float);  // face tables (HLLC-Batten; at most)
        b += SRC32_N * ne * sizeof(float);                                   // its source terms
    }
    else b += SRC_N * ne * D;                                     // interphase source terms
    const size_t tc = (ne + TILE - 1) / TILE, tg = (ng + TILE - 1) / TILE + 1;
    b += sizeof(Params) + sizeof(State) + sizeof(Mesh) + (tc + tg) * sizeof(Tile) + 4 * I + 4 * tc * D;
    return b + 64 * 4096;  // allocation granularity of the ~60 buffers
}

}  // extern "C"
