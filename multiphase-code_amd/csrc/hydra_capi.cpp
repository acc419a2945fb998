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
#include <vector>

#include "hydra_dev.h"
#include "hydra_hip.h"

using namespace hydra;

struct hp_ctx {
    hipStream_t st = nullptr;
    Params P{};
    Mesh m{};
    State s{};
    std::vector<void *> bufs;
    double *dtpart = nullptr;
    int nparts = 0;
    double *endpart = nullptr;
    int nendparts = 0;
    std::string err;
    // timing of the last hp_step
    std::vector<hipEvent_t> fev;  // start/stop pairs around flux launches
    hipEvent_t t0 = nullptr, t1 = nullptr;
    int nflux_timed = 0;
};

namespace {

thread_local std::string g_create_err;

int fail(hp_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    else g_create_err = msg;
    return code;
}

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
    double xmax = -1.0e30, ymax = -1.0e30, zmax = -1.0e30, ymin = 1.0e30, zmin = 1.0e30;
    for (int n = 0; n < nodes; n++) {
        xmax = std::max(xmax, x[n]);
        ymax = std::max(ymax, y[n]);
        zmax = std::max(zmax, z[n]);
        ymin = std::min(ymin, y[n]);
        zmin = std::min(zmin, z[n]);
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
    ALLOC(s.q, (size_t)Q_N * ntot) ALLOC(s.grad, (size_t)3 * G_N * ne) ALLOC(s.phi, ne)
    ALLOC(s.vis, 8 * (size_t)ne) ALLOC(s.d24, 10 * (size_t)ne) ALLOC(s.d2q, 10 * (size_t)ne)
    ALLOC(s.anu, 2 * (size_t)ne) ALLOC(s.rhs, 10 * (size_t)ne) ALLOC(s.scal, SC_N) ALLOC(s.itot, 1)
    if (params->precision == 1) {
        ALLOC(s.g32, (size_t)3 * G_N * ne) ALLOC(s.phi32, ne)
    }
    c->nparts = (ne + 255) / 256;
    ALLOC(c->dtpart, c->nparts)
    c->nendparts = (ne + 255) / 256;
    ALLOC(c->endpart, 4 * (size_t)c->nendparts)
#undef ALLOC
    m.x = dx; m.y = dy; m.z = dz; m.nod = dnod; m.nc = dnc; m.ncp = dncp; m.iblank = dib;
    m.gpar = dgpar; m.gtype = dgtype; m.gside = dgside; m.gdd_win = dgwin;
    s.dl = m.dl;
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
                (rc = h2d(dgside, gside.data(), sizeof(int) * ng)) || (rc = h2d(dgwin, gwin.data(), sizeof(int) * ng))))) {
        g_create_err = c->err;
        hp_destroy(c);
        return rc;
    }
    launch_geometry(m, xmax, ymax, zmax, ymin, zmin, c->st);
    e = hipStreamSynchronize(c->st);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
        g_create_err = std::string("geometry: ") + hipGetErrorString(e);
        hp_destroy(c);
        return HP_ERR_HIP;
    }
    *out = c;
    return HP_OK;
}

int hp_destroy(hp_ctx *c)
{
    if (!c) return HP_OK;
    if (c->st) (void)hipStreamSynchronize(c->st);
    for (void *p : c->bufs) (void)hipFree(p);
    for (hipEvent_t e : c->fev) (void)hipEventDestroy(e);
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
    const int slot[6] = {Q_UG, Q_VG, Q_WG, Q_PG, Q_TG, Q_ROG};
    for (int k = 0; k < 6 && ng > 0; k++)
        HIPCHK(c, hipMemcpyAsync(c->s.q + (size_t)slot[k] * c->s.nt + ne, gp + (size_t)k * ng, sizeof(double) * ng,
                                 hipMemcpyHostToDevice, c->st));
    // main.F90:456-476: viscous fluxes and dissipation start at zero
    HIPCHK(c, hipMemsetAsync(c->s.vis, 0, sizeof(double) * 8 * ne, c->st));
    HIPCHK(c, hipMemsetAsync(c->s.d24, 0, sizeof(double) * 10 * ne, c->st));
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

int hp_primitive(hp_ctx *c)
{
    if (!c) return fail(nullptr, HP_ERR_ARG, "null context");
    launch_eos(c->P, c->s, 1, nullptr, c->st);
    return check_launch(c, "primitive");
}

int hp_dt_min(hp_ctx *c, double *dtmin)
{
    if (!c || !dtmin) return fail(c, HP_ERR_ARG, "hp_dt_min: bad argument");
    launch_dt_from_prims(c->P, c->s, c->dtpart, c->st);
    if (int rc = check_launch(c, "dt")) return rc;
    std::vector<double> part(c->nparts);
    HIPCHK(c, hipMemcpyAsync(part.data(), c->dtpart, sizeof(double) * c->nparts, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    double r = 1.0e8;
    for (double v : part) r = v < r ? v : r;
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
    launch_smooth(c->P, c->s, c->m, c->st);
    return check_launch(c, "smooth");
}

int hp_viscous(hp_ctx *c)
{
    if (!c) return fail(nullptr, HP_ERR_ARG, "null context");
    launch_bc(c->P, c->s, c->m, 1, c->st);
    launch_viscous(c->P, c->s, c->m, c->st);
    return check_launch(c, "viscous");
}

static void flux_solver(hp_ctx *c, int stage)
{
    launch_eos(c->P, c->s, 1, nullptr, c->st);  // primitiveplusbc, interior
    launch_bc(c->P, c->s, c->m, 0, c->st);      // primitiveplusbc, ghosts
    if (c->P.myorder == 2) launch_grad(c->P, c->s, c->m, c->st);
    launch_flux(c->P, c->s, c->m, stage, c->st);
}

int hp_flux(hp_ctx *c)
{
    if (!c) return fail(nullptr, HP_ERR_ARG, "null context");
    flux_solver(c, 0);
    return check_launch(c, "flux");
}

int hp_rk_stage(hp_ctx *c, int stage)
{
    if (!c || stage < 1 || stage > 4) return fail(c, HP_ERR_ARG, "hp_rk_stage: stage must be 1..4");
    launch_rk_update(c->s, c->m, stage, c->st);
    return check_launch(c, "rk_stage");
}

int hp_rk(hp_ctx *c)
{
    if (!c) return fail(nullptr, HP_ERR_ARG, "null context");
    launch_bc(c->P, c->s, c->m, 1, c->st);
    launch_viscous(c->P, c->s, c->m, c->st);
    for (int st = 1; st <= 4; st++) {
        flux_solver(c, 0);
        launch_rk_update(c->s, c->m, st, c->st);
    }
    return check_launch(c, "rk");
}

int hp_end_step(hp_ctx *c, double *diffs)
{
    if (!c) return fail(nullptr, HP_ERR_ARG, "null context");
    launch_eos(c->P, c->s, 1, nullptr, c->st);
    launch_end_step(c->s, diffs ? c->endpart : nullptr, c->st);
    if (int rc = check_launch(c, "end_step")) return rc;
    if (diffs) {
        std::vector<double> part(4 * (size_t)c->nendparts);
        HIPCHK(c, hipMemcpyAsync(part.data(), c->endpart, sizeof(double) * part.size(), hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        double d[4] = {0, 0, 0, 0};
        for (int b = 0; b < c->nendparts; b++)
            for (int k = 0; k < 4; k++) d[k] = std::max(d[k], part[(size_t)4 * b + k]);
        // main.F90:653 prints diffu, diffv, diffw, diffp = variables 2,3,4,1
        diffs[0] = d[1]; diffs[1] = d[2]; diffs[2] = d[3]; diffs[3] = d[0];
    }
    return HP_OK;
}

int hp_step(hp_ctx *c, int nsteps)
{
    if (!c || nsteps < 0) return fail(c, HP_ERR_ARG, "hp_step: bad argument");
    const int need = 8 * nsteps;
    while ((int)c->fev.size() < need) {
        hipEvent_t e;
        HIPCHK(c, hipEventCreate(&e));
        c->fev.push_back(e);
    }
    if (!c->t0) {
        HIPCHK(c, hipEventCreate(&c->t0));
        HIPCHK(c, hipEventCreate(&c->t1));
    }
    HIPCHK(c, hipEventRecord(c->t0, c->st));
    int ev = 0;
    for (int n = 0; n < nsteps; n++) {
        // main.F90:515-558: primitive, total+1, dt = cfl*min(...), time+dt
        launch_eos(c->P, c->s, 1, c->dtpart, c->st);
        launch_dt_final(c->P, c->s, c->dtpart, c->nparts, c->st);
        launch_smooth(c->P, c->s, c->m, c->st);          // :563
        launch_bc(c->P, c->s, c->m, 1, c->st);           // sr_RK.F90:19
        launch_viscous(c->P, c->s, c->m, c->st);
        for (int st = 1; st <= 4; st++) {
            launch_eos(c->P, c->s, 1, nullptr, c->st);
            launch_bc(c->P, c->s, c->m, 0, c->st);
            if (c->P.myorder == 2) launch_grad(c->P, c->s, c->m, c->st);
            HIPCHK(c, hipEventRecord(c->fev[ev++], c->st));
            launch_flux(c->P, c->s, c->m, st, c->st);
            HIPCHK(c, hipEventRecord(c->fev[ev++], c->st));
        }
        launch_eos(c->P, c->s, 1, nullptr, c->st);       // main.F90:595
        launch_end_step(c->s, nullptr, c->st);           // :619-625
    }
    HIPCHK(c, hipEventRecord(c->t1, c->st));
    c->nflux_timed = ev / 2;
    return check_launch(c, "step");
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

int hp_probe_cell(hp_ctx *c, int cell, double *prims)
{
    if (!c || !prims || cell < 1 || cell > c->s.nt) return fail(c, HP_ERR_ARG, "hp_probe_cell: bad argument");
    for (int k = 0; k < Q_N; k++)
        HIPCHK(c, hipMemcpyAsync(prims + k, c->s.q + (size_t)k * c->s.nt + (cell - 1), sizeof(double),
                                 hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return HP_OK;
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
    double *p;
    size_t n;
    if (int rc = array_view(c, which, &p, &n)) return rc;
    if (count != n) return fail(c, HP_ERR_ARG, "hp_get_array: count mismatch");
    HIPCHK(c, hipMemcpyAsync(host, p, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return HP_OK;
}

int hp_set_array(hp_ctx *c, int which, const double *host, size_t count)
{
    if (!c || !host) return fail(c, HP_ERR_ARG, "hp_set_array: bad argument");
    double *p;
    size_t n;
    if (int rc = array_view(c, which, &p, &n)) return rc;
    if (count != n) return fail(c, HP_ERR_ARG, "hp_set_array: count mismatch");
    HIPCHK(c, hipMemcpyAsync(p, host, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return HP_OK;
}

}  // extern "C"
