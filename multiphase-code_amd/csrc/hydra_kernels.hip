// hydra_kernels.hip -- FP64 HIP kernels of the HydraPhase per-step finite
// volume update for MI355X (gfx950, wave64).
//
// One thread owns one tetrahedral cell and evaluates that cell's own four
// faces, as the reference does (sr_hllcbatten.F90:59-556): the HLLC flux
// carries non-antisymmetric regularisers (+1.0E-6 in the star-state
// denominators, +1.0E-8 in the contact speed), so a face-once scatter would
// not reproduce the reference's residuals.  Residuals are gathered in
// registers, never scattered: no atomics, bitwise-deterministic results.
//
// Expressions follow the reference's left-to-right FP64 evaluation, with its
// REAL*4 literals promoted to double; the library is built with
// -ffp-contract=off so no multiply-add is fused.  Citations are relative to
// HydraPhase_60_FPC_5mill/src/.
//
// Data layout in HBM (structure of arrays, variable-major, leading dimension
// neles unless noted): conserved cn/cu/cn1 (10), primitives q (14 x ntot:
// interior cells then boundary ghosts), gradients (33), limiter phi, viscous
// fluxes (8), JST d24 (10).  Boundary ghosts are reached through nc >= neles;
// the JST routine reads a ghost through its parent (nc_par) because the
// reference copies parent values into the ghost right before using them
// (sr_smooth.F90:19-41, :164-187).

#include <hip/hip_runtime.h>

#include "hydra_dev.h"
#include "hydra_math.h"

#define F1E8 ((double)1.0e-8f)
#define F1E6 ((double)1.0e-6f)
#define F0P1 ((double)0.1f)

namespace hydra {

// four floats as the compiler's native vector (non-temporal stores of
// 16-byte pieces: the stage's records, entries and state are written once
// and read by the next launch, so they bypass the caches the gathers use)
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }

// mod_MatrixOps.F90:4-9
__device__ __forceinline__ double det3(double a11, double a21, double a31, double a12, double a22,
                                       double a32, double a13, double a23, double a33)
{
    return a11 * (a22 * a33 - a23 * a32) - a21 * (a12 * a33 - a13 * a32) + a31 * (a12 * a23 - a13 * a22);
}

// mod_MatrixOps.F90:12-18
__device__ __forceinline__ double fvenkat(double delp, double delm, double epssq)
{
    return (((delp * delp + epssq) * delm + 2.0 * delm * delm * delp) /
            (delp * delp + 2.0 * delm * delm + delp * delm + epssq)) / delm;
}

#define QV(s, i) q[(size_t)(s) * nt + (i)]

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2 cache.  A tile loop gives the blocks of XCD x (blockIdx % 8 == x) the
// x-th contiguous eighth of the tile table, so tiles that are neighbours in
// memory order meet in one L2; within its eighth an XCD's blocks stride by
// gridDim/8.  Grids are multiples of 8 blocks.
__device__ __forceinline__ void tile_span(int n, int &t, int &end, int &step)
{
    const int x = (int)(blockIdx.x & 7u);
    step = (int)(gridDim.x >> 3);
    t = (int)(((long long)n * x) >> 3) + (int)(blockIdx.x >> 3);
    end = (int)(((long long)n * (x + 1)) >> 3);
}
// The streaming per-cell kernels (no LDS, no barrier) take blocks of
// blockDim.x / TILE tiles, one per wave (the batch's tiles_per_block): the
// block's waves are its virtual blocks, consecutive in the XCD's eighth.
__device__ __forceinline__ void tile_span_w(int n, int &t, int &end, int &step)
{
    // the wave's tile within the block: wave-uniform, kept scalar so the
    // tile's case is read by scalar loads as in the one-tile blocks
    const int tpb = (int)(blockDim.x / TILE), sub = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / TILE));
    const int x = (int)(blockIdx.x & 7u);
    step = (int)(gridDim.x >> 3) * tpb;
    t = (int)(((long long)n * x) >> 3) + (int)(blockIdx.x >> 3) * tpb + sub;
    end = (int)(((long long)n * (x + 1)) >> 3);
}
// MODE (a template parameter of every tile kernel): 0 tables, grid with a
// block for every tile the tables can hold (at most one tile per block);
// 1 tables, persistent grid (blocks loop over tiles); 2 one case carried in
// the launch arguments (B.P1, B.S1, B.M1), a block per tile, no table reads.
#define TILE_LOOP(NT, N1)                                                    \
    int t_ = 0, e_ = 0, st_ = 1;                                             \
    tile_span(MODE == 2 ? ((N1) + TILE - 1) / TILE : (NT), t_, e_, st_);     \
    for (int t = t_; t < e_; t = (MODE == 1) ? t + st_ : e_)
// the same over blocks of several tiles (tile_span_w; lane = threadIdx.x % TILE)
#define TILE_LOOP_W(NT, N1)                                                  \
    int t_ = 0, e_ = 0, st_ = 1;                                             \
    tile_span_w(MODE == 2 ? ((N1) + TILE - 1) / TILE : (NT), t_, e_, st_);   \
    for (int t = t_; t < e_; t = (MODE == 1) ? t + st_ : e_)

// the case of tile t and its cell (or ghost) range; by value, so stores
// through the state pointers cannot alias it
struct TileCase {
    Params P;
    State S;
    Mesh M;
    int first, count;
};

template <int MODE>
__device__ __forceinline__ TileCase tile_case(const Batch &B, const Tile *tab, int t, int n1)
{
    TileCase c;
    if (MODE == 2) {
        c.P = B.P1;
        c.S = B.S1;
        c.M = B.M1;
        c.first = t * TILE;
        c.count = n1 - c.first < TILE ? n1 - c.first : TILE;
    } else {
        const Tile T = tab[t];
        c.P = B.P[T.cs];
        c.S = B.S[T.cs];
        c.M = B.M[T.cs];
        c.first = T.first;
        c.count = T.count;
    }
    if (B.swap) {
        double *t = c.S.q;
        c.S.q = c.S.q2;
        c.S.q2 = t;
    }
    return c;
}

// lanes 4c..4c+3 of a wave cooperate on cell c: one face each
__device__ __forceinline__ double quad_get(double v, int qbase, int j) { return __shfl(v, qbase + j, 64); }

__device__ __forceinline__ double quad_sum4(double v, int qbase)
{
    // ((t0 + t1) + t2) + t3 in lane order
    const double t0 = quad_get(v, qbase, 0), t1 = quad_get(v, qbase, 1), t2 = quad_get(v, qbase, 2),
                 t3 = quad_get(v, qbase, 3);
    return t0 + t1 + t2 + t3;
}

// Lane J of this lane's quad by a DPP quad permutation (a VALU move; no
// memory or LDS traffic).  Callers run with all four lanes of the quad active.
template <int J>
__device__ __forceinline__ int quad_lane32(int v)
{
    static_assert(J >= 0 && J < 4, "quad lane");
    return __builtin_amdgcn_update_dpp(0, v, J * 0x55, 0xf, 0xf, false);
}
template <int J>
__device__ __forceinline__ double quad_lane(double v)
{
    return __hiloint2double(quad_lane32<J>(__double2hiint(v)), quad_lane32<J>(__double2loint(v)));
}
template <int J>
__device__ __forceinline__ float quad_lane(float v)
{
    return __int_as_float(quad_lane32<J>(__float_as_int(v)));
}

// Lane j ^ D of this lane's quad (D = 1 or 2), a DPP quad permutation
template <int D>
__device__ __forceinline__ float quad_xor(float v)
{
    static_assert(D == 1 || D == 2, "quad xor");
    constexpr int ctrl = D == 1 ? 0xB1 : 0x4E;  // quad_perm [1,0,3,2] / [2,3,0,1]
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xf, 0xf, false));
}
__device__ __forceinline__ float sel(bool c, float a, float b) { return c ? a : b; }

// In-quad transpose of sixteen-byte items: lane j (j = lane & 3) holds
// a[q] = item (j, q) of its quad and ends with a[c] = item (c, j), by two
// butterfly stages (the partner lanes j ^ 1, then j ^ 2).  All four lanes
// of the quad take part.
__device__ __forceinline__ void quad_transpose(float4 (&a)[4], int j)
{
    float4 t[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const bool c = (j ^ q) & 1;
        const float4 o = a[q ^ 1];
        t[q] = make_float4(sel(c, quad_xor<1>(o.x), a[q].x), sel(c, quad_xor<1>(o.y), a[q].y),
                           sel(c, quad_xor<1>(o.z), a[q].z), sel(c, quad_xor<1>(o.w), a[q].w));
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const bool c = (j ^ q) & 2;
        const float4 o = t[q ^ 2];
        a[q] = make_float4(sel(c, quad_xor<2>(o.x), t[q].x), sel(c, quad_xor<2>(o.y), t[q].y),
                           sel(c, quad_xor<2>(o.z), t[q].z), sel(c, quad_xor<2>(o.w), t[q].w));
    }
}

// The R32_N-float record of this lane's cell cq, read a quad at a time:
// each load instruction takes four consecutive sixteen-byte pieces of one
// of the quad's four records (one 64-byte line per quad), and in-quad
// transposes hand every lane its own record.  All four lanes take part.
__device__ __forceinline__ void quad_record32(const float4 *rec, int cq, int j, float (&r)[R32_N])
{
    int c[4];
    c[0] = quad_lane32<0>(cq); c[1] = quad_lane32<1>(cq); c[2] = quad_lane32<2>(cq); c[3] = quad_lane32<3>(cq);
#pragma unroll
    for (int k = 0; k < R32_N / 16; k++) {
        float4 a[4];
#pragma unroll
        for (int q = 0; q < 4; q++) a[q] = rec[(size_t)(R32_N / 4) * c[q] + 4 * k + j];
        quad_transpose(a, j);
#pragma unroll
        for (int p = 0; p < 4; p++) {
            r[16 * k + 4 * p] = a[p].x; r[16 * k + 4 * p + 1] = a[p].y;
            r[16 * k + 4 * p + 2] = a[p].z; r[16 * k + 4 * p + 3] = a[p].w;
        }
    }
}

// Planes 0..NP-1 of one cell shared by the four lanes of its quad: lane f
// loads planes f, f+4, f+8, ... (the last clamped to NP-1) into v, so the
// quad issues a quarter of the loads; quad_plane<P> hands plane P to every
// lane of the quad.
template <int NP, typename T>
__device__ __forceinline__ void quad_load(const T *a, size_t stride, int i, int f, T (&v)[(NP + 3) / 4])
{
#pragma unroll
    for (int j = 0; j < (NP + 3) / 4; j++) {
        int p = 4 * j + f;
        if (p > NP - 1) p = NP - 1;
        v[j] = a[(size_t)p * stride + i];
    }
}
template <int P, typename T, int N>
__device__ __forceinline__ T quad_plane(const T (&v)[N])
{
    static_assert(P >= 0 && (P >> 2) < N, "plane");
    return quad_lane<P & 3>(v[P >> 2]);
}

// The row layout of the viscous kernels: the wave's 16 cells c hold their
// four faces in lanes c, c+16, c+32, c+48, so a load instruction's lanes
// take one face of 16 consecutive cells (measured faster there; the flux
// keeps quads, whose many shared planes move by DPP).  Lane J of this
// lane's cell by a lane shuffle.
template <int J>
__device__ __forceinline__ double row_lane(double v)
{
    return __shfl(v, (int)(threadIdx.x & 15) + 16 * J, 64);
}
template <int P, int N>
__device__ __forceinline__ double row_plane(const double (&v)[N])
{
    static_assert(P >= 0 && (P >> 2) < N, "plane");
    return row_lane<P & 3>(v[P >> 2]);
}

// The four rows' values of this lane's column (b[r] = the value row r
// holds), by gfx950's row swaps: v_permlane16_swap exchanges rows 1 and 3
// of one operand with rows 0 and 2 of the other, v_permlane32_swap the
// upper half of one with the lower half of the other.  Three VALU swaps per
// 32-bit word hand every row all four values, with no LDS traffic.  Every
// value moves within its column, so a column whose lanes are all inactive
// leaves the others' results untouched.
__device__ __forceinline__ void row_spread(int v, int (&b)[4])
{
    const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);      // [r0 r0 r2 r2], [r1 r1 r3 r3]
    const auto a = __builtin_amdgcn_permlane32_swap(p[0], p[0], false, false);  // [r0 r0 r0 r0], [r2 r2 r2 r2]
    const auto c = __builtin_amdgcn_permlane32_swap(p[1], p[1], false, false);  // [r1 r1 r1 r1], [r3 r3 r3 r3]
    b[0] = (int)a[0];
    b[1] = (int)c[0];
    b[2] = (int)a[1];
    b[3] = (int)c[1];
}
__device__ __forceinline__ void row_spread(double v, double (&b)[4])
{
    int h[4], l[4];
    row_spread(__double2hiint(v), h);
    row_spread(__double2loint(v), l);
#pragma unroll
    for (int r = 0; r < 4; r++) b[r] = __hiloint2double(h[r], l[r]);
}

// Row J's value of this lane's column alone (two swaps per 32-bit word)
template <int J>
__device__ __forceinline__ int row_take(int v)
{
    static_assert(J >= 0 && J < 4, "row");
    const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    const int h = (J & 1) ? (int)p[1] : (int)p[0];
    const auto a = __builtin_amdgcn_permlane32_swap(h, h, false, false);
    return (J & 2) ? (int)a[1] : (int)a[0];
}
template <int J>
__device__ __forceinline__ double row_take(double v)
{
    return __hiloint2double(row_take<J>(__double2hiint(v)), row_take<J>(__double2loint(v)));
}
// plane P of the planes a column loaded with quad_load (row r holds
// planes r, r+4, ...)
template <int P, int N>
__device__ __forceinline__ double col_plane(const double (&v)[N])
{
    static_assert(P >= 0 && (P >> 2) < N, "plane");
    return row_take<P & 3>(v[P >> 2]);
}

// ---------------------------------------------------------------------------
// geometry (sr_geometry.F90:147-289), once per mesh
// ---------------------------------------------------------------------------
__global__ void k_geom_cell(Mesh m)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.ne) return;
    const int ne = m.ne;
    const double *x = m.x, *y = m.y, *z = m.z;
    int n[4];
    for (int k = 0; k < 4; k++) n[k] = m.nod[(size_t)k * ne + i];
    // area vectors, face node triples of sr_geometry.F90:149-158
    const int tri[4][3] = {{1, 2, 3}, {3, 2, 0}, {1, 3, 0}, {2, 1, 0}};
    for (int f = 0; f < 4; f++) {
        const int n1 = n[tri[f][0]], n2 = n[tri[f][1]], n3 = n[tri[f][2]];
        double dxyA = (x[n1] - x[n2]) * (y[n1] + y[n2]);
        double dxyB = (x[n2] - x[n3]) * (y[n2] + y[n3]);
        double dxyC = (x[n3] - x[n1]) * (y[n3] + y[n1]);
        double dyzA = (y[n1] - y[n2]) * (z[n1] + z[n2]);
        double dyzB = (y[n2] - y[n3]) * (z[n2] + z[n3]);
        double dyzC = (y[n3] - y[n1]) * (z[n3] + z[n1]);
        double dzxA = (z[n1] - z[n2]) * (x[n1] + x[n2]);
        double dzxB = (z[n2] - z[n3]) * (x[n2] + x[n3]);
        double dzxC = (z[n3] - z[n1]) * (x[n3] + x[n1]);
        m.sc[(size_t)(3 * f + 0) * ne + i] = 0.5 * (dyzA + dyzB + dyzC);
        m.sc[(size_t)(3 * f + 1) * ne + i] = 0.5 * (dzxA + dzxB + dzxC);
        m.sc[(size_t)(3 * f + 2) * ne + i] = 0.5 * (dxyA + dxyB + dxyC);
    }
    const double x1 = x[n[0]], x2 = x[n[1]], x3 = x[n[2]], x4 = x[n[3]];
    const double y1 = y[n[0]], y2 = y[n[1]], y3 = y[n[2]], y4 = y[n[3]];
    const double z1 = z[n[0]], z2 = z[n[1]], z3 = z[n[2]], z4 = z[n[3]];
    m.xc[i] = 0.25 * (x1 + x2 + x3 + x4);
    m.yc[i] = 0.25 * (y1 + y2 + y3 + y4);
    m.zc[i] = 0.25 * (z1 + z2 + z3 + z4);
    double d12 = sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2) + (z1 - z2) * (z1 - z2));
    double d13 = sqrt((x1 - x3) * (x1 - x3) + (y1 - y3) * (y1 - y3) + (z1 - z3) * (z1 - z3));
    double d14 = sqrt((x1 - x4) * (x1 - x4) + (y1 - y4) * (y1 - y4) + (z1 - z4) * (z1 - z4));
    double d23 = sqrt((x2 - x3) * (x2 - x3) + (y2 - y3) * (y2 - y3) + (z2 - z3) * (z2 - z3));
    double d42 = sqrt((x4 - x2) * (x4 - x2) + (y4 - y2) * (y4 - y2) + (z4 - z2) * (z4 - z2));
    double d43 = sqrt((x4 - x3) * (x4 - x3) + (y4 - y3) * (y4 - y3) + (z4 - z3) * (z4 - z3));
    double dl = dmin(dmin(dmin(dmin(dmin(d12, d13), d14), d23), d42), d43);
    m.dl[i] = dl;
    double v6 = det3(x2, x3, x4, y2, y3, y4, z2, z3, z4) - det3(x1, x3, x4, y1, y3, y4, z1, z3, z4) +
                det3(x1, x2, x4, y1, y2, y4, z1, z2, z4) - det3(x1, x2, x3, y1, y2, y3, z1, z2, z3);
    m.vol[i] = fabs(v6) / 6.0;
    // limiter threshold (0.3*dl)**3.0d0, evaluated as (t*t)*t (sr_VK.F90:34)
    const double t = 0.3 * dl;
    m.epssq[i] = t * t * t;
    // face centroids: flux/limiter variant (sum order of sr_hllcbatten.F90:79-119,
    // times 0.3333D0) and viscous variant (sr_viscousgp.F90:499-530, / 3.0d0)
    const int fl[4][3] = {{1, 3, 2}, {2, 3, 0}, {3, 1, 0}, {1, 2, 0}};
    const int fv[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    for (int f = 0; f < 4; f++) {
        const int a = n[fl[f][0]], b = n[fl[f][1]], d = n[fl[f][2]];
        m.fc0[(size_t)(3 * f + 0) * ne + i] = (x[a] + x[b] + x[d]) * 0.3333;
        m.fc0[(size_t)(3 * f + 1) * ne + i] = (y[a] + y[b] + y[d]) * 0.3333;
        m.fc0[(size_t)(3 * f + 2) * ne + i] = (z[a] + z[b] + z[d]) * 0.3333;
        const int a3 = n[fv[f][0]], b3 = n[fv[f][1]], d3 = n[fv[f][2]];
        m.fc3[(size_t)(3 * f + 0) * ne + i] = (x[a3] + x[b3] + x[d3]) / 3.0;
        m.fc3[(size_t)(3 * f + 1) * ne + i] = (y[a3] + y[b3] + y[d3]) / 3.0;
        m.fc3[(size_t)(3 * f + 2) * ne + i] = (z[a3] + z[b3] + z[d3]) / 3.0;
    }
}

// ghost centroid mirrored through the boundary face; ghost volume = parent's
// (sr_geometry.F90:245-289).  dd, the distance used by the relaxing BCs, is
// written by the ghost the host marked as the last one of its parent with
// type 3/51/53 in list order (sr_geometry.F90:295-347 overwrites).
__global__ void k_geom_ghost(Mesh m, double xmax, double ymax, double zmax, double ymin, double zmin)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= m.ng) return;
    const int ne = m.ne, np = m.gpar[g], ng = ne + g, ns = m.gside[g];
    const double *x = m.x, *y = m.y, *z = m.z;
    int n[4];
    for (int k = 0; k < 4; k++) n[k] = m.nod[(size_t)k * ne + np];
    const int tri[4][3] = {{1, 2, 3}, {3, 2, 0}, {1, 3, 0}, {2, 1, 0}};
    const int n1 = n[tri[ns - 1][0]], n2 = n[tri[ns - 1][1]], n3 = n[tri[ns - 1][2]];
    double xcf = (x[n1] + x[n2] + x[n3]) / 3.0;
    double ycf = (y[n1] + y[n2] + y[n3]) / 3.0;
    double zcf = (z[n1] + z[n2] + z[n3]) / 3.0;
    m.xc[ng] = 2.0 * xcf - m.xc[np];
    m.yc[ng] = 2.0 * ycf - m.yc[np];
    m.zc[ng] = 2.0 * zcf - m.zc[np];
    m.vol[ng] = m.vol[np];
    if (m.gdd_win[g]) {
        const int nt = m.gtype[g];
        if (nt == 3) {
            double c = (x[n[0]] + x[n[1]] + x[n[2]] + x[n[3]]) * 0.25;
            m.dd[np] = fabs(xmax - c);
        } else if (nt == 51) {
            double c = (y[n[0]] + y[n[1]] + y[n[2]] + y[n[3]]) * 0.25;
            m.dd[np] = c > 0.0 ? fabs(ymax - c) : fabs(ymin - c);
        } else if (nt == 53) {
            double c = (z[n[0]] + z[n[1]] + z[n[2]] + z[n[3]]) * 0.25;
            m.dd[np] = c > F0P1 ? fabs(zmax - c) : fabs(zmin - c);
        }
    }
}

// ---------------------------------------------------------------------------
// primitive recovery: mixture-pressure quadratic, phase clipping and cubic
// blending, conserved rewrite (sr_primitive.F90:23-161,
// sr_primitiveplusbc.F90:36-172).  Optionally the CFL bound of
// main.F90:541-545 reduced to one minimum per block with wave shuffles.
// ---------------------------------------------------------------------------
// Primitive recovery of one cell from its conserved variables c: the
// primitives v (Q_ slots) and the conserved variables rewritten from them, nv
__device__ __forceinline__ void eos_prims(const Params &P, const double (&c)[10], double (&v)[Q_N], double (&nv)[10])
{
    const double c1 = c[0], c2 = c[1], c3 = c[2], c4 = c[3], c5 = c[4], c6 = c[5], c7 = c[6], c8 = c[7],
                 c9 = c[8], c10 = c[9];
    double ug = c2 / c1, vg = c3 / c1, wg = c4 / c1;
    double up = c7 / c6, vp = c8 / c6, wp = c9 / c6;
    double eg = c5 / c1 - 0.5 * (ug * ug + vg * vg + wg * wg);
    double ep = c10 / c6 - 0.5 * (up * up + vp * vp + wp * wp);
    const double gm1 = P.gamag - 1.0, pm1 = P.gamap - 1.0;
    const double sb = P.foursigmabyd + P.pinf * P.gamap;
    double bbb = sb - (c1 * gm1 * eg) - (c6 * pm1 * ep);
    double ccc = (c1 * gm1 * eg * sb);
    double pg = (-bbb + sqrt(bbb * bbb + 4.0 * ccc)) * 0.5;
    double pp = pg + P.foursigmabyd;
    double phig = (c1 * gm1 * eg) / pg;
    double phip = 1.0 - phig;
    double rog = c1 / phig;
    double rop = c6 / phip;
    double tg = pg / (rog * P.rrg);
    double tp = ((pp + P.pinf) * P.gamap) / (pm1 * rop * P.cpp);
    if (phig < P.epslnmin) {
        phig = P.epslnmin; phip = 1.0 - phig;
        ug = up; vg = vp; wg = wp; tg = tp; pg = pp;
    }
    if (phip < P.epslnmin) {
        phip = P.epslnmin; phig = 1.0 - phip;
        up = ug; vp = vg; wp = wg; tp = tg; pp = pg;
    }
    if (phig >= P.epslnmin && phig <= P.epslnmax) {
        double psig = (phig - P.epslnmin) / (P.epslnmax - P.epslnmin);
        double fg = -psig * psig * (2.0 * psig - 3.0);
        ug = fg * ug + (1.0 - fg) * up;
        vg = fg * vg + (1.0 - fg) * vp;
        wg = fg * wg + (1.0 - fg) * wp;
        tg = fg * tg + (1.0 - fg) * tp;
        pg = fg * pg + (1.0 - fg) * pp;
    }
    if (phip >= P.epslnmin && phip <= P.epslnmax) {
        double psig = (phip - P.epslnmin) / (P.epslnmax - P.epslnmin);
        double fg = -psig * psig * (2.0 * psig - 3.0);
        up = fg * up + (1.0 - fg) * ug;
        vp = fg * vp + (1.0 - fg) * vg;
        wp = fg * wp + (1.0 - fg) * wg;
        tp = fg * tp + (1.0 - fg) * tg;
        pp = fg * pp + (1.0 - fg) * pg;
    }
    rog = pg / (P.rrg * tg);
    rop = (pp + P.pinf) * P.gamap / (pm1 * P.cpp * tp);
    eg = pg / (gm1 * rog);
    ep = (pp + P.gamap * P.pinf) / (pm1 * rop);
    const double egt = eg + 0.5 * (ug * ug + vg * vg + wg * wg);
    const double ept = ep + 0.5 * (up * up + vp * vp + wp * wp);
    v[Q_UG] = ug; v[Q_VG] = vg; v[Q_WG] = wg; v[Q_PG] = pg; v[Q_PP] = pp;
    v[Q_TG] = tg; v[Q_PHIG] = phig; v[Q_UP] = up; v[Q_VP] = vp; v[Q_WP] = wp;
    v[Q_PHIP] = phip; v[Q_TP] = tp; v[Q_ROP] = rop; v[Q_ROG] = rog;
    const double mg = phig * rog, mp = phip * rop;
    nv[0] = mg; nv[1] = mg * ug; nv[2] = mg * vg; nv[3] = mg * wg; nv[4] = mg * egt;
    nv[5] = mp; nv[6] = mp * up; nv[7] = mp * vp; nv[8] = mp * wp; nv[9] = mp * ept;
}

// cu (may be null): also store the rewritten conserved variables there, the
// end-of-step cu = cn (main.F90:625) fused into the last primitive recovery
__device__ __forceinline__ void eos_cell(const Params &P, double *cn, int ne, int i, double *q, int nt,
                                         double *dtloc, const double *dl, double *cu = nullptr)
{
    double c[10], v[Q_N], nv[10];
    // streamed once: non-temporal loads and stores
    for (int k = 0; k < 10; k++) c[k] = __builtin_nontemporal_load(&cn[(size_t)k * ne + i]);
    eos_prims(P, c, v, nv);
    for (int k = 0; k < Q_N; k++) __builtin_nontemporal_store(v[k], &QV(k, i));
    for (int k = 0; k < 10; k++) __builtin_nontemporal_store(nv[k], &cn[(size_t)k * ne + i]);
    if (cu)
        for (int k = 0; k < 10; k++) cu[(size_t)k * ne + i] = nv[k];
    if (dtloc) {
        // main.F90:541-545
        const double pg = v[Q_PG], pp = v[Q_PP], rog = v[Q_ROG], rop = v[Q_ROP];
        const double ug = v[Q_UG], vg = v[Q_VG], wg = v[Q_WG], up = v[Q_UP], vp = v[Q_VP], wp = v[Q_WP];
        double ag = sqrt((P.gamag * pg) / rog);
        double ap = sqrt((P.gamap * (pp + P.pinf)) / rop);
        double qgi = sqrt(ug * ug + vg * vg + wg * wg);
        double qpi = sqrt(up * up + vp * vp + wp * wp);
        *dtloc = dl[i] / (dmax(ag, ap) + dmax(qgi, qpi));
    }
}

__device__ __forceinline__ double wave_min(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = dmin(v, __shfl_xor(v, off, 64));
    return v;
}

// A tile's share of the CFL minimum, which the reference takes over the
// cells in order (dtmin = MIN(dtmin, bound), main.F90:545; the serial
// oracle folds dtmin < bound ? dtmin : bound).  Without a NaN bound the
// fold takes the tile's minimum m as dmin(dtmin, m): part[4t] = m,
// part[4t+1] = 0.  A NaN bound restarts the fold, so a tile holding one
// yields the minimum of the bounds after its last NaN (NaN if that is the
// tile's last cell) whatever came before: part[4t+1] = 1.  Lane l holds
// the bound of cell l of the tile's count.
__device__ __forceinline__ void tile_dt_part(const Batch &B, int t, double loc, int l, int count)
{
    const bool valid = l < count;
    const unsigned long long nanm = __ballot(valid && !(loc == loc));
    if (nanm == 0ull) {
        const double v = wave_min(loc);
        if (l == 0) {
            B.part[(size_t)4 * t] = v;
            B.part[(size_t)4 * t + 1] = 0.0;
        }
    } else {
        const int j = 63 - __clzll((long long)nanm);  // the tile's last NaN
        double v = wave_min((valid && l > j) ? loc : __builtin_inf());
        if (j == count - 1) v = __builtin_nan("");
        if (l == 0) {
            B.part[(size_t)4 * t] = v;
            B.part[(size_t)4 * t + 1] = 1.0;
        }
    }
}





// ---------------------------------------------------------------------------
// boundary ghosts (sr_primitiveplusbc.F90:198-592; viscous variant
// sr_viscousgp.F90:71-446).  Each ghost owns its slot, so the five
// reference loops collapse into one launch with a switch on the type.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void bc_ghost(const Params &P, const State &s, const Mesh &m, int g, int viscous)
{
    const int ne = m.ne, nt = s.nt, np = m.gpar[g], ng = ne + g, type = m.gtype[g];
    double *q = s.q;
    const double dt = s.scal[SC_DT];
    const int total = s.itot[0];
    if (type == 2) {
        QV(Q_PHIP, ng) = P.phipzero;
        QV(Q_PHIG, ng) = 1.0 - P.phipzero;
        QV(Q_UG, ng) = P.ugmin; QV(Q_VG, ng) = P.vgmin; QV(Q_WG, ng) = P.wgmin;
        const double tg = P.tgmin;
        QV(Q_TG, ng) = tg;
        const double pg = QV(Q_PG, np);
        QV(Q_PG, ng) = pg;
        QV(Q_ROG, ng) = pg / (P.rrg * tg);
        QV(Q_UP, ng) = 0.99 * P.ugmin; QV(Q_VP, ng) = 0.99 * P.vgmin; QV(Q_WP, ng) = 0.99 * P.wgmin;
        QV(Q_TP, ng) = P.tgmin;
        QV(Q_PP, ng) = QV(Q_PP, np);
        QV(Q_ROP, ng) = P.gamap * (pg + P.pinf) / ((P.gamap - 1.0) * P.cpp * P.tgmin);
        return;
    }
    if (type == 5) {
        QV(Q_PHIG, ng) = QV(Q_PHIG, np);
        QV(Q_UG, ng) = -QV(Q_UG, np); QV(Q_VG, ng) = -QV(Q_VG, np); QV(Q_WG, ng) = -QV(Q_WG, np);
        const double tg = QV(Q_TG, np), tp = QV(Q_TP, np), pg = QV(Q_PG, np), pp = QV(Q_PP, np);
        QV(Q_TG, ng) = tg;
        QV(Q_PHIP, ng) = QV(Q_PHIP, np);
        QV(Q_UP, ng) = -QV(Q_UP, np); QV(Q_VP, ng) = -QV(Q_VP, np); QV(Q_WP, ng) = -QV(Q_WP, np);
        QV(Q_TP, ng) = tp;
        QV(Q_PG, ng) = pg; QV(Q_PP, ng) = pp;
        QV(Q_ROG, ng) = pg / (P.rrg * tg);
        QV(Q_ROP, ng) = P.gamap * (pp + P.pinf) / ((P.gamap - 1.0) * P.cpp * tp);
        return;
    }
    const double dd = m.dd[np];
    if (type == 3) {
        QV(Q_PHIG, ng) = QV(Q_PHIG, np);
        QV(Q_PHIP, ng) = QV(Q_PHIP, np);
        const double ugp = QV(Q_UG, np), vgp = QV(Q_VG, np), wgp = QV(Q_WG, np), tgp = QV(Q_TG, np);
        const double wT = ugp;
        const double wU = ugp + sqrt(P.gamag * QV(Q_PG, np) / QV(Q_ROG, np));
        const double alphaT = wT * dt / dd;
        const double alphaU = wU * dt / dd;
        const double fracT = 1.0 / (1.0 + alphaT);
        const double fracU = 1.0 / (1.0 + alphaU);
        double ugg = QV(Q_UG, ng), vgg = QV(Q_VG, ng), wgg = QV(Q_WG, ng), tgg = QV(Q_TG, ng);
        if (total == 1) { ugg = ugp; vgg = vgp; wgg = wgp; tgg = tgp; }
        if (ugp > ugg) {
            ugg = ugg * fracU + (1.0 - fracU) * (2.0 * ugp - ugg);
            vgg = vgg * fracT + (1.0 - fracT) * (2.0 * vgp - vgg);
            wgg = wgg * fracT + (1.0 - fracT) * (2.0 * wgp - wgg);
            tgg = tgg * fracT + (1.0 - fracT) * (2.0 * tgp - tgg);
        } else {
            ugg = ugp; vgg = vgp; wgg = wgp; tgg = tgp;
        }
        QV(Q_UG, ng) = ugg; QV(Q_VG, ng) = vgg; QV(Q_WG, ng) = wgg; QV(Q_TG, ng) = tgg;
        QV(Q_PG, ng) = P.pmin;
        QV(Q_ROG, ng) = P.pmin / (P.rrg * tgg);
        QV(Q_VP, ng) = QV(Q_VP, np); QV(Q_UP, ng) = QV(Q_UP, np); QV(Q_WP, ng) = QV(Q_WP, np);
        QV(Q_PP, ng) = P.pmin;
        const double tp = QV(Q_TP, np);
        QV(Q_TP, ng) = tp;
        QV(Q_ROP, ng) = P.gamap * (P.pmin + P.pinf) / ((P.gamap - 1.0) * P.cpp * tp);
        return;
    }
    if (type == 51 || type == 53) {
        // tangential relaxation toward the mirrored state, normal velocity
        // reflected (:425-591).  51 reflects v, 53 reflects w; in the viscous
        // copy the 53 branch negates the ghost's own w (sr_viscousgp.F90:408).
        const bool y = (type == 51);
        const double ugp = QV(Q_UG, np), vgp = QV(Q_VG, np), wgp = QV(Q_WG, np), tgp = QV(Q_TG, np);
        const double nrm = y ? vgp : wgp;
        const double alphaT = nrm * dt / dd;
        const double fracT = 1.0 / (1.0 + alphaT);
        double ugg = QV(Q_UG, ng), vgg = QV(Q_VG, ng), wgg = QV(Q_WG, ng), tgg = QV(Q_TG, ng);
        if (y) {
            if (vgp > vgg) {
                vgg = -vgp;
                ugg = ugg * fracT + (1.0 - fracT) * (2.0 * ugp - ugg);
                wgg = wgg * fracT + (1.0 - fracT) * (2.0 * wgp - wgg);
                tgg = tgg * fracT + (1.0 - fracT) * (2.0 * tgp - tgg);
            } else {
                ugg = ugp; vgg = -vgp; wgg = wgp; tgg = tgp;
            }
        } else {
            if (wgp > wgg) {
                wgg = viscous ? -wgg : -wgp;
                ugg = ugg * fracT + (1.0 - fracT) * (2.0 * ugp - ugg);
                vgg = vgg * fracT + (1.0 - fracT) * (2.0 * vgp - vgg);
                tgg = tgg * fracT + (1.0 - fracT) * (2.0 * tgp - tgg);
            } else {
                ugg = ugp; vgg = vgp; wgg = -wgp; tgg = tgp;
            }
        }
        QV(Q_UG, ng) = ugg; QV(Q_VG, ng) = vgg; QV(Q_WG, ng) = wgg; QV(Q_TG, ng) = tgg;
        const double pg = QV(Q_PG, np), pp = QV(Q_PP, np), tp = QV(Q_TP, np);
        QV(Q_PG, ng) = pg;
        QV(Q_PHIG, ng) = QV(Q_PHIG, np);
        QV(Q_PHIP, ng) = QV(Q_PHIP, np);
        QV(Q_UP, ng) = QV(Q_UP, np);
        QV(Q_VP, ng) = y ? -QV(Q_VP, np) : QV(Q_VP, np);
        QV(Q_WP, ng) = y ? QV(Q_WP, np) : -QV(Q_WP, np);
        QV(Q_TP, ng) = tp;
        QV(Q_PP, ng) = pp;
        QV(Q_ROG, ng) = pg / (P.rrg * tgg);
        QV(Q_ROP, ng) = P.gamap * (pp + P.pinf) / ((P.gamap - 1.0) * P.cpp * tp);
    }
}

// ---------------------------------------------------------------------------
// JST dissipation (sr_smooth.F90): pass 1 switches anu + undivided Laplacian
// d2q, pass 2 the blended 2nd/4th difference d24.  Ghost values are the
// parent's (nc_par).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double switch_sum(const double *a, int i, const int *nb, double *den)
{
    const double c = a[i];
    double n = fabs(a[nb[0]] - c) + fabs(a[nb[1]] - c) + fabs(a[nb[2]] - c) + fabs(a[nb[3]] - c);
    *den = fabs(a[nb[0]] + c) + fabs(a[nb[1]] + c) + fabs(a[nb[2]] + c) + fabs(a[nb[3]] + c);
    return n;
}

__device__ __forceinline__ void smooth1_cell(const State &s, const Mesh &m, int i)
{
    const int ne = m.ne;
    if (i >= ne) return;
    int nb[4];
    for (int k = 0; k < 4; k++) nb[k] = m.ncp[(size_t)k * ne + i];
    const double *q = s.q;
    const size_t nt = s.nt;
    double den, num;
    num = switch_sum(q + (size_t)Q_PG * nt, i, nb, &den);
    const double anupg = num / den;
    num = switch_sum(s.cn + (size_t)4 * ne, i, nb, &den);
    double a1 = anupg + num / (den + F1E6);
    num = switch_sum(s.cn, i, nb, &den);
    a1 = a1 + num / (den + F1E6);
    num = switch_sum(q + (size_t)Q_PP * nt, i, nb, &den);
    const double anupp = num / den;
    num = switch_sum(s.cn + (size_t)5 * ne, i, nb, &den);
    double a2 = anupp + num / (den + F1E6);
    num = switch_sum(s.cn + (size_t)9 * ne, i, nb, &den);
    a2 = a2 + num / (den + F1E6);
    // every store after every load: the stores could alias cn for all the
    // compiler knows, and the switch sums' loads of cn 0, 4, 5, 9 are reused
    double d2[10];
    for (int k = 0; k < 10; k++) {
        const double *c = s.cn + (size_t)k * ne;
        const double ci = c[i];
        d2[k] = c[nb[0]] - ci + c[nb[1]] - ci + c[nb[2]] - ci + c[nb[3]] - ci;
    }
    s.anu[i] = a1;
    s.anu[(size_t)ne + i] = a2;
    for (int k = 0; k < 10; k++) s.d2q[(size_t)k * ne + i] = d2[k];
}

__device__ __forceinline__ void smooth2_cell(const Params &P, const State &s, const Mesh &m, int i)
{
    const int ne = m.ne;
    if (i >= ne) return;
    int nb[4];
    for (int k = 0; k < 4; k++) nb[k] = m.ncp[(size_t)k * ne + i];
    const double dt = s.scal[SC_DT];
    const double vi = m.vol[i];
    double d24n[4];
    for (int f = 0; f < 4; f++) d24n[f] = 0.5 * (vi / dt + m.vol[nb[f]] / dt);
    for (int set = 0; set < 2; set++) {
        const double *anu = s.anu + (size_t)set * ne;
        double d2c[4], d4c[4];
        for (int f = 0; f < 4; f++) {
            const double e2 = P.co2 * dmax(anu[i], anu[nb[f]]);
            const double e4 = dmax(0.0, P.co4 - e2);
            d2c[f] = e2 * d24n[f];
            d4c[f] = e4 * d24n[f];
        }
        for (int k = 5 * set; k < 5 * set + 5; k++) {
            const double *c = s.cn + (size_t)k * ne;
            const double *w = s.d2q + (size_t)k * ne;
            const double ci = c[i], wi = w[i];
            double d2 = d2c[0] * (c[nb[0]] - ci) + d2c[1] * (c[nb[1]] - ci) + d2c[2] * (c[nb[2]] - ci) +
                        d2c[3] * (c[nb[3]] - ci);
            double d4 = d4c[0] * (w[nb[0]] - wi) + d4c[1] * (w[nb[1]] - wi) + d4c[2] * (w[nb[2]] - wi) +
                        d4c[3] * (w[nb[3]] - wi);
            if (s.d24f)
                s.d24f[(size_t)k * ne + i] = (float)(d2 - d4);
            else
                s.d24[(size_t)k * ne + i] = d2 - d4;
        }
    }
}

// ---------------------------------------------------------------------------
// viscous face fluxes (sr_viscousgp.F90:450-1139, AMUP_CONSTANT build):
// least-squares gradient at each face centroid from the cell's 4 neighbours
// plus the cell (boundary face) or plus the neighbour's 4 neighbours
// (interior face); Newtonian stress and Fourier heat flux of both phases.
// ---------------------------------------------------------------------------
// row_lane with a lane index that is a constant once the caller's loop is
// unrolled
__device__ __forceinline__ double row_from(double v, int k)
{
    switch (k) {
    case 0: return row_lane<0>(v);
    case 1: return row_lane<1>(v);
    case 2: return row_lane<2>(v);
    default: return row_lane<3>(v);
    }
}

// Face-centred LSQ gradients of the eight face variables and the two
// phases' stress and heat-flux contributions to face f (sr_viscousgp.F90:
// 450-804).  The stencil of an interior face is the four face neighbours of
// cell i, then the four of the cell across face f (m2); that of a boundary
// face is cell i's neighbours and cell i itself (m2[0] = i).  Points 0-3 are
// the same for the cell's four faces: the lane of face k holds point k's
// centroid (xn, yn, zn) and primitives (qn) and hands them over by lane
// shuffles (row_from), so the four lanes run the stencil together, and a
// boundary-face lane drops points 5-7 by selects, leaving every sum the
// five-point one.
__device__ __forceinline__ void viscous_stencil(const Params &P, const double *q, size_t nt, const Mesh &m,
                                                const int (&m2)[4], bool ghost, double xn, double yn, double zn,
                                                const double (&qn)[8], double xf, double yf, double zf,
                                                double scx, double scy, double scz, const double (&fv)[8],
                                                double phigf, double phipf, double (&acc)[8])
{
    const int slot[8] = {Q_UG, Q_VG, Q_WG, Q_TG, Q_UP, Q_VP, Q_WP, Q_TP};
    const double tgf = fv[3];
    const double amug = P.as * sqrt(tgf) / (1.0 + P.ts / tgf);
    const double akg = amug * P.cpg / P.prg;
    const double amup = 0.001;
    const double akp = amup * P.cpp / P.prp;
    double dx[8], dy[8], dz[8];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        dx[k] = row_from(xn, k) - xf;
        dy[k] = row_from(yn, k) - yf;
        dz[k] = row_from(zn, k) - zf;
    }
#pragma unroll
    for (int k = 4; k < 8; k++) {
        dx[k] = m.xc[m2[k - 4]] - xf;
        dy[k] = m.yc[m2[k - 4]] - yf;
        dz[k] = m.zc[m2[k - 4]] - zf;
    }
    double a11 = dx[0] * dx[0], a22 = dy[0] * dy[0], a33 = dz[0] * dz[0];
    double a12 = dx[0] * dy[0], a13 = dx[0] * dz[0], a23 = dy[0] * dz[0];
#pragma unroll
    for (int k = 1; k < 8; k++) {
        const bool use = !(ghost && k >= 5);
        const double n11 = a11 + dx[k] * dx[k], n22 = a22 + dy[k] * dy[k], n33 = a33 + dz[k] * dz[k];
        const double n12 = a12 + dx[k] * dy[k], n13 = a13 + dx[k] * dz[k], n23 = a23 + dy[k] * dz[k];
        a11 = use ? n11 : a11; a22 = use ? n22 : a22; a33 = use ? n33 : a33;
        a12 = use ? n12 : a12; a13 = use ? n13 : a13; a23 = use ? n23 : a23;
    }
    const double dnr = det3(a11, a12, a13, a12, a22, a23, a13, a23, a33);
#pragma unroll
    for (int ph = 0; ph < 2; ph++) {
        const int o = 4 * ph;
        double gr[4][3];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            double dv = row_from(qn[o + v], 0) - fv[o + v];
            double b1 = dv * dx[0], b2 = dv * dy[0], b3 = dv * dz[0];
#pragma unroll
            for (int k = 1; k < 8; k++) {
                const bool use = !(ghost && k >= 5);
                const double qk = k < 4 ? row_from(qn[o + v], k) : QV(slot[o + v], m2[k - 4]);
                dv = qk - fv[o + v];
                const double n1 = b1 + dv * dx[k], n2 = b2 + dv * dy[k], n3 = b3 + dv * dz[k];
                b1 = use ? n1 : b1; b2 = use ? n2 : b2; b3 = use ? n3 : b3;
            }
            gr[v][0] = det3(b1, b2, b3, a12, a22, a23, a13, a23, a33) / dnr;
            gr[v][1] = det3(a11, a12, a13, b1, b2, b3, a13, a23, a33) / dnr;
            gr[v][2] = det3(a11, a12, a13, a12, a22, a23, b1, b2, b3) / dnr;
        }
        // Newtonian stress and Fourier heat flux (:761-804)
        const double mu = ph ? amup : amug, ak = ph ? akp : akg;
        const double uf = fv[o], vf = fv[o + 1], wf = fv[o + 2];
        const double dux = gr[0][0], duy = gr[0][1], duz = gr[0][2];
        const double dvx = gr[1][0], dvy = gr[1][1], dvz = gr[1][2];
        const double dwx = gr[2][0], dwy = gr[2][1], dwz = gr[2][2];
        const double dtx = gr[3][0], dty = gr[3][1], dtz = gr[3][2];
        const double delv = dux + dvy + dwz;
        const double txx = mu * (2.0 * dux - P.c23 * delv);
        const double tyy = mu * (2.0 * dvy - P.c23 * delv);
        const double tzz = mu * (2.0 * dwz - P.c23 * delv);
        const double txy = mu * (dvx + duy);
        const double txz = mu * (dwx + duz);
        const double tyz = mu * (dvz + dwy);
        const double qxx = uf * txx + vf * txy + wf * txz + ak * dtx;
        const double qyy = uf * txy + vf * tyy + wf * tyz + ak * dty;
        const double qzz = uf * txz + vf * tyz + wf * tzz + ak * dtz;
        const double phf = ph ? phipf : phigf;
        acc[o + 0] = phf * (txx * scx + txy * scy + txz * scz);
        acc[o + 1] = phf * (txy * scx + tyy * scy + tyz * scz);
        acc[o + 2] = phf * (txz * scx + tyz * scy + tzz * scz);
        acc[o + 3] = phf * (qxx * scx + qyy * scy + qzz * scz);
    }
}

// One cell per quad, one face per lane; the quad sums the four face
// contributions in face order, vis = (((0+V1)+V2)+V3)+V4, as the reference's
// accumulation into zeroed arrays (sr_viscousgp.F90:41-48, :797-804).
__device__ __forceinline__ void viscous_face(const Params &P, const State &s, const Mesh &m, int i, int f,
                                             int qbase)
{
    const int ne = m.ne;
    if (m.iblank[i] != 1) return;
    const double *q = s.q;
    const size_t nt = s.nt;
    const int slot[8] = {Q_UG, Q_VG, Q_WG, Q_TG, Q_UP, Q_VP, Q_WP, Q_TP};
    double acc[8];
    {
        const double scx = m.sc[(size_t)(3 * f) * ne + i], scy = m.sc[(size_t)(3 * f + 1) * ne + i],
                     scz = m.sc[(size_t)(3 * f + 2) * ne + i];
        const double xf = m.fc3[(size_t)(3 * f) * ne + i], yf = m.fc3[(size_t)(3 * f + 1) * ne + i],
                     zf = m.fc3[(size_t)(3 * f + 2) * ne + i];
        // lane f's neighbour is point f of every face stencil of the quad
        const int neigh = m.nc[(size_t)f * ne + i];
        double qn[8];
        for (int v = 0; v < 8; v++) qn[v] = QV(slot[v], neigh);
        const double xn = m.xc[neigh], yn = m.yc[neigh], zn = m.zc[neigh];
        double vq[(Q_N + 3) / 4];
        quad_load<Q_N>(q, nt, i, f, vq);
        double fv[8];
        fv[0] = (row_plane<Q_UG>(vq) + qn[0]) * 0.5;
        fv[1] = (row_plane<Q_VG>(vq) + qn[1]) * 0.5;
        fv[2] = (row_plane<Q_WG>(vq) + qn[2]) * 0.5;
        fv[3] = (row_plane<Q_TG>(vq) + qn[3]) * 0.5;
        fv[4] = (row_plane<Q_UP>(vq) + qn[4]) * 0.5;
        fv[5] = (row_plane<Q_VP>(vq) + qn[5]) * 0.5;
        fv[6] = (row_plane<Q_WP>(vq) + qn[6]) * 0.5;
        fv[7] = (row_plane<Q_TP>(vq) + qn[7]) * 0.5;
        const double phigf = (row_plane<Q_PHIG>(vq) + QV(Q_PHIG, neigh)) * 0.5;
        const double phipf = (row_plane<Q_PHIP>(vq) + QV(Q_PHIP, neigh)) * 0.5;
        const bool ghost = neigh >= ne;
        const int far = ghost ? i : neigh;
        int m2[4];
        for (int k = 0; k < 4; k++) m2[k] = ghost ? i : m.nc[(size_t)k * ne + far];
        viscous_stencil(P, q, nt, m, m2, ghost, xn, yn, zn, qn, xf, yf, zf, scx, scy, scz, fv, phigf, phipf, acc);
    }
    for (int v = 0; v < 8; v++) {
        const double t0 = __shfl(acc[v], qbase, 64), t1 = __shfl(acc[v], qbase + 16, 64),
                     t2 = __shfl(acc[v], qbase + 32, 64), t3 = __shfl(acc[v], qbase + 48, 64);
        if ((v & 3) == f) s.vis[(size_t)v * ne + i] = 0.0 + t0 + t1 + t2 + t3;
    }
}

// ---------------------------------------------------------------------------
// cell gradients (sr_derivative.F90:45-322, inverse-squared-distance LSQ over
// the 4 face neighbours, Cramer's rule) fused with the Venkatakrishnan
// limiter (sr_VK.F90:32-274), which only needs this cell's gradients.
// ---------------------------------------------------------------------------
// Interphase source of cell i (sr_source.F90:46-155) from the stage's
// primitives: drag, interface pressure times the gradient of phig (an LSQ fit
// over the four face neighbours with 1/distance weights) and heat exchange.
// Stored as the five terms times the cell volume that the stage update
// subtracts (gas) or adds (liquid), each formed as the reference forms it.
__device__ __forceinline__ void source_cell(const Params &P, const State &s, const Mesh &m, int i)
{
    const int ne = m.ne;
    if (m.iblank[i] != 1) return;
    const double *q = s.q;
    const size_t nt = s.nt;
    int nb[4];
    for (int k = 0; k < 4; k++) nb[k] = m.nc[(size_t)k * ne + i];
    const double xi = m.xc[i], yi = m.yc[i], zi = m.zc[i];
    double dx[4], dy[4], dz[4], wt[4];
    for (int k = 0; k < 4; k++) {
        dx[k] = m.xc[nb[k]] - xi;
        dy[k] = m.yc[nb[k]] - yi;
        dz[k] = m.zc[nb[k]] - zi;
        wt[k] = dx[k] * dx[k] + dy[k] * dy[k] + dz[k] * dz[k];
    }
    const double phig = QV(Q_PHIG, i), phip = QV(Q_PHIP, i), rog = QV(Q_ROG, i), rop = QV(Q_ROP, i);
    const double ug = QV(Q_UG, i), vg = QV(Q_VG, i), wg = QV(Q_WG, i), up = QV(Q_UP, i), vp = QV(Q_VP, i),
                 wp = QV(Q_WP, i), tg = QV(Q_TG, i), tp = QV(Q_TP, i), pg = QV(Q_PG, i);
    const double voli = m.vol[i], dl = m.dl[i];
    const double amug = P.as * sqrt(tg) / (1.0 + P.ts / tg);
    const double du = ug - up, dv = vg - vp, dw = wg - wp;
    const double drag = 0.44 * phig * (phig - 1.0) * rog * sqrt(du * du + dv * dv + dw * dw) * dl;
    const double fdgx = drag * du, fdgy = drag * dv, fdgz = drag * dw;
    const double qsource = 6.0 * phip * amug * P.cpg * 2.0 * (tp - tg) / (P.prg * (dl * dl));
    const double delpstar = 2.0 * (phig * rog * phip * rop) / (phig * rop + phip * rog) * (du * du + dv * dv + dw * dw);
    const double pint = pg - dmin(delpstar, 0.01 * pg);
    const double den = phig * rog + (1.0 - phig) * rop;
    const Divisor DN = divisor(den);
    const double uint = divide(phig * rog * ug + (1.0 - phig) * rop * up, DN);
    const double vint = divide(phig * rog * vg + (1.0 - phig) * rop * vp, DN);
    const double wint = divide(phig * rog * wg + (1.0 - phig) * rop * wp, DN);
    // face order k = 0..3, summed left to right
    Divisor W[4];
    double dph[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        W[k] = divisor(sqrt(wt[k]));
        dph[k] = QV(Q_PHIG, nb[k]) - phig;
    }
#define SUM4(ex) (divide(ex(0), W[0]) + divide(ex(1), W[1]) + divide(ex(2), W[2]) + divide(ex(3), W[3]))
#define XX(k) (dx[k] * dx[k])
#define YY(k) (dy[k] * dy[k])
#define ZZ(k) (dz[k] * dz[k])
#define XY(k) (dx[k] * dy[k])
#define XZ(k) (dx[k] * dz[k])
#define YZ(k) (dy[k] * dz[k])
#define PX(k) (dph[k] * dx[k])
#define PY(k) (dph[k] * dy[k])
#define PZ(k) (dph[k] * dz[k])
    const double a1 = SUM4(XX), a2 = SUM4(YY), a3 = SUM4(ZZ), b1 = SUM4(XY), b2 = SUM4(XZ), b3 = SUM4(YZ);
    const double c1 = SUM4(PX), c2 = SUM4(PY), c3 = SUM4(PZ);
#undef SUM4
#undef XX
#undef YY
#undef ZZ
#undef XY
#undef XZ
#undef YZ
#undef PX
#undef PY
#undef PZ
    const double dnr = det3(a1, b1, b2, b1, a2, b3, b2, b3, a3);
    const Divisor D = divisor(dnr);
    const double dphx = divide(det3(c1, c2, c3, b1, a2, b3, b2, b3, a3), D);
    const double dphy = divide(det3(a1, b1, b2, c1, c2, c3, b2, b3, a3), D);
    const double dphz = divide(det3(a1, b1, b2, b1, a2, b3, c1, c2, c3), D);
    const double work = pint * (uint * dphx + vint * dphy + wint * dphz) + uint * fdgx + vint * fdgy +
                        wint * fdgz + qsource;
    double *o = s.src;
    o[(size_t)SRC_MX * ne + i] = (fdgx + pint * dphx + phig * rog * P.gravity) * voli;
    o[(size_t)SRC_MY * ne + i] = (fdgy + pint * dphy) * voli;
    o[(size_t)SRC_MZ * ne + i] = (fdgz + pint * dphz) * voli;
    o[(size_t)SRC_E * ne + i] = work * voli;
    o[(size_t)SRC_MXP * ne + i] = (fdgx + pint * dphx - phip * rop * P.gravity) * voli;
}

__device__ __forceinline__ void grad_cell(const State &s, const Mesh &m, int i)
{
    const int ne = m.ne;
    if (m.iblank[i] != 1) return;
    const double *q = s.q;
    const size_t nt = s.nt;
    int nb[4];
    for (int k = 0; k < 4; k++) nb[k] = m.nc[(size_t)k * ne + i];
    const double xi = m.xc[i], yi = m.yc[i], zi = m.zc[i];
    double dx[4], dy[4], dz[4], wt[4];
    for (int k = 0; k < 4; k++) {
        dx[k] = m.xc[nb[k]] - xi;
        dy[k] = m.yc[nb[k]] - yi;
        dz[k] = m.zc[nb[k]] - zi;
        wt[k] = dx[k] * dx[k] + dy[k] * dy[k] + dz[k] * dz[k];
    }
    Divisor W[4];
    for (int k = 0; k < 4; k++) W[k] = divisor(wt[k]);
    const double a1 = divide(dx[0] * dx[0], W[0]) + divide(dx[1] * dx[1], W[1]) + divide(dx[2] * dx[2], W[2]) + divide(dx[3] * dx[3], W[3]);
    const double a2 = divide(dy[0] * dy[0], W[0]) + divide(dy[1] * dy[1], W[1]) + divide(dy[2] * dy[2], W[2]) + divide(dy[3] * dy[3], W[3]);
    const double a3 = divide(dz[0] * dz[0], W[0]) + divide(dz[1] * dz[1], W[1]) + divide(dz[2] * dz[2], W[2]) + divide(dz[3] * dz[3], W[3]);
    const double b1 = divide(dx[0] * dy[0], W[0]) + divide(dx[1] * dy[1], W[1]) + divide(dx[2] * dy[2], W[2]) + divide(dx[3] * dy[3], W[3]);
    const double b2 = divide(dx[0] * dz[0], W[0]) + divide(dx[1] * dz[1], W[1]) + divide(dx[2] * dz[2], W[2]) + divide(dx[3] * dz[3], W[3]);
    const double b3 = divide(dz[0] * dy[0], W[0]) + divide(dz[1] * dy[1], W[1]) + divide(dz[2] * dy[2], W[2]) + divide(dz[3] * dy[3], W[3]);
    const double dnr = det3(a1, b1, b2, b1, a2, b3, b2, b3, a3);
    const Divisor D = divisor(dnr);
    // 13 scalars; the first 11 feed the face reconstruction and are stored,
    // rog/rop (and pg/pp) feed the limiter below
    const int slot[13] = {Q_UG, Q_VG, Q_WG, Q_TG, Q_UP, Q_VP, Q_WP, Q_TP, Q_PG, Q_PP, Q_PHIG, Q_ROG, Q_ROP};
    double glim[4][3];  // pg, pp, rog, rop
    double dpl[4], dml[4];  // their neighbour maxima and minima less the cell's value
#pragma unroll
    for (int sI = 0; sI < 13; sI++) {
        const int sl = slot[sI];
        const double qi = QV(sl, i);
        double qn[4], dq[4];
        for (int k = 0; k < 4; k++) {
            qn[k] = QV(sl, nb[k]);
            dq[k] = qn[k] - qi;
        }
        const double c1 = divide(dq[0] * dx[0], W[0]) + divide(dq[1] * dx[1], W[1]) + divide(dq[2] * dx[2], W[2]) + divide(dq[3] * dx[3], W[3]);
        double c2;
        if (sl == Q_TP)  // sr_derivative.F90:225, last term without /wt4
            c2 = divide(dq[0] * dy[0], W[0]) + divide(dq[1] * dy[1], W[1]) + divide(dq[2] * dy[2], W[2]) + dq[3] * dy[3];
        else
            c2 = divide(dq[0] * dy[0], W[0]) + divide(dq[1] * dy[1], W[1]) + divide(dq[2] * dy[2], W[2]) + divide(dq[3] * dy[3], W[3]);
        const double c3 = divide(dq[0] * dz[0], W[0]) + divide(dq[1] * dz[1], W[1]) + divide(dq[2] * dz[2], W[2]) + divide(dq[3] * dz[3], W[3]);
        const double gx = divide(det3(c1, c2, c3, b1, a2, b3, b2, b3, a3), D);
        const double gy = divide(det3(a1, b1, b2, c1, c2, c3, b2, b3, a3), D);
        const double gz = divide(det3(a1, b1, b2, b1, a2, b3, c1, c2, c3), D);
        if (sI < 11) {
            __builtin_nontemporal_store(gx, &s.grad[(size_t)(3 * sI + 0) * ne + i]);
            __builtin_nontemporal_store(gy, &s.grad[(size_t)(3 * sI + 1) * ne + i]);
            __builtin_nontemporal_store(gz, &s.grad[(size_t)(3 * sI + 2) * ne + i]);
        }
        const int li = (sl == Q_PG) ? 0 : (sl == Q_PP) ? 1 : (sl == Q_ROG) ? 2 : (sl == Q_ROP) ? 3 : -1;
        if (li >= 0) {
            glim[li][0] = gx; glim[li][1] = gy; glim[li][2] = gz;
            // the limiter's extrema (sr_VK.F90:89-90, dmax1/dmin1 over the
            // cell and its neighbours in that order) from the values just read
            double vmax = qi, vmin = qi;
            for (int k = 0; k < 4; k++) {
                vmax = dmax(vmax, qn[k]);
                vmin = dmin(vmin, qn[k]);
            }
            dpl[li] = vmax - qi;
            dml[li] = vmin - qi;
        }
    }
    // limiter (sr_VK.F90:34-272)
    const double epssq = m.epssq[i];
    double r0[4][3];
    for (int f = 0; f < 4; f++) {
        r0[f][0] = m.fc0[(size_t)(3 * f) * ne + i] - xi;
        r0[f][1] = m.fc0[(size_t)(3 * f + 1) * ne + i] - yi;
        r0[f][2] = m.fc0[(size_t)(3 * f + 2) * ne + i] - zi;
    }
    double ph = 0.0;
    for (int v = 0; v < 4; v++) {
        for (int f = 0; f < 4; f++) {
            const double delf = r0[f][0] * glim[v][0] + r0[f][1] * glim[v][1] + r0[f][2] * glim[v][2];
            double pv;
            if (delf > 0.0)
                pv = fvenkat(dpl[v], delf, epssq);
            else if (delf < 0.0)
                pv = fvenkat(dml[v], delf, epssq);
            else
                pv = 1.0;
            ph = (v == 0 && f == 0) ? pv : dmin(ph, pv);
        }
    }
    __builtin_nontemporal_store(ph, &s.phi[i]);
}

// ---------------------------------------------------------------------------
// face loop + interphase source + optional RK stage update
// ---------------------------------------------------------------------------
struct FState {
    double rog, ug, vg, wg, tg, phig, rop, up, vp, wp, tp, pg, pp, phip;
};

__device__ __forceinline__ void load_fstate(const double *q, size_t nt, int i, FState &s)
{
    s.rog = QV(Q_ROG, i); s.ug = QV(Q_UG, i); s.vg = QV(Q_VG, i); s.wg = QV(Q_WG, i);
    s.tg = QV(Q_TG, i); s.phig = QV(Q_PHIG, i); s.rop = QV(Q_ROP, i); s.up = QV(Q_UP, i);
    s.vp = QV(Q_VP, i); s.wp = QV(Q_WP, i); s.tp = QV(Q_TP, i); s.pg = QV(Q_PG, i); s.pp = QV(Q_PP, i);
}

// sr_hllcbatten.F90:228-318
__device__ __forceinline__ void clip_blend(const Params &P, FState &s)
{
    if (s.phig < P.epslnmin) {
        s.phig = P.epslnmin; s.phip = 1.0 - s.phig;
        s.ug = s.up; s.vg = s.vp; s.wg = s.wp; s.tg = s.tp; s.pg = s.pp;
    }
    if (s.phip < P.epslnmin) {
        s.phip = P.epslnmin; s.phig = 1.0 - s.phip;
        s.up = s.ug; s.vp = s.vg; s.wp = s.wg; s.tp = s.tg; s.pp = s.pg;
    }
    if (s.phig >= P.epslnmin && s.phig <= P.epslnmax) {
        const double psig = (s.phig - P.epslnmin) / (P.epslnmax - P.epslnmin);
        const double fg = -psig * psig * (2.0 * psig - 3.0);
        s.ug = fg * s.ug + (1.0 - fg) * s.up;
        s.vg = fg * s.vg + (1.0 - fg) * s.vp;
        s.wg = fg * s.wg + (1.0 - fg) * s.wp;
        s.tg = fg * s.tg + (1.0 - fg) * s.tp;
        s.pg = fg * s.pg + (1.0 - fg) * s.pp;
    }
    if (s.phip >= P.epslnmin && s.phip <= P.epslnmax) {
        const double psig = (s.phip - P.epslnmin) / (P.epslnmax - P.epslnmin);
        const double fg = -psig * psig * (2.0 * psig - 3.0);
        s.up = fg * s.up + (1.0 - fg) * s.ug;
        s.vp = fg * s.vp + (1.0 - fg) * s.vg;
        s.wp = fg * s.wp + (1.0 - fg) * s.wg;
        s.tp = fg * s.tp + (1.0 - fg) * s.tg;
        s.pp = fg * s.pp + (1.0 - fg) * s.pg;
    }
    s.rog = s.pg / (P.rrg * s.tg);
    s.rop = P.gamap * (s.pp + P.pinf) / ((P.gamap - 1.0) * P.cpp * s.tp);
}

// One phase of the HLLC-Batten flux (sr_hllcbatten.F90:338-529): each phase
// has its own wave speeds and contact.  g = gas (energy from R T/(gamma-1)),
// otherwise liquid (stiffened gas: cp T/gamma + p_inf/rho).
template <bool GAS>
__device__ __forceinline__ void hllc_phase(const Params &P, double phL, double roL, double uL, double vL,
                                           double wL, double tL, double pL, double phR, double roR, double uR,
                                           double vR, double wR, double tR, double pR, double enx, double eny,
                                           double enz, double area, double *F)
{
    double EL, ER, aL, aR;
    if (GAS) {
        EL = P.rrg * tL / (P.gamag - 1.0) + 0.5 * (uL * uL + vL * vL + wL * wL);
        ER = P.rrg * tR / (P.gamag - 1.0) + 0.5 * (uR * uR + vR * vR + wR * wR);
        aL = sqrt(P.gamag * pL / roL);
        aR = sqrt(P.gamag * pR / roR);
    } else {
        EL = (P.cpp * tL) / P.gamap + P.pinf / roL + 0.5 * (uL * uL + vL * vL + wL * wL);
        ER = (P.cpp * tR) / P.gamap + P.pinf / roR + 0.5 * (uR * uR + vR * vR + wR * wR);
        aL = sqrt(P.gamap * (pL + P.pinf) / roL);
        aR = sqrt(P.gamap * (pR + P.pinf) / roR);
    }
    const double HL = EL + pL / roL, HR = ER + pR / roR;
    const double qL = uL * enx + vL * eny + wL * enz;
    const double qR = uR * enx + vR * eny + wR * enz;
    const double SL = dmin(dmin(qL - aL, qR - aR), F1E8);
    const double SR = dmax(dmax(qL + aL, qR + aR), F1E8);
    const double snr = pR - pL + roL * qL * (SL - qL) - roR * qR * (SR - qR);
    const double sdr = roL * (SL - qL) - roR * (SR - qR);
    const double SS = snr / (sdr + F1E8);
    const double pst = dmax(pL + roL * (qL - SL) * (qL - SS), F1E6 * dmin(pL, pR));
    const double mL = phL * roL, mR = phR * roR;
    if (SL > 0.0) {
        F[0] = mL * qL * area;
        F[1] = (mL * qL * uL + phL * pL * enx) * area;
        F[2] = (mL * qL * vL + phL * pL * eny) * area;
        F[3] = (mL * qL * wL + phL * pL * enz) * area;
        F[4] = mL * qL * HL * area;
    } else if (SL <= 0.0 && SS > 0.0) {
        const double dn = (SS - SL) + F1E6;
        const double c1 = mL, c2 = mL * uL, c3 = mL * vL, c4 = mL * wL, c5 = mL * EL;
        const double s1 = (qL - SL) * c1 / dn;
        const double s2 = ((qL - SL) * c2 + phL * (pL * enx - pst * enx)) / dn;
        const double s3 = ((qL - SL) * c3 + phL * (pL * eny - pst * eny)) / dn;
        const double s4 = ((qL - SL) * c4 + phL * (pL * enz - pst * enz)) / dn;
        const double s5 = ((qL - SL) * c5 + phL * (pL * qL - pst * SS)) / dn;
        F[0] = (mL * qL + SL * (s1 - c1)) * area;
        F[1] = (mL * qL * uL + phL * pL * enx + SL * (s2 - c2)) * area;
        F[2] = (mL * qL * vL + phL * pL * eny + SL * (s3 - c3)) * area;
        F[3] = (mL * qL * wL + phL * pL * enz + SL * (s4 - c4)) * area;
        F[4] = (mL * qL * HL + SL * (s5 - c5)) * area;
    } else if (SS <= 0.0 && SR >= 0.0) {
        const double dn = (SS - SR) + F1E6;
        const double c1 = mR, c2 = mR * uR, c3 = mR * vR, c4 = mR * wR, c5 = mR * ER;
        const double s1 = (qR - SR) * c1 / dn;
        const double s2 = ((qR - SR) * c2 + phR * (pR * enx - pst * enx)) / dn;
        const double s3 = ((qR - SR) * c3 + phR * (pR * eny - pst * eny)) / dn;
        const double s4 = ((qR - SR) * c4 + phR * (pR * enz - pst * enz)) / dn;
        const double s5 = ((qR - SR) * c5 + phR * (pR * qR - pst * SS)) / dn;
        F[0] = (mR * qR + SR * (s1 - c1)) * area;
        F[1] = (mR * qR * uR + phR * pR * enx + SR * (s2 - c2)) * area;
        F[2] = (mR * qR * vR + phR * pR * eny + SR * (s3 - c3)) * area;
        F[3] = (mR * qR * wR + phR * pR * enz + SR * (s4 - c4)) * area;
        F[4] = (mR * qR * HR + SR * (s5 - c5)) * area;
    } else if (SR < 0.0) {
        F[0] = mR * qR * area;
        F[1] = (mR * qR * uR + phR * pR * enx) * area;
        F[2] = (mR * qR * vR + phR * pR * eny) * area;
        F[3] = (mR * qR * wR + phR * pR * enz) * area;
        F[4] = mR * qR * HR * area;
    }
}

// HLLC-Zien flux (sr_hllczien.F90:327-487): mixture wave speeds and one
// contact shared by both phases.
__device__ __forceinline__ void hllc_zien(const Params &P, const FState &L, const FState &R, double enx,
                                          double eny, double enz, double area, double *F)
{
    const double EGL = P.rrg * L.tg / (P.gamag - 1.0) + 0.5 * (L.ug * L.ug + L.vg * L.vg + L.wg * L.wg);
    const double HGL = EGL + L.pg / L.rog;
    const double QGL = L.ug * enx + L.vg * eny + L.wg * enz;
    const double AGL = sqrt(P.gamag * L.pg / L.rog);
    const double EGR = P.rrg * R.tg / (P.gamag - 1.0) + 0.5 * (R.ug * R.ug + R.vg * R.vg + R.wg * R.wg);
    const double HGR = EGR + R.pg / R.rog;
    const double QGR = R.ug * enx + R.vg * eny + R.wg * enz;
    const double AGR = sqrt(P.gamag * R.pg / R.rog);
    const double EPL = (P.cpp * L.tp) / P.gamap + P.pinf / L.rop + 0.5 * (L.up * L.up + L.vp * L.vp + L.wp * L.wp);
    const double HPL = EPL + L.pp / L.rop;
    const double QPL = L.up * enx + L.vp * eny + L.wp * enz;
    const double APL = sqrt(P.gamap * (L.pp + P.pinf) / L.rop);
    const double EPR = (P.cpp * R.tp) / P.gamap + P.pinf / R.rop + 0.5 * (R.up * R.up + R.vp * R.vp + R.wp * R.wp);
    const double HPR = EPR + R.pp / R.rop;
    const double QPR = R.up * enx + R.vp * eny + R.wp * enz;
    const double APR = sqrt(P.gamap * (R.pp + P.pinf) / R.rop);
    const double SL = dmin(dmin(dmin(dmin(QGL - AGL, QPL - APL), QGR - AGR), QPR - APR), F1E8);
    const double SR = dmax(dmax(dmax(dmax(QGL + AGL, QPL + APL), QGR + AGR), QPR + APR), F1E8);
    const double ROL = L.phig * L.rog + L.phip * L.rop;
    const double ROQL = L.phig * L.rog * QGL + L.phip * L.rop * QPL;
    const double QL = ROQL / ROL;
    const double PRL = L.phig * L.pg + L.phip * L.pp;
    const double ROR = R.phig * R.rog + R.phip * R.rop;
    const double ROQR = R.phig * R.rog * QGR + R.phip * R.rop * QPR;
    const double QR = ROQR / ROR;
    const double PRR = R.phig * R.pg + R.phip * R.pp;
    const double SSTAR = (PRR - PRL + ROQL * (SL - QL) - ROQR * (SR - QR)) / (ROL * (SL - QL) - ROR * (SR - QR) + F1E8);
    const double PGSTAR = dmax(L.pg + L.rog * (QGL - SL) * (QGL - SSTAR), F1E6 * dmin(L.pg, R.pg));
    const double PPSTAR = dmax(L.pp + L.rop * (QPL - SL) * (QPL - SSTAR), F1E6 * dmin(L.pp, R.pp));
    const double mgL = L.phig * L.rog, mpL = L.phip * L.rop, mgR = R.phig * R.rog, mpR = R.phip * R.rop;
    if (SL >= 0.0) {
        F[0] = mgL * QGL * area;
        F[1] = (mgL * QGL * L.ug + L.phig * L.pg * enx) * area;
        F[2] = (mgL * QGL * L.vg + L.phig * L.pg * eny) * area;
        F[3] = (mgL * QGL * L.wg + L.phig * L.pg * enz) * area;
        F[4] = mgL * QGL * HGL * area;
        F[5] = mpL * QPL * area;
        F[6] = (mpL * QPL * L.up + L.phip * L.pp * enx) * area;
        F[7] = (mpL * QPL * L.vp + L.phip * L.pp * eny) * area;
        F[8] = (mpL * QPL * L.wp + L.phip * L.pp * enz) * area;
        F[9] = mpL * QPL * HPL * area;
    } else if (SL <= 0.0 && SSTAR > 0.0) {
        const double dn = (SSTAR - SL) + F1E6;
        const double cg[5] = {mgL, mgL * L.ug, mgL * L.vg, mgL * L.wg, mgL * EGL};
        const double cp[5] = {mpL, mpL * L.up, mpL * L.vp, mpL * L.wp, mpL * EPL};
        const double sg[5] = {(QGL - SL) * cg[0] / dn,
                              ((QGL - SL) * cg[1] + L.phig * (L.pg * enx - PGSTAR * enx)) / dn,
                              ((QGL - SL) * cg[2] + L.phig * (L.pg * eny - PGSTAR * eny)) / dn,
                              ((QGL - SL) * cg[3] + L.phig * (L.pg * enz - PGSTAR * enz)) / dn,
                              ((QGL - SL) * cg[4] + L.phig * (L.pg * QGL - PGSTAR * SSTAR)) / dn};
        const double sp[5] = {(QPL - SL) * cp[0] / dn,
                              ((QPL - SL) * cp[1] + L.phip * (L.pp * enx - PPSTAR * enx)) / dn,
                              ((QPL - SL) * cp[2] + L.phip * (L.pp * eny - PPSTAR * eny)) / dn,
                              ((QPL - SL) * cp[3] + L.phip * (L.pp * enz - PPSTAR * enz)) / dn,
                              ((QPL - SL) * cp[4] + L.phip * (L.pp * QPL - PPSTAR * SSTAR)) / dn};
        F[0] = (mgL * QGL + SL * (sg[0] - cg[0])) * area;
        F[1] = (mgL * QGL * L.ug + L.phig * L.pg * enx + SL * (sg[1] - cg[1])) * area;
        F[2] = (mgL * QGL * L.vg + L.phig * L.pg * eny + SL * (sg[2] - cg[2])) * area;
        F[3] = (mgL * QGL * L.wg + L.phig * L.pg * enz + SL * (sg[3] - cg[3])) * area;
        F[4] = (mgL * QGL * HGL + SL * (sg[4] - cg[4])) * area;
        F[5] = (mpL * QPL + SL * (sp[0] - cp[0])) * area;
        F[6] = (mpL * QPL * L.up + L.phip * L.pp * enx + SL * (sp[1] - cp[1])) * area;
        F[7] = (mpL * QPL * L.vp + L.phip * L.pp * eny + SL * (sp[2] - cp[2])) * area;
        F[8] = (mpL * QPL * L.wp + L.phip * L.pp * enz + SL * (sp[3] - cp[3])) * area;
        F[9] = (mpL * QPL * HPL + SL * (sp[4] - cp[4])) * area;
    } else if (SSTAR <= 0.0 && SR >= 0.0) {
        const double dn = (SSTAR - SR) + F1E6;
        const double cg[5] = {mgR, mgR * R.ug, mgR * R.vg, mgR * R.wg, mgR * EGR};
        const double cp[5] = {mpR, mpR * R.up, mpR * R.vp, mpR * R.wp, mpR * EPR};
        const double sg[5] = {(QGR - SR) * cg[0] / dn,
                              ((QGR - SR) * cg[1] + R.phig * (R.pg * enx - PGSTAR * enx)) / dn,
                              ((QGR - SR) * cg[2] + R.phig * (R.pg * eny - PGSTAR * eny)) / dn,
                              ((QGR - SR) * cg[3] + R.phig * (R.pg * enz - PGSTAR * enz)) / dn,
                              ((QGR - SR) * cg[4] + R.phig * (R.pg * QGR - PGSTAR * SSTAR)) / dn};
        const double sp[5] = {(QPR - SR) * cp[0] / dn,
                              ((QPR - SR) * cp[1] + R.phip * (R.pp * enx - PPSTAR * enx)) / dn,
                              ((QPR - SR) * cp[2] + R.phip * (R.pp * eny - PPSTAR * eny)) / dn,
                              ((QPR - SR) * cp[3] + R.phip * (R.pp * enz - PPSTAR * enz)) / dn,
                              ((QPR - SR) * cp[4] + R.phip * (R.pp * QPR - PPSTAR * SSTAR)) / dn};
        F[0] = (mgR * QGR + SR * (sg[0] - cg[0])) * area;
        F[1] = (mgR * QGR * R.ug + R.phig * R.pg * enx + SR * (sg[1] - cg[1])) * area;
        F[2] = (mgR * QGR * R.vg + R.phig * R.pg * eny + SR * (sg[2] - cg[2])) * area;
        F[3] = (mgR * QGR * R.wg + R.phig * R.pg * enz + SR * (sg[3] - cg[3])) * area;
        F[4] = (mgR * QGR * HGR + SR * (sg[4] - cg[4])) * area;
        F[5] = (mpR * QPR + SR * (sp[0] - cp[0])) * area;
        F[6] = (mpR * QPR * R.up + R.phip * R.pp * enx + SR * (sp[1] - cp[1])) * area;
        F[7] = (mpR * QPR * R.vp + R.phip * R.pp * eny + SR * (sp[2] - cp[2])) * area;
        F[8] = (mpR * QPR * R.wp + R.phip * R.pp * enz + SR * (sp[3] - cp[3])) * area;
        F[9] = (mpR * QPR * HPR + SR * (sp[4] - cp[4])) * area;
    } else if (SR <= 0.0) {
        F[0] = mgR * QGR * area;
        F[1] = (mgR * QGR * R.ug + R.phig * R.pg * enx) * area;
        F[2] = (mgR * QGR * R.vg + R.phig * R.pg * eny) * area;
        F[3] = (mgR * QGR * R.wg + R.phig * R.pg * enz) * area;
        F[4] = mgR * QGR * HGR * area;
        F[5] = mpR * QPR * area;
        F[6] = (mpR * QPR * R.up + R.phip * R.pp * enx) * area;
        F[7] = (mpR * QPR * R.vp + R.phip * R.pp * eny) * area;
        F[8] = (mpR * QPR * R.wp + R.phip * R.pp * enz) * area;
        F[9] = mpR * QPR * HPR * area;
    }
}

#define GRAD(s, d, c) s_grad[(size_t)(3 * (s) + (d)) * ne + (c)]

// the residual r of element o (variable-major offset): stored in rhs
// (stage 0, per-routine checks) or applied as RK stage `stage`
// (sr_RK.F90:29-161), whose new value is returned and, with `store`,
// written to cn; FP64 in every precision
template <typename TD>
__device__ __forceinline__ double rk_finish(const State &s, const TD *d24, size_t o, double r, int stage,
                                            double dtbyvol, bool store)
{
    if (stage == 0) {
        s.rhs[o] = r;
        return 0.0;
    }
    // the stage's state is streamed through once: non-temporal loads and
    // stores keep it out of the caches the gathers use
    const double rr = r - (double)__builtin_nontemporal_load(&d24[o]);
    const double cu = __builtin_nontemporal_load(&s.cu[o]);
    double cn = __builtin_nontemporal_load(&s.cn[o]);
    if (stage == 1) {
        cn = cu - (11.0 / 20.0) * dtbyvol * rr;
        __builtin_nontemporal_store(cn, &s.cn1[o]);
    } else if (stage == 2) {
        cn = (3.0 / 8.0) * cu + (5.0 / 8.0) * (cn - (11.0 / 20.0) * dtbyvol * rr);
    } else if (stage == 3) {
        cn = (4.0 / 9.0) * cu + (5.0 / 9.0) * (cn - (11.0 / 20.0) * dtbyvol * rr);
    } else {
        cn = (111.0 / 1331.0) * cu + (260.0 / 1331.0) * __builtin_nontemporal_load(&s.cn1[o]) +
             (960.0 / 1331.0) * (cn - (11.0 / 20.0) * dtbyvol * rr);
    }
    if (store) s.cn[o] = cn;
    return cn;
}

// RK stage update (sr_RK.F90:29-161) of the quad's cell i from its
// face-summed residual and the stored interphase source terms (source_cell;
// FP64 form).  Lane f of the quad finishes the conserved variables k = f,
// f+4, f+8, whose face sums are rhs[0], rhs[1], rhs[2].  stage >= 1: the
// new values go to cv[0..2] and, with `store`, to cn (cn1 at stage 1);
// stage == 0: store the residual in rhs (per-routine checks).
__device__ __forceinline__ void source_and_update(const State &s, const Mesh &m, int i, int f,
                                                  const double *rhs, int stage, bool store, double (&cv)[3])
{
    const int ne = m.ne;
    const double *vis = s.vis, *src = s.src;
    const double dtbyvol = (stage > 0) ? s.scal[SC_DT] / m.vol[i] : 0.0;
    for (int j = 0; j < 3; j++) {
        const int k = f + 4 * j;
        if (k >= 10) break;
        double r = rhs[j];
        switch (k) {
        case 1: r = r - __builtin_nontemporal_load(&src[(size_t)SRC_MX * ne + i]) - __builtin_nontemporal_load(&vis[(size_t)0 * ne + i]); break;
        case 2: r = r - __builtin_nontemporal_load(&src[(size_t)SRC_MY * ne + i]) - __builtin_nontemporal_load(&vis[(size_t)1 * ne + i]); break;
        case 3: r = r - __builtin_nontemporal_load(&src[(size_t)SRC_MZ * ne + i]) - __builtin_nontemporal_load(&vis[(size_t)2 * ne + i]); break;
        case 4: r = r - __builtin_nontemporal_load(&src[(size_t)SRC_E * ne + i]) - __builtin_nontemporal_load(&vis[(size_t)3 * ne + i]); break;
        case 6: r = r + __builtin_nontemporal_load(&src[(size_t)SRC_MXP * ne + i]) - __builtin_nontemporal_load(&vis[(size_t)4 * ne + i]); break;
        case 7: r = r + __builtin_nontemporal_load(&src[(size_t)SRC_MY * ne + i]) - __builtin_nontemporal_load(&vis[(size_t)5 * ne + i]); break;
        case 8: r = r + __builtin_nontemporal_load(&src[(size_t)SRC_MZ * ne + i]) - __builtin_nontemporal_load(&vis[(size_t)6 * ne + i]); break;
        case 9: r = r + __builtin_nontemporal_load(&src[(size_t)SRC_E * ne + i]) - __builtin_nontemporal_load(&vis[(size_t)7 * ne + i]); break;
        default: break;
        }
        cv[j] = rk_finish(s, s.d24, (size_t)k * ne + i, r, stage, dtbyvol, store);
    }
}

// Fused step: the primitive recovery of cell i right after its stage
// update, i.e. the EOS pass the next stage (or the end of the step) would
// run on it.  The quads of a tile leave their cells' new conserved
// variables in LDS (c: variable k of the thread's cell at c[k * TILE]); one
// thread per cell recovers the primitives into the other primitive buffer
// q2 (the next stage's q) and stores the rewritten conserved variables in
// cn, and in cu as well at the end of the step (`end`).  An overset donor
// also keeps its stage update, unrewritten, in cnd: that is what the donor
// gather after this launch interpolates.
template <int STRIDE = TILE>
__device__ __forceinline__ void eos_staged(const Params &P, const State &s, int i, const double *c_lds, bool end,
                                           double (&d)[4])
{
    double c[10], v[Q_N], nv[10];
    for (int k = 0; k < 10; k++) c[k] = c_lds[k * STRIDE];
    eos_prims(P, c, v, nv);
    const int ne = s.ne;
    const size_t nt = s.nt;
    // streamed stores (non-temporal), as the stage update's
    for (int k = 0; k < Q_N; k++) __builtin_nontemporal_store(v[k], &s.q2[(size_t)k * nt + i]);
    const unsigned char fl = s.dflag ? s.dflag[i] : 0;
    if (fl & 1)
        for (int k = 0; k < 10; k++) s.cnd[(size_t)k * ne + i] = c[k];
    for (int k = 0; k < 10; k++) __builtin_nontemporal_store(nv[k], &s.cn[(size_t)k * ne + i]);
    if (end && !(fl & 2)) {
        // the monitor (main.F90:603-612) against the step's starting cu, and
        // cu = cn; an overset receptor's final value, its monitor difference
        // and its cu come from the gather that follows, which reads its cu
        for (int k = 0; k < 4; k++) d[k] = fabs(nv[k] - s.cu[(size_t)k * ne + i]);
        for (int k = 0; k < 10; k++) __builtin_nontemporal_store(nv[k], &s.cu[(size_t)k * ne + i]);
    }
}

// One cell per column of lanes, one face per lane (sr_hllcbatten.F90:77-555):
// the wave's 16 cells c hold their four faces in lanes c, c+16, c+32, c+48
// (rows of 16 lanes), so each load instruction takes one face of 16
// consecutive cells.  Each lane reconstructs, clips and solves the Riemann
// problem of its own face; the column then sums the four face fluxes in face
// order (row_spread), so every residual is the reference's
// (((0+F1)+F2)+F3)+F4 without atomics, and the lane of face f finishes the
// conserved variables k = f, f+4, f+8.  Cell i's primitives and gradients
// are loaded once per column (row f loads planes f, f+4, ...) and handed to
// the four rows by row swaps.
template <int FLUX>
__device__ __forceinline__ void flux_update(const Params &P, const State &s, const Mesh &m, int i, int f,
                                            int stage, bool store, double (&cv)[3])
{
    const int ne = m.ne;
    const double *q = s.q;
    const size_t nt = s.nt;
    const double *s_grad = s.grad;
    const double xi = m.xc[i], yi = m.yc[i], zi = m.zc[i];
    const int neigh = m.nc[(size_t)f * ne + i];
    double F[10];
    {
        // the face's own-cell geometry is read once per stage (non-temporal)
        const double scx = __builtin_nontemporal_load(&m.sc[(size_t)(3 * f) * ne + i]), scy = __builtin_nontemporal_load(&m.sc[(size_t)(3 * f + 1) * ne + i]),
                     scz = __builtin_nontemporal_load(&m.sc[(size_t)(3 * f + 2) * ne + i]);
        const double xf = __builtin_nontemporal_load(&m.fc0[(size_t)(3 * f) * ne + i]), yf = __builtin_nontemporal_load(&m.fc0[(size_t)(3 * f + 1) * ne + i]),
                     zf = __builtin_nontemporal_load(&m.fc0[(size_t)(3 * f + 2) * ne + i]);
        const double area = sqrt(scx * scx + scy * scy + scz * scz);
        const double enx = scx / area, eny = scy / area, enz = scz / area;
        const int ntype = (neigh >= ne) ? m.gtype[neigh - ne] : 0;
        FState L, R;
        {
            // cell i's primitives: shared by the column
            double v[(Q_N + 3) / 4];
            quad_load<Q_N>(q, nt, i, f, v);
            L.rog = col_plane<Q_ROG>(v); L.ug = col_plane<Q_UG>(v); L.vg = col_plane<Q_VG>(v);
            L.wg = col_plane<Q_WG>(v); L.tg = col_plane<Q_TG>(v); L.phig = col_plane<Q_PHIG>(v);
            L.rop = col_plane<Q_ROP>(v); L.up = col_plane<Q_UP>(v); L.vp = col_plane<Q_VP>(v);
            L.wp = col_plane<Q_WP>(v); L.tp = col_plane<Q_TP>(v); L.pg = col_plane<Q_PG>(v);
            L.pp = col_plane<Q_PP>(v);
        }
        load_fstate(q, nt, neigh, R);
        if (P.myorder == 2) {
            // MUSCL reconstruction with one limiter per face pair
            // (sr_hllcbatten.F90:166-224), for faces with a cell on the far
            // side.  The column shares cell i's gradients, so its four lanes
            // stay together here: a face with a ghost on the far side forms
            // the terms from cell i's own data and discards them.
            const bool rec = neigh < ne;
            const int nb = rec ? neigh : i;
            double g[(3 * G_N + 3) / 4];
            quad_load<3 * G_N>(s_grad, (size_t)ne, i, f, g);
            const double phimin = dmin(s.phi[i], s.phi[nb]);
            const double xf0 = xf - xi, yf0 = yf - yi, zf0 = zf - zi;
            const double xfn = xf - m.xc[nb], yfn = yf - m.yc[nb], zfn = zf - m.zc[nb];
#define REC(fld, sl)                                                                                  \
    {                                                                                                 \
        const double lv = L.fld + phimin * (xf0 * col_plane<3 * (sl)>(g) +                            \
                                            yf0 * col_plane<3 * (sl) + 1>(g) +                        \
                                            zf0 * col_plane<3 * (sl) + 2>(g));                        \
        const double rv = R.fld + phimin * (xfn * GRAD(sl, 0, nb) + yfn * GRAD(sl, 1, nb) +           \
                                            zfn * GRAD(sl, 2, nb));                                   \
        if (rec) { L.fld = lv; R.fld = rv; }                                                          \
    }
            REC(phig, G_PHIG) REC(ug, G_UG) REC(vg, G_VG) REC(wg, G_WG) REC(tg, G_TG) REC(pg, G_PG)
            REC(up, G_UP) REC(vp, G_VP) REC(wp, G_WP) REC(tp, G_TP) REC(pp, G_PP)
#undef REC
        }
        L.phip = 1.0 - L.phig;
        R.phip = 1.0 - R.phig;
        clip_blend(P, L);
        clip_blend(P, R);
        if (ntype == 5) {
            // slip wall: pressure force only (sr_hllcbatten.F90:324-336); the
            // untouched components add +0.0, an identity on a sum that
            // started from +0.0
            F[0] = 0.0; F[4] = 0.0; F[5] = 0.0; F[9] = 0.0;
            F[1] = L.phig * L.pg * enx * area;
            F[2] = L.phig * L.pg * eny * area;
            F[3] = L.phig * L.pg * enz * area;
            F[6] = L.phip * L.pp * enx * area;
            F[7] = L.phip * L.pp * eny * area;
            F[8] = L.phip * L.pp * enz * area;
        } else if (FLUX == 1) {
            hllc_zien(P, L, R, enx, eny, enz, area, F);
        } else {
            hllc_phase<true>(P, L.phig, L.rog, L.ug, L.vg, L.wg, L.tg, L.pg, R.phig, R.rog, R.ug, R.vg, R.wg,
                             R.tg, R.pg, enx, eny, enz, area, F);
            hllc_phase<false>(P, L.phip, L.rop, L.up, L.vp, L.wp, L.tp, L.pp, R.phip, R.rop, R.up, R.vp,
                              R.wp, R.tp, R.pp, enx, eny, enz, area, F + 5);
        }
    }
    // the lane of face f keeps the face sums of its own variables k = f, f+4, f+8
    double rhs[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 10; k++) {
        double t[4];
        row_spread(F[k], t);
        const double sum = 0.0 + t[0] + t[1] + t[2] + t[3];
        if ((k & 3) == f) rhs[k >> 2] = sum;
    }
    source_and_update(s, m, i, f, rhs, stage, store, cv);
}

// Face fluxes and stage update of cell i (flux_update) for an active cell.
// feos (fused step): the update is not stored; lane f returns the cell's
// new conserved variables f, f+4, f+8 in cv (a frozen cell its unchanged
// ones) for the primitive recovery (eos_staged).  All four lanes of a quad
// share i, so the quad takes one branch.
template <int FLUX>
__device__ __forceinline__ void flux_face(const Params &P, const State &s, const Mesh &m, int i, int f,
                                          int stage, int feos, double (&cv)[3])
{
    if (m.iblank[i] == 1) {
        flux_update<FLUX>(P, s, m, i, f, stage, feos == 0, cv);
    } else if (feos) {
        const int ne = m.ne;
        cv[0] = s.cn[(size_t)f * ne + i];
        cv[1] = s.cn[(size_t)(f + 4) * ne + i];
        if (f + 8 < 10) cv[2] = s.cn[(size_t)(f + 8) * ne + i];
    }
}

// stage update from a stored residual (sr_RK.F90), for the per-routine path
__device__ __forceinline__ void rk_cell(const State &s, const Mesh &m, int i, int stage)
{
    const int ne = m.ne;
    if (m.iblank[i] != 1) return;
    const double dtbyvol = s.scal[SC_DT] / m.vol[i];
    for (int j = 0; j < 10; j++) {
        const size_t o = (size_t)j * ne + i;
        const double r = s.rhs[o] - (s.d24f ? (double)s.d24f[o] : s.d24[o]);
        const double cu = s.cu[o];
        double cn = s.cn[o];
        if (stage == 1) {
            cn = cu - (11.0 / 20.0) * dtbyvol * r;
            s.cn1[o] = cn;
        } else if (stage == 2) {
            cn = (3.0 / 8.0) * cu + (5.0 / 8.0) * (cn - (11.0 / 20.0) * dtbyvol * r);
        } else if (stage == 3) {
            cn = (4.0 / 9.0) * cu + (5.0 / 9.0) * (cn - (11.0 / 20.0) * dtbyvol * r);
        } else {
            cn = (111.0 / 1331.0) * cu + (260.0 / 1331.0) * s.cn1[o] +
                 (960.0 / 1331.0) * (cn - (11.0 / 20.0) * dtbyvol * r);
        }
        s.cn[o] = cn;
    }
}

// end of step: |cn-cu| of variables 1..4 for the monitor maximum
// (main.F90:603-612, a max is order independent) and cu = cn (:619-625)
__device__ __forceinline__ void end_cell(const State &s, int i, double *d)
{
    const int ne = s.ne;
    for (int k = 0; k < 10; k++) {
        const size_t o = (size_t)k * ne + i;
        const double c = s.cn[o];
        if (k < 4) d[k] = fabs(c - s.cu[o]);
        s.cu[o] = c;
    }
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
static inline int nblk(int n, int b) { return (n + b - 1) / b; }

void launch_geometry(const Mesh &m, double xmax, double ymax, double zmax, double ymin, double zmin,
                     hipStream_t st)
{
    hipLaunchKernelGGL(k_geom_cell, dim3(nblk(m.ne, 128)), dim3(128), 0, st, m);
    if (m.ng > 0)
        hipLaunchKernelGGL(k_geom_ghost, dim3(nblk(m.ng, 128)), dim3(128), 0, st, m, xmax, ymax, zmax, ymin, zmin);
}

}  // namespace hydra

namespace hydra {

// CFL bound of one cell from the current primitives (main.F90:525-546)
__device__ __forceinline__ double dt_cell(const Params &P, const State &s, int i)
{
    return cfl_bound(s.q, (size_t)s.nt, s.dl, i, P.gamag, P.gamap, P.pinf);
}

}  // namespace hydra

// ===========================================================================
// FP32-bulk variant (hp_params.precision = 1).  Gradients, limiter, face
// reconstruction, the HLLC and viscous face fluxes and the interphase
// source terms run in FP32; every difference of two FP64 fields (neighbour
// minus cell, face minus centroid, neighbour maximum minus cell, gas minus
// liquid velocity) is formed in FP64 before it is rounded, so FP32 sees
// small well-scaled numbers.  Face fluxes and source terms are widened to
// FP64 where they join the residual; the residual sums, the RK update, the
// EOS inversion, boundary conditions, dt reduction, smoothing and donor
// gather are the FP64 code above.
// Multiply-adds may contract here and divisions use the hardware
// reciprocal: these kernels are measured against the FP64 path, not
// required to match it bit for bit.
// ===========================================================================

namespace hydra {

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqr(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fmn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fmx(float a, float b) { return a > b ? a : b; }

__device__ __forceinline__ float det3f(float a11, float a21, float a31, float a12, float a22, float a32,
                                       float a13, float a23, float a33)
{
#pragma clang fp contract(on)
    return a11 * (a22 * a33 - a23 * a32) - a21 * (a12 * a33 - a13 * a32) + a31 * (a12 * a23 - a13 * a22);
}

__device__ __forceinline__ float fvenkatf(float delp, float delm, float epssq)
{
#pragma clang fp contract(on)
    const float num = (delp * delp + epssq) * delm + 2.0f * delm * delm * delp;
    const float den = delp * delp + 2.0f * delm * delm + delp * delm + epssq;
    return num * frcp(den * delm);
}

__device__ __forceinline__ void cell_source32(const Params &P, const State &s, const Mesh &m, int i,
                                              const float (&dx)[4], const float (&dy)[4], const float (&dz)[4],
                                              const float (&dph)[4], float (&T)[SRC32_N]);

// gradients (sr_derivative.F90:45-322, incl. the cTp2 defect) and limiter
// (sr_VK.F90:32-274) in FP32, one thread per cell, written with the cell's
// primitives as its record (R32_*); with src32 (HLLC-Batten) also the
// cell's interphase source terms (sr_source.F90:46-155), from the centroid
// offsets and phig differences the gradients use
__device__ __forceinline__ void grad_cell_f32(const Params &P, const State &s, const Mesh &m, int i, float *lrec)
{
#pragma clang fp contract(on)
    const int ne = m.ne;
    const double *q = s.q;
    const size_t nt = s.nt;
    // the cell's record: its primitives whatever its iblank; gradients and
    // limiter of an updated cell (a frozen one keeps zeros)
    float r[R32_N];
    r[0] = (float)QV(Q_ROG, i); r[1] = (float)QV(Q_UG, i); r[2] = (float)QV(Q_VG, i);
    r[3] = (float)QV(Q_WG, i); r[4] = (float)QV(Q_TG, i); r[5] = (float)QV(Q_PHIG, i);
    r[6] = (float)QV(Q_ROP, i); r[7] = (float)QV(Q_UP, i); r[8] = (float)QV(Q_VP, i);
    r[9] = (float)QV(Q_WP, i); r[10] = (float)QV(Q_TP, i); r[11] = (float)QV(Q_PG, i);
    r[12] = (float)QV(Q_PP, i);
#pragma unroll
    for (int k = R32_GRAD; k < R32_N; k++) r[k] = 0.0f;
    // the record goes to the wave's LDS staging (lrec: this lane's R32_N
    // floats) a 16-byte group at a time, as soon as the group's four values
    // are known; the wave then writes the tile's records out contiguously
    float4 *dst = reinterpret_cast<float4 *>(lrec);
    auto put = [&](int j) { dst[j] = make_float4(r[4 * j], r[4 * j + 1], r[4 * j + 2], r[4 * j + 3]); };
#pragma unroll
    for (int j = 0; j < 3; j++) put(j);
    if (m.iblank[i] != 1) {
#pragma unroll
        for (int j = 3; j < R32_N / 4; j++) put(j);
        return;
    }
    int nb[4];
    for (int k = 0; k < 4; k++) nb[k] = m.nc[(size_t)k * ne + i];
    // the neighbours' centroid offsets, rounded at set-up (grad32_geo)
    const float *cg = s.cgeo32;
    float dx[4], dy[4], dz[4], iw[4];
    for (int k = 0; k < 4; k++) {
        dx[k] = cg[(size_t)k * ne + i];
        dy[k] = cg[(size_t)(4 + k) * ne + i];
        dz[k] = cg[(size_t)(8 + k) * ne + i];
        iw[k] = frcp(dx[k] * dx[k] + dy[k] * dy[k] + dz[k] * dz[k]);
    }
    float a1 = 0.0f, a2 = 0.0f, a3 = 0.0f, b1 = 0.0f, b2 = 0.0f, b3 = 0.0f;
    for (int k = 0; k < 4; k++) {
        a1 += dx[k] * dx[k] * iw[k];
        a2 += dy[k] * dy[k] * iw[k];
        a3 += dz[k] * dz[k] * iw[k];
        b1 += dx[k] * dy[k] * iw[k];
        b2 += dx[k] * dz[k] * iw[k];
        b3 += dz[k] * dy[k] * iw[k];
    }
    if (s.src32) {
        // the source first, while the cell's FP64 primitives are in
        // registers; its phig differences are the gradient's (slot Q_PHIG)
        float dph[4], T[SRC32_N];
        const double phi_i = QV(Q_PHIG, i);
        for (int k = 0; k < 4; k++) dph[k] = (float)(QV(Q_PHIG, nb[k]) - phi_i);
        cell_source32(P, s, m, i, dx, dy, dz, dph, T);
#pragma unroll
        for (int j = 0; j < SRC32_N; j++) __builtin_nontemporal_store(T[j], &s.src32[(size_t)j * ne + i]);
    }
    const float idn = frcp(det3f(a1, b1, b2, b1, a2, b3, b2, b3, a3));
    const int slot[13] = {Q_UG, Q_VG, Q_WG, Q_TG, Q_UP, Q_VP, Q_WP, Q_TP, Q_PG, Q_PP, Q_PHIG, Q_ROG, Q_ROP};
    float glim[4][3], dpl[4], dml[4];
#pragma unroll
    for (int sI = 0; sI < 13; sI++) {
        const int sl = slot[sI];
        const double qi = QV(sl, i);
        double qn[4];
        for (int k = 0; k < 4; k++) qn[k] = QV(sl, nb[k]);
        float c1 = 0.0f, c2 = 0.0f, c3 = 0.0f;
        for (int k = 0; k < 4; k++) {
            const float dq = (float)(qn[k] - qi);
            c1 += dq * dx[k] * iw[k];
            c2 += (sl == Q_TP && k == 3) ? dq * dy[k] : dq * dy[k] * iw[k];  // sr_derivative.F90:225
            c3 += dq * dz[k] * iw[k];
        }
        const float gx = det3f(c1, c2, c3, b1, a2, b3, b2, b3, a3) * idn;
        const float gy = det3f(a1, b1, b2, c1, c2, c3, b2, b3, a3) * idn;
        const float gz = det3f(a1, b1, b2, b1, a2, b3, c1, c2, c3) * idn;
        if (sI < 11) {
            r[R32_GRAD + 3 * sI + 0] = gx;
            r[R32_GRAD + 3 * sI + 1] = gy;
            r[R32_GRAD + 3 * sI + 2] = gz;
            // groups whose last value is this slot's (index R32_GRAD + 3 sI + 2)
            // or an earlier one, and not complete before it
#pragma unroll
            for (int j = 3; j < R32_N / 4 - 1; j++)
                if (4 * j + 3 <= R32_GRAD + 3 * sI + 2 && 4 * j + 3 > R32_GRAD + 3 * sI - 1) put(j);
        }
        const int li = (sl == Q_PG) ? 0 : (sl == Q_PP) ? 1 : (sl == Q_ROG) ? 2 : (sl == Q_ROP) ? 3 : -1;
        if (li >= 0) {
            glim[li][0] = gx; glim[li][1] = gy; glim[li][2] = gz;
            // the limiter's extrema from the values just read
            double vmax = qi, vmin = qi;
            for (int k = 0; k < 4; k++) {
                vmax = dmax(vmax, qn[k]);
                vmin = dmin(vmin, qn[k]);
            }
            dpl[li] = (float)(vmax - qi);
            dml[li] = (float)(vmin - qi);
        }
    }
    const float epssq = (float)m.epssq[i];
    float r0[4][3];
    for (int f = 0; f < 4; f++) {
        r0[f][0] = cg[(size_t)(12 + f) * ne + i];
        r0[f][1] = cg[(size_t)(16 + f) * ne + i];
        r0[f][2] = cg[(size_t)(20 + f) * ne + i];
    }
    float ph = 1.0f;
    for (int v = 0; v < 4; v++) {
        const float dp = dpl[v], dm = dml[v];
        for (int f = 0; f < 4; f++) {
            const float delf = r0[f][0] * glim[v][0] + r0[f][1] * glim[v][1] + r0[f][2] * glim[v][2];
            const float pv = delf > 0.0f ? fvenkatf(dp, delf, epssq) : delf < 0.0f ? fvenkatf(dm, delf, epssq) : 1.0f;
            ph = (v == 0 && f == 0) ? pv : fmn(ph, pv);
        }
    }
    r[R32_PHI] = ph;
    put(R32_N / 4 - 1);
}

// the source terms alone (first-order FP32-bulk HLLC-Batten)
__device__ __forceinline__ void source_cell_f32(const Params &P, const State &s, const Mesh &m, int i)
{
    if (m.iblank[i] != 1) return;
    const int ne = m.ne;
    const double *q = s.q;
    const size_t nt = s.nt;
    const double xi = m.xc[i], yi = m.yc[i], zi = m.zc[i], phi_i = QV(Q_PHIG, i);
    float dx[4], dy[4], dz[4], dph[4];
    for (int k = 0; k < 4; k++) {
        const int nb = m.nc[(size_t)k * ne + i];
        dx[k] = (float)(m.xc[nb] - xi);
        dy[k] = (float)(m.yc[nb] - yi);
        dz[k] = (float)(m.zc[nb] - zi);
        dph[k] = (float)(QV(Q_PHIG, nb) - phi_i);
    }
    float T[SRC32_N];
    cell_source32(P, s, m, i, dx, dy, dz, dph, T);
#pragma unroll
    for (int j = 0; j < SRC32_N; j++) s.src32[(size_t)j * ne + i] = T[j];
}

struct FState32 {
    float rog, ug, vg, wg, tg, phig, rop, up, vp, wp, tp, pg, pp, phip;
};

__device__ __forceinline__ void load_fstate32(const double *q, size_t nt, int i, FState32 &s)
{
    s.rog = (float)QV(Q_ROG, i); s.ug = (float)QV(Q_UG, i); s.vg = (float)QV(Q_VG, i);
    s.wg = (float)QV(Q_WG, i); s.tg = (float)QV(Q_TG, i); s.phig = (float)QV(Q_PHIG, i);
    s.rop = (float)QV(Q_ROP, i); s.up = (float)QV(Q_UP, i); s.vp = (float)QV(Q_VP, i);
    s.wp = (float)QV(Q_WP, i); s.tp = (float)QV(Q_TP, i); s.pg = (float)QV(Q_PG, i);
    s.pp = (float)QV(Q_PP, i);
}

// sr_hllcbatten.F90:228-318 in FP32
__device__ __forceinline__ void clip_blend32(const Params &P, FState32 &s)
{
#pragma clang fp contract(on)
    const float emin = (float)P.epslnmin, emax = (float)P.epslnmax, ide = frcp(emax - emin);
    if (s.phig < emin) {
        s.phig = emin; s.phip = 1.0f - s.phig;
        s.ug = s.up; s.vg = s.vp; s.wg = s.wp; s.tg = s.tp; s.pg = s.pp;
    }
    if (s.phip < emin) {
        s.phip = emin; s.phig = 1.0f - s.phip;
        s.up = s.ug; s.vp = s.vg; s.wp = s.wg; s.tp = s.tg; s.pp = s.pg;
    }
    if (s.phig >= emin && s.phig <= emax) {
        const float psig = (s.phig - emin) * ide;
        const float fg = -psig * psig * (2.0f * psig - 3.0f);
        s.ug = fg * s.ug + (1.0f - fg) * s.up;
        s.vg = fg * s.vg + (1.0f - fg) * s.vp;
        s.wg = fg * s.wg + (1.0f - fg) * s.wp;
        s.tg = fg * s.tg + (1.0f - fg) * s.tp;
        s.pg = fg * s.pg + (1.0f - fg) * s.pp;
    }
    if (s.phip >= emin && s.phip <= emax) {
        const float psig = (s.phip - emin) * ide;
        const float fg = -psig * psig * (2.0f * psig - 3.0f);
        s.up = fg * s.up + (1.0f - fg) * s.ug;
        s.vp = fg * s.vp + (1.0f - fg) * s.vg;
        s.wp = fg * s.wp + (1.0f - fg) * s.wg;
        s.tp = fg * s.tp + (1.0f - fg) * s.tg;
        s.pp = fg * s.pp + (1.0f - fg) * s.pg;
    }
    s.rog = s.pg * frcp((float)P.rrg * s.tg);
    s.rop = (float)P.gamap * (s.pp + (float)P.pinf) * frcp((float)((P.gamap - 1.0) * P.cpp) * s.tp);
}

// one phase of the HLLC-Batten flux (sr_hllcbatten.F90:338-529) in FP32
template <bool GAS>
__device__ __forceinline__ void hllc_phase32(const Params &P, float phL, float roL, float uL, float vL, float wL,
                                             float tL, float pL, float phR, float roR, float uR, float vR, float wR,
                                             float tR, float pR, float enx, float eny, float enz, float area,
                                             float (&Fa)[10])
{
#pragma clang fp contract(on)
    const float irL = frcp(roL), irR = frcp(roR);
    float EL, ER, aL, aR;
    if (GAS) {
        const float cv = (float)(P.rrg / (P.gamag - 1.0)), g = (float)P.gamag;
        EL = cv * tL + 0.5f * (uL * uL + vL * vL + wL * wL);
        ER = cv * tR + 0.5f * (uR * uR + vR * vR + wR * wR);
        aL = fsqr(g * pL * irL);
        aR = fsqr(g * pR * irR);
    } else {
        const float cv = (float)(P.cpp / P.gamap), g = (float)P.gamap, pinf = (float)P.pinf;
        EL = cv * tL + pinf * irL + 0.5f * (uL * uL + vL * vL + wL * wL);
        ER = cv * tR + pinf * irR + 0.5f * (uR * uR + vR * vR + wR * wR);
        aL = fsqr(g * (pL + pinf) * irL);
        aR = fsqr(g * (pR + pinf) * irR);
    }
    const float HL = EL + pL * irL, HR = ER + pR * irR;
    const float qL = uL * enx + vL * eny + wL * enz;
    const float qR = uR * enx + vR * eny + wR * enz;
    const float SL = fmn(fmn(qL - aL, qR - aR), 1.0e-8f);
    const float SR = fmx(fmx(qL + aL, qR + aR), 1.0e-8f);
    const float snr = pR - pL + roL * qL * (SL - qL) - roR * qR * (SR - qR);
    const float sdr = roL * (SL - qL) - roR * (SR - qR);
    const float SS = snr * frcp(sdr + 1.0e-8f);
    const float pst = fmx(pL + roL * (qL - SL) * (qL - SS), 1.0e-6f * fmn(pL, pR));
    const float mL = phL * roL, mR = phR * roR;
    // the branch's five components, stored once below (gas 0-4, liquid 5-9)
    float r0, r1, r2, r3, r4;
    if (SL > 0.0f) {
        r0 = mL * qL * area;
        r1 = (mL * qL * uL + phL * pL * enx) * area;
        r2 = (mL * qL * vL + phL * pL * eny) * area;
        r3 = (mL * qL * wL + phL * pL * enz) * area;
        r4 = mL * qL * HL * area;
    } else if (SS > 0.0f) {
        const float idn = frcp((SS - SL) + 1.0e-6f);
        const float c1 = mL, c2 = mL * uL, c3 = mL * vL, c4 = mL * wL, c5 = mL * EL, d = qL - SL;
        const float s1 = d * c1 * idn;
        const float s2 = (d * c2 + phL * (pL - pst) * enx) * idn;
        const float s3 = (d * c3 + phL * (pL - pst) * eny) * idn;
        const float s4 = (d * c4 + phL * (pL - pst) * enz) * idn;
        const float s5 = (d * c5 + phL * (pL * qL - pst * SS)) * idn;
        r0 = (mL * qL + SL * (s1 - c1)) * area;
        r1 = (mL * qL * uL + phL * pL * enx + SL * (s2 - c2)) * area;
        r2 = (mL * qL * vL + phL * pL * eny + SL * (s3 - c3)) * area;
        r3 = (mL * qL * wL + phL * pL * enz + SL * (s4 - c4)) * area;
        r4 = (mL * qL * HL + SL * (s5 - c5)) * area;
    } else if (SR >= 0.0f) {
        const float idn = frcp((SS - SR) + 1.0e-6f);
        const float c1 = mR, c2 = mR * uR, c3 = mR * vR, c4 = mR * wR, c5 = mR * ER, d = qR - SR;
        const float s1 = d * c1 * idn;
        const float s2 = (d * c2 + phR * (pR - pst) * enx) * idn;
        const float s3 = (d * c3 + phR * (pR - pst) * eny) * idn;
        const float s4 = (d * c4 + phR * (pR - pst) * enz) * idn;
        const float s5 = (d * c5 + phR * (pR * qR - pst * SS)) * idn;
        r0 = (mR * qR + SR * (s1 - c1)) * area;
        r1 = (mR * qR * uR + phR * pR * enx + SR * (s2 - c2)) * area;
        r2 = (mR * qR * vR + phR * pR * eny + SR * (s3 - c3)) * area;
        r3 = (mR * qR * wR + phR * pR * enz + SR * (s4 - c4)) * area;
        r4 = (mR * qR * HR + SR * (s5 - c5)) * area;
    } else {
        r0 = mR * qR * area;
        r1 = (mR * qR * uR + phR * pR * enx) * area;
        r2 = (mR * qR * vR + phR * pR * eny) * area;
        r3 = (mR * qR * wR + phR * pR * enz) * area;
        r4 = mR * qR * HR * area;
    }
    constexpr int o = GAS ? 0 : 5;
    Fa[o] = r0; Fa[o + 1] = r1; Fa[o + 2] = r2; Fa[o + 3] = r3; Fa[o + 4] = r4;
}

// HLLC-Zien flux (sr_hllczien.F90:327-487) in FP32
__device__ __forceinline__ void hllc_zien32(const Params &P, const FState32 &L, const FState32 &R, float enx,
                                            float eny, float enz, float area, float *F)
{
#pragma clang fp contract(on)
    const float cvg = (float)(P.rrg / (P.gamag - 1.0)), gg = (float)P.gamag;
    const float cvp = (float)(P.cpp / P.gamap), gp = (float)P.gamap, pinf = (float)P.pinf;
    const float irgL = frcp(L.rog), irgR = frcp(R.rog), irpL = frcp(L.rop), irpR = frcp(R.rop);
    const float EGL = cvg * L.tg + 0.5f * (L.ug * L.ug + L.vg * L.vg + L.wg * L.wg);
    const float HGL = EGL + L.pg * irgL;
    const float QGL = L.ug * enx + L.vg * eny + L.wg * enz;
    const float AGL = fsqr(gg * L.pg * irgL);
    const float EGR = cvg * R.tg + 0.5f * (R.ug * R.ug + R.vg * R.vg + R.wg * R.wg);
    const float HGR = EGR + R.pg * irgR;
    const float QGR = R.ug * enx + R.vg * eny + R.wg * enz;
    const float AGR = fsqr(gg * R.pg * irgR);
    const float EPL = cvp * L.tp + pinf * irpL + 0.5f * (L.up * L.up + L.vp * L.vp + L.wp * L.wp);
    const float HPL = EPL + L.pp * irpL;
    const float QPL = L.up * enx + L.vp * eny + L.wp * enz;
    const float APL = fsqr(gp * (L.pp + pinf) * irpL);
    const float EPR = cvp * R.tp + pinf * irpR + 0.5f * (R.up * R.up + R.vp * R.vp + R.wp * R.wp);
    const float HPR = EPR + R.pp * irpR;
    const float QPR = R.up * enx + R.vp * eny + R.wp * enz;
    const float APR = fsqr(gp * (R.pp + pinf) * irpR);
    const float SL = fmn(fmn(fmn(fmn(QGL - AGL, QPL - APL), QGR - AGR), QPR - APR), 1.0e-8f);
    const float SR = fmx(fmx(fmx(fmx(QGL + AGL, QPL + APL), QGR + AGR), QPR + APR), 1.0e-8f);
    const float ROL = L.phig * L.rog + L.phip * L.rop;
    const float ROQL = L.phig * L.rog * QGL + L.phip * L.rop * QPL;
    const float QL = ROQL * frcp(ROL);
    const float PRL = L.phig * L.pg + L.phip * L.pp;
    const float ROR = R.phig * R.rog + R.phip * R.rop;
    const float ROQR = R.phig * R.rog * QGR + R.phip * R.rop * QPR;
    const float QR = ROQR * frcp(ROR);
    const float PRR = R.phig * R.pg + R.phip * R.pp;
    const float SSTAR = (PRR - PRL + ROQL * (SL - QL) - ROQR * (SR - QR)) *
                        frcp(ROL * (SL - QL) - ROR * (SR - QR) + 1.0e-8f);
    const float PGSTAR = fmx(L.pg + L.rog * (QGL - SL) * (QGL - SSTAR), 1.0e-6f * fmn(L.pg, R.pg));
    const float PPSTAR = fmx(L.pp + L.rop * (QPL - SL) * (QPL - SSTAR), 1.0e-6f * fmn(L.pp, R.pp));
    const bool left = SL >= 0.0f || SSTAR > 0.0f;
    const FState32 &S = left ? L : R;
    const float Sw = left ? SL : SR;
    const float QG = left ? QGL : QGR, QP = left ? QPL : QPR, HG = left ? HGL : HGR, HP = left ? HPL : HPR;
    const float EG = left ? EGL : EGR, EP = left ? EPL : EPR;
    const float mg = S.phig * S.rog, mp = S.phip * S.rop;
    F[0] = mg * QG * area;
    F[1] = (mg * QG * S.ug + S.phig * S.pg * enx) * area;
    F[2] = (mg * QG * S.vg + S.phig * S.pg * eny) * area;
    F[3] = (mg * QG * S.wg + S.phig * S.pg * enz) * area;
    F[4] = mg * QG * HG * area;
    F[5] = mp * QP * area;
    F[6] = (mp * QP * S.up + S.phip * S.pp * enx) * area;
    F[7] = (mp * QP * S.vp + S.phip * S.pp * eny) * area;
    F[8] = (mp * QP * S.wp + S.phip * S.pp * enz) * area;
    F[9] = mp * QP * HP * area;
    const bool star = left ? !(SL >= 0.0f) : (SR >= 0.0f);
    if (star) {
        const float idn = frcp((SSTAR - Sw) + 1.0e-6f);
        const float dg = QG - Sw, dp = QP - Sw;
        const float cg[5] = {mg, mg * S.ug, mg * S.vg, mg * S.wg, mg * EG};
        const float cp[5] = {mp, mp * S.up, mp * S.vp, mp * S.wp, mp * EP};
        const float sg[5] = {dg * cg[0] * idn, (dg * cg[1] + S.phig * (S.pg - PGSTAR) * enx) * idn,
                             (dg * cg[2] + S.phig * (S.pg - PGSTAR) * eny) * idn,
                             (dg * cg[3] + S.phig * (S.pg - PGSTAR) * enz) * idn,
                             (dg * cg[4] + S.phig * (S.pg * QG - PGSTAR * SSTAR)) * idn};
        const float sp[5] = {dp * cp[0] * idn, (dp * cp[1] + S.phip * (S.pp - PPSTAR) * enx) * idn,
                             (dp * cp[2] + S.phip * (S.pp - PPSTAR) * eny) * idn,
                             (dp * cp[3] + S.phip * (S.pp - PPSTAR) * enz) * idn,
                             (dp * cp[4] + S.phip * (S.pp * QP - PPSTAR * SSTAR)) * idn};
        for (int k = 0; k < 5; k++) {
            F[k] += Sw * (sg[k] - cg[k]) * area;
            F[5 + k] += Sw * (sp[k] - cp[k]) * area;
        }
    }
}

__device__ __forceinline__ float quad_sum4f(float v, int qbase)
{
    const float t0 = __shfl(v, qbase, 64), t1 = __shfl(v, qbase + 1, 64), t2 = __shfl(v, qbase + 2, 64),
                t3 = __shfl(v, qbase + 3, 64);
    return t0 + t1 + t2 + t3;
}

// interphase source (sr_source.F90:46-155) in FP32 for the FP32-bulk form:
// differences of FP64 fields are formed in FP64 (the faces' centroid offsets
// dx, dy, dz and phig differences dph come in rounded from them); the
// source terms are widened to FP64 where they join the FP64 residual, and
// the RK update is the FP64 one.  Src32 holds cell i's terms from the sums
// of its faces' phig LSQ (a1..c3).
struct Src32 {
    float fdgx, fdgy, fdgz, pint, dphx, dphy, dphz, work, phig, rog, phip, rop, grav;
};
__device__ __forceinline__ Src32 source_f32(const Params &P, const State &s, const Mesh &m, int i, float a1,
                                            float a2, float a3, float b1, float b2, float b3, float c1, float c2,
                                            float c3)
{
#pragma clang fp contract(on)
    const double *q = s.q;
    const size_t nt = s.nt;
    const double phig_d = QV(Q_PHIG, i);
    const float phig = (float)phig_d, phip = (float)QV(Q_PHIP, i), rog = (float)QV(Q_ROG, i),
                rop = (float)QV(Q_ROP, i);
    const double ug_d = QV(Q_UG, i), vg_d = QV(Q_VG, i), wg_d = QV(Q_WG, i), up_d = QV(Q_UP, i),
                 vp_d = QV(Q_VP, i), wp_d = QV(Q_WP, i), tg_d = QV(Q_TG, i), tp_d = QV(Q_TP, i);
    const float ug = (float)ug_d, vg = (float)vg_d, wg = (float)wg_d, up = (float)up_d, vp = (float)vp_d,
                wp = (float)wp_d, tg = (float)tg_d, pg = (float)QV(Q_PG, i);
    const float dl = (float)m.dl[i];
    const float amug = (float)P.as * fsqr(tg) * frcp(1.0f + (float)P.ts * frcp(tg));
    const float du = (float)(ug_d - up_d), dv = (float)(vg_d - vp_d), dw = (float)(wg_d - wp_d);
    const float q2 = du * du + dv * dv + dw * dw;
    const float drag = 0.44f * phig * (phig - 1.0f) * rog * fsqr(q2) * dl;
    Src32 r;
    r.fdgx = drag * du;
    r.fdgy = drag * dv;
    r.fdgz = drag * dw;
    const float qsource = 6.0f * phip * amug * (float)P.cpg * 2.0f * (float)(tp_d - tg_d) *
                          frcp((float)P.prg * (dl * dl));
    const float delpstar = 2.0f * (phig * rog * phip * rop) * frcp(phig * rop + phip * rog) * q2;
    r.pint = pg - fmn(delpstar, 0.01f * pg);
    const float iden = frcp(phig * rog + (1.0f - phig) * rop);
    const float uint = (phig * rog * ug + (1.0f - phig) * rop * up) * iden;
    const float vint = (phig * rog * vg + (1.0f - phig) * rop * vp) * iden;
    const float wint = (phig * rog * wg + (1.0f - phig) * rop * wp) * iden;
    const float idn = frcp(det3f(a1, b1, b2, b1, a2, b3, b2, b3, a3));
    r.dphx = det3f(c1, c2, c3, b1, a2, b3, b2, b3, a3) * idn;
    r.dphy = det3f(a1, b1, b2, c1, c2, c3, b2, b3, a3) * idn;
    r.dphz = det3f(a1, b1, b2, b1, a2, b3, c1, c2, c3) * idn;
    r.work = r.pint * (uint * r.dphx + vint * r.dphy + wint * r.dphz) + uint * r.fdgx + vint * r.fdgy +
             wint * r.fdgz + qsource;
    r.grav = (float)P.gravity;
    r.phig = phig;
    r.rog = rog;
    r.phip = phip;
    r.rop = rop;
    return r;
}

// The five FP32 terms the stage update takes from Src32 (SRC32_N order:
// variable 1, variables 2 and 7, 3 and 8, 4 and 9, variable 6), each one
// expression, so every form that computes them rounds alike
__device__ __forceinline__ void src_terms32(const Src32 &S, float (&T)[SRC32_N])
{
#pragma clang fp contract(on)
    T[0] = S.fdgx + S.pint * S.dphx + S.phig * S.rog * S.grav;
    T[1] = S.fdgy + S.pint * S.dphy;
    T[2] = S.fdgz + S.pint * S.dphz;
    T[3] = S.work;
    T[4] = S.fdgx + S.pint * S.dphx - S.phip * S.rop * S.grav;
}

// the source's 1/|d| least-squares sums over a cell's four faces
// (sr_source.F90:60-95) from the faces' centroid offsets and phig
// differences: each face's terms as their own statements (no contraction
// across faces), then the sums in face order
__device__ __forceinline__ void src_lsq32(const float (&dx)[4], const float (&dy)[4], const float (&dz)[4],
                                          const float (&dph)[4], float (&sum)[9])
{
#pragma clang fp contract(on)
    float t[9][4];
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const float iw = frcp(fsqr(dx[f] * dx[f] + dy[f] * dy[f] + dz[f] * dz[f]));
        t[0][f] = dx[f] * dx[f] * iw;
        t[1][f] = dy[f] * dy[f] * iw;
        t[2][f] = dz[f] * dz[f] * iw;
        t[3][f] = dx[f] * dy[f] * iw;
        t[4][f] = dx[f] * dz[f] * iw;
        t[5][f] = dy[f] * dz[f] * iw;
        t[6][f] = dph[f] * dx[f] * iw;
        t[7][f] = dph[f] * dy[f] * iw;
        t[8][f] = dph[f] * dz[f] * iw;
    }
#pragma unroll
    for (int j = 0; j < 9; j++) sum[j] = t[j][0] + t[j][1] + t[j][2] + t[j][3];
}

// cell i's source terms (SRC32_N) from its four faces' offsets and phig
// differences
__device__ __forceinline__ void cell_source32(const Params &P, const State &s, const Mesh &m, int i,
                                              const float (&dx)[4], const float (&dy)[4], const float (&dz)[4],
                                              const float (&dph)[4], float (&T)[SRC32_N])
{
    float sum[9];
    src_lsq32(dx, dy, dz, dph, sum);
    const Src32 S = source_f32(P, s, m, i, sum[0], sum[1], sum[2], sum[3], sum[4], sum[5], sum[6], sum[7], sum[8]);
    src_terms32(S, T);
}

// the stage update of conserved variable k of cell i from its face sum r:
// the source term (T, SRC32_N order) and viscous sum joined in FP64, then
// the RK stage
__device__ __forceinline__ double update_f32(const State &s, int ne, int i, int k, double r,
                                             const float (&T)[SRC32_N], double voli, int stage, double dtbyvol,
                                             bool store)
{
#pragma clang fp contract(on)
    const float *vis = s.visf;
    switch (k) {
    case 1: r = r - (double)T[0] * voli - vis[(size_t)0 * ne + i]; break;
    case 2: r = r - (double)T[1] * voli - vis[(size_t)1 * ne + i]; break;
    case 3: r = r - (double)T[2] * voli - vis[(size_t)2 * ne + i]; break;
    case 4: r = r - (double)T[3] * voli - vis[(size_t)3 * ne + i]; break;
    case 6: r = r + (double)T[4] * voli - vis[(size_t)4 * ne + i]; break;
    case 7: r = r + (double)T[1] * voli - vis[(size_t)5 * ne + i]; break;
    case 8: r = r + (double)T[2] * voli - vis[(size_t)6 * ne + i]; break;
    case 9: r = r + (double)T[3] * voli - vis[(size_t)7 * ne + i]; break;
    default: break;
    }
    return rk_finish(s, s.d24f, (size_t)k * ne + i, r, stage, dtbyvol, store);
}

// the quad form: lane f holds face f's offsets and finishes the conserved
// variables k = f, f+4, f+8 from their face sums rhs[0..2]
__device__ __forceinline__ void source_and_update_f32(const Params &P, const State &s, const Mesh &m, int i,
                                                      int f, int qbase, float dx, float dy, float dz, float dph,
                                                      const double *rhs, int stage, bool store, double (&cv)[3])
{
#pragma clang fp contract(on)
    float a1, a2, a3, b1, b2, b3, c1, c2, c3;
    {
        const float iw = frcp(fsqr(dx * dx + dy * dy + dz * dz));
        a1 = quad_sum4f(dx * dx * iw, qbase);
        a2 = quad_sum4f(dy * dy * iw, qbase);
        a3 = quad_sum4f(dz * dz * iw, qbase);
        b1 = quad_sum4f(dx * dy * iw, qbase);
        b2 = quad_sum4f(dx * dz * iw, qbase);
        b3 = quad_sum4f(dy * dz * iw, qbase);
        c1 = quad_sum4f(dph * dx * iw, qbase);
        c2 = quad_sum4f(dph * dy * iw, qbase);
        c3 = quad_sum4f(dph * dz * iw, qbase);
    }
    const Src32 S = source_f32(P, s, m, i, a1, a2, a3, b1, b2, b3, c1, c2, c3);
    float T[SRC32_N];
    src_terms32(S, T);
    const double voli = m.vol[i];
    const double dtbyvol = (stage > 0) ? s.scal[SC_DT] / voli : 0.0;
    for (int j = 0; j < 3; j++) {
        const int k = f + 4 * j;
        if (k >= 10) break;
        cv[j] = update_f32(s, m.ne, i, k, rhs[j], T, voli, stage, dtbyvol, store);
    }
}

// the cell form (HLLC-Batten, after k_face32): one thread per cell reads its
// four face entries and sums the fluxes in face order (widened to FP64),
// takes its source terms from the gradient pass (src32) and finishes all
// ten conserved variables
__device__ __forceinline__ void cell_update_f32(const State &s, const Mesh &m, int i, int stage, bool store,
                                                double (&cv)[10])
{
#pragma clang fp contract(on)
    const int ne = m.ne;
    float F[4][10];
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const int code = s.fslot[(size_t)f * ne + i];
        const float4 *src = reinterpret_cast<const float4 *>(s.fflux) + (size_t)FACE32_N / 4 * (code >> 1);
        const float4 u0 = src[0], u1 = src[1];
        const float2 u2 = *reinterpret_cast<const float2 *>(src + 2);
        const float sg = (code & 1) ? -1.0f : 1.0f;
        F[f][0] = sg * u0.x; F[f][1] = sg * u0.y; F[f][2] = sg * u0.z; F[f][3] = sg * u0.w;
        F[f][4] = sg * u1.x; F[f][5] = sg * u1.y; F[f][6] = sg * u1.z; F[f][7] = sg * u1.w;
        F[f][8] = sg * u2.x; F[f][9] = sg * u2.y;
    }
    float T[SRC32_N];
#pragma unroll
    for (int j = 0; j < SRC32_N; j++) T[j] = s.src32[(size_t)j * ne + i];
    const double voli = m.vol[i];
    const double dtbyvol = (stage > 0) ? s.scal[SC_DT] / voli : 0.0;
#pragma unroll
    for (int k = 0; k < 10; k++) {
        const double r = 0.0 + (double)F[0][k] + (double)F[1][k] + (double)F[2][k] + (double)F[3][k];
        cv[k] = update_f32(s, ne, i, k, r, T, voli, stage, dtbyvol, store);
    }
}

// a face state from a record's primitives (R32_PRIM order)
__device__ __forceinline__ void rec_state32(const float (&r)[R32_N], FState32 &st)
{
    st.rog = r[0]; st.ug = r[1]; st.vg = r[2]; st.wg = r[3]; st.tg = r[4]; st.phig = r[5]; st.rop = r[6];
    st.up = r[7]; st.vp = r[8]; st.wp = r[9]; st.tp = r[10]; st.pg = r[11]; st.pp = r[12];
}

// A face's geometry as the FP32 face loop uses it: the area vector rounded
// from FP64, the face centroid's offsets from cell i's and the neighbour's
// centroids and the neighbour's offset from cell i, each difference formed
// in FP64 and rounded (the offsets to the face centroid only where a cell
// lies on the far side)
struct FaceGeo32 {
    float scx, scy, scz, xf0, yf0, zf0, xfn, yfn, zfn, dx, dy, dz;
};
__device__ __forceinline__ FaceGeo32 face_geo32(const Mesh &m, int i, int f, int neigh)
{
    const int ne = m.ne;
    FaceGeo32 g;
    g.scx = (float)m.sc[(size_t)(3 * f) * ne + i];
    g.scy = (float)m.sc[(size_t)(3 * f + 1) * ne + i];
    g.scz = (float)m.sc[(size_t)(3 * f + 2) * ne + i];
    g.xf0 = g.yf0 = g.zf0 = g.xfn = g.yfn = g.zfn = 0.0f;
    if (neigh < ne) {
        const double xf = m.fc0[(size_t)(3 * f) * ne + i], yf = m.fc0[(size_t)(3 * f + 1) * ne + i],
                     zf = m.fc0[(size_t)(3 * f + 2) * ne + i];
        g.xf0 = (float)(xf - m.xc[i]);
        g.yf0 = (float)(yf - m.yc[i]);
        g.zf0 = (float)(zf - m.zc[i]);
        g.xfn = (float)(xf - m.xc[neigh]);
        g.yfn = (float)(yf - m.yc[neigh]);
        g.zfn = (float)(zf - m.zc[neigh]);
    }
    g.dx = (float)(m.xc[neigh] - m.xc[i]);
    g.dy = (float)(m.yc[neigh] - m.yc[i]);
    g.dz = (float)(m.zc[neigh] - m.zc[i]);
    return g;
}

// One side's FP32 face state from its record r, reconstructed with phimin
// and the side's face-centroid offset (xo, yo, zo) on interior faces
__device__ __forceinline__ void side_state32(const float (&r)[R32_N], bool interior, float phimin, float xo,
                                             float yo, float zo, FState32 &S)
{
#pragma clang fp contract(on)
    rec_state32(r, S);
    if (interior) {
#define REC32S(fld, sl)                                                                              \
    S.fld += phimin * (xo * r[R32_GRAD + 3 * (sl)] + yo * r[R32_GRAD + 3 * (sl) + 1] +               \
                       zo * r[R32_GRAD + 3 * (sl) + 2]);
        REC32S(phig, G_PHIG) REC32S(ug, G_UG) REC32S(vg, G_VG) REC32S(wg, G_WG) REC32S(tg, G_TG)
        REC32S(pg, G_PG) REC32S(up, G_UP) REC32S(vp, G_VP) REC32S(wp, G_WP) REC32S(tp, G_TP)
        REC32S(pp, G_PP)
#undef REC32S
    }
}

// the flux of a face from both sides' FP32 face states (defined below)
template <int FLUX>
__device__ __forceinline__ void flux_from_states32(const Params &P, const Mesh &m, int neigh, const FaceGeo32 &g,
                                                   FState32 &L, FState32 &R, float (&F)[10]);

// The FP32 flux of cell i's face towards neigh (the face loop of
// hllcbatten / hllczien) with the face's geometry g: both sides' face states
// (the records' primitives, reconstructed with the smaller limiter on
// interior faces; this lane reads both records, twelve 16-byte loads each)
// or, in first order, the primitives, then clipping and blending and the
// wall flux or the HLLC flux, into F.
template <int FLUX>
__device__ __forceinline__ void face_flux_f32(const Params &P, const State &s, const Mesh &m, int i, int neigh,
                                              const FaceGeo32 &g, float (&F)[10])
{
    FState32 L, R;
    if (P.myorder == 2) {
        float a[R32_N], b[R32_N];
        const float4 *ra = reinterpret_cast<const float4 *>(s.rec32) + (size_t)(R32_N / 4) * i;
        const float4 *rb = reinterpret_cast<const float4 *>(s.rec32) + (size_t)(R32_N / 4) * neigh;
#pragma unroll
        for (int j = 0; j < R32_N / 4; j++) {
            const float4 u = ra[j], v = rb[j];
            a[4 * j] = u.x; a[4 * j + 1] = u.y; a[4 * j + 2] = u.z; a[4 * j + 3] = u.w;
            b[4 * j] = v.x; b[4 * j + 1] = v.y; b[4 * j + 2] = v.z; b[4 * j + 3] = v.w;
        }
        const bool interior = neigh < m.ne;
        const float phimin = fmn(a[R32_PHI], b[R32_PHI]);
        side_state32(a, interior, phimin, g.xf0, g.yf0, g.zf0, L);
        side_state32(b, interior, phimin, g.xfn, g.yfn, g.zfn, R);
    } else {
        load_fstate32(s.q, s.nt, i, L);
        load_fstate32(s.q, s.nt, neigh, R);
    }
    flux_from_states32<FLUX>(P, m, neigh, g, L, R, F);
}

// The flux of a face from both sides' FP32 face states: clipping and
// blending, then the wall flux (pressure terms only) or the HLLC flux.
template <int FLUX>
__device__ __forceinline__ void flux_from_states32(const Params &P, const Mesh &m, int neigh, const FaceGeo32 &g,
                                                   FState32 &L, FState32 &R, float (&F)[10])
{
#pragma clang fp contract(on)
    const int ne = m.ne;
    const float scx = g.scx, scy = g.scy, scz = g.scz;
    const float area = fsqr(scx * scx + scy * scy + scz * scz), ia = frcp(area);
    const float enx = scx * ia, eny = scy * ia, enz = scz * ia;
    const int ntype = (neigh >= ne) ? m.gtype[neigh - ne] : 0;
    L.phip = 1.0f - L.phig;
    R.phip = 1.0f - R.phig;
    clip_blend32(P, L);
    clip_blend32(P, R);
    // the HLLC flux everywhere, the wall flux (pressure terms only) selected
    // on wall faces: every component is written once, so F stays in registers
    float G[10];
    if (FLUX == 1) {
        hllc_zien32(P, L, R, enx, eny, enz, area, G);
    } else {
        hllc_phase32<true>(P, L.phig, L.rog, L.ug, L.vg, L.wg, L.tg, L.pg, R.phig, R.rog, R.ug, R.vg, R.wg,
                           R.tg, R.pg, enx, eny, enz, area, G);
        hllc_phase32<false>(P, L.phip, L.rop, L.up, L.vp, L.wp, L.tp, L.pp, R.phip, R.rop, R.up, R.vp,
                            R.wp, R.tp, R.pp, enx, eny, enz, area, G);
    }
    const bool wall = ntype == 5;
    F[0] = wall ? 0.0f : G[0];
    F[1] = wall ? L.phig * L.pg * enx * area : G[1];
    F[2] = wall ? L.phig * L.pg * eny * area : G[2];
    F[3] = wall ? L.phig * L.pg * enz * area : G[3];
    F[4] = wall ? 0.0f : G[4];
    F[5] = wall ? 0.0f : G[5];
    F[6] = wall ? L.phip * L.pp * enx * area : G[6];
    F[7] = wall ? L.phip * L.pp * eny * area : G[7];
    F[8] = wall ? L.phip * L.pp * enz * area : G[8];
    F[9] = wall ? 0.0f : G[9];
}

// the source's phig difference across the face, formed in FP64 and rounded
// (it changes sign with the face's orientation, exactly, as the centroid
// offset does)
__device__ __forceinline__ float face_dphig32(const State &s, int i, int neigh)
{
    const double *q = s.q;
    const size_t nt = s.nt;
    return (float)(QV(Q_PHIG, neigh) - QV(Q_PHIG, i));
}

// face loop in FP32, one cell per quad of lanes as in k_flux, each face
// evaluated from the cell's side (HLLC-Zien; HLLC-Batten runs k_face32 and
// k_cells32 instead); the quad's face fluxes are widened to FP64 and summed
// in face order, then the FP64 source and stage update run unchanged.
template <int FLUX>
__device__ __forceinline__ void flux_update_f32(const Params &P, const State &s, const Mesh &m, int i, int f,
                                                int qbase, int stage, bool store, double (&cv)[3])
{
#pragma clang fp contract(on)
    float F[10];
    const int neigh = m.nc[(size_t)f * m.ne + i];
    const FaceGeo32 g = face_geo32(m, i, f, neigh);
    face_flux_f32<FLUX>(P, s, m, i, neigh, g, F);
    const float dx = g.dx, dy = g.dy, dz = g.dz, dph = face_dphig32(s, i, neigh);
    double rhs[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 10; k++) {
        const double t0 = (double)__shfl(F[k], qbase, 64), t1 = (double)__shfl(F[k], qbase + 1, 64),
                     t2 = (double)__shfl(F[k], qbase + 2, 64), t3 = (double)__shfl(F[k], qbase + 3, 64);
        const double sum = 0.0 + t0 + t1 + t2 + t3;
        if ((k & 3) == f) rhs[k >> 2] = sum;
    }
    source_and_update_f32(P, s, m, i, f, qbase, dx, dy, dz, dph, rhs, stage, store, cv);
}

// flux_face with the FP32-bulk face loop
template <int FLUX>
__device__ __forceinline__ void flux_face_f32(const Params &P, const State &s, const Mesh &m, int i, int f,
                                              int qbase, int stage, int feos, double (&cv)[3])
{
    if (m.iblank[i] == 1) {
        flux_update_f32<FLUX>(P, s, m, i, f, qbase, stage, feos == 0, cv);
    } else if (feos) {
        const int ne = m.ne;
        cv[0] = s.cn[(size_t)f * ne + i];
        cv[1] = s.cn[(size_t)(f + 4) * ne + i];
        if (f + 8 < 10) cv[2] = s.cn[(size_t)(f + 8) * ne + i];
    }
}

// viscous_stencil in FP32, with the quad's shared points as there:
// differences of FP64 fields are formed in FP64 and rounded once; sums,
// determinants and stresses run in FP32.  A boundary-face lane zeroes the
// offsets of points 5-7, so they add nothing.
__device__ __forceinline__ void viscous_stencil_f32(const Params &P, const double *q, size_t nt, const Mesh &m,
                                                    const int (&m2)[4], bool ghost, double xn, double yn,
                                                    double zn, const double (&qn)[8], double xf, double yf,
                                                    double zf, float scx, float scy, float scz,
                                                    const double (&fv)[8], float phigf, float phipf,
                                                    float (&acc)[8])
{
#pragma clang fp contract(on)
    const int slot[8] = {Q_UG, Q_VG, Q_WG, Q_TG, Q_UP, Q_VP, Q_WP, Q_TP};
    const float tgf = (float)fv[3];
    const float amug = (float)P.as * fsqr(tgf) * frcp(1.0f + (float)P.ts * frcp(tgf));
    const float akg = amug * (float)(P.cpg / P.prg);
    const float amup = 0.001f;
    const float akp = amup * (float)(P.cpp / P.prp);
    const float c23 = (float)P.c23;
    float dx[8], dy[8], dz[8];
    float a11 = 0.0f, a22 = 0.0f, a33 = 0.0f, a12 = 0.0f, a13 = 0.0f, a23 = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const bool use = !(ghost && k >= 5);
        const double px = k < 4 ? row_from(xn, k) : m.xc[m2[k - 4]];
        const double py = k < 4 ? row_from(yn, k) : m.yc[m2[k - 4]];
        const double pz = k < 4 ? row_from(zn, k) : m.zc[m2[k - 4]];
        dx[k] = use ? (float)(px - xf) : 0.0f;
        dy[k] = use ? (float)(py - yf) : 0.0f;
        dz[k] = use ? (float)(pz - zf) : 0.0f;
        a11 += dx[k] * dx[k];
        a22 += dy[k] * dy[k];
        a33 += dz[k] * dz[k];
        a12 += dx[k] * dy[k];
        a13 += dx[k] * dz[k];
        a23 += dy[k] * dz[k];
    }
    const float idn = frcp(det3f(a11, a12, a13, a12, a22, a23, a13, a23, a33));
#pragma unroll
    for (int ph = 0; ph < 2; ph++) {
        const int o = 4 * ph;
        float gr[4][3];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            float b1 = 0.0f, b2 = 0.0f, b3 = 0.0f;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const double qk = k < 4 ? row_from(qn[o + v], k) : QV(slot[o + v], m2[k - 4]);
                const float dv = (float)(qk - fv[o + v]);
                b1 += dv * dx[k];
                b2 += dv * dy[k];
                b3 += dv * dz[k];
            }
            gr[v][0] = det3f(b1, b2, b3, a12, a22, a23, a13, a23, a33) * idn;
            gr[v][1] = det3f(a11, a12, a13, b1, b2, b3, a13, a23, a33) * idn;
            gr[v][2] = det3f(a11, a12, a13, a12, a22, a23, b1, b2, b3) * idn;
        }
        const float mu = ph ? amup : amug, ak = ph ? akp : akg;
        const float uf = (float)fv[o], vf = (float)fv[o + 1], wf = (float)fv[o + 2];
        const float delv = gr[0][0] + gr[1][1] + gr[2][2];
        const float txx = mu * (2.0f * gr[0][0] - c23 * delv);
        const float tyy = mu * (2.0f * gr[1][1] - c23 * delv);
        const float tzz = mu * (2.0f * gr[2][2] - c23 * delv);
        const float txy = mu * (gr[1][0] + gr[0][1]);
        const float txz = mu * (gr[2][0] + gr[0][2]);
        const float tyz = mu * (gr[1][2] + gr[2][1]);
        const float qxx = uf * txx + vf * txy + wf * txz + ak * gr[3][0];
        const float qyy = uf * txy + vf * tyy + wf * tyz + ak * gr[3][1];
        const float qzz = uf * txz + vf * tyz + wf * tzz + ak * gr[3][2];
        const float phf = ph ? phipf : phigf;
        acc[o + 0] = phf * (txx * scx + txy * scy + txz * scz);
        acc[o + 1] = phf * (txy * scx + tyy * scy + tyz * scz);
        acc[o + 2] = phf * (txz * scx + tyz * scy + tzz * scz);
        acc[o + 3] = phf * (qxx * scx + qyy * scy + qzz * scz);
    }
}

// viscous face flux (sr_viscousgp.F90:450-1139) in FP32, one face per lane
__device__ __forceinline__ void viscous_face_f32(const Params &P, const State &s, const Mesh &m, int i, int f,
                                                 int qbase)
{
    const int ne = m.ne;
    if (m.iblank[i] != 1) return;
    const double *q = s.q;
    const size_t nt = s.nt;
    const int slot[8] = {Q_UG, Q_VG, Q_WG, Q_TG, Q_UP, Q_VP, Q_WP, Q_TP};
    float acc[8];
    {
        const float scx = (float)m.sc[(size_t)(3 * f) * ne + i], scy = (float)m.sc[(size_t)(3 * f + 1) * ne + i],
                    scz = (float)m.sc[(size_t)(3 * f + 2) * ne + i];
        const double xf = m.fc3[(size_t)(3 * f) * ne + i], yf = m.fc3[(size_t)(3 * f + 1) * ne + i],
                     zf = m.fc3[(size_t)(3 * f + 2) * ne + i];
        const int neigh = m.nc[(size_t)f * ne + i];
        double qn[8];
        for (int v = 0; v < 8; v++) qn[v] = QV(slot[v], neigh);
        const double xn = m.xc[neigh], yn = m.yc[neigh], zn = m.zc[neigh];
        double vq[(Q_N + 3) / 4];
        quad_load<Q_N>(q, nt, i, f, vq);
        double fv[8];
        fv[0] = (row_plane<Q_UG>(vq) + qn[0]) * 0.5;
        fv[1] = (row_plane<Q_VG>(vq) + qn[1]) * 0.5;
        fv[2] = (row_plane<Q_WG>(vq) + qn[2]) * 0.5;
        fv[3] = (row_plane<Q_TG>(vq) + qn[3]) * 0.5;
        fv[4] = (row_plane<Q_UP>(vq) + qn[4]) * 0.5;
        fv[5] = (row_plane<Q_VP>(vq) + qn[5]) * 0.5;
        fv[6] = (row_plane<Q_WP>(vq) + qn[6]) * 0.5;
        fv[7] = (row_plane<Q_TP>(vq) + qn[7]) * 0.5;
        const float phigf = (float)((row_plane<Q_PHIG>(vq) + QV(Q_PHIG, neigh)) * 0.5);
        const float phipf = (float)((row_plane<Q_PHIP>(vq) + QV(Q_PHIP, neigh)) * 0.5);
        const bool ghost = neigh >= ne;
        const int far = ghost ? i : neigh;
        int m2[4];
        for (int k = 0; k < 4; k++) m2[k] = ghost ? i : m.nc[(size_t)k * ne + far];
        viscous_stencil_f32(P, q, nt, m, m2, ghost, xn, yn, zn, qn, xf, yf, zf, scx, scy, scz, fv, phigf,
                            phipf, acc);
    }
    for (int v = 0; v < 8; v++) {
        const double t0 = (double)__shfl(acc[v], qbase, 64), t1 = (double)__shfl(acc[v], qbase + 16, 64),
                     t2 = (double)__shfl(acc[v], qbase + 32, 64), t3 = (double)__shfl(acc[v], qbase + 48, 64);
        // summed in FP64, kept in FP32 (the form's fluxes)
        if ((v & 3) == f) s.visf[(size_t)v * ne + i] = (float)(0.0 + t0 + t1 + t2 + t3);
    }
}

}  // namespace hydra

// ===========================================================================
// Kernels: one launch covers every case of a batch.  Each block walks its
// share of the tile table (tile_span); per-cell kernels run TILE threads, one
// cell each, and the face kernels run 4*TILE threads, four lanes per cell.
// ===========================================================================
namespace hydra {

__device__ __forceinline__ double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = dmax(v, __shfl_xor(v, off, 64));
    return v;
}

// primitive (sr_primitive.F90:18-178) on interior cells; want_dt: also the
// CFL bound per tile (main.F90:541-545)
// KIND 0: primitives; 1: and the CFL bound per tile
template <int MODE, int KIND>
__global__ void __launch_bounds__(4 * TILE) k_eos(Batch B)
{
    const bool want_dt = KIND == 1;
    const int l = (int)(threadIdx.x % TILE);
    TILE_LOOP_W(B.cnt[0], B.n1)
    {
        const TileCase C = tile_case<MODE>(B, B.ct, t, B.n1);
        const State &s = C.S;
        double loc = 1.0e8;
        if (l < C.count)
            eos_cell(C.P, s.cn, s.ne, C.first + l, s.q, s.nt, want_dt ? &loc : nullptr, s.dl);
        if (want_dt) tile_dt_part(B, t, loc, l, C.count);
    }
}

// per case slot (one block each): dtmin over its tiles, dt = cfl*dtmin,
// time += dt, total += 1 (main.F90:520-523, :555-556).  Without a NaN bound
// the minimum does not depend on the order it is taken in; with one the
// tiles are folded in cell order (tile_dt_part): each thread folds a
// contiguous run of tiles, then the first thread folds the runs in order.
__global__ void __launch_bounds__(1024) k_dt_final(Batch B)
{
    const int c = blockIdx.x;
    const int t0 = B.slot_t0[c], t1 = B.slot_t1[c];
    if (t0 >= t1) return;
    int nan_tiles = 0;
    for (int t = t0 + (int)threadIdx.x; t < t1; t += (int)blockDim.x) nan_tiles |= B.part[(size_t)4 * t + 1] != 0.0;
    if (__syncthreads_or(nan_tiles)) {
        __shared__ double run_m[1024];
        __shared__ int run_h[1024];
        const int n = t1 - t0, per = (n + (int)blockDim.x - 1) / (int)blockDim.x;
        const int a = t0 + (int)threadIdx.x * per, e = min(a + per, t1);
        double m = 0.0;
        int h = -1;  // -1: empty run
        for (int t = a; t < e; t++) {
            const double tm = B.part[(size_t)4 * t];
            if (B.part[(size_t)4 * t + 1] != 0.0) { m = tm; h = 1; }
            else if (h < 0) { m = tm; h = 0; }
            else m = dmin(m, tm);
        }
        run_m[threadIdx.x] = m;
        run_h[threadIdx.x] = h;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = 1.0e8;
            for (int k = 0; k < (int)blockDim.x; k++)
                if (run_h[k] == 1) r = run_m[k];
                else if (run_h[k] == 0) r = r < run_m[k] ? r : run_m[k];
            const State &s = B.S[c];
            const double dt = r * B.P[c].cfl;
            s.scal[SC_DTMIN] = r;
            s.scal[SC_DT] = dt;
            s.scal[SC_TIME] = s.scal[SC_TIME] + dt;
            s.itot[0] = s.itot[0] + 1;
        }
        return;
    }
    double v = 1.0e8;
    // eight independent loads in flight per thread (the partials of a
    // 5M-cell case are 78k tiles; one load at a time left this launch
    // waiting on latency)
    const int bd = (int)blockDim.x;
    for (int t = t0 + (int)threadIdx.x; t < t1; t += 8 * bd) {
        double w[8];
#pragma unroll
        for (int j = 0; j < 8; j++) w[j] = (t + j * bd < t1) ? B.part[(size_t)4 * (t + j * bd)] : 1.0e8;
#pragma unroll
        for (int j = 0; j < 8; j++) v = dmin(v, w[j]);
    }
    __shared__ double red[16];
    v = wave_min(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = dmin(r, red[w]);
        const State &s = B.S[c];
        const double dt = r * B.P[c].cfl;
        s.scal[SC_DTMIN] = r;
        s.scal[SC_DT] = dt;
        s.scal[SC_TIME] = s.scal[SC_TIME] + dt;
        s.itot[0] = s.itot[0] + 1;
    }
}

// CFL bound per tile from the current primitives (hp_dt_min)
template <int MODE>
__global__ void __launch_bounds__(TILE) k_dt_from_prims(Batch B)
{
    TILE_LOOP(B.cnt[0], B.n1)
    {
        const TileCase C = tile_case<MODE>(B, B.ct, t, B.n1);
        double loc = 1.0e8;
        if ((int)threadIdx.x < C.count) loc = dt_cell(C.P, C.S, C.first + (int)threadIdx.x);
        tile_dt_part(B, t, loc, (int)threadIdx.x, C.count);
    }
}

template <int MODE>
__global__ void __launch_bounds__(TILE) k_bc(Batch B, int viscous)
{
    TILE_LOOP(B.cnt[1], B.n1g)
    {
        const TileCase C = tile_case<MODE>(B, B.gt, t, B.n1g);
        if ((int)threadIdx.x < C.count) {
            const int g = C.first + (int)threadIdx.x;
            bc_ghost(C.P, C.S, C.M, g, viscous);
            // The fused step alternates two primitive buffers; a ghost's state
            // goes to both, so the relaxing BCs (types 3, 51, 53) always read
            // the ghost's latest value, and the buffer a step ends in holds
            // the last stage's ghosts.
            const size_t nt = C.S.nt, o = (size_t)C.S.ne + g;
            for (int k = 0; k < Q_N; k++) C.S.q2[(size_t)k * nt + o] = C.S.q[(size_t)k * nt + o];
            if (C.S.rec32) {
                // the ghost's FP32-bulk record: its primitives (R32_PRIM order)
                const double *q = C.S.q;
                float4 *dst = reinterpret_cast<float4 *>(C.S.rec32) + (size_t)(R32_N / 4) * o;
                dst[0] = make_float4((float)QV(Q_ROG, o), (float)QV(Q_UG, o), (float)QV(Q_VG, o), (float)QV(Q_WG, o));
                dst[1] = make_float4((float)QV(Q_TG, o), (float)QV(Q_PHIG, o), (float)QV(Q_ROP, o), (float)QV(Q_UP, o));
                dst[2] = make_float4((float)QV(Q_VP, o), (float)QV(Q_WP, o), (float)QV(Q_TP, o), (float)QV(Q_PG, o));
                dst[3] = make_float4((float)QV(Q_PP, o), 0.0f, 0.0f, 0.0f);
            }
        }
    }
}

template <int MODE>
__global__ void __launch_bounds__(4 * TILE) k_smooth1(Batch B)
{
    const int l = (int)(threadIdx.x % TILE);
    TILE_LOOP_W(B.cnt[0], B.n1)
    {
        const TileCase C = tile_case<MODE>(B, B.ct, t, B.n1);
        if (l < C.count) smooth1_cell(C.S, C.M, C.first + l);
    }
}

template <int MODE>
__global__ void __launch_bounds__(4 * TILE) k_smooth2(Batch B)
{
    const int l = (int)(threadIdx.x % TILE);
    TILE_LOOP_W(B.cnt[0], B.n1)
    {
        const TileCase C = tile_case<MODE>(B, B.ct, t, B.n1);
        if (l < C.count) smooth2_cell(C.P, C.S, C.M, C.first + l);
    }
}

// F32: FP32-bulk gradients and limiter (one wave per tile).  FP64: GRAD
// runs two waves per tile, the first computing gradients and limiter, the
// second the tile's interphase source terms, so that neither path's
// registers add to the other's; without GRAD (first order) one wave computes
// the source.
template <int MODE, bool GRAD>
__global__ void __launch_bounds__(GRAD ? 2 * TILE : TILE) k_grad_limiter(Batch B)
{
    TILE_LOOP(B.cnt[0], B.n1)
    {
        const TileCase C = tile_case<MODE>(B, B.ct, t, B.n1);
        const int lane = (int)(threadIdx.x & 63u);
        if (lane < C.count) {
            if (GRAD && threadIdx.x < TILE)
                grad_cell(C.S, C.M, C.first + lane);
            else
                source_cell(C.P, C.S, C.M, C.first + lane);
        }
    }
}

// The FP32-bulk form's gradient pass: records, and HLLC-Batten's source
// terms (without GRAD, first order: the source terms alone).  The
// single-case launch form takes 134 VGPRs (three waves per SIMD); held to
// four waves (128 VGPRs) it runs no faster on the 5.2M-tet case.
template <int MODE, bool GRAD>
__global__ void __launch_bounds__(TILE) k_grad32(Batch B)
{
    TILE_LOOP(B.cnt[0], B.n1)
    {
        const TileCase C = tile_case<MODE>(B, B.ct, t, B.n1);
        const int lane = (int)threadIdx.x;
        if (!GRAD) {
            if (lane < C.count) source_cell_f32(C.P, C.S, C.M, C.first + lane);
        } else {
            // records staged in LDS, then stored as one contiguous run of
            // the tile's count x R32_N floats (consecutive cells' records
            // are adjacent)
            __shared__ float4 stage32[TILE * R32_N / 4];
            if (lane < C.count)
                grad_cell_f32(C.P, C.S, C.M, C.first + lane, reinterpret_cast<float *>(stage32) + R32_N * lane);
            __syncthreads();
            float4 *out = reinterpret_cast<float4 *>(C.S.rec32) + (size_t)(R32_N / 4) * C.first;
            for (int k = lane; k < C.count * (R32_N / 4); k += TILE)
                __builtin_nontemporal_store(*reinterpret_cast<const f32x4 *>(&stage32[k]),
                                            reinterpret_cast<f32x4 *>(&out[k]));
            if (MODE == 1) __syncthreads();  // the next tile reuses the staging
        }
    }
}

// four lanes of a block per cell C.first + l, one face each, in rows
template <int MODE, bool F32>
__global__ void __launch_bounds__(4 * TILE) k_viscous(Batch B)
{
    TILE_LOOP(B.cnt[0], B.n1)
    {
        const TileCase C = tile_case<MODE>(B, B.ct, t, B.n1);
        // row layout: lane c + 16 f of wave w holds face f of cell 16 w + c
        const int l = (int)((threadIdx.x >> 6) * 16 + (threadIdx.x & 15)), f = (int)((threadIdx.x >> 4) & 3),
                  qbase = (int)(threadIdx.x & 15);
        if (l < C.count) {
            if (F32)
                viscous_face_f32(C.P, C.S, C.M, C.first + l, f, qbase);
            else
                viscous_face(C.P, C.S, C.M, C.first + l, f, qbase);
        }
    }
}

// FP32-bulk HLLC-Batten: the flux of every face listed for the tile's
// cells (its boundary faces and those towards higher-numbered cells; at
// most 4 per cell), one face per lane, into the face's entry of fflux
// (FACE32_N floats: the 10 components, 2 pad), which k_cells32 then reads
// for both cells of the face.
// A quad of lanes reads its four faces' records together (quad_record32),
// one side at a time, each side reconstructed as soon as its record is in:
// at 5.0M cells 0.73 ms at 117 VGPRs, four waves per SIMD, against 0.78 ms
// with each lane reading its own faces' records (then held to three blocks
// per CU by a 48 KB LDS reservation, under which the compiler kept more
// record loads in flight).
template <int MODE>
__global__ void __launch_bounds__(4 * TILE) k_face32(Batch B)
{
    TILE_LOOP(B.cnt[0], B.n1)
    {
        const TileCase C = tile_case<MODE>(B, B.ct, t, B.n1);
        const int base = C.S.fbase[C.first];
        const int nf = C.S.fbase[C.first + C.count] - base;
        const int l = (int)threadIdx.x;
        // a quad runs when its first face is listed; lanes past the list
        // redo the last face (nothing stored), so a quad's record reads
        // have all four lanes
        if ((l & ~3) < nf) {
            const int lf = l < nf ? l : nf - 1;
            const int fc = C.S.flist[base + lf], i = fc >> 2;
            const int neigh = C.M.nc[(size_t)(fc & 3) * C.M.ne + i];
            const float3 *gs = reinterpret_cast<const float3 *>(C.S.fgeom) + (size_t)3 * (base + lf);
            const float3 g0 = gs[0], g1 = gs[1], g2 = gs[2];
            const FaceGeo32 g = {g0.x, g0.y, g0.z, g1.x, g1.y, g1.z, g2.x, g2.y, g2.z, 0.0f, 0.0f, 0.0f};
            float F[10];
            if (C.P.myorder == 2) {
                const float4 *rec = reinterpret_cast<const float4 *>(C.S.rec32);
                const int j = l & 3;
                const bool interior = neigh < C.M.ne;
                // the far side's limiter first (its record's last piece), so
                // each side is reconstructed as soon as its record is in
                const float phib = rec[(size_t)(R32_N / 4) * neigh + (R32_PHI >> 2)].z;
                FState32 L, R;
                float phimin;
                {
                    float a[R32_N];
                    quad_record32(rec, i, j, a);
                    phimin = fmn(a[R32_PHI], phib);
                    side_state32(a, interior, phimin, g.xf0, g.yf0, g.zf0, L);
                }
                {
                    float b[R32_N];
                    quad_record32(rec, neigh, j, b);
                    side_state32(b, interior, phimin, g.xfn, g.yfn, g.zfn, R);
                }
                flux_from_states32<2>(C.P, C.M, neigh, g, L, R, F);
            } else {
                face_flux_f32<2>(C.P, C.S, C.M, i, neigh, g, F);
            }
            if (l < nf) {
                f32x4 *dst = reinterpret_cast<f32x4 *>(C.S.fflux) + (size_t)FACE32_N / 4 * (base + l);
                __builtin_nontemporal_store((f32x4){F[0], F[1], F[2], F[3]}, &dst[0]);
                __builtin_nontemporal_store((f32x4){F[4], F[5], F[6], F[7]}, &dst[1]);
                __builtin_nontemporal_store((f32x4){F[8], F[9], 0.0f, 0.0f}, &dst[2]);
            }
        }
    }
}

// the FP32-bulk HLLC-Batten face table's geometry (State::fgeom), once
__global__ void k_face32_geom(Mesh m, State s, int nf)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nf) return;
    const int fc = s.flist[k], i = fc >> 2, f = fc & 3;
    const FaceGeo32 g = face_geo32(m, i, f, m.nc[(size_t)f * m.ne + i]);
    float3 *d = reinterpret_cast<float3 *>(s.fgeom) + (size_t)3 * k;
    d[0] = make_float3(g.scx, g.scy, g.scz);
    d[1] = make_float3(g.xf0, g.yf0, g.zf0);
    d[2] = make_float3(g.xfn, g.yfn, g.zfn);
}

void launch_face32_geometry(const Mesh &m, const State &s, int nf, hipStream_t st)
{
    if (nf > 0) hipLaunchKernelGGL(k_face32_geom, dim3(nblk(nf, 256)), dim3(256), 0, st, m, s, nf);
}

// the FP32-bulk gradient pass's geometry (State::cgeo32), once: each
// neighbour's centroid offset and each face centroid's offset from the
// cell's centroid, differences formed in FP64 and rounded
__global__ void k_grad32_geom(Mesh m, State s)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, ne = m.ne;
    if (i >= ne) return;
    const double xi = m.xc[i], yi = m.yc[i], zi = m.zc[i];
    float *cg = s.cgeo32;
    for (int k = 0; k < 4; k++) {
        const int nb = m.nc[(size_t)k * ne + i];
        cg[(size_t)k * ne + i] = (float)(m.xc[nb] - xi);
        cg[(size_t)(4 + k) * ne + i] = (float)(m.yc[nb] - yi);
        cg[(size_t)(8 + k) * ne + i] = (float)(m.zc[nb] - zi);
        cg[(size_t)(12 + k) * ne + i] = (float)(m.fc0[(size_t)(3 * k) * ne + i] - xi);
        cg[(size_t)(16 + k) * ne + i] = (float)(m.fc0[(size_t)(3 * k + 1) * ne + i] - yi);
        cg[(size_t)(20 + k) * ne + i] = (float)(m.fc0[(size_t)(3 * k + 2) * ne + i] - zi);
    }
}

void launch_grad32_geometry(const Mesh &m, const State &s, hipStream_t st)
{
    if (s.cgeo32 && m.ne > 0) hipLaunchKernelGGL(k_grad32_geom, dim3(nblk(m.ne, 256)), dim3(256), 0, st, m, s);
}

// FP32-bulk HLLC-Batten after k_face32: one thread per cell sums its face
// entries, adds the source terms the gradient pass stored (src32) and runs
// the stage update (cell_update_f32) and,
// in the fused step (feos), the cell's primitive recovery (eos_staged), as
// k_flux does for the four-lanes-per-cell forms
template <int MODE>
__global__ void __launch_bounds__(4 * TILE) k_cells32(Batch B, int stage, int feos)
{
    const int l = (int)(threadIdx.x % TILE);
    TILE_LOOP_W(B.cnt[0], B.n1)
    {
        const TileCase C = tile_case<MODE>(B, B.ct, t, B.n1);
        double d[4] = {0.0, 0.0, 0.0, 0.0};
        if (l < C.count) {
            const int i = C.first + l, ne = C.M.ne;
            double cv[10];
            if (C.M.iblank[i] == 1) {
                cell_update_f32(C.S, C.M, i, stage, feos == 0, cv);
            } else if (feos) {
                for (int k = 0; k < 10; k++) cv[k] = C.S.cn[(size_t)k * ne + i];
            }
            if (feos) eos_staged<1>(C.P, C.S, i, cv, feos == 2, d);
        }
        if (feos == 2) {
            // the tile's monitor maxima (a max does not depend on the order)
            for (int k = 0; k < 4; k++) {
                const double v = wave_max(d[k]);
                if (l == 0) B.part[(size_t)4 * t + k] = v;
            }
        }
    }
}

// The FP32-bulk face loop of k_flux (HLLC-Zien, each lane reading its own
// faces' records) runs fastest held to three blocks per CU by a 48 KB
// static LDS reservation: the compiler then schedules with more registers
// (more gathers in flight per wave).  At 4.8M cells: 2.22 ms with it (150
// VGPRs), 3.01 ms without (109 VGPRs, four waves per SIMD); before the
// fused primitive recovery 2.07 ms against 2.49 unconstrained and 2.84 with
// a reservation allowing four blocks; the same 48 KB requested as dynamic
// LDS at launch, which the compiler does not see, measured 3.46 ms.  The
// FP64 flux is the other way round: held to three or two blocks it slows
// down.  The fused step's staging of a tile's new conserved variables
// (eos_staged) uses the first 10 x TILE doubles of the reservation.
constexpr int FLUX_F32_LDS_CAP = 48 * 1024;
constexpr int FLUX_STAGE_DOUBLES = 10 * TILE;
static_assert(FLUX_STAGE_DOUBLES * 8 <= FLUX_F32_LDS_CAP, "staging fits the FP32 reservation");

// feos (fused step): 0 none; 1 each cell's primitive recovery right after
// its stage update, by the block's first wave from the tile's updated cells
// staged in LDS (eos_staged); 2 as 1 plus cu = cn (last stage)
template <int MODE, int FLUX, bool F32>
__global__ void __launch_bounds__(4 * TILE) k_flux(Batch B, int stage, int feos)
{
    __shared__ double stage_lds[F32 ? FLUX_F32_LDS_CAP / 8 : FLUX_STAGE_DOUBLES];
    if (F32 && stage < 0) stage_lds[threadIdx.x] = 0.0;  // never taken (stage >= 0): keeps the FP32 reservation
    TILE_LOOP(B.cnt[0], B.n1)
    {
        const TileCase C = tile_case<MODE>(B, B.ct, t, B.n1);
        // FP64: rows (lane c + 16 f of wave w holds face f of cell 16 w + c);
        // FP32 bulk: quads (lanes 4 c .. 4 c + 3 of the block hold cell c)
        const int l = F32 ? (int)(threadIdx.x >> 2) : (int)((threadIdx.x >> 6) * 16 + (threadIdx.x & 15));
        const int f = F32 ? (int)(threadIdx.x & 3) : (int)((threadIdx.x >> 4) & 3);
        const int qbase = (int)(threadIdx.x & 63) & ~3;
        if (l < C.count) {
            double cv[3] = {0.0, 0.0, 0.0};
            if (F32)
                flux_face_f32<FLUX>(C.P, C.S, C.M, C.first + l, f, qbase, stage, feos, cv);
            else
                flux_face<FLUX>(C.P, C.S, C.M, C.first + l, f, stage, feos, cv);
            if (feos) {
                stage_lds[f * TILE + l] = cv[0];
                stage_lds[(f + 4) * TILE + l] = cv[1];
                if (f + 8 < 10) stage_lds[(f + 8) * TILE + l] = cv[2];
            }
        }
        if (feos) {
            __syncthreads();
            double d[4] = {0.0, 0.0, 0.0, 0.0};
            if ((int)threadIdx.x < C.count)
                eos_staged(C.P, C.S, C.first + (int)threadIdx.x, stage_lds + threadIdx.x, feos == 2, d);
            if (feos == 2 && threadIdx.x < TILE) {
                // the tile's monitor maxima (a max does not depend on the order)
                for (int k = 0; k < 4; k++) {
                    const double v = wave_max(d[k]);
                    if (threadIdx.x == 0) B.part[(size_t)4 * t + k] = v;
                }
            }
            if (MODE == 1 && t + st_ < e_) __syncthreads();  // the next tile reuses the staging
        }
    }
}

template <int MODE>
__global__ void __launch_bounds__(4 * TILE) k_rk_update(Batch B, int stage)
{
    const int l = (int)(threadIdx.x % TILE);
    TILE_LOOP_W(B.cnt[0], B.n1)
    {
        const TileCase C = tile_case<MODE>(B, B.ct, t, B.n1);
        if (l < C.count) rk_cell(C.S, C.M, C.first + l, stage);
    }
}

// cu = cn; want_diffs: the tile's max|cn-cu| of variables 1..4 into part
template <int MODE>
__global__ void __launch_bounds__(4 * TILE) k_end_step(Batch B, int want_diffs)
{
    const int l = (int)(threadIdx.x % TILE);
    TILE_LOOP_W(B.cnt[0], B.n1)
    {
        const TileCase C = tile_case<MODE>(B, B.ct, t, B.n1);
        double d[4] = {0.0, 0.0, 0.0, 0.0};
        if (l < C.count) end_cell(C.S, C.first + l, d);
        if (want_diffs) {
            for (int k = 0; k < 4; k++) {
                const double v = wave_max(d[k]);
                if (l == 0) B.part[(size_t)4 * t + k] = v;
            }
        }
    }
}

// Health of each case slot's interior cells: flags[slot] |= 1 for a
// non-finite conserved variable or primitive, |= 2 for pg <= 0 or
// pp + pinf <= 0 (the host clears flags first).  An OR of flags: the result
// does not depend on the order the tiles run in.
template <int MODE>
__global__ void __launch_bounds__(TILE) k_health(Batch B, int *flags)
{
    TILE_LOOP(B.cnt[0], B.n1)
    {
        const TileCase C = tile_case<MODE>(B, B.ct, t, B.n1);
        const int slot = (MODE == 2) ? 0 : B.ct[t].cs;
        int bad = 0;
        if ((int)threadIdx.x < C.count) {
            const State &s = C.S;
            const int i = C.first + (int)threadIdx.x, ne = s.ne;
            const size_t nt = s.nt;
            for (int k = 0; k < 10; k++)
                if (!isfinite(s.cn[(size_t)k * ne + i])) bad |= HP_HEALTH_NONFINITE;
            for (int k = 0; k < Q_N; k++)
                if (!isfinite(s.q[(size_t)k * nt + i])) bad |= HP_HEALTH_NONFINITE;
            if (s.q[(size_t)Q_PG * nt + i] <= 0.0 || s.q[(size_t)Q_PP * nt + i] + C.P.pinf <= 0.0)
                bad |= HP_HEALTH_PRESSURE;
        }
        const bool any1 = __any(bad & HP_HEALTH_NONFINITE), any2 = __any(bad & HP_HEALTH_PRESSURE);
        if (threadIdx.x == 0 && (any1 || any2))
            atomicOr(&flags[slot], (any1 ? HP_HEALTH_NONFINITE : 0) | (any2 ? HP_HEALTH_PRESSURE : 0));
    }
}

// MODE of a launch: 2 for a context on its own, 1 for a persistent grid
// (G.persist bit of the family), 0 otherwise
#define PICK(bit, K, ...) \
    (G.single ? K<2, ##__VA_ARGS__> : (G.persist & (bit)) ? K<1, ##__VA_ARGS__> : K<0, ##__VA_ARGS__>)

void launch_eos(const Batch &B, const Grid &G, int want_dt, hipStream_t st)
{
    if (want_dt)
        hipLaunchKernelGGL(PICK(8, k_eos, 1), dim3(G.cellw), dim3(TILE * G.tpb), 0, st, B);
    else
        hipLaunchKernelGGL(PICK(8, k_eos, 0), dim3(G.cellw), dim3(TILE * G.tpb), 0, st, B);
}


void launch_dt_final(const Batch &B, const Grid &G, hipStream_t st)
{
    hipLaunchKernelGGL(k_dt_final, dim3(G.slots), dim3(1024), 0, st, B);
}

void launch_dt_from_prims(const Batch &B, const Grid &G, hipStream_t st)
{
    hipLaunchKernelGGL(PICK(1, k_dt_from_prims), dim3(G.cell), dim3(TILE), 0, st, B);
}

void launch_bc(const Batch &B, const Grid &G, int viscous, hipStream_t st)
{
    hipLaunchKernelGGL(PICK(4, k_bc), dim3(G.ghost), dim3(TILE), 0, st, B, viscous);
}

void launch_smooth(const Batch &B, const Grid &G, hipStream_t st)
{
    hipLaunchKernelGGL(PICK(8, k_smooth1), dim3(G.cellw), dim3(TILE * G.tpb), 0, st, B);
    hipLaunchKernelGGL(PICK(8, k_smooth2), dim3(G.cellw), dim3(TILE * G.tpb), 0, st, B);
}

void launch_viscous(const Batch &B, const Grid &G, hipStream_t st)
{
    if (G.precision == 1)
        hipLaunchKernelGGL(PICK(2, k_viscous, true), dim3(G.quad), dim3(4 * TILE), 0, st, B);
    else
        hipLaunchKernelGGL(PICK(2, k_viscous, false), dim3(G.quad), dim3(4 * TILE), 0, st, B);
}

void launch_grad(const Batch &B, const Grid &G, hipStream_t st)
{
    if (G.precision == 1) {
        // HLLC-Batten's source terms come from this pass (with the gradients
        // in second order); the HLLC-Zien flux forms its source itself
        if (G.myorder == 2)
            hipLaunchKernelGGL(PICK(1, k_grad32, true), dim3(G.cell), dim3(TILE), 0, st, B);
        else if (G.myflux == 2)
            hipLaunchKernelGGL(PICK(1, k_grad32, false), dim3(G.cell), dim3(TILE), 0, st, B);
    } else if (G.myorder == 2) {
        hipLaunchKernelGGL(PICK(1, k_grad_limiter, true), dim3(G.cell), dim3(2 * TILE), 0, st, B);
    } else {
        hipLaunchKernelGGL(PICK(1, k_grad_limiter, false), dim3(G.cell), dim3(TILE), 0, st, B);
    }
}

void launch_flux(const Batch &B, const Grid &G, int stage, int feos, hipStream_t st)
{
    const dim3 g(G.quad), b(4 * TILE);
    if (G.precision == 1) {
        if (G.myflux == 1) hipLaunchKernelGGL(PICK(2, k_flux, 1, true), g, b, 0, st, B, stage, feos);
        else {
            // each face's flux once, then one thread per cell: its sums,
            // source, update and primitive recovery
            hipLaunchKernelGGL(PICK(2, k_face32), g, b, 0, st, B);
            hipLaunchKernelGGL(PICK(8, k_cells32), dim3(G.cellw), dim3(TILE * G.tpb), 0, st, B, stage, feos);
        }
    } else {
        if (G.myflux == 1) hipLaunchKernelGGL(PICK(2, k_flux, 1, false), g, b, 0, st, B, stage, feos);
        else hipLaunchKernelGGL(PICK(2, k_flux, 2, false), g, b, 0, st, B, stage, feos);
    }
}

void launch_rk_update(const Batch &B, const Grid &G, int stage, hipStream_t st)
{
    hipLaunchKernelGGL(PICK(8, k_rk_update), dim3(G.cellw), dim3(TILE * G.tpb), 0, st, B, stage);
}

void launch_end_step(const Batch &B, const Grid &G, int want_diffs, hipStream_t st)
{
    hipLaunchKernelGGL(PICK(8, k_end_step), dim3(G.cellw), dim3(TILE * G.tpb), 0, st, B, want_diffs);
}

// self-check of the shared-divisor division: pairs where divide() and the
// hardware a / b differ in any bit (NaNs compare equal to NaNs)
__global__ void k_check_division(const double *a, const double *b, size_t n, unsigned long long *bad)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
        const double x = divide(a[k], divisor(b[k])), y = a[k] / b[k];
        if (__double_as_longlong(x) != __double_as_longlong(y) && !(isnan(x) && isnan(y))) atomicAdd(bad, 1ull);
    }
}

void launch_check_division(const double *a, const double *b, size_t n, unsigned long long *bad, hipStream_t st)
{
    hipLaunchKernelGGL(k_check_division, dim3(1024), dim3(256), 0, st, a, b, n, bad);
}

void launch_health(const Batch &B, const Grid &G, int *flags, hipStream_t st)
{
    hipLaunchKernelGGL(PICK(1, k_health), dim3(G.cell), dim3(TILE), 0, st, B, flags);
}

#undef PICK

// overset: each receptor's conserved variables from its four donors,
// w1*q1 + w2*q2 + w3*q3 + w4*q4 in that order (receptors are never donors,
// so the reads never see this launch's writes)
__global__ void __launch_bounds__(DONOR_BLOCK) k_donor_gather(Params P, State s, Donors D, int eos)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    double dm[4] = {0.0, 0.0, 0.0, 0.0};
    if (r < D.n) {
        const int4 d = D.ids[r];
        const double2 wa = D.w[2 * r], wb = D.w[2 * r + 1];
        const int i = D.recv[r], ne = s.ne;
        // fused step: the donors' stage update as the flux launch kept it (cnd);
        // their cn already holds the rewrite of their primitive recovery
        const double *src = eos ? s.cnd : s.cn;
        for (int k = 0; k < 10; k++) {
            const double *c = src + (size_t)k * ne;
            s.cn[(size_t)k * ne + i] = wa.x * c[d.x] + wa.y * c[d.y] + wb.x * c[d.z] + wb.y * c[d.w];
        }
        // fused step: the receptor's primitive recovery into s.q (the next
        // stage's buffer), with cu = cn at the end of the step (eos == 2)
        // and the receptor's monitor differences against the old cu
        double cu0[4];
        if (eos == 2)
            for (int k = 0; k < 4; k++) cu0[k] = s.cu[(size_t)k * ne + i];
        if (eos) eos_cell(P, s.cn, ne, i, s.q, s.nt, nullptr, s.dl, eos == 2 ? s.cu : nullptr);
        if (eos == 2)
            for (int k = 0; k < 4; k++) dm[k] = fabs(s.cn[(size_t)k * ne + i] - cu0[k]);
    }
    if (eos == 2) {
        __shared__ double red[4][DONOR_BLOCK / 64];
        for (int k = 0; k < 4; k++) {
            const double v = wave_max(dm[k]);
            if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v;
        }
        __syncthreads();
        if (threadIdx.x < 4) {
            double v = red[threadIdx.x][0];
            for (int w = 1; w < DONOR_BLOCK / 64; w++) v = dmax(v, red[threadIdx.x][w]);
            D.part[(size_t)4 * blockIdx.x + threadIdx.x] = v;
        }
    }
}

void launch_donor_gather(const Params &P, const State &s, const Donors &D, int eos, hipStream_t st)
{
    if (D.n > 0)
        hipLaunchKernelGGL(k_donor_gather, dim3((D.n + DONOR_BLOCK - 1) / DONOR_BLOCK), dim3(DONOR_BLOCK), 0, st,
                           P, s, D, eos);
}

// per case slot (one block each): the monitor maxima of variables 1..4
// over the slot's tiles (and a lone case's receptor blocks)
__global__ void __launch_bounds__(1024) k_monitor_final(Batch B, const double *dpart, int ndpart, double *out)
{
    const int c = blockIdx.x;
    const int t0 = B.slot_t0[c], t1 = B.slot_t1[c];
    if (t0 >= t1) return;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    const int bd = (int)blockDim.x;
    for (int t = t0 + (int)threadIdx.x; t < t1; t += 4 * bd) {
        double4 w[4];
#pragma unroll
        for (int j = 0; j < 4; j++)
            w[j] = (t + j * bd < t1) ? reinterpret_cast<const double4 *>(B.part)[t + j * bd]
                                     : make_double4(0.0, 0.0, 0.0, 0.0);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            v[0] = dmax(v[0], w[j].x); v[1] = dmax(v[1], w[j].y);
            v[2] = dmax(v[2], w[j].z); v[3] = dmax(v[3], w[j].w);
        }
    }
    for (int b = (int)threadIdx.x; b < ndpart; b += (int)blockDim.x)
        for (int k = 0; k < 4; k++) v[k] = dmax(v[k], dpart[(size_t)4 * b + k]);
    __shared__ double red[4][16];
    for (int k = 0; k < 4; k++) {
        const double w = wave_max(v[k]);
        if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double r = red[threadIdx.x][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = dmax(r, red[threadIdx.x][w]);
        double *dst = out ? out : B.S[c].mon;
        dst[threadIdx.x] = r;
    }
}

void launch_monitor_final(const Batch &B, const Grid &G, const Donors *D, double *out, hipStream_t st)
{
    const int nd = (D && D->n > 0) ? (D->n + DONOR_BLOCK - 1) / DONOR_BLOCK : 0;
    hipLaunchKernelGGL(k_monitor_final, dim3(G.slots), dim3(1024), 0, st, B, nd ? D->part : nullptr, nd, out);
}

}  // namespace hydra
