// 2D scalar row-tile kernels for the fused time steps and their transposes.
//
// Layout: a 2D scalar field of padded extents (P1 rows, P2 columns), column
// index fastest.  A wave owns a strip of 256 columns (each lane 4 consecutive
// columns, loaded as one 16-byte access) and RX consecutive rows; a block
// stacks BY waves along the rows.  Row neighbours (x +- 1) come from the rows
// the lane already holds (two halo rows per lane are loaded), column
// neighbours from the adjacent lanes by shuffles (lanes 0 and 63 load their
// outer one).  Every lane covers the full padded row range, so ghost nodes are
// handled in the same pass (copied, as the generic kernels do).
//
// Arithmetic is the same, operation for operation, as the generic kernels in
// time_step.hip and adjoint.hip; results are identical bit for bit.
#include "tile2d.hpp"

namespace pde {

constexpr int RX = 4;     // rows per lane
constexpr int RX_ADJ = 4; // rows per lane of the wave transpose
constexpr int BY2 = 4;    // waves per block (stacked along rows)

enum Op2 { OP_HEAT = STEP2D_HEAT, OP_WAVE = STEP2D_WAVE, OP_WAVE_ADJ = STEP2D_WAVE_ADJ,
           OP_HEAT_ADJ = STEP2D_HEAT_ADJ };

constexpr int rows_of(int op) { return op == OP_WAVE_ADJ ? RX_ADJ : RX; }

struct Args2 {
    const float *a;      // HEAT/WAVE: u2; ADJ: gun (HEAT_ADJ: g)
    const float *b;      // WAVE: v;  WAVE_ADJ: gvn
    const float *c2;     // WAVE / WAVE_ADJ: coefficient (may be null)
    const float *u2;     // WAVE_ADJ: forward state for the c2 derivative
    float *o1;           // HEAT/WAVE: u_next; ADJ: gu2
    float *o2;           // WAVE: v_next (may alias b); WAVE_ADJ: gv
    float *gc2;          // WAVE_ADJ: accumulated c2 gradient (may be null)
    float wm, wc, wp, alpha, dt;
    int first;
    const int64_t *step; // first = (*step == 0) when set; n = *step
    pde_point_set src;   // WAVE: v[point] += scale * data[n][slot] before the step
    pde_point_set rec;   // WAVE: data[n][slot] = u_next[point]; WAVE_ADJ: gun += ...
    int64_t *step_next;  // when set, *step_next = n + step_delta (block 0, thread 0)
    int64_t step_delta;
};

template <int OP>
__global__ __launch_bounds__(64 * BY2) void step2d_kernel(Args2 A, Tile2 t)
{
    constexpr int R = rows_of(OP);
    // Tiles are numbered down each column strip (vertically adjacent tiles
    // consecutive) and spread so that blocks b, b + 8, b + 16, ... (one XCD:
    // blocks go round-robin over the 8 XCDs) take consecutive tiles: a tile's
    // halo rows, loaded by its neighbours above and below, are L2 hits.  The
    // grid is padded to a multiple of 8; padding blocks exit.
    const uint32_t ntiles = t.tq * t.tr;
    const uint32_t per = gridDim.x >> 3;
    const uint32_t L = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
    if (L >= ntiles) return;
    const uint32_t qt = L / t.tr;
    const uint32_t rt = L % t.tr;
    const int lane = threadIdx.x;
    const int64_t q = (int64_t)qt * 64 + lane;
    const bool lane_ok = q < t.nq;
    const int64_t k0 = (lane_ok ? q : t.nq - 1) * 4;
    const int64_t x0 = ((int64_t)rt * BY2 + threadIdx.y) * R;
    int first = A.first;
    int64_t n = 0;
    if (A.step) {
        n = *A.step;
        first = (n == 0);
    }
    if (A.step_next && blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0)
        *A.step_next = n + A.step_delta;
    const bool has_src = A.src.row_start != nullptr;
    const bool has_rec = A.rec.row_start != nullptr;
    const float dt = A.dt;

    // interior flags of the lane's 4 columns and their column neighbours
    bool kin[4], lin[4], rin[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t k = k0 + i;
        kin[i] = inside(k, t.g, t.P2);
        lin[i] = inside(k - 1, t.g, t.P2);
        rin[i] = inside(k + 1, t.g, t.P2);
    }

    if (OP == OP_HEAT) {
        heat_rows<R>(A.a, A.o1, t, x0, k0, lane_ok, lane, 64, A.wm, A.wc, A.wp, A.alpha, dt);
        return;
    }

    if (OP == OP_WAVE) {
        v4 rows[R + 2];
#pragma unroll
        for (int r = 0; r < R + 2; ++r) rows[r] = ld_row4(A.a, t, x0 - 1 + r, k0);
        const int64_t kw0 = (int64_t)qt * 256;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t x = x0 + r;
            float left, right;
            col_edges(A.a, t, x, k0, rows[r + 1], lane, 64, left, right);
            if (x >= t.P1) break;                 // the same on every lane
            const bool xin = inside(x, t.g, t.P1);
            v4 un = rows[r + 1];
            v4 vo = ld_row4(A.b, t, x, k0);
            if (has_src) {
                int a, b;
                float none = 0.0f;
                pts_range(A.src, x, t.P1, a, b);
                pts_add(A.src, a, b, k0, n, lane, vo, -1, none);
            }
            v4 cc = {1.0f, 1.0f, 1.0f, 1.0f};
            if (A.c2) cc = ld_row4(A.c2, t, x, k0);
            v4 vn = vo;
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (xin && kin[i]) {
                    const float acc = acc5(rows[r], rows[r + 1], rows[r + 2], left, right, i,
                                           A.wm, A.wc, A.wp);
                    float rhs = A.alpha * acc;
                    if (A.c2) rhs = el(cc, i) * rhs;
                    const float vnew = el(vo, i) + dt * rhs;
                    set_el(vn, i, vnew);
                    set_el(un, i, el(rows[r + 1], i) + dt * (first ? el(vo, i) : vnew));
                }
            if (lane_ok) {
                st_row4(A.o2, t, x, k0, vn);
                st_row4(A.o1, t, x, k0, un);
            }
            if (has_rec) {
                int a, b;
                pts_range(A.rec, x, t.P1, a, b);
                pts_store(A.rec, a, b, kw0, n, lane, un);
            }
        }
        return;
    }

    if (OP == OP_HEAT_ADJ) {
        // gu2 = g + dt * (alpha * acc^T(g on interior nodes))
        v4 rows[R + 2];
#pragma unroll
        for (int r = 0; r < R + 2; ++r) rows[r] = ld_row4(A.a, t, x0 - 1 + r, k0);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t x = x0 + r;
            float left, right;
            col_edges(A.a, t, x, k0, rows[r + 1], lane, 64, left, right);
            if (!lane_ok || x >= t.P1) continue;
            const bool xin = inside(x, t.g, t.P1);
            const bool upin = inside(x - 1, t.g, t.P1), dnin = inside(x + 1, t.g, t.P1);
            v4 o = rows[r + 1];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float acc = acct5(rows[r], rows[r + 1], rows[r + 2], left, right, i, xin,
                                        upin, dnin, kin[i], lin[i], rin[i], A.wm, A.wc, A.wp);
                set_el(o, i, el(rows[r + 1], i) + dt * (A.alpha * acc));
            }
            st_row4(A.o1, t, x, k0, o);
        }
        return;
    }

    // OP_WAVE_ADJ.  h = c2 * gvt on interior nodes (0 elsewhere) with
    // gvt = gvn (+ dt * gun unless first); gu2 = gun + dt * (alpha * acc^T(h));
    // gv = gvt (gvn + dt * gun when first) on interior nodes, gvn on ghosts;
    // gc2 += dt * (gvt * (alpha * acc(u2))) on interior nodes.
    // Rolling window over rows x-1, x, x+1: each input row is loaded once.
    struct RowIn {
        v4 gn, gv, h, u;
        float he;   // lanes 0 and 63: h at the column just outside the strip
    };
    auto load_row = [&](int64_t x) {
        RowIn R_;
        R_.gn = ld_row4(A.a, t, x, k0);
        R_.gv = ld_row4(A.b, t, x, k0);
        v4 cc = {1.0f, 1.0f, 1.0f, 1.0f};
        if (A.c2) cc = ld_row4(A.c2, t, x, k0);
        const bool xin = inside(x, t.g, t.P1);
        // the outer column of the strip's edge lanes (0 outside the interior)
        const int64_t kk = lane == 0 ? k0 - 1 : k0 + 4;
        const bool e_ok = (lane == 0 || lane == 63) && kk >= 0 && kk < t.P2 && xin &&
                          inside(kk, t.g, t.P2);
        float ea = 0.0f, eb = 0.0f, ec = 1.0f;
        if (e_ok) {
            const int64_t e = x * t.P2 + kk;
            ea = A.a[e];
            eb = A.b[e];
            if (A.c2) ec = A.c2[e];
        }
        if (has_rec) {
            int a, b;
            pts_range(A.rec, x, t.P1, a, b);
            pts_add(A.rec, a, b, k0, n, lane, R_.gn, e_ok ? kk : -1, ea);
        }
        v4 h = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (xin && kin[i]) {
                const float tv = first ? el(R_.gv, i) : el(R_.gv, i) + dt * el(R_.gn, i);
                set_el(h, i, A.c2 ? el(cc, i) * tv : tv);
            }
        R_.h = h;
        R_.he = 0.0f;
        if (e_ok) {
            const float tv = first ? eb : eb + dt * ea;
            R_.he = A.c2 ? ec * tv : tv;
        }
        R_.u = A.gc2 ? ld_row4(A.u2, t, x, k0) : v4{0.0f, 0.0f, 0.0f, 0.0f};
        return R_;
    };
    v4 hup = load_row(x0 - 1).h;
    v4 uup = A.gc2 ? ld_row4(A.u2, t, x0 - 1, k0) : v4{0.0f, 0.0f, 0.0f, 0.0f};
    RowIn mid = load_row(x0);
#pragma unroll 1
    for (int r = 0; r < R; ++r) {
        const int64_t x = x0 + r;
        const RowIn dn = load_row(x + 1);
        float hl, hrt, ul = 0.0f, urt = 0.0f;
        col_edges(nullptr, t, -1, k0, mid.h, lane, 64, hl, hrt);
        if (A.gc2) col_edges(A.u2, t, x, k0, mid.u, lane, 64, ul, urt);
        if (lane == 0) hl = mid.he;
        else if (lane == 63) hrt = mid.he;
        if (lane_ok && x < t.P1) {
            const bool xin = inside(x, t.g, t.P1);
            const bool upin = inside(x - 1, t.g, t.P1), dnin = inside(x + 1, t.g, t.P1);
            v4 gu2 = mid.gn, gv = mid.gv;
            v4 gc = {0.0f, 0.0f, 0.0f, 0.0f};
            if (A.gc2) gc = ld_row4(A.gc2, t, x, k0);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float acc = acct5(hup, mid.h, dn.h, hl, hrt, i, xin, upin, dnin, kin[i],
                                        lin[i], rin[i], A.wm, A.wc, A.wp);
                set_el(gu2, i, el(mid.gn, i) + dt * (A.alpha * acc));
                if (xin && kin[i]) {
                    const float tv = first ? el(mid.gv, i) : el(mid.gv, i) + dt * el(mid.gn, i);
                    set_el(gv, i, first ? el(mid.gv, i) + dt * el(mid.gn, i) : tv);
                    if (A.gc2) {
                        const float a = acc5(uup, mid.u, dn.u, ul, urt, i, A.wm, A.wc, A.wp);
                        set_el(gc, i, el(gc, i) + dt * (tv * (A.alpha * a)));
                    }
                }
            }
            st_row4(A.o1, t, x, k0, gu2);
            st_row4(A.o2, t, x, k0, gv);
            if (A.gc2) st_row4(A.gc2, t, x, k0, gc);
        }
        hup = mid.h;
        uup = mid.u;
        mid = dn;
    }
}

static int launch_args(int op, const Geom &G, const Args2 &A, int ghost, hipStream_t s)
{
    Tile2 t;
    t.P1 = G.P[1];
    t.P2 = G.P[2];
    t.g = ghost;
    t.nq = (t.P2 + 3) / 4;
    t.tq = (uint32_t)((t.nq + 63) / 64);
    const int rows = op == STEP2D_WAVE_ADJ ? RX_ADJ : RX;
    t.tr = (uint32_t)((t.P1 + BY2 * rows - 1) / (BY2 * rows));
    const int64_t nb = ((int64_t)t.tq * t.tr + 7) / 8 * 8;
    if (nb > 0x7fffffff) return PDE_ESHAPE;
    const dim3 grid((unsigned)nb), block(64, BY2);
    switch (op) {
    case STEP2D_HEAT: hipLaunchKernelGGL(step2d_kernel<OP_HEAT>, grid, block, 0, s, A, t); break;
    case STEP2D_WAVE: hipLaunchKernelGGL(step2d_kernel<OP_WAVE>, grid, block, 0, s, A, t); break;
    case STEP2D_WAVE_ADJ: hipLaunchKernelGGL(step2d_kernel<OP_WAVE_ADJ>, grid, block, 0, s, A, t); break;
    default: hipLaunchKernelGGL(step2d_kernel<OP_HEAT_ADJ>, grid, block, 0, s, A, t); break;
    }
    return launch_status();
}

// Launches the 2D row-tile kernel for op on a 2D scalar field.  Returns
// STEP2D_NA when the layout is not one (the caller then uses its generic
// kernel), else a launch status.
int step2d_launch(int op, const Geom &G, int ncomp, const float *a, const float *b,
                  const float *c2, const float *u2, float *o1, float *o2, float *gc2, float wm,
                  float wc, float wp, float alpha, float dt, int first, const int64_t *step,
                  int ghost, hipStream_t s)
{
    if (G.ndim != 2 || ncomp != 1) return STEP2D_NA;
    Args2 A{a, b, c2, u2, o1, o2, gc2, wm, wc, wp, alpha, dt, first, step, {}, {}, nullptr, 0};
    return launch_args(op, G, A, ghost, s);
}

}  // namespace pde

using namespace pde;

static bool point_set_ok(const pde_point_set *p)
{
    return !p || (p->row_start && p->col && p->slot && p->data && p->stride >= 0);
}

extern "C" int pde_wave_step_fused_f32(const float *u, const float *v, const float *c2,
                                       float *un, float *vn, int64_t nx, int64_t ny, int64_t nz,
                                       int32_t ghost, float wm, float wc, float wp, float alpha,
                                       float dt, const int64_t *step, int64_t *step_next,
                                       const pde_point_set *sources,
                                       const pde_point_set *receivers, void *stream)
{
    if (!u || !v || !un || !vn || u == un || un == vn || un == v || vn == u) return PDE_EINVAL;
    if (!step || !step_next || step == step_next) return PDE_EINVAL;
    if (!point_set_ok(sources) || !point_set_ok(receivers)) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim != 2) return PDE_ESHAPE;
    Args2 A{u, v, c2, nullptr, un, vn, nullptr, wm, wc, wp, alpha, dt, 0, step,
            sources ? *sources : pde_point_set{}, receivers ? *receivers : pde_point_set{},
            step_next, 1};
    return launch_args(STEP2D_WAVE, G, A, ghost, as_stream(stream));
}

extern "C" int pde_wave_step_adjoint_fused_f32(const float *gun, const float *gvn,
                                               const float *c2, const float *u2, float *gu2,
                                               float *gv, float *gc2, int64_t nx, int64_t ny,
                                               int64_t nz, int32_t ghost, float wm, float wc,
                                               float wp, float alpha, float dt,
                                               const int64_t *step, int64_t *step_next,
                                               const pde_point_set *receivers, void *stream)
{
    if (!gun || !gvn || !gu2 || !gv) return PDE_EINVAL;
    if (gc2 && !u2) return PDE_EINVAL;
    const float *ins[4] = {gun, gvn, c2, u2};
    float *outs[3] = {gu2, gv, gc2};
    for (float *o : outs)
        for (const float *i : ins)
            if (o && i && o == i) return PDE_EINVAL;
    if (gu2 == gv || (gc2 && (gc2 == gu2 || gc2 == gv))) return PDE_EINVAL;
    if (!step || !step_next || step == step_next || !point_set_ok(receivers)) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim != 2) return PDE_ESHAPE;
    Args2 A{gun, gvn, c2, u2, gu2, gv, gc2, wm, wc, wp, alpha, dt, 0, step, pde_point_set{},
            receivers ? *receivers : pde_point_set{}, step_next, -1};
    return launch_args(STEP2D_WAVE_ADJ, G, A, ghost, as_stream(stream));
}
