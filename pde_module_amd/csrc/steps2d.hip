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
constexpr int BY2 = 4;    // waves per block (stacked along rows)

enum Op2 { OP_HEAT = STEP2D_HEAT, OP_WAVE = STEP2D_WAVE, OP_WAVE_ADJ = STEP2D_WAVE_ADJ,
           OP_HEAT_ADJ = STEP2D_HEAT_ADJ };

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
    const int64_t *step; // first = (*step == 0) when set
};

template <int OP>
__global__ __launch_bounds__(64 * BY2) void step2d_kernel(Args2 A, Tile2 t)
{
    uint32_t b = blockIdx.x;
    const uint32_t qt = b % t.tq;
    const uint32_t rt = b / t.tq;
    const int lane = threadIdx.x;
    const int64_t q = (int64_t)qt * 64 + lane;
    const bool lane_ok = q < t.nq;
    const int64_t k0 = (lane_ok ? q : t.nq - 1) * 4;
    const int64_t x0 = ((int64_t)rt * BY2 + threadIdx.y) * RX;
    int first = A.first;
    if (A.step) first = (*A.step == 0);
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
        heat_rows<RX>(A.a, A.o1, t, x0, k0, lane_ok, lane, 64, A.wm, A.wc, A.wp, A.alpha, dt);
        return;
    }

    if (OP == OP_WAVE) {
        v4 rows[RX + 2];
#pragma unroll
        for (int r = 0; r < RX + 2; ++r) rows[r] = ld_row4(A.a, t, x0 - 1 + r, k0);
#pragma unroll
        for (int r = 0; r < RX; ++r) {
            const int64_t x = x0 + r;
            float left, right;
            col_edges(A.a, t, x, k0, rows[r + 1], lane, 64, left, right);
            if (!lane_ok || x >= t.P1) continue;
            const bool xin = inside(x, t.g, t.P1);
            v4 un = rows[r + 1];
            {
                const v4 vo = ld_row4(A.b, t, x, k0);
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
                st_row4(A.o2, t, x, k0, vn);
                st_row4(A.o1, t, x, k0, un);
            }
        }
        return;
    }

    if (OP == OP_HEAT_ADJ) {
        // gu2 = g + dt * (alpha * acc^T(g on interior nodes))
        v4 rows[RX + 2];
#pragma unroll
        for (int r = 0; r < RX + 2; ++r) rows[r] = ld_row4(A.a, t, x0 - 1 + r, k0);
#pragma unroll
        for (int r = 0; r < RX; ++r) {
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
    };
    auto load_row = [&](int64_t x) {
        RowIn R;
        R.gn = ld_row4(A.a, t, x, k0);
        R.gv = ld_row4(A.b, t, x, k0);
        v4 cc = {1.0f, 1.0f, 1.0f, 1.0f};
        if (A.c2) cc = ld_row4(A.c2, t, x, k0);
        const bool xin = inside(x, t.g, t.P1);
        v4 h = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (xin && kin[i]) {
                const float tv = first ? el(R.gv, i) : el(R.gv, i) + dt * el(R.gn, i);
                set_el(h, i, A.c2 ? el(cc, i) * tv : tv);
            }
        R.h = h;
        R.u = A.gc2 ? ld_row4(A.u2, t, x, k0) : v4{0.0f, 0.0f, 0.0f, 0.0f};
        return R;
    };
    // h at a column outside the lane's strip (lanes 0 and 63)
    auto h_at = [&](int64_t x, int64_t kk) {
        if (kk < 0 || kk >= t.P2 || !inside(x, t.g, t.P1) || !inside(kk, t.g, t.P2)) return 0.0f;
        const int64_t e = x * t.P2 + kk;
        const float tv = first ? A.b[e] : A.b[e] + dt * A.a[e];
        return A.c2 ? A.c2[e] * tv : tv;
    };
    v4 hup = load_row(x0 - 1).h;
    v4 uup = A.gc2 ? ld_row4(A.u2, t, x0 - 1, k0) : v4{0.0f, 0.0f, 0.0f, 0.0f};
    RowIn mid = load_row(x0);
#pragma unroll 1
    for (int r = 0; r < RX; ++r) {
        const int64_t x = x0 + r;
        const RowIn dn = load_row(x + 1);
        float hl, hrt, ul = 0.0f, urt = 0.0f;
        col_edges(nullptr, t, -1, k0, mid.h, lane, 64, hl, hrt);
        if (A.gc2) col_edges(A.u2, t, x, k0, mid.u, lane, 64, ul, urt);
        if (lane == 0) hl = h_at(x, k0 - 1);
        if (lane == 63) hrt = h_at(x, k0 + 4);
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
                        const float a =
                            acc5(uup, mid.u, dn.u, ul, urt, i, A.wm, A.wc, A.wp);
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

// Launches the 2D row-tile kernel for op on a 2D scalar field.  Returns
// STEP2D_NA when the layout is not one (the caller then uses its generic
// kernel), else a launch status.
int step2d_launch(int op, const Geom &G, int ncomp, const float *a, const float *b,
                  const float *c2, const float *u2, float *o1, float *o2, float *gc2, float wm,
                  float wc, float wp, float alpha, float dt, int first, const int64_t *step,
                  int ghost, hipStream_t s)
{
    if (G.ndim != 2 || ncomp != 1) return STEP2D_NA;
    Tile2 t;
    t.P1 = G.P[1];
    t.P2 = G.P[2];
    t.g = ghost;
    t.nq = (t.P2 + 3) / 4;
    t.tq = (uint32_t)((t.nq + 63) / 64);
    t.tr = (uint32_t)((t.P1 + BY2 * RX - 1) / (BY2 * RX));
    const int64_t nb = (int64_t)t.tq * t.tr;
    if (nb > 0x7fffffff) return PDE_ESHAPE;
    Args2 A{a, b, c2, u2, o1, o2, gc2, wm, wc, wp, alpha, dt, first, step};
    const dim3 grid((unsigned)nb), block(64, BY2);
    switch (op) {
    case STEP2D_HEAT: hipLaunchKernelGGL(step2d_kernel<OP_HEAT>, grid, block, 0, s, A, t); break;
    case STEP2D_WAVE: hipLaunchKernelGGL(step2d_kernel<OP_WAVE>, grid, block, 0, s, A, t); break;
    case STEP2D_WAVE_ADJ: hipLaunchKernelGGL(step2d_kernel<OP_WAVE_ADJ>, grid, block, 0, s, A, t); break;
    default: hipLaunchKernelGGL(step2d_kernel<OP_HEAT_ADJ>, grid, block, 0, s, A, t); break;
    }
    return launch_status();
}

}  // namespace pde
