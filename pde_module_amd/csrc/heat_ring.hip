// The heat step with its boundary in one launch (2D scalar fields).
//
// The two-launch form applies the boundary in place (pde_bc_apply_*) and then
// the fused step (pde_heat_step_f32).  Here the launch is split like the
// time-reversed wave pair kernel (wave_reverse.hip): nodes at least `depth`
// from the edge run as row tiles of the plain step (the boundary writes and
// reads no node they read, so they see u2 = u), and the outer `depth` rows
// and columns run as ring pieces that stage themselves grown by two in LDS,
// evaluate every boundary op writing into the piece grown by one from the
// unmodified values (the in-place apply's semantics), write the written
// scalars of their own rectangle back into u, and step their own nodes.  Same
// values as the two launches, bit for bit, including u after the call.
#include "tile2d.hpp"

namespace pde {

constexpr int HR_ROWS = 4;      // rows per lane of the interior tiles
constexpr int HR_BY = 4;        // waves per workgroup
constexpr int HR_CAP = 448;     // nodes of a ring piece grown by two

struct HeatRArgs {
    float *u, *un;
    const pde_bc_op *ops;
    int64_t nops;   // entries whose op slot is not below nops are skipped
    const float *t;
    float wm, wc, wp, alpha, dt;
    pde_ring_plan ring;
};

struct HeatRGeom {
    int64_t P1, P2, g, D;
    uint32_t nstrips, nrt, nint, nring;
};

__device__ __forceinline__ void heat_interior(const HeatRArgs &A, const HeatRGeom &pg,
                                              uint32_t ib)
{
    const uint32_t per = pg.nint >> 3;
    const uint32_t L = (ib & 7u) * per + (ib >> 3);
    if (L >= pg.nstrips * pg.nrt) return;
    const uint32_t rt = L / pg.nstrips, qt = L % pg.nstrips;
    const int lane = threadIdx.x;
    const int64_t D = pg.D, kend = pg.P2 - D, xend = pg.P1 - D;
    const int64_t x0 = D + ((int64_t)rt * HR_BY + threadIdx.y) * HR_ROWS;
    if (x0 >= xend) return;   // the same on every lane of the wave
    Tile2 t;
    t.P1 = pg.P1;
    t.P2 = pg.P2;
    t.g = pg.g;
    t.nq = (t.P2 + 3) / 4;
    t.tq = t.tr = 0;
    int64_t k0 = D + (int64_t)qt * 256 + 4 * lane;
    const bool lane_ok = k0 < kend;
    if (k0 >= t.P2) k0 = t.P2 - 1;   // lanes past the field (never stored)
    v4 rows[HR_ROWS + 2];
#pragma unroll
    for (int r = 0; r < HR_ROWS + 2; ++r) rows[r] = ld_row4(A.u, t, x0 - 1 + r, k0);
    // every row's column edges before any store (u and un may alias as far as
    // the compiler knows, so a load after a store would wait for its turn)
    float left[HR_ROWS], right[HR_ROWS];
#pragma unroll
    for (int r = 0; r < HR_ROWS; ++r)
        col_edges(A.u, t, x0 + r, k0, rows[r + 1], lane, 64, left[r], right[r]);
#pragma unroll
    for (int r = 0; r < HR_ROWS; ++r) {
        const int64_t x = x0 + r;
        if (!lane_ok || x >= xend) continue;
        v4 o;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float acc = acc5(rows[r], rows[r + 1], rows[r + 2], left[r], right[r], i, A.wm,
                                   A.wc, A.wp);
            set_el(o, i, el(rows[r + 1], i) + A.dt * (A.alpha * acc));
        }
        if (k0 + 4 <= kend) {
            st_row4(A.un, t, x, k0, o);
        } else {
            float *p = A.un + x * t.P2 + k0;
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (k0 + i < kend) p[i] = el(o, i);
        }
    }
}

__device__ __forceinline__ void heat_ring(const HeatRArgs &A, const HeatRGeom &pg, int pid,
                                          float *lds)
{
    // the time value and the piece record load together; each thread's first
    // boundary entry loads with the staging of the field and its op before the
    // barrier, so the chain before the boundary evaluation is three loads deep
    const float tv = A.t ? A.t[0] : 0.0f;
    const pde_ring_piece pc = A.ring.pieces[pid];
    const int64_t P1 = pg.P1, P2 = pg.P2, g = pg.g;
    const int tid = threadIdx.y * 64 + threadIdx.x;
    constexpr int NT = 64 * HR_BY;
    const int64_t xI0 = pc.r0 - 2 > 0 ? pc.r0 - 2 : 0, xI1 = pc.r1 + 2 < P1 ? pc.r1 + 2 : P1;
    const int64_t kI0 = pc.c0 - 2 > 0 ? pc.c0 - 2 : 0, kI1 = pc.c1 + 2 < P2 ? pc.c1 + 2 : P2;
    const int wI = (int)(kI1 - kI0), hI = (int)(xI1 - xI0);
    const int nI = wI * hI;
    if (wI <= 0 || hI <= 0 || nI > HR_CAP) return;   // the host never builds such a piece
    float *U = lds, *U2 = lds + HR_CAP;
    // node indices fit in 32 bits (P1 * P2 <= 2^30, checked at launch)
    const uint32_t P2u = (uint32_t)P2;
    auto lidx = [&](int64_t node) -> int {
        const uint32_t n = (uint32_t)node;
        const int64_t x = n / P2u, k = n - (uint32_t)x * P2u;
        if (x < xI0 || x >= xI1 || k < kI0 || k >= kI1) return -1;
        return (int)((x - xI0) * wI + (k - kI0));
    };
    const bool has0 = tid < pc.w_end - pc.w_beg;
    pde_ring_entry w0{};
    if (has0) w0 = A.ring.w[pc.w_beg + tid];
    for (int i = tid; i < nI; i += NT) {
        const int64_t x = xI0 + i / wI, k = kI0 + i % wI;
        const float v = A.u[x * P2 + k];
        U[i] = v;
        U2[i] = v;
    }
    pde_bc_op op0{};
    if (has0) op0 = A.ops[w0.a];
    __syncthreads();
    // the boundary on the piece grown by one, every value from the untouched U
    auto eval = [&](const pde_ring_entry &w, const pde_bc_op &op) {
        const int l = lidx(w.node);
        if (l < 0) return;
        const float v = op_value_tv(op, tv, A.t);
        const int ls = op.kind == PDE_BC_SET ? 0 : lidx(op.src);
        const float in = ls >= 0 ? U[ls] : 0.0f;
        float r;
        switch (op.kind) {
        case PDE_BC_SET: r = v; break;
        case PDE_BC_COPY: r = in; break;
        case PDE_BC_GHOST_D: r = 2.0f * v - in; break;
        default: r = op.coef * v + in; break;
        }
        U2[l] = r;
    };
    if (has0) eval(w0, op0);
    for (int e = pc.w_beg + tid + NT; e < pc.w_end; e += NT) {
        const pde_ring_entry w = A.ring.w[e];
        eval(w, A.ops[w.a]);
    }
    __syncthreads();
    const float wm = A.wm, wc = A.wc, wp = A.wp, dt = A.dt, alpha = A.alpha;
    for (int i = tid; i < nI; i += NT) {
        const int64_t x = xI0 + i / wI, k = kI0 + i % wI;
        if (x < pc.r0 || x >= pc.r1 || k < pc.c0 || k >= pc.c1) continue;
        const float c = U2[i];
        float o = c;
        if (inside(x, g, P1) && inside(k, g, P2)) {
            float acc = 0.0f;
            acc += wm * U2[i - wI];
            acc += wc * c;
            acc += wp * U2[i + wI];
            acc += wm * U2[i - 1];
            acc += wc * c;
            acc += wp * U2[i + 1];
            o = c + dt * (alpha * acc);
        }
        A.un[x * P2 + k] = o;
    }
    // the written scalars of the piece itself back into u
    auto back = [&](const pde_ring_entry &w) {
        const int l = lidx(w.node);
        if (l >= 0 && (w.flags & 2)) A.u[w.node] = U2[l];
    };
    if (has0) back(w0);
    for (int e = pc.w_beg + tid + NT; e < pc.w_end; e += NT) back(A.ring.w[e]);
}

__global__ __launch_bounds__(64 * HR_BY) void heat_ring_kernel(HeatRArgs A, HeatRGeom pg)
{
    __shared__ float lds[2 * HR_CAP];
    if (blockIdx.x >= pg.nint) {
        const int pid = (int)(blockIdx.x - pg.nint);
        if (pid < A.ring.npieces) heat_ring(A, pg, pid, lds);
        return;
    }
    heat_interior(A, pg, blockIdx.x);
}

}  // namespace pde

using namespace pde;

extern "C" int pde_heat_step_ring_f32(float *u, float *un, int64_t nx, int64_t ny, int64_t nz,
                                      int32_t ghost, const pde_bc_op *ops, int64_t nops,
                                      const float *t, float wm, float wc, float wp, float alpha,
                                      float dt, const pde_ring_plan *ring, void *stream)
{
    if (!u || !un || u == un || nops < 0 || (nops > 0 && !ops)) return PDE_EINVAL;
    if (!ring || ring->npieces < 1 || !ring->pieces || !ring->w) return PDE_EINVAL;
    // every plan entry's op slot is below the plan's nops (host-built): the
    // table passed must be the one the plan was built for
    if (ring->nops != nops) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim != 2) return PDE_ESHAPE;
    HeatRGeom pg;
    pg.P1 = G.P[1];
    pg.P2 = G.P[2];
    pg.g = ghost;
    pg.D = ring->depth;
    if (pg.D < 4 || pg.D > 8 || pg.D < ghost + 2) return PDE_EINVAL;
    if (pg.P1 < 2 * pg.D || pg.P2 < 2 * pg.D || pg.P1 * pg.P2 > 0x3fffffff) return PDE_ESHAPE;
    const int64_t ns = (pg.P2 - 2 * pg.D + 255) / 256;
    const int64_t nr = (pg.P1 - 2 * pg.D + HR_BY * HR_ROWS - 1) / (HR_BY * HR_ROWS);
    const int64_t nint = (ns * nr + 7) / 8 * 8;
    const int64_t nring = ((int64_t)ring->npieces + 7) / 8 * 8;
    if (nint + nring > 0x7fffffff) return PDE_ESHAPE;
    pg.nstrips = (uint32_t)ns;
    pg.nrt = (uint32_t)nr;
    pg.nint = (uint32_t)nint;
    pg.nring = (uint32_t)nring;
    HeatRArgs A{u, un, ops, nops, t, wm, wc, wp, alpha, dt, *ring};
    hipLaunchKernelGGL(heat_ring_kernel, dim3((unsigned)(nint + nring)), dim3(64, HR_BY), 0,
                       as_stream(stream), A, pg);
    return launch_status();
}
