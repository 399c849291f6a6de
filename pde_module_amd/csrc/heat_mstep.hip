// Several heat steps per launch on LDS-resident tiles, for one grid or a
// batch of independent 2D grids (pde_heat_multistep_f32).
//
// A workgroup owns one tile: a region of at most HT_ROWS x HT_PITCH cells of
// one grid's padded field, loaded once into LDS.  It then runs `steps`
// sub-steps there, each the pair the time loop makes
// (examples/2D_Heat_Equation.py:58-62): the boundary ops whose written scalar
// lies in the region (two-phase: every value from the pre-boundary cells,
// then the stores), then u_next = u2 + dt*(alpha*lap(u2)) on interior nodes
// and u2 elsewhere, with the arithmetic of heat_rows (tile2d.hpp).  Cells on
// a side where the region was cut out of a larger grid lose one valid layer
// per sub-step; the tile's output rectangle lies `steps` cells inside every
// cut side, so after the last sub-step it holds exactly the state the
// one-step launches reach, and is stored to un.  Sides on the grid's edge
// are never cut, so a grid that fits one region is stepped whole.
//
// The host builds the tiles (outputs covering every cell of every grid
// once) and each tile's op list, and checks that every op's source cell
// stays valid as long as its destination does; see ring/multistep planning
// in pde_module_amd/multistep.py.
#include "pde_common.hpp"

namespace pde {

constexpr int HT_PITCH = PDE_HEAT_TILE_PITCH;   // LDS row pitch (cells): region width cap
constexpr int HT_ROWS = PDE_HEAT_TILE_ROWS;     // region height cap
constexpr int HT_TX = 64, HT_TY = 16;           // 1024 threads
constexpr int HT_W = HT_PITCH / HT_TX;          // columns per thread
constexpr int HT_H = HT_ROWS / HT_TY;           // rows per thread (a contiguous band)
constexpr int HT_OPT = PDE_HEAT_TILE_MAXOPS / (HT_TX * HT_TY);   // ops per thread

__global__ __launch_bounds__(HT_TX *HT_TY) void heat_mstep_kernel(
    const float *__restrict__ u, float *__restrict__ un, const pde_batch_grid *__restrict__ grids,
    int64_t ngrids, const pde_heat_tile *__restrict__ tiles,
    const pde_heat_tile_op *__restrict__ tops, const pde_bc_op *__restrict__ ops,
    const float *__restrict__ t, int steps)
{
    __shared__ float A[HT_ROWS * HT_PITCH];
    const pde_heat_tile T = tiles[blockIdx.x];
    if (T.grid < 0 || T.grid >= ngrids) return;
    const pde_batch_grid G = grids[T.grid];
    const int rh = T.rh, rw = T.rw;
    // a malformed record (never produced by the host planner) cannot reach
    // outside the LDS array or the grid
    if (rh < 1 || rh > HT_ROWS || rw < 1 || rw > HT_PITCH || T.r0 < 0 || T.c0 < 0 ||
        T.r0 + rh > G.p1 || T.c0 + rw > G.p2 || T.op_end - T.op_beg > PDE_HEAT_TILE_MAXOPS ||
        T.op_end < T.op_beg)
        return;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * HT_TX + tx;
    const int64_t P2 = G.p2;
    const float *ug = u + G.offset;

    for (int lr = ty; lr < rh; lr += HT_TY) {
        const float *row = ug + (int64_t)(T.r0 + lr) * P2 + T.c0;
        for (int lc = tx; lc < rw; lc += HT_TX) A[lr * HT_PITCH + lc] = row[lc];
    }

    // this thread's boundary ops: value (constant over the sub-steps: one
    // time array per launch), kind, source and destination cells
    const int nop = T.op_end - T.op_beg;
    const float tv = t ? t[0] : 0.0f;
    float ov[HT_OPT];
    int ok[HT_OPT], osrc[HT_OPT], odst[HT_OPT];
    float oc[HT_OPT];
#pragma unroll
    for (int j = 0; j < HT_OPT; ++j) {
        const int e = tid + j * HT_TX * HT_TY;
        odst[j] = -1;
        if (e < nop) {
            const pde_heat_tile_op m = tops[T.op_beg + e];
            const pde_bc_op op = ops[m.op];
            const int cap = rh * HT_PITCH;
            if (m.ldst >= 0 && m.ldst < cap && m.lsrc < cap &&
                (op.kind == PDE_BC_SET || m.lsrc >= 0)) {
                odst[j] = m.ldst;
                osrc[j] = m.lsrc;
                ok[j] = op.kind;
                ov[j] = op_value_tv(op, tv, t);
                oc[j] = op.coef;
            }
        }
    }

    // rows of this thread's band and their position in the grid
    const int lr0 = ty * HT_H;
    const int g = G.ghost, P1 = G.p1;
    const float wm = G.wm, wc = G.wc, wp = G.wp, alpha = G.alpha, dt = G.dt;
    __syncthreads();
    for (int s = 0; s < steps; ++s) {
        if (nop > 0) {
            float val[HT_OPT];
#pragma unroll
            for (int j = 0; j < HT_OPT; ++j) {
                if (odst[j] < 0) continue;
                const float in = ok[j] == PDE_BC_SET ? 0.0f : A[osrc[j]];
                float r;
                switch (ok[j]) {
                case PDE_BC_SET: r = ov[j]; break;
                case PDE_BC_COPY: r = in; break;
                case PDE_BC_GHOST_D: r = 2.0f * ov[j] - in; break;
                default: r = oc[j] * ov[j] + in; break;
                }
                val[j] = r;
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < HT_OPT; ++j)
                if (odst[j] >= 0) A[odst[j]] = val[j];
            __syncthreads();
        }
        // the step on u2, each thread walking its band down every column
        float nv[HT_W][HT_H];
#pragma unroll
        for (int w = 0; w < HT_W; ++w) {
            const int lc = tx + w * HT_TX;
            const bool col_ok = lc < rw;
            const int k = T.c0 + lc;
            const bool kin = lc > 0 && lc < rw - 1 && k >= g && k < G.p2 - g;
            float up = col_ok && lr0 > 0 && lr0 - 1 < rh ? A[(lr0 - 1) * HT_PITCH + lc] : 0.0f;
            float mid = col_ok && lr0 < rh ? A[lr0 * HT_PITCH + lc] : 0.0f;
#pragma unroll
            for (int h = 0; h < HT_H; ++h) {
                const int lr = lr0 + h;
                const float dn = col_ok && lr + 1 < rh ? A[(lr + 1) * HT_PITCH + lc] : 0.0f;
                float o = mid;
                const int x = T.r0 + lr;
                if (kin && lr > 0 && lr < rh - 1 && x >= g && x < P1 - g) {
                    const float zl = A[lr * HT_PITCH + lc - 1];
                    const float zr = A[lr * HT_PITCH + lc + 1];
                    float acc = 0.0f;
                    acc += wm * up;
                    acc += wc * mid;
                    acc += wp * dn;
                    acc += wm * zl;
                    acc += wc * mid;
                    acc += wp * zr;
                    o = mid + dt * (alpha * acc);
                }
                nv[w][h] = o;
                up = mid;
                mid = dn;
            }
        }
        __syncthreads();
#pragma unroll
        for (int w = 0; w < HT_W; ++w) {
            const int lc = tx + w * HT_TX;
#pragma unroll
            for (int h = 0; h < HT_H; ++h) {
                const int lr = lr0 + h;
                if (lc < rw && lr < rh) A[lr * HT_PITCH + lc] = nv[w][h];
            }
        }
        __syncthreads();
    }

    float *ung = un + G.offset;
    const int orow0 = T.o_r0 - T.r0, orow1 = T.o_r1 - T.r0;
    const int ocol0 = T.o_c0 - T.c0, ocol1 = T.o_c1 - T.c0;
    if (orow0 < 0 || orow1 > rh || ocol0 < 0 || ocol1 > rw) return;
    for (int lr = orow0 + ty; lr < orow1; lr += HT_TY) {
        float *row = ung + (int64_t)(T.r0 + lr) * P2 + T.c0;
        for (int lc = ocol0 + tx; lc < ocol1; lc += HT_TX) row[lc] = A[lr * HT_PITCH + lc];
    }
}

}  // namespace pde

using namespace pde;

extern "C" int pde_heat_multistep_f32(const float *u, float *un, const pde_batch_grid *grids,
                                      int64_t ngrids, const pde_heat_tile *tiles, int64_t ntiles,
                                      const pde_heat_tile_op *tops, const pde_bc_op *ops,
                                      const float *t, int32_t steps, void *stream)
{
    if (!u || !un || u == un || !grids || ngrids < 1 || ntiles < 0 || steps < 1) return PDE_EINVAL;
    if (ntiles == 0) return PDE_OK;
    if (!tiles || ntiles > 0x7fffffff || ngrids > 0x7fffffff) return PDE_EINVAL;
    hipLaunchKernelGGL(heat_mstep_kernel, dim3((unsigned)ntiles), dim3(HT_TX, HT_TY), 0,
                       as_stream(stream), u, un, grids, ngrids, tiles, tops, ops, t, steps);
    return launch_status();
}
