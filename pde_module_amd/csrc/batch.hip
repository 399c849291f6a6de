// Batched heat ensembles: many independent 2D scalar grids of uneven sizes
// packed into one buffer and stepped by one launch per phase.
//
// * pde_bc_apply_batch_inplace_f32: one workgroup per grid applies that grid's
//   slice of the packed boundary-op table in place (gather into LDS, barrier,
//   scatter): the same two-phase semantics as pde_bc_apply_inplace_f32, with
//   no grid-wide dependency, so one launch serves every grid.
// * pde_heat_batch_step_f32: a flat list of tiles over all grids; a block
//   takes tiles with a grid-stride loop (persistent when launched with fewer
//   blocks than tiles).  A tile is (grid, first row, first 4-column chunk,
//   segment width): its waves split into row segments of cw lanes (64, 32 or
//   16), so narrow grids keep most lanes busy.  The per-lane arithmetic is
//   heat_rows, shared with the single-grid kernel, so every grid's result is
//   bit-identical to stepping it alone.
#include "tile2d.hpp"

namespace pde {

constexpr int BATCH_BC_CAP = BC_BLOCK_CAP;   // ops per grid staged in LDS
static_assert(sizeof(pde_batch_grid) == 64, "pde_batch_grid layout");

__global__ __launch_bounds__(1024) void bc_batch_inplace_kernel(float *__restrict__ u,
                                                                const pde_batch_grid *grids,
                                                                const pde_bc_op *__restrict__ ops,
                                                                const float *t)
{
    __shared__ float vals[BATCH_BC_CAP];
    const pde_batch_grid G = grids[blockIdx.x];
    const int64_t n = G.op_end - G.op_begin;
    if (n <= 0 || n > BATCH_BC_CAP) return;   // rejected on the host
    bc_inplace_block(u, ops + G.op_begin, (int)n, t, vals);
}

template <int RXT, int BYT>
__global__ __launch_bounds__(64 * BYT) void heat_batch_kernel(const float *__restrict__ u,
                                                              float *__restrict__ un,
                                                              const pde_batch_grid *grids,
                                                              int64_t ngrids,
                                                              const int32_t *__restrict__ tiles,
                                                              int64_t ntiles)
{
    for (int64_t ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const int gi = tiles[4 * ti];
        const int cw = tiles[4 * ti + 3];
        if (gi < 0 || gi >= ngrids || !(cw == 64 || cw == 32 || cw == 16)) continue;
        const pde_batch_grid G = grids[gi];
        Tile2 t;
        t.P1 = G.p1;
        t.P2 = G.p2;
        t.g = G.ghost;
        t.nq = (t.P2 + 3) / 4;
        const int lane = threadIdx.x;
        const int seg = lane / cw, sl = lane - seg * cw;
        const int64_t q = (int64_t)tiles[4 * ti + 2] + sl;
        const bool lane_ok = q < t.nq;
        const int64_t k0 = (lane_ok ? q : t.nq - 1) * 4;
        const int64_t x0 =
            (int64_t)tiles[4 * ti + 1] + ((int64_t)threadIdx.y * (64 / cw) + seg) * RXT;
        heat_rows<RXT>(u + G.offset, un + G.offset, t, x0, k0, lane_ok, sl, cw, G.wm, G.wc, G.wp,
                       G.alpha, G.dt);
    }
}

}  // namespace pde

using namespace pde;

extern "C" int pde_bc_apply_batch_inplace_f32(float *u, const pde_batch_grid *grids,
                                              int64_t ngrids, const pde_bc_op *ops,
                                              const float *t, void *stream)
{
    if (!u || !grids || ngrids < 0) return PDE_EINVAL;
    if (ngrids == 0) return PDE_OK;
    if (!ops || ngrids > 0x7fffffff) return PDE_EINVAL;
    hipLaunchKernelGGL(bc_batch_inplace_kernel, dim3((unsigned)ngrids), dim3(1024), 0,
                       as_stream(stream), u, grids, ops, t);
    return launch_status();
}

extern "C" int pde_batch_bc_capacity(void) { return BATCH_BC_CAP; }

#define PDE_BATCH_CASE(R, B)                                                                    \
    if (rows_per_lane == R && waves == B) {                                                     \
        hipLaunchKernelGGL((heat_batch_kernel<R, B>), dim3(nb), dim3(64, B), 0, s, u, un,       \
                           grids, ngrids, tiles, ntiles);                                       \
        return launch_status();                                                                 \
    }

extern "C" int pde_heat_batch_step_f32(const float *u, float *un, const pde_batch_grid *grids,
                                       int64_t ngrids, const int32_t *tiles, int64_t ntiles,
                                       int32_t rows_per_lane, int32_t waves, int64_t blocks,
                                       void *stream)
{
    if (!u || !un || u == un || !grids || ngrids < 0 || ntiles < 0 || blocks < 0)
        return PDE_EINVAL;
    if (ntiles == 0) return PDE_OK;
    if (!tiles) return PDE_EINVAL;
    const int64_t b = blocks == 0 || blocks > ntiles ? ntiles : blocks;
    if (b > 0x7fffffff) return PDE_ESHAPE;
    const unsigned nb = (unsigned)b;
    hipStream_t s = as_stream(stream);
    PDE_BATCH_CASE(1, 1) PDE_BATCH_CASE(1, 2) PDE_BATCH_CASE(1, 4)
    PDE_BATCH_CASE(2, 1) PDE_BATCH_CASE(2, 2) PDE_BATCH_CASE(2, 4)
    PDE_BATCH_CASE(4, 1) PDE_BATCH_CASE(4, 2) PDE_BATCH_CASE(4, 4)
    PDE_BATCH_CASE(8, 1) PDE_BATCH_CASE(8, 2) PDE_BATCH_CASE(8, 4)
    return PDE_EINVAL;
}
