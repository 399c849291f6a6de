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
// stays valid as long as its destination does (pde_module_amd/multistep.py).
#include "mstep.hpp"

namespace pde {

constexpr int HT_PITCH = MS_PITCH;
constexpr int HT_ROWS = MS_ROWS;
constexpr int HT_NT = MS_NT;
constexpr int HT_RPB = MS_RPB;
constexpr int HT_OPT = PDE_HEAT_TILE_MAXOPS / HT_NT;   // ops per thread

// 16 bytes of global memory at 4-byte alignment (rows start anywhere)
typedef float ht_f4u __attribute__((ext_vector_type(4), aligned(4)));
typedef float ht_f4a __attribute__((ext_vector_type(4)));   // 16-byte aligned

// The heat update of one cell from its u2 neighbours, in heat_rows' order
// (tile2d.hpp acc5: rows x-1, x, x+1, then columns k-1, k, k+1).
__device__ __forceinline__ float ht_cell(float up, float mid, float dn, float zl, float zr,
                                         float wm, float wc, float wp, float alpha, float dt)
{
    float acc = 0.0f;
    acc += wm * up;
    acc += wc * mid;
    acc += wp * dn;
    acc += wm * zl;
    acc += wc * mid;
    acc += wp * zr;
    return mid + dt * (alpha * acc);
}

// Thread layout and DPP neighbours: mstep.hpp.
__global__ __launch_bounds__(HT_NT) void heat_mstep_kernel(
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
    if (!ms_tile_ok(T, G)) return;
    const int tid = threadIdx.x;
    const int64_t P2 = G.p2;
    const float *ug = u + G.offset;

    // the region, 4 columns per thread and load, every load issued before
    // the first is stored (a row's last chunk may be partial: its cells one
    // by one afterwards)
    constexpr int Q = HT_PITCH / 4;
    constexpr int LPT = HT_ROWS * Q / HT_NT;   // chunks per thread
    {
        ht_f4u v[LPT];
#pragma unroll
        for (int j = 0; j < LPT; ++j) {
            const int c = tid + j * HT_NT, lr = c / Q, lc = 4 * (c - lr * Q);
            if (lr < rh && lc + 4 <= rw)
                v[j] = *reinterpret_cast<const ht_f4u *>(ug + (int64_t)(T.r0 + lr) * P2 + T.c0 + lc);
        }
#pragma unroll
        for (int j = 0; j < LPT; ++j) {
            const int c = tid + j * HT_NT, lr = c / Q, lc = 4 * (c - lr * Q);
            if (lr >= rh || lc >= rw) continue;
            if (lc + 4 <= rw) {
                *reinterpret_cast<ms_f4 *>(&A[lr * HT_PITCH + lc]) =
                    ms_f4{v[j].x, v[j].y, v[j].z, v[j].w};
            } else {
                const float *src = ug + (int64_t)(T.r0 + lr) * P2 + T.c0 + lc;
                for (int i = 0; lc + i < rw; ++i) A[lr * HT_PITCH + lc + i] = src[i];
            }
        }
    }

    // this thread's boundary ops, folded for the sub-steps (one time array per
    // launch): SET r = b; COPY r = in + (-0); GHOST_D r = 2v + (-in);
    // GHOST_N r = coef*v + in -- each the same rounding as op_eval's
    const int nop = T.op_end - T.op_beg;
    const float tv = t ? t[0] : 0.0f;
    int odst[HT_OPT], osrc[HT_OPT];
    float ob[HT_OPT];
    bool oneg[HT_OPT];
#pragma unroll
    for (int j = 0; j < HT_OPT; ++j) {
        const int e = tid + j * HT_NT;
        odst[j] = -1;
        osrc[j] = -1;
        ob[j] = 0.0f;
        oneg[j] = false;
        if (e < nop) {
            const pde_heat_tile_op m = tops[T.op_beg + e];
            const pde_bc_op op = ops[m.op];
            const int cap = rh * HT_PITCH;
            if (m.ldst >= 0 && m.ldst < cap && m.lsrc < cap &&
                (op.kind == PDE_BC_SET || m.lsrc >= 0)) {
                const float v = op_value_tv(op, tv, t);
                odst[j] = m.ldst;
                osrc[j] = op.kind == PDE_BC_SET ? -1 : m.lsrc;
                switch (op.kind) {
                case PDE_BC_SET: ob[j] = v; break;
                case PDE_BC_COPY: ob[j] = -0.0f; break;
                case PDE_BC_GHOST_D: ob[j] = 2.0f * v; oneg[j] = true; break;
                default: ob[j] = op.coef * v; break;
                }
            }
        }
    }

    // this thread's chunk: band, rows, columns and which of its cells step
    const MsLayout L = ms_layout(tid, rh, rw);
    const int rpb = L.rpb, c4 = L.c4, lc0 = L.lc0, lr0 = L.lr0;
    const bool active = L.active;
    const int g = G.ghost;
    bool cm[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int lc = lc0 + i, k = T.c0 + lc;
        cm[i] = lc > 0 && lc < rw - 1 && k >= g && k < G.p2 - g;
    }
    bool rm[HT_RPB];
#pragma unroll
    for (int h = 0; h < HT_RPB; ++h) {
        const int lr = lr0 + h, x = T.r0 + lr;
        rm[h] = active && h < rpb && lr > 0 && lr < rh - 1 && x >= g && x < G.p1 - g;
    }
    const float wm = G.wm, wc = G.wc, wp = G.wp, alpha = G.alpha, dt = G.dt;
    auto row4 = [&](int lr) { return ms_row4(A, lr, c4); };
    __syncthreads();
    for (int s = 0; s < steps; ++s) {
        if (nop > 0) {
            float val[HT_OPT];
#pragma unroll
            for (int j = 0; j < HT_OPT; ++j) {
                const float in = osrc[j] >= 0 ? A[osrc[j]] : 0.0f;
                val[j] = osrc[j] < 0 ? ob[j] : ob[j] + (oneg[j] ? -in : in);
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < HT_OPT; ++j)
                if (odst[j] >= 0) A[odst[j]] = val[j];
            __syncthreads();
        }
        ms_f4 nv[HT_RPB];
        ms_f4 up = row4(lr0 - 1), mid = row4(lr0);
#pragma unroll
        for (int h = 0; h < HT_RPB; ++h) {
            const ms_f4 dn = row4(lr0 + h + 1);
            // every lane shifts (no divergence around the DPP moves)
            const float left = ms_from_left(mid.w), right = ms_from_right(mid.x);
            ms_f4 o = mid;
            if (rm[h]) {
                const float n0 = ht_cell(up.x, mid.x, dn.x, left, mid.y, wm, wc, wp, alpha, dt);
                const float n1 = ht_cell(up.y, mid.y, dn.y, mid.x, mid.z, wm, wc, wp, alpha, dt);
                const float n2 = ht_cell(up.z, mid.z, dn.z, mid.y, mid.w, wm, wc, wp, alpha, dt);
                const float n3 = ht_cell(up.w, mid.w, dn.w, mid.z, right, wm, wc, wp, alpha, dt);
                o.x = cm[0] ? n0 : mid.x;
                o.y = cm[1] ? n1 : mid.y;
                o.z = cm[2] ? n2 : mid.z;
                o.w = cm[3] ? n3 : mid.w;
            }
            nv[h] = o;
            up = mid;
            mid = dn;
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int h = 0; h < HT_RPB; ++h)
                if (h < rpb && lr0 + h < rh) ms_store4(A, lr0 + h, c4, nv[h]);
        }
        __syncthreads();
    }

    float *ung = un + G.offset;
    const int orow0 = T.o_r0 - T.r0, orow1 = T.o_r1 - T.r0;
    const int ocol0 = T.o_c0 - T.c0, ocol1 = T.o_c1 - T.c0;
    if (orow0 < 0 || orow1 > rh || ocol0 < 0 || ocol1 > rw) return;
    // the output rectangle, row by row in 16-byte-aligned chunks of 4
    // columns (streaming stores; unaligned ones would write partial 32-byte
    // sectors twice): chunk 0 of a row is its head up to the first aligned
    // column, the rest aligned chunks, the last one possibly partial
    const int ow = ocol1 - ocol0, oq = 1 + (ow + 3) / 4;
    if (ow <= 0 || orow1 <= orow0) return;
    for (int c = tid; c < (orow1 - orow0) * oq; c += HT_NT) {
        const int r = c / oq, q = c - r * oq, lr = orow0 + r;
        float *row = ung + (int64_t)(T.r0 + lr) * P2 + T.c0;
        const int head = (int)((4 - ((reinterpret_cast<uintptr_t>(row + ocol0) >> 2) & 3)) & 3);
        const int h = head < ow ? head : ow;
        const int lc = q == 0 ? ocol0 : ocol0 + h + 4 * (q - 1);
        const int n = q == 0 ? h : (ocol1 - lc < 4 ? ocol1 - lc : 4);
        if (n <= 0) continue;
        const float *a = &A[lr * HT_PITCH + lc];
        if (n == 4) {
            __builtin_nontemporal_store(ht_f4a{a[0], a[1], a[2], a[3]},
                                        reinterpret_cast<ht_f4a *>(row + lc));
        } else {
            for (int i = 0; i < n; ++i) row[lc + i] = a[i];
        }
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
    hipLaunchKernelGGL(heat_mstep_kernel, dim3((unsigned)ntiles), dim3(HT_NT), 0,
                       as_stream(stream), u, un, grids, ngrids, tiles, tops, ops, t, steps);
    return launch_status();
}
