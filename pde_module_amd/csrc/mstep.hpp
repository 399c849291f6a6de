// Shared pieces of the kernels that run several steps (or sweeps) per launch
// on LDS-resident tiles (heat_mstep.hip, jacobi_mstep.hip): the tile's LDS
// array geometry, whole-wave DPP lane shifts and the thread layout.
//
// Thread layout: a band of `lanes` consecutive threads (a power of two, at
// most 32, at least the region's float4 columns) covers one band of up to
// MS_RPB rows, thread c4 of the band the 4 columns 4*c4 .. 4*c4+3; bands never
// straddle a wave, so the DPP shifts bring each chunk its left and right
// neighbours (the band's first and last chunks only border region-edge
// cells, which are never updated).
#pragma once

#include "pde_common.hpp"

namespace pde {

constexpr int MS_PITCH = PDE_HEAT_TILE_PITCH;   // LDS row pitch (cells): region width cap
constexpr int MS_ROWS = PDE_HEAT_TILE_ROWS;     // region height cap
constexpr int MS_NT = 1024;                     // threads per workgroup
constexpr int MS_RPB = 4;                       // rows per thread band (at most)

// Whole-wave shifts by one lane (DPP wave_shr:1 / wave_shl:1): lane l receives
// lane l-1's value (from_left) or lane l+1's (from_right); 0 past the ends.
__device__ __forceinline__ float ms_from_left(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}

__device__ __forceinline__ float ms_from_right(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}

struct ms_f4 {
    float x, y, z, w;
};

// This thread's chunk of a tile region (rh rows, rw columns).
struct MsLayout {
    int ncol4, rpb, c4, lc0, lr0;
    bool active;
};

__device__ __forceinline__ MsLayout ms_layout(int tid, int rh, int rw)
{
    MsLayout L;
    L.ncol4 = (rw + 3) >> 2;
    int lanes = 1;
    while (lanes < L.ncol4) lanes <<= 1;
    const int bands = MS_NT / lanes;
    L.rpb = (rh + bands - 1) / bands;   // <= MS_RPB for regions of at most MS_ROWS rows
    const int band = tid / lanes;
    L.c4 = tid - band * lanes;
    L.lc0 = 4 * L.c4;
    L.lr0 = band * L.rpb;
    L.active = L.c4 < L.ncol4 && L.lr0 < rh && L.rpb <= MS_RPB;
    return L;
}

// float4 chunk c4 of region row lr (rows outside [0, MS_ROWS) read a clamped
// row; the caller masks what it uses)
__device__ __forceinline__ ms_f4 ms_row4(const float *A, int lr, int c4)
{
    const int r = lr < 0 ? 0 : (lr > MS_ROWS - 1 ? MS_ROWS - 1 : lr);
    const int c = c4 < MS_PITCH / 4 ? c4 : MS_PITCH / 4 - 1;
    return reinterpret_cast<const ms_f4 *>(A)[r * (MS_PITCH / 4) + c];
}

__device__ __forceinline__ void ms_store4(float *A, int lr, int c4, const ms_f4 &v)
{
    reinterpret_cast<ms_f4 *>(A)[lr * (MS_PITCH / 4) + c4] = v;
}

__device__ __forceinline__ float ms_el(const ms_f4 &v, int i)
{
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
}

__device__ __forceinline__ void ms_set(ms_f4 &v, int i, float x)
{
    if (i == 0) v.x = x;
    else if (i == 1) v.y = x;
    else if (i == 2) v.z = x;
    else v.w = x;
}

// A tile record the host planner built cannot reach outside the LDS array
// or its grid (checked again here so a malformed record only skips a tile).
__device__ __forceinline__ bool ms_tile_ok(const pde_heat_tile &T, const pde_batch_grid &G)
{
    return T.rh >= 1 && T.rh <= MS_ROWS && T.rw >= 1 && T.rw <= MS_PITCH && T.r0 >= 0 &&
           T.c0 >= 0 && T.r0 + T.rh <= G.p1 && T.c0 + T.rw <= G.p2 &&
           T.op_end - T.op_beg <= PDE_HEAT_TILE_MAXOPS && T.op_end >= T.op_beg &&
           T.o_r0 >= T.r0 && T.o_r1 <= T.r0 + T.rh && T.o_c0 >= T.c0 && T.o_c1 <= T.c0 + T.rw;
}

}  // namespace pde
