// Shared host/device helpers for libpdehip.so (gfx950).
//
// Canonical geometry: axes of size 1 are not grid axes, and because their
// extent is 1 dropping them leaves every flat index unchanged.  The host
// therefore folds (nx, ny, nz) into up to three *eligible* extents P[0..2],
// right-aligned (P[0] = 1 for a 2D grid, P[0] = P[1] = 1 for 1D), so one
// kernel family covers 1D/2D/3D: canonical axis 2 is always the fastest
// (node stride 1), axis 1 has node stride P[2], axis 0 node stride P[1]*P[2].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pde_hip.h"

namespace pde {

struct Geom {
    int ndim;       // number of eligible axes (0..3)
    int64_t P[3];   // canonical padded extents, right-aligned
    int64_t E[3];   // interior extents (P - 2g on eligible axes, 1 otherwise)
    int64_t G[3];   // interior start (g on eligible axes, 0 otherwise)
    int64_t S[3];   // node strides (S[2] = 1)
    int64_t nodes;  // P[0]*P[1]*P[2]
};

// Returns PDE_OK, PDE_EINVAL or PDE_ESHAPE.
inline int make_geom(int64_t nx, int64_t ny, int64_t nz, int g, Geom *G)
{
    if (nx < 1 || ny < 1 || nz < 1 || g < 0) return PDE_EINVAL;
    int64_t in[3] = {nx, ny, nz};
    int64_t el[3];
    int nd = 0;
    for (int a = 0; a < 3; ++a)
        if (in[a] > 1) el[nd++] = in[a];
    G->ndim = nd;
    for (int a = 0; a < 3; ++a) {
        int src = a - (3 - nd);
        G->P[a] = src >= 0 ? el[src] : 1;
        bool e = src >= 0;
        G->G[a] = e ? g : 0;
        G->E[a] = e ? G->P[a] - 2 * (int64_t)g : 1;
        if (e && (g < 1 || G->E[a] < 1)) return PDE_ESHAPE;
    }
    G->S[2] = 1;
    G->S[1] = G->P[2];
    G->S[0] = G->P[1] * G->P[2];
    G->nodes = G->P[0] * G->P[1] * G->P[2];
    return PDE_OK;
}

inline int launch_status()
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PDE_OK : (int)e;
}

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned clamp_grid(int64_t n, int64_t cap)
{
    return (unsigned)(n < 1 ? 1 : (n > cap ? cap : n));
}

// Blocks are dispatched round-robin over the 8 XCDs of an MI355X, each with
// its own L2.  Remapping the hardware block id so that runs of consecutive
// logical tiles land on one XCD keeps tiles that share halo rows in the same
// L2.  Identity when the grid is not a multiple of 8.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nblocks)
{
    if (nblocks & 7u) return b;
    const uint32_t per = nblocks >> 3;
    return (b & 7u) * per + (b >> 3);
}

// Value of one resolved boundary op (PDE_BC_* kinds) read from `in`; t is the
// device time scalar of ramp values (null = 0).
// The op's boundary value with the time tv = t[0] already loaded (0 for a
// null t, the reference's default time array zeros(1),
// functional_boundary.py:237).
__device__ __forceinline__ float op_value_tv(const pde_bc_op &op, float tv, const float *t)
{
    if (op.vkind == PDE_BCV_SINE) return op.value * sinf(op.tmax * tv);
    if (op.vkind == PDE_BCV_ARRAY) return t ? t[1 + op.slot] : 0.0f;
    if (op.tmax > 0.0f) {
        // wp.min(t / T_max, 1) * const  (boundary_functions.py:85-89)
        float ramp = tv / op.tmax;
        ramp = ramp < 1.0f ? ramp : 1.0f;
        return ramp * op.value;
    }
    return op.value;
}

__device__ __forceinline__ float op_value(const pde_bc_op &op, const float *t)
{
    return op_value_tv(op, t ? t[0] : 0.0f, t);
}

__device__ __forceinline__ float op_eval(const pde_bc_op &op, const float *__restrict__ in,
                                         const float *t)
{
    const float v = op_value(op, t);
    switch (op.kind) {
    case PDE_BC_SET: return v;
    case PDE_BC_COPY: return in[op.src];
    case PDE_BC_GHOST_D: return 2.0f * v - in[op.src];
    default: return op.coef * v + in[op.src];
    }
}

// In-place boundary application by one workgroup of 1024 threads: every op's
// value is computed from the untouched field into LDS (vals, n <= its
// capacity), then the workgroup barrier, then the stores.  Each thread's ops
// (tid, tid + 1024, ...) are loaded BC_UNROLL at a time before any field value,
// so a pass costs two memory round trips, not two per op.
constexpr int BC_BLOCK_CAP = 16384;   // ops (64 KiB of LDS)
constexpr int BC_CHUNK = 256;         // ops per conflict-closed chunk (one workgroup)
constexpr int BC_UNROLL = 8;

__device__ __forceinline__ void bc_inplace_block(float *__restrict__ u,
                                                 const pde_bc_op *__restrict__ ops, int n,
                                                 const float *t, float *vals)
{
    for (int base = 0; base < n; base += 1024 * BC_UNROLL) {
        pde_bc_op op[BC_UNROLL];
#pragma unroll
        for (int j = 0; j < BC_UNROLL; ++j) {
            const int m = base + j * 1024 + (int)threadIdx.x;
            if (m < n) op[j] = ops[m];
        }
#pragma unroll
        for (int j = 0; j < BC_UNROLL; ++j) {
            const int m = base + j * 1024 + (int)threadIdx.x;
            if (m < n) vals[m] = op_eval(op[j], u, t);
        }
    }
    __syncthreads();
    for (int base = 0; base < n; base += 1024 * BC_UNROLL) {
        int64_t dst[BC_UNROLL];
#pragma unroll
        for (int j = 0; j < BC_UNROLL; ++j) {
            const int m = base + j * 1024 + (int)threadIdx.x;
            if (m < n) dst[j] = ops[m].dst;
        }
#pragma unroll
        for (int j = 0; j < BC_UNROLL; ++j) {
            const int m = base + j * 1024 + (int)threadIdx.x;
            if (m < n) u[dst[j]] = vals[m];
        }
    }
}

// 2D scalar row-tile kernels of the fused steps and their transposes
// (steps2d.hip).  a/b/c2/u2 are inputs, o1/o2/gc2 outputs:
//   HEAT:     a = u2, o1 = u_next
//   WAVE:     a = u2, b = v, c2, o1 = u_next, o2 = v_next
//   HEAT_ADJ: a = g_next, o1 = g_u2
//   WAVE_ADJ: a = g_un, b = g_vn, c2, u2, o1 = g_u2, o2 = g_v, gc2 (+=)
enum { STEP2D_HEAT = 0, STEP2D_WAVE = 1, STEP2D_WAVE_ADJ = 2, STEP2D_HEAT_ADJ = 3 };
constexpr int STEP2D_NA = -100;
int step2d_launch(int op, const Geom &G, int ncomp, const float *a, const float *b,
                  const float *c2, const float *u2, float *o1, float *o2, float *gc2, float wm,
                  float wc, float wp, float alpha, float dt, int first, const int64_t *step,
                  int ghost, hipStream_t s);

}  // namespace pde
