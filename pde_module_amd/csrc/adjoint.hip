// Adjoint (transpose) kernels for reverse-mode differentiation.
//
// Each operator is a linear map from the whole padded field to its interior
// nodes (output ghosts are not part of the range).  Its transpose is written
// as a gather: for every padded node m, over axes ascending and s = -1, 0, +1,
// the term w_s * g[m - s e_a] of every interior node m - s e_a whose stencil
// reads m, accumulated from 0 and scaled as the forward scales.  Ghost values
// of the incoming gradient are never read.  Built with -ffp-contract=off like
// the forward kernels, so a serial restatement in the same order reproduces
// every value bit for bit.
#include "pde_common.hpp"

namespace pde {

// Padded-row geometry shared by the transpose kernels: a thread owns one
// float of a padded row (P2 * comps floats), rows indexed by (i0, i1).
struct AdjGeom {
    int64_t S0, S1, C;   // float strides of canonical axes 0 and 1; floats per node
    int64_t P0, P1, R;   // padded planes, rows, floats per row
    int64_t lo[3], hi[3];  // interior node ranges per canonical axis
};

inline AdjGeom adj_geom(const Geom &G, int64_t C)
{
    AdjGeom a;
    a.S0 = G.S[0] * C;
    a.S1 = G.S[1] * C;
    a.C = C;
    a.P0 = G.P[0];
    a.P1 = G.P[1];
    a.R = G.P[2] * C;
    for (int d = 0; d < 3; ++d) {
        a.lo[d] = G.G[d];
        a.hi[d] = G.G[d] + G.E[d];
    }
    return a;
}

__device__ __forceinline__ bool in_range(int64_t v, int64_t lo, int64_t hi) { return v >= lo && v < hi; }

// acc of the transposed 3-point-per-axis stencil at padded float e with node
// coordinates (i0, i1, k); `fetch(e2)` returns the incoming value at float e2
// (only called for interior nodes).  s = -1 reads m + e, s = +1 reads m - e.
template <int NDIM, typename F>
__device__ __forceinline__ float adj_acc(const AdjGeom &g, int64_t e, int64_t i0, int64_t i1,
                                         int64_t k, float wm, float wc, float wp, F fetch)
{
    const bool in0 = in_range(i0, g.lo[0], g.hi[0]);
    const bool in1 = in_range(i1, g.lo[1], g.hi[1]);
    const bool in2 = in_range(k, g.lo[2], g.hi[2]);
    float acc = 0.0f;
    if (NDIM >= 3 && in1 && in2) {
        if (in_range(i0 + 1, g.lo[0], g.hi[0])) acc += wm * fetch(e + g.S0);
        if (in0) acc += wc * fetch(e);
        if (in_range(i0 - 1, g.lo[0], g.hi[0])) acc += wp * fetch(e - g.S0);
    }
    if (NDIM >= 2 && in0 && in2) {
        if (in_range(i1 + 1, g.lo[1], g.hi[1])) acc += wm * fetch(e + g.S1);
        if (in1) acc += wc * fetch(e);
        if (in_range(i1 - 1, g.lo[1], g.hi[1])) acc += wp * fetch(e - g.S1);
    }
    if (NDIM >= 1 && in0 && in1) {
        if (in_range(k + 1, g.lo[2], g.hi[2])) acc += wm * fetch(e + g.C);
        if (in2) acc += wc * fetch(e);
        if (in_range(k - 1, g.lo[2], g.hi[2])) acc += wp * fetch(e - g.C);
    }
    return acc;
}

#define PDE_ADJ_LOOP_BEGIN                                                        \
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;                  \
    if (r >= g.R) return;                                                       \
    const int64_t k = r / g.C;                                                  \
    for (int64_t i0 = blockIdx.z; i0 < g.P0; i0 += gridDim.z)                   \
        for (int64_t i1 = blockIdx.y; i1 < g.P1; i1 += gridDim.y) {             \
            const int64_t e = i0 * g.S0 + i1 * g.S1 + r;
#define PDE_ADJ_LOOP_END }

// Laplacian transpose: gin[m] = alpha * acc^T(gout).
template <int NDIM>
__global__ __launch_bounds__(256) void lap_adj_kernel(const float *__restrict__ gout,
                                                      float *__restrict__ gin, AdjGeom g,
                                                      float wm, float wc, float wp, float alpha)
{
    PDE_ADJ_LOOP_BEGIN
    const float acc = adj_acc<NDIM>(g, e, i0, i1, k, wm, wc, wp,
                                    [&](int64_t x) { return gout[x]; });
    gin[e] = alpha * acc;
    PDE_ADJ_LOOP_END
}

// Grad transpose.  Geometry over the *input* field (comps = N): thread per
// (padded node, r); gout holds N x D per node.  gin[m][r] = alpha * sum over
// axes t (ascending) and s of w_s * gout[m - s e_t][r][t].
template <int NDIM>
__global__ __launch_bounds__(256) void grad_adj_kernel(const float *__restrict__ gout,
                                                       float *__restrict__ gin, AdjGeom g,
                                                       float wm, float wc, float wp, float alpha)
{
    PDE_ADJ_LOOP_BEGIN
    const bool in0 = in_range(i0, g.lo[0], g.hi[0]);
    const bool in1 = in_range(i1, g.lo[1], g.hi[1]);
    const bool in2 = in_range(k, g.lo[2], g.hi[2]);
    const float *go = gout;
    float acc = 0.0f;
    int t = 0;
    if (NDIM >= 3) {
        if (in1 && in2) {
            if (in_range(i0 + 1, g.lo[0], g.hi[0])) acc += wm * go[(e + g.S0) * NDIM + t];
            if (in0) acc += wc * go[e * NDIM + t];
            if (in_range(i0 - 1, g.lo[0], g.hi[0])) acc += wp * go[(e - g.S0) * NDIM + t];
        }
        ++t;
    }
    if (NDIM >= 2) {
        if (in0 && in2) {
            if (in_range(i1 + 1, g.lo[1], g.hi[1])) acc += wm * go[(e + g.S1) * NDIM + t];
            if (in1) acc += wc * go[e * NDIM + t];
            if (in_range(i1 - 1, g.lo[1], g.hi[1])) acc += wp * go[(e - g.S1) * NDIM + t];
        }
        ++t;
    }
    if (in0 && in1) {
        if (in_range(k + 1, g.lo[2], g.hi[2])) acc += wm * go[(e + g.C) * NDIM + t];
        if (in2) acc += wc * go[e * NDIM + t];
        if (in_range(k - 1, g.lo[2], g.hi[2])) acc += wp * go[(e - g.C) * NDIM + t];
    }
    gin[e] = alpha * acc;
    PDE_ADJ_LOOP_END
}

// Divergence transpose.  Geometry over the output field (comps = rows):
// thread per (padded node, r); writes gin[m][r][t] = alpha * sum_s w_s *
// gout[m - s e_t][r] for each axis t.
template <int NDIM>
__global__ __launch_bounds__(256) void div_adj_kernel(const float *__restrict__ gout,
                                                      float *__restrict__ gin, AdjGeom g,
                                                      float wm, float wc, float wp, float alpha)
{
    PDE_ADJ_LOOP_BEGIN
    const bool in0 = in_range(i0, g.lo[0], g.hi[0]);
    const bool in1 = in_range(i1, g.lo[1], g.hi[1]);
    const bool in2 = in_range(k, g.lo[2], g.hi[2]);
    float *o = gin + e * NDIM;
    int t = 0;
    if (NDIM >= 3) {
        float acc = 0.0f;
        if (in1 && in2) {
            if (in_range(i0 + 1, g.lo[0], g.hi[0])) acc += wm * gout[e + g.S0];
            if (in0) acc += wc * gout[e];
            if (in_range(i0 - 1, g.lo[0], g.hi[0])) acc += wp * gout[e - g.S0];
        }
        o[t++] = alpha * acc;
    }
    if (NDIM >= 2) {
        float acc = 0.0f;
        if (in0 && in2) {
            if (in_range(i1 + 1, g.lo[1], g.hi[1])) acc += wm * gout[e + g.S1];
            if (in1) acc += wc * gout[e];
            if (in_range(i1 - 1, g.lo[1], g.hi[1])) acc += wp * gout[e - g.S1];
        }
        o[t++] = alpha * acc;
    }
    {
        float acc = 0.0f;
        if (in0 && in1) {
            if (in_range(k + 1, g.lo[2], g.hi[2])) acc += wm * gout[e + g.C];
            if (in2) acc += wc * gout[e];
            if (in_range(k - 1, g.lo[2], g.hi[2])) acc += wp * gout[e - g.C];
        }
        o[t] = alpha * acc;
    }
    PDE_ADJ_LOOP_END
}

// Heat step transpose: g_u2 = g + dt * (alpha * acc^T(g)).
template <int NDIM>
__global__ __launch_bounds__(256) void heat_adj_kernel(const float *__restrict__ gn,
                                                       float *__restrict__ gu2, AdjGeom g,
                                                       float wm, float wc, float wp, float alpha,
                                                       float dt)
{
    PDE_ADJ_LOOP_BEGIN
    const float acc = adj_acc<NDIM>(g, e, i0, i1, k, wm, wc, wp,
                                    [&](int64_t x) { return gn[x]; });
    gu2[e] = gn[e] + dt * (alpha * acc);
    PDE_ADJ_LOOP_END
}

// u/v wave step transpose (scalar field).  Forward, on interior nodes:
//   vn = v + dt * (c2 * (alpha * acc(u2)));  un = u2 + dt * (first ? v : vn)
// and vn = v, un = u2 on ghosts.  With gvt = gvn (+ dt * gun unless first) on
// interior nodes:
//   gv  = gvt, or gvn + dt * gun when first (interior); gvn on ghosts
//   gu2 = gun + dt * (alpha * acc^T(h)),  h = c2 * gvt on interior nodes
//   gc2 += dt * (gvt * (alpha * acc(u2)))  on interior nodes (if gc2 != NULL)
template <int NDIM>
__global__ __launch_bounds__(256) void wave_adj_kernel(
    const float *__restrict__ gun, const float *__restrict__ gvn, const float *__restrict__ c2,
    const float *__restrict__ u2, float *__restrict__ gu2, float *__restrict__ gv,
    float *__restrict__ gc2, AdjGeom g, float wm, float wc, float wp, float alpha, float dt,
    int first, const int64_t *__restrict__ step)
{
    if (step) first = (*step == 0);
    PDE_ADJ_LOOP_BEGIN
    auto gvt_at = [&](int64_t x) {
        return first ? gvn[x] : gvn[x] + dt * gun[x];
    };
    auto h_at = [&](int64_t x) {
        const float t = gvt_at(x);
        return c2 ? c2[x] * t : t;
    };
    const bool inner = in_range(i0, g.lo[0], g.hi[0]) && in_range(i1, g.lo[1], g.hi[1]) &&
                       in_range(k, g.lo[2], g.hi[2]);
    const float acc = adj_acc<NDIM>(g, e, i0, i1, k, wm, wc, wp, h_at);
    gu2[e] = gun[e] + dt * (alpha * acc);
    if (inner) {
        const float t = gvt_at(e);
        gv[e] = first ? gvn[e] + dt * gun[e] : t;
        if (gc2) {
            const float c = u2[e];
            float a = 0.0f;
            if (NDIM >= 3) {
                a += wm * u2[e - g.S0];
                a += wc * c;
                a += wp * u2[e + g.S0];
            }
            if (NDIM >= 2) {
                a += wm * u2[e - g.S1];
                a += wc * c;
                a += wp * u2[e + g.S1];
            }
            a += wm * u2[e - g.C];
            a += wc * c;
            a += wp * u2[e + g.C];
            gc2[e] = gc2[e] + dt * (t * (alpha * a));
        }
    } else {
        gv[e] = gvn[e];
    }
    PDE_ADJ_LOOP_END
}

// Boundary transpose.  The linear part of a boundary apply is out[dst] =
// sign * in[src] (sign 0 for a constant write) on written scalars and
// out = in elsewhere; its transpose zeroes every dst, then each distinct
// src gathers: gin[src] = (src is also a dst ? 0 : gout[src]) + sum over its
// writes, in table order, of sign * gout[dst].
__global__ __launch_bounds__(256) void bc_adj_zero_kernel(float *__restrict__ gin,
                                                          const int64_t *__restrict__ idx,
                                                          int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) gin[idx[i]] = 0.0f;
}

__global__ __launch_bounds__(256) void bc_adj_gather_kernel(
    const float *__restrict__ gout, float *__restrict__ gin, const int64_t *__restrict__ src,
    const int32_t *__restrict__ flags, const int64_t *__restrict__ off,
    const int64_t *__restrict__ cdst, const float *__restrict__ csign, int64_t nsrc)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nsrc) return;
    const int64_t s = src[i];
    float acc = (flags[i] & 1) ? 0.0f : gout[s];
    for (int64_t q = off[i]; q < off[i + 1]; ++q) acc += csign[q] * gout[cdst[q]];
    gin[s] = acc;
}

// In-place form: gather the new source values into staging (reading the
// untouched g), then zero every dst, then write the staged values.
__global__ __launch_bounds__(256) void bc_adj_gather_stage_kernel(
    const float *__restrict__ g, const int64_t *__restrict__ src,
    const int32_t *__restrict__ flags, const int64_t *__restrict__ off,
    const int64_t *__restrict__ cdst, const float *__restrict__ csign, int64_t nsrc,
    float *__restrict__ staging)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nsrc) return;
    float acc = (flags[i] & 1) ? 0.0f : g[src[i]];
    for (int64_t q = off[i]; q < off[i + 1]; ++q) acc += csign[q] * g[cdst[q]];
    staging[i] = acc;
}

// Second phase of the in-place transpose: threads [0, nsrc) store the staged
// source values, threads [nsrc, nsrc + nzero) zero the written scalars that
// are not sources (the two index sets are disjoint).
__global__ __launch_bounds__(256) void bc_adj_write_kernel(float *__restrict__ g,
                                                           const int64_t *__restrict__ src,
                                                           const float *__restrict__ staging,
                                                           int64_t nsrc,
                                                           const int64_t *__restrict__ zero_idx,
                                                           int64_t nzero)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < nsrc) g[src[i]] = staging[i];
    else if (i < nsrc + nzero) g[zero_idx[i - nsrc]] = 0.0f;
}

// The in-place transpose by one workgroup of 1024 threads for small tables
// (latency-bound; one CU cannot stream a large one): staged values (same accumulation order as the stage kernel: own
// term, then csign[q] * g[cdst[q]] for q ascending) go to LDS; after the
// barrier the sources are stored and the zero_idx scalars cleared.  Each
// thread's sources are processed BC_UNROLL at a time, stage by stage, so the
// dependent loads of different sources overlap.
__global__ __launch_bounds__(1024) void bc_adj_inplace_block_kernel(
    float *__restrict__ g, const int64_t *__restrict__ zero_idx, int64_t nzero,
    const int64_t *__restrict__ src, const int32_t *__restrict__ flags,
    const int64_t *__restrict__ off, const int64_t *__restrict__ cdst,
    const float *__restrict__ csign, int64_t nsrc)
{
    __shared__ float vals[BC_BLOCK_CAP];
    const int n = (int)nsrc;
    for (int base = 0; base < n; base += 1024 * BC_UNROLL) {
        int64_t sj[BC_UNROLL], o0[BC_UNROLL], o1[BC_UNROLL], d0[BC_UNROLL];
        int32_t fj[BC_UNROLL];
        float acc[BC_UNROLL], c0[BC_UNROLL];
#pragma unroll
        for (int j = 0; j < BC_UNROLL; ++j) {
            const int m = base + j * 1024 + (int)threadIdx.x;
            if (m < n) {
                sj[j] = src[m];
                fj[j] = flags[m];
                o0[j] = off[m];
                o1[j] = off[m + 1];
            }
        }
#pragma unroll
        for (int j = 0; j < BC_UNROLL; ++j) {
            const int m = base + j * 1024 + (int)threadIdx.x;
            if (m < n) {
                acc[j] = (fj[j] & 1) ? 0.0f : g[sj[j]];
                if (o1[j] > o0[j]) {
                    d0[j] = cdst[o0[j]];
                    c0[j] = csign[o0[j]];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < BC_UNROLL; ++j) {
            const int m = base + j * 1024 + (int)threadIdx.x;
            if (m < n) {
                float a = acc[j];
                if (o1[j] > o0[j]) a += c0[j] * g[d0[j]];
                for (int64_t q = o0[j] + 1; q < o1[j]; ++q) a += csign[q] * g[cdst[q]];
                vals[m] = a;
            }
        }
    }
    __syncthreads();
    for (int base = 0; base < n; base += 1024 * BC_UNROLL) {
        int64_t sj[BC_UNROLL];
#pragma unroll
        for (int j = 0; j < BC_UNROLL; ++j) {
            const int m = base + j * 1024 + (int)threadIdx.x;
            if (m < n) sj[j] = src[m];
        }
#pragma unroll
        for (int j = 0; j < BC_UNROLL; ++j) {
            const int m = base + j * 1024 + (int)threadIdx.x;
            if (m < n) g[sj[j]] = vals[m];
        }
    }
    for (int64_t i = threadIdx.x; i < nzero; i += 1024) g[zero_idx[i]] = 0.0f;
}

// The in-place transpose over conflict-closed chunks (host-built): chunk c
// holds sources [item_start[c], item_start[c+1]) (at most BC_CHUNK) and zero
// entries [zero_start[c], zero_start[c+1]); every scalar a chunk's sources
// read that the transpose writes belongs to the same chunk.  One workgroup
// per chunk: stage, barrier, store and zero.
__global__ __launch_bounds__(BC_CHUNK) void bc_adj_chunk_inplace_kernel(
    float *__restrict__ g, const int64_t *__restrict__ src, const int32_t *__restrict__ flags,
    const int64_t *__restrict__ off, const int64_t *__restrict__ cdst,
    const float *__restrict__ csign, const int64_t *__restrict__ item_start,
    const int64_t *__restrict__ zero_idx, const int64_t *__restrict__ zero_start)
{
    const int64_t ib = item_start[blockIdx.x], ie = item_start[blockIdx.x + 1];
    const int64_t i = ib + threadIdx.x;
    const bool on = i < ie && ie - ib <= BC_CHUNK;
    float acc = 0.0f;
    int64_t s = 0;
    if (on) {
        s = src[i];
        acc = (flags[i] & 1) ? 0.0f : g[s];
        for (int64_t q = off[i]; q < off[i + 1]; ++q) acc += csign[q] * g[cdst[q]];
    }
    __syncthreads();
    if (on) g[s] = acc;
    for (int64_t z = zero_start[blockIdx.x] + threadIdx.x; z < zero_start[blockIdx.x + 1];
         z += BC_CHUNK)
        g[zero_idx[z]] = 0.0f;
}

__global__ __launch_bounds__(256) void scale_kernel(const float *in, float *out, int64_t n, float s)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = s * in[i];
}

static dim3 adj_grid(const AdjGeom &g)
{
    return dim3(clamp_grid((g.R + 255) / 256, 0x7fffffff), clamp_grid(g.P1, 65535),
                clamp_grid(g.P0, 65535));
}

}  // namespace pde

using namespace pde;

#define PDE_ADJ_DISPATCH(KERNEL, ND, GRID, ...)                                               \
    switch (ND) {                                                                            \
    case 1: hipLaunchKernelGGL(KERNEL<1>, GRID, dim3(256), 0, s, __VA_ARGS__); break;        \
    case 2: hipLaunchKernelGGL(KERNEL<2>, GRID, dim3(256), 0, s, __VA_ARGS__); break;        \
    default: hipLaunchKernelGGL(KERNEL<3>, GRID, dim3(256), 0, s, __VA_ARGS__); break;       \
    }

extern "C" int pde_laplacian_adjoint_f32(const float *gout, float *gin, int64_t nx, int64_t ny,
                                         int64_t nz, int32_t ncomp, int32_t ghost, float wm,
                                         float wc, float wp, float alpha, void *stream)
{
    if (!gout || !gin || gout == gin || ncomp < 1) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim == 0) return PDE_ESHAPE;
    AdjGeom g = adj_geom(G, ncomp);
    hipStream_t s = as_stream(stream);
    PDE_ADJ_DISPATCH(lap_adj_kernel, G.ndim, adj_grid(g), gout, gin, g, wm, wc, wp, alpha)
    return launch_status();
}

extern "C" int pde_grad_adjoint_f32(const float *gout, float *gin, int64_t nx, int64_t ny,
                                    int64_t nz, int32_t ncomp, int32_t ghost, float wm, float wc,
                                    float wp, float alpha, void *stream)
{
    if (!gout || !gin || gout == gin || ncomp < 1) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim == 0) return PDE_ESHAPE;
    AdjGeom g = adj_geom(G, ncomp);
    hipStream_t s = as_stream(stream);
    PDE_ADJ_DISPATCH(grad_adj_kernel, G.ndim, adj_grid(g), gout, gin, g, wm, wc, wp, alpha)
    return launch_status();
}

extern "C" int pde_divergence_adjoint_f32(const float *gout, float *gin, int64_t nx, int64_t ny,
                                          int64_t nz, int32_t nrows, int32_t ghost, float wm,
                                          float wc, float wp, float alpha, void *stream)
{
    if (!gout || !gin || gout == gin || nrows < 1) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim == 0) return PDE_ESHAPE;
    AdjGeom g = adj_geom(G, nrows);
    hipStream_t s = as_stream(stream);
    PDE_ADJ_DISPATCH(div_adj_kernel, G.ndim, adj_grid(g), gout, gin, g, wm, wc, wp, alpha)
    return launch_status();
}

extern "C" int pde_heat_step_adjoint_f32(const float *gn, float *gu2, int64_t nx, int64_t ny,
                                         int64_t nz, int32_t ncomp, int32_t ghost, float wm,
                                         float wc, float wp, float alpha, float dt, void *stream)
{
    if (!gn || !gu2 || gn == gu2 || ncomp < 1) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim == 0) return PDE_ESHAPE;
    hipStream_t s = as_stream(stream);
    st = step2d_launch(STEP2D_HEAT_ADJ, G, ncomp, gn, nullptr, nullptr, nullptr, gu2, nullptr,
                       nullptr, wm, wc, wp, alpha, dt, 0, nullptr, ghost, s);
    if (st != STEP2D_NA) return st;
    AdjGeom g = adj_geom(G, ncomp);
    PDE_ADJ_DISPATCH(heat_adj_kernel, G.ndim, adj_grid(g), gn, gu2, g, wm, wc, wp, alpha, dt)
    return launch_status();
}

static int wave_adj_impl(const float *gun, const float *gvn, const float *c2, const float *u2,
                         float *gu2, float *gv, float *gc2, int64_t nx, int64_t ny, int64_t nz,
                         int32_t ghost, float wm, float wc, float wp, float alpha, float dt,
                         int32_t first_step, const int64_t *step, void *stream)
{
    if (!gun || !gvn || !gu2 || !gv) return PDE_EINVAL;
    if (gc2 && !u2) return PDE_EINVAL;
    const float *ins[4] = {gun, gvn, c2, u2};
    float *outs[3] = {gu2, gv, gc2};
    for (float *o : outs)
        for (const float *i : ins)
            if (o && i && o == i) return PDE_EINVAL;
    if (gu2 == gv || (gc2 && (gc2 == gu2 || gc2 == gv))) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim == 0) return PDE_ESHAPE;
    hipStream_t s = as_stream(stream);
    const int f = first_step ? 1 : 0;
    st = step2d_launch(STEP2D_WAVE_ADJ, G, 1, gun, gvn, c2, u2, gu2, gv, gc2, wm, wc, wp, alpha,
                       dt, f, step, ghost, s);
    if (st != STEP2D_NA) return st;
    AdjGeom g = adj_geom(G, 1);
    PDE_ADJ_DISPATCH(wave_adj_kernel, G.ndim, adj_grid(g), gun, gvn, c2, u2, gu2, gv, gc2, g, wm,
                     wc, wp, alpha, dt, f, step)
    return launch_status();
}

extern "C" int pde_wave_step_adjoint_f32(const float *gun, const float *gvn, const float *c2,
                                         const float *u2, float *gu2, float *gv, float *gc2,
                                         int64_t nx, int64_t ny, int64_t nz, int32_t ghost,
                                         float wm, float wc, float wp, float alpha, float dt,
                                         int32_t first_step, void *stream)
{
    return wave_adj_impl(gun, gvn, c2, u2, gu2, gv, gc2, nx, ny, nz, ghost, wm, wc, wp, alpha, dt,
                         first_step, nullptr, stream);
}

extern "C" int pde_wave_step_adjoint_ctr_f32(const float *gun, const float *gvn, const float *c2,
                                             const float *u2, float *gu2, float *gv, float *gc2,
                                             int64_t nx, int64_t ny, int64_t nz, int32_t ghost,
                                             float wm, float wc, float wp, float alpha, float dt,
                                             const int64_t *step, void *stream)
{
    if (!step) return PDE_EINVAL;
    return wave_adj_impl(gun, gvn, c2, u2, gu2, gv, gc2, nx, ny, nz, ghost, wm, wc, wp, alpha, dt,
                         0, step, stream);
}

extern "C" int pde_copy_f32(const float *in, float *out, int64_t n, void *stream);

extern "C" int pde_bc_adjoint_f32(const float *gout, float *gin, int64_t n,
                                  const int64_t *zero_idx, int64_t nzero,
                                  const int64_t *src_idx, const int32_t *src_flags,
                                  const int64_t *off, const int64_t *cdst, const float *csign,
                                  int64_t nsrc, void *stream)
{
    if (!gout || !gin || gout == gin || n < 0 || nzero < 0 || nsrc < 0) return PDE_EINVAL;
    if ((nzero > 0 && !zero_idx) || (nsrc > 0 && (!src_idx || !src_flags || !off || !cdst || !csign)))
        return PDE_EINVAL;
    int st = pde_copy_f32(gout, gin, n, stream);
    if (st) return st;
    hipStream_t s = as_stream(stream);
    if (nzero > 0) {
        hipLaunchKernelGGL(bc_adj_zero_kernel, dim3((unsigned)((nzero + 255) / 256)), dim3(256), 0,
                           s, gin, zero_idx, nzero);
        st = launch_status();
        if (st) return st;
    }
    if (nsrc > 0) {
        hipLaunchKernelGGL(bc_adj_gather_kernel, dim3((unsigned)((nsrc + 255) / 256)), dim3(256),
                           0, s, gout, gin, src_idx, src_flags, off, cdst, csign, nsrc);
        st = launch_status();
    }
    return st;
}

extern "C" int pde_bc_adjoint_inplace_f32(float *g, const int64_t *zero_idx, int64_t nzero,
                                          const int64_t *src_idx, const int32_t *src_flags,
                                          const int64_t *off, const int64_t *cdst,
                                          const float *csign, int64_t nsrc, float *staging,
                                          void *stream)
{
    if (!g || nzero < 0 || nsrc < 0) return PDE_EINVAL;
    if ((nzero > 0 && !zero_idx) ||
        (nsrc > 0 && (!src_idx || !src_flags || !off || !cdst || !csign || !staging)))
        return PDE_EINVAL;
    hipStream_t s = as_stream(stream);
    int st = PDE_OK;
    if (nsrc <= 2048 && nzero <= 4096) {
        if (nsrc + nzero == 0) return PDE_OK;
        hipLaunchKernelGGL(bc_adj_inplace_block_kernel, dim3(1), dim3(1024), 0, s, g, zero_idx,
                           nzero, src_idx, src_flags, off, cdst, csign, nsrc);
        return launch_status();
    }
    if (nsrc > 0) {
        hipLaunchKernelGGL(bc_adj_gather_stage_kernel, dim3((unsigned)((nsrc + 255) / 256)),
                           dim3(256), 0, s, g, src_idx, src_flags, off, cdst, csign, nsrc, staging);
        st = launch_status();
        if (st) return st;
    }
    if (nsrc + nzero > 0) {
        hipLaunchKernelGGL(bc_adj_write_kernel, dim3((unsigned)((nsrc + nzero + 255) / 256)),
                           dim3(256), 0, s, g, src_idx, staging, nsrc, zero_idx, nzero);
        st = launch_status();
    }
    return st;
}

extern "C" int pde_scale_f32(const float *in, float *out, int64_t n, float scale, void *stream)
{
    if (!in || !out || n < 0) return PDE_EINVAL;
    if (n == 0) return PDE_OK;
    hipLaunchKernelGGL(scale_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       as_stream(stream), in, out, n, scale);
    return launch_status();
}

extern "C" int pde_bc_adjoint_chunked_inplace_f32(float *g, const int64_t *src_idx,
                                                  const int32_t *src_flags, const int64_t *off,
                                                  const int64_t *cdst, const float *csign,
                                                  const int64_t *item_start,
                                                  const int64_t *zero_idx,
                                                  const int64_t *zero_start, int64_t nchunks,
                                                  void *stream)
{
    if (!g || nchunks < 0) return PDE_EINVAL;
    if (nchunks == 0) return PDE_OK;
    if (!src_idx || !src_flags || !off || !cdst || !csign || !item_start || !zero_idx ||
        !zero_start || nchunks > 0x7fffffff)
        return PDE_EINVAL;
    hipLaunchKernelGGL(bc_adj_chunk_inplace_kernel, dim3((unsigned)nchunks), dim3(BC_CHUNK), 0,
                       as_stream(stream), g, src_idx, src_flags, off, cdst, csign, item_start,
                       zero_idx, zero_start);
    return launch_status();
}
