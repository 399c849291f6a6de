// Explicit time stepping: ForwardEuler, the device copy, and the fused heat
// and u/v wave steps (boundary-applied state in, next state out, one HBM pass).
#include "pde_common.hpp"

namespace pde {

__device__ __forceinline__ bool aligned16(const void *p)
{
    return ((uintptr_t)p & 15u) == 0;
}

typedef float cp_f4 __attribute__((ext_vector_type(4)));

// out = in; float4 body when both pointers are 16-byte aligned.  Every byte
// is read once and written once, so both sides stream (non-temporal loads
// and stores: 84% of the HBM peak at 543 MB against 77% with plain ones).
__global__ __launch_bounds__(256) void copy_kernel(const float *__restrict__ in,
                                                   float *__restrict__ out, int64_t n)
{
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t i = q * 4;
    if (i + 3 < n && aligned16(in) && aligned16(out)) {
        const cp_f4 v = __builtin_nontemporal_load(reinterpret_cast<const cp_f4 *>(in) + q);
        __builtin_nontemporal_store(v, reinterpret_cast<cp_f4 *>(out) + q);
        return;
    }
    for (int64_t e = i; e < n && e < i + 4; ++e) out[e] = in[e];
}

// ForwardEuler (pm/time_step/forwardEuler.py:79): new = cur + dt * k.  No
// __restrict__: out may alias cur or k (element-wise, same index).  Every
// byte is read or written once: streaming loads and stores, as in the copy.
__global__ __launch_bounds__(256) void euler_kernel(const float *cur, const float *k, float *out,
                                                    int64_t n, float dt)
{
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t i = q * 4;
    if (i + 3 < n && aligned16(cur) && aligned16(k) && aligned16(out)) {
        const cp_f4 a = __builtin_nontemporal_load(reinterpret_cast<const cp_f4 *>(cur) + q);
        const cp_f4 b = __builtin_nontemporal_load(reinterpret_cast<const cp_f4 *>(k) + q);
        cp_f4 r;
        r.x = a.x + dt * b.x;
        r.y = a.y + dt * b.y;
        r.z = a.z + dt * b.z;
        r.w = a.w + dt * b.w;
        __builtin_nontemporal_store(r, reinterpret_cast<cp_f4 *>(out) + q);
        return;
    }
    for (int64_t e = i; e < n && e < i + 4; ++e) out[e] = cur[e] + dt * k[e];
}

// Whole-array geometry for the fused steps: a thread owns one float of a
// padded row; rows are (i0, i1).  Interior tests are per canonical axis.
struct StepGeom {
    int64_t S0, S1;   // float strides of canonical axes 0, 1
    int64_t C;        // components per node (axis-2 float stride)
    int64_t P0, P1, R;  // padded planes, rows, floats per row (P2 * C)
    int64_t lo0, hi0, lo1, hi1, lo2, hi2;  // interior node ranges
};

inline StepGeom step_geom(const Geom &G, int64_t C)
{
    StepGeom s;
    s.S0 = G.S[0] * C;
    s.S1 = G.S[1] * C;
    s.C = C;
    s.P0 = G.P[0];
    s.P1 = G.P[1];
    s.R = G.P[2] * C;
    s.lo0 = G.G[0];
    s.hi0 = G.G[0] + G.E[0];
    s.lo1 = G.G[1];
    s.hi1 = G.G[1] + G.E[1];
    s.lo2 = G.G[2];
    s.hi2 = G.G[2] + G.E[2];
    return s;
}

template <int NDIM>
__device__ __forceinline__ float lap_acc(const float *__restrict__ u, int64_t e,
                                         const StepGeom &g, float wm, float wc, float wp)
{
    const float c = u[e];
    float acc = 0.0f;
    if (NDIM >= 3) {
        acc += wm * u[e - g.S0];
        acc += wc * c;
        acc += wp * u[e + g.S0];
    }
    if (NDIM >= 2) {
        acc += wm * u[e - g.S1];
        acc += wc * c;
        acc += wp * u[e + g.S1];
    }
    if (NDIM >= 1) {
        acc += wm * u[e - g.C];
        acc += wc * c;
        acc += wp * u[e + g.C];
    }
    return acc;
}

// Fused heat step: u_next = u2 + dt * (alpha * lap_acc(u2)) on non-ghost
// nodes, u_next = u2 on ghosts.  Same arithmetic as Laplacian followed by
// ForwardEuler in the reference loop.
template <int NDIM>
__global__ __launch_bounds__(256) void heat_step_kernel(const float *__restrict__ u,
                                                        float *__restrict__ un, StepGeom g,
                                                        float wm, float wc, float wp,
                                                        float alpha, float dt)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= g.R) return;
    const int64_t k = r / g.C;
    const bool in2 = k >= g.lo2 && k < g.hi2;
    for (int64_t i0 = blockIdx.z; i0 < g.P0; i0 += gridDim.z)
        for (int64_t i1 = blockIdx.y; i1 < g.P1; i1 += gridDim.y) {
            const int64_t e = i0 * g.S0 + i1 * g.S1 + r;
            const bool inner = in2 && i0 >= g.lo0 && i0 < g.hi0 && i1 >= g.lo1 && i1 < g.hi1;
            if (inner) {
                const float acc = lap_acc<NDIM>(u, e, g, wm, wc, wp);
                un[e] = u[e] + dt * (alpha * acc);
            } else {
                un[e] = u[e];
            }
        }
}

// Fused u/v wave step (scalar field).  v and vn may alias (same index only).
template <int NDIM>
__global__ __launch_bounds__(256) void wave_step_kernel(const float *__restrict__ u,
                                                        const float *v, const float *c2,
                                                        float *__restrict__ un, float *vn,
                                                        StepGeom g, float wm, float wc,
                                                        float wp, float alpha, float dt,
                                                        int first, const int64_t *step)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= g.R) return;
    if (step) first = (*step == 0);
    const bool in2 = r >= g.lo2 && r < g.hi2;
    for (int64_t i0 = blockIdx.z; i0 < g.P0; i0 += gridDim.z)
        for (int64_t i1 = blockIdx.y; i1 < g.P1; i1 += gridDim.y) {
            const int64_t e = i0 * g.S0 + i1 * g.S1 + r;
            const bool inner = in2 && i0 >= g.lo0 && i0 < g.hi0 && i1 >= g.lo1 && i1 < g.hi1;
            const float vo = v[e];
            if (inner) {
                const float acc = lap_acc<NDIM>(u, e, g, wm, wc, wp);
                float rhs = alpha * acc;
                if (c2) rhs = c2[e] * rhs;
                const float vnew = vo + dt * rhs;
                vn[e] = vnew;
                un[e] = u[e] + dt * (first ? vo : vnew);
            } else {
                vn[e] = vo;
                un[e] = u[e];
            }
        }
}

static dim3 step_grid(const StepGeom &g)
{
    return dim3(clamp_grid((g.R + 255) / 256, 0x7fffffff), clamp_grid(g.P1, 65535),
                clamp_grid(g.P0, 65535));
}

}  // namespace pde

using namespace pde;

extern "C" int pde_copy_f32(const float *in, float *out, int64_t n, void *stream)
{
    if (!in || !out || n < 0) return PDE_EINVAL;
    if (n == 0) return PDE_OK;
    const int64_t threads = (n + 3) / 4;
    if ((threads + 255) / 256 > 0x7fffffff) return PDE_ESHAPE;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                       as_stream(stream), in, out, n);
    return launch_status();
}

extern "C" int pde_forward_euler_f32(const float *cur, const float *k, float *out, int64_t n,
                                     float dt, void *stream)
{
    if (!cur || !k || !out || n < 0) return PDE_EINVAL;
    if (n == 0) return PDE_OK;
    const int64_t threads = (n + 3) / 4;
    if ((threads + 255) / 256 > 0x7fffffff) return PDE_ESHAPE;
    hipLaunchKernelGGL(euler_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                       as_stream(stream), cur, k, out, n, dt);
    return launch_status();
}

extern "C" int pde_heat_step_f32(const float *u, float *un, int64_t nx, int64_t ny, int64_t nz,
                                 int32_t ncomp, int32_t ghost, float wm, float wc, float wp,
                                 float alpha, float dt, void *stream)
{
    if (!u || !un || u == un || ncomp < 1) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim == 0) return PDE_ESHAPE;
    hipStream_t s = as_stream(stream);
    st = step2d_launch(STEP2D_HEAT, G, ncomp, u, nullptr, nullptr, nullptr, un, nullptr, nullptr,
                       wm, wc, wp, alpha, dt, 0, nullptr, ghost, s);
    if (st != STEP2D_NA) return st;
    StepGeom g = step_geom(G, ncomp);
    const dim3 grid = step_grid(g), block(256);
    switch (G.ndim) {
    case 1: hipLaunchKernelGGL(heat_step_kernel<1>, grid, block, 0, s, u, un, g, wm, wc, wp, alpha, dt); break;
    case 2: hipLaunchKernelGGL(heat_step_kernel<2>, grid, block, 0, s, u, un, g, wm, wc, wp, alpha, dt); break;
    default: hipLaunchKernelGGL(heat_step_kernel<3>, grid, block, 0, s, u, un, g, wm, wc, wp, alpha, dt); break;
    }
    return launch_status();
}

static int wave_step_impl(const float *u, const float *v, const float *c2, float *un, float *vn,
                          int64_t nx, int64_t ny, int64_t nz, int32_t ghost, float wm, float wc,
                          float wp, float alpha, float dt, int32_t first_step,
                          const int64_t *step, void *stream)
{
    if (!u || !v || !un || !vn || u == un || un == vn || un == v || vn == u) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim == 0) return PDE_ESHAPE;
    hipStream_t s = as_stream(stream);
    const int f = first_step ? 1 : 0;
    const int64_t *ctr = step;
    st = step2d_launch(STEP2D_WAVE, G, 1, u, v, c2, nullptr, un, vn, nullptr, wm, wc, wp, alpha,
                       dt, f, ctr, ghost, s);
    if (st != STEP2D_NA) return st;
    StepGeom g = step_geom(G, 1);
    const dim3 grid = step_grid(g), block(256);
    switch (G.ndim) {
    case 1: hipLaunchKernelGGL(wave_step_kernel<1>, grid, block, 0, s, u, v, c2, un, vn, g, wm, wc, wp, alpha, dt, f, ctr); break;
    case 2: hipLaunchKernelGGL(wave_step_kernel<2>, grid, block, 0, s, u, v, c2, un, vn, g, wm, wc, wp, alpha, dt, f, ctr); break;
    default: hipLaunchKernelGGL(wave_step_kernel<3>, grid, block, 0, s, u, v, c2, un, vn, g, wm, wc, wp, alpha, dt, f, ctr); break;
    }
    return launch_status();
}

extern "C" int pde_wave_step_f32(const float *u, const float *v, const float *c2, float *un,
                                 float *vn, int64_t nx, int64_t ny, int64_t nz, int32_t ghost,
                                 float wm, float wc, float wp, float alpha, float dt,
                                 int32_t first_step, void *stream)
{
    return wave_step_impl(u, v, c2, un, vn, nx, ny, nz, ghost, wm, wc, wp, alpha, dt, first_step,
                          nullptr, stream);
}

extern "C" int pde_wave_step_ctr_f32(const float *u, const float *v, const float *c2, float *un,
                                     float *vn, int64_t nx, int64_t ny, int64_t nz, int32_t ghost,
                                     float wm, float wc, float wp, float alpha, float dt,
                                     const int64_t *step, void *stream)
{
    if (!step) return PDE_EINVAL;
    return wave_step_impl(u, v, c2, un, vn, nx, ny, nz, ghost, wm, wc, wp, alpha, dt, 0, step,
                          stream);
}

extern "C" int pde_version(void) { return 100; }

extern "C" const char *pde_status_string(int status)
{
    if (status == PDE_OK) return "ok";
    if (status == PDE_EINVAL) return "invalid argument";
    if (status == PDE_ESHAPE) return "grid too small for the ghost width or launch limits";
    if (status == PDE_ETOOBIG)
        return "table larger than the allocation-free entry point's scratch array";
    if (status > 0) return hipGetErrorString((hipError_t)status);
    return "unknown status";
}

// ------------------------------------------------------------ point samples
// Receivers and point sources of a time loop: gather field values at a list
// of flat indices, and add scaled values at a list of distinct flat indices.
namespace pde {

__global__ __launch_bounds__(256) void gather_kernel(const float *__restrict__ f,
                                                     const int64_t *__restrict__ idx,
                                                     float *__restrict__ out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = f[idx[i]];
}

__global__ __launch_bounds__(256) void scatter_add_kernel(float *__restrict__ f,
                                                          const int64_t *__restrict__ idx,
                                                          const float *__restrict__ vals,
                                                          int64_t n, float scale)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) f[idx[i]] = f[idx[i]] + scale * vals[i];
}

}  // namespace pde

extern "C" int pde_gather_f32(const float *field, const int64_t *idx, float *out, int64_t n,
                              void *stream)
{
    if (!field || !out || n < 0 || (n > 0 && !idx)) return PDE_EINVAL;
    if (n == 0) return PDE_OK;
    hipLaunchKernelGGL(gather_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       as_stream(stream), field, idx, out, n);
    return launch_status();
}

extern "C" int pde_scatter_add_f32(float *field, const int64_t *idx, const float *vals, int64_t n,
                                   float scale, void *stream)
{
    if (!field || n < 0 || (n > 0 && (!idx || !vals))) return PDE_EINVAL;
    if (n == 0) return PDE_OK;
    hipLaunchKernelGGL(scatter_add_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       as_stream(stream), field, idx, vals, n, scale);
    return launch_status();
}

// Step-counter forms: the row of `vals` / `out` is chosen by a device-resident
// step counter, so one captured graph of a loop segment can be replayed for
// every segment.
namespace pde {

__global__ __launch_bounds__(256) void gather_step_kernel(const float *__restrict__ f,
                                                          const int64_t *__restrict__ idx,
                                                          float *__restrict__ out, int64_t n,
                                                          int64_t stride,
                                                          const int64_t *__restrict__ step)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[*step * stride + i] = f[idx[i]];
}

__global__ __launch_bounds__(256) void scatter_add_step_kernel(
    float *__restrict__ f, const int64_t *__restrict__ idx, const float *__restrict__ vals,
    int64_t n, int64_t stride, float scale, const int64_t *__restrict__ step)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) f[idx[i]] = f[idx[i]] + scale * vals[*step * stride + i];
}

__global__ void counter_add_kernel(int64_t *c, int64_t delta)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) c[0] = c[0] + delta;
}

__global__ void counter_copy_kernel(const int64_t *src, int64_t *dst, int64_t delta)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) dst[0] = src[0] + delta;
}

}  // namespace pde

extern "C" int pde_gather_step_f32(const float *field, const int64_t *idx, float *out, int64_t n,
                                   int64_t stride, const int64_t *step, void *stream)
{
    if (!field || !out || !step || n < 0 || stride < 0 || (n > 0 && !idx)) return PDE_EINVAL;
    if (n == 0) return PDE_OK;
    hipLaunchKernelGGL(gather_step_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       as_stream(stream), field, idx, out, n, stride, step);
    return launch_status();
}

extern "C" int pde_scatter_add_step_f32(float *field, const int64_t *idx, const float *vals,
                                        int64_t n, int64_t stride, float scale,
                                        const int64_t *step, void *stream)
{
    if (!field || !step || n < 0 || stride < 0 || (n > 0 && (!idx || !vals))) return PDE_EINVAL;
    if (n == 0) return PDE_OK;
    hipLaunchKernelGGL(scatter_add_step_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       as_stream(stream), field, idx, vals, n, stride, scale, step);
    return launch_status();
}

extern "C" int pde_counter_add(int64_t *counter, int64_t delta, void *stream)
{
    if (!counter) return PDE_EINVAL;
    hipLaunchKernelGGL(counter_add_kernel, dim3(1), dim3(64), 0, as_stream(stream), counter,
                       delta);
    return launch_status();
}

extern "C" int pde_counter_copy(const int64_t *src, int64_t *dst, int64_t delta, void *stream)
{
    if (!src || !dst) return PDE_EINVAL;
    hipLaunchKernelGGL(counter_copy_kernel, dim3(1), dim3(64), 0, as_stream(stream), src, dst,
                       delta);
    return launch_status();
}
