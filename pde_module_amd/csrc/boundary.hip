// Ghost-cell boundary conditions from a host-resolved op table.
//
// The reference applies a GridBoundary as copy(in -> out) followed by one
// launch per boundary group (pm/FDM/boundary/functional_boundary.py:278-291)
// or one table launch over (boundary node, component)
// (pm/deprecated/gridBoundary.py:231-242).  Every write reads only `in`, so
// the final value of each written scalar is the last group's formula applied
// to `in`; the host resolves that into one op per destination and a single
// launch here reproduces either semantics exactly.
#include "pde_common.hpp"

#include <mutex>

// Staging of pde_bc_apply_inplace_noalloc_f32: one copy per GPU, for the
// process's lifetime.  In host code this name is only a handle for
// hipGetSymbolAddress: kernels receive the per-device address it returns.
static __device__ __align__(256) float scratch[PDE_BC_SCRATCH_FLOATS];

namespace pde {

__global__ __launch_bounds__(256) void bc_apply_kernel(const float *__restrict__ in,
                                                       float *__restrict__ out,
                                                       const pde_bc_op *__restrict__ ops,
                                                       int64_t nops, const float *t)
{
    const int64_t m = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (m >= nops) return;
    const pde_bc_op op = ops[m];
    out[op.dst] = op_eval(op, in, t);
}

__global__ __launch_bounds__(256) void bc_gather_kernel(const float *__restrict__ u,
                                                        const pde_bc_op *__restrict__ ops,
                                                        int64_t nops, const float *t,
                                                        float *__restrict__ staging)
{
    const int64_t m = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (m >= nops) return;
    staging[m] = op_eval(ops[m], u, t);
}

__global__ __launch_bounds__(256) void bc_scatter_kernel(float *__restrict__ u,
                                                         const pde_bc_op *__restrict__ ops,
                                                         int64_t nops,
                                                         const float *__restrict__ staging)
{
    const int64_t m = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (m >= nops) return;
    u[ops[m].dst] = staging[m];
}

// Small tables (up to 4096 ops: latency-bound) gather and scatter in one
// launch of one workgroup, with a workgroup barrier between the phases.
__global__ __launch_bounds__(1024) void bc_inplace_small_kernel(float *__restrict__ u,
                                                                const pde_bc_op *__restrict__ ops,
                                                                int64_t nops, const float *t)
{
    __shared__ float vals[BC_BLOCK_CAP];
    bc_inplace_block(u, ops, (int)nops, t, vals);
}

// Conflict-closed chunks (host-built): no op reads a scalar that an op of
// another chunk writes, so each workgroup applies its chunk in place on its
// own (all reads, workgroup barrier, all writes) and one full-chip launch
// applies the whole table.
// With SAVE, op m also records the value its destination held before the
// apply in save[n * nops + m], n = *step: the pre-boundary field on the
// written scalars, which a time-reversed walk needs to rebuild each state.
template <bool SAVE>
__global__ __launch_bounds__(BC_CHUNK) void bc_chunk_inplace_kernel(
    float *__restrict__ u, const pde_bc_op *__restrict__ ops,
    const int64_t *__restrict__ chunk_start, const float *t, float *__restrict__ save,
    int64_t nops, const int64_t *__restrict__ step)
{
    const int64_t b = chunk_start[blockIdx.x], e = chunk_start[blockIdx.x + 1];
    const int64_t m = b + threadIdx.x;
    const bool on = m < e && e - b <= BC_CHUNK;
    float v = 0.0f;
    int64_t dst = 0;
    if (on) {
        const pde_bc_op op = ops[m];
        v = op_eval(op, u, t);
        dst = op.dst;
        if (SAVE) save[*step * nops + m] = u[dst];
    }
    __syncthreads();
    if (on) u[dst] = v;
}

}  // namespace pde

using namespace pde;

extern "C" int pde_copy_f32(const float *in, float *out, int64_t n, void *stream);

extern "C" int pde_bc_apply_f32(const float *in, float *out, int64_t n, const pde_bc_op *ops,
                                int64_t nops, const float *t, void *stream)
{
    if (!in || !out || n < 0 || nops < 0 || (nops > 0 && !ops) || in == out) return PDE_EINVAL;
    int st = pde_copy_f32(in, out, n, stream);
    if (st) return st;
    if (nops == 0) return PDE_OK;
    hipLaunchKernelGGL(bc_apply_kernel, dim3((unsigned)((nops + 255) / 256)), dim3(256), 0,
                       as_stream(stream), in, out, ops, nops, t);
    return launch_status();
}

extern "C" int pde_bc_apply_inplace_f32(float *u, const pde_bc_op *ops, int64_t nops,
                                        const float *t, float *staging, void *stream)
{
    if (!u || nops < 0 || (nops > 0 && !ops)) return PDE_EINVAL;
    if (nops == 0) return PDE_OK;
    hipStream_t s = as_stream(stream);
    if (nops <= 4096) {
        hipLaunchKernelGGL(bc_inplace_small_kernel, dim3(1), dim3(1024), 0, s, u, ops, nops, t);
        return launch_status();
    }
    if (!staging) return PDE_EINVAL;
    const dim3 grid((unsigned)((nops + 255) / 256)), block(256);
    hipLaunchKernelGGL(bc_gather_kernel, grid, block, 0, s, u, ops, nops, t, staging);
    int st = launch_status();
    if (st) return st;
    hipLaunchKernelGGL(bc_scatter_kernel, grid, block, 0, s, u, ops, nops, staging);
    return launch_status();
}

namespace {

constexpr int kScratchDevices = 64;
std::mutex scratch_mu;
float *scratch_ptr[kScratchDevices] = {};

// The current device's copy of `scratch` (looked up once per device).
int scratch_of_current_device(float **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= kScratchDevices) return PDE_EINVAL;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        if (!p) return PDE_EINVAL;
        scratch_ptr[dev] = static_cast<float *>(p);
    }
    *out = scratch_ptr[dev];
    return PDE_OK;
}

}  // namespace

extern "C" int pde_bc_apply_inplace_noalloc_f32(float *u, const pde_bc_op *ops, int64_t nops,
                                                const float *t, void *stream)
{
    if (!u || nops < 0 || (nops > 0 && !ops)) return PDE_EINVAL;
    if (nops > PDE_BC_SCRATCH_FLOATS) return PDE_ETOOBIG;
    if (nops <= 4096)      // one workgroup stages in LDS: no scratch needed
        return pde_bc_apply_inplace_f32(u, ops, nops, t, nullptr, stream);
    float *staging = nullptr;
    int st = scratch_of_current_device(&staging);
    if (st) return st;
    return pde_bc_apply_inplace_f32(u, ops, nops, t, staging, stream);
}

extern "C" int pde_bc_apply_chunked_inplace_f32(float *u, const pde_bc_op *ops,
                                                const int64_t *chunk_start, int64_t nchunks,
                                                const float *t, void *stream)
{
    if (!u || nchunks < 0) return PDE_EINVAL;
    if (nchunks == 0) return PDE_OK;
    if (!ops || !chunk_start || nchunks > 0x7fffffff) return PDE_EINVAL;
    hipLaunchKernelGGL(bc_chunk_inplace_kernel<false>, dim3((unsigned)nchunks), dim3(BC_CHUNK),
                       0, as_stream(stream), u, ops, chunk_start, t, nullptr, 0,
                       (const int64_t *)nullptr);
    return launch_status();
}

extern "C" int pde_bc_apply_chunked_inplace_save_f32(float *u, const pde_bc_op *ops,
                                                     const int64_t *chunk_start, int64_t nchunks,
                                                     int64_t nops, const float *t, float *save,
                                                     const int64_t *step, void *stream)
{
    if (!u || nchunks < 0 || nops < 0 || !save || !step) return PDE_EINVAL;
    if (nchunks == 0) return PDE_OK;
    if (!ops || !chunk_start || nchunks > 0x7fffffff) return PDE_EINVAL;
    hipLaunchKernelGGL(bc_chunk_inplace_kernel<true>, dim3((unsigned)nchunks), dim3(BC_CHUNK), 0,
                       as_stream(stream), u, ops, chunk_start, t, save, nops, step);
    return launch_status();
}
