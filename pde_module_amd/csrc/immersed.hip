// Staircase immersed boundary: Dirichlet values and Neumann running means on
// the solid boundary nodes, fill of the interior solid nodes (reference
// pm/FDM/boundary/immersedBoundary.py:214-300).  One launch covers both: the
// first nops threads take the (node, component) records, the rest the
// interior solid nodes.  Reads touch fluid nodes only and writes solid nodes
// only, so the same kernel serves in place.
#include "pde_common.hpp"

namespace pde {

struct IbArgs {
    int64_t stride[3];   // node strides of the eligible axes, ascending
    int32_t ndim;
    int32_t ncomp;
    float dx;
    float fill;
};

__global__ __launch_bounds__(256) void immersed_kernel(const float *in, float *out,
                                                       const pde_ib_op *__restrict__ ops,
                                                       int64_t nops,
                                                       const int64_t *__restrict__ fill_nodes,
                                                       int64_t nfill, IbArgs a)
{
    const int64_t total = nops + nfill;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * 256) {
        if (i < nops) {
            const pde_ib_op op = ops[i];
            const int64_t e = op.node * a.ncomp + op.comp;
            if (op.kind == 1) {
                out[e] = op.value;
            } else if (op.kind == 2) {
                float avg = 0.0f, n = 1.0f;
                for (int d = 0; d < a.ndim; ++d)
                    for (int side = 0; side < 2; ++side) {
                        if (!(op.mask & (1 << (2 * d + side)))) continue;
                        const int s = side ? 1 : -1;
                        const float c =
                            in[(op.node + s * a.stride[d]) * a.ncomp + op.comp] -
                            ((float)s * op.value) * a.dx;
                        avg += (c - avg) / n;
                        n += 1.0f;
                    }
                out[e] = avg;
            }
        } else {
            const int64_t node = fill_nodes[i - nops];
            for (int c = 0; c < a.ncomp; ++c) out[node * a.ncomp + c] = a.fill;
        }
    }
}

}  // namespace pde

using namespace pde;

extern "C" int pde_immersed_apply_f32(const float *in, float *out, const pde_ib_op *ops,
                                      int64_t nops, const int64_t *fill_nodes, int64_t nfill,
                                      float fill_value, int32_t ncomp, int32_t ndim,
                                      int64_t stride0, int64_t stride1, int64_t stride2,
                                      float dx, void *stream)
{
    if (!in || !out || nops < 0 || nfill < 0 || (nops > 0 && !ops) || (nfill > 0 && !fill_nodes) ||
        ncomp < 1 || ndim < 1 || ndim > 3)
        return PDE_EINVAL;
    const int64_t total = nops + nfill;
    if (total == 0) return PDE_OK;
    IbArgs a;
    a.stride[0] = stride0;
    a.stride[1] = stride1;
    a.stride[2] = stride2;
    a.ndim = ndim;
    a.ncomp = ncomp;
    a.dx = dx;
    a.fill = fill_value;
    hipLaunchKernelGGL(immersed_kernel, dim3(clamp_grid((total + 255) / 256, 65536)), dim3(256), 0,
                       as_stream(stream), in, out, ops, nops, fill_nodes, nfill, a);
    return launch_status();
}
