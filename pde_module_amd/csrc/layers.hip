// Absorbing boundary layers: the sponge (far-field relaxation) and the viscous
// damping layer (reference pm/FDM/kernel/farfield.py:8-78,
// viscousDampingLayer.py:8-76).  Both act on a host-built list of layer nodes
// with their distance levels; the damping profile per level comes from the
// host, so the kernel is one multiply and one subtract per scalar.
#include "pde_common.hpp"

namespace pde {

struct LayerArgs {
    float ff[PDE_LAYER_MAXC];   // far-field value per component
    int32_t ncomp;
    int32_t has_ff;
    float scale;
};

__global__ __launch_bounds__(256) void damping_layer_kernel(const float *__restrict__ in,
                                                            float *__restrict__ out,
                                                            const pde_layer_point *__restrict__ pts,
                                                            int64_t npts,
                                                            const float *__restrict__ profile,
                                                            LayerArgs a)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < npts;
         i += (int64_t)gridDim.x * 256) {
        const pde_layer_point p = pts[i];
        // damp_factor = sigma_max * (d / n) ** p; the output is -damp_factor * (...)
        const float f = a.scale * profile[p.level - 1];
        const int64_t e = p.node * a.ncomp;
        for (int c = 0; c < a.ncomp; ++c) {
            float v = in[e + c];
            if (a.has_ff) v = v - a.ff[c];
            out[e + c] = -f * v;
        }
    }
}

}  // namespace pde

using namespace pde;

extern "C" int pde_damping_layer_f32(const float *in, float *out, const pde_layer_point *pts,
                                     int64_t npts, const float *profile, int32_t nlevels,
                                     int32_t ncomp, const float *farfield, float scale,
                                     void *stream)
{
    if (!in || !out || (npts > 0 && (!pts || !profile)) || npts < 0 || nlevels < 1 ||
        ncomp < 1 || ncomp > PDE_LAYER_MAXC)
        return PDE_EINVAL;
    if (npts == 0) return PDE_OK;
    LayerArgs a;
    for (int c = 0; c < PDE_LAYER_MAXC; ++c) a.ff[c] = (farfield && c < ncomp) ? farfield[c] : 0.0f;
    a.ncomp = ncomp;
    a.has_ff = farfield != nullptr;
    a.scale = scale;
    const unsigned blocks = clamp_grid((npts + 255) / 256, 65536);
    hipLaunchKernelGGL(damping_layer_kernel, dim3(blocks), dim3(256), 0, as_stream(stream), in, out,
                       pts, npts, profile, a);
    return launch_status();
}
