// Node-wise glue between the stencil operators: the element operations the
// reference builds from Warp functions (outer product, scalar x vector,
// matrix x vector, vector / scalar), array arithmetic chains such as
// `dp - u_conv + u_diff` fused into one pass, and the affine map of a scalar
// field.  Every op runs over the whole padded array (ghosts included), as the
// reference's element-wise launches do.  One rounding per + - * / (the build
// uses -ffp-contract=off and IEEE division).
#include "pde_common.hpp"

namespace pde {

// out[n][i][j] = a[n][i] * b[n][j]   (wp.outer)
__global__ __launch_bounds__(256) void outer_kernel(const float *__restrict__ a,
                                                    const float *__restrict__ b,
                                                    float *__restrict__ out, int64_t total,
                                                    int la, int lb)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    const int64_t per = (int64_t)la * lb;
    const int64_t n = t / per;
    const int r = (int)(t - n * per);
    const int i = r / lb, j = r - i * lb;
    out[t] = a[n * la + i] * b[n * lb + j];
}

// out[n][k] = s[n] * b[n][k]
__global__ __launch_bounds__(256) void scalar_mul_kernel(const float *__restrict__ s,
                                                         const float *__restrict__ b,
                                                         float *__restrict__ out, int64_t total,
                                                         int lb)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    out[t] = s[t / lb] * b[t];
}

// out[n][k] = b[n][k] / s[n]
__global__ __launch_bounds__(256) void div_scalar_kernel(const float *__restrict__ b,
                                                         const float *__restrict__ s,
                                                         float *__restrict__ out, int64_t total,
                                                         int lb)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    out[t] = b[t] / s[t / lb];
}

// out[n][i] = m[n][i][0] * v[n][0] + m[n][i][1] * v[n][1] + ...  (columns
// ascending, accumulated left to right)
__global__ __launch_bounds__(256) void matvec_kernel(const float *__restrict__ m,
                                                     const float *__restrict__ v,
                                                     float *__restrict__ out, int64_t total,
                                                     int rows, int cols)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    const int64_t n = t / rows;
    const int i = (int)(t - n * rows);
    const float *mr = m + (n * rows + i) * cols;
    const float *vr = v + n * cols;
    float acc = mr[0] * vr[0];
    for (int j = 1; j < cols; ++j) acc = acc + mr[j] * vr[j];
    out[t] = acc;
}

// out = c0*x0 + c1*x1 + ... (k terms, left to right); a coefficient of +-1
// multiplies exactly, so `a - b + c` is reproduced bit for bit.  out may
// alias any x (same index only).
__global__ __launch_bounds__(256) void lincomb_kernel(const float *x0, const float *x1,
                                                      const float *x2, const float *x3, float c0,
                                                      float c1, float c2, float c3, int k,
                                                      float *out, int64_t n)
{
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t i0 = q * 4;
    if (i0 >= n) return;
    const int64_t i1 = i0 + 4 < n ? i0 + 4 : n;
    for (int64_t i = i0; i < i1; ++i) {
        float acc = c0 * x0[i];
        if (k > 1) acc = acc + c1 * x1[i];
        if (k > 2) acc = acc + c2 * x2[i];
        if (k > 3) acc = acc + c3 * x3[i];
        out[i] = acc;
    }
}

// out[n][c] = scale * (x[n][c] - shift) on component `comp` (all components
// when comp < 0); other components copied.
__global__ __launch_bounds__(256) void affine_kernel(const float *x, float *out, int64_t total,
                                                     int ncomp, int comp, float scale,
                                                     float shift)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    const int c = (int)(t % ncomp);
    const float v = x[t];
    out[t] = (comp < 0 || c == comp) ? scale * (v - shift) : v;
}

static inline bool grid_ok(int64_t total, unsigned *blocks)
{
    const int64_t b = (total + 255) / 256;
    if (b > 0x7fffffff) return false;
    *blocks = (unsigned)b;
    return true;
}

}  // namespace pde

using namespace pde;

extern "C" int pde_outer_f32(const float *a, const float *b, float *out, int64_t nodes,
                             int32_t la, int32_t lb, void *stream)
{
    if (!a || !b || !out || nodes < 0 || la < 1 || lb < 1) return PDE_EINVAL;
    if (out == a || out == b) return PDE_EINVAL;
    const int64_t total = nodes * la * lb;
    if (total == 0) return PDE_OK;
    unsigned nb;
    if (!grid_ok(total, &nb)) return PDE_ESHAPE;
    hipLaunchKernelGGL(outer_kernel, dim3(nb), dim3(256), 0, as_stream(stream), a, b, out, total,
                       la, lb);
    return launch_status();
}

extern "C" int pde_scalar_mul_f32(const float *s, const float *b, float *out, int64_t nodes,
                                  int32_t lb, void *stream)
{
    if (!s || !b || !out || nodes < 0 || lb < 1 || out == s) return PDE_EINVAL;
    const int64_t total = nodes * lb;
    if (total == 0) return PDE_OK;
    unsigned nb;
    if (!grid_ok(total, &nb)) return PDE_ESHAPE;
    hipLaunchKernelGGL(scalar_mul_kernel, dim3(nb), dim3(256), 0, as_stream(stream), s, b, out,
                       total, lb);
    return launch_status();
}

extern "C" int pde_div_scalar_f32(const float *b, const float *s, float *out, int64_t nodes,
                                  int32_t lb, void *stream)
{
    if (!s || !b || !out || nodes < 0 || lb < 1 || out == s) return PDE_EINVAL;
    const int64_t total = nodes * lb;
    if (total == 0) return PDE_OK;
    unsigned nb;
    if (!grid_ok(total, &nb)) return PDE_ESHAPE;
    hipLaunchKernelGGL(div_scalar_kernel, dim3(nb), dim3(256), 0, as_stream(stream), b, s, out,
                       total, lb);
    return launch_status();
}

extern "C" int pde_matvec_f32(const float *m, const float *v, float *out, int64_t nodes,
                              int32_t rows, int32_t cols, void *stream)
{
    if (!m || !v || !out || nodes < 0 || rows < 1 || cols < 1) return PDE_EINVAL;
    if (out == m || out == v) return PDE_EINVAL;
    const int64_t total = nodes * rows;
    if (total == 0) return PDE_OK;
    unsigned nb;
    if (!grid_ok(total, &nb)) return PDE_ESHAPE;
    hipLaunchKernelGGL(matvec_kernel, dim3(nb), dim3(256), 0, as_stream(stream), m, v, out, total,
                       rows, cols);
    return launch_status();
}

extern "C" int pde_lincomb_f32(const float *x0, const float *x1, const float *x2, const float *x3,
                               float c0, float c1, float c2, float c3, int32_t k, float *out,
                               int64_t n, void *stream)
{
    if (k < 1 || k > 4 || !out || n < 0 || !x0) return PDE_EINVAL;
    if ((k > 1 && !x1) || (k > 2 && !x2) || (k > 3 && !x3)) return PDE_EINVAL;
    if (n == 0) return PDE_OK;
    unsigned nb;
    if (!grid_ok((n + 3) / 4, &nb)) return PDE_ESHAPE;
    hipLaunchKernelGGL(lincomb_kernel, dim3(nb), dim3(256), 0, as_stream(stream), x0, x1, x2, x3,
                       c0, c1, c2, c3, k, out, n);
    return launch_status();
}

extern "C" int pde_affine_f32(const float *x, float *out, int64_t nodes, int32_t ncomp,
                              int32_t comp, float scale, float shift, void *stream)
{
    if (!x || !out || nodes < 0 || ncomp < 1 || comp >= ncomp) return PDE_EINVAL;
    const int64_t total = nodes * ncomp;
    if (total == 0) return PDE_OK;
    unsigned nb;
    if (!grid_ok(total, &nb)) return PDE_ESHAPE;
    hipLaunchKernelGGL(affine_kernel, dim3(nb), dim3(256), 0, as_stream(stream), x, out, total,
                       ncomp, comp < 0 ? -1 : comp, scale, shift);
    return launch_status();
}
