// Kernels of the projection-method cavity: the Jacobi sweep of the pressure
// Poisson equation and its transpose, and the transposed node-wise
// matrix-vector product that the convection term's gradient needs.
//
// Jacobi sweep on the interior nodes of a scalar field (1..3 eligible axes):
//   acc = 0; over eligible axes ascending: acc += p[n - e]; acc += p[n + e];
//   out[n] = inv * (acc - h2 * rhs[n]),   inv = 1 / (2 * ndim), h2 = dx^2,
// one Jacobi iteration of (sum of neighbours - 2 ndim p) / dx^2 = rhs.  Ghosts
// are read (apply the pressure boundary first) and never written.
#include "tile2d.hpp"

namespace pde {

struct JacGeom {
    int64_t S[3];          // node strides of canonical axes 0..2
    int64_t E[3], G[3];    // interior extents and starts
    int ndim;
};

static JacGeom jac_geom(const Geom &G)
{
    JacGeom j;
    for (int a = 0; a < 3; ++a) {
        j.S[a] = G.S[a];
        j.E[a] = G.E[a];
        j.G[a] = G.G[a];
    }
    j.ndim = G.ndim;
    return j;
}

__device__ __forceinline__ bool interior(const JacGeom &g, int64_t i0, int64_t i1, int64_t k)
{
    return i0 >= g.G[0] && i0 < g.G[0] + g.E[0] && i1 >= g.G[1] && i1 < g.G[1] + g.E[1] &&
           k >= g.G[2] && k < g.G[2] + g.E[2];
}

// thread per interior node: x over the contiguous axis, y/z loops over the others
__global__ __launch_bounds__(256) void jacobi_kernel(const float *__restrict__ p,
                                                     const float *__restrict__ rhs,
                                                     float *__restrict__ out, JacGeom g, float h2,
                                                     float inv)
{
    const int64_t k = g.G[2] + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= g.G[2] + g.E[2]) return;
    for (int64_t i0 = g.G[0] + blockIdx.z; i0 < g.G[0] + g.E[0]; i0 += gridDim.z)
        for (int64_t i1 = g.G[1] + blockIdx.y; i1 < g.G[1] + g.E[1]; i1 += gridDim.y) {
            const int64_t e = i0 * g.S[0] + i1 * g.S[1] + k;
            float acc = 0.0f;
            for (int a = 3 - g.ndim; a < 3; ++a) {
                acc += p[e - g.S[a]];
                acc += p[e + g.S[a]];
            }
            out[e] = inv * (acc - h2 * rhs[e]);
        }
}

// The sweep of a field whose ghosts a Neumann-0 legacy boundary would fill
// first ('ALL', value 0: ghost b - s e = coef * 0 + p[b + s e], coef = -2 dx
// on '-' faces and +2 dx on '+' faces): a neighbour in the ghost layer is read
// as (-0 or +0) + its mirror across the wall, the same FP32 value the boundary
// writes, so one launch replaces boundary + sweep and p's ghosts are not read.
__global__ __launch_bounds__(256) void jacobi_neumann0_kernel(const float *__restrict__ p,
                                                              const float *__restrict__ rhs,
                                                              float *__restrict__ out, JacGeom g,
                                                              float h2, float inv)
{
    const int64_t k = g.G[2] + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= g.G[2] + g.E[2]) return;
    for (int64_t i0 = g.G[0] + blockIdx.z; i0 < g.G[0] + g.E[0]; i0 += gridDim.z)
        for (int64_t i1 = g.G[1] + blockIdx.y; i1 < g.G[1] + g.E[1]; i1 += gridDim.y) {
            const int64_t e = i0 * g.S[0] + i1 * g.S[1] + k;
            const int64_t c[3] = {i0, i1, k};
            float acc = 0.0f;
            for (int a = 3 - g.ndim; a < 3; ++a) {
                const float lo = c[a] - 1 < g.G[a] ? -0.0f + p[e + g.S[a]] : p[e - g.S[a]];
                const float hi = c[a] + 1 >= g.G[a] + g.E[a] ? 0.0f + p[e - g.S[a]] : p[e + g.S[a]];
                acc += lo;
                acc += hi;
            }
            out[e] = inv * (acc - h2 * rhs[e]);
        }
}

// Transpose: for every padded node m, acc over axes ascending of g[m + e]
// (m is the -1 neighbour of m + e) then g[m - e], each when that node is
// interior; gp[m] = inv * acc.  grhs[n] = -h2 * (inv * g[n]) on interior n,
// 0 on ghosts.
__global__ __launch_bounds__(256) void jacobi_adjoint_kernel(const float *__restrict__ gout,
                                                             float *__restrict__ gp,
                                                             float *__restrict__ grhs, JacGeom g,
                                                             int64_t P0, int64_t P1, int64_t P2,
                                                             float h2, float inv)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= P2) return;
    for (int64_t i0 = blockIdx.z; i0 < P0; i0 += gridDim.z)
        for (int64_t i1 = blockIdx.y; i1 < P1; i1 += gridDim.y) {
            const int64_t e = i0 * g.S[0] + i1 * g.S[1] + k;
            const int64_t c[3] = {i0, i1, k};
            float acc = 0.0f;
            for (int a = 3 - g.ndim; a < 3; ++a) {
                int64_t q[3] = {c[0], c[1], c[2]};
                q[a] = c[a] + 1;
                if (interior(g, q[0], q[1], q[2])) acc += gout[e + g.S[a]];
                q[a] = c[a] - 1;
                if (interior(g, q[0], q[1], q[2])) acc += gout[e - g.S[a]];
            }
            gp[e] = inv * acc;
            if (grhs) grhs[e] = interior(g, i0, i1, k) ? -h2 * (inv * gout[e]) : 0.0f;
        }
}

// Transpose of the folded sweep (jacobi_neumann0_kernel) as a map of p's
// interior and rhs: thread per padded node m.  On interior m, acc over axes
// ascending of g[m + e] then g[m - e] (each when interior), gp[m] = inv * acc,
// then per axis ascending the wall mirrors: + inv * g[m - e] when m - e is
// interior and m - 2e a ghost (the low wall's ghost of m - e mirrors m), then
// + inv * g[m + e] when m + e is interior and m + 2e a ghost.  That is the
// order of the unfolded transpose (pde_jacobi_adjoint_f32, whose ghost
// entries inv * g[wall node] the boundary's transpose then adds to the
// mirror).  gp = 0 on ghosts; grhs[m] += -h2 * (inv * g[m]) on interior m.
__global__ __launch_bounds__(256) void jacobi_neumann0_adjoint_kernel(
    const float *__restrict__ gout, float *__restrict__ gp, float *__restrict__ grhs, JacGeom g,
    int64_t P0, int64_t P1, int64_t P2, float h2, float inv)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= P2) return;
    for (int64_t i0 = blockIdx.z; i0 < P0; i0 += gridDim.z)
        for (int64_t i1 = blockIdx.y; i1 < P1; i1 += gridDim.y) {
            const int64_t e = i0 * g.S[0] + i1 * g.S[1] + k;
            if (!interior(g, i0, i1, k)) {
                gp[e] = 0.0f;
                continue;
            }
            const int64_t c[3] = {i0, i1, k};
            float acc = 0.0f;
            for (int a = 3 - g.ndim; a < 3; ++a) {
                if (c[a] + 1 < g.G[a] + g.E[a]) acc += gout[e + g.S[a]];
                if (c[a] - 1 >= g.G[a]) acc += gout[e - g.S[a]];
            }
            float v = inv * acc;
            for (int a = 3 - g.ndim; a < 3; ++a) {
                if (c[a] - 1 >= g.G[a] && c[a] - 2 < g.G[a]) v = v + inv * gout[e - g.S[a]];
                if (c[a] + 1 < g.G[a] + g.E[a] && c[a] + 2 >= g.G[a] + g.E[a])
                    v = v + inv * gout[e + g.S[a]];
            }
            gp[e] = v;
            grhs[e] = grhs[e] + -h2 * (inv * gout[e]);
        }
}

// 2D grids (and batches of them, one instance per blockIdx.y): the folded
// sweep and its transpose as row tiles.  A lane owns 4 consecutive columns
// (16-byte accesses) of JR rows; a block stacks JBY waves along the rows and
// covers a 256-column strip; the rows above and below come from the lane's
// own loads, the columns beside its 4 from its neighbours by shuffles (the
// strip's edge lanes load theirs).  Tile order as in the fused 2D steps
// (vertically adjacent tiles on one XCD).  Same arithmetic, in the same
// order, as jacobi_neumann0_kernel / jacobi_neumann0_adjoint_kernel.
constexpr int JR = 2;
constexpr int JBY = 4;

struct JTile {
    Tile2 t;
    int64_t stride;   // floats per instance
};

__device__ __forceinline__ uint32_t jtile_index()
{
    const uint32_t per = gridDim.x >> 3;
    return (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
}

__global__ __launch_bounds__(64 * JBY) void jacobi_n0_rows_kernel(const float *__restrict__ p,
                                                                  const float *__restrict__ rhs,
                                                                  float *__restrict__ out,
                                                                  JTile J, float h2, float inv)
{
    const Tile2 &t = J.t;
    const uint32_t L = jtile_index();
    if (L >= t.tq * t.tr) return;
    const int64_t off = (int64_t)blockIdx.y * J.stride;
    p += off;
    rhs += off;
    out += off;
    const uint32_t qt = L / t.tr, rt = L % t.tr;
    const int lane = threadIdx.x;
    const int64_t q = (int64_t)qt * 64 + lane;
    const bool lane_ok = q < t.nq;
    const int64_t k0 = (lane_ok ? q : t.nq - 1) * 4;
    const int64_t x0 = ((int64_t)rt * JBY + threadIdx.y) * JR;
    const int64_t lo1 = t.g, hi1 = t.P1 - t.g, lo2 = t.g, hi2 = t.P2 - t.g;
    v4 rows[JR + 2];
#pragma unroll
    for (int r = 0; r < JR + 2; ++r) rows[r] = ld_row4(p, t, x0 - 1 + r, k0);
#pragma unroll
    for (int r = 0; r < JR; ++r) {
        const int64_t x = x0 + r;
        float left, right;
        col_edges(p, t, x, k0, rows[r + 1], lane, 64, left, right);
        if (!lane_ok || x < lo1 || x >= hi1) continue;
        const v4 rh = ld_row4(rhs, t, x, k0);
        v4 o = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t k = k0 + i;
            const float up = el(rows[r], i), dn = el(rows[r + 2], i);
            const float zl = i == 0 ? left : el(rows[r + 1], i > 0 ? i - 1 : 0);
            const float zr = i == 3 ? right : el(rows[r + 1], i < 3 ? i + 1 : 3);
            const float a_lo = x - 1 < lo1 ? -0.0f + dn : up;
            const float a_hi = x + 1 >= hi1 ? 0.0f + up : dn;
            const float b_lo = k - 1 < lo2 ? -0.0f + zr : zl;
            const float b_hi = k + 1 >= hi2 ? 0.0f + zl : zr;
            float acc = 0.0f;
            acc += a_lo;
            acc += a_hi;
            acc += b_lo;
            acc += b_hi;
            set_el(o, i, inv * (acc - h2 * el(rh, i)));
        }
        if (k0 >= lo2 && k0 + 4 <= hi2) {
            st_row4(out, t, x, k0, o);
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (k0 + i >= lo2 && k0 + i < hi2) out[x * t.P2 + k0 + i] = el(o, i);
        }
    }
}

__global__ __launch_bounds__(64 * JBY) void jacobi_n0_adj_rows_kernel(
    const float *__restrict__ g, float *__restrict__ gp, float *__restrict__ grhs, JTile J,
    float h2, float inv)
{
    const Tile2 &t = J.t;
    const uint32_t L = jtile_index();
    if (L >= t.tq * t.tr) return;
    const int64_t off = (int64_t)blockIdx.y * J.stride;
    g += off;
    gp += off;
    grhs += off;
    const uint32_t qt = L / t.tr, rt = L % t.tr;
    const int lane = threadIdx.x;
    const int64_t q = (int64_t)qt * 64 + lane;
    const bool lane_ok = q < t.nq;
    const int64_t k0 = (lane_ok ? q : t.nq - 1) * 4;
    const int64_t x0 = ((int64_t)rt * JBY + threadIdx.y) * JR;
    const int64_t lo1 = t.g, hi1 = t.P1 - t.g, lo2 = t.g, hi2 = t.P2 - t.g;
    v4 rows[JR + 2];
#pragma unroll
    for (int r = 0; r < JR + 2; ++r) rows[r] = ld_row4(g, t, x0 - 1 + r, k0);
#pragma unroll
    for (int r = 0; r < JR; ++r) {
        const int64_t x = x0 + r;
        float left, right;
        col_edges(g, t, x, k0, rows[r + 1], lane, 64, left, right);
        if (!lane_ok || x >= t.P1) continue;
        const bool xin = x >= lo1 && x < hi1;
        v4 o = {0.0f, 0.0f, 0.0f, 0.0f};
        v4 gr = {0.0f, 0.0f, 0.0f, 0.0f};
        if (xin) gr = ld_row4(grhs, t, x, k0);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t k = k0 + i;
            if (!xin || k < lo2 || k >= hi2) continue;
            const float up = el(rows[r], i), mid = el(rows[r + 1], i), dn = el(rows[r + 2], i);
            const float zl = i == 0 ? left : el(rows[r + 1], i > 0 ? i - 1 : 0);
            const float zr = i == 3 ? right : el(rows[r + 1], i < 3 ? i + 1 : 3);
            float acc = 0.0f;
            if (x + 1 < hi1) acc += dn;
            if (x - 1 >= lo1) acc += up;
            if (k + 1 < hi2) acc += zr;
            if (k - 1 >= lo2) acc += zl;
            float v = inv * acc;
            if (x - 1 >= lo1 && x - 2 < lo1) v = v + inv * up;
            if (x + 1 < hi1 && x + 2 >= hi1) v = v + inv * dn;
            if (k - 1 >= lo2 && k - 2 < lo2) v = v + inv * zl;
            if (k + 1 < hi2 && k + 2 >= hi2) v = v + inv * zr;
            set_el(o, i, v);
            set_el(gr, i, el(gr, i) + -h2 * (inv * mid));
        }
        st_row4(gp, t, x, k0, o);
        if (xin) {
            if (k0 >= lo2 && k0 + 4 <= hi2) {
                st_row4(grhs, t, x, k0, gr);
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (k0 + i >= lo2 && k0 + i < hi2) grhs[x * t.P2 + k0 + i] = el(gr, i);
            }
        }
    }
}

// Two folded sweeps in one pass: out = J(J(p)), the same values as two
// launches.  The intermediate sweep lives in registers: a lane computes it on
// JR + 2 rows (its own and one above and below) of its 4 columns, and the
// strips overlap by one lane on each side (lanes 0 and 63 compute only the
// intermediate), so every column neighbour of the second sweep comes from a
// shuffle.  A strip covers 62 * 4 = 248 output columns.
constexpr int J2_COLS = 248;

__device__ __forceinline__ v4 j2_load(const float *__restrict__ f, const Tile2 &t, int64_t x,
                                      int64_t k0)
{
    const v4 z = {0.0f, 0.0f, 0.0f, 0.0f};
    return (k0 < 0 || k0 >= t.P2) ? z : ld_row4(f, t, x, k0);
}

// One folded sweep at the 4 columns of a lane (rows up / mid / dn, column
// neighbours zl of element 0 and zr of element 3), interior nodes only.
__device__ __forceinline__ v4 j_sweep4(const v4 &up, const v4 &mid, const v4 &dn, float zl0,
                                       float zr3, const v4 &rh, int64_t x, int64_t k0,
                                       int64_t lo1, int64_t hi1, int64_t lo2, int64_t hi2,
                                       float h2, float inv)
{
    v4 o = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t k = k0 + i;
        if (x < lo1 || x >= hi1 || k < lo2 || k >= hi2) continue;
        const float u = el(up, i), d = el(dn, i);
        const float zl = i == 0 ? zl0 : el(mid, i > 0 ? i - 1 : 0);
        const float zr = i == 3 ? zr3 : el(mid, i < 3 ? i + 1 : 3);
        const float a_lo = x - 1 < lo1 ? -0.0f + d : u;
        const float a_hi = x + 1 >= hi1 ? 0.0f + u : d;
        const float b_lo = k - 1 < lo2 ? -0.0f + zr : zl;
        const float b_hi = k + 1 >= hi2 ? 0.0f + zl : zr;
        float acc = 0.0f;
        acc += a_lo;
        acc += a_hi;
        acc += b_lo;
        acc += b_hi;
        set_el(o, i, inv * (acc - h2 * el(rh, i)));
    }
    return o;
}

__global__ __launch_bounds__(64 * JBY) void jacobi_n0_rows2_kernel(const float *__restrict__ p,
                                                                   const float *__restrict__ rhs,
                                                                   float *__restrict__ out,
                                                                   JTile J, float h2, float inv)
{
    const Tile2 &t = J.t;
    const uint32_t L = jtile_index();
    if (L >= t.tq * t.tr) return;
    const int64_t off = (int64_t)blockIdx.y * J.stride;
    p += off;
    rhs += off;
    out += off;
    const uint32_t qt = L / t.tr, rt = L % t.tr;
    const int lane = threadIdx.x;
    const int64_t k0 = (int64_t)qt * J2_COLS + 4 * (int64_t)(lane - 1);
    const int64_t x0 = ((int64_t)rt * JBY + threadIdx.y) * JR;
    const int64_t lo1 = t.g, hi1 = t.P1 - t.g, lo2 = t.g, hi2 = t.P2 - t.g;
    v4 pr[JR + 4], rh[JR + 2];
#pragma unroll
    for (int r = 0; r < JR + 4; ++r) pr[r] = j2_load(p, t, x0 - 2 + r, k0);
#pragma unroll
    for (int r = 0; r < JR + 2; ++r) rh[r] = j2_load(rhs, t, x0 - 1 + r, k0);
    // the intermediate sweep on rows x0-1 .. x0+JR; lane 0's left and lane
    // 63's right column neighbour from memory (0 outside the field)
    v4 q1[JR + 2];
#pragma unroll
    for (int r = 0; r < JR + 2; ++r) {
        const int64_t x = x0 - 1 + r;
        float zl = __shfl_up(pr[r + 1].w, 1, 64), zr = __shfl_down(pr[r + 1].x, 1, 64);
        if (lane == 0 || lane == 63) {
            const int64_t k = lane == 0 ? k0 - 1 : k0 + 4;
            const float v = (k >= 0 && k < t.P2 && x >= 0 && x < t.P1) ? p[x * t.P2 + k] : 0.0f;
            if (lane == 0) zl = v;
            else zr = v;
        }
        q1[r] = j_sweep4(pr[r], pr[r + 1], pr[r + 2], zl, zr, rh[r], x, k0, lo1, hi1, lo2, hi2,
                         h2, inv);
    }
    // the second sweep on rows x0 .. x0+JR-1, lanes 1 .. 62
#pragma unroll
    for (int r = 0; r < JR; ++r) {
        const int64_t x = x0 + r;
        const float zl = __shfl_up(q1[r + 1].w, 1, 64), zr = __shfl_down(q1[r + 1].x, 1, 64);
        const v4 o = j_sweep4(q1[r], q1[r + 1], q1[r + 2], zl, zr, rh[r + 1], x, k0, lo1, hi1,
                              lo2, hi2, h2, inv);
        if (lane == 0 || lane == 63 || x < lo1 || x >= hi1 || k0 >= t.P2) continue;
        if (k0 >= lo2 && k0 + 4 <= hi2) {
            st_row4(out, t, x, k0, o);
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (k0 + i >= lo2 && k0 + i < hi2) out[x * t.P2 + k0 + i] = el(o, i);
        }
    }
}

// One transposed folded sweep at the 4 columns of a lane (gp; 0 off the
// interior), as jacobi_neumann0_adjoint_kernel computes it.
__device__ __forceinline__ v4 j_adj4(const v4 &up, const v4 &mid, const v4 &dn, float zl0,
                                     float zr3, int64_t x, int64_t k0, int64_t lo1, int64_t hi1,
                                     int64_t lo2, int64_t hi2, float inv)
{
    v4 o = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t k = k0 + i;
        if (x < lo1 || x >= hi1 || k < lo2 || k >= hi2) continue;
        const float u = el(up, i), d = el(dn, i);
        const float zl = i == 0 ? zl0 : el(mid, i > 0 ? i - 1 : 0);
        const float zr = i == 3 ? zr3 : el(mid, i < 3 ? i + 1 : 3);
        float acc = 0.0f;
        if (x + 1 < hi1) acc += d;
        if (x - 1 >= lo1) acc += u;
        if (k + 1 < hi2) acc += zr;
        if (k - 1 >= lo2) acc += zl;
        float v = inv * acc;
        if (x - 1 >= lo1 && x - 2 < lo1) v = v + inv * u;
        if (x + 1 < hi1 && x + 2 >= hi1) v = v + inv * d;
        if (k - 1 >= lo2 && k - 2 < lo2) v = v + inv * zl;
        if (k + 1 < hi2 && k + 2 >= hi2) v = v + inv * zr;
        set_el(o, i, v);
    }
    return o;
}

// Two transposed sweeps in one pass (the same overlapping strips as
// jacobi_n0_rows2_kernel): gp = T(T(g)), grhs += (rhs part of T at g), then
// += (that at T(g)), in that order: the values of two launches.
__global__ __launch_bounds__(64 * JBY) void jacobi_n0_adj_rows2_kernel(
    const float *__restrict__ g, float *__restrict__ gp, float *__restrict__ grhs, JTile J,
    float h2, float inv)
{
    const Tile2 &t = J.t;
    const uint32_t L = jtile_index();
    if (L >= t.tq * t.tr) return;
    const int64_t off = (int64_t)blockIdx.y * J.stride;
    g += off;
    gp += off;
    grhs += off;
    const uint32_t qt = L / t.tr, rt = L % t.tr;
    const int lane = threadIdx.x;
    const int64_t k0 = (int64_t)qt * J2_COLS + 4 * (int64_t)(lane - 1);
    const int64_t x0 = ((int64_t)rt * JBY + threadIdx.y) * JR;
    const int64_t lo1 = t.g, hi1 = t.P1 - t.g, lo2 = t.g, hi2 = t.P2 - t.g;
    v4 gr[JR + 4];
#pragma unroll
    for (int r = 0; r < JR + 4; ++r) gr[r] = j2_load(g, t, x0 - 2 + r, k0);
    v4 t1[JR + 2];
#pragma unroll
    for (int r = 0; r < JR + 2; ++r) {
        const int64_t x = x0 - 1 + r;
        float zl = __shfl_up(gr[r + 1].w, 1, 64), zr = __shfl_down(gr[r + 1].x, 1, 64);
        if (lane == 0 || lane == 63) {
            const int64_t k = lane == 0 ? k0 - 1 : k0 + 4;
            const float v = (k >= 0 && k < t.P2 && x >= 0 && x < t.P1) ? g[x * t.P2 + k] : 0.0f;
            if (lane == 0) zl = v;
            else zr = v;
        }
        t1[r] = j_adj4(gr[r], gr[r + 1], gr[r + 2], zl, zr, x, k0, lo1, hi1, lo2, hi2, inv);
    }
#pragma unroll
    for (int r = 0; r < JR; ++r) {
        const int64_t x = x0 + r;
        const float zl = __shfl_up(t1[r + 1].w, 1, 64), zr = __shfl_down(t1[r + 1].x, 1, 64);
        const v4 o = j_adj4(t1[r], t1[r + 1], t1[r + 2], zl, zr, x, k0, lo1, hi1, lo2, hi2, inv);
        if (lane == 0 || lane == 63 || x >= t.P1 || k0 >= t.P2) continue;
        st_row4(gp, t, x, k0, o);
        if (x < lo1 || x >= hi1) continue;
        v4 acc = ld_row4(grhs, t, x, k0);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t k = k0 + i;
            if (k < lo2 || k >= hi2) continue;
            float a = el(acc, i) + -h2 * (inv * el(gr[r + 2], i));
            a = a + -h2 * (inv * el(t1[r + 1], i));
            set_el(acc, i, a);
        }
        if (k0 >= lo2 && k0 + 4 <= hi2) {
            st_row4(grhs, t, x, k0, acc);
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (k0 + i >= lo2 && k0 + i < hi2) grhs[x * t.P2 + k0 + i] = el(acc, i);
        }
    }
}

template <bool ADJ>
static int jacobi_rows2_launch(const float *p, const float *rhs, float *out, float *grhs,
                               int64_t batch, int64_t P1, int64_t P2, int64_t ghost, float h2,
                               float inv, hipStream_t s)
{
    JTile J;
    J.t.P1 = P1;
    J.t.P2 = P2;
    J.t.g = ghost;
    J.t.nq = (P2 + 3) / 4;
    J.t.tq = (uint32_t)((P2 + J2_COLS - 1) / J2_COLS);
    J.t.tr = (uint32_t)((P1 + JBY * JR - 1) / (JBY * JR));
    J.stride = P1 * P2;
    const int64_t nb = ((int64_t)J.t.tq * J.t.tr + 7) / 8 * 8;
    if (nb > 0x7fffffff || batch > 65535) return PDE_ESHAPE;
    if (ADJ)
        hipLaunchKernelGGL(jacobi_n0_adj_rows2_kernel, dim3((unsigned)nb, (unsigned)batch),
                           dim3(64, JBY), 0, s, p, out, grhs, J, h2, inv);
    else
        hipLaunchKernelGGL(jacobi_n0_rows2_kernel, dim3((unsigned)nb, (unsigned)batch),
                           dim3(64, JBY), 0, s, p, rhs, out, J, h2, inv);
    return launch_status();
}

// Launch of either row-tile kernel over `batch` instances of a 2D grid
// (padded extents P1 x P2); PDE_ESHAPE beyond the grid limits.
template <bool ADJ>
static int jacobi_rows_launch(const float *a, float *b, float *c, const float *rhs,
                              int64_t batch, int64_t P1, int64_t P2, int64_t ghost, float h2,
                              float inv, hipStream_t s)
{
    JTile J;
    J.t.P1 = P1;
    J.t.P2 = P2;
    J.t.g = ghost;
    J.t.nq = (P2 + 3) / 4;
    J.t.tq = (uint32_t)((J.t.nq + 63) / 64);
    J.t.tr = (uint32_t)((P1 + JBY * JR - 1) / (JBY * JR));
    J.stride = P1 * P2;
    const int64_t nb = ((int64_t)J.t.tq * J.t.tr + 7) / 8 * 8;
    if (nb > 0x7fffffff || batch > 65535) return PDE_ESHAPE;
    const dim3 grid((unsigned)nb, (unsigned)batch), block(64, JBY);
    if (ADJ)
        hipLaunchKernelGGL(jacobi_n0_adj_rows_kernel, grid, block, 0, s, a, b, c, J, h2, inv);
    else
        hipLaunchKernelGGL(jacobi_n0_rows_kernel, grid, block, 0, s, a, rhs, b, J, h2, inv);
    return launch_status();
}

// The cavity's predictor in one pass over a 2D vec2 field (thread per padded
// node): us = u2 + dt * (-(grad u2) u2 + nu * lap(u2)), each piece with the
// operation order of its own kernel (grad_generic: per component, the row
// axis then the column axis, acc = 0, += wm u[n-e], += wc u[n], += wp u[n+e],
// times 1; matvec_kernel: m0 v0 then + m1 v1; lap_generic<2>: six terms then
// times nu; lincomb_kernel: (-1) conv then + 1 diff; euler_kernel: u2 + dt F).
// Off the interior those kernels leave grad and lap at 0, so there
// us = u2 + dt * ((-1) (0 u0 + 0 u1) + 1 * 0), computed the same way.
__global__ __launch_bounds__(256) void cavity_predict_kernel(
    const float *__restrict__ u2, float *__restrict__ us, int64_t P1, int64_t P2, int64_t g,
    float lm, float lc, float lp, float gm, float gc, float gp, float nu, float dt)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= P2) return;
    const int64_t nb = (int64_t)blockIdx.z * P1 * P2;   // nodes of the instances before
    u2 += 2 * nb;
    us += 2 * nb;
    const int64_t S1 = 2 * P2, S2 = 2;
    for (int64_t x = blockIdx.y; x < P1; x += gridDim.y) {
        const int64_t e = (x * P2 + k) * 2;
        const float v0 = u2[e], v1 = u2[e + 1];
        const bool in = x >= g && x < P1 - g && k >= g && k < P2 - g;
        float du[2][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}}, diff[2] = {0.0f, 0.0f};
        if (in) {
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const float ctr = u2[e + c];
                const float xm = u2[e + c - S1], xp = u2[e + c + S1];
                const float ym = u2[e + c - S2], yp = u2[e + c + S2];
                float a = 0.0f;
                a += gm * xm;
                a += gc * ctr;
                a += gp * xp;
                du[c][0] = 1.0f * a;
                a = 0.0f;
                a += gm * ym;
                a += gc * ctr;
                a += gp * yp;
                du[c][1] = 1.0f * a;
                a = 0.0f;
                a += lm * xm;
                a += lc * ctr;
                a += lp * xp;
                a += lm * ym;
                a += lc * ctr;
                a += lp * yp;
                diff[c] = nu * a;
            }
        }
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            float conv = du[c][0] * v0;
            conv = conv + du[c][1] * v1;
            float F = -1.0f * conv;
            F = F + 1.0f * diff[c];
            us[e + c] = u2[e + c] + dt * F;
        }
    }
}

// The cavity's correction in one pass (thread per padded node):
// u_out = us2 + (-dt) * grad(B_p(p)) with B_p the legacy Neumann-0 boundary
// on every face folded in (a face-ghost neighbour of an interior node reads
// (-0 or +0) + its mirror, the value the boundary writes; see
// jacobi_neumann0_kernel), grad with grad_generic's order (per axis, acc = 0,
// += wm p[n-e], += wc p[n], += wp p[n+e], times 1) and 0 off the interior,
// and the update with euler_kernel's (us2 + (-dt) * gp) on every node.
__global__ __launch_bounds__(256) void cavity_correct_kernel(
    const float *__restrict__ p, const float *__restrict__ us2, float *__restrict__ u_out,
    int64_t P1, int64_t P2, int64_t g, float gm, float gc, float gp, float ndt)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= P2) return;
    const int64_t nb = (int64_t)blockIdx.z * P1 * P2;
    p += nb;
    us2 += 2 * nb;
    u_out += 2 * nb;
    for (int64_t x = blockIdx.y; x < P1; x += gridDim.y) {
        const int64_t n = x * P2 + k;
        float d[2] = {0.0f, 0.0f};
        if (x >= g && x < P1 - g && k >= g && k < P2 - g) {
            const float pc = p[n];
            const float xm = x - 1 < g ? -0.0f + p[n + P2] : p[n - P2];
            const float xp = x + 1 >= P1 - g ? 0.0f + p[n - P2] : p[n + P2];
            const float ym = k - 1 < g ? -0.0f + p[n + 1] : p[n - 1];
            const float yp = k + 1 >= P2 - g ? 0.0f + p[n - 1] : p[n + 1];
            float a = 0.0f;
            a += gm * xm;
            a += gc * pc;
            a += gp * xp;
            d[0] = 1.0f * a;
            a = 0.0f;
            a += gm * ym;
            a += gc * pc;
            a += gp * yp;
            d[1] = 1.0f * a;
        }
        u_out[2 * n] = us2[2 * n] + ndt * d[0];
        u_out[2 * n + 1] = us2[2 * n + 1] + ndt * d[1];
    }
}

// Transpose of the predictor's interior part in one pass (thread per padded
// node of a 2D vec2 field): from gus (the gradient of us) and u2,
//   gF = dt gus, gconv = -1 gF, glap = nu * lap^T(gF), gdu = outer(gconv, u2),
//   gmv = (grad u2)^T gconv, ggr = 1 * grad^T(gdu),
//   gu2 = 1 gus + 1 glap + 1 gmv + 1 ggr,
// each with the operation order of its own kernel (scale_kernel,
// lap_adj_kernel / grad_adj_kernel's gathers with interior masks,
// outer_kernel, matvec_t_kernel, lincomb_kernel); grad u2 is recomputed with
// grad_generic's order and is 0 off the interior.
__global__ __launch_bounds__(256) void cavity_adj_predict_kernel(
    const float *__restrict__ gus, const float *__restrict__ u2, float *__restrict__ gu2,
    int64_t P1, int64_t P2, int64_t g, float lm, float lc, float lp, float gm, float gc,
    float gp, float nu, float dt)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= P2) return;
    const int64_t nb = (int64_t)blockIdx.z * P1 * P2;
    gus += 2 * nb;
    u2 += 2 * nb;
    gu2 += 2 * nb;
    typedef float f2 __attribute__((ext_vector_type(2)));
    const f2 *G2 = reinterpret_cast<const f2 *>(gus);
    const f2 *U2 = reinterpret_cast<const f2 *>(u2);
    auto inx = [&](int64_t v) { return v >= g && v < P1 - g; };
    auto ink = [&](int64_t v) { return v >= g && v < P2 - g; };
    const int64_t km = k > 0 ? k - 1 : k, kp = k + 1 < P2 ? k + 1 : k;
    for (int64_t x = blockIdx.y; x < P1; x += gridDim.y) {
        const bool ix = inx(x), iy = ink(k);
        const bool xp_in = inx(x + 1), xm_in = inx(x - 1), kp_in = ink(k + 1), km_in = ink(k - 1);
        // the node and its four neighbours, each loaded once (clamped to the
        // array; a value is used only where the node it stands for is interior)
        const int64_t xm = x > 0 ? x - 1 : x, xq = x + 1 < P1 ? x + 1 : x;
        const int64_t n = x * P2 + k;
        const f2 g0 = G2[n], gxm = G2[xm * P2 + k], gxp = G2[xq * P2 + k];
        const f2 gym = G2[x * P2 + km], gyp = G2[x * P2 + kp];
        const f2 u0 = U2[n], uxm = U2[xm * P2 + k], uxp = U2[xq * P2 + k];
        const f2 uym = U2[x * P2 + km], uyp = U2[x * P2 + kp];
        auto cmp = [](const f2 &v, int c) { return c == 0 ? v.x : v.y; };
        // forward grad u2 here (du[c][t]), 0 off the interior
        float du[2][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}};
        if (ix && iy) {
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                float a = 0.0f;
                a += gm * cmp(uxm, c);
                a += gc * cmp(u0, c);
                a += gp * cmp(uxp, c);
                du[c][0] = 1.0f * a;
                a = 0.0f;
                a += gm * cmp(uym, c);
                a += gc * cmp(u0, c);
                a += gp * cmp(uyp, c);
                du[c][1] = 1.0f * a;
            }
        }
        // gF = dt gus, gconv = -1 gF, gdu(c, t) = gconv(c) u2(t), at any node
        auto gF = [&](const f2 &gv, int c) { return dt * cmp(gv, c); };
        auto gcv = [&](const f2 &gv, int c) { return -1.0f * gF(gv, c); };
        const float gc0 = gcv(g0, 0), gc1 = gcv(g0, 1);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            float al = 0.0f;
            if (iy) {
                if (xp_in) al += lm * gF(gxp, c);
                if (ix) al += lc * gF(g0, c);
                if (xm_in) al += lp * gF(gxm, c);
            }
            if (ix) {
                if (kp_in) al += lm * gF(gyp, c);
                if (iy) al += lc * gF(g0, c);
                if (km_in) al += lp * gF(gym, c);
            }
            const float glap = nu * al;
            float gmv = du[0][c] * gc0;
            gmv = gmv + du[1][c] * gc1;
            float ag = 0.0f;
            if (iy) {
                if (xp_in) ag += gm * (gcv(gxp, c) * uxp.x);
                if (ix) ag += gc * (gcv(g0, c) * u0.x);
                if (xm_in) ag += gp * (gcv(gxm, c) * uxm.x);
            }
            if (ix) {
                if (kp_in) ag += gm * (gcv(gyp, c) * uyp.y);
                if (iy) ag += gc * (gcv(g0, c) * u0.y);
                if (km_in) ag += gp * (gcv(gym, c) * uym.y);
            }
            const float ggr = 1.0f * ag;
            float acc = 1.0f * cmp(g0, c);
            acc = acc + 1.0f * glap;
            acc = acc + 1.0f * gmv;
            acc = acc + 1.0f * ggr;
            gu2[2 * n + c] = acc;
        }
    }
}

// Transpose of the correction's pressure part, up to the boundary's
// transpose: gpf = 1 * grad^T((-dt) gu_out) on a 2D scalar pressure field,
// with scale_kernel's product and grad_adj_kernel's gather (N = 1) order.
__global__ __launch_bounds__(256) void cavity_adj_correct_kernel(
    const float *__restrict__ gu_out, float *__restrict__ gpf, int64_t P1, int64_t P2, int64_t g,
    float gm, float gc, float gp, float ndt)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= P2) return;
    const int64_t nb = (int64_t)blockIdx.z * P1 * P2;
    gu_out += 2 * nb;
    gpf += nb;
    auto inx = [&](int64_t v) { return v >= g && v < P1 - g; };
    auto ink = [&](int64_t v) { return v >= g && v < P2 - g; };
    auto gg = [&](int64_t node, int t) { return ndt * gu_out[2 * node + t]; };
    for (int64_t x = blockIdx.y; x < P1; x += gridDim.y) {
        const int64_t n = x * P2 + k;
        const bool ix = inx(x), iy = ink(k);
        float acc = 0.0f;
        if (iy) {
            if (inx(x + 1)) acc += gm * gg(n + P2, 0);
            if (ix) acc += gc * gg(n, 0);
            if (inx(x - 1)) acc += gp * gg(n - P2, 0);
        }
        if (ix) {
            if (ink(k + 1)) acc += gm * gg(n + 1, 1);
            if (iy) acc += gc * gg(n, 1);
            if (ink(k - 1)) acc += gp * gg(n - 1, 1);
        }
        gpf[n] = 1.0f * acc;
    }
}

// gus2 = 1 gu_out + 1 div^T(grhs) (alpha = 1/dt) on a 2D vec2 field, with
// div_adj_kernel's order (rows = 1) and lincomb_kernel's sum.
__global__ __launch_bounds__(256) void cavity_adj_div_kernel(
    const float *__restrict__ grhs, const float *__restrict__ gu_out, float *__restrict__ gus2,
    int64_t P1, int64_t P2, int64_t g, float gm, float gc, float gp, float alpha)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= P2) return;
    const int64_t nb = (int64_t)blockIdx.z * P1 * P2;
    grhs += nb;
    gu_out += 2 * nb;
    gus2 += 2 * nb;
    auto inx = [&](int64_t v) { return v >= g && v < P1 - g; };
    auto ink = [&](int64_t v) { return v >= g && v < P2 - g; };
    for (int64_t x = blockIdx.y; x < P1; x += gridDim.y) {
        const int64_t n = x * P2 + k;
        const bool ix = inx(x), iy = ink(k);
        float a0 = 0.0f, a1 = 0.0f;
        if (iy) {
            if (inx(x + 1)) a0 += gm * grhs[n + P2];
            if (ix) a0 += gc * grhs[n];
            if (inx(x - 1)) a0 += gp * grhs[n - P2];
        }
        if (ix) {
            if (ink(k + 1)) a1 += gm * grhs[n + 1];
            if (iy) a1 += gc * grhs[n];
            if (ink(k - 1)) a1 += gp * grhs[n - 1];
        }
        float v0 = 1.0f * gu_out[2 * n];
        v0 = v0 + 1.0f * (alpha * a0);
        float v1 = 1.0f * gu_out[2 * n + 1];
        v1 = v1 + 1.0f * (alpha * a1);
        gus2[2 * n] = v0;
        gus2[2 * n + 1] = v1;
    }
}

// out[n][j] = sum_i m[n][i][j] * g[n][i], i ascending (transpose of matvec).
__global__ __launch_bounds__(256) void matvec_t_kernel(const float *__restrict__ m,
                                                       const float *__restrict__ gv,
                                                       float *__restrict__ out, int64_t total,
                                                       int rows, int cols)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    const int64_t n = t / cols;
    const int j = (int)(t - n * cols);
    const float *mn = m + n * rows * cols + j;
    const float *gn = gv + n * rows;
    float acc = mn[0] * gn[0];
    for (int i = 1; i < rows; ++i) acc = acc + mn[i * cols] * gn[i];
    out[t] = acc;
}

}  // namespace pde

using namespace pde;

extern "C" int pde_jacobi_f32(const float *p, const float *rhs, float *out, int64_t nx,
                              int64_t ny, int64_t nz, int32_t ghost, float h2, float inv,
                              void *stream)
{
    if (!p || !rhs || !out || out == p || out == rhs) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim == 0) return PDE_ESHAPE;
    const JacGeom g = jac_geom(G);
    const dim3 grid(clamp_grid((G.E[2] + 255) / 256, 0x7fffffff), clamp_grid(G.E[1], 65535),
                    clamp_grid(G.E[0], 65535));
    hipLaunchKernelGGL(jacobi_kernel, grid, dim3(256), 0, as_stream(stream), p, rhs, out, g, h2,
                       inv);
    return launch_status();
}

extern "C" int pde_jacobi_neumann0_f32(const float *p, const float *rhs, float *out, int64_t nx,
                                       int64_t ny, int64_t nz, int32_t ghost, float h2, float inv,
                                       void *stream)
{
    if (!p || !rhs || !out || out == p || out == rhs) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim == 0) return PDE_ESHAPE;
    for (int a = 3 - G.ndim; a < 3; ++a)
        if (G.E[a] < 2) return PDE_ESHAPE;    // a wall node's mirror must be interior
    if (G.ndim == 2)
        return jacobi_rows_launch<false>(p, out, nullptr, rhs, 1, G.P[1], G.P[2], ghost, h2, inv,
                                         as_stream(stream));
    const JacGeom g = jac_geom(G);
    const dim3 grid(clamp_grid((G.E[2] + 255) / 256, 0x7fffffff), clamp_grid(G.E[1], 65535),
                    clamp_grid(G.E[0], 65535));
    hipLaunchKernelGGL(jacobi_neumann0_kernel, grid, dim3(256), 0, as_stream(stream), p, rhs, out,
                       g, h2, inv);
    return launch_status();
}

extern "C" int pde_jacobi_adjoint_f32(const float *gout, float *gp, float *grhs, int64_t nx,
                                      int64_t ny, int64_t nz, int32_t ghost, float h2, float inv,
                                      void *stream)
{
    if (!gout || !gp || gp == gout || grhs == gout || grhs == gp) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim == 0) return PDE_ESHAPE;
    const JacGeom g = jac_geom(G);
    const dim3 grid(clamp_grid((G.P[2] + 255) / 256, 0x7fffffff), clamp_grid(G.P[1], 65535),
                    clamp_grid(G.P[0], 65535));
    hipLaunchKernelGGL(jacobi_adjoint_kernel, grid, dim3(256), 0, as_stream(stream), gout, gp, grhs,
                       g, G.P[0], G.P[1], G.P[2], h2, inv);
    return launch_status();
}

extern "C" int pde_jacobi_neumann0_adjoint_f32(const float *gout, float *gp, float *grhs,
                                               int64_t nx, int64_t ny, int64_t nz, int32_t ghost,
                                               float h2, float inv, void *stream)
{
    if (!gout || !gp || !grhs || gp == gout || grhs == gout || grhs == gp) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim == 0) return PDE_ESHAPE;
    for (int a = 3 - G.ndim; a < 3; ++a)
        if (G.E[a] < 2) return PDE_ESHAPE;
    if (G.ndim == 2)
        return jacobi_rows_launch<true>(gout, gp, grhs, nullptr, 1, G.P[1], G.P[2], ghost, h2, inv,
                                        as_stream(stream));
    const JacGeom g = jac_geom(G);
    const dim3 grid(clamp_grid((G.P[2] + 255) / 256, 0x7fffffff), clamp_grid(G.P[1], 65535),
                    clamp_grid(G.P[0], 65535));
    hipLaunchKernelGGL(jacobi_neumann0_adjoint_kernel, grid, dim3(256), 0, as_stream(stream), gout,
                       gp, grhs, g, G.P[0], G.P[1], G.P[2], h2, inv);
    return launch_status();
}

// Batches of 1D/2D instances stored one after another (stride = the padded
// node count of one instance): the batch index takes the canonical axis 0,
// which the folded geometry of such an instance leaves at size 1 and which
// no stencil term reads, so each instance is swept exactly as alone.
static int batch_geom(int64_t batch, int64_t nx, int64_t ny, int64_t nz, int32_t ghost,
                      JacGeom *out, int64_t P[3])
{
    if (batch > 65535) return PDE_ESHAPE;
    if (batch < 1) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim == 0 || G.ndim == 3 || G.P[0] != 1) return PDE_ESHAPE;
    for (int a = 3 - G.ndim; a < 3; ++a)
        if (G.E[a] < 2) return PDE_ESHAPE;
    JacGeom g = jac_geom(G);
    g.S[0] = nx * ny * nz;
    g.G[0] = 0;
    g.E[0] = batch;
    *out = g;
    P[0] = batch;
    P[1] = G.P[1];
    P[2] = G.P[2];
    return PDE_OK;
}

extern "C" int pde_jacobi_neumann0_batch_f32(const float *p, const float *rhs, float *out,
                                             int64_t batch, int64_t nx, int64_t ny, int64_t nz,
                                             int32_t ghost, float h2, float inv, void *stream)
{
    if (!p || !rhs || !out || out == p || out == rhs) return PDE_EINVAL;
    JacGeom g;
    int64_t P[3];
    int st = batch_geom(batch, nx, ny, nz, ghost, &g, P);
    if (st) return st;
    if (g.ndim == 2)
        return jacobi_rows_launch<false>(p, out, nullptr, rhs, batch, P[1], P[2], ghost, h2, inv,
                                         as_stream(stream));
    const dim3 grid(clamp_grid((g.E[2] + 255) / 256, 0x7fffffff), clamp_grid(g.E[1], 65535),
                    clamp_grid(batch, 65535));
    hipLaunchKernelGGL(jacobi_neumann0_kernel, grid, dim3(256), 0, as_stream(stream), p, rhs, out,
                       g, h2, inv);
    return launch_status();
}

extern "C" int pde_jacobi_neumann0_adjoint_batch_f32(const float *gout, float *gp, float *grhs,
                                                     int64_t batch, int64_t nx, int64_t ny,
                                                     int64_t nz, int32_t ghost, float h2,
                                                     float inv, void *stream)
{
    if (!gout || !gp || !grhs || gp == gout || grhs == gout || grhs == gp) return PDE_EINVAL;
    JacGeom g;
    int64_t P[3];
    int st = batch_geom(batch, nx, ny, nz, ghost, &g, P);
    if (st) return st;
    if (g.ndim == 2)
        return jacobi_rows_launch<true>(gout, gp, grhs, nullptr, batch, P[1], P[2], ghost, h2, inv,
                                        as_stream(stream));
    const dim3 grid(clamp_grid((P[2] + 255) / 256, 0x7fffffff), clamp_grid(P[1], 65535),
                    clamp_grid(P[0], 65535));
    hipLaunchKernelGGL(jacobi_neumann0_adjoint_kernel, grid, dim3(256), 0, as_stream(stream), gout,
                       gp, grhs, g, P[0], P[1], P[2], h2, inv);
    return launch_status();
}

extern "C" int pde_jacobi_neumann0_x2_f32(const float *p, const float *rhs, float *out,
                                          int64_t batch, int64_t nx, int64_t ny, int64_t nz,
                                          int32_t ghost, float h2, float inv, void *stream)
{
    if (!p || !rhs || !out || out == p || out == rhs) return PDE_EINVAL;
    JacGeom g;
    int64_t P[3];
    int st = batch_geom(batch, nx, ny, nz, ghost, &g, P);
    if (st) return st;
    if (g.ndim != 2) return PDE_ESHAPE;
    return jacobi_rows2_launch<false>(p, rhs, out, nullptr, batch, P[1], P[2], ghost, h2, inv,
                                      as_stream(stream));
}

extern "C" int pde_jacobi_neumann0_adjoint_x2_f32(const float *gout, float *gp, float *grhs,
                                                  int64_t batch, int64_t nx, int64_t ny,
                                                  int64_t nz, int32_t ghost, float h2, float inv,
                                                  void *stream)
{
    if (!gout || !gp || !grhs || gp == gout || grhs == gout || grhs == gp) return PDE_EINVAL;
    JacGeom g;
    int64_t P[3];
    int st = batch_geom(batch, nx, ny, nz, ghost, &g, P);
    if (st) return st;
    if (g.ndim != 2) return PDE_ESHAPE;
    return jacobi_rows2_launch<true>(gout, nullptr, gp, grhs, batch, P[1], P[2], ghost, h2, inv,
                                     as_stream(stream));
}

// Grid of the cavity kernels for `batch` instances of a 2D grid stored one
// after another (thread per column, rows on grid.y, instances on grid.z).
static int cavity_grid(int64_t batch, int64_t nx, int64_t ny, int64_t nz, int32_t ghost,
                       bool mirror, Geom *G, dim3 *grid)
{
    if (batch < 1) return PDE_EINVAL;
    if (batch > 65535) return PDE_ESHAPE;
    int st = make_geom(nx, ny, nz, ghost, G);
    if (st) return st;
    if (G->ndim != 2) return PDE_ESHAPE;
    if (mirror && (G->E[1] < 2 || G->E[2] < 2)) return PDE_ESHAPE;   // a wall node's mirror
    *grid = dim3(clamp_grid((G->P[2] + 255) / 256, 0x7fffffff), clamp_grid(G->P[1], 65535),
                 (unsigned)batch);
    return PDE_OK;
}

extern "C" int pde_cavity_predict_batch_f32(const float *u2, float *us, int64_t batch,
                                            int64_t nx, int64_t ny, int64_t nz, int32_t ghost,
                                            float lap_wm, float lap_wc, float lap_wp,
                                            float grad_wm, float grad_wc, float grad_wp, float nu,
                                            float dt, void *stream)
{
    if (!u2 || !us || u2 == us) return PDE_EINVAL;
    Geom G;
    dim3 grid;
    int st = cavity_grid(batch, nx, ny, nz, ghost, false, &G, &grid);
    if (st) return st;
    hipLaunchKernelGGL(cavity_predict_kernel, grid, dim3(256), 0, as_stream(stream), u2, us,
                       G.P[1], G.P[2], (int64_t)ghost, lap_wm, lap_wc, lap_wp, grad_wm, grad_wc,
                       grad_wp, nu, dt);
    return launch_status();
}

extern "C" int pde_cavity_predict_f32(const float *u2, float *us, int64_t nx, int64_t ny,
                                      int64_t nz, int32_t ghost, float lap_wm, float lap_wc,
                                      float lap_wp, float grad_wm, float grad_wc, float grad_wp,
                                      float nu, float dt, void *stream)
{
    return pde_cavity_predict_batch_f32(u2, us, 1, nx, ny, nz, ghost, lap_wm, lap_wc, lap_wp,
                                        grad_wm, grad_wc, grad_wp, nu, dt, stream);
}

extern "C" int pde_cavity_correct_batch_f32(const float *p, const float *us2, float *u_out,
                                            int64_t batch, int64_t nx, int64_t ny, int64_t nz,
                                            int32_t ghost, float grad_wm, float grad_wc,
                                            float grad_wp, float dt, void *stream)
{
    if (!p || !us2 || !u_out || u_out == p) return PDE_EINVAL;
    Geom G;
    dim3 grid;
    int st = cavity_grid(batch, nx, ny, nz, ghost, true, &G, &grid);
    if (st) return st;
    hipLaunchKernelGGL(cavity_correct_kernel, grid, dim3(256), 0, as_stream(stream), p, us2,
                       u_out, G.P[1], G.P[2], (int64_t)ghost, grad_wm, grad_wc, grad_wp, -dt);
    return launch_status();
}

extern "C" int pde_cavity_correct_f32(const float *p, const float *us2, float *u_out,
                                      int64_t nx, int64_t ny, int64_t nz, int32_t ghost,
                                      float grad_wm, float grad_wc, float grad_wp, float dt,
                                      void *stream)
{
    return pde_cavity_correct_batch_f32(p, us2, u_out, 1, nx, ny, nz, ghost, grad_wm, grad_wc,
                                        grad_wp, dt, stream);
}

extern "C" int pde_cavity_adjoint_predict_batch_f32(const float *gus, const float *u2,
                                                    float *gu2, int64_t batch, int64_t nx,
                                                    int64_t ny, int64_t nz, int32_t ghost,
                                                    float lap_wm, float lap_wc, float lap_wp,
                                                    float grad_wm, float grad_wc,
                                                    float grad_wp, float nu, float dt,
                                                    void *stream)
{
    if (!gus || !u2 || !gu2 || gu2 == gus || gu2 == u2) return PDE_EINVAL;
    Geom G;
    dim3 grid;
    int st = cavity_grid(batch, nx, ny, nz, ghost, false, &G, &grid);
    if (st) return st;
    hipLaunchKernelGGL(cavity_adj_predict_kernel, grid, dim3(256), 0, as_stream(stream), gus, u2,
                       gu2, G.P[1], G.P[2], (int64_t)ghost, lap_wm, lap_wc, lap_wp, grad_wm,
                       grad_wc, grad_wp, nu, dt);
    return launch_status();
}

extern "C" int pde_cavity_adjoint_predict_f32(const float *gus, const float *u2, float *gu2,
                                              int64_t nx, int64_t ny, int64_t nz, int32_t ghost,
                                              float lap_wm, float lap_wc, float lap_wp,
                                              float grad_wm, float grad_wc, float grad_wp,
                                              float nu, float dt, void *stream)
{
    return pde_cavity_adjoint_predict_batch_f32(gus, u2, gu2, 1, nx, ny, nz, ghost, lap_wm,
                                                lap_wc, lap_wp, grad_wm, grad_wc, grad_wp, nu,
                                                dt, stream);
}

extern "C" int pde_cavity_adjoint_correct_batch_f32(const float *gu_out, float *gpf,
                                                    int64_t batch, int64_t nx, int64_t ny,
                                                    int64_t nz, int32_t ghost, float grad_wm,
                                                    float grad_wc, float grad_wp, float dt,
                                                    void *stream)
{
    // gpf gathers gu_out's neighbours: it may not alias it
    if (!gu_out || !gpf || gpf == gu_out) return PDE_EINVAL;
    Geom G;
    dim3 grid;
    int st = cavity_grid(batch, nx, ny, nz, ghost, false, &G, &grid);
    if (st) return st;
    hipLaunchKernelGGL(cavity_adj_correct_kernel, grid, dim3(256), 0, as_stream(stream), gu_out,
                       gpf, G.P[1], G.P[2], (int64_t)ghost, grad_wm, grad_wc, grad_wp, -dt);
    return launch_status();
}

extern "C" int pde_cavity_adjoint_correct_f32(const float *gu_out, float *gpf, int64_t nx,
                                              int64_t ny, int64_t nz, int32_t ghost,
                                              float grad_wm, float grad_wc, float grad_wp,
                                              float dt, void *stream)
{
    return pde_cavity_adjoint_correct_batch_f32(gu_out, gpf, 1, nx, ny, nz, ghost, grad_wm,
                                                grad_wc, grad_wp, dt, stream);
}

extern "C" int pde_cavity_adjoint_div_batch_f32(const float *grhs, const float *gu_out,
                                                float *gus2, int64_t batch, int64_t nx,
                                                int64_t ny, int64_t nz, int32_t ghost,
                                                float grad_wm, float grad_wc, float grad_wp,
                                                float alpha, void *stream)
{
    if (!grhs || !gu_out || !gus2 || gus2 == grhs) return PDE_EINVAL;
    Geom G;
    dim3 grid;
    int st = cavity_grid(batch, nx, ny, nz, ghost, false, &G, &grid);
    if (st) return st;
    hipLaunchKernelGGL(cavity_adj_div_kernel, grid, dim3(256), 0, as_stream(stream), grhs, gu_out,
                       gus2, G.P[1], G.P[2], (int64_t)ghost, grad_wm, grad_wc, grad_wp, alpha);
    return launch_status();
}

extern "C" int pde_cavity_adjoint_div_f32(const float *grhs, const float *gu_out, float *gus2,
                                          int64_t nx, int64_t ny, int64_t nz, int32_t ghost,
                                          float grad_wm, float grad_wc, float grad_wp,
                                          float alpha, void *stream)
{
    return pde_cavity_adjoint_div_batch_f32(grhs, gu_out, gus2, 1, nx, ny, nz, ghost, grad_wm,
                                            grad_wc, grad_wp, alpha, stream);
}

extern "C" int pde_matvec_t_f32(const float *m, const float *g, float *out, int64_t nodes,
                                int32_t rows, int32_t cols, void *stream)
{
    if (!m || !g || !out || nodes < 0 || rows < 1 || cols < 1) return PDE_EINVAL;
    if (out == m || out == g) return PDE_EINVAL;
    const int64_t total = nodes * cols;
    if (total == 0) return PDE_OK;
    if ((total + 255) / 256 > 0x7fffffff) return PDE_ESHAPE;
    hipLaunchKernelGGL(matvec_t_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       as_stream(stream), m, g, out, total, rows, cols);
    return launch_status();
}
