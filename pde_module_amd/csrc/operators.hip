// Stencil operators: Laplacian, Grad, Divergence, the fused 3D operator pass
// and the periodic ghost fill.
//
// Every kernel keeps the reference's accumulation order (acc starts at 0, axes
// ascending, s = -1, 0, +1) and the library is compiled with
// -ffp-contract=off, so each result is the same FP32 value a serial CPU
// restatement of pm/FDM/kernel/{laplacian,grad,divergence}.py produces.
#include "pde_common.hpp"

namespace pde {

// ---------------------------------------------------------------- Laplacian

// Element geometry of one interior "row" family: a thread owns one element
// r of the flattened (fastest axis x components) interior row; rows are
// indexed by (i0, i1).  Strides are in floats.
struct RowGeom {
    int64_t base;  // float offset of the first interior element
    int64_t R;     // interior elements per row (E[2] * comps)
    int64_t E0, E1;
    int64_t S0, S1, S2;  // float strides of canonical axes 0, 1, 2
};

inline RowGeom row_geom(const Geom &G, int64_t comps)
{
    RowGeom r;
    r.S0 = G.S[0] * comps;
    r.S1 = G.S[1] * comps;
    r.S2 = comps;
    r.base = G.G[0] * r.S0 + G.G[1] * r.S1 + G.G[2] * r.S2;
    r.R = G.E[2] * comps;
    r.E0 = G.E[0];
    r.E1 = G.E[1];
    return r;
}

// Generic Laplacian: any ncomp, 0..3 eligible axes.  Canonical axis 0 is a
// grid axis only for NDIM == 3, axis 1 for NDIM >= 2, axis 2 for NDIM >= 1.
template <int NDIM>
__global__ __launch_bounds__(256) void lap_generic(const float *__restrict__ in,
                                                   float *__restrict__ out, RowGeom g,
                                                   float wm, float wc, float wp, float alpha)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= g.R) return;
    for (int64_t i0 = blockIdx.z; i0 < g.E0; i0 += gridDim.z)
        for (int64_t i1 = blockIdx.y; i1 < g.E1; i1 += gridDim.y) {
            const int64_t e = g.base + i0 * g.S0 + i1 * g.S1 + r;
            const float c = in[e];
            float acc = 0.0f;
            if (NDIM >= 3) {
                acc += wm * in[e - g.S0];
                acc += wc * c;
                acc += wp * in[e + g.S0];
            }
            if (NDIM >= 2) {
                acc += wm * in[e - g.S1];
                acc += wc * c;
                acc += wp * in[e + g.S1];
            }
            if (NDIM >= 1) {
                acc += wm * in[e - g.S2];
                acc += wc * c;
                acc += wp * in[e + g.S2];
            }
            out[e] = alpha * acc;
        }
}

// 3D scalar Laplacian: x-marching row tiles in an XCD band order.
//
// A block owns TB rows (y; LB_TB = 4 by default, 8 for Grad and Div) x 512
// points (z) of NXC consecutive planes (x; LB_NXC = 4 by default, 2 for Grad,
// 16 for Div on large grids) and walks them one plane per step.  Each lane holds 4 consecutive z
// points, so every global access is a 16-byte load or store (unaligned by the
// ghost offset, which the hardware serves).  Planes x-1, x, x+1 ride in
// registers; plane x+2, the next plane's tile-edge halo rows and its row-end
// columns are loaded before the step's barrier, so their latency overlaps the
// current plane.  y neighbours come through LDS from the rows of the block (the
// tile's top and bottom rows load their outer halo row), z neighbours by
// shuffles (a wave's end lanes load their outer column).  Outputs are written
// with non-temporal stores: nothing in the launch reads them back.
//
// Block order: hardware block b runs on XCD b % 8 (blocks are dealt
// round-robin over the 8 XCDs, each with its own L2).  Within every x-chunk
// the tiles are split into 8 contiguous bands, one per XCD, so the halo rows
// that vertically adjacent tiles share are read through one L2, while all XCDs
// sweep the same x-chunks at the same time.  Short x-chunks (4 planes) keep
// many independent blocks per CU to cover the per-plane barrier.  A padding
// slot of the last band exits at once.
//
// PER (periodic) replaces every neighbour index that leaves the interior by
// the opposite interior index instead of the ghost, which gives the values of
// a periodic ghost fill followed by the plain kernel without touching ghosts.
//
// Lanes whose 4 points run past the interior use scalar loads and stores;
// lanes and rows entirely outside it still take part in the barriers.
struct Band3 {
    int64_t SX, SY;        // plane and row pitch (floats)
    int64_t E0, E1, E2;    // interior extents
    int64_t g;
    uint32_t ty;           // tiles along y
    uint32_t T;            // tiles per x-chunk (z tiles x y tiles)
    uint32_t per;          // tiles per XCD band (ceil(T / 8))
};

constexpr int LB_TB = 4;       // rows per block
constexpr int LB_NXC = 4;      // planes per block
constexpr int LB_TZ = 512;     // z points per block (128 lanes x 4)

typedef float pde_f4 __attribute__((ext_vector_type(4)));
typedef float pde_f4u __attribute__((ext_vector_type(4), aligned(4)));

__device__ __forceinline__ pde_f4 ld4(const float *p)
{
    const pde_f4u v = *reinterpret_cast<const pde_f4u *>(p);
    return pde_f4{v.x, v.y, v.z, v.w};
}

__device__ __forceinline__ void st4(float *p, pde_f4 v)
{
    *reinterpret_cast<pde_f4u *>(p) = pde_f4u{v.x, v.y, v.z, v.w};
}

__device__ __forceinline__ void st4_nt(float *p, pde_f4 v)
{
    __builtin_nontemporal_store(pde_f4u{v.x, v.y, v.z, v.w}, reinterpret_cast<pde_f4u *>(p));
}

// Loads the n (1..3) valid points of a partial lane; the rest read as 0.
__device__ __forceinline__ pde_f4 ld4_part(const float *p, int n)
{
    pde_f4 v = {0.0f, 0.0f, 0.0f, 0.0f};
    v.x = p[0];
    if (n > 1) v.y = p[1];
    if (n > 2) v.z = p[2];
    return v;
}

// mode 2: a full lane, 1: a partial lane with np points, 0: idle.
__device__ __forceinline__ pde_f4 ld4_mode(const float *p, int mode, int np)
{
    return mode == 2 ? ld4(p) : ld4_part(p, np);
}

// The same with a streaming (non-temporal) load for the full lane: for rows
// no other block reads.
__device__ __forceinline__ pde_f4 ld4_mode_nt(const float *p, int mode, int np)
{
    if (mode != 2) return ld4_part(p, np);
    const pde_f4u v = __builtin_nontemporal_load(reinterpret_cast<const pde_f4u *>(p));
    return pde_f4{v.x, v.y, v.z, v.w};
}

__device__ __forceinline__ float lap7(float xm, float xc, float xp, float ym, float yp, float zm,
                                      float zp, float wm, float wc, float wp, float alpha)
{
    float acc = 0.0f;
    acc += wm * xm;
    acc += wc * xc;
    acc += wp * xp;
    acc += wm * ym;
    acc += wc * xc;
    acc += wp * yp;
    acc += wm * zm;
    acc += wc * xc;
    acc += wp * zp;
    return alpha * acc;
}

// Interior index of the neighbour i (-1 .. E) of an interior index along an
// axis of interior extent E: the ghost index itself, or the wrapped one.
template <bool PER>
__device__ __forceinline__ int64_t nb_index(int64_t i, int64_t E)
{
    if (!PER) return i;
    return i < 0 ? i + E : (i >= E ? i - E : i);
}

// Weights of the band kernels: Laplacian (w, alpha) and central difference
// (c, alpha of Grad and of Div).
struct BandW {
    float lm, lc, lp, la;
    float cm, cc, cp, ga, da;
};

enum { BAND_LAP = 1, BAND_GRAD = 2, BAND_DIV = 4 };

// Loads the 3 * n floats of n (1..3) vec3 nodes of a partial lane into the
// three 16-byte words of a full lane's layout; the rest read as 0.
__device__ __forceinline__ void ld12_part(const float *p, int n, pde_f4 &a, pde_f4 &b, pde_f4 &c)
{
    float t[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) t[i] = i < 3 * n ? p[i] : 0.0f;
    a = pde_f4{t[0], t[1], t[2], t[3]};
    b = pde_f4{t[4], t[5], t[6], t[7]};
    c = pde_f4{t[8], t[9], t[10], t[11]};
}

__device__ __forceinline__ void ld12_mode(const float *p, int mode, int np, pde_f4 &a, pde_f4 &b,
                                          pde_f4 &c)
{
    if (mode == 2) {
        a = ld4(p);
        b = ld4(p + 4);
        c = ld4(p + 8);
    } else {
        ld12_part(p, np, a, b, c);
    }
}

// Components of the 4 vec3 nodes held in three 16-byte words (a, b, c).
__device__ __forceinline__ pde_f4 comp_x(pde_f4 a, pde_f4 b, pde_f4 c) { return pde_f4{a.x, a.w, b.z, c.y}; }
__device__ __forceinline__ pde_f4 comp_y(pde_f4 a, pde_f4 b, pde_f4 c) { return pde_f4{a.y, b.x, b.w, c.z}; }
__device__ __forceinline__ pde_f4 comp_z(pde_f4 a, pde_f4 b, pde_f4 c) { return pde_f4{a.z, b.y, c.x, c.w}; }

// The y components of the n valid nodes of a vec3 row segment (strided loads).
__device__ __forceinline__ pde_f4 ld_comp_y(const float *p, int mode, int np)
{
    pde_f4 v = {0.0f, 0.0f, 0.0f, 0.0f};
    const int n = mode == 2 ? 4 : np;
    v.x = p[1];
    if (n > 1) v.y = p[4];
    if (n > 2) v.z = p[7];
    if (n > 3) v.w = p[10];
    return v;
}

__device__ __forceinline__ float el4(const pde_f4 &v, int i)
{
    return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}

// acc = ((0 + wm*m) + wc*c) + wp*p: one axis of a central stencil.
__device__ __forceinline__ float axis3(float m, float c, float p, float wm, float wc, float wp)
{
    float acc = 0.0f;
    acc += wm * m;
    acc += wc * c;
    acc += wp * p;
    return acc;
}

// 3D band kernel family: OPS selects Laplacian and/or Grad of a scalar s and
// Divergence of a vec3 v, all on one grid, in one pass (the fused operator
// step).  The scalar rides as in the Laplacian above: planes x-1, x, x+1 and
// the prefetched x+2 in registers, y neighbours through LDS, z neighbours by
// shuffles.  For the vector, a lane holds its 4 nodes as three 16-byte words;
// only the x components of plane x-1 are kept, the y components of the plane
// go through LDS (the tile's edge rows load the halo row's y components), the
// z components cross lanes by shuffles.  Grad writes each node's 3 slots, a
// lane's 12 floats as three 16-byte stores.
// The vector path's three planes of 12 floats make the Div and fused variants
// register-heavy (84 and 137 VGPRs: 5 and 3 waves per SIMD); capping them at
// 6 and 4 waves spills to scratch and measured slower (519 vs 454 us, 1136 vs
// 957 us at 512^3), so no occupancy is forced.
template <int OPS, bool PER, int TB = LB_TB, int NXC = LB_NXC>
__global__ __launch_bounds__(128 * TB) void ops3d_band(const float *__restrict__ s,
                                                          const float *__restrict__ v,
                                                          float *__restrict__ lap,
                                                          float *__restrict__ grad,
                                                          float *__restrict__ div, Band3 m,
                                                          BandW W)
{
    constexpr bool SC = (OPS & (BAND_LAP | BAND_GRAD)) != 0;   // scalar path
    constexpr bool VE = (OPS & BAND_DIV) != 0;                 // vector path
    __shared__ pde_f4 rows[SC ? 2 : 1][TB][128];
    __shared__ pde_f4 rowsv[VE ? 2 : 1][TB][128];
    // Grad output staging: a wave's 64 x 12 floats, so that each of its three
    // store instructions writes 1 KB contiguously
    __shared__ pde_f4 gstage[(OPS & BAND_GRAD) ? 2 * TB : 1][192];
    const uint32_t b = blockIdx.x, slot = b >> 3;
    const uint32_t xt = slot / m.per;
    const uint32_t lt = (b & 7u) * m.per + slot % m.per;
    if (lt >= m.T) return;                                // whole block: no barrier skipped
    const uint32_t yt = lt % m.ty, zt = lt / m.ty;
    const int q = threadIdx.x;          // lane within the 128-lane row group
    const int w = threadIdx.y;          // row within the tile
    const int lane = q & 63;
    const int64_t k0 = (int64_t)zt * LB_TZ + 4 * q;         // first interior z of the lane
    const int64_t nk = m.E2 - k0;                            // valid points from k0
    const int mode = nk >= 4 ? 2 : (nk > 0 ? 1 : 0);
    const int np = (int)(nk < 4 ? (nk > 0 ? nk : 0) : 4);
    // the last tile along y is shifted up to end at the last row, so every
    // tile has TB rows (E1 >= TB); rows it shares with the tile before
    // are written twice with the same values
    int64_t j0 = (int64_t)yt * TB;
    if (j0 > m.E1 - TB) j0 = m.E1 - TB;
    const int64_t j = j0 + w;
    const bool act = mode != 0;
    const bool top = w == 0, bot = w == TB - 1;
    // z neighbours from memory: lane 0 its left one, a wave's last lane or the
    // lane holding the row's last point its right one.  One of them is
    // prefetched with the planes; the right one of a lane 0 that also holds
    // the row's last point (a z tile of at most 4 points) is loaded when used.
    const bool zl_ld = act && lane == 0;
    const bool ze_ld = zl_ld || (act && (lane == 63 || nk <= 4));
    // every lane of the wave holds 4 points (uniform over the wave)
    const bool wave_full = __all(mode == 2);
    const int64_t g = m.g;
    const int64_t x0 = (int64_t)xt * NXC;
    const int nx = (int)(m.E0 - x0 < NXC ? m.E0 - x0 : NXC);
    // node offsets within a plane: the lane's centre, its tile-edge halo row,
    // its memory-read z neighbour
    const int64_t rc = (j + g) * m.SY + (k0 + g);
    const int64_t rh = (nb_index<PER>(top ? j - 1 : j + 1, m.E1) + g) * m.SY + (k0 + g);
    const int64_t rz = (j + g) * m.SY + nb_index<PER>(zl_ld ? k0 - 1 : k0 + np, m.E2) + g;
    auto plane = [&](int64_t x) -> int64_t { return (nb_index<PER>(x, m.E0) + g) * m.SX; };
    const pde_f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    // scalar state
    pde_f4 xm = zero, xc = zero, xp = zero, h = zero;
    float ze = 0.0f;
    // vector state: x components of plane x-1, planes x and x+1 as 3 words,
    // the halo row's y components, the memory-read z neighbour's z component
    pde_f4 vxm = zero, vca = zero, vcb = zero, vcc = zero, vpa = zero, vpb = zero, vpc = zero;
    pde_f4 vh = zero;
    float vze = 0.0f;
    if (act) {
        const int64_t p0 = plane(x0), pm1 = plane(x0 - 1), pp1 = plane(x0 + 1);
        if (SC) {
            xm = ld4_mode(s + pm1 + rc, mode, np);
            xc = ld4_mode(s + p0 + rc, mode, np);
            // rows 1 .. TB-2 of the chunk's inner planes are read by this block
            // alone (the edge rows are the neighbouring tiles' halos, the
            // outer planes the neighbouring chunks'): streaming loads in
            // 8-row tiles (Grad: 392 against 417 us at 512^3; 4-row tiles
            // and the vector loads of Div measured no gain or a loss)
            xp = TB >= 8 && !top && !bot && nx > 2 ? ld4_mode_nt(s + pp1 + rc, mode, np)
                                                   : ld4_mode(s + pp1 + rc, mode, np);
            if (top || bot) h = ld4_mode(s + p0 + rh, mode, np);
            if (ze_ld) ze = s[p0 + rz];
        }
        if (VE) {
            pde_f4 a, bb, c;
            ld12_mode(v + 3 * (pm1 + rc), mode, np, a, bb, c);
            vxm = comp_x(a, bb, c);
            ld12_mode(v + 3 * (p0 + rc), mode, np, vca, vcb, vcc);
            ld12_mode(v + 3 * (pp1 + rc), mode, np, vpa, vpb, vpc);
            if (top || bot) vh = ld_comp_y(v + 3 * (p0 + rh), mode, np);
            if (ze_ld) vze = v[3 * (p0 + rz) + 2];
        }
    }
    int buf = 0;
    for (int x = 0; x < nx; ++x) {
        const int64_t xa = x0 + x;                  // interior plane of this step
        pde_f4 xpp = zero, hn = zero, vqa = zero, vqb = zero, vqc = zero, vhn = zero;
        float zen = 0.0f, vzen = 0.0f;
        if (act && x + 1 < nx) {                     // x + 1 < nx is uniform over the block
            const int64_t p1 = plane(xa + 1), p2 = plane(xa + 2);
            if (SC) {
                xpp = TB >= 8 && !top && !bot && x + 3 < nx
                          ? ld4_mode_nt(s + p2 + rc, mode, np)
                          : ld4_mode(s + p2 + rc, mode, np);
                if (top || bot) hn = ld4_mode(s + p1 + rh, mode, np);
                if (ze_ld) zen = s[p1 + rz];
            }
            if (VE) {
                ld12_mode(v + 3 * (p2 + rc), mode, np, vqa, vqb, vqc);
                if (top || bot) vhn = ld_comp_y(v + 3 * (p1 + rh), mode, np);
                if (ze_ld) vzen = v[3 * (p1 + rz) + 2];
            }
        }
        const pde_f4 vy = comp_y(vca, vcb, vcc), vz = comp_z(vca, vcb, vcc);
        if (SC) rows[buf][w][q] = xc;
        if (VE) rowsv[buf][w][q] = vy;
        __syncthreads();
        float l = 0.0f, r = 0.0f, vl = 0.0f, vr = 0.0f;
        if (SC) {
            l = __shfl_up(xc.w, 1, 64);
            r = __shfl_down(xc.x, 1, 64);
        }
        if (VE) {
            vl = __shfl_up(vz.w, 1, 64);
            vr = __shfl_down(vz.x, 1, 64);
        }
        if (act) {
            const int64_t pc = (xa + g) * m.SX + rc;   // the lane's first output node
            const bool late_r = zl_ld && nk <= 4;      // lane 0 holding the row's last point
            const int64_t late = (xa + g) * m.SX + (j + g) * m.SY + nb_index<PER>(k0 + np, m.E2) + g;
            if (SC) {
                if (zl_ld) l = ze;
                else if (ze_ld) r = ze;
                if (late_r) r = s[late];
                const pde_f4 lo = top ? h : rows[buf][w > 0 ? w - 1 : 0][q];
                const pde_f4 hi = bot ? h : rows[buf][w + 1 < TB ? w + 1 : w][q];
                float o[4], gx[4], gy[4], gz[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const float c = el4(xc, t);
                    const float zm = t == 0 ? l : el4(xc, t > 0 ? t - 1 : 0);
                    const float zp = t == (mode == 2 ? 3 : np - 1) ? r : el4(xc, t < 3 ? t + 1 : 3);
                    if (OPS & BAND_LAP)
                        o[t] = lap7(el4(xm, t), c, el4(xp, t), el4(lo, t), el4(hi, t), zm, zp,
                                    W.lm, W.lc, W.lp, W.la);
                    if (OPS & BAND_GRAD) {
                        gx[t] = W.ga * axis3(el4(xm, t), c, el4(xp, t), W.cm, W.cc, W.cp);
                        gy[t] = W.ga * axis3(el4(lo, t), c, el4(hi, t), W.cm, W.cc, W.cp);
                        gz[t] = W.ga * axis3(zm, c, zp, W.cm, W.cc, W.cp);
                    }
                }
                if (mode == 2) {
                    if (OPS & BAND_LAP) st4_nt(lap + pc, pde_f4{o[0], o[1], o[2], o[3]});
                    if (OPS & BAND_GRAD) {
                        const pde_f4 g0 = {gx[0], gy[0], gz[0], gx[1]};
                        const pde_f4 g1 = {gy[1], gz[1], gx[2], gy[2]};
                        const pde_f4 g2 = {gz[2], gx[3], gy[3], gz[3]};
                        if (wave_full) {
                            // through LDS: lane i then stores floats 4i.. of
                            // each KB of the wave's 3 KB output
                            pde_f4 *st = gstage[2 * w + (q >> 6)];
                            st[3 * lane] = g0;
                            st[3 * lane + 1] = g1;
                            st[3 * lane + 2] = g2;
                            __builtin_amdgcn_wave_barrier();
                            const pde_f4 s0 = st[lane], s1 = st[64 + lane], s2 = st[128 + lane];
                            float *gp = grad + 3 * (pc - 4 * lane) + 4 * lane;
                            st4_nt(gp, s0);
                            st4_nt(gp + 256, s1);
                            st4_nt(gp + 512, s2);
                        } else {
                            // a wave with partial lanes (the row's last z tile)
                            float *gp = grad + 3 * pc;
                            st4_nt(gp, g0);
                            st4_nt(gp + 4, g1);
                            st4_nt(gp + 8, g2);
                        }
                    }
                } else {
                    for (int t = 0; t < np; ++t) {
                        if (OPS & BAND_LAP) lap[pc + t] = o[t];
                        if (OPS & BAND_GRAD) {
                            grad[3 * (pc + t)] = gx[t];
                            grad[3 * (pc + t) + 1] = gy[t];
                            grad[3 * (pc + t) + 2] = gz[t];
                        }
                    }
                }
            }
            if (VE) {
                if (zl_ld) vl = vze;
                else if (ze_ld) vr = vze;
                if (late_r) vr = v[3 * late + 2];
                const pde_f4 lo = top ? vh : rowsv[buf][w > 0 ? w - 1 : 0][q];
                const pde_f4 hi = bot ? vh : rowsv[buf][w + 1 < TB ? w + 1 : w][q];
                const pde_f4 vxc = comp_x(vca, vcb, vcc), vxp = comp_x(vpa, vpb, vpc);
                float d[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const float zm = t == 0 ? vl : el4(vz, t > 0 ? t - 1 : 0);
                    const float zp = t == (mode == 2 ? 3 : np - 1) ? vr : el4(vz, t < 3 ? t + 1 : 3);
                    float acc = 0.0f;
                    acc += W.cm * el4(vxm, t);
                    acc += W.cc * el4(vxc, t);
                    acc += W.cp * el4(vxp, t);
                    acc += W.cm * el4(lo, t);
                    acc += W.cc * el4(vy, t);
                    acc += W.cp * el4(hi, t);
                    acc += W.cm * zm;
                    acc += W.cc * el4(vz, t);
                    acc += W.cp * zp;
                    d[t] = W.da * acc;
                }
                if (mode == 2)
                    st4_nt(div + pc, pde_f4{d[0], d[1], d[2], d[3]});
                else
                    for (int t = 0; t < np; ++t) div[pc + t] = d[t];
            }
        }
        if (SC) {
            xm = xc;
            xc = xp;
            xp = xpp;
            h = hn;
            ze = zen;
        }
        if (VE) {
            vxm = comp_x(vca, vcb, vcc);
            vca = vpa;
            vcb = vpb;
            vcc = vpc;
            vpa = vqa;
            vpb = vqb;
            vpc = vqc;
            vh = vhn;
            vze = vzen;
        }
        buf ^= 1;
    }
}

// 3D scalar Laplacian through an LDS plane ring (rows of 256 * NW points).
//
// The same tiles, chunks and XCD band order as the band kernels (RT_TB = 4
// rows x 256 * NW z points per block, chunks of RT_NXC_LONG = 32 planes where
// that still gives 2048 blocks, else RT_NXC = 12), but the
// planes never pass through registers on their way in: every wave streams its
// row segment of each plane (16 bytes per lane) straight into LDS with
// global_load_lds_dwordx4, RT_PF planes ahead of the plane being computed.
// A ring slot holds one plane of the tile: its TB rows, the two tile-edge halo
// rows, and in each row the two z neighbours of the segment's ends (position 3
// and 4 + 256 * NW; the points start at position 4, so every lane's 16-byte
// LDS read is aligned).  The plane before the current one (in march order)
// stays in registers from the step before, so a slot is free once its plane
// has been the current one: the ring holds planes x and x+1 in use plus RT_PF
// planes in flight (RT_PF + 2 slots, 62 KB for NW = 2: two blocks per CU).
//
// Each wave counts its own outstanding vector-memory operations: per plane it
// issues one row load, one z-end load on the first and last wave of a row, one
// halo-row load on the tile's first and last rows, and per computed plane one
// store (stores count in vmcnt on this architecture).  Before computing plane
// x it waits until everything up to plane x + 1's loads has landed, then a
// barrier publishes the other waves' loads; the slot refilled after that
// barrier (plane x + 1 + RT_PF) held plane x - 1, which every wave finished
// reading before it.  The compiler is kept from moving LDS reads across the
// barrier by memory clobbers and emits no wait of its own for the ring.
//
// Odd chunks march down (planes in descending order, the x-1 and x+1 operands
// swapped back so the arithmetic is unchanged): the chunks on either side of
// a chunk boundary then load its two planes at about the same time, so the
// second load finds them in L2 instead of reading them again from memory.
//
// Requires E2 a multiple of 256 * NW (every lane holds 4 points) and
// E1 >= RT_TB; the arithmetic and its order are the band kernel's.
//
// Prefetch depth and rows per block are template parameters with these
// defaults.  At 512^3 periodic (NW = 2, 32-plane chunks): PF 3 / TB 4
// 180-181 us; PF 4 (75 KB, still two blocks per CU) 181; PF 2 184; TB 2
// (PF 3..6) 194-201, its extra halo rows costing more than the third block
// per CU gains; TB 8 / PF 2 (one block per CU) 185; PF 4 with 64-plane
// chunks 187.
constexpr int RT_TB = 4;
constexpr int RT_NXC = 12;
constexpr int RT_NXC_LONG = 32;
constexpr int RT_PF = 3;

struct Ring3 {
    int64_t SX, SY;
    int32_t E0, E1, E2, g;
    uint32_t tz, T, per;   // z tiles, tiles per chunk, tiles per XCD band
    int32_t nxc;           // planes per chunk
};

template <bool PER>
__device__ __forceinline__ int wrap_index(int i, int E)
{
    if (!PER) return i;
    return i < 0 ? i + E : (i >= E ? i - E : i);
}

// s_waitcnt vmcnt(n) for a wave-uniform n <= 31 (the count is an immediate:
// vmcnt bits 3:0 and 5:4 sit at 3:0 and 15:14, expcnt and lgkmcnt stay at
// their maxima).  Larger counts wait for everything.
#define PDE_VMCNT(n) (0x0f70 | ((n) & 15) | (((n) >> 4) << 14))
#define PDE_WAIT_CASE(n) \
    case n: __builtin_amdgcn_s_waitcnt(PDE_VMCNT(n)); break;
__device__ __forceinline__ void wait_vm(int n)
{
    switch (n) {
    PDE_WAIT_CASE(0) PDE_WAIT_CASE(1) PDE_WAIT_CASE(2) PDE_WAIT_CASE(3)
    PDE_WAIT_CASE(4) PDE_WAIT_CASE(5) PDE_WAIT_CASE(6) PDE_WAIT_CASE(7)
    PDE_WAIT_CASE(8) PDE_WAIT_CASE(9) PDE_WAIT_CASE(10) PDE_WAIT_CASE(11)
    PDE_WAIT_CASE(12) PDE_WAIT_CASE(13) PDE_WAIT_CASE(14) PDE_WAIT_CASE(15)
    PDE_WAIT_CASE(16) PDE_WAIT_CASE(17) PDE_WAIT_CASE(18) PDE_WAIT_CASE(19)
    PDE_WAIT_CASE(20) PDE_WAIT_CASE(21) PDE_WAIT_CASE(22) PDE_WAIT_CASE(23)
    PDE_WAIT_CASE(24) PDE_WAIT_CASE(25) PDE_WAIT_CASE(26) PDE_WAIT_CASE(27)
    PDE_WAIT_CASE(28) PDE_WAIT_CASE(29) PDE_WAIT_CASE(30) PDE_WAIT_CASE(31)
    default: __builtin_amdgcn_s_waitcnt(PDE_VMCNT(0)); break;
    }
}
#undef PDE_WAIT_CASE

typedef __attribute__((address_space(3))) void *lds_ptr;

// Cache policy operand of the direct-to-LDS loads: non-temporal (streaming).
constexpr int CPOL_NT = 2;

template <bool PER, int NW, int PF = RT_PF, int TB = RT_TB>
__global__ __launch_bounds__(64 * NW * TB) void lap3d_ring(const float *__restrict__ s,
                                                           float *__restrict__ out, Ring3 m,
                                                           float wm, float wc, float wp,
                                                           float alpha)
{
    const int NXC = m.nxc;
    constexpr int R = PF + 2;                // ring slots; march plane p in slot p % R
    static_assert(PF >= 1 && PF + (PF - 1) * 4 <= 31, "wait counts must fit wait_vm");
    constexpr int ZT = 256 * NW;             // z points per tile
    constexpr int RW = ZT + 8;               // floats per LDS row
    __shared__ __attribute__((aligned(16))) float ring[R][TB + 2][RW];
    const uint32_t b = blockIdx.x, slot = b >> 3;
    const uint32_t xt = slot / m.per;
    const uint32_t lt = (b & 7u) * m.per + slot % m.per;
    if (lt >= m.T) return;                   // whole block: no barrier skipped
    // z tiles of one row band are neighbours in the XCD band (the segments
    // read each other's end points through one L2)
    const uint32_t zt = lt % m.tz, yt = lt / m.tz;
    // half and w are uniform over a wave (64 * NW threads along x): scalars
    const int lane = threadIdx.x & 63;
    const int half = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    int j0 = (int)yt * TB;                // the last tile ends at the last row
    if (j0 > m.E1 - TB) j0 = m.E1 - TB;
    const int j = j0 + w;
    const int k0 = (int)zt * ZT;
    const int x0 = (int)xt * NXC;
    const int nx = m.E0 - x0 < NXC ? m.E0 - x0 : NXC;
    const bool top = w == 0, bot = w == TB - 1;
    const bool zlo = half == 0, zhi = half == NW - 1;
    const int64_t g = m.g;
    const int c0 = 4 + 256 * half + 4 * lane;             // the lane's LDS column
    // offsets within a plane: the lane's 4 points, its halo row (tile edges),
    // the row segment's z neighbours
    const int64_t rc = (j + g) * m.SY + g + (c0 - 4 + k0);
    const int64_t rh = (wrap_index<PER>(top ? j - 1 : j + 1, m.E1) + g) * m.SY + g + (c0 - 4 + k0);
    const int64_t rzl = (j + g) * m.SY + g + wrap_index<PER>(k0 - 1, m.E2);
    const int64_t rzh = (j + g) * m.SY + g + wrap_index<PER>(k0 + ZT, m.E2);
    const bool down = (xt & 1u) != 0;        // this chunk marches down
    // base of the pl-th plane of the march (-1 .. nx)
    auto plane_of = [&](int pl) -> const float * {
        const int ap = down ? nx - 1 - pl : pl;     // chunk-relative plane
        return s + (wrap_index<PER>(x0 + ap, m.E0) + g) * m.SX;
    };
    auto issue = [&](int pl) {
        const int sl = pl % R;
        const float *pb = plane_of(pl);
        // rows no other block reads (the tile's inner rows of the chunk's
        // inner planes) stream past L2 (non-temporal policy): 184-186
        // against 188 us at 512^3; on every row load 190 us
        if (!top && !bot && pl > 0 && pl < nx - 1)
            __builtin_amdgcn_global_load_lds((const void *)(pb + rc),
                                             (lds_ptr)&ring[sl][w + 1][4 + 256 * half], 16, 0,
                                             CPOL_NT);
        else
            __builtin_amdgcn_global_load_lds((const void *)(pb + rc),
                                             (lds_ptr)&ring[sl][w + 1][4 + 256 * half], 16, 0, 0);
        if (zlo && lane == 0)
            __builtin_amdgcn_global_load_lds((const void *)(pb + rzl),
                                             (lds_ptr)&ring[sl][w + 1][3], 4, 0, 0);
        if (zhi && lane == 0)
            __builtin_amdgcn_global_load_lds((const void *)(pb + rzh),
                                             (lds_ptr)&ring[sl][w + 1][4 + ZT], 4, 0, 0);
        if (top || bot)
            __builtin_amdgcn_global_load_lds((const void *)(pb + rh),
                                             (lds_ptr)&ring[sl][top ? 0 : TB + 1][4 + 256 * half],
                                             16, 0, 0);
    };
    const int per = 1 + (zlo ? 1 : 0) + (zhi ? 1 : 0) + ((top || bot) ? 1 : 0);
    const int npro = PF < nx ? PF : nx;
    // march plane -1: a register load issued before every ring load, so the
    // wait for it waits for nothing else
    pde_f4 prev = ld4(plane_of(-1) + rc);
    for (int p = 0; p <= npro; ++p) issue(p);
    for (int x = 0; x < nx; ++x) {
        // issued after plane x + 1's ring loads: the stores of planes
        // max(x - PF, 0) .. x - 1 and the ring loads of planes x + 2 .. last
        const int last = x + PF < nx ? x + PF : nx;
        const int ahead = last - (x + 1) > 0 ? last - (x + 1) : 0;
        wait_vm(__builtin_amdgcn_readfirstlane((x < PF ? x : PF) + ahead * per));
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        // the slot refilled held plane x - 1, which every wave finished
        // reading before the barrier
        if (x + 1 + PF <= nx) issue(x + 1 + PF);
        const int sc = x % R, sn = (x + 1) % R;
        const pde_f4 xc = *reinterpret_cast<const pde_f4 *>(&ring[sc][w + 1][c0]);
        const pde_f4 lo = *reinterpret_cast<const pde_f4 *>(&ring[sc][w][c0]);
        const pde_f4 hi = *reinterpret_cast<const pde_f4 *>(&ring[sc][w + 2][c0]);
        const pde_f4 xn = *reinterpret_cast<const pde_f4 *>(&ring[sn][w + 1][c0]);
        const float l = ring[sc][w + 1][c0 - 1];
        const float r = ring[sc][w + 1][c0 + 4];
        const pde_f4 xm = down ? xn : prev, xp = down ? prev : xn;
        float o[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const float zm = t == 0 ? l : el4(xc, t > 0 ? t - 1 : 0);
            const float zp = t == 3 ? r : el4(xc, t < 3 ? t + 1 : 3);
            o[t] = lap7(el4(xm, t), el4(xc, t), el4(xp, t), el4(lo, t), el4(hi, t), zm, zp, wm, wc,
                        wp, alpha);
        }
        st4_nt(out + (x0 + (down ? nx - 1 - x : x) + g) * m.SX + rc,
               pde_f4{o[0], o[1], o[2], o[3]});
        prev = xc;
    }
}

template <bool PER, int NW, int PF = RT_PF, int TB = RT_TB>
static int lap3d_ring_launch(const float *in, float *out, const Geom &G, int ghost, float wm,
                             float wc, float wp, float alpha, hipStream_t s,
                             int nxc_long = RT_NXC_LONG)
{
    Ring3 m;
    m.SX = G.S[0];
    m.SY = G.S[1];
    m.E0 = (int32_t)G.E[0];
    m.E1 = (int32_t)G.E[1];
    m.E2 = (int32_t)G.E[2];
    m.g = ghost;
    const int64_t ty = (G.E[1] + TB - 1) / TB, tz = G.E[2] / (256 * NW);
    const int64_t T = tz * ty;
    const int64_t per = (T + 7) / 8;
    // long chunks re-read fewer chunk-boundary planes (512^3: 180-182 us
    // with 32 planes, 178-181 with 64, 184-187 with 12, 185 with 48, which
    // leaves a short last chunk); short ones keep small grids' blocks many
    const int64_t nxc = 8 * per * ((G.E[0] + nxc_long - 1) / nxc_long) >= 2048
                            ? nxc_long
                            : RT_NXC;
    const int64_t nb = 8 * per * ((G.E[0] + nxc - 1) / nxc);
    if (nb > 0x7fffffff) return PDE_ESHAPE;
    m.nxc = (int32_t)nxc;
    m.tz = (uint32_t)tz;
    m.T = (uint32_t)T;
    m.per = (uint32_t)per;
    hipLaunchKernelGGL((lap3d_ring<PER, NW, PF, TB>), dim3((unsigned)nb), dim3(64 * NW, TB), 0,
                       s, in, out, m, wm, wc, wp, alpha);
    return launch_status();
}

// 3D vec3 Divergence through an LDS plane ring (the Laplacian ring's scheme
// for a vector field).  A block owns TB rows x 256 points (one wave per row,
// 4 points per lane) of a chunk of planes (DR_TB = 4 rows of DR_NXC = 16
// planes by default; div3d_ring_run picks the shape); a ring slot holds the tile's
// rows of one plane as the field stores them (3 floats per point, 3 KB per
// row, three 16-byte direct loads per lane), its two halo rows, and the z
// components of each row segment's two outer neighbours, placed where the
// lane arithmetic looks for them (position 3 and 4 + 3 * 256 + 2; the points
// start at position 4).  The x components of the plane before the current one
// stay in registers; the plane after gives its x components, the halo rows
// their y components, and each lane's outer z neighbours are single floats
// beside its 12.  DR_PF = 2 planes in flight, DR_PF + 2 slots (74 KB at 4
// rows: two blocks per CU).  Per plane a wave issues three row loads and two z-end
// loads, the edge rows' waves three more for their halo row; one store per
// computed plane.  Odd chunks march down as in the Laplacian ring.
// Requires E2 % 256 == 0 and E1 >= TB; operation order of the band kernel.
constexpr int DR_TB = 4;
constexpr int DR_NXC = 16;
constexpr int DR_PF = 2;

template <int PF = DR_PF, int TB = DR_TB>
__global__ __launch_bounds__(64 * TB) void div3d_ring(const float *__restrict__ v,
                                                      float *__restrict__ out, Ring3 m,
                                                      float cm, float cc, float cp, float alpha)
{
    const int NXC = m.nxc;
    constexpr int R = PF + 2;                // march plane p in slot p % R
    static_assert(PF >= 1 && PF + (PF - 1) * 8 <= 31, "wait counts must fit wait_vm");
    constexpr int ZT = 256;
    constexpr int RW = 3 * ZT + 8;
    __shared__ __attribute__((aligned(16))) float ring[R][TB + 2][RW];
    const uint32_t b = blockIdx.x, slot = b >> 3;
    const uint32_t xt = slot / m.per;
    const uint32_t lt = (b & 7u) * m.per + slot % m.per;
    if (lt >= m.T) return;                   // whole block: no barrier skipped
    // z tiles of one row band are neighbours in the XCD band (the segments
    // read each other's end points through one L2)
    const uint32_t zt = lt % m.tz, yt = lt / m.tz;
    const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    int j0 = (int)yt * TB;
    if (j0 > m.E1 - TB) j0 = m.E1 - TB;
    const int j = j0 + w;
    const int k0 = (int)zt * ZT;
    const int x0 = (int)xt * NXC;
    const int nx = m.E0 - x0 < NXC ? m.E0 - x0 : NXC;
    const bool top = w == 0, bot = w == TB - 1;
    const bool down = (xt & 1u) != 0;
    const int64_t g = m.g;
    const int64_t n0 = (j + g) * m.SY + g + k0;                  // first node of the segment
    const int64_t nh = ((top ? j - 1 : j + 1) + g) * m.SY + g + k0;
    auto plane_of = [&](int pl) -> const float * {
        const int ap = down ? nx - 1 - pl : pl;
        return v + 3 * ((x0 + ap + g) * m.SX);
    };
    auto issue = [&](int pl) {
        const int sl = pl % R;
        const float *pb = plane_of(pl);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            // rows no other block reads stream past L2 (374 against 377 us)
            if (!top && !bot && pl > 0 && pl < nx - 1)
                __builtin_amdgcn_global_load_lds((const void *)(pb + 3 * n0 + 256 * c + 4 * lane),
                                                 (lds_ptr)&ring[sl][w + 1][4 + 256 * c], 16, 0,
                                                 CPOL_NT);
            else
                __builtin_amdgcn_global_load_lds((const void *)(pb + 3 * n0 + 256 * c + 4 * lane),
                                                 (lds_ptr)&ring[sl][w + 1][4 + 256 * c], 16, 0, 0);
        }
        if (lane == 0) {
            __builtin_amdgcn_global_load_lds((const void *)(pb + 3 * (n0 - 1) + 2),
                                             (lds_ptr)&ring[sl][w + 1][3], 4, 0, 0);
            __builtin_amdgcn_global_load_lds((const void *)(pb + 3 * (n0 + ZT) + 2),
                                             (lds_ptr)&ring[sl][w + 1][4 + 3 * ZT + 2], 4, 0, 0);
        }
        if (top || bot) {
            const int hr = top ? 0 : TB + 1;
#pragma unroll
            for (int c = 0; c < 3; ++c)
                __builtin_amdgcn_global_load_lds((const void *)(pb + 3 * nh + 256 * c + 4 * lane),
                                                 (lds_ptr)&ring[sl][hr][4 + 256 * c], 16, 0, 0);
        }
    };
    const int per = 5 + ((top || bot) ? 3 : 0);
    const int npro = PF < nx ? PF : nx;
    // x components of march plane -1: register loads issued before every ring
    // load, so the wait for them waits for nothing else
    const float *p1 = plane_of(-1) + 3 * n0 + 12 * lane;
    pde_f4 prev = comp_x(ld4(p1), ld4(p1 + 4), ld4(p1 + 8));
    for (int p = 0; p <= npro; ++p) issue(p);
    const int c0 = 4 + 12 * lane;
    for (int x = 0; x < nx; ++x) {
        // issued after plane x + 1's ring loads: the stores of planes
        // max(x - PF, 0) .. x - 1 and the ring loads of planes x + 2 .. last
        const int last = x + PF < nx ? x + PF : nx;
        const int ahead = last - (x + 1) > 0 ? last - (x + 1) : 0;
        wait_vm(__builtin_amdgcn_readfirstlane((x < PF ? x : PF) + ahead * per));
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        if (x + 1 + PF <= nx) issue(x + 1 + PF);
        const int sc = x % R, sn = (x + 1) % R;
        auto ld3 = [&](const float *p, pde_f4 &a3, pde_f4 &b3, pde_f4 &c3) {
            a3 = *reinterpret_cast<const pde_f4 *>(p);
            b3 = *reinterpret_cast<const pde_f4 *>(p + 4);
            c3 = *reinterpret_cast<const pde_f4 *>(p + 8);
        };
        pde_f4 va, vb, vc;
        ld3(&ring[sc][w + 1][c0], va, vb, vc);
        const pde_f4 vxc = comp_x(va, vb, vc), vy = comp_y(va, vb, vc), vz = comp_z(va, vb, vc);
        ld3(&ring[sn][w + 1][c0], va, vb, vc);
        const pde_f4 vxn = comp_x(va, vb, vc);
        ld3(&ring[sc][w][c0], va, vb, vc);
        const pde_f4 lo = comp_y(va, vb, vc);
        ld3(&ring[sc][w + 2][c0], va, vb, vc);
        const pde_f4 hi = comp_y(va, vb, vc);
        const float l = ring[sc][w + 1][c0 - 1];
        const float r = ring[sc][w + 1][c0 + 14];
        const pde_f4 vxl = down ? vxn : prev, vxr = down ? prev : vxn;
        float d[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const float zm = t == 0 ? l : el4(vz, t > 0 ? t - 1 : 0);
            const float zp = t == 3 ? r : el4(vz, t < 3 ? t + 1 : 3);
            float acc = 0.0f;
            acc += cm * el4(vxl, t);
            acc += cc * el4(vxc, t);
            acc += cp * el4(vxr, t);
            acc += cm * el4(lo, t);
            acc += cc * el4(vy, t);
            acc += cp * el4(hi, t);
            acc += cm * zm;
            acc += cc * el4(vz, t);
            acc += cp * zp;
            d[t] = alpha * acc;
        }
        const int ap = down ? nx - 1 - x : x;
        st4_nt(out + (x0 + ap + g) * m.SX + n0 + 4 * lane, pde_f4{d[0], d[1], d[2], d[3]});
        prev = vxc;
    }
}

template <int PF = DR_PF, int TB = DR_TB>
static int div3d_ring_launch(const float *in, float *out, const Geom &G, int ghost, float wm,
                             float wc, float wp, float alpha, hipStream_t s, int nxc = DR_NXC)
{
    Ring3 m;
    m.SX = G.S[0];
    m.SY = G.S[1];
    m.E0 = (int32_t)G.E[0];
    m.E1 = (int32_t)G.E[1];
    m.E2 = (int32_t)G.E[2];
    m.g = ghost;
    const int64_t ty = (G.E[1] + TB - 1) / TB, tz = G.E[2] / 256;
    const int64_t T = tz * ty;
    const int64_t per = (T + 7) / 8;
    const int64_t nb = 8 * per * ((G.E[0] + nxc - 1) / nxc);
    if (nb > 0x7fffffff) return PDE_ESHAPE;
    m.nxc = nxc;
    m.tz = (uint32_t)tz;
    m.T = (uint32_t)T;
    m.per = (uint32_t)per;
    hipLaunchKernelGGL((div3d_ring<PF, TB>), dim3((unsigned)nb), dim3(64, TB), 0, s, in, out, m,
                       wm, wc, wp, alpha);
    return launch_status();
}

// The Divergence ring's tile shape: 8-row tiles of 32-plane chunks (one
// block per CU) where those still give 2048 blocks, else 4-row tiles of 16
// planes.  At 512^3: 366 us against 371-372 for 4 rows with 16, 24 or 32
// planes and 371 for 8 rows with 16; a third plane in flight (one block per
// CU at 4 rows) 458-480.
static int div3d_ring_run(const float *in, float *out, const Geom &G, int ghost, float wm,
                          float wc, float wp, float alpha, hipStream_t s)
{
    const int64_t per8 = ((G.E[2] / 256) * ((G.E[1] + 7) / 8) + 7) / 8;
    if (G.E[1] >= 8 && 8 * per8 * ((G.E[0] + 31) / 32) >= 2048)
        return div3d_ring_launch<DR_PF, 8>(in, out, G, ghost, wm, wc, wp, alpha, s, 32);
    return div3d_ring_launch(in, out, G, ghost, wm, wc, wp, alpha, s);
}

// Generic periodic Laplacian (any ncomp, 1..3 eligible axes): thread per
// interior element; a neighbour index that leaves the interior wraps to the
// opposite interior index.
template <int NDIM>
__global__ __launch_bounds__(256) void lap_generic_periodic(const float *__restrict__ in,
                                                            float *__restrict__ out, RowGeom g,
                                                            int64_t comps, float wm, float wc,
                                                            float wp, float alpha)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= g.R) return;
    const int64_t k = r / comps, E2 = g.R / comps;
    const int64_t rm = (k == 0 ? r + g.R : r) - comps;             // wrapped z neighbours
    const int64_t rp = (k == E2 - 1 ? r - g.R : r) + comps;
    for (int64_t i0 = blockIdx.z; i0 < g.E0; i0 += gridDim.z)
        for (int64_t i1 = blockIdx.y; i1 < g.E1; i1 += gridDim.y) {
            const int64_t row = g.base + i0 * g.S0 + i1 * g.S1;
            const int64_t e = row + r;
            const float c = in[e];
            float acc = 0.0f;
            if (NDIM >= 3) {
                const int64_t a = (i0 == 0 ? g.E0 - 1 : i0 - 1) - i0;
                const int64_t b = (i0 == g.E0 - 1 ? 0 : i0 + 1) - i0;
                acc += wm * in[e + a * g.S0];
                acc += wc * c;
                acc += wp * in[e + b * g.S0];
            }
            if (NDIM >= 2) {
                const int64_t a = (i1 == 0 ? g.E1 - 1 : i1 - 1) - i1;
                const int64_t b = (i1 == g.E1 - 1 ? 0 : i1 + 1) - i1;
                acc += wm * in[e + a * g.S1];
                acc += wc * c;
                acc += wp * in[e + b * g.S1];
            }
            if (NDIM >= 1) {
                acc += wm * in[row + rm];
                acc += wc * c;
                acc += wp * in[row + rp];
            }
            out[e] = alpha * acc;
        }
}

// ------------------------------------------------------------------ Grad

// Thread per (node, component r) of an interior row; writes the D slots of
// row r of the node's N x D output matrix.
template <int NDIM>
__global__ __launch_bounds__(256) void grad_generic(const float *__restrict__ in,
                                                    float *__restrict__ out, RowGeom g,
                                                    float wm, float wc, float wp, float alpha)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= g.R) return;
    for (int64_t i0 = blockIdx.z; i0 < g.E0; i0 += gridDim.z)
        for (int64_t i1 = blockIdx.y; i1 < g.E1; i1 += gridDim.y) {
            const int64_t e = g.base + i0 * g.S0 + i1 * g.S1 + r;
            const float c = in[e];
            float *o = out + e * NDIM;
            int t = 0;
            if (NDIM >= 3) {
                float acc = 0.0f;
                acc += wm * in[e - g.S0];
                acc += wc * c;
                acc += wp * in[e + g.S0];
                o[t++] = alpha * acc;
            }
            if (NDIM >= 2) {
                float acc = 0.0f;
                acc += wm * in[e - g.S1];
                acc += wc * c;
                acc += wp * in[e + g.S1];
                o[t++] = alpha * acc;
            }
            if (NDIM >= 1) {
                float acc = 0.0f;
                acc += wm * in[e - g.S2];
                acc += wc * c;
                acc += wp * in[e + g.S2];
                o[t] = alpha * acc;
            }
        }
}

// ------------------------------------------------------------- Divergence

// Thread per (node, row r).  Input element (node, r, t) sits at float
// (node*rows + r)*D + t; `g` is the row geometry of the *output* (comps =
// rows), so the input offset of output element e is e*D (+ t).
template <int NDIM>
__global__ __launch_bounds__(256) void div_generic(const float *__restrict__ in,
                                                   float *__restrict__ out, RowGeom g,
                                                   float wm, float wc, float wp, float alpha)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= g.R) return;
    const int64_t D = NDIM;
    for (int64_t i0 = blockIdx.z; i0 < g.E0; i0 += gridDim.z)
        for (int64_t i1 = blockIdx.y; i1 < g.E1; i1 += gridDim.y) {
            const int64_t e = g.base + i0 * g.S0 + i1 * g.S1 + r;
            const float *p = in + e * D;
            float acc = 0.0f;
            int t = 0;
            if (NDIM >= 3) {
                acc += wm * p[t - g.S0 * D];
                acc += wc * p[t];
                acc += wp * p[t + g.S0 * D];
                ++t;
            }
            if (NDIM >= 2) {
                acc += wm * p[t - g.S1 * D];
                acc += wc * p[t];
                acc += wp * p[t + g.S1 * D];
                ++t;
            }
            if (NDIM >= 1) {
                acc += wm * p[t - g.S2 * D];
                acc += wc * p[t];
                acc += wp * p[t + g.S2 * D];
            }
            out[e] = alpha * acc;
        }
}

// ------------------------------------------------------- fused lap/grad/div

struct March3 {
    int64_t SX, SY;        // plane and row pitch (floats)
    int64_t E0, E1, E2;    // interior extents
    int64_t g;
    uint32_t tz, ty;       // tile counts along z and y
};

// One x-marching pass over a scalar s and a vec3 v: lap(s), grad(s), div(v).
// s's x-1 / x / x+1 and v.x's x-1 / x+1 ride in registers.
template <int BY, int NXC>
__global__ __launch_bounds__(64 * BY) void fused3d_march(
    const float *__restrict__ s, const float *__restrict__ v, float *__restrict__ lap,
    float *__restrict__ grad, float *__restrict__ div, March3 m, float lm, float lc, float lp,
    float cm, float cc, float cp, float al, float ag, float ad)
{
    uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t zt = b % m.tz;
    b /= m.tz;
    const uint32_t yt = b % m.ty;
    const uint32_t xt = b / m.ty;
    const int64_t k = (int64_t)zt * 64 + threadIdx.x;
    const int64_t j = (int64_t)yt * BY + threadIdx.y;
    if (k >= m.E2 || j >= m.E1) return;
    const int64_t x0 = (int64_t)xt * NXC;
    int64_t nx = m.E0 - x0;
    if (nx > NXC) nx = NXC;
    int64_t e = (x0 + m.g) * m.SX + (j + m.g) * m.SY + (k + m.g);
    float sm = s[e - m.SX], sc = s[e];
    float vxm = v[3 * (e - m.SX)], vxc = v[3 * e];
#pragma unroll 2
    for (int64_t x = 0; x < nx; ++x, e += m.SX) {
        const float sp = s[e + m.SX];
        const float sym = s[e - m.SY], syp = s[e + m.SY];
        const float szm = s[e - 1], szp = s[e + 1];
        const float vxp = v[3 * (e + m.SX)];
        const float vym = v[3 * (e - m.SY) + 1], vyc = v[3 * e + 1], vyp = v[3 * (e + m.SY) + 1];
        const float vzm = v[3 * (e - 1) + 2], vzc = v[3 * e + 2], vzp = v[3 * (e + 1) + 2];
        float acc = 0.0f;
        acc += lm * sm;
        acc += lc * sc;
        acc += lp * sp;
        acc += lm * sym;
        acc += lc * sc;
        acc += lp * syp;
        acc += lm * szm;
        acc += lc * sc;
        acc += lp * szp;
        lap[e] = al * acc;
        float gx = 0.0f, gy = 0.0f, gz = 0.0f;
        gx += cm * sm;
        gx += cc * sc;
        gx += cp * sp;
        gy += cm * sym;
        gy += cc * sc;
        gy += cp * syp;
        gz += cm * szm;
        gz += cc * sc;
        gz += cp * szp;
        grad[3 * e] = ag * gx;
        grad[3 * e + 1] = ag * gy;
        grad[3 * e + 2] = ag * gz;
        float d = 0.0f;
        d += cm * vxm;
        d += cc * vxc;
        d += cp * vxp;
        d += cm * vym;
        d += cc * vyc;
        d += cp * vyp;
        d += cm * vzm;
        d += cc * vzc;
        d += cp * vzp;
        div[e] = ad * d;
        sm = sc;
        sc = sp;
        vxm = vxc;
        vxc = vxp;
    }
}

// --------------------------------------------------------- periodic fill

// Ghost layer number l (0 .. 2g-1) along an axis of padded extent P and
// interior extent E: for l < g the ghost index l takes index l + E; for l >= g
// the ghost index P - 2g + l takes index P - 2g + l - E.  With E >= g every
// source is an interior index, so one launch never reads what it writes.
__device__ __forceinline__ void ghost_pair(int64_t l, int64_t g, int64_t P, int64_t E,
                                           int64_t *dst, int64_t *src)
{
    if (l < g) {
        *dst = l;
        *src = l + E;
    } else {
        *dst = P - 2 * g + l;
        *src = *dst - E;
    }
}

// Axis 0 (planes): thread per float of a ghost plane; coalesced.
__global__ __launch_bounds__(256) void periodic_planes(float *__restrict__ u, int64_t plane,
                                                       int64_t g, int64_t P, int64_t E)
{
    const int64_t f = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (f >= plane) return;
    int64_t d, s;
    ghost_pair(blockIdx.y, g, P, E, &d, &s);
    u[d * plane + f] = u[s * plane + f];
}

// Axis 1 (rows): thread per float of a ghost row, grid.y = ghost layer,
// grid.z strides over planes; coalesced.
__global__ __launch_bounds__(256) void periodic_rows(float *__restrict__ u, int64_t row,
                                                     int64_t nplanes, int64_t g, int64_t P,
                                                     int64_t E)
{
    const int64_t f = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (f >= row) return;
    int64_t d, s;
    ghost_pair(blockIdx.y, g, P, E, &d, &s);
    const int64_t plane = P * row;
    for (int64_t i = blockIdx.z; i < nplanes; i += gridDim.z)
        u[i * plane + d * row + f] = u[i * plane + s * row + f];
}

// Axis 2 (the contiguous axis): thread per (row, ghost layer), the layers of
// a row on neighbouring lanes.  Each ghost node costs one dependent load and
// store; with a thread per row walking its 2g layers the round trips were
// serial (20 us at 514^3 against 11-12 for this layout).
__global__ __launch_bounds__(256) void periodic_cols(float *__restrict__ u, int64_t nrows,
                                                     int C, int g, int64_t P, int64_t E)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t r = t / (2 * g);
    if (r >= nrows) return;
    const int l = (int)(t - r * 2 * g);
    float *row = u + r * P * C;
    int64_t d, s;
    ghost_pair(l, g, P, E, &d, &s);
    for (int c = 0; c < C; ++c) row[d * C + c] = row[s * C + c];
}

// All ghost cells of a 3D field (three eligible axes) in one launch: ghost
// cell (x, y, z) takes interior cell (wrap x, wrap y, wrap z), which is the
// value the plane, row and column passes above leave there one after the
// other (each pass copies from cells the passes before it made final).
// Every source is an interior cell, so no thread reads what another writes.
// Threads [0, np) cover the 2g ghost planes whole, [np, np + nr) the 2g ghost
// rows of every interior plane, the rest the 2g ghost columns of every
// interior row (the host keeps each segment below 2^31 elements).
struct PFill3 {
    uint32_t P1, P2, E0, E1, E2, C, g;
    uint32_t np, nr, nc;
};

__device__ __forceinline__ uint32_t pf_wrap(uint32_t i, uint32_t g, uint32_t E)
{
    return i < g ? i + E : (i >= g + E ? i - E : i);
}

// ghost layer l (0 .. 2g-1) of an axis with padded extent P
__device__ __forceinline__ uint32_t pf_layer(uint32_t l, uint32_t g, uint32_t P)
{
    return l < g ? l : P - 2 * g + l;
}

__global__ __launch_bounds__(256) void periodic_fill3(float *__restrict__ u, PFill3 m)
{
    const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t g = m.g, C = m.C;
    uint32_t x, y, z, c;
    if (t < m.np) {
        const uint32_t i = (uint32_t)t, plane = m.P1 * m.P2 * C;
        const uint32_t l = i / plane, f = i - l * plane;
        x = pf_layer(l, g, m.E0 + 2 * g);
        y = f / (m.P2 * C);
        const uint32_t f2 = f - y * (m.P2 * C);
        z = f2 / C;
        c = f2 - z * C;
    } else if (t < (uint64_t)m.np + m.nr) {
        const uint32_t i = (uint32_t)(t - m.np), row = m.P2 * C;
        const uint32_t k = i / row, f = i - k * row;
        const uint32_t xi = k / (2 * g), l = k - xi * (2 * g);
        x = g + xi;
        y = pf_layer(l, g, m.P1);
        z = f / C;
        c = f - z * C;
    } else {
        const uint64_t i64 = t - m.np - m.nr;
        if (i64 >= m.nc) return;
        const uint32_t i = (uint32_t)i64, per = 2 * g * C;
        const uint32_t r = i / per, f = i - r * per;
        const uint32_t l = f / C;
        c = f - l * C;
        const uint32_t xi = r / m.E1;
        x = g + xi;
        y = g + (r - xi * m.E1);
        z = pf_layer(l, g, m.P2);
    }
    const uint32_t sx = pf_wrap(x, g, m.E0), sy = pf_wrap(y, g, m.E1), sz = pf_wrap(z, g, m.E2);
    const uint64_t P1 = m.P1, P2 = m.P2;
    u[((x * P1 + y) * P2 + z) * C + c] = u[((sx * P1 + sy) * P2 + sz) * C + c];
}

// ------------------------------------------------------------ dispatchers

static dim3 row_grid(const RowGeom &rg)
{
    return dim3(clamp_grid((rg.R + 255) / 256, 0x7fffffff), clamp_grid(rg.E1, 65535),
                clamp_grid(rg.E0, 65535));
}

}  // namespace pde

namespace pde {

// The band kernel needs at least LB_TB interior rows (every tile has LB_TB
// rows); smaller grids take the generic kernels.
static bool lap3d_band_fits(const Geom &G) { return G.E[1] >= LB_TB; }

// Blocks of a band launch with TB-row tiles and NXC-plane chunks.
static int64_t band_blocks(const Geom &G, int64_t TB, int64_t NXC)
{
    const int64_t T = ((G.E[2] + LB_TZ - 1) / LB_TZ) * ((G.E[1] + TB - 1) / TB);
    return 8 * ((T + 7) / 8) * ((G.E[0] + NXC - 1) / NXC);
}

template <int OPS, bool PER, int TB = LB_TB, int NXC = LB_NXC>
static int ops3d_band_launch(const float *s_, const float *v, float *lap, float *grad,
                             float *div, const Geom &G, int ghost, const BandW &W,
                             hipStream_t s)
{
    if (G.E[1] < TB) return PDE_ESHAPE;
    Band3 m;
    m.SX = G.S[0];
    m.SY = G.S[1];
    m.E0 = G.E[0];
    m.E1 = G.E[1];
    m.E2 = G.E[2];
    m.g = ghost;
    const int64_t tz = (G.E[2] + LB_TZ - 1) / LB_TZ;
    const int64_t ty = (G.E[1] + TB - 1) / TB;
    const int64_t T = tz * ty;
    const int64_t per = (T + 7) / 8;
    const int64_t nxt = (G.E[0] + NXC - 1) / NXC;
    const int64_t nb = 8 * per * nxt;
    if (ty > 0xffffffffLL || T > 0xffffffffLL || nb > 0x7fffffff) return PDE_ESHAPE;
    m.ty = (uint32_t)ty;
    m.T = (uint32_t)T;
    m.per = (uint32_t)per;
    hipLaunchKernelGGL((ops3d_band<OPS, PER, TB, NXC>), dim3((unsigned)nb), dim3(128, TB), 0, s, s_, v,
                       lap, grad, div, m, W);
    return launch_status();
}

// The LDS plane rings take 3D grids whose rows split into whole wave
// segments (zt points) and that have at least tb rows; extents stay below
// 2^31.
static bool ring_fits(const Geom &G, int64_t zt, int64_t tb)
{
    return G.E[2] % zt == 0 && G.E[1] >= tb && G.E[0] < 0x7fffffffLL &&
           G.E[1] < 0x7fffffffLL && G.E[2] < 0x7fffffffLL;
}

template <bool PER>
static int lap3d_band_launch(const float *in, float *out, const Geom &G, int ghost, float wm,
                             float wc, float wp, float alpha, hipStream_t s)
{
    if (ring_fits(G, 256, RT_TB))
        return G.E[2] % 512 == 0
                   ? lap3d_ring_launch<PER, 2>(in, out, G, ghost, wm, wc, wp, alpha, s)
                   : lap3d_ring_launch<PER, 1>(in, out, G, ghost, wm, wc, wp, alpha, s);
    BandW W = {wm, wc, wp, alpha, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    return ops3d_band_launch<BAND_LAP, PER>(in, nullptr, out, nullptr, nullptr, G, ghost, W, s);
}

}  // namespace pde

using namespace pde;

extern "C" int pde_laplacian_f32(const float *in, float *out, int64_t nx, int64_t ny,
                                 int64_t nz, int32_t ncomp, int32_t ghost, float wm,
                                 float wc, float wp, float alpha, void *stream)
{
    if (!in || !out || ncomp < 1) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    hipStream_t s = as_stream(stream);
    if (G.ndim == 3 && ncomp == 1 && lap3d_band_fits(G))
        return lap3d_band_launch<false>(in, out, G, ghost, wm, wc, wp, alpha, s);
    RowGeom rg = row_geom(G, ncomp);
    dim3 grid = row_grid(rg), block(256);
    switch (G.ndim) {
    case 0: hipLaunchKernelGGL(lap_generic<0>, grid, block, 0, s, in, out, rg, wm, wc, wp, alpha); break;
    case 1: hipLaunchKernelGGL(lap_generic<1>, grid, block, 0, s, in, out, rg, wm, wc, wp, alpha); break;
    case 2: hipLaunchKernelGGL(lap_generic<2>, grid, block, 0, s, in, out, rg, wm, wc, wp, alpha); break;
    default: hipLaunchKernelGGL(lap_generic<3>, grid, block, 0, s, in, out, rg, wm, wc, wp, alpha); break;
    }
    return launch_status();
}

// Generic (non-marching) Laplacian, exported for tests and benchmarks that
// compare the marching kernel with the straightforward one.
extern "C" int pde_laplacian_generic_f32(const float *in, float *out, int64_t nx, int64_t ny,
                                         int64_t nz, int32_t ncomp, int32_t ghost, float wm,
                                         float wc, float wp, float alpha, void *stream)
{
    if (!in || !out || ncomp < 1) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    hipStream_t s = as_stream(stream);
    RowGeom rg = row_geom(G, ncomp);
    dim3 grid = row_grid(rg), block(256);
    switch (G.ndim) {
    case 0: hipLaunchKernelGGL(lap_generic<0>, grid, block, 0, s, in, out, rg, wm, wc, wp, alpha); break;
    case 1: hipLaunchKernelGGL(lap_generic<1>, grid, block, 0, s, in, out, rg, wm, wc, wp, alpha); break;
    case 2: hipLaunchKernelGGL(lap_generic<2>, grid, block, 0, s, in, out, rg, wm, wc, wp, alpha); break;
    default: hipLaunchKernelGGL(lap_generic<3>, grid, block, 0, s, in, out, rg, wm, wc, wp, alpha); break;
    }
    return launch_status();
}

// Periodic Laplacian: as pde_laplacian_f32 on a field whose ghost layers were
// filled periodically (pde_periodic_fill_f32), without reading or writing any
// ghost.  Every eligible interior extent must be >= 1.
extern "C" int pde_laplacian_periodic_f32(const float *in, float *out, int64_t nx, int64_t ny,
                                          int64_t nz, int32_t ncomp, int32_t ghost, float wm,
                                          float wc, float wp, float alpha, void *stream)
{
    if (!in || !out || ncomp < 1) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim == 0) return PDE_ESHAPE;
    hipStream_t s = as_stream(stream);
    if (G.ndim == 3 && ncomp == 1 && lap3d_band_fits(G))
        return lap3d_band_launch<true>(in, out, G, ghost, wm, wc, wp, alpha, s);
    RowGeom rg = row_geom(G, ncomp);
    dim3 grid = row_grid(rg), block(256);
    const int64_t C = ncomp;
    switch (G.ndim) {
    case 1: hipLaunchKernelGGL(lap_generic_periodic<1>, grid, block, 0, s, in, out, rg, C, wm, wc, wp, alpha); break;
    case 2: hipLaunchKernelGGL(lap_generic_periodic<2>, grid, block, 0, s, in, out, rg, C, wm, wc, wp, alpha); break;
    default: hipLaunchKernelGGL(lap_generic_periodic<3>, grid, block, 0, s, in, out, rg, C, wm, wc, wp, alpha); break;
    }
    return launch_status();
}

extern "C" int pde_grad_f32(const float *in, float *out, int64_t nx, int64_t ny, int64_t nz,
                            int32_t ncomp, int32_t ghost, float wm, float wc, float wp,
                            float alpha, void *stream)
{
    if (!in || !out || ncomp < 1) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim == 0) return PDE_ESHAPE;
    hipStream_t s = as_stream(stream);
    if (G.ndim == 3 && ncomp == 1 && lap3d_band_fits(G)) {
        const BandW W = {0.0f, 0.0f, 0.0f, 0.0f, wm, wc, wp, alpha, 0.0f};
        // 8-row tiles where the grid has 8 rows (fewer halo rows per output
        // row; measured 415 vs 439 us at 512^3), else 4-row tiles.  Grad
        // writes three floats per node read: 2-plane chunks (twice the blocks
        // of 4-plane ones) measured 384 against 393 us at 512^3, 1 plane 478
        if (G.E[1] >= 8)
            return ops3d_band_launch<BAND_GRAD, false, 8, 2>(in, nullptr, nullptr, out, nullptr, G,
                                                             ghost, W, s);
        return ops3d_band_launch<BAND_GRAD, false>(in, nullptr, nullptr, out, nullptr, G, ghost,
                                                   W, s);
    }
    RowGeom rg = row_geom(G, ncomp);
    dim3 grid = row_grid(rg), block(256);
    switch (G.ndim) {
    case 1: hipLaunchKernelGGL(grad_generic<1>, grid, block, 0, s, in, out, rg, wm, wc, wp, alpha); break;
    case 2: hipLaunchKernelGGL(grad_generic<2>, grid, block, 0, s, in, out, rg, wm, wc, wp, alpha); break;
    default: hipLaunchKernelGGL(grad_generic<3>, grid, block, 0, s, in, out, rg, wm, wc, wp, alpha); break;
    }
    return launch_status();
}

extern "C" int pde_divergence_f32(const float *in, float *out, int64_t nx, int64_t ny,
                                  int64_t nz, int32_t nrows, int32_t ghost, float wm,
                                  float wc, float wp, float alpha, void *stream)
{
    if (!in || !out || nrows < 1) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim == 0) return PDE_ESHAPE;
    hipStream_t s = as_stream(stream);
    if (G.ndim == 3 && nrows == 1 && ring_fits(G, 256, DR_TB))
        return div3d_ring_run(in, out, G, ghost, wm, wc, wp, alpha, s);
    if (G.ndim == 3 && nrows == 1 && lap3d_band_fits(G)) {
        const BandW W = {0.0f, 0.0f, 0.0f, 0.0f, wm, wc, wp, 0.0f, alpha};
        if (G.E[1] >= 8) {   // 8-row tiles: 450 vs 456 us at 512^3
            // the vec3 planes fill L2 three times as fast as a scalar's, so a
            // chunk's boundary planes are mostly re-read from memory by the
            // next chunk: 16-plane chunks re-read fewer of them (434 against
            // 448 us at 512^3; 32 planes 433, 64 planes 431) wherever they
            // still give every CU 4 blocks
            if (band_blocks(G, 8, 16) >= 1024)
                return ops3d_band_launch<BAND_DIV, false, 8, 16>(nullptr, in, nullptr, nullptr,
                                                                 out, G, ghost, W, s);
            return ops3d_band_launch<BAND_DIV, false, 8>(nullptr, in, nullptr, nullptr, out, G,
                                                         ghost, W, s);
        }
        return ops3d_band_launch<BAND_DIV, false>(nullptr, in, nullptr, nullptr, out, G, ghost, W,
                                                  s);
    }
    RowGeom rg = row_geom(G, nrows);
    dim3 grid = row_grid(rg), block(256);
    switch (G.ndim) {
    case 1: hipLaunchKernelGGL(div_generic<1>, grid, block, 0, s, in, out, rg, wm, wc, wp, alpha); break;
    case 2: hipLaunchKernelGGL(div_generic<2>, grid, block, 0, s, in, out, rg, wm, wc, wp, alpha); break;
    default: hipLaunchKernelGGL(div_generic<3>, grid, block, 0, s, in, out, rg, wm, wc, wp, alpha); break;
    }
    return launch_status();
}

extern "C" int pde_divergence_batch_f32(const float *in, float *out, int64_t batch, int64_t nx,
                                        int64_t ny, int64_t nz, int32_t nrows, int32_t ghost,
                                        float wm, float wc, float wp, float alpha, void *stream)
{
    if (!in || !out || nrows < 1 || batch < 1) return PDE_EINVAL;
    if (batch > 65535) return PDE_ESHAPE;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim == 0 || G.ndim == 3) return PDE_ESHAPE;
    // canonical axis 0 is not a grid axis of a 1D or 2D grid: it walks the
    // instances (stride = one instance's nodes)
    RowGeom rg = row_geom(G, nrows);
    rg.E0 = batch;
    dim3 grid = row_grid(rg), block(256);
    hipStream_t s = as_stream(stream);
    if (G.ndim == 1)
        hipLaunchKernelGGL(div_generic<1>, grid, block, 0, s, in, out, rg, wm, wc, wp, alpha);
    else
        hipLaunchKernelGGL(div_generic<2>, grid, block, 0, s, in, out, rg, wm, wc, wp, alpha);
    return launch_status();
}

extern "C" int pde_fused_lap_grad_div_f32(const float *s_, const float *v, float *lap,
                                          float *grad, float *div, int64_t nx, int64_t ny,
                                          int64_t nz, int32_t ghost, float lm, float lc,
                                          float lp, float cm, float cc, float cp, float al,
                                          float ag, float ad, void *stream)
{
    if (!s_ || !v || !lap || !grad || !div) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim != 3) return PDE_EINVAL;
    if (lap3d_band_fits(G)) {
        const BandW W = {lm, lc, lp, al, cm, cc, cp, ag, ad};
        hipStream_t s = as_stream(stream);
        // Grids the Divergence ring takes: the scalar half (Laplacian and Grad
        // of s, one band launch with Grad's 8-row 2-plane tiles) and the
        // vector half (the ring) as two launches.  Each input is still read
        // once and each output written once (the same 36 bytes per node), and
        // the pair measured 872 against 961 us for the single fused band
        // launch at 512^3, whose scalar and vector paths together need 137
        // VGPRs (3 waves per SIMD).
        if (G.E[1] >= 8 && ring_fits(G, 256, DR_TB)) {
            st = ops3d_band_launch<BAND_LAP | BAND_GRAD, false, 8, 2>(s_, nullptr, lap, grad,
                                                                    nullptr, G, ghost, W, s);
            if (st) return st;
            return div3d_ring_run(v, div, G, ghost, cm, cc, cp, ad, s);
        }
        return ops3d_band_launch<BAND_LAP | BAND_GRAD | BAND_DIV, false>(s_, v, lap, grad, div, G,
                                                                        ghost, W, s);
    }
    constexpr int BY = 4, NXC = 64;
    March3 m;
    m.SX = G.S[0];
    m.SY = G.S[1];
    m.E0 = G.E[0];
    m.E1 = G.E[1];
    m.E2 = G.E[2];
    m.g = ghost;
    m.tz = (uint32_t)((G.E[2] + 63) / 64);
    m.ty = (uint32_t)((G.E[1] + BY - 1) / BY);
    const int64_t tx = (G.E[0] + NXC - 1) / NXC;
    const int64_t nb = (int64_t)m.tz * m.ty * tx;
    if (nb > 0x7fffffff) return PDE_ESHAPE;
    hipLaunchKernelGGL((fused3d_march<BY, NXC>), dim3((unsigned)nb), dim3(64, BY), 0,
                       as_stream(stream), s_, v, lap, grad, div, m, lm, lc, lp, cm, cc, cp, al,
                       ag, ad);
    return launch_status();
}

extern "C" int pde_periodic_fill_f32(float *u, int64_t nx, int64_t ny, int64_t nz,
                                     int32_t ncomp, int32_t ghost, void *stream)
{
    if (!u || ncomp < 1) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    for (int A = 3 - G.ndim; A < 3; ++A)
        if (G.E[A] < ghost) return PDE_ESHAPE;  // a ghost's source must be interior
    hipStream_t s = as_stream(stream);
    const int64_t C = ncomp;
    if (G.ndim == 3) {
        // one launch (periodic_fill3) where every segment's element count
        // fits 32 bits (fill + Laplacian at 512^3: 199.1 us against 200.8-201.4
        // with the three passes; the columns' scattered 64-byte sectors are
        // most of the fill's time either way)
        const int64_t g2 = 2 * (int64_t)ghost;
        const int64_t np = g2 * G.P[1] * G.P[2] * C, nr = G.E[0] * g2 * G.P[2] * C,
                      nc = G.E[0] * G.E[1] * g2 * C;
        const int64_t lim = 0x7fffffff;
        if (np < lim && nr < lim && nc < lim && G.P[0] * G.P[1] * G.P[2] < lim) {
            PFill3 m;
            m.P1 = (uint32_t)G.P[1];
            m.P2 = (uint32_t)G.P[2];
            m.E0 = (uint32_t)G.E[0];
            m.E1 = (uint32_t)G.E[1];
            m.E2 = (uint32_t)G.E[2];
            m.C = (uint32_t)C;
            m.g = (uint32_t)ghost;
            m.np = (uint32_t)np;
            m.nr = (uint32_t)nr;
            m.nc = (uint32_t)nc;
            const int64_t nb = (np + nr + nc + 255) / 256;
            hipLaunchKernelGGL(periodic_fill3, dim3((unsigned)nb), dim3(256), 0, s, u, m);
            return launch_status();
        }
    }
    for (int A = 3 - G.ndim; A < 3; ++A) {
        const int64_t P = G.P[A], E = G.E[A];
        const unsigned layers = (unsigned)(2 * ghost);
        if (A == 0) {
            const int64_t plane = G.S[0] * C;
            hipLaunchKernelGGL(periodic_planes, dim3((unsigned)((plane + 255) / 256), layers),
                               dim3(256), 0, s, u, plane, (int64_t)ghost, P, E);
        } else if (A == 1) {
            const int64_t row = G.S[1] * C;
            hipLaunchKernelGGL(periodic_rows,
                               dim3((unsigned)((row + 255) / 256), layers,
                                    clamp_grid(G.P[0], 65535)),
                               dim3(256), 0, s, u, row, G.P[0], (int64_t)ghost, P, E);
        } else {
            const int64_t nrows = G.P[0] * G.P[1];
            const int64_t nb = (nrows * 2 * ghost + 255) / 256;
            if (nb > 0x7fffffff) return PDE_ESHAPE;
            hipLaunchKernelGGL(periodic_cols, dim3((unsigned)nb), dim3(256), 0, s, u, nrows,
                               (int)C, (int)ghost, P, E);
        }
        st = launch_status();
        if (st) return st;
    }
    return PDE_OK;
}
