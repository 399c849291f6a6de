// Time-reversed transposed wave steps, two per launch.
//
// The u/v wave step (plain Euler on step 0, symplectic afterwards) is
// invertible on interior nodes: u_next = u2 + dt*v_next gives u2 = u_next -
// dt*v_next, and v_next = v + dt*c2*alpha*lap(u2) gives v back.  The boundary
// is not invertible, so the forward records the scalars it overwrites (the
// save rows of pde_bc_apply_chunked_inplace_save_f32) and the walk puts them
// back.  A gradient pass can then rebuild each step's state from the next one
// while it runs the transposed step, instead of recomputing the segment from a
// checkpoint and reading the states back.
//
// One launch covers steps n and n-1.  Reads: u, v, g_u, g_v, c2 and gc2 at
// step n+1; writes the same five at step n-1 (44 bytes per node for two
// steps).  Two kinds of workgroups share the launch:
//  * interior tiles (nodes at least `depth` from the edge): a wave owns 64
//    lanes of 4 columns and walks PR_ROWS output rows; step-n rows are
//    computed one row and one lane beyond the output (strips overlap by one
//    lane per side, 248 output columns each), so the second step needs no
//    exchange.  The boundary touches no node these tiles compute, so none of
//    its work appears here.  Waves whose nodes are all interior take a fast
//    path (no interior tests, DPP lane shifts, 32-bit offsets; adjacent waves
//    walk in opposite directions so shared halo rows are loaded together);
//    the rest take a general path with the interior tests of steps2d.hip.
//  * ring pieces (the outer `depth` rows and columns, up to 32 + depth nodes
//    along the edge), launched after the tiles: the piece grown by two nodes
//    is staged in LDS; step n is computed on the piece grown by one, the
//    saved boundary scalars and the boundary transpose are applied there,
//    then step n-1 on the piece itself, with its own boundary fix-ups before
//    the stores.  The host checks that every boundary transpose a piece needs
//    stays inside it.
// Arithmetic mirrors steps2d.hip's transposed step operation by operation, so
// g_u and g_v equal two fused transposed steps plus boundary transposes bit
// for bit; only gc2 sees the rounding of the rebuilt states.
#include "tile2d.hpp"

namespace pde {

constexpr int PR_ROWS = 9;     // output rows per wave (a multiple of 3; 9: 42.1 us per pair at
                               // 2048^2 against 44.4 with 6, 42.4 with 3, 43.8 with 12)
constexpr int PR_BY = 4;       // waves per workgroup, stacked along rows
constexpr int PR_COLS = 248;   // output columns per strip (lanes 1 .. 62)
constexpr int RING_CAP = 448;  // nodes of a ring piece grown by two
constexpr int RING_ARR = 12;   // LDS arrays of a ring piece

struct PairArgs {
    const float *u, *v, *gu, *gv, *c2;
    float *uo, *vo, *guo, *gvo, *gc2;
    float wm, wc, wp, alpha, dt;
    const int64_t *step;
    int64_t *step_next;
    pde_point_set src, rec;
    const float *save;
    int64_t nops;
    pde_ring_plan ring;
};

struct PairGeom {
    int64_t P1, P2, g, D;
    uint32_t nstrips, nrt;   // interior strips and row tiles
    uint32_t nring;          // ring workgroups (a multiple of 8)
    uint32_t nint;           // interior workgroups (a multiple of 8)
};

// Subtracts scale * data[n][slot] of the points [a, b) from the elements of v
// in columns k0 .. k0+3, walking the list last to first (the forward added
// them first to last).  Every lane of the wave calls it with the same range.
__device__ __forceinline__ void pts_sub(const pde_point_set &P, int a, int b, int64_t k0,
                                        int64_t n, int lane, v4 &v)
{
    for (int top = b; top > a; top -= 64) {
        const int base = top - 64 > a ? top - 64 : a;
        const int p = base + lane;
        int col = -1;
        float val = 0.0f;
        if (p < top) {
            col = P.col[p];
            val = P.scale * P.data[n * P.stride + P.slot[p]];
        }
        for (int i = top - base - 1; i >= 0; --i) {
            const int64_t c = __builtin_amdgcn_readlane(col, i);
            const float d = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(val), i));
            const int64_t dd = c - k0;
            if (dd >= 0 && dd < 4) set_el(v, (int)dd, el(v, (int)dd) - d);
        }
    }
}

// ------------------------------------------------------------ interior tiles

struct Row1 {            // a step-(n+1) row of a lane's 4 columns
    v4 u2, h;            // rebuilt u2 of step n; h = c2 * (g_v + dt*g_u) (0 off the interior)
    v4 V, GU, GV, C;
    float eu2, eh;       // lanes 0 / 63: u2 and h one column outside the strip
};

struct RowN {            // a step-n row
    v4 u2, h;            // u2 of step n-1 (u_n - dt*v_n) and its h
    v4 vn, gun, gvn, C, T1;
};

__device__ __forceinline__ Row1 pair_load1(const PairArgs &A, const Tile2 &t, int64_t x,
                                           int64_t k0, int lane, const bool *kin, int64_t n,
                                           bool has_rec)
{
    Row1 r;
    const v4 U = ld_row4(A.u, t, x, k0);
    r.V = ld_row4(A.v, t, x, k0);
    r.GU = ld_row4(A.gu, t, x, k0);
    r.GV = ld_row4(A.gv, t, x, k0);
    r.C = A.c2 ? ld_row4(A.c2, t, x, k0) : v4{1.0f, 1.0f, 1.0f, 1.0f};
    const bool xin = inside(x, t.g, t.P1);
    const int64_t kk = lane == 0 ? k0 - 1 : k0 + 4;
    const bool e_ok = (lane == 0 || lane == 63) && kk >= 0 && kk < t.P2 && x >= 0 && x < t.P1;
    float eU = 0.0f, eV = 0.0f, eGU = 0.0f, eGV = 0.0f, eC = 1.0f;
    if (e_ok) {
        const int64_t e = x * t.P2 + kk;
        eU = A.u[e];
        eV = A.v[e];
        eGU = A.gu[e];
        eGV = A.gv[e];
        if (A.c2) eC = A.c2[e];
    }
    if (has_rec) {
        int a, b;
        pts_range(A.rec, x, t.P1, a, b);
        pts_add(A.rec, a, b, k0, n, lane, r.GU, e_ok ? kk : -1, eGU);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const bool in_ = xin && kin[i];
        const float uu = el(U, i);
        set_el(r.u2, i, in_ ? uu - A.dt * el(r.V, i) : uu);
        const float tv = el(r.GV, i) + A.dt * el(r.GU, i);
        set_el(r.h, i, in_ ? el(r.C, i) * tv : 0.0f);
    }
    const bool ein = e_ok && xin && inside(kk, t.g, t.P2);
    r.eu2 = ein ? eU - A.dt * eV : eU;
    r.eh = ein ? eC * (eGV + A.dt * eGU) : 0.0f;
    return r;
}

// Step n at row x from step-(n+1) rows x-1, x, x+1.
__device__ __forceinline__ RowN pair_leveln(const PairArgs &A, const Tile2 &t, const Row1 &up,
                                            const Row1 &mid, const Row1 &dn, int64_t x,
                                            int64_t k0, int lane, const bool *kin,
                                            const bool *lin, const bool *rin, int64_t n,
                                            bool has_src, bool has_rec)
{
    float ul = __shfl_up(mid.u2.w, 1, 64), ur = __shfl_down(mid.u2.x, 1, 64);
    float hl = __shfl_up(mid.h.w, 1, 64), hr = __shfl_down(mid.h.x, 1, 64);
    if (lane == 0) {
        ul = mid.eu2;
        hl = mid.eh;
    } else if (lane == 63) {
        ur = mid.eu2;
        hr = mid.eh;
    }
    const bool xin = inside(x, t.g, t.P1);
    const bool upin = inside(x - 1, t.g, t.P1), dnin = inside(x + 1, t.g, t.P1);
    RowN r;
    r.C = mid.C;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const bool in_ = xin && kin[i];
        const float Lu = acc5(up.u2, mid.u2, dn.u2, ul, ur, i, A.wm, A.wc, A.wp);
        const float rhs = el(mid.C, i) * (A.alpha * Lu);
        set_el(r.vn, i, in_ ? el(mid.V, i) - A.dt * rhs : el(mid.V, i));
        const float tv = el(mid.GV, i) + A.dt * el(mid.GU, i);
        const float at = acct5(up.h, mid.h, dn.h, hl, hr, i, xin, upin, dnin, kin[i], lin[i],
                               rin[i], A.wm, A.wc, A.wp);
        set_el(r.gun, i, el(mid.GU, i) + A.dt * (A.alpha * at));
        set_el(r.gvn, i, in_ ? tv : el(mid.GV, i));
        set_el(r.T1, i, in_ ? A.dt * (tv * (A.alpha * Lu)) : 0.0f);
    }
    if (has_src) {
        int a, b;
        pts_range(A.src, x, t.P1, a, b);
        pts_sub(A.src, a, b, k0, n, lane, r.vn);
    }
    if (has_rec) {
        int a, b;
        float none = 0.0f;
        pts_range(A.rec, x, t.P1, a, b);
        pts_add(A.rec, a, b, k0, n - 1, lane, r.gun, -1, none);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const bool in_ = xin && kin[i];
        const float un = el(mid.u2, i);
        set_el(r.u2, i, in_ ? un - A.dt * el(r.vn, i) : un);
        const float tv = el(r.gvn, i) + A.dt * el(r.gun, i);
        set_el(r.h, i, in_ ? el(r.C, i) * tv : 0.0f);
    }
    return r;
}

__device__ __forceinline__ void st_masked(float *f, const Tile2 &t, int64_t x, int64_t k0,
                                          int64_t kend, const v4 &v)
{
    if (k0 + 4 <= kend) {
        st_row4(f, t, x, k0, v);
        return;
    }
    float *p = f + x * t.P2 + k0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (k0 + i < kend) p[i] = el(v, i);
}

// Step n-1 at row x from step-n rows x-1, x, x+1; stores when `store`.
__device__ __forceinline__ void pair_out(const PairArgs &A, const Tile2 &t, const RowN &up,
                                         const RowN &mid, const RowN &dn, int64_t x, int64_t k0,
                                         int64_t kend, const bool *kin, const bool *lin,
                                         const bool *rin, int64_t n, int lane, bool has_src,
                                         bool store, v4 gc)
{
    // lanes 0 and 63 read no neighbour outside the strip: their results are
    // never stored
    const float ul = __shfl_up(mid.u2.w, 1, 64), ur = __shfl_down(mid.u2.x, 1, 64);
    const float hl = __shfl_up(mid.h.w, 1, 64), hr = __shfl_down(mid.h.x, 1, 64);
    const bool xin = inside(x, t.g, t.P1);
    const bool upin = inside(x - 1, t.g, t.P1), dnin = inside(x + 1, t.g, t.P1);
    v4 vo, guo, gvo;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const bool in_ = xin && kin[i];
        const float Lu = acc5(up.u2, mid.u2, dn.u2, ul, ur, i, A.wm, A.wc, A.wp);
        const float rhs = el(mid.C, i) * (A.alpha * Lu);
        set_el(vo, i, in_ ? el(mid.vn, i) - A.dt * rhs : el(mid.vn, i));
        const float at = acct5(up.h, mid.h, dn.h, hl, hr, i, xin, upin, dnin, kin[i], lin[i],
                               rin[i], A.wm, A.wc, A.wp);
        set_el(guo, i, el(mid.gun, i) + A.dt * (A.alpha * at));
        const float tv = el(mid.gvn, i) + A.dt * el(mid.gun, i);
        set_el(gvo, i, in_ ? tv : el(mid.gvn, i));
        if (in_) set_el(gc, i, (el(gc, i) + el(mid.T1, i)) + A.dt * (tv * (A.alpha * Lu)));
    }
    if (has_src) {
        int a, b;
        pts_range(A.src, x, t.P1, a, b);
        pts_sub(A.src, a, b, k0, n - 1, lane, vo);
    }
    if (!store) return;
    st_masked(A.uo, t, x, k0, kend, mid.u2);
    st_masked(A.vo, t, x, k0, kend, vo);
    st_masked(A.guo, t, x, k0, kend, guo);
    st_masked(A.gvo, t, x, k0, kend, gvo);
    if (A.gc2) st_masked(A.gc2, t, x, k0, kend, gc);
}

// ------------------------------------------------------- interior fast path
// Waves whose rows and columns (with the one-column edges of lanes 0 and 63)
// are all interior and hold no point: every node is updated, so the interior
// tests, the point walks and the masked stores drop out.  Same operations in
// the same order as the general path.

// Whole-wave shifts by one lane (DPP wave_shr:1 / wave_shl:1): lane l receives
// lane l-1's value (from_left) or lane l+1's (from_right); the lane past the
// end receives 0.
__device__ __forceinline__ float from_left(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}

__device__ __forceinline__ float from_right(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}

// 16 bytes at byte offset b (< 2^32: the entry checks the field size) from a
// base pointer, so the access is a scalar base plus a 32-bit lane offset.
__device__ __forceinline__ v4 ldv(const float *base, uint32_t b)
{
    const v4u q = *reinterpret_cast<const v4u *>(reinterpret_cast<const char *>(base) + b);
    return v4{q.x, q.y, q.z, q.w};
}

__device__ __forceinline__ float ld1(const float *base, uint32_t b)
{
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + b);
}

__device__ __forceinline__ void stv(float *base, uint32_t b, const v4 &v)
{
    *reinterpret_cast<v4u *>(reinterpret_cast<char *>(base) + b) = v4u{v.x, v.y, v.z, v.w};
}

// acct5 with every node interior
__device__ __forceinline__ float acct5_in(const v4 &up, const v4 &mid, const v4 &dn, float left,
                                          float right, int i, float wm, float wc, float wp)
{
    const float zr = i == 3 ? right : el(mid, i + 1);
    const float zl = i == 0 ? left : el(mid, i - 1);
    float acc = 0.0f;
    acc += wm * el(dn, i);
    acc += wc * el(mid, i);
    acc += wp * el(up, i);
    acc += wm * zr;
    acc += wc * el(mid, i);
    acc += wp * zl;
    return acc;
}

struct F1 {              // a step-(n+1) row, every node interior
    v4 u2, h, V, GU, gvt, C;
};

struct FN {              // a step-n row: u2 / h of step n-1, v_n, g_u, gvt of step n, and
    v4 u2, h, vn, gun, gvt, gcp;   // gc2 with step n's term added
};

// Row x of step n+1 as loaded (the lane's 4 columns at byte offset o)...
struct Raw {
    v4 U, V, GU, GV, C;
};

template <bool C2>
__device__ __forceinline__ void fast_raw(const PairArgs &A, uint32_t o, Raw &w)
{
    w.U = ldv(A.u, o);
    w.V = ldv(A.v, o);
    w.GU = ldv(A.gu, o);
    w.GV = ldv(A.gv, o);
    w.C = C2 ? ldv(A.c2, o) : v4{1.0f, 1.0f, 1.0f, 1.0f};
}

// ... and its step-n quantities, with the receivers' step-n gradients added
// to g_u when the wave's rows hold any.
template <bool C2>
__device__ __forceinline__ void fast_finish1(const PairArgs &A, Raw &w, int64_t x, int64_t k0,
                                             int lane, int64_t n, int64_t P1, bool has_rec,
                                             F1 &r)
{
    if (has_rec) {
        int a, b;
        float none = 0.0f;
        pts_range(A.rec, x, P1, a, b);
        pts_add(A.rec, a, b, k0, n, lane, w.GU, -1, none);
    }
    const float dt = A.dt;
    r.V = w.V;
    r.GU = w.GU;
    r.C = w.C;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        set_el(r.u2, i, el(w.U, i) - dt * el(w.V, i));
        const float tv = el(w.GV, i) + dt * el(w.GU, i);
        set_el(r.gvt, i, tv);
        set_el(r.h, i, C2 ? el(w.C, i) * tv : tv);
    }
}

template <bool C2>
__device__ __forceinline__ void fast_leveln(const PairArgs &A, const F1 &up, const F1 &mid,
                                            const F1 &dn, int64_t x, int64_t k0, int lane,
                                            int64_t n, int64_t P1, bool has_src, bool has_rec,
                                            const v4 &gc0, FN &r)
{
    // lanes 0 and 63 take 0 beyond the strip: only their inner columns,
    // which never read it, reach a stored value
    const float ul = from_left(mid.u2.w), ur = from_right(mid.u2.x);
    const float hl = from_left(mid.h.w), hr = from_right(mid.h.x);
    const float dt = A.dt, alpha = A.alpha;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float Lu = acc5(up.u2, mid.u2, dn.u2, ul, ur, i, A.wm, A.wc, A.wp);
        const float rhs = C2 ? el(mid.C, i) * (alpha * Lu) : alpha * Lu;
        set_el(r.vn, i, el(mid.V, i) - dt * rhs);
        const float at = acct5_in(up.h, mid.h, dn.h, hl, hr, i, A.wm, A.wc, A.wp);
        set_el(r.gun, i, el(mid.GU, i) + dt * (alpha * at));
        set_el(r.gcp, i, el(gc0, i) + dt * (el(mid.gvt, i) * (alpha * Lu)));
    }
    if (has_src) {
        int a, b;
        pts_range(A.src, x, P1, a, b);
        pts_sub(A.src, a, b, k0, n, lane, r.vn);
    }
    if (has_rec) {
        int a, b;
        float none = 0.0f;
        pts_range(A.rec, x, P1, a, b);
        pts_add(A.rec, a, b, k0, n - 1, lane, r.gun, -1, none);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        set_el(r.u2, i, el(mid.u2, i) - dt * el(r.vn, i));
        const float tv2 = el(mid.gvt, i) + dt * el(r.gun, i);
        set_el(r.gvt, i, tv2);
        set_el(r.h, i, C2 ? el(mid.C, i) * tv2 : tv2);
    }
}

// Stores the lane's 4 values at columns k0 .. k0+3 limited to [lo, hi).
__device__ __forceinline__ void stv_cols(float *base, uint32_t o, const v4 &v, int64_t k0,
                                         int64_t lo, int64_t hi)
{
    if (k0 >= lo && k0 + 4 <= hi) {
        stv(base, o, v);
        return;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (k0 + i >= lo && k0 + i < hi)
            *reinterpret_cast<float *>(reinterpret_cast<char *>(base) + o + 4u * i) = el(v, i);
}

template <bool C2>
__device__ __forceinline__ void fast_out(const PairArgs &A, const FN &up, const FN &mid,
                                         const FN &dn, uint32_t o, int64_t x, int64_t k0,
                                         int lane, int64_t n, int64_t P1, bool has_src,
                                         bool store, int64_t lo, int64_t hi)
{
    const v4 C = C2 ? ldv(A.c2, o) : v4{1.0f, 1.0f, 1.0f, 1.0f};   // reloaded: L2
    const float ul = from_left(mid.u2.w), ur = from_right(mid.u2.x);
    const float hl = from_left(mid.h.w), hr = from_right(mid.h.x);
    const float dt = A.dt, alpha = A.alpha;
    v4 vo, guo, gc;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float Lu = acc5(up.u2, mid.u2, dn.u2, ul, ur, i, A.wm, A.wc, A.wp);
        const float rhs = C2 ? el(C, i) * (alpha * Lu) : alpha * Lu;
        set_el(vo, i, el(mid.vn, i) - dt * rhs);
        const float at = acct5_in(up.h, mid.h, dn.h, hl, hr, i, A.wm, A.wc, A.wp);
        set_el(guo, i, el(mid.gun, i) + dt * (alpha * at));
        set_el(gc, i, el(mid.gcp, i) + dt * (el(mid.gvt, i) * (alpha * Lu)));
    }
    if (has_src) {
        int a, b;
        pts_range(A.src, x, P1, a, b);
        pts_sub(A.src, a, b, k0, n - 1, lane, vo);
    }
    if (!store) return;
    stv_cols(A.uo, o, mid.u2, k0, lo, hi);
    stv_cols(A.vo, o, vo, k0, lo, hi);
    stv_cols(A.guo, o, guo, k0, lo, hi);
    stv_cols(A.gvo, o, mid.gvt, k0, lo, hi);
    if (A.gc2) stv_cols(A.gc2, o, gc, k0, lo, hi);
}

// Rows x0 .. x0+RW-1 (RW a multiple of 3: the row slots rotate with period 3,
// so the loop body is unrolled three times and nothing moves between slots),
// walked downwards (DIR = 1) or upwards (DIR = -1); rows below own_x and
// columns outside [lo, hi) are computed, not stored.  Vertically adjacent
// waves walk in opposite directions, so the halo rows two waves share are
// loaded by both at the same moment (start or end of their walks) and the
// second load hits L2.
template <bool C2, int RW, int DIR>
__device__ __forceinline__ void pair_fast(const PairArgs &A, int64_t P1, int64_t P2, int64_t x0,
                                          int64_t own_x, int64_t k0, int lane, bool lane_out,
                                          int64_t lo, int64_t hi, int64_t n, bool has_src,
                                          bool has_rec)
{
    static_assert(RW % 3 == 0, "rows per wave must be a multiple of 3");
    const uint32_t rowb = (uint32_t)P2 * 4u;
    const uint32_t kb = (uint32_t)k0 * 4u;
    auto off = [&](int64_t x) { return (uint32_t)x * rowb + kb; };
    auto gcrow = [&](int64_t x) {
        return (A.gc2 && lane_out && x >= own_x) ? ldv(A.gc2, off(x))
                                                 : v4{0.0f, 0.0f, 0.0f, 0.0f};
    };
    // step n at row y from rows y - DIR (prev), y, y + DIR (next); g = gc2 row y
    auto LV = [&](const F1 &prev, const F1 &mid, const F1 &next, int64_t y, const v4 &g,
                  FN &r) {
        if (DIR > 0)
            fast_leveln<C2>(A, prev, mid, next, y, k0, lane, n, P1, has_src, has_rec, g, r);
        else
            fast_leveln<C2>(A, next, mid, prev, y, k0, lane, n, P1, has_src, has_rec, g, r);
    };
    auto OUT = [&](const FN &prev, const FN &mid, const FN &next, int64_t y) {
        const bool st = lane_out && y >= own_x;
        if (DIR > 0)
            fast_out<C2>(A, prev, mid, next, off(y), y, k0, lane, n, P1, has_src, st, lo, hi);
        else
            fast_out<C2>(A, next, mid, prev, off(y), y, k0, lane, n, P1, has_src, st, lo, hi);
    };
    auto LD = [&](int64_t y, F1 &r) {
        Raw w;
        fast_raw<C2>(A, off(y), w);
        fast_finish1<C2>(A, w, y, k0, lane, n, P1, has_rec, r);
    };
    const int64_t xs = DIR > 0 ? x0 : x0 + RW - 1;   // first output row of the walk
    const v4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    F1 a, b, c;
    FN p, q, s;
    LD(xs - 2 * DIR, a);
    LD(xs - DIR, b);
    LD(xs, c);
    LV(a, b, c, xs - DIR, zero, p);
    LD(xs + DIR, a);
    LV(b, c, a, xs, gcrow(xs), q);
#pragma unroll 1
    for (int r = 0; r < RW; r += 3) {
        const int64_t x = xs + r * DIR;
        LD(x + 2 * DIR, b);
        LV(c, a, b, x + DIR, gcrow(x + DIR), s);
        OUT(p, q, s, x);
        LD(x + 3 * DIR, c);
        LV(a, b, c, x + 2 * DIR, gcrow(x + 2 * DIR), p);
        OUT(q, s, p, x + DIR);
        LD(x + 4 * DIR, a);
        LV(b, c, a, x + 3 * DIR, gcrow(x + 3 * DIR), q);
        OUT(s, p, q, x + 2 * DIR);
    }
}

// Whether point set P has a point in rows [x0, x1) (the same on every lane).
__device__ __forceinline__ bool pts_any(const pde_point_set &P, int64_t x0, int64_t x1,
                                        int64_t P1)
{
    if (!P.row_start) return false;
    x0 = x0 < 0 ? 0 : x0;
    x1 = x1 > P1 ? P1 : x1;
    if (x1 <= x0) return false;
    return __builtin_amdgcn_readfirstlane(P.row_start[x0]) !=
           __builtin_amdgcn_readfirstlane(P.row_start[x1]);
}

template <int RW>
__device__ __forceinline__ void pair_interior(const PairArgs &A, const PairGeom &pg, uint32_t ib, int64_t n)
{
    const uint32_t per = pg.nint >> 3;
    const uint32_t L = (ib & 7u) * per + (ib >> 3);
    if (L >= pg.nstrips * pg.nrt) return;
    // tiles numbered across the strips first: each XCD (a contiguous run of
    // L) takes a band of rows over every strip, so the edge strips (general
    // path, slower) spread over all XCDs and vertically adjacent tiles, which
    // share halo rows, mostly sit in one L2
    const uint32_t rt = L / pg.nstrips, qt = L % pg.nstrips;
    const int lane = threadIdx.x;
    const int64_t D = pg.D, g = pg.g;
    const int64_t kend = pg.P2 - D, xend = pg.P1 - D;
    // the wave owns output rows [x0, x0 + RW) and columns [own_k, hi) of the
    // interior region
    const int64_t x0 = D + ((int64_t)rt * PR_BY + threadIdx.y) * RW;
    if (x0 >= xend) return;   // the same on every lane of the wave
    const int64_t own_k = D + (int64_t)qt * PR_COLS;
    const int64_t hi = own_k + PR_COLS < kend ? own_k + PR_COLS : kend;
    {
        // Fast layout: a last strip or row band that would leave the region
        // is moved back to end at its edge (its lanes and rows before the
        // owned ones are computed, not stored).  Lanes 0 and 63 need only
        // their two inner columns correct, so the strip needs interior nodes
        // from 2 columns before its first output to 2 after its last.
        const int64_t ks = own_k + PR_COLS > kend && kend - PR_COLS >= D ? kend - PR_COLS : own_k;
        const int64_t xs = x0 + RW > xend && xend - RW >= D ? xend - RW : x0;
        const bool fast = ks - 4 >= 0 && ks - 2 >= g && ks + 249 < pg.P2 - g &&
                          ks + 251 <= pg.P2 - 1 && xs - 2 >= g && xs + RW + 1 < pg.P1 - g;
        if (fast) {
            const bool has_src = pts_any(A.src, xs - 2, xs + RW + 2, pg.P1);
            const bool has_rec = pts_any(A.rec, xs - 2, xs + RW + 2, pg.P1);
            const int64_t k0 = ks + 4 * (lane - 1);
            const bool lane_out = lane >= 1 && lane <= 62;
            // waves alternate their walk direction (odd waves upwards)
            const bool up = threadIdx.y & 1;
            if (A.c2 && !up)
                pair_fast<true, RW, 1>(A, pg.P1, pg.P2, xs, x0, k0, lane, lane_out, own_k, hi, n,
                                       has_src, has_rec);
            else if (A.c2)
                pair_fast<true, RW, -1>(A, pg.P1, pg.P2, xs, x0, k0, lane, lane_out, own_k, hi,
                                        n, has_src, has_rec);
            else if (!up)
                pair_fast<false, RW, 1>(A, pg.P1, pg.P2, xs, x0, k0, lane, lane_out, own_k, hi,
                                        n, has_src, has_rec);
            else
                pair_fast<false, RW, -1>(A, pg.P1, pg.P2, xs, x0, k0, lane, lane_out, own_k, hi,
                                         n, has_src, has_rec);
            return;
        }
    }
    Tile2 t;
    t.P1 = pg.P1;
    t.P2 = pg.P2;
    t.g = pg.g;
    t.nq = (t.P2 + 3) / 4;
    t.tq = t.tr = 0;
    int64_t k0 = own_k + 4 * (lane - 1);   // >= 0: D >= 4
    if (k0 >= t.P2) k0 = t.P2 - 1;   // lanes past the field (never stored)
    const bool lane_out = lane >= 1 && lane <= 62 && k0 < kend;
    bool kin[4], lin[4], rin[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t k = k0 + i;
        kin[i] = inside(k, t.g, t.P2);
        lin[i] = inside(k - 1, t.g, t.P2);
        rin[i] = inside(k + 1, t.g, t.P2);
    }
    // point lists are walked only by waves whose rows hold a point
    const bool has_src = pts_any(A.src, x0 - 2, x0 + RW + 2, t.P1);
    const bool has_rec = pts_any(A.rec, x0 - 2, x0 + RW + 2, t.P1);

    Row1 ra = pair_load1(A, t, x0 - 2, k0, lane, kin, n, has_rec);
    Row1 rb = pair_load1(A, t, x0 - 1, k0, lane, kin, n, has_rec);
    Row1 rc = pair_load1(A, t, x0, k0, lane, kin, n, has_rec);
    RowN p = pair_leveln(A, t, ra, rb, rc, x0 - 1, k0, lane, kin, lin, rin, n, has_src, has_rec);
    ra = rb;
    rb = rc;
    rc = pair_load1(A, t, x0 + 1, k0, lane, kin, n, has_rec);
    RowN q = pair_leveln(A, t, ra, rb, rc, x0, k0, lane, kin, lin, rin, n, has_src, has_rec);
#pragma unroll 1
    for (int r = 0; r < RW; ++r) {
        const int64_t x = x0 + r;
        const bool store = lane_out && x < pg.P1 - D;
        v4 gc = {0.0f, 0.0f, 0.0f, 0.0f};
        if (A.gc2 && store) gc = ld_row4(A.gc2, t, x, k0);   // issued before the row's work
        ra = rb;
        rb = rc;
        rc = pair_load1(A, t, x + 2, k0, lane, kin, n, has_rec);
        const RowN s =
            pair_leveln(A, t, ra, rb, rc, x + 1, k0, lane, kin, lin, rin, n, has_src, has_rec);
        pair_out(A, t, p, q, s, x, k0, kend, kin, lin, rin, n, lane, has_src, store, gc);
        p = q;
        q = s;
    }
}

// ---------------------------------------------------------------- ring pieces

// One thread walks row x's point list ([ka, kb) columns only) and adds
// (sign > 0) or subtracts, in list order or reverse list order, scale *
// data[n][slot] to row[column - k_base].
__device__ __forceinline__ void ring_pts(const pde_point_set &P, int64_t x, int64_t P1,
                                         int64_t ka, int64_t kb, int64_t n, float *row,
                                         int64_t k_base, bool add)
{
    if (x < 0 || x >= P1) return;
    const int a = P.row_start[x], b = P.row_start[x + 1];
    if (add) {
        for (int p = a; p < b; ++p) {
            const int64_t c = P.col[p];
            if (c >= ka && c < kb) row[c - k_base] = row[c - k_base] + P.scale * P.data[n * P.stride + P.slot[p]];
        }
    } else {
        for (int p = b - 1; p >= a; --p) {
            const int64_t c = P.col[p];
            if (c >= ka && c < kb) row[c - k_base] = row[c - k_base] - P.scale * P.data[n * P.stride + P.slot[p]];
        }
    }
}

__device__ __forceinline__ void pair_ring(const PairArgs &A, const PairGeom &pg, int pid, int64_t n, float *lds)
{
    const pde_ring_piece pc = A.ring.pieces[pid];
    const int64_t P1 = pg.P1, P2 = pg.P2, g = pg.g;
    const int tid = threadIdx.y * 64 + threadIdx.x;
    constexpr int NT = 64 * PR_BY;
    // staged region I (the piece grown by 2), step-n region W (grown by 1)
    const int64_t xI0 = pc.r0 - 2 > 0 ? pc.r0 - 2 : 0, xI1 = pc.r1 + 2 < P1 ? pc.r1 + 2 : P1;
    const int64_t kI0 = pc.c0 - 2 > 0 ? pc.c0 - 2 : 0, kI1 = pc.c1 + 2 < P2 ? pc.c1 + 2 : P2;
    const int64_t xW0 = pc.r0 - 1 > 0 ? pc.r0 - 1 : 0, xW1 = pc.r1 + 1 < P1 ? pc.r1 + 1 : P1;
    const int64_t kW0 = pc.c0 - 1 > 0 ? pc.c0 - 1 : 0, kW1 = pc.c1 + 1 < P2 ? pc.c1 + 1 : P2;
    const int wI = (int)(kI1 - kI0), hI = (int)(xI1 - xI0);
    const int nI = wI * hI;
    if (wI <= 0 || hI <= 0 || nI > RING_CAP) return;   // the host never builds such a piece
    float *U = lds, *V = U + RING_CAP, *GU = V + RING_CAP, *GV = GU + RING_CAP;
    float *C = GV + RING_CAP, *H = C + RING_CAP, *UN = H + RING_CAP, *VN = UN + RING_CAP;
    float *GUN = VN + RING_CAP, *GVN = GUN + RING_CAP, *T1 = GVN + RING_CAP, *ST = T1 + RING_CAP;
    auto lidx = [&](int64_t node) -> int {
        const int64_t x = node / P2, k = node - x * P2;
        if (x < xI0 || x >= xI1 || k < kI0 || k >= kI1) return -1;
        return (int)((x - xI0) * wI + (k - kI0));
    };
    const float dt = A.dt, alpha = A.alpha, wm = A.wm, wc = A.wc, wp = A.wp;

    for (int i = tid; i < nI; i += NT) {
        const int64_t x = xI0 + i / wI, k = kI0 + i % wI, e = x * P2 + k;
        U[i] = A.u[e];
        V[i] = A.v[e];
        GU[i] = A.gu[e];
        GV[i] = A.gv[e];
        C[i] = A.c2 ? A.c2[e] : 1.0f;
    }
    __syncthreads();
    if (A.rec.row_start && tid < hI)
        ring_pts(A.rec, xI0 + tid, P1, kI0, kI1, n, GU + tid * wI, kI0, true);
    __syncthreads();
    for (int i = tid; i < nI; i += NT) {
        const int64_t x = xI0 + i / wI, k = kI0 + i % wI;
        const bool in_ = inside(x, g, P1) && inside(k, g, P2);
        const float uu = U[i];
        U[i] = in_ ? uu - dt * V[i] : uu;
        const float tv = GV[i] + dt * GU[i];
        H[i] = in_ ? C[i] * tv : 0.0f;
    }
    __syncthreads();
    // step n on W
    for (int i = tid; i < nI; i += NT) {
        const int64_t x = xI0 + i / wI, k = kI0 + i % wI;
        if (x < xW0 || x >= xW1 || k < kW0 || k >= kW1) continue;
        const bool xin = inside(x, g, P1), upin = inside(x - 1, g, P1), dnin = inside(x + 1, g, P1);
        const bool kin = inside(k, g, P2), lin = inside(k - 1, g, P2), rin = inside(k + 1, g, P2);
        const bool in_ = xin && kin;
        const bool hu = x > 0, hd = x + 1 < P1, hlf = k > 0, hrt = k + 1 < P2;
        const float c = U[i];
        float Lu = 0.0f;
        Lu += wm * (hu ? U[i - wI] : 0.0f);
        Lu += wc * c;
        Lu += wp * (hd ? U[i + wI] : 0.0f);
        Lu += wm * (hlf ? U[i - 1] : 0.0f);
        Lu += wc * c;
        Lu += wp * (hrt ? U[i + 1] : 0.0f);
        const float rhs = C[i] * (alpha * Lu);
        VN[i] = in_ ? V[i] - dt * rhs : V[i];
        const float tv = GV[i] + dt * GU[i];
        float at = 0.0f;
        if (kin) {
            if (dnin) at += wm * H[i + wI];
            if (xin) at += wc * H[i];
            if (upin) at += wp * H[i - wI];
        }
        if (xin) {
            if (rin) at += wm * H[i + 1];
            if (kin) at += wc * H[i];
            if (lin) at += wp * H[i - 1];
        }
        GUN[i] = GU[i] + dt * (alpha * at);
        GVN[i] = in_ ? tv : GV[i];
        T1[i] = in_ ? dt * (tv * (alpha * Lu)) : 0.0f;
        UN[i] = c;
    }
    __syncthreads();
    // step n fix-ups: sources out of v, saved boundary scalars back into u,
    // boundary transpose of g_u (staged, then written)
    if (A.src.row_start && tid < (int)(xW1 - xW0))
        ring_pts(A.src, xW0 + tid, P1, kW0, kW1, n, VN + (xW0 + tid - xI0) * wI, kI0, false);
    for (int e = pc.w_beg + tid; e < pc.w_end; e += NT) {
        const pde_ring_entry w = A.ring.w[e];
        const int l = lidx(w.node);
        if (l >= 0) UN[l] = A.save[n * A.nops + w.a];
    }
    for (int e = pc.t_beg + tid; e < pc.t_end; e += NT) {
        const pde_ring_entry w = A.ring.t[e];
        const int l = lidx(w.node);
        float acc = (w.flags & 1) || l < 0 ? 0.0f : GUN[l];
        for (int q = w.a; q < w.b; ++q) {
            const int m = lidx(A.ring.cdst[q]);
            acc += A.ring.csign[q] * (m >= 0 ? GUN[m] : 0.0f);
        }
        if (e - pc.t_beg < RING_CAP) ST[e - pc.t_beg] = acc;
    }
    __syncthreads();
    for (int e = pc.t_beg + tid; e < pc.t_end; e += NT) {
        const int l = lidx(A.ring.t[e].node);
        if (l >= 0 && e - pc.t_beg < RING_CAP) GUN[l] = ST[e - pc.t_beg];
    }
    for (int e = pc.z_beg + tid; e < pc.z_end; e += NT) {
        const int l = lidx(A.ring.z[e].node);
        if (l >= 0) GUN[l] = 0.0f;
    }
    __syncthreads();
    if (A.rec.row_start && tid < (int)(xW1 - xW0))
        ring_pts(A.rec, xW0 + tid, P1, kW0, kW1, n - 1, GUN + (xW0 + tid - xI0) * wI, kI0, true);
    __syncthreads();
    for (int i = tid; i < nI; i += NT) {
        const int64_t x = xI0 + i / wI, k = kI0 + i % wI;
        if (x < xW0 || x >= xW1 || k < kW0 || k >= kW1) continue;
        const bool in_ = inside(x, g, P1) && inside(k, g, P2);
        const float un = UN[i];
        UN[i] = in_ ? un - dt * VN[i] : un;
        const float tv = GVN[i] + dt * GUN[i];
        H[i] = in_ ? C[i] * tv : 0.0f;
    }
    __syncthreads();
    // step n-1 on the piece
    for (int i = tid; i < nI; i += NT) {
        const int64_t x = xI0 + i / wI, k = kI0 + i % wI;
        if (x < pc.r0 || x >= pc.r1 || k < pc.c0 || k >= pc.c1) continue;
        const bool xin = inside(x, g, P1), upin = inside(x - 1, g, P1), dnin = inside(x + 1, g, P1);
        const bool kin = inside(k, g, P2), lin = inside(k - 1, g, P2), rin = inside(k + 1, g, P2);
        const bool in_ = xin && kin;
        const bool hu = x > 0, hd = x + 1 < P1, hlf = k > 0, hrt = k + 1 < P2;
        const float c = UN[i];
        float Lu = 0.0f;
        Lu += wm * (hu ? UN[i - wI] : 0.0f);
        Lu += wc * c;
        Lu += wp * (hd ? UN[i + wI] : 0.0f);
        Lu += wm * (hlf ? UN[i - 1] : 0.0f);
        Lu += wc * c;
        Lu += wp * (hrt ? UN[i + 1] : 0.0f);
        const float rhs = C[i] * (alpha * Lu);
        const float vout = in_ ? VN[i] - dt * rhs : VN[i];
        float at = 0.0f;
        if (kin) {
            if (dnin) at += wm * H[i + wI];
            if (xin) at += wc * H[i];
            if (upin) at += wp * H[i - wI];
        }
        if (xin) {
            if (rin) at += wm * H[i + 1];
            if (kin) at += wc * H[i];
            if (lin) at += wp * H[i - 1];
        }
        const float guout = GUN[i] + dt * (alpha * at);
        const float tv = GVN[i] + dt * GUN[i];
        const int64_t e = x * P2 + k;
        A.gvo[e] = in_ ? tv : GVN[i];
        if (A.gc2 && in_) A.gc2[e] = (A.gc2[e] + T1[i]) + dt * (tv * (alpha * Lu));
        U[i] = c;
        V[i] = vout;
        GU[i] = guout;
    }
    __syncthreads();
    // step n-1 fix-ups on the piece
    if (A.src.row_start && tid < pc.r1 - pc.r0)
        ring_pts(A.src, pc.r0 + tid, P1, pc.c0, pc.c1, n - 1, V + (pc.r0 + tid - xI0) * wI, kI0,
                 false);
    for (int e = pc.w_beg + tid; e < pc.w_end; e += NT) {
        const pde_ring_entry w = A.ring.w[e];
        const int l = lidx(w.node);
        if (l >= 0 && (w.flags & 2)) U[l] = A.save[(n - 1) * A.nops + w.a];
    }
    for (int e = pc.t_beg + tid; e < pc.t_end; e += NT) {
        const pde_ring_entry w = A.ring.t[e];
        if (!(w.flags & 2)) continue;
        const int l = lidx(w.node);
        float acc = (w.flags & 1) || l < 0 ? 0.0f : GU[l];
        for (int q = w.a; q < w.b; ++q) {
            const int m = lidx(A.ring.cdst[q]);
            acc += A.ring.csign[q] * (m >= 0 ? GU[m] : 0.0f);
        }
        if (e - pc.t_beg < RING_CAP) ST[e - pc.t_beg] = acc;
    }
    __syncthreads();
    for (int e = pc.t_beg + tid; e < pc.t_end; e += NT) {
        const pde_ring_entry w = A.ring.t[e];
        const int l = lidx(w.node);
        if ((w.flags & 2) && l >= 0 && e - pc.t_beg < RING_CAP) GU[l] = ST[e - pc.t_beg];
    }
    for (int e = pc.z_beg + tid; e < pc.z_end; e += NT) {
        const pde_ring_entry w = A.ring.z[e];
        const int l = lidx(w.node);
        if ((w.flags & 2) && l >= 0) GU[l] = 0.0f;
    }
    __syncthreads();
    for (int i = tid; i < nI; i += NT) {
        const int64_t x = xI0 + i / wI, k = kI0 + i % wI;
        if (x < pc.r0 || x >= pc.r1 || k < pc.c0 || k >= pc.c1) continue;
        const int64_t e = x * P2 + k;
        A.uo[e] = U[i];
        A.vo[e] = V[i];
        A.guo[e] = GU[i];
    }
}

// 3 workgroups per CU (168 VGPRs, 3 waves per SIMD; a few registers spill):
// the tiles are latency-bound, and the third wave measures faster than the
// spill-free 173-register build at 2 waves per SIMD.
__global__ __launch_bounds__(64 * PR_BY, 3) void wave_pair_kernel(PairArgs A, PairGeom pg)
{
    __shared__ float lds[RING_ARR * RING_CAP];
    const int64_t n = *A.step;
    if (blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) *A.step_next = n - 2;
    // ring pieces after the interior tiles: they fill the slots the tiles'
    // last round leaves free
    if (blockIdx.x >= pg.nint) {
        const int pid = (int)(blockIdx.x - pg.nint);
        if (pid < A.ring.npieces) pair_ring(A, pg, pid, n, lds);
        return;
    }
    pair_interior<PR_ROWS>(A, pg, blockIdx.x, n);
}

}  // namespace pde

using namespace pde;

static bool point_set_valid(const pde_point_set *p)
{
    return !p || (p->row_start && p->col && p->slot && p->data && p->stride >= 0);
}

extern "C" int pde_wave_reverse_adjoint_pair_f32(
    const float *u, const float *v, const float *gu, const float *gv, const float *c2, float *uo,
    float *vo, float *guo, float *gvo, float *gc2, int64_t nx, int64_t ny, int64_t nz,
    int32_t ghost, float wm, float wc, float wp, float alpha, float dt, const int64_t *step,
    int64_t *step_next, const pde_point_set *sources, const pde_point_set *receivers,
    const float *save, int64_t nops, const pde_ring_plan *ring, void *stream)
{
    if (!u || !v || !gu || !gv || !uo || !vo || !guo || !gvo) return PDE_EINVAL;
    const float *ins[5] = {u, v, gu, gv, c2};
    float *outs[5] = {uo, vo, guo, gvo, gc2};
    for (float *o : outs)
        for (const float *i : ins)
            if (o && i && o == i) return PDE_EINVAL;
    for (int a = 0; a < 5; ++a)
        for (int b = a + 1; b < 5; ++b)
            if (outs[a] && outs[a] == outs[b]) return PDE_EINVAL;
    if (!step || !step_next || step == step_next) return PDE_EINVAL;
    if (!point_set_valid(sources) || !point_set_valid(receivers)) return PDE_EINVAL;
    if (nops < 0 || (nops > 0 && !save)) return PDE_EINVAL;
    if (!ring || ring->npieces < 1 || !ring->pieces) return PDE_EINVAL;
    if ((ring->w == nullptr) || (ring->t == nullptr) || (ring->z == nullptr)) return PDE_EINVAL;
    // save rows and plan entries are indexed by op slot: the plan must be the
    // one built for this table (every entry's slot is below ring->nops)
    if (ring->nops != nops) return PDE_EINVAL;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim != 2) return PDE_ESHAPE;
    PairGeom pg;
    pg.P1 = G.P[1];
    pg.P2 = G.P[2];
    pg.g = ghost;
    pg.D = ring->depth;
    if (pg.D < 4 || pg.D > 8 || pg.D < ghost + 2) return PDE_EINVAL;
    // byte offsets within the field fit 32 bits
    if (pg.P1 < 2 * pg.D || pg.P2 < 2 * pg.D || pg.P1 * pg.P2 > 0x3fffffff) return PDE_ESHAPE;
    const int64_t ns = (pg.P2 - 2 * pg.D + PR_COLS - 1) / PR_COLS;
    const int64_t nr = (pg.P1 - 2 * pg.D + PR_BY * PR_ROWS - 1) / (PR_BY * PR_ROWS);
    const int64_t nint = (ns * nr + 7) / 8 * 8;
    const int64_t nring = ((int64_t)ring->npieces + 7) / 8 * 8;
    if (nint + nring > 0x7fffffff) return PDE_ESHAPE;
    pg.nstrips = (uint32_t)ns;
    pg.nrt = (uint32_t)nr;
    pg.nint = (uint32_t)nint;
    pg.nring = (uint32_t)nring;
    PairArgs A{u, v, gu, gv, c2, uo, vo, guo, gvo, gc2, wm, wc, wp, alpha, dt, step, step_next,
               sources ? *sources : pde_point_set{}, receivers ? *receivers : pde_point_set{},
               save, nops, *ring};
    hipLaunchKernelGGL(wave_pair_kernel, dim3((unsigned)(nint + nring)), dim3(64, PR_BY), 0,
                       as_stream(stream), A, pg);
    return launch_status();
}
