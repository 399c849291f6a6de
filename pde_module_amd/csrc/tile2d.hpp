// Row-tile helpers for 2D scalar fields (padded extents P1 rows x P2 columns,
// column index fastest).  A lane owns 4 consecutive columns of a row, loaded
// as one 16-byte access (unaligned dwordx4 is legal on gfx950); lanes are
// grouped in row segments of `cw` lanes (64, 32 or 16) that cover 4 * cw
// consecutive columns, and column neighbours cross lanes by shuffles within a
// segment (the segment's edge lanes load their outer neighbour).
#pragma once
#include "pde_common.hpp"

namespace pde {

typedef float v4 __attribute__((ext_vector_type(4)));
typedef float v4u __attribute__((ext_vector_type(4), aligned(4)));

struct Tile2 {
    int64_t P1, P2;       // padded rows, columns
    int64_t g;            // ghost width
    int64_t nq;           // 4-column chunks per row (ceil(P2 / 4))
    uint32_t tq, tr;      // tiles along columns and rows (single-grid launches)
};

__device__ __forceinline__ bool inside(int64_t v, int64_t g, int64_t P) { return v >= g && v < P - g; }

// Loads the 4 columns k0 .. k0+3 of row x (clamped to the array: columns past
// P2 read as 0; rows outside [0, P1) read as 0).
__device__ __forceinline__ v4 ld_row4(const float *__restrict__ f, const Tile2 &t, int64_t x,
                                      int64_t k0)
{
    v4 r = {0.0f, 0.0f, 0.0f, 0.0f};
    if (x < 0 || x >= t.P1) return r;
    const float *p = f + x * t.P2 + k0;
    if (k0 + 4 <= t.P2) {
        const v4u q = *reinterpret_cast<const v4u *>(p);
        return v4{q.x, q.y, q.z, q.w};
    }
    const int64_t n = t.P2 - k0;
    r.x = p[0];
    if (n > 1) r.y = p[1];
    if (n > 2) r.z = p[2];
    return r;
}

__device__ __forceinline__ void st_row4(float *__restrict__ f, const Tile2 &t, int64_t x,
                                        int64_t k0, v4 v)
{
    float *p = f + x * t.P2 + k0;
    if (k0 + 4 <= t.P2) {
        *reinterpret_cast<v4u *>(p) = v4u{v.x, v.y, v.z, v.w};
        return;
    }
    const int64_t n = t.P2 - k0;
    p[0] = v.x;
    if (n > 1) p[1] = v.y;
    if (n > 2) p[2] = v.z;
}

__device__ __forceinline__ float el(const v4 &v, int i)
{
    return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}

__device__ __forceinline__ void set_el(v4 &v, int i, float s)
{
    if (i == 0) v.x = s;
    else if (i == 1) v.y = s;
    else if (i == 2) v.z = s;
    else v.w = s;
}

// Point lists are walked by the whole wave: the row index is the same on
// every lane, the wave loads up to 64 of the row's points at once (one per
// lane), then each lane walks them in list order through v_readlane and
// applies those in its own columns.  A row's cost is one round of loads
// however many points it holds, and sums at a node keep the list order.

// [a, b): the list range of padded row x (empty outside the field).
__device__ __forceinline__ void pts_range(const pde_point_set &P, int64_t x, int64_t P1, int &a,
                                          int &b)
{
    a = b = 0;
    if (x < 0 || x >= P1) return;
    a = __builtin_amdgcn_readfirstlane(P.row_start[x]);
    b = __builtin_amdgcn_readfirstlane(P.row_start[x + 1]);
}

// Adds scale * data[n][slot] of the points [a, b) to the elements of v in
// columns k0 .. k0+3, and to ex when the point's column is kex (kex < 0:
// none).  Every lane of the wave calls it with the same range.
__device__ __forceinline__ void pts_add(const pde_point_set &P, int a, int b, int64_t k0,
                                        int64_t n, int lane, v4 &v, int64_t kex, float &ex)
{
    for (int base = a; base < b; base += 64) {
        const int p = base + lane;
        int col = -1;
        float val = 0.0f;
        if (p < b) {
            col = P.col[p];
            val = P.scale * P.data[n * P.stride + P.slot[p]];
        }
        const int cnt = b - base < 64 ? b - base : 64;
        for (int i = 0; i < cnt; ++i) {
            const int64_t c = __builtin_amdgcn_readlane(col, i);
            const float d = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(val), i));
            const int64_t dd = c - k0;
            if (dd >= 0 && dd < 4) set_el(v, (int)dd, el(v, (int)dd) + d);
            if (c == kex) ex = ex + d;
        }
    }
}

// data[n][slot] = v at the point's column for the points [a, b) whose column
// lies in the wave's strip [kw0, kw0 + 256) (lane l holds columns
// kw0 + 4l .. kw0 + 4l + 3).  Every lane of the wave calls it.
__device__ __forceinline__ void pts_store(const pde_point_set &P, int a, int b, int64_t kw0,
                                          int64_t n, int lane, const v4 &v)
{
    for (int base = a; base < b; base += 64) {
        const int p = base + lane;
        const int64_t d = p < b ? (int64_t)P.col[p] - kw0 : -1;
        const bool mine = d >= 0 && d < 256;
        const int from = mine ? (int)(d >> 2) : lane;
        const float e0 = __shfl(v.x, from, 64), e1 = __shfl(v.y, from, 64);
        const float e2 = __shfl(v.z, from, 64), e3 = __shfl(v.w, from, 64);
        if (mine) {
            const int e = (int)(d & 3);
            P.data[n * P.stride + P.slot[p]] = e == 0 ? e0 : (e == 1 ? e1 : (e == 2 ? e2 : e3));
        }
    }
}

// Column neighbours of the 4 values of a lane: left of element 0 and right of
// element 3, from the neighbouring lanes of its row segment (sl = lane index
// in the segment of cw lanes), or memory at the segment's edges.
__device__ __forceinline__ void col_edges(const float *__restrict__ f, const Tile2 &t, int64_t x,
                                          int64_t k0, const v4 &c, int sl, int cw, float &left,
                                          float &right)
{
    left = __shfl_up(c.w, 1, cw);
    right = __shfl_down(c.x, 1, cw);
    // both segment edges load in one instruction (the other lanes idle)
    const bool lo = sl == 0, hi = sl == cw - 1;
    if (lo || hi) {
        const int64_t k = lo ? k0 - 1 : k0 + 4;
        const float v = (k >= 0 && k < t.P2 && x >= 0 && x < t.P1) ? f[x * t.P2 + k] : 0.0f;
        if (lo) left = v;
        if (hi) right = v;
    }
}

// Laplacian sum at element i of a lane: rows up/mid/down, column edges.
__device__ __forceinline__ float acc5(const v4 &up, const v4 &mid, const v4 &dn, float left,
                                      float right, int i, float wm, float wc, float wp)
{
    const float c = el(mid, i);
    const float zl = i == 0 ? left : el(mid, i - 1);
    const float zr = i == 3 ? right : el(mid, i + 1);
    float acc = 0.0f;
    acc += wm * el(up, i);
    acc += wc * c;
    acc += wp * el(dn, i);
    acc += wm * zl;
    acc += wc * c;
    acc += wp * zr;
    return acc;
}

// Transposed sum at element i: the same neighbours in mirrored order
// (s = -1 reads +1, s = +1 reads -1), each counted only when the neighbour
// node is interior; h is 0 outside the interior, so that is implied except
// for the centre term, which needs node (x, k) itself interior.
__device__ __forceinline__ float acct5(const v4 &up, const v4 &mid, const v4 &dn, float left,
                                       float right, int i, bool x_in, bool up_in, bool dn_in,
                                       bool k_in, bool left_in, bool right_in, float wm,
                                       float wc, float wp)
{
    float acc = 0.0f;
    if (k_in) {
        if (dn_in) acc += wm * el(dn, i);
        if (x_in) acc += wc * el(mid, i);
        if (up_in) acc += wp * el(up, i);
    }
    if (x_in) {
        const float zr = i == 3 ? right : el(mid, i + 1);
        const float zl = i == 0 ? left : el(mid, i - 1);
        if (right_in) acc += wm * zr;
        if (k_in) acc += wc * el(mid, i);
        if (left_in) acc += wp * zl;
    }
    return acc;
}

// One lane's share of the fused heat step u_next = u2 + dt * (alpha *
// acc(u2)) on interior nodes, u_next = u2 elsewhere: rows x0 .. x0+RXT-1 of
// its 4 columns k0 .. k0+3.  Every lane of the wave calls it (shuffles).
template <int RXT>
__device__ __forceinline__ void heat_rows(const float *__restrict__ u, float *__restrict__ un,
                                          const Tile2 &t, int64_t x0, int64_t k0, bool lane_ok,
                                          int sl, int cw, float wm, float wc, float wp,
                                          float alpha, float dt)
{
    bool kin[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) kin[i] = inside(k0 + i, t.g, t.P2);
    v4 rows[RXT + 2];
#pragma unroll
    for (int r = 0; r < RXT + 2; ++r) rows[r] = ld_row4(u, t, x0 - 1 + r, k0);
    // every row's column edges before any store, so their loads issue together
    float left[RXT], right[RXT];
#pragma unroll
    for (int r = 0; r < RXT; ++r)
        col_edges(u, t, x0 + r, k0, rows[r + 1], sl, cw, left[r], right[r]);
#pragma unroll
    for (int r = 0; r < RXT; ++r) {
        const int64_t x = x0 + r;
        if (!lane_ok || x >= t.P1) continue;
        const bool xin = inside(x, t.g, t.P1);
        v4 o = rows[r + 1];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (xin && kin[i]) {
                const float acc = acc5(rows[r], rows[r + 1], rows[r + 2], left[r], right[r], i,
                                       wm, wc, wp);
                set_el(o, i, el(rows[r + 1], i) + dt * (alpha * acc));
            }
        st_row4(un, t, x, k0, o);
    }
}

}  // namespace pde
