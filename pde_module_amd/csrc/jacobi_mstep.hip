// Several Jacobi sweeps of the projection cavity's pressure equation per
// launch, on LDS-resident tiles (pde_jacobi_neumann0_multisweep_f32), and
// the same for the transposed sweeps (pde_jacobi_neumann0_adjoint_multisweep_f32).
//
// The sweep is the folded Neumann-0 one of jacobi_neumann0_kernel
// (flow.hip): on interior nodes p' = inv * (acc - h2 * rhs), acc over the axes
// of the two neighbours, a neighbour in the ghost layer read as (-0 or +0)
// plus its mirror across the wall; ghosts are neither read nor written.  The
// transposed sweep is jacobi_neumann0_adjoint_kernel's: g' = inv * (sum of
// the interior neighbours) plus the wall mirrors' terms on interior nodes, 0
// on ghosts, and rhs's gradient gathering -h2 * (inv * g) of every sweep's
// input, in sweep order.  Same arithmetic in the same order as those
// kernels, so any number of sweeps per launch gives their values bit for bit.
//
// Tiles: each instance's nodes are cut into output rectangles of L1 x L2
// nodes (the last ones shorter; one piece along an axis that fits a
// region), computed from the block index; a tile's region is its output
// grown by `sweeps` nodes on every side, inside the grid.  A workgroup loads
// its region once into LDS and sweeps it there: between two buffers, one
// barrier per sweep, or (the forward of large batches) in one buffer, each
// sweep's rows held in registers between two barriers, which halves the LDS
// and raises the occupancy.  Cells on a side where the region was cut out of its
// grid go stale one layer per sweep, so after the last sweep the output
// rectangle holds exactly what the one-sweep launches reach, and is stored.
//
// Every thread owns 4 columns (one 16-byte LDS chunk) of RPB consecutive
// rows (2 in 64 x 64 regions, 4 in 128 x 128 ones); a band of PITCH/4 lanes
// covers the pitch, the neighbours beside a chunk come from the adjacent
// lanes by whole-wave DPP shifts.  All cells are swept with the interior
// formula, without masks: a region's edge cells are either on a cut side
// (stale anyway) or ghost nodes, which the folded sweep never reads (their
// garbage is never stored: the forward stores interior nodes, the transpose
// stores 0 on ghosts).  Only interior nodes within two cells of a wall need
// the wall terms; a wave holding any such node takes the general loop
// (selects on the wall flags), every other wave the plain one.  Row -1 and
// the row past the last of each buffer are padding that the first and last
// bands read (their garbage reaches region-edge cells only).
#include "mstep.hpp"

namespace pde {

template <int PITCH, int NT, int RPB>
struct JmCfg {
    static constexpr int LANES = PITCH / 4;           // threads per band (16 or 32)
    static constexpr int BANDS = NT / LANES;
    static constexpr int ROWS = BANDS * RPB;          // region height cap
    static constexpr int BUF = (ROWS + 2) * PITCH;    // one buffer, padding rows included
    static_assert(LANES == 16 || LANES == 32, "bands of 16 or 32 lanes (within a wave)");
    static_assert(ROWS == PITCH, "square regions");
};

// 16-byte chunks (the buffers start 16-byte aligned; PITCH is a multiple of 4)
typedef float jm_v4 __attribute__((ext_vector_type(4)));

template <int PITCH>
__device__ __forceinline__ ms_f4 jm_rd(const float *B, int lr, int c4)
{
    const jm_v4 v = *static_cast<const jm_v4 *>(
        __builtin_assume_aligned(B + (lr + 1) * PITCH + 4 * c4, 16));
    return ms_f4{v.x, v.y, v.z, v.w};
}

template <int PITCH>
__device__ __forceinline__ void jm_wr(float *B, int lr, int c4, const ms_f4 &v)
{
    float *p = static_cast<float *>(__builtin_assume_aligned(B + (lr + 1) * PITCH + 4 * c4, 16));
    *reinterpret_cast<ms_f4 *>(p) = v;
}

// Wall flags of one node (only read where EDGE): interior node next to the
// low / high wall along rows (x) and columns (k), and (transpose) the
// neighbour of such a node, whose wall mirror adds a term.
struct JmWall {
    bool xlo, xhi, klo, khi;        // the -1 / +1 neighbour is a ghost
    bool mxlo, mxhi, mklo, mkhi;    // the -1 / +1 neighbour is next to its wall
};

// Forward folded sweep of one node: u/d the rows above/below, zl/zr the
// columns left/right, hr = h2 * rhs.
template <bool EDGE>
__device__ __forceinline__ float jm_fwd(float u, float d, float zl, float zr, float hr,
                                        float inv, const JmWall &w)
{
    float a_lo = u, a_hi = d, b_lo = zl, b_hi = zr;
    if (EDGE) {
        a_lo = w.xlo ? -0.0f + d : u;
        a_hi = w.xhi ? 0.0f + u : d;
        b_lo = w.klo ? -0.0f + zr : zl;
        b_hi = w.khi ? 0.0f + zl : zr;
    }
    float acc = 0.0f;
    acc += a_lo;
    acc += a_hi;
    acc += b_lo;
    acc += b_hi;
    return inv * (acc - hr);
}

// Transposed sweep of one node (acc never holds -0, so adding +0 in place of
// skipping a ghost neighbour leaves it unchanged; the mirror terms select).
template <bool EDGE>
__device__ __forceinline__ float jm_adj(float u, float d, float zl, float zr, float inv,
                                        const JmWall &w)
{
    float acc = 0.0f;
    if (EDGE) {
        acc += w.xhi ? 0.0f : d;
        acc += w.xlo ? 0.0f : u;
        acc += w.khi ? 0.0f : zr;
        acc += w.klo ? 0.0f : zl;
    } else {
        acc += d;
        acc += u;
        acc += zr;
        acc += zl;
    }
    float v = inv * acc;
    if (EDGE) {
        v = w.mxlo ? v + inv * u : v;
        v = w.mxhi ? v + inv * d : v;
        v = w.mklo ? v + inv * zl : v;
        v = w.mkhi ? v + inv * zr : v;
    }
    return v;
}

struct JmGeom {
    int lo1, hi1, lo2, hi2;   // interior rows [lo1, hi1), columns [lo2, hi2) of the grid
    int r0, c0;               // region origin in the grid
};

__device__ __forceinline__ JmWall jm_wall(const JmGeom &g, int x, int k)
{
    JmWall w;
    w.xlo = x - 1 < g.lo1;
    w.xhi = x + 1 >= g.hi1;
    w.klo = k - 1 < g.lo2;
    w.khi = k + 1 >= g.hi2;
    w.mxlo = x - 1 >= g.lo1 && x - 2 < g.lo1;
    w.mxhi = x + 1 < g.hi1 && x + 2 >= g.hi1;
    w.mklo = k - 1 >= g.lo2 && k - 2 < g.lo2;
    w.mkhi = k + 1 < g.hi2 && k + 2 >= g.hi2;
    return w;
}

// `sweeps` sweeps from buffer A; returns the buffer holding the result.  Two
// buffers (ONE false): each sweep writes B, one barrier, the buffers swap.
// One buffer (ONE true, B unused): each sweep keeps its rows in registers,
// a barrier, writes them back over A, a second barrier.
template <bool ADJ, bool EDGE, int PITCH, int RPB, bool ONE>
__device__ __forceinline__ float *jm_sweeps(float *A, float *B, int sweeps, int lr0, int c4,
                                            float (&r)[RPB][4], float h2, float inv,
                                            const JmGeom &g)
{
    const int k0 = g.c0 + 4 * c4;
    for (int s = 0; s < sweeps; ++s) {
        ms_f4 nv[ONE ? RPB : 1];
        ms_f4 up = jm_rd<PITCH>(A, lr0 - 1, c4), mid = jm_rd<PITCH>(A, lr0, c4);
#pragma unroll
        for (int h = 0; h < RPB; ++h) {
            const ms_f4 dn = jm_rd<PITCH>(A, lr0 + h + 1, c4);
            const float left = ms_from_left(mid.w), right = ms_from_right(mid.x);
            const int x = g.r0 + lr0 + h;
            const float zl[4] = {left, mid.x, mid.y, mid.z};
            const float zr[4] = {mid.y, mid.z, mid.w, right};
            const float uu[4] = {up.x, up.y, up.z, up.w};
            const float dd[4] = {dn.x, dn.y, dn.z, dn.w};
            const float cc[4] = {mid.x, mid.y, mid.z, mid.w};
            float o[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                JmWall w{};
                if (EDGE) w = jm_wall(g, x, k0 + i);
                if (!ADJ) {
                    o[i] = jm_fwd<EDGE>(uu[i], dd[i], zl[i], zr[i], r[h][i], inv, w);
                } else {
                    o[i] = jm_adj<EDGE>(uu[i], dd[i], zl[i], zr[i], inv, w);
                    r[h][i] = r[h][i] + -h2 * (inv * cc[i]);
                }
            }
            if constexpr (ONE)
                nv[h] = ms_f4{o[0], o[1], o[2], o[3]};
            else
                jm_wr<PITCH>(B, lr0 + h, c4, ms_f4{o[0], o[1], o[2], o[3]});
            up = mid;
            mid = dn;
        }
        __syncthreads();
        if constexpr (ONE) {
#pragma unroll
            for (int h = 0; h < RPB; ++h) jm_wr<PITCH>(A, lr0 + h, c4, nv[h]);
            __syncthreads();
        } else {
            float *t = A;
            A = B;
            B = t;
        }
    }
    return A;
}

// The launch's tiling of each instance: outputs of L1 x L2 nodes (the last
// ones shorter), m2 of them along the columns; regions grown by `halo` on
// every side inside the grid.  Instance b at offset b * stride.
struct JmPlan {
    int P1, P2, ghost;
    int L1, L2, m2, halo;
    int64_t stride;
};

struct JmTile {
    int r0, c0, rh, rw;               // region origin and extents
    int o_r0, o_r1, o_c0, o_c1;       // output rectangle (grid indices)
};

__device__ __forceinline__ JmTile jm_tile(const JmPlan &P, int t)
{
    JmTile T;
    const int j1 = t / P.m2, j2 = t - j1 * P.m2;
    T.o_r0 = j1 * P.L1;
    T.o_r1 = min(P.P1, T.o_r0 + P.L1);
    T.o_c0 = j2 * P.L2;
    T.o_c1 = min(P.P2, T.o_c0 + P.L2);
    T.r0 = max(0, T.o_r0 - P.halo);
    T.c0 = max(0, T.o_c0 - P.halo);
    T.rh = min(P.P1, T.o_r1 + P.halo) - T.r0;
    T.rw = min(P.P2, T.o_c1 + P.halo) - T.c0;
    return T;
}

// MINW: the waves per SIMD the register allocation must allow.
template <bool ADJ, int PITCH, int NT, int RPB, bool ONE, int MINW>
__global__ __launch_bounds__(NT, MINW) void jacobi_mstep_kernel(
    const float *__restrict__ in, const float *__restrict__ in2, float *__restrict__ out,
    const float *__restrict__ rhs, float *__restrict__ grhs, JmPlan PL, float h2, float inv,
    int sweeps, int accumulate)
{
    using C = JmCfg<PITCH, NT, RPB>;
    __shared__ __align__(16) float S[(ONE ? 1 : 2) * C::BUF];
    const JmTile T = jm_tile(PL, blockIdx.x);
    const int rh = T.rh, rw = T.rw;
    // the host sizes the plan for the LDS arrays; a tile outside them is skipped
    if (!(rh >= 1 && rh <= C::ROWS && rw >= 1 && rw <= PITCH && T.o_r0 < T.o_r1 &&
          T.o_c0 < T.o_c1))
        return;
    const int tid = threadIdx.x;
    const int64_t P2 = PL.P2;
    const int64_t off = (int64_t)blockIdx.y * PL.stride;
    JmGeom g;
    g.lo1 = PL.ghost;
    g.hi1 = PL.P1 - PL.ghost;
    g.lo2 = PL.ghost;
    g.hi2 = PL.P2 - PL.ghost;
    g.r0 = T.r0;
    g.c0 = T.c0;
    float *A = S, *B = ONE ? S : S + C::BUF;

    const int band = tid / C::LANES, c4 = tid - band * C::LANES, lr0 = band * RPB;
    const int lc0 = 4 * c4;
    auto interior = [&](int x, int k) {
        return x >= g.lo1 && x < g.hi1 && k >= g.lo2 && k < g.hi2;
    };
    // Stage the region row by row (consecutive threads, consecutive
    // columns): the field into buffer A; with two buffers also h2 * rhs
    // (forward) or the running rhs gradient (transpose; 0 when not
    // accumulating) of interior nodes into buffer B, 0 elsewhere, from where
    // each thread takes its own chunk; with one buffer each thread loads its
    // chunk of those straight into registers.
    // Each load reads a valid address (out-of-region ones the grid's first
    // node, then discarded).
    constexpr int PER = C::ROWS * PITCH / NT;
    const float *ig = in + off;
    const float *ig2 = in2 ? in2 + off : ig;      // the field is in + in2 when in2 is given
    const float *rg = ADJ ? grhs + off : rhs + off;
    float r[RPB][4];
    bool edge = false;
    if constexpr (!ONE) {
        // every load issued before the first is used
        float f[PER], f2[PER], q[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int e = tid + NT * j, lr = e / PITCH, lc = e % PITCH;
            const int x = T.r0 + lr, k = T.c0 + lc;
            const bool in_region = lr < rh && lc < rw;
            const int64_t a = in_region ? (int64_t)x * P2 + k : 0;
            f[j] = ig[a];
            f2[j] = in2 ? ig2[a] : 0.0f;
            q[j] = ADJ && !accumulate ? 0.0f : rg[in_region && interior(x, k) ? a : 0];
        }
        if (in2) {
#pragma unroll
            for (int j = 0; j < PER; ++j) f[j] = f[j] + f2[j];   // pde_lincomb_f32 (1, 1)
        }
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int e = tid + NT * j, lr = e / PITCH, lc = e % PITCH;
            const int x = T.r0 + lr, k = T.c0 + lc;
            const bool in_region = lr < rh && lc < rw;
            const bool in_int = in_region && interior(x, k);
            A[(lr + 1) * PITCH + lc] = in_region ? f[j] : 0.0f;
            B[(lr + 1) * PITCH + lc] = in_int ? (ADJ ? q[j] : h2 * q[j]) : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < RPB; ++h) {
            const ms_f4 v = jm_rd<PITCH>(B, lr0 + h, c4);
            r[h][0] = v.x;
            r[h][1] = v.y;
            r[h][2] = v.z;
            r[h][3] = v.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int e = tid + NT * j, lr = e / PITCH, lc = e % PITCH;
            const bool in_region = lr < rh && lc < rw;
            const int64_t a = in_region ? (int64_t)(T.r0 + lr) * P2 + T.c0 + lc : 0;
            float f = ig[a];
            if (in2) f = f + ig2[a];                  // pde_lincomb_f32 (1, 1)
            A[(lr + 1) * PITCH + lc] = in_region ? f : 0.0f;
        }
#pragma unroll
        for (int h = 0; h < RPB; ++h) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int lr = lr0 + h, lc = lc0 + i, x = T.r0 + lr, k = T.c0 + lc;
                const bool in_int = lr < rh && lc < rw && interior(x, k);
                const float q =
                    ADJ && !accumulate ? 0.0f : rg[in_int ? (int64_t)x * P2 + k : 0];
                r[h][i] = in_int ? (ADJ ? q : h2 * q) : 0.0f;
            }
        }
    }
    // whether any of the thread's interior nodes lies within two of a wall
#pragma unroll
    for (int h = 0; h < RPB; ++h) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int lr = lr0 + h, lc = lc0 + i, x = T.r0 + lr, k = T.c0 + lc;
            edge = edge || (lr < rh && lc < rw && interior(x, k) &&
                            (x - 2 < g.lo1 || x + 2 >= g.hi1 || k - 2 < g.lo2 || k + 2 >= g.hi2));
        }
    }
    __syncthreads();   // buffer B is the first sweep's destination; A is complete
    const bool wave_edge = __ballot(edge) != 0;   // uniform over the wave
    float *R = wave_edge
                   ? jm_sweeps<ADJ, true, PITCH, RPB, ONE>(A, B, sweeps, lr0, c4, r, h2, inv, g)
                   : jm_sweeps<ADJ, false, PITCH, RPB, ONE>(A, B, sweeps, lr0, c4, r, h2, inv, g);
    float *W = R == A ? B : A;   // two buffers: the last sweep's source, free after its barrier

    // The output rectangle, row by row: the forward stores interior nodes
    // only (the folded sweep never writes ghosts); the transpose every node
    // (0 on ghosts) and the rhs gradient of its interior nodes: with two
    // buffers staged through the free one and stored row by row, with one
    // each thread its own chunk from registers.
    if (ADJ && !ONE) {
#pragma unroll
        for (int h = 0; h < RPB; ++h)
            jm_wr<PITCH>(W, lr0 + h, c4, ms_f4{r[h][0], r[h][1], r[h][2], r[h][3]});
        __syncthreads();
    }
    float *og = out + off;
    float *gg = ADJ ? grhs + off : nullptr;
    const int orow0 = T.o_r0 - T.r0, orow1 = T.o_r1 - T.r0;
    const int ocol0 = T.o_c0 - T.c0, ocol1 = T.o_c1 - T.c0;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int e = tid + NT * j, lr = e / PITCH, lc = e % PITCH;
        if (lr < orow0 || lr >= orow1 || lc < ocol0 || lc >= ocol1) continue;
        const int x = T.r0 + lr, k = T.c0 + lc;
        const int64_t a = (int64_t)x * P2 + k;
        const bool in_ = interior(x, k);
        if (in_) og[a] = R[(lr + 1) * PITCH + lc];
        if (ADJ) {
            if (!in_) og[a] = 0.0f;
            else if (!ONE) gg[a] = W[(lr + 1) * PITCH + lc];
        }
    }
    if (ADJ && ONE) {
#pragma unroll
        for (int h = 0; h < RPB; ++h) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int lr = lr0 + h, lc = lc0 + i;
                if (lr < orow0 || lr >= orow1 || lc < ocol0 || lc >= ocol1) continue;
                const int x = T.r0 + lr, k = T.c0 + lc;
                if (interior(x, k)) gg[(int64_t)x * P2 + k] = r[h][i];
            }
        }
    }
}

// The tiling of one axis of P nodes for regions of at most `cap` nodes and
// `halo` stale layers: one piece when it fits, else the fewest equal pieces
// whose regions (piece + halo on both sides) fit.
static void jm_axis(int P, int cap, int halo, int *L, int *m)
{
    if (P <= cap) {
        *L = P;
        *m = 1;
        return;
    }
    const int Lmax = cap - 2 * halo;
    const int pieces = (P + Lmax - 1) / Lmax;
    *L = (P + pieces - 1) / pieces;
    *m = (P + *L - 1) / *L;
}

// 64 x 64 regions: 512 threads of 2 rows each (more waves per tile than 4
// rows per thread, less LDS traffic per node than 1).
constexpr int JM_NT64 = 512;

template <bool ADJ>
static int jm_launch(const float *in, const float *in2, float *out, const float *rhs,
                     float *grhs, int64_t batch, int64_t nx, int64_t ny, int64_t nz,
                     int32_t ghost, float h2, float inv, int32_t sweeps, int32_t accumulate,
                     hipStream_t stream)
{
    if (sweeps < 1 || sweeps > PDE_JACOBI_MAX_SWEEPS || batch < 1) return PDE_EINVAL;
    if (batch > 65535) return PDE_ESHAPE;
    Geom G;
    int st = make_geom(nx, ny, nz, ghost, &G);
    if (st) return st;
    if (G.ndim != 2 || G.E[1] < 2 || G.E[2] < 2) return PDE_ESHAPE;
    if (G.P[1] > 0x3fffffff || G.P[2] > 0x3fffffff) return PDE_ESHAPE;
    JmPlan PL;
    PL.P1 = (int)G.P[1];
    PL.P2 = (int)G.P[2];
    PL.ghost = ghost;
    PL.halo = sweeps;
    PL.stride = G.P[1] * G.P[2];
    // 64 x 64 regions (512 threads, two LDS buffers, up to four workgroups
    // per CU) while the halo leaves them room and the grids are small enough
    // to want many workgroups; 128 x 128 (1024 threads) otherwise -- except
    // the forward of batches over 2^22 nodes, which keeps 64 x 64 regions in
    // one LDS buffer (17 KB, registers for 6 waves per SIMD).  Measured at
    // 1026^2, 10 sweeps per launch, graph-replayed cavity steps: 16 seeds
    // batched, forward step 0.895 against 0.980 ms (128 x 128); the one-buffer
    // form for one grid 0.119 against 0.110 ms per forward step.
    const bool fits64 = (PL.P1 <= 64 && PL.P2 <= 64) || 2 * sweeps + 16 <= 64;
    const bool big = batch * PL.stride > (int64_t)1 << 22;
    const bool one = !ADJ && fits64 && big;
    const bool small = (PL.P1 <= 64 && PL.P2 <= 64) || (fits64 && !big) || one;
    const int cap = small ? 64 : 128;
    int m1;
    jm_axis(PL.P1, cap, sweeps, &PL.L1, &m1);
    jm_axis(PL.P2, cap, sweeps, &PL.L2, &PL.m2);
    const int64_t tiles = (int64_t)m1 * PL.m2;
    if (tiles > 0x7fffffff) return PDE_ESHAPE;
    const dim3 grid((unsigned)tiles, (unsigned)batch);
    if (one) {
        if constexpr (!ADJ)
            hipLaunchKernelGGL((jacobi_mstep_kernel<false, 64, JM_NT64, 2, true, 6>), grid,
                               dim3(JM_NT64), 0, stream, in, in2, out, rhs, grhs, PL, h2, inv,
                               sweeps, accumulate);
    } else if (small)
        hipLaunchKernelGGL((jacobi_mstep_kernel<ADJ, 64, JM_NT64, 2, false, 1>), grid,
                           dim3(JM_NT64), 0, stream, in, in2, out, rhs, grhs, PL, h2, inv, sweeps,
                           accumulate);
    else
        hipLaunchKernelGGL((jacobi_mstep_kernel<ADJ, 128, 1024, 4, false, 1>), grid, dim3(1024),
                           0, stream, in, in2, out, rhs, grhs, PL, h2, inv, sweeps, accumulate);
    return launch_status();
}

}  // namespace pde

using namespace pde;

extern "C" int pde_jacobi_neumann0_multisweep_f32(const float *p, const float *rhs, float *out,
                                                  int64_t batch, int64_t nx, int64_t ny,
                                                  int64_t nz, int32_t ghost, float h2, float inv,
                                                  int32_t sweeps, void *stream)
{
    if (!p || !rhs || !out || p == out || rhs == out) return PDE_EINVAL;
    return jm_launch<false>(p, nullptr, out, rhs, nullptr, batch, nx, ny, nz, ghost, h2, inv,
                            sweeps, 0, as_stream(stream));
}

extern "C" int pde_jacobi_neumann0_adjoint_multisweep_f32(const float *g, const float *g2,
                                                          float *gp, float *grhs, int64_t batch,
                                                          int64_t nx, int64_t ny, int64_t nz,
                                                          int32_t ghost, float h2, float inv,
                                                          int32_t sweeps, int32_t accumulate,
                                                          void *stream)
{
    if (!g || !gp || !grhs || g == gp || g == grhs || gp == grhs) return PDE_EINVAL;
    if (g2 && (g2 == gp || g2 == grhs)) return PDE_EINVAL;
    if (accumulate != 0 && accumulate != 1) return PDE_EINVAL;
    return jm_launch<true>(g, g2, gp, nullptr, grhs, batch, nx, ny, nz, ghost, h2, inv, sweeps,
                           accumulate, as_stream(stream));
}
