/*
 * pde_hip.h — C ABI of libpdehip.so, the MI355X (gfx950) kernels behind
 * pde_module_amd's uniform-grid finite-difference modules.
 *
 * Every entry point:
 *   - works on caller-owned device pointers (FP32, C-order, AoS);
 *   - launches asynchronously on `stream` (a hipStream_t passed as void*;
 *     NULL = the legacy default stream), so a caller may capture it in a graph;
 *   - never allocates, frees or synchronises;
 *   - returns PDE_OK (0), a negative PDE_E* argument error, or a positive
 *     hipError_t from the launch.
 *
 * Field layout (reference pm/mesh/mesh_functions.py:108-159): padded extents
 * (nx, ny, nz), size-1 axes are not grid axes; element n = (i*ny + j)*nz + k
 * holds a vector of ncomp floats, or an N x D row-major matrix.  Eligible axes
 * are those of size > 1, ascending (pm/FDM/utils.py:6-27).  Operators write the
 * interior nodes only (index in [g, P-g) on eligible axes) and leave output
 * ghosts untouched (pm/stencil/stencil.py:38-56: outputs are wp.empty).
 *
 * Arithmetic keeps the reference's operation order with one rounding per + and
 * per * (the library is built with -ffp-contract=off), so results are
 * reproducible bit for bit by a CPU restatement in the same order.
 */
#ifndef PDE_HIP_H
#define PDE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PDE_OK 0
#define PDE_EINVAL (-1)   /* null pointer, bad count or unsupported layout */
#define PDE_ESHAPE (-2)   /* grid too small for the ghost width / stencil  */
#define PDE_ETOOBIG (-3)  /* larger than a fixed-size scratch array         */

/* ---------------------------------------------------------------- operators */

/* Laplacian of a vec(ncomp) field: out[n] = alpha * sum over eligible axes
 * (ascending) of (w_minus*in[n-e] + w_center*in[n] + w_plus*in[n+e]),
 * accumulated from 0 term by term.
 * Replaces pm/FDM/functional/_laplacian.py:5-36 (wp.launch of the kernel built
 * by pm/FDM/kernel/laplacian.py:7-61), called from Laplacian.forward
 * (pm/FDM/module/laplacian.py:54-73). */
int pde_laplacian_f32(const float *in, float *out, int64_t nx, int64_t ny, int64_t nz,
                      int32_t ncomp, int32_t ghost, float w_minus, float w_center,
                      float w_plus, float alpha, void *stream);

/* The same Laplacian through the straightforward one-thread-per-element
 * kernel only (pde_laplacian_f32 switches 3D scalar fields with at least 4
 * interior rows to an x-marching tile kernel).  Same results bit for bit; kept so the two can be compared and
 * timed side by side.  Replaces the same reference call as pde_laplacian_f32. */
int pde_laplacian_generic_f32(const float *in, float *out, int64_t nx, int64_t ny,
                              int64_t nz, int32_t ncomp, int32_t ghost, float w_minus,
                              float w_center, float w_plus, float alpha, void *stream);

/* Periodic Laplacian: the values pde_laplacian_f32 gives after
 * pde_periodic_fill_f32, computed from interior values only (a neighbour index
 * that leaves the interior wraps to the opposite interior index); no ghost is
 * read or written.  No reference counterpart (the reference has no periodic
 * boundary); it serves periodic grids such as the 3D operator benchmark,
 * whose per-step ghost fill it makes unnecessary. */
int pde_laplacian_periodic_f32(const float *in, float *out, int64_t nx, int64_t ny,
                               int64_t nz, int32_t ncomp, int32_t ghost, float w_minus,
                               float w_center, float w_plus, float alpha, void *stream);

/* Central gradient of a vec(ncomp) field into an ncomp x D matrix per node
 * (D = number of eligible axes; for ncomp == 1 this is also the vec(D)
 * layout): out[n][r][t] = alpha * (w_minus*in[n-e_t][r] + w_center*in[n][r] +
 * w_plus*in[n+e_t][r]).
 * Replaces pm/FDM/functional/_grad.py:5-36 with the kernel of
 * pm/FDM/kernel/grad.py:8-102, called from Grad.forward (module/grad.py:68-88). */
int pde_grad_f32(const float *in, float *out, int64_t nx, int64_t ny, int64_t nz,
                 int32_t ncomp, int32_t ghost, float w_minus, float w_center,
                 float w_plus, float alpha, void *stream);

/* Divergence of an nrows x D matrix field (nrows == 1: a vec(D) field) into a
 * vec(nrows) field: out[n][r] = alpha * sum over axes t, then s in (-1,0,1),
 * of w_s * in[n + s e_t][r][t].
 * Replaces pm/FDM/functional/_divergence.py:4-24 with the kernel of
 * pm/FDM/kernel/divergence.py:7-108.  The reference module launches it twice
 * per call with the same result (module/divergence.py:98-111); one launch here. */
int pde_divergence_f32(const float *in, float *out, int64_t nx, int64_t ny, int64_t nz,
                       int32_t nrows, int32_t ghost, float w_minus, float w_center,
                       float w_plus, float alpha, void *stream);

/* pde_divergence_f32 for `batch` (<= 65535) instances of a 1D or 2D grid
 * stored one after another, one launch: instance b's values are those of
 * pde_divergence_f32 on its own arrays.  PDE_ESHAPE for 3D grids.  No
 * reference counterpart (the projection cavity's seeds stepped together). */
int pde_divergence_batch_f32(const float *in, float *out, int64_t batch, int64_t nx, int64_t ny,
                             int64_t nz, int32_t nrows, int32_t ghost, float w_minus,
                             float w_center, float w_plus, float alpha, void *stream);

/* Fused 3D operator pass over one scalar field s and one vec(3) field v on the
 * same padded grid: lap = alpha_lap * Laplacian(s), grad = alpha_grad * Grad(s)
 * (vec3 per node), div = alpha_div * Divergence(v), each bit-identical to the
 * separate entry points above with the same weights.  No reference
 * counterpart as a single call: it replaces Laplacian, Grad and Divergence
 * launched back to back on the same grid (pm/FDM/module/{laplacian,grad,
 * divergence}.py).  Requires three eligible axes. */
int pde_fused_lap_grad_div_f32(const float *s, const float *v, float *lap, float *grad,
                               float *div, int64_t nx, int64_t ny, int64_t nz,
                               int32_t ghost, float wl_minus, float wl_center,
                               float wl_plus, float wc_minus, float wc_center,
                               float wc_plus, float alpha_lap, float alpha_grad,
                               float alpha_div, void *stream);

/* ------------------------------------------------------------- time stepping */

/* out[i] = cur[i] + dt * k[i] for i < n (n = number of floats, ghosts
 * included).  out may alias cur or k.
 * Replaces ForwardEuler.forward (pm/time_step/forwardEuler.py:41-61, kernel
 * :70-79). */
int pde_forward_euler_f32(const float *cur, const float *k, float *out, int64_t n,
                          float dt, void *stream);

/* Plain device copy of n floats (out = in).  Replaces the wp.copy in
 * GridBoundary.copy_array (pm/FDM/boundary/functional_boundary.py:260-263,
 * pm/deprecated/gridBoundary.py:193-196); also the bandwidth yardstick. */
int pde_copy_f32(const float *in, float *out, int64_t n, void *stream);

/* Fused explicit heat step on a vec(ncomp) field (reference loop
 * examples/2D_Heat_Equation.py:58-62: u2 = BC(u); L = lap(u2, alpha);
 * u_next = u2 + dt * L).  `u` must already hold u2 (the boundary ops applied
 * in place, see pde_bc_apply_inplace_f32); writes, for every non-ghost node,
 * u_next[n] = u2[n] + dt * (alpha * acc(u2, n)) with acc the Laplacian sum,
 * and copies u2's ghost values into u_next's ghosts.  u_next must not alias u. */
int pde_heat_step_f32(const float *u, float *u_next, int64_t nx, int64_t ny, int64_t nz,
                      int32_t ncomp, int32_t ghost, float w_minus, float w_center,
                      float w_plus, float alpha, float dt, void *stream);

/* Fused u/v wave step on a scalar field (reference loop
 * examples/wave2d/standingWave.py:72-76 with a per-node coefficient):
 *   v_next[n] = v[n] + dt * (c2[n] * (alpha * acc(u2, n)))
 *   u_next[n] = u2[n] + dt * (first_step ? v[n] : v_next[n])
 * for non-ghost nodes; `u` holds u2 (boundary already applied).  c2 == NULL
 * means a uniform coefficient of 1 (then the scale is alpha alone).  v_next
 * may alias v; u_next must not alias u. */
int pde_wave_step_f32(const float *u, const float *v, const float *c2, float *u_next,
                      float *v_next, int64_t nx, int64_t ny, int64_t nz, int32_t ghost,
                      float w_minus, float w_center, float w_plus, float alpha, float dt,
                      int32_t first_step, void *stream);

/* pde_wave_step_f32 with first_step = (*step == 0), read from device memory
 * (a step counter advanced by pde_counter_add), so one captured graph of a loop
 * segment serves every segment.  No reference counterpart (graph-capturable
 * form of the step above). */
int pde_wave_step_ctr_f32(const float *u, const float *v, const float *c2, float *u_next,
                          float *v_next, int64_t nx, int64_t ny, int64_t nz, int32_t ghost,
                          float w_minus, float w_center, float w_plus, float alpha, float dt,
                          const int64_t *step, void *stream);

/* out[step * stride + i] = field[idx[i]] for i < n, step read from device
 * memory (receiver records of a captured loop).  No reference counterpart. */
int pde_gather_step_f32(const float *field, const int64_t *idx, float *out, int64_t n,
                        int64_t stride, const int64_t *step, void *stream);

/* field[idx[i]] += scale * vals[step * stride + i] for i < n (distinct
 * indices), step read from device memory (point sources of a captured loop).
 * No reference counterpart. */
int pde_scatter_add_step_f32(float *field, const int64_t *idx, const float *vals, int64_t n,
                             int64_t stride, float scale, const int64_t *step, void *stream);

/* *counter += delta on the device (advances the step counter inside a
 * captured loop).  No reference counterpart. */
int pde_counter_add(int64_t *counter, int64_t delta, void *stream);

/* out[i] = field[idx[i]] for i < n (receiver samples of a time loop).  No
 * reference counterpart: the reference's examples read fields back with
 * .numpy() (e.g. examples/2D_Heat_Equation.py:64-65). */
int pde_gather_f32(const float *field, const int64_t *idx, float *out, int64_t n, void *stream);

/* field[idx[i]] += scale * vals[i] for i < n; the indices must be distinct
 * (point sources of a time loop, and the receiver transpose).  No reference
 * counterpart (the reference's loops add forcing with Warp array operators,
 * examples/wave2d/Interference.py:143). */
int pde_scatter_add_f32(float *field, const int64_t *idx, const float *vals, int64_t n,
                        float scale, void *stream);

/* ------------------------------------------------------------------ boundary */

/* One resolved boundary write.  The value v is read at launch from the
 * time array t (device memory: t[0] is the time, t[1 + k] the k-th value a
 * host-evaluated boundary function produced for this call), by value kind:
 *   PDE_BCV_CONST:  v = value, or min(t[0]/tmax, 1) * value when tmax > 0
 *                   (the reference's get_constant_func / get_ramp_func,
 *                   pm/FDM/boundary/boundary_functions.py:9-92)
 *   PDE_BCV_SINE:   v = value * sin(tmax * t[0])   (amplitude, angular
 *                   frequency: the examples' sin_wave,
 *                   examples/wave2d/Interference.py:53-62)
 *   PDE_BCV_ARRAY:  v = t[1 + slot]   (an arbitrary boundary function,
 *                   evaluated by the host layer on the device before the
 *                   launch; pm/FDM/boundary/functional_boundary.py:376 and
 *                   pm/deprecated/func_boundary.py:118)
 * and written by kind:
 *   PDE_BC_SET:     out[dst] = v
 *   PDE_BC_COPY:    out[dst] = in[src]
 *   PDE_BC_GHOST_D: out[dst] = 2*v - in[src]            (Dirichlet ghost)
 *   PDE_BC_GHOST_N: out[dst] = coef*v + in[src]         (Neumann ghost, coef =
 *                                                       -sign(s)*2*dx)
 * A NULL t reads as time 0 and every array value as 0; the host passes an
 * array of 1 + (largest slot + 1) floats whenever a table has array values. */
#define PDE_BC_SET 0
#define PDE_BC_COPY 1
#define PDE_BC_GHOST_D 2
#define PDE_BC_GHOST_N 3

#define PDE_BCV_CONST 0
#define PDE_BCV_SINE 1
#define PDE_BCV_ARRAY 2

typedef struct pde_bc_op {
    int64_t dst;
    int64_t src;
    int16_t kind;       /* PDE_BC_* */
    int16_t vkind;      /* PDE_BCV_* */
    union {
        float value;    /* PDE_BCV_CONST, PDE_BCV_SINE */
        int32_t slot;   /* PDE_BCV_ARRAY */
    };
    float coef;
    float tmax;         /* PDE_BCV_CONST: ramp end time (0: none); PDE_BCV_SINE: omega */
} pde_bc_op; /* 32 bytes */

/* out = in (n floats; in and out must not alias), then every op applied with
 * all reads from `in`.  The host resolves
 * a GridBoundary's groups into `ops` so that each dst appears once; the result
 * equals the reference's copy + sequential group launches.
 * Replaces GridBoundary.forward of pm/FDM/boundary/functional_boundary.py:265-294
 * (kernel :356-400) and of pm/deprecated/gridBoundary.py:218-246 (kernel
 * :267-302).  t may be NULL when no op has tmax > 0. */
int pde_bc_apply_f32(const float *in, float *out, int64_t n, const pde_bc_op *ops,
                     int64_t nops, const float *t, void *stream);

/* The same ops applied in place to u: every op's value is first gathered from
 * the unmodified u into staging[nops], then scattered, so reads still see the
 * pre-boundary field.  staging holds nops floats of caller-owned device
 * memory.  No reference counterpart as one call: it is the copy-free form of
 * GridBoundary used by the fused steps. */
int pde_bc_apply_inplace_f32(float *u, const pde_bc_op *ops, int64_t nops, const float *t,
                             float *staging, void *stream);

/* pde_bc_apply_inplace_f32 with no staging argument and no allocation: a
 * table of more than 4096 ops stages its values in a file-scope device array
 * of PDE_BC_SCRATCH_FLOATS floats (4 MiB) that each GPU the library runs on
 * holds until the process ends (its address is looked up once per device, on
 * the first call with that device current).  Same values as
 * pde_bc_apply_inplace_f32.  PDE_ETOOBIG when nops > PDE_BC_SCRATCH_FLOATS.
 * Every call on a GPU shares that GPU's array (each slot is written before it
 * is read, so nothing needs clearing): calls that overlap on one GPU, on
 * different streams or threads, race.  No counterpart in the reference: it is
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). */
#define PDE_BC_SCRATCH_FLOATS (1 << 20)
int pde_bc_apply_inplace_noalloc_f32(float *u, const pde_bc_op *ops, int64_t nops,
                                     const float *t, void *stream);

/* Periodic ghost fill, in place: for each eligible axis ascending, ghost layer
 * i < g takes interior layer i + E and ghost layer i >= P - g takes i - E
 * (E = P - 2g), over the full extent of the other axes.  No reference
 * counterpart (the reference has no periodic boundary); it supplies the
 * periodic grids the 3D operator benchmark runs on. */
int pde_periodic_fill_f32(float *u, int64_t nx, int64_t ny, int64_t nz, int32_t ncomp,
                          int32_t ghost, void *stream);

/* ------------------------------------------------------------------ adjoints
 *
 * Transposes for reverse-mode differentiation.  None has a reference
 * counterpart (PDE_Module has no autodiff: no wp.Tape anywhere); they are the
 * exact transposes of the forward entry points above, each forward operator
 * viewed as a linear map from the whole padded field to its interior nodes.
 * Every one writes the whole padded input-gradient field and never reads the
 * ghost values of the incoming gradient.  Input and output must not alias. */

/* gin[m] = alpha * sum over axes (ascending), s = -1, 0, +1, of
 * w_s * gout[m - s e] for each interior node m - s e (transpose of
 * pde_laplacian_f32). */
int pde_laplacian_adjoint_f32(const float *gout, float *gin, int64_t nx, int64_t ny,
                              int64_t nz, int32_t ncomp, int32_t ghost, float w_minus,
                              float w_center, float w_plus, float alpha, void *stream);

/* Transpose of pde_grad_f32: gout holds ncomp x D per node, gin ncomp per node. */
int pde_grad_adjoint_f32(const float *gout, float *gin, int64_t nx, int64_t ny, int64_t nz,
                         int32_t ncomp, int32_t ghost, float w_minus, float w_center,
                         float w_plus, float alpha, void *stream);

/* Transpose of pde_divergence_f32: gout holds nrows per node, gin nrows x D. */
int pde_divergence_adjoint_f32(const float *gout, float *gin, int64_t nx, int64_t ny,
                               int64_t nz, int32_t nrows, int32_t ghost, float w_minus,
                               float w_center, float w_plus, float alpha, void *stream);

/* Transpose of pde_heat_step_f32 (as a map u2 -> u_next):
 * gu2 = g + dt * (alpha * laplacian_transpose_sum(g)). */
int pde_heat_step_adjoint_f32(const float *g, float *gu2, int64_t nx, int64_t ny, int64_t nz,
                              int32_t ncomp, int32_t ghost, float w_minus, float w_center,
                              float w_plus, float alpha, float dt, void *stream);

/* Transpose of pde_wave_step_f32 in (u2, v) and its derivative in c2.  With
 * gvt = gvn + dt*gun (gvn when first_step) on interior nodes:
 *   gv  = gvt (gvn + dt*gun when first_step) on interior nodes, gvn on ghosts;
 *   gu2 = gun + dt * (alpha * laplacian_transpose_sum(c2 * gvt));
 *   gc2 += dt * (gvt * (alpha * acc(u2))) on interior nodes (skipped when
 *   gc2 is NULL; u2 is the boundary-applied state the forward step read).
 * c2 == NULL means the uniform coefficient 1. */
int pde_wave_step_adjoint_f32(const float *gun, const float *gvn, const float *c2,
                              const float *u2, float *gu2, float *gv, float *gc2, int64_t nx,
                              int64_t ny, int64_t nz, int32_t ghost, float w_minus,
                              float w_center, float w_plus, float alpha, float dt,
                              int32_t first_step, void *stream);

/* pde_wave_step_adjoint_f32 with first_step = (*step == 0) read from device
 * memory (graph-capturable form; see pde_wave_step_ctr_f32). */
int pde_wave_step_adjoint_ctr_f32(const float *gun, const float *gvn, const float *c2,
                                  const float *u2, float *gu2, float *gv, float *gc2, int64_t nx,
                                  int64_t ny, int64_t nz, int32_t ghost, float w_minus,
                                  float w_center, float w_plus, float alpha, float dt,
                                  const int64_t *step, void *stream);

/* ------------------------------------------- fused wave loop steps (2D grids) */
/* Point sources and receivers of a 2D scalar field, listed by padded row:
 * the points of row x are [row_start[x], row_start[x+1]) (row_start has
 * p1 + 1 entries, p1 = padded rows); col[p] is the point's padded column and
 * slot[p] its column in `data`, whose row n belongs to step n (row stride
 * `stride` floats).  Sources read data (v[point] += scale * data[n][slot]
 * before the step, points in list order); receivers write it
 * (data[n][slot] = u_next[point]). */
typedef struct pde_point_set {
    const int32_t *row_start;
    const int32_t *col;
    const int32_t *slot;
    float *data;
    int64_t stride;
    float scale;
    int32_t reserved;
} pde_point_set;

/* One step of the wave loop on a 2D scalar field in one launch: the source
 * injection, the fused u/v step of pde_wave_step_ctr_f32 and the receiver
 * samples, with n = *step and *step_next = n + 1 written by the kernel
 * (step and step_next are two device counters used alternately, so no
 * kernel reads a counter another one of the same step writes).  sources and
 * receivers are host pointers to point sets (NULL: none).  Same values as
 * pde_scatter_add_step_f32 + pde_wave_step_ctr_f32 + pde_gather_step_f32 +
 * pde_counter_add.  PDE_ESHAPE unless the field is 2D.  No reference
 * counterpart (the reference loop is examples/wave2d/standingWave.py:70-76,
 * without sources or receivers). */
int pde_wave_step_fused_f32(const float *u, const float *v, const float *c2, float *u_next,
                            float *v_next, int64_t nx, int64_t ny, int64_t nz, int32_t ghost,
                            float w_minus, float w_center, float w_plus, float alpha, float dt,
                            const int64_t *step, int64_t *step_next,
                            const pde_point_set *sources, const pde_point_set *receivers,
                            void *stream);

/* The transpose of one such step in one launch: the receiver gradients
 * (receivers->data rows, scale 1) added to g_un at the receivers, then
 * pde_wave_step_adjoint_ctr_f32, with n = *step and *step_next = n - 1. */
int pde_wave_step_adjoint_fused_f32(const float *gun, const float *gvn, const float *c2,
                                    const float *u2, float *gu2, float *gv, float *gc2,
                                    int64_t nx, int64_t ny, int64_t nz, int32_t ghost,
                                    float w_minus, float w_center, float w_plus, float alpha,
                                    float dt, const int64_t *step, int64_t *step_next,
                                    const pde_point_set *receivers, void *stream);

/* *dst = *src + delta on the device (moves a step counter between the two
 * alternating counters of the fused steps).  No reference counterpart. */
int pde_counter_copy(const int64_t *src, int64_t *dst, int64_t delta, void *stream);

/* Boundary ring of a 2D field for pde_wave_reverse_adjoint_pair_f32 (host
 * built, device arrays).  The ring is every node less than `depth` from the
 * field's edge; it is cut into pieces (output rectangles rows [r0, r1) x
 * columns [c0, c1), at most 32 along the edge).  Entries refer to flat node
 * indices; for a piece they cover its rectangle grown by one node:
 *   w[w_beg, w_end):  boundary-written nodes (node, a = save slot)
 *   t[t_beg, t_end):  boundary-transpose sources (node, a..b = range of
 *                     cdst/csign, flags bit 0: the source is itself written)
 *   z[z_beg, z_end):  written nodes that are no source (node)
 * flags bit 1 marks entries inside the piece's own rectangle. */
typedef struct pde_ring_piece {
    int32_t r0, r1, c0, c1;
    int32_t w_beg, w_end, t_beg, t_end, z_beg, z_end;
    int32_t reserved[2];
} pde_ring_piece; /* 48 bytes */

typedef struct pde_ring_entry {
    int32_t node, a, b, flags;
} pde_ring_entry; /* 16 bytes */

typedef struct pde_ring_plan {
    int32_t depth, npieces;
    int64_t nops; /* ops in the table the plan was built for; every entry's op
                     slot is below it, and the entry points refuse a table of
                     any other size (PDE_EINVAL) */
    const pde_ring_piece *pieces;
    const pde_ring_entry *w, *t, *z;
    const int32_t *cdst;
    const float *csign;
} pde_ring_plan;

/* The heat step of pde_heat_step_f32 with the in-place boundary folded into
 * the same launch, on a 2D scalar field: u is left boundary-applied (the ops
 * of pde_bc_apply_inplace_f32, t as there) and un = u2 + dt*(alpha*lap(u2))
 * on interior nodes, u2 elsewhere; the same values as the two launches.  ring
 * is the field's boundary ring (pde_ring_plan, built by the host for the op
 * table and checked so that each piece's ops read inside the piece and no
 * scalar another piece writes; PDE_EINVAL unless ring->nops == nops).  No
 * reference counterpart as one call: it
 * replaces GridBoundary + Laplacian + ForwardEuler of
 * examples/2D_Heat_Equation.py:58-62 with one launch. */
int pde_heat_step_ring_f32(float *u, float *un, int64_t nx, int64_t ny, int64_t nz,
                           int32_t ghost, const pde_bc_op *ops, int64_t nops, const float *t,
                           float w_minus, float w_center, float w_plus, float alpha, float dt,
                           const pde_ring_plan *ring, void *stream);

/* Two transposed wave steps, n then n - 1 (n = *step, *step_next = n - 2),
 * walking the forward states backwards instead of reading them: from the
 * step-(n+1) state (u, v: pre-boundary u, the v the step wrote) it rebuilds
 * u2 = u - dt*v and v_n = v - dt*c2*alpha*lap(u2) - source on interior
 * nodes, restores the boundary-written scalars from save (rows n and n-1, as
 * recorded by pde_bc_apply_chunked_inplace_save_f32), and runs the transposes
 * of pde_wave_step_adjoint_fused_f32 and of the boundary in between, so
 * (uo, vo) is the step-(n-1) state and (guo, gvo) the gradient with respect
 * to it; gc2 (may be NULL) accumulates both steps' c2 terms in place.  Nodes
 * at least ring->depth from the edge run as register row tiles, the ring as
 * LDS pieces, all in one launch.  The gradients g_u, g_v are those of two
 * pde_wave_step_adjoint_fused_f32 + boundary-transpose steps bit for bit; gc2
 * differs from the recomputed-state result only through the rounding of the
 * rebuilt states.  Never the first step of the loop (n >= 1).  ring must be
 * the plan of the op table whose scalars save holds (PDE_EINVAL unless
 * ring->nops == nops).  No reference counterpart. */
int pde_wave_reverse_adjoint_pair_f32(const float *u, const float *v, const float *gu,
                                      const float *gv, const float *c2, float *uo, float *vo,
                                      float *guo, float *gvo, float *gc2, int64_t nx,
                                      int64_t ny, int64_t nz, int32_t ghost, float w_minus,
                                      float w_center, float w_plus, float alpha, float dt,
                                      const int64_t *step, int64_t *step_next,
                                      const pde_point_set *sources,
                                      const pde_point_set *receivers, const float *save,
                                      int64_t nops, const pde_ring_plan *ring, void *stream);

/* pde_bc_apply_inplace_f32 in one full-chip launch.  The host orders the ops
 * in conflict-closed chunks: chunk c is ops [chunk_start[c], chunk_start[c+1])
 * (at most 256), and no op reads a scalar that an op of another chunk writes;
 * each workgroup then applies its chunk with the two-phase semantics on its
 * own.  No reference counterpart (the in-place form of the boundary apply). */
int pde_bc_apply_chunked_inplace_f32(float *u, const pde_bc_op *ops, const int64_t *chunk_start,
                                     int64_t nchunks, const float *t, void *stream);

/* pde_bc_apply_chunked_inplace_f32 that also records, for op m (chunk order),
 * the value its destination held before the apply in save[n * nops + m],
 * n = *step.  The wave loop's time-reversed gradient rebuilds each step's
 * pre-boundary field from these.  No reference counterpart. */
int pde_bc_apply_chunked_inplace_save_f32(float *u, const pde_bc_op *ops,
                                          const int64_t *chunk_start, int64_t nchunks,
                                          int64_t nops, const float *t, float *save,
                                          const int64_t *step, void *stream);

/* pde_bc_adjoint_inplace_f32 in one full-chip launch over conflict-closed
 * chunks: chunk c stages sources [item_start[c], item_start[c+1]) (at most
 * 256; src_idx/src_flags/off in chunk order, off indexing cdst/csign) and
 * clears zero_idx[zero_start[c] .. zero_start[c+1]); every written scalar
 * that a chunk's sources read belongs to the same chunk. */
int pde_bc_adjoint_chunked_inplace_f32(float *g, const int64_t *src_idx, const int32_t *src_flags,
                                       const int64_t *off, const int64_t *cdst,
                                       const float *csign, const int64_t *item_start,
                                       const int64_t *zero_idx, const int64_t *zero_start,
                                       int64_t nchunks, void *stream);

/* Transpose of the linear part of pde_bc_apply_f32: gin = gout, then
 * gin[zero_idx[i]] = 0 (the written scalars; the host passes those that are
 * not sources, the others are overwritten next), then for each distinct
 * source src_idx[i] (flag bit 0 set when it is itself written):
 * gin[src] = (flag ? 0 : gout[src]) + sum over q in [off[i], off[i+1]) of
 * csign[q] * gout[cdst[q]].  The host builds these arrays from the op table. */
int pde_bc_adjoint_f32(const float *gout, float *gin, int64_t n, const int64_t *zero_idx,
                       int64_t nzero, const int64_t *src_idx, const int32_t *src_flags,
                       const int64_t *off, const int64_t *cdst, const float *csign,
                       int64_t nsrc, void *stream);

/* The same transpose applied to g in place: the new source values are first
 * gathered into staging[nsrc] from the untouched g, then, in one launch, the
 * staged values are stored and the zero_idx scalars zeroed.  Here zero_idx
 * must not contain any src_idx entry (the written scalars that are not
 * sources; a source that is also written gets its staged value, whose flag
 * already drops its own term).  No full-field copy. */
int pde_bc_adjoint_inplace_f32(float *g, const int64_t *zero_idx, int64_t nzero,
                               const int64_t *src_idx, const int32_t *src_flags,
                               const int64_t *off, const int64_t *cdst, const float *csign,
                               int64_t nsrc, float *staging, void *stream);

/* out[i] = scale * in[i] (the dt * g factor of ForwardEuler's transpose). */
int pde_scale_f32(const float *in, float *out, int64_t n, float scale, void *stream);

/* ----------------------------------------------------------- batched grids */
/* Many independent 2D scalar heat grids of uneven sizes, packed one after the
 * other in one float buffer, stepped together.  The reference runs one grid
 * per module call (examples/2D_Heat_Equation.py:58-62: boundary, Laplacian,
 * ForwardEuler); an ensemble there is a Python loop over instances.  These
 * entry points have no reference counterpart: they are the batched form of
 * pde_bc_apply_inplace_f32 + pde_heat_step_f32, giving every grid the same
 * bits as stepping it alone. */
typedef struct pde_batch_grid {
    int64_t offset;    /* first float of the grid's padded field in the buffer */
    int64_t op_begin;  /* its boundary ops: [op_begin, op_end) of the packed  */
    int64_t op_end;    /* table (dst/src are indices into the whole buffer)   */
    int32_t p1, p2;    /* padded rows, padded columns (column index fastest)  */
    int32_t ghost;     /* ghost width g (interior = [g, p - g) on both axes)  */
    int32_t reserved0;
    float wm, wc, wp;  /* Laplacian weights 1/dx^2, -2/dx^2, 1/dx^2           */
    float alpha, dt;
    float reserved1;
} pde_batch_grid; /* 64 bytes */

/* Boundary of every grid, in place: workgroup i applies grid i's ops
 * (at most pde_batch_bc_capacity() of them) with the two-phase semantics of
 * pde_bc_apply_inplace_f32.  t: device time scalar for ramp values, or NULL. */
int pde_bc_apply_batch_inplace_f32(float *u, const pde_batch_grid *grids, int64_t ngrids,
                                   const pde_bc_op *ops, const float *t, void *stream);

/* Largest number of boundary ops of one grid in pde_bc_apply_batch_inplace_f32. */
int pde_batch_bc_capacity(void);

/* Fused heat step of every grid (un = u2 + dt * alpha * lap(u2) on interior
 * nodes, u2 elsewhere) over a tile list: tiles[4*i .. 4*i+3] = (grid, first
 * row, first 4-column chunk, segment width cw in {64, 32, 16}); a tile covers
 * waves * (64 / cw) * rows_per_lane rows and 4 * cw columns.  rows_per_lane
 * in {1, 2, 4, 8}, waves in {1, 2, 4}; blocks = 0 launches one block per
 * tile, else `blocks` blocks loop over the tiles. */
int pde_heat_batch_step_f32(const float *u, float *un, const pde_batch_grid *grids,
                            int64_t ngrids, const int32_t *tiles, int64_t ntiles,
                            int32_t rows_per_lane, int32_t waves, int64_t blocks, void *stream);

/* Several heat steps per launch on LDS-resident tiles.  A tile is a region
 * of one grid (rows [r0, r0+rh), columns [c0, c0+rw) of its padded field, at
 * most PDE_HEAT_TILE_ROWS x PDE_HEAT_TILE_PITCH cells, inside the grid) that
 * one workgroup loads once and steps `steps` times, writing back its output
 * rectangle [o_r0, o_r1) x [o_c0, o_c1).  Its boundary ops are
 * tops[op_beg .. op_end) (at most PDE_HEAT_TILE_MAXOPS): region-local cells
 * (row * PDE_HEAT_TILE_PITCH + column) of the written scalar and of its
 * source (-1 for a PDE_BC_SET), and the index of the full op in ops (its kind
 * and value; PDE_BCV_ARRAY values are not allowed: they are evaluated once
 * per step on the host side). */
#define PDE_HEAT_TILE_PITCH 128
#define PDE_HEAT_TILE_ROWS 128
#define PDE_HEAT_TILE_MAXOPS 2048

typedef struct pde_heat_tile {
    int32_t grid;               /* index into grids                          */
    int32_t r0, c0, rh, rw;     /* LDS region                                */
    int32_t o_r0, o_r1;         /* output rows    [o_r0, o_r1)               */
    int32_t o_c0, o_c1;         /* output columns [o_c0, o_c1)               */
    int32_t op_beg, op_end;     /* the tile's boundary ops in tops           */
    int32_t reserved;
} pde_heat_tile; /* 48 bytes */

typedef struct pde_heat_tile_op {
    int32_t ldst, lsrc;         /* region-local cells (lsrc -1 for a SET)    */
    int32_t op;                 /* index into ops                            */
    int32_t reserved;
} pde_heat_tile_op; /* 16 bytes */

/* `steps` heat steps of every grid (each the in-place boundary of
 * pde_bc_apply_inplace_f32, t as there, then pde_heat_step_f32), from u into
 * un (u is not written; u != un); the same bits as `steps` pairs of those
 * launches when the tiles' outputs cover every cell of every grid exactly
 * once, each output lies `steps` cells inside every side where its region was
 * cut from the grid, and every op's source cell stays inside the valid part
 * of the region as long as its written cell does (the host planner builds and
 * checks this).  The sub-steps share one time array.  No reference
 * counterpart as one call: it replaces `steps` iterations of GridBoundary +
 * Laplacian + ForwardEuler (examples/2D_Heat_Equation.py:58-62), for one grid
 * or a batch. */
int pde_heat_multistep_f32(const float *u, float *un, const pde_batch_grid *grids,
                           int64_t ngrids, const pde_heat_tile *tiles, int64_t ntiles,
                           const pde_heat_tile_op *tops, const pde_bc_op *ops, const float *t,
                           int32_t steps, void *stream);

/* ------------------------------------------------------------ node-wise glue */
/* The element-wise operations between stencil outputs.  `nodes` counts padded
 * nodes (ghosts included: the reference's element-wise launches cover the
 * whole array); per-node values are contiguous (AoS). */

/* out[n][i][j] = a[n][i] * b[n][j]; a is (nodes, la), b (nodes, lb), out
 * (nodes, la, lb).  Replaces OuterProduct (pm/FDM/module/elementWiseOps.py:10-49,
 * element op wp.outer at :18) launched through elementWiseOp
 * (pm/FDM/functional/_elementWiseOps.py:5-31). */
int pde_outer_f32(const float *a, const float *b, float *out, int64_t nodes, int32_t la,
                  int32_t lb, void *stream);

/* out[n][k] = s[n] * b[n][k]; s is a (nodes, 1) field.  Replaces
 * scalarVectorMult (pm/FDM/module/elementWiseOps.py:52-89; element op
 * pm/FDM/kernel/elementWiseOps.py:6-23).  out may alias b. */
int pde_scalar_mul_f32(const float *s, const float *b, float *out, int64_t nodes, int32_t lb,
                       void *stream);

/* out[n][k] = b[n][k] / s[n] (IEEE division).  Replaces the MapWise map
 * `lambda m, rho: m / rho[0]` (examples/LDC/LDC_RK2.py:117,
 * pm/stencil/mapWise.py:8-44).  out may alias b. */
int pde_div_scalar_f32(const float *b, const float *s, float *out, int64_t nodes, int32_t lb,
                       void *stream);

/* out[n][i] = sum_j m[n][i][j] * v[n][j], j ascending; m is (nodes, rows,
 * cols) row-major.  Replaces the Warp array product `du * u` of a mat(N, D)
 * field and a vec(D) field (examples/LDC/LDC_No_Graph.py:87). */
int pde_matvec_f32(const float *m, const float *v, float *out, int64_t nodes, int32_t rows,
                   int32_t cols, void *stream);

/* out[i] = c0*x0[i] + c1*x1[i] + ... over k (1..4) terms, left to right, for
 * i < n floats.  With coefficients +-1 this is exactly the chain of Warp
 * array operators `a - b + c + d` (examples/LDC/LDC_No_Graph.py:90,
 * examples/LDC/LDC_RK2.py:133) in one pass.  out may alias any x. */
int pde_lincomb_f32(const float *x0, const float *x1, const float *x2, const float *x3, float c0,
                    float c1, float c2, float c3, int32_t k, float *out, int64_t n,
                    void *stream);

/* out[n][c] = scale * (x[n][c] - shift) on component comp (every component
 * when comp < 0), other components copied.  Replaces the MapWise map
 * get_p_op, rho -> c2 * (rho - density) (examples/LDC/LDC_RK2.py:111-116).
 * out may alias x. */
int pde_affine_f32(const float *x, float *out, int64_t nodes, int32_t ncomp, int32_t comp,
                   float scale, float shift, void *stream);

/* ------------------------------------------------- projection-method cavity
 *
 * Kernels of the incompressible lid-driven cavity by pressure projection (an
 * explicit predictor from the operators above, a pressure Poisson equation
 * solved by a fixed number of Jacobi sweeps, then the velocity correction).
 * None has a reference counterpart: the reference's cavity examples use the
 * artificial-compressibility method (examples/LDC/LDC_No_Graph.py) instead. */

/* One Jacobi sweep on the interior nodes of a scalar field:
 *   acc = 0; over eligible axes ascending: acc += p[n-e]; acc += p[n+e];
 *   out[n] = inv * (acc - h2 * rhs[n])      (inv = 1/(2 ndim), h2 = dx^2).
 * Ghosts of p are read (apply the pressure boundary first), those of out are
 * not written.  out must not alias p or rhs. */
int pde_jacobi_f32(const float *p, const float *rhs, float *out, int64_t nx, int64_t ny,
                   int64_t nz, int32_t ghost, float h2, float inv, void *stream);

/* pde_jacobi_f32 on a field whose ghosts a legacy Neumann-0 boundary on every
 * face would fill first (ghost = coef * 0 + mirror, coef = -2 dx on '-' faces,
 * +2 dx on '+' faces): each ghost-layer neighbour is read as that value from
 * its mirror across the wall, so boundary + sweep is one launch with the same
 * FP32 results; p's ghosts are neither read nor written.  Every eligible
 * interior extent must be >= 2. */
int pde_jacobi_neumann0_f32(const float *p, const float *rhs, float *out, int64_t nx, int64_t ny,
                            int64_t nz, int32_t ghost, float h2, float inv, void *stream);

/* Transpose of pde_jacobi_f32: gp[m] = inv * (sum over axes ascending of
 * gout[m+e] then gout[m-e], each taken when that node is interior) on every
 * padded node; grhs[n] = -h2 * (inv * gout[n]) on interior nodes, 0 on ghosts
 * (grhs may be NULL). */
int pde_jacobi_adjoint_f32(const float *gout, float *gp, float *grhs, int64_t nx, int64_t ny,
                           int64_t nz, int32_t ghost, float h2, float inv, void *stream);

/* Transpose of pde_jacobi_neumann0_f32 as a linear map of (p's interior, rhs):
 * gp = 0 on ghosts; on interior m, inv * (sum over axes ascending of g[m + e]
 * then g[m - e], each when interior), then per axis ascending + inv * g[m - e]
 * when m - e is the low-wall node (m - 2e a ghost) and + inv * g[m + e] when
 * m + e is the high-wall node (the folded ghost mirrors); grhs[m] +=
 * -h2 * (inv * g[m]) on interior m (accumulated; ghosts untouched).  The
 * same values as pde_jacobi_adjoint_f32 followed by the Neumann-0 boundary's
 * transpose, in one launch.  Every eligible axis needs 2+ interior nodes.  No
 * reference counterpart (the reference has no adjoints); it serves the
 * gradient of the projection cavity. */
int pde_jacobi_neumann0_adjoint_f32(const float *gout, float *gp, float *grhs, int64_t nx,
                                    int64_t ny, int64_t nz, int32_t ghost, float h2, float inv,
                                    void *stream);

/* pde_jacobi_neumann0_f32 / pde_jacobi_neumann0_adjoint_f32 over `batch`
 * instances of the same 1D or 2D grid stored one after another (instance b
 * at offset b * nx * ny * nz in every array): one launch for all, each
 * instance the same values as swept alone.  PDE_ESHAPE for 3D grids.  No
 * reference counterpart; it serves the projection cavity stepped over many
 * seeds at once. */
int pde_jacobi_neumann0_batch_f32(const float *p, const float *rhs, float *out, int64_t batch,
                                  int64_t nx, int64_t ny, int64_t nz, int32_t ghost, float h2,
                                  float inv, void *stream);
int pde_jacobi_neumann0_adjoint_batch_f32(const float *gout, float *gp, float *grhs,
                                          int64_t batch, int64_t nx, int64_t ny, int64_t nz,
                                          int32_t ghost, float h2, float inv, void *stream);

/* Two consecutive pde_jacobi_neumann0_f32 sweeps in one launch, over `batch`
 * instances of a 2D grid stored one after another: out = J(J(p, rhs), rhs)
 * with the same values as two launches (the intermediate sweep never leaves
 * registers); out's ghosts untouched.  PDE_ESHAPE unless the grid is 2D with
 * 2+ interior nodes per axis.  No reference counterpart; it serves the
 * projection cavity's pressure sweeps. */
int pde_jacobi_neumann0_x2_f32(const float *p, const float *rhs, float *out, int64_t batch,
                               int64_t nx, int64_t ny, int64_t nz, int32_t ghost, float h2,
                               float inv, void *stream);

/* Two consecutive pde_jacobi_neumann0_adjoint_f32 sweeps in one launch (same
 * batching and limits): gp = T(T(gout)), grhs accumulated with the first
 * sweep's contribution, then the second's, as two launches would. */
int pde_jacobi_neumann0_adjoint_x2_f32(const float *gout, float *gp, float *grhs, int64_t batch,
                                       int64_t nx, int64_t ny, int64_t nz, int32_t ghost,
                                       float h2, float inv, void *stream);

/* Most sweeps one multi-sweep Jacobi launch runs. */
#define PDE_JACOBI_MAX_SWEEPS 48

/* `sweeps` (1 .. PDE_JACOBI_MAX_SWEEPS) folded Neumann-0 sweeps
 * (pde_jacobi_neumann0_f32) in one launch, for `batch` instances of a 2D
 * grid stored one after another (arguments as pde_jacobi_neumann0_x2_f32):
 * each workgroup loads a tile of the grid into LDS (the node's output
 * rectangle grown by `sweeps` nodes on every side inside the grid), sweeps
 * it there and stores the rectangle.  out = the interior after `sweeps`
 * sweeps from p (out != p, != rhs; out's ghosts are not written).  The values
 * of `sweeps` launches of pde_jacobi_neumann0_f32 bit for bit.  PDE_ESHAPE
 * for grids that are not 2D (or have an interior side under 2) or batch >
 * 65535.  No reference counterpart (the projection cavity's pressure solve). */
int pde_jacobi_neumann0_multisweep_f32(const float *p, const float *rhs, float *out,
                                       int64_t batch, int64_t nx, int64_t ny, int64_t nz,
                                       int32_t ghost, float h2, float inv, int32_t sweeps,
                                       void *stream);

/* `sweeps` transposed sweeps (pde_jacobi_neumann0_adjoint_f32) in one
 * launch, tiled as pde_jacobi_neumann0_multisweep_f32: gp = the transpose
 * applied `sweeps` times to g, or to g + g2 when g2 is not NULL (the values
 * of pde_lincomb_f32 with coefficients 1, 1), 0 on ghosts; on interior nodes
 * grhs gathers each sweep's -h2 * (inv * input), in sweep order, starting
 * from grhs (accumulate 1) or from 0 (accumulate 0: grhs's interior is
 * overwritten, its ghosts are not written).  gp and grhs distinct from each
 * other and from g and g2.  The values of `sweeps` launches of
 * pde_jacobi_neumann0_adjoint_f32 bit for bit (on a zeroed grhs for
 * accumulate 0).  No reference counterpart. */
int pde_jacobi_neumann0_adjoint_multisweep_f32(const float *g, const float *g2, float *gp,
                                               float *grhs, int64_t batch, int64_t nx,
                                               int64_t ny, int64_t nz, int32_t ghost, float h2,
                                               float inv, int32_t sweeps, int32_t accumulate,
                                               void *stream);

/* The projection cavity's predictor on a 2D vec2 field in one launch:
 * us = u2 + dt * (-(grad u2) u2 + nu * lap(u2)) with the values of
 * pde_grad_f32 (alpha 1) + pde_matvec_f32 + pde_laplacian_f32 (alpha nu) +
 * pde_lincomb_f32 (-1, +1) + pde_forward_euler_f32 on zero-ghost work
 * buffers, every padded node written.  u2's ghosts are read.  No reference
 * counterpart (the reference's cavities use glue-operator chains,
 * examples/LDC/LDC_No_Graph.py:87,90). */
int pde_cavity_predict_f32(const float *u2, float *us, int64_t nx, int64_t ny, int64_t nz,
                           int32_t ghost, float lap_wm, float lap_wc, float lap_wp,
                           float grad_wm, float grad_wc, float grad_wp, float nu, float dt,
                           void *stream);

/* The projection cavity's correction on a 2D grid in one launch:
 * u_out = us2 - dt * grad(B_p(p)) (u_out, us2 vec2; p scalar) with the
 * values of a legacy vonNeumann_BC('ALL', 0) applied to p, pde_grad_f32
 * (alpha 1) on a zero-ghost buffer and pde_forward_euler_f32 (-dt) over every
 * node; p's ghosts are not read.  u_out may alias us2, not p.  Every
 * interior extent must be >= 2.  No reference counterpart. */
int pde_cavity_correct_f32(const float *p, const float *us2, float *u_out, int64_t nx, int64_t ny,
                           int64_t nz, int32_t ghost, float grad_wm, float grad_wc,
                           float grad_wp, float dt, void *stream);

/* Transpose of the predictor's interior part (the map u2 -> us of
 * pde_cavity_predict_f32) in one launch: gu2 = gus + nu lap^T(dt gus) +
 * (grad u2)^T (-dt gus) + grad^T(outer(-dt gus, u2)), with the values of
 * pde_scale_f32, pde_laplacian_adjoint_f32, pde_grad_f32, pde_outer_f32,
 * pde_matvec_t_f32, pde_grad_adjoint_f32 and pde_lincomb_f32 run in that
 * chain; every padded node written.  No reference counterpart. */
int pde_cavity_adjoint_predict_f32(const float *gus, const float *u2, float *gu2, int64_t nx,
                                   int64_t ny, int64_t nz, int32_t ghost, float lap_wm,
                                   float lap_wc, float lap_wp, float grad_wm, float grad_wc,
                                   float grad_wp, float nu, float dt, void *stream);

/* The correction's transpose in the pressure, up to the boundary's
 * transpose, in one launch: gpf = grad^T(-dt * gu_out) (gpf scalar, gu_out
 * vec2, 2D), the values of pde_scale_f32(-dt) + pde_grad_adjoint_f32
 * (alpha 1); every node of gpf written; gpf may not alias gu_out.  No
 * reference counterpart. */
int pde_cavity_adjoint_correct_f32(const float *gu_out, float *gpf, int64_t nx, int64_t ny,
                                   int64_t nz, int32_t ghost, float grad_wm, float grad_wc,
                                   float grad_wp, float dt, void *stream);

/* gus2 = gu_out + div^T(grhs) (alpha = the divergence's scale, 1/dt in the
 * cavity) on a 2D grid (grhs scalar, gu_out and gus2 vec2) in one launch:
 * the values of pde_divergence_adjoint_f32 + pde_lincomb_f32 (1, 1); gus2
 * may alias gu_out.  No reference counterpart. */
int pde_cavity_adjoint_div_f32(const float *grhs, const float *gu_out, float *gus2, int64_t nx,
                               int64_t ny, int64_t nz, int32_t ghost, float grad_wm,
                               float grad_wc, float grad_wp, float alpha, void *stream);

/* The five cavity kernels above for `batch` (<= 65535) instances of the 2D
 * grid stored one after another (every array holds batch instances; the
 * batch goes before nx): one launch each, instance b's values those of the
 * single-instance entry point on its own arrays.  No reference counterpart
 * (the projection cavity's seeds stepped together). */
int pde_cavity_predict_batch_f32(const float *u2, float *us, int64_t batch, int64_t nx,
                                 int64_t ny, int64_t nz, int32_t ghost, float lap_wm,
                                 float lap_wc, float lap_wp, float grad_wm, float grad_wc,
                                 float grad_wp, float nu, float dt, void *stream);
int pde_cavity_correct_batch_f32(const float *p, const float *us2, float *u_out, int64_t batch,
                                 int64_t nx, int64_t ny, int64_t nz, int32_t ghost,
                                 float grad_wm, float grad_wc, float grad_wp, float dt,
                                 void *stream);
int pde_cavity_adjoint_predict_batch_f32(const float *gus, const float *u2, float *gu2,
                                         int64_t batch, int64_t nx, int64_t ny, int64_t nz,
                                         int32_t ghost, float lap_wm, float lap_wc,
                                         float lap_wp, float grad_wm, float grad_wc,
                                         float grad_wp, float nu, float dt, void *stream);
int pde_cavity_adjoint_correct_batch_f32(const float *gu_out, float *gpf, int64_t batch,
                                         int64_t nx, int64_t ny, int64_t nz, int32_t ghost,
                                         float grad_wm, float grad_wc, float grad_wp, float dt,
                                         void *stream);
int pde_cavity_adjoint_div_batch_f32(const float *grhs, const float *gu_out, float *gus2,
                                     int64_t batch, int64_t nx, int64_t ny, int64_t nz,
                                     int32_t ghost, float grad_wm, float grad_wc,
                                     float grad_wp, float alpha, void *stream);

/* out[n][j] = sum_i m[n][i][j] * g[n][i], i ascending: the transpose of
 * pde_matvec_f32 in its vector argument (the convection term's gradient). */
int pde_matvec_t_f32(const float *m, const float *g, float *out, int64_t nodes, int32_t rows,
                     int32_t cols, void *stream);

/* --------------------------------------------------------- immersed boundary
 *
 * Staircase immersed boundary (reference pm/FDM/boundary/immersedBoundary.py:
 * create_staircase_boundary_kernel :236-300 and create_fill_solids_inplace
 * :214-233, launched by ImmersedBoundary.forward :177-211).  One record per
 * (solid boundary node, component), resolved on the host from the bitmask and
 * the (node, component) type/value tables:
 *   kind 1 (Dirichlet): out[node][comp] = value
 *   kind 2 (Neumann):   running mean, over the fluid neighbours in axis order
 *                       (eligible axes ascending, side -1 then +1), of
 *                       in[node + s*stride[d]][comp] - (s * value) * dx:
 *                       avg += (c - avg) / n; n += 1, starting from avg = 0,
 *                       n = 1.  Bit 2d + (s > 0) of `mask` marks a fluid
 *                       neighbour.
 * Then every interior solid node (all components) takes fill_value.  Reads
 * touch fluid nodes only and writes solid nodes only, so in == out (in place)
 * is allowed. */
typedef struct pde_ib_op {
    int64_t node;     /* flat padded node index */
    float value;
    int32_t comp;
    int32_t kind;     /* 1 Dirichlet, 2 Neumann, anything else: no write */
    int32_t mask;     /* fluid neighbours, bit 2*d + (side > 0) */
} pde_ib_op;

int pde_immersed_apply_f32(const float *in, float *out, const pde_ib_op *ops, int64_t nops,
                           const int64_t *fill_nodes, int64_t nfill, float fill_value,
                           int32_t ncomp, int32_t ndim, int64_t stride0, int64_t stride1,
                           int64_t stride2, float dx, void *stream);

/* ------------------------------------------------------------ absorbing layers
 *
 * Sponge (far-field relaxation) and viscous damping layers over the nodes
 * within num_layers of an eligible face (reference
 * pm/FDM/kernel/farfield.py:8-78 and viscousDampingLayer.py:8-76, launched by
 * pm/FDM/functional/_farfield.py:5-41 and _viscousDampingLayer.py:5-38).  The
 * host resolves each layer node's distance level d (1..num_layers, measured
 * along the axis of the nearest face, lowest axis on ties) and the profile
 * prof[d-1] = fp32((fp32(d) / fp32(num_layers)) ** p), the power taken in FP64
 * and rounded once. */
typedef struct pde_layer_point {
    int64_t node;     /* flat padded node index */
    int32_t level;    /* 1..nlevels */
    int32_t reserved;
} pde_layer_point;

#define PDE_LAYER_MAXC 16   /* most floats per element (vec16 / mat4x4) */

/* For every point: f = scale * prof[level-1];
 *   out[node][c] = -f * (in[node][c] - farfield[c])   (farfield != NULL, FarField)
 *   out[node][c] = -f * in[node][c]                   (farfield == NULL,
 *                                                      ViscousDampingLayer)
 * farfield is a HOST array of ncomp floats, passed by value to the kernel (so
 * a captured graph keeps it).  Nodes not in the list are not written.  Every
 * level must lie in 1..nlevels. */
int pde_damping_layer_f32(const float *in, float *out, const pde_layer_point *pts, int64_t npts,
                          const float *profile, int32_t nlevels, int32_t ncomp,
                          const float *farfield, float scale, void *stream);

/* ------------------------------------------------------------------- library */

/* Library version (major*10000 + minor*100 + patch). */
int pde_version(void);

/* Human-readable name of a status code returned by any entry point. */
const char *pde_status_string(int status);

#ifdef __cplusplus
}
#endif

#endif /* PDE_HIP_H */
