/*
 * ORACLE — test infrastructure only.  Nothing in the shipped package links,
 * loads or calls this file; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker / CPU baseline.
 *
 * A plain-C FP32 restatement of PDE_Module's uniform-grid finite-difference
 * kernels.  Every function follows the reference's arithmetic in the same
 * operation order (one rounding per + and per *, no fused multiply-add: this
 * file must be compiled with -ffp-contract=off and without -ffast-math), so that
 * a GPU kernel that keeps the same order can be compared bit for bit.
 *
 * Layout (reference pm/mesh/mesh_functions.py:108-159, pm/utils/utils.py:35-52):
 *   a field is a C-order array of padded shape (X, Y, Z) whose element is a
 *   vector of C floats (AoS, component innermost) or an N x D row-major matrix.
 *   Eligible axes are those of size > 1, ascending (pm/FDM/utils.py:6-27).
 *   Operators write interior nodes only: index in [g, P-g) on eligible axes
 *   (pm/FDM/functional/_laplacian.py:26-35); output ghosts are left untouched.
 *
 * Each public function takes `nthreads` (1 = the serial reference-equivalent
 * CPU path, >1 = OpenMP over the slowest axis).  Results do not depend on it.
 */
#include <stdint.h>
#include <string.h>
#include <math.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int64_t P[3];   /* padded extents */
    int64_t lo[3];  /* interior start per axis */
    int64_t hi[3];  /* interior end (exclusive) per axis */
    int64_t sn[3];  /* node strides per axis */
    int elig[3];    /* axis has size > 1 */
    int ndim;       /* number of eligible axes */
    int ax[3];      /* eligible axes, ascending */
} geom_t;

static geom_t make_geom(int64_t X, int64_t Y, int64_t Z, int g)
{
    geom_t G;
    G.P[0] = X; G.P[1] = Y; G.P[2] = Z;
    G.sn[0] = Y * Z; G.sn[1] = Z; G.sn[2] = 1;
    G.ndim = 0;
    for (int a = 0; a < 3; ++a) {
        G.elig[a] = G.P[a] > 1;
        G.lo[a] = G.elig[a] ? g : 0;
        G.hi[a] = G.elig[a] ? G.P[a] - g : G.P[a];
        if (G.elig[a]) G.ax[G.ndim++] = a;
    }
    return G;
}

/* Laplacian — reference pm/FDM/kernel/laplacian.py:34-59:
 *   acc = 0; for each eligible axis (ascending), for s in (-1, 0, +1):
 *   acc += w[s] * in[n + s e_axis]   (componentwise);  out[n] = alpha * acc.
 * The centre weight is applied once per axis (kernel/laplacian.py:40-45).   */
void orc_laplacian_f32(const float *in, float *out, int64_t X, int64_t Y, int64_t Z,
                       int C, int g, float wm, float wc, float wp, float alpha,
                       int nthreads)
{
    geom_t G = make_geom(X, Y, Z, g);
    int64_t se[3];
    for (int a = 0; a < 3; ++a) se[a] = G.sn[a] * C;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = G.lo[0]; i < G.hi[0]; ++i)
        for (int64_t j = G.lo[1]; j < G.hi[1]; ++j)
            for (int64_t k = G.lo[2]; k < G.hi[2]; ++k)
                for (int c = 0; c < C; ++c) {
                    int64_t e = ((i * Y + j) * Z + k) * C + c;
                    float acc = 0.0f;
                    for (int t = 0; t < G.ndim; ++t) {
                        int64_t s = se[G.ax[t]];
                        acc += wm * in[e - s];
                        acc += wc * in[e];
                        acc += wp * in[e + s];
                    }
                    out[e] = alpha * acc;
                }
}

/* Grad — reference pm/FDM/kernel/grad.py:78-100.  Input vec N.  Output slot
 * (r, t) of an N x D matrix (or slot t of a D-vector when N == 1 and the
 * output is a vector; same memory layout).  Each slot starts from 0 and sums
 * its own axis's three terms in s order, then is scaled by alpha.  The slot is
 * the eligible-axis ordinal t, which equals the reference's grid-axis slot for
 * the leading-axis layouts the reference supports (grad.py:87-96).          */
void orc_grad_f32(const float *in, float *out, int64_t X, int64_t Y, int64_t Z,
                  int N, int g, float wm, float wc, float wp, float alpha, int nthreads)
{
    geom_t G = make_geom(X, Y, Z, g);
    const int D = G.ndim;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = G.lo[0]; i < G.hi[0]; ++i)
        for (int64_t j = G.lo[1]; j < G.hi[1]; ++j)
            for (int64_t k = G.lo[2]; k < G.hi[2]; ++k) {
                int64_t n = (i * Y + j) * Z + k;
                for (int r = 0; r < N; ++r)
                    for (int t = 0; t < D; ++t) {
                        int64_t s = G.sn[G.ax[t]];
                        float acc = 0.0f;
                        acc += wm * in[(n - s) * N + r];
                        acc += wc * in[n * N + r];
                        acc += wp * in[(n + s) * N + r];
                        out[(n * N + r) * D + t] = alpha * acc;
                    }
            }
}

/* Divergence — reference pm/FDM/kernel/divergence.py:85-106.
 * matrix_in == 0: input vec D, output vec 1: acc over (axis t asc, s asc) of
 *   w[s] * in[n + s e_t][t].
 * matrix_in == 1: input N x D matrix, output vec N: per row r the same sum of
 *   w[s] * M[n + s e_t][r][t].
 * Then alpha * acc.                                                           */
void orc_divergence_f32(const float *in, float *out, int64_t X, int64_t Y, int64_t Z,
                        int N, int matrix_in, int g, float wm, float wc, float wp,
                        float alpha, int nthreads)
{
    geom_t G = make_geom(X, Y, Z, g);
    const int D = G.ndim;
    const int rows = matrix_in ? N : 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = G.lo[0]; i < G.hi[0]; ++i)
        for (int64_t j = G.lo[1]; j < G.hi[1]; ++j)
            for (int64_t k = G.lo[2]; k < G.hi[2]; ++k) {
                int64_t n = (i * Y + j) * Z + k;
                for (int r = 0; r < rows; ++r) {
                    float acc = 0.0f;
                    for (int t = 0; t < D; ++t) {
                        int64_t s = G.sn[G.ax[t]];
                        int64_t off = (int64_t)r * D + t;
                        acc += wm * in[(n - s) * rows * D + off];
                        acc += wc * in[n * rows * D + off];
                        acc += wp * in[(n + s) * rows * D + off];
                    }
                    out[n * rows + r] = alpha * acc;
                }
            }
}

/* ForwardEuler — reference pm/time_step/forwardEuler.py:55-60,79:
 * new[i] = cur[i] + dt * k[i] over every scalar of the flattened array.     */
void orc_forward_euler_f32(const float *cur, const float *k, float *out, int64_t n,
                           float dt, int nthreads)
{
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; ++i) out[i] = cur[i] + dt * k[i];
}

/* The u/v wave step on a boundary-applied u2 — the reference loop
 * examples/wave2d/standingWave.py:72-76 (v += dt * c2 * lap(u2); u = u2 + dt * v,
 * the first step reading the old v), as np_oracle.wave_step_core fuses it: on
 * interior nodes acc = the Laplacian sum (axes ascending, s = -1, 0, +1),
 * rhs = alpha * acc, times c2[n] when c2 is given, vn = v + dt * rhs,
 * un = u2 + dt * (first ? v : vn); elsewhere vn = v and un = u2.  Scalar
 * fields (one component).                                                    */
void orc_wave_step_f32(const float *u2, const float *v, const float *c2, float *un, float *vn,
                       int64_t X, int64_t Y, int64_t Z, int g, float wm, float wc, float wp,
                       float alpha, float dt, int first, int nthreads)
{
    geom_t G = make_geom(X, Y, Z, g);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < X; ++i)
        for (int64_t j = 0; j < Y; ++j)
            for (int64_t k = 0; k < Z; ++k) {
                const int64_t e = (i * Y + j) * Z + k;
                const int in = i >= G.lo[0] && i < G.hi[0] && j >= G.lo[1] && j < G.hi[1] &&
                               k >= G.lo[2] && k < G.hi[2];
                if (!in) {
                    vn[e] = v[e];
                    un[e] = u2[e];
                    continue;
                }
                float acc = 0.0f;
                for (int t = 0; t < G.ndim; ++t) {
                    const int64_t s = G.sn[G.ax[t]];
                    acc += wm * u2[e - s];
                    acc += wc * u2[e];
                    acc += wp * u2[e + s];
                }
                float rhs = alpha * acc;
                if (c2) rhs = c2[e] * rhs;
                const float vv = v[e] + dt * rhs;
                vn[e] = vv;
                un[e] = u2[e] + dt * (first ? v[e] : vv);
            }
}

static inline float neg_sign_times_two(int s) { return (s > 0) ? -2.0f : 2.0f; }

/* Legacy GridBoundary — reference pm/deprecated/gridBoundary.py:193-196 and
 * the table kernel :267-302.  out = copy(in); then for each boundary entry
 * (b, c) and each axis with interior direction s != 0, ghost = b - s e_a and
 * adj = b + s e_a:
 *   type 1 (Dirichlet): out[b][c] = v;  out[ghost][c] = 2 v - in[adj][c]
 *   type 2 (Neumann):   out[ghost][c] = ((-sign(s) * 2) * dx) * v + in[adj][c]
 * All reads come from `in`.  ijk/interior are nb x 3 int32; type is nb x C
 * int8; value is nb x C float32.                                              */
void orc_bc_legacy_f32(const float *in, float *out, int64_t X, int64_t Y, int64_t Z,
                       int C, const int32_t *ijk, const int32_t *interior,
                       const int8_t *type, const float *value, int64_t nb, float dx)
{
    const int64_t sn[3] = {Y * Z, Z, 1};
    memcpy(out, in, (size_t)(X * Y * Z * C) * sizeof(float));
    for (int64_t b = 0; b < nb; ++b) {
        const int64_t node = ijk[3 * b] * sn[0] + ijk[3 * b + 1] * sn[1] + ijk[3 * b + 2];
        for (int c = 0; c < C; ++c) {
            const float v = value[b * C + c];
            const int8_t ty = type[b * C + c];
            for (int a = 0; a < 3; ++a) {
                const int s = interior[3 * b + a];
                if (s == 0) continue;
                const int64_t ghost = node - s * sn[a];
                const int64_t adj = node + s * sn[a];
                if (ty == 1) {
                    out[node * C + c] = v;
                    out[ghost * C + c] = 2.0f * v - in[adj * C + c];
                } else if (ty == 2) {
                    out[ghost * C + c] = neg_sign_times_two(s) * dx * v + in[adj * C + c];
                }
            }
        }
    }
}

/* Functional GridBoundary — reference pm/FDM/boundary/functional_boundary.py:
 * copy (:260-263), then groups in insertion order (:278-291), each applying
 * the kernel at :356-400 to its flat C-order node ids:
 *   val = in[node] with the selected components replaced by the group value
 *     (constant: fp32(const), boundary_functions.py:42-44; ramp:
 *      min(t / tmax, 1) * const, boundary_functions.py:85-89);
 *   Dirichlet: out[node] = val (the whole vector);
 *   for each eligible axis with node[a] == g (s = +1) or == P-1-g (s = -1):
 *     for each selected component c (in the order given):
 *       Dirichlet: out[ghost][c] = 2 val[c] - in[adj][c]
 *       Neumann:   out[ghost][c] = ((-sign(s) * 2) * dx) * val[c] + in[adj][c]
 * Group gi owns ids[off[gi] .. off[gi+1]) and comps[coff[gi] .. coff[gi+1]).
 * gtmax[gi] <= 0 marks a constant group, > 0 a ramp with that T_max.         */
void orc_bc_functional_f32(const float *in, float *out, int64_t X, int64_t Y, int64_t Z,
                           int C, int g, int ngroups, const int64_t *off,
                           const int32_t *ids, const int64_t *coff, const int32_t *comps,
                           const int8_t *gtype, const float *gconst, const float *gtmax,
                           float t, float dx)
{
    geom_t G = make_geom(X, Y, Z, g);
    memcpy(out, in, (size_t)(X * Y * Z * C) * sizeof(float));
    float val[64];
    for (int gi = 0; gi < ngroups; ++gi) {
        float v = gconst[gi];
        if (gtmax[gi] > 0.0f) {
            float ramp = t / gtmax[gi];
            ramp = (ramp < 1.0f) ? ramp : 1.0f;   /* wp.min(t/T_max, 1) */
            v = ramp * gconst[gi];
        }
        for (int64_t q = off[gi]; q < off[gi + 1]; ++q) {
            const int64_t node = ids[q];
            int64_t idx[3];
            idx[0] = node / (Y * Z);
            idx[1] = (node / Z) % Y;
            idx[2] = node % Z;
            for (int c = 0; c < C; ++c) val[c] = in[node * C + c];
            for (int64_t m = coff[gi]; m < coff[gi + 1]; ++m) val[comps[m]] = v;
            if (gtype[gi] == 1)
                for (int c = 0; c < C; ++c) out[node * C + c] = val[c];
            for (int t2 = 0; t2 < G.ndim; ++t2) {
                const int a = G.ax[t2];
                int s = 0;
                if (idx[a] == g) s = 1;
                else if (idx[a] == G.P[a] - 1 - g) s = -1;
                if (s == 0) continue;
                const int64_t ghost = node - s * G.sn[a];
                const int64_t adj = node + s * G.sn[a];
                for (int64_t m = coff[gi]; m < coff[gi + 1]; ++m) {
                    const int c = comps[m];
                    float nv;
                    if (gtype[gi] == 1) nv = 2.0f * val[c] - in[adj * C + c];
                    else nv = neg_sign_times_two(s) * dx * val[c] + in[adj * C + c];
                    out[ghost * C + c] = nv;
                }
            }
        }
    }
}

/* Periodic ghost fill (no counterpart in the reference, which has no periodic
 * boundary): in place, for each eligible axis in ascending order, ghost layer
 * i < g takes layer i + E and ghost layer i >= P - g takes layer i - E, where
 * E = P - 2g, over the full extent of the other axes (so edges and corners
 * are filled by the later axes).                                            */
void orc_periodic_fill_f32(float *u, int64_t X, int64_t Y, int64_t Z, int C, int g)
{
    geom_t G = make_geom(X, Y, Z, g);
    for (int t = 0; t < G.ndim; ++t) {
        const int a = G.ax[t];
        const int64_t E = G.P[a] - 2 * g;
        for (int64_t i = 0; i < X; ++i)
            for (int64_t j = 0; j < Y; ++j)
                for (int64_t k = 0; k < Z; ++k) {
                    int64_t idx[3] = {i, j, k};
                    int64_t src;
                    if (idx[a] < g) src = idx[a] + E;
                    else if (idx[a] >= G.P[a] - g) src = idx[a] - E;
                    else continue;
                    int64_t n = (i * Y + j) * Z + k;
                    int64_t m = n + (src - idx[a]) * G.sn[a];
                    for (int c = 0; c < C; ++c) u[n * C + c] = u[m * C + c];
                }
    }
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
