/* SPDX-License-Identifier: Apache-2.0
 * Test oracle: a restatement in C of OpenFAWE's src/openfawe/aero/vortex_numba.py and vortex_tree.py (Copyright
 * 2026 Yinong Tian, Apache License 2.0, itself adapted from OpenFAST/OLAF; see NOTICE). */
/*
 * CPU oracle for the free-wake induced-velocity path.  TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C, the reference algorithm of OpenFAWE's
 * Biot-Savart segment sum (direct and tree).  It exists to check the HIP
 * product: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * may load it, and only to check results or as the timed CPU baseline.  The product
 * library (openfawe_amd/libofw_hip.so) never links or calls it.
 *
 * Every function cites the reference file:line it follows.  Arithmetic is
 * written operation by operation in the reference's order and the file is
 * compiled with -ffp-contract=off, so FP64 results follow the reference's
 * IEEE operation sequence (libm hypot/expm1 as on the reference's CPU path).
 *
 * Parity pinning (what checks this restatement, see tests/test_oracle_*.py):
 *   pair_velocity      closed forms K01/K02, the Lamb-Oseen cutoff boundary,
 *                      the recorded welib fixture (tests/golden), 96-point
 *                      Gauss-Legendre quadrature, invariances, and A1 against
 *                      an independent vectorised NumPy restatement.
 *   direct_sum         A1 against the NumPy restatement; thread-count bitwise.
 *   exclusion_csr      brute force (sorted unique pairs) on random inputs.
 *   build_tree*        serial DFS build == level build + renumber (the
 *                      reference's own T1 check), a pure-Python serial build,
 *                      and structural invariants (permutation, stable sort on
 *                      the longest axis, (a+b)//2 cut, numbering rule).
 *   node_moments       single-segment closed forms, closed ring M0 = 0, the
 *                      parallel-axis identity between parent and children.
 *   far_velocity       equals multipole_velocity_explicit (rtol 1e-12), ring
 *                      dipole limits, convergence to the direct sum.
 *   tree_velocity      theta -> 0 equals the direct sum, exclusions never leak.
 *   pair_matrix /
 *   assemble_columns   bitwise against per-column direct sums.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define FLAG_OK 0
#define FLAG_SINGULAR 1
#define FLAG_NONFINITE 2

/* vortex_numba.py:28-30 */
static double AXIS_EPS;
static double LAMB_CUT;
static double FOUR_PI;
static double INV_FOUR_PI; /* vortex_tree.py:47 */

static void init_constants(void) {
    AXIS_EPS = 8.0 * DBL_EPSILON;
    LAMB_CUT = sqrt(10.0 / 1.25643);
    FOUR_PI = 4.0 * M_PI;
    INV_FOUR_PI = 1.0 / (4.0 * M_PI);
}

int ofo_abi_version(void) { init_constants(); return 1; }

static void set_threads(int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}

/* vortex_numba.py:37-39 */
static inline double norm3(double x, double y, double z) { return hypot(hypot(x, y), z); }

/* vortex_numba.py:42-109 (pair_velocity).  Returns the flag; writes u[3]. */
int ofo_pair_velocity(double px, double py, double pz, double sx, double sy, double sz,
                      double dx, double dy, double dz, double length, double gamma,
                      int model, double radius, int raise_singular, double *u) {
    if (AXIS_EPS == 0.0) init_constants();
    u[0] = u[1] = u[2] = 0.0;
    double ax = (px - sx) / length;
    double ay = (py - sy) / length;
    double az = (pz - sz) / length;
    double t = (ax * dx + ay * dy) + az * dz;
    double cx = dy * az - dz * ay;
    double cy = dz * ax - dx * az;
    double cz = dx * ay - dy * ax;
    double h = norm3(cx, cy, cz);
    double na = norm3(ax, ay, az);
    double nb = norm3(ax - dx, ay - dy, az - dz);
    if (!(isfinite(h) && isfinite(na) && isfinite(nb) && isfinite(t))) return FLAG_NONFINITE;
    double big = na > nb ? na : nb;           /* max(na, nb) */
    double scale = 1.0 > big ? 1.0 : big;     /* max(1.0, ...) */
    if (h <= AXIS_EPS * scale) {
        if (raise_singular && t >= 0.0 && t <= 1.0) return FLAG_SINGULAR;
        return FLAG_OK;
    }
    if (gamma == 0.0) return FLAG_OK;
    double product = na * nb;
    double dot = h * h + t * (t - 1.0);
    double plus = product + dot;
    if (dot < 0.0) plus = (h * h) / (product - dot);
    double denominator = product * plus;
    double rc = radius / length;
    double factor = 1.0;
    if (radius > 0.0 && model != 0) {
        if (model == 5) { /* Rosenhead-Moore, vortex_numba.py:74-83 */
            na = hypot(na, rc);
            nb = hypot(nb, rc);
            product = na * nb;
            double transverse2 = h * h + rc * rc;
            dot = transverse2 + t * (t - 1.0);
            plus = product + dot;
            if (dot < 0.0) plus = transverse2 / (product - dot);
            denominator = product * plus;
        } else if (model == 4) { /* offset, :84-85 */
            denominator = denominator + rc * rc;
        } else {
            double ratio = h / rc;
            if (!isfinite(ratio)) return FLAG_NONFINITE; /* :88-91 */
            if (model == 1) {                            /* Rankine :92-94 */
                double limited = ratio < 1.0 ? ratio : 1.0;
                factor = limited * limited;
            } else if (model == 2) {                     /* Lamb-Oseen :95-97 */
                if (ratio <= LAMB_CUT) factor = -expm1(-1.25643 * (ratio * ratio));
            } else if (model == 3) {                     /* Vatistas n=2 :98-104 */
                if (ratio < 1.0) {
                    double q = ratio * ratio;
                    factor = q / hypot(1.0, q);
                } else {
                    double inverse = 1.0 / ratio;
                    factor = 1.0 / hypot(1.0, inverse * inverse);
                }
            }
        }
    }
    /* :105, evaluated left to right as Python does */
    double coefficient = (gamma / length) / FOUR_PI * factor * (na + nb) / denominator;
    if (!(isfinite(denominator) && isfinite(coefficient) && isfinite(product) && isfinite(dot)))
        return FLAG_NONFINITE;
    u[0] = coefficient * cx;
    u[1] = coefficient * cy;
    u[2] = coefficient * cz;
    return FLAG_OK;
}

/* Batch helper for tests: one pair per row. */
void ofo_pair_values(int64_t k, const double *p, const double *s, const double *d, const double *len,
                     const double *gamma, int model, const double *radius, int raise_singular,
                     double *out, int8_t *flags) {
    for (int64_t i = 0; i < k; i++)
        flags[i] = (int8_t)ofo_pair_velocity(p[3 * i], p[3 * i + 1], p[3 * i + 2], s[3 * i], s[3 * i + 1],
                                             s[3 * i + 2], d[3 * i], d[3 * i + 1], d[3 * i + 2], len[i],
                                             gamma[i], model, radius[i], raise_singular, out + 3 * i);
}

/* vortex_numba.py:112-147 (_sum_point) driven by :159-174 (parallel over targets). */
void ofo_direct_sum(int64_t n, const double *points, int64_t m, const double *starts,
                    const double *directions, const double *lengths, const double *gamma,
                    const double *radius, int model, int raise_singular, const int64_t *ex_ptr,
                    const int64_t *ex_seg, double *out, int8_t *flags, double *scale, int threads) {
    if (AXIS_EPS == 0.0) init_constants();
    set_threads(threads);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t p = 0; p < n; p++) {
        double px = points[3 * p], py = points[3 * p + 1], pz = points[3 * p + 2];
        double ux = 0.0, uy = 0.0, uz = 0.0, total = 0.0;
        int flag = FLAG_OK;
        int64_t k = ex_ptr[p], stop = ex_ptr[p + 1];
        for (int64_t s = 0; s < m; s++) {
            if (k < stop && ex_seg[k] == s) { k++; continue; }
            double v[3];
            int f = ofo_pair_velocity(px, py, pz, starts[3 * s], starts[3 * s + 1], starts[3 * s + 2],
                                      directions[3 * s], directions[3 * s + 1], directions[3 * s + 2],
                                      lengths[s], gamma[s], model, radius[s], raise_singular, v);
            if (f != FLAG_OK) { if (flag == FLAG_OK) flag = f; continue; }
            ux += v[0]; uy += v[1]; uz += v[2];
            if (scale) total += sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        }
        if (!(isfinite(ux) && isfinite(uy) && isfinite(uz)) && flag == FLAG_OK) flag = FLAG_NONFINITE;
        out[3 * p] = ux; out[3 * p + 1] = uy; out[3 * p + 2] = uz;
        flags[p] = (int8_t)flag;
        if (scale) scale[p] = total; /* sum_s |u_s|: the A1 normaliser (test_vortex_numba.py:21-35) */
    }
}

/* vortex_numba.py:193-225 (_exclusion_csr_compiled).  pointer has point_count+1 entries, segs_out
 * has room for e entries; returns the number of unique pairs. */
int64_t ofo_exclusion_csr(int64_t point_count, int64_t e, const int64_t *pts, const int64_t *segs,
                          int64_t *pointer, int64_t *segs_out) {
    int64_t *start = (int64_t *)calloc((size_t)point_count + 1, sizeof(int64_t));
    int64_t *cursor = (int64_t *)malloc(((size_t)point_count + 1) * sizeof(int64_t));
    int64_t *buffer = (int64_t *)malloc(((size_t)e + 1) * sizeof(int64_t));
    for (int64_t k = 0; k < e; k++) start[pts[k] + 1] += 1;
    for (int64_t p = 0; p < point_count; p++) start[p + 1] += start[p];
    for (int64_t p = 0; p < point_count; p++) cursor[p] = start[p];
    for (int64_t k = 0; k < e; k++) { int64_t p = pts[k]; buffer[cursor[p]] = segs[k]; cursor[p]++; }
    pointer[0] = 0;
    int64_t mm = 0;
    for (int64_t p = 0; p < point_count; p++) {
        int64_t a = start[p], b = start[p + 1];
        for (int64_t i = a + 1; i < b; i++) {
            int64_t value = buffer[i], j = i - 1;
            while (j >= a && buffer[j] > value) { buffer[j + 1] = buffer[j]; j--; }
            buffer[j + 1] = value;
        }
        for (int64_t i = a; i < b; i++)
            if (i == a || buffer[i] != buffer[i - 1]) buffer[mm++] = buffer[i];
        pointer[p + 1] = mm;
    }
    memcpy(segs_out, buffer, (size_t)mm * sizeof(int64_t));
    free(start); free(cursor); free(buffer);
    return mm;
}

/* ---------------------------------------------------------------- tree build */

/* Stable merge sort of idx[0..n) by keys[idx[.]] -- the permutation that
 * np.argsort(kind="mergesort") returns for finite keys (ties keep input order,
 * -0.0 == 0.0 under C comparison as under NumPy's). */
static void stable_argsort(const double *keys, int64_t n, int64_t *perm, int64_t *tmp) {
    for (int64_t i = 0; i < n; i++) perm[i] = i;
    for (int64_t w = 1; w < n; w *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            int64_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            int64_t i = lo, j = mid, o = lo;
            while (i < mid && j < hi) {
                if (keys[perm[j]] < keys[perm[i]]) tmp[o++] = perm[j++];
                else tmp[o++] = perm[i++];
            }
            while (i < mid) tmp[o++] = perm[i++];
            while (j < hi) tmp[o++] = perm[j++];
        }
        memcpy(perm, tmp, (size_t)n * sizeof(int64_t));
    }
}

/* The per-node split step shared by vortex_tree.py:73-101 (serial) and :177-203 (_split_level):
 * returns 1 and re-orders order[a:b) when the node splits, 0 otherwise. */
static int split_node(const double *mid, int64_t *order, int64_t a, int64_t b, int64_t leaf_size,
                      int *axis_out, int64_t *scratch) {
    if (b - a <= leaf_size) return 0;
    double low[3] = {INFINITY, INFINITY, INFINITY}, high[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t k = a; k < b; k++) {
        int64_t s = order[k];
        for (int axis = 0; axis < 3; axis++) {
            double v = mid[3 * s + axis];
            if (v < low[axis]) low[axis] = v;
            if (v > high[axis]) high[axis] = v;
        }
    }
    int axis = 0;
    double extent = high[0] - low[0];
    for (int c = 1; c < 3; c++)
        if (high[c] - low[c] > extent) { extent = high[c] - low[c]; axis = c; }
    if (!(extent > 0.0)) return 0; /* coincident midpoints stay one leaf */
    int64_t len = b - a;
    int64_t *indices = scratch, *perm = scratch + len, *tmp = scratch + 2 * len;
    double *keys = (double *)(scratch + 3 * len);
    for (int64_t k = 0; k < len; k++) { indices[k] = order[a + k]; keys[k] = mid[3 * indices[k] + axis]; }
    stable_argsort(keys, len, perm, tmp);
    for (int64_t k = 0; k < len; k++) order[a + k] = indices[perm[k]];
    *axis_out = axis;
    return 1;
}

int64_t ofo_tree_capacity(int64_t n, int64_t leaf_size) {
    return 4 * (n / (leaf_size > 1 ? leaf_size : 1) + 1) + 1; /* vortex_tree.py:59, :224 */
}

/* vortex_tree.py:51-115 (build_tree_serial).  Arrays sized by ofo_tree_capacity.  Returns count. */
int64_t ofo_build_tree_serial(int64_t n, const double *mid, int64_t leaf_size, int64_t *order,
                              int64_t *lo, int64_t *hi, int64_t *left, int64_t *right) {
    int64_t capacity = ofo_tree_capacity(n, leaf_size);
    for (int64_t i = 0; i < n; i++) order[i] = i;
    for (int64_t i = 0; i < capacity; i++) { left[i] = -1; right[i] = -1; }
    int64_t *stack = (int64_t *)malloc((size_t)capacity * sizeof(int64_t));
    int64_t *scratch = (int64_t *)malloc(((size_t)4 * n + 4) * sizeof(int64_t));
    lo[0] = 0; hi[0] = n;
    int64_t count = 1, top = 1;
    stack[0] = 0;
    while (top > 0) {
        top--;
        int64_t node = stack[top], a = lo[node], b = hi[node];
        int axis;
        if (!split_node(mid, order, a, b, leaf_size, &axis, scratch)) continue;
        int64_t middle = (a + b) / 2; /* (a+b)//2, nonnegative */
        int64_t first = count, second = count + 1;
        count += 2;
        lo[first] = a; hi[first] = middle;
        lo[second] = middle; hi[second] = b;
        left[node] = first; right[node] = second;
        stack[top] = second; stack[top + 1] = first;
        top += 2;
    }
    free(stack); free(scratch);
    return count;
}

/* vortex_tree.py:271-306 (_renumber). */
static void renumber(const int64_t *lo, const int64_t *hi, const int64_t *left, const int64_t *right,
                     int64_t count, int64_t *out_lo, int64_t *out_hi, int64_t *out_left, int64_t *out_right) {
    int64_t *number = (int64_t *)malloc((size_t)count * sizeof(int64_t));
    int64_t *stack = (int64_t *)malloc((size_t)count * sizeof(int64_t));
    for (int64_t i = 0; i < count; i++) { out_left[i] = -1; out_right[i] = -1; }
    number[0] = 0; out_lo[0] = lo[0]; out_hi[0] = hi[0];
    int64_t assigned = 1, top = 1;
    stack[0] = 0;
    while (top > 0) {
        top--;
        int64_t node = stack[top];
        if (left[node] < 0) continue;
        int64_t first = assigned, second = assigned + 1;
        assigned += 2;
        number[left[node]] = first; number[right[node]] = second;
        out_left[number[node]] = first; out_right[number[node]] = second;
        out_lo[first] = lo[left[node]]; out_hi[first] = hi[left[node]];
        out_lo[second] = lo[right[node]]; out_hi[second] = hi[right[node]];
        stack[top] = right[node]; stack[top + 1] = left[node];
        top += 2;
    }
    free(number); free(stack);
}

/* vortex_tree.py:208-268 (build_tree with _split_level in parallel over a level's nodes, _next_level,
 * then _renumber).  Arrays sized by ofo_tree_capacity.  Returns count. */
int64_t ofo_build_tree(int64_t n, const double *mid, int64_t leaf_size, int64_t *order, int64_t *out_lo,
                       int64_t *out_hi, int64_t *out_left, int64_t *out_right, int threads) {
    set_threads(threads);
    int64_t capacity = ofo_tree_capacity(n, leaf_size);
    int64_t *lo = (int64_t *)malloc((size_t)capacity * sizeof(int64_t));
    int64_t *hi = (int64_t *)malloc((size_t)capacity * sizeof(int64_t));
    int64_t *left = (int64_t *)malloc((size_t)capacity * sizeof(int64_t));
    int64_t *right = (int64_t *)malloc((size_t)capacity * sizeof(int64_t));
    int64_t *level = (int64_t *)malloc((size_t)capacity * sizeof(int64_t));
    int64_t *following = (int64_t *)malloc((size_t)capacity * sizeof(int64_t));
    char *split = (char *)malloc((size_t)capacity);
    for (int64_t i = 0; i < n; i++) order[i] = i;
    for (int64_t i = 0; i < capacity; i++) { left[i] = -1; right[i] = -1; }
    lo[0] = 0; hi[0] = n;
    int64_t count = 1, width = 1;
    level[0] = 0;
    while (width > 0) {
#pragma omp parallel
        {
            int64_t *scratch = (int64_t *)malloc(((size_t)4 * n + 4) * sizeof(int64_t));
#pragma omp for schedule(dynamic, 1)
            for (int64_t i = 0; i < width; i++) {
                int axis;
                split[i] = (char)split_node(mid, order, lo[level[i]], hi[level[i]], leaf_size, &axis, scratch);
            }
            free(scratch);
        }
        int64_t j = 0;
        for (int64_t i = 0; i < width; i++) { /* _next_level, vortex_tree.py:244-268 */
            if (!split[i]) continue;
            int64_t node = level[i], a = lo[node], b = hi[node], middle = (a + b) / 2;
            int64_t first = count, second = count + 1;
            count += 2;
            lo[first] = a; hi[first] = middle; lo[second] = middle; hi[second] = b;
            left[node] = first; right[node] = second;
            following[j] = first; following[j + 1] = second;
            j += 2;
        }
        memcpy(level, following, (size_t)j * sizeof(int64_t));
        width = j;
    }
    renumber(lo, hi, left, right, count, out_lo, out_hi, out_left, out_right);
    free(lo); free(hi); free(left); free(right); free(level); free(following); free(split);
    return count;
}

/* ------------------------------------------------------------ node moments */

/* vortex_tree.py:118-167 (node_moments_serial; parallel over nodes as :320-376, which is bitwise equal).
 * centre (c,3), bound (c), core (c), m0 (c,3), m1 (c,3,3), m2 (c,3,3,3). */
void ofo_node_moments(int64_t count, const int64_t *order, const int64_t *lo, const int64_t *hi,
                      const double *starts, const double *ends, const double *alpha,
                      const double *length_vectors, const double *radius, const double *gamma,
                      double *centre, double *bound, double *core, double *m0, double *m1, double *m2,
                      int threads) {
    set_threads(threads);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t node = 0; node < count; node++) {
        double low[3] = {INFINITY, INFINITY, INFINITY}, high[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int64_t k = lo[node]; k < hi[node]; k++) {
            int64_t s = order[k];
            for (int axis = 0; axis < 3; axis++) {
                double vals[2] = {starts[3 * s + axis], ends[3 * s + axis]};
                for (int q = 0; q < 2; q++) {
                    if (vals[q] < low[axis]) low[axis] = vals[q];
                    if (vals[q] > high[axis]) high[axis] = vals[q];
                }
            }
        }
        double *c = centre + 3 * node;
        for (int axis = 0; axis < 3; axis++) c[axis] = 0.5 * (low[axis] + high[axis]);
        double *M0 = m0 + 3 * node, *M1 = m1 + 9 * node, *M2 = m2 + 27 * node;
        for (int i = 0; i < 3; i++) M0[i] = 0.0;
        for (int i = 0; i < 9; i++) M1[i] = 0.0;
        for (int i = 0; i < 27; i++) M2[i] = 0.0;
        double r2 = 0.0, largest = 0.0;
        for (int64_t k = lo[node]; k < hi[node]; k++) {
            int64_t s = order[k];
            for (int q = 0; q < 2; q++) {
                const double *point = q == 0 ? starts + 3 * s : ends + 3 * s;
                double e0 = point[0] - c[0], e1 = point[1] - c[1], e2 = point[2] - c[2];
                double qq = e0 * e0 + e1 * e1 + e2 * e2;
                if (qq > r2) r2 = qq;
            }
            if (radius[s] > largest) largest = radius[s];
            if (gamma[s] == 0.0) continue;
            double d[3];
            for (int axis = 0; axis < 3; axis++) d[axis] = 0.5 * (starts[3 * s + axis] + ends[3 * s + axis]) - c[axis];
            const double *al = alpha + 3 * s, *lv = length_vectors + 3 * s;
            for (int i = 0; i < 3; i++) {
                M0[i] += al[i];
                for (int j = 0; j < 3; j++) {
                    M1[3 * i + j] += al[i] * d[j];
                    for (int l = 0; l < 3; l++)
                        M2[9 * i + 3 * j + l] += al[i] * (d[j] * d[l] + lv[j] * lv[l] / 12.0);
                }
            }
        }
        bound[node] = sqrt(r2) * (1.0 + 1e-12);
        core[node] = largest;
    }
}

/* vortex_tree.py:425-447 (node_contractions). */
void ofo_node_contractions(int64_t count, const double *m1, const double *m2, double *w1, double *t2, double *w2) {
    for (int64_t node = 0; node < count; node++) {
        const double *M1 = m1 + 9 * node, *M2 = m2 + 27 * node;
        double *W1 = w1 + 3 * node, *T2 = t2 + 3 * node, *W2 = w2 + 9 * node;
#define A1(i, j) M1[3 * (i) + (j)]
#define A2(i, j, k) M2[9 * (i) + 3 * (j) + (k)]
        W1[0] = A1(1, 2) - A1(2, 1);
        W1[1] = A1(2, 0) - A1(0, 2);
        W1[2] = A1(0, 1) - A1(1, 0);
        for (int i = 0; i < 3; i++) T2[i] = A2(i, 0, 0) + A2(i, 1, 1) + A2(i, 2, 2);
        for (int k = 0; k < 3; k++) {
            W2[0 * 3 + k] = A2(1, 2, k) - A2(2, 1, k);
            W2[1 * 3 + k] = A2(2, 0, k) - A2(0, 2, k);
            W2[2 * 3 + k] = A2(0, 1, k) - A2(1, 0, k);
        }
#undef A1
#undef A2
    }
}

/* vortex_tree.py:379-383 */
static inline void cross_add(double scale, double a0, double a1, double a2, double b0, double b1, double b2,
                             double *out) {
    out[0] += scale * (a1 * b2 - a2 * b1);
    out[1] += scale * (a2 * b0 - a0 * b2);
    out[2] += scale * (a0 * b1 - a1 * b0);
}

/* vortex_tree.py:386-422 (multipole_velocity_explicit); adds into out. */
void ofo_multipole_explicit(double rx, double ry, double rz, const double *m0, const double *m1,
                            const double *m2, int order, double *out) {
    if (AXIS_EPS == 0.0) init_constants();
    double r[3] = {rx, ry, rz};
    double r2 = rx * rx + ry * ry + rz * rz;
    double distance = sqrt(r2);
    double inv3 = 1.0 / (r2 * distance);
    double acc[3] = {0.0, 0.0, 0.0};
    cross_add(inv3, m0[0], m0[1], m0[2], rx, ry, rz, acc);
    if (order >= 1) {
        double inv5 = inv3 / r2;
        for (int j = 0; j < 3; j++) {
            double g[3];
            for (int c = 0; c < 3; c++) g[c] = -3.0 * r[j] * r[c] * inv5;
            g[j] += inv3;
            cross_add(-1.0, m1[0 * 3 + j], m1[1 * 3 + j], m1[2 * 3 + j], g[0], g[1], g[2], acc);
        }
    }
    if (order >= 2) {
        double inv5 = inv3 / r2;
        double inv7 = inv5 / r2;
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                double g[3];
                for (int c = 0; c < 3; c++) {
                    g[c] = 15.0 * r[j] * r[k] * r[c] * inv7;
                    if (j == k) g[c] -= 3.0 * r[c] * inv5;
                }
                g[j] -= 3.0 * r[k] * inv5;
                g[k] -= 3.0 * r[j] * inv5;
                cross_add(0.5, m2[0 * 9 + 3 * j + k], m2[1 * 9 + 3 * j + k], m2[2 * 9 + 3 * j + k], g[0], g[1],
                          g[2], acc);
            }
    }
    out[0] += INV_FOUR_PI * acc[0];
    out[1] += INV_FOUR_PI * acc[1];
    out[2] += INV_FOUR_PI * acc[2];
}

/* vortex_tree.py:450-497 (far_velocity) for one node's moment blocks; writes u[3]. */
void ofo_far_velocity(double rx, double ry, double rz, const double *m0, const double *m1, const double *m2,
                      const double *w1, const double *t2, const double *w2, int order, double *u) {
    if (AXIS_EPS == 0.0) init_constants();
    double r2 = rx * rx + ry * ry + rz * rz;
    double inv3 = 1.0 / (r2 * sqrt(r2));
    double inv5 = inv3 / r2;
    double a0 = m0[0], a1 = m0[1], a2 = m0[2];
    double u0 = inv3 * (a1 * rz - a2 * ry);
    double u1 = inv3 * (a2 * rx - a0 * rz);
    double u2 = inv3 * (a0 * ry - a1 * rx);
    if (order >= 1) {
        double v0 = m1[0] * rx + m1[1] * ry + m1[2] * rz;
        double v1 = m1[3] * rx + m1[4] * ry + m1[5] * rz;
        double v2 = m1[6] * rx + m1[7] * ry + m1[8] * rz;
        u0 += 3.0 * inv5 * (v1 * rz - v2 * ry) - inv3 * w1[0];
        u1 += 3.0 * inv5 * (v2 * rx - v0 * rz) - inv3 * w1[1];
        u2 += 3.0 * inv5 * (v0 * ry - v1 * rx) - inv3 * w1[2];
    }
    if (order >= 2) {
        double inv7 = inv5 / r2;
        double q0 = 0.0, q1 = 0.0, q2 = 0.0;
        double r[3] = {rx, ry, rz};
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                double rr = r[j] * r[k];
                q0 += m2[0 * 9 + 3 * j + k] * rr;
                q1 += m2[1 * 9 + 3 * j + k] * rr;
                q2 += m2[2 * 9 + 3 * j + k] * rr;
            }
        double b0 = 7.5 * inv7 * q0 - 1.5 * inv5 * t2[0];
        double b1 = 7.5 * inv7 * q1 - 1.5 * inv5 * t2[1];
        double b2 = 7.5 * inv7 * q2 - 1.5 * inv5 * t2[2];
        double s0 = w2[0] * rx + w2[1] * ry + w2[2] * rz;
        double s1 = w2[3] * rx + w2[4] * ry + w2[5] * rz;
        double s2 = w2[6] * rx + w2[7] * ry + w2[8] * rz;
        u0 += (b1 * rz - b2 * ry) - 3.0 * inv5 * s0;
        u1 += (b2 * rx - b0 * rz) - 3.0 * inv5 * s1;
        u2 += (b0 * ry - b1 * rx) - 3.0 * inv5 * s2;
    }
    u[0] = INV_FOUR_PI * u0;
    u[1] = INV_FOUR_PI * u1;
    u[2] = INV_FOUR_PI * u2;
}

/* vortex_tree.py:571-580 (_positions): sorted tree positions of each point's excluded segments. */
void ofo_positions(int64_t point_count, const int64_t *ex_ptr, const int64_t *ex_segments,
                   const int64_t *inverse, int64_t *positions) {
    for (int64_t p = 0; p < point_count; p++) {
        int64_t a = ex_ptr[p], b = ex_ptr[p + 1];
        for (int64_t k = a; k < b; k++) positions[k] = inverse[ex_segments[k]];
        for (int64_t i = a + 1; i < b; i++) { /* insertion sort: a point has few exclusions */
            int64_t v = positions[i], j = i - 1;
            while (j >= a && positions[j] > v) { positions[j + 1] = positions[j]; j--; }
            positions[j + 1] = v;
        }
    }
}

#define STACK 512 /* vortex_tree.py:48 */

/* vortex_tree.py:508-568 (tree_velocity). */
void ofo_tree_velocity(int64_t n, const double *points, const double *starts, const double *directions,
                       const double *lengths, const double *gamma, const double *radius, int model,
                       int raise_singular, const int64_t *order, const int64_t *lo, const int64_t *hi,
                       const int64_t *left, const int64_t *right, const double *centre, const double *bound,
                       const double *core, const double *m0, const double *m1, const double *m2,
                       const double *w1, const double *t2, const double *w2, double theta, int expansion_order,
                       double clearance, const int64_t *ex_ptr, const int64_t *ex_pos, double *out,
                       int8_t *flags, int64_t *far_counts, int64_t *near_counts, int threads) {
    if (AXIS_EPS == 0.0) init_constants();
    set_threads(threads);
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t p = 0; p < n; p++) {
        double x0 = points[3 * p], x1 = points[3 * p + 1], x2 = points[3 * p + 2];
        double acc[3] = {0.0, 0.0, 0.0};
        int flag = FLAG_OK;
        int64_t far = 0, near = 0;
        int64_t e0 = ex_ptr[p], e1 = ex_ptr[p + 1];
        int64_t stack[STACK];
        stack[0] = 0;
        int top = 1;
        while (top > 0) {
            top--;
            int64_t node = stack[top];
            double rx = x0 - centre[3 * node], ry = x1 - centre[3 * node + 1], rz = x2 - centre[3 * node + 2];
            double distance = sqrt(rx * rx + ry * ry + rz * rz);
            int accept = bound[node] < theta * distance && distance - bound[node] >= clearance * core[node];
            if (accept) /* _excluded_in, :500-505 */
                for (int64_t k = e0; k < e1; k++)
                    if (ex_pos[k] >= lo[node] && ex_pos[k] < hi[node]) { accept = 0; break; }
            if (accept) {
                double f[3];
                ofo_far_velocity(rx, ry, rz, m0 + 3 * node, m1 + 9 * node, m2 + 27 * node, w1 + 3 * node,
                                 t2 + 3 * node, w2 + 9 * node, expansion_order, f);
                acc[0] += f[0]; acc[1] += f[1]; acc[2] += f[2];
                far++;
            } else if (left[node] < 0) {
                for (int64_t k = lo[node]; k < hi[node]; k++) {
                    int skip = 0;
                    for (int64_t e = e0; e < e1; e++)
                        if (ex_pos[e] == k) { skip = 1; break; }
                    if (skip) continue;
                    int64_t s = order[k];
                    double v[3];
                    near++;
                    int f = ofo_pair_velocity(x0, x1, x2, starts[3 * s], starts[3 * s + 1], starts[3 * s + 2],
                                              directions[3 * s], directions[3 * s + 1], directions[3 * s + 2],
                                              lengths[s], gamma[s], model, radius[s], raise_singular, v);
                    if (f != FLAG_OK) { if (flag == FLAG_OK) flag = f; continue; }
                    acc[0] += v[0]; acc[1] += v[1]; acc[2] += v[2];
                }
            } else {
                /* the reference does not check its 512-deep stack on the CPU; a balanced median tree
                 * needs about 2 entries per level, far below it */
                stack[top] = right[node];
                stack[top + 1] = left[node];
                top += 2;
            }
        }
        if (!(isfinite(acc[0]) && isfinite(acc[1]) && isfinite(acc[2])) && flag == FLAG_OK) flag = FLAG_NONFINITE;
        out[3 * p] = acc[0]; out[3 * p + 1] = acc[1]; out[3 * p + 2] = acc[2];
        flags[p] = (int8_t)flag;
        far_counts[p] = far;
        if (near_counts) near_counts[p] = near;
    }
}

/* vortex_tree.py:588-626 (tree_sum), after the model / theta checks that the caller makes.  The tree
 * arrays are returned to the caller (sized by ofo_tree_capacity) so tests can inspect them.
 * Returns the node count (0 for empty input). */
int64_t ofo_tree_sum(int64_t n, const double *points, int64_t m, const double *starts, const double *directions,
                     const double *lengths, const double *gamma, const double *radius, int model,
                     int raise_singular, const int64_t *ex_ptr, const int64_t *ex_seg, double theta,
                     int expansion_order, int64_t leaf_size, double clearance, double *out, int8_t *flags,
                     int64_t *far_counts, int64_t *near_counts, int threads) {
    if (AXIS_EPS == 0.0) init_constants();
    for (int64_t i = 0; i < 3 * n; i++) out[i] = 0.0;
    for (int64_t i = 0; i < n; i++) { flags[i] = 0; far_counts[i] = 0; if (near_counts) near_counts[i] = 0; }
    if (m == 0 || n == 0) return 0;
    double *lv = (double *)malloc((size_t)m * 3 * sizeof(double));
    double *ends = (double *)malloc((size_t)m * 3 * sizeof(double));
    double *alpha = (double *)malloc((size_t)m * 3 * sizeof(double));
    double *mid = (double *)malloc((size_t)m * 3 * sizeof(double));
    for (int64_t s = 0; s < m; s++)
        for (int c = 0; c < 3; c++) { /* vortex_tree.py:608-611 */
            lv[3 * s + c] = directions[3 * s + c] * lengths[s];
            ends[3 * s + c] = starts[3 * s + c] + lv[3 * s + c];
            alpha[3 * s + c] = gamma[s] * lv[3 * s + c];
            mid[3 * s + c] = 0.5 * (starts[3 * s + c] + ends[3 * s + c]);
        }
    int64_t capacity = ofo_tree_capacity(m, leaf_size);
    int64_t *order = (int64_t *)malloc((size_t)m * sizeof(int64_t));
    int64_t *lo = (int64_t *)malloc((size_t)capacity * sizeof(int64_t));
    int64_t *hi = (int64_t *)malloc((size_t)capacity * sizeof(int64_t));
    int64_t *left = (int64_t *)malloc((size_t)capacity * sizeof(int64_t));
    int64_t *right = (int64_t *)malloc((size_t)capacity * sizeof(int64_t));
    int64_t count = ofo_build_tree(m, mid, leaf_size, order, lo, hi, left, right, threads);
    double *centre = (double *)malloc((size_t)count * 3 * sizeof(double));
    double *bound = (double *)malloc((size_t)count * sizeof(double));
    double *core = (double *)malloc((size_t)count * sizeof(double));
    double *m0 = (double *)malloc((size_t)count * 3 * sizeof(double));
    double *m1 = (double *)malloc((size_t)count * 9 * sizeof(double));
    double *m2 = (double *)malloc((size_t)count * 27 * sizeof(double));
    double *w1 = (double *)malloc((size_t)count * 3 * sizeof(double));
    double *t2 = (double *)malloc((size_t)count * 3 * sizeof(double));
    double *w2 = (double *)malloc((size_t)count * 9 * sizeof(double));
    ofo_node_moments(count, order, lo, hi, starts, ends, alpha, lv, radius, gamma, centre, bound, core, m0, m1,
                     m2, threads);
    ofo_node_contractions(count, m1, m2, w1, t2, w2);
    int64_t *inverse = (int64_t *)malloc((size_t)m * sizeof(int64_t));
    for (int64_t k = 0; k < m; k++) inverse[order[k]] = k;
    int64_t e = ex_ptr[n];
    int64_t *positions = (int64_t *)malloc(((size_t)e + 1) * sizeof(int64_t));
    if (e) ofo_positions(n, ex_ptr, ex_seg, inverse, positions);
    ofo_tree_velocity(n, points, starts, directions, lengths, gamma, radius, model, raise_singular, order, lo, hi,
                      left, right, centre, bound, core, m0, m1, m2, w1, t2, w2, theta, expansion_order, clearance,
                      ex_ptr, positions, out, flags, far_counts, near_counts, threads);
    free(lv); free(ends); free(alpha); free(mid); free(order); free(lo); free(hi); free(left); free(right);
    free(centre); free(bound); free(core); free(m0); free(m1); free(m2); free(w1); free(t2); free(w2);
    free(inverse); free(positions);
    return count;
}

/* vortex_numba.py:271-287 (pair_matrix): unit-circulation pair velocities out[p, s, :]. */
void ofo_pair_matrix(int64_t n, const double *points, int64_t m, const double *starts, const double *directions,
                     const double *lengths, const double *radius, int model, int raise_singular, double *out,
                     int8_t *flags, int threads) {
    set_threads(threads);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++) {
        int flag = FLAG_OK;
        for (int64_t s = 0; s < m; s++) {
            double v[3];
            int f = ofo_pair_velocity(points[3 * p], points[3 * p + 1], points[3 * p + 2], starts[3 * s],
                                      starts[3 * s + 1], starts[3 * s + 2], directions[3 * s],
                                      directions[3 * s + 1], directions[3 * s + 2], lengths[s], 1.0, model,
                                      radius[s], raise_singular, v);
            if (f != FLAG_OK && flag == FLAG_OK) flag = f;
            double *o = out + 3 * (p * m + s);
            o[0] = v[0]; o[1] = v[1]; o[2] = v[2];
        }
        flags[p] = (int8_t)flag;
    }
}

/* vortex_numba.py:290-308 (assemble_columns). */
void ofo_assemble_columns(int64_t n, int64_t m, const double *pairs, int64_t columns, const int64_t *column_ptr,
                          const int64_t *column_segment, const double *column_sign, double *out) {
    for (int64_t p = 0; p < n; p++)
        for (int64_t j = 0; j < columns; j++) {
            double ux = 0.0, uy = 0.0, uz = 0.0;
            for (int64_t k = column_ptr[j]; k < column_ptr[j + 1]; k++) {
                int64_t s = column_segment[k];
                double sign = column_sign[k];
                const double *v = pairs + 3 * (p * m + s);
                ux += sign * v[0]; uy += sign * v[1]; uz += sign * v[2];
            }
            double *o = out + 3 * (p * columns + j);
            o[0] = ux; o[1] = uy; o[2] = uz;
        }
}
