/* ORACLE — test infrastructure only.  H-matrix parts of the reference:
 *   orc_hmat_matvec      HmatSolver_Pararllel (scripts/Functions_Hmatrix.jl:631-646):
 *                        y[rcv] += B * x[src] block by block (each row's sum in
 *                        block order, parallel over rows);
 *                        a low-rank block (PartialQR) applies as U * (V * x)
 *                        (Functions_Hmatrix.jl:637-638).
 *   orc_sim_hmat         main_H (scripts/Functions_RSFDFN3DMain_H.jl:2-383) through
 *                        the shared loop of sim.c: Elastic_Load_Disp =
 *                        H_load * (0 - DispOld) ./ -K_Self (:241-243),
 *                        EffNormalStressMatrixProduct = H_N * DispOld (:245-248),
 *                        external stress at TOld (:266-271).
 *   orc_build_hierarchy  BuildHierarchy (Functions_Hmatrix.jl:180-212) line by
 *                        line, including its interval test for blocks already
 *                        covered at a coarser level (:231-236). */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

int orc_set_threads(int n) {
    const int prev = omp_get_max_threads();
    if (n > 0) omp_set_num_threads(n);
    return prev;
}

/* Parallel over receiver rows with the serial summation order: t = V x of
 * every low-rank block first (each as in the one-thread loop), then each row
 * adds its blocks' contributions (each a sum over columns, or over the rank,
 * in ascending order) in block order, as y[rcv] += B * x[src] block by block
 * does.  The result is the same for any thread count. */
void orc_hmat_matvec(const orc_hmat_t *h, const double *x, double *y) {
    const int64_t n = h->n, nb = h->nblk;
    int64_t *toff = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nb + 1));
    toff[0] = 0;
    for (int64_t q = 0; q < nb; ++q) toff[q + 1] = toff[q] + (h->kind[q] == 0 ? 0 : h->rank[q]);
    double *t = (double *)malloc(sizeof(double) * (size_t)(toff[nb] > 0 ? toff[nb] : 1));
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t q = 0; q < nb; ++q) {
        if (h->kind[q] == 0) continue;
        const int64_t s0 = h->range[4 * q] - 1, k = h->range[4 * q + 1] - s0;
        const int64_t r = h->rank[q];
        const double *V = h->b[q], *xs = x + s0;
        for (int64_t p = 0; p < r; ++p) {
            double acc = 0.0;
            for (int64_t j = 0; j < k; ++j) acc += V[p + j * r] * xs[j];
            t[toff[q] + p] = acc;
        }
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i0 = 0; i0 < n; i0 += 64) {
        const int64_t i1 = i0 + 64 < n ? i0 + 64 : n;
        for (int64_t i = i0; i < i1; ++i) y[i] = 0.0;
        for (int64_t q = 0; q < nb; ++q) {
            const int64_t s0 = h->range[4 * q] - 1, s1 = h->range[4 * q + 1] - 1;
            const int64_t r0 = h->range[4 * q + 2] - 1, r1 = h->range[4 * q + 3] - 1;
            if (r1 < i0 || r0 >= i1) continue;
            const int64_t m = r1 - r0 + 1, k = s1 - s0 + 1;
            const int64_t lo = r0 > i0 ? r0 : i0, hi = r1 + 1 < i1 ? r1 + 1 : i1;
            const double *xs = x + s0;
            if (h->kind[q] == 0) {
                const double *A = h->a[q];
                for (int64_t i = lo; i < hi; ++i) {
                    double acc = 0.0;
                    for (int64_t j = 0; j < k; ++j) acc += A[(i - r0) + j * m] * xs[j];
                    y[i] += acc;
                }
            } else {
                const int64_t r = h->rank[q];
                const double *U = h->a[q], *tq = t + toff[q];
                for (int64_t i = lo; i < hi; ++i) {
                    double acc = 0.0;
                    for (int64_t p = 0; p < r; ++p) acc += U[(i - r0) + p * m] * tq[p];
                    y[i] += acc;
                }
            }
        }
    }
    free(t);
    free(toff);
}

typedef struct {
    const orc_hmat_t *S, *N;
    const double *k_self;
    double *tmp;
} hctx_t;

static void hmat_cb(const void *ctx, const double *x, double *TL, double *SigProd, int planar) {
    const hctx_t *c = (const hctx_t *)ctx;
    const int64_t n = c->S->n;
    for (int64_t i = 0; i < n; ++i) c->tmp[i] = 0.0 - x[i];
    orc_hmat_matvec(c->S, c->tmp, TL);
    for (int64_t i = 0; i < n; ++i) TL[i] = TL[i] / -c->k_self[i];
    if (!planar) orc_hmat_matvec(c->N, x, SigProd);
}

int orc_sim_hmat(const orc_hmat_t *shear_loading, const orc_hmat_t *normal, const orc_sim_in_t *in,
                 orc_sim_out_t *out) {
    hctx_t c;
    c.S = shear_loading;
    c.N = normal;
    c.k_self = in->k_self;
    c.tmp = (double *)malloc(sizeof(double) * (size_t)in->n);
    const int st = orc_sim_run(in, hmat_cb, &c, out);
    free(c.tmp);
    return st;
}

int64_t orc_build_hierarchy(int64_t n_nodes, const int64_t *node_begin, const int64_t *node_end,
                            const int64_t *node_level, const double *node_ctr, const double *node_diam,
                            int64_t total_levels, double crit, int64_t loading_begin, int64_t loading_end,
                            int64_t cap, int64_t *range_SR, int64_t *admissible) {
    int64_t nb = 0;
    if (loading_begin > 0) { /* :190-194 */
        if (nb >= cap) return -1;
        range_SR[0] = loading_begin;
        range_SR[1] = loading_end;
        range_SR[2] = 1;
        range_SR[3] = loading_begin - 1;
        admissible[0] = 0;
        nb = 1;
    }
    int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_nodes > 0 ? n_nodes : 1));
    for (int64_t level = 1; level <= total_levels; ++level) {
        int64_t B = 0;
        for (int64_t q = 0; q < n_nodes; ++q)
            if (node_level[q] == level) cur[B++] = q;
        for (int64_t ii = 0; ii < B; ++ii)
            for (int64_t jj = 0; jj < B; ++jj) {
                const int64_t i = cur[ii], j = cur[jj];
                int done = 0;
                for (int64_t k = 0; k < nb; ++k)
                    if (range_SR[4 * k] <= node_begin[i] && node_begin[i] < range_SR[4 * k + 1] &&
                        range_SR[4 * k + 2] <= node_begin[j] && node_begin[j] < range_SR[4 * k + 3])
                        done = 1;
                if (done) continue;
                double d2 = 0.0;
                for (int c = 0; c < 3; ++c) {
                    const double d = node_ctr[3 * i + c] - node_ctr[3 * j + c];
                    d2 += d * d;
                }
                const double dist = sqrt(d2);
                const double diam = node_diam[i] > node_diam[j] ? node_diam[i] : node_diam[j];
                int add = -1;
                if (dist / diam > crit)
                    add = 1;
                else if (level == total_levels)
                    add = 0;
                if (add >= 0) {
                    if (nb >= cap) {
                        free(cur);
                        return -1;
                    }
                    range_SR[4 * nb] = node_begin[i];
                    range_SR[4 * nb + 1] = node_end[i];
                    range_SR[4 * nb + 2] = node_begin[j];
                    range_SR[4 * nb + 3] = node_end[j];
                    admissible[nb] = add;
                    ++nb;
                }
            }
    }
    free(cur);
    return nb;
}
