/* ORACLE — test infrastructure only.  Event catalog of the reference,
 * restated sequentially:
 *   orc_event_fragments  get_event_fragments  scripts/Functions_Plot.jl:430-510
 *   orc_merge_fragments  merge_fragments      scripts/Functions_Plot.jl:513-579
 * Arrays are column-major like the Julia matrices (History_Disp n_rec x n with
 * leading dimension ld, FaultCenter n x 3, locations count x 3). */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

typedef struct {
    double t;
    int64_t idx;
} orc_key_t;

/* Julia's isless on non-NaN Float64 (-0.0 before 0.0), ties by original index:
 * the stable sortperm of :498. */
static int key_cmp(const void *pa, const void *pb) {
    const orc_key_t *a = (const orc_key_t *)pa, *b = (const orc_key_t *)pb;
    if (a->t < b->t) return -1;
    if (a->t > b->t) return 1;
    if (signbit(a->t) != signbit(b->t)) return signbit(a->t) ? -1 : 1;
    return a->idx < b->idx ? -1 : (a->idx > b->idx);
}

/* Returns the fragment count, or -1 if it exceeds cap.  Outputs are in the
 * reference's final (time-sorted) order; element[] is the 0-based FaultIdx. */
int64_t orc_event_fragments(int64_t n_rec, int64_t n, const double *time, const double *disp, int64_t ld,
                            const double *center, const double *area, double mu, double crit, int64_t cap,
                            double *ftime, double *floc, double *fmom, int32_t *felem) {
    /* element-major scan, :467-494 */
    double *t0 = malloc(sizeof(double) * (cap > 0 ? cap : 1));
    double *m0 = malloc(sizeof(double) * (cap > 0 ? cap : 1));
    int64_t *e0 = malloc(sizeof(int64_t) * (cap > 0 ? cap : 1));
    int64_t cnt = 0;
    for (int64_t f = 0; f < n; ++f) {
        int on = 0;
        for (int64_t t = 1; t < n_rec; ++t) {
            const double rate = (disp[t + f * ld] - disp[t - 1 + f * ld]) / (time[t] - time[t - 1]);
            if (rate > crit) {
                const double inc = disp[t + f * ld] - disp[t - 1 + f * ld];
                const double m = inc * area[f] * mu;
                if (on == 0) {
                    on = 1;
                    if (cnt >= cap) {
                        free(t0);
                        free(m0);
                        free(e0);
                        return -1;
                    }
                    t0[cnt] = time[t];
                    e0[cnt] = f;
                    m0[cnt] = m;
                    ++cnt;
                } else {
                    m0[cnt - 1] = m0[cnt - 1] + m;
                }
            } else {
                on = 0;
            }
        }
    }
    /* stable sort by time, :498-501 */
    orc_key_t *k = malloc(sizeof(orc_key_t) * (cnt > 0 ? cnt : 1));
    for (int64_t i = 0; i < cnt; ++i) {
        k[i].t = t0[i];
        k[i].idx = i;
    }
    qsort(k, (size_t)cnt, sizeof(orc_key_t), key_cmp);
    for (int64_t i = 0; i < cnt; ++i) {
        const int64_t j = k[i].idx, e = e0[j];
        ftime[i] = t0[j];
        fmom[i] = m0[j];
        felem[i] = (int32_t)e;
        for (int c = 0; c < 3; ++c) floc[i + cnt * c] = center[e + n * c];
    }
    free(k);
    free(t0);
    free(m0);
    free(e0);
    return cnt;
}

/* Julia's generic 2-norm of a 3-vector in the unscaled range: sqrt of the
 * sum of squares in index order. */
static double norm3(double x, double y, double z) { return sqrt(x * x + y * y + z * z); }

/* Returns the merged event count.  mtime, mw: count entries; mloc: count x 3
 * with leading dimension F (the caller sizes by the fragment count);
 * fid: 0-based event of each fragment (EventNumberMerged_forFragment - 1). */
int64_t orc_merge_fragments(int64_t F, const double *ftime, const double *floc, const double *fmom, double mt,
                            double md, double *mtime, double *mloc, double *mmom, double *mw, int32_t *fid) {
    int64_t E = 0;
    for (int64_t i = 0; i < F; ++i) { /* :524-566, 0-based */
        int64_t back = 1;
        for (;;) {
            const int64_t j = i - back;
            if (i == 0) {
                E = 1;
                mtime[0] = ftime[i];
                for (int c = 0; c < 3; ++c) mloc[0 + F * c] = floc[i + F * c];
                fid[i] = 0;
                break;
            } else if (j < 0) {
                mtime[E] = ftime[i];
                for (int c = 0; c < 3; ++c) mloc[E + F * c] = floc[i + F * c];
                fid[i] = (int32_t)E;
                ++E;
                break;
            } else if (ftime[i] - ftime[j] < mt) {
                if (norm3(floc[i] - floc[j], floc[i + F] - floc[j + F], floc[i + 2 * F] - floc[j + 2 * F]) < md) {
                    fid[i] = fid[j];
                    break;
                }
                ++back;
            } else {
                mtime[E] = ftime[i];
                for (int c = 0; c < 3; ++c) mloc[E + F * c] = floc[i + F * c];
                fid[i] = (int32_t)E;
                ++E;
                break;
            }
        }
    }
    for (int64_t e = 0; e < E; ++e) mmom[e] = 0.0; /* :572-577 */
    for (int64_t i = 0; i < F; ++i) mmom[fid[i]] = mmom[fid[i]] + fmom[i];
    for (int64_t e = 0; e < E; ++e) mw[e] = (log10(mmom[e]) - 9.05) / 1.5;
    return E;
}
