// H-matrix structure on the host: cluster tree and block partition.
//
// Restates GroupAndSort / GroupAndSort_AllLevel / BuildHierarchy
// (scripts/Functions_Hmatrix.jl:1-212):
//   * level 1 splits all non-loading elements into two clusters; at each
//     further level a cluster is split in two when (end - begin) >=
//     MinimumElementsToCut, else carried over unchanged (:143-147);
//   * the split is 2-means on element centres (kmeans(k = 2, maxiter = 200),
//     :15); the reference's Clustering.jl seeds k-means++ from an unseeded
//     RNG, so its partition is not reproducible.  Here the two seeds are
//     deterministic (the centre farthest from the cluster mean, then the
//     centre farthest from that one) followed by Lloyd iterations until the
//     assignment is stable;
//   * children are ordered by the distance of their mean to ArrangePoint
//     (:21-30) and elements are sorted by child with their previous relative
//     order kept;
//   * a cluster's diameter is the diagonal of the bounding box of its rows'
//     Input_Segment columns 1:3 (:46-51): the centre for rectangles, the
//     first vertex for triangles;
//   * blocks: an optional dense loading block (loading sources onto all
//     non-loading receivers, :218-222), then level by level every pair
//     (i, j) of that level's clusters not covered by an admissible block of
//     a coarser level: admissible (|c_i - c_j| / max(d_i, d_j) > crit) ->
//     low-rank block, else on the last level -> dense block (:224-258).
//     Coverage is tracked through the parent pairs; the reference's
//     interval test (:233-234) is equivalent except for single-element
//     clusters, which its half-open comparison never reports as covered.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "qdfn_internal.h"

namespace qdfn {
namespace {

struct Node {
    int64_t begin, end;  // 0-based inclusive, sorted order
    int64_t level, parent;
    double c[3];
    double diam;
};

struct Tree {
    const double* ctr;   // n x 3 column-major (original order)
    const double* dpts;  // n x 3 column-major (original order)
    int64_t n;
    std::vector<int64_t> order;  // sorted position -> original index
    std::vector<Node> nodes;
};

void mean_of(const Tree& T, int64_t b, int64_t e, const std::vector<int64_t>& ids, double m[3]) {
    double s[3] = {0, 0, 0};
    for (int64_t q = b; q <= e; ++q) {
        const int64_t o = ids[q];
        for (int k = 0; k < 3; ++k) s[k] += T.ctr[o + k * T.n];
    }
    const double cnt = (double)(e - b + 1);
    for (int k = 0; k < 3; ++k) m[k] = s[k] / cnt;
}

double diameter_of(const Tree& T, int64_t b, int64_t e) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t q = b; q <= e; ++q) {
        const int64_t o = T.order[q];
        for (int k = 0; k < 3; ++k) {
            const double v = T.dpts[o + k * T.n];
            lo[k] = std::min(lo[k], v);
            hi[k] = std::max(hi[k], v);
        }
    }
    double s = 0;
    for (int k = 0; k < 3; ++k) s += (hi[k] - lo[k]) * (hi[k] - lo[k]);
    return std::sqrt(s);
}

double dist2(const double* a, const double* b) {
    double s = 0;
    for (int k = 0; k < 3; ++k) s += (a[k] - b[k]) * (a[k] - b[k]);
    return s;
}

// 2-means of the centres at sorted positions [b, e]; assign[q - b] in {0, 1}.
// Returns false when every centre falls in one cluster.
bool two_means(const Tree& T, int64_t b, int64_t e, int64_t max_iter, std::vector<int>& assign,
               double cent[2][3]) {
    const int64_t cnt = e - b + 1;
    auto P = [&](int64_t q, int k) { return T.ctr[T.order[q] + k * T.n]; };
    double m[3];
    mean_of(T, b, e, T.order, m);
    int64_t s0 = b;
    double best = -1;
    for (int64_t q = b; q <= e; ++q) {
        const double p[3] = {P(q, 0), P(q, 1), P(q, 2)};
        const double d = dist2(p, m);
        if (d > best) {
            best = d;
            s0 = q;
        }
    }
    const double p0[3] = {P(s0, 0), P(s0, 1), P(s0, 2)};
    int64_t s1 = b;
    best = -1;
    for (int64_t q = b; q <= e; ++q) {
        const double p[3] = {P(q, 0), P(q, 1), P(q, 2)};
        const double d = dist2(p, p0);
        if (d > best) {
            best = d;
            s1 = q;
        }
    }
    if (best <= 0) return false;  // all centres coincide
    for (int k = 0; k < 3; ++k) {
        cent[0][k] = P(s0, k);
        cent[1][k] = P(s1, k);
    }
    assign.assign(cnt, -1);
    for (int64_t it = 0; it < std::max<int64_t>(max_iter, 1); ++it) {
        bool changed = false;
        for (int64_t q = b; q <= e; ++q) {
            const double p[3] = {P(q, 0), P(q, 1), P(q, 2)};
            const int a = dist2(p, cent[1]) < dist2(p, cent[0]) ? 1 : 0;
            if (a != assign[q - b]) {
                assign[q - b] = a;
                changed = true;
            }
        }
        double s[2][3] = {{0, 0, 0}, {0, 0, 0}};
        int64_t c[2] = {0, 0};
        for (int64_t q = b; q <= e; ++q) {
            const int a = assign[q - b];
            ++c[a];
            for (int k = 0; k < 3; ++k) s[a][k] += P(q, k);
        }
        if (c[0] == 0 || c[1] == 0) return false;
        for (int a = 0; a < 2; ++a)
            for (int k = 0; k < 3; ++k) cent[a][k] = s[a][k] / (double)c[a];
        if (!changed) break;
    }
    return true;
}

// Split node `pi` (level l) into children at level l + 1 appended to T.nodes.
void split_node(Tree& T, int64_t pi, int64_t level, bool may_split, const qdfn_hmat_params& p) {
    const Node par = T.nodes[pi];
    const int64_t b = par.begin, e = par.end;
    std::vector<int> assign;
    double cent[2][3];
    const bool two = may_split && e > b && two_means(T, b, e, p.kmeans_max_iter, assign, cent);
    if (!two) {
        Node ch;
        ch.begin = b;
        ch.end = e;
        ch.level = level;
        ch.parent = pi;
        mean_of(T, b, e, T.order, ch.c);
        ch.diam = diameter_of(T, b, e);
        T.nodes.push_back(ch);
        return;
    }
    // order the two clusters by the distance of their centre to ArrangePoint
    const double d0 = dist2(cent[0], p.arrange_point), d1 = dist2(cent[1], p.arrange_point);
    const int first = (d1 < d0) ? 1 : 0;
    std::vector<int64_t> seg(T.order.begin() + b, T.order.begin() + e + 1), sorted;
    sorted.reserve(seg.size());
    for (int pass = 0; pass < 2; ++pass) {
        const int want = pass == 0 ? first : 1 - first;
        for (size_t q = 0; q < seg.size(); ++q)
            if (assign[q] == want) sorted.push_back(seg[q]);
    }
    std::copy(sorted.begin(), sorted.end(), T.order.begin() + b);
    int64_t cnt_first = 0;
    for (int a : assign) cnt_first += (a == first);
    const int64_t ranges[2][2] = {{b, b + cnt_first - 1}, {b + cnt_first, e}};
    for (int c = 0; c < 2; ++c) {
        Node ch;
        ch.begin = ranges[c][0];
        ch.end = ranges[c][1];
        ch.level = level;
        ch.parent = pi;
        mean_of(T, ch.begin, ch.end, T.order, ch.c);
        ch.diam = diameter_of(T, ch.begin, ch.end);
        T.nodes.push_back(ch);
    }
}

}  // namespace
}  // namespace qdfn

using namespace qdfn;

extern "C" int qdfn_hmat_structure(const qdfn_hmat_params* p, int64_t n, const double* ctr, const double* dpts,
                                   int64_t* perm, int64_t cap_nodes, int64_t* node_info, double* node_geo,
                                   int64_t* n_nodes, int64_t cap_blocks, int64_t* range_SR, int64_t* admissible,
                                   int64_t* n_blocks) {
    if (!p || !ctr || !dpts || !perm || !n_nodes || !n_blocks || n < 2)
        return fail(QDFN_ERR_ARG, "qdfn_hmat_structure: bad arguments");
    const int64_t L = p->n_loading, N0 = n - L;
    if (L < 0 || N0 < 2 || p->max_levels < 1 || p->min_elements_to_cut < 0)
        return fail(QDFN_ERR_ARG, "qdfn_hmat_structure: need >= 2 non-loading elements and >= 1 level");
    Tree T;
    T.ctr = ctr;
    T.dpts = dpts;
    T.n = n;
    T.order.resize(n);
    std::iota(T.order.begin(), T.order.end(), 0);
    // root (not reported): all non-loading elements
    Node root;
    root.begin = 0;
    root.end = N0 - 1;
    root.level = 0;
    root.parent = -1;
    mean_of(T, 0, N0 - 1, T.order, root.c);
    root.diam = diameter_of(T, 0, N0 - 1);
    T.nodes.push_back(root);
    std::vector<int64_t> lvl_begin(p->max_levels + 2, 0), lvl_end(p->max_levels + 2, 0);
    lvl_begin[0] = 0;
    lvl_end[0] = 1;
    for (int64_t l = 1; l <= p->max_levels; ++l) {
        lvl_begin[l] = (int64_t)T.nodes.size();
        for (int64_t pi = lvl_begin[l - 1]; pi < lvl_end[l - 1]; ++pi) {
            const Node& par = T.nodes[pi];
            const bool may = (l == 1) || (par.end - par.begin >= p->min_elements_to_cut);
            split_node(T, pi, l, may, *p);
        }
        lvl_end[l] = (int64_t)T.nodes.size();
    }
    // block partition, level by level in the reference's order
    std::vector<int64_t> blk;  // 4 per block, 0-based inclusive
    std::vector<int64_t> adm;
    if (L > 0) {
        blk.insert(blk.end(), {N0, n - 1, 0, N0 - 1});
        adm.push_back(0);
    }
    const double crit = p->dist_diam_ratio;
    std::vector<uint8_t> act_prev, adm_prev, act, admc;  // per level: B x B flags
    int64_t Bp = 0;
    for (int64_t l = 1; l <= p->max_levels; ++l) {
        const int64_t b0 = lvl_begin[l], B = lvl_end[l] - lvl_begin[l];
        act.assign(B * B, 0);
        admc.assign(B * B, 0);
        for (int64_t i = 0; i < B; ++i) {
            const Node& ni = T.nodes[b0 + i];
            for (int64_t j = 0; j < B; ++j) {
                const Node& nj = T.nodes[b0 + j];
                bool active = true;
                if (l > 1) {
                    const int64_t pi = ni.parent - lvl_begin[l - 1], pj = nj.parent - lvl_begin[l - 1];
                    active = act_prev[pi * Bp + pj] && !adm_prev[pi * Bp + pj];
                }
                if (!active) continue;
                act[i * B + j] = 1;
                const double dist = std::sqrt(dist2(ni.c, nj.c));
                const double diam = std::max(ni.diam, nj.diam);
                if (dist / diam > crit) {
                    admc[i * B + j] = 1;
                    blk.insert(blk.end(), {ni.begin, ni.end, nj.begin, nj.end});
                    adm.push_back(1);
                } else if (l == p->max_levels) {
                    blk.insert(blk.end(), {ni.begin, ni.end, nj.begin, nj.end});
                    adm.push_back(0);
                }
            }
        }
        act_prev.swap(act);
        adm_prev.swap(admc);
        Bp = B;
    }
    const int64_t nn = (int64_t)T.nodes.size() - 1, nb = (int64_t)adm.size();
    *n_nodes = nn;
    *n_blocks = nb;
    if (nn > cap_nodes || nb > cap_blocks) return fail(QDFN_ERR_ARG, "qdfn_hmat_structure: capacity too small");
    for (int64_t i = 0; i < n; ++i) perm[i] = T.order[i];
    for (int64_t q = 0; q < nn; ++q) {
        const Node& nd = T.nodes[q + 1];
        if (node_info) {
            node_info[4 * q + 0] = nd.begin + 1;
            node_info[4 * q + 1] = nd.end + 1;
            node_info[4 * q + 2] = nd.level;
            node_info[4 * q + 3] = nd.parent - 1;  // index into the reported list, -1 for level 1
        }
        if (node_geo) {
            for (int k = 0; k < 3; ++k) node_geo[4 * q + k] = nd.c[k];
            node_geo[4 * q + 3] = nd.diam;
        }
    }
    for (int64_t q = 0; q < nb; ++q) {
        if (range_SR)
            for (int k = 0; k < 4; ++k) range_SR[4 * q + k] = blk[4 * q + k] + 1;
        if (admissible) admissible[q] = adm[q];
    }
    return QDFN_OK;
}
