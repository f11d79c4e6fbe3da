// Host-side (CPU, OpenMP) near-pair finder.
//
// Reference: fast_find_nearfield.get_nearfield (tectosaur/mesh/
// fast_find_nearfield.cpp:22-141, an octree dual traversal).  The pair set
// is defined by the leaf test at :76-83:
//     include (i, j)  iff  dist2(c_i, c_j) <= pow((r_i + r_j) * thr, 2)
// with dist2 = ((0 + dx*dx) + dy*dy) + dz*dz (fmm/geometry.hpp:7-13,61-64).
// This implementation finds exactly that set with a uniform grid over the
// B centroids (cell >= the largest possible reach), evaluates the same
// predicate with the same floating-point operation order (this file is
// compiled with -ffp-contract=off), and returns the pairs sorted by (i, j).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/tct_hip.h"

namespace tct {
int set_error(int code, const char *fmt, ...);
}

namespace {

inline double dist2(const double *a, const double *b) {
    double out = 0;
    for (int d = 0; d < 3; d++) {
        double t = a[d] - b[d];
        out += t * t;
    }
    return out;
}

struct Grid {
    double lo[3];
    double cell;
    int64_t dims[3];
    std::vector<int64_t> start;  // [n_cells + 1]
    std::vector<int64_t> items;  // B indices grouped by cell

    int64_t coord(double v, int d) const {
        int64_t c = (int64_t)std::floor((v - lo[d]) / cell);
        return std::min<int64_t>(std::max<int64_t>(c, 0), dims[d] - 1);
    }
};

void build_grid(Grid &g, int64_t n, const double *c, double cell) {
    double hi[3];
    for (int d = 0; d < 3; d++) {
        g.lo[d] = INFINITY;
        hi[d] = -INFINITY;
    }
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            g.lo[d] = std::min(g.lo[d], c[3 * i + d]);
            hi[d] = std::max(hi[d], c[3 * i + d]);
        }
    // keep the cell count near n so empty cells cost little
    double ext = 0;
    for (int d = 0; d < 3; d++) ext = std::max(ext, hi[d] - g.lo[d]);
    cell = std::max(cell, ext / 1024.0);
    if (!(cell > 0)) cell = 1.0;
    for (;;) {
        int64_t total = 1;
        for (int d = 0; d < 3; d++) {
            g.dims[d] = (int64_t)std::floor((hi[d] - g.lo[d]) / cell) + 1;
            total *= g.dims[d];
        }
        if (total <= std::max<int64_t>(8 * n, 64)) break;
        cell *= 1.5;
    }
    g.cell = cell;
    const int64_t nc = g.dims[0] * g.dims[1] * g.dims[2];
    g.start.assign(nc + 1, 0);
    std::vector<int64_t> cid(n);
    for (int64_t i = 0; i < n; i++) {
        int64_t cx = g.coord(c[3 * i], 0), cy = g.coord(c[3 * i + 1], 1),
                cz = g.coord(c[3 * i + 2], 2);
        cid[i] = (cz * g.dims[1] + cy) * g.dims[0] + cx;
        g.start[cid[i] + 1]++;
    }
    for (int64_t k = 0; k < nc; k++) g.start[k + 1] += g.start[k];
    g.items.resize(n);
    std::vector<int64_t> fill(g.start.begin(), g.start.end() - 1);
    for (int64_t i = 0; i < n; i++) g.items[fill[cid[i]]++] = i;
}

}  // namespace

// Threads of the OpenMP regions that the host utilities start from the
// calling thread (a server's preparation workers keep a few cores free for
// the thread that drives the device).
extern "C" int tct_set_host_threads(int n) {
    if (n < 1) return tct::set_error(TCT_EARG, "set_host_threads: n %d", n);
    omp_set_num_threads(n);
    return TCT_OK;
}

extern "C" int tct_find_near_pairs(int64_t n_a, const double *ca, const double *ra, int64_t n_b,
                                   const double *cb, const double *rb, double threshold,
                                   int64_t *out, int64_t cap, int64_t *n_out) {
    if (n_a < 0 || n_b < 0 || !n_out || (n_a && (!ca || !ra)) || (n_b && (!cb || !rb)) ||
        !(threshold >= 0))
        return tct::set_error(TCT_EARG, "find_near_pairs: bad arguments");
    *n_out = 0;
    if (n_a == 0 || n_b == 0) return TCT_OK;
    double rmax_a = 0, rmax_b = 0;
    for (int64_t i = 0; i < n_a; i++) rmax_a = std::max(rmax_a, ra[i]);
    for (int64_t j = 0; j < n_b; j++) rmax_b = std::max(rmax_b, rb[j]);
    Grid g;
    build_grid(g, n_b, cb, threshold * (rmax_a + rmax_b) * 1.0000001);

    // per-thread pair lists over contiguous i ranges, concatenated in order
    std::vector<std::vector<int64_t>> per_i(n_a);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n_a; i++) {
        const double *ci = ca + 3 * i;
        const double reach = threshold * (ra[i] + rmax_b) * 1.0000001;
        int64_t lo[3], hi[3];
        for (int d = 0; d < 3; d++) {
            lo[d] = g.coord(ci[d] - reach, d);
            hi[d] = g.coord(ci[d] + reach, d);
        }
        std::vector<int64_t> &mine = per_i[i];
        for (int64_t z = lo[2]; z <= hi[2]; z++)
            for (int64_t y = lo[1]; y <= hi[1]; y++)
                for (int64_t x = lo[0]; x <= hi[0]; x++) {
                    const int64_t cell = (z * g.dims[1] + y) * g.dims[0] + x;
                    for (int64_t k = g.start[cell]; k < g.start[cell + 1]; k++) {
                        const int64_t j = g.items[k];
                        const double lim = std::pow((ra[i] + rb[j]) * threshold, 2);
                        if (dist2(ci, cb + 3 * j) > lim) continue;
                        mine.push_back(j);
                    }
                }
        std::sort(mine.begin(), mine.end());
    }
    int64_t total = 0;
    for (int64_t i = 0; i < n_a; i++) total += (int64_t)per_i[i].size();
    *n_out = total;
    if (!out || cap < total) return TCT_OK;  // caller retries with cap >= *n_out
    int64_t k = 0;
    for (int64_t i = 0; i < n_a; i++)
        for (int64_t j : per_i[i]) {
            out[2 * k] = i;
            out[2 * k + 1] = j;
            k++;
        }
    return TCT_OK;
}

// Reference: split_adjacent_close (tectosaur/mesh/fast_find_nearfield.cpp
// :178-217).  For each pair count the shared vertices (a pair with three
// shared vertices is coincident and dropped); shared corners are recorded
// in (corner_a, corner_b) row-major order.  Output order follows the input.
// near [cap][2], va [cap][4], ea [cap][6] (cap >= n_pairs suffices);
// counts[3] receives the class sizes.
extern "C" int tct_split_adjacent_close(int64_t n_pairs, const int64_t *pairs,
                                        const int64_t *tris_a, const int64_t *tris_b,
                                        int64_t *near, int64_t *va, int64_t *ea,
                                        int64_t *counts) {
    if (n_pairs < 0 || !counts || (n_pairs > 0 && (!pairs || !tris_a || !tris_b || !near ||
                                                    !va || !ea)))
        return tct::set_error(TCT_EARG, "split_adjacent_close: bad arguments");
    std::vector<int8_t> cls(n_pairs);
    std::vector<int8_t> corner(4 * n_pairs);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n_pairs; k++) {
        const int64_t *a = tris_a + 3 * pairs[2 * k];
        const int64_t *b = tris_b + 3 * pairs[2 * k + 1];
        int found = 0;
        for (int d1 = 0; d1 < 3; d1++)
            for (int d2 = 0; d2 < 3; d2++) {
                if (a[d1] != b[d2]) continue;
                if (found < 2) {
                    corner[4 * k + 2 * found] = (int8_t)d1;
                    corner[4 * k + 2 * found + 1] = (int8_t)d2;
                }
                found++;
            }
        cls[k] = (int8_t)(found >= 3 ? 3 : found);
    }
    int64_t nn = 0, nv = 0, ne = 0;
    for (int64_t k = 0; k < n_pairs; k++) {
        const int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
        const int8_t *c = corner.data() + 4 * k;
        switch (cls[k]) {
        case 0:
            near[2 * nn] = i;
            near[2 * nn + 1] = j;
            nn++;
            break;
        case 1:
            va[4 * nv] = i;
            va[4 * nv + 1] = j;
            va[4 * nv + 2] = c[0];
            va[4 * nv + 3] = c[1];
            nv++;
            break;
        case 2:
            ea[6 * ne] = i;
            ea[6 * ne + 1] = j;
            for (int q = 0; q < 4; q++) ea[6 * ne + 2 + q] = c[q];
            ne++;
            break;
        default:
            break;
        }
    }
    counts[0] = nn;
    counts[1] = nv;
    counts[2] = ne;
    return TCT_OK;
}
