// Device near-pair finder and adjacency classifier.
//
// Reference: fast_find_nearfield.get_nearfield and split_adjacent_close
// (tectosaur/mesh/fast_find_nearfield.cpp:22-182).  The pair set is the
// reference's leaf test
//     include (i, j)  iff  dist2(c_i, c_j) <= ((r_i + r_j) thr)^2
// with dist2 = ((0 + dx dx) + dy dy) + dz dz, evaluated here with the same
// operation order and no contraction, so the set is bit-identical to the
// host finder's (host_mesh.cpp); pairs come out sorted by (i, j).
//
// Layout and passes (all on one stream, no host round trip between them):
//   1. a uniform grid over the B centroids: cell id per point, counts per
//      cell (atomics), exclusive scan, scatter of point ids into cells;
//   2. count pass: one thread per A point scans the cells its reach
//      touches and counts matches; exclusive scan -> row offsets;
//   3. fill pass (a second call, after the caller has read the total):
//      the same scan writes j's into the row, then sorts the row in place
//      (rows are short: tens of entries).
// The classifier is one thread per pair (shared-vertex test in the
// reference's row-major corner order) followed by three order-preserving
// compactions (flag, scan, scatter).
#include <hipcub/hipcub.hpp>

#include "tct_common.h"

namespace tct {

struct GridDev {
    double lo[3];
    double cell;
    int dims[3];
};

__device__ __forceinline__ int grid_coord(double v, const GridDev &g, int d) {
    long long c = (long long)floor((v - g.lo[d]) / g.cell);
    c = c < 0 ? 0 : c;
    return (int)(c > g.dims[d] - 1 ? g.dims[d] - 1 : c);
}

// The reference's predicate, operation for operation (no FMA contraction).
__device__ __forceinline__ bool near_test(const double *ci, const double *cj, double ri,
                                          double rj, double thr) {
#pragma clang fp contract(off)
    double d2 = 0.0;
    const double t0 = ci[0] - cj[0];
    d2 += t0 * t0;
    const double t1 = ci[1] - cj[1];
    d2 += t1 * t1;
    const double t2 = ci[2] - cj[2];
    d2 += t2 * t2;
    const double s = (ri + rj) * thr;
    return !(d2 > s * s);
}

__global__ void fill_i32_kernel(int64_t n, int32_t *p, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_i64_kernel(int64_t n, int64_t *p, int64_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void grid_count_kernel(int n_b, const double *__restrict__ cb, GridDev g,
                                  int32_t *__restrict__ cid, int32_t *__restrict__ cell_count) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_b) return;
    const int x = grid_coord(cb[3 * (size_t)j], g, 0);
    const int y = grid_coord(cb[3 * (size_t)j + 1], g, 1);
    const int z = grid_coord(cb[3 * (size_t)j + 2], g, 2);
    const int c = (z * g.dims[1] + y) * g.dims[0] + x;
    cid[j] = c;
    atomicAdd(&cell_count[c], 1);
}

__global__ void grid_scatter_kernel(int n_b, const int32_t *__restrict__ cid,
                                    int32_t *__restrict__ cursor, int32_t *__restrict__ items) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_b) return;
    items[atomicAdd(&cursor[cid[j]], 1)] = j;
}

// FILL = false: count[i] = number of matches.  FILL = true: write the
// matches of row i at out[offsets[i]..] and sort them.
template <bool FILL>
__global__ void pair_scan_kernel(int n_a, const double *__restrict__ ca,
                                 const double *__restrict__ ra, const double *__restrict__ cb,
                                 const double *__restrict__ rb, double thr, double rmax_b, GridDev g,
                                 const int32_t *__restrict__ cell_start,
                                 const int32_t *__restrict__ items, int64_t *__restrict__ counts,
                                 const int64_t *__restrict__ offsets, int64_t *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_a) return;
    const double ci[3] = {ca[3 * (size_t)i], ca[3 * (size_t)i + 1], ca[3 * (size_t)i + 2]};
    const double r_i = ra[i];
    const double reach = thr * (r_i + rmax_b) * 1.0000001;
    int lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        lo[d] = grid_coord(ci[d] - reach, g, d);
        hi[d] = grid_coord(ci[d] + reach, g, d);
    }
    int64_t k = FILL ? offsets[i] : 0;
    const int64_t k0 = k;
    // the row's slot ends at offsets[i + 1]; a fill that finds more matches
    // than the count did (inputs changed between the calls) stops there
    const int64_t k_end = FILL ? offsets[i + 1] : 0;
    for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
            for (int x = lo[0]; x <= hi[0]; ++x) {
                const int c = (z * g.dims[1] + y) * g.dims[0] + x;
                for (int m = cell_start[c]; m < cell_start[c + 1]; ++m) {
                    const int j = items[m];
                    if (!near_test(ci, cb + 3 * (size_t)j, r_i, rb[j], thr)) continue;
                    if (FILL) {
                        if (k >= k_end) continue;
                        out[2 * k + 1] = j;
                    }
                    ++k;
                }
            }
    if (!FILL) {
        counts[i] = k;
        return;
    }
    if (k > k_end) k = k_end;
    // insertion sort of the row's j's, then the i column
    for (int64_t a = k0 + 1; a < k; ++a) {
        const int64_t v = out[2 * a + 1];
        int64_t b = a - 1;
        while (b >= k0 && out[2 * b + 1] > v) {
            out[2 * (b + 1) + 1] = out[2 * b + 1];
            --b;
        }
        out[2 * (b + 1) + 1] = v;
    }
    for (int64_t a = k0; a < k; ++a) out[2 * a] = i;
}

// cls: 0 near, 1 vertex-adjacent, 2 edge-adjacent, 3 coincident (dropped);
// corner[4] = the first two shared (corner_a, corner_b) in row-major order.
__global__ void classify_kernel(int64_t n, const int64_t *__restrict__ pairs,
                                const int64_t *__restrict__ tris_a,
                                const int64_t *__restrict__ tris_b, int32_t *__restrict__ flags,
                                int32_t *__restrict__ cls, int32_t *__restrict__ corner) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int64_t a = pairs[2 * p], b = pairs[2 * p + 1];
    int found = 0;
    int c[4] = {0, 0, 0, 0};
#pragma unroll
    for (int d1 = 0; d1 < 3; ++d1)
#pragma unroll
        for (int d2 = 0; d2 < 3; ++d2) {
            if (tris_a[3 * a + d1] != tris_b[3 * b + d2]) continue;
            if (found < 2) {
                c[2 * found] = d1;
                c[2 * found + 1] = d2;
            }
            ++found;
        }
    const int k = found >= 3 ? 3 : found;
    cls[p] = k;
#pragma unroll
    for (int q = 0; q < 4; ++q) corner[4 * p + q] = c[q];
#pragma unroll
    for (int q = 0; q < 3; ++q) flags[(size_t)q * n + p] = (k == q);
}

__global__ void compact_kernel(int64_t n, const int64_t *__restrict__ pairs,
                               const int32_t *__restrict__ cls, const int32_t *__restrict__ corner,
                               const int64_t *__restrict__ pos, int64_t *__restrict__ near,
                               int64_t *__restrict__ va, int64_t *__restrict__ ea,
                               int64_t *__restrict__ counts) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int k = cls[p];
    if (p == n - 1) {
        // totals: the exclusive position of the last pair plus its own flag
#pragma unroll
        for (int q = 0; q < 3; ++q) counts[q] = pos[(size_t)q * n + p] + (k == q ? 1 : 0);
    }
    if (k == 3) return;
    const int64_t r = pos[(size_t)k * n + p];
    const int64_t a = pairs[2 * p], b = pairs[2 * p + 1];
    const int32_t *c = corner + 4 * p;
    if (k == 0) {
        near[2 * r] = a;
        near[2 * r + 1] = b;
    } else if (k == 1) {
        int64_t *o = va + 4 * r;
        o[0] = a;
        o[1] = b;
        o[2] = c[0];
        o[3] = c[1];
    } else {
        int64_t *o = ea + 6 * r;
        o[0] = a;
        o[1] = b;
        o[2] = c[0];
        o[3] = c[1];
        o[4] = c[2];
        o[5] = c[3];
    }
}

// flags (int32, 0/1) -> exclusive positions (int64): widening on the fly.
struct WidenI32 {
    __host__ __device__ int64_t operator()(int32_t v) const { return v; }
};

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct NearScratch {
    size_t cid, cell_count, cell_start, cursor, items, counts, temp, temp_bytes, total;
};

static int near_scratch_layout(int64_t n_a, int64_t n_b, int64_t n_cells, NearScratch &s) {
    size_t t1 = 0, t2 = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, t1, (int32_t *)nullptr, (int32_t *)nullptr,
                                         (int)(n_cells + 1)) != hipSuccess ||
        hipcub::DeviceScan::ExclusiveSum(nullptr, t2, (int64_t *)nullptr, (int64_t *)nullptr,
                                         (int)(n_a + 1)) != hipSuccess)
        return set_error(TCT_EHIP, "near pairs: scan size query failed");
    size_t off = 0;
    s.cid = off;
    off += align256(4 * (size_t)n_b);
    s.cell_count = off;
    off += align256(4 * (size_t)(n_cells + 1));
    s.cell_start = off;
    off += align256(4 * (size_t)(n_cells + 1));
    s.cursor = off;
    off += align256(4 * (size_t)(n_cells + 1));
    s.items = off;
    off += align256(4 * (size_t)n_b);
    s.counts = off;
    off += align256(8 * (size_t)(n_a + 1));
    s.temp = off;
    s.temp_bytes = std::max<size_t>(std::max(t1, t2), 1);
    off += align256(s.temp_bytes);
    s.total = off;
    return TCT_OK;
}

static int make_grid(const double *lo, double cell, const int64_t *dims, GridDev &g,
                     int64_t &n_cells) {
    TCT_CHECK_ARG(lo && dims && cell > 0, "near pairs: bad grid");
    n_cells = 1;
    for (int d = 0; d < 3; ++d) {
        TCT_CHECK_ARG(dims[d] >= 1 && dims[d] < (1 << 20), "near pairs: grid dims[%d] = %lld", d,
                      (long long)dims[d]);
        g.lo[d] = lo[d];
        g.dims[d] = (int)dims[d];
        n_cells *= dims[d];
    }
    TCT_CHECK_ARG(n_cells < (1LL << 30), "near pairs: %lld grid cells", (long long)n_cells);
    g.cell = cell;
    return TCT_OK;
}

}  // namespace tct

using namespace tct;

extern "C" int tct_near_grid_params(int64_t n_b, const double *cb, double reach_max, double *lo,
                                    double *cell, int64_t *dims) {
    TCT_CHECK_ARG(n_b > 0 && cb && lo && cell && dims && reach_max >= 0,
                  "near_grid_params: bad arguments");
    double hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = INFINITY;
        hi[d] = -INFINITY;
    }
    for (int64_t i = 0; i < n_b; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], cb[3 * i + d]);
            hi[d] = std::max(hi[d], cb[3 * i + d]);
        }
    double ext = 0;
    for (int d = 0; d < 3; ++d) ext = std::max(ext, hi[d] - lo[d]);
    double c = std::max(reach_max, ext / 1024.0);
    if (!(c > 0)) c = 1.0;
    for (;;) {
        int64_t total = 1;
        for (int d = 0; d < 3; ++d) {
            dims[d] = (int64_t)std::floor((hi[d] - lo[d]) / c) + 1;
            total *= dims[d];
        }
        if (total <= std::max<int64_t>(8 * n_b, 64)) break;
        c *= 1.5;
    }
    *cell = c;
    return TCT_OK;
}

extern "C" int tct_near_pairs_dev_scratch(int64_t n_a, int64_t n_b, const int64_t *dims,
                                          size_t *bytes) {
    TCT_CHECK_ARG(n_a > 0 && n_b > 0 && n_a < (1LL << 31) && n_b < (1LL << 31) && dims && bytes,
                  "near_pairs_dev_scratch: bad arguments");
    int64_t n_cells = dims[0] * dims[1] * dims[2];
    TCT_CHECK_ARG(n_cells > 0 && n_cells < (1LL << 30), "near_pairs_dev_scratch: cells %lld",
                  (long long)n_cells);
    NearScratch s;
    int rc = near_scratch_layout(n_a, n_b, n_cells, s);
    if (rc != TCT_OK) return rc;
    *bytes = s.total;
    return TCT_OK;
}

static int near_pairs_common(int64_t n_a, const double *ca, const double *ra, int64_t n_b,
                             const double *cb, const double *rb, double thr, const double *lo,
                             double cell, const int64_t *dims, void *scratch,
                             size_t scratch_bytes, GridDev &g, NearScratch &s) {
    TCT_CHECK_ARG(n_a > 0 && n_b > 0 && n_a < (1LL << 31) && n_b < (1LL << 31),
                  "near pairs: n_a %lld n_b %lld", (long long)n_a, (long long)n_b);
    TCT_CHECK_ARG(ca && ra && cb && rb && scratch && thr >= 0, "near pairs: bad arguments");
    int64_t n_cells = 0;
    int rc = make_grid(lo, cell, dims, g, n_cells);
    if (rc != TCT_OK) return rc;
    rc = near_scratch_layout(n_a, n_b, n_cells, s);
    if (rc != TCT_OK) return rc;
    TCT_CHECK_ARG(scratch_bytes >= s.total, "near pairs: scratch %zu bytes < %zu", scratch_bytes,
                  s.total);
    return TCT_OK;
}

extern "C" int tct_near_pairs_dev_count(int64_t n_a, const double *ca, const double *ra,
                                        int64_t n_b, const double *cb, const double *rb,
                                        double thr, double rmax_b, const double *lo, double cell,
                                        const int64_t *dims, void *scratch, size_t scratch_bytes,
                                        int64_t *offsets, void *stream) {
    GridDev g;
    NearScratch s;
    int rc = near_pairs_common(n_a, ca, ra, n_b, cb, rb, thr, lo, cell, dims, scratch,
                               scratch_bytes, g, s);
    if (rc != TCT_OK) return rc;
    TCT_CHECK_ARG(offsets, "near pairs: null offsets");
    hipStream_t st = as_stream(stream);
    char *base = (char *)scratch;
    int32_t *cid = (int32_t *)(base + s.cid);
    int32_t *cell_count = (int32_t *)(base + s.cell_count);
    int32_t *cell_start = (int32_t *)(base + s.cell_start);
    int32_t *cursor = (int32_t *)(base + s.cursor);
    int32_t *items = (int32_t *)(base + s.items);
    int64_t *counts = (int64_t *)(base + s.counts);
    const int64_t n_cells = (int64_t)g.dims[0] * g.dims[1] * g.dims[2];
    hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)ceil_div(n_cells + 1, 256)), dim3(256), 0,
                       st, n_cells + 1, cell_count, 0);
    hipLaunchKernelGGL(grid_count_kernel, dim3((unsigned)ceil_div(n_b, 256)), dim3(256), 0, st,
                       (int)n_b, cb, g, cid, cell_count);
    size_t tb = s.temp_bytes;
    if (hipcub::DeviceScan::ExclusiveSum(base + s.temp, tb, cell_count, cell_start,
                                         (int)(n_cells + 1), st) != hipSuccess)
        return set_error(TCT_EHIP, "near pairs: cell scan failed");
    if (hipMemcpyAsync(cursor, cell_start, 4 * (size_t)n_cells, hipMemcpyDeviceToDevice, st) !=
        hipSuccess)
        return set_error(TCT_EHIP, "near pairs: cursor copy failed");
    hipLaunchKernelGGL(grid_scatter_kernel, dim3((unsigned)ceil_div(n_b, 256)), dim3(256), 0, st,
                       (int)n_b, cid, cursor, items);
    double rmax = rmax_b;
    hipLaunchKernelGGL((pair_scan_kernel<false>), dim3((unsigned)ceil_div(n_a, 128)), dim3(128), 0,
                       st, (int)n_a, ca, ra, cb, rb, thr, rmax, g, cell_start, items, counts,
                       (const int64_t *)nullptr, (int64_t *)nullptr);
    // counts[n_a] is the scan's tail: zero it so offsets[n_a] = total
    hipLaunchKernelGGL(fill_i64_kernel, dim3(1), dim3(64), 0, st, (int64_t)1, counts + n_a,
                       (int64_t)0);
    tb = s.temp_bytes;
    if (hipcub::DeviceScan::ExclusiveSum(base + s.temp, tb, counts, offsets, (int)(n_a + 1), st) !=
        hipSuccess)
        return set_error(TCT_EHIP, "near pairs: row scan failed");
    TCT_CHECK_LAUNCH("near_pairs_dev_count");
    return TCT_OK;
}

extern "C" int tct_near_pairs_dev_fill(int64_t n_a, const double *ca, const double *ra,
                                       int64_t n_b, const double *cb, const double *rb,
                                       double thr, double rmax_b, const double *lo, double cell,
                                       const int64_t *dims, void *scratch, size_t scratch_bytes,
                                       const int64_t *offsets, int64_t *out, void *stream) {
    GridDev g;
    NearScratch s;
    int rc = near_pairs_common(n_a, ca, ra, n_b, cb, rb, thr, lo, cell, dims, scratch,
                               scratch_bytes, g, s);
    if (rc != TCT_OK) return rc;
    TCT_CHECK_ARG(offsets && out, "near pairs: null output");
    char *base = (char *)scratch;
    hipLaunchKernelGGL((pair_scan_kernel<true>), dim3((unsigned)ceil_div(n_a, 128)), dim3(128), 0,
                       as_stream(stream), (int)n_a, ca, ra, cb, rb, thr, rmax_b, g,
                       (const int32_t *)(base + s.cell_start), (const int32_t *)(base + s.items),
                       (int64_t *)nullptr, offsets, out);
    TCT_CHECK_LAUNCH("near_pairs_dev_fill");
    return TCT_OK;
}

extern "C" int tct_split_adjacent_dev_scratch(int64_t n_pairs, size_t *bytes) {
    TCT_CHECK_ARG(n_pairs > 0 && n_pairs < (1LL << 31) && bytes, "split_adjacent_dev_scratch");
    size_t t = 0;
    hipcub::TransformInputIterator<int64_t, WidenI32, const int32_t *> it(nullptr, WidenI32());
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, t, it, (int64_t *)nullptr, (int)n_pairs) !=
        hipSuccess)
        return set_error(TCT_EHIP, "split_adjacent: scan size query failed");
    *bytes = align256(3 * 4 * (size_t)n_pairs) + align256(4 * (size_t)n_pairs) +
             align256(16 * (size_t)n_pairs) + align256(3 * 8 * (size_t)n_pairs) +
             align256(std::max<size_t>(t, 1));
    return TCT_OK;
}

extern "C" int tct_split_adjacent_close_dev(int64_t n_pairs, const int64_t *pairs,
                                            const int64_t *tris_a, const int64_t *tris_b,
                                            int64_t *near, int64_t *va, int64_t *ea,
                                            int64_t *counts, void *scratch, size_t scratch_bytes,
                                            void *stream) {
    TCT_CHECK_ARG(n_pairs > 0 && n_pairs < (1LL << 31), "split_adjacent_dev: n_pairs %lld",
                  (long long)n_pairs);
    TCT_CHECK_ARG(pairs && tris_a && tris_b && near && va && ea && counts && scratch,
                  "split_adjacent_dev: null pointer");
    size_t need = 0;
    int rc = tct_split_adjacent_dev_scratch(n_pairs, &need);
    if (rc != TCT_OK) return rc;
    TCT_CHECK_ARG(scratch_bytes >= need, "split_adjacent_dev: scratch %zu < %zu", scratch_bytes,
                  need);
    hipStream_t st = as_stream(stream);
    const size_t n = (size_t)n_pairs;
    char *p = (char *)scratch;
    int32_t *flags = (int32_t *)p;
    p += align256(3 * 4 * n);
    int32_t *cls = (int32_t *)p;
    p += align256(4 * n);
    int32_t *corner = (int32_t *)p;
    p += align256(16 * n);
    int64_t *pos = (int64_t *)p;
    p += align256(3 * 8 * n);
    void *temp = p;
    const size_t temp_bytes = need - (size_t)(p - (char *)scratch);
    const unsigned nb = (unsigned)ceil_div(n_pairs, 256);
    hipLaunchKernelGGL(classify_kernel, dim3(nb), dim3(256), 0, st, n_pairs, pairs, tris_a,
                       tris_b, flags, cls, corner);
    for (int q = 0; q < 3; ++q) {
        size_t t = temp_bytes;
        hipcub::TransformInputIterator<int64_t, WidenI32, const int32_t *> it(flags + q * n,
                                                                              WidenI32());
        if (hipcub::DeviceScan::ExclusiveSum(temp, t, it, pos + q * n, (int)n_pairs, st) !=
            hipSuccess)
            return set_error(TCT_EHIP, "split_adjacent_dev: scan failed");
    }
    hipLaunchKernelGGL(compact_kernel, dim3(nb), dim3(256), 0, st, n_pairs, pairs, cls, corner,
                       pos, near, va, ea, counts);
    TCT_CHECK_LAUNCH("split_adjacent_close_dev");
    return TCT_OK;
}
