// Farfield direct N-body for the elastic Kelvin kernels U, T, A, H on gfx950.
//
// Reference semantics: farfield_pts{K} (tectosaur/farfield_direct.cl:15-113)
// and the quadrature-point loop of farfield_tris_to_tris{K}
// (tectosaur/matrix_free.cl:47-101): out_i = sum_j K(y_j - x_i) s_j with
// pairs at r2 == 0 skipped.
//
// Design (MI355X / CDNA4, wave64):
//  * One lane owns P observation points (coalesced SoA loads, kept in VGPRs).
//  * Source points are wave-uniform: each wave walks the same source list,
//    so the 12-word source record is fetched with scalar loads into SGPRs
//    and feeds every lane's FMAs directly -- no LDS staging and no barriers.
//  * The kernel body is the algebraically factored form of the reference's
//    vector code (source strengths pre-scaled by the kernel constant, per-
//    source dot products precomputed in the record), pure VALU, one
//    v_rsq_f32 per pair.  It is not a dense contraction, so no MFMA.
//  * FP32 path: coordinates are stored relative to an FP64 origin per group
//    of 64 consecutive points.  Per 64-point source tile the lane forms
//    x' = x_local - (O_src - O_obs) once, so D = y_local - x' is accurate to
//    FP32 relative to |D| + the tile extent rather than to the mesh extent.
//    Accumulation is FP32 within a source tile and FP64 across tiles.
//  * The source list can be split into chunks along grid.y; each chunk
//    writes its own partial sum and the reduce/scatter kernels add the
//    chunks in a fixed order (deterministic results).
#include <type_traits>

#include "far_pair.h"

namespace tct {
// Words of a source record that the pair body of KID reads.
template <int KID>
constexpr int used_words() {
    return KID == TCT_RH ? 12 : KID == TCT_RT ? 12 : KID == TCT_H ? 11 : KID == TCT_T ? 10 : 9;
}
// Source records of a segment-table launch are read through the constant
// address space: the kernel never writes them, and a uniform load from there
// is a scalar load although the pointer comes from the segment table.  The
// single-problem launch reads them through its restrict-qualified argument
// (scalar loads as well): with the constant address space there the
// compiler ran the two pairs of a pipeline stage one after the other
// (45 VGPRs instead of 62), 3% slower on H.
typedef const __attribute__((address_space(4))) float *rec_ptr;
template <bool SEG>
using RecPtr = typename std::conditional<SEG, rec_ptr, const float *>::type;

// Four floats as a native vector (LDS staging moves these).
typedef float f4 __attribute__((ext_vector_type(4)));

// U consecutive records (the first NW of their RW words) into registers.
template <int U, int NW, int RW, typename PTR>
__device__ __forceinline__ void load_recs(float (&q)[U][NW], PTR r) {
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int w = 0; w < NW; ++w) q[u][w] = r[u * RW + w];
}

// An empty statement that reads the last loaded word (so the scalar loads
// before it are waited for here), orders memory accesses around it, and
// passes the lane's coordinates through, so the pairs that use them are not
// moved above it.
template <int P>
__device__ __forceinline__ void fence_recs(float (&xs)[P][3], float last) {
#pragma unroll
    for (int p = 0; p < P; ++p)
        asm volatile("" : "+v"(xs[p][0]), "+v"(xs[p][1]), "+v"(xs[p][2]) : "s"(last) : "memory");
}

// Keeps a value in a VGPR from here on (an instruction that reads an SGPR
// is never dual-issued).
__device__ __forceinline__ void pin_vgpr(float &x) { asm volatile("" : "+v"(x)); }

// The (observation block, source chunk) of this workgroup.  Plain order:
// grid = (observation blocks, chunks).  XCD order (xcd_chunks > 0): a 1-D
// grid of observation blocks x chunks, chunks a multiple of 8; workgroups
// are dispatched round-robin over the 8 XCDs by linear id, so workgroup L
// runs on XCD L mod 8, and XCD x is given chunks x, x + 8, ..., each for
// every observation block before the next: the workgroups resident on one
// XCD at a time read one chunk of source records, which stays in that
// XCD's L2.  Any dispatch order gives the same results (the map is a
// bijection); only cache reuse depends on it.
struct BlockChunk {
    int ob, c, chunks;
};
__device__ __forceinline__ BlockChunk block_chunk(int xcd_chunks, int obs_blocks) {
    if (xcd_chunks == 0) return {(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.y};
    const int L = (int)blockIdx.x;
    const int k = L >> 3;
    return {k % obs_blocks, (k / obs_blocks) * 8 + (L & 7), xcd_chunks};
}

// ---------------------------------------------------------------------------
// FP32 N-body.  grid = (ceil(n_obs_pad / (blockDim*P)), chunks), or the XCD
// order of block_chunk.
// loc32 [6][n_obs_pad]; obs_org [n_obs_pad/G][3]; rec [n_src_pad][RW];
// src_org [n_src_pad/G][3]; partial [chunks][NO][n_obs_pad].
// sep2: a source tile whose origin is farther than sqrt(sep2) from the
// wave's observation group origins holds no pair with r2 <= eps2, so its
// pairs skip the test (the same values: the test passes them through).
// The last tile (it holds the padding records) always keeps the test.
// Several problems' N-bodies in one launch use a segment table passed by
// value: segment s owns linear workgroups [blk0, next blk0), laid out as the
// plain order of one problem's launch (observation block fastest, then
// chunk).  One kernel and precision for all segments; each has its own
// layout, chunks, constants and r2 test.
struct FarSeg {
    const float *loc32;
    const double *obs_org;
    const float *rec;
    const double *src_org;
    double *partial;
    double sep2;
    KConst<float> k;
    int n_obs_pad, n_tiles, chunks, obs_blocks, blk0;
};
constexpr int FAR_SEG_MAX = 24;
struct FarSegTable {
    int n;
    FarSeg s[FAR_SEG_MAX];
};
struct NoSegs {
    int n;
};
template <bool SEG>
using SegArg = typename std::conditional<SEG, FarSegTable, NoSegs>::type;

// SEG = false: one problem, the arguments as named, workgroups mapped by
// block_chunk.  SEG = true: the arguments come from this workgroup's
// segment of `segs` (the named ones are ignored).  The body is the kernel's
// own, not a shared device function: inlined into a kernel, a function body
// let the compiler move the next records' scalar loads above the wait for
// the current ones (they are invariant loads), which undid the pipelining.
// LDS = true: each wave stages its current 64-point source tile in LDS
// (one coalesced vector load of the next tile per lane, issued a tile ahead)
// and reads every record with broadcast LDS loads into VGPRs; the kernel
// constants are copied to VGPRs too.  gfx950 issues two VALU instructions
// of one SIMD in the same cycle only when neither reads an SGPR, so pair
// bodies whose record words and constants sit in VGPRs can be dual-issued
// throughout, where records in SGPRs (LDS = false: wave-uniform scalar
// loads) keep a third of the body single-issue.
template <int KID, int P, bool SEG, bool LDS>
__global__ __launch_bounds__(256) void far_p2p_f32_kernel(
    int n_obs_pad, const float *__restrict__ loc32, const double *__restrict__ obs_org,
    int n_tiles, const float *__restrict__ rec, const double *__restrict__ src_org,
    KConst<float> k, double sep2, double *__restrict__ partial, int xcd_chunks,
    int obs_blocks, SegArg<SEG> segs) {
    int ob, c, chunks;
    if constexpr (SEG) {
        const int L = (int)blockIdx.x;
        int s = 0;
        while (s + 1 < segs.n && L >= segs.s[s + 1].blk0) ++s;
        const FarSeg &g = segs.s[s];
        const int l = L - g.blk0;
        ob = l % g.obs_blocks;
        c = l / g.obs_blocks;
        chunks = g.chunks;
        n_obs_pad = g.n_obs_pad;
        loc32 = g.loc32;
        obs_org = g.obs_org;
        n_tiles = g.n_tiles;
        rec = g.rec;
        src_org = g.src_org;
        k = g.k;
        sep2 = g.sep2;
        partial = g.partial;
    } else {
        const BlockChunk bc = block_chunk(xcd_chunks, obs_blocks);
        ob = bc.ob;
        c = bc.c;
        chunks = bc.chunks;
    }
    constexpr int RW = rec_words<KID>();
    constexpr int NO = n_out<KID>();
    const int bd = blockDim.x;
    const int base = ob * bd * P;
    const int t0 = (int)(((long)n_tiles * c) / chunks);
    const int t1 = (int)(((long)n_tiles * (c + 1)) / chunks);

    float xl[P][3], m[P][3];
    double oo[P][3], acc64[P][NO];
    int idx[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        int i = base + p * bd + (int)threadIdx.x;
        idx[p] = i;
        int ic = i < n_obs_pad ? i : n_obs_pad - 1;
        int g = ic / G;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            xl[p][d] = loc32[(size_t)d * n_obs_pad + ic];
            m[p][d] = needs_obs_normal<KID>() ? loc32[(size_t)(3 + d) * n_obs_pad + ic] : 0.0f;
            oo[p][d] = obs_org[3 * g + d];
        }
#pragma unroll
        for (int a = 0; a < NO; ++a) acc64[p][a] = 0.0;
    }

    // LDS staging: the wave's tile buffer and the next tile's words held
    // in registers while the current tile computes
    constexpr int NV = LDS ? G * RW / 4 / 64 : 1;   // 4-word vectors per lane per tile
    extern __shared__ f4 lds_tiles[];
    f4 *wtile = lds_tiles + (LDS ? (threadIdx.x >> 6) * (G * RW / 4) : 0);
    const int lane = (int)(threadIdx.x & 63);
    f4 pre[NV];
    KConst<float> kv = k;
    if constexpr (LDS) {
        pin_vgpr(kv.c0);
        pin_vgpr(kv.c1);
        pin_vgpr(kv.c2);
        pin_vgpr(kv.c3);
        pin_vgpr(kv.c4);
        pin_vgpr(kv.eps2);
        // every chunk holds at least one tile (src_chunks <= n_tiles), so
        // tile t0 exists
        const f4 *src = (const f4 *)(rec + (size_t)t0 * G * RW);
#pragma unroll
        for (int v = 0; v < NV; ++v) pre[v] = src[v * 64 + lane];
    }

    for (int t = t0; t < t1; ++t) {
        if constexpr (LDS) {
            // the previous tile's LDS reads were issued before these writes,
            // and a wave's LDS operations complete in order
#pragma unroll
            for (int v = 0; v < NV; ++v) wtile[v * 64 + lane] = pre[v];
            // the next tile's words (the last tile loads itself again: an
            // unconditional load keeps pre in registers, where a guarded one
            // was placed in scratch memory)
            const int tn = t + 1 < t1 ? t + 1 : t;
            const f4 *src = (const f4 *)(rec + (size_t)tn * G * RW);
#pragma unroll
            for (int v = 0; v < NV; ++v) pre[v] = src[v * 64 + lane];
        }
        const double os0 = src_org[3 * t], os1 = src_org[3 * t + 1], os2 = src_org[3 * t + 2];
        float xs[P][3], acc[P][NO], am[P][n_mterm_arr<KID>()];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            xs[p][0] = xl[p][0] - (float)(os0 - oo[p][0]);
            xs[p][1] = xl[p][1] - (float)(os1 - oo[p][1]);
            xs[p][2] = xl[p][2] - (float)(os2 - oo[p][2]);
#pragma unroll
            for (int a = 0; a < NO; ++a) acc[p][a] = 0.0f;
#pragma unroll
            for (int a = 0; a < n_mterm_arr<KID>(); ++a) am[p][a] = 0.0f;
        }
        const RecPtr<SEG> r = (RecPtr<SEG>)(rec + (size_t)t * G * RW);
        bool close = t == n_tiles - 1;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const double d0 = os0 - oo[p][0], d1 = os1 - oo[p][1], d2 = os2 - oo[p][2];
            close = close || d0 * d0 + d1 * d1 + d2 * d2 <= sep2;
        }
        if constexpr (LDS) {
            constexpr int NQ = (used_words<KID>() + 3) / 4;   // float4 words read per record
            const f4 *tile = wtile;
            if (__any(close)) {
#pragma unroll 2
                for (int j = 0; j < G; ++j) {
                    float q[4 * NQ];
#pragma unroll
                    for (int v = 0; v < NQ; ++v) {
                        const f4 f = tile[j * (RW / 4) + v];
                        q[4 * v] = f.x;
                        q[4 * v + 1] = f.y;
                        q[4 * v + 2] = f.z;
                        q[4 * v + 3] = f.w;
                    }
#pragma unroll
                    for (int p = 0; p < P; ++p)
                        pair_eval<KID, float, true>(xs[p], m[p], q, kv, acc[p], am[p]);
                }
            } else {
#pragma unroll 2
                for (int j = 0; j < G; ++j) {
                    float q[4 * NQ];
#pragma unroll
                    for (int v = 0; v < NQ; ++v) {
                        const f4 f = tile[j * (RW / 4) + v];
                        q[4 * v] = f.x;
                        q[4 * v + 1] = f.y;
                        q[4 * v + 2] = f.z;
                        q[4 * v + 3] = f.w;
                    }
#pragma unroll
                    for (int p = 0; p < P; ++p)
                        pair_eval<KID, float, false>(xs[p], m[p], q, kv, acc[p], am[p]);
                }
            }
        } else if (__any(close)) {
            // the same pipelined loop with the r2 test (see below)
            constexpr int NW = used_words<KID>();
            constexpr int U = (NW > 12 || KID == TCT_RH) ? 1 : 2;
            float qa[U][NW], qb[U][NW];
            load_recs<U, NW, RW>(qa, r);
#pragma unroll 1
            for (int j = 0; j < G; j += 2 * U) {
                fence_recs<P>(xs, qa[U - 1][NW - 1]);
                load_recs<U, NW, RW>(qb, r + (j + U) * RW);
#pragma unroll
                for (int u = 0; u < U; ++u)
#pragma unroll
                    for (int p = 0; p < P; ++p)
                        pair_eval<KID, float, true>(xs[p], m[p], qa[u], k, acc[p], am[p]);
                fence_recs<P>(xs, qb[U - 1][NW - 1]);
                load_recs<U, NW, RW>(qa, r + min(j + 2 * U, G - U) * RW);
#pragma unroll
                for (int u = 0; u < U; ++u)
#pragma unroll
                    for (int p = 0; p < P; ++p)
                        pair_eval<KID, float, true>(xs[p], m[p], qb[u], k, acc[p], am[p]);
            }
        } else {
            // Records of the next U source points are loaded while the
            // current U compute: scalar loads return out of order, so a wait
            // for one waits for all; fence_recs places that wait before the
            // next loads are issued and the pairs that need it after it.
            constexpr int NW = used_words<KID>();
            constexpr int U = (NW > 12 || KID == TCT_RH) ? 1 : 2;
            float qa[U][NW], qb[U][NW];
            load_recs<U, NW, RW>(qa, r);
#pragma unroll 1
            for (int j = 0; j < G; j += 2 * U) {
                fence_recs<P>(xs, qa[U - 1][NW - 1]);
                load_recs<U, NW, RW>(qb, r + (j + U) * RW);
#pragma unroll
                for (int u = 0; u < U; ++u)
#pragma unroll
                    for (int p = 0; p < P; ++p)
                        pair_eval<KID, float, false>(xs[p], m[p], qa[u], k, acc[p], am[p]);
                fence_recs<P>(xs, qb[U - 1][NW - 1]);
                load_recs<U, NW, RW>(qa, r + min(j + 2 * U, G - U) * RW);
#pragma unroll
                for (int u = 0; u < U; ++u)
#pragma unroll
                    for (int p = 0; p < P; ++p)
                        pair_eval<KID, float, false>(xs[p], m[p], qb[u], k, acc[p], am[p]);
            }
        }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            fold_mterm<KID, float>(m[p], kv, acc[p], am[p]);
#pragma unroll
            for (int a = 0; a < NO; ++a) acc64[p][a] += (double)acc[p][a];
        }
    }

    double *out = partial + (size_t)c * NO * n_obs_pad;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        if (idx[p] < n_obs_pad) {
#pragma unroll
            for (int a = 0; a < NO; ++a) out[(size_t)a * n_obs_pad + idx[p]] = acc64[p][a];
        }
    }
}

// FP64 N-body ("plain FP64" path): absolute coordinates, FP64 throughout.
// soa64 [6][n_obs_pad]; rec64 [n_src_pad][RW].
template <int KID, int P>
__global__ __launch_bounds__(256) void far_p2p_f64_kernel(int n_obs_pad,
                                                          const double *__restrict__ soa64,
                                                          int n_tiles,
                                                          const double *__restrict__ rec,
                                                          KConst<double> k,
                                                          double *__restrict__ partial,
                                                          int xcd_chunks, int obs_blocks) {
    constexpr int RW = rec_words<KID>();
    constexpr int NO = n_out<KID>();
    const int bd = blockDim.x;
    const BlockChunk bc = block_chunk(xcd_chunks, obs_blocks);
    const int base = bc.ob * bd * P;
    const int chunks = bc.chunks;
    const int c = bc.c;
    const int t0 = (int)(((long)n_tiles * c) / chunks);
    const int t1 = (int)(((long)n_tiles * (c + 1)) / chunks);

    double x[P][3], m[P][3], acc[P][NO], am[P][n_mterm_arr<KID>()];
    int idx[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        int i = base + p * bd + (int)threadIdx.x;
        idx[p] = i;
        int ic = i < n_obs_pad ? i : n_obs_pad - 1;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            x[p][d] = soa64[(size_t)d * n_obs_pad + ic];
            m[p][d] = needs_obs_normal<KID>() ? soa64[(size_t)(3 + d) * n_obs_pad + ic] : 0.0;
        }
#pragma unroll
        for (int a = 0; a < NO; ++a) acc[p][a] = 0.0;
#pragma unroll
        for (int a = 0; a < n_mterm_arr<KID>(); ++a) am[p][a] = 0.0;
    }
    const double *r = rec + (size_t)t0 * G * RW;
    const int nj = (t1 - t0) * G;
#pragma unroll 2
    for (int j = 0; j < nj; ++j) {
        const double *q = r + (size_t)j * RW;
#pragma unroll
        for (int p = 0; p < P; ++p) pair_eval<KID, double>(x[p], m[p], q, k, acc[p], am[p]);
    }
#pragma unroll
    for (int p = 0; p < P; ++p) fold_mterm<KID, double>(m[p], k, acc[p], am[p]);
    double *out = partial + (size_t)c * NO * n_obs_pad;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        if (idx[p] < n_obs_pad) {
#pragma unroll
            for (int a = 0; a < NO; ++a) out[(size_t)a * n_obs_pad + idx[p]] = acc[p][a];
        }
    }
}

// ---------------------------------------------------------------------------
// Layout packing.  One wave per 64-point group.
constexpr double FAR_AWAY = 1.0e12;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ void pack_obs_kernel(int64_t n, const double *__restrict__ pts,
                                const double *__restrict__ ns, int64_t n_pad,
                                float *__restrict__ loc32, double *__restrict__ org,
                                double *__restrict__ soa64) {
    const int64_t g = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t i = g * G + lane;
    const bool valid = i < n;
    double p[3], nn[3];
    for (int d = 0; d < 3; ++d) {
        p[d] = valid ? pts[3 * i + d] : 0.0;
        nn[d] = valid ? ns[3 * i + d] : 0.0;
    }
    const double cnt = wave_sum(valid ? 1.0 : 0.0);
    double o[3];
    for (int d = 0; d < 3; ++d) o[d] = wave_sum(p[d]) / cnt;
    if (lane == 0)
        for (int d = 0; d < 3; ++d) org[3 * g + d] = o[d];
    for (int d = 0; d < 3; ++d) {
        loc32[d * n_pad + i] = valid ? (float)(p[d] - o[d]) : 0.0f;
        loc32[(3 + d) * n_pad + i] = (float)nn[d];
        soa64[d * n_pad + i] = p[d];
        soa64[(3 + d) * n_pad + i] = nn[d];
    }
}

// rw words per record: [y n 0 ...]; padding points far away, zero strength.
__global__ void pack_src_kernel(int64_t n, const double *__restrict__ pts,
                                const double *__restrict__ ns, int64_t n_pad, int rw,
                                float *__restrict__ rec32, double *__restrict__ org,
                                double *__restrict__ rec64) {
    const int64_t g = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t i = g * G + lane;
    const bool valid = i < n;
    double p[3], nn[3];
    for (int d = 0; d < 3; ++d) {
        p[d] = valid ? pts[3 * i + d] : 0.0;
        nn[d] = valid ? ns[3 * i + d] : 0.0;
    }
    const double cnt = wave_sum(valid ? 1.0 : 0.0);
    double o[3];
    for (int d = 0; d < 3; ++d) o[d] = wave_sum(p[d]) / cnt;
    if (lane == 0)
        for (int d = 0; d < 3; ++d) org[3 * g + d] = o[d];
    float *r32 = rec32 + i * rw;
    double *r64 = rec64 + i * rw;
    for (int d = 0; d < 3; ++d) {
        r32[d] = valid ? (float)(p[d] - o[d]) : (float)FAR_AWAY;
        r64[d] = valid ? p[d] : FAR_AWAY;
        r32[3 + d] = (float)nn[d];
        r64[3 + d] = nn[d];
    }
    for (int d = 6; d < rw; ++d) {
        r32[d] = 0.0f;
        r64[d] = 0.0;
    }
}

// Strength scale and per-source extras for the point kernels.
struct StrengthConsts {
    double scale;     // multiplies the raw strength
    double x0, x1;    // e0 = x0 (n.s'), e1 = x1 (n.s')
};

inline StrengthConsts strength_consts(int kid, const double *params) {
    ElasticConsts c = elastic_consts(params);
    StrengthConsts s{1.0, 0.0, 0.0};
    switch (kid) {
    case TCT_U: s.scale = c.CsU1; break;
    case TCT_T: s.scale = c.CsT1; s.x0 = 1.0; break;
    case TCT_A: s.scale = c.CsT1; break;
    case TCT_H: s.scale = c.CsH0; s.x0 = c.CsH2; s.x1 = 3.0 * c.CsH1; break;
    }
    return s;
}

template <typename T>
__device__ __forceinline__ void write_strength(T *r, double s0, double s1, double s2,
                                               const StrengthConsts &sc) {
    const double n0 = (double)r[3], n1 = (double)r[4], n2 = (double)r[5];
    const double sn = n0 * s0 + n1 * s1 + n2 * s2;
    r[6] = (T)s0;
    r[7] = (T)s1;
    r[8] = (T)s2;
    r[9] = (T)(sc.x0 * sn);
    r[10] = (T)(sc.x1 * sn);
}

template <typename T>
__global__ void strength_pts_kernel(int64_t n, int64_t n_pad, const double *__restrict__ vec,
                                    StrengthConsts sc, T *__restrict__ rec) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pad) return;
    double s[3] = {0.0, 0.0, 0.0};
    if (i < n)
        for (int d = 0; d < 3; ++d) s[d] = sc.scale * vec[3 * i + d];
    write_strength<T>(rec + i * R, s[0], s[1], s[2], sc);
}

template <typename T>
__global__ void gather_tris_kernel(int64_t n_pts, int64_t n_pad, int nq,
                                   const int32_t *__restrict__ tri_of_pt,
                                   const double *__restrict__ qphi, const double *__restrict__ wj,
                                   const double *__restrict__ v, StrengthConsts sc,
                                   T *__restrict__ rec) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pad) return;
    double s[3] = {0.0, 0.0, 0.0};
    if (i < n_pts) {
        const int64_t t = tri_of_pt[i];
        const int q = (int)(i % nq);
        const double f = sc.scale * wj[i];
        const double *vt = v + 9 * t;
        const double p0 = qphi[3 * q], p1 = qphi[3 * q + 1], p2 = qphi[3 * q + 2];
        for (int d = 0; d < 3; ++d) s[d] = f * (p0 * vt[d] + p1 * vt[3 + d] + p2 * vt[6 + d]);
    }
    write_strength<T>(rec + i * R, s[0], s[1], s[2], sc);
}

__device__ __forceinline__ int levi3(int i, int j, int k) { return (i - j) * (j - k) * (k - i) / 2; }

// Surface-curl kernels: per source point the record fields that carry the
// source triangle's surface curl S[d][I] = sum_b curl[b][I] v[9t+3b+d]
// (kernels.cl:100-206) and the interpolated input.  tri_curl [n_tris][3][3]
// (basis, component), indexed by DOF-space triangle.
template <int KID, typename T>
__global__ void gather_curl_kernel(int64_t n_pts, int64_t n_pad, int nq,
                                   const int32_t *__restrict__ tri_of_pt,
                                   const double *__restrict__ qphi, const double *__restrict__ wj,
                                   const double *__restrict__ tri_curl,
                                   const double *__restrict__ v, ElasticConsts c,
                                   T *__restrict__ rec) {
    constexpr int RW = rec_words<KID>();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pad) return;
    T *r = rec + i * RW;
    double S[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double in[3] = {0, 0, 0};
    double f = 0.0;
    if (i < n_pts) {
        const int64_t t = tri_of_pt[i];
        const int q = (int)(i % nq);
        f = wj[i];
        const double *vt = v + 9 * t;
        const double *cu = tri_curl + 9 * t;
        for (int d = 0; d < 3; ++d) {
            in[d] = qphi[3 * q] * vt[d] + qphi[3 * q + 1] * vt[3 + d] + qphi[3 * q + 2] * vt[6 + d];
            for (int I = 0; I < 3; ++I)
                S[d][I] = cu[I] * vt[d] + cu[3 + I] * vt[3 + d] + cu[6 + I] * vt[6 + d];
        }
    }
    if constexpr (KID == TCT_RT) {
        for (int tt = 0; tt < 3; ++tt) {
            double a = 0.0;
            for (int d = 0; d < 3; ++d)
                for (int I = 0; I < 3; ++I) a += levi3(I, tt, d) * S[d][I];
            r[9 + tt] = (T)(c.CsRT1 * f * a);
            r[6 + tt] = (T)(c.CsRT2 * f * in[tt]);
        }
    } else if constexpr (KID == TCT_RA) {
        for (int d = 0; d < 3; ++d) r[6 + d] = (T)(f * in[d]);
    } else {
        // G only: the kernels apply P(G) = CsRH1 tr(G) I + CsRH2 G^T - 2 G
        // to their sums of G / r
        const double g = f * c.CsRH0;
        for (int d = 0; d < 3; ++d)
            for (int mm = 0; mm < 3; ++mm) r[3 + d * 3 + mm] = (T)(g * S[d][mm]);
    }
}

__global__ void reduce_pts_kernel(int64_t n, int64_t n_obs_pad, int chunks,
                                  const double *__restrict__ partial, double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int d = 0; d < 3; ++d) {
        double s = 0.0;
        for (int c = 0; c < chunks; ++c) s += partial[((size_t)c * 3 + d) * n_obs_pad + i];
        out[3 * i + d] = s;
    }
}

// partial[0][r] = sum_c partial[c][r] over r < stride (one chunk's length),
// in fixed chunk order; each thread touches only its own r, so in place is
// safe.  Lets the Galerkin scatters read one chunk with their scattered,
// triangle-ordered accesses.
__global__ void chunk_sum_kernel(int64_t stride, int chunks, double *__restrict__ partial) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= stride) return;
    double s = partial[r];
    for (int c = 1; c < chunks; ++c) s += partial[(size_t)c * stride + r];
    partial[r] = s;
}

__global__ void scatter_tris_kernel(int64_t n_tris, int nq, const int32_t *__restrict__ pt_of_tri,
                                    const double *__restrict__ qphi,
                                    const double *__restrict__ wj_obs, int64_t n_obs_pad,
                                    int chunks, const double *__restrict__ partial,
                                    double *__restrict__ out) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n_tris * 9) return;
    const int64_t t = gid / 9;
    const int bd = (int)(gid % 9);
    const int b = bd / 3, d = bd % 3;
    const int64_t p0 = (int64_t)pt_of_tri[t] * nq;
    double s = 0.0;
    for (int q = 0; q < nq; ++q) {
        const int64_t p = p0 + q;
        double r = 0.0;
        for (int c = 0; c < chunks; ++c) r += partial[((size_t)c * 3 + d) * n_obs_pad + p];
        s += wj_obs[p] * qphi[3 * q + b] * r;
    }
    out[gid] = s;
}

// Galerkin scatter of the 9-moment kernels (RA, RH): applies the observation
// triangle's surface curls.  tri_curl [n_tris][3][3] by DOF-space triangle.
// One thread per triangle: each quadrature point's 9 moments are summed over
// the chunks once and contracted with all three basis curls.
template <int KID>
__global__ void scatter_curl_kernel(int64_t n_tris, int nq, const int32_t *__restrict__ pt_of_tri,
                                    const double *__restrict__ qphi,
                                    const double *__restrict__ wj_obs,
                                    const double *__restrict__ tri_curl, int64_t n_obs_pad,
                                    int chunks, const double *__restrict__ partial,
                                    ElasticConsts c, double *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tris) return;
    double cu[3][3];
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int I = 0; I < 3; ++I) cu[b][I] = tri_curl[9 * t + 3 * b + I];
    const int64_t p0 = (int64_t)pt_of_tri[t] * nq;
    double s[9];
#pragma unroll
    for (int o = 0; o < 9; ++o) s[o] = 0.0;
    for (int q = 0; q < nq; ++q) {
        const int64_t p = p0 + q;
        double M[9];
#pragma unroll
        for (int a = 0; a < 9; ++a) {
            double r = 0.0;
            for (int ch = 0; ch < chunks; ++ch) r += partial[((size_t)ch * 9 + a) * n_obs_pad + p];
            M[a] = r;
        }
        const double w = wj_obs[p];
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double val;
                if constexpr (KID == TCT_RH) {
                    val = M[k * 3] * cu[b][0] + M[k * 3 + 1] * cu[b][1] + M[k * 3 + 2] * cu[b][2];
                } else {
                    // sum_I curl[I] sum_s eps_{I s k} (CsRT1 v1_s + CsRT0 v2_s) - CsRT2 phi_b v3_k
                    double e = 0.0;
#pragma unroll
                    for (int I = 0; I < 3; ++I)
#pragma unroll
                        for (int ss = 0; ss < 3; ++ss)
                            e += cu[b][I] * levi3(I, ss, k) *
                                 (c.CsRT1 * M[ss] + c.CsRT0 * M[3 + ss]);
                    val = e - c.CsRT2 * qphi[3 * q + b] * M[6 + k];
                }
                s[b * 3 + k] += w * val;
            }
    }
#pragma unroll
    for (int o = 0; o < 9; ++o) out[9 * t + o] = s[o];
}

// ---------------------------------------------------------------------------
// LDS bytes of a workgroup of `block` threads with staged source tiles.
template <int KID>
constexpr size_t tile_lds_bytes(int block) {
    return (size_t)(block / 64) * G * rec_words<KID>() * sizeof(float);
}

template <int KID, int P>
static void launch_f32(dim3 grid, int block, hipStream_t st, int n_obs_pad, const float *loc32,
                       const double *obs_org, int n_tiles, const float *rec,
                       const double *src_org, KConst<float> k, double sep2, double *partial,
                       int xcd_chunks, int obs_blocks, bool lds) {
    if (lds)
        hipLaunchKernelGGL((far_p2p_f32_kernel<KID, P, false, true>), grid, dim3(block),
                           tile_lds_bytes<KID>(block), st, n_obs_pad, loc32, obs_org, n_tiles,
                           rec, src_org, k, sep2, partial, xcd_chunks, obs_blocks, NoSegs{0});
    else
        hipLaunchKernelGGL((far_p2p_f32_kernel<KID, P, false, false>), grid, dim3(block), 0, st,
                           n_obs_pad, loc32, obs_org, n_tiles, rec, src_org, k, sep2, partial,
                           xcd_chunks, obs_blocks, NoSegs{0});
}

template <int KID, int P>
static void launch_f64(dim3 grid, int block, hipStream_t st, int n_obs_pad, const double *soa64,
                       int n_tiles, const double *rec, KConst<double> k, double *partial,
                       int xcd_chunks, int obs_blocks) {
    hipLaunchKernelGGL((far_p2p_f64_kernel<KID, P>), grid, dim3(block), 0, st, n_obs_pad, soa64,
                       n_tiles, rec, k, partial, xcd_chunks, obs_blocks);
}


// sep^2 beyond which a source tile holds no pair with r2 <= eps2: points of
// groups whose origins are d apart are at least d - group_extent apart
// (group_extent = E_obs + E_src); INFINITY keeps the test on every tile.
static double tile_sep2(double eps2, double group_extent, int flags) {
    if (eps2 > 0.0 && group_extent >= 0.0 && std::isfinite(group_extent) && !(flags & 1)) {
        const double e = std::sqrt(eps2);
        const double sep = group_extent * (1.0 + 1e-3) + 5.0 * e;
        return sep * sep;
    }
    return INFINITY;
}

template <int KID, int P>
static void launch_f32_batched(unsigned blocks, int block, hipStream_t st, const FarSegTable &t) {
    hipLaunchKernelGGL((far_p2p_f32_kernel<KID, P, true, false>), dim3(blocks), dim3(block), 0, st, 0,
                       (const float *)nullptr, (const double *)nullptr, 0, (const float *)nullptr,
                       (const double *)nullptr, KConst<float>{}, 0.0, (double *)nullptr, 0, 1, t);
}

}  // namespace tct

using namespace tct;

extern "C" int tct_far_p2p_batched(int kernel, const double *params_host, int block,
                                   int obs_per_lane, int n_seg, const tct_far_segment *segs,
                                   void *stream) {
    TCT_CHECK_ARG(kernel >= TCT_U && kernel <= TCT_RH, "far_p2p_batched: kernel id %d", kernel);
    TCT_CHECK_ARG(block == 64 || block == 128 || block == 256, "far_p2p_batched: block %d",
                  block);
    TCT_CHECK_ARG(obs_per_lane == 1 || obs_per_lane == 2 || obs_per_lane == 4,
                  "far_p2p_batched: obs_per_lane %d", obs_per_lane);
    TCT_CHECK_ARG(n_seg >= 0 && (n_seg == 0 || (segs && params_host)),
                  "far_p2p_batched: bad segment list");
    hipStream_t st = as_stream(stream);
    const KConst<float> kc0 = kernel_consts<float>(kernel, params_host, 0.0);
    for (int s0 = 0; s0 < n_seg; s0 += FAR_SEG_MAX) {
        FarSegTable t;
        t.n = n_seg - s0 < FAR_SEG_MAX ? n_seg - s0 : FAR_SEG_MAX;
        int64_t blocks = 0;
        for (int i = 0; i < t.n; ++i) {
            const tct_far_segment &a = segs[s0 + i];
            const int si = s0 + i;
            TCT_CHECK_ARG(a.n_obs_pad > 0 && a.n_obs_pad % G == 0 && a.n_src_pad > 0 &&
                              a.n_src_pad % G == 0 && a.n_obs_pad < (1LL << 31) &&
                              a.n_src_pad < (1LL << 31),
                          "far_p2p_batched: segment %d sizes must be positive multiples of %d",
                          si, G);
            TCT_CHECK_ARG(a.loc32 && a.obs_org && a.rec && a.src_org && a.out_partial,
                          "far_p2p_batched: segment %d null pointer", si);
            TCT_CHECK_ARG(a.eps2 >= 0.0 && std::isfinite(a.eps2) && !std::isnan(a.group_extent),
                          "far_p2p_batched: segment %d eps2 / group_extent", si);
            TCT_CHECK_ARG((a.flags & ~1) == 0, "far_p2p_batched: segment %d flags 0x%x", si,
                          (unsigned)a.flags);
            const int n_tiles = (int)(a.n_src_pad / G);
            TCT_CHECK_ARG(a.src_chunks >= 1 && a.src_chunks <= n_tiles,
                          "far_p2p_batched: segment %d src_chunks %d not in [1, %d]", si,
                          a.src_chunks, n_tiles);
            const int64_t ob = ceil_div(a.n_obs_pad, (int64_t)block * obs_per_lane);
            FarSeg &g = t.s[i];
            g.loc32 = a.loc32;
            g.obs_org = a.obs_org;
            g.rec = a.rec;
            g.src_org = a.src_org;
            g.partial = a.out_partial;
            g.sep2 = tile_sep2(a.eps2, a.group_extent, a.flags);
            g.k = kc0;
            g.k.eps2 = (float)a.eps2;
            g.n_obs_pad = (int)a.n_obs_pad;
            g.n_tiles = n_tiles;
            g.chunks = a.src_chunks;
            g.obs_blocks = (int)ob;
            g.blk0 = (int)blocks;
            blocks += ob * a.src_chunks;
            TCT_CHECK_ARG(blocks < (1LL << 31), "far_p2p_batched: grid too large");
        }
        if (blocks == 0) continue;
#define TCT_B(KID, P) launch_f32_batched<KID, P>((unsigned)blocks, block, st, t)
#define TCT_BP(KID)                                     \
    switch (obs_per_lane) {                             \
    case 1: TCT_B(KID, 1); break;                       \
    case 2: TCT_B(KID, 2); break;                       \
    default: TCT_B(KID, 4); break;                      \
    }
        switch (kernel) {
        case TCT_U: TCT_BP(TCT_U); break;
        case TCT_T: TCT_BP(TCT_T); break;
        case TCT_A: TCT_BP(TCT_A); break;
        case TCT_H: TCT_BP(TCT_H); break;
        case TCT_RT: TCT_BP(TCT_RT); break;
        case TCT_RA: TCT_BP(TCT_RA); break;
        default: TCT_BP(TCT_RH); break;
        }
#undef TCT_BP
#undef TCT_B
        TCT_CHECK_LAUNCH("far_p2p_batched");
    }
    return TCT_OK;
}

extern "C" int tct_far_pack_obs(int64_t n, const double *pts, const double *ns, int64_t n_pad,
                                float *loc32, double *org, double *soa64, void *stream) {
    TCT_CHECK_ARG(n > 0 && n_pad >= n && n_pad % G == 0, "pack_obs: bad sizes n=%lld n_pad=%lld",
                  (long long)n, (long long)n_pad);
    TCT_CHECK_ARG(pts && ns && loc32 && org && soa64, "pack_obs: null pointer");
    hipLaunchKernelGGL(pack_obs_kernel, dim3((unsigned)(n_pad / G)), dim3(G), 0,
                       as_stream(stream), n, pts, ns, n_pad, loc32, org, soa64);
    TCT_CHECK_LAUNCH("pack_obs");
    return TCT_OK;
}

extern "C" int tct_far_pack_src_ex(int64_t n, const double *pts, const double *ns, int64_t n_pad,
                                   int rec_words, float *rec32, double *org, double *rec64,
                                   void *stream) {
    TCT_CHECK_ARG(n > 0 && n_pad >= n && n_pad % G == 0, "pack_src: bad sizes n=%lld n_pad=%lld",
                  (long long)n, (long long)n_pad);
    TCT_CHECK_ARG(rec_words == R, "pack_src: rec_words %d (records are %d words)", rec_words, R);
    TCT_CHECK_ARG(pts && ns && rec32 && org && rec64, "pack_src: null pointer");
    hipLaunchKernelGGL(pack_src_kernel, dim3((unsigned)(n_pad / G)), dim3(G), 0,
                       as_stream(stream), n, pts, ns, n_pad, rec_words, rec32, org, rec64);
    TCT_CHECK_LAUNCH("pack_src");
    return TCT_OK;
}

extern "C" int tct_far_pack_src(int64_t n, const double *pts, const double *ns, int64_t n_pad,
                                float *rec32, double *org, double *rec64, void *stream) {
    return tct_far_pack_src_ex(n, pts, ns, n_pad, R, rec32, org, rec64, stream);
}

static int check_point_kernel(int kernel) {
    if (kernel < TCT_U || kernel > TCT_H)
        return set_error(TCT_EUNSUPPORTED, "point kernel id %d not in U/T/A/H", kernel);
    return TCT_OK;
}

extern "C" int tct_far_strength_pts(int kernel, int prec, int64_t n, const double *vec,
                                    const double *params_host, void *rec, void *stream) {
    if (int e = check_point_kernel(kernel)) return e;
    TCT_CHECK_ARG(n >= 0 && vec && rec && params_host, "strength_pts: bad arguments");
    const int64_t n_pad = ceil_div(n, G) * G;
    StrengthConsts sc = strength_consts(kernel, params_host);
    const unsigned nb = (unsigned)ceil_div(n_pad, 256);
    if (nb == 0) return TCT_OK;
    if (prec == TCT_F32)
        hipLaunchKernelGGL(strength_pts_kernel<float>, dim3(nb), dim3(256), 0, as_stream(stream),
                           n, n_pad, vec, sc, (float *)rec);
    else
        hipLaunchKernelGGL(strength_pts_kernel<double>, dim3(nb), dim3(256), 0,
                           as_stream(stream), n, n_pad, vec, sc, (double *)rec);
    TCT_CHECK_LAUNCH("strength_pts");
    return TCT_OK;
}

extern "C" int tct_far_gather_tris(int kernel, int prec, int64_t n_pts, int nq,
                                   const int32_t *tri_of_pt, const double *qphi, const double *wj,
                                   const double *v, const double *params_host, void *rec,
                                   void *stream) {
    if (int e = check_point_kernel(kernel)) return e;
    TCT_CHECK_ARG(n_pts >= 0 && nq > 0 && n_pts % nq == 0, "gather_tris: n_pts %% nq != 0");
    TCT_CHECK_ARG(tri_of_pt && qphi && wj && v && rec && params_host, "gather_tris: null ptr");
    const int64_t n_pad = ceil_div(n_pts, G) * G;
    StrengthConsts sc = strength_consts(kernel, params_host);
    const unsigned nb = (unsigned)ceil_div(n_pad, 256);
    if (nb == 0) return TCT_OK;
    if (prec == TCT_F32)
        hipLaunchKernelGGL(gather_tris_kernel<float>, dim3(nb), dim3(256), 0, as_stream(stream),
                           n_pts, n_pad, nq, tri_of_pt, qphi, wj, v, sc, (float *)rec);
    else
        hipLaunchKernelGGL(gather_tris_kernel<double>, dim3(nb), dim3(256), 0, as_stream(stream),
                           n_pts, n_pad, nq, tri_of_pt, qphi, wj, v, sc, (double *)rec);
    TCT_CHECK_LAUNCH("gather_tris");
    return TCT_OK;
}

extern "C" int tct_far_gather_curl(int kernel, int prec, int64_t n_pts, int nq,
                                   const int32_t *tri_of_pt, const double *qphi, const double *wj,
                                   const double *tri_curl, const double *v,
                                   const double *params_host, void *rec, void *stream) {
    TCT_CHECK_ARG(kernel >= TCT_RT && kernel <= TCT_RH, "gather_curl: kernel %d not RT/RA/RH",
                  kernel);
    TCT_CHECK_ARG(n_pts >= 0 && nq > 0 && n_pts % nq == 0, "gather_curl: n_pts %% nq != 0");
    TCT_CHECK_ARG(tri_of_pt && qphi && wj && tri_curl && v && rec && params_host,
                  "gather_curl: null pointer");
    const int64_t n_pad = ceil_div(n_pts, G) * G;
    const ElasticConsts c = elastic_consts(params_host);
    const unsigned nb = (unsigned)ceil_div(n_pad, 256);
    if (nb == 0) return TCT_OK;
    hipStream_t st = as_stream(stream);
#define TCT_GC(KID)                                                                            \
    if (prec == TCT_F32)                                                                       \
        hipLaunchKernelGGL((gather_curl_kernel<KID, float>), dim3(nb), dim3(256), 0, st, n_pts, \
                           n_pad, nq, tri_of_pt, qphi, wj, tri_curl, v, c, (float *)rec);      \
    else                                                                                       \
        hipLaunchKernelGGL((gather_curl_kernel<KID, double>), dim3(nb), dim3(256), 0, st,      \
                           n_pts, n_pad, nq, tri_of_pt, qphi, wj, tri_curl, v, c, (double *)rec)
    if (kernel == TCT_RT) {
        TCT_GC(TCT_RT);
    } else if (kernel == TCT_RA) {
        TCT_GC(TCT_RA);
    } else {
        TCT_GC(TCT_RH);
    }
#undef TCT_GC
    TCT_CHECK_LAUNCH("gather_curl");
    return TCT_OK;
}

extern "C" int tct_far_p2p(int kernel, int prec, int64_t n_obs_pad, const float *loc32,
                           const double *obs_org, const double *obs_soa64, int64_t n_src_pad,
                           const void *rec, const double *src_org, const double *params_host,
                           double eps2, double group_extent, const tct_tile_cfg *cfg,
                           double *out_partial, void *stream) {
    TCT_CHECK_ARG(kernel >= TCT_U && kernel <= TCT_RH, "far_p2p: kernel id %d", kernel);
    TCT_CHECK_ARG(eps2 >= 0.0 && std::isfinite(eps2), "far_p2p: eps2 must be finite and >= 0");
    TCT_CHECK_ARG(!std::isnan(group_extent), "far_p2p: group_extent is NaN");
    TCT_CHECK_ARG(n_obs_pad > 0 && n_obs_pad % G == 0 && n_src_pad > 0 && n_src_pad % G == 0,
                  "far_p2p: sizes must be positive multiples of %d", G);
    TCT_CHECK_ARG(n_obs_pad < (1LL << 31) && n_src_pad < (1LL << 31), "far_p2p: too many points");
    TCT_CHECK_ARG(rec && out_partial && params_host, "far_p2p: null pointer");
    TCT_CHECK_ARG(prec == TCT_F64 || (loc32 && obs_org && src_org), "far_p2p: f32 layout null");
    TCT_CHECK_ARG(prec == TCT_F32 || obs_soa64, "far_p2p: f64 layout null");
    tct_tile_cfg c{256, 2, 1, 0};
    if (cfg) c = *cfg;
    TCT_CHECK_ARG(c.block == 64 || c.block == 128 || c.block == 256, "far_p2p: block %d", c.block);
    TCT_CHECK_ARG((c.flags & ~7) == 0, "far_p2p: unknown flags 0x%x", (unsigned)c.flags);
    TCT_CHECK_ARG(c.obs_per_lane == 1 || c.obs_per_lane == 2 || c.obs_per_lane == 4,
                  "far_p2p: obs_per_lane %d", c.obs_per_lane);
    const int n_tiles = (int)(n_src_pad / G);
    TCT_CHECK_ARG(c.src_chunks >= 1 && c.src_chunks <= n_tiles && c.src_chunks <= 65535,
                  "far_p2p: src_chunks %d not in [1, %d]", c.src_chunks, n_tiles);
    const int64_t obs_blocks = ceil_div(n_obs_pad, (int64_t)c.block * c.obs_per_lane);
    const bool xcd = (c.flags & 2) != 0;
    TCT_CHECK_ARG(!xcd || c.src_chunks % 8 == 0,
                  "far_p2p: XCD chunk order needs src_chunks a multiple of 8 (got %d)",
                  c.src_chunks);
    TCT_CHECK_ARG(obs_blocks * c.src_chunks < (1LL << 31), "far_p2p: grid too large");
    const int xcd_chunks = xcd ? c.src_chunks : 0;
    const int ob = (int)obs_blocks;
    dim3 grid = xcd ? dim3((unsigned)(obs_blocks * c.src_chunks), 1u)
                    : dim3((unsigned)obs_blocks, (unsigned)c.src_chunks);
    hipStream_t st = as_stream(stream);
    const int no = (int)n_obs_pad;
    // group_extent = E_obs + E_src, E the largest distance of a point from
    // its group origin, so points of groups whose origins are d apart are at
    // least d - group_extent apart; beyond sep (with room for the rounding
    // of the local coordinates) no pair can reach r2 <= eps2 and the tile
    // runs without the test.  The bound uses only the caller's geometry, so
    // any eps2 is safe.  eps2 = 0, a negative group_extent or bit 0 of
    // cfg.flags keeps the test everywhere.
    const double sep2 = tile_sep2(eps2, group_extent, c.flags);
#define TCT_F32_CASE(KID, P)                                                                 \
    launch_f32<KID, P>(grid, c.block, st, no, loc32, obs_org, n_tiles, (const float *)rec,    \
                       src_org, kernel_consts<float>(KID, params_host, eps2), sep2, out_partial, \
                       xcd_chunks, ob, (c.flags & 4) != 0)
#define TCT_F64_CASE(KID, P)                                                              \
    launch_f64<KID, P>(grid, c.block, st, no, obs_soa64, n_tiles, (const double *)rec,     \
                       kernel_consts<double>(KID, params_host, 0.0), out_partial, xcd_chunks, ob)
#define TCT_P_SWITCH(MAC, KID)                          \
    switch (c.obs_per_lane) {                           \
    case 1: MAC(KID, 1); break;                         \
    case 2: MAC(KID, 2); break;                         \
    default: MAC(KID, 4); break;                        \
    }
#define TCT_K_SWITCH(MAC)                               \
    switch (kernel) {                                   \
    case TCT_U: TCT_P_SWITCH(MAC, TCT_U); break;        \
    case TCT_T: TCT_P_SWITCH(MAC, TCT_T); break;        \
    case TCT_A: TCT_P_SWITCH(MAC, TCT_A); break;        \
    case TCT_H: TCT_P_SWITCH(MAC, TCT_H); break;        \
    case TCT_RT: TCT_P_SWITCH(MAC, TCT_RT); break;      \
    case TCT_RA: TCT_P_SWITCH(MAC, TCT_RA); break;      \
    default: TCT_P_SWITCH(MAC, TCT_RH); break;          \
    }
    if (prec == TCT_F32) {
        TCT_K_SWITCH(TCT_F32_CASE)
    } else {
        TCT_K_SWITCH(TCT_F64_CASE)
    }
#undef TCT_K_SWITCH
#undef TCT_P_SWITCH
#undef TCT_F64_CASE
#undef TCT_F32_CASE
    TCT_CHECK_LAUNCH("far_p2p");
    return TCT_OK;
}

extern "C" int tct_far_reduce_pts(int64_t n, int64_t n_obs_pad, int chunks, const double *partial,
                                  double *out, void *stream) {
    TCT_CHECK_ARG(n >= 0 && n <= n_obs_pad && chunks >= 1 && partial && out,
                  "reduce_pts: bad arguments");
    if (n == 0) return TCT_OK;
    hipLaunchKernelGGL(reduce_pts_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                       as_stream(stream), n, n_obs_pad, chunks, partial, out);
    TCT_CHECK_LAUNCH("reduce_pts");
    return TCT_OK;
}

static void sum_chunks(hipStream_t st, int64_t stride, int chunks, double *partial) {
    hipLaunchKernelGGL(chunk_sum_kernel, dim3((unsigned)ceil_div(stride, 256)), dim3(256), 0, st,
                       stride, chunks, partial);
}

extern "C" int tct_far_scatter_tris(int64_t n_tris, int nq, const int32_t *pt_of_tri,
                                    const double *qphi, const double *wj_obs, int64_t n_obs_pad,
                                    int chunks, double *partial, double *out, void *stream) {
    TCT_CHECK_ARG(n_tris >= 0 && nq > 0 && chunks >= 1 && n_tris * nq <= n_obs_pad,
                  "scatter_tris: bad sizes");
    TCT_CHECK_ARG(pt_of_tri && qphi && wj_obs && partial && out, "scatter_tris: null pointer");
    if (n_tris == 0) return TCT_OK;
    hipStream_t st = as_stream(stream);
    if (chunks > 1) sum_chunks(st, 3 * n_obs_pad, chunks, partial);
    hipLaunchKernelGGL(scatter_tris_kernel, dim3((unsigned)ceil_div(n_tris * 9, 256)), dim3(256),
                       0, st, n_tris, nq, pt_of_tri, qphi, wj_obs, n_obs_pad, 1, partial, out);
    TCT_CHECK_LAUNCH("scatter_tris");
    return TCT_OK;
}

extern "C" int tct_far_scatter_curl(int kernel, int64_t n_tris, int nq, const int32_t *pt_of_tri,
                                    const double *qphi, const double *wj_obs,
                                    const double *tri_curl, int64_t n_obs_pad, int chunks,
                                    double *partial, const double *params_host, double *out,
                                    void *stream) {
    TCT_CHECK_ARG(kernel == TCT_RA || kernel == TCT_RH, "scatter_curl: kernel %d not RA/RH",
                  kernel);
    TCT_CHECK_ARG(n_tris >= 0 && nq > 0 && chunks >= 1 && n_tris * nq <= n_obs_pad,
                  "scatter_curl: bad sizes");
    TCT_CHECK_ARG(pt_of_tri && qphi && wj_obs && tri_curl && partial && out && params_host,
                  "scatter_curl: null pointer");
    if (n_tris == 0) return TCT_OK;
    const ElasticConsts c = elastic_consts(params_host);
    const unsigned nb = (unsigned)ceil_div(n_tris, 256);
    hipStream_t st = as_stream(stream);
    if (chunks > 1) sum_chunks(st, 9 * n_obs_pad, chunks, partial);
    if (kernel == TCT_RH)
        hipLaunchKernelGGL(scatter_curl_kernel<TCT_RH>, dim3(nb), dim3(256), 0, st, n_tris, nq,
                           pt_of_tri, qphi, wj_obs, tri_curl, n_obs_pad, 1, partial, c, out);
    else
        hipLaunchKernelGGL(scatter_curl_kernel<TCT_RA>, dim3(nb), dim3(256), 0, st, n_tris, nq,
                           pt_of_tri, qphi, wj_obs, tri_curl, n_obs_pad, 1, partial, c, out);
    TCT_CHECK_LAUNCH("scatter_curl");
    return TCT_OK;
}
