/*
 * tct_hip.h -- C ABI of libtct_hip.so, the MI355X (gfx950) farfield /
 * nearfield hot path of tectosaur.
 *
 * Conventions (all entry points):
 *   - return 0 on success, a negative TCT_E* code on failure; the message of
 *     the last failure on the calling thread is in tct_last_error();
 *   - every pointer marked "dev" is device memory owned by the caller;
 *     every output is caller-allocated; nothing here allocates device memory
 *     except where a comment says so;
 *   - "stream" is a hipStream_t passed as void*; work is enqueued on it and
 *     the call returns without synchronising (except where noted);
 *   - indices are int32 on the device, int64 on the host.
 *
 * Kernel ids follow the reference's kernel names (tectosaur/kernels/
 * __init__.py:20-258).  DOF layout is v[9*tri + 3*basis + component]
 * (tectosaur/matrix_free.cl:53-56); pair blocks are
 * [b_obs*27 + d_obs*9 + b_src*3 + d_src] (tectosaur/assemble.cl:94-97).
 */
#ifndef TCT_HIP_H
#define TCT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TCT_OK 0
#define TCT_EARG (-1)
#define TCT_EHIP (-2)
#define TCT_EUNSUPPORTED (-3)

enum tct_kernel {
    TCT_U = 0,  /* elasticU3  */
    TCT_T = 1,  /* elasticT3  */
    TCT_A = 2,  /* elasticA3  */
    TCT_H = 3,  /* elasticH3  */
    TCT_RT = 4, /* elasticRT3 */
    TCT_RA = 5, /* elasticRA3 */
    TCT_RH = 6  /* elasticRH3 */
};

enum tct_prec { TCT_F32 = 0, TCT_F64 = 1 };

/* Launch shape of the farfield N-body kernel (chosen by the autotuner).
 *   block:        threads per workgroup (64, 128 or 256);
 *   obs_per_lane: observation points held per lane (1, 2 or 4);
 *   src_chunks:   the source points are split into this many contiguous
 *                 chunks, one grid row each; partial sums are written per
 *                 chunk and summed in a fixed order by the reduce/scatter
 *                 step, so results are bit-reproducible;
 *   flags:        bit 0 keeps the FP32 kernel's r2 <= eps2 test on every
 *                 source tile (by default tiles too far from a wave's
 *                 observation groups to hold such a pair skip it; the
 *                 results are the same); bit 1 orders the workgroups by
 *                 XCD (a 1-D grid; src_chunks a multiple of 8; XCD x works
 *                 through chunks x, x + 8, ... for every observation block,
 *                 so one chunk of source records at a time stays in each
 *                 XCD's L2; the same results); bit 2 stages each wave's
 *                 64-point source tile in LDS and reads the records into
 *                 VGPRs (gfx950 dual-issues two VALU instructions only when
 *                 neither reads an SGPR; the same results, faster for H,
 *                 slower where the LDS reads outweigh the gain); other bits
 *                 must be 0. */
typedef struct {
    int32_t block;
    int32_t obs_per_lane;
    int32_t src_chunks;
    int32_t flags;
} tct_tile_cfg;

/* Points per origin group of the FP32 layout (obs and src).  FP32
 * coordinates are stored relative to an FP64 origin per group of this many
 * consecutive points. */
#define TCT_GROUP 64
/* Floats (FP32) / doubles (FP64) per source record:
 * [y0 y1 y2 n0 n1 n2 s0 s1 s2 e0 e1 e2]. */
#define TCT_REC 12

const char *tct_last_error(void);
int tct_abi_version(void);

/* ---------------- farfield: layout packing (one-time per point set) ----
 * Replaces the host->device uploads of tectosaur/farfield.py:26-31 and
 * tectosaur/ops/sparse_farfield_op.py:27-35.
 *
 * tct_far_pack_obs: pts/ns [n][3] f64 (dev) ->
 *   loc32  [6][n_pad] f32: x, y, z relative to the group origin, then the
 *          unit normal (dev, n_pad = round_up(n, TCT_GROUP));
 *   org    [n_pad/TCT_GROUP][3] f64 group origins (mean of the group);
 *   soa64  [6][n_pad] f64 absolute coordinates + normals (for TCT_F64);
 *   padding points are placed far away with zero normal. */
int tct_far_pack_obs(int64_t n, const double *pts, const double *ns, int64_t n_pad,
                     float *loc32, double *org, double *soa64, void *stream);

/* tct_far_pack_src: geometry part of the source records:
 *   rec32 [n_pad][TCT_REC] f32 (y local to its group origin, unit normal),
 *   org   [n_pad/TCT_GROUP][3] f64,
 *   rec64 [n_pad][TCT_REC] f64 (absolute).  Strength fields are zeroed. */
int tct_far_pack_src(int64_t n, const double *pts, const double *ns, int64_t n_pad,
                     float *rec32, double *org, double *rec64, void *stream);

/* tct_far_pack_src_ex: as tct_far_pack_src with the record size given
 * (rec_words must be TCT_REC: every kernel's record, elasticRH3's
 * [y, G(9)] included, is TCT_REC words); words past the normal are zeroed. */
int tct_far_pack_src_ex(int64_t n, const double *pts, const double *ns, int64_t n_pad,
                        int rec_words, float *rec32, double *org, double *rec64, void *stream);

/* ---------------- farfield: per-matvec strengths -------------------------
 * tct_far_strength_pts: point strengths vec [n][3] f64 (the "input" of
 * farfield_pts{K}, tectosaur/farfield_direct.cl:70-72) -> record strength
 * and kernel-specific extra fields, for precision prec. */
int tct_far_strength_pts(int kernel, int prec, int64_t n, const double *vec,
                         const double *params_host, void *rec, void *stream);

/* tct_far_gather_tris: triangle DOF input v[n_tris*9] f64 -> per-quadrature
 * point strengths s = J_t w_q sum_b phi_b(q) v[9t+3b+:] (the interpolation
 * of tectosaur/matrix_free.cl:84-90 with the factor of :83 folded in).
 *   tri_of_pt [n_pts] int32: DOF-space triangle of each source point
 *             (points of one triangle are consecutive, nq of them);
 *   qphi      [nq][3] f64 basis values at the rule's points;
 *   wj        [n_pts] f64: J_t * w_q per point. */
int tct_far_gather_tris(int kernel, int prec, int64_t n_pts, int nq, const int32_t *tri_of_pt,
                        const double *qphi, const double *wj, const double *v,
                        const double *params_host, void *rec, void *stream);

/* tct_far_gather_curl: the surface-curl kernels RT, RA, RH (farfield_tris_to_tris
 * with the vector codes of tectosaur/kernels/kernels.cl:100-206).  Per source
 * point it writes what the point-pair form needs from the source triangle's
 * surface curl S[d][I] = sum_b curl[b][I] v[9t+3b+d] and the interpolated
 * input, scaled by w_q J_t.  tri_curl [n_src_tris][3][3] (basis, component)
 * holds the basis surface curls of each DOF-space source triangle
 * (integral_primitives.cl:95-112). */
int tct_far_gather_curl(int kernel, int prec, int64_t n_pts, int nq, const int32_t *tri_of_pt,
                        const double *qphi, const double *wj, const double *tri_curl,
                        const double *v, const double *params_host, void *rec, void *stream);

/* ---------------- farfield: the N-body kernel ---------------------------
 * Reference: farfield_pts{K} (tectosaur/farfield_direct.cl:15-113), the
 * hot loop of farfield_tris_to_tris{K} (tectosaur/matrix_free.cl:47-101).
 * out_partial [cfg.src_chunks][3][n_obs_pad] f64 (dev): per-chunk sums over
 * the chunk's source points of K(y_j - x_i) s_j, pairs with r2 == 0
 * skipped.  TCT_F32: FP32 evaluation and FP32 accumulation within each
 * 64-point source tile, FP64 accumulation across tiles; pairs with
 * r2 <= eps2 are skipped.  The packing uses eps2 = (8 * 2^-24 *
 * group_extent)^2, a bound on the rounding of one point stored relative to
 * two different group origins, so a point paired with itself is skipped.
 * group_extent = E_obs + E_src, E the largest distance of a point from its
 * packed group origin: source tiles whose origin is farther than
 * group_extent + 5 sqrt(eps2) (with a rounding margin) from the wave's
 * observation group origins hold no pair with r2 <= eps2 and run without
 * the test (same results).  A negative group_extent, eps2 = 0 or cfg->flags
 * bit 0 keeps the test on every tile (eps2 = 0 tests r2 == 0 only).
 * TCT_F64: FP64 throughout, absolute coordinates, exact r2 == 0 test (eps2
 * and group_extent ignored).
 * Kernels U, T, A, H (3 sums per point) and the surface-curl forms RT (3),
 * RA and RH (9 moments per point, applied to the observation curls by
 * tct_far_scatter_curl); out_partial is [src_chunks][3 or 9][n_obs_pad]. */
int tct_far_p2p(int kernel, int prec, int64_t n_obs_pad, const float *loc32,
                const double *obs_org, const double *obs_soa64, int64_t n_src_pad,
                const void *rec, const double *src_org, const double *params_host,
                double eps2, double group_extent, const tct_tile_cfg *cfg,
                double *out_partial, void *stream);

/* Several problems' FP32 N-bodies in one launch (no reference counterpart:
 * the reference launches farfield_tris_to_tris{K} once per operator and
 * matvec, tectosaur/ops/sparse_farfield_op.py:50-59; a server batches the
 * inner matvecs of many solves with this).  Each segment is one
 * tct_far_p2p call's FP32 arguments (layouts from tct_far_pack_obs /
 * tct_far_pack_src_ex and a strength gather; flags bit 0 only); all
 * segments share the kernel, the constants (params_host) and the launch
 * shape (block, obs_per_lane).  A segment's out_partial holds exactly what
 * tct_far_p2p with the same src_chunks and flags writes (bit for bit: the
 * per-point summation order depends only on the chunks). */
typedef struct {
    int64_t n_obs_pad;
    const float *loc32;
    const double *obs_org;
    int64_t n_src_pad;
    const float *rec;
    const double *src_org;
    double eps2;
    double group_extent;
    int32_t src_chunks;
    int32_t flags;
    double *out_partial;
} tct_far_segment;

int tct_far_p2p_batched(int kernel, const double *params_host, int block, int obs_per_lane,
                        int n_seg, const tct_far_segment *segs, void *stream);

/* Sum of the chunk partials: out [n][3] f64 (point order of the packing). */
int tct_far_reduce_pts(int64_t n, int64_t n_obs_pad, int chunks, const double *partial,
                       double *out, void *stream);

/* Galerkin scatter (tectosaur/matrix_free.cl:94-98): out[9t+3b+d] =
 *   sum_q wj_obs[p] phi_b(q) sum_chunks partial[c][d][p],  p = pt_of_tri[t]*nq + q.
 * pt_of_tri [n_tris] int32: first-point position (divided by nq) of each
 * DOF-space obs triangle in the packed order.  With chunks > 1 the chunks
 * are first summed in place into chunk 0 (partial is consumed). */
int tct_far_scatter_tris(int64_t n_tris, int nq, const int32_t *pt_of_tri, const double *qphi,
                         const double *wj_obs, int64_t n_obs_pad, int chunks, double *partial,
                         double *out, void *stream);

/* Galerkin scatter of the 9-moment kernels RA and RH: out[9t+3b+k] from the
 * per-point moments and the observation triangle's basis surface curls
 * tri_curl [n_obs_tris][3][3].  partial is consumed as in
 * tct_far_scatter_tris. */
int tct_far_scatter_curl(int kernel, int64_t n_tris, int nq, const int32_t *pt_of_tri,
                         const double *qphi, const double *wj_obs, const double *tri_curl,
                         int64_t n_obs_pad, int chunks, double *partial,
                         const double *params_host, double *out, void *stream);

/* ---------------- treecode farfield ----------------------------------------
 * Reference: the treecode TSFMM (tectosaur/fmm/tsfmm.py:30-264: p2m :160-177,
 * m2m :178-192, m2p :193-214, p2p :215-240) with the elasticU3/RT3/RA3/RH3
 * expansions of tectosaur/fmm/ts_kernels.cl, used through FMMFarfieldOp
 * (tectosaur/ops/sparse_farfield_op.py:73-118).  Own formulation: Cartesian
 * Taylor moments of 1/r, per-kernel Laplace charges, scaled by the node
 * radius.  The tree is a binary tree over the 64-point Morton tiles of the
 * packed source set (tct_far_pack_src_ex layout): node t < n_tiles is tile
 * t; parents follow their children.
 *   node_cr [n_nodes][4] f64: centre (3) and radius (> 0) of every node;
 *   moments [n_nodes][C][n_terms(order)] in the precision of the records,
 *     C = 4 (U, RA), 13 (RT), 12 (RH), n_terms(p) = (p+1)(p+2)(p+3)/6.
 * Kernels: TCT_U, TCT_RT, TCT_RA, TCT_RH; orders 0..TCT_TREE_MAX_ORDER for
 * p2m/m2m, and 2-8 for eval (others return TCT_EUNSUPPORTED). */
#define TCT_TREE_MAX_ORDER 8

/* Tile moments (nodes 0..n_tiles-1) from the packed source records (the
 * first n_src are real points, the rest padding): FP32 records are local to
 * src_org per tile, FP64 records absolute (src_org unused). */
int tct_tree_p2m(int kernel, int prec, int order, int64_t n_src, int64_t n_tiles, const void *rec,
                 const double *src_org, const double *node_cr, void *moments, void *stream);

/* One upward level: parents [p_off, p_off + n_par) from children
 * c_off + 2i and c_off + 2i + 1 (the latter when 2i + 1 < n_child). */
int tct_tree_m2m(int kernel, int prec, int order, int p_off, int n_par, int c_off, int n_child,
                 const double *node_cr, void *moments, void *stream);

/* Evaluation over interaction-list segments: list int32 entries e >= 0
 * (m2p from node e) or e < 0 (p2p with source tile -(e + 1)); seg
 * [n_seg][3] int32 = (observation tile, first entry, end entry), one
 * wavefront each.  obs = loc32 (FP32, with obs_org) or soa64 (FP64);
 * rec/src_org as for tct_far_p2p.  seg_partial [n_seg][3 or 9][64] f64. */
int tct_tree_eval(int kernel, int prec, int order, int n_obs_pad, const void *obs,
                  const double *obs_org, const void *rec, const double *src_org,
                  const double *node_cr, const void *moments, int n_seg, const int *seg,
                  const int *list, const double *params_host, double eps2, double *seg_partial,
                  void *stream);

/* out_partial [3 or 9][n_obs_pad] = sum of each observation tile's segments
 * [seg_off[t], seg_off[t + 1]) in order (the layout of tct_far_p2p with one
 * chunk, for tct_far_reduce_pts / scatter_tris / scatter_curl). */
int tct_tree_reduce(int kernel, int n_obs_tiles, int n_obs_pad, const int *seg_off,
                    const double *seg_partial, double *out_partial, void *stream);

/* ---------------- nearfield pair integrator (FP64) ----------------------
 * Reference: single_pairs{Z,N} / single_pairsN_adj (tectosaur/assemble.cl
 * :24-103) driven by PairsIntegrator (tectosaur/nearfield/
 * pairs_integrator.py:27-93).
 *   qx [nq][4] f64 (obs xhat, obs yhat, src xhat, src yhat), qw [nq] f64;
 *   pts [n_pts][3] f64, tris [n_tris][3] int32;
 *   pairs [n_pairs][pair_width] int32, pair_width 2 (obs, src) or 5
 *         (obs, src, obs_clicks, src_clicks, src_flip) -- width 5 rotates
 *         and de-rotates exactly as assemble.cl:77-101;
 *   check0: the r = 0 test (the "Z" variant): 0 none; 1 skips points with
 *           r2 == 0 and, on pairs that are one triangle in one frame, the
 *           points with equal obs and src parameters (where the positions
 *           coincide exactly); 2 skips r2 == 0 only, for rules the caller
 *           knows to have no point of equal obs and src parameters;
 *   out [n_pairs][81] f64.
 * lanes_per_pair (1, 4, 16 or 64) sets how many lanes share one pair's
 * quadrature points; 0 picks from nq and n_pairs (the fewest lanes that
 * still fill the device; for elasticRH3 with nq >= 512, where that would be
 * several lanes, one lane per pair and rule chunk instead); -C (elasticRH3,
 * C <= nq, and 1 <= C <= 11, C a multiple of 4 up to 44 or a multiple of 8
 * up to 88) runs one lane per pair and each of C rule chunks, their sums
 * added in chunk order (out rows serve as the chunks' workspace, 11 slots
 * of 7 values; beyond 11 chunks the four or eight waves of a workgroup take
 * as many consecutive chunks of the same 64 pairs and add them in LDS
 * before they fill a slot). */
int tct_pairs_integrate(int kernel, int64_t nq, const double *qx, const double *qw,
                        const double *pts, const int32_t *tris, int64_t n_pairs,
                        const int32_t *pairs, int pair_width, int check0,
                        const double *params_host, int lanes_per_pair, double *out,
                        void *stream);

/* out[i] = a[i] - b[i] over n doubles (nearfield correction, tectosaur/
 * nearfield/nearfield_op.py:135-141). */
int tct_sub_f64(int64_t n, const double *a, const double *b, double *out, void *stream);

/* ---------------- interior evaluation (FP64) -----------------------------
 * Point observers against source triangles, for InteriorOp
 * (tectosaur/interior.py:29-119).
 *
 * tct_interior_pairs: interior_pairs (tectosaur/assemble.cl:140-186) via
 * integrate_pt_tri (tectosaur/integral_primitives.cl:216-254), called by
 * interior_pairs_quad (tectosaur/interior.py:121-142).
 *   kernel: U, T, A, H or RT (no observation surface curl);
 *   qx [nq][2] f64 (src xhat, yhat), qw [nq] f64;
 *   obs_pts, obs_ns [n_obs][3] f64; src_pts [n][3] f64, src_tris [n][3] int32;
 *   pairs [n_pairs][3] int32 = (obs point, src triangle, src clicks);
 *   check0: skip r2 == 0 (the point-farfield form; interior_pairs does not);
 *   finite_part: write only column b_src = clicks (others zero);
 *   out [n_pairs][27] f64, index d_obs*9 + b_src*3 + d_src, times J_src.
 * lanes_per_pair (1, 4, 16 or 64) or 0 to pick from nq. */
int tct_interior_pairs(int kernel, int64_t nq, const double *qx, const double *qw,
                       const double *obs_pts, const double *obs_ns, const double *src_pts,
                       const int32_t *src_tris, int64_t n_pairs, const int32_t *pairs,
                       int check0, int finite_part, const double *params_host,
                       int lanes_per_pair, double *out, void *stream);

/* tct_interior_pairs_sinh: as tct_interior_pairs for near-singular pairs,
 * with a polar sinh rule about the point of the triangle closest to the
 * projection of the observation point (no reference counterpart: the
 * reference uses a fixed Gauss rule, which loses accuracy as the point
 * nears the triangle).  tq, tw [n_t] and rq, rw [n_r]: Gauss rules on
 * [0, 1] for the angle and the sinh variable (6 n_t n_r points per pair).
 * pairs and out as tct_interior_pairs (physical weights, J included). */
int tct_interior_pairs_sinh(int kernel, int n_t, const double *tq, const double *tw, int n_r,
                            const double *rq, const double *rw, const double *obs_pts,
                            const double *obs_ns, const double *src_pts,
                            const int32_t *src_tris, int64_t n_pairs, const int32_t *pairs,
                            const double *params_host, double *out, void *stream);

/* tct_interior_corners: interior_corners (tectosaur/interior_corners.cl
 * :23-175), called by InteriorOp.build_vertex_mat (tectosaur/interior.py
 * :66-95): observation point at vertex `clicks` of the source triangle.
 *   qt, qw [nq]: the angle rule on [0, pi/2]; pairs [n_pairs][3] int32 =
 *   (obs point, src triangle, vertex); out [n_pairs][27] f64 as above.
 * Uses elasticH3's constants whatever the caller's kernel, as that file
 * does. */
int tct_interior_corners(int64_t nq, const double *qt, const double *qw, const double *obs_ns,
                         const double *src_pts, const int32_t *src_tris, int64_t n_pairs,
                         const int32_t *pairs, const double *params_host, double *out,
                         void *stream);

/* tct_bsr39_spmv: y[3i+d] (+)= sum_k data[k][d][j] x[9 indices[k] + j] over
 * the 3 x 9 blocks of block-row i -- the pair matrices of make_pairs_mat
 * (tectosaur/interior.py:14-26) applied in InteriorOp.dot (:108-114).
 * indptr [n_block_rows+1] int32, indices [n_blocks] int32,
 * data [n_blocks][27] f64. */
int tct_bsr39_spmv(int64_t n_block_rows, const int32_t *indptr, const int32_t *indices,
                   const double *data, const double *x, double *y, int accumulate,
                   void *stream);

/* tct_limit_combine: out[i][c] = sum_k w[i][k] vals[k][i][c] -- the
 * extrapolation to the surface of a field sampled at n_eps offsets along
 * each point's normal (w: per-point least-squares fit weights of the
 * polynomial + log model of tectosaur tests/richardson_error.py:4-21).
 * w [n_pts][n_eps], vals [n_eps][n_pts][comps], out [n_pts][comps], f64.
 * No reference counterpart in the library (the fit exists there only as
 * that test script). */
int tct_limit_combine(int n_eps, int64_t n_pts, int comps, const double *w, const double *vals,
                      double *out, void *stream);

/* ---------------- block sparse matvec (FP64) -----------------------------
 * Reference: bsrmv9 / bcoomv9 (tectosaur/util/fast_sparse.cpp:71-163).
 * y[9r+i] (+)= sum_k A_k[i][j] x[9 col_k + j] over the blocks of block-row r.
 * indptr [n_block_rows+1] int32, indices [n_blocks] int32,
 * data [n_blocks][81] f64.  accumulate != 0 adds into y. */
int tct_bsr9_spmv(int64_t n_block_rows, const int32_t *indptr, const int32_t *indices,
                  const double *data, const double *x, double *y, int accumulate,
                  void *stream);

/* As tct_bsr9_spmv with the block data stored in FP32 (accumulation and
 * vectors FP64): half the bytes, for operators used as an inner (inexact)
 * operator of mixed-precision refinement.  No reference counterpart (the
 * reference's float_type=float32 nearfield, tectosaur/util/sparse.py:29-33,
 * is the same idea on the host). */
int tct_bsr9_spmv_f32(int64_t n_block_rows, const int32_t *indptr, const int32_t *indices,
                      const float *data, const double *x, double *y, int accumulate,
                      void *stream);

/* CSR SpMV y (+)= A x, FP64 (the constraint matrix cm and cm^T of
 * tectosaur/constraints.py:10-17 applied each GMRES iteration). */
int tct_csr_spmv(int64_t n_rows, const int32_t *indptr, const int32_t *indices,
                 const double *data, const double *x, double *y, int accumulate, void *stream);

/* ---------------- Krylov basis (device GMRES) -----------------------------
 * The Arnoldi step of the GMRES the reference calls on the host
 * (scipy.sparse.linalg.gmres from tectosaur/simple_solver.py:14-34): CGS2
 * against V [m][n] (row stride ld >= n), every reduction in a fixed order
 * and no library workspace (solves on different streams share nothing).
 * tct_krylov_scratch: doubles of scratch for basis size up to m.
 * tct_cgs2: w <- w - V^T h1, then again with h2; h [m + 1] = (h1 + h2, |w|);
 *   vnext (and vnext2), when not NULL, receive w / |w| (the next basis
 *   vector, without a host round trip).
 * tct_krylov_combine: out <- beta out + alpha V[:m]^T c. */
int tct_krylov_scratch(int m, int64_t n, int64_t *n_doubles);
int tct_cgs2(int m, int64_t n, int64_t ld, const double *V, double *w, double *h, double *scratch,
             int64_t scratch_doubles, double *vnext, double *vnext2, void *stream);
int tct_krylov_combine(int m, int64_t n, int64_t ld, const double *V, const double *c,
                       double alpha, double beta, double *out, void *stream);

/* tct_cgs2 without the scratch argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The partial sums use
 * a static device array of TCT_CGS2_NOALLOC_MAX_DOUBLES doubles (32 MiB,
 * held by each device the library runs on until the process ends); a call
 * whose tct_krylov_scratch(m, n) exceeds it returns TCT_EARG.  Nothing is
 * allocated and nothing needs clearing; results are bit-identical to
 * tct_cgs2.  Calls that overlap on one device (concurrent streams or
 * threads) share the array and race: serialise them. */
#define TCT_CGS2_NOALLOC_MAX_DOUBLES (1LL << 22)
int tct_cgs2_noalloc(int m, int64_t n, int64_t ld, const double *V, double *w, double *h,
                     double *vnext, double *vnext2, void *stream);

/* ---------------- measurement ---------------------------------------------
 * tct_calib_valu: VALU issue-rate calibration (no reference counterpart; used
 * by bench.py --valu-calibration).  blocks x 256 lanes each run 16
 * independent pairs of dependent FMA chains for iters steps (32 iters
 * FMA instructions per lane); out [blocks * 256] f32 receives each lane's
 * sum.  mode 0: FP32, every operand in a VGPR; 1: FP32, one operand a
 * kernel argument (SGPR); 2 and 3: the same in FP64. */
int tct_calib_valu(int mode, int blocks, int iters, float *out, void *stream);

/* Serving-loop plumbing (no reference counterpart: the reference's solves
 * are one at a time on the host).
 * tct_arnoldi_step: one Arnoldi step of a solve whose operator application
 *   is a captured graph: launch graph_exec (a hipGraphExec_t that reads the
 *   graph input buffer and writes w) on stream, then tct_cgs2.
 * tct_tick_gather: the results of one tick of many solves.  main_stream
 *   first waits for each streams[i] (through events[i], re-recorded there),
 *   then the n_seg device segments src[k][0, len[k]) are concatenated into
 *   staging (staging_doubles long) and copied to host (pinned) memory, and
 *   done_event (may be NULL) is recorded on main_stream. */
int tct_arnoldi_step(const void *graph_exec, int m, int64_t n, int64_t ld, const double *V,
                     double *w, double *h, double *scratch, int64_t scratch_doubles,
                     double *vnext, double *vnext2, void *stream);
int tct_tick_gather(int n_streams, void *const *streams, void *const *events, int n_seg,
                    const double *const *src, const int *len, double *staging,
                    int64_t staging_doubles, double *host, void *main_stream, void *done_event);

/* tct_set_host_threads: OpenMP threads for the host utilities below when
 * called from the calling thread (no reference counterpart; the reference's
 * OpenMP uses its default). */
int tct_set_host_threads(int n);

/* ---------------- host-side mesh utilities (CPU, OpenMP) -----------------
 * Reference: fast_find_nearfield.get_nearfield + split_adjacent_close
 * (tectosaur/mesh/fast_find_nearfield.cpp:22-217).
 * tct_find_near_pairs: all (i, j) with |c_i - c_j|^2 <= ((r_i + r_j) thr)^2
 * (c: centroids [n][3], r: bounding radii).  *n_out receives the pair
 * count; the pairs (sorted by i, then j) are written to out [cap][2] int64
 * only when cap >= *n_out, otherwise nothing is written and the caller
 * retries with a larger buffer. */
int tct_find_near_pairs(int64_t n_a, const double *ca, const double *ra, int64_t n_b,
                        const double *cb, const double *rb, double threshold, int64_t *out,
                        int64_t cap, int64_t *n_out);

/* ---------------- device near-pair finder and classifier ----------------
 * Reference: fast_find_nearfield.get_nearfield and split_adjacent_close
 * (tectosaur/mesh/fast_find_nearfield.cpp:22-182), called through
 * find_close_or_touching (tectosaur/mesh/find_near_adj.py:29-35).  Same pair
 * set and order as tct_find_near_pairs (bit-identical predicate).
 *
 * tct_near_grid_params (host): grid over the B centroids cb [n_b][3] (host)
 *   for a largest reach reach_max = thr (rmax_a + rmax_b) 1.0000001:
 *   lo[3], cell, dims[3].
 * tct_near_pairs_dev_scratch (host): scratch bytes for n_a, n_b and dims.
 * tct_near_pairs_dev_count: ca, ra [n_a][3]/[n_a], cb, rb (device f64);
 *   builds the grid in scratch and writes offsets [n_a + 1] int64 (device;
 *   offsets[n_a] = number of pairs).
 * tct_near_pairs_dev_fill: same arguments and scratch, after _count on the
 *   same stream: out [offsets[n_a]][2] int64 (device), sorted by (i, j).
 * tct_split_adjacent_close_dev: pairs [n][2] int64, tris_a/tris_b [.][3]
 *   int64 (device) -> near [n][2], va [n][4], ea [n][6] (capacity n) and
 *   counts [3] int64 (device), in input order; coincident pairs dropped.
 *   Scratch size from tct_split_adjacent_dev_scratch. */
int tct_near_grid_params(int64_t n_b, const double *cb, double reach_max, double *lo,
                         double *cell, int64_t *dims);
int tct_near_pairs_dev_scratch(int64_t n_a, int64_t n_b, const int64_t *dims, size_t *bytes);
int tct_near_pairs_dev_count(int64_t n_a, const double *ca, const double *ra, int64_t n_b,
                             const double *cb, const double *rb, double thr, double rmax_b,
                             const double *lo, double cell, const int64_t *dims, void *scratch,
                             size_t scratch_bytes, int64_t *offsets, void *stream);
int tct_near_pairs_dev_fill(int64_t n_a, const double *ca, const double *ra, int64_t n_b,
                            const double *cb, const double *rb, double thr, double rmax_b,
                            const double *lo, double cell, const int64_t *dims, void *scratch,
                            size_t scratch_bytes, const int64_t *offsets, int64_t *out,
                            void *stream);
int tct_split_adjacent_dev_scratch(int64_t n_pairs, size_t *bytes);
int tct_split_adjacent_close_dev(int64_t n_pairs, const int64_t *pairs, const int64_t *tris_a,
                                 const int64_t *tris_b, int64_t *near, int64_t *va, int64_t *ea,
                                 int64_t *counts, void *scratch, size_t scratch_bytes,
                                 void *stream);

/* tct_split_adjacent_close: classify near pairs by shared vertices
 * (tectosaur/mesh/fast_find_nearfield.cpp:178-217).  pairs [n][2] index
 * tris_a / tris_b ([*][3] int64).  Writes, in input order, near [*][2]
 * (no shared vertex), va [*][4] (i, j, corner_i, corner_j) and ea [*][6]
 * (i, j, two corner pairs); coincident pairs are dropped.  Each output
 * needs room for n_pairs rows; counts[3] receives the three sizes. */
int tct_split_adjacent_close(int64_t n_pairs, const int64_t *pairs, const int64_t *tris_a,
                             const int64_t *tris_b, int64_t *near, int64_t *va, int64_t *ea,
                             int64_t *counts);

/* ---------------- host-side constraint matrix (CPU) ---------------------
 * Reference: build_constraint_matrix / reduce_constraints
 * (tectosaur/fast_constraints.cpp:120-300, tectosaur/constraints.py:10-17).
 * Constraint i is sum_k term_vals[.] u[term_dofs[.]] = rhs_i with n_terms[i]
 * terms taken in order.  Writes u = cm x + rhs_mat rhs as COO triplets:
 * cm (cm_rows, cm_cols, cm_vals; *cm_n entries) and rhs_mat (r_rows, r_cols,
 * r_vals; *r_n entries); *n_eliminated dofs are removed (cm has
 * n_total_dofs - *n_eliminated columns).  If cap_cm < *cm_n or cap_r < *r_n
 * nothing is written and the caller retries with the reported sizes. */
int tct_build_constraint_matrix(int64_t n_cs, const int64_t *n_terms, const int64_t *term_dofs,
                                const double *term_vals, int64_t n_total_dofs, int64_t cap_cm,
                                int64_t *cm_rows, int64_t *cm_cols, double *cm_vals,
                                int64_t *cm_n, int64_t cap_r, int64_t *r_rows, int64_t *r_cols,
                                double *r_vals, int64_t *r_n, int64_t *n_eliminated);

#ifdef __cplusplus
}
#endif

#endif /* TCT_HIP_H */
