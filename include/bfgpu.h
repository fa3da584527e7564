/*
 * bfgpu.h -- C ABI of the MI355X (gfx950) butterfly far-field library.
 *
 * The boundary replaces the Julia call sites of ButterflyFactorizations.jl
 * listed beside each entry point.  All entry points return an int status
 * (BFGPU_OK = 0); on failure bfgpu_last_error() returns a message for the
 * calling thread.  Indices are 0-based.  Complex data is interleaved
 * (re, im) in either complex FP32 (precision 1) or complex FP64 (precision 2).
 * Matrices are column-major.  Host buffers are borrowed for the duration of
 * a call; device memory belongs to the handle that allocated it.  All GPU
 * work of a context runs on one HIP stream; calls on one context must be
 * serialised by the caller.
 */
#ifndef BFGPU_H
#define BFGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    BFGPU_OK = 0,
    BFGPU_ERR_ARG = 1,      /* invalid argument (bad size, null handle, ...) */
    BFGPU_ERR_HIP = 2,      /* a HIP runtime call failed */
    BFGPU_ERR_NOMEM = 3,    /* device or host allocation failed */
    BFGPU_ERR_STATE = 4,    /* object in the wrong state for the call */
    BFGPU_ERR_NUMERIC = 5   /* numerical failure (e.g. singular R11) */
};

enum { BFGPU_C64 = 1, BFGPU_C128 = 2 };                        /* precision */
enum { BFGPU_OP_N = 0, BFGPU_OP_T = 1, BFGPU_OP_C = 2 };       /* y = op(A) x */
enum { BFGPU_TREE_BALANCED = 0, BFGPU_TREE_MAXPOINTS = 1 };    /* tree split rule */
enum { BFGPU_RANK_BUTTERFLY = 0, BFGPU_RANK_GEOMETRIC = 1, BFGPU_RANK_LEARNED = 2 };
enum { BFGPU_XY_HOST = 0, BFGPU_XY_DEVICE = 1 };               /* matvec flags */

typedef struct bfgpu_ctx bfgpu_ctx;
typedef struct bfgpu_tree bfgpu_tree;
typedef struct bfgpu_bf bfgpu_bf;
typedef struct bfgpu_mlp bfgpu_mlp;
typedef struct bfgpu_near bfgpu_near;
typedef struct bfgpu_kernel bfgpu_kernel;

/* Message of the last failed call on this thread ("" if none). */
const char *bfgpu_last_error(void);
/* Library version string. */
const char *bfgpu_version(void);

/* ---- context (device + stream) ---- */
int bfgpu_ctx_create(int device, bfgpu_ctx **out);
int bfgpu_ctx_destroy(bfgpu_ctx *ctx);
/* Launch on an external stream (e.g. torch's current stream); NULL is the
 * null (default) stream, which torch's default stream is.  reset_stream
 * returns to the context's own (non-blocking) stream. */
int bfgpu_ctx_set_stream(bfgpu_ctx *ctx, void *hip_stream);
int bfgpu_ctx_reset_stream(bfgpu_ctx *ctx);
int bfgpu_ctx_get_stream(bfgpu_ctx *ctx, void **hip_stream);
int bfgpu_ctx_sync(bfgpu_ctx *ctx);

/* ---- cluster trees ----
 * The tree contract of src/treeinterface.jl:7-164 (cluster_center,
 * cluster_radius = halfsize*sqrt(3) as ext/BFH2Trees/BFH2Trees.jl:15-16,
 * cluster_children, cluster_isleaf, cluster_root = node 0, cluster_values,
 * cluster_parent).  Node ids are breadth-first; children are contiguous;
 * values(node) = perm[val_begin[node] .. val_end[node]). */
typedef struct {
    int32_t n_nodes;
    int32_t n_values;
    const int32_t *parent;
    const int32_t *first_child;
    const int32_t *n_children;
    const int32_t *val_begin;
    const int32_t *val_end;
    const double *center;   /* 3 per node */
    const double *halfsize; /* per node */
    const int32_t *perm;    /* n_values */
    const int32_t *level;   /* depth per node */
    const double *points;   /* 3 per value, original point order */
} bfgpu_tree_view;

/* Build a 2^dims-ary box tree (dims 2: quadtree splitting x,y; dims 3:
 * octree) over n points (xyz, 3 per point).  BALANCED subdivides every node
 * to depth ceil(log_{2^dims}(n / leaf_size)); MAXPOINTS subdivides while a
 * node holds more than leaf_size points.  Empty children are not created.
 * Stands in for the H2Trees TwoNTree the reference's tests build. */
int bfgpu_tree_build(const double *xyz, int64_t n, int32_t dims, int32_t leaf_size,
                     int32_t mode, bfgpu_tree **out);
/* Adopt a tree built elsewhere (e.g. exported from H2Trees); arrays copied. */
int bfgpu_tree_create(const double *xyz, int32_t n_values, int32_t n_nodes,
                      const int32_t *parent, const int32_t *first_child,
                      const int32_t *n_children, const int32_t *val_begin,
                      const int32_t *val_end, const double *center, const double *halfsize,
                      const int32_t *perm, bfgpu_tree **out);
int bfgpu_tree_view_get(const bfgpu_tree *t, bfgpu_tree_view *out);
int bfgpu_tree_destroy(bfgpu_tree *t);

/* ---- interaction lists ----
 * nearandfar (src/Matrixassembly/intlists.jl:30-217) with isFarFunctor
 * (kind 0, param alpha; ext/BFH2Trees/h2admissible.jl:11-40) or
 * CenterDistanceAdmissibility (kind 1, param beta; :74-91).  far/near
 * receive (observer node, source node) pairs; pass NULL to count only. */
int bfgpu_nearandfar(const bfgpu_tree *test, const bfgpu_tree *trial, int32_t kind,
                     double param, int32_t minbflvl, int32_t unbalancedints,
                     int32_t leafcomp, int32_t leafimbalance, int32_t *far, int64_t *nfar,
                     int32_t *near, int64_t *nnear);

/* ---- near field ----
 * Near-field blocks of the global operator: for every (observer node,
 * source node) pair, the dense block G(vals(O), vals(S)) assembled on the
 * device in `precision` (replaces the near assembly of
 * src/Matrixassembly/petrovgalerkinbf.jl:55-119, a BlockSparseMatrix of
 * blocks filled by the :near kernel matrix).  Coincident points (r = 0)
 * take *self_value (complex, 2 doubles; NULL: 0), the point kernel's
 * value on the diagonal.  Both trees must carry points.  The operator is
 * rows(test) x rows(trial), like the far field. */
int bfgpu_near_assemble(bfgpu_ctx *ctx, const bfgpu_tree *test, const bfgpu_tree *trial, double k,
                        int32_t precision, const double *self_value, int64_t n_pairs,
                        const int32_t *pairs, bfgpu_near **out);
/* The near field of a kernel-matrix plug-in: blocks filled by the plug-in's
 * near-field quadrature (bfgpu_kernel_rwg_set_near must have been called);
 * the trees' values index its test / trial functions.  Otherwise as
 * bfgpu_near_assemble; its adjoint / transpose products read the stored
 * blocks (two-pass), since the quadrature is not symmetric under the swap. */
int bfgpu_near_assemble_kernel(bfgpu_ctx *ctx, const bfgpu_tree *test, const bfgpu_tree *trial,
                               const bfgpu_kernel *kern, double k, int32_t precision, int64_t n_pairs,
                               const int32_t *pairs, bfgpu_near **out);
/* y = alpha * op(N) * x + beta * y (the near mul! of
 * src/matrixalgebra/matrixvector.jl:31-37 and its adjoint/transpose,
 * src/matrixalgebra/matrix_adj_tr.jl); arguments as bfgpu_matvec. */
int bfgpu_near_matvec(bfgpu_near *n, int32_t op, int64_t nrhs, const void *x, int64_t ldx, void *y,
                      int64_t ldy, const double *alpha, const double *beta, int32_t flags);
/* 8 int64: rows, cols, pairs, items (observer nodes), stored entries,
 * device bytes, precision, 1 when the product writes y in one pass. */
int bfgpu_near_info(const bfgpu_near *n, int64_t *info);
/* Block of pair `pair` (input order): mn receives (m, n); data (may be
 * NULL) receives the m x n column-major block in the stored precision. */
int bfgpu_near_block(const bfgpu_near *n, int64_t pair, int32_t *mn, void *data);
/* Transposed / adjoint products: on (default) applies them through a
 * transposed copy of the blocks (the kernel is symmetric, G(y,x) = G(x,y)),
 * assembled on the first such product when device memory allows (the copy
 * doubles the near field's memory); off frees the copy and uses the
 * two-pass product over the stored blocks.  No counterpart in the reference,
 * whose adjoint views re-read the same blocks (matrix_adj_tr.jl). */
int bfgpu_near_set_transposed_copy(bfgpu_near *n, int32_t on);
/* Complex FP32 products of at least min_rhs right-hand sides run on the
 * matrix cores (default 8; 0: vector ALU only).  Applies when every row is
 * produced by one item (the leaf case).  No counterpart in the reference. */
int bfgpu_near_set_mfma(bfgpu_near *n, int32_t min_rhs);
int bfgpu_near_destroy(bfgpu_near *n);

/* ---- kernel-matrix plug-in: EFIE on RWG functions ----
 * The Maxwell single layer between RWG test and trial functions, far
 * interactions by double numerical quadrature (the BEAST plug-in of
 * ext/BFBEAST/kernelmatrix.jl:22-82 with DoubleNumQStrat):
 *   Z[m,n] = sum_a sum_b G(x_a, y_b) (alpha f_a . g_b + beta d_a e_b),
 * G = exp(-ik|x-y|)/(4 pi |x-y|), k from bfgpu_opts.k.  test_samples holds,
 * per test function, s_test samples of 7 doubles (x, y, z, w A f_m(x) (3),
 * w A div f_m): the quadrature points of both triangles with weights w and
 * areas A folded in; trial_samples likewise.  alpha, beta: complex (2
 * doubles).  Pass the handle as bfgpu_opts.kernel; the trees' values are
 * function indices. */
int bfgpu_kernel_rwg_create(bfgpu_ctx *ctx, const double *alpha, const double *beta, int64_t n_test,
                            int32_t s_test, const double *test_samples, int64_t n_trial, int32_t s_trial,
                            const double *trial_samples, bfgpu_kernel **out);
/* One m x n block Z[rows, cols] (column-major, `precision`) computed on the
 * device by the plug-in's fill kernel; for checks. */
int bfgpu_kernel_fill(const bfgpu_kernel *kern, double k, int32_t precision, int32_t m, const int32_t *rows,
                      int32_t n, const int32_t *cols, void *Z);
/* Near-field quadrature of the plug-in, the :near kernel matrix of
 * ext/BFBEAST/kernelmatrix.jl:26-36 (BEAST.defaultquadstrat; BEAST is absent,
 * so the scheme is this restatement): per pair of triangles of a test and a
 * trial function, when their centroids are closer than near_factor times
 * the longer of their longest edges, singularity extraction -- the static
 * kernel 1/(4 pi R) integrated in closed form over the trial triangle
 * (Wilton, Rao, Glisson, Schaubert, Al-Bundak and Butler, IEEE Trans. AP
 * 32(3), 1984), the rest (exp(-ikR) - 1)/(4 pi R) by the triangle rule of
 * degree near_inner, the outer integral by the rule of degree near_outer --
 * otherwise the double rule of degrees (far_outer, far_inner); degrees 1-5.
 * test_geom / trial_geom: per function, for its plus then minus triangle,
 * 13 doubles: the three vertices, the free vertex v and c with f(x) =
 * c (x - v) on that triangle (so div f = 2 c). */
int bfgpu_kernel_rwg_set_near(bfgpu_kernel *kern, const double *test_geom, const double *trial_geom,
                              int32_t near_outer, int32_t near_inner, int32_t far_outer, int32_t far_inner,
                              double near_factor);
/* One m x n block of the near-field quadrature (column-major, `precision`),
 * computed on the device; for checks. */
int bfgpu_kernel_fill_near(const bfgpu_kernel *kern, double k, int32_t precision, int32_t m, const int32_t *rows,
                           int32_t n, const int32_t *cols, void *Z);
int bfgpu_kernel_destroy(bfgpu_kernel *kern);

/* ---- factorization ---- */
typedef struct {
    double k;            /* wavenumber of G(x,y) = exp(-ik|x-y|)/(4 pi |x-y|) */
    double tol;          /* tau; each compression uses tau / L (bfassembly.jl:74) */
    int32_t adaptive;    /* adaptive row doubling (compressors.jl:134-151) */
    int32_t rank_mode;   /* BFGPU_RANK_* */
    double c_tau;        /* ButterflyRankEstimator constant (default 4.0) */
    int32_t rmin;        /* estimator floor (default 10) */
    double c_geo, c_eps; /* GeometricRankEstimator constants */
    int32_t precision;   /* BFGPU_C64 or BFGPU_C128 */
    int32_t oversample;  /* rows added to a learned rank prediction */
    uint64_t seed;       /* row-sample seed */
    const bfgpu_mlp *mlp;/* rank predictor for BFGPU_RANK_LEARNED */
    int64_t workspace_bytes; /* most device bytes of kernel-matrix tiles per
                                batch of compressions (0: min(32 GiB, 30 % of
                                the free device memory)) */
    const bfgpu_kernel *kernel; /* kernel-matrix plug-in (NULL: the Helmholtz
                                   point kernel above); tree values index its
                                   test / trial functions */
    /* Residual judge of BFGPU_RANK_LEARNED compressions (no counterpart in the
     * reference, whose adaptive loop compressors.jl:124-151 is the fallback
     * it triggers): after the pivoted QR accepts a compression whose row
     * sample was sized by the prediction, its interpolative decomposition T
     * is checked on the next judge_rows rows of the same sample prefix.  If
     * max_j |Zc[:,j] - Zc[:,skel] T[:,j]| > judge_slack * eps * max_j
     * |Zc[:,j]| (the pivoted QR's own stopping rule on rows it never saw),
     * the compression is redone by plain adaptive RRQR from
     * max(ceil(c_tau ln(1/eps)), 10, 2 * rows) rows.  Compressions of more
     * than 256 columns, or whose sample holds every observer, are not
     * judged.  judge_rows 0: no judge; judge_slack <= 0: 4. */
    int32_t judge_rows;
    double judge_slack;
} bfgpu_opts;

/* assemble_BF (src/ButterflyFactorization/bfassembly.jl:48-200) for n_pairs
 * admissible blocks (observer root, source root) at once; n_pairs = 1 is
 * the single-block constructor, several pairs form the far field of
 * PetrovGalerkinBF (src/Matrixassembly/petrovgalerkinbf.jl:121-161).
 * Kernel assembly, pivoted QR and interpolation matrices run on the GPU. */
int bfgpu_factorize(bfgpu_ctx *ctx, const bfgpu_tree *test, const bfgpu_tree *trial,
                    int32_t n_pairs, const int32_t *pairs, const bfgpu_opts *opts,
                    bfgpu_bf **out);

/* Build a device factorization from blocks in the reference layout
 * (ButterflyFactorization, bfstructs.jl:21-147): table rows are (stage,
 * obs_out, src_out, obs_in, src_in, m, n, identity, out_ptr, in_ptr) with
 * stage 0 = Q, 1..n_stages-2 = R levels, n_stages-1 = P; data_ofs gives the
 * first complex128 entry of each block's column-major data (-1 identity).
 * Used to apply factorizations built elsewhere. */
int bfgpu_bf_from_host(bfgpu_ctx *ctx, const bfgpu_tree *test, const bfgpu_tree *trial,
                       int32_t n_stages, int64_t n_blocks, const int32_t *table,
                       const int64_t *data_ofs, const double *data, int32_t precision,
                       bfgpu_bf **out);

typedef struct {
    int32_t n_bf;          /* factorizations held */
    int32_t precision;
    int64_t rows, cols;    /* size(BF) (bfdims.jl:54-64); far field: tree sizes */
    int32_t max_stages;
    int64_t n_blocks;      /* reference-layout blocks over all factorizations */
    int64_t n_data;        /* complex entries of factor data */
    int64_t factor_bytes;  /* device bytes of packed factors */
    int64_t slot_len;      /* skeleton-vector entries per right-hand side */
    int64_t n_comp;        /* compressions performed */
    int64_t n_retry;       /* adaptive doublings performed */
    double t_factor_ms;    /* wall time of the factorization call */
    /* its parts: host planning and block emission, row-sample preparation,
     * kernel fill + pivoted QR (GPU, synchronised), interpolation matrices
     * (GPU), launch-table packing */
    double t_plan_ms, t_sample_ms, t_qr_ms, t_interp_ms, t_pack_ms;
    int64_t device_bytes;  /* device memory the handle holds: factors, launch
                              tables, skeleton-vector and staging buffers */
    int64_t n_judged;      /* learned ranks: compressions checked by the judge */
    int64_t n_fallback;    /* of these, redone by plain adaptive RRQR */
} bfgpu_bf_info;
int bfgpu_bf_info_get(const bfgpu_bf *bf, bfgpu_bf_info *out);

/* recompress_BF (src/Butterflyalgebra/algrecomp.jl:19-148): a new operator
 * whose R levels are recompressed algebraically at relative tolerance tau
 * (right recompression, then left recompression through the adjoint);
 * every factorization of the operator, on the GPU.  Q and P are kept. */
int bfgpu_bf_recompress(bfgpu_bf *bf, double tau, bfgpu_bf **out);

/* adjoint(BF) / transpose(BF) constructed explicitly
 * (src/Butterflyalgebra/bfadjtr.jl:43-89): a new operator over the swapped
 * trees whose factorizations have their stages reversed, block keys swapped
 * and blocks conjugate-transposed (op BFGPU_OP_C) or transposed
 * (BFGPU_OP_T).  Its forward product equals op(A) x through the lazy
 * views of bfgpu_matvec, and runs on the forward kernels. */
int bfgpu_bf_materialize(bfgpu_bf *bf, int32_t op, bfgpu_bf **out);

/* Serialized factorizations (no single counterpart in the reference: the
 * reference keeps factorizations in memory only; this lets a factorization
 * be cached between runs).  save writes every factorization's blocks in the
 * reference layout (bfstructs.jl:21-147) with the skeleton and compression
 * logs and the factor buffers in their stored precision; load reads it
 * onto the device for the same observer / source trees (checked by a
 * fingerprint of their numbering) and rebuilds the launch tables. */
int bfgpu_bf_save(const bfgpu_bf *bf, const char *path);
int bfgpu_bf_load(bfgpu_ctx *ctx, const bfgpu_tree *test, const bfgpu_tree *trial, const char *path,
                  bfgpu_bf **out);

/* Export factorization `index` in the reference layout (complex128 data);
 * sizes from bfgpu_bf_counts. */
int bfgpu_bf_counts(const bfgpu_bf *bf, int32_t index, int32_t *n_stages, int64_t *n_blocks,
                    int64_t *n_data, int64_t *n_skel, int64_t *n_skel_idx, int64_t *n_comp,
                    int32_t roots[2]);
int bfgpu_bf_export(const bfgpu_bf *bf, int32_t index, int32_t *table, int64_t *data_ofs,
                    double *data);
/* Skeleton records (stage, snode, onode, ofs, len) + point indices, and
 * compression records (stage, snode, onode, guess, rows, rank). */
int bfgpu_bf_export_logs(const bfgpu_bf *bf, int32_t index, int32_t *skel_rec,
                         int32_t *skel_idx, int32_t *comp_rec);
/* Judge verdict of each compression record (same order as comp_rec): 0 not
 * judged, 1 accepted by the residual judge, 2 rejected and redone by plain
 * adaptive RRQR (bfgpu_opts.judge_rows). */
int bfgpu_bf_export_judge(const bfgpu_bf *bf, int32_t index, int32_t *judge);
/* Features of each compression record (6 floats, same order as comp_rec):
 * source and observer radii in wavelengths, centre distance in
 * wavelengths, |cos| of the separation angle to z, level, log10(1/eps).
 * These are the inputs of the learned rank predictor. */
int bfgpu_bf_export_features(const bfgpu_bf *bf, int32_t index, float *feat);
/* Every compression record of every factorization, in factorization order
 * (the rank logs of compressors.jl:52-69 over a whole operator): 7 int32
 * per record (factorization index, stage, snode, onode, guess, rows, rank)
 * and its 6 features (as bfgpu_bf_export_features).  With rec and feat
 * NULL only the count is returned. */
int bfgpu_bf_compressions(const bfgpu_bf *bf, int64_t *n, int32_t *rec, float *feat);
/* Per-stage statistics of the packed tables: blocks, identity blocks, max
 * rank, factor bytes (show, src/showfcts.jl:81-124). 4 int64 per stage. */
int bfgpu_bf_stage_stats(const bfgpu_bf *bf, int32_t *n_stages, int64_t *stats);
/* Launch tables of one direction (op N, or T/C which share theirs), per
 * level in launch order: items, contributions, most contributions of an
 * item, factor entries, algorithmic vector bytes per RHS, kernel variant,
 * tiled shape (packed as for BFGPU_OPT_TILE), partial-sum items (0 when the
 * level has no split items), distinct input entries read, output entries
 * written (twice for y, which is read and written).  10 int64 per level;
 * the last two give the level's algorithmic vector traffic. */
int bfgpu_bf_launch_stats(const bfgpu_bf *bf, int32_t op, int32_t *n_levels, int64_t *stats);
/* Streaming-kernel tables of the forward levels (BFGPU_OPT_STREAM; no
 * counterpart in the reference): per level, chunk KiB (0: not built),
 * chunks, tasks in chunks, tasks left to the dense kernel, factor entries
 * copied that belong to no block, input entries copied.  6 int64 per
 * level. */
int bfgpu_bf_stream_stats(const bfgpu_bf *bf, int32_t *n_levels, int64_t *stats);

/* Convert the factor storage to another precision (e.g. FP64 -> FP32). */
int bfgpu_bf_convert(bfgpu_bf *bf, int32_t precision, bfgpu_bf **out);

/* y = alpha * op(A) * x + beta * y for nrhs right-hand sides
 * (mul!, src/Butterflyalgebra/bfvector.jl:200-296; bfmatrix.jl:10-50;
 * adjoint/transpose of bfadjtr.jl:43-89 applied without materialising;
 * far loop of src/matrixalgebra/matrixvector.jl:39-62).  x, y have the
 * factorization's precision; flags BFGPU_XY_DEVICE means x and y are device
 * pointers (work is enqueued on the context stream), otherwise host
 * pointers (the call copies and synchronises).  alpha, beta: complex (2
 * doubles). */
int bfgpu_matvec(bfgpu_bf *bf, int32_t op, int64_t nrhs, const void *x, int64_t ldx, void *y,
                 int64_t ldy, const double *alpha, const double *beta, int32_t flags);

/* Allocation-free form of bfgpu_matvec: x and y are device pointers (work
 * is enqueued on the context stream) and the call allocates no device or
 * host memory.  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The skeleton vectors of the call live
 * in a static device array of BFGPU_NOALLOC_SCRATCH_BYTES per GPU, shared
 * by every call on that GPU: calls that overlap on one GPU race.  A call
 * whose skeleton vectors (bfgpu_bf_info.slot_len * nrhs * element size)
 * exceed it fails with BFGPU_ERR_ARG.  One launch per level, one column per
 * right-hand side (the matrix-core path across right-hand sides, the fused
 * sweep and graph replay are not used); results equal bfgpu_matvec's to
 * rounding (identical launches at one right-hand side). */
#define BFGPU_NOALLOC_SCRATCH_BYTES (64 * 1024 * 1024)
int bfgpu_matvec_noalloc(bfgpu_bf *bf, int32_t op, int64_t nrhs, const void *x, int64_t ldx, void *y,
                         int64_t ldy, const double *alpha, const double *beta);

/* Same as bfgpu_matvec with device x/y, repeated `iters` times with HIP
 * events around every stage launch on the context stream; stage_ms (length
 * n_stages) receives the mean duration per stage, total_ms the mean per
 * matvec; stage_bytes receives the algorithmic bytes per stage. */
int bfgpu_matvec_timed(bfgpu_bf *bf, int32_t op, int64_t nrhs, const void *x, int64_t ldx,
                       void *y, int64_t ldy, int32_t iters, double *stage_ms,
                       int64_t *stage_bytes, int32_t *n_stages, double *total_ms);

/* Per-launch form of bfgpu_matvec_timed: mean duration (HIP events on the
 * context stream) of every kernel launch of one matvec, in launch order;
 * n_launch receives the count, the first `cap` launches are written.
 * desc (6 int32 per launch): level, table (0 level table, 1 partial sums,
 * 2 input permutation, 3 fused sweep), kernel variant, output kind (0
 * slot, 1 rows of y), packed launch shape (as BFGPU_OPT_TILE; matrix-core
 * tables, variant 5: the most rows of their tiles; the matrix-free last
 * level on the matrix cores, variant 10: 1 when FP32 is emulated on bf16),
 * right-hand sides per lane. */
int bfgpu_matvec_launch_times(bfgpu_bf *bf, int32_t op, int64_t nrhs, const void *x, int64_t ldx,
                              void *y, int64_t ldy, int32_t iters, int32_t cap, int32_t *n_launch,
                              double *launch_ms, int32_t *desc);
/* Launch tuning of every table (what the autotuner chose), to save and
 * restore: 8 int32 per level = (variant, packed tile shape) of the forward
 * table, its partial-sum table, the transposed table and its partial-sum
 * table, (-1, -1) where a level has none.  set_tuning validates all
 * entries before applying any. */
int bfgpu_bf_get_tuning(const bfgpu_bf *bf, int32_t *n, int32_t *vals);
int bfgpu_bf_set_tuning(bfgpu_bf *bf, int32_t n, const int32_t *vals);

/* enable = 1: forward matvecs with 1..8 right-hand sides run as one fused,
 * dependency-driven launch over all levels; 0 (default) selects one launch
 * per level.  Both give the same results to rounding. */
int bfgpu_bf_set_fused(bfgpu_bf *bf, int32_t enable);
/* Tuning options (used by the launch autotuner):
 *   BFGPU_OPT_SWEEP      0: one launch per level, 1: fused sweep with dynamic
 *                        task tickets, 2: fused sweep, cooperative launch
 *   BFGPU_OPT_WG_PER_CU  workgroups per CU of the fused sweep (0 = occupancy)
 *   BFGPU_OPT_GRAPH      1: replay per-level launches from a captured HIP
 *                        graph when x, y, alpha, beta repeat
 *   BFGPU_OPT_VARIANT    kernel of every launch table: 0 workgroup per task,
 *                        1 wavefront per task (forward; transposed tables
 *                        take 0), 2 tiled (several tasks per wavefront),
 *                        3 dense (forward: lanes packed to each block's row
 *                        count; transposed: a lane per output column, or
 *                        with csplit 1 lanes on the rows of the block
 *                        columns and a group per column); get:
 *                        the common value, or -1 when levels differ
 *   BFGPU_OPT_TILE       tiled shape of every table, packed as rp_shift |
 *                        g_shift << 4 | csplit << 8 | wg_per_cu << 12 |
 *                        xcd << 20 (2^rp_shift outputs x 2^g_shift lane
 *                        groups per task, 2 <= rp_shift + g_shift <= 6; groups
 *                        split the inner index (csplit 0) or the
 *                        contributions (1); wg_per_cu 1..64; xcd 1: each
 *                        XCD walks one contiguous range of tasks; qw4 <<
 *                        24: 4 instead of 8 loads per lane and pass for
 *                        the dense kernel at one RHS); also selects
 *                        variant 2
 *   BFGPU_OPT_DENSE      dense shape of every table, packed as for
 *                        BFGPU_OPT_TILE with rp_shift = log2(lanes per
 *                        task) and g_shift = 0 (the transposed tables use
 *                        the lanes per task, the XCD mapping, csplit 1 for
 *                        grouped rows and qw4 with it); selects variant 3
 *   BFGPU_OPT_RHS_LANES  forward matvecs with at least this many right-hand
 *                        sides (default 32; 0 = never) put the right-hand
 *                        sides across lanes, 64 per sweep, with the factor
 *                        entries read once per sweep
 *   BFGPU_OPT_MFMA       with right-hand sides across lanes, complex FP32
 *                        factorizations multiply on the matrix cores
 *                        (v_mfma_f32_16x16x4_f32; small blocks that read the
 *                        same inputs stacked into tiles); 16, 32 (default),
 *                        48 or 64: the most rows of a tile (taller tiles share
 *                        each input load between more rows and hold more
 *                        registers); 1: on, height unchanged; 0: vector-ALU
 *                        kernel; get: 0 or the tile height
 *   BFGPU_OPT_RHS_CHUNK  most right-hand sides per pass over the levels;
 *                        0 (default): as many as the skeleton-vector buffer
 *                        of one pass allows in 90 % of the free device
 *                        memory (at most 64 with right-hand sides across
 *                        lanes)
 *   BFGPU_OPT_PFREE      1 (default): one-RHS complex FP32 products (forward,
 *                        adjoint, transpose) compute the last level's P
 *                        entries (point-kernel entries between observer-
 *                        leaf points and skeleton points) from the points
 *                        instead of reading them;
 *                        0: read the stored P blocks.  Applies to operators
 *                        factorized here with the point kernel (and their
 *                        recompressed or converted copies).
 *   BFGPU_OPT_SLOT_CAP   most KiB the skeleton vectors of one pass over the
 *                        levels may take (0, default: 90 % of the free
 *                        device memory); when 16 interleaved columns do not
 *                        fit, right-hand sides go one column each, in
 *                        passes of as many as fit.
 *   BFGPU_OPT_TCOPY      1 (default): the autotuner may run adjoint and
 *                        transpose levels with the forward dense kernel over
 *                        a transposed copy of the stored blocks (tuning
 *                        variant 8 of a transposed table), built on the
 *                        first autotune (or on this setting) when device
 *                        memory holds it with a tenth and 1 GiB to spare;
 *                        the matrix-free level is not copied; 0 releases
 *                        the copy.  get: 1 while the copy exists.
 *   BFGPU_OPT_STREAM     4 or 8: every forward table runs the streaming
 *                        kernel at one right-hand side (variant 9): runs of
 *                        consecutive single-block tasks whose blocks lie
 *                        back to back are copied, 4 or 8 KiB at a time, into
 *                        the wavefront's LDS with 1 KiB coalesced loads and
 *                        computed there; other tasks, and products of
 *                        several right-hand sides, take the dense kernel
 *                        (the table's dense shape, bit 28 of the packed
 *                        shape selects 4 KiB chunks); 0: variant 3 (dense).
 *                        get: the chunk KiB when every forward table
 *                        streams, else 0.
 *   BFGPU_OPT_MFMA_EMUL  1 (default): with right-hand sides across lanes,
 *                        the matrix-free last level emulates its FP32
 *                        products on the bf16 matrix cores (bf16x3: each
 *                        operand as a high and a low bf16 half, three
 *                        products per real product, FP32 accumulation;
 *                        about 17 bits per operand, products within about
 *                        1e-5 of the FP32 ones); 0: FP32 matrix cores. */
enum {
    BFGPU_OPT_SWEEP = 1,
    BFGPU_OPT_WG_PER_CU = 2,
    BFGPU_OPT_GRAPH = 3,
    BFGPU_OPT_VARIANT = 4,
    BFGPU_OPT_TILE = 5,
    BFGPU_OPT_DENSE = 6,
    BFGPU_OPT_RHS_LANES = 7,
    BFGPU_OPT_MFMA = 8,
    BFGPU_OPT_RHS_CHUNK = 9,
    BFGPU_OPT_PFREE = 10,
    BFGPU_OPT_SLOT_CAP = 11,
    BFGPU_OPT_TCOPY = 12,
    BFGPU_OPT_STREAM = 13,
    BFGPU_OPT_MFMA_EMUL = 14
};
int bfgpu_bf_set_option(bfgpu_bf *bf, int32_t key, int32_t value);
int bfgpu_bf_get_option(const bfgpu_bf *bf, int32_t key, int32_t *value);
/* Autotune the launch shape of every level for nrhs right-hand sides, in
 * both directions (forward, and adjoint/transpose): times each kernel
 * variant and tiled shape over `iters` launches and keeps the fastest.
 * stage_ms / variants (length = number of launch stages, may be NULL)
 * receive the forward choice. */
int bfgpu_bf_autotune(bfgpu_bf *bf, int64_t nrhs, int32_t iters, double *stage_ms,
                      int32_t *variants);
/* Synchronise and report a fault flagged by the fused sweep since the last
 * check (BFGPU_ERR_STATE), e.g. after device-pointer matvecs. */
int bfgpu_bf_check(bfgpu_bf *bf);

int bfgpu_bf_destroy(bfgpu_bf *bf);

/* ---- learned rank predictor (MLP 6 -> H -> H -> H -> 1, ReLU, FP32) ---- */
int bfgpu_mlp_create(bfgpu_ctx *ctx, int32_t n_in, int32_t hidden, int32_t n_layers,
                     uint64_t seed, bfgpu_mlp **out);
int bfgpu_mlp_destroy(bfgpu_mlp *m);
/* Parameters as one flat FP32 host array (W1,b1,...,WL,bL; W row-major
 * out x in). */
int bfgpu_mlp_n_params(const bfgpu_mlp *m, int64_t *n);
int bfgpu_mlp_get_params(const bfgpu_mlp *m, float *params);
int bfgpu_mlp_set_params(bfgpu_mlp *m, const float *params);
/* Input standardisation x' = (x - mu) / sigma (n_in each) and output map
 * rank = y * out_scale + out_shift; targets of loss_grad/train are ranks. */
int bfgpu_mlp_set_norm(bfgpu_mlp *m, const float *mu, const float *sigma, float out_shift,
                       float out_scale);
/* Forward for n samples (features n x n_in row-major, host) -> out (n). */
int bfgpu_mlp_forward(bfgpu_mlp *m, int64_t n, const float *features, float *out);
/* Loss (mean squared error) and gradient w.r.t. the parameters for one
 * batch (host arrays); grads has n_params entries. */
int bfgpu_mlp_loss_grad(bfgpu_mlp *m, int64_t n, const float *features, const float *target,
                        float *loss, float *grads);
/* Train with Adam for `epochs` passes over the data in minibatches. */
int bfgpu_mlp_train(bfgpu_mlp *m, int64_t n, const float *features, const float *target,
                    int32_t epochs, int32_t batch, float lr, uint64_t seed, float *final_loss);

#ifdef __cplusplus
}
#endif
#endif
