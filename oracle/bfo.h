/*
 * bfo.h -- CPU FP64 oracle for the butterfly far-field path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * ButterflyFactorizations.jl algorithm (kernel fill, rank estimation, partial
 * pivoted QR compression, the level-by-level factorization constructor, the
 * flat-workspace matvec and the adjoint/transpose), used by the tests and by
 * the CPU-baseline leg of the benchmark, as the reference results.  Nothing in the
 * product library links, loads or calls it.
 *
 * Conventions: indices are 0-based int32; complex numbers are interleaved
 * (re, im) doubles (C99 double _Complex); matrices are column-major with
 * leading dimension equal to the row count.
 */
#ifndef BFO_H
#define BFO_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A cluster tree, borrowed arrays.  Node 0 is the root; the children of a
 * node are the contiguous ids [first_child, first_child + n_children); the
 * values (point indices) of a node are perm[val_begin .. val_end). */
typedef struct {
    int32_t n_nodes;
    int32_t n_values;
    const int32_t *parent;
    const int32_t *first_child;
    const int32_t *n_children;
    const int32_t *val_begin;
    const int32_t *val_end;
    const double *center;   /* 3 * n_nodes */
    const double *halfsize; /* n_nodes; radius = halfsize * sqrt(3) */
    const int32_t *perm;    /* n_values */
} bfo_tree;

typedef struct {
    double k;          /* wavenumber */
    double tol;        /* user tolerance tau; compressions use tau / L */
    int32_t adaptive;  /* adaptive row doubling */
    int32_t rank_mode; /* 0: butterfly estimator, 1: geometric, 2: fixed table */
    double c_tau;      /* butterfly estimator constant */
    int32_t rmin;      /* estimator floor */
    double c_geo, c_eps; /* geometric estimator constants */
    uint64_t seed;     /* row-sample seed */
    const int32_t *fixed_guess; /* rank_mode 2: per-compression guess, consumed in order */
    int32_t n_fixed;
    int32_t judge_rows; /* > 0: residual judge on this many held-out rows (bfo_judge) */
    double judge_slack; /* judge accepts when e_max <= judge_slack * eps * z_max */
} bfo_opts;

typedef struct bfo_bf bfo_bf;

/* ---- kernel ---- */
void bfo_helmholtz_fill(double k, const double *obs_xyz, const int32_t *rows, int32_t nr,
                        const double *src_xyz, const int32_t *cols, int32_t nc,
                        double *Z /* complex, nr x nc */);

/* ---- rank estimators ---- */
int32_t bfo_rank_butterfly(double eps, double c_tau, int32_t rmin, double *x1, double *x2);
int32_t bfo_rank_geometric(double k, const double *cs, const double *co, double as,
                           double ao, double eps, double C, double Ceps, int32_t rmin,
                           double *x1, double *x2);

/* ---- row sample ---- */
uint64_t bfo_mix64(uint64_t z);
uint64_t bfo_sample_key(uint64_t seed, int32_t bf_obs_root, int32_t bf_src_root,
                        int32_t stage, int32_t snode, int32_t onode);
void bfo_sample_prefix(uint64_t key, int32_t n, int32_t m, int32_t *out);

/* ---- pivoted QR and interpolation matrix ---- */
int32_t bfo_cpqr(int32_t m, int32_t n, double *A /* complex, in-out */, double rtol,
                 int32_t *piv, double *tau /* complex, min(m,n) */);
void bfo_interp(int32_t m, int32_t n, const double *R, int32_t rank, const int32_t *piv,
                double *T /* complex, rank x n */);
int32_t bfo_compress(double k, const double *obs_xyz, const int32_t *obs_idx, int32_t n_obs,
                     const double *src_xyz, const int32_t *src_idx, int32_t n_src,
                     int32_t guess, double eps, int32_t adaptive, uint64_t key,
                     double *T_out /* complex, cap n_src*n_src or larger */, int32_t *skel_out,
                     int32_t *rows_used);

/* Residual of an interpolative decomposition on held-out rows of the same
 * sample prefix (rows rows .. rows + min(judge_rows, n_obs - rows) - 1):
 * e_max = max column norm of K(check, src) - K(check, skel) T, z_max = max
 * column norm of K(check, src). */
void bfo_judge(double k, const double *obs_xyz, const int32_t *obs_idx, int32_t n_obs,
               const double *src_xyz, const int32_t *src_idx, int32_t n_src, uint64_t key, int32_t rows,
               int32_t judge_rows, const double *T, const int32_t *skel, int32_t r, double *e_max,
               double *z_max);

/* ---- factorization ---- */
bfo_bf *bfo_assemble(const bfo_tree *obsT, const double *obs_xyz, const bfo_tree *srcT,
                     const double *src_xyz, int32_t no, int32_t ns, const bfo_opts *opts);
void bfo_bf_free(bfo_bf *bf);
/* Materialised adjoint (conj=1) or transpose (conj=0), as the reference does. */
bfo_bf *bfo_bf_transform(const bfo_bf *bf, int32_t conj);

/* Export.  n_stages = number of R levels + 2 (Q, R^1..R^{L-1}, P).  Block
 * table rows: stage, obs_out, src_out, obs_in, src_in, m, n, identity,
 * out_ptr, in_ptr (10 int32 each). */
int32_t bfo_bf_n_stages(const bfo_bf *bf);
int64_t bfo_bf_n_blocks(const bfo_bf *bf);
int64_t bfo_bf_n_data(const bfo_bf *bf); /* complex entries */
void bfo_bf_export(const bfo_bf *bf, int32_t *table, int64_t *data_ofs, double *data);
/* Skeleton log: one record per K update in order (stage, snode, onode, ofs, len). */
int64_t bfo_bf_n_skel(const bfo_bf *bf);
int64_t bfo_bf_n_skel_idx(const bfo_bf *bf);
void bfo_bf_export_skel(const bfo_bf *bf, int32_t *rec /* 5 per record */, int32_t *idx);
/* Sample log: one record per compression (stage, snode, onode, guess, rows, rank). */
int64_t bfo_bf_n_comp(const bfo_bf *bf);
void bfo_bf_export_comp(const bfo_bf *bf, int32_t *rec /* 6 per record */);
/* Per compression record: 0 not judged, 1 accepted by the residual judge,
 * 2 rejected and redone by plain adaptive RRQR. */
void bfo_bf_export_judge(const bfo_bf *bf, int32_t *v);
void bfo_bf_dims(const bfo_bf *bf, int64_t *rows, int64_t *cols);

/* Import a factorization from the same block table / data format. */
bfo_bf *bfo_bf_import(const bfo_tree *obsT, const bfo_tree *srcT, int32_t n_stages,
                      int64_t n_blocks, const int32_t *table, const int64_t *data_ofs,
                      const double *data);

/* y = alpha * BF * X + beta * y, column by column, flat-workspace semantics. */
void bfo_matvec(const bfo_bf *bf, int64_t nrhs, const double *x, int64_t ldx, double *y,
                int64_t ldy, const double *alpha, const double *beta);

/* Global far operator: y = sum_b BF_b x (beta applied once). */
void bfo_matvec_many(bfo_bf *const *bfs, int32_t nbf, int64_t nrhs, const double *x,
                     int64_t ldx, double *y, int64_t ldy, const double *alpha,
                     const double *beta, int32_t nthreads);

/* ---- admissibility and interaction lists ---- */
int32_t bfo_isfar_aabb(const bfo_tree *srcT, const bfo_tree *tstT, int32_t s, int32_t o,
                       double alpha);
int32_t bfo_center_distance(const bfo_tree *srcT, const bfo_tree *tstT, int32_t s, int32_t o,
                            double beta);
/* kind 0: isFarFunctor(alpha), 1: CenterDistanceAdmissibility(beta).  Fills
 * far/near pair lists (obs, src) and returns their counts through nfar, nnear;
 * pass NULL buffers to count only. */
void bfo_nearandfar(const bfo_tree *srcT, const bfo_tree *tstT, int32_t kind, double param,
                    int32_t minbflvl, int32_t unbalancedints, int32_t leafcomp,
                    int32_t leafimbalance, int32_t *far, int64_t *nfar, int32_t *near,
                    int64_t *nnear);

#ifdef __cplusplus
}
#endif
#endif
