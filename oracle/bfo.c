/*
 * bfo.c -- CPU FP64 oracle for the butterfly far-field path.
 *
 * TEST INFRASTRUCTURE ONLY: the tests and the benchmark's CPU-baseline leg
 * use this file as the reference the GPU results are compared with.  The product library never links or calls it.
 *
 * Each function cites the reference ButterflyFactorizations.jl file:line it
 * restates.  Parts the reference delegates to third-party code are restated
 * from the published algorithm and say so:
 *   - LowRankApprox.pqr (column-pivoted QR with a relative tolerance) is
 *     restated as LAPACK xGEQP3/xLAQP2 (Businger-Golub pivoting with the
 *     Drmac-Bujanovic norm-downdating safeguard, LAPACK Working Note 176),
 *     truncated at the first step k with |R_kk| <= rtol * |R_11|.
 *   - Julia's randperm in the row sample is replaced by a documented,
 *     seeded, counter-based sparse Fisher-Yates prefix (bfo_sample_prefix).
 */
#include "bfo.h"

#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex cplx;

static void *xmalloc(size_t n) {
    void *p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "bfo: out of memory (%zu bytes)\n", n);
        abort();
    }
    return p;
}
static void *xrealloc(void *p, size_t n) {
    void *q = realloc(p, n ? n : 1);
    if (!q) {
        fprintf(stderr, "bfo: out of memory (%zu bytes)\n", n);
        abort();
    }
    return q;
}

/* ------------------------------------------------------------------------ */
/* Kernel: G(x, y) = exp(-i k |x - y|) / (4 pi |x - y|).                     */
/* Restates the kernel plug-in contract (km)(Z, rows, cols) of               */
/* src/kernelmatrix/abstractkernelmatrix.jl:1-7 (fill in place; empty lists  */
/* leave Z untouched as ext/BFBEAST/kernelmatrix.jl:74-82 zero-fills).       */
/* ------------------------------------------------------------------------ */
static cplx helm(double k, const double *x, const double *y) {
    double dx = x[0] - y[0], dy = x[1] - y[1], dz = x[2] - y[2];
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    double s = sin(k * r), c = cos(k * r);
    double g = 1.0 / (4.0 * M_PI * r);
    return (c * g) - I * (s * g);
}

void bfo_helmholtz_fill(double k, const double *obs_xyz, const int32_t *rows, int32_t nr,
                        const double *src_xyz, const int32_t *cols, int32_t nc, double *Zd) {
    cplx *Z = (cplx *)Zd;
    for (int32_t j = 0; j < nc; ++j) {
        const double *y = src_xyz + 3 * (int64_t)cols[j];
        for (int32_t i = 0; i < nr; ++i)
            Z[i + (int64_t)j * nr] = helm(k, obs_xyz + 3 * (int64_t)rows[i], y);
    }
}

/* ------------------------------------------------------------------------ */
/* Rank estimators: src/ButterflyFactorization/rankestimation.jl:93-98      */
/* (estimate_rank_butterfly) and :70-88 / ext/BFH2Trees/h2rankest.jl:8-38    */
/* (estimate_rank_3d on trees).                                              */
/* ------------------------------------------------------------------------ */
int32_t bfo_rank_butterfly(double eps, double c_tau, int32_t rmin, double *x1, double *x2) {
    double l = log(1.0 / eps);
    int32_t R = (int32_t)ceil(c_tau * l);
    if (x1) *x1 = 1.0;
    if (x2) *x2 = l;
    return R > rmin ? R : rmin;
}

int32_t bfo_rank_geometric(double k, const double *cs, const double *co, double as,
                           double ao, double eps, double C, double Ceps, int32_t rmin,
                           double *x1, double *x2) {
    double d = sqrt((cs[0] - co[0]) * (cs[0] - co[0]) + (cs[1] - co[1]) * (cs[1] - co[1]) +
                    (cs[2] - co[2]) * (cs[2] - co[2]));
    double ds = d > 1e-4 ? d : 1e-4;
    double a = k * (as * ao) / ds;
    double v1 = a * a, v2 = log(1.0 / eps);
    int32_t R = (int32_t)ceil(C * v1 + Ceps * v2);
    if (x1) *x1 = v1;
    if (x2) *x2 = v2;
    return R > rmin ? R : rmin;
}

/* ------------------------------------------------------------------------ */
/* Row sample.  The reference draws shuffled_obs = obs_index[randperm(n)]    */
/* and takes growing prefixes of it (compressors.jl:101,119,144).  Here the  */
/* permutation prefix is produced by a sparse Fisher-Yates shuffle driven by */
/* the SplitMix64 finaliser used as a counter-based generator:               */
/*   u_i = mix64(key + (i + 1) * 0x9E3779B97F4A7C15)                          */
/*   j_i = i + floor((u_i >> 32) * (n - i) / 2^32)                           */
/*   swap(a[i], a[j_i]) on the identity permutation a.                        */
/* A longer prefix extends a shorter one, as the reference's doubling does.  */
/* ------------------------------------------------------------------------ */
uint64_t bfo_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t bfo_sample_key(uint64_t seed, int32_t bf_obs_root, int32_t bf_src_root,
                        int32_t stage, int32_t snode, int32_t onode) {
    uint64_t h = bfo_mix64(seed + 0x9E3779B97F4A7C15ULL);
    h = bfo_mix64(h ^ (uint64_t)(uint32_t)bf_obs_root);
    h = bfo_mix64(h ^ ((uint64_t)(uint32_t)bf_src_root << 1));
    h = bfo_mix64(h ^ (uint64_t)(uint32_t)stage);
    h = bfo_mix64(h ^ (uint64_t)(uint32_t)snode);
    h = bfo_mix64(h ^ (uint64_t)(uint32_t)onode);
    return h;
}

/* open-addressing map int32 -> int32 used by the sparse shuffle */
typedef struct {
    int32_t *key, *val;
    int32_t cap;
} imap;

static void imap_init(imap *m, int32_t n) {
    int32_t cap = 16;
    while (cap < 4 * n) cap <<= 1;
    m->cap = cap;
    m->key = (int32_t *)xmalloc(sizeof(int32_t) * cap);
    m->val = (int32_t *)xmalloc(sizeof(int32_t) * cap);
    for (int32_t i = 0; i < cap; ++i) m->key[i] = -1;
}
static int32_t *imap_slot(imap *m, int32_t k) {
    uint32_t h = (uint32_t)((uint64_t)(uint32_t)k * 0x9E3779B1u) & (uint32_t)(m->cap - 1);
    while (m->key[h] != -1 && m->key[h] != k) h = (h + 1) & (uint32_t)(m->cap - 1);
    return (int32_t *)&m->key[h];
}
static int32_t imap_get(imap *m, int32_t k) {
    int32_t *s = imap_slot(m, k);
    if (*s == -1) return k;
    return m->val[s - m->key];
}
static void imap_set(imap *m, int32_t k, int32_t v) {
    int32_t *s = imap_slot(m, k);
    *s = k;
    m->val[s - m->key] = v;
}

void bfo_sample_prefix(uint64_t key, int32_t n, int32_t m, int32_t *out) {
    if (m > n) m = n;
    imap mp;
    imap_init(&mp, m);
    for (int32_t i = 0; i < m; ++i) {
        uint64_t u = bfo_mix64(key + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL);
        uint32_t span = (uint32_t)(n - i);
        int32_t j = i + (int32_t)(((u >> 32) * (uint64_t)span) >> 32);
        int32_t ai = imap_get(&mp, i), aj = imap_get(&mp, j);
        imap_set(&mp, i, aj);
        imap_set(&mp, j, ai);
        out[i] = aj;
    }
    free(mp.key);
    free(mp.val);
}

/* ------------------------------------------------------------------------ */
/* Column-pivoted Householder QR with relative-tolerance truncation.        */
/* Stands in for LowRankApprox.pqr(Z; rtol) (compressors.jl:129).  Restates */
/* LAPACK ZGEQP3/ZLAQP2 (unblocked): pivot = first column of maximal partial */
/* norm; Householder reflector in ZLARFG form (beta real, tau complex,       */
/* v(1) = 1); trailing update with H^H; partial norms downdated with the     */
/* LAPACK Working Note 176 test (recompute when the downdate loses more than */
/* half the digits).  Stops before step k (k >= 1) when the pivot column's    */
/* remaining norm is <= rtol * |R_11|; returns that k as the rank.           */
/* A is overwritten: R in the upper triangle of the first rank rows.        */
/* piv[j] = original column now at position j.                              */
/* ------------------------------------------------------------------------ */
static double cnrm2(int32_t n, const cplx *x) {
    /* scaled two-norm */
    double scale = 0.0, ssq = 1.0;
    for (int32_t i = 0; i < n; ++i) {
        double v[2] = {creal(x[i]), cimag(x[i])};
        for (int t = 0; t < 2; ++t) {
            if (v[t] != 0.0) {
                double a = fabs(v[t]);
                if (scale < a) {
                    ssq = 1.0 + ssq * (scale / a) * (scale / a);
                    scale = a;
                } else {
                    ssq += (a / scale) * (a / scale);
                }
            }
        }
    }
    return scale * sqrt(ssq);
}

int32_t bfo_cpqr(int32_t m, int32_t n, double *Ad, double rtol, int32_t *piv, double *taud) {
    cplx *A = (cplx *)Ad;
    cplx *tau = (cplx *)taud;
    int32_t mn = m < n ? m : n;
    double *vn1 = (double *)xmalloc(sizeof(double) * (n > 0 ? n : 1));
    double *vn2 = (double *)xmalloc(sizeof(double) * (n > 0 ? n : 1));
    const double tol3z = sqrt(DBL_EPSILON / 2); /* sqrt(DLAMCH('E')) */
    for (int32_t j = 0; j < n; ++j) {
        piv[j] = j;
        vn1[j] = cnrm2(m, A + (int64_t)j * m);
        vn2[j] = vn1[j];
    }
    int32_t rank = mn;
    double r11 = 0.0;
    for (int32_t k = 0; k < mn; ++k) {
        /* pivot: first maximal partial norm */
        int32_t p = k;
        for (int32_t j = k + 1; j < n; ++j)
            if (vn1[j] > vn1[p]) p = j;
        /* truncation test on the exact remaining norm of the pivot column */
        double pn = cnrm2(m - k, A + (int64_t)p * m + k);
        if (k == 0) r11 = pn;
        if (pn == 0.0 || (k > 0 && pn <= rtol * r11)) {
            rank = k;
            break;
        }
        if (p != k) {
            for (int32_t i = 0; i < m; ++i) {
                cplx t = A[i + (int64_t)p * m];
                A[i + (int64_t)p * m] = A[i + (int64_t)k * m];
                A[i + (int64_t)k * m] = t;
            }
            int32_t ti = piv[p];
            piv[p] = piv[k];
            piv[k] = ti;
            vn1[p] = vn1[k];
            vn2[p] = vn2[k];
        }
        /* reflector (ZLARFG) */
        cplx *a = A + (int64_t)k * m;
        cplx alpha = a[k];
        double xnorm = cnrm2(m - k - 1, a + k + 1);
        cplx tk = 0.0;
        if (!(xnorm == 0.0 && cimag(alpha) == 0.0)) {
            double beta = hypot(hypot(creal(alpha), cimag(alpha)), xnorm);
            if (creal(alpha) >= 0.0) beta = -beta;
            tk = (beta - creal(alpha)) / beta - I * (cimag(alpha) / beta);
            cplx scal = 1.0 / (alpha - beta);
            for (int32_t i = k + 1; i < m; ++i) a[i] *= scal;
            alpha = beta;
        }
        a[k] = alpha;
        tau[k] = tk;
        /* apply H^H = I - conj(tau) v v^H to the trailing columns */
        if (tk != 0.0) {
            for (int32_t j = k + 1; j < n; ++j) {
                cplx *c = A + (int64_t)j * m;
                cplx w = c[k];
                for (int32_t i = k + 1; i < m; ++i) w += conj(a[i]) * c[i];
                w *= conj(tk);
                c[k] -= w;
                for (int32_t i = k + 1; i < m; ++i) c[i] -= a[i] * w;
            }
        }
        /* downdate partial norms (ZLAQP2, LAWN 176) */
        for (int32_t j = k + 1; j < n; ++j) {
            if (vn1[j] != 0.0) {
                double t = cabs(A[k + (int64_t)j * m]) / vn1[j];
                t = 1.0 - t * t;
                if (t < 0.0) t = 0.0;
                double t2 = t * (vn1[j] / vn2[j]) * (vn1[j] / vn2[j]);
                if (t2 <= tol3z) {
                    if (k + 1 < m) {
                        vn1[j] = cnrm2(m - k - 1, A + (int64_t)j * m + k + 1);
                        vn2[j] = vn1[j];
                    } else {
                        vn1[j] = 0.0;
                        vn2[j] = 0.0;
                    }
                } else {
                    vn1[j] *= sqrt(t);
                }
            }
        }
    }
    free(vn1);
    free(vn2);
    return rank;
}

/* Interpolation matrix T = R11^{-1} [R11 R12] P^T (rank x n): the reference
 * forms R11 \ (Q1^H Z) (compressors.jl:155-160); Q1^H Z P equals the first
 * rank rows of R, so T[:, piv[j]] = e_j for j < rank and
 * T[:, piv[j]] = R11^{-1} R[0:rank, j] for j >= rank. */
void bfo_interp(int32_t m, int32_t n, const double *Rd, int32_t rank, const int32_t *piv,
                double *Td) {
    const cplx *R = (const cplx *)Rd;
    cplx *T = (cplx *)Td;
    cplx *col = (cplx *)xmalloc(sizeof(cplx) * (rank > 0 ? rank : 1));
    for (int32_t j = 0; j < n; ++j) {
        cplx *t = T + (int64_t)piv[j] * rank;
        if (j < rank) {
            for (int32_t i = 0; i < rank; ++i) t[i] = (i == j) ? 1.0 : 0.0;
            continue;
        }
        for (int32_t i = 0; i < rank; ++i) col[i] = R[i + (int64_t)j * m];
        for (int32_t i = rank - 1; i >= 0; --i) {
            cplx s = col[i];
            for (int32_t q = i + 1; q < rank; ++q) s -= R[i + (int64_t)q * m] * col[q];
            col[i] = s / R[i + (int64_t)i * m];
        }
        for (int32_t i = 0; i < rank; ++i) t[i] = col[i];
    }
    free(col);
}

/* PartialQR functor, compressors.jl:81-165: sample n_otilde = min(max(guess,
 * 10), n_obs) rows, fill Z = K(rows, src), pivoted QR at rtol eps, double the
 * rows while adaptive and r > floor(0.8 rows) and rows < n_obs, then return
 * the interpolation matrix and skeleton src[piv[0:r]].  T_out must hold
 * r * n_src complex entries (r <= min(rows, n_src)).  Returns r. */
int32_t bfo_compress(double k, const double *obs_xyz, const int32_t *obs_idx, int32_t n_obs,
                     const double *src_xyz, const int32_t *src_idx, int32_t n_src,
                     int32_t guess, double eps, int32_t adaptive, uint64_t key,
                     double *T_out, int32_t *skel_out, int32_t *rows_used) {
    if (n_obs == 0 || n_src == 0) {
        if (rows_used) *rows_used = 0;
        return 0;
    }
    int32_t rows = guess > 10 ? guess : 10;
    if (rows > n_obs) rows = n_obs;
    int32_t *perm = NULL, *ridx = NULL, *piv = (int32_t *)xmalloc(sizeof(int32_t) * n_src);
    cplx *Z = NULL, *tau = NULL;
    int32_t r;
    for (;;) {
        perm = (int32_t *)xrealloc(perm, sizeof(int32_t) * rows);
        ridx = (int32_t *)xrealloc(ridx, sizeof(int32_t) * rows);
        bfo_sample_prefix(key, n_obs, rows, perm);
        for (int32_t i = 0; i < rows; ++i) ridx[i] = obs_idx[perm[i]];
        Z = (cplx *)xrealloc(Z, sizeof(cplx) * (size_t)rows * n_src);
        tau = (cplx *)xrealloc(tau, sizeof(cplx) * (size_t)(rows < n_src ? rows : n_src));
        bfo_helmholtz_fill(k, obs_xyz, ridx, rows, src_xyz, src_idx, n_src, (double *)Z);
        r = bfo_cpqr(rows, n_src, (double *)Z, eps, piv, (double *)tau);
        if (adaptive && r > (int32_t)floor(0.8 * rows) && rows < n_obs) {
            rows = rows * 2 < n_obs ? rows * 2 : n_obs;
            continue;
        }
        break;
    }
    bfo_interp(rows, n_src, (double *)Z, r, piv, T_out);
    for (int32_t i = 0; i < r; ++i) skel_out[i] = src_idx[piv[i]];
    if (rows_used) *rows_used = rows;
    free(perm);
    free(ridx);
    free(piv);
    free(Z);
    free(tau);
    return r;
}

/* Residual judge of a compression whose row sample was sized by a rank
 * prediction (no counterpart in the reference; its adaptive loop,
 * compressors.jl:124-151, is the fallback it triggers): the interpolative
 * decomposition T (r x n_src, skeleton columns skel) found on the first
 * `rows` sampled rows is checked on the next m = min(judge_rows, n_obs -
 * rows) rows of the same sample prefix, rows the QR never saw.  With
 * Z_c = K(check rows, src) and E = Z_c - K(check rows, skel) T:
 *   e_max = max_j |E[:, j]|_2,   z_max = max_j |Z_c[:, j]|_2,
 * the held-out form of the pivoted QR's own stopping rule (every remaining
 * column norm within rtol of the largest column norm). */
void bfo_judge(double k, const double *obs_xyz, const int32_t *obs_idx, int32_t n_obs,
               const double *src_xyz, const int32_t *src_idx, int32_t n_src, uint64_t key, int32_t rows,
               int32_t judge_rows, const double *Td, const int32_t *skel, int32_t r, double *e_max,
               double *z_max) {
    int32_t m = n_obs - rows < judge_rows ? n_obs - rows : judge_rows;
    *e_max = 0.0;
    *z_max = 0.0;
    if (m <= 0 || n_src <= 0) return;
    const cplx *T = (const cplx *)Td;
    int32_t *perm = (int32_t *)xmalloc(sizeof(int32_t) * (size_t)(rows + m));
    int32_t *cr = (int32_t *)xmalloc(sizeof(int32_t) * (size_t)m);
    bfo_sample_prefix(key, n_obs, rows + m, perm);
    for (int32_t i = 0; i < m; ++i) cr[i] = obs_idx[perm[rows + i]];
    cplx *Z = (cplx *)xmalloc(sizeof(cplx) * (size_t)m * n_src);
    cplx *Zs = (cplx *)xmalloc(sizeof(cplx) * (size_t)m * (r > 0 ? r : 1));
    bfo_helmholtz_fill(k, obs_xyz, cr, m, src_xyz, src_idx, n_src, (double *)Z);
    if (r > 0) bfo_helmholtz_fill(k, obs_xyz, cr, m, src_xyz, skel, r, (double *)Zs);
    for (int32_t j = 0; j < n_src; ++j) {
        double en = 0.0, zn = 0.0;
        for (int32_t i = 0; i < m; ++i) {
            cplx z = Z[i + (int64_t)j * m], e = z;
            for (int32_t q = 0; q < r; ++q) e -= Zs[i + (int64_t)q * m] * T[q + (int64_t)j * r];
            en += creal(e) * creal(e) + cimag(e) * cimag(e);
            zn += creal(z) * creal(z) + cimag(z) * cimag(z);
        }
        en = sqrt(en);
        zn = sqrt(zn);
        if (en > *e_max) *e_max = en;
        if (zn > *z_max) *z_max = zn;
    }
    free(perm);
    free(cr);
    free(Z);
    free(Zs);
}

/* ------------------------------------------------------------------------ */
/* Growable arrays                                                           */
/* ------------------------------------------------------------------------ */
typedef struct {
    int32_t *v;
    int64_t n, cap;
} ivec;
static void iv_push(ivec *a, int32_t x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 16;
        a->v = (int32_t *)xrealloc(a->v, sizeof(int32_t) * a->cap);
    }
    a->v[a->n++] = x;
}
static void iv_free(ivec *a) {
    free(a->v);
    a->v = NULL;
    a->n = a->cap = 0;
}

/* ------------------------------------------------------------------------ */
/* Tree helpers: src/auxillaries.jl:263-281 (treelevels) and :340-350       */
/* (traverseandpad: shallow leaves are pushed down to the last level).       */
/* ------------------------------------------------------------------------ */
typedef struct {
    int32_t nlev;
    ivec *lev;
} levels_t;

static int isleaf(const bfo_tree *t, int32_t v) { return t->n_children[v] == 0; }
static int32_t nvals(const bfo_tree *t, int32_t v) { return t->val_end[v] - t->val_begin[v]; }
static const int32_t *vals(const bfo_tree *t, int32_t v) { return t->perm + t->val_begin[v]; }

static levels_t traverseandpad(const bfo_tree *t, int32_t root) {
    levels_t L = {0, NULL};
    ivec cur = {0};
    iv_push(&cur, root);
    while (cur.n > 0) {
        L.lev = (ivec *)xrealloc(L.lev, sizeof(ivec) * (L.nlev + 1));
        L.lev[L.nlev++] = cur;
        ivec nxt = {0};
        for (int64_t i = 0; i < cur.n; ++i) {
            int32_t v = cur.v[i];
            for (int32_t c = 0; c < t->n_children[v]; ++c) iv_push(&nxt, t->first_child[v] + c);
        }
        cur = nxt;
    }
    iv_free(&cur);
    /* for l in 2:(length-1): leaves at level l are appended to level l+1 */
    for (int32_t l = 1; l < L.nlev - 1; ++l)
        for (int64_t i = 0; i < L.lev[l].n; ++i)
            if (isleaf(t, L.lev[l].v[i])) iv_push(&L.lev[l + 1], L.lev[l].v[i]);
    return L;
}
static void levels_free(levels_t *L) {
    for (int32_t i = 0; i < L->nlev; ++i) iv_free(&L->lev[i]);
    free(L->lev);
}

/* ------------------------------------------------------------------------ */
/* Skeleton map K[(S, O)] -> index list (bfassembly.jl:77, fastkey :8-10)   */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint64_t *key;
    int32_t **val;
    int32_t *len;
    int64_t cap, n;
} kmap;
static uint64_t fastkey(int32_t s, int32_t o) {
    return ((uint64_t)(uint32_t)s << 32) | (uint64_t)(uint32_t)o;
}
static void km_init(kmap *m) {
    m->cap = 1024;
    m->n = 0;
    m->key = (uint64_t *)xmalloc(sizeof(uint64_t) * m->cap);
    m->val = (int32_t **)xmalloc(sizeof(int32_t *) * m->cap);
    m->len = (int32_t *)xmalloc(sizeof(int32_t) * m->cap);
    for (int64_t i = 0; i < m->cap; ++i) m->key[i] = UINT64_MAX;
}
static int64_t km_find(const kmap *m, uint64_t k) {
    uint64_t h = bfo_mix64(k) & (uint64_t)(m->cap - 1);
    while (m->key[h] != UINT64_MAX && m->key[h] != k) h = (h + 1) & (uint64_t)(m->cap - 1);
    return (int64_t)h;
}
static void km_put(kmap *m, uint64_t k, const int32_t *v, int32_t n);
static void km_grow(kmap *m) {
    kmap o = *m;
    m->cap *= 2;
    m->n = 0;
    m->key = (uint64_t *)xmalloc(sizeof(uint64_t) * m->cap);
    m->val = (int32_t **)xmalloc(sizeof(int32_t *) * m->cap);
    m->len = (int32_t *)xmalloc(sizeof(int32_t) * m->cap);
    for (int64_t i = 0; i < m->cap; ++i) m->key[i] = UINT64_MAX;
    for (int64_t i = 0; i < o.cap; ++i)
        if (o.key[i] != UINT64_MAX) {
            int64_t h = km_find(m, o.key[i]);
            m->key[h] = o.key[i];
            m->val[h] = o.val[i];
            m->len[h] = o.len[i];
            m->n++;
        }
    free(o.key);
    free(o.val);
    free(o.len);
}
static void km_put(kmap *m, uint64_t k, const int32_t *v, int32_t n) {
    if (2 * (m->n + 1) > m->cap) km_grow(m);
    int64_t h = km_find(m, k);
    int32_t *copy = (int32_t *)xmalloc(sizeof(int32_t) * (n > 0 ? n : 1));
    if (n) memcpy(copy, v, sizeof(int32_t) * n);
    if (m->key[h] == UINT64_MAX) {
        m->key[h] = k;
        m->n++;
    } else {
        free(m->val[h]);
    }
    m->val[h] = copy;
    m->len[h] = n;
}
static const int32_t *km_get(const kmap *m, uint64_t k, int32_t *n) {
    int64_t h = km_find(m, k);
    if (m->key[h] == UINT64_MAX) {
        fprintf(stderr, "bfo: missing skeleton key (%d,%d)\n", (int)(k >> 32),
                (int)(k & 0xffffffffu));
        abort();
    }
    *n = m->len[h];
    return m->val[h];
}
static void km_free(kmap *m) {
    for (int64_t i = 0; i < m->cap; ++i)
        if (m->key[i] != UINT64_MAX) free(m->val[i]);
    free(m->key);
    free(m->val);
    free(m->len);
}

/* ------------------------------------------------------------------------ */
/* Factorization data model: ButterflyBlock / ButterflyLevel /               */
/* ButterflyFactorization (src/ButterflyFactorization/bfstructs.jl:21-147). */
/* ------------------------------------------------------------------------ */
typedef struct {
    int32_t obs_out, src_out, obs_in, src_in;
    int32_t m, n, ident;
    cplx *data; /* NULL for identity */
    int32_t out_ptr, in_ptr;
} blk;

typedef struct {
    blk *b;
    int64_t n, cap;
} blist;

static void bl_push(blist *l, blk b) {
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 16;
        l->b = (blk *)xrealloc(l->b, sizeof(blk) * l->cap);
    }
    l->b[l->n++] = b;
}

typedef struct {
    int32_t stage, s, o, len;
    int64_t ofs;
} skrec;

struct bfo_bf {
    int32_t nst; /* stages: 0 = Q, 1..nst-2 = R levels, nst-1 = P */
    blist *st;
    const bfo_tree *obsT, *srcT; /* observer (test) and source (trial) trees */
    int32_t no, ns;
    int32_t nslots_level_max;
    /* logs */
    skrec *sk;
    int64_t nsk, capsk;
    ivec skidx;
    ivec comp; /* 6 ints per compression */
    ivec judge; /* per compression: 0 not judged, 1 accepted, 2 redone by plain adaptive RRQR */
};

static blk mkblk(int32_t oo, int32_t so, int32_t oi, int32_t si, int32_t m, int32_t n,
                 cplx *data) {
    blk b = {oo, so, oi, si, m, n, data == NULL, data, -1, -1};
    return b;
}

static void log_skel(bfo_bf *bf, int32_t stage, int32_t s, int32_t o, const int32_t *idx,
                     int32_t n) {
    if (bf->nsk == bf->capsk) {
        bf->capsk = bf->capsk ? 2 * bf->capsk : 64;
        bf->sk = (skrec *)xrealloc(bf->sk, sizeof(skrec) * bf->capsk);
    }
    skrec r = {stage, s, o, n, bf->skidx.n};
    bf->sk[bf->nsk++] = r;
    for (int32_t i = 0; i < n; ++i) iv_push(&bf->skidx, idx[i]);
}

/* wire out_ptr / in_ptr exactly as the inner constructor does
 * (bfstructs.jl:107-145): one shared key -> slot map per level, slots
 * numbered in first-use order. */
typedef struct {
    uint64_t *key;
    int32_t *val;
    int64_t cap, n;
} pmap;
static void pm_init(pmap *m, int64_t n) {
    m->cap = 16;
    while (m->cap < 4 * n) m->cap <<= 1;
    m->n = 0;
    m->key = (uint64_t *)xmalloc(sizeof(uint64_t) * m->cap);
    m->val = (int32_t *)xmalloc(sizeof(int32_t) * m->cap);
    for (int64_t i = 0; i < m->cap; ++i) m->key[i] = UINT64_MAX;
}
static int32_t pm_get(pmap *m, uint64_t k) {
    uint64_t h = bfo_mix64(k) & (uint64_t)(m->cap - 1);
    while (m->key[h] != UINT64_MAX && m->key[h] != k) h = (h + 1) & (uint64_t)(m->cap - 1);
    if (m->key[h] == UINT64_MAX) {
        m->key[h] = k;
        m->val[h] = (int32_t)m->n++;
    }
    return m->val[h];
}
static void pm_free(pmap *m) {
    free(m->key);
    free(m->val);
}

static void wire_ptrs(bfo_bf *bf) {
    int32_t nlev = bf->nst - 1; /* slot levels 1..L */
    pmap *pm = (pmap *)xmalloc(sizeof(pmap) * nlev);
    for (int32_t l = 0; l < nlev; ++l) pm_init(&pm[l], bf->st[l].n + bf->st[l + 1].n + 16);
    for (int32_t s = 0; s < bf->nst; ++s) {
        for (int64_t i = 0; i < bf->st[s].n; ++i) {
            blk *b = &bf->st[s].b[i];
            if (s > 0) b->in_ptr = pm_get(&pm[s - 1], fastkey(b->obs_in, b->src_in));
            else b->in_ptr = -1;
            if (s < bf->nst - 1) b->out_ptr = pm_get(&pm[s], fastkey(b->obs_out, b->src_out));
            else b->out_ptr = -1;
        }
    }
    bf->nslots_level_max = 0;
    for (int32_t l = 0; l < nlev; ++l) {
        if (pm[l].n > bf->nslots_level_max) bf->nslots_level_max = (int32_t)pm[l].n;
        pm_free(&pm[l]);
    }
    free(pm);
}

/* one compression with the configured rank estimator; appends the skeleton
 * to *skel (resized) and returns T (r x n_src) */
static cplx *do_compress(bfo_bf *bf, const bfo_opts *o, const double *obs_xyz,
                         const double *src_xyz, int32_t stage, int32_t S, int32_t O,
                         const int32_t *src_idx, int32_t n_src, const int32_t *obs_idx,
                         int32_t n_obs, double eps, int32_t **skel, int32_t *r_out,
                         int64_t *fixed_pos) {
    int32_t guess;
    if (o->rank_mode == 1) {
        const bfo_tree *sT = bf->srcT, *oT = bf->obsT;
        guess = bfo_rank_geometric(o->k, sT->center + 3 * S, oT->center + 3 * O,
                                   sT->halfsize[S] * sqrt(3.0), oT->halfsize[O] * sqrt(3.0),
                                   eps, o->c_geo, o->c_eps, o->rmin, NULL, NULL);
    } else if (o->rank_mode == 2) {
        guess = (*fixed_pos < o->n_fixed) ? o->fixed_guess[*fixed_pos] : 10;
        (*fixed_pos)++;
    } else {
        guess = bfo_rank_butterfly(eps, o->c_tau, o->rmin, NULL, NULL);
    }
    int32_t cap = n_src;
    cplx *T = (cplx *)xmalloc(sizeof(cplx) * (size_t)(cap > 0 ? cap : 1) * (n_src > 0 ? n_src : 1));
    *skel = (int32_t *)xrealloc(*skel, sizeof(int32_t) * (n_src > 0 ? n_src : 1));
    uint64_t key = bfo_sample_key(o->seed, bf->no, bf->ns, stage, S, O);
    int32_t rows = 0;
    int32_t r = bfo_compress(o->k, obs_xyz, obs_idx, n_obs, src_xyz, src_idx, n_src, guess, eps,
                             o->adaptive, key, (double *)T, *skel, &rows);
    int32_t verdict = 0;
    /* judged: samples that left observers unseen, at most 256 columns */
    if (o->judge_rows > 0 && n_src > 0 && n_src <= 256 && rows < n_obs) {
        double e_max = 0.0, z_max = 0.0;
        bfo_judge(o->k, obs_xyz, obs_idx, n_obs, src_xyz, src_idx, n_src, key, rows, o->judge_rows,
                  (const double *)T, *skel, r, &e_max, &z_max);
        verdict = 1;
        if (!(e_max <= o->judge_slack * eps * z_max)) {
            /* plain adaptive RRQR from the butterfly estimator's rows (its
             * floor of 10), and at least twice the rows the prediction used */
            int32_t g = bfo_rank_butterfly(eps, o->c_tau, 10, NULL, NULL);
            if (g < 2 * rows) g = 2 * rows;
            r = bfo_compress(o->k, obs_xyz, obs_idx, n_obs, src_xyz, src_idx, n_src, g, eps, o->adaptive, key,
                             (double *)T, *skel, &rows);
            verdict = 2;
        }
    }
    iv_push(&bf->judge, verdict);
    iv_push(&bf->comp, stage);
    iv_push(&bf->comp, S);
    iv_push(&bf->comp, O);
    iv_push(&bf->comp, guess);
    iv_push(&bf->comp, rows);
    iv_push(&bf->comp, r);
    *r_out = r;
    return T;
}

/* split the columns of T (r x |U|) by the child skeleton lengths: the R
 * blocks are column slices of one compression (bfassembly.jl:296-307) */
static cplx *colslice(const cplx *T, int32_t r, int32_t c0, int32_t nc) {
    cplx *d = (cplx *)xmalloc(sizeof(cplx) * (size_t)(r > 0 ? r : 1) * (nc > 0 ? nc : 1));
    if (r > 0 && nc > 0) memcpy(d, T + (int64_t)c0 * r, sizeof(cplx) * (size_t)r * nc);
    return d;
}

/* U(S, O): concatenation of the children's skeletons, or the node's own
 * skeleton for a leaf (build_U_skeletons, bfassembly.jl:206-230). */
static int32_t *build_U(const bfo_tree *sT, const kmap *K, int32_t S, int32_t O, int32_t *n) {
    int32_t tot = 0, len;
    if (isleaf(sT, S)) {
        const int32_t *v = km_get(K, fastkey(S, O), &len);
        int32_t *u = (int32_t *)xmalloc(sizeof(int32_t) * (len > 0 ? len : 1));
        if (len) memcpy(u, v, sizeof(int32_t) * len);
        *n = len;
        return u;
    }
    for (int32_t c = 0; c < sT->n_children[S]; ++c) {
        km_get(K, fastkey(sT->first_child[S] + c, O), &len);
        tot += len;
    }
    int32_t *u = (int32_t *)xmalloc(sizeof(int32_t) * (tot > 0 ? tot : 1));
    int32_t p = 0;
    for (int32_t c = 0; c < sT->n_children[S]; ++c) {
        const int32_t *v = km_get(K, fastkey(sT->first_child[S] + c, O), &len);
        memcpy(u + p, v, sizeof(int32_t) * len);
        p += len;
    }
    *n = tot;
    return u;
}

typedef struct {
    int32_t s, o, n;
    int32_t *v;
} kupd;
typedef struct {
    kupd *u;
    int64_t n, cap;
} kulist;
static void ku_push(kulist *l, int32_t s, int32_t o, const int32_t *v, int32_t n) {
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 16;
        l->u = (kupd *)xrealloc(l->u, sizeof(kupd) * l->cap);
    }
    int32_t *c = (int32_t *)xmalloc(sizeof(int32_t) * (n > 0 ? n : 1));
    if (n) memcpy(c, v, sizeof(int32_t) * n);
    kupd k = {s, o, n, c};
    l->u[l->n++] = k;
}

/* Emit the R blocks of one compression whose input is U(Svert, Overt):
 * one block per source child (column slice), or the whole T for a leaf
 * Svert.  obs_out is the compression's row node. */
static void emit_R(const bfo_tree *sT, const kmap *K, blist *lvl, int32_t Oout, int32_t Svert,
                   int32_t Overt, const cplx *T, int32_t r, int32_t nU) {
    if (isleaf(sT, Svert)) {
        bl_push(lvl, mkblk(Oout, Svert, Overt, Svert, r, nU, colslice(T, r, 0, nU)));
        return;
    }
    int32_t last = 0, ks;
    for (int32_t c = 0; c < sT->n_children[Svert]; ++c) {
        int32_t Sc = sT->first_child[Svert] + c;
        km_get(K, fastkey(Sc, Overt), &ks);
        bl_push(lvl, mkblk(Oout, Svert, Overt, Sc, r, ks, colslice(T, r, last, ks)));
        last += ks;
    }
}

/* assemble_BF, bfassembly.jl:48-200 with the builders of :206-481. */
bfo_bf *bfo_assemble(const bfo_tree *obsT, const double *obs_xyz, const bfo_tree *srcT,
                     const double *src_xyz, int32_t no, int32_t ns, const bfo_opts *o) {
    levels_t tS = traverseandpad(srcT, ns), tO = traverseandpad(obsT, no);
    int32_t LS = tS.nlev, LO = tO.nlev, L = LS > LO ? LS : LO;
    double eps = o->tol / L;
    bfo_bf *bf = (bfo_bf *)calloc(1, sizeof(bfo_bf));
    bf->nst = L + 1;
    bf->st = (blist *)calloc(bf->nst, sizeof(blist));
    bf->obsT = obsT;
    bf->srcT = srcT;
    bf->no = no;
    bf->ns = ns;
    kmap K;
    km_init(&K);
    int64_t fixed_pos = 0;
    int32_t *skel = NULL;

    /* 1. Q: each source leaf against all observer values of `no` (:82-102) */
    {
        ivec *lv = &tS.lev[LS - 1];
        for (int64_t i = 0; i < lv->n; ++i) {
            int32_t Sl = lv->v[i], r;
            cplx *T = do_compress(bf, o, obs_xyz, src_xyz, 0, Sl, no, vals(srcT, Sl),
                                  nvals(srcT, Sl), vals(obsT, no), nvals(obsT, no), eps, &skel,
                                  &r, &fixed_pos);
            bl_push(&bf->st[0], mkblk(no, Sl, no, Sl, r, nvals(srcT, Sl), T));
            km_put(&K, fastkey(Sl, no), skel, r);
            log_skel(bf, 0, Sl, no, skel, r);
        }
    }

    /* 2. R levels (:107-181) */
    for (int32_t l = 1; l <= L - 1; ++l) {
        int sf = l >= LS, of = l >= LO;
        blist *lvl = &bf->st[l];
        kulist upd = {0};
        if (!sf && !of) {
            /* build_nonfrozen_R_blocks (:232-349); interactions are
             * product(treeO[l], treeS[LS-l]) with the observer index fastest */
            ivec *Os = &tO.lev[l - 1], *Ss = &tS.lev[LS - l - 1];
            for (int64_t si = 0; si < Ss->n; ++si)
                for (int64_t oi = 0; oi < Os->n; ++oi) {
                    int32_t Ov = Os->v[oi], Sv = Ss->v[si], nU, r;
                    int32_t *U = build_U(srcT, &K, Sv, Ov, &nU);
                    if (!isleaf(obsT, Ov)) {
                        for (int32_t c = 0; c < obsT->n_children[Ov]; ++c) {
                            int32_t Oc = obsT->first_child[Ov] + c;
                            cplx *T = do_compress(bf, o, obs_xyz, src_xyz, l, Sv, Oc, U, nU,
                                                  vals(obsT, Oc), nvals(obsT, Oc), eps, &skel,
                                                  &r, &fixed_pos);
                            ku_push(&upd, Sv, Oc, skel, r);
                            emit_R(srcT, &K, lvl, Oc, Sv, Ov, T, r, nU);
                            free(T);
                        }
                    } else if (!isleaf(srcT, Sv)) {
                        cplx *T = do_compress(bf, o, obs_xyz, src_xyz, l, Sv, Ov, U, nU,
                                              vals(obsT, Ov), nvals(obsT, Ov), eps, &skel, &r,
                                              &fixed_pos);
                        emit_R(srcT, &K, lvl, Ov, Sv, Ov, T, r, nU);
                        ku_push(&upd, Sv, Ov, skel, r);
                        free(T);
                    } else {
                        int32_t len;
                        const int32_t *kv = km_get(&K, fastkey(Sv, Ov), &len);
                        bl_push(lvl, mkblk(Ov, Sv, Ov, Sv, len, len, NULL));
                        ku_push(&upd, Sv, Ov, kv, len);
                    }
                    free(U);
                }
        } else if (sf && !of) {
            /* build_sourcefrozen_R_blocks (:351-412) */
            ivec *Os = &tO.lev[l - 1];
            int32_t Sv = ns;
            for (int64_t oi = 0; oi < Os->n; ++oi) {
                int32_t Ov = Os->v[oi];
                if (!isleaf(obsT, Ov)) {
                    for (int32_t c = 0; c < obsT->n_children[Ov]; ++c) {
                        int32_t Oc = obsT->first_child[Ov] + c, nS, r;
                        const int32_t *src = km_get(&K, fastkey(Sv, Ov), &nS);
                        int32_t *srcc = (int32_t *)xmalloc(sizeof(int32_t) * (nS > 0 ? nS : 1));
                        if (nS) memcpy(srcc, src, sizeof(int32_t) * nS);
                        cplx *T = do_compress(bf, o, obs_xyz, src_xyz, l, Sv, Oc, srcc, nS,
                                              vals(obsT, Oc), nvals(obsT, Oc), eps, &skel, &r,
                                              &fixed_pos);
                        bl_push(lvl, mkblk(Oc, Sv, Ov, Sv, r, nS, T));
                        ku_push(&upd, Sv, Oc, skel, r);
                        free(srcc);
                    }
                } else {
                    int32_t len;
                    const int32_t *kv = km_get(&K, fastkey(Sv, Ov), &len);
                    bl_push(lvl, mkblk(Ov, Sv, Ov, Sv, len, len, NULL));
                    ku_push(&upd, Sv, Ov, kv, len);
                }
            }
        } else if (!sf && of) {
            /* build_observerfrozen_R_blocks (:414-481) over treeO[LO] */
            ivec *Os = &tO.lev[LO - 1], *Ss = &tS.lev[LS - l - 1];
            for (int64_t si = 0; si < Ss->n; ++si)
                for (int64_t oi = 0; oi < Os->n; ++oi) {
                    int32_t Ov = Os->v[oi], Sv = Ss->v[si], nU, r;
                    if (!isleaf(srcT, Sv)) {
                        int32_t *U = build_U(srcT, &K, Sv, Ov, &nU);
                        cplx *T = do_compress(bf, o, obs_xyz, src_xyz, l, Sv, Ov, U, nU,
                                              vals(obsT, Ov), nvals(obsT, Ov), eps, &skel, &r,
                                              &fixed_pos);
                        emit_R(srcT, &K, lvl, Ov, Sv, Ov, T, r, nU);
                        ku_push(&upd, Sv, Ov, skel, r);
                        free(T);
                        free(U);
                    } else {
                        int32_t len;
                        const int32_t *kv = km_get(&K, fastkey(Sv, Ov), &len);
                        bl_push(lvl, mkblk(Ov, Sv, Ov, Sv, len, len, NULL));
                        ku_push(&upd, Sv, Ov, kv, len);
                    }
                }
        }
        /* sequential K updates between levels (:177-180) */
        for (int64_t i = 0; i < upd.n; ++i) {
            km_put(&K, fastkey(upd.u[i].s, upd.u[i].o), upd.u[i].v, upd.u[i].n);
            log_skel(bf, l, upd.u[i].s, upd.u[i].o, upd.u[i].v, upd.u[i].n);
            free(upd.u[i].v);
        }
        free(upd.u);
    }

    /* 3. P: dense kernel blocks K(values(Oleaf), K[(ns, Oleaf)]) (:186-197) */
    {
        ivec *lv = &tO.lev[LO - 1];
        blist *P = &bf->st[bf->nst - 1];
        for (int64_t i = 0; i < lv->n; ++i) {
            int32_t Ol = lv->v[i], nc;
            const int32_t *col = km_get(&K, fastkey(ns, Ol), &nc);
            int32_t nr = nvals(obsT, Ol);
            cplx *Z = (cplx *)xmalloc(sizeof(cplx) * (size_t)(nr > 0 ? nr : 1) * (nc > 0 ? nc : 1));
            bfo_helmholtz_fill(o->k, obs_xyz, vals(obsT, Ol), nr, src_xyz, col, nc, (double *)Z);
            bl_push(P, mkblk(Ol, ns, Ol, ns, nr, nc, Z));
        }
    }
    free(skel);
    km_free(&K);
    levels_free(&tS);
    levels_free(&tO);
    wire_ptrs(bf);
    return bf;
}

void bfo_bf_free(bfo_bf *bf) {
    if (!bf) return;
    for (int32_t s = 0; s < bf->nst; ++s) {
        for (int64_t i = 0; i < bf->st[s].n; ++i) free(bf->st[s].b[i].data);
        free(bf->st[s].b);
    }
    free(bf->st);
    free(bf->sk);
    iv_free(&bf->skidx);
    iv_free(&bf->comp);
    iv_free(&bf->judge);
    free(bf);
}

/* adjoint / transpose: bfadjtr.jl:13-24 (transform_block swaps the keys and
 * materialises op(data)), :43-89 (P' becomes Q, R levels reversed, Q'
 * becomes P), reverse_tree :32-35 (the trees swap roles). */
bfo_bf *bfo_bf_transform(const bfo_bf *bf, int32_t conjugate) {
    bfo_bf *t = (bfo_bf *)calloc(1, sizeof(bfo_bf));
    t->nst = bf->nst;
    t->st = (blist *)calloc(t->nst, sizeof(blist));
    t->obsT = bf->srcT;
    t->srcT = bf->obsT;
    t->no = bf->ns;
    t->ns = bf->no;
    for (int32_t s = 0; s < bf->nst; ++s) {
        const blist *src = &bf->st[bf->nst - 1 - s];
        for (int64_t i = 0; i < src->n; ++i) {
            const blk *b = &src->b[i];
            cplx *d = NULL;
            if (!b->ident) {
                d = (cplx *)xmalloc(sizeof(cplx) * (size_t)(b->m > 0 ? b->m : 1) * (b->n > 0 ? b->n : 1));
                for (int32_t c = 0; c < b->n; ++c)
                    for (int32_t r = 0; r < b->m; ++r) {
                        cplx v = b->data[r + (int64_t)c * b->m];
                        d[c + (int64_t)r * b->n] = conjugate ? conj(v) : v;
                    }
            }
            blk nb = mkblk(b->src_in, b->obs_in, b->src_out, b->obs_out, b->n, b->m, d);
            nb.ident = b->ident;
            bl_push(&t->st[s], nb);
        }
    }
    wire_ptrs(t);
    return t;
}

int32_t bfo_bf_n_stages(const bfo_bf *bf) { return bf->nst; }
int64_t bfo_bf_n_blocks(const bfo_bf *bf) {
    int64_t n = 0;
    for (int32_t s = 0; s < bf->nst; ++s) n += bf->st[s].n;
    return n;
}
int64_t bfo_bf_n_data(const bfo_bf *bf) {
    int64_t n = 0;
    for (int32_t s = 0; s < bf->nst; ++s)
        for (int64_t i = 0; i < bf->st[s].n; ++i)
            if (!bf->st[s].b[i].ident) n += (int64_t)bf->st[s].b[i].m * bf->st[s].b[i].n;
    return n;
}
void bfo_bf_export(const bfo_bf *bf, int32_t *table, int64_t *data_ofs, double *data) {
    int64_t q = 0, d = 0;
    for (int32_t s = 0; s < bf->nst; ++s)
        for (int64_t i = 0; i < bf->st[s].n; ++i, ++q) {
            const blk *b = &bf->st[s].b[i];
            int32_t *row = table + 10 * q;
            row[0] = s;
            row[1] = b->obs_out;
            row[2] = b->src_out;
            row[3] = b->obs_in;
            row[4] = b->src_in;
            row[5] = b->m;
            row[6] = b->n;
            row[7] = b->ident;
            row[8] = b->out_ptr;
            row[9] = b->in_ptr;
            if (b->ident) {
                data_ofs[q] = -1;
            } else {
                data_ofs[q] = d;
                if (data) memcpy(data + 2 * d, b->data, sizeof(cplx) * (size_t)b->m * b->n);
                d += (int64_t)b->m * b->n;
            }
        }
}
int64_t bfo_bf_n_skel(const bfo_bf *bf) { return bf->nsk; }
int64_t bfo_bf_n_skel_idx(const bfo_bf *bf) { return bf->skidx.n; }
void bfo_bf_export_skel(const bfo_bf *bf, int32_t *rec, int32_t *idx) {
    for (int64_t i = 0; i < bf->nsk; ++i) {
        rec[5 * i + 0] = bf->sk[i].stage;
        rec[5 * i + 1] = bf->sk[i].s;
        rec[5 * i + 2] = bf->sk[i].o;
        rec[5 * i + 3] = (int32_t)bf->sk[i].ofs;
        rec[5 * i + 4] = bf->sk[i].len;
    }
    if (bf->skidx.n) memcpy(idx, bf->skidx.v, sizeof(int32_t) * bf->skidx.n);
}
int64_t bfo_bf_n_comp(const bfo_bf *bf) { return bf->comp.n / 6; }
void bfo_bf_export_judge(const bfo_bf *bf, int32_t *v) {
    if (bf->judge.n) memcpy(v, bf->judge.v, sizeof(int32_t) * bf->judge.n);
}
void bfo_bf_export_comp(const bfo_bf *bf, int32_t *rec) {
    if (bf->comp.n) memcpy(rec, bf->comp.v, sizeof(int32_t) * bf->comp.n);
}

/* size(BF) = (max(values(test root)), max(values(trial root))) in 1-based
 * terms (bfdims.jl:54-64); returned here as a count. */
void bfo_bf_dims(const bfo_bf *bf, int64_t *rows, int64_t *cols) {
    int32_t mr = -1, mc = -1;
    for (int32_t i = 0; i < nvals(bf->obsT, bf->no); ++i)
        if (vals(bf->obsT, bf->no)[i] > mr) mr = vals(bf->obsT, bf->no)[i];
    for (int32_t i = 0; i < nvals(bf->srcT, bf->ns); ++i)
        if (vals(bf->srcT, bf->ns)[i] > mc) mc = vals(bf->srcT, bf->ns)[i];
    *rows = mr + 1;
    *cols = mc + 1;
}

bfo_bf *bfo_bf_import(const bfo_tree *obsT, const bfo_tree *srcT, int32_t n_stages,
                      int64_t n_blocks, const int32_t *table, const int64_t *data_ofs,
                      const double *data) {
    bfo_bf *bf = (bfo_bf *)calloc(1, sizeof(bfo_bf));
    bf->nst = n_stages;
    bf->st = (blist *)calloc(n_stages, sizeof(blist));
    bf->obsT = obsT;
    bf->srcT = srcT;
    for (int64_t q = 0; q < n_blocks; ++q) {
        const int32_t *row = table + 10 * q;
        cplx *d = NULL;
        if (!row[7]) {
            size_t cnt = (size_t)row[5] * row[6];
            d = (cplx *)xmalloc(sizeof(cplx) * (cnt ? cnt : 1));
            if (cnt) memcpy(d, data + 2 * data_ofs[q], sizeof(cplx) * cnt);
        }
        blk b = mkblk(row[1], row[2], row[3], row[4], row[5], row[6], d);
        b.ident = row[7];
        bl_push(&bf->st[row[0]], b);
    }
    bf->no = bf->st[0].n ? bf->st[0].b[0].obs_out : 0;
    bf->ns = bf->st[n_stages - 1].n ? bf->st[n_stages - 1].b[0].src_out : 0;
    wire_ptrs(bf);
    return bf;
}

/* ------------------------------------------------------------------------ */
/* Flat-workspace matvec, bfvector.jl:200-296: y <- beta y; Q gathers        */
/* x[values(src_in)] into level-1 slots; each R level zeroes its output     */
/* slots then accumulates (identity blocks add); P adds alpha * data * v    */
/* into y[values(obs_out)].  Multi-RHS is column by column (bfmatrix.jl).  */
/* ------------------------------------------------------------------------ */
static void matvec1(const bfo_bf *bf, const cplx *x, cplx *y, cplx alpha, cplx beta) {
    int64_t nr, nc;
    bfo_bf_dims(bf, &nr, &nc);
    if (beta != 1.0) {
        for (int64_t i = 0; i < nr; ++i) y[i] = (beta == 0.0) ? 0.0 : beta * y[i];
    }
    int32_t nlev = bf->nst - 1;
    cplx **buf = (cplx **)calloc(nlev, sizeof(cplx *));
    int32_t **lens = (int32_t **)calloc(nlev, sizeof(int32_t *));
    int32_t *nsl = (int32_t *)calloc(nlev, sizeof(int32_t));
    /* slot capacity per level: count and max length */
    int32_t maxlen = 1;
    for (int32_t s = 0; s < bf->nst; ++s)
        for (int64_t i = 0; i < bf->st[s].n; ++i) {
            const blk *b = &bf->st[s].b[i];
            if (b->m > maxlen) maxlen = b->m;
            if (b->n > maxlen) maxlen = b->n;
            if (s < nlev && b->out_ptr + 1 > nsl[s]) nsl[s] = b->out_ptr + 1;
        }
    for (int32_t l = 0; l < nlev; ++l) {
        buf[l] = (cplx *)calloc((size_t)(nsl[l] > 0 ? nsl[l] : 1) * maxlen, sizeof(cplx));
        lens[l] = (int32_t *)calloc(nsl[l] > 0 ? nsl[l] : 1, sizeof(int32_t));
    }
    /* Q */
    for (int64_t i = 0; i < bf->st[0].n; ++i) {
        const blk *b = &bf->st[0].b[i];
        const int32_t *idx = vals(bf->srcT, b->src_in);
        int32_t n = nvals(bf->srcT, b->src_in);
        cplx *out = buf[0] + (int64_t)b->out_ptr * maxlen;
        if (b->ident) {
            for (int32_t r = 0; r < n; ++r) out[r] = x[idx[r]];
            lens[0][b->out_ptr] = n;
        } else {
            for (int32_t r = 0; r < b->m; ++r) out[r] = 0.0;
            for (int32_t c = 0; c < b->n; ++c) {
                cplx xv = x[idx[c]];
                const cplx *col = b->data + (int64_t)c * b->m;
                for (int32_t r = 0; r < b->m; ++r) out[r] += col[r] * xv;
            }
            lens[0][b->out_ptr] = b->m;
        }
    }
    /* R levels */
    for (int32_t s = 1; s < bf->nst - 1; ++s) {
        cplx *bi = buf[s - 1], *bo = buf[s];
        for (int64_t i = 0; i < bf->st[s].n; ++i) {
            const blk *b = &bf->st[s].b[i];
            int32_t len = b->ident ? lens[s - 1][b->in_ptr] : b->m;
            cplx *out = bo + (int64_t)b->out_ptr * maxlen;
            for (int32_t r = 0; r < len; ++r) out[r] = 0.0;
            lens[s][b->out_ptr] = len;
        }
        for (int64_t i = 0; i < bf->st[s].n; ++i) {
            const blk *b = &bf->st[s].b[i];
            const cplx *in = bi + (int64_t)b->in_ptr * maxlen;
            cplx *out = bo + (int64_t)b->out_ptr * maxlen;
            if (b->ident) {
                int32_t len = lens[s - 1][b->in_ptr];
                for (int32_t r = 0; r < len; ++r) out[r] += in[r];
            } else {
                for (int32_t c = 0; c < b->n; ++c) {
                    cplx xv = in[c];
                    const cplx *col = b->data + (int64_t)c * b->m;
                    for (int32_t r = 0; r < b->m; ++r) out[r] += col[r] * xv;
                }
            }
        }
    }
    /* P */
    {
        int32_t s = bf->nst - 1;
        cplx *bi = buf[s - 1];
        for (int64_t i = 0; i < bf->st[s].n; ++i) {
            const blk *b = &bf->st[s].b[i];
            const cplx *in = bi + (int64_t)b->in_ptr * maxlen;
            const int32_t *idx = vals(bf->obsT, b->obs_out);
            if (b->ident) {
                int32_t len = lens[s - 1][b->in_ptr];
                for (int32_t r = 0; r < len; ++r) y[idx[r]] += alpha * in[r];
            } else {
                for (int32_t r = 0; r < b->m; ++r) {
                    cplx acc = 0.0;
                    for (int32_t c = 0; c < b->n; ++c) acc += b->data[r + (int64_t)c * b->m] * in[c];
                    y[idx[r]] += alpha * acc;
                }
            }
        }
    }
    for (int32_t l = 0; l < nlev; ++l) {
        free(buf[l]);
        free(lens[l]);
    }
    free(buf);
    free(lens);
    free(nsl);
}

void bfo_matvec(const bfo_bf *bf, int64_t nrhs, const double *x, int64_t ldx, double *y,
                int64_t ldy, const double *alpha, const double *beta) {
    cplx a = alpha ? alpha[0] + I * alpha[1] : 1.0;
    cplx b = beta ? beta[0] + I * beta[1] : 0.0;
    for (int64_t j = 0; j < nrhs; ++j)
        matvec1(bf, (const cplx *)x + j * ldx, (cplx *)y + j * ldy, a, b);
}

/* Far part of mul!(y, A::PetrovGalerkinBF, x, alpha, beta)
 * (src/matrixalgebra/matrixvector.jl:39-62): per-thread y buffers, static
 * split over the factorizations, serial reduction. */
void bfo_matvec_many(bfo_bf *const *bfs, int32_t nbf, int64_t nrhs, const double *x,
                     int64_t ldx, double *y, int64_t ldy, const double *alpha,
                     const double *beta, int32_t nthreads) {
    cplx a = alpha ? alpha[0] + I * alpha[1] : 1.0;
    cplx b = beta ? beta[0] + I * beta[1] : 0.0;
    int64_t nrow = 0;
    for (int32_t i = 0; i < nbf; ++i) {
        int64_t r, c;
        bfo_bf_dims(bfs[i], &r, &c);
        if (r > nrow) nrow = r;
    }
    if (nthreads < 1) nthreads = 1;
    for (int64_t j = 0; j < nrhs; ++j) {
        cplx *yj = (cplx *)y + j * ldy;
        const cplx *xj = (const cplx *)x + j * ldx;
        if (b != 1.0)
            for (int64_t i = 0; i < nrow; ++i) yj[i] = (b == 0.0) ? 0.0 : b * yj[i];
        cplx *tb = (cplx *)calloc((size_t)nthreads * (nrow > 0 ? nrow : 1), sizeof(cplx));
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
        for (int32_t i = 0; i < nbf; ++i) {
            int tid = 0;
#ifdef _OPENMP
            tid = omp_get_thread_num();
#endif
            matvec1(bfs[i], xj, tb + (int64_t)tid * nrow, 1.0, 1.0);
        }
        for (int32_t t = 0; t < nthreads; ++t)
            for (int64_t i = 0; i < nrow; ++i) yj[i] += a * tb[(int64_t)t * nrow + i];
        free(tb);
    }
}

/* ------------------------------------------------------------------------ */
/* Admissibility: ext/BFH2Trees/h2admissible.jl:11-40 (isFarFunctor on      */
/* boxes: AABB gap > alpha * max radius, touching nodes are near) and        */
/* :74-91 (CenterDistanceAdmissibility: |c_s - c_o| > beta (r_s + r_o)).     */
/* Radii follow ext/BFH2Trees/BFH2Trees.jl:15-16 (halfsize * sqrt(3)).       */
/* ------------------------------------------------------------------------ */
int32_t bfo_isfar_aabb(const bfo_tree *srcT, const bfo_tree *tstT, int32_t s, int32_t o,
                       double alpha) {
    const double *oc = tstT->center + 3 * o, *sc = srcT->center + 3 * s;
    double ol = tstT->halfsize[o] * sqrt(3.0), sl = srcT->halfsize[s] * sqrt(3.0);
    double W = sl > ol ? sl : ol;
    double target = alpha * W, mind = 0.0;
    for (int i = 0; i < 3; ++i) {
        double d = fabs(oc[i] - sc[i]) - (sl + ol);
        if (d > 0.0) mind += d * d;
    }
    if (mind < 1e-10) return 0;
    return sqrt(mind) > target;
}

int32_t bfo_center_distance(const bfo_tree *srcT, const bfo_tree *tstT, int32_t s, int32_t o,
                            double beta) {
    const double *oc = tstT->center + 3 * o, *sc = srcT->center + 3 * s;
    double ol = tstT->halfsize[o] * sqrt(3.0), sl = srcT->halfsize[s] * sqrt(3.0);
    double d = sqrt((sc[0] - oc[0]) * (sc[0] - oc[0]) + (sc[1] - oc[1]) * (sc[1] - oc[1]) +
                    (sc[2] - oc[2]) * (sc[2] - oc[2]));
    if (d < 1e-5) return 0;
    return d > beta * (sl + ol);
}

typedef struct {
    const bfo_tree *s, *t;
    int32_t kind;
    double param;
    int32_t minbflvl, unbal, leafcomp, leafimb;
    int32_t *far, *near;
    int64_t nfar, nnear;
} nf_ctx;

static void nf_push(int32_t *buf, int64_t *n, int32_t o, int32_t s) {
    if (buf) {
        buf[2 * *n] = o;
        buf[2 * *n + 1] = s;
    }
    (*n)++;
}

/* process_nodes!, src/Matrixassembly/intlists.jl:75-217 */
static void process_nodes(nf_ctx *c, int32_t o, int32_t s, int32_t h) {
    int lo = isleaf(c->t, o), ls = isleaf(c->s, s);
    int any = lo || ls, both = lo && ls;
    int adm = c->kind == 0 ? bfo_isfar_aabb(c->s, c->t, s, o, c->param)
                           : bfo_center_distance(c->s, c->t, s, o, c->param);
    if (h <= c->minbflvl && !c->leafcomp) {
        nf_push(c->near, &c->nnear, o, s);
        return;
    }
    if (adm) {
        if (any && !c->leafcomp) nf_push(c->near, &c->nnear, o, s);
        else nf_push(c->far, &c->nfar, o, s);
        return;
    }
    if (both || (any && !c->leafimb)) {
        nf_push(c->near, &c->nnear, o, s);
        return;
    }
    if (c->unbal) {
        /* largernode (treeinterface.jl:153-154): radius_o >= radius_s */
        int larger = c->t->halfsize[o] >= c->s->halfsize[s];
        if ((larger && !lo) || ls) {
            for (int32_t k = 0; k < c->t->n_children[o]; ++k)
                process_nodes(c, c->t->first_child[o] + k, s, h - 1);
        } else {
            for (int32_t k = 0; k < c->s->n_children[s]; ++k)
                process_nodes(c, o, c->s->first_child[s] + k, h - 1);
        }
    } else {
        if (lo) {
            for (int32_t k = 0; k < c->s->n_children[s]; ++k)
                process_nodes(c, o, c->s->first_child[s] + k, h - 1);
        } else if (ls) {
            for (int32_t k = 0; k < c->t->n_children[o]; ++k)
                process_nodes(c, c->t->first_child[o] + k, s, h - 1);
        } else {
            for (int32_t a = 0; a < c->t->n_children[o]; ++a)
                for (int32_t b = 0; b < c->s->n_children[s]; ++b)
                    process_nodes(c, c->t->first_child[o] + a, c->s->first_child[s] + b, h - 1);
        }
    }
}

static int32_t tree_height(const bfo_tree *t) {
    levels_t L = traverseandpad(t, 0);
    int32_t h = L.nlev;
    levels_free(&L);
    return h;
}

/* nearandfar, src/Matrixassembly/intlists.jl:30-59 */
void bfo_nearandfar(const bfo_tree *srcT, const bfo_tree *tstT, int32_t kind, double param,
                    int32_t minbflvl, int32_t unbalancedints, int32_t leafcomp,
                    int32_t leafimbalance, int32_t *far, int64_t *nfar, int32_t *near,
                    int64_t *nnear) {
    nf_ctx c = {srcT, tstT, kind, param, minbflvl, unbalancedints, leafcomp, leafimbalance,
                far, near, 0, 0};
    int32_t ht = tree_height(tstT), hs = tree_height(srcT);
    process_nodes(&c, 0, 0, ht < hs ? ht : hs);
    *nfar = c.nfar;
    *nnear = c.nnear;
}
