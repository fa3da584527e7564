// Near-field part of the global operator (petrovgalerkinbf.jl:55-119; the
// near product of matrixvector.jl:31-37).  Every near pair (observer node O,
// source node S) holds the dense kernel block G(vals(O), vals(S)), as the
// reference's BlockSparseMatrix does.
//
// Layout in HBM: the pairs are grouped by observer node into items; the
// blocks of one item sit side by side as one column-major m x ncol matrix
// (m = |vals(O)|, ncol = sum of |vals(S)| over the item's pairs), so a whole
// item streams as one contiguous array.  Flat columns carry the original
// source index of their point (colx), flat rows the original observer index
// (rowx).  Assembly is one fill launch with one task per item.
//
// Forward product: one workgroup per row chunk (<= 64 rows) of an item, the
// four wavefronts take contiguous quarters of the columns, a lane holds one
// row (or, for items of at most 32 rows, one row of one of 64 / m columns
// loaded together), the inputs of 64 columns are loaded once (one lane per
// column) and handed to the row lanes by readlane or permute; the
// wavefronts are summed through LDS in fixed order.  When the items' rows
// are disjoint and cover every row, the workgroup writes y; otherwise it
// writes per-item partial rows and a second pass sums, for every row of y,
// the partial rows of that point in item order.  Transposed product: a
// wavefront takes 8 columns at a time, lanes over rows, and folds the 8
// column sums out of the 64 lanes by a halving exchange (3 steps that halve
// the values per lane, then 3 full steps); items of at most 32 rows load
// 64 / m columns at once and sum each column's rows through LDS.  The
// per-column results are summed per point of the source side by the same
// second pass.  Every sum has a fixed order: results do not depend on
// scheduling.
#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "device.hpp"
#include "internal.hpp"

namespace bfg {
void set_last_error(const std::string &m);

namespace {

struct NearTask {       // one row chunk of one item (forward) or one item (transposed)
    int64_t fofs;       // entry (row0, 0) of the item's column-major matrix
    int32_t m;          // leading dimension: rows of the item
    int32_t rows;       // rows of the chunk (forward: <= 64; transposed: m)
    int32_t ncol;       // columns of the item
    int32_t col0;       // first flat column of the item
    int32_t rflat;      // flat row of the chunk's first row
    int32_t pad;
};
static_assert(sizeof(NearTask) == 32, "NearTask layout");

template <typename R> __device__ __forceinline__ R bcast(R v, int l);
template <> __device__ __forceinline__ float bcast<float>(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
template <> __device__ __forceinline__ double bcast<double>(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}
template <typename R> __device__ __forceinline__ C<R> shfl_xor_c(C<R> v, int s) {
    return C<R>{__shfl_xor(v.x, s, 64), __shfl_xor(v.y, s, 64)};
}

enum { NEAR_PART = 0, NEAR_Y_SET = 1, NEAR_Y_ACC = 2 };

// mode NEAR_PART: part[r ldo + rflat + i] = acc; NEAR_Y_SET: y = alpha acc;
// NEAR_Y_ACC: y = beta y + alpha acc (y rows through rowx); CONJ multiplies
// by the conjugate entries (the adjoint through a transposed copy).  A chunk of
// `rows` <= 32 rows spans whole columns (rows = m): P = 64 / rows columns are
// loaded per instruction, lane = (column sub, row), so one load reads P
// consecutive columns as one contiguous run; the sub-columns are summed with
// the wavefronts at the end.  With P = 1 a lane is a row and the inputs are
// broadcast by readlane; with P > 1 by cross-lane permute.
template <typename R, int NR, bool CONJ>
__global__ __launch_bounds__(256) void near_fwd_kernel(const NearTask *__restrict__ tasks,
                                                       const C<R> *__restrict__ A,
                                                       const int32_t *__restrict__ colx,
                                                       const int32_t *__restrict__ rowx,
                                                       const C<R> *__restrict__ x, int64_t ldx,
                                                       C<R> *__restrict__ out, int64_t ldo, int nrhs,
                                                       C<R> alpha, C<R> beta, int mode) {
    __shared__ C<R> red[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const NearTask t = tasks[blockIdx.x];
    const int rb = blockIdx.y * NR, nr = min(NR, nrhs - rb);
    const int P = t.rows <= 32 ? 64 / t.rows : 1;
    const int sub = lane / t.rows, i = lane - sub * t.rows;
    const bool act = sub < P;
    const C<R> *a = A + t.fofs + min(i, t.rows - 1);
    const int q = (((t.ncol + 3) >> 2) + 7) & ~7;
    const int cb = w * q, ce = min(t.ncol, cb + q);
    C<R> acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = C<R>{R(0), R(0)};
    for (int c0 = cb; c0 < ce; c0 += 64) {
        const int cn = min(64, ce - c0);
        const bool okc = lane < cn;
        const int64_t xi = colx[t.col0 + c0 + min(lane, cn - 1)];
        C<R> xs[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const C<R> v = x[xi + int64_t(min(rb + r, nrhs - 1)) * ldx];
            xs[r] = C<R>{okc && r < nr ? v.x : R(0), okc && r < nr ? v.y : R(0)};
        }
        if (P == 1) {
            for (int cc = 0; cc < cn; cc += 8) {
                C<R> av[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) av[u] = a[int64_t(c0 + min(cc + u, cn - 1)) * t.m];
#pragma unroll
                for (int u = 0; u < 8; ++u)
#pragma unroll
                    for (int r = 0; r < NR; ++r)
                    {
                        const C<R> xv{bcast(xs[r].x, cc + u), bcast(xs[r].y, cc + u)};
                        if (CONJ) cfmac(acc[r], av[u], xv);
                        else cfma(acc[r], av[u], xv);
                    }
            }
        } else {
            for (int cc = 0; cc < cn; cc += 8 * P) {
                C<R> av[8];
                int kk[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    kk[u] = cc + u * P + sub;
                    av[u] = a[int64_t(c0 + min(kk[u], cn - 1)) * t.m];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const bool ok = act && kk[u] < cn;
                    const int src = min(kk[u], 63);
#pragma unroll
                    for (int r = 0; r < NR; ++r) {
                        const R vx = __shfl(xs[r].x, src, 64), vy = __shfl(xs[r].y, src, 64);
                        const C<R> xv{ok ? vx : R(0), ok ? vy : R(0)};
                        if (CONJ) cfmac(acc[r], av[u], xv);
                        else cfma(acc[r], av[u], xv);
                    }
                }
            }
        }
    }
    // sum over wavefronts and sub-columns, in fixed order, one RHS at a time
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        if (r >= nr) break;
        __syncthreads();
        red[w][lane] = act ? acc[r] : C<R>{R(0), R(0)};
        __syncthreads();
        if (w == 0 && lane < t.rows) {
            C<R> s{R(0), R(0)};
            for (int ww = 0; ww < 4; ++ww)
                for (int sb = 0; sb < P; ++sb) s = cadd(s, red[ww][sb * t.rows + lane]);
            if (mode == NEAR_PART) {
                out[int64_t(rb + r) * ldo + t.rflat + lane] = s;
            } else {
                C<R> *yp = out + int64_t(rb + r) * ldo + rowx[t.rflat + lane];
                *yp = mode == NEAR_Y_SET ? cmul(alpha, s) : cadd(cmul(beta, *yp), cmul(alpha, s));
            }
        }
    }
}

// part[r ldo + col0 + c] = sum_i op(A[i, c]) x[rowx[rflat + i]], op the
// conjugate (CONJ) or plain entry.
template <typename R, int NR, bool CONJ>
__global__ __launch_bounds__(256) void near_tr_kernel(const NearTask *__restrict__ tasks,
                                                      const C<R> *__restrict__ A,
                                                      const int32_t *__restrict__ rowx,
                                                      const C<R> *__restrict__ x, int64_t ldx,
                                                      C<R> *__restrict__ part, int64_t ldo, int nrhs) {
    __shared__ C<R> red[4][8][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const NearTask t = tasks[blockIdx.x];
    const int rb = blockIdx.y * NR, nr = min(NR, nrhs - rb);
    if (t.m <= 32) {
        // short columns: lane = (column sub, row), P = 64 / m columns per
        // load; the products go through this wavefront's LDS rows and one
        // lane per (u, sub) sums the m rows of that column in row order
        const int m = t.m, P = 64 / m, sub = lane / m, i = lane - sub * m;
        const bool act = sub < P;
        const int64_t xi = rowx[t.rflat + i];
        C<R> xv[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const C<R> v = x[xi + int64_t(min(rb + r, nrhs - 1)) * ldx];
            xv[r] = C<R>{act && r < nr ? v.x : R(0), act && r < nr ? v.y : R(0)};
        }
        C<R>(*buf)[64] = red[w];
        for (int c0 = w * 8 * P; c0 < t.ncol; c0 += 32 * P) {
            C<R> av[8];
            int kk[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                kk[u] = c0 + u * P + sub;
                av[u] = A[t.fofs + int64_t(min(kk[u], t.ncol - 1)) * m + i];
            }
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                if (r >= nr) break;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const C<R> pr = CONJ ? cmulc(av[u], xv[r]) : cmul(av[u], xv[r]);
                    buf[u][lane] = act && kk[u] < t.ncol ? pr : C<R>{R(0), R(0)};
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                for (int L = lane; L < 8 * P; L += 64) {  // 8 P columns: more than 64 when m < 8
                    const int u = L / P, sb = L - u * P, c = c0 + u * P + sb;
                    C<R> sum{R(0), R(0)};
                    for (int ii = 0; ii < m; ++ii) sum = cadd(sum, buf[u][sb * m + ii]);
                    if (c < t.ncol) part[int64_t(rb + r) * ldo + t.col0 + c] = sum;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
        return;
    }
    for (int g = w; g * 8 < t.ncol; g += 4) {
        const int c0 = g * 8;
        C<R> p[8][NR];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int r = 0; r < NR; ++r) p[u][r] = C<R>{R(0), R(0)};
        for (int r0 = 0; r0 < t.m; r0 += 64) {
            const int i = r0 + lane;
            const bool ok = i < t.m;
            const int ic = min(i, t.m - 1);
            const int64_t xi = rowx[t.rflat + ic];
            C<R> xv[NR];
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const C<R> v = x[xi + int64_t(min(rb + r, nrhs - 1)) * ldx];
                xv[r] = C<R>{ok && r < nr ? v.x : R(0), ok && r < nr ? v.y : R(0)};
            }
            C<R> av[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) av[u] = A[t.fofs + int64_t(min(c0 + u, t.ncol - 1)) * t.m + ic];
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    if (CONJ) cfmac(p[u][r], av[u], xv[r]);
                    else cfma(p[u][r], av[u], xv[r]);
                }
        }
        // halving exchange: after the steps over lane bits 5, 4, 3 a lane
        // holds one column, (lane >> 3) & 7, summed over the 8 lanes that
        // differ in those bits; lane bits 2, 1, 0 are then summed in full
#pragma unroll
        for (int st = 0; st < 3; ++st) {
            const int s = 32 >> st, K = 8 >> st, H = K >> 1;
            const bool hi = (lane & s) != 0;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (u >= H) break;
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    const C<R> lo_v = p[u][r], hi_v = p[u + H][r];
                    const C<R> recv = shfl_xor_c(hi ? lo_v : hi_v, s);
                    p[u][r] = cadd(hi ? hi_v : lo_v, recv);
                }
            }
        }
#pragma unroll
        for (int s = 4; s > 0; s >>= 1)
#pragma unroll
            for (int r = 0; r < NR; ++r) p[0][r] = cadd(p[0][r], shfl_xor_c(p[0][r], s));
        const int c = c0 + ((lane >> 3) & 7);
        if ((lane & 7) == 0 && c < t.ncol) {
#pragma unroll
            for (int r = 0; r < NR; ++r)
                if (r < nr) part[int64_t(rb + r) * ldo + t.col0 + c] = p[0][r];
        }
    }
}

// Many right-hand sides, complex FP32 on the matrix cores.  The inputs of
// a pass (at most 64 RHS) are gathered RHS-interleaved in source-tree order
// (xg[p ns + j], ns = 16 NT), so the columns of one source leaf are
// consecutive rows of xg.  One workgroup per 16-row chunk of an item, its
// four wavefronts on quarters of the columns (summed through LDS in fixed
// order at the end); per 4 columns: one A load (row li, column lk at lane
// li + 16 lk), one B load per 16-RHS group (column lk, RHS li), four
// v_mfma_f32_16x16x4_f32 per group into real and imaginary accumulators
// (conjugated entries for CONJ); the next 4 columns load while the current
// ones multiply.  Lane
// li + 16 lk then holds rows 4 lk .. 4 lk + 3 of RHS li; rows are disjoint
// across chunks (direct items), so each result is written once.
typedef float near_v4f __attribute__((ext_vector_type(4)));

template <int NT, bool CONJ>
__global__ __launch_bounds__(256) void near_mfma_kernel(const NearTask *__restrict__ tasks, int32_t n_tasks,
                                                        const C<float> *__restrict__ A,
                                                        const int32_t *__restrict__ ctree,
                                                        const int32_t *__restrict__ rowx,
                                                        const C<float> *__restrict__ xg, int32_t ns,
                                                        C<float> *__restrict__ y, int64_t ldy, int nc,
                                                        C<float> alpha, C<float> beta, int beta_zero) {
    __shared__ near_v4f red[3][2 * NT][64];
    const int l = threadIdx.x & 63, li = l & 15, lk = l >> 4;
    const int w = threadIdx.x >> 6;
    const NearTask tk = tasks[blockIdx.x];
    const bool rok = li < tk.rows;
    const C<float> *ap = A + tk.fofs + min(li, tk.rows - 1);
    const int32_t *cp = ctree + tk.col0;
    // wavefront w takes columns [cb, ce), a quarter rounded to 4
    const int q = (((tk.ncol + 3) >> 2) + 3) & ~3;
    const int cb = w * q, ce = min(tk.ncol, cb + q);
    near_v4f accr[NT], acci[NT];
#pragma unroll
    for (int g = 0; g < NT; ++g) accr[g] = acci[g] = near_v4f{0.f, 0.f, 0.f, 0.f};
    if (cb < ce) {
        auto load_a = [&](int k) -> C<float> {
            const C<float> v = ap[int64_t(min(k, ce - 1)) * tk.m];
            return (rok && k < ce) ? v : C<float>{0.f, 0.f};
        };
        C<float> a = load_a(cb + lk), b[NT];
        {
            const C<float> *xp = xg + int64_t(cp[min(cb + lk, ce - 1)]) * ns + li;
#pragma unroll
            for (int g = 0; g < NT; ++g) b[g] = cb + lk < ce ? xp[16 * g] : C<float>{0.f, 0.f};
        }
        for (int k0 = cb; k0 < ce; k0 += 4) {
            const int kn = k0 + 4 + lk;
            const C<float> an = load_a(kn);
            C<float> bn[NT];
            const C<float> *xp = xg + int64_t(cp[min(kn, ce - 1)]) * ns + li;
#pragma unroll
            for (int g = 0; g < NT; ++g) bn[g] = kn < ce ? xp[16 * g] : C<float>{0.f, 0.f};
            const float nai = -a.y;
#pragma unroll
            for (int g = 0; g < NT; ++g) {
                accr[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b[g].x, accr[g], 0, 0, 0);
                acci[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b[g].y, acci[g], 0, 0, 0);
                accr[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(CONJ ? a.y : nai, b[g].y, accr[g], 0, 0, 0);
                acci[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(CONJ ? nai : a.y, b[g].x, acci[g], 0, 0, 0);
            }
            a = an;
#pragma unroll
            for (int g = 0; g < NT; ++g) b[g] = bn[g];
        }
    }
    // the four wavefronts' tiles summed in fixed order
    if (w > 0) {
#pragma unroll
        for (int g = 0; g < NT; ++g) {
            red[w - 1][2 * g][l] = accr[g];
            red[w - 1][2 * g + 1][l] = acci[g];
        }
    }
    __syncthreads();
    if (w != 0) return;
#pragma unroll
    for (int g = 0; g < NT; ++g)
#pragma unroll
        for (int ww = 0; ww < 3; ++ww) {
            accr[g] += red[ww][2 * g][l];
            acci[g] += red[ww][2 * g + 1][l];
        }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int r = 4 * lk + v;
        if (r >= tk.rows) continue;
        const int64_t yrow = rowx[tk.rflat + r];
#pragma unroll
        for (int g = 0; g < NT; ++g) {
            const int j = 16 * g + li;
            if (j >= nc) continue;
            const C<float> val{accr[g][v], acci[g][v]};
            C<float> *yp = y + int64_t(j) * ldy + yrow;
            *yp = beta_zero ? cmul(alpha, val) : cadd(cmul(beta, *yp), cmul(alpha, val));
        }
    }
}

// xg[p ns + j] = x[perm[p] + j ldx] for j < nc, 0 for nc <= j < ns
__global__ __launch_bounds__(256) void near_gather_kernel(const C<float> *__restrict__ x, int64_t ldx,
                                                          const int32_t *__restrict__ perm, int32_t n,
                                                          C<float> *__restrict__ xg, int32_t ns, int nc) {
    const int64_t total = int64_t(n) * ns;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += int64_t(gridDim.x) * 256) {
        const int64_t p = e / ns;
        const int j = int(e - p * ns);
        xg[e] = j < nc ? x[perm[p] + int64_t(j) * ldx] : C<float>{0.f, 0.f};
    }
}

// y[r ldy + v] = beta y + alpha sum_{j in [ptr[v], ptr[v+1])} part[r ldp + idx[j]]
// (beta exactly zero: y is not read)
template <typename R>
__global__ __launch_bounds__(256) void near_sum_kernel(const int64_t *__restrict__ ptr,
                                                       const int32_t *__restrict__ idx, int32_t n_out,
                                                       const C<R> *__restrict__ part, int64_t ldp,
                                                       C<R> *__restrict__ y, int64_t ldy, int nrhs,
                                                       C<R> alpha, C<R> beta, int beta_zero) {
    const int64_t total = int64_t(n_out) * nrhs;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += int64_t(gridDim.x) * 256) {
        const int64_t r = e / n_out, v = e - r * n_out;
        C<R> s{R(0), R(0)};
        for (int64_t j = ptr[v]; j < ptr[v + 1]; ++j) s = cadd(s, part[r * ldp + idx[j]]);
        C<R> *yp = y + r * ldy + v;
        *yp = beta_zero ? cmul(alpha, s) : cadd(cmul(beta, *yp), cmul(alpha, s));
    }
}

struct Csr {
    DevBuf ptr, idx;
    int32_t n = 0;
};

}  // namespace

struct Near {
    Ctx *ctx = nullptr;
    const Tree *test = nullptr, *trial = nullptr;
    int32_t precision = BFGPU_C128;
    int64_t rows = 0, cols = 0, n_pairs = 0, n_items = 0, entries = 0;
    int64_t n_rflat = 0, n_cflat = 0;
    bool direct = false;           // forward items write disjoint rows covering y
    DevBuf data, fwd_tasks, tr_tasks, colx, rowx;
    int32_t n_fwd = 0, n_tr = 0;
    Csr rsum, csum;                // rows of y <- flat rows; columns <- flat columns
    DevBuf part, stage;            // partial sums; host-pointer staging
    // per pair: item, first column within the item, m, n (block export)
    std::vector<int64_t> pair_fofs;
    std::vector<int32_t> pair_m, pair_n;
    // transposed copy for op T / C: the blocks of the swapped pairs over the
    // swapped trees, G(y, x) = G(x, y); built on the first T / C product,
    // skipped (two-pass product instead) when device memory is short
    double k = 0, self_re = 0, self_im = 0;
    std::vector<int32_t> pairs;
    // matrix-core products (complex FP32, direct items): 16-row chunk tasks,
    // source-tree position of every flat column, the source tree's
    // permutation, the RHS-interleaved inputs of a pass
    DevBuf mtasks, ctree, perm_src, xg;
    int32_t n_mtasks = 0;
    int32_t mfma_min_rhs = 8;        // 0: vector ALU only
    std::unique_ptr<Near> tcopy;
    bool tcopy_failed = false, use_tcopy = true;
    // kernel-matrix plug-in whose near quadrature fills the blocks (null:
    // the point kernel); its blocks are not symmetric under the swap, so
    // T / C products take the two-pass product over the stored blocks
    const KernelRWG *kern = nullptr;
};

namespace {

template <class T> void put(DevBuf &d, const std::vector<T> &h, hipStream_t s) {
    d.alloc(sizeof(T) * std::max<size_t>(h.size(), 1));
    if (!h.empty()) BFG_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
}

// CSR of the flat entries of each original index in [0, n): entries in
// increasing flat order
void build_csr(Csr &c, const std::vector<int32_t> &owner, int32_t n, hipStream_t s) {
    std::vector<int64_t> ptr(size_t(n) + 1, 0);
    for (int32_t v : owner) ++ptr[size_t(v) + 1];
    for (int32_t v = 0; v < n; ++v) ptr[v + 1] += ptr[v];
    std::vector<int32_t> idx(owner.size());
    std::vector<int64_t> pos(ptr.begin(), ptr.end() - 1);
    for (size_t j = 0; j < owner.size(); ++j) idx[size_t(pos[size_t(owner[j])]++)] = int32_t(j);
    c.n = n;
    put(c.ptr, ptr, s);
    put(c.idx, idx, s);
    BFG_HIP(hipStreamSynchronize(s));  // the host arrays end here
}

template <typename R>
void run_sum(hipStream_t s, const Csr &c, const void *part, int64_t ldp, void *y, int64_t ldy, int64_t nrhs,
             const double *alpha, const double *beta) {
    const int64_t total = int64_t(c.n) * nrhs;
    if (total <= 0) return;
    const bool bz = beta[0] == 0.0 && beta[1] == 0.0;
    dim3 grid(unsigned(std::min<int64_t>((total + 255) / 256, int64_t(1) << 20)));
    hipLaunchKernelGGL(near_sum_kernel<R>, grid, dim3(256), 0, s, c.ptr.as<int64_t>(), c.idx.as<int32_t>(), c.n,
                       static_cast<const C<R> *>(part), ldp, static_cast<C<R> *>(y), ldy, int(nrhs),
                       C<R>{R(alpha[0]), R(alpha[1])}, C<R>{R(beta[0]), R(beta[1])}, bz ? 1 : 0);
    BFG_HIP(hipGetLastError());
}

template <typename R, int NR, bool CONJ>
void launch_fwd(hipStream_t s, const Near &nf, int32_t ntask, const void *x, int64_t ldx, void *out, int64_t ldo,
                int64_t nrhs, const double *alpha, const double *beta, int mode) {
    for (int64_t r0 = 0; r0 < nrhs; r0 += int64_t(NR) * 65535) {
        const int64_t nr = std::min<int64_t>(nrhs - r0, int64_t(NR) * 65535);
        dim3 grid(unsigned(ntask), unsigned((nr + NR - 1) / NR));
        hipLaunchKernelGGL((near_fwd_kernel<R, NR, CONJ>), grid, dim3(256), 0, s, nf.fwd_tasks.as<NearTask>(),
                           nf.data.as<C<R>>(), nf.colx.as<int32_t>(), nf.rowx.as<int32_t>(),
                           static_cast<const C<R> *>(x) + r0 * ldx, ldx, static_cast<C<R> *>(out) + r0 * ldo,
                           ldo, int(nr), C<R>{R(alpha[0]), R(alpha[1])}, C<R>{R(beta[0]), R(beta[1])}, mode);
        BFG_HIP(hipGetLastError());
    }
}

template <typename R, int NR, bool CONJ>
void launch_tr(hipStream_t s, const Near &nf, const void *x, int64_t ldx, void *part, int64_t ldo, int64_t nrhs) {
    for (int64_t r0 = 0; r0 < nrhs; r0 += int64_t(NR) * 65535) {
        const int64_t nr = std::min<int64_t>(nrhs - r0, int64_t(NR) * 65535);
        dim3 grid(unsigned(nf.n_tr), unsigned((nr + NR - 1) / NR));
        hipLaunchKernelGGL((near_tr_kernel<R, NR, CONJ>), grid, dim3(256), 0, s, nf.tr_tasks.as<NearTask>(),
                           nf.data.as<C<R>>(), nf.rowx.as<int32_t>(), static_cast<const C<R> *>(x) + r0 * ldx,
                           ldx, static_cast<C<R> *>(part) + r0 * ldo, ldo, int(nr));
        BFG_HIP(hipGetLastError());
    }
}

template <typename R, bool CONJ>
void fwd_nr(hipStream_t s, const Near &nf, const void *x, int64_t ldx, void *out, int64_t ldo, int64_t nrhs,
            const double *alpha, const double *beta, int mode) {
    if (nrhs == 1) launch_fwd<R, 1, CONJ>(s, nf, nf.n_fwd, x, ldx, out, ldo, nrhs, alpha, beta, mode);
    else if (nrhs == 2) launch_fwd<R, 2, CONJ>(s, nf, nf.n_fwd, x, ldx, out, ldo, nrhs, alpha, beta, mode);
    else if (nrhs <= 4) launch_fwd<R, 4, CONJ>(s, nf, nf.n_fwd, x, ldx, out, ldo, nrhs, alpha, beta, mode);
    else launch_fwd<R, 8, CONJ>(s, nf, nf.n_fwd, x, ldx, out, ldo, nrhs, alpha, beta, mode);
}

// y = alpha N x + beta y (conj(N) with CONJ), device x and y
template <int NT, bool CONJ>
void launch_mfma(hipStream_t s, const Near &nf, int32_t ns, void *y, int64_t ldy, int nc, const double *alpha,
                 const double *beta, bool bz) {
    hipLaunchKernelGGL((near_mfma_kernel<NT, CONJ>), dim3(unsigned(nf.n_mtasks)), dim3(256), 0, s,
                       nf.mtasks.as<NearTask>(), nf.n_mtasks, nf.data.as<C<float>>(), nf.ctree.as<int32_t>(),
                       nf.rowx.as<int32_t>(), nf.xg.as<C<float>>(), ns, static_cast<C<float> *>(y), ldy, nc,
                       C<float>{float(alpha[0]), float(alpha[1])}, C<float>{float(beta[0]), float(beta[1])},
                       bz ? 1 : 0);
    BFG_HIP(hipGetLastError());
}

// complex FP32 on the matrix cores, passes of at most 64 RHS
template <bool CONJ>
void apply_mfma(Near &nf, int64_t nrhs, const void *x, int64_t ldx, void *y, int64_t ldy, const double *alpha,
                const double *beta) {
    hipStream_t s = nf.ctx->stream;
    const bool bz = beta[0] == 0.0 && beta[1] == 0.0;
    const int32_t nsrc = nf.trial->n_values;
    nf.xg.reserve(size_t(std::max(nsrc, 1)) * 64 * sizeof(C<float>));
    for (int64_t r0 = 0; r0 < nrhs; r0 += 64) {
        const int nc = int(std::min<int64_t>(64, nrhs - r0));
        const int nt = (nc + 15) / 16;
        const int32_t ns = 16 * nt;
        const int64_t total = int64_t(nsrc) * ns;
        if (total > 0) {
            dim3 grid(unsigned(std::min<int64_t>((total + 255) / 256, int64_t(1) << 20)));
            hipLaunchKernelGGL(near_gather_kernel, grid, dim3(256), 0, s, static_cast<const C<float> *>(x) + r0 * ldx,
                               ldx, nf.perm_src.as<int32_t>(), nsrc, nf.xg.as<C<float>>(), ns, nc);
            BFG_HIP(hipGetLastError());
        }
        C<float> *yp = static_cast<C<float> *>(y) + r0 * ldy;
        if (nt == 1) launch_mfma<1, CONJ>(s, nf, ns, yp, ldy, nc, alpha, beta, bz);
        else if (nt == 2) launch_mfma<2, CONJ>(s, nf, ns, yp, ldy, nc, alpha, beta, bz);
        else if (nt == 3) launch_mfma<3, CONJ>(s, nf, ns, yp, ldy, nc, alpha, beta, bz);
        else launch_mfma<4, CONJ>(s, nf, ns, yp, ldy, nc, alpha, beta, bz);
    }
}

template <typename R, bool CONJ>
void apply_forward(Near &nf, int64_t nrhs, const void *x, int64_t ldx, void *y, int64_t ldy, const double *alpha,
                   const double *beta) {
    hipStream_t s = nf.ctx->stream;
    const bool bz = beta[0] == 0.0 && beta[1] == 0.0;
    if (sizeof(R) == 4 && nf.direct && nf.n_mtasks > 0 && nf.mfma_min_rhs > 0 && nrhs >= nf.mfma_min_rhs) {
        apply_mfma<CONJ>(nf, nrhs, x, ldx, y, ldy, alpha, beta);
        return;
    }
    if (nf.direct && nf.n_fwd > 0) {
        fwd_nr<R, CONJ>(s, nf, x, ldx, y, ldy, nrhs, alpha, beta, bz ? NEAR_Y_SET : NEAR_Y_ACC);
        return;
    }
    const int64_t ldp = std::max<int64_t>(nf.n_rflat, 1);
    nf.part.reserve(size_t(ldp) * size_t(nrhs) * sizeof(C<R>));
    if (nf.n_fwd > 0) {
        const double one[2] = {1, 0}, zero[2] = {0, 0};
        fwd_nr<R, CONJ>(s, nf, x, ldx, nf.part.p, ldp, nrhs, one, zero, NEAR_PART);
    }
    run_sum<R>(s, nf.rsum, nf.part.p, ldp, y, ldy, nrhs, alpha, beta);
}

void near_assemble(Near &nf, double k, double self_re, double self_im, int64_t n_pairs, const int32_t *pairs);

// the transposed copy, built on first use; null when it does not fit
Near *transposed_copy(Near &nf) {
    if (!nf.use_tcopy || nf.tcopy_failed || nf.kern) return nullptr;
    if (nf.tcopy) return nf.tcopy.get();
    size_t fr = 0, tot = 0;
    BFG_HIP(hipMemGetInfo(&fr, &tot));
    const size_t es = nf.precision == BFGPU_C128 ? 16 : 8;
    // the copy holds as many entries as the original; keep a tenth of the
    // device free for the caller
    if (size_t(nf.entries) * es + (size_t(64) << 20) > fr / 10 * 9) {
        nf.tcopy_failed = true;
        return nullptr;
    }
    auto t = std::make_unique<Near>();
    t->ctx = nf.ctx;
    t->test = nf.trial;
    t->trial = nf.test;
    t->precision = nf.precision;
    t->use_tcopy = false;
    t->mfma_min_rhs = nf.mfma_min_rhs;
    std::vector<int32_t> sw(nf.pairs.size());
    for (size_t p = 0; p + 1 < nf.pairs.size(); p += 2) {
        sw[p] = nf.pairs[p + 1];
        sw[p + 1] = nf.pairs[p];
    }
    try {
        near_assemble(*t, nf.k, nf.self_re, nf.self_im, int64_t(sw.size() / 2), sw.data());
    } catch (const Error &e) {
        // device memory ran short after all: keep the two-pass product
        if (e.code != BFGPU_ERR_NOMEM) throw;
        nf.tcopy_failed = true;
        return nullptr;
    }
    nf.tcopy = std::move(t);
    return nf.tcopy.get();
}

// y = alpha op(N) x + beta y, device x and y
template <typename R>
void near_apply(Near &nf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y, int64_t ldy,
                const double *alpha, const double *beta) {
    hipStream_t s = nf.ctx->stream;
    const size_t es = sizeof(C<R>);
    if (op == BFGPU_OP_N) {
        apply_forward<R, false>(nf, nrhs, x, ldx, y, ldy, alpha, beta);
        return;
    }
    if (Near *t = transposed_copy(nf)) {
        if (op == BFGPU_OP_C) apply_forward<R, true>(*t, nrhs, x, ldx, y, ldy, alpha, beta);
        else apply_forward<R, false>(*t, nrhs, x, ldx, y, ldy, alpha, beta);
        return;
    }
    const int64_t ldp = std::max<int64_t>(nf.n_cflat, 1);
    nf.part.reserve(size_t(ldp) * size_t(nrhs) * es);
    if (nf.n_tr > 0) {
        const bool cj = op == BFGPU_OP_C;
        if (nrhs == 1) {
            if (cj) launch_tr<R, 1, true>(s, nf, x, ldx, nf.part.p, ldp, nrhs);
            else launch_tr<R, 1, false>(s, nf, x, ldx, nf.part.p, ldp, nrhs);
        } else if (nrhs == 2) {
            if (cj) launch_tr<R, 2, true>(s, nf, x, ldx, nf.part.p, ldp, nrhs);
            else launch_tr<R, 2, false>(s, nf, x, ldx, nf.part.p, ldp, nrhs);
        } else {
            if (cj) launch_tr<R, 4, true>(s, nf, x, ldx, nf.part.p, ldp, nrhs);
            else launch_tr<R, 4, false>(s, nf, x, ldx, nf.part.p, ldp, nrhs);
        }
    }
    run_sum<R>(s, nf.csum, nf.part.p, ldp, y, ldy, nrhs, alpha, beta);
}

void near_matvec(Near &nf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y, int64_t ldy,
                 const double *alpha, const double *beta, int flags) {
    BFG_REQUIRE(op == BFGPU_OP_N || op == BFGPU_OP_T || op == BFGPU_OP_C, "near_matvec: bad op");
    BFG_REQUIRE(nrhs >= 0, "near_matvec: nrhs < 0");
    if (nrhs == 0) return;
    BFG_REQUIRE(x && y, "near_matvec: null x or y");
    const bool fwd = op == BFGPU_OP_N;
    const int64_t in_len = fwd ? nf.cols : nf.rows, out_len = fwd ? nf.rows : nf.cols;
    BFG_REQUIRE(ldx >= in_len && ldy >= out_len, "near_matvec: leading dimension too small (DimensionMismatch)");
    BFG_HIP(hipSetDevice(nf.ctx->device));
    auto apply = [&](const void *dx, void *dy) {
        if (nf.precision == BFGPU_C128) near_apply<double>(nf, op, nrhs, dx, ldx, dy, ldy, alpha, beta);
        else near_apply<float>(nf, op, nrhs, dx, ldx, dy, ldy, alpha, beta);
    };
    if (flags & BFGPU_XY_DEVICE) {
        apply(x, y);
        return;
    }
    const size_t es = nf.precision == BFGPU_C128 ? 16 : 8;
    const size_t xb = size_t(ldx) * size_t(nrhs - 1) * es + size_t(in_len) * es;
    const size_t yb = size_t(ldy) * size_t(nrhs - 1) * es + size_t(out_len) * es;
    nf.stage.reserve(xb + yb);
    char *dx = nf.stage.as<char>(), *dy = dx + xb;
    hipStream_t s = nf.ctx->stream;
    BFG_HIP(hipMemcpyAsync(dx, x, xb, hipMemcpyHostToDevice, s));
    BFG_HIP(hipMemcpyAsync(dy, y, yb, hipMemcpyHostToDevice, s));
    apply(dx, dy);
    BFG_HIP(hipMemcpyAsync(y, dy, yb, hipMemcpyDeviceToHost, s));
    BFG_HIP(hipStreamSynchronize(s));
}

void near_assemble(Near &nf, double k, double self_re, double self_im, int64_t n_pairs, const int32_t *pairs) {
    const Tree &te = *nf.test, &tr = *nf.trial;
    hipStream_t s = nf.ctx->stream;
    BFG_HIP(hipSetDevice(nf.ctx->device));
    nf.rows = te.n_values;
    nf.cols = tr.n_values;
    nf.n_pairs = n_pairs;
    nf.k = k;
    nf.self_re = self_re;
    nf.self_im = self_im;
    for (int64_t p = 0; p < n_pairs; ++p)
        BFG_REQUIRE(pairs[2 * p] >= 0 && pairs[2 * p] < te.n_nodes && pairs[2 * p + 1] >= 0 &&
                        pairs[2 * p + 1] < tr.n_nodes,
                    "near_assemble: node id out of range (pair " + std::to_string(p) + ")");
    // items: observer nodes in increasing id order; each lists its pairs in
    // input order
    std::vector<int64_t> order(static_cast<size_t>(n_pairs));
    for (int64_t p = 0; p < n_pairs; ++p) order[size_t(p)] = p;
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return pairs[2 * a] < pairs[2 * b]; });
    std::vector<NearTask> fwd, trt;
    std::vector<FillTask> fill;
    std::vector<int64_t> fill_end;
    std::vector<int32_t> rowx, colx, ctree;
    std::vector<NearTask> mt;
    nf.pair_fofs.assign(size_t(n_pairs), 0);
    nf.pair_m.assign(size_t(n_pairs), 0);
    nf.pair_n.assign(size_t(n_pairs), 0);
    int64_t fofs = 0;
    for (size_t a = 0; a < order.size();) {
        const int32_t O = pairs[2 * order[a]];
        size_t b = a;
        while (b < order.size() && pairs[2 * order[b]] == O) ++b;
        const int32_t m = te.nvals(O);
        const int64_t col0 = int64_t(colx.size()), rflat = int64_t(rowx.size());
        for (size_t q = a; q < b; ++q) {
            const int64_t p = order[q];
            const int32_t S = pairs[2 * p + 1];
            nf.pair_fofs[size_t(p)] = fofs + (int64_t(colx.size()) - col0) * m;
            nf.pair_m[size_t(p)] = m;
            nf.pair_n[size_t(p)] = tr.nvals(S);
            colx.insert(colx.end(), tr.vals(S), tr.vals(S) + tr.nvals(S));
            for (int32_t j = tr.val_begin[S]; j < tr.val_end[S]; ++j) ctree.push_back(j);
        }
        const int64_t ncol = int64_t(colx.size()) - col0;
        BFG_REQUIRE(ncol < (int64_t(1) << 31) && int64_t(colx.size()) < (int64_t(1) << 31),
                    "near_assemble: too many near columns for 32-bit indices");
        if (m == 0 || ncol == 0) {
            // an empty block contributes nothing: no flat rows or columns
            // (the second pass would read partials no task writes)
            colx.resize(size_t(col0));
            ctree.resize(size_t(col0));
        } else {
            rowx.insert(rowx.end(), te.vals(O), te.vals(O) + m);
            BFG_REQUIRE(int64_t(rowx.size()) < (int64_t(1) << 31), "near_assemble: too many near rows");
            for (int32_t r0 = 0; r0 < m; r0 += 64)
                fwd.push_back(NearTask{fofs + r0, m, std::min(64, m - r0), int32_t(ncol), int32_t(col0),
                                       int32_t(rflat + r0), 0});
            for (int32_t r0 = 0; r0 < m; r0 += 16)
                mt.push_back(NearTask{fofs + r0, m, std::min(16, m - r0), int32_t(ncol), int32_t(col0),
                                      int32_t(rflat + r0), 0});
            trt.push_back(NearTask{fofs, m, m, int32_t(ncol), int32_t(col0), int32_t(rflat), 0});
            fill.push_back(FillTask{fofs, m, int32_t(ncol), int32_t(rflat), int32_t(col0)});
            fofs += int64_t(m) * ncol;
            fill_end.push_back(fofs);
        }
        ++nf.n_items;
        a = b;
    }
    BFG_REQUIRE(mt.size() < (size_t(1) << 31) && trt.size() < (size_t(1) << 31), "near_assemble: too many items");
    if (nf.use_tcopy) nf.pairs.assign(pairs, pairs + 2 * n_pairs);
    nf.entries = fofs;
    nf.n_rflat = int64_t(rowx.size());
    nf.n_cflat = int64_t(colx.size());
    nf.n_fwd = int32_t(fwd.size());
    nf.n_tr = int32_t(trt.size());
    nf.n_mtasks = int32_t(mt.size());
    // forward writes y directly when every row of y is produced by exactly
    // one item row
    {
        std::vector<int32_t> cover(size_t(te.n_values), 0);
        for (int32_t v : rowx) ++cover[size_t(v)];
        nf.direct = std::all_of(cover.begin(), cover.end(), [](int32_t c) { return c == 1; });
    }
    const size_t es = nf.precision == BFGPU_C128 ? 16 : 8;
    nf.data.alloc(size_t(std::max<int64_t>(fofs, 1)) * es);
    put(nf.fwd_tasks, fwd, s);
    put(nf.tr_tasks, trt, s);
    put(nf.rowx, rowx, s);
    put(nf.colx, colx, s);
    put(nf.mtasks, mt, s);
    put(nf.ctree, ctree, s);
    put(nf.perm_src, tr.perm, s);
    build_csr(nf.rsum, rowx, te.n_values, s);
    build_csr(nf.csum, colx, tr.n_values, s);
    if (!fill.empty()) {
        DevBuf d_obs, d_src, d_fill, d_end;
        put(d_obs, te.points, s);
        put(d_src, tr.points, s);
        put(d_fill, fill, s);
        put(d_end, fill_end, s);
        if (nf.kern)
            launch_fill_rwg_near(s, nf.precision, *nf.kern, nf.rowx.as<int32_t>(), nf.colx.as<int32_t>(),
                                 d_fill.as<FillTask>(), d_end.as<int64_t>(), int32_t(fill.size()), fofs, nf.data.p, k);
        else
            launch_fill(s, nf.precision, d_obs.as<double>(), d_src.as<double>(), nf.rowx.as<int32_t>(),
                        nf.colx.as<int32_t>(), d_fill.as<FillTask>(), d_end.as<int64_t>(), int32_t(fill.size()),
                        fofs, nf.data.p, k, self_re, self_im);
        BFG_HIP(hipStreamSynchronize(s));
    } else {
        BFG_HIP(hipStreamSynchronize(s));
    }
}

}  // namespace
}  // namespace bfg

using namespace bfg;

struct bfgpu_near { bfg::Near n; };

#define NF_API_BEGIN try {
#define NF_API_END                        \
    }                                     \
    catch (const Error &e) {              \
        set_last_error(e.what());         \
        return e.code;                    \
    }                                     \
    catch (const std::bad_alloc &) {      \
        set_last_error("out of memory");  \
        return BFGPU_ERR_NOMEM;           \
    }                                     \
    catch (const std::exception &e) {     \
        set_last_error(e.what());         \
        return BFGPU_ERR_STATE;           \
    }                                     \
    return BFGPU_OK;

extern "C" int bfgpu_near_assemble_kernel(bfgpu_ctx *ctx, const bfgpu_tree *test, const bfgpu_tree *trial,
                                          const bfgpu_kernel *kern, double k, int32_t precision, int64_t n_pairs,
                                          const int32_t *pairs, bfgpu_near **out) {
    NF_API_BEGIN
    BFG_REQUIRE(ctx && test && trial && kern && out, "near_assemble_kernel: null argument");
    BFG_REQUIRE(n_pairs >= 0 && (n_pairs == 0 || pairs), "near_assemble_kernel: bad pair list");
    BFG_REQUIRE(precision == BFGPU_C64 || precision == BFGPU_C128, "near_assemble_kernel: precision");
    const KernelRWG *kr = kernel_of(kern);
    BFG_REQUIRE(kernel_functions(*kr, true) >= test->t.n_values && kernel_functions(*kr, false) >= trial->t.n_values,
                "near_assemble_kernel: the plug-in has fewer functions than the trees have values");
    BFG_REQUIRE(kernel_has_near(*kr), "near_assemble_kernel: call bfgpu_kernel_rwg_set_near first");
    auto h = std::make_unique<bfgpu_near>();
    h->n.ctx = &ctx->c;
    h->n.test = &test->t;
    h->n.trial = &trial->t;
    h->n.precision = precision;
    h->n.kern = kr;
    h->n.use_tcopy = false;
    near_assemble(h->n, k, 0.0, 0.0, n_pairs, pairs);
    *out = h.release();
    NF_API_END
}

extern "C" int bfgpu_near_assemble(bfgpu_ctx *ctx, const bfgpu_tree *test, const bfgpu_tree *trial, double k,
                                   int32_t precision, const double *self_value, int64_t n_pairs,
                                   const int32_t *pairs, bfgpu_near **out) {
    NF_API_BEGIN
    BFG_REQUIRE(ctx && test && trial && out, "near_assemble: null argument");
    BFG_REQUIRE(n_pairs >= 0 && (n_pairs == 0 || pairs), "near_assemble: bad pair list");
    BFG_REQUIRE(precision == BFGPU_C64 || precision == BFGPU_C128, "near_assemble: precision");
    BFG_REQUIRE(test->t.points.size() == size_t(3) * size_t(test->t.n_values) &&
                    trial->t.points.size() == size_t(3) * size_t(trial->t.n_values),
                "near_assemble: trees carry no points");
    auto h = std::make_unique<bfgpu_near>();
    h->n.ctx = &ctx->c;
    h->n.test = &test->t;
    h->n.trial = &trial->t;
    h->n.precision = precision;
    near_assemble(h->n, k, self_value ? self_value[0] : 0.0, self_value ? self_value[1] : 0.0, n_pairs, pairs);
    *out = h.release();
    NF_API_END
}

extern "C" int bfgpu_near_matvec(bfgpu_near *h, int32_t op, int64_t nrhs, const void *x, int64_t ldx, void *y,
                                 int64_t ldy, const double *alpha, const double *beta, int32_t flags) {
    NF_API_BEGIN
    BFG_REQUIRE(h, "near_matvec: null handle");
    static const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    near_matvec(h->n, op, nrhs, x, ldx, y, ldy, alpha ? alpha : one, beta ? beta : zero, flags);
    NF_API_END
}

extern "C" int bfgpu_near_info(const bfgpu_near *h, int64_t *info) {
    NF_API_BEGIN
    BFG_REQUIRE(h && info, "near_info: null argument");
    const Near &n = h->n;
    info[0] = n.rows;
    info[1] = n.cols;
    info[2] = n.n_pairs;
    info[3] = n.n_items;
    info[4] = n.entries;
    auto bytes = [](const Near &q) {
        return int64_t(q.data.bytes + q.fwd_tasks.bytes + q.tr_tasks.bytes + q.colx.bytes + q.rowx.bytes +
                       q.rsum.ptr.bytes + q.rsum.idx.bytes + q.csum.ptr.bytes + q.csum.idx.bytes + q.part.bytes +
                       q.mtasks.bytes + q.ctree.bytes + q.perm_src.bytes + q.xg.bytes);
    };
    info[5] = bytes(n) + (n.tcopy ? bytes(*n.tcopy) : 0);
    info[6] = n.precision;
    info[7] = n.direct ? 1 : 0;
    NF_API_END
}

extern "C" int bfgpu_near_block(const bfgpu_near *h, int64_t pair, int32_t *mn, void *data) {
    NF_API_BEGIN
    BFG_REQUIRE(h && mn, "near_block: null argument");
    const Near &n = h->n;
    BFG_REQUIRE(pair >= 0 && pair < n.n_pairs, "near_block: pair out of range");
    mn[0] = n.pair_m[size_t(pair)];
    mn[1] = n.pair_n[size_t(pair)];
    if (!data) return BFGPU_OK;
    const size_t es = n.precision == BFGPU_C128 ? 16 : 8;
    const size_t cnt = size_t(mn[0]) * size_t(mn[1]);
    if (cnt == 0) return BFGPU_OK;
    BFG_HIP(hipSetDevice(n.ctx->device));
    BFG_HIP(hipStreamSynchronize(n.ctx->stream));
    BFG_HIP(hipMemcpy(data, static_cast<const char *>(n.data.p) + size_t(n.pair_fofs[size_t(pair)]) * es, cnt * es,
                      hipMemcpyDeviceToHost));
    NF_API_END
}

extern "C" int bfgpu_near_set_transposed_copy(bfgpu_near *h, int32_t on) {
    NF_API_BEGIN
    BFG_REQUIRE(h, "near_set_transposed_copy: null handle");
    BFG_HIP(hipSetDevice(h->n.ctx->device));
    BFG_HIP(hipStreamSynchronize(h->n.ctx->stream));
    h->n.use_tcopy = on != 0;
    h->n.tcopy_failed = false;
    if (!on) h->n.tcopy.reset();
    NF_API_END
}

extern "C" int bfgpu_near_set_mfma(bfgpu_near *h, int32_t min_rhs) {
    NF_API_BEGIN
    BFG_REQUIRE(h && min_rhs >= 0, "near_set_mfma: null handle or negative count");
    h->n.mfma_min_rhs = min_rhs;
    if (h->n.tcopy) h->n.tcopy->mfma_min_rhs = min_rhs;
    NF_API_END
}

extern "C" int bfgpu_near_destroy(bfgpu_near *h) {
    NF_API_BEGIN
    if (h) {
        BFG_HIP(hipSetDevice(h->n.ctx->device));
        BFG_HIP(hipStreamSynchronize(h->n.ctx->stream));
    }
    delete h;
    NF_API_END
}
