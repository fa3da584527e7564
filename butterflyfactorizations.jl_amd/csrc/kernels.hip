// kernels.hip -- gfx950 kernels of the butterfly far-field path:
//   * Helmholtz kernel-matrix tiles over index lists,
//   * batched column-pivoted Householder QR with relative-tolerance stop,
//   * interpolation matrices T = R11^{-1} [R11 R12] P^T (FP64 solve),
//   * level stages of the butterfly matvec (forward and transposed).
// All kernels use 256-thread workgroups = 4 wavefronts of 64 lanes.
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "device.hpp"
#include "internal.hpp"

namespace bfg {


// ---------------------------------------------------------------------------
// Helmholtz tiles.  Z[i + m j] = exp(-i k r) / (4 pi r), r = |obs[rows[i]] -
// src[cols[j]]|, for every task; one thread per entry, task found by binary
// search over the running entry counts.  The distance and the phase
// reduction k r mod 2 pi are computed in FP64; FP32 tiles take sincos of the
// reduced phase in FP32.  Coincident points (r = 0, only in near-field
// blocks) take the given self value.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void fill_kernel(const double *__restrict__ obs,
                                                   const double *__restrict__ src,
                                                   const int32_t *__restrict__ rows,
                                                   const int32_t *__restrict__ cols,
                                                   const FillTask *__restrict__ tasks,
                                                   const int64_t *__restrict__ task_end,
                                                   int32_t n_tasks, int64_t total, C<R> *Z,
                                                   double k, double self_re, double self_im) {
    // grid-stride: a dispatch holds fewer than 2^32 work-items, the tiles of
    // one level can hold more entries
    int32_t lo = 0;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += int64_t(gridDim.x) * 256) {
    int32_t hi = n_tasks - 1;  // entries only grow: search from the last task
    while (lo < hi) {
        int32_t mid = (lo + hi) >> 1;
        if (task_end[mid] > e) hi = mid;
        else lo = mid + 1;
    }
    const FillTask t = tasks[lo];
    int64_t local = e - (task_end[lo] - int64_t(t.m) * t.n);
    int32_t i = int32_t(local % t.m), j = int32_t(local / t.m);
    const double *a = obs + 3 * int64_t(rows[t.rofs + i]);
    const double *b = src + 3 * int64_t(cols[t.cofs + j]);
    double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    double g = 1.0 / (4.0 * M_PI * r);
    C<R> v;
    if constexpr (sizeof(R) == 8) {
        double s, c;
        sincos(k * r, &s, &c);
        v = {c * g, -s * g};
    } else {
        double ph = k * r;
        ph -= 6.283185307179586 * rint(ph * 0.15915494309189535);
        float s, c;
        sincosf(float(ph), &s, &c);
        v = {float(c * g), float(-s * g)};
    }
    if (r == 0.0) v = {R(self_re), R(self_im)};  // coincident points
    Z[t.zofs + local] = v;
    }
}

void launch_fill(hipStream_t s, int prec, const double *obs_pts, const double *src_pts,
                 const int32_t *rows, const int32_t *cols, const FillTask *tasks,
                 const int64_t *task_end, int32_t n_tasks, int64_t total, void *Z, double k,
                 double self_re, double self_im) {
    if (total <= 0 || n_tasks <= 0) return;
    dim3 grid(unsigned(std::min<int64_t>((total + 255) / 256, int64_t(1) << 22)));
    if (prec == BFGPU_C128)
        hipLaunchKernelGGL(fill_kernel<double>, grid, dim3(256), 0, s, obs_pts, src_pts, rows,
                           cols, tasks, task_end, n_tasks, total, (C<double> *)Z, k, self_re, self_im);
    else
        hipLaunchKernelGGL(fill_kernel<float>, grid, dim3(256), 0, s, obs_pts, src_pts, rows,
                           cols, tasks, task_end, n_tasks, total, (C<float> *)Z, k, self_re, self_im);
    BFG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Batched column-pivoted Householder QR, one workgroup per block.
// Same algorithm as LAPACK xGEQP3/xLAQP2 (unblocked): pivot = first column
// of largest partial norm; reflector in xLARFG form; trailing update with
// H^H; partial norms (FP64) downdated with the LAPACK Working Note 176 test
// and recomputed when the downdate has cancelled.  Stops before step k >= 1
// when the pivot column's remaining norm is <= rtol * |R_11|.
// LDS: vn1[n], vn2[n] (FP64).  Z is updated in place in global memory.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void cpqr_kernel(C<R> *Z, const QrTask *__restrict__ tasks,
                                                   int32_t *piv_all, int32_t *rank_out,
                                                   double rtol, double *tau_out) {
    extern __shared__ double sm[];
    const QrTask tk = tasks[blockIdx.x];
    const int m = tk.m, n = tk.n;
    C<R> *A = Z + tk.zofs;
    int32_t *piv = piv_all + tk.pofs;
    double *vn1 = sm, *vn2 = sm + n;
    __shared__ double rv[4];
    __shared__ int ri[4];
    __shared__ double s_pn;
    __shared__ double s_beta, s_tau_x, s_tau_y, s_scal_x, s_scal_y;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    for (int j = wave; j < n; j += 4) {
        double s = 0.0;
        for (int i = lane; i < m; i += 64) {
            C<R> a = A[i + int64_t(j) * m];
            s += double(a.x) * a.x + double(a.y) * a.y;
        }
        s = wave_sum(s);
        if (lane == 0) { vn1[j] = sqrt(s); vn2[j] = vn1[j]; }
    }
    for (int j = tid; j < n; j += 256) piv[j] = j;
    __syncthreads();

    const int mn = m < n ? m : n;
    int rank = mn;
    double r11 = 0.0;
    // recompute threshold sqrt(eps) of the precision the columns are stored
    // in (LAPACK xLAMCH('E') = 2^-53 / 2^-24): FP32 columns lose accuracy in
    // the downdate long before FP64 ones do
    const double tol3z = sizeof(R) == 8 ? 1.0536712127723509e-08 : 2.44140625e-04;
    for (int k = 0; k < mn; ++k) {
        // pivot: first maximal vn1 over [k, n)
        double bv = -1.0;
        int bi = n;
        for (int j = k + tid; j < n; j += 256) {
            double v = vn1[j];
            if (v > bv || (v == bv && j < bi)) { bv = v; bi = j; }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            double ov = __shfl_xor(bv, o, 64);
            int oi = __shfl_xor(bi, o, 64);
            if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
        }
        if (lane == 0) { rv[wave] = bv; ri[wave] = bi; }
        __syncthreads();
        int p;
        {
            double b = rv[0];
            p = ri[0];
            for (int w = 1; w < 4; ++w)
                if (rv[w] > b || (rv[w] == b && ri[w] < p)) { b = rv[w]; p = ri[w]; }
        }
        // no pivot among the columns (only when norms are NaN): stop here
        if (p >= n) { rank = k; break; }
        // exact remaining norm of the pivot column
        if (wave == 0) {
            double s = 0.0;
            for (int i = k + lane; i < m; i += 64) {
                C<R> a = A[i + int64_t(p) * m];
                s += double(a.x) * a.x + double(a.y) * a.y;
            }
            s = wave_sum(s);
            if (lane == 0) s_pn = sqrt(s);
        }
        __syncthreads();
        double pn = s_pn;
        if (k == 0) r11 = pn;
        if (!(pn > 0.0) || (k > 0 && pn <= rtol * r11)) { rank = k; break; }
        if (p != k) {
            for (int i = tid; i < m; i += 256) {
                C<R> t = A[i + int64_t(p) * m];
                A[i + int64_t(p) * m] = A[i + int64_t(k) * m];
                A[i + int64_t(k) * m] = t;
            }
            if (tid == 0) {
                int32_t t = piv[p]; piv[p] = piv[k]; piv[k] = t;
                vn1[p] = vn1[k];
                vn2[p] = vn2[k];
            }
        }
        __syncthreads();
        // reflector from column k
        C<R> *a = A + int64_t(k) * m;
        if (wave == 0) {
            double s = 0.0;
            for (int i = k + 1 + lane; i < m; i += 64) s += double(a[i].x) * a[i].x + double(a[i].y) * a[i].y;
            s = wave_sum(s);
            if (lane == 0) {
                double xnorm = sqrt(s);
                double ar = a[k].x, ai = a[k].y;
                if (xnorm == 0.0 && ai == 0.0) {
                    s_beta = ar; s_tau_x = 0.0; s_tau_y = 0.0; s_scal_x = 1.0; s_scal_y = 0.0;
                } else {
                    double beta = hypot(hypot(ar, ai), xnorm);
                    if (ar >= 0.0) beta = -beta;
                    s_beta = beta;
                    s_tau_x = (beta - ar) / beta;
                    s_tau_y = -ai / beta;
                    // 1 / (alpha - beta)
                    double dr = ar - beta, di = ai, den = dr * dr + di * di;
                    s_scal_x = dr / den;
                    s_scal_y = -di / den;
                }
            }
        }
        __syncthreads();
        if (tau_out && tid == 0) {
            tau_out[2 * (int64_t(tk.pofs) + k)] = s_tau_x;
            tau_out[2 * (int64_t(tk.pofs) + k) + 1] = s_tau_y;
        }
        const C<R> scal = {R(s_scal_x), R(s_scal_y)};
        const C<R> ctau = {R(s_tau_x), R(-s_tau_y)};  // conj(tau)
        const bool apply = !(s_tau_x == 0.0 && s_tau_y == 0.0);
        if (apply)
            for (int i = k + 1 + tid; i < m; i += 256) a[i] = cmul(a[i], scal);
        __syncthreads();
        if (tid == 0) a[k] = C<R>{R(s_beta), R(0)};
        // trailing update, one wave per column
        for (int j = k + 1 + wave; j < n; j += 4) {
            C<R> *c = A + int64_t(j) * m;
            if (apply) {
                C<R> w = {R(0), R(0)};
                for (int i = k + lane; i < m; i += 64) {
                    C<R> v = (i == k) ? C<R>{R(1), R(0)} : a[i];
                    cfmac(w, v, c[i]);
                }
                w.x = wave_sum(w.x);
                w.y = wave_sum(w.y);
                w = cmul(ctau, w);
                for (int i = k + lane; i < m; i += 64) {
                    C<R> v = (i == k) ? C<R>{R(1), R(0)} : a[i];
                    c[i] = csub(c[i], cmul(v, w));
                }
            }
            // norm downdate (lane 0 holds row k)
            int recompute = 0;
            double nv = 0.0;
            if (lane == 0 && vn1[j] != 0.0) {
                C<R> ck = c[k];
                double t = sqrt(double(ck.x) * ck.x + double(ck.y) * ck.y) / vn1[j];
                t = 1.0 - t * t;
                if (t < 0.0) t = 0.0;
                double r = vn1[j] / vn2[j];
                double t2 = t * r * r;
                if (t2 <= tol3z) recompute = 1;
                else nv = vn1[j] * sqrt(t);
            }
            recompute = __shfl(recompute, 0, 64);
            if (recompute) {
                double s = 0.0;
                // same lane -> row map as the update, so every lane reads its own writes
                for (int i = k + lane; i < m; i += 64)
                    if (i > k) s += double(c[i].x) * c[i].x + double(c[i].y) * c[i].y;
                s = wave_sum(s);
                if (lane == 0) { vn1[j] = sqrt(s); vn2[j] = vn1[j]; }
            } else if (lane == 0 && vn1[j] != 0.0) {
                vn1[j] = nv;
            }
        }
        __syncthreads();
    }
    if (tid == 0) rank_out[blockIdx.x] = rank;
}

// ---------------------------------------------------------------------------
// The same pivoted QR for blocks of at most 64 columns that fit in LDS: one
// wavefront per block, `wpb` blocks per workgroup, the block copied into the
// wavefront's LDS region and factored there, then written back.  Lane j owns
// column j (partial norms in FP64 registers, trailing update and downdate of
// its column); row sums are wave reductions.  No workgroup barriers.  The
// steps, the pivot rule (first maximal partial norm), the reflector, the
// downdate test and the stopping rule are those of cpqr_kernel.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void cpqr_wave_kernel(C<R> *Z, const QrTask *__restrict__ tasks, int32_t n_tasks,
                                                        int32_t *piv_all, int32_t *rank_out, double rtol,
                                                        int32_t region, double *tau_out) {
    extern __shared__ char smraw[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t task = int64_t(blockIdx.x) * (blockDim.x >> 6) + wave;
    if (task >= n_tasks) return;  // whole wavefront; no barriers below
    const QrTask tk = tasks[task];
    const int m = tk.m, n = tk.n;
    C<R> *G = Z + tk.zofs;
    // LDS copy with an odd column stride (lanes walking a row of the block
    // then spread over the banks instead of hitting one)
    const int ld = m | 1;
    C<R> *A = reinterpret_cast<C<R> *>(smraw) + int64_t(wave) * region;
    for (int e = lane; e < m * n; e += 64) {
        const int j = e / m;
        A[j * ld + (e - j * m)] = G[e];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const bool colok = lane < n;
    C<R> *cj = A + int64_t(colok ? lane : 0) * ld;  // this lane's column
    double vn1 = 0.0, vn2 = 0.0;
    if (colok) {
        double s = 0.0;
        for (int i = 0; i < m; ++i) s += double(cj[i].x) * cj[i].x + double(cj[i].y) * cj[i].y;
        vn1 = vn2 = sqrt(s);
    }
    int pv = lane;
    const int mn = m < n ? m : n;
    int rank = mn;
    double r11 = 0.0;
    const double tol3z = sizeof(R) == 8 ? 1.0536712127723509e-08 : 2.44140625e-04;
    auto sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    for (int k = 0; k < mn; ++k) {
        // pivot: first maximal partial norm over columns [k, n)
        double bv = (colok && lane >= k) ? vn1 : -1.0;
        int bi = lane;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double ov = __shfl_xor(bv, o, 64);
            const int oi = __shfl_xor(bi, o, 64);
            if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
        }
        const int p = __shfl(bi, 0, 64);
        // no pivot among the columns (only when norms are NaN): stop here
        if (p >= n || !(__shfl(bv, 0, 64) >= 0.0)) { rank = k; break; }
        // exact remaining norm of the pivot column
        const C<R> *ap = A + int64_t(p) * ld;
        double s = 0.0;
        for (int i = k + lane; i < m; i += 64) s += double(ap[i].x) * ap[i].x + double(ap[i].y) * ap[i].y;
        const double pn = sqrt(wave_sum(s));
        if (k == 0) r11 = pn;
        if (!(pn > 0.0) || (k > 0 && pn <= rtol * r11)) { rank = k; break; }
        if (p != k) {
            sync();  // the norm's reads of column p precede the swap's writes
            C<R> *ak = A + int64_t(k) * ld;
            C<R> *aq = A + int64_t(p) * ld;
            for (int i = lane; i < m; i += 64) {
                const C<R> t = aq[i];
                aq[i] = ak[i];
                ak[i] = t;
            }
            const int pk = __shfl(pv, k, 64), pp = __shfl(pv, p, 64);
            const double v1k = __shfl(vn1, k, 64), v1p = __shfl(vn1, p, 64);
            const double v2k = __shfl(vn2, k, 64), v2p = __shfl(vn2, p, 64);
            if (lane == p) { pv = pk; vn1 = v1k; vn2 = v2k; }
            if (lane == k) { pv = pp; vn1 = v1p; vn2 = v2p; }
            sync();
        }
        // reflector from column k (scalars in FP64, every lane the same)
        C<R> *a = A + int64_t(k) * ld;
        double xs = 0.0;
        for (int i = k + 1 + lane; i < m; i += 64) xs += double(a[i].x) * a[i].x + double(a[i].y) * a[i].y;
        const double xnorm = sqrt(wave_sum(xs));
        const double ar = a[k].x, ai = a[k].y;
        double beta, tx, ty, sx, sy;
        if (xnorm == 0.0 && ai == 0.0) {
            beta = ar; tx = 0.0; ty = 0.0; sx = 1.0; sy = 0.0;
        } else {
            beta = hypot(hypot(ar, ai), xnorm);
            if (ar >= 0.0) beta = -beta;
            tx = (beta - ar) / beta;
            ty = -ai / beta;
            const double dr = ar - beta, di = ai, den = dr * dr + di * di;
            sx = dr / den;
            sy = -di / den;
        }
        const bool apply = !(tx == 0.0 && ty == 0.0);
        if (tau_out && lane == 0) {
            tau_out[2 * (int64_t(tk.pofs) + k)] = tx;
            tau_out[2 * (int64_t(tk.pofs) + k) + 1] = ty;
        }
        const C<R> scal = {R(sx), R(sy)};
        const C<R> ctau = {R(tx), R(-ty)};  // conj(tau)
        sync();  // every lane has read a[k] before it is overwritten
        if (apply)
            for (int i = k + 1 + lane; i < m; i += 64) a[i] = cmul(a[i], scal);
        if (lane == 0) a[k] = C<R>{R(beta), R(0)};
        sync();
        // trailing update and norm downdate of this lane's column
        if (colok && lane > k) {
            if (apply) {
                C<R> w = cj[k];  // v_k = 1
                for (int i = k + 1; i < m; ++i) cfmac(w, a[i], cj[i]);
                w = cmul(ctau, w);
                cj[k] = csub(cj[k], w);
                for (int i = k + 1; i < m; ++i) cj[i] = csub(cj[i], cmul(a[i], w));
            }
            if (vn1 != 0.0) {
                const C<R> ck = cj[k];
                double t = sqrt(double(ck.x) * ck.x + double(ck.y) * ck.y) / vn1;
                t = 1.0 - t * t;
                if (t < 0.0) t = 0.0;
                const double rr = vn1 / vn2;
                if (t * rr * rr <= tol3z) {
                    double s2 = 0.0;
                    for (int i = k + 1; i < m; ++i) s2 += double(cj[i].x) * cj[i].x + double(cj[i].y) * cj[i].y;
                    vn1 = vn2 = sqrt(s2);
                } else {
                    vn1 = vn1 * sqrt(t);
                }
            }
        }
        sync();
    }
    for (int e = lane; e < m * n; e += 64) {
        const int j = e / m;
        G[e] = A[j * ld + (e - j * m)];
    }
    if (colok) piv_all[tk.pofs + lane] = pv;
    if (lane == 0) rank_out[task] = rank;
}

bool cpqr_wave_ok(int prec, int32_t m, int32_t n) {
    return n <= 64 && m <= (prec == BFGPU_C128 ? 32 : 64);
}

void launch_cpqr(hipStream_t s, int prec, void *Z, const QrTask *tasks, int32_t n_tasks,
                 int32_t *piv, int32_t *rank, double rtol, int32_t max_n, int32_t max_m, double *tau) {
    if (n_tasks <= 0) return;
    const size_t es = prec == BFGPU_C128 ? 16 : 8;
    const size_t region = size_t(std::max(max_m, 1) | 1) * size_t(std::max(max_n, 1));  // entries per block
    const size_t per_wave = region * es;  // at most 65 x 64 x 8 or 33 x 64 x 16 bytes
    if (cpqr_wave_ok(prec, max_m, max_n)) {
        // blocks per workgroup: as many as 64 KiB of LDS hold, at most 4
        const int wpb = int(std::max<size_t>(1, std::min<size_t>(4, (64 * 1024) / per_wave)));
        const size_t lds = per_wave * size_t(wpb);
        // chunks of tasks: a dispatch holds fewer than 2^32 work-items
        for (int32_t t0 = 0; t0 < n_tasks; t0 += TASK_CHUNK) {
            const int32_t nt = std::min(TASK_CHUNK, n_tasks - t0);
            dim3 grid(unsigned((int64_t(nt) + wpb - 1) / wpb));
            if (prec == BFGPU_C128)
                hipLaunchKernelGGL(cpqr_wave_kernel<double>, grid, dim3(64 * wpb), lds, s, (C<double> *)Z, tasks + t0,
                                   nt, piv, rank + t0, rtol, int32_t(region), tau);
            else
                hipLaunchKernelGGL(cpqr_wave_kernel<float>, grid, dim3(64 * wpb), lds, s, (C<float> *)Z, tasks + t0,
                                   nt, piv, rank + t0, rtol, int32_t(region), tau);
            BFG_HIP(hipGetLastError());
        }
        return;
    }
    size_t lds = sizeof(double) * 2 * size_t(max_n > 0 ? max_n : 1);
    BFG_REQUIRE(lds <= 64 * 1024, "cpqr: too many columns for the norm arrays");
    for (int32_t t0 = 0; t0 < n_tasks; t0 += TASK_CHUNK) {
        const int32_t nt = std::min(TASK_CHUNK, n_tasks - t0);
        if (prec == BFGPU_C128)
            hipLaunchKernelGGL(cpqr_kernel<double>, dim3(nt), dim3(256), lds, s,
                               (C<double> *)Z, tasks + t0, piv, rank + t0, rtol, tau);
        else
            hipLaunchKernelGGL(cpqr_kernel<float>, dim3(nt), dim3(256), lds, s,
                               (C<float> *)Z, tasks + t0, piv, rank + t0, rtol, tau);
        BFG_HIP(hipGetLastError());
    }
}

// ---------------------------------------------------------------------------
// Interpolation matrices.  For each task: T[:, piv[j]] = e_j (j < rank) and
// T[:, piv[j]] = R11^{-1} R[0:rank, j] (j >= rank), solved in FP64 by one
// wave per column (lanes hold the right-hand side rows, column-oriented back
// substitution).  T is column-major rank x n.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void interp_kernel(const C<R> *Z,
                                                     const InterpTask *__restrict__ tasks,
                                                     const int32_t *__restrict__ piv_all,
                                                     C<R> *T) {
    const InterpTask tk = tasks[blockIdx.x];
    const int m = tk.m, n = tk.n, r = tk.rank;
    const C<R> *A = Z + tk.zofs;
    const int32_t *piv = piv_all + tk.pofs;
    C<R> *out = T + tk.tofs;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int MAXR = 4;  // rank <= 256
    for (int j = wave; j < n; j += 4) {
        C<R> *t = out + int64_t(piv[j]) * r;
        if (j < r) {
            for (int i = lane; i < r; i += 64) t[i] = (i == j) ? C<R>{R(1), R(0)} : C<R>{R(0), R(0)};
            continue;
        }
        C<double> b[MAXR];
#pragma unroll
        for (int q = 0; q < MAXR; ++q) {
            int i = lane + 64 * q;
            C<R> v = (i < r) ? A[i + int64_t(j) * m] : C<R>{R(0), R(0)};
            b[q] = {double(v.x), double(v.y)};
        }
        for (int i = r - 1; i >= 0; --i) {
            // x_i = b_i / R_ii, then b_q -= R_qi x_i for q < i
            int owner = i & 63, slot = i >> 6;
            C<double> bi = {0.0, 0.0};
#pragma unroll
            for (int q = 0; q < MAXR; ++q)
                if (q == slot) bi = b[q];
            bi.x = __shfl(bi.x, owner, 64);
            bi.y = __shfl(bi.y, owner, 64);
            C<R> d = A[i + int64_t(i) * m];
            double dr = d.x, di = d.y, den = dr * dr + di * di;
            C<double> xi = {(bi.x * dr + bi.y * di) / den, (bi.y * dr - bi.x * di) / den};
#pragma unroll
            for (int q = 0; q < MAXR; ++q) {
                int row = lane + 64 * q;
                if (row < i) {
                    C<R> rv = A[row + int64_t(i) * m];
                    C<double> rq = {double(rv.x), double(rv.y)};
                    b[q] = csub(b[q], cmul(rq, xi));
                } else if (row == i) {
                    b[q] = xi;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < MAXR; ++q) {
            int i = lane + 64 * q;
            if (i < r) t[i] = C<R>{R(b[q].x), R(b[q].y)};
        }
    }
}

// Interpolation matrices of compressions of rank at most RB: one wavefront
// per compression, a lane per column, the column's right-hand side
// R[0:r, j] held in registers and solved by the same column-oriented FP64
// back substitution (x_i = b_i / R_ii, then b_q -= R_qi x_i for q < i, i
// from r - 1 down) as interp_kernel, so the results are the same; the
// entries of R11 are read by every lane at once (one address per load).
template <typename R, int RB>
__global__ __launch_bounds__(256) void interp_small_kernel(const C<R> *Z, const InterpTask *__restrict__ tasks,
                                                           int32_t n_tasks, const int32_t *__restrict__ piv_all,
                                                           C<R> *T) {
    const int64_t t = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (t >= n_tasks) return;
    const int lane = threadIdx.x & 63;
    const InterpTask tk = tasks[t];
    const int m = tk.m, n = tk.n, r = tk.rank;
    const C<R> *A = Z + tk.zofs;
    const int32_t *piv = piv_all + tk.pofs;
    C<R> *out = T + tk.tofs;
    for (int j = lane; j < n; j += 64) {
        C<R> *tc = out + int64_t(piv[j]) * r;
        if (j < r) {
            for (int i = 0; i < r; ++i) tc[i] = (i == j) ? C<R>{R(1), R(0)} : C<R>{R(0), R(0)};
            continue;
        }
        C<double> b[RB];
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            const C<R> v = i < r ? A[i + int64_t(j) * m] : C<R>{R(0), R(0)};
            b[i] = C<double>{double(v.x), double(v.y)};
        }
#pragma unroll
        for (int i = RB - 1; i >= 0; --i) {
            if (i >= r) continue;
            const C<R> d = A[i + int64_t(i) * m];
            const double dr = d.x, di = d.y, den = dr * dr + di * di;
            const C<double> xi = {(b[i].x * dr + b[i].y * di) / den, (b[i].y * dr - b[i].x * di) / den};
            b[i] = xi;
#pragma unroll
            for (int q = 0; q < i; ++q) {
                const C<R> rv = A[q + int64_t(i) * m];
                b[q] = csub(b[q], cmul(C<double>{double(rv.x), double(rv.y)}, xi));
            }
        }
#pragma unroll
        for (int i = 0; i < RB; ++i)
            if (i < r) tc[i] = C<R>{R(b[i].x), R(b[i].y)};
    }
}

void launch_interp_small(hipStream_t s, int prec, const void *Z, const InterpTask *tasks, int32_t n_tasks,
                         const int32_t *piv, void *T) {
    if (n_tasks <= 0) return;
    const int64_t nb = (int64_t(n_tasks) + 3) / 4;
    BFG_REQUIRE(nb < (int64_t(1) << 31), "interp: too many tasks");
    if (prec == BFGPU_C128)
        hipLaunchKernelGGL((interp_small_kernel<double, INTERP_SMALL_RANK>), dim3(unsigned(nb)), dim3(256), 0, s,
                           (const C<double> *)Z, tasks, n_tasks, piv, (C<double> *)T);
    else
        hipLaunchKernelGGL((interp_small_kernel<float, INTERP_SMALL_RANK>), dim3(unsigned(nb)), dim3(256), 0, s,
                           (const C<float> *)Z, tasks, n_tasks, piv, (C<float> *)T);
    BFG_HIP(hipGetLastError());
}

void launch_interp(hipStream_t s, int prec, const void *Z, const InterpTask *tasks,
                   int32_t n_tasks, const int32_t *piv, void *T, int32_t max_rank) {
    if (n_tasks <= 0) return;
    BFG_REQUIRE(max_rank <= 256, "interp: rank above 256 is not supported");
    for (int32_t t0 = 0; t0 < n_tasks; t0 += TASK_CHUNK) {
        const int32_t nt = std::min(TASK_CHUNK, n_tasks - t0);
        if (prec == BFGPU_C128)
            hipLaunchKernelGGL(interp_kernel<double>, dim3(nt), dim3(256), 0, s,
                               (const C<double> *)Z, tasks + t0, piv, (C<double> *)T);
        else
            hipLaunchKernelGGL(interp_kernel<float>, dim3(nt), dim3(256), 0, s,
                               (const C<float> *)Z, tasks + t0, piv, (C<float> *)T);
        BFG_HIP(hipGetLastError());
    }
}

// ---------------------------------------------------------------------------
// Row samples on the device: one wavefront per compression.  Step i of the
// sparse Fisher-Yates prefix draws j = i + floor(hi32(u) (n - i) / 2^32),
// u = SplitMix64(key + (i + 1) 0x9E3779B97F4A7C15), takes a_j (the value
// at position j, j if untouched) and moves a_i to position j; position i is
// never read again (later steps read positions > i), so only moved values
// are kept: a swap map of at most one entry per step, entry e in lane
// e % 64, register e / 64, found by ballot.  Lane i % 64 keeps a_j of step
// i, and the sample is written at the end in one pass.
// ---------------------------------------------------------------------------
__device__ inline uint64_t dmix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void sample_rows_kernel(const SampleTask *__restrict__ tasks,
                                                          const FillTask *__restrict__ fill, int32_t n_tasks,
                                                          const int32_t *__restrict__ perm,
                                                          const int32_t *__restrict__ big, int32_t *__restrict__ rows) {
    const int64_t q = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (q >= n_tasks) return;
    const int lane = threadIdx.x & 63;
    const SampleTask t = tasks[q];
    const FillTask f = fill[q];
    const int32_t m = f.m;
    int32_t mc = 0, hofs = 0;
    if (t.held >= 0) {
        const FillTask h = fill[n_tasks + t.held];
        mc = h.m;
        hofs = h.rofs;
    }
    const int32_t total = m + mc;
    auto dst = [&](int32_t i) { return i < m ? int64_t(f.rofs) + i : int64_t(hofs) + (i - m); };
    if (t.big >= 0) {
        for (int32_t i = lane; i < total; i += 64) rows[dst(i)] = big[int64_t(t.big) + i];
        return;
    }
    constexpr int NS = SAMPLE_DEVICE_MAX / 64;
    int32_t mk[NS], mv[NS], res[NS];
#pragma unroll
    for (int c = 0; c < NS; ++c) {
        mk[c] = -1;
        mv[c] = 0;
        res[c] = 0;
    }
    int32_t cnt = 0;
    for (int32_t i = 0; i < total; ++i) {
        const uint64_t u = dmix64(t.key + uint64_t(i + 1) * 0x9E3779B97F4A7C15ULL);
        const uint32_t span = uint32_t(t.n_obs - i);
        const int32_t j = i + int32_t(((u >> 32) * uint64_t(span)) >> 32);
        int32_t ai = i, aj = j, js = -1, jl = 0;
#pragma unroll
        for (int c = 0; c < NS; ++c) {
            if (64 * c >= cnt) break;
            const uint64_t bi = __ballot(mk[c] == i);
            if (bi) ai = __builtin_amdgcn_readlane(mv[c], int(__builtin_ctzll(bi)));
            const uint64_t bj = __ballot(mk[c] == j);
            if (bj) {
                jl = int(__builtin_ctzll(bj));
                aj = __builtin_amdgcn_readlane(mv[c], jl);
                js = c;
            }
        }
        if (j != i) {
            const int e = js >= 0 ? 64 * js + jl : cnt++;
#pragma unroll
            for (int c = 0; c < NS; ++c)
                if (c == (e >> 6) && lane == (e & 63)) {
                    mk[c] = j;
                    mv[c] = ai;
                }
        }
#pragma unroll
        for (int c = 0; c < NS; ++c)
            if (c == (i >> 6) && lane == (i & 63)) res[c] = aj;
    }
#pragma unroll
    for (int c = 0; c < NS; ++c) {
        const int32_t i = lane + 64 * c;
        if (i < total) rows[dst(i)] = perm[int64_t(t.obs0) + res[c]];
    }
}

void launch_sample_rows(hipStream_t s, const SampleTask *tasks, const FillTask *fill, int32_t n_tasks,
                        const int32_t *perm, const int32_t *big, int32_t *rows) {
    if (n_tasks <= 0) return;
    const int64_t nb = (int64_t(n_tasks) + 3) / 4;
    BFG_REQUIRE(nb < (int64_t(1) << 31), "sample: too many tasks");
    hipLaunchKernelGGL(sample_rows_kernel, dim3(unsigned(nb)), dim3(256), 0, s, tasks, fill, n_tasks, perm, big, rows);
    BFG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Residual judge of compressions sized by a rank prediction.  For each
// held-out row z (1 x n) the row elimination
//     s_k = z[piv[k]],  then for q < r:  w_q = s_q / R[q, q],
//     s_k -= w_q R[q, k]  (k > q)
// leaves s_k = z[piv[k]] - z[P1] R11^{-1} R[0:r, k], the residual of the
// interpolative decomposition at column piv[k] (k >= r), with no
// interpolation matrix formed.  Per task
//     e_max = max_{k >= r} |E[:, k]|_2,   z_max = max_k |Zc[:, k]|_2,
// in FP64.  One wavefront per task, lane per column (four chunks of 64,
// n <= JUDGE_MAX_COLS), w_q broadcast from the lane that owns column q;
// four tasks per workgroup, whole waves leave together.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void judge_kernel(const C<R> *__restrict__ Z, const int32_t *__restrict__ jq,
                                                    const QrTask *__restrict__ qr, const FillTask *__restrict__ held,
                                                    const int32_t *__restrict__ rank, int32_t n_tasks,
                                                    const int32_t *__restrict__ piv_all, double *__restrict__ out) {
    const int64_t t = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (t >= n_tasks) return;
    const int lane = threadIdx.x & 63;
    const int32_t q = jq[t];
    const QrTask qt = qr[q];
    const FillTask ht = held[t];
    const C<R> *Rm = Z + qt.zofs;
    const C<R> *Zc = Z + ht.zofs;
    const int32_t *piv = piv_all + qt.pofs;
    const int mc = ht.m, ldr = qt.m, n = qt.n, r = rank[q];
    if (n <= 64) {
        // one column per lane; eight held-out rows at a time, so each
        // entry of R is loaded once per eight rows
        constexpr int RBK = 8;
        const bool col = lane < n;
        const int k = col ? lane : 0;
        double en1 = 0.0, zn1 = 0.0;
        const int pk = piv[k];
        for (int i0 = 0; i0 < mc; i0 += RBK) {
            C<double> sv[RBK];
#pragma unroll
            for (int b = 0; b < RBK; ++b) {
                sv[b] = C<double>{0.0, 0.0};
                if (col && i0 + b < mc) {
                    const C<R> v = Zc[i0 + b + int64_t(pk) * mc];
                    sv[b] = C<double>{double(v.x), double(v.y)};
                    zn1 += sv[b].x * sv[b].x + sv[b].y * sv[b].y;
                }
            }
            for (int q = 0; q < r; ++q) {
                const C<R> d = Rm[q + int64_t(q) * ldr];
                const double dr = d.x, di = d.y, inv = 1.0 / (dr * dr + di * di);
                const C<R> rv = Rm[q + int64_t(k) * ldr];
                const C<double> rq = {double(rv.x), double(rv.y)};
#pragma unroll
                for (int b = 0; b < RBK; ++b) {
                    const double sx = __shfl(sv[b].x, q, 64), sy = __shfl(sv[b].y, q, 64);
                    const C<double> w = {(sx * dr + sy * di) * inv, (sy * dr - sx * di) * inv};
                    if (k > q) sv[b] = csub(sv[b], cmul(w, rq));
                }
            }
            if (col && k >= r)
#pragma unroll
                for (int b = 0; b < RBK; ++b)
                    if (i0 + b < mc) en1 += sv[b].x * sv[b].x + sv[b].y * sv[b].y;
        }
        double emax = col && k >= r ? sqrt(en1) : 0.0, zmax = col ? sqrt(zn1) : 0.0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            emax = fmax(emax, __shfl_xor(emax, o, 64));
            zmax = fmax(zmax, __shfl_xor(zmax, o, 64));
        }
        if (lane == 0) {
            out[2 * t] = emax;
            out[2 * t + 1] = zmax;
        }
        return;
    }
    constexpr int NC = JUDGE_MAX_COLS / 64;
    double en[NC], zn[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) en[c] = zn[c] = 0.0;
    for (int i = 0; i < mc; ++i) {
        C<double> sv[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int k = lane + 64 * c;
            sv[c] = C<double>{0.0, 0.0};
            if (k < n) {
                const C<R> v = Zc[i + int64_t(piv[k]) * mc];
                sv[c] = C<double>{double(v.x), double(v.y)};
                zn[c] += sv[c].x * sv[c].x + sv[c].y * sv[c].y;
            }
        }
        for (int q = 0; q < r; ++q) {
            const int oc = q >> 6, ol = q & 63;
            C<double> sq = {0.0, 0.0};
#pragma unroll
            for (int c = 0; c < NC; ++c)
                if (c == oc) sq = sv[c];
            sq.x = __shfl(sq.x, ol, 64);
            sq.y = __shfl(sq.y, ol, 64);
            const C<R> d = Rm[q + int64_t(q) * ldr];
            const double dr = d.x, di = d.y, den = dr * dr + di * di;
            const C<double> w = {(sq.x * dr + sq.y * di) / den, (sq.y * dr - sq.x * di) / den};
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const int k = lane + 64 * c;
                if (k > q && k < n) {
                    const C<R> rv = Rm[q + int64_t(k) * ldr];
                    sv[c] = csub(sv[c], cmul(w, C<double>{double(rv.x), double(rv.y)}));
                }
            }
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int k = lane + 64 * c;
            if (k >= r && k < n) en[c] += sv[c].x * sv[c].x + sv[c].y * sv[c].y;
        }
    }
    double emax = 0.0, zmax = 0.0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        emax = fmax(emax, sqrt(en[c]));
        zmax = fmax(zmax, sqrt(zn[c]));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        emax = fmax(emax, __shfl_xor(emax, o, 64));
        zmax = fmax(zmax, __shfl_xor(zmax, o, 64));
    }
    if (lane == 0) {
        out[2 * t] = emax;
        out[2 * t + 1] = zmax;
    }
}

void launch_judge(hipStream_t s, int prec, const void *Z, const int32_t *jq, const QrTask *qr,
                  const FillTask *held, const int32_t *rank, int32_t n_tasks, const int32_t *piv, double *out) {
    if (n_tasks <= 0) return;
    const int64_t nb = (int64_t(n_tasks) + 3) / 4;
    BFG_REQUIRE(nb < (int64_t(1) << 31), "judge: too many tasks");
    if (prec == BFGPU_C128)
        hipLaunchKernelGGL(judge_kernel<double>, dim3(unsigned(nb)), dim3(256), 0, s, (const C<double> *)Z, jq, qr,
                           held, rank, n_tasks, piv, out);
    else
        hipLaunchKernelGGL(judge_kernel<float>, dim3(unsigned(nb)), dim3(256), 0, s, (const C<float> *)Z, jq, qr,
                           held, rank, n_tasks, piv, out);
    BFG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Forward matvec stage (y = A x), one workgroup per (task, group of NR
// right-hand sides).  A task is an output vector (a slot, or rows of y
// reached through a permutation) with its first contribution inline and
// further contributions in `contribs`.  Thread t works on row i = t % rows
// and column group g = t / rows (G = 256 / rows groups); group g owns
// columns g, g + G, ...  For one column index q the active threads read G
// consecutive columns of the column-major block -- one contiguous range --
// and the factor loads of a pass are issued together with the input loads,
// so a task costs one memory round trip per pass.  Partial sums meet in LDS
// and are added in fixed group order (deterministic).
// ---------------------------------------------------------------------------
template <typename R, int INK, int OUTK, int NR>
__global__ __launch_bounds__(256) void fwd_kernel(
    const MvTask *__restrict__ tasks, const MvContrib *__restrict__ contribs,
    const C<R> *__restrict__ F, const C<R> *__restrict__ x, int64_t ldx, C<R> *__restrict__ y,
    int64_t ldy, C<R> *__restrict__ slots, int64_t slot_len, const int32_t *__restrict__ perm_in,
    const int32_t *__restrict__ perm_out, C<R> alpha, C<R> beta, int fuse_beta, int nrhs) {
    constexpr int QMAX = NR <= 2 ? 8 : 4;
    __shared__ C<R> red[NR][256];
    const MvTask tk = tasks[blockIdx.x];
    const int t = threadIdx.x;
    const int rb = blockIdx.y * NR;
    const int nr = min(NR, nrhs - rb);

    auto in_val = [&](int r, int idx) -> C<R> {
        if (r >= nr) return C<R>{R(0), R(0)};
        if (INK == IN_GATHER) return x[int64_t(rb + r) * ldx + perm_in[idx]];
        return slots[int64_t(rb + r) * slot_len + idx];
    };

    for (int row0 = 0; row0 < tk.out_len; row0 += 256) {
        const int rows = min(256, tk.out_len - row0);
        const int G = 256 / rows;
        const int i = t % rows, g = t / rows;
        const bool act = g < G;
        C<R> acc[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] = C<R>{R(0), R(0)};
        for (int ci = tk.c_begin - 1; ci < tk.c_end; ++ci) {
            MvContrib c;
            if (ci < tk.c_begin) c = MvContrib{tk.fofs, tk.m, tk.n, tk.in_ofs, 0};
            else c = contribs[ci];
            if (c.fofs < 0) {  // identity block
                if (g == 0)
#pragma unroll
                    for (int r = 0; r < NR; ++r) acc[r] = cadd(acc[r], in_val(r, c.in_ofs + row0 + i));
                continue;
            }
            const C<R> *A = F + c.fofs + row0 + i;
            for (int jb = g; jb < c.n; jb += G * QMAX) {
                C<R> a[QMAX], xv[NR][QMAX];
#pragma unroll
                for (int q = 0; q < QMAX; ++q) {
                    const int j = jb + q * G;
                    const bool ok = act && j < c.n;
                    a[q] = ok ? A[int64_t(j) * c.m] : C<R>{R(0), R(0)};
#pragma unroll
                    for (int r = 0; r < NR; ++r) xv[r][q] = ok ? in_val(r, c.in_ofs + j) : C<R>{R(0), R(0)};
                }
#pragma unroll
                for (int q = 0; q < QMAX; ++q)
#pragma unroll
                    for (int r = 0; r < NR; ++r) cfma(acc[r], a[q], xv[r][q]);
            }
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) red[r][t] = acc[r];
        __syncthreads();
        if (t < rows) {
            for (int r = 0; r < nr; ++r) {
                C<R> s = red[r][t];
                for (int q = 1; q < G; ++q) s = cadd(s, red[r][q * rows + t]);
                if (OUTK == OUT_SLOT) {
                    slots[int64_t(rb + r) * slot_len + tk.out_ofs + row0 + t] = s;
                } else {
                    C<R> *yp = y + int64_t(rb + r) * ldy + perm_out[tk.out_ofs + row0 + t];
                    C<R> old = *yp;
                    *yp = cadd(fuse_beta ? cmul(beta, old) : old, cmul(alpha, s));
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Forward stage, one wavefront per task (4 tasks per workgroup), for stages
// of small blocks (leaf-level Q/P blocks of a few hundred entries).  Rows
// are padded to rp = 2^ceil(log2 rows) <= 64 lanes; lane = g * rp + i works
// on row i and column group g of G = 64 / rp groups, so one load
// instruction of the wave reads G consecutive columns (contiguous).  Taller
// blocks are walked in 64-row chunks with G = 1.  Groups are summed with
// xor-shuffles (no LDS, no barriers).
// ---------------------------------------------------------------------------
template <typename R, int INK, int OUTK, int NR>
__global__ __launch_bounds__(256) void fwdw_kernel(
    const MvTask *__restrict__ tasks, int32_t n_tasks, const MvContrib *__restrict__ contribs,
    const C<R> *__restrict__ F, const C<R> *__restrict__ x, int64_t ldx, C<R> *__restrict__ y,
    int64_t ldy, C<R> *__restrict__ slots, int64_t slot_len, const int32_t *__restrict__ perm_in,
    const int32_t *__restrict__ perm_out, C<R> alpha, C<R> beta, int fuse_beta, int nrhs) {
    constexpr int QW = NR <= 2 ? 8 : 4;
    const int lane = threadIdx.x & 63;
    const int task = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (task >= n_tasks) return;  // whole wave leaves; no barriers below
    const MvTask tk = tasks[task];
    const int rb = blockIdx.y * NR;
    const int nr = min(NR, nrhs - rb);
    auto in_val = [&](int r, int idx) -> C<R> {
        if (r >= nr) return C<R>{R(0), R(0)};
        if (INK == IN_GATHER) return x[int64_t(rb + r) * ldx + perm_in[idx]];
        return slots[int64_t(rb + r) * slot_len + idx];
    };
    for (int row0 = 0; row0 < tk.out_len; row0 += 64) {
        const int rows = min(64, tk.out_len - row0);
        int rp = 1;
        while (rp < rows) rp <<= 1;
        const int G = 64 / rp;
        const int i = lane & (rp - 1), g = lane / rp;
        const bool rowok = i < rows;
        C<R> acc[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] = C<R>{R(0), R(0)};
        for (int ci = tk.c_begin - 1; ci < tk.c_end; ++ci) {
            MvContrib c;
            if (ci < tk.c_begin) c = MvContrib{tk.fofs, tk.m, tk.n, tk.in_ofs, 0};
            else c = contribs[ci];
            if (c.fofs < 0) {
                if (g == 0 && rowok)
#pragma unroll
                    for (int r = 0; r < NR; ++r) acc[r] = cadd(acc[r], in_val(r, c.in_ofs + row0 + i));
                continue;
            }
            const C<R> *A = F + c.fofs + row0 + i;
            for (int p0 = 0; p0 < c.n; p0 += G * QW) {
                C<R> a[QW], xv[NR][QW];
#pragma unroll
                for (int q = 0; q < QW; ++q) {
                    const int j = p0 + g + q * G;
                    const bool ok = rowok && j < c.n;
                    a[q] = ok ? A[int64_t(j) * c.m] : C<R>{R(0), R(0)};
#pragma unroll
                    for (int r = 0; r < NR; ++r) xv[r][q] = ok ? in_val(r, c.in_ofs + j) : C<R>{R(0), R(0)};
                }
#pragma unroll
                for (int q = 0; q < QW; ++q)
#pragma unroll
                    for (int r = 0; r < NR; ++r) cfma(acc[r], a[q], xv[r][q]);
            }
        }
        for (int s = rp; s < 64; s <<= 1)
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                acc[r].x += __shfl_xor(acc[r].x, s, 64);
                acc[r].y += __shfl_xor(acc[r].y, s, 64);
            }
        if (g == 0 && rowok) {
            for (int r = 0; r < nr; ++r) {
                if (OUTK == OUT_SLOT) {
                    slots[int64_t(rb + r) * slot_len + tk.out_ofs + row0 + i] = acc[r];
                } else {
                    C<R> *yp = y + int64_t(rb + r) * ldy + perm_out[tk.out_ofs + row0 + i];
                    C<R> old = *yp;
                    *yp = cadd(fuse_beta ? cmul(beta, old) : old, cmul(alpha, acc[r]));
                }
            }
        }
    }
}

template <typename R, int NR>
static void launch_fwdw_t(hipStream_t s, const StageTable &t, const void *factors, const void *x,
                          int64_t ldx, void *y, int64_t ldy, void *slots, int64_t slot_len,
                          const int32_t *perm_in, const int32_t *perm_out, int64_t nrhs,
                          const double *alpha, const double *beta, bool fuse) {
    C<R> a = {R(alpha[0]), R(alpha[1])}, b = {R(beta[0]), R(beta[1])};
    auto F = (const C<R> *)factors;
    auto X = (const C<R> *)x;
    auto Y = (C<R> *)y;
    auto S = (C<R> *)slots;
    auto P = t.contribs.as<const MvContrib>();
    for (int32_t t0 = 0; t0 < t.n_items; t0 += TASK_CHUNK) {
        const int32_t nt = std::min(TASK_CHUNK, t.n_items - t0);
        dim3 grid(unsigned((nt + 3) / 4), unsigned((nrhs + NR - 1) / NR));
        auto T = t.tasks.as<const MvTask>() + t0;
#define BFG_FWDW(INK, OUTK)                                                                \
    hipLaunchKernelGGL((fwdw_kernel<R, INK, OUTK, NR>), grid, dim3(256), 0, s, T, nt, P, F, X, \
                       ldx, Y, ldy, S, slot_len, perm_in, perm_out, a, b, int(fuse), int(nrhs))
        if (t.in_kind == IN_SLOT && t.out_kind == OUT_SLOT) BFG_FWDW(IN_SLOT, OUT_SLOT);
        else if (t.in_kind == IN_GATHER && t.out_kind == OUT_SLOT) BFG_FWDW(IN_GATHER, OUT_SLOT);
        else if (t.in_kind == IN_SLOT && t.out_kind == OUT_SCATTER) BFG_FWDW(IN_SLOT, OUT_SCATTER);
        else BFG_FWDW(IN_GATHER, OUT_SCATTER);
#undef BFG_FWDW
        BFG_HIP(hipGetLastError());
    }
}

template <typename R, int NR>
static void launch_fwd_t(hipStream_t s, const StageTable &t, const void *factors, const void *x,
                         int64_t ldx, void *y, int64_t ldy, void *slots, int64_t slot_len,
                         const int32_t *perm_in, const int32_t *perm_out, int64_t nrhs,
                         const double *alpha, const double *beta, bool fuse) {
    C<R> a = {R(alpha[0]), R(alpha[1])}, b = {R(beta[0]), R(beta[1])};
    auto F = (const C<R> *)factors;
    auto X = (const C<R> *)x;
    auto Y = (C<R> *)y;
    auto S = (C<R> *)slots;
    auto P = t.contribs.as<const MvContrib>();
    for (int32_t t0 = 0; t0 < t.n_items; t0 += TASK_CHUNK) {
        const int32_t nt = std::min(TASK_CHUNK, t.n_items - t0);
        dim3 grid(static_cast<unsigned>(nt), static_cast<unsigned>((nrhs + NR - 1) / NR));
        auto T = t.tasks.as<const MvTask>() + t0;
#define BFG_FWD(INK, OUTK)                                                                 \
    hipLaunchKernelGGL((fwd_kernel<R, INK, OUTK, NR>), grid, dim3(256), 0, s, T, P, F, X, ldx, \
                       Y, ldy, S, slot_len, perm_in, perm_out, a, b, int(fuse), int(nrhs))
        if (t.in_kind == IN_SLOT && t.out_kind == OUT_SLOT) BFG_FWD(IN_SLOT, OUT_SLOT);
        else if (t.in_kind == IN_GATHER && t.out_kind == OUT_SLOT) BFG_FWD(IN_GATHER, OUT_SLOT);
        else if (t.in_kind == IN_SLOT && t.out_kind == OUT_SCATTER) BFG_FWD(IN_SLOT, OUT_SCATTER);
        else BFG_FWD(IN_GATHER, OUT_SCATTER);
#undef BFG_FWD
        BFG_HIP(hipGetLastError());
    }
}

// Task descriptors of one wave's task group (tpw consecutive MvTaskC, at
// most 16 = 128 dwords): lane j holds dwords 2 j and 2 j + 1, read with one
// coalesced load; zero past the last task (an empty task, m = n = 0).
__device__ __forceinline__ uint2 load_task_group(const MvTaskC *__restrict__ tasks, int32_t n_tasks, int64_t gi,
                                                 int tpw, int lane) {
    const int64_t d = gi * tpw * 8 + 2 * lane;
    if (2 * lane < tpw * 8 && d < int64_t(n_tasks) * 8) return reinterpret_cast<const uint2 *>(tasks)[d >> 1];
    return make_uint2(0u, 0u);
}

struct TaskView {
    int64_t fofs;
    int in_ofs, out_ofs, m, n, c_begin, c_end;
};

// Fields of task `sub` of the group held in v; every lane of the wave must
// take part (the fields are cross-lane permutes).
__device__ __forceinline__ TaskView unpack_task(uint2 v, int sub) {
    auto f = [&](int k) { return __shfl(int((k & 1) ? v.y : v.x), sub * 4 + (k >> 1), 64); };
    TaskView t;
    t.fofs = int64_t((uint64_t(uint32_t(f(1))) << 32) | uint64_t(uint32_t(f(0))));
    t.in_ofs = f(2);
    t.out_ofs = f(3);
    t.m = f(4);
    t.n = f(5);
    t.c_begin = f(6);
    t.c_end = f(7);
    return t;
}

// ---------------------------------------------------------------------------
// Tiled stage kernel, forward (OP 0) and transposed (OP 1, conjugated when
// `conj`).  A wavefront holds 64 >> (rp_shift + g_shift) tasks; the lanes of
// one task are RP = 2^rp_shift output entries times G = 2^g_shift groups.
//   CSPLIT = false: group g takes the inner indices k = g, g + G, ... of
//     every contribution (one long block: leaf-level Q).
//   CSPLIT = true: group g takes contributions g, g + G, ... whole (many
//     short blocks: observer-leaf sums, transposed levels).
// Each pass issues QW factor and QW input loads per lane, so RP x G x QW
// entries of every task of the wave are in flight at once; groups meet by
// xor-shuffles inside the task's lanes (fixed order, deterministic).
// Waves are persistent: a wave walks task groups with a grid stride and
// loads the next group's descriptors (and each group the descriptor after
// the next contribution) before using the current ones.  Inputs come from
// the slot buffer at in_shift (IN_GATHER inputs are permuted there first by
// gather_kernel); outputs are slots, or rows of y through a permutation.
// Element (i, k) of a forward block is F[fofs + i + k m]; the transposed
// product reads F[fofs + k + i m] (column i of the block).
// ---------------------------------------------------------------------------
template <typename R, int OP, int OUTK, int NR, int QW, bool CSPLIT>
__global__ __launch_bounds__(256) void tile_kernel(
    const MvTaskC *__restrict__ tasks, int32_t n_tasks, const MvContrib *__restrict__ contribs,
    const C<R> *__restrict__ F, C<R> *__restrict__ slots, int64_t slot_len, int64_t in_shift,
    C<R> *__restrict__ y, int64_t ldy, const int32_t *__restrict__ perm_out, C<R> alpha, C<R> beta,
    int fuse_beta, int nrhs, int rp_shift, int g_shift, int conj, int xcd_map) {
    const int lane = threadIdx.x & 63;
    const int RP = 1 << rp_shift, G = 1 << g_shift;
    const int lpt_shift = rp_shift + g_shift;
    const int tpw = 64 >> lpt_shift;
    const int sub = lane >> lpt_shift;
    const int i = lane & (RP - 1);
    const int g = (lane >> rp_shift) & (G - 1);
    const int rb = blockIdx.y * NR;
    const int nr = min(NR, nrhs - rb);
    const R csign = conj ? R(-1) : R(1);
    const C<R> *in = slots + in_shift + int64_t(rb) * slot_len;
    const MvContrib none{-1, 0, 0, 0, 0};
    // task groups [grp_lo, grp_hi) walked with stride n_waves.  xcd_map:
    // workgroups are dispatched round-robin over the 8 XCDs, so workgroup b
    // runs on XCD b % 8; each XCD takes one contiguous eighth of the groups
    // (neighbouring blocks and the inputs they share stay in one L2)
    const int64_t n_groups = (int64_t(n_tasks) + tpw - 1) / tpw;
    int64_t grp, grp_hi, n_waves;
    if (xcd_map && gridDim.x >= 8) {  // every XCD holds a workgroup
        const int64_t nwg = gridDim.x, x = blockIdx.x % 8;
        n_waves = ((nwg - x + 7) / 8) * 4;
        grp = n_groups * x / 8 + (int64_t(blockIdx.x) / 8) * 4 + (threadIdx.x >> 6);
        grp_hi = n_groups * (x + 1) / 8;
    } else {
        n_waves = int64_t(gridDim.x) * 4;
        grp = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
        grp_hi = n_groups;
    }

    // contribution q of a task: q = 0 is the inline one
    auto contrib = [&](const TaskView &tk, int q) -> MvContrib {
        if (q == 0) return MvContrib{tk.fofs, tk.m, tk.n, tk.in_ofs, 0};
        const int ci = tk.c_begin + q - 1;
        return ci < tk.c_end ? contribs[ci] : none;
    };

    uint2 tv = load_task_group(tasks, n_tasks, grp, tpw, lane);
    while (grp < grp_hi) {
        const int64_t ngrp = grp + n_waves;
        const uint2 nv = ngrp < grp_hi ? load_task_group(tasks, n_tasks, ngrp, tpw, lane) : make_uint2(0u, 0u);
        TaskView tk = unpack_task(tv, sub);
        const int out_len = OP == 0 ? tk.m : tk.n;
        const int nc = out_len > 0 ? 1 + tk.c_end - tk.c_begin : 0;
        for (int row0 = 0; row0 < out_len; row0 += RP) {
            const int row = row0 + i;
            const bool rowok = row < out_len;
            C<R> acc[NR];
#pragma unroll
            for (int r = 0; r < NR; ++r) acc[r] = C<R>{R(0), R(0)};
            // Regular contribution: loads carry no predicate (row and inner
            // index clamped into the block); inputs of clamped entries are
            // zeroed by a select.
            auto apply = [&](const MvContrib &c, int kbeg, int kstep) {
                if (c.fofs < 0) {  // identity (m == n) or none (m == 0)
                    if (c.m > 0 && kbeg == 0 && rowok)
#pragma unroll
                        for (int r = 0; r < NR; ++r)
                            if (r < nr) acc[r] = cadd(acc[r], in[int64_t(r) * slot_len + c.in_ofs + row]);
                    return;
                }
                const int K = OP == 0 ? c.n : c.m, last = K - 1;
                const int rowc = min(row, out_len - 1);
                const C<R> *pa = F + c.fofs + (OP == 0 ? rowc : rowc * c.m);
                const int ks = OP == 0 ? c.m : 1;
                const C<R> *px = in + c.in_ofs;
                for (int k0 = kbeg; k0 < K; k0 += kstep * QW) {
                    C<R> a[QW], xv[NR][QW];
#pragma unroll
                    for (int q = 0; q < QW; ++q) {
                        const int k = k0 + q * kstep;
                        const int kk = min(k, last);
                        const bool ok = rowok && k < K;
                        a[q] = pa[kk * ks];
                        a[q].y *= csign;
#pragma unroll
                        for (int r = 0; r < NR; ++r) {
                            const C<R> v = px[int64_t(r < nr ? r : nr - 1) * slot_len + kk];
                            xv[r][q] = C<R>{ok && r < nr ? v.x : R(0), ok && r < nr ? v.y : R(0)};
                        }
                    }
#pragma unroll
                    for (int q = 0; q < QW; ++q)
#pragma unroll
                        for (int r = 0; r < NR; ++r) cfma(acc[r], a[q], xv[r][q]);
                }
            };
            if (CSPLIT) {
                // group g: contributions g, g + G, ...; descriptor of the next
                // one loaded before the current one is applied
                MvContrib cur = g < nc ? contrib(tk, g) : none;
                for (int q = g; q < nc; q += G) {
                    const MvContrib nx = q + G < nc ? contrib(tk, q + G) : none;
                    apply(cur, 0, 1);
                    cur = nx;
                }
            } else {
                MvContrib cur = contrib(tk, 0);
                for (int q = 0; q < nc; ++q) {
                    const MvContrib nx = q + 1 < nc ? contrib(tk, q + 1) : none;
                    apply(cur, g, G);
                    cur = nx;
                }
            }
            for (int s = RP; s < (RP << g_shift); s <<= 1)
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    acc[r].x += __shfl_xor(acc[r].x, s, 64);
                    acc[r].y += __shfl_xor(acc[r].y, s, 64);
                }
            if (g == 0 && rowok) {
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    if (r >= nr) break;
                    if (OUTK == OUT_SLOT) {
                        slots[int64_t(rb + r) * slot_len + tk.out_ofs + row] = acc[r];
                    } else {
                        C<R> *yp = y + int64_t(rb + r) * ldy + perm_out[tk.out_ofs + row];
                        const C<R> old = *yp;
                        *yp = cadd(fuse_beta ? cmul(beta, old) : old, cmul(alpha, acc[r]));
                    }
                }
            }
        }
        tv = nv;
        grp = ngrp;
    }
}

// ---------------------------------------------------------------------------
// Dense forward kernel.  As the tiled kernel, but the lanes of a task are
// packed to the block's own row count, two rows per lane: with LPT =
// 2^lpt_shift lanes per task and r = min(2 LPT, rows left) rows, r2 =
// ceil(r / 2) lanes form one group and lane l holds rows 2 i, 2 i + 1 with
// i = l mod r2 of group g = l / r2 (G = LPT / r2 groups; lanes l >= G r2
// idle).  One 16-byte load per lane reads both rows of a column, so the
// active lanes of one load read G r consecutive entries of a column-major
// block.  Groups split the inner index (CSPLIT false) or the contributions
// (CSPLIT true).  Factor loads carry no predicate: columns are clamped into
// the block (the second row of an odd block reads one entry past it: the
// factor buffers are allocated with a tail pad, and that row is dropped).
// The inputs of one pass are loaded once, one lane per column, and handed
// to the row lanes by cross-lane permutes; inputs of clamped columns and
// idle lanes are zeroed by a select.  Groups are summed by a shuffle tree
// over the stride r2, in fixed order.
// ---------------------------------------------------------------------------
template <typename R, int OUTK, int NR, int QW, bool CSPLIT>
__global__ __launch_bounds__(256) void dense_kernel(
    const MvTaskC *__restrict__ tasks, int32_t n_tasks, const MvContrib *__restrict__ contribs,
    const C<R> *__restrict__ F, C<R> *__restrict__ slots, int64_t slot_len, int64_t in_shift,
    C<R> *__restrict__ y, int64_t ldy, const int32_t *__restrict__ perm_out, C<R> alpha, C<R> beta,
    int fuse_beta, int nrhs, int lpt_shift, int xcd_map, int conj) {
    const int lane = threadIdx.x & 63;
    const int LPT = 1 << lpt_shift;
    const int tpw = 64 >> lpt_shift;
    const int sub = lane >> lpt_shift;
    const int l = lane & (LPT - 1);
    const int base = sub << lpt_shift;
    const int rb = blockIdx.y * NR;
    const int nr = min(NR, nrhs - rb);
    const C<R> *in = slots + in_shift + int64_t(rb) * slot_len;
    const MvContrib none{-1, 0, 0, 0, 0};
    const int64_t n_groups = (int64_t(n_tasks) + tpw - 1) / tpw;
    int64_t grp, grp_hi, n_waves;
    if (xcd_map && gridDim.x >= 8) {  // see tile_kernel
        const int64_t nwg = gridDim.x, x = blockIdx.x % 8;
        n_waves = ((nwg - x + 7) / 8) * 4;
        grp = n_groups * x / 8 + (int64_t(blockIdx.x) / 8) * 4 + (threadIdx.x >> 6);
        grp_hi = n_groups * (x + 1) / 8;
    } else {
        n_waves = int64_t(gridDim.x) * 4;
        grp = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
        grp_hi = n_groups;
    }
    auto contrib = [&](const TaskView &tk, int q) -> MvContrib {
        if (q == 0) return MvContrib{tk.fofs, tk.m, tk.n, tk.in_ofs, 0};
        const int ci = tk.c_begin + q - 1;
        return ci < tk.c_end ? contribs[ci] : none;
    };
    struct Pair { C<R> a, b; };

    uint2 tv = load_task_group(tasks, n_tasks, grp, tpw, lane);
    while (grp < grp_hi) {
        const int64_t ngrp = grp + n_waves;
        const uint2 nv = ngrp < grp_hi ? load_task_group(tasks, n_tasks, ngrp, tpw, lane) : make_uint2(0u, 0u);
        const TaskView tk = unpack_task(tv, sub);
        const int out_len = tk.m;
        const int nc = out_len > 0 ? 1 + tk.c_end - tk.c_begin : 0;
        for (int row0 = 0; row0 < out_len; row0 += 2 * LPT) {
            const int r = min(2 * LPT, out_len - row0);
            const int r2 = (r + 1) >> 1;
            // G = LPT / r2 and g = l / r2 through a reciprocal: exact for
            // operands below 2^7 (the quotient is off by less than 1/r2^2)
            const float inv = 1.0f / float(r2);
            const int G = int((float(LPT) + 0.5f) * inv);
            const int g = int((float(l) + 0.5f) * inv);
            const int i2 = l - g * r2;
            const bool act = g < G;
            const int row = row0 + 2 * i2;
            const bool two = 2 * i2 + 1 < r;
            C<R> acc0[NR], acc1[NR];
#pragma unroll
            for (int q = 0; q < NR; ++q) acc0[q] = acc1[q] = C<R>{R(0), R(0)};
            auto ident = [&](const MvContrib &c) {
#pragma unroll
                for (int rr = 0; rr < NR; ++rr)
                    if (rr < nr) {
                        const C<R> *p = in + int64_t(rr) * slot_len + c.in_ofs + row;
                        acc0[rr] = cadd(acc0[rr], p[0]);
                        if (two) acc1[rr] = cadd(acc1[rr], p[1]);
                    }
            };
            // pass over columns k0 + kc, kc = kc0 + q kcs < w; the inputs
            // are held by lanes xbase + j (column k0 + j, j < w)
            auto pass = [&](const C<R> *pa, const C<R> *px, int m, int K, int k0, int w, int xbase, int xl,
                            int kc0, int kcs) {
                const int last = K - 1;
                const int kx = k0 + xl;
                const bool okx = xl < w && kx < K;
                C<R> xs[NR];
#pragma unroll
                for (int rr = 0; rr < NR; ++rr) {
                    const C<R> v = px[int64_t(rr < nr ? rr : nr - 1) * slot_len + min(kx, last)];
                    xs[rr] = C<R>{okx ? v.x : R(0), okx ? v.y : R(0)};
                }
                Pair a[QW];
                C<R> xv[NR][QW];
#pragma unroll
                for (int q = 0; q < QW; ++q) {
                    const int kc = kc0 + q * kcs;
                    const int k = k0 + kc;
                    const bool ok = act && kc < w && k < K;
                    __builtin_memcpy(&a[q], __builtin_assume_aligned(pa + min(k, last) * m, sizeof(C<R>)),
                                     sizeof(Pair));
                    if (conj) {  // conjugated entries (adjoint over a transposed copy)
                        a[q].a.y = -a[q].a.y;
                        a[q].b.y = -a[q].b.y;
                    }
                    const int src = xbase + min(kc, w - 1);
#pragma unroll
                    for (int rr = 0; rr < NR; ++rr) {
                        const R vx = __shfl(xs[rr].x, src, 64), vy = __shfl(xs[rr].y, src, 64);
                        xv[rr][q] = C<R>{ok && rr < nr ? vx : R(0), ok && rr < nr ? vy : R(0)};
                    }
                }
#pragma unroll
                for (int q = 0; q < QW; ++q)
#pragma unroll
                    for (int rr = 0; rr < NR; ++rr) {
                        cfma(acc0[rr], a[q].a, xv[rr][q]);
                        cfma(acc1[rr], a[q].b, xv[rr][q]);
                    }
            };
            if (CSPLIT) {
                // group g takes contributions g, g + G, ...; its r2 lanes hold
                // the inputs of min(QW, r2) columns per pass
                const int w = min(QW, r2);
                MvContrib cur = g < nc ? contrib(tk, g) : none;
                for (int q = g; q < nc; q += G) {
                    const MvContrib nx = q + G < nc ? contrib(tk, q + G) : none;
                    if (cur.fofs < 0) {
                        if (cur.m > 0 && act) ident(cur);
                    } else {
                        for (int k0 = 0; k0 < cur.n; k0 += w)
                            pass(F + cur.fofs + row, in + cur.in_ofs, cur.m, cur.n, k0, w, base + g * r2, i2, 0, 1);
                    }
                    cur = nx;
                }
            } else {
                // groups split the columns: lane (g, i) takes columns
                // k0 + g + G q of a window of min(G QW, LPT) columns
                const int w = min(G * QW, LPT);
                MvContrib cur = contrib(tk, 0);
                for (int q = 0; q < nc; ++q) {
                    const MvContrib nx = q + 1 < nc ? contrib(tk, q + 1) : none;
                    if (cur.fofs < 0) {
                        if (cur.m > 0 && g == 0 && act) ident(cur);
                    } else {
                        for (int k0 = 0; k0 < cur.n; k0 += w)
                            pass(F + cur.fofs + row, in + cur.in_ofs, cur.m, cur.n, k0, w, base, l, g, G);
                    }
                    cur = nx;
                }
            }
            for (int s = 1; s < G; s <<= 1) {
                const int src = base + min(l + s * r2, LPT - 1);
                const bool take = g % (2 * s) == 0 && g + s < G;
#pragma unroll
                for (int rr = 0; rr < NR; ++rr) {
                    const R o0x = __shfl(acc0[rr].x, src, 64), o0y = __shfl(acc0[rr].y, src, 64);
                    const R o1x = __shfl(acc1[rr].x, src, 64), o1y = __shfl(acc1[rr].y, src, 64);
                    if (take) {
                        acc0[rr].x += o0x;
                        acc0[rr].y += o0y;
                        acc1[rr].x += o1x;
                        acc1[rr].y += o1y;
                    }
                }
            }
            if (g == 0) {
#pragma unroll
                for (int rr = 0; rr < NR; ++rr) {
                    if (rr >= nr) break;
                    if (OUTK == OUT_SLOT) {
                        C<R> *o = slots + int64_t(rb + rr) * slot_len + tk.out_ofs + row;
                        o[0] = acc0[rr];
                        if (two) o[1] = acc1[rr];
                    } else {
                        C<R> *yr = y + int64_t(rb + rr) * ldy;
                        C<R> *yp = yr + perm_out[tk.out_ofs + row];
                        C<R> old = *yp;
                        *yp = cadd(fuse_beta ? cmul(beta, old) : old, cmul(alpha, acc0[rr]));
                        if (two) {
                            yp = yr + perm_out[tk.out_ofs + row + 1];
                            old = *yp;
                            *yp = cadd(fuse_beta ? cmul(beta, old) : old, cmul(alpha, acc1[rr]));
                        }
                    }
                }
            }
        }
        tv = nv;
        grp = ngrp;
    }
}

// ---------------------------------------------------------------------------
// Streaming forward kernel (one right-hand side, single-block tasks).
// Blocks of consecutive tasks lie back to back in the factor buffer, so a
// chunk of tasks is one contiguous range of at most NLD KiB.  Each wavefront
// owns an LDS region; per chunk it
//   1. copies the range with aligned 16-byte loads (word q * 64 + lane in
//      round q: every load instruction reads 1 KiB of consecutive bytes),
//      and the chunk's input vectors (an entry per lane and round; the
//      entry's vector is the number of vector starts at or below it in the
//      chunk's xmask, minus one; lane v holds vector v's slot and LDS
//      offsets), into registers, then into its LDS region (a task that reads
//      several vectors finds them one after the other);
//   2. computes the chunk from LDS: lane (g, rho) takes row pair rho of the
//      chunk (rows 2 i, 2 i + 1 of the task that holds it: the number of
//      task starts at or below rho in rmask, minus one) over the columns
//      k = g, g + G, ...; the G groups meet in a fixed shuffle tree and
//      group 0 writes the outputs.
// Entries are conjugated when CONJ (the adjoint over the transposed copy of
// the blocks).  The copy of chunk c + 1 is issued before chunk c is
// computed, its task and vector descriptors one chunk earlier and its chunk
// record two chunks earlier, so
// the loads of the next chunk are in flight while the current one is
// computed.  No barriers: a wavefront reads only what it wrote (LDS
// instructions of one wavefront complete in order).  The second row of an
// odd-height block reads one entry past it (the region has a one-word pad)
// and is dropped.
// ---------------------------------------------------------------------------
typedef unsigned int bfg_u32x4 __attribute__((ext_vector_type(4)));

template <typename R>
__device__ __forceinline__ C<R> lds_c(const C<R> *p) {  // one entry, 8-byte aligned
    using V = typename std::conditional<sizeof(R) == 4, float2, double2>::type;
    const V v = *reinterpret_cast<const V *>(__builtin_assume_aligned(p, 8));
    return C<R>{v.x, v.y};
}

template <typename R, int NLD, bool CONJ>
__global__ __launch_bounds__(128) void stream_kernel(const StreamChunk *__restrict__ chunks, int32_t n_chunks,
                                                     const StreamTask *__restrict__ tasks,
                                                     const StreamRun *__restrict__ runs,
                                                     const C<R> *__restrict__ F, C<R> *__restrict__ slots,
                                                     int64_t in_shift, int xcd_map) {
    constexpr int NW = NLD * 64;  // words per chunk
    constexpr int XCAP = stream_xcap<R>();
    constexpr int XLD = XCAP / 64;
    static_assert(XCAP % 64 == 0 && XLD <= 4, "stream_kernel: input region");
    __shared__ uint4 fbuf[2][NW + 1];
    __shared__ C<R> xbuf[2][XCAP];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    uint4 *fw = fbuf[wv];
    const C<R> *fl = reinterpret_cast<const C<R> *>(fbuf[wv]);
    C<R> *xl = xbuf[wv];
    const uint4 *FW = reinterpret_cast<const uint4 *>(F);
    const C<R> *in = slots + in_shift;
    const uint64_t at_or_below = ~uint64_t(0) >> (63 - lane);  // bits 0..lane
    int64_t c, c_hi, stride;
    if (xcd_map && gridDim.x >= 8) {  // workgroup b runs on XCD b % 8: each XCD walks one eighth
        const int64_t nwg = gridDim.x, x = blockIdx.x % 8;
        stride = ((nwg - x + 7) / 8) * 2;
        c = int64_t(n_chunks) * x / 8 + (int64_t(blockIdx.x) / 8) * 2 + wv;
        c_hi = int64_t(n_chunks) * (x + 1) / 8;
    } else {
        stride = int64_t(gridDim.x) * 2;
        c = int64_t(blockIdx.x) * 2 + wv;
        c_hi = n_chunks;
    }
    if (c >= c_hi) return;

    struct Rec {
        int64_t f0w;
        int32_t nw, X, t0, nt, H, G, r0, nr;
        uint64_t rm, xm[XLD];
    };
    struct Desc {
        uint4 t;  // lane t: task t of the chunk
        int2 v;   // lane v: input vector v of the chunk (slot offset, LDS offset)
    };
    auto rec = [&](int64_t ci) -> Rec {  // nt = 0 past the wave's last chunk
        Rec r;
        r.f0w = 0;
        r.nw = r.X = r.t0 = r.nt = r.r0 = r.nr = 0;
        r.H = r.G = 1;
        r.rm = 1;
#pragma unroll
        for (int q = 0; q < XLD; ++q) r.xm[q] = 0;
        if (ci < c_hi) {
            const StreamChunk &k = chunks[ci];
            r.f0w = int64_t(k.w0 & ((uint64_t(1) << 40) - 1));
            r.nw = int32_t((k.w0 >> 40) & 2047);
            r.X = int32_t((k.w0 >> 51) & 2047);
            r.t0 = k.t0;
            r.nt = int32_t(k.info & 127);
            r.H = int32_t((k.info >> 7) & 127);
            r.G = int32_t((k.info >> 14) & 127);
            r.rm = k.rmask;
#pragma unroll
            for (int q = 0; q < XLD; ++q) r.xm[q] = k.xmask[q];
            r.r0 = k.r0;
            r.nr = k.nruns;
        }
        return r;
    };
    auto desc = [&](const Rec &r) -> Desc {
        Desc d;
        d.t = lane < r.nt ? reinterpret_cast<const uint4 *>(tasks)[r.t0 + lane] : make_uint4(0u, 0u, 0u, 0u);
        d.v = lane < r.nr ? reinterpret_cast<const int2 *>(runs)[r.r0 + lane] : make_int2(0, 0);
        return d;
    };
    auto issue = [&](const Rec &r, const Desc &d, uint4(&V)[NLD], C<R>(&XV)[XLD]) {
#pragma unroll
        for (int q = 0; q < NLD; ++q) {
            const int wi = q * 64 + lane;
            // factor words are read once: non-temporal, so that they do not
            // evict the input vectors other chunks read again
            if (wi < r.nw) {
                const bfg_u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const bfg_u32x4 *>(FW) + r.f0w + wi);
                V[q] = make_uint4(v.x, v.y, v.z, v.w);
            }
        }
        int before = 0;  // input vectors starting in earlier rounds
#pragma unroll
        for (int q = 0; q < XLD; ++q) {
            if (q * 64 >= r.X) break;
            const int e = q * 64 + lane;
            const int v = before + __popcll(r.xm[q] & at_or_below) - 1;
            before += __popcll(r.xm[q]);
            const int kv = __shfl(d.v.y, v, 64), lv = __shfl(d.v.x, v, 64);
            if (e < r.X) XV[q] = in[int64_t(lv) + (e - kv)];
        }
    };
    auto store = [&](const Rec &r, const uint4(&V)[NLD], const C<R>(&XV)[XLD]) {
#pragma unroll
        for (int q = 0; q < NLD; ++q)
            if (q * 64 < r.nw) fw[q * 64 + lane] = V[q];
#pragma unroll
        for (int q = 0; q < XLD; ++q)
            if (q * 64 < r.X) xl[q * 64 + lane] = XV[q];
    };
    auto compute = [&](const Rec &r, const Desc &dd) {
        const uint4 &d = dd.t;
        const int H = r.H, G = r.G;
        // g = lane / H through a reciprocal: exact for operands below 2^7
        const int g = int((float(lane) + 0.5f) * __builtin_amdgcn_rcpf(float(H)));
        const int rho = lane - g * H;
        const int t = __popcll(r.rm & (~uint64_t(0) >> (63 - rho))) - 1;
        const uint32_t tz = __shfl(d.z, t, 64), tw = __shfl(d.w, t, 64);
        const int oo = __shfl(int(d.y), t, 64);
        const int a = int(tz & 2047), xo = int((tz >> 11) & 1023);
        const int m = int(tw & 255), hp = int((tw >> 8) & 127), n = int(tw >> 15);
        const int i = 2 * (rho - hp);
        C<R> acc0{R(0), R(0)}, acc1{R(0), R(0)};
        if (g < G) {
            const int step = G * m;
            const C<R> *pa = fl + a + i + g * m;
            const C<R> *px = xl + xo + g;
            int k = g;
            for (; k + 3 * G < n; k += 4 * G) {
                C<R> a0[4], a1[4], xv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    a0[u] = lds_c(pa + u * step);
                    a1[u] = lds_c(pa + u * step + 1);
                    xv[u] = lds_c(px + u * G);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (CONJ) {
                        cfmac(acc0, a0[u], xv[u]);
                        cfmac(acc1, a1[u], xv[u]);
                    } else {
                        cfma(acc0, a0[u], xv[u]);
                        cfma(acc1, a1[u], xv[u]);
                    }
                }
                pa += 4 * step;
                px += 4 * G;
            }
            for (; k < n; k += G) {
                const C<R> x = lds_c(px);
                if (CONJ) {
                    cfmac(acc0, lds_c(pa), x);
                    cfmac(acc1, lds_c(pa + 1), x);
                } else {
                    cfma(acc0, lds_c(pa), x);
                    cfma(acc1, lds_c(pa + 1), x);
                }
                pa += step;
                px += G;
            }
        }
        for (int sft = 1; sft < G; sft <<= 1) {
            const int src = min(lane + sft * H, 63);
            const bool take = (g & (2 * sft - 1)) == 0 && g + sft < G;
            const R o0x = __shfl(acc0.x, src, 64), o0y = __shfl(acc0.y, src, 64);
            const R o1x = __shfl(acc1.x, src, 64), o1y = __shfl(acc1.y, src, 64);
            if (take) {
                acc0.x += o0x;
                acc0.y += o0y;
                acc1.x += o1x;
                acc1.y += o1y;
            }
        }
        if (g == 0) {  // lane rho < H
            C<R> *o = slots + int64_t(oo) + i;
            o[0] = acc0;
            if (i + 1 < m) o[1] = acc1;
        }
    };

    uint4 V[NLD];
    C<R> XV[XLD];
#pragma unroll
    for (int q = 0; q < NLD; ++q) V[q] = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
    for (int q = 0; q < XLD; ++q) XV[q] = C<R>{R(0), R(0)};
    Rec r0 = rec(c), r1 = rec(c + stride), r2 = rec(c + 2 * stride);
    Desc d0 = desc(r0), d1 = desc(r1);
    issue(r0, d0, V, XV);
    for (;;) {
        store(r0, V, XV);
        if (r1.nt > 0) issue(r1, d1, V, XV);
        const Desc d2 = desc(r2);
        const Rec r3 = rec(c + 3 * stride);
        compute(r0, d0);
        if (r1.nt == 0) break;
        r0 = r1;
        r1 = r2;
        r2 = r3;
        d0 = d1;
        d1 = d2;
        c += stride;
    }
}

// ---------------------------------------------------------------------------
// Dense transposed kernel (adjoint / transpose): out[j] = sum_i op(A[i, j])
// in[i].  Lane l of a task owns output column j = j0 + l (LPT outputs per
// pass), so its accumulator persists over all contributions and no
// cross-lane reduction is needed; the lane walks its block column, which
// is contiguous, with 16-byte loads of two entries.  The inputs of an inner
// chunk of LPT entries are loaded once, one lane per entry, and handed to
// every lane of the task by cross-lane permutes (XL false), or stored by
// those lanes into the wavefront's own LDS rows and read back by every lane
// of the task as 16-byte broadcasts, two entries per read (XL true: one LDS
// read where the permutes take four).  Loads carry no predicate (column
// clamped into the block; the entry after the last row of a column reads
// the next column or the buffer's tail pad and meets a zeroed input).
// ---------------------------------------------------------------------------
template <typename R, int OUTK, int NR, bool XL>
__global__ __launch_bounds__(256) void denset_kernel(
    const MvTaskC *__restrict__ tasks, int32_t n_tasks, const MvContrib *__restrict__ contribs,
    const C<R> *__restrict__ F, C<R> *__restrict__ slots, int64_t slot_len, int64_t in_shift,
    C<R> *__restrict__ y, int64_t ldy, const int32_t *__restrict__ perm_out, C<R> alpha, C<R> beta,
    int fuse_beta, int nrhs, int lpt_shift, int xcd_map, int conj) {
    constexpr int UP = 4;  // pairs per unrolled step (8 inner entries)
    const int lane = threadIdx.x & 63;
    const int LPT = 1 << lpt_shift;
    const int tpw = 64 >> lpt_shift;
    const int sub = lane >> lpt_shift;
    const int l = lane & (LPT - 1);
    const int base = sub << lpt_shift;
    const int rb = blockIdx.y * NR;
    const int nr = min(NR, nrhs - rb);
    const R csign = conj ? R(-1) : R(1);
    const C<R> *in = slots + in_shift + int64_t(rb) * slot_len;
    const MvContrib none{-1, 0, 0, 0, 0};
    const int64_t n_groups = (int64_t(n_tasks) + tpw - 1) / tpw;
    int64_t grp, grp_hi, n_waves;
    if (xcd_map && gridDim.x >= 8) {  // see tile_kernel
        const int64_t nwg = gridDim.x, x = blockIdx.x % 8;
        n_waves = ((nwg - x + 7) / 8) * 4;
        grp = n_groups * x / 8 + (int64_t(blockIdx.x) / 8) * 4 + (threadIdx.x >> 6);
        grp_hi = n_groups * (x + 1) / 8;
    } else {
        n_waves = int64_t(gridDim.x) * 4;
        grp = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
        grp_hi = n_groups;
    }
    auto contrib = [&](const TaskView &tk, int q) -> MvContrib {
        if (q == 0) return MvContrib{tk.fofs, tk.m, tk.n, tk.in_ofs, 0};
        const int ci = tk.c_begin + q - 1;
        return ci < tk.c_end ? contribs[ci] : none;
    };
    struct Pair { C<R> a, b; };
    // inputs of the current chunk, per wavefront: [rhs][lane] (XL only)
    __shared__ __align__(32) C<R> xlds[XL ? 4 * 64 * NR : 1];
    C<R> *xrow = xlds + (XL ? (threadIdx.x >> 6) * 64 * NR : 0);

    uint2 tv = load_task_group(tasks, n_tasks, grp, tpw, lane);
    while (grp < grp_hi) {
        const int64_t ngrp = grp + n_waves;
        const uint2 nv = ngrp < grp_hi ? load_task_group(tasks, n_tasks, ngrp, tpw, lane) : make_uint2(0u, 0u);
        const TaskView tk = unpack_task(tv, sub);
        const int out_len = tk.n;
        const int nc = out_len > 0 ? 1 + tk.c_end - tk.c_begin : 0;
        for (int j0 = 0; j0 < out_len; j0 += LPT) {
            const int j = j0 + l;
            const bool jok = j < out_len;
            const int jc = min(j, out_len - 1);
            C<R> acc[NR];
#pragma unroll
            for (int q = 0; q < NR; ++q) acc[q] = C<R>{R(0), R(0)};
            MvContrib cur = contrib(tk, 0);
            for (int q = 0; q < nc; ++q) {
                const MvContrib nx = q + 1 < nc ? contrib(tk, q + 1) : none;
                if (cur.fofs < 0) {  // identity: out[j] += in[j]
                    if (cur.m > 0 && jok)
#pragma unroll
                        for (int rr = 0; rr < NR; ++rr)
                            if (rr < nr) acc[rr] = cadd(acc[rr], in[int64_t(rr) * slot_len + cur.in_ofs + j]);
                } else {
                    const C<R> *col = F + cur.fofs + int64_t(jc) * cur.m;
                    const C<R> *px = in + cur.in_ofs;
                    for (int i0 = 0; i0 < cur.m; i0 += LPT) {
                        const int iw = min(LPT, cur.m - i0);
                        C<R> xs[NR];
#pragma unroll
                        for (int rr = 0; rr < NR; ++rr) {
                            const C<R> v = px[int64_t(rr < nr ? rr : nr - 1) * slot_len + i0 + min(l, iw - 1)];
                            xs[rr] = C<R>{l < iw ? v.x : R(0), l < iw ? v.y : R(0)};
                        }
                        if (XL) {
                            // the task's LDS row holds this chunk's inputs (zero past iw);
                            // earlier reads of the row were issued before these stores
#pragma unroll
                            for (int rr = 0; rr < NR; ++rr) xrow[rr * 64 + lane] = xs[rr];
                            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                            __builtin_amdgcn_wave_barrier();
                            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                        }
                        for (int ii = 0; ii < iw; ii += 2 * UP) {
                            Pair p[UP];
#pragma unroll
                            for (int u = 0; u < UP; ++u)
                                __builtin_memcpy(&p[u],
                                                 __builtin_assume_aligned(col + i0 + min(ii + 2 * u, iw - 1),
                                                                          sizeof(C<R>)),
                                                 sizeof(Pair));
#pragma unroll
                            for (int u = 0; u < UP; ++u) {
                                const int e0 = ii + 2 * u, e1 = e0 + 1;
                                if (XL) {
                                    // entries e0 < iw and e1 <= LPT - 1 of the task's row (e0
                                    // even; the row is zero from iw on); steps past iw add nothing
                                    if (e0 >= iw) continue;
#pragma unroll
                                    for (int rr = 0; rr < NR; ++rr) {
                                        Pair xp;
                                        __builtin_memcpy(&xp, &xrow[rr * 64 + base + e0], sizeof(Pair));
                                        cfma(acc[rr], C<R>{p[u].a.x, p[u].a.y * csign}, xp.a);
                                        cfma(acc[rr], C<R>{p[u].b.x, p[u].b.y * csign}, xp.b);
                                    }
                                    continue;
                                }
                                const int s0 = base + min(e0, iw - 1), s1 = base + min(e1, iw - 1);
#pragma unroll
                                for (int rr = 0; rr < NR; ++rr) {
                                    const R x0r = __shfl(xs[rr].x, s0, 64), x0i = __shfl(xs[rr].y, s0, 64);
                                    const R x1r = __shfl(xs[rr].x, s1, 64), x1i = __shfl(xs[rr].y, s1, 64);
                                    const bool ok0 = e0 < iw, ok1 = e1 < iw;
                                    const C<R> x0{ok0 ? x0r : R(0), ok0 ? x0i : R(0)};
                                    const C<R> x1{ok1 ? x1r : R(0), ok1 ? x1i : R(0)};
                                    cfma(acc[rr], C<R>{p[u].a.x, p[u].a.y * csign}, x0);
                                    cfma(acc[rr], C<R>{p[u].b.x, p[u].b.y * csign}, x1);
                                }
                            }
                        }
                    }
                }
                cur = nx;
            }
            if (jok) {
#pragma unroll
                for (int rr = 0; rr < NR; ++rr) {
                    if (rr >= nr) break;
                    if (OUTK == OUT_SLOT) {
                        slots[int64_t(rb + rr) * slot_len + tk.out_ofs + j] = acc[rr];
                    } else {
                        C<R> *yp = y + int64_t(rb + rr) * ldy + perm_out[tk.out_ofs + j];
                        const C<R> old = *yp;
                        *yp = cadd(fuse_beta ? cmul(beta, old) : old, cmul(alpha, acc[rr]));
                    }
                }
            }
        }
        tv = nv;
        grp = ngrp;
    }
}

// ---------------------------------------------------------------------------
// Matrix-free last level (forward, one RHS, complex FP32).  The P blocks of
// the point kernel are kernel entries G(x_i, y_c) between the points of an
// observer leaf and the skeleton points of its P columns, so the level is
// computed from the points instead of read: one workgroup per observer
// leaf, the four wavefronts on quarters of its flat column list, a lane per
// column, rows in chunks of 16 (8 or 4 for a leaf's last few rows)
// accumulated per lane (row data are uniform loads).  Coordinates are local
// to the leaf centre c: rows hold u = x - c and |u|^2, columns d = y - c and
// the phase of |d| in revolutions (from FP64), so the distance enters as
//   r = |d| + delta,  delta = (|u|^2 - 2 d.u) / (r + |d|),
// and the phase as frac(phase(|d|) + k delta / (2 pi)): its FP32 error is
// relative to the leaf's size, not to the distance, and does not grow with
// the electrical size of the geometry.  r^2 = |d|^2 + (|u|^2 - 2 d.u) reuses
// the numerator (one add per entry instead of three differences and a
// square norm); 1 / (4 pi r) is the reciprocal square root of r^2
// (1 / (4 pi) folded into the column's input value); the entry is VALU
// bound (about 14 full-rate and 4 transcendental instructions);
// sine and cosine are the hardware ones on revolutions.  The 16 row sums
// leave the 64 lanes by a halving exchange and the wavefronts meet in LDS,
// in fixed order.
// ---------------------------------------------------------------------------
// One chunk of NR rows (16; 8 or 4 for a leaf's last few rows) of an item.
template <int NR>
__device__ __forceinline__ void pfree_rows(const int4 it, int r0, int cb, int ce, int lane, int w,
                                           const float4 *__restrict__ cdat, const int32_t *__restrict__ colin,
                                           const float4 *__restrict__ rdat, const C<float> *__restrict__ slots,
                                           float kr, C<float> (*red)[16], C<float> *__restrict__ y,
                                           const int32_t *__restrict__ perm_out, C<float> alpha, C<float> beta,
                                           int fuse_beta) {
    constexpr float inv4pi = 0.07957747154594767f;
    constexpr int NS = NR == 16 ? 4 : NR == 8 ? 3 : 2;  // exchange steps: 2^NS = NR rows
    const int nrow = min(NR, it.y - r0);
    float ux[NR], uy[NR], uz[NR], uu[NR];
#pragma unroll
    for (int rr = 0; rr < NR; ++rr) {  // rows past nrow repeat the last one, unused
        const float4 p = rdat[it.x + r0 + min(rr, nrow - 1)];
        ux[rr] = p.x;
        uy[rr] = p.y;
        uz[rr] = p.z;
        uu[rr] = p.w;
    }
    C<float> acc[NR];
#pragma unroll
    for (int rr = 0; rr < NR; ++rr) acc[rr] = C<float>{0.f, 0.f};
    for (int c0 = cb; c0 < ce; c0 += 64) {
        const int c = c0 + lane;
        const int cc = min(c, ce - 1);
        const float4 d = cdat[cc];
        const C<float> xv0 = slots[colin[cc]];
        const float sc = c < ce ? inv4pi : 0.f;
        const C<float> xv{xv0.x * sc, xv0.y * sc};
        const float D2 = fmaf(d.x, d.x, fmaf(d.y, d.y, d.z * d.z));
        const float D = D2 * __builtin_amdgcn_rsqf(D2);
        const float mx = -2.f * d.x, my = -2.f * d.y, mz = -2.f * d.z;
#pragma unroll
        for (int rr = 0; rr < NR; ++rr) {
            const float num = fmaf(mx, ux[rr], fmaf(my, uy[rr], fmaf(mz, uz[rr], uu[rr])));
            const float r2 = D2 + num;  // |d - u|^2 = |d|^2 - 2 d.u + |u|^2
            const float rs = __builtin_amdgcn_rsqf(r2);
            const float delta = num * __builtin_amdgcn_rcpf(fmaf(r2, rs, D));
            const float rev = __builtin_amdgcn_fractf(fmaf(delta, kr, d.w));
            const float cs = __builtin_amdgcn_cosf(rev), sn = __builtin_amdgcn_sinf(rev);
            const float tx = rs * xv.x, ty = rs * xv.y;
            acc[rr].x = fmaf(cs, tx, fmaf(sn, ty, acc[rr].x));  // (cos - i sin) (tx + i ty)
            acc[rr].y = fmaf(cs, ty, fmaf(-sn, tx, acc[rr].y));
        }
    }
    // halving exchange: after the steps over lane bits 5 .. 6 - NS a lane
    // holds row (lane >> (6 - NS)) & (NR - 1) summed over 2^NS lanes; the
    // lower bits in full
#pragma unroll
    for (int st = 0; st < NS; ++st) {
        const int s = 32 >> st, K = NR >> st, H = K >> 1;
        const bool hi = (lane & s) != 0;
#pragma unroll
        for (int u = 0; u < H; ++u) {
            const C<float> lo_v = acc[u], hi_v = acc[u + H];
            const C<float> send = hi ? lo_v : hi_v;
            const C<float> recv{__shfl_xor(send.x, s, 64), __shfl_xor(send.y, s, 64)};
            acc[u] = cadd(hi ? hi_v : lo_v, recv);
        }
    }
#pragma unroll
    for (int s = 32 >> NS; s > 0; s >>= 1)
        acc[0] = cadd(acc[0], C<float>{__shfl_xor(acc[0].x, s, 64), __shfl_xor(acc[0].y, s, 64)});
    __syncthreads();  // red is free (previous chunk read)
    if ((lane & ((1 << (6 - NS)) - 1)) == 0) red[w][(lane >> (6 - NS)) & (NR - 1)] = acc[0];
    __syncthreads();
    if (w == 0 && lane < nrow) {
        C<float> sum = red[0][lane];
        sum = cadd(sum, red[1][lane]);
        sum = cadd(sum, red[2][lane]);
        sum = cadd(sum, red[3][lane]);
        C<float> *yp = y + perm_out[it.x + r0 + lane];
        *yp = cadd(fuse_beta ? cmul(beta, *yp) : *yp, cmul(alpha, sum));
    }
}

__global__ __launch_bounds__(256) void pfree_kernel(const int4 *__restrict__ items,
                                                    const float4 *__restrict__ cdat,
                                                    const int32_t *__restrict__ colin,
                                                    const float4 *__restrict__ rdat,
                                                    const C<float> *__restrict__ slots, float kr,
                                                    C<float> *__restrict__ y, const int32_t *__restrict__ perm_out,
                                                    C<float> alpha, C<float> beta, int fuse_beta) {
    __shared__ C<float> red[4][16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int4 it = items[blockIdx.x];  // first row (tree position), rows, columns [z, w)
    const int ncol = it.w - it.z;
    const int q = (((ncol + 3) >> 2) + 63) & ~63;
    const int cb = it.z + w * q, ce = min(it.w, cb + q);
    for (int r0 = 0; r0 < it.y;) {  // chunks of 16 rows; at most 8 / 4 left: one chunk of 8 / 4
        if (it.y - r0 > 8) {
            pfree_rows<16>(it, r0, cb, ce, lane, w, cdat, colin, rdat, slots, kr, red, y, perm_out, alpha, beta,
                           fuse_beta);
            r0 += 16;
        } else if (it.y - r0 <= 4) {
            pfree_rows<4>(it, r0, cb, ce, lane, w, cdat, colin, rdat, slots, kr, red, y, perm_out, alpha, beta,
                          fuse_beta);
            r0 += 4;
        } else {
            pfree_rows<8>(it, r0, cb, ce, lane, w, cdat, colin, rdat, slots, kr, red, y, perm_out, alpha, beta,
                          fuse_beta);
            r0 += 8;
        }
    }
}

void launch_pfree(hipStream_t s, const BF &bf, const void *slots, void *y, int64_t ldy, const double *alpha,
                  const double *beta, bool fuse) {
    (void)ldy;
    if (bf.pf_n_items == 0) return;
    hipLaunchKernelGGL(pfree_kernel, dim3(unsigned(bf.pf_n_items)), dim3(256), 0, s, bf.pf_items.as<const int4>(),
                       bf.pf_col.as<const float4>(), bf.pf_in.as<const int32_t>(), bf.pf_row.as<const float4>(),
                       static_cast<const C<float> *>(slots),
                       float(bf.pf_k * 0.15915494309189535), static_cast<C<float> *>(y),
                       bf.perm_test.as<const int32_t>(), C<float>{float(alpha[0]), float(alpha[1])},
                       C<float>{float(beta[0]), float(beta[1])}, fuse ? 1 : 0);
    BFG_HIP(hipGetLastError());
}

// Matrix-free last level of the transposed / adjoint product (one RHS,
// complex FP32): the first step of the transposed sweep writes every P
// column's skeleton entry, sum over the leaf's rows of G(x_i, y_c) x_i
// (conjugated for the adjoint).  One workgroup per observer leaf, a thread
// per column (256 columns per pass, each written by exactly one thread), the
// leaf's rows (local coordinates and input values) staged in LDS 256 at a
// time and read as broadcasts.  Entries as in the forward kernel.
template <bool CONJ>
__global__ __launch_bounds__(256) void pfree_t_kernel(const int4 *__restrict__ items,
                                                      const float4 *__restrict__ cdat,
                                                      const int32_t *__restrict__ colin,
                                                      const float4 *__restrict__ rdat,
                                                      const C<float> *__restrict__ x,
                                                      const int32_t *__restrict__ perm_in, float kr,
                                                      C<float> *__restrict__ slots) {
    __shared__ float4 su[256];
    __shared__ C<float> sx[256];
    const int t = threadIdx.x;
    const int4 it = items[blockIdx.x];  // first row (tree position), rows, columns [z, w)
    constexpr float inv4pi = 0.07957747154594767f;
    constexpr float sgn = CONJ ? 1.f : -1.f;  // exp(-i k r), conjugated for the adjoint
    for (int cb = it.z; cb < it.w; cb += 256) {
        const int c = cb + t, cc = min(c, it.w - 1);
        const float4 d = cdat[cc];
        const float D2 = fmaf(d.x, d.x, fmaf(d.y, d.y, d.z * d.z));
        const float D = D2 * __builtin_amdgcn_rsqf(D2);
        const float mx = -2.f * d.x, my = -2.f * d.y, mz = -2.f * d.z;
        C<float> acc{0.f, 0.f};
        for (int r0 = 0; r0 < it.y; r0 += 256) {
            const int nr = min(256, it.y - r0);
            __syncthreads();  // the previous rows are read
            if (t < nr) {
                su[t] = rdat[it.x + r0 + t];
                const C<float> xv = x[perm_in[it.x + r0 + t]];
                sx[t] = C<float>{xv.x * inv4pi, xv.y * inv4pi};
            }
            __syncthreads();
#pragma unroll 4
            for (int r = 0; r < nr; ++r) {
                const float4 u = su[r];
                const C<float> xv = sx[r];
                const float num = fmaf(mx, u.x, fmaf(my, u.y, fmaf(mz, u.z, u.w)));
                const float r2 = D2 + num;  // |d - u|^2 = |d|^2 - 2 d.u + |u|^2
                const float rs = __builtin_amdgcn_rsqf(r2);
                const float delta = num * __builtin_amdgcn_rcpf(fmaf(r2, rs, D));
                const float rev = __builtin_amdgcn_fractf(fmaf(delta, kr, d.w));
                const float cs = __builtin_amdgcn_cosf(rev), sn = sgn * __builtin_amdgcn_sinf(rev);
                const float tx = rs * xv.x, ty = rs * xv.y;
                acc.x = fmaf(cs, tx, fmaf(-sn, ty, acc.x));  // (cos + i sn) (tx + i ty)
                acc.y = fmaf(cs, ty, fmaf(sn, tx, acc.y));
            }
        }
        if (c < it.w) slots[colin[c]] = acc;
    }
}

void launch_pfree_t(hipStream_t s, const BF &bf, const void *x, void *slots, bool conj) {
    if (bf.pf_n_items == 0) return;
    auto k = conj ? pfree_t_kernel<true> : pfree_t_kernel<false>;
    hipLaunchKernelGGL(k, dim3(unsigned(bf.pf_n_items)), dim3(256), 0, s, bf.pf_items.as<const int4>(),
                       bf.pf_col.as<const float4>(), bf.pf_in.as<const int32_t>(), bf.pf_row.as<const float4>(),
                       static_cast<const C<float> *>(x), bf.perm_test.as<const int32_t>(),
                       float(bf.pf_k * 0.15915494309189535), static_cast<C<float> *>(slots));
    BFG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Multi-RHS forward kernel: right-hand sides across lanes.  Lane r < nc of
// a wavefront is right-hand side r; a wavefront takes one task at a time
// (persistent, grid stride; the task index is wave-uniform, so task and
// contribution descriptors are scalar loads).  Skeleton vectors are
// interleaved: entry e of right-hand side r at slots[e * ns + r], so one
// input row of a block is one coalesced load across the lanes.  The block's
// entries are loaded spread over the lanes (64 entries per coalesced load)
// and broadcast to every lane with readlane as they are used; input rows
// are loaded KC at a time so their latencies overlap.  Output rows are
// accumulated in chunks of RM in registers; each factor entry feeds nc
// complex FMAs.
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ C<R> lane_bcast(C<R> v, int lane) {
    if constexpr (sizeof(R) == 4) {
        return C<R>{__builtin_bit_cast(R, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v.x), lane)),
                    __builtin_bit_cast(R, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v.y), lane))};
    } else {
        const int64_t bx = __builtin_bit_cast(int64_t, v.x), by = __builtin_bit_cast(int64_t, v.y);
        const int64_t rx = (int64_t(uint32_t(__builtin_amdgcn_readlane(int(bx), lane)))) |
                           (int64_t(__builtin_amdgcn_readlane(int(bx >> 32), lane)) << 32);
        const int64_t ry = (int64_t(uint32_t(__builtin_amdgcn_readlane(int(by), lane)))) |
                           (int64_t(__builtin_amdgcn_readlane(int(by >> 32), lane)) << 32);
        return C<R>{__builtin_bit_cast(R, rx), __builtin_bit_cast(R, ry)};
    }
}

template <typename R, int OUTK, int RM>
__global__ __launch_bounds__(256) void rhs_kernel(
    const MvTaskC *__restrict__ tasks, int32_t n_tasks, const MvContrib *__restrict__ contribs,
    const C<R> *__restrict__ F, C<R> *__restrict__ slots, int32_t ns, int64_t in_shift,
    C<R> *__restrict__ y, int64_t ldy, const int32_t *__restrict__ perm_out, C<R> alpha, C<R> beta,
    int fuse_beta, int nc) {
    constexpr int KC = 8;
    const int r = threadIdx.x & 63;
    const bool lane_ok = r < nc;
    const int rr = lane_ok ? r : 0;  // a valid column of the slot rows for idle lanes
    const int64_t n_waves = int64_t(gridDim.x) * 4;
    const C<R> *in = slots + in_shift * ns;
    // the wave index is uniform across the wavefront: say so, so that the
    // descriptors it selects are scalar loads
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    for (int64_t t = int64_t(blockIdx.x) * 4 + wave; t < n_tasks; t += n_waves) {
        const MvTaskC tk = tasks[t];
        const int out_len = tk.m;
        for (int i0 = 0; i0 < out_len; i0 += RM) {
            const int rows = min(RM, out_len - i0);
            C<R> acc[RM];
#pragma unroll
            for (int i = 0; i < RM; ++i) acc[i] = C<R>{R(0), R(0)};
            for (int q = tk.c_begin - 1; q < tk.c_end; ++q) {
                const MvContrib c = q < tk.c_begin ? MvContrib{tk.fofs, tk.m, tk.n, tk.in_ofs, 0} : contribs[q];
                if (c.fofs < 0) {  // identity
#pragma unroll
                    for (int i = 0; i < RM; ++i)
                        if (i < rows) acc[i] = cadd(acc[i], in[int64_t(c.in_ofs + i0 + i) * ns + rr]);
                    continue;
                }
                // entries (i0 + i, k), i < rows, column-major with stride m:
                // a window of 64 consecutive entries per lane-distributed load
                const C<R> *a = F + c.fofs;
                int64_t win = -1;
                C<R> aw = C<R>{R(0), R(0)};
                for (int k0 = 0; k0 < c.n; k0 += KC) {
                    C<R> xv[KC];
#pragma unroll
                    for (int kk = 0; kk < KC; ++kk) {
                        const int k = min(k0 + kk, c.n - 1);
                        xv[kk] = in[int64_t(c.in_ofs + k) * ns + rr];
                    }
#pragma unroll
                    for (int kk = 0; kk < KC; ++kk) {
                        const int k = k0 + kk;
                        if (k >= c.n) break;
#pragma unroll
                        for (int i = 0; i < RM; ++i) {
                            if (i >= rows) break;
                            const int64_t e = int64_t(k) * c.m + i0 + i;
                            if ((e >> 6) != win) {  // next 64-entry window of the block
                                win = e >> 6;
                                const int64_t ew = (win << 6) + r;
                                aw = F[c.fofs + min<int64_t>(ew, int64_t(c.m) * c.n - 1)];
                            }
                            cfma(acc[i], lane_bcast(aw, int(e & 63)), xv[kk]);
                        }
                    }
                }
                (void)a;
            }
            if (lane_ok) {
#pragma unroll
                for (int i = 0; i < RM; ++i) {
                    if (i >= rows) break;
                    if (OUTK == OUT_SLOT) {
                        slots[int64_t(tk.out_ofs + i0 + i) * ns + r] = acc[i];
                    } else {
                        C<R> *yp = y + int64_t(r) * ldy + perm_out[tk.out_ofs + i0 + i];
                        const C<R> old = *yp;
                        *yp = cadd(fuse_beta ? cmul(beta, old) : old, cmul(alpha, acc[i]));
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Multi-RHS forward kernel on the matrix cores (complex FP32).  Skeleton
// vectors are interleaved as for rhs_kernel (entry e of right-hand side r at
// slots[e * ns + r], ns = 16 NT).  One wavefront per tile of at most 64
// output rows: either rows [row0, row0 + rows) of one item, summed over its
// contributions, or whole items stacked on top of each other that read the
// same input entries (one contribution each, same in_ofs and n), so small
// blocks fill the rows of the product; a tile holds RB row blocks of 16
// (RB = 1..4, up to 64 rows), which share every input load.  Per step of 4 input entries the
// lanes load A[i][k] (lane i + 16 k: row i of the tile, entry k) and, per
// group g of 16 right-hand sides, B[k][j] (lane j + 16 k; column j of group
// g is right-hand side rhs_col(g, j): groups pair up so that a lane's two
// right-hand sides are adjacent, one 16-byte load), and issue
// v_mfma_f32_16x16x4_f32 four times per group (re*re, -im*im into the real
// part, re*im, im*re into the imaginary part); lane j + 16 q then holds
// rows 4q..4q+3 of column j.  Padding rows and entries load zeros; lanes of
// right-hand sides >= nc compute on whatever the slot buffer holds there and
// are never stored (a column of the product depends on its own column of
// B only).
// ---------------------------------------------------------------------------
typedef float bfg_v4f __attribute__((ext_vector_type(4)));

// Occupancy target per RHS group count and tile height: 16-row tiles take
// 6 / 5 / 4 waves per SIMD for <= 32 / 48 / 64 RHS, the most that fit the
// accumulators without more than a couple of spilled registers (32 RHS:
// 28.5 -> 27.2 ms on the 262,144-point sphere against the compiler's own
// choice of 4); taller tiles hold RB times the accumulators and get the
// occupancy that holds them (the 16-row target spilled 121-318 registers).
constexpr int mfma_waves(int NT, int RB) {
    return RB == 1 ? (NT <= 2 ? 6 : NT == 3 ? 5 : 4) : (NT * RB <= 3 ? 4 : 2);
}
// Right-hand side of group g in lane column li of the matrix-core kernel:
// groups pair up, lane li holding RHS 32 h + 2 li and 32 h + 2 li + 1 of pair
// h (one 16-byte load or store per pair, 256 contiguous bytes per 16
// lanes; skeleton-vector rows are 16-byte aligned, ns a multiple of 16
// entries); an odd last group holds RHS 32 h + li.
template <int NT>
__device__ __forceinline__ int rhs_col(int g, int li) {
    return (g < (NT & ~1)) ? 32 * (g >> 1) + 2 * li + (g & 1) : 32 * (NT >> 1) + li;
}
// The lane's NT right-hand sides of one entry row p (zero when !ok).
template <int NT>
__device__ __forceinline__ void rhs_load(C<float> *b, const C<float> *p, int li, bool ok) {
#pragma unroll
    for (int h = 0; h < NT / 2; ++h) {
        const float4 v = ok ? *reinterpret_cast<const float4 *>(p + 32 * h + 2 * li) : float4{0.f, 0.f, 0.f, 0.f};
        b[2 * h] = C<float>{v.x, v.y};
        b[2 * h + 1] = C<float>{v.z, v.w};
    }
    if (NT & 1) b[NT - 1] = ok ? p[32 * (NT >> 1) + li] : C<float>{0.f, 0.f};
}

template <int OUTK, int NT, int RB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(mfma_waves(NT, RB)))) void mfma_rhs_kernel(
    const int4 *__restrict__ tiles, int32_t n_tiles, const MvTaskC *__restrict__ mtasks,
    const MvContrib *__restrict__ contribs, const C<float> *__restrict__ F, C<float> *__restrict__ slots,
    int32_t ns, int64_t in_shift, C<float> *__restrict__ y, int64_t ldy, const int32_t *__restrict__ perm_out,
    C<float> alpha, C<float> beta, int fuse_beta, int nc) {
    const int l = threadIdx.x & 63;
    const int li = l & 15, lk = l >> 4;
    const int64_t n_waves = int64_t(gridDim.x) * 4;
    const C<float> *in = slots + in_shift * ns;
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    int64_t t = int64_t(blockIdx.x) * 4 + wave;
    int4 tl_next = t < n_tiles ? tiles[t] : int4{0, 0, 0, 0};
    for (; t < n_tiles; t += n_waves) {
        const int4 tl = tl_next;  // first stacked task, task count, first row, rows (<= 16 RB)
        if (t + n_waves < n_tiles) tl_next = tiles[t + n_waves];
        // the tile's tasks, one per lane (at most 64)
        int64_t qf = 0;
        int32_t qm = 0, qo = 0;
        if (l < tl.y) {
            qf = mtasks[tl.x + l].fofs;
            qm = mtasks[tl.x + l].m;
            qo = mtasks[tl.x + l].out_ofs;
        }
        const MvTaskC tk0 = mtasks[tl.x];
        // row 16 rb + li of the tile: its block and output entry
        int64_t rbase[RB];
        int32_t rld[RB], rout[RB];
        bool rok[RB];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            rbase[rb] = 0;
            rld[rb] = 1;
            rout[rb] = tk0.out_ofs + tl.z + 16 * rb + li;
            rok[rb] = 16 * rb + li < tl.w;
        }
        if (tl.y > 1) {
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) rok[rb] = false;
            int start = 0;
            for (int q = 0; q < tl.y; ++q) {
                const int32_t mq = __builtin_amdgcn_readlane(qm, q);
                const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint64_t(qf)), q));
                const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint64_t(qf) >> 32), q));
                const int64_t fq = int64_t((uint64_t(hi) << 32) | lo);
                const int32_t oq = __builtin_amdgcn_readlane(qo, q);
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) {
                    const int ri = 16 * rb + li;
                    if (ri >= start && ri < start + mq) {
                        rok[rb] = true;
                        rbase[rb] = fq + (ri - start);
                        rld[rb] = mq;
                        rout[rb] = oq + (ri - start);
                    }
                }
                start += mq;
            }
        }
        bfg_v4f accr[RB][NT], acci[RB][NT];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int g = 0; g < NT; ++g) {
                accr[rb][g] = bfg_v4f{0.f, 0.f, 0.f, 0.f};
                acci[rb][g] = bfg_v4f{0.f, 0.f, 0.f, 0.f};
            }
        // contributions; the first is inline in the task, the next one's
        // descriptor is loaded while the current one is multiplied
        int64_t cf = tk0.fofs;
        int32_t cm = tk0.m, cn = tk0.n, ci = tk0.in_ofs;
        for (int q = tk0.c_begin - 1; q < tk0.c_end; ++q) {
            int64_t nf = 0;
            int32_t nm = 0, nn = 0, ni = 0;
            if (q + 1 < tk0.c_end) {
                nf = contribs[q + 1].fofs;
                nm = contribs[q + 1].m;
                nn = contribs[q + 1].n;
                ni = contribs[q + 1].in_ofs;
            }
            if (cf < 0) {  // identity (single-task tiles only): add the input rows
#pragma unroll
                for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int r = 16 * rb + 4 * lk + v;
                        if (r < tl.w) {
#pragma unroll
                            for (int g = 0; g < NT; ++g) {
                                const C<float> xv = in[int64_t(ci + tl.z + r) * ns + rhs_col<NT>(g, li)];
                                accr[rb][g][v] += xv.x;
                                acci[rb][g][v] += xv.y;
                            }
                        }
                    }
            } else {
                const C<float> *ap[RB];
                int64_t ald[RB];
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) {
                    ap[rb] = F + (tl.y == 1 ? cf + tl.z + 16 * rb + li : rbase[rb]);
                    ald[rb] = tl.y == 1 ? cm : rld[rb];
                }
                const C<float> *xin = in + int64_t(ci) * ns;
                // step k0 multiplies entries k0..k0+3 while k0+4..k0+7 load
                C<float> a[RB], b[NT];
#pragma unroll
                for (int rb = 0; rb < RB; ++rb)
                    a[rb] = (rok[rb] && lk < cn) ? ap[rb][int64_t(lk) * ald[rb]] : C<float>{0.f, 0.f};
                rhs_load<NT>(b, xin + int64_t(lk) * ns, li, lk < cn);
                for (int k0 = 0; k0 < cn; k0 += 4) {
                    const int kn = k0 + 4 + lk;
                    C<float> an[RB], bn[NT];
#pragma unroll
                    for (int rb = 0; rb < RB; ++rb)
                        an[rb] = (rok[rb] && kn < cn) ? ap[rb][int64_t(kn) * ald[rb]] : C<float>{0.f, 0.f};
                    rhs_load<NT>(bn, xin + int64_t(kn) * ns, li, kn < cn);
#pragma unroll
                    for (int rb = 0; rb < RB; ++rb) {
                        const float nai = -a[rb].y;
#pragma unroll
                        for (int g = 0; g < NT; ++g)
                            accr[rb][g] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rb].x, b[g].x, accr[rb][g], 0, 0, 0);
#pragma unroll
                        for (int g = 0; g < NT; ++g)
                            acci[rb][g] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rb].x, b[g].y, acci[rb][g], 0, 0, 0);
#pragma unroll
                        for (int g = 0; g < NT; ++g)
                            accr[rb][g] = __builtin_amdgcn_mfma_f32_16x16x4f32(nai, b[g].y, accr[rb][g], 0, 0, 0);
#pragma unroll
                        for (int g = 0; g < NT; ++g)
                            acci[rb][g] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rb].y, b[g].x, acci[rb][g], 0, 0, 0);
                    }
#pragma unroll
                    for (int rb = 0; rb < RB; ++rb) a[rb] = an[rb];
#pragma unroll
                    for (int g = 0; g < NT; ++g) b[g] = bn[g];
                }
            }
            cf = nf;
            cm = nm;
            cn = nn;
            ci = ni;
        }
        // lane li + 16 lk holds rows 16 rb + 4 lk + v of right-hand sides
        // rhs_col(g, li)
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = 4 * lk + v;
                const int orow = __shfl(rout[rb], r, 64);
                if (__shfl(int(rok[rb]), r, 64) == 0) continue;
                if (OUTK == OUT_SLOT) {  // pairs of RHS: one 16-byte store each
#pragma unroll
                    for (int h = 0; h < NT / 2; ++h) {
                        const int j = 32 * h + 2 * li;
                        if (j + 1 < nc) {
                            *reinterpret_cast<float4 *>(slots + int64_t(orow) * ns + j) =
                                float4{accr[rb][2 * h][v], acci[rb][2 * h][v], accr[rb][2 * h + 1][v],
                                       acci[rb][2 * h + 1][v]};
                        } else if (j < nc) {
                            slots[int64_t(orow) * ns + j] = C<float>{accr[rb][2 * h][v], acci[rb][2 * h][v]};
                        }
                    }
                    if (NT & 1) {
                        const int j = 32 * (NT >> 1) + li;
                        if (j < nc) slots[int64_t(orow) * ns + j] = C<float>{accr[rb][NT - 1][v], acci[rb][NT - 1][v]};
                    }
                    continue;
                }
#pragma unroll
                for (int g = 0; g < NT; ++g) {
                    const int j = rhs_col<NT>(g, li);
                    if (j >= nc) continue;
                    const C<float> val = C<float>{accr[rb][g][v], acci[rb][g][v]};
                    if (OUTK == OUT_SLOT) {
                        slots[int64_t(orow) * ns + j] = val;
                    } else {
                        C<float> *yp = y + int64_t(j) * ldy + perm_out[orow];
                        const C<float> old = *yp;
                        *yp = cadd(fuse_beta ? cmul(beta, old) : old, cmul(alpha, val));
                    }
                }
            }
    }
}

void launch_mfma_rhs(hipStream_t s, const StageTable &t, const void *factors, void *slots, int32_t ns,
                     int64_t in_shift, void *y, int64_t ldy, const int32_t *perm_out, int nc, const double *alpha,
                     const double *beta, bool fuse, int n_cu) {
    if (t.n_mtiles == 0) return;
    BFG_REQUIRE(ns % 16 == 0 && ns >= 16 && ns <= 64 && nc >= 1 && nc <= ns,
                "matvec: matrix-core chunk of 16, 32, 48 or 64 right-hand sides");
    const C<float> al{float(alpha[0]), float(alpha[1])}, be{float(beta[0]), float(beta[1])};
    // tiles are stored by row blocks (1..4 of 16 rows): one launch per kind
    for (int rb = 1; rb <= 4; ++rb) {
        const int32_t t0 = t.mtile_ofs[rb - 1], nt = t.mtile_ofs[rb] - t0;
        if (nt <= 0) continue;
        const int64_t wgs = (int64_t(nt) + 3) / 4;
        const int64_t cap = int64_t(std::max(n_cu, 1)) * 16;
        dim3 grid(unsigned(std::max<int64_t>(1, std::min(wgs, cap))));
        const int4 *T = t.mtiles.as<const int4>() + t0;
#define BFG_MF(OUTK, NT, RBV)                                                                                  \
    hipLaunchKernelGGL((mfma_rhs_kernel<OUTK, NT, RBV>), grid, dim3(256), 0, s, T, nt, t.mtasks.as<const MvTaskC>(), \
                       t.contribs.as<const MvContrib>(), (const C<float> *)factors, (C<float> *)slots, ns, in_shift, \
                       (C<float> *)y, ldy, perm_out, al, be, int(fuse), nc)
#define BFG_MF_RB(OUTK, NT)            \
    switch (rb) {                      \
    case 1: BFG_MF(OUTK, NT, 1); break; \
    case 2: BFG_MF(OUTK, NT, 2); break; \
    case 3: BFG_MF(OUTK, NT, 3); break; \
    default: BFG_MF(OUTK, NT, 4); break; \
    }
#define BFG_MF_NT(OUTK)                 \
    switch (ns / 16) {                  \
    case 1: BFG_MF_RB(OUTK, 1); break;  \
    case 2: BFG_MF_RB(OUTK, 2); break;  \
    case 3: BFG_MF_RB(OUTK, 3); break;  \
    default: BFG_MF_RB(OUTK, 4); break; \
    }
        if (t.out_kind == OUT_SLOT) {
            BFG_MF_NT(OUT_SLOT);
        } else {
            BFG_MF_NT(OUT_SCATTER);
        }
#undef BFG_MF_NT
#undef BFG_MF_RB
#undef BFG_MF
        BFG_HIP(hipGetLastError());
    }
}

// ---------------------------------------------------------------------------
// Matrix-free last level on the matrix cores (forward, right-hand sides
// across lanes, complex FP32).  The last level's blocks are point-kernel
// entries between an observer leaf's points and the skeleton points of its
// columns, so, as in pfree_kernel, they are computed instead of read: one
// wavefront per observer leaf (largest first), all of the leaf's row blocks
// of 16 at once (at most 4: a leaf of more than 64 rows is taken 64 rows at
// a time).  Per step of 4 columns, lane i + 16 k computes
// the entry of row 16 rb + i and column k of every row block rb (leaf-local
// coordinates, as pfree_kernel) -- the A operand -- and loads, per group of
// 16 right-hand sides, the column's skeleton entries (lane j + 16 k) -- the
// B operand, read once for all row blocks; four v_mfma_f32_16x16x4_f32 per
// row block and group.  Column data and inputs of the next step are loaded
// while the current one is multiplied.  Rows past the leaf and columns past
// its end carry zero entries; right-hand sides >= nc are never stored.
// ---------------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NT <= 2 ? 4 : 2))) void pfree_mfma_kernel(
    const int4 *__restrict__ items, const int32_t *__restrict__ order, int32_t n_items,
    const float4 *__restrict__ cdat, const int32_t *__restrict__ colin, const float4 *__restrict__ rdat,
    const C<float> *__restrict__ slots, int32_t ns, float kr, C<float> *__restrict__ y, int64_t ldy,
    const int32_t *__restrict__ perm_out, C<float> alpha, C<float> beta, int fuse_beta, int nc) {
    constexpr float inv4pi = 0.07957747154594767f;
    const int l = threadIdx.x & 63;
    const int li = l & 15, lk = l >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    const int64_t n_waves = int64_t(gridDim.x) * 4;
    for (int64_t t = int64_t(blockIdx.x) * 4 + wave; t < n_items; t += n_waves)
    for (int r0 = 0; r0 < items[order[t]].y; r0 += 64) {  // rows in groups of 64 (leaves hold at most 64 as a rule)
        const int4 it = items[order[t]];  // first row (tree position), rows, columns [z, w)
        const int nrb = min(4, (it.y - r0 + 15) >> 4);  // row blocks of this group, 1..4
        float ux[4], uy[4], uz[4], uu[4];
        bool rok[4];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
            const int row = r0 + 16 * rb + li;
            rok[rb] = row < it.y;
            const float4 p = rdat[it.x + min(row, it.y - 1)];
            ux[rb] = p.x;
            uy[rb] = p.y;
            uz[rb] = p.z;
            uu[rb] = p.w;
        }
        bfg_v4f accr[4][NT], acci[4][NT];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int g = 0; g < NT; ++g) {
                accr[rb][g] = bfg_v4f{0.f, 0.f, 0.f, 0.f};
                acci[rb][g] = bfg_v4f{0.f, 0.f, 0.f, 0.f};
            }
        auto fetch = [&](int c0, float4 &d, C<float>(&b)[NT], bool &ok) {
            const int c = c0 + lk;
            ok = c < it.w;
            const int cc = min(c, it.w - 1);
            d = cdat[cc];
            rhs_load<NT>(b, slots + int64_t(colin[cc]) * ns, li, true);
        };
        float4 d;
        C<float> b[NT];
        bool cok;
        fetch(it.z, d, b, cok);
        for (int c0 = it.z; c0 < it.w; c0 += 4) {
            float4 dn = d;
            C<float> bn[NT];
            bool nok = false;
#pragma unroll
            for (int g = 0; g < NT; ++g) bn[g] = b[g];
            if (c0 + 4 < it.w) fetch(c0 + 4, dn, bn, nok);
            const float D2 = fmaf(d.x, d.x, fmaf(d.y, d.y, d.z * d.z));
            const float D = D2 * __builtin_amdgcn_rsqf(D2);
            const float mx = -2.f * d.x, my = -2.f * d.y, mz = -2.f * d.z;
#pragma unroll
            for (int rb = 0; rb < 4; ++rb) {
                if (rb >= nrb) break;
                const float num = fmaf(mx, ux[rb], fmaf(my, uy[rb], fmaf(mz, uz[rb], uu[rb])));
                const float r2 = D2 + num;  // |d - u|^2 = |d|^2 - 2 d.u + |u|^2
                const float rs = __builtin_amdgcn_rsqf(r2);
                const float delta = num * __builtin_amdgcn_rcpf(fmaf(r2, rs, D));
                const float rev = __builtin_amdgcn_fractf(fmaf(delta, kr, d.w));
                const float amp = cok && rok[rb] ? rs * inv4pi : 0.f;
                const float ar = amp * __builtin_amdgcn_cosf(rev), ai = -amp * __builtin_amdgcn_sinf(rev);
#pragma unroll
                for (int g = 0; g < NT; ++g)
                    accr[rb][g] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b[g].x, accr[rb][g], 0, 0, 0);
#pragma unroll
                for (int g = 0; g < NT; ++g)
                    acci[rb][g] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b[g].y, acci[rb][g], 0, 0, 0);
#pragma unroll
                for (int g = 0; g < NT; ++g)
                    accr[rb][g] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, b[g].y, accr[rb][g], 0, 0, 0);
#pragma unroll
                for (int g = 0; g < NT; ++g)
                    acci[rb][g] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, b[g].x, acci[rb][g], 0, 0, 0);
            }
            d = dn;
            cok = nok;
#pragma unroll
            for (int g = 0; g < NT; ++g) b[g] = bn[g];
        }
        // lane li + 16 lk holds rows 16 rb + 4 lk + v of right-hand side li (+ 16 g)
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
            if (rb >= nrb) break;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = r0 + 16 * rb + 4 * lk + v;
                if (row >= it.y) continue;
                const int64_t yr = perm_out[it.x + row];
#pragma unroll
                for (int g = 0; g < NT; ++g) {
                    const int j = rhs_col<NT>(g, li);
                    if (j >= nc) continue;
                    C<float> *yp = y + int64_t(j) * ldy + yr;
                    const C<float> old = *yp;
                    *yp = cadd(fuse_beta ? cmul(beta, old) : old,
                               cmul(alpha, C<float>{accr[rb][g][v], acci[rb][g][v]}));
                }
            }
        }
    }
}

// FP32 products emulated on the bf16 matrix cores ("bf16x3"): x = h + l
// with h = bf16(x) and l = bf16(x - h) (|x - h - l| <= 2^-16 |x|), and a
// product a b taken as a_h b_h + a_h b_l + a_l b_h (the dropped a_l b_l is
// below 2^-16 |a b|), accumulated in FP32 by v_mfma_f32_16x16x16_bf16
// (lane i + 16 q holds A[i][4q .. 4q+3] and B[4q .. 4q+3][i]; rows
// 4q .. 4q+3 of the product, as for the FP32 form).  A complex product
// takes 12 such instructions per 16 columns (6 into the real part, 6 into
// the imaginary part; -Im b is a negated bf16 pair of the shared B operand)
// against 16 FP32 16x16x4 instructions; each operand carries about 17 bits
// instead of 24.
typedef short bfg_v4s __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void bf16_split(float x, short &h, short &l) {
    const __bf16 hb = __bf16(x);
    const __bf16 lb = __bf16(x - float(hb));
    h = __builtin_bit_cast(short, hb);
    l = __builtin_bit_cast(short, lb);
}

// The matrix-free last level on the bf16 matrix cores: pfree_mfma_kernel
// with steps of 16 columns (lane i + 16 k computes the entries of columns
// 4 k .. 4 k + 3 of row i of each row block, and splits them and the
// columns' skeleton entries into bf16 halves).
template <int NT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NT <= 2 ? 3 : 2))) void pfree_mfma3_kernel(
    const int4 *__restrict__ items, const int32_t *__restrict__ order, int32_t n_items,
    const float4 *__restrict__ cdat, const int32_t *__restrict__ colin, const float4 *__restrict__ rdat,
    const C<float> *__restrict__ slots, int32_t ns, float kr, C<float> *__restrict__ y, int64_t ldy,
    const int32_t *__restrict__ perm_out, C<float> alpha, C<float> beta, int fuse_beta, int nc) {
    constexpr float inv4pi = 0.07957747154594767f;
    const int l = threadIdx.x & 63;
    const int li = l & 15, lk = l >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    const int64_t n_waves = int64_t(gridDim.x) * 4;
    for (int64_t t = int64_t(blockIdx.x) * 4 + wave; t < n_items; t += n_waves)
    for (int r0 = 0; r0 < items[order[t]].y; r0 += 64) {
        const int4 it = items[order[t]];
        const int nrb = min(4, (it.y - r0 + 15) >> 4);
        float ux[4], uy[4], uz[4], uu[4];
        bool rok[4];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
            const int row = r0 + 16 * rb + li;
            rok[rb] = row < it.y;
            const float4 p = rdat[it.x + min(row, it.y - 1)];
            ux[rb] = p.x;
            uy[rb] = p.y;
            uz[rb] = p.z;
            uu[rb] = p.w;
        }
        bfg_v4f accr[4][NT], acci[4][NT];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int g = 0; g < NT; ++g) {
                accr[rb][g] = bfg_v4f{0.f, 0.f, 0.f, 0.f};
                acci[rb][g] = bfg_v4f{0.f, 0.f, 0.f, 0.f};
            }
        for (int c0 = it.z; c0 < it.w; c0 += 16) {
            float4 dd[4];
            bool okc[4];
            // B halves per group: Re, Im, -Im (hi and lo)
            bfg_v4s brh[NT], brl[NT], bih[NT], bil[NT], nbih[NT], nbil[NT];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int c = c0 + 4 * lk + u;
                okc[u] = c < it.w;
                const int cc = min(c, it.w - 1);
                dd[u] = cdat[cc];
                C<float> vv[NT];
                rhs_load<NT>(vv, slots + int64_t(colin[cc]) * ns, li, true);
#pragma unroll
                for (int g = 0; g < NT; ++g) {
                    const C<float> v = vv[g];
                    short h, lo;
                    bf16_split(v.x, h, lo);
                    brh[g][u] = h;
                    brl[g][u] = lo;
                    bf16_split(v.y, h, lo);
                    bih[g][u] = h;
                    bil[g][u] = lo;
                    nbih[g][u] = short(h ^ short(0x8000));
                    nbil[g][u] = short(lo ^ short(0x8000));
                }
            }
#pragma unroll
            for (int rb = 0; rb < 4; ++rb) {
                if (rb >= nrb) break;
                bfg_v4s arh, arl, aih, ail;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const float4 d = dd[u];
                    const float D2 = fmaf(d.x, d.x, fmaf(d.y, d.y, d.z * d.z));
                    const float D = D2 * __builtin_amdgcn_rsqf(D2);
                    const float num =
                        fmaf(-2.f * d.x, ux[rb], fmaf(-2.f * d.y, uy[rb], fmaf(-2.f * d.z, uz[rb], uu[rb])));
                    const float r2 = D2 + num;  // |d - u|^2 = |d|^2 - 2 d.u + |u|^2
                    const float rs = __builtin_amdgcn_rsqf(r2);
                    const float delta = num * __builtin_amdgcn_rcpf(fmaf(r2, rs, D));
                    const float rev = __builtin_amdgcn_fractf(fmaf(delta, kr, d.w));
                    const float amp = okc[u] && rok[rb] ? rs * inv4pi : 0.f;
                    short h, lo;
                    bf16_split(amp * __builtin_amdgcn_cosf(rev), h, lo);
                    arh[u] = h;
                    arl[u] = lo;
                    bf16_split(-amp * __builtin_amdgcn_sinf(rev), h, lo);
                    aih[u] = h;
                    ail[u] = lo;
                }
#pragma unroll
                for (int g = 0; g < NT; ++g) {
                    bfg_v4f r = accr[rb][g], i = acci[rb][g];
                    r = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh, brh[g], r, 0, 0, 0);
                    i = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh, bih[g], i, 0, 0, 0);
                    r = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh, brl[g], r, 0, 0, 0);
                    i = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh, bil[g], i, 0, 0, 0);
                    r = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arl, brh[g], r, 0, 0, 0);
                    i = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arl, bih[g], i, 0, 0, 0);
                    r = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aih, nbih[g], r, 0, 0, 0);
                    i = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aih, brh[g], i, 0, 0, 0);
                    r = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aih, nbil[g], r, 0, 0, 0);
                    i = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aih, brl[g], i, 0, 0, 0);
                    r = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ail, nbih[g], r, 0, 0, 0);
                    i = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ail, brh[g], i, 0, 0, 0);
                    accr[rb][g] = r;
                    acci[rb][g] = i;
                }
            }
        }
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
            if (rb >= nrb) break;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = r0 + 16 * rb + 4 * lk + v;
                if (row >= it.y) continue;
                const int64_t yr = perm_out[it.x + row];
#pragma unroll
                for (int g = 0; g < NT; ++g) {
                    const int j = rhs_col<NT>(g, li);
                    if (j >= nc) continue;
                    C<float> *yp = y + int64_t(j) * ldy + yr;
                    const C<float> old = *yp;
                    *yp = cadd(fuse_beta ? cmul(beta, old) : old,
                               cmul(alpha, C<float>{accr[rb][g][v], acci[rb][g][v]}));
                }
            }
        }
    }
}

void launch_pfree_mfma(hipStream_t s, const BF &bf, const void *slots, int32_t ns, void *y, int64_t ldy, int nc,
                       const double *alpha, const double *beta, bool fuse, int n_cu) {
    if (bf.pf_n_items == 0) return;
    BFG_REQUIRE(bf.pf_order_built, "matvec: matrix-free matrix-core tables missing");
    BFG_REQUIRE(ns % 16 == 0 && ns >= 16 && ns <= 64 && nc >= 1 && nc <= ns,
                "matvec: matrix-core chunk of 16, 32, 48 or 64 right-hand sides");
    const C<float> al{float(alpha[0]), float(alpha[1])}, be{float(beta[0]), float(beta[1])};
    // persistent: as many workgroups as are resident (one wave per SIMD
    // each; 5 / 4 / 3 / 2 waves per SIMD for 16 / 32 / 48 / 64 RHS), the
    // items (largest first) dealt round robin
    const int64_t wgs = (int64_t(bf.pf_n_items) + 3) / 4;
    const int per_cu = bf.mfma_emul ? (ns <= 32 ? 3 : 2) : ns <= 16 ? 5 : ns <= 32 ? 4 : ns <= 48 ? 3 : 2;
    const int64_t cap = int64_t(std::max(n_cu, 1)) * per_cu;
    dim3 grid(unsigned(std::max<int64_t>(1, std::min(wgs, cap))));
    const float kr = float(bf.pf_k * 0.15915494309189535);
#define BFG_PFM(NT)                                                                                            \
    hipLaunchKernelGGL((bf.mfma_emul ? pfree_mfma3_kernel<NT> : pfree_mfma_kernel<NT>), grid, dim3(256), 0, s,     \
                       bf.pf_items.as<const int4>(),                                                             \
                       bf.pf_order.as<const int32_t>(), bf.pf_n_items, bf.pf_col.as<const float4>(),              \
                       bf.pf_in.as<const int32_t>(), bf.pf_row.as<const float4>(),                               \
                       static_cast<const C<float> *>(slots), ns, kr, static_cast<C<float> *>(y), ldy,             \
                       bf.perm_test.as<const int32_t>(), al, be, fuse ? 1 : 0, nc)
    switch (ns / 16) {
    case 1: BFG_PFM(1); break;
    case 2: BFG_PFM(2); break;
    case 3: BFG_PFM(3); break;
    default: BFG_PFM(4); break;
    }
#undef BFG_PFM
    BFG_HIP(hipGetLastError());
}

// dst[v * ns + r] = x[r][perm[v]] for v in [lo, hi), r < nc: inputs of the
// first level in tree order, right-hand sides interleaved.
template <typename R>
__global__ void gather_rhs_kernel(const C<R> *__restrict__ x, int64_t ldx, const int32_t *__restrict__ perm,
                                  int32_t lo, int32_t n, C<R> *__restrict__ dst, int32_t ns, int nc) {
    const int r = threadIdx.x & 63;
    const int64_t v = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (v >= n || r >= nc) return;
    dst[(lo + v) * ns + r] = x[int64_t(r) * ldx + perm[lo + v]];
}

void launch_rhs(hipStream_t s, int prec, const StageTable &t, const void *factors, void *slots, int32_t ns,
                int64_t in_shift, void *y, int64_t ldy, const int32_t *perm_out, int nc, const double *alpha,
                const double *beta, bool fuse, int n_cu) {
    if (t.n_items == 0) return;
    BFG_REQUIRE(nc >= 1 && nc <= 64 && nc <= ns, "matvec: right-hand-side chunk of 1..64");
    const int64_t wgs = (int64_t(t.n_items) + 3) / 4;
    const int64_t cap = int64_t(std::max(n_cu, 1)) * 16;
    dim3 grid(unsigned(std::max<int64_t>(1, std::min(wgs, cap))));
    auto T = t.tasks_c.as<const MvTaskC>();
    auto P = t.contribs.as<const MvContrib>();
#define BFG_RHS(RT, OUTK)                                                                                     \
    hipLaunchKernelGGL((rhs_kernel<RT, OUTK, 8>), grid, dim3(256), 0, s, T, t.n_items, P, (const C<RT> *)factors, \
                       (C<RT> *)slots, ns, in_shift, (C<RT> *)y, ldy, perm_out, C<RT>{RT(alpha[0]), RT(alpha[1])}, \
                       C<RT>{RT(beta[0]), RT(beta[1])}, int(fuse), nc)
    if (prec == BFGPU_C128) {
        if (t.out_kind == OUT_SLOT) BFG_RHS(double, OUT_SLOT);
        else BFG_RHS(double, OUT_SCATTER);
    } else {
        if (t.out_kind == OUT_SLOT) BFG_RHS(float, OUT_SLOT);
        else BFG_RHS(float, OUT_SCATTER);
    }
#undef BFG_RHS
    BFG_HIP(hipGetLastError());
}

void launch_gather_rhs(hipStream_t s, int prec, const void *x, int64_t ldx, const int32_t *perm, int32_t lo,
                       int32_t hi, void *dst, int32_t ns, int nc) {
    const int32_t n = hi - lo;
    if (n <= 0 || nc <= 0) return;
    dim3 grid(unsigned((n + 3) / 4));
    if (prec == BFGPU_C128)
        hipLaunchKernelGGL(gather_rhs_kernel<double>, grid, dim3(256), 0, s, (const C<double> *)x, ldx, perm, lo, n,
                           (C<double> *)dst, ns, nc);
    else
        hipLaunchKernelGGL(gather_rhs_kernel<float>, grid, dim3(256), 0, s, (const C<float> *)x, ldx, perm, lo, n,
                           (C<float> *)dst, ns, nc);
    BFG_HIP(hipGetLastError());
}

// dst[r][lo + v] = x[r][perm[lo + v]] for v < n: the inputs of an IN_GATHER
// stage in tree order, so the tile kernel reads them contiguously.
template <typename R>
__global__ void gather_kernel(const C<R> *__restrict__ x, int64_t ldx, const int32_t *__restrict__ perm,
                              int32_t lo, int32_t n, C<R> *__restrict__ dst, int64_t ld_dst) {
    const int32_t v = int32_t(blockIdx.x) * 256 + int32_t(threadIdx.x);
    if (v >= n) return;
    const int64_t r = blockIdx.y;
    dst[r * ld_dst + lo + v] = x[r * ldx + perm[lo + v]];
}

void launch_gather(hipStream_t s, int prec, const void *x, int64_t ldx, const int32_t *perm,
                   int32_t lo, int32_t hi, void *dst, int64_t ld_dst, int64_t nrhs) {
    const int32_t n = hi - lo;
    if (n <= 0 || nrhs <= 0) return;
    dim3 grid(unsigned((n + 255) / 256), unsigned(nrhs));
    if (prec == BFGPU_C128)
        hipLaunchKernelGGL(gather_kernel<double>, grid, dim3(256), 0, s, (const C<double> *)x, ldx, perm, lo, n,
                           (C<double> *)dst, ld_dst);
    else
        hipLaunchKernelGGL(gather_kernel<float>, grid, dim3(256), 0, s, (const C<float> *)x, ldx, perm, lo, n,
                           (C<float> *)dst, ld_dst);
    BFG_HIP(hipGetLastError());
}

template <typename R, int OP, int NR, int QW>
static void launch_tile_t(hipStream_t s, const StageTable &t, const void *factors, void *slots,
                          int64_t slot_len, int64_t in_shift, void *y, int64_t ldy, const int32_t *perm_out,
                          int64_t nrhs, const double *alpha, const double *beta, bool fuse, int n_cu, int conj) {
    C<R> a = {R(alpha[0]), R(alpha[1])}, b = {R(beta[0]), R(beta[1])};
    const int lpt = t.tile.rp_shift + t.tile.g_shift;
    BFG_REQUIRE(tile_ok(t.tile), "matvec: bad tile shape");
    const int64_t tpw = 64 >> lpt;
    const int64_t wgs = ((int64_t(t.n_items) + tpw - 1) / tpw + 3) / 4;
    const int64_t cap = int64_t(std::max(n_cu, 1)) * std::max<int64_t>(t.tile.wg_per_cu, 1);
    dim3 grid(unsigned(std::max<int64_t>(1, std::min(wgs, cap))), unsigned((nrhs + NR - 1) / NR));
    auto F = (const C<R> *)factors;
    auto S = (C<R> *)slots;
    auto Y = (C<R> *)y;
    auto T = t.tasks_c.as<const MvTaskC>();
    auto P = t.contribs.as<const MvContrib>();
#define BFG_TILE(OUTK, CS)                                                                              \
    hipLaunchKernelGGL((tile_kernel<R, OP, OUTK, NR, QW, CS>), grid, dim3(256), 0, s, T, t.n_items, P, F, S, \
                       slot_len, in_shift, Y, ldy, perm_out, a, b, int(fuse), int(nrhs), int(t.tile.rp_shift),  \
                       int(t.tile.g_shift), conj, int(t.tile.xcd))
    if (t.out_kind == OUT_SLOT) {
        if (t.tile.csplit) BFG_TILE(OUT_SLOT, true);
        else BFG_TILE(OUT_SLOT, false);
    } else {
        if (t.tile.csplit) BFG_TILE(OUT_SCATTER, true);
        else BFG_TILE(OUT_SCATTER, false);
    }
#undef BFG_TILE
    BFG_HIP(hipGetLastError());
}

template <typename R, int NR, int QW>
static void launch_dense_t(hipStream_t s, const StageTable &t, const void *factors, void *slots,
                           int64_t slot_len, int64_t in_shift, void *y, int64_t ldy, const int32_t *perm_out,
                           int64_t nrhs, const double *alpha, const double *beta, bool fuse, int n_cu,
                           int tab, int conj) {
    C<R> a = {R(alpha[0]), R(alpha[1])}, b = {R(beta[0]), R(beta[1])};
    const int lpt = t.tile.rp_shift;
    BFG_REQUIRE(tile_ok(t.tile) && t.tile.g_shift == 0, "matvec: bad dense shape");
    const int32_t n_tasks = tab == DENSE_REST ? t.n_srest : t.n_items;
    if (n_tasks == 0) return;
    const int64_t tpw = 64 >> lpt;
    const int64_t wgs = ((int64_t(n_tasks) + tpw - 1) / tpw + 3) / 4;
    const int64_t cap = int64_t(std::max(n_cu, 1)) * std::max<int64_t>(t.tile.wg_per_cu, 1);
    dim3 grid(unsigned(std::max<int64_t>(1, std::min(wgs, cap))), unsigned((nrhs + NR - 1) / NR));
    auto F = (const C<R> *)factors;
    auto S = (C<R> *)slots;
    auto Y = (C<R> *)y;
    auto T = (tab == DENSE_SWAPPED ? t.tasks_sw : tab == DENSE_REST ? t.s_rest : t.tasks_c).as<const MvTaskC>();
    auto P = (tab == DENSE_SWAPPED || (tab == DENSE_REST && t.stream_copy) ? t.contribs_sw : t.contribs)
                 .as<const MvContrib>();
    BFG_REQUIRE(T != nullptr && P != nullptr, "matvec: dense tables missing");
#define BFG_DENSE(OUTK, CS)                                                                              \
    hipLaunchKernelGGL((dense_kernel<R, OUTK, NR, QW, CS>), grid, dim3(256), 0, s, T, n_tasks, P, F, S,  \
                       slot_len, in_shift, Y, ldy, perm_out, a, b, int(fuse), int(nrhs), lpt, int(t.tile.xcd), conj)
    if (t.out_kind == OUT_SLOT) {
        if (t.tile.csplit) BFG_DENSE(OUT_SLOT, true);
        else BFG_DENSE(OUT_SLOT, false);
    } else {
        if (t.tile.csplit) BFG_DENSE(OUT_SCATTER, true);
        else BFG_DENSE(OUT_SCATTER, false);
    }
#undef BFG_DENSE
    BFG_HIP(hipGetLastError());
}

template <typename R>
static void launch_dense_nr(hipStream_t s, const StageTable &t, const void *factors, void *slots,
                            int64_t slot_len, int64_t in_shift, void *y, int64_t ldy, const int32_t *perm_out,
                            int64_t nrhs, const double *alpha, const double *beta, bool fuse, int n_cu,
                            int sw, int conj) {
    if (nrhs == 1) {
        if (t.tile.qw4) launch_dense_t<R, 1, 4>(s, t, factors, slots, slot_len, in_shift, y, ldy, perm_out, nrhs, alpha, beta, fuse, n_cu, sw, conj);
        else launch_dense_t<R, 1, 8>(s, t, factors, slots, slot_len, in_shift, y, ldy, perm_out, nrhs, alpha, beta, fuse, n_cu, sw, conj);
    } else if (nrhs == 2) {
        launch_dense_t<R, 2, 4>(s, t, factors, slots, slot_len, in_shift, y, ldy, perm_out, nrhs, alpha, beta, fuse, n_cu, sw, conj);
    } else if (nrhs <= 4) {
        launch_dense_t<R, 4, 2>(s, t, factors, slots, slot_len, in_shift, y, ldy, perm_out, nrhs, alpha, beta, fuse, n_cu, sw, conj);
    } else {
        launch_dense_t<R, 8, 2>(s, t, factors, slots, slot_len, in_shift, y, ldy, perm_out, nrhs, alpha, beta, fuse, n_cu, sw, conj);
    }
}

template <typename R, int NR, bool XL>
static void launch_denset_t(hipStream_t s, const StageTable &t, const void *factors, void *slots,
                            int64_t slot_len, int64_t in_shift, void *y, int64_t ldy, const int32_t *perm_out,
                            int64_t nrhs, const double *alpha, const double *beta, bool fuse, int n_cu, int conj) {
    C<R> a = {R(alpha[0]), R(alpha[1])}, b = {R(beta[0]), R(beta[1])};
    const int lpt = t.tile.rp_shift;
    BFG_REQUIRE(tile_ok(t.tile) && t.tile.g_shift == 0, "matvec: bad dense shape");
    const int64_t tpw = 64 >> lpt;
    const int64_t wgs = ((int64_t(t.n_items) + tpw - 1) / tpw + 3) / 4;
    const int64_t cap = int64_t(std::max(n_cu, 1)) * std::max<int64_t>(t.tile.wg_per_cu, 1);
    dim3 grid(unsigned(std::max<int64_t>(1, std::min(wgs, cap))), unsigned((nrhs + NR - 1) / NR));
    auto T = t.tasks_c.as<const MvTaskC>();
    auto P = t.contribs.as<const MvContrib>();
#define BFG_DENSET(OUTK)                                                                                    \
    hipLaunchKernelGGL((denset_kernel<R, OUTK, NR, XL>), grid, dim3(256), 0, s, T, t.n_items, P, (const C<R> *)factors, \
                       (C<R> *)slots, slot_len, in_shift, (C<R> *)y, ldy, perm_out, a, b, int(fuse), int(nrhs), lpt,  \
                       int(t.tile.xcd), conj)
    if (t.out_kind == OUT_SLOT) BFG_DENSET(OUT_SLOT);
    else BFG_DENSET(OUT_SCATTER);
#undef BFG_DENSET
    BFG_HIP(hipGetLastError());
}

void launch_denset(hipStream_t s, int prec, int op, const StageTable &t, const void *factors, void *slots,
                   int64_t slot_len, int64_t in_shift, void *y, int64_t ldy, const int32_t *perm_out,
                   int64_t nrhs, const double *alpha, const double *beta, bool fuse_beta, int n_cu) {
    if (t.n_items == 0 || nrhs == 0) return;
    BFG_REQUIRE(nrhs < 65536 * 8, "matvec: nrhs too large");
    const int conj = op == BFGPU_OP_C ? 1 : 0;
#define BFG_DT(RT, NRv)                                                                                    \
    do {                                                                                                   \
        if (t.tile.csplit)                                                                                 \
            launch_denset_t<RT, NRv, true>(s, t, factors, slots, slot_len, in_shift, y, ldy, perm_out, nrhs, \
                                           alpha, beta, fuse_beta, n_cu, conj);                           \
        else                                                                                               \
            launch_denset_t<RT, NRv, false>(s, t, factors, slots, slot_len, in_shift, y, ldy, perm_out, nrhs, \
                                            alpha, beta, fuse_beta, n_cu, conj);                          \
    } while (0)
    if (prec == BFGPU_C128) {
        if (nrhs == 1) BFG_DT(double, 1);
        else if (nrhs == 2) BFG_DT(double, 2);
        else if (nrhs <= 4) BFG_DT(double, 4);
        else BFG_DT(double, 8);
    } else {
        if (nrhs == 1) BFG_DT(float, 1);
        else if (nrhs == 2) BFG_DT(float, 2);
        else if (nrhs <= 4) BFG_DT(float, 4);
        else BFG_DT(float, 8);
    }
#undef BFG_DT
}

template <typename R, int NLD, bool CONJ>
static void launch_stream_t(hipStream_t s, const StageTable &t, const void *factors, void *slots, int64_t in_shift,
                            int n_cu) {
    const int64_t wgs = (int64_t(t.n_schunks) + 1) / 2;
    const int64_t cap = int64_t(std::max(n_cu, 1)) * std::max<int64_t>(t.tile.wg_per_cu, 1);
    dim3 grid(unsigned(std::max<int64_t>(1, std::min(wgs, cap))));
    hipLaunchKernelGGL((stream_kernel<R, NLD, CONJ>), grid, dim3(128), 0, s, t.s_chunks.as<const StreamChunk>(),
                       t.n_schunks, t.s_tasks.as<const StreamTask>(), t.s_runs.as<const StreamRun>(),
                       (const C<R> *)factors, (C<R> *)slots, in_shift, int(t.tile.xcd));
    BFG_HIP(hipGetLastError());
}

template <typename R>
static void launch_stream_r(hipStream_t s, const StageTable &t, const void *factors, void *slots, int64_t in_shift,
                            int n_cu, bool conj) {
    if (t.stream_nld == 4) {
        if (conj) launch_stream_t<R, 4, true>(s, t, factors, slots, in_shift, n_cu);
        else launch_stream_t<R, 4, false>(s, t, factors, slots, in_shift, n_cu);
    } else {
        if (conj) launch_stream_t<R, 8, true>(s, t, factors, slots, in_shift, n_cu);
        else launch_stream_t<R, 8, false>(s, t, factors, slots, in_shift, n_cu);
    }
}

void launch_stream(hipStream_t s, int prec, const StageTable &t, const void *factors, void *slots,
                   int64_t in_shift, int n_cu, bool conj) {
    BFG_REQUIRE(t.stream_built && (t.stream_nld == 4 || t.stream_nld == 8), "matvec: streaming tables missing");
    if (t.n_schunks == 0) return;
    if (prec == BFGPU_C128) launch_stream_r<double>(s, t, factors, slots, in_shift, n_cu, conj);
    else launch_stream_r<float>(s, t, factors, slots, in_shift, n_cu, conj);
}

void launch_dense(hipStream_t s, int prec, const StageTable &t, const void *factors, void *slots,
                  int64_t slot_len, int64_t in_shift, void *y, int64_t ldy, const int32_t *perm_out,
                  int64_t nrhs, const double *alpha, const double *beta, bool fuse_beta, int n_cu,
                  int tab, int conj) {
    if ((tab == DENSE_REST ? t.n_srest : t.n_items) == 0 || nrhs == 0) return;
    BFG_REQUIRE(nrhs < 65536 * 8, "matvec: nrhs too large");
    if (prec == BFGPU_C128)
        launch_dense_nr<double>(s, t, factors, slots, slot_len, in_shift, y, ldy, perm_out, nrhs, alpha, beta,
                                fuse_beta, n_cu, tab, conj);
    else
        launch_dense_nr<float>(s, t, factors, slots, slot_len, in_shift, y, ldy, perm_out, nrhs, alpha, beta,
                               fuse_beta, n_cu, tab, conj);
}

// ---------------------------------------------------------------------------
// Skeletons of a batch: a thread per compression maps its pivot prefix to
// column ids.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void skeleton_gather_kernel(const QrTask *__restrict__ qr,
                                                              const FillTask *__restrict__ fill,
                                                              const int32_t *__restrict__ colpool,
                                                              const int32_t *__restrict__ piv,
                                                              const int32_t *__restrict__ skofs, int32_t n,
                                                              int32_t *__restrict__ out) {
    for (int32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const int32_t o = skofs[q], cnt = skofs[q + 1] - o;
        const int32_t *p = piv + qr[q].pofs;
        const int32_t *cols = colpool + fill[q].cofs;
        for (int32_t i = 0; i < cnt; ++i) out[int64_t(o) + i] = cols[p[i]];
    }
}

void launch_skeleton_gather(hipStream_t s, const QrTask *qr, const FillTask *fill, const int32_t *colpool,
                            const int32_t *piv, const int32_t *skofs, int32_t n, int32_t *out) {
    if (n <= 0) return;
    const unsigned grid = unsigned(std::min<int64_t>((int64_t(n) + 255) / 256, 8192));
    hipLaunchKernelGGL(skeleton_gather_kernel, dim3(grid), dim3(256), 0, s, qr, fill, colpool, piv, skofs, n, out);
    BFG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Launch-table tasks from items and contributions: task i is item i with its
// first contribution inline (the host keeps the same table for its own
// passes; only items and contributions cross the bus).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void build_tasks_kernel(const MvItem *__restrict__ items, int32_t n_items,
                                                          const MvContrib *__restrict__ con, int32_t n_con,
                                                          MvTask *__restrict__ tasks, MvTaskC *__restrict__ tc) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_items; i += gridDim.x * blockDim.x) {
        const MvItem it = items[i];
        MvContrib c0{-1, 0, 0, 0, 0};
        if (it.c_begin >= 0 && it.c_begin < n_con) c0 = con[it.c_begin];
        tasks[i] = MvTask{c0.fofs, c0.m, c0.n, c0.in_ofs, it.out_ofs, it.out_len, it.c_begin + 1, it.c_end, -1, 0, -1, 0};
        tc[i] = MvTaskC{c0.fofs, c0.in_ofs, it.out_ofs, c0.m, c0.n, it.c_begin + 1, it.c_end};
    }
}

void launch_build_tasks(hipStream_t s, const MvItem *items, int32_t n_items, const MvContrib *con, int32_t n_con,
                        MvTask *tasks, MvTaskC *tc) {
    if (n_items <= 0) return;
    const unsigned grid = unsigned(std::min<int64_t>((int64_t(n_items) + 255) / 256, 4096));
    hipLaunchKernelGGL(build_tasks_kernel, dim3(grid), dim3(256), 0, s, items, n_items, con, n_con, tasks, tc);
    BFG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Block transposition for the transposed copy: a workgroup per block, entries
// read in column-major order (coalesced) and written transposed.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void transpose_blocks_kernel(const TBlock *__restrict__ blocks, int64_t n_blocks,
                                                               const C<R> *__restrict__ src, C<R> *__restrict__ dst) {
    for (int64_t q = blockIdx.x; q < n_blocks; q += gridDim.x) {
        const TBlock b = blocks[q];
        const int m = b.m, n = b.n;
        const int64_t total = int64_t(m) * n;
        for (int64_t e = threadIdx.x; e < total; e += blockDim.x) {
            const int k = int(e / m), i = int(e - int64_t(k) * m);
            dst[b.dofs + int64_t(i) * n + k] = src[b.sofs + e];
        }
    }
}

void launch_transpose_blocks(hipStream_t s, int prec, const TBlock *blocks, int64_t n_blocks, const void *src,
                             void *dst) {
    if (n_blocks <= 0) return;
    dim3 grid(unsigned(std::min<int64_t>(n_blocks, int64_t(1) << 20)));
    if (prec == BFGPU_C128)
        hipLaunchKernelGGL(transpose_blocks_kernel<double>, grid, dim3(256), 0, s, blocks, n_blocks,
                           (const C<double> *)src, (C<double> *)dst);
    else
        hipLaunchKernelGGL(transpose_blocks_kernel<float>, grid, dim3(256), 0, s, blocks, n_blocks,
                           (const C<float> *)src, (C<float> *)dst);
    BFG_HIP(hipGetLastError());
}

template <typename R, int OP>
static void launch_tile_nr(hipStream_t s, const StageTable &t, const void *factors, void *slots,
                           int64_t slot_len, int64_t in_shift, void *y, int64_t ldy, const int32_t *perm_out,
                           int64_t nrhs, const double *alpha, const double *beta, bool fuse, int n_cu, int conj) {
    if (nrhs == 1) launch_tile_t<R, OP, 1, 8>(s, t, factors, slots, slot_len, in_shift, y, ldy, perm_out, nrhs, alpha, beta, fuse, n_cu, conj);
    else if (nrhs == 2) launch_tile_t<R, OP, 2, 8>(s, t, factors, slots, slot_len, in_shift, y, ldy, perm_out, nrhs, alpha, beta, fuse, n_cu, conj);
    else if (nrhs <= 4) launch_tile_t<R, OP, 4, 4>(s, t, factors, slots, slot_len, in_shift, y, ldy, perm_out, nrhs, alpha, beta, fuse, n_cu, conj);
    else launch_tile_t<R, OP, 8, 4>(s, t, factors, slots, slot_len, in_shift, y, ldy, perm_out, nrhs, alpha, beta, fuse, n_cu, conj);
}

void launch_tile(hipStream_t s, int prec, int op, const StageTable &t, const void *factors, void *slots,
                 int64_t slot_len, int64_t in_shift, void *y, int64_t ldy, const int32_t *perm_out,
                 int64_t nrhs, const double *alpha, const double *beta, bool fuse_beta, int n_cu) {
    if (t.n_items == 0 || nrhs == 0) return;
    BFG_REQUIRE(nrhs < 65536 * 8, "matvec: nrhs too large");
    const int conj = op == BFGPU_OP_C ? 1 : 0;
    if (prec == BFGPU_C128) {
        if (op == BFGPU_OP_N) launch_tile_nr<double, 0>(s, t, factors, slots, slot_len, in_shift, y, ldy, perm_out, nrhs, alpha, beta, fuse_beta, n_cu, 0);
        else launch_tile_nr<double, 1>(s, t, factors, slots, slot_len, in_shift, y, ldy, perm_out, nrhs, alpha, beta, fuse_beta, n_cu, conj);
    } else {
        if (op == BFGPU_OP_N) launch_tile_nr<float, 0>(s, t, factors, slots, slot_len, in_shift, y, ldy, perm_out, nrhs, alpha, beta, fuse_beta, n_cu, 0);
        else launch_tile_nr<float, 1>(s, t, factors, slots, slot_len, in_shift, y, ldy, perm_out, nrhs, alpha, beta, fuse_beta, n_cu, conj);
    }
}

template <typename R>
static void launch_fwd(hipStream_t s, const StageTable &t, const void *factors, const void *x,
                       int64_t ldx, void *y, int64_t ldy, void *slots, int64_t slot_len,
                       const int32_t *perm_in, const int32_t *perm_out, int64_t nrhs,
                       const double *alpha, const double *beta, bool fuse) {
    if (t.variant == 1) {
        if (nrhs == 1) launch_fwdw_t<R, 1>(s, t, factors, x, ldx, y, ldy, slots, slot_len, perm_in, perm_out, nrhs, alpha, beta, fuse);
        else if (nrhs == 2) launch_fwdw_t<R, 2>(s, t, factors, x, ldx, y, ldy, slots, slot_len, perm_in, perm_out, nrhs, alpha, beta, fuse);
        else if (nrhs <= 4) launch_fwdw_t<R, 4>(s, t, factors, x, ldx, y, ldy, slots, slot_len, perm_in, perm_out, nrhs, alpha, beta, fuse);
        else launch_fwdw_t<R, 8>(s, t, factors, x, ldx, y, ldy, slots, slot_len, perm_in, perm_out, nrhs, alpha, beta, fuse);
        return;
    }
    if (nrhs == 1) launch_fwd_t<R, 1>(s, t, factors, x, ldx, y, ldy, slots, slot_len, perm_in, perm_out, nrhs, alpha, beta, fuse);
    else if (nrhs == 2) launch_fwd_t<R, 2>(s, t, factors, x, ldx, y, ldy, slots, slot_len, perm_in, perm_out, nrhs, alpha, beta, fuse);
    else if (nrhs <= 4) launch_fwd_t<R, 4>(s, t, factors, x, ldx, y, ldy, slots, slot_len, perm_in, perm_out, nrhs, alpha, beta, fuse);
    else launch_fwd_t<R, 8>(s, t, factors, x, ldx, y, ldy, slots, slot_len, perm_in, perm_out, nrhs, alpha, beta, fuse);
}

// ---------------------------------------------------------------------------
// Fused forward sweep.  All tasks of all stages (Q, R^1..R^{L-1}, P) in one
// cooperative launch; workgroup w runs tasks w, w + W, w + 2W, ... (tasks are
// ordered by stage, so every wait targets a task with a smaller index, and
// all W workgroups are co-resident).  A task of stage s > 0 waits until the
// counter of its input group has collected `wait_cnt` signals for this
// launch's epoch; its factor loads are issued before the wait.  Skeleton
// vectors written and read inside the launch use agent-scope (device-
// coherent) atomic loads and stores, so producers and consumers on
// different XCDs agree without cache flushes: a producer's stores complete
// (s_waitcnt) before it increments the consumer's counter.  Waits are
// bounded; an expired wait raises the error flag and the task proceeds, so
// every wave reaches the end of the kernel.
// ---------------------------------------------------------------------------
struct SweepArgs {
    const MvTask *tasks;
    const MvContrib *contribs;
    int32_t n_tasks, n_stages;
    const void *F[SWEEP_MAX_STAGES];
    const void *x;
    int64_t ldx;
    void *y;
    int64_t ldy;
    void *slots;
    int64_t slot_len;
    const int32_t *perm_in, *perm_out;
    double alpha[2], beta[2];
    int32_t fuse, nrhs;
    int32_t *counters;
    int32_t epoch;
    int32_t *err;
    int32_t *ticket;  // dynamic mode: [0] next task, [1] finished workgroups
    int32_t dyn;
};

template <typename R> __device__ __forceinline__ C<R> ld_coh(const C<R> *p) {
    const uint64_t *q = reinterpret_cast<const uint64_t *>(p);
    if constexpr (sizeof(R) == 4) {
        uint64_t v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        C<R> r;
        __builtin_memcpy(&r, &v, 8);
        return r;
    } else {
        uint64_t a = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint64_t b = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        C<R> r;
        __builtin_memcpy(&r.x, &a, 8);
        __builtin_memcpy(&r.y, &b, 8);
        return r;
    }
}

template <typename R> __device__ __forceinline__ void st_coh(C<R> *p, C<R> v) {
    uint64_t *q = reinterpret_cast<uint64_t *>(p);
    if constexpr (sizeof(R) == 4) {
        uint64_t u;
        __builtin_memcpy(&u, &v, 8);
        __hip_atomic_store(q, u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        uint64_t a, b;
        __builtin_memcpy(&a, &v.x, 8);
        __builtin_memcpy(&b, &v.y, 8);
        __hip_atomic_store(q, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(q + 1, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <typename R, int NR>
__global__ __launch_bounds__(256) void sweep_kernel(SweepArgs args) {
    constexpr int QMAX = NR <= 2 ? 8 : 4;
    constexpr int64_t SPIN_MAX = int64_t(1) << 26;
    __shared__ C<R> red[NR][256];
    const int t = threadIdx.x;
    const int nr = min(NR, args.nrhs);
    const C<R> *x = static_cast<const C<R> *>(args.x);
    C<R> *y = static_cast<C<R> *>(args.y);
    C<R> *slots = static_cast<C<R> *>(args.slots);
    const C<R> alpha = {R(args.alpha[0]), R(args.alpha[1])}, beta = {R(args.beta[0]), R(args.beta[1])};

    // Static mode (cooperative launch): tasks w, w + W, ...  Dynamic mode
    // (ordinary launch): tasks are claimed in order from a ticket counter;
    // a claimed task only waits on smaller tickets, which running
    // workgroups hold, so no co-residency is needed.  The last workgroup to
    // finish resets the counter for the next launch.
    __shared__ int s_task;
    auto claim = [&](int prev) -> int {
        if (!args.dyn) return prev < 0 ? int(blockIdx.x) : prev + int(gridDim.x);
        if (t == 0) s_task = __hip_atomic_fetch_add(args.ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        int v = s_task;
        __syncthreads();
        return v;
    };
    for (int task = claim(-1); task < args.n_tasks; task = claim(task)) {
        const MvTask tk = args.tasks[task];
        const int stage = tk.pad;
        const bool gather = stage == 0, scatter = stage == args.n_stages - 1;
        const C<R> *F = static_cast<const C<R> *>(args.F[stage]);
        auto in_val = [&](int r, int idx) -> C<R> {
            if (r >= nr) return C<R>{R(0), R(0)};
            if (gather) return x[int64_t(r) * args.ldx + args.perm_in[idx]];
            return ld_coh(slots + int64_t(r) * args.slot_len + idx);
        };
        // prefetch the first pass of the first contribution before waiting
        const int rows0 = min(256, tk.out_len);
        const bool pre = rows0 > 0 && tk.fofs >= 0;
        C<R> a0[QMAX];
        if (pre) {
            const int G0 = 256 / rows0, i0 = t % rows0, g0 = t / rows0;
#pragma unroll
            for (int q = 0; q < QMAX; ++q) {
                const int j = g0 + q * G0;
                a0[q] = (g0 < G0 && j < tk.n) ? F[tk.fofs + i0 + int64_t(j) * tk.m] : C<R>{R(0), R(0)};
            }
        }
        if (tk.wait_grp >= 0) {  // uniform: every thread reaches the barrier
            if (t == 0) {
                const int32_t target = args.epoch * tk.wait_cnt;
                int64_t spins = 0;
                while (__hip_atomic_load(args.counters + tk.wait_grp, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT) < target) {
                    __builtin_amdgcn_s_sleep(1);
                    if (++spins > SPIN_MAX) {
                        __hip_atomic_fetch_or(args.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                }
            }
            __syncthreads();
        }
        for (int row0 = 0; row0 < tk.out_len; row0 += 256) {
            const int rows = min(256, tk.out_len - row0);
            const int G = 256 / rows;
            const int i = t % rows, g = t / rows;
            const bool act = g < G;
            C<R> acc[NR];
#pragma unroll
            for (int r = 0; r < NR; ++r) acc[r] = C<R>{R(0), R(0)};
            for (int ci = tk.c_begin - 1; ci < tk.c_end; ++ci) {
                MvContrib c;
                if (ci < tk.c_begin) c = MvContrib{tk.fofs, tk.m, tk.n, tk.in_ofs, 0};
                else c = args.contribs[ci];
                if (c.fofs < 0) {  // identity block
                    if (g == 0)
#pragma unroll
                        for (int r = 0; r < NR; ++r) acc[r] = cadd(acc[r], in_val(r, c.in_ofs + row0 + i));
                    continue;
                }
                const C<R> *A = F + c.fofs + row0 + i;
                for (int p0 = 0; p0 < c.n; p0 += G * QMAX) {
                    C<R> a[QMAX], xv[NR][QMAX];
                    const bool use_pre = pre && row0 == 0 && ci < tk.c_begin && p0 == 0;
#pragma unroll
                    for (int q = 0; q < QMAX; ++q) {
                        const int j = p0 + g + q * G;
                        const bool ok = act && j < c.n;
                        a[q] = use_pre ? a0[q] : (ok ? A[int64_t(j) * c.m] : C<R>{R(0), R(0)});
#pragma unroll
                        for (int r = 0; r < NR; ++r) xv[r][q] = ok ? in_val(r, c.in_ofs + j) : C<R>{R(0), R(0)};
                    }
#pragma unroll
                    for (int q = 0; q < QMAX; ++q)
#pragma unroll
                        for (int r = 0; r < NR; ++r) cfma(acc[r], a[q], xv[r][q]);
                }
            }
#pragma unroll
            for (int r = 0; r < NR; ++r) red[r][t] = acc[r];
            __syncthreads();
            if (t < rows) {
                for (int r = 0; r < nr; ++r) {
                    C<R> s = red[r][t];
                    for (int q = 1; q < G; ++q) s = cadd(s, red[r][q * rows + t]);
                    if (!scatter) {
                        st_coh(slots + int64_t(r) * args.slot_len + tk.out_ofs + row0 + t, s);
                    } else {
                        C<R> *yp = y + int64_t(r) * args.ldy + args.perm_out[tk.out_ofs + row0 + t];
                        C<R> old = *yp;
                        *yp = cadd(args.fuse ? cmul(beta, old) : old, cmul(alpha, s));
                    }
                }
            }
            __syncthreads();
        }
        if (tk.sig_grp >= 0) {
            __builtin_amdgcn_s_waitcnt(0);  // this thread's stores have completed
            __syncthreads();
            if (t == 0)
                __hip_atomic_fetch_add(args.counters + tk.sig_grp, 1, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if (args.dyn && t == 0) {
        int d = __hip_atomic_fetch_add(args.ticket + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (d == int(gridDim.x) - 1) {  // every other workgroup has claimed its last ticket
            __hip_atomic_store(args.ticket, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(args.ticket + 1, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

template <typename R, int NR>
static void launch_sweep_t(hipStream_t s, const SweepTables &sw, const std::vector<DevBuf> &factors,
                           const void *x, int64_t ldx, void *y, int64_t ldy, void *slots,
                           int64_t slot_len, const int32_t *perm_in, const int32_t *perm_out,
                           int64_t nrhs, const double *alpha, const double *beta, bool fuse, int n_cu) {
    SweepArgs a{};
    a.tasks = sw.tasks.as<const MvTask>();
    a.contribs = sw.contribs.as<const MvContrib>();
    a.n_tasks = sw.n_tasks;
    a.n_stages = int32_t(factors.size());
    BFG_REQUIRE(a.n_stages <= SWEEP_MAX_STAGES, "sweep: too many stages");
    for (int i = 0; i < a.n_stages; ++i) a.F[i] = factors[i].p;
    a.x = x;
    a.ldx = ldx;
    a.y = y;
    a.ldy = ldy;
    a.slots = slots;
    a.slot_len = slot_len;
    a.perm_in = perm_in;
    a.perm_out = perm_out;
    a.alpha[0] = alpha[0]; a.alpha[1] = alpha[1];
    a.beta[0] = beta[0]; a.beta[1] = beta[1];
    a.fuse = fuse;
    a.nrhs = int32_t(nrhs);
    a.counters = sw.counters.as<int32_t>();
    a.epoch = sw.epoch;
    a.err = sw.err.as<int32_t>();
    a.ticket = sw.counters.as<int32_t>() + sw.n_groups;
    a.dyn = sw.mode == 0;
    int per_cu = 0;
    BFG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sweep_kernel<R, NR>, 256, 0));
    BFG_REQUIRE(per_cu >= 1, "sweep: kernel does not fit on a CU");
    int grid = std::min<int64_t>(int64_t(std::min(per_cu, sw.wg_per_cu > 0 ? sw.wg_per_cu : per_cu)) * n_cu,
                                 sw.n_tasks);
    grid = std::max(grid, 1);
    if (a.dyn) {
        hipLaunchKernelGGL((sweep_kernel<R, NR>), dim3(grid), dim3(256), 0, s, a);
        BFG_HIP(hipGetLastError());
    } else {
        void *kargs[] = {&a};
        BFG_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(sweep_kernel<R, NR>), dim3(grid),
                                           dim3(256), kargs, 0, s));
    }
}

void launch_sweep(hipStream_t s, int prec, const SweepTables &sw, const std::vector<DevBuf> &factors,
                  const void *x, int64_t ldx, void *y, int64_t ldy, void *slots, int64_t slot_len,
                  const int32_t *perm_in, const int32_t *perm_out, int64_t nrhs, const double *alpha,
                  const double *beta, bool fuse, int n_cu) {
    BFG_REQUIRE(nrhs >= 1 && nrhs <= 8, "sweep: 1..8 right-hand sides");
#define BFG_SW(R, NR) launch_sweep_t<R, NR>(s, sw, factors, x, ldx, y, ldy, slots, slot_len, perm_in, perm_out, nrhs, alpha, beta, fuse, n_cu)
    if (prec == BFGPU_C128) {
        if (nrhs == 1) BFG_SW(double, 1);
        else if (nrhs == 2) BFG_SW(double, 2);
        else if (nrhs <= 4) BFG_SW(double, 4);
        else BFG_SW(double, 8);
    } else {
        if (nrhs == 1) BFG_SW(float, 1);
        else if (nrhs == 2) BFG_SW(float, 2);
        else if (nrhs <= 4) BFG_SW(float, 4);
        else BFG_SW(float, 8);
    }
#undef BFG_SW
}

__global__ void zero_i32_kernel(int32_t *p, int64_t n) {
    int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

void launch_zero_i32(hipStream_t s, int32_t *p, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(zero_i32_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, p, n);
    BFG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Matvec stage.  One workgroup per (item, right-hand side).  An item is an
// output vector (slot, or rows of y reached through a permutation) that
// receives the sum of its contributions op(A_c) in_c; in_c is a slot or a
// gather of x through a permutation.
//   OP = 0 (N): out[i] = sum_c sum_j A_c[i,j] in_c[j].  Lanes walk rows
//     (column-major blocks -> coalesced), the 4 waves split the columns,
//     partial sums meet in LDS in fixed wave order (deterministic).
//   OP = 1 (T) / 2 (C): out[j] = sum_c sum_i op(A_c[i,j]) in_c[i].  Wave w
//     owns columns j = w mod 4, lanes walk rows, wave reduction per column.
// Identity contributions (fofs < 0) add in_c.
// ---------------------------------------------------------------------------
template <typename R, int OP, int INK, int OUTK>
__global__ __launch_bounds__(256) void stage_kernel(
    const MvItem *__restrict__ items, const MvContrib *__restrict__ contribs,
    const C<R> *__restrict__ F, const C<R> *__restrict__ x, int64_t ldx, C<R> *__restrict__ y,
    int64_t ldy, C<R> *__restrict__ slots, int64_t slot_len, const int32_t *__restrict__ perm_in,
    const int32_t *__restrict__ perm_out, C<R> alpha, C<R> beta, int fuse_beta, int max_in) {
    extern __shared__ char smraw[];
    C<R> *xin = reinterpret_cast<C<R> *>(smraw);               // max_in
    C<R> *oacc = xin + max_in;                                 // OP != 0: out_len
    __shared__ C<R> red[4][64];
    const MvItem it = items[blockIdx.x];
    const int64_t rhs = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const C<R> *xr = x + rhs * ldx;
    C<R> *yr = y + rhs * ldy;
    C<R> *sr = slots + rhs * slot_len;

    auto load_in = [&](const MvContrib &c, int len) {
        for (int j = tid; j < len; j += 256) {
            if (INK == IN_GATHER) xin[j] = xr[perm_in[c.in_ofs + j]];
            else xin[j] = sr[c.in_ofs + j];
        }
    };
    auto store_out = [&](int i, C<R> v) {
        if (OUTK == OUT_SLOT) {
            sr[it.out_ofs + i] = v;
        } else {
            int64_t idx = perm_out[it.out_ofs + i];
            C<R> old = yr[idx];
            C<R> base = fuse_beta ? cmul(beta, old) : old;
            yr[idx] = cadd(base, cmul(alpha, v));
        }
    };

    if (OP == 0) {
        for (int r0 = 0; r0 < it.out_len; r0 += 64) {
            const int i = r0 + lane;
            C<R> acc = {R(0), R(0)};
            for (int ci = it.c_begin; ci < it.c_end; ++ci) {
                const MvContrib c = contribs[ci];
                __syncthreads();
                load_in(c, c.n);
                __syncthreads();
                if (c.fofs < 0) {
                    if (wave == 0 && i < c.m) acc = cadd(acc, xin[i]);
                    continue;
                }
                const int jb = (c.n * wave) >> 2, je = (c.n * (wave + 1)) >> 2;
                if (i < c.m) {
                    const C<R> *Ac = F + c.fofs + i;
                    for (int j = jb; j < je; ++j) cfma(acc, Ac[int64_t(j) * c.m], xin[j]);
                }
            }
            red[wave][lane] = acc;
            __syncthreads();
            if (wave == 0 && i < it.out_len) {
                C<R> s = red[0][lane];
                s = cadd(s, red[1][lane]);
                s = cadd(s, red[2][lane]);
                s = cadd(s, red[3][lane]);
                store_out(i, s);
            }
            __syncthreads();
        }
    } else {
        for (int j = tid; j < it.out_len; j += 256) oacc[j] = C<R>{R(0), R(0)};
        for (int ci = it.c_begin; ci < it.c_end; ++ci) {
            const MvContrib c = contribs[ci];
            __syncthreads();
            load_in(c, c.m);
            __syncthreads();
            if (c.fofs < 0) {
                for (int j = tid; j < c.n; j += 256) oacc[j] = cadd(oacc[j], xin[j]);
                continue;
            }
            for (int j = wave; j < c.n; j += 4) {
                const C<R> *Ac = F + c.fofs + int64_t(j) * c.m;
                C<R> s = {R(0), R(0)};
                for (int i = lane; i < c.m; i += 64) {
                    if (OP == 2) cfmac(s, Ac[i], xin[i]);
                    else cfma(s, Ac[i], xin[i]);
                }
                s.x = wave_sum(s.x);
                s.y = wave_sum(s.y);
                if (lane == 0) oacc[j] = cadd(oacc[j], s);
            }
        }
        __syncthreads();
        for (int j = tid; j < it.out_len; j += 256) store_out(j, oacc[j]);
    }
}

template <typename R, int OP>
static void launch_stage_t(hipStream_t s, const StageTable &t, const void *factors,
                           const void *x, int64_t ldx, void *y, int64_t ldy, void *slots,
                           int64_t slot_len, const int32_t *perm_in, const int32_t *perm_out,
                           int64_t nrhs, const double *alpha, const double *beta, bool fuse) {
    C<R> a = {R(alpha[0]), R(alpha[1])}, b = {R(beta[0]), R(beta[1])};
    int max_in = t.max_in > 0 ? t.max_in : 1;
    size_t lds = sizeof(C<R>) * (size_t(max_in) + (OP != 0 ? size_t(t.max_out) : 0));
    BFG_REQUIRE(lds <= 64 * 1024, "matvec stage: vectors too long for LDS staging");
    auto F = (const C<R> *)factors;
    auto X = (const C<R> *)x;
    auto Y = (C<R> *)y;
    auto S = (C<R> *)slots;
    auto P = t.contribs.as<const MvContrib>();
    for (int32_t t0 = 0; t0 < t.n_items; t0 += TASK_CHUNK) {
        const int32_t nt = std::min(TASK_CHUNK, t.n_items - t0);
        dim3 grid(static_cast<unsigned>(nt), static_cast<unsigned>(nrhs));
        auto I = t.items.as<const MvItem>() + t0;
#define BFG_STAGE(INK, OUTK)                                                                \
    hipLaunchKernelGGL((stage_kernel<R, OP, INK, OUTK>), grid, dim3(256), lds, s, I, P, F, X, \
                       ldx, Y, ldy, S, slot_len, perm_in, perm_out, a, b, int(fuse), max_in)
        if (t.in_kind == IN_SLOT && t.out_kind == OUT_SLOT) BFG_STAGE(IN_SLOT, OUT_SLOT);
        else if (t.in_kind == IN_GATHER && t.out_kind == OUT_SLOT) BFG_STAGE(IN_GATHER, OUT_SLOT);
        else if (t.in_kind == IN_SLOT && t.out_kind == OUT_SCATTER) BFG_STAGE(IN_SLOT, OUT_SCATTER);
        else BFG_STAGE(IN_GATHER, OUT_SCATTER);
#undef BFG_STAGE
        BFG_HIP(hipGetLastError());
    }
}

void launch_stage(hipStream_t s, int prec, int op, const StageTable &t, const void *factors,
                  const void *x, int64_t ldx, void *y, int64_t ldy, void *slots,
                  int64_t slot_len, const int32_t *perm_in, const int32_t *perm_out,
                  int64_t nrhs, const double *alpha, const double *beta, bool fuse_beta) {
    if (t.n_items == 0 || nrhs == 0) return;
    BFG_REQUIRE(nrhs < 65536, "matvec: nrhs too large");
    if (prec == BFGPU_C128) {
        if (op == BFGPU_OP_N) launch_fwd<double>(s, t, factors, x, ldx, y, ldy, slots, slot_len, perm_in, perm_out, nrhs, alpha, beta, fuse_beta);
        else if (op == BFGPU_OP_T) launch_stage_t<double, 1>(s, t, factors, x, ldx, y, ldy, slots, slot_len, perm_in, perm_out, nrhs, alpha, beta, fuse_beta);
        else launch_stage_t<double, 2>(s, t, factors, x, ldx, y, ldy, slots, slot_len, perm_in, perm_out, nrhs, alpha, beta, fuse_beta);
    } else {
        if (op == BFGPU_OP_N) launch_fwd<float>(s, t, factors, x, ldx, y, ldy, slots, slot_len, perm_in, perm_out, nrhs, alpha, beta, fuse_beta);
        else if (op == BFGPU_OP_T) launch_stage_t<float, 1>(s, t, factors, x, ldx, y, ldy, slots, slot_len, perm_in, perm_out, nrhs, alpha, beta, fuse_beta);
        else launch_stage_t<float, 2>(s, t, factors, x, ldx, y, ldy, slots, slot_len, perm_in, perm_out, nrhs, alpha, beta, fuse_beta);
    }
}

// y[:, r] = beta * y[:, r] over n rows (beta == 0 writes zeros).
template <typename R>
__global__ void scale_kernel(C<R> *y, int64_t n, int64_t ldy, C<R> beta) {
    int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    C<R> *p = y + int64_t(blockIdx.y) * ldy + i;
    *p = (beta.x == R(0) && beta.y == R(0)) ? C<R>{R(0), R(0)} : cmul(beta, *p);
}

void launch_scale(hipStream_t s, int prec, void *y, int64_t n, int64_t ldy, int64_t nrhs,
                  const double *beta) {
    if (n <= 0 || nrhs <= 0) return;
    dim3 grid(unsigned((n + 255) / 256), unsigned(nrhs));
    if (prec == BFGPU_C128)
        hipLaunchKernelGGL(scale_kernel<double>, grid, dim3(256), 0, s, (C<double> *)y, n, ldy,
                           C<double>{beta[0], beta[1]});
    else
        hipLaunchKernelGGL(scale_kernel<float>, grid, dim3(256), 0, s, (C<float> *)y, n, ldy,
                           C<float>{float(beta[0]), float(beta[1])});
    BFG_HIP(hipGetLastError());
}

template <typename A, typename B>
__global__ void convert_kernel(const A *in, B *out, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        out[i] = B(in[i]);
}

void launch_convert(hipStream_t s, const void *in, int in_prec, void *out, int out_prec,
                    int64_t n) {
    // n complex entries -> 2 n reals
    int64_t nr = 2 * n;
    if (nr <= 0) return;
    dim3 grid(unsigned(std::min<int64_t>((nr + 255) / 256, int64_t(1) << 22)));
    if (in_prec == BFGPU_C128 && out_prec == BFGPU_C64)
        hipLaunchKernelGGL((convert_kernel<double, float>), grid, dim3(256), 0, s,
                           (const double *)in, (float *)out, nr);
    else if (in_prec == BFGPU_C64 && out_prec == BFGPU_C128)
        hipLaunchKernelGGL((convert_kernel<float, double>), grid, dim3(256), 0, s,
                           (const float *)in, (double *)out, nr);
    else
        BFG_HIP(hipMemcpyAsync(out, in, size_t(n) * (in_prec == BFGPU_C128 ? 16 : 8),
                               hipMemcpyDeviceToDevice, s));
    BFG_HIP(hipGetLastError());
}

}  // namespace bfg
