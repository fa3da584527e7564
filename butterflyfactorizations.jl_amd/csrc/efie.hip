// EFIE kernel-matrix plug-in on RWG functions (the BEAST plug-in of
// ext/BFBEAST/kernelmatrix.jl:22-82: Maxwell single layer between
// Raviart-Thomas / RWG test and trial functions, far interactions by a
// double numerical quadrature, DoubleNumQStrat).
//
//   Z[m, n] = sum_a sum_b G(x_a, y_b) (alpha f_a . g_b + beta d_a e_b)
//
// over the quadrature samples of test function m (points x_a on its two
// triangles, f_a = w A f_m(x_a), d_a = w A div f_m) and of trial function n
// (y_b, g_b, e_b), G(x, y) = exp(-ik|x-y|)/(4 pi |x-y|).  The host builds the
// samples (the RWG geometry and the rule weights); the device holds them as
// [function][sample][x, y, z, fx, fy, fz, div] doubles, so one entry reads
// two contiguous runs.  One thread per entry, FP64 arithmetic throughout,
// stored in the factorization's precision.  Coincident sample points (only
// possible for overlapping supports, never in admissible blocks) add nothing.
#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "device.hpp"
#include "internal.hpp"

namespace bfg {
void set_last_error(const std::string &m);

struct KernelRWG {
    Ctx *ctx = nullptr;
    double alpha[2] = {0, 0}, beta[2] = {0, 0};
    int64_t n_test = 0, n_trial = 0;
    int32_t s_test = 0, s_trial = 0;  // samples per function (2 triangles x rule points)
    DevBuf test, trial;
};

namespace {

template <typename R>
__global__ __launch_bounds__(256) void fill_rwg_kernel(const double *__restrict__ ts, int32_t s_test,
                                                       const double *__restrict__ ss, int32_t s_trial,
                                                       const int32_t *__restrict__ rows,
                                                       const int32_t *__restrict__ cols,
                                                       const FillTask *__restrict__ tasks,
                                                       const int64_t *__restrict__ task_end, int32_t n_tasks,
                                                       int64_t total, C<R> *Z, double k, double ar, double ai,
                                                       double br, double bi) {
    int32_t lo = 0;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += int64_t(gridDim.x) * 256) {
        int32_t hi = n_tasks - 1;  // entries only grow: search from the last task
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (task_end[mid] > e) hi = mid;
            else lo = mid + 1;
        }
        const FillTask t = tasks[lo];
        const int64_t local = e - (task_end[lo] - int64_t(t.m) * t.n);
        const int32_t i = int32_t(local % t.m), j = int32_t(local / t.m);
        const double *a = ts + int64_t(rows[t.rofs + i]) * s_test * 7;
        const double *b = ss + int64_t(cols[t.cofs + j]) * s_trial * 7;
        double sr = 0.0, si = 0.0;  // sum of G * (alpha f.g + beta d e)
        for (int p = 0; p < s_test; ++p) {
            const double *ap = a + 7 * p;
            for (int q = 0; q < s_trial; ++q) {
                const double *bq = b + 7 * q;
                const double dx = ap[0] - bq[0], dy = ap[1] - bq[1], dz = ap[2] - bq[2];
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                if (r == 0.0) continue;
                const double fg = ap[3] * bq[3] + ap[4] * bq[4] + ap[5] * bq[5];
                const double de = ap[6] * bq[6];
                const double wr = ar * fg + br * de, wi = ai * fg + bi * de;
                double s, c;
                sincos(k * r, &s, &c);
                const double g = 1.0 / (4.0 * M_PI * r);
                const double gr = c * g, gi = -s * g;
                sr += gr * wr - gi * wi;
                si += gr * wi + gi * wr;
            }
        }
        Z[t.zofs + local] = C<R>{R(sr), R(si)};
    }
}

}  // namespace

void launch_fill_rwg(hipStream_t s, int prec, const KernelRWG &kr, const int32_t *rows, const int32_t *cols,
                     const FillTask *tasks, const int64_t *task_end, int32_t n_tasks, int64_t total, void *Z,
                     double k) {
    if (total <= 0 || n_tasks <= 0) return;
    dim3 grid(unsigned(std::min<int64_t>((total + 255) / 256, int64_t(1) << 22)));
    if (prec == BFGPU_C128)
        hipLaunchKernelGGL(fill_rwg_kernel<double>, grid, dim3(256), 0, s, kr.test.as<double>(), kr.s_test,
                           kr.trial.as<double>(), kr.s_trial, rows, cols, tasks, task_end, n_tasks, total,
                           static_cast<C<double> *>(Z), k, kr.alpha[0], kr.alpha[1], kr.beta[0], kr.beta[1]);
    else
        hipLaunchKernelGGL(fill_rwg_kernel<float>, grid, dim3(256), 0, s, kr.test.as<double>(), kr.s_test,
                           kr.trial.as<double>(), kr.s_trial, rows, cols, tasks, task_end, n_tasks, total,
                           static_cast<C<float> *>(Z), k, kr.alpha[0], kr.alpha[1], kr.beta[0], kr.beta[1]);
    BFG_HIP(hipGetLastError());
}

}  // namespace bfg

using namespace bfg;

struct bfgpu_kernel { bfg::KernelRWG k; };

#define KR_API_BEGIN try {
#define KR_API_END                        \
    }                                     \
    catch (const Error &e) {              \
        set_last_error(e.what());         \
        return e.code;                    \
    }                                     \
    catch (const std::exception &e) {     \
        set_last_error(e.what());         \
        return BFGPU_ERR_STATE;           \
    }                                     \
    return BFGPU_OK;

extern "C" int bfgpu_kernel_rwg_create(bfgpu_ctx *ctx, const double *alpha, const double *beta, int64_t n_test,
                                       int32_t s_test, const double *test_samples, int64_t n_trial,
                                       int32_t s_trial, const double *trial_samples, bfgpu_kernel **out) {
    KR_API_BEGIN
    BFG_REQUIRE(ctx && alpha && beta && out && test_samples && trial_samples, "kernel_rwg_create: null argument");
    BFG_REQUIRE(n_test > 0 && n_trial > 0 && n_test < (int64_t(1) << 31) && n_trial < (int64_t(1) << 31),
                "kernel_rwg_create: function counts");
    BFG_REQUIRE(s_test >= 1 && s_test <= 64 && s_trial >= 1 && s_trial <= 64,
                "kernel_rwg_create: 1 to 64 samples per function");
    BFG_HIP(hipSetDevice(ctx->c.device));
    auto h = std::make_unique<bfgpu_kernel>();
    KernelRWG &k = h->k;
    k.ctx = &ctx->c;
    k.alpha[0] = alpha[0];
    k.alpha[1] = alpha[1];
    k.beta[0] = beta[0];
    k.beta[1] = beta[1];
    k.n_test = n_test;
    k.n_trial = n_trial;
    k.s_test = s_test;
    k.s_trial = s_trial;
    const size_t bt = sizeof(double) * 7 * size_t(n_test) * size_t(s_test);
    const size_t bs = sizeof(double) * 7 * size_t(n_trial) * size_t(s_trial);
    k.test.alloc(bt);
    k.trial.alloc(bs);
    BFG_HIP(hipMemcpy(k.test.p, test_samples, bt, hipMemcpyHostToDevice));
    BFG_HIP(hipMemcpy(k.trial.p, trial_samples, bs, hipMemcpyHostToDevice));
    *out = h.release();
    KR_API_END
}

extern "C" int bfgpu_kernel_fill(const bfgpu_kernel *h, double k, int32_t precision, int32_t m,
                                 const int32_t *rows, int32_t n, const int32_t *cols, void *Z) {
    KR_API_BEGIN
    BFG_REQUIRE(h && rows && cols && Z, "kernel_fill: null argument");
    BFG_REQUIRE(precision == BFGPU_C64 || precision == BFGPU_C128, "kernel_fill: precision");
    BFG_REQUIRE(m >= 1 && n >= 1, "kernel_fill: empty block");
    const KernelRWG &kr = h->k;
    for (int32_t i = 0; i < m; ++i)
        BFG_REQUIRE(rows[i] >= 0 && rows[i] < kr.n_test, "kernel_fill: row index out of range");
    for (int32_t j = 0; j < n; ++j)
        BFG_REQUIRE(cols[j] >= 0 && cols[j] < kr.n_trial, "kernel_fill: column index out of range");
    BFG_HIP(hipSetDevice(kr.ctx->device));
    hipStream_t s = kr.ctx->stream;
    const size_t es = precision == BFGPU_C128 ? 16 : 8;
    const int64_t total = int64_t(m) * n;
    DevBuf dr, dc, dt, de, dz;
    dr.alloc(sizeof(int32_t) * size_t(m));
    dc.alloc(sizeof(int32_t) * size_t(n));
    const FillTask task{0, m, n, 0, 0};
    dt.alloc(sizeof(FillTask));
    de.alloc(sizeof(int64_t));
    dz.alloc(size_t(total) * es);
    BFG_HIP(hipMemcpyAsync(dr.p, rows, sizeof(int32_t) * size_t(m), hipMemcpyHostToDevice, s));
    BFG_HIP(hipMemcpyAsync(dc.p, cols, sizeof(int32_t) * size_t(n), hipMemcpyHostToDevice, s));
    BFG_HIP(hipMemcpyAsync(dt.p, &task, sizeof(FillTask), hipMemcpyHostToDevice, s));
    BFG_HIP(hipMemcpyAsync(de.p, &total, sizeof(int64_t), hipMemcpyHostToDevice, s));
    launch_fill_rwg(s, precision, kr, dr.as<int32_t>(), dc.as<int32_t>(), dt.as<FillTask>(), de.as<int64_t>(), 1,
                    total, dz.p, k);
    BFG_HIP(hipMemcpyAsync(Z, dz.p, size_t(total) * es, hipMemcpyDeviceToHost, s));
    BFG_HIP(hipStreamSynchronize(s));
    KR_API_END
}

extern "C" int bfgpu_kernel_destroy(bfgpu_kernel *h) {
    KR_API_BEGIN
    if (h) {
        BFG_HIP(hipSetDevice(h->k.ctx->device));
        BFG_HIP(hipStreamSynchronize(h->k.ctx->stream));
    }
    delete h;
    KR_API_END
}

namespace bfg {
const KernelRWG *kernel_of(const bfgpu_kernel *h) { return h ? &h->k : nullptr; }
int64_t kernel_functions(const KernelRWG &k, bool test) { return test ? k.n_test : k.n_trial; }
}  // namespace bfg
