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
#include <cmath>
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
    // near-field quadrature (bfgpu_kernel_rwg_set_near): per function and
    // side 13 doubles (triangle vertices, free vertex, c), the rule degrees
    // and the singularity-extraction distance factor
    bool has_near = false;
    DevBuf test_geom, trial_geom;
    int32_t near_outer = 5, near_inner = 5, far_outer = 5, far_inner = 5;
    double near_factor = 3.0;
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

// ---------------------------------------------------------------------------
// Near-field quadrature.  Symmetric triangle rules of degree 1-5 (Strang &
// Fix; Dunavant 1985): barycentric points and weights summing to 1, the
// points of degree d at RULE_OFS[d] .. RULE_OFS[d] + RULE_N[d].
// ---------------------------------------------------------------------------
constexpr double A4 = 0.445948490915965, W4A = 0.223381589678011, B4 = 0.091576213509771,
                 W4B = 0.109951743655322;
constexpr double A5 = 0.470142064105115, W5A = 0.132394152788506, B5 = 0.101286507323456,
                 W5B = 0.125939180544827;
__constant__ double RULE_B[23][3] = {
    {1.0 / 3, 1.0 / 3, 1.0 / 3},                                                   // degree 1
    {4.0 / 6, 1.0 / 6, 1.0 / 6}, {1.0 / 6, 4.0 / 6, 1.0 / 6}, {1.0 / 6, 1.0 / 6, 4.0 / 6},  // 2
    {1.0 / 3, 1.0 / 3, 1.0 / 3}, {0.6, 0.2, 0.2}, {0.2, 0.6, 0.2}, {0.2, 0.2, 0.6},   // 3
    {1 - 2 * A4, A4, A4}, {A4, 1 - 2 * A4, A4}, {A4, A4, 1 - 2 * A4},                // 4
    {1 - 2 * B4, B4, B4}, {B4, 1 - 2 * B4, B4}, {B4, B4, 1 - 2 * B4},
    {1.0 / 3, 1.0 / 3, 1.0 / 3},                                                   // 5
    {1 - 2 * A5, A5, A5}, {A5, 1 - 2 * A5, A5}, {A5, A5, 1 - 2 * A5},
    {1 - 2 * B5, B5, B5}, {B5, 1 - 2 * B5, B5}, {B5, B5, 1 - 2 * B5}};
__constant__ double RULE_W[23] = {1.0,
                                  1.0 / 3, 1.0 / 3, 1.0 / 3,
                                  -27.0 / 48, 25.0 / 48, 25.0 / 48, 25.0 / 48,
                                  W4A, W4A, W4A, W4B, W4B, W4B,
                                  0.225, W5A, W5A, W5A, W5B, W5B, W5B};
__constant__ int RULE_OFS[6] = {0, 0, 1, 4, 8, 14};
__constant__ int RULE_N[6] = {0, 1, 3, 4, 6, 7};

struct V3 {
    double x, y, z;
};
__device__ __forceinline__ V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 vadd(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 vscale(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 vcross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double vnorm(V3 a) { return sqrt(vdot(a, a)); }

// One side of an RWG function: triangle p[3], f(x) = c (x - free).
struct Side {
    V3 p[3], free;
    double c;
};
__device__ __forceinline__ Side load_side(const double *g) {
    Side s;
    for (int v = 0; v < 3; ++v) s.p[v] = V3{g[3 * v], g[3 * v + 1], g[3 * v + 2]};
    s.free = V3{g[9], g[10], g[11]};
    s.c = g[12];
    return s;
}
__device__ __forceinline__ double side_area(const Side &s) {
    return 0.5 * vnorm(vcross(vsub(s.p[1], s.p[0]), vsub(s.p[2], s.p[0])));
}
__device__ __forceinline__ double longest_edge(const Side &s) {
    return fmax(vnorm(vsub(s.p[1], s.p[0])), fmax(vnorm(vsub(s.p[2], s.p[1])), vnorm(vsub(s.p[0], s.p[2]))));
}
__device__ __forceinline__ V3 rule_point(const Side &s, int q) {
    return vadd(vadd(vscale(s.p[0], RULE_B[q][0]), vscale(s.p[1], RULE_B[q][1])), vscale(s.p[2], RULE_B[q][2]));
}

// Closed-form potential integrals of triangle t at x (Wilton, Rao, Glisson,
// Schaubert, Al-Bundak and Butler, IEEE Trans. AP 32(3), 1984):
//   im1 = int_t 1/|x - y| dS_y,  iv = int_t (rho_y - rho_x)/|x - y| dS_y,
// rho the projection onto the plane of t; per edge a -> b with unit tangent
// l and outward in-plane normal u = l x n: l+- = (b|a - rho).l,
// P0 = (a - rho).u, R0^2 = P0^2 + d^2, R+- = sqrt(R0^2 + l+-^2),
// f = ln((R+ + l+)/(R- + l-)) in the form without cancellation,
// im1 = sum P0 f - |d| (atan(P0 l+/(R0^2 + |d| R+)) - atan(P0 l-/(R0^2 + |d| R-))),
// iv = 1/2 sum u (R0^2 f + l+ R+ - l- R-).  Also returns rho.
__device__ void potential_integrals(const Side &t, V3 x, double &im1, V3 &iv, V3 &rho) {
    V3 n = vcross(vsub(t.p[1], t.p[0]), vsub(t.p[2], t.p[0]));
    n = vscale(n, 1.0 / vnorm(n));
    const double d = vdot(vsub(x, t.p[0]), n), ad = fabs(d);
    rho = vsub(x, vscale(n, d));
    const double scale = longest_edge(t);
    im1 = 0.0;
    iv = V3{0.0, 0.0, 0.0};
    for (int i = 0; i < 3; ++i) {
        const V3 a = t.p[i], b = t.p[(i + 1) % 3];
        V3 l = vsub(b, a);
        l = vscale(l, 1.0 / vnorm(l));
        const V3 u = vcross(l, n);
        const double lp = vdot(vsub(b, rho), l), lm = vdot(vsub(a, rho), l);
        const double p0 = vdot(vsub(a, rho), u);
        const double r02 = p0 * p0 + d * d;
        const double rp = sqrt(r02 + lp * lp), rm = sqrt(r02 + lm * lm);
        double f = 0.0, at = 0.0;
        if (r02 > (1e-14 * scale) * (1e-14 * scale)) {
            if (lm >= 0.0) f = log((rp + lp) / (rm + lm));
            else if (lp <= 0.0) f = log((rm - lm) / (rp - lp));
            else f = log((rp + lp) * (rm - lm) / r02);
            at = atan2(p0 * lp, r02 + ad * rp) - atan2(p0 * lm, r02 + ad * rm);
        }
        im1 += p0 * f - ad * at;
        iv = vadd(iv, vscale(u, 0.5 * (r02 * f + lp * rp - lm * rm)));
    }
}

// (exp(-ikr) - 1) / (4 pi r), with its limit -ik/(4 pi) at r = 0
__device__ __forceinline__ C<double> g_smooth(double k, double r) {
    if (r == 0.0) return C<double>{0.0, -k / (4.0 * M_PI)};
    double s, c;
    sincos(0.5 * k * r, &s, &c);
    const double g = 1.0 / (4.0 * M_PI * r);
    return C<double>{-2.0 * s * s * g, -2.0 * s * c * g};  // cos kr - 1 = -2 sin^2(kr/2), sin kr = 2 sin cos
}

// Entry part of test side m against trial side n: singularity extraction
// (outer rule over m; static part in closed form over n, the rest by the
// inner rule) or the double rule.
__device__ C<double> side_pair(const Side &m, const Side &n, double k, C<double> alpha, C<double> beta, bool near,
                               int outer, int inner) {
    const double am = side_area(m), an = side_area(n);
    const int o0 = RULE_OFS[outer], no = RULE_N[outer], i0 = RULE_OFS[inner], ni = RULE_N[inner];
    double sr = 0.0, si = 0.0;
    for (int a = 0; a < no; ++a) {
        const V3 x = rule_point(m, o0 + a);
        const V3 fx = vscale(vsub(x, m.free), m.c);
        const double wa = RULE_W[o0 + a] * am;
        // vec = int_n (y - free_n) G dS_y, sca = int_n G dS_y (complex)
        C<double> vx = {0, 0}, vy = {0, 0}, vz = {0, 0}, sca = {0, 0};
        if (near) {
            double im1;
            V3 iv, rho;
            potential_integrals(n, x, im1, iv, rho);
            const double q = 1.0 / (4.0 * M_PI);
            const V3 st = vscale(vadd(iv, vscale(vsub(rho, n.free), im1)), q);
            vx.x = st.x;
            vy.x = st.y;
            vz.x = st.z;
            sca.x = im1 * q;
        }
        for (int b = 0; b < ni; ++b) {
            const V3 y = rule_point(n, i0 + b);
            const double r = vnorm(vsub(x, y));
            const double wb = RULE_W[i0 + b] * an;
            C<double> g;
            if (near) {
                g = g_smooth(k, r);
            } else {
                double s, c;
                sincos(k * r, &s, &c);
                const double gg = 1.0 / (4.0 * M_PI * r);
                g = C<double>{c * gg, -s * gg};
            }
            g.x *= wb;
            g.y *= wb;
            const V3 dy = vsub(y, n.free);
            vx = C<double>{vx.x + g.x * dy.x, vx.y + g.y * dy.x};
            vy = C<double>{vy.x + g.x * dy.y, vy.y + g.y * dy.y};
            vz = C<double>{vz.x + g.x * dy.z, vz.y + g.y * dy.z};
            sca = cadd(sca, g);
        }
        // alpha c_n fx . vec + beta (2 c_m)(2 c_n) sca, times wa
        const C<double> fv = {fx.x * vx.x + fx.y * vy.x + fx.z * vz.x, fx.x * vx.y + fx.y * vy.y + fx.z * vz.y};
        const C<double> t1 = cmul(alpha, C<double>{n.c * fv.x, n.c * fv.y});
        const double dd = 4.0 * m.c * n.c;
        const C<double> t2 = cmul(beta, C<double>{dd * sca.x, dd * sca.y});
        sr += wa * (t1.x + t2.x);
        si += wa * (t1.y + t2.y);
    }
    return C<double>{sr, si};
}

// One thread per entry of the listed blocks (task layout as fill_rwg_kernel):
// the sum over the 2 x 2 triangle pairs of the test and trial function,
// each by singularity extraction when its centroids are closer than
// near_factor times the longer longest edge, else by the double rule.
template <typename R>
__global__ __launch_bounds__(256) void fill_rwg_near_kernel(const double *__restrict__ tg,
                                                            const double *__restrict__ sg,
                                                            const int32_t *__restrict__ rows,
                                                            const int32_t *__restrict__ cols,
                                                            const FillTask *__restrict__ tasks,
                                                            const int64_t *__restrict__ task_end, int32_t n_tasks,
                                                            int64_t total, C<R> *Z, double k, double ar, double ai,
                                                            double br, double bi, int near_outer, int near_inner,
                                                            int far_outer, int far_inner, double near_factor) {
    int32_t lo = 0;
    const C<double> alpha = {ar, ai}, beta = {br, bi};
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
        const double *a = tg + int64_t(rows[t.rofs + i]) * 26;
        const double *b = sg + int64_t(cols[t.cofs + j]) * 26;
        C<double> z = {0.0, 0.0};
        for (int p = 0; p < 2; ++p) {
            const Side sm = load_side(a + 13 * p);
            const V3 cm = vscale(vadd(vadd(sm.p[0], sm.p[1]), sm.p[2]), 1.0 / 3.0);
            const double hm = longest_edge(sm);
            for (int q = 0; q < 2; ++q) {
                const Side sn = load_side(b + 13 * q);
                const V3 cn = vscale(vadd(vadd(sn.p[0], sn.p[1]), sn.p[2]), 1.0 / 3.0);
                const bool near = vnorm(vsub(cm, cn)) < near_factor * fmax(hm, longest_edge(sn));
                z = cadd(z, side_pair(sm, sn, k, alpha, beta, near, near ? near_outer : far_outer,
                                      near ? near_inner : far_inner));
            }
        }
        Z[t.zofs + local] = C<R>{R(z.x), R(z.y)};
    }
}

}  // namespace

void launch_fill_rwg_near(hipStream_t s, int prec, const KernelRWG &kr, const int32_t *rows, const int32_t *cols,
                          const FillTask *tasks, const int64_t *task_end, int32_t n_tasks, int64_t total, void *Z,
                          double k) {
    BFG_REQUIRE(kr.has_near, "near field of the RWG plug-in: call bfgpu_kernel_rwg_set_near first");
    if (total <= 0 || n_tasks <= 0) return;
    dim3 grid(unsigned(std::min<int64_t>((total + 255) / 256, int64_t(1) << 22)));
    if (prec == BFGPU_C128)
        hipLaunchKernelGGL(fill_rwg_near_kernel<double>, grid, dim3(256), 0, s, kr.test_geom.as<double>(),
                           kr.trial_geom.as<double>(), rows, cols, tasks, task_end, n_tasks, total,
                           static_cast<C<double> *>(Z), k, kr.alpha[0], kr.alpha[1], kr.beta[0], kr.beta[1],
                           kr.near_outer, kr.near_inner, kr.far_outer, kr.far_inner, kr.near_factor);
    else
        hipLaunchKernelGGL(fill_rwg_near_kernel<float>, grid, dim3(256), 0, s, kr.test_geom.as<double>(),
                           kr.trial_geom.as<double>(), rows, cols, tasks, task_end, n_tasks, total,
                           static_cast<C<float> *>(Z), k, kr.alpha[0], kr.alpha[1], kr.beta[0], kr.beta[1],
                           kr.near_outer, kr.near_inner, kr.far_outer, kr.far_inner, kr.near_factor);
    BFG_HIP(hipGetLastError());
}

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

extern "C" int bfgpu_kernel_rwg_set_near(bfgpu_kernel *h, const double *test_geom, const double *trial_geom,
                                         int32_t near_outer, int32_t near_inner, int32_t far_outer,
                                         int32_t far_inner, double near_factor) {
    KR_API_BEGIN
    BFG_REQUIRE(h && test_geom && trial_geom, "kernel_rwg_set_near: null argument");
    for (int32_t d : {near_outer, near_inner, far_outer, far_inner})
        BFG_REQUIRE(d >= 1 && d <= 5, "kernel_rwg_set_near: rule degrees 1 to 5");
    BFG_REQUIRE(near_factor >= 0.0 && std::isfinite(near_factor), "kernel_rwg_set_near: near_factor >= 0");
    KernelRWG &k = h->k;
    const size_t bt = sizeof(double) * 26 * size_t(k.n_test), bs = sizeof(double) * 26 * size_t(k.n_trial);
    for (int64_t f = 0; f < k.n_test + k.n_trial; ++f) {
        const double *g = f < k.n_test ? test_geom + 26 * f : trial_geom + 26 * (f - k.n_test);
        for (int q = 0; q < 26; ++q) BFG_REQUIRE(std::isfinite(g[q]), "kernel_rwg_set_near: non-finite geometry");
        for (int side = 0; side < 2; ++side) {
            const double *t = g + 13 * side;
            const double ux = t[3] - t[0], uy = t[4] - t[1], uz = t[5] - t[2];
            const double vx = t[6] - t[0], vy = t[7] - t[1], vz = t[8] - t[2];
            const double cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
            BFG_REQUIRE(cx * cx + cy * cy + cz * cz > 0.0, "kernel_rwg_set_near: degenerate triangle");
        }
    }
    BFG_HIP(hipSetDevice(k.ctx->device));
    k.test_geom.alloc(bt);
    k.trial_geom.alloc(bs);
    BFG_HIP(hipMemcpy(k.test_geom.p, test_geom, bt, hipMemcpyHostToDevice));
    BFG_HIP(hipMemcpy(k.trial_geom.p, trial_geom, bs, hipMemcpyHostToDevice));
    k.near_outer = near_outer;
    k.near_inner = near_inner;
    k.far_outer = far_outer;
    k.far_inner = far_inner;
    k.near_factor = near_factor;
    k.has_near = true;
    KR_API_END
}

extern "C" int bfgpu_kernel_fill_near(const bfgpu_kernel *h, double k, int32_t precision, int32_t m,
                                      const int32_t *rows, int32_t n, const int32_t *cols, void *Z) {
    KR_API_BEGIN
    BFG_REQUIRE(h && rows && cols && Z, "kernel_fill_near: null argument");
    BFG_REQUIRE(precision == BFGPU_C64 || precision == BFGPU_C128, "kernel_fill_near: precision");
    BFG_REQUIRE(m >= 1 && n >= 1, "kernel_fill_near: empty block");
    const KernelRWG &kr = h->k;
    BFG_REQUIRE(kr.has_near, "kernel_fill_near: call bfgpu_kernel_rwg_set_near first");
    for (int32_t i = 0; i < m; ++i)
        BFG_REQUIRE(rows[i] >= 0 && rows[i] < kr.n_test, "kernel_fill_near: row index out of range");
    for (int32_t j = 0; j < n; ++j)
        BFG_REQUIRE(cols[j] >= 0 && cols[j] < kr.n_trial, "kernel_fill_near: column index out of range");
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
    launch_fill_rwg_near(s, precision, kr, dr.as<int32_t>(), dc.as<int32_t>(), dt.as<FillTask>(), de.as<int64_t>(),
                         1, total, dz.p, k);
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
bool kernel_has_near(const KernelRWG &k) { return k.has_near; }
}  // namespace bfg
