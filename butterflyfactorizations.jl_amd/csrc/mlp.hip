// mlp.hip -- learned skeleton-rank predictor: a ReLU MLP in FP32 with its
// own forward, backward and Adam kernels.  It replaces the sample-size guess
// of the rank estimators (src/ButterflyFactorization/rankestimation.jl) and
// plays the role of the reference's rank calibration
// (evaluation/ranklogging.jl:16-61, a least-squares fit on logged ranks).
//
// Layout: parameters are one flat FP32 array W1, b1, ..., WL, bL with W_l
// row-major (out x in).  Activations are row-major (samples x features).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>

#include "internal.hpp"

struct bfgpu_mlp {
    bfg::Ctx *ctx = nullptr;
    int32_t n_in = 0, hidden = 0, n_layers = 0;
    std::vector<int32_t> dims;          // n_in, H, ..., H, 1
    std::vector<int64_t> w_ofs, b_ofs;  // offsets into params
    int64_t n_params = 0;
    bfg::DevBuf params, grads, adam_m, adam_v, norm;  // norm: mu[n_in], 1/sigma[n_in]
    float out_shift = 0.f, out_scale = 1.f;
    std::vector<float> mu, sigma;
    bfg::DevBuf ws;                     // activations + gradients workspace
    int64_t adam_t = 0;
};

namespace bfg {

constexpr int TM = 64, TN = 64, TK = 16;

// C[M x N] (ldc) = sum_k A'(m,k) B'(k,n); A'(m,k) = TA ? A[k*lda+m] : A[m*lda+k],
// B'(k,n) = TB ? B[n*ldb+k] : B[k*ldb+n].  EPI 0: plain, 1: + bias[n],
// 2: relu(. + bias[n]), 3: . * (mask[m*ldc+n] > 0).
template <int TA, int TB, int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(int M, int N, int K, const float *__restrict__ A,
                                                   int lda, const float *__restrict__ B, int ldb,
                                                   float *__restrict__ C, int ldc,
                                                   const float *__restrict__ bias,
                                                   const float *__restrict__ mask) {
    __shared__ float As[TK][TM + 1];
    __shared__ float Bs[TK][TN + 1];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += TK) {
        for (int e = tid; e < TM * TK; e += 256) {
            int mm = e % TM, kk = e / TM;
            int m = m0 + mm, k = k0 + kk;
            float v = 0.f;
            if (m < M && k < K) v = TA ? A[int64_t(k) * lda + m] : A[int64_t(m) * lda + k];
            As[kk][mm] = v;
        }
        for (int e = tid; e < TN * TK; e += 256) {
            int nn = e % TN, kk = e / TN;
            int n = n0 + nn, k = k0 + kk;
            float v = 0.f;
            if (n < N && k < K) v = TB ? B[int64_t(n) * ldb + k] : B[int64_t(k) * ldb + n];
            Bs[kk][nn] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int m = m0 + ty * 4 + i, n = n0 + tx * 4 + j;
            if (m >= M || n >= N) continue;
            float v = acc[i][j];
            if (EPI == 1 || EPI == 2) v += bias[n];
            if (EPI == 2) v = v > 0.f ? v : 0.f;
            if (EPI == 3) v = mask[int64_t(m) * ldc + n] > 0.f ? v : 0.f;
            C[int64_t(m) * ldc + n] = v;
        }
}

template <int TA, int TB, int EPI>
static void gemm(hipStream_t s, int M, int N, int K, const float *A, int lda, const float *B, int ldb,
                 float *C, int ldc, const float *bias = nullptr, const float *mask = nullptr) {
    if (M <= 0 || N <= 0) return;
    dim3 grid(unsigned((N + TN - 1) / TN), unsigned((M + TM - 1) / TM));
    hipLaunchKernelGGL((gemm_kernel<TA, TB, EPI>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C,
                       ldc, bias, mask);
    BFG_HIP(hipGetLastError());
}

// out[n] = sum_m X[m * ld + n] (bias gradient); one wave per column.
__global__ void colsum_kernel(const float *X, int M, int N, int ld, float *out) {
    int n = blockIdx.x;
    float s = 0.f;
    for (int m = threadIdx.x; m < M; m += 64) s += X[int64_t(m) * ld + n];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (threadIdx.x == 0) out[n] = s;
}

// x_norm = (x - mu) / sigma over row-major n x d features
__global__ void normalize_kernel(const float *x, const float *norm, int64_t n, int d, float *out) {
    int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (e >= n * d) return;
    int j = int(e % d);
    out[e] = (x[e] - norm[j]) * norm[d + j];
}

// Squared-error loss in normalised target units and its gradient.
__global__ void mse_kernel(const float *y, const float *t, int64_t n, float shift, float inv_scale,
                           float *dy, float *loss_part) {
    __shared__ float red[4];
    int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    float l = 0.f;
    if (i < n) {
        float d = y[i] - (t[i] - shift) * inv_scale;
        dy[i] = 2.f * d / float(n);
        l = d * d;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void adam_kernel(float *p, const float *g, float *m, float *v, int64_t n, float lr,
                            float b1, float b2, float c1, float c2, float eps) {
    int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.f - b1) * gi;
    float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---- host side ----
struct MlpRun {
    bfgpu_mlp &m;
    int64_t n;
    // workspace views
    float *x0, *y, *dy, *lossp;
    std::vector<float *> act, grad;  // per hidden layer output
    MlpRun(bfgpu_mlp &mm, int64_t nn) : m(mm), n(nn) {
        int L = m.n_layers;
        int64_t H = m.hidden;
        int64_t need = n * m.n_in + 2 * n + (n + 255) / 256 + 2 * int64_t(L - 1) * n * H + 64;
        m.ws.reserve(size_t(need) * sizeof(float));
        float *p = m.ws.as<float>();
        x0 = p; p += n * m.n_in;
        y = p; p += n;
        dy = p; p += n;
        lossp = p; p += (n + 255) / 256;
        for (int l = 0; l < L - 1; ++l) { act.push_back(p); p += n * H; }
        for (int l = 0; l < L - 1; ++l) { grad.push_back(p); p += n * H; }
    }
    const float *W(int l) const { return m.params.as<float>() + m.w_ofs[l]; }
    const float *b(int l) const { return m.params.as<float>() + m.b_ofs[l]; }
    void forward(hipStream_t s, const float *x_raw) {
        hipLaunchKernelGGL(normalize_kernel, dim3(unsigned((n * m.n_in + 255) / 256)), dim3(256), 0, s,
                           x_raw, m.norm.as<float>(), n, m.n_in, x0);
        BFG_HIP(hipGetLastError());
        const float *in = x0;
        for (int l = 0; l < m.n_layers; ++l) {
            int din = m.dims[l], dout = m.dims[l + 1];
            float *out = (l == m.n_layers - 1) ? y : act[l];
            if (l == m.n_layers - 1)
                gemm<0, 1, 1>(s, int(n), dout, din, in, din, W(l), din, out, dout, b(l));
            else
                gemm<0, 1, 2>(s, int(n), dout, din, in, din, W(l), din, out, dout, b(l));
            in = out;
        }
    }
    void backward(hipStream_t s) {
        float *G = m.grads.as<float>();
        const float *dz = dy;
        for (int l = m.n_layers - 1; l >= 0; --l) {
            int din = m.dims[l], dout = m.dims[l + 1];
            const float *ain = (l == 0) ? x0 : act[l - 1];
            // dW (dout x din) = dz^T ain
            gemm<1, 0, 0>(s, dout, din, int(n), dz, dout, ain, din, G + m.w_ofs[l], din);
            hipLaunchKernelGGL(colsum_kernel, dim3(unsigned(dout)), dim3(64), 0, s, dz, int(n), dout, dout,
                               G + m.b_ofs[l]);
            BFG_HIP(hipGetLastError());
            if (l > 0) {
                // dz_{l-1} = (dz W_l) * (act_{l-1} > 0)
                gemm<0, 0, 3>(s, int(n), din, dout, dz, dout, W(l), din, grad[l - 1], din, nullptr,
                              act[l - 1]);
                dz = grad[l - 1];
            }
        }
    }
};

void mlp_upload_norm(bfgpu_mlp &m) {
    std::vector<float> h(2 * m.n_in);
    for (int i = 0; i < m.n_in; ++i) {
        h[i] = m.mu[i];
        h[m.n_in + i] = 1.f / (m.sigma[i] > 0.f ? m.sigma[i] : 1.f);
    }
    BFG_HIP(hipMemcpy(m.norm.p, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
}

// Fused inference for the shape the rank predictor ships in: n_in <= 8
// inputs, hidden width 128, four layers (two hidden-to-hidden).  FP32 on
// the matrix cores (v_mfma_f32_16x16x4f32) with the transposed product
//     a_{l+1}^T (128 x 16 samples) = relu(W_l a_l^T + b_l),
// 16 samples per operand, two sample groups per wavefront.  The 16 x 16
// output block of a layer leaves lane (sample s = lane % 16, g = lane / 16)
// holding outputs 16 ob + 4 g + r (register r), which is the B operand of
// the next layer's step (ib = ob, r) for the input order
// k = 16 ib + 4 g + r -- activations never leave registers; the weights of
// the two hidden-to-hidden layers sit in LDS (128 KiB) in that order, one
// 16-byte read per lane feeding four steps.  Only the summation order
// differs from the layer-by-layer GEMMs.
constexpr int MI_H = 128, MI_NB = MI_H / 16, MI_S = 2, MI_THREADS = 512;
typedef float mi_f4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(MI_THREADS) void mlp_infer_kernel(const float *__restrict__ x, int64_t n, int n_in,
                                                               const float *__restrict__ norm,
                                                               const float *__restrict__ P, int64_t w0,
                                                               int64_t b0, int64_t w1, int64_t b1, int64_t w2,
                                                               int64_t b2, int64_t w3, int64_t b3,
                                                               float *__restrict__ y) {
    __shared__ mi_f4 Wm[2][MI_NB * MI_NB * 64];
    __shared__ float W0s[MI_NB * 2 * 64];
    __shared__ float bs[3][MI_H];
    __shared__ float w3s[MI_H];
    for (int e = threadIdx.x; e < MI_NB * MI_NB * 64; e += MI_THREADS) {
        const int ln = e & 63, ib = (e >> 6) % MI_NB, ob = (e >> 6) / MI_NB;
        const int64_t row = int64_t(ob * 16 + (ln & 15)) * MI_H + ib * 16 + 4 * (ln >> 4);
        Wm[0][e] = mi_f4{P[w1 + row], P[w1 + row + 1], P[w1 + row + 2], P[w1 + row + 3]};
        Wm[1][e] = mi_f4{P[w2 + row], P[w2 + row + 1], P[w2 + row + 2], P[w2 + row + 3]};
    }
    for (int e = threadIdx.x; e < MI_NB * 2 * 64; e += MI_THREADS) {
        const int ln = e & 63, t = (e >> 6) & 1, ob = e >> 7;
        const int k = 4 * t + (ln >> 4);
        W0s[e] = k < n_in ? P[w0 + int64_t(ob * 16 + (ln & 15)) * n_in + k] : 0.f;
    }
    for (int e = threadIdx.x; e < MI_H; e += MI_THREADS) {
        bs[0][e] = P[b0 + e];
        bs[1][e] = P[b1 + e];
        bs[2][e] = P[b2 + e];
        w3s[e] = P[w3 + e];
    }
    const float bias3 = P[b3];
    __syncthreads();
    const int lane = threadIdx.x & 63, s = lane & 15, g = lane >> 4;
    const int64_t nw = int64_t(gridDim.x) * (MI_THREADS / 64);
    const int64_t ntile = (n + 16 * MI_S - 1) / (16 * MI_S);
    for (int64_t tile = int64_t(blockIdx.x) * (MI_THREADS / 64) + (threadIdx.x >> 6); tile < ntile; tile += nw) {
        mi_f4 act[MI_S][MI_NB];
        float xin[MI_S][2];
#pragma unroll
        for (int q = 0; q < MI_S; ++q) {
            const int64_t smp = tile * (16 * MI_S) + q * 16 + s;
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int k = 4 * t + g;
                xin[q][t] = (smp < n && k < n_in) ? (x[smp * n_in + k] - norm[k]) * norm[n_in + k] : 0.f;
            }
        }
#pragma unroll
        for (int ob = 0; ob < MI_NB; ++ob) {
            const float wa = W0s[(ob * 2 + 0) * 64 + lane], wb = W0s[(ob * 2 + 1) * 64 + lane];
#pragma unroll
            for (int q = 0; q < MI_S; ++q) {
                mi_f4 c = {0.f, 0.f, 0.f, 0.f};
                c = __builtin_amdgcn_mfma_f32_16x16x4f32(wa, xin[q][0], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x4f32(wb, xin[q][1], c, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; ++r) act[q][ob][r] = fmaxf(c[r] + bs[0][ob * 16 + 4 * g + r], 0.f);
            }
        }
#pragma unroll
        for (int L = 0; L < 2; ++L) {
            mi_f4 nxt[MI_S][MI_NB];
#pragma unroll
            for (int q = 0; q < MI_S; ++q)
#pragma unroll
                for (int ob = 0; ob < MI_NB; ++ob) nxt[q][ob] = mi_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ib = 0; ib < MI_NB; ++ib) {
                mi_f4 w[MI_NB];
#pragma unroll
                for (int ob = 0; ob < MI_NB; ++ob) w[ob] = Wm[L][(ob * MI_NB + ib) * 64 + lane];
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int ob = 0; ob < MI_NB; ++ob)
#pragma unroll
                        for (int q = 0; q < MI_S; ++q)
                            nxt[q][ob] = __builtin_amdgcn_mfma_f32_16x16x4f32(w[ob][r], act[q][ib][r], nxt[q][ob],
                                                                              0, 0, 0);
                // keep the next block's weight reads from being hoisted
                // (registers: activations, accumulators, one block of weights)
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int q = 0; q < MI_S; ++q)
#pragma unroll
                for (int ob = 0; ob < MI_NB; ++ob)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        act[q][ob][r] = fmaxf(nxt[q][ob][r] + bs[L + 1][ob * 16 + 4 * g + r], 0.f);
        }
#pragma unroll
        for (int q = 0; q < MI_S; ++q) {
            float p = 0.f;
#pragma unroll
            for (int ob = 0; ob < MI_NB; ++ob)
#pragma unroll
                for (int r = 0; r < 4; ++r) p = fmaf(w3s[ob * 16 + 4 * g + r], act[q][ob][r], p);
            p += __shfl_xor(p, 16, 64);
            p += __shfl_xor(p, 32, 64);
            const int64_t smp = tile * (16 * MI_S) + q * 16 + s;
            if (g == 0 && smp < n) y[smp] = p + bias3;
        }
    }
}

static bool fused_infer_ok(const bfgpu_mlp &m) {
    return m.n_layers == 4 && m.hidden == MI_H && m.n_in >= 1 && m.n_in <= 8;
}

// Forward pass over n samples in chunks of at most INFER_CHUNK, so that the
// activation workspace stays bounded (n can be every compression of a
// level); raw network outputs into y (host).
constexpr int64_t INFER_CHUNK = int64_t(1) << 20;
static void forward_chunked(bfgpu_mlp &m, int64_t n, const float *features, float *y) {
    hipStream_t s = m.ctx->stream;
    const char *gemm_env = std::getenv("BFGPU_MLP_GEMM");  // 1: layer-by-layer GEMMs (for comparisons)
    if (fused_infer_ok(m) && !(gemm_env && std::atoi(gemm_env) != 0)) {
        // fused kernel: no activation workspace, chunks of 4M samples
        const int64_t ch = std::min(n, int64_t(4) * INFER_CHUNK);
        DevBuf dx, dy;
        dx.alloc(sizeof(float) * size_t(ch) * m.n_in);
        dy.alloc(sizeof(float) * size_t(ch));
        int dev = 0, n_cu = 0;
        BFG_HIP(hipGetDevice(&dev));
        BFG_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
        for (int64_t i0 = 0; i0 < n; i0 += ch) {
            const int64_t nc = std::min(ch, n - i0);
            BFG_HIP(hipMemcpyAsync(dx.p, features + i0 * m.n_in, sizeof(float) * size_t(nc) * m.n_in,
                                   hipMemcpyHostToDevice, s));
            const int64_t tiles = (nc + 16 * MI_S - 1) / (16 * MI_S);
            const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(n_cu, (tiles + 7) / 8));
            hipLaunchKernelGGL(mlp_infer_kernel, dim3(unsigned(grid)), dim3(MI_THREADS), 0, s, dx.as<float>(), nc,
                               int(m.n_in), m.norm.as<float>(), m.params.as<float>(), m.w_ofs[0], m.b_ofs[0],
                               m.w_ofs[1], m.b_ofs[1], m.w_ofs[2], m.b_ofs[2], m.w_ofs[3], m.b_ofs[3],
                               dy.as<float>());
            BFG_HIP(hipGetLastError());
            BFG_HIP(hipMemcpyAsync(y + i0, dy.p, sizeof(float) * size_t(nc), hipMemcpyDeviceToHost, s));
            BFG_HIP(hipStreamSynchronize(s));
        }
        return;
    }
    DevBuf dx;
    const int64_t ch = std::min(n, INFER_CHUNK);
    dx.alloc(sizeof(float) * size_t(ch) * m.n_in);
    for (int64_t i0 = 0; i0 < n; i0 += ch) {
        const int64_t nc = std::min(ch, n - i0);
        BFG_HIP(hipMemcpyAsync(dx.p, features + i0 * m.n_in, sizeof(float) * size_t(nc) * m.n_in,
                               hipMemcpyHostToDevice, s));
        MlpRun r(m, nc);
        r.forward(s, dx.as<float>());
        BFG_HIP(hipMemcpyAsync(y + i0, r.y, sizeof(float) * size_t(nc), hipMemcpyDeviceToHost, s));
        BFG_HIP(hipStreamSynchronize(s));
    }
}

int32_t mlp_predict_ranks(const bfgpu_mlp *mc, const float *feat, int64_t n, int32_t *out) {
    bfgpu_mlp &m = const_cast<bfgpu_mlp &>(*mc);
    if (n == 0) return 0;
    hvec<float> y(static_cast<size_t>(n));
    forward_chunked(m, n, feat, y.data());
    parallel_for(n, [&](int64_t i) {
        const float rk = y[size_t(i)] * m.out_scale + m.out_shift;
        out[size_t(i)] = std::max(1, int32_t(std::ceil(rk)));
    }, 1 << 16);
    return 0;
}

}  // namespace bfg

using namespace bfg;

#define MLP_API_BEGIN try {
#define MLP_API_END                               \
    }                                             \
    catch (const Error &e) {                      \
        set_last_error(e.what());                 \
        return e.code;                            \
    }                                             \
    catch (const std::exception &e) {             \
        set_last_error(e.what());                 \
        return BFGPU_ERR_STATE;                   \
    }                                             \
    return BFGPU_OK;

extern "C" int bfgpu_mlp_create(bfgpu_ctx *ctx, int32_t n_in, int32_t hidden, int32_t n_layers,
                                uint64_t seed, bfgpu_mlp **out) {
    MLP_API_BEGIN
    BFG_REQUIRE(ctx && out, "mlp_create: null argument");
    BFG_REQUIRE(n_in >= 1 && n_in <= 64 && hidden >= 1 && hidden <= 4096 && n_layers >= 2 && n_layers <= 16,
                "mlp_create: size out of range");
    auto *m = new bfgpu_mlp{};
    try {
        m->ctx = &ctx->c;
        m->n_in = n_in;
        m->hidden = hidden;
        m->n_layers = n_layers;
        m->dims.push_back(n_in);
        for (int l = 0; l < n_layers - 1; ++l) m->dims.push_back(hidden);
        m->dims.push_back(1);
        int64_t o = 0;
        for (int l = 0; l < n_layers; ++l) {
            m->w_ofs.push_back(o);
            o += int64_t(m->dims[l]) * m->dims[l + 1];
            m->b_ofs.push_back(o);
            o += m->dims[l + 1];
        }
        m->n_params = o;
        // He-uniform initialisation from a seeded generator
        std::mt19937_64 g(seed);
        std::vector<float> p(size_t(o), 0.f);
        for (int l = 0; l < n_layers; ++l) {
            float lim = std::sqrt(6.f / float(m->dims[l]));
            std::uniform_real_distribution<float> U(-lim, lim);
            for (int64_t i = 0; i < int64_t(m->dims[l]) * m->dims[l + 1]; ++i) p[m->w_ofs[l] + i] = U(g);
        }
        m->params.alloc(sizeof(float) * size_t(o));
        m->grads.alloc(sizeof(float) * size_t(o));
        m->adam_m.alloc(sizeof(float) * size_t(o));
        m->adam_v.alloc(sizeof(float) * size_t(o));
        BFG_HIP(hipMemcpy(m->params.p, p.data(), sizeof(float) * size_t(o), hipMemcpyHostToDevice));
        BFG_HIP(hipMemset(m->adam_m.p, 0, sizeof(float) * size_t(o)));
        BFG_HIP(hipMemset(m->adam_v.p, 0, sizeof(float) * size_t(o)));
        m->mu.assign(n_in, 0.f);
        m->sigma.assign(n_in, 1.f);
        m->norm.alloc(sizeof(float) * 2 * size_t(n_in));
        mlp_upload_norm(*m);
    } catch (...) {
        delete m;
        throw;
    }
    *out = m;
    MLP_API_END
}

extern "C" int bfgpu_mlp_destroy(bfgpu_mlp *m) {
    delete m;
    return BFGPU_OK;
}

extern "C" int bfgpu_mlp_n_params(const bfgpu_mlp *m, int64_t *n) {
    MLP_API_BEGIN
    BFG_REQUIRE(m && n, "mlp_n_params: null argument");
    *n = m->n_params;
    MLP_API_END
}

extern "C" int bfgpu_mlp_get_params(const bfgpu_mlp *m, float *p) {
    MLP_API_BEGIN
    BFG_REQUIRE(m && p, "mlp_get_params: null argument");
    BFG_HIP(hipStreamSynchronize(m->ctx->stream));
    BFG_HIP(hipMemcpy(p, m->params.p, sizeof(float) * size_t(m->n_params), hipMemcpyDeviceToHost));
    MLP_API_END
}

extern "C" int bfgpu_mlp_set_params(bfgpu_mlp *m, const float *p) {
    MLP_API_BEGIN
    BFG_REQUIRE(m && p, "mlp_set_params: null argument");
    BFG_HIP(hipMemcpy(m->params.p, p, sizeof(float) * size_t(m->n_params), hipMemcpyHostToDevice));
    MLP_API_END
}

extern "C" int bfgpu_mlp_set_norm(bfgpu_mlp *m, const float *mu, const float *sigma, float out_shift,
                                  float out_scale) {
    MLP_API_BEGIN
    BFG_REQUIRE(m && mu && sigma, "mlp_set_norm: null argument");
    BFG_REQUIRE(out_scale != 0.f, "mlp_set_norm: out_scale must be non-zero");
    m->mu.assign(mu, mu + m->n_in);
    m->sigma.assign(sigma, sigma + m->n_in);
    m->out_shift = out_shift;
    m->out_scale = out_scale;
    mlp_upload_norm(*m);
    MLP_API_END
}

extern "C" int bfgpu_mlp_forward(bfgpu_mlp *m, int64_t n, const float *features, float *out) {
    MLP_API_BEGIN
    BFG_REQUIRE(m && features && out && n >= 0, "mlp_forward: bad argument");
    BFG_REQUIRE(n < (int64_t(1) << 30), "mlp_forward: too many samples");
    if (n == 0) return BFGPU_OK;
    forward_chunked(*m, n, features, out);
    for (int64_t i = 0; i < n; ++i) out[i] = out[i] * m->out_scale + m->out_shift;
    MLP_API_END
}

static float loss_grad_dev(bfgpu_mlp &m, int64_t n, const float *dx, const float *dt) {
    hipStream_t s = m.ctx->stream;
    MlpRun r(m, n);
    r.forward(s, dx);
    int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(mse_kernel, dim3(unsigned(nb)), dim3(256), 0, s, r.y, dt, n, m.out_shift,
                       1.f / m.out_scale, r.dy, r.lossp);
    BFG_HIP(hipGetLastError());
    r.backward(s);
    std::vector<float> lp(static_cast<size_t>(nb));
    BFG_HIP(hipMemcpyAsync(lp.data(), r.lossp, sizeof(float) * size_t(nb), hipMemcpyDeviceToHost, s));
    BFG_HIP(hipStreamSynchronize(s));
    double l = 0.0;
    for (float v : lp) l += v;
    return float(l / double(n));
}

extern "C" int bfgpu_mlp_loss_grad(bfgpu_mlp *m, int64_t n, const float *features, const float *target,
                                   float *loss, float *grads) {
    MLP_API_BEGIN
    BFG_REQUIRE(m && features && target && n >= 1, "mlp_loss_grad: bad argument");
    BFG_REQUIRE(n < (int64_t(1) << 30), "mlp_loss_grad: too many samples");
    hipStream_t s = m->ctx->stream;
    DevBuf dx, dt;
    dx.alloc(sizeof(float) * size_t(n) * m->n_in);
    dt.alloc(sizeof(float) * size_t(n));
    BFG_HIP(hipMemcpyAsync(dx.p, features, sizeof(float) * size_t(n) * m->n_in, hipMemcpyHostToDevice, s));
    BFG_HIP(hipMemcpyAsync(dt.p, target, sizeof(float) * size_t(n), hipMemcpyHostToDevice, s));
    float l = loss_grad_dev(*m, n, dx.as<float>(), dt.as<float>());
    if (loss) *loss = l;
    if (grads)
        BFG_HIP(hipMemcpy(grads, m->grads.p, sizeof(float) * size_t(m->n_params), hipMemcpyDeviceToHost));
    MLP_API_END
}

extern "C" int bfgpu_mlp_train(bfgpu_mlp *m, int64_t n, const float *features, const float *target,
                               int32_t epochs, int32_t batch, float lr, uint64_t seed, float *final_loss) {
    MLP_API_BEGIN
    BFG_REQUIRE(m && features && target && n >= 1 && epochs >= 1 && batch >= 1, "mlp_train: bad argument");
    BFG_REQUIRE(n < (int64_t(1) << 30), "mlp_train: too many samples");
    hipStream_t s = m->ctx->stream;
    const int d = m->n_in;
    std::vector<int64_t> idx(static_cast<size_t>(n));
    std::iota(idx.begin(), idx.end(), 0);
    std::mt19937_64 g(seed);
    std::vector<float> fx(static_cast<size_t>(n) * d), ft(static_cast<size_t>(n));
    DevBuf dx, dt;
    dx.alloc(sizeof(float) * fx.size());
    dt.alloc(sizeof(float) * ft.size());
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float last = 0.f;
    for (int e = 0; e < epochs; ++e) {
        std::shuffle(idx.begin(), idx.end(), g);
        for (int64_t i = 0; i < n; ++i) {
            std::memcpy(&fx[size_t(i) * d], features + idx[i] * d, sizeof(float) * d);
            ft[size_t(i)] = target[idx[i]];
        }
        BFG_HIP(hipMemcpyAsync(dx.p, fx.data(), sizeof(float) * fx.size(), hipMemcpyHostToDevice, s));
        BFG_HIP(hipMemcpyAsync(dt.p, ft.data(), sizeof(float) * ft.size(), hipMemcpyHostToDevice, s));
        double el = 0.0;
        int64_t nbat = 0;
        for (int64_t b0 = 0; b0 < n; b0 += batch) {
            int64_t nb = std::min<int64_t>(batch, n - b0);
            float l = loss_grad_dev(*m, nb, dx.as<float>() + b0 * d, dt.as<float>() + b0);
            ++m->adam_t;
            float c1 = 1.f - std::pow(b1, float(m->adam_t)), c2 = 1.f - std::pow(b2, float(m->adam_t));
            hipLaunchKernelGGL(adam_kernel, dim3(unsigned((m->n_params + 255) / 256)), dim3(256), 0, s,
                               m->params.as<float>(), m->grads.as<float>(), m->adam_m.as<float>(),
                               m->adam_v.as<float>(), m->n_params, lr, b1, b2, c1, c2, eps);
            BFG_HIP(hipGetLastError());
            el += l;
            ++nbat;
        }
        last = float(el / double(nbat));
    }
    BFG_HIP(hipStreamSynchronize(s));
    if (final_loss) *final_loss = last;
    MLP_API_END
}
