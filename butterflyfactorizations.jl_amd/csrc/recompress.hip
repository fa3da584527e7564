// Algebraic recompression of butterfly factorizations
// (src/Butterflyalgebra/algrecomp.jl:19-148).  Right recompression: from
// the last R level down to the second, the blocks reading one skeleton
// vector are stacked, A = [B_1; B_2; ...] = Q T by column-pivoted QR
// truncated at tau (T = R[:, invperm(p)]), each B_j becomes its slice of
// the thin Q, and the blocks of the level before that write the vector
// become T * B (T for an identity block).  Left recompression is the right
// one of the adjoint (levels reversed, keys swapped, blocks conjugate-
// transposed), taken back.  Q and P are not touched.  All factorizations of
// an operator run together; the stacking, QR, forming Q, T and the block
// products are GPU kernels, the grouping by keys is host bookkeeping.
#include <algorithm>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "device.hpp"
#include "internal.hpp"

namespace bfg {
void set_last_error(const std::string &m);
void pack_tables(BF &bf);

namespace {

// copy of one m x n block (column-major, leading dimension ld_src) to dst
// (leading dimension ld_dst), or of its conjugate transpose (conj_t 1) or
// transpose (conj_t 2), n x m
struct CopyTask {
    const void *src;
    void *dst;
    int32_t m, n, ld_src, ld_dst;
    int32_t conj_t, pad;
    int64_t end;  // running entry count
};

template <typename R>
__global__ __launch_bounds__(256) void copy_blocks_kernel(const CopyTask *__restrict__ tasks, int32_t n_tasks,
                                                          int64_t total) {
    int32_t lo = 0;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += int64_t(gridDim.x) * 256) {
        int32_t hi = n_tasks - 1;
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (tasks[mid].end > e) hi = mid;
            else lo = mid + 1;
        }
        const CopyTask t = tasks[lo];
        const int64_t local = e - (t.end - int64_t(t.m) * t.n);
        const int32_t i = int32_t(local % t.m), j = int32_t(local / t.m);
        const C<R> v = static_cast<const C<R> *>(t.src)[i + int64_t(j) * t.ld_src];
        if (t.conj_t == 1) static_cast<C<R> *>(t.dst)[j + int64_t(i) * t.ld_dst] = C<R>{v.x, -v.y};
        else if (t.conj_t == 2) static_cast<C<R> *>(t.dst)[j + int64_t(i) * t.ld_dst] = v;
        else static_cast<C<R> *>(t.dst)[i + int64_t(j) * t.ld_dst] = v;
    }
}

// one stacked matrix after pivoted QR: m x n at zofs (ld m), rank r,
// pivots and reflector scalars at pofs; T (r x n, ld r) at tofs
struct QTask {
    int64_t zofs, tofs;
    int32_t m, n, r, pofs;
};

// T[:, piv[j]] = R[0:r, j] (upper trapezoid of the factored matrix)
template <typename R>
__global__ __launch_bounds__(256) void t_extract_kernel(const C<R> *Z, const QTask *__restrict__ tasks,
                                                        const int32_t *__restrict__ piv_all, C<R> *T) {
    const QTask tk = tasks[blockIdx.x];
    const C<R> *A = Z + tk.zofs;
    const int32_t *piv = piv_all + tk.pofs;
    C<R> *out = T + tk.tofs;
    for (int64_t e = threadIdx.x; e < int64_t(tk.r) * tk.n; e += 256) {
        const int i = int(e % tk.r), j = int(e / tk.r);
        out[i + int64_t(piv[j]) * tk.r] = i <= j ? A[i + int64_t(j) * tk.m] : C<R>{R(0), R(0)};
    }
}

// Thin Q (m x r) formed in place in the first r columns from the
// reflectors below the diagonal and their scalars (as LAPACK ZUNG2R:
// Q = H_1 ... H_r [I; 0], H_k = I - tau_k v_k v_k^H, v_k(k) = 1).  Sums in
// FP64; one wave per column of the update.
template <typename R>
__global__ __launch_bounds__(256) void orgqr_kernel(C<R> *Z, const QTask *__restrict__ tasks,
                                                    const double *__restrict__ tau_all) {
    const QTask tk = tasks[blockIdx.x];
    const int m = tk.m, r = tk.r;
    C<R> *A = Z + tk.zofs;
    const double *tau = tau_all + 2 * int64_t(tk.pofs);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int k = r - 1; k >= 0; --k) {
        const double tr = tau[2 * k], ti = tau[2 * k + 1];
        const C<R> *v = A + int64_t(k) * m;
        for (int j = k + 1 + wave; j < r; j += 4) {
            C<R> *c = A + int64_t(j) * m;
            double wx = 0.0, wy = 0.0;
            for (int i = k + lane; i < m; i += 64) {
                const double vx = i == k ? 1.0 : double(v[i].x), vy = i == k ? 0.0 : double(v[i].y);
                const double cx = c[i].x, cy = c[i].y;
                wx += vx * cx + vy * cy;  // conj(v) c
                wy += vx * cy - vy * cx;
            }
            wx = wave_sum(wx);
            wy = wave_sum(wy);
            const double ux = tr * wx - ti * wy, uy = tr * wy + ti * wx;  // tau w
            for (int i = k + lane; i < m; i += 64) {
                const double vx = i == k ? 1.0 : double(v[i].x), vy = i == k ? 0.0 : double(v[i].y);
                c[i].x = R(double(c[i].x) - (vx * ux - vy * uy));
                c[i].y = R(double(c[i].y) - (vx * uy + vy * ux));
            }
        }
        __syncthreads();  // column k's reflector is read above, overwritten below
        C<R> *ck = A + int64_t(k) * m;
        for (int i = tid; i < m; i += 256) {
            if (i < k) {
                ck[i] = C<R>{R(0), R(0)};
            } else if (i == k) {
                ck[i] = C<R>{R(1.0 - tr), R(-ti)};
            } else {
                const double vx = ck[i].x, vy = ck[i].y;
                ck[i] = C<R>{R(-(tr * vx - ti * vy)), R(-(tr * vy + ti * vx))};
            }
        }
        __syncthreads();
    }
}

// C (r x nb) = A (r x n) B (n x nb); B null: C = A (A times identity)
struct GemmTask {
    const void *a, *b;
    void *c;
    int32_t r, n, nb, pad;
    int64_t end;  // running output-entry count
};

template <typename R>
__global__ __launch_bounds__(256) void gemm_blocks_kernel(const GemmTask *__restrict__ tasks, int32_t n_tasks,
                                                          int64_t total) {
    int32_t lo = 0;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += int64_t(gridDim.x) * 256) {
        int32_t hi = n_tasks - 1;
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (tasks[mid].end > e) hi = mid;
            else lo = mid + 1;
        }
        const GemmTask t = tasks[lo];
        const int64_t local = e - (t.end - int64_t(t.r) * t.nb);
        const int32_t i = int32_t(local % t.r), j = int32_t(local / t.r);
        const C<R> *a = static_cast<const C<R> *>(t.a);
        C<R> acc{R(0), R(0)};
        if (!t.b) {
            acc = a[i + int64_t(j) * t.r];
        } else {
            const C<R> *b = static_cast<const C<R> *>(t.b) + int64_t(j) * t.n;
            for (int k = 0; k < t.n; ++k) cfma(acc, a[i + int64_t(k) * t.r], b[k]);
        }
        static_cast<C<R> *>(t.c)[i + int64_t(j) * t.r] = acc;
    }
}

template <class T> void upload_vec(DevBuf &d, const std::vector<T> &h, hipStream_t s) {
    d.reserve(sizeof(T) * std::max<size_t>(h.size(), 1));
    if (!h.empty()) BFG_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
}

unsigned stride_grid(int64_t total) {
    return unsigned(std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, int64_t(1) << 22)));
}

// A block of an R level during recompression; data is device memory.
struct RB {
    int32_t oo, so, oi, si, m, n;
    bool ident;
    void *data;
};
using Levels = std::vector<std::vector<RB>>;  // R levels of one factorization

struct Recompressor {
    hipStream_t s;
    int prec;
    size_t es;
    std::vector<DevBuf> store;  // device memory of the blocks written so far

    void *alloc(size_t bytes) {
        store.emplace_back();
        store.back().alloc(std::max<size_t>(bytes, 16));
        return store.back().p;
    }

    void copy(const std::vector<CopyTask> &t0) {
        if (t0.empty()) return;
        std::vector<CopyTask> t = t0;
        int64_t e = 0;
        for (CopyTask &c : t) c.end = (e += int64_t(c.m) * c.n);
        if (e == 0) return;
        DevBuf d;
        upload_vec(d, t, s);
        if (prec == BFGPU_C128)
            hipLaunchKernelGGL(copy_blocks_kernel<double>, dim3(stride_grid(e)), dim3(256), 0, s,
                               d.as<const CopyTask>(), int32_t(t.size()), e);
        else
            hipLaunchKernelGGL(copy_blocks_kernel<float>, dim3(stride_grid(e)), dim3(256), 0, s,
                               d.as<const CopyTask>(), int32_t(t.size()), e);
        BFG_HIP(hipGetLastError());
        BFG_HIP(hipStreamSynchronize(s));
    }

    void gemm(std::vector<GemmTask> &t) {
        int64_t e = 0;
        for (GemmTask &g : t) g.end = (e += int64_t(g.r) * g.nb);
        if (e == 0) return;
        DevBuf d;
        upload_vec(d, t, s);
        if (prec == BFGPU_C128)
            hipLaunchKernelGGL(gemm_blocks_kernel<double>, dim3(stride_grid(e)), dim3(256), 0, s,
                               d.as<const GemmTask>(), int32_t(t.size()), e);
        else
            hipLaunchKernelGGL(gemm_blocks_kernel<float>, dim3(stride_grid(e)), dim3(256), 0, s,
                               d.as<const GemmTask>(), int32_t(t.size()), e);
        BFG_HIP(hipGetLastError());
        BFG_HIP(hipStreamSynchronize(s));
    }

    // The R levels of the adjoint: levels reversed, keys swapped, blocks
    // conjugate-transposed (bfadjtr.jl:13-58).
    void adjoint(std::vector<Levels> &all) {
        std::vector<CopyTask> ct;
        size_t bytes = 0;
        for (const Levels &R : all)
            for (const auto &lv : R)
                for (const RB &b : lv)
                    if (!b.ident) bytes += size_t(b.m) * b.n * es;
        char *base = static_cast<char *>(alloc(bytes));
        size_t off = 0;
        for (Levels &R : all) {
            Levels out(R.size());
            for (size_t l = 0; l < R.size(); ++l) {
                for (const RB &b : R[R.size() - 1 - l]) {
                    RB a{b.si, b.oi, b.so, b.oo, b.n, b.m, b.ident, nullptr};
                    if (!b.ident) {
                        a.data = base + off;
                        off += size_t(b.m) * b.n * es;
                        ct.push_back(CopyTask{b.data, a.data, b.m, b.n, b.m, b.n, 1, 0, 0});
                    }
                    out[l].push_back(a);
                }
            }
            R.swap(out);
        }
        copy(ct);
    }

    // recompress_BF_right (algrecomp.jl:46-148) over every factorization.
    void right(std::vector<Levels> &all, double tau) {
        int32_t max_lr = 0;
        for (const Levels &R : all) max_lr = std::max<int32_t>(max_lr, int32_t(R.size()));
        for (int32_t t = 0; t + 2 <= max_lr; ++t) {
            // groups: blocks of one level reading one vector, as (bf, key)
            struct Group {
                int32_t bf, lold;
                std::vector<int32_t> idx;  // blocks of R[lold]
                int32_t rows = 0, n = 0;
                int64_t zofs = 0;
            };
            std::vector<Group> groups;
            for (int32_t b = 0; b < int32_t(all.size()); ++b) {
                const Levels &R = all[b];
                const int32_t lold = int32_t(R.size()) - 1 - t;
                if (lold < 1) continue;
                std::map<std::pair<int32_t, int32_t>, std::vector<int32_t>> by_col;
                for (int32_t i = 0; i < int32_t(R[lold].size()); ++i)
                    by_col[{R[lold][i].oi, R[lold][i].si}].push_back(i);
                for (auto &kv : by_col) {
                    bool ident = false;
                    for (int32_t i : kv.second) ident |= R[lold][i].ident;
                    if (ident) continue;  // the reference cannot stack an identity block
                    Group g;
                    g.bf = b;
                    g.lold = lold;
                    g.idx = kv.second;
                    g.n = R[lold][g.idx[0]].n;
                    for (int32_t i : g.idx) {
                        BFG_REQUIRE(R[lold][i].n == g.n, "recompress: blocks reading one vector differ in width");
                        g.rows += R[lold][i].m;
                    }
                    groups.push_back(std::move(g));
                }
            }
            if (groups.empty()) continue;
            // stack
            int64_t zt = 0, pt = 0;
            std::vector<QrTask> qt;
            std::vector<CopyTask> ct;
            for (Group &g : groups) {
                g.zofs = zt;
                zt += int64_t(g.rows) * g.n;
            }
            DevBuf dZ;
            dZ.alloc(size_t(std::max<int64_t>(zt, 1)) * es);
            char *Zb = static_cast<char *>(dZ.p);
            for (Group &g : groups) {
                const Levels &R = all[g.bf];
                int32_t row = 0;
                for (int32_t i : g.idx) {
                    const RB &b = R[g.lold][i];
                    ct.push_back(CopyTask{b.data, Zb + (size_t(g.zofs) + row) * es, b.m, b.n, b.m, g.rows, 0, 0, 0});
                    row += b.m;
                }
                qt.push_back(QrTask{g.zofs, g.rows, g.n, int32_t(pt), 0});
                pt += g.n;
                BFG_REQUIRE(pt < (int64_t(1) << 31), "recompress: too many columns in one step");
            }
            copy(ct);
            // pivoted QR, one launch per kernel class
            DevBuf dpiv, drank, dtau, dq;
            dpiv.alloc(sizeof(int32_t) * size_t(std::max<int64_t>(pt, 1)));
            dtau.alloc(2 * sizeof(double) * size_t(std::max<int64_t>(pt, 1)));
            BFG_HIP(hipMemsetAsync(dtau.p, 0, dtau.bytes, s));
            std::vector<int32_t> rank(qt.size(), 0);
            for (int cls = 0; cls < 2; ++cls) {
                std::vector<QrTask> sub;
                std::vector<int32_t> where;
                int32_t mm = 1, mn = 1;
                for (size_t q = 0; q < qt.size(); ++q)
                    if ((cpqr_wave_ok(prec, qt[q].m, qt[q].n) ? 1 : 0) == cls) {
                        sub.push_back(qt[q]);
                        where.push_back(int32_t(q));
                        mm = std::max(mm, qt[q].m);
                        mn = std::max(mn, qt[q].n);
                    }
                if (sub.empty()) continue;
                upload_vec(dq, sub, s);
                drank.reserve(sizeof(int32_t) * sub.size());
                launch_cpqr(s, prec, dZ.p, dq.as<QrTask>(), int32_t(sub.size()), dpiv.as<int32_t>(),
                            drank.as<int32_t>(), tau, mn, mm, dtau.as<double>());
                std::vector<int32_t> r(sub.size());
                BFG_HIP(hipMemcpyAsync(r.data(), drank.p, sizeof(int32_t) * r.size(), hipMemcpyDeviceToHost, s));
                BFG_HIP(hipStreamSynchronize(s));
                for (size_t q = 0; q < where.size(); ++q) rank[where[q]] = r[q];
            }
            // T (r x n) per group, then Q in place
            std::vector<QTask> tt(groups.size());
            int64_t tofs = 0;
            for (size_t q = 0; q < groups.size(); ++q) {
                tt[q] = QTask{groups[q].zofs, tofs, groups[q].rows, groups[q].n, rank[q], qt[q].pofs};
                tofs += int64_t(rank[q]) * groups[q].n;
            }
            char *Tb = static_cast<char *>(alloc(size_t(std::max<int64_t>(tofs, 1)) * es));
            DevBuf dtt;
            upload_vec(dtt, tt, s);
            for (int32_t q0 = 0; q0 < int32_t(tt.size()); q0 += TASK_CHUNK) {
                const int32_t nq = std::min<int32_t>(TASK_CHUNK, int32_t(tt.size()) - q0);
                if (prec == BFGPU_C128) {
                    hipLaunchKernelGGL(t_extract_kernel<double>, dim3(nq), dim3(256), 0, s, dZ.as<const C<double>>(),
                                       dtt.as<const QTask>() + q0, dpiv.as<const int32_t>(), (C<double> *)Tb);
                    hipLaunchKernelGGL(orgqr_kernel<double>, dim3(nq), dim3(256), 0, s, dZ.as<C<double>>(),
                                       dtt.as<const QTask>() + q0, dtau.as<const double>());
                } else {
                    hipLaunchKernelGGL(t_extract_kernel<float>, dim3(nq), dim3(256), 0, s, dZ.as<const C<float>>(),
                                       dtt.as<const QTask>() + q0, dpiv.as<const int32_t>(), (C<float> *)Tb);
                    hipLaunchKernelGGL(orgqr_kernel<float>, dim3(nq), dim3(256), 0, s, dZ.as<C<float>>(),
                                       dtt.as<const QTask>() + q0, dtau.as<const double>());
                }
                BFG_HIP(hipGetLastError());
            }
            BFG_HIP(hipStreamSynchronize(s));
            // blocks of R[lold]: their slices of Q; blocks of R[lold - 1]
            // writing the vector: T times their data
            size_t qbytes = 0, ubytes = 0;
            std::vector<std::map<std::pair<int32_t, int32_t>, int32_t>> t_of(all.size());
            for (size_t q = 0; q < groups.size(); ++q) {
                const Group &g = groups[q];
                qbytes += size_t(g.rows) * rank[q] * es;
                const RB &b0 = all[g.bf][g.lold][g.idx[0]];
                t_of[g.bf][{b0.oi, b0.si}] = int32_t(q);
            }
            for (size_t b = 0; b < all.size(); ++b) {
                const int32_t lold = int32_t(all[b].size()) - 1 - t;
                if (lold < 1) continue;
                for (const RB &w : all[b][lold - 1]) {
                    auto it = t_of[b].find({w.oo, w.so});
                    if (it != t_of[b].end()) ubytes += size_t(rank[it->second]) * w.n * es;
                }
            }
            char *Qn = static_cast<char *>(alloc(qbytes));
            char *Un = static_cast<char *>(alloc(ubytes));
            size_t qo = 0, uo = 0;
            ct.clear();
            std::vector<GemmTask> gt;
            for (size_t q = 0; q < groups.size(); ++q) {
                Group &g = groups[q];
                const int32_t r = rank[q];
                int32_t row = 0;
                for (int32_t i : g.idx) {
                    RB &b = all[g.bf][g.lold][i];
                    void *dst = Qn + qo;
                    ct.push_back(CopyTask{Zb + (size_t(g.zofs) + row) * es, dst, b.m, r, g.rows, b.m, 0, 0, 0});
                    qo += size_t(b.m) * r * es;
                    row += b.m;
                    b.n = r;
                    b.data = dst;
                }
            }
            for (size_t b = 0; b < all.size(); ++b) {
                const int32_t lold = int32_t(all[b].size()) - 1 - t;
                if (lold < 1) continue;
                for (RB &w : all[b][lold - 1]) {
                    auto it = t_of[b].find({w.oo, w.so});
                    if (it == t_of[b].end()) continue;
                    const int32_t q = it->second;
                    const int32_t r = rank[q];
                    BFG_REQUIRE(w.m == groups[q].n, "recompress: writer and reader of a vector differ in length");
                    void *dst = Un + uo;
                    uo += size_t(r) * w.n * es;
                    gt.push_back(GemmTask{Tb + size_t(tt[q].tofs) * es, w.ident ? nullptr : w.data, dst, r, w.m,
                                          w.n, 0, 0});
                    w.m = r;
                    w.ident = false;
                    w.data = dst;
                }
            }
            copy(ct);
            gemm(gt);
        }
    }
};

}  // namespace

// The recompressed operator: a new BF over the same trees and precision.
void recompress(const BF &src, double tau, BF &out) {
    const size_t es = src.precision == BFGPU_C128 ? 16 : 8;
    Recompressor rc{src.ctx->stream, src.precision, es, {}};
    const int32_t nL = src.n_stage_launch;
    auto level_of = [&](const OneBF &ob, size_t s) { return s + 1 == ob.st.size() ? nL - 1 : int32_t(s); };
    // R levels of every factorization, pointing into the source buffers
    std::vector<Levels> all(src.bfs.size());
    for (size_t b = 0; b < src.bfs.size(); ++b) {
        const OneBF &ob = src.bfs[b];
        all[b].resize(ob.st.size() >= 2 ? ob.st.size() - 2 : 0);
        for (size_t s = 1; s + 1 < ob.st.size(); ++s)
            for (const Block &k : ob.st[s].b)
                all[b][s - 1].push_back(RB{k.obs_out, k.src_out, k.obs_in, k.src_in, k.m, k.n, k.ident != 0,
                                           k.ident ? nullptr
                                                   : static_cast<char *>(src.factors[level_of(ob, s)].p) +
                                                         size_t(k.fofs) * es});
    }
    rc.right(all, tau);
    rc.adjoint(all);
    rc.right(all, tau);
    rc.adjoint(all);
    // the new operator: Q and P copied, R levels from the recompressed blocks
    out.ctx = src.ctx;
    out.test = src.test;
    out.trial = src.trial;
    out.precision = src.precision;
    out.n_stage_launch = nL;
    out.n_comp = src.n_comp;
    out.n_retry = src.n_retry;
    out.n_judged = src.n_judged;
    out.n_fallback = src.n_fallback;
    out.bfs = src.bfs;  // Q and P unchanged: P still holds kernel entries
    out.pf_k = src.pf_k;
    out.use_pfree = src.use_pfree;
    out.factor_elems.assign(size_t(nL), 0);
    struct Put {  // a block's data and its place in the new level buffer
        const void *from;
        int32_t L, m, n;
        int64_t fofs;
    };
    std::vector<Put> puts;
    for (size_t b = 0; b < out.bfs.size(); ++b) {
        OneBF &ob = out.bfs[b];
        for (size_t s = 0; s < ob.st.size(); ++s) {
            const int32_t L = level_of(ob, s);
            std::vector<Block> &blk = ob.st[s].b;
            const bool isR = s >= 1 && s + 1 < ob.st.size();
            for (size_t i = 0; i < blk.size(); ++i) {
                Block &k = blk[i];
                const void *from = nullptr;
                if (isR) {
                    const RB &r = all[b][s - 1][i];
                    BFG_REQUIRE(r.oo == k.obs_out && r.so == k.src_out && r.oi == k.obs_in && r.si == k.src_in,
                                "recompress: block order changed");
                    k.m = r.m;
                    k.n = r.n;
                    k.ident = r.ident ? 1 : 0;
                    from = r.data;
                } else if (!k.ident) {
                    from = static_cast<const char *>(src.factors[L].p) + size_t(k.fofs) * es;
                }
                if (k.ident) {
                    k.fofs = -1;
                    continue;
                }
                k.fofs = out.factor_elems[L];
                out.factor_elems[L] += int64_t(k.m) * k.n;
                puts.push_back(Put{from, L, k.m, k.n, k.fofs});
            }
        }
        ob.stage_data_base.assign(ob.st.size(), 0);
    }
    out.factors.resize(size_t(nL));
    for (int32_t L = 0; L < nL; ++L) {
        out.factors[L].alloc(size_t(out.factor_elems[L] + FACTOR_PAD) * es);
        BFG_HIP(hipMemsetAsync(static_cast<char *>(out.factors[L].p) + size_t(out.factor_elems[L]) * es, 0,
                               size_t(FACTOR_PAD) * es, rc.s));
    }
    std::vector<CopyTask> ct;
    ct.reserve(puts.size());
    for (const Put &p : puts)
        ct.push_back(CopyTask{p.from, static_cast<char *>(out.factors[p.L].p) + size_t(p.fofs) * es, p.m, p.n, p.m,
                              p.m, 0, 0, 0});
    rc.copy(ct);
    rc.store.clear();
    pack_tables(out);
}

// The adjoint (conj) or transpose of every factorization as a factorization
// of its own over the swapped trees (bfadjtr.jl:43-89): stages reversed,
// keys swapped (transform_block, :13-24), blocks conjugate-transposed or
// transposed, identity blocks kept.  Applying it forward equals applying the
// source through its transposed launch tables.
void materialize_transpose(const BF &src, bool conj, BF &out) {
    const size_t es = src.precision == BFGPU_C128 ? 16 : 8;
    Recompressor rc{src.ctx->stream, src.precision, es, {}};
    const int32_t nL = src.n_stage_launch;
    auto level_of = [&](size_t ns, size_t s) { return s + 1 == ns ? nL - 1 : int32_t(s); };
    out.ctx = src.ctx;
    out.test = src.trial;
    out.trial = src.test;
    out.precision = src.precision;
    out.n_stage_launch = nL;
    out.n_comp = src.n_comp;
    out.n_retry = src.n_retry;
    out.n_judged = src.n_judged;
    out.n_fallback = src.n_fallback;
    out.bfs.resize(src.bfs.size());
    out.factor_elems.assign(size_t(nL), 0);
    struct Put {
        const void *from;
        int32_t L, m, n;  // source block m x n
        int64_t fofs;     // destination offset in level L
    };
    std::vector<Put> puts;
    for (size_t b = 0; b < src.bfs.size(); ++b) {
        const OneBF &ob = src.bfs[b];
        OneBF &a = out.bfs[b];
        const size_t ns = ob.st.size();
        a.no = ob.ns;
        a.ns = ob.no;
        a.st.resize(ns);
        a.stage_data_base.assign(ns, 0);
        for (size_t sp = 0; sp < ns; ++sp) {
            const size_t s = ns - 1 - sp;
            const int32_t Ls = level_of(ns, s), Ld = level_of(ns, sp);
            for (const Block &k : ob.st[s].b) {
                Block t = k;
                t.obs_out = k.src_in;
                t.src_out = k.obs_in;
                t.obs_in = k.src_out;
                t.src_in = k.obs_out;
                t.m = k.n;
                t.n = k.m;
                t.out_ptr = t.in_ptr = -1;
                if (k.ident) {
                    t.fofs = -1;
                } else {
                    t.fofs = out.factor_elems[Ld];
                    out.factor_elems[Ld] += int64_t(k.m) * k.n;
                    puts.push_back(Put{static_cast<const char *>(src.factors[Ls].p) + size_t(k.fofs) * es, Ld,
                                       k.m, k.n, t.fofs});
                }
                a.st[sp].b.push_back(t);
            }
        }
    }
    out.factors.resize(size_t(nL));
    for (int32_t L = 0; L < nL; ++L) {
        out.factors[L].alloc(size_t(out.factor_elems[L] + FACTOR_PAD) * es);
        BFG_HIP(hipMemsetAsync(static_cast<char *>(out.factors[L].p) + size_t(out.factor_elems[L]) * es, 0,
                               size_t(FACTOR_PAD) * es, rc.s));
    }
    // copies in bounded batches (the task list of a large operator is big)
    std::vector<CopyTask> ct;
    constexpr size_t BATCH = size_t(1) << 22;
    for (size_t i = 0; i < puts.size(); ++i) {
        const Put &p = puts[i];
        ct.push_back(CopyTask{p.from, static_cast<char *>(out.factors[p.L].p) + size_t(p.fofs) * es, p.m, p.n, p.m,
                              p.n, conj ? 1 : 2, 0, 0});
        if (ct.size() == BATCH || i + 1 == puts.size()) {
            rc.copy(ct);
            ct.clear();
        }
    }
    pack_tables(out);
}

}  // namespace bfg

using namespace bfg;

extern "C" int bfgpu_bf_materialize(bfgpu_bf *h, int32_t op, bfgpu_bf **out) {
    try {
        BFG_REQUIRE(h && out, "bf_materialize: null argument");
        BFG_REQUIRE(op == BFGPU_OP_T || op == BFGPU_OP_C, "bf_materialize: op must be BFGPU_OP_T or BFGPU_OP_C");
        BFG_HIP(hipSetDevice(h->b.ctx->device));
        auto n = std::make_unique<bfgpu_bf>();
        materialize_transpose(h->b, op == BFGPU_OP_C, n->b);
        *out = n.release();
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return BFGPU_ERR_STATE;
    }
    return BFGPU_OK;
}

extern "C" int bfgpu_bf_recompress(bfgpu_bf *h, double tau, bfgpu_bf **out) {
    try {
        BFG_REQUIRE(h && out, "bf_recompress: null argument");
        BFG_REQUIRE(tau >= 0.0, "bf_recompress: tau must be >= 0");
        BFG_HIP(hipSetDevice(h->b.ctx->device));
        auto n = std::make_unique<bfgpu_bf>();
        recompress(h->b, tau, n->b);
        *out = n.release();
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return BFGPU_ERR_STATE;
    }
    return BFGPU_OK;
}
