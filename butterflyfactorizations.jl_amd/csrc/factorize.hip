// factorize.hip -- level-by-level butterfly factorization driven from the
// host, with every kernel-matrix tile, pivoted QR and interpolation matrix
// computed on the GPU in batches (all compressions of one level, over every
// factorization of the batch, per launch).  Host-side planning and
// emission run in parallel over the factorizations of a batch.
//
// Restates assemble_BF (src/ButterflyFactorization/bfassembly.jl:48-200)
// with build_U_skeletons (:206-230), build_nonfrozen_R_blocks (:232-349),
// build_sourcefrozen_R_blocks (:351-412), build_observerfrozen_R_blocks
// (:414-481) and the PartialQR functor (compressors.jl:81-165).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <exception>
#include <map>
#include <mutex>
#include <numeric>
#include <thread>
#include <unordered_map>

#include <sys/mman.h>
#include <sys/resource.h>
#include <type_traits>
#include <malloc.h>
#include <climits>

#include "internal.hpp"

namespace bfg {

int32_t mlp_predict_ranks(const bfgpu_mlp *m, const float *feat, int64_t n, int32_t *out);

int host_threads() {
    static int n = [] {
        int h = int(std::thread::hardware_concurrency());
        if (const char *e = std::getenv("OMP_NUM_THREADS")) {
            int v = std::atoi(e);
            if (v > 0) h = std::min(h > 0 ? h : v, v);
        }
        return std::max(1, std::min(h, 64));
    }();
    return n;
}

// ---- row sample: sparse Fisher-Yates prefix over a SplitMix64 counter
// stream (the documented replacement of compressors.jl:101 randperm) ----
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

namespace {
struct BigPool {
    std::mutex mu;
    std::multimap<size_t, void *> free_blocks;    // mapped size -> block
    std::unordered_map<void *, size_t> mapped;    // block -> mapped size
    size_t cached = 0;
    const size_t cap = [] {
        const char *e = std::getenv("BFGPU_HEAP_KEEP");
        return (e && std::atoi(e) == 0) ? size_t(0) : size_t(32) << 30;
    }();
};
BigPool &big_pool() {
    static BigPool *p = new BigPool;  // outlives static destructors of arrays that use it
    return *p;
}
constexpr size_t HUGE_PAGE = size_t(2) << 20;
}  // namespace

void *host_big_alloc(size_t bytes) {
    const size_t rb = (bytes + HUGE_PAGE - 1) & ~(HUGE_PAGE - 1);
    BigPool &P = big_pool();
    {
        std::lock_guard<std::mutex> g(P.mu);
        // a kept block of at least the size, at most a quarter larger
        auto it = P.free_blocks.lower_bound(rb);
        if (it != P.free_blocks.end() && it->first <= rb + rb / 4) {
            void *p = it->second;
            P.cached -= it->first;
            P.free_blocks.erase(it);
            return p;
        }
    }
    void *raw = mmap(nullptr, rb + HUGE_PAGE, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (raw == MAP_FAILED) throw std::bad_alloc();
    const uintptr_t a = (uintptr_t(raw) + HUGE_PAGE - 1) & ~uintptr_t(HUGE_PAGE - 1);
    if (a > uintptr_t(raw)) munmap(raw, a - uintptr_t(raw));
    const uintptr_t end = uintptr_t(raw) + rb + HUGE_PAGE;
    if (end > a + rb) munmap(reinterpret_cast<void *>(a + rb), end - (a + rb));
    void *p = reinterpret_cast<void *>(a);
    static const bool thp = [] {
        const char *e = std::getenv("BFGPU_HOST_THP");
        return !(e && std::atoi(e) == 0);
    }();
    if (thp) madvise(p, rb, MADV_HUGEPAGE);
    std::lock_guard<std::mutex> g(P.mu);
    P.mapped[p] = rb;
    return p;
}

void host_big_free(void *p, size_t bytes) {
    (void)bytes;
    if (!p) return;
    BigPool &P = big_pool();
    size_t rb = 0;
    {
        std::lock_guard<std::mutex> g(P.mu);
        auto it = P.mapped.find(p);
        if (it == P.mapped.end()) return;  // not ours: cannot happen for arrays of this allocator
        rb = it->second;
        if (P.cached + rb <= P.cap) {
            P.free_blocks.emplace(rb, p);
            P.cached += rb;
            return;
        }
        P.mapped.erase(it);
    }
    munmap(p, rb);
}

uint64_t sample_key(uint64_t seed, int32_t no, int32_t ns, int32_t stage, int32_t s, int32_t o) {
    uint64_t h = mix64(seed + 0x9E3779B97F4A7C15ULL);
    h = mix64(h ^ uint64_t(uint32_t(no)));
    h = mix64(h ^ (uint64_t(uint32_t(ns)) << 1));
    h = mix64(h ^ uint64_t(uint32_t(stage)));
    h = mix64(h ^ uint64_t(uint32_t(s)));
    h = mix64(h ^ uint64_t(uint32_t(o)));
    return h;
}

// Same permutation prefix as a swap map over the identity; the map is an
// open-addressing table reused per thread (generation-stamped, no clears).
void sample_prefix(uint64_t key, int32_t n, int32_t m, int32_t *out) {
    if (m > n) m = n;
    if (m <= 0) return;
    thread_local std::vector<int32_t> tk, tv;
    thread_local std::vector<uint32_t> tg;
    thread_local uint32_t gen = 0;
    size_t cap = 16;
    while (cap < size_t(4) * size_t(m)) cap <<= 1;
    if (tk.size() < cap) {
        tk.assign(cap, 0);
        tv.assign(cap, 0);
        tg.assign(cap, 0);
        gen = 0;
    }
    if (++gen == 0) {
        std::fill(tg.begin(), tg.end(), 0u);
        gen = 1;
    }
    const size_t mask = cap - 1;
    auto slot = [&](int32_t k) {
        size_t h = size_t((uint64_t(uint32_t(k)) * 0x9E3779B97F4A7C15ULL) >> 32) & mask;
        while (tg[h] == gen && tk[h] != k) h = (h + 1) & mask;
        return h;
    };
    auto get = [&](int32_t k) {
        size_t h = slot(k);
        return tg[h] == gen ? tv[h] : k;
    };
    auto set = [&](int32_t k, int32_t v) {
        size_t h = slot(k);
        tg[h] = gen;
        tk[h] = k;
        tv[h] = v;
    };
    for (int32_t i = 0; i < m; ++i) {
        uint64_t u = mix64(key + uint64_t(i + 1) * 0x9E3779B97F4A7C15ULL);
        uint32_t span = uint32_t(n - i);
        int32_t j = i + int32_t(((u >> 32) * uint64_t(span)) >> 32);
        int32_t ai = get(i), aj = get(j);
        set(i, aj);
        set(j, ai);
        out[i] = aj;
    }
}

static inline uint64_t fastkey(int32_t s, int32_t o) {
    return (uint64_t(uint32_t(s)) << 32) | uint64_t(uint32_t(o));
}

namespace {

struct Comp {
    int32_t bf, stage, S, O;
    const int32_t *obs;   // candidate observer point ids
    int32_t n_obs;
    int32_t n_cols;       // source point ids: cols[0, n_cols), in the level's column pool or
    const int32_t *cols;  // the tree's value list
    double eps;
    uint64_t key;
    int32_t guess = 10, rows = 0, rank = 0;
    int32_t judge = 0;      // 0 not judged, 1 accepted by the residual judge, 2 redone by plain RRQR
    int64_t skel_ofs = -1;  // skeleton (rank source ids) in the level's skeleton pool
    int64_t tofs = 0;
    float feat[6];
};

// Skeleton index of one factorization: (source node, observer node) ->
// (offset, length) of the skeleton in the factorization's skeleton log;
// open addressing, doubled at half load.
struct SkelMap {
    std::vector<uint64_t> k;  // key + 1 (0: empty)
    std::vector<uint64_t> v;  // offset << 32 | length
    size_t n = 0;
    static size_t hash(uint64_t key) {
        const uint64_t x = key * 0x9E3779B97F4A7C15ULL;
        return size_t(x ^ (x >> 29));
    }
    size_t find(uint64_t key) const {
        const size_t mask = k.size() - 1;
        size_t h = hash(key) & mask;
        while (k[h] != 0 && k[h] != key + 1) h = (h + 1) & mask;
        return h;
    }
    void put(uint64_t key, int64_t ofs, int32_t len) {
        BFG_REQUIRE(ofs >= 0 && ofs < (int64_t(1) << 31), "factorize: skeleton log exceeds 2^31 entries");
        if (2 * (n + 1) > k.size()) {
            std::vector<uint64_t> ok, ov;
            ok.swap(k);
            ov.swap(v);
            k.assign(std::max<size_t>(16, 2 * ok.size()), 0);
            v.assign(k.size(), 0);
            for (size_t i = 0; i < ok.size(); ++i)
                if (ok[i]) {
                    const size_t h = find(ok[i] - 1);
                    k[h] = ok[i];
                    v[h] = ov[i];
                }
        }
        const size_t h = find(key);
        if (k[h] == 0) {
            k[h] = key + 1;
            ++n;
        }
        v[h] = (uint64_t(uint32_t(ofs)) << 32) | uint32_t(len);
    }
    // offset and length of key's skeleton; the key must be present
    std::pair<int64_t, int32_t> at(uint64_t key) const {
        const size_t h = k.empty() ? 0 : find(key);
        BFG_REQUIRE(!k.empty() && k[h] == key + 1, "factorize: skeleton of a block missing");
        return {int64_t(v[h] >> 32), int32_t(uint32_t(v[h]))};
    }
    void swap(SkelMap &o) {
        k.swap(o.k);
        v.swap(o.v);
        std::swap(n, o.n);
    }
};

// Per-factorization assembly state.
struct Build;

// The planning state of a finished factorization (millions of small
// containers over all its factorizations) is released by a background
// thread, so the call returns without waiting for the frees; the next
// release, and the end of the process, wait for the previous one.
class Reaper {
    std::mutex mu_;
    std::thread th_;

public:
    void post(std::vector<Build> &&b);
    ~Reaper() {
        if (th_.joinable()) th_.join();
    }
};

// Node lists of the levels of a tree under a root, flat: level i is
// v[ofs[i], ofs[i + 1]) (two allocations per tree instead of one per level).
struct Levels {
    std::vector<int32_t> v, ofs;
    struct Span {
        const int32_t *p;
        size_t n;
        const int32_t *begin() const { return p; }
        const int32_t *end() const { return p + n; }
        size_t size() const { return n; }
        int32_t operator[](size_t i) const { return p[i]; }
    };
    Levels() = default;
    explicit Levels(const std::vector<std::vector<int32_t>> &lv) {
        ofs.resize(lv.size() + 1, 0);
        for (size_t i = 0; i < lv.size(); ++i) ofs[i + 1] = ofs[i] + int32_t(lv[i].size());
        v.reserve(size_t(ofs.back()));
        for (const auto &l : lv) v.insert(v.end(), l.begin(), l.end());
    }
    size_t size() const { return ofs.empty() ? 0 : ofs.size() - 1; }
    Span operator[](size_t i) const { return Span{v.data() + ofs[i], size_t(ofs[i + 1] - ofs[i])}; }
    Span back() const { return (*this)[size() - 1]; }
    void swap(Levels &o) {
        v.swap(o.v);
        ofs.swap(o.ofs);
    }
};

struct Build {
    int32_t no, ns, LS, LO, L;
    double eps;
    Levels tS, tO;
    SkelMap K;  // skeletons so far, as spans of the factorization's skeleton log
    // release o's containers here (called from worker threads, so the
    // many small frees of a large batch run in parallel)
    void swap_out(Build &o) {
        tS.swap(o.tS);
        tO.swap(o.tO);
        K.swap(o.K);
    }
};

void Reaper::post(std::vector<Build> &&b) {
    std::lock_guard<std::mutex> g(mu_);
    if (th_.joinable()) th_.join();
    th_ = std::thread([v = std::move(b)]() mutable {
        try {
            parallel_for(int64_t(v.size()), [&](int64_t i) { Build().swap_out(v[i]); }, 4096);
        } catch (...) {  // the remaining frees happen when v goes out of scope
        }
    });
}

Reaper &reaper() {
    static Reaper r;
    return r;
}

template <class T> void upload(DevBuf &d, const std::vector<T> &h, hipStream_t s) {
    d.reserve(sizeof(T) * std::max<size_t>(h.size(), 1));
    if (!h.empty()) BFG_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
}

using clk = std::chrono::steady_clock;
inline double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

struct Engine {
    Ctx &ctx;
    const Tree &test, &trial;
    const bfgpu_opts &o;
    int prec;
    size_t es;
    DevBuf d_test_pts, d_trial_pts;
    DevBuf d_rows, d_cols, d_fill, d_fill_end, d_qr, d_interp, d_piv, d_rank, d_Z;
    DevBuf d_test_perm, d_stask, d_big;  // row samples drawn on the device
    DevBuf d_skofs, d_skel;               // skeletons gathered on the device
    // residual judge: tasks and results
    DevBuf d_jq, d_jres;
    int64_t n_comp = 0, n_retry = 0, n_rank0 = 0, n_judged = 0, n_fallback = 0;
    double t_host = 0, t_qr = 0, t_interp = 0;  // seconds: sample prep, fill+QR, interpolation
    double t_judge = 0;                         // seconds: held-out rows, fill, judge (part of t_qr)
    double t_judge_gpu = 0;                     // of which the judge kernel and its copies
    double t_guess = 0, t_trim = 0;             // seconds: rank guesses, factor-buffer trims
    // host lists of a level's batches, kept across levels: their capacity is
    // reused instead of hundreds of MB being unmapped and faulted in again
    // per level (unmapping stalls every worker thread)
    hvec<int32_t> pend, later, retry, acc;
    hvec<int32_t> hrows;  // host-drawn samples (prefixes too long for the device)
    hvec<SampleTask> stask;
    hvec<int64_t> bigofs;
    hvec<int32_t> hrank, jq, jpos;
    hvec<FillTask> ft;
    hvec<int64_t> fe;
    hvec<int64_t> rofs, cofs, zo, jr, jz, jrf, jzf, jcnt, tofs_run, it_pos, it_big, sk_run;
    hvec<QrTask> qt;
    hvec<InterpTask> it;
    hvec<float> gfeat;
    hvec<int32_t> gpred;
    hvec<int64_t> e_pos;  // per compression of a level: its place among the pending ones
    hvec<double> j_res;  // residual judge
    std::vector<uint8_t> j_fail;
    hvec<int32_t> e_skofs;  // per batch: skeleton offsets
    hvec<uint8_t> e_dbl;  // per batch: doubled, held-out rows
    hvec<int32_t> e_jm;

    Engine(Ctx &c, const Tree &te, const Tree &tr, const bfgpu_opts &op)
        : ctx(c), test(te), trial(tr), o(op), prec(op.precision),
          es(op.precision == BFGPU_C128 ? 16 : 8) {
        upload(d_test_pts, test.points, ctx.stream);
        upload(d_trial_pts, trial.points, ctx.stream);
        upload(d_test_perm, test.perm, ctx.stream);
    }

    // Kernel-matrix entries of the listed blocks: the plug-in's fill when
    // one is given, else the Helmholtz point kernel.
    void fill(const int32_t *rows, const int32_t *cols, const FillTask *tasks, const int64_t *task_end,
              int32_t n_tasks, int64_t total, void *Z) {
        if (const KernelRWG *kr = kernel_of(o.kernel))
            launch_fill_rwg(ctx.stream, prec, *kr, rows, cols, tasks, task_end, n_tasks, total, Z, o.k);
        else
            launch_fill(ctx.stream, prec, d_test_pts.as<double>(), d_trial_pts.as<double>(), rows, cols, tasks,
                        task_end, n_tasks, total, Z, o.k);
    }

    // Shrink a level's factor buffer to what it holds (the growth policy
    // below leaves up to a third of it unused) and zero its tail pad: the
    // matvec kernels clamp their loads into the blocks and may read the pad
    // entry after the last one, which then meets a zero weight and must be
    // finite.
    void trim_pool(DevBuf &pool, int64_t &cap_elems, int64_t used) {
        const auto t0 = clk::now();
        struct Acc {
            double &t;
            clk::time_point a;
            ~Acc() { t += secs(a, clk::now()); }
        } acc{t_trim, t0};
        const int64_t need = used + FACTOR_PAD;
        if (used == 0 && cap_elems == 0) return;  // no blocks on this level
        BFG_REQUIRE(cap_elems >= need, "factorize: factor buffer without room for its tail pad");
        if (cap_elems > need + need / 32) {
            DevBuf nb;
            nb.alloc(size_t(need) * es);
            BFG_HIP(hipMemcpyAsync(nb.p, pool.p, size_t(used) * es, hipMemcpyDeviceToDevice, ctx.stream));
            BFG_HIP(hipStreamSynchronize(ctx.stream));
            pool = std::move(nb);
            cap_elems = need;
        }
        BFG_HIP(hipMemsetAsync(static_cast<char *>(pool.p) + size_t(used) * es, 0, size_t(FACTOR_PAD) * es,
                               ctx.stream));
        BFG_HIP(hipStreamSynchronize(ctx.stream));
    }

    void release() {
        for (DevBuf *d : {&d_rows, &d_cols, &d_fill, &d_fill_end, &d_qr, &d_interp, &d_piv, &d_rank, &d_Z,
                          &d_test_perm, &d_stask, &d_big, &d_skofs, &d_skel,
                          &d_jq, &d_jres})
            *d = DevBuf();
        Engine *e = this;
        auto drop = [](auto &v) { std::remove_reference_t<decltype(v)>().swap(v); };
        for (auto *v : {&e->pend, &e->later, &e->retry, &e->acc, &e->hrank, &e->jq, &e->jpos, &e->e_jm, &e->e_skofs})
            drop(*v);
        drop(gpred);
        drop(e_pos);
        for (auto *v : {&e->rofs, &e->cofs, &e->zo, &e->jr, &e->jz, &e->jrf, &e->jzf, &e->jcnt, &e->tofs_run,
                        &e->it_pos, &e->it_big, &e->sk_run})
            drop(*v);
        drop(hrows);
        drop(stask);
        drop(bigofs);
        drop(ft);
        drop(fe);
        drop(qt);
        drop(it);
        drop(gfeat);
        drop(j_res);
        drop(j_fail);
        drop(e_dbl);
    }

    // out[0] = 0, out[i + 1] = out[i] + v(i): chunk sums in parallel, a
    // serial pass over the chunk totals, then the chunks offset in parallel
    template <class V, class F> static void prefix_sum(int64_t n, V &out, F v) {
        out.resize(size_t(n) + 1);
        out[0] = 0;
        constexpr int64_t CH = int64_t(1) << 15;
        const int64_t nch = (n + CH - 1) / CH;
        std::vector<int64_t> part(size_t(nch) + 1, 0);
        parallel_for(nch, [&](int64_t c) {
            int64_t sum = 0;
            for (int64_t i = c * CH; i < std::min(n, (c + 1) * CH); ++i) {
                sum += v(i);
                out[size_t(i) + 1] = sum;
            }
            part[size_t(c) + 1] = sum;
        }, 1);
        for (int64_t c = 0; c < nch; ++c) part[size_t(c) + 1] += part[size_t(c)];
        parallel_for(nch, [&](int64_t c) {
            const int64_t base = part[size_t(c)];
            if (base)
                for (int64_t i = c * CH; i < std::min(n, (c + 1) * CH); ++i) out[size_t(i) + 1] += base;
        }, 1);
    }

    // Grow a device buffer of the engine; kernels still queued (the
    // interpolation of the previous batch is not waited for) may read the
    // old one, so the stream is drained before it is freed.
    void need(DevBuf &d, size_t n) {
        if (d.p && n <= d.bytes) return;
        BFG_HIP(hipStreamSynchronize(ctx.stream));
        d.reserve(n);
    }
    template <class T, class A> void upload(DevBuf &d, const std::vector<T, A> &h, hipStream_t s) {
        upload_n(d, h.data(), h.size(), s);
    }
    template <class T> void upload_n(DevBuf &d, const T *h, size_t n, hipStream_t s) {
        need(d, sizeof(T) * std::max<size_t>(n, 1));
        if (n) BFG_HIP(hipMemcpyAsync(d.p, h, sizeof(T) * n, hipMemcpyHostToDevice, s));
    }

    void ensure_pool(DevBuf &pool, int64_t &cap_elems, int64_t need) {
        if (need <= cap_elems) return;
        int64_t nc = std::max<int64_t>(need, cap_elems * 3 / 2 + 1024);
        DevBuf nb;
        nb.alloc(size_t(nc) * es);
        if (cap_elems > 0)
            BFG_HIP(hipMemcpyAsync(nb.p, pool.p, size_t(cap_elems) * es, hipMemcpyDeviceToDevice, ctx.stream));
        BFG_HIP(hipStreamSynchronize(ctx.stream));
        pool = std::move(nb);
        cap_elems = nc;
    }

    void guesses(std::vector<Comp> &cs) {
        if (o.rank_mode == BFGPU_RANK_LEARNED) {
            BFG_REQUIRE(o.mlp != nullptr, "factorize: learned ranks need an MLP");
            hvec<float> &f = gfeat;
            f.resize(6 * cs.size());
            parallel_for(int64_t(cs.size()), [&](int64_t i) {
                for (int q = 0; q < 6; ++q) f[6 * size_t(i) + q] = cs[size_t(i)].feat[q];
            }, 1 << 14);
            hvec<int32_t> &pr = gpred;
            pr.resize(cs.size());
            mlp_predict_ranks(o.mlp, f.data(), int64_t(cs.size()), pr.data());
            parallel_for(int64_t(cs.size()), [&](int64_t i) {
                cs[size_t(i)].guess = std::max(pr[size_t(i)], 1) + o.oversample;
            }, 1 << 14);
            return;
        }
        parallel_for(int64_t(cs.size()), [&](int64_t i) {
            Comp &c = cs[size_t(i)];
            if (o.rank_mode == BFGPU_RANK_GEOMETRIC) {
                const double *csrc = &trial.center[3 * c.S], *cobs = &test.center[3 * c.O];
                double d = std::sqrt((csrc[0] - cobs[0]) * (csrc[0] - cobs[0]) +
                                     (csrc[1] - cobs[1]) * (csrc[1] - cobs[1]) +
                                     (csrc[2] - cobs[2]) * (csrc[2] - cobs[2]));
                double ds = std::max(d, 1e-4);
                double a = o.k * (trial.radius(c.S) * test.radius(c.O)) / ds;
                int32_t R = int32_t(std::ceil(o.c_geo * a * a + o.c_eps * std::log(1.0 / c.eps)));
                c.guess = std::max(R, o.rmin);
            } else {
                int32_t R = int32_t(std::ceil(o.c_tau * std::log(1.0 / c.eps)));
                c.guess = std::max(R, o.rmin);
            }
        }, 1 << 14);
    }

    // Held-out rows the residual judge checks compression c on: the next
    // judge_rows rows of its sample prefix (fewer when the observers run
    // out), for learned ranks whose sample left observers unseen and whose
    // columns fit the judge kernel; 0 when c is not judged.
    int32_t judge_m(const Comp &c) const {
        if (o.rank_mode != BFGPU_RANK_LEARNED || o.judge_rows <= 0 || c.judge != 0 || c.rows >= c.n_obs ||
            c.n_cols == 0 || c.n_cols > JUDGE_MAX_COLS)
            return 0;
        return std::min(o.judge_rows, c.n_obs - c.rows);
    }

    // Residual judge of the held-out tiles of the current batch (jq[a]:
    // batch index of tile a, the tile's fill task at np + a) whose
    // compression the pivoted QR accepted (dbl[q] == 0; doubled ones are
    // judged after their retry): a compression passes when the largest
    // residual column norm of its interpolative decomposition on the
    // held-out tile is at most judge_slack * eps times the tile's largest
    // column norm; one that fails is redone by plain adaptive RRQR from
    // max(ceil(c_tau ln(1/eps)), 10, 2 * rows) rows.  The QR tiles, pivots,
    // ranks, held-out tiles and their descriptors are still on the device:
    // only the batch indices go up and two numbers per tile come back.
    void judge(std::vector<Comp> &cs, const hvec<int32_t> &pend, const hvec<uint8_t> &dbl, size_t np,
               hvec<int32_t> &retry) {
        auto t0 = clk::now();
        const int64_t nj = int64_t(jq.size());
        if (nj == 0) return;
        upload(d_jq, jq, ctx.stream);
        need(d_jres, sizeof(double) * 2 * size_t(nj));
        launch_judge(ctx.stream, prec, d_Z.p, d_jq.as<int32_t>(), d_qr.as<QrTask>(), d_fill.as<FillTask>() + np,
                     d_rank.as<int32_t>(), int32_t(nj), d_piv.as<int32_t>(), d_jres.as<double>());
        hvec<double> &res = j_res;
        res.resize(2 * static_cast<size_t>(nj));
        BFG_HIP(hipMemcpyAsync(res.data(), d_jres.p, sizeof(double) * res.size(), hipMemcpyDeviceToHost, ctx.stream));
        BFG_HIP(hipStreamSynchronize(ctx.stream));
        auto t1 = clk::now();
        const double slack = o.judge_slack > 0.0 ? o.judge_slack : 4.0;
        // per tile: 0 accepted, 1 rejected, 2 not judged (QR doubled the rows)
        std::vector<uint8_t> &verdict = j_fail;
        verdict.resize(static_cast<size_t>(nj));
        parallel_for(nj, [&](int64_t a) {
            const int32_t q = jq[size_t(a)];
            if (dbl[size_t(q)]) {
                verdict[size_t(a)] = 2;
                return;
            }
            Comp &c = cs[pend[q]];
            const bool fail = !(res[2 * a] <= slack * c.eps * res[2 * a + 1]);
            verdict[size_t(a)] = fail ? 1 : 0;
            c.judge = fail ? 2 : 1;
        }, 4096);
        // rejected ones (about one in a hundred), in batch order
        int64_t skipped = 0;
        for (int64_t a = 0; a < nj; ++a) {
            if (verdict[size_t(a)] == 0) continue;
            if (verdict[size_t(a)] == 2) {
                ++skipped;
                continue;
            }
            Comp &c = cs[pend[jq[size_t(a)]]];
            ++n_fallback;
            const int32_t plain = std::max(int32_t(std::ceil(o.c_tau * std::log(1.0 / c.eps))), 10);
            c.rows = std::min(std::max(plain, 2 * c.rows), c.n_obs);
            retry.push_back(pend[jq[size_t(a)]]);
        }
        n_judged += nj - skipped;
        t_judge_gpu += secs(t0, t1);
        t_judge += secs(t0, clk::now());
    }

    // Run every compression of one level in batches: fill -> pivoted QR ->
    // (adaptive doubling of the rows that saturated) -> interpolation
    // matrices written into `pool` (the level's factor buffer), skeletons
    // appended to `skp` (the level's skeleton pool, Comp::skel_ofs).
    // Every compression's columns lie in colpool[0, colpool_len) (the
    // trial tree's values at Q, the level's column lists at R): uploaded once
    // per level, the fill tasks index it.
    void run(std::vector<Comp> &cs, DevBuf &pool, int64_t &pool_cap, int64_t &pool_used, std::vector<int32_t> &skp,
             const int32_t *colpool, int64_t colpool_len) {
        const auto tg = clk::now();
        BFG_REQUIRE(colpool_len >= 0 && colpool_len < (int64_t(1) << 31), "factorize: column lists exceed 2^31 entries");
        parallel_for(int64_t(cs.size()), [&](int64_t i) {
            const Comp &c = cs[size_t(i)];
            BFG_REQUIRE(c.n_cols == 0 || (c.cols >= colpool && c.cols + c.n_cols <= colpool + colpool_len),
                        "factorize: compression columns outside the level's column lists");
        }, 1 << 14);
        upload_n(d_cols, colpool, size_t(colpool_len), ctx.stream);
        guesses(cs);
        // the compressions with observers and columns, in level order (the
        // others have rank 0 and no factor entries)
        const int64_t ncs = int64_t(cs.size());
        auto live = [&](int64_t i) { return cs[size_t(i)].n_obs > 0 && cs[size_t(i)].n_cols > 0; };
        prefix_sum(ncs, e_pos, [&](int64_t i) { return int64_t(live(i)); });
        pend.resize(size_t(e_pos[size_t(ncs)]));
        parallel_for(ncs, [&](int64_t i) {
            Comp &c = cs[size_t(i)];
            if (!live(i)) {
                c.rank = 0;
                c.rows = 0;
                c.tofs = pool_used;
                return;
            }
            c.rows = std::min(std::max(c.guess, 10), c.n_obs);
            pend[size_t(e_pos[size_t(i)])] = int32_t(i);
        }, 1 << 14);
        n_comp += ncs;
        t_guess += secs(tg, clk::now());
        // kernel-matrix tiles of one batch: at most 32 GiB or 30 % of the
        // free device memory (besides what the tile buffer already holds)
        int64_t z_budget = o.workspace_bytes;
        if (z_budget <= 0) {
            size_t free_b = 0, total_b = 0;
            BFG_HIP(hipMemGetInfo(&free_b, &total_b));
            z_budget = std::max<int64_t>(int64_t(1) << 30,
                                         std::min<int64_t>(int64_t(32) << 30, int64_t(0.3 * double(free_b + d_Z.bytes))));
        }
        while (!pend.empty()) {
            auto th = clk::now();
            {
                size_t nb = 0;
                int64_t zb = 0;
                for (; nb < pend.size(); ++nb) {
                    const Comp &c = cs[pend[nb]];
                    const int64_t z = int64_t(c.rows + judge_m(c)) * int64_t(c.n_cols) * es;
                    if (nb > 0 && zb + z > z_budget) break;
                    zb += z;
                }
                later.assign(pend.begin() + nb, pend.end());
                pend.resize(nb);
            }
            const size_t np = pend.size();
            // sizes of every tile, then running offsets (parallel prefix
            // sums): main tiles, then the judge's held-out tiles
            hvec<int32_t> &jm = e_jm;
            jm.resize(np);
            parallel_for(int64_t(np), [&](int64_t q) { jm[q] = judge_m(cs[pend[q]]); }, 4096);
            prefix_sum(int64_t(np), rofs, [&](int64_t q) { return int64_t(cs[pend[q]].rows); });
            prefix_sum(int64_t(np), cofs, [&](int64_t q) { return int64_t(cs[pend[q]].n_cols); });
            prefix_sum(int64_t(np), zo, [&](int64_t q) {
                return int64_t(cs[pend[q]].rows) * int64_t(cs[pend[q]].n_cols);
            });
            prefix_sum(int64_t(np), jrf, [&](int64_t q) { return int64_t(jm[q]); });
            prefix_sum(int64_t(np), jzf, [&](int64_t q) { return int64_t(jm[q]) * int64_t(cs[pend[q]].n_cols); });
            prefix_sum(int64_t(np), jcnt, [&](int64_t q) { return int64_t(jm[q] > 0); });
            const size_t nj = size_t(jcnt[np]);
            const int64_t zmain = zo[np];
            int64_t zofs = zmain + jzf[np];
            ft.resize(np + nj);
            fe.resize(np + nj);
            qt.resize(np);
            jq.resize(nj);
            jr.resize(nj + 1);
            jz.resize(nj);
            jpos.resize(np);
            std::atomic<int32_t> amax_n{1}, amax_m{1};
            parallel_for(int64_t(np), [&](int64_t q) {
                const Comp &c = cs[pend[q]];
                const int32_t n = c.n_cols;
                const int32_t cp = c.n_cols ? int32_t(c.cols - colpool) : 0;  // columns in the level's lists
                ft[q] = FillTask{zo[q], c.rows, n, int32_t(rofs[q]), cp};
                fe[q] = zo[q + 1];
                qt[q] = QrTask{zo[q], c.rows, n, int32_t(cofs[q]), 0};
                int32_t v = amax_n.load(std::memory_order_relaxed);
                while (n > v && !amax_n.compare_exchange_weak(v, n)) {
                }
                v = amax_m.load(std::memory_order_relaxed);
                while (c.rows > v && !amax_m.compare_exchange_weak(v, c.rows)) {
                }
                if (jm[q] == 0) {
                    jpos[q] = -1;
                    return;
                }
                const int64_t a = jcnt[q];
                jpos[q] = int32_t(a);
                jq[a] = int32_t(q);
                jr[a] = jrf[q];
                jz[a] = zmain + jzf[q];
                ft[np + a] = FillTask{jz[a], jm[q], n, int32_t(rofs[np] + jrf[q]), cp};
                fe[np + a] = jz[a] + int64_t(jm[q]) * n;
            }, 4096);
            jr[nj] = jrf[np];
            const int32_t max_n = amax_n.load(), max_m = amax_m.load();
            BFG_REQUIRE(rofs[np] + jr[nj] < (int64_t(1) << 31) && cofs[np] < (int64_t(1) << 31),
                        "factorize: level too large for 32-bit index lists");
            const int32_t pofs = int32_t(cofs[np]);
            // row samples: drawn on the device, except prefixes longer than
            // the device's swap map holds (drawn here, copied there);
            // BFGPU_SAMPLE_DEVICE_MAX (read per batch) lowers the limit, so
            // that tests can send every prefix through the host's draw
            int64_t dev_max = SAMPLE_DEVICE_MAX;
            if (const char *e = std::getenv("BFGPU_SAMPLE_DEVICE_MAX"))
                dev_max = std::max<int64_t>(0, std::min<int64_t>(SAMPLE_DEVICE_MAX, std::atoll(e)));
            stask.resize(np);
            hvec<int64_t> &bo = bigofs;
            prefix_sum(int64_t(np), bo, [&](int64_t q) {
                const int64_t t = int64_t(cs[pend[q]].rows) + jm[q];
                return t > dev_max ? t : int64_t(0);
            });
            BFG_REQUIRE(bo[np] < (int64_t(1) << 31), "factorize: host-drawn samples exceed 2^31 entries");
            hrows.resize(size_t(std::max<int64_t>(bo[np], 1)));
            parallel_for(int64_t(np), [&](int64_t q) {
                const Comp &c = cs[pend[q]];
                const bool host = bo[q + 1] > bo[q];
                stask[q] = SampleTask{c.key, test.val_begin[c.O], c.n_obs, jpos[q], host ? int32_t(bo[q]) : -1};
                if (!host) return;
                int32_t *r = hrows.data() + bo[q];
                const int32_t t = int32_t(bo[q + 1] - bo[q]);
                sample_prefix(c.key, c.n_obs, t, r);
                for (int32_t i = 0; i < t; ++i) r[i] = c.obs[r[i]];
            }, 1024);
            auto tq = clk::now();
            t_host += secs(th, tq);
            upload(d_fill, ft, ctx.stream);
            upload(d_fill_end, fe, ctx.stream);
            upload(d_qr, qt, ctx.stream);
            upload(d_stask, stask, ctx.stream);
            if (bo[np] > 0) upload_n(d_big, hrows.data(), size_t(bo[np]), ctx.stream);
            need(d_rows, sizeof(int32_t) * size_t(std::max<int64_t>(rofs[np] + jr[nj], 1)));
            launch_sample_rows(ctx.stream, d_stask.as<SampleTask>(), d_fill.as<FillTask>(), int32_t(np),
                               d_test_perm.as<int32_t>(), d_big.as<int32_t>(), d_rows.as<int32_t>());
            need(d_Z, size_t(zofs) * es);
            need(d_piv, sizeof(int32_t) * size_t(std::max(pofs, 1)));
            need(d_rank, sizeof(int32_t) * np);
            fill(d_rows.as<int32_t>(), d_cols.as<int32_t>(), d_fill.as<FillTask>(), d_fill_end.as<int64_t>(),
                 int32_t(ft.size()), zofs, d_Z.p);
            // QR launches: one per (tolerance, kernel) group -- the batch may
            // hold factorizations of different depth, and the blocks that
            // fit a wavefront's LDS take the wavefront kernel while larger
            // ones take the workgroup kernel, whatever else is in the batch
            // (one group, the common case, is found without building lists)
            const double eps0 = cs[pend[0]].eps;
            const bool ok0 = cpqr_wave_ok(prec, qt[0].m, qt[0].n);
            std::atomic<bool> one_group{true};
            parallel_for(int64_t(np), [&](int64_t q) {
                if (cs[pend[q]].eps != eps0 || cpqr_wave_ok(prec, qt[q].m, qt[q].n) != ok0)
                    one_group.store(false, std::memory_order_relaxed);
            }, 1 << 14);
            std::map<std::pair<double, int>, std::vector<int32_t>> groups;
            if (!one_group.load())
                for (size_t q = 0; q < np; ++q)
                    groups[{cs[pend[q]].eps, cpqr_wave_ok(prec, qt[q].m, qt[q].n) ? 1 : 0}].push_back(int32_t(q));
            hrank.assign(np, 0);
            if (one_group.load()) {
                launch_cpqr(ctx.stream, prec, d_Z.p, d_qr.as<QrTask>(), int32_t(np),
                            d_piv.as<int32_t>(), d_rank.as<int32_t>(), eps0, max_n, max_m);
                BFG_HIP(hipMemcpyAsync(hrank.data(), d_rank.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost, ctx.stream));
            } else {
                std::vector<QrTask> sub;
                DevBuf d_sub, d_subrank;
                for (const auto &gr : groups) {
                    const std::vector<int32_t> &where = gr.second;
                    sub.resize(where.size());
                    int32_t gn = 1, gm = 1;
                    for (size_t i = 0; i < where.size(); ++i) {
                        sub[i] = qt[where[i]];
                        gn = std::max(gn, sub[i].n);
                        gm = std::max(gm, sub[i].m);
                    }
                    upload(d_sub, sub, ctx.stream);
                    need(d_subrank, sizeof(int32_t) * sub.size());
                    launch_cpqr(ctx.stream, prec, d_Z.p, d_sub.as<QrTask>(), int32_t(sub.size()),
                                d_piv.as<int32_t>(), d_subrank.as<int32_t>(), gr.first.first, gn, gm);
                    std::vector<int32_t> r(sub.size());
                    BFG_HIP(hipMemcpyAsync(r.data(), d_subrank.p, sizeof(int32_t) * r.size(), hipMemcpyDeviceToHost, ctx.stream));
                    BFG_HIP(hipStreamSynchronize(ctx.stream));
                    for (size_t q = 0; q < where.size(); ++q) hrank[where[q]] = r[q];
                }
                // the judge reads the ranks in batch order
                BFG_HIP(hipMemcpyAsync(d_rank.p, hrank.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice, ctx.stream));
            }
            BFG_HIP(hipStreamSynchronize(ctx.stream));  // ranks on the host
            auto ti = clk::now();
            t_qr += secs(tq, ti);
            // adaptive doubling (compressors.jl:134) decided per compression
            // in parallel; the retried ones in batch order
            hvec<uint8_t> &dbl = e_dbl;
            dbl.assign(np, 0);
            parallel_for(int64_t(np), [&](int64_t q) {
                Comp &c = cs[pend[q]];
                const int32_t r = hrank[q];
                dbl[q] = o.adaptive && r > int32_t(std::floor(0.8 * c.rows)) && c.rows < c.n_obs;
                if (!dbl[q]) c.rank = r;
            }, 4096);
            retry.clear();
            acc.clear();
            acc.reserve(np);
            for (size_t q = 0; q < np; ++q) {
                Comp &c = cs[pend[q]];
                if (dbl[q]) {
                    c.rows = std::min(2 * c.rows, c.n_obs);
                    retry.push_back(pend[q]);
                    ++n_retry;
                    continue;
                }
                if (c.rank == 0 && n_rank0++ < 4 && std::getenv("BFGPU_TIMING"))
                    std::fprintf(stderr, "[bfgpu] rank-0 compression: %d sampled rows of %d observers x %d columns\n",
                                 c.rows, c.n_obs, c.n_cols);
                acc.push_back(int32_t(q));
            }
            // residual judge of the compressions whose rows a prediction
            // sized: rejected ones are redone by plain adaptive RRQR in a
            // later batch
            const double tj0 = t_judge;
            if (!jq.empty()) judge(cs, pend, dbl, np, retry);
            t_qr += t_judge - tj0;  // reported with fill + QR
            // interpolation matrices of the accepted compressions, in batch
            // order, at running offsets of the level's factor buffer
            const int64_t na = int64_t(acc.size());
            auto kept = [&](int64_t a) {
                const Comp &c = cs[pend[acc[size_t(a)]]];
                return !(c.judge == 2 && c.rows != ft[acc[size_t(a)]].m);  // rejected: redone with more rows
            };
            prefix_sum(na, tofs_run, [&](int64_t a) {
                const Comp &c = cs[pend[acc[size_t(a)]]];
                return kept(a) ? int64_t(c.rank) * int64_t(c.n_cols) : int64_t(0);
            });
            // low ranks (almost all: the lane-per-column kernel) first, the
            // rest after them
            auto small = [&](int64_t a) { return cs[pend[acc[size_t(a)]]].rank <= INTERP_SMALL_RANK; };
            prefix_sum(na, it_pos, [&](int64_t a) { return int64_t(kept(a) && small(a)); });
            prefix_sum(na, it_big, [&](int64_t a) { return int64_t(kept(a) && !small(a)); });
            const int64_t n_small = it_pos[size_t(na)];
            it.resize(size_t(n_small + it_big[size_t(na)]));
            std::atomic<int32_t> amax_rank{0};
            parallel_for(na, [&](int64_t a) {
                if (!kept(a)) return;
                const int32_t q = acc[size_t(a)];
                Comp &c = cs[pend[q]];
                c.tofs = pool_used + tofs_run[size_t(a)];
                const int64_t at = small(a) ? it_pos[size_t(a)] : n_small + it_big[size_t(a)];
                it[size_t(at)] = InterpTask{qt[q].zofs, c.tofs, qt[q].m, qt[q].n, c.rank, qt[q].pofs};
                int32_t v = amax_rank.load(std::memory_order_relaxed);
                while (c.rank > v && !amax_rank.compare_exchange_weak(v, c.rank)) {
                }
            }, 4096);
            pool_used += tofs_run[size_t(na)];
            const int32_t max_rank = amax_rank.load();
            // skeletons (pivot prefixes) at running offsets of the level's
            // skeleton pool, in parallel; compressions being retried with more
            // rows take theirs from the later batch
            auto done = [&](int64_t q) { return cs[pend[q]].rows == ft[q].m; };
            prefix_sum(int64_t(np), sk_run, [&](int64_t q) { return done(q) ? int64_t(cs[pend[q]].rank) : int64_t(0); });
            // gathered on the device (column ids of the pivot prefixes), only
            // they come back
            const int64_t sk0 = int64_t(skp.size());
            BFG_REQUIRE(sk_run[np] < (int64_t(1) << 31), "factorize: skeletons of one batch exceed 2^31 entries");
            skp.resize(size_t(sk0 + sk_run[np]));
            hvec<int32_t> &so = e_skofs;
            so.resize(np + 1);
            parallel_for(int64_t(np) + 1, [&](int64_t q) {
                so[size_t(q)] = int32_t(sk_run[size_t(q)]);
                if (q < int64_t(np) && done(q)) cs[pend[q]].skel_ofs = sk0 + sk_run[size_t(q)];
            }, 1 << 14);
            if (sk_run[np] > 0) {
                upload(d_skofs, so, ctx.stream);
                need(d_skel, sizeof(int32_t) * size_t(sk_run[np]));
                launch_skeleton_gather(ctx.stream, d_qr.as<QrTask>(), d_fill.as<FillTask>(), d_cols.as<int32_t>(),
                                       d_piv.as<int32_t>(), d_skofs.as<int32_t>(), int32_t(np), d_skel.as<int32_t>());
                BFG_HIP(hipMemcpyAsync(skp.data() + sk0, d_skel.p, sizeof(int32_t) * size_t(sk_run[np]),
                                       hipMemcpyDeviceToHost, ctx.stream));
                BFG_HIP(hipStreamSynchronize(ctx.stream));
            }
            if (!it.empty()) {
                ensure_pool(pool, pool_cap, pool_used + FACTOR_PAD);
                upload(d_interp, it, ctx.stream);
                launch_interp_small(ctx.stream, prec, d_Z.p, d_interp.as<InterpTask>(), int32_t(n_small),
                                    d_piv.as<int32_t>(), pool.p);
                launch_interp(ctx.stream, prec, d_Z.p, d_interp.as<InterpTask>() + n_small,
                              int32_t(int64_t(it.size()) - n_small), d_piv.as<int32_t>(), pool.p, max_rank);
                // no wait: the host goes on with the next batch or level
                // (sampling, planning) while the interpolation matrices are
                // formed; everything that next touches the tiles, pivots or
                // the pool is ordered after them on the same stream
            }
            t_interp += secs(ti, clk::now()) - (t_judge - tj0);
            retry.insert(retry.end(), later.begin(), later.end());
            pend.swap(retry);
        }
    }
};

// Geometric features of a compression (source node S, observer node O) for
// the learned rank predictor: box radii in wavelengths, centre distance in
// wavelengths, |cos| of the angle between the separation and the z axis,
// level, and log10(1/eps).
void features(const Tree &test, const Tree &trial, double k, int32_t S, int32_t O, int32_t stage,
              double eps, float *f) {
    double lam = 2.0 * M_PI / k;
    const double *cs = &trial.center[3 * S], *co = &test.center[3 * O];
    double d[3] = {co[0] - cs[0], co[1] - cs[1], co[2] - cs[2]};
    double dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    f[0] = float(trial.radius(S) / lam);
    f[1] = float(test.radius(O) / lam);
    f[2] = float(dist / lam);
    f[3] = float(dist > 0 ? std::fabs(d[2]) / dist : 0.0);
    f[4] = float(stage);
    f[5] = float(std::log10(1.0 / eps));
}

// An action of one level: a compression (rule `kind` 0 = column slices per
// source child or the whole T for a leaf source, 2 = source frozen) or an
// identity block (kind 1).
struct Act {
    int32_t kind, comp;      // comp: index into the factorization's own list
    int32_t Oout, Sv, Ov;    // emission keys
    int32_t len;             // identity length
};

}  // namespace

void compression_features(const Tree &test, const Tree &trial, double k, int32_t S, int32_t O,
                          int32_t stage, double eps, float *f) {
    features(test, trial, k, S, O, stage, eps, f);
}

// Factorize n_pairs blocks; fills bf.bfs (reference layout) and bf.factors.
void factorize_all(BF &bf, const bfgpu_opts &o, const std::vector<std::pair<int32_t, int32_t>> &pairs) {
    // Host heap: a factorization writes several GB of small per-block
    // records from 16 threads; with glibc's defaults the per-thread heaps
    // are trimmed and grown again around that, and the page-table updates
    // stall every worker (block emission took 0.4 s or 3 s from one run to
    // the next).  On the first factorization the process keeps freed memory
    // for reuse instead (trim threshold INT_MAX, 64 MiB top pad, blocks up
    // to 32 MiB from the heaps); BFGPU_HEAP_KEEP=0 leaves glibc's defaults.
    static const bool heap_keep = [] {
        const char *e = std::getenv("BFGPU_HEAP_KEEP");
        if (e && std::atoi(e) == 0) return false;
        mallopt(M_TRIM_THRESHOLD, INT_MAX);
        mallopt(M_TOP_PAD, 64 << 20);
        mallopt(M_MMAP_THRESHOLD, 32 << 20);
        return true;
    }();
    (void)heap_keep;
    auto t0 = clk::now();
    auto minflt = [] {
        rusage u{};
        getrusage(RUSAGE_SELF, &u);
        return int64_t(u.ru_minflt);
    };
    const int64_t flt0 = minflt();
    bf.pf_k = o.kernel ? 0.0 : o.k;  // P blocks of the point kernel can be recomputed
    Ctx &ctx = *bf.ctx;
    const Tree &test = *bf.test, &trial = *bf.trial;
    Engine E(ctx, test, trial, o);
    const int64_t nB = int64_t(pairs.size());
    std::vector<Build> B(nB);
    bf.bfs.assign(nB, OneBF{});
    for (int64_t b = 0; b < nB; ++b)
        BFG_REQUIRE(pairs[b].first >= 0 && pairs[b].first < test.n_nodes && pairs[b].second >= 0 &&
                        pairs[b].second < trial.n_nodes,
                    "factorize: root node out of range");
    parallel_for(nB, [&](int64_t b) {
        Build &x = B[b];
        x.no = pairs[b].first;
        x.ns = pairs[b].second;
        x.tS = Levels(traverse_and_pad(trial, x.ns));
        x.tO = Levels(traverse_and_pad(test, x.no));
        x.LS = int32_t(x.tS.size());
        x.LO = int32_t(x.tO.size());
        x.L = std::max(x.LS, x.LO);
        x.eps = o.tol / x.L;
        bf.bfs[b].no = x.no;
        bf.bfs[b].ns = x.ns;
        bf.bfs[b].st.resize(x.L + 1);
    });
    const double t_setup = secs(t0, clk::now());
    int32_t Lmax = 1;
    for (const Build &x : B) Lmax = std::max(Lmax, x.L);
    bf.n_stage_launch = Lmax + 1;
    bf.factors.clear();
    bf.factors.resize(bf.n_stage_launch);
    bf.factor_elems.assign(bf.n_stage_launch, 0);

    auto mkcomp = [&](int32_t b, int32_t stage, int32_t S, int32_t O, const int32_t *cols, int32_t n_cols) {
        Comp c;
        c.bf = b;
        c.stage = stage;
        c.S = S;
        c.O = O;
        c.obs = test.vals(O);
        c.n_obs = test.nvals(O);
        c.cols = cols;
        c.n_cols = n_cols;
        c.eps = B[b].eps;
        c.key = sample_key(o.seed, B[b].no, B[b].ns, stage, S, O);
        features(test, trial, o.k, S, O, stage, c.eps, c.feat);
        return c;
    };
    auto log_comp = [&](OneBF &ob, const Comp &c) {
        CompRec r{c.stage, c.S, c.O, c.guess, c.rows, c.rank, {}, c.judge};
        for (int q = 0; q < 6; ++q) r.feat[q] = c.feat[q];
        ob.comp.push_back(r);
    };
    // A skeleton update of one level, applied after the level's blocks (the
    // reference's sequential update): the skeleton at `ofs` of the
    // factorization's own log (an identity block) or of the level's pool.
    struct Upd {
        uint64_t key;
        int64_t ofs;
        int32_t len;
        bool from_log;
    };
    // Append u's skeleton to the factorization's skeleton log (a record of
    // stage `stage`) and point K at it.
    auto apply_upd = [&](OneBF &ob, Build &x, int32_t stage, const Upd &u, const std::vector<int32_t> &skp) {
        const int64_t so = int64_t(ob.skel_idx.size());
        BFG_REQUIRE(so + u.len < (int64_t(1) << 31), "factorize: skeleton log exceeds 2^31 entries");
        ob.skel.push_back(SkelRec{stage, int32_t(u.key >> 32), int32_t(u.key & 0xffffffffu), int32_t(so), u.len});
        ob.skel_idx.resize(size_t(so + u.len));
        if (u.len > 0) {
            if (u.from_log)  // an earlier entry of the same log: [ofs, ofs + len) lies before so
                std::copy_n(ob.skel_idx.begin() + u.ofs, u.len, ob.skel_idx.begin() + so);
            else
                std::copy_n(skp.begin() + u.ofs, u.len, ob.skel_idx.begin() + so);
        }
        x.K.put(u.key, so, u.len);
    };
    double t_plan = 0, t_emit = 0, t_drop = 0;
    // per-level arrays, reused level to level (capacity kept): the level's
    // skeletons (Comp::skel_ofs), compressions, actions and column pool
    std::vector<int32_t> skp;
    std::vector<Comp> cs;
    std::vector<Act> acts;
    std::vector<int32_t> up;

    // ---- stage 0: Q (bfassembly.jl:82-102) ----
    {
        auto tp = clk::now();
        std::vector<int64_t> off(nB + 1, 0);
        for (int64_t b = 0; b < nB; ++b) off[b + 1] = off[b] + int64_t(B[b].tS.back().size());
        cs.assign(static_cast<size_t>(off[nB]), Comp{});
        parallel_for(nB, [&](int64_t b) {
            const auto &lv = B[b].tS.back();
            for (size_t i = 0; i < lv.size(); ++i) {
                const int32_t Sl = lv[i];
                cs[off[b] + int64_t(i)] = mkcomp(int32_t(b), 0, Sl, B[b].no, trial.vals(Sl), trial.nvals(Sl));
            }
        }, 1024);
        t_plan += secs(tp, clk::now());
        int64_t cap = 0, used = 0;
        skp.clear();
        E.run(cs, bf.factors[0], cap, used, skp, trial.perm.data(), int64_t(trial.perm.size()));
        E.trim_pool(bf.factors[0], cap, used);
        bf.factor_elems[0] = used;
        auto te = clk::now();
        parallel_for(nB, [&](int64_t b) {
            OneBF &ob = bf.bfs[b];
            Build &x = B[b];
            ob.st[0].b.reserve(size_t(off[b + 1] - off[b]));
            for (int64_t i = off[b]; i < off[b + 1]; ++i) {
                const Comp &c = cs[i];
                log_comp(ob, c);
                ob.st[0].b.push_back(Block{x.no, c.S, x.no, c.S, c.rank, c.n_cols, 0, c.tofs, -1, -1});
                apply_upd(ob, x, 0, Upd{fastkey(c.S, x.no), c.skel_ofs, c.skel_ofs >= 0 ? c.rank : 0, false}, skp);
            }
        }, 1024);
        t_emit += secs(te, clk::now());
        const auto td = clk::now();
        cs.clear();
        t_drop += secs(td, clk::now());
    }

    // ---- R levels (bfassembly.jl:107-181) ----
    for (int32_t l = 1; l <= Lmax - 1; ++l) {
        auto tp = clk::now();
        // Two walks over the actions of every factorization: the first counts
        // compressions, actions and column entries, the second writes them at
        // their offsets in the level's arrays.  The columns U of a
        // compression (the skeletons of the source children under the
        // observer parent, or the frozen source's) are written once per
        // (Svert, Overt) and shared by the observer children's compressions.
        std::vector<int64_t> nc(nB + 1, 0), na(nB + 1, 0), nu(nB + 1, 0);
        auto walk = [&](int64_t b, bool fill) {
            Build &x = B[b];
            int64_t kc = 0, ka = 0, ku = 0;
            if (l <= x.L - 1) {
                const int32_t *sk = bf.bfs[b].skel_idx.data();
                Comp *C = fill ? cs.data() + nc[b] : nullptr;
                Act *A = fill ? acts.data() + na[b] : nullptr;
                int32_t *U = fill ? up.data() + nu[b] : nullptr;
                auto put = [&](uint64_t key) {
                    const auto sp = x.K.at(key);
                    if (fill) std::copy_n(sk + sp.first, sp.second, U + ku);
                    ku += sp.second;
                };
                // the columns of (Sv, Ov): offset within this factorization's U, length
                auto U_of = [&](int32_t Sv, int32_t Ov) {
                    const int64_t o = ku;
                    if (trial.isleaf(Sv)) put(fastkey(Sv, Ov));
                    else
                        for (int32_t c = 0; c < trial.n_children[Sv]; ++c)
                            put(fastkey(trial.first_child[Sv] + c, Ov));
                    return std::pair<int64_t, int32_t>(o, int32_t(ku - o));
                };
                auto frozen = [&](int32_t Sv, int32_t Ov) {
                    const int64_t o = ku;
                    put(fastkey(Sv, Ov));
                    return std::pair<int64_t, int32_t>(o, int32_t(ku - o));
                };
                auto add_comp = [&](int32_t S, int32_t O, std::pair<int64_t, int32_t> u, int32_t kind, int32_t Oout,
                                    int32_t Sv, int32_t Ov) {
                    if (fill) {
                        C[kc] = mkcomp(int32_t(b), l, S, O, U + u.first, u.second);
                        A[ka] = Act{kind, int32_t(kc), Oout, Sv, Ov, 0};
                    }
                    ++kc;
                    ++ka;
                };
                auto add_ident = [&](int32_t Ov, int32_t Sv) {
                    if (fill) A[ka] = Act{1, -1, Ov, Sv, Ov, x.K.at(fastkey(Sv, Ov)).second};
                    ++ka;
                };
                const bool sf = l >= x.LS, of = l >= x.LO;
                if (!sf && !of) {
                    const auto &Os = x.tO[l - 1];
                    const auto &Ss = x.tS[x.LS - l - 1];
                    for (int32_t Sv : Ss)
                        for (int32_t Ov : Os) {
                            if (!test.isleaf(Ov)) {
                                const auto u = U_of(Sv, Ov);
                                for (int32_t c = 0; c < test.n_children[Ov]; ++c) {
                                    const int32_t Oc = test.first_child[Ov] + c;
                                    add_comp(Sv, Oc, u, 0, Oc, Sv, Ov);
                                }
                            } else if (!trial.isleaf(Sv)) {
                                add_comp(Sv, Ov, U_of(Sv, Ov), 0, Ov, Sv, Ov);
                            } else {
                                add_ident(Ov, Sv);
                            }
                        }
                } else if (sf && !of) {
                    const auto &Os = x.tO[l - 1];
                    const int32_t Sv = x.ns;
                    for (int32_t Ov : Os) {
                        if (!test.isleaf(Ov)) {
                            // source frozen: the input is the single slot (Ov, ns)
                            const auto u = frozen(Sv, Ov);
                            for (int32_t c = 0; c < test.n_children[Ov]; ++c) {
                                const int32_t Oc = test.first_child[Ov] + c;
                                add_comp(Sv, Oc, u, 2, Oc, Sv, Ov);
                            }
                        } else {
                            add_ident(Ov, Sv);
                        }
                    }
                } else if (!sf && of) {
                    const auto &Os = x.tO[x.LO - 1];
                    const auto &Ss = x.tS[x.LS - l - 1];
                    for (int32_t Sv : Ss)
                        for (int32_t Ov : Os) {
                            if (!trial.isleaf(Sv)) add_comp(Sv, Ov, U_of(Sv, Ov), 0, Ov, Sv, Ov);
                            else add_ident(Ov, Sv);
                        }
                }
            }
            if (!fill) {
                nc[b + 1] = kc;
                na[b + 1] = ka;
                nu[b + 1] = ku;
            }
        };
        parallel_for(nB, [&](int64_t b) { walk(b, false); }, 1024);
        for (int64_t b = 0; b < nB; ++b) {
            nc[b + 1] += nc[b];
            na[b + 1] += na[b];
            nu[b + 1] += nu[b];
        }
        cs.resize(size_t(nc[nB]));
        acts.resize(size_t(na[nB]));
        up.resize(size_t(nu[nB]));
        parallel_for(nB, [&](int64_t b) { walk(b, true); }, 1024);
        t_plan += secs(tp, clk::now());
        int64_t cap = 0, used = 0;
        skp.clear();
        E.run(cs, bf.factors[l], cap, used, skp, up.data(), int64_t(up.size()));
        E.trim_pool(bf.factors[l], cap, used);
        bf.factor_elems[l] = used;
        // emit blocks in action order; K updates in the same order, applied
        // after the level (the reference's sequential update, :177-180)
        auto te = clk::now();
        parallel_for(nB, [&](int64_t b) {
            Build &x = B[b];
            if (l > x.L - 1) return;
            OneBF &ob = bf.bfs[b];
            auto &st = ob.st[l].b;
            {  // the level's blocks: one per identity or whole-T compression, else one per source child
                size_t nblk = 0;
                for (int64_t ai = na[b]; ai < na[b + 1]; ++ai) {
                    const Act &a = acts[size_t(ai)];
                    nblk += (a.kind != 0 || trial.isleaf(a.Sv)) ? 1 : size_t(trial.n_children[a.Sv]);
                }
                st.reserve(nblk);
            }
            thread_local std::vector<Upd> upd;
            upd.clear();
            for (int64_t ai = na[b]; ai < na[b + 1]; ++ai) {
                const Act &a = acts[size_t(ai)];
                if (a.kind == 1) {
                    st.push_back(Block{a.Oout, a.Sv, a.Ov, a.Sv, a.len, a.len, 1, -1, -1, -1});
                    const auto sp = x.K.at(fastkey(a.Sv, a.Ov));
                    upd.push_back(Upd{fastkey(a.Sv, a.Ov), sp.first, sp.second, true});
                    continue;
                }
                const Comp &c = cs[size_t(nc[b] + a.comp)];
                log_comp(ob, c);
                const int32_t nU = c.n_cols;
                if (a.kind == 2 || trial.isleaf(a.Sv)) {
                    // whole T: source frozen (Oc, ns) <- (Ov, ns), or leaf Svert
                    st.push_back(Block{a.Oout, a.Sv, a.Ov, a.Sv, c.rank, nU, 0, c.tofs, -1, -1});
                } else {
                    int32_t last = 0;
                    for (int32_t q = 0; q < trial.n_children[a.Sv]; ++q) {
                        const int32_t Sc = trial.first_child[a.Sv] + q;
                        const int32_t ks = x.K.at(fastkey(Sc, a.Ov)).second;
                        st.push_back(Block{a.Oout, a.Sv, a.Ov, Sc, c.rank, ks, 0,
                                           c.tofs + int64_t(c.rank) * last, -1, -1});
                        last += ks;
                    }
                }
                upd.push_back(Upd{fastkey(c.S, c.O), c.skel_ofs, c.skel_ofs >= 0 ? c.rank : 0, false});
            }
            for (const Upd &u : upd) apply_upd(ob, x, l, u, skp);
        }, 1024);
        t_emit += secs(te, clk::now());
        const auto td = clk::now();
        cs.clear();
        acts.clear();
        up.clear();
        t_drop += secs(td, clk::now());
    }

    // ---- P (bfassembly.jl:186-197): dense kernel blocks ----
    {
        auto tp = clk::now();
        const int32_t sP = Lmax;
        std::vector<int64_t> noff(nB + 1, 0);  // P blocks per factorization
        for (int64_t b = 0; b < nB; ++b) noff[b + 1] = noff[b] + int64_t(B[b].tO.back().size());
        const int64_t nP = noff[nB];
        std::vector<int64_t> rofs(nP + 1, 0), cofs(nP + 1, 0), zo(nP + 1, 0);
        parallel_for(nB, [&](int64_t b) {
            const Build &x = B[b];
            const auto &lv = x.tO.back();
            for (size_t i = 0; i < lv.size(); ++i) {
                int64_t q = noff[b] + int64_t(i);
                rofs[q + 1] = test.nvals(lv[i]);
                cofs[q + 1] = int64_t(x.K.at(fastkey(x.ns, lv[i])).second);
            }
        });
        for (int64_t q = 0; q < nP; ++q) {
            zo[q + 1] = zo[q] + rofs[q + 1] * cofs[q + 1];
            rofs[q + 1] += rofs[q];
            cofs[q + 1] += cofs[q];
        }
        BFG_REQUIRE(rofs[nP] < (int64_t(1) << 31) && cofs[nP] < (int64_t(1) << 31),
                    "factorize: P stage too large for 32-bit index lists");
        std::vector<int32_t> hrows(static_cast<size_t>(rofs[nP])), hcols(static_cast<size_t>(cofs[nP]));
        std::vector<FillTask> ftall(static_cast<size_t>(nP));
        const bool kernel_entries = o.kernel == nullptr;  // point kernel: P is recomputable
        parallel_for(nB, [&](int64_t b) {
            Build &x = B[b];
            auto &st = bf.bfs[b].st[x.L].b;
            const auto &lv = x.tO.back();
            if (kernel_entries)
                bf.bfs[b].pcols.reserve(size_t(cofs[noff[b + 1]] - cofs[noff[b]]));
            for (size_t i = 0; i < lv.size(); ++i) {
                int64_t q = noff[b] + int64_t(i);
                int32_t Ol = lv[i];
                const auto sp = x.K.at(fastkey(x.ns, Ol));
                const int32_t *col = bf.bfs[b].skel_idx.data() + sp.first;
                FillTask f{zo[q], test.nvals(Ol), sp.second, int32_t(rofs[q]), int32_t(cofs[q])};
                std::copy(test.vals(Ol), test.vals(Ol) + f.m, hrows.begin() + rofs[q]);
                std::copy_n(col, sp.second, hcols.begin() + cofs[q]);
                ftall[q] = f;
                st.push_back(Block{Ol, x.ns, Ol, x.ns, f.m, f.n, 0, zo[q], -1, -1});
                if (kernel_entries) bf.bfs[b].pcols.insert(bf.bfs[b].pcols.end(), col, col + sp.second);
            }
        });
        std::vector<FillTask> ft;
        std::vector<int64_t> fe;
        ft.reserve(size_t(nP));
        fe.reserve(size_t(nP));
        for (const FillTask &f : ftall)
            if (int64_t(f.m) * f.n > 0) {
                ft.push_back(f);
                fe.push_back(f.zofs + int64_t(f.m) * f.n);
            }
        const int64_t zofs = zo[nP];
        t_plan += secs(tp, clk::now());
        bf.factors[sP].alloc(size_t(zofs + FACTOR_PAD) * E.es);
        BFG_HIP(hipMemsetAsync(static_cast<char *>(bf.factors[sP].p) + size_t(zofs) * E.es, 0,
                               size_t(FACTOR_PAD) * E.es, ctx.stream));
        bf.factor_elems[sP] = zofs;
        upload(E.d_rows, hrows, ctx.stream);
        upload(E.d_cols, hcols, ctx.stream);
        upload(E.d_fill, ft, ctx.stream);
        upload(E.d_fill_end, fe, ctx.stream);
        E.fill(E.d_rows.as<int32_t>(), E.d_cols.as<int32_t>(), E.d_fill.as<FillTask>(), E.d_fill_end.as<int64_t>(),
               int32_t(ft.size()), zofs, bf.factors[sP].p);
        BFG_HIP(hipStreamSynchronize(ctx.stream));
    }
    bf.n_comp = E.n_comp;
    bf.n_retry = E.n_retry;
    bf.n_judged = E.n_judged;
    bf.n_fallback = E.n_fallback;
    bf.precision = o.precision;
    for (int64_t b = 0; b < nB; ++b) bf.bfs[b].stage_data_base.assign(bf.bfs[b].st.size(), 0);
    auto tfree = clk::now();
    E.release();  // tile workspace and index lists: not needed for packing
    for (auto *v : {&skp, &up}) std::vector<int32_t>().swap(*v);
    std::vector<Comp>().swap(cs);
    std::vector<Act>().swap(acts);
    reaper().post(std::move(B));
    const double t_free = secs(tfree, clk::now());
    auto tpk = clk::now();
    pack_tables(bf);
    auto t1 = clk::now();
    bf.t_factor_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    bf.t_plan_ms = 1e3 * (t_plan + t_emit);
    bf.t_sample_ms = 1e3 * E.t_host;
    bf.t_qr_ms = 1e3 * E.t_qr;
    bf.t_interp_ms = 1e3 * E.t_interp;
    bf.t_pack_ms = 1e3 * secs(tpk, t1);
    if (std::getenv("BFGPU_TIMING"))
        std::fprintf(stderr,
                     "[bfgpu] factorize %.3fs: setup %.3fs, plan %.3fs, guess %.3fs, sample-prep %.3fs, "
                     "fill+qr %.3fs, interp %.3fs, trim %.3fs, emit %.3fs, drop %.3fs, free %.3fs, pack %.3fs "
                     "(%lld compressions, %lld retries, "
                     "%lld of rank 0, judge %.3fs (kernel and copies %.3fs): %lld judged, %lld fallbacks, "
                     "%d threads; %lld page faults)\n",
                     bf.t_factor_ms * 1e-3, t_setup, t_plan, E.t_guess, E.t_host, E.t_qr, E.t_interp, E.t_trim,
                     t_emit, t_drop, t_free, secs(tpk, t1),
                     (long long)E.n_comp, (long long)E.n_retry, (long long)E.n_rank0, E.t_judge, E.t_judge_gpu,
                     (long long)E.n_judged, (long long)E.n_fallback, host_threads(), (long long)(minflt() - flt0));
}

}  // namespace bfg
