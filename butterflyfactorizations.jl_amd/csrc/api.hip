// api.hip -- C ABI entry points (include/bfgpu.h): error strings, contexts,
// factorization handles, export, conversion and matvec.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>

#include <cmath>

#include "internal.hpp"

namespace bfg {
static thread_local std::string g_last_error;
void set_last_error(const std::string &m) { g_last_error = m; }

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}
void DevBuf::alloc(size_t n) {
    release();
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
        p = nullptr;
        throw Error(BFGPU_ERR_NOMEM, "hipMalloc(" + std::to_string(n) + "): " + hipGetErrorString(e));
    }
    bytes = n;
}
void DevBuf::reserve(size_t n) {
    if (n <= bytes && p) return;
    alloc(std::max(n, bytes + bytes / 2));
}

void factorize_all(BF &bf, const bfgpu_opts &o, const std::vector<std::pair<int32_t, int32_t>> &pairs);
void matvec(BF &bf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y, int64_t ldy,
            const double *alpha, const double *beta, int flags);
void matvec_timed(BF &bf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y,
                  int64_t ldy, int32_t iters, double *stage_ms, int64_t *stage_bytes,
                  int32_t *n_stages, double *total_ms);
void matvec_launch_times(BF &bf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y, int64_t ldy,
                         int32_t iters, int32_t cap, int32_t *n_launch, double *launch_ms, int32_t *desc);
void check_sweep_error(BF &bf);
void release_graph(BF &bf);
void autotune(BF &bf, int64_t nrhs, int32_t iters, double *stage_ms, int32_t *variants);
void noalloc_init(int device);
void matvec_noalloc(BF &bf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y, int64_t ldy,
                    const double *alpha, const double *beta);
}  // namespace bfg

using namespace bfg;


#define BFG_API_BEGIN try {
#define BFG_API_END                                  \
    }                                                \
    catch (const Error &e) {                         \
        set_last_error(e.what());                    \
        return e.code;                               \
    }                                                \
    catch (const std::bad_alloc &) {                 \
        set_last_error("host allocation failed");    \
        return BFGPU_ERR_NOMEM;                      \
    }                                                \
    catch (const std::exception &e) {                \
        set_last_error(e.what());                    \
        return BFGPU_ERR_STATE;                      \
    }                                                \
    return BFGPU_OK;

extern "C" const char *bfgpu_last_error(void) { return g_last_error.c_str(); }
extern "C" const char *bfgpu_version(void) { return "bfgpu 0.1 (gfx950)"; }

extern "C" int bfgpu_ctx_create(int device, bfgpu_ctx **out) {
    BFG_API_BEGIN
    BFG_REQUIRE(out, "ctx_create: null out");
    int n = 0;
    BFG_HIP(hipGetDeviceCount(&n));
    BFG_REQUIRE(device >= 0 && device < n, "ctx_create: no such device");
    BFG_HIP(hipSetDevice(device));
    auto *h = new bfgpu_ctx{};
    h->c.device = device;
    hipError_t e = hipStreamCreateWithFlags(&h->c.own, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete h;
        throw Error(BFGPU_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    h->c.stream = h->c.own;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) h->c.n_cu = prop.multiProcessorCount;
    noalloc_init(device);  // device address of the allocation-free matvec's scratch array
    *out = h;
    BFG_API_END
}

extern "C" int bfgpu_ctx_destroy(bfgpu_ctx *ctx) {
    if (!ctx) return BFGPU_OK;
    if (ctx->c.own) (void)hipStreamDestroy(ctx->c.own);
    delete ctx;
    return BFGPU_OK;
}

extern "C" int bfgpu_ctx_set_stream(bfgpu_ctx *ctx, void *s) {
    BFG_API_BEGIN
    BFG_REQUIRE(ctx, "ctx_set_stream: null ctx");
    ctx->c.stream = static_cast<hipStream_t>(s);
    BFG_API_END
}

extern "C" int bfgpu_ctx_reset_stream(bfgpu_ctx *ctx) {
    BFG_API_BEGIN
    BFG_REQUIRE(ctx, "ctx_reset_stream: null ctx");
    ctx->c.stream = ctx->c.own;
    BFG_API_END
}

extern "C" int bfgpu_ctx_get_stream(bfgpu_ctx *ctx, void **s) {
    BFG_API_BEGIN
    BFG_REQUIRE(ctx && s, "ctx_get_stream: null argument");
    *s = ctx->c.stream;
    BFG_API_END
}

extern "C" int bfgpu_ctx_sync(bfgpu_ctx *ctx) {
    BFG_API_BEGIN
    BFG_REQUIRE(ctx, "ctx_sync: null ctx");
    BFG_HIP(hipStreamSynchronize(ctx->c.stream));
    BFG_API_END
}

static void check_opts(const bfgpu_opts *o) {
    BFG_REQUIRE(o, "factorize: null opts");
    BFG_REQUIRE(o->precision == BFGPU_C64 || o->precision == BFGPU_C128, "factorize: precision must be 1 or 2");
    BFG_REQUIRE(o->tol > 0.0 && o->tol < 1.0, "factorize: tol must lie in (0, 1)");
    BFG_REQUIRE(o->k > 0.0, "factorize: wavenumber must be positive");
    BFG_REQUIRE(o->rank_mode >= 0 && o->rank_mode <= 2, "factorize: unknown rank_mode");
    BFG_REQUIRE(o->workspace_bytes >= 0, "factorize: workspace_bytes must be >= 0 (0: automatic)");
    BFG_REQUIRE(o->judge_rows >= 0 && o->judge_rows <= 4096, "factorize: judge_rows must lie in [0, 4096]");
    BFG_REQUIRE(std::isfinite(o->judge_slack), "factorize: judge_slack must be finite");
}

extern "C" int bfgpu_factorize(bfgpu_ctx *ctx, const bfgpu_tree *test, const bfgpu_tree *trial,
                               int32_t n_pairs, const int32_t *pairs, const bfgpu_opts *opts,
                               bfgpu_bf **out) {
    BFG_API_BEGIN
    BFG_REQUIRE(ctx && test && trial && pairs && out, "factorize: null argument");
    BFG_REQUIRE(n_pairs >= 1, "factorize: need at least one block");
    check_opts(opts);
    if (const KernelRWG *kr = kernel_of(opts->kernel))
        BFG_REQUIRE(kernel_functions(*kr, true) >= test->t.n_values && kernel_functions(*kr, false) >= trial->t.n_values,
                    "factorize: the kernel plug-in has fewer functions than the trees have values");
    BFG_HIP(hipSetDevice(ctx->c.device));
    std::vector<std::pair<int32_t, int32_t>> pr;
    for (int32_t i = 0; i < n_pairs; ++i) pr.emplace_back(pairs[2 * i], pairs[2 * i + 1]);
    auto *h = new bfgpu_bf{};
    h->b.ctx = &ctx->c;
    h->b.test = &test->t;
    h->b.trial = &trial->t;
    h->b.precision = opts->precision;
    try {
        factorize_all(h->b, *opts, pr);
    } catch (...) {
        delete h;
        throw;
    }
    *out = h;
    BFG_API_END
}

extern "C" int bfgpu_bf_from_host(bfgpu_ctx *ctx, const bfgpu_tree *test, const bfgpu_tree *trial,
                                  int32_t n_stages, int64_t n_blocks, const int32_t *table,
                                  const int64_t *data_ofs, const double *data, int32_t precision,
                                  bfgpu_bf **out) {
    BFG_API_BEGIN
    BFG_REQUIRE(ctx && test && trial && table && data_ofs && out, "bf_from_host: null argument");
    BFG_REQUIRE(n_stages >= 2, "bf_from_host: need at least Q and P stages");
    BFG_REQUIRE(precision == BFGPU_C64 || precision == BFGPU_C128, "bf_from_host: bad precision");
    BFG_HIP(hipSetDevice(ctx->c.device));
    auto *h = new bfgpu_bf{};
    try {
        BF &bf = h->b;
        bf.ctx = &ctx->c;
        bf.test = &test->t;
        bf.trial = &trial->t;
        bf.precision = precision;
        bf.bfs.resize(1);
        OneBF &ob = bf.bfs[0];
        ob.st.resize(n_stages);
        bf.n_stage_launch = n_stages;
        std::vector<std::vector<double>> hd(n_stages);
        std::vector<float> hf;
        for (int64_t q = 0; q < n_blocks; ++q) {
            const int32_t *r = table + 10 * q;
            int32_t s = r[0];
            BFG_REQUIRE(s >= 0 && s < n_stages, "bf_from_host: stage out of range");
            BFG_REQUIRE(r[5] >= 0 && r[6] >= 0, "bf_from_host: negative block size");
            Block b{r[1], r[2], r[3], r[4], r[5], r[6], r[7] ? 1 : 0, -1, -1, -1};
            const Tree &ot = (s == n_stages - 1) ? test->t : test->t;
            (void)ot;
            if (!b.ident) {
                BFG_REQUIRE(data && data_ofs[q] >= 0, "bf_from_host: missing block data");
                b.fofs = int64_t(hd[s].size() / 2);
                const double *src = data + 2 * data_ofs[q];
                hd[s].insert(hd[s].end(), src, src + 2 * int64_t(b.m) * b.n);
            } else {
                BFG_REQUIRE(b.m == b.n, "bf_from_host: identity block must be square");
            }
            ob.st[s].b.push_back(b);
        }
        BFG_REQUIRE(!ob.st[0].b.empty() && !ob.st[n_stages - 1].b.empty(), "bf_from_host: empty Q or P stage");
        ob.no = ob.st[0].b[0].obs_out;
        ob.ns = ob.st[n_stages - 1].b[0].src_out;
        for (const Block &b : ob.st[0].b)
            BFG_REQUIRE(b.src_in >= 0 && b.src_in < trial->t.n_nodes, "bf_from_host: Q source node out of range");
        for (const Block &b : ob.st[n_stages - 1].b) {
            BFG_REQUIRE(b.obs_out >= 0 && b.obs_out < test->t.n_nodes, "bf_from_host: P observer node out of range");
            BFG_REQUIRE(b.ident || b.m == test->t.nvals(b.obs_out), "bf_from_host: P rows differ from the leaf size");
        }
        for (const Block &b : ob.st[0].b)
            BFG_REQUIRE(b.ident || b.n == trial->t.nvals(b.src_in), "bf_from_host: Q columns differ from the leaf size");
        bf.factors.resize(n_stages);
        bf.factor_elems.assign(n_stages, 0);
        for (int32_t s = 0; s < n_stages; ++s) {
            int64_t ne = int64_t(hd[s].size() / 2);
            bf.factor_elems[s] = ne;
            const size_t es = precision == BFGPU_C128 ? 16 : 8;
            bf.factors[s].alloc(size_t(ne + FACTOR_PAD) * es);
            BFG_HIP(hipMemset(static_cast<char *>(bf.factors[s].p) + size_t(ne) * es, 0, size_t(FACTOR_PAD) * es));
            if (precision == BFGPU_C128) {
                if (ne) BFG_HIP(hipMemcpy(bf.factors[s].p, hd[s].data(), size_t(ne) * 16, hipMemcpyHostToDevice));
            } else {
                hf.assign(hd[s].begin(), hd[s].end());
                if (ne) BFG_HIP(hipMemcpy(bf.factors[s].p, hf.data(), size_t(ne) * 8, hipMemcpyHostToDevice));
            }
        }
        ob.stage_data_base.assign(n_stages, 0);
        pack_tables(bf);
    } catch (...) {
        delete h;
        throw;
    }
    *out = h;
    BFG_API_END
}

extern "C" int bfgpu_bf_info_get(const bfgpu_bf *h, bfgpu_bf_info *o) {
    BFG_API_BEGIN
    BFG_REQUIRE(h && o, "bf_info_get: null argument");
    const BF &bf = h->b;
    std::memset(o, 0, sizeof(*o));
    o->n_bf = int32_t(bf.bfs.size());
    o->precision = bf.precision;
    o->rows = bf.rows;
    o->cols = bf.cols;
    o->max_stages = bf.n_stage_launch;
    o->n_blocks = bf.n_blocks;
    o->n_data = bf.n_data;
    const int es = bf.precision == BFGPU_C128 ? 16 : 8;
    for (int64_t e : bf.factor_elems) o->factor_bytes += e * es;
    o->slot_len = bf.slot_len;
    o->n_comp = bf.n_comp;
    o->n_retry = bf.n_retry;
    o->n_judged = bf.n_judged;
    o->n_fallback = bf.n_fallback;
    o->t_factor_ms = bf.t_factor_ms;
    o->t_plan_ms = bf.t_plan_ms;
    o->t_sample_ms = bf.t_sample_ms;
    o->t_qr_ms = bf.t_qr_ms;
    o->t_interp_ms = bf.t_interp_ms;
    o->t_pack_ms = bf.t_pack_ms;
    int64_t dev = 0;
    for (const DevBuf &d : bf.factors) dev += int64_t(d.bytes);
    auto tables = [&](const StageTable &t) {
        for (const DevBuf *d : {&t.items, &t.contribs, &t.tasks, &t.tasks_c, &t.mtiles, &t.mtasks})
            dev += int64_t(d->bytes);
    };
    for (const std::vector<StageTable> *v : {&bf.fwd, &bf.tr})
        for (const StageTable &t : *v) {
            tables(t);
            if (t.pre) tables(*t.pre);
        }
    for (const DevBuf *d : {&bf.perm_test, &bf.perm_trial, &bf.slots, &bf.xy_host_stage, &bf.sweep.tasks,
                            &bf.sweep.contribs, &bf.sweep.counters, &bf.sweep.err})
        dev += int64_t(d->bytes);
    o->device_bytes = dev;
    BFG_API_END
}

extern "C" int bfgpu_bf_counts(const bfgpu_bf *h, int32_t index, int32_t *n_stages, int64_t *n_blocks,
                               int64_t *n_data, int64_t *n_skel, int64_t *n_skel_idx, int64_t *n_comp,
                               int32_t roots[2]) {
    BFG_API_BEGIN
    BFG_REQUIRE(h, "bf_counts: null handle");
    BFG_REQUIRE(index >= 0 && index < int32_t(h->b.bfs.size()), "bf_counts: index out of range");
    const OneBF &ob = h->b.bfs[index];
    int64_t nb = 0, nd = 0;
    for (const auto &st : ob.st) {
        nb += int64_t(st.b.size());
        for (const Block &b : st.b)
            if (!b.ident) nd += int64_t(b.m) * b.n;
    }
    if (n_stages) *n_stages = int32_t(ob.st.size());
    if (n_blocks) *n_blocks = nb;
    if (n_data) *n_data = nd;
    if (n_skel) *n_skel = int64_t(ob.skel.size());
    if (n_skel_idx) *n_skel_idx = int64_t(ob.skel_idx.size());
    if (n_comp) *n_comp = int64_t(ob.comp.size());
    if (roots) { roots[0] = ob.no; roots[1] = ob.ns; }
    BFG_API_END
}

extern "C" int bfgpu_bf_export(const bfgpu_bf *h, int32_t index, int32_t *table, int64_t *data_ofs,
                               double *data) {
    BFG_API_BEGIN
    BFG_REQUIRE(h && table && data_ofs, "bf_export: null argument");
    const BF &bf = h->b;
    BFG_REQUIRE(index >= 0 && index < int32_t(bf.bfs.size()), "bf_export: index out of range");
    const OneBF &ob = bf.bfs[index];
    const int32_t nst = int32_t(ob.st.size()), Lmax = bf.n_stage_launch - 1;
    const int es = bf.precision == BFGPU_C128 ? 16 : 8;
    // factors never change after construction: download once and keep a
    // host copy when it is small; large factorizations copy block by block
    int64_t total_bytes = 0;
    for (int64_t e : bf.factor_elems) total_bytes += e * es;
    const bool cached = total_bytes <= (int64_t(2) << 30);
    std::vector<std::vector<double>> &host = bf.export_cache;
    if (data && cached && host.empty()) {
        host.resize(bf.n_stage_launch);
        BFG_HIP(hipStreamSynchronize(bf.ctx->stream));
        for (int32_t L = 0; L < bf.n_stage_launch; ++L) {
            int64_t ne = bf.factor_elems[L];
            host[L].resize(size_t(2 * ne));
            if (!ne) continue;
            if (es == 16) {
                BFG_HIP(hipMemcpy(host[L].data(), bf.factors[L].p, size_t(ne) * 16, hipMemcpyDeviceToHost));
            } else {
                std::vector<float> tmp(size_t(2 * ne));
                BFG_HIP(hipMemcpy(tmp.data(), bf.factors[L].p, size_t(ne) * 8, hipMemcpyDeviceToHost));
                std::copy(tmp.begin(), tmp.end(), host[L].begin());
            }
        }
    }
    if (data && !cached) BFG_HIP(hipStreamSynchronize(bf.ctx->stream));
    int64_t q = 0, d = 0;
    for (int32_t s = 0; s < nst; ++s) {
        int32_t L = (s == nst - 1) ? Lmax : s;
        for (const Block &b : ob.st[s].b) {
            int32_t *r = table + 10 * q;
            r[0] = s; r[1] = b.obs_out; r[2] = b.src_out; r[3] = b.obs_in; r[4] = b.src_in;
            r[5] = b.m; r[6] = b.n; r[7] = b.ident; r[8] = b.out_ptr; r[9] = b.in_ptr;
            if (b.ident) {
                data_ofs[q] = -1;
            } else {
                data_ofs[q] = d;
                if (data) {
                    const int64_t cnt = int64_t(b.m) * b.n;
                    BFG_REQUIRE(b.fofs + cnt <= bf.factor_elems[L], "bf_export: block outside its stage buffer");
                    if (cached) {
                        std::memcpy(data + 2 * d, host[L].data() + 2 * b.fofs, sizeof(double) * 2 * size_t(cnt));
                    } else if (es == 16) {
                        BFG_HIP(hipMemcpy(data + 2 * d, static_cast<const char *>(bf.factors[L].p) + size_t(b.fofs) * 16,
                                          size_t(cnt) * 16, hipMemcpyDeviceToHost));
                    } else {
                        std::vector<float> tmp(size_t(2 * cnt));
                        BFG_HIP(hipMemcpy(tmp.data(), static_cast<const char *>(bf.factors[L].p) + size_t(b.fofs) * 8,
                                          size_t(cnt) * 8, hipMemcpyDeviceToHost));
                        std::copy(tmp.begin(), tmp.end(), data + 2 * d);
                    }
                }
                d += int64_t(b.m) * b.n;
            }
            ++q;
        }
    }
    BFG_API_END
}

extern "C" int bfgpu_bf_export_logs(const bfgpu_bf *h, int32_t index, int32_t *skel_rec,
                                    int32_t *skel_idx, int32_t *comp_rec) {
    BFG_API_BEGIN
    BFG_REQUIRE(h, "bf_export_logs: null handle");
    BFG_REQUIRE(index >= 0 && index < int32_t(h->b.bfs.size()), "bf_export_logs: index out of range");
    const OneBF &ob = h->b.bfs[index];
    if (skel_rec)
        for (size_t i = 0; i < ob.skel.size(); ++i) {
            const SkelRec &r = ob.skel[i];
            int32_t *o = skel_rec + 5 * i;
            o[0] = r.stage; o[1] = r.s; o[2] = r.o; o[3] = r.ofs; o[4] = r.len;
        }
    if (skel_idx && !ob.skel_idx.empty())
        std::memcpy(skel_idx, ob.skel_idx.data(), sizeof(int32_t) * ob.skel_idx.size());
    if (comp_rec)
        for (size_t i = 0; i < ob.comp.size(); ++i) {
            const CompRec &r = ob.comp[i];
            int32_t *o = comp_rec + 6 * i;
            o[0] = r.stage; o[1] = r.s; o[2] = r.o; o[3] = r.guess; o[4] = r.rows; o[5] = r.rank;
        }
    BFG_API_END
}

extern "C" int bfgpu_bf_export_judge(const bfgpu_bf *h, int32_t index, int32_t *judge) {
    BFG_API_BEGIN
    BFG_REQUIRE(h && judge, "bf_export_judge: null argument");
    BFG_REQUIRE(index >= 0 && index < int32_t(h->b.bfs.size()), "bf_export_judge: index out of range");
    const OneBF &ob = h->b.bfs[index];
    for (size_t i = 0; i < ob.comp.size(); ++i) judge[i] = ob.comp[i].judge;
    BFG_API_END
}

extern "C" int bfgpu_bf_export_features(const bfgpu_bf *h, int32_t index, float *feat) {
    BFG_API_BEGIN
    BFG_REQUIRE(h && feat, "bf_export_features: null argument");
    BFG_REQUIRE(index >= 0 && index < int32_t(h->b.bfs.size()), "bf_export_features: index out of range");
    const OneBF &ob = h->b.bfs[index];
    for (size_t i = 0; i < ob.comp.size(); ++i)
        for (int q = 0; q < 6; ++q) feat[6 * i + q] = ob.comp[i].feat[q];
    BFG_API_END
}

extern "C" int bfgpu_bf_compressions(const bfgpu_bf *h, int64_t *n, int32_t *rec, float *feat) {
    BFG_API_BEGIN
    BFG_REQUIRE(h && n, "bf_compressions: null argument");
    const BF &bf = h->b;
    const int64_t nb = int64_t(bf.bfs.size());
    std::vector<int64_t> ofs(nb + 1, 0);
    for (int64_t b = 0; b < nb; ++b) ofs[b + 1] = ofs[b] + int64_t(bf.bfs[b].comp.size());
    *n = ofs[nb];
    if (!rec && !feat) return BFGPU_OK;
    parallel_for(nb, [&](int64_t b) {
        const auto &cs = bf.bfs[b].comp;
        for (size_t i = 0; i < cs.size(); ++i) {
            const int64_t q = ofs[b] + int64_t(i);
            const CompRec &c = cs[i];
            if (rec) {
                int32_t *r = rec + 7 * q;
                r[0] = int32_t(b);
                r[1] = c.stage;
                r[2] = c.s;
                r[3] = c.o;
                r[4] = c.guess;
                r[5] = c.rows;
                r[6] = c.rank;
            }
            if (feat)
                for (int k = 0; k < 6; ++k) feat[6 * q + k] = c.feat[k];
        }
    }, 1024);
    BFG_API_END
}

extern "C" int bfgpu_bf_stage_stats(const bfgpu_bf *h, int32_t *n_stages, int64_t *stats) {
    BFG_API_BEGIN
    BFG_REQUIRE(h && n_stages, "bf_stage_stats: null argument");
    const BF &bf = h->b;
    *n_stages = bf.n_stage_launch;
    if (!stats) return BFGPU_OK;
    const int es = bf.precision == BFGPU_C128 ? 16 : 8;
    const int32_t Lmax = bf.n_stage_launch - 1;
    for (int32_t L = 0; L <= Lmax; ++L) {
        int64_t nb = 0, ni = 0, mr = 0, by = 0;
        for (const OneBF &ob : bf.bfs) {
            int32_t nst = int32_t(ob.st.size());
            for (int32_t s = 0; s < nst; ++s) {
                if (((s == nst - 1) ? Lmax : s) != L) continue;
                for (const Block &b : ob.st[s].b) {
                    ++nb;
                    if (b.ident) { ++ni; continue; }
                    mr = std::max<int64_t>(mr, (s == nst - 1) ? b.n : b.m);
                    by += int64_t(b.m) * b.n * es;
                }
            }
        }
        stats[4 * L + 0] = nb;
        stats[4 * L + 1] = ni;
        stats[4 * L + 2] = mr;
        stats[4 * L + 3] = by;
    }
    BFG_API_END
}

extern "C" int bfgpu_bf_stream_stats(const bfgpu_bf *h, int32_t *n_levels, int64_t *stats) {
    BFG_API_BEGIN
    BFG_REQUIRE(h && n_levels, "bf_stream_stats: null argument");
    const BF &bf = h->b;
    *n_levels = bf.n_stage_launch;
    if (!stats) return BFGPU_OK;
    for (int32_t L = 0; L < bf.n_stage_launch; ++L) {
        int64_t *o = stats + 6 * L;
        std::fill(o, o + 6, int64_t(0));
        for (const StageTable *u : {&bf.fwd[L], static_cast<const StageTable *>(bf.fwd[L].pre.get())}) {
            if (!u || !u->stream_built) continue;
            o[0] = u->stream_nld;
            o[1] += u->n_schunks;
            o[2] += u->n_stasks;
            o[3] += u->n_srest;
            o[4] += u->stream_gap;
            o[5] += u->stream_xent;
        }
    }
    BFG_API_END
}

extern "C" int bfgpu_bf_launch_stats(const bfgpu_bf *h, int32_t op, int32_t *n_levels, int64_t *stats) {
    BFG_API_BEGIN
    BFG_REQUIRE(h && n_levels, "bf_launch_stats: null argument");
    BFG_REQUIRE(op == BFGPU_OP_N || op == BFGPU_OP_T || op == BFGPU_OP_C, "bf_launch_stats: bad op");
    const BF &bf = h->b;
    const int32_t nL = bf.n_stage_launch;
    *n_levels = nL;
    if (!stats) return BFGPU_OK;
    const bool fwd = op == BFGPU_OP_N;
    std::vector<uint8_t> mark;
    for (int32_t i = 0; i < nL; ++i) {
        const int32_t L = fwd ? i : nL - 1 - i;
        const StageTable &t = fwd ? bf.fwd[L] : bf.tr[L];
        int64_t *o = stats + 10 * i;
        // distinct input entries the level reads (partial sums excluded)
        // and output entries it writes (read and written for y)
        const int64_t span = t.in_kind == IN_SLOT ? bf.slot_len : std::max(bf.test->n_values, bf.trial->n_values);
        mark.assign(size_t(span), 0);
        auto mark_tab = [&](const StageTable &u) {
            parallel_for(int64_t(u.h_contribs.size()), [&](int64_t c) {
                const MvContrib &k = u.h_contribs[c];
                const int32_t len = fwd ? k.n : k.m;
                if (u.in_kind == IN_SLOT && k.in_ofs >= bf.part_ofs && k.in_ofs < bf.gather_ofs) return;
                std::fill(mark.begin() + k.in_ofs, mark.begin() + k.in_ofs + len, uint8_t(1));
            }, 4096);
        };
        mark_tab(t);
        if (t.pre) mark_tab(*t.pre);
        int64_t uniq = 0;
        for (uint8_t m : mark) uniq += m;
        int64_t outs = 0;
        for (const MvItem &it : t.h_items) outs += it.out_len;
        o[8] = uniq;
        o[9] = outs * (t.out_kind == OUT_SCATTER ? 2 : 1);
        o[0] = t.n_items;
        o[1] = t.n_contrib + (t.pre ? t.pre->n_contrib : 0);
        o[2] = std::max(t.max_nc, t.pre ? t.pre->max_nc : 0);
        o[3] = t.factor_elems + (t.pre ? t.pre->factor_elems : 0);
        o[4] = t.alg_bytes_1 + (t.pre ? t.pre->alg_bytes_1 : 0);
        o[5] = t.variant;
        o[6] = pack_tile(t.tile);
        o[7] = t.pre ? t.pre->n_items : 0;
    }
    BFG_API_END
}

extern "C" int bfgpu_bf_convert(bfgpu_bf *h, int32_t precision, bfgpu_bf **out) {
    BFG_API_BEGIN
    BFG_REQUIRE(h && out, "bf_convert: null argument");
    BFG_REQUIRE(precision == BFGPU_C64 || precision == BFGPU_C128, "bf_convert: bad precision");
    const BF &src = h->b;
    auto *n = new bfgpu_bf{};
    try {
        BF &bf = n->b;
        bf.ctx = src.ctx;
        bf.test = src.test;
        bf.trial = src.trial;
        bf.bfs = src.bfs;
        bf.n_stage_launch = src.n_stage_launch;
        bf.factor_elems = src.factor_elems;
        bf.precision = precision;
        bf.n_comp = src.n_comp;
        bf.n_retry = src.n_retry;
        bf.n_judged = src.n_judged;
        bf.n_fallback = src.n_fallback;
        bf.t_factor_ms = src.t_factor_ms;
        bf.pf_k = src.pf_k;
        bf.use_pfree = src.use_pfree;
        bf.factors.resize(src.factors.size());
        const int es = precision == BFGPU_C128 ? 16 : 8;
        for (size_t L = 0; L < src.factors.size(); ++L) {
            bf.factors[L].alloc(size_t(src.factor_elems[L] + FACTOR_PAD) * es);
            BFG_HIP(hipMemset(static_cast<char *>(bf.factors[L].p) + size_t(src.factor_elems[L]) * es, 0,
                              size_t(FACTOR_PAD) * es));
            launch_convert(src.ctx->stream, src.factors[L].p, src.precision, bf.factors[L].p, precision,
                           src.factor_elems[L]);
        }
        BFG_HIP(hipStreamSynchronize(src.ctx->stream));
        pack_tables(bf);
    } catch (...) {
        delete n;
        throw;
    }
    *out = n;
    BFG_API_END
}

extern "C" int bfgpu_matvec(bfgpu_bf *h, int32_t op, int64_t nrhs, const void *x, int64_t ldx,
                            void *y, int64_t ldy, const double *alpha, const double *beta,
                            int32_t flags) {
    BFG_API_BEGIN
    BFG_REQUIRE(h, "matvec: null handle");
    static const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    matvec(h->b, op, nrhs, x, ldx, y, ldy, alpha ? alpha : one, beta ? beta : zero, flags);
    BFG_API_END
}

extern "C" int bfgpu_matvec_noalloc(bfgpu_bf *h, int32_t op, int64_t nrhs, const void *x, int64_t ldx,
                                    void *y, int64_t ldy, const double *alpha, const double *beta) {
    BFG_API_BEGIN
    BFG_REQUIRE(h, "matvec_noalloc: null handle");
    static const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    matvec_noalloc(h->b, op, nrhs, x, ldx, y, ldy, alpha ? alpha : one, beta ? beta : zero);
    BFG_API_END
}

extern "C" int bfgpu_matvec_timed(bfgpu_bf *h, int32_t op, int64_t nrhs, const void *x, int64_t ldx,
                                  void *y, int64_t ldy, int32_t iters, double *stage_ms,
                                  int64_t *stage_bytes, int32_t *n_stages, double *total_ms) {
    BFG_API_BEGIN
    BFG_REQUIRE(h && x && y, "matvec_timed: null argument");
    BFG_REQUIRE(op == BFGPU_OP_N || op == BFGPU_OP_T || op == BFGPU_OP_C, "matvec_timed: bad op");
    const bool fwd = op == BFGPU_OP_N;
    BFG_REQUIRE(ldx >= (fwd ? h->b.cols : h->b.rows) && ldy >= (fwd ? h->b.rows : h->b.cols),
                "matvec_timed: leading dimension too small");
    matvec_timed(h->b, op, nrhs, x, ldx, y, ldy, iters, stage_ms, stage_bytes, n_stages, total_ms);
    BFG_API_END
}

extern "C" int bfgpu_matvec_launch_times(bfgpu_bf *h, int32_t op, int64_t nrhs, const void *x, int64_t ldx,
                                         void *y, int64_t ldy, int32_t iters, int32_t cap, int32_t *n_launch,
                                         double *launch_ms, int32_t *desc) {
    BFG_API_BEGIN
    BFG_REQUIRE(h && x && y && n_launch, "matvec_launch_times: null argument");
    BFG_REQUIRE(op == BFGPU_OP_N || op == BFGPU_OP_T || op == BFGPU_OP_C, "matvec_launch_times: bad op");
    BFG_REQUIRE(nrhs >= 1 && cap >= 0, "matvec_launch_times: nrhs >= 1, cap >= 0");
    const bool fwd = op == BFGPU_OP_N;
    BFG_REQUIRE(ldx >= (fwd ? h->b.cols : h->b.rows) && ldy >= (fwd ? h->b.rows : h->b.cols),
                "matvec_launch_times: leading dimension too small");
    matvec_launch_times(h->b, op, nrhs, x, ldx, y, ldy, iters, cap, n_launch, launch_ms, desc);
    BFG_API_END
}

// Tuning state: per level, the forward table, its partial-sum table, the
// transposed table and its partial-sum table, each (variant, packed shape);
// (-1, -1) where a level has no partial-sum table.
extern "C" int bfgpu_bf_get_tuning(const bfgpu_bf *h, int32_t *n, int32_t *vals) {
    BFG_API_BEGIN
    BFG_REQUIRE(h && n, "bf_get_tuning: null argument");
    const BF &bf = h->b;
    *n = 8 * bf.n_stage_launch;
    if (!vals) return BFGPU_OK;
    for (int32_t L = 0; L < bf.n_stage_launch; ++L) {
        const StageTable *t[4] = {&bf.fwd[L], bf.fwd[L].pre.get(), &bf.tr[L], bf.tr[L].pre.get()};
        for (int q = 0; q < 4; ++q) {
            vals[8 * L + 2 * q] = t[q] ? t[q]->variant : -1;
            vals[8 * L + 2 * q + 1] = t[q] ? int32_t(pack_tile(t[q]->tile)) : -1;
        }
    }
    BFG_API_END
}

extern "C" int bfgpu_bf_set_tuning(bfgpu_bf *h, int32_t n, const int32_t *vals) {
    BFG_API_BEGIN
    BFG_REQUIRE(h && vals, "bf_set_tuning: null argument");
    BF &bf = h->b;
    BFG_REQUIRE(n == 8 * bf.n_stage_launch, "bf_set_tuning: wrong length (levels differ)");
    // the transposed copy, where a transposed table asks for it
    for (int32_t L = 0; L < bf.n_stage_launch; ++L)
        for (int q = 2; q < 4; ++q)
            if (vals[8 * L + 2 * q] == 8 || vals[8 * L + 2 * q] == 9) ensure_tcopy(bf);
    // validate everything before changing anything
    for (int pass = 0; pass < 2; ++pass)
        for (int32_t L = 0; L < bf.n_stage_launch; ++L) {
            StageTable *t[4] = {&bf.fwd[L], bf.fwd[L].pre.get(), &bf.tr[L], bf.tr[L].pre.get()};
            for (int q = 0; q < 4; ++q) {
                const int32_t v = vals[8 * L + 2 * q], p = vals[8 * L + 2 * q + 1];
                if (!t[q]) {
                    BFG_REQUIRE(v == -1 && p == -1, "bf_set_tuning: partial-sum tables differ");
                    continue;
                }
                const TileShape sh = unpack_tile(p);
                BFG_REQUIRE(((v >= 0 && v <= 3) || (v == 8 && q >= 2) || v == 9) && !(q >= 2 && v == 1) && p >= 0 &&
                                p < TILE_PACK_LIMIT && tile_ok(sh) && !(v >= 3 && sh.g_shift != 0),
                            "bf_set_tuning: bad variant or shape");
                BFG_REQUIRE((v != 8 && !(v == 9 && q >= 2)) || (bf.tcopy_built && (bf.factor_elems[L] == 0 ||
                                                          (L < int32_t(bf.factors_t.size()) && bf.factors_t[L].p))),
                            "bf_set_tuning: the transposed copy of a level is not available (tcopy off, or "
                            "device memory short)");
                if (pass == 1) {
                    if (v == 9)
                        build_stream(*t[q], bf.precision == BFGPU_C128 ? 16 : 8, sh.snld4 ? 4 : 8,
                                     bf.factor_elems[L], q >= 2);
                    t[q]->variant = v;
                    t[q]->tile = sh;
                }
            }
        }
    release_graph(bf);
    BFG_API_END
}

extern "C" int bfgpu_bf_set_fused(bfgpu_bf *h, int32_t enable) {
    BFG_API_BEGIN
    BFG_REQUIRE(h, "bf_set_fused: null handle");
    h->b.use_sweep = enable != 0;
    if (h->b.use_sweep && !h->b.sweep.built) build_sweep(h->b);
    BFG_API_END
}

extern "C" int bfgpu_bf_set_option(bfgpu_bf *h, int32_t key, int32_t value) {
    BFG_API_BEGIN
    BFG_REQUIRE(h, "bf_set_option: null handle");
    BF &bf = h->b;
    switch (key) {
    case BFGPU_OPT_SWEEP:
        BFG_REQUIRE(value >= 0 && value <= 2, "bf_set_option: sweep mode 0..2");
        bf.use_sweep = value != 0;
        if (value) bf.sweep.mode = value == 1 ? 0 : 1;
        if (bf.use_sweep && !bf.sweep.built) build_sweep(bf);
        break;
    case BFGPU_OPT_WG_PER_CU:
        BFG_REQUIRE(value >= 0 && value <= 32, "bf_set_option: workgroups per CU 0..32");
        bf.sweep.wg_per_cu = value;
        break;
    case BFGPU_OPT_GRAPH:
        bf.use_graph = value != 0;
        release_graph(bf);
        break;
    case BFGPU_OPT_VARIANT:
        BFG_REQUIRE(value >= 0 && value <= 3, "bf_set_option: kernel variant 0..3");
        for (int dir = 0; dir < 2; ++dir)
            for (StageTable &t : dir == 0 ? bf.fwd : bf.tr) {
                const int32_t v = dir == 1 ? (value == 1 ? 0 : value) : value;
                if (value == 3) {  // dense: as many lanes per task as the tiled shape, >= 16
                    for (StageTable *u : {&t, t.pre.get()}) {
                        if (!u) continue;
                        u->tile.rp_shift = int8_t(std::max<int>(4, u->tile.rp_shift + u->tile.g_shift));
                        u->tile.g_shift = 0;
                    }
                }
                t.variant = v;
                if (t.pre) t.pre->variant = v;
            }
        release_graph(bf);
        break;
    case BFGPU_OPT_TILE: {
        const TileShape sh = unpack_tile(value);
        BFG_REQUIRE(value >= 0 && value < TILE_PACK_LIMIT && tile_ok(sh), "bf_set_option: bad tile shape");
        for (int dir = 0; dir < 2; ++dir)
            for (StageTable &t : dir == 0 ? bf.fwd : bf.tr) {
                t.variant = 2;
                t.tile = sh;
                if (t.pre) {
                    t.pre->variant = 2;
                    t.pre->tile = sh;
                }
            }
        release_graph(bf);
        break;
    }
    case BFGPU_OPT_RHS_LANES:
        BFG_REQUIRE(value >= 0, "bf_set_option: rhs_lanes is a minimum number of right-hand sides (0: off)");
        bf.rhs_lanes_min = value;
        release_graph(bf);
        break;
    case BFGPU_OPT_MFMA:
        BFG_REQUIRE(value == 0 || value == 1 || (value >= 16 && value <= MTILE_ROWS && value % 16 == 0),
                    "bf_set_option: mfma is 0, 1 or a tile height of 16, 32, 48 or 64 rows");
        bf.use_mfma = value != 0;
        if (value >= 16) bf.mfma_rows = value;
        release_graph(bf);
        break;
    case BFGPU_OPT_RHS_CHUNK:
        BFG_REQUIRE(value >= 0, "bf_set_option: rhs_chunk is a number of right-hand sides (0: automatic)");
        bf.rhs_chunk = value;
        release_graph(bf);
        break;
    case BFGPU_OPT_SLOT_CAP:
        BFG_REQUIRE(value >= 0, "bf_set_option: slot_cap is a number of KiB (0: as device memory allows)");
        bf.slot_cap_kib = value;
        break;
    case BFGPU_OPT_TCOPY:
        BFG_REQUIRE(value == 0 || value == 1, "bf_set_option: tcopy is 0 or 1");
        if (value == 0) {
            release_tcopy(bf);
            bf.tcopy_mode = 0;
        } else {
            bf.tcopy_mode = 1;
            bf.tcopy_tried = false;
            ensure_tcopy(bf);
        }
        break;
    case BFGPU_OPT_STREAM: {
        BFG_REQUIRE(value == 0 || value == 4 || value == 8, "bf_set_option: stream is 0, 4 or 8");
        const int es = bf.precision == BFGPU_C128 ? 16 : 8;
        for (int32_t L = 0; L < bf.n_stage_launch; ++L)
            for (StageTable *u : {&bf.fwd[L], bf.fwd[L].pre.get()}) {
                if (!u) continue;
                if (value == 0) {
                    if (u->variant == 9) u->variant = 3;
                    continue;
                }
                if (u->variant != 3 && u->variant != 9) {  // dense shape for the tasks that do not stream
                    u->tile.rp_shift = int8_t(std::max<int>(4, std::min<int>(6, u->tile.rp_shift + u->tile.g_shift)));
                    u->tile.g_shift = 0;
                }
                u->tile.snld4 = int8_t(value == 4 ? 1 : 0);
                build_stream(*u, es, value, bf.factor_elems[L]);
                u->variant = 9;
            }
        release_graph(bf);
        break;
    }
    case BFGPU_OPT_MFMA_EMUL:
        BFG_REQUIRE(value == 0 || value == 1, "bf_set_option: mfma_emul is 0 or 1");
        bf.mfma_emul = value != 0;
        release_graph(bf);
        break;
    case BFGPU_OPT_PFREE:
        BFG_REQUIRE(value == 0 || value == 1, "bf_set_option: pfree is 0 or 1");
        bf.use_pfree = value != 0;
        release_graph(bf);
        break;
    case BFGPU_OPT_DENSE: {
        const TileShape sh = unpack_tile(value);
        BFG_REQUIRE(value >= 0 && value < TILE_PACK_LIMIT && tile_ok(sh) && sh.g_shift == 0,
                    "bf_set_option: bad dense shape");
        for (int dir = 0; dir < 2; ++dir)
            for (StageTable &t : dir == 0 ? bf.fwd : bf.tr)
                for (StageTable *u : {&t, t.pre.get()})
                    if (u) {
                        u->variant = 3;
                        u->tile = sh;
                    }
        release_graph(bf);
        break;
    }
    default:
        BFG_REQUIRE(false, "bf_set_option: unknown key");
    }
    BFG_API_END
}

extern "C" int bfgpu_bf_get_option(const bfgpu_bf *h, int32_t key, int32_t *value) {
    BFG_API_BEGIN
    BFG_REQUIRE(h && value, "bf_get_option: null argument");
    const BF &bf = h->b;
    switch (key) {
    case BFGPU_OPT_SWEEP: *value = bf.use_sweep ? (bf.sweep.mode == 0 ? 1 : 2) : 0; break;
    case BFGPU_OPT_WG_PER_CU: *value = bf.sweep.wg_per_cu; break;
    case BFGPU_OPT_GRAPH: *value = bf.use_graph ? 1 : 0; break;
    case BFGPU_OPT_VARIANT: {
        int32_t v = bf.fwd.empty() ? 0 : bf.fwd[0].variant;
        for (int dir = 0; dir < 2; ++dir)
            for (const StageTable &t : dir == 0 ? bf.fwd : bf.tr) {
                const int32_t w = dir == 1 ? (v == 1 ? 0 : v) : v;
                if (t.variant != w || (t.pre && t.pre->variant != w)) v = -1;
            }
        *value = v;
        break;
    }
    case BFGPU_OPT_TILE:
    case BFGPU_OPT_DENSE: *value = bf.fwd.empty() ? 0 : int32_t(pack_tile(bf.fwd[0].tile)); break;
    case BFGPU_OPT_RHS_LANES: *value = bf.rhs_lanes_min; break;
    case BFGPU_OPT_MFMA: *value = bf.use_mfma ? bf.mfma_rows : 0; break;
    case BFGPU_OPT_RHS_CHUNK: *value = bf.rhs_chunk; break;
    case BFGPU_OPT_SLOT_CAP: *value = bf.slot_cap_kib; break;
    case BFGPU_OPT_TCOPY: *value = bf.tcopy_built ? 1 : 0; break;
    case BFGPU_OPT_MFMA_EMUL: *value = bf.mfma_emul ? 1 : 0; break;
    case BFGPU_OPT_STREAM: {
        int32_t v = -1;
        for (const StageTable &t : bf.fwd)
            for (const StageTable *u : {&t, static_cast<const StageTable *>(t.pre.get())}) {
                if (!u) continue;
                const int32_t w = u->variant == 9 ? (u->tile.snld4 ? 4 : 8) : 0;
                v = v < 0 || v == w ? w : 0;
            }
        *value = std::max(v, 0);
        break;
    }
    case BFGPU_OPT_PFREE: *value = bf.use_pfree && bf.pfree_ok ? 1 : 0; break;
    default: BFG_REQUIRE(false, "bf_get_option: unknown key");
    }
    BFG_API_END
}

extern "C" int bfgpu_bf_autotune(bfgpu_bf *h, int64_t nrhs, int32_t iters, double *stage_ms,
                                 int32_t *variants) {
    BFG_API_BEGIN
    BFG_REQUIRE(h, "bf_autotune: null handle");
    autotune(h->b, nrhs, iters, stage_ms, variants);
    BFG_API_END
}

extern "C" int bfgpu_bf_check(bfgpu_bf *h) {
    BFG_API_BEGIN
    BFG_REQUIRE(h, "bf_check: null handle");
    check_sweep_error(h->b);
    BFG_API_END
}

extern "C" int bfgpu_bf_destroy(bfgpu_bf *h) {
    if (h && h->b.ctx) (void)hipStreamSynchronize(h->b.ctx->stream);
    delete h;
    return BFGPU_OK;
}
