// matvec.hip -- packing of reference-layout butterfly blocks into per-stage
// launch tables and the level-by-level application y = alpha op(A) x + beta y.
//
// Semantics follow the flat-workspace mul! (src/Butterflyalgebra/
// bfvector.jl:200-296): Q gathers x over the source leaves into level-1
// skeleton vectors, every R level sums its blocks into the next level's
// vectors, P adds alpha * P v into y over the observer leaves.  Several
// factorizations batched together are summed as the far loop of
// src/matrixalgebra/matrixvector.jl:39-62 does, without per-thread copies.
// The adjoint and transpose (bfadjtr.jl:43-89) run the same tables backwards
// with transposed (conjugated) block access instead of materialised copies.
#include <algorithm>
#include <cstring>
#include <map>
#include <unordered_map>

#include "internal.hpp"

namespace bfg {

namespace {
inline uint64_t key2(int32_t a, int32_t b) { return (uint64_t(uint32_t(a)) << 32) | uint32_t(b); }

struct SlotMap {
    std::unordered_map<uint64_t, std::pair<int64_t, int32_t>> m;  // key -> (offset, len)
};

template <class T> void upload_tab(DevBuf &d, const std::vector<T> &h) {
    d.alloc(sizeof(T) * std::max<size_t>(h.size(), 1));
    if (!h.empty()) BFG_HIP(hipMemcpy(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
}

// merge adjacent contributions that read consecutive inputs with
// consecutive column slices of one column-major block
// (only within one item: out[first..] holds this item's contributions)
void merge_contribs(std::vector<MvContrib> &c, int32_t b, int32_t e, std::vector<MvContrib> &out) {
    const size_t first = out.size();
    for (int32_t i = b; i < e; ++i) {
        const MvContrib &x = c[i];
        if (out.size() > first) {
            MvContrib &p = out.back();
            if (p.fofs >= 0 && x.fofs >= 0 && p.m == x.m && p.fofs + int64_t(p.m) * p.n == x.fofs &&
                p.in_ofs + p.n == x.in_ofs) {
                p.n += x.n;
                continue;
            }
        }
        out.push_back(x);
    }
}
}  // namespace

void build_sweep(BF &bf, const std::vector<std::vector<std::vector<int32_t>>> &fin);

void pack_tables(BF &bf) {
    const Tree &test = *bf.test, &trial = *bf.trial;
    const int32_t nL = bf.n_stage_launch, Lmax = nL - 1;
    const int es = bf.precision == BFGPU_C128 ? 16 : 8;
    const size_t nb = bf.bfs.size();
    // ---- slot offsets, consumer order, all factorizations in one buffer ----
    std::vector<std::vector<SlotMap>> slots(nb);
    int64_t counter = 0;
    for (size_t b = 0; b < nb; ++b) {
        OneBF &ob = bf.bfs[b];
        const int32_t ns = int32_t(ob.st.size());
        slots[b].resize(std::max(ns - 1, 1));
        for (int32_t s = 0; s + 1 < ns; ++s) {
            std::unordered_map<uint64_t, int32_t> len;
            for (const Block &k : ob.st[s].b) len[key2(k.obs_out, k.src_out)] = k.m;
            SlotMap &sm = slots[b][s];
            for (const Block &k : ob.st[s + 1].b) {
                uint64_t kk = key2(k.obs_in, k.src_in);
                if (sm.m.count(kk)) continue;
                auto it = len.find(kk);
                BFG_REQUIRE(it != len.end(), "pack: block reads a skeleton vector no block writes");
                sm.m[kk] = {counter, it->second};
                counter += it->second;
            }
            for (const Block &k : ob.st[s].b) {  // written but never read
                uint64_t kk = key2(k.obs_out, k.src_out);
                if (!sm.m.count(kk)) { sm.m[kk] = {counter, k.m}; counter += k.m; }
            }
        }
        // reference slot ids: first-use order per level (bfstructs.jl:107-145)
        std::vector<std::unordered_map<uint64_t, int32_t>> ids(std::max(ns - 1, 1));
        auto idof = [&](int32_t lvl, uint64_t kk) {
            auto &m = ids[lvl];
            auto it = m.find(kk);
            if (it != m.end()) return it->second;
            int32_t v = int32_t(m.size());
            m[kk] = v;
            return v;
        };
        for (int32_t s = 0; s < ns; ++s)
            for (Block &k : ob.st[s].b) {
                k.in_ptr = s > 0 ? idof(s - 1, key2(k.obs_in, k.src_in)) : -1;
                k.out_ptr = s < ns - 1 ? idof(s, key2(k.obs_out, k.src_out)) : -1;
            }
    }
    bf.slot_len = std::max<int64_t>(counter, 1);
    BFG_REQUIRE(counter < (int64_t(1) << 31), "pack: skeleton vectors exceed 2^31 entries");

    // ---- size (bfdims.jl:54-64) and coverage ----
    if (nb == 1) {
        int32_t mr = -1, mc = -1;
        for (int32_t i = 0; i < test.nvals(bf.bfs[0].no); ++i) mr = std::max(mr, test.vals(bf.bfs[0].no)[i]);
        for (int32_t i = 0; i < trial.nvals(bf.bfs[0].ns); ++i) mc = std::max(mc, trial.vals(bf.bfs[0].ns)[i]);
        bf.rows = mr + 1;
        bf.cols = mc + 1;
    } else {
        bf.rows = test.n_values;
        bf.cols = trial.n_values;
    }

    auto launch_of = [&](const OneBF &ob, int32_t s) {
        return s == int32_t(ob.st.size()) - 1 ? Lmax : s;
    };

    bf.fwd.clear();
    bf.tr.clear();
    bf.fwd.resize(nL);
    bf.tr.resize(nL);
    std::vector<std::vector<std::vector<int32_t>>> fin(nL);
    int64_t covered_rows = 0, covered_cols = 0;
    for (int32_t L = 0; L < nL; ++L) {
        StageTable &F = bf.fwd[L], &T = bf.tr[L];
        F.in_kind = (L == 0) ? IN_GATHER : IN_SLOT;
        F.out_kind = (L == Lmax) ? OUT_SCATTER : OUT_SLOT;
        T.in_kind = (L == Lmax) ? IN_GATHER : IN_SLOT;
        T.out_kind = (L == 0) ? OUT_SCATTER : OUT_SLOT;
        // forward items keyed by output target, transposed by input source
        std::map<std::pair<int64_t, int64_t>, int32_t> fidx, tidx;
        std::vector<std::vector<MvContrib>> fc, tc;
        std::vector<MvItem> fi, ti;
        for (size_t b = 0; b < nb; ++b) {
            const OneBF &ob = bf.bfs[b];
            for (int32_t s = 0; s < int32_t(ob.st.size()); ++s) {
                if (launch_of(ob, s) != L) continue;
                const bool isQ = (s == 0), isP = (s == int32_t(ob.st.size()) - 1);
                for (const Block &k : ob.st[s].b) {
                    // output side (forward target / transposed source)
                    int64_t o_ofs;
                    std::pair<int64_t, int64_t> okey;
                    if (isP) {
                        o_ofs = test.val_begin[k.obs_out];
                        okey = {-1, k.obs_out};
                    } else {
                        o_ofs = slots[b][s].m.at(key2(k.obs_out, k.src_out)).first;
                        okey = {int64_t(b), o_ofs};
                    }
                    // input side
                    int64_t i_ofs;
                    std::pair<int64_t, int64_t> ikey;
                    if (isQ) {
                        i_ofs = trial.val_begin[k.src_in];
                        ikey = {-1, k.src_in};
                    } else {
                        i_ofs = slots[b][s - 1].m.at(key2(k.obs_in, k.src_in)).first;
                        ikey = {int64_t(b), i_ofs};
                    }
                    int64_t fofs = k.ident ? -1 : k.fofs;
                    // forward: out[m] += A (m x n) in[n]
                    auto fit = fidx.find(okey);
                    if (fit == fidx.end()) {
                        fit = fidx.emplace(okey, int32_t(fi.size())).first;
                        fi.push_back(MvItem{int32_t(o_ofs), k.m, 0, 0});
                        fc.emplace_back();
                    }
                    fc[fit->second].push_back(MvContrib{fofs, k.m, k.n, int32_t(i_ofs), 0});
                    // transposed: in[n] += op(A)^T out[m]
                    auto tit = tidx.find(ikey);
                    if (tit == tidx.end()) {
                        tit = tidx.emplace(ikey, int32_t(ti.size())).first;
                        ti.push_back(MvItem{int32_t(i_ofs), k.n, 0, 0});
                        tc.emplace_back();
                    }
                    tc[tit->second].push_back(MvContrib{fofs, k.m, k.n, int32_t(o_ofs), 0});
                    if (k.ident) F.factor_elems += 0;
                    else F.factor_elems += int64_t(k.m) * k.n;
                }
            }
        }
        T.factor_elems = F.factor_elems;
        auto finish = [&](StageTable &S, std::vector<MvItem> &items, std::vector<std::vector<MvContrib>> &cl,
                          bool fwd) {
            S.h_items.clear();
            S.h_contribs.clear();
            int64_t vec = 0;
            for (size_t i = 0; i < items.size(); ++i) {
                MvItem it = items[i];
                it.c_begin = int32_t(S.h_contribs.size());
                if (fwd) merge_contribs(cl[i], 0, int32_t(cl[i].size()), S.h_contribs);
                else S.h_contribs.insert(S.h_contribs.end(), cl[i].begin(), cl[i].end());
                it.c_end = int32_t(S.h_contribs.size());
                S.h_items.push_back(it);
                S.max_out = std::max(S.max_out, it.out_len);
                vec += it.out_len * ((S.out_kind == OUT_SCATTER) ? 2 : 1);
            }
            for (const MvContrib &c : S.h_contribs) {
                int32_t in_len = fwd ? c.n : c.m;
                S.max_in = std::max(S.max_in, in_len);
                vec += in_len;
            }
            S.n_items = int32_t(S.h_items.size());
            S.n_contrib = int32_t(S.h_contribs.size());
            S.alg_bytes_1 = vec * es;
            S.h_tasks.clear();
            for (const MvItem &it : S.h_items) {
                const MvContrib &c0 = S.h_contribs[it.c_begin];
                S.h_tasks.push_back(MvTask{c0.fofs, c0.m, c0.n, c0.in_ofs, it.out_ofs, it.out_len,
                                           it.c_begin + 1, it.c_end, -1, 0, -1, 0});
            }
            // default forward variant until the autotuner measures: a
            // wavefront per task for small blocks, a workgroup otherwise
            S.variant = (S.n_items > 0 && S.factor_elems / S.n_items < 4096) ? 1 : 0;
            upload_tab(S.items, S.h_items);
            upload_tab(S.contribs, S.h_contribs);
            upload_tab(S.tasks, S.h_tasks);
        };
        // input slots of every forward item (before merging), for the sweep
        fin[L].resize(fc.size());
        for (size_t i = 0; i < fc.size(); ++i) {
            for (const MvContrib &c : fc[i]) fin[L][i].push_back(c.in_ofs);
            std::sort(fin[L][i].begin(), fin[L][i].end());
            fin[L][i].erase(std::unique(fin[L][i].begin(), fin[L][i].end()), fin[L][i].end());
        }
        finish(F, fi, fc, true);
        finish(T, ti, tc, false);
        if (L == Lmax)
            for (const MvItem &it : F.h_items) covered_rows += it.out_len;
        if (L == 0)
            for (const MvItem &it : T.h_items) covered_cols += it.out_len;
    }
    bf.y_covered = covered_rows == bf.rows;
    bf.x_covered = covered_cols == bf.cols;
    upload_tab(bf.perm_test, test.perm);
    upload_tab(bf.perm_trial, trial.perm);
    build_sweep(bf, fin);
}

// Fused-sweep tables: every forward task of every stage in stage order.
// Items that read the same set of skeleton vectors form one dependency
// group with one counter; each produced vector is read by exactly one group
// (butterfly data flow), so its producer signals that group's counter once.
void build_sweep(BF &bf, const std::vector<std::vector<std::vector<int32_t>>> &fin) {
    SweepTables &sw = bf.sweep;
    sw.ok = false;
    const int32_t nL = bf.n_stage_launch;
    if (nL > SWEEP_MAX_STAGES) return;
    std::map<std::vector<int32_t>, int32_t> gid;
    std::unordered_map<int32_t, int32_t> consumer;  // slot offset -> group
    std::vector<int32_t> gcnt;
    std::vector<std::vector<int32_t>> wait(nL);
    for (int32_t L = 1; L < nL; ++L) {
        wait[L].resize(fin[L].size());
        for (size_t i = 0; i < fin[L].size(); ++i) {
            const auto &set = fin[L][i];
            auto it = gid.find(set);
            int32_t g;
            if (it == gid.end()) {
                g = int32_t(gcnt.size());
                gid.emplace(set, g);
                gcnt.push_back(int32_t(set.size()));
                for (int32_t s : set) {
                    auto c = consumer.find(s);
                    if (c != consumer.end() && c->second != g) return;  // not a butterfly data flow
                    consumer[s] = g;
                }
            } else {
                g = it->second;
            }
            wait[L][i] = g;
        }
    }
    std::vector<MvTask> tasks;
    std::vector<MvContrib> contribs;
    int32_t max_wait = 1;
    for (int32_t L = 0; L < nL; ++L) {
        const StageTable &S = bf.fwd[L];
        const int32_t base = int32_t(contribs.size());
        contribs.insert(contribs.end(), S.h_contribs.begin(), S.h_contribs.end());
        for (size_t i = 0; i < S.h_tasks.size(); ++i) {
            MvTask t = S.h_tasks[i];
            t.c_begin += base;
            t.c_end += base;
            t.pad = L;
            t.wait_grp = L > 0 ? wait[L][i] : -1;
            t.wait_cnt = L > 0 ? gcnt[wait[L][i]] : 0;
            max_wait = std::max(max_wait, t.wait_cnt);
            t.sig_grp = -1;
            if (L < nL - 1) {
                auto c = consumer.find(t.out_ofs);
                if (c != consumer.end()) t.sig_grp = c->second;
            }
            tasks.push_back(t);
        }
    }
    sw.n_tasks = int32_t(tasks.size());
    sw.n_groups = int32_t(gcnt.size());
    sw.max_wait = max_wait;
    upload_tab(sw.tasks, tasks);
    upload_tab(sw.contribs, contribs);
    sw.counters.alloc(sizeof(int32_t) * size_t(sw.n_groups + 2));
    BFG_HIP(hipMemset(sw.counters.p, 0, sizeof(int32_t) * size_t(sw.n_groups + 2)));
    sw.err.alloc(sizeof(int32_t));
    BFG_HIP(hipMemset(sw.err.p, 0, sizeof(int32_t)));
    sw.epoch = 0;
    sw.ok = true;
}

void check_sweep_error(BF &bf);

bool sweep_ok(const BF &bf, int64_t nrhs) {
    return bf.use_sweep && bf.sweep.ok && nrhs >= 1 && nrhs <= 8 && bf.ctx->n_cu > 0;
}

void release_graph(BF &bf) {
    if (bf.graph_exec) (void)hipGraphExecDestroy(bf.graph_exec);
    bf.graph_exec = nullptr;
    bf.graph_key.clear();
}

BF::~BF() { release_graph(*this); }

static void stage_launches(BF &bf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y,
                           int64_t ldy, const double *alpha, const double *beta,
                           std::vector<hipEvent_t> *ev);

static void stage_seq(BF &bf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y,
                      int64_t ldy, const double *alpha, const double *beta,
                      std::vector<hipEvent_t> *ev) {
    hipStream_t s = bf.ctx->stream;
    const int es = bf.precision == BFGPU_C128 ? 16 : 8;
    size_t need = size_t(bf.slot_len) * size_t(nrhs) * es;
    if (need > bf.slots_cap) {
        release_graph(bf);  // a captured graph holds the old slot buffer
        bf.slots.alloc(need);
        bf.slots_cap = need;
    }
    const bool fused = op == BFGPU_OP_N && sweep_ok(bf, nrhs);
    if (ev || !bf.use_graph || fused) {
        stage_launches(bf, op, nrhs, x, ldx, y, ldy, alpha, beta, ev);
        return;
    }
    // per-level launches replayed from a captured graph while the call
    // signature repeats (iterative solvers reuse their vectors)
    auto bits = [](double v) { int64_t b; std::memcpy(&b, &v, 8); return b; };
    std::vector<int64_t> key = {op, nrhs, int64_t(reinterpret_cast<intptr_t>(x)), ldx,
                                int64_t(reinterpret_cast<intptr_t>(y)), ldy, bits(alpha[0]), bits(alpha[1]),
                                bits(beta[0]), bits(beta[1]), int64_t(reinterpret_cast<intptr_t>(s))};
    if (!bf.graph_exec || key != bf.graph_key) {
        release_graph(bf);
        BFG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipGraph_t g = nullptr;
        try {
            stage_launches(bf, op, nrhs, x, ldx, y, ldy, alpha, beta, nullptr);
        } catch (...) {
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        BFG_HIP(hipStreamEndCapture(s, &g));
        hipError_t e = hipGraphInstantiate(&bf.graph_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        BFG_HIP(e);
        bf.graph_key = key;
    }
    BFG_HIP(hipGraphLaunch(bf.graph_exec, s));
}

static void stage_launches(BF &bf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y,
                           int64_t ldy, const double *alpha, const double *beta,
                           std::vector<hipEvent_t> *ev) {
    hipStream_t s = bf.ctx->stream;
    const int32_t nL = bf.n_stage_launch;
    const bool fwd = (op == BFGPU_OP_N);
    const int64_t out_len = fwd ? bf.rows : bf.cols;
    const bool fuse = fwd ? bf.y_covered : bf.x_covered;
    static const double one[2] = {1.0, 0.0};
    if (!fuse) launch_scale(s, bf.precision, y, out_len, ldy, nrhs, beta);
    const double *b_eff = fuse ? beta : one;
    const int32_t *pt = bf.perm_test.as<const int32_t>(), *pr = bf.perm_trial.as<const int32_t>();
    int q = 0;
    if (ev) BFG_HIP(hipEventRecord((*ev)[q++], s));
    if (fwd && sweep_ok(bf, nrhs)) {
        SweepTables &sw = bf.sweep;
        if (int64_t(sw.epoch + 1) * sw.max_wait >= (int64_t(1) << 30)) {
            launch_zero_i32(s, sw.counters.as<int32_t>(), sw.n_groups);
            sw.epoch = 0;
        }
        ++sw.epoch;
        launch_sweep(s, bf.precision, sw, bf.factors, x, ldx, y, ldy, bf.slots.p, bf.slot_len, pr, pt,
                     nrhs, alpha, b_eff, fuse, bf.ctx->n_cu);
        if (ev) BFG_HIP(hipEventRecord((*ev)[q++], s));
        return;
    }
    for (int32_t i = 0; i < nL; ++i) {
        int32_t L = fwd ? i : nL - 1 - i;
        const StageTable &t = fwd ? bf.fwd[L] : bf.tr[L];
        const int32_t *pin = fwd ? (L == 0 ? pr : nullptr) : (L == nL - 1 ? pt : nullptr);
        const int32_t *pout = fwd ? (L == nL - 1 ? pt : nullptr) : (L == 0 ? pr : nullptr);
        launch_stage(s, bf.precision, op, t, bf.factors[L].p, x, ldx, y, ldy, bf.slots.p,
                     bf.slot_len, pin, pout, nrhs, alpha, b_eff, fuse);
        if (ev) BFG_HIP(hipEventRecord((*ev)[q++], s));
    }
}

void matvec(BF &bf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y, int64_t ldy,
            const double *alpha, const double *beta, int flags) {
    BFG_REQUIRE(op == BFGPU_OP_N || op == BFGPU_OP_T || op == BFGPU_OP_C, "matvec: bad op");
    BFG_REQUIRE(nrhs >= 0, "matvec: nrhs < 0");
    if (nrhs == 0) return;
    BFG_REQUIRE(x && y, "matvec: null x or y");
    const bool fwd = (op == BFGPU_OP_N);
    const int64_t in_len = fwd ? bf.cols : bf.rows, out_len = fwd ? bf.rows : bf.cols;
    BFG_REQUIRE(ldx >= in_len && ldy >= out_len, "matvec: leading dimension too small (DimensionMismatch)");
    const int es = bf.precision == BFGPU_C128 ? 16 : 8;
    hipStream_t s = bf.ctx->stream;
    if (flags & BFGPU_XY_DEVICE) {
        stage_seq(bf, op, nrhs, x, ldx, y, ldy, alpha, beta, nullptr);
        return;
    }
    size_t xb = size_t(in_len) * nrhs * es, yb = size_t(out_len) * nrhs * es;
    bf.xy_host_stage.reserve(xb + yb);
    char *dx = bf.xy_host_stage.as<char>(), *dy = dx + xb;
    BFG_HIP(hipMemcpy2DAsync(dx, size_t(in_len) * es, x, size_t(ldx) * es, size_t(in_len) * es,
                             size_t(nrhs), hipMemcpyHostToDevice, s));
    BFG_HIP(hipMemcpy2DAsync(dy, size_t(out_len) * es, y, size_t(ldy) * es, size_t(out_len) * es,
                             size_t(nrhs), hipMemcpyHostToDevice, s));
    stage_seq(bf, op, nrhs, dx, in_len, dy, out_len, alpha, beta, nullptr);
    BFG_HIP(hipMemcpy2DAsync(y, size_t(ldy) * es, dy, size_t(out_len) * es, size_t(out_len) * es,
                             size_t(nrhs), hipMemcpyDeviceToHost, s));
    BFG_HIP(hipStreamSynchronize(s));
    check_sweep_error(bf);
}

void matvec_timed(BF &bf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y,
                  int64_t ldy, int32_t iters, double *stage_ms, int64_t *stage_bytes,
                  int32_t *n_stages, double *total_ms) {
    BFG_REQUIRE(iters >= 1, "matvec_timed: iters >= 1");
    const int32_t nL = bf.n_stage_launch;
    const int es = bf.precision == BFGPU_C128 ? 16 : 8;
    const bool fwd = (op == BFGPU_OP_N);
    const bool fused = fwd && sweep_ok(bf, nrhs);
    const int32_t nS = fused ? 1 : nL;  // timed launches per matvec
    std::vector<hipEvent_t> ev(nS + 1);
    for (auto &e : ev) BFG_HIP(hipEventCreate(&e));
    std::vector<double> acc(nS, 0.0);
    double tot = 0.0;
    static const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    try {
        for (int32_t it = 0; it < iters; ++it) {
            stage_seq(bf, op, nrhs, x, ldx, y, ldy, one, zero, &ev);
            BFG_HIP(hipEventSynchronize(ev[nS]));
            for (int32_t i = 0; i < nS; ++i) {
                float ms = 0.f;
                BFG_HIP(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
                acc[i] += ms;
            }
            float ms = 0.f;
            BFG_HIP(hipEventElapsedTime(&ms, ev[0], ev[nS]));
            tot += ms;
        }
    } catch (...) {
        for (auto &e : ev) (void)hipEventDestroy(e);
        throw;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    check_sweep_error(bf);
    for (int32_t i = 0; i < nS; ++i) {
        if (stage_ms) stage_ms[i] = acc[i] / iters;
        if (!stage_bytes) continue;
        if (fused) {
            int64_t b = 0;
            for (const StageTable &t : bf.fwd) b += t.factor_elems * es + t.alg_bytes_1 * nrhs;
            stage_bytes[i] = b;
        } else {
            int32_t L = fwd ? i : nL - 1 - i;
            const StageTable &t = fwd ? bf.fwd[L] : bf.tr[L];
            stage_bytes[i] = t.factor_elems * es + t.alg_bytes_1 * nrhs;
        }
    }
    if (n_stages) *n_stages = nS;
    if (total_ms) *total_ms = tot / iters;
}

// Raise (and clear) the fused sweep's error flag: a dependency wait that
// expired means the results of the launch are not to be trusted.
void check_sweep_error(BF &bf) {
    if (!bf.sweep.ok) return;
    int32_t e = 0;
    BFG_HIP(hipStreamSynchronize(bf.ctx->stream));
    BFG_HIP(hipMemcpy(&e, bf.sweep.err.p, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (e) {
        BFG_HIP(hipMemset(bf.sweep.err.p, 0, sizeof(int32_t)));
        throw Error(BFGPU_ERR_STATE, "fused sweep: a dependency wait timed out");
    }
}

}  // namespace bfg

namespace bfg {

// Launch-shape autotuner: for every forward stage, time each kernel
// variant on device buffers and keep the fastest (per-stage choice is
// stored in the stage table and used by all later matvecs).
void autotune(BF &bf, int64_t nrhs, int32_t iters, double *stage_ms, int32_t *variants) {
    BFG_REQUIRE(nrhs >= 1 && iters >= 1, "autotune: nrhs, iters >= 1");
    hipStream_t s = bf.ctx->stream;
    const int es = bf.precision == BFGPU_C128 ? 16 : 8;
    DevBuf dx, dy;
    dx.alloc(size_t(bf.cols) * nrhs * es);
    dy.alloc(size_t(bf.rows) * nrhs * es);
    BFG_HIP(hipMemsetAsync(dx.p, 0, dx.bytes, s));
    BFG_HIP(hipMemsetAsync(dy.p, 0, dy.bytes, s));
    size_t need = size_t(bf.slot_len) * size_t(nrhs) * es;
    if (need > bf.slots_cap) {
        release_graph(bf);
        bf.slots.alloc(need);
        bf.slots_cap = need;
    }
    BFG_HIP(hipMemsetAsync(bf.slots.p, 0, need, s));
    release_graph(bf);
    const int32_t nL = bf.n_stage_launch;
    const int32_t *pt = bf.perm_test.as<const int32_t>(), *pr = bf.perm_trial.as<const int32_t>();
    static const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    hipEvent_t e0, e1;
    BFG_HIP(hipEventCreate(&e0));
    BFG_HIP(hipEventCreate(&e1));
    try {
        for (int32_t L = 0; L < nL; ++L) {
            StageTable &t = bf.fwd[L];
            const int32_t *pin = L == 0 ? pr : nullptr, *pout = L == nL - 1 ? pt : nullptr;
            double best = 1e30;
            int32_t bestv = t.variant;
            for (int32_t v = 0; v <= 1; ++v) {
                t.variant = v;
                launch_stage(s, bf.precision, BFGPU_OP_N, t, bf.factors[L].p, dx.p, bf.cols, dy.p, bf.rows,
                             bf.slots.p, bf.slot_len, pin, pout, nrhs, one, zero, true);
                BFG_HIP(hipEventRecord(e0, s));
                for (int32_t k = 0; k < iters; ++k)
                    launch_stage(s, bf.precision, BFGPU_OP_N, t, bf.factors[L].p, dx.p, bf.cols, dy.p,
                                 bf.rows, bf.slots.p, bf.slot_len, pin, pout, nrhs, one, zero, true);
                BFG_HIP(hipEventRecord(e1, s));
                BFG_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                BFG_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms / iters < best) { best = ms / iters; bestv = v; }
            }
            t.variant = bestv;
            if (stage_ms) stage_ms[L] = best;
            if (variants) variants[L] = bestv;
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

}  // namespace bfg
