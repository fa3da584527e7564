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
#include <array>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>

#include <sys/resource.h>

#include "internal.hpp"

namespace bfg {

namespace {
inline uint64_t key2(int32_t a, int32_t b) { return (uint64_t(uint32_t(a)) << 32) | uint32_t(b); }

// Open-addressing map uint64 -> int32 for the small key sets of one
// factorization; reused across calls (generation stamps, no clearing).
struct KeyMap {
    std::vector<uint64_t> k;
    std::vector<int32_t> v;
    std::vector<uint32_t> g;
    uint32_t gen = 0;
    size_t mask = 0;
    void reset(size_t n) {
        size_t cap = 16;
        while (cap < 2 * n + 2) cap <<= 1;
        if (k.size() < cap) {
            k.assign(cap, 0);
            v.assign(cap, 0);
            g.assign(cap, 0);
            gen = 0;
        }
        if (++gen == 0) {
            std::fill(g.begin(), g.end(), 0u);
            gen = 1;
        }
        mask = cap - 1;
    }
    size_t slot(uint64_t key) const {
        uint64_t x = key * 0x9E3779B97F4A7C15ULL;
        size_t h = size_t(x ^ (x >> 29)) & mask;
        while (g[h] == gen && k[h] != key) h = (h + 1) & mask;
        return h;
    }
    int32_t *get(uint64_t key) {
        size_t h = slot(key);
        return g[h] == gen ? &v[h] : nullptr;
    }
    // insert `val` unless present; returns the stored value and whether
    // this call inserted it
    std::pair<int32_t *, bool> emplace(uint64_t key, int32_t val) {
        size_t h = slot(key);
        if (g[h] == gen) return {&v[h], false};
        g[h] = gen;
        k[h] = key;
        v[h] = val;
        return {&v[h], true};
    }
};

// Map from a factorization's local slot offsets [0, n) to ordinals: a
// directly indexed array with generation stamps (no hashing, no clearing).
struct SlotMap {
    std::vector<uint32_t> g;
    std::vector<int32_t> v;
    uint32_t gen = 0;
    void reset(size_t n) {
        if (g.size() < n) {
            g.assign(n, 0u);
            v.assign(n, 0);
            gen = 0;
        }
        if (++gen == 0) {
            std::fill(g.begin(), g.end(), 0u);
            gen = 1;
        }
    }
    // the ordinal stored for `key`, inserting `val` when absent
    std::pair<int32_t, bool> emplace(uint32_t key, int32_t val) {
        if (g[key] == gen) return {v[key], false};
        g[key] = gen;
        v[key] = val;
        return {val, true};
    }
};

template <class T, class A> void upload_tab(DevBuf &d, const std::vector<T, A> &h) {
    d.alloc(sizeof(T) * std::max<size_t>(h.size(), 1));
    if (!h.empty()) BFG_HIP(hipMemcpy(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
}

// merge adjacent contributions of one item that read consecutive inputs
// with consecutive column slices of one column-major block; returns the
// merged count (the merged list is a prefix of c[0, n))
int32_t merge_contribs(MvContrib *c, int32_t n) {
    int32_t w = 0;
    for (int32_t i = 0; i < n; ++i) {
        const MvContrib x = c[i];
        if (w > 0) {
            MvContrib &p = c[w - 1];
            if (p.fofs >= 0 && x.fofs >= 0 && p.m == x.m && p.fofs + int64_t(p.m) * p.n == x.fofs &&
                p.in_ofs + p.n == x.in_ofs) {
                p.n += x.n;
                continue;
            }
        }
        c[w++] = x;
    }
    return w;
}

// Flat numbering of every block of every factorization: block i of stage s
// of factorization b is block first[sf[b] + s] + i.
struct BlockIndex {
    std::vector<int64_t> sf;     // per factorization: offset into first
    std::vector<int64_t> first;  // per (factorization, stage): first flat index; + one end entry
    int64_t total = 0;
};

BlockIndex index_blocks(const BF &bf) {
    BlockIndex ix;
    const int64_t nb = int64_t(bf.bfs.size());
    ix.sf.assign(nb + 1, 0);
    for (int64_t b = 0; b < nb; ++b) ix.sf[b + 1] = ix.sf[b] + int64_t(bf.bfs[b].st.size()) + 1;
    ix.first.assign(size_t(ix.sf[nb]), 0);
    int64_t g = 0;
    for (int64_t b = 0; b < nb; ++b) {
        const OneBF &ob = bf.bfs[b];
        int64_t *f = &ix.first[ix.sf[b]];
        for (size_t s = 0; s < ob.st.size(); ++s) {
            f[s] = g;
            g += int64_t(ob.st[s].b.size());
        }
        f[ob.st.size()] = g;
    }
    ix.total = g;
    return ix;
}
}  // namespace

// Every range a launch table names lies inside the buffer it indexes (slot
// buffer, permuted x / y, the launch's factor data); checked on the host so
// that an inconsistent block set is an error, not a device fault.
static void validate_tables(const BF &bf) {
    const int32_t nL = bf.n_stage_launch, Lmax = nL - 1;
    const int64_t nx = bf.trial->n_values, ny = bf.test->n_values;
    for (int32_t L = 0; L < nL; ++L)
        for (int dir = 0; dir < 4; ++dir) {
            const bool fwd = (dir & 1) == 0;
            const StageTable &M = fwd ? bf.fwd[L] : bf.tr[L];
            if (dir >= 2 && !M.pre) continue;
            const StageTable &S = dir >= 2 ? *M.pre : M;
            const int64_t out_lim = S.out_kind == OUT_SLOT ? bf.slot_len : (fwd ? ny : nx);
            const int64_t in_lim = S.in_kind == IN_SLOT ? bf.slot_len : (fwd ? nx : ny);
            const int64_t fe = L < int32_t(bf.factor_elems.size()) ? bf.factor_elems[L] : 0;
            BFG_REQUIRE(fe == 0 || (L < int32_t(bf.factors.size()) &&
                                    bf.factors[L].bytes >= size_t(fe + FACTOR_PAD) *
                                                               (bf.precision == BFGPU_C128 ? 16 : 8)),
                        "pack: factor buffer without its tail pad");
            std::atomic<int64_t> bad{INT64_MAX};  // first item out of bounds
            parallel_for(int64_t(S.h_items.size()), [&](int64_t i) {
                const MvItem &it = S.h_items[i];
                bool ok = it.out_ofs >= 0 && it.out_len >= 0 && int64_t(it.out_ofs) + it.out_len <= out_lim &&
                          it.c_begin >= 0 && it.c_begin < it.c_end && it.c_end <= S.n_contrib;
                for (int32_t c = ok ? it.c_begin : 0; ok && c < it.c_end; ++c) {
                    const MvContrib &k = S.h_contribs[c];
                    const int32_t rows = fwd ? k.m : k.n, cols = fwd ? k.n : k.m;
                    ok = k.m >= 0 && k.n >= 0 && rows == it.out_len && k.in_ofs >= 0 &&
                         int64_t(k.in_ofs) + cols <= in_lim &&
                         (k.fofs < 0 ? k.m == k.n : k.fofs + int64_t(k.m) * k.n <= fe);
                }
                if (!ok) {
                    int64_t cur = bad.load();
                    while (i < cur && !bad.compare_exchange_weak(cur, i)) {
                    }
                }
            }, 4096);
            if (bad.load() != INT64_MAX) {
                const MvItem &it = S.h_items[size_t(bad.load())];
                std::string d = "item " + std::to_string(bad.load()) + " out " + std::to_string(it.out_ofs) + "+" +
                                std::to_string(it.out_len) + " of " + std::to_string(out_lim) + ", contributions [" +
                                std::to_string(it.c_begin) + ", " + std::to_string(it.c_end) + ") of " +
                                std::to_string(S.n_contrib);
                for (int32_t c = std::max(it.c_begin, 0); c < it.c_end && c < S.n_contrib && c < it.c_begin + 4; ++c) {
                    const MvContrib &k = S.h_contribs[c];
                    d += "; block " + std::to_string(k.m) + "x" + std::to_string(k.n) + " at " + std::to_string(k.fofs) +
                         " of " + std::to_string(fe) + ", input " + std::to_string(k.in_ofs) + " of " +
                         std::to_string(in_lim);
                }
                BFG_REQUIRE(false, "pack: launch table " + std::to_string(L) + (fwd ? " (forward" : " (transposed") +
                                       (dir >= 2 ? ", partial sums)" : ")") + " indexes outside its buffers: " + d);
            }
            if (S.in_kind == IN_GATHER) {
                // the permuted copy of the inputs covers [in_lo, in_hi)
                const std::vector<int32_t> &perm = fwd ? bf.trial->perm : bf.test->perm;
                const int64_t lim = fwd ? bf.cols : bf.rows;
                BFG_REQUIRE(S.in_lo >= 0 && S.in_lo <= S.in_hi && S.in_hi <= int64_t(perm.size()),
                            "pack: gather range outside the tree");
                for (int32_t v = S.in_lo; v < S.in_hi; ++v)
                    BFG_REQUIRE(perm[v] >= 0 && perm[v] < lim, "pack: gather range reads outside x");
            }
        }
    (void)Lmax;
}

// Slot layout: every skeleton vector of every factorization in one buffer;
// per factorization and level in the order its consumers first read them
// (the column slices one compression reads are then adjacent, so their
// contributions merge into one), then vectors written but never read.
// Launch tables: forward items are output vectors (a slot, or an observer
// leaf's rows of y shared by all factorizations), transposed items input
// vectors; each item lists its contributions.
void pack_tables(BF &bf) {
    const Tree &test = *bf.test, &trial = *bf.trial;
    const int32_t nL = bf.n_stage_launch, Lmax = nL - 1;
    const int es = bf.precision == BFGPU_C128 ? 16 : 8;
    const int64_t nb = int64_t(bf.bfs.size());
    using clk = std::chrono::steady_clock;
    const bool timing = std::getenv("BFGPU_TIMING") != nullptr;
    const bool timing_detail = timing && std::atoi(std::getenv("BFGPU_TIMING")) >= 2;
    double t_group = 0, t_split = 0, t_finish = 0, t_up = 0, t_g_items = 0, t_g_fe = 0, t_g_merge = 0;
    auto tp0 = clk::now();
    const hipStream_t ps = bf.ctx ? bf.ctx->stream : hipStream_t(nullptr);
    auto minflt = [] {
        rusage u{};
        getrusage(RUSAGE_SELF, &u);
        return int64_t(u.ru_minflt);
    };
    const int64_t flt0 = minflt();
    auto since = [](clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); };
    const BlockIndex ix = index_blocks(bf);
    double t_keys = 0, t_slots = 0, t_pf = 0;
    // block keys and roots index the trees' node arrays below
    // The blocks of one factorization and stage sit in their own heap array
    // behind two pointers: loops over factorizations request every stage's
    // blocks of factorization b + 4 (its stage array of b + 8) ahead of use.
    auto prefetch_all = [&](int64_t b) {
        if (b + 8 < nb) __builtin_prefetch(bf.bfs[size_t(b + 8)].st.data());
        if (b + 4 < nb)
            for (const StageBlocks &st : bf.bfs[size_t(b + 4)].st) __builtin_prefetch(st.b.data());
    };
    parallel_for(nb, [&](int64_t b) {
        prefetch_all(b);
        const OneBF &ob = bf.bfs[b];
        BFG_REQUIRE(ob.no >= 0 && ob.no < test.n_nodes && ob.ns >= 0 && ob.ns < trial.n_nodes,
                    "pack: root node outside the trees");
        for (const StageBlocks &st : ob.st)
            for (const Block &k : st.b)
                BFG_REQUIRE(k.obs_out >= 0 && k.obs_out < test.n_nodes && k.obs_in >= 0 && k.obs_in < test.n_nodes &&
                                k.src_out >= 0 && k.src_out < trial.n_nodes && k.src_in >= 0 &&
                                k.src_in < trial.n_nodes,
                            "pack: block key outside the trees");
    }, 256);
    t_keys = since(tp0);
    auto tsl = clk::now();
    // ---- slot offsets (per factorization, local), reference slot ids ----
    hvec<int32_t> sl_out(static_cast<size_t>(ix.total)), sl_in(static_cast<size_t>(ix.total));
    parallel_for(ix.total, [&](int64_t g) {
        sl_out[size_t(g)] = -1;
        sl_in[size_t(g)] = -1;
    }, 1 << 16);
    std::vector<int64_t> slot_base(nb + 1, 0);
    parallel_for(nb, [&](int64_t b) {
        prefetch_all(b);
        thread_local KeyMap len, sm, ids;
        OneBF &ob = bf.bfs[b];
        const int32_t ns = int32_t(ob.st.size());
        const int64_t *f = &ix.first[ix.sf[b]];
        int64_t cnt = 0;
        for (int32_t s = 0; s + 1 < ns; ++s) {
            std::vector<Block> &P = ob.st[s].b, &Cn = ob.st[s + 1].b;
            len.reset(P.size());
            for (const Block &k : P) *len.emplace(key2(k.obs_out, k.src_out), k.m).first = k.m;
            sm.reset(P.size() + Cn.size());
            for (size_t i = 0; i < Cn.size(); ++i) {
                const uint64_t kk = key2(Cn[i].obs_in, Cn[i].src_in);
                auto r = sm.emplace(kk, int32_t(cnt));
                if (r.second) {
                    const int32_t *l = len.get(kk);
                    BFG_REQUIRE(l, "pack: block reads a skeleton vector no block writes");
                    cnt += *l;
                }
                sl_in[size_t(f[s + 1]) + i] = *r.first;
            }
            for (size_t i = 0; i < P.size(); ++i) {
                auto r = sm.emplace(key2(P[i].obs_out, P[i].src_out), int32_t(cnt));
                if (r.second) cnt += P[i].m;
                sl_out[size_t(f[s]) + i] = *r.first;
            }
            BFG_REQUIRE(cnt < (int64_t(1) << 31), "pack: skeleton vectors exceed 2^31 entries");
            // reference slot ids of level s: first use in stage order, the
            // writers of level s before its readers (bfstructs.jl:107-145)
            ids.reset(P.size() + Cn.size());
            int32_t nid = 0;
            for (Block &k : P) {
                auto r = ids.emplace(key2(k.obs_out, k.src_out), nid);
                nid += r.second;
                k.out_ptr = *r.first;
            }
            for (Block &k : Cn) {
                auto r = ids.emplace(key2(k.obs_in, k.src_in), nid);
                nid += r.second;
                k.in_ptr = *r.first;
            }
        }
        for (Block &k : ob.st[0].b) k.in_ptr = -1;
        for (Block &k : ob.st[ns - 1].b) k.out_ptr = -1;
        slot_base[b + 1] = cnt;
    });
    for (int64_t b = 0; b < nb; ++b) slot_base[b + 1] += slot_base[b];
    const int64_t counter = slot_base[nb];
    {
        std::vector<int64_t> nd(nb, 0);
        parallel_for(nb, [&](int64_t b) {
            for (const auto &st : bf.bfs[b].st)
                for (const Block &k : st.b)
                    if (!k.ident) nd[b] += int64_t(k.m) * k.n;
        });
        bf.n_blocks = ix.total;
        bf.n_data = 0;
        for (int64_t v : nd) bf.n_data += v;
    }
    BFG_REQUIRE(counter < (int64_t(1) << 31), "pack: skeleton vectors exceed 2^31 entries");
    bf.part_ofs = counter;
    t_slots = since(tsl);

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

    // stage of factorization b launched at L (Q = 0, R^l = l, P = Lmax)
    auto stage_at = [&](int64_t b, int32_t L) -> int32_t {
        const int32_t ns = int32_t(bf.bfs[b].st.size());
        if (L == Lmax) return ns - 1;
        return L < ns - 1 ? L : -1;
    };
    // the blocks of one stage only
    auto prefetch_blocks = [&](int64_t b, int32_t L) {
        if (b + 8 < nb) __builtin_prefetch(bf.bfs[size_t(b + 8)].st.data());
        if (b + 4 < nb) {
            const int32_t s = stage_at(b + 4, L);
            if (s >= 0) __builtin_prefetch(bf.bfs[size_t(b + 4)].st[size_t(s)].b.data());
        }
    };
    hvec<int32_t> litem(size_t(std::max<int64_t>(ix.total, 1)));  // written before it is read
    // Group the blocks of launch L into items.  fwd: keyed by output,
    // contributions read the input; transposed: the other way round.
    auto group = [&](int32_t L, bool fwd, StageTable &S) {
        const bool node_keyed = fwd ? (L == Lmax) : (L == 0);
        const auto tg_start = clk::now();
        auto tg0 = tg_start;
        hvec<MvItem> items;
        hvec<MvContrib> con;
        auto describe = [&](int64_t b, int32_t s, const Block &k, int64_t g, int64_t &out_ofs,
                            int64_t &in_ofs) {
            const bool isQ = s == 0, isP = s == int32_t(bf.bfs[b].st.size()) - 1;
            const int64_t o = isP ? int64_t(test.val_begin[k.obs_out]) : slot_base[b] + sl_out[g];
            const int64_t i = isQ ? int64_t(trial.val_begin[k.src_in]) : slot_base[b] + sl_in[g];
            out_ofs = fwd ? o : i;
            in_ofs = fwd ? i : o;
        };
        if (node_keyed) {
            // items: the leaves of the shared side, in node order; each
            // lists its blocks in factorization order.  Two passes over
            // ranges of factorizations: count per (range, node), then each
            // range writes from its own start position in every item.
            const Tree &t = fwd ? test : trial;
            const int64_t NV = t.n_nodes;
            const int64_t nr = std::min<int64_t>(nb, 4 * int64_t(host_threads()));
            auto rbeg = [&](int64_t r) { return nb * r / nr; };
            std::vector<int64_t> pos(size_t(nr * NV), 0);
            parallel_for(nr, [&](int64_t r) {
                int64_t *c = &pos[size_t(r * NV)];
                for (int64_t b = rbeg(r); b < rbeg(r + 1); ++b) {
                    prefetch_blocks(b, L);
                    const int32_t s = stage_at(b, L);
                    if (s < 0) continue;
                    for (const Block &k : bf.bfs[b].st[s].b) ++c[fwd ? k.obs_out : k.src_in];
                }
            }, 1);
            std::vector<int32_t> item_of(size_t(NV), -1);
            int64_t total = 0;
            for (int64_t v = 0; v < NV; ++v) {
                const int64_t start = total;
                for (int64_t r = 0; r < nr; ++r) {
                    const int64_t n = pos[size_t(r * NV + v)];
                    pos[size_t(r * NV + v)] = total;
                    total += n;
                }
                if (total > start) {
                    BFG_REQUIRE(total < (int64_t(1) << 31), "pack: too many blocks in one launch");
                    item_of[v] = int32_t(items.size());
                    items.push_back(MvItem{0, 0, int32_t(start), int32_t(total)});
                }
            }
            con.resize(size_t(total));
            parallel_for(nr, [&](int64_t r) {
                int64_t *c = &pos[size_t(r * NV)];
                for (int64_t b = rbeg(r); b < rbeg(r + 1); ++b) {
                    prefetch_blocks(b, L);
                    const int32_t s = stage_at(b, L);
                    if (s < 0) continue;
                    const auto &blk = bf.bfs[b].st[s].b;
                    const int64_t g0 = ix.first[ix.sf[b] + s];
                    for (size_t q = 0; q < blk.size(); ++q) {
                        const Block &k = blk[q];
                        const int32_t v = fwd ? k.obs_out : k.src_in;
                        int64_t oo, io;
                        describe(b, s, k, g0 + int64_t(q), oo, io);
                        const int64_t p = c[v]++;
                        MvItem &it = items[size_t(item_of[size_t(v)])];
                        if (p == it.c_begin) {  // the item's first block sets its output
                            it.out_ofs = int32_t(oo);
                            it.out_len = fwd ? k.m : k.n;
                        }
                        con[size_t(p)] = MvContrib{k.ident ? -1 : k.fofs, k.m, k.n, int32_t(io), 0};
                    }
                }
            }, 1);
        } else {
            // items: vectors of one factorization, first-use order within it
            std::vector<int64_t> nit(nb + 1, 0), nct(nb + 1, 0);
            parallel_for(nb, [&](int64_t b) {
                prefetch_blocks(b, L);
                const int32_t s = stage_at(b, L);
                if (s < 0) return;
                thread_local SlotMap km;
                const auto &blk = bf.bfs[b].st[s].b;
                const int64_t g0 = ix.first[ix.sf[b] + s];
                km.reset(size_t(slot_base[b + 1] - slot_base[b]));
                int32_t n = 0, kept = 0;
                for (size_t q = 0; q < blk.size(); ++q) {
                    const int64_t g = g0 + int64_t(q);
                    // a block that writes an empty vector (rank 0) writes
                    // nothing; its vector shares its offset with the next
                    // one, so it must not key an item
                    if ((fwd ? blk[q].m : blk[q].n) == 0) {
                        litem[g] = -1;
                        continue;
                    }
                    auto r = km.emplace(uint32_t(fwd ? sl_out[g] : sl_in[g]), n);
                    n += r.second;
                    litem[g] = r.first;
                    ++kept;
                }
                nit[b + 1] = n;
                nct[b + 1] = kept;
            });
            for (int64_t b = 0; b < nb; ++b) {
                nit[b + 1] += nit[b];
                nct[b + 1] += nct[b];
            }
            BFG_REQUIRE(nct[nb] < (int64_t(1) << 31), "pack: too many blocks in one launch");
            items.resize(size_t(nit[nb]));
            con.resize(size_t(nct[nb]));
            parallel_for(nb, [&](int64_t b) {
                prefetch_blocks(b, L);
                const int32_t s = stage_at(b, L);
                if (s < 0) return;
                thread_local std::vector<int32_t> cur;
                const auto &blk = bf.bfs[b].st[s].b;
                const int64_t g0 = ix.first[ix.sf[b] + s];
                const int32_t n = int32_t(nit[b + 1] - nit[b]);
                cur.assign(size_t(n) + 1, 0);
                for (size_t q = 0; q < blk.size(); ++q)
                    if (litem[g0 + int64_t(q)] >= 0) ++cur[size_t(litem[g0 + int64_t(q)]) + 1];
                for (int32_t j = 0; j < n; ++j) cur[j + 1] += cur[j];
                for (int32_t j = 0; j < n; ++j) {
                    MvItem &it = items[size_t(nit[b] + j)];
                    it.c_begin = it.c_end = int32_t(nct[b] + cur[j]);
                }
                for (size_t q = 0; q < blk.size(); ++q) {
                    const Block &k = blk[q];
                    if (litem[g0 + int64_t(q)] < 0) continue;
                    int64_t oo, io;
                    describe(b, s, k, g0 + int64_t(q), oo, io);
                    MvItem &it = items[size_t(nit[b] + litem[g0 + int64_t(q)])];
                    if (it.c_end == it.c_begin) {
                        it.out_ofs = int32_t(oo);
                        it.out_len = fwd ? k.m : k.n;
                    }
                    con[size_t(it.c_end++)] = MvContrib{k.ident ? -1 : k.fofs, k.m, k.n, int32_t(io), 0};
                }
            });
        }
        t_g_items += since(tg0);
        tg0 = clk::now();
        {  // factor entries read, summed in parallel over chunks
            constexpr int64_t CH = int64_t(1) << 16;
            const int64_t nk = int64_t(con.size()), nch = (nk + CH - 1) / CH;
            std::vector<int64_t> part(size_t(std::max<int64_t>(nch, 1)), 0);
            parallel_for(nch, [&](int64_t q) {
                int64_t f = 0;
                for (int64_t i = q * CH; i < std::min(nk, (q + 1) * CH); ++i)
                    if (con[size_t(i)].fofs >= 0) f += int64_t(con[size_t(i)].m) * con[size_t(i)].n;
                part[size_t(q)] = f;
            }, 1);
            int64_t fe = 0;
            for (int64_t f : part) fe += f;
            S.factor_elems = fe;
        }
        t_g_fe += since(tg0);
        tg0 = clk::now();
        if (fwd) {
            // merge within each item, then compact
            std::vector<int32_t> kept(items.size() + 1, 0);
            parallel_for(int64_t(items.size()), [&](int64_t i) {
                MvItem &it = items[i];
                kept[i + 1] = merge_contribs(con.data() + it.c_begin, it.c_end - it.c_begin);
            }, 4096);
            for (size_t i = 0; i < items.size(); ++i) kept[i + 1] += kept[i];
            hvec<MvContrib> out(static_cast<size_t>(kept[items.size()]));
            parallel_for(int64_t(items.size()), [&](int64_t i) {
                MvItem &it = items[i];
                std::copy(con.begin() + it.c_begin, con.begin() + it.c_begin + (kept[i + 1] - kept[i]),
                          out.begin() + kept[i]);
                it.c_begin = kept[i];
                it.c_end = kept[i + 1];
            }, 4096);
            con.swap(out);
        }
        S.h_items.swap(items);
        S.h_contribs.swap(con);
        t_g_merge += since(tg0);
        if (timing_detail)
            std::fprintf(stderr, "[bfgpu] group L=%d %s%s: %zu items, %zu contributions, %.3fs\n", L,
                         fwd ? "forward" : "transposed", node_keyed ? " (leaf-keyed)" : "", S.h_items.size(),
                         S.h_contribs.size(), since(tg_start));
    };
    // Split items whose contribution lists are long (observer-leaf sums over
    // many factorizations) into chunks of about CHUNK factor entries: a
    // `pre` table writes one partial sum per chunk into slots after the
    // skeleton vectors, the item then adds its partials (identity
    // contributions).  Returns the partial-sum entries used.
    auto split = [&](StageTable &S) -> int64_t {
        constexpr int64_t CHUNK = 4096;
        auto elems = [](const MvContrib &c) { return c.fofs < 0 ? int64_t(c.m) : int64_t(c.m) * c.n; };
        const int64_t ni = int64_t(S.h_items.size());
        // chunk boundaries of item i: contributions [start, c] close a chunk
        // when its entries reach CHUNK or the item ends
        auto is_split = [&](const MvItem &it) {
            if (it.c_end - it.c_begin <= 2) return false;
            int64_t total = 0;
            for (int32_t c = it.c_begin; c < it.c_end; ++c) total += elems(S.h_contribs[c]);
            return total > 2 * CHUNK;
        };
        auto n_chunks = [&](const MvItem &it) {
            int64_t n = 0, acc = 0;
            for (int32_t c = it.c_begin; c < it.c_end; ++c) {
                acc += elems(S.h_contribs[c]);
                if (acc < CHUNK && c + 1 < it.c_end) continue;
                ++n;
                acc = 0;
            }
            return n;
        };
        // first pass (parallel): per item, its chunks (0: kept whole)
        std::vector<int64_t> nchk(size_t(ni) + 1, 0);
        parallel_for(ni, [&](int64_t i) {
            const MvItem &it = S.h_items[size_t(i)];
            nchk[size_t(i) + 1] = is_split(it) ? n_chunks(it) : 0;
        }, 1024);
        // running offsets: main contributions, partial items, partial contributions, partial entries
        std::vector<int64_t> omc(size_t(ni) + 1, 0), opi(size_t(ni) + 1, 0), opc(size_t(ni) + 1, 0),
            opo(size_t(ni) + 1, 0);
        for (int64_t i = 0; i < ni; ++i) {
            const MvItem &it = S.h_items[size_t(i)];
            const int64_t k = nchk[size_t(i) + 1], nc = it.c_end - it.c_begin;
            omc[size_t(i) + 1] = omc[size_t(i)] + (k ? k : nc);
            opi[size_t(i) + 1] = opi[size_t(i)] + k;
            opc[size_t(i) + 1] = opc[size_t(i)] + (k ? nc : 0);
            opo[size_t(i) + 1] = opo[size_t(i)] + k * it.out_len;
        }
        if (opi[size_t(ni)] == 0) return 0;
        BFG_REQUIRE(bf.part_ofs + opo[size_t(ni)] < (int64_t(1) << 31) && omc[size_t(ni)] < (int64_t(1) << 31) &&
                        opc[size_t(ni)] < (int64_t(1) << 31),
                    "pack: partial sums exceed 2^31 entries");
        hvec<MvItem> mi(static_cast<size_t>(ni)), pi(static_cast<size_t>(opi[size_t(ni)]));
        hvec<MvContrib> mc(static_cast<size_t>(omc[size_t(ni)])), pc(static_cast<size_t>(opc[size_t(ni)]));
        // second pass (parallel): write every item at its offsets
        parallel_for(ni, [&](int64_t i) {
            const MvItem &it = S.h_items[size_t(i)];
            MvItem m = it;
            m.c_begin = int32_t(omc[size_t(i)]);
            m.c_end = int32_t(omc[size_t(i) + 1]);
            mi[size_t(i)] = m;
            if (nchk[size_t(i) + 1] == 0) {
                std::copy(S.h_contribs.begin() + it.c_begin, S.h_contribs.begin() + it.c_end,
                          mc.begin() + omc[size_t(i)]);
                return;
            }
            int64_t q = opi[size_t(i)], wm = omc[size_t(i)], wp = opc[size_t(i)], po = bf.part_ofs + opo[size_t(i)];
            int32_t start = it.c_begin;
            int64_t acc = 0;
            for (int32_t c = it.c_begin; c < it.c_end; ++c) {
                acc += elems(S.h_contribs[c]);
                if (acc < CHUNK && c + 1 < it.c_end) continue;
                const int64_t n = c + 1 - start;
                pi[size_t(q)] = MvItem{int32_t(po), it.out_len, int32_t(wp), int32_t(wp + n)};
                std::copy(S.h_contribs.begin() + start, S.h_contribs.begin() + c + 1, pc.begin() + wp);
                mc[size_t(wm)] = MvContrib{-1, it.out_len, it.out_len, int32_t(po), 0};
                ++q;
                ++wm;
                wp += n;
                po += it.out_len;
                start = c + 1;
                acc = 0;
            }
        }, 1024);
        S.pre = std::make_unique<StageTable>();
        S.pre->in_kind = S.in_kind;
        S.pre->out_kind = OUT_SLOT;
        S.pre->h_items.swap(pi);
        S.pre->h_contribs.swap(pc);
        S.h_items.swap(mi);
        S.h_contribs.swap(mc);
        return opo[size_t(ni)];
    };
    // p-quantile of a histogram over power-of-two buckets: the smallest
    // 2^b with at least q of the counted values <= 2^b
    auto pow2_quantile = [](const std::vector<int64_t> &h, double q) {
        int64_t tot = 0;
        for (int64_t c : h) tot += c;
        int64_t run = 0;
        for (size_t b = 0; b < h.size(); ++b) {
            run += h[b];
            if (run >= q * double(tot)) return int(b);
        }
        return int(h.size()) - 1;
    };
    auto log2up = [](int64_t v) {
        int b = 0;
        while ((int64_t(1) << b) < v) ++b;
        return b;
    };
    std::function<void(StageTable &, bool)> finish = [&](StageTable &S, bool fwd) {
        // statistics of the table, in parallel over chunks of items and of
        // contributions, partials merged in chunk order
        struct Part {
            int64_t vec = 0, fe = 0, nc_sum = 0;
            int32_t max_out = 0, max_in = 0, max_nc = 0, in_lo = INT32_MAX, in_hi = 0;
            int64_t hrow[32] = {}, hk[32] = {};
        };
        constexpr int64_t PCH = 1 << 16;
        const int64_t ni = int64_t(S.h_items.size()), nk = int64_t(S.h_contribs.size());
        const int64_t nci = (ni + PCH - 1) / PCH, nck = (nk + PCH - 1) / PCH;
        std::vector<Part> parts(size_t(nci + nck));
        parallel_for(nci + nck, [&](int64_t q) {
            Part &p = parts[size_t(q)];
            if (q < nci) {
                for (int64_t i = q * PCH; i < std::min(ni, (q + 1) * PCH); ++i) {
                    const MvItem &it = S.h_items[i];
                    p.max_out = std::max(p.max_out, it.out_len);
                    p.vec += int64_t(it.out_len) * ((S.out_kind == OUT_SCATTER) ? 2 : 1);
                    ++p.hrow[size_t(log2up(std::max(it.out_len, 1)))];
                    p.max_nc = std::max(p.max_nc, it.c_end - it.c_begin);
                    p.nc_sum += it.c_end - it.c_begin;
                }
            } else {
                const int64_t c0 = (q - nci) * PCH;
                for (int64_t i = c0; i < std::min(nk, c0 + PCH); ++i) {
                    const MvContrib &c = S.h_contribs[i];
                    const int32_t in_len = fwd ? c.n : c.m;
                    p.max_in = std::max(p.max_in, in_len);
                    p.vec += in_len;
                    if (c.fofs >= 0) {
                        p.fe += int64_t(c.m) * c.n;
                        ++p.hk[size_t(log2up(std::max(in_len, 1)))];
                    }
                    p.in_lo = std::min(p.in_lo, c.in_ofs);
                    p.in_hi = std::max(p.in_hi, c.in_ofs + in_len);
                }
            }
        }, 1);
        int64_t vec = 0;
        S.max_out = S.max_in = 0;
        S.factor_elems = 0;
        std::vector<int64_t> hrow(32, 0), hk(32, 0);
        int64_t nc_sum = 0;
        S.max_nc = 0;
        S.in_lo = INT32_MAX;
        S.in_hi = 0;
        for (const Part &p : parts) {
            vec += p.vec;
            S.factor_elems += p.fe;
            nc_sum += p.nc_sum;
            S.max_out = std::max(S.max_out, p.max_out);
            S.max_in = std::max(S.max_in, p.max_in);
            S.max_nc = std::max(S.max_nc, p.max_nc);
            S.in_lo = std::min(S.in_lo, p.in_lo);
            S.in_hi = std::max(S.in_hi, p.in_hi);
            for (int b = 0; b < 32; ++b) {
                hrow[b] += p.hrow[b];
                hk[b] += p.hk[b];
            }
        }
        if (S.in_hi <= S.in_lo) S.in_lo = S.in_hi = 0;
        // tiled-kernel shape: rows cover 90% of the outputs; the lane groups
        // split the contributions when items have several, else the inner
        // index (QW = 8 loads per lane and pass)
        const int rb = std::min(6, pow2_quantile(hrow, 0.9));
        const double avg_nc = S.h_items.empty() ? 1.0 : double(nc_sum) / double(S.h_items.size());
        S.tile.rp_shift = int8_t(rb);
        S.tile.csplit = avg_nc >= 1.5 ? 1 : 0;
        int gb = S.tile.csplit ? log2up(int64_t(std::min(avg_nc, 8.0) + 0.5))
                               : std::max(0, pow2_quantile(hk, 0.9) - 3);
        S.tile.g_shift = int8_t(std::max(2 - rb, std::min(gb, 6 - rb)));
        S.tile.wg_per_cu = 8;
        S.n_items = int32_t(S.h_items.size());
        S.n_contrib = int32_t(S.h_contribs.size());
        S.alg_bytes_1 = vec * es;
        S.h_tasks.resize(S.h_items.size());
        parallel_for(int64_t(S.h_items.size()), [&](int64_t i) {
            const MvItem &it = S.h_items[i];
            const MvContrib &c0 = S.h_contribs[it.c_begin];
            S.h_tasks[i] = MvTask{c0.fofs, c0.m, c0.n, c0.in_ofs, it.out_ofs, it.out_len,
                                  it.c_begin + 1, it.c_end, -1, 0, -1, 0};
        }, 4096);
        // default forward variant until the autotuner measures: a
        // wavefront per task for small blocks, a workgroup otherwise
        S.variant = fwd ? ((S.n_items > 0 && S.factor_elems / S.n_items < 4096) ? 1 : 0) : 0;
        auto tu = clk::now();
        upload_tab(S.items, S.h_items);
        upload_tab(S.contribs, S.h_contribs);
        // the tasks (as the host's h_tasks) are derived on the device
        S.tasks.alloc(sizeof(MvTask) * std::max<size_t>(S.h_items.size(), 1));
        S.tasks_c.alloc(sizeof(MvTaskC) * std::max<size_t>(S.h_items.size(), 1));
        launch_build_tasks(ps, S.items.as<MvItem>(), S.n_items, S.contribs.as<MvContrib>(), S.n_contrib,
                           S.tasks.as<MvTask>(), S.tasks_c.as<MvTaskC>());
        t_up += since(tu);
        if (S.pre) finish(*S.pre, fwd);
    };

    bf.fwd.clear();
    bf.tr.clear();
    bf.fwd.resize(nL);
    bf.tr.resize(nL);
    int64_t covered_rows = 0, covered_cols = 0, part_used = 0;
    for (int32_t L = 0; L < nL; ++L) {
        StageTable &F = bf.fwd[L], &T = bf.tr[L];
        F.in_kind = (L == 0) ? IN_GATHER : IN_SLOT;
        F.out_kind = (L == Lmax) ? OUT_SCATTER : OUT_SLOT;
        T.in_kind = (L == Lmax) ? IN_GATHER : IN_SLOT;
        T.out_kind = (L == 0) ? OUT_SCATTER : OUT_SLOT;
        auto ta = clk::now();
        group(L, true, F);
        group(L, false, T);
        t_group += since(ta);
        ta = clk::now();
        if (L == Lmax) part_used = std::max(part_used, split(F));
        if (L == 0) part_used = std::max(part_used, split(T));
        t_split += since(ta);
        ta = clk::now();
        finish(F, true);
        finish(T, false);
        t_finish += since(ta);
        if (L == Lmax)
            for (const MvItem &it : F.h_items) covered_rows += it.out_len;
        if (L == 0)
            for (const MvItem &it : T.h_items) covered_cols += it.out_len;
    }
    bf.y_covered = covered_rows == bf.rows;
    bf.x_covered = covered_cols == bf.cols;

    // ---- matrix-free last level: per observer leaf, its P columns ----
    auto tpf = clk::now();
    bf.pfree_ok = bf.pfree_t_ok = false;
    bf.pf_n_items = 0;
    bool pf = bf.precision == BFGPU_C64 && bf.pf_k > 0.0 && nb > 0;
    double t_pf_check = 0, t_pf_cols = 0, t_pf_up = 0;
    if (pf) {
        std::atomic<bool> ok{true};
        parallel_for(nb, [&](int64_t b) {
            const OneBF &ob = bf.bfs[b];
            int64_t nc = 0;
            for (const Block &k : ob.st.back().b)  // rows: whole observer leaves (their centres frame them)
                nc += k.ident || (k.m > 0 && !test.isleaf(k.obs_out)) ? -(int64_t(1) << 40) : int64_t(k.n);
            if (!(ob.st.size() >= 2 && nc == int64_t(ob.pcols.size())))  // P reads skeleton slots
                ok.store(false, std::memory_order_relaxed);
        }, 1024);
        pf = ok.load();
    }
    t_pf_check = since(tpf);
    if (pf) {
        const int64_t NV = test.n_nodes;
        const int64_t nr = std::min<int64_t>(nb, 4 * int64_t(host_threads()));
        auto rbeg = [&](int64_t r) { return nb * r / nr; };
        std::vector<int64_t> pos(size_t(nr * NV), 0);
        parallel_for(nr, [&](int64_t r) {
            int64_t *c = &pos[size_t(r * NV)];
            for (int64_t b = rbeg(r); b < rbeg(r + 1); ++b)
                for (const Block &k : bf.bfs[b].st.back().b) c[k.obs_out] += k.n;
        }, 1);
        std::vector<int4> items;
        int64_t total = 0;
        for (int64_t v = 0; v < NV; ++v) {
            const int64_t start = total;
            for (int64_t r = 0; r < nr; ++r) {
                const int64_t n = pos[size_t(r * NV + v)];
                pos[size_t(r * NV + v)] = total;
                total += n;
            }
            if (total > start) items.push_back(int4{test.val_begin[v], test.nvals(int32_t(v)), int32_t(start),
                                                     int32_t(total)});
        }
        if (total < (int64_t(1) << 31) && !items.empty()) {
            // columns: the skeleton point relative to the observer leaf's
            // centre and the phase of its distance (revolutions, from FP64);
            // rows: the point relative to its leaf's centre and its square norm
            const double krev = bf.pf_k / (2.0 * M_PI);
            hvec<float4> col(static_cast<size_t>(total));
            hvec<int32_t> cin(static_cast<size_t>(total));
            parallel_for(nr, [&](int64_t r) {
                int64_t *c = &pos[size_t(r * NV)];
                for (int64_t b = rbeg(r); b < rbeg(r + 1); ++b) {
                    const OneBF &ob = bf.bfs[b];
                    const int32_t s = int32_t(ob.st.size()) - 1;
                    const int64_t g0 = ix.first[ix.sf[b] + s];
                    int64_t po = 0;
                    for (size_t q = 0; q < ob.st[s].b.size(); ++q) {
                        const Block &k = ob.st[s].b[q];
                        {  // blocks of empty leaves too: the transposed product writes their slots
                            const int64_t in0 = slot_base[b] + sl_in[size_t(g0) + q];
                            const double *ctr = &test.center[3 * size_t(k.obs_out)];
                            int64_t &p = c[k.obs_out];
                            for (int32_t j = 0; j < k.n; ++j, ++p) {
                                const double *y = &trial.points[3 * size_t(ob.pcols[size_t(po + j)])];
                                const double d0 = y[0] - ctr[0], d1 = y[1] - ctr[1], d2 = y[2] - ctr[2];
                                const double ph = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2) * krev;
                                col[size_t(p)] = float4{float(d0), float(d1), float(d2), float(ph - std::floor(ph))};
                                cin[size_t(p)] = int32_t(in0 + j);
                            }
                        }
                        po += k.n;
                    }
                }
            }, 1);
            std::vector<float4> row(size_t(std::max(test.n_values, 1)), float4{0.f, 0.f, 0.f, 0.f});
            parallel_for(NV, [&](int64_t v) {
                if (!test.isleaf(int32_t(v))) return;
                const double *ctr = &test.center[3 * size_t(v)];
                for (int32_t i = test.val_begin[v]; i < test.val_end[v]; ++i) {
                    const double *x = &test.points[3 * size_t(test.perm[i])];
                    const double u0 = x[0] - ctr[0], u1 = x[1] - ctr[1], u2 = x[2] - ctr[2];
                    row[size_t(i)] = float4{float(u0), float(u1), float(u2), float(u0 * u0 + u1 * u1 + u2 * u2)};
                }
            }, 256);
            {  // the transposed kernel writes each column's slot: one writer
               // per slot.  Slots of different factorizations are disjoint
               // ([slot_base[b], slot_base[b + 1])), so it suffices that the
               // last-level blocks of each factorization read disjoint
               // ranges of its skeleton vectors.
                std::atomic<bool> once{true};
                parallel_for(nb, [&](int64_t b) {
                    thread_local std::vector<std::pair<int32_t, int32_t>> rg;
                    const OneBF &ob = bf.bfs[b];
                    const int32_t s = int32_t(ob.st.size()) - 1;
                    const int64_t g0 = ix.first[ix.sf[b] + s];
                    rg.clear();
                    for (size_t q = 0; q < ob.st[s].b.size(); ++q)
                        if (ob.st[s].b[q].n > 0) rg.emplace_back(sl_in[size_t(g0) + q], ob.st[s].b[q].n);
                    std::sort(rg.begin(), rg.end());
                    for (size_t q = 1; q < rg.size(); ++q)
                        if (rg[q].first < rg[q - 1].first + rg[q - 1].second) once.store(false, std::memory_order_relaxed);
                }, 256);
                bf.pfree_t_ok = once.load();
            }
            bf.pf_order_built = false;
            t_pf_cols = since(tpf) - t_pf_check;
            upload_tab(bf.pf_items, items);
            upload_tab(bf.pf_col, col);
            upload_tab(bf.pf_in, cin);
            upload_tab(bf.pf_row, row);
            t_pf_up = since(tpf) - t_pf_check - t_pf_cols;
            bf.pf_n_items = int32_t(items.size());
            bf.pfree_ok = true;
        }
    }
    t_pf = since(tpf);
    bf.gather_ofs = bf.part_ofs + part_used;
    bf.slot_len = bf.gather_ofs + std::max(test.n_values, trial.n_values);
    BFG_REQUIRE(bf.slot_len < (int64_t(1) << 31), "pack: slot buffer exceeds 2^31 entries");
    auto tv = clk::now();
    validate_tables(bf);
    const double t_valid = since(tv);
    upload_tab(bf.perm_test, test.perm);
    upload_tab(bf.perm_trial, trial.perm);
    BFG_HIP(hipStreamSynchronize(ps));  // device-built tasks complete
    bf.sweep = SweepTables{};
    if (bf.use_sweep) build_sweep(bf);
    if (timing)
        std::fprintf(stderr, "[bfgpu] pack %.3fs: index+keys %.3fs, slots %.3fs, group %.3fs (items %.3fs, entries %.3fs, "
                             "merge %.3fs), split %.3fs, "
                             "finish+upload %.3fs (upload %.3fs), matrix-free tables %.3fs (check %.3fs, columns %.3fs, "
                             "upload %.3fs), validate %.3fs; %lld page faults\n",
                     since(tp0), t_keys, t_slots, t_group, t_g_items, t_g_fe, t_g_merge, t_split, t_finish, t_up, t_pf, t_pf_check, t_pf_cols,
                     t_pf_up, t_valid, (long long)(minflt() - flt0));
}

// Streaming-kernel tables of a launch table (variant 9): the forward tables
// over the stored blocks, or (over_copy) the transposed tables over the
// transposed copy of the blocks, where an item's blocks lie back to back.
// A task streams when its output is a skeleton vector, it has no identity
// block, its blocks have its output's rows (at most 128) and lie back to
// back in the buffer (one block of m x sum n), and it reads at most 64
// input vectors.  Consecutive streaming tasks whose blocks follow each
// other in the buffer with gaps of at most 16 entries are cut into chunks of
// at most nld KiB of 16-byte words, 64 tasks, STREAM_HMAX row pairs, 64
// distinct input vectors and stream_xcap input entries (a task that reads
// the same vectors as the one before it shares their LDS copy).  Ranges of
// tasks are chunked in parallel (a chunk never spans two ranges); every
// other task goes to a dense-kernel list.
void build_stream(StageTable &S, int es, int nld, int64_t factor_elems, bool over_copy) {
    if (S.stream_built && S.stream_nld == nld && S.stream_copy == over_copy) return;
    BFG_REQUIRE(nld == 4 || nld == 8, "pack: stream chunks of 4 or 8 KiB");
    BFG_REQUIRE(es == 8 || es == 16, "pack: entry size 8 or 16 bytes");
    BFG_REQUIRE(!over_copy || S.h_cofs.size() == S.h_contribs.size(), "pack: transposed copy offsets missing");
    const int64_t E = 16 / es;                   // entries per word
    const int64_t cape = int64_t(nld) * 64 * E;  // entries per chunk
    const int32_t xcap = es == 8 ? stream_xcap<float>() : stream_xcap<double>();
    constexpr int64_t GAP = 16;
    const int64_t n = int64_t(S.h_items.size());
    const bool slot_out = S.out_kind == OUT_SLOT;
    // contribution k of the table as seen by the kernel: block offset, rows,
    // columns, input
    auto cf = [&](int32_t k) { return over_copy ? S.h_cofs[size_t(k)] : S.h_contribs[size_t(k)].fofs; };
    auto cm = [&](int32_t k) { return over_copy ? S.h_contribs[size_t(k)].n : S.h_contribs[size_t(k)].m; };
    auto cn = [&](int32_t k) { return over_copy ? S.h_contribs[size_t(k)].m : S.h_contribs[size_t(k)].n; };
    struct TV {
        bool ok;
        int64_t f0;
        int32_t m, ntot;
    };
    auto view = [&](int64_t i) {
        const MvItem &it = S.h_items[size_t(i)];
        TV v{false, -1, it.out_len, 0};
        if (!slot_out || it.c_end <= it.c_begin || it.c_end - it.c_begin > 64 || v.m < 1 || v.m > 2 * STREAM_HMAX)
            return v;
        v.f0 = cf(it.c_begin);
        int64_t f = v.f0;
        for (int32_t k = it.c_begin; k < it.c_end; ++k) {
            if (cf(k) < 0 || cf(k) != f || cm(k) != v.m || cn(k) < 1) return v;
            f += int64_t(v.m) * cn(k);
            v.ntot += cn(k);
        }
        v.ok = int64_t(v.m) * v.ntot + (E - 1) <= cape && v.ntot <= xcap;
        return v;
    };
    struct Part {
        hvec<StreamChunk> ch;
        hvec<StreamTask> tk;
        hvec<StreamRun> rn;
        hvec<MvTaskC> rest;
        int64_t gap = 0, xent = 0;
    };
    const int64_t nr = std::max<int64_t>(1, std::min<int64_t>(4 * int64_t(host_threads()), n / 4096 + 1));
    std::vector<Part> parts(static_cast<size_t>(nr));
    parallel_for(nr, [&](int64_t r) {
        Part &P = parts[size_t(r)];
        const int64_t lo = n * r / nr, hi = n * (r + 1) / nr;
        bool open = false;
        int64_t f0w = 0, fend = 0;
        int32_t t0 = 0, nt = 0, H = 0, X = 0, r0 = 0;
        uint64_t rmask = 0, xmask[4] = {0, 0, 0, 0};
        int64_t prev = -1;      // previous task of the chunk (its input vectors)
        int32_t prev_xo = 0;
        auto same_inputs = [&](int64_t a, int64_t b) {
            const MvItem &x = S.h_items[size_t(a)], &y = S.h_items[size_t(b)];
            if (x.c_end - x.c_begin != y.c_end - y.c_begin) return false;
            for (int32_t q = 0; q < x.c_end - x.c_begin; ++q)
                if (S.h_contribs[size_t(x.c_begin + q)].in_ofs != S.h_contribs[size_t(y.c_begin + q)].in_ofs ||
                    cn(x.c_begin + q) != cn(y.c_begin + q))
                    return false;
            return true;
        };
        auto close = [&]() {
            if (!open) return;
            const int64_t nw = (fend - f0w * E + E - 1) / E;
            StreamChunk k;
            k.w0 = uint64_t(f0w) | uint64_t(nw) << 40 | uint64_t(X) << 51;
            k.t0 = t0;  // part-relative until the parts are joined
            k.info = uint32_t(nt) | uint32_t(H) << 7 | uint32_t(64 / H) << 14;
            k.rmask = rmask;
            for (int q = 0; q < 4; ++q) k.xmask[q] = xmask[q];
            k.r0 = r0;  // part-relative until the parts are joined
            k.nruns = int32_t(P.rn.size()) - r0;
            P.ch.push_back(k);
            P.xent += X;
            open = false;
        };
        for (int64_t i = lo; i < hi; ++i) {
            const MvItem &it = S.h_items[size_t(i)];
            const TV v = view(i);
            if (!v.ok) {
                close();
                const MvTask &t = S.h_tasks[size_t(i)];
                if (over_copy)
                    P.rest.push_back(MvTaskC{it.c_begin < it.c_end ? cf(it.c_begin) : t.fofs, t.in_ofs, t.out_ofs,
                                             t.n, t.m, t.c_begin, t.c_end});
                else
                    P.rest.push_back(MvTaskC{t.fofs, t.in_ofs, t.out_ofs, t.m, t.n, t.c_begin, t.c_end});
                continue;
            }
            const int32_t h = (v.m + 1) / 2;
            const int32_t nb = it.c_end - it.c_begin;
            const int64_t end = v.f0 + int64_t(v.m) * v.ntot;
            const bool share = open && prev >= 0 && same_inputs(prev, i);
            const int32_t nruns = open ? int32_t(P.rn.size()) - r0 : 0;
            const bool fits = open && v.f0 >= fend && v.f0 - fend <= GAP && end - f0w * E <= cape && nt < 64 &&
                              H + h <= STREAM_HMAX &&
                              (share || (X + v.ntot <= xcap && nruns + nb <= 64));
            if (!fits) {
                close();
                open = true;
                f0w = v.f0 / E;
                fend = f0w * E;
                t0 = int32_t(P.tk.size());
                r0 = int32_t(P.rn.size());
                nt = H = X = 0;
                rmask = 0;
                xmask[0] = xmask[1] = xmask[2] = xmask[3] = 0;
                prev = -1;
            }
            P.gap += v.f0 - fend;
            int32_t xo;
            if (open && prev >= 0 && same_inputs(prev, i)) {
                xo = prev_xo;
            } else {  // the task's input vectors, listed at LDS entries X, X + n_0, ...
                xo = X;
                for (int32_t k = it.c_begin; k < it.c_end; ++k) {
                    xmask[X / 64] |= uint64_t(1) << (X % 64);
                    P.rn.push_back(StreamRun{S.h_contribs[size_t(k)].in_ofs, X});
                    X += cn(k);
                }
            }
            rmask |= uint64_t(1) << H;
            StreamTask st;
            st.lin = 0;
            st.out_ofs = it.out_ofs;
            st.ax = uint32_t(v.f0 - f0w * E) | uint32_t(xo) << 11;
            st.mhn = uint32_t(v.m) | uint32_t(H) << 8 | uint32_t(v.ntot) << 15;
            P.tk.push_back(st);
            ++nt;
            H += h;
            fend = end;
            prev = i;
            prev_xo = xo;
        }
        close();
    }, 1);
    int64_t n_ch = 0, n_tk = 0, n_rest = 0, n_rn = 0, gap = 0, xent = 0;
    for (const Part &P : parts) {
        n_ch += int64_t(P.ch.size());
        n_tk += int64_t(P.tk.size());
        n_rn += int64_t(P.rn.size());
        n_rest += int64_t(P.rest.size());
        gap += P.gap;
        xent += P.xent;
    }
    BFG_REQUIRE(n_tk < (int64_t(1) << 31) && n_ch < (int64_t(1) << 31), "pack: too many streaming tasks");
    hvec<StreamChunk> ch(static_cast<size_t>(n_ch));
    hvec<StreamTask> tk(static_cast<size_t>(n_tk));
    BFG_REQUIRE(n_rn < (int64_t(1) << 31), "pack: too many streaming input vectors");
    hvec<MvTaskC> rest(static_cast<size_t>(n_rest));
    hvec<StreamRun> rn(static_cast<size_t>(n_rn));
    {
        int64_t oc = 0, ot = 0, orr = 0, orn = 0;
        std::vector<int64_t> bc(static_cast<size_t>(nr)), bt(static_cast<size_t>(nr)), br(static_cast<size_t>(nr)),
            bn(static_cast<size_t>(nr));
        for (int64_t r = 0; r < nr; ++r) {
            bc[size_t(r)] = oc;
            bt[size_t(r)] = ot;
            br[size_t(r)] = orr;
            bn[size_t(r)] = orn;
            oc += int64_t(parts[size_t(r)].ch.size());
            ot += int64_t(parts[size_t(r)].tk.size());
            orr += int64_t(parts[size_t(r)].rest.size());
            orn += int64_t(parts[size_t(r)].rn.size());
        }
        std::atomic<bool> bad{false};
        parallel_for(nr, [&](int64_t r) {
            const Part &P = parts[size_t(r)];
            for (size_t q = 0; q < P.ch.size(); ++q) {
                StreamChunk k = P.ch[q];
                k.t0 += int32_t(bt[size_t(r)]);
                k.r0 += int32_t(bn[size_t(r)]);
                const int64_t f0 = int64_t(k.w0 & ((uint64_t(1) << 40) - 1)) * E;
                const int64_t nw = int64_t((k.w0 >> 40) & 2047);
                if (f0 < 0 || f0 + nw * E > factor_elems + FACTOR_PAD || nw > cape / E || k.nruns < 1 ||
                    k.nruns > 64)
                    bad.store(true);
                ch[size_t(bc[size_t(r)]) + q] = k;
            }
            std::copy(P.tk.begin(), P.tk.end(), tk.begin() + bt[size_t(r)]);
            std::copy(P.rn.begin(), P.rn.end(), rn.begin() + bn[size_t(r)]);
            std::copy(P.rest.begin(), P.rest.end(), rest.begin() + br[size_t(r)]);
        }, 1);
        BFG_REQUIRE(!bad.load(), "pack: streaming chunk outside its factor buffer");
    }
    upload_tab(S.s_chunks, ch);
    upload_tab(S.s_tasks, tk);
    upload_tab(S.s_runs, rn);
    upload_tab(S.s_rest, rest);
    S.stream_copy = over_copy;
    S.n_schunks = int32_t(n_ch);
    S.n_stasks = int32_t(n_tk);
    S.n_srest = int32_t(n_rest);
    S.stream_gap = gap;
    S.stream_xent = xent;
    S.stream_nld = nld;
    S.stream_built = true;
}

// Order of the matrix-free last level's items for the matrix-core kernel:
// decreasing work (row blocks of 16 times columns), so that the largest
// leaves start first; built on first use (host sort of the leaves).
void build_pf_order(BF &bf) {
    if (bf.pf_order_built || !bf.pfree_ok) return;
    const int32_t n = bf.pf_n_items;
    std::vector<int4> items(size_t(std::max(n, 1)));
    if (n > 0) BFG_HIP(hipMemcpy(items.data(), bf.pf_items.p, sizeof(int4) * size_t(n), hipMemcpyDeviceToHost));
    std::vector<int32_t> ord(static_cast<size_t>(n));
    for (int32_t i = 0; i < n; ++i) ord[size_t(i)] = i;
    auto work = [&](int32_t i) { const int4 &t = items[size_t(i)]; return int64_t((t.y + 15) / 16) * (t.w - t.z); };
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) { return work(a) > work(b); });
    upload_tab(bf.pf_order, ord);
    bf.pf_order_built = true;
}

// Fused-sweep tables: every forward task of every stage in stage order.
// Items that read the same set of skeleton vectors form one dependency
// group with one counter; each produced vector is read by exactly one group
// (butterfly data flow), so its producer signals that group's counter once.
// The vectors an item reads are the producer outputs inside the input
// ranges of its (merged) contributions.
// Matrix-core tiles of a forward table.  Items with one factor-block
// contribution and at most `tile_rows` rows that read the same input
// entries (same in_ofs and n) are stacked, whole, into tiles of at most
// `tile_rows` rows; every other item (several contributions, an identity
// block, more rows) gets tiles of `tile_rows` rows of its own.  Items are taken in order
// of their first input entry, so consecutive tiles read neighbouring
// inputs; the tasks are copied in that order.
void build_mfma(StageTable &S, int32_t tile_rows) {
    if (S.mfma_built && S.mfma_rows == tile_rows) return;
    BFG_REQUIRE(tile_rows >= 16 && tile_rows <= MTILE_ROWS && tile_rows % 16 == 0, "pack: tile rows 16..64");
    const int32_t n = int32_t(S.h_items.size());
    struct Key { int32_t in_ofs, n, idx; };
    std::vector<Key> keys(size_t(std::max(n, 0)));
    std::vector<uint8_t> stack(size_t(std::max(n, 0)));
    parallel_for(int64_t(n), [&](int64_t i) {
        const MvItem &it = S.h_items[i];
        const MvContrib &c = S.h_contribs[it.c_begin];
        keys[i] = Key{c.in_ofs, c.n, int32_t(i)};
        stack[i] = it.c_end - it.c_begin == 1 && c.fofs >= 0 && it.out_len <= tile_rows;
    }, 4096);
    std::sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
        return a.in_ofs != b.in_ofs ? a.in_ofs < b.in_ofs : a.n != b.n ? a.n < b.n : a.idx < b.idx;
    });
    std::vector<MvTaskC> items(keys.size());
    std::vector<int4> tiles;
    tiles.reserve(keys.size() / 2 + 16);
    int4 open{0, 0, 0, 0};  // open stacked tile (count 0: none)
    int32_t open_in = 0, open_n = 0;
    auto close = [&] {
        if (open.y > 0) tiles.push_back(open);
        open = int4{0, 0, 0, 0};
    };
    for (int32_t p = 0; p < int32_t(keys.size()); ++p) {
        const int32_t i = keys[p].idx;
        const MvTask &h = S.h_tasks[i];
        items[p] = MvTaskC{h.fofs, h.in_ofs, h.out_ofs, h.m, h.n, h.c_begin, h.c_end};
        const int32_t m = S.h_items[i].out_len;
        if (m <= 0) continue;
        if (stack[i]) {
            if (open.y > 0 && (keys[p].in_ofs != open_in || keys[p].n != open_n || open.w + m > tile_rows)) close();
            if (open.y == 0) {
                open = int4{p, 0, 0, 0};
                open_in = keys[p].in_ofs;
                open_n = keys[p].n;
            }
            ++open.y;
            open.w += m;
        } else {
            close();
            for (int32_t r0 = 0; r0 < m; r0 += tile_rows) tiles.push_back(int4{p, 1, r0, std::min(tile_rows, m - r0)});
        }
    }
    close();
    BFG_REQUIRE(tiles.size() < size_t(INT32_MAX), "pack: too many matrix-core tiles");
    // by number of 16-row blocks (one launch per kind), stable within a kind
    auto blocks = [](const int4 &t) { return (t.w + 15) / 16; };
    std::stable_sort(tiles.begin(), tiles.end(), [&](const int4 &a, const int4 &b) { return blocks(a) < blocks(b); });
    for (int rb = 0; rb <= 4; ++rb) S.mtile_ofs[rb] = 0;
    for (const int4 &t : tiles) ++S.mtile_ofs[blocks(t)];
    for (int rb = 1; rb <= 4; ++rb) S.mtile_ofs[rb] += S.mtile_ofs[rb - 1];
    upload_tab(S.mtasks, items);
    upload_tab(S.mtiles, tiles);
    S.n_mtiles = int32_t(tiles.size());
    S.mfma_built = true;
    S.mfma_rows = tile_rows;
}

void build_sweep(BF &bf) {
    SweepTables &sw = bf.sweep;
    const int32_t mode = sw.mode, wg = sw.wg_per_cu;
    sw = SweepTables{};
    sw.mode = mode;
    sw.wg_per_cu = wg;
    sw.built = true;
    const int32_t nL = bf.n_stage_launch;
    if (nL > SWEEP_MAX_STAGES) return;
    for (const StageTable &S : bf.fwd)
        if (S.pre) return;  // partial sums need their own launch
    std::vector<int32_t> starts;  // slot offsets of every produced vector
    for (int32_t L = 0; L + 1 < nL; ++L)
        for (const MvItem &it : bf.fwd[L].h_items) starts.push_back(it.out_ofs);
    std::sort(starts.begin(), starts.end());
    std::map<std::vector<int32_t>, int32_t> gid;
    std::unordered_map<int32_t, int32_t> consumer;  // slot offset -> group
    std::vector<int32_t> gcnt;
    std::vector<std::vector<int32_t>> wait(nL);
    std::vector<int32_t> set;
    for (int32_t L = 1; L < nL; ++L) {
        const StageTable &S = bf.fwd[L];
        wait[L].resize(S.h_items.size());
        for (size_t i = 0; i < S.h_items.size(); ++i) {
            const MvItem &it = S.h_items[i];
            set.clear();
            for (int32_t c = it.c_begin; c < it.c_end; ++c) {
                const MvContrib &k = S.h_contribs[c];
                auto lo = std::lower_bound(starts.begin(), starts.end(), k.in_ofs);
                auto hi = std::lower_bound(starts.begin(), starts.end(), k.in_ofs + k.n);
                set.insert(set.end(), lo, hi);
            }
            std::sort(set.begin(), set.end());
            set.erase(std::unique(set.begin(), set.end()), set.end());
            auto f = gid.find(set);
            int32_t g;
            if (f == gid.end()) {
                g = int32_t(gcnt.size());
                gid.emplace(set, g);
                gcnt.push_back(int32_t(set.size()));
                for (int32_t s : set) {
                    auto c = consumer.find(s);
                    if (c != consumer.end() && c->second != g) return;  // not a butterfly data flow
                    consumer[s] = g;
                }
            } else {
                g = f->second;
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

// HIP events around every kernel launch of a matvec (timing runs only):
// ev[0] before the first launch, ev[i + 1] after launch i, described by
// rec[i] = {level, table (0 level table, 1 partial sums, 2 input
// permutation, 3 fused sweep), variant, output kind, packed launch shape,
// right-hand sides per lane}.
struct Timeline {
    std::vector<hipEvent_t> ev;
    std::vector<std::array<int32_t, 6>> rec;
    size_t n = 0;
    void start(hipStream_t s) {
        n = 0;
        rec.clear();
        mark(s);
    }
    void mark(hipStream_t s) {
        if (ev.size() <= n) {
            hipEvent_t e;
            BFG_HIP(hipEventCreate(&e));
            ev.push_back(e);
        }
        BFG_HIP(hipEventRecord(ev[n++], s));
    }
    void launched(hipStream_t s, std::array<int32_t, 6> r) {
        rec.push_back(r);
        mark(s);
    }
    ~Timeline() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};

static int32_t nr_per_lane(int64_t nrhs) { return nrhs <= 1 ? 1 : nrhs == 2 ? 2 : nrhs <= 4 ? 4 : 8; }

static void stage_launches(BF &bf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y,
                           int64_t ldy, const double *alpha, const double *beta, Timeline *tl, bool lanes,
                           void *slots, bool sweep_allowed = true);

// One launch table of level L: the tiled kernel (after permuting an
// IN_GATHER input into the slot buffer) or the per-task kernels.
static void run_table(BF &bf, int op, const StageTable &t, int32_t L, int64_t nrhs, const void *x,
                      int64_t ldx, void *y, int64_t ldy, const double *alpha, const double *beta,
                      bool fuse, void *slots, Timeline *tl = nullptr, int32_t table = 0) {
    hipStream_t s = bf.ctx->stream;
    const bool fwd = op == BFGPU_OP_N;
    const int32_t *pt = bf.perm_test.as<const int32_t>(), *pr = bf.perm_trial.as<const int32_t>();
    const int32_t *pin = t.in_kind == IN_GATHER ? (fwd ? pr : pt) : nullptr;
    const int32_t *pout = t.out_kind == OUT_SCATTER ? (fwd ? pt : pr) : nullptr;
    if (t.variant >= 2) {
        int64_t shift = 0;
        if (t.in_kind == IN_GATHER) {
            const int es = bf.precision == BFGPU_C128 ? 16 : 8;
            launch_gather(s, bf.precision, x, ldx, pin, t.in_lo, t.in_hi,
                          static_cast<char *>(slots) + size_t(bf.gather_ofs) * es, bf.slot_len, nrhs);
            shift = bf.gather_ofs;
            if (tl) tl->launched(s, {L, 2, 0, 0, 0, 1});
        }
        const void *ft = L < int32_t(bf.factors_t.size()) ? bf.factors_t[L].p : nullptr;
        if (t.variant == 9 && fwd && nrhs == 1) {
            // streaming kernel over its chunks, dense kernel over the other tasks
            launch_stream(s, bf.precision, t, bf.factors[L].p, slots, shift, bf.ctx->n_cu);
            launch_dense(s, bf.precision, t, bf.factors[L].p, slots, bf.slot_len, shift, y, ldy, pout, nrhs,
                         alpha, beta, fuse, bf.ctx->n_cu, DENSE_REST);
        } else if (t.variant == 9 && !fwd && nrhs == 1) {
            // the same over the transposed copy (its items' blocks back to back)
            BFG_REQUIRE(bf.tcopy_built && t.stream_built && t.stream_copy && (ft || bf.factor_elems[L] == 0),
                        "matvec: transposed copy of the blocks or its streaming tables missing");
            const int conj = op == BFGPU_OP_C ? 1 : 0;
            launch_stream(s, bf.precision, t, ft ? ft : bf.factors[L].p, slots, shift, bf.ctx->n_cu, conj != 0);
            launch_dense(s, bf.precision, t, ft ? ft : bf.factors[L].p, slots, bf.slot_len, shift, y, ldy, pout,
                         nrhs, alpha, beta, fuse, bf.ctx->n_cu, DENSE_REST, conj);
        } else if ((t.variant == 3 || t.variant == 9) && fwd) {
            launch_dense(s, bf.precision, t, bf.factors[L].p, slots, bf.slot_len, shift, y, ldy, pout, nrhs,
                         alpha, beta, fuse, bf.ctx->n_cu);
        } else if (t.variant == 8 || t.variant == 9) {  // forward dense kernel over the transposed copy
            BFG_REQUIRE(!fwd && bf.tcopy_built && (ft || bf.factor_elems[L] == 0),
                        "matvec: transposed copy of the blocks missing");
            launch_dense(s, bf.precision, t, ft ? ft : bf.factors[L].p, slots, bf.slot_len, shift, y, ldy, pout,
                         nrhs, alpha, beta, fuse, bf.ctx->n_cu, DENSE_SWAPPED, op == BFGPU_OP_C ? 1 : 0);
        } else if (t.variant == 3) {
            launch_denset(s, bf.precision, op, t, bf.factors[L].p, slots, bf.slot_len, shift, y, ldy, pout,
                          nrhs, alpha, beta, fuse, bf.ctx->n_cu);
        } else {
            launch_tile(s, bf.precision, op, t, bf.factors[L].p, slots, bf.slot_len, shift, y, ldy, pout,
                        nrhs, alpha, beta, fuse, bf.ctx->n_cu);
        }
    } else {
        launch_stage(s, bf.precision, op, t, bf.factors[L].p, x, ldx, y, ldy, slots, bf.slot_len, pin,
                     pout, nrhs, alpha, beta, fuse);
    }
    const int32_t ran = t.variant == 9 && nrhs != 1 ? (fwd ? 3 : 8) : t.variant;  // streams at one RHS only
    if (tl) tl->launched(s, {L, table, ran, t.out_kind, int32_t(pack_tile(t.tile)), nr_per_lane(nrhs)});
}

static void stage_seq_pass(BF &bf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y, int64_t ldy,
                           const double *alpha, const double *beta, Timeline *tl, bool lanes, bool graph_ok);

static void stage_seq(BF &bf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y,
                      int64_t ldy, const double *alpha, const double *beta, Timeline *tl) {
    const int es = bf.precision == BFGPU_C128 ? 16 : 8;
    // right-hand sides per pass: all of them (at most 64 per sweep with
    // right-hand sides across lanes), fewer when the caller caps them or
    // when the slot buffer of one pass would not fit in 90 % of the device
    // memory it may take
    bool lanes = op == BFGPU_OP_N && bf.rhs_lanes_min > 0 && nrhs >= bf.rhs_lanes_min;
    int64_t full = lanes ? std::min<int64_t>(64, nrhs) : nrhs;
    int64_t per = full;
    if (bf.rhs_chunk > 0) per = std::min<int64_t>(per, bf.rhs_chunk);
    auto cols_of = [&](int64_t c) { return size_t(lanes ? (c + 15) / 16 * 16 : c); };
    const size_t col_bytes = size_t(std::max<int64_t>(bf.slot_len, 1)) * es;
    const size_t cap = bf.slot_cap_kib > 0 ? size_t(bf.slot_cap_kib) << 10 : SIZE_MAX;
    if (cols_of(per) * col_bytes > std::min(bf.slots_cap, cap)) {
        size_t free_b = 0, total_b = 0;
        BFG_HIP(hipMemGetInfo(&free_b, &total_b));
        const size_t avail = std::min(cap, size_t(0.9 * double(free_b + bf.slots_cap)));
        if (lanes && cols_of(1) * col_bytes > avail) {
            // not even 16 interleaved columns fit: one column per RHS
            lanes = false;
            full = nrhs;
            per = bf.rhs_chunk > 0 ? std::min<int64_t>(nrhs, bf.rhs_chunk) : nrhs;
        }
        while (per > 1 && cols_of(per) * col_bytes > avail) per = lanes && per > 16 ? per - 16 : (per + 1) / 2;
        BFG_REQUIRE(cols_of(per) * col_bytes <= avail, "matvec: not enough device memory for the skeleton vectors");
    }
    if (per >= full) {
        stage_seq_pass(bf, op, nrhs, x, ldx, y, ldy, alpha, beta, tl, lanes, true);
        return;
    }
    for (int64_t c0 = 0; c0 < nrhs; c0 += per)
        stage_seq_pass(bf, op, std::min(per, nrhs - c0), static_cast<const char *>(x) + size_t(c0) * size_t(ldx) * es,
                       ldx, static_cast<char *>(y) + size_t(c0) * size_t(ldy) * es, ldy, alpha, beta, tl, lanes,
                       false);
}

// One pass over nrhs columns (lanes: right-hand sides across lanes, up to
// 64 per sweep); graph_ok: the per-level launches may be replayed from a
// captured graph.
static void stage_seq_pass(BF &bf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y, int64_t ldy,
                           const double *alpha, const double *beta, Timeline *tl, bool lanes, bool graph_ok) {
    hipStream_t s = bf.ctx->stream;
    const int es = bf.precision == BFGPU_C128 ? 16 : 8;
    // slot buffer: one column per RHS, or (right-hand sides across lanes)
    // up to 64 interleaved RHS per sweep
    const size_t cols = lanes ? size_t((std::min<int64_t>(64, nrhs) + 15) / 16 * 16) : size_t(nrhs);
    const size_t need = size_t(bf.slot_len) * cols * es;
    if (need > bf.slots_cap) {
        release_graph(bf);  // a captured graph holds the old slot buffer
        bf.slots.alloc(need);
        bf.slots_cap = need;
    }
    if (lanes && bf.use_mfma && bf.precision == BFGPU_C64) {  // before any capture: the builds upload
        for (StageTable &t : bf.fwd) {
            if (t.pre) build_mfma(*t.pre, bf.mfma_rows);
            build_mfma(t, bf.mfma_rows);
        }
        if (bf.use_pfree) build_pf_order(bf);
    }
    const bool fused = op == BFGPU_OP_N && !lanes && sweep_ok(bf, nrhs);
    if (tl || !bf.use_graph || fused || !graph_ok) {
        stage_launches(bf, op, nrhs, x, ldx, y, ldy, alpha, beta, tl, lanes, bf.slots.p);
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
        // capture on the context's own stream (the caller's stream may be
        // the null stream, which cannot be captured; capturing runs
        // nothing), replay on the caller's stream
        hipStream_t cap = bf.ctx->own;
        BFG_REQUIRE(cap != nullptr, "matvec: graph capture needs the context's own stream");
        BFG_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        hipGraph_t g = nullptr;
        bf.ctx->stream = cap;
        try {
            stage_launches(bf, op, nrhs, x, ldx, y, ldy, alpha, beta, nullptr, lanes, bf.slots.p);
        } catch (...) {
            bf.ctx->stream = s;
            (void)hipStreamEndCapture(cap, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        bf.ctx->stream = s;
        BFG_HIP(hipStreamEndCapture(cap, &g));
        hipError_t e = hipGraphInstantiate(&bf.graph_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        BFG_HIP(e);
        bf.graph_key = key;
    }
    BFG_HIP(hipGraphLaunch(bf.graph_exec, s));
}

static void stage_launches(BF &bf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y,
                           int64_t ldy, const double *alpha, const double *beta, Timeline *tl, bool lanes,
                           void *slots, bool sweep_allowed) {
    hipStream_t s = bf.ctx->stream;
    const int32_t nL = bf.n_stage_launch;
    const bool fwd = (op == BFGPU_OP_N);
    const int64_t out_len = fwd ? bf.rows : bf.cols;
    const bool fuse = fwd ? bf.y_covered : bf.x_covered;
    static const double one[2] = {1.0, 0.0};
    if (!fuse) launch_scale(s, bf.precision, y, out_len, ldy, nrhs, beta);
    const double *b_eff = fuse ? beta : one;
    const int32_t *pt = bf.perm_test.as<const int32_t>(), *pr = bf.perm_trial.as<const int32_t>();
    if (tl) tl->start(s);
    if (fwd && !lanes && sweep_allowed && sweep_ok(bf, nrhs)) {
        SweepTables &sw = bf.sweep;
        if (int64_t(sw.epoch + 1) * sw.max_wait >= (int64_t(1) << 30)) {
            launch_zero_i32(s, sw.counters.as<int32_t>(), sw.n_groups);
            sw.epoch = 0;
        }
        ++sw.epoch;
        launch_sweep(s, bf.precision, sw, bf.factors, x, ldx, y, ldy, slots, bf.slot_len, pr, pt,
                     nrhs, alpha, b_eff, fuse, bf.ctx->n_cu);
        if (tl) tl->launched(s, {-1, 3, 0, 0, 0, nr_per_lane(nrhs)});
        return;
    }
    if (fwd && lanes) {
        // right-hand sides across lanes, up to 64 per sweep, skeleton
        // vectors interleaved (entry-major) in the slot buffer
        const int es = bf.precision == BFGPU_C128 ? 16 : 8;
        const bool mfma = bf.use_mfma && bf.precision == BFGPU_C64;
        if (mfma)
            for (const StageTable &t : bf.fwd)
                BFG_REQUIRE(t.mfma_built && t.mfma_rows == bf.mfma_rows && (!t.pre || t.pre->mfma_built),
                            "matvec: matrix-core tiles not built");
        for (int64_t c0 = 0; c0 < nrhs; c0 += 64) {
            const int nc = int(std::min<int64_t>(64, nrhs - c0));
            const int32_t ns = (nc + 15) / 16 * 16;
            const char *xc = static_cast<const char *>(x) + size_t(c0) * size_t(ldx) * es;
            char *yc = static_cast<char *>(y) + size_t(c0) * size_t(ldy) * es;
            const StageTable &t0 = bf.fwd[0];
            launch_gather_rhs(s, bf.precision, xc, ldx, pr, t0.in_lo, t0.in_hi,
                              static_cast<char *>(slots) + size_t(bf.gather_ofs) * ns * es, ns, nc);
            if (tl) tl->launched(s, {0, 2, 4, 0, 0, nc});
            for (int32_t L = 0; L < nL; ++L) {
                const StageTable &t = bf.fwd[L];
                if (mfma && L == nL - 1 && bf.pfree_ok && bf.use_pfree && bf.pf_order_built) {
                    // last level's entries from the points, on the matrix cores
                    launch_pfree_mfma(s, bf, slots, ns, yc, ldy, nc, alpha, b_eff, fuse, bf.ctx->n_cu);
                    if (tl) tl->launched(s, {L, 0, 10, OUT_SCATTER, bf.mfma_emul ? 1 : 0, nc});
                    continue;
                }
                const StageTable *pre = t.pre.get();
                for (const StageTable *u : {pre, &t}) {
                    if (!u) continue;
                    const int64_t shift = u->in_kind == IN_GATHER ? bf.gather_ofs : 0;
                    if (mfma) {
                        launch_mfma_rhs(s, *u, bf.factors[L].p, slots, ns, shift, yc, ldy, pt, nc, alpha,
                                        b_eff, fuse, bf.ctx->n_cu);
                    } else {
                        launch_rhs(s, bf.precision, *u, bf.factors[L].p, slots, ns, shift, yc, ldy, pt, nc,
                                   alpha, b_eff, fuse, bf.ctx->n_cu);
                    }
                    if (tl) tl->launched(s, {L, u == &t ? 0 : 1, mfma ? 5 : 4, u->out_kind, mfma ? bf.mfma_rows : 0, nc});
                }
            }
        }
        return;
    }
    for (int32_t i = 0; i < nL; ++i) {
        int32_t L = fwd ? i : nL - 1 - i;
        const StageTable &t = fwd ? bf.fwd[L] : bf.tr[L];
        if (fwd && L == nL - 1 && nrhs == 1 && bf.pfree_ok && bf.use_pfree) {
            launch_pfree(s, bf, slots, y, ldy, alpha, b_eff, fuse);  // P entries from the points
            if (tl) tl->launched(s, {L, 0, 6, OUT_SCATTER, 0, 1});
            continue;
        }
        if (!fwd && L == nL - 1 && nrhs == 1 && bf.pfree_t_ok && bf.use_pfree) {
            launch_pfree_t(s, bf, x, slots, op == BFGPU_OP_C);  // skeleton entries from the points
            if (tl) tl->launched(s, {L, 0, 7, OUT_SLOT, 0, 1});
            continue;
        }
        if (t.pre) run_table(bf, op, *t.pre, L, nrhs, x, ldx, y, ldy, alpha, b_eff, fuse, slots, tl, 1);
        run_table(bf, op, t, L, nrhs, x, ldx, y, ldy, alpha, b_eff, fuse, slots, tl, 0);
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

// ---- allocation-free matvec ----
// The skeleton vectors of bfgpu_matvec_noalloc live in one static device
// array instead of the handle's slot buffer; every GPU that loads the
// library holds its own copy until the process ends.  On the host the
// array's name is a stand-in, not a device address: the device address is
// looked up once per device (noalloc_init, called when a context is
// created on that device, with the device current) and only that pointer
// is ever handed to kernels.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(BFGPU_NOALLOC_SCRATCH_BYTES);
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex noalloc_mu;
static void *noalloc_ptr[NOALLOC_MAX_DEVICES] = {};
static std::string noalloc_err[NOALLOC_MAX_DEVICES];

void noalloc_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return;
    std::lock_guard<std::mutex> g(noalloc_mu);
    if (noalloc_ptr[device]) return;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != device) e = hipSetDevice(device);
    void *p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur >= 0 && cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) {
        // the other entry points do not need the array: keep the reason for
        // bfgpu_matvec_noalloc to report
        noalloc_err[device] = std::string("hipGetSymbolAddress(scratch): ") +
                              (e != hipSuccess ? hipGetErrorString(e) : "null address");
        return;
    }
    noalloc_ptr[device] = p;
}

// y = alpha op(A) x + beta y with device x, y: the per-level launches of
// matvec with one column per right-hand side (never right-hand sides
// across lanes, whose matrix-core tiles are built on first use, never the
// fused sweep or a captured graph), skeleton vectors in the scratch array.
// Every skeleton-vector entry a launch reads is written earlier in the same
// call (the gather of x, then each level's outputs), so the array needs no
// clearing; calls that overlap on one GPU share it and race.
void matvec_noalloc(BF &bf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y, int64_t ldy,
                    const double *alpha, const double *beta) {
    BFG_REQUIRE(op == BFGPU_OP_N || op == BFGPU_OP_T || op == BFGPU_OP_C, "matvec_noalloc: bad op");
    BFG_REQUIRE(nrhs >= 0, "matvec_noalloc: nrhs < 0");
    if (nrhs == 0) return;
    BFG_REQUIRE(x && y, "matvec_noalloc: null x or y");
    const bool fwd = (op == BFGPU_OP_N);
    const int64_t in_len = fwd ? bf.cols : bf.rows, out_len = fwd ? bf.rows : bf.cols;
    BFG_REQUIRE(ldx >= in_len && ldy >= out_len, "matvec_noalloc: leading dimension too small (DimensionMismatch)");
    const int dev = bf.ctx->device;
    BFG_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "matvec_noalloc: device id out of range");
    void *slots = nullptr;
    {
        std::lock_guard<std::mutex> g(noalloc_mu);
        slots = noalloc_ptr[dev];
        if (!slots)
            throw Error(BFGPU_ERR_STATE, "matvec_noalloc: no scratch array on device " + std::to_string(dev) +
                                             (noalloc_err[dev].empty() ? std::string(" (no context created on it)")
                                                                       : ": " + noalloc_err[dev]));
    }
    const int64_t es = bf.precision == BFGPU_C128 ? 16 : 8;
    const int64_t max_rhs = int64_t(NOALLOC_SCRATCH_BYTES) / (std::max<int64_t>(bf.slot_len, 1) * es);
    BFG_REQUIRE(nrhs <= max_rhs,
                "matvec_noalloc: the skeleton vectors of " + std::to_string(nrhs) + " right-hand sides need " +
                    std::to_string(std::max<int64_t>(bf.slot_len, 1) * es * nrhs) + " bytes, more than the " +
                    std::to_string(NOALLOC_SCRATCH_BYTES) + "-byte scratch array (at most " +
                    std::to_string(max_rhs) + " right-hand sides for this operator)");
    stage_launches(bf, op, nrhs, x, ldx, y, ldy, alpha, beta, nullptr, false, slots, false);
}

// Mean duration of every launch over `iters` device matvecs (alpha 1,
// beta 0); launches are the same sequence in every iteration.
static void time_launches(BF &bf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y, int64_t ldy,
                          int32_t iters, std::vector<double> &ms, std::vector<std::array<int32_t, 6>> &rec,
                          double &total_ms) {
    BFG_REQUIRE(iters >= 1, "matvec_timed: iters >= 1");
    static const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    Timeline tl;
    total_ms = 0.0;
    for (int32_t it = 0; it < iters; ++it) {
        stage_seq(bf, op, nrhs, x, ldx, y, ldy, one, zero, &tl);
        const size_t nl = tl.rec.size();
        BFG_HIP(hipEventSynchronize(tl.ev[nl]));
        if (it == 0) {
            ms.assign(nl, 0.0);
            rec = tl.rec;
        }
        BFG_REQUIRE(nl == ms.size(), "matvec_timed: launch sequence changed between iterations");
        for (size_t i = 0; i < nl; ++i) {
            float e = 0.f;
            BFG_HIP(hipEventElapsedTime(&e, tl.ev[i], tl.ev[i + 1]));
            ms[i] += e;
        }
        float e = 0.f;
        BFG_HIP(hipEventElapsedTime(&e, tl.ev[0], tl.ev[nl]));
        total_ms += e;
    }
    for (double &v : ms) v /= iters;
    total_ms /= iters;
    check_sweep_error(bf);
}

void matvec_timed(BF &bf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y,
                  int64_t ldy, int32_t iters, double *stage_ms, int64_t *stage_bytes,
                  int32_t *n_stages, double *total_ms) {
    const int32_t nL = bf.n_stage_launch;
    const int es = bf.precision == BFGPU_C128 ? 16 : 8;
    const bool fwd = (op == BFGPU_OP_N);
    std::vector<double> ms;
    std::vector<std::array<int32_t, 6>> rec;
    double tot = 0.0;
    time_launches(bf, op, nrhs, x, ldx, y, ldy, iters, ms, rec, tot);
    const bool fused = !rec.empty() && rec[0][1] == 3;
    const int32_t nS = fused ? 1 : nL;  // reported entries: levels in launch order
    std::vector<double> acc(nS, 0.0);
    for (size_t i = 0; i < rec.size(); ++i) {
        const int32_t L = rec[i][0];
        acc[fused ? 0 : (fwd ? L : nL - 1 - L)] += ms[i];
    }
    for (int32_t i = 0; i < nS; ++i) {
        if (stage_ms) stage_ms[i] = acc[i];
        if (!stage_bytes) continue;
        if (fused) {
            int64_t b = 0;
            for (const StageTable &t : bf.fwd) b += t.factor_elems * es + t.alg_bytes_1 * nrhs;
            stage_bytes[i] = b;
        } else {
            int32_t L = fwd ? i : nL - 1 - i;
            const StageTable &t = fwd ? bf.fwd[L] : bf.tr[L];
            stage_bytes[i] = t.factor_elems * es + t.alg_bytes_1 * nrhs;
            if (t.pre) stage_bytes[i] += t.pre->factor_elems * es + t.pre->alg_bytes_1 * nrhs;
        }
    }
    if (n_stages) *n_stages = nS;
    if (total_ms) *total_ms = tot;
}

void matvec_launch_times(BF &bf, int op, int64_t nrhs, const void *x, int64_t ldx, void *y, int64_t ldy,
                         int32_t iters, int32_t cap, int32_t *n_launch, double *launch_ms, int32_t *desc) {
    std::vector<double> ms;
    std::vector<std::array<int32_t, 6>> rec;
    double tot = 0.0;
    time_launches(bf, op, nrhs, x, ldx, y, ldy, iters, ms, rec, tot);
    *n_launch = int32_t(ms.size());
    for (int32_t i = 0; i < std::min<int32_t>(cap, int32_t(ms.size())); ++i) {
        if (launch_ms) launch_ms[i] = ms[i];
        if (desc)
            for (int q = 0; q < 6; ++q) desc[6 * i + q] = rec[i][q];
    }
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

// Transposed copy of the stored blocks: every non-identity contribution of a
// level's transposed table (and of its partial-sum table) is one block,
// transposed in place of its own offset in the copy; the swapped task
// tables read each block as n x m.  The level whose transposed product is
// matrix-free is not copied.  Built once, when device memory holds the copy
// with a tenth and 1 GiB to spare.
bool ensure_tcopy(BF &bf) {
    if (bf.tcopy_built) return true;
    if (bf.tcopy_mode == 0 || bf.tcopy_tried) return false;
    bf.tcopy_tried = true;
    const int32_t nL = bf.n_stage_launch, Lmax = nL - 1;
    const size_t es = bf.precision == BFGPU_C128 ? 16 : 8;
    auto copied = [&](int32_t L) {
        return L < int32_t(bf.factor_elems.size()) && bf.factor_elems[L] > 0 && !(L == Lmax && bf.pfree_t_ok);
    };
    size_t bytes = 0;
    for (int32_t L = 0; L < nL; ++L)
        if (copied(L)) bytes += size_t(bf.factor_elems[L] + FACTOR_PAD) * es;
    if (bytes == 0) return false;
    size_t fr = 0, tot = 0;
    BFG_HIP(hipMemGetInfo(&fr, &tot));
    if (bytes + bytes / 10 + (size_t(1) << 30) > fr) return false;
    hipStream_t s = bf.ctx->stream;
    std::vector<DevBuf> ft(static_cast<size_t>(nL));
    DevBuf d_blk;
    for (int32_t L = 0; L < nL; ++L) {
        if (!copied(L)) continue;
        ft[L].alloc(size_t(bf.factor_elems[L] + FACTOR_PAD) * es);
        BFG_HIP(hipMemsetAsync(static_cast<char *>(ft[L].p) + size_t(bf.factor_elems[L]) * es, 0,
                               size_t(FACTOR_PAD) * es, s));
        // the blocks of the level's transposed tables (main, then partial
        // sums) at running offsets in their contribution order: every
        // transposed item's blocks lie back to back, item after item
        hvec<TBlock> blk;
        int64_t run = 0;
        for (StageTable *u : {&bf.tr[L], bf.tr[L].pre.get()}) {
            if (!u) continue;
            u->h_cofs.resize(u->h_contribs.size());
            for (size_t q = 0; q < u->h_contribs.size(); ++q) {
                const MvContrib &c = u->h_contribs[q];
                if (c.fofs < 0 || int64_t(c.m) * c.n == 0) {
                    u->h_cofs[q] = c.fofs < 0 ? -1 : run;
                    continue;
                }
                BFG_REQUIRE(c.fofs + int64_t(c.m) * c.n <= bf.factor_elems[L],
                            "transposed copy: block outside its level's factors");
                blk.push_back(TBlock{c.fofs, run, c.m, c.n});
                u->h_cofs[q] = run;
                run += int64_t(c.m) * c.n;
            }
        }
        BFG_REQUIRE(run <= bf.factor_elems[L], "transposed copy: blocks overlap");
        if (blk.empty()) continue;
        d_blk.reserve(blk.size() * sizeof(TBlock));
        BFG_HIP(hipMemcpyAsync(d_blk.p, blk.data(), blk.size() * sizeof(TBlock), hipMemcpyHostToDevice, s));
        launch_transpose_blocks(s, bf.precision, d_blk.as<const TBlock>(), int64_t(blk.size()), bf.factors[L].p,
                                ft[L].p);
        BFG_HIP(hipStreamSynchronize(s));
    }
    for (int32_t L = 0; L < nL; ++L)
        for (StageTable *u : {&bf.tr[L], bf.tr[L].pre.get()}) {
            if (!u) continue;
            const bool cp = copied(L) && u->h_cofs.size() == u->h_contribs.size();
            auto cofs = [&](size_t q) { return cp ? u->h_cofs[q] : u->h_contribs[q].fofs; };
            hvec<MvTaskC> tc(u->h_tasks.size());
            parallel_for(int64_t(tc.size()), [&](int64_t i) {
                const MvTask &t = u->h_tasks[i];
                const int64_t f = u->h_items[i].c_begin < u->h_items[i].c_end ? cofs(size_t(u->h_items[i].c_begin))
                                                                            : t.fofs;
                tc[i] = MvTaskC{f, t.in_ofs, t.out_ofs, t.n, t.m, t.c_begin, t.c_end};
            }, 4096);
            hvec<MvContrib> cc(u->h_contribs.size());
            parallel_for(int64_t(cc.size()), [&](int64_t i) {
                const MvContrib &c = u->h_contribs[i];
                cc[i] = MvContrib{cofs(size_t(i)), c.n, c.m, c.in_ofs, c.pad};
            }, 4096);
            upload_tab(u->tasks_sw, tc);
            upload_tab(u->contribs_sw, cc);
        }
    bf.factors_t.swap(ft);
    bf.tcopy_built = true;
    release_graph(bf);
    return true;
}

void release_tcopy(BF &bf) {
    for (int32_t L = 0; L < int32_t(bf.tr.size()); ++L)
        for (StageTable *u : {&bf.tr[L], bf.tr[L].pre.get()}) {
            if (!u) continue;
            if (u->variant == 8 || u->variant == 9) {  // back to a lane per output column, same lanes per task
                u->variant = 3;
                u->tile.csplit = 0;
                u->tile.qw4 = 0;
                u->tile.snld4 = 0;
            }
            u->tasks_sw = DevBuf();
            u->contribs_sw = DevBuf();
            u->h_cofs = hvec<int64_t>();
            if (u->stream_copy) {
                u->stream_built = false;
                u->s_chunks = DevBuf();
                u->s_tasks = DevBuf();
                u->s_runs = DevBuf();
                u->s_rest = DevBuf();
            }
        }
    bf.factors_t.clear();
    bf.tcopy_built = false;
    bf.tcopy_tried = false;
    release_graph(bf);
}

// Launch-shape autotuner: for every launch table of both directions
// (forward, and the adjoint whose tables the transpose shares), time each
// kernel variant and tiled shape on device buffers and keep the fastest;
// the choice is stored in the table and used by all later matvecs.
// stage_ms / variants report the forward levels.
void autotune(BF &bf, int64_t nrhs, int32_t iters, double *stage_ms, int32_t *variants) {
    BFG_REQUIRE(nrhs >= 1 && iters >= 1, "autotune: nrhs, iters >= 1");
    hipStream_t s = bf.ctx->stream;
    const int es = bf.precision == BFGPU_C128 ? 16 : 8;
    const int64_t nv = std::max(bf.rows, bf.cols);
    DevBuf dx, dy;
    dx.alloc(size_t(nv) * nrhs * es);
    dy.alloc(size_t(nv) * nrhs * es);
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
    ensure_tcopy(bf);  // when the mode and device memory allow
    const int32_t nL = bf.n_stage_launch;
    static const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    hipEvent_t e0, e1;
    BFG_HIP(hipEventCreate(&e0));
    BFG_HIP(hipEventCreate(&e1));
    auto time_table = [&](int op, StageTable &t, int32_t L) {
        run_table(bf, op, t, L, nrhs, dx.p, nv, dy.p, nv, one, zero, true, bf.slots.p);
        BFG_HIP(hipEventRecord(e0, s));
        for (int32_t k = 0; k < iters; ++k) run_table(bf, op, t, L, nrhs, dx.p, nv, dy.p, nv, one, zero, true, bf.slots.p);
        BFG_HIP(hipEventRecord(e1, s));
        BFG_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        BFG_HIP(hipEventElapsedTime(&ms, e0, e1));
        return double(ms) / iters;
    };
    auto tune = [&](int op, StageTable &t, int32_t L) {
        if (t.n_items == 0) return 0.0;
        const int32_t v0 = t.variant;
        const TileShape t0 = t.tile;
        double best = 1e30;
        int32_t bv = v0;
        TileShape bs = t0;
        auto consider = [&](int32_t v, TileShape sh) {
            t.variant = v;
            t.tile = sh;
            const double ms = time_table(op, t, L);
            if (ms < best) { best = ms; bv = v; bs = sh; }
        };
        consider(0, t0);
        if (op == BFGPU_OP_N) consider(1, t0);
        for (int rp = std::max(0, t0.rp_shift - 1); rp <= std::min(6, t0.rp_shift + 1); ++rp)
            for (int gs = 0; gs <= std::min(3, 6 - rp); ++gs)
                for (int cs = 0; cs <= (t.max_nc >= 2 ? 1 : 0); ++cs) {
                    if (cs == 1 && gs == 0) continue;  // same as cs == 0 with one group
                    if (rp + gs < 2) continue;          // at least 4 lanes per task
                    TileShape sh = t0;
                    sh.rp_shift = int8_t(rp);
                    sh.g_shift = int8_t(gs);
                    sh.csplit = int8_t(cs);
                    consider(2, sh);
                }
        if (op != BFGPU_OP_N)  // transposed dense: a lane per output column, inputs by permutes or LDS
            for (int lp = 2; lp <= 6; ++lp)
                for (int xl = 0; xl <= 1; ++xl) {
                    TileShape sh = t0;
                    sh.rp_shift = int8_t(lp);
                    sh.g_shift = 0;
                    sh.csplit = int8_t(xl);
                    sh.qw4 = 0;
                    consider(3, sh);
                }
        // the forward dense kernel over the transposed copy of the blocks
        const bool tc = op != BFGPU_OP_N && bf.tcopy_built &&
                        (bf.factor_elems[L] == 0 || (L < int32_t(bf.factors_t.size()) && bf.factors_t[L].p));
        if (op == BFGPU_OP_N || tc)
            for (int lp = 2; lp <= 6; ++lp)
                for (int cs = 0; cs <= (t.max_nc >= 2 ? 1 : 0); ++cs)
                    for (int q4 = 0; q4 <= 1; ++q4) {
                        TileShape sh = t0;
                        sh.rp_shift = int8_t(lp);
                        sh.g_shift = 0;
                        sh.csplit = int8_t(cs);
                        sh.qw4 = int8_t(q4);
                        consider(op == BFGPU_OP_N ? 3 : 8, sh);
                    }
        // the streaming kernel (one RHS; forward over the stored blocks,
        // transposed over the transposed copy): chunks of 8 and 4 KiB; its
        // other tasks keep the best dense shape
        if (nrhs == 1 && t.out_kind == OUT_SLOT && (op == BFGPU_OP_N || tc)) {
            const bool dense_best = bv == 3 || bv == 8;
            TileShape base = dense_best ? bs : t0;
            if (!dense_best) {
                base.rp_shift = int8_t(std::max<int>(4, std::min<int>(6, base.rp_shift + base.g_shift)));
                base.g_shift = 0;
                base.csplit = 0;
            }
            for (int nld : {8, 4}) {
                build_stream(t, es, nld, bf.factor_elems[L], op != BFGPU_OP_N);
                for (int wg : {4, 8}) {
                    TileShape sh = base;
                    sh.snld4 = int8_t(nld == 4 ? 1 : 0);
                    sh.wg_per_cu = int16_t(wg);
                    sh.xcd = 0;
                    consider(9, sh);
                }
            }
        }
        if (bv >= 2) {  // the other task-to-XCD mapping, then other occupancies, for the best shape
            TileShape sh = bs;
            sh.xcd = int8_t(1 - sh.xcd);
            consider(bv, sh);
            const TileShape b0 = bs;
            for (int wg : {2, 3, 4, 5, 6, 12, 16}) {
                if (wg == b0.wg_per_cu) continue;
                sh = b0;
                sh.wg_per_cu = int16_t(wg);
                consider(bv, sh);
            }
        }
        if (bv == 9) build_stream(t, es, bs.snld4 ? 4 : 8, bf.factor_elems[L], op != BFGPU_OP_N);
        t.variant = bv;
        t.tile = bs;
        return best;
    };
    try {
        for (int dir = 0; dir < 2; ++dir) {
            const int op = dir == 0 ? BFGPU_OP_N : BFGPU_OP_C;
            for (int32_t L = 0; L < nL; ++L) {
                StageTable &t = dir == 0 ? bf.fwd[L] : bf.tr[L];
                double ms = 0.0;
                if (t.pre) ms += tune(op, *t.pre, L);
                ms += tune(op, t, L);
                if (dir == 0) {
                    if (stage_ms) stage_ms[L] = ms;
                    if (variants) variants[L] = t.variant;
                }
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    BFG_HIP(hipStreamSynchronize(s));
}

}  // namespace bfg
