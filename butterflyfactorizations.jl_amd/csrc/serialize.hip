// Serialized factorizations: an operator written to a file and read back
// onto the device, so that a factorization can be cached between runs or
// handed to another program.  The file holds what the launch tables are
// built from -- every factorization's blocks in the reference layout
// (bfstructs.jl:21-147: keys, sizes, identity flag, reference slot ids)
// with their offsets into the per-level factor buffers, the skeleton and
// compression logs -- and the factor buffers themselves in their stored
// precision.  Loading uploads the buffers and rebuilds the tables, so a
// loaded operator applies exactly as the one that was saved.
//
// Layout (little endian):
//   "BFGPUSV1" | u32 version = 2 | i32 precision | i32 launch levels
//   | u64 observer-tree fingerprint | u64 source-tree fingerprint
//   | i64 compressions | i64 adaptive doublings | i64 factorizations
//   | per factorization: i32 no, ns, stages | per stage: i64 blocks, Block[]
//     | i64 n, SkelRec[] | i64 n, i32[] skeleton indices | i64 n, CompRec[]
//   | per launch level: i64 entries, entries x (8 or 16) bytes of factors
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "internal.hpp"

namespace bfg {
void set_last_error(const std::string &m);
void pack_tables(BF &bf);

namespace {
constexpr char MAGIC[8] = {'B', 'F', 'G', 'P', 'U', 'S', 'V', '1'};
constexpr uint32_t VERSION = 2;  // 2: compression records carry the judge verdict

// FNV-1a over the arrays that fix a tree's node numbering and values.
uint64_t fingerprint(const Tree &t) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void *p, size_t n) {
        const unsigned char *c = static_cast<const unsigned char *>(p);
        for (size_t i = 0; i < n; ++i) {
            h ^= c[i];
            h *= 1099511628211ull;
        }
    };
    mix(&t.n_nodes, sizeof(t.n_nodes));
    mix(&t.n_values, sizeof(t.n_values));
    for (const std::vector<int32_t> *v : {&t.parent, &t.first_child, &t.n_children, &t.val_begin, &t.val_end, &t.perm})
        mix(v->data(), v->size() * sizeof(int32_t));
    return h;
}

struct File {
    FILE *f = nullptr;
    std::string path;
    File(const char *p, const char *mode) : f(std::fopen(p, mode)), path(p) {
        BFG_REQUIRE(f, "cannot open " + path);
    }
    ~File() {
        if (f) std::fclose(f);
    }
    void put(const void *p, size_t n) {
        if (n) BFG_REQUIRE(std::fwrite(p, 1, n, f) == n, "write failed: " + path);
    }
    void get(void *p, size_t n) {
        if (n) BFG_REQUIRE(std::fread(p, 1, n, f) == n, "truncated file: " + path);
    }
    template <class T> void put_vec(const std::vector<T> &v) {
        const int64_t n = int64_t(v.size());
        put(&n, sizeof(n));
        put(v.data(), v.size() * sizeof(T));
    }
    template <class T> void get_vec(std::vector<T> &v, int64_t limit) {
        int64_t n = 0;
        get(&n, sizeof(n));
        BFG_REQUIRE(n >= 0 && n <= limit, "corrupt file (array length): " + path);
        v.resize(size_t(n));
        get(v.data(), v.size() * sizeof(T));
    }
};
}  // namespace
}  // namespace bfg

using namespace bfg;

#define SV_API_BEGIN try {
#define SV_API_END                        \
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

extern "C" int bfgpu_bf_save(const bfgpu_bf *h, const char *path) {
    SV_API_BEGIN
    BFG_REQUIRE(h && path, "bf_save: null argument");
    const BF &bf = h->b;
    File f(path, "wb");
    f.put(MAGIC, sizeof(MAGIC));
    f.put(&VERSION, sizeof(VERSION));
    f.put(&bf.precision, sizeof(int32_t));
    f.put(&bf.n_stage_launch, sizeof(int32_t));
    const uint64_t fo = fingerprint(*bf.test), fs = fingerprint(*bf.trial);
    f.put(&fo, 8);
    f.put(&fs, 8);
    f.put(&bf.n_comp, 8);
    f.put(&bf.n_retry, 8);
    const int64_t nb = int64_t(bf.bfs.size());
    f.put(&nb, 8);
    for (const OneBF &ob : bf.bfs) {
        const int32_t hdr[3] = {ob.no, ob.ns, int32_t(ob.st.size())};
        f.put(hdr, sizeof(hdr));
        for (const StageBlocks &st : ob.st) f.put_vec(st.b);
        f.put_vec(ob.skel);
        f.put_vec(ob.skel_idx);
        f.put_vec(ob.comp);
    }
    // factor buffers, one level at a time through a bounded host buffer
    const size_t es = bf.precision == BFGPU_C128 ? 16 : 8;
    std::vector<char> buf;
    for (int32_t L = 0; L < bf.n_stage_launch; ++L) {
        const int64_t ne = bf.factor_elems[L];
        f.put(&ne, 8);
        const size_t total = size_t(ne) * es, step = size_t(1) << 28;
        buf.resize(std::min(total, step));
        for (size_t o = 0; o < total; o += step) {
            const size_t n = std::min(step, total - o);
            BFG_HIP(hipMemcpy(buf.data(), static_cast<const char *>(bf.factors[L].p) + o, n, hipMemcpyDeviceToHost));
            f.put(buf.data(), n);
        }
    }
    BFG_REQUIRE(std::fflush(f.f) == 0, std::string("write failed: ") + path);
    SV_API_END
}

extern "C" int bfgpu_bf_load(bfgpu_ctx *ctx, const bfgpu_tree *test, const bfgpu_tree *trial, const char *path,
                             bfgpu_bf **out) {
    SV_API_BEGIN
    BFG_REQUIRE(ctx && test && trial && path && out, "bf_load: null argument");
    BFG_HIP(hipSetDevice(ctx->c.device));
    File f(path, "rb");
    char magic[8];
    uint32_t version = 0;
    f.get(magic, 8);
    BFG_REQUIRE(std::memcmp(magic, MAGIC, 8) == 0, std::string("not a saved factorization: ") + path);
    f.get(&version, 4);
    BFG_REQUIRE(version == VERSION, "unsupported file version " + std::to_string(version));
    auto h = std::make_unique<bfgpu_bf>();
    BF &bf = h->b;
    bf.ctx = &ctx->c;
    bf.test = &test->t;
    bf.trial = &trial->t;
    f.get(&bf.precision, 4);
    f.get(&bf.n_stage_launch, 4);
    BFG_REQUIRE(bf.precision == BFGPU_C64 || bf.precision == BFGPU_C128, "corrupt file (precision)");
    BFG_REQUIRE(bf.n_stage_launch >= 2 && bf.n_stage_launch <= 64, "corrupt file (levels)");
    uint64_t fo = 0, fs = 0;
    f.get(&fo, 8);
    f.get(&fs, 8);
    BFG_REQUIRE(fo == fingerprint(test->t) && fs == fingerprint(trial->t),
                "bf_load: the trees differ from the ones the factorization was built on");
    f.get(&bf.n_comp, 8);
    f.get(&bf.n_retry, 8);
    int64_t nb = 0;
    f.get(&nb, 8);
    BFG_REQUIRE(nb >= 1 && nb < (int64_t(1) << 40), "corrupt file (factorizations)");
    bf.bfs.resize(size_t(nb));
    const int64_t lim = int64_t(1) << 40;
    for (OneBF &ob : bf.bfs) {
        int32_t hdr[3];
        f.get(hdr, sizeof(hdr));
        BFG_REQUIRE(hdr[2] >= 2 && hdr[2] <= bf.n_stage_launch, "corrupt file (stages)");
        BFG_REQUIRE(hdr[0] >= 0 && hdr[0] < test->t.n_nodes && hdr[1] >= 0 && hdr[1] < trial->t.n_nodes,
                    "corrupt file (roots)");
        ob.no = hdr[0];
        ob.ns = hdr[1];
        ob.st.resize(size_t(hdr[2]));
        for (StageBlocks &st : ob.st) f.get_vec(st.b, lim);
        f.get_vec(ob.skel, lim);
        f.get_vec(ob.skel_idx, lim);
        f.get_vec(ob.comp, lim);
        for (const CompRec &c : ob.comp) {
            bf.n_judged += c.judge != 0;
            bf.n_fallback += c.judge == 2;
        }
        ob.stage_data_base.assign(ob.st.size(), 0);
    }
    const size_t es = bf.precision == BFGPU_C128 ? 16 : 8;
    bf.factors.resize(size_t(bf.n_stage_launch));
    bf.factor_elems.assign(size_t(bf.n_stage_launch), 0);
    std::vector<char> buf;
    for (int32_t L = 0; L < bf.n_stage_launch; ++L) {
        int64_t ne = 0;
        f.get(&ne, 8);
        BFG_REQUIRE(ne >= 0 && ne < lim, "corrupt file (factor entries)");
        bf.factor_elems[L] = ne;
        bf.factors[L].alloc(size_t(ne + FACTOR_PAD) * es);
        BFG_HIP(hipMemset(static_cast<char *>(bf.factors[L].p) + size_t(ne) * es, 0, size_t(FACTOR_PAD) * es));
        const size_t total = size_t(ne) * es, step = size_t(1) << 28;
        buf.resize(std::min(total, step));
        for (size_t o = 0; o < total; o += step) {
            const size_t n = std::min(step, total - o);
            f.get(buf.data(), n);
            BFG_HIP(hipMemcpy(static_cast<char *>(bf.factors[L].p) + o, buf.data(), n, hipMemcpyHostToDevice));
        }
    }
    // Block keys index the trees' node arrays in the table packer, and block
    // data must lie in its level's buffer: both are checked here, before
    // the packer reads any of them; the packer checks the slot, input and
    // output ranges it derives before any launch.
    const Tree &te = test->t, &tr = trial->t;
    for (const OneBF &ob : bf.bfs)
        for (size_t s = 0; s < ob.st.size(); ++s) {
            const int32_t L = (s + 1 == ob.st.size()) ? bf.n_stage_launch - 1 : int32_t(s);
            const bool first = s == 0, last = s + 1 == ob.st.size();
            for (const Block &b : ob.st[s].b) {
                BFG_REQUIRE(b.obs_out >= 0 && b.obs_out < te.n_nodes && b.obs_in >= 0 && b.obs_in < te.n_nodes &&
                                b.src_out >= 0 && b.src_out < tr.n_nodes && b.src_in >= 0 && b.src_in < tr.n_nodes,
                            "corrupt file (block key outside the trees)");
                BFG_REQUIRE(b.m >= 0 && b.n >= 0 &&
                                (b.ident ? b.m == b.n : b.fofs >= 0 && b.fofs + int64_t(b.m) * b.n <= bf.factor_elems[L]),
                            "corrupt file (block outside its factor buffer)");
                // Q blocks read the values of their source node, P blocks
                // write the values of their observer node
                BFG_REQUIRE(!first || b.ident || b.n == tr.nvals(b.src_in),
                            "corrupt file (Q block columns differ from its source node's values)");
                BFG_REQUIRE(!last || b.ident || b.m == te.nvals(b.obs_out),
                            "corrupt file (P block rows differ from its observer node's values)");
            }
        }
    pack_tables(bf);
    *out = h.release();
    SV_API_END
}
