// internal.hpp -- shared declarations of the gfx950 butterfly library.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <exception>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bfgpu.h"

namespace bfg {

// Error carried from deep inside the library to the C ABI boundary.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &m);

#define BFG_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw ::bfg::Error(BFGPU_ERR_HIP, std::string(#call) + ": " +                 \
                                                  hipGetErrorString(e_) + " at " __FILE__ \
                                                  ":" + std::to_string(__LINE__));         \
    } while (0)

#define BFG_REQUIRE(cond, msg)                                                     \
    do {                                                                           \
        if (!(cond)) throw ::bfg::Error(BFGPU_ERR_ARG, std::string(msg));          \
    } while (0)

// ---------------------------------------------------------------------------
// Cluster tree (host).  Node 0 is the root; children contiguous; the values
// of a node are perm[val_begin, val_end).
// ---------------------------------------------------------------------------
struct Tree {
    int32_t n_nodes = 0, n_values = 0;
    std::vector<int32_t> parent, first_child, n_children, val_begin, val_end, perm, level;
    std::vector<double> center, halfsize, points;
    bool isleaf(int32_t v) const { return n_children[v] == 0; }
    int32_t nvals(int32_t v) const { return val_end[v] - val_begin[v]; }
    const int32_t *vals(int32_t v) const { return perm.data() + val_begin[v]; }
    double radius(int32_t v) const { return halfsize[v] * 1.7320508075688772; }
};

// treelevels + traverseandpad (src/auxillaries.jl:263-281, :340-350).
std::vector<std::vector<int32_t>> traverse_and_pad(const Tree &t, int32_t root);
int32_t tree_height(const Tree &t);

// ---------------------------------------------------------------------------
// Device buffer (RAII).
// ---------------------------------------------------------------------------
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release();
    void alloc(size_t n);              // exact size (frees old)
    void reserve(size_t n);            // grow-only
    template <class T> T *as() const { return static_cast<T *>(p); }
};

struct Ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    int n_cu = 0;
};

// ---------------------------------------------------------------------------
// Factorization in the reference block layout + packed device tables.
// ---------------------------------------------------------------------------
struct Block {
    int32_t obs_out, src_out, obs_in, src_in;
    int32_t m, n, ident;
    int64_t fofs;      // element offset into the stage's factor buffer (-1 identity)
    int32_t out_ptr, in_ptr;  // reference slot ids (first use order per level)
};

struct SkelRec { int32_t stage, s, o, ofs, len; };
// judge: 0 not judged, 1 accepted by the residual judge, 2 rejected and
// redone by plain adaptive RRQR
struct CompRec { int32_t stage, s, o, guess, rows, rank; float feat[6]; int32_t judge; };

// One stage of one factorization in reference layout.
struct StageBlocks { std::vector<Block> b; };

struct OneBF {
    int32_t no = 0, ns = 0;
    std::vector<StageBlocks> st;      // Q, R^1..R^{L-1}, P
    std::vector<SkelRec> skel;
    std::vector<int32_t> skel_idx;
    std::vector<CompRec> comp;
    std::vector<int64_t> stage_data_base;  // per stage: element base in that stage's buffer
    // source point ids of the columns of the P blocks, block after block
    // (P blocks are kernel entries G(observer leaf, these points)); empty
    // when the P blocks are not known to be kernel entries
    std::vector<int32_t> pcols;
};

// Work item of one stage launch: an output target with a list of
// contributions (factor blocks) summed into it.
// Host memory for large plain arrays (HOST_BIG_MIN bytes and more): 2 MiB
// aligned anonymous mappings advised for transparent huge pages, and freed
// blocks kept for reuse (up to a cap), so that the temporaries of one level,
// launch table or call do not fault fresh pages in again for the next.
// BFGPU_HEAP_KEEP=0 returns every block to the system when it is freed.
constexpr size_t HOST_BIG_MIN = size_t(4) << 20;
void *host_big_alloc(size_t bytes);
void host_big_free(void *p, size_t bytes);

// Host arrays of plain records whose elements are all written after
// allocation: resize() leaves new elements uninitialized (no serial
// zero-fill of hundreds of MB before a parallel fill); large ones come from
// host_big_alloc.
template <class T> struct NoInitAlloc : std::allocator<T> {
    template <class U> struct rebind { using other = NoInitAlloc<U>; };
    NoInitAlloc() = default;
    template <class U> NoInitAlloc(const NoInitAlloc<U> &) {}
    T *allocate(size_t n) {
        if (n * sizeof(T) >= HOST_BIG_MIN) return static_cast<T *>(host_big_alloc(n * sizeof(T)));
        return std::allocator<T>::allocate(n);
    }
    void deallocate(T *p, size_t n) {
        if (n * sizeof(T) >= HOST_BIG_MIN) host_big_free(p, n * sizeof(T));
        else std::allocator<T>::deallocate(p, n);
    }
    template <class U> void construct(U *p) noexcept { ::new (static_cast<void *>(p)) U; }
    template <class U, class... A> void construct(U *p, A &&...a) {
        ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
    }
};
template <class T> using hvec = std::vector<T, NoInitAlloc<T>>;

struct MvItem {
    int32_t out_ofs;   // slot offset, or value offset into the scatter perm
    int32_t out_len;
    int32_t c_begin, c_end;  // contributions [c_begin, c_end)
};
struct MvContrib {
    int64_t fofs;      // element offset of the column-major block (-1: identity)
    int32_t m, n;      // block rows / cols
    int32_t in_ofs;    // slot offset, or value offset into the gather perm
    int32_t pad;
};

// Forward task: an item with its first contribution inline (one descriptor
// load per workgroup); contributions [c_begin, c_end) follow in contribs.
struct MvTask {
    int64_t fofs;      // first contribution
    int32_t m, n, in_ofs;
    int32_t out_ofs, out_len;
    int32_t c_begin, c_end;
    int32_t wait_grp, wait_cnt, sig_grp;  // dependency counters (fused sweep)
    int32_t pad;
};

// Compact task of the tiled and dense kernels (32 bytes, read by a wave as
// one coalesced load of its task group): first contribution inline,
// further ones in contribs[c_begin, c_end).
struct MvTaskC {
    int64_t fofs;
    int32_t in_ofs, out_ofs;
    int32_t m, n;
    int32_t c_begin, c_end;
};
static_assert(sizeof(MvTaskC) == 32, "MvTaskC is 8 dwords");

// Streaming forward kernel (one right-hand side).  A chunk is a run of
// consecutive single-block tasks whose blocks lie back to back in the
// level's factor buffer; a wavefront copies the chunk's factor range (nw
// 16-byte words from word f0w) and the distinct input vectors its tasks read
// (X entries) into its own LDS region and computes the chunk from there.
//   w0    = f0w | nw << 40 | X << 51
//   info  = nt | H << 7 | G << 14: tasks, row pairs, lane groups per row pair
//   rmask = bit p set where a task's row pairs start (its first is pair p)
//   xmask = bit e (word e / 64) set where an input vector starts in the
//           chunk's LDS input region (entry e)
//   input vectors: runs[r0 .. r0 + nruns), in LDS order
struct StreamChunk {
    uint64_t w0;
    int32_t t0;       // first task (index into the stream task array)
    uint32_t info;
    uint64_t rmask;
    uint64_t xmask[4];
    int32_t r0, nruns;
};
static_assert(sizeof(StreamChunk) == 64, "StreamChunk is 16 dwords");
// Task t of a chunk: out_ofs its output slot; a, xo the LDS offsets
// (entries) of its block (m x n: its blocks side by side) and of its input
// (its input vectors one after the other); hp the chunk's row pairs before
// it.
//   ax  = a | xo << 11
//   mhn = m | hp << 8 | n << 15
struct StreamTask {
    int32_t lin, out_ofs;  // lin unused
    uint32_t ax, mhn;
};
static_assert(sizeof(StreamTask) == 16, "StreamTask is 4 dwords");
// Input vector of a chunk: slot offset, LDS offset (entries).
struct StreamRun {
    int32_t lin, key;
};
// Input entries per chunk in LDS (1.5 KiB in FP32, 2 KiB in FP64; a
// multiple of 64 and at most 256, the bits of xmask)
template <typename R> constexpr int stream_xcap() { return sizeof(R) == 4 ? 192 : 128; }
constexpr int STREAM_HMAX = 32;  // row pairs per chunk: at least two lane groups per row pair

enum InKind { IN_SLOT = 0, IN_GATHER = 1 };
enum OutKind { OUT_SLOT = 0, OUT_SCATTER = 1 };

// Launch shape of the tiled stage kernel: 2^rp_shift output entries times
// 2^g_shift lane groups per task; groups split the inner index (csplit 0)
// or the contributions (csplit 1); persistent grid of wg_per_cu
// workgroups per CU.  For the transposed dense kernel csplit 1 selects
// grouped rows (lanes on the rows of the block columns).
struct TileShape {
    int8_t rp_shift = 3, g_shift = 0, csplit = 0;
    int16_t wg_per_cu = 8;
    int8_t xcd = 0;  // 1: each XCD walks one contiguous range of tasks
    int8_t qw4 = 0;  // dense kernels, one RHS: 4 loads per lane and pass (else 8)
    int8_t snld4 = 0;  // streaming kernel: chunks of 4 KiB (else 8 KiB)
};
inline int64_t pack_tile(const TileShape &t) {
    return int64_t(t.rp_shift) | int64_t(t.g_shift) << 4 | int64_t(t.csplit) << 8 | int64_t(t.wg_per_cu) << 12 |
           int64_t(t.xcd) << 20 | int64_t(t.qw4) << 24 | int64_t(t.snld4) << 28;
}
constexpr int64_t TILE_PACK_LIMIT = int64_t(1) << 29;  // packed shapes are below this
inline TileShape unpack_tile(int64_t v) {
    TileShape t;
    t.rp_shift = int8_t(v & 15);
    t.g_shift = int8_t((v >> 4) & 15);
    t.csplit = int8_t((v >> 8) & 15);
    t.wg_per_cu = int16_t((v >> 12) & 255);
    t.xcd = int8_t((v >> 20) & 15);
    t.qw4 = int8_t((v >> 24) & 15);
    t.snld4 = int8_t((v >> 28) & 1);
    return t;
}
// at least 4 lanes per task (a wave's task group is at most 16 descriptors)
inline bool tile_ok(const TileShape &t) {
    return t.rp_shift >= 0 && t.g_shift >= 0 && t.rp_shift + t.g_shift >= 2 && t.rp_shift + t.g_shift <= 6 &&
           t.csplit <= 1 &&
           t.wg_per_cu >= 1 && t.wg_per_cu <= 64 && t.xcd <= 1 && t.qw4 <= 1 && t.snld4 <= 1;
}

struct StageTable {
    int32_t n_items = 0, n_contrib = 0;
    int32_t in_kind = IN_SLOT, out_kind = OUT_SLOT;
    int32_t max_in = 0, max_out = 0;   // longest input / output vector
    int64_t alg_bytes_1 = 0;           // algorithmic bytes per RHS (factor bytes separate)
    int64_t factor_elems = 0;          // factor entries read
    // kernel: 0 workgroup per task, 1 wavefront per task (forward only),
    // 2 tiled (several tasks per wavefront, persistent), 3 dense (forward:
    // lanes packed to each block's row count; transposed: a lane per output
    // column; tile.rp_shift is the log2 of the lanes per task), 8
    // (transposed only) the forward dense kernel over the transposed copy of
    // the blocks (BF::factors_t) with tasks_sw / contribs_sw
    int32_t variant = 0;
    TileShape tile;
    int32_t in_lo = 0, in_hi = 0;      // IN_GATHER: value range read through the permutation
    int32_t max_nc = 0;                // most contributions of one item
    DevBuf items, contribs, tasks, tasks_c;
    // transposed tables over the transposed copy: tasks_c / contribs with
    // every block's rows and columns swapped (built with the copy)
    DevBuf tasks_sw, contribs_sw;
    hvec<int64_t> h_cofs;  // per contribution: its block's offset in the transposed copy (-1 identity)
    hvec<MvItem> h_items;
    hvec<MvContrib> h_contribs;
    hvec<MvTask> h_tasks;
    // items split into partial sums: `pre` writes the partials into slots,
    // this table's combine items add them (identity contributions)
    std::unique_ptr<StageTable> pre;
    // matrix-core tiles of the multi-RHS kernel (built on first use):
    // int4 {first, count, row0, rows} over mtasks, the tasks in stacked
    // order; tiles sorted by their number of 16-row blocks
    bool mfma_built = false;
    int32_t mfma_rows = 0;  // tile height the tiles were built for
    int32_t n_mtiles = 0;
    int32_t mtile_ofs[5] = {0, 0, 0, 0, 0};  // tiles of 1, 2, 3, 4 row blocks: [ofs[rb-1], ofs[rb])
    DevBuf mtiles, mtasks;
    // streaming kernel (variant 9, forward, one RHS; built on first use for
    // chunks of stream_nld 1 KiB rounds): chunks over the tasks it takes,
    // and the other tasks (several contributions, identity blocks, blocks
    // too large for a chunk) as a dense-kernel task list
    bool stream_built = false;
    bool stream_copy = false;  // built over the transposed copy (transposed tables)
    int32_t stream_nld = 0;
    int32_t n_schunks = 0, n_stasks = 0, n_srest = 0;
    int64_t stream_gap = 0;   // entries copied that belong to no block
    int64_t stream_xent = 0;  // input entries copied (distinct per chunk)
    DevBuf s_chunks, s_tasks, s_runs, s_rest;
};

constexpr int SWEEP_MAX_STAGES = 48;
struct SweepTables {
    bool built = false;     // tables built (on first enable)
    bool ok = false;        // built and the data flow admits the sweep
    int32_t n_tasks = 0, n_groups = 0, max_wait = 1;
    DevBuf tasks, contribs, counters, err;  // counters: n_groups + ticket pair
    int32_t epoch = 0;
    int32_t mode = 0;       // 0: dynamic tickets, ordinary launch; 1: cooperative, static
    int32_t wg_per_cu = 0;  // workgroups per CU (0: occupancy limit)
};

struct BF {
    Ctx *ctx = nullptr;
    int32_t precision = BFGPU_C128;
    const Tree *test = nullptr;   // observer tree
    const Tree *trial = nullptr;  // source tree
    std::vector<OneBF> bfs;
    int32_t n_stage_launch = 0;   // Q, R^1..R^{Lmax-1}, P
    std::vector<DevBuf> factors;  // per launch stage: packed factor data
    std::vector<int64_t> factor_elems;
    int64_t slot_len = 0;         // per RHS: skeleton vectors, partial sums, permuted inputs
    int64_t part_ofs = 0;         // first partial-sum entry
    int64_t gather_ofs = 0;       // permuted inputs: value v at gather_ofs + v
    int64_t rows = 0, cols = 0;
    bool y_covered = true;        // forward P covers every row of y
    bool x_covered = true;        // transposed Q covers every column
    std::vector<StageTable> fwd, tr;
    SweepTables sweep;            // fused forward sweep
    bool use_sweep = false;       // measured slower than per-level launches; opt-in
    bool use_graph = false;       // replay per-level launches from a HIP graph
    int32_t rhs_lanes_min = 32;   // forward matvecs with >= this many RHS: RHS across lanes (0: never)
    bool use_mfma = true;         // RHS across lanes in complex FP32: matrix-core kernel
    int32_t mfma_rows = 32;       // most rows of its tiles (16, 32, 48 or 64)
    bool mfma_emul = true;        // matrix-free last level: FP32 products emulated on the bf16 matrix cores (bf16x3)
    int32_t rhs_chunk = 0;        // most right-hand sides per pass (0: as device memory allows)
    int32_t slot_cap_kib = 0;     // most KiB of skeleton vectors per pass (0: 90 % of the free memory)
    // matrix-free last level (forward, one RHS, complex FP32): per observer
    // leaf, its rows and the flat list of P columns (source point, input
    // slot); the entries G(x, y) are computed from the points
    bool pfree_ok = false, use_pfree = true;
    int32_t pf_n_items = 0;
    double pf_k = 0.0;            // wavenumber of the point kernel
    DevBuf pf_items, pf_col, pf_in, pf_row;  // matrix-free last level tables
    // matrix-core form of the matrix-free last level (right-hand sides
    // across lanes): items in order of decreasing work, built on first use
    bool pf_order_built = false;
    DevBuf pf_order;
    bool pfree_t_ok = false;                 // every P column's skeleton slot written once
    // transposed copy of the stored blocks for the adjoint / transpose
    // (variant 8 of the transposed tables): per level, each block n x m
    // column-major at its own offset; levels whose transposed product is
    // matrix-free are not copied.  Mode 0 off, 1 built on first use when
    // device memory allows.
    std::vector<DevBuf> factors_t;
    int32_t tcopy_mode = 1;
    bool tcopy_built = false, tcopy_tried = false;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<int64_t> graph_key;  // op, nrhs, x, ldx, y, ldy, alpha, beta, stream
    ~BF();
    DevBuf perm_test, perm_trial;
    DevBuf slots;                 // slot buffer (slot_len * nrhs)
    size_t slots_cap = 0;
    DevBuf xy_host_stage;         // staging for host x / y
    mutable std::vector<std::vector<double>> export_cache;  // host copy of factors (export)
    int64_t n_comp = 0, n_retry = 0;
    int64_t n_judged = 0, n_fallback = 0;  // learned ranks: judged compressions, plain fallbacks
    int64_t n_blocks = 0, n_data = 0;  // all blocks, factor entries (set by pack_tables)
    double t_factor_ms = 0.0;
    double t_plan_ms = 0.0, t_sample_ms = 0.0, t_qr_ms = 0.0, t_interp_ms = 0.0, t_pack_ms = 0.0;
};

// Most rows of a matrix-core tile (4 blocks of 16).
constexpr int32_t MTILE_ROWS = 64;

// Most tasks per dispatch of the kernels that give every task a workgroup
// (a dispatch holds fewer than 2^32 work-items in x).
constexpr int32_t TASK_CHUNK = 1 << 23;

// Entries allocated past the end of every factor buffer: the dense kernel
// reads two rows with one load, one entry beyond an odd-height block.
constexpr int64_t FACTOR_PAD = 16;

// Packing of reference-layout blocks into launch tables (matvec.hip).
void pack_tables(BF &bf);
// Fused-sweep tables, built on first use (matvec.hip).
void build_sweep(BF &bf);
// Order of the matrix-free last level's items for its matrix-core kernel
// (matvec.hip), built on first use.
void build_pf_order(BF &bf);
// Streaming-kernel tables of a forward table (matvec.hip); es: entry bytes.
void build_stream(StageTable &S, int es, int nld, int64_t factor_elems, bool over_copy = false);

// Host worker threads (OMP_NUM_THREADS caps the hardware count).
int host_threads();

// parallel_for over [0, n) in chunks of `chunk`; the first exception thrown
// by any worker is rethrown on the caller.
template <class F> void parallel_for(int64_t n, F &&f, int64_t chunk = 64) {
    int nt = int(std::min<int64_t>(host_threads(), (n + chunk - 1) / chunk));
    if (nt <= 1) {
        for (int64_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int64_t> next{0};
    std::exception_ptr err;
    std::mutex mu;
    auto work = [&]() {
        try {
            for (;;) {
                int64_t b = next.fetch_add(chunk);
                if (b >= n) break;
                int64_t e = std::min(n, b + chunk);
                for (int64_t i = b; i < e; ++i) f(i);
            }
        } catch (...) {
            std::lock_guard<std::mutex> g(mu);
            if (!err) err = std::current_exception();
            next.store(n);
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work);
    work();
    for (auto &t : th) t.join();
    if (err) std::rethrow_exception(err);
}

// Kernel launchers (kernels.hip).
struct FillTask {       // one kernel-matrix block
    int64_t zofs;       // element offset of the column-major block
    int32_t m, n;
    int32_t rofs, cofs; // offsets into the row / column point-index lists
};
// EFIE / RWG plug-in (efie.hip)
struct KernelRWG;
const KernelRWG *kernel_of(const bfgpu_kernel *h);
int64_t kernel_functions(const KernelRWG &k, bool test);
bool kernel_has_near(const KernelRWG &k);
void launch_fill_rwg(hipStream_t s, int prec, const KernelRWG &kr, const int32_t *rows, const int32_t *cols,
                     const FillTask *tasks, const int64_t *task_end, int32_t n_tasks, int64_t total, void *Z,
                     double k);
// the plug-in's near-field quadrature (singularity extraction), same tasks
void launch_fill_rwg_near(hipStream_t s, int prec, const KernelRWG &kr, const int32_t *rows, const int32_t *cols,
                          const FillTask *tasks, const int64_t *task_end, int32_t n_tasks, int64_t total, void *Z,
                          double k);
void launch_fill(hipStream_t s, int prec, const double *obs_pts, const double *src_pts,
                 const int32_t *rows, const int32_t *cols, const FillTask *tasks,
                 const int64_t *task_end, int32_t n_tasks, int64_t total, void *Z, double k,
                 double self_re = 0.0, double self_im = 0.0);

// Row sample of one compression, drawn on the device (kernels.hip): the
// same sparse Fisher-Yates prefix over the SplitMix64 counter stream as the
// host's sample_prefix, of length fill[q].m (+ the held-out tile's rows),
// mapped through perm[obs0 ..] and written at the fill tasks' row offsets.
struct SampleTask {
    uint64_t key;           // the compression's sample stream
    int32_t obs0, n_obs;    // its observers: perm[obs0 .. obs0 + n_obs)
    int32_t held;           // its held-out tile among the batch's (-1: none)
    int32_t big;            // offset of host-drawn rows in `big` (-1: drawn on the device)
};
// Longest prefix drawn on the device (swap-map entries: four per lane).
constexpr int32_t SAMPLE_DEVICE_MAX = 256;
void launch_sample_rows(hipStream_t s, const SampleTask *tasks, const FillTask *fill, int32_t n_tasks,
                        const int32_t *perm, const int32_t *big, int32_t *rows);

struct QrTask {
    int64_t zofs;       // Z block (m x n, ld = m)
    int32_t m, n;
    int32_t pofs;       // pivot array offset (n entries)
    int32_t pad;
};
// Skeletons of a batch (kernels.hip): for compression q, its first
// skofs[q + 1] - skofs[q] pivots mapped to column ids,
// out[skofs[q] + i] = colpool[fill[q].cofs + piv[qr[q].pofs + i]].
void launch_skeleton_gather(hipStream_t s, const QrTask *qr, const FillTask *fill, const int32_t *colpool,
                            const int32_t *piv, const int32_t *skofs, int32_t n, int32_t *out);
// Blocks the wavefront QR kernel takes (the rest: a workgroup per block).
bool cpqr_wave_ok(int prec, int32_t m, int32_t n);
void launch_cpqr(hipStream_t s, int prec, void *Z, const QrTask *tasks, int32_t n_tasks,
                 int32_t *piv, int32_t *rank, double rtol, int32_t max_n, int32_t max_m,
                 double *tau = nullptr);

struct InterpTask {
    int64_t zofs;       // R (in Z workspace, ld = m)
    int64_t tofs;       // T output (rank x n, column-major) in the factor buffer
    int32_t m, n, rank, pofs;
};
void launch_interp(hipStream_t s, int prec, const void *Z, const InterpTask *tasks,
                   int32_t n_tasks, const int32_t *piv, void *T, int32_t max_rank);
// Compressions of rank at most INTERP_SMALL_RANK: a lane per column, the
// same back substitution in registers (kernels.hip).
constexpr int32_t INTERP_SMALL_RANK = 16;
void launch_interp_small(hipStream_t s, int prec, const void *Z, const InterpTask *tasks, int32_t n_tasks,
                         const int32_t *piv, void *T);

// Residual judge (kernels.hip) of held-out tiles t = 0 .. n_tasks-1 of one
// batch: compression q = jq[t] has R of its pivoted QR at qr[q].zofs
// (leading dimension qr[q].m, columns in pivot order, rank[q] steps taken,
// pivots at qr[q].pofs) and its held[t].m held-out rows at held[t].zofs
// (held[t].m x qr[q].n, column-major, columns in the original order).
// Most columns of a judged compression (the judge kernel keeps four
// 64-lane chunks of a held-out row in registers).
constexpr int32_t JUDGE_MAX_COLS = 256;
// out[2 t], out[2 t + 1] = (e_max, z_max): the largest column norm of the
// interpolative decomposition's residual on the held-out rows, and of the
// held-out tile itself.
void launch_judge(hipStream_t s, int prec, const void *Z, const int32_t *jq, const QrTask *qr,
                  const FillTask *held, const int32_t *rank, int32_t n_tasks, const int32_t *piv, double *out);

void launch_stage(hipStream_t s, int prec, int op, const StageTable &t, const void *factors,
                  const void *x, int64_t ldx, void *y, int64_t ldy, void *slots,
                  int64_t slot_len, const int32_t *perm_in, const int32_t *perm_out,
                  int64_t nrhs, const double *alpha, const double *beta, bool fuse_beta);
void launch_tile(hipStream_t s, int prec, int op, const StageTable &t, const void *factors, void *slots,
                 int64_t slot_len, int64_t in_shift, void *y, int64_t ldy, const int32_t *perm_out,
                 int64_t nrhs, const double *alpha, const double *beta, bool fuse_beta, int n_cu);
// Forward dense kernel over table t; with swapped true it reads the
// swapped task tables (t.tasks_sw, t.contribs_sw) and `factors` is the
// transposed copy; conj conjugates the factor entries.
// tab: the table's own tasks (DENSE_TASKS), the tasks over the transposed
// copy (DENSE_SWAPPED), or the tasks the streaming kernel leaves (DENSE_REST)
enum DenseTab { DENSE_TASKS = 0, DENSE_SWAPPED = 1, DENSE_REST = 2 };
void launch_dense(hipStream_t s, int prec, const StageTable &t, const void *factors, void *slots,
                  int64_t slot_len, int64_t in_shift, void *y, int64_t ldy, const int32_t *perm_out,
                  int64_t nrhs, const double *alpha, const double *beta, bool fuse_beta, int n_cu,
                  int tab = DENSE_TASKS, int conj = 0);
// Streaming forward kernel over the chunks of a table (one RHS); the
// table's other tasks are not launched here.
void launch_stream(hipStream_t s, int prec, const StageTable &t, const void *factors, void *slots,
                   int64_t in_shift, int n_cu, bool conj = false);
// Device tasks (MvTask and MvTaskC) of a table from its uploaded items and
// contributions.
void launch_build_tasks(hipStream_t s, const MvItem *items, int32_t n_items, const MvContrib *con, int32_t n_con,
                        MvTask *tasks, MvTaskC *tc);
// Block transposition: for block q, dst[dofs + i n + k] = src[sofs + k m + i].
struct TBlock {
    int64_t sofs, dofs;
    int32_t m, n;
};
void launch_transpose_blocks(hipStream_t s, int prec, const TBlock *blocks, int64_t n_blocks, const void *src,
                             void *dst);
// Transposed copy of the stored blocks (built once when mode and device
// memory allow; true when built) and its release.
bool ensure_tcopy(BF &bf);
void release_tcopy(BF &bf);
void launch_denset(hipStream_t s, int prec, int op, const StageTable &t, const void *factors, void *slots,
                   int64_t slot_len, int64_t in_shift, void *y, int64_t ldy, const int32_t *perm_out,
                   int64_t nrhs, const double *alpha, const double *beta, bool fuse_beta, int n_cu);
void launch_rhs(hipStream_t s, int prec, const StageTable &t, const void *factors, void *slots, int32_t ns,
                int64_t in_shift, void *y, int64_t ldy, const int32_t *perm_out, int nc, const double *alpha,
                const double *beta, bool fuse, int n_cu);
void launch_mfma_rhs(hipStream_t s, const StageTable &t, const void *factors, void *slots, int32_t ns,
                     int64_t in_shift, void *y, int64_t ldy, const int32_t *perm_out, int nc, const double *alpha,
                     const double *beta, bool fuse, int n_cu);
// Matrix-free last level (kernels.hip): y rows of every observer leaf from
// the P columns, entries computed from the points.
// Matrix-free last level with right-hand sides across lanes (ns = 16, 32,
// 48 or 64 interleaved slot columns, nc of them used) on the matrix cores.
void launch_pfree_mfma(hipStream_t s, const BF &bf, const void *slots, int32_t ns, void *y, int64_t ldy, int nc,
                       const double *alpha, const double *beta, bool fuse, int n_cu);
void launch_pfree(hipStream_t s, const BF &bf, const void *slots, void *y, int64_t ldy, const double *alpha,
                  const double *beta, bool fuse);
void launch_pfree_t(hipStream_t s, const BF &bf, const void *x, void *slots, bool conj);
// Matrix-core tiles of a forward table, at most `rows` rows each (matvec.hip).
void build_mfma(StageTable &t, int32_t rows);
void launch_gather_rhs(hipStream_t s, int prec, const void *x, int64_t ldx, const int32_t *perm, int32_t lo,
                       int32_t hi, void *dst, int32_t ns, int nc);
void launch_gather(hipStream_t s, int prec, const void *x, int64_t ldx, const int32_t *perm,
                   int32_t lo, int32_t hi, void *dst, int64_t ld_dst, int64_t nrhs);
// Fused forward sweep: every stage in one cooperative launch, tasks ordered
// by stage, dependencies through per-group counters.
void launch_sweep(hipStream_t s, int prec, const SweepTables &sw, const std::vector<DevBuf> &factors,
                  const void *x, int64_t ldx, void *y, int64_t ldy, void *slots, int64_t slot_len,
                  const int32_t *perm_in, const int32_t *perm_out, int64_t nrhs, const double *alpha,
                  const double *beta, bool fuse_beta, int n_cu);
void launch_zero_i32(hipStream_t s, int32_t *p, int64_t n);

void launch_scale(hipStream_t s, int prec, void *y, int64_t n, int64_t ldy, int64_t nrhs,
                  const double *beta);
void launch_convert(hipStream_t s, const void *in, int in_prec, void *out, int out_prec,
                    int64_t n);

// features of one compression for the learned rank predictor
void compression_features(const Tree &test, const Tree &trial, double k, int32_t S, int32_t O,
                          int32_t stage, double eps, float *f);

}  // namespace bfg

// opaque handles of the C ABI
struct bfgpu_ctx { bfg::Ctx c; };
struct bfgpu_tree { bfg::Tree t; };
struct bfgpu_bf { bfg::BF b; };
