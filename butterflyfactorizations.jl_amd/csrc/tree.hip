// tree.hip -- cluster trees, level lists and interaction lists (host code).
#include <algorithm>
#include <cmath>
#include <functional>

#include "internal.hpp"

namespace bfg {

// Build a 2^dims-ary box tree breadth first.  Children of a node are created
// in child-code order (bit d set when coordinate d >= box centre) and empty
// children are skipped; the node's point range is stably partitioned by
// child code so that every node owns a contiguous slice of perm.
static Tree build_tree(const double *xyz, int64_t n, int dims, int leaf, int mode) {
    BFG_REQUIRE(n > 0 && n < (int64_t(1) << 31), "tree: point count out of range");
    BFG_REQUIRE(dims == 2 || dims == 3, "tree: dims must be 2 or 3");
    BFG_REQUIRE(leaf >= 1, "tree: leaf_size must be >= 1");
    Tree t;
    t.n_values = int32_t(n);
    t.points.assign(xyz, xyz + 3 * n);
    t.perm.resize(n);
    for (int64_t i = 0; i < n; ++i) t.perm[i] = int32_t(i);
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) { lo[d] = xyz[d]; hi[d] = xyz[d]; }
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], xyz[3 * i + d]);
            hi[d] = std::max(hi[d], xyz[3 * i + d]);
        }
    double half = 0.0, c0[3];
    for (int d = 0; d < 3; ++d) {
        c0[d] = 0.5 * (lo[d] + hi[d]);
        if (d < dims) half = std::max(half, 0.5 * (hi[d] - lo[d]));
    }
    if (half <= 0.0) half = 1.0;
    half *= 1.0 + 1e-12;
    int nchild = 1 << dims;
    int32_t depth_max = 0;
    if (mode == BFGPU_TREE_BALANCED) {
        double cnt = double(n);
        while (cnt > leaf) { cnt /= nchild; ++depth_max; }
    } else {
        depth_max = 40;
    }
    auto push = [&](int32_t par, const double *c, double h, int32_t b, int32_t e, int32_t lv) {
        t.parent.push_back(par);
        t.first_child.push_back(0);
        t.n_children.push_back(0);
        t.val_begin.push_back(b);
        t.val_end.push_back(e);
        t.center.insert(t.center.end(), c, c + 3);
        t.halfsize.push_back(h);
        t.level.push_back(lv);
    };
    push(-1, c0, half, 0, int32_t(n), 0);
    std::vector<int32_t> tmp;
    std::vector<uint8_t> code;
    for (size_t v = 0; v < t.parent.size(); ++v) {
        int32_t b = t.val_begin[v], e = t.val_end[v], lv = t.level[v];
        bool split = (mode == BFGPU_TREE_BALANCED) ? (lv < depth_max) : (e - b > leaf && lv < depth_max);
        if (!split || e - b == 0) continue;
        double c[3] = {t.center[3 * v], t.center[3 * v + 1], t.center[3 * v + 2]};
        double h = t.halfsize[v];
        int32_t cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        code.resize(e - b);
        for (int32_t i = b; i < e; ++i) {
            const double *p = xyz + 3 * int64_t(t.perm[i]);
            int cc = 0;
            for (int d = 0; d < dims; ++d)
                if (p[d] >= c[d]) cc |= 1 << d;
            code[i - b] = uint8_t(cc);
            cnt[cc]++;
        }
        int32_t start[8], s = b;
        for (int q = 0; q < nchild; ++q) { start[q] = s; s += cnt[q]; }
        tmp.assign(t.perm.begin() + b, t.perm.begin() + e);
        int32_t pos[8];
        std::copy(start, start + 8, pos);
        for (int32_t i = b; i < e; ++i) t.perm[pos[code[i - b]]++] = tmp[i - b];
        int32_t first = int32_t(t.parent.size()), nc = 0;
        for (int q = 0; q < nchild; ++q) {
            if (cnt[q] == 0) continue;
            double cc[3] = {c[0], c[1], c[2]};
            for (int d = 0; d < dims; ++d) cc[d] += ((q >> d) & 1) ? 0.5 * h : -0.5 * h;
            push(int32_t(v), cc, 0.5 * h, start[q], start[q] + cnt[q], lv + 1);
            ++nc;
        }
        t.first_child[v] = first;
        t.n_children[v] = nc;
    }
    t.n_nodes = int32_t(t.parent.size());
    return t;
}

std::vector<std::vector<int32_t>> traverse_and_pad(const Tree &t, int32_t root) {
    std::vector<std::vector<int32_t>> lv;
    std::vector<int32_t> cur{root};
    while (!cur.empty()) {
        lv.push_back(cur);
        std::vector<int32_t> nxt;
        for (int32_t v : cur)
            for (int32_t c = 0; c < t.n_children[v]; ++c) nxt.push_back(t.first_child[v] + c);
        cur.swap(nxt);
    }
    for (size_t l = 1; l + 1 < lv.size(); ++l)
        for (size_t i = 0; i < lv[l].size(); ++i)
            if (t.isleaf(lv[l][i])) lv[l + 1].push_back(lv[l][i]);
    return lv;
}

int32_t tree_height(const Tree &t) { return int32_t(traverse_and_pad(t, 0).size()); }

}  // namespace bfg

using namespace bfg;


extern "C" int bfgpu_tree_build(const double *xyz, int64_t n, int32_t dims, int32_t leaf_size,
                                int32_t mode, bfgpu_tree **out) {
    try {
        BFG_REQUIRE(xyz && out, "tree_build: null argument");
        BFG_REQUIRE(mode == BFGPU_TREE_BALANCED || mode == BFGPU_TREE_MAXPOINTS,
                    "tree_build: unknown mode");
        auto *h = new bfgpu_tree{build_tree(xyz, n, dims, leaf_size, mode)};
        *out = h;
        return BFGPU_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return BFGPU_ERR_NOMEM;
    }
}

extern "C" int bfgpu_tree_create(const double *xyz, int32_t n_values, int32_t n_nodes,
                                 const int32_t *parent, const int32_t *first_child,
                                 const int32_t *n_children, const int32_t *val_begin,
                                 const int32_t *val_end, const double *center,
                                 const double *halfsize, const int32_t *perm, bfgpu_tree **out) {
    try {
        BFG_REQUIRE(xyz && parent && first_child && n_children && val_begin && val_end &&
                        center && halfsize && perm && out,
                    "tree_create: null argument");
        BFG_REQUIRE(n_values > 0 && n_nodes > 0, "tree_create: empty tree");
        Tree t;
        t.n_values = n_values;
        t.n_nodes = n_nodes;
        t.points.assign(xyz, xyz + 3 * int64_t(n_values));
        t.parent.assign(parent, parent + n_nodes);
        t.first_child.assign(first_child, first_child + n_nodes);
        t.n_children.assign(n_children, n_children + n_nodes);
        t.val_begin.assign(val_begin, val_begin + n_nodes);
        t.val_end.assign(val_end, val_end + n_nodes);
        t.center.assign(center, center + 3 * int64_t(n_nodes));
        t.halfsize.assign(halfsize, halfsize + n_nodes);
        t.perm.assign(perm, perm + n_values);
        t.level.assign(n_nodes, 0);
        BFG_REQUIRE(t.parent[0] == -1, "tree_create: node 0 must be the root");
        for (int32_t v = 0; v < n_nodes; ++v) {
            BFG_REQUIRE(t.val_begin[v] >= 0 && t.val_end[v] <= n_values &&
                            t.val_begin[v] <= t.val_end[v],
                        "tree_create: value range out of bounds");
            BFG_REQUIRE(t.n_children[v] >= 0 &&
                            (t.n_children[v] == 0 ||
                             (t.first_child[v] > v && t.first_child[v] + t.n_children[v] <= n_nodes)),
                        "tree_create: children must be later contiguous ids");
            for (int32_t c = 0; c < t.n_children[v]; ++c) {
                int32_t ch = t.first_child[v] + c;
                BFG_REQUIRE(t.parent[ch] == v, "tree_create: parent/children mismatch");
                t.level[ch] = t.level[v] + 1;
            }
        }
        for (int32_t i = 0; i < n_values; ++i)
            BFG_REQUIRE(t.perm[i] >= 0 && t.perm[i] < n_values, "tree_create: perm out of range");
        *out = new bfgpu_tree{std::move(t)};
        return BFGPU_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return BFGPU_ERR_NOMEM;
    }
}

extern "C" int bfgpu_tree_view_get(const bfgpu_tree *h, bfgpu_tree_view *o) {
    if (!h || !o) {
        set_last_error("tree_view_get: null argument");
        return BFGPU_ERR_ARG;
    }
    const Tree &t = h->t;
    o->n_nodes = t.n_nodes;
    o->n_values = t.n_values;
    o->parent = t.parent.data();
    o->first_child = t.first_child.data();
    o->n_children = t.n_children.data();
    o->val_begin = t.val_begin.data();
    o->val_end = t.val_end.data();
    o->center = t.center.data();
    o->halfsize = t.halfsize.data();
    o->perm = t.perm.data();
    o->level = t.level.data();
    o->points = t.points.data();
    return BFGPU_OK;
}

extern "C" int bfgpu_tree_destroy(bfgpu_tree *h) {
    delete h;
    return BFGPU_OK;
}

// ---------------------------------------------------------------------------
// Interaction lists: nearandfar / process_nodes! (intlists.jl:30-217) with
// isFarFunctor on boxes (h2admissible.jl:11-40) or the centre-distance rule
// (h2admissible.jl:74-91).
// ---------------------------------------------------------------------------
namespace {
struct NF {
    const Tree *s, *t;
    int32_t kind;
    double param;
    int32_t minbflvl, unbal, leafcomp, leafimb;
    int32_t *far, *near;
    int64_t nfar = 0, nnear = 0;
};

bool admissible(const NF &c, int32_t o, int32_t s) {
    const double *oc = &c.t->center[3 * o], *sc = &c.s->center[3 * s];
    double ol = c.t->radius(o), sl = c.s->radius(s);
    if (c.kind == 0) {
        double W = std::max(sl, ol), mind = 0.0;
        for (int i = 0; i < 3; ++i) {
            double d = std::fabs(oc[i] - sc[i]) - (sl + ol);
            if (d > 0.0) mind += d * d;
        }
        if (mind < 1e-10) return false;
        return std::sqrt(mind) > c.param * W;
    }
    double d = std::sqrt((sc[0] - oc[0]) * (sc[0] - oc[0]) + (sc[1] - oc[1]) * (sc[1] - oc[1]) +
                         (sc[2] - oc[2]) * (sc[2] - oc[2]));
    if (d < 1e-5) return false;
    return d > c.param * (sl + ol);
}

void push(int32_t *buf, int64_t &n, int32_t o, int32_t s) {
    if (buf) { buf[2 * n] = o; buf[2 * n + 1] = s; }
    ++n;
}

void process(NF &c, int32_t o, int32_t s, int32_t h) {
    bool lo = c.t->isleaf(o), ls = c.s->isleaf(s);
    bool any = lo || ls, both = lo && ls;
    bool adm = admissible(c, o, s);
    if (h <= c.minbflvl && !c.leafcomp) { push(c.near, c.nnear, o, s); return; }
    if (adm) {
        if (any && !c.leafcomp) push(c.near, c.nnear, o, s);
        else push(c.far, c.nfar, o, s);
        return;
    }
    if (both || (any && !c.leafimb)) { push(c.near, c.nnear, o, s); return; }
    auto split_o = [&]() {
        for (int32_t k = 0; k < c.t->n_children[o]; ++k) process(c, c.t->first_child[o] + k, s, h - 1);
    };
    auto split_s = [&]() {
        for (int32_t k = 0; k < c.s->n_children[s]; ++k) process(c, o, c.s->first_child[s] + k, h - 1);
    };
    if (c.unbal) {
        bool larger = c.t->radius(o) >= c.s->radius(s);
        if ((larger && !lo) || ls) split_o();
        else split_s();
    } else if (lo) {
        split_s();
    } else if (ls) {
        split_o();
    } else {
        for (int32_t a = 0; a < c.t->n_children[o]; ++a)
            for (int32_t b = 0; b < c.s->n_children[s]; ++b)
                process(c, c.t->first_child[o] + a, c.s->first_child[s] + b, h - 1);
    }
}
}  // namespace

extern "C" int bfgpu_nearandfar(const bfgpu_tree *test, const bfgpu_tree *trial, int32_t kind,
                                double param, int32_t minbflvl, int32_t unbalancedints,
                                int32_t leafcomp, int32_t leafimbalance, int32_t *far,
                                int64_t *nfar, int32_t *near, int64_t *nnear) {
    try {
        BFG_REQUIRE(test && trial && nfar && nnear, "nearandfar: null argument");
        BFG_REQUIRE(kind == 0 || kind == 1, "nearandfar: kind must be 0 or 1");
        NF c{&trial->t, &test->t, kind, param, minbflvl, unbalancedints, leafcomp, leafimbalance,
             far, near};
        int32_t h = std::min(tree_height(test->t), tree_height(trial->t));
        process(c, 0, 0, h);
        *nfar = c.nfar;
        *nnear = c.nnear;
        return BFGPU_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    }
}
