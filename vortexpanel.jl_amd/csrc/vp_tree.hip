// Fast summation of the free wake's blob sums (blob_velocities, src/freewake/step.jl:176-187, with blob_velocity,
// src/kernels/blob.jl:21-25).  The reference sums directly and defers fast summation (docs/claude_handoff.md:177);
// this is an opt-in approximation beside the direct sums of vp_field.hip, not a replacement for them.
//
// Method: a cluster-particle treecode with barycentric Lagrange interpolation on the target side (the cluster-particle
// form of Wang, Krasny & Tlupova, "A kernel-independent treecode based on barycentric Lagrange interpolation",
// Commun. Comput. Phys. 28, 2020; the barycentric formula of Berrut & Trefethen, SIAM Review 46, 2004).  The
// particles carry their own core radii (the free wake's cores differ from particle to particle), so sources are
// never merged into expansions: every source keeps its exact kernel.  Instead, the field of a source cluster S far
// from a target cluster T is smooth over T's box, so it is evaluated only at the (n+1)^3 Chebyshev points of the
// second kind of that box (the proxies) and interpolated to T's targets.  A pair (T, S) is well separated when
// r_T <= theta * dist(c_T, box_S): every source of S is at least r_T / theta from T's centre.
//
// Both point sets are sorted along a Morton curve on the device (63-bit codes, hipCUB's radix sort, the points gathered
// in sorted order); on the host (C++) each is split top-down at the highest differing code bit of the sorted codes
// (binary radix tree), to leaves of at most 64 points, whose tight boxes the device computes; the host unites them up
// the tree and pads every box so that no side is thinner than 1% of the longest.  A dual traversal from (root, root)
// sorts every (target, source) pair into exactly one of: a far pair (T with more targets than proxies, interpolated)
// or the direct sums of T's leaves (each leaf's near sources are the complement of its ancestors' far sources).  Device: tree_proxy evaluates the far
// pairs at the proxies (a block per work item of about 2k sources of one target cluster, one proxy per lane),
// tree_proxy_reduce adds each cluster's items in order, tree_near sums a target leaf's near sources directly in work
// items of at most 4k sources (a block of four waves each), and tree_finish takes one target leaf per block: its
// items' sums in order, then the interpolants of its far ancestors.
// Results do not depend on scheduling (no atomics; fixed summation orders).  precision 1 evaluates the kernel in FP32
// relative to a nearby origin (a far cluster's centre, a leaf's first target), as the direct FP32 sums do; the
// interpolation and every running sum stay FP64.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "vp_capi_internal.hpp"
#include "vp_device.hpp"

namespace vp {

namespace {

constexpr int TREE_LEAF = 64;              // points per leaf: a target leaf is one wave
constexpr int TREE_MAX_DEGREE = 7;         // (n+1)^3 <= 512 proxies, one per lane of a proxy block
constexpr long long ITEM_SOURCES = 2048;   // sources per proxy work item
constexpr int NEAR_ITEM = 4096;            // near sources per near work item
constexpr double BOX_MIN_SIDE = 1e-2;      // a box side is at least this fraction of its longest side

#define TCHECK(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct TNode {
    int a, b;        // points [a, b) of the sorted order
    int c0, c1;      // children (-1: leaf)
    int parent;
    double lo[3], hi[3];  // padded box
    double rad;           // half diagonal of the padded box
    bool ok;              // the box has a positive extent (a cluster of identical points is never interpolated)
};

struct Tree {
    std::vector<TNode> nodes;  // nodes[0] is the root
    std::vector<int> leaves;   // leaf node ids by increasing a
};

__host__ __device__ inline uint64_t spread21(uint64_t v)  // bits 0..20 of v to bits 0, 3, 6, ..., 60
{
    v &= 0x1fffffull;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

// The Morton frame of a point set: its bounding box's low corner and 2^21 - 1 cells over its longest side
struct Frame {
    double lo[3];
    double scale;
};

Frame frame_of(const double *x, int n)
{
    Frame f{{0, 0, 0}, 1.0};
    if (n <= 0) return f;
    double lo[3] = {x[0], x[1], x[2]}, hi[3] = {x[0], x[1], x[2]};
    bool finite = true;
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            const double v = x[3 * (size_t)i + d];
            finite = finite && std::isfinite(v);
            lo[d] = std::min(lo[d], v);
            hi[d] = std::max(hi[d], v);
        }
    if (!finite) throw std::invalid_argument("non-finite point coordinates");
    double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
    if (!(ext > 0)) ext = 1.0;
    for (int d = 0; d < 3; d++) f.lo[d] = lo[d];
    f.scale = 2097151.0 / ext;
    return f;
}

__host__ __device__ inline uint64_t morton(const double *p, const Frame &f)
{
    uint64_t c = 0;
    for (int d = 0; d < 3; d++) {
        double q = (p[d] - f.lo[d]) * f.scale;
        q = q < 0.0 ? 0.0 : (q > 2097151.0 ? 2097151.0 : q);
        c |= spread21((uint64_t)q) << (2 - d);
    }
    return c;
}

// LSD radix sort of (key, index) by 11-bit digits (six passes cover the 63-bit codes; the histograms of every pass are
// counted in one read, and a pass whose digit is the same for every key is skipped)
void radix_sort(std::vector<uint64_t> &key, std::vector<int> &idx)
{
    constexpr int BITS = 11, NB = 1 << BITS, PASSES = 6;
    const size_t n = key.size();
    std::vector<uint64_t> k2(n);
    std::vector<int> i2(n);
    std::vector<uint32_t> cnt((size_t)PASSES * NB, 0);
    for (size_t i = 0; i < n; i++)
        for (int p = 0; p < PASSES; p++) cnt[(size_t)p * NB + ((key[i] >> (BITS * p)) & (NB - 1))]++;
    for (int p = 0; p < PASSES; p++) {
        uint32_t *c = cnt.data() + (size_t)p * NB;
        bool trivial = false;
        uint32_t sum = 0;
        for (int d = 0; d < NB; d++) {
            if (c[d] == n) trivial = true;
            const uint32_t t = c[d];
            c[d] = sum;
            sum += t;
        }
        if (trivial) continue;
        const int sh = BITS * p;
        for (size_t i = 0; i < n; i++) {
            const uint32_t q = c[(key[i] >> sh) & (NB - 1)]++;
            k2[q] = key[i];
            i2[q] = idx[i];
        }
        key.swap(k2);
        idx.swap(i2);
    }
}

// Topology from the sorted codes: every node's points, children and parent, and the leaves by increasing a (the
// boxes are set by tree_boxes)
Tree tree_topology(const uint64_t *code, int n)
{
    Tree t;
    if (n <= 0) return t;
    t.nodes.push_back(TNode{0, n, -1, -1, -1, {}, {}, 0.0, false});
    std::vector<int> stack{0};
    while (!stack.empty()) {  // depth first, the lower half first, so leaves come out by increasing a
        const int id = stack.back();
        stack.pop_back();
        const int a = t.nodes[id].a, b = t.nodes[id].b;
        if (b - a <= TREE_LEAF) {
            t.leaves.push_back(id);
            continue;
        }
        int s;
        const uint64_t ca = code[a], cb = code[b - 1];
        if (ca == cb) {
            s = a + (b - a) / 2;
        } else {
            const int bit = 63 - __builtin_clzll(ca ^ cb);
            const uint64_t key = ((ca >> bit) | 1ull) << bit;  // the shared prefix, then this bit set
            s = (int)(std::lower_bound(code + a, code + b, key) - code);
        }
        const int c0 = (int)t.nodes.size();
        t.nodes[id].c0 = c0;
        t.nodes[id].c1 = c0 + 1;
        t.nodes.push_back(TNode{a, s, -1, -1, id, {}, {}, 0.0, false});
        t.nodes.push_back(TNode{s, b, -1, -1, id, {}, {}, 0.0, false});
        stack.push_back(c0 + 1);
        stack.push_back(c0);
    }
    return t;
}

// Boxes from the leaves' tight boxes (6 per leaf in t.leaves' order: lo, hi): every other node's from its children,
// then every box padded to a minimum side, and its radius.  (Every box holds its own points; a padded child box may
// stick out of its padded parent, which nothing needs.)
void tree_boxes(Tree &t, const double *leaf_box)
{
    for (size_t l = 0; l < t.leaves.size(); l++) {
        TNode &q = t.nodes[t.leaves[l]];
        for (int d = 0; d < 3; d++) {
            q.lo[d] = leaf_box[6 * l + d];
            q.hi[d] = leaf_box[6 * l + 3 + d];
        }
    }
    for (int id = (int)t.nodes.size() - 1; id >= 0; id--) {  // children are created after their parent
        TNode &q = t.nodes[id];
        if (q.c0 < 0) continue;
        const TNode &u = t.nodes[q.c0], &v = t.nodes[q.c1];
        for (int d = 0; d < 3; d++) {
            q.lo[d] = std::min(u.lo[d], v.lo[d]);
            q.hi[d] = std::max(u.hi[d], v.hi[d]);
        }
    }
    for (TNode &q : t.nodes) {
        double hw[3], m = 0.0;
        for (int d = 0; d < 3; d++) m = std::max(m, hw[d] = 0.5 * (q.hi[d] - q.lo[d]));
        q.ok = m > 0;
        for (int d = 0; d < 3; d++) {
            const double c = 0.5 * (q.lo[d] + q.hi[d]);
            hw[d] = std::max(hw[d], BOX_MIN_SIDE * m);
            q.lo[d] = c - hw[d];
            q.hi[d] = c + hw[d];
        }
        q.rad = std::sqrt(hw[0] * hw[0] + hw[1] * hw[1] + hw[2] * hw[2]);
    }
}

// The leaves' tight boxes from the points in sorted order (pos, 3 per point), 6 per leaf
std::vector<double> leaf_boxes_host(const Tree &t, const double *pos)
{
    std::vector<double> box(6 * t.leaves.size());
    for (size_t l = 0; l < t.leaves.size(); l++) {
        const TNode &q = t.nodes[t.leaves[l]];
        double *o = &box[6 * l];
        for (int d = 0; d < 3; d++) o[d] = o[3 + d] = pos[3 * (size_t)q.a + d];
        for (int i = q.a + 1; i < q.b; i++)
            for (int d = 0; d < 3; d++) {
                o[d] = std::min(o[d], pos[3 * (size_t)i + d]);
                o[3 + d] = std::max(o[3 + d], pos[3 * (size_t)i + d]);
            }
    }
    return box;
}

// The same tree sorted on the host (the coverage check, which runs without a device)
Tree build_tree(const double *x, int n)
{
    if (n <= 0) return Tree{};
    const Frame f = frame_of(x, n);
    std::vector<uint64_t> code(n);
    std::vector<int> perm(n);
    for (int i = 0; i < n; i++) {
        code[i] = morton(x + 3 * (size_t)i, f);
        perm[i] = i;
    }
    radix_sort(code, perm);
    std::vector<double> pos((size_t)n * 3);
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) pos[3 * (size_t)i + d] = x[3 * (size_t)perm[i] + d];
    Tree t = tree_topology(code.data(), n);
    tree_boxes(t, leaf_boxes_host(t, pos.data()).data());
    return t;
}

double box_dist(const double *c, const TNode &s)  // distance from the point c to the box of s
{
    double d2 = 0.0;
    for (int d = 0; d < 3; d++) {
        const double e = c[d] < s.lo[d] ? s.lo[d] - c[d] : (c[d] > s.hi[d] ? c[d] - s.hi[d] : 0.0);
        d2 += e * e;
    }
    return std::sqrt(d2);
}

struct Range {
    int a, b;
};

struct Lists {
    int P = 0;
    // far clusters: target node ids, their padded boxes, their source ranges (CSR), and the proxy work items
    std::vector<int> far_node;
    std::vector<double> far_box;  // 6 per far cluster: lo, hi
    std::vector<int> far_item_off;  // per far cluster: its items [off, off+1)
    std::vector<int> item_node, item_off;
    std::vector<Range> item_rng;
    // per target leaf: its points, near source ranges (CSR) and far ancestors (CSR)
    std::vector<int> leaf_a, leaf_b, near_off, anc_off, anc;
    std::vector<Range> near_rng;
    // each range's start within its group's concatenated sources (an item's, a leaf's near ones), and each group's
    // total: the kernels stage a group's sources in full tiles across its ranges
    std::vector<int> item_vst, item_tot, near_vst, near_tot;
    // near work items of at most NEAR_ITEM sources: (leaf, first range, v0, v1), and each leaf's items (CSR)
    std::vector<int> near_item, leaf_item_off;
    long long far_pairs = 0, near_pairs = 0, proxy_evals = 0, near_evals = 0, interp_evals = 0;
};

void virtual_starts(const std::vector<int> &off, const std::vector<Range> &rng, std::vector<int> &vst,
                    std::vector<int> &tot)
{
    vst.assign(rng.size(), 0);
    tot.assign(off.empty() ? 0 : off.size() - 1, 0);
    for (size_t g = 0; g + 1 < off.size(); g++) {
        int at = 0;
        for (int r = off[g]; r < off[g + 1]; r++) {
            vst[r] = at;
            at += rng[r].b - rng[r].a;
        }
        tot[g] = at;
    }
}

void merge_ranges(std::vector<Range> &r)  // sorted by start, touching ranges joined
{
    std::sort(r.begin(), r.end(), [](const Range &p, const Range &q) { return p.a < q.a; });
    size_t m = 0;
    for (size_t i = 0; i < r.size(); i++) {
        if (m && r[m - 1].b == r[i].a)
            r[m - 1].b = r[i].b;
        else
            r[m++] = r[i];
    }
    r.resize(m);
}

// f on a second thread while g runs on this one; an exception from either is rethrown on this thread
template <typename F, typename G> void run_both(F &&f, G &&g)
{
    std::exception_ptr ef;
    std::thread th([&] {
        try {
            f();
        } catch (...) {
            ef = std::current_exception();
        }
    });
    try {
        g();
    } catch (...) {
        th.join();
        throw;
    }
    th.join();
    if (ef) std::rethrow_exception(ef);
}

// f(begin, end) over [0, n) in contiguous chunks on up to 16 host threads; an exception is rethrown on this thread
template <typename F> void parallel_for(int n, F &&f)
{
    const int nth = std::max(1, std::min({16, (int)std::thread::hardware_concurrency(), n / 64}));
    if (nth == 1) {
        if (n > 0) f(0, n);
        return;
    }
    std::vector<std::exception_ptr> errs(nth);
    std::vector<std::thread> pool;
    auto run = [&](int t) {
        try {
            f((int)((long long)n * t / nth), (int)((long long)n * (t + 1) / nth));
        } catch (...) {
            errs[t] = std::current_exception();
        }
    };
    for (int t = 1; t < nth; t++) pool.emplace_back(run, t);
    run(0);
    for (auto &th : pool) th.join();
    for (auto &e : errs)
        if (e) std::rethrow_exception(e);
}

struct FarRec {  // a far pair: target cluster (node id) and the source cluster's points
    int t;
    Range s;
};

// One step of the dual traversal from (ti, si): a far pair is recorded; otherwise the larger cluster is split (target
// children with no more targets than proxies are dropped: their sources not taken by an ancestor are summed directly)
inline void traverse_step(const Tree &tt, const Tree &st, int P, double theta, int ti, int si,
                          std::vector<std::pair<int, int>> &next, std::vector<FarRec> &out)
{
    const TNode &T = tt.nodes[ti], &S = st.nodes[si];
    double c[3];
    for (int d = 0; d < 3; d++) c[d] = 0.5 * (T.lo[d] + T.hi[d]);
    if (T.ok && T.rad <= theta * box_dist(c, S)) {
        out.push_back(FarRec{ti, Range{S.a, S.b}});
        return;
    }
    const bool tleaf = T.c0 < 0, sleaf = S.c0 < 0;
    if (!tleaf && (sleaf || T.rad >= S.rad)) {
        for (int ch : {T.c1, T.c0})
            if (tt.nodes[ch].b - tt.nodes[ch].a > P) next.push_back({ch, si});
    } else if (!sleaf) {
        next.push_back({ti, S.c1});
        next.push_back({ti, S.c0});
    }
}

Lists build_lists(const Tree &tt, const Tree &st, double theta, int degree)
{
    Lists L;
    const int n1 = degree + 1;
    L.P = n1 * n1 * n1;
    const int nleaf = (int)tt.leaves.size();
    const int nb = st.nodes.empty() ? 0 : st.nodes[0].b;
    std::vector<std::vector<Range>> farl;
    std::vector<int> far_of_node(tt.nodes.size(), -1);
    // Far pairs only: a target cluster with no more targets than proxies never interpolates, so the traversal stops
    // there, and everything its ancestors do not take at their proxies is summed directly.  The traversal is expanded
    // breadth-first to a few hundred independent pairs, which host threads then finish depth-first; the far pairs are
    // grouped by target node and each cluster's ranges sorted, so the lists do not depend on the threads' schedule.
    std::vector<FarRec> recs;
    if (!tt.nodes.empty() && !st.nodes.empty() && tt.nodes[0].b - tt.nodes[0].a > L.P) {
        std::vector<std::pair<int, int>> front{{0, 0}}, kids;
        size_t head = 0;
        while (head < front.size() && front.size() - head < 512) {
            kids.clear();
            traverse_step(tt, st, L.P, theta, front[head].first, front[head].second, kids, recs);
            head++;
            front.insert(front.end(), kids.begin(), kids.end());
        }
        const size_t nwork = front.size() - head;
        const int nth = (int)std::max<size_t>(1, std::min<size_t>({16, std::thread::hardware_concurrency(), nwork / 8}));
        std::vector<std::vector<FarRec>> outs(nth);
        std::vector<std::exception_ptr> errs(nth);
        std::atomic<size_t> next{0};
        auto worker = [&](int t) {
            try {
                std::vector<std::pair<int, int>> stack;
                for (size_t w; (w = next.fetch_add(1)) < nwork;) {
                    stack.push_back(front[head + w]);
                    while (!stack.empty()) {
                        const auto [ti, si] = stack.back();
                        stack.pop_back();
                        traverse_step(tt, st, L.P, theta, ti, si, stack, outs[t]);
                    }
                }
            } catch (...) {
                errs[t] = std::current_exception();
                next.store(nwork);
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nth; t++) pool.emplace_back(worker, t);
        worker(0);
        for (auto &th : pool) th.join();
        for (auto &e : errs)
            if (e) std::rethrow_exception(e);
        for (auto &o : outs) recs.insert(recs.end(), o.begin(), o.end());
    }
    // far clusters in increasing node id, each with its source ranges (sorted by merge_ranges below)
    std::vector<int> cnt(tt.nodes.size() + 1, 0);
    for (const FarRec &r : recs) cnt[r.t + 1]++;
    for (size_t t = 0; t < tt.nodes.size(); t++)
        if (cnt[t + 1]) {
            far_of_node[t] = (int)L.far_node.size();
            L.far_node.push_back((int)t);
            farl.emplace_back();
            farl.back().reserve(cnt[t + 1]);
        }
    for (const FarRec &r : recs) {
        farl[far_of_node[r.t]].push_back(r.s);
        L.far_pairs++;
        L.proxy_evals += (long long)L.P * (r.s.b - r.s.a);
    }
    // far clusters: boxes and work items of about ITEM_SOURCES sources
    const int nfar = (int)L.far_node.size();
    L.far_item_off.push_back(0);
    L.item_off.push_back(0);
    parallel_for(nfar, [&](int k0, int k1) {
        for (int k = k0; k < k1; k++) merge_ranges(farl[k]);
    });
    for (int k = 0; k < nfar; k++) {
        const TNode &T = tt.nodes[L.far_node[k]];
        for (int d = 0; d < 3; d++) L.far_box.push_back(T.lo[d]);
        for (int d = 0; d < 3; d++) L.far_box.push_back(T.hi[d]);
        long long acc = 0;
        for (Range r : farl[k]) {
            while (r.a < r.b) {
                const int take = (int)std::min<long long>(r.b - r.a, ITEM_SOURCES - acc);
                L.item_rng.push_back(Range{r.a, r.a + take});
                r.a += take;
                acc += take;
                if (acc == ITEM_SOURCES) {
                    L.item_node.push_back(k);
                    L.item_off.push_back((int)L.item_rng.size());
                    acc = 0;
                }
            }
        }
        if (acc) {
            L.item_node.push_back(k);
            L.item_off.push_back((int)L.item_rng.size());
        }
        L.far_item_off.push_back((int)L.item_node.size());
    }
    // target leaves, in order: points, far ancestors (nearest first), and the near ranges: the complement in [0, nb)
    // of the union of the ancestors' far ranges (disjoint by construction of the traversal), carried down the tree
    // merged (a leaf's union has about as many ranges as its complement).  The target tree is cut into subtrees of
    // at most a 64th of its targets, walked on host threads, their leaves joined in order.
    struct Frag {
        std::vector<int> leaf_a, leaf_b, anc, anc_n, near_n;
        std::vector<Range> near_rng;
        long long interp = 0, pairs = 0, evals = 0;
    };
    auto visit = [&](auto &&self, int id, const std::vector<Range> &up, std::vector<int> &path, Frag &f) -> void {
        const TNode &T = tt.nodes[id];
        const int k = far_of_node[id];
        std::vector<Range> cur;
        const std::vector<Range> *u = &up;
        if (k >= 0) {
            cur.reserve(up.size() + farl[k].size());
            cur.insert(cur.end(), up.begin(), up.end());
            cur.insert(cur.end(), farl[k].begin(), farl[k].end());
            merge_ranges(cur);
            u = &cur;
            path.push_back(k);
        }
        if (T.c0 >= 0) {
            self(self, T.c0, *u, path, f);
            self(self, T.c1, *u, path, f);
        } else {
            f.leaf_a.push_back(T.a);
            f.leaf_b.push_back(T.b);
            for (auto it = path.rbegin(); it != path.rend(); ++it) f.anc.push_back(*it);
            f.anc_n.push_back((int)path.size());
            f.interp += (long long)L.P * (T.b - T.a) * (long long)path.size();
            const size_t r0 = f.near_rng.size();
            int at = 0;
            for (const Range &r : *u) {
                if (r.a > at) f.near_rng.push_back(Range{at, r.a});
                at = std::max(at, r.b);
            }
            if (at < nb) f.near_rng.push_back(Range{at, nb});
            f.near_n.push_back((int)(f.near_rng.size() - r0));
            for (size_t r = r0; r < f.near_rng.size(); r++) {
                f.pairs++;
                f.evals += (long long)(T.b - T.a) * (f.near_rng[r].b - f.near_rng[r].a);
            }
        }
        if (k >= 0) path.pop_back();
    };
    struct Task {
        int node;
        std::vector<int> path;
        std::vector<Range> uni;
    };
    std::vector<Task> tasks;
    if (nleaf) {
        const int cut = std::max(TREE_LEAF, (tt.nodes[0].b - tt.nodes[0].a) / 64);
        std::vector<int> path;
        auto split = [&](auto &&self, int id, const std::vector<Range> &up) -> void {
            const TNode &T = tt.nodes[id];
            if (T.c0 < 0 || T.b - T.a <= cut) {
                tasks.push_back(Task{id, path, up});
                return;
            }
            const int k = far_of_node[id];
            std::vector<Range> cur = up;
            if (k >= 0) {
                cur.insert(cur.end(), farl[k].begin(), farl[k].end());
                merge_ranges(cur);
                path.push_back(k);
            }
            self(self, T.c0, cur);
            self(self, T.c1, cur);
            if (k >= 0) path.pop_back();
        };
        split(split, 0, std::vector<Range>{});
    }
    std::vector<Frag> frags(tasks.size());
    parallel_for((int)tasks.size(), [&](int t0, int t1) {
        for (int t = t0; t < t1; t++) {
            std::vector<int> path = tasks[t].path;
            visit(visit, tasks[t].node, tasks[t].uni, path, frags[t]);
        }
    });
    L.near_off.push_back(0);
    L.anc_off.push_back(0);
    for (const Frag &f : frags) {
        L.leaf_a.insert(L.leaf_a.end(), f.leaf_a.begin(), f.leaf_a.end());
        L.leaf_b.insert(L.leaf_b.end(), f.leaf_b.begin(), f.leaf_b.end());
        L.anc.insert(L.anc.end(), f.anc.begin(), f.anc.end());
        for (int c : f.anc_n) L.anc_off.push_back(L.anc_off.back() + c);
        L.near_rng.insert(L.near_rng.end(), f.near_rng.begin(), f.near_rng.end());
        for (int c : f.near_n) L.near_off.push_back(L.near_off.back() + c);
        L.interp_evals += f.interp;
        L.near_pairs += f.pairs;
        L.near_evals += f.evals;
    }
    virtual_starts(L.item_off, L.item_rng, L.item_vst, L.item_tot);
    virtual_starts(L.near_off, L.near_rng, L.near_vst, L.near_tot);
    L.leaf_item_off.assign(nleaf + 1, 0);
    for (int l = 0; l < nleaf; l++)
        L.leaf_item_off[l + 1] = L.leaf_item_off[l] + (L.near_tot[l] + NEAR_ITEM - 1) / NEAR_ITEM;
    L.near_item.resize(4 * (size_t)L.leaf_item_off[nleaf]);
    parallel_for(nleaf, [&](int l0, int l1) {
        for (int l = l0; l < l1; l++) {
            int r = L.near_off[l];
            int *o = &L.near_item[4 * (size_t)L.leaf_item_off[l]];
            for (int v0 = 0; v0 < L.near_tot[l]; v0 += NEAR_ITEM, o += 4) {
                while (r + 1 < L.near_off[l + 1] && L.near_vst[r + 1] <= v0) r++;
                o[0] = l;
                o[1] = r;
                o[2] = v0;
                o[3] = std::min(v0 + NEAR_ITEM, L.near_tot[l]);
            }
        }
    });
    return L;
}

// ------------------------------------------------------------------------------------------------ device -----------

// Chebyshev points of the second kind on [-1, 1] and their barycentric weights (-1)^j, halved at the ends
template <int N1> __device__ __forceinline__ double cheb_pt(int j) { return cospi((double)j / (double)(N1 - 1)); }
template <int N1> __device__ __forceinline__ double cheb_w(int j)
{
    const double s = (j & 1) ? -1.0 : 1.0;
    return (j == 0 || j == N1 - 1) ? 0.5 * s : s;
}

// Stages the sources at positions [base, base + nk) of a group's concatenated ranges r0..r1 (starts vst) into S, one
// per thread, their positions taken relative to o (formed in FP64); rb, the range holding base, only moves forward
// and is the same on every thread
template <typename T>
__device__ __forceinline__ void stage_sources(const double *__restrict__ src, const Range *__restrict__ rng,
                                              const int *__restrict__ vst, int r1, int &rb, int base, int nk,
                                              T (*S)[7], V3<double> o)
{
    while (rb + 1 < r1 && vst[rb + 1] <= base) rb++;
    if ((int)threadIdx.x < nk) {
        const int v = base + (int)threadIdx.x;
        int r = rb;
        while (r + 1 < r1 && vst[r + 1] <= v) r++;
        const size_t q = 7 * (size_t)(rng[r].a + (v - vst[r]));
        S[threadIdx.x][0] = (T)(src[q] - o.x);
        S[threadIdx.x][1] = (T)(src[q + 1] - o.y);
        S[threadIdx.x][2] = (T)(src[q + 2] - o.z);
        for (int c = 3; c < 7; c++) S[threadIdx.x][c] = (T)src[q + c];
    }
}

// Threads of a proxy block: one proxy per lane, whole waves
template <int N1> constexpr int proxy_block() { return ((N1 * N1 * N1 + 63) / 64) * 64; }

// One work item (sources of one far cluster) at that cluster's proxies: partial[item][proxy][3].  The item's sources
// pass through LDS in full tiles of one per thread across its ranges, read by every lane as a broadcast.  T = float
// evaluates relative to the cluster's box centre and sums a tile in FP32, carrying the running sum in FP64; T = double
// evaluates in absolute coordinates.
template <int N1, typename T>
__global__ __launch_bounds__(512) void tree_proxy(const double *__restrict__ src, const int *__restrict__ item_node,
                                                  const int *__restrict__ item_off, const Range *__restrict__ item_rng,
                                                  const int *__restrict__ item_vst, const int *__restrict__ item_tot,
                                                  const double *__restrict__ box, double *__restrict__ partial)
{
    constexpr int P = N1 * N1 * N1, BLK = proxy_block<N1>();
    __shared__ T S[BLK][7];
    const int it = blockIdx.x;
    const int k = item_node[it];
    const double *bx = box + 6 * (size_t)k;
    const V3<double> o = sizeof(T) == 4 ? mk3(0.5 * (bx[0] + bx[3]), 0.5 * (bx[1] + bx[4]), 0.5 * (bx[2] + bx[5]))
                                        : mk3(0.0, 0.0, 0.0);
    const int p = min((int)threadIdx.x, P - 1);
    const int i = p / (N1 * N1), j = (p / N1) % N1, l = p % N1;
    const V3<T> px = mk3((T)(bx[0] + 0.5 * (bx[3] - bx[0]) * (1.0 + cheb_pt<N1>(i)) - o.x),
                         (T)(bx[1] + 0.5 * (bx[4] - bx[1]) * (1.0 + cheb_pt<N1>(j)) - o.y),
                         (T)(bx[2] + 0.5 * (bx[5] - bx[2]) * (1.0 + cheb_pt<N1>(l)) - o.z));
    V3<double> acc = mk3(0.0, 0.0, 0.0);
    const int r1 = item_off[it + 1], tot = item_tot[it];
    int rb = item_off[it];
    for (int base = 0; base < tot; base += BLK) {
        const int nk = min(BLK, tot - base);
        __syncthreads();
        stage_sources(src, item_rng, item_vst, r1, rb, base, nk, S, o);
        __syncthreads();
        V3<T> a = mk3(T(0), T(0), T(0));
#pragma unroll 4
        for (int s = 0; s < nk; s++)
            a = a + blob_vel(px - mk3(S[s][0], S[s][1], S[s][2]), mk3(S[s][3], S[s][4], S[s][5]), S[s][6]);
        acc = acc + mk3((double)a.x, (double)a.y, (double)a.z);
    }
    if ((int)threadIdx.x < P) {
        double *q = partial + 3 * ((size_t)it * P + threadIdx.x);
        q[0] = acc.x;
        q[1] = acc.y;
        q[2] = acc.z;
    }
}

// proxy_u[k][p] = sum of cluster k's items in item order
__global__ void tree_proxy_reduce(const double *__restrict__ partial, const int *__restrict__ far_item_off, int P,
                                  int nfar, double *__restrict__ proxy_u)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long long)nfar * P) return;
    const int k = (int)(g / P), p = (int)(g % P);
    double u0 = 0.0, u1 = 0.0, u2 = 0.0;
    for (int it = far_item_off[k]; it < far_item_off[k + 1]; it++) {
        const double *o = partial + 3 * ((size_t)it * P + p);
        u0 += o[0];
        u1 += o[1];
        u2 += o[2];
    }
    double *o = proxy_u + 3 * ((size_t)k * P + p);
    o[0] = u0;
    o[1] = u1;
    o[2] = u2;
}

// barycentric weights of t in [-1, 1] at the N1 Chebyshev points (exactly 1 at a point t equals)
template <int N1> __device__ __forceinline__ void bary(double t, double *w)
{
    int hit = -1;
    double sum = 0.0;
#pragma unroll
    for (int j = 0; j < N1; j++) {
        const double d = t - cheb_pt<N1>(j);
        if (d == 0.0) hit = j;
        w[j] = cheb_w<N1>(j) / d;
        sum += w[j];
    }
    if (hit >= 0) {
#pragma unroll
        for (int j = 0; j < N1; j++) w[j] = j == hit ? 1.0 : 0.0;
    } else {
        const double is = 1.0 / sum;
#pragma unroll
        for (int j = 0; j < N1; j++) w[j] *= is;
    }
}

// Near sums of one work item: a target leaf's targets on the lanes of each of a block's four waves, against at most
// NEAR_ITEM of the leaf's near sources (positions [v0, v1) of its concatenated ranges, starting in range rb).  The
// sources pass through LDS in full tiles of 256, wave w taking every fourth; the four waves' sums are added in wave
// order into partial[item][lane].  T = float evaluates relative to the leaf's first target (the digits of x - xp kept
// where the kernel is largest) and sums a tile in FP32, the running sum in FP64.
template <typename T>
__global__ __launch_bounds__(256) void tree_near(const double *__restrict__ src, const double *__restrict__ tx,
                                                 const int *__restrict__ leaf_a, const int *__restrict__ leaf_b,
                                                 const int4 *__restrict__ items, const int *__restrict__ near_off,
                                                 const Range *__restrict__ near_rng, const int *__restrict__ near_vst,
                                                 double *__restrict__ partial)
{
    __shared__ T S[256][7];
    __shared__ double red[3][64][3];
    const int4 itm = items[blockIdx.x];  // leaf, first range, v0, v1
    const int leaf = itm.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int a = leaf_a[leaf], b = leaf_b[leaf];
    const V3<double> o = sizeof(T) == 4 ? ld3(tx + 3 * (size_t)a) : mk3(0.0, 0.0, 0.0);
    const V3<double> xd = ld3(tx + 3 * (size_t)min(a + lane, b - 1));
    const V3<T> x = mk3((T)(xd.x - o.x), (T)(xd.y - o.y), (T)(xd.z - o.z));
    V3<double> acc = mk3(0.0, 0.0, 0.0);
    const int r1 = near_off[leaf + 1];
    int rb = itm.y;
    for (int base = itm.z; base < itm.w; base += 256) {
        const int nk = min(256, itm.w - base);
        __syncthreads();
        stage_sources(src, near_rng, near_vst, r1, rb, base, nk, S, o);
        __syncthreads();
        V3<T> t = mk3(T(0), T(0), T(0));
#pragma unroll 4
        for (int s = w; s < nk; s += 4)
            t = t + blob_vel(x - mk3(S[s][0], S[s][1], S[s][2]), mk3(S[s][3], S[s][4], S[s][5]), S[s][6]);
        acc = acc + mk3((double)t.x, (double)t.y, (double)t.z);
    }
    if (w > 0) {
        red[w - 1][lane][0] = acc.x;
        red[w - 1][lane][1] = acc.y;
        red[w - 1][lane][2] = acc.z;
    }
    __syncthreads();
    if (w == 0) {
        for (int v = 0; v < 3; v++) acc = acc + mk3(red[v][lane][0], red[v][lane][1], red[v][lane][2]);
        double *q = partial + 3 * ((size_t)blockIdx.x * 64 + lane);
        q[0] = acc.x;
        q[1] = acc.y;
        q[2] = acc.z;
    }
}

// One target leaf per block of four waves: its near items' partial sums in item order, then the interpolants of its
// far ancestors (shared out over the waves, their sums added in wave order); u in input order
template <int N1>
__global__ __launch_bounds__(256) void tree_finish(const double *__restrict__ tx, const int *__restrict__ leaf_a,
                                                   const int *__restrict__ leaf_b, const int *__restrict__ leaf_item_off,
                                                   const double *__restrict__ partial, const int *__restrict__ anc_off,
                                                   const int *__restrict__ anc, const double *__restrict__ box,
                                                   const double *__restrict__ proxy_u, const int *__restrict__ perm,
                                                   double *__restrict__ u)
{
    constexpr int P = N1 * N1 * N1;
    __shared__ double red[3][64][3];
    const int leaf = blockIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int a = leaf_a[leaf], b = leaf_b[leaf];
    const int i = min(a + lane, b - 1);
    const V3<double> x = ld3(tx + 3 * (size_t)i);
    V3<double> acc = mk3(0.0, 0.0, 0.0);
    if (w == 0)
        for (int it = leaf_item_off[leaf]; it < leaf_item_off[leaf + 1]; it++)
            acc = acc + ld3(partial + 3 * ((size_t)it * 64 + lane));
    for (int r = anc_off[leaf] + w; r < anc_off[leaf + 1]; r += 4) {
        const int k = anc[r];
        const double *bx = box + 6 * (size_t)k;
        double wx[N1], wy[N1], wz[N1];
        bary<N1>((2.0 * x.x - bx[0] - bx[3]) / (bx[3] - bx[0]), wx);
        bary<N1>((2.0 * x.y - bx[1] - bx[4]) / (bx[4] - bx[1]), wy);
        bary<N1>((2.0 * x.z - bx[2] - bx[5]) / (bx[5] - bx[2]), wz);
        const double *U = proxy_u + 3 * (size_t)k * P;
        V3<double> f = mk3(0.0, 0.0, 0.0);
        for (int p0 = 0; p0 < N1; p0++) {
            V3<double> g = mk3(0.0, 0.0, 0.0);
            for (int p1 = 0; p1 < N1; p1++) {
                V3<double> h = mk3(0.0, 0.0, 0.0);
#pragma unroll
                for (int p2 = 0; p2 < N1; p2++) h = h + wz[p2] * ld3(U + 3 * ((p0 * N1 + p1) * N1 + p2));
                g = g + wy[p1] * h;
            }
            f = f + wx[p0] * g;
        }
        acc = acc + f;
    }
    if (w > 0) {
        red[w - 1][lane][0] = acc.x;
        red[w - 1][lane][1] = acc.y;
        red[w - 1][lane][2] = acc.z;
    }
    __syncthreads();
    if (w == 0 && a + lane < b) {
        for (int v = 0; v < 3; v++) acc = acc + mk3(red[v][lane][0], red[v][lane][1], red[v][lane][2]);
        double *o = u + 3 * (size_t)perm[i];
        o[0] = acc.x;
        o[1] = acc.y;
        o[2] = acc.z;
    }
}

// Morton codes of n points and their indices (the sort's keys and values)
__global__ void tree_codes(const double *__restrict__ x, int n, Frame f, uint64_t *__restrict__ code,
                           int *__restrict__ idx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    code[i] = morton(x + 3 * (size_t)i, f);
    idx[i] = i;
}

// The sources in sorted order: position, strength, core radius squared (7 per source), and their positions alone
__global__ void tree_gather_src(const int *__restrict__ perm, const double *__restrict__ bx,
                                const double *__restrict__ ba, const double *__restrict__ bd, int n,
                                double *__restrict__ src, double *__restrict__ pos)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int j = perm[s];
    for (int d = 0; d < 3; d++) {
        src[7 * (size_t)s + d] = pos[3 * (size_t)s + d] = bx[3 * (size_t)j + d];
        src[7 * (size_t)s + 3 + d] = ba[3 * (size_t)j + d];
    }
    src[7 * (size_t)s + 6] = bd[j] * bd[j];
}

__global__ void tree_gather_pts(const int *__restrict__ perm, const double *__restrict__ x, int n,
                                double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int d = 0; d < 3; d++) out[3 * (size_t)i + d] = x[3 * (size_t)perm[i] + d];
}

// The leaves' tight boxes from the points in sorted order: a thread per leaf of at most TREE_LEAF points, 6 per leaf
__global__ void tree_leaf_boxes(const double *__restrict__ pos, const int2 *__restrict__ leaf, int nleaf,
                                double *__restrict__ box)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nleaf) return;
    const int2 r = leaf[l];
    double lo[3], hi[3];
    for (int d = 0; d < 3; d++) lo[d] = hi[d] = pos[3 * (size_t)r.x + d];
    for (int i = r.x + 1; i < r.y; i++)
        for (int d = 0; d < 3; d++) {
            const double v = pos[3 * (size_t)i + d];
            lo[d] = fmin(lo[d], v);
            hi[d] = fmax(hi[d], v);
        }
    for (int d = 0; d < 3; d++) {
        box[6 * (size_t)l + d] = lo[d];
        box[6 * (size_t)l + 3 + d] = hi[d];
    }
}

template <typename T> T *tnew(size_t n, std::vector<void *> &owned)
{
    T *p = nullptr;
    TCHECK(hipMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(T)));
    owned.push_back(p);
    return p;
}

template <typename T> T *tput(const std::vector<T> &h, hipStream_t st, std::vector<void *> &owned)
{
    T *p = tnew<T>(h.size(), owned);
    if (!h.empty()) TCHECK(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return p;
}

struct TOwned {
    std::vector<void *> v;
    hipStream_t st;
    ~TOwned()
    {
        (void)hipStreamSynchronize(st);
        for (void *p : v) (void)hipFree(p);
    }
};

struct DevLists {
    const double *box = nullptr;
    double *proxy = nullptr, *part = nullptr;
    const int *item_node = nullptr, *item_off = nullptr, *far_item_off = nullptr;
    const Range *item_rng = nullptr, *near_rng = nullptr;
    const int *leaf_a = nullptr, *leaf_b = nullptr, *near_off = nullptr, *anc_off = nullptr, *anc = nullptr;
    const int *item_vst = nullptr, *item_tot = nullptr, *near_vst = nullptr, *near_tot = nullptr;
    const int4 *near_item = nullptr;
    const int *leaf_item_off = nullptr;
    double *near_part = nullptr;
    int nnear = 0;
    int nfar = 0, nitem = 0, nleaf = 0, P = 0;
};

DevLists upload_lists(hipStream_t st, const Lists &L, std::vector<void *> &owned)
{
    DevLists d;
    d.P = L.P;
    d.nfar = (int)L.far_node.size();
    d.nitem = (int)L.item_node.size();
    d.nleaf = (int)L.leaf_a.size();
    d.box = tput(L.far_box, st, owned);
    d.proxy = tnew<double>((size_t)d.nfar * L.P * 3, owned);
    d.part = tnew<double>((size_t)d.nitem * L.P * 3, owned);
    d.item_node = tput(L.item_node, st, owned);
    d.item_off = tput(L.item_off, st, owned);
    d.item_rng = tput(L.item_rng, st, owned);
    d.far_item_off = tput(L.far_item_off, st, owned);
    d.leaf_a = tput(L.leaf_a, st, owned);
    d.leaf_b = tput(L.leaf_b, st, owned);
    d.near_off = tput(L.near_off, st, owned);
    d.near_rng = tput(L.near_rng, st, owned);
    d.anc_off = tput(L.anc_off, st, owned);
    d.anc = tput(L.anc, st, owned);
    d.item_vst = tput(L.item_vst, st, owned);
    d.item_tot = tput(L.item_tot, st, owned);
    d.near_vst = tput(L.near_vst, st, owned);
    d.near_tot = tput(L.near_tot, st, owned);
    d.nnear = (int)L.near_item.size() / 4;
    d.near_item = reinterpret_cast<const int4 *>(tput(L.near_item, st, owned));
    d.leaf_item_off = tput(L.leaf_item_off, st, owned);
    d.near_part = tnew<double>((size_t)d.nnear * 64 * 3, owned);
    return d;
}

template <int N1, typename T>
void tree_launch(hipStream_t st, const DevLists &d, const double *d_src, const double *d_tx, const int *d_perm,
                 double *d_u)
{
    if (d.nitem) {
        tree_proxy<N1, T><<<d.nitem, proxy_block<N1>(), 0, st>>>(d_src, d.item_node, d.item_off, d.item_rng,
                                                                  d.item_vst, d.item_tot, d.box, d.part);
        TCHECK(hipGetLastError());
        const long long nr = (long long)d.nfar * d.P;
        tree_proxy_reduce<<<(unsigned)((nr + 255) / 256), 256, 0, st>>>(d.part, d.far_item_off, d.P, d.nfar, d.proxy);
        TCHECK(hipGetLastError());
    }
    if (d.nnear) {
        tree_near<T><<<d.nnear, 256, 0, st>>>(d_src, d_tx, d.leaf_a, d.leaf_b, d.near_item, d.near_off, d.near_rng,
                                               d.near_vst, d.near_part);
        TCHECK(hipGetLastError());
    }
    if (d.nleaf) {
        tree_finish<N1><<<d.nleaf, 256, 0, st>>>(d_tx, d.leaf_a, d.leaf_b, d.leaf_item_off, d.near_part, d.anc_off,
                                                 d.anc, d.box, d.proxy, d_perm, d_u);
        TCHECK(hipGetLastError());
    }
}

template <typename T>
void launch_degree(int degree, hipStream_t st, const DevLists &d, const double *d_src, const double *d_tx,
                   const int *d_perm, double *d_u)
{
    switch (degree + 1) {
    case 2: tree_launch<2, T>(st, d, d_src, d_tx, d_perm, d_u); break;
    case 3: tree_launch<3, T>(st, d, d_src, d_tx, d_perm, d_u); break;
    case 4: tree_launch<4, T>(st, d, d_src, d_tx, d_perm, d_u); break;
    case 5: tree_launch<5, T>(st, d, d_src, d_tx, d_perm, d_u); break;
    case 6: tree_launch<6, T>(st, d, d_src, d_tx, d_perm, d_u); break;
    case 7: tree_launch<7, T>(st, d, d_src, d_tx, d_perm, d_u); break;
    default: tree_launch<8, T>(st, d, d_src, d_tx, d_perm, d_u); break;
    }
}

void check_tree_args(int nb, int np, const vp_tree_options &o)
{
    if (o.precision != 0 && o.precision != 1) throw std::invalid_argument("precision must be 0 (fp64) or 1 (mixed)");
    if (!(o.theta > 0.0 && o.theta < 1.0)) throw std::invalid_argument("theta must lie in (0, 1)");
    if (o.degree < 1 || o.degree > TREE_MAX_DEGREE)
        throw std::invalid_argument("degree must lie in 1.." + std::to_string(TREE_MAX_DEGREE));
    if (nb < 0 || np < 0) throw std::invalid_argument("negative size");
}

}  // namespace

// Every (target, source) pair must be covered exactly once: for each target leaf, its near ranges and the source
// ranges of its far ancestors tile [0, nb).  Returns the number of leaves that fail (host only, no device work).
long long tree_coverage_host(int nb, const double *bx, int np, const double *xs, const vp_tree_options &o,
                             vp_tree_stats *stats)
{
    check_tree_args(nb, np, o);
    const auto t0 = std::chrono::steady_clock::now();
    const Tree st = build_tree(bx, nb), tt = build_tree(xs, np);
    const Lists L = build_lists(tt, st, o.theta, o.degree);
    const double build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    // far cluster -> its ranges, from the items
    std::vector<std::vector<Range>> farr(L.far_node.size());
    for (size_t it = 0; it < L.item_node.size(); it++)
        for (int r = L.item_off[it]; r < L.item_off[it + 1]; r++) farr[L.item_node[it]].push_back(L.item_rng[r]);
    long long bad = 0;
    for (size_t l = 0; l < L.leaf_a.size(); l++) {
        std::vector<Range> all(L.near_rng.begin() + L.near_off[l], L.near_rng.begin() + L.near_off[l + 1]);
        for (int r = L.anc_off[l]; r < L.anc_off[l + 1]; r++)
            all.insert(all.end(), farr[L.anc[r]].begin(), farr[L.anc[r]].end());
        std::sort(all.begin(), all.end(), [](const Range &p, const Range &q) { return p.a < q.a; });
        int at = 0;
        bool okl = true;
        for (const Range &r : all) {
            if (r.a != at || r.b <= r.a) okl = false;
            at = r.b;
        }
        if (at != nb) okl = false;
        bad += okl ? 0 : 1;
    }
    if (stats) {
        stats->far_clusters = (long long)L.far_node.size();
        stats->far_pairs = L.far_pairs;
        stats->near_pairs = L.near_pairs;
        stats->proxy_evals = L.proxy_evals;
        stats->near_evals = L.near_evals;
        stats->interp_evals = L.interp_evals;
        stats->build_ms = build_ms;
    }
    return bad;
}

void blob_velocities_tree_host(Ctx *c, int nb, const double *bx, const double *ba, const double *bd, int np,
                               const double *xs, const vp_tree_options &o, double *u, vp_tree_stats *stats)
{
    check_tree_args(nb, np, o);
    hipStream_t st = c->stream;
    const auto t0 = std::chrono::steady_clock::now();
    TOwned ow{{}, st};
    // on the device: Morton codes, the sort, the sources and targets gathered in sorted order
    Frame fs, ft;
    run_both([&] { fs = frame_of(bx, nb); }, [&] { ft = frame_of(xs, np); });
    const size_t ns = std::max(nb, 1), nt = std::max(np, 1);
    double *d_bx = tnew<double>(ns * 3, ow.v), *d_ba = tnew<double>(ns * 3, ow.v), *d_bd = tnew<double>(ns, ow.v);
    double *d_xs = tnew<double>(nt * 3, ow.v);
    if (nb) {
        TCHECK(hipMemcpyAsync(d_bx, bx, sizeof(double) * 3 * nb, hipMemcpyHostToDevice, st));
        TCHECK(hipMemcpyAsync(d_ba, ba, sizeof(double) * 3 * nb, hipMemcpyHostToDevice, st));
        TCHECK(hipMemcpyAsync(d_bd, bd, sizeof(double) * nb, hipMemcpyHostToDevice, st));
    }
    TCHECK(hipMemcpyAsync(d_xs, xs, sizeof(double) * 3 * np, hipMemcpyHostToDevice, st));
    uint64_t *d_kin = tnew<uint64_t>(ns + nt, ow.v), *d_kout = tnew<uint64_t>(ns + nt, ow.v);
    int *d_iin = tnew<int>(ns + nt, ow.v), *d_perm = tnew<int>(ns + nt, ow.v);
    double *d_src = tnew<double>(ns * 7, ow.v), *d_spos = tnew<double>(ns * 3, ow.v), *d_tx = tnew<double>(nt * 3, ow.v);
    size_t tmp_s = 0, tmp_t = 0;
    TCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_s, d_kin, d_kout, d_iin, d_perm, std::max(nb, 1), 0, 63, st));
    TCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_t, d_kin, d_kout, d_iin, d_perm, np, 0, 63, st));
    void *d_tmp = tnew<char>(std::max(tmp_s, tmp_t), ow.v);
    if (nb) {
        tree_codes<<<(unsigned)((nb + 255) / 256), 256, 0, st>>>(d_bx, nb, fs, d_kin, d_iin);
        TCHECK(hipGetLastError());
        TCHECK(hipcub::DeviceRadixSort::SortPairs(d_tmp, tmp_s, d_kin, d_kout, d_iin, d_perm, nb, 0, 63, st));
        tree_gather_src<<<(unsigned)((nb + 255) / 256), 256, 0, st>>>(d_perm, d_bx, d_ba, d_bd, nb, d_src, d_spos);
        TCHECK(hipGetLastError());
    }
    tree_codes<<<(unsigned)((np + 255) / 256), 256, 0, st>>>(d_xs, np, ft, d_kin + ns, d_iin + ns);
    TCHECK(hipGetLastError());
    TCHECK(hipcub::DeviceRadixSort::SortPairs(d_tmp, tmp_t, d_kin + ns, d_kout + ns, d_iin + ns, d_perm + ns, np, 0, 63,
                                              st));
    tree_gather_pts<<<(unsigned)((np + 255) / 256), 256, 0, st>>>(d_perm + ns, d_xs, np, d_tx);
    TCHECK(hipGetLastError());
    // on the host: the topologies from the sorted codes; the leaves' boxes on the device; then the lists
    std::vector<uint64_t> hks(nb), hkt(np);
    if (nb) TCHECK(hipMemcpyAsync(hks.data(), d_kout, sizeof(uint64_t) * nb, hipMemcpyDeviceToHost, st));
    TCHECK(hipMemcpyAsync(hkt.data(), d_kout + ns, sizeof(uint64_t) * np, hipMemcpyDeviceToHost, st));
    TCHECK(hipStreamSynchronize(st));
    const auto t1 = std::chrono::steady_clock::now();
    Tree stree, ttree;
    run_both([&] { stree = tree_topology(hks.data(), nb); }, [&] { ttree = tree_topology(hkt.data(), np); });
    const size_t nls = stree.leaves.size(), nlt = ttree.leaves.size();
    std::vector<int> hr(2 * (nls + nlt));
    for (size_t l = 0; l < nls; l++) {
        hr[2 * l] = stree.nodes[stree.leaves[l]].a;
        hr[2 * l + 1] = stree.nodes[stree.leaves[l]].b;
    }
    for (size_t l = 0; l < nlt; l++) {
        hr[2 * (nls + l)] = ttree.nodes[ttree.leaves[l]].a;
        hr[2 * (nls + l) + 1] = ttree.nodes[ttree.leaves[l]].b;
    }
    const int2 *d_lr = reinterpret_cast<const int2 *>(tput(hr, st, ow.v));
    double *d_lbox = tnew<double>(6 * (nls + nlt), ow.v);
    if (nls) tree_leaf_boxes<<<(unsigned)((nls + 255) / 256), 256, 0, st>>>(d_spos, d_lr, (int)nls, d_lbox);
    if (nlt)
        tree_leaf_boxes<<<(unsigned)((nlt + 255) / 256), 256, 0, st>>>(d_tx, d_lr + nls, (int)nlt, d_lbox + 6 * nls);
    TCHECK(hipGetLastError());
    std::vector<double> hbox(6 * (nls + nlt));
    TCHECK(hipMemcpyAsync(hbox.data(), d_lbox, sizeof(double) * hbox.size(), hipMemcpyDeviceToHost, st));
    TCHECK(hipStreamSynchronize(st));
    run_both([&] { tree_boxes(stree, hbox.data()); }, [&] { tree_boxes(ttree, hbox.data() + 6 * nls); });
    const Lists L = build_lists(ttree, stree, o.theta, o.degree);
    const auto t2 = std::chrono::steady_clock::now();
    const int *d_tperm = d_perm + ns;
    double *d_u = tnew<double>((size_t)np * 3, ow.v);
    const DevLists d = upload_lists(st, L, ow.v);
    TCHECK(hipStreamSynchronize(st));
    const auto t3 = std::chrono::steady_clock::now();
    auto ms = [](auto p, auto q) { return std::chrono::duration<double, std::milli>(q - p).count(); };
    hipEvent_t e0, e1;
    TCHECK(hipEventCreate(&e0));
    TCHECK(hipEventCreate(&e1));
    TCHECK(hipEventRecord(e0, st));
    if (o.precision == 1)
        launch_degree<float>(o.degree, st, d, d_src, d_tx, d_tperm, d_u);
    else
        launch_degree<double>(o.degree, st, d, d_src, d_tx, d_tperm, d_u);
    TCHECK(hipEventRecord(e1, st));
    TCHECK(hipMemcpyAsync(u, d_u, sizeof(double) * (size_t)np * 3, hipMemcpyDeviceToHost, st));
    TCHECK(hipStreamSynchronize(st));
    float kms = 0.f;
    TCHECK(hipEventElapsedTime(&kms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (stats) {
        stats->far_clusters = (long long)L.far_node.size();
        stats->far_pairs = L.far_pairs;
        stats->near_pairs = L.near_pairs;
        stats->proxy_evals = L.proxy_evals;
        stats->near_evals = L.near_evals;
        stats->interp_evals = L.interp_evals;
        stats->build_ms = ms(t0, t3);
        stats->device_ms = kms;
        stats->sort_ms = ms(t0, t1);
        stats->lists_ms = ms(t1, t2);
    }
}

}  // namespace vp
