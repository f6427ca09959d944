// Host binned-SAH BVH builder and flattening to the device node layouts.
//
// Replaces the Embree build of the reference (include/nanogi/rt.hpp:
// 2087-2143; RTC_SCENE_STATIC | RTC_SCENE_INCOHERENT).  Top-down binned SAH
// (Wald, "On fast construction of SAH-based bounding volume hierarchies",
// 2007) over triangle centroids, subtrees built in parallel; node order is
// depth-first pre-order and independent of thread scheduling, so the build
// is deterministic.  The binary tree is then either emitted as 64-byte
// binary records or, built down to single triangles, collapsed into 8-wide
// 128-byte records by an SAH-optimal dynamic programme.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <thread>

#include "ngi_bvh.h"
#include "ngi_internal.h"

namespace ngi {
namespace {

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], b.lo[k]), hi[k] = std::max(hi[k], b.hi[k]);
    }
    void grow(const float p[3]) {
        for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], p[k]), hi[k] = std::max(hi[k], p[k]);
    }
    double area() const {
        if (!(hi[0] >= lo[0])) return 0.0;
        const double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

struct Node {
    Box box;
    int32_t left = -1, right = -1;   // children (node indices) for internal nodes
    uint32_t start = 0, count = 0;   // leaf range (count > 0)
};

struct Ctx {
    std::vector<Box> tri_box;
    std::vector<float> centroid;   // 3 per triangle
    std::vector<uint32_t> refs;
    int max_leaf = 8;
    int bins = 32;
    std::atomic<int> spare_threads{0};
    double ct = 1.2, ci = 1.0;     // relative traversal / intersection cost
};

// SAH cost of a wide-node visit in units of one triangle test, for the
// collapse into 8-wide nodes (default of ngi_bvh_opts.traversal_cost there).
// Measured on the 1.8M-triangle sphere field / the nave (ptdirect, samples/s
// on one MI355X): 1,454 / 593 M at 1.0, 1,441 / 599 M at 1.5, 1,440 / 596 M
// at 2.0, 1,420 / 593 M at 2.5, 1,396 M (C3) at 3.0; the greedy collapse
// (open the child of largest area until eight) 1,451 / 582 M.
#ifndef NGI_BVH8_NODE_COST
#define NGI_BVH8_NODE_COST 1.0
#endif

const int kMedianDepth = 36;        // beyond this depth: object-median splits (bounds depth by 36 + log2(n))

void build_rec(Ctx& c, uint32_t b, uint32_t e, int depth, std::vector<Node>& out);

int32_t build_child(Ctx& c, uint32_t b, uint32_t e, int depth, std::vector<Node>& out) {
    const int32_t idx = (int32_t)out.size();
    build_rec(c, b, e, depth, out);
    return idx;
}

void build_rec(Ctx& c, uint32_t b, uint32_t e, int depth, std::vector<Node>& out) {
    const uint32_t n = e - b;
    const size_t self = out.size();
    out.push_back(Node());
    Box bounds, cb;
    for (uint32_t i = b; i < e; ++i) {
        bounds.grow(c.tri_box[c.refs[i]]);
        cb.grow(&c.centroid[3 * (size_t)c.refs[i]]);
    }
    out[self].box = bounds;
    auto make_leaf = [&]() {
        out[self].start = b;
        out[self].count = n;
    };
    if (n == 1) return make_leaf();

    uint32_t mid = b;
    bool split_found = false;
    if (depth < kMedianDepth) {
        // Binned SAH.
        const int B = c.bins;
        double best_cost = INFINITY;
        int best_axis = -1, best_bin = -1;
        for (int ax = 0; ax < 3; ++ax) {
            const float ext = cb.hi[ax] - cb.lo[ax];
            if (!(ext > 0)) continue;
            const double scale = B / (double)ext;
            std::vector<Box> bb(B);
            std::vector<uint32_t> bc(B, 0);
            for (uint32_t i = b; i < e; ++i) {
                const uint32_t t = c.refs[i];
                int k = (int)((c.centroid[3 * (size_t)t + ax] - cb.lo[ax]) * scale);
                k = std::min(std::max(k, 0), B - 1);
                bb[k].grow(c.tri_box[t]);
                bc[k]++;
            }
            std::vector<double> right_area(B);
            std::vector<uint32_t> right_count(B);
            Box acc;
            uint32_t cnt = 0;
            for (int k = B - 1; k > 0; --k) {
                acc.grow(bb[k]);
                cnt += bc[k];
                right_area[k] = acc.area();
                right_count[k] = cnt;
            }
            acc = Box();
            cnt = 0;
            for (int k = 0; k < B - 1; ++k) {
                acc.grow(bb[k]);
                cnt += bc[k];
                if (cnt == 0 || right_count[k + 1] == 0) continue;
                const double cost = acc.area() * cnt + right_area[k + 1] * right_count[k + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = ax;
                    best_bin = k;
                }
            }
        }
        const double parent_area = bounds.area();
        const double leaf_cost = c.ci * n;
        const double split_cost = c.ct + (parent_area > 0 ? c.ci * best_cost / parent_area : INFINITY);
        if (best_axis >= 0) {
            if (n <= (uint32_t)c.max_leaf && leaf_cost <= split_cost) return make_leaf();
            const int ax = best_axis;
            const double scale = c.bins / (double)(cb.hi[ax] - cb.lo[ax]);
            auto it = std::partition(c.refs.begin() + b, c.refs.begin() + e, [&](uint32_t t) {
                int k = (int)((c.centroid[3 * (size_t)t + ax] - cb.lo[ax]) * scale);
                k = std::min(std::max(k, 0), c.bins - 1);
                return k <= best_bin;
            });
            mid = (uint32_t)(it - c.refs.begin());
            split_found = mid > b && mid < e;
        }
    }
    if (!split_found) {
        if (n <= (uint32_t)c.max_leaf) return make_leaf();
        // Object median along the widest centroid axis (index order if all
        // centroids coincide).
        int ax = 0;
        for (int k = 1; k < 3; ++k)
            if (cb.hi[k] - cb.lo[k] > cb.hi[ax] - cb.lo[ax]) ax = k;
        mid = b + n / 2;
        std::nth_element(c.refs.begin() + b, c.refs.begin() + mid, c.refs.begin() + e, [&](uint32_t x, uint32_t y) {
            const float cx = c.centroid[3 * (size_t)x + ax], cy = c.centroid[3 * (size_t)y + ax];
            return cx < cy || (cx == cy && x < y);
        });
    }
    // Children: the right subtree may be built on another thread into its
    // own vector and appended afterwards (pre-order is preserved).
    bool parallel = false;
    if (n > 32768) {
        if (c.spare_threads.fetch_sub(1) > 0) parallel = true;
        else c.spare_threads.fetch_add(1);
    }
    if (parallel) {
        std::vector<Node> right_nodes;
        auto fut = std::async(std::launch::async, [&]() { build_rec(c, mid, e, depth + 1, right_nodes); });
        const int32_t l = build_child(c, b, mid, depth + 1, out);
        fut.get();
        c.spare_threads.fetch_add(1);
        const int32_t off = (int32_t)out.size();
        for (Node nd : right_nodes) {
            if (nd.count == 0) nd.left += off, nd.right += off;
            out.push_back(nd);
        }
        out[self].left = l;
        out[self].right = off;
    } else {
        const int32_t l = build_child(c, b, mid, depth + 1, out);
        const int32_t r = build_child(c, mid, e, depth + 1, out);
        out[self].left = l;
        out[self].right = r;
    }
}

void depth_stats(const std::vector<Node>& nodes, int32_t i, uint32_t d, uint32_t& max_depth) {
    max_depth = std::max(max_depth, d);
    if (nodes[i].count) return;
    depth_stats(nodes, nodes[i].left, d + 1, max_depth);
    depth_stats(nodes, nodes[i].right, d + 1, max_depth);
}

// ---------------- binary records ----------------
void emit_bvh2(const std::vector<Node>& nodes, Bvh& out) {
    // Internal nodes in pre-order get consecutive record indices.
    std::vector<int32_t> rec(nodes.size(), -1);
    int32_t nrec = 0;
    for (size_t i = 0; i < nodes.size(); ++i)
        if (nodes[i].count == 0) rec[i] = nrec++;
    const bool root_leaf = nodes[0].count != 0;
    if (root_leaf) nrec = 1;
    out.nodes.assign((size_t)nrec * sizeof(Bvh2Node), 0);
    Bvh2Node* R = reinterpret_cast<Bvh2Node*>(out.nodes.data());
    auto ref = [&](int32_t i) -> int32_t { return nodes[i].count ? ngi_leaf_ref(nodes[i].start, nodes[i].count) : rec[i]; };
    auto fill = [&](Bvh2Node& r, const Node& a, int32_t ra, const Node& b, int32_t rb) {
        r.c0lox = a.box.lo[0], r.c0hix = a.box.hi[0], r.c0loy = a.box.lo[1], r.c0hiy = a.box.hi[1];
        r.c1lox = b.box.lo[0], r.c1hix = b.box.hi[0], r.c1loy = b.box.lo[1], r.c1hiy = b.box.hi[1];
        r.c0loz = a.box.lo[2], r.c0hiz = a.box.hi[2], r.c1loz = b.box.lo[2], r.c1hiz = b.box.hi[2];
        r.child0 = ra, r.child1 = rb;
    };
    if (root_leaf) {
        fill(R[0], nodes[0], ref(0), nodes[0], ref(0));
    } else {
        for (size_t i = 0; i < nodes.size(); ++i)
            if (nodes[i].count == 0) fill(R[rec[i]], nodes[nodes[i].left], ref(nodes[i].left), nodes[nodes[i].right], ref(nodes[i].right));
    }
    out.num_nodes = (uint32_t)nrec;
    out.node_bytes = sizeof(Bvh2Node);
}

// ---------------- 8-wide records ----------------
// Collapse: each wide node adopts up to eight descendants of a binary node
// by repeatedly opening the internal child with the largest surface area.
struct WideTmp {
    int32_t bin;                 // binary node this wide node represents
    std::vector<int32_t> kids;   // binary node indices of the children
    std::vector<uint8_t> leaf;   // per kid: a leaf over the kid's whole subtree
};

// Optimal collapse of the binary tree into 8-wide nodes (after Ylitie,
// Karras, Laine, "Efficient Incoherent Ray Traversal on GPUs Through
// Compressed Wide BVHs", HPG 2017, section 3.1).  For every binary node n and
// i = 1..8 child slots, C(n, i) is the least SAH cost of n's subtree spread
// over at most i slots of one wide node:
//   leaf      A(n) * tris(n)                     (tris(n) <= max_leaf)
//   wide node A(n) * c_node + D(n, 8)           (one slot holding a new node)
//   opened    D(n, i) = min_k C(l, k) + C(r, i - k), i >= 2
// and C(n, i) <= C(n, i - 1).  c_node is the cost of a wide-node visit in
// units of one triangle test.
struct Collapse {
    struct How {
        uint8_t kind;   // 0 leaf, 1 wide node, 2 opened
        uint8_t kl, kr; // opened: slots given to the left and right child
    };
    std::vector<float> cost;      // C(n, i) at [8 n + i - 1]
    std::vector<How> how;
    std::vector<uint8_t> split8;  // k of the best D(n, 8)
    std::vector<uint32_t> first, tris;

    void run(const std::vector<Node>& nodes, int max_leaf, double c_node) {
        const size_t n = nodes.size();
        cost.assign(8 * n, 0.0f);
        how.assign(8 * n, How{0, 0, 0});
        split8.assign(n, 1);
        first.assign(n, 0);
        tris.assign(n, 0);
        const double inv_root = 1.0 / std::max(nodes[0].box.area(), 1e-30);
        for (size_t q = n; q-- > 0;) {   // pre-order: children after their parent
            const Node& nd = nodes[q];
            const double A = nd.box.area() * inv_root;
            if (nd.count) {
                first[q] = nd.start, tris[q] = nd.count;
                const float c = (float)(A * nd.count);
                for (int i = 0; i < 8; ++i) cost[8 * q + i] = c, how[8 * q + i] = How{0, 0, 0};
                continue;
            }
            const size_t l = (size_t)nd.left, r = (size_t)nd.right;
            first[q] = first[l], tris[q] = tris[l] + tris[r];
            float D[9];
            uint8_t Dk[9];
            for (int j = 2; j <= 8; ++j) {
                D[j] = INFINITY, Dk[j] = 1;
                for (int k = 1; k < j; ++k) {
                    const float c = cost[8 * l + k - 1] + cost[8 * r + (j - k) - 1];
                    if (c < D[j]) D[j] = c, Dk[j] = (uint8_t)k;
                }
            }
            split8[q] = Dk[8];
            const float leafc = tris[q] <= (uint32_t)max_leaf ? (float)(A * tris[q]) : INFINITY;
            const float widec = (float)(A * c_node) + D[8];
            float best = leafc;
            How h{0, 0, 0};
            if (widec < best) best = widec, h = How{1, 0, 0};
            cost[8 * q] = best, how[8 * q] = h;
            for (int i = 2; i <= 8; ++i) {
                float c = cost[8 * q + i - 2];
                How hi = how[8 * q + i - 2];
                if (D[i] < c) c = D[i], hi = How{2, Dk[i], (uint8_t)(i - Dk[i])};
                cost[8 * q + i - 1] = c, how[8 * q + i - 1] = hi;
            }
        }
    }
    // Slot contents for binary node q given i slots: (node, as-leaf) pairs.
    void gather(int32_t q, int i, const std::vector<Node>& nodes, WideTmp& w) const {
        const How h = how[8 * (size_t)q + i - 1];
        if (h.kind == 2) {
            gather(nodes[q].left, h.kl, nodes, w);
            gather(nodes[q].right, h.kr, nodes, w);
            return;
        }
        w.kids.push_back(q);
        w.leaf.push_back(h.kind == 0 ? 1 : 0);
    }
};

// Quantise child boxes conservatively on a per-axis power-of-two grid.
void quantise(const Box& parent, const std::vector<Box>& kids, uint8_t valid, Bvh8Node& r) {
    int ebias[3];
    for (int a = 0; a < 3; ++a) {
        const float lo = parent.lo[a], hi = parent.hi[a];
        const float ext = hi - lo;
        int e = -126;
        if (ext > 0) {
            // smallest e with 255 * 2^e >= ext
            e = (int)std::ceil(std::log2((double)ext / 255.0));
            while (std::ldexp(255.0, e) < (double)ext) ++e;
            e = std::max(e, -126);
        }
        ebias[a] = e;
    }
    r.ox = parent.lo[0], r.oy = parent.lo[1], r.oz = parent.lo[2];
    r.ex = (uint8_t)(ebias[0] + 127), r.ey = (uint8_t)(ebias[1] + 127), r.ez = (uint8_t)(ebias[2] + 127);
    const float org[3] = {r.ox, r.oy, r.oz};
    uint8_t* qlo[3] = {r.qlox, r.qloy, r.qloz};
    uint8_t* qhi[3] = {r.qhix, r.qhiy, r.qhiz};
    for (size_t k = 0; k < 8; ++k) {
        for (int a = 0; a < 3; ++a) {
            if (!((valid >> k) & 1)) {
                qlo[a][k] = 0, qhi[a][k] = 0;     // unused slot (masked by `valid`)
                continue;
            }
            const double scale = std::ldexp(1.0, -ebias[a]);
            double l = std::floor(((double)kids[k].lo[a] - org[a]) * scale);
            double h = std::ceil(((double)kids[k].hi[a] - org[a]) * scale);
            l = std::min(std::max(l, 0.0), 255.0);
            h = std::min(std::max(h, 0.0), 255.0);
            // Guard against float rounding of the dequantised bound.
            while (l > 0 && (float)(org[a] + (float)std::ldexp(l, ebias[a])) > kids[k].lo[a]) l -= 1;
            while (h < 255 && (float)(org[a] + (float)std::ldexp(h, ebias[a])) < kids[k].hi[a]) h += 1;
            qlo[a][k] = (uint8_t)l, qhi[a][k] = (uint8_t)h;
        }
    }
}

void emit_bvh8(const std::vector<Node>& nodes, int max_leaf, double c_node, Bvh& out) {
    Collapse col;
    col.run(nodes, max_leaf, c_node);
    std::vector<WideTmp> wide;
    auto make = [&](int32_t bin) {
        WideTmp w;
        w.bin = bin;
        if (nodes[bin].count || col.how[8 * (size_t)bin].kind == 0) {
            // Root leaf: a single leaf child over the whole tree.
            w.kids.push_back(bin);
            w.leaf.push_back(1);
        } else {
            col.gather(nodes[bin].left, col.split8[bin], nodes, w);
            col.gather(nodes[bin].right, 8 - col.split8[bin], nodes, w);
        }
        return w;
    };
    // Pre-order numbering: assign indices depth-first so a node's first
    // internal child directly follows it.
    std::vector<int32_t> wide_index_of_bin(nodes.size(), -1);
    std::vector<int32_t> st = {0};
    while (!st.empty()) {
        int32_t bin = st.back();
        st.pop_back();
        wide_index_of_bin[bin] = (int32_t)wide.size();
        wide.push_back(make(bin));
        const WideTmp& w = wide.back();
        for (int k = (int)w.kids.size() - 1; k >= 0; --k)
            if (!w.leaf[k]) st.push_back(w.kids[k]);
    }
    out.nodes.assign(wide.size() * sizeof(Bvh8Node), 0);
    Bvh8Node* R = reinterpret_cast<Bvh8Node*>(out.nodes.data());
    for (size_t i = 0; i < wide.size(); ++i) {
        const WideTmp& w = wide[i];
        const Box& pb = nodes[w.bin].box;
        // Greedy slot assignment: cost(child, slot) = -<c - p, s(slot)>,
        // s(slot) = (+-1, +-1, +-1) from the slot's octant bits.
        const int nk = (int)w.kids.size();
        int slot_of[8], kid_in[8];
        for (int k = 0; k < 8; ++k) slot_of[k] = -1, kid_in[k] = -1;
        double cost[8][8];
        for (int k = 0; k < nk; ++k) {
            const Box& cb = nodes[w.kids[k]].box;
            for (int sl = 0; sl < 8; ++sl) {
                double c = 0;
                for (int a = 0; a < 3; ++a) {
                    const double off = 0.5 * ((double)cb.lo[a] + cb.hi[a]) - 0.5 * ((double)pb.lo[a] + pb.hi[a]);
                    c -= off * (((sl >> a) & 1) ? 1.0 : -1.0);
                }
                cost[k][sl] = c;
            }
        }
        for (int it = 0; it < nk; ++it) {
            int bk = -1, bs = -1;
            double bc = INFINITY;
            for (int k = 0; k < nk; ++k) {
                if (slot_of[k] >= 0) continue;
                for (int sl = 0; sl < 8; ++sl)
                    if (kid_in[sl] < 0 && cost[k][sl] < bc) bc = cost[k][sl], bk = k, bs = sl;
            }
            slot_of[bk] = bs;
            kid_in[bs] = bk;
        }
        Bvh8Node& r = R[i];
        std::vector<Box> kb(8);
        uint8_t valid = 0;
        for (int sl = 0; sl < 8; ++sl)
            if (kid_in[sl] >= 0) {
                kb[sl] = nodes[w.kids[kid_in[sl]]].box;
                valid |= (uint8_t)(1u << sl);
            }
        quantise(pb, kb, valid, r);
        r.valid = valid;
        for (int sl = 0; sl < 8; ++sl) {
            if (kid_in[sl] < 0) {
                r.child[sl] = NGI_BVH_EMPTY;
                continue;
            }
            const int32_t kid = w.kids[kid_in[sl]];
            r.child[sl] = w.leaf[kid_in[sl]] ? ngi_leaf_ref(col.first[kid], col.tris[kid]) : wide_index_of_bin[kid];
        }
    }
    // Depth of the wide tree bounds the traversal stack (one entry per level).
    std::vector<uint32_t> wdepth(wide.size(), 0);
    uint32_t wmax = 0;
    for (size_t i = 0; i < wide.size(); ++i)   // pre-order: parents precede children
        for (int sl = 0; sl < 8; ++sl) {
            const int32_t c = R[i].child[sl];
            if (c >= 0 && c != NGI_BVH_EMPTY) {
                wdepth[c] = wdepth[i] + 1;
                wmax = std::max(wmax, wdepth[c]);
            }
        }
    out.max_depth = wmax;
    out.num_nodes = (uint32_t)wide.size();
    out.node_bytes = sizeof(Bvh8Node);
    uint32_t leaves = 0;
    for (const WideTmp& w : wide)
        for (uint8_t f : w.leaf) leaves += f;
    out.num_leaves = leaves;
    out.sah_cost = nodes[0].count ? col.cost[0] : (double)col.cost[7] + c_node;   // root: a wide node of cost c_node
}

}  // namespace

int bvh_build(const HostScene& s, const ngi_bvh_opts& o, Bvh& out) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t nt = (uint32_t)(s.faces.size() / 3);
    out = Bvh();
    const int width = o.width ? o.width : NGI_BVH_8;
    if (width != NGI_BVH_2 && width != NGI_BVH_8) return fail(NGI_ERR_INVALID_ARG, "BVH width must be 2 or 8");
    out.width = width;
    if (nt == 0) {
        out.build_ms = 0;
        return NGI_OK;
    }
    if (nt >= (1u << 27)) return fail(NGI_ERR_INVALID_ARG, "too many triangles (%u)", nt);
    Ctx c;
    // Wide BVHs: the binary tree goes down to single triangles and the
    // optimal collapse chooses the leaves (at most max_leaf triangles each).
    const int max_leaf = o.max_leaf_size > 0 ? std::min(o.max_leaf_size, 16) : 4;
    c.max_leaf = width == NGI_BVH_8 ? 1 : max_leaf;
    c.bins = o.sah_bins > 0 ? std::min(std::max(o.sah_bins, 4), 256) : 32;
    if (o.traversal_cost < 0 || o.traversal_cost > 1e3) return fail(NGI_ERR_INVALID_ARG, "traversal_cost out of range");
    if (o.traversal_cost > 0) c.ct = o.traversal_cost;
    int threads = o.num_threads > 0 ? o.num_threads : (int)std::thread::hardware_concurrency();
    c.spare_threads = std::max(threads - 1, 0);
    c.tri_box.resize(nt);
    c.centroid.resize(3 * (size_t)nt);
    c.refs.resize(nt);
    for (uint32_t t = 0; t < nt; ++t) {
        Box bx;
        for (int k = 0; k < 3; ++k) {
            const uint32_t v = s.faces[3 * (size_t)t + k];
            const float p[3] = {(float)s.positions[3 * (size_t)v], (float)s.positions[3 * (size_t)v + 1], (float)s.positions[3 * (size_t)v + 2]};
            bx.grow(p);
        }
        c.tri_box[t] = bx;
        for (int k = 0; k < 3; ++k) c.centroid[3 * (size_t)t + k] = 0.5f * (bx.lo[k] + bx.hi[k]);
        c.refs[t] = t;
    }
    std::vector<Node> nodes;
    nodes.reserve(2 * (size_t)nt);
    build_rec(c, 0, nt, 0, nodes);
    uint32_t depth = 0;
    depth_stats(nodes, 0, 0, depth);
    double sah = 0;
    const double root_area = std::max(nodes[0].box.area(), 1e-30);
    uint32_t leaves = 0;
    for (const Node& nd : nodes) {
        if (nd.count) {
            sah += c.ci * nd.count * nd.box.area() / root_area;
            leaves++;
        } else {
            sah += c.ct * nd.box.area() / root_area;
        }
    }
    out.leaf_to_tri = c.refs;
    out.num_leaves = leaves;
    out.max_depth = depth;
    out.sah_cost = sah;
    if (width == NGI_BVH_2) {
        emit_bvh2(nodes, out);
        if (out.max_depth >= 64) return fail(NGI_ERR_INVALID_SCENE, "binary BVH depth %u exceeds the traversal stack", out.max_depth);
    } else {
        emit_bvh8(nodes, max_leaf, o.traversal_cost > 0 ? o.traversal_cost : NGI_BVH8_NODE_COST, out);
        if (out.max_depth >= 32) return fail(NGI_ERR_INVALID_SCENE, "wide BVH depth %u exceeds the traversal stack", out.max_depth);
        if (out.num_nodes >= (1u << 24)) return fail(NGI_ERR_INVALID_SCENE, "too many wide BVH nodes (%u)", out.num_nodes);
    }
    out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return NGI_OK;
}

}  // namespace ngi

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace ngi;

extern "C" {

int ngi_scene_build(ngi_scene* s, const ngi_bvh_opts* opts) {
    clear_error();
    if (!s) return fail(NGI_ERR_INVALID_ARG, "null scene");
    ngi_bvh_opts o;
    memset(&o, 0, sizeof(o));
    if (opts) o = *opts;
    device_release(*s);
    int st = bvh_build(*s, o, s->bvh);
    if (st) return st;
    s->built = true;
    return NGI_OK;
}

int ngi_scene_bvh_info(const ngi_scene* s, ngi_bvh_info* out) {
    clear_error();
    if (!s || !out) return fail(NGI_ERR_INVALID_ARG, "null argument");
    if (!s->built) return fail(NGI_ERR_INVALID_ARG, "scene is not built");
    out->width = s->bvh.width;
    out->num_nodes = s->bvh.num_nodes;
    out->num_leaves = s->bvh.num_leaves;
    out->num_triangles = (uint32_t)(s->faces.size() / 3);
    out->max_depth = s->bvh.max_depth;
    out->node_bytes = s->bvh.node_bytes;
    out->sah_cost = s->bvh.sah_cost;
    out->build_ms = s->bvh.build_ms;
    return NGI_OK;
}

int ngi_scene_bvh_export(const ngi_scene* s, void* nodes, uint32_t* leaf_to_tri) {
    clear_error();
    if (!s) return fail(NGI_ERR_INVALID_ARG, "null scene");
    if (!s->built) return fail(NGI_ERR_INVALID_ARG, "scene is not built");
    if (nodes && !s->bvh.nodes.empty()) memcpy(nodes, s->bvh.nodes.data(), s->bvh.nodes.size());
    if (leaf_to_tri && !s->bvh.leaf_to_tri.empty()) memcpy(leaf_to_tri, s->bvh.leaf_to_tri.data(), s->bvh.leaf_to_tri.size() * 4);
    return NGI_OK;
}

}  // extern "C"
