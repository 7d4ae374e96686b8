// Host BVH8 builder: binned-SAH binary BVH, SAH-optimal collapse to 8-wide
// nodes by dynamic programming over the number of child slots a subtree may
// take (the collapse of Ylitie, Karras and Laine, HPG 2017), octant slot
// assignment and conservative 8-bit quantisation (layout in device/bvh.h).
#include "bvh_build.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>

namespace rm {
namespace {

struct Aabb {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Aabb& b) {
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    void grow(const float p[3]) {
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    double area() const {
        const double dx = std::max(0.0, (double)hi[0] - lo[0]);
        const double dy = std::max(0.0, (double)hi[1] - lo[1]);
        const double dz = std::max(0.0, (double)hi[2] - lo[2]);
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
    bool valid() const { return lo[0] <= hi[0]; }
};

struct Node2 {
    Aabb box;
    int32_t left = -1, right = -1;
    uint32_t first = 0, count = 0;
};

constexpr int kBins = 32;

struct Builder2 {
    const std::vector<Aabb>& tb;
    const std::vector<std::array<float, 3>>& cent;
    std::vector<uint32_t>& idx;
    std::vector<Node2> nodes;
    std::atomic<int> next{0};
    std::atomic<int> threads_left{0};
    uint32_t max_leaf = 3;  // primitives per leaf

    Builder2(const std::vector<Aabb>& tb_, const std::vector<std::array<float, 3>>& c, std::vector<uint32_t>& i)
        : tb(tb_), cent(c), idx(i) {}

    int alloc(int n) { return next.fetch_add(n); }

    void make_leaf(int ni, uint32_t begin, uint32_t end) {
        nodes[ni].left = nodes[ni].right = -1;
        nodes[ni].first = begin;
        nodes[ni].count = end - begin;
    }

    void build(int ni, uint32_t begin, uint32_t end) {
        Aabb box, cbox;
        for (uint32_t i = begin; i < end; i++) {
            box.grow(tb[idx[i]]);
            cbox.grow(cent[idx[i]].data());
        }
        nodes[ni].box = box;
        const uint32_t n = end - begin;
        if (n == 1) {
            make_leaf(ni, begin, end);
            return;
        }
        // binned SAH over the three axes
        double best_cost = INFINITY;
        int best_axis = -1, best_split = -1;
        const double parent_area = box.area();
        for (int a = 0; a < 3; a++) {
            const float ext = cbox.hi[a] - cbox.lo[a];
            if (!(ext > 0.0f)) continue;
            Aabb bb[kBins];
            uint32_t bc[kBins] = {0};
            const float k = (float)kBins * (1.0f - 1e-6f) / ext;
            for (uint32_t i = begin; i < end; i++) {
                int b = (int)((cent[idx[i]][a] - cbox.lo[a]) * k);
                b = std::min(std::max(b, 0), kBins - 1);
                bc[b]++;
                bb[b].grow(tb[idx[i]]);
            }
            double ra[kBins];
            uint32_t rc[kBins];
            Aabb acc;
            uint32_t cnt = 0;
            for (int b = kBins - 1; b > 0; b--) {
                acc.grow(bb[b]);
                cnt += bc[b];
                ra[b] = acc.valid() ? acc.area() : 0.0;
                rc[b] = cnt;
            }
            Aabb lacc;
            uint32_t lcnt = 0;
            for (int b = 0; b < kBins - 1; b++) {
                lacc.grow(bb[b]);
                lcnt += bc[b];
                if (lcnt == 0 || rc[b + 1] == 0) continue;
                const double cost = 1.0 + (lacc.area() * lcnt + ra[b + 1] * rc[b + 1]) / std::max(parent_area, 1e-30);
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = a;
                    best_split = b;
                }
            }
        }
        if (n <= max_leaf && (best_axis < 0 || (double)n <= best_cost)) {
            make_leaf(ni, begin, end);
            return;
        }
        uint32_t mid;
        if (best_axis < 0) {
            mid = begin + n / 2;  // coincident centroids: split the range
        } else {
            const int a = best_axis;
            const float ext = cbox.hi[a] - cbox.lo[a];
            const float k = (float)kBins * (1.0f - 1e-6f) / ext;
            auto it = std::partition(idx.begin() + begin, idx.begin() + end, [&](uint32_t t) {
                int b = (int)((cent[t][a] - cbox.lo[a]) * k);
                b = std::min(std::max(b, 0), kBins - 1);
                return b <= best_split;
            });
            mid = (uint32_t)(it - idx.begin());
            if (mid == begin || mid == end) mid = begin + n / 2;
        }
        const int l = alloc(2);
        nodes[ni].left = l;
        nodes[ni].right = l + 1;
        if (n > 65536 && threads_left.fetch_sub(1) > 0) {
            std::thread th([&, l, begin, mid] { build(l, begin, mid); });
            build(l + 1, mid, end);
            th.join();
        } else {
            build(l, begin, mid);
            build(l + 1, mid, end);
        }
    }
};

// Quantisation of one node; returns false if a child box cannot be covered.
// Decoded child boxes enclose the exact ones strictly (the grid origin sits
// one ulp below the node box): a ray parallel to an axis that lies in a face
// of a child's geometry bounds then starts strictly inside that slab, where
// the slab test's 0 * (1 / 1e-30) would otherwise close it at t = 0.
bool quantise(Bvh8Node& out, const Aabb& node_box, const Aabb* child, const bool* used) {
    out.px = std::nextafter(node_box.lo[0], -INFINITY);
    out.py = std::nextafter(node_box.lo[1], -INFINITY);
    out.pz = std::nextafter(node_box.lo[2], -INFINITY);
    const float p[3] = {out.px, out.py, out.pz};
    uint8_t e[3];
    for (int a = 0; a < 3; a++) {
        const double ext = (double)node_box.hi[a] - (double)p[a];
        int ee = 1;
        if (ext > 0.0) {
            ee = (int)std::ceil(std::log2(ext / 255.0)) + 127;
            ee = std::max(1, std::min(254, ee));
        }
        for (;; ee++) {
            if (ee > 254) return false;
            const float scale = exp_scale((uint8_t)ee);
            bool ok = true;
            for (int c = 0; c < 8 && ok; c++) {
                if (!used[c]) continue;
                const double rlo = ((double)child[c].lo[a] - p[a]) / scale;
                const double rhi = ((double)child[c].hi[a] - p[a]) / scale;
                int qlo = (int)std::floor(rlo), qhi = (int)std::ceil(rhi);
                qlo = std::max(0, std::min(255, qlo));
                qhi = std::max(0, std::min(255, qhi));
                while (qlo > 0 && decode_q(p[a], (uint8_t)qlo, scale) >= child[c].lo[a]) qlo--;
                while (qhi < 255 && decode_q(p[a], (uint8_t)qhi, scale) <= child[c].hi[a]) qhi++;
                if (decode_q(p[a], (uint8_t)qlo, scale) >= child[c].lo[a] ||
                    decode_q(p[a], (uint8_t)qhi, scale) <= child[c].hi[a]) {
                    ok = false;
                    break;
                }
                uint8_t* lo_arr = a == 0 ? out.qlo_x : (a == 1 ? out.qlo_y : out.qlo_z);
                uint8_t* hi_arr = a == 0 ? out.qhi_x : (a == 1 ? out.qhi_y : out.qhi_z);
                lo_arr[c] = (uint8_t)qlo;
                hi_arr[c] = (uint8_t)qhi;
            }
            if (ok) break;
        }
        e[a] = (uint8_t)ee;
    }
    out.ex = e[0];
    out.ey = e[1];
    out.ez = e[2];
    return true;
}

// Collapse costs: a node visit tests 8 boxes, a triangle one watertight
// test.  0.6 measured best on the 5.2M-triangle sphere field among 0.15,
// 0.3, 0.6 and 1.0 (within 2% of each other).
constexpr double kCostNode = 1.0;
constexpr double kCostTri = 0.6;

// cost[8 n + i - 1]: least SAH cost of BVH2 subtree n spread over at most i
// child slots of its BVH8 parent.  how[8 n + i - 1]: for i = 1, 0 = one leaf
// slot, 1 = one internal child; for i > 1, k in 1..i-1 = k slots to the left
// child and i - k to the right, or 0 = the best of i - 1 slots.
struct Collapse {
    std::vector<float> cost;
    std::vector<uint8_t> how;
    std::vector<uint32_t> first, count;  // primitive range of every BVH2 node
    uint32_t max_leaf = 3;               // primitives per leaf slot

    void run(const std::vector<Node2>& N2, int n2) {
        cost.assign((size_t)n2 * 8, 0.0f);
        how.assign((size_t)n2 * 8, 0);
        first.assign(n2, 0);
        count.assign(n2, 0);
        // children have larger indices than their parents
        for (int n = n2 - 1; n >= 0; n--) {
            const Node2& nd = N2[n];
            const double area = nd.box.area();
            float* c = &cost[(size_t)n * 8];
            uint8_t* h = &how[(size_t)n * 8];
            if (nd.left < 0) {
                first[n] = nd.first;
                count[n] = nd.count;
                for (int i = 0; i < 8; i++) {
                    c[i] = (float)(area * kCostTri * nd.count);
                    h[i] = 0;
                }
                continue;
            }
            first[n] = first[nd.left];
            count[n] = count[nd.left] + count[nd.right];
            const float* cl = &cost[(size_t)nd.left * 8];
            const float* cr = &cost[(size_t)nd.right * 8];
            // split of i slots between the children, i = 2..8
            double dist[9];
            uint8_t dk[9];
            for (int i = 2; i <= 8; i++) {
                dist[i] = INFINITY;
                dk[i] = 1;
                for (int k = 1; k < i; k++) {
                    const double v = (double)cl[k - 1] + (double)cr[i - k - 1];
                    if (v < dist[i]) {
                        dist[i] = v;
                        dk[i] = (uint8_t)k;
                    }
                }
            }
            const double internal = area * kCostNode + dist[8];
            const double leaf = count[n] <= max_leaf ? area * kCostTri * count[n] : INFINITY;
            c[0] = (float)std::min(internal, leaf);
            h[0] = leaf <= internal ? 0 : 1;
            for (int i = 2; i <= 8; i++) {
                if (dist[i] < (double)c[i - 2]) {
                    c[i - 1] = (float)dist[i];
                    h[i - 1] = dk[i];
                } else {
                    c[i - 1] = c[i - 2];
                    h[i - 1] = 0;
                }
            }
        }
    }

    // The child slots of BVH2 node n given i slots: (node, is_leaf).
    void unpack(const std::vector<Node2>& N2, int n, int i, std::vector<std::pair<int, bool>>& out) const {
        while (i > 1 && how[(size_t)n * 8 + i - 1] == 0) i--;
        if (i == 1) {
            out.push_back({n, N2[n].left < 0 || how[(size_t)n * 8] == 0});
            return;
        }
        const int k = how[(size_t)n * 8 + i - 1];
        unpack(N2, N2[n].left, k, out);
        unpack(N2, N2[n].right, i - k, out);
    }
};

}  // namespace

Bvh8Build build_bvh8(const std::vector<BuildTri>& tris, int threads, std::string* error, int max_leaf) {
    const auto t0 = std::chrono::steady_clock::now();
    Bvh8Build res;
    const size_t n = tris.size();
    if (n == 0) {
        // an empty tree: root with no children, every ray misses
        Bvh8Node root;
        std::memset(&root, 0, sizeof(root));
        res.nodes.push_back(root);
        res.depth = 1;
        return res;
    }
    if (n >= (1u << 31)) {
        if (error) *error = "too many triangles";
        return res;
    }
    std::vector<Aabb> tb(n);
    std::vector<std::array<float, 3>> cent(n);
    for (size_t i = 0; i < n; i++) {
        const BuildTri& t = tris[i];
        const float* vs[3] = {&t.v0.x, &t.v1.x, &t.v2.x};
        Aabb b;
        for (int v = 0; v < 3; v++) b.grow(vs[v]);
        tb[i] = b;
        for (int a = 0; a < 3; a++) cent[i][a] = 0.5f * (b.lo[a] + b.hi[a]);
    }
    std::vector<uint32_t> idx(n);
    for (size_t i = 0; i < n; i++) idx[i] = (uint32_t)i;

    Builder2 b2(tb, cent, idx);
    b2.max_leaf = (uint32_t)std::max(1, std::min(max_leaf, 3));
    b2.nodes.resize(2 * n + 1);
    int hw = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
    b2.threads_left = std::max(0, std::min(hw, 64) - 1);
    b2.next = 1;
    b2.build(0, 0, (uint32_t)n);
    res.bvh2_nodes = b2.next.load();
    const std::vector<Node2>& N2 = b2.nodes;

    // Collapse: each BVH8 node lists up to 8 BVH2 subtrees as children, as
    // the cost table distributes them.
    Collapse col;
    col.max_leaf = b2.max_leaf;
    col.run(N2, res.bvh2_nodes);
    struct Pending {
        int n2;      // BVH2 node this BVH8 node expands
        int index;   // BVH8 node index
        int depth;
    };
    std::vector<Pending> queue;
    res.nodes.reserve(n / 2 + 8);
    res.nodes.emplace_back();
    // A leaf root is wrapped: the BVH8 root gets it as its only (leaf) child.
    queue.push_back({0, 0, 1});
    size_t head = 0;
    int max_depth = 1;
    std::vector<std::pair<int, bool>> slots;
    while (head < queue.size()) {
        const Pending pd = queue[head++];
        max_depth = std::max(max_depth, pd.depth);
        int ch[8];
        bool leaf[8];
        int nc = 0;
        slots.clear();
        if (N2[pd.n2].left < 0 || (pd.n2 == 0 && col.how[0] == 0)) {
            slots.push_back({pd.n2, true});
        } else {
            // the node's own 8 slots: the best split of 8 between its children
            int best_k = 1;
            double best = INFINITY;
            const float* cl = &col.cost[(size_t)N2[pd.n2].left * 8];
            const float* cr = &col.cost[(size_t)N2[pd.n2].right * 8];
            for (int k = 1; k < 8; k++) {
                const double v = (double)cl[k - 1] + (double)cr[8 - k - 1];
                if (v < best) {
                    best = v;
                    best_k = k;
                }
            }
            col.unpack(N2, N2[pd.n2].left, best_k, slots);
            col.unpack(N2, N2[pd.n2].right, 8 - best_k, slots);
        }
        for (const auto& sl : slots) {
            ch[nc] = sl.first;
            leaf[nc] = sl.second;
            nc++;
        }
        // octant slot assignment: slot k takes the child nearest corner k
        Aabb node_box;
        for (int c = 0; c < nc; c++) node_box.grow(N2[ch[c]].box);
        double pc[3], cc[8][3];
        for (int a = 0; a < 3; a++) pc[a] = 0.5 * ((double)node_box.lo[a] + node_box.hi[a]);
        for (int c = 0; c < nc; c++)
            for (int a = 0; a < 3; a++) cc[c][a] = 0.5 * ((double)N2[ch[c]].box.lo[a] + N2[ch[c]].box.hi[a]) - pc[a];
        struct Cost { double v; int c, s; };
        std::vector<Cost> costs;
        for (int c = 0; c < nc; c++)
            for (int s = 0; s < 8; s++) {
                double v = 0;
                for (int a = 0; a < 3; a++) v -= cc[c][a] * (((s >> a) & 1) ? 1.0 : -1.0);
                costs.push_back({v, c, s});
            }
        std::sort(costs.begin(), costs.end(), [](const Cost& x, const Cost& y) {
            if (x.v != y.v) return x.v < y.v;
            if (x.c != y.c) return x.c < y.c;
            return x.s < y.s;
        });
        int slot_child[8];
        bool child_done[8] = {false}, slot_used[8] = {false};
        for (int s = 0; s < 8; s++) slot_child[s] = -1;
        for (const Cost& k : costs) {
            if (child_done[k.c] || slot_used[k.s]) continue;
            child_done[k.c] = slot_used[k.s] = true;
            slot_child[k.s] = k.c;
        }
        Bvh8Node out;
        std::memset(&out, 0, sizeof(out));
        Aabb cb[8];
        bool used[8] = {false};
        int n_internal = 0;
        for (int s = 0; s < 8; s++) {
            if (slot_child[s] < 0) continue;
            used[s] = true;
            cb[s] = N2[ch[slot_child[s]]].box;
            if (!leaf[slot_child[s]]) n_internal++;
        }
        if (!quantise(out, node_box, cb, used)) {
            if (error) *error = "BVH8 quantisation failed";
            return Bvh8Build();
        }
        out.child_base = (uint32_t)res.nodes.size();
        out.tri_base = (uint32_t)res.tris.size();
        uint32_t tri_off = 0;
        int rank = 0;
        for (int s = 0; s < 8; s++) {
            if (slot_child[s] < 0) continue;
            const int c2 = ch[slot_child[s]];
            if (!leaf[slot_child[s]]) {
                out.imask |= (uint8_t)(1u << s);
                out.meta[s] = (uint8_t)((1u << 5) | (24u + s));
                queue.push_back({c2, (int)out.child_base + rank, pd.depth + 1});
                rank++;
            } else {
                const uint32_t cnt = col.count[c2];
                out.meta[s] = (uint8_t)((((1u << cnt) - 1u) << 5) | tri_off);
                for (uint32_t t = 0; t < cnt; t++) {
                    const BuildTri& bt = tris[idx[col.first[c2] + t]];
                    BvhTri r;
                    r.v0x = bt.v0.x; r.v0y = bt.v0.y; r.v0z = bt.v0.z; r.inst = bt.inst;
                    r.v1x = bt.v1.x; r.v1y = bt.v1.y; r.v1z = bt.v1.z; r.prim = bt.prim;
                    r.v2x = bt.v2.x; r.v2y = bt.v2.y; r.v2z = bt.v2.z; r.pad = 0;
                    res.tris.push_back(r);
                }
                tri_off += cnt;
            }
        }
        for (int r = 0; r < n_internal; r++) res.nodes.emplace_back();
        res.nodes[pd.index] = out;
    }
    res.depth = max_depth;
    if (max_depth >= 60) {
        if (error) *error = "BVH8 too deep for the traversal stack";
        return Bvh8Build();
    }
    res.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return res;
}

}  // namespace rm
