// Binned-SAH hierarchy builder (host, multi-threaded) and flattening into the
// node layouts of the traversal kernels.
//
// Cost model follows pine's BuildSAHFull (src/impl/accel/bvh.cpp:145-284):
// split cost = 1 + (A_left*N_left + A_right*N_right) / A_node, leaf cost = N.
// Pine sweeps every sorted split position with a naive quicksort, which grows
// roughly as n^1.5 and takes minutes at a million triangles; binning the
// centroids into `bins` buckets per axis gives the same kind of tree in
// O(n log n).  The hit a ray reports does not depend on the tree (closest t,
// ties to the smaller prim id), so the tree is free to differ from pine's.
#include "bvh_build.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>

namespace pg {
namespace {

struct Box {
    float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f};
    float hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    void grow(const float* l, const float* h) {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], l[k]);
            hi[k] = std::max(hi[k], h[k]);
        }
    }
    void grow_pt(const float* p) { grow(p, p); }
    float area() const {
        float d[3];
        for (int k = 0; k < 3; k++)
            d[k] = std::max(hi[k] - lo[k], 0.0f);
        return 2.0f * (d[0] * d[1] + d[0] * d[2] + d[1] * d[2]);
    }
    bool empty() const { return lo[0] > hi[0]; }
};

struct BNode {
    Box box;
    int32_t left = -1, right = -1;  // left < 0 => leaf
    int32_t start = 0, count = 0;
    int32_t kind = 0;
};

inline void centroid(const BuildPrim& p, float* c) {
    for (int k = 0; k < 3; k++)
        c[k] = (p.lo[k] + p.hi[k]) * 0.5f;
}

class Builder {
  public:
    Builder(std::vector<BuildPrim>& prims, const BuildParams& params)
        : prims_(prims), params_(params), nodes_(std::max<size_t>(2 * prims.size(), 1)) {
        int hw = (int)std::thread::hardware_concurrency();
        max_threads_ = params.threads > 0 ? params.threads : std::max(1, hw);
    }

    int32_t build() {
        active_ = 1;
        return build_range(0, (int32_t)prims_.size(), 0);
    }
    const std::vector<BNode>& nodes() const { return nodes_; }

  private:
    int32_t alloc() { return next_.fetch_add(1); }

    int32_t make_leaf(int32_t s, int32_t e, const Box& box, int kind) {
        int32_t id = alloc();
        BNode& n = nodes_[id];
        n.box = box;
        n.start = s;
        n.count = e - s;
        n.kind = kind;
        return id;
    }

    int32_t make_inner(const Box& box, int32_t s, int32_t mid, int32_t e, int depth) {
        int32_t id = alloc();
        int32_t l, r;
        bool spawn = (e - s) > 65536 && active_.load() < max_threads_;
        if (spawn) {
            active_++;
            std::thread th([&]() { l = build_range(s, mid, depth + 1); });
            r = build_range(mid, e, depth + 1);
            th.join();
            active_--;
        } else {
            l = build_range(s, mid, depth + 1);
            r = build_range(mid, e, depth + 1);
        }
        BNode& n = nodes_[id];
        n.box = box;
        n.left = l;
        n.right = r;
        return id;
    }

    int32_t median_split(int32_t s, int32_t e, const Box& box, const Box& cbox, int depth) {
        int axis = 0;
        float best = -1.0f;
        for (int k = 0; k < 3; k++)
            if (cbox.hi[k] - cbox.lo[k] > best) {
                best = cbox.hi[k] - cbox.lo[k];
                axis = k;
            }
        int32_t mid = s + (e - s) / 2;
        std::nth_element(prims_.begin() + s, prims_.begin() + mid, prims_.begin() + e,
                         [axis](const BuildPrim& a, const BuildPrim& b) {
                             float ca = a.lo[axis] + a.hi[axis], cb = b.lo[axis] + b.hi[axis];
                             return ca < cb || (ca == cb && a.id < b.id);
                         });
        return make_inner(box, s, mid, e, depth);
    }

    int32_t build_range(int32_t s, int32_t e, int depth) {
        const int32_t n = e - s;
        Box box, cbox;
        int32_t nkind1 = 0;
        for (int32_t i = s; i < e; i++) {
            box.grow(prims_[i].lo, prims_[i].hi);
            float c[3];
            centroid(prims_[i], c);
            cbox.grow_pt(c);
            nkind1 += prims_[i].kind;
        }
        const bool mixed = nkind1 != 0 && nkind1 != n;
        const int leaf_max = std::min(std::max(params_.leaf_max, 1), 16);

        if (n == 1)
            return make_leaf(s, e, box, prims_[s].kind);
        if (mixed && n <= leaf_max) {
            // Leaves hold one record kind: split by kind first.
            int32_t mid = (int32_t)(std::stable_partition(prims_.begin() + s, prims_.begin() + e,
                                                          [](const BuildPrim& p) { return p.kind == 0; }) -
                                    prims_.begin());
            return make_inner(box, s, mid, e, depth);
        }
        // Keep every root-to-leaf path within kMaxDepth levels: once the levels
        // left could not fit a balanced subtree, split at the object median.
        int need = 0;
        for (int32_t m = (n + leaf_max - 1) / leaf_max; m > 1; m = (m + 1) / 2)
            need++;
        if (depth + need >= kMaxDepth - 1) {
            if (n <= leaf_max && !mixed)
                return make_leaf(s, e, box, prims_[s].kind);
            return median_split(s, e, box, cbox, depth);
        }

        const int B = std::min(std::max(params_.bins, 4), 256);
        float best_cost = 3.4e38f;
        int best_axis = -1, best_bin = -1;
        const float area = box.area();
        for (int axis = 0; axis < 3; axis++) {
            float ext = cbox.hi[axis] - cbox.lo[axis];
            if (!(ext > 0.0f))
                continue;
            Box bb[256];
            int32_t cnt[256] = {};
            const float scale = (float)B / ext;
            for (int32_t i = s; i < e; i++) {
                float c = (prims_[i].lo[axis] + prims_[i].hi[axis]) * 0.5f;
                int b = std::min(B - 1, (int)((c - cbox.lo[axis]) * scale));
                cnt[b]++;
                bb[b].grow(prims_[i].lo, prims_[i].hi);
            }
            float right_area[256];
            int32_t right_cnt[256];
            Box acc;
            int32_t ac = 0;
            for (int b = B - 1; b > 0; b--) {
                if (!bb[b].empty())
                    acc.grow(bb[b].lo, bb[b].hi);
                ac += cnt[b];
                right_area[b] = acc.area();
                right_cnt[b] = ac;
            }
            Box lacc;
            int32_t lc = 0;
            for (int b = 0; b < B - 1; b++) {
                if (!bb[b].empty())
                    lacc.grow(bb[b].lo, bb[b].hi);
                lc += cnt[b];
                if (lc == 0 || right_cnt[b + 1] == 0)
                    continue;
                float cost = 1.0f + (lacc.area() * lc + right_area[b + 1] * right_cnt[b + 1]) / area;
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = axis;
                    best_bin = b;
                }
            }
        }
        if (best_axis < 0) {
            // All centroids coincide.
            if (n <= leaf_max && !mixed)
                return make_leaf(s, e, box, prims_[s].kind);
            return median_split(s, e, box, cbox, depth);
        }
        if (n <= leaf_max && !mixed && (float)n <= best_cost)
            return make_leaf(s, e, box, prims_[s].kind);

        const float ext = cbox.hi[best_axis] - cbox.lo[best_axis];
        const float scale = (float)B / ext;
        const float clo = cbox.lo[best_axis];
        auto mid_it = std::partition(prims_.begin() + s, prims_.begin() + e, [&](const BuildPrim& p) {
            float c = (p.lo[best_axis] + p.hi[best_axis]) * 0.5f;
            int b = std::min(B - 1, (int)((c - clo) * scale));
            return b <= best_bin;
        });
        int32_t mid = (int32_t)(mid_it - prims_.begin());
        if (mid == s || mid == e)
            return median_split(s, e, box, cbox, depth);
        return make_inner(box, s, mid, e, depth);
    }

    std::vector<BuildPrim>& prims_;
    BuildParams params_;
    std::vector<BNode> nodes_;
    std::atomic<int32_t> next_{0};
    std::atomic<int> active_{0};
    int max_threads_ = 1;
};

// Flattening ------------------------------------------------------------------

struct Flattener {
    const std::vector<BNode>& bn;
    const std::vector<BuildPrim>& prims;
    FlatBVH& out;
    float inflate;
    int width;

    int32_t leaf_code(const BNode& n) {
        const bool ana = n.kind != 0;
        std::vector<int32_t>& order = ana ? out.ana_order : out.tri_order;
        int32_t start = (int32_t)order.size();
        for (int32_t i = 0; i < n.count; i++)
            order.push_back(prims[n.start + i].id);
        return make_leaf(ana ? 1 : 0, n.count, start);
    }

    // Children of a node after collapsing to `width`: expands the inner child of
    // largest area until the node is full.
    int gather(int32_t id, int32_t* kids) {
        int cnt = 0;
        kids[cnt++] = bn[id].left;
        kids[cnt++] = bn[id].right;
        while (cnt < width) {
            int pick = -1;
            float best = -1.0f;
            for (int i = 0; i < cnt; i++) {
                const BNode& c = bn[kids[i]];
                if (c.left >= 0 && c.box.area() > best) {
                    best = c.box.area();
                    pick = i;
                }
            }
            if (pick < 0)
                break;
            int32_t expand = kids[pick];
            kids[pick] = bn[expand].left;
            kids[cnt++] = bn[expand].right;
        }
        return cnt;
    }

    void put_box(float* node, int slot, const Box& b) {
        float lo[3], hi[3];
        for (int k = 0; k < 3; k++) {
            lo[k] = b.lo[k] - inflate;
            hi[k] = b.hi[k] + inflate;
        }
        if (width == 2) {
            float* f = node + slot * 6;
            f[0] = lo[0], f[1] = lo[1], f[2] = lo[2], f[3] = hi[0], f[4] = hi[1], f[5] = hi[2];
        } else {
            node[0 + slot] = lo[0];
            node[4 + slot] = lo[1];
            node[8 + slot] = lo[2];
            node[12 + slot] = hi[0];
            node[16 + slot] = hi[1];
            node[20 + slot] = hi[2];
        }
    }
    void put_child(float* node, int slot, int32_t code) {
        int off = width == 2 ? 12 : 24;
        memcpy(node + off + slot, &code, 4);
    }

    // Emits node `id` (inner) at flat index `at`, recursing in depth-first order.
    void emit(int32_t id, int64_t at, int depth) {
        const int stride = width == 2 ? 16 : 32;
        out.max_depth = std::max(out.max_depth, depth);
        int32_t kids[8];
        int cnt = gather(id, kids);
        int32_t codes[8];
        int64_t inner_at[8];
        for (int i = 0; i < cnt; i++) {
            const BNode& c = bn[kids[i]];
            if (c.left < 0) {
                codes[i] = leaf_code(c);
                inner_at[i] = -1;
            } else {
                inner_at[i] = out.n_nodes++;
                out.nodes.resize(out.n_nodes * stride, 0.0f);
                codes[i] = (int32_t)inner_at[i];
            }
        }
        float* node = &out.nodes[at * stride];
        for (int i = 0; i < width; i++) {
            if (i < cnt) {
                put_box(node, i, bn[kids[i]].box);
                put_child(node, i, codes[i]);
            } else {
                Box empty;  // lo > hi: never hit by the slab test
                if (width == 2) {
                    float* f = node + i * 6;
                    f[0] = f[1] = f[2] = 1.0f;
                    f[3] = f[4] = f[5] = -1.0f;
                } else {
                    node[0 + i] = node[4 + i] = node[8 + i] = 1.0f;
                    node[12 + i] = node[16 + i] = node[20 + i] = -1.0f;
                }
                put_child(node, i, kEmptyChild);
            }
        }
        for (int i = 0; i < cnt; i++)
            if (inner_at[i] >= 0)
                emit(kids[i], inner_at[i], depth + 1);
    }
};

}  // namespace

FlatBVH build_bvh(std::vector<BuildPrim>& prims, const BuildParams& params, float inflate) {
    FlatBVH out;
    out.width = params.width == 4 ? 4 : 2;
    const int stride = out.width == 2 ? 16 : 32;
    if (prims.empty()) {
        out.n_nodes = 1;
        out.nodes.assign(stride, 0.0f);
        for (int i = 0; i < out.width; i++) {
            int32_t code = kEmptyChild;
            memcpy(&out.nodes[(out.width == 2 ? 12 : 24) + i], &code, 4);
        }
        return out;
    }
    Builder b(prims, params);
    int32_t root = b.build();
    std::vector<BNode> nodes = b.nodes();
    if (nodes[root].left < 0) {
        // Single leaf: give the root an inner node so traversal starts uniformly.
        BNode wrap;
        wrap.box = nodes[root].box;
        wrap.left = root;
        // Second child: an empty leaf range is never reached because its box is
        // inverted below; reuse the same leaf to keep codes valid.
        wrap.right = root;
        nodes.push_back(wrap);
        root = (int32_t)nodes.size() - 1;
        Flattener f{nodes, prims, out, inflate, out.width};
        out.n_nodes = 1;
        out.nodes.assign(stride, 0.0f);
        // Emit with only one child.
        const BNode& leaf = nodes[nodes[root].left];
        f.put_box(out.nodes.data(), 0, leaf.box);
        f.put_child(out.nodes.data(), 0, f.leaf_code(leaf));
        for (int i = 1; i < out.width; i++) {
            float* node = out.nodes.data();
            if (out.width == 2) {
                float* q = node + i * 6;
                q[0] = q[1] = q[2] = 1.0f;
                q[3] = q[4] = q[5] = -1.0f;
            } else {
                node[0 + i] = node[4 + i] = node[8 + i] = 1.0f;
                node[12 + i] = node[16 + i] = node[20 + i] = -1.0f;
            }
            f.put_child(node, i, kEmptyChild);
        }
        out.max_depth = 1;
        return out;
    }
    Flattener f{nodes, prims, out, inflate, out.width};
    out.n_nodes = 1;
    out.nodes.assign(stride, 0.0f);
    f.emit(root, 0, 1);
    return out;
}

}  // namespace pg
