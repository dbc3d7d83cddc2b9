// refcpu — TEST INFRASTRUCTURE ONLY (see ref_math.hpp header).
//
// The reference's binary BVH, src/accelerator/bvh.cpp:
//   * build: binned SAH on the centroid axis of maximum extent, 12 buckets,
//     cost 1 + (nL*SA_L + nR*SA_R)/SA, split at the cheapest bucket boundary
//     iff count > max_leaf (no leaf-vs-split cost test), <=2 prims split at
//     the median with nth_element, degenerate centroid bounds -> leaf
//     (bvh.cpp:191-288); depth-first flattening into 32-byte nodes, left child
//     stored next (bvh.cpp:290-306).
//   * traversal: stack of node indices, near child first by dir_is_neg[axis],
//     slab test Bounds3::intersect_p with the (1 + 1e-7) widening and
//     "tMin < ray.t_max && tMax > 0" acceptance (geometry.hpp:837-864),
//     every primitive of a leaf tested (bvh.cpp:112-185).
#pragma once
#include "ref_math.hpp"
#include <vector>
#include <stack>
#include <algorithm>

namespace refcpu {

struct Bounds3 {
    V3f lo{std::numeric_limits<float>::max(), std::numeric_limits<float>::max(), std::numeric_limits<float>::max()};
    V3f hi{std::numeric_limits<float>::lowest(), std::numeric_limits<float>::lowest(), std::numeric_limits<float>::lowest()};
    void grow(const V3f& p) { lo = vmin(lo, p); hi = vmax(hi, p); }
    void grow(const Bounds3& b) { lo = vmin(lo, b.lo); hi = vmax(hi, b.hi); }
    V3f diag() const { return hi - lo; }
    float area() const { V3f d = diag(); return 2 * (d.x * d.y + d.x * d.z + d.y * d.z); }
    int max_extent() const {
        V3f d = diag();
        if (d.x > d.y && d.x > d.z) return 0;
        if (d.y > d.z) return 1;
        return 2;
    }
    V3f offset(const V3f& p) const {
        V3f o = p - lo;
        if (hi.x > lo.x) o.x /= hi.x - lo.x;
        if (hi.y > lo.y) o.y /= hi.y - lo.y;
        if (hi.z > lo.z) o.z /= hi.z - lo.z;
        return o;
    }
    const V3f& operator[](int i) const { return i == 0 ? lo : hi; }
};

// bvh.cpp:51-64 — 32-byte flattened node.
struct RefNode {
    Bounds3 b;
    int offset;          // first primitive (leaf) or second child (interior)
    uint16_t count;      // > 0 means leaf
    uint8_t axis;
    uint8_t pad;
};
static_assert(sizeof(RefNode) == 32, "reference node layout is 32 bytes");

class RefBVH {
public:
    std::vector<RefNode> nodes;
    std::vector<int> prims;   // reordered primitive indices
    int max_leaf = 3;

    void build(const std::vector<Bounds3>& prim_bounds, int max_leaf_prims) {
        max_leaf = max_leaf_prims;
        nodes.clear();
        prims.clear();
        size_t n = prim_bounds.size();
        if (!n) return;
        std::vector<Info> info(n);
        for (size_t i = 0; i < n; ++i) {
            info[i].index = int(i);
            info[i].b = prim_bounds[i];
            info[i].c = prim_bounds[i].lo * 0.5f + prim_bounds[i].hi * 0.5f;
        }
        build_nodes.clear();
        build_nodes.reserve(2 * n);
        prims.reserve(n);
        int root = recursive_build(info, 0, n);
        nodes.resize(build_nodes.size());
        size_t off = 0;
        flatten(root, off);
    }

    template <class HitFn>   // HitFn(prim, ...) -> bool, may shrink tmax
    bool traverse(const V3f& o, const V3f& d, float& tmax, HitFn&& hit_fn, bool any_hit,
                  uint64_t* nodes_visited = nullptr, uint64_t* prims_tested = nullptr) const {
        if (nodes.empty()) return false;
        // bvh.cpp:115,152 — reciprocal in double, stored as float
        V3f inv(float(1.0 / d.x), float(1.0 / d.y), float(1.0 / d.z));
        int neg[3] = {inv.x < 0, inv.y < 0, inv.z < 0};
        bool hit = false;
        std::stack<size_t> s;
        s.push(0);
        while (!s.empty()) {
            size_t ni = s.top();
            s.pop();
            const RefNode& nd = nodes[ni];
            if (nodes_visited) ++*nodes_visited;
            if (!slab(nd.b, o, inv, neg, tmax)) continue;
            if (nd.count > 0) {
                for (int i = 0; i < nd.count; ++i) {
                    if (prims_tested) ++*prims_tested;
                    if (hit_fn(prims[nd.offset + i], tmax)) {
                        hit = true;
                        if (any_hit) return true;
                    }
                }
            } else {
                if (neg[nd.axis]) { s.push(ni + 1); s.push(size_t(nd.offset)); }
                else { s.push(size_t(nd.offset)); s.push(ni + 1); }
            }
        }
        return hit;
    }

    // geometry.hpp:837-864
    static bool slab(const Bounds3& b, const V3f& o, const V3f& inv, const int neg[3], float t_max) {
        float tMin = (b[neg[0]].x - o.x) * inv.x;
        float tMax = (b[1 - neg[0]].x - o.x) * inv.x;
        float tyMin = (b[neg[1]].y - o.y) * inv.y;
        float tyMax = (b[1 - neg[1]].y - o.y) * inv.y;
        tMax = float(tMax * (1 + 0.0000001));
        tyMax = float(tyMax * (1 + 0.0000001));
        if (tMin > tyMax || tyMin > tMax) return false;
        if (tyMin > tMin) tMin = tyMin;
        if (tyMax < tMax) tMax = tyMax;
        float tzMin = (b[neg[2]].z - o.z) * inv.z;
        float tzMax = (b[1 - neg[2]].z - o.z) * inv.z;
        tzMax = float(tzMax * (1 + 0.0000001));
        if (tMin > tzMax || tzMin > tMax) return false;
        if (tzMin > tMin) tMin = tzMin;
        if (tzMax < tMax) tMax = tzMax;
        return (tMin < t_max) && (tMax > 0);
    }

private:
    struct Info { int index; Bounds3 b; V3f c; };
    struct BNode { Bounds3 b; int left = -1, right = -1, axis = 0; int first = 0, count = 0; };
    std::vector<BNode> build_nodes;

    int recursive_build(std::vector<Info>& info, size_t start, size_t end) {
        int id = int(build_nodes.size());
        build_nodes.emplace_back();
        Bounds3 bounds;
        for (size_t i = start; i < end; ++i) bounds.grow(info[i].b);
        size_t count = end - start;
        auto make_leaf = [&]() {
            BNode& nd = build_nodes[id];
            nd.first = int(prims.size());
            for (size_t i = start; i < end; ++i) prims.push_back(info[i].index);
            nd.count = int(count);
            nd.b = bounds;
        };
        if (count == 1) { make_leaf(); return id; }
        Bounds3 cb;
        for (size_t i = start; i < end; ++i) cb.grow(info[i].c);
        size_t mid = (start + end) >> 1;
        int dim = cb.max_extent();
        if (cb.hi[dim] == cb.lo[dim]) { make_leaf(); return id; }
        if (count <= 2) {
            std::nth_element(info.begin() + start, info.begin() + mid, info.begin() + end,
                             [dim](const Info& a, const Info& b) { return a.c[dim] < b.c[dim]; });
        } else {
            constexpr int NB = 12;
            struct Bucket { size_t count = 0; Bounds3 b; } bk[NB];
            for (size_t i = start; i < end; ++i) {
                int b = int(NB * cb.offset(info[i].c)[dim]);
                if (b == NB) b = NB - 1;
                bk[b].count++;
                bk[b].b.grow(info[i].b);
            }
            float cost[NB - 1];
            for (int i = 0; i < NB - 1; ++i) {
                Bounds3 lb, rb;
                int lc = 0, rc = 0;
                for (int j = 0; j <= i; ++j) { lb.grow(bk[j].b); lc += int(bk[j].count); }
                for (int j = i + 1; j < NB; ++j) { rb.grow(bk[j].b); rc += int(bk[j].count); }
                cost[i] = 1 + (lc * lb.area() + rc * rb.area()) / bounds.area();
            }
            int minb = int(std::min_element(cost, cost + NB - 1) - cost);
            if (count > size_t(max_leaf)) {
                auto pm = std::partition(info.begin() + start, info.begin() + end, [&](const Info& pi) {
                    int b = int(NB * cb.offset(pi.c)[dim]);
                    if (b == NB) b = NB - 1;
                    return b <= minb;
                });
                mid = size_t(pm - info.begin());
            } else {
                make_leaf();
                return id;
            }
        }
        int l = recursive_build(info, start, mid);
        int r = recursive_build(info, mid, end);
        BNode& nd = build_nodes[id];
        nd.left = l; nd.right = r; nd.axis = dim;
        nd.b = build_nodes[l].b;
        nd.b.grow(build_nodes[r].b);
        nd.count = 0;
        return id;
    }

    size_t flatten(int bi, size_t& off) {
        const BNode& bn = build_nodes[bi];
        size_t my = off++;
        RefNode& ln = nodes[my];
        ln.b = bn.b;
        ln.pad = 0;
        if (bn.count > 0) {
            ln.count = uint16_t(bn.count);
            ln.offset = bn.first;
            ln.axis = 0;
        } else {
            ln.axis = uint8_t(bn.axis);
            ln.count = 0;
            flatten(bn.left, off);
            ln.offset = int(flatten(bn.right, off));
        }
        return my;
    }
};

}  // namespace refcpu
