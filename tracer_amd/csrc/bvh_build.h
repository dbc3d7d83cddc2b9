// Host-side BVH construction for the device traversal kernels.
//
// Replaces BVHAccel::build (src/accelerator/bvh.cpp:78-105,191-306): the
// reference builds a binary binned-SAH tree over one axis and flattens it to
// 32-byte nodes visited through a std::stack.  Here a binned SAH binary tree
// (all three axes) is collapsed into 4-wide nodes of 128 bytes — one L2
// cache line on CDNA4 — whose child boxes are stored as six float4 lanes so a
// wavefront lane reads a whole node with eight dwordx4 loads, and triangles
// are copied into leaf order (three float4 each) so a leaf is one contiguous
// run of memory.
#pragma once
#include <cstdint>
#include <vector>

namespace trhip {

// 4-wide node.  Child c: box [lo, hi] in the SoA lanes; child[c] >= 0 is an
// inner node index, child[c] < 0 is a leaf: ~child = (first_tri << 4) | count
// with 1 <= count <= 15.  Unused slots hold an empty box (lo = +inf,
// hi = -inf) and child = -1 (a leaf of zero triangles).
struct alignas(16) Node4 {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    int32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(Node4) == 128, "Node4 must be one 128-byte line");

// Leaf-ordered triangle: v0.xyz + original primitive index (int bits),
// v1.xyz + 0, v2.xyz + 0.
struct alignas(16) TriPacked {
    float v0[3]; int32_t prim;
    float v1[3]; int32_t pad1;
    float v2[3]; int32_t pad2;
};
static_assert(sizeof(TriPacked) == 48, "TriPacked is three float4");

struct BvhBuildOptions {
    int max_leaf = 4;      // split while a node holds more triangles than this
    int threads = 0;       // 0 = hardware concurrency
};

struct BvhBuildResult {
    std::vector<Node4> nodes;
    std::vector<TriPacked> tris;
    int depth = 0;                 // max 4-wide depth
    double sah_cost = 0;           // SAH cost of the binary tree (node = 1, tri = 1)
};

// positions: 3*nv floats, indices: 3*nt ints.  Throws std::runtime_error on
// bad input.
BvhBuildResult build_bvh4(const float* positions, int64_t nv, const int32_t* indices, int64_t nt,
                          const BvhBuildOptions& opts);

}  // namespace trhip
