// Host-side hierarchy construction: binned-SAH binary tree, optionally
// collapsed to 4-wide nodes, then flattened into the HBM layouts the traversal
// kernels read.
#pragma once

#include <cstdint>
#include <vector>

namespace pg {

struct BuildPrim {
    float lo[3], hi[3];
    int32_t id;    // global prim id
    int32_t kind;  // 0 = triangle record, 1 = analytic record
};

struct BuildParams {
    int width = 2;      // 2 or 4
    int leaf_max = 4;   // prims per leaf, 1..8
    int bins = 32;      // SAH bins per axis
    int threads = 0;    // 0 = hardware concurrency
};

// A leaf reference packs (kind, count, first record) into a negative int32:
//   bit31 = 1, bit30 = kind, bits 26..29 = count-1, bits 0..25 = first record.
constexpr int32_t kLeafBit = int32_t(0x80000000u);
constexpr int32_t kLeafAnalytic = 0x40000000;
constexpr int kLeafCountShift = 26;
constexpr int32_t kLeafStartMask = (1 << 26) - 1;
constexpr int32_t kEmptyChild = 0x7fffffff;  // unused slot of a wide node

// Levels of the binary tree (root = 1); bounds the traversal stack.
constexpr int kMaxDepth = 32;

inline int32_t make_leaf(int kind, int count, int32_t start) {
    return kLeafBit | (kind ? kLeafAnalytic : 0) | ((count - 1) << kLeafCountShift) | start;
}

// Flattened hierarchy ready for upload.
struct FlatBVH {
    int width = 2;
    std::vector<float> nodes;        // width==2: 16 floats/node; width==4: 32 floats/node
    std::vector<int32_t> tri_order;  // record slot -> global prim id (triangles)
    std::vector<int32_t> ana_order;  // record slot -> global prim id (analytic)
    int max_depth = 0;
    int64_t n_nodes = 0;
};

// Builds over prims (reordered in place).  `inflate` is the absolute amount every
// child box is grown by so that slab tests in float stay conservative.
FlatBVH build_bvh(std::vector<BuildPrim>& prims, const BuildParams& params, float inflate);

}  // namespace pg
