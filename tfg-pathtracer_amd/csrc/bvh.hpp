// Wide BVH for the MI355X traversal kernels.
//
// Replaces the reference's fixed-depth binary tree (BVH.hpp:29-62, built by
// buildAux/divideSAH at BVH.hpp:291-460): a binned-SAH binary tree with
// SAH-terminated leaves is collapsed into W-wide nodes (W = 4 or 8) whose
// child boxes are stored structure-of-arrays so one node is W*32 bytes of
// aligned 16-byte loads.  Leaf triangles are stored contiguously in leaf
// order as three float4 (v0, e1 = v1-v0, e2 = v2-v0, original index).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace eleven {

// Device node layout.  child[i] >= 0: internal node index; child[i] == -1:
// empty slot (its box is inverted so it never intersects); child[i] <= -2:
// leaf, k = -child-2, first triangle = k >> 4, triangle count = (k & 15) + 1.
template <int W>
struct alignas(16) WideNode {
    float lox[W], loy[W], loz[W];
    float hix[W], hiy[W], hiz[W];
    int32_t child[W];
    int32_t pad[W];
};
static_assert(sizeof(WideNode<4>) == 128, "BVH4 node must be 128 B");
static_assert(sizeof(WideNode<8>) == 256, "BVH8 node must be 256 B");

// Quantized wide node, the layout the device traverses.  Per axis a, child
// box planes lie on a grid: plane = p[a] + q * 2^(e[a]-127) with q an 8-bit
// integer.  p[a] is a multiple of the grid step with |p/step| + 255 < 2^24,
// so every plane p + q*step is exactly representable in FP32 and the device
// decodes it exactly (one fma).  Quantization rounds outwards (floor for lo,
// ceil for hi), so each quantized box contains the (padded) FP32 box.
//   p[3], exps (e_x | e_y << 8 | e_z << 16)          16 B
//   q[axis*2 + {0 lo, 1 hi}][W/4] bytes, child i in byte i%4 of dword i/4
//   child[W] (same encoding as WideNode)
// W = 4: 56 B used, padded to 64 (4 loads); W = 8: 96 B (6 loads).
template <int W>
struct QNode;
template <>
struct alignas(16) QNode<4> {
    float p[3];
    uint32_t exps;
    uint32_t q[6][1];
    int32_t child[4];
    int32_t pad[2];
    uint32_t plane(int as, int i) const { return (q[as][0] >> (8 * i)) & 0xffu; }
    void setPlane(int as, int i, uint32_t v) { q[as][0] |= v << (8 * i); }
};
static_assert(sizeof(QNode<4>) == 64, "quantized BVH4 node must be 64 B");
template <>
struct alignas(16) QNode<8> {
    float p[3];
    uint32_t exps;
    uint32_t q[6][2];
    int32_t child[8];
    uint32_t plane(int as, int i) const { return (q[as][i / 4] >> (8 * (i % 4))) & 0xffu; }
    void setPlane(int as, int i, uint32_t v) { q[as][i / 4] |= v << (8 * (i % 4)); }
};
static_assert(sizeof(QNode<8>) == 96, "quantized BVH8 node must be 96 B");

#ifndef EL_LEAF_MAX
#define EL_LEAF_MAX 8
#endif
constexpr int kMaxLeafTris = EL_LEAF_MAX;   // <= 16 fits the leaf encoding
constexpr int kLeafFirstLimit = 1 << 27;

inline int32_t encodeLeaf(int first, int count) { return -(2 + ((first << 4) | (count - 1))); }

struct Bvh2Node {
    float lo[3], hi[3];
    int left = -1, right = -1;   // internal: children
    int first = 0, count = 0;    // leaf: range in the leaf-order array
    bool leaf() const { return count > 0; }
};

struct Bvh2 {
    std::vector<Bvh2Node> nodes;   // node 0 is the root
    std::vector<int> order;        // leaf order -> original triangle index
};

struct WideBvh {
    int width = 4;
    std::vector<unsigned char> nodes;   // nodeCount * sizeof(QNode<W>)
    size_t nodeStride = 0;              // bytes per node as stored
    size_t nodeCount = 0;
    std::vector<float> triStream;       // 12 floats per leaf-order triangle
    std::vector<int> order;             // leaf order -> original index
    int stackBound = 0;                 // worst-case traversal stack depth
    int depth = 0;                      // wide-tree depth (root = 1)
    float sceneLo[3] = {0, 0, 0}, sceneHi[3] = {0, 0, 0};
};

// vertices: [n][3][3] floats.  Uses up to `threads` host threads (0 = the
// hardware threads, capped by OMP_NUM_THREADS when it is set).
Bvh2 buildBvh2(const float* vertices, int n, int threads = 0);
// Collapse to width 4 or 8 and pack for the device; boxes are padded
// outwards by `pad` (absolute, in scene units) so the slab test is
// conservative with respect to the Moller-Trumbore acceptance region.
WideBvh collapseWide(const Bvh2& b, const float* vertices, int n, int width);

// Quantizes one FP32 wide node (boxes already padded); returns false if a
// child box is not finite.
template <int W>
bool quantizeNode(const WideNode<W>& in, QNode<W>& out);

}  // namespace eleven
