// Wide BVH check (bvh.cpp collapseWide): decodes every stored node the way
// the traversal does (quantized planes, child codes: >= 0 inner node, <= -2
// leaf (first << 4 | count - 1), -1 empty) and verifies that
//   * the leaves reached from the root cover the leaf order exactly once,
//     and the leaf order is a permutation of the triangles;
//   * every triangle lies inside the decoded box of each child slot on its
//     path from the root (so the slab test never culls it);
//   * empty slots decode inverted (never hit);
//   * the triangle stream holds, at leaf position i, the vertices and the
//     original index of triangle order[i].
// Scenes: uniform soups, clustered soups with tiny and huge triangles, a
// single triangle, and a count that is not a multiple of the leaf size.
#include "bvh.hpp"
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>
using namespace eleven;

static int bad = 0;
#define CHECK(c, ...)                          \
    do {                                       \
        if (!(c)) {                            \
            if (bad < 10) printf(__VA_ARGS__); \
            bad++;                             \
        }                                      \
    } while (0)

struct Box {
    float lo[3], hi[3];
};

template <int W>
static void walk(const WideBvh& w, const std::vector<float>& v, int node, std::vector<Box>& path,
                 std::vector<int>& seen, int depth) {
    CHECK(depth <= w.depth, "depth %d above the reported %d\n", depth, w.depth);
    QNode<W> n;
    std::memcpy(&n, w.nodes.data() + (size_t)node * sizeof(QNode<W>), sizeof n);
    for (int i = 0; i < W; i++) {
        Box b;
        for (int a = 0; a < 3; a++) {
            const float s = std::ldexp(1.0f, (int)((n.exps >> (8 * a)) & 0xff) - 127);
            const uint32_t ql = n.plane(a * 2, i);
            const uint32_t qh = n.plane(a * 2 + 1, i);
            b.lo[a] = std::fma((float)ql, s, n.p[a]);
            b.hi[a] = std::fma((float)qh, s, n.p[a]);
        }
        const int32_t c = n.child[i];
        if (c == -1) {
            CHECK(b.lo[0] > b.hi[0] || b.lo[1] > b.hi[1] || b.lo[2] > b.hi[2], "empty slot not inverted\n");
            continue;
        }
        path.push_back(b);
        if (c >= 0) {
            CHECK((size_t)c < w.nodeCount, "inner child %d out of range\n", c);
            if ((size_t)c < w.nodeCount) walk<W>(w, v, c, path, seen, depth + 1);
        } else {
            const int k = -c - 2, first = k >> 4, count = (k & 15) + 1;
            CHECK(count <= kMaxLeafTris, "leaf of %d triangles\n", count);
            for (int t = first; t < first + count; t++) {
                CHECK(t >= 0 && (size_t)t < seen.size(), "leaf triangle %d out of range\n", t);
                if (t < 0 || (size_t)t >= seen.size()) continue;
                seen[t]++;
                const int o = w.order[t];
                for (int cc = 0; cc < 3; cc++)
                    for (int a = 0; a < 3; a++) {
                        const float x = v[(size_t)o * 9 + cc * 3 + a];
                        for (const Box& pb : path)
                            CHECK(pb.lo[a] <= x && x <= pb.hi[a], "triangle %d outside an ancestor box\n", o);
                    }
            }
        }
        path.pop_back();
    }
}

template <int W>
static void check(const std::vector<float>& v, int n) {
    Bvh2 b2 = buildBvh2(v.data(), n, 4);
    WideBvh w = collapseWide(b2, v.data(), n, W);
    CHECK(w.nodeStride == sizeof(QNode<W>), "node stride %zu\n", w.nodeStride);
    CHECK(w.order.size() == (size_t)n, "order size %zu vs %d\n", w.order.size(), n);
    std::vector<int> perm(n, 0);
    for (int o : w.order)
        if (o >= 0 && o < n) perm[o]++;
    for (int i = 0; i < n; i++) CHECK(perm[i] == 1, "triangle %d appears %d times in the leaf order\n", i, perm[i]);
    std::vector<int> seen(n, 0);
    std::vector<Box> path;
    walk<W>(w, v, 0, path, seen, 1);
    for (int i = 0; i < n; i++) CHECK(seen[i] == 1, "leaf position %d reached %d times\n", i, seen[i]);
    for (int i = 0; i < n; i++) {
        const float* s = &w.triStream[(size_t)i * 12];
        const float* t = &v[(size_t)w.order[i] * 9];
        int32_t id;
        std::memcpy(&id, &s[9], 4);
        CHECK(id == w.order[i] && s[0] == t[0] && s[1] == t[1] && s[2] == t[2] && s[3] == t[3] - t[0] &&
                  s[8] == t[8] - t[2],
              "triangle stream entry %d\n", i);
    }
}

int main() {
    std::mt19937 g(7);
    std::uniform_real_distribution<float> U(-1, 1);
    const int sizes[] = {1, 7, 1000, 20011};
    for (int n : sizes)
        for (int mode = 0; mode < 2; mode++) {
            std::vector<float> v((size_t)n * 9);
            for (int i = 0; i < n; i++) {
                float c[3], r = mode == 0 ? 0.05f : std::ldexp(1.0f, (int)(g() % 24) - 20);
                for (int a = 0; a < 3; a++) c[a] = mode == 0 ? U(g) * 10 : (float)(g() % 4) * 100 + U(g);
                for (int k = 0; k < 3; k++)
                    for (int a = 0; a < 3; a++) v[(size_t)i * 9 + k * 3 + a] = c[a] + U(g) * r;
            }
            check<4>(v, n);
            check<8>(v, n);
        }
    printf("bad=%d\n", bad);
    return bad != 0;
}
