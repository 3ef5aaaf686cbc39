// Binned-SAH binary BVH build (parallel over subtrees) and collapse into
// SoA wide nodes.  Own design; the reference's build (BVH.hpp:187-460) uses
// a fixed depth of 18 with 14 bins and no leaf-size termination, which makes
// every ray descend 18 levels regardless of scene size.
#include "bvh.hpp"
#include "host_parallel.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>
#include <stdexcept>
#include <thread>

namespace eleven {

namespace {

#ifndef EL_SAH_BINS
#define EL_SAH_BINS 32
#endif
#ifndef EL_SAH_NODE_COST
#define EL_SAH_NODE_COST 1.0f
#endif
constexpr int kBins = EL_SAH_BINS;
constexpr float kTraversalCost = EL_SAH_NODE_COST;   // relative to one triangle test
constexpr int kParallelThreshold = 65536;

struct Box {
    float lo[3], hi[3];
    Box() { reset(); }
    void reset() {
        for (int a = 0; a < 3; a++) {
            lo[a] = std::numeric_limits<float>::infinity();
            hi[a] = -std::numeric_limits<float>::infinity();
        }
    }
    void grow(const float* l, const float* h) {
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], l[a]);
            hi[a] = std::max(hi[a], h[a]);
        }
    }
    void growPoint(const float* p) { grow(p, p); }
    void grow(const Box& b) { grow(b.lo, b.hi); }
    float area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0) || !(dy >= 0) || !(dz >= 0)) return 0.f;
        return 2.f * (dx * dy + dx * dz + dy * dz);
    }
};

struct Prim {
    float lo[3], hi[3], c[3];
};

struct Builder {
    const std::vector<Prim>& prims;
    std::vector<int>& idx;
    std::vector<Bvh2Node>& nodes;
    std::atomic<int> nodeCount{1};
    int maxSpawnDepth;

    Builder(const std::vector<Prim>& p, std::vector<int>& i, std::vector<Bvh2Node>& n, int spawnDepth)
        : prims(p), idx(i), nodes(n), maxSpawnDepth(spawnDepth) {}

    void makeLeaf(int ni, int first, int count, const Box& b) {
        Bvh2Node& nd = nodes[ni];
        std::memcpy(nd.lo, b.lo, sizeof nd.lo);
        std::memcpy(nd.hi, b.hi, sizeof nd.hi);
        nd.first = first;
        nd.count = count;
        nd.left = nd.right = -1;
    }

    void build(int ni, int first, int count, int depth) {
        Box b, cb;
        for (int i = first; i < first + count; i++) {
            const Prim& p = prims[idx[i]];
            b.grow(p.lo, p.hi);
            cb.growPoint(p.c);
        }
        if (count <= 1) {
            makeLeaf(ni, first, count, b);
            return;
        }
        // best binned split over the three axes
        float bestCost = std::numeric_limits<float>::infinity();
        int bestAxis = -1, bestBin = -1;
        for (int a = 0; a < 3; a++) {
            float ext = cb.hi[a] - cb.lo[a];
            if (!(ext > 0)) continue;
            float scale = kBins / ext;
            Box bb[kBins];
            int bc[kBins] = {0};
            for (int i = first; i < first + count; i++) {
                const Prim& p = prims[idx[i]];
                int k = std::min(kBins - 1, (int)((p.c[a] - cb.lo[a]) * scale));
                bc[k]++;
                bb[k].grow(p.lo, p.hi);
            }
            float rightArea[kBins];
            int rightCount[kBins];
            Box acc;
            int cnt = 0;
            for (int k = kBins - 1; k > 0; k--) {
                acc.grow(bb[k]);
                cnt += bc[k];
                rightArea[k] = acc.area();
                rightCount[k] = cnt;
            }
            acc.reset();
            cnt = 0;
            for (int k = 1; k < kBins; k++) {
                acc.grow(bb[k - 1]);
                cnt += bc[k - 1];
                if (cnt == 0 || rightCount[k] == 0) continue;
                float cost = acc.area() * cnt + rightArea[k] * rightCount[k];
                if (cost < bestCost) {
                    bestCost = cost;
                    bestAxis = a;
                    bestBin = k;
                }
            }
        }
        float parentArea = b.area();
        float leafCost = (float)count;
        float splitCost = parentArea > 0 ? kTraversalCost + bestCost / parentArea : kTraversalCost + count;
        int mid;
        if (bestAxis < 0) {
            // all centroids coincide: leaf if it fits, else split the range
            if (count <= kMaxLeafTris) {
                makeLeaf(ni, first, count, b);
                return;
            }
            mid = first + count / 2;
        } else {
            if (count <= kMaxLeafTris && leafCost <= splitCost) {
                makeLeaf(ni, first, count, b);
                return;
            }
            float ext = cb.hi[bestAxis] - cb.lo[bestAxis];
            float scale = kBins / ext;
            float lo = cb.lo[bestAxis];
            auto it = std::partition(idx.begin() + first, idx.begin() + first + count, [&](int pi) {
                int k = std::min(kBins - 1, (int)((prims[pi].c[bestAxis] - lo) * scale));
                return k < bestBin;
            });
            mid = (int)(it - idx.begin());
            if (mid == first || mid == first + count) mid = first + count / 2;
        }
        int l = nodeCount.fetch_add(2);
        int r = l + 1;
        Bvh2Node& nd = nodes[ni];
        std::memcpy(nd.lo, b.lo, sizeof nd.lo);
        std::memcpy(nd.hi, b.hi, sizeof nd.hi);
        nd.left = l;
        nd.right = r;
        nd.count = 0;
        if (depth < maxSpawnDepth && count > kParallelThreshold) {
            std::thread t([&, l, first, mid, depth] { build(l, first, mid - first, depth + 1); });
            build(r, mid, first + count - mid, depth + 1);
            t.join();
        } else {
            build(l, first, mid - first, depth + 1);
            build(r, mid, first + count - mid, depth + 1);
        }
    }
};

}  // namespace

Bvh2 buildBvh2(const float* v, int n, int threads) {
    Bvh2 out;
    std::vector<Prim> prims((size_t)n);
    for (int i = 0; i < n; i++) {
        const float* t = v + (size_t)i * 9;
        Prim& p = prims[i];
        for (int a = 0; a < 3; a++) {
            p.lo[a] = std::min(t[a], std::min(t[3 + a], t[6 + a]));
            p.hi[a] = std::max(t[a], std::max(t[3 + a], t[6 + a]));
            p.c[a] = 0.5f * (p.lo[a] + p.hi[a]);
        }
    }
    out.order.resize((size_t)n);
    for (int i = 0; i < n; i++) out.order[i] = i;
    out.nodes.resize((size_t)std::max(1, 2 * n));
    if (threads <= 0) threads = (int)hostThreads();
    int spawnDepth = 0;
    while ((1 << spawnDepth) < threads) spawnDepth++;
    Builder bld(prims, out.order, out.nodes, spawnDepth);
    if (n == 0) {
        Box e;
        bld.makeLeaf(0, 0, 0, e);
    } else {
        bld.build(0, 0, n, 0);
    }
    out.nodes.resize((size_t)bld.nodeCount.load());
    return out;
}

namespace {

template <int W>
void packWide(const Bvh2& b, const float pad, WideBvh& out) {
    std::vector<WideNode<W>> wn;
    wn.reserve(b.nodes.size() / 2 + 1);
    struct Item {
        int bvh2;        // internal BVH2 node that this wide node expands
        int wide;        // index of the wide node
        int stackDepth;  // stack entries pushed by the ancestors
        int depth;
    };
    auto emptyNode = [] {
        WideNode<W> n;
        for (int i = 0; i < W; i++) {
            n.lox[i] = n.loy[i] = n.loz[i] = std::numeric_limits<float>::infinity();
            n.hix[i] = n.hiy[i] = n.hiz[i] = -std::numeric_limits<float>::infinity();
            n.child[i] = -1;
            n.pad[i] = 0;
        }
        return n;
    };
    wn.push_back(emptyNode());
    std::deque<Item> q;
    int maxStack = 0, maxDepth = 1;
    const Bvh2Node& root = b.nodes[0];
    if (root.left < 0) {   // the root is a leaf (or the scene is empty)
        if (root.count > 0) {
            WideNode<W>& n = wn[0];
            n.lox[0] = root.lo[0] - pad;
            n.loy[0] = root.lo[1] - pad;
            n.loz[0] = root.lo[2] - pad;
            n.hix[0] = root.hi[0] + pad;
            n.hiy[0] = root.hi[1] + pad;
            n.hiz[0] = root.hi[2] + pad;
            n.child[0] = encodeLeaf(root.first, root.count);
        }
    } else {
        q.push_back({0, 0, 0, 1});
    }
    while (!q.empty()) {
        Item it = q.front();
        q.pop_front();
        maxDepth = std::max(maxDepth, it.depth);
        // gather up to W children by repeatedly opening the largest internal child
        int kids[W] = {b.nodes[it.bvh2].left, b.nodes[it.bvh2].right};
        int nk = 2;
        while (nk < W) {
            int best = -1;
            float bestA = -1;
            for (int k = 0; k < nk; k++) {
                const Bvh2Node& c = b.nodes[kids[k]];
                if (c.leaf()) continue;
                Box bx;
                bx.grow(c.lo, c.hi);
                float a = bx.area();
                if (a > bestA) {
                    bestA = a;
                    best = k;
                }
            }
            if (best < 0) break;
            int open = kids[best];
            kids[best] = b.nodes[open].left;
            kids[nk++] = b.nodes[open].right;
        }
        int sd = it.stackDepth + (nk - 1);
        maxStack = std::max(maxStack, sd);
        WideNode<W> node = emptyNode();
        for (int k = 0; k < nk; k++) {
            const Bvh2Node& c = b.nodes[kids[k]];
            node.lox[k] = c.lo[0] - pad;
            node.loy[k] = c.lo[1] - pad;
            node.loz[k] = c.lo[2] - pad;
            node.hix[k] = c.hi[0] + pad;
            node.hiy[k] = c.hi[1] + pad;
            node.hiz[k] = c.hi[2] + pad;
            if (c.leaf()) {
                if (c.first >= kLeafFirstLimit) throw std::runtime_error("too many triangles for the leaf encoding");
                node.child[k] = encodeLeaf(c.first, c.count);
            } else {
                int wi = (int)wn.size();
                wn.push_back(emptyNode());
                node.child[k] = wi;
                q.push_back({kids[k], wi, sd, it.depth + 1});
            }
        }
        wn[it.wide] = node;
    }
    out.nodeCount = wn.size();
    out.nodeStride = sizeof(QNode<W>);
    std::vector<QNode<W>> qn(wn.size());
    std::atomic<bool> finite{true};
    parallelFor(wn.size(), [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; i++)
            if (!quantizeNode<W>(wn[i], qn[i])) finite = false;
    });
    if (!finite) throw std::runtime_error("non-finite bounding box (vertex data)");
    out.nodes.resize(qn.size() * sizeof(QNode<W>));
    std::memcpy(out.nodes.data(), qn.data(), out.nodes.size());
    out.stackBound = std::max(1, maxStack);
    out.depth = maxDepth;
}

}  // namespace

template <int W>
bool quantizeNode(const WideNode<W>& in, QNode<W>& out) {
    std::memset(&out, 0, sizeof out);
    const float* lo[3] = {in.lox, in.loy, in.loz};
    const float* hi[3] = {in.hix, in.hiy, in.hiz};
    uint8_t qlo[3][W], qhi[3][W];
    for (int a = 0; a < 3; a++) {
        double L = std::numeric_limits<double>::infinity(), H = -L;
        for (int i = 0; i < W; i++) {
            if (in.child[i] == -1) continue;
            if (!std::isfinite(lo[a][i]) || !std::isfinite(hi[a][i])) return false;
            L = std::min(L, (double)lo[a][i]);
            H = std::max(H, (double)hi[a][i]);
        }
        int e = 0;
        double p = 0, step = 1;
        if (L <= H) {
            // smallest step with the extent inside 253 steps, then coarser
            // until the grid origin's multiple stays below 2^24 - 256
            double ext = std::max(H - L, 1e-30);
            e = std::max(-126, (int)std::ceil(std::log2(ext / 253.0)));
            for (;; e++) {
                if (e > 127) return false;
                step = std::ldexp(1.0, e);
                double k = std::floor(L / step);
                if (std::fabs(k) + 256.0 >= 16777216.0) continue;
                p = k * step;
                if (std::ceil((H - p) / step) > 255.0) continue;
                break;
            }
        }
        out.p[a] = (float)p;   // exact: |p / step| < 2^24
        out.exps |= (uint32_t)(e + 127) << (8 * a);
        for (int i = 0; i < W; i++) {
            if (in.child[i] == -1 || !(L <= H)) {
                qlo[a][i] = 255;   // inverted: never intersected
                qhi[a][i] = 0;
                continue;
            }
            double ql = std::floor(((double)lo[a][i] - p) / step);
            double qh = std::ceil(((double)hi[a][i] - p) / step);
            qlo[a][i] = (uint8_t)std::min(255.0, std::max(0.0, ql));
            qhi[a][i] = (uint8_t)std::min(255.0, std::max(0.0, qh));
        }
    }
    for (int a = 0; a < 3; a++)
        for (int i = 0; i < W; i++) {
            out.setPlane(a * 2 + 0, i, qlo[a][i]);
            out.setPlane(a * 2 + 1, i, qhi[a][i]);
        }
    for (int i = 0; i < W; i++) out.child[i] = in.child[i];
    return true;
}

template bool quantizeNode<4>(const WideNode<4>&, QNode<4>&);
template bool quantizeNode<8>(const WideNode<8>&, QNode<8>&);

WideBvh collapseWide(const Bvh2& b, const float* v, int n, int width) {
    if (width != 4 && width != 8) throw std::runtime_error("BVH width must be 4 or 8");
    WideBvh out;
    out.width = width;
    for (int a = 0; a < 3; a++) {
        out.sceneLo[a] = std::numeric_limits<float>::infinity();
        out.sceneHi[a] = -std::numeric_limits<float>::infinity();
    }
    for (size_t i = 0; i < (size_t)n * 3; i++)
        for (int a = 0; a < 3; a++) {
            out.sceneLo[a] = std::min(out.sceneLo[a], v[i * 3 + a]);
            out.sceneHi[a] = std::max(out.sceneHi[a], v[i * 3 + a]);
        }
    float diag = 0, mag = 0;
    if (n > 0) {
        for (int a = 0; a < 3; a++) {
            float e = out.sceneHi[a] - out.sceneLo[a];
            diag += e * e;
            mag = std::max(mag, std::max(std::fabs(out.sceneLo[a]), std::fabs(out.sceneHi[a])));
        }
        diag = std::sqrt(diag);
    }
    // Outward padding: covers slab-test rounding and the Moller-Trumbore
    // acceptance region (hits accepted slightly outside the exact triangle).
    const float pad = std::max(diag * (1.0f / 65536.0f), mag * (1.0f / 1048576.0f)) + 1e-30f;
    if (width == 4)
        packWide<4>(b, pad, out);
    else
        packWide<8>(b, pad, out);
    out.order = b.order;
    out.triStream.assign((size_t)n * 12, 0.f);
    parallelFor((size_t)n, [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; i++) {
            int o = b.order[i];
            const float* t = v + (size_t)o * 9;
            float* s = &out.triStream[i * 12];
            s[0] = t[0];
            s[1] = t[1];
            s[2] = t[2];
            s[3] = t[3] - t[0];
            s[4] = t[4] - t[1];
            s[5] = t[5] - t[2];
            s[6] = t[6] - t[0];
            s[7] = t[7] - t[1];
            s[8] = t[8] - t[2];
            int32_t oi = o;
            std::memcpy(&s[9], &oi, 4);
        }
    });
    return out;
}

}  // namespace eleven
