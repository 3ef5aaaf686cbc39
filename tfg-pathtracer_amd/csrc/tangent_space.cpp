// Per-corner tangent frames of a triangle mesh by Mikkelsen's method
// (M. S. Mikkelsen, "Simulation of Wrinkled Surfaces Revisited", 2008, and
// the MikkTSpace convention used by the reference: genTangSpaceDefault over
// each mesh object, ObjLoader.hpp:167-168, mikktspaceCallback.hpp:27-99),
// written from the published algorithm:
//
//  1. Corners with bit-identical position, normal and uv are one vertex.
//     Triangles with two corners on one vertex are degenerate and set aside.
//  2. Each good triangle gets its first-order uv derivatives: with
//     d1 = p1 - p0, d2 = p2 - p0, (s, t) uv deltas t21, t31 and the signed uv
//     area A = t21.s t31.t - t21.t t31.s,
//         Os = t31.t d1 - t21.t d2,   Ot = -t31.s d1 + t21.s d2,
//     normalised and signed by sign(A) ("orientation preserving" if A > 0),
//     with magnitudes |Os| / |A|, |Ot| / |A|.  A triangle whose uv area or
//     derivatives vanish joins any group ("group with any").
//  3. Neighbours: the triangle across each edge is one that holds the same
//     two vertices in the opposite winding (edges sorted by (lower vertex,
//     higher vertex, triangle); the first unmatched candidate pairs).
//  4. Groups: around each vertex, the fan of triangles reachable through
//     shared edges and of the same orientation forms one group (depth-first
//     from the first triangle met in triangle order; an any-group triangle
//     takes the orientation of the first group that reaches it).
//  5. Each corner's frame: the derivatives of every triangle of its group
//     (those whose projected directions are not opposite to its own, i.e.
//     all of them at the default 180-degree threshold), projected onto the
//     tangent plane of the vertex normal, normalised, weighted by the angle
//     of the triangle at that vertex, summed in ascending triangle order and
//     normalised.  Corners that no group reaches keep (1, 0, 0), sign -1.
//  6. Degenerate triangles copy the frame of the first corner (in triangle
//     order) on the same vertex.
// The sign of a corner is +1 for an orientation-preserving group; the
// reference stores one sign per triangle, the last corner's.
// All arithmetic is FP32 in the order written (the build compiles host code
// without contraction).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "scene.hpp"

namespace eleven {
namespace {

struct V3 {
    float x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 scale(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float len(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline bool nonzero(float f) { return std::fabs(f) > FLT_MIN; }
inline bool nonzero(V3 a) { return nonzero(a.x) || nonzero(a.y) || nonzero(a.z); }
inline V3 unit(V3 a) { return scale(1.0f / len(a), a); }
inline V3 unitIfNonzero(V3 a) { return nonzero(a) ? unit(a) : a; }
inline V3 projectOff(V3 n, V3 a) { return sub(a, scale(dot(n, a), n)); }

constexpr int kOrientPreserving = 1, kGroupWithAny = 2;
constexpr float kPi = 3.14159265358979323846f;

struct TriInfo {
    int neighbour[3] = {-1, -1, -1};   // across edge (i, i + 1)
    int group[3] = {-1, -1, -1};
    V3 os{0, 0, 0}, ot{0, 0, 0};
    float magS = 0, magT = 0;
    int flags = kGroupWithAny;
    int face = 0;                     // original triangle index
};

struct Group {
    int vertex;
    bool orient;
    std::vector<int> tris;            // in the order they joined
};

struct Frame {
    V3 os{1, 0, 0};
    bool orient = false;
};

struct CornerKey {
    uint32_t w[8];
    bool operator==(const CornerKey& o) const { return std::memcmp(w, o.w, sizeof w) == 0; }
};
struct CornerHash {
    size_t operator()(const CornerKey& k) const {
        uint64_t h = 1469598103934665603ull;
        for (uint32_t v : k.w) h = (h ^ v) * 1099511628211ull;
        return (size_t)h;
    }
};
inline uint32_t bitsOf(float f) {
    if (f == 0.0f) f = 0.0f;          // -0 == +0 as the float comparison the method uses
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

class Mikk {
public:
    explicit Mikk(std::vector<Tri>& t) : tris(t) {}

    void run() {
        const int n = (int)tris.size();
        if (n == 0) return;
        weld();
        // good triangles first, in their order; degenerate ones after
        std::vector<int> order;
        for (int f = 0; f < n; f++)
            if (!degenerate(f)) order.push_back(f);
        const int good = (int)order.size();
        for (int f = 0; f < n; f++)
            if (degenerate(f)) order.push_back(f);
        info.assign(good, TriInfo{});
        list.assign(3 * (size_t)n, 0);
        for (int k = 0; k < n; k++)
            for (int i = 0; i < 3; i++) list[3 * (size_t)k + i] = vid[3 * (size_t)order[k] + i];
        for (int k = 0; k < good; k++) derivatives(k, order[k]);
        neighbours(good);
        buildGroups(good);
        frames.assign(3 * (size_t)n, Frame{});
        evalGroups();
        // degenerate triangles: the frame of the first corner on the same vertex
        for (int k = good; k < n; k++)
            for (int i = 0; i < 3; i++) {
                const int v = list[3 * (size_t)k + i];
                for (size_t j = 0; j < 3 * (size_t)n; j++)
                    if (list[j] == v) {
                        frames[3 * (size_t)order[k] + i] = frames[3 * (size_t)order[j / 3] + j % 3];
                        break;
                    }
            }
        for (int f = 0; f < n; f++) {
            for (int i = 0; i < 3; i++) {
                const Frame& fr = frames[3 * (size_t)f + i];
                tris[f].tangents[i] = Vector3(fr.os.x, fr.os.y, fr.os.z);
                tris[f].tangentsSign = fr.orient ? 1.0f : -1.0f;   // one per triangle: the last corner's
            }
        }
    }

private:
    std::vector<Tri>& tris;
    std::vector<int> vid;             // corner -> representative corner (vertex id)
    std::vector<int> list;            // reordered triangles' corner vertex ids
    std::vector<TriInfo> info;        // good triangles, reordered
    std::vector<Group> groups;
    std::vector<Frame> frames;        // by original corner

    V3 pos(int c) const {
        const Vector3& p = tris[c / 3].vertices[c % 3];
        return {p.x, p.y, p.z};
    }
    V3 nrm(int c) const {
        const Vector3& p = tris[c / 3].normals[c % 3];
        return {p.x, p.y, p.z};
    }
    V3 tex(int c) const {
        const Vector3& p = tris[c / 3].uv[c % 3];
        return {p.x, p.y, 1.0f};
    }
    bool degenerate(int f) const {
        const int a = vid[3 * (size_t)f], b = vid[3 * (size_t)f + 1], c = vid[3 * (size_t)f + 2];
        return a == b || a == c || b == c;
    }

    void weld() {
        const size_t nc = 3 * tris.size();
        vid.resize(nc);
        std::unordered_map<CornerKey, int, CornerHash> seen;
        seen.reserve(nc);
        for (size_t c = 0; c < nc; c++) {
            const V3 p = pos((int)c), q = nrm((int)c), t = tex((int)c);
            const float v[8] = {p.x, p.y, p.z, q.x, q.y, q.z, t.x, t.y};
            bool nan = false;
            CornerKey k;
            for (int i = 0; i < 8; i++) {
                nan |= std::isnan(v[i]);
                k.w[i] = bitsOf(v[i]);
            }
            if (nan) {   // NaN equals nothing: its own vertex
                vid[c] = (int)c;
                continue;
            }
            auto it = seen.emplace(k, (int)c).first;
            vid[c] = it->second;
        }
    }

    void derivatives(int k, int f) {
        TriInfo& ti = info[k];
        ti.face = f;
        const int* v = &list[3 * (size_t)k];
        const V3 p1 = pos(v[0]), p2 = pos(v[1]), p3 = pos(v[2]);
        const V3 t1 = tex(v[0]), t2 = tex(v[1]), t3 = tex(v[2]);
        const float t21x = t2.x - t1.x, t21y = t2.y - t1.y, t31x = t3.x - t1.x, t31y = t3.y - t1.y;
        const V3 d1 = sub(p2, p1), d2 = sub(p3, p1);
        const float area = t21x * t31y - t21y * t31x;
        const V3 os = sub(scale(t31y, d1), scale(t21y, d2));
        const V3 ot = add(scale(-t31x, d1), scale(t21x, d2));
        if (area > 0) ti.flags |= kOrientPreserving;
        if (nonzero(area)) {
            const float absArea = std::fabs(area);
            const float lenS = len(os), lenT = len(ot);
            const float s = (ti.flags & kOrientPreserving) ? 1.0f : -1.0f;
            if (nonzero(lenS)) ti.os = scale(s / lenS, os);
            if (nonzero(lenT)) ti.ot = scale(s / lenT, ot);
            ti.magS = lenS / absArea;
            ti.magT = lenT / absArea;
            if (nonzero(ti.magS) && nonzero(ti.magT)) ti.flags &= ~kGroupWithAny;
        }
    }

    // the edge of triangle k holding vertices {a, b}: its number and its two
    // vertices in the triangle's winding
    void edgeOf(int k, int a, int b, int& e, int& v0, int& v1) const {
        const int* t = &list[3 * (size_t)k];
        if (t[0] == a || t[0] == b) {
            if (t[1] == a || t[1] == b) {
                e = 0, v0 = t[0], v1 = t[1];
            } else {
                e = 2, v0 = t[2], v1 = t[0];
            }
        } else {
            e = 1, v0 = t[1], v1 = t[2];
        }
    }

    void neighbours(int good) {
        struct Edge {
            int lo, hi, tri;
        };
        std::vector<Edge> edges;
        edges.reserve(3 * (size_t)good);
        for (int k = 0; k < good; k++)
            for (int i = 0; i < 3; i++) {
                const int a = list[3 * (size_t)k + i], b = list[3 * (size_t)k + (i + 1) % 3];
                edges.push_back({std::min(a, b), std::max(a, b), k});
            }
        std::sort(edges.begin(), edges.end(), [](const Edge& x, const Edge& y) {
            if (x.lo != y.lo) return x.lo < y.lo;
            if (x.hi != y.hi) return x.hi < y.hi;
            return x.tri < y.tri;
        });
        for (size_t i = 0; i < edges.size(); i++) {
            const Edge& A = edges[i];
            int eA, a0, a1;
            edgeOf(A.tri, A.lo, A.hi, eA, a0, a1);
            if (info[A.tri].neighbour[eA] != -1) continue;
            for (size_t j = i + 1; j < edges.size() && edges[j].lo == A.lo && edges[j].hi == A.hi; j++) {
                int eB, b0, b1;
                edgeOf(edges[j].tri, edges[j].lo, edges[j].hi, eB, b1, b0);   // reversed winding
                if (a0 == b0 && a1 == b1 && info[edges[j].tri].neighbour[eB] == -1) {
                    info[A.tri].neighbour[eA] = edges[j].tri;
                    info[edges[j].tri].neighbour[eB] = A.tri;
                    break;
                }
            }
        }
    }

    // Adds triangle k to group g through its corner on the group's vertex,
    // then its two neighbours sharing that vertex.
    void join(int k, int g) {
        TriInfo& ti = info[k];
        const int vtx = groups[g].vertex;
        int i = 0;
        while (i < 3 && list[3 * (size_t)k + i] != vtx) i++;
        if (i == 3 || ti.group[i] == g || ti.group[i] != -1) return;
        if ((ti.flags & kGroupWithAny) && ti.group[0] == -1 && ti.group[1] == -1 && ti.group[2] == -1)
            ti.flags = (ti.flags & ~kOrientPreserving) | (groups[g].orient ? kOrientPreserving : 0);
        if (((ti.flags & kOrientPreserving) != 0) != groups[g].orient) return;
        groups[g].tris.push_back(k);
        ti.group[i] = g;
        const int left = ti.neighbour[i], right = ti.neighbour[i > 0 ? i - 1 : 2];
        if (left >= 0) join(left, g);
        if (right >= 0) join(right, g);
    }

    void buildGroups(int good) {
        for (int k = 0; k < good; k++)
            for (int i = 0; i < 3; i++) {
                TriInfo& ti = info[k];
                if ((ti.flags & kGroupWithAny) || ti.group[i] != -1) continue;
                const int g = (int)groups.size();
                groups.push_back({list[3 * (size_t)k + i], (ti.flags & kOrientPreserving) != 0, {k}});
                ti.group[i] = g;
                const int left = ti.neighbour[i], right = ti.neighbour[i > 0 ? i - 1 : 2];
                if (left >= 0) join(left, g);
                if (right >= 0) join(right, g);
            }
    }

    int cornerOf(int k, int vtx) const {
        for (int i = 0; i < 3; i++)
            if (list[3 * (size_t)k + i] == vtx) return i;
        return -1;
    }

    // Angle-weighted frame of the triangles `members` at vertex vtx.
    V3 evalFrame(const std::vector<int>& members, int vtx) const {
        V3 sum{0, 0, 0};
        for (int k : members) {
            const TriInfo& ti = info[k];
            if (ti.flags & kGroupWithAny) continue;
            const int i = cornerOf(k, vtx);
            const int* t = &list[3 * (size_t)k];
            const V3 n = nrm(t[i]);
            const V3 os = unitIfNonzero(projectOff(n, ti.os));
            const V3 p0 = pos(t[i > 0 ? i - 1 : 2]), p1 = pos(t[i]), p2 = pos(t[i < 2 ? i + 1 : 0]);
            const V3 e1 = unitIfNonzero(projectOff(n, sub(p0, p1)));
            const V3 e2 = unitIfNonzero(projectOff(n, sub(p2, p1)));
            float c = dot(e1, e2);
            c = c > 1 ? 1 : (c < -1 ? -1 : c);
            const float angle = std::acos(c);
            sum = add(sum, scale(angle, os));
        }
        return unitIfNonzero(sum);
    }

    void evalGroups() {
        const float thresCos = (float)std::cos((180.0f * kPi) / 180.0f);   // -1: threshold disabled
        for (const Group& G : groups) {
            std::vector<std::vector<int>> uniq;
            std::vector<V3> uniqFrame;
            for (int k : G.tris) {
                const int i = cornerOf(k, G.vertex);
                const V3 n = nrm(list[3 * (size_t)k + i]);
                const TriInfo& tk = info[k];
                const V3 os = unitIfNonzero(projectOff(n, tk.os)), ot = unitIfNonzero(projectOff(n, tk.ot));
                std::vector<int> members;
                for (int t : G.tris) {
                    const TriInfo& tt = info[t];
                    const V3 os2 = unitIfNonzero(projectOff(n, tt.os)), ot2 = unitIfNonzero(projectOff(n, tt.ot));
                    const bool any = ((tk.flags | tt.flags) & kGroupWithAny) != 0;
                    if (any || t == k || (dot(os, os2) > thresCos && dot(ot, ot2) > thresCos)) members.push_back(t);
                }
                std::sort(members.begin(), members.end());
                size_t u = 0;
                while (u < uniq.size() && uniq[u] != members) u++;
                if (u == uniq.size()) {
                    uniq.push_back(members);
                    uniqFrame.push_back(evalFrame(members, G.vertex));
                }
                Frame& out = frames[3 * (size_t)tk.face + i];
                out.os = uniqFrame[u];
                out.orient = G.orient;
            }
        }
    }
};

}  // namespace

void mikkTangents(std::vector<Tri>& tris) { Mikk(tris).run(); }

}  // namespace eleven
