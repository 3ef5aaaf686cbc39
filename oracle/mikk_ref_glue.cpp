// Glue that drives the reference's vendored MikkTSpace implementation
// (src/tfg-pathtracer/mikktspace.cpp, compiled where it lies by
// oracle/Makefile into oracle/_ref/) over flat triangle arrays, the way the
// reference's loader does (mikktspaceCallback.hpp:27-99, called once per
// mesh object from ObjLoader.hpp:167-168): positions, smooth-shading vertex
// normals and uvs per triangle corner in; each corner's tangent out, and per
// triangle the sign of the last corner written (Tri::tangentsSign holds one
// value per triangle and every corner overwrites it).
//
// Test infrastructure only: tests compare the library's own tangent
// generator against it.  Nothing of the reference is built into the product.
#include <mikktspace.h>

namespace {
struct Mesh {
    int ntri;
    const float *pos, *nrm, *uv;
    float *tan, *sign;
};
const Mesh& M(const SMikkTSpaceContext* c) { return *static_cast<const Mesh*>(c->m_pUserData); }
int numFaces(const SMikkTSpaceContext* c) { return M(c).ntri; }
int numVerts(const SMikkTSpaceContext*, const int) { return 3; }
void position(const SMikkTSpaceContext* c, float out[], const int f, const int v) {
    for (int k = 0; k < 3; k++) out[k] = M(c).pos[9 * f + 3 * v + k];
}
void normal(const SMikkTSpaceContext* c, float out[], const int f, const int v) {
    for (int k = 0; k < 3; k++) out[k] = M(c).nrm[9 * f + 3 * v + k];
}
void texcoord(const SMikkTSpaceContext* c, float out[], const int f, const int v) {
    out[0] = M(c).uv[6 * f + 2 * v];
    out[1] = M(c).uv[6 * f + 2 * v + 1];
}
void setBasic(const SMikkTSpaceContext* c, const float t[], const float s, const int f, const int v) {
    const Mesh& m = M(c);
    for (int k = 0; k < 3; k++) m.tan[9 * f + 3 * v + k] = t[k];
    m.sign[f] = s;
}
}  // namespace

extern "C" int mikk_ref_tangents(int ntri, const float* pos, const float* nrm, const float* uv, float* tan,
                                 float* sign) {
    Mesh m{ntri, pos, nrm, uv, tan, sign};
    SMikkTSpaceInterface iface{};
    iface.m_getNumFaces = numFaces;
    iface.m_getNumVerticesOfFace = numVerts;
    iface.m_getPosition = position;
    iface.m_getNormal = normal;
    iface.m_getTexCoord = texcoord;
    iface.m_setTSpaceBasic = setBasic;
    SMikkTSpaceContext ctx{};
    ctx.m_pInterface = &iface;
    ctx.m_pUserData = &m;
    return genTangSpaceDefault(&ctx) ? 0 : -1;
}
