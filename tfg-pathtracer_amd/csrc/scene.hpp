// Host scene model: the reference's Scene / Camera / Material / PointLight /
// Tri / MeshObject / Texture / HDRI structs, field for field, as plain host
// C++ (no device code here).  Field names and defaults follow
//   Scene.hpp:21-33, Camera.hpp:10-23, Material.hpp:8-37, PointLight.hpp:6-18,
//   Tri.hpp:10-19, MeshObject.hpp:10-19, Texture.hpp:12-33, HDRI.hpp:10-16
// so code written against the reference's scene API reads the same here.
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace eleven {

constexpr float kPi = 3.14159265358979323846f;  // Math.hpp:7

struct Vector3 {
    float x = 0, y = 0, z = 0;
    Vector3() = default;
    Vector3(float a, float b, float c) : x(a), y(b), z(c) {}
    explicit Vector3(float a) : x(a), y(a), z(a) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

enum Filter { NO_FILTER = 0, BILINEAR = 1 };   // Texture.hpp:8
enum CS { LINEAR = 0, sRGB = 1 };              // Texture.hpp:9

struct Texture {
    std::vector<float> data;   // RGB FP32, data[3*(y*width+x)+c]
    Filter filter = NO_FILTER;
    Vector3 color;
    std::string path;
    int width = 1;
    int height = 1;
    float xTile = 1, yTile = 1;
    float xOffset = 0, yOffset = 0;
};

struct HDRI {
    Texture texture;
    std::vector<float> cdf;    // width*height + 1 entries
    float radianceSum = 0;
};

struct Camera {
    unsigned int xRes = 1280;
    unsigned int yRes = 720;
    float focalLength = 35 * 0.001f;
    float sensorWidth = 35 * 0.001f;
    float sensorHeight = 35 * 0.001f * (720.0f / 1280.0f);
    float aperture = 2.8f;
    float focusDistance = 1000000;
    Vector3 rotation;
    bool bokeh = false;
    Vector3 position;

    Camera() = default;
    Camera(unsigned int xr, unsigned int yr) : xRes(xr), yRes(yr) {
        setSensorWidth(sensorWidth);
    }
    void setSensorWidth(float size) {   // Camera.hpp:31-33
        sensorHeight = size * ((float)yRes / (float)xRes);
    }
};

struct Material {
    std::string name;
    int albedoTextureID = -1;
    int emissionTextureID = -1;
    int roughnessTextureID = -1;
    int metallicTextureID = -1;
    int normalTextureID = -1;
    int opacityTextureID = -1;
    Vector3 albedo = Vector3(0.5f, 0.5f, 0.5f);
    Vector3 emission = Vector3(0.f);
    Vector3 opacity = Vector3(1.f);
    float roughness = 1;
    float metallic = 0;
    float clearcoatGloss = 0;
    float clearcoat = 0;
    float anisotropic = 0;
    float eta = 0;
    float transmission = 0;
    float specular = 0.5f;
    float specularTint = 0;
    float sheenTint = 0;
    float subsurface = 0;
    float sheen = 0;
};

struct PointLight {
    Vector3 position;
    Vector3 radiance;
};

struct Tri {
    Vector3 vertices[3];
    Vector3 uv[3];
    Vector3 normals[3];
    Vector3 tangents[3];
    float tangentsSign = 0;
    int objectID = 0;
};

// Mikkelsen tangent frames of one mesh object's triangles (tangent_space.cpp):
// sets every corner's tangent and each triangle's tangentsSign.
void mikkTangents(std::vector<Tri>& tris);

struct MeshObject {
    std::string name;
    std::string matName;
    unsigned int triFirst = 0;   // offset into Scene::tris (the reference keeps a Tri*)
    unsigned int triCount = 0;
    int materialID = 0;
    int objectID = 0;
};

struct Scene {
    std::vector<Material> materials;
    std::vector<MeshObject> meshObjects;
    std::vector<Texture> textures;
    std::vector<Tri> tris;
    std::vector<PointLight> pointLights;
    HDRI hdri;
    Camera camera;

    int materialCount() const { return (int)materials.size(); }
    int textureCount() const { return (int)textures.size(); }
    int meshObjectCount() const { return (int)meshObjects.size(); }
    int triCount() const { return (int)tris.size(); }
    int pointLightCount() const { return (int)pointLights.size(); }

    // Scene.hpp:75-93: the mesh gets the next object id and its tris are
    // appended with that id.
    void addMeshObject(MeshObject mo, const std::vector<Tri>& meshTris);
};

// ---- loading (SceneLoader.hpp / ObjLoader.hpp / Texture.hpp / HDRI.hpp) ----
// All loaders throw eleven::LoadError with a message on failure.
struct LoadError {
    std::string msg;
};

Scene loadScene(const std::string& dir);
Texture loadTexture(const std::string& path, CS colorSpace);
// The value an 8-bit texel byte i loads as, per colour space (stb_image.h
// ldr-to-hdr as modified by the reference: fastPow(i / 255.0f, gamma)).
void ldrTable(CS colorSpace, float out[256]);
HDRI loadHDRIFile(const std::string& path);
HDRI makeConstantHDRI(Vector3 color);
// HDRI::generateCDF (HDRI.hpp:107-128): sequential FP32 sums, row-major.
void generateCDF(HDRI& h);
// Texture::getValueFromCoordinates (Texture.hpp:95-106) with tiling/offset.
Vector3 texelAt(const Texture& t, int x, int y);

// Ankerl's approximate pow (published bit trick, 2012): replace the high
// 32 bits of the IEEE-754 double by b*(hi - 1072632447) + 1072632447 and
// zero the low word.  The reference converts LDR texels with it.
double fastPow(double a, double b);

// Images: 8-bit RGBA writers used by the CLI (main.cpp:150-165 saves the
// clamped, gamma-encoded beauty pass).
bool writePNG(const std::string& path, int w, int h, const unsigned char* rgba);
bool writeBMP(const std::string& path, int w, int h, const unsigned char* rgba);

}  // namespace eleven
