// C-ABI over the host scene loader: el_scene_load / el_scene_view_get.
// The loaded eleven::Scene (Scene.hpp mirror) is flattened once into the
// structure-of-arrays el_scene_view that el_upload_scene consumes.
#include <cstring>
#include <string>
#include <vector>

#include "../../include/eleven.h"
#include "hdri_table.hpp"
#include "scene.hpp"

using namespace eleven;

struct el_scene {
    Scene scene;
    std::vector<float> vertices, uvs, normals, tangents, sign;
    std::vector<int32_t> objectId, meshMaterial;
    std::vector<el_material> materials;
    std::vector<el_texture> textures;
    std::vector<el_point_light> lights;
    el_scene_view view{};
};

static thread_local std::string g_sceneError;

static el_texture texView(const Texture& t) {
    el_texture v{};
    v.width = t.width;
    v.height = t.height;
    v.data = t.data.data();
    v.x_tile = t.xTile;
    v.y_tile = t.yTile;
    v.x_offset = t.xOffset;
    v.y_offset = t.yOffset;
    v.filter = (int32_t)t.filter;
    return v;
}

static void flatten(el_scene* s) {
    const Scene& sc = s->scene;
    size_t n = sc.tris.size();
    s->vertices.resize(n * 9);
    s->uvs.resize(n * 6);
    s->normals.resize(n * 9);
    s->tangents.resize(n * 9);
    s->sign.resize(n);
    s->objectId.resize(n);
    for (size_t i = 0; i < n; i++) {
        const Tri& t = sc.tris[i];
        for (int k = 0; k < 3; k++) {
            s->vertices[i * 9 + 3 * k + 0] = t.vertices[k].x;
            s->vertices[i * 9 + 3 * k + 1] = t.vertices[k].y;
            s->vertices[i * 9 + 3 * k + 2] = t.vertices[k].z;
            s->normals[i * 9 + 3 * k + 0] = t.normals[k].x;
            s->normals[i * 9 + 3 * k + 1] = t.normals[k].y;
            s->normals[i * 9 + 3 * k + 2] = t.normals[k].z;
            s->tangents[i * 9 + 3 * k + 0] = t.tangents[k].x;
            s->tangents[i * 9 + 3 * k + 1] = t.tangents[k].y;
            s->tangents[i * 9 + 3 * k + 2] = t.tangents[k].z;
            s->uvs[i * 6 + 2 * k + 0] = t.uv[k].x;
            s->uvs[i * 6 + 2 * k + 1] = t.uv[k].y;
        }
        s->sign[i] = t.tangentsSign;
        s->objectId[i] = t.objectID;
    }
    s->meshMaterial.clear();
    for (const MeshObject& m : sc.meshObjects) s->meshMaterial.push_back(m.materialID);
    if (s->meshMaterial.empty()) s->meshMaterial.push_back(0);
    s->materials.clear();
    for (const Material& m : sc.materials) {
        el_material e{};
        e.albedo_tex = m.albedoTextureID;
        e.emission_tex = m.emissionTextureID;
        e.roughness_tex = m.roughnessTextureID;
        e.metallic_tex = m.metallicTextureID;
        e.normal_tex = m.normalTextureID;
        e.opacity_tex = m.opacityTextureID;
        e.albedo = {m.albedo.x, m.albedo.y, m.albedo.z};
        e.emission = {m.emission.x, m.emission.y, m.emission.z};
        e.opacity = {m.opacity.x, m.opacity.y, m.opacity.z};
        e.roughness = m.roughness;
        e.metallic = m.metallic;
        e.clearcoat_gloss = m.clearcoatGloss;
        e.clearcoat = m.clearcoat;
        e.anisotropic = m.anisotropic;
        e.eta = m.eta;
        e.transmission = m.transmission;
        e.specular = m.specular;
        e.specular_tint = m.specularTint;
        e.sheen_tint = m.sheenTint;
        e.subsurface = m.subsurface;
        e.sheen = m.sheen;
        s->materials.push_back(e);
    }
    s->textures.clear();
    for (const Texture& t : sc.textures) s->textures.push_back(texView(t));
    s->lights.clear();
    for (const PointLight& l : sc.pointLights)
        s->lights.push_back({{l.position.x, l.position.y, l.position.z}, {l.radiance.x, l.radiance.y, l.radiance.z}});

    el_scene_view& v = s->view;
    v.tri_count = (uint32_t)n;
    v.vertices = s->vertices.data();
    v.uvs = s->uvs.data();
    v.normals = s->normals.data();
    v.tangents = s->tangents.data();
    v.tangent_sign = s->sign.data();
    v.object_id = s->objectId.data();
    v.mesh_count = (uint32_t)s->meshMaterial.size();
    v.mesh_material = s->meshMaterial.data();
    v.material_count = (uint32_t)s->materials.size();
    v.materials = s->materials.data();
    v.texture_count = (uint32_t)s->textures.size();
    v.textures = s->textures.data();
    v.hdri = texView(sc.hdri.texture);
    v.light_count = (uint32_t)s->lights.size();
    v.lights = s->lights.data();
    const Camera& c = sc.camera;
    v.camera.x_res = c.xRes;
    v.camera.y_res = c.yRes;
    v.camera.focal_length = c.focalLength;
    v.camera.sensor_width = c.sensorWidth;
    v.camera.sensor_height = c.sensorHeight;
    v.camera.aperture = c.aperture;
    v.camera.focus_distance = c.focusDistance;
    v.camera.rotation = {c.rotation.x, c.rotation.y, c.rotation.z};
    v.camera.position = {c.position.x, c.position.y, c.position.z};
    v.camera.bokeh = c.bokeh ? 1 : 0;
}

extern "C" {

int el_scene_load(const char* dir, el_scene** out) {
    if (!dir || !out) {
        g_sceneError = "null argument";
        return EL_ERR_ARG;
    }
    *out = nullptr;
    el_scene* s = new (std::nothrow) el_scene();
    if (!s) {
        g_sceneError = "out of memory";
        return EL_ERR_NOMEM;
    }
    try {
        s->scene = loadScene(dir);
        flatten(s);
    } catch (const LoadError& e) {
        g_sceneError = e.msg;
        delete s;
        return EL_ERR_IO;
    } catch (const std::bad_alloc&) {
        g_sceneError = "out of memory";
        delete s;
        return EL_ERR_NOMEM;
    } catch (const std::exception& e) {
        g_sceneError = e.what();
        delete s;
        return EL_ERR_IO;
    }
    *out = s;
    return EL_OK;
}

int el_scene_view_get(const el_scene* s, el_scene_view* out) {
    if (!s || !out) return EL_ERR_ARG;
    *out = s->view;
    return EL_OK;
}

int el_scene_hdri_cdf(const el_scene* s, const float** cdf, float* radianceSum) {
    if (!s || !cdf || !radianceSum) return EL_ERR_ARG;
    *cdf = s->scene.hdri.cdf.data();
    *radianceSum = s->scene.hdri.radianceSum;
    return EL_OK;
}

void el_scene_free(el_scene* s) { delete s; }

int el_hdri_lookup(const el_texture* hdri, const float* r, uint32_t n, int32_t* pixel) {
    if (!hdri || !hdri->data || hdri->width <= 0 || hdri->height <= 0 || (n && (!r || !pixel))) return EL_ERR_ARG;
    try {
        HdriSampler s = buildHdriSampler(hdri->data, hdri->width, hdri->height);
        for (uint32_t i = 0; i < n; i++) {
            if (!(r[i] >= 0.f && r[i] < 1.f)) return EL_ERR_RANGE;
            pixel[i] = hdriLookup(s.cdf.data(), s.guide.data(), s.map2.data(), s.log2G, r[i]);
        }
    } catch (const std::bad_alloc&) {
        return EL_ERR_NOMEM;
    }
    return EL_OK;
}

const char* el_scene_last_error(void) { return g_sceneError.c_str(); }

}  // extern "C"
