// kittlespt/pod_types.hpp — the reference's POD scene and settings types
// (KittlesPT/src/pod_types.hpp:9-219) with their defaults, over small local
// vector/matrix types in place of glm (column-major mat4 like glm::mat4).
#pragma once
#include <cstdint>
#include <string>
#include <string_view>
#include <vector>

namespace KittlesPT {

struct vec2 { float x = 0, y = 0; };
struct vec3 {
    float x = 0, y = 0, z = 0;
    vec3() = default;
    vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    explicit vec3(float v) : x(v), y(v), z(v) {}
};
// column-major: m[c][r]
struct mat4 {
    float m[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    float* data() { return &m[0][0]; }
    const float* data() const { return &m[0][0]; }
    float* operator[](int c) { return m[c]; }   // column c, like glm::mat4
    const float* operator[](int c) const { return m[c]; }
};

// pod_types.hpp:9-33
struct ExposureValues {
    ExposureValues() = default;
    ExposureValues(float aperture_f_num, int32_t iso, float shutter_secs, int32_t max_iso = 6400,
                   int32_t min_iso = 100, float max_shutter_secs = 0.0333f, float min_shutter_secs = 0.00025f)
        : aperture_f_num(aperture_f_num), ISO(static_cast<float>(iso)), shutter_speed_secs(shutter_secs),
          MAX_ISO(max_iso), MIN_ISO(min_iso), MIN_SHUTTER_SECS(min_shutter_secs), MAX_SHUTTER_SECS(max_shutter_secs) {}
    float aperture_f_num = 0.0f;
    float ISO = 0.0f;
    float shutter_speed_secs = 0.0f;
    int32_t MAX_ISO = 6400;
    int32_t MIN_ISO = 100;
    float MIN_SHUTTER_SECS = 0.00025f;
    float MAX_SHUTTER_SECS = 0.0333f;
};

// pod_types.hpp:37-44
struct ProceduralEnvironmentSettings {
    float sun_angular_diameter_rad = 0.0087f;
    float sun_phi_rad = 4.18879f;
    float sun_theta_rad = 17.0f * 0.017453292519943295f;
    float sun_emission_factor = 1.0f;
};

// pod_types.hpp:46-57
struct RendererSettings {
    int32_t integrator_max_ray_depth = 3;
    bool integrator_use_temporal_accumulation = false;
    bool bloom_generate_bloom = false;
    bool bloom_use_karis_average = false;
    float bloom_final_blend = 0.175f;
    float bloom_internal_blend = 0.65f;
    bool tonemapper_enable_auto_exposure = false;
    bool postprocess_enable_effects = false;
    bool upscale_enable = false;
};

// pod_types.hpp:60-123
struct MaterialSceneEntity {
    MaterialSceneEntity() = default;
    MaterialSceneEntity(std::string_view name, int32_t albedo_texture_id, const vec3& albedo, int32_t orm_texture_id,
                        float metallicity, float roughness, int32_t transmission_texture_id, float transmission,
                        float ior, int32_t emission_texture_id, const vec3& emission_factor,
                        float emission_scale_nits, int32_t normal_texture_id, float normal_scale)
        : name(name), albedo_texture_id(albedo_texture_id), albedo_factor(albedo), ORM_texture_id(orm_texture_id),
          metallic_factor(metallicity), roughness_factor(roughness), transmission_texture_id(transmission_texture_id),
          transmission_factor(transmission), ior(ior), emission_texture_id(emission_texture_id),
          emission_factor(emission_factor), emission_scale_nits(emission_scale_nits),
          normal_texture_id(normal_texture_id), normal_scale(normal_scale) {}
    std::string name;
    int32_t albedo_texture_id = -1;
    vec3 albedo_factor = vec3(1.0f);
    int32_t ORM_texture_id = -1;
    float metallic_factor = 0.0f;
    float roughness_factor = 0.5f;
    int32_t transmission_texture_id = -1;
    float transmission_factor = 0.0f;
    float ior = 1.45f;
    int32_t emission_texture_id = -1;
    vec3 emission_factor = vec3(1.0f);
    float emission_scale_nits = 0.0f;
    int32_t normal_texture_id = -1;
    float normal_scale = 1.0f;
    bool isEmissive() const {
        return emission_factor.x * emission_scale_nits > 0.0f || emission_factor.y * emission_scale_nits > 0.0f ||
               emission_factor.z * emission_scale_nits > 0.0f;
    }
};

// pod_types.hpp:125-147
struct TriangleSceneEntity {
    TriangleSceneEntity(vec3 p0, vec3 p1, vec3 p2, vec3 n0, vec3 n1, vec3 n2, vec2 t0, vec2 t1, vec2 t2,
                        int32_t material_id)
        : p0(p0), p1(p1), p2(p2), n0(n0), n1(n1), n2(n2), t0(t0), t1(t1), t2(t2), material_id(material_id) {}
    vec3 p0, p1, p2;
    vec3 n0, n1, n2;
    vec2 t0, t1, t2;
    int32_t material_id = -1;
};

// pod_types.hpp:149-161
struct MeshSceneEntity {
    MeshSceneEntity(std::string_view name, const mat4& model_transform) : name(name), model_matrix(model_transform) {}
    void addShape(const TriangleSceneEntity& shape) { shape_entities.push_back(shape); }
    std::string name;
    mat4 model_matrix;
    std::vector<TriangleSceneEntity> shape_entities;
};

// pod_types.hpp:163-175
struct TextureSceneEntity {
    TextureSceneEntity(std::string_view name, const uint8_t* data, uint32_t width, uint32_t height, uint32_t chl_count)
        : name(name), pixels_data(data, data + (size_t)width * height * chl_count), channels_count(chl_count),
          width(width), height(height) {}
    std::string name;
    std::vector<uint8_t> pixels_data;
    uint32_t channels_count = 0;
    uint32_t width = 0, height = 0;
};

// pod_types.hpp:177-185
struct CameraSceneEntity {
    CameraSceneEntity(std::string_view name, mat4 view_transform, float yfov_rads)
        : name(name), y_fov_radians(yfov_rads), view_matrix(view_transform) {}
    std::string name;
    float y_fov_radians = 0.0f;
    mat4 view_matrix;
};

// An instance of a mesh (this library's extension: the reference draws each
// mesh once, at its model matrix).
struct InstanceSceneEntity {
    int32_t mesh_index = 0;
    mat4 model_matrix;
};

// pod_types.hpp:192-219
struct BasicScene {
    void addTexture(const TextureSceneEntity& t) { texture_entities.push_back(t); }
    void addMaterial(const MaterialSceneEntity& m) { material_entities.push_back(m); }
    void addMesh(const MeshSceneEntity& m) { mesh_entities.push_back(m); }
    void addCamera(const CameraSceneEntity& c) { camera_entities.push_back(c); }
    void addInstance(const InstanceSceneEntity& i) { instance_entities.push_back(i); }
    std::vector<MaterialSceneEntity> material_entities;
    std::vector<MeshSceneEntity> mesh_entities;
    std::vector<CameraSceneEntity> camera_entities;
    std::vector<TextureSceneEntity> texture_entities;
    std::vector<InstanceSceneEntity> instance_entities;   // empty: one instance per mesh
};

}  // namespace KittlesPT
