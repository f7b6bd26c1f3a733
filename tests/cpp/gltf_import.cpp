// A C++ application of include/kittlespt/model_importer.hpp: imports a glTF
// file into a BasicScene and dumps it for tests/test_gltf.py to compare with
// the Python importer's scene.
//
// usage: gltf_import <file.gltf|.glb> <dump path>
// Dump (little-endian): int32 meshes, materials, textures, cameras; per mesh
// int32 triangles, float[16] model, per triangle float[24] (p0 p1 p2 n0 n1 n2
// t0 t1 t2) + int32 material; per material float[18] (albedo id, albedo rgb,
// ORM id, metallic, roughness, transmission id, transmission, ior, emission
// id, emission rgb, nits, normal id, normal scale, 0); per texture uint32
// width, height, channels + pixel bytes; per camera float[16] view + yfov.
// On a failed import it prints the importer's error and exits with 3.
#include <cstdio>
#include <vector>

#include "kittlespt/model_importer.hpp"

template <typename T>
static void put(std::FILE* f, const T& v) { std::fwrite(&v, sizeof(T), 1, f); }
static void put3(std::FILE* f, const KittlesPT::vec3& v) { put(f, v.x); put(f, v.y); put(f, v.z); }

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    KittlesPT::BasicScene scene;
    KittlesPT::ModelImporter imp;
    if (!imp.loadGLTFfromFile(argv[1], &scene)) {
        std::printf("import failed: %s\n", imp.lastError().c_str());
        return 3;
    }
    std::FILE* f = std::fopen(argv[2], "wb");
    if (!f) return 2;
    put(f, (int32_t)scene.mesh_entities.size());
    put(f, (int32_t)scene.material_entities.size());
    put(f, (int32_t)scene.texture_entities.size());
    put(f, (int32_t)scene.camera_entities.size());
    for (const auto& m : scene.mesh_entities) {
        put(f, (int32_t)m.shape_entities.size());
        std::fwrite(m.model_matrix.data(), sizeof(float), 16, f);
        for (const auto& t : m.shape_entities) {
            put3(f, t.p0); put3(f, t.p1); put3(f, t.p2);
            put3(f, t.n0); put3(f, t.n1); put3(f, t.n2);
            put(f, t.t0.x); put(f, t.t0.y); put(f, t.t1.x); put(f, t.t1.y); put(f, t.t2.x); put(f, t.t2.y);
            put(f, (int32_t)t.material_id);
        }
    }
    for (const auto& m : scene.material_entities) {
        const float v[18] = {(float)m.albedo_texture_id, m.albedo_factor.x, m.albedo_factor.y, m.albedo_factor.z,
                             (float)m.ORM_texture_id, m.metallic_factor, m.roughness_factor,
                             (float)m.transmission_texture_id, m.transmission_factor, m.ior,
                             (float)m.emission_texture_id, m.emission_factor.x, m.emission_factor.y,
                             m.emission_factor.z, m.emission_scale_nits, (float)m.normal_texture_id, m.normal_scale,
                             0.0f};
        std::fwrite(v, sizeof(float), 18, f);
    }
    for (const auto& t : scene.texture_entities) {
        put(f, t.width); put(f, t.height); put(f, t.channels_count);
        std::fwrite(t.pixels_data.data(), 1, t.pixels_data.size(), f);
    }
    for (const auto& c : scene.camera_entities) {
        std::fwrite(c.view_matrix.data(), sizeof(float), 16, f);
        put(f, c.y_fov_radians);
    }
    std::fclose(f);
    return 0;
}
