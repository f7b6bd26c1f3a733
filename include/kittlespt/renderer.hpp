// kittlespt/renderer.hpp — KittlesPT::Renderer (KittlesPT/src/renderer.hpp:13-58)
// as a header-only C++ class over the C-ABI of libkpt_hip.so (kpt_abi.h).
//
// Same method names, arguments and side effects as the reference (setters
// reset accumulation as renderer.cu:223, 264-266, 293, 304, 351-353 do).
// GL textures are replaced by host readback (readOutput / readDebug).  Errors:
// methods that the reference declares void throw KittlesPT::RendererError
// (the reference calls exit(99), error_check.cu:4-13); bool setters return
// false on a bad index as the reference does.
#pragma once
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../kpt_abi.h"
#include "pod_types.hpp"

namespace KittlesPT {

class RendererError : public std::runtime_error {
public:
    RendererError(int code, const std::string& msg) : std::runtime_error(msg), code(code) {}
    int code;
};

class Renderer {
public:
    // renderer.cu:22-46
    void initialize(int device = -1) {
        if (m_ctx) return;
        check(kpt_ctx_create(device, &m_ctx), "kpt_ctx_create");
        kpt_get_settings(m_ctx, &m_settings);
        m_settings.integrator = KPT_INTEGRATOR_WAVEFRONT;
        check(kpt_set_settings(m_ctx, &m_settings), "kpt_set_settings");
    }
    // renderer.cu:48-53
    void shutdown() {
        if (m_ctx) kpt_ctx_destroy(m_ctx);
        m_ctx = nullptr;
    }
    ~Renderer() { shutdown(); }

    // renderer.cu:55-101: nothing happens when the output size and the render
    // size (output / 2.5 with upscale_enable) are already these
    void resizeResolution(uint32_t width, uint32_t height) {
        const bool up = m_rs.upscale_enable;
        const int32_t rw = up ? (int32_t)((float)width / 2.5f) : (int32_t)width;
        const int32_t rh = up ? (int32_t)((float)height / 2.5f) : (int32_t)height;
        int32_t res[4] = {0, 0, 0, 0};
        check(kpt_get_resolution(m_ctx, res), "kpt_get_resolution");
        if ((int32_t)width == res[2] && (int32_t)height == res[3] && rw == res[0] && rh == res[1]) return;
        check(kpt_resize(m_ctx, (int32_t)width, (int32_t)height), "kpt_resize");
        m_width = width;
        m_height = height;
    }

    // renderer.cu:103-195: one path-traced sample per pixel per call, then
    // modulation, auto exposure, bloom and the AgX post pass
    void executeRendering(float delta_time_ms) {
        check(kpt_set_frame_delta(m_ctx, delta_time_ms), "kpt_set_frame_delta");
        check(kpt_render(m_ctx, 1), "kpt_render");
    }

    // replaces getRenderTargetTexture(GLuint) (renderer.cu:197-200): the
    // post-processed output (non-linear sRGB, RGBA32F), row-major,
    // width*height*4 floats
    std::vector<float> readOutput() const { return read(4); }
    // the modulated linear render buffer (render_texture)
    std::vector<float> readRender() const { return read(0); }
    // replaces getDebugRenderTargetTexture (renderer.cu:202-205): the debug
    // texture the path-trace kernel writes, (velocity.x, velocity.y, 0, 1)
    // per pixel (kernels.cu:447), taken from the V-buffer's velocity
    std::vector<float> readDebug() const {
        std::vector<float> v = read(3);
        for (size_t i = 0; i + 3 < v.size(); i += 4) {
            v[i + 2] = 0.0f;
            v[i + 3] = 1.0f;
        }
        return v;
    }
    // the V-buffer itself (velocity xy, primitive and instance ids as floats)
    std::vector<float> readVBuffer() const { return read(3); }

    size_t getMaterialsCount() const { return m_materials.size(); }
    size_t getMeshCount() const { return m_models.size(); }

    // renderer.cu:356-436 (+ setMeshTransform for every instance)
    void loadScene(const BasicScene& s) {
        std::vector<float> pos, nrm, uv;
        std::vector<int32_t> mids, first, count;
        std::vector<float> mesh_model;
        for (const MeshSceneEntity& m : s.mesh_entities) {
            first.push_back((int32_t)mids.size());
            count.push_back((int32_t)m.shape_entities.size());
            mesh_model.insert(mesh_model.end(), m.model_matrix.data(), m.model_matrix.data() + 16);
            for (const TriangleSceneEntity& t : m.shape_entities) {
                const vec3* P[3] = {&t.p0, &t.p1, &t.p2};
                const vec3* N[3] = {&t.n0, &t.n1, &t.n2};
                const vec2* T[3] = {&t.t0, &t.t1, &t.t2};
                for (int v = 0; v < 3; ++v) {
                    pos.insert(pos.end(), {P[v]->x, P[v]->y, P[v]->z});
                    nrm.insert(nrm.end(), {N[v]->x, N[v]->y, N[v]->z});
                    uv.insert(uv.end(), {T[v]->x, T[v]->y});
                }
                mids.push_back(t.material_id);
            }
        }
        std::vector<kpt_material> mats;
        for (const MaterialSceneEntity& m : s.material_entities) mats.push_back(toAbi(m));
        std::vector<kpt_texture_desc> texs;
        std::vector<uint8_t> pixels;
        for (const TextureSceneEntity& t : s.texture_entities) {
            kpt_texture_desc d{t.width, t.height, t.channels_count, 0, (uint64_t)pixels.size()};
            texs.push_back(d);
            pixels.insert(pixels.end(), t.pixels_data.begin(), t.pixels_data.end());
        }
        std::vector<int32_t> inst_mesh;
        std::vector<float> inst_model;
        for (const InstanceSceneEntity& i : s.instance_entities) {
            inst_mesh.push_back(i.mesh_index);
            inst_model.insert(inst_model.end(), i.model_matrix.data(), i.model_matrix.data() + 16);
        }
        kpt_scene_desc d{};
        d.num_triangles = (int32_t)mids.size();
        d.positions = pos.data();
        d.normals = nrm.data();
        d.uvs = uv.data();
        d.material_ids = mids.data();
        d.num_meshes = (int32_t)first.size();
        d.mesh_first = first.data();
        d.mesh_count = count.data();
        d.mesh_model = mesh_model.data();
        d.num_instances = (int32_t)inst_mesh.size();
        d.instance_mesh = inst_mesh.data();
        d.instance_model = inst_model.data();
        d.num_materials = (int32_t)mats.size();
        d.materials = mats.data();
        d.num_textures = (int32_t)texs.size();
        d.textures = texs.data();
        d.pixels = pixels.data();
        d.num_pixel_bytes = pixels.size();
        check(kpt_upload_scene(m_ctx, &d), "kpt_upload_scene");
        m_materials = s.material_entities;
        m_models.clear();
        if (s.instance_entities.empty())
            for (const MeshSceneEntity& m : s.mesh_entities) m_models.push_back(m.model_matrix);
        else
            for (const InstanceSceneEntity& i : s.instance_entities) m_models.push_back(i.model_matrix);
    }

    // renderer.cu:207-226
    bool setMaterial(uint32_t index, const MaterialSceneEntity& material) {
        if (index >= m_materials.size()) return false;
        kpt_material m = toAbi(material);
        if (kpt_set_material(m_ctx, (int32_t)index, &m) != KPT_OK) return false;
        m_materials[index] = material;
        return true;
    }
    // renderer.cu:228-248 (bounds-checked: throws instead of reading garbage)
    MaterialSceneEntity getMaterial(uint32_t index) const {
        if (index >= m_materials.size()) throw RendererError(KPT_ERR_INVALID_ARG, "material index out of range");
        MaterialSceneEntity m = m_materials[index];
        m.name = "unpreserved_name";
        return m;
    }

    // renderer.cu:250-269.  Matrices: KittlesPT::mat4 or any column-major
    // 4x4 type indexed m[column][row] over contiguous floats (glm::mat4).
    template <class M>
    bool setMeshTransform(uint32_t index, const M& model) {
        if (index >= m_models.size()) return false;
        mat4 m = toMat4(model);
        if (kpt_set_instance_transform(m_ctx, (int32_t)index, m.data()) != KPT_OK) return false;
        m_models[index] = m;
        return true;
    }
    template <class M = mat4>
    M getMeshTransform(uint32_t index) const {
        if (index >= m_models.size()) throw RendererError(KPT_ERR_INVALID_ARG, "mesh index out of range");
        M out;
        std::memcpy(&out[0][0], m_models[index].data(), 16 * sizeof(float));
        return out;
    }

    // renderer.cu:290-299
    void setProceduralEnvironmentData(const ProceduralEnvironmentSettings& e) {
        m_env = e;
        push();
    }
    ProceduralEnvironmentSettings getProceduralEnvironmentData() const { return m_env; }

    // renderer.cu:301-310
    void setRendererSettings(const RendererSettings& s) {
        m_rs = s;
        push();
    }
    RendererSettings getRendererSettings() const { return m_rs; }

    // renderer.cu:342-354
    template <class M>
    void setView(const M& projection, const M& view) {
        mat4 p = toMat4(projection), v = toMat4(view);
        check(kpt_set_view(m_ctx, p.data(), v.data()), "kpt_set_view");
    }

    // renderer.cu:312-329
    void setExposure(const ExposureValues& v, float ev_comp, float white_point_ev, float black_point_ev) {
        kpt_exposure e{v.aperture_f_num, v.ISO, v.shutter_speed_secs, v.MAX_ISO, v.MIN_ISO,
                       v.MIN_SHUTTER_SECS, v.MAX_SHUTTER_SECS, ev_comp, white_point_ev, black_point_ev};
        check(kpt_set_exposure(m_ctx, &e), "kpt_set_exposure");
    }
    // renderer.cu:331-334: the auto-exposure program's current values
    ExposureValues getExposure() const {
        kpt_exposure e{};
        check(kpt_get_exposure(m_ctx, &e), "kpt_get_exposure");
        ExposureValues v;
        v.aperture_f_num = e.aperture_f_num;
        v.ISO = e.iso;
        v.shutter_speed_secs = e.shutter_speed_secs;
        v.MAX_ISO = e.max_iso;
        v.MIN_ISO = e.min_iso;
        v.MIN_SHUTTER_SECS = e.min_shutter_secs;
        v.MAX_SHUTTER_SECS = e.max_shutter_secs;
        return v;
    }

    // renderer.cu:336-340
    void resetAccumulation() { check(kpt_reset_accumulation(m_ctx), "kpt_reset_accumulation"); }

    // extensions of this library
    kpt_ctx* context() const { return m_ctx; }

    // Batched cameras (no reference counterpart): n cameras of any
    // resolutions rendered as one set of paths, `spp` samples each; returns
    // every camera's per-pixel estimate (RGBA32F), cameras back to back.  The
    // renderer's own frame, accumulation and post chain are untouched.
    struct CameraDesc {
        mat4 projection;
        mat4 view;
        uint32_t width = 0, height = 0;
    };
    std::vector<float> executeBatch(const CameraDesc* cams, size_t n, int spp = 1) {
        std::vector<kpt_view> v(n);
        size_t total = 0;
        for (size_t k = 0; k < n; ++k) {
            std::memcpy(v[k].projection, cams[k].projection.data(), 16 * sizeof(float));
            std::memcpy(v[k].view, cams[k].view.data(), 16 * sizeof(float));
            v[k].width = (int32_t)cams[k].width;
            v[k].height = (int32_t)cams[k].height;
            total += (size_t)cams[k].width * cams[k].height;
        }
        std::vector<float> out(total * 4);
        check(kpt_render_views_host(m_ctx, v.data(), (int32_t)n, spp, out.data()), "kpt_render_views_host");
        return out;
    }
    void setSeedMode(int mode) { m_settings.seed_mode = mode; push(); }
    // The reference chooses its upscaler at compile time (USE_FSR,
    // kernels.cuh:3; off by default): with upscale_enable, true selects FSR
    // 1.0 (EASU + RCAS) instead of the bilinear upscale.
    void setFsrUpscale(bool fsr) { m_settings.upscale_fsr = fsr ? 1 : 0; push(); }

private:
    // buffers 0-3 at the render size, 4 (output) at the output size
    std::vector<float> read(int which) const {
        int32_t res[4] = {0, 0, 0, 0};
        check(kpt_get_resolution(m_ctx, res), "kpt_get_resolution");
        const size_t n = which == 4 ? (size_t)res[2] * res[3] : (size_t)res[0] * res[1];
        std::vector<float> out(n * 4);
        check(kpt_readback(m_ctx, which, out.data(), (int64_t)(out.size() * sizeof(float))), "kpt_readback");
        return out;
    }
    template <class M>
    static mat4 toMat4(const M& src) {
        mat4 m;
        std::memcpy(m.data(), &src[0][0], 16 * sizeof(float));
        return m;
    }
    static kpt_material toAbi(const MaterialSceneEntity& m) {
        kpt_material k{};
        k.albedo_texture_id = m.albedo_texture_id;
        k.albedo[0] = m.albedo_factor.x; k.albedo[1] = m.albedo_factor.y; k.albedo[2] = m.albedo_factor.z;
        k.orm_texture_id = m.ORM_texture_id;
        k.metallic = m.metallic_factor;
        k.roughness = m.roughness_factor;
        k.transmission_texture_id = m.transmission_texture_id;
        k.transmission = m.transmission_factor;
        k.ior = m.ior;
        k.emission_texture_id = m.emission_texture_id;
        k.emission_factor[0] = m.emission_factor.x; k.emission_factor[1] = m.emission_factor.y;
        k.emission_factor[2] = m.emission_factor.z;
        k.emission_scale_nits = m.emission_scale_nits;
        k.normal_texture_id = m.normal_texture_id;
        k.normal_scale = m.normal_scale;
        return k;
    }
    void push() {
        m_settings.max_depth = m_rs.integrator_max_ray_depth;
        m_settings.use_temporal_accumulation = m_rs.integrator_use_temporal_accumulation ? 1 : 0;
        m_settings.bloom_generate_bloom = m_rs.bloom_generate_bloom ? 1 : 0;
        m_settings.bloom_use_karis_average = m_rs.bloom_use_karis_average ? 1 : 0;
        m_settings.bloom_final_blend = m_rs.bloom_final_blend;
        m_settings.bloom_internal_blend = m_rs.bloom_internal_blend;
        m_settings.enable_auto_exposure = m_rs.tonemapper_enable_auto_exposure ? 1 : 0;
        m_settings.postprocess_enable_effects = m_rs.postprocess_enable_effects ? 1 : 0;
        m_settings.upscale_enable = m_rs.upscale_enable ? 1 : 0;
        m_settings.sun_angular_diameter_rad = m_env.sun_angular_diameter_rad;
        m_settings.sun_phi_rad = m_env.sun_phi_rad;
        m_settings.sun_theta_rad = m_env.sun_theta_rad;
        m_settings.sun_emission_factor = m_env.sun_emission_factor;
        check(kpt_set_settings(m_ctx, &m_settings), "kpt_set_settings");
    }
    void check(int rc, const char* what) const {
        if (rc != KPT_OK) throw RendererError(rc, std::string(what) + ": " + kpt_last_error(m_ctx));
    }

    kpt_ctx* m_ctx = nullptr;
    kpt_settings m_settings{};
    RendererSettings m_rs;
    ProceduralEnvironmentSettings m_env;
    uint32_t m_width = 0, m_height = 0;
    std::vector<MaterialSceneEntity> m_materials;
    std::vector<mat4> m_models;
};

}  // namespace KittlesPT
