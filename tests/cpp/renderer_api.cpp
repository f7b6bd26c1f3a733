// Drives KittlesPT::Renderer (include/kittlespt/renderer.hpp) the way an
// application of the reference drives KittlesPT/src/renderer.hpp: build a
// BasicScene from entities, loadScene, setRendererSettings,
// setProceduralEnvironmentData, resizeResolution, setView, executeRendering,
// then read the output.  The scene comes from a flat little-endian dump so a
// test can render the same scene through the Python binding and compare.
//
// usage: renderer_api <scene.bin> <out.bin>
// dump: i32 T M I NM NT W H spp depth seed_mode integrator,
//       f32 sun[4] (angular diameter, phi, theta, emission factor),
//       f32 positions[T*9] normals[T*9] uvs[T*6], i32 mids[T],
//       i32 mesh_first[M] mesh_count[M], f32 mesh_model[M*16],
//       i32 inst_mesh[I], f32 inst_model[I*16],
//       f32 materials[NM*18] (MaterialSceneEntity fields in declaration order,
//       texture ids stored as floats), u32 tex[NT*3] (w,h,channels),
//       u64 pixel_bytes, u8 pixels, f32 proj[16] view[16]
// out: W*H*4 f32 render estimate, then W*H*4 f32 V-buffer, then the
//      executeBatch estimates of the same view at W x H and W/2 x H/2.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "kittlespt/renderer.hpp"

using namespace KittlesPT;

static FILE* g_in;
template <class T>
static std::vector<T> rd(size_t n) {
    std::vector<T> v(n);
    if (n && fread(v.data(), sizeof(T), n, g_in) != n) {
        fprintf(stderr, "short read\n");
        exit(3);
    }
    return v;
}
#define CHECK(c)                                               \
    do {                                                       \
        if (!(c)) {                                            \
            fprintf(stderr, "check failed: %s\n", #c);         \
            return 4;                                          \
        }                                                      \
    } while (0)

int main(int argc, char** argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s scene.bin out.bin\n", argv[0]);
        return 2;
    }
    g_in = fopen(argv[1], "rb");
    if (!g_in) return 2;
    std::vector<int32_t> hdr = rd<int32_t>(11);
    const int T = hdr[0], M = hdr[1], I = hdr[2], NM = hdr[3], NT = hdr[4];
    const int W = hdr[5], H = hdr[6], spp = hdr[7], depth = hdr[8];
    std::vector<float> sun = rd<float>(4);
    std::vector<float> pos = rd<float>((size_t)T * 9), nrm = rd<float>((size_t)T * 9), uv = rd<float>((size_t)T * 6);
    std::vector<int32_t> mids = rd<int32_t>(T), mfirst = rd<int32_t>(M), mcount = rd<int32_t>(M);
    std::vector<float> mmodel = rd<float>((size_t)M * 16);
    std::vector<int32_t> imesh = rd<int32_t>(I);
    std::vector<float> imodel = rd<float>((size_t)I * 16);
    std::vector<float> mats = rd<float>((size_t)NM * 18);
    std::vector<uint32_t> tex = rd<uint32_t>((size_t)NT * 3);
    uint64_t npix = rd<uint64_t>(1)[0];
    std::vector<uint8_t> pixels = rd<uint8_t>(npix);
    std::vector<float> proj = rd<float>(16), view = rd<float>(16);
    fclose(g_in);

    BasicScene scene;
    size_t off = 0;
    for (int t = 0; t < NT; ++t) {
        const uint32_t w = tex[t * 3], h = tex[t * 3 + 1], c = tex[t * 3 + 2];
        scene.addTexture(TextureSceneEntity("tex", pixels.data() + off, w, h, c));
        off += (size_t)w * h * c;
    }
    for (int m = 0; m < NM; ++m) {
        const float* f = &mats[(size_t)m * 18];
        scene.addMaterial(MaterialSceneEntity("mat", (int32_t)f[0], vec3(f[1], f[2], f[3]), (int32_t)f[4], f[5], f[6],
                                              (int32_t)f[7], f[8], f[9], (int32_t)f[10], vec3(f[11], f[12], f[13]),
                                              f[14], (int32_t)f[15], f[16]));
    }
    for (int m = 0; m < M; ++m) {
        mat4 model;
        for (int k = 0; k < 16; ++k) model.data()[k] = mmodel[(size_t)m * 16 + k];
        MeshSceneEntity mesh("mesh", model);
        for (int t = mfirst[m]; t < mfirst[m] + mcount[m]; ++t) {
            const float* p = &pos[(size_t)t * 9];
            const float* n = &nrm[(size_t)t * 9];
            const float* u = &uv[(size_t)t * 6];
            mesh.addShape(TriangleSceneEntity(vec3(p[0], p[1], p[2]), vec3(p[3], p[4], p[5]), vec3(p[6], p[7], p[8]),
                                              vec3(n[0], n[1], n[2]), vec3(n[3], n[4], n[5]), vec3(n[6], n[7], n[8]),
                                              vec2{u[0], u[1]}, vec2{u[2], u[3]}, vec2{u[4], u[5]}, mids[t]));
        }
        scene.addMesh(mesh);
    }
    for (int i = 0; i < I; ++i) {
        InstanceSceneEntity inst;
        inst.mesh_index = imesh[i];
        for (int k = 0; k < 16; ++k) inst.model_matrix.data()[k] = imodel[(size_t)i * 16 + k];
        scene.addInstance(inst);
    }

    Renderer r;
    try {
        r.initialize(0);
        r.setSeedMode(hdr[9]);
        if (hdr[10] != KPT_INTEGRATOR_WAVEFRONT) {
            kpt_settings s;
            kpt_get_settings(r.context(), &s);
            s.integrator = hdr[10];
            kpt_set_settings(r.context(), &s);
        }
        r.loadScene(scene);
        CHECK(r.getMaterialsCount() == (size_t)NM);
        CHECK(r.getMeshCount() == (size_t)(I ? I : M));
        RendererSettings rs = r.getRendererSettings();
        rs.integrator_max_ray_depth = depth;
        r.setRendererSettings(rs);
        ProceduralEnvironmentSettings env;
        env.sun_angular_diameter_rad = sun[0];
        env.sun_phi_rad = sun[1];
        env.sun_theta_rad = sun[2];
        env.sun_emission_factor = sun[3];
        r.setProceduralEnvironmentData(env);
        // setter round trips and out-of-range behaviour of the reference API
        CHECK(r.setMaterial(0, r.getMaterial(0)));
        CHECK(!r.setMaterial((uint32_t)NM, MaterialSceneEntity()));
        CHECK(r.setMeshTransform(0, r.getMeshTransform(0)));
        CHECK(!r.setMeshTransform((uint32_t)r.getMeshCount(), mat4()));
        r.resizeResolution((uint32_t)W, (uint32_t)H);
        mat4 P, V;
        for (int k = 0; k < 16; ++k) {
            P.data()[k] = proj[k];
            V.data()[k] = view[k];
        }
        r.setView(P, V);
        r.resetAccumulation();
        for (int s = 0; s < spp; ++s) r.executeRendering(16.0f);
        std::vector<float> img = r.readOutput(), vb = r.readDebug();
        FILE* o = fopen(argv[2], "wb");
        if (!o) return 2;
        fwrite(img.data(), sizeof(float), img.size(), o);
        fwrite(vb.data(), sizeof(float), vb.size(), o);
        // the batched-camera extension: this view at W x H and at W/2 x H/2
        Renderer::CameraDesc cams[2];
        cams[0].projection = P; cams[0].view = V; cams[0].width = (uint32_t)W; cams[0].height = (uint32_t)H;
        cams[1].projection = P; cams[1].view = V; cams[1].width = (uint32_t)(W / 2); cams[1].height = (uint32_t)(H / 2);
        std::vector<float> batch = r.executeBatch(cams, 2, spp);
        CHECK(batch.size() == (size_t)(W * H + (W / 2) * (H / 2)) * 4);
        fwrite(batch.data(), sizeof(float), batch.size(), o);
        fclose(o);
        bool threw = false;
        try {
            (void)r.getMaterial((uint32_t)NM);
        } catch (const RendererError&) {
            threw = true;
        }
        CHECK(threw);
        r.shutdown();
    } catch (const RendererError& e) {
        fprintf(stderr, "RendererError %d: %s\n", e.code, e.what());
        return 5;
    }
    printf("ok %d x %d, %d spp\n", W, H, spp);
    return 0;
}
