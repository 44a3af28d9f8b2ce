// HipPathIntegrator.cpp — flattens SceneView / GeometryPool / materials / lights into the
// ahip_scene_desc of include/anacapa_hip.h and maps ahip status codes onto the IIntegrator
// contract (renderFrame returns false on failure so RenderSession renders tiles instead,
// RenderSession.cpp:546-547).  The merge into a Film lives in HipPathIntegratorFilm.cpp.
#include "HipPathIntegrator.h"

#include <anacapa/accel/IAccelerationStructure.h>
#include <anacapa/shading/ILight.h>
#include <anacapa/shading/IMaterial.h>
#include <anacapa/shading/ShadingContext.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/anacapa_hip.h"
#include "shading/Lambertian.h"
#include "shading/StandardSurface.h"
#include "shading/lights/AreaLight.h"
#include "shading/lights/DirectionalLight.h"
#include "shading/lights/DomeLight.h"

namespace anacapa {

namespace {

void put3(float* d, Vec3f v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }

ShadingContext probeCtx() {
    SurfaceInteraction si;
    si.n = si.ng = {0, 0, 1};
    return ShadingContext(si, {0, 0, 1});
}

// Material flattening (the same dynamic_cast walk as CudaPathIntegrator.cu:42-105).
bool toAhip(const IMaterial* mat, ahip_material& m) {
    m = ahip_material{};
    m.roughness_texture = -1;
    ShadingContext ctx = probeCtx();
    if (auto* em = dynamic_cast<const EmissiveMaterial*>(mat)) {
        m.type = AHIP_MAT_EMISSIVE;
        put3(m.albedo, em->reflectance(ctx));
        put3(m.emission, em->Le(ctx, {0, 0, 1}));
        return true;
    }
    if (auto* ss = dynamic_cast<const StandardSurfaceMaterial*>(mat)) {
        const auto& p = ss->params();
        if (p.base_color.hasTexture() || p.metalness.hasTexture() || p.roughness.hasTexture() ||
            p.specular.hasTexture() || p.has_normal_map || p.opacity.hasTexture())
            return false;   // textures are served by the CPU integrator
        m.type = AHIP_MAT_STANDARD;
        put3(m.albedo, p.base_color.value);
        put3(m.emission, p.emission_color.value);
        m.emission_weight = p.emission;
        m.base = p.base;
        m.metalness = p.metalness.value;
        m.roughness = p.roughness.value;
        m.specular = p.specular.value;
        put3(m.specular_color, p.specular_color);
        m.specular_ior = p.specular_IOR;
        m.coat = p.coat;
        m.coat_roughness = p.coat_roughness;
        m.transmission = p.transmission;           // glass (StandardSurface.h:248-347)
        m.alpha_mask = p.alphaMask ? 1u : 0u;      // constant opacity (StandardSurface.h:197-217)
        m.opacity = p.opacity.value;
        m.opacity_texture = -1;
        return true;
    }
    if (dynamic_cast<const LambertianMaterial*>(mat)) {
        m.type = AHIP_MAT_LAMBERT;
        put3(m.albedo, mat->reflectance(ctx));
        return true;
    }
    return false;
}

}  // namespace

HipPathIntegrator::HipPathIntegrator(int device, uint32_t maxDepth, uint32_t integrator, float fireflyClamp)
    : m_maxDepth(maxDepth), m_integrator(integrator), m_fireflyClamp(fireflyClamp) {
    if (ahip_create(device, &m_ctx) != AHIP_OK)
        std::fprintf(stderr, "[warn] HipPathIntegrator: %s\n", ahip_last_error(m_ctx));
}

HipPathIntegrator::~HipPathIntegrator() { ahip_destroy(m_ctx); }

bool HipPathIntegrator::isValid() const { return ahip_is_valid(m_ctx) != 0; }
bool HipPathIntegrator::prepared() const {
    std::lock_guard<std::mutex> lk(m_mutex);
    return m_prepared;
}
std::string HipPathIntegrator::name() const { return ahip_device_name(m_ctx); }
std::string HipPathIntegrator::lastError() const { return ahip_last_error(m_ctx); }

void HipPathIntegrator::clearAccum() {
    std::lock_guard<std::mutex> lk(m_mutex);
    ahip_clear_accum(m_ctx);
}

void HipPathIntegrator::prepare(const SceneView& scene) {
    std::lock_guard<std::mutex> lk(m_mutex);
    m_prepared = false;
    if (!isValid() || !scene.accel) return;
    const GeometryPool& pool = scene.accel->pool();
    std::vector<uint32_t> vOff{0}, iOff{0}, nCnt, uCnt;
    std::vector<int32_t> meshMat;
    std::vector<float> pos, nrm, uvs;
    std::vector<uint32_t> idx;
    std::vector<ahip_material> mats;
    for (uint32_t m = 0; m < pool.numMeshes(); ++m) {
        const MeshDesc& mesh = pool.mesh(m);
        if (mesh.hasMotion()) return;   // motion blur stays on the CPU path
        for (const Vec3f& p : mesh.positions) pos.insert(pos.end(), {p.x, p.y, p.z});
        nrm.resize(pos.size(), 0.f);
        uvs.resize(pos.size() / 3 * 2, 0.f);
        for (size_t k = 0; k < mesh.normals.size() && k < mesh.positions.size(); ++k)
            put3(&nrm[3 * (vOff.back() + k)], mesh.normals[k]);
        for (size_t k = 0; k < mesh.uvs.size() && k < mesh.positions.size(); ++k) {
            uvs[2 * (vOff.back() + k)] = mesh.uvs[k].x;
            uvs[2 * (vOff.back() + k) + 1] = mesh.uvs[k].y;
        }
        nCnt.push_back((uint32_t)std::min(mesh.normals.size(), mesh.positions.size()));
        uCnt.push_back((uint32_t)std::min(mesh.uvs.size(), mesh.positions.size()));
        idx.insert(idx.end(), mesh.indices.begin(), mesh.indices.end());
        vOff.push_back((uint32_t)(pos.size() / 3));
        iOff.push_back((uint32_t)idx.size());
        const IMaterial* mat = m < scene.materials.size() ? scene.materials[m] : nullptr;
        if (!mat) {
            meshMat.push_back(-1);
            continue;
        }
        ahip_material am;
        if (!toAhip(mat, am)) return;
        meshMat.push_back((int32_t)mats.size());
        mats.push_back(am);
    }
    std::vector<ahip_light> lights;
    int32_t envIndex = -1;
    const DomeLight* dome = nullptr;
    for (const ILight* l : scene.lights) {
        ahip_light al{};
        if (auto* area = dynamic_cast<const AreaLight*>(l)) {
            // centre and half-axes recovered from sampleLe's uniform quad parameterisation
            LightLeSample c = area->sampleLe({0.5f, 0.5f}, {0.5f, 0.5f});
            LightLeSample u1 = area->sampleLe({1.f, 0.5f}, {0.5f, 0.5f});
            LightLeSample v1 = area->sampleLe({0.5f, 1.f}, {0.5f, 0.5f});
            al.type = AHIP_LIGHT_AREA;
            put3(al.center, c.pos);
            put3(al.u_axis, u1.pos - c.pos);
            put3(al.v_axis, v1.pos - c.pos);
            put3(al.Le, c.Le);
        } else if (auto* d = dynamic_cast<const DomeLight*>(l)) {
            if (d != scene.envLight) return;
            dome = d;
            al.type = AHIP_LIGHT_DOME;
            al.intensity = d->intensity();
            Vec3f r0, r1, r2;
            d->getRotation(r0, r1, r2);
            put3(al.rot, r0); put3(al.rot + 3, r1); put3(al.rot + 6, r2);
            envIndex = (int32_t)lights.size();
            // the scene sphere of sampleLe (DomeLight.h:142-168): at uPos = (0, 0) the origin is the
            // disk's centre, centre + envDir * radius, and at (1, 0) one radius from it along t
            for (float v : {0.5f, 0.25f, 0.75f, 0.125f}) {
                const LightLeSample a = d->sampleLe({0.f, 0.f}, {0.5f, v});
                const LightLeSample b = d->sampleLe({1.f, 0.f}, {0.5f, v});
                if (!(a.pdfPos > 0.f) || !(a.pdfDir > 0.f)) continue;
                const float radius = (b.pos - a.pos).length();
                put3(al.scene_center, a.pos + a.dir * radius);   // dir = -envDir
                al.scene_radius = radius;
                break;
            }
        } else if (auto* dl = dynamic_cast<const DirectionalLight*>(l)) {
            // direction and radiance from its own sample at the cone's axis (u.x = 0 gives
            // cos theta = 1); the angular radius from the sample at the cone's rim (u.x = 1), 0 when
            // the light is hard and returns its axis for every u (DirectionalLight.h:50-76)
            const LightSample axis = dl->sample({0.f, 0.f, 0.f}, {0.f, 0.f, 1.f}, {0.f, 0.f});
            const LightSample rim = dl->sample({0.f, 0.f, 0.f}, {0.f, 0.f, 1.f}, {1.f, 0.f});
            al.type = AHIP_LIGHT_DIRECTIONAL;
            put3(al.direction, axis.wi);
            put3(al.Le, axis.Li);
            const float c = std::min(1.f, std::max(-1.f, dot(axis.wi, rim.wi)));
            al.half_angle = (rim.wi.x == axis.wi.x && rim.wi.y == axis.wi.y && rim.wi.z == axis.wi.z)
                                ? 0.f : std::acos(c) * 180.f / 3.14159265358979323846f;
            // the scene sphere of sampleLe (DirectionalLight.h:88-125): at uPos = (0, 0) the origin is
            // the disk's centre, centre + dirToLight * 100 radii, and at (1, 0) one radius from it
            const LightLeSample a = dl->sampleLe({0.f, 0.f}, {0.f, 0.f});
            const LightLeSample b = dl->sampleLe({1.f, 0.f}, {0.f, 0.f});
            const float radius = (b.pos - a.pos).length();
            put3(al.scene_center, a.pos - axis.wi * (radius * 100.f));
            al.scene_radius = radius;
        } else {
            return;   // other lights stay on the CPU path
        }
        lights.push_back(al);
    }
    ahip_scene_desc sd{};
    sd.num_meshes = (uint32_t)pool.numMeshes();
    sd.mesh_vertex_offset = vOff.data();
    sd.mesh_index_offset = iOff.data();
    sd.mesh_normal_count = nCnt.data();
    sd.mesh_uv_count = uCnt.data();
    sd.mesh_material = meshMat.data();
    sd.positions = pos.data();
    sd.normals = nrm.data();
    sd.uvs = uvs.data();
    sd.indices = idx.data();
    sd.num_vertices = (uint32_t)(pos.size() / 3);
    sd.num_indices = (uint32_t)idx.size();
    sd.num_materials = (uint32_t)mats.size();
    sd.materials = mats.data();
    sd.num_lights = (uint32_t)lights.size();
    sd.lights = lights.data();
    sd.env_light = envIndex;
    put3(sd.env_radiance, scene.envRadiance);
    if (dome) {
        sd.env_width = dome->envWidth();
        sd.env_height = dome->envHeight();
        sd.env_pixels = dome->pixels();
    }
    if (ahip_prepare(m_ctx, &sd) != AHIP_OK) {
        std::fprintf(stderr, "[warn] HipPathIntegrator::prepare: %s\n", ahip_last_error(m_ctx));
        return;
    }
    // the integrator is checked against the scene now (ahip_set_params refuses BDPT for scenes it
    // does not take), so that RenderSession can keep its CPU integrator
    const ahip_render_params p{m_maxDepth, 2, 0, 0, 0, m_integrator, m_fireflyClamp};
    if (ahip_set_params(m_ctx, &p) != AHIP_OK) {
        std::fprintf(stderr, "[warn] HipPathIntegrator::prepare: %s\n", ahip_last_error(m_ctx));
        return;
    }
    m_prepared = true;
}

bool HipPathIntegrator::ensureCamera(const SceneView& scene, uint32_t w, uint32_t h) {
    Camera cam = scene.camera.value_or(Camera::makePinhole({0.f, 0.f, -2.5f}, {0.f, 0.f, 1.f}, {0.f, 1.f, 0.f},
                                                           50.f, w, h));   // PathIntegrator.cpp:14-20
    if (cam.apertureRadius > 0.f) return false;   // thin lens stays on the CPU path
    ahip_camera c{};
    put3(c.origin, cam.origin);
    put3(c.lower_left, cam.lowerLeftCorner);
    put3(c.horizontal, cam.horizontal);
    put3(c.vertical, cam.vertical);
    c.width = w;
    c.height = h;
    ahip_render_params p{m_maxDepth, 2, 0, 0, 0, m_integrator, m_fireflyClamp};
    return ahip_set_camera(m_ctx, &c) == AHIP_OK && ahip_set_params(m_ctx, &p) == AHIP_OK;
}

bool HipPathIntegrator::renderFrame(const SceneView& scene, uint32_t W, uint32_t H, uint32_t sampleStart,
                                    uint32_t sampleCount, Film& film) {
    std::lock_guard<std::mutex> lk(m_mutex);
    if (!m_prepared || !ensureCamera(scene, W, H)) return false;
    std::vector<float> acc((size_t)W * H * 5);
    if (ahip_render_tile(m_ctx, 0, 0, W, H, sampleStart, sampleCount, acc.data()) != AHIP_OK) return false;
    mergeAccumIntoFilm(film, acc.data(), 0, 0, W, H);
    return true;
}

bool HipPathIntegrator::renderAdaptive(const SceneView& scene, uint32_t W, uint32_t H, uint32_t tileSize,
                                       uint32_t totalSpp, uint32_t baseSpp, Film& film) {
    std::lock_guard<std::mutex> lk(m_mutex);
    if (!m_prepared || !ensureCamera(scene, W, H)) return false;
    if (ahip_clear_accum(m_ctx) != AHIP_OK) return false;
    if (ahip_render_adaptive(m_ctx, W, H, tileSize, totalSpp, baseSpp, nullptr, nullptr) != AHIP_OK) return false;
    std::vector<float> acc((size_t)W * H * 5);
    if (ahip_download_accum(m_ctx, acc.data()) != AHIP_OK) return false;
    mergeAccumIntoFilm(film, acc.data(), 0, 0, W, H);
    return true;
}

void HipPathIntegrator::renderTile(const SceneView& scene, const TileRequest& tile, uint32_t W, uint32_t H,
                                   ISampler&, TileBuffer& out) {
    std::lock_guard<std::mutex> lk(m_mutex);
    if (!m_prepared || !ensureCamera(scene, W, H)) return;
    const size_t n = (size_t)tile.width * tile.height;
    std::vector<float> acc(n * 5), alb(n * 3), nrm(n * 3);
    if (ahip_render_tile_aovs(m_ctx, tile.x0, tile.y0, tile.width, tile.height, tile.sampleStart, tile.sampleCount,
                              acc.data(), alb.data(), nrm.data()) != AHIP_OK)
        return;
    for (uint32_t y = 0; y < tile.height; ++y)
        for (uint32_t x = 0; x < tile.width; ++x) {
            const size_t i = (size_t)y * tile.width + x;
            const float* p = &acc[5 * i];
            float w = p[3] > 0.f ? p[3] : 1.f;
            out.add(x, y, p[0] / w, p[1] / w, p[2] / w, w);
            out.addLumSq(x, y, p[4]);
            // the denoising AOVs PathIntegrator::renderTile records (:58-66); BDPTIntegrator::renderTile
            // records them only for pixels whose camera subpaths reached a surface (:567-580), which
            // the device reports as zero albedo and zero normal (a surface hit leaves one of them
            // nonzero unless every hit surface is black and their normals cancel exactly)
            const bool hitSurface = alb[3 * i] != 0.f || alb[3 * i + 1] != 0.f || alb[3 * i + 2] != 0.f ||
                                    nrm[3 * i] != 0.f || nrm[3 * i + 1] != 0.f || nrm[3 * i + 2] != 0.f;
            if (tile.sampleCount > 0 && (m_integrator == AHIP_INTEGRATOR_PATH || hitSurface)) {
                out.addAlbedo(x, y, makeSpectrum(alb[3 * i], alb[3 * i + 1], alb[3 * i + 2]));
                out.addNormal(x, y, Vec3f{nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]});
            }
        }
}

}  // namespace anacapa
