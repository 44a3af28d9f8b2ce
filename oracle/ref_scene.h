// ref_scene.h — ORACLE / TEST INFRASTRUCTURE ONLY: a reference SceneView built from an
// ahip_scene_desc with the reference's own classes (GeometryPool, BVHBackend, LambertianMaterial,
// EmissiveMaterial, AreaLight, DirectionalLight), shared by the reference drivers in oracle/.
// Include after the reference headers those classes come from (accel/BVHBackend.h,
// shading/Lambertian.h, shading/lights/AreaLight.h, shading/lights/DirectionalLight.h).
#pragma once
#include <memory>
#include <vector>

#include "../include/anacapa_hip.h"

using namespace anacapa;

namespace {

struct RefScene {
    GeometryPool pool;
    std::unique_ptr<BVHBackend> accel;
    std::vector<std::unique_ptr<IMaterial>> mats;
    std::vector<std::unique_ptr<ILight>> lights;
    SceneView view;
};

inline Vec3f v3(const float* p) { return Vec3f{p[0], p[1], p[2]}; }

// The pinhole Camera of an ahip_camera (IIntegrator.h:27-74).
inline Camera toCamera(const ahip_camera* c) {
    Camera cam;
    cam.origin = v3(c->origin);
    cam.lowerLeftCorner = v3(c->lower_left);
    cam.horizontal = v3(c->horizontal);
    cam.vertical = v3(c->vertical);
    cam.basisU = Vec3f{1, 0, 0};
    cam.basisV = Vec3f{0, 1, 0};
    cam.imageWidth = c->width;
    cam.imageHeight = c->height;
    return cam;
}

// Meshes into a GeometryPool committed to a BVHBackend, Lambert / Emissive materials by mesh, quad
// area lights and a constant background; nullptr for anything else (curves: CurveBrute.cpp needs
// spdlog; StandardSurface and DomeLight need OpenImageIO-only sources).  `standard`, when given,
// makes the IMaterial of a standard_surface material (nullptr: the scene is refused).
// `dome`, when given, makes the ILight of the environment dome from the descriptor's pixels.
using StandardFactory = std::unique_ptr<IMaterial> (*)(const ahip_material&);
using DomeFactory = std::unique_ptr<ILight> (*)(const ahip_scene_desc*, const ahip_light&);
inline std::unique_ptr<RefScene> buildRefScene(const ahip_scene_desc* d, StandardFactory standard = nullptr,
                                               DomeFactory dome = nullptr) {
    if (d->num_strands) return nullptr;
    auto s = std::make_unique<RefScene>();
    for (uint32_t m = 0; m < d->num_meshes; ++m) {
        MeshDesc mesh;
        uint32_t v0 = d->mesh_vertex_offset[m], v1 = d->mesh_vertex_offset[m + 1];
        uint32_t i0 = d->mesh_index_offset[m], i1 = d->mesh_index_offset[m + 1];
        for (uint32_t v = v0; v < v1; ++v) mesh.positions.push_back(v3(d->positions + 3 * v));
        uint32_t nn = d->mesh_normal_count ? d->mesh_normal_count[m] : 0;
        for (uint32_t v = 0; v < nn; ++v) mesh.normals.push_back(v3(d->normals + 3 * (v0 + v)));
        uint32_t nu = d->mesh_uv_count ? d->mesh_uv_count[m] : 0;
        for (uint32_t v = 0; v < nu; ++v)
            mesh.uvs.push_back(Vec2f{d->uvs[2 * (v0 + v)], d->uvs[2 * (v0 + v) + 1]});
        for (uint32_t i = i0; i < i1; ++i) mesh.indices.push_back(d->indices[i]);
        s->pool.addMesh(std::move(mesh));
    }
    for (uint32_t i = 0; i < d->num_materials; ++i) {
        const ahip_material& m = d->materials[i];
        if (m.type == AHIP_MAT_LAMBERT)
            s->mats.push_back(std::make_unique<LambertianMaterial>(v3(m.albedo)));
        else if (m.type == AHIP_MAT_EMISSIVE)
            s->mats.push_back(std::make_unique<EmissiveMaterial>(v3(m.albedo), v3(m.emission)));
        else if (m.type == AHIP_MAT_STANDARD && standard && !d->tri_albedo) {
            auto p = standard(m);
            if (!p) return nullptr;
            s->mats.push_back(std::move(p));
        } else
            return nullptr;
    }
    for (uint32_t m = 0; m < d->num_meshes; ++m) {
        int32_t mi = d->mesh_material[m];
        s->view.materials.push_back(mi >= 0 ? s->mats[mi].get() : nullptr);
    }
    for (uint32_t i = 0; i < d->num_lights; ++i) {
        const ahip_light& l = d->lights[i];
        if (l.type == AHIP_LIGHT_DOME && dome && d->env_light == (int32_t)i && d->env_pixels) {
            s->lights.push_back(dome(d, l));
            s->view.lights.push_back(s->lights.back().get());
            s->view.envLight = s->lights.back().get();   // the dome is also the environment
            continue;
        }
        if (l.type == AHIP_LIGHT_DIRECTIONAL) {   // the scene sphere only places BDPT's emission disk
            s->lights.push_back(std::make_unique<DirectionalLight>(v3(l.direction), v3(l.Le), l.scene_radius,
                                                                   v3(l.scene_center), l.half_angle));
            s->view.lights.push_back(s->lights.back().get());
            continue;
        }
        if (l.type != AHIP_LIGHT_AREA) return nullptr;
        s->lights.push_back(std::make_unique<AreaLight>(v3(l.center), v3(l.u_axis), v3(l.v_axis), v3(l.Le)));
        s->view.lights.push_back(s->lights.back().get());
    }
    if (d->env_light >= 0 && !s->view.envLight) return nullptr;
    s->view.envRadiance = v3(d->env_radiance);
    s->accel = std::make_unique<BVHBackend>(s->pool);
    s->accel->commit();
    s->view.accel = s->accel.get();
    return s;
}

}  // namespace
