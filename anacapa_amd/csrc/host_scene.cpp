// host_scene.cpp — materials, lights, dome distributions, Halton tables and the pinhole camera,
// computed on the host.  Compiled by g++ with the reference's Release flags (like bvh_build.cpp)
// so the float arithmetic of the dome's distribution build and the camera set-up contracts as
// the reference's does.
#include <algorithm>
#include <cmath>
#include <cstring>

#define __host__
#define __device__
#include "device_types.h"
#include "host_scene.h"

namespace ahip {

namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr float k2Pi = 6.28318530717958647692f;

// First 128 primes (the Halton bases of src/sampling/HaltonSampler.h:68-79).
void firstPrimes(uint32_t* out, int n) {
    int k = 0;
    for (uint32_t c = 2; k < n; ++c) {
        bool prime = true;
        for (uint32_t d = 2; d * d <= c; ++d)
            if (c % d == 0) { prime = false; break; }
        if (prime) out[k++] = c;
    }
}

struct V {
    float x, y, z;
};
V sub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V mul(V a, float s) { return {a.x * s, a.y * s, a.z * s}; }
V crs(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float lenv(V a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
V nrm(V a) { return mul(a, 1.f / lenv(a)); }
V nrmSafe(V a) {
    float l = lenv(a);
    return l > 1e-8f ? mul(a, 1.f / l) : V{0.f, 0.f, 1.f};
}
V ld(const float* p) { return {p[0], p[1], p[2]}; }
void st(float* p, V v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

// Distribution1D::build (src/shading/lights/DomeLight.h:25-37)
// First index i of cdf[0..n) with cdf[i] >= u (n if none), the device search's predicate.
uint32_t lowerBoundHost(const float* cdf, uint32_t n, float u) {
    uint32_t lo = 0, count = n;
    while (count > 0u) {
        uint32_t step = count >> 1;
        if (cdf[lo + step] < u) { lo += step + 1u; count -= step + 1u; }
        else count = step;
    }
    return lo;
}

// Guide table size: a power of two of at least twice the bins, so k / M and u * M are exact.
uint32_t guideSize(uint32_t bins) {
    uint32_t m = 1;
    while (m < 2u * bins && m < (1u << 20)) m <<= 1;
    return m;
}

// guide[k] = lowerBound(cdf, k / M) for k = 0..M.  The search is monotone in u, so for u in
// [k/M, (k+1)/M) its result lies in [guide[k], guide[k+1]] and a search restricted to that range
// returns the same index.
void buildGuide(const float* cdf, uint32_t n, uint32_t M, uint32_t* guide) {
    const float inv = 1.f / (float)M;
    for (uint32_t k = 0; k <= M; ++k) guide[k] = lowerBoundHost(cdf, n, (float)k * inv);
}

float buildDist(const float* f, uint32_t n, float* func, float* cdf) {
    std::memcpy(func, f, n * sizeof(float));
    cdf[0] = 0.f;
    for (uint32_t i = 0; i < n; ++i) cdf[i + 1] = cdf[i] + func[i];
    float integral = cdf[n];
    if (integral > 0.f) {
        float inv = 1.f / integral;
        for (uint32_t i = 0; i <= n; ++i) cdf[i] *= inv;
    }
    return integral;
}

}  // namespace

void fillHaltonTables(HaltonTables& t) {
    firstPrimes(t.prime, 128);
    for (int i = 0; i < 128; ++i) {
        const uint32_t b = t.prime[i];
        t.invPrime[i] = 1.f / static_cast<float>(b);
        // A 32-bit multiplier when one exists (Granlund & Montgomery, Theorem 4.2 with N = 32):
        // m = ceil(2^(32+s) / b) and m * b - 2^(32+s) <= 2^s give n / b == mulhi(m, n) >> s for
        // every 32-bit n (flag bit 16).  Otherwise the 33-bit multiplier of their figure 4.1.
        bool simple = false;
        for (uint32_t sh = 0; sh < 32u && !simple; ++sh) {
            const uint64_t p2 = 1ull << (32u + sh);
            const uint64_t m = (p2 + b - 1u) / b;
            if (m < (1ull << 32) && m * b - p2 <= (1ull << sh)) {
                t.magic[i] = static_cast<uint32_t>(m);
                t.shifts[i] = sh | 0x10000u;
                simple = true;
            }
        }
        if (simple) continue;
        uint32_t l = 0;
        while ((1ull << l) < b) ++l;                       // l = ceil(log2 b)
        t.magic[i] = static_cast<uint32_t>((((1ull << l) - b) << 32) / b + 1u);
        const uint32_t sh1 = l < 1u ? l : 1u, sh2 = l > 0u ? l - 1u : 0u;
        t.shifts[i] = sh1 | (sh2 << 8);
    }
}

// Camera::makePinhole (include/anacapa/integrator/IIntegrator.h:49-74)
int makePinholeCamera(const float from[3], const float at[3], const float up[3], float vfov, uint32_t width,
                      uint32_t height, ahip_camera* out) {
    float theta = vfov * 3.14159265f / 180.f;
    float h = std::tan(theta * 0.5f);
    float aspect = static_cast<float>(width) / static_cast<float>(height);
    V f = ld(from), a = ld(at), u0 = ld(up);
    V w = nrm(sub(f, a));
    V u = nrm(crs(u0, w));
    V v = crs(w, u);
    float vpH = 2.f * h, vpW = aspect * vpH;
    st(out->origin, f);
    st(out->horizontal, mul(u, vpW));
    st(out->vertical, mul(v, vpH));
    V llc = sub(sub(sub(f, mul(u, vpW * 0.5f)), mul(v, vpH * 0.5f)), w);
    st(out->lower_left, llc);
    out->width = width;
    out->height = height;
    return AHIP_OK;
}

// fresnelDielectric (StandardSurface.h:84-92): exact Fresnel reflectance of a dielectric interface.
static float fresnelDielectricHost(float cosI, float eta) {
    float sin2I = std::max(0.f, 1.f - cosI * cosI);
    float sin2T = sin2I / (eta * eta);
    if (sin2T >= 1.f) return 1.f;
    float cosT = std::sqrt(1.f - sin2T);
    float rs = (cosI - eta * cosT) / (cosI + eta * cosT);
    float rp = (eta * cosI - cosT) / (eta * cosI + cosT);
    return 0.5f * (rs * rs + rp * rp);
}

int convertMaterial(const ahip_material& m, uint32_t numTextures, DMat& D, std::string& err) {
    D = DMat{};
    D.type = m.type;
    D.roughTex = -1;
    D.opacityTex = -1;
    std::memcpy(D.albedo, m.albedo, 12);
    if (m.type == AHIP_MAT_LAMBERT || m.type == AHIP_MAT_EMISSIVE) {
        if (m.type == AHIP_MAT_EMISSIVE) std::memcpy(D.Le, m.emission, 12);
    } else if (m.type == AHIP_MAT_HAIR) {
        std::memcpy(D.hairSigma, m.hair_sigma_a, 12);   // MarschnerHair.h:193-199
        // the constructor's Marschner constants (MarschnerHair.h:200-216): longitudinal variances
        // of R, TT, TRT, the azimuthal logistic scale and the cuticle tilts
        D.hairEta = m.hair_eta;
        float bm = std::clamp(m.hair_beta_m, 1e-3f, 1.f);
        float v0 = 0.726f * bm + 0.812f * bm * bm + 3.7f * std::pow(bm, 20.f);
        D.hairV[0] = v0 * v0;
        D.hairV[1] = D.hairV[0] * 0.25f;
        D.hairV[2] = D.hairV[0] * 4.0f;
        float bn = std::clamp(m.hair_beta_n, 1e-3f, 1.f);
        D.hairS = 0.626657069f * (0.265f * bn + 1.194f * bn * bn + 5.372f * std::pow(bn, 22.f));
        const float kPiF = 3.14159265358979323846f;
        D.hairAlpha[0] = -m.hair_alpha * kPiF / 180.f;
        D.hairAlpha[1] = m.hair_alpha * kPiF / 180.f * 0.5f;
        D.hairAlpha[2] = -m.hair_alpha * kPiF / 180.f * 1.5f;
    } else if (m.type == AHIP_MAT_STANDARD) {
        D.base = m.base;
        D.metal = m.metalness;
        D.spec = m.specular;
        D.rough = m.roughness;
        std::memcpy(D.specColor, m.specular_color, 12);
        D.coat = m.coat;
        float ca = std::max(1e-4f, m.coat_roughness * m.coat_roughness);
        D.coatA2 = ca * ca;
        float r = (1.f - m.specular_ior) / (1.f + m.specular_ior);
        D.f0d = r * r;
        float rc = (1.f - 1.5f) / (1.f + 1.5f);
        D.coatF0 = rc * rc;
        float rr = m.roughness, mm = m.metalness, ss = m.specular, tr = m.transmission;
        // isDelta: smooth dielectric glass, else a smooth metal or specular surface (:162-169)
        D.isDelta = ((rr < 0.001f && tr > 0.001f && mm < 0.001f) || (rr < 0.001f && (mm > 0.999f || ss > 0.f)))
                        ? 1u : 0u;
        // the glass branches (:248-347, 410-474) and transmittanceColor (:219-224) test the
        // transmission and metalness thresholds in these two forms
        D.glass = (tr > 0.001f && mm < 0.001f) ? 1u : 0u;
        D.trans = (!(tr < 0.001f) && !(mm > 0.001f)) ? 1u : 0u;
        D.eta = m.specular_ior;
        D.trScale = tr * (1.f - fresnelDielectricHost(1.f, m.specular_ior));
        if (m.alpha_mask) {
            D.alphaMask = 1u;
            D.opacity = m.opacity;
            D.opacityTex = -1;
            if (m.opacity_texture >= 0) {
                if ((uint32_t)m.opacity_texture >= numTextures) {
                    err = "opacity_texture out of range";
                    return AHIP_E_INVALID_ARG;
                }
                D.opacityTex = m.opacity_texture;
            }
        }
        D.specularFlag = rr < 0.001f ? 1u : 0u;
        if (m.emission_weight > 0.f)
            for (int k = 0; k < 3; ++k) D.Le[k] = m.emission[k] * m.emission_weight;
        if (m.roughness_texture >= 0) {
            if ((uint32_t)m.roughness_texture >= numTextures) {
                err = "roughness_texture out of range";
                return AHIP_E_INVALID_ARG;
            }
            D.roughTex = m.roughness_texture;
        }
    } else {
        err = "unknown material type";
        return AHIP_E_INVALID_ARG;
    }
    D.hasLe = (D.Le[0] != 0.f || D.Le[1] != 0.f || D.Le[2] != 0.f) ? 1u : 0u;
    return AHIP_OK;
}

int buildHostScene(const ahip_scene_desc* d, HostScene& out, std::string& err) {
    out = HostScene{};
    if (d->num_meshes && (!d->mesh_vertex_offset || !d->mesh_index_offset || !d->mesh_material || !d->positions ||
                          !d->indices)) {
        err = "scene arrays missing";
        return AHIP_E_INVALID_ARG;
    }
    if (d->num_meshes && d->mesh_vertex_offset[d->num_meshes] > d->num_vertices) {
        err = "mesh_vertex_offset beyond num_vertices";
        return AHIP_E_INVALID_ARG;
    }
    // textures
    for (uint32_t i = 0; i < d->num_textures; ++i) {
        uint32_t w = d->texture_width[i], h = d->texture_height[i];
        if (!w || !h || !d->texture_pixels[i]) {
            err = "empty texture";
            return AHIP_E_INVALID_ARG;
        }
        out.texPixels.emplace_back(d->texture_pixels[i], d->texture_pixels[i] + (size_t)w * h);
        out.texW.push_back(w);
        out.texH.push_back(h);
    }
    // materials (Lambertian.h:29-99, StandardSurface.h:152-187, MarschnerHair.h:190-216)
    std::vector<bool> emits(d->num_materials, false);
    for (uint32_t i = 0; i < d->num_materials; ++i) {
        DMat D{};
        int st = convertMaterial(d->materials[i], d->num_textures, D, err);
        if (st) return st;
        emits[i] = D.hasLe != 0u;
        if (D.type == AHIP_MAT_HAIR) out.hasHair = true;
        if (D.trans || D.alphaMask) out.hasTransparent = true;
        if (D.alphaMask) out.hasAlpha = true;
        out.mats.push_back(D);
    }
    out.meshMat.assign(d->mesh_material, d->mesh_material + d->num_meshes);
    for (int32_t mi : out.meshMat) {
        if (mi >= (int32_t)d->num_materials) {
            err = "mesh_material out of range";
            return AHIP_E_INVALID_ARG;
        }
        if (mi >= 0 && emits[mi]) out.hasEmitters = true;
    }
    // lights (AreaLight.h:123-131)
    for (uint32_t i = 0; i < d->num_lights; ++i) {
        const ahip_light& l = d->lights[i];
        DLight L{};
        L.type = l.type;
        if (l.type == AHIP_LIGHT_AREA) {
            std::memcpy(L.center, l.center, 12);
            std::memcpy(L.uA, l.u_axis, 12);
            std::memcpy(L.vA, l.v_axis, 12);
            std::memcpy(L.Le, l.Le, 12);
            V c = crs(ld(l.u_axis), ld(l.v_axis));
            st(L.normal, nrmSafe(c));
            L.area = lenv(c) * 4.f;
            L.uLen = lenv(ld(l.u_axis));
            L.vLen = lenv(ld(l.v_axis));
        } else if (l.type == AHIP_LIGHT_DIRECTIONAL) {
            // DirectionalLight(dirToLight, Le, ..., halfAngleDeg) (DirectionalLight.h:33-45)
            const V dl = ld(l.direction);
            const float len = lenv(dl);
            if (!(len > 1e-8f)) {
                L.dir[0] = 0.f; L.dir[1] = 0.f; L.dir[2] = 1.f;   // safeNormalize's fallback
            } else {
                st(L.dir, V{dl.x * (1.f / len), dl.y * (1.f / len), dl.z * (1.f / len)});
            }
            std::memcpy(L.Le, l.Le, 12);
            if (!(l.half_angle >= 0.f) || !(l.half_angle <= 90.f)) {
                err = "directional light half_angle must be in [0, 90] degrees";
                return AHIP_E_INVALID_ARG;
            }
            L.cosCone = std::cos(l.half_angle * 3.14159265358979323846f / 180.f);
        } else if (l.type == AHIP_LIGHT_DOME) {
            if ((int32_t)i != d->env_light) {
                err = "a dome light must be the scene's env_light";
                return AHIP_E_INVALID_ARG;
            }
        } else {
            err = "unknown light type";
            return AHIP_E_INVALID_ARG;
        }
        if (l.type != AHIP_LIGHT_AREA) {
            if (!(l.scene_radius >= 0.f) || !std::isfinite(l.scene_radius) || !std::isfinite(l.scene_center[0]) ||
                !std::isfinite(l.scene_center[1]) || !std::isfinite(l.scene_center[2])) {
                err = "a light's scene_radius must be finite and >= 0 and its scene_center finite";
                return AHIP_E_INVALID_ARG;
            }
            std::memcpy(L.sceneC, l.scene_center, 12);
            L.sceneR = l.scene_radius;
        }
        out.lights.push_back(L);
    }
    if (d->env_light >= 0) {
        if ((uint32_t)d->env_light >= d->num_lights || d->lights[d->env_light].type != AHIP_LIGHT_DOME) {
            err = "env_light must index a dome light";
            return AHIP_E_INVALID_ARG;
        }
        // DomeLight::buildDistribution (DomeLight.h:228-259); an absent image is the 1x1 white fallback
        HostDome& D = out.dome;
        const ahip_light& l = d->lights[d->env_light];
        D.intensity = l.intensity;
        std::memcpy(D.rot, l.rot, sizeof(D.rot));
        std::vector<float> rgb;
        if (d->env_pixels && d->env_width && d->env_height) {
            D.W = d->env_width;
            D.H = d->env_height;
            rgb.assign(d->env_pixels, d->env_pixels + (size_t)D.W * D.H * 3);
        } else {
            D.W = D.H = 1;
            rgb = {1.f, 1.f, 1.f};
        }
        D.px4.resize((size_t)D.W * D.H * 4);
        for (size_t i = 0; i < (size_t)D.W * D.H; ++i) {
            D.px4[4 * i] = rgb[3 * i];
            D.px4[4 * i + 1] = rgb[3 * i + 1];
            D.px4[4 * i + 2] = rgb[3 * i + 2];
            D.px4[4 * i + 3] = 0.f;
        }
        D.condFunc.resize((size_t)D.W * D.H);
        D.condCdf.resize((size_t)(D.W + 1) * D.H);
        D.condInt.resize(D.H);
        std::vector<float> row(D.W), rowSums(D.H);
        for (uint32_t r = 0; r < D.H; ++r) {
            float theta = kPi * (r + 0.5f) / static_cast<float>(D.H);
            float s = std::sin(theta);
            for (uint32_t c = 0; c < D.W; ++c) {
                const float* p = &rgb[((size_t)r * D.W + c) * 3];
                float lum = 0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2];
                row[c] = lum * s;
            }
            D.condInt[r] = buildDist(row.data(), D.W, &D.condFunc[(size_t)r * D.W], &D.condCdf[(size_t)r * (D.W + 1)]);
            rowSums[r] = D.condInt[r];
        }
        D.margFunc.resize(D.H);
        D.margCdf.resize(D.H + 1);
        D.margInt = buildDist(rowSums.data(), D.H, D.margFunc.data(), D.margCdf.data());
        D.margM = guideSize(D.H);
        D.condM = guideSize(D.W);
        D.margGuide.resize(D.margM + 1);
        buildGuide(D.margCdf.data(), D.H + 1, D.margM, D.margGuide.data());
        D.condGuide.resize((size_t)(D.condM + 1) * D.H);
        for (uint32_t r = 0; r < D.H; ++r)
            buildGuide(&D.condCdf[(size_t)r * (D.W + 1)], D.W + 1, D.condM, &D.condGuide[(size_t)r * (D.condM + 1)]);
        out.hasDome = true;
        (void)k2Pi;
    }
    return AHIP_OK;
}

}  // namespace ahip
