// host_scene.h — host-side conversion of ahip_scene_desc into the device records.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/anacapa_hip.h"
#include "device_types.h"

namespace ahip {

struct HaltonTables;

struct HostDome {
    uint32_t W = 0, H = 0;
    std::vector<float> px4;          // W*H*4 (rgb + pad)
    std::vector<float> margFunc, margCdf, condFunc, condCdf, condInt;
    // guide tables of the CDF searches: guide[k] = first index i with cdf[i] >= k / M (M a power
    // of two), so a sample u in [k/M, (k+1)/M) is found between guide[k] and guide[k+1]
    std::vector<uint32_t> margGuide, condGuide;
    uint32_t margM = 0, condM = 0;
    float margInt = 0.f;
    float intensity = 1.f;
    float rot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
};

struct HostScene {
    std::vector<DMat> mats;
    std::vector<int32_t> meshMat;
    std::vector<DLight> lights;
    std::vector<std::vector<float>> texPixels;
    std::vector<uint32_t> texW, texH;
    HostDome dome;
    bool hasDome = false;
    bool hasEmitters = false;
    bool hasHair = false;
    bool hasTransparent = false;   // a glass or alpha-masked material
    bool hasAlpha = false;         // an alpha-masked material
};

// One material's device record (Lambert / Emissive / StandardSurface, glass and alpha masks
// included / hair).  Returns an AHIP_E_* status.
int convertMaterial(const ahip_material& m, uint32_t numTextures, DMat& out, std::string& err);

// Validates the description and builds materials, lights, textures and the dome's sampling
// distributions.  Returns an AHIP_E_* status.
int buildHostScene(const ahip_scene_desc* d, HostScene& out, std::string& err);
void fillHaltonTables(HaltonTables& t);
int makePinholeCamera(const float from[3], const float at[3], const float up[3], float vfov, uint32_t w, uint32_t h,
                      ahip_camera* out);

}  // namespace ahip
