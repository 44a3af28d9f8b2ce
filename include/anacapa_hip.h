/*
 * anacapa_hip.h — C ABI of the MI355X (gfx950) wavefront Path+NEE backend.
 *
 * This library is a drop-in for the renderer's GPU backend slot, the role
 * CudaPathIntegrator plays in the reference (src/gpu/cuda/CudaPathIntegrator.h:21-52,
 * behind IIntegrator, include/anacapa/integrator/IIntegrator.h:172-198).
 * Every entry point takes plain pointers and sizes; all device memory is owned
 * by the context.  Every call returns an int status (0 = OK, see AHIP_E_*);
 * the message of the last failure is available from ahip_last_error().
 * Calls on one context are serialised internally; they are synchronous at the
 * ABI boundary (results are complete when the call returns).
 */
#ifndef ANACAPA_HIP_H
#define ANACAPA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AHIP_ABI_VERSION 4

/* status codes */
#define AHIP_OK              0
#define AHIP_E_INVALID_ARG   1
#define AHIP_E_NO_DEVICE     2
#define AHIP_E_HIP           3
#define AHIP_E_NOT_PREPARED  4
#define AHIP_E_UNSUPPORTED   5
#define AHIP_E_INJECTED      6
#define AHIP_E_OUT_OF_MEMORY 7

/* material types (reference: src/shading/Lambertian.h:29-99,
 * src/shading/StandardSurface.h:115-691, src/shading/MarschnerHair.h) */
#define AHIP_MAT_LAMBERT   0
#define AHIP_MAT_EMISSIVE  1
#define AHIP_MAT_STANDARD  2
#define AHIP_MAT_HAIR      3   /* MarschnerHairMaterial; the reference runtime evaluates Kajiya-Kay */

/* light types (reference: src/shading/lights/AreaLight.h:118-226,
 * src/shading/lights/DomeLight.h:78-388) */
#define AHIP_LIGHT_AREA 0
#define AHIP_LIGHT_DOME 1
#define AHIP_LIGHT_DIRECTIONAL 2   /* DirectionalLight (src/shading/lights/DirectionalLight.h:24-160): a
                                      delta light at infinity; the path tracer only (BDPT refuses it) */

/* ahip_render_params.flags */
#define AHIP_FLAG_FIX_DOUBLE_COSINE 1u  /* opt out of the reference's extra cos in NEE's BSDF branch
                                          (src/integrator/PathIntegrator.cpp:257-258) */
#define AHIP_FLAG_HAIR_MARSCHNER    2u  /* hair materials evaluate and sample the Marschner R/TT/TRT
                                          model MarschnerHairMaterial carries below its Kajiya-Kay
                                          returns (src/shading/MarschnerHair.h:243-257, 284-375,
                                          490-556) instead of the Kajiya-Kay BSDF it runs (:240,
                                          266-282); off = the reference's runtime behaviour */

typedef struct ahip_material {
    uint32_t type;              /* AHIP_MAT_* */
    float    albedo[3];         /* Lambert/Emissive albedo; StandardSurface base_color */
    float    emission[3];       /* Emissive Le; StandardSurface emission_color */
    float    emission_weight;   /* StandardSurface `emission` (0 = none) */
    float    base;              /* StandardSurface Params (StandardSurface.h:117-148) */
    float    metalness;
    float    roughness;
    float    specular;
    float    specular_color[3];
    float    specular_ior;
    float    coat;
    float    coat_roughness;
    float    transmission;      /* > 0.001 with metalness < 0.001: dielectric glass (smooth below
                                   roughness 0.001, else GGX reflection + microfacet transmission,
                                   StandardSurface.h:248-347, 410-474); shadow rays step through
                                   it with the tint transmittanceColor gives (:207-225) */
    int32_t  roughness_texture; /* -1 = constant roughness; otherwise index of a texture */
    float    opacity;           /* Params::opacity (StandardSurface.h:141), read when alpha_mask */
    int32_t  opacity_texture;   /* -1 = constant opacity; otherwise index of a texture (evalTOV) */
    float    hair_sigma_a[3];   /* MarschnerHairMaterial::Params (MarschnerHair.h:193-199) */
    float    hair_eta;
    float    hair_beta_m;
    float    hair_beta_n;
    float    hair_alpha;        /* degrees */
    uint32_t alpha_mask;        /* StandardSurface Params::alphaMask (StandardSurface.h:143): evalOpacity
                                   = opacity (:197-201), a stochastic pass-through in the bounce loop
                                   (PathIntegrator.cpp:104-118) and a tint 1 - opacity for shadow
                                   rays (:213-217); 0 = opaque (evalOpacity = 1) */
} ahip_material;

typedef struct ahip_light {
    uint32_t type;              /* AHIP_LIGHT_* */
    float    center[3];         /* area: quad centre */
    float    u_axis[3];         /* area: half-extent vectors */
    float    v_axis[3];
    float    Le[3];             /* area: emitted radiance */
    float    intensity;         /* dome: radiance scale */
    float    rot[9];            /* dome: world->envmap rows (DomeLight.h:105-112) */
    float    direction[3];      /* directional: dirToLight, from a surface toward the light (normalised
                                   on prepare as the constructor does); Le: its radiance */
    float    half_angle;        /* directional: angular radius in degrees (0 = a hard, exact direction;
                                   DirectionalLight.h:33-45, 55-74) */
    float    scene_center[3];   /* dome, directional: the scene's bounding sphere, where BDPT's light
                                   subpaths start (sampleLe: a disk of radius scene_radius facing the
                                   scene; DomeLight.h:84-92, 142-168, DirectionalLight.h:28-41, 88-125);
                                   centre = the geometry bounds' centroid, radius = the half
                                   diagonal x 2 for RenderSession's dome (RenderSession.cpp:407-416),
                                   x 1.5 for the USD loader's lights (USDLoader.cpp:1730-1744).
                                   The path tracer does not read them. */
    float    scene_radius;      /* > 0 for BDPT with a dome or directional light */
} ahip_light;

/* Flat scene.  Mirrors GeometryPool / SceneView (include/anacapa/accel/GeometryPool.h:118-197,
 * IIntegrator.h:150-157): meshes are ranges of one vertex array and one index array;
 * indices are local to their mesh. Materials are indexed by mesh (SceneView::materials). */
typedef struct ahip_scene_desc {
    uint32_t        num_meshes;
    const uint32_t* mesh_vertex_offset;   /* [num_meshes+1] first vertex of each mesh */
    const uint32_t* mesh_index_offset;    /* [num_meshes+1] first index (multiple of 3) */
    const uint32_t* mesh_normal_count;    /* [num_meshes] normals present per mesh (<= vertices), may be NULL */
    const uint32_t* mesh_uv_count;        /* [num_meshes] uvs present per mesh, may be NULL */
    const int32_t*  mesh_material;        /* [num_meshes] material index, -1 = none */
    const float*    positions;            /* [num_vertices*3] world space */
    const float*    normals;              /* [num_vertices*3] or NULL */
    const float*    uvs;                  /* [num_vertices*2] or NULL */
    const uint32_t* indices;              /* [num_indices] */
    uint32_t        num_vertices;
    uint32_t        num_indices;

    uint32_t             num_materials;
    const ahip_material* materials;
    uint32_t             num_lights;
    const ahip_light*    lights;          /* the dome (if any) is also the environment light */
    int32_t              env_light;       /* index of the dome in lights[], -1 = constant background */
    float                env_radiance[3]; /* SceneView::envRadiance when env_light < 0 */
    uint32_t             env_width;       /* dome equirect RGB f32 image, row 0 = zenith */
    uint32_t             env_height;
    const float*         env_pixels;

    uint32_t             num_textures;    /* single-channel f32 textures (roughness) */
    const uint32_t*      texture_width;
    const uint32_t*      texture_height;
    const float* const*  texture_pixels;

    /* Hair strands (CurvePool / StrandDesc, include/anacapa/accel/CurvePool.h:27-77): cubic
     * Bezier segments with shared end points, strand s owns control points
     * [strand_cv_offset[s], strand_cv_offset[s+1]) (3*segments+1 of them). */
    uint32_t        num_strands;
    const uint32_t* strand_cv_offset;     /* [num_strands+1] */
    const float*    curve_cvs;            /* [total_cvs*3] world space */
    const float*    strand_width;         /* [num_strands*2] root and tip diameter (StrandDesc::widthAt) */
    const uint32_t* strand_mesh_index;    /* [num_strands] StrandDesc::materialIndex, reported as meshID */
    const float*    strand_color;         /* [num_strands*3] StrandDesc::color, may be NULL (white) */
    /* Per-triangle base colour, [num_indices/3 * 3] in commit order (mesh by mesh, triangle by
     * triangle), replacing the albedo / base_color of the triangle's Lambert or standard_surface
     * material; NULL = the materials' own.  Equivalent to one material per triangle in the
     * reference; it is the albedo parameter vector of ahip_adjoint. */
    const float*    tri_albedo;
    /* Per-CV strand diameters (StrandDesc::widths, CurvePool.h:30), used instead of strand_width
     * when non-NULL: strand s owns strand_widths[strand_width_offset[s], strand_width_offset[s+1]).
     * StrandDesc::widthAt (CurvePool.h:45-54): no value gives 0.001, one a constant, two a
     * root-to-tip ramp, more a piecewise-linear ramp over their index. */
    const uint32_t* strand_width_offset;  /* [num_strands+1] */
    const float*    strand_widths;
} ahip_scene_desc;

/* Camera fields of the reference pinhole Camera (IIntegrator.h:27-74). */
typedef struct ahip_camera {
    float    origin[3];
    float    lower_left[3];
    float    horizontal[3];
    float    vertical[3];
    uint32_t width;
    uint32_t height;
} ahip_camera;

/* ahip_render_params.integrator (RenderSettings::integrator, src/render/RenderSession.h:45-47) */
#define AHIP_INTEGRATOR_PATH 0u   /* PathIntegrator (src/integrator/PathIntegrator.cpp) */
#define AHIP_INTEGRATOR_BDPT 1u   /* BDPTIntegrator (src/integrator/BDPTIntegrator.cpp): scenes of
                                     triangles with Lambert / Emissive / StandardSurface materials
                                     without a roughness texture (glass and alpha masks included) lit by area
                                     lights, the dome and directional lights (each with its
                                     scene_radius > 0); no constant env_radiance, no per-triangle
                                     albedo; ahip_render_tile,
                                     ahip_render_tile_aovs and ahip_render_frame (without AOVs) */

typedef struct ahip_render_params {
    uint32_t max_depth;   /* PathIntegrator maxDepth, inclusive (PathIntegrator.cpp:80); BDPT: bounces
                             per subpath (BDPTIntegrator.cpp:97, 240) */
    uint32_t min_depth;   /* Russian roulette from bounce >= min_depth (PathIntegrator.cpp:156) */
    uint64_t seed;        /* HaltonSampler seed (HaltonSampler.h:24) */
    uint32_t flags;       /* AHIP_FLAG_* */
    uint32_t curve_depth; /* de Casteljau depth of the ribbon test, 1..6; 0 = the reference's 6
                             (MAX_DEPTH, src/accel/CurveBrute.cpp:194); other depths are not pinned by
                             the reference, which fixes 6 */
    uint32_t integrator;  /* AHIP_INTEGRATOR_*; 0 = the path tracer */
    float firefly_clamp;  /* BDPT: largest luminance of one s, t >= 2 strategy's contribution, 0 = off
                             (RenderSettings::fireflyClamp, RenderSession.h:47: the reference's default 10) */
} ahip_render_params;

/* One camera view of the batched persistent form. */
typedef struct ahip_view {
    ahip_camera camera;
    uint32_t    sample_start;
    uint32_t    sample_count;
    uint32_t    max_depth;
    uint32_t    pad;
    float*      accum;        /* host output [width*height*5] r,g,b,w,lumSq sums, may be NULL */
} ahip_view;

typedef struct ahip_stats {
    uint64_t rays_closest;       /* continuation + MIS rays traced (closest hit) */
    uint64_t rays_shadow;        /* NEE shadow rays traced (any hit) */
    uint64_t samples;            /* camera samples rendered */
    uint64_t nodes_visited;      /* filled by ahip_count_traffic only */
    uint64_t prims_tested;
    double   trace_ms;           /* sum of HIP-event durations of the traversal kernels */
    double   shade_ms;           /* sum of HIP-event durations of generate/shade/accumulate */
    uint64_t trace_launches;
    uint64_t shade_launches;
    double   render_ms;          /* wall time of the last render call on the stream (HIP events) */
    uint64_t passes;             /* wavefront passes run */
    uint64_t graph_passes;       /* of which replayed from a captured HIP graph */
    uint64_t graph_captures;     /* pass graphs captured (a graph is reused while its launch arguments hold) */
} ahip_stats;

typedef struct ahip_context ahip_context;

/* --- lifetime (CudaPathIntegrator ctor/isValid, CudaPathIntegrator.cu:211-227; IBackend.h:13-20) --- */
int         ahip_create(int device_ordinal, ahip_context** out);
void        ahip_destroy(ahip_context* ctx);
int         ahip_is_valid(const ahip_context* ctx);
const char* ahip_device_name(const ahip_context* ctx);
const char* ahip_backend_type(void);
const char* ahip_last_error(const ahip_context* ctx);
int         ahip_abi_version(void);

/* --- scene upload (CudaPathIntegrator::prepare, CudaPathIntegrator.cu:240-317):
 *     builds the SAH-12 BVH of BVHBackend::commit (src/accel/BVHBackend.cpp:17-228)
 *     and copies geometry, materials, lights and the HDRI into HBM. --- */
int ahip_prepare(ahip_context* ctx, const ahip_scene_desc* scene);
/* BVH sizes after prepare: nodes (64-B records) and leaf-ordered triangles. */
int ahip_bvh_info(ahip_context* ctx, uint32_t* num_nodes, uint32_t* num_tris, uint32_t* max_depth);
/* Dump the binary (reference-layout) BVH built on the host: per node
 * {bmin[3], dataA, bmax[3], dataB} exactly as BVHNode (BVHBackend.h:25-39), and the
 * leaf primitive order (BVHBackend m_primIndices). Either pointer may be NULL. */
int ahip_bvh_dump(ahip_context* ctx, float* nodes32, uint32_t* prim_indices);

/* --- rendering (IIntegrator::renderFrame / renderTile, IIntegrator.h:181-197) --- */
int ahip_set_camera(ahip_context* ctx, const ahip_camera* cam);
int ahip_set_params(ahip_context* ctx, const ahip_render_params* params);
int ahip_clear_accum(ahip_context* ctx);    /* CudaPathIntegrator::clearAccum (.h:47) */
/* Accumulate samples [sample_start, sample_start+sample_count) of every pixel into the
 * persistent device buffer (CudaPathIntegrator::renderFrame, .cu:379-444). */
int ahip_render_frame(ahip_context* ctx, uint32_t width, uint32_t height,
                      uint32_t sample_start, uint32_t sample_count);
/* Copy the accumulation buffer: [width*height*5] = r,g,b sums, weight, sum of luminance^2
 * (GpuAccumPixel, src/gpu/cuda/shaders/SharedTypes.h:127-130). */
int ahip_download_accum(ahip_context* ctx, float* out);
/* Render one tile into out[w*h*5] with the same layout (CudaPathIntegrator::renderTile, .cu:449-489). */
int ahip_render_tile(ahip_context* ctx, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                     uint32_t sample_start, uint32_t sample_count, float* out);
/* Resolve the accumulation buffer on the device, each output optional ([w*h*3] or [w*h]):
 * srgb8 — Film::writePNG's pixels (exposure 2^exposure, clamp, sRGB curve, 8 bits; Film.cpp:233-256);
 * rgb — PixelAccumulator::resolve (sums / weight, Film.h:46-55); variance — Film::varianceAt
 * (Film.cpp:80-90); albedo, normal — the first-hit AOV means (normal scaled to unit length), which
 * need ahip_set_aovs before rendering.  Only the requested outputs cross the bus. */
int ahip_resolve(ahip_context* ctx, float exposure, uint8_t* srgb8, float* rgb, float* variance, float* albedo,
                 float* normal);
/* Image writers for resolved buffers (host only), in place of the reference's OpenImageIO output:
 * ahip_write_png — 8-bit RGB, written beside the target and renamed into place (Film.cpp:257-275);
 * ahip_write_exr — float RGB layers, channels "<prefix>R/G/B" (the beauty with prefix "", AOVs as
 * "albedo." / "normals.", Film.cpp:158-228), uncompressed scanlines. */
int ahip_write_png(const char* path, uint32_t width, uint32_t height, const uint8_t* rgb);
int ahip_write_exr(const char* path, uint32_t width, uint32_t height, uint32_t num_layers,
                   const char* const* layer_prefixes, const float* const* layers);
/* Batched persistent form: many views share one BVH; one graph-captured bounce loop. */
int ahip_render_batch(ahip_context* ctx, const ahip_view* views, uint32_t num_views);
/* Adaptive two-pass render into the accumulation buffer (the caller clears it first, as
 * RenderSession does before the base pass).  The reference's GPU path refines uniformly
 * (RenderSession.cpp:577-586); this restates its CPU tile path instead (RenderSession.cpp:549-653):
 * base pass [0, base) over the frame (base_spp 0 = min(32, max(16, total/16)), capped at total),
 * then each tile_size tile (row-major, RenderSession::scheduleTiles) scored by the mean of
 * Film::varianceAt over its pixels (Film.cpp:80-90), and the remaining samples handed out as
 * half uniformly plus half by log(1 + score) share, capped at the remainder; the tiles are
 * rendered as one job list over [base, base + extra_i).  extra_per_tile and tile_variance
 * ([tiles], may be NULL) receive the allocation and the scores. */
int ahip_render_adaptive(ahip_context* ctx, uint32_t width, uint32_t height, uint32_t tile_size,
                         uint32_t total_spp, uint32_t base_spp, uint32_t* extra_per_tile, float* tile_variance);

/* --- denoising AOVs: first-hit reflectance (IMaterial::reflectance) and shading normal (si.n),
 *     captured where PathIntegrator::Li captures them (PathIntegrator.cpp:121-126) and recorded per
 *     pixel like PathIntegrator::renderTile records them in TileBuffer (PathIntegrator.cpp:40-66,
 *     Film.h:108-118).  A camera ray that misses, or hits a mesh without material, reports zeros.
 *     The CUDA backend computes no AOVs; these calls extend the GPU path to the CPU path's
 *     outputs. --- */
/* Turn AOV accumulation by ahip_render_frame on (1) or off (0); the sums live beside the radiance
 * accumulation buffer and are cleared by ahip_clear_accum. */
int ahip_set_aovs(ahip_context* ctx, int enable);
/* [width*height*6]: per pixel, sums over the accumulated samples of the first-hit reflectance
 * (r, g, b) and of the first-hit normal (x, y, z); the sample count is ahip_download_accum's
 * weight. */
int ahip_download_aovs(ahip_context* ctx, float* out);
/* ahip_render_tile plus the tile's AOVs as PathIntegrator::renderTile forms them
 * (PathIntegrator.cpp:58-66): albedo[w*h*3] = the sum of the samples' reflectance times
 * 1/sample_count; normal[w*h*3] = the sum of the samples' normals times 1/sample_count, scaled to
 * unit length when that length exceeds 1e-6. */
int ahip_render_tile_aovs(ahip_context* ctx, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                          uint32_t sample_start, uint32_t sample_count, float* out, float* albedo,
                          float* normal);

/* --- acceleration boundary (IAccelerationStructure::traceBatch / occludedBatch,
 *     include/anacapa/accel/IAccelerationStructure.h:75-87) ---
 * rays: [n*8] = origin xyz, direction xyz, tMin, tMax (Ray, include/anacapa/core/Types.h:253-266).
 * hits: one ahip_hit per ray; mesh_id == ~0u on a miss. */
typedef struct ahip_hit {
    float    t, u, v;      /* ray parameter and barycentrics of the closest hit (curves: u = 0,
                              v = strand parameter, SurfaceInteraction::uv) */
    uint32_t mesh_id;      /* SurfaceInteraction::meshID (IAccelerationStructure.h:21) */
    uint32_t prim_id;      /* triangle index within its mesh / segment index within its strand */
    uint32_t tri_index;    /* commit-order triangle index, or 0x80000000 | strand index for a curve */
} ahip_hit;
int ahip_trace_rays(ahip_context* ctx, const float* rays, uint32_t n, ahip_hit* hits);
int ahip_occluded_rays(ahip_context* ctx, const float* rays, uint32_t n, uint8_t* occluded);
/* ahip_trace_rays_noalloc — no counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the closest hits of ahip_trace_rays, bit for bit, with no allocation of any kind.
 * `rays` and `hits` are DEVICE pointers (the layouts above) that the caller owns.  The traversal
 * stack past its LDS part and the overflow flag live in one fixed device array per GPU, 64 MiB,
 * held until the process ends: a scene whose BVH needs more than AHIP_NOALLOC_MAX_STACK stack
 * entries beyond the LDS ones (a tree deeper than 12 two-level records), or n above
 * AHIP_NOALLOC_MAX_RAYS, is refused with AHIP_E_UNSUPPORTED / AHIP_E_INVALID_ARG.  Calls share
 * that array, so calls that overlap on one GPU (from different contexts or threads) race; calls
 * through one context are serialised.  Synchronous on the context's stream. */
#define AHIP_NOALLOC_MAX_STACK 32u
#define AHIP_NOALLOC_MAX_RAYS  (1u << 30)
int ahip_trace_rays_noalloc(ahip_context* ctx, const float* rays, uint32_t n, ahip_hit* hits);

/* --- sampler (HaltonSampler::get1D, src/sampling/HaltonSampler.h:43-50): dimension dim of
 *     sample idx at pixel (px,py); dims >= 128 come from the per-pixel PCG stream. --- */
int ahip_halton_eval(ahip_context* ctx, uint64_t seed, const uint32_t* px, const uint32_t* py,
                     const uint32_t* idx, const uint32_t* dim, uint32_t n, float* out);

/* --- material probe: IMaterial::evaluate and IMaterial::sample (include/anacapa/shading/IMaterial.h:
 *     52-100) of one material on the device, for the per-lobe parity checks.  flags:
 *     AHIP_FLAG_HAIR_MARSCHNER as for rendering.  in[n*19] = normal, tangent, strand colour, wo, wi,
 *     u.x, u.y, uComponent, curve (0/1: a curve row takes the tangent as its fiber, ShadingContext.h:
 *     39-44); out[n*16] = evaluate(wo, wi): f rgb, pdf, 0 | sample(wo, u, uComponent): wi, f rgb, pdf,
 *     valid | hair effectiveSigmaA rgb, else transmittanceColor rgb (StandardSurface.h:207-225: glass
 *     and alpha-masked surfaces; black for opaque ones).  The frame is the integrator's for a hit
 *     seen along -wo. --- */
int ahip_material_probe(ahip_context* ctx, const ahip_material* mat, uint32_t flags, uint32_t n, const float* in,
                        float* out);

/* --- differentiation (no counterpart in the reference, which has no adjoint) ---
 * The differentiated estimator is the fixed-seed path-replay estimator with detached sampling:
 * every random decision (Halton dims, lobe choice, sampled direction, light choice, Russian
 * roulette) and every pdf / MIS weight is taken at the *sampling* parameters (the prepared
 * scene: its tri_albedo and the roughness texture's texels), while the BSDF values f are taken at
 * the *evaluation* parameters set here.  Its image is smooth in the evaluation parameters, equals
 * the ordinary render when both sets agree, and its derivative is what ahip_adjoint returns.
 *
 * ahip_set_eval_params: evaluation parameters.  tri_albedo [num_tris*3] commit order (NULL =
 * the sampling values); texture = index of the differentiable roughness texture among the
 * scene's textures; texels [w*h] of that texture (NULL = the sampling values). */
int ahip_set_eval_params(ahip_context* ctx, const float* tri_albedo, uint32_t texture, const float* texels);

/* Image of the detached estimator for the current camera, [height*width*5] like
 * ahip_render_tile (r, g, b sums, weight, luminance-squared sum), full film. */
int ahip_replay(ahip_context* ctx, uint32_t sample_start, uint32_t sample_count, float* out);

/* Gradient of sum(dL_dimage * I) with I = per-pixel mean radiance of the detached estimator
 * (full film of the current camera, samples [sample_start, +sample_count)), with respect to the
 * evaluation parameters at their current values.  dL_dimage [height*width*3]; d_albedo
 * [num_tris*3] commit order and d_roughness_tex [texture w*h] are overwritten (either may be NULL).
 * Paths are regenerated from their Halton indices, not stored; per-path state is O(max_depth). */
int ahip_adjoint(ahip_context* ctx, uint32_t width, uint32_t height, uint32_t sample_start,
                 uint32_t sample_count, const float* dL_dimage,
                 float* d_albedo, float* d_roughness_tex);

/* --- measurement --- */
int ahip_get_stats(ahip_context* ctx, ahip_stats* out);
/* Per-bounce occupancy since the last ahip_reset_stats: rays[3*b + k] = continuation (k = 0),
 * shadow (1) and BSDF-sampled MIS (2) rays of bounce b summed over passes; ms[b] = HIP-event time of
 * bounce b's traversal and shade launches (launches outside graphs only).  Either array may be
 * NULL; *num_bounces = bounces recorded (entries beyond `capacity` are not written). */
int ahip_get_bounce_stats(ahip_context* ctx, uint32_t capacity, uint64_t* rays, double* ms, uint32_t* num_bounces);
int ahip_reset_stats(ahip_context* ctx);
/* Re-run the last render_frame workload with per-ray node/primitive counters
 * (a separate, untimed pass) and fill nodes_visited/prims_tested. */
int ahip_count_traffic(ahip_context* ctx, uint32_t width, uint32_t height,
                       uint32_t sample_start, uint32_t sample_count);
/* Launch-shape knobs (autotuner): paths per wavefront pass and tile width of the
 * pixel->path swizzle; 0 keeps the current value. */
int ahip_set_launch(ahip_context* ctx, uint32_t max_paths, uint32_t tile);

/* --- failure injection for the CPU-fallback tests (RenderSession.cpp:441-464, 546-547) --- */
int ahip_fail_next_launch(ahip_context* ctx);

/* --- host helper: reference pinhole camera (Camera::makePinhole, IIntegrator.h:49-74) --- */
int ahip_camera_pinhole(const float from[3], const float at[3], const float up[3],
                        float vfov_degrees, uint32_t width, uint32_t height, ahip_camera* out);

/* --- USD-lite scene ingest (host only; no device context needed) ---------------------------
 * Reads one text USD layer (.usda) without the USD SDK into a flat scene for ahip_prepare,
 * restating the reference's loadUSD (src/scene/usd/USDLoader.h:26-29, USDLoader.cpp:1166-1833,
 * returning LoadedScene, include/anacapa/scene/SceneLoader.h:20-33): meshes fan-triangulated
 * into world space with three vertices per triangle, GeomSubset and displayColor materials,
 * UsdPreviewSurface constants as standard_surface, RectLight / SphereLight / DiskLight and
 * emissive meshes as area lights, the pinhole camera (explicit path, RenderSettings camera, or
 * the first camera), upAxis Z turned to Y-up.  A layer that needs composition (references,
 * payloads, sublayers, variants), animation, image textures, DistantLight / DomeLight or a
 * thin-lens camera is refused with AHIP_E_UNSUPPORTED; malformed text with AHIP_E_INVALID_ARG.
 * The message goes to err (err_len bytes, may be NULL).  The scene owns every array the
 * descriptor points to; both stay valid until ahip_usd_free. */
typedef struct ahip_usd_scene ahip_usd_scene;
int ahip_usd_load(const char* path, uint32_t film_width, uint32_t film_height, const char* camera_path,
                  ahip_usd_scene** out, char* err, size_t err_len);
int ahip_usd_load_text(const char* text, size_t len, uint32_t film_width, uint32_t film_height,
                       const char* camera_path, ahip_usd_scene** out, char* err, size_t err_len);
void ahip_usd_free(ahip_usd_scene* scene);
/* LoadedScene::geomPool + sceneView as the descriptor ahip_prepare takes. */
const ahip_scene_desc* ahip_usd_scene_desc(const ahip_usd_scene* scene);
/* LoadedScene::camera (buildCamera, USDLoader.cpp:1084-1161): eye, target, up, vertical field of
 * view for the film size given at load, and the pinhole built from them; any output may be NULL.
 * AHIP_E_NOT_PREPARED when the layer has no camera (the caller picks its default). */
int ahip_usd_camera(const ahip_usd_scene* scene, float from[3], float at[3], float up[3],
                    float* vfov_degrees, ahip_camera* camera);
/* MeshDesc::name of mesh i (the prim name), NULL past the end. */
const char* ahip_usd_mesh_name(const ahip_usd_scene* scene, uint32_t mesh);

/* --- hair curves in interchange bases (host only) ------------------------------------------
 * The curve handling of the reference's Alembic reader without the Alembic SDK
 * (loadAlembicCurves, src/scene/alembic/AlembicLoader.h:42-45; processICurves and the basis
 * conversions, AlembicLoader.cpp:58-306): one ICurves sample (per-curve vertex counts, points,
 * widths, per-curve colours, the object's world matrix) becomes endpoint-sharing cubic Bezier
 * strands with per-CV diameters, ready for ahip_scene_desc's strand arrays. */
#define AHIP_CURVE_LINEAR      0   /* kLinear */
#define AHIP_CURVE_BEZIER      1   /* kCubic, kBezierBasis (3n+1 CVs, used as is) */
#define AHIP_CURVE_BSPLINE     2   /* kCubic, kBsplineBasis (uniform) */
#define AHIP_CURVE_CATMULL_ROM 3   /* kCubic, kCatmullromBasis */

/* AlembicCurveOptions (AlembicLoader.h:12-20). */
typedef struct ahip_curve_options {
    float    default_width;   /* diameter when the curves carry no widths (0.005 in the reference) */
    float    width_scale;     /* multiplies authored widths (1.0) */
    uint32_t material_index;  /* StrandDesc::materialIndex of every strand (baseMaterialIndex) */
    uint32_t pad;
} ahip_curve_options;

/* The converted strands, in ahip_scene_desc's strand layout; owned by the curve set. */
typedef struct ahip_strands {
    uint32_t        num_strands;
    uint32_t        num_cvs;
    uint32_t        num_skipped;          /* curves dropped for too few or mis-counted CVs */
    uint32_t        pad;
    const uint32_t* strand_cv_offset;     /* [num_strands+1] */
    const float*    curve_cvs;            /* [num_cvs*3] world space */
    const uint32_t* strand_width_offset;  /* [num_strands+1] */
    const float*    strand_widths;        /* [num_cvs] per-CV diameters */
    const uint32_t* strand_mesh_index;    /* [num_strands] */
    const float*    strand_color;         /* [num_strands*3], white where no colours were given */
} ahip_strands;

typedef struct ahip_curve_set ahip_curve_set;
/* widths: num_widths == total points (per vertex), == num_curves (per curve), == 1 (constant),
 * 0 (none: default_width); colors: [num_curves*3] or NULL; world: row-major 4x4 acting on column
 * vectors (Mat4f, Types.h:144-150) or NULL for identity. */
int ahip_curves_to_bezier(uint32_t basis, uint32_t num_curves, const uint32_t* counts, const float* points,
                          const float* widths, uint32_t num_widths, const float* colors, const float* world,
                          const ahip_curve_options* opts, ahip_curve_set** out, char* err, size_t err_len);
const ahip_strands* ahip_curve_set_strands(const ahip_curve_set* set);
void ahip_curve_set_free(ahip_curve_set* set);

#ifdef __cplusplus
}
#endif
#endif /* ANACAPA_HIP_H */
