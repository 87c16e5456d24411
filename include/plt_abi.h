/* plt_abi.h — C ABI of the MI355X-native PLT generalized-ray path-tracing hot path.
 *
 * Boundary: this library replaces the device side of the reference's PLT material interface
 * (Source/Falcor/Rendering/Materials/IMaterialInstance.slang:122-184: eval / sample / pdf /
 * lambdaPdf) and the PLTPT render pass's GPU work (Source/RenderPasses/PLTPT/PLTPT.cpp:357-499,
 * two DXR dispatches per tile). Host callers pass plain device pointers, sizes and a
 * hipStream_t (as void*). No C++ or torch types cross this boundary, no exceptions escape it:
 * every entry point returns a plt_status and plt_last_error() describes the last failure of
 * the calling thread.
 *
 * Units follow the reference (Source/Falcor/Rendering/PLT.slang:80-82): wavelengths in nm,
 * wavenumbers in 1/um, beam distances in m, coherence in mm, surface heights/lengths in um.
 */
#ifndef PLT_ABI_H
#define PLT_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct plt_context plt_context;
typedef void* plt_stream; /* hipStream_t; NULL = the default (null) stream */

/* Status codes. Replaces the reference's C++ exceptions / logError
 * (PLTPT.cpp:173, :409; Core/Errors.h). */
typedef enum plt_status {
    PLT_OK = 0,
    PLT_WARN_UNKNOWN_KEY = 1, /* PLTPT.cpp:173 logs an error and continues; so do we */
    PLT_EINVAL = -1,
    PLT_ENOMEM = -2,
    PLT_EHIP = -3,
    PLT_ENOTIMPL = -4,
    PLT_ESTATE = -5
} plt_status;

/* Thread-local description of the last failure (empty string if none). */
const char* plt_last_error(void);
/* ABI version: major*10000 + minor*100 + patch. */
int plt_abi_version(void);

/* ---------------------------------------------------------------------------------------
 * Context. One device, one process; not thread-safe (external synchronisation), as the
 * reference's single render thread (PLTPT.cpp:487-495).
 * ------------------------------------------------------------------------------------- */
int plt_context_create(int device, plt_context** out);
int plt_context_destroy(plt_context* ctx);

/* ---------------------------------------------------------------------------------------
 * Spectral profiles. Layout of Source/Falcor/Scene/Material/BasicMaterialData.slang:43-54
 * (SpectralProfile, kMaxBins = 128), evaluated as Scene::evalSpectrum (Scene/Scene.slang:251-295).
 * ------------------------------------------------------------------------------------- */
#define PLT_SPECTRAL_PROFILE_MAX_BINS 128
typedef struct plt_spectral_profile {
    float rgb[3];
    float intensity;
    uint32_t bins;
    float min_wavelength;
    float max_wavelength;
    float pdf[PLT_SPECTRAL_PROFILE_MAX_BINS];
    float icdf[PLT_SPECTRAL_PROFILE_MAX_BINS];
} plt_spectral_profile;

/* Host-side construction of a profile from a tabulated spectrum (npts >= 2, wavelengths
 * non-decreasing): PiecewiseLinearSpectrum -> SampledSpectrum (5 nm bins) -> genSpectralProfile
 * (Utils/Color/Spectrum.h:88-123, Utils/Color/SampledSpectrum.h:64-74, Scene/SceneBuilder.cpp:911-960).
 * compute_icdf != 0 normalises pdf and fills icdf (emission spectra); 0 keeps raw values (n, k). */
int plt_spectral_profile_build(uint32_t npts, const float* wavelengths, const float* values, int compute_icdf,
                               plt_spectral_profile* out);
/* Spectral tables from files (host only). A `.spd` file holds "wavelength value" lines; blank and
 * '#' lines are skipped and the first line that does not parse ends the table
 * (PiecewiseLinearSpectrum::fromFile, Utils/Color/Spectrum.cpp:89-112). plt_spectrum_load_spd
 * writes up to `capacity` entries and the table's size to *count (capacity 0 with NULL arrays
 * queries the size). */
int plt_spectrum_load_spd(const char* path, uint32_t capacity, float* wavelengths, float* values, uint32_t* count);
/* The profile of a constant spectrum (SampledSpectrum<float>(value): one bin over 380-720 nm). */
int plt_spectral_profile_constant(float value, plt_spectral_profile* out);
/* SceneBuilder::addSpectralProfileFromMaterial (SceneBuilder.cpp:972-983): the n and k profiles of
 * <tables_dir>/ior/<name>.eta.spd and .k.spd (no inverse CDF); name "none" gives n = 1, k = 0. */
int plt_spectral_profiles_from_material(const char* tables_dir, const char* name, plt_spectral_profile* n_out,
                                        plt_spectral_profile* k_out);
/* SceneBuilder::addSpectralProfileForEmitterType (SceneBuilder.cpp:984-1000): the emission profile
 * (with inverse CDF) of <tables_dir>/emission/<name>.spd scaled by `scale`; PLT_EINVAL if absent. */
int plt_spectral_profile_from_emitter(const char* tables_dir, const char* name, float scale, plt_spectral_profile* out);
/* The profile of a SampledSpectrum given directly (n values spaced uniformly over
 * [lambda_start, lambda_end], no resampling): genSpectralProfile on it, as SceneBuilder does for
 * spectra made from RGB (SpectrumUtils::rgbToSpectrum, Utils/Color/SpectrumUtils.h:154-215). */
int plt_spectral_profile_build_sampled(uint32_t n, float lambda_start, float lambda_end, const float* values,
                                       int compute_icdf, plt_spectral_profile* out);
/* Uploads `count` host profiles to the device; profile ids are their indices. */
int plt_context_set_spectral_profiles(plt_context* ctx, const plt_spectral_profile* profiles, uint32_t count);

/* ---------------------------------------------------------------------------------------
 * Sample generator (Utils/Sampling/UniformSampleGenerator.slang:51-69: SplitMix64 seeding of
 * xoshiro128**, Utils/Sampling/Pseudorandom/Xoshiro.slang:52-66). Writes, for stream s in
 * [0, n_streams) with pixel = (s % width, s / width), `draws` raw uint32 values to
 * out[d * n_streams + s]. Device pointer `out`.
 * ------------------------------------------------------------------------------------- */
int plt_sample_generator_stream(uint32_t n_streams, uint32_t width, uint32_t sample_number, uint32_t draws,
                                uint32_t* out, plt_stream stream);
/* The same for either generator type: type 1 = UniformSampleGenerator (as above), type 0 =
 * TinyUniformSampleGenerator (Utils/Sampling/TinyUniformSampleGenerator.slang:55-66: LCG
 * (Pseudorandom/LCG.slang:45-61) seeded with blockCipherTEA(interleave_32bit(pixel),
 * sampleNumber).x, Utils/Math/HashUtils.slang:59-71). */
int plt_sample_generator_stream_typed(int type, uint32_t n_streams, uint32_t width, uint32_t sample_number,
                                      uint32_t draws, uint32_t* out, plt_stream stream);

/* ---------------------------------------------------------------------------------------
 * Batched BSDF queries (the IMaterialInstance eval + pdf pair that next-event estimation runs,
 * Source/RenderPasses/PLTPT/pltpt_solve.rt.slang:219-243 and :289-292).
 *
 * A batch of n queries is PLT_Q_PLANES structure-of-arrays planes of n 32-bit words
 * (plane p of query i at word p*n + i). Directions are in the local shading frame (N = +z):
 * wi = sd.V = direction towards the light (the beam arrives along -wi), wo = direction
 * towards the eye. The incident beam carries one wavelength with Stokes (S0..S3) and an
 * isotropic coherence matrix whose coherence area at the query wavelength equals
 * PLT_Q_COH_AREA (um^2) at 1 mm from its source. Evaluation runs on the NonDelta lobe.
 *
 * PLT_Q_MAT   : bits 31..28 material type (plt_material_type); for conductors bits 13..0 the
 *               n profile id and bits 27..14 the k profile id.
 * PLT_Q_PARAM : two IEEE binary16 values, as the reference stores material parameters
 *               (BasicMaterialData.slang:70-83). Dielectric: lo = IoR, hi = dispersion B
 *               (Cauchy, PLTDielectricMaterial.cpp:87-101). Conductor: lo = exterior IoR,
 *               hi = PSD exponent gamma (PLTConductorMaterial.cpp:54-58).
 * PLT_Q_ROUGH_Q : RMS surface height q (um), the coherent/specular split
 *               alpha = exp(-((|cos_i|+|cos_o|) q k)^2) (FractalSurface.slang:77-88).
 * PLT_Q_ROUGH_L : correlation length l (um): Gaussian PSD sigma2 = l^2 (GaussianSurface.slang:97-109),
 *               K-correlation T = min(l^2, 5000) (FractalSurface.slang:44-51,92-95).
 *
 * Results: PLT_R_PLANES planes of n floats: outgoing Stokes S0..S3 and pdf(wi -> wo).
 * ------------------------------------------------------------------------------------- */
enum {
    PLT_Q_WI_X = 0, PLT_Q_WI_Y, PLT_Q_WI_Z,
    PLT_Q_WO_X, PLT_Q_WO_Y, PLT_Q_WO_Z,
    PLT_Q_LAMBDA, PLT_Q_MAT, PLT_Q_PARAM,
    PLT_Q_ROUGH_Q, PLT_Q_ROUGH_L, PLT_Q_COH_AREA,
    PLT_Q_S0, PLT_Q_S1, PLT_Q_S2, PLT_Q_S3,
    PLT_Q_PLANES
};
enum { PLT_R_S0 = 0, PLT_R_S1, PLT_R_S2, PLT_R_S3, PLT_R_PDF, PLT_R_PLANES };

typedef enum plt_material_type {
    /* The reference's MaterialType values (Scene/Material/MaterialTypes.slang:36-53); Unknown (0)
     * and Standard (1) are not PLT materials and are refused. */
    PLT_MAT_DIFFUSE = 2,                   /* PLTDiffuseBSDF.slang:39-155 (Lambertian; the emitter material) */
    PLT_MAT_CONDUCTOR = 3,                 /* PLTConductorBSDF.slang:44-242 (K-correlation PSD, tabulated n,k) */
    PLT_MAT_DIELECTRIC = 4,                /* PLTDielectricBSDF.slang:43-361 (Gaussian PSD) */
    PLT_MAT_OPAQUE_DIELECTRIC = 5,         /* PLTOpaqueDielectricBSDF.slang:43-212 */
    PLT_MAT_THIN_DIELECTRIC = 6,           /* PLTThinDielectricBSDF.slang:43-269 (+ Birefringence.slang) */
    PLT_MAT_COATED_CONDUCTOR = 7,          /* PLTCoatedConductorBSDF.slang:45-313 */
    PLT_MAT_COATED_OPAQUE_DIELECTRIC = 8,  /* PLTCoatedOpaqueDielectricBSDF.slang:45-312 */
    PLT_MAT_GRATED_CONDUCTOR = 9,          /* PLTDiffractionGratedConductorBSDF.slang:45-290 */
    PLT_MAT_GRATED_OPAQUE_DIELECTRIC = 10, /* PLTDiffractionGratedOpaqueDielectricBSDF.slang:45-283 */
    PLT_MAT_MULTILAYER_STACK = 11          /* PLTMultiLayeredStackBSDF.slang:42-201 */
} plt_material_type;

/* Synthetic batch: query i draws 14 numbers from UniformSampleGenerator(pixel=(i & 0xFFFF,
 * i >> 16), sample_number=seed): type (u<0.5: dielectric), IoR U[1.3,1.8] (binary16, Abbe 100
 * dispersion), q U[0.01,0.3] um, l U[1,20] um, coherence area U[1,100] um^2, wi and wo
 * cosine-hemisphere (concentric), transmission side for dielectrics (u<0.5 flips wo.z),
 * wavelength U[380,780] nm, Stokes direction uniform on the Poincare sphere with degree of
 * polarisation U[0,1]. Conductors use profile ids 0 (n) and 1 (k). Device pointer `queries`. */
int plt_bsdf_generate_queries(plt_context* ctx, uint32_t n, uint32_t seed, uint32_t* queries, plt_stream stream);

/* Evaluates a batch (device pointers). Material-divergent batches are accepted as is. */
int plt_bsdf_eval_queries(plt_context* ctx, uint32_t n, const uint32_t* queries, float* results, plt_stream stream);

/* ---------------------------------------------------------------------------------------
 * Materials. The reference's BasicMaterialData parameter block (Scene/Material/
 * BasicMaterialData.slang:63-107): binary16 fields stored as raw bits, spectral-profile ids
 * (0xFFFFFFFF = none). Field use per type follows each *Material.slang setupMaterialInstance:
 *   dielectric   : specular = (dispersion B, roughness, -, ext IoR), ior, transmission
 *   conductor    : specular = (gamma, roughness, -, -), ior = ext IoR, n/k profiles
 *   diffuse      : base_color, emission profile
 *   opaque diel. : base_color, specular.y = roughness
 *   coated opaque: base_color, specular = (-, roughness, thickness scale, -), data1 = (thickness,
 *                  -, -, coat IoR), ior = ext IoR
 *   grated cond. : base_color, specular = (height scale, power mult), data1 = (pitch x, pitch y,
 *                  direction, type|lobes bits), data2 = (height, roughness), ior, n/k profiles.
 *                  Radial gratings (type bit 0x10, DiffractionGrating.slangh:37) orient the grating
 *                  by the texture coordinates, which these meshes do not carry: PLT_ENOTIMPL.
 *   grated opaque: as the grated conductor over an opaque dielectric base: base_color = albedo,
 *                  no ior or n/k profiles (PLTDiffractionGratedOpaqueDielectricMaterial.slang:40-70)
 *   coated cond. : base_color = specular colour, specular = (-, roughness, thickness scale, -),
 *                  data1 = (thickness um, -, -, coat IoR), ior = ext IoR, n/k profiles
 *   thin diel.   : base_color = (birefringence, optic axis x, optic axis y, alpha),
 *                  specular = (birefringence scale, -, -, thickness mm), ior = mean IoR
 *   multilayer   : base_color = base albedo, data1 = (thickness 1 um, thickness 2 um, -, layer
 *                  count as a raw uint16), data2 = (IoR 1, IoR 2), ior = ext IoR
 * ------------------------------------------------------------------------------------- */
#define PLT_NO_PROFILE 0xFFFFFFFFu
typedef struct plt_material {
    uint32_t type; /* plt_material_type */
    uint16_t base_color[4];
    uint16_t specular[4];
    uint16_t data1[4];
    uint16_t data2[4];
    uint16_t transmission[3];
    uint16_t ior;
    uint32_t emission_profile;
    uint32_t ior_n_profile;
    uint32_t ior_k_profile;
} plt_material;

/* Batched material interface: the reference's IMaterialInstance eval / pdf / lambdaPdf and
 * sample (Rendering/Materials/IMaterialInstance.slang:122-184) for `n` queries of any material of
 * the device array `materials`, as SoA float planes (`n` floats each; integer fields are stored as
 * their bit patterns). Queries live in a local frame with N = +z. eval: wi = sd.V (towards the
 * light), wo = toEye, the beam arrives along -wi with 4 HWSS wavelengths and Stokes vectors and
 * the given coherence area (um^2); results are the 4 outgoing Stokes vectors, pdf(toLight = wi
 * | sd.V = wo) at wavelength 0 and lambdaPdf. sample: query i draws from the sample generator
 * (pixel (i, 0), sample number `frame`). Profiles are the context's. A material index out of
 * range, a type outside the PLT MaterialType values 2-11 or a radial grating gives zeros (eval) or
 * valid = 0 (sample). */
enum {
    PLT_MQ_WI_X = 0, PLT_MQ_WI_Y, PLT_MQ_WI_Z, PLT_MQ_WO_X, PLT_MQ_WO_Y, PLT_MQ_WO_Z,
    PLT_MQ_LOBE, PLT_MQ_MAT, PLT_MQ_LAMBDA /* 4 planes */, PLT_MQ_S = PLT_MQ_LAMBDA + 4 /* 16 planes, [j*4 + c] */,
    PLT_MQ_AREA = PLT_MQ_S + 16, PLT_MQ_PLANES
};
enum { PLT_MR_S = 0 /* 16 planes */, PLT_MR_PDF = 16, PLT_MR_LAMBDA_PDF /* 4 planes */, PLT_MR_PLANES = PLT_MR_LAMBDA_PDF + 4 };
enum { PLT_MSQ_WI_X = 0, PLT_MSQ_WI_Y, PLT_MSQ_WI_Z, PLT_MSQ_MAT, PLT_MSQ_LAMBDA, PLT_MSQ_PLANES };
enum { PLT_MSR_WO_X = 0, PLT_MSR_WO_Y, PLT_MSR_WO_Z, PLT_MSR_PDF, PLT_MSR_WEIGHT, PLT_MSR_LOBE, PLT_MSR_VALID, PLT_MSR_PLANES };
int plt_material_eval_batch(plt_context* ctx, const plt_material* materials, uint32_t n_materials, uint32_t n,
                            const float* queries, float* results, plt_stream stream);
int plt_material_sample_batch(plt_context* ctx, const plt_material* materials, uint32_t n_materials, uint32_t n,
                              uint32_t frame, const float* queries, float* results, plt_stream stream);

/* Pinhole camera (Scene/Camera/Camera.slang:46-78, Camera.cpp:187-193). */
typedef struct plt_camera {
    float pos[3];
    float target[3];
    float up[3];
    float fov_y;          /* radians */
    float aspect;         /* width / height */
    float focal_distance;
    float jitter[2];      /* sample offset as a fraction of the image width / height (the reference's
                             data.jitterX/Y: a pixel offset in [-0.5, 0.5) divided by the frame size);
                             0 = pixel centre. Replaced per frame by a jitter pattern
                             (plt_pass_set_sample_pattern). */
    float polarizer;      /* camera polarizer strength in [0, 1] (CameraData.slang:67-68, Camera.cpp:315-321):
                             0 = none; p mixes (1 - p) I + p MuellerPolarizer(polarizer_theta) into every
                             measured beam (pltpt_solve.rt.slang:111-126) */
    float polarizer_theta;/* the reference's polarizer angle (radians); MuellerPolarizer(theta) takes
                             cos/sin of theta itself, so the transmission axis lies at theta / 2 */
} plt_camera;

/* ---------------------------------------------------------------------------------------
 * Scenes. A scene is triangle meshes (world space, one material per mesh), spectral profiles,
 * materials and a camera, with a software BVH kept in HBM, built on the host (binned SAH) or on
 * the GPU (plt_scene_build_bvh). Replaces the reference's DXR BLAS/TLAS build (Scene/Scene.cpp:3092,
 * :3624).
 * ------------------------------------------------------------------------------------- */
typedef struct plt_scene plt_scene;

/* Procedural workload scene: a ground quad (PLTDiffuse), `sphere_count` icospheres of
 * subdivision `subdivision` (20*4^s triangles each) at uniform random positions on the quad,
 * resting on it, radius U[0.2, 1.0], materials drawn uniformly from the first
 * `material_kinds` (1-8) of {dielectric, conductor, coated opaque dielectric, grated conductor,
 * birefringent thin dielectric (thickness U[0.2, 2] um, birefringence U[0.01, 0.2]), multilayer
 * stack (20 layers 1.38 / 2.3, quarter-wave at 550 nm), coated conductor, opaque dielectric}
 * (grating pitch U[0.5, 2] um), one quad area light (PLTDiffuse emitter, 6500 K blackbody
 * tabulated in 5 nm bins), and a pinhole camera of the given aspect. Deterministic in `seed`. */
typedef struct plt_procedural_scene_desc {
    uint32_t seed;
    uint32_t sphere_count;
    uint32_t subdivision;
    uint32_t material_kinds;
    float ground_size;
    float aspect;
} plt_procedural_scene_desc;

/* ctx == NULL builds a host-only scene (arrays and BVH; plt_scene_trace then returns PLT_ESTATE).
 * The BVH is the host builder's (plt_scene_build_bvh rebuilds it on the device). */
int plt_scene_create_procedural(plt_context* ctx, const plt_procedural_scene_desc* desc, plt_scene** out);

/* A scene from caller arrays: what the reference's SceneBuilder hands the renderer for triangle
 * meshes (Scene/SceneBuilder.cpp: vertex positions and normals, 32-bit indices, per-mesh material
 * ids, the material table and the spectral profiles built by genSpectralProfile,
 * SceneBuilder.cpp:911-960) plus the camera. Triangles are in mesh order: tri_mesh[t] names the
 * mesh of triangle t. Normals are stored as binary16, as Falcor's packed vertex data stores them
 * (Scene/SceneTypes.slang:219-242). Arrays are copied; the BVH is built here (binned SAH).
 * Returns PLT_EINVAL for an index, mesh, material or profile id out of range. */
typedef struct plt_scene_desc {
    uint32_t n_vertices, n_triangles, n_meshes, n_materials, n_profiles;
    const float* positions;       /* 3 x n_vertices */
    const float* normals;         /* 3 x n_vertices */
    const uint32_t* indices;      /* 3 x n_triangles */
    const uint32_t* tri_mesh;     /* n_triangles */
    const uint32_t* mesh_material;/* n_meshes */
    const plt_material* materials;
    const plt_spectral_profile* profiles;
    plt_camera camera;
} plt_scene_desc;
int plt_scene_create(plt_context* ctx, const plt_scene_desc* desc, plt_scene** out);
int plt_scene_destroy(plt_scene* scene);

/* Host view of a scene's arrays (valid until plt_scene_destroy). */
typedef struct plt_scene_arrays {
    uint32_t n_vertices, n_triangles, n_meshes, n_materials, n_profiles, n_bvh_nodes;
    const float* positions;          /* 3 per vertex, world space */
    const float* normals;            /* 3 per vertex, binary16-rounded (SceneTypes.slang:240-241) */
    const uint32_t* indices;         /* 3 per triangle */
    const uint32_t* tri_mesh;        /* mesh id per triangle */
    const uint32_t* mesh_first_tri;  /* first triangle of each mesh */
    const uint32_t* mesh_material;   /* material id per mesh */
    const plt_material* materials;
    const plt_spectral_profile* profiles;
    plt_camera camera;
    uint32_t bvh_max_depth;          /* levels of the 4-wide BVH (root = 1) */
} plt_scene_arrays;
int plt_scene_get_arrays(const plt_scene* scene, plt_scene_arrays* out);

/* BVH builders. The reference builds its acceleration structures on the device
 * (Scene/Scene.cpp:3092-3400 buildBlas, :3624-3700 buildTlas); both builders here write the same
 * 4-wide node and leaf-order triangle layout, and closest hits do not depend on the tree, so a
 * scene renders bit-identically with either (only the speed differs).
 *   PLT_BVH_HOST_SAH: binned SAH on the host's cores (what plt_scene_create runs);
 *   PLT_BVH_DEVICE_PLOC: on the scene's GPU, parallel locally-ordered clustering of the triangles
 *     sorted along a Morton curve (Meister and Bittner 2018), collapsed to 4-wide nodes with leaves
 *     of at most four triangles. */
typedef enum plt_bvh_builder { PLT_BVH_HOST_SAH = 0, PLT_BVH_DEVICE_PLOC = 1 } plt_bvh_builder;
typedef struct plt_bvh_info {
    uint32_t builder;     /* the builder whose tree the scene holds */
    uint32_t n_nodes;     /* 4-wide nodes */
    uint32_t max_depth;   /* levels of 4-wide nodes (root = 1) */
    uint32_t iterations;  /* clustering iterations (device builder; 0 for the host's) */
    float build_ms;       /* wall-clock time of the build (the device builder's with its allocations) */
    uint32_t pad[3];
} plt_bvh_info;
/* Rebuilds the tree of a scene created with a context and replaces the one it holds (after waiting
 * for the device: passes bound to the scene read the tree at their next execute). The device
 * builder runs on `stream` and synchronizes it. PLT_EINVAL when the device tree would need more
 * levels than the traversal stack holds (33); the scene then keeps its tree. */
int plt_scene_build_bvh(plt_scene* scene, int builder, plt_stream stream, plt_bvh_info* info);
/* plt_scene_create with a choice of builder: PLT_BVH_DEVICE_PLOC builds no host tree (it falls
 * back to the host builder if the device tree would be too deep; info->builder says which ran).
 * ctx must not be NULL for the device builder. info may be NULL. */
int plt_scene_create_ex(plt_context* ctx, const plt_scene_desc* desc, int builder, plt_scene** out, plt_bvh_info* info);
/* The tree a scene holds: info, and (arrays not NULL) device copies to host memory of its
 * n_nodes x 128 B nodes and n_triangles x 48 B leaf-order triangle records (v0 with the triangle
 * id's bits in w, v1, v2). Sizes smaller than these return PLT_EINVAL. */
int plt_scene_get_bvh(const plt_scene* scene, plt_bvh_info* info, void* nodes, size_t nodes_bytes, void* tris,
                      size_t tris_bytes);

/* Analytic light sources: the reference's LightData (Scene/Lights/LightData.slang:36-66), with
 * the fields its light classes set (Scene/Lights/Light.cpp:115-400): position, direction, spot
 * opening/penumbra angles, cosine of the subtended half-angle (PLT beam solid angle), surface
 * area (PLT beam area for point lights, shape area for area lights), the spectral profile of the
 * intensity, and for area and distant lights the local-to-world matrix and its inverse transpose
 * (row-major 4x4). Area lights keep the position the reference leaves them (the origin unless
 * set); their beams are sourced there, as the reference sources them at LightData::posW. */
typedef enum plt_light_type {
    PLT_LIGHT_POINT = 0, PLT_LIGHT_DIRECTIONAL = 1, PLT_LIGHT_DISTANT = 2, PLT_LIGHT_RECT = 3, PLT_LIGHT_DISC = 4,
    PLT_LIGHT_SPHERE = 5
} plt_light_type;
typedef struct plt_light {
    float pos[3];
    uint32_t type;
    float dir[3];
    float opening_angle;
    float cos_opening_angle;
    float cos_subtended_angle;
    float penumbra_angle;
    uint32_t intensity_profile;
    float surface_area;
    float pad[3];
    float trans[16];
    float trans_it[16];
} plt_light;
/* Replaces the scene's analytic lights (Scene::addLight). A pass reads them when the scene is
 * bound (plt_pass_set_scene). Returns PLT_EINVAL for a bad type or profile id. */
int plt_scene_set_lights(plt_scene* scene, const plt_light* lights, uint32_t count);
/* Replaces the scene's camera (the reference's camera setters between frames, Camera.cpp:150-330,
 * among them setPolarizer / setPolarizerTheta): a pass bound to the scene renders its next execute
 * with it. */
int plt_scene_set_camera(plt_scene* scene, const plt_camera* camera);

/* Environment map light (Scene/Lights/EnvMap.{h,cpp}, EnvMap.slang, EnvMapData.slang) with the
 * importance sampling of Rendering/Lights/EnvMapSampler.{cpp,slang}: a lat-long RGBA32F image of
 * width x height texels (host memory, row 0 at the +y pole, u = 0.5 towards -z; bilinear lookup
 * wrapping in u and clamped in v), the local-to-world rotation (row-major 3x3, NULL = identity;
 * the reference builds it from its rotation angles), and radiance scale intensity x tint (tint
 * NULL = white). The 512 x 512 hierarchical importance map is built here. width or height 0
 * removes the map. A pass reads it when the scene is bound (plt_pass_set_scene). */
int plt_scene_set_env_map(plt_scene* scene, uint32_t width, uint32_t height, const float* rgba, const float* transform,
                          float intensity, const float* tint);

/* ---------------------------------------------------------------------------------------
 * Ray casting against a scene (device pointers). Rays: 8 planes of n floats (origin xyz,
 * direction xyz, tmin, tmax). Closest hit (mode 0, the reference's TraceRay with a
 * closest-hit shader) writes 4 planes of n 32-bit words: triangle id (0xFFFFFFFF on miss),
 * barycentrics b1, b2 (the weights of vertices 1 and 2, DXR convention) and t. Any hit
 * (mode 1, RAY_FLAG_ACCEPT_FIRST_HIT_AND_END_SEARCH, ShadowRays.slangh:65-88) writes one
 * plane: 1 if any triangle is hit with tmin <= t <= tmax, else 0. Intersection is watertight
 * and deterministic; among equal distances the smallest triangle id wins.
 * ------------------------------------------------------------------------------------- */
int plt_scene_trace(plt_scene* scene, uint32_t n, const float* rays, uint32_t* hits, int mode, plt_stream stream);

/* ---------------------------------------------------------------------------------------
 * The PLTPT render pass (Source/RenderPasses/PLTPT/PLTPT.h:42-121). Replaces the reference's
 * plugin class (registerPlugin / RenderPass::create/execute/setScene, PLTPT.cpp:37-46,139-573).
 * Options are the reference's dictionary keys with numeric values (booleans as 0/1, enums as
 * their integer value) and the reference's defaults; an unknown key returns
 * PLT_WARN_UNKNOWN_KEY and changes nothing (PLTPT.cpp:173 logs and continues).
 * ------------------------------------------------------------------------------------- */
typedef struct plt_pass plt_pass;

/* type must be "PLTPT". ctx may be NULL for a pass that only holds options (set_scene and
 * execute then return PLT_ESTATE). */
int plt_pass_create(plt_context* ctx, const char* type, plt_pass** out);
int plt_pass_destroy(plt_pass* pass);
int plt_pass_set_option(plt_pass* pass, const char* key, double value);
int plt_pass_get_option(const plt_pass* pass, const char* key, double* value);
/* i-th option key (NULL past the end), for getScriptingDictionary-style enumeration. */
const char* plt_pass_option_name(uint32_t i);
/* Binds a scene (NULL unbinds) and resets the frame counter to 0 (PLTPT.cpp:501-552). */
int plt_pass_set_scene(plt_pass* pass, plt_scene* scene);
/* Frame counter: seeds UniformSampleGenerator(pixel, frame) of the next execute. */
int plt_pass_set_frame(plt_pass* pass, uint32_t frame);

/* Camera jitter pattern of the G-buffer pass that runs before PLTPT in the reference graph
 * (scripts/PLTPathTracer.py:16: GBufferRT samplePattern Halton, sampleCount 16;
 * GBufferBase.h:39-45, GBufferBase.cpp:202-234, Camera.cpp:63-70): each execute draws the next
 * offset in [-0.5, 0.5)^2 pixels and uses offset / (width, height) as the camera jitter.
 * Center (the default) keeps the camera's own jitter. Setting a pattern, or a scene, restarts it. */
enum { PLT_SAMPLE_PATTERN_CENTER = 0, PLT_SAMPLE_PATTERN_DIRECTX = 1, PLT_SAMPLE_PATTERN_HALTON = 2,
       PLT_SAMPLE_PATTERN_STRATIFIED = 3 };
int plt_pass_set_sample_pattern(plt_pass* pass, int pattern, uint32_t sample_count);
/* The camera jitter (fractions of width and height) the last execute used. */
int plt_pass_get_jitter(const plt_pass* pass, float* xy);
/* The first n offsets (pixels, xy pairs) of a fresh generator of the pattern (host memory), as
 * CPUSampleGenerator::next returns them (HaltonSamplePattern.cpp:56-70, DxSamplePattern.h,
 * StratifiedSamplePattern.cpp:67-88). */
int plt_sample_pattern_generate(int pattern, uint32_t sample_count, uint32_t n, float* xy);

/* I/O of one execute (PLTPT.cpp:65-78): caller-owned device buffers of width*height float4
 * (RGBA32F): color (spectral radiance -> linear Rec.709), albedo and shading normal of the first
 * non-delta vertex, the normal in view space (the rotation of the camera's right-handed look-at
 * view matrix, pltpt_sample.rt.slang:111). vbuffer is accepted and ignored, as in the reference. */
typedef struct plt_render_data {
    uint32_t width, height;
    float* color;
    float* albedo;
    float* normal;
    const void* vbuffer;
} plt_render_data;

/* Renders one frame (sample pass + solve pass) on `stream`, then increments the frame counter.
 * sampleGenerator selects the reference's generators (SampleGeneratorType.slangh:32-36): 1 (the
 * default) = UniformSampleGenerator (xoshiro128**), 0 = TinyUniformSampleGenerator (32-bit LCG
 * seeded by blockCipherTEA); NEEUsePerTileSG = 1 draws the NEE light-type choice from a second
 * generator seeded per tileSize x tileSize tile (pltpt_solve.rt.slang:410-412,
 * LightSampling.slangh:181-185). Options whose code is not built (emissiveSampler other than 0
 * Uniform or 2 Power, debug views, MNEEMaxOccluders above 8, HWSS other than 1 or 4) return
 * PLT_ENOTIMPL. */
int plt_pass_execute(plt_pass* pass, const plt_render_data* data, plt_stream stream);

/* Copies the bounce records of the last execute to host memory `dst` (synchronous). Layout:
 * (maxBounces+1) x 9 planes of width*height 32-bit words, plane (bounce*9 + field), fields:
 * triangle id (0xFFFFFFFF = miss), b1, b2, octahedral wi.x, wi.y, sampled lobe, path flags,
 * path contribution 1/(pdf*thp), last non-delta pdf (the reference's BounceData, Path.slang:113-142).
 * Returns PLT_EINVAL if `bytes` is smaller than the record buffer. */
int plt_pass_copy_bounce_records(plt_pass* pass, void* dst, size_t bytes);

/* Test hook (no counterpart in the reference): the n-th device upload of the next scene bindings
 * (plt_pass_set_scene) fails as an allocation failure would; n < 0 turns it off. */
int plt_debug_fail_upload(int n);

/* Schedule of the sample pass (no counterpart in the reference, whose sample pass is one DXR ray
 * generation shader): PLT_SCHEDULE_MEGAKERNEL (default, the faster on MI355X) runs one kernel in
 * which each lane follows its path; PLT_SCHEDULE_WAVEFRONT runs it as per-bounce kernels over
 * queues of live paths -- a traversal-only kernel, a counting sort of the hits by material type,
 * and a shading kernel over the material-sorted queue. Both produce identical bounce records and
 * outputs. */
enum { PLT_SCHEDULE_MEGAKERNEL = 0, PLT_SCHEDULE_WAVEFRONT = 1 };
int plt_pass_set_schedule(plt_pass* pass, int schedule);

/* Schedule of the solve pass (no counterpart in the reference, whose solve pass is one DXR ray
 * generation shader): PLT_SOLVE_MEGAKERNEL runs one kernel in which each lane replays its path,
 * samples lights and carries every sourced beam to the camera; PLT_SOLVE_SPLIT runs a
 * light-sampling kernel (light samples, shadow rays, the random part of manifold walks; it appends
 * one light event per beam), a manifold-walk kernel (the walks' Newton iterations and visibility,
 * from a queue), and an unwind kernel (each event's beam through the stored vertices, then per-path
 * sums in the megakernel's order), then the megakernel for the paths whose events did not fit the
 * event area (solve_events_per_path events and solve_hits_per_path manifold-walk vertices per path,
 * pooled over regions of 64 paths; defaults 8 and 4). Both write identical colours. */
enum { PLT_SOLVE_MEGAKERNEL = 0, PLT_SOLVE_SPLIT = 1 };

/* Launch configuration of the pass (no counterpart in the reference, whose two DXR dispatches per
 * tile are fixed): the sample-pass schedule, the megakernel's block size (64, 128 or 256 threads),
 * the wavefront schedule's persistent grid (64-thread blocks per CU, 1..64), graph replay, and the
 * solve-pass schedule with its event-area sizes (1..64 events and 0..64 walk vertices per path).
 * With `graph` set, each execute captures the sample pass (with its per-bounce wavefront launches)
 * and the solve pass into HIP graphs on a stream of the pass's own, updates the instantiated graphs
 * with the new frame's arguments and launches them on the caller's stream. Every configuration
 * renders identical outputs. */
typedef struct plt_launch_config {
    uint32_t schedule;              /* PLT_SCHEDULE_* */
    uint32_t sample_block;          /* 64, 128 or 256 */
    uint32_t wf_blocks_per_cu;      /* 1..64 */
    uint32_t graph;                 /* 0 or 1 */
    uint32_t solve_schedule;        /* PLT_SOLVE_* */
    uint32_t solve_events_per_path; /* 1..64 */
    uint32_t solve_hits_per_path;   /* 0..64 */
} plt_launch_config;
int plt_pass_set_launch_config(plt_pass* pass, const plt_launch_config* config);
int plt_pass_get_launch_config(const plt_pass* pass, plt_launch_config* config);
int plt_pass_set_graph(plt_pass* pass, int enable);
/* Autotuner: renders `frames` frames (from the current frame counter, outputs in `data`) with the
 * current configuration and with each candidate -- the sample megakernel at 64, 128 and 256 threads
 * per block, the wavefront sample schedule at 4, 8, 16 and 32 blocks per CU, compared by the sample
 * pass's own time; then, with the fastest sample pass, the solve megakernel and the split solve
 * pass at 4, 8 and 16 events per path; then the fastest of these replayed from graphs (taken only
 * when more than 1 % faster), compared by the frame time -- keeps the fastest (never one slower
 * than the starting configuration as
 * measured), and restores the frame counter and the jitter pattern. Writes the chosen configuration
 * and the mean frame time of the starting and the chosen one (ms). */
int plt_pass_autotune(plt_pass* pass, const plt_render_data* data, uint32_t frames, plt_stream stream,
                      plt_launch_config* best, double* default_ms, double* best_ms);

/* Kernel timing, the counterpart of the reference's profiler scopes around its sample and solve
 * dispatches (FALCOR_PROFILE in Falcor passes; PLTPT.cpp:487-495 issues the two dispatches):
 * while enabled, every execute records HIP events on its stream around the sample and solve
 * kernels. plt_pass_get_kernel_times waits for them and returns the summed milliseconds of each
 * kernel and the number of executes since the previous call (or since enabling), then restarts
 * the sums. Enabling or disabling restarts them too. */
int plt_pass_set_profiling(plt_pass* pass, int enable);
int plt_pass_get_kernel_times(plt_pass* pass, double* sample_ms, double* solve_ms, uint32_t* executes);
/* The same sums with the solve pass resolved into its kernels: the split pass's light-sampling
 * kernel, its manifold-walk kernel, the rest of its unwind (the unwind of the non-walk events runs
 * on a second stream beside the walk kernel; this is the walks' unwind, the join and the per-path
 * sums), and the solve megakernel (the whole solve pass of the megakernel schedule, the split
 * pass's fallback launch, and the whole split pass when it is replayed from a graph, whose launches
 * are not timed one by one). */
int plt_pass_get_kernel_times_split(plt_pass* pass, double* sample_ms, double* light_ms, double* walk_ms, double* unwind_ms,
                                    double* megakernel_ms, uint32_t* executes);
/* Counts of the last execute's split solve pass, read back from the device (waits for the device):
 * light events, manifold-walk vertices stored with them, the sum over events of (vertex index + 1)
 * -- the stored vertices the unwind kernel replays for them -- the paths left to the megakernel
 * because their region's event area was full, the manifold walks the walk kernel ran, those of
 * them that failed (their events add nothing) and the rays the walk kernel cast (line-search steps
 * and visibility). PLT_ESTATE if the last execute used the megakernel. */
int plt_pass_get_solve_stats(plt_pass* pass, uint64_t* events, uint64_t* walk_vertices, uint64_t* event_vertices,
                             uint64_t* overflowed_paths, uint64_t* walks, uint64_t* failed_walks, uint64_t* walk_rays);

/* ---------------------------------------------------------------------------------------
 * Learned BSDF surrogate. No counterpart in the reference: a small MLP (22 inputs, three hidden
 * layers of 128 ReLU units, 16 outputs = 38,032 parameters) that stands in for the exact Mueller
 * matrix of the costly first-order scattering (SPM) lobes -- the coated opaque dielectric, the
 * coated conductor and the grated conductor's base lobe -- trained on exact samples of the
 * material layer above. (The multilayer stack has no scattering lobe, only its delta reflection,
 * whose Re(G)^2 oscillates too fast in wavelength and angle for a network this size; it stays
 * exact.) Forward, backward and Adam run on the matrix cores (BF16 operands, FP32 accumulation,
 * FP32 master weights).
 *
 * Inputs X [n x 22] float32, one sample per row (produced by plt_surrogate_generate): wi (3),
 * wo (3) in the local frame (N = +z), (lambda_nm - 550) / 170, (q_um - 0.325) / 0.275,
 * (sqrt(T_um2) - 0.46) / 0.41 of the K-correlation surface, (coherence area_um2 - 50.5) / 49.5, a
 * one-hot family (coated opaque dielectric, coated conductor, grated conductor), the lobe's derived
 * inputs (ln(k l |wi_t + wo_t| + 1e-3) - 1) / 3, (ln(k q) - 1) / 1.2, (ln(k l) - 2) / 1.5,
 * (ln(|wi_t + wo_t| + 1e-4) + 1.5) / 2.5 with k = 2 pi / lambda_um, and the material: coated --
 * 2 albedo(lambda) - 1 (0 for the conductor), coat IoR and thickness mapped from [1.3, 1.8] and
 * [0.2, 5] um to [-1, 1], 0, 0; grated -- pitch from [0.5, 2] um, the grating direction (cos,
 * sin), height from [0.05, 0.5] um, 0.
 * Outputs Y [n x 16]: the network's raw (standardised) targets; plt_surrogate_eval decodes them
 * into Mueller matrices with the statistics given to plt_surrogate_set_stats.
 * ------------------------------------------------------------------------------------- */
typedef struct plt_surrogate plt_surrogate;
#define PLT_SURROGATE_INPUTS 22
#define PLT_SURROGATE_OUTPUTS 16

/* max_batch: samples per internal batch (a multiple of 32); seed: weight initialisation. */
int plt_surrogate_create(plt_context* ctx, uint32_t max_batch, uint32_t seed, plt_surrogate** out);
void plt_surrogate_destroy(plt_surrogate* s);
/* Parameter count (38,032) and host copies in the order W1 [128x22], b1, W2 [128x128], b2, W3, b3,
 * W4 [16x128], b4 (row-major, out x in). Setting parameters resets the Adam state. */
uint32_t plt_surrogate_param_count(void);
int plt_surrogate_params(plt_surrogate* s, float* host, int set);
/* Y = MLP(X) for n samples (device pointers; any n). */
int plt_surrogate_forward(plt_surrogate* s, uint32_t n, const float* X, float* Y, plt_stream stream);
/* Mean-squared-error loss over the 16 outputs against targets T [n x 16] and its gradient with
 * respect to every parameter (host copy in the parameter order when grads_host is not NULL);
 * n a multiple of 32, at most max_batch. */
int plt_surrogate_gradients(plt_surrogate* s, uint32_t n, const float* X, const float* T, float* grads_host, float* loss,
                            plt_stream stream);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the gradient of a batch. */
int plt_surrogate_train_step(plt_surrogate* s, uint32_t n, const float* X, const float* T, float lr, float* loss,
                             plt_stream stream);
/* Target statistics (targets are y0 = ln M00, yk = asinh(Mk / M00) for k = 1..15, standardised with
 * mean/std) and the training box of the inputs: a sample with any input outside [x_lo, x_hi]
 * is out of distribution and is evaluated exactly. */
int plt_surrogate_set_stats(plt_surrogate* s, const float* y_mean, const float* y_std, const float* x_lo, const float* x_hi);
/* The training data pipeline on the device. From n exact samples (inputs X [n x 22], Mueller
 * matrices M [n x 16] as plt_surrogate_generate writes them): the training targets T [n x 16]
 * (ln M00 and asinh(Mk / M00), standardised by the valid rows' mean and standard deviation, the
 * deviations divided by the output weights 4, 2, 2, 2, 1, ...), the valid rows' indices in row
 * order in idx (the first *n_valid entries; a row is valid when M is finite and M00 > 1e-12), and
 * the statistics and the valid inputs' box set as plt_surrogate_set_stats would. Waits for the
 * stream. */
int plt_surrogate_prepare(plt_surrogate* s, uint32_t n, const float* X, const float* M, float* T, uint32_t* idx, uint32_t* n_valid,
                          plt_stream stream);
/* The order plt_surrogate_train draws epoch `epoch` in for n samples and `seed`: out[pos] is the
 * sample taken at position pos (a permutation of [0, n)). */
int plt_surrogate_shuffle(uint32_t n, uint32_t seed, uint32_t epoch, uint32_t* out, plt_stream stream);
/* The statistics set (host copies: target mean [16] and scale [16], input box lo [22], hi [22]). */
int plt_surrogate_get_stats(const plt_surrogate* s, float* y_mean, float* y_std, float* x_lo, float* x_hi);
/* Trains for `epochs` passes over the n_valid samples idx[0..n_valid) of (X, T) (from
 * plt_surrogate_prepare) in batches of `batch` (a multiple of 32, at most max_batch; the last
 * n_valid % batch of each epoch's order are left out): each epoch draws a fresh shuffle (a keyed
 * bijection of [0, n_valid) computed on the device, seeded by `seed` and the epoch), Adam with the
 * cosine-decayed learning rate lr (1 + cos(pi step / steps)) / 2 (at least 1e-6). With `graph` set,
 * one training step is captured as a HIP graph and replayed (every step-dependent value lives in
 * device state); without, the same launches run eagerly (identical results). Writes each epoch's
 * mean batch loss to epoch_loss [epochs]; waits for the stream. */
int plt_surrogate_train(plt_surrogate* s, uint32_t n_valid, const uint32_t* idx, const float* X, const float* T, uint32_t batch,
                        uint32_t epochs, float lr, uint32_t seed, int graph, float* epoch_loss, plt_stream stream);

/* Exact samples for training and judging (the material layer of the path tracer, FP32 with FP64
 * multilayer steps): per sample i a material (written to materials[i], binary16 fields as
 * plt_material), its geometry (PLT_SG_* planes of n floats), the inputs X[i] and the exact
 * Mueller matrix mueller[i] (row-major 4x4, hero wavelength: S_out = M S_in in the beam frames of
 * plt_material_eval_batch). Sample i draws from the sample generator of pixel (i & 0xFFFF, i >> 16)
 * at frame `seed`. The conductors use spectral profiles 0 (n) and 1 (k) of the context. */
enum {
    PLT_SG_WI_X = 0, PLT_SG_WI_Y, PLT_SG_WI_Z, PLT_SG_WO_X, PLT_SG_WO_Y, PLT_SG_WO_Z, PLT_SG_LAMBDA, PLT_SG_LOBE,
    PLT_SG_AREA, PLT_SG_PLANES
};
int plt_surrogate_generate(plt_context* ctx, uint32_t n, uint32_t seed, plt_material* materials, float* geometry, float* X,
                           float* mueller, plt_stream stream);
/* Outgoing Stokes vectors stokes_out [n x 4] for incident stokes_in [n x 4]: mode 0 = the
 * surrogate, with the exact evaluation for out-of-distribution samples (their count is added to
 * *fallbacks when not NULL, synchronously); mode 1 = exact evaluation of every sample (the
 * surrogate's judge). */
int plt_surrogate_eval(plt_surrogate* s, uint32_t n, const plt_material* materials, const float* geometry, const float* X,
                       const float* stokes_in, float* stokes_out, int mode, uint32_t* fallbacks, plt_stream stream);

/* The surrogate in the path tracer (C4: the render with the learned lobes). With a trained
 * surrogate set (its training statistics set), every execute of the pass runs the split solve pass
 * with the surrogate: the unwind of every light event first appends the network's inputs for each
 * unwind step whose lobe the surrogate stands for -- the scattering lobes of the coated opaque
 * dielectric and the coated conductor and the grated conductor's base lobe, one row per HWSS
 * wavelength, with the coherence area the beam presents -- then one batched inference runs on the
 * matrix cores, then the unwind runs again with the network's Mueller matrices at those steps.
 * Steps whose inputs fall outside the training box are evaluated exactly, as are paths the split
 * pass leaves to its megakernel fallback. NULL restores the exact render. The surrogate must
 * outlive its use by the pass. Mode PLT_SURROGATE_TEACHER runs the same passes with the exact
 * lobe's matrix, taken in the network's frames and output encoding, in place of the network's
 * outputs (no inference): its images differ from the exact render only by rounding, which checks
 * the query / apply plumbing apart from the network's error. */
enum { PLT_SURROGATE_NETWORK = 0, PLT_SURROGATE_TEACHER = 1 };
int plt_pass_set_surrogate(plt_pass* pass, plt_surrogate* s, int mode);
/* Counts of the last execute with the surrogate (waits for the device): query rows appended (one
 * per HWSS wavelength of every eligible step; rows beyond the query buffer, four per path, are not
 * evaluated), and eligible steps evaluated exactly (outside the training box or beyond the buffer). */
int plt_pass_get_surrogate_stats(plt_pass* pass, uint64_t* rows, uint64_t* exact_steps);
/* Copies up to max_rows of the last execute's query rows (22 inputs each, in append order; rows of
 * HWSS components a beam does not carry are not written) to host memory, for analysing which
 * inputs leave the training box; *rows receives the number copied. */
int plt_pass_copy_surrogate_queries(plt_pass* pass, float* dst, uint64_t max_rows, uint64_t* rows);
/* The matrix-core GEMM the MLP is built from: C [M x N] (FP32) = A [M x K] * Bt [N x K]^T with
 * bf16 row-major operands, K a multiple of 32 (for tests of the kernel). */
int plt_mlp_gemm_nt(uint32_t M, uint32_t N, uint32_t K, const void* A, const void* Bt, float* C, plt_stream stream);

/* ---------------------------------------------------------------------------------------
 * Temporal accumulation: the AccumulatePass that follows PLTPT in the reference's render graph
 * (Source/RenderPasses/AccumulatePass/AccumulatePass.{h,cpp}, Accumulate.cs.slang;
 * scripts/PLTPathTracer.py:19-24 uses Double for PLTPT.color and Single for albedo and normal).
 * Replaces the plugin class's create / execute / reset / setEnabled / setMaxAccumulatedFrames
 * (AccumulatePass.cpp:31-47,88-116,144-301,370-382; AccumulatePass.h:66-67). Options are its dictionary keys and
 * defaults: enabled 1, autoReset 1, precisionMode 1 (Single), maxFrameCount 0 (unlimited),
 * overflowMode 0 (Stop), outputSize 0 (Default) and outputFormat 0 (RGBA32Float) -- other
 * output sizes or formats return PLT_ENOTIMPL. Setting precisionMode, maxFrameCount or
 * overflowMode resets the frame count (as the reference's UI and setMaxAccumulatedFrames do);
 * setting `enabled` behaves as setEnabled (a change resets).
 * ------------------------------------------------------------------------------------- */
typedef struct plt_accumulate plt_accumulate;
enum { PLT_ACCUM_DOUBLE = 0, PLT_ACCUM_SINGLE = 1, PLT_ACCUM_SINGLE_COMPENSATED = 2 };
enum { PLT_ACCUM_OVERFLOW_STOP = 0, PLT_ACCUM_OVERFLOW_RESET = 1, PLT_ACCUM_OVERFLOW_EMA = 2 };
/* Input texel formats: RGBA32Float, RGBA32Uint, RGBA32Sint (the shader's _INPUT_FORMAT,
 * Accumulate.cs.slang:48-59, AccumulatePass.cpp:245-260). */
enum { PLT_ACCUM_INPUT_FLOAT = 0, PLT_ACCUM_INPUT_UINT = 1, PLT_ACCUM_INPUT_SINT = 2 };

int plt_accumulate_create(plt_context* ctx, plt_accumulate** out);
int plt_accumulate_destroy(plt_accumulate* pass);
int plt_accumulate_set_option(plt_accumulate* pass, const char* key, double value);
int plt_accumulate_get_option(const plt_accumulate* pass, const char* key, double* value);
const char* plt_accumulate_option_name(uint32_t i);
int plt_accumulate_set_enabled(plt_accumulate* pass, int enabled);
int plt_accumulate_reset(plt_accumulate* pass);
/* Frames accumulated so far (mFrameCount). */
int plt_accumulate_get_frame_count(const plt_accumulate* pass, uint32_t* count);
/* One execute on `stream`: src is width*height texels of src_format (16 B each, 16-byte aligned),
 * dst width*height float4 (the running average). refresh_flags != 0 reports what the reference
 * reads from the render data's refresh flags and scene/camera updates (a reset when autoReset is
 * on; AccumulatePass.cpp:146-174). With maxFrameCount reached and overflowMode Stop, dst is not
 * written. A change of width or height resets. */
int plt_accumulate_execute(plt_accumulate* pass, const void* src, int src_format, uint32_t width, uint32_t height,
                           float* dst, uint32_t refresh_flags, plt_stream stream);

/* ---------------------------------------------------------------------------------------
 * Tone mapping: the ToneMapper at the end of the reference's render graph
 * (Source/RenderPasses/ToneMapper/ToneMapper.{h,cpp}, Luminance.ps.slang, ToneMapping.ps.slang;
 * scripts/PLTPathTracer.py:26-27 sets autoExposure 1, exposureCompensation 0). Options are the
 * keys of ToneMapper::getScriptingDictionary and its property exposureValue
 * (ToneMapper.cpp:57-76,99-127,181-200), applied through the reference's setters with their clamps and
 * coupled updates (ToneMapper.cpp:406-515); outputSize other than 0 (Default) or outputFormat
 * other than 0 (RGBA32Float) return PLT_ENOTIMPL.
 * ------------------------------------------------------------------------------------- */
typedef struct plt_tonemapper plt_tonemapper;
/* ToneMapperOperator (ToneMapperParams.slang:33-41). */
enum {
    PLT_TONEMAP_LINEAR = 0, PLT_TONEMAP_REINHARD = 1, PLT_TONEMAP_REINHARD_MODIFIED = 2, PLT_TONEMAP_HEJI_HABLE_ALU = 3,
    PLT_TONEMAP_HABLE_UC2 = 4, PLT_TONEMAP_ACES = 5
};

int plt_tonemapper_create(plt_context* ctx, plt_tonemapper** out);
int plt_tonemapper_destroy(plt_tonemapper* pass);
int plt_tonemapper_set_option(plt_tonemapper* pass, const char* key, double value);
int plt_tonemapper_get_option(const plt_tonemapper* pass, const char* key, double* value);
const char* plt_tonemapper_option_name(uint32_t i);
/* setScene with scene metadata (ToneMapper.cpp:395-405): each value that is not NaN is applied
 * through its setter when useSceneMetadata is on. */
int plt_tonemapper_set_scene_metadata(plt_tonemapper* pass, float film_iso, float f_number, float shutter_speed);
/* The 3x3 colour transform (row-major) the next execute applies: white balance x
 * 2^exposureCompensation x manual exposure (ToneMapper.cpp:527-547). */
int plt_tonemapper_get_color_transform(const plt_tonemapper* pass, float* out9);
/* Tone-maps src (width*height float4, 16-byte aligned) into dst (same size) on `stream`. With
 * autoExposure on, mean_log_luminance (a device float, may be NULL) receives the mean log2
 * luminance the exposure used. */
int plt_tonemapper_execute(plt_tonemapper* pass, const float* src, uint32_t width, uint32_t height, float* dst,
                           float* mean_log_luminance, plt_stream stream);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): plt_tonemapper_execute
 * with the same kernels and results, whose auto-exposure reduction works in a fixed 512 KiB
 * workspace compiled into the library (one per GPU, held until the process ends) instead of a
 * buffer it allocates. Calls no allocating routine. Images whose reduction needs more than 65,536
 * tiles (bit_floor(width) / 32 x bit_floor(height) / 32, i.e. either bit_floor above 8192) return
 * PLT_EINVAL. Every call on one GPU shares that workspace: calls that overlap on one GPU (other
 * streams, other threads) race and must be serialised by the caller. */
int plt_tonemapper_execute_noalloc(plt_tonemapper* pass, const float* src, uint32_t width, uint32_t height, float* dst,
                                   float* mean_log_luminance, plt_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* PLT_ABI_H */
