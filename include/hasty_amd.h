/* C ABI of the MI355X path tracer.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes
 * (0 = OK, nonzero = error; hasty_last_error() returns the message of the
 * calling thread's last failure).  No C++ exception crosses this boundary.
 *
 * Films are row-major RGB float32 arrays of width*height*3 floats holding
 * SUMS over samples, exactly like the reference's accumulation buffers
 * (core/include/Hasty/Image.h:302-313); estimate = sum / number of samples
 * (apps/ViewerUtils.cpp:95-121).
 *
 * Calls on one scene must be serialised by the caller (the reference's
 * threads only share the scene read-only and hand out sample indices under a
 * mutex, core/include/Hasty/RenderJob.h:53-60).  All device memory belongs to
 * the scene; every call is synchronous on the scene's own HIP stream.
 */
#ifndef HASTY_AMD_H
#define HASTY_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct HastyScene HastyScene;

/* RenderSettings (core/include/Hasty/RenderJob.h:15-30; JSON keys
 * core/src/RenderJob.cpp:10-20). */
typedef struct HastyRenderSettings {
  int32_t roulette_start_depth;
  float roulette_q;
  uint32_t width;
  uint32_t height;
  uint32_t num_threads;
  int32_t max_depth;
  int32_t num_samples;
  int32_t use_mis;
} HastyRenderSettings;

typedef struct HastySceneInfo {
  HastyRenderSettings settings;
  uint32_t num_triangles;
  uint32_t num_shapes;
  uint32_t num_materials;
  uint32_t num_light_triangles;
  int32_t has_environment;
  uint32_t bvh_nodes;      /* nodes of the BVH in use (32 x bvh_width bytes each) */
  int32_t bvh_depth;       /* its inner levels */
  uint64_t device_bytes;   /* scene + BVH bytes resident in HBM */
  int32_t bvh_width;       /* children per node: 4 or 8 */
  float camera_exposure;   /* camera "exposure" (Camera.cpp:22): display scale 2^exposure */
} HastySceneInfo;

/* Replaces RenderJob::loadJSON (core/src/RenderJob.cpp:21-28) together with
 * Scene::Scene and its rtcCommitScene (core/src/Scene.cpp:12-195): parses
 * renderJob.json -> scene JSON -> .obj/.mtl (+ environment map and image
 * textures in PFM, PPM / PGM, PNG or OpenEXR), builds
 * the 4-wide BVH on the host and uploads everything to the current device.
 * Errors use the reference's messages (malformed input, a face without a
 * material, an emissive face with a non-black Kd/Ks, ...). */
int hasty_scene_load(const char* render_job_path, HastyScene** out_scene);

/* ---- scene from data the caller already holds ----
 * What Scene::Scene (core/src/Scene.cpp:12-195) has after tinyobj and OIIO
 * loaded the files, handed over as flat arrays instead of re-parsed, so an
 * app that builds its own Scene keeps exactly its inputs. */

/* PrincipledBRDF parameters after the MTL mapping of Scene.cpp:65-118:
 * albedo = Kd, roughness = 1 - sqrt(Ns / 1000), metallic = Ka[0],
 * specular = Ks[0], ior = Ni, anisotropy = aniso, transmission = 1 - d,
 * emission = Ke (radiance Ke / pi).  tinted_emitter: the MTL's Kd or Ks is
 * non-black (the reference refuses a face using an emissive material that
 * is, Scene.cpp:166-173). */
typedef struct HastyMaterialDesc {
  float albedo[3];
  float roughness;
  float metallic;
  float specular;
  float ior;
  float anisotropy;
  float transmission;
  int32_t vary_ior;  /* world.vary_glass_ior && transmission > 0 */
  float emission[3];
  int32_t tinted_emitter;
} HastyMaterialDesc;

/* Camera (core/include/Hasty/Camera.h:14-60) with its frame as setLookAt
 * leaves it (Camera.cpp:31-39). */
typedef struct HastyCameraDesc {
  float position[3];
  float forward[3];
  float right[3];
  float up[3];
  float fov_degree;  /* 45 in the reference */
  float aperture;
  float focal_distance;
  float blade_rotation;
  int32_t num_blades;
  float exposure;
} HastyCameraDesc;

/* An image texture as TextureTraits::loadTexture leaves it
 * (core/src/Texture.cpp:21-49): width*height texels of 1 float (roughness,
 * metallic) or 3 floats (albedo, already decoded from sRGB; normal map),
 * linear, row 0 at the top. */
typedef struct HastyTextureDesc {
  uint32_t width;
  uint32_t height;
  uint32_t channels; /* 1 or 3 */
  const float* pixels;
} HastyTextureDesc;

/* Triangles are listed shape by shape, faces in order: geom_ids[i] (the
 * tinyobj shape) never decreases and prim_ids[i] counts the faces of its
 * shape from 0, so (geom_id, prim_id) are the ids Embree would report.
 * Index arrays hold 3 entries per triangle into the attribute arrays
 * (tinyobj::attrib_t), -1 = absent; normal_indices / texcoord_indices may be
 * NULL.  env_rgb is width*height RGB float32 with row 0 at the top (OIIO's
 * order); without an environment the background colour is used. */
typedef struct HastySceneDesc {
  HastyRenderSettings settings;
  HastyCameraDesc camera;
  uint32_t num_vertices;
  const float* vertices;   /* 3 per vertex */
  uint32_t num_normals;
  const float* normals;    /* 3 per normal */
  uint32_t num_texcoords;
  const float* texcoords;  /* 2 per texcoord */
  uint32_t num_triangles;
  const int32_t* vertex_indices;
  const int32_t* normal_indices;
  const int32_t* texcoord_indices;
  const uint32_t* geom_ids;
  const uint32_t* prim_ids;
  const int32_t* material_ids;
  const uint8_t* smooth;   /* smoothing group != 0 (NULL: none) */
  uint32_t num_materials;
  const HastyMaterialDesc* materials;
  int32_t has_environment;
  uint32_t env_width;
  uint32_t env_height;
  const float* env_rgb;
  float background[3];
  /* Image textures (may be 0 / NULL).  material_textures holds 4 texture
   * indices per material -- albedo (map_Kd), roughness (map_Ns), metallic
   * (map_Pm), normal map (bump), -1 = the material's constant
   * (core/src/Scene.cpp:73-114) -- or is NULL for an untextured scene. */
  uint32_t num_textures;
  const HastyTextureDesc* textures;
  const int32_t* material_textures;
} HastySceneDesc;

/* Builds the BVH on the host and uploads the scene to the current device,
 * like hasty_scene_load after parsing.  The descriptor's arrays are copied;
 * the caller keeps ownership. */
int hasty_scene_create(const HastySceneDesc* desc, HastyScene** out_scene);

/* Frees the scene and its device memory (Scene::~Scene, Scene.cpp:196-200). */
void hasty_scene_destroy(HastyScene* scene);

int hasty_scene_info(const HastyScene* scene, HastySceneInfo* out_info);

/* Overrides the job's settings (e.g. a smaller film for a check render).
 * width/height/max_depth/roulette/use_mis take effect on the next render. */
int hasty_scene_set_settings(HastyScene* scene, const HastyRenderSettings* settings);

/* Path tracing: one call renders samples [first_sample, first_sample +
 * num_samples) of every pixel and ADDS them, in sample order, into the three
 * films -- what renderThread's pass loop does (core/src/PathTracing.cpp:
 * 98-160, declared core/include/Hasty/PathTracing.h:63).  Sample k uses the
 * Halton (2,3) point of index k+1 as its shared sub-pixel offset
 * (RenderJob.h:53-60).  Random numbers come from one counter-based stream per
 * (seed, pixel, sample).  Samples whose radiance is NaN are skipped in all
 * three films (PathTracing.cpp:132-135).  Host pointers. */
int hasty_render_pt(HastyScene* scene, uint32_t seed, uint32_t first_sample, uint32_t num_samples, float* color_sum,
                    float* normal_sum, float* albedo_sum);

/* Same, with the films already resident in device memory (HBM) of the
 * scene's device. */
int hasty_render_pt_device(HastyScene* scene, uint32_t seed, uint32_t first_sample, uint32_t num_samples,
                           float* d_color_sum, float* d_normal_sum, float* d_albedo_sum);

/* Batched Scene::rayHit (core/src/Scene.cpp:281-303): closest hit of n rays
 * (origin xyz, direction xyz; tnear 1e-6, tfar +inf).  Outputs per ray:
 * geom_prim[2i] = geomID (OBJ shape index) and geom_prim[2i+1] = primID, or
 * 0xFFFFFFFF twice for a miss; tri[i] = global triangle index (-1 miss);
 * tuv[3i..] = t, u, v with barycentrics (1-u-v, u, v).  Host pointers. */
int hasty_trace_closest(HastyScene* scene, const float* rays, uint32_t n, uint32_t* geom_prim, int32_t* tri,
                        float* tuv);

/* One ray of a batched query: origin and direction (tnear 1e-6, tfar +inf,
 * the values Scene::rayHit sets, core/src/Scene.cpp:281-291). */
typedef struct HastyRay {
  float org[3];
  float dir[3];
} HastyRay;

/* The RTCHit fields Scene::rayHit reads (core/src/Scene.cpp:293-302,
 * core/include/Hasty/Scene.h:20-46): geom_id (OBJ shape index) and prim_id
 * (face within the shape), both 0xFFFFFFFF for a miss (RTC_INVALID_GEOMETRY_ID);
 * the hit distance t (+inf for a miss); barycentrics (1-u-v, u, v); and the
 * geometric normal normalize((v1 - v0) x (v2 - v0)), zero for a miss. */
typedef struct HastyHit {
  uint32_t geom_id;
  uint32_t prim_id;
  float t;
  float u;
  float v;
  float ng[3];
} HastyHit;

/* Batched Scene::rayHit (core/src/Scene.cpp:281-303, declared
 * core/include/Hasty/Scene.h:68) with the full hit record, host pointers. */
int hasty_trace_hits(HastyScene* scene, const HastyRay* rays, uint32_t n, HastyHit* hits);

/* Same with rays and hits in device memory of the scene's device.  The
 * traversal's stack overflow buffer is the scene's (allocated on first use).
 * Synchronous on the scene's stream. */
int hasty_trace_hits_device(HastyScene* scene, const HastyRay* d_rays, uint32_t n, HastyHit* d_hits);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the query of hasty_trace_hits_device with its traversal
 * stack overflow in a static 9 MiB device array of the library instead of a
 * scene buffer, on a fixed grid of 64 blocks.  Allocates nothing (no
 * hipMalloc / hipFree); n at most 2^31 - 1; hits 16-byte aligned.  Every call
 * on one GPU shares the array: calls that overlap on one GPU race.
 * Synchronous on the scene's stream. */
int hasty_trace_hits_device_noalloc(HastyScene* scene, const HastyRay* d_rays, uint32_t n, HastyHit* d_hits);

/* Per-kernel timing with HIP events on the scene's stream (flags bit 0;
 * with bit 2 only the closest-hit extension launches are timed, which keeps
 * the events' own cost out of a timed render) and device counters (flags
 * bit 1, see hasty_counters). */
typedef struct HastyKernelStat {
  char name[32];
  uint64_t launches;
  double total_ms;
} HastyKernelStat;
int hasty_set_profiling(HastyScene* scene, int flags);
/* Writes up to max_stats entries, returns the number written (negative on error). */
int hasty_kernel_stats(HastyScene* scene, HastyKernelStat* stats, int max_stats);
int hasty_reset_stats(HastyScene* scene);

/* Device counters accumulated since the last reset (profiling flag bit 1):
 * out[0] extension rays, out[1] shadow rays, out[2] BVH node visits,
 * out[3] triangle tests, out[4] paths whose media stack overflowed. */
int hasty_counters(HastyScene* scene, uint64_t* out5);
/* The first n (at most 7) of: the five counters above, then the shadow
 * rays' share of the node visits (out[5]) and triangle tests (out[6]). */
int hasty_counters_ex(HastyScene* scene, uint64_t* out, uint32_t n);

/* Launch-shape knobs: path slots in flight (paths are regenerated into
 * freed slots; default 2^23, at least 2^16 unless the film is smaller) and
 * the grid sizes of the traversal and shading kernels.  0 keeps the current
 * value.  Images do not depend on these. */
int hasty_set_launch(HastyScene* scene, uint32_t path_slots, uint32_t trace_blocks, uint32_t shade_blocks);

/* Launch configuration of a scene's path pool (also used for batches whose
 * first job is this scene).  chunk = bounces between host checks (even);
 * use_graph = replay each chunk as one captured hipGraph; bvh_width =
 * children per BVH node, 4 or 8 (the tree of that width is built and
 * uploaded the first time it is asked for; images do not depend on it).
 * Fields that are 0 keep their value (except use_graph). */
typedef struct HastyLaunch {
  uint32_t path_slots;
  uint32_t trace_blocks;
  uint32_t shade_blocks;
  uint32_t chunk;
  int32_t use_graph;
  uint32_t bvh_width;
} HastyLaunch;
int hasty_get_launch(const HastyScene* scene, HastyLaunch* out);
int hasty_set_launch_ex(HastyScene* scene, const HastyLaunch* launch);

/* ---- batching: many render jobs through one pool of path slots ----
 * No reference counterpart (the reference renders one job per process,
 * apps/PathTracingViewer.cpp).  Each job's films receive exactly what
 * hasty_render_pt_device(scene, seed, first, n, ...) adds for it alone; the
 * paths of all jobs are in flight together, so every wavefront stage is one
 * launch for the whole batch.  Uses the path pool, stream and launch settings
 * of jobs[0].scene; all scenes must be on one device. */
typedef struct HastyBatchJob {
  HastyScene* scene;
  uint32_t seed;
  uint32_t first_sample;
  uint32_t num_samples;
  float* d_color;
  float* d_normal;
  float* d_albedo;
} HastyBatchJob;
int hasty_render_batch_device(const HastyBatchJob* jobs, uint32_t n_jobs);

/* Autotuner (no reference counterpart): times short renders (trial_samples
 * samples of every job into scratch films) over path slots, grid sizes,
 * chunk length, graph replay and BVH node width by coordinate descent,
 * applies the fastest configuration to jobs[0].scene (the width to every
 * job's scene) and returns it in *chosen (may be NULL). */
int hasty_autotune_batch(const HastyBatchJob* jobs, uint32_t n_jobs, uint32_t trial_samples, HastyLaunch* chosen);
int hasty_autotune(HastyScene* scene, uint32_t seed, uint32_t trial_samples, HastyLaunch* chosen);

/* ---- learned path guide (no counterpart in the reference) ----
 * A small MLP (BF16 weights on MFMA, FP32 accumulation) maps (position
 * normalised to the scene box, normal, outgoing direction) to a mixture of 8
 * von Mises-Fisher lobes (raw outputs per lobe: mean direction 3, log
 * concentration 1, logit 1; then a selection logit; 48 floats with
 * padding).  Trained with Adam on
 * the loss -F log q(w_i).  The guide lives on its scene's device and stream. */
typedef struct HastyGuide HastyGuide;
uint32_t hasty_guide_num_params(void);
int hasty_guide_create(HastyScene* scene, uint32_t seed, HastyGuide** out);
void hasty_guide_destroy(HastyGuide* guide);
int hasty_guide_set_box(HastyGuide* guide, const float* lo3, const float* hi3);
int hasty_guide_set_params(HastyGuide* guide, const float* params);  /* resets Adam */
int hasty_guide_get_params(HastyGuide* guide, float* params);
/* queries: n x 9 floats (position, normal, wo) -> raw: n x 48 floats */
int hasty_guide_infer(HastyGuide* guide, const float* queries, uint32_t n, float* raw);
/* samples: n x 13 floats (position, normal, wo, wi, weight F).  gradient of
 * the mean loss (params-sized) without changing the guide */
int hasty_guide_gradient(HastyGuide* guide, const float* samples, uint32_t n, float* grad, float* loss);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on the mean loss of the samples */
int hasty_guide_train_step(HastyGuide* guide, const float* samples, uint32_t n, float lr, float* loss);
/* Learned guide/BSDF selection (no counterpart in the reference): raw output
 * 40 is the logit z of the probability c = lo + (hi - lo) sigmoid(z) of
 * drawing a guided vertex's direction from the guide.  With 0 < lo < hi < 1
 * guided renders use c instead of the scene's fixed fraction, and samples
 * that carry the BSDF density pB of wi and the density pS wi was drawn with
 * add the second moment F^2 pS / (c q + (1 - c) pB) of the vertex estimate
 * to the loss, with q held fixed and trained through the last layer's row
 * 40 only (online training records pB and pS).  lo = hi = 0 (the default)
 * switches it off. */
int hasty_guide_set_selection(HastyGuide* guide, float lo, float hi);
/* hasty_guide_gradient with samples of `stride` floats: 13 as there, or 15
 * with the BSDF density of wi and the density wi was drawn with last (0:
 * none) */
int hasty_guide_gradient_ex(HastyGuide* guide, const float* samples, uint32_t n, uint32_t stride, float* grad,
                            float* loss);

/* Guided rendering of a scene: hasty_render_pt[_device] then draws the
 * continuation direction at glossy and diffuse vertices from the guide with
 * probability `fraction` and from the BSDF otherwise, weighting by the
 * mixture pdf (one-sample MIS; NEE's MIS uses the same mixture), so the
 * estimate stays unbiased.  When train_lr > 0 sample indices are rendered
 * in groups of about 2^20 paths, each group followed by a pass of Adam
 * steps (at least one per sample index of the group, more passes if needed)
 * over the radiance samples its paths produced (up to 2^21; weight =
 * luminance of the radiance the continuation brought over the throughput
 * before the sample, normalised to mean 1 and capped), in minibatches of a
 * pseudo-random order; without training all indices go in one call.  guide NULL returns to plain
 * rendering.  The guide must belong to the scene; the scene must outlive its
 * use. */
int hasty_scene_set_guide(HastyScene* scene, HastyGuide* guide, float fraction, float train_lr);
/* Online-training minibatch (default 16384, in [16, 2^21]) and weight cap
 * (default 20).  No reference counterpart. */
int hasty_scene_set_guide_training(HastyScene* scene, uint32_t minibatch, float weight_cap);
/* Photon-traced training samples (no counterpart in the reference): after
 * each training step's camera paths, `photons` light paths leave the surface
 * lights through concentrated lobes (as tracePhoton, PhotonMapping.cpp:
 * 191-265); each landing on a guided diffuse receiver after such a bounce
 * adds a sample (its position, the direction the light came from, weight =
 * flux luminance), the photon samples scaled to `share` x the camera
 * samples' weight sum.  photons = 0 (the default) turns them off. */
int hasty_scene_set_guide_photons(HastyScene* scene, uint32_t photons, float share);
/* Only path vertices of depth below max_depth are guided (default 255: all),
 * and only while every channel of the path throughput is at most
 * max_throughput (default 3e38: no limit); other vertices sample the BSDF
 * alone and skip the network. */
int hasty_scene_set_guide_depth(HastyScene* scene, uint32_t max_depth, float max_throughput);
/* loss and sample count of the last online training step */
int hasty_scene_guide_stats(const HastyScene* scene, float* last_loss, uint32_t* last_samples);

/* ---- progressive photon mapping (caustic photon map + MIS path tracing) ----
 * Replaces renderPhotonThread (core/src/PhotonMapping.cpp:814-933, declared
 * core/include/Hasty/PhotonMapping.h:124) for one PMRenderJob
 * (PhotonMapping.h:33-121).  A job keeps the per-pixel statistics and the
 * radius image on the device between passes; pass k uses the k-th sample
 * offset of the job (RenderJob.h:53-60) and random streams keyed by
 * (seed, photon or pixel index, k).  The films receive the CURRENT estimate
 * of every pixel (the reference sets numSamples = 1 after each pass,
 * PhotonMapping.cpp:918-923), not sums. */
typedef struct HastySppm HastySppm;

typedef struct HastySppmParams {
  float alpha;           /* radius reduction, PMRenderJob::alpha (PhotonMapping.h:65), default 0.4 */
  float footprint_scale; /* initial radius / pixel footprint (PhotonMapping.cpp:840), default 10 */
  double store_fraction; /* stop emitting at store_fraction*W*H stored photons (PhotonMapping.cpp:834), 0.5 */
  double emit_fraction;  /* ... or emit_fraction*W*H emitted photons (PhotonMapping.cpp:835), 10 */
  float initial_radius;  /* > 0: constant initial radius (no counterpart; the reference uses the
                            footprint estimate, PhotonMapping.cpp:838-841), default 0 */
  /* No counterpart (the reference sums FP32, PhotonMapping.cpp:271-296):
   * nonzero = each gather rounds its photon terms to int64 units of
   * 2^-flux_frac_bits and sums them in integers, so the gathered flux does
   * not depend on the order the photons are summed in (shared or per-lane
   * gathers, cell size, a split of the photons over GPUs).  A saturated sum
   * is counted in HastySppmInfo.flux_overflows.  Default 0 (FP32). */
  int32_t fixed_point_flux;
  int32_t flux_frac_bits;            /* 0: 24; at most 40 */
  uint32_t gather_share_candidates;  /* 0: 512; a wave gathers its queries together when one of them has
                                        more candidate photons than this (no effect on the result with
                                        fixed-point flux) */
  float gather_union_factor;         /* > 0: such a wave reads the union of its queries' cells once, each
                                        lane summing its own photons in the reference's order (FP32 results
                                        of the per-lane gather), when the union box holds at most this x the
                                        queries' mean cells; 0 (default): the shared gather, faster on dense
                                        maps (config-4 gathers 87 against 102 ms per pass) */
} HastySppmParams;

typedef struct HastySppmInfo {
  uint64_t emitted;     /* photons emitted in the last pass (PhotonMap::emittedCount) */
  uint64_t stored;      /* photons stored in the last pass (PhotonMap::photonCount) */
  float cell_size;      /* computeCellSize of the last pass (PhotonMapping.cpp:522-576) */
  uint32_t passes_done;
  uint32_t emit_limit;
  uint32_t store_limit;
  uint32_t flux_overflows; /* gathers of the last pass whose fixed-point flux sum saturated */
  /* gathers of the last pass (photonMapFluxEstimate calls, PhotonMapping.cpp:271-296): queries,
   * candidate photons in the cells they visit, photons within the radius */
  uint64_t gather_queries;
  uint64_t gather_candidates;
  uint64_t gather_found;
} HastySppmInfo;

/* params may be NULL (reference defaults).  The scene must outlive the job. */
int hasty_sppm_create(HastyScene* scene, uint32_t seed, const HastySppmParams* params, HastySppm** out);
void hasty_sppm_destroy(HastySppm* job);
/* One pass into device films (3*W*H floats each). */
int hasty_sppm_pass_device(HastySppm* job, uint32_t pass, float* d_color, float* d_normal, float* d_albedo);
/* One pass into host films. */
int hasty_sppm_pass(HastySppm* job, uint32_t pass, float* color, float* normal, float* albedo);
int hasty_sppm_info(const HastySppm* job, HastySppmInfo* info);
/* PMRenderJobPixel planes (PhotonMapping.h:12-31), W*H floats each: radius,
 * N, Ne, flux r/g/b, LSum r/g/b, Nd (int32 bits). */
int hasty_sppm_pixels(HastySppm* job, float* out_10_planes);
/* The radius image (W*H floats, row-major). */
int hasty_sppm_radii(HastySppm* job, float* out);
/* Photons of the last pass in emission order, 10 floats each: position,
 * dot(n_g, w_i), w_i, flux (Photon, PhotonMapping.cpp:11-23). */
int hasty_sppm_photons(HastySppm* job, float* out, uint64_t max_photons);

/* ---- the output step (apps/ViewerUtils.cpp:15-152) on device films ----
 * computeEstimate (ViewerUtils.cpp:95-121): the films, each W*H*3 floats of
 * the scene's size holding sums of num_samples[k] samples, added in order
 * from the first film with samples (the reference replaces its image until
 * the running total is nonzero), divided by 1e-7 + the total, and scaled by
 * 2^exposure as the viewers scale the colour estimate
 * (PathTracingViewer.cpp:120, 160; pass 0 for normal / albedo films).
 * Device pointers on the scene's device; 1 <= n_films <= 64. */
int hasty_estimate_device(HastyScene* scene, const float* const* d_films, const uint32_t* num_samples,
                          uint32_t n_films, float exposure, float* d_out);
/* preparePresentableImage (ViewerUtils.cpp:15-65): each pixel of a W*H*3
 * image as R << 24 | G << 16 | B << 8 | 255 with channel
 * uint(255 * pow(clamp(v, 0, 1), 1/2.2)). */
int hasty_present_device(HastyScene* scene, const float* d_image, uint32_t* d_rgba);

/* Device films with snapshots for progressive display -- the renderThread /
 * DoubleBuffer pattern (PathTracing.cpp:98-160, core/include/Hasty/Image.h:
 * 18-45) with the GPU holding the sums:
 *   hasty_films_render adds samples like hasty_render_pt_device;
 *   hasty_films_snapshot copies the three films on the scene's stream after
 *     the work queued so far and starts their copy to pinned host memory on
 *     a second stream, so the next render overlaps the readback;
 *   hasty_films_read hands the last finished snapshot to host buffers (the
 *     DoubleBuffer's write buffer) with its sample count: 0 = copied, 4 =
 *     nothing new, or (wait == 0) its copy has not finished;
 *   hasty_films_device returns the films' device pointers (for
 *     hasty_estimate_device) and the samples added so far.
 * A film set belongs to one scene and must not outlive it; calls on one film
 * set from several threads are serialised inside. */
typedef struct HastyFilms HastyFilms;
int hasty_films_create(HastyScene* scene, HastyFilms** out_films);
void hasty_films_destroy(HastyFilms* films);
int hasty_films_render(HastyFilms* films, uint32_t seed, uint32_t first_sample, uint32_t num_samples);
int hasty_films_snapshot(HastyFilms* films);
int hasty_films_read(HastyFilms* films, int wait, float* color, float* normal, float* albedo,
                     uint32_t* num_samples);
int hasty_films_device(HastyFilms* films, float** d_color, float** d_normal, float** d_albedo,
                       uint32_t* num_samples);
/* The viewers' last step (PathTracingViewer.cpp:157-162) over the film sets
 * of one scene, one per render thread: computeEstimate of their colour,
 * normal and albedo films on the GPU (colour times 2^exposure), copied to
 * host buffers of W*H*3 floats (for writeEXR, hasty_image_write_exr). */
int hasty_films_estimate(HastyFilms* const* films, uint32_t n_films, float exposure, float* color, float* normal,
                         float* albedo);

/* Image files, on the host (no device work): reads a PFM, PPM / PGM, PNG or
 * OpenEXR file into float pixels, row 0 at the top, channels interleaved --
 * what OIIO hands readImage1f / readImage3f (core/src/Image.cpp:74-138).
 * With pixels NULL only the size is returned; otherwise capacity must hold
 * width*height*channels floats. */
int hasty_image_read(const char* path, uint32_t* width, uint32_t* height, uint32_t* channels, float* pixels,
                     uint64_t capacity);
/* writeEXR (core/src/Image.cpp:140-152): an uncompressed OpenEXR file of
 * FLOAT channels (stored B, G, R, the format's order) from width*height RGB
 * floats, row 0 at the top. */
int hasty_image_write_exr(const char* path, uint32_t width, uint32_t height, const float* rgb);

const char* hasty_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* HASTY_AMD_H */
