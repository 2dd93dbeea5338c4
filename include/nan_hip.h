/*
 * nan_hip.h - C ABI of libnan_hip.so, the MI355X (gfx950) texture-space
 * path tracer.
 *
 * Conventions follow the reference's native helpers: extern "C", int status
 * (0 = success), an optional caller error buffer (char* err, size_t cap;
 * surface_probe_sampler.h:120-129) plus nan_last_error()
 * (vertex_baking_utils.h:11-23), opaque handles with explicit destroy
 * functions (static_shadow_tree_encoder.h:93-115).  Device pointers are
 * plain pointers into HIP device memory (for instance torch tensors);
 * `stream` is a hipStream_t passed as void* (NULL = default stream).  The
 * library never frees caller memory.
 */
#ifndef NAN_HIP_H
#define NAN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NAN_API __attribute__((visibility("default")))

#define NAN_HIP_ABI_VERSION 3

/* Scene input: the buffers the reference's Scene uploads (scene.py:330-437).
 *   vertices        float[V][8]  object-space position, normal, uv
 *   indices         uint32[3T]   mesh-local, meshes concatenated
 *   mesh_descs      uint32[M][4] vertex_count, index_count, vertex_offset, index_offset
 *   instance_descs  uint32[I][4] mesh_id, material_id, transform_id, flags (bit0 odd-negative)
 *   transforms      float[X][16] row-major object-to-world
 *   inv_transposes  float[X][16] row-major transpose(inverse(M))
 *   materials       uint32[Mat][12] base_color.rgb (f32), flags (bit0 double-sided),
 *                                  emissive.rgb (f32), alpha_mode (0 opaque, 1 mask, 2 blend),
 *                                  base-colour, emissive and normal texture index
 *                                  (0xFFFFFFFF: none), alpha_cutoff (f32)
 *                                  (MaterialDesc, scene.slang; material.py:27-67)
 *   lighting        float[16]    sun dir.xyz, cos_half_angle, color*intensity.rgb, pad,
 *                                transmittance.rgb, pad, sky radiance.rgb, pad
 *   textures        uint32[Tx][4] first texel, width, height, flags (bit0: RGB is sRGB)
 *   texels          uint32[N]    RGBA8 texels of every texture, rows top to bottom,
 *                                R in the low byte (texture_manager.py:70-160)
 * Textures are sampled like the reference's SampleLevel(linear_sampler, uv, 0) on
 * mip 0: bilinear, wrap addressing, sRGB decoded per texel before filtering
 * (scene.slang:124-131, 313-368).  Materials whose alpha_mode is not opaque are
 * alpha tested on every candidate hit (scene.slang:134-149, 1017-1172). */
typedef struct NanSceneDesc {
    const float *vertices;
    uint32_t vertex_count;
    const uint32_t *indices;
    uint32_t index_count;
    const uint32_t *mesh_descs;
    uint32_t mesh_count;
    const uint32_t *instance_descs;
    uint32_t instance_count;
    const float *transforms;
    const float *inv_transposes;
    uint32_t transform_count;
    const uint32_t *materials;
    uint32_t material_count;
    const float *lighting;
    const uint32_t *textures;  /* NULL with texture_count 0 */
    uint32_t texture_count;
    const uint32_t *texels;
    uint64_t texel_count;
} NanSceneDesc;

/* BVH build options (0 selects the default). */
typedef struct NanBvhOptions {
    uint32_t sah_bins;       /* binned SAH bins, default 16 */
    uint32_t max_leaf_size;  /* triangles per leaf, 1..8, default 4 */
    uint32_t build_threads;  /* host threads for the build, default hardware concurrency */
} NanBvhOptions;

typedef struct NanSceneInfo {
    uint64_t triangle_count;
    uint64_t node_count;       /* 4-wide 64-byte nodes */
    uint64_t leaf_count;
    uint64_t device_bytes;     /* BVH + triangle + shading buffers */
    uint32_t max_stack;        /* worst-case traversal stack entries */
    uint32_t stack_depth;      /* stack compiled into the kernels in use */
    double build_ms;
    float sah_cost;
} NanSceneInfo;

typedef struct NanScene NanScene;
typedef struct NanTracer NanTracer;

/* Counter slots written by the *_counted entry points (uint64 each). */
enum {
    NAN_COUNTER_EXTEND_RAYS = 0,
    NAN_COUNTER_SHADOW_RAYS = 1,
    NAN_COUNTER_NODE_VISITS = 2,   /* extend (closest-hit) rays */
    NAN_COUNTER_TRI_TESTS = 3,     /* extend (closest-hit) rays */
    NAN_COUNTER_SLOT_UPDATES = 4,
    NAN_COUNTER_SHADOW_NODE_VISITS = 5,
    NAN_COUNTER_SHADOW_TRI_TESTS = 6,
    NAN_COUNTER_COUNT = 8
};

NAN_API int nan_version(void);
NAN_API const char *nan_last_error(void);

/* Builds the BVH on the host and uploads scene buffers (replaces Scene.__init__
 * BLAS/TLAS builds, scene.py:440-533). */
NAN_API int nan_scene_create(const NanSceneDesc *desc, const NanBvhOptions *options, NanScene **out,
                             char *err, size_t cap);
NAN_API void nan_scene_destroy(NanScene *scene);
NAN_API int nan_scene_info(const NanScene *scene, NanSceneInfo *out);
/* Replaces the lighting record (scene.slang:7-17 uniforms). */
NAN_API int nan_scene_set_lighting(NanScene *scene, const float *lighting16);
/* Attaches the sky-view table (float[200][200][4], device memory the caller keeps
 * alive; NULL detaches): rays that leave the scene then read Scene.sample_sky_lut
 * (scene.slang:469-518) times the sun's outer-space illuminance instead of the
 * constant sky of the lighting record.  Batches copy the table of each scene at
 * nan_batch_create. */
NAN_API int nan_scene_set_sky_view_lut(NanScene *scene, const float *sky_view_lut);
/* Russian roulette on the scene's paths (enable != 0); no counterpart in the
 * reference, whose paths always run to max_bounces (texture_space_path_tracer.
 * slang:57-84).  A path that would take another bounce survives with
 * probability min(1, max component of its throughput) and divides its
 * throughput by it, drawing one random number; the estimate stays unbiased
 * and paths end after 1 to max_bounces bounces.  Batches copy the flag of each
 * scene at nan_batch_create.  The learned forms refuse such scenes. */
NAN_API int nan_scene_set_russian_roulette(NanScene *scene, uint32_t enable);
/* Frame index of the stochastic alpha test of blend materials (Scene.frame_index,
 * scene.py:163, 1213-1216; read at scene.slang:146-148), written on `stream` by a
 * kernel, so a captured graph may hold the write.  Batches hold a frame index of
 * their own (nan_batch_set_frame_index). */
NAN_API int nan_scene_set_frame_index(NanScene *scene, uint32_t frame_index, void *stream);

/* Closest hit and occlusion for rays float[n][8] (origin, tmin, dir, tmax) in device
 * memory; writes hits float[n][4] (t, b1, b2, 0; t < 0 on a miss), ids uint32[n][2]
 * (instance, primitive; 0xFFFFFFFF on a miss) and occluded uint32[n].  Replaces
 * Scene.intersect / get_visibility (scene.slang:1017-1172). */
NAN_API int nan_trace_rays(NanScene *scene, const float *rays, uint64_t n, float *hits, uint32_t *ids,
                           uint32_t *occluded, void *stream);

/* Uploads a Mesh Colors layout (mesh_colors.py:299-337: face uint2, instance uint4,
 * side uint2 tables) and creates a tracer over scene (TextureSpacePathTracer.__init__,
 * texture_space_path_tracer.py:8-31).  The wavefront working set (per-path state,
 * queues, counters, spill area, second stream) is allocated here, sized from the
 * layout, so the tracer's first nan_tst_execute allocates nothing and may be
 * captured into a graph. */
NAN_API int nan_tracer_create(NanScene *scene, const uint32_t *face_table, uint32_t face_count,
                              const uint32_t *instance_table, const uint32_t *side_table,
                              uint32_t instance_count, uint64_t total_surface_texels,
                              uint64_t total_payload_count, NanTracer **out, char *err, size_t cap);
NAN_API void nan_tracer_destroy(NanTracer *tracer);
/* Device copies of the layout tables, for the resolve and filter passes
 * (texture_space_path_tracer.py:26-31 face/instance/side buffers). */
NAN_API int nan_tracer_tables(const NanTracer *tracer, const void **face_table, const void **instance_table,
                              const void **side_table);

/* Kernel organisation of nan_tst_execute; both give bit-identical payloads.
 * NAN_TRACER_WAVEFRONT (default): staged ray-gen / extend / shade / shadow /
 * finish kernels with persistent traversal; NAN_TRACER_MEGAKERNEL: one thread
 * per texel runs the whole path (the reference's dispatch shape).  When at
 * most 1/8 of the front texels are double-sided, the wavefront traces those
 * texels' paths on a second stream of the tracer's own, forked from the
 * caller's stream by an event and joined back before the call returns (also
 * under graph capture); work queued afterwards on the caller's stream sees
 * the whole frame.  Stage timing (nan_tracer_profile) keeps one stream. */
enum { NAN_TRACER_MEGAKERNEL = 0, NAN_TRACER_WAVEFRONT = 1 };
NAN_API int nan_tracer_set_mode(NanTracer *tracer, int mode);

/* Per-stage kernel timing with HIP events on the launch stream.  While enabled,
 * every tracer kernel launch is bracketed by events; nan_tracer_stage_times
 * waits for them and returns milliseconds summed per stage (double[NAN_STAGE_COUNT])
 * and launch counts (uint32[NAN_STAGE_COUNT], may be NULL), then clears them.
 * No reference counterpart (the reference has no GPU timers on this path). */
enum { NAN_STAGE_GENERATE = 0, NAN_STAGE_EXTEND, NAN_STAGE_SHADE, NAN_STAGE_SHADOW, NAN_STAGE_FINISH,
       NAN_STAGE_MEGAKERNEL,
       /* learned form (nan_tst_execute_learned): training paths, cache training
        * step, first bounce of every texel sample, cache queries */
       NAN_STAGE_LEARNED_PATHS, NAN_STAGE_CACHE_TRAIN, NAN_STAGE_FIRST_BOUNCE, NAN_STAGE_CACHE_QUERY,
       NAN_STAGE_COUNT };
NAN_API int nan_tracer_profile(NanTracer *tracer, int enable);
NAN_API int nan_tracer_stage_times(NanTracer *tracer, double *ms, uint32_t *launches);

/* Launch shape of the wavefront traversal kernels: persistent blocks per CU
 * (0 = occupancy limit), LDS stack ring entries per lane (16, 24 or 32; deeper
 * entries spill to HBM) and traversal steps between per-lane ray refills.
 * Results do not depend on it.  A new launch shape resizes the tracer's
 * working set here (allocation, device synchronisation), so this is not
 * called during graph capture; execute calls never allocate.  No reference
 * counterpart (autotuner hook). */
NAN_API int nan_tracer_tune(NanTracer *tracer, uint32_t blocks_per_cu, uint32_t lds_stack, uint32_t refill_steps);

/* Measurement hook (no reference counterpart): a VALU-bound kernel of
 * independent FP32 FMA chains (no memory traffic), `trips` loop trips of 32
 * FMAs per thread, launched to fill every SIMD and timed with HIP events on
 * the default stream (the second of two launches).  Returns the
 * milliseconds and the FMA wave-instructions of the timed launch: their
 * ratio is the VALU issue ceiling of plain FP32 code, the denominator of the
 * traversal kernels' measured SQ_INSTS_VALU rate. */
NAN_API int nan_calibrate_valu(uint32_t trips, double *ms_out, double *fma_wave_insts_out);

/* Many independent scenes traced by one set of launches: concatenates the
 * scenes' BVHs, shading records and layouts (each tracer must be over the
 * scene at the same index) into a batched tracer (wavefront mode).  Each
 * scene keeps its own cache, RNG streams and lighting, so every slot equals a
 * one-at-a-time run of that scene.  No reference counterpart (the reference
 * dispatches per instance of one scene, texture_space_path_tracer.py:42-59). */
NAN_API int nan_batch_create(NanScene *const *scenes, NanTracer *const *tracers, uint32_t count, NanTracer **out,
                             char *err, size_t cap);
/* Per-scene payload buffers (device pointers, one per scene); required before
 * nan_tst_execute on a batch, whose payload argument is then ignored. */
NAN_API int nan_batch_set_payloads(NanTracer *batch, void *const *payloads);
/* Per-scene path length (0 = the launch's max_bounces); the launch's
 * max_bounces must be at least the largest. */
NAN_API int nan_batch_set_bounces(NanTracer *batch, const uint32_t *max_bounces);
/* Frame index of the stochastic alpha test for every scene of the batch (as
 * nan_scene_set_frame_index; the batch's scenes share it, starting at 0). */
NAN_API int nan_batch_set_frame_index(NanTracer *batch, uint32_t frame_index, void *stream);

/* Device-frame mode for graph replay: while enabled, nan_tst_execute and
 * nan_tst_execute_learned take the iteration and reset flag from a device
 * counter (set here to iteration/reset) and advance it at the end of each
 * call, so one captured hipGraph of a call replays successive iterations.
 * Wavefront mode only.  No reference counterpart. */
NAN_API int nan_tracer_device_frame(NanTracer *tracer, int enable, uint32_t iteration, uint32_t reset);

/* Random streams (rng_mode).  NAN_RNG_REFERENCE is the reference's LCG
 * seeded per (texel, instance, sample) (common.slang:164-210,
 * texture_space_path_tracer.slang:169-171).  NAN_RNG_PCG gives every
 * (texel, instance) its own PCG sequence (RXS-M-XS output, 32-bit state, a
 * per-texel odd increment) started from a hash of the sample index; no
 * reference counterpart. */
enum { NAN_RNG_REFERENCE = 0, NAN_RNG_PCG = 1 };

/* One progressive iteration over every texel of every instance
 * (TextureSpacePathTracer.execute, texture_space_path_tracer.py:33-59; kernel
 * texture_space_path_tracer.slang:140-216).  payload: 16-byte slots
 * {f32 r,g,b; u32 count}[total_payload_count] in device memory.  Any
 * samples_per_texel; in wavefront mode max_bounces up to 127 (checked before
 * the first launch, so a refused call leaves the payload untouched). */
NAN_API int nan_tst_execute(NanTracer *tracer, void *payload, uint32_t iteration, uint32_t samples_per_texel,
                            uint32_t max_bounces, uint32_t reset, uint32_t rng_mode, void *stream);
/* Same, adding per-launch work counters (NAN_COUNTER_*) into counters
 * (uint64[NAN_COUNTER_COUNT], device memory, caller zeroes). */
NAN_API int nan_tst_execute_counted(NanTracer *tracer, void *payload, uint32_t iteration,
                                    uint32_t samples_per_texel, uint32_t max_bounces, uint32_t reset,
                                    uint32_t rng_mode, uint64_t *counters, void *stream);
/* Texel ray generation only (texture_space_path_tracer.slang:142-199) with the
 * reference's stream: per flat front texel writes uint32[4] = instance, primitive,
 * bits of bary.x, bits of bary.y and uint32 RNG state after the jitter draws
 * (rng_state may be NULL). */
NAN_API int nan_tst_raygen(NanTracer *tracer, uint32_t iteration, uint32_t samples_per_texel, uint32_t *out,
                           uint32_t *rng_state, void *stream);

/* RGB9E5 freeze of n slots (mesh_colors_rgb9e5_pack.slang:7-16). */
NAN_API int nan_rgb9e5_pack(const void *payload, uint32_t *packed, uint64_t n, void *stream);
/* Decode n words to float[n][4] (mesh_colors.slang:246-253). */
NAN_API int nan_rgb9e5_unpack(const uint32_t *packed, float *out, uint64_t n, void *stream);

/* Primary-ray resolve into out float[h][w][4] (mesh_colors_resolve.slang:29-73;
 * frozen != 0 reads RGB9E5 words, mesh_colors_frozen_resolve.slang:92-135).
 * camera16 (host memory): position, jitter_x, image_u, jitter_y, image_v, near, image_w, far. */
NAN_API int nan_resolve(NanTracer *tracer, const void *cache, int frozen, const float *camera16, float *out,
                        uint32_t width, uint32_t height, void *stream);

/* One pass of the Mesh Colors surface filter (MeshColorsSurfaceFilter.execute,
 * mesh_colors_surface_filter.py:96-128; kernel mesh_colors_surface_filter.slang:90-253):
 * every slot of the tracer's layout gets its one-ring, normal-weighted mean
 * from src into dst (payload slots, distinct buffers; sample counts copied).
 * adjacency uint32[F][4] (packed edge records of mesh_colors_adjacency.py:20-57,
 * adjacent faces instance-local), face_normals float[F][4] (world-space unit
 * normals); device memory.  neighbor_weight = exp(-0.5 / spatial_sigma^2). */
NAN_API int nan_surface_filter(NanTracer *tracer, const uint32_t *adjacency, const float *face_normals,
                               const void *src, void *dst, float neighbor_weight, float normal_sigma_radians,
                               void *stream);

/* Atmosphere lookup tables (AtmosphereTransmittanceLUT, AtmosphereMultiScatteringLUT,
 * AtmosphereSkyViewLUT .generate, atmosphere.py:4-225; kernels atmosphere.slang:143-397).
 * Each table is a float4 image in device memory, rows [y][x]: transmittance 256x64,
 * multiple scattering 32x32 (reads the transmittance table), sky view 200x200 (reads
 * both, for a sun direction given on the host). */
NAN_API int nan_atmosphere_transmittance_lut(float *out, void *stream);
NAN_API int nan_atmosphere_multiscattering_lut(const float *transmittance, float *out, void *stream);
NAN_API int nan_atmosphere_sky_view_lut(const float *transmittance, const float *multiscattering,
                                        const float *sun_direction3, float *out, void *stream);
/* Scene.get_sun_transmittance_at_view (scene.slang:997-1009): the transmittance table
 * at the view point and sun direction (host float[3]) into out4 (device float[4]). */
NAN_API int nan_atmosphere_view_transmittance(const float *transmittance, const float *sun_direction3, float *out4,
                                              void *stream);
/* Scene.sample_sky_lut (scene.slang:469-518) for n directions dirs float[n][4]
 * (device) into out float[n][4] (device), with the sun of lighting16 (host). */
NAN_API int nan_sky_view_lookup(const float *sky_view_lut, const float *lighting16, const float *dirs, uint64_t n,
                                float *out, void *stream);

/* Screen accumulation (accumulator.slang:3-29) and LPM tone map + Rec.709
 * (tone_mapper.slang:8-90) over n float4 pixels. */
NAN_API int nan_accumulate(const float *input, float *accum, float *output, uint64_t n, int reset, void *stream);
NAN_API int nan_tonemap(const float *input, float *output, uint64_t n, float exposure, void *stream);

/* ------------------------------------------------------------------------
 * Learned radiance cache.  No reference counterpart: the reference has no
 * learned cache; these entry points serve the learned form of the texel
 * tracer (paths end in the cache after one bounce) and its tests.
 *
 * Network: multiresolution hash-grid encoding (levels x 2 features, 2^log2
 * entries per level, FP16) of the position in the box [bbox_min, bbox_max],
 * concatenated with normal and albedo, then hidden_layers ReLU layers of
 * width and a linear RGB output, FP16 on MFMA with FP32 accumulation, no
 * biases.  Trained with Adam on the relative L2 loss
 * sum_c (p_c - t_c)^2 / (p_c^2 + 0.01) (p_c not differentiated in the
 * denominator) averaged over valid records and channels.
 * ------------------------------------------------------------------------ */
typedef struct NanNrc NanNrc;

typedef struct NanNrcConfig {
    uint32_t levels;              /* 1..16 */
    uint32_t features_per_level;  /* 2 */
    uint32_t log2_table_size;     /* entries per level, 8..24 */
    float base_resolution;        /* coarsest grid resolution */
    float max_resolution;         /* finest grid resolution */
    float bbox_min[3], bbox_max[3];
    uint32_t width;               /* 64 */
    uint32_t hidden_layers;       /* 4 */
    float learning_rate, beta1, beta2, epsilon;
    uint32_t seed;                /* parameter initialisation */
    uint32_t dense_table_adam;    /* 0: table entries with zero gradient are not updated */
    uint32_t minibatches;         /* Adam steps per training call (records split evenly), >= 1 */
} NanNrcConfig;

NAN_API int nan_nrc_create(const NanNrcConfig *config, NanNrc **out, char *err, size_t cap);
NAN_API void nan_nrc_destroy(NanNrc *nrc);
/* Number of MLP weights and hash-table parameters. */
NAN_API int nan_nrc_param_counts(const NanNrc *nrc, uint64_t *mlp_weights, uint64_t *table_params);
/* FP32 master parameters (host memory, either may be NULL): weights per layer
 * as [out][in] (48->64, 3 x 64->64, 64->16), table as [level][entry][2]. */
NAN_API int nan_nrc_get_params(const NanNrc *nrc, float *weights, float *table);
NAN_API int nan_nrc_set_params(NanNrc *nrc, const float *weights, const float *table);
/* Predicted RGB radiance (clamped at 0) for n queries: position, normal and
 * albedo float[n][3], out float[n][3]; device memory. */
NAN_API int nan_nrc_infer(NanNrc *nrc, const float *pos, const float *nrm, const float *alb, float *out, uint32_t n,
                          void *stream);
/* nan_nrc_infer without any allocation: no reference counterpart; for callers
 * that cannot allocate device memory (code under graph capture, tight serving
 * loops, harnesses that forbid it).  Same kernels and results as
 * nan_nrc_infer, with the scratch (encoded input, raw outputs) in a 7 MiB
 * device array of the library's own that each GPU holds until the process
 * ends.  At most 65,536 samples per call (more: an error, nothing launched).
 * All calls on one GPU share that array, so calls that overlap on one GPU
 * (other streams, unordered threads) race. */
NAN_API int nan_nrc_infer_noalloc(NanNrc *nrc, const float *pos, const float *nrm, const float *alb, float *out,
                                  uint32_t n, void *stream);
/* Training on n records with targets float[n][3]: forward, loss, backward and
 * Adam, once per minibatch (config.minibatches consecutive slices); adds the
 * last step's mean loss to *loss_out (device float, may be NULL). */
NAN_API int nan_nrc_train(NanNrc *nrc, const float *pos, const float *nrm, const float *alb, const float *target,
                          uint32_t n, float *loss_out, void *stream);
/* Which kernels a training step runs: fused = 1 (the default) takes the
 * three-launch step for minibatches of up to 32,768 records, on dense and
 * hashed levels alike (encoding, forward, loss and backward in one launch;
 * the weights' Adam beside the table-gradient atomics; the table's lazy
 * Adam over each dense level's lattice and each hashed level's table),
 * fused = 0 the seven-launch step.  Both sum table gradients with packed
 * FP16 atomics.  Larger minibatches always take the seven-launch step. */
NAN_API int nan_nrc_set_fused_training(NanNrc *nrc, int fused);
/* Testing hooks: the raw first four network outputs float[n][4], and the
 * gradients of the mean loss for targets float[n][3] with respect to the MLP
 * weights (float[mlp_weights], layout as nan_nrc_get_params) and to the
 * network input float[n][48] (hash features, normal, albedo, padding);
 * parameters are not updated.  The backward runs the kernels a training step
 * of this cache runs (see nan_nrc_set_fused_training). */
NAN_API int nan_nrc_forward_raw(NanNrc *nrc, const float *pos, const float *nrm, const float *alb, float *out,
                                uint32_t n, void *stream);
NAN_API int nan_nrc_backward_raw(NanNrc *nrc, const float *pos, const float *nrm, const float *alb,
                                 const float *target, uint32_t n, float *weight_grad, float *input_grad,
                                 void *stream);
/* Mean loss and valid record count of the last training step (host floats;
 * waits for the device). */
NAN_API int nan_nrc_stats(NanNrc *nrc, float *loss_and_count2);
/* Testing hook: the hash-table gradient of the mean loss as the Adam step
 * receives it (packed FP16 atomic sums, float[table_params], layout as
 * nan_nrc_get_params); parameters are not updated. */
NAN_API int nan_nrc_table_grad(NanNrc *nrc, const float *pos, const float *nrm, const float *alb,
                               const float *target, uint32_t n, float *table_grad, void *stream);
/* Measurement hook: encodes n queries, then launches the fused MLP forward
 * kernel alone `repeats` times between two HIP events on `stream` and
 * returns the average milliseconds per launch (waits for the device). */
NAN_API int nan_nrc_time_forward(NanNrc *nrc, const float *pos, const float *nrm, const float *alb, uint32_t n,
                                 uint32_t repeats, float *ms_out, void *stream);
/* Encoded network input float[n][48] (hash features, normal, albedo, zeros). */
NAN_API int nan_nrc_encode(NanNrc *nrc, const float *pos, const float *nrm, const float *alb, float *out, uint32_t n,
                           void *stream);

/* Learned form of nan_tst_execute: first trains the cache on train_samples
 * plain path suffixes (texels drawn uniformly, plain max_bounces paths whose
 * radiance leaving the first hit, emission excluded, is the target), then
 * runs one iteration in which every texel sample takes the sun term and one
 * cosine-sampled bounce as nan_tst_execute does and ends in the cache:
 * E = sun + (az / pdf) * (Le(y1) + max(0, cache(y1))).  With
 * NAN_LEARNED_EMITTER_NEE in flags, emissive triangles are also light-sampled
 * (uniform in area, one shadow ray) at x0 and at every suffix vertex that has
 * a successor, combined with the bounce-sampled emission by the power
 * heuristic; the estimate's expectation is unchanged.  With
 * NAN_LEARNED_RAMPED_MEAN, the k-th iteration since the reset enters the slot
 * with weight proportional to k instead of 1 (the cache improves with the
 * iterations).  With NAN_LEARNED_TEXEL_QUERY, paths end in the cache at x0
 * itself: training targets are the irradiance of plain paths from x0 less
 * the sun term (and, with NAN_LEARNED_EMITTER_NEE, with bounce-sampled
 * emission MIS-weighted against light sampling), and each texel sample is
 * E = sun + NEE(emitters) + max(0, cache(x0)), with no bounce ray; the cache
 * then fits the irradiance field over the surfaces (biased by its fit, no
 * per-texel sampling noise in the indirect term).  counters as in
 * nan_tst_execute_counted (may be NULL).  NAN_LEARNED_RNG_PCG selects the
 * PCG streams of NAN_RNG_PCG (default: the reference's).  With
 * NAN_LEARNED_OVERLAP (one-bounce form, per-texel kernels, no counters) the
 * call first copies the parameters its queries read, then the cache trains
 * on the previous call's training records on a second stream, beside this
 * call's tracing (texel samples and training paths in one launch) and its
 * queries of the copy, and this call's records are kept for the next call:
 * the queries see the cache as the previous call's training left it
 * (trained on records up to two frames old) and never wait for the
 * training.  Before the first such call the kept records are zeros (none
 * valid), so that call's step changes no parameter and only advances Adam's
 * step count.  The work is forked from and joined back into `stream`.
 * No reference counterpart. */
enum { NAN_LEARNED_EMITTER_NEE = 1, NAN_LEARNED_RAMPED_MEAN = 2, NAN_LEARNED_TEXEL_QUERY = 4,
       NAN_LEARNED_RNG_PCG = 8, NAN_LEARNED_OVERLAP = 16 };
NAN_API int nan_tst_execute_learned(NanTracer *tracer, NanNrc *nrc, void *payload, uint32_t iteration,
                                    uint32_t samples_per_texel, uint32_t max_bounces, uint32_t reset,
                                    uint32_t train_samples, uint32_t flags, uint64_t *counters, void *stream);
/* Sizes the learned forms' working set (training records, cache queries and
 * the network's batch scratch) for train_samples training paths and
 * samples_per_texel samples per texel; growing them synchronises the device,
 * so this must not be called during graph capture.  nan_tst_execute_learned grows the working
 * set itself when called outside a capture; inside one it allocates nothing
 * and fails if the working set is too small, so a caller capturing a first
 * learned iteration reserves it here beforehand.  No reference counterpart. */
NAN_API int nan_tst_learned_reserve(NanTracer *tracer, NanNrc *nrc, uint32_t samples_per_texel,
                                    uint32_t train_samples);
/* The training records of one iteration of nan_tst_execute_learned (n plain
 * paths from uniformly drawn texels): features pos/nrm/alb float[n][3], target
 * float[n][3], valid uint32[n], device memory; built by the wavefront queues
 * in wavefront mode (when no emitter is light-sampled), else by the per-texel
 * kernel, with identical results.  Test hook; no reference counterpart. */
NAN_API int nan_tst_learned_records(NanTracer *tracer, uint32_t iteration, uint32_t n, uint32_t max_bounces,
                                    uint32_t flags, float *pos, float *nrm, float *alb, float *target,
                                    uint32_t *valid, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* NAN_HIP_H */
