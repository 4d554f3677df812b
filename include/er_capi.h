/*
 * er_capi.h - the C-ABI drop-in boundary of the MI355X single-scattering volume tracer.
 *
 * Every entry point returns an int status (ER_OK == 0). A non-zero status carries the
 * exception level the reference would have thrown (Source/exception.h:27-55,
 * Source/enums.h:69-75) and er_last_error() returns its message.
 *
 * The descriptor structs are plain-old-data mirrors of the reference's bindable classes.
 * They hold only int32, float and pointer fields, so C, C++, ctypes and cgo lay them out
 * identically; er_abi_sizeof() lets a binding check that.
 */
#ifndef ER_CAPI_H
#define ER_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__) || defined(__clang__)
#define ER_API __attribute__((visibility("default")))
#else
#define ER_API
#endif

/* ---- constants ------------------------------------------------------------------------ */

#define ER_MAX_NO_TF_NODES   128   /* Source/defines.h:77; a PLF holds at most 127 nodes (plf.h:53) */
#define ER_MAX_INDICES       256   /* Source/vector.h:554-579 (Indices: 256 ints + Count) */
#define ER_LIST_CAPACITY     32    /* Source/list.cuh:27 (MaxSize); 31 live objects per kind */
#define ER_TF_LUT_NODES      256   /* node-table transfer function (er_set_tf_lut), 4 channels per node */

/* status codes: 0 = success, otherwise the reference's ExceptionLevel + 1 */
#define ER_OK        0
#define ER_INFO      1   /* Enums::Info    */
#define ER_WARNING   2   /* Enums::Warning: unknown resource ID (list.cuh:150-164) */
#define ER_ERROR     3   /* Enums::Error:   device failure (wrapper.cuh:38-46), bad argument */
#define ER_FATAL     4   /* Enums::Fatal    */

/* Enums::ShapeType (Source/enums.h:52-61) */
#define ER_SHAPE_PLANE    0
#define ER_SHAPE_DISK     1
#define ER_SHAPE_RING     2
#define ER_SHAPE_BOX      3
#define ER_SHAPE_SPHERE   4
#define ER_SHAPE_CYLINDER 5

/* Enums::ShadingMode (Source/enums.h:63-71) */
#define ER_SHADING_BRDF_ONLY           0
#define ER_SHADING_PHASE_ONLY          1
#define ER_SHADING_HYBRID              2
#define ER_SHADING_MODULATION          3
#define ER_SHADING_THRESHOLD           4
#define ER_SHADING_GRADIENT_MAGNITUDE  5

/* Enums::GradientMode (Source/enums.h:73-78) */
#define ER_GRADIENT_FORWARD  0
#define ER_GRADIENT_CENTRAL  1
#define ER_GRADIENT_FILTERED 2

/* Enums::EmissionUnit (Source/enums.h:87-92) */
#define ER_UNIT_POWER     0
#define ER_UNIT_LUX       1
#define ER_UNIT_INTENSITY 2

/* Enums::TextureType / ProceduralType (Source/enums.h:40-50) */
#define ER_TEXTURE_PROCEDURAL 0
#define ER_TEXTURE_BITMAP     1
#define ER_PROCEDURAL_UNIFORM  0
#define ER_PROCEDURAL_CHECKER  1
#define ER_PROCEDURAL_GRADIENT 2

/* Estimator modes (new). COMPAT reproduces the live reference estimator bit-for-bit in RNG
 * use and integer decisions; FULL enables the designed-but-dead features (shading-mode
 * selection, both halves of MIS, clipping objects, bilateral display filter). */
#define ER_ESTIMATOR_COMPAT 0
#define ER_ESTIMATOR_FULL   1

/* Free-flight samplers (new). RAYMARCH = Source/raymarching.h:30-113 (jittered fixed-step
 * tracking); WOODCOCK = delta tracking with a piecewise constant majorant, DS^2 times the
 * largest opacity a sample can take in each 8-cell macrocell (full estimator only; design
 * from the commented sampler at Source/raymarching.h:115-273). */
#define ER_SAMPLER_RAYMARCH 0
#define ER_SAMPLER_WOODCOCK 1

/* ---- descriptors ---------------------------------------------------------------------- */

/* PiecewiseLinearFunction<128> (Source/plf.h:26-92 over Source/pf.h:26-71). Initialise
 * with er_plf_init (node_range = {FLT_MAX, FLT_MIN}, count = 0) and fill with
 * er_plf_add_node, or fill the fields directly. */
typedef struct er_plf {
    float node_range[2];
    float position[ER_MAX_NO_TF_NODES];
    float value[ER_MAX_NO_TF_NODES];
    int   count;
} er_plf;

/* ColorTransferFunction1D (Source/transferfunction.h:117-154): one PLF per XYZ channel. */
typedef struct er_color_tf {
    er_plf plf[3];
} er_color_tf;

/* Camera (Source/camera.h:98-109): the caller-set fields. The derived fields
 * (N, U, V, Screen, InvScreen; camera.h:61-96) are computed at bind time. */
typedef struct er_camera {
    int   film_size[2];
    float pos[3];
    float target[3];
    float up[3];
    float focal_distance;   /* -1 => |target - pos| (camera.h:69-70) */
    float aperture_size;
    float clip_near;
    float clip_far;
    float exposure;
    float gamma;
    float fov;              /* degrees */
} er_camera;

/* RenderSettings (Source/rendersettings.h:27-131) plus the settings this build adds. */
typedef struct er_render_settings {
    /* TraversalSettings (rendersettings.h:33-54) */
    float step_factor_primary;     /* default 0.1 */
    float step_factor_shadow;      /* default 0.1 */
    int   shadows;                 /* default 1   */
    float max_shadow_distance;     /* default 1.0 */
    /* ShadingSettings (rendersettings.h:66-98) */
    int   shading_type;            /* ER_SHADING_*; ignored in COMPAT (the live code is BRDF-only) */
    float density_scale;           /* default 100 */
    int   opacity_modulated;       /* default 1 */
    int   gradient_computation;    /* ER_GRADIENT_*; default central */
    float gradient_threshold;      /* default 0.5 */
    float gradient_factor;         /* default 0.5 */
    /* new */
    int   estimator;               /* ER_ESTIMATOR_* */
    int   sampler;                 /* ER_SAMPLER_*   */
    int   bilateral;               /* FULL only: bilateral filter on the RGBA8 display image */
    int   bilateral_radius;        /* taps either side (Source/filter.h:34-40 KernelRadius) */
    float bilateral_sigma_spatial; /* KernelD[k] = exp(-k^2 / (2 s^2)) */
    float bilateral_sigma_range;   /* GaussSimilarity[d] = exp(-d^2 / (2 s^2)), d = |dY| in u8 units */
} er_render_settings;

/* ErTracer (Source/ertracer.h:33-117). */
typedef struct er_tracer_desc {
    int          id;                /* in/out: -1 before the first bind (erbindable.h:31-36,65) */
    int          enabled;
    er_plf       opacity;           /* Opacity1D    */
    er_color_tf  diffuse;           /* Diffuse1D    */
    er_color_tf  specular;          /* Specular1D   */
    er_plf       glossiness;        /* Glossiness1D */
    er_color_tf  emission;          /* Emission1D   */
    er_camera    camera;
    er_render_settings settings;
    int          no_iterations;     /* copied on (re)bind; 0 restarts accumulation */
    int          volume_id;
    int          light_ids[ER_MAX_INDICES];
    int          light_count;
    int          object_ids[ER_MAX_INDICES];
    int          object_count;
    int          clipping_ids[ER_MAX_INDICES];
    int          clipping_count;
} er_tracer_desc;

/* Shape (Source/shape.h:47-112). The caller supplies both tm and inv_tm; area is derived
 * at bind time (Shape::Update, shape.h:91-103). Matrices are row-major NN[row][col]. */
typedef struct er_shape {
    float tm[4][4];
    float inv_tm[4][4];
    int   one_sided;
    int   type;                     /* ER_SHAPE_* */
    float size[3];
    float area;                     /* output of bind */
    float inner_radius;
    float outer_radius;
} er_shape;

/* ErVolume (Source/ervolume.h:28-74). voxels: resolution[0]*resolution[1]*resolution[2]
 * uint16, x fastest (index z*Rx*Ry + y*Rx + x, Source/buffer3d.h:210-214). The voxels are
 * uploaded when voxels_dirty is set (Buffer3D dirty-gated copy, buffer3d.h:52-65); bind
 * clears it, so an unchanged rebind skips the upload, and a rebind without new voxels keeps
 * the volume's resolution (spacing and normalisation come from the descriptor). */
typedef struct er_volume_desc {
    int             id;
    int             enabled;
    int             resolution[3];
    float           spacing[3];
    int             normalize_size; /* 0 => zero-size volume, as the reference (volume.h:106-117) */
    int             voxels_dirty;
    const uint16_t* voxels;
} er_volume_desc;

/* ErLight (Source/erlight.h:27-66). */
typedef struct er_light_desc {
    int      id;
    int      enabled;
    int      visible;               /* visible to camera rays (lights.h:102) */
    int      texture_id;            /* < 0 => emits nothing (textures.h:66-67) */
    float    multiplier;
    int      unit;                  /* ER_UNIT_*; LUX divides by the area (lights.h:55-56) */
    er_shape shape;
} er_light_desc;

/* Procedural (Source/procedural.h:27-60). */
typedef struct er_procedural {
    int         type;               /* ER_PROCEDURAL_* */
    float       uniform_color[3];
    float       checker_color1[3];
    float       checker_color2[3];
    er_color_tf gradient;
} er_procedural;

/* ErTexture (Source/ertexture.h:27-82). */
typedef struct er_texture_desc {
    int           id;
    int           enabled;
    int           type;             /* ER_TEXTURE_* */
    float         output_level;
    int           bitmap_id;
    float         offset[2];
    float         repeat[2];
    int           flip[2];
    er_procedural procedural;
} er_texture_desc;

/* ErClippingObject (Source/erclippingobject.h:27-58). */
typedef struct er_clipping_desc {
    int      id;
    int      enabled;
    int      invert;
    er_shape shape;
} er_clipping_desc;

/* ErObject (Source/erobject.h:27-67). */
typedef struct er_object_desc {
    int      id;
    int      enabled;
    int      diffuse_texture_id;
    int      specular_texture_id;
    int      glossiness_texture_id;
    float    ior;
    er_shape shape;
} er_object_desc;

/* ErBitmap (Source/erbitmap.h:28-61): RGBA8 pixels, row-major y*W + x. */
typedef struct er_bitmap_desc {
    int            id;
    int            enabled;
    int            resolution[2];
    int            pixels_dirty;
    const uint8_t* pixels;
} er_bitmap_desc;

/* Per-stage timing of er_render_passes (hipEvent-measured on the tracer's stream). */
typedef struct er_timing {
    float total_ms;        /* first launch to last completion */
    float trace_ms;        /* sum over passes of primary start to combine end (passes overlap) */
    float post_ms;         /* sum over passes of filter + accumulate + tonemap (+ bilateral) */
    int   passes;
    float primary_ms;      /* sum over passes of the primary-march kernel */
} er_timing;

/* Sample counters of one pass (er_set_counting): the unit counts of the algorithmic-bytes
 * model, B = 16 * (primary + shadow + shading) + 64 bytes per path. */
typedef struct er_counters {
    uint64_t paths;
    uint64_t primary_samples;   /* iterations of raymarching.h:55-68 */
    uint64_t shadow_samples;    /* iterations of raymarching.h:97-110 */
    uint64_t shading_samples;   /* 1 (transport.h:119) + gradient samples per volume scatter */
    uint64_t primary_fetches;   /* samples that read voxels (the rest are in empty bricks) */
    uint64_t shadow_fetches;
    uint64_t scatter_events;
    uint64_t shadow_rays;
    uint64_t march_iterations;  /* iterations of the march loops that computed a sample position
                                   (Woodcock: majorant segments walked) */
    uint64_t skipped_samples;   /* samples advanced over inside empty macrocells without a position */
    uint64_t skips_undone;      /* macrocell skips rejected by the exactness check */
    uint64_t reserved;
} er_counters;

/* ---- helpers ------------------------------------------------------------------------- */

ER_API int  er_abi_version(void);
/* sizeof of each descriptor, in the order tracer, volume, light, texture, clipping, object,
 * bitmap, plf, shape, timing, counters, batch item, tune; returns how many it wrote. */
ER_API int  er_abi_sizeof(int* sizes, int n);
ER_API void er_plf_init(er_plf* plf);
ER_API void er_plf_add_node(er_plf* plf, float position, float value);  /* plf.h:50-66 */
ER_API void er_tracer_desc_init(er_tracer_desc* d);   /* ErTracer() defaults */
ER_API void er_volume_desc_init(er_volume_desc* d);
ER_API void er_light_desc_init(er_light_desc* d);
ER_API void er_texture_desc_init(er_texture_desc* d);
ER_API void er_clipping_desc_init(er_clipping_desc* d);
ER_API void er_object_desc_init(er_object_desc* d);
ER_API void er_bitmap_desc_init(er_bitmap_desc* d);
ER_API int  er_last_error(char* buf, size_t size);    /* message of the last non-OK status */

/* ---- the reference entry points (Source/exposurerender.h:32-42) ----------------------- */

/* BindTracer/BindVolume/... (exposurerender.h:32-38; core.cu:56-124; list.cuh:59-106).
 * bind != 0: first bind assigns d->id, a rebind with the same id updates in place;
 * bind == 0: unbinds (an unknown id is ignored, list.cuh:89-93). A 32nd live object of
 * one kind is ignored and d->id stays -1 (list.cuh:63-67). */
ER_API int er_bind_tracer(er_tracer_desc* d, int bind);
ER_API int er_bind_volume(er_volume_desc* d, int bind);
ER_API int er_bind_light(er_light_desc* d, int bind);
ER_API int er_bind_object(er_object_desc* d, int bind);
ER_API int er_bind_clipping_object(er_clipping_desc* d, int bind);
ER_API int er_bind_texture(er_texture_desc* d, int bind);
ER_API int er_bind_bitmap(er_bitmap_desc* d, int bind);

/* RenderEstimate (exposurerender.h:39; core.cu:126-136): one progressive pass - trace,
 * 3x3 Gaussian frame filter, running average, tonemap - then NoIterations++. Synchronous. */
ER_API int er_render_estimate(int tracer_id);

/* GetEstimate (exposurerender.h:40; core.cu:138-143): W*H RGBA8, row-major y*W + x.
 * size must be at least W*H*4. With the bilateral filter on, the filtered image. */
ER_API int er_get_estimate(int tracer_id, uint8_t* rgba, size_t size);

/* GetAutoFocusDistance (exposurerender.h:41): ComputeAutoFocusDistance (autofocus.cuh:26-75),
 * which the reference leaves commented out (core.cu:145-148) - the mean distance from the
 * camera of the volume events of 100 rays through film point (u, v), each jittered by 0.01 of
 * a concentric-disk sample in screen space. The RNG stream is a fixed function of the tracer
 * ID and the film point (the reference used rand()), exposed for checking through
 * er_autofocus_seeds. */
ER_API int er_get_autofocus_distance(int tracer_id, int film_u, int film_v, float* distance);
/* the two RNG seeds er_get_autofocus_distance starts from (SplitMix64 of the tracer ID and
 * film point, zero halves replaced by 1) */
ER_API int er_autofocus_seeds(int tracer_id, int film_u, int film_v, uint32_t* s0, uint32_t* s1);

/* GetNoIterations (exposurerender.h:42; stub in core.cu:150-153): passes since the last bind. */
ER_API int er_get_no_iterations(int tracer_id, int* no_iterations);

/* ---- extensions (tests, benchmarks, hosts that manage passes themselves) -------------- */

ER_API int er_init(void);
ER_API int er_shutdown(void);     /* unbinds everything and frees device memory */
ER_API int er_synchronize(void);  /* waits for all work on the device */

/* Per-pixel RNG seeds (FrameBuffer::RandomSeeds1/2, framebuffer.h:100-101). count = W*H. */
ER_API int er_get_seeds(int tracer_id, uint32_t* seeds1, uint32_t* seeds2, size_t count);
ER_API int er_set_seeds(int tracer_id, const uint32_t* seeds1, const uint32_t* seeds2, size_t count);
/* Deterministic per-pixel seeds: SplitMix64(seed ^ pixel) split into two non-zero halves. */
ER_API int er_seed_tracer(int tracer_id, uint64_t seed);

/* Float readbacks: the per-pass unfiltered frame estimate (SingleScattering output) and the
 * running estimate; XYZA float, W*H*16 bytes. */
ER_API int er_get_frame_xyza(int tracer_id, float* xyza, size_t size);
ER_API int er_get_running_xyza(int tracer_id, float* xyza, size_t size);

/* Runs `passes` progressive passes back-to-back with one host synchronisation at the end
 * (same result as `passes` calls of er_render_estimate). timing may be NULL; with timing the
 * tracer keeps four events per pass of its longest timed call until it is released. */
ER_API int er_render_passes(int tracer_id, int passes, er_timing* timing);

/* Sample counting: while on, each pass accumulates counters (a slower kernel variant). */
ER_API int er_set_counting(int tracer_id, int on);
ER_API int er_get_counters(int tracer_id, er_counters* counters);
ER_API int er_reset_counters(int tracer_id);

/* Maximum gradient magnitude of the tracer's volume (full-mode hybrid shading normaliser,
 * Source/gradientmagnitude.cuh:46-70). */
ER_API int er_get_gradient_magnitude_max(int tracer_id, float* gm_max);

/* -log(u) table entry used for the free-path threshold, for u = m * 2^-23. */
ER_API int er_neglog_table_entry(uint32_t m, float* value);

/* Capacity of each resource list (Source/list.cuh:27 MaxSize = 32: at most capacity - 1 live
 * objects per kind, further binds are ignored). The default is the reference's; hosts that
 * keep many scenes bound at once (batches) raise it. A single tracer still uses at most
 * ER_LIST_CAPACITY distinct lights and textures. */
ER_API int er_set_list_capacity(int capacity);

/* ---- launch autotuner (new; no counterpart in the reference) --------------------------- */

/* A tracer's launch configuration. None of it changes the image: paths are per pixel and the
 * empty-space skipping is exact, so every configuration renders the same bits. */
typedef struct er_tune {
    int   brick;                  /* macrocell (brick) edge in cells: 2, 4, 8, 16, 32 */
    int   tile_w, tile_h;         /* primary tile, 64 pixels: 4x16, 8x8, 16x4 or 32x2 */
    int   blocks_per_cu;          /* persistent 256-lane primary workgroups per CU */
    int   shadow_waves_per_cu;    /* persistent 64-lane shadow-march waves per CU (default 4) */
    int   heavy_tiles;            /* most expensive tiles of a pass whose rays the next pass marches
                                     with several lanes per ray (heavy_rays_per_wave; default 0,
                                     with 64 rays per wave only) */
    int   rays_per_wave;          /* primary rays one 64-lane wave marches at once: 64 (one lane per
                                     ray, default), 16 or 4 (groups of 4 or 16 lanes per ray, each
                                     round trip reading 24 or 32 consecutive samples of the ray), 1
                                     (the whole wave per ray); Woodcock tracking always takes 64.
                                     The reference runs one thread per pixel
                                     (singlescattering.cuh:34-38). */
    int   shadow_rays_per_wave;   /* shadow rays one wave marches at once: 64 (one lane per ray,
                                     default), 16 or 4 (groups of 4 or 16 lanes per ray); Woodcock
                                     tracking always takes 64 */
    int   heavy_rays_per_wave;    /* rays of a heavy tile one wave marches at once: 4 (groups of
                                     16 lanes per ray, default), 16 (groups of 4 lanes) or 1 (the
                                     whole wave per ray) */
    float ms_per_pass;            /* er_autotune: the chosen configuration's time per pass */
    float default_ms_per_pass;    /* er_autotune: the starting configuration's time per pass */
    int   trials;                 /* er_autotune: configurations timed (0: taken from the cache) */
} er_tune;

#define ER_TUNE_FORCE 1           /* er_autotune flag: measure even if the scene is cached */

/* Times `passes` passes of each candidate (passes <= 0: as many as make a trial of about
 * 30 ms, from one probe) from the tracer's current seeds, median of three, and keeps a candidate
 * that is at least 3% faster: bricks, then rays per wave, then shadow rays per wave, then tile
 * shapes, then workgroups per CU, then shadow waves per CU, then heavy tiles (0, 256 or 768,
 * each with 16, 4 or 1 rays per wave), then their rays per wave alone; a second sweep over the
 * same axes follows when the first changed anything, then a closer look at the tile shapes,
 * workgroups per CU and shadow waves per CU (median of six, 1.5% faster to replace). The
 * tracer's seeds, running estimate, display, counters and pass count are restored afterwards.
 * Results are cached per scene (descriptor, node table and volume contents). */
ER_API int er_autotune(int tracer_id, int passes, int flags, er_tune* out);
ER_API int er_set_tune(int tracer_id, const er_tune* tune);
ER_API int er_get_tune(int tracer_id, er_tune* tune);

/* ---- batches (new; no counterpart in the reference) ----------------------------------- */

typedef struct er_batch_item {
    int tracer_id;
    int passes;
} er_batch_item;

/* Renders items[i].passes progressive passes of each tracer - the same result, bit for bit,
 * as er_render_passes on each tracer in turn - with one pass round of every tracer per launch
 * sequence: a persistent primary kernel over the tiles of all tracers still rendering, then
 * shading, shadow-march, combine and post kernels over all of their queues. All rounds are
 * captured in one HIP graph, cached for the next call with the same jobs and launch shape.
 * Each tracer may appear once, and all must use the same estimator mode; samplers may differ
 * (with a Woodcock job in the batch every march takes one lane per ray). The batch marches
 * with the jobs' common rays per wave (one lane per ray when they differ). timing may be NULL
 * (trace_ms then covers whole rounds). */
ER_API int er_render_batch(const er_batch_item* items, int n, er_timing* timing);
/* One configuration for a batch: the er_autotune search over bricks, rays per wave, shadow
 * rays per wave, tile shapes and workgroups per CU, each candidate applied to every job (a
 * brick only where it fits the job's volume) and timed as `rounds` batch rounds (default 4);
 * every job's render state is restored and the jobs keep the chosen configuration. */
ER_API int er_autotune_batch(const er_batch_item* items, int n, int rounds, er_tune* out);

/* ---- differentiable transfer function (new; no counterpart in the reference) ----------- */

/* Node-table transfer function. lut holds ER_TF_LUT_NODES x 4 floats, node-major: opacity,
 * diffuse X, Y, Z. Node n sits at intensity n * range / (ER_TF_LUT_NODES - 1); between nodes
 * the value is linear, beyond the last node it is the last node's, and negative values clamp
 * to zero. While set, it replaces the tracer's Opacity1D and Diffuse1D (specular, glossiness
 * and emission keep their PLFs). lut == NULL returns to the PLFs. Restarts accumulation
 * (NoIterations = 0), as a tracer rebind does. */
ER_API int er_set_tf_lut(int tracer_id, const float* lut, float range);

/* Path-replay adjoint of `passes` progressive passes rendered from the per-pixel seeds
 * (seeds1, seeds2; count = W*H) with the current node table, starting from NoIterations = 0.
 * Input: dL/d(running XYZ estimate after the passes), W*H*3 floats. Output: dL/d(node table),
 * ER_TF_LUT_NODES*4 doubles (summed in FP64). Each path is replayed from its seed with the same
 * decisions as the forward pass; shadow visibility and the primary free flight are
 * differentiated through their expected values (transmittance exp(-DS * sum sigma dt) on the
 * same samples, and the log-probability of the sampled scatter sample). Requires the full
 * estimator, phase-function-only shading and the ray-march sampler. The tracer's own seeds,
 * frames and running estimate are left unchanged. */
ER_API int er_adjoint_lut(int tracer_id, const uint32_t* seeds1, const uint32_t* seeds2, size_t count, int passes,
                          const float* dL_drunning_xyz, double* dL_dlut);

/* No counterpart in the reference. er_adjoint_lut for callers that cannot allocate device
 * memory (code under HIP graph capture, tight optimisation or serving loops, hosts that forbid
 * hipMalloc): same arguments, same result up to FP32 summation order, but every work buffer
 * lives in one static device array that the library holds per GPU until the process ends
 * (about 95 MiB), so no call allocates or frees. Limits: films of at most 512 x 512 pixels
 * (larger ones are rejected with ER_ERROR); the tracer's device tables must already exist for
 * its current volume, film and launch configuration (a change of those needs one render pass or
 * one er_adjoint_lut call first, else ER_ERROR). Passes run one at a time on the tracer's
 * stream. Every call on a GPU shares that GPU's array: calls that overlap on one GPU race. */
ER_API int er_adjoint_lut_noalloc(int tracer_id, const uint32_t* seeds1, const uint32_t* seeds2, size_t count,
                                  int passes, const float* dL_drunning_xyz, double* dL_dlut);

/* The smooth estimator the adjoint differentiates, evaluated with the node table eval_lut
 * while every decision (scatter sample, light choice, directions) is taken with the current
 * table: the running XYZ estimate (W*H*3 floats) after `passes` passes from the given seeds.
 * With eval_lut equal to the current table its expectation equals the forward estimate's;
 * central differences of it in eval_lut check er_adjoint_lut. */
ER_API int er_adjoint_value(int tracer_id, const uint32_t* seeds1, const uint32_t* seeds2, size_t count, int passes,
                            const float* eval_lut, float* running_xyz);

/* ---- volume ingestion (new: the reference leaves loading to its absent application) ---- */

/* element types of a volume file */
#define ER_ELEMENT_UINT8    0
#define ER_ELEMENT_INT8     1
#define ER_ELEMENT_UINT16   2
#define ER_ELEMENT_INT16    3
#define ER_ELEMENT_UINT32   4
#define ER_ELEMENT_INT32    5
#define ER_ELEMENT_FLOAT32  6
#define ER_ELEMENT_FLOAT64  7

/* MetaImage header (.mhd): resolution (x, y, z), spacing and element type. The data may be
 * LOCAL or in a separate raw file; compressed data and file lists are not supported. */
ER_API int er_mhd_info(const char* path, int resolution[3], float spacing[3], int* element_type);
/* Reads a MetaImage volume into `voxels` (x fastest, what er_volume_desc.voxels takes).
 * uint8 / uint16 values are kept; other integer types are shifted by their minimum, floating
 * point is rescaled from [min, max] to [0, 65535]; u16 = (value - offset) * scale, clamped. */
ER_API int er_mhd_read(const char* path, uint16_t* voxels, size_t count, float* offset, float* scale);
/* The same for a headerless raw file of resolution[0] * [1] * [2] elements after
 * header_bytes bytes (-1: the data is the end of the file); msb: big-endian elements. */
ER_API int er_raw_read(const char* path, const int resolution[3], int element_type, int msb, long long header_bytes,
                       uint16_t* voxels, size_t count, float* offset, float* scale);

/* ---- synthetic phantoms (seeded, deterministic) --------------------------------------- */

/* Isotropic Gaussian blobs plus uniform noise in [-noise, noise]; clamp to [0,65535], round.
 * Blob centres uniform in the unit cube, sigma uniform in [sigma_lo, sigma_hi] of the extent,
 * peak uniform in [peak_lo, peak_hi]. */
ER_API int er_phantom_blobs(uint16_t* out, int rx, int ry, int rz, int n_blobs,
                            float sigma_lo, float sigma_hi, float peak_lo, float peak_hi,
                            float noise, uint64_t seed);
/* Nested/overlapping ellipsoids (n_ellipsoids) and, when n_tubes > 0, cylinders along random
 * axes; intensities drawn from `n_bands` tissue-like bands in [0, max_intensity]; Gaussian
 * noise of standard deviation noise_sigma. */
ER_API int er_phantom_ellipsoids(uint16_t* out, int rx, int ry, int rz, int n_ellipsoids,
                                 int n_tubes, float axis_lo, float axis_hi, int n_bands,
                                 float max_intensity, float noise_sigma, uint64_t seed);

#ifdef __cplusplus
}
#endif

#endif /* ER_CAPI_H */
