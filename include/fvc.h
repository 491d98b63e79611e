/*
 * fvc.h -- C ABI of the MI355X-native FrustumVolumeCaching render path.
 *
 * One shared library (libfvc.so) exposes the renderer / cache-renderer of the
 * reference behind plain C: opaque handles, POD structs, int status codes.
 * No C++ exception and no torch type crosses this boundary.
 *
 * Reference interface each entry point replaces (paths relative to the
 * reference repository root):
 *   SceneInfo / StepsizeInfo / CameraInfo / RaymarchInfo   include/common.h:74-218
 *   NerfNetwork<half>::{init, set_params, inference,
 *                       inferenceLatents, inferenceHead}   include/nerf_network.h:88-308
 *   PerformanceRenderer::{render, resizeRenderbuffers}     src/performance_renderer.cu:186-286
 *   CacheRenderer::{resizeCache, initCacheBlockwise,
 *                   swapCache, setCacheSettings}           src/cache_renderer.cu:832-1057,
 *                                                           include/cache_renderer.h:42-66
 *   CachePerformanceRenderer::render                       src/cache_performance_renderer.cu:1059-1263
 *   Cache (brick indirection, pool, known_ranges_to)       include/cache.h:12-74
 *
 * All device pointers are HIP device pointers on the renderer's device; all
 * other pointers are host pointers.  Every function returns FVC_OK (0) or a
 * negative status; fvc_last_error_message() gives the text (thread-local).
 */
#ifndef FVC_H
#define FVC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference throws C++ exceptions instead) ---- */
#define FVC_OK                   0
#define FVC_E_INVALID          (-1)  /* bad argument / null pointer */
#define FVC_E_UNSUPPORTED      (-2)  /* std::invalid_argument in my_fully_fused_mlp.cu:1386-1411 */
#define FVC_E_NOT_INITIALIZED  (-3)  /* "Cache is not initialized!" cache_performance_renderer.cu:1063-1064 */
#define FVC_E_HIP              (-4)  /* MY_CUDA_CHECK_THROW cuda_helper.h:20-27 */
#define FVC_E_OUT_OF_MEMORY    (-5)
#define FVC_E_NO_DEVICE        (-6)

/* ---- POD mirrors of the reference types ---- */
typedef struct { float x, y; } fvc_float2;
typedef struct { float x, y, z; } fvc_float3;
typedef struct { float x, y, z, w; } fvc_float4;

/* CameraMatrix (common.h:22-62): three rows of a 3x4 rigid transform. */
typedef struct { fvc_float4 m0, m1, m2; } fvc_camera_matrix;

/* ContractionType (common.h:66-72) */
enum {
  FVC_CONTRACTION_NONE = 0,
  FVC_CONTRACTION_AABB = 1,
  FVC_CONTRACTION_WARP_AABB_L2 = 2,
  FVC_CONTRACTION_WARP_AABB_LINF = 3
};

/* SceneInfo (common.h:74-91).  bools are int32 here. */
typedef struct {
  int32_t grid_resolution;        /* must be 128 (performance_renderer.cu:162-163) */
  int32_t grid_nlvl;              /* occupancy cascades */
  int32_t vals_per_lvl;           /* grid_resolution^3 */
  int32_t normalized;             /* positions already in [0,1] of the inner cascade */
  fvc_float3 aabb_from, aabb_to;  /* outer (ray-bounding) AABB */
  fvc_float3 contraction_aabb_from, contraction_aabb_to;
  int32_t contraction_type;
  int32_t is_open_gl;
  float alpha_thre;
} fvc_scene_info;

/* StepsizeInfo (common.h:93-164).  Fill the first six fields, then call
 * fvc_stepsize_setup() (setAdditionalInfo, common.h:109-128). */
typedef struct {
  float cone_angle, log1p_c, stepsize, near_, near_i, far_;
  float dt_min, dt_max, t_min, t_max, i_min, i_max;
} fvc_stepsize_info;

/* CameraInfo (common.h:166-195); depth-of-field fields are not supported. */
typedef struct {
  int32_t res_x, res_y;
  int32_t is_open_gl;
  fvc_float2 focal, principal;
  fvc_camera_matrix cam2world, world2cam;
} fvc_camera_info;

/* RaymarchInfo (common.h:210-218).  Only deterministic = 1 and
 * motion_blur = 0 are supported (the eval and cache-fill settings,
 * eval.cu:157,182,190). */
typedef struct {
  fvc_camera_info cam_info;
  fvc_stepsize_info stepsize_info;
  int32_t sample_index;           /* running-mean index of the uchar4 write */
  int32_t motion_blur;
  int32_t deterministic;
} fvc_raymarch_info;

/* Network description: the subset of the tcnn JSON config the render path
 * needs (nerf_network.h:88-118).  Supported: HashGrid L=16,F=2; density MLP
 * 32->64->64->16 (ReLU, no bias); head (n_latents + SH4(16) + pad 1.0 = 32)
 * ->64->3 (padded 16).  Anything else returns FVC_E_UNSUPPORTED.
 *
 * Plain models have one of two latent layouts:
 *  - the reference's (NerfNetwork::init, nerf_network.h:88-118, 334-342):
 *    latent_offset = 0, first_latent_is_density = 1, n_latents =
 *    base_n_output (1..16): the head reads output rows 0..n-1, log sigma
 *    first, and the cache stores the density in latent 0
 *    (DensityIntermediates); cache rendering needs n_latents <= 8
 *    (cache_performance_renderer.cu:964-967);
 *  - separate density (the synthetic workloads): latent_offset = 1,
 *    first_latent_is_density = 0, n_latents = 8, base_n_output >= 9: row 0
 *    is log sigma, rows 1..8 the latents, the cache keeps the density apart.
 *
 * split_head = 1 describes the view-dependent split model
 * (NerfNetwork::init_viewdep, nerf_network.h:120-158): the density MLP's 16
 * outputs and SH4 of the initial view direction (32 inputs) feed
 * mlp_first_head (64 wide, 1 or 2 hidden layers) whose outputs are the
 * n_latents = 8 latents; the colour head is the reference's custom kernel
 * (my_fully_fused_mlp.cu:677-895): inputs [8 latents, SH4(view dir),
 * SH4(view dir) - SH4(initial view dir), frequency encoding of the position
 * (freq_encoding_degree 4: 24 values; 0: none, 8 padding 1.0)] -> 64 -> 3.
 * first_latent_is_density = 1: latent 0 is log sigma and the cache stores
 * sigma in its place (DensityIntermediates, cache_renderer.cu:481-486),
 * latent_offset = 0; else latent_offset = 1.  FP16 only. */
/* MLP operand formats (no counterpart in the reference, which runs FP16):
 * FP16; BF16; FP8-E4M3 with per-layer scales; BF16X3: every operand split
 * into two bfloat16 terms (hi = bf16(v), lo = bf16(v - hi)) and each product
 * taken as hi*hi + hi*lo + lo*hi on three BF16 MFMAs (the lo*lo term
 * dropped; Henry, Tang, Heinecke 2019), FP32 accumulation throughout. */
enum { FVC_PREC_FP16 = 0, FVC_PREC_BF16 = 1, FVC_PREC_FP8 = 2, FVC_PREC_BF16X3 = 3 };
typedef struct {
  int32_t n_levels;               /* 16 */
  int32_t n_features_per_level;   /* 2 */
  int32_t log2_hashmap_size;      /* 19 .. 22 */
  int32_t base_resolution;        /* 16 */
  float   per_level_scale;        /* e.g. (2048/16)^(1/15) */
  int32_t base_width;             /* 64 */
  int32_t base_n_hidden_layers;   /* 2 */
  int32_t base_n_output;          /* 1 (log sigma) + n_latents, <= 16 */
  int32_t head_width;             /* 64 */
  int32_t head_n_hidden_layers;   /* 1 */
  int32_t n_latents;              /* rows of the density-MLP output the head reads as latents */
  int32_t latent_offset;          /* first such row (nerf_network.h:334-337): 0 = log sigma is latent 0 */
  int32_t sh_degree;              /* 4 */
  int32_t precision;              /* FVC_PREC_* (MLP arithmetic) */
  int32_t split_head;             /* 1: view-dependent split model (mlp_first_head) */
  int32_t first_head_width;       /* 64 */
  int32_t first_head_n_hidden_layers; /* 1 or 2 */
  int32_t first_head_n_input;     /* 32 = 16 density-MLP outputs + SH4 of the initial view direction */
  int32_t first_latent_is_density;/* 1 (default of the reference) or 0 */
  int32_t first_sh_degree;        /* 4 */
  int32_t freq_encoding_degree;   /* 4 or 0 */
  int32_t pad_;
} fvc_net_desc;

/* DensityInterpolVariant (cache_renderer.h:15-22) */
enum {
  FVC_INTERPOL_DENSITY = 0, FVC_INTERPOL_ALPHA = 1, FVC_INTERPOL_INTERMEDIATES = 2,
  FVC_INTERPOL_DENSITY_INTERMEDIATES = 3, FVC_INTERPOL_ALPHA_INTERMEDIATES = 4
};
/* InterpolFunction (common.h:221-227) */
enum { FVC_FN_LINEAR = 0, FVC_FN_SMOOTHSTEP = 1, FVC_FN_SMOOTHERSTEP = 2, FVC_FN_NEAREST = 3 };

/* CacheSettings (cache_renderer.h:24-32).  interpol_variant must match the
 * model: DENSITY for models whose latents exclude the density (separate
 * density row), DENSITY_INTERMEDIATES for models whose first latent is the
 * density (the reference's plain models and split models with
 * first_latent_is_density; the variant eval.cu:216 picks).  interpol_function
 * LINEAR, SMOOTHSTEP, SMOOTHERSTEP or NEAREST weights the 8 lookup corners
 * (applyInterpolationFunction, raymarch_common.h:393-402).
 * use_inter_brick_interpolation = 0 clamps the lookup point into its brick
 * (froxel2data, raymarch_common.h:341-357, as the reference's basic cache
 * renderer applies it; its performance renderer ignores the flag).
 * use_hw_interpol, reweight_intermediates and
 * disable_reweighting_first_intermediate are ignored, as by the reference's
 * performance renderer.  A renderer starts with the variant of its model,
 * linear weights and inter-brick interpolation (cache_renderer.h:55-63). */
typedef struct {
  int32_t interpol_variant;
  int32_t interpol_function;
  int32_t use_hw_interpol;        /* ignored: no texture hardware is used */
  int32_t reweight_intermediates;
  int32_t disable_reweighting_first_intermediate;
  int32_t use_inter_brick_interpolation;
} fvc_cache_settings;

/* CacheRenderer::InitStats (cache_renderer.h:104-109) + fill diagnostics */
typedef struct {
  float   samples_ppx;
  int32_t n_bricks_set;
  int32_t n_bricks_reserved;      /* pool capacity in bricks */
  int32_t overflow;               /* 1 if a brick request exceeded capacity */
  int32_t n_subiters;
  int64_t n_samples_evaluated;
} fvc_init_stats;

/* RenderStats (performance_renderer.h:45-49, cache_performance_renderer.h:123-131) */
typedef struct {
  float    samples_ppx;
  float    resample_samples_evaluated;  /* per pixel */
  float    cache_samples_evaluated;     /* per pixel */
  uint64_t n_resamples;
  uint64_t n_cache_hits;
  int32_t  n_subiters;
  int32_t  pad_;
} fvc_render_stats;

/* Differentiable render statistics */
typedef struct {
  int64_t n_rays;
  int64_t n_samples;              /* samples evaluated forward (all occupied samples marched) */
  int64_t n_cache_samples;        /* of which from the cache (cached render) */
} fvc_diff_stats;

/* Frustum cache geometry (Cache::frustumDims / brickArrayDims, cache.h:70-73) */
typedef struct {
  int32_t frustum_x, frustum_y, frustum_z;
  int32_t bricks_x, bricks_y, bricks_z;
  int32_t capacity_bricks;
  int32_t n_bricks_used;
  int32_t initialized;
  int32_t brick_size;             /* 8 */
  int32_t latent_format;          /* FVC_LATENT_FP16 / FVC_LATENT_FP8 */
} fvc_cache_info;

/* Per-kernel timing of the renderer's launches (HIP events; the reference's
 * CudaTimer per TimingState, renderer.h:79-105) with the algorithmic work
 * they did, accumulated while profiling is enabled. */
typedef struct {
  char    name[32];
  int64_t launches;
  double  total_ms;
  double  units;        /* samples, steps or pixels, per kernel */
  double  alg_bytes;    /* algorithmic bytes (no cache effects) */
  double  alg_flops;
} fvc_kernel_stats;

typedef struct fvc_model fvc_model;
typedef struct fvc_renderer fvc_renderer;

/* ---- library ---- */
const char* fvc_last_error_message(void);
int fvc_version(void);
/* StepsizeInfo::setAdditionalInfo (common.h:109-128); host only. */
int fvc_stepsize_setup(fvc_stepsize_info* info, const fvc_scene_info* scene);
/* Number of fp16 parameters of each block of the nerfacc export layout
 * (pos_enc.dat, mlp_base.dat, mlp_head.dat; loader.cu:232-285). Host only. */
int fvc_net_param_counts(const fvc_net_desc* net, int64_t* n_pos_enc, int64_t* n_mlp_base, int64_t* n_mlp_head);
/* The same with mlp_first_head.dat for split models (loader.cu:251-262). */
int fvc_net_param_counts_viewdep(const fvc_net_desc* net, int64_t* n_pos_enc, int64_t* n_mlp_base,
                                 int64_t* n_mlp_first_head, int64_t* n_mlp_head);

/* ---- model: NerfNetwork<half> + occupancy bitfield (loader.cu:20-55, 287-301) ----
 * occupancy_linear_u8: host [nlvl][res][res][res] (x slowest), >0 = occupied;
 * converted to the Morton bitfield the renderers read. params_*: host fp16 bits. */
int fvc_model_create(const fvc_scene_info* scene, const fvc_net_desc* net,
                     const uint8_t* occupancy_linear_u8,
                     const uint16_t* params_pos_enc, int64_t n_pos_enc,
                     const uint16_t* params_mlp_base, int64_t n_mlp_base,
                     const uint16_t* params_mlp_head, int64_t n_mlp_head,
                     fvc_model** out_model);
/* Split model: NerfNetwork::set_params(pos_enc, mlp_base, mlp_first_head,
 * mlp_head) (nerf_network.h:286-293) after init_viewdep (loader.cu:254-262). */
int fvc_model_create_viewdep(const fvc_scene_info* scene, const fvc_net_desc* net,
                             const uint8_t* occupancy_linear_u8,
                             const uint16_t* params_pos_enc, int64_t n_pos_enc,
                             const uint16_t* params_mlp_base, int64_t n_mlp_base,
                             const uint16_t* params_mlp_first_head, int64_t n_mlp_first_head,
                             const uint16_t* params_mlp_head, int64_t n_mlp_head,
                             fvc_model** out_model);
void fvc_model_destroy(fvc_model* model);
/* ---- model export: load_nerfacc_export (loader.cu:232-285) ----
 * <dir>/config.json ("scene", "mlp_base", "mlp_head"; parse_helper.h:25-63)
 * and raw .dat files: pos_enc.dat (fp16 table; optional, else appended to
 * mlp_base.dat), mlp_base.dat, mlp_head.dat (fp16), occupancy_grid.dat
 * (uint8 [nlvl][res][res][res]).  read_config parses config.json, applies
 * StepsizeInfo::setAdditionalInfo and checks the file sizes; counts[4] =
 * element counts of pos_enc, mlp_base, mlp_head, occupancy.  Host only. */
int fvc_nerfacc_read_config(const char* dir, fvc_scene_info* scene_out, fvc_stepsize_info* stepsize_out,
                            fvc_net_desc* net_out, int64_t* counts_out);
/* Read the four data blocks into caller buffers sized by read_config.  Host only. */
int fvc_nerfacc_read_data(const char* dir, uint16_t* pos_enc, uint16_t* mlp_base, uint16_t* mlp_head,
                          uint8_t* occupancy_linear_u8);
/* read_data for split models: mlp_first_head.dat as well (loader.cu:251),
 * sized by fvc_net_param_counts_viewdep.  Host only. */
int fvc_nerfacc_read_data_viewdep(const char* dir, uint16_t* pos_enc, uint16_t* mlp_base, uint16_t* mlp_first_head,
                                  uint16_t* mlp_head, uint8_t* occupancy_linear_u8);
/* read_config + read_data + fvc_model_create (or _viewdep when config.json
 * has "mlp_first_head"); scene / stepsize / net may be NULL. */
int fvc_model_load_nerfacc(const char* dir, fvc_scene_info* scene_out, fvc_stepsize_info* stepsize_out,
                           fvc_net_desc* net_out, fvc_model** out_model);
/* Copy the Morton bitfield (nlvl*res^3/8 bytes) to host. */
int fvc_model_get_bitfield(fvc_model* model, uint8_t* out_host, int64_t n_bytes);

/* NerfNetwork::inferenceLatents (nerf_network.h:173-202): pos_dev float[n][3]
 * in [0,1]^3 -> out_dev half[n][16] (row 0 log sigma, rows 1.. latents). */
int fvc_network_latents(fvc_model* model, const float* pos_dev, int64_t n, uint16_t* out_dev, void* stream);
/* NerfNetwork::inference (nerf_network.h:160-171): pos, dir01 float[n][3] ->
 * out half[n][4] = (r, g, b raw, log sigma) (output rows 0-3 of the reference). */
int fvc_network_inference(fvc_model* model, const float* pos_dev, const float* dir_dev, int64_t n, uint16_t* out_dev, void* stream);
/* NerfNetwork::inferenceHead (nerf_network.h:204-266): head_in half[n][16]
 * (row 0 log sigma, rows 1..8 latents), dir01 float[n][3] -> out half[n][4]. */
int fvc_network_head(fvc_model* model, const uint16_t* head_in_dev, const float* dir_dev, int64_t n, uint16_t* out_dev, void* stream);
/* Split models: NerfNetwork::inferenceLatents with the initial view
 * direction (nerf_network.h:173-196): pos, init_dir01 float[n][3] ->
 * half[n][16] = the 8 head latents (as the colour head reads them from
 * row latent_offset) in rows 0..7, log sigma in row 8, zeros after. */
int fvc_network_latents_viewdep(fvc_model* model, const float* pos_dev, const float* init_dir_dev, int64_t n,
                                uint16_t* out_dev, void* stream);
/* Split models: inferenceHead with the custom head (nerf_network.h:215-219,
 * my_fully_fused_mlp.cu:1363-1421): head_in half[n][16] in the layout of
 * fvc_network_latents_viewdep, pos (frequency encoding), dir01, init_dir01
 * float[n][3] -> half[n][4] = (r, g, b raw, density row: latent 0 when
 * first_latent_is_density, else row 8). */
int fvc_network_head_viewdep(fvc_model* model, const uint16_t* head_in_dev, const float* pos_dev,
                             const float* dir_dev, const float* init_dir_dev, int64_t n, uint16_t* out_dev,
                             void* stream);
/* Multiresolution hash-grid encoding alone: pos float[n][3] -> half[n][32]. */
int fvc_network_encode(fvc_model* model, const float* pos_dev, int64_t n, uint16_t* out_dev, void* stream);

/* ---- renderer: PerformanceRenderer + CacheRenderer + CachePerformanceRenderer ----
 * double_buffered = 0 mirrors CachePerformanceRenderer(&cache, &cache)
 * (eval.cu:224); 1 gives a separate fill target swapped by fvc_cache_swap. */
int fvc_renderer_create(fvc_model* model, int double_buffered, fvc_renderer** out_renderer);
void fvc_renderer_destroy(fvc_renderer* renderer);
/* Use the caller's HIP stream for render calls (NULL = renderer's own stream). */
int fvc_renderer_set_stream(fvc_renderer* renderer, void* stream);
/* Sub-iteration batching (the reference's TARGET_BATCH_SIZE = 2^21 and
 * MAX_N_STEPS_PER_SUBITER = 8, performance_renderer.h:54-58; here 2^25 and 32
 * by default): samples targeted per launch, and the cap on samples per ray
 * per launch (1..64).  Changes launch sizes only, never the image. */
int fvc_renderer_set_batching(fvc_renderer* renderer, int64_t target_samples, int32_t max_samples_per_ray);
/* Renderer::resizeRenderbuffers (renderer.h:112) */
int fvc_resize_renderbuffers(fvc_renderer* renderer, int width, int height);
/* PerformanceRenderer::render (performance_renderer.cu:214-286):
 * rgba_dev uchar4[H][W] (running mean over sample_index, as the reference). */
int fvc_render_full(fvc_renderer* renderer, const fvc_raymarch_info* rm, uint8_t* rgba_dev, fvc_render_stats* stats);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): fvc_render_full for views of at most 256 x 256 pixels, with
 * every per-frame device buffer (ray state, sample streams, counters) in one
 * static device array of the library (about 30 MB, held by each GPU the
 * library runs on until the process ends) instead of the renderer's
 * buffers; no device or pinned memory is allocated.  rgba_dev (uchar4, may
 * be NULL) as fvc_render_full; result_dev (float4 rgb + T per pixel, may be
 * NULL) receives the float result.  Larger views return FVC_E_INVALID.
 * Calls that overlap on one GPU race on the array (one at a time per GPU). */
int fvc_render_full_noalloc(fvc_renderer* renderer, const fvc_raymarch_info* rm, uint8_t* rgba_dev, float* result_dev,
                            fvc_render_stats* stats);
/* CacheRenderer::initCacheBlockwise (cache_renderer.cu:931-1051) into the fill target. */
int fvc_cache_fill(fvc_renderer* renderer, const fvc_raymarch_info* rm, fvc_init_stats* stats);
/* CachePerformanceRenderer::render (cache_performance_renderer.cu:1059-1263). */
int fvc_render_cached(fvc_renderer* renderer, const fvc_raymarch_info* rm, uint8_t* rgba_dev, fvc_render_stats* stats);
/* Batched multi-view render (no counterpart in the reference, which renders
 * one view per call): n_views views of any resolution sharing the stepsize
 * info, rgba[v] the device uchar4 image of view v (null entries skip the
 * write).  All views' rays form one alive list and one sample stream, so
 * each sub-iteration is a single launch per kernel over every view's
 * samples.  Stats are over all views; fvc_get_last_result returns the
 * views' results concatenated. */
int fvc_render_full_batch(fvc_renderer* renderer, const fvc_raymarch_info* rms, int32_t n_views, uint8_t* const* rgba,
                          fvc_render_stats* stats);
int fvc_render_cached_batch(fvc_renderer* renderer, const fvc_raymarch_info* rms, int32_t n_views,
                            uint8_t* const* rgba, fvc_render_stats* stats);
/* Mixed-scene full-render batch (no counterpart in the reference): view v
 * renders the scene of models[v] (null: the renderer's model).  The models
 * must share the renderer's network (description and parameters equal) and
 * scene info (box, cascades, contraction); their occupancy grids may differ,
 * so rays of one packed stream march different grids.  FVC_E_UNSUPPORTED
 * otherwise, or for more than 64 distinct models. */
int fvc_render_full_batch_scenes(fvc_renderer* renderer, fvc_model* const* models, const fvc_raymarch_info* rms,
                                 int32_t n_views, uint8_t* const* rgba, fvc_render_stats* stats);
/* ---- background refill (double-buffered renderers) ----
 * The viewer's cache update (application.cu:168-216) as library calls: the
 * fill of the back cache runs on its own thread and stream while frames
 * render from the front cache. */
#define FVC_FILL_IDLE    0
#define FVC_FILL_RUNNING 1
#define FVC_FILL_DONE    2   /* finished, not yet published */
typedef struct {
  float   min_hit_ratio;          /* refill when cache samples / samples < this (m_cache_miss_limit) */
  float   max_distance;           /* ... or the camera moved farther than this from the cache pose */
  int32_t enabled;
  int32_t pad_;
} fvc_refill_policy;
/* start filling the back cache from rm in the background */
int fvc_cache_fill_async(fvc_renderer* renderer, const fvc_raymarch_info* rm);
/* state of the background fill (FVC_FILL_*); wait = 1 blocks until it ends;
 * stats of a finished fill */
int fvc_cache_fill_poll(fvc_renderer* renderer, int wait, int32_t* state, fvc_init_stats* stats);
/* swap a finished background fill to the front (*swapped = 1), else no-op */
int fvc_cache_publish(fvc_renderer* renderer, int32_t* swapped);
int fvc_renderer_set_refill_policy(fvc_renderer* renderer, const fvc_refill_policy* policy);
/* render_cached with the policy: publishes a finished fill first, renders,
 * and starts a background fill at rm when the policy triggers
 * (*refill_started = 1).  The first call fills synchronously. */
int fvc_render_cached_auto(fvc_renderer* renderer, const fvc_raymarch_info* rm, uint8_t* rgba_dev,
                           fvc_render_stats* stats, int32_t* refill_started);

/* CacheRenderer::swapCache (cache_renderer.cu:1053-1057); a background fill
 * still running is waited for and published instead. */
int fvc_cache_swap(fvc_renderer* renderer);
/* CacheRenderer::setCacheSettings (cache_renderer.h:66) */
int fvc_cache_set_settings(fvc_renderer* renderer, const fvc_cache_settings* settings);
/* CachePerformanceRenderer::RenderMode (cache_performance_renderer.h:23-28),
 * the public _render_mode member (:121): ONLY_CACHE drops the samples that
 * would be resampled, ONLY_RESAMPLED the samples that would come from the
 * cache (cache_performance_renderer.cu:117-118, 125-126). */
#define FVC_RENDER_BOTH           0
#define FVC_RENDER_ONLY_CACHE     1
#define FVC_RENDER_ONLY_RESAMPLED 2
int fvc_cache_set_render_mode(fvc_renderer* renderer, int mode);
/* Latent storage of the brick pool for the next fills (no counterpart in
 * the reference, which stores half4 latents): FVC_LATENT_FP16, or
 * FVC_LATENT_FP8 = OCP E4M3 codes of l / s with one scale s = absmax / 448
 * per brick (a 16-byte voxel instead of 32). */
#define FVC_LATENT_FP16 0
#define FVC_LATENT_FP8  1
int fvc_cache_set_latent_format(fvc_renderer* renderer, int format);
/* The used part of a cache's pool decoded to floats, [n_bricks][512][10]:
 * 8 latents, density, isset (either latent format). */
int fvc_cache_get_pool_decoded(fvc_renderer* renderer, int which, float* out_host, int64_t n_floats);
/* Pool capacity override in bricks (0 = reference estimate n_rays/8 rounded
 * up to 16384, cache_renderer.cu:832-846 + cache.cu:143-146); at most
 * 2^23 - 1 bricks (FVC_E_INVALID above; the estimate is clamped to it). */
int fvc_cache_set_capacity(fvc_renderer* renderer, int32_t capacity_bricks);

/* ---- autotuning (no counterpart in the reference) ---- */
#define FVC_TUNE_MAX 64
typedef struct {
  int32_t mlp_rows;               /* chosen MLP rows per wavefront: 16, 32 or 64 */
  int32_t max_steps;              /* chosen samples per ray per launch */
  int64_t target_batch;           /* chosen target samples per launch */
  float   ms_best;                /* frame time of the choice (fastest of five renders) */
  float   ms_default;             /* ... of the settings before tuning */
  int32_t n_tried;
  int32_t pad_;
  float   ms[FVC_TUNE_MAX];       /* every candidate: time, rows, steps, target */
  int32_t rows[FVC_TUNE_MAX];
  int32_t steps[FVC_TUNE_MAX];
  int64_t targets[FVC_TUNE_MAX];
} fvc_tune_result;
/* Times full (cached = 0) or cached renders of rm over the candidate MLP
 * rows per wavefront and sub-iteration chunk sizes, keeps the fastest for
 * this renderer and returns the table.  Images are not written. */
int fvc_renderer_autotune(fvc_renderer* renderer, const fvc_raymarch_info* rm, int cached, fvc_tune_result* result);
/* Candidate sets of a sweep: MLP rows per wavefront (16, 32, 64), samples
 * per ray per launch (1..64) and target samples per launch (2^10..2^28; a
 * batch scales it by its pixels over its largest view's); at most
 * FVC_TUNE_MAX combinations. */
typedef struct {
  int32_t n_rows, rows[3];
  int32_t n_steps, steps[3];
  int32_t n_targets, pad_;
  int64_t targets[8];
} fvc_tune_space;
/* fvc_renderer_autotune over a batch of views (full: the views may name
 * models as in fvc_render_full_batch_scenes, models may be NULL) and a
 * candidate space (NULL: rows {16, 32, 64} x targets {2^23, 2^24, 2^25} x
 * steps {16, 32, 64}). */
int fvc_renderer_autotune_batch(fvc_renderer* renderer, fvc_model* const* models, const fvc_raymarch_info* rms,
                                int32_t n_views, int cached, const fvc_tune_space* space, fvc_tune_result* result);
/* Graph mode (no counterpart in the reference): each full or cached frame
 * is recorded with hipStreamBeginCapture into a hipGraph (ray setup and as
 * many sub-iterations as the previous frame needed, plus one; kernels read
 * the live counts on the device and idle when no ray is left) and replayed
 * with one launch; the executable graph is updated in place from frame to
 * frame.  A frame needing more sub-iterations continues with ordinary
 * launches.  Profiling (fvc_profile_enable) renders without graphs. */
int fvc_renderer_set_graph(fvc_renderer* renderer, int enable);
/* MLP rows per wavefront (16, 32 or 64) for the network kernels (the cache
 * heads always take 64 samples per wave, one per lane). */
int fvc_renderer_set_mlp_rows(fvc_renderer* renderer, int32_t rows);

/* ---- differentiable full render (training mode) ----
 * No counterpart entry point in the reference, whose renderers are
 * inference-only; the forward is PerformanceRenderer::render's per-ray
 * semantics (performance_renderer.cu:15-125) for a pixel list, with the
 * caller's FP32 parameter blocks (nerfacc layout: pos_enc [entries][2],
 * mlp_base W1|W2|W3, mlp_head H1|H2, row-major [out][in]) and FP32
 * arithmetic.  All pointers are device pointers.  pix: pixel indices y*W+x
 * of rm's camera; rgb [n][3] = composited colour + T (white background).
 * The parameter blocks must stay unchanged until the backward call.
 * backward: adds d(sum(d_rgb * rgb))/d(params) of the last forward to the
 * three gradient blocks (same layouts; the caller zeroes them). */
int fvc_diff_forward(fvc_renderer* renderer, const fvc_raymarch_info* rm, const int32_t* pix, int32_t n_rays,
                     const float* pos_enc, const float* mlp_base, const float* mlp_head, float* rgb,
                     fvc_diff_stats* stats);
int fvc_diff_backward(fvc_renderer* renderer, const float* d_rgb, float* g_pos_enc, float* g_mlp_base,
                      float* g_mlp_head);
/* The same for CachePerformanceRenderer::render's per-ray semantics
 * (cache_performance_renderer.cu:22-313, 789-908) against the renderer's
 * render cache: resampled samples take the network, cache samples the
 * trilinear lookup of the brick pool and the colour MLP.  The backward also
 * adds, through the adjoint of the lookup, the gradients of each voxel's 8
 * latents and density to g_pool, float [n_bricks_used][512][10] (per voxel
 * in the pool's order: 8 latents, density, isset; isset, a mask, gets
 * none). */
int fvc_diff_cached_forward(fvc_renderer* renderer, const fvc_raymarch_info* rm, const int32_t* pix, int32_t n_rays,
                            const float* pos_enc, const float* mlp_base, const float* mlp_head, float* rgb,
                            fvc_diff_stats* stats);
int fvc_diff_cached_backward(fvc_renderer* renderer, const float* d_rgb, float* g_pos_enc, float* g_mlp_base,
                             float* g_mlp_head, float* g_pool);

/* ---- profiling ---- */
/* Enable (1) or disable (0) per-kernel timing; enabling clears the table. */
int fvc_profile_enable(fvc_renderer* renderer, int enable);
int fvc_profile_get(fvc_renderer* renderer, fvc_kernel_stats* out, int32_t max_entries, int32_t* n_entries);

/* ---- inspection (bit-exact parity tests) ---- */
/* Float result of the last render before the 8-bit write: rgb + transmittance
 * per pixel, float[H][W][4] (BaseRayResult, renderer_common.h:74-78). */
int fvc_get_last_result(fvc_renderer* renderer, float* out_dev, int64_t n_floats);
/* Cache getters: which 0 = the render (front) cache, 1 = the fill target.
 * Reading the fill target while a background fill runs waits for that fill
 * and returns its result; the render cache can be read meanwhile. */
int fvc_cache_get_info(fvc_renderer* renderer, int which /*0 = render cache, 1 = fill target*/, fvc_cache_info* info);
/* Indirection table int32[bz][by][bx]: pool slot or -1 (cache.h:31-33). */
int fvc_cache_get_brick_table(fvc_renderer* renderer, int which, int32_t* out_host, int64_t n);
/* known_ranges_to float[H][W] (cache.h:36) */
int fvc_cache_get_known_ranges(fvc_renderer* renderer, int which, float* out_host, int64_t n);
/* Voxel records of the first n_bricks pool slots: per voxel
 * {half latents[8]; float density; float isset; float znext[2]} = 32 bytes,
 * voxel order z,y,x inside a brick (x fastest, as the reference's bricks;
 * the pool itself keeps z fastest); znext repeats density and isset of the
 * +z neighbour in the brick column (zeros at z = 7). */
int fvc_cache_get_pool(fvc_renderer* renderer, int which, void* out_host, int64_t n_bytes);

/* Ray marcher alone (sample_performance_kernel, performance_renderer.cu:15-69,
 * without the network): per pixel the number of occupied samples and the
 * first max_per_ray sample start times t0; host outputs. */
int fvc_debug_march(fvc_renderer* renderer, const fvc_raymarch_info* rm, int32_t max_per_ray,
                    int32_t* counts_host, float* t0_host);
/* The cache head's trilinear lookup alone (generateNetworkInput_cache_shfl_
 * perf_kernel_cg, cache_performance_renderer.cu:678-787) at n froxel points
 * (host float[n][3]) of cache `which`, with the renderer's cache settings:
 * host out float[n][9] = the head inputs as half values, [density row,
 * latents 0..7]. */
int fvc_debug_cache_lookup(fvc_renderer* renderer, int which, const float* froxel_host, int64_t n, float* out_host);
/* CachePerformanceRenderer::render with inspection records (host outputs):
 * codes int8[H*W][max_codes] = the class of each occupied sample of the
 * pixel's ray in march order as sample_cache_performance_kernel decides it
 * (cache_performance_renderer.cu:96-160; 0 skip, 1 resample, 2 cache, -1
 * past the samples classified), classified int32[H*W] = occupied samples
 * classified, composited int32[H*W][2] = resampled and cached samples
 * composited up to and including the one that ends the ray. */
int fvc_debug_render_cached(fvc_renderer* renderer, const fvc_raymarch_info* rm, int32_t max_codes, int8_t* codes_host,
                            int32_t* classified_host, int32_t* composited_host, fvc_render_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* FVC_H */
