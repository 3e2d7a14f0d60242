/* cindergs.h — C ABI of the MI355X (gfx950) CinderGS render / cage-deformation
 * hot path.  Plain pointers and sizes only; every entry point returns a
 * cgs_status (0 = OK) and leaves a message for cgs_last_error().
 *
 * Each entry point names the reference interface (Rust, CinderGS 2.0.0) it
 * replaces.  Paths are relative to the reference repository root.
 *
 * Threading: a cgs_context is bound to one device and one HIP stream and is not
 * thread-safe (the reference's GpuPipelineState methods all take &mut self).
 * Calls are asynchronous on the context stream except the readbacks,
 * cgs_execute_frame (which waits for its counters) and cgs_synchronize.
 */
#ifndef CINDERGS_H
#define CINDERGS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* GpuError {Runtime, NoAdapter, UnsupportedTarget, Execution}
 * (src/render/gpu.rs:14-30) and GpuBridgeError {Readback, ShortReadback}
 * (src/render/gpu_bridge.rs:28-48) as status codes. */
typedef enum cgs_status {
  CGS_OK = 0,
  CGS_ERR_RUNTIME = 1,
  CGS_ERR_NO_DEVICE = 2,
  CGS_ERR_UNSUPPORTED = 3,
  CGS_ERR_EXECUTION = 4,
  CGS_ERR_SHORT_READBACK = 5,
  CGS_ERR_INVALID_ARG = 6,
  CGS_ERR_FORMAT = 7 /* malformed or unsupported input file (PlyError, ply_loader.rs:157-222) */
} cgs_status;

typedef struct cgs_context cgs_context;

/* CameraUniforms, 96 bytes (src/render/camera.rs:9-36). view_matrix is the
 * column-major world-to-camera matrix. */
typedef struct cgs_camera_uniforms {
  float view_matrix[16];
  float focal_x, focal_y;
  float principal_x, principal_y;
  float viewport_width, viewport_height;
  float near_plane;
  float pad;
} cgs_camera_uniforms;

/* CompositorUniforms, 48 bytes (src/render/rasterizer/mod.rs:16-37). */
typedef struct cgs_compositor_uniforms {
  uint32_t screen_width, screen_height;
  uint32_t tiles_x, tiles_y;
  float background_color[4];
  float camera_position[3];
  uint32_t sh_degree;
} cgs_compositor_uniforms;

/* The scalar arguments of GpuPipelineState::execute_frame_gpu
 * (src/render/gpu_state.rs:613-626). */
typedef struct cgs_frame_params {
  float strain_split_threshold;
  float mip_filter_3d_var;
  float min_opacity;
  float min_radius_px;
  uint32_t profile_stages;
  uint32_t force_stage3;
} cgs_frame_params;

/* GpuStageTimings (src/render/gpu_state.rs:513-534) plus the number of Stage 3
 * re-runs caused by growing the tile-entry arena. */
typedef struct cgs_stage_timings {
  double stage1_deform_ms;
  double stage2_project_ms;
  double stage3_bin_sort_ms;
  double stage4_composite_ms;
  uint32_t stage1_skipped;
  uint32_t stage3_cached;
  uint32_t used_hardware_timestamps;
  uint32_t stage3_reruns;
  uint64_t active_gaussians;
  uint64_t surviving_gaussians;
  uint64_t total_tile_entries;
} cgs_stage_timings;

/* ---- lifetime (GpuCompute::new + GpuPipelineState::new, gpu.rs:63-100,
 *      gpu_state.rs:553-562) ---- */
int cgs_create(int device, cgs_context** out);
void cgs_destroy(cgs_context* ctx);
const char* cgs_last_error(void);
int cgs_device_name(cgs_context* ctx, char* buf, size_t len);
int cgs_synchronize(cgs_context* ctx);

/* GpuPipelineState::invalidate / invalidate_stage3 (gpu_state.rs:565-573);
 * both also drop the context's sort cache (see cgs_set_sort_hysteresis) */
int cgs_invalidate(cgs_context* ctx);
int cgs_invalidate_stage3(cgs_context* ctx);

/* ---- temporal hysteresis for Stage 3: SortCache (src/render/pipeline.rs:
 *      40-63, 66-193) ----
 * A host object with the reference's methods.  It keeps the view matrix, the
 * flattened cage vertices (xyz per vertex), the frame size and, for the
 * *_for_frame forms, the splat count and the bits of focal_x, focal_y, the
 * split threshold and the mip variance of the last committed binning.  A
 * later frame takes it back when the view moved by at most 1e-5 (largest
 * absolute element difference), no cage coordinate moved by more than 1e-5
 * and the rest is equal.  The binning (TileBinningOutputs) stays with the
 * caller; the object records whether one is held: take_* clears that, a
 * commit sets it.  No device work. */
typedef struct cgs_sort_cache cgs_sort_cache;
typedef struct cgs_sort_frame {
  uint64_t splat_count;
  float focal_x, focal_y, split_threshold, mip_var;
} cgs_sort_frame;
int cgs_sort_cache_create(cgs_sort_cache** out);   /* SortCache::default */
void cgs_sort_cache_destroy(cgs_sort_cache* cache);
int cgs_sort_cache_can_reuse(const cgs_sort_cache* cache, const float view[16],
                             const float* cage_xyz, uint64_t num_cage_vertices, uint32_t width,
                             uint32_t height, uint32_t* reusable);              /* 85-101 */
int cgs_sort_cache_take_if_reusable(cgs_sort_cache* cache, const float view[16],
                                    const float* cage_xyz, uint64_t num_cage_vertices,
                                    uint32_t width, uint32_t height, uint32_t* reused); /* 103-117 */
int cgs_sort_cache_take_if_reusable_for_frame(cgs_sort_cache* cache, const float view[16],
                                              const float* cage_xyz, uint64_t num_cage_vertices,
                                              uint32_t width, uint32_t height,
                                              const cgs_sort_frame* frame,
                                              uint32_t* reused);                /* 119-145 */
int cgs_sort_cache_commit(cgs_sort_cache* cache, const float view[16], const float* cage_xyz,
                          uint64_t num_cage_vertices, uint32_t width, uint32_t height); /* 147-161 */
int cgs_sort_cache_commit_for_frame(cgs_sort_cache* cache, const float view[16],
                                    const float* cage_xyz, uint64_t num_cage_vertices,
                                    uint32_t width, uint32_t height,
                                    const cgs_sort_frame* frame);               /* 163-184 */
int cgs_sort_cache_invalidate(cgs_sort_cache* cache);                           /* 186-192 */
int cgs_sort_cache_reuse_count(const cgs_sort_cache* cache, uint64_t* count);   /* reuse_count */

/* The same hysteresis inside cgs_execute_frame, as the reference's CPU
 * fallback frame applies it (execute_cpu_fallback_frame, pipeline.rs:539-600):
 * with enable = 1 each frame asks the context's sort cache
 * (take_if_reusable_for_frame with the frame's view matrix, cage vertices,
 * size, splat count, focal lengths, split threshold and mip variance); on a
 * hit Stages 1-2 run as usual, Stage 3 is skipped and Stage 4 composites the
 * new projection in the previous frame's tile order (stage3_cached = 1); a
 * splat culled in the new frame but present in that order contributes
 * nothing.  Every frame then commits.  enable = 0 (the default) restores the
 * exact Stage 3 cache key of gpu_state.rs:454-499.  Either call starts a
 * fresh cache (reuse count 0). */
int cgs_set_sort_hysteresis(cgs_context* ctx, uint32_t enable);
int cgs_sort_hysteresis_reuse_count(cgs_context* ctx, uint64_t* count);

/* GpuSceneState::sync_scene_and_cages, scene part (gpu_state.rs:182-213):
 * GaussianSceneSoa (src/scene/soa.rs:30-45) — positions[3n] xyz, covariances
 * [6n] (xx,xy,xz,yy,yz,zz), opacities[n] (activated), sh[48n] per splat
 * [dc r,g,b, rest_r 15, rest_g 15, rest_b 15].  Uploads only when the
 * fingerprint (count, pointers, 64-sample FNV-1a, gpu_state.rs:65-79) changes. */
int cgs_sync_scene(cgs_context* ctx, const float* positions, const float* covariances,
                   const float* opacities, const float* sh_coeffs, uint64_t count);

/* QuantizedGaussianSceneSoa (src/scene/soa.rs:429-500): covariances as 3 u32
 * (f16 pairs, low half first, pack_f16_pair soa.rs:348-352), SH as
 * ((bases*3+1)/2) u32 words per splat, channel-major and degree-trimmed.
 * Rendered by the f16 projection (project.rs:1344-1679). */
int cgs_sync_scene_f16(cgs_context* ctx, const float* positions, const uint32_t* cov_packed,
                       const float* opacities, const uint32_t* sh_packed, uint32_t sh_degree,
                       uint64_t count);

/* GpuSceneState::sync_scene_and_cages, cage part (gpu_state.rs:215-343), with
 * the CageScene already flattened by the caller: tet_elements[4*num_tets]
 * (global vertex indices), inv_dm[9*num_tets] (TetPrecomputed, col-major),
 * vertices[3*num_vertices], bind_tet[count] (global tet index, 0xFFFFFFFF =
 * unbound), bind_weights[4*count].  num_tets == 0 detaches the cages.
 * Topology is re-uploaded when its fingerprint changes; vertices when they
 * differ from the last call (marks Stage 1 dirty). */
int cgs_sync_cages(cgs_context* ctx, const uint32_t* tet_elements, const float* inv_dm,
                   uint32_t num_tets, const float* vertices, uint32_t num_vertices,
                   const uint32_t* bind_tet, const float* bind_weights, uint64_t count);

/* GpuPipelineState::execute_frame_gpu (gpu_state.rs:613-1243): Stages 1-4 into
 * the device framebuffer (RGBA8 words, rows padded to ceil(W/64)*64). */
int cgs_execute_frame(cgs_context* ctx, const cgs_camera_uniforms* camera,
                      const cgs_compositor_uniforms* compositor,
                      const cgs_frame_params* params, cgs_stage_timings* timings);

/* readback_framebuffer_rgba8 (gpu_state.rs:1296-1328): tightly packed
 * width*height*4 bytes. */
int cgs_readback_rgba8(cgs_context* ctx, uint8_t* dst, uint32_t width, uint32_t height);

/* Replaces copy_framebuffer_to_texture (gpu_state.rs:1247-1292): the device
 * address of the framebuffer and its row stride in u32 words. */
int cgs_framebuffer_device_ptr(cgs_context* ctx, void** ptr, uint32_t* padded_row_u32);

/* Per-launch device times of the last cgs_execute_frame run with
 * profile_stages >= 2 (the reference's measure_stage, gpu_state.rs:575-608,
 * at kernel granularity), or of the last cgs_mvc_deform_scene,
 * cgs_activate_raw (and cgs_ply_load), cgs_quantize_scene, cgs_export_raw or
 * cgs_export_deformed_scene call: names[32*i] (NUL-terminated) and ms[i]. */
int cgs_kernel_times(cgs_context* ctx, char* names, double* ms, uint32_t cap, uint32_t* count);

/* ---- launch tuning ----
 * Launch shapes that leave every output bit-identical (keys, ranges, RGBA8):
 * radix partition size and digit width of the Stage 3 sort, the tile sort,
 * splats staged per compositor step, and the heaviest-tiles-first compositor
 * order.  The reference fixes its shapes in code (radix_sort.rs:292-368:
 * 4-bit digits, 256-key blocks; compositor.rs:721-883: batches of 256;
 * TILE_SIZE = 16, tiles/mod.rs:10); here they are per-context settings that
 * cgs_autotune picks by timing.  The tile size is one of them but changes the
 * frame: with tiles of w x h pixels every output is the reference's algorithm
 * run at that tile size (tile rectangles, tile count, and with it the split of
 * the 32-bit key between tile and depth bits).  Defaults:
 * {16384, 8, 64, 1, 0, 16, 16}; the radix fields matter only with
 * tile_sort = 0. */
typedef struct cgs_tuning {
  uint32_t radix_partition_keys; /* 16384, 12288, 10240, 8192 or 4096 keys per
                                    sort partition (1024 threads x 16 or 12
                                    keys; 640, 512 or 256 threads x 16) */
  uint32_t radix_digit_bits;     /* 8 or 4 (32-bit keys; the 64-bit path uses 8) */
  uint32_t composite_batch;      /* 64, 128, 256 or 512 splats per compositor
                                    step (8 x 8 tiles: 64; 512: 16 x 16 tiles).
                                    64 on larger tiles: each 8 x 8 quadrant's
                                    wave walks the tile's list on its own in
                                    batches of 64, with no block barrier; the
                                    others stage each batch once per block */
  uint32_t tile_order;           /* 1: heaviest tiles composited first */
  uint32_t tile_sort;            /* 1: entries written into per-tile buckets and each
                                    tile sorted by (key, splat id) in LDS; 0: one
                                    stable LSD radix sort of all entries (the
                                    reference's method).  Frames of more than 32768
                                    tiles always take the radix sort. */
  uint32_t tile_w, tile_h;       /* tile size in pixels: 16 x 16 (the reference's),
                                    8 x 8, 16 x 32 or 32 x 16 */
} cgs_tuning;

int cgs_set_tuning(cgs_context* ctx, const cgs_tuning* tuning);
int cgs_get_tuning(cgs_context* ctx, cgs_tuning* tuning);

/* cgs_autotune flags: also sweep the other tile sizes (whose frames differ
 * from the current size's; each is checked against its own first frame) */
#define CGS_TUNE_TILE_SHAPES 1u

/* Times the cgs_tuning choices of the current tile size (all sizes with
 * CGS_TUNE_TILE_SHAPES) on `reps` frames of the given view (after one untimed
 * frame each), checks that each gives the same framebuffer and tile entries as
 * the current setting (as the first choice of its tile size), and keeps the
 * fastest.  ms[i] (optional, n_ms entries) gets the mean device time of choice
 * i in the order of cgs_tuning_candidate, or -1 for a choice not swept; *best
 * gets the chosen setting. */
int cgs_autotune(cgs_context* ctx, const cgs_camera_uniforms* camera,
                 const cgs_compositor_uniforms* compositor, const cgs_frame_params* params,
                 uint32_t reps, uint32_t flags, cgs_tuning* best, double* ms, uint32_t n_ms);
/* candidate i (< cgs_tuning_candidates()) of the autotuner's sweep */
uint32_t cgs_tuning_candidates(void);
int cgs_tuning_candidate(uint32_t i, cgs_tuning* out);

/* ---- batched multi-view rendering ----
 * The reference renders one view per GpuPipelineState::execute_frame_gpu call
 * (gpu_state.rs:613-1243) from a per-view loop in its callers
 * (benchmark.rs:231-366, headless.rs:140-210).  A cgs_batch renders many views
 * of several resident scenes (each a cgs_context after cgs_sync_scene /
 * cgs_sync_cages) in one call: Stage 1 once per scene, then Stages 2-4 of
 * every view from one captured HIP graph, replayed while the views and the
 * buffers are unchanged.  Each view's RGBA8 equals what cgs_execute_frame
 * gives for the same scene, camera and parameters.  A batch is bound to one
 * device and is not thread-safe; it must not run concurrently with frames of
 * the contexts it reads. */
typedef struct cgs_batch cgs_batch;

typedef struct cgs_batch_stats {
  uint32_t views;              /* views rendered */
  uint32_t lanes;              /* concurrent branches of the graph */
  uint32_t graph_built;        /* 1 when this call captured the graph */
  uint32_t reruns;             /* replays after a lane's tile-entry arena grew */
  uint64_t total_tile_entries; /* summed over the views */
  double gpu_ms;               /* device time of the graph replay(s) */
  uint32_t presort_views;      /* views sorted with the splats presorted by depth */
  uint32_t reserved;
} cgs_batch_stats;

/* lanes in [1, 16]: views run on that many concurrent graph branches, each
 * with its own frame arena. */
int cgs_batch_create(int device, uint32_t lanes, cgs_batch** out);
void cgs_batch_destroy(cgs_batch* batch);

/* View v renders scenes[view_scene[v]] with cameras[v] and compositors[v];
 * params are shared (profile_stages and force_stage3 are ignored: every view
 * runs Stage 3).  Waits for the views to finish.  With the LSD sort
 * (tile_sort 0), a view whose last replay in the same batch had more than 4
 * tile entries per splat slot is sorted with the splats presorted by depth
 * (its keys, ids, ranges and pixels are the same). */
int cgs_execute_views_batched(cgs_batch* batch, cgs_context* const* scenes, uint32_t n_scenes,
                              const uint32_t* view_scene, const cgs_camera_uniforms* cameras,
                              const cgs_compositor_uniforms* compositors, uint32_t n_views,
                              const cgs_frame_params* params, cgs_batch_stats* stats);

/* launch tuning of every lane (see cgs_tuning) */
int cgs_batch_set_tuning(cgs_batch* batch, const cgs_tuning* tuning);
/* 1 (default): Stage 2 of all views of a scene runs as one pass over the
 * scene (each splat's record read once, its SH once), writing every view's
 * projected records (13 words per splat slot per view); 0: one projection per
 * view in its lane.  The pixels are the same either way. */
int cgs_batch_set_shared_projection(cgs_batch* batch, uint32_t enable);

/* View v of the last batch: device framebuffer and row stride in u32 words,
 * tightly packed readback, and its tile-entry / survivor counts. */
int cgs_batch_framebuffer(cgs_batch* batch, uint32_t view, void** ptr, uint32_t* padded_row_u32);
int cgs_batch_readback_rgba8(cgs_batch* batch, uint32_t view, uint8_t* dst, uint32_t width,
                             uint32_t height);
int cgs_batch_view_counts(cgs_batch* batch, uint32_t view, uint32_t* tile_entries,
                          uint32_t* survivors);

/* ---- per-stage access (the stage wrappers of src/render/gpu_bridge.rs) ---- */

/* Stage 2 output in the reference layout: attrs[10*n] = [mu_x, mu_y, a, b, c,
 * alpha_2d, r, g, b, tz] and extents[n] (ext_x | ext_y << 16, 0 = culled)
 * (project.rs:1271-1284); n = active splats of the last frame. */
int cgs_read_projected(cgs_context* ctx, float* attrs10, uint32_t* extents, uint64_t n);

/* Stage 3 output of the last frame: per-splat tile counts and exclusive
 * offsets (needs cgs_set_capture(ctx, 1) before the frame), sorted keys
 * (packed 32-bit keys, or tile ids on the 64-bit path), sorted splat ids and
 * tile ranges[2*tiles].  Any pointer may be null.  *n_entries gets E. */
int cgs_read_stage3(cgs_context* ctx, uint32_t* counts, uint32_t* offsets, uint32_t* sorted_keys,
                    uint32_t* sorted_ids, uint32_t* ranges, uint64_t* n_entries);
int cgs_set_capture(cgs_context* ctx, uint32_t enable);

/* Stage 1 output of the last deformation (deform_and_split_compact_kernel,
 * deformation.rs:872-1211): positions[3*m], covariances[6*m], opacities[m],
 * split_meta[2*splits] (parent, tet) with m = count + splits. */
int cgs_read_deformed(cgs_context* ctx, float* positions, float* covariances, float* opacities,
                      uint32_t* split_meta, uint32_t* split_count);
/* compute_tet_transforms_kernel outputs (deformation.rs:682-865). */
int cgs_read_tet_transforms(cgs_context* ctx, float* jacobians, float* rotations,
                            float* max_stretch);

/* The deformed scene of the last cgs_execute_frame with cages as a
 * GaussianSceneSoa for export, as the CPU path's deform_scene_cpu writes it
 * (src/deformation/cpu_reference.rs:97-137, 166-182): the Stage 1 positions,
 * covariances and opacities the renderer used, and each splat's SH (a split
 * child's: its parent's) rotated by the polar rotation of its tet with
 * ShRotator (src/render/rasterizer/sh_rotation.rs:396-446), whose matrices
 * are built once per tet; unbound splats keep their SH.  Bands above
 * sh_degree are copied as they are.  Outputs in device memory with room for
 * `capacity` records (d_sh 16-byte aligned); *count receives the number of
 * records (rest splats + split children).  With all four outputs null the
 * call only sets *count.  The f16 storage variant has no cage deformation:
 * CGS_ERR_UNSUPPORTED. */
int cgs_export_deformed_scene(cgs_context* ctx, uint32_t sh_degree, float* d_pos, float* d_cov,
                              float* d_opa, float* d_sh, uint64_t capacity, uint64_t* count);

/* TetMesh::bind_gaussians / bind_gaussians_with_bounds (src/deformation/
 * mod.rs:263-323, bind_point 267-310) on the GPU: for each of the n positions
 * (xyz), the first tetrahedron in index order whose barycentric penalty
 * sum(max(-w_k, 0)) is <= 1e-4, else the one of smallest penalty; with bounds
 * (min xyz, max xyz; NULL for none) a position outside them, or whose best
 * penalty exceeds 0.15, is unbound (tet 0xFFFFFFFF, weights 0).  rest_vertices
 * [3*num_vertices], elements [4*num_tets], inv_dm [9*num_tets] column-major
 * (TetMesh::new, 202-227); out_tet [n], out_weights [4*n] (w0..w3).  Host
 * arrays; returns when the results are written. */
int cgs_bind_tets(cgs_context* ctx, const float* rest_vertices, uint32_t num_vertices,
                  const uint32_t* elements, const float* inv_dm, uint32_t num_tets,
                  const float* positions, uint64_t n, const float* bounds, uint32_t* out_tet,
                  float* out_weights);

/* PolyCageMesh::bind_gaussians (src/deformation/mvc.rs:498-541): exact 3D mean
 * value coordinates (compute_mvc_weights, 366-471) of n points against a
 * closed outward-wound triangle cage, their central-difference gradients
 * (step clamp(1e-3 * bbox diagonal, 1e-5, 1e-2), 485-495) and the inside test
 * |winding number| >= 0.5 (326-363); rows of points outside are zero.  Cage
 * arrays on the host (vertices [3*nv], triangles [3*nt], nv*24 + nt*12 <= 64 KiB);
 * d_points [3n], d_weights [n*nv], d_gradients [n*nv*3] (or NULL), d_inside
 * [n] u32 (or NULL) in device memory (cgs_malloc). */
int cgs_mvc_bind_exact(cgs_context* ctx, const float* cage_vertices, uint32_t num_vertices,
                       const uint32_t* triangles, uint32_t num_triangles, const float* d_points,
                       uint64_t n, float* d_weights, float* d_gradients, uint32_t* d_inside);
/* PolyCageMesh::winding_number (mvc.rs:326-347) of n device points into d_out [n]. */
int cgs_mvc_winding(cgs_context* ctx, const float* cage_vertices, uint32_t num_vertices,
                    const uint32_t* triangles, uint32_t num_triangles, const float* d_points,
                    uint64_t n, float* d_out);
/* evaluate_position (mvc.rs:473-481) and MvcBindings::local_jacobian (66-78):
 * x' = sum_j w_j v'_j and F = sum_j v'_j grad(w_j)^T (column-major) for n
 * points of a binding table; a point whose d_inside flag is 0 keeps its rest
 * position (and F = I).  deformed_vertices on the host; the rest in device
 * memory; d_gradients / d_out_jacobians may be NULL together. */
int cgs_mvc_deform(cgs_context* ctx, const float* d_weights, const float* d_gradients,
                   const uint32_t* d_inside, const float* d_rest_positions,
                   const float* deformed_vertices, uint32_t num_vertices, uint64_t n,
                   float* d_out_positions, float* d_out_jacobians);

/* PolyCageMesh::deform_scene_cpu (mvc.rs:564-749) on the GPU: for each of
 * the n splats of a binding table, x' = sum_j w_j v'_j and F = sum_j v'_j
 * grad(w_j)^T; a bound splat gets Sigma' = F Sigma F^T and its SH (degree
 * sh_degree, record layout of cgs_sync_scene) rotated by the polar rotation
 * of F (higham_polar_decomposition, polar.rs:13-42; ShRotator,
 * sh_rotation.rs:372-456), keeping its rest opacity; an unbound one keeps its
 * rest record.  With split_threshold > 0 a bound splat whose largest stretch
 * of F exceeds it is split along the major axis of Sigma'
 * (split_gaussian_along_major_axis, math/covariance.rs:343-381): the first
 * child replaces it and the second is appended after the n records in parent
 * order.  The 3D mip filter (288-304) applies to every unsplit record and to
 * both children.  Inputs and outputs in device memory (outputs hold up to 2n
 * records; d_out_sh 16-byte aligned, as cgs_malloc returns), deformed_vertices
 * [3*num_vertices] on the host; the weight gradients are required; *out_count
 * receives n + the number of splits. */
int cgs_mvc_deform_scene(cgs_context* ctx, const float* d_weights, const float* d_gradients,
                         const uint32_t* d_inside, const float* deformed_vertices,
                         uint32_t num_vertices, uint64_t n, const float* d_rest_pos,
                         const float* d_rest_cov, const float* d_rest_opa, const float* d_rest_sh,
                         uint32_t sh_degree, float split_threshold, float mip_filter_3d_var,
                         float* d_out_pos, float* d_out_cov, float* d_out_opa, float* d_out_sh,
                         uint64_t* out_count);

/* ---- scene ingest and quantised storage ---------------------------------------
 * PlyLoader (src/scene/ply_loader.rs:601-716), GaussianSceneSoa::push_raw /
 * from_raw_vertices_par (soa.rs:113-189), QuantizedGaussianSceneSoa::
 * from_soa_with_degree (454-500) and to_raw_vertices (306-328). */

/* PlyFormat (ply_loader.rs:12-16): only the binary ones load */
#define CGS_PLY_BINARY_LE 1u
#define CGS_PLY_BINARY_BE 2u
/* PlyError variants (ply_loader.rs:157-222), in cgs_ply_layout.error */
#define CGS_PLY_OK 0u
#define CGS_PLY_MISSING_END_HEADER 1u
#define CGS_PLY_INVALID_MAGIC 2u
#define CGS_PLY_INVALID_HEADER_UTF8 3u
#define CGS_PLY_UNSUPPORTED_FORMAT 4u
#define CGS_PLY_MISSING_FORMAT 5u
#define CGS_PLY_MISSING_VERTEX_ELEMENT 6u
#define CGS_PLY_UNSUPPORTED_LAYOUT 7u
#define CGS_PLY_UNSUPPORTED_LIST 8u
#define CGS_PLY_UNSUPPORTED_ELEMENT 9u  /* error_detail[0] = record count */
#define CGS_PLY_MISSING_PROPERTY 10u
#define CGS_PLY_INVALID_PROPERTY_TYPE 11u
#define CGS_PLY_UNSUPPORTED_SH_LAYOUT 12u /* error_detail[0] = f_rest count */
#define CGS_PLY_TRUNCATED_PAYLOAD 13u    /* error_detail = needed, available bytes */

/* PlyHeader (ply_loader.rs:125-140) as a decode table: for each of the 62
 * floats of a RawPlyVertex (raw_vertex.rs:16-37: position, normal, f_dc,
 * f_rest 15 per channel, opacity, log-scale, quaternion w x y z), the byte
 * offset in a vertex record and the scalar type of the property that fills it
 * (slot_type 0: absent, the float stays 0; 1-8: int8, uint8, int16, uint16,
 * int32, uint32, float32, float64).  A compact SH layout (9 or 24 f_rest) is
 * scattered per channel as parse_vertex_payload does (540-598). */
typedef struct cgs_ply_layout {
  uint64_t vertex_count;
  uint64_t payload_offset;   /* bytes before the vertex payload */
  uint32_t vertex_stride;    /* bytes per vertex record */
  uint32_t format;           /* CGS_PLY_BINARY_LE / _BE */
  uint32_t sh_degree;        /* 0 / 9 / 24 / 45 f_rest give 0..3 */
  uint32_t canonical;        /* the 62 float32 properties in RawPlyVertex order, LE */
  uint32_t num_properties;
  uint32_t error;            /* CGS_PLY_* of the last failure, 0 on success */
  uint64_t error_detail[2];
  uint32_t slot_offset[62];
  uint8_t slot_type[62];
  uint8_t pad[2];
} cgs_ply_layout;

/* parse_header (ply_loader.rs:290-468) of a PLY file in host memory; on a
 * malformed header returns CGS_ERR_FORMAT with layout->error set and the
 * reference's message in cgs_last_error(). */
int cgs_ply_parse_header(const uint8_t* bytes, uint64_t len, cgs_ply_layout* layout);

/* PlyLoader::load_from_bytes (622-654) into device memory: the header is
 * parsed on the host (layout may be NULL, or the result of
 * cgs_ply_parse_header on the same bytes), the vertex payload is copied to
 * the device as is, and one kernel decodes and activates every record:
 * d_pos [3n], d_cov [6n], d_opa [n], d_sh [48n] (n = vertex_count).  A payload
 * shorter than n * stride gives CGS_PLY_TRUNCATED_PAYLOAD. */
int cgs_ply_load(cgs_context* ctx, const uint8_t* bytes, uint64_t len,
                 const cgs_ply_layout* layout, float* d_pos, float* d_cov, float* d_opa,
                 float* d_sh);

/* from_raw_vertices_par (soa.rs:137-189) on a payload already in device memory
 * (d_records: n records of layout->vertex_stride bytes, 4-byte aligned). */
int cgs_activate_raw(cgs_context* ctx, const void* d_records, const cgs_ply_layout* layout,
                     uint64_t n, float* d_pos, float* d_cov, float* d_opa, float* d_sh);

/* QuantizedGaussianSceneSoa::from_soa_with_degree (soa.rs:454-500) on the
 * device: per splat three f16 pairs of the covariance (pack_f16_pair, round
 * to nearest even), the octahedral rotation of its principal axes
 * (decompose_covariance_3d, encode_octahedral_rotation, 369-402), and the
 * degree-trimmed f16 SH words ((degree + 1)^2 * 3 halves, rounded up to
 * whole words; 2 / 6 / 14 / 24 per splat for degree 0..3).  Positions and
 * opacities are unchanged and stay with the caller. */
int cgs_quantize_scene(cgs_context* ctx, const float* d_cov, const float* d_sh, uint64_t n,
                       uint32_t sh_degree, uint32_t* d_cov_f16, uint32_t* d_rot_oct,
                       uint32_t* d_sh_f16);

/* GaussianSceneSoa::to_raw_vertices (soa.rs:306-328) on the device: n
 * RawPlyVertex records (248 bytes: logit opacity, log-scales and the WXYZ
 * quaternion of decompose_covariance_3d, zero normals) into d_records, the
 * payload of PlyLoader::write_binary_ply_bytes (674-705). */
int cgs_export_raw(cgs_context* ctx, const float* d_pos, const float* d_cov,
                   const float* d_opa, const float* d_sh, uint64_t n, void* d_records);

/* ---- MVC surrogate network (no counterpart in the reference) ------------------
 * A coordinate MLP that stands in for exact MVC when binding many points to one
 * cage: x -> [sin 2pi Bx, cos 2pi Bx] (B 32 x 3, N(0, sigma^2)) -> 4 x
 * (Linear 256 + ReLU) -> Linear n_out -> w = z + (1 - sum z) / n_out.  BF16 MFMA
 * GEMMs with FP32 accumulation, FP32 master weights, Adam.  Flat parameter
 * layout: layer l (0..4) is a row-major [rows_l x (cols_l + 32)] block at
 * offsets_l, column cols_l holding the bias (see cgs_mlp_layout); rows =
 * {256, 256, 256, 256, n_out rounded up to 32}, cols = {64, 256, 256, 256, 256}. */
typedef struct cgs_mlp cgs_mlp;
int cgs_mlp_create(cgs_context* ctx, uint32_t n_out, float fourier_sigma, uint64_t seed,
                   cgs_mlp** out);
void cgs_mlp_destroy(cgs_mlp* net);
uint64_t cgs_mlp_param_count(const cgs_mlp* net);
int cgs_mlp_layout(const cgs_mlp* net, uint32_t* rows, uint32_t* cols, uint64_t* offsets);
/* flat parameters [param_count] and the Fourier matrix [32 x 3] (either NULL) */
int cgs_mlp_get_params(cgs_mlp* net, float* params, float* fourier);
int cgs_mlp_set_params(cgs_mlp* net, const float* params, const float* fourier);
/* steps of Adam (lr, betas 0.9/0.999, eps 1e-8) on batches of `batch` (a
 * multiple of 64) drawn from the n device pairs (d_x [3n], d_targets
 * [n*n_out]); losses [steps] receives each step's mean squared error (or NULL) */
int cgs_mlp_train(cgs_mlp* net, const float* d_x, const float* d_targets, uint64_t n,
                  uint32_t steps, uint32_t batch, float lr, float* losses);
/* loss and parameter gradient (flat layout, host) of the first n pairs (n a
 * multiple of 64), without an update */
int cgs_mlp_gradient(cgs_mlp* net, const float* d_x, const float* d_targets, uint32_t n,
                     float* grad, float* loss);
/* weights [n*n_out] and, if d_grad, their input derivatives [n*n_out*3]; with a
 * near field set, the raw outputs (the far targets of the network's kind,
 * as cgs_mvc_far_targets scales them) and their derivatives */
int cgs_mlp_infer(cgs_mlp* net, const float* d_x, uint64_t n, float* d_w, float* d_grad);
/* Near/far split (r0 < r1, both 0 to switch it off).  MVC's sums over the
 * cage triangles split with phi_t(x) = smoothstep((r1 - |x - c_t|) / (r1 - r0)),
 * c_t the triangle's centroid: the near sums u_near = sum_t phi_t mu_t are
 * evaluated exactly at bind time, the network learns the far share
 * q = (u - u_near) / sum u (scaled by n_out; targets from cgs_mvc_far_share,
 * plain squared error, no affine normalisation) and the binding's weights are
 * w = (1 - sum q) u_near / sum u_near + q (|sum u_near| floored at a small tau,
 * sign kept), projected onto linear
 * precision. */
int cgs_mlp_set_near_field(cgs_mlp* net, float r0, float r1);
/* What a near/far network learns (no counterpart in the reference, which has
 * no learned binding).  CGS_FAR_SHARE (the default): the far share q above.
 * CGS_FAR_SUMS: the far sums u_far = u - u_near themselves, times
 * num_vertices x diagonal / (4 pi) (targets from cgs_mvc_far_targets), and
 * the binding's weights are w = (u_near + u_far) / sum (u_near + u_far),
 * projected onto linear precision: the network's error is then divided by
 * the whole MVC sum, which is large next to the cage, where the far share
 * changes fastest. */
#define CGS_FAR_SHARE 0u
#define CGS_FAR_SUMS 1u
int cgs_mlp_set_far_kind(cgs_mlp* net, uint32_t kind);
/* Training targets for a near/far network: q x num_vertices per point,
 * d_out [n x num_vertices] (zero rows for points on a vertex or a face). */
int cgs_mvc_far_share(cgs_context* ctx, const float* cage_vertices, uint32_t num_vertices,
                      const uint32_t* triangles, uint32_t num_triangles, float near_r0,
                      float near_r1, const float* d_points, uint64_t n, float* d_out);
/* Training targets of either kind (cgs_mlp_set_far_kind): kind CGS_FAR_SHARE
 * is cgs_mvc_far_share; CGS_FAR_SUMS writes u_far x num_vertices x diagonal
 * / (4 pi), diagonal the cage's bounding-box diagonal. */
int cgs_mvc_far_targets(cgs_context* ctx, const float* cage_vertices, uint32_t num_vertices,
                        const uint32_t* triangles, uint32_t num_triangles, float near_r0,
                        float near_r1, const float* d_points, uint64_t n, uint32_t kind,
                        float* d_out);
/* The near/far combination of cgs_mlp_bind with the far shares given by the
 * caller: d_q [n x num_vertices] (as cgs_mvc_far_share writes them), d_dq
 * [3n x num_vertices] (row 3p + axis: their derivatives) or NULL.  Outputs as
 * cgs_mvc_bind_exact plus d_residual [n]: |sum w v - x| before the
 * linear-precision projection (+inf where the near sums meet a vertex or a
 * face).  Exact far shares give exact MVC weights. */
int cgs_mvc_nearfar_bind(cgs_context* ctx, const float* cage_vertices, uint32_t num_vertices,
                         const uint32_t* triangles, uint32_t num_triangles, float near_r0,
                         float near_r1, const float* d_points, uint64_t n, const float* d_q,
                         const float* d_dq, float* d_weights, float* d_gradients,
                         uint32_t* d_inside, float* d_residual);
/* The same with far targets of either kind (as cgs_mvc_far_targets writes
 * them for `kind`). */
int cgs_mvc_nearfar_bind_kind(cgs_context* ctx, const float* cage_vertices,
                              uint32_t num_vertices, const uint32_t* triangles,
                              uint32_t num_triangles, float near_r0, float near_r1,
                              const float* d_points, uint64_t n, const float* d_q,
                              const float* d_dq, uint32_t kind, float* d_weights,
                              float* d_gradients, uint32_t* d_inside, float* d_residual);
/* The surrogate for PolyCageMesh::bind_gaussians (mvc.rs:498-541): same outputs
 * as cgs_mvc_bind_exact (zero rows outside), the network's weights and input
 * derivatives (with a near field: combined with the exact near sums) for
 * inside points whose residual |sum_j w_j v_j - x| is at most
 * residual_tol x (cage bounding-box diagonal), exact MVC for the others;
 * fallback_count receives how many points took the exact path. */
int cgs_mlp_bind(cgs_context* ctx, cgs_mlp* net, const float* cage_vertices,
                 uint32_t num_vertices, const uint32_t* triangles, uint32_t num_triangles,
                 const float* d_points, uint64_t n, float residual_tol, float* d_weights,
                 float* d_gradients, uint32_t* d_inside, uint64_t* fallback_count);

/* Inject a projected splat set (reference layout, as cgs_read_projected) and
 * run Stage 3 / Stage 4 on it (the split of execute_frame_gpu used by
 * test_gpu_stage3_binning_and_ranges_match_cpu, gpu_parity_tests.rs:664-917). */
int cgs_load_projected(cgs_context* ctx, const float* attrs10, const uint32_t* extents,
                       uint64_t n);
int cgs_run_stage3(cgs_context* ctx, const cgs_camera_uniforms* camera,
                   const cgs_compositor_uniforms* compositor, uint64_t* n_entries);
int cgs_run_stage4(cgs_context* ctx, const cgs_compositor_uniforms* compositor);

/* ---- device-buffer primitives ---- */
int cgs_malloc(cgs_context* ctx, uint64_t bytes, void** dptr);
int cgs_free(cgs_context* ctx, void* dptr);
int cgs_memcpy_h2d(cgs_context* ctx, void* dst, const void* src, uint64_t bytes);
int cgs_memcpy_d2h(cgs_context* ctx, void* dst, const void* src, uint64_t bytes);
int cgs_memcpy_d2d(cgs_context* ctx, void* dst, const void* src, uint64_t bytes);

/* gpu_exclusive_scan_total (src/render/kernels/scan.rs:312-333). */
int cgs_exclusive_scan_u32(cgs_context* ctx, const uint32_t* d_in, uint32_t* d_out, uint64_t n,
                           uint32_t* total);
/* No counterpart in the reference.  cgs_exclusive_scan_u32 for callers that
 * cannot allocate device memory: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  Same outputs (d_out and the total).
 * Its scratch (the total, the partition ticket and one 8-byte look-back status
 * word per 4096 elements: 512 KiB) is a fixed array in the library's device
 * image, one copy per GPU, held until the process ends; no call allocates.
 * n <= CGS_SCAN_NOALLOC_MAX_N, else CGS_ERR_INVALID_ARG.  With total NULL the
 * call is asynchronous on the context stream; otherwise it waits and reads the
 * total.  Every context on one GPU shares that GPU's scratch: calls that
 * overlap on one GPU (other streams, other threads) race. */
#define CGS_SCAN_NOALLOC_MAX_N (1u << 28)
int cgs_exclusive_scan_u32_noalloc(cgs_context* ctx, const uint32_t* d_in, uint32_t* d_out,
                                   uint64_t n, uint32_t* total);
/* gpu_radix_sort_u32_pairs (radix_sort.rs:292-368): stable sort of (key, value)
 * on the low sort_bits key bits, in place. */
int cgs_radix_sort_u32_pairs(cgs_context* ctx, uint32_t* d_keys, uint32_t* d_vals, uint64_t n,
                             uint32_t sort_bits);
/* gpu_radix_sort_u64_pairs (radix_sort.rs:377-477): stable sort of entries by
 * (tile id on tile_bits, depth bits).  d_tile_ids and d_splat_ids receive the
 * sorted order (what u64_sorted_outputs returns, radix_sort.rs:270-286);
 * d_depths is left sorted by depth alone (it serves as the first key). */
int cgs_radix_sort_u64_pairs(cgs_context* ctx, uint32_t* d_tile_ids, uint32_t* d_depths,
                             uint32_t* d_splat_ids, uint64_t n, uint32_t tile_bits);

#ifdef __cplusplus
}
#endif

#endif /* CINDERGS_H */
