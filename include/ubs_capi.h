/*
 * ubs_capi.h -- C ABI of libubs_hip.so, the gfx950 (MI355X) implementation
 * of the differentiable Beta-splat tile rasteriser that universal-beta-
 * splatting reaches through its gsplat fork.
 *
 * Conventions
 *   - Every pointer is a device pointer to contiguous row-major data, fp32
 *     unless the type says otherwise.  Shapes are given in the comments.
 *   - The caller allocates every output and workspace; no entry point
 *     allocates device memory.  Gradient outputs are zeroed by the callee.
 *   - `stream` is a hipStream_t (NULL = default stream).  Work is enqueued
 *     asynchronously; nothing synchronises with the host.
 *   - Return value: UBS_OK or an error code; ubs_last_error() gives the
 *     message of the calling thread's last failure.
 *   - Reentrant.  The only process-wide state is the error string (thread
 *     local) and one-time kernel attribute setup.
 *
 * Each entry point cites the reference binding it replaces (file:line under
 * submodules/gsplat/cuda/csrc/ of the reference).
 */
#ifndef UBS_CAPI_H
#define UBS_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UBS_OK 0
#define UBS_ERR_INVALID 1   /* bad argument: shape, size or unsupported option */
#define UBS_ERR_HIP 2       /* HIP runtime error */
#define UBS_ERR_WORKSPACE 3 /* workspace smaller than required */

/* Library version (major*10000 + minor*100 + patch). */
int ubs_version(void);

/* Message of the calling thread's last failed call ("" if none). */
const char *ubs_last_error(void);

/* ------------------------------------------------------------------------
 * Parameter activations and the view/time query (one pass per primitive).
 * ------------------------------------------------------------------------ */

/* Leaves xyz [N,3], mean_rest [N,D-3], scale_raw [N,D], opacity_raw [N,1],
 * beta_raw [N,D-2]; optional quats [N,4] (w,x,y,z); cam_pos device [3].
 * -> mean [N,D] = cat(xyz, mean_rest), scale [N,D] = softplus(scale_raw)
 *    (x above 20), opacity [N,1] = sigmoid, beta [N,D-2] = 4 exp(beta_raw),
 *    query [N,D-3] = normalize(xyz - cam_pos) (, timestamp for D = 7; may be
 *    NULL, requires D = 6 or 7), rot [N,3,3] = R(quats / |quats|) (may be NULL).
 * Replaces the torch elementwise ops BetaModel runs before the rasteriser:
 * scene/beta_model.py:36-52 and :103-121 (activations), :675-690 (query). */
int ubs_beta_activate_fwd(int64_t N, int64_t D, const float *xyz, const float *mean_rest, const float *scale_raw,
                          const float *opacity_raw, const float *beta_raw, const float *quats, const float *cam_pos,
                          float timestamp, float *mean, float *scale, float *opacity, float *beta, float *query,
                          float *rot, void *stream);

/* Backward of ubs_beta_activate_fwd (the query gets no gradient, as the
 * conditioning backward drops it).  Any of v_mean, v_scale, v_opacity,
 * v_beta, v_rot may be NULL (zero); v_quats may be NULL (not computed).
 * -> v_xyz [N,3], v_mean_rest [N,D-3], v_scale_raw [N,D], v_opacity_raw [N,1],
 *    v_beta_raw [N,D-2], v_quats [N,4].  Same reference lines as above. */
int ubs_beta_activate_bwd(int64_t N, int64_t D, const float *scale_raw, const float *opacity_raw,
                          const float *beta_raw, const float *quats, const float *v_mean, const float *v_scale,
                          const float *v_opacity, const float *v_beta, const float *v_rot, float *v_xyz,
                          float *v_mean_rest, float *v_scale_raw, float *v_opacity_raw, float *v_beta_raw,
                          float *v_quats, void *stream);

/* ------------------------------------------------------------------------
 * Covariance construction.
 * ------------------------------------------------------------------------ */

/* R = I + A from the skew parameters (a01, a02, a12).
 * l_triangle [N,3] -> rotmat [N,3,3].
 * Replaces l_triangle_to_rotmat_fwd_tensor, l_triagnle_to_rotmat_fwd.cu:38-67. */
int ubs_l_triangle_to_rotmat_fwd(int64_t N, const float *l_triangle, float *rotmat, void *stream);

/* v_rotmat [N,3,3] -> v_l_triangle [N,3].
 * Replaces l_triangle_to_rotmat_bwd_tensor, l_triagnle_to_rotmat_bwd.cu:27-58. */
int ubs_l_triangle_to_rotmat_bwd(int64_t N, const float *v_rotmat, float *v_l_triangle, void *stream);

/* Sigma = L L^T, L[:3,:3] = rot diag(scale[:3]), L[i,i] = scale[i] (i>=3),
 * L[rest_i[m], rest_j[m]] = l_triangle[3+m].  rest_i/rest_j: device int32
 * [M_rest], grouped by row with ascending columns.  3 <= D <= 11.
 * covar: [N,D,D], or [N,3,3] when spatial_block != 0 or D == 3.
 * Replaces rot_scale_l_triangle_to_covar_fwd_tensor,
 * rot_scale_l_triangle_to_covar_fwd.cu:191-235. */
int ubs_rot_scale_l_triangle_to_covar_fwd(int64_t N, int64_t D, const float *rot, const float *scale,
                                          const float *l_triangle, const int32_t *rest_i,
                                          const int32_t *rest_j, int64_t M_rest, int spatial_block,
                                          float *covar, void *stream);

/* Backward: v_covar [N,dim,dim] -> v_rot [N,3,3], v_scale [N,D],
 * v_l_triangle [N,D(D-1)/2] (first three entries zero).
 * Replaces rot_scale_l_triangle_to_covar_bwd_tensor,
 * rot_scale_l_triangle_to_covar_bwd.cu:238-293. */
int ubs_rot_scale_l_triangle_to_covar_bwd(int64_t N, int64_t D, const float *rot, const float *scale,
                                          const float *l_triangle, const int32_t *rest_i,
                                          const int32_t *rest_j, int64_t M_rest, int spatial_block,
                                          const float *v_covar, float *v_rot, float *v_scale,
                                          float *v_l_triangle, void *stream);

/* ------------------------------------------------------------------------
 * Conditioning of a D-dim Beta primitive on its C = D-3 query dims
 * (view direction, time).  4 <= D <= 11.
 * ------------------------------------------------------------------------ */

/* means [N,D], covars [N,D,D], opacities [N,1], betas [N,C], query [N,C]
 * -> out_means [N,3], out_covars [N,3,3], out_opacities [N,1].
 * Replaces cond_mean_convariance_opacity_fwd_tensor,
 * cond_mean_convariance_opacity_fwd.cu:292-349. */
int ubs_cond_mean_convariance_opacity_fwd(int64_t N, int64_t D, const float *means, const float *covars,
                                          const float *opacities, const float *betas,
                                          const float *query, float *out_means, float *out_covars,
                                          float *out_opacities, void *stream);

/* Backward (the query receives no gradient, as in the reference):
 * v_means [N,3], v_covars [N,3,3], v_opacities [N,1] -> g_means [N,D],
 * g_covars [N,D,D], g_opacities [N,1], g_betas [N,C].
 * Replaces cond_mean_convariance_opacity_bwd_tensor,
 * cond_mean_convariance_opacity_bwd.cu:565-630. */
int ubs_cond_mean_convariance_opacity_bwd(int64_t N, int64_t D, const float *means, const float *covars,
                                          const float *opacities, const float *betas,
                                          const float *query, const float *v_means,
                                          const float *v_covars, const float *v_opacities,
                                          float *g_means, float *g_covars, float *g_opacities,
                                          float *g_betas, void *stream);

/* ------------------------------------------------------------------------
 * EWA projection.
 * ------------------------------------------------------------------------ */

/* Layout of the `covars` argument of the projection entry points. */
#define UBS_COVAR_TRIU6 0  /* [N,6]: c00 c01 c02 c11 c12 c22 (the reference layout) */
#define UBS_COVAR_FULL33 1 /* [N,3,3] row-major; only the upper triangle is read, and the
                              backward writes the upper-triangle-only gradient (what the
                              gather at rendering.py:54-56 back-propagates) */

/* means [N,3]; either covars (layout covar_layout) or quats [N,4] + scales
 * [N,3]; viewmats [C,4,4] world->camera; Ks [C,3,3].
 * Outputs radii int32 [C,N], means2d [C,N,2], depths [C,N], conics [C,N,3],
 * compensations [C,N] (may be NULL).  Culled entries get radii = 0 and leave
 * the other outputs untouched.
 * Replaces fully_fused_projection_fwd_tensor, fully_fused_projection_fwd.cu:179-256. */
int ubs_fully_fused_projection_fwd(uint32_t C, uint32_t N, const float *means, const float *covars,
                                   int covar_layout, const float *quats, const float *scales, const float *viewmats,
                                   const float *Ks, int32_t width, int32_t height, float eps2d,
                                   float near_plane, float far_plane, float radius_clip, int ortho,
                                   int32_t *radii, float *means2d, float *depths, float *conics,
                                   float *compensations, void *stream);

/* Backward.  v_compensations and compensations may be NULL.  Outputs (zeroed
 * here): v_means [N,3]; v_covars (same layout as covars) when covars given,
 * else v_quats [N,4] and v_scales [N,3]; v_viewmats [C,4,4] or NULL.
 * Replaces fully_fused_projection_bwd_tensor, fully_fused_projection_bwd.cu:242-348. */
int ubs_fully_fused_projection_bwd(uint32_t C, uint32_t N, const float *means, const float *covars,
                                   int covar_layout, const float *quats, const float *scales, const float *viewmats,
                                   const float *Ks, int32_t width, int32_t height, float eps2d,
                                   int ortho, const int32_t *radii, const float *conics,
                                   const float *compensations, const float *v_means2d,
                                   const float *v_depths, const float *v_conics,
                                   const float *v_compensations, float *v_means, float *v_covars,
                                   float *v_quats, float *v_scales, float *v_viewmats, void *stream);

/* ------------------------------------------------------------------------
 * Fused Beta projection: leaves of a D-dim Beta scene (D = 6 or 7) -> the
 * per-view 2-D primitives, for C views in one launch.  Replaces, per view,
 * the chain BetaModel.render runs before the blend (scene/beta_model.py:
 * 660-722): the activations and query (:36-52, :103-121, :675-690),
 * l_triangle_to_rotmat_fwd_tensor (l_triagnle_to_rotmat_fwd.cu:38-67) or the
 * quaternion rotation, rot_scale_l_triangle_to_covar_fwd_tensor
 * (rot_scale_l_triangle_to_covar_fwd.cu:191-235),
 * cond_mean_convariance_opacity_fwd_tensor
 * (cond_mean_convariance_opacity_fwd.cu:292-349), the [N,3,3] -> [N,6]
 * gather and repeat of rendering.py:54-56,100-101 and
 * fully_fused_projection_fwd_tensor (fully_fused_projection_fwd.cu:179-256).
 * With C = 1 every output equals that chain's bit for bit.
 * ------------------------------------------------------------------------ */

/* Leaves as BetaModel stores them, pre-activation: xyz [N,3], mean_rest
 * [N,D-3], scale_raw [N,D], l_triangle [N,D(D-1)/2] (canonical rest pattern of
 * beta_model.py:69-73), opacity_raw [N,1], beta_raw [N,D-2]; quats [N,4] or
 * NULL (then the rotation is I + skew(l_triangle[:, :3])).
 * Per view: cam_pos [C,3] (camera centre; NULL -> taken from viewmats as
 * ubs_camera_centers does), timestamps [C] (7-D; NULL -> 0),
 * viewmats [C,4,4], Ks [C,3,3]; image size width x height, or per view
 * image_wh int32 [C,2] when not NULL.
 * Outputs [C,N]: radii int32 (0 = culled; the other 2-D outputs of a culled
 * entry are left untouched), means2d [C,N,2], depths, conics [C,N,3],
 * opacities (conditioned) and betas (beta_0) for every entry. */
int ubs_beta_project_fwd(int64_t N, int64_t D, uint32_t C, const float *xyz, const float *mean_rest,
                         const float *scale_raw, const float *l_triangle, const float *opacity_raw,
                         const float *beta_raw, const float *quats, const float *cam_pos, const float *timestamps,
                         const float *viewmats, const float *Ks, int32_t width, int32_t height,
                         const int32_t *image_wh, float eps2d, float near_plane, float far_plane, float radius_clip,
                         int ortho, int32_t *radii, float *means2d, float *depths, float *conics, float *opacities,
                         float *betas, void *stream);

/* ubs_beta_project_fwd that also writes, for every visible entry c N + n,
 * the blend's 64-byte record (means2d, conic, opacity, beta_0, rgb[n] [N,3])
 * into records (the first ubs_rasterize_workspace_bytes(C, N, 0) bytes of
 * the blend's workspace), for ubs_rasterize_to_pixels_fwd_prepacked.  No
 * counterpart in the reference. */
int ubs_beta_project_fwd_rec(int64_t N, int64_t D, uint32_t C, const float *xyz, const float *mean_rest,
                             const float *scale_raw, const float *l_triangle, const float *opacity_raw,
                             const float *beta_raw, const float *quats, const float *cam_pos,
                             const float *timestamps, const float *viewmats, const float *Ks, int32_t width,
                             int32_t height, const int32_t *image_wh, float eps2d, float near_plane,
                             float far_plane, float radius_clip, int ortho, int32_t *radii, float *means2d,
                             float *depths, float *conics, float *opacities, float *betas, const float *rgb,
                             void *records, void *stream);

/* Camera centres [C,3] of world->camera viewmats [C,4,4] (row-major):
 * -(R^T t), with the FMA order the fused projection uses when its cam_pos is
 * NULL.  The reference takes the centre from its Camera object
 * (scene/beta_model.py:676, viewpoint_camera.camera_center). */
int ubs_camera_centers(uint32_t C, const float *viewmats, float *centres, void *stream);

/* Backward: the saved radii and conics and the gradients [C,N] of means2d,
 * depths (NULL -> zeros), conics, opacities and betas -> the leaf gradients (each written
 * once, summed over the views): v_xyz [N,3], v_mean_rest [N,D-3],
 * v_scale_raw [N,D], v_l_triangle [N,D(D-1)/2], v_opacity_raw [N,1],
 * v_beta_raw [N,D-2], v_quats [N,4] (required when quats is given).
 * Replaces the backward of the chain above (fully_fused_projection_bwd.cu:
 * 242-348, cond_mean_convariance_opacity_bwd.cu:565-630,
 * rot_scale_l_triangle_to_covar_bwd.cu:238-293, l_triagnle_to_rotmat_bwd.cu:
 * 27-58 and torch's activation backward); follows ubs_set_cond_backward_exact. */
/* The same projection writing view c's rows at c * row_stride (row_stride
 * >= N): several scenes' views share one [C, row_stride] batch of projected
 * rows, each scene's call writing its views' first N entries; the caller
 * zero-fills the radii of the rows past a scene's N, which then count as
 * culled.  No counterpart in the reference (the render server's batches
 * across resident scenes). */
int ubs_beta_project_fwd_rows(int64_t N, int64_t D, uint32_t C, const float *xyz, const float *mean_rest,
                              const float *scale_raw, const float *l_triangle, const float *opacity_raw,
                              const float *beta_raw, const float *quats, const float *cam_pos,
                              const float *timestamps, const float *viewmats, const float *Ks, int32_t width,
                              int32_t height, const int32_t *image_wh, float eps2d, float near_plane,
                              float far_plane, float radius_clip, int ortho, int32_t *radii, float *means2d,
                              float *depths, float *conics, float *opacities, float *betas, int64_t row_stride,
                              void *stream);

/* One view's scene in a batch over several resident scenes: the leaf
 * tensors as ubs_beta_project_fwd takes them (quats may be NULL), rgb [N,3],
 * N (0 marks an empty slot) and D (6 or 7). */
typedef struct ubs_view_scene {
    const float *xyz, *mean_rest, *scale_raw, *l_triangle, *opacity_raw, *beta_raw, *quats, *rgb;
    int64_t N, D;
} ubs_view_scene;

/* Forward projection of C views of several scenes (the render server's
 * batches): table (device memory, C entries) names each view's scene; view
 * c's rows are entries c * row_stride + n (row_stride >= the largest N),
 * image_wh [C,2] per view, and colors [C, row_stride, 3] receives each view's
 * scene colours.  Entries past a view's N get radius 0.  Per view the
 * results equal ubs_beta_project_fwd's for that scene bit for bit.  No
 * counterpart in the reference (BetaModel.view renders one request at a
 * time, scene/beta_model.py:725-831). */
int ubs_beta_project_fwd_table(uint32_t C, int64_t row_stride, const ubs_view_scene *table, const float *cam_pos,
                               const float *timestamps, const float *viewmats, const float *Ks,
                               const int32_t *image_wh, float eps2d, float near_plane, float far_plane,
                               float radius_clip, int32_t *radii, float *means2d, float *depths, float *conics,
                               float *opacities, float *betas, float *colors, void *stream);

int ubs_beta_project_bwd(int64_t N, int64_t D, uint32_t C, const float *xyz, const float *mean_rest,
                         const float *scale_raw, const float *l_triangle, const float *opacity_raw,
                         const float *beta_raw, const float *quats, const float *cam_pos, const float *timestamps,
                         const float *viewmats, const float *Ks, int32_t width, int32_t height,
                         const int32_t *image_wh, float eps2d, int ortho, const int32_t *radii, const float *conics,
                         const float *v_means2d, const float *v_depths, const float *v_conics,
                         const float *v_opacities, const float *v_betas, float *v_xyz, float *v_mean_rest,
                         float *v_scale_raw, float *v_l_triangle, float *v_opacity_raw, float *v_beta_raw,
                         float *v_quats, void *stream);

/* ------------------------------------------------------------------------
 * Tile intersection (isect_tiles.cu:99-285 split at its host sync).
 * ------------------------------------------------------------------------ */

/* Workspace bytes for ubs_isect_tiles_count. */
size_t ubs_isect_count_workspace_bytes(uint32_t C, uint32_t N);

/* Pass 1: tiles_per_gauss int32 [C,N] and their inclusive cumulative sum
 * cum_tiles int64 [C*N]; *n_isects_dev (device int64) receives the total.
 * The host reads n_isects to size pass 2. */
int ubs_isect_tiles_count(uint32_t C, uint32_t N, const float *means2d, const int32_t *radii,
                          uint32_t tile_size, uint32_t tile_width, uint32_t tile_height,
                          int32_t *tiles_per_gauss, int64_t *cum_tiles, int64_t *n_isects_dev,
                          void *workspace, size_t workspace_bytes, void *stream);

/* Pass 2: isect_ids int64 [n_isects] = cam | tile | depth bits (unsorted,
 * primitive-major) and flatten_ids int32 [n_isects]. */
int ubs_isect_tiles_emit(uint32_t C, uint32_t N, const float *means2d, const int32_t *radii,
                         const float *depths, const int64_t *cum_tiles, uint32_t tile_size,
                         uint32_t tile_width, uint32_t tile_height, int64_t *isect_ids,
                         int32_t *flatten_ids, void *stream);

/* Pass 1 that also writes the bit range of the depths [C,N] of the entries
 * with radii > 0, as uint32 bit patterns, into device memory range_dev[2]:
 * range_dev[0] = ~min, range_dev[1] = max (both 0 when nothing is visible).
 * Depth bytes on which min and max agree are the same in every isect id, so
 * the sort can skip their passes (ubs_radix_sort_pairs_skip).  range_dev may
 * be NULL (then depths is unused).  No counterpart in the reference (it sorts
 * all 32 + prefix bits, isect_tiles.cu:224-278). */
int ubs_isect_tiles_count_range(uint32_t C, uint32_t N, const float *means2d, const int32_t *radii,
                                const float *depths, uint32_t tile_size, uint32_t tile_width,
                                uint32_t tile_height, int32_t *tiles_per_gauss, int64_t *cum_tiles,
                                int64_t *n_isects_dev, uint32_t *range_dev, void *workspace,
                                size_t workspace_bytes, void *stream);

/* The same pass writing the three int64 words the host reads in one copy:
 * info[0] = n_isects, info[1] = the depth range as above (two uint32),
 * info[2] = *status_in (a status word the caller accumulates; NULL -> 0);
 * with depth_keys / depth_vals (both or neither) it also writes what
 * ubs_isect_depth_keys writes, for the depth-first emission's sort; with
 * pair_records (float4 [C*N], may be NULL) the pair's (x, y, radius bits,
 * depth) for ubs_isect_tiles_emit_ordered_rec.  No counterpart in the
 * reference. */
int ubs_isect_tiles_count_info(uint32_t C, uint32_t N, const float *means2d, const int32_t *radii,
                               const float *depths, uint32_t tile_size, uint32_t tile_width,
                               uint32_t tile_height, int32_t *tiles_per_gauss, int64_t *cum_tiles,
                               int64_t *info, const int64_t *status_in, int64_t *depth_keys,
                               int32_t *depth_vals, void *pair_records, void *workspace,
                               size_t workspace_bytes, void *stream);

/* ubs_isect_tiles_emit_ordered_cap reading each pair's centre, radius and
 * depth from the 16-byte pair records ubs_isect_tiles_count_info wrote (one
 * gather per pair in depth order instead of three).  No counterpart in the
 * reference. */
int ubs_isect_tiles_emit_ordered_rec(uint32_t C, uint32_t N, const void *pair_records, const int32_t *order,
                                     const int64_t *cum, uint32_t tile_size, uint32_t tile_width,
                                     uint32_t tile_height, int64_t *isect_ids, int32_t *flatten_ids,
                                     int64_t cap, void *stream);

/* Bits the (cam, tile) prefix of an isect id occupies: 32 + this = end_bit. */
uint32_t ubs_isect_prefix_bits(uint32_t C, uint32_t tile_width, uint32_t tile_height);

/* ------------------------------------------------------------------------
 * Onesweep LSD radix sort of (int64 key, int32 value) pairs on key bits
 * [begin_bit, end_bit), stable (what cub::DeviceRadixSort::SortPairs does at
 * isect_tiles.cu:224-278).  Double-buffered: the result lands in
 * (keys, vals) or (keys_alt, vals_alt); *result_in_alt tells which.
 * n <= 2^30 - 1.
 * ------------------------------------------------------------------------ */
size_t ubs_radix_sort_workspace_bytes(int64_t n, int begin_bit, int end_bit);
int ubs_radix_sort_pairs(int64_t n, int64_t *keys, int64_t *keys_alt, int32_t *vals,
                         int32_t *vals_alt, int begin_bit, int end_bit, void *workspace,
                         size_t workspace_bytes, int *result_in_alt, void *stream);
/* The same sort without the passes whose bit set in skip_passes (bit p =
 * key bits [begin_bit + 8p, begin_bit + 8p + 8)): the caller asserts that
 * every key has the same digit there, so the pass would be the identity (the
 * isect path knows it for the high bytes of the depths from
 * ubs_isect_tiles_count_range).  No counterpart in the reference. */
int ubs_radix_sort_pairs_skip(int64_t n, int64_t *keys, int64_t *keys_alt, int32_t *vals,
                              int32_t *vals_alt, int begin_bit, int end_bit, uint32_t skip_passes,
                              void *workspace, size_t workspace_bytes, int *result_in_alt, void *stream);
/* ubs_radix_sort_pairs_skip that also sets *flag_dev = 1 (device memory,
 * never cleared by the sort) when a look-back wait gives up, so a caller can
 * collect the status of several sorts in one word and read it with a later
 * host read.  No counterpart in the reference. */
int ubs_radix_sort_pairs_flag(int64_t n, int64_t *keys, int64_t *keys_alt, int32_t *vals,
                              int32_t *vals_alt, int begin_bit, int end_bit, uint32_t skip_passes,
                              void *workspace, size_t workspace_bytes, int *result_in_alt, uint32_t *flag_dev,
                              void *stream);

/* Depth-first emission: the same intersection entries in an order that a
 * stable sort on the camera/tile bits alone ([32, 32 + tile_bits +
 * cam_bits)) turns into the full stable sort's order, bit for bit.  The
 * (camera, primitive) pairs are sorted stably by depth once (C*N pairs
 * instead of n_isects entries), then each emits its tiles in that order.
 *   ubs_isect_depth_keys: keys int64 [C*N] = depth bit pattern of each
 *     visible pair (0xFFFFFFFF when culled), vals int32 [C*N] = flat id;
 *     sort them on bits [0, 32) -> order (the sorted vals).
 *   ubs_isect_scan_ordered: cum int64 [C*N] = inclusive prefix of
 *     tiles_per_gauss[order[k]]; workspace ubs_isect_ordered_workspace_bytes.
 *   ubs_isect_tiles_emit_ordered: the entries of pair order[k] at
 *     [cum[k-1], cum[k]), its tiles in row-major order.
 * Replaces the emission and the depth passes of the sort of
 * isect_tiles.cu:76-96 and :224-278. */
int ubs_isect_depth_keys(uint32_t C, uint32_t N, const int32_t *radii, const float *depths, int64_t *keys,
                         int32_t *vals, void *stream);
size_t ubs_isect_ordered_workspace_bytes(uint32_t C, uint32_t N);
int ubs_isect_scan_ordered(uint32_t C, uint32_t N, const int32_t *tiles_per_gauss, const int32_t *order,
                           int64_t *cum, void *workspace, size_t workspace_bytes, void *stream);
int ubs_isect_tiles_emit_ordered(uint32_t C, uint32_t N, const float *means2d, const int32_t *radii,
                                 const float *depths, const int32_t *order, const int64_t *cum, uint32_t tile_size,
                                 uint32_t tile_width, uint32_t tile_height, int64_t *isect_ids, int32_t *flatten_ids,
                                 void *stream);

/* Device-sized intersection lists (no host read of n_isects): the caller
 * allocates the lists for a capacity `cap` and passes the count that
 * ubs_isect_tiles_count[_range] left in device memory (n_dev).
 *   ubs_isect_tiles_emit_ordered_cap: ubs_isect_tiles_emit_ordered storing
 *     only the entries at index < cap (a count above cap is an overflow the
 *     caller detects from *n_dev after the work and repeats with more room);
 *   ubs_radix_sort_pairs_dev: ubs_radix_sort_pairs of min(*n_dev, n_cap)
 *     keys, workspace ubs_radix_sort_workspace_bytes(n_cap, ...); every pass
 *     runs, so *result_in_alt is known on return;
 *   ubs_isect_offset_encode_dev: ubs_isect_offset_encode of min(*n_dev,
 *     n_cap) sorted ids.
 * The launches of a batch then have fixed shapes and no host
 * synchronisation, so a caller can capture them in a hipGraph and replay
 * it.  No counterpart in the reference (it reads n_isects on the host,
 * isect_tiles.cu:180-181). */
int ubs_isect_tiles_emit_ordered_cap(uint32_t C, uint32_t N, const float *means2d, const int32_t *radii,
                                     const float *depths, const int32_t *order, const int64_t *cum, uint32_t ts,
                                     uint32_t tw, uint32_t th, int64_t *isect_ids, int32_t *flatten_ids,
                                     int64_t cap, void *stream);
int ubs_radix_sort_pairs_dev(int64_t n_cap, const int64_t *n_dev, int64_t *keys, int64_t *keys_alt, int32_t *vals,
                             int32_t *vals_alt, int begin_bit, int end_bit, void *workspace, size_t ws_bytes,
                             int *result_in_alt, void *stream);
int ubs_isect_offset_encode_dev(int64_t n_cap, const int64_t *n_dev, const int64_t *isect_ids, uint32_t C,
                                uint32_t tile_width, uint32_t tile_height, int32_t *offsets, void *stream);

/* Depth sort inside each (camera, tile) segment of isect ids that are
 * already sorted stably on their camera/tile bits (ubs_radix_sort_pairs_skip
 * with begin_bit 32): keys_in/vals_in [n] -> keys_out/vals_out [n], each
 * segment [tile_offsets[L], tile_offsets[L+1]) (the last to n) sorted stably
 * on the 32 depth bits.  The two steps together give the order of the full
 * stable sort on bits [0, 32 + tile_bits + cam_bits), bit for bit.
 * tile_offsets: ubs_isect_offset_encode of keys_in (the same for keys_out).
 * skip_passes: bit p leaves out the pass over depth byte p (every key shares
 * it).  keys_in/vals_in are scratch for segments longer than 8192.
 * Replaces the depth passes of the SortPairs call of isect_tiles.cu:224-278. */
int ubs_sort_tile_segments(int64_t n, int64_t *keys_in, int32_t *vals_in, int64_t *keys_out, int32_t *vals_out,
                           const int32_t *tile_offsets, uint32_t C, uint32_t tile_width, uint32_t tile_height,
                           uint32_t skip_passes, void *stream);

/* The same stable sort (all passes of [begin_bit, end_bit)) with no
 * workspace argument and no allocation: its histograms, tile counters and
 * look-back slots live in a 24 MiB device array of the library, one per GPU,
 * held until the process ends.  Sorts whose scratch would not fit (more than
 * 16777216 keys at 8 passes; ubs_radix_sort_workspace_bytes tells) are
 * rejected with UBS_ERR_INVALID before anything is enqueued.  status_dev
 * (device uint32, may be NULL) receives the status word that
 * ubs_radix_sort_status reads for the workspace form.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (two streams, two
 * host threads) race and must be serialised by the caller.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under hipGraph capture, serving loops, harnesses that forbid
 * hipMalloc). */
int ubs_radix_sort_pairs_noalloc(int64_t n, int64_t *keys, int64_t *keys_alt, int32_t *vals, int32_t *vals_alt,
                                 int begin_bit, int end_bit, int *result_in_alt, uint32_t *status_dev,
                                 void *stream);

/* Reads the sort's device-side status word (0 = fine, nonzero = a look-back
 * wait hit its spin limit and the result is invalid).  Synchronises stream. */
int ubs_radix_sort_status(const void *workspace, int *status, void *stream);

/* offsets int32 [C, tile_height, tile_width]: first sorted index of each
 * (camera, tile); empty tiles get the next start.  n_isects = 0 -> zeros.
 * Replaces isect_offset_encode_tensor, isect_tiles.cu:335-368. */
int ubs_isect_offset_encode(int64_t n_isects, const int64_t *isect_ids, uint32_t C,
                            uint32_t tile_width, uint32_t tile_height, int32_t *offsets,
                            void *stream);

/* ------------------------------------------------------------------------
 * Tile blending.  tile_size in {8, 16, 32}; any channel count >= 1.
 * ------------------------------------------------------------------------ */

/* Workspace of the blend entry points: room for one 64-byte record per
 * (camera, primitive) -- x, y, conic, opacity, beta and, for blends of at
 * most 4 channels in one pass, the colours -- that the forward and the
 * backward gather per intersection instead of five separate arrays, and in
 * the backward as much again for the packed gradients.  A NULL workspace
 * selects the unpacked path (same results within rounding, more gather
 * traffic). */
size_t ubs_rasterize_workspace_bytes(uint32_t C, uint32_t N, int backward);

/* The same workspace with room, after it, for the order in which the blend
 * deals its tiles to the GPU: heaviest first (by their list
 * lengths), where ubs_rasterize_workspace_bytes gives raster order.  Any
 * blend entry point given at least this many bytes computes that order from
 * tile_offsets and uses it; results are the same as in raster order (the
 * backward's float atomics add in another order: same within rounding).
 * No counterpart in the reference (its launch grid is in raster order,
 * rasterize_to_pixels_fwd.cu:235-236). */
size_t ubs_rasterize_workspace_bytes_tiles(uint32_t C, uint32_t N, uint32_t tile_width, uint32_t tile_height,
                                           int backward);

/* means2d [C,N,2], conics [C,N,3], colors [C,N,channels], opacities [C,N],
 * betas [C,N], backgrounds [C,channels] or NULL, masks uint8 [C,th,tw] or
 * NULL, tile_offsets [C,th,tw], flatten_ids [n_isects]
 * -> render_colors [C,H,W,channels], render_alphas [C,H,W,1],
 *    last_ids int32 [C,H,W], T_final [C,H,W] (transmittance after the last
 *    contributor, kept exact for the backward; may be NULL).
 * workspace: NULL or ubs_rasterize_workspace_bytes(C, N, 0) bytes.
 * Replaces rasterize_to_pixels_fwd_tensor, rasterize_to_pixels_fwd.cu:193-361. */
int ubs_rasterize_to_pixels_fwd(uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels,
                                const float *means2d, const float *conics, const float *colors,
                                const float *opacities, const float *betas,
                                const float *backgrounds, const uint8_t *masks, uint32_t width,
                                uint32_t height, uint32_t tile_size, uint32_t tile_width,
                                uint32_t tile_height, const int32_t *tile_offsets,
                                const int32_t *flatten_ids, float *render_colors,
                                float *render_alphas, int32_t *last_ids, float *T_final,
                                void *workspace, size_t workspace_bytes, void *stream);

/* ubs_rasterize_to_pixels_fwd whose workspace already holds the packed
 * records (3 colour channels, tile_size 8 or 16: the 64-byte ones) of every
 * visible entry, written by ubs_beta_project_fwd_rec: the packing pass is
 * left out.  No counterpart in the reference. */
int ubs_rasterize_to_pixels_fwd_prepacked(uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels,
                                          const float *means2d, const float *conics, const float *colors,
                                          const float *opacities, const float *betas, const float *backgrounds,
                                          const uint8_t *masks, uint32_t W, uint32_t H, uint32_t tile_size,
                                          uint32_t tw, uint32_t th, const int32_t *tile_offsets,
                                          const int32_t *flatten_ids, float *render_colors, float *render_alphas,
                                          int32_t *last_ids, float *T_final, void *workspace,
                                          size_t workspace_bytes, void *stream);

/* Backward.  T_final may be NULL (then 1 - render_alphas is used, as the
 * reference does); v_render_alphas may be NULL (a zero gradient, without
 * the caller filling one).  Outputs (zeroed here): v_means2d [C,N,2], v_conics
 * [C,N,3], v_colors [C,N,channels], v_opacities [C,N], v_betas [C,N].
 * workspace: NULL or ubs_rasterize_workspace_bytes(C, N, 1) bytes.
 * Replaces rasterize_to_pixels_bwd_tensor, rasterize_to_pixels_bwd.cu:278-479. */
int ubs_rasterize_to_pixels_bwd(uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels,
                                const float *means2d, const float *conics, const float *colors,
                                const float *opacities, const float *betas,
                                const float *backgrounds, const uint8_t *masks, uint32_t width,
                                uint32_t height, uint32_t tile_size, uint32_t tile_width,
                                uint32_t tile_height, const int32_t *tile_offsets,
                                const int32_t *flatten_ids, const float *render_alphas,
                                const int32_t *last_ids, const float *T_final,
                                const float *v_render_colors, const float *v_render_alphas,
                                float *v_means2d, float *v_conics, float *v_colors,
                                float *v_opacities, float *v_betas, void *workspace,
                                size_t workspace_bytes, void *stream);

/* Stored-transmittance variant of the blend pair (the recompute-vs-store
 * choice for per-pixel transmittance; no counterpart in the reference,
 * whose backward only recovers T by division, rasterize_to_pixels_bwd.cu:
 * 202-203).  The forward also writes each pixel's T after every batch of
 * tile_size^2 entries of its tile's list into T_ckpt; the backward restarts
 * each batch from it, so the division recovery runs over one batch at a
 * time.  T_ckpt holds ubs_transmittance_checkpoint_floats(...) floats and
 * needs no initialisation; the pair must see the same tile lists. */
size_t ubs_transmittance_checkpoint_floats(uint32_t C, int64_t n_isects, uint32_t tile_width,
                                           uint32_t tile_height, uint32_t tile_size);

/* ubs_rasterize_to_pixels_fwd over lists sized on the device: n_cap is the
 * lists' capacity, *n_dev (device int64) the intersection count (clamped to
 * n_cap).  No counterpart in the reference (see ubs_radix_sort_pairs_dev). */
int ubs_rasterize_to_pixels_fwd_dev(uint32_t C, uint32_t N, int64_t n_cap, const int64_t *n_dev,
                                    uint32_t channels, const float *means2d, const float *conics,
                                    const float *colors, const float *opacities, const float *betas,
                                    const float *backgrounds, const uint8_t *masks, uint32_t W, uint32_t H,
                                    uint32_t tile_size, uint32_t tw, uint32_t th, const int32_t *tile_offsets,
                                    const int32_t *flatten_ids, float *render_colors, float *render_alphas,
                                    int32_t *last_ids, float *T_final, void *workspace, size_t workspace_bytes,
                                    void *stream);

int ubs_rasterize_to_pixels_fwd_ckpt(uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels,
                                     const float *means2d, const float *conics, const float *colors,
                                     const float *opacities, const float *betas,
                                     const float *backgrounds, const uint8_t *masks, uint32_t width,
                                     uint32_t height, uint32_t tile_size, uint32_t tile_width,
                                     uint32_t tile_height, const int32_t *tile_offsets,
                                     const int32_t *flatten_ids, float *render_colors,
                                     float *render_alphas, int32_t *last_ids, float *T_final,
                                     float *T_ckpt, size_t ckpt_floats, void *workspace,
                                     size_t workspace_bytes, void *stream);

int ubs_rasterize_to_pixels_bwd_ckpt(uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels,
                                     const float *means2d, const float *conics, const float *colors,
                                     const float *opacities, const float *betas,
                                     const float *backgrounds, const uint8_t *masks, uint32_t width,
                                     uint32_t height, uint32_t tile_size, uint32_t tile_width,
                                     uint32_t tile_height, const int32_t *tile_offsets,
                                     const int32_t *flatten_ids, const float *render_alphas,
                                     const int32_t *last_ids, const float *T_final,
                                     const float *T_ckpt, size_t ckpt_floats,
                                     const float *v_render_colors, const float *v_render_alphas,
                                     float *v_means2d, float *v_conics, float *v_colors,
                                     float *v_opacities, float *v_betas, void *workspace,
                                     size_t workspace_bytes, void *stream);

/* ubs_rasterize_to_pixels_bwd (T_ckpt NULL) or _bwd_ckpt that reuses the
 * forward's packed primitive records and tile order instead of packing and
 * ordering again: fwd_workspace is the workspace the forward of the same
 * inputs was given, of ubs_rasterize_workspace_bytes_tiles(C, N, tile_width,
 * tile_height, 0) bytes (read only); workspace then holds the packed
 * gradients only, ubs_rasterize_workspace_bytes(C, N, 0) bytes.  Same
 * results as the backward given one workspace.  No counterpart in the
 * reference (its backward reads the separate arrays,
 * rasterize_to_pixels_bwd.cu:138-151). */
int ubs_rasterize_to_pixels_bwd_reuse(uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels,
                                      const float *means2d, const float *conics, const float *colors,
                                      const float *opacities, const float *betas, const float *backgrounds,
                                      const uint8_t *masks, uint32_t width, uint32_t height, uint32_t tile_size,
                                      uint32_t tile_width, uint32_t tile_height, const int32_t *tile_offsets,
                                      const int32_t *flatten_ids, const float *render_alphas,
                                      const int32_t *last_ids, const float *T_final, const float *T_ckpt,
                                      size_t ckpt_floats, const float *v_render_colors,
                                      const float *v_render_alphas, float *v_means2d, float *v_conics,
                                      float *v_colors, float *v_opacities, float *v_betas,
                                      const void *fwd_workspace, size_t fwd_workspace_bytes, void *workspace,
                                      size_t workspace_bytes, void *stream);

/* ------------------------------------------------------------------------
 * Image loss of the training step: (1 - lambda) L1 + lambda (1 - SSIM).
 * Replaces torch's l1_loss (utils/loss_utils.py:18-19) and the external
 * fused_ssim the reference calls at train.py:117-121, whose arithmetic is
 * the SSIM of utils/loss_utils.py:45-80 (11x11 Gaussian window, sigma 1.5,
 * zero padding, C1 = 0.01^2, C2 = 0.03^2).
 * ------------------------------------------------------------------------ */

/* Workspace bytes for images [B,H,W,C]. */
size_t ubs_image_loss_workspace_bytes(int64_t B, int64_t H, int64_t W, int64_t C);

/* img, gt float32 [B,H,W,C]; *loss (device float) = (1 - lambda) mean|img - gt|
 * + lambda (1 - mean SSIM), the means over every element.  Leaves in the
 * workspace what ubs_image_loss_bwd needs.  lambda_dssim in [0, 1]. */
int ubs_image_loss_fwd(int64_t B, int64_t H, int64_t W, int64_t C, const float *img, const float *gt,
                       float lambda_dssim, float *loss, void *workspace, size_t workspace_bytes, void *stream);

/* v_img [B,H,W,C] = *v_loss (device float) x d loss / d img, from the
 * workspace of the matching forward call. */
int ubs_image_loss_bwd(int64_t B, int64_t H, int64_t W, int64_t C, const float *img, const float *gt,
                       float lambda_dssim, const float *v_loss, const void *workspace, size_t workspace_bytes,
                       float *v_img, void *stream);

/* ------------------------------------------------------------------------
 * Optimiser and MCMC step after the backward.
 * ------------------------------------------------------------------------ */

/* Replaces torch.optim.Adam(groups, lr=0.0, eps=1e-15).step() over the
 * scene's leaf groups (scene/beta_model.py:239-276, train.py:170): one
 * launch updates up to 8 float32 tensors in place.  Host arrays of
 * n_tensors entries: device pointers of each parameter, its gradient and
 * its two moment buffers, element counts, learning rates and step counts
 * (the count after this update, >= 1; bias corrections use it).  Tensors
 * of size 0 are skipped. */
int ubs_adam_step(int n_tensors, float *const *params, const float *const *grads, float *const *exp_avgs,
                  float *const *exp_avg_sqs, const int64_t *sizes, const double *lrs, const int64_t *steps,
                  double beta1, double beta2, double eps, void *stream);

/* Replaces the covariance-shaped position noise of train.py:159-168
 * (get_xyz_covariance, scene/beta_model.py:143-152, then torch.bmm):
 * xyz[N,3] += Sigma_xyz (normal * (1 - sigmoid(opacity_raw))^100 * scale),
 * scale = noise_lr * xyz_lr; normal [N,3] holds the standard normal draws. */
int ubs_xyz_noise(int64_t N, int64_t D, float *xyz, const float *l_triangle, const float *scale_raw,
                  const float *opacity_raw, const float *normal, float scale, void *stream);

/* Replaces the row updates of relocate_gs / add_new_gs / _update_params
 * and replace_tensors_to_optimizer(inds) (scene/beta_model.py:506-658) for
 * K sampled (src, dst, ratio) triples: each leaf's dst row takes the src
 * row, the opacity leaf (one column, pre-activation) the split opacity
 * logit(clamp(1 - (1 - sigmoid(o))^(1/(ratio+1)), min_opacity, 1 - 2^-23));
 * then each src row takes that opacity and its Adam moments are zeroed
 * (exp_avgs / exp_avg_sqs may be null, or hold null entries).  dst rows must
 * not be src rows.  leaves/widths/exp_avgs/exp_avg_sqs are host arrays of
 * n_leaves (<= 8) entries; rows are contiguous float32 of width columns. */
int ubs_mcmc_relocate(int64_t K, int n_leaves, float *const *leaves, const int *widths, int opacity_leaf,
                      float *const *exp_avgs, float *const *exp_avg_sqs, const int64_t *src, const int64_t *dst,
                      const int64_t *ratio, float min_opacity, void *stream);

/* ------------------------------------------------------------------------
 * Process-wide numerical switches (no counterpart in the reference).
 * ------------------------------------------------------------------------ */

/* Blend backward: recover transmittance in fp64 (1) or fp32 (0, default). */
int ubs_set_backward_fp64_transmittance(int enable);

/* Conditioning backward: 0 (default) reproduces the reference's gradient,
 * whose Cholesky backward (cond_mean_convariance_opacity_bwd.cu:508-518)
 * solves with L where L^T is needed, so the gradient of the conditioned
 * opacity w.r.t. the query-query covariance block is not the true one;
 * 1 computes the exact gradient. */
int ubs_set_cond_backward_exact(int enable);

#ifdef __cplusplus
}
#endif

#endif /* UBS_CAPI_H */
