/*
 * libslamtb_hip.so -- C ABI of the MI355X (gfx950) ICP Gauss-Newton step.
 *
 * Replaces the CUDA half of the reference's pybind11 module `_cslamtb`
 * (src/_cslamtb.cpp:22-53). Conventions of every entry point:
 *   - all array pointers are device pointers to C-contiguous buffers that the
 *     caller owns; the library allocates nothing and keeps no state;
 *   - `stream` is a hipStream_t (NULL = the default stream); work is only
 *     enqueued, the library never synchronises the device;
 *   - the return value is STB_OK (0) or an error code, and stb_last_error()
 *     gives the message of the last failure on the calling thread;
 *   - 3x4 / 4x4 matrices are row-major; `rt_rows` is 3 or 4 (only the first
 *     three rows are read), `kcam` is a row-major 3x3 matrix.
 *   - masks are one byte per element (torch.bool), non-zero = valid.
 */
#ifndef SLAMTB_HIP_H_
#define SLAMTB_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum stb_status {
  STB_OK = 0,
  STB_ERR_INVALID = 1, /* bad argument: shape, dtype, null pointer */
  STB_ERR_LAUNCH = 2,  /* a kernel launch failed */
  STB_ERR_UNSUPPORTED = 3
};

enum stb_dtype {
  STB_F32 = 0,
  STB_F64 = 1,
  STB_BF16 = 2,
  STB_F16 = 3,
  STB_I32 = 4,
  STB_U8 = 5,
  STB_I64 = 6,
  STB_I16 = 7
};

/* Last error message of the calling thread ("" if none). */
const char* stb_last_error(void);
/* ABI version, bumped on any signature change. */
int stb_abi_version(void);

/* ------------------------------------------------------------------------ */
/* Drop-in entry points (same argument meaning as the reference bindings).   */
/* ------------------------------------------------------------------------ */

/* CorrespondenceMap.compute_correspondence_map
 * (include/slamtb/matching.hpp:11-21, src/correspondence_map.cu:64-102).
 * dtype: STB_F32 or STB_F64 for points, normals, rt_cam and kcam.
 * src_mask is applied only when src_mask_len == n_src (else ignored, as in
 * src/correspondence_map.cu:40-41). corresp_map [height*width] int64 holds
 * (source_index << 32) | float_bits(distance) and is min-merged in place;
 * among equal distances the lowest source index is kept. */
int stb_compute_correspondence_map(
    int dtype, const void* src_points, const void* src_normals,
    const uint8_t* src_mask, int64_t src_mask_len, int64_t n_src,
    const void* rt_cam, int rt_rows, const void* tgt_points,
    const void* tgt_normals, const uint8_t* tgt_mask, int32_t height,
    int32_t width, const void* kcam, int64_t* corresp_map,
    double distance_thresh, double normal_angle_thresh, void* stream);

/* ICPJacobian.estimate_geometric (include/slamtb/icp_jacobian.hpp:72-77,
 * src/icp_jacobian_geometric.cu:163-214). Target maps are float32 for both
 * dtypes (src/icp_jacobian_geometric.cu:22-23). Writes per-pixel
 * JtJ_partial [n,6,6], Jr_partial [n,6], squared_residual [n] in `dtype`;
 * *match_count (device int32) is zeroed and then counts the matches. */
int stb_icp_estimate_geometric(
    int dtype, const void* src_points, const uint8_t* src_mask, int64_t n_src,
    const void* rt_cam, int rt_rows, const float* tgt_points,
    const float* tgt_normals, int32_t height, int32_t width, const void* kcam,
    const int64_t* corresp_map, void* JtJ_partial, void* Jr_partial,
    void* squared_residual, int32_t* match_count, void* stream);

/* ICPJacobian.get_information_matrix (include/slamtb/icp_jacobian.hpp:79-85,
 * src/icp_jacobian_geometric.cu:216-249). GtG_partial [n,6,6] in `dtype`,
 * parameter order [rotation, translation]. */
int stb_icp_get_information_matrix(
    int dtype, const void* src_points, const uint8_t* src_mask, int64_t n_src,
    const void* rt_cam, int rt_rows, const float* tgt_points, int32_t height,
    int32_t width, const void* kcam, const int64_t* corresp_map,
    void* GtG_partial, void* stream);

/* ICPJacobian.estimate_feature (include/slamtb/icp_jacobian.hpp:140-146,
 * src/icp_jacobian_feature.cu:113-169). src_feats [channels, n_src],
 * tgt_feats [channels, height, width], both in `dtype`. As in the reference
 * the per-pixel J keeps only the last channel while the squared residual
 * sums every channel; pixels above residual_thresh^2 are zeroed, uncounted. */
int stb_icp_estimate_feature(
    int dtype, const void* src_points, const void* src_feats,
    const uint8_t* src_mask, int64_t n_src, const void* rt_cam, int rt_rows,
    const void* tgt_feats, int32_t channels, int32_t height, int32_t width,
    const void* kcam, const int64_t* corresp_map, float residual_thresh,
    void* JtJ_partial, void* Jr_partial, void* squared_residual,
    int32_t* match_count, void* stream);

/* Processing.bilateral_depth_filter (src/processing.cpp:9,
 * src/bilateral_depth_filter.cu:79-110). dtype: STB_F32, STB_F64, STB_I32,
 * STB_I16 or STB_U8 for input and result [height, width]. */
int stb_bilateral_depth_filter(int dtype, const void* input,
                               const uint8_t* mask, void* result,
                               int32_t height, int32_t width,
                               int32_t filter_width, float sigma_color,
                               float sigma_space, float depth_scale,
                               void* stream);

/* Processing.estimate_normals (src/processing.cpp:10, src/normals.cu:190-208).
 * method 0 = CentralDifferences, 1 = Average8 (unsupported here). */
int stb_estimate_normals(int dtype, const void* xyz, const uint8_t* mask,
                         void* out_normals, int32_t height, int32_t width,
                         int method, void* stream);

/* Processing.downsample_xyz (src/processing.cpp:11, src/downsample.cu:75-98).
 * `normalize` is accepted and ignored, as in the reference; method 0 =
 * Nearest. dst is [dst_height, dst_width, 3]. */
int stb_downsample_xyz(int dtype, const void* src, const uint8_t* mask,
                       int32_t height, int32_t width, float scale, void* dst,
                       int32_t dst_height, int32_t dst_width, int normalize,
                       int method, void* stream);

/* Processing.downsample_mask (src/processing.cpp:12, src/downsample.cu:119-131). */
int stb_downsample_mask(const uint8_t* mask, int32_t height, int32_t width,
                        float scale, uint8_t* dst, int32_t dst_height,
                        int32_t dst_width, void* stream);

/* Processing.downsample_features (src/processing.cpp:13,
 * src/downsample.cu:171-191). src [channels, height, width]. */
int stb_downsample_features(int dtype, const void* src, int32_t channels,
                            int32_t height, int32_t width, float scale,
                            void* dst, int32_t dst_height, int32_t dst_width,
                            void* stream);

/* Processing.erode_mask (src/processing.cpp:14, src/erode_mask.cu:42-54). */
int stb_erode_mask(const uint8_t* in_mask, uint8_t* out_mask, int32_t height,
                   int32_t width, void* stream);

/* Separable normalised 3x3 Gaussian blur with reflect borders over
 * [channels, height, width]; stands in for the kornia GaussianBlur2d call of
 * feature_pyramid (slamtb/processing.py:148-153). */
int stb_gaussian_blur3x3(int dtype, const void* src, int32_t channels,
                         int32_t height, int32_t width, double sigma,
                         void* dst, void* stream);

/* depth_image_to_uvz + KCamera.backproject (slamtb/_utils.py:11-37,
 * slamtb/camera.py:127-149): points [height, width, 3] float32 from a depth
 * image of type `depth_dtype` (STB_F32, STB_F64, STB_I32, STB_I16, STB_U8). */
int stb_depth_to_points(int depth_dtype, const void* depth, int32_t height,
                        int32_t width, float fx, float fy, float cx, float cy,
                        float depth_scale, float depth_bias, float* points,
                        void* stream);

/* ExpRtToMatrixOp.forward (src/se3.cpp:9, src/exp_rt_to_matrix.cu:41-52):
 * exp_rt [batch, 6] = [t; w] -> matrix [batch, 3, 4] = [R(w) | t]. */
int stb_exp_rt_to_matrix_forward(int dtype, const void* exp_rt, int64_t batch,
                                 void* matrix, void* stream);

/* ExpRtToMatrixOp.backward (src/se3.cpp:10, src/exp_rt_to_matrix.cu:104-121). */
int stb_exp_rt_to_matrix_backward(int dtype, const void* d_matrix,
                                  const void* exp_rt, const void* y_matrix,
                                  int64_t batch, void* d_exp_rt, void* stream);

/* MatrixToExpRtOp.forward (src/se3.cpp:14, src/matrix_to_exp_rt.cu:38-50). */
int stb_matrix_to_exp_rt_forward(int dtype, const void* matrix, int64_t batch,
                                 void* exp_rt, void* stream);

/* FPCLMatcherOp.forward (src/matching.cpp:8, src/fpcl_matcher.cu:
 * 85-123): for each of n_src source points, projected with kcam, the target
 * pixel that corresp_map (reference layout, [height, width] int64) assigns to
 * that point gives out_points/out_normals [n_src, 3] and the bilinear
 * feature sample out_feats [channels, n_src]; match_mask [n_src] (bytes)
 * flags the matched points. Unmatched rows are written as zeros (the
 * reference leaves them unwritten). */
int stb_fpcl_match_forward(int32_t dtype, const void* tgt_points,
                           const void* tgt_normals, const void* tgt_feats,
                           int32_t channels, int32_t height, int32_t width,
                           const void* src_points, int64_t n_src,
                           const void* kcam, const int64_t* corresp_map,
                           void* out_points, void* out_normals,
                           void* out_feats, uint8_t* match_mask, void* stream);

/* FPCLMatcherOp.backward (src/matching.cpp:9, src/fpcl_matcher.cu:
 * 187-229): dx_points [n_src, 3] = Dx_Project^T (sum_c dI_c/du dl_c,
 * sum_c dI_c/dv dl_c) for matched points, 0 otherwise; the feature
 * derivatives are central differences of step grad_precision, which must
 * lie in (0, 0.5) so that every shifted sample stays inside the image; a
 * flagged point whose rounded projection is outside the image gets 0. */
int stb_fpcl_match_backward(int32_t dtype, const void* tgt_feats,
                            int32_t channels, int32_t height, int32_t width,
                            const void* src_points, int64_t n_src,
                            const uint8_t* match_mask, const void* dl_feats,
                            const void* kcam, double grad_precision,
                            void* dx_points, void* stream);

/* ProjectOp.forward (src/camera.cpp:11, src/camera.cu:34-60): points [n, 3]
 * -> uv [n, 2] with u = fx x / z + cx, v = fy y / z + cy; `intrinsics` is a
 * row-major 3x3 or 2x3 matrix. */
int stb_project_forward(int32_t dtype, const void* points, int64_t n,
                        const void* intrinsics, void* out_uv, void* stream);

/* ProjectOp.backward (src/camera.cpp:12, src/camera.cu:89-117). */
int stb_project_backward(int32_t dtype, const void* dy_uv, const void* points,
                         int64_t n, const void* intrinsics, void* dx_points,
                         void* stream);

/* RigidTransformOp.transform_inplace (src/camera.cpp:43, src/camera.cu:
 * 134-174): out[n,3] = R p + t for points[n,3]; matrix is the device-resident
 * row-major 3x4 block (rows 0..2 of a [3,4] or [4,4] tensor, row stride 4).
 * out may equal points (in place). Also the forward pass of RigidTransform's
 * `@` (slamtb/camera.py:333-337). */
int stb_rigid_transform_points(int32_t dtype, const void* matrix,
                               const void* points, int64_t n, void* out,
                               void* stream);

/* RigidTransformOp.transform_normals_inplace (src/camera.cpp:44,
 * src/camera.cu:176-240): out[n,3] = N n with N the cofactor
 * inverse-transpose of R (include/slamtb/camera.hpp:99-116). */
int stb_rigid_transform_normals(int32_t dtype, const void* matrix,
                                const void* normals, int64_t n, void* out,
                                void* stream);

/* Exact k-nearest-neighbour search, the device replacement of the host
 * k-d tree query of KDTreeLayer.query (slamtb/spatial/kdtree_layer.py:40-60,
 * pykdtree KDTree.query with sqr_dists=True): for each of n query points the
 * k (1..16) nearest of m target points ([m,3] / [n,3], dtype STB_F32 or
 * STB_F64) with squared distance strictly below distance_upper_bound^2
 * (pass +inf for no bound), sorted by squared distance and then by target
 * index; out_sq_dist [n,k] and out_index [n,k] hold (inf, m) in unused
 * slots. Exact brute force (the reference's k-d tree searched approximately,
 * eps = 0.5). At most 2^28 queries per call. */
int stb_knn(int32_t dtype, const void* target, int64_t m, const void* query,
            int64_t n, int32_t k, double distance_upper_bound,
            void* out_sq_dist, int64_t* out_index, void* stream);

/* Uniform grid for stb_knn_grid: the cell of each of m points, (z * ny + y)
 * * nx + x with x = floor((p.x - lo_x) / cell) etc., or nx * ny * nz for a
 * point outside the grid or not finite. */
int stb_grid_cells(int32_t dtype, const void* points, int64_t m, double lo_x,
                   double lo_y, double lo_z, double cell, int32_t nx, int32_t ny,
                   int32_t nz, int32_t* out_cells, void* stream);

/* stb_knn with a finite distance bound over a uniform grid (cell >= bound):
 * entries [m,4] holds the targets stably sorted by cell as (x, y, z, index),
 * the index's int32 bits in the fourth float for STB_F32 and its value for
 * STB_F64; cell_start [nx*ny*nz + 1] the first sorted position of each cell.
 * query_order (device [n] permutation of the queries, or NULL) is the order
 * in which threads take the queries: ordered by grid cell (stb_grid_cells
 * of the queries, stably sorted), a wavefront's queries scan the same cells
 * and its candidate loads coalesce. Each query's row of the outputs is its
 * own either way. Same result as stb_knn (ties broken by original index).
 * m must be below 2^31 (the search keeps 31-bit indices); at most 2^28
 * queries per call. */
int stb_knn_grid(int32_t dtype, const void* entries, const int64_t* cell_start,
                 int32_t nx, int32_t ny, int32_t nz,
                 double lo_x, double lo_y, double lo_z, double cell, int64_t m,
                 const void* query, int64_t n, const int64_t* query_order,
                 int32_t k, double distance_upper_bound, void* out_sq_dist,
                 int64_t* out_index, void* stream);

/* Test hook of stb_knn_grid, no counterpart in the reference: calls with at
 * most max_split_queries queries split each query's candidates over four
 * lanes, larger calls give each query one lane (default 2^20). Returns the
 * previous value. Both give the same results; tests lower it to run the
 * one-lane form on small inputs. */
int64_t stb_knn_grid_set_lane_threshold(int64_t max_split_queries);

/* NearestNeighborsOp.forward (src/nearest_neighbors.cpp:7,
 * src/nearest_neighbors.cu:13-79): out_features [channels, n] = sum over the
 * neighbours of features[:, index] * w, w = (1 / (1 + d)) / sum of 1 / (1 + d)
 * over the finite distances; features is [channels, m]. */
int stb_nn_features_forward(int32_t dtype, const void* nn_distances,
                            const int64_t* nn_index, int64_t n, int32_t k,
                            const void* features, int32_t channels, int64_t m,
                            void* out_features, void* stream);

/* NearestNeighborsOp.compute_epsilon_distances (src/nearest_neighbors.cpp:9,
 * src/nearest_neighbors.cu:86-164): epsilon_distances [n, 6, k], the
 * normalised weights 1 / (1 + |target - (source +- h e_axis)|^2) of each
 * query's neighbours for the six offsets (h = 0.0005 in the reference, as a
 * float). */
int stb_nn_epsilon_distances(int32_t dtype, const void* target_xyz, int64_t m,
                             const void* source_xyz, const int64_t* nn_index,
                             int64_t n, int32_t k, double h,
                             void* epsilon_distances, void* stream);

/* NearestNeighborsOp.backward (src/nearest_neighbors.cpp:8,
 * src/nearest_neighbors.cu:166-295): dl_xyz [n,3] = sum over channels of the
 * central difference (step h) of the weighted feature average times
 * dl_features [channels, n]. */
int stb_nn_features_backward(int32_t dtype, const void* epsilon_distances,
                             const int64_t* nn_index, int64_t n, int32_t k,
                             const void* features, int32_t channels, int64_t m,
                             const void* dl_features, double h, void* dl_xyz,
                             void* stream);

/* Voxel id of each point (HashVolume::Hash, include/slamtb/hash_volume.hpp:
 * 67-77): int((p - min) / voxel_size) per axis, combined as
 * depth * width * height + row * width + col in int arithmetic, with width
 * and height the truncated extents of the volume. */
int stb_voxel_ids(const float* points, int64_t n, float min_x, float min_y,
                  float min_z, float voxel_size, int32_t width, int32_t height,
                  int32_t* out_ids, void* stream);

/* SurfelVolume.merge + to_surfel_cloud (src/surfel_volume.cpp:13-14,
 * src/surfel_volume.cu:12-123): per voxel the mean of its members' points,
 * normals, colours (truncated to uint8), confidences, radii and features
 * [channels, n]; times are 0, as the reference leaves them. `order` lists the
 * point indices stably sorted by voxel id and `starts` [num_voxels + 1] the
 * start of each voxel's run in it. */
int stb_voxel_merge(const float* points, const float* normals,
                    const uint8_t* colors, const float* confidences,
                    const float* radii, const float* features, int32_t channels,
                    int64_t n, const int64_t* order, const int64_t* starts,
                    int64_t num_voxels, float* out_points, float* out_normals,
                    uint8_t* out_colors, float* out_confidences, float* out_radii,
                    int32_t* out_times, float* out_features, void* stream);

/* Backward of out = R p + t (no native counterpart: the reference
 * differentiates its torch matmul, slamtb/camera.py:333-337):
 * dx_points[n,3] = R^T dy (skipped when NULL) and d_matrix[3,4] =
 * sum_i dy_i [p_i; 1]^T, summed in FP64 in a fixed order. workspace is device
 * memory of at least stb_rigid_backward_workspace_bytes() bytes. */
size_t stb_rigid_backward_workspace_bytes(void);
int stb_rigid_transform_backward(int32_t dtype, const void* matrix,
                                 const void* points, int64_t n, const void* dy,
                                 void* dx_points, void* d_matrix,
                                 void* workspace, size_t workspace_bytes,
                                 void* stream);

/* ------------------------------------------------------------------------ */
/* Fused Gauss-Newton path used by slamtb.registration.ICPOdometry.          */
/* It replaces the per-iteration sequence of icp.py:210-256 (two maps, two   */
/* per-pixel partial kernels, torch sums, host Cholesky and SE3 update)      */
/* with two kernels per iteration and no host round trip.                    */
/* ------------------------------------------------------------------------ */

/* One pyramid level of one frame pair. Storage dtype (STB_F32, STB_BF16,
 * STB_F16) applies to points, normals and features; masks are bytes.
 * Source arrays are [n_src, 3] / [channels, n_src]; target arrays are
 * [height, width, 3] / [channels, height, width]. Set src_feats/tgt_feats to
 * NULL (or channels to 0) for a geometric-only level.
 * feat_layout STB_FEAT_INTERLEAVED stores the features channel-innermost
 * instead: source [n_src, channels], target [height, width, channels]
 * (the reference's feature maps are channel-planar; this layout has no
 * counterpart there). A texel's channels are then read 16 bytes at a time:
 * with BF16 coordinates, whose 8-bit mantissa scatters the projections of
 * neighbouring pixels, that touches far fewer cache lines per pixel than
 * one line per channel and tap. Results are bit-identical to the planar
 * layout. It needs 16-bit storage and an even channel count of at most 32,
 * and only the fused Gauss-Newton entry points (single-pair and batched)
 * accept it. */
#define STB_FEAT_PLANAR 0
#define STB_FEAT_INTERLEAVED 1

typedef struct stb_icp_level {
  const void* src_points;
  const void* src_normals;
  const uint8_t* src_mask;
  const void* src_feats;
  const void* tgt_points;
  const void* tgt_normals;
  const uint8_t* tgt_mask;
  const void* tgt_feats;
  int64_t n_src;
  int32_t height;
  int32_t width;
  int32_t channels;
  int32_t storage;
  float fx, fy, cx, cy;
  int32_t feat_layout;  /* STB_FEAT_PLANAR (0) or STB_FEAT_INTERLEAVED */
  int32_t reserved;     /* 0 */
} stb_icp_level;

/* Options of ICPOdometry (slamtb/registration/icp.py:137-150). */
typedef struct stb_icp_params {
  double geom_weight;
  double feat_weight;
  double distance_thresh;
  double normal_angle_thresh;
  float feat_residual_thresh;
  int32_t use_geom;     /* geom_weight > 0 */
  int32_t use_feat;     /* feat_weight > 0 and both feature maps given */
  int32_t feat_mode;    /* 0: last channel's J (reference), 1: all channels */
  int32_t grid_update;  /* workgroups of the Jacobian kernel (0 = default) */
  int32_t grid_corresp; /* workgroups of the correspondence kernel (0 = default) */
} stb_icp_params;

/* Device-resident result, at offset 0 of the workspace. Fields mirror
 * RegistrationResult (slamtb/registration/result.py:9-46). */
typedef struct stb_icp_state {
  double transform[16];   /* current pose, row-major 4x4 */
  double best_transform[16];
  double best_jtj[36];    /* combined JtJ of the last accepted iteration */
  double information[36]; /* G^T G at the final pose */
  double best_residual;   /* residual / match_count */
  double best_match_count;
  double last_jtj[36];    /* combined system of the most recent iteration */
  double last_jr[6];
  double last_geom_r2, last_feat_r2, last_geom_count, last_feat_count;
  float rt[12];           /* float copy of the pose rows (transform.to(f32)) */
  float nm[9];            /* inverse-transpose of rt's rotation */
  int32_t stopped;        /* 1 after a failed Cholesky (the loop's `break`) */
  int32_t iterations;     /* accepted iterations */
  int32_t map_epoch;      /* key tag the next iteration writes (255..0) */
  int32_t map_sel;        /* correspondence map the next iteration uses */
} stb_icp_state;

/* sizeof(stb_icp_state) as compiled into the library (for bindings that
 * mirror the struct). No counterpart in the reference. */
size_t stb_icp_state_bytes(void);

/* Bytes of workspace needed for a level of n_src source and height*width
 * target pixels (state, two correspondence maps, block partials, the loop's
 * solve record). No counterpart in the reference. */
size_t stb_icp_workspace_bytes(int64_t n_src, int32_t height, int32_t width);

/* Resets the state to the 4x4 row-major FP64 pose `transform0` (device
 * pointer, NULL = identity) and clears both correspondence maps. With
 * `chain` != 0 the start pose is instead the workspace's own best_transform
 * when its previous run accepted at least one iteration, else identity: the
 * hand-over between pyramid levels of MultiscaleRegistration.estimate_frame
 * (slamtb/registration/multiscale.py:52-65). */
int stb_icp_begin(void* workspace, size_t workspace_bytes,
                  const double* transform0, int32_t chain, int64_t n_src,
                  int32_t height, int32_t width, void* stream);

/* Enqueues `num_iters` Gauss-Newton iterations in 2 * num_iters + 1
 * launches. Each iteration is a correspondence launch and a Jacobian launch
 * that leaves one partial per block (summed over the block in FP64, stored
 * in FP32); the partials are summed in FP64 in a fixed order, the 6x6
 * system solved by Cholesky in FP64 and exp_rt_to_matrix(x)
 * left-multiplied into the pose by the next iteration's correspondence
 * launch (and, after the last iteration, by one solve launch). A failed
 * factorisation sets `stopped`, after which the remaining iterations change
 * nothing. The state a loop leaves is bit-identical to that of the same
 * number of stb_icp_correspond + stb_icp_update pairs. Which of the two
 * correspondence maps an iteration uses is kept in the state, so a captured
 * graph of any number of iterations replays correctly. */
int stb_icp_iterate(const stb_icp_level* level, const stb_icp_params* params,
                    void* workspace, size_t workspace_bytes, int32_t num_iters,
                    void* stream);

/* The two halves of one iteration, for callers that schedule or time them
 * separately: stb_icp_correspond builds the correspondence map at the
 * current pose; stb_icp_update reduces the Jacobians over it, solves and
 * updates the pose. correspond + update == one stb_icp_iterate step. */
int stb_icp_correspond(const stb_icp_level* level, const stb_icp_params* params,
                       void* workspace, size_t workspace_bytes, void* stream);
int stb_icp_update(const stb_icp_level* level, const stb_icp_params* params,
                   void* workspace, size_t workspace_bytes, void* stream);

/* The Jacobian launch of one iteration alone, over the correspondence map
 * of the last stb_icp_linearize call (a loop's own map holds tagged keys
 * that only its iterations read): per-block partials into the workspace,
 * with no reduction, no pose update and both maps untouched, so that
 * repeated calls do identical work. For timing the dominant launch of an
 * iteration on its own. */
int stb_icp_jacobians(const stb_icp_level* level, const stb_icp_params* params,
                      void* workspace, size_t workspace_bytes, void* stream);

/* One linearisation at the current pose without a pose update: fills
 * last_jtj / last_jr / last_*_r2 / last_*_count with the reduced system of
 * the active terms (weights and the feature term's sign applied as in the
 * update). The fused counterpart of one _JacobianStep call per term
 * (slamtb/registration/icp.py:30-82). */
int stb_icp_linearize(const stb_icp_level* level, const stb_icp_params* params,
                      void* workspace, size_t workspace_bytes, void* stream);

/* Fills state.information with sum of G^T G at the current pose
 * (src/icp_jacobian_geometric.cu:88-159 summed). */
int stb_icp_information(const stb_icp_level* level,
                        const stb_icp_params* params, void* workspace,
                        size_t workspace_bytes, void* stream);

/* No-allocation variant of stb_icp_begin + stb_icp_iterate (+
 * stb_icp_information) for one level. No counterpart in the reference: it
 * is for callers that cannot allocate device memory or pass a workspace
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The workspace is a static device array of
 * stb_icp_noalloc_scratch_bytes() (16 MiB) that every GPU holds until the
 * process ends, which bounds the target to about 1280x720 pixels (larger
 * targets are refused with STB_ERR_INVALID). Resets the pose to
 * `transform0` (device 4x4 FP64, NULL = identity), runs `num_iters`
 * iterations with the same kernels and launch shapes as the workspace entry
 * points (the same results, bit for bit), adds the information matrix when
 * `with_information` and the geometric term is on, and copies the final
 * state to `out_state` (device memory, one stb_icp_state). Allocates
 * nothing. Every call on one GPU shares that GPU's array and each call
 * starts by clearing what it reads: calls that overlap on one GPU (other
 * streams, other host threads) race and must be serialised by the caller. */
size_t stb_icp_noalloc_scratch_bytes(void);
int stb_icp_estimate_noalloc(const stb_icp_level* level,
                             const stb_icp_params* params,
                             const double* transform0, int32_t num_iters,
                             int32_t with_information, stb_icp_state* out_state,
                             void* stream);

/* ------------------------------------------------------------------------ */
/* Batched Gauss-Newton over many frame pairs of uneven resolution.          */
/* No reference counterpart: the reference registers one pair per call      */
/* (slamtb/registration/multiscale.py:30-67); this runs every pair of a      */
/* batch through the same two persistent launches per iteration.            */
/* ------------------------------------------------------------------------ */

typedef struct stb_icp_batch_sizes {
  int32_t num_pairs;
  int32_t total_tiles;
  int32_t tile_px;      /* source pixels per tile (multiple of 256) */
  int32_t storage;      /* storage dtype shared by every pair */
  int64_t total_map;    /* sum of target pixels over the pairs */
  size_t plan_bytes;    /* size of the plan image */
  size_t scratch_bytes; /* device scratch: two maps per pair, tile partials */
  int32_t max_channels; /* most feature channels of any pair */
  int32_t feat_layout;  /* shared by every pair (stb_icp_level.feat_layout) */
} stb_icp_batch_sizes;

/* Host only. Builds the plan image of one pyramid level for `num_pairs`
 * pairs (levels[p] as in stb_icp_level, device pointers inside) cut into
 * tiles of `tile_px` source pixels, and fills `sizes`. With plan_host ==
 * NULL only `sizes` is filled. The caller copies the image to device memory
 * and passes that device copy as `plan` below. */
int stb_icp_batch_plan(int32_t num_pairs, const stb_icp_level* levels,
                       int32_t tile_px, void* plan_host, size_t plan_capacity,
                       stb_icp_batch_sizes* sizes);

/* `states` is a device array of num_pairs stb_icp_state. Resets every pair
 * to transforms0[p] (device [num_pairs, 4, 4] FP64, NULL = identity) or,
 * with `chain`, to its own last accepted pose; clears the maps. */
int stb_icp_batch_begin(const void* plan, const stb_icp_batch_sizes* sizes,
                        void* states, const double* transforms0, int32_t chain,
                        void* scratch, size_t scratch_bytes, void* stream);

/* num_iters iterations of every pair: per iteration one correspondence and
 * one Jacobian/reduction launch over all tiles; the workgroup completing a
 * pair solves and updates that pair. A pair whose factorisation fails stops
 * while the others continue. */
int stb_icp_batch_iterate(const void* plan, const stb_icp_batch_sizes* sizes,
                          const stb_icp_params* params, void* states,
                          void* scratch, size_t scratch_bytes,
                          int32_t num_iters, void* stream);

/* Per-pair stb_icp_linearize / stb_icp_information over a batch. */
int stb_icp_batch_linearize(const void* plan, const stb_icp_batch_sizes* sizes,
                            const stb_icp_params* params, void* states,
                            void* scratch, size_t scratch_bytes, void* stream);
int stb_icp_batch_information(const void* plan,
                              const stb_icp_batch_sizes* sizes,
                              const stb_icp_params* params, void* states,
                              void* scratch, size_t scratch_bytes,
                              void* stream);

/* ------------------------------------------------------------------------ */
/* Adjoint of one Gauss-Newton step (differentiable ICP). No reference       */
/* counterpart; the reference's differentiable registration is AutogradICP   */
/* (slamtb/registration/autogradicp.py:95-374), built on the same            */
/* correspondence and exp_rt backward ops.                                   */
/* ------------------------------------------------------------------------ */

/* Per-pair gradient buffers (device pointers, FP32, accumulated into). */
typedef struct stb_icp_grads {
  float* d_src_points;  /* [n_src, 3] */
  float* d_src_feats;   /* [channels, n_src] */
  float* d_tgt_points;  /* [height * width, 3] */
  float* d_tgt_normals; /* [height * width, 3] */
  float* d_tgt_feats;   /* [channels, height * width] */
} stb_icp_grads;

/* With every pair's pose reset to T_k (stb_icp_batch_begin), rebuilds the
 * correspondences at T_k and, for each pair p with coef[p] = [lambda (6);
 * x (6)] (lambda = A_k^-1 dL/dx_k, x = the step's solution), adds the
 * pixels' contributions to grads[p] and writes dL/dT_k (3x4 rows, the part
 * that flows through the transformed source points) to dT[p] (12 doubles,
 * fixed-order FP64 sum). FP32 storage only. `grads` is a device array of
 * num_pairs structs. */
int stb_icp_batch_adjoint(const void* plan, const stb_icp_batch_sizes* sizes,
                          const stb_icp_params* params, void* states,
                          const stb_icp_grads* grads, const double* coef,
                          double* dT, void* scratch, size_t scratch_bytes,
                          void* stream);

/* The normal equations of one Gauss-Newton step of every pair at its current
 * pose (reset with stb_icp_batch_begin), recomputed for the adjoint: the
 * same correspondences, feature gate and bilinear cells as the forward's
 * FP32 arithmetic, every smooth quantity (transformed point, u and v,
 * bilinear fractions, J, r) and every sum in FP64. systems [num_pairs, 27]
 * holds the weighted JtJ upper triangle (21, row-major) and the weighted Jr
 * (6), combined as stb_icp_iterate combines them. FP32 storage only. No
 * reference counterpart: the forward's FP32 per-pixel positions round u
 * and v by up to 3e-5 px at 640x480, which moves the sampled feature
 * derivatives by ~1e-4, and the adjoint's solves need the system without
 * that rounding. */
int stb_icp_batch_system64(const void* plan, const stb_icp_batch_sizes* sizes,
                           const stb_icp_params* params, void* states,
                           double* systems, void* scratch, size_t scratch_bytes,
                           void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SLAMTB_HIP_H_ */
