/*
 * meshtryoshka_hip.h - C ABI of the MI355X (gfx950) extract -> rasterize -> composite
 * kernels.
 *
 * Conventions (every entry point):
 *   - Pointers are device pointers to contiguous row-major arrays owned by the
 *     caller. Native code never allocates or frees device memory: operations that
 *     need scratch space take a caller-owned workspace sized by the matching
 *     *_workspace_bytes query.
 *   - Element counts are int64_t. Array shapes are stated per argument.
 *   - All work is enqueued on `stream` (a hipStream_t; NULL = the default stream).
 *     Nothing synchronises the host.
 *   - The return value is MT_OK (0) or an MT_ERR_* code; mt_get_last_error_string()
 *     describes the last failure on the calling host thread.
 *
 * Each entry point cites the reference interface it replaces (path:line inside the
 * reference repository dcharatan/meshtryoshka). The reference binds Slang kernels by
 * keyword (slangtorch `launchRaw`) and two CUB pybind functions; here the same
 * operations are plain C functions.
 */
#ifndef MESHTRYOSHKA_HIP_H_
#define MESHTRYOSHKA_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* mt_stream_t;

enum {
  MT_OK = 0,
  MT_ERR_INVALID_ARGUMENT = 1,
  MT_ERR_HIP = 2,
  MT_ERR_WORKSPACE_TOO_SMALL = 3,
};

/* Bit layout of a 64-bit sort key (triangle_rasterization/keys.py:4-22): offsets and
 * lengths of the camera, shell and tile fields; the depth field is always 32 bits. */
typedef struct mt_key_layout {
  int32_t camera_offset, camera_bits;
  int32_t shell_offset, shell_bits;
  int32_t tile_offset, tile_bits;
  int32_t depth_offset;
} mt_key_layout;

/* Human-readable description of the last error on this host thread. */
const char* mt_get_last_error_string(void);

/* Library version (major * 10000 + minor * 100 + patch). */
int32_t mt_version(void);

/* ---------------------------------------------------------------------------------
 * Prefix sums
 * ------------------------------------------------------------------------------- */

/* Workspace for the single-pass (decoupled look-back) scans below, for n elements. */
size_t mt_scan_workspace_bytes(int64_t n);

/* In-place exclusive prefix sum of int32 data[n] (int32 wrap-around).
 * Replaces compute_exclusive_cumsum (triangle_extraction/compute_exclusive_cumsum.cu:8-39,
 * bound at compute_exclusive_cumsum.py:19-21). */
int mt_exclusive_scan_i32(int32_t* data, int64_t n, void* workspace, size_t workspace_bytes,
                          mt_stream_t stream);

/* Largest n accepted by mt_exclusive_scan_i32_noalloc (2^28 elements = 16384 tiles). */
#define MT_SCAN_NOALLOC_MAX_ELEMENTS (1LL << 28)

/* mt_exclusive_scan_i32 without a workspace and without any device allocation: the
 * look-back state (128 KiB + 256 B) lives in a device array inside the library, one copy
 * per GPU, held until the process ends. No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). n <= MT_SCAN_NOALLOC_MAX_ELEMENTS, else
 * MT_ERR_INVALID_ARGUMENT. `stream` must belong to the current device. Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (concurrent streams or host
 * threads) race and must be serialised by the caller. Each call clears the state it uses
 * with a fill kernel before scanning. */
int mt_exclusive_scan_i32_noalloc(int32_t* data, int64_t n, mt_stream_t stream);

/* out[i] = sum_{j <= i} in[j] as int64. Replaces the int64 torch.cumsum of per-(camera,
 * face) tile counts in generate_keys (triangle_rasterization/generate_keys.py:48). */
int mt_inclusive_scan_i32_to_i64(const int32_t* in, int64_t* out, int64_t n, void* workspace,
                                 size_t workspace_bytes, mt_stream_t stream);

/* ---------------------------------------------------------------------------------
 * Voxel structure (triangle_extraction/__init__.py:31-62)
 * ------------------------------------------------------------------------------- */

/* occupancy int32[I, J, KP] (bit b of word (i, j, kp) = cell (i, j, 32 kp + b)).
 * Writes vertex_occupancy int32[I+1, J+1, KP+1] and vertex_counts int32[(I+1)(J+1)(KP+1)]
 * (popcounts). Replaces compute_vertex_occupancy.slang:9-68 (.py:34-61). */
int mt_compute_vertex_occupancy(const int32_t* occupancy, int32_t I, int32_t J, int32_t KP,
                                int32_t* vertex_occupancy, int32_t* vertex_counts,
                                mt_stream_t stream);

/* Writes voxel_counts int32[(I+1)(J+1)(KP+1)] (popcounts of occupancy, 0 on the padded
 * high faces). Replaces count_occupancy.slang:3-20 (.py:26-44). */
int mt_count_occupancy(const int32_t* occupancy, int32_t I, int32_t J, int32_t KP,
                       int32_t* voxel_counts, mt_stream_t stream);

/* Builds the Voxels structure from the exclusive offsets of the two count arrays.
 * vertices f32[num_vertices, 3]; neighbors int32[7, num_voxels];
 * lower_corners int32[4, num_vertices]; upper_corners int32[4, num_voxels];
 * indices int32[num_voxels]. Replaces create_voxels.slang:3-276 (.py:47-96). */
int mt_create_voxels(const int32_t* occupancy, const int32_t* voxel_offsets,
                     const int32_t* vertex_occupancy, const int32_t* vertex_offsets, int32_t I,
                     int32_t J, int32_t KP, int64_t num_voxels, int64_t num_vertices,
                     float* vertices, int32_t* neighbors, int32_t* lower_corners,
                     int32_t* upper_corners, int32_t* indices, mt_stream_t stream);

/* ---------------------------------------------------------------------------------
 * Marching cubes over level sets (triangle_extraction/__init__.py:75-157)
 * ------------------------------------------------------------------------------- */

/* counts int32[S * num_columns] (level-set major): crossings of the column's owned
 * +x/+y/+z edges. Replaces countTriangleVertices (count_triangle_vertices.slang:3-37,
 * .py:23-50). lower_corners int32[4, num_columns]; sdf f32[num grid vertices]. */
int mt_count_triangle_vertices(const float* sdf, const int32_t* lower_corners,
                               int64_t num_columns, const float* level_sets, int32_t S,
                               int32_t* counts, mt_stream_t stream);

/* mt_count_triangle_vertices that also writes axes uint8[S * num_columns]: bit a set when
 * the column's axis-a edge crosses the level set (the column then has a vertex of type a,
 * at its first vertex plus the set bits below a). Batched training step; no counterpart in
 * the reference (its create_triangle_faces.slang:3-78 searches the vertex types instead):
 * mt_create_triangle_faces_axes_i64 reads these bits. */
int mt_count_triangle_vertices_axes(const float* sdf, const int32_t* lower_corners,
                                    int64_t num_columns, const float* level_sets, int32_t S,
                                    int32_t* counts, uint8_t* axes, mt_stream_t stream);

/* Forward of CreateTriangleVertices (create_triangle_vertices.slang:145-199,
 * .py:31-102). vertex_offsets int32[S * num_columns + 1] (exclusive). grid_sh
 * f32[num_sh, num_grid_vertices, 3]. Outputs f32[V, 3], f32[V], f32[num_sh, V, 3],
 * uint8[V] with V = num_triangle_vertices, the capacity of the output arrays (the
 * plain path passes the exact count). Vertices at or beyond the capacity are not
 * written; if any exist and `overflow` is not NULL, *overflow is set to 1. */
int mt_create_triangle_vertices_forward(
    const float* grid_vertices, const float* grid_sdf, const float* grid_sh,
    int64_t num_grid_vertices, int32_t num_sh, const int32_t* lower_corners,
    int64_t num_columns, const float* level_sets, int32_t S, const int32_t* vertex_offsets,
    int64_t num_triangle_vertices, float* out_vertices, float* out_sdf, float* out_sh,
    uint8_t* out_types, int32_t* overflow, mt_stream_t stream);

/* Workspace for the atomic-free path of mt_create_triangle_vertices_backward. */
size_t mt_create_triangle_vertices_backward_workspace_bytes(int64_t num_grid_vertices);

/* Adjoint of the forward (create_triangle_vertices.slang:201-291, .py:104-170) into
 * zero-initialised grid gradients. Any grad_* input may be NULL (no gradient flows from
 * that output); any grid_*_grad output may be NULL (not requested). With a workspace of
 * at least the queried size and num_sh in {1, 4, 9, 16}, every grid vertex gathers its
 * gradient from the columns that reference it (deterministic, no atomics); otherwise, or
 * if two columns share an origin or an axis end, the columns scatter with atomics. */
int mt_create_triangle_vertices_backward(
    const float* grid_sdf, int64_t num_grid_vertices, int32_t num_sh,
    const int32_t* lower_corners, int64_t num_columns, const float* level_sets, int32_t S,
    const int32_t* vertex_offsets, int64_t num_triangle_vertices, const float* grad_vertices,
    const float* grad_sdf, const float* grad_sh, float* grid_vertices_grad,
    float* grid_sdf_grad, float* grid_sh_grad, void* workspace, size_t workspace_bytes,
    mt_stream_t stream);

/* Batched training step (no single counterpart in the reference: fuses
 * create_triangle_vertices.slang:145-199 with evaluate_spherical_harmonics.slang:187-226).
 * Writes out_vertices / out_sdf / out_types exactly like
 * mt_create_triangle_vertices_forward and out_colors f32[B, num_triangle_vertices, 3] for
 * the B cameras centred at camera_positions f32[B, 3]; each triangle vertex's SH
 * coefficients are interpolated from grid_sh in registers and never stored. Colours are
 * bit-identical to mt_create_triangle_vertices_forward + mt_evaluate_sh_forward. parts: 1 =
 * the vertices only, 2 = the colours only (the vertices are recomputed, not written), 3 =
 * both; split, the colours can run on another stream beside the work that needs only the
 * vertices. Adding 4 to parts reads grid_sh as SH rows (mt_pack_sh_rows' layout, 16-byte
 * aligned) instead of [num_sh, num_grid_vertices, 3]: the colours are the same bits.
 * colors_stride 0 writes out_colors as f32[B, num_triangle_vertices, 3]; s >= B writes
 * f32[num_triangle_vertices, s, 3] (a vertex's cameras together: one contiguous record per
 * vertex, which the batched step's colour adjoint reads once). column_words (may be NULL):
 * also writes uint32[S * num_columns], word s * num_columns + c = (vertex_offsets[s *
 * num_columns + c] << 3) | the column's crossing axes at level set s (bit a: axis a), for
 * mt_create_triangle_faces_words_i64; needs num_triangle_vertices <=
 * MT_COLUMN_WORD_MAX_VERTICES. */
#define MT_COLUMN_WORD_MAX_VERTICES (1LL << 29)
int mt_create_triangle_vertices_colors_forward(
    const float* grid_vertices, const float* grid_sdf, const float* grid_sh,
    int64_t num_grid_vertices, int32_t num_sh, int32_t sh_degree, const int32_t* lower_corners,
    int64_t num_columns, const float* level_sets, int32_t S, const int32_t* vertex_offsets,
    int64_t num_triangle_vertices, const float* camera_positions, int32_t B, float* out_vertices,
    float* out_sdf, uint8_t* out_types, float* out_colors, int32_t* overflow, int32_t parts,
    int32_t colors_stride, uint32_t* column_words, mt_stream_t stream);

/* Floats per SH row for an SH degree (0..3; -1 otherwise): 3 (deg+1)^2 rounded up to a
 * multiple of 4. No counterpart in the reference (a layout of this library). */
int64_t mt_sh_row_floats(int32_t sh_degree);

/* grid_sh f32[num_sh, num_grid_vertices, 3] (the reference's spherical_harmonics layout,
 * meshtryoshka/model/tesseract.py:239) -> rows f32[num_grid_vertices, mt_sh_row_floats]:
 * each grid vertex's first (sh_degree+1)^2 coefficients x 3 channels contiguous
 * (coefficient-major, then channel), zero-padded. No counterpart in the reference: the
 * fused colour pass (parts | 4) gathers one row per edge end instead of one entry from
 * each coefficient plane. */
int mt_pack_sh_rows(const float* grid_sh, int64_t num_grid_vertices, int32_t num_sh,
                    int32_t sh_degree, float* rows, mt_stream_t stream);

/* Workspace for mt_create_triangle_vertices_colors_backward. */
size_t mt_create_triangle_vertices_colors_backward_workspace_bytes(int64_t num_grid_vertices,
                                                                   int64_t num_columns,
                                                                   int32_t sh_degree);

/* Adjoint of the fused forward above with respect to grid_sdf and grid_sh (fuses
 * evaluate_spherical_harmonics.slang:228-270 with create_triangle_vertices.slang:201-291).
 * Overwrites (accumulate = 0; no zero-fill needed) or adds to (accumulate = 1)
 * grid_sdf_grad f32[num_grid_vertices] and grid_sh_grad f32[num_sh, num_grid_vertices, 3]
 * from grad_sdf f32[V] and
 * grad_colors f32[B, V, 3], using the forward's vertices f32[V, 3] and colors (relu is
 * active where the colour is > 0). Each column visits its crossings once; the origin's
 * and the axis ends' shares go through workspace rows to the grid vertices. When each grid
 * vertex is the origin / +axis end of at most one column (as extract_voxels guarantees),
 * each grid vertex gathers its rows and is written once; when columns share grid vertices
 * (a multi-component grid whose coincident vertices were welded) the rows are added with
 * atomics instead, decided on the device. status is not written (may be NULL).
 * num_sh in {1, 4, 9, 16}. vertex_touched (uint8[V] or NULL): the flags
 * mt_intersect_and_composite_backward set; vertices without the flag have zero gradients
 * and are skipped. clear_gradients = 1 zeroes the grad_sdf and grad_colors entries it
 * reads (every nonzero entry when the flags are given), so that a caller that keeps them
 * zero between steps needs no zero-fill (no counterpart in the reference). maps_ready = 1
 * reuses the grid-vertex maps (the first part of the workspace) that an earlier call built
 * over the same lower_corners, num_columns and num_grid_vertices, with the workspace not
 * written since except by such calls; 0 rebuilds them. colors_stride: the layout of colors
 * and grad_colors, as in the forward (0: f32[B, V, 3]; s >= B: f32[V, s, 3]). */
int mt_create_triangle_vertices_colors_backward(
    const float* grid_sdf, int64_t num_grid_vertices, int32_t num_sh, int32_t sh_degree,
    const int32_t* lower_corners, int64_t num_columns, const float* level_sets, int32_t S,
    const int32_t* vertex_offsets, int64_t num_triangle_vertices, const float* vertices,
    const float* colors, const float* camera_positions, int32_t B, float* grad_sdf,
    float* grad_colors, float* grid_sdf_grad, float* grid_sh_grad, void* workspace,
    size_t workspace_bytes, int32_t accumulate, const uint8_t* vertex_touched,
    int32_t clear_gradients, int32_t maps_ready, int32_t* status, int32_t colors_stride,
    mt_stream_t stream);

/* SDF regularizers (replace compute_sdf_regularizers.slang:18-128, .py:26-139). Per voxel:
 * the Eikonal loss at the voxel centre, (|(voxel_x dx, voxel_y dy, voxel_z dz)| - 1)^2 with
 * (dx, dy, dz) the trilinear gradient there, split into eikonal_loss_pos / _neg by the sign
 * of |.| - 1, and the curvature loss (the 7-point Laplacian at the voxel's +xyz corner over
 * its present neighbours)^2; each f32[num_voxels]. voxel_neighbors int32[7, num_voxels],
 * voxel_lower_corners int32[4, num_columns], voxel_upper_corners int32[4, num_voxels];
 * (voxel_x, voxel_y, voxel_z) = (k, j, i) of the grid size. The backward adds
 * d(sum of grad_* x loss) / d grid_sdf into grid_sdf_grad (float atomics; zero it first
 * for a plain gradient). */
int mt_compute_sdf_regularizers_forward(const float* grid_sdf, const int32_t* voxel_neighbors,
                                        const int32_t* voxel_lower_corners, int64_t num_columns,
                                        const int32_t* voxel_upper_corners, int64_t num_voxels,
                                        int32_t voxel_x, int32_t voxel_y, int32_t voxel_z,
                                        float* eikonal_loss_pos, float* eikonal_loss_neg,
                                        float* curvature_loss, mt_stream_t stream);
int mt_compute_sdf_regularizers_backward(
    const float* grid_sdf, const int32_t* voxel_neighbors, const int32_t* voxel_lower_corners,
    int64_t num_columns, const int32_t* voxel_upper_corners, int64_t num_voxels, int32_t voxel_x,
    int32_t voxel_y, int32_t voxel_z, const float* grad_eikonal_loss_pos,
    const float* grad_eikonal_loss_neg, const float* grad_curvature_loss, float* grid_sdf_grad,
    mt_stream_t stream);

/* Coarse-to-fine subdivision. Packed grids are int32[I, J, KP], bit b of word (i, j, kp)
 * being cell k = 32 kp + b.
 * mt_write_occupancy: sets the cells of voxel_indices int32[n] (linear cell indices) in
 * occupancy (num_words words, zeroed by the caller); indices outside the grid are ignored;
 * min_shell / max_shell are unused, as in the reference (write_occupancy.slang:3-20,
 * .py:25-46).
 * mt_dilate_occupancy: box dilation by 0 < dilation < 32 cells along every axis into out
 * (dilate_occupancy.slang:9-94, .py:27-48).
 * mt_upscale_occupancy: cell (i, j, k) of the finer grid int32[I2, J2, KP2] (I2 >= I, ...)
 * is occupied where the coarse cell under any of its 8 corners is, corners mapped with
 * integer arithmetic (upscale_occupancy.slang:1-74, .py:27-53).
 * mt_upscale_parameters: for every occupied vertex of the new vertex-occupancy grid
 * (int32[NI, NJ, NKP] with its exclusive offsets), the old vertices (vertex-occupancy grid
 * int32[OI, OJ, OKP], offsets, SDF f32[old_num_vertices], SH f32[num_sh, old_num_vertices,
 * 3]) around its position, trilinearly weighted over the occupied ones and normalised,
 * into new_sdf f32[new_num_vertices] and new_sh f32[num_sh, new_num_vertices, 3]
 * (upscale_parameters.slang:54-198, .py:34-99). num_sh in {1, 4, 9, 16}. */
int mt_write_occupancy(const int32_t* voxel_indices, int64_t n, int32_t* occupancy,
                       int64_t num_words, int32_t min_shell, int32_t max_shell,
                       mt_stream_t stream);
int mt_dilate_occupancy(const int32_t* occupancy, int32_t I, int32_t J, int32_t KP,
                        int32_t dilation, int32_t* out, mt_stream_t stream);
int mt_upscale_occupancy(const int32_t* occupancy, int32_t I, int32_t J, int32_t KP,
                         int32_t* out, int32_t I2, int32_t J2, int32_t KP2, mt_stream_t stream);
int mt_upscale_parameters(const int32_t* old_vertex_occupancy, const int32_t* old_vertex_offsets,
                          int32_t OI, int32_t OJ, int32_t OKP, const float* old_sdf,
                          const float* old_sh, int64_t old_num_vertices, int32_t num_sh,
                          const int32_t* new_vertex_occupancy, const int32_t* new_vertex_offsets,
                          int32_t NI, int32_t NJ, int32_t NKP, float* new_sdf, float* new_sh,
                          int64_t new_num_vertices, mt_stream_t stream);

/* Vertex welding. A point p = points[i] f32[n, 3] is quantised per coordinate as
 * q = rint(p * fl32(1 / tol)) (the reciprocal in double, rounded to float32, then a float32
 * product rounded half to even: torch.round(points / tol) as ATen evaluates it on the device).
 * mt_quantized_max_abs: max |q| over all coordinates into the device int64 *max_abs (non-finite
 * values or magnitudes of 2^62 or more count as 2^62). Replaces index_vertices.py:17-18.
 * mt_index_vertices: out int32[n], out[i] = the smallest input index j with q_j == q_i.
 * max_abs must be at least the device value mt_quantized_max_abs wrote (the caller reads it
 * back, as the reference does); bits = bit_length(max_abs) + 1 and 3 bits must not exceed 63
 * (MT_ERR_INVALID_ARGUMENT, message "need <3 bits> bits, exceed 63 bits total"). The three
 * coordinates are packed into one key as q + max_abs, so negative coordinates are packed
 * without aliasing. Replaces index_vertices (index_vertices.py:6-30).
 * mt_index_keys: out[i] = the smallest input index j with keys[j] == keys[i] (any int64 keys).
 * Replaces index_keys (index_vertices.py:33-51).
 * Both sort (key, input index) pairs with mt_sort_keys; the workspace is
 * mt_index_vertices_workspace_bytes(n) for either; n < 2^31. */
int mt_quantized_max_abs(const float* points, int64_t n, double tol, int64_t* max_abs,
                         mt_stream_t stream);
size_t mt_index_vertices_workspace_bytes(int64_t n);
int mt_index_vertices(const float* points, int64_t n, double tol, int64_t max_abs, int32_t* out,
                      void* workspace, size_t workspace_bytes, mt_stream_t stream);
int mt_index_keys(const int64_t* keys, int64_t n, int32_t* out, void* workspace,
                  size_t workspace_bytes, mt_stream_t stream);

/* face_counts int32[S * num_voxels] (3 x triangles) and cell_codes uint8[S, num_voxels].
 * Replaces count_triangle_faces.slang:13-50 (.py:28-63). upper_corners
 * int32[4, num_voxels]. */
int mt_count_triangle_faces(const float* sdf, const int32_t* lower_corners, int64_t num_columns,
                            const int32_t* upper_corners, int64_t num_voxels,
                            const float* level_sets, int32_t S, int32_t* face_counts,
                            uint8_t* cell_codes, mt_stream_t stream);

/* faces int32[F, 3] and voxel_indices int32[F] from the exclusive face offsets
 * int32[S * num_voxels + 1]; F = face_capacity (faces at or beyond it are dropped and
 * flagged in *overflow when not NULL). vertex_types holds vertex_capacity entries; a
 * vertex range reaching past it yields corner -1 and raises the flag.
 * Replaces create_triangle_faces.slang:3-78 (.py:28-65). */
int mt_create_triangle_faces(const int32_t* neighbors, const int32_t* indices,
                             int64_t num_voxels, int64_t num_columns,
                             const int32_t* vertex_offsets, const uint8_t* vertex_types,
                             const uint8_t* cell_codes, int32_t S, const int32_t* face_offsets,
                             int32_t* faces, int32_t* voxel_indices, int64_t face_capacity,
                             int64_t vertex_capacity, int32_t* overflow, mt_stream_t stream);

/* mt_create_triangle_faces reading int64 exclusive face offsets int64[S * num_voxels + 1]
 * (from mt_inclusive_scan_i32_to_i64 of the face counts), for meshes with 2^31 or more
 * face corners, where the reference's int32 compute_exclusive_cumsum
 * (compute_exclusive_cumsum.cu:8-39) wraps around; face ids stay int32
 * (face_capacity < 2^31). Same outputs as mt_create_triangle_faces wherever the int32
 * offsets do not wrap. Replaces create_triangle_faces.slang:3-78 (.py:28-65). */
int mt_create_triangle_faces_i64(const int32_t* neighbors, const int32_t* indices,
                                 int64_t num_voxels, int64_t num_columns,
                                 const int32_t* vertex_offsets, const uint8_t* vertex_types,
                                 const uint8_t* cell_codes, int32_t S, const int64_t* face_offsets,
                                 int32_t* faces, int32_t* voxel_indices, int64_t face_capacity,
                                 int64_t vertex_capacity, int32_t* overflow, mt_stream_t stream);

/* mt_create_triangle_faces_i64 for the batched training step: each corner's vertex comes
 * from its owner column's crossing-axis bits (column_axes uint8[S * num_columns], from
 * mt_count_triangle_vertices_axes) instead of a search of the column's vertex types, one
 * dependent load fewer per corner; no voxel indices. faces are the same as
 * mt_create_triangle_faces_i64's. Replaces create_triangle_faces.slang:3-78. */
int mt_create_triangle_faces_axes_i64(const int32_t* neighbors, int64_t num_voxels,
                                      int64_t num_columns, const int32_t* vertex_offsets,
                                      const uint8_t* column_axes, const uint8_t* cell_codes,
                                      int32_t S, const int64_t* face_offsets, int32_t* faces,
                                      int64_t face_capacity, int64_t vertex_capacity,
                                      int32_t* overflow, mt_stream_t stream);

/* mt_create_triangle_faces_axes_i64 with each (level set, column)'s first vertex and
 * crossing axes in one word (column_words, as mt_create_triangle_vertices_colors_forward
 * writes them): one 4-byte load per corner instead of an offset and an axes byte.
 * vertex_capacity <= MT_COLUMN_WORD_MAX_VERTICES. Same faces. Replaces
 * create_triangle_faces.slang:3-78 in the batched training step. */
int mt_create_triangle_faces_words_i64(const int32_t* neighbors, int64_t num_voxels,
                                       int64_t num_columns, const uint32_t* column_words,
                                       const uint8_t* cell_codes, int32_t S,
                                       const int64_t* face_offsets, int32_t* faces,
                                       int64_t face_capacity, int64_t vertex_capacity,
                                       int32_t* overflow, mt_stream_t stream);

/* ---------------------------------------------------------------------------------
 * Rasterization (triangle_rasterization/__init__.py:22-329)
 * ------------------------------------------------------------------------------- */

/* vertices f32[V, 3]; extrinsics f32[B, 4, 4]; intrinsics f32[B, 3, 3];
 * out_xy f32[B, V, 2]; out_depth f32[B, V] (may be NULL: not written). Replaces
 * project_vertices.slang:61-92 (.py:31-56). */
int mt_project_vertices(const float* vertices, int64_t V, const float* extrinsics,
                        const float* intrinsics, int32_t B, float* out_xy, float* out_depth,
                        mt_stream_t stream);

/* mt_project_vertices over the first min(*num_live, V) vertices only (num_live: int32 on
 * the device, e.g. the last entry of a vertex-offset scan); the outputs keep the row stride
 * V and their entries past the live count are not written. For callers whose vertex
 * arrays have a capacity above the live count (the batched training step); the reference
 * projects exactly its V vertices (project_vertices.py:31-56). */
int mt_project_vertices_live(const float* vertices, int64_t V, const int32_t* num_live,
                             const float* extrinsics, const float* intrinsics, int32_t B,
                             float* out_xy, float* out_depth, mt_stream_t stream);

/* Back-face / near-plane culling and tile bounds. faces int32[F, 3]; outputs
 * tile_min int32[B, F, 2] (row, col) inclusive, tile_max int32[B, F, 2] exclusive,
 * counts int32[B, F]. If num_faces_x3 is not NULL, only faces below (*num_faces_x3) / 3
 * are live (F is then the capacity) and the rest get empty bounds. A face with a corner
 * outside [0, V) is culled (empty bounds).
 * Replaces bound_triangles.slang:144-192 (.py:35-71). */
int mt_bound_triangles(const float* xy, const float* depths, int64_t V, const int32_t* faces,
                       int64_t F, int32_t B, int32_t tile_h, int32_t tile_w, int32_t row_min,
                       int32_t col_min, int32_t row_max, int32_t col_max, float near_plane,
                       int32_t* tile_min, int32_t* tile_max, int32_t* counts,
                       const int64_t* num_faces_x3, mt_stream_t stream);

/* One key per (camera, face, covered tile), written at the inclusive offsets
 * int64[B * F] of the tile counts. keys int64[K], values int32[K] (face ids) with
 * K = key_capacity; keys at or beyond it are dropped and flagged in *overflow when not
 * NULL. Replaces generate_keys.slang:4-68 (.py:34-70). */
int mt_generate_keys(const float* depths, int64_t V, const int32_t* faces, int64_t F,
                     const int32_t* shell_face_boundaries, int32_t S, const int32_t* tile_min,
                     const int32_t* tile_max, const int64_t* offsets, int32_t B,
                     int32_t grid_cols, mt_key_layout layout, int64_t* keys, int32_t* values,
                     int64_t key_capacity, int32_t* overflow, mt_stream_t stream);

/* Batched training step (no single counterpart in the reference: fuses
 * bound_triangles.slang:144-192 with generate_keys.slang:4-68 and the tile-count cumsum of
 * generate_keys.py:40-47). Faces are taken in blocks of 256 per camera. block_counts
 * int32[B, ceil(F / 256)] = number of keys of every (camera, face block), computed with
 * mt_bound_triangles' arithmetic (same culling, same num_faces_x3 semantics). cam_dims
 * int32[B, 4] (height, width, grid_rows, grid_cols), or NULL: camera b's tile grid is
 * [0, grid_rows_b) x [0, grid_cols_b) instead of [row_min, row_max) x [col_min, col_max).
 * cull = 1 narrows each face's tiles to those where some pixel can pass the barycentric
 * test: the candidate pixel rectangle (the bound the composites use per tile, here over the
 * face's whole tile box) excludes the others, and when it holds at most 16 pixels each is
 * tested with the reference's float operations. Keys of tiles where no pixel can pass are
 * dropped, which changes no image or gradient of the fused composites; cull = 0 keeps every
 * tile of the bounding box. */
int mt_key_block_counts(const float* xy, const float* depths, int64_t V, const int32_t* faces,
                        int64_t F, const int64_t* num_faces_x3, int32_t B, int32_t tile_h,
                        int32_t tile_w, int32_t row_min, int32_t col_min, int32_t row_max,
                        int32_t col_max, float near_plane, const int32_t* cam_dims,
                        int32_t cull, int32_t* block_counts, mt_stream_t stream);

/* Keys of every (camera, face, covered tile) like mt_generate_keys, with each face's
 * tile rectangle recomputed from the projected geometry (no stored bounds) and each
 * block's first key at block_offsets[q - 1] (0 for q = 0), block_offsets = the inclusive
 * int64 scan of mt_key_block_counts' output (mt_inclusive_scan_i32_to_i64). With cull = 0
 * the keys and values are identical to mt_bound_triangles + inclusive scan +
 * mt_generate_keys; cull (the same value as for mt_key_block_counts) and cam_dims as in
 * mt_key_block_counts (tile ids then use each camera's grid_cols). */
int mt_generate_keys_from_geometry(const float* xy, const float* depths, int64_t V,
                                   const int32_t* faces, int64_t F, const int64_t* num_faces_x3,
                                   const int32_t* shell_face_boundaries, int32_t S, int32_t B,
                                   int32_t tile_h, int32_t tile_w, int32_t row_min,
                                   int32_t col_min, int32_t row_max, int32_t col_max,
                                   float near_plane, const int32_t* cam_dims, int32_t cull,
                                   const int64_t* block_offsets, int32_t grid_cols,
                                   mt_key_layout layout, int64_t* keys,
                                   int32_t* values, int64_t key_capacity, int32_t* overflow,
                                   mt_stream_t stream);

/* The three calls above in one pass (batched training step and serving; no single
 * counterpart in the reference: fuses bound_triangles.slang:144-192, the tile-count cumsum
 * of generate_keys.py:40-47 and generate_keys.slang:4-68). A workgroup takes one unit of
 * (face, camera) pairs (a block of faces for a slice of the cameras), counts the unit's
 * keys, reserves an output range with one atomic add and writes the keys into the scratch
 * arrays; a scan of the unit counts and a copy kernel then place every unit's keys in unit
 * order in keys / values. So each face's tile rectangle is computed once and no workgroup
 * waits for another. For every camera, the keys and values are those of
 * mt_key_block_counts + scan + mt_generate_keys_from_geometry with the same arguments, in
 * the same order; the cameras' sequences are interleaved unit by unit, which a stable sort
 * cannot tell apart because keys of different cameras differ in the camera field (B > 1
 * therefore needs 2^camera_bits >= B). Tile grids: at most 65535 rows and columns
 * (row_min, col_min >= 0 and row_max, col_max <= 65535; the same for the cam_dims grids,
 * which the caller checks). A unit holds 1024 (face, camera) pairs and counts its keys in
 * int32, so 1024 x (row_max - row_min) x (col_max - col_min) must stay below 2^31; with
 * cam_dims, row_max and col_max (otherwise unused) bound every camera's grid for this
 * check. *n_keys (device) = the total key count (also when it exceeds key_capacity, which
 * sets *overflow; the keys below the capacity are then each a generated key or key 0 with
 * value 0, so that the sort, the delineation and the composites stay within their arrays
 * while the caller discards the result). scratch_keys / scratch_values: key_capacity
 * entries each, separate from
 * keys and values (their contents are overwritten). Workspace:
 * mt_generate_keys_single_pass_workspace_bytes(F, B). */
size_t mt_generate_keys_single_pass_workspace_bytes(int64_t F, int32_t B);
int mt_generate_keys_single_pass(const float* xy, const float* depths, int64_t V,
                                 const int32_t* faces, int64_t F, const int64_t* num_faces_x3,
                                 const int32_t* shell_face_boundaries, int32_t S, int32_t B,
                                 int32_t tile_h, int32_t tile_w, int32_t row_min, int32_t col_min,
                                 int32_t row_max, int32_t col_max, float near_plane,
                                 const int32_t* cam_dims, int32_t cull, int32_t grid_cols,
                                 mt_key_layout layout, int64_t* keys, int32_t* values,
                                 int64_t key_capacity, int64_t* n_keys, int32_t* overflow,
                                 int64_t* scratch_keys, int32_t* scratch_values,
                                 void* workspace, size_t workspace_bytes, mt_stream_t stream);


/* Workspace for mt_sort_keys on n pairs. */
size_t mt_sort_workspace_bytes(int64_t n);

/* Stable ascending sort of (signed int64 key, int32 value) pairs over all 64 key bits
 * (hand-written reduce-then-scan LSD radix sort: per 8-bit digit a count pass, a scan of
 * the (digit, tile) counts and a scatter pass; passes whose digit is constant over all
 * keys are skipped on the device). Inputs are not modified. If n_device is not NULL the device value
 * min(*n_device, n) is the number of pairs and n only sizes the launch (capacity).
 * Replaces sort_keys (sort_keys.cu:215-270, sort_keys.py:18-29). */
int mt_sort_keys(const int64_t* keys_in, const int32_t* values_in, int64_t n, int64_t* keys_out,
                 int32_t* values_out, void* workspace, size_t workspace_bytes,
                 const int64_t* n_device, mt_stream_t stream);

/* Tile-bucketed sort, part 1 (batched step and serving; no single counterpart in the
 * reference, whose sort_keys.cu:215-270 sorts all bits at once): mt_sort_keys over the key
 * bits [low_bits, 64) only (signed order on bit 63), stable, so keys with equal high bits
 * keep their input order. Same workspace as mt_sort_keys. With the fused layout and
 * low_bits = the tile field's offset, the keys come out grouped by (camera, tile); a
 * delineation then gives each tile's range, and mt_sort_key_ranges orders each range by
 * the low bits (depth, shell): together the stable sort by the whole key, i.e.
 * mt_sort_keys' output. */
int mt_sort_keys_high(const int64_t* keys_in, const int32_t* values_in, int64_t n,
                      int64_t* keys_out, int32_t* values_out, void* workspace,
                      size_t workspace_bytes, const int64_t* n_device, int32_t low_bits,
                      mt_stream_t stream);

/* Tile-bucketed sort, part 2: stably sorts, in place, each range [ranges[2r], ranges[2r+1])
 * (r < num_ranges < 2^31; disjoint; empty ranges allowed) of keys / values by the key bits
 * [0, low_bits) (unsigned), 1 <= low_bits <= 63. A first kernel lists the ranges by size;
 * ranges of up to 1024, 2048, 4096, 8192 and 16384 keys are then sorted by one workgroup of
 * 64, 128, 256, 512 and 1024 threads with the keys in registers and LDS (radix passes of 8 bits over
 * the bits that vary inside the range); longer ranges pass through tmp_keys / tmp_values
 * (n entries each, separate arrays, their contents overwritten). workspace: at least
 * mt_sort_key_ranges_workspace_bytes(num_ranges) bytes (the size lists; no initial
 * contents needed). No counterpart in the reference (its sort_keys.cu:215-270 sorts the
 * whole key at once); part 1 and part 2 together replace it. */
size_t mt_sort_key_ranges_workspace_bytes(int64_t num_ranges);
int mt_sort_key_ranges(int64_t* keys, int32_t* values, int64_t n, const int64_t* ranges,
                       int64_t num_ranges, int32_t low_bits, int64_t* tmp_keys,
                       int32_t* tmp_values, void* workspace, size_t workspace_bytes,
                       mt_stream_t stream);

/* [start, end) of every (camera[, shell], tile) run of the sorted keys into the
 * zero-initialised int64[B, S', T, 2] (S' = S if separate_shells else 1). If n_device
 * is not NULL, min(*n_device, n) keys are live. With tile_base int64[B + 1] (cameras of
 * different resolutions, S' = 1 only) camera b's tiles start at pair tile_base[b] and T is
 * ignored. Replaces delineate_tiles.slang:12-58 (.py:24-55). */
int mt_delineate_tiles(const int64_t* sorted_keys, int64_t n, mt_key_layout layout,
                       int32_t separate_shells, int32_t B, int32_t S_out, int32_t T,
                       int64_t* tile_ranges, const int64_t* n_device, const int64_t* tile_base,
                       mt_stream_t stream);

/* out f32[B, 3] = inverse(extrinsics[b])[:3, 3] (computed in double precision).
 * Replaces torch.inverse(extrinsics)[:, :3, 3] (evaluate_spherical_harmonics.py:38). */
int mt_camera_positions(const float* extrinsics, int32_t B, float* out, mt_stream_t stream);

/* colors f32[B, V, 3] from sh f32[num_sh, V, 3] (first (deg+1)^2 coefficients used).
 * Replaces evaluate_spherical_harmonics_forward (evaluate_spherical_harmonics.slang:
 * 187-222, .py:25-60). */
int mt_evaluate_sh_forward(const float* vertices, int64_t V, const float* sh, int32_t num_sh,
                           int32_t sh_degree, const float* camera_positions, int32_t B,
                           float* colors, mt_stream_t stream);

/* Accumulates into zero-initialised grad_sh f32[num_sh, V, 3].
 * Replaces evaluate_spherical_harmonics_backward (slang:224-270, .py:62-97). */
int mt_evaluate_sh_backward(const float* vertices, int64_t V, const float* sh, int32_t num_sh,
                            int32_t sh_degree, const float* camera_positions, int32_t B,
                            const float* grad_colors, float* grad_sh, mt_stream_t stream);

/* Arguments shared by the fused intersect-and-composite kernels. */
typedef struct mt_composite_args {
  const float* xy;            /* f32[B, V, 2] projected vertices */
  const float* colors;        /* f32[B, V, 3] (or f32[V, colors_stride, 3]) */
  const float* sdf;           /* f32[V] */
  int64_t num_vertices;
  const int32_t* faces;       /* int32[F, 3] */
  const float* sharpness;     /* f32[1] (device) */
  const int64_t* sorted_keys; /* int64[K] */
  const int32_t* sorted_values; /* int32[K] */
  const int64_t* tile_ranges; /* int64[B, T, 2] */
  /* A tile holds at most 512 pixels (tile_h * tile_w <= 512). */
  int32_t num_cameras, num_shells, grid_rows, grid_cols, tile_h, tile_w, height, width;
  mt_key_layout layout;
  int32_t chunk; /* triangles staged in LDS per round (<= 512): 0 = tile pixel count */
  /* Optional (may be NULL): uint64[2] incremented by the number of (pixel, triangle)
   * barycentric tests executed and of accepted hits. */
  unsigned long long* stats;
  /* Optional camera table for cameras of different resolutions (all NULL / 0: every
   * camera is height x width on a grid_rows x grid_cols tile grid). cam_dims int32[B, 4] =
   * (height, width, grid_rows, grid_cols) per camera; tile_base int64[B + 1] = first tile
   * of each camera in tile_ranges (pairs, S' = 1); pixel_base int64[B + 1] = first pixel of
   * each camera: its planes [4, H_b, W_b] start at out + 4 * pixel_base[b];
   * total_tiles = tile_base[B] (host copy). */
  const int32_t* cam_dims;
  const int64_t* tile_base;
  const int64_t* pixel_base;
  int64_t total_tiles;
  /* Optional (may be NULL), backward only: uint8[V], set to 1 for every vertex that
   * receives a gradient (the caller zeroes it); lets the colour adjoint skip the others. */
  uint8_t* vertex_touched;
  /* Layout of colors and colors_grad: 0 = f32[B, V, 3] (camera-major, the reference's);
   * s >= B = f32[V, s, 3] (each vertex's cameras together, the batched step's layout). */
  int32_t colors_stride;
} mt_composite_args;

/* out f32[B, 4, H, W] (rgb, transmittance). Replaces intersect_and_composite_forward
 * (intersect_and_composite.slang:224-339, .py:25-80). */
int mt_intersect_and_composite_forward(const mt_composite_args* args, float* out,
                                       mt_stream_t stream);

/* Recomputes the traversal and accumulates into zero-initialised colors_grad
 * f32[B, V, 3], sdf_grad f32[V], sharpness_grad f32[1]. reference_quirk = 1 scatters each
 * delayed sample gradient with the next triangle's barycentrics exactly as the
 * reference does; 0 uses the sample's own triangle (the true adjoint).
 * Replaces intersect_and_composite_backward (slang:341-564, .py:82-147). */
int mt_intersect_and_composite_backward(const mt_composite_args* args, const float* out,
                                        const float* grad_out, float* colors_grad,
                                        float* sdf_grad, float* sharpness_grad,
                                        int32_t reference_quirk, mt_stream_t stream);

/* Training loss of the batched step: loss += mean((clip(rgb + T * background, 0, 1) -
 * target)^2) over B x 3 x H x W, and grad_out f32[B, 4, H, W] = its gradient with
 * respect to out f32[B, 4, H, W] (rgb, transmittance). The reference computes this loss
 * with torch ops in the model (model_meshtryoshka.py:249-255, L2 part). background f32[3],
 * target f32[B, 3, H, W]; *loss (float64, accumulated over the blocks' partial sums) must
 * be zeroed by the caller. With pixel_base int64[B + 1]
 * (cameras of different resolutions; H, W ignored) camera b's planes start at
 * out + 4 * pixel_base[b] and target + 3 * pixel_base[b], total_pixels = pixel_base[B].
 * mean_pixels > 0 sets the pixel count of the mean (a batch rendered in camera groups adds
 * each group's share of one mean); 0 means the pixels of this call. */
int mt_image_l2_loss(const float* out, const float* background, const float* target, int32_t B,
                     int32_t H, int32_t W, const int64_t* pixel_base, int64_t total_pixels,
                     int64_t mean_pixels, double* loss, float* grad_out, mt_stream_t stream);

/* Explicit mode: first hit per (camera, shell, pixel) in depth order. tile_ranges
 * int64[B, S, T, 2]; out_uv f32[B*S, H, W, 2]; out_index int32[B*S, H, W] (-1 = miss).
 * Replaces intersect.slang:85-182 (.py:33-70). */
int mt_intersect(const int32_t* sorted_values, const int64_t* tile_ranges, const float* xy,
                 int64_t V, const int32_t* faces, int32_t B, int32_t S, int32_t grid_rows,
                 int32_t grid_cols, int32_t tile_h, int32_t tile_w, int32_t height,
                 int32_t width, float* out_uv, int32_t* out_index, mt_stream_t stream);

/* (rgb, sdf) = u a + v b + (1 - u - v) c per pixel; misses write 0.
 * uv f32[B*S, H, W, 2]; index int32[B*S, H, W]; colors f32[B, V, 3].
 * Replaces interpolate.slang:19-76 (.py:21-75). */
int mt_interpolate_forward(const float* uv, const int32_t* index, const int32_t* faces,
                           const float* sdf, const float* colors, int64_t V, int32_t B,
                           int32_t S, int32_t H, int32_t W, float* out_sdf, float* out_colors,
                           mt_stream_t stream);

/* Adjoint of the above into zero-initialised sdf_grad f32[V], colors_grad f32[B, V, 3]
 * (interpolate.slang auto-derivative, .py:76-121). */
int mt_interpolate_backward(const float* uv, const int32_t* index, const int32_t* faces,
                            int64_t V, int32_t B, int32_t S, int32_t H, int32_t W,
                            const float* grad_sdf, const float* grad_colors, float* sdf_grad,
                            float* colors_grad, mt_stream_t stream);

/* Per-pixel front-to-back composite over S masked layers. sdf f32[P, S],
 * colors f32[P, S, 3], mask uint8[P, S]; out_colors f32[P, 3], out_trans f32[P].
 * Replaces composite_forward (composite.slang:60-104, .py:22-61). */
int mt_composite_forward(const float* sdf, const float* colors, const uint8_t* mask, int64_t P,
                         int32_t S, const float* sharpness, float* out_colors, float* out_trans,
                         mt_stream_t stream);

/* Adjoint into zero-initialised sdf_grad f32[P, S], colors_grad f32[P, S, 3] and
 * sharpness_grad f32[1]. reference_quirk = 1 sends each delayed SDF gradient to layer
 * s - 1 and the last one to layer S - 1 as the reference does; 0 sends them to the
 * previous and the last unmasked layer. Replaces composite_backward (composite.slang:
 * 106-218, .py:63-101). */
int mt_composite_backward(const float* sdf, const float* colors, const uint8_t* mask, int64_t P,
                          int32_t S, const float* sharpness, const float* out_colors,
                          const float* out_trans, const float* grad_colors,
                          const float* grad_trans, float* sdf_grad, float* colors_grad,
                          float* sharpness_grad, int32_t reference_quirk, mt_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* MESHTRYOSHKA_HIP_H_ */
