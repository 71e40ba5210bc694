// SDF regularizers of the training step: the Eikonal loss at each voxel's centre and a
// 7-point Laplacian curvature loss (compute_sdf_regularizers.slang:18-128), forward and a
// hand-written adjoint. One thread per voxel; the floating-point expressions keep the
// reference's operation order (the library is compiled without FP contraction), so the
// forward losses are reproducible bit for bit.
#include "common.h"

namespace mt {
namespace {

constexpr int kBlock = 256;

// The voxel's eight corner values and the (at most three) present far-side neighbours of
// its +xyz corner (the lower corners 1..3 of the subvoxel at offset (1, 1, 1)).
struct VoxelSdf {
  float d000, d001, d010, d100, d111, d110, d101, d011;
  int64_t i000, i001, i010, i100, i111, i110, i101, i011;
  float far[3];
  int64_t ifar[3];
  int n_far;
};

__device__ __forceinline__ VoxelSdf load_voxel(const float* sdf, const int32_t* neighbors,
                                               const int32_t* lower, int64_t n_lower,
                                               const int32_t* upper, int64_t n, int64_t v) {
  VoxelSdf s;
  s.i000 = lower[0 * n_lower + v];  // origin
  s.i001 = lower[1 * n_lower + v];  // +x
  s.i010 = lower[2 * n_lower + v];  // +y
  s.i100 = lower[3 * n_lower + v];  // +z
  s.i111 = upper[0 * n + v];        // +xyz
  s.i110 = upper[1 * n + v];        // +yz
  s.i101 = upper[2 * n + v];        // +xz
  s.i011 = upper[3 * n + v];        // +xy
  s.d000 = sdf[s.i000];
  s.d001 = sdf[s.i001];
  s.d010 = sdf[s.i010];
  s.d100 = sdf[s.i100];
  s.d111 = sdf[s.i111];
  s.d110 = sdf[s.i110];
  s.d101 = sdf[s.i101];
  s.d011 = sdf[s.i011];
  // compute_sdf_regularizers.slang:110-117. A voxel without that subvoxel (never the case
  // for extract_voxels' output, which the reference's kernel assumes) counts no far side.
  const int64_t sub = neighbors[6 * n + v];
  s.n_far = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int64_t c = sub >= 0 && sub < n_lower ? lower[(i + 1) * n_lower + sub] : -1;
    s.ifar[i] = c;
    s.far[i] = c >= 0 ? sdf[c] : 0.f;
    s.n_far += c >= 0 ? 1 : 0;
  }
  return s;
}

// Gradient of the trilinear interpolant at the voxel centre (all weights 1/2 * 1/2,
// compute_sdf_regularizers.slang:64-71), in the reference's order of terms.
struct CentreGradient {
  float dx, dy, dz;
};

__device__ __forceinline__ CentreGradient centre_gradient(const VoxelSdf& s) {
  const float tx = 0.5f, ty = 0.5f, tz = 0.5f;
  const float tx1 = 1.0f - tx, ty1 = 1.0f - ty, tz1 = 1.0f - tz;
  CentreGradient g;
  g.dx = ((ty1 * tz1 * (s.d001 - s.d000) + ty * tz1 * (s.d011 - s.d010)) +
          ty1 * tz * (s.d101 - s.d100)) +
         ty * tz * (s.d111 - s.d110);
  g.dy = ((tx1 * tz1 * (s.d010 - s.d000) + tx * tz1 * (s.d011 - s.d001)) +
          tx1 * tz * (s.d110 - s.d100)) +
         tx * tz * (s.d111 - s.d101);
  g.dz = ((tx1 * ty1 * (s.d100 - s.d000) + tx * ty1 * (s.d101 - s.d001)) +
          tx1 * ty * (s.d110 - s.d010)) +
         tx * ty * (s.d111 - s.d011);
  return g;
}

// stable_norm (compute_sdf_regularizers.slang:3-16).
__device__ __forceinline__ float stable_norm(float a, float b, float c) {
  const float scale = fmaxf(fmaxf(fabsf(a), fabsf(b)), fabsf(c));
  if (scale < 1e-10f) return 0.0f;
  const float inv = 1.0f / scale;
  const float sa = inv * a, sb = inv * b, sc = inv * c;
  const float sum = (sa * sa + sb * sb) + sc * sc;
  return scale * sqrtf(sum);
}

// Curvature sum (compute_sdf_regularizers.slang:105-124): the six Laplacian neighbours of
// the +xyz corner that are present, minus their count times that corner.
__device__ __forceinline__ float curvature_sum(const VoxelSdf& s) {
  float sum = (s.d110 + s.d101) + s.d011;
  int counted = 3;
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (s.ifar[i] >= 0) {
      sum += s.far[i];
      ++counted;
    }
  return sum - static_cast<float>(counted) * s.d111;
}

__global__ void __launch_bounds__(kBlock) regularizers_forward_kernel(
    const float* sdf, const int32_t* neighbors, const int32_t* lower, int64_t n_lower,
    const int32_t* upper, int64_t n, float sx, float sy, float sz, float* eik_pos,
    float* eik_neg, float* curvature) {
  const int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const VoxelSdf s = load_voxel(sdf, neighbors, lower, n_lower, upper, n, v);
  const CentreGradient g = centre_gradient(s);
  const float delta = stable_norm(sx * g.dx, sy * g.dy, sz * g.dz) - 1.0f;
  // One sample (NUM_EIKONAL_SAMPLES = 1): the sums divided by 1 are the squares.
  eik_pos[v] = delta > 0.0f ? delta * delta : 0.0f;
  eik_neg[v] = delta < 0.0f ? delta * delta : 0.0f;
  const float c = curvature_sum(s);
  curvature[v] = c * c;
}

// Adjoint: recomputes the forward terms and scatters d loss / d corner with float atomics
// (a grid vertex is a corner of up to eight voxels and a Laplacian neighbour of up to six).
// d stable_norm / d a = a / |(a, b, c)| (the norm's own derivative; 0 below the 1e-10 cut).
__global__ void __launch_bounds__(kBlock) regularizers_backward_kernel(
    const float* sdf, const int32_t* neighbors, const int32_t* lower, int64_t n_lower,
    const int32_t* upper, int64_t n, float sx, float sy, float sz, const float* g_pos,
    const float* g_neg, const float* g_curv, float* grad) {
  const int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const VoxelSdf s = load_voxel(sdf, neighbors, lower, n_lower, upper, n, v);
  const CentreGradient g = centre_gradient(s);
  const float a = sx * g.dx, b = sy * g.dy, c = sz * g.dz;
  const float norm = stable_norm(a, b, c);
  const float delta = norm - 1.0f;
  float g_delta = 0.0f;
  if (delta > 0.0f) g_delta += 2.0f * delta * g_pos[v];
  if (delta < 0.0f) g_delta += 2.0f * delta * g_neg[v];
  float gx = 0.f, gy = 0.f, gz = 0.f;  // d loss / d (dx, dy, dz)
  if (g_delta != 0.0f && norm > 0.0f) {
    const float k = g_delta / norm;
    gx = k * a * sx;
    gy = k * b * sy;
    gz = k * c * sz;
  }
  // Every centre-gradient weight is 1/4.
  const float qx = 0.25f * gx, qy = 0.25f * gy, qz = 0.25f * gz;
  const float cs = curvature_sum(s);
  const float gc = 2.0f * cs * g_curv[v];
  const float c000 = -qx - qy - qz;
  const float c001 = qx - qy - qz;
  const float c010 = -qx + qy - qz;
  const float c100 = -qx - qy + qz;
  const float c011 = qx + qy - qz + gc;
  const float c101 = qx - qy + qz + gc;
  const float c110 = -qx + qy + qz + gc;
  const float c111 = qx + qy + qz - static_cast<float>(3 + s.n_far) * gc;
  atomicAdd(grad + s.i000, c000);
  atomicAdd(grad + s.i001, c001);
  atomicAdd(grad + s.i010, c010);
  atomicAdd(grad + s.i100, c100);
  atomicAdd(grad + s.i011, c011);
  atomicAdd(grad + s.i101, c101);
  atomicAdd(grad + s.i110, c110);
  atomicAdd(grad + s.i111, c111);
  if (gc != 0.0f) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
      if (s.ifar[i] >= 0) atomicAdd(grad + s.ifar[i], gc);
  }
}

}  // namespace
}  // namespace mt

extern "C" int mt_compute_sdf_regularizers_forward(
    const float* grid_sdf, const int32_t* voxel_neighbors, const int32_t* voxel_lower_corners,
    int64_t num_columns, const int32_t* voxel_upper_corners, int64_t num_voxels, int32_t voxel_x,
    int32_t voxel_y, int32_t voxel_z, float* eikonal_loss_pos, float* eikonal_loss_neg,
    float* curvature_loss, mt_stream_t stream) {
  MT_REQUIRE(num_voxels >= 0 && num_columns >= num_voxels,
             "mt_compute_sdf_regularizers_forward: bad sizes");
  if (num_voxels == 0) return MT_OK;
  mt::regularizers_forward_kernel<<<mt::blocks_for(num_voxels, mt::kBlock), mt::kBlock, 0,
                                    static_cast<hipStream_t>(stream)>>>(
      grid_sdf, voxel_neighbors, voxel_lower_corners, num_columns, voxel_upper_corners,
      num_voxels, static_cast<float>(voxel_x), static_cast<float>(voxel_y),
      static_cast<float>(voxel_z), eikonal_loss_pos, eikonal_loss_neg, curvature_loss);
  return mt::check_launch("mt_compute_sdf_regularizers_forward");
}

extern "C" int mt_compute_sdf_regularizers_backward(
    const float* grid_sdf, const int32_t* voxel_neighbors, const int32_t* voxel_lower_corners,
    int64_t num_columns, const int32_t* voxel_upper_corners, int64_t num_voxels, int32_t voxel_x,
    int32_t voxel_y, int32_t voxel_z, const float* grad_eikonal_loss_pos,
    const float* grad_eikonal_loss_neg, const float* grad_curvature_loss, float* grid_sdf_grad,
    mt_stream_t stream) {
  MT_REQUIRE(num_voxels >= 0 && num_columns >= num_voxels,
             "mt_compute_sdf_regularizers_backward: bad sizes");
  if (num_voxels == 0) return MT_OK;
  mt::regularizers_backward_kernel<<<mt::blocks_for(num_voxels, mt::kBlock), mt::kBlock, 0,
                                     static_cast<hipStream_t>(stream)>>>(
      grid_sdf, voxel_neighbors, voxel_lower_corners, num_columns, voxel_upper_corners,
      num_voxels, static_cast<float>(voxel_x), static_cast<float>(voxel_y),
      static_cast<float>(voxel_z), grad_eikonal_loss_pos, grad_eikonal_loss_neg,
      grad_curvature_loss, grid_sdf_grad);
  return mt::check_launch("mt_compute_sdf_regularizers_backward");
}
