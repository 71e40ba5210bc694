// Real spherical harmonics up to degree 3 with the reference's constants and signs
// (evaluate_spherical_harmonics.slang:124-185), and the camera-to-vertex view direction.
// Shared by the SH kernels and the fused vertex + colour kernels of the batched step, so
// both evaluate colours with the same operations in the same order.
#pragma once

#include <hip/hip_runtime.h>

namespace mt {

constexpr float kC0 = 0.28209479177387814f;
constexpr float kC1 = 0.4886025119029199f;
constexpr float kC2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                          -1.0925484305920792f, 0.5462742152960396f};
constexpr float kC3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                          0.3731763325901154f,  -0.4570457994644658f, 1.445305721320277f,
                          -0.5900435899266435f};

// Basis values with the reference's signs; only the first (deg+1)^2 are written.
__device__ __forceinline__ void sh_basis(float x, float y, float z, int deg, float* out) {
  out[0] = kC0;
  if (deg > 0) {
    out[1] = -kC1 * y;
    out[2] = kC1 * z;
    out[3] = -kC1 * x;
  }
  if (deg > 1) {
    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
    out[4] = kC2[0] * xy;
    out[5] = kC2[1] * yz;
    out[6] = kC2[2] * (2.0f * zz - xx - yy);
    out[7] = kC2[3] * xz;
    out[8] = kC2[4] * (xx - yy);
    if (deg > 2) {
      out[9] = kC3[0] * y * (3.0f * xx - yy);
      out[10] = kC3[1] * xy * z;
      out[11] = kC3[2] * y * (4.0f * zz - xx - yy);
      out[12] = kC3[3] * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
      out[13] = kC3[4] * x * (4.0f * zz - xx - yy);
      out[14] = kC3[5] * z * (xx - yy);
      out[15] = kC3[6] * x * (xx - 3.0f * yy);
    }
  }
}

__device__ __forceinline__ float3 vertex_at(const float* verts, int64_t v) {
  return make_float3(verts[3 * v + 0], verts[3 * v + 1], verts[3 * v + 2]);
}

// Unit direction from the camera centre to the vertex.
__device__ __forceinline__ float3 view_dir(float3 p, const float* cam) {
  float dx = p.x - cam[0];
  float dy = p.y - cam[1];
  float dz = p.z - cam[2];
  const float inv = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
  return make_float3(dx * inv, dy * inv, dz * inv);
}

// Colour of one channel from the coefficients: relu(sum_k basis_k coef_k + 0.5), summed in
// coefficient order (evaluate_spherical_harmonics.slang:187-226).
template <int kN>
__device__ __forceinline__ float sh_channel_pre(const float* basis, const float (*coef)[3], int c) {
  float acc = basis[0] * coef[0][c];
#pragma unroll
  for (int k = 1; k < kN; ++k) acc += basis[k] * coef[k][c];
  return acc + 0.5f;
}

// The same sum for one channel whose coefficients are in coef[0..kN).
template <int kN>
__device__ __forceinline__ float sh_pre(const float* basis, const float* coef) {
  float acc = basis[0] * coef[0];
#pragma unroll
  for (int k = 1; k < kN; ++k) acc += basis[k] * coef[k];
  return acc + 0.5f;
}

}  // namespace mt
