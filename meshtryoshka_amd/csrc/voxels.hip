// Voxel structure from a bit-packed occupancy grid.
//
// Lattice positions are ranked in row-major (i, j, k) order, k = 32 kp + bit. A vertex
// is a corner of an occupied cell, a voxel an occupied cell, a subvoxel a vertex that is
// not a voxel (numbered after all voxels); a column is a voxel or a subvoxel. One thread
// owns one 32-bit word of the vertex-occupancy grid and derives every rank it needs from
// the word's exclusive offsets plus a popcount of the bits below the position, so no
// per-bit running counters are carried. Threads are flattened over (i, j, kp) with kp
// fastest so that a wave touches contiguous words.
#include "common.h"

namespace mt {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t low_mask(int b) {  // bits [0, b), b in [0, 32]
  return b >= 32 ? 0xFFFFFFFFu : ((1u << b) - 1u);
}

__global__ void vertex_occupancy_kernel(const uint32_t* occ, int I, int J, int KP,
                                        int32_t* vocc_out, int32_t* vcount_out) {
  const int64_t n = int64_t(I + 1) * (J + 1) * (KP + 1);
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int kp = static_cast<int>(idx % (KP + 1));
  const int64_t ij = idx / (KP + 1);
  const int j = static_cast<int>(ij % (J + 1));
  const int i = static_cast<int>(ij / (J + 1));
  uint32_t result = 0;
  for (int di = 0; di < 2; ++di) {
    for (int dj = 0; dj < 2; ++dj) {
      const int ci = i - di, cj = j - dj;
      if (ci < 0 || cj < 0 || ci >= I || cj >= J) continue;
      const uint32_t* row = occ + (int64_t(ci) * J + cj) * KP;
      const uint32_t cur = (kp < KP) ? row[kp] : 0u;
      const uint32_t prev = (kp > 0) ? row[kp - 1] : 0u;
      // Cell k or cell k - 1 along the bit axis (the carry comes from the previous word).
      result |= cur | (cur << 1) | (prev >> 31);
    }
  }
  vocc_out[idx] = static_cast<int32_t>(result);
  vcount_out[idx] = __popc(result);
}

__global__ void count_occupancy_kernel(const uint32_t* occ, int I, int J, int KP,
                                       int32_t* counts) {
  const int64_t n = int64_t(I + 1) * (J + 1) * (KP + 1);
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int kp = static_cast<int>(idx % (KP + 1));
  const int64_t ij = idx / (KP + 1);
  const int j = static_cast<int>(ij % (J + 1));
  const int i = static_cast<int>(ij / (J + 1));
  int32_t c = 0;
  if (i < I && j < J && kp < KP) c = __popc(occ[(int64_t(i) * J + j) * KP + kp]);
  counts[idx] = c;
}

// The four lattice rows a word's cells touch: (i, j), (i, j + 1), (i + 1, j), (i + 1, j + 1).
struct Row {
  uint32_t vocc;   // vertex-occupancy word
  uint32_t cocc;   // cell-occupancy word (0 outside the cell grid)
  uint32_t socc;   // subvoxel bits = vocc & ~cocc
  int64_t vbase;   // vertices before this word
  int64_t cbase;   // voxels before this word
  int64_t sbase;   // column index of the first subvoxel at or after this word
  uint32_t next_cocc;  // cell word at kp + 1 (0 outside the grid)
};

__device__ __forceinline__ int64_t vert_at(const Row& r, int b) {
  return r.vbase + __popc(r.vocc & low_mask(b));
}

// Column index of the (vertex-occupied) position at bit q of this row, q in [0, 32];
// q == 32 is bit 0 of the next word.
__device__ __forceinline__ int64_t column_at(const Row& r, int q) {
  bool cell;
  if (q < 32) {
    cell = (r.cocc >> q) & 1u;
  } else {
    cell = r.next_cocc & 1u;
  }
  return cell ? r.cbase + __popc(r.cocc & low_mask(q)) : r.sbase + __popc(r.socc & low_mask(q));
}

__global__ void create_voxels_kernel(const uint32_t* occ, const int32_t* voxel_offsets,
                                     const uint32_t* vocc, const int32_t* vertex_offsets, int I,
                                     int J, int KP, int64_t num_voxels, int64_t num_vertices,
                                     float* vertices, int32_t* neighbors, int32_t* lower,
                                     int32_t* upper, int32_t* indices) {
  const int64_t n = int64_t(I + 1) * (J + 1) * (KP + 1);
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const uint32_t vword = vocc[idx];
  if (vword == 0) return;
  const int kp = static_cast<int>(idx % (KP + 1));
  const int64_t ij = idx / (KP + 1);
  const int j = static_cast<int>(ij % (J + 1));
  const int i = static_cast<int>(ij / (J + 1));

  Row rows[2][2];
  for (int di = 0; di < 2; ++di) {
    for (int dj = 0; dj < 2; ++dj) {
      Row& r = rows[di][dj];
      const int ii = i + di, jj = j + dj;
      if (ii > I || jj > J) {
        r = Row{0u, 0u, 0u, 0, 0, 0, 0u};
        continue;
      }
      const int64_t w = (int64_t(ii) * (J + 1) + jj) * (KP + 1) + kp;
      const bool in_cells = ii < I && jj < J;
      r.vocc = vocc[w];
      r.cocc = (in_cells && kp < KP) ? occ[(int64_t(ii) * J + jj) * KP + kp] : 0u;
      r.next_cocc = (in_cells && kp + 1 < KP) ? occ[(int64_t(ii) * J + jj) * KP + kp + 1] : 0u;
      r.socc = r.vocc & ~r.cocc;
      r.vbase = vertex_offsets[w];
      r.cbase = voxel_offsets[w];
      r.sbase = num_voxels + r.vbase - r.cbase;
    }
  }
  const Row& r00 = rows[0][0];
  const Row& r01 = rows[0][1];  // +y
  const Row& r10 = rows[1][0];  // +z
  const Row& r11 = rows[1][1];  // +yz
  // Vertex occupancy of the next word in this row, for the +x corner of bit 31.
  const uint32_t next_vocc = (kp < KP) ? vocc[idx + 1] : 0u;

  const float inv_i = 1.0f / float(I);
  const float inv_j = 1.0f / float(J);
  const float inv_k = 1.0f / float(32 * KP);
  const int32_t k_max = 32 * KP;

  uint32_t bits = vword;
  while (bits) {
    const int b = __ffs(bits) - 1;
    bits &= bits - 1;
    const int32_t k = 32 * kp + b;
    const int64_t v = vert_at(r00, b);
    vertices[3 * v + 0] = float(k) * inv_k;
    vertices[3 * v + 1] = float(j) * inv_j;
    vertices[3 * v + 2] = float(i) * inv_i;

    const bool is_voxel = (r00.cocc >> b) & 1u;
    const int64_t col = column_at(r00, b);
    const bool x_occ = (b < 31) ? ((r00.vocc >> (b + 1)) & 1u) : (next_vocc & 1u);
    const bool y_occ = (r01.vocc >> b) & 1u;
    const bool z_occ = (r10.vocc >> b) & 1u;
    lower[0 * num_vertices + col] = static_cast<int32_t>(v);
    lower[1 * num_vertices + col] = (is_voxel || x_occ) ? static_cast<int32_t>(v + 1) : -64;
    lower[2 * num_vertices + col] = (is_voxel || y_occ) ? static_cast<int32_t>(vert_at(r01, b)) : -64;
    lower[3 * num_vertices + col] = (is_voxel || z_occ) ? static_cast<int32_t>(vert_at(r10, b)) : -64;
    if (!is_voxel) continue;

    upper[0 * num_voxels + col] = static_cast<int32_t>(vert_at(r11, b) + 1);  // +xyz
    upper[1 * num_voxels + col] = static_cast<int32_t>(vert_at(r11, b));      // +yz
    upper[2 * num_voxels + col] = static_cast<int32_t>(vert_at(r10, b) + 1);  // +xz
    upper[3 * num_voxels + col] = static_cast<int32_t>(vert_at(r01, b) + 1);  // +xy
    indices[col] = k_max * (J * i + j) + k;

    neighbors[0 * num_voxels + col] = static_cast<int32_t>(column_at(r00, b + 1));  // +x
    neighbors[1 * num_voxels + col] = static_cast<int32_t>(column_at(r01, b));      // +y
    neighbors[2 * num_voxels + col] = static_cast<int32_t>(column_at(r10, b));      // +z
    neighbors[3 * num_voxels + col] = static_cast<int32_t>(column_at(r01, b + 1));  // +xy
    neighbors[4 * num_voxels + col] = static_cast<int32_t>(column_at(r11, b));      // +yz
    neighbors[5 * num_voxels + col] = static_cast<int32_t>(column_at(r10, b + 1));  // +xz
    neighbors[6 * num_voxels + col] = static_cast<int32_t>(column_at(r11, b + 1));  // +xyz
  }
}

int check_grid(int I, int J, int KP, const char* name) {
  MT_REQUIRE(I > 0 && J > 0 && KP > 0, "%s: grid dimensions must be positive", name);
  MT_REQUIRE(int64_t(I + 1) * (J + 1) * (KP + 1) < (int64_t(1) << 40), "%s: grid too large", name);
  return MT_OK;
}

}  // namespace
}  // namespace mt

extern "C" int mt_compute_vertex_occupancy(const int32_t* occupancy, int32_t I, int32_t J,
                                           int32_t KP, int32_t* vertex_occupancy,
                                           int32_t* vertex_counts, mt_stream_t stream) {
  if (int rc = mt::check_grid(I, J, KP, "mt_compute_vertex_occupancy")) return rc;
  const int64_t n = int64_t(I + 1) * (J + 1) * (KP + 1);
  mt::vertex_occupancy_kernel<<<mt::blocks_for(n, mt::kBlock), mt::kBlock, 0,
                                static_cast<hipStream_t>(stream)>>>(
      reinterpret_cast<const uint32_t*>(occupancy), I, J, KP, vertex_occupancy, vertex_counts);
  return mt::check_launch("mt_compute_vertex_occupancy");
}

extern "C" int mt_count_occupancy(const int32_t* occupancy, int32_t I, int32_t J, int32_t KP,
                                  int32_t* voxel_counts, mt_stream_t stream) {
  if (int rc = mt::check_grid(I, J, KP, "mt_count_occupancy")) return rc;
  const int64_t n = int64_t(I + 1) * (J + 1) * (KP + 1);
  mt::count_occupancy_kernel<<<mt::blocks_for(n, mt::kBlock), mt::kBlock, 0,
                               static_cast<hipStream_t>(stream)>>>(
      reinterpret_cast<const uint32_t*>(occupancy), I, J, KP, voxel_counts);
  return mt::check_launch("mt_count_occupancy");
}

extern "C" int mt_create_voxels(const int32_t* occupancy, const int32_t* voxel_offsets,
                                const int32_t* vertex_occupancy, const int32_t* vertex_offsets,
                                int32_t I, int32_t J, int32_t KP, int64_t num_voxels,
                                int64_t num_vertices, float* vertices, int32_t* neighbors,
                                int32_t* lower_corners, int32_t* upper_corners, int32_t* indices,
                                mt_stream_t stream) {
  if (int rc = mt::check_grid(I, J, KP, "mt_create_voxels")) return rc;
  MT_REQUIRE(num_vertices < (int64_t(1) << 31), "mt_create_voxels: too many vertices for int32");
  if (num_vertices == 0) return MT_OK;
  const int64_t n = int64_t(I + 1) * (J + 1) * (KP + 1);
  mt::create_voxels_kernel<<<mt::blocks_for(n, mt::kBlock), mt::kBlock, 0,
                             static_cast<hipStream_t>(stream)>>>(
      reinterpret_cast<const uint32_t*>(occupancy), voxel_offsets,
      reinterpret_cast<const uint32_t*>(vertex_occupancy), vertex_offsets, I, J, KP, num_voxels,
      num_vertices, vertices, neighbors, lower_corners, upper_corners, indices);
  return mt::check_launch("mt_create_voxels");
}
