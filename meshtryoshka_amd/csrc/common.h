// Shared helpers for the gfx950 kernels: error plumbing, launch geometry, wave64
// primitives and the key-layout bit fields.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/meshtryoshka_hip.h"

namespace mt {

constexpr int kWave = 64;

// Last error message, per host thread (the C ABI returns codes; the message explains).
void set_error(const char* fmt, ...);

inline int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: launch failed: %s", what, hipGetErrorString(e));
    return MT_ERR_HIP;
  }
  return MT_OK;
}

#define MT_REQUIRE(cond, ...)          \
  do {                                 \
    if (!(cond)) {                     \
      ::mt::set_error(__VA_ARGS__);    \
      return MT_ERR_INVALID_ARGUMENT;  \
    }                                  \
  } while (0)

inline unsigned blocks_for(int64_t n, int block) {
  int64_t b = (n + block - 1) / block;
  return static_cast<unsigned>(b < 1 ? 1 : b);
}

// Remap a flat block id so that the blocks that the hardware dispatches to one XCD
// (round-robin: block b runs on XCD b % 8) cover one contiguous range of work items.
// Neighbouring tiles then share that XCD's L2.
__device__ __forceinline__ uint32_t xcd_swizzle(uint32_t bid, uint32_t nblocks) {
  constexpr uint32_t kXcds = 8;
  const uint32_t per = nblocks / kXcds;
  const uint32_t full = per * kXcds;
  if (bid >= full) return bid;  // tail blocks keep their id
  return (bid % kXcds) * per + bid / kXcds;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const uint32_t lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Wave64 inclusive prefix sum of an integer (Hillis-Steele over shuffles).
template <typename T>
__device__ __forceinline__ T wave_inclusive_sum(T x) {
  const int lane = __lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    T y = __shfl_up(x, d, kWave);
    if (lane >= d) x += y;
  }
  return x;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) x += __shfl_xor(x, d, kWave);
  return x;
}

__device__ __forceinline__ float wave_sum_f(float x) {
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) x += __shfl_xor(x, d, kWave);
  return x;
}

// Single-pass scans (scan.hip), callable from other kernels' host code.
size_t scan_workspace_bytes_i64(int64_t n);
int inclusive_scan_i32_to_i64(const int32_t* in, int64_t* out, int64_t n, void* ws,
                              size_t ws_bytes, hipStream_t stream);

// Key bit fields (keys.slang:1-34). A depth always has 32 bits.
struct KeyFields {
  int camera_offset, camera_bits, shell_offset, shell_bits, tile_offset, tile_bits, depth_offset;
};

__host__ __device__ __forceinline__ KeyFields to_fields(const mt_key_layout& l) {
  return KeyFields{l.camera_offset, l.camera_bits, l.shell_offset, l.shell_bits,
                   l.tile_offset,   l.tile_bits,   l.depth_offset};
}

__device__ __forceinline__ uint64_t field_mask(int bits) {
  return bits >= 64 ? ~0ull : ((1ull << bits) - 1ull);
}

__device__ __forceinline__ uint64_t pack_key(const KeyFields& f, uint32_t camera, uint32_t shell,
                                             uint32_t tile, float depth) {
  const uint64_t c = (uint64_t(camera) & field_mask(f.camera_bits)) << f.camera_offset;
  const uint64_t s = (uint64_t(shell) & field_mask(f.shell_bits)) << f.shell_offset;
  const uint64_t t = (uint64_t(tile) & field_mask(f.tile_bits)) << f.tile_offset;
  const uint64_t d = uint64_t(__float_as_uint(depth)) << f.depth_offset;
  return c | s | t | d;
}

__device__ __forceinline__ uint32_t key_field(uint64_t key, int offset, int bits) {
  return static_cast<uint32_t>((key >> offset) & field_mask(bits));
}

}  // namespace mt
