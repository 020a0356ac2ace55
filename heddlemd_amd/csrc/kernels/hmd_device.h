// Device helpers shared by every gfx950 kernel: periodic boundaries,
// fixed-point conversion, wave64 primitives.
#pragma once
#ifndef __HIPCC_RTC__   // hiprtc provides the runtime declarations itself
#include <hip/hip_runtime.h>
#endif
#include "hmd_params.h"

#define HMD_WAVE 64

// ------------------------------------------------------------------ PBC --
// Cartesian -> fractional by back-substitution and minimum image, operation
// for operation as kernels/pbc.cuh:17-43 (lower-triangular cell).  Fused
// multiply-add contraction is disabled so the f32 result is one exactly
// specified function (the CPU oracle evaluates the same sequence), which is
// what makes the neighbour pair set bit-exact.
__device__ inline void hmd_min_image(float& dx, float& dy, float& dz,
                                     const HmdLattice& L) {
#pragma clang fp contract(off)
  float sc = dz / L.lz;
  float sb = (dy - sc * L.yz) / L.ly;
  float sa = (dx - sb * L.xy - sc * L.xz) / L.lx;
  float ka = floorf(sa + 0.5f);
  float kb = floorf(sb + 0.5f);
  float kc = floorf(sc + 0.5f);
  dx -= ka * L.lx + kb * L.xy + kc * L.xz;
  dy -= kb * L.ly + kc * L.yz;
  dz -= kc * L.lz;
}

// Wrap a position into the primary image and return the image triple
// (kernels/pbc.cuh:74-92).
__device__ inline void hmd_wrap(float& x, float& y, float& z, int& ka_out,
                                int& kb_out, int& kc_out, const HmdLattice& L) {
#pragma clang fp contract(off)
  float sc = z / L.lz;
  float sb = (y - sc * L.yz) / L.ly;
  float sa = (x - sb * L.xy - sc * L.xz) / L.lx;
  float ka = floorf(sa + 0.5f);
  float kb = floorf(sb + 0.5f);
  float kc = floorf(sc + 0.5f);
  x -= ka * L.lx + kb * L.xy + kc * L.xz;
  y -= kb * L.ly + kc * L.yz;
  z -= kc * L.lz;
  ka_out = (int)ka;
  kb_out = (int)kb;
  kc_out = (int)kc;
}

__device__ inline void hmd_cart_to_frac(float x, float y, float z,
                                        const HmdLattice& L, float& sa,
                                        float& sb, float& sc) {
#pragma clang fp contract(off)
  sc = z / L.lz;
  sb = (y - sc * L.yz) / L.ly;
  sa = (x - sb * L.xy - sc * L.xz) / L.lx;
}

// Bin of r^2 in the real-space Ewald table and the position t = 1 + s in
// [1, 2) inside it (s in [0, 1) across the bin), read straight from the f32
// bits of r^2: the exponent and the top HMD_COUL_BITS mantissa bits give the
// bin (2^-HMD_COUL_BITS of an octave of r^2 wide), the next mantissa bits,
// under the exponent of 1.0, give t.  The table's cubics are in t, so s is
// never formed.  r^2 is clamped to the table ([u_lo, u_hi] as bits; NaN and
// values past the table clamp to its last bin), so the index is always
// inside it.  No square root or reciprocal.
__device__ inline uint32_t hmd_coul_bin(float r2, uint32_t u_lo, uint32_t u_hi, float& t) {
  // one v_med3_u32 for the clamp (u_lo <= u_hi; the compiler does not fold
  // min(max()) into it without knowing that)
  uint32_t u;
  asm("v_med3_u32 %0, %1, %2, %3" : "=v"(u) : "v"(__float_as_uint(r2)), "s"(u_lo), "v"(u_hi));
  t = __uint_as_float(((u << HMD_COUL_BITS) & 0x007FFFFFu) | 0x3F800000u);
  return (u >> (23u - HMD_COUL_BITS)) - (u_lo >> (23u - HMD_COUL_BITS));
}

// Exact f32 squared min-image distance (the neighbour pair-set test).
__device__ inline float hmd_r2_exact(float xi, float yi, float zi, float xj,
                                     float yj, float zj, const HmdLattice& L) {
#pragma clang fp contract(off)
  float dx = xi - xj, dy = yi - yj, dz = zi - zj;
  hmd_min_image(dx, dy, dz, L);
  return dx * dx + dy * dy + dz * dz;
}

// Orthorhombic variant that multiplies by the reciprocal box length instead
// of dividing.  Its image index k can differ from the exact one only when
// s = d/L lies within a rounding of a half-integer, i.e. |d| ~ L/2; for every
// other d the subtraction d - k*L is the same operation on the same operands
// and r^2 is bit-identical.  The caller uses it only when L/2 exceeds the
// search radius with margin, so the |d| ~ L/2 cases are out of range under
// either formula and the accept/reject decision is exact.
__device__ inline float hmd_r2_ortho(float xi, float yi, float zi, float xj,
                                     float yj, float zj, const HmdLattice& L,
                                     float ilx, float ily, float ilz) {
#pragma clang fp contract(off)
  float dx = xi - xj, dy = yi - yj, dz = zi - zj;
  float kc = floorf(dz * ilz + 0.5f);
  float kb = floorf(dy * ily + 0.5f);
  float ka = floorf(dx * ilx + 0.5f);
  dx -= ka * L.lx + kb * L.xy + kc * L.xz;
  dy -= kb * L.ly + kc * L.yz;
  dz -= kc * L.lz;
  return dx * dx + dy * dy + dz * dz;
}

// ---------------------------------------------------------- fixed point --
__device__ inline long long hmd_to_fixed(float v, int shift) {
  return __float2ll_rn(ldexpf(v, shift));
}
__device__ inline float hmd_from_fixed(long long v, int shift) {
  return (float)ldexp((double)v, -shift);
}
__device__ inline void hmd_atomic_add_fixed(long long* p, long long v) {
  if (v != 0)
    atomicAdd(reinterpret_cast<unsigned long long*>(p),
              static_cast<unsigned long long>(v));
}

// ---------------------------------------------------------- wave helpers --
__device__ inline uint32_t hmd_lane() { return __lane_id(); }

__device__ inline float hmd_shfl_xor(float v, int mask) {
  return __shfl_xor(v, mask, HMD_WAVE);
}

__device__ inline uint32_t hmd_readlane(uint32_t v, uint32_t lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}

// Sum v[0..7] over the wave; lane l returns the total of v[l >> 3].
__device__ inline float hmd_transpose_reduce8(const float v[8], uint32_t lane) {
  const bool b5 = lane & 32u, b4 = lane & 16u, b3 = lane & 8u;
  float a[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    float send = b5 ? v[t] : v[t + 4];
    float keep = b5 ? v[t + 4] : v[t];
    a[t] = keep + hmd_shfl_xor(send, 32);
  }
  float b[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    float send = b4 ? a[t] : a[t + 2];
    float keep = b4 ? a[t + 2] : a[t];
    b[t] = keep + hmd_shfl_xor(send, 16);
  }
  float send = b3 ? b[0] : b[1];
  float keep = b3 ? b[1] : b[0];
  float c = keep + hmd_shfl_xor(send, 8);
  c += hmd_shfl_xor(c, 4);
  c += hmd_shfl_xor(c, 2);
  c += hmd_shfl_xor(c, 1);
  return c;
}

// DPP row operations (rows of 16 lanes): lane l receives the value of lane
// l + n (row_shl:n), l - n (row_shr:n), or (l + n) mod 16 within its row
// (row_ror:n).  Lanes whose source lies outside the row, or is disabled in
// EXEC, receive 0: callers execute them with every lane enabled.
template <int CTRL>
__device__ inline float hmd_dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
#define HMD_DPP_ROW_SHL(n) (0x100 + (n))
#define HMD_DPP_ROW_SHR(n) (0x110 + (n))
#define HMD_DPP_ROW_ROR(n) (0x120 + (n))

// Sum v[0..7] over the wave; the total of v[m] is left in lane 8m + 7 (other
// lanes hold partial sums).  A transposed butterfly: every exchange stage
// halves the number of values a lane carries.  The half-wave and row
// exchanges are gfx950's v_permlane32_swap / v_permlane16_swap, which trade
// the upper half (odd rows) of one register for the lower half (even rows) of
// another in one instruction; the in-row stages are DPP operands of the adds.
__device__ inline float hmd_wave_sum8(const float v[8], uint32_t lane) {
  float a[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    // lanes 0-31: v[t] of lanes l and l+32; lanes 32-63: v[t+4] likewise
    const auto s = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[t]),
                                                    __float_as_uint(v[t + 4]), false, false);
    a[t] = __uint_as_float(s[0]) + __uint_as_float(s[1]);
  }
  float b[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    // even rows: a[t] of rows 2q and 2q+1; odd rows: a[t+2] likewise
    const auto s = __builtin_amdgcn_permlane16_swap(__float_as_uint(a[t]),
                                                    __float_as_uint(a[t + 2]), false, false);
    b[t] = __uint_as_float(s[0]) + __uint_as_float(s[1]);
  }
  // lanes with bit 3 clear keep b[0] and take the partner's b[0] (lane l^8)
  const bool b3 = (lane & 8u) != 0u;
  const float send = b3 ? b[0] : b[1];
  const float keep = b3 ? b[1] : b[0];
  float c = keep + hmd_dpp<HMD_DPP_ROW_ROR(8)>(send);
  c += hmd_dpp<HMD_DPP_ROW_SHR(1)>(c);
  c += hmd_dpp<HMD_DPP_ROW_SHR(2)>(c);
  c += hmd_dpp<HMD_DPP_ROW_SHR(4)>(c);
  // DPP reads a disabled source lane as 0: an opaque use here keeps every
  // stage above in the converged code of the caller, so that none of them is
  // sunk into a branch that only the lanes 8m + 7 take
  asm volatile("" : "+v"(c));
  return c;
}

__device__ inline float hmd_readlane_f(float v, uint32_t lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)lane));
}

// Device-side predicate for kernels captured into the per-step graph that
// must run only when a rebuild is pending (status word bit set earlier in the
// same step by another kernel on the same stream).
__device__ inline bool hmd_gate(const uint32_t* status, uint32_t mask) {
  return status == nullptr || (*(volatile const uint32_t*)status & mask) != 0u;
}

// Batched launches: one launch runs the blocks of several independent
// instances of a kernel (the replicas of a batch), each with its own
// parameter block.  Instance r owns blocks block_start[r] .. block_start[r+1]
// of the launch; every block finds its instance by a wave-uniform binary
// search and runs the kernel body with its block index inside the instance
// and the instance's own block count.
__device__ inline uint32_t hmd_batch_member(const HmdBatch& b) {
  uint32_t lo = 0, hi = b.n;
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (b.block_start[mid] <= blockIdx.x) lo = mid;
    else hi = mid;
  }
  return lo;
}

// The kernel NAME (one instance, the parameter struct by value) and its
// batched twin NAME_batch, both running NAME_body(p, block, blocks).
#define HMD_KERNEL(NAME, PT, ...)                                              \
  extern "C" __global__ void __launch_bounds__(__VA_ARGS__) NAME(const PT p) { \
    NAME##_body(p, blockIdx.x, gridDim.x);                                     \
  }                                                                            \
  extern "C" __global__ void __launch_bounds__(__VA_ARGS__)                    \
      NAME##_batch(const HmdBatch b) {                                         \
    const uint32_t r = hmd_batch_member(b);                                    \
    const uint32_t b0 = b.block_start[r];                                      \
    NAME##_body(*reinterpret_cast<const PT*>(b.params + (size_t)r * b.stride), \
                blockIdx.x - b0, b.block_start[r + 1] - b0);                   \
  }
