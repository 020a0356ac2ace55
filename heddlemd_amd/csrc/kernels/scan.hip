// Deterministic exclusive prefix sum of u32 arrays (integer, so any
// association order gives the same bits).  Replaces the reference's
// multi-level Hillis-Steele scan (kernels/neighbor.cu:126-200,
// src/gpu/kernels.rs:1844-1950) with a 1024-thread, 4096-element block scan
// built on wave64 shuffles; two launches cover 16.7 M elements.
#include "hmd_device.h"

#define HMD_SCAN_THREADS 1024u
#define HMD_SCAN_PER_THREAD 4u
#define HMD_SCAN_TILE (HMD_SCAN_THREADS * HMD_SCAN_PER_THREAD)

__device__ inline uint32_t hmd_wave_incl_scan(uint32_t x, uint32_t lane) {
#pragma unroll
  for (uint32_t off = 1; off < HMD_WAVE; off <<= 1) {
    uint32_t t = __shfl_up(x, off, HMD_WAVE);
    if (lane >= off) x += t;
  }
  return x;
}

// Scan one 4096-element tile per block.  block_sums[b] receives the tile
// total.  With total_slot != 0 (single-block use) out[len] = total.
__device__ __forceinline__ void hmd_scan_tiles_body(const ScanParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  if (p.gate && !hmd_gate(p.gate, p.gate_mask)) return;
  __shared__ uint32_t s_wave[HMD_SCAN_THREADS / HMD_WAVE];
  const uint32_t t = threadIdx.x, lane = hmd_lane(), wave = t / HMD_WAVE;
  const uint32_t base = hmd_bx * HMD_SCAN_TILE + t * HMD_SCAN_PER_THREAD;
  uint32_t v[HMD_SCAN_PER_THREAD];
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < HMD_SCAN_PER_THREAD; ++k) {
    v[k] = (base + k < p.len) ? p.in[base + k] : 0u;
    sum += v[k];
  }
  uint32_t incl = hmd_wave_incl_scan(sum, lane);
  if (lane == HMD_WAVE - 1) s_wave[wave] = incl;
  __syncthreads();
  if (wave == 0) {
    uint32_t w = lane < HMD_SCAN_THREADS / HMD_WAVE ? s_wave[lane] : 0u;
    uint32_t wi = hmd_wave_incl_scan(w, lane);
    if (lane < HMD_SCAN_THREADS / HMD_WAVE) s_wave[lane] = wi - w;  // exclusive
  }
  __syncthreads();
  uint32_t run = incl - sum + s_wave[wave];
#pragma unroll
  for (uint32_t k = 0; k < HMD_SCAN_PER_THREAD; ++k) {
    if (base + k < p.len) p.out[base + k] = run;
    run += v[k];
  }
  if (t == HMD_SCAN_THREADS - 1) {
    if (p.block_sums) p.block_sums[hmd_bx] = run;
    if (p.total_slot) p.out[p.len] = run;
  }
}
HMD_KERNEL(hmd_scan_tiles, ScanParams, 1024)

// Second level: add the scanned tile offsets (in) to every element of out;
// the last block also writes out[len] = total (block_sums holds the
// unscanned tile totals).
__device__ __forceinline__ void hmd_scan_add_body(const ScanParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  if (p.gate && !hmd_gate(p.gate, p.gate_mask)) return;
  const uint32_t tile = hmd_bx;
  const uint32_t off = p.in[tile];
  const uint32_t base = tile * HMD_SCAN_TILE;
  for (uint32_t k = threadIdx.x; k < HMD_SCAN_TILE; k += blockDim.x)
    if (base + k < p.len) p.out[base + k] += off;
  if (p.total_slot && tile == hmd_gx - 1 && threadIdx.x == 0)
    p.out[p.len] = off + p.block_sums[tile];
}
HMD_KERNEL(hmd_scan_add, ScanParams, 1024)
