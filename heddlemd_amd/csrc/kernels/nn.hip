// Learned coarse-mesh SPME correction on gfx950: per-atom radial charge
// descriptors, and the pair-scalar MLP (51 -> 64 -> 64 -> 1, SiLU) forward
// and backward passes on the bf16 matrix cores with f32 accumulation.
//
// There is no counterpart in the reference (it has no learned component).
// The model: s = r/r_nn, phi_k(s) = exp(-(15 s - k)^2 / 2) (k = 0..15),
// fc(s) = (1 + cos(pi s))/2; D_i = sum_j q_j phi(s_ij) fc(s_ij);
// x_ij = [phi, one-hot(species_i + species_j), D_i, D_j];
// g_ij = s_out fc(s_ij) (N(x_ij) + N(x_ji)) / 2; dF_i = sum_j g_ij r^_ij.
//
// Traversal: the force field's packed half list (one wave per contiguous
// share of the work list, as the pair kernel).  The descriptor pass keeps
// the i-atom sums in registers and reduces the j-atom sums across the wave
// with the transposed butterfly.  The network passes compact the in-range
// pairs of each entry into an LDS queue and evaluate them 32 pairs (64
// rows: both orientations) at a time:
//  * layer 1 is Z1^T = W1 X^T and layer 2 Z2^T = W2 H1^T as chains of
//    v_mfma_f32_16x16x32_bf16 (A = weights, B = activations); each lane
//    builds its own slice of X^T in registers (no staging);
//  * the layer-1 result registers (hidden h = 16t + 4g + r of row 16m + l%16,
//    g = l/16) are exactly the layer-2 B operand once the hidden units are
//    taken in the order pi(32k + 8g + e) = 16(2k + e/4) + 4g + e%4, so the
//    packed W2 columns follow pi and the activations never leave registers;
//  * per-pair forces go to LDS i64 fixed-point accumulators for the entry's
//    64 i- and 64 j-atoms, flushed with global i64 atomics (order-free sums:
//    bit-reproducible);
//  * the backward pass recomputes the forward per batch, chains dH1^T =
//    W2^T dZ2^T the same way, and accumulates the weight gradients
//    dW = dZ^T (rows) x A (rows) on the matrix cores after a transpose of
//    the two operands through LDS; per-wave partials are summed in a fixed
//    order by hmd_nn_grad_reduce.
#include "hmd_device.h"

typedef __bf16 hmd_bf16x8 __attribute__((ext_vector_type(8)));
typedef float hmd_f32x4 __attribute__((ext_vector_type(4)));

#define HMD_NN_FORCE_SHIFT 40
#define HMD_NN_LOG2E 1.4426950408889634f
#define HMD_NN_QCAP 1088u   // 16 j-slots x 64 lanes + the < 32 pairs carried over
#define HMD_NN_TSTRIDE 72u  // transpose buffer row stride (bf16), 144 B

__device__ inline hmd_f32x4 hmd_mfma16(hmd_bf16x8 a, hmd_bf16x8 b, hmd_f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ inline float hmd_sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-z * HMD_NN_LOG2E));
}

__device__ inline void hmd_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ inline uint16_t hmd_bf16_bits(float v) {
  const __bf16 b = (__bf16)v;
  return __builtin_bit_cast(uint16_t, b);
}

typedef uint32_t hmd_u32x4 __attribute__((ext_vector_type(4)));

// Bits of element e of a bf16x8 operand, read from its packed 32-bit words
// (element 2w is the low half of word w, element 2w + 1 the high half).
__device__ inline uint16_t hmd_bf16x8_bits(hmd_bf16x8 v, int e) {
  const hmd_u32x4 w = __builtin_bit_cast(hmd_u32x4, v);
  const uint32_t d = w[e >> 1];
  return (uint16_t)((e & 1) ? (d >> 16) : (d & 0xFFFFu));
}

// Internal input order of the 64 (51 used) network inputs: rbf 0..15, the
// first atom's descriptor 16..31, the second's 32..47, one-hot 48..50.
// Returns the canonical index (rbf 0..15, one-hot 16..18, D_first 19..34,
// D_second 35..50) or -1 for padding.
__device__ inline int hmd_nn_canon_of_internal(uint32_t f) {
  if (f < 16u) return (int)f;
  if (f < 32u) return (int)(f + 3u);
  if (f < 48u) return (int)(f + 3u);
  if (f < 51u) return (int)(f - 32u);
  return -1;
}

// Hidden-unit order of the chained layer: B-operand k-index -> hidden unit.
__device__ inline uint32_t hmd_nn_pi(uint32_t k) {
  const uint32_t kap = k >> 5, g = (k >> 3) & 3u, e = k & 7u;
  return 16u * (2u * kap + (e >> 2)) + 4u * g + (e & 3u);
}

__device__ inline float hmd_nn_cutoff(float s) {
  // cos(pi s) through the native cosine, whose argument is in revolutions
  return 0.5f * (1.0f + __builtin_amdgcn_cosf(0.5f * s));
}

__device__ inline float hmd_nn_rbf(float u15, float k) {
  const float t = u15 - k;
  return __builtin_amdgcn_exp2f(-0.5f * HMD_NN_LOG2E * t * t);
}

// phi_{k0 + e}(x), e = 0..7, for x = 15 s, by the Gaussian product
// recurrence phi_{k+1} = phi_k exp(x - k - 1/2), whose ratio itself shrinks
// by e^-1 per step: two exponentials instead of eight.  From k0 = 0 the
// start value underflows for x > 13.2, where phi_0..7 are below 6e-9; from
// k0 = 8 it never does.
__device__ inline void hmd_nn_rbf8(float u15, float k0, float out[8]) {
  const float t = u15 - k0;
  float ph = __builtin_amdgcn_exp2f(-0.5f * HMD_NN_LOG2E * t * t);
  float r = __builtin_amdgcn_exp2f(HMD_NN_LOG2E * (t - 0.5f));
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    out[e] = ph;
    ph *= r;
    r *= 0.36787944117144233f;
  }
}

__device__ inline uint32_t hmd_nn_species(uint32_t mask, uint32_t type) {
  return (mask >> (type & 31u)) & 1u;
}

// ---------------------------------------------------------- packing -----
// One thread per bf16 operand element (3 sets x 4 x 2 x 64 x 8).
extern "C" __global__ void __launch_bounds__(256)
hmd_nn_pack(const NnPackParams p) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const float* W1 = p.master;
  const float* b1 = W1 + 64 * 51;
  const float* W2 = b1 + 64;
  const float* b2 = W2 + 64 * 64;
  const float* w3 = b2 + 64;
  const float* b3 = w3 + 64;
  if (tid < 3u * 4u * 2u * 64u * 8u) {
    const uint32_t e = tid & 7u, lane = (tid >> 3) & 63u, kap = (tid >> 9) & 1u;
    const uint32_t t = (tid >> 10) & 3u, set = tid >> 12;
    const uint32_t row = 16u * t + (lane & 15u);
    const uint32_t k = 32u * kap + 8u * (lane >> 4) + e;
    float v = 0.0f;
    if (set == 0u) {
      const int cin = hmd_nn_canon_of_internal(k);
      v = cin >= 0 ? W1[row * 51u + (uint32_t)cin] : 0.0f;
    } else if (set == 1u) {
      v = W2[row * 64u + hmd_nn_pi(k)];
    } else {
      v = W2[hmd_nn_pi(k) * 64u + row];
    }
    p.ops[tid] = hmd_bf16_bits(v);
  }
  if (tid < 64u) {
    p.vec[tid] = b1[tid];
    p.vec[64 + tid] = b2[tid];
    p.vec[128 + tid] = (float)(__bf16)w3[tid];
  }
  if (tid == 0u) p.vec[192] = b3[0];
}

// Single-image shift of a cluster (compare jit_composed.rs:1396-1430, as in
// the pair kernel): valid when every atom within r_cut + r_skin/2 of the
// cluster lies within half a box of its centre; positions are then moved to
// the image nearest the centre and pair differences need no minimum image.
struct HmdNnShift {
  bool on;
  float cx, cy, cz;
};

__device__ inline HmdNnShift hmd_nn_cluster_shift(const NnPassParams& p, uint32_t c) {
  const HmdLattice& L = p.L;
  HmdNnShift sh{false, 0.0f, 0.0f, 0.0f};
  if (L.xy != 0.0f || L.xz != 0.0f || L.yz != 0.0f) return sh;
  const float4* bb = reinterpret_cast<const float4*>(p.bbox);
  const float4 cb = bb[2 * c], hb = bb[2 * c + 1];
  sh.on = (hb.x + p.shift_margin < 0.5f * L.lx) && (hb.y + p.shift_margin < 0.5f * L.ly) &&
          (hb.z + p.shift_margin < 0.5f * L.lz);
  sh.cx = cb.x; sh.cy = cb.y; sh.cz = cb.z;
  return sh;
}

__device__ inline float4 hmd_nn_apply_shift(float4 a, const HmdNnShift& sh, const HmdLattice& L) {
  if (sh.on) {
    a.x -= L.lx * floorf((a.x - sh.cx) * (1.0f / L.lx) + 0.5f);
    a.y -= L.ly * floorf((a.y - sh.cy) * (1.0f / L.ly) + 0.5f);
    a.z -= L.lz * floorf((a.z - sh.cz) * (1.0f / L.lz) + 0.5f);
  }
  return a;
}

// ------------------------------------------------------- descriptors ----
// The pair kernel's traversal: one wave per contiguous share of the work
// list, lane = j-atom of the entry (shifted once to the image nearest the
// i-cluster's centre), the i-atoms broadcast from a per-wave LDS table one
// per step, over the 8-atom i-sub-clusters the entry's mask admits.  Per
// pair the 16 basis values come from two exponentials each half (the
// product recurrence of hmd_nn_rbf8) with the cutoff folded into their
// start; the j-atom's sums stay in the lane for the whole entry; the
// i-atom's 16 sums of the step are reduced over the wave by two transposed
// butterflies (totals of component m and 8 + m in lane 8m + 7) and added to
// the wave's i accumulators in LDS.  Sums go to i64 fixed point (2^32): the
// j side per entry, the i side per cluster share of the wave.
extern "C" __global__ void __launch_bounds__(256)
hmd_nn_descriptors(const NnPassParams p) {
#pragma clang fp contract(off)
  __shared__ float4 s_ip[4][HMD_CLUSTER];
  // i accumulators, component k of atom l at k * 65 + l (the odd stride
  // keeps the 8 lanes that add one atom's components on different banks)
  __shared__ float s_di[4][16 * 65];
  const uint32_t lane = hmd_lane();
  const uint32_t wv = threadIdx.x / HMD_WAVE;
  float* di = s_di[wv];
#pragma unroll
  for (int k = 0; k < 16; ++k) di[k * 65 + lane] = 0.0f;
  const uint32_t waves_per_block = blockDim.x / HMD_WAVE;
  const uint32_t n_waves = gridDim.x * waves_per_block;
  const uint32_t gw = blockIdx.x * waves_per_block + wv;
  const uint32_t nw = *p.n_work;
  const uint32_t per = (nw + n_waves - 1) / n_waves;
  const uint32_t w0 = gw * per;
  const uint32_t w1 = min(w0 + per, nw);
  if (w0 >= w1) return;
  const float4* ps = reinterpret_cast<const float4*>(p.posq_sorted);
  const HmdLattice L = p.L;
  long long* acc = p.desc_acc;
  const uint32_t stride = p.n_pad;

  uint32_t cur_c = HMD_SENTINEL, si = 0;
  bool vi = false;
  HmdNnShift sh{false, 0.0f, 0.0f, 0.0f};
  auto flush_i = [&]() {
    hmd_wave_sync();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const float v = di[k * 65 + lane];
      di[k * 65 + lane] = 0.0f;
      if (vi) hmd_atomic_add_fixed(&acc[k * stride + si], hmd_to_fixed(v, HMD_NN_DESC_SHIFT));
    }
  };
  for (uint32_t w = w0; w < w1; ++w) {
    const uint32_t slot = p.work[w];
    const uint32_t c = slot / p.cap_entries;
    const uint32_t c0 = c * HMD_CLUSTER;
    if (c != cur_c) {
      if (cur_c != HMD_SENTINEL) flush_i();
      cur_c = c;
      si = c0 + lane;
      vi = si < p.n;
      sh = hmd_nn_cluster_shift(p, c);
      float4 a = hmd_nn_apply_shift(ps[si], sh, L);
      if (!vi) a.x = a.y = a.z = 1.0e18f;   // padding lanes fail every cutoff test
      hmd_wave_sync();
      s_ip[wv][lane] = a;
      hmd_wave_sync();
    }
    const uint32_t my_j = p.entry_j[(size_t)slot * HMD_ENTRY + lane];
    const bool valid_j = my_j != HMD_SENTINEL;
    float4 pj = hmd_nn_apply_shift(ps[valid_j ? my_j : 0u], sh, L);
    if (!valid_j) pj.x = pj.y = pj.z = 1.0e18f;
    // entries whose j-atoms all lie in later clusters need only the cutoff
    // test (i < j holds); the cluster's own entry also needs i < j
    const bool later = __ballot(valid_j && my_j < c0 + HMD_CLUSTER) == 0ull;
    uint32_t pend = __builtin_amdgcn_readfirstlane(p.entry_mask[slot]) & 0xFFu;
    float dj[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) dj[k] = 0.0f;
    while (pend) {
      const uint32_t sb = (uint32_t)__builtin_ctz(pend);
      pend &= pend - 1u;
#pragma unroll 2
      for (uint32_t u = 0; u < 8; ++u) {
        const uint32_t i = 8u * sb + u;
        const float4 a = s_ip[wv][i];
        float dx = a.x - pj.x, dy = a.y - pj.y, dz = a.z - pj.z;
        if (!sh.on) hmd_min_image(dx, dy, dz, L);
        const float r2 = dx * dx + dy * dy + dz * dz;
        const bool ok = (r2 < p.r_cut2) && (r2 > 0.0f) && (later || c0 + i < my_j);
        if (__ballot(ok) == 0ull) continue;
        const float s = __builtin_sqrtf(ok ? r2 : 0.0f) * p.inv_r_cut;
        const float fc = ok ? hmd_nn_cutoff(s) : 0.0f;
        float w[16];
        hmd_nn_rbf8(15.0f * s, 0.0f, w);
        hmd_nn_rbf8(15.0f * s, 8.0f, w + 8);
        float vlo[8], vhi[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const float wl = w[k] * fc, wh = w[8 + k] * fc;
          dj[k] = fmaf(a.w, wl, dj[k]);
          dj[8 + k] = fmaf(a.w, wh, dj[8 + k]);
          vlo[k] = pj.w * wl;
          vhi[k] = pj.w * wh;
        }
        const float tlo = hmd_wave_sum8(vlo, lane), thi = hmd_wave_sum8(vhi, lane);
        if ((lane & 7u) == 7u) {
          const uint32_t m = lane >> 3;
          di[m * 65 + i] += tlo;          // one lane per address: program order
          di[(8u + m) * 65 + i] += thi;
        }
      }
    }
    if (valid_j) {
#pragma unroll
      for (int k = 0; k < 16; ++k)
        hmd_atomic_add_fixed(&acc[k * stride + my_j], hmd_to_fixed(dj[k], HMD_NN_DESC_SHIFT));
    }
  }
  flush_i();
}

// Fixed point -> f32 descriptors in sorted order (and canonical order when
// asked); the accumulators are cleared for the next evaluation.
extern "C" __global__ void __launch_bounds__(256)
hmd_nn_desc_finalize(const NnDescFinalizeParams p) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= p.n_pad) return;
  const uint32_t id = p.sorted_id[s];
#pragma unroll
  for (uint32_t k = 0; k < 16; ++k) {
    const float v = hmd_from_fixed(p.desc_acc[k * p.n_pad + s], HMD_NN_DESC_SHIFT);
    p.desc_acc[k * p.n_pad + s] = 0;
    p.desc[s * 16 + k] = v;
    if (p.desc_canon && s < p.n) p.desc_canon[id * 16 + k] = v;
  }
}

// ------------------------------------------------ network passes --------
struct HmdNnShared {
  hmd_bf16x8 ops[3][8][HMD_WAVE];          // W1, W2 (, W2^T) operands
  float vec[HMD_NN_HID * 3 + 1];           // b1, b2, w3, b3
};

// Per-wave tables of the current entry: its i-atoms (0..63) and j-atoms
// (64..127), positions shifted to the image nearest the cluster centre when
// the cluster qualifies, and the queue of in-range pairs.
struct HmdNnTables {
  float4 pos[2 * HMD_WAVE];
  uint32_t sp[2 * HMD_WAVE];               // species
  hmd_bf16x8 dsc[2 * HMD_WAVE][2];         // descriptors, bf16
  uint16_t queue[HMD_NN_QCAP];             // (i lane) | (j slot << 6)
};

struct HmdNnWave : HmdNnTables {           // backward pass
  float4 res[2 * HMD_WAVE];                // dL/d(dF) of the same atoms
};

struct HmdNnFwdWave : HmdNnTables {        // forward pass: per-batch pair table
  hmd_bf16x8 phi[32][2];                   // radial basis, bf16
  float4 geo[32];                          // fc, unit vector i - j
  uint32_t info[32];                       // i | j << 6 | one-hot << 12 | valid << 14
};

// Geometry and inputs of the two pairs lane l works on in a batch
// (pairs l%16 and 16 + l%16 of the batch).
struct HmdNnPairs {
  bool valid[2];
  uint32_t ia[2], jb[2];                   // wave-table rows of i and j
  float fc[2];
  float rh[2][3];
  hmd_bf16x8 B[4][2];                      // X^T operands: row tiles m, k-steps
};

__device__ inline void hmd_nn_gather_pairs(const NnPassParams& p, const HmdNnTables& W,
                                           uint32_t base, uint32_t cnt, uint32_t lane,
                                           HmdNnPairs& P) {
  const uint32_t g = lane >> 4, pl = lane & 15u;
  float phi[2][8];
  uint32_t oh[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint32_t idx = 16u * h + pl;
    P.valid[h] = idx < cnt;
    const uint32_t q = P.valid[h] ? W.queue[base + idx] : 0u;
    const uint32_t il = q & 63u, jk = q >> 6;
    P.ia[h] = il;
    P.jb[h] = HMD_WAVE + jk;
    const float4 a = W.pos[il], b = W.pos[HMD_WAVE + jk];
    float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    hmd_min_image(dx, dy, dz, p.L);
    const float r2 = P.valid[h] ? dx * dx + dy * dy + dz * dz : 1.0f;
    const float inv_r = __builtin_amdgcn_rsqf(r2);
    const float s = r2 * inv_r * p.inv_r_cut;
    P.fc[h] = P.valid[h] ? hmd_nn_cutoff(s) : 0.0f;
    P.rh[h][0] = dx * inv_r;
    P.rh[h][1] = dy * inv_r;
    P.rh[h][2] = dz * inv_r;
    oh[h] = W.sp[il] + W.sp[HMD_WAVE + jk];
    const float u15 = 15.0f * s;
    const float k0 = (float)(8u * (g & 1u));
    hmd_nn_rbf8(u15, k0, phi[h]);
  }
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int h = m & 1;
    const bool ij = m < 2;
    const uint32_t rowA = ij ? P.ia[h] : P.jb[h];
    const uint32_t rowB = ij ? P.jb[h] : P.ia[h];
    hmd_bf16x8 b0, b1;
    if (g < 2u) {
#pragma unroll
      for (int e = 0; e < 8; ++e) b0[e] = (__bf16)phi[h][e];
      b1 = W.dsc[rowB][g];
    } else {
      b0 = W.dsc[rowA][g - 2u];
#pragma unroll
      for (int e = 0; e < 8; ++e)
        b1[e] = (__bf16)((g == 2u && (uint32_t)e == oh[h]) ? 1.0f : 0.0f);
    }
    if (!P.valid[h]) {
#pragma unroll
      for (int e = 0; e < 8; ++e) { b0[e] = (__bf16)0.0f; b1[e] = (__bf16)0.0f; }
    }
    P.B[m][0] = b0;
    P.B[m][1] = b1;
  }
}

// Layers 1 and 2 for the 64 rows of a batch: z1 (pre-activation, bias
// included), the bf16 layer-2 operands, and z2 (bias included).
__device__ inline void hmd_nn_layers(const HmdNnShared& S, const HmdNnPairs& P, uint32_t lane,
                                     float z1[4][4][4], hmd_bf16x8 H1[4][2],
                                     float z2[4][4][4]) {
  const uint32_t g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const hmd_bf16x8 a0 = S.ops[0][2 * t][lane], a1 = S.ops[0][2 * t + 1][lane];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      hmd_f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
      c = hmd_mfma16(a0, P.B[m][0], c);
      c = hmd_mfma16(a1, P.B[m][1], c);
#pragma unroll
      for (int r = 0; r < 4; ++r) z1[t][m][r] = c[r] + S.vec[16 * t + 4 * g + r];
    }
  }
#pragma unroll
  for (int m = 0; m < 4; ++m) {
#pragma unroll
    for (int kap = 0; kap < 2; ++kap) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const float z = z1[2 * kap + (e >> 2)][m][e & 3];
        H1[m][kap][e] = (__bf16)(z * hmd_sigmoid(z));
      }
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const hmd_bf16x8 a0 = S.ops[1][2 * t][lane], a1 = S.ops[1][2 * t + 1][lane];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      hmd_f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
      c = hmd_mfma16(a0, H1[m][0], c);
      c = hmd_mfma16(a1, H1[m][1], c);
#pragma unroll
      for (int r = 0; r < 4; ++r) z2[t][m][r] = c[r] + S.vec[64 + 16 * t + 4 * g + r];
    }
  }
}

// Network output y of the four rows (row tiles m) of lane l, reduced over
// the four lane groups.
__device__ inline void hmd_nn_output(const HmdNnShared& S, const float z2[4][4][4],
                                     uint32_t lane, float y[4]) {
  const uint32_t g = lane >> 4;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    float acc = 0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float z = z2[t][m][r];
        acc += S.vec[128 + 16 * t + 4 * g + r] * (z * hmd_sigmoid(z));
      }
    acc += hmd_shfl_xor(acc, 16);
    acc += hmd_shfl_xor(acc, 32);
    y[m] = acc + S.vec[192];
  }
}

// Load the wave tables of a new i-cluster (lane = i-atom; a is shifted).
__device__ inline void hmd_nn_load_i(const NnPassParams& p, HmdNnTables& W, uint32_t si,
                                     float4 a, uint32_t lane) {
  W.pos[lane] = a;
  W.sp[lane] = hmd_nn_species(p.species_mask, p.type_sorted[si]);
  const float* dr = p.desc + (size_t)si * 16;
  hmd_bf16x8 d0, d1;
#pragma unroll
  for (int e = 0; e < 8; ++e) { d0[e] = (__bf16)dr[e]; d1[e] = (__bf16)dr[8 + e]; }
  W.dsc[lane][0] = d0;
  W.dsc[lane][1] = d1;
}

__device__ inline void hmd_nn_load_j(const NnPassParams& p, HmdNnTables& W, uint32_t jj,
                                     float4 b, uint32_t lane) {
  W.pos[HMD_WAVE + lane] = b;
  W.sp[HMD_WAVE + lane] = hmd_nn_species(p.species_mask, p.type_sorted[jj]);
  const float* dr = p.desc + (size_t)jj * 16;
  hmd_bf16x8 d0, d1;
#pragma unroll
  for (int e = 0; e < 8; ++e) { d0[e] = (__bf16)dr[e]; d1[e] = (__bf16)dr[8 + e]; }
  W.dsc[HMD_WAVE + lane][0] = d0;
  W.dsc[HMD_WAVE + lane][1] = d1;
}

__device__ inline float4 hmd_nn_resid(const NnPassParams& p, uint32_t s) {
  return make_float4(p.resid[s], p.resid[p.n_pad + s], p.resid[2 * p.n_pad + s], 0.0f);
}

// Append the in-range pairs of j-slots [kb, kb + 16) to the wave queue.
__device__ inline uint32_t hmd_nn_compact(const NnPassParams& p, HmdNnTables& W, uint32_t my_j,
                                          float4 pj, float xi, float yi, float zi, bool vi,
                                          uint32_t si, uint32_t c, bool shift, uint32_t kb,
                                          uint32_t qn, uint32_t lane) {
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll 4
  for (uint32_t u = 0; u < 16; ++u) {
    const uint32_t k = kb + u;
    const uint32_t j = hmd_readlane(my_j, k);
    if (j == HMD_SENTINEL) break;
    float dx = xi - hmd_readlane_f(pj.x, k);
    float dy = yi - hmd_readlane_f(pj.y, k);
    float dz = zi - hmd_readlane_f(pj.z, k);
    if (!shift) hmd_min_image(dx, dy, dz, p.L);
    const float r2 = dx * dx + dy * dy + dz * dz;
    const bool ok = vi && ((j >> 6) != c || si < j) && (r2 < p.r_cut2) && (r2 > 0.0f);
    const uint64_t m = __ballot(ok);
    if (ok) W.queue[qn + (uint32_t)__builtin_popcountll(m & lt)] = (uint16_t)(lane | (k << 6));
    qn += (uint32_t)__builtin_popcountll(m);
  }
  return qn;
}

__device__ inline void hmd_nn_load_shared(const NnPassParams& p, HmdNnShared& S, uint32_t sets) {
  const hmd_bf16x8* ops = reinterpret_cast<const hmd_bf16x8*>(p.ops);
  for (uint32_t k = threadIdx.x; k < sets * 8u * HMD_WAVE; k += blockDim.x)
    (&S.ops[0][0][0])[k] = ops[k];
  for (uint32_t k = threadIdx.x; k < HMD_NN_HID * 3 + 1; k += blockDim.x) S.vec[k] = p.vec[k];
  __syncthreads();
}

// Forward: dF into the fixed-point accumulators (sorted order).
//
// A batch is 32 pairs.  Its geometry and radial basis are computed once per
// pair (lanes l and l + 32 take pair l % 32, each half of the basis) into a
// per-wave LDS table; then the four row tiles (pairs 0-15 and 16-31, each in
// both orientations) go through the two layers one tile at a time, so only
// one tile's activations are live.
__device__ inline void hmd_nn_forward_batch(const NnPassParams& p, const HmdNnShared& S,
                                            HmdNnFwdWave& W, long long (*acc)[3][HMD_WAVE],
                                            uint32_t base, uint32_t cnt, bool shift,
                                            const float (&bias1)[4][4], const float (&bias2)[4][4],
                                            const float (&w3)[4][4], uint32_t lane) {
  const uint32_t g = lane >> 4, pl = lane & 15u;
  {
    const uint32_t pr = lane & 31u, hf = lane >> 5;
    const bool valid = pr < cnt;
    const uint32_t q = valid ? W.queue[base + pr] : 0u;
    const uint32_t il = q & 63u, jk = q >> 6;
    const float4 a = W.pos[il], b = W.pos[HMD_WAVE + jk];
    float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    if (!shift) hmd_min_image(dx, dy, dz, p.L);
    const float r2 = valid ? dx * dx + dy * dy + dz * dz : 1.0f;
    const float inv_r = __builtin_amdgcn_rsqf(r2);
    const float sc = r2 * inv_r * p.inv_r_cut;
    const float u15 = 15.0f * sc;
    const float k0 = (float)(8u * hf);
    hmd_bf16x8 ph;
    float phf[8];
    hmd_nn_rbf8(u15, k0, phf);
#pragma unroll
    for (int e = 0; e < 8; ++e) ph[e] = (__bf16)phf[e];
    W.phi[pr][hf] = ph;
    if (hf == 0u) {
      const float fc = valid ? hmd_nn_cutoff(sc) : 0.0f;
      W.geo[pr] = make_float4(fc, dx * inv_r, dy * inv_r, dz * inv_r);
      const uint32_t oh = W.sp[il] + W.sp[HMD_WAVE + jk];
      W.info[pr] = il | (jk << 6) | (oh << 12) | ((valid ? 1u : 0u) << 14);
    }
  }
  hmd_wave_sync();
  const uint32_t info0 = W.info[pl], info1 = W.info[16 + pl];
  float y[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const uint32_t info = (m & 1) ? info1 : info0;
    const uint32_t pr = 16u * (m & 1) + pl;
    const bool valid = (info >> 14) & 1u;
    const uint32_t il = info & 63u, jrow = HMD_WAVE + ((info >> 6) & 63u);
    const uint32_t oh = (info >> 12) & 3u;
    const uint32_t rowA = (m < 2) ? il : jrow;
    const uint32_t rowB = (m < 2) ? jrow : il;
    hmd_bf16x8 b0, b1;
    if (g < 2u) {
      b0 = W.phi[pr][g];
      b1 = W.dsc[rowB][g];
    } else {
      b0 = W.dsc[rowA][g - 2u];
#pragma unroll
      for (int e = 0; e < 8; ++e)
        b1[e] = (__bf16)((g == 2u && (uint32_t)e == oh) ? 1.0f : 0.0f);
    }
    if (!valid) {
#pragma unroll
      for (int e = 0; e < 8; ++e) { b0[e] = (__bf16)0.0f; b1[e] = (__bf16)0.0f; }
    }
    float h[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      hmd_f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
      c = hmd_mfma16(S.ops[0][2 * t][lane], b0, c);
      c = hmd_mfma16(S.ops[0][2 * t + 1][lane], b1, c);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float z = c[r] + bias1[t][r];
        h[t][r] = z * hmd_sigmoid(z);
      }
    }
    hmd_bf16x8 H0, H1;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      H0[e] = (__bf16)h[e >> 2][e & 3];
      H1[e] = (__bf16)h[2 + (e >> 2)][e & 3];
    }
    float acc_y = 0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      hmd_f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
      c = hmd_mfma16(S.ops[1][2 * t][lane], H0, c);
      c = hmd_mfma16(S.ops[1][2 * t + 1][lane], H1, c);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float z = c[r] + bias2[t][r];
        acc_y += w3[t][r] * (z * hmd_sigmoid(z));
      }
    }
    acc_y += hmd_shfl_xor(acc_y, 16);
    acc_y += hmd_shfl_xor(acc_y, 32);
    y[m] = acc_y + S.vec[192];
  }
  // lanes of group 0 apply pair l%16, lanes of group 1 pair 16 + l%16
  if (g < 2u) {
    const uint32_t info = g ? info1 : info0;
    if ((info >> 14) & 1u) {
      const float4 geo = W.geo[16u * g + pl];
      const float ysum = g ? (y[1] + y[3]) : (y[0] + y[2]);
      const float gp = p.force_scale * geo.x * 0.5f * ysum;
      const uint32_t ia = info & 63u, jk = (info >> 6) & 63u;
      const float rh[3] = {geo.y, geo.z, geo.w};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const long long f = hmd_to_fixed(gp * rh[d], HMD_NN_FORCE_SHIFT);
        atomicAdd(reinterpret_cast<unsigned long long*>(&acc[0][d][ia]),
                  static_cast<unsigned long long>(f));
        atomicAdd(reinterpret_cast<unsigned long long*>(&acc[1][d][jk]),
                  static_cast<unsigned long long>(-f));
      }
    }
  }
  hmd_wave_sync();
}

extern "C" __global__ void __launch_bounds__(256)
hmd_nn_forward(const NnPassParams p) {
  __shared__ HmdNnShared S;
  __shared__ HmdNnFwdWave Wv[4];
  __shared__ long long s_acc[4][2][3][HMD_WAVE];
  hmd_nn_load_shared(p, S, 2);
  const uint32_t lane = hmd_lane();
  const uint32_t wv = threadIdx.x / HMD_WAVE;
  const uint32_t n_waves = gridDim.x * (blockDim.x / HMD_WAVE);
  const uint32_t gw = blockIdx.x * (blockDim.x / HMD_WAVE) + wv;
  const uint32_t nw = *p.n_work;
  const uint32_t per = (nw + n_waves - 1) / n_waves;
  const uint32_t w0 = gw * per;
  const uint32_t w1 = min(w0 + per, nw);
  if (w0 >= w1) return;
  HmdNnFwdWave& W = Wv[wv];
  long long(*acc)[3][HMD_WAVE] = s_acc[wv];
  const float4* ps = reinterpret_cast<const float4*>(p.posq_sorted);
  const uint32_t stride = p.n_pad;
  const uint32_t g = lane >> 4;
  float bias1[4][4], bias2[4][4], w3[4][4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      bias1[t][r] = S.vec[16 * t + 4 * g + r];
      bias2[t][r] = S.vec[64 + 16 * t + 4 * g + r];
      w3[t][r] = S.vec[128 + 16 * t + 4 * g + r];
    }
#pragma unroll
  for (int d = 0; d < 3; ++d) { acc[0][d][lane] = 0; acc[1][d][lane] = 0; }

  uint32_t cur_c = HMD_SENTINEL, si = 0;
  bool vi = false;
  HmdNnShift sh{false, 0.0f, 0.0f, 0.0f};
  float4 ai = make_float4(0, 0, 0, 0);
  for (uint32_t w = w0; w < w1; ++w) {
    const uint32_t slot = p.work[w];
    const uint32_t c = slot / p.cap_entries;
    hmd_wave_sync();
    if (c != cur_c) {
      if (cur_c != HMD_SENTINEL) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const long long v = acc[0][d][lane];
          acc[0][d][lane] = 0;
          if (vi) hmd_atomic_add_fixed(&p.force_acc[d * stride + si], v);
        }
      }
      cur_c = c;
      si = c * HMD_CLUSTER + lane;
      vi = si < p.n;
      sh = hmd_nn_cluster_shift(p, c);
      ai = hmd_nn_apply_shift(ps[si], sh, p.L);
      hmd_nn_load_i(p, W, si, ai, lane);
    }
    const uint32_t my_j = p.entry_j[(size_t)slot * HMD_ENTRY + lane];
    const uint32_t jj = my_j != HMD_SENTINEL ? my_j : 0u;
    const float4 pj = hmd_nn_apply_shift(ps[jj], sh, p.L);
    hmd_nn_load_j(p, W, jj, pj, lane);
    hmd_wave_sync();
    uint32_t qn = 0;
    for (uint32_t kb = 0; kb < HMD_ENTRY; kb += 16) {
      if (hmd_readlane(my_j, kb) == HMD_SENTINEL) break;
      qn = hmd_nn_compact(p, W, my_j, pj, ai.x, ai.y, ai.z, vi, si, c, sh.on, kb, qn, lane);
      hmd_wave_sync();
      uint32_t head = 0;
      for (; qn - head >= 32u; head += 32u)
        hmd_nn_forward_batch(p, S, W, acc, head, 32u, sh.on, bias1, bias2, w3, lane);
      const uint32_t rem = qn - head;
      const uint16_t carry = lane < rem ? W.queue[head + lane] : (uint16_t)0;
      hmd_wave_sync();
      if (lane < rem) W.queue[lane] = carry;
      qn = rem;
      hmd_wave_sync();
    }
    if (qn > 0) hmd_nn_forward_batch(p, S, W, acc, 0, qn, sh.on, bias1, bias2, w3, lane);
    hmd_wave_sync();
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const long long v = acc[1][d][lane];
      acc[1][d][lane] = 0;
      if (my_j != HMD_SENTINEL) hmd_atomic_add_fixed(&p.force_acc[d * stride + my_j], v);
    }
  }
  hmd_wave_sync();
  if (vi) {
#pragma unroll
    for (int d = 0; d < 3; ++d) hmd_atomic_add_fixed(&p.force_acc[d * stride + si], acc[0][d][lane]);
  }
}

// Backward: weight gradients of sum over rows of dy * y, with dy from the
// residual dL/d(dF) of the two atoms; per-wave partials.
struct HmdNnGrad {
  hmd_f32x4 w1[4][4], w2[4][4];    // [hidden tile][input tile]
  float b1[4][4], b2[4][4], w3[4][4], b3;
};

__device__ inline void hmd_nn_backward_batch(const NnPassParams& p, const HmdNnShared& S,
                                             HmdNnWave& W, uint16_t (*T0)[HMD_NN_TSTRIDE],
                                             uint16_t (*T1)[HMD_NN_TSTRIDE], HmdNnGrad& G,
                                             uint32_t base, uint32_t cnt, uint32_t lane) {
  HmdNnPairs P;
  hmd_nn_gather_pairs(p, W, base, cnt, lane, P);
  float z1[4][4][4], z2[4][4][4];
  hmd_bf16x8 H1[4][2];
  hmd_nn_layers(S, P, lane, z1, H1, z2);
  const uint32_t g = lane >> 4, pl = lane & 15u;
  float dy[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const float4 ri = W.res[P.ia[h]], rj = W.res[P.jb[h]];
    const float dg = (ri.x - rj.x) * P.rh[h][0] + (ri.y - rj.y) * P.rh[h][1] +
                     (ri.z - rj.z) * P.rh[h][2];
    dy[h] = P.valid[h] ? dg * p.force_scale * P.fc[h] * 0.5f : 0.0f;
  }
  if (g == 0u) G.b3 += 2.0f * (dy[0] + dy[1]);
  // output layer and layer-2 error signal (kept in z2)
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float z = z2[t][m][r];
        const float sg = hmd_sigmoid(z);
        const float dyv = dy[m & 1];
        G.w3[t][r] += dyv * (z * sg);
        const float dz = S.vec[128 + 16 * t + 4 * g + r] * dyv * sg * (1.0f + z * (1.0f - sg));
        G.b2[t][r] += dz;
        z2[t][m][r] = dz;
      }
  // dH1^T = W2^T dZ2^T (chained), then layer-1 error signal (kept in z1)
  hmd_bf16x8 D2[4][2];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int kap = 0; kap < 2; ++kap)
#pragma unroll
      for (int e = 0; e < 8; ++e) D2[m][kap][e] = (__bf16)z2[2 * kap + (e >> 2)][m][e & 3];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const hmd_bf16x8 a0 = S.ops[2][2 * t][lane], a1 = S.ops[2][2 * t + 1][lane];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      hmd_f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
      c = hmd_mfma16(a0, D2[m][0], c);
      c = hmd_mfma16(a1, D2[m][1], c);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float z = z1[t][m][r];
        const float sg = hmd_sigmoid(z);
        const float dz = c[r] * sg * (1.0f + z * (1.0f - sg));
        G.b1[t][r] += dz;
        z1[t][m][r] = dz;
      }
    }
  }
  // dW2 += dZ2^T H1 over the 64 rows: both operands transposed through LDS
  hmd_wave_sync();
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t hrow = 16u * t + 4u * g + r, col = 16u * m + pl;
        T0[hrow][col] = hmd_bf16_bits(z2[t][m][r]);
        T1[hrow][col] = hmd_bf16x8_bits(H1[m][t >> 1], 4 * (t & 1) + r);
      }
  hmd_wave_sync();
#pragma unroll
  for (int to = 0; to < 4; ++to) {
    const hmd_bf16x8 a0 = *reinterpret_cast<const hmd_bf16x8*>(&T0[16 * to + pl][8 * g]);
    const hmd_bf16x8 a1 = *reinterpret_cast<const hmd_bf16x8*>(&T0[16 * to + pl][32 + 8 * g]);
#pragma unroll
    for (int ti = 0; ti < 4; ++ti) {
      const hmd_bf16x8 b0 = *reinterpret_cast<const hmd_bf16x8*>(&T1[16 * ti + pl][8 * g]);
      const hmd_bf16x8 b1 = *reinterpret_cast<const hmd_bf16x8*>(&T1[16 * ti + pl][32 + 8 * g]);
      G.w2[to][ti] = hmd_mfma16(a0, b0, G.w2[to][ti]);
      G.w2[to][ti] = hmd_mfma16(a1, b1, G.w2[to][ti]);
    }
  }
  // dW1 += dZ1^T X over the rows
  hmd_wave_sync();
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        T0[16u * t + 4u * g + r][16u * m + pl] = hmd_bf16_bits(z1[t][m][r]);
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int kap = 0; kap < 2; ++kap)
#pragma unroll
      for (int e = 0; e < 8; ++e)
        T1[32u * kap + 8u * g + e][16u * m + pl] = hmd_bf16x8_bits(P.B[m][kap], e);
  hmd_wave_sync();
#pragma unroll
  for (int th = 0; th < 4; ++th) {
    const hmd_bf16x8 a0 = *reinterpret_cast<const hmd_bf16x8*>(&T0[16 * th + pl][8 * g]);
    const hmd_bf16x8 a1 = *reinterpret_cast<const hmd_bf16x8*>(&T0[16 * th + pl][32 + 8 * g]);
#pragma unroll
    for (int tf = 0; tf < 4; ++tf) {
      const hmd_bf16x8 b0 = *reinterpret_cast<const hmd_bf16x8*>(&T1[16 * tf + pl][8 * g]);
      const hmd_bf16x8 b1 = *reinterpret_cast<const hmd_bf16x8*>(&T1[16 * tf + pl][32 + 8 * g]);
      G.w1[th][tf] = hmd_mfma16(a0, b0, G.w1[th][tf]);
      G.w1[th][tf] = hmd_mfma16(a1, b1, G.w1[th][tf]);
    }
  }
  hmd_wave_sync();
}

extern "C" __global__ void __launch_bounds__(256)
hmd_nn_backward(const NnPassParams p) {
  __shared__ HmdNnShared S;
  __shared__ HmdNnWave Wv[4];
  __shared__ __attribute__((aligned(16))) uint16_t s_t[4][2][HMD_NN_HID][HMD_NN_TSTRIDE];
  hmd_nn_load_shared(p, S, 3);
  const uint32_t lane = hmd_lane();
  const uint32_t wv = threadIdx.x / HMD_WAVE;
  const uint32_t n_waves = gridDim.x * (blockDim.x / HMD_WAVE);
  const uint32_t gw = blockIdx.x * (blockDim.x / HMD_WAVE) + wv;
  const uint32_t nw = *p.n_work;
  const uint32_t per = (nw + n_waves - 1) / n_waves;
  const uint32_t w0 = gw * per;
  const uint32_t w1 = min(w0 + per, nw);
  HmdNnWave& W = Wv[wv];
  const float4* ps = reinterpret_cast<const float4*>(p.posq_sorted);

  HmdNnGrad G;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      G.w1[a][b] = hmd_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      G.w2[a][b] = hmd_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      G.b1[a][b] = G.b2[a][b] = G.w3[a][b] = 0.0f;
    }
  G.b3 = 0.0f;

  uint32_t cur_c = HMD_SENTINEL, si = 0;
  bool vi = false;
  HmdNnShift sh{false, 0.0f, 0.0f, 0.0f};
  float4 ai = make_float4(0, 0, 0, 0);
  for (uint32_t w = w0; w < w1; ++w) {
    const uint32_t slot = p.work[w];
    const uint32_t c = slot / p.cap_entries;
    if (c != cur_c) {
      hmd_wave_sync();
      cur_c = c;
      si = c * HMD_CLUSTER + lane;
      vi = si < p.n;
      sh = hmd_nn_cluster_shift(p, c);
      ai = hmd_nn_apply_shift(ps[si], sh, p.L);
      hmd_nn_load_i(p, W, si, ai, lane);
      W.res[lane] = hmd_nn_resid(p, si);
    }
    const uint32_t my_j = p.entry_j[(size_t)slot * HMD_ENTRY + lane];
    const uint32_t jj = my_j != HMD_SENTINEL ? my_j : 0u;
    hmd_wave_sync();
    const float4 pj = hmd_nn_apply_shift(ps[jj], sh, p.L);
    hmd_nn_load_j(p, W, jj, pj, lane);
    W.res[HMD_WAVE + lane] = hmd_nn_resid(p, jj);
    hmd_wave_sync();
    uint32_t qn = 0;
    for (uint32_t kb = 0; kb < HMD_ENTRY; kb += 16) {
      if (hmd_readlane(my_j, kb) == HMD_SENTINEL) break;
      qn = hmd_nn_compact(p, W, my_j, pj, ai.x, ai.y, ai.z, vi, si, c, sh.on, kb, qn, lane);
      hmd_wave_sync();
      uint32_t head = 0;
      for (; qn - head >= 32u; head += 32u)
        hmd_nn_backward_batch(p, S, W, s_t[wv][0], s_t[wv][1], G, head, 32u, lane);
      const uint32_t rem = qn - head;
      const uint16_t carry = lane < rem ? W.queue[head + lane] : (uint16_t)0;
      hmd_wave_sync();
      if (lane < rem) W.queue[lane] = carry;
      qn = rem;
      hmd_wave_sync();
    }
    if (qn > 0) hmd_nn_backward_batch(p, S, W, s_t[wv][0], s_t[wv][1], G, 0, qn, lane);
  }

  float* out = p.partials + (size_t)gw * HMD_NN_PARTIAL;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        out[((a * 4 + b) * 4 + r) * 64 + lane] = G.w1[a][b][r];
        out[4096 + ((a * 4 + b) * 4 + r) * 64 + lane] = G.w2[a][b][r];
      }
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      out[8192 + (t * 4 + r) * 64 + lane] = G.b1[t][r];
      out[9216 + (t * 4 + r) * 64 + lane] = G.b2[t][r];
      out[10240 + (t * 4 + r) * 64 + lane] = G.w3[t][r];
    }
  out[11264 + lane] = G.b3;
}

// Canonical gradient from the per-wave partials: one workgroup per
// parameter; threads take waves w = tid, tid + 256, ... (and the lanes that
// hold the parameter), then a fixed LDS tree: FP64, a fixed summation order.
extern "C" __global__ void __launch_bounds__(256)
hmd_nn_grad_reduce(const NnGradReduceParams p) {
  __shared__ double s_sum[256];
  const uint32_t k = blockIdx.x;
  uint32_t off = 0, n_lanes = 1;
  if (k < 64u * 51u) {                                   // W1[h][cin]
    const uint32_t h = k / 51u, cin = k % 51u;
    const uint32_t f = cin < 16u ? cin : (cin < 19u ? cin + 32u : cin - 3u);
    const uint32_t lane = 16u * ((h & 15u) >> 2) + (f & 15u);
    off = (((h >> 4) * 4u + (f >> 4)) * 4u + (h & 3u)) * 64u + lane;
  } else if (k < 64u * 51u + 64u) {                      // b1[h]
    const uint32_t h = k - 64u * 51u;
    off = 8192u + ((h >> 4) * 4u + (h & 3u)) * 64u + 16u * ((h & 15u) >> 2);
    n_lanes = 16;
  } else if (k < 64u * 51u + 64u + 4096u) {              // W2[o][i]
    const uint32_t q = k - (64u * 51u + 64u), o = q >> 6, i = q & 63u;
    const uint32_t lane = 16u * ((o & 15u) >> 2) + (i & 15u);
    off = 4096u + (((o >> 4) * 4u + (i >> 4)) * 4u + (o & 3u)) * 64u + lane;
  } else if (k < 64u * 51u + 128u + 4096u) {             // b2[h]
    const uint32_t h = k - (64u * 51u + 64u + 4096u);
    off = 9216u + ((h >> 4) * 4u + (h & 3u)) * 64u + 16u * ((h & 15u) >> 2);
    n_lanes = 16;
  } else if (k < 64u * 51u + 192u + 4096u) {             // w3[h]
    const uint32_t h = k - (64u * 51u + 128u + 4096u);
    off = 10240u + ((h >> 4) * 4u + (h & 3u)) * 64u + 16u * ((h & 15u) >> 2);
    n_lanes = 16;
  } else {                                               // b3
    off = 11264u;
    n_lanes = 64;
  }
  double acc = 0.0;
  for (uint32_t w = threadIdx.x; w < p.n_waves; w += blockDim.x) {
    const float* src = p.partials + (size_t)w * HMD_NN_PARTIAL + off;
    for (uint32_t l = 0; l < n_lanes; ++l) acc += (double)src[l];
  }
  s_sum[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) s_sum[threadIdx.x] += s_sum[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) p.grad[k] += s_sum[0];
}

// ------------------------------------------------------------ helpers ---
extern "C" __global__ void __launch_bounds__(256)
hmd_nn_finalize(const NnFinalizeParams p) {
  const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= p.n) return;
  const uint32_t s = p.slot_of[a];
#pragma unroll
  for (uint32_t d = 0; d < 3; ++d) {
    const float v = hmd_from_fixed(p.force_acc[d * p.n_pad + s], HMD_NN_FORCE_SHIFT);
    p.force_acc[d * p.n_pad + s] = 0;
    p.df[d * p.n + a] = v;
    if (p.add_to) p.add_to[d * p.n + a] += v;
  }
}

extern "C" __global__ void __launch_bounds__(256)
hmd_nn_residual(const NnResidualParams p) {
  __shared__ double s[256];
  const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (a < p.n) {
    const uint32_t slot = p.slot_of[a];
#pragma unroll
    for (uint32_t d = 0; d < 3; ++d) {
      const float r = p.df[d * p.n + a] - p.target[d * p.n + a];
      p.resid[d * p.n_pad + slot] = p.scale * r;
      acc += (double)r * (double)r;
    }
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) p.partials[blockIdx.x] = s[0] * p.loss_scale;
}

// Adam (Kingma & Ba 2015, Algorithm 1) on the f32 master copy; the
// accumulated FP64 gradient is cleared.
extern "C" __global__ void __launch_bounds__(256)
hmd_nn_adam(const NnAdamParams p) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= HMD_NN_PARAMS) return;
  const float g = (float)p.grad[k];
  p.grad[k] = 0.0;
  const float m = p.beta1 * p.m[k] + (1.0f - p.beta1) * g;
  const float v = p.beta2 * p.v[k] + (1.0f - p.beta2) * g * g;
  p.m[k] = m;
  p.v[k] = v;
  const float mh = m / p.bc1, vh = v / p.bc2;
  p.master[k] -= p.lr * mh / (sqrtf(vh) + p.eps);
}

extern "C" __global__ void __launch_bounds__(256)
hmd_nn_record(const NnRecordParams p) {
  const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= p.n) return;
  reinterpret_cast<float4*>(p.frame_posq)[a] = reinterpret_cast<const float4*>(p.posq)[a];
#pragma unroll
  for (uint32_t d = 0; d < 3; ++d)
    p.frame_target[d * p.n + a] = p.fine_f[d * p.n + a] - p.coarse_f[d * p.n + a];
}

extern "C" __global__ void __launch_bounds__(256)
hmd_nn_diff(const NnDiffParams p) {
  __shared__ double s[2][256];
  const uint32_t chunk = (p.n + gridDim.x - 1) / gridDim.x;
  const uint32_t beg = blockIdx.x * chunk, end = min(beg + chunk, p.n);
  double e = 0.0, b2 = 0.0;
  for (uint32_t i = beg + threadIdx.x; i < end; i += blockDim.x) {
#pragma unroll
    for (uint32_t d = 0; d < 3; ++d) {
      const double a = p.a[d * p.n + i], b = p.b[d * p.n + i];
      e += (a - b) * (a - b);
      b2 += b * b;
    }
  }
  s[0][threadIdx.x] = e;
  s[1][threadIdx.x] = b2;
  __syncthreads();
  for (uint32_t st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      s[0][threadIdx.x] += s[0][threadIdx.x + st];
      s[1][threadIdx.x] += s[1][threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    p.partials[blockIdx.x] = s[0][0];
    p.partials[gridDim.x + blockIdx.x] = s[1][0];
  }
}
