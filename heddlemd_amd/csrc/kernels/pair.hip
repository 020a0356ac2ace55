// Fused LJ 12-6 (+C1 switch) + real-space Ewald pair kernel for gfx950.
//
// Physics is the reference's composed pair kernel: LjPairFunctor
// (src/forces/lj.rs:395-445), SpmeRealPairFunctor with the Abramowitz &
// Stegun 7.1.26 erfc (src/forces/spme.rs:712-760), exclusion scale 0 for
// listed partners, branch-free max-cutoff mask, Newton's third law, and the
// half/half energy and virial split (src/forces/jit_composed.rs:1302-1643).
//
// Mapping (wave64, VALU FP32; no MFMA -- there is no dense contraction):
//  * one wave = one 64-atom i-cluster, one i-atom per lane, held in VGPRs;
//  * an entry's 64 j-atoms are gathered with one load per lane (and shifted
//    to the image nearest the cluster centre) into a per-wave LDS table,
//    then broadcast one per step by a wave-uniform LDS read, so each step
//    evaluates 64 pairs with no global memory access and no shuffles;
//  * slot combinations (LJ, Coulomb, energies) are compile-time variants,
//    and LJ is skipped for j-types with no LJ partner (uniform branch);
//  * the j-side reaction forces of 8 consecutive j-atoms are summed across
//    the wave by a transposed butterfly (8 values -> 1 per lane group), then
//    8 lanes issue the fixed-point atomics;
//  * i-forces stay in registers across every entry of the same cluster in the
//    wave's contiguous share of the work list and are flushed once;
//  * accumulation is i64 fixed point (2^40), so the per-atom totals do not
//    depend on the order in which waves finish: runs are bit-reproducible.
#include "hmd_device.h"

#define HMD_MAX_TYPES 16u
#define HMD_TWO_OVER_SQRT_PI 1.1283791670955126f
#define HMD_LOG2E 1.4426950408889634f

struct HmdPairState {
  float xi, yi, zi, kqi;
  uint32_t si, ti;
  bool vi;
  float fix, fiy, fiz, eiv, wiv;
};

// One entry: the 64 j-atoms of the entry are gathered once (one lane per
// j), shifted to the image nearest the cluster centre when SHIFT, and then
// broadcast one per step with readlane.
template <bool FEV, bool LJ, bool COUL, bool SHIFT>
__device__ inline void hmd_pair_entry(const PairParams& p, const float2* s_lj,
                                      float4* s_jp, uint2* s_jt, HmdPairState& st,
                                      uint32_t slot, float cx, float cy, float cz,
                                      uint32_t lane) {
  const float4* ps = reinterpret_cast<const float4*>(p.posq_sorted);
  const HmdLattice L = p.L;
  const size_t ebase = (size_t)slot * HMD_ENTRY;
  const uint32_t my_j = p.entry_j[ebase + lane];
  const uint64_t my_ex = p.entry_excl[ebase + lane];
  const uint32_t jj_l = (my_j != HMD_SENTINEL) ? my_j : 0u;
  float4 pj = ps[jj_l];
  const uint32_t tj_l = p.type_sorted[jj_l];
  if (SHIFT) {
    pj.x -= L.lx * floorf((pj.x - cx) * (1.0f / L.lx) + 0.5f);
    pj.y -= L.ly * floorf((pj.y - cy) * (1.0f / L.ly) + 0.5f);
    pj.z -= L.lz * floorf((pj.z - cz) * (1.0f / L.lz) + 0.5f);
  }
  const float c_gauss = -(p.alpha * p.alpha) * HMD_LOG2E;
  const float two_alpha_sqrtpi = p.alpha * HMD_TWO_OVER_SQRT_PI;
  long long* acc = p.acc;
  const uint32_t stride = p.acc_stride;
  // The entry's j-atoms go to a per-wave LDS table; each step reads one
  // record at a wave-uniform address (an LDS broadcast, no VALU work).
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  s_jp[lane] = pj;
  s_jt[lane] = make_uint2(my_j, tj_l);
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();

  for (uint32_t kb = 0; kb < HMD_ENTRY; kb += 8) {
    if (hmd_readlane(my_j, kb) == HMD_SENTINEL) break;
    float jfx[8], jfy[8], jfz[8], jfe[8], jfw[8];
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u) {
      const uint32_t k = kb + u;
      const float4 bj = s_jp[k];
      const uint2 jt2 = s_jt[k];
      const uint32_t j = __builtin_amdgcn_readfirstlane(jt2.x);   // wave-uniform
      const float xj = bj.x, yj = bj.y, zj = bj.z;
      float dx = st.xi - xj, dy = st.yi - yj, dz = st.zi - zj;
      if (!SHIFT) hmd_min_image(dx, dy, dz, L);
      const float r2 = dx * dx + dy * dy + dz * dz;
      const bool excluded = (my_ex >> k) & 1ull;
      const bool ok = (j != HMD_SENTINEL) && st.vi && (st.si < j) && (r2 <= p.rc2) && !excluded;
      const float r2s = ok ? r2 : 1.0f;
      const float inv_r = __builtin_amdgcn_rsqf(r2s);
      const float inv_r2 = inv_r * inv_r;
      float factor = 0.0f, energy = 0.0f;
      if (COUL) {
        const float qj = bj.w;
        const float qq = st.kqi * qj;
        const float ar = p.alpha * (r2s * inv_r);
        const float gauss = __builtin_amdgcn_exp2f(r2s * c_gauss);
        const float t = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * ar);
        const float erfc_ar =
            ((((1.061405429f * t - 1.453152027f) * t + 1.421413741f) * t -
              0.284496736f) * t + 0.254829592f) * t * gauss;
        const float qq_ir2 = qq * inv_r2;
        factor = qq_ir2 * (erfc_ar * inv_r + two_alpha_sqrtpi * gauss);
        if (FEV) energy = qq * inv_r * erfc_ar;
      }
      if (LJ) {
        const uint32_t tj = __builtin_amdgcn_readfirstlane(jt2.y);
        if ((p.lj_type_mask >> tj) & 1u) {                  // uniform branch
          const float2 cc = s_lj[st.ti * p.n_types + tj];
          const float ir6 = inv_r2 * inv_r2 * inv_r2;
          float e_lj = (cc.y * ir6 - cc.x) * ir6;
          float f_lj = (12.0f * cc.y * ir6 - 6.0f * cc.x) * ir6 * inv_r2;
          float tau = (r2s - p.lj_rs2) * p.lj_inv_delta;
          tau = (r2s > p.lj_rs2) ? tau : 0.0f;
          const float omt = 1.0f - tau;
          const float sw = omt * omt * (1.0f + 2.0f * tau);
          const float chain = 12.0f * tau * omt * p.lj_inv_delta;
          f_lj = sw * f_lj + chain * e_lj;
          factor += f_lj;
          if (FEV) energy += sw * e_lj;
        }
      }
      factor = ok ? factor : 0.0f;
      const float fx = factor * dx, fy = factor * dy, fz = factor * dz;
      st.fix += fx; st.fiy += fy; st.fiz += fz;
      jfx[u] = -fx; jfy[u] = -fy; jfz[u] = -fz;
      if (FEV) {
        const float he = ok ? 0.5f * energy : 0.0f;
        const float hw = 0.5f * factor * r2s;
        st.eiv += he; st.wiv += hw;
        jfe[u] = he; jfw[u] = hw;
      }
    }
    const float rx = hmd_transpose_reduce8(jfx, lane);
    const float ry = hmd_transpose_reduce8(jfy, lane);
    const float rz = hmd_transpose_reduce8(jfz, lane);
    float re = 0.0f, rw = 0.0f;
    if (FEV) {
      re = hmd_transpose_reduce8(jfe, lane);
      rw = hmd_transpose_reduce8(jfw, lane);
    }
    const uint32_t jt = __shfl(my_j, (int)(kb + (lane >> 3)), HMD_WAVE);
    if ((lane & 7u) == 0u && jt != HMD_SENTINEL) {
      hmd_atomic_add_fixed(&acc[jt], hmd_to_fixed(rx, HMD_FIXED_SHIFT));
      hmd_atomic_add_fixed(&acc[stride + jt], hmd_to_fixed(ry, HMD_FIXED_SHIFT));
      hmd_atomic_add_fixed(&acc[2 * stride + jt], hmd_to_fixed(rz, HMD_FIXED_SHIFT));
      if (FEV) {
        hmd_atomic_add_fixed(&acc[3 * stride + jt], hmd_to_fixed(re, HMD_FIXED_SHIFT));
        hmd_atomic_add_fixed(&acc[4 * stride + jt], hmd_to_fixed(rw, HMD_FIXED_SHIFT));
      }
    }
  }
}

template <bool FEV>
__device__ inline void hmd_flush_i(const PairParams& p, const HmdPairState& st) {
  if (!st.vi) return;
  long long* acc = p.acc;
  const uint32_t stride = p.acc_stride, si = st.si;
  hmd_atomic_add_fixed(&acc[si], hmd_to_fixed(st.fix, HMD_FIXED_SHIFT));
  hmd_atomic_add_fixed(&acc[stride + si], hmd_to_fixed(st.fiy, HMD_FIXED_SHIFT));
  hmd_atomic_add_fixed(&acc[2 * stride + si], hmd_to_fixed(st.fiz, HMD_FIXED_SHIFT));
  if (FEV) {
    hmd_atomic_add_fixed(&acc[3 * stride + si], hmd_to_fixed(st.eiv, HMD_FIXED_SHIFT));
    hmd_atomic_add_fixed(&acc[4 * stride + si], hmd_to_fixed(st.wiv, HMD_FIXED_SHIFT));
  }
}

template <bool FEV, bool LJ, bool COUL>
__device__ inline void hmd_pair_body(const PairParams& p) {
  __shared__ float2 s_lj[HMD_MAX_TYPES * HMD_MAX_TYPES];
  __shared__ float4 s_jp[4][HMD_ENTRY];
  __shared__ uint2 s_jt[4][HMD_ENTRY];
  const uint32_t nt2 = p.n_types * p.n_types;
  for (uint32_t k = threadIdx.x; k < nt2; k += blockDim.x)
    s_lj[k] = reinterpret_cast<const float2*>(p.lj_c6c12)[k];
  __syncthreads();

  const uint32_t lane = hmd_lane();
  const uint32_t waves_per_block = blockDim.x / HMD_WAVE;
  const uint32_t n_waves = gridDim.x * waves_per_block;
  const uint32_t gw = blockIdx.x * waves_per_block + threadIdx.x / HMD_WAVE;
  const uint32_t nw = *p.n_work;
  const uint32_t per = (nw + n_waves - 1) / n_waves;
  const uint32_t w0 = gw * per;
  const uint32_t w1 = min(w0 + per, nw);
  if (w0 >= w1) return;

  const float4* ps = reinterpret_cast<const float4*>(p.posq_sorted);
  const float4* bb = reinterpret_cast<const float4*>(p.bbox);
  const HmdLattice L = p.L;
  const bool ortho = (L.xy == 0.0f) && (L.xz == 0.0f) && (L.yz == 0.0f);

  HmdPairState st{};
  uint32_t cur_c = HMD_SENTINEL;
  bool shift = false;
  float cx = 0, cy = 0, cz = 0;
  for (uint32_t w = w0; w < w1; ++w) {
    const uint32_t slot = p.work[w];
    const uint32_t c = slot / p.cap_entries;
    if (c != cur_c) {
      if (cur_c != HMD_SENTINEL) hmd_flush_i<FEV>(p, st);
      cur_c = c;
      st.si = c * HMD_CLUSTER + lane;
      st.vi = st.si < p.n;
      const float4 a = ps[st.si];
      st.xi = a.x; st.yi = a.y; st.zi = a.z;
      st.kqi = p.k_coulomb * a.w;
      st.ti = p.type_sorted[st.si];
      const float4 cb = bb[2 * c], hb = bb[2 * c + 1];
      // Single-image fast path (compare jit_composed.rs:1396-1430): valid
      // when every atom that can interact with the cluster lies within half
      // a box length of its centre, so shifting i and j to the image nearest
      // the centre yields the minimum-image displacement of every in-cutoff
      // pair.  The margin covers motion since the rebuild.
      shift = ortho && (hb.x + p.shift_margin < 0.5f * L.lx) &&
              (hb.y + p.shift_margin < 0.5f * L.ly) &&
              (hb.z + p.shift_margin < 0.5f * L.lz);
      cx = cb.x; cy = cb.y; cz = cb.z;
      if (shift) {
        st.xi -= L.lx * floorf((st.xi - cx) * (1.0f / L.lx) + 0.5f);
        st.yi -= L.ly * floorf((st.yi - cy) * (1.0f / L.ly) + 0.5f);
        st.zi -= L.lz * floorf((st.zi - cz) * (1.0f / L.lz) + 0.5f);
      }
      st.fix = st.fiy = st.fiz = st.eiv = st.wiv = 0.0f;
    }
    const uint32_t wv = threadIdx.x / HMD_WAVE;
    if (shift)
      hmd_pair_entry<FEV, LJ, COUL, true>(p, s_lj, s_jp[wv], s_jt[wv], st, slot, cx, cy, cz,
                                          lane);
    else
      hmd_pair_entry<FEV, LJ, COUL, false>(p, s_lj, s_jp[wv], s_jt[wv], st, slot, cx, cy, cz,
                                           lane);
  }
  hmd_flush_i<FEV>(p, st);
}

#define HMD_PAIR_VARIANT(NAME, FEV, LJ, COUL)                                  \
  extern "C" __global__ void __launch_bounds__(256) NAME(const PairParams p) { \
    hmd_pair_body<FEV, LJ, COUL>(p);                                          \
  }
HMD_PAIR_VARIANT(hmd_pair_force_f, false, true, true)
HMD_PAIR_VARIANT(hmd_pair_force_fev, true, true, true)
HMD_PAIR_VARIANT(hmd_pair_force_lj_f, false, true, false)
HMD_PAIR_VARIANT(hmd_pair_force_lj_fev, true, true, false)
HMD_PAIR_VARIANT(hmd_pair_force_coul_f, false, false, true)
HMD_PAIR_VARIANT(hmd_pair_force_coul_fev, true, false, true)

// Fixed point -> f32, sorted -> canonical order, plus the slow-class (SPME)
// forces: the reference's finalize_packed_forces + combine_class_totals
// (kernels/neighbor.cu:728-760, kernels/forces.cu:13-41).  The pair
// accumulators are cleared for the next evaluation as they are read.
extern "C" __global__ void __launch_bounds__(256)
hmd_pair_finalize(const PairFinalizeParams p) {
  const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= p.n) return;
  const uint32_t s = p.slot_of[a];
  const uint32_t st = p.acc_stride;
  long long* acc = p.acc_clear;
  float f[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    f[d] = hmd_from_fixed(p.acc[d * st + s], HMD_FIXED_SHIFT);
    acc[d * st + s] = 0;
    if (p.slow_f) f[d] += p.slow_f[d * p.n + a];
    p.f[d * p.n + a] = f[d];
  }
  if (p.fev) {
    float e = hmd_from_fixed(p.acc[3 * st + s], HMD_FIXED_SHIFT);
    float w = hmd_from_fixed(p.acc[4 * st + s], HMD_FIXED_SHIFT);
    acc[3 * st + s] = 0;
    acc[4 * st + s] = 0;
    if (p.slow_e) e += p.slow_e[a];
    if (p.slow_w) w += p.slow_w[a];
    p.e[a] = e;
    p.w[a] = w;
  }
}
