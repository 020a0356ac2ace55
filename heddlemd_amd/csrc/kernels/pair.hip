// Fused LJ 12-6 (+C1 switch) + real-space Ewald pair kernel for gfx950.
//
// Physics is the reference's composed pair kernel: LjPairFunctor
// (src/forces/lj.rs:395-445), SpmeRealPairFunctor with the Abramowitz &
// Stegun 7.1.26 erfc (src/forces/spme.rs:712-760), exclusion scale 0 for
// listed partners, branch-free max-cutoff mask, Newton's third law, and the
// half/half energy and virial split (src/forces/jit_composed.rs:1302-1643).
// The real-space Ewald force and energy functions of r (the A&S erfc
// formula, evaluated in FP64 on the host) are read from a 512-bin cubic
// Hermite table in LDS (relative interpolation error below 3e-7 for
// r >= 0.09 nm), which replaces the exp, reciprocal and polynomial of the
// closed form by one LDS gather and three FMAs per pair.
//
// Mapping (wave64, VALU FP32; no MFMA -- there is no dense contraction):
//  * one wave = one 64-atom i-cluster, one i-atom per lane, held in VGPRs;
//  * an entry's 64 j-atoms are gathered with one load per lane (and shifted
//    to the image nearest the cluster centre) into a per-wave LDS table,
//    then broadcast one per step by a wave-uniform LDS read, so each step
//    evaluates 64 pairs with no global memory access and no shuffles;
//  * slot combinations (LJ, Coulomb, energies) are compile-time variants,
//    and LJ is skipped for j-types with no LJ partner (uniform branch);
//  * entries whose j-atoms all lie in later clusters and carry no exclusion
//    (all but the first one or two of a cluster) run a step variant with the
//    cutoff test alone;
//  * the j-side reaction forces of 8 consecutive j-atoms are summed across
//    the wave by a transposed butterfly on the gfx950 lane swaps
//    (v_permlane32_swap, v_permlane16_swap) and DPP row operations; the 3
//    (with energy and virial 5) components of the 8 j-atoms are then packed
//    into 24 (40) lanes that convert to fixed point and issue their atomics
//    together;
//  * i-forces stay in registers across every entry of the same cluster in the
//    wave's contiguous share of the work list and are flushed once;
//  * accumulation is i64 fixed point (2^40), so the per-atom totals do not
//    depend on the order in which waves finish: runs are bit-reproducible.
#include "hmd_device.h"

#define HMD_MAX_TYPES 16u

struct HmdPairState {
  float xi, yi, zi, kqi;
  uint32_t si, ti, ti_row;  // sorted index; type; type * n_types (LJ table row)
  bool vi;
  float fix, fiy, fiz, eiv, wiv;
};

// Steps over the j-atoms of one entry (64 pairs per step).  PLAIN entries
// (every valid j-atom in a later cluster, no exclusion bit in any lane) need
// only the cutoff test: i < j holds by construction, the i-atoms are all
// valid (only the last cluster is partial, and it has no later cluster), and
// empty slots carry a far-away position so that they fail the cutoff test.
template <bool FEV, bool LJ, bool COUL, bool SHIFT, bool PLAIN>
__device__ inline void hmd_pair_steps(const PairParams& p, const float4* s_lj,
                                      const float4* s_ct, const float4* s_jp,
                                      const uint2* s_jt,
                                      HmdPairState& st, uint32_t my_j, uint64_t my_ex,
                                      uint64_t lj_bits, uint32_t lane) {
  const HmdLattice L = p.L;
  const float x_max = (float)HMD_COUL_BINS - 0.01f;   // last bin: beyond r_cut, masked
  const float lj_rs2_d = p.lj_rs2 * p.lj_inv_delta;
  const float lj_12d = 12.0f * p.lj_inv_delta;
  constexpr uint32_t NC = FEV ? 5u : 3u;        // reaction components per j
  const uint32_t grp = lane & 7u;
  const uint32_t comp = 7u - grp;               // component packed into this lane
  long long* acc = p.acc;
  const uint32_t stride = p.acc_stride;

  for (uint32_t kb = 0; kb < HMD_ENTRY; kb += 8) {
    if (hmd_readlane(my_j, kb) == HMD_SENTINEL) break;
    // jr[c][u]: this lane's contribution to component c of j-atom kb + u
    // (forces with the i-side sign; the j-side reaction is their negation)
    float jr[NC][8];
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u) {
      const uint32_t k = kb + u;
      const float4 bj = s_jp[k];
      float dx = st.xi - bj.x, dy = st.yi - bj.y, dz = st.zi - bj.z;
      if (!SHIFT) hmd_min_image(dx, dy, dz, L);
      const float r2 = dx * dx + dy * dy + dz * dz;
      bool ok = r2 <= p.rc2;
      if (!PLAIN) {
        const uint32_t j = __builtin_amdgcn_readfirstlane(s_jt[k].x);   // wave-uniform
        const bool excluded = (my_ex >> k) & 1ull;
        ok = ok && (j != HMD_SENTINEL) && st.vi && (st.si < j) && !excluded;
      }
      // PLAIN: every pair has r2 > 0 (distinct atoms; empty slots far away),
      // so the masked-out pairs evaluate to finite values without a guard
      const float r2s = PLAIN ? r2 : (ok ? r2 : 1.0f);
      const float inv_r = __builtin_amdgcn_rsqf(r2s);
      const float inv_r2 = inv_r * inv_r;
      float factor = 0.0f, energy = 0.0f;
      if (COUL) {
        // tabulated erfc kernel: bin b = floor(r/h), cubic in t = r/h - b
        const float qq = st.kqi * bj.w;
        const float x = fminf(r2s * inv_r * p.coul_inv_h, x_max);
        const float t = __builtin_amdgcn_fractf(x);
        const uint32_t b = (uint32_t)x;
        const float4 g = s_ct[b];
        factor = qq * (((g.w * t + g.z) * t + g.y) * t + g.x);
        if (FEV) {
          const float4 e = s_ct[HMD_COUL_BINS + b];
          energy = qq * (((e.w * t + e.z) * t + e.y) * t + e.x);
        }
      }
      if (LJ) {
        if ((lj_bits >> k) & 1ull) {                        // uniform branch
          // (c6, c12, 6 c6, 12 c12)
          const float4 cc = s_lj[st.ti_row + s_jt[k].y];
          const float ir6 = inv_r2 * inv_r2 * inv_r2;
          const float e_lj = (cc.y * ir6 - cc.x) * ir6;
          const float f_lj = (cc.w * ir6 - cc.z) * ir6 * inv_r2;
          // C1 switch: tau = (r^2 - rs^2)/(rc^2 - rs^2) clamped at 0,
          // S = (1 - tau)^2 (1 + 2 tau), force gains -dS/dr / r * E
          const float tau = fmaxf(r2s * p.lj_inv_delta - lj_rs2_d, 0.0f);
          const float omt = 1.0f - tau;
          const float sw = omt * omt * (1.0f + 2.0f * tau);
          const float chain = (lj_12d * tau) * omt;
          factor += sw * f_lj + chain * e_lj;
          if (FEV) energy += sw * e_lj;
        }
      }
#ifdef HMD_USER_FUNCTOR
      {
        // a user pair term composed in at run time: force / r and energy
        // from (types, r^2, 1/r); masked with the built-in terms below
        float f_u = 0.0f, e_u = 0.0f;
        HMD_USER_FUNCTOR{}(p.user_params, st.ti, s_jt[k].y, r2s, inv_r, f_u, e_u);
        factor += f_u;
        if (FEV) energy += e_u;
      }
#endif
      factor = ok ? factor : 0.0f;
      // the i-side sums use their own FMAs, so that the j-side values are
      // dead once the lane swaps consume them (no copies)
      st.fix = fmaf(factor, dx, st.fix);
      st.fiy = fmaf(factor, dy, st.fiy);
      st.fiz = fmaf(factor, dz, st.fiz);
      jr[0][u] = factor * dx; jr[1][u] = factor * dy; jr[2][u] = factor * dz;
      if (FEV) {
        const float he = ok ? 0.5f * energy : 0.0f;
        const float hw = 0.5f * factor * r2s;
        st.eiv += he; st.wiv += hw;
        jr[3][u] = he; jr[4][u] = hw;
      }
    }
    // Wave sums of the 8 j-atoms' components (totals in lanes 8m + 7), then
    // component c of j-atom kb + m moves to lane 8m + 7 - c, so that one
    // conversion and one atomic per lane cover every component.
    float tot[NC];
#pragma unroll
    for (uint32_t c = 0; c < NC; ++c) tot[c] = hmd_wave_sum8(jr[c], lane);
    float sh[NC];
    sh[0] = tot[0];
    sh[1] = hmd_dpp<HMD_DPP_ROW_SHL(1)>(tot[1]);
    sh[2] = hmd_dpp<HMD_DPP_ROW_SHL(2)>(tot[2]);
    if (FEV) {
      sh[3 % NC] = hmd_dpp<HMD_DPP_ROW_SHL(3)>(tot[3 % NC]);
      sh[4 % NC] = hmd_dpp<HMD_DPP_ROW_SHL(4)>(tot[4 % NC]);
    }
    // the moves run with all lanes enabled (an opaque use pins them here,
    // ahead of the branch below, where a disabled source lane would read 0)
#pragma unroll
    for (uint32_t c = 0; c < NC; ++c) asm volatile("" : "+v"(sh[c]));
    float v = sh[0];
#pragma unroll
    for (uint32_t c = 1; c < NC; ++c) v = (comp == c) ? sh[c] : v;
    const uint32_t jt = s_jt[kb + (lane >> 3)].x;
    if (comp < NC && jt != HMD_SENTINEL) {
      v = (comp < 3u) ? -v : v;                    // forces react; E, W halves add
      hmd_atomic_add_fixed(&acc[comp * stride + jt], hmd_to_fixed(v, HMD_FIXED_SHIFT));
    }
  }
}

// One entry: the 64 j-atoms are gathered once (one lane per j), shifted to
// the image nearest the cluster centre when SHIFT, and stored in a per-wave
// LDS table that the steps read at wave-uniform addresses (LDS broadcasts).
template <bool FEV, bool LJ, bool COUL, bool SHIFT>
__device__ inline void hmd_pair_entry(const PairParams& p, const float4* s_lj,
                                      const float4* s_ct, float4* s_jp, uint2* s_jt, HmdPairState& st,
                                      uint32_t slot, float cx, float cy, float cz,
                                      uint32_t c_end, bool has_ex, uint32_t lane) {
  const float4* ps = reinterpret_cast<const float4*>(p.posq_sorted);
  const HmdLattice L = p.L;
  const size_t ebase = (size_t)slot * HMD_ENTRY;
  const uint32_t my_j = p.entry_j[ebase + lane];
  // masks exist only for the cluster's leading entries that hold a partner
  const uint64_t my_ex = has_ex ? p.entry_excl[ebase + lane] : 0ull;
  const bool valid_j = my_j != HMD_SENTINEL;
  const uint32_t jj_l = valid_j ? my_j : 0u;
  float4 pj = ps[jj_l];
  const uint32_t tj_l = p.type_sorted[jj_l];
  if (SHIFT) {
    pj.x -= L.lx * floorf((pj.x - cx) * (1.0f / L.lx) + 0.5f);
    pj.y -= L.ly * floorf((pj.y - cy) * (1.0f / L.ly) + 0.5f);
    pj.z -= L.lz * floorf((pj.z - cz) * (1.0f / L.lz) + 0.5f);
    if (!valid_j) pj.x = pj.y = pj.z = 1.0e18f;   // fails every cutoff test
  }
  const uint64_t lj_bits = LJ ? __ballot(valid_j && ((p.lj_type_mask >> tj_l) & 1u)) : 0ull;
  const bool plain = SHIFT && __ballot(my_ex != 0ull || my_j < c_end) == 0ull;
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  s_jp[lane] = pj;
  s_jt[lane] = make_uint2(my_j, tj_l);
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (SHIFT && plain)
    hmd_pair_steps<FEV, LJ, COUL, true, true>(p, s_lj, s_ct, s_jp, s_jt, st, my_j, my_ex,
                                               lj_bits, lane);
  else
    hmd_pair_steps<FEV, LJ, COUL, SHIFT, false>(p, s_lj, s_ct, s_jp, s_jt, st, my_j, my_ex,
                                                 lj_bits, lane);
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
__device__ inline void hmd_pair_body(const PairParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  __shared__ float4 s_lj[HMD_MAX_TYPES * HMD_MAX_TYPES];   // (c6, c12, 6 c6, 12 c12)
  // erfc table: force coefficients, then (energy variants) energy coefficients
  __shared__ float4 s_ct[COUL ? (FEV ? 2 * HMD_COUL_BINS : HMD_COUL_BINS) : 1];
  __shared__ float4 s_jp[4][HMD_ENTRY];
  __shared__ uint2 s_jt[4][HMD_ENTRY];
  const uint32_t nt2 = p.n_types * p.n_types;
  for (uint32_t k = threadIdx.x; k < nt2; k += blockDim.x) {
    const float2 cc = reinterpret_cast<const float2*>(p.lj_c6c12)[k];
    s_lj[k] = make_float4(cc.x, cc.y, 6.0f * cc.x, 12.0f * cc.y);
  }
  if (COUL) {
    const float4* ct = reinterpret_cast<const float4*>(p.coul_table);
    for (uint32_t k = threadIdx.x; k < (FEV ? 2u : 1u) * HMD_COUL_BINS; k += blockDim.x)
      s_ct[k] = ct[k];
  }
  __syncthreads();

  const uint32_t lane = hmd_lane();
  const uint32_t waves_per_block = blockDim.x / HMD_WAVE;
  const uint32_t n_waves = hmd_gx * waves_per_block;
  const uint32_t gw = hmd_bx * waves_per_block + threadIdx.x / HMD_WAVE;
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
  uint32_t cur_nex = 0;
  bool shift = false;
  float cx = 0, cy = 0, cz = 0;
  for (uint32_t w = w0; w < w1; ++w) {
    const uint32_t slot = p.work[w];
    const uint32_t c = slot / p.cap_entries;
    if (c != cur_c) {
      if (cur_c != HMD_SENTINEL) hmd_flush_i<FEV>(p, st);
      cur_c = c;
      cur_nex = p.excl_entries[c];
      st.si = c * HMD_CLUSTER + lane;
      st.vi = st.si < p.n;
      const float4 a = ps[st.si];
      st.xi = a.x; st.yi = a.y; st.zi = a.z;
      st.kqi = p.k_coulomb * a.w;
      st.ti = p.type_sorted[st.si];
      st.ti_row = st.ti * p.n_types;
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
    const bool has_ex = slot - c * p.cap_entries < cur_nex;
    if (shift)
      hmd_pair_entry<FEV, LJ, COUL, true>(p, s_lj, s_ct, s_jp[wv], s_jt[wv], st, slot, cx, cy, cz,
                                          (c + 1u) * HMD_CLUSTER, has_ex, lane);
    else
      hmd_pair_entry<FEV, LJ, COUL, false>(p, s_lj, s_ct, s_jp[wv], s_jt[wv], st, slot, cx, cy, cz,
                                           (c + 1u) * HMD_CLUSTER, has_ex, lane);
  }
  hmd_flush_i<FEV>(p, st);
}

// Minimum waves per SIMD the register allocation must allow (1: no limit).
#ifndef HMD_PAIR_WAVES_PER_EU
#define HMD_PAIR_WAVES_PER_EU 1
#endif
#define HMD_PAIR_VARIANT(NAME, FEV, LJ, COUL)                                  \
  __device__ __forceinline__ void NAME##_body(const PairParams& p,             \
                                              const uint32_t hmd_bx,           \
                                              const uint32_t hmd_gx) {         \
    hmd_pair_body<FEV, LJ, COUL>(p, hmd_bx, hmd_gx);                           \
  }                                                                            \
  HMD_KERNEL(NAME, PairParams, 256, HMD_PAIR_WAVES_PER_EU)
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
__device__ __forceinline__ void hmd_pair_finalize_body(const PairFinalizeParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  const uint32_t a = hmd_bx * blockDim.x + threadIdx.x;
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
HMD_KERNEL(hmd_pair_finalize, PairFinalizeParams, 256)

// Scaled exclusions: one thread per listed pair with a non-zero scale, the
// same per-pair arithmetic as the fused kernel (full minimum image, the
// cutoff mask, LJ 12-6 with the C1 switch, the tabulated real-space Ewald
// term) times the pair's scales, the exclusion-scaled sum of the reference
// (src/forces/jit_composed.rs:939-972 with heddle_jit_exclusion_scale
// :1260-1272).  Fixed-point atomics make the totals independent of order.
__device__ __forceinline__ void hmd_pair_scaled_body(const PairScaledParams& p,
                                                     const uint32_t hmd_bx,
                                                     const uint32_t hmd_gx) {
  const uint32_t k = hmd_bx * blockDim.x + threadIdx.x;
  if (k >= p.n_pairs) return;
  const uint32_t i = p.pair_i[k], j = p.pair_j[k];
  const float4 a = reinterpret_cast<const float4*>(p.posq)[i];
  const float4 b = reinterpret_cast<const float4*>(p.posq)[j];
  float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  hmd_min_image(dx, dy, dz, p.L);
  const float r2 = dx * dx + dy * dy + dz * dz;
  if (!(r2 <= p.rc2) || r2 == 0.0f) return;
  const float inv_r = __builtin_amdgcn_rsqf(r2);
  const float inv_r2 = inv_r * inv_r;
  float factor = 0.0f, energy = 0.0f;
  if (p.coul_on && p.s_coul[k] != 0.0f) {
    const float qq = p.k_coulomb * a.w * b.w;
    const float x = fminf(r2 * inv_r * p.coul_inv_h, (float)HMD_COUL_BINS - 0.01f);
    const float t = __builtin_amdgcn_fractf(x);
    const uint32_t bin = (uint32_t)x;
    const float4 g = reinterpret_cast<const float4*>(p.coul_table)[bin];
    factor += p.s_coul[k] * (qq * (((g.w * t + g.z) * t + g.y) * t + g.x));
    if (p.fev) {
      const float4 e = reinterpret_cast<const float4*>(p.coul_table)[HMD_COUL_BINS + bin];
      energy += p.s_coul[k] * (qq * (((e.w * t + e.z) * t + e.y) * t + e.x));
    }
  }
  if (p.lj_on && p.s_lj[k] != 0.0f) {
    const float2 cc =
        reinterpret_cast<const float2*>(p.lj_c6c12)[p.type[i] * p.n_types + p.type[j]];
    const float ir6 = inv_r2 * inv_r2 * inv_r2;
    const float e_lj = (cc.y * ir6 - cc.x) * ir6;
    const float f_lj = (12.0f * cc.y * ir6 - 6.0f * cc.x) * ir6 * inv_r2;
    const float tau = fmaxf((r2 - p.lj_rs2) * p.lj_inv_delta, 0.0f);
    const float omt = 1.0f - tau;
    const float sw = omt * omt * (1.0f + 2.0f * tau);
    const float chain = (12.0f * p.lj_inv_delta * tau) * omt;
    factor += p.s_lj[k] * (sw * f_lj + chain * e_lj);
    if (p.fev) energy += p.s_lj[k] * (sw * e_lj);
  }
  const uint32_t si = p.slot_of[i], sj = p.slot_of[j], st = p.acc_stride;
  const float f3[3] = {factor * dx, factor * dy, factor * dz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const long long v = hmd_to_fixed(f3[d], HMD_FIXED_SHIFT);
    hmd_atomic_add_fixed(&p.acc[d * st + si], v);
    hmd_atomic_add_fixed(&p.acc[d * st + sj], -v);
  }
  if (p.fev) {
    const long long he = hmd_to_fixed(0.5f * energy, HMD_FIXED_SHIFT);
    const long long hw = hmd_to_fixed(0.5f * factor * r2, HMD_FIXED_SHIFT);
    hmd_atomic_add_fixed(&p.acc[3 * st + si], he);
    hmd_atomic_add_fixed(&p.acc[3 * st + sj], he);
    hmd_atomic_add_fixed(&p.acc[4 * st + si], hw);
    hmd_atomic_add_fixed(&p.acc[4 * st + sj], hw);
  }
}
HMD_KERNEL(hmd_pair_scaled, PairScaledParams, 256)
