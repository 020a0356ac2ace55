// Fused LJ 12-6 (+C1 switch) + real-space Ewald pair kernel for gfx950.
//
// Physics is the reference's composed pair kernel: LjPairFunctor
// (src/forces/lj.rs:395-445), SpmeRealPairFunctor with the Abramowitz &
// Stegun 7.1.26 erfc (src/forces/spme.rs:712-760), exclusion scale 0 for
// listed partners, branch-free max-cutoff mask, Newton's third law, and the
// half/half energy and virial split (src/forces/jit_composed.rs:1302-1643).
// The real-space Ewald force and energy functions (the A&S erfc formula,
// evaluated in FP64 on the host) are read from a cubic Hermite table in LDS
// over octaves of r^2, 128 bins per octave, whose bin and in-bin coordinate
// come from the f32 bits of r^2 (hmd_coul_bin): one clamp, two shifts and a
// bit insert replace the square root, the exp and the polynomial of the
// closed form, and no reciprocal is taken unless the step has LJ.
//
// Mapping (wave64, VALU FP32; no MFMA -- there is no dense contraction):
//  * one wave walks a contiguous share of the work list; an entry holds 64
//    j-atoms of one 64-atom i-cluster's list, one j-atom per lane, gathered
//    once (and shifted to the image nearest the cluster centre);
//  * the i-cluster's atoms sit in a per-wave LDS table and are broadcast one
//    per step by a wave-uniform LDS read, so each step evaluates 64 pairs
//    with no global memory access and no shuffles;
//  * an entry visits only the 8-atom i-sub-clusters that some of its j-atoms
//    can reach (the entry's 8-bit mask from the list build, whose placement
//    groups j-atoms of similar reach into the same entries): about two
//    thirds of the steps of visiting every i-atom;
//  * a step (one i-atom against the entry's 64 j-atoms) with no pair inside
//    the cutoff is skipped by a wave-uniform branch after the distance test;
//  * slot combinations (LJ, Coulomb, energies) are compile-time variants,
//    and LJ is skipped for i-types with no LJ partner (uniform branch);
//  * entries whose j-atoms all lie in later clusters and carry no exclusion
//    (most of a cluster's list) run a step variant with the cutoff test
//    alone; exclusion rows are transposed once per entry (lane = j-atom,
//    bits over the i-atoms);
//  * the i-side values of a sub-cluster's 8 atoms are summed across the wave
//    by a transposed butterfly on the gfx950 lane swaps (v_permlane32_swap,
//    v_permlane16_swap) and DPP row operations, packed one component per
//    lane and added to the wave's i accumulators in LDS; the j reaction stays
//    in the lane and is flushed once per entry;
//  * accumulation is i64 fixed point (2^40), each entry's i- and j-side sums
//    converted separately, so the per-atom totals depend neither on the
//    order in which waves finish nor on how the work list is split among
//    them: runs are bit-reproducible across launch shapes and GPUs.
#include "hmd_device.h"

#define HMD_MAX_TYPES 16u

// 64 x 64 bit-matrix transpose across the wave: lane r holds row r (bit k =
// column k) on entry and column r (bit k = row k) on return.  Six stages of
// block swaps with the lane 2^s apart.
__device__ inline uint64_t hmd_transpose64(uint64_t x, uint32_t lane) {
  const uint64_t masks[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                             0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
  uint32_t j = 32;
#pragma unroll
  for (int k = 0; k < 6; ++k, j >>= 1) {
    const uint64_t m = masks[k];
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)x, (int)j, HMD_WAVE);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(x >> 32), (int)j, HMD_WAVE);
    const uint64_t y = ((uint64_t)hi << 32) | lo;
    x = (lane & j) ? ((x & ~m) | ((y >> j) & m)) : ((x & m) | ((y << j) & ~m));
  }
  return x;
}

// The i-atoms of the sub-clusters an entry reaches, 64 pairs per step: lane
// k holds j-atom k of the entry; step u of sub-cluster b takes i-atom 8b + u
// from the wave's LDS table (a broadcast read).  PLAIN entries (every valid
// j-atom in a later cluster, no exclusion) need only the cutoff test: i < j
// holds by construction, the i-atoms are all valid (only the last cluster is
// partial, and it has no later cluster), and empty slots carry a far-away
// position so that they fail the cutoff test.  The j reaction accumulates in
// the lane (fj); the i-side values of a sub-cluster's 8 atoms are summed over
// the wave by the transposed butterfly and added to the wave's i
// accumulators in LDS.
template <bool FEV, bool LJ, bool COUL, bool SHIFT, bool PLAIN>
__device__ inline void hmd_pair_isteps(const PairParams& p, const float4* s_lj,
                                       const float4* s_ct, const float4* s_ip,
                                       const uint32_t* s_it, float* s_fi, float4 pj,
                                       uint32_t tj, float4 cc_single, uint32_t my_j,
                                       uint64_t exT,
                                       uint32_t emask, uint32_t c, uint64_t lj_ibits,
                                       float (&fj)[5], uint32_t lane) {
  // Every fused multiply-add below is written out: the solo kernel and its
  // batched twin then round identically (left to the compiler, a*b + c*d
  // may be fused on either side, differently in the two).
#pragma clang fp contract(off)
  const HmdLattice L = p.L;
  const float4* s_ctb = s_ct;
  const float lj_rs2_d = p.lj_rs2 * p.lj_inv_delta;
  const float lj_12d = 12.0f * p.lj_inv_delta;
  constexpr uint32_t NC = FEV ? 5u : 3u;        // components per atom
  const uint32_t comp = 7u - (lane & 7u);       // component packed into this lane
  const bool valid_j = my_j != HMD_SENTINEL;
  const uint32_t c0 = c * HMD_CLUSTER;
  uint32_t pend = emask;
  while (pend) {
    const uint32_t sb = (uint32_t)__builtin_ctz(pend);   // i-sub-cluster
    pend &= pend - 1u;
    // ir[c][u]: this lane's contribution to component c of i-atom 8 sb + u
    float ir[NC][8];
    // LJ flags of the sub-cluster's 8 i-atoms, in a scalar register
    const uint32_t lj8 = __builtin_amdgcn_readfirstlane((uint32_t)(lj_ibits >> (8u * sb)) & 0xFFu);
    // (x, y, z, k_C q) of the i-atoms, each read one step ahead
    float4 a_next = s_ip[8u * sb];
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u) {
      const uint32_t i = 8u * sb + u;
      const float4 a = a_next;
      if (u < 7) a_next = s_ip[i + 1];
      float dx = a.x - pj.x, dy = a.y - pj.y, dz = a.z - pj.z;
      if (!SHIFT) hmd_min_image(dx, dy, dz, L);
      const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
      bool ok = r2 <= p.rc2;
      if (!PLAIN) {
        const uint32_t si = c0 + i;
        ok = ok && valid_j && (si < p.n) && (si < my_j) && !((exT >> i) & 1ull);
      }
      // A step with no pair inside the cutoff in any lane (about a fifth of
      // the steps an entry's sub-cluster mask admits) is skipped as a whole:
      // a wave-uniform branch.
      if (__ballot(ok) == 0ull) {
#pragma unroll
        for (uint32_t cc = 0; cc < NC; ++cc) ir[cc][u] = 0.0f;
        continue;
      }
      // PLAIN: every pair has r2 > 0 (distinct atoms; empty slots far away),
      // so the masked-out pairs evaluate to finite values without a guard
      const float r2s = PLAIN ? r2 : (ok ? r2 : 1.0f);
      float factor = 0.0f, energy = 0.0f;
      if (COUL) {
        // tabulated erfc kernel: octave-table bin of r^2, cubic in its t
        const float qq = a.w * pj.w;
        float t;
        const uint32_t b = hmd_coul_bin(r2s, p.coul_u_lo, p.coul_u_hi, t);
        const float4 g = s_ctb[b];
        factor = qq * fmaf(fmaf(fmaf(g.w, t, g.z), t, g.y), t, g.x);
        if (FEV) {
          const float4 e = s_ctb[HMD_COUL_BINS + b];
          energy = qq * fmaf(fmaf(fmaf(e.w, t, e.z), t, e.y), t, e.x);
        }
      }
      if (LJ) {
        if ((lj8 >> u) & 1u) {                               // scalar branch
          // (c6, c12, 6 c6, 12 c12): held per entry when one type carries
          // all the LJ (the i-atom's type is then that one), else looked up
          float4 cc = cc_single;
          if (p.lj_single_type >= p.n_types) {
            const uint32_t ti = __builtin_amdgcn_readfirstlane(s_it[i]);
            cc = s_lj[ti * p.n_types + tj];
          }
          const float inv_r2 = __builtin_amdgcn_rcpf(r2s);
          const float ir6 = inv_r2 * inv_r2 * inv_r2;
          const float e_lj = fmaf(cc.y, ir6, -cc.x) * ir6;
          const float f_lj = fmaf(cc.w, ir6, -cc.z) * ir6 * inv_r2;
          // C1 switch: tau = (r^2 - rs^2)/(rc^2 - rs^2) clamped at 0,
          // S = (1 - tau)^2 (1 + 2 tau), force gains -dS/dr / r * E
          const float tau = fmaxf(fmaf(r2s, p.lj_inv_delta, -lj_rs2_d), 0.0f);
          const float omt = 1.0f - tau;
          const float sw = omt * omt * fmaf(2.0f, tau, 1.0f);
          const float chain = (lj_12d * tau) * omt;
          factor += fmaf(sw, f_lj, chain * e_lj);
          if (FEV) energy = fmaf(sw, e_lj, energy);
        }
      }
#ifdef HMD_USER_FUNCTOR
      {
        // a user pair term composed in at run time: force / r and energy
        // from (types, r^2, 1/r); masked with the built-in terms below
        float f_u = 0.0f, e_u = 0.0f;
        const uint32_t ti = __builtin_amdgcn_readfirstlane(s_it[i]);
        HMD_USER_FUNCTOR{}(p.user_params, ti, tj, r2s, __builtin_amdgcn_rsqf(r2s), f_u, e_u);
        factor += f_u;
        if (FEV) energy += e_u;
      }
#endif
      factor = ok ? factor : 0.0f;
      // the j reaction uses its own FMAs, so that the i-side values are dead
      // once the lane swaps consume them (no copies)
      fj[0] = fmaf(-factor, dx, fj[0]);
      fj[1] = fmaf(-factor, dy, fj[1]);
      fj[2] = fmaf(-factor, dz, fj[2]);
      ir[0][u] = factor * dx; ir[1][u] = factor * dy; ir[2][u] = factor * dz;
      if (FEV) {
        const float he = ok ? 0.5f * energy : 0.0f;
        const float hw = 0.5f * factor * r2s;
        fj[3] += he; fj[4] += hw;
        ir[3 % NC][u] = he; ir[4 % NC][u] = hw;
      }
    }
    // Wave sums of the 8 i-atoms' components (totals in lanes 8m + 7), then
    // component c of i-atom 8 sb + m moves to lane 8m + 7 - c, so that one
    // LDS add per lane covers every component.
    float tot[NC];
#pragma unroll
    for (uint32_t cc = 0; cc < NC; ++cc) tot[cc] = hmd_wave_sum8(ir[cc], lane);
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
    for (uint32_t cc = 0; cc < NC; ++cc) asm volatile("" : "+v"(sh[cc]));
    float v = sh[0];
#pragma unroll
    for (uint32_t cc = 1; cc < NC; ++cc) v = (comp == cc) ? sh[cc] : v;
    // LDS atomic add without return: no wait on the old value; one lane per
    // address, in program order, so the sums are the same every run
    if (comp < NC) atomicAdd(&s_fi[comp * HMD_CLUSTER + 8u * sb + (lane >> 3)], v);
  }
}

// One entry: lane k gathers j-atom k (shifted to the image nearest the
// cluster centre when SHIFT), the i-sub-clusters it reaches come from the
// entry's mask, and the entry's exclusion rows (lane = i-atom, bits over the
// j-slots) are transposed so that each j lane holds its bits over the
// i-atoms.  The j reaction is flushed once per entry.
template <bool FEV, bool LJ, bool COUL, bool SHIFT>
__device__ inline void hmd_pair_entry(const PairParams& p, const float4* s_lj,
                                      const float4* s_ct, const float4* s_ip,
                                      const uint32_t* s_it, float* s_fi, uint32_t slot,
                                      uint32_t c, float cx, float cy, float cz,
                                      uint64_t lj_ibits, uint32_t lane) {
#pragma clang fp contract(off)   // as in hmd_pair_isteps: fused multiply-adds written out
  const float4* ps = reinterpret_cast<const float4*>(p.posq_sorted);
  const HmdLattice L = p.L;
  const size_t ebase = (size_t)slot * HMD_ENTRY;
  const uint32_t my_j = p.entry_j[ebase + lane];
  // bits 0-7: i-sub-clusters reached; bit 8: the entry holds an exclusion
  const uint32_t emask = __builtin_amdgcn_readfirstlane(p.entry_mask[slot]);
  const uint64_t ex = (emask & 0x100u) ? p.entry_excl[ebase + lane] : 0ull;
  const bool valid_j = my_j != HMD_SENTINEL;
  const uint32_t jj_l = valid_j ? my_j : 0u;
  float4 pj = ps[jj_l];
  const uint32_t tj = p.type_sorted[jj_l];
  if (SHIFT) {
    pj.x = fmaf(-L.lx, floorf(fmaf(pj.x - cx, 1.0f / L.lx, 0.5f)), pj.x);
    pj.y = fmaf(-L.ly, floorf(fmaf(pj.y - cy, 1.0f / L.ly, 0.5f)), pj.y);
    pj.z = fmaf(-L.lz, floorf(fmaf(pj.z - cz, 1.0f / L.lz, 0.5f)), pj.z);
    if (!valid_j) pj.x = pj.y = pj.z = 1.0e18f;   // fails every cutoff test
  }
  const bool any_ex = __ballot(ex != 0ull) != 0ull;
  const bool plain = SHIFT && !any_ex && __ballot(my_j < (c + 1u) * HMD_CLUSTER) == 0ull;
  const uint64_t exT = any_ex ? hmd_transpose64(ex, lane) : 0ull;
  float fj[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  const float4 cc_single = (LJ && p.lj_single_type < p.n_types)
                               ? s_lj[p.lj_single_type * p.n_types + tj]
                               : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (SHIFT && plain)
    hmd_pair_isteps<FEV, LJ, COUL, true, true>(p, s_lj, s_ct, s_ip, s_it, s_fi, pj, tj,
                                                cc_single, my_j, exT, emask & 0xFFu, c,
                                                lj_ibits, fj, lane);
  else
    hmd_pair_isteps<FEV, LJ, COUL, SHIFT, false>(p, s_lj, s_ct, s_ip, s_it, s_fi, pj, tj,
                                                  cc_single, my_j, exT, emask & 0xFFu, c,
                                                  lj_ibits, fj, lane);
  if (valid_j) {
    constexpr uint32_t NC = FEV ? 5u : 3u;
    long long* acc = p.acc;
#pragma unroll
    for (uint32_t cc = 0; cc < NC; ++cc)
      hmd_atomic_add_fixed(&acc[cc * p.acc_stride + my_j], hmd_to_fixed(fj[cc], HMD_FIXED_SHIFT));
  }
}

// The wave's i accumulators (f32 sums over one entry) to fixed point, added
// to the lane's i64 running totals (integer adds: any grouping of entries
// gives the same bits), and cleared for the next entry.
template <bool FEV>
__device__ inline void hmd_fold_i(float* s_fi, long long (&fix)[5], uint32_t lane) {
  constexpr uint32_t NC = FEV ? 5u : 3u;
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (uint32_t cc = 0; cc < NC; ++cc) {
    fix[cc] += hmd_to_fixed(s_fi[cc * HMD_CLUSTER + lane], HMD_FIXED_SHIFT);
    s_fi[cc * HMD_CLUSTER + lane] = 0.0f;
  }
}

// The running totals of the wave's share of one cluster to the accumulators.
template <bool FEV>
__device__ inline void hmd_flush_i(const PairParams& p, long long (&fix)[5], uint32_t si,
                                   bool vi) {
  constexpr uint32_t NC = FEV ? 5u : 3u;
#pragma unroll
  for (uint32_t cc = 0; cc < NC; ++cc) {
    if (vi) hmd_atomic_add_fixed(&p.acc[cc * p.acc_stride + si], fix[cc]);
    fix[cc] = 0;
  }
}

template <bool FEV, bool LJ, bool COUL>
__device__ inline void hmd_pair_body(const PairParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
#pragma clang fp contract(off)   // as in hmd_pair_isteps: fused multiply-adds written out
  __shared__ float4 s_lj[HMD_MAX_TYPES * HMD_MAX_TYPES];   // (c6, c12, 6 c6, 12 c12)
  // erfc table: force coefficients, then (energy variants) energy coefficients
  __shared__ float4 s_ct[COUL ? (FEV ? 2 * HMD_COUL_BINS : HMD_COUL_BINS) : 1];
  // per wave: the current i-cluster (shifted positions with k_C q, types)
  // and its f32 force (energy, virial) accumulators
  __shared__ float4 s_ip[4][HMD_CLUSTER];
  __shared__ uint32_t s_it[4][HMD_CLUSTER];
  __shared__ float s_fi[4][(FEV ? 5 : 3) * HMD_CLUSTER];
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
  const uint32_t wv = threadIdx.x / HMD_WAVE;
  const uint32_t lane = hmd_lane();
  float* fi = s_fi[wv];
#pragma unroll
  for (uint32_t cc = 0; cc < (FEV ? 5u : 3u); ++cc) fi[cc * HMD_CLUSTER + lane] = 0.0f;
  __syncthreads();

  const uint32_t waves_per_block = blockDim.x / HMD_WAVE;
  const uint32_t n_waves = hmd_gx * waves_per_block;
  const uint32_t gw = hmd_bx * waves_per_block + wv;
  const uint32_t nw = *p.n_work;
  const uint32_t per = (nw + n_waves - 1) / n_waves;
  const uint32_t w0 = gw * per;
  const uint32_t w1 = min(w0 + per, nw);
  if (w0 >= w1) return;

  const float4* ps = reinterpret_cast<const float4*>(p.posq_sorted);
  const float4* bb = reinterpret_cast<const float4*>(p.bbox);
  const HmdLattice L = p.L;
  const bool ortho = (L.xy == 0.0f) && (L.xz == 0.0f) && (L.yz == 0.0f);

  uint32_t cur_c = HMD_SENTINEL;
  uint32_t si = 0;
  bool vi = false;
  uint64_t lj_ibits = 0;   // i-atoms of the cluster whose type has an LJ partner
  bool shift = false;
  float cx = 0, cy = 0, cz = 0;
  long long fix[5] = {0, 0, 0, 0, 0};
  for (uint32_t w = w0; w < w1; ++w) {
    const uint32_t slot = p.work[w];
    const uint32_t c = slot / p.cap_entries;
    if (c != cur_c) {
      if (cur_c != HMD_SENTINEL) hmd_flush_i<FEV>(p, fix, si, vi);
      cur_c = c;
      si = c * HMD_CLUSTER + lane;
      vi = si < p.n;
      float4 a = ps[si];
      const uint32_t ti = p.type_sorted[si];
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
        a.x = fmaf(-L.lx, floorf(fmaf(a.x - cx, 1.0f / L.lx, 0.5f)), a.x);
        a.y = fmaf(-L.ly, floorf(fmaf(a.y - cy, 1.0f / L.ly, 0.5f)), a.y);
        a.z = fmaf(-L.lz, floorf(fmaf(a.z - cz, 1.0f / L.lz, 0.5f)), a.z);
        if (!vi) a.x = a.y = a.z = 1.0e18f;   // padding lanes fail every cutoff test
      }
      a.w *= p.k_coulomb;
      {
        const uint64_t b = __ballot(vi && ((p.lj_type_mask >> ti) & 1u));
        // held in scalar registers: the per-step LJ test is a scalar branch
        // (readfirstlane returns int: widen through uint32_t, or a set bit 31
        // would sign-extend over the upper half)
        lj_ibits = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(b >> 32)) << 32) |
                   (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)b);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
      s_ip[wv][lane] = a;
      s_it[wv][lane] = ti;
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    if (shift)
      hmd_pair_entry<FEV, LJ, COUL, true>(p, s_lj, s_ct, s_ip[wv], s_it[wv], fi, slot, c, cx, cy,
                                          cz, lj_ibits, lane);
    else
      hmd_pair_entry<FEV, LJ, COUL, false>(p, s_lj, s_ct, s_ip[wv], s_it[wv], fi, slot, c, cx,
                                           cy, cz, lj_ibits, lane);
    // the entry's i-side sums go to fixed point on their own: the totals are
    // then the same however the work list is split among waves (launch
    // shape, CU count), not only run to run
    hmd_fold_i<FEV>(fi, fix, lane);
  }
  hmd_flush_i<FEV>(p, fix, si, vi);
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

// A slot's per-atom results added into caller-owned outputs (+=): the pair
// slots' fixed-point accumulators (cleared for the next evaluation as they
// are read) or the reciprocal slot's f32 arrays.  Starting from zeroed
// outputs, the pair slot then the reciprocal slot give the bytes the
// finalize above writes (each sum is one f32 rounding of the same two
// values).
__device__ __forceinline__ void hmd_slot_add_body(const SlotAddParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  const uint32_t a = hmd_bx * blockDim.x + threadIdx.x;
  if (a >= p.n) return;
  float v[5];
  if (p.acc) {
    const uint32_t s = p.slot_of[a];
    const uint32_t st = p.acc_stride;
    const uint32_t nc = p.fev ? 5u : 3u;
    for (uint32_t d = 0; d < nc; ++d) {
      v[d] = hmd_from_fixed(p.acc[d * st + s], HMD_FIXED_SHIFT);
      p.acc_clear[d * st + s] = 0;
    }
  } else {
    for (uint32_t d = 0; d < 3; ++d) v[d] = p.src_f[d * p.n + a];
    if (p.fev) { v[3] = p.src_e[a]; v[4] = p.src_w[a]; }
  }
  p.out_x[a] = p.out_x[a] + v[0];
  p.out_y[a] = p.out_y[a] + v[1];
  p.out_z[a] = p.out_z[a] + v[2];
  if (p.fev) {
    if (p.out_e) p.out_e[a] = p.out_e[a] + v[3];
    if (p.out_w) p.out_w[a] = p.out_w[a] + v[4];
  }
}
HMD_KERNEL(hmd_slot_add, SlotAddParams, 256)

// Scaled exclusions: one thread per listed pair with a non-zero scale, the
// same per-pair arithmetic as the fused kernel (full minimum image, the
// cutoff mask, LJ 12-6 with the C1 switch, the tabulated real-space Ewald
// term) times the pair's scales, the exclusion-scaled sum of the reference
// (src/forces/jit_composed.rs:939-972 with heddle_jit_exclusion_scale
// :1260-1272).  Fixed-point atomics make the totals independent of order.
__device__ __forceinline__ void hmd_pair_scaled_body(const PairScaledParams& p,
                                                     const uint32_t hmd_bx,
                                                     const uint32_t hmd_gx) {
#pragma clang fp contract(off)   // fused multiply-adds written out, as in the fused kernel
  const uint32_t k = hmd_bx * blockDim.x + threadIdx.x;
  if (k >= p.n_pairs) return;
  const uint32_t i = p.pair_i[k], j = p.pair_j[k];
  const float4 a = reinterpret_cast<const float4*>(p.posq)[i];
  const float4 b = reinterpret_cast<const float4*>(p.posq)[j];
  float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  hmd_min_image(dx, dy, dz, p.L);
  const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
  if (!(r2 <= p.rc2) || r2 == 0.0f) return;
  const float inv_r2 = __builtin_amdgcn_rcpf(r2);
  float factor = 0.0f, energy = 0.0f;
  if (p.coul_on && p.s_coul[k] != 0.0f) {
    const float qq = p.k_coulomb * a.w * b.w;
    float t;
    const uint32_t bin = hmd_coul_bin(r2, p.coul_u_lo, p.coul_u_hi, t);
    const float4 g = reinterpret_cast<const float4*>(p.coul_table)[bin];
    factor += p.s_coul[k] * (qq * fmaf(fmaf(fmaf(g.w, t, g.z), t, g.y), t, g.x));
    if (p.fev) {
      const float4 e = reinterpret_cast<const float4*>(p.coul_table)[HMD_COUL_BINS + bin];
      energy += p.s_coul[k] * (qq * fmaf(fmaf(fmaf(e.w, t, e.z), t, e.y), t, e.x));
    }
  }
  if (p.lj_on && p.s_lj[k] != 0.0f) {
    const float2 cc =
        reinterpret_cast<const float2*>(p.lj_c6c12)[p.type[i] * p.n_types + p.type[j]];
    const float ir6 = inv_r2 * inv_r2 * inv_r2;
    const float e_lj = fmaf(cc.y, ir6, -cc.x) * ir6;
    const float f_lj = fmaf(12.0f * cc.y, ir6, -6.0f * cc.x) * ir6 * inv_r2;
    const float tau = fmaxf((r2 - p.lj_rs2) * p.lj_inv_delta, 0.0f);
    const float omt = 1.0f - tau;
    const float sw = omt * omt * fmaf(2.0f, tau, 1.0f);
    const float chain = (12.0f * p.lj_inv_delta * tau) * omt;
    factor += p.s_lj[k] * fmaf(sw, f_lj, chain * e_lj);
    if (p.fev) energy += p.s_lj[k] * (sw * e_lj);
  }
#ifdef HMD_USER_FUNCTOR
  // the composed term's own exclusion_scale(i, j) (src/forces/
  // jit_composed.rs:939-972): s_user x the term for each listed pair
  if (p.s_user && p.s_user[k] != 0.0f) {
    float f_u = 0.0f, e_u = 0.0f;
    HMD_USER_FUNCTOR{}(p.user_params, p.type[i], p.type[j], r2, __builtin_amdgcn_rsqf(r2), f_u,
                       e_u);
    factor += p.s_user[k] * f_u;
    if (p.fev) energy += p.s_user[k] * e_u;
  }
#endif
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
