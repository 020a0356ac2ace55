// SPME charge spread and force gather for gfx950, deterministic
// owner-computes mesh accumulation (no global atomics on the mesh).
//
// Reference behaviour: kernels/spme_spread_gather.cuh (Cox-de Boor weights
// :36-86, per-particle setup :90-126, spread :162-221, gather :227-334) and
// kernels/spme_recip.cu:208-253.  The reference spreads with p threads per
// atom and one global i64 atomic per mesh contribution.  Here:
//
//  * every step the atoms are binned by the mesh brick (b0 x b1 x b2 points)
//    that holds their primary point g = floor(u);
//  * one workgroup owns one brick: it walks the atoms of its own bin and of
//    the +1 neighbour bins along each axis (the only atoms whose p^3 support
//    reaches the brick, since b >= p), accumulates their contributions for
//    the points it owns in LDS as i64 fixed point (2^32, round to nearest,
//    as spme_spread_gather.cuh:211-212), and writes the brick once as f32.
//    Integer LDS sums make the result independent of atom order;
//  * the gather stages the brick's potential plus its (p-1) halo in LDS and
//    evaluates each atom of the brick's own bin in one thread.
#include "hmd_device.h"

template <int P>
__device__ inline void hmd_bspline(float x, float w[P]) {
  // M_P(x - 0 .. x - (P-1)) by Cox-de Boor from the M_1 indicators; returns
  // w[d] = M_P(d + t) for x = t, d = 0..P-1, evaluated as in
  // spme_spread_gather.cuh:36-54 (one recursion per support point).
#pragma unroll
  for (int d = 0; d < P; ++d) {
    float xd = x + (float)d;
    float v[P];
#pragma unroll
    for (int i = 0; i < P; ++i) {
      float xi = xd - (float)i;
      v[i] = (xi >= 0.0f && xi < 1.0f) ? 1.0f : 0.0f;
    }
#pragma unroll
    for (int order = 2; order <= P; ++order) {
      float inv = 1.0f / (float)(order - 1);
#pragma unroll
      for (int i = 0; i < P - order + 1; ++i) {
        float xi = xd - (float)i;
        v[i] = xi * inv * v[i] + ((float)order - xi) * inv * v[i + 1];
      }
    }
    w[d] = v[0];
  }
}

template <int P>
__device__ inline void hmd_bspline_deriv(float x, float w[P], float dw[P]) {
  // M_P(d + t) and M_P'(d + t) = M_{P-1}(d + t) - M_{P-1}(d + t - 1)
  // (spme_spread_gather.cuh:58-86).
#pragma unroll
  for (int d = 0; d < P; ++d) {
    float xd = x + (float)d;
    float v[P + 1];
#pragma unroll
    for (int i = 0; i < P + 1; ++i) {
      float xi = xd - (float)i;
      v[i] = (xi >= 0.0f && xi < 1.0f) ? 1.0f : 0.0f;
    }
#pragma unroll
    for (int order = 2; order < P; ++order) {
      float inv = 1.0f / (float)(order - 1);
#pragma unroll
      for (int i = 0; i < P - order + 1; ++i) {
        float xi = xd - (float)i;
        v[i] = xi * inv * v[i] + ((float)order - xi) * inv * v[i + 1];
      }
    }
    dw[d] = v[0] - v[1];
    float inv = 1.0f / (float)(P - 1);
    w[d] = xd * inv * v[0] + ((float)P - xd) * inv * v[1];
  }
}

// Order 4 in closed form (the cubic cardinal B-spline on its four unit
// intervals; the same function as the recursion, fewer operations):
// w[d] = M_4(d + t), dw[d] = M_4'(d + t).
template <>
__device__ inline void hmd_bspline<4>(float t, float w[4]) {
  const float t2 = t * t, t3 = t2 * t, u = 1.0f - t;
  const float s6 = 1.0f / 6.0f;
  w[0] = t3 * s6;
  w[1] = (1.0f + 3.0f * t + 3.0f * t2 - 3.0f * t3) * s6;
  w[2] = (4.0f - 6.0f * t2 + 3.0f * t3) * s6;
  w[3] = u * u * u * s6;
}

template <>
__device__ inline void hmd_bspline_deriv<4>(float t, float w[4], float dw[4]) {
  hmd_bspline<4>(t, w);
  const float t2 = t * t, u = 1.0f - t;
  dw[0] = 0.5f * t2;
  dw[1] = 0.5f * (1.0f + 2.0f * t - 3.0f * t2);
  dw[2] = 0.5f * (3.0f * t2 - 4.0f * t);
  dw[3] = -0.5f * u * u;
}

// u = (s + 1/2) n, primary point g = floor(u) mod n, t = u - floor(u)
// (spme_spread_gather.cuh:90-126).
__device__ inline void hmd_spme_setup(float4 a, const HmdLattice& L,
                                      uint32_t n0, uint32_t n1, uint32_t n2,
                                      float t[3], int g[3]) {
  float x = a.x, y = a.y, z = a.z;
  int ia, ib, ic;
  hmd_wrap(x, y, z, ia, ib, ic, L);
  float s[3];
  hmd_cart_to_frac(x, y, z, L, s[0], s[1], s[2]);
  const uint32_t nn[3] = {n0, n1, n2};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float u = (s[d] + 0.5f) * (float)nn[d];
    float f = floorf(u);
    t[d] = u - f;
    int gi = (int)f % (int)nn[d];
    if (gi < 0) gi += (int)nn[d];
    g[d] = gi;
  }
}

// Bin each atom by the brick holding its primary point.  Atoms are visited
// in the neighbour list's Hilbert order, so a wave's 64 atoms fall into a few
// bins: one atomic per (wave, bin) instead of one per atom.  The primary
// point is computed from the same f32 position the spread and the gather
// read (posq, not the list's image-shifted copy): an atom within a rounding
// of a box face can wrap to the other face from a shifted copy, and the
// gather would then look for its stencil outside the brick it is binned in.
__device__ inline uint32_t hmd_spme_bin_of(const SpmeBinParams& p, float4 pq) {
  float t[3];
  int g[3];
  hmd_spme_setup(pq, p.L, p.n0, p.n1, p.n2, t, g);
  return ((uint32_t)g[0] / p.b0 * p.nb1 + (uint32_t)g[1] / p.b1) * p.nb2 +
         (uint32_t)g[2] / p.b2;
}

__device__ __forceinline__ void hmd_spme_bin_body(const SpmeBinParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  const uint32_t s = hmd_bx * blockDim.x + threadIdx.x;
  const bool valid = s < p.n;
  uint32_t bin = HMD_SENTINEL;
  if (valid) {
    bin = hmd_spme_bin_of(p, reinterpret_cast<const float4*>(p.posq)[p.sorted_id[s]]);
    p.atom_bin[s] = bin;
  }
  uint64_t pending = __ballot(valid);
  while (pending) {
    const uint32_t leader = (uint32_t)__builtin_ctzll(pending);
    const uint32_t b = hmd_readlane(bin, leader);
    const uint64_t same = __ballot(valid && bin == b);
    if (hmd_lane() == leader) atomicAdd(&p.bin_count[b], (uint32_t)__builtin_popcountll(same));
    pending &= ~same;
  }
}
HMD_KERNEL(hmd_spme_bin, SpmeBinParams, 256)

// Scatter canonical atom ids into their bins.  The count array is driven
// back to zero (one atomicSub per (wave, bin) claims a block of positions).
__device__ __forceinline__ void hmd_spme_bin_scatter_body(const SpmeBinParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  const uint32_t s = hmd_bx * blockDim.x + threadIdx.x;
  const bool valid = s < p.n;
  const uint32_t lane = hmd_lane();
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint32_t bin = valid ? p.atom_bin[s] : HMD_SENTINEL;
  uint64_t pending = __ballot(valid);
  while (pending) {
    const uint32_t leader = (uint32_t)__builtin_ctzll(pending);
    const uint32_t b = hmd_readlane(bin, leader);
    const uint64_t same = __ballot(valid && bin == b);
    const uint32_t cnt = (uint32_t)__builtin_popcountll(same);
    uint32_t old = 0;
    if (lane == leader) old = atomicSub(&p.bin_count[b], cnt);
    old = hmd_readlane(old, leader);
    if (valid && bin == b) {
      const uint32_t rank = (uint32_t)__builtin_popcountll(same & lt_mask);
      p.bin_atoms[p.bin_start[b] + old - cnt + rank] = p.sorted_id[s];
    }
    pending &= ~same;
  }
}
HMD_KERNEL(hmd_spme_bin_scatter, SpmeBinParams, 256)

// A contribution v (in units of 2^-24 e, |v| < 2^31) as the i64 count of
// 2^-32 e rounded to nearest, from two i32 roundings: r1 = rint(v), the
// remainder v - r1 (exact in f32, |.| <= 1/2) scaled by 2^8 and rounded,
// and r1 * 2^8 + r2 -- the same value as rounding v * 2^8 directly, without
// the long f32 -> i64 conversion sequence.
__device__ inline long long hmd_fixed32_of24(float v) {
  const float f1 = rintf(v);
  const int r1 = (int)f1;
  const int r2 = (int)rintf((v - f1) * 256.0f);
  return ((long long)r1 << 8) + (long long)r2;
}

template <int P>
__device__ inline void hmd_spread_body(const SpmeSpreadParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  extern __shared__ long long s_acc[];
  const uint32_t bsz = p.b0 * p.b1 * p.b2;
  const uint32_t B = hmd_bx;
  const uint32_t B2 = B % p.nb2, B1 = (B / p.nb2) % p.nb1, B0 = B / (p.nb2 * p.nb1);
  const int o0 = (int)(B0 * p.b0), o1 = (int)(B1 * p.b1), o2 = (int)(B2 * p.b2);
  for (uint32_t k = threadIdx.x; k < bsz; k += blockDim.x) s_acc[k] = 0;
  __syncthreads();

  for (uint32_t nb = 0; nb < 8; ++nb) {
    // With a single brick along an axis the +1 neighbour is the brick
    // itself: visit it once.
    if ((p.nb0 == 1 && (nb & 4u)) || (p.nb1 == 1 && (nb & 2u)) ||
        (p.nb2 == 1 && (nb & 1u)))
      continue;
    const uint32_t c0 = (B0 + ((nb >> 2) & 1u)) % p.nb0;
    const uint32_t c1 = (B1 + ((nb >> 1) & 1u)) % p.nb1;
    const uint32_t c2 = (B2 + (nb & 1u)) % p.nb2;
    const uint32_t bin = (c0 * p.nb1 + c1) * p.nb2 + c2;
    const uint32_t beg = p.bin_start[bin], end = p.bin_start[bin + 1];
    for (uint32_t k = beg + threadIdx.x; k < end; k += blockDim.x) {
      const uint32_t a = p.bin_atoms[k];
      const float4 pq = reinterpret_cast<const float4*>(p.posq)[a];
      if (pq.w == 0.0f) continue;
      float t[3];
      int g[3];
      hmd_spme_setup(pq, p.L, p.n0, p.n1, p.n2, t, g);
      // Which of the P support planes along each axis fall in this brick;
      // an atom of a neighbour bin usually reaches it along no plane of
      // some axis, and is then skipped before its weights are computed.
      uint32_t ma = 0, mb = 0, mc = 0;
#pragma unroll
      for (int d = 0; d < P; ++d) {
        int la = g[0] - d - o0, lb = g[1] - d - o1, lc = g[2] - d - o2;
        la += (la < 0) ? (int)p.n0 : 0;
        lb += (lb < 0) ? (int)p.n1 : 0;
        lc += (lc < 0) ? (int)p.n2 : 0;
        ma |= (la < (int)p.b0) ? (1u << d) : 0u;
        mb |= (lb < (int)p.b1) ? (1u << d) : 0u;
        mc |= (lc < (int)p.b2) ? (1u << d) : 0u;
      }
      if (ma == 0u || mb == 0u || mc == 0u) continue;
      float wa[P], wb[P], wc[P];
      hmd_bspline<P>(t[0], wa);
      hmd_bspline<P>(t[1], wb);
      hmd_bspline<P>(t[2], wc);
      int lc_of[P];
#pragma unroll
      for (int dc = 0; dc < P; ++dc) {
        int lc = g[2] - dc - o2;
        lc_of[dc] = lc + ((lc < 0) ? (int)p.n2 : 0);
      }
      const float q24 = pq.w * 16777216.0f;   // charge in 2^-24 units
#pragma unroll
      for (int da = 0; da < P; ++da) {
        if (!((ma >> da) & 1u)) continue;
        int la = g[0] - da - o0;
        la += (la < 0) ? (int)p.n0 : 0;
        const float qa = q24 * wa[da];
#pragma unroll
        for (int db = 0; db < P; ++db) {
          if (!((mb >> db) & 1u)) continue;
          int lb = g[1] - db - o1;
          lb += (lb < 0) ? (int)p.n1 : 0;
          const float qab = qa * wb[db];
          long long* row = s_acc + ((uint32_t)la * p.b1 + (uint32_t)lb) * p.b2;
#pragma unroll
          for (int dc = 0; dc < P; ++dc) {
            if (!((mc >> dc) & 1u)) continue;
            atomicAdd(reinterpret_cast<unsigned long long*>(&row[lc_of[dc]]),
                      (unsigned long long)hmd_fixed32_of24(qab * wc[dc]));
          }
        }
      }
    }
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < bsz; k += blockDim.x) {
    const uint32_t lc = k % p.b2, lb = (k / p.b2) % p.b1, la = k / (p.b2 * p.b1);
    const size_t gi = ((size_t)(o0 + la) * p.n1 + (o1 + lb)) * p.n2 + (o2 + lc);
    p.rho[gi] = hmd_from_fixed(s_acc[k], HMD_SPREAD_SHIFT);
  }
}

template <int P>
__device__ inline void hmd_gather_body(const SpmeGatherParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  extern __shared__ float s_v[];
  const uint32_t h0 = p.b0 + P - 1, h1 = p.b1 + P - 1, h2 = p.b2 + P - 1;
  const uint32_t B = hmd_bx;
  const uint32_t B2 = B % p.nb2, B1 = (B / p.nb2) % p.nb1, B0 = B / (p.nb2 * p.nb1);
  const int o0 = (int)(B0 * p.b0) - (P - 1);
  const int o1 = (int)(B1 * p.b1) - (P - 1);
  const int o2 = (int)(B2 * p.b2) - (P - 1);
  for (uint32_t k = threadIdx.x; k < h0 * h1 * h2; k += blockDim.x) {
    const uint32_t lc = k % h2, lb = (k / h2) % h1, la = k / (h2 * h1);
    int ga = o0 + (int)la, gb = o1 + (int)lb, gc = o2 + (int)lc;
    ga += ga < 0 ? (int)p.n0 : 0;
    gb += gb < 0 ? (int)p.n1 : 0;
    gc += gc < 0 ? (int)p.n2 : 0;
    s_v[k] = p.V[((size_t)ga * p.n1 + gb) * p.n2 + gc];
  }
  __syncthreads();

  const HmdLattice L = p.L;
  // Rows of H^{-T} for the lower-triangular cell
  // (spme_spread_gather.cuh:128-148).
  const float ilx = 1.0f / L.lx, ily = 1.0f / L.ly, ilz = 1.0f / L.lz;
  const float ba_x = ilx, ba_y = -L.xy * ilx * ily,
              ba_z = (L.xy * L.yz - L.xz * L.ly) * ilx * ily * ilz;
  const float bb_y = ily, bb_z = -L.yz * ily * ilz;
  const float bc_z = ilz;
  const float w_share = (float)(p.w_recip[0] / (double)p.n);

  const uint32_t beg = p.bin_start[B], end = p.bin_start[B + 1];
  for (uint32_t k = beg + threadIdx.x; k < end; k += blockDim.x) {
    const uint32_t a = p.bin_atoms[k];
    const float4 pq = reinterpret_cast<const float4*>(p.posq)[a];
    float t[3];
    int g[3];
    hmd_spme_setup(pq, L, p.n0, p.n1, p.n2, t, g);
    float wa[P], wb[P], wc[P], dwa[P], dwb[P], dwc[P];
    hmd_bspline_deriv<P>(t[0], wa, dwa);
    hmd_bspline_deriv<P>(t[1], wb, dwb);
    hmd_bspline_deriv<P>(t[2], wc, dwc);
    // local halo coordinates of the primary point
    const int la0 = g[0] - o0, lb0 = g[1] - o1, lc0 = g[2] - o2;
    float phi = 0.0f, gra = 0.0f, grb = 0.0f, grc = 0.0f;
#pragma unroll
    for (int da = 0; da < P; ++da) {
#pragma unroll
      for (int db = 0; db < P; ++db) {
        const float* row = &s_v[((uint32_t)(la0 - da) * h1 + (uint32_t)(lb0 - db)) * h2];
        float sv = 0.0f, sdc = 0.0f;
#pragma unroll
        for (int dc = 0; dc < P; ++dc) {
          const float v = row[lc0 - dc];
          sv += v * wc[dc];
          sdc += v * dwc[dc];
        }
        phi += wa[da] * wb[db] * sv;
        gra += dwa[da] * wb[db] * sv;
        grb += wa[da] * dwb[db] * sv;
        grc += wa[da] * wb[db] * sdc;
      }
    }
    const float ga = (float)p.n0 * gra, gb = (float)p.n1 * grb, gc = (float)p.n2 * grc;
    p.f[a] = -pq.w * (ga * ba_x);
    p.f[p.n + a] = -pq.w * (ga * ba_y + gb * bb_y);
    p.f[2 * p.n + a] = -pq.w * (ga * ba_z + gb * bb_z + gc * bc_z);
    if (p.fev) {
      p.e[a] = 0.5f * pq.w * phi - p.u_self[a];
      p.w[a] = w_share;
    }
  }
}

#define HMD_SPME_ORDER(P)                                                    \
  __device__ __forceinline__ void hmd_spme_spread_p##P##_body(              \
      const SpmeSpreadParams& p, const uint32_t hmd_bx, const uint32_t hmd_gx) { \
    hmd_spread_body<P>(p, hmd_bx, hmd_gx);                                   \
  }                                                                          \
  __device__ __forceinline__ void hmd_spme_gather_p##P##_body(              \
      const SpmeGatherParams& p, const uint32_t hmd_bx, const uint32_t hmd_gx) { \
    hmd_gather_body<P>(p, hmd_bx, hmd_gx);                                   \
  }                                                                          \
  HMD_KERNEL(hmd_spme_spread_p##P, SpmeSpreadParams, 256)                    \
  HMD_KERNEL(hmd_spme_gather_p##P, SpmeGatherParams, 256)

HMD_SPME_ORDER(4)
HMD_SPME_ORDER(5)
HMD_SPME_ORDER(6)
HMD_SPME_ORDER(8)
