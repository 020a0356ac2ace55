// Cell list + packed cluster-pair neighbour list for gfx950.
//
// Reference behaviour (taylor-a-barnes/heddlemd): kernels/neighbor.cu and
// src/forces/neighbor_list.rs.  The list holds every unordered pair whose
// exact f32 minimum-image squared distance is <= (r_c + r_skin)^2
// (kernels/neighbor.cu:536-553; hmd_nbl_pairs decodes exactly that set), plus
// some farther pairs that the force kernels mask by their cutoff tests.  The
// construction is wave64-native:
//
//  * atoms are sorted along a Hilbert curve over a power-of-two grid of
//    cells (about 0.2-0.4 nm), ties broken by atom id, so 64 consecutive
//    sorted atoms form a spatially compact cluster and one wave owns one
//    cluster (one atom per lane);
//  * candidate j-clusters come from a bounding-box prune of clusters with a
//    higher index (half list, Newton's third law in the force kernel);
//  * j-atoms are filtered per atom against the i-cluster's eight 8-atom
//    sub-boxes and compacted in increasing sorted order with a wave ballot +
//    prefix popcount, so each cluster's list -- and every entry of 64
//    j-slots -- is deterministic;
//  * each entry carries, per i-lane, a 64-bit mask of excluded j-slots built
//    from the exclusion CSR, so the force kernel needs no per-pair search.
//
// Every rebuild kernel is captured into the per-step graph and gated on the
// device status word, so a displacement-triggered rebuild happens in the step
// that needs it without a host round trip.
#include "hmd_device.h"

// Spread 10 bits of x so that consecutive bits are 3 apart (Morton code).
__device__ inline uint32_t hmd_spread3(uint32_t x) {
  x &= 0x3ffu;
  x = (x | (x << 16)) & 0x030000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

// Position of cell (x, y, z) along the 3D Hilbert curve of a 2^bits grid:
// the coordinates are transformed into the curve's "transpose" form
// (Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707, 2004)
// and their bits interleaved.  Unlike the Morton (Z) order, consecutive
// cells are always face neighbours, so runs of 64 atoms form compact
// clusters with no long jumps.
__device__ inline uint32_t hmd_hilbert3(uint32_t x, uint32_t y, uint32_t z, uint32_t bits) {
  uint32_t X[3] = {x, y, z};
  const uint32_t M = 1u << (bits - 1u);
  for (uint32_t Q = M; Q > 1u; Q >>= 1) {
    const uint32_t P = Q - 1u;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        const uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  X[1] ^= X[0];
  X[2] ^= X[1];
  uint32_t t = 0;
  for (uint32_t Q = M; Q > 1u; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1u;
  X[0] ^= t;
  X[1] ^= t;
  X[2] ^= t;
  return (hmd_spread3(X[0]) << 2) | (hmd_spread3(X[1]) << 1) | hmd_spread3(X[2]);
}

// Fine cell of a (wrapped) position, clamped like
// parallelepiped_cell_indices (kernels/neighbor.cu:11-36).
__device__ inline void hmd_fine_cell(float x, float y, float z,
                                     const HmdLattice& L, uint32_t g0,
                                     uint32_t g1, uint32_t g2, uint32_t& c0,
                                     uint32_t& c1, uint32_t& c2) {
  int ia, ib, ic;
  hmd_wrap(x, y, z, ia, ib, ic, L);
  float sa, sb, sc;
  hmd_cart_to_frac(x, y, z, L, sa, sb, sc);
  int a = (int)floorf((sa + 0.5f) * (float)g0);
  int b = (int)floorf((sb + 0.5f) * (float)g1);
  int c = (int)floorf((sc + 0.5f) * (float)g2);
  a = a < 0 ? 0 : (a >= (int)g0 ? (int)g0 - 1 : a);
  b = b < 0 ? 0 : (b >= (int)g1 ? (int)g1 - 1 : b);
  c = c < 0 ? 0 : (c >= (int)g2 ? (int)g2 - 1 : c);
  c0 = (uint32_t)a;
  c1 = (uint32_t)b;
  c2 = (uint32_t)c;
}

// Skin trigger (kernels/neighbor.cu:50-76): any atom whose min-image
// displacement from its rebuild-time position exceeds r_skin/2 sets bit 0.
// The squared displacement is the exact f32 min-image distance (no
// contraction), so the trigger is one specified function of the positions.
__device__ __forceinline__ void hmd_nbl_displacement_body(const NblDispParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
#pragma clang fp contract(off)
  uint32_t i = hmd_bx * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  float4 a = reinterpret_cast<const float4*>(p.posq)[i];
  float4 r = reinterpret_cast<const float4*>(p.ref_pos)[i];
  float dx = a.x - r.x, dy = a.y - r.y, dz = a.z - r.z;
  hmd_min_image(dx, dy, dz, p.L);
  float d2 = dx * dx + dy * dy + dz * dz;
  if (d2 > p.threshold_sq) atomicOr(p.status, HMD_STATUS_DISPLACED);
}
HMD_KERNEL(hmd_nbl_displacement, NblDispParams, 256)

// Fine-cell Hilbert key + histogram.
__device__ __forceinline__ void hmd_nbl_keys_body(const NblKeyParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  if (!hmd_gate(p.status, HMD_STATUS_DISPLACED)) return;
  uint32_t i = hmd_bx * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  float4 a = reinterpret_cast<const float4*>(p.posq)[i];
  uint32_t c0, c1, c2;
  hmd_fine_cell(a.x, a.y, a.z, p.L, p.g0, p.g1, p.g2, c0, c1, c2);
  const uint32_t key = hmd_hilbert3(c0, c1, c2, p.bits);
  p.atom_key[i] = key;
  atomicAdd(&p.cell_count[key], 1u);
}
HMD_KERNEL(hmd_nbl_keys, NblKeyParams, 256)

// Scatter into cells.  The count array doubles as the cursor and is driven
// back to zero, ready for the next rebuild.  Order inside a cell is fixed
// afterwards by hmd_nbl_sort_cells.
__device__ __forceinline__ void hmd_nbl_scatter_body(const NblKeyParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  if (!hmd_gate(p.status, HMD_STATUS_DISPLACED)) return;
  uint32_t i = hmd_bx * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  uint32_t key = p.atom_key[i];
  uint32_t k = atomicSub(&p.cell_count[key], 1u) - 1u;
  p.sorted_id[p.cell_start[key] + k] = i;
}
HMD_KERNEL(hmd_nbl_scatter, NblKeyParams, 256)

// Canonical (cell, atom id) order: insertion sort inside each fine cell
// (a cell holds a couple of atoms; compare kernels/neighbor.cu:218-238).
__device__ __forceinline__ void hmd_nbl_sort_cells_body(const NblSortParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  if (!hmd_gate(p.status, HMD_STATUS_DISPLACED)) return;
  uint32_t c = hmd_bx * blockDim.x + threadIdx.x;
  if (c >= p.n_keys) return;
  uint32_t s = p.cell_start[c], e = p.cell_start[c + 1];
  for (uint32_t k = s + 1; k < e; ++k) {
    uint32_t key = p.sorted_id[k];
    uint32_t pos = k;
    while (pos > s && p.sorted_id[pos - 1] > key) {
      p.sorted_id[pos] = p.sorted_id[pos - 1];
      --pos;
    }
    p.sorted_id[pos] = key;
  }
}
HMD_KERNEL(hmd_nbl_sort_cells, NblSortParams, 256)

// Sorted views.  Every step it refreshes the sorted positions
// (kernels/neighbor.cu:251-285 does the same refresh every step in
// ForceField::run); in a rebuilding step it also writes the inverse map, the
// sorted types, the rebuild-time reference positions and the padding lanes.
// The sorted positions carry the lattice images that the rebuild's cell
// assignment removed (hmd_wrap of the rebuild-time position, kept per slot
// and applied every step with the same operations): at the rebuild they are
// the wrapped positions the cells were taken from, so every cluster's box is
// compact even for atoms that a constraint step left just outside the primary
// cell, and between rebuilds they move continuously.
__device__ __forceinline__ void hmd_nbl_gather_body(const NblGatherParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
#pragma clang fp contract(off)
  const bool rebuild = hmd_gate(p.status, HMD_STATUS_DISPLACED);
  uint32_t s = hmd_bx * blockDim.x + threadIdx.x;
  if (s >= p.n_pad) return;
  float4* ps = reinterpret_cast<float4*>(p.posq_sorted);
  if (s >= p.n) {
    if (rebuild) {
      // Padding lanes: far away and uncharged; their slots never pass the
      // validity test of the force kernel.
      ps[s] = make_float4(1.0e18f, 1.0e18f, 1.0e18f, 0.0f);
      p.type_sorted[s] = 0u;
    }
    return;
  }
  uint32_t id = p.sorted_id[s];
  const float4 a = reinterpret_cast<const float4*>(p.posq)[id];
  int4* img = reinterpret_cast<int4*>(p.image);
  int4 im;
  if (rebuild) {
    float x = a.x, y = a.y, z = a.z;
    hmd_wrap(x, y, z, im.x, im.y, im.z, p.L);
    im.w = 0;
    img[s] = im;
    p.slot_of[id] = s;
    p.type_sorted[s] = p.type[id];
    reinterpret_cast<float4*>(p.ref_pos)[id] = a;
  } else {
    im = img[s];
  }
  const float ka = (float)im.x, kb = (float)im.y, kc = (float)im.z;
  float4 w = a;
  w.x -= ka * p.L.lx + kb * p.L.xy + kc * p.L.xz;
  w.y -= kb * p.L.ly + kc * p.L.yz;
  w.z -= kc * p.L.lz;
  ps[s] = w;
}
HMD_KERNEL(hmd_nbl_gather, NblGatherParams, 256)

// Per-cluster axis-aligned bounding box (centre and half extents) of the
// valid atoms, one wave per cluster (compare compute_block_bbox,
// kernels/neighbor.cu:291-361), plus the boxes of the cluster's eight
// 8-atom sub-clusters (consecutive sorted atoms) against which the list build
// tests each j-atom.
__device__ __forceinline__ void hmd_nbl_bbox_body(const NblBboxParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  if (!hmd_gate(p.status, HMD_STATUS_DISPLACED)) return;
  uint32_t c = hmd_bx * (blockDim.x / HMD_WAVE) + threadIdx.x / HMD_WAVE;
  if (c >= p.n_clusters) return;
  uint32_t lane = hmd_lane();
  uint32_t s = c * HMD_CLUSTER + lane;
  bool valid = s < p.n;
  float4 a = reinterpret_cast<const float4*>(p.posq_sorted)[s];
  float lo[3] = {valid ? a.x : 3.0e38f, valid ? a.y : 3.0e38f, valid ? a.z : 3.0e38f};
  float hi[3] = {valid ? a.x : -3.0e38f, valid ? a.y : -3.0e38f, valid ? a.z : -3.0e38f};
  float4* sub = reinterpret_cast<float4*>(p.sub_bbox);
#pragma unroll
  for (int off = 1; off <= 32; off <<= 1) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = fminf(lo[d], hmd_shfl_xor(lo[d], off));
      hi[d] = fmaxf(hi[d], hmd_shfl_xor(hi[d], off));
    }
    if (off == 4 && (lane & 7u) == 0u) {   // 8-lane groups reduced
      const uint32_t k = c * 8u + lane / 8u;
      sub[2 * k] = make_float4(0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]),
                               0.5f * (lo[2] + hi[2]), 0.0f);
      sub[2 * k + 1] = make_float4(0.5f * (hi[0] - lo[0]), 0.5f * (hi[1] - lo[1]),
                                   0.5f * (hi[2] - lo[2]), 0.0f);
    }
  }
  if (lane == 0) {
    float4* b = reinterpret_cast<float4*>(p.bbox);
    b[2 * c] = make_float4(0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]),
                           0.5f * (lo[2] + hi[2]), 0.0f);
    b[2 * c + 1] = make_float4(0.5f * (hi[0] - lo[0]), 0.5f * (hi[1] - lo[1]),
                               0.5f * (hi[2] - lo[2]), 0.0f);
  }
  // The atom's exclusion partners as sorted slots (a chain of dependent
  // loads: id, CSR row, partner, its slot), resolved here for every atom at
  // once so that the list build reads them with one load.
  uint32_t b0 = 0, b1 = 0;
  if (valid) {
    const uint32_t id = p.sorted_id[s];
    b0 = p.excl_offsets[id];
    b1 = p.excl_offsets[id + 1];
  }
#pragma unroll
  for (uint32_t m = 0; m < 8; ++m) {
    uint32_t slot = HMD_SENTINEL;
    if (m < b1 - b0) slot = p.slot_of[p.excl_partners[b0 + m]];
    p.excl_slot[(size_t)m * p.n_pad + s] = slot;
  }
}
HMD_KERNEL(hmd_nbl_bbox, NblBboxParams, 256)

// Conservative min-image gap between two boxes (compare
// packed_block_bbox_dist_sq, kernels/neighbor.cu:365-388).  Used only as a
// prune against r_search^2 with a relative margin, so the image may be
// chosen with a reciprocal multiply.
__device__ inline float hmd_box_gap2(float4 ca, float4 ha, float4 cb, float4 hb,
                                     const HmdLattice& L, bool ortho, float ilx,
                                     float ily, float ilz) {
  float dx = ca.x - cb.x, dy = ca.y - cb.y, dz = ca.z - cb.z;
  if (ortho) {
    dx -= L.lx * rintf(dx * ilx);
    dy -= L.ly * rintf(dy * ily);
    dz -= L.lz * rintf(dz * ilz);
  } else {
    hmd_min_image(dx, dy, dz, L);
  }
  float ex = fmaxf(fabsf(dx) - (ha.x + hb.x), 0.0f);
  float ey = fmaxf(fabsf(dy) - (ha.y + hb.y), 0.0f);
  float ez = fmaxf(fabsf(dz) - (ha.z + hb.z), 0.0f);
  return ex * ex + ey * ey + ez * ez;
}

// Build the packed list of one i-cluster per workgroup of four waves.
//
// Candidate j-clusters are those with index >= c whose boxes lie within the
// search radius (one ballot per 64 j-clusters, the four waves taking turns).
// After a prefix over the per-chunk masks the candidates are listed in
// increasing index order in LDS, and the four waves share them round robin:
// lane k tests j-atom k of the candidate against the i-cluster's eight 8-atom
// sub-boxes (consecutive sorted atoms, so sub-box s holds i-atoms 8s..8s+7)
// and keeps it when a sub-box holding an i-atom with a lower sorted index lies
// within r_search of it; the next candidate's j-atoms are loaded while the
// current one is tested.  Each kept j-atom carries its 8-bit mask of reachable
// sub-boxes.
//
// Placement.  The force kernel walks an entry's 64 j-atoms (one per lane)
// over the i-atoms of the sub-boxes that any of them reaches, so entries whose
// j-atoms share their masks cost least.  A cluster's list is therefore ordered
// by a key of the mask: 0 for j-atoms of the cluster itself (first, with most
// of the exclusion partners), else 1 + lowest + highest reachable sub-box (a
// mask is close to a contiguous run along the Hilbert curve, and the key
// groups similar runs), and by candidate then lane within a key: a stable
// counting sort from per-(candidate, key) counts and ranks taken with ballots
// in the hit test.  A cluster with more candidates than one round holds
// keeps plain candidate order.  Either way the list depends only on the
// positions (deterministic).  All waves place the kept atoms in an LDS copy of
// the list (positions past its lds_slots go straight to the entry array),
// written out as whole entries, together with each entry's mask (the union of
// its j-atoms' masks).  Each lane of wave 0 keeps the sorted slots of its
// atom's exclusion partners (at most HMD_MAX_INLINE_EXCL) in registers and
// finds their list positions, from which the per-(entry, i-lane) exclusion
// bit masks are written.
#define HMD_MAX_INLINE_EXCL 8
#define HMD_NBL_CAND_CAP 128u       // candidates per round (one round: key-sorted list)
#define HMD_NBL_KEYS 16u            // placement keys: own cluster, 1 + lo + hi
#define HMD_NBL_EMASK_CAP 256u      // entries whose masks are kept (later ones: all 8)
#ifndef HMD_NBL_PREFETCH
#define HMD_NBL_PREFETCH 3       // candidates' j-atoms loaded ahead per wave
#endif

typedef __attribute__((address_space(4))) const float4 hmd_const_f4;

// Inclusive prefix sum over the 64 lanes.
__device__ inline uint32_t hmd_wave_scan_incl(uint32_t v, uint32_t lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = (uint32_t)__shfl_up((int)v, (unsigned)o, HMD_WAVE);
    if (lane >= (uint32_t)o) v += t;
  }
  return v;
}

// Placement key of a kept j-atom (see above).
__device__ inline uint32_t hmd_nbl_key(uint32_t m8, bool own) {
  return own ? 0u : 1u + (uint32_t)__builtin_ctz(m8) + (31u - (uint32_t)__builtin_clz(m8));
}

#ifndef HMD_NBL_BUILD_WAVES_PER_EU
#define HMD_NBL_BUILD_WAVES_PER_EU 3
#endif
#ifndef HMD_NBL_WAVES
#define HMD_NBL_WAVES 4u              // waves per i-cluster workgroup
#endif
#ifndef HMD_NBL_BOX_BATCH
#define HMD_NBL_BOX_BATCH 4u          // candidate chunks loaded together
#endif
__device__ __forceinline__ void hmd_nbl_build_body(const NblBuildParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  if (!hmd_gate(p.status, HMD_STATUS_DISPLACED)) return;
  __shared__ uint32_t s_cand[HMD_NBL_CAND_CAP];
  __shared__ uint64_t s_hit[HMD_NBL_CAND_CAP];
  __shared__ uint32_t s_off[HMD_NBL_CAND_CAP];
  // per kept j-atom: sub-box mask | (rank among the candidate's hits of its key) << 8
  __shared__ uint16_t s_lane[HMD_NBL_CAND_CAP][HMD_ENTRY];
  // per (candidate, key): hit count, then list position of the first such hit
  __shared__ uint16_t s_koff[HMD_NBL_CAND_CAP][HMD_NBL_KEYS];
  __shared__ uint32_t s_emask[HMD_NBL_EMASK_CAP];
  __shared__ uint32_t s_fill, s_nex;
  // dynamic: candidate masks per chunk of 64 j-clusters, their prefix, and
  // the head of the cluster's j list
  extern __shared__ uint64_t s_dyn[];
  uint64_t* s_mask = s_dyn;                                   // [max_chunks]
  uint64_t* s_nowrap = s_dyn + p.max_chunks;                  // [max_chunks]
  uint32_t* s_pre = reinterpret_cast<uint32_t*>(s_dyn + 2 * p.max_chunks);  // [max_chunks + 1]
  uint32_t* s_list = s_pre + ((p.max_chunks + 2) & ~1u);      // [lds_slots]
  const uint32_t c = hmd_bx;
  const uint32_t wave = threadIdx.x / HMD_WAVE;
  const uint32_t lane = hmd_lane();
  const float4* ps = reinterpret_cast<const float4*>(p.posq_sorted);
  const float4* bb = reinterpret_cast<const float4*>(p.bbox);
  const bool ortho = p.orthorhombic_fast != 0;
  const float ilx = 1.0f / p.L.lx, ily = 1.0f / p.L.ly, ilz = 1.0f / p.L.lz;
  const float rs2_prune = p.r_search_sq * 1.0002f;   // conservative prune
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint32_t n_chunks = (p.n_clusters - c + HMD_WAVE - 1) / HMD_WAVE;
  // The i-cluster's sub-boxes are read at wave-uniform addresses in the hit
  // tests: read them through the constant address space (scalar loads into
  // SGPRs) rather than as LDS broadcasts.
  const hmd_const_f4* csub = (const hmd_const_f4*)(reinterpret_cast<const float4*>(p.sub_bbox) + 16 * c);

  for (uint32_t e = threadIdx.x; e < HMD_NBL_EMASK_CAP; e += blockDim.x) s_emask[e] = 0u;
  // wave 0: the sorted slots of each i-atom's exclusion partners (resolved
  // by hmd_nbl_bbox), and their positions in the cluster's list once placed
  const uint32_t si = c * HMD_CLUSTER + lane;
  uint32_t ex_slot[HMD_MAX_INLINE_EXCL], ex_pos[HMD_MAX_INLINE_EXCL];
#pragma unroll
  for (uint32_t m = 0; m < HMD_MAX_INLINE_EXCL; ++m) {
    ex_slot[m] = HMD_SENTINEL;
    ex_pos[m] = HMD_SENTINEL;
  }
  if (wave == 0 && si < p.n) {
#pragma unroll
    for (uint32_t m = 0; m < HMD_MAX_INLINE_EXCL; ++m)
      ex_slot[m] = p.excl_slot[(size_t)m * p.n_pad + si];
  }
  // candidate masks, 64 j-clusters per ballot; each wave loads the boxes of
  // four chunks before testing them, so that the loads' latencies overlap
  {
    const float4 ci = bb[2 * c], hi = bb[2 * c + 1];
    for (uint32_t ch0 = wave; ch0 < n_chunks; ch0 += HMD_NBL_BOX_BATCH * HMD_NBL_WAVES) {
      float4 cj[HMD_NBL_BOX_BATCH], hj[HMD_NBL_BOX_BATCH];
#pragma unroll
      for (uint32_t u = 0; u < HMD_NBL_BOX_BATCH; ++u) {
        const uint32_t jc = c + (ch0 + HMD_NBL_WAVES * u) * HMD_WAVE + lane;
        const bool in = ch0 + HMD_NBL_WAVES * u < n_chunks && jc < p.n_clusters;
        cj[u] = in ? bb[2 * jc] : make_float4(0, 0, 0, 0);
        hj[u] = in ? bb[2 * jc + 1] : make_float4(-1.0f, -1.0f, -1.0f, 0);
      }
#pragma unroll
      for (uint32_t u = 0; u < HMD_NBL_BOX_BATCH; ++u) {
        const uint32_t ch = ch0 + HMD_NBL_WAVES * u;
        if (ch >= n_chunks) break;
        const uint32_t jc = c + ch * HMD_WAVE + lane;
        bool cand = false, nowrap = false;
        if (jc < p.n_clusters) {
          cand = hmd_box_gap2(ci, hi, cj[u], hj[u], p.L, ortho, ilx, ily, ilz) <= rs2_prune;
          // No wrap: every atom pair of the two clusters is closer than 0.49 L
          // along each axis (orthorhombic box), so the minimum image of each
          // difference is the difference itself (image index 0).
          nowrap = ortho && fabsf(cj[u].x - ci.x) + hj[u].x + hi.x < 0.49f * p.L.lx &&
                   fabsf(cj[u].y - ci.y) + hj[u].y + hi.y < 0.49f * p.L.ly &&
                   fabsf(cj[u].z - ci.z) + hj[u].z + hi.z < 0.49f * p.L.lz;
        }
        const uint64_t m = __ballot(cand);
        const uint64_t nw = __ballot(nowrap);
        if (lane == 0) {
          s_mask[ch] = m;
          s_nowrap[ch] = nw;
        }
      }
    }
  }
  __syncthreads();
  if (wave == 0) {
    uint32_t carry = 0;
    for (uint32_t b = 0; b < n_chunks; b += HMD_WAVE) {
      const uint32_t ch = b + lane;
      const uint32_t cnt = ch < n_chunks ? (uint32_t)__builtin_popcountll(s_mask[ch]) : 0u;
      const uint32_t incl = hmd_wave_scan_incl(cnt, lane);
      if (ch < n_chunks) s_pre[ch] = carry + incl - cnt;
      carry += hmd_readlane(incl, HMD_WAVE - 1);
    }
    if (lane == 0) {
      s_pre[n_chunks] = carry;
      s_fill = 0;
    }
  }
  __syncthreads();
  const uint32_t total = s_pre[n_chunks];
  const bool sorted = total <= HMD_NBL_CAND_CAP;   // one round: key-sorted placement

  const size_t base_entry = (size_t)c * p.cap_entries;
  const uint32_t cap_slots = p.cap_entries * HMD_ENTRY;
  const float4 zero = make_float4(0, 0, 0, 0);

  for (uint32_t base = 0; base < total; base += HMD_NBL_CAND_CAP) {
    const uint32_t cnt = min(HMD_NBL_CAND_CAP, total - base);
    for (uint32_t ch = wave; ch < n_chunks; ch += HMD_NBL_WAVES) {
      const uint64_t m = s_mask[ch];
      if ((m >> lane) & 1ull) {
        const uint32_t pos = s_pre[ch] + (uint32_t)__builtin_popcountll(m & lt_mask);
        if (pos >= base && pos < base + cnt)
          s_cand[pos - base] = (c + ch * HMD_WAVE + lane) |
                               ((uint32_t)((s_nowrap[ch] >> lane) & 1ull) << 31);
      }
    }
    __syncthreads();

    // hit masks, candidates shared round robin, next j-atom prefetched
    // (j-atoms of the next HMD_NBL_PREFETCH candidates of this wave in flight)
    uint32_t q = wave;
    float4 pj_pre[HMD_NBL_PREFETCH];
#pragma unroll
    for (uint32_t d = 0; d < HMD_NBL_PREFETCH; ++d) {
      const uint32_t qq = q + HMD_NBL_WAVES * d;
      pj_pre[d] = zero;
      if (qq < cnt) {
        const uint32_t sj = (s_cand[qq] & 0x7FFFFFFFu) * HMD_CLUSTER + lane;
        if (sj < p.n) pj_pre[d] = ps[sj];
      }
    }
    for (; q < cnt; q += HMD_NBL_WAVES) {
      const uint32_t cw = s_cand[q];
      const uint32_t jcl = cw & 0x7FFFFFFFu;
      const bool nowrap = (cw >> 31) != 0u;
      const uint32_t sj = jcl * HMD_CLUSTER + lane;
      const float4 pj = pj_pre[0];
#pragma unroll
      for (uint32_t d = 0; d + 1 < HMD_NBL_PREFETCH; ++d) pj_pre[d] = pj_pre[d + 1];
      {
        const uint32_t qn = q + HMD_NBL_WAVES * HMD_NBL_PREFETCH;
        pj_pre[HMD_NBL_PREFETCH - 1] = zero;
        if (qn < cnt) {
          const uint32_t sn = (s_cand[qn] & 0x7FFFFFFFu) * HMD_CLUSTER + lane;
          if (sn < p.n) pj_pre[HMD_NBL_PREFETCH - 1] = ps[sn];
        }
      }
      // Sub-box mask: which 8-atom i-sub-clusters this j-atom may reach (the
      // nearest point of the sub-box within the search radius, with a relative
      // margin of 2e-4).  The j-atom is listed when it may reach one that holds
      // an i-atom below it in sorted order: every pair within r_search is
      // listed, together with some farther ones, which the force kernels mask
      // by their own cutoff tests.  Empty sub-boxes (past the last atom) have
      // negative extents and never pass.
      uint32_t smask = 0;
      if (sj < p.n) {
        if (nowrap) {
#pragma unroll
          for (uint32_t k = 0; k < 8; ++k) {
            const float4 cs = csub[2 * k], hs = csub[2 * k + 1];
            const float ex = fmaxf(fabsf(cs.x - pj.x) - hs.x, 0.0f);
            const float ey = fmaxf(fabsf(cs.y - pj.y) - hs.y, 0.0f);
            const float ez = fmaxf(fabsf(cs.z - pj.z) - hs.z, 0.0f);
            smask |= (ex * ex + ey * ey + ez * ez <= rs2_prune) ? (1u << k) : 0u;
          }
        } else {
#pragma unroll
          for (uint32_t k = 0; k < 8; ++k) {
            const float g2 = hmd_box_gap2(csub[2 * k], csub[2 * k + 1], pj, zero, p.L, ortho,
                                          ilx, ily, ilz);
            smask |= (g2 <= rs2_prune) ? (1u << k) : 0u;
          }
        }
      }
      // own cluster: only sub-clusters holding an i-atom below this lane's j
      const bool own = jcl == c;
      const uint32_t lim = own ? ((1u << ((lane + 7u) / 8u)) - 1u) : 0xFFu;
      const uint32_t m8 = smask & lim;
      const bool hit = m8 != 0u;
      const uint64_t H = __ballot(hit);
      // ranks among this candidate's hits of the same key, and per-key counts
      const uint32_t key = hit ? hmd_nbl_key(m8, own) : HMD_NBL_KEYS;
      if (lane < HMD_NBL_KEYS) s_koff[q][lane] = 0;
      uint32_t rank = 0;
      uint64_t pending = sorted ? H : 0ull;
      while (pending) {
        const uint32_t leader = (uint32_t)__builtin_ctzll(pending);
        const uint32_t k = hmd_readlane(key, leader);
        const uint64_t B = __ballot(key == k);
        if (key == k) rank = (uint32_t)__builtin_popcountll(B & lt_mask);
        if (lane == 0) s_koff[q][k] = (uint16_t)__builtin_popcountll(B);
        pending &= ~B;
      }
      s_lane[q][lane] = (uint16_t)(m8 | (rank << 8));
      if (lane == 0) s_hit[q] = H;
    }
    __syncthreads();

    // list positions (wave 0): key-major counting sort of the one round, or
    // hit counts in candidate order
    if (wave == 0) {
      uint32_t carry = s_fill;
      if (sorted) {
        // per key (lane < 16): exclusive prefix over the candidates, then the
        // keys' bases by a prefix over the lanes
        uint32_t run = 0;
        for (uint32_t k = 0; k < cnt; ++k) {
          if (lane < HMD_NBL_KEYS) {
            const uint32_t t = s_koff[k][lane];
            s_koff[k][lane] = (uint16_t)run;
            run += t;
          }
        }
        const uint32_t tot = lane < HMD_NBL_KEYS ? run : 0u;
        const uint32_t incl = hmd_wave_scan_incl(tot, lane);
        const uint32_t kbase = carry + incl - tot;
        for (uint32_t k = 0; k < cnt; ++k)
          if (lane < HMD_NBL_KEYS) s_koff[k][lane] = (uint16_t)(s_koff[k][lane] + kbase);
        carry += hmd_readlane(incl, HMD_WAVE - 1);
      } else {
        for (uint32_t b = 0; b < cnt; b += HMD_WAVE) {
          const uint32_t k = b + lane;
          const uint32_t h = k < cnt ? (uint32_t)__builtin_popcountll(s_hit[k]) : 0u;
          const uint32_t incl = hmd_wave_scan_incl(h, lane);
          if (k < cnt) s_off[k] = carry + incl - h;
          carry += hmd_readlane(incl, HMD_WAVE - 1);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (lane == 0) s_fill = carry;
      // exclusion partners in this round's candidates: the candidates are in
      // increasing cluster order, so a binary search finds a partner's one
#pragma unroll
      for (uint32_t m = 0; m < HMD_MAX_INLINE_EXCL; ++m) {
        const uint32_t es = ex_slot[m];
        if (es == HMD_SENTINEL) continue;
        const uint32_t pc = es / HMD_CLUSTER, pb = es % HMD_CLUSTER;
        uint32_t lo = 0, hi = cnt;
        while (lo < hi) {
          const uint32_t mid = (lo + hi) / 2;
          if ((s_cand[mid] & 0x7FFFFFFFu) < pc) lo = mid + 1;
          else hi = mid;
        }
        if (lo < cnt && (s_cand[lo] & 0x7FFFFFFFu) == pc) {
          const uint64_t H = s_hit[lo];
          if ((H >> pb) & 1ull) {
            if (sorted) {
              const uint32_t v = s_lane[lo][pb];
              ex_pos[m] = s_koff[lo][hmd_nbl_key(v & 0xFFu, pc == c)] + (v >> 8);
            } else {
              const uint64_t below = (pb == 0) ? 0ull : (H & (~0ull >> (64 - pb)));
              ex_pos[m] = s_off[lo] + (uint32_t)__builtin_popcountll(below);
            }
          }
        }
      }
    }
    __syncthreads();

    // placement, all waves: a kept j-atom goes to its list position, in the
    // LDS list or, past the slots staged there, straight into the entry
    // array; its mask joins its entry's
    for (uint32_t k = wave; k < cnt; k += HMD_NBL_WAVES) {
      const uint64_t H = s_hit[k];
      if ((H >> lane) & 1ull) {
        const uint32_t cw = s_cand[k] & 0x7FFFFFFFu;
        const uint32_t v = s_lane[k][lane];
        const uint32_t m8 = v & 0xFFu;
        const uint32_t pos = sorted ? s_koff[k][hmd_nbl_key(m8, cw == c)] + (v >> 8)
                                    : s_off[k] + (uint32_t)__builtin_popcountll(H & lt_mask);
        const uint32_t sj = cw * HMD_CLUSTER + lane;
        if (pos < p.lds_slots) s_list[pos] = sj;
        else if (pos < cap_slots) p.entry_j[base_entry * HMD_ENTRY + pos] = sj;
        const uint32_t e = pos / HMD_ENTRY;
        if (e < HMD_NBL_EMASK_CAP) atomicOr(&s_emask[e], m8);
      }
    }
    __syncthreads();
  }
  const uint32_t fill = s_fill;
  const uint32_t n_entries = (fill + HMD_ENTRY - 1) / HMD_ENTRY;
  const uint32_t n_written = min(n_entries, p.cap_entries);
  // whole entries out of LDS, one coalesced 256-byte row per wave and entry
  // (scattered partial-row stores of each candidate's hits straight to
  // memory took half of the build's time); empty slots get the sentinel.
  // Slots past the LDS staging were written by the placement; only their
  // empty tail is left.
  for (uint32_t e = wave; e < n_written; e += HMD_NBL_WAVES) {
    const uint32_t pos = e * HMD_ENTRY + lane;
    if (pos < p.lds_slots)
      p.entry_j[base_entry * HMD_ENTRY + pos] = pos < fill ? s_list[pos] : HMD_SENTINEL;
    else if (pos >= fill)
      p.entry_j[base_entry * HMD_ENTRY + pos] = HMD_SENTINEL;
  }
  if (wave == 0) {
    // Exclusion rows (lane = i-atom, bits over the entry's j-slots) for the
    // entries up to the last one that holds a partner; an entry that holds
    // one is flagged by bit 8 of its mask (the force kernel loads rows only
    // for flagged entries; entries past the kept masks are all flagged, and
    // their rows always written).
    uint32_t e_hi = 0;
#pragma unroll
    for (uint32_t m = 0; m < HMD_MAX_INLINE_EXCL; ++m)
      if (ex_pos[m] != HMD_SENTINEL) e_hi = max(e_hi, ex_pos[m] / HMD_ENTRY + 1u);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) e_hi = max(e_hi, (uint32_t)__shfl_xor((int)e_hi, o, HMD_WAVE));
    const uint32_t n_ex = min(e_hi, n_written);
    for (uint32_t e = 0; e < n_ex; ++e) {
      uint64_t mask = 0;
#pragma unroll
      for (uint32_t m = 0; m < HMD_MAX_INLINE_EXCL; ++m)
        if (ex_pos[m] != HMD_SENTINEL && ex_pos[m] / HMD_ENTRY == e)
          mask |= 1ull << (ex_pos[m] % HMD_ENTRY);
      const bool any = __ballot(mask != 0ull) != 0ull;
      if (any || e >= HMD_NBL_EMASK_CAP) p.entry_excl[(base_entry + e) * HMD_ENTRY + lane] = mask;
      if (any && lane == 0 && e < HMD_NBL_EMASK_CAP) s_emask[e] |= 0x100u;
    }
    if (lane == 0) {
      s_nex = n_ex;
      p.entry_count[c] = n_written;
      p.excl_entries[c] = n_ex;
      atomicMax(&p.status[2], n_entries);   // high-water mark: the host resizes
      if (n_entries > p.cap_entries) atomicOr(p.status, HMD_STATUS_ENTRY_OVERFLOW);
    }
  }
  __syncthreads();
  const uint32_t n_ex = s_nex;
  for (uint32_t e = threadIdx.x; e < n_written; e += blockDim.x)
    p.entry_mask[base_entry + e] =
        e < HMD_NBL_EMASK_CAP ? s_emask[e] : (0xFFu | (e < n_ex ? 0x100u : 0u));
}
HMD_KERNEL(hmd_nbl_build, NblBuildParams, 64 * HMD_NBL_WAVES, HMD_NBL_BUILD_WAVES_PER_EU)

// Flatten (cluster, entry) into the work list the force kernel walks, one
// wave per cluster (entry_offset is the exclusive scan of entry_count with
// the total at [n_clusters]).
__device__ __forceinline__ void hmd_nbl_worklist_body(const NblWorkParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  if (!hmd_gate(p.status, HMD_STATUS_DISPLACED)) return;
  const uint32_t c = hmd_bx * (blockDim.x / HMD_WAVE) + threadIdx.x / HMD_WAVE;
  if (c >= p.n_clusters) return;
  const uint32_t cnt = p.entry_count[c];
  const uint32_t off = p.entry_offset[c];
  for (uint32_t k = hmd_lane(); k < cnt; k += HMD_WAVE) p.work[off + k] = c * p.cap_entries + k;
}
HMD_KERNEL(hmd_nbl_worklist, NblWorkParams, 256)

// Close the rebuild: work-list length, status flags and generation (a kernel
// of its own, after every block of the work-list kernel has read the gate).
__device__ __forceinline__ void hmd_nbl_close_body(const NblWorkParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  if (!hmd_gate(p.status, HMD_STATUS_DISPLACED)) return;
  if (threadIdx.x == 0) {
    p.n_work[0] = p.entry_offset[p.n_clusters];
    uint32_t* st = p.status;
    const uint32_t s = st[0];
    st[0] = (s & ~HMD_STATUS_DISPLACED) | HMD_STATUS_REBUILT;
    st[1] += 1u;   // rebuild generation
  }
}
HMD_KERNEL(hmd_nbl_close, NblWorkParams, 64)

// Decode the list into the exact pair set (canonical (min, max) ids of every
// listed pair whose f32 min-image distance is <= r_search), one wave per
// work item.  This is the PairSnapshot of src/forces/neighbor_list.rs:
// 1358-1445 restricted to pairs inside the search radius.
__device__ __forceinline__ void hmd_nbl_pairs_body(const NblPairsParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  uint32_t w = hmd_bx * (blockDim.x / HMD_WAVE) + threadIdx.x / HMD_WAVE;
  if (w >= *p.n_work) return;
  uint32_t lane = hmd_lane();
  uint32_t slot = p.work[w];
  uint32_t c = slot / p.cap_entries;
  uint32_t si = c * HMD_CLUSTER + lane;
  bool vi = si < p.n;
  const float4* ps = reinterpret_cast<const float4*>(p.posq_sorted);
  float4 a = ps[si];
  uint32_t ida = vi ? p.sorted_id[si] : 0u;
  for (uint32_t k = 0; k < HMD_ENTRY; ++k) {
    uint32_t sj = p.entry_j[(size_t)slot * HMD_ENTRY + k];
    if (sj == HMD_SENTINEL) break;
    float4 b = ps[sj];
    bool ok = vi && si < sj &&
              hmd_r2_exact(a.x, a.y, a.z, b.x, b.y, b.z, p.L) <= p.r_search_sq;
    if (ok) {
      uint32_t idb = p.sorted_id[sj];
      unsigned long long o = atomicAdd(p.count, 1ull);
      if (o < p.cap) {
        p.pairs[2 * o] = min(ida, idb);
        p.pairs[2 * o + 1] = max(ida, idb);
      }
    }
  }
}
HMD_KERNEL(hmd_nbl_pairs, NblPairsParams, 256)
