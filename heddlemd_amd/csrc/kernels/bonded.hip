// Bonded terms: harmonic and Morse bonds, harmonic angles and periodic
// dihedrals, the four intramolecular potentials of src/forces/
// harmonic_bond.rs:316-360, morse.rs:330-360, angle.rs:350-410 and
// dihedral.rs:440-530, which the reference composes into one module per
// arity (src/forces/jit_composed.rs:1845-2406).
//
// One thread per term.  The per-term arithmetic is the reference functors',
// in f32, with the displacements r_i - r_j, r_k - r_j, r_k - r_l under the
// minimum image of kernels/pbc.cuh.  The reference writes one force (and
// energy and virial share) per term endpoint and gathers them per atom in a
// fixed order; here each endpoint's force goes to fixed point (2^-40) and is
// added with an integer atomic into the same per-atom accumulators the pair
// kernel fills, which the finalize converts: integer addition is
// associative, so the sum does not depend on which thread adds first and no
// staging array or gather pass is needed.  Energy and virial shares: u/2,
// w/2 per bond endpoint, u/3, w/3 per angle endpoint, u/4, w/4 per dihedral
// endpoint, as the reference's entry points split them.
#include "hmd_device.h"
#include "hmd_params.h"

__device__ inline void hmd_bonded_add(const BondedParams& p, uint32_t atom, float fx, float fy,
                                      float fz, float e, float w) {
  const uint32_t s = p.slot_of[atom];
  long long* a = p.acc;
  const uint32_t st = p.acc_stride;
  hmd_atomic_add_fixed(&a[s], hmd_to_fixed(fx, HMD_FIXED_SHIFT));
  hmd_atomic_add_fixed(&a[st + s], hmd_to_fixed(fy, HMD_FIXED_SHIFT));
  hmd_atomic_add_fixed(&a[2 * st + s], hmd_to_fixed(fz, HMD_FIXED_SHIFT));
  if (p.fev) {
    hmd_atomic_add_fixed(&a[3 * st + s], hmd_to_fixed(e, HMD_FIXED_SHIFT));
    hmd_atomic_add_fixed(&a[4 * st + s], hmd_to_fixed(w, HMD_FIXED_SHIFT));
  }
}

__device__ inline float3 hmd_bonded_disp(const BondedParams& p, uint32_t a, uint32_t b) {
  const float4 pa = reinterpret_cast<const float4*>(p.posq)[a];
  const float4 pb = reinterpret_cast<const float4*>(p.posq)[b];
  float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
  hmd_min_image(dx, dy, dz, p.L);
  return make_float3(dx, dy, dz);
}

// Bond: U = k (r - r0)^2 / 2 (harmonic) or De (1 - e^{-a (r - r0)})^2
// (Morse); F_i = fmag (r_i - r_j), F_j = -F_i, W = fmag r^2.
__device__ inline void hmd_bonded_bond(const BondedParams& p, uint32_t t) {
  const uint32_t* q = p.terms + 3u * t;
  const uint32_t i = q[0], j = q[1];
  const float4 c = p.types[q[2]];
  const float3 d = hmd_bonded_disp(p, i, j);
  const float r2 = d.x * d.x + d.y * d.y + d.z * d.z;
  float fmag = 0.0f, u = 0.0f, w = 0.0f;
  if (r2 != 0.0f) {
    const float r = sqrtf(r2);
    if (p.kind == HMD_BONDED_HARMONIC_BOND) {
      if (r >= 1.0e-7f) {
        const float dr = r - c.y;
        fmag = -c.x * dr / r;
        u = 0.5f * c.x * dr * dr;
        w = fmag * r2;
      }
    } else {
      const float e = expf(-c.y * (r - c.z));
      const float ome = 1.0f - e;
      fmag = -2.0f * c.x * c.y * ome * e / r;
      u = c.x * ome * ome;
      w = fmag * r2;
    }
  }
  hmd_bonded_add(p, i, fmag * d.x, fmag * d.y, fmag * d.z, 0.5f * u, 0.5f * w);
  hmd_bonded_add(p, j, -fmag * d.x, -fmag * d.y, -fmag * d.z, 0.5f * u, 0.5f * w);
}

// Angle i-j-k: U = k_theta (theta - theta0)^2 / 2, theta from atan2 of
// |b1||b2| sin and b1.b2; F_j = -(F_i + F_k).
__device__ inline void hmd_bonded_angle(const BondedParams& p, uint32_t t) {
  const uint32_t* q = p.terms + 4u * t;
  const uint32_t i = q[0], j = q[1], k = q[2];
  const float4 c = p.types[q[3]];
  const float3 a = hmd_bonded_disp(p, i, j);
  const float3 b = hmd_bonded_disp(p, k, j);
  float fi[3] = {0.0f, 0.0f, 0.0f}, fk[3] = {0.0f, 0.0f, 0.0f}, u = 0.0f, w = 0.0f;
  const float a2 = a.x * a.x + a.y * a.y + a.z * a.z;
  const float b2 = b.x * b.x + b.y * b.y + b.z * b.z;
  if (a2 != 0.0f && b2 != 0.0f) {
    const float la = sqrtf(a2), lb = sqrtf(b2);
    const float inv_ab = 1.0f / (la * lb);
    const float dot = a.x * b.x + a.y * b.y + a.z * b.z;
    float ct = dot * inv_ab;
    ct = fminf(fmaxf(ct, -1.0f), 1.0f);
    const float s2 = 1.0f - ct * ct;
    const float st = sqrtf(s2 > 0.0f ? s2 : 0.0f);
    if (st >= 1.0e-7f) {
      const float theta = atan2f(la * lb * st, dot);
      const float dth = theta - c.y;
      const float g = -c.x * dth / st;
      const float ia2 = 1.0f / a2, ib2 = 1.0f / b2;
      const float av[3] = {a.x, a.y, a.z}, bv[3] = {b.x, b.y, b.z};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        fi[d] = g * (ct * ia2 * av[d] - inv_ab * bv[d]);
        fk[d] = g * (ct * ib2 * bv[d] - inv_ab * av[d]);
      }
      u = 0.5f * c.x * dth * dth;
      w = (a.x * fi[0] + a.y * fi[1] + a.z * fi[2]) + (b.x * fk[0] + b.y * fk[1] + b.z * fk[2]);
    }
  }
  const float es = u * (1.0f / 3.0f), ws = w * (1.0f / 3.0f);
  hmd_bonded_add(p, i, fi[0], fi[1], fi[2], es, ws);
  hmd_bonded_add(p, j, -(fi[0] + fk[0]), -(fi[1] + fk[1]), -(fi[2] + fk[2]), es, ws);
  hmd_bonded_add(p, k, fk[0], fk[1], fk[2], es, ws);
}

// Dihedral i-j-k-l: U = k_phi (1 + cos(n phi - phi0)), phi from atan2 of
// |b2| b1.(b2 x b3) and (b1 x b2).(b2 x b3) with b1 = r_i - r_j,
// b2 = r_k - r_j, b3 = r_k - r_l (Blondel & Karplus, J. Comput. Chem. 17,
// 1132 (1996) gradients).
__device__ inline void hmd_bonded_dihedral(const BondedParams& p, uint32_t t) {
  const uint32_t* q = p.terms + 5u * t;
  const uint32_t i = q[0], j = q[1], k = q[2], l = q[3];
  const float4 c = p.types[q[4]];
  const float3 b1 = hmd_bonded_disp(p, i, j);
  const float3 b2 = hmd_bonded_disp(p, k, j);
  const float3 b3 = hmd_bonded_disp(p, k, l);
  const float mx = b1.y * b2.z - b1.z * b2.y, my = b1.z * b2.x - b1.x * b2.z,
              mz = b1.x * b2.y - b1.y * b2.x;
  const float nx = b2.y * b3.z - b2.z * b3.y, ny = b2.z * b3.x - b2.x * b3.z,
              nz = b2.x * b3.y - b2.y * b3.x;
  const float m2 = mx * mx + my * my + mz * mz;
  const float n2 = nx * nx + ny * ny + nz * nz;
  const float bb = b2.x * b2.x + b2.y * b2.y + b2.z * b2.z;
  float fi[3] = {0, 0, 0}, fj[3] = {0, 0, 0}, fk[3] = {0, 0, 0}, fl[3] = {0, 0, 0};
  float u = 0.0f, w = 0.0f;
  if (!(m2 < 1.0e-14f || n2 < 1.0e-14f || bb < 1.0e-14f)) {
    const float lb = sqrtf(bb);
    const float phi = atan2f(lb * (b1.x * nx + b1.y * ny + b1.z * nz),
                             mx * nx + my * ny + mz * nz);
    const float nm = (float)(uint32_t)c.z;
    const float delta = nm * phi - c.y;
    const float fphi = c.x * nm * sinf(delta);
    const float ci = fphi * lb / m2, cl = -fphi * lb / n2;
    fi[0] = ci * mx; fi[1] = ci * my; fi[2] = ci * mz;
    fl[0] = cl * nx; fl[1] = cl * ny; fl[2] = cl * nz;
    const float ib = 1.0f / bb;
    const float s = (b1.x * b2.x + b1.y * b2.y + b1.z * b2.z) * ib;
    const float tt = (b3.x * b2.x + b3.y * b2.y + b3.z * b2.z) * ib;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      fj[d] = (s - 1.0f) * fi[d] - tt * fl[d];
      fk[d] = -s * fi[d] + (tt - 1.0f) * fl[d];
    }
    u = c.x * (1.0f + cosf(delta));
    w = (b1.x * fi[0] + b1.y * fi[1] + b1.z * fi[2]) + (b2.x * fk[0] + b2.y * fk[1] + b2.z * fk[2]) +
        ((b2.x - b3.x) * fl[0] + (b2.y - b3.y) * fl[1] + (b2.z - b3.z) * fl[2]);
  }
  const float es = 0.25f * u, ws = 0.25f * w;
  hmd_bonded_add(p, i, fi[0], fi[1], fi[2], es, ws);
  hmd_bonded_add(p, j, fj[0], fj[1], fj[2], es, ws);
  hmd_bonded_add(p, k, fk[0], fk[1], fk[2], es, ws);
  hmd_bonded_add(p, l, fl[0], fl[1], fl[2], es, ws);
}

__device__ __forceinline__ void hmd_bonded_terms_body(const BondedParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  const uint32_t t = hmd_bx * blockDim.x + threadIdx.x;
  if (t >= p.n_terms) return;
  if (p.kind == HMD_BONDED_HARMONIC_ANGLE) hmd_bonded_angle(p, t);
  else if (p.kind == HMD_BONDED_PERIODIC_DIHEDRAL) hmd_bonded_dihedral(p, t);
  else hmd_bonded_bond(p, t);
}
HMD_KERNEL(hmd_bonded_terms, BondedParams, 256)
