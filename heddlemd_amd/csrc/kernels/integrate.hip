// Integrator, constraint and reduction kernels for gfx950.
//
// Reference behaviour: kernels/integrate.cu (velocity Verlet, f32 and the
// compensated (f32 hi, f64 lo) "lossless" variants, :24-314),
// kernels/settle.cu (SETTLE, :47-510), kernels/langevin.cu + philox.cuh
// (BAOAB A and O steps, :13-80; Philox-4x32-10, :25-150) and the KE / virial
// reductions of kernels/nose_hoover.cu:22-104 and kernels/barostat.cu:11-85,
// which here accumulate in FP64 with a fixed reduction tree.
#include "hmd_device.h"

// -------------------------------------------------------- velocity Verlet --
// The arithmetic is that of integrate.cu:24-170 as compiled with FMA
// contraction on: a = F / m (IEEE divide), v + a dt/2 and x + v dt fused
// into one rounding each; the compensated (f32 hi + f64 lo) variant adds the
// f32 product a dt/2 to (hi + lo) in f64 and fuses (v_hi + v_lo) dt into the
// f64 position sum.  Written with explicit fmaf / fma so the rounding does
// not depend on the compiler's contraction choices.

// v += F/m dt/2; x += v dt; wrap + image count (integrate.cu:24-170).
__device__ inline void hmd_vv_kick_drift_atom(const VvParams& p, const uint32_t i) {
#pragma clang fp contract(off)
  const float m = p.m[i];
  const float half_dt = p.dt * 0.5f;
  float4 pq = reinterpret_cast<float4*>(p.posq)[i];
  float a[3] = {p.f[i] / m, p.f[p.n + i] / m, p.f[2 * p.n + i] / m};
  float x[3] = {pq.x, pq.y, pq.z};
  if (p.compensated) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const size_t k = (size_t)d * p.n + i;
      double ext_v = ((double)p.v[k] + p.vel_lo[k]) + (double)(a[d] * half_dt);
      float vh = (float)ext_v;
      double vl = ext_v - (double)vh;
      p.v[k] = vh;
      p.vel_lo[k] = vl;
      double ext_x = fma((double)vh + vl, (double)p.dt, (double)x[d] + p.pos_lo[k]);
      float xh = (float)ext_x;
      p.pos_lo[k] = ext_x - (double)xh;
      x[d] = xh;
    }
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const size_t k = (size_t)d * p.n + i;
      float v = fmaf(a[d], half_dt, p.v[k]);
      p.v[k] = v;
      x[d] = fmaf(v, p.dt, x[d]);
    }
  }
  int ka, kb, kc;
  hmd_wrap(x[0], x[1], x[2], ka, kb, kc, p.L);
  p.images[i] += ka;
  p.images[p.n + i] += kb;
  p.images[2 * p.n + i] += kc;
  pq.x = x[0];
  pq.y = x[1];
  pq.z = x[2];
  reinterpret_cast<float4*>(p.posq)[i] = pq;
}

// v += F/m dt/2 (integrate.cu:172-215).
__device__ inline void hmd_vv_kick_atom(const VvParams& p, const uint32_t i) {
#pragma clang fp contract(off)
  const float m = p.m[i];
  const float half_dt = p.dt * 0.5f;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const size_t k = (size_t)d * p.n + i;
    const float a = p.f[k] / m;
    if (p.compensated) {
      double ext_v = ((double)p.v[k] + p.vel_lo[k]) + (double)(a * half_dt);
      float vh = (float)ext_v;
      p.v[k] = vh;
      p.vel_lo[k] = ext_v - (double)vh;
    } else {
      p.v[k] = fmaf(a, half_dt, p.v[k]);
    }
  }
}

// ------------------------------------------------------------------ SETTLE --
// Every f32 operation rounds once (no contraction), so that the result does
// not depend on how the compiler fuses multiplies and adds in the context a
// group function is inlined into (one kernel per sub-step or the fused step
// kernels).
__device__ inline float hmd_csqrt(float x) { return sqrtf(x > 0.0f ? x : 0.0f); }

// Snapshot of the pre-drift positions (settle.cu:47-67).
__device__ inline void hmd_settle_snapshot_group(const SettleParams& p, const uint32_t g) {
  const float4* ps = reinterpret_cast<const float4*>(p.posq);
  float4* sn = reinterpret_cast<float4*>(p.snapshot);
#pragma unroll
  for (int a = 0; a < 3; ++a) sn[3 * g + a] = ps[p.group_atoms[3 * g + a]];
}

// Apply a position correction to atom `idx`: f32 path or, in compensated
// mode, to the extended (hi + lo) value, re-split.
__device__ inline void hmd_apply_pos(const SettleParams& p, uint32_t idx, float4& pq,
                                     const float cur[3], const float corr[3],
                                     float inv_dt) {
#pragma clang fp contract(off)
  float out[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const size_t k = (size_t)d * p.n + idx;
    if (p.pos_lo) {
      double ext = (double)cur[d] + p.pos_lo[k] + (double)corr[d];
      float h = (float)ext;
      p.pos_lo[k] = ext - (double)h;
      out[d] = h;
      double ev = (double)p.v[k] + p.vel_lo[k] + (double)(corr[d] * inv_dt);
      float vh = (float)ev;
      p.vel_lo[k] = ev - (double)vh;
      p.v[k] = vh;
    } else {
      out[d] = cur[d] + corr[d];
      p.v[k] = p.v[k] + corr[d] * inv_dt;
    }
  }
  pq.x = out[0];
  pq.y = out[1];
  pq.z = out[2];
}

// Analytic rigid-water position reset (Miyamoto & Kollman, J. Comput.
// Chem. 13(8), 952-962, 1992) with the half-step velocity fix and the
// position-level constraint virial, as settle.cu:145-335.
__device__ inline void hmd_settle_positions_group(const SettleParams& p, const uint32_t g,
                                                  const float dt) {
#pragma clang fp contract(off)
  const HmdLattice L = p.L;
  float4* ps = reinterpret_cast<float4*>(p.posq);
  const float4* sn = reinterpret_cast<const float4*>(p.snapshot);
  const uint32_t idx[3] = {p.group_atoms[3 * g], p.group_atoms[3 * g + 1],
                           p.group_atoms[3 * g + 2]};
  float4 pq[3] = {ps[idx[0]], ps[idx[1]], ps[idx[2]]};
  float4 s0 = sn[3 * g], s1 = sn[3 * g + 1], s2 = sn[3 * g + 2];
  float cur[3][3] = {{pq[0].x, pq[0].y, pq[0].z},
                     {pq[1].x, pq[1].y, pq[1].z},
                     {pq[2].x, pq[2].y, pq[2].z}};
  float snp[3][3] = {{s0.x, s0.y, s0.z}, {s1.x, s1.y, s1.z}, {s2.x, s2.y, s2.z}};
  // Reference bond vectors and per-atom drifts, minimum-imaged.
  float b0[3] = {snp[1][0] - snp[0][0], snp[1][1] - snp[0][1], snp[1][2] - snp[0][2]};
  float c0[3] = {snp[2][0] - snp[0][0], snp[2][1] - snp[0][1], snp[2][2] - snp[0][2]};
  hmd_min_image(b0[0], b0[1], b0[2], L);
  hmd_min_image(c0[0], c0[1], c0[2], L);
  float xp[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float dx = cur[a][0] - snp[a][0], dy = cur[a][1] - snp[a][1], dz = cur[a][2] - snp[a][2];
    hmd_min_image(dx, dy, dz, L);
    xp[a][0] = dx; xp[a][1] = dy; xp[a][2] = dz;
  }
  const float m0 = p.m_o, m1 = p.m_h, m2 = p.m_h;
  const float inv_total = 1.0f / (m0 + m1 + m2);
  const float ra = p.ra, rb = p.rb, rc = p.rc;
  // Unconstrained positions relative to the reference oxygen, and the COM.
  float unc[3][3], com[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    unc[0][d] = xp[0][d];
    unc[1][d] = b0[d] + xp[1][d];
    unc[2][d] = c0[d] + xp[2][d];
    com[d] = (unc[0][d] * m0 + unc[1][d] * m1 + unc[2][d] * m2) * inv_total;
  }
  float a1[3], b1[3], c1[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    a1[d] = unc[0][d] - com[d];
    b1[d] = unc[1][d] - com[d];
    c1[d] = unc[2][d] - com[d];
  }
  // Frame: Z' normal to the reference plane, X' from the displaced O.
  float ez[3] = {b0[1] * c0[2] - b0[2] * c0[1], b0[2] * c0[0] - b0[0] * c0[2],
                 b0[0] * c0[1] - b0[1] * c0[0]};
  float ex[3] = {a1[1] * ez[2] - a1[2] * ez[1], a1[2] * ez[0] - a1[0] * ez[2],
                 a1[0] * ez[1] - a1[1] * ez[0]};
  float ey[3] = {ez[1] * ex[2] - ez[2] * ex[1], ez[2] * ex[0] - ez[0] * ex[2],
                 ez[0] * ex[1] - ez[1] * ex[0]};
  const float lx_ = hmd_csqrt(ex[0] * ex[0] + ex[1] * ex[1] + ex[2] * ex[2]);
  const float ly_ = hmd_csqrt(ey[0] * ey[0] + ey[1] * ey[1] + ey[2] * ey[2]);
  const float lz_ = hmd_csqrt(ez[0] * ez[0] + ez[1] * ez[1] + ez[2] * ez[2]);
  const float iex = lx_ > 0.0f ? 1.0f / lx_ : 0.0f;
  const float iey = ly_ > 0.0f ? 1.0f / ly_ : 0.0f;
  const float iez = lz_ > 0.0f ? 1.0f / lz_ : 0.0f;
#pragma unroll
  for (int d = 0; d < 3; ++d) { ex[d] *= iex; ey[d] *= iey; ez[d] *= iez; }
  auto dot = [](const float* u, const float* w) { return u[0] * w[0] + u[1] * w[1] + u[2] * w[2]; };
  const float b0x = dot(ex, b0), b0y = dot(ey, b0), c0x = dot(ex, c0), c0y = dot(ey, c0);
  const float a1z = dot(ez, a1);
  const float b1x = dot(ex, b1), b1y = dot(ey, b1), b1z = dot(ez, b1);
  const float c1x = dot(ex, c1), c1y = dot(ey, c1), c1z = dot(ez, c1);
  // Canonical triangle tilted by (phi, psi), then in-plane rotation theta.
  const float sinphi = a1z / ra;
  const float cosphi = hmd_csqrt(1.0f - sinphi * sinphi);
  const float sinpsi = (b1z - c1z) / (2.0f * rc * cosphi);
  const float cospsi = hmd_csqrt(1.0f - sinpsi * sinpsi);
  const float a2y = ra * cosphi;
  float b2x = -rc * cospsi;
  const float b2y = -rb * cosphi - rc * sinpsi * sinphi;
  const float c2y = -rb * cosphi + rc * sinpsi * sinphi;
  const float hh = 4.0f * b2x * b2x + (b2y - c2y) * (b2y - c2y) + (b1z - c1z) * (b1z - c1z);
  const float d2sq = 4.0f * rc * rc;
  const float delta = 2.0f * b2x + hmd_csqrt(4.0f * b2x * b2x - hh + d2sq);
  b2x -= 0.5f * delta;
  const float al = b2x * (b0x - c0x) + b0y * b2y + c0y * c2y;
  const float be = b2x * (c0y - b0y) + b0x * b2y + c0x * c2y;
  const float ga = b0x * b1y - b1x * b0y + c0x * c1y - c1x * c0y;
  const float ab2 = al * al + be * be;
  const float iab2 = ab2 > 0.0f ? 1.0f / ab2 : 0.0f;
  const float sinth = (al * ga - be * hmd_csqrt(ab2 - ga * ga)) * iab2;
  const float costh = hmd_csqrt(1.0f - sinth * sinth);
  const float pr[3][3] = {{-a2y * sinth, a2y * costh, a1z},
                          {b2x * costh - b2y * sinth, b2x * sinth + b2y * costh, b1z},
                          {-b2x * costh - c2y * sinth, -b2x * sinth + c2y * costh, c1z}};
  const float inv_dt = dt != 0.0f ? 1.0f / dt : 0.0f;
  const float inv_dt2 = inv_dt * inv_dt;
  const float mass[3] = {m0, m1, m2};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float rcm[3], corr[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      rcm[d] = ex[d] * pr[a][0] + ey[d] * pr[a][1] + ez[d] * pr[a][2];
      corr[d] = (com[d] + rcm[d]) - unc[a][d];
    }
    hmd_apply_pos(p, idx[a], pq[a], cur[a], corr, inv_dt);
    ps[idx[a]] = pq[a];
    if (p.constraint_virial)
      p.constraint_virial[3 * g + a] =
          mass[a] * inv_dt2 * (corr[0] * rcm[0] + corr[1] * rcm[1] + corr[2] * rcm[2]);
  }
}

// Analytic velocity projection: 3x3 Cramer solve for the bond impulses
// (settle.cu:376-490).
__device__ inline void hmd_settle_velocities_group(const SettleParams& p, const uint32_t g,
                                                   const float dt) {
#pragma clang fp contract(off)
  const HmdLattice L = p.L;
  const float4* ps = reinterpret_cast<const float4*>(p.posq);
  const uint32_t idx[3] = {p.group_atoms[3 * g], p.group_atoms[3 * g + 1],
                           p.group_atoms[3 * g + 2]};
  float px[3], py[3], pz[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float4 q = ps[idx[a]];
    px[a] = q.x; py[a] = q.y; pz[a] = q.z;
  }
#pragma unroll
  for (int a = 1; a < 3; ++a) {
    float dx = px[a] - px[0], dy = py[a] - py[0], dz = pz[a] - pz[0];
    hmd_min_image(dx, dy, dz, L);
    px[a] = px[0] + dx; py[a] = py[0] + dy; pz[a] = pz[0] + dz;
  }
  const float io = 1.0f / p.m_o, ih = 1.0f / p.m_h;
  float v[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) v[a][d] = p.v[(size_t)d * p.n + idx[a]];
  const float r1[3] = {px[0] - px[1], py[0] - py[1], pz[0] - pz[1]};
  const float r2[3] = {px[0] - px[2], py[0] - py[2], pz[0] - pz[2]};
  const float r3[3] = {px[1] - px[2], py[1] - py[2], pz[1] - pz[2]};
  auto dot = [](const float* u, const float* w) { return u[0] * w[0] + u[1] * w[1] + u[2] * w[2]; };
  const float d11 = dot(r1, r1), d22 = dot(r2, r2), d33 = dot(r3, r3);
  const float d12 = dot(r1, r2), d13 = dot(r1, r3), d23 = dot(r2, r3);
  float v01[3], v02[3], v12[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    v01[d] = v[0][d] - v[1][d];
    v02[d] = v[0][d] - v[2][d];
    v12[d] = v[1][d] - v[2][d];
  }
  const float rhs[3] = {-dot(v01, r1), -dot(v02, r2), -dot(v12, r3)};
  const float M[3][3] = {{d11 * (io + ih), d12 * io, -d13 * ih},
                         {d12 * io, d22 * (io + ih), d23 * ih},
                         {-d13 * ih, d23 * ih, 2.0f * d33 * ih}};
  auto det3 = [](const float A[3][3]) {
    return A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
           A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
           A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
  };
  const float det = det3(M);
  if (det == 0.0f) return;
  const float idet = 1.0f / det;
  float gg[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float A[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) A[i][j] = (j == c) ? rhs[i] : M[i][j];
    gg[c] = det3(A) * idet;
  }
  float dv[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    dv[0][d] = (gg[0] * r1[d] + gg[1] * r2[d]) * io;
    dv[1][d] = (-gg[0] * r1[d] + gg[2] * r3[d]) * ih;
    dv[2][d] = (-gg[1] * r2[d] - gg[2] * r3[d]) * ih;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const size_t k = (size_t)d * p.n + idx[a];
      if (p.vel_lo) {
        double ev = (double)v[a][d] + p.vel_lo[k] + (double)dv[a][d];
        float vh = (float)ev;
        p.vel_lo[k] = ev - (double)vh;
        p.v[k] = vh;
      } else {
        p.v[k] = v[a][d] + dv[a][d];
      }
    }
  if (p.constraint_virial && dt > 0.0f) {
    const float inv_dt = 1.0f / dt;
    const float tm = p.m_o + 2.0f * p.m_h;
    const float cx = (p.m_o * px[0] + p.m_h * px[1] + p.m_h * px[2]) / tm;
    const float cy = (p.m_o * py[0] + p.m_h * py[1] + p.m_h * py[2]) / tm;
    const float cz = (p.m_o * pz[0] + p.m_h * pz[1] + p.m_h * pz[2]) / tm;
    const float mass[3] = {p.m_o, p.m_h, p.m_h};
#pragma unroll
    for (int a = 0; a < 3; ++a)
      p.constraint_virial[3 * g + a] +=
          mass[a] * inv_dt * (dv[a][0] * (px[a] - cx) + dv[a][1] * (py[a] - cy) +
                              dv[a][2] * (pz[a] - cz));
  }
}

// The group's constraint virial (position level from the last reset plus
// velocity level from the last projection) added to its atoms' virials:
// settle_virial_scatter (kernels/settle.cu:498-510).  Groups are disjoint,
// so the adds need no atomics.
__device__ inline void hmd_settle_virial_group(const SettleParams& p, const uint32_t g) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const uint32_t i = p.group_atoms[3 * g + a];
    p.particle_virial[i] = p.particle_virial[i] + p.constraint_virial[3 * g + a];
  }
}

// --------------------------------------------------------------- Langevin --
// Philox-4x32-10 (Salmon et al., SC'11), as kernels/philox.cuh:25-70.
__device__ inline void hmd_philox(uint32_t k0, uint32_t k1, uint32_t c[4]) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(c[0], 0xD2511F53u), lo0 = c[0] * 0xD2511F53u;
    const uint32_t hi2 = __umulhi(c[2], 0xCD9E8D57u), lo2 = c[2] * 0xCD9E8D57u;
    const uint32_t n0 = hi2 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo2; c[2] = n2; c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Box-Muller cos branch on the first two words with (x + 1/2) 2^-32
// uniforms, f64 internals (philox.cuh:136-150).
__device__ inline double hmd_philox_gaussian_f64(uint32_t sl, uint32_t sh, uint32_t cl,
                                                 uint32_t ch, uint32_t pid, uint32_t axis) {
  uint32_t c[4] = {cl, ch, pid, axis};
  hmd_philox(sl, sh, c);
  const double u1 = ((double)c[0] + 0.5) * (1.0 / 4294967296.0);
  const double u2 = ((double)c[1] + 0.5) * (1.0 / 4294967296.0);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__device__ inline float hmd_philox_gaussian(uint32_t sl, uint32_t sh, uint32_t cl,
                                            uint32_t ch, uint32_t pid, uint32_t axis) {
  return (float)hmd_philox_gaussian_f64(sl, sh, cl, ch, pid, axis);
}

// A step: x += v dt/2 (one rounding, as the reference's contracted
// x + v * half_dt), wrap (langevin.cu:13-50).
__device__ inline void hmd_lan_drift_half_atom(const LangevinParams& p, const uint32_t i) {
#pragma clang fp contract(off)
  float4 pq = reinterpret_cast<float4*>(p.posq)[i];
  const float h = p.dt * 0.5f;
  float x = fmaf(p.v[i], h, pq.x), y = fmaf(p.v[p.n + i], h, pq.y),
        z = fmaf(p.v[2 * p.n + i], h, pq.z);
  int ka, kb, kc;
  hmd_wrap(x, y, z, ka, kb, kc, p.L);
  p.images[i] += ka;
  p.images[p.n + i] += kb;
  p.images[2 * p.n + i] += kc;
  pq.x = x; pq.y = y; pq.z = z;
  reinterpret_cast<float4*>(p.posq)[i] = pq;
}

// O step: v = a v + sigma xi with sigma = sqrt((1 - a^2) kT / m) in f32
// (1 - a^2 and a v + (sigma xi) each rounded once, as the reference's
// contracted expressions), xi from Philox keyed by the seed with counter
// (draw_lo, draw_hi, particle id, axis) (langevin.cu:52-80).
__device__ inline void hmd_lan_ou_atom(const LangevinParams& p, const uint32_t i) {
#pragma clang fp contract(off)
  const unsigned long long step = *p.step;
  const uint32_t cl = (uint32_t)(step & 0xFFFFFFFFull), ch = (uint32_t)(step >> 32);
  const float sig = sqrtf(fmaf(-p.a, p.a, 1.0f) * p.kt / p.m[i]);
#pragma unroll
  for (uint32_t d = 0; d < 3; ++d) {
    const float xi = hmd_philox_gaussian(p.seed_lo, p.seed_hi, cl, ch, i, d);
    const size_t k = (size_t)d * p.n + i;
    p.v[k] = fmaf(p.a, p.v[k], sig * xi);
  }
}

// ---------------------------------------------------- one kernel per op --
// The reference's sub-steps one launch each (the standalone operator entry
// points and the unfused integrator plan).
__device__ __forceinline__ void hmd_vv_kick_drift_body(const VvParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  const uint32_t i = hmd_bx * blockDim.x + threadIdx.x;
  if (i < p.n) hmd_vv_kick_drift_atom(p, i);
}
HMD_KERNEL(hmd_vv_kick_drift, VvParams, 256)
__device__ __forceinline__ void hmd_vv_kick_body(const VvParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  const uint32_t i = hmd_bx * blockDim.x + threadIdx.x;
  if (i < p.n) hmd_vv_kick_atom(p, i);
}
HMD_KERNEL(hmd_vv_kick, VvParams, 256)
__device__ __forceinline__ void hmd_settle_snapshot_body(const SettleParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  const uint32_t g = hmd_bx * blockDim.x + threadIdx.x;
  if (g < p.n_groups) hmd_settle_snapshot_group(p, g);
}
HMD_KERNEL(hmd_settle_snapshot, SettleParams, 256)
__device__ __forceinline__ void hmd_settle_positions_body(const SettleParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  const uint32_t g = hmd_bx * blockDim.x + threadIdx.x;
  if (g < p.n_groups) hmd_settle_positions_group(p, g, p.dt);
}
HMD_KERNEL(hmd_settle_positions, SettleParams, 256)
__device__ __forceinline__ void hmd_settle_velocities_body(const SettleParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  const uint32_t g = hmd_bx * blockDim.x + threadIdx.x;
  if (g < p.n_groups) hmd_settle_velocities_group(p, g, p.dt);
}
HMD_KERNEL(hmd_settle_velocities, SettleParams, 256)
__device__ __forceinline__ void hmd_settle_virial_body(const SettleParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  const uint32_t g = hmd_bx * blockDim.x + threadIdx.x;
  if (g < p.n_groups) hmd_settle_virial_group(p, g);
}
HMD_KERNEL(hmd_settle_virial, SettleParams, 256)
__device__ __forceinline__ void hmd_lan_drift_half_body(const LangevinParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  const uint32_t i = hmd_bx * blockDim.x + threadIdx.x;
  if (i < p.n) hmd_lan_drift_half_atom(p, i);
}
HMD_KERNEL(hmd_lan_drift_half, LangevinParams, 256)
__device__ __forceinline__ void hmd_lan_ou_body(const LangevinParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  const uint32_t i = hmd_bx * blockDim.x + threadIdx.x;
  if (i < p.n) hmd_lan_ou_atom(p, i);
}
HMD_KERNEL(hmd_lan_ou, LangevinParams, 256)

// ------------------------------------------------------ fused step kernels --
// The integrator's sub-steps between two force evaluations in one launch.
// Thread u < n_groups owns SETTLE group u and runs every sub-step for its
// three atoms in plan order (the constraint couples only the atoms of one
// group, so no other thread's data is read); thread n_groups + k owns free
// atom free_atoms[k] (atom k when free_atoms is null).  Each atom sees the
// same operations, in the same order and with the same rounding, as in the
// one-kernel-per-sub-step plan.
__device__ inline bool hmd_unit(const StepParams& p, uint32_t u, uint32_t& g, uint32_t& atom) {
  if (u < p.st.n_groups) { g = u; return true; }
  const uint32_t k = u - p.st.n_groups;
  g = HMD_SENTINEL;
  atom = p.free_atoms ? p.free_atoms[k] : k;
  return k < p.n_free;
}

// The neighbour list's skin check (hmd_nbl_displacement, the same f32
// arithmetic) on an atom's final position of the step, folded into the
// kernel that wrote it: one dependent launch less before the force
// evaluation.
__device__ inline void hmd_step_disp_check(const NblDispParams& d, uint32_t atom) {
#pragma clang fp contract(off)
  if (!d.status) return;
  const float4 a = reinterpret_cast<const float4*>(d.posq)[atom];
  const float4 r = reinterpret_cast<const float4*>(d.ref_pos)[atom];
  float dx = a.x - r.x, dy = a.y - r.y, dz = a.z - r.z;
  hmd_min_image(dx, dy, dz, d.L);
  const float d2 = dx * dx + dy * dy + dz * dz;
  if (d2 > d.threshold_sq) atomicOr(d.status, HMD_STATUS_DISPLACED);
}

// velocity Verlet: [snapshot] kick-drift [SETTLE positions] [skin check]
__device__ __forceinline__ void hmd_step_vv_first_body(const StepParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  const uint32_t u = hmd_bx * blockDim.x + threadIdx.x;
  uint32_t g, atom;
  if (!hmd_unit(p, u, g, atom)) return;
  if (g != HMD_SENTINEL) {
    hmd_settle_snapshot_group(p.st, g);
#pragma unroll 1
    for (int a = 0; a < 3; ++a) hmd_vv_kick_drift_atom(p.vv, p.st.group_atoms[3 * g + a]);
    hmd_settle_positions_group(p.st, g, p.dt);
#pragma unroll 1
    for (int a = 0; a < 3; ++a) hmd_step_disp_check(p.disp, p.st.group_atoms[3 * g + a]);
  } else {
    hmd_vv_kick_drift_atom(p.vv, atom);
    hmd_step_disp_check(p.disp, atom);
  }
}
HMD_KERNEL(hmd_step_vv_first, StepParams, 256)

// Langevin BAOAB up to the force evaluation: B [SETTLE v] [snapshot] A
// [SETTLE x, dt/2] O [SETTLE v] [snapshot] A [SETTLE x, dt/2] [skin check]
__device__ __forceinline__ void hmd_step_baoa_body(const StepParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  const uint32_t u = hmd_bx * blockDim.x + threadIdx.x;
  uint32_t g, atom;
  if (!hmd_unit(p, u, g, atom)) return;
  const float h = 0.5f * p.dt;
  if (g != HMD_SENTINEL) {
    const uint32_t* ga = p.st.group_atoms + 3 * g;
#pragma unroll 1
    for (int a = 0; a < 3; ++a) hmd_vv_kick_atom(p.vv, ga[a]);
    hmd_settle_velocities_group(p.st, g, p.dt);
    hmd_settle_snapshot_group(p.st, g);
#pragma unroll 1
    for (int a = 0; a < 3; ++a) hmd_lan_drift_half_atom(p.lan, ga[a]);
    hmd_settle_positions_group(p.st, g, h);
#pragma unroll 1
    for (int a = 0; a < 3; ++a) hmd_lan_ou_atom(p.lan, ga[a]);
    hmd_settle_velocities_group(p.st, g, p.dt);
    hmd_settle_snapshot_group(p.st, g);
#pragma unroll 1
    for (int a = 0; a < 3; ++a) hmd_lan_drift_half_atom(p.lan, ga[a]);
    hmd_settle_positions_group(p.st, g, h);
#pragma unroll 1
    for (int a = 0; a < 3; ++a) hmd_step_disp_check(p.disp, ga[a]);
  } else {
    hmd_vv_kick_atom(p.vv, atom);
    hmd_lan_drift_half_atom(p.lan, atom);
    hmd_lan_ou_atom(p.lan, atom);
    hmd_lan_drift_half_atom(p.lan, atom);
    hmd_step_disp_check(p.disp, atom);
  }
}
HMD_KERNEL(hmd_step_baoa, StepParams, 256)

// after the force evaluation: kick [SETTLE velocities]; the Langevin draw
// counter (read by this step's O sub-step) advances here
__device__ __forceinline__ void hmd_step_kick_last_body(const StepParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  const uint32_t u = hmd_bx * blockDim.x + threadIdx.x;
  if (u == 0 && p.counter) p.counter[0] += 1ull;
  uint32_t g, atom;
  if (!hmd_unit(p, u, g, atom)) return;
  if (g != HMD_SENTINEL) {
#pragma unroll 1
    for (int a = 0; a < 3; ++a) hmd_vv_kick_atom(p.vv, p.st.group_atoms[3 * g + a]);
    hmd_settle_velocities_group(p.st, g, p.dt);
    if (p.st.particle_virial) hmd_settle_virial_group(p.st, g);
  } else {
    hmd_vv_kick_atom(p.vv, atom);
  }
}
HMD_KERNEL(hmd_step_kick_last, StepParams, 256)

__device__ __forceinline__ void hmd_counter_increment_body(const CounterParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  if (threadIdx.x == 0 && hmd_bx == 0) p.counter[0] += 1ull;
}
HMD_KERNEL(hmd_counter_increment, CounterParams, 64)

// -------------------------------------------------------------- reductions --
// Fixed-shape FP64 reduction: block b sums a fixed contiguous chunk in a
// fixed per-thread order and a fixed LDS tree; hmd_reduce_final sums the
// block partials the same way.  Mode 0: sum x[0..n); mode 1: kinetic energy
// 0.5 m |v|^2 (nose_hoover.cu:22-104 and barostat.cu:11-85 restated in
// FP64).
__device__ __forceinline__ void hmd_reduce_partial_body(const ReduceParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  __shared__ double s[256];
  const uint32_t chunk = (p.n + hmd_gx - 1) / hmd_gx;
  const uint32_t beg = hmd_bx * chunk, end = min(beg + chunk, p.n);
  double acc = 0.0;
  for (uint32_t i = beg + threadIdx.x; i < end; i += blockDim.x) {
    if (p.mode == 0) {
      acc += (double)p.x[i];
    } else {
      const double vx = p.v[i], vy = p.v[p.n + i], vz = p.v[2 * p.n + i];
      acc += 0.5 * (double)p.m[i] * (vx * vx + vy * vy + vz * vz);
    }
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) p.partials[hmd_bx] = s[0];
}
HMD_KERNEL(hmd_reduce_partial, ReduceParams, 256)

__device__ __forceinline__ void hmd_reduce_final_body(const ReduceParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  __shared__ double s[256];
  double acc = 0.0;
  for (uint32_t i = threadIdx.x; i < p.n_blocks; i += blockDim.x) acc += p.partials[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) p.out[0] = s[0];
}
HMD_KERNEL(hmd_reduce_final, ReduceParams, 256)

// ------------------------------------------------------------------- CSVR --
// Stochastic velocity rescaling (Bussi, Donadio & Parrinello, J. Chem. Phys.
// 126, 014101 (2007)) as kernels/nose_hoover.cu:181-360 applies it after a
// step: with the kinetic energy K, N_f = g_dof and the normals xi_i of
// Philox (seed, draw counter, i, 0),
//   K' = c K + (kT/2)(1 - c)(sum_{i=1}^{N_f-1} xi_i^2 + xi_0^2)
//        + 2 xi_0 sqrt(c (1 - c) K kT/2),
// v <- sqrt(K'/K) v, and the injected energy K' - K accumulates for the
// conserved quantity.  All on the device (capturable); the draw counter
// advances once per application.  Sums: a fixed chunk per block, a fixed
// lane stride and LDS tree per block, a fixed tree over the blocks.
__device__ inline double hmd_block_sum(double v, double* s) {
  s[threadIdx.x] = v;
  __syncthreads();
  for (uint32_t st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  return s[0];
}

__device__ __forceinline__ void hmd_csvr_sample_body(const CsvrParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  __shared__ double s[256];
  const unsigned long long ctr = *p.counter;
  const uint32_t cl = (uint32_t)(ctr & 0xFFFFFFFFull), ch = (uint32_t)(ctr >> 32);
  const uint32_t m = p.g_dof > 0u ? p.g_dof - 1u : 0u;   // draws i = 1 .. g_dof - 1
  const uint32_t chunk = (m + hmd_gx - 1u) / hmd_gx;
  const uint32_t beg = hmd_bx * chunk, end = min(beg + chunk, m);
  double acc = 0.0;
  for (uint32_t j = beg + threadIdx.x; j < end; j += blockDim.x) {
    const double xi = hmd_philox_gaussian_f64(p.seed_lo, p.seed_hi, cl, ch, j + 1u, 0u);
    acc += xi * xi;
  }
  const double t = hmd_block_sum(acc, s);
  if (threadIdx.x == 0) p.partials[hmd_bx] = t;
}
HMD_KERNEL(hmd_csvr_sample, CsvrParams, 256)

__device__ __forceinline__ void hmd_csvr_finish_body(const CsvrParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  __shared__ double s[256];
  double acc = 0.0;
  for (uint32_t i = threadIdx.x; i < p.n_partials; i += blockDim.x) acc += p.partials[i];
  const double sum = hmd_block_sum(acc, s);
  if (threadIdx.x == 0) {
    const unsigned long long ctr = *p.counter;
    const uint32_t cl = (uint32_t)(ctr & 0xFFFFFFFFull), ch = (uint32_t)(ctr >> 32);
    const double r = hmd_philox_gaussian_f64(p.seed_lo, p.seed_hi, cl, ch, 0u, 0u);
    const double k = *p.ke;
    const double cross = k > 0.0 ? 2.0 * r * sqrt(p.c * p.one_minus_c * k * p.kt_half) : 0.0;
    double kn = p.c * k + p.kt_half * p.one_minus_c * (sum + r * r) + cross;
    if (!isfinite(kn) || kn <= 0.0) kn = k;
    const double f = (k > 0.0 && kn != k) ? sqrt(kn / k) : 1.0;
    p.factor[0] = (float)f;
    p.injection[0] += kn - k;
    *p.counter = ctr + 1ull;
  }
}
HMD_KERNEL(hmd_csvr_finish, CsvrParams, 256)

// v <- factor v (f32, as rescale_velocities_device_factor); in compensated
// mode the extended velocity hi + lo is scaled in f64 and split again.
__device__ __forceinline__ void hmd_csvr_rescale_body(const CsvrParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  const uint32_t i = hmd_bx * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  const float f = p.factor[0];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const size_t k = (size_t)d * p.n + i;
    if (p.vel_lo) {
      const double ev = ((double)p.v[k] + p.vel_lo[k]) * (double)f;
      const float vh = (float)ev;
      p.vel_lo[k] = ev - (double)vh;
      p.v[k] = vh;
    } else {
      p.v[k] = p.v[k] * f;
    }
  }
}
HMD_KERNEL(hmd_csvr_rescale, CsvrParams, 256)

// ------------------------------------------------------- c-rescale --
// Stochastic cell rescaling (Bernetti & Bussi, J. Chem. Phys. 153, 114107
// (2020)), one isotropic move per step as kernels/barostat.cu:169-227 makes
// it, in FP64 on one thread: with the kinetic energy K, the virial sum W and
// the volume V,
//   P   = (2K + W) / (3V),
//   mu3 = 1 - beta (dt / tau)(P_target - P) + sqrt(2 beta kT dt / (tau V)) r,
// r the Philox normal (seed, draw counter, 0, 0); mu3 is floored at mu3_min
// and mu = cbrt(mu3).  The energy P_target (V mu3 - V) accumulates for the
// conserved quantity and the draw counter advances once.
__device__ __forceinline__ void hmd_crescale_mu_body(const CRescaleParams& p,
    const uint32_t hmd_bx, const uint32_t hmd_gx) {
  if (threadIdx.x != 0 || hmd_bx != 0) return;
  const unsigned long long ctr = *p.counter;
  const uint32_t cl = (uint32_t)(ctr & 0xFFFFFFFFull), ch = (uint32_t)(ctr >> 32);
  const double v = p.v_pre;
  const double pressure = (2.0 * p.ke[0] + p.w[0]) / (3.0 * v);
  const double r = hmd_philox_gaussian_f64(p.seed_lo, p.seed_hi, cl, ch, 0u, 0u);
  const double drift = -p.beta * (p.dt / p.tau) * (p.p_target - pressure);
  const double noise = sqrt(2.0 * p.beta * p.kt * p.dt / (p.tau * v));
  double mu3 = 1.0 + drift + noise * r;
  mu3 = mu3 > p.mu3_min ? mu3 : p.mu3_min;
  const double v_post = v * mu3;
  p.state[0] = cbrt(mu3);
  p.state[1] = pressure;
  p.state[2] = v_post;
  p.state[3] += p.p_target * (v_post - v);
  *p.counter = ctr + 1ull;
}
HMD_KERNEL(hmd_crescale_mu, CRescaleParams, 64)

// x <- (float) mu x for every atom (rescale_positions_device_factor,
// kernels/barostat.cu:93-107); fractional coordinates and image counts are
// unchanged.  In compensated mode the extended position hi + lo is scaled in
// f64 and split again.
__device__ __forceinline__ void hmd_crescale_positions_body(const CRescaleParams& p,
    const uint32_t hmd_bx, const uint32_t hmd_gx) {
  const uint32_t i = hmd_bx * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  const float f = (float)p.state[0];
  float4 pq = reinterpret_cast<float4*>(p.posq)[i];
  float x[3] = {pq.x, pq.y, pq.z};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (p.pos_lo) {
      const size_t k = (size_t)d * p.n + i;
      const double e = ((double)x[d] + p.pos_lo[k]) * (double)f;
      x[d] = (float)e;
      p.pos_lo[k] = e - (double)x[d];
    } else {
      x[d] = x[d] * f;
    }
  }
  pq.x = x[0]; pq.y = x[1]; pq.z = x[2];
  reinterpret_cast<float4*>(p.posq)[i] = pq;
}
HMD_KERNEL(hmd_crescale_positions, CRescaleParams, 256)

// ------------------------------------------------------- steepest descent --
// x += h F, with |h F| capped at max_step per atom, then wrap (setup-time
// relaxation of synthetic lattices; the reference relaxes its examples with
// its steepest-descent minimizer, src/minimizer/).
__device__ __forceinline__ void hmd_sd_step_body(const SdParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  const uint32_t i = hmd_bx * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  float4 pq = reinterpret_cast<float4*>(p.posq)[i];
  float d[3] = {p.f[i] * p.h, p.f[p.n + i] * p.h, p.f[2 * p.n + i] * p.h};
  const float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const float s = len > p.max_step ? p.max_step / len : 1.0f;
  float x = pq.x + d[0] * s, y = pq.y + d[1] * s, z = pq.z + d[2] * s;
  int ka, kb, kc;
  hmd_wrap(x, y, z, ka, kb, kc, p.L);
  pq.x = x; pq.y = y; pq.z = z;
  reinterpret_cast<float4*>(p.posq)[i] = pq;
}
HMD_KERNEL(hmd_sd_step, SdParams, 256)
