// LDS-resident 3D real<->complex FFT and fused influence-function solve for
// SPME on gfx950.  Replaces the reference's closed-library R2C / C2R calls
// and its separate influence kernels (src/forces/spme.rs:423-500,
// kernels/spme_recip.cu:34-114, :296-402).
//
// The transform is five line passes over a complex mesh of shape
// (n0, n1, n2/2 + 1) (last axis fastest, as the reference's R2C layout):
//   1. c axis, real -> complex (an n2/2 complex FFT plus the standard
//      even/odd split),
//   2. b axis forward,
//   3. a axis forward, multiply by the influence function G(k) while
//      accumulating the Hermitian-weighted virial and energy, a axis
//      inverse -- one kernel, the line never leaves LDS,
//   4. b axis inverse,
//   5. c axis, complex -> real (unnormalised, like the reference's C2R).
// Each workgroup loads several lines into LDS with coalesced accesses, runs
// a mixed-radix (2, 3, 4, 5) Stockham FFT there, and writes them back.
#include "hmd_device.h"

#define HMD_FFT_THREADS 256u

__device__ inline float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by sign * i
__device__ inline float2 cmul_si(float2 a, int sign) {
  return sign < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

// One radix-R butterfly: y_k = sum_r x_r e^{sign 2 pi i r k / R}.
__device__ inline void hmd_dft(float2* v, uint32_t R, int sign) {
  if (R == 2) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if (R == 4) {
    float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    float2 t2 = cadd(v[1], v[3]), t3 = cmul_si(csub(v[1], v[3]), sign);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  } else if (R == 3) {
    const float s3 = 0.86602540378443865f;
    float2 t = cadd(v[1], v[2]);
    float2 u = cmul_si(csub(v[1], v[2]), sign);
    float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
    v[0] = cadd(v[0], t);
    v[1] = make_float2(m.x + s3 * u.x, m.y + s3 * u.y);
    v[2] = make_float2(m.x - s3 * u.x, m.y - s3 * u.y);
  } else {  // R == 5
    const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
    const float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;
    float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    float2 x0 = v[0];
    float2 m1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
    float2 m2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
    float2 n1 = cmul_si(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y), sign);
    float2 n2 = cmul_si(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y), sign);
    v[0] = cadd(x0, cadd(a1, a2));
    v[1] = cadd(m1, n1);
    v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[3] = csub(m2, n2);
  }
}

// Mixed-radix Stockham autosort FFT of `lines` lines of length n stored
// line-major in LDS with line stride ls (n + 1: the pad spreads the strided
// loads of the b and a passes over the LDS banks); buf1 is scratch.  Returns the buffer holding the
// result.  Twiddles come from tw[m] = e^{-2 pi i m / n} (conjugated for the
// inverse).
__device__ float2* hmd_fft_lds(float2* buf0, float2* buf1, uint32_t n,
                               uint32_t ls, uint32_t lines,
                               const uint32_t* radix, uint32_t n_stages,
                               const float2* tw, int sign) {
  float2* src = buf0;
  float2* dst = buf1;
  uint32_t Ns = 1;
  for (uint32_t st = 0; st < n_stages; ++st) {
    const uint32_t R = radix[st];
    const uint32_t nb = n / R;
    const uint32_t step = n / (Ns * R);
    for (uint32_t b = threadIdx.x; b < lines * nb; b += blockDim.x) {
      const uint32_t l = b / nb, j = b - l * nb;
      const uint32_t jm = j % Ns;
      float2 v[5];
      for (uint32_t r = 0; r < R; ++r) {
        float2 x = src[l * ls + j + r * nb];
        if (r > 0) {
          float2 w = tw[(jm * r * step) % n];
          if (sign > 0) w = cconj(w);
          x = cmul(x, w);
        }
        v[r] = x;
      }
      hmd_dft(v, R, sign);
      const uint32_t idx = (j / Ns) * Ns * R + jm;
      for (uint32_t r = 0; r < R; ++r) dst[l * ls + idx + r * Ns] = v[r];
    }
    __syncthreads();
    float2* t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  return src;
}

__device__ inline size_t hmd_line_base(const FftParams& p, uint32_t l) {
  return (size_t)(l / p.line_minor) * p.major_stride +
         (size_t)(l % p.line_minor) * p.minor_stride;
}

// Pass 1: real -> complex along c.  data = real mesh (n0*n1 lines of n2
// reals), out = complex mesh.  n = n2/2.
__device__ __forceinline__ void hmd_fft_r2c_c_body(const FftParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  extern __shared__ float2 s_buf[];
  const uint32_t n = p.n, L = p.lines_per_block, ls = n + 1;
  float2* b0 = s_buf;
  float2* b1 = s_buf + L * ls;
  const uint32_t l0 = hmd_bx * L;
  const uint32_t lines = min(L, p.n_lines - l0);
  const float2* in = reinterpret_cast<const float2*>(p.data);
  for (uint32_t k = threadIdx.x; k < lines * n; k += blockDim.x) {
    const uint32_t l = k / n, m = k - l * n;
    b0[l * ls + m] = in[(size_t)l0 * n + k];   // lines are contiguous
  }
  __syncthreads();
  float2* z = hmd_fft_lds(b0, b1, n, ls, lines, p.radix, p.n_stages,
                          reinterpret_cast<const float2*>(p.twiddle), -1);
  const uint32_t nch = n + 1;
  float2* out = reinterpret_cast<float2*>(p.out);
  const float2* wr = reinterpret_cast<const float2*>(p.twiddle_r2c);
  for (uint32_t k = threadIdx.x; k < lines * nch; k += blockDim.x) {
    const uint32_t l = k / nch, m = k - l * nch;
    float2 zk = z[l * ls + (m % n)];
    float2 zc = cconj(z[l * ls + ((n - m) % n)]);
    float2 e = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y + zc.y));
    float2 o = cmul_si(make_float2(0.5f * (zk.x - zc.x), 0.5f * (zk.y - zc.y)), -1);
    out[(size_t)(l0 + l) * nch + m] = cadd(e, cmul(wr[m], o));
  }
}
HMD_KERNEL(hmd_fft_r2c_c, FftParams, 256)

// Pass 5: complex -> real along c (unnormalised).  data = complex mesh,
// out = real mesh.  n = n2/2.
__device__ __forceinline__ void hmd_fft_c2r_c_body(const FftParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  extern __shared__ float2 s_buf[];
  const uint32_t n = p.n, L = p.lines_per_block, ls = n + 1;
  float2* b0 = s_buf;
  float2* b1 = s_buf + L * ls;
  const uint32_t l0 = hmd_bx * L;
  const uint32_t lines = min(L, p.n_lines - l0);
  const uint32_t nch = n + 1;
  const float2* in = reinterpret_cast<const float2*>(p.data);
  const float2* wr = reinterpret_cast<const float2*>(p.twiddle_r2c);
  for (uint32_t k = threadIdx.x; k < lines * n; k += blockDim.x) {
    const uint32_t l = k / n, m = k - l * n;
    const float2* x = in + (size_t)(l0 + l) * nch;
    float2 xk = x[m], xc = cconj(x[n - m]);
    float2 ee = cadd(xk, xc);
    float2 oo = cmul(csub(xk, xc), cconj(wr[m]));
    b0[l * ls + m] = cadd(ee, cmul_si(oo, +1));
  }
  __syncthreads();
  float2* z = hmd_fft_lds(b0, b1, n, ls, lines, p.radix, p.n_stages,
                          reinterpret_cast<const float2*>(p.twiddle), +1);
  float2* out = reinterpret_cast<float2*>(p.out);
  for (uint32_t k = threadIdx.x; k < lines * n; k += blockDim.x) {
    const uint32_t l = k / n, m = k - l * n;
    out[(size_t)l0 * n + k] = z[l * ls + m];
  }
}
HMD_KERNEL(hmd_fft_c2r_c, FftParams, 256)

// Influence function G(k) and virial factor at mesh index (ka, kb, kc)
// (kernels/spme_recip.cu:34-114).  Evaluated in FP64: exp(-k^2/4a^2) turns
// the relative rounding of its argument into a relative error of G that
// grows with k^2/4a^2, which in f32 dominated the reciprocal energy error;
// the rest of the mesh pipeline is f32.
__device__ inline void hmd_influence(const FftParams& p, uint32_t ka, uint32_t kb,
                                     uint32_t kc, double bprod, double& G, double& vf) {
  const HmdLattice L = p.L;
  const double lx = L.lx, ly = L.ly, lz = L.lz, xy = L.xy, xz = L.xz, yz = L.yz;
  const double ilx = 1.0 / lx, ily = 1.0 / ly, ilz = 1.0 / lz;
  const double ra_x = ilx, ra_y = -xy * ilx * ily, ra_z = (xy * yz - xz * ly) * ilx * ily * ilz;
  const double rb_y = ily, rb_z = -yz * ily * ilz, rc_z = ilz;
  const double ma = (ka <= p.n0 / 2) ? (double)ka : (double)ka - (double)p.n0;
  const double mb = (kb <= p.n1 / 2) ? (double)kb : (double)kb - (double)p.n1;
  const double mc = (kc <= p.n2 / 2) ? (double)kc : (double)kc - (double)p.n2;
  const double two_pi = 6.283185307179586;
  const double kx = two_pi * (ma * ra_x);
  const double ky = two_pi * (ma * ra_y + mb * rb_y);
  const double kz = two_pi * (ma * ra_z + mb * rb_z + mc * rc_z);
  const double k2 = kx * kx + ky * ky + kz * kz;
  if (k2 == 0.0) {
    G = 0.0;
    vf = 0.0;
    return;
  }
  const double vol = lx * ly * lz;
  const double a2 = (double)p.alpha * (double)p.alpha;
  G = ((double)p.k_coulomb / vol) * (12.566370614359172 / k2) * exp(-k2 / (4.0 * a2)) * bprod;
  vf = G * (1.0 - k2 / (2.0 * a2));
}

// Influence table: (G, vf) of every complex mesh point, evaluated in FP64 and
// rounded once to f32, in the layout of the complex mesh.  The box is fixed
// for the life of a force field, so the a-axis pass reads this table instead
// of evaluating exp per point every step.
__device__ __forceinline__ void hmd_spme_gv_table_body(const FftParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  const size_t i = (size_t)hmd_bx * blockDim.x + threadIdx.x;
  const size_t total = (size_t)p.n0 * p.n1 * p.nch;
  if (i >= total) return;
  const uint32_t kc = (uint32_t)(i % p.nch);
  const uint32_t kb = (uint32_t)((i / p.nch) % p.n1);
  const uint32_t ka = (uint32_t)(i / ((size_t)p.nch * p.n1));
  double G, vf;
  hmd_influence(p, ka, kb, kc, p.bfac0[ka] * p.bfac1[kb] * p.bfac2[kc], G, vf);
  reinterpret_cast<float2*>(const_cast<float*>(p.gv))[i] = make_float2((float)G, (float)vf);
}
HMD_KERNEL(hmd_spme_gv_table, FftParams, 256)

// Passes 2-4: complex FFT along a strided axis (b or a).  With
// apply_influence the a-axis line is transformed forward, multiplied by
// G(k), and transformed back in place; per-block virial / energy partial
// sums go to partials[block] in a fixed order (deterministic).
__device__ __forceinline__ void hmd_fft_axis_body(const FftParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  extern __shared__ float2 s_buf[];
  __shared__ double s_red[2][HMD_FFT_THREADS];
  const uint32_t n = p.n, L = p.lines_per_block, ls = n + 1;
  float2* b0 = s_buf;
  float2* b1 = s_buf + L * ls;
  const uint32_t l0 = hmd_bx * L;
  const uint32_t lines = min(L, p.n_lines - l0);
  float2* data = reinterpret_cast<float2*>(p.data);
  // element k of line l sits at base(l) + k * elem_stride; thread mapping
  // (k outer, l inner) keeps consecutive threads on consecutive lines, which
  // are adjacent in memory (minor_stride = 1).
  for (uint32_t t = threadIdx.x; t < lines * n; t += blockDim.x) {
    const uint32_t k = t / lines, l = t - k * lines;
    b0[l * ls + k] = data[hmd_line_base(p, l0 + l) + (size_t)k * p.elem_stride];
  }
  __syncthreads();
  const float2* tw = reinterpret_cast<const float2*>(p.twiddle);
  float2* res = hmd_fft_lds(b0, b1, n, ls, lines, p.radix, p.n_stages, tw, p.sign);
  if (p.apply_influence) {
    double vir = 0.0, ene = 0.0;
    for (uint32_t t = threadIdx.x; t < lines * n; t += blockDim.x) {
      const uint32_t l = t / n, ka = t - l * n;
      const uint32_t gl = l0 + l;                 // line = (kb, kc)
      const uint32_t kb = gl / p.nch, kc = gl - kb * p.nch;
      double G, vf;
      hmd_influence(p, ka, kb, kc, p.bfac0[ka] * p.bfac1[kb] * p.bfac2[kc], G, vf);
      float2 x = res[l * ls + ka];
      const double hw = (kc == 0u || (p.n2 % 2u == 0u && 2u * kc == p.n2)) ? 1.0 : 2.0;
      const double r2 = (double)x.x * x.x + (double)x.y * x.y;
      vir += hw * vf * r2;
      ene += hw * G * r2;
      const float Gf = (float)G;
      res[l * ls + ka] = make_float2(Gf * x.x, Gf * x.y);
    }
    s_red[0][threadIdx.x] = vir;
    s_red[1][threadIdx.x] = ene;
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) {
        s_red[0][threadIdx.x] += s_red[0][threadIdx.x + s];
        s_red[1][threadIdx.x] += s_red[1][threadIdx.x + s];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      p.partials[2 * hmd_bx] = s_red[0][0];
      p.partials[2 * hmd_bx + 1] = s_red[1][0];
    }
    float2* other = (res == b0) ? b1 : b0;
    __syncthreads();
    res = hmd_fft_lds(res, other, n, ls, lines, p.radix, p.n_stages, tw, +1);
  }
  for (uint32_t t = threadIdx.x; t < lines * n; t += blockDim.x) {
    const uint32_t k = t / lines, l = t - k * lines;
    data[hmd_line_base(p, l0 + l) + (size_t)k * p.elem_stride] = res[l * ls + k];
  }
}
HMD_KERNEL(hmd_fft_axis, FftParams, 256)

// Fixed-order reduction of the per-block partials: out[0] = W_recip =
// 0.5 * sum hw vf |rho_hat|^2, out[1] = U_recip = 0.5 * sum hw G |rho_hat|^2
// (kernels/spme_recip.cu:376-402 with the 0.5 half-sum factor).
__device__ __forceinline__ void hmd_spme_reduce_body(const SpmeReduceParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  __shared__ double s[2][256];
  double a = 0.0, b = 0.0;
  for (uint32_t i = threadIdx.x; i < p.n_partials; i += blockDim.x) {
    a += p.partials[2 * i];
    b += p.partials[2 * i + 1];
  }
  s[0][threadIdx.x] = a;
  s[1][threadIdx.x] = b;
  __syncthreads();
  for (uint32_t st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      s[0][threadIdx.x] += s[0][threadIdx.x + st];
      s[1][threadIdx.x] += s[1][threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    p.out[0] = 0.5 * s[0][0];
    p.out[1] = 0.5 * s[1][0];
  }
}
HMD_KERNEL(hmd_spme_reduce, SpmeReduceParams, 256)

// ---------------------------------------------------------------------------
// Power-of-two line lengths n = N1 N2 (16 ... 256): compile-time two-stage
// transform, n = N2 n1 + n2 -> k = k1 + N1 k2:
//   stage 1: for each n2, an N1-point DFT over n1 in registers, times the
//            twiddle e^{s 2 pi i n2 k1 / n}, stored at N1 n2 + k1;
//   stage 2: for each k1, an N2-point DFT over n2 in registers, in place,
//            giving X[k1 + N1 k2] in natural order.
// The block works on a tile of W lines ("columns") of one slab of the mesh,
// tile[k * W + w], loaded and stored with coalesced accesses; no runtime
// radix, no integer division in the butterflies, no scratch arrays.

template <int SG>
__device__ inline float2 hmd_mul_si(float2 a) {   // a * (SG i)
  return SG < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

template <int SG>
__device__ inline void hmd_dft4(float2* v) {
  const float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  const float2 t2 = cadd(v[1], v[3]), t3 = hmd_mul_si<SG>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <int SG>
__device__ inline void hmd_dft8(float2* v) {
  float2 e[4] = {v[0], v[2], v[4], v[6]};
  float2 o[4] = {v[1], v[3], v[5], v[7]};
  hmd_dft4<SG>(e);
  hmd_dft4<SG>(o);
  const float h = 0.70710678118654752f;   // cos(pi/4)
  const float2 o1 = cmul(o[1], make_float2(h, SG * h));
  const float2 o2 = hmd_mul_si<SG>(o[2]);
  const float2 o3 = cmul(o[3], make_float2(-h, SG * h));
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

template <int SG>
__device__ inline void hmd_dft16(float2* v) {
  float2 e[8], o[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    e[k] = v[2 * k];
    o[k] = v[2 * k + 1];
  }
  hmd_dft8<SG>(e);
  hmd_dft8<SG>(o);
  // cos and sin of 2 pi k / 16
  const float C[8] = {1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
                      0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f};
  const float S[8] = {0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f,
                      1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float2 t = cmul(o[k], make_float2(C[k], SG * S[k]));
    v[k] = cadd(e[k], t);
    v[k + 8] = csub(e[k], t);
  }
}

template <int N, int SG>
__device__ inline void hmd_dft(float2* v) {
  if constexpr (N == 4) hmd_dft4<SG>(v);
  else if constexpr (N == 8) hmd_dft8<SG>(v);
  else hmd_dft16<SG>(v);
}

// In-place transform of the W columns of tile[k * S + w] (k < N1 N2, row
// stride S >= W); requires W * N2 <= 2 * blockDim.x.
// tw[m] = e^{-2 pi i m / n}.
template <int N1, int N2, int SG>
__device__ inline void hmd_fft_cols_s(float2* tile, uint32_t W, uint32_t S, const float2* tw) {
  constexpr uint32_t N = (uint32_t)(N1 * N2);
  float2 v[2][N1];
  const uint32_t ntask1 = W * (uint32_t)N2;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const uint32_t task = threadIdx.x + (uint32_t)a * blockDim.x;
    if (task < ntask1) {
      const uint32_t w = task % W, n2 = task / W;
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) v[a][n1] = tile[((uint32_t)N2 * n1 + n2) * S + w];
      hmd_dft<N1, SG>(v[a]);
#pragma unroll
      for (int k1 = 1; k1 < N1; ++k1) {
        float2 t = tw[(n2 * (uint32_t)k1) & (N - 1u)];
        if (SG > 0) t = cconj(t);
        v[a][k1] = cmul(v[a][k1], t);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const uint32_t task = threadIdx.x + (uint32_t)a * blockDim.x;
    if (task < ntask1) {
      const uint32_t w = task % W, n2 = task / W;
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) tile[((uint32_t)N1 * n2 + k1) * S + w] = v[a][k1];
    }
  }
  __syncthreads();
  const uint32_t ntask2 = W * (uint32_t)N1;
  for (uint32_t task = threadIdx.x; task < ntask2; task += blockDim.x) {
    const uint32_t w = task % W, k1 = task / W;
    float2 u[N2];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) u[n2] = tile[((uint32_t)N1 * n2 + k1) * S + w];
    hmd_dft<N2, SG>(u);
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) tile[(k1 + (uint32_t)N1 * k2) * S + w] = u[k2];
  }
  __syncthreads();
}

// The same with the rows packed (S = W).
template <int N1, int N2, int SG>
__device__ inline void hmd_fft_cols(float2* tile, uint32_t W, const float2* tw) {
  hmd_fft_cols_s<N1, N2, SG>(tile, W, W, tw);
}

// The N twiddles of a transform copied into LDS, where the butterflies read
// them (a global read per twiddle put an L2 round trip into every stage);
// visible after the caller's next __syncthreads.
template <uint32_t N>
__device__ inline const float2* hmd_lds_twiddles(float2* s_tw, const float* tw) {
  for (uint32_t k = threadIdx.x; k < N; k += blockDim.x)
    s_tw[k] = reinterpret_cast<const float2*>(tw)[k];
  return s_tw;
}

// Tile fill with the loads of HMD_FFT_BATCH elements per thread issued before
// any of them is waited for: a plain strided loop waits for each load before
// its LDS store, which serialises the memory latency (~10 round trips per
// tile).  src(t) returns element t, dst(t) its LDS index.
#ifndef HMD_FFT_BATCH
#define HMD_FFT_BATCH 8
#endif
template <typename Src, typename Dst>
__device__ inline void hmd_fill_tile(float2* tile, uint32_t total, Src src, Dst dst) {
  for (uint32_t t0 = threadIdx.x; t0 < total; t0 += HMD_FFT_BATCH * blockDim.x) {
    float2 v[HMD_FFT_BATCH];
#pragma unroll
    for (int u = 0; u < HMD_FFT_BATCH; ++u) {
      const uint32_t t = t0 + (uint32_t)u * blockDim.x;
      v[u] = (t < total) ? src(t) : make_float2(0.0f, 0.0f);
    }
#pragma unroll
    for (int u = 0; u < HMD_FFT_BATCH; ++u) {
      const uint32_t t = t0 + (uint32_t)u * blockDim.x;
      if (t < total) tile[dst(t)] = v[u];
    }
  }
}

// Passes 2-4 for power-of-two axis lengths.  Block = (outer index, column
// tile): columns are the line_minor lines of one slab that are adjacent in
// memory (the c index), elem_stride steps along the transformed axis,
// major_stride along the outer axis.  With apply_influence (a axis: outer =
// b, column = c) the line is multiplied by G(k) between the forward and the
// inverse transform, with fixed-order virial / energy block partials.
template <int N1, int N2>
__device__ inline void hmd_fft_axis_slab(const FftParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  extern __shared__ float2 s_tile[];
  __shared__ double s_red[2][HMD_FFT_THREADS];
  constexpr uint32_t N = (uint32_t)(N1 * N2);
  __shared__ float2 s_tw[N];
  const float2* tw = hmd_lds_twiddles<N>(s_tw, p.twiddle);
  const uint32_t W = p.lines_per_block;
  const uint32_t n_tiles = (p.line_minor + W - 1) / W;
  const uint32_t outer = hmd_bx / n_tiles;
  const uint32_t c0 = (hmd_bx - outer * n_tiles) * W;
  const uint32_t Wc = min(W, p.line_minor - c0);
  float2* data = reinterpret_cast<float2*>(p.data);
  const size_t base = (size_t)outer * p.major_stride + c0;
  hmd_fill_tile(
      s_tile, N * W,
      [&](uint32_t t) {
        const uint32_t k = t / W, w = t - k * W;
        return (w < Wc) ? data[base + (size_t)k * p.elem_stride + w] : make_float2(0.0f, 0.0f);
      },
      [](uint32_t t) { return t; });
  // (G, vf) of this tile from the influence table, behind the data tile
  float2* s_gv = s_tile + N * W;
  if (p.apply_influence) {
    const float2* gv = reinterpret_cast<const float2*>(p.gv);
    hmd_fill_tile(
        s_gv, N * W,
        [&](uint32_t t) {
          const uint32_t k = t / W, w = t - k * W;
          return (w < Wc) ? gv[base + (size_t)k * p.elem_stride + w] : make_float2(0.0f, 0.0f);
        },
        [](uint32_t t) { return t; });
  }
  __syncthreads();
  if (p.sign < 0)
    hmd_fft_cols<N1, N2, -1>(s_tile, W, tw);
  else
    hmd_fft_cols<N1, N2, +1>(s_tile, W, tw);
  if (p.apply_influence) {
    double vir = 0.0, ene = 0.0;
    for (uint32_t t = threadIdx.x; t < N * W; t += blockDim.x) {
      const uint32_t w = t % W;
      if (w >= Wc) continue;
      const uint32_t kc = c0 + w;
      const float2 g = s_gv[t];          // (G, vf)
      const float2 x = s_tile[t];
      const double hw = (kc == 0u || (p.n2 % 2u == 0u && 2u * kc == p.n2)) ? 1.0 : 2.0;
      const double r2 = (double)x.x * x.x + (double)x.y * x.y;
      vir += hw * (double)g.y * r2;
      ene += hw * (double)g.x * r2;
      s_tile[t] = make_float2(g.x * x.x, g.x * x.y);
    }
    s_red[0][threadIdx.x] = vir;
    s_red[1][threadIdx.x] = ene;
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) {
        s_red[0][threadIdx.x] += s_red[0][threadIdx.x + s];
        s_red[1][threadIdx.x] += s_red[1][threadIdx.x + s];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      p.partials[2 * hmd_bx] = s_red[0][0];
      p.partials[2 * hmd_bx + 1] = s_red[1][0];
    }
    hmd_fft_cols<N1, N2, +1>(s_tile, W, tw);
  }
  for (uint32_t t = threadIdx.x; t < N * W; t += blockDim.x) {
    const uint32_t k = t / W, w = t - k * W;
    if (w < Wc) data[base + (size_t)k * p.elem_stride + w] = s_tile[t];
  }
}

// Pass 1 for n2 = 2N: W contiguous real lines per block viewed as N complex
// values each, transformed, then split into the n2/2 + 1 half spectrum.
template <int N1, int N2>
__device__ inline void hmd_fft_r2c_slab(const FftParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  extern __shared__ float2 s_tile[];
  constexpr uint32_t N = (uint32_t)(N1 * N2);
  __shared__ float2 s_tw[N];
  const float2* tw = hmd_lds_twiddles<N>(s_tw, p.twiddle);
  const uint32_t W = p.lines_per_block;
  const uint32_t l0 = hmd_bx * W;
  const uint32_t Wc = min(W, p.n_lines - l0);
  const float2* in = reinterpret_cast<const float2*>(p.data);
  hmd_fill_tile(
      s_tile, N * W,
      [&](uint32_t t) {
        const uint32_t w = t / N, m = t - w * N;
        return (w < Wc) ? in[(size_t)(l0 + w) * N + m] : make_float2(0.0f, 0.0f);
      },
      [&](uint32_t t) {
        const uint32_t w = t / N, m = t - w * N;
        return m * W + w;
      });
  __syncthreads();
  __shared__ float2 s_wr[N + 1];
  const float2* wr = hmd_lds_twiddles<N + 1>(s_wr, p.twiddle_r2c);
  hmd_fft_cols<N1, N2, -1>(s_tile, W, tw);   // ends with __syncthreads
  const uint32_t nch = N + 1;
  float2* out = reinterpret_cast<float2*>(p.out);
  for (uint32_t t = threadIdx.x; t < Wc * nch; t += blockDim.x) {
    const uint32_t w = t / nch, m = t - w * nch;
    const float2 zk = s_tile[(m & (N - 1u)) * W + w];
    const float2 zc = cconj(s_tile[((N - m) & (N - 1u)) * W + w]);
    const float2 e = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y + zc.y));
    const float2 o = cmul_si(make_float2(0.5f * (zk.x - zc.x), 0.5f * (zk.y - zc.y)), -1);
    out[(size_t)(l0 + w) * nch + m] = cadd(e, cmul(wr[m], o));
  }
}

// Pass 5 for n2 = 2N (unnormalised complex -> real).
template <int N1, int N2>
__device__ inline void hmd_fft_c2r_slab(const FftParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  extern __shared__ float2 s_tile[];
  constexpr uint32_t N = (uint32_t)(N1 * N2);
  __shared__ float2 s_tw[N];
  const float2* tw = hmd_lds_twiddles<N>(s_tw, p.twiddle);
  const uint32_t W = p.lines_per_block;
  const uint32_t l0 = hmd_bx * W;
  const uint32_t Wc = min(W, p.n_lines - l0);
  const uint32_t nch = N + 1;
  const float2* in = reinterpret_cast<const float2*>(p.data);
  const float2* wr = reinterpret_cast<const float2*>(p.twiddle_r2c);
  hmd_fill_tile(
      s_tile, N * W,
      [&](uint32_t t) {
        const uint32_t w = t / N, m = t - w * N;
        float2 z = make_float2(0.0f, 0.0f);
        if (w < Wc) {
          const float2* x = in + (size_t)(l0 + w) * nch;
          const float2 xk = x[m], xc = cconj(x[N - m]);
          const float2 ee = cadd(xk, xc);
          const float2 oo = cmul(csub(xk, xc), cconj(wr[m]));
          z = cadd(ee, cmul_si(oo, +1));
        }
        return z;
      },
      [&](uint32_t t) {
        const uint32_t w = t / N, m = t - w * N;
        return m * W + w;
      });
  __syncthreads();
  hmd_fft_cols<N1, N2, +1>(s_tile, W, tw);
  float2* out = reinterpret_cast<float2*>(p.out);
  for (uint32_t t = threadIdx.x; t < Wc * N; t += blockDim.x) {
    const uint32_t w = t / N, m = t - w * N;
    out[(size_t)(l0 + w) * N + m] = s_tile[m * W + w];
  }
}

#define HMD_FFT_P2(N, N1, N2)                                                       \
  __device__ __forceinline__ void hmd_fft_axis_##N##_body(                         \
      const FftParams& p, const uint32_t hmd_bx, const uint32_t hmd_gx) {           \
    hmd_fft_axis_slab<N1, N2>(p, hmd_bx, hmd_gx);                                   \
  }                                                                                 \
  __device__ __forceinline__ void hmd_fft_r2c_##N##_body(                          \
      const FftParams& p, const uint32_t hmd_bx, const uint32_t hmd_gx) {           \
    hmd_fft_r2c_slab<N1, N2>(p, hmd_bx, hmd_gx);                                    \
  }                                                                                 \
  __device__ __forceinline__ void hmd_fft_c2r_##N##_body(                          \
      const FftParams& p, const uint32_t hmd_bx, const uint32_t hmd_gx) {           \
    hmd_fft_c2r_slab<N1, N2>(p, hmd_bx, hmd_gx);                                    \
  }                                                                                 \
  HMD_KERNEL(hmd_fft_axis_##N, FftParams, 256)                                      \
  HMD_KERNEL(hmd_fft_r2c_##N, FftParams, 256)                                       \
  HMD_KERNEL(hmd_fft_c2r_##N, FftParams, 256)
HMD_FFT_P2(16, 4, 4)
HMD_FFT_P2(32, 8, 4)
HMD_FFT_P2(64, 8, 8)
HMD_FFT_P2(128, 16, 8)
HMD_FFT_P2(256, 16, 16)

// ---------------------------------------------------------------------------
// Slab path, for meshes whose (b, c) slab of the complex mesh fits in LDS
// (128^3: 66 KiB, transposed, of gfx950's 160 KiB per workgroup): three
// kernels instead of five passes, each reading and writing the mesh once.
//   slab_fwd  (one workgroup per a): the real rows of the slab, transposed
//             into LDS as N_c = n2/2 complex values per row; the c-axis FFT
//             and the real -> half-spectrum split; the b-axis FFT of the
//             n2/2 + 1 columns; the complex slab written out;
//   slab_solve (one workgroup per b): the (a, c) plane, the a-axis FFT, the
//             influence multiply with its fixed-order virial / energy
//             partial (one per workgroup), the inverse a-axis FFT;
//   slab_inv  (one workgroup per a): inverse b-axis FFT, the half-spectrum
//             -> complex pre-twiddle, inverse c-axis FFT, real rows out.
// The arithmetic per line is that of the five-pass path (same butterflies,
// twiddles and split formulas); only the grouping of lines into workgroups
// differs.  Rows of the transposed slab are padded by one element.

// Sums of v over the block in a fixed order: wave butterflies, then the
// waves' results in wave order (the result is valid in thread 0).
__device__ inline double hmd_block_sum_fixed(double v, double* s_w) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, HMD_WAVE);
  const uint32_t wv = threadIdx.x / HMD_WAVE;
  if ((threadIdx.x & (HMD_WAVE - 1)) == 0) s_w[wv] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (uint32_t k = 0; k < blockDim.x / HMD_WAVE; ++k) t += s_w[k];
  __syncthreads();
  return t;
}

// In-place transform of R contiguous rows tile[r * S + j] (j < N1 N2).
template <int N1, int N2, int SG>
__device__ inline void hmd_fft_rows_s(float2* tile, uint32_t R, uint32_t S, const float2* tw) {
  constexpr uint32_t N = (uint32_t)(N1 * N2);
  const uint32_t ntask1 = R * (uint32_t)N2;
  for (uint32_t t0 = 0; t0 < ntask1; t0 += 2 * blockDim.x) {
    float2 v[2][N1];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const uint32_t task = t0 + threadIdx.x + (uint32_t)a * blockDim.x;
      if (task < ntask1) {
        const uint32_t r = task / (uint32_t)N2, n2 = task - r * (uint32_t)N2;
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) v[a][n1] = tile[r * S + (uint32_t)N2 * n1 + n2];
        hmd_dft<N1, SG>(v[a]);
#pragma unroll
        for (int k1 = 1; k1 < N1; ++k1) {
          float2 t = tw[(n2 * (uint32_t)k1) & (N - 1u)];
          if (SG > 0) t = cconj(t);
          v[a][k1] = cmul(v[a][k1], t);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const uint32_t task = t0 + threadIdx.x + (uint32_t)a * blockDim.x;
      if (task < ntask1) {
        const uint32_t r = task / (uint32_t)N2, n2 = task - r * (uint32_t)N2;
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) tile[r * S + (uint32_t)N1 * n2 + k1] = v[a][k1];
      }
    }
    __syncthreads();
  }
  const uint32_t ntask2 = R * (uint32_t)N1;
  for (uint32_t task = threadIdx.x; task < ntask2; task += blockDim.x) {
    const uint32_t r = task / (uint32_t)N1, k1 = task - r * (uint32_t)N1;
    float2 u[N2];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) u[n2] = tile[r * S + (uint32_t)N1 * n2 + k1];
    hmd_dft<N2, SG>(u);
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) tile[r * S + k1 + (uint32_t)N1 * k2] = u[k2];
  }
  __syncthreads();
}

// One LDS buffer T of NC + 1 rows of NB (+1 pad) points: row m holds point m
// of every c line, then, after the split, row k holds spectrum point k of
// every b line, so the b-axis transform runs along the rows.
template <int B1, int B2, int C1, int C2>
__device__ inline void hmd_fft_slab_fwd(const FftParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  extern __shared__ float2 s_lds[];
  constexpr uint32_t NB = (uint32_t)(B1 * B2), NC = (uint32_t)(C1 * C2), NCH = NC + 1;
  constexpr uint32_t TS = NB + 1;
  __shared__ float2 s_twc[NC], s_twb[NB], s_wr[NC + 1];
  const float2* twc = hmd_lds_twiddles<NC>(s_twc, p.twiddle);
  const float2* twb = hmd_lds_twiddles<NB>(s_twb, p.twiddle2);
  const float2* wr = hmd_lds_twiddles<NC + 1>(s_wr, p.twiddle_r2c);
  float2* T = s_lds;
  const uint32_t a = hmd_bx;
  const float2* in = reinterpret_cast<const float2*>(p.data) + (size_t)a * NB * NC;
  hmd_fill_tile(T, NB * NC, [&](uint32_t t) { return in[t]; },
                [&](uint32_t t) {
                  const uint32_t b = t / NC, m = t - b * NC;
                  return m * TS + b;
                });
  __syncthreads();
  hmd_fft_cols_s<C1, C2, -1>(T, NB, TS, twc);
  // real -> half spectrum in place: pair j produces points j and NC - j (and
  // point NC from j = 0) from rows j and NC - j only
  for (uint32_t t = threadIdx.x; t < (NC / 2 + 1) * NB; t += blockDim.x) {
    const uint32_t j = t / NB, b = t - j * NB;
    const uint32_t jp = (NC - j) & (NC - 1u);           // row of the partner point
    const float2 zj = T[j * TS + b], zp = T[jp * TS + b];
    const float2 zc = cconj(zp);
    const float2 e = make_float2(0.5f * (zj.x + zc.x), 0.5f * (zj.y + zc.y));
    const float2 o = cmul_si(make_float2(0.5f * (zj.x - zc.x), 0.5f * (zj.y - zc.y)), -1);
    const float2 xj = cadd(e, cmul(wr[j], o));
    // point NC - j: zk = zp, zc = conj(zj)
    const float2 zc2 = cconj(zj);
    const float2 e2 = make_float2(0.5f * (zp.x + zc2.x), 0.5f * (zp.y + zc2.y));
    const float2 o2 = cmul_si(make_float2(0.5f * (zp.x - zc2.x), 0.5f * (zp.y - zc2.y)), -1);
    const float2 xp = cadd(e2, cmul(wr[NC - j], o2));
    T[j * TS + b] = xj;
    if (j != NC / 2) T[(NC - j) * TS + b] = xp;
  }
  __syncthreads();
  hmd_fft_rows_s<B1, B2, -1>(T, NCH, TS, twb);
  float2* out = reinterpret_cast<float2*>(p.out) + (size_t)a * NB * NCH;
  for (uint32_t t = threadIdx.x; t < NB * NCH; t += blockDim.x) {
    const uint32_t b = t / NCH, k = t - b * NCH;
    out[t] = T[k * TS + b];
  }
}

template <int A1, int A2>
__device__ inline void hmd_fft_slab_solve(const FftParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  extern __shared__ float2 s_lds[];
  __shared__ double s_w[2][1024 / HMD_WAVE];
  constexpr uint32_t NA = (uint32_t)(A1 * A2);
  __shared__ float2 s_twa[NA];
  const float2* tw = hmd_lds_twiddles<NA>(s_twa, p.twiddle);
  const uint32_t nch = p.nch, n1 = p.n1, b = hmd_bx;
  float2* Al = s_lds;              // [NA][nch]: the (a, c) plane at this b
  float2* mesh = reinterpret_cast<float2*>(p.data);
  hmd_fill_tile(Al, NA * nch,
                [&](uint32_t t) {
                  const uint32_t a = t / nch, k = t - a * nch;
                  return mesh[((size_t)a * n1 + b) * nch + k];
                },
                [](uint32_t t) { return t; });
  __syncthreads();
  hmd_fft_cols_s<A1, A2, -1>(Al, nch, nch, tw);
  const float2* gv = reinterpret_cast<const float2*>(p.gv);
  double vir = 0.0, ene = 0.0;
  // (G, vf) read HMD_FFT_BATCH at a time per thread, every load in flight
  // before the first is used
  const uint32_t total = NA * nch;
  for (uint32_t t0 = threadIdx.x; t0 < total; t0 += HMD_FFT_BATCH * blockDim.x) {
    float2 g[HMD_FFT_BATCH];
#pragma unroll
    for (int u = 0; u < HMD_FFT_BATCH; ++u) {
      const uint32_t t = t0 + (uint32_t)u * blockDim.x;
      const uint32_t a = t / nch, k = t - a * nch;
      g[u] = t < total ? gv[((size_t)a * n1 + b) * nch + k] : make_float2(0.0f, 0.0f);
    }
#pragma unroll
    for (int u = 0; u < HMD_FFT_BATCH; ++u) {
      const uint32_t t = t0 + (uint32_t)u * blockDim.x;
      if (t >= total) continue;
      const uint32_t k = t % nch;
      const float2 x = Al[t];
      const double hw = (k == 0u || (p.n2 % 2u == 0u && 2u * k == p.n2)) ? 1.0 : 2.0;
      const double r2 = (double)x.x * x.x + (double)x.y * x.y;
      vir += hw * (double)g[u].y * r2;
      ene += hw * (double)g[u].x * r2;
      Al[t] = make_float2(g[u].x * x.x, g[u].x * x.y);
    }
  }
  const double vs = hmd_block_sum_fixed(vir, s_w[0]);
  const double es = hmd_block_sum_fixed(ene, s_w[1]);
  if (threadIdx.x == 0) {
    p.partials[2 * b] = vs;
    p.partials[2 * b + 1] = es;
  }
  hmd_fft_cols_s<A1, A2, +1>(Al, nch, nch, tw);
  for (uint32_t t = threadIdx.x; t < NA * nch; t += blockDim.x) {
    const uint32_t a = t / nch, k = t - a * nch;
    mesh[((size_t)a * n1 + b) * nch + k] = Al[t];
  }
}

template <int B1, int B2, int C1, int C2>
__device__ inline void hmd_fft_slab_inv(const FftParams& p, const uint32_t hmd_bx,
    const uint32_t hmd_gx) {
  extern __shared__ float2 s_lds[];
  constexpr uint32_t NB = (uint32_t)(B1 * B2), NC = (uint32_t)(C1 * C2), NCH = NC + 1;
  constexpr uint32_t TS = NB + 1;
  __shared__ float2 s_twc[NC], s_twb[NB], s_wr[NC + 1];
  const float2* twc = hmd_lds_twiddles<NC>(s_twc, p.twiddle);
  const float2* twb = hmd_lds_twiddles<NB>(s_twb, p.twiddle2);
  const float2* wr = hmd_lds_twiddles<NC + 1>(s_wr, p.twiddle_r2c);
  float2* T = s_lds;
  const uint32_t a = hmd_bx;
  const float2* in = reinterpret_cast<const float2*>(p.data) + (size_t)a * NB * NCH;
  hmd_fill_tile(T, NB * NCH, [&](uint32_t t) { return in[t]; },
                [&](uint32_t t) {
                  const uint32_t b = t / NCH, k = t - b * NCH;
                  return k * TS + b;
                });
  __syncthreads();
  hmd_fft_rows_s<B1, B2, +1>(T, NCH, TS, twb);
  // half spectrum -> complex c lines in place: pair j produces points j and
  // NC - j from rows j and NC - j (row NC is read by j = 0)
  for (uint32_t t = threadIdx.x; t < (NC / 2 + 1) * NB; t += blockDim.x) {
    const uint32_t j = t / NB, b = t - j * NB;
    const float2 xj = T[j * TS + b], xp = T[(NC - j) * TS + b];
    const float2 xc = cconj(xp);
    const float2 ee = cadd(xj, xc);
    const float2 oo = cmul(csub(xj, xc), cconj(wr[j]));
    const float2 zj = cadd(ee, cmul_si(oo, +1));
    // point NC - j: xk = xp, xc = conj(xj)
    const float2 xc2 = cconj(xj);
    const float2 ee2 = cadd(xp, xc2);
    const float2 oo2 = cmul(csub(xp, xc2), cconj(wr[NC - j]));
    const float2 zp = cadd(ee2, cmul_si(oo2, +1));
    T[j * TS + b] = zj;
    if (j != 0 && j != NC / 2) T[(NC - j) * TS + b] = zp;
  }
  __syncthreads();
  hmd_fft_cols_s<C1, C2, +1>(T, NB, TS, twc);
  float2* out = reinterpret_cast<float2*>(p.out) + (size_t)a * NB * NC;
  for (uint32_t t = threadIdx.x; t < NB * NC; t += blockDim.x) {
    const uint32_t b = t / NC, m = t - b * NC;
    out[t] = T[m * TS + b];
  }
}

#define HMD_FFT_SLAB(TAG, A1, A2, B1, B2, C1, C2)                                   \
  __device__ __forceinline__ void hmd_fft_slab_fwd_##TAG##_body(                   \
      const FftParams& p, const uint32_t hmd_bx, const uint32_t hmd_gx) {           \
    hmd_fft_slab_fwd<B1, B2, C1, C2>(p, hmd_bx, hmd_gx);                            \
  }                                                                                 \
  __device__ __forceinline__ void hmd_fft_slab_solve_##TAG##_body(                 \
      const FftParams& p, const uint32_t hmd_bx, const uint32_t hmd_gx) {           \
    hmd_fft_slab_solve<A1, A2>(p, hmd_bx, hmd_gx);                                  \
  }                                                                                 \
  __device__ __forceinline__ void hmd_fft_slab_inv_##TAG##_body(                   \
      const FftParams& p, const uint32_t hmd_bx, const uint32_t hmd_gx) {           \
    hmd_fft_slab_inv<B1, B2, C1, C2>(p, hmd_bx, hmd_gx);                            \
  }                                                                                 \
  HMD_KERNEL(hmd_fft_slab_fwd_##TAG, FftParams, 1024)                               \
  HMD_KERNEL(hmd_fft_slab_solve_##TAG, FftParams, 1024)                             \
  HMD_KERNEL(hmd_fft_slab_inv_##TAG, FftParams, 1024)
HMD_FFT_SLAB(128, 16, 8, 16, 8, 8, 8)
HMD_FFT_SLAB(64, 8, 8, 8, 8, 8, 4)
