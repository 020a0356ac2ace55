"""Pair-term fragments for the run-time composer (hmd_forcefield_set_pair_fragment)."""

# A Morse pair term between every pair of atoms within the cutoff:
#   E = D (1 - exp(-a (r - r0)))^2 - D,  F/r = -dE/dr / r
# params = (D, a, r0).
MORSE = r"""
struct MorsePair {
  __device__ void operator()(const float* prm, uint32_t ti, uint32_t tj, float r2,
                             float inv_r, float& f_over_r, float& energy) const {
    const float D = prm[0], a = prm[1], r0 = prm[2];
    const float r = r2 * inv_r;
    const float e = expf(-a * (r - r0));
    const float om = 1.0f - e;
    energy = D * om * om - D;
    f_over_r = -2.0f * D * a * om * e * inv_r;
  }
};
"""

# The zero term: the composed kernels must reproduce the built-in ones.
ZERO = r"""
struct ZeroPair {
  __device__ void operator()(const float*, uint32_t, uint32_t, float, float,
                             float& f_over_r, float& energy) const {
    f_over_r = 0.0f;
    energy = 0.0f;
  }
};
"""
