/*
 * heddle_oracle.c -- CPU restatement of HeddleMD's nonbonded hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * Nothing in the product library links it, calls it or falls back to it.
 *
 * Every routine restates one piece of the reference algorithm (Rust host +
 * CUDA kernels of taylor-a-barnes/heddlemd) and cites the reference file:line
 * it follows.  Arithmetic is FP64 except where a routine exists to reproduce
 * an exact FP32 decision (the neighbour pair-set distance test), which is
 * compiled with -ffp-contract=off so that every f32 operation rounds exactly
 * once, as on the GPU.
 *
 * Units are Hartree atomic units throughout (Bohr, Hartree, electron mass,
 * elementary charge, hbar/E_h), as in the reference (src/units/).  The
 * Coulomb prefactor k_C is a parameter (1 in atomic units).
 *
 * Lattice layout (all routines): L[6] = {lx, ly, lz, xy, xz, yz}, the
 * lower-triangular cell used by kernels/pbc.cuh:17-43.
 *
 * Pinning (what checks each routine; see tests/):
 *   orc_min_image_f32 / orc_pair_set_f32  -- brute force is the definition;
 *       tests also check translation invariance and a hand-built KAT.
 *   orc_pair_forces                       -- LJ closed-form KAT (r=1.5,
 *       sigma=1.3, eps=2) from tests/forces_framework.rs:590-627; finite
 *       differences of the oracle energy; Newton-3 (sum F = 0).
 *   orc_ewald_recip_direct                -- finite differences; two- and
 *       four-charge fixtures vs SPME (tests/spme_pipeline.rs:101-189).
 *   orc_spme_*                            -- single-particle spread vs the
 *       B-spline closed form (tests/spme_pipeline.rs:320-382); sum(rho) =
 *       sum(q); G[0] = 0; FFT vs numpy; energy vs direct Ewald within 5e-3.
 *   orc_settle_*                          -- rigidity, COM velocity, zero
 *       bond-velocity invariants (tests/settle.rs).
 *   orc_philox4x32_10                     -- Random123 published KATs.
 *   orc_vv_kick_drift_f32 / orc_vv_kick_f32 -- the exact-value cases of
 *       tests/integration.rs and tests/integration_lossless.rs; constant-
 *       force kinematics in closed form (the compensated sum to f64
 *       rounding); wrap image counts; a perturbed variant (unfused drift)
 *       must fail the fused-rounding case.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

/* ------------------------------------------------------------------ */
/* Periodic boundaries                                                 */
/* ------------------------------------------------------------------ */

/* f32 minimum image, operation for operation as kernels/pbc.cuh:17-43
 * (triclinic_cart_to_frac + triclinic_min_image).  Compiled without FMA
 * contraction so that it is the exact f32 function the GPU pair-set test
 * evaluates. */
static void mi_f32(float *dx, float *dy, float *dz, const float *L)
{
    float sc = *dz / L[2];
    float sb = (*dy - sc * L[5]) / L[1];
    float sa = (*dx - sb * L[3] - sc * L[4]) / L[0];
    float ka = floorf(sa + 0.5f);
    float kb = floorf(sb + 0.5f);
    float kc = floorf(sc + 0.5f);
    *dx -= ka * L[0] + kb * L[3] + kc * L[4];
    *dy -= kb * L[1] + kc * L[5];
    *dz -= kc * L[2];
}

/* FP64 minimum image (same construction, kernels/pbc.cuh:30-43). */
static void mi_f64(double *dx, double *dy, double *dz, const double *L)
{
    double sc = *dz / L[2];
    double sb = (*dy - sc * L[5]) / L[1];
    double sa = (*dx - sb * L[3] - sc * L[4]) / L[0];
    double ka = floor(sa + 0.5), kb = floor(sb + 0.5), kc = floor(sc + 0.5);
    *dx -= ka * L[0] + kb * L[3] + kc * L[4];
    *dy -= kb * L[1] + kc * L[5];
    *dz -= kc * L[2];
}

/* Exact f32 squared min-image distance of one pair. */
float orc_r2_f32(const float *pi, const float *pj, const float *L)
{
    float dx = pi[0] - pj[0];
    float dy = pi[1] - pj[1];
    float dz = pi[2] - pj[2];
    mi_f32(&dx, &dy, &dz, L);
    return dx * dx + dy * dy + dz * dz;
}

/* orc_r2_f32 of n pairs (a_i, b_i), posq layout (f32[4n] each): the skin
 * trigger's displacement test (kernels/neighbor.cu:50-76) per atom. */
void orc_r2_f32_many(int n, const float *a, const float *b, const float *L, float *out)
{
    for (int i = 0; i < n; ++i) out[i] = orc_r2_f32(a + 4 * i, b + 4 * i, L);
}

/* Wrap positions into the primary image (kernels/pbc.cuh:74-92), f64. */
void orc_wrap_f64(int n, double *pos, const double *L)
{
    for (int i = 0; i < n; ++i) {
        double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        mi_f64(&x, &y, &z, L);
        pos[3 * i] = x; pos[3 * i + 1] = y; pos[3 * i + 2] = z;
    }
}

/* ------------------------------------------------------------------ */
/* Neighbour pair set (a1-a7)                                          */
/* ------------------------------------------------------------------ */

/* Brute-force canonical pair set: every unordered pair (i<j) whose f32
 * min-image squared distance is <= r_search_sq, with the distance test of
 * find_blocks_with_interactions (kernels/neighbor.cu:536-553).
 * pos is f32[4n] (posq layout; w ignored).  Pairs are written as
 * (i, j) u32 in ascending (i, j) order; returns the total count even when
 * it exceeds cap (only the first cap pairs are stored). */
int64_t orc_pair_set_f32(int n, const float *posq, const float *L,
                         float r_search_sq, uint32_t *pairs, int64_t cap)
{
    int64_t count = 0;
    for (int i = 0; i < n; ++i) {
        const float *pi = posq + 4 * i;
        for (int j = i + 1; j < n; ++j) {
            float r2 = orc_r2_f32(pi, posq + 4 * j, L);
            if (r2 <= r_search_sq) {
                if (count < cap) {
                    pairs[2 * count] = (uint32_t)i;
                    pairs[2 * count + 1] = (uint32_t)j;
                }
                ++count;
            }
        }
    }
    return count;
}

/* Cell-list form of orc_pair_set_f32 for large orthorhombic boxes: cells
 * at least 1.001 r_search wide (>= 3 per axis, else the brute force), every
 * candidate pair from the 27 neighbouring cells given the same exact f32
 * test, output in the same ascending (i, j) order.  Pinned against the
 * brute force in the tests. */
static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

int64_t orc_pair_set_f32_cells(int n, const float *posq, const float *L,
                               float r_search_sq, uint32_t *pairs, int64_t cap)
{
    const double rs = sqrt((double)r_search_sq) * 1.001;
    int nc[3];
    for (int d = 0; d < 3; ++d) {
        nc[d] = (int)floor((double)L[d] / rs);
        if (nc[d] < 3 || L[3] != 0.0f || L[4] != 0.0f || L[5] != 0.0f)
            return orc_pair_set_f32(n, posq, L, r_search_sq, pairs, cap);
    }
    const size_t ncell = (size_t)nc[0] * nc[1] * nc[2];
    int *cell = malloc(sizeof(int) * n);
    int *start = calloc(ncell + 1, sizeof(int));
    int *members = malloc(sizeof(int) * n);
    for (int i = 0; i < n; ++i) {
        int ci[3];
        for (int d = 0; d < 3; ++d) {
            double s = (double)posq[4 * i + d] / (double)L[d] + 0.5;
            s -= floor(s);
            int k = (int)(s * nc[d]);
            ci[d] = k >= nc[d] ? nc[d] - 1 : k;
        }
        cell[i] = (ci[0] * nc[1] + ci[1]) * nc[2] + ci[2];
        start[cell[i] + 1]++;
    }
    for (size_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
    int *fill = calloc(ncell, sizeof(int));
    for (int i = 0; i < n; ++i) members[start[cell[i]] + fill[cell[i]]++] = i;
    free(fill);
    /* pass 1: count per i; pass 2: write at the prefix offsets, j ascending */
    int64_t *cnt = calloc((size_t)n + 1, sizeof(int64_t));
    for (int pass = 0; pass < 2; ++pass) {
        #pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < n; ++i) {
            const float *pi = posq + 4 * i;
            int c = cell[i];
            int ca = c / (nc[1] * nc[2]), cb = (c / nc[2]) % nc[1], cc = c % nc[2];
            int nbr[27], nn = 0;
            for (int da = -1; da <= 1; ++da)
            for (int db = -1; db <= 1; ++db)
            for (int dc = -1; dc <= 1; ++dc) {
                int a = (ca + da + nc[0]) % nc[0], b = (cb + db + nc[1]) % nc[1];
                int k = (cc + dc + nc[2]) % nc[2];
                int id = (a * nc[1] + b) * nc[2] + k, dup = 0;
                for (int m = 0; m < nn; ++m) dup |= nbr[m] == id;
                if (!dup) nbr[nn++] = id;
            }
            /* candidates j > i, ascending */
            int cap_j = 0;
            for (int m = 0; m < nn; ++m) cap_j += start[nbr[m] + 1] - start[nbr[m]];
            int *js = malloc(sizeof(int) * (cap_j + 1));
            int k = 0;
            for (int m = 0; m < nn; ++m)
                for (int t = start[nbr[m]]; t < start[nbr[m] + 1]; ++t)
                    if (members[t] > i) js[k++] = members[t];
            qsort(js, (size_t)k, sizeof(int), cmp_int);
            int64_t o = pass ? cnt[i] : 0;
            for (int t = 0; t < k; ++t) {
                float r2 = orc_r2_f32(pi, posq + 4 * js[t], L);
                if (r2 <= r_search_sq) {
                    if (pass && o < cap) {
                        pairs[2 * o] = (uint32_t)i;
                        pairs[2 * o + 1] = (uint32_t)js[t];
                    }
                    ++o;
                }
            }
            if (!pass) cnt[i + 1] = o;
            free(js);
        }
        if (!pass)
            for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    }
    int64_t total = cnt[n];
    free(cnt); free(cell); free(start); free(members);
    return total;
}

/* ------------------------------------------------------------------ */
/* Pair forces: LJ 12-6 (+C1 switch) + real-space Ewald (a8)           */
/* ------------------------------------------------------------------ */

/* Abramowitz & Stegun 7.1.26 erfc, the f32 formula of the reference's
 * SPME-real fragment (src/forces/spme.rs:738-749), evaluated in f64. */
static double erfc_as(double x)
{
    double t = 1.0 / (1.0 + 0.3275911 * x);
    return (0.254829592 + (-0.284496736 + (1.421413741 + (-1.453152027
             + 1.061405429 * t) * t) * t) * t) * t * exp(-x * x);
}

typedef struct {
    /* particles */
    int n;
    const double *pos;        /* [3n] */
    const double *q;          /* [n]  */
    const int *type;          /* [n]  */
    const double *L;          /* [6]  */
    /* LJ table, row-major [nt*nt] */
    int n_types;
    const double *lj_sigma;
    const double *lj_epsilon;
    const double *lj_cutoff;
    const double *lj_switch;
    int lj_on;
    /* SPME real-space */
    int coul_on;
    double k_coulomb, alpha, r_cut_real;
    int erfc_mode;            /* 0 = libm erfc, 1 = A&S 7.1.26 */
    /* exclusions: CSR over atoms, scales per fragment */
    const int *excl_offsets;  /* [n+1] or NULL */
    const int *excl_partners;
    const double *excl_lj_scale;
    const double *excl_coul_scale;
    double max_cutoff;
} orc_pair_params;

static double excl_lookup(const int *off, const int *par, const double *sc,
                          int i, int j)
{
    if (!off) return 1.0;
    for (int m = off[i]; m < off[i + 1]; ++m)
        if (par[m] == j) return sc[m];
    return 1.0;
}

/* One pair's (factor, energy, virial) summed over the active fragments,
 * each fragment's contribution multiplied by its own exclusion scale, as the
 * composer does (src/forces/jit_composed.rs:941-990).  factor is F/r so the
 * force on i is factor * (r_i - r_j). */
static void pair_eval(const orc_pair_params *P, int i, int j, double r2,
                      double *factor, double *energy, double *virial)
{
    double f = 0.0, e = 0.0, w = 0.0;
    double r = sqrt(r2);
    double inv_r2 = 1.0 / r2;
    if (P->lj_on) {
        /* src/forces/lj.rs:395-445 */
        int p = P->type[i] * P->n_types + P->type[j];
        double sig = P->lj_sigma[p], eps = P->lj_epsilon[p];
        double rc = P->lj_cutoff[p], rs = P->lj_switch[p];
        if (r2 <= rc * rc) {
            double sr2 = sig * sig * inv_r2;
            double sr6 = sr2 * sr2 * sr2;
            double sr12 = sr6 * sr6;
            double fl = 24.0 * eps * inv_r2 * (2.0 * sr12 - sr6);
            double el = 4.0 * eps * (sr12 - sr6);
            double rs2 = rs * rs, rc2 = rc * rc, delta = rc2 - rs2;
            if (delta > 0.0 && r2 > rs2) {
                double tau = (r2 - rs2) / delta;
                double omt = 1.0 - tau;
                double s = omt * omt * (1.0 + 2.0 * tau);
                double chain = 12.0 * tau * omt / delta;
                fl = s * fl + chain * el;
                el = s * el;
            }
            double sc = excl_lookup(P->excl_offsets, P->excl_partners,
                                    P->excl_lj_scale, i, j);
            f += sc * fl; e += sc * el; w += sc * fl * r2;
        }
    }
    if (P->coul_on && r2 <= P->r_cut_real * P->r_cut_real) {
        /* src/forces/spme.rs:733-755 */
        double qq = P->q[i] * P->q[j];
        double ar = P->alpha * r;
        double g = exp(-ar * ar);
        double ec = P->erfc_mode ? erfc_as(ar) : erfc(ar);
        double el = P->k_coulomb * qq * ec / r;
        double fl = P->k_coulomb * qq * inv_r2
                    * (ec / r + 2.0 * P->alpha / sqrt(ORC_PI) * g);
        double sc = excl_lookup(P->excl_offsets, P->excl_partners,
                                P->excl_coul_scale, i, j);
        f += sc * fl; e += sc * el; w += sc * fl * r2;
    }
    *factor = f; *energy = e; *virial = w;
}

/* Brute-force O(N^2) pair forces over every unordered pair within the max
 * cutoff (FP64 min-image).  Outputs are overwritten: F[3n], E[n], W[n] with
 * the half/half energy and virial split of jit_composed.rs:1554-1562. */
void orc_pair_forces(const orc_pair_params *P, double *F, double *E, double *W)
{
    int n = P->n;
    memset(F, 0, sizeof(double) * 3 * n);
    memset(E, 0, sizeof(double) * n);
    memset(W, 0, sizeof(double) * n);
    double rc2 = P->max_cutoff * P->max_cutoff;
    for (int i = 0; i < n; ++i) {
        for (int j = i + 1; j < n; ++j) {
            double dx = P->pos[3 * i] - P->pos[3 * j];
            double dy = P->pos[3 * i + 1] - P->pos[3 * j + 1];
            double dz = P->pos[3 * i + 2] - P->pos[3 * j + 2];
            mi_f64(&dx, &dy, &dz, P->L);
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > rc2) continue;
            double f, e, w;
            pair_eval(P, i, j, r2, &f, &e, &w);
            F[3 * i] += f * dx; F[3 * i + 1] += f * dy; F[3 * i + 2] += f * dz;
            F[3 * j] -= f * dx; F[3 * j + 1] -= f * dy; F[3 * j + 2] -= f * dz;
            E[i] += 0.5 * e; E[j] += 0.5 * e;
            W[i] += 0.5 * w; W[j] += 0.5 * w;
        }
    }
}

/* Cell-list variant of orc_pair_forces for large systems (orthorhombic
 * boxes with >= 3 cells per axis).  Same arithmetic per pair; each atom
 * accumulates only its own force (both directions are visited), so the
 * loop parallelises over i without atomics. */
void orc_pair_forces_cells(const orc_pair_params *P, double *F, double *E,
                           double *W)
{
    int n = P->n;
    const double *L = P->L;
    double rc = P->max_cutoff, rc2 = rc * rc;
    int nc[3];
    for (int d = 0; d < 3; ++d) {
        nc[d] = (int)floor(L[d] / rc);
        if (nc[d] < 3) { orc_pair_forces(P, F, E, W); return; }
    }
    int ncell = nc[0] * nc[1] * nc[2];
    int *cell = malloc(sizeof(int) * n);
    int *head = malloc(sizeof(int) * ncell);
    int *next = malloc(sizeof(int) * n);
    for (int c = 0; c < ncell; ++c) head[c] = -1;
    for (int i = n - 1; i >= 0; --i) {
        int ci[3];
        for (int d = 0; d < 3; ++d) {
            double s = P->pos[3 * i + d] / L[d] + 0.5;
            s -= floor(s);
            int k = (int)(s * nc[d]);
            if (k >= nc[d]) k = nc[d] - 1;
            ci[d] = k;
        }
        int c = (ci[0] * nc[1] + ci[1]) * nc[2] + ci[2];
        cell[i] = c;
        next[i] = head[c];
        head[c] = i;
    }
    #pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        int c = cell[i];
        int ca = c / (nc[1] * nc[2]), cb = (c / nc[2]) % nc[1], cc = c % nc[2];
        double fx = 0, fy = 0, fz = 0, ee = 0, ww = 0;
        for (int da = -1; da <= 1; ++da)
        for (int db = -1; db <= 1; ++db)
        for (int dc = -1; dc <= 1; ++dc) {
            int a = (ca + da + nc[0]) % nc[0];
            int b = (cb + db + nc[1]) % nc[1];
            int k = (cc + dc + nc[2]) % nc[2];
            for (int j = head[(a * nc[1] + b) * nc[2] + k]; j >= 0; j = next[j]) {
                if (j == i) continue;
                double dx = P->pos[3 * i] - P->pos[3 * j];
                double dy = P->pos[3 * i + 1] - P->pos[3 * j + 1];
                double dz = P->pos[3 * i + 2] - P->pos[3 * j + 2];
                mi_f64(&dx, &dy, &dz, L);
                double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 > rc2) continue;
                double f, e, w;
                pair_eval(P, i, j, r2, &f, &e, &w);
                fx += f * dx; fy += f * dy; fz += f * dz;
                ee += 0.5 * e; ww += 0.5 * w;
            }
        }
        F[3 * i] = fx; F[3 * i + 1] = fy; F[3 * i + 2] = fz;
        E[i] = ee; W[i] = ww;
    }
    free(cell); free(head); free(next);
}

/* ------------------------------------------------------------------ */
/* Reciprocal lattice                                                  */
/* ------------------------------------------------------------------ */

/* Rows of H^{-T} for the lower-triangular cell
 * (kernels/spme_spread_gather.cuh:128-148). */
static void recip_rows(const double *L, double b[3][3])
{
    double lx = L[0], ly = L[1], lz = L[2], xy = L[3], xz = L[4], yz = L[5];
    b[0][0] = 1.0 / lx;
    b[0][1] = -xy / (lx * ly);
    b[0][2] = (xy * yz - xz * ly) / (lx * ly * lz);
    b[1][0] = 0.0; b[1][1] = 1.0 / ly; b[1][2] = -yz / (ly * lz);
    b[2][0] = 0.0; b[2][1] = 0.0; b[2][2] = 1.0 / lz;
}

/* ------------------------------------------------------------------ */
/* Direct (explicit) Ewald reciprocal sum                              */
/* ------------------------------------------------------------------ */

/* Explicit Ewald reciprocal energy and forces truncated to |m_d| <= m_max
 * (the test oracle of tests/spme_pipeline.rs:46-83, extended to forces by
 * differentiating the structure factor).  E_i is the per-atom share
 * 0.5*q_i*phi_i (without self energy).  Returns U_recip. */
double orc_ewald_recip_direct(int n, const double *pos, const double *q,
                              const double *L, double alpha, double k_coulomb,
                              int m_max, double *F, double *E)
{
    double b[3][3];
    recip_rows(L, b);
    double vol = L[0] * L[1] * L[2];
    double inv4a2 = 1.0 / (4.0 * alpha * alpha);
    memset(F, 0, sizeof(double) * 3 * n);
    memset(E, 0, sizeof(double) * n);
    double *cs = malloc(sizeof(double) * 2 * n);
    double total = 0.0;
    for (int ma = -m_max; ma <= m_max; ++ma)
    for (int mb = -m_max; mb <= m_max; ++mb)
    for (int mc = -m_max; mc <= m_max; ++mc) {
        if (ma == 0 && mb == 0 && mc == 0) continue;
        double kv[3];
        for (int d = 0; d < 3; ++d)
            kv[d] = 2.0 * ORC_PI * (ma * b[0][d] + mb * b[1][d] + mc * b[2][d]);
        double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        double A = (4.0 * ORC_PI / k2) * exp(-k2 * inv4a2);
        double C = 0.0, D = 0.0;
        for (int i = 0; i < n; ++i) {
            double th = kv[0] * pos[3 * i] + kv[1] * pos[3 * i + 1] + kv[2] * pos[3 * i + 2];
            double c = cos(th), s = sin(th);
            cs[2 * i] = c; cs[2 * i + 1] = s;
            C += q[i] * c; D += q[i] * s;
        }
        total += A * (C * C + D * D);
        double pref = k_coulomb / vol * A;
        for (int i = 0; i < n; ++i) {
            double c = cs[2 * i], s = cs[2 * i + 1];
            double g = pref * q[i] * (C * s - D * c);
            F[3 * i] += g * kv[0]; F[3 * i + 1] += g * kv[1]; F[3 * i + 2] += g * kv[2];
            E[i] += 0.5 * pref * q[i] * (C * c + D * s);
        }
    }
    free(cs);
    return k_coulomb / (2.0 * vol) * total;
}

/* ------------------------------------------------------------------ */
/* SPME (a11-a18), FP64 restatement                                    */
/* ------------------------------------------------------------------ */

/* Cardinal B-spline M_p(x) by Cox-de Boor (src/forces/spme.rs:566-585). */
double orc_bspline(int p, double x)
{
    double v[32];
    for (int i = 0; i < p; ++i) {
        double xi = x - i;
        v[i] = (xi >= 0.0 && xi < 1.0) ? 1.0 : 0.0;
    }
    for (int order = 2; order <= p; ++order) {
        double inv = 1.0 / (order - 1.0);
        for (int i = 0; i < p - order + 1; ++i) {
            double xi = x - i;
            v[i] = xi * inv * v[i] + (order - xi) * inv * v[i + 1];
        }
    }
    return v[0];
}

/* M_p(x) and M_p'(x) = M_{p-1}(x) - M_{p-1}(x-1)
 * (kernels/spme_spread_gather.cuh:58-86). */
static void bspline_wd(int p, double x, double *w, double *dw)
{
    *w = orc_bspline(p, x);
    *dw = orc_bspline(p - 1, x) - orc_bspline(p - 1, x - 1.0);
}

/* |b(k)|^2 per axis with the near-zero neighbour averaging of
 * src/forces/spme.rs:524-562.  Output f64[n]. */
void orc_b_factors(int n, int p, double *out)
{
    double *mod = malloc(sizeof(double) * n);
    for (int k = 0; k < n; ++k) {
        double th = 2.0 * ORC_PI * k / n, re = 0.0, im = 0.0;
        for (int j = 0; j < p - 1; ++j) {
            double m = orc_bspline(p, j + 1.0);
            re += m * cos(th * j);
            im += m * sin(th * j);
        }
        mod[k] = re * re + im * im;
    }
    for (int k = 0; k < n; ++k)
        if (mod[k] < 1.0e-7)
            mod[k] = 0.5 * (mod[(k + n - 1) % n] + mod[(k + 1) % n]);
    for (int k = 0; k < n; ++k) out[k] = mod[k] > 0.0 ? 1.0 / mod[k] : 0.0;
    free(mod);
}

/* Fractional setup of one particle: u = (s + 1/2) n, g = floor(u), t = u - g
 * (kernels/spme_spread_gather.cuh:90-126), f64. */
static void spme_setup(const double *p, const double *L, const int *nk,
                       double t[3], int g[3])
{
    double x = p[0], y = p[1], z = p[2];
    mi_f64(&x, &y, &z, L);                  /* wrap into primary image */
    double sc = z / L[2];
    double sb = (y - sc * L[5]) / L[1];
    double sa = (x - sb * L[3] - sc * L[4]) / L[0];
    double s[3] = {sa, sb, sc};
    for (int d = 0; d < 3; ++d) {
        double u = (s[d] + 0.5) * nk[d];
        double f = floor(u);
        t[d] = u - f;
        int gi = (int)f % nk[d];
        if (gi < 0) gi += nk[d];
        g[d] = gi;
    }
}

/* Charge spread: rho[(g_a-d_a, g_b-d_b, g_c-d_c) mod n] +=
 * q*M_p(d_a+t_a)*M_p(d_b+t_b)*M_p(d_c+t_c)
 * (kernels/spme_spread_gather.cuh:162-221), exact f64 sum. */
void orc_spme_spread(int n, const double *pos, const double *q, const double *L,
                     const int *nk, int p, double *rho)
{
    int M = nk[0] * nk[1] * nk[2];
    memset(rho, 0, sizeof(double) * M);
    for (int i = 0; i < n; ++i) {
        double t[3]; int g[3];
        spme_setup(pos + 3 * i, L, nk, t, g);
        double w[3][32];
        for (int d = 0; d < 3; ++d)
            for (int k = 0; k < p; ++k) w[d][k] = orc_bspline(p, k + t[d]);
        for (int a = 0; a < p; ++a) {
            int ga = (g[0] - a + nk[0]) % nk[0];
            for (int b = 0; b < p; ++b) {
                int gb = (g[1] - b + nk[1]) % nk[1];
                for (int c = 0; c < p; ++c) {
                    int gc = (g[2] - c + nk[2]) % nk[2];
                    rho[(ga * nk[1] + gb) * nk[2] + gc] += q[i] * w[0][a] * w[1][b] * w[2][c];
                }
            }
        }
    }
}

/* In-place complex FFT of length n with stride: recursive mixed radix over
 * the smallest prime factors (2, 3, 5, 7; a plain DFT for anything else).
 * tw[k] = exp(sign * 2 pi i k / n_top) for the top-level length n_top, and
 * a sub-transform of length n uses every (n_top / n)-th entry.  sign = -1
 * is the forward transform, +1 the unnormalised inverse: the discrete
 * Fourier transform that cuFFT computes at src/forces/spme.rs:483-500. */
static void fft_rec(int n, int n_top, const double *twr, const double *twi,
                    double *re, double *im, int stride, double *scratch)
{
    if (n == 1) return;
    int r = 0;
    for (int f = 2; f <= 7; ++f) if (n % f == 0) { r = f; break; }
    if (r == 0) r = n;
    int m = n / r, step = n_top / n;
    double *sre = scratch, *sim = scratch + n, *rest = scratch + 2 * n;
    for (int s = 0; s < r; ++s)
        for (int j = 0; j < m; ++j) {
            sre[s * m + j] = re[(j * r + s) * stride];
            sim[s * m + j] = im[(j * r + s) * stride];
        }
    for (int s = 0; s < r; ++s) fft_rec(m, n_top, twr, twi, sre + s * m, sim + s * m, 1, rest);
    for (int k = 0; k < n; ++k) {
        double ar = 0.0, ai = 0.0;
        for (int s = 0; s < r; ++s) {
            int t = (int)(((long)s * k) % n) * step;
            double c = twr[t], sn = twi[t];
            double xr = sre[s * m + (k % m)], xi = sim[s * m + (k % m)];
            ar += xr * c - xi * sn;
            ai += xr * sn + xi * c;
        }
        re[k * stride] = ar;
        im[k * stride] = ai;
    }
}

static void fft_line(int n, int sign, double *re, double *im, int stride)
{
    double *twr = malloc(sizeof(double) * n), *twi = malloc(sizeof(double) * n);
    double *scratch = malloc(sizeof(double) * 4 * n + 64);
    for (int k = 0; k < n; ++k) {
        double ang = sign * 2.0 * ORC_PI * k / n;
        twr[k] = cos(ang); twi[k] = sin(ang);
    }
    fft_rec(n, n, twr, twi, re, im, stride, scratch);
    free(twr); free(twi); free(scratch);
}

/* 3D complex FFT over a row-major (n0, n1, n2) array, in place. */
void orc_fft3d(const int *nk, int sign, double *re, double *im)
{
    int n0 = nk[0], n1 = nk[1], n2 = nk[2];
    #pragma omp parallel for schedule(static)
    for (int ab = 0; ab < n0 * n1; ++ab)
        fft_line(n2, sign, re + (long)ab * n2, im + (long)ab * n2, 1);
    #pragma omp parallel for schedule(static)
    for (int ac = 0; ac < n0 * n2; ++ac) {
        int a = ac / n2, c = ac % n2;
        fft_line(n1, sign, re + (long)a * n1 * n2 + c, im + (long)a * n1 * n2 + c, n2);
    }
    #pragma omp parallel for schedule(static)
    for (int bc = 0; bc < n1 * n2; ++bc)
        fft_line(n0, sign, re + bc, im + bc, n1 * n2);
}

/* Influence function and virial factor at one full-grid index
 * (kernels/spme_recip.cu:34-114), f64. */
static void influence(const double *L, const int *nk, const double *ba,
                      const double *bb, const double *bc, double alpha,
                      double k_coulomb, int ka, int kb, int kc,
                      double *G, double *vf)
{
    double b[3][3];
    recip_rows(L, b);
    double vol = L[0] * L[1] * L[2];
    double ma = (ka <= nk[0] / 2) ? ka : ka - nk[0];
    double mb = (kb <= nk[1] / 2) ? kb : kb - nk[1];
    double mc = (kc <= nk[2] / 2) ? kc : kc - nk[2];
    double kx = 2 * ORC_PI * (ma * b[0][0]);
    double ky = 2 * ORC_PI * (ma * b[0][1] + mb * b[1][1]);
    double kz = 2 * ORC_PI * (ma * b[0][2] + mb * b[1][2] + mc * b[2][2]);
    double k2 = kx * kx + ky * ky + kz * kz;
    if (k2 == 0.0) { *G = 0.0; *vf = 0.0; return; }
    double g = k_coulomb / vol * (4 * ORC_PI / k2) * exp(-k2 / (4 * alpha * alpha))
               * ba[ka] * bb[kb] * bc[kc];
    *G = g;
    *vf = g * (1.0 - k2 / (2 * alpha * alpha));
}

/* Full SPME reciprocal evaluation in FP64: spread -> FFT -> influence ->
 * inverse FFT -> gather (src/forces/spme.rs:403-503 and
 * kernels/spme_spread_gather.cuh:227-334).  Outputs per-atom F[3n],
 * E[n] = 0.5 q phi - u_self, and returns U_recip (without self energy);
 * *w_recip receives the scalar reciprocal virial
 * 0.5 * sum hw*vf*|rho_hat|^2 (kernels/spme_recip.cu:296-402).
 * rho_out / v_out (f64[M]) may be NULL. */
double orc_spme_recip(int n, const double *pos, const double *q, const double *L,
                      const int *nk, int p, double alpha, double k_coulomb,
                      double *F, double *E, double *w_recip,
                      double *rho_out, double *v_out)
{
    int M = nk[0] * nk[1] * nk[2];
    double *re = calloc(M, sizeof(double)), *im = calloc(M, sizeof(double));
    orc_spme_spread(n, pos, q, L, nk, p, re);
    if (rho_out) memcpy(rho_out, re, sizeof(double) * M);
    double *ba = malloc(sizeof(double) * nk[0]);
    double *bb = malloc(sizeof(double) * nk[1]);
    double *bc = malloc(sizeof(double) * nk[2]);
    orc_b_factors(nk[0], p, ba);
    orc_b_factors(nk[1], p, bb);
    orc_b_factors(nk[2], p, bc);
    orc_fft3d(nk, -1, re, im);
    double energy = 0.0, wsum = 0.0;
    for (int ka = 0; ka < nk[0]; ++ka)
    for (int kb = 0; kb < nk[1]; ++kb)
    for (int kc = 0; kc < nk[2]; ++kc) {
        int idx = (ka * nk[1] + kb) * nk[2] + kc;
        double G, vf;
        influence(L, nk, ba, bb, bc, alpha, k_coulomb, ka, kb, kc, &G, &vf);
        double r2 = re[idx] * re[idx] + im[idx] * im[idx];
        energy += G * r2;
        wsum += vf * r2;     /* full grid: the Hermitian weight is implicit */
        re[idx] *= G; im[idx] *= G;
    }
    orc_fft3d(nk, +1, re, im);
    if (v_out) memcpy(v_out, re, sizeof(double) * M);
    if (w_recip) *w_recip = 0.5 * wsum;
    double b[3][3];
    recip_rows(L, b);
    double self_pref = k_coulomb * alpha / sqrt(ORC_PI);
    for (int i = 0; i < n; ++i) {
        double t[3]; int g[3];
        spme_setup(pos + 3 * i, L, nk, t, g);
        double w[3][32], dw[3][32];
        for (int d = 0; d < 3; ++d)
            for (int k = 0; k < p; ++k) bspline_wd(p, k + t[d], &w[d][k], &dw[d][k]);
        double phi = 0, ga = 0, gb = 0, gc = 0;
        for (int a = 0; a < p; ++a) {
            int ia = (g[0] - a + nk[0]) % nk[0];
            for (int bq = 0; bq < p; ++bq) {
                int ib = (g[1] - bq + nk[1]) % nk[1];
                for (int c = 0; c < p; ++c) {
                    int ic = (g[2] - c + nk[2]) % nk[2];
                    double v = re[(ia * nk[1] + ib) * nk[2] + ic];
                    phi += v * w[0][a] * w[1][bq] * w[2][c];
                    ga += v * dw[0][a] * w[1][bq] * w[2][c];
                    gb += v * w[0][a] * dw[1][bq] * w[2][c];
                    gc += v * w[0][a] * w[1][bq] * dw[2][c];
                }
            }
        }
        ga *= nk[0]; gb *= nk[1]; gc *= nk[2];
        for (int d = 0; d < 3; ++d)
            F[3 * i + d] = -q[i] * (ga * b[0][d] + gb * b[1][d] + gc * b[2][d]);
        E[i] = 0.5 * q[i] * phi - self_pref * q[i] * q[i];
    }
    free(re); free(im); free(ba); free(bb); free(bc);
    return 0.5 * energy;
}

/* ------------------------------------------------------------------ */
/* SETTLE (a19), FP64 restatement of kernels/settle.cu                  */
/* ------------------------------------------------------------------ */

static double csqrt0(double x) { return sqrt(x > 0.0 ? x : 0.0); }

/* Canonical geometry (src/integrator/settle.rs:83-93). out = {ra, rb, rc}. */
void orc_settle_geometry(double d_oh, double d_hh, double m_o, double m_h,
                         double *out)
{
    double total = m_o + 2.0 * m_h;
    double h = sqrt(d_oh * d_oh - (d_hh * 0.5) * (d_hh * 0.5));
    out[0] = (2.0 * m_h / total) * h;
    out[1] = (m_o / total) * h;
    out[2] = d_hh * 0.5;
}

/* Miyamoto-Kollman position reset of one water (kernels/settle.cu:145-335).
 * snp/cur: 3 atoms x 3 coords (O, H1, H2); updates cur and v in place and
 * writes the position-level constraint virial per atom (vir[3]). */
void orc_settle_positions_one(const double *snp, double *cur, double *v,
                              double m_o, double m_h, double ra, double rb,
                              double rc, double dt, const double *L,
                              double *vir)
{
    double b0[3], c0[3], xp[3][3];
    for (int d = 0; d < 3; ++d) { b0[d] = snp[3 + d] - snp[d]; c0[d] = snp[6 + d] - snp[d]; }
    mi_f64(&b0[0], &b0[1], &b0[2], L);
    mi_f64(&c0[0], &c0[1], &c0[2], L);
    for (int a = 0; a < 3; ++a) {
        double d0 = cur[3 * a] - snp[3 * a], d1 = cur[3 * a + 1] - snp[3 * a + 1],
               d2 = cur[3 * a + 2] - snp[3 * a + 2];
        mi_f64(&d0, &d1, &d2, L);
        xp[a][0] = d0; xp[a][1] = d1; xp[a][2] = d2;
    }
    double m[3] = {m_o, m_h, m_h};
    double invM = 1.0 / (m_o + 2.0 * m_h);
    double unc[3][3], com[3];
    for (int d = 0; d < 3; ++d) {
        unc[0][d] = xp[0][d];
        unc[1][d] = b0[d] + xp[1][d];
        unc[2][d] = c0[d] + xp[2][d];
        com[d] = (unc[0][d] * m[0] + unc[1][d] * m[1] + unc[2][d] * m[2]) * invM;
    }
    double a1[3], b1[3], c1[3];
    for (int d = 0; d < 3; ++d) { a1[d] = unc[0][d] - com[d]; b1[d] = unc[1][d] - com[d]; c1[d] = unc[2][d] - com[d]; }
    double ez[3] = {b0[1] * c0[2] - b0[2] * c0[1], b0[2] * c0[0] - b0[0] * c0[2], b0[0] * c0[1] - b0[1] * c0[0]};
    double ex[3] = {a1[1] * ez[2] - a1[2] * ez[1], a1[2] * ez[0] - a1[0] * ez[2], a1[0] * ez[1] - a1[1] * ez[0]};
    double ey[3] = {ez[1] * ex[2] - ez[2] * ex[1], ez[2] * ex[0] - ez[0] * ex[2], ez[0] * ex[1] - ez[1] * ex[0]};
    double lx_ = csqrt0(ex[0] * ex[0] + ex[1] * ex[1] + ex[2] * ex[2]);
    double ly_ = csqrt0(ey[0] * ey[0] + ey[1] * ey[1] + ey[2] * ey[2]);
    double lz_ = csqrt0(ez[0] * ez[0] + ez[1] * ez[1] + ez[2] * ez[2]);
    for (int d = 0; d < 3; ++d) {
        ex[d] = lx_ > 0 ? ex[d] / lx_ : 0; ey[d] = ly_ > 0 ? ey[d] / ly_ : 0;
        ez[d] = lz_ > 0 ? ez[d] / lz_ : 0;
    }
#define DOT(u, w) ((u)[0] * (w)[0] + (u)[1] * (w)[1] + (u)[2] * (w)[2])
    double b0x = DOT(ex, b0), b0y = DOT(ey, b0), c0x = DOT(ex, c0), c0y = DOT(ey, c0);
    double a1z = DOT(ez, a1);
    double b1x = DOT(ex, b1), b1y = DOT(ey, b1), b1z = DOT(ez, b1);
    double c1x = DOT(ex, c1), c1y = DOT(ey, c1), c1z = DOT(ez, c1);
#undef DOT
    double sinphi = a1z / ra;
    double cosphi = csqrt0(1.0 - sinphi * sinphi);
    double sinpsi = (b1z - c1z) / (2.0 * rc * cosphi);
    double cospsi = csqrt0(1.0 - sinpsi * sinpsi);
    double a2y = ra * cosphi;
    double b2x = -rc * cospsi;
    double b2y = -rb * cosphi - rc * sinpsi * sinphi;
    double c2y = -rb * cosphi + rc * sinpsi * sinphi;
    double hh = 4.0 * b2x * b2x + (b2y - c2y) * (b2y - c2y) + (b1z - c1z) * (b1z - c1z);
    double d2sq = 4.0 * rc * rc;
    double delta = 2.0 * b2x + csqrt0(4.0 * b2x * b2x - hh + d2sq);
    b2x -= 0.5 * delta;
    double al = b2x * (b0x - c0x) + b0y * b2y + c0y * c2y;
    double be = b2x * (c0y - b0y) + b0x * b2y + c0x * c2y;
    double ga = b0x * b1y - b1x * b0y + c0x * c1y - c1x * c0y;
    double ab2 = al * al + be * be;
    double sinth = ab2 > 0 ? (al * ga - be * csqrt0(ab2 - ga * ga)) / ab2 : 0.0;
    double costh = csqrt0(1.0 - sinth * sinth);
    double pr[3][3] = {
        {-a2y * sinth, a2y * costh, a1z},
        {b2x * costh - b2y * sinth, b2x * sinth + b2y * costh, b1z},
        {-b2x * costh - c2y * sinth, -b2x * sinth + c2y * costh, c1z}};
    double invdt = dt != 0 ? 1.0 / dt : 0.0, invdt2 = dt != 0 ? 1.0 / (dt * dt) : 0.0;
    for (int a = 0; a < 3; ++a) {
        double rcm[3], corr[3];
        for (int d = 0; d < 3; ++d) {
            rcm[d] = ex[d] * pr[a][0] + ey[d] * pr[a][1] + ez[d] * pr[a][2];
            corr[d] = (com[d] + rcm[d]) - unc[a][d];
            v[3 * a + d] += corr[d] * invdt;
            cur[3 * a + d] += corr[d];
        }
        vir[a] = m[a] * invdt2 * (corr[0] * rcm[0] + corr[1] * rcm[1] + corr[2] * rcm[2]);
    }
}

/* Analytic velocity projection of one water (kernels/settle.cu:376-490):
 * solve the 3x3 system for bond impulses so that (v_i - v_j).r_ij = 0 for
 * the three bonds. pos/v: 3 atoms x 3. */
void orc_settle_velocities_one(const double *pos_in, double *v, double m_o,
                               double m_h, const double *L, double dt,
                               double *vir)
{
    double p[3][3];
    for (int a = 0; a < 3; ++a) for (int d = 0; d < 3; ++d) p[a][d] = pos_in[3 * a + d];
    for (int a = 1; a < 3; ++a) {
        double dx = p[a][0] - p[0][0], dy = p[a][1] - p[0][1], dz = p[a][2] - p[0][2];
        mi_f64(&dx, &dy, &dz, L);
        p[a][0] = p[0][0] + dx; p[a][1] = p[0][1] + dy; p[a][2] = p[0][2] + dz;
    }
    double io = 1.0 / m_o, ih = 1.0 / m_h;
    double r1[3], r2[3], r3[3];
    for (int d = 0; d < 3; ++d) { r1[d] = p[0][d] - p[1][d]; r2[d] = p[0][d] - p[2][d]; r3[d] = p[1][d] - p[2][d]; }
#define DOT(u, w) ((u)[0] * (w)[0] + (u)[1] * (w)[1] + (u)[2] * (w)[2])
    double d11 = DOT(r1, r1), d22 = DOT(r2, r2), d33 = DOT(r3, r3);
    double d12 = DOT(r1, r2), d13 = DOT(r1, r3), d23 = DOT(r2, r3);
    double v01[3], v02[3], v12[3];
    for (int d = 0; d < 3; ++d) { v01[d] = v[d] - v[3 + d]; v02[d] = v[d] - v[6 + d]; v12[d] = v[3 + d] - v[6 + d]; }
    double b1 = DOT(v01, r1), b2 = DOT(v02, r2), b3 = DOT(v12, r3);
#undef DOT
    double M[3][3] = {{d11 * (io + ih), d12 * io, -d13 * ih},
                      {d12 * io, d22 * (io + ih), d23 * ih},
                      {-d13 * ih, d23 * ih, 2.0 * d33 * ih}};
    double rhs[3] = {-b1, -b2, -b3};
    double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1])
               - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0])
               + M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    if (det == 0.0) return;
    double g[3];
    for (int c = 0; c < 3; ++c) {
        double A[3][3];
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) A[i][j] = (j == c) ? rhs[i] : M[i][j];
        g[c] = (A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1])
              - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0])
              + A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0])) / det;
    }
    double dv[3][3];
    for (int d = 0; d < 3; ++d) {
        dv[0][d] = (g[0] * r1[d] + g[1] * r2[d]) * io;
        dv[1][d] = (-g[0] * r1[d] + g[2] * r3[d]) * ih;
        dv[2][d] = (-g[1] * r2[d] - g[2] * r3[d]) * ih;
        for (int a = 0; a < 3; ++a) v[3 * a + d] += dv[a][d];
    }
    /* Velocity-level constraint virial per atom (dt > 0):
     * (m_a dv_a / dt) . (r_a - r_com), the impulse of the projection as a
     * force over the step (kernels/settle.cu, settle_velocities). */
    if (vir && dt > 0.0) {
        double m[3] = {m_o, m_h, m_h}, com[3];
        for (int d = 0; d < 3; ++d)
            com[d] = (m_o * p[0][d] + m_h * p[1][d] + m_h * p[2][d]) / (m_o + 2.0 * m_h);
        for (int a = 0; a < 3; ++a) {
            double w = 0.0;
            for (int d = 0; d < 3; ++d) w += dv[a][d] * (p[a][d] - com[d]);
            vir[a] = m[a] / dt * w;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Velocity Verlet (a20), kernels/integrate.cu:24-198                  */
/* ------------------------------------------------------------------ */

/* One f64 VV step for free particles under a force callback-free constant
 * force field: used by the kinematics tests.  x += v dt + F/m dt^2/2 via
 * kick(dt/2) -> drift(dt) -> kick(dt/2) with the same F. */
void orc_vv_constant_force(int n, double *x, double *v, const double *F,
                           const double *m, double dt, int steps)
{
    for (int s = 0; s < steps; ++s)
        for (int i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) {
                v[3 * i + d] += F[3 * i + d] / m[i] * 0.5 * dt;
                x[3 * i + d] += v[3 * i + d] * dt;
                v[3 * i + d] += F[3 * i + d] / m[i] * 0.5 * dt;
            }
}

/* f32 velocity-Verlet halves, bit-exact with the device kernels.
 * kernels/integrate.cu:24-215 (vv_kick_drift_body / vv_kick_body) compiled
 * by nvcc with its default FMA contraction: a = F / m is an IEEE divide,
 * v + a dt/2 and x + v dt round once (fmaf); the compensated ("lossless")
 * variant forms the f32 product a dt/2, adds it to (v_hi + v_lo) in f64,
 * re-splits, then rounds (x_hi + x_lo) + (v_hi + v_lo) dt once (fma).  The
 * wrap is triclinic_wrap_with_image (kernels/pbc.cuh:74-91) with every f32
 * operation rounded once.  Arrays are SoA [3][n]; posq is float4[n]. */
/* float -> int as the GPU converts (saturating, NaN -> 0); a plain C cast
 * of NaN or an out-of-range value is undefined. */
static int f2i_sat(float f)
{
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f <= -2147483648.0f) return INT32_MIN;
    return (int)f;
}

static void wrap_image_f32(float *x, float *y, float *z, int *ka, int *kb, int *kc,
                           const float *L)
{
    float sc = *z / L[2];
    float sb = (*y - sc * L[5]) / L[1];
    float sa = (*x - sb * L[3] - sc * L[4]) / L[0];
    float fa = floorf(sa + 0.5f), fb = floorf(sb + 0.5f), fc = floorf(sc + 0.5f);
    *x -= fa * L[0] + fb * L[3] + fc * L[4];
    *y -= fb * L[1] + fc * L[5];
    *z -= fc * L[2];
    *ka = f2i_sat(fa);
    *kb = f2i_sat(fb);
    *kc = f2i_sat(fc);
}

void orc_vv_kick_drift_f32(int n, float *posq, float *v, double *x_lo, double *v_lo,
                           const float *f, const float *m, int *images, const float *L,
                           float dt, int compensated)
{
    const float half_dt = dt * 0.5f;
    for (int i = 0; i < n; ++i) {
        float x[3] = {posq[4 * i], posq[4 * i + 1], posq[4 * i + 2]};
        for (int d = 0; d < 3; ++d) {
            const size_t k = (size_t)d * n + i;
            const float a = f[k] / m[i];
            if (compensated) {
                double ext_v = ((double)v[k] + v_lo[k]) + (double)(a * half_dt);
                float vh = (float)ext_v;
                double vl = ext_v - (double)vh;
                v[k] = vh;
                v_lo[k] = vl;
                double ext_x = fma((double)vh + vl, (double)dt, (double)x[d] + x_lo[k]);
                float xh = (float)ext_x;
                x_lo[k] = ext_x - (double)xh;
                x[d] = xh;
            } else {
                v[k] = fmaf(a, half_dt, v[k]);
                x[d] = fmaf(v[k], dt, x[d]);
            }
        }
        int ka, kb, kc;
        wrap_image_f32(&x[0], &x[1], &x[2], &ka, &kb, &kc, L);
        images[i] += ka;
        images[n + i] += kb;
        images[2 * n + i] += kc;
        posq[4 * i] = x[0];
        posq[4 * i + 1] = x[1];
        posq[4 * i + 2] = x[2];
    }
}

void orc_vv_kick_f32(int n, float *v, double *v_lo, const float *f, const float *m,
                     float dt, int compensated)
{
    const float half_dt = dt * 0.5f;
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            const size_t k = (size_t)d * n + i;
            const float a = f[k] / m[i];
            if (compensated) {
                double ext_v = ((double)v[k] + v_lo[k]) + (double)(a * half_dt);
                float vh = (float)ext_v;
                v[k] = vh;
                v_lo[k] = ext_v - (double)vh;
            } else {
                v[k] = fmaf(a, half_dt, v[k]);
            }
        }
}

/* ------------------------------------------------------------------ */
/* Philox-4x32-10 (a21), kernels/philox.cuh:25-150, philox.rs:21-73     */
/* ------------------------------------------------------------------ */

void orc_philox4x32_10(uint32_t k0, uint32_t k1, const uint32_t *ctr, uint32_t *out)
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)c0 * 0xD2511F53u;
        uint64_t p1 = (uint64_t)c2 * 0xCD9E8D57u;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Box-Muller cos branch on the first two Philox words
 * (kernels/philox.cuh:136-150). */
double orc_philox_gaussian(uint32_t seed_lo, uint32_t seed_hi, uint32_t step_lo,
                           uint32_t step_hi, uint32_t pid, uint32_t axis)
{
    uint32_t ctr[4] = {step_lo, step_hi, pid, axis}, o[4];
    orc_philox4x32_10(seed_lo, seed_hi, ctr, o);
    double u1 = ((double)o[0] + 0.5) / 4294967296.0;
    double u2 = ((double)o[1] + 0.5) / 4294967296.0;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* CSVR (kernels/nose_hoover.cu:181-256, Bussi-Donadio-Parrinello 2007):
 * the new kinetic energy from the old one and the Philox normals
 * xi_i = philox_gaussian(seed, counter, i, 0), i = 0 .. g_dof - 1:
 * K' = c K + kt_half (1 - c)(sum_{i>=1} xi_i^2 + xi_0^2)
 *      + 2 xi_0 sqrt(c (1 - c) K kt_half), K' = K if not finite or <= 0.
 * Returns K' (the rescale factor is sqrt(K'/K)). */
double orc_csvr_knew(double k_old, uint32_t seed_lo, uint32_t seed_hi, uint64_t counter,
                     uint32_t g_dof, double c, double kt_half)
{
    uint32_t cl = (uint32_t)(counter & 0xFFFFFFFFu), ch = (uint32_t)(counter >> 32);
    double s = 0.0;
    for (uint32_t i = 1; i < g_dof; ++i) {
        double xi = orc_philox_gaussian(seed_lo, seed_hi, cl, ch, i, 0);
        s += xi * xi;
    }
    double r = orc_philox_gaussian(seed_lo, seed_hi, cl, ch, 0, 0);
    double cross = k_old > 0.0 ? 2.0 * r * sqrt(c * (1.0 - c) * k_old * kt_half) : 0.0;
    double kn = c * k_old + kt_half * (1.0 - c) * (s + r * r) + cross;
    if (!isfinite(kn) || kn <= 0.0) kn = k_old;
    return kn;
}

/* C-rescale barostat move (kernels/barostat.cu:169-227; Bernetti & Bussi,
 * J. Chem. Phys. 153, 114107 (2020)) in FP64: with the kinetic energy k,
 * the virial sum w and the volume v,
 *   P   = (2k + w) / (3v),
 *   mu3 = 1 - beta (dt/tau)(P_target - P) + sqrt(2 beta kT dt/(tau v)) r,
 * r = philox_gaussian(seed, counter, 0, 0), mu3 floored at mu3_min.
 * out: [mu = cbrt(mu3), P, v mu3, P_target (v mu3 - v)]. */
void orc_crescale_mu(double k, double w, double v, uint32_t seed_lo, uint32_t seed_hi,
                     uint64_t counter, double p_target, double tau, double beta, double kt,
                     double dt, double mu3_min, double out[4])
{
    uint32_t cl = (uint32_t)(counter & 0xFFFFFFFFu), ch = (uint32_t)(counter >> 32);
    double pressure = (2.0 * k + w) / (3.0 * v);
    double r = orc_philox_gaussian(seed_lo, seed_hi, cl, ch, 0, 0);
    double mu3 = 1.0 - beta * (dt / tau) * (p_target - pressure)
               + sqrt(2.0 * beta * kt * dt / (tau * v)) * r;
    if (!(mu3 > mu3_min)) mu3 = mu3_min;
    out[0] = cbrt(mu3);
    out[1] = pressure;
    out[2] = v * mu3;
    out[3] = p_target * (v * mu3 - v);
}

/* Langevin BAOAB A and O steps in f32 (kernels/langevin.cu:13-80) with
 * the reference's contracted roundings: x + v dt/2 once (fmaf);
 * sigma = sqrt(fl(1 - a^2) kT / m) with 1 - a^2 rounded once; v = a v +
 * fl(sigma xi) rounded once; xi = (float) of the f64 Box-Muller normal.  The
 * f64 log / cos come from the host libm: a last-bit difference from the
 * device's can move xi by one f32 ulp. */
void orc_lan_drift_half_f32(int n, float *posq, const float *v, int *images,
                            const float *L, float dt)
{
    const float h = dt * 0.5f;
    for (int i = 0; i < n; ++i) {
        float x = fmaf(v[i], h, posq[4 * i]);
        float y = fmaf(v[n + i], h, posq[4 * i + 1]);
        float z = fmaf(v[2 * n + i], h, posq[4 * i + 2]);
        int ka, kb, kc;
        wrap_image_f32(&x, &y, &z, &ka, &kb, &kc, L);
        images[i] += ka;
        images[n + i] += kb;
        images[2 * n + i] += kc;
        posq[4 * i] = x;
        posq[4 * i + 1] = y;
        posq[4 * i + 2] = z;
    }
}

void orc_lan_ou_step_f32(int n, float *v, const float *m, uint64_t seed, float a, float kt,
                         uint64_t draw_counter)
{
    const uint32_t sl = (uint32_t)seed, sh = (uint32_t)(seed >> 32);
    const uint32_t cl = (uint32_t)draw_counter, ch = (uint32_t)(draw_counter >> 32);
    for (int i = 0; i < n; ++i) {
        const float sig = sqrtf(fmaf(-a, a, 1.0f) * kt / m[i]);
        for (int d = 0; d < 3; ++d) {
            const float xi = (float)orc_philox_gaussian(sl, sh, cl, ch, (uint32_t)i, (uint32_t)d);
            v[(size_t)d * n + i] = fmaf(a, v[(size_t)d * n + i], sig * xi);
        }
    }
}

/* ------------------------------------------------------------------ */
/* Reductions (a22)                                                     */
/* ------------------------------------------------------------------ */

double orc_kinetic_energy(int n, const double *v, const double *m)
{
    double ke = 0.0;
    for (int i = 0; i < n; ++i)
        ke += 0.5 * m[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1]
                            + v[3 * i + 2] * v[3 * i + 2]);
    return ke;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- bonded terms ----------------------------------------------------------
 * Harmonic bond U = k (r - r0)^2 / 2 (src/forces/harmonic_bond.rs:316-360),
 * Morse U = De (1 - exp(-a (r - r0)))^2 (morse.rs:330-360), harmonic angle
 * U = k (theta - theta0)^2 / 2 at the middle atom (angle.rs:350-410) and
 * periodic dihedral U = k (1 + cos(n phi - phi0)) with phi = atan2(|b2| b1 .
 * (b2 x b3), (b1 x b2) . (b2 x b3)), b1 = r_i - r_j, b2 = r_k - r_j,
 * b3 = r_k - r_l (dihedral.rs:440-530), in FP64.  kind 1..4 as above;
 * terms [n_terms][arity + 1]; params [n_types][3].  Adds each endpoint's
 * force to f[atom][3] and the term's energy and virial (sum of (r_a - r_j)
 * . F_a) in equal shares to e[atom], w[atom].  Returns the total energy. */
double orc_bonded(int kind, const double *pos, const double *L, int n_terms,
                  const uint32_t *terms, const double *params, double *f, double *e,
                  double *w)
{
    int arity = kind <= 2 ? 2 : kind == 3 ? 3 : 4;
    double total = 0.0;
    for (int t = 0; t < n_terms; ++t) {
        const uint32_t *q = terms + (size_t)(arity + 1) * t;
        const double *c = params + 3 * (size_t)q[arity];
        double d[3][3];            /* displacements: a -> (q[0]-q[1]), b, c */
        double fa[4][3] = {{0}};
        double u = 0.0;
        int pairs[3][2] = {{0, 1}, {2, 1}, {2, 3}};
        int nd = arity == 2 ? 1 : arity == 3 ? 2 : 3;
        for (int m = 0; m < nd; ++m) {
            const double *pa = pos + 3 * (size_t)q[pairs[m][0]];
            const double *pb = pos + 3 * (size_t)q[pairs[m][1]];
            d[m][0] = pa[0] - pb[0]; d[m][1] = pa[1] - pb[1]; d[m][2] = pa[2] - pb[2];
            mi_f64(&d[m][0], &d[m][1], &d[m][2], L);
        }
        if (arity == 2) {
            double r = sqrt(d[0][0] * d[0][0] + d[0][1] * d[0][1] + d[0][2] * d[0][2]);
            double g = 0.0;        /* -dU/dr */
            if (r > 0.0) {
                if (kind == 1) {
                    u = 0.5 * c[0] * (r - c[1]) * (r - c[1]);
                    g = -c[0] * (r - c[1]);
                } else {
                    double ex = exp(-c[1] * (r - c[2]));
                    u = c[0] * (1.0 - ex) * (1.0 - ex);
                    g = -2.0 * c[0] * c[1] * (1.0 - ex) * ex;
                }
                for (int k = 0; k < 3; ++k) { fa[0][k] = g * d[0][k] / r; fa[1][k] = -fa[0][k]; }
            }
        } else if (arity == 3) {
            double *a = d[0], *b = d[1];
            double la = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
            double lb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
            double dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
            double cx[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                            a[0] * b[1] - a[1] * b[0]};
            double lc = sqrt(cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2]);
            if (la > 0.0 && lb > 0.0 && lc > 1e-7 * la * lb) {
                double theta = atan2(lc, dot), ct = dot / (la * lb), st = lc / (la * lb);
                double dth = theta - c[1];
                u = 0.5 * c[0] * dth * dth;
                /* dtheta/dr_i = (cos a/|a|^2 - b/(|a||b|)) / sin, likewise k */
                double g = -c[0] * dth / st;
                for (int k = 0; k < 3; ++k) {
                    fa[0][k] = g * (ct * a[k] / (la * la) - b[k] / (la * lb));
                    fa[2][k] = g * (ct * b[k] / (lb * lb) - a[k] / (la * lb));
                    fa[1][k] = -(fa[0][k] + fa[2][k]);
                }
            }
        } else {
            double *b1 = d[0], *b2 = d[1], *b3 = d[2];
            double m[3] = {b1[1] * b2[2] - b1[2] * b2[1], b1[2] * b2[0] - b1[0] * b2[2],
                           b1[0] * b2[1] - b1[1] * b2[0]};
            double nn[3] = {b2[1] * b3[2] - b2[2] * b3[1], b2[2] * b3[0] - b2[0] * b3[2],
                            b2[0] * b3[1] - b2[1] * b3[0]};
            double m2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
            double n2 = nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2];
            double bb = b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2];
            if (m2 >= 1e-14 && n2 >= 1e-14 && bb >= 1e-14) {
                double lb = sqrt(bb);
                double phi = atan2(lb * (b1[0] * nn[0] + b1[1] * nn[1] + b1[2] * nn[2]),
                                   m[0] * nn[0] + m[1] * nn[1] + m[2] * nn[2]);
                double delta = c[2] * phi - c[1];
                u = c[0] * (1.0 + cos(delta));
                double fphi = c[0] * c[2] * sin(delta);     /* -dU/dphi */
                double s = (b1[0] * b2[0] + b1[1] * b2[1] + b1[2] * b2[2]) / bb;
                double tt = (b3[0] * b2[0] + b3[1] * b2[1] + b3[2] * b2[2]) / bb;
                for (int k = 0; k < 3; ++k) {
                    fa[0][k] = fphi * lb / m2 * m[k];
                    fa[3][k] = -fphi * lb / n2 * nn[k];
                    fa[1][k] = (s - 1.0) * fa[0][k] - tt * fa[3][k];
                    fa[2][k] = -s * fa[0][k] + (tt - 1.0) * fa[3][k];
                }
            }
        }
        /* virial about atom q[1]: r_i - r_j = d[0], r_k - r_j = d[1],
         * r_l - r_j = d[1] - d[2] */
        double wv = 0.0;
        for (int k = 0; k < 3; ++k) {
            wv += d[0][k] * fa[0][k];
            if (arity >= 3) wv += d[1][k] * fa[2][k];
            if (arity == 4) wv += (d[1][k] - d[2][k]) * fa[3][k];
        }
        for (int a2 = 0; a2 < arity; ++a2) {
            for (int k = 0; k < 3; ++k) f[3 * (size_t)q[a2] + k] += fa[a2][k];
            e[q[a2]] += u / arity;
            w[q[a2]] += wv / arity;
        }
        total += u;
    }
    return total;
}
