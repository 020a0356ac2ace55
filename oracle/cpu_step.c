/* Same-arithmetic CPU baseline of the MD step -- TEST / BENCH INFRASTRUCTURE.
 *
 * The step the GPU path takes, restated for the host cores (C11 + OpenMP) so
 * that bench.py can time it beside the GPU: f32 forces (LJ 12-6 with the C1
 * switch + real-space Ewald from the same FP64-built cubic erfc table), a
 * cell list and a half Verlet list with skin (displacement-triggered
 * rebuilds at r_skin/2, kernels/neighbor.cu:50-76), SPME order 4 in f32
 * (closed-form cubic B-splines, slab-parallel spread, a radix-2 complex 3D
 * FFT, the influence function G(k) of kernels/spme_recip.cu:34-114 built in
 * FP64, gather), compensated (f32, f64) velocity Verlet
 * (kernels/integrate.cu:24-314) and SETTLE (the FP64 oracle's per-molecule
 * reset and projection on the f32 state).  Only bench.py's cpu_baseline leg
 * uses it; nothing in the product does.  Orthorhombic boxes, power-of-two
 * meshes. */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

void orc_settle_positions_one(const double *snp, double *cur, double *v, double m_o,
                              double m_h, double ra, double rb, double rc, double dt,
                              const double *L, double *vir);
void orc_settle_velocities_one(const double *pos_in, double *v, double m_o, double m_h,
                               const double *L, double dt, double *vir);

#define CS_BINS 512
#define CS_PI 3.14159265358979323846

typedef struct {
    float re, im;
} cs_cpx;

typedef struct {
    int n, n_types, n_groups, compensated;
    float L[3];
    float *x, *v, *f, *q, *m;       /* x, v, f: [n][3] */
    double *xlo, *vlo;              /* compensated low parts [n][3] */
    int *type;
    float *c6, *c12;                /* [n_types^2] */
    float rc2, rs2_sw, inv_delta, kc, inv_h, dt;
    float tab[2][CS_BINS][4];       /* cubic Hermite: force factor, energy */
    int *ex_off, *ex_par;           /* exclusion CSR */
    /* neighbour list */
    float r_search, skin_half2;
    float *xref;
    int *nl_off, *nl_j, nl_cap;
    int n_rebuilds;
    /* SPME */
    int K[3];
    float *G;                       /* [K0 K1 K2] influence */
    cs_cpx *mesh;
    float *twr, *twi;               /* twiddles per axis (max K) */
    int kmax;
    /* SETTLE */
    int *groups;                    /* [n_groups][3] */
    double m_o, m_h, ra, rb, rcs;
    double *snap;                   /* [n_groups][9] */
    /* per-thread force buffers */
    int n_threads;
    float *fthr;                    /* [n_threads][n][3] */
} cs_md;

static double erfc_as(double x)
{
    double t = 1.0 / (1.0 + 0.3275911 * x);
    double p = t * (0.254829592 + t * (-0.284496736 + t * (1.421413741 +
               t * (-1.453152027 + t * 1.061405429))));
    return p * exp(-x * x);
}

static void build_table(cs_md *s, double alpha, double rc)
{
    const int nb = CS_BINS;
    const double h = rc / (nb - 1);
    for (int fk = 0; fk < 2; ++fk) {
        double *v = malloc(sizeof(double) * (nb + 1)), *dv = malloc(sizeof(double) * (nb + 1));
        for (int k = 0; k <= nb; ++k) {
            double r0 = (k > 0 ? (double)k : 0.5) * h, d = 1e-3 * h, vals[5];
            for (int t = -2; t <= 2; ++t) {
                double r = r0 + t * d;
                double e = erfc_as(alpha * r) / r;
                vals[t + 2] = fk == 0 ? (e + 2.0 * alpha / sqrt(CS_PI) * exp(-alpha * alpha * r * r)) / (r * r) : e;
            }
            v[k] = vals[2];
            dv[k] = h * (vals[0] - 8 * vals[1] + 8 * vals[3] - vals[4]) / (12 * d);
        }
        for (int k = 0; k < nb; ++k) {
            s->tab[fk][k][0] = (float)v[k];
            s->tab[fk][k][1] = (float)dv[k];
            s->tab[fk][k][2] = (float)(3.0 * (v[k + 1] - v[k]) - 2.0 * dv[k] - dv[k + 1]);
            s->tab[fk][k][3] = (float)(2.0 * (v[k] - v[k + 1]) + dv[k] + dv[k + 1]);
        }
        free(v); free(dv);
    }
    s->inv_h = (float)(1.0 / h);
}

static double bmod2(int n, int k)
{
    /* |b(k)|^-2 for order 4: M_4 at 1, 2, 3 = 1/6, 2/3, 1/6 */
    const double m[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0};
    double th = 2.0 * CS_PI * k / n, re = 0, im = 0;
    for (int j = 0; j < 3; ++j) { re += m[j] * cos(th * j); im += m[j] * sin(th * j); }
    return re * re + im * im;
}

static void build_influence(cs_md *s, double alpha)
{
    const int *K = s->K;
    double *b[3];
    for (int d = 0; d < 3; ++d) {
        b[d] = malloc(sizeof(double) * K[d]);
        for (int k = 0; k < K[d]; ++k) {
            double m = bmod2(K[d], k);
            if (m < 1e-7) m = 0.5 * (bmod2(K[d], (k + K[d] - 1) % K[d]) + bmod2(K[d], (k + 1) % K[d]));
            b[d][k] = m > 0 ? 1.0 / m : 0.0;
        }
    }
    const double vol = (double)s->L[0] * s->L[1] * s->L[2];
    #pragma omp parallel for schedule(static)
    for (int a = 0; a < K[0]; ++a)
        for (int bb = 0; bb < K[1]; ++bb)
            for (int c = 0; c < K[2]; ++c) {
                double ma = a <= K[0] / 2 ? a : a - K[0], mb = bb <= K[1] / 2 ? bb : bb - K[1];
                double mc = c <= K[2] / 2 ? c : c - K[2];
                double kx = 2 * CS_PI * ma / s->L[0], ky = 2 * CS_PI * mb / s->L[1];
                double kz = 2 * CS_PI * mc / s->L[2];
                double k2 = kx * kx + ky * ky + kz * kz;
                double g = k2 > 0 ? s->kc / vol * 4 * CS_PI / k2 * exp(-k2 / (4 * alpha * alpha)) *
                                        b[0][a] * b[1][bb] * b[2][c] : 0.0;
                s->G[((size_t)a * K[1] + bb) * K[2] + c] = (float)g;
            }
    for (int d = 0; d < 3; ++d) free(b[d]);
}

static inline float mi1(float d, float L) { return d - L * rintf(d / L); }

/* Cell list + half Verlet list at r_search; reference positions kept. */
static void build_list(cs_md *s)
{
    const int n = s->n;
    int nc[3];
    for (int d = 0; d < 3; ++d) {
        nc[d] = (int)floorf(s->L[d] / (s->r_search * 1.001f));
        if (nc[d] < 3) nc[d] = 1;
    }
    const int ncell = nc[0] * nc[1] * nc[2];
    int *cell = malloc(sizeof(int) * n), *start = calloc(ncell + 1, sizeof(int));
    int *mem = malloc(sizeof(int) * n), *fill = calloc(ncell, sizeof(int));
    for (int i = 0; i < n; ++i) {
        int c[3];
        for (int d = 0; d < 3; ++d) {
            float u = s->x[3 * i + d] / s->L[d] + 0.5f;
            u -= floorf(u);
            c[d] = (int)(u * nc[d]);
            if (!(c[d] >= 0 && c[d] < nc[d])) c[d] = c[d] < 0 ? 0 : nc[d] - 1;   /* NaN, edge */
        }
        cell[i] = (c[0] * nc[1] + c[1]) * nc[2] + c[2];
        start[cell[i] + 1]++;
    }
    for (int c = 0; c < ncell; ++c) start[c + 1] += start[c];
    for (int i = 0; i < n; ++i) mem[start[cell[i]] + fill[cell[i]]++] = i;
    const float rs2 = s->r_search * s->r_search;
    int *cnt = calloc(n + 1, sizeof(int));
    for (int pass = 0; pass < 2; ++pass) {
        #pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < n; ++i) {
            int ci = cell[i], ca = ci / (nc[1] * nc[2]), cb = (ci / nc[2]) % nc[1], cc = ci % nc[2];
            int o = pass ? s->nl_off[i] : 0;
            int seen[27], ns = 0;
            for (int da = -1; da <= 1; ++da)
            for (int db = -1; db <= 1; ++db)
            for (int dc = -1; dc <= 1; ++dc) {
                int a = (ca + da + nc[0]) % nc[0], b = (cb + db + nc[1]) % nc[1];
                int k = (cc + dc + nc[2]) % nc[2], id = (a * nc[1] + b) * nc[2] + k, dup = 0;
                for (int t = 0; t < ns; ++t) dup |= seen[t] == id;
                if (dup) continue;
                seen[ns++] = id;
                for (int t = start[id]; t < start[id + 1]; ++t) {
                    int j = mem[t];
                    if (j <= i) continue;
                    float dx = mi1(s->x[3 * i] - s->x[3 * j], s->L[0]);
                    float dy = mi1(s->x[3 * i + 1] - s->x[3 * j + 1], s->L[1]);
                    float dz = mi1(s->x[3 * i + 2] - s->x[3 * j + 2], s->L[2]);
                    if (dx * dx + dy * dy + dz * dz <= rs2) {
                        if (pass) s->nl_j[o] = j;
                        ++o;
                    }
                }
            }
            if (!pass) cnt[i + 1] = o;
        }
        if (!pass) {
            for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
            memcpy(s->nl_off, cnt, sizeof(int) * (n + 1));
            if (cnt[n] > s->nl_cap) {
                free(s->nl_j);
                s->nl_cap = cnt[n] + cnt[n] / 4;
                s->nl_j = malloc(sizeof(int) * s->nl_cap);
            }
        }
    }
    memcpy(s->xref, s->x, sizeof(float) * 3 * n);
    s->n_rebuilds++;
    free(cell); free(start); free(mem); free(fill); free(cnt);
}

static int needs_rebuild(const cs_md *s)
{
    int any = 0;
    #pragma omp parallel for reduction(|:any) schedule(static)
    for (int i = 0; i < s->n; ++i) {
        float dx = mi1(s->x[3 * i] - s->xref[3 * i], s->L[0]);
        float dy = mi1(s->x[3 * i + 1] - s->xref[3 * i + 1], s->L[1]);
        float dz = mi1(s->x[3 * i + 2] - s->xref[3 * i + 2], s->L[2]);
        any |= dx * dx + dy * dy + dz * dz > s->skin_half2;
    }
    return any;
}

static inline int excluded(const cs_md *s, int i, int j)
{
    for (int t = s->ex_off[i]; t < s->ex_off[i + 1]; ++t)
        if (s->ex_par[t] == j) return 1;
    return 0;
}

static void pair_forces(cs_md *s)
{
    const int n = s->n;
    #pragma omp parallel
    {
        float *fb = s->fthr + (size_t)omp_get_thread_num() * 3 * n;
        memset(fb, 0, sizeof(float) * 3 * n);
        #pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < n; ++i) {
            float fx = 0, fy = 0, fz = 0;
            const float xi = s->x[3 * i], yi = s->x[3 * i + 1], zi = s->x[3 * i + 2];
            const float kqi = s->kc * s->q[i];
            const int ti = s->type[i] * s->n_types;
            for (int t = s->nl_off[i]; t < s->nl_off[i + 1]; ++t) {
                const int j = s->nl_j[t];
                float dx = mi1(xi - s->x[3 * j], s->L[0]);
                float dy = mi1(yi - s->x[3 * j + 1], s->L[1]);
                float dz = mi1(zi - s->x[3 * j + 2], s->L[2]);
                float r2 = dx * dx + dy * dy + dz * dz;
                if (r2 > s->rc2 || excluded(s, i, j)) continue;
                float inv_r = 1.0f / sqrtf(r2), inv_r2 = inv_r * inv_r;
                float xx = r2 * inv_r * s->inv_h;
                if (xx > CS_BINS - 0.01f) xx = CS_BINS - 0.01f;
                int b = (int)xx;
                float u = xx - b;
                const float *g = s->tab[0][b];
                float factor = kqi * s->q[j] * (((g[3] * u + g[2]) * u + g[1]) * u + g[0]);
                const int p = ti + s->type[j];
                if (s->c6[p] != 0.0f || s->c12[p] != 0.0f) {
                    float ir6 = inv_r2 * inv_r2 * inv_r2;
                    float e_lj = (s->c12[p] * ir6 - s->c6[p]) * ir6;
                    float f_lj = (12.0f * s->c12[p] * ir6 - 6.0f * s->c6[p]) * ir6 * inv_r2;
                    float tau = fmaxf((r2 - s->rs2_sw) * s->inv_delta, 0.0f);
                    float omt = 1.0f - tau;
                    factor += omt * omt * (1.0f + 2.0f * tau) * f_lj + 12.0f * s->inv_delta * tau * omt * e_lj;
                }
                fx += factor * dx; fy += factor * dy; fz += factor * dz;
                fb[3 * j] -= factor * dx; fb[3 * j + 1] -= factor * dy; fb[3 * j + 2] -= factor * dz;
            }
            fb[3 * i] += fx; fb[3 * i + 1] += fy; fb[3 * i + 2] += fz;
        }
        #pragma omp for schedule(static)
        for (int k = 0; k < 3 * n; ++k) {
            float acc = 0.0f;
            for (int t = 0; t < s->n_threads; ++t) acc += s->fthr[(size_t)t * 3 * n + k];
            s->f[k] = acc;
        }
    }
}

static inline void bspline4(float t, float w[4], float dw[4])
{
    /* M_4(t + 3 - k) for k = 0..3 (the point g - k gets weight w[k]) */
    float t2 = t * t, t3 = t2 * t, u = 1.0f - t;
    w[0] = t3 / 6.0f;
    w[1] = (1.0f + 3.0f * t + 3.0f * t2 - 3.0f * t3) / 6.0f;
    w[2] = (4.0f - 6.0f * t2 + 3.0f * t3) / 6.0f;
    w[3] = u * u * u / 6.0f;
    dw[0] = 0.5f * t2;
    dw[1] = 0.5f + t - 1.5f * t2;
    dw[2] = -2.0f * t + 1.5f * t2;
    dw[3] = -0.5f * u * u;
}

static void fft_axis(cs_md *s, int axis, int sign)
{
    const int *K = s->K;
    const int n = K[axis];
    const size_t stride = axis == 2 ? 1 : (axis == 1 ? (size_t)K[2] : (size_t)K[1] * K[2]);
    const int o1 = axis == 0 ? K[1] : K[0], o2 = axis == 2 ? K[1] : K[2];
    int logn = 0;
    while ((1 << logn) < n) ++logn;
    #pragma omp parallel
    {
        cs_cpx *buf = malloc(sizeof(cs_cpx) * n);
        #pragma omp for collapse(2) schedule(static)
        for (int a = 0; a < o1; ++a)
            for (int b = 0; b < o2; ++b) {
                size_t base;
                if (axis == 0) base = (size_t)a * K[2] + b;
                else if (axis == 1) base = (size_t)a * K[1] * K[2] + b;
                else base = ((size_t)a * K[1] + b) * K[2];
                for (int k = 0; k < n; ++k) {
                    int r = 0;
                    for (int bit = 0; bit < logn; ++bit) r |= ((k >> bit) & 1) << (logn - 1 - bit);
                    buf[r] = s->mesh[base + (size_t)k * stride];
                }
                for (int len = 2; len <= n; len <<= 1) {
                    const int step = s->kmax / len;
                    for (int i0 = 0; i0 < n; i0 += len)
                        for (int k = 0; k < len / 2; ++k) {
                            float wr = s->twr[k * step], wi = sign * s->twi[k * step];
                            cs_cpx u = buf[i0 + k], v = buf[i0 + k + len / 2];
                            cs_cpx t = {v.re * wr - v.im * wi, v.re * wi + v.im * wr};
                            buf[i0 + k] = (cs_cpx){u.re + t.re, u.im + t.im};
                            buf[i0 + k + len / 2] = (cs_cpx){u.re - t.re, u.im - t.im};
                        }
                }
                for (int k = 0; k < n; ++k) s->mesh[base + (size_t)k * stride] = buf[k];
            }
        free(buf);
    }
}

static void spme(cs_md *s)
{
    const int n = s->n, *K = s->K;
    const size_t M = (size_t)K[0] * K[1] * K[2];
    memset(s->mesh, 0, sizeof(cs_cpx) * M);
    /* spread: atoms of even then odd 8-plane slabs along a (no two threads
     * write the same point: supports span 4 planes) */
    const int slab = 8, nslab = K[0] / slab;
    for (int parity = 0; parity < 2; ++parity) {
        #pragma omp parallel for schedule(dynamic, 1)
        for (int sl = parity; sl < nslab; sl += 2) {
            for (int i = 0; i < n; ++i) {
                float u0 = (s->x[3 * i] / s->L[0] + 0.5f) * K[0];
                int g0 = (int)floorf(u0);
                int gw = ((g0 % K[0]) + K[0]) % K[0];
                if (gw / slab != sl) continue;
                float w[3][4], dw[4], t[3];
                int g[3];
                for (int d = 0; d < 3; ++d) {
                    float uu = (s->x[3 * i + d] / s->L[d] + 0.5f) * K[d];
                    float fl = floorf(uu);
                    t[d] = uu - fl;
                    g[d] = (((int)fl % K[d]) + K[d]) % K[d];
                    bspline4(t[d], w[d], dw);
                }
                for (int a = 0; a < 4; ++a) {
                    int ia = (g[0] - a + K[0]) % K[0];
                    for (int b = 0; b < 4; ++b) {
                        int ib = (g[1] - b + K[1]) % K[1];
                        float qab = s->q[i] * w[0][a] * w[1][b];
                        for (int c = 0; c < 4; ++c) {
                            int ic = (g[2] - c + K[2]) % K[2];
                            s->mesh[((size_t)ia * K[1] + ib) * K[2] + ic].re += qab * w[2][c];
                        }
                    }
                }
            }
        }
    }
    for (int d = 0; d < 3; ++d) fft_axis(s, d, -1);
    #pragma omp parallel for schedule(static)
    for (size_t k = 0; k < M; ++k) { s->mesh[k].re *= s->G[k]; s->mesh[k].im *= s->G[k]; }
    for (int d = 0; d < 3; ++d) fft_axis(s, d, +1);
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        float w[3][4], dwv[3][4];
        int g[3];
        for (int d = 0; d < 3; ++d) {
            float uu = (s->x[3 * i + d] / s->L[d] + 0.5f) * K[d];
            float fl = floorf(uu);
            g[d] = (((int)fl % K[d]) + K[d]) % K[d];
            bspline4(uu - fl, w[d], dwv[d]);
        }
        float ga = 0, gb = 0, gc = 0;
        for (int a = 0; a < 4; ++a) {
            int ia = (g[0] - a + K[0]) % K[0];
            for (int b = 0; b < 4; ++b) {
                int ib = (g[1] - b + K[1]) % K[1];
                for (int c = 0; c < 4; ++c) {
                    int ic = (g[2] - c + K[2]) % K[2];
                    float v = s->mesh[((size_t)ia * K[1] + ib) * K[2] + ic].re;
                    ga += v * dwv[0][a] * w[1][b] * w[2][c];
                    gb += v * w[0][a] * dwv[1][b] * w[2][c];
                    gc += v * w[0][a] * w[1][b] * dwv[2][c];
                }
            }
        }
        s->f[3 * i] -= s->q[i] * ga * K[0] / s->L[0];
        s->f[3 * i + 1] -= s->q[i] * gb * K[1] / s->L[1];
        s->f[3 * i + 2] -= s->q[i] * gc * K[2] / s->L[2];
    }
}

static void forces(cs_md *s)
{
    if (needs_rebuild(s)) build_list(s);
    pair_forces(s);
    spme(s);
}

/* compensated or f32 kick (v += F/m dt/2) and drift (x += v dt), wrapped */
static void kick(cs_md *s)
{
    const float h = 0.5f * s->dt;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < 3 * s->n; ++k) {
        float a = s->f[k] / s->m[k / 3];
        if (s->compensated) {
            double ev = (double)s->v[k] + s->vlo[k] + (double)a * (double)h;
            s->v[k] = (float)ev;
            s->vlo[k] = ev - (double)s->v[k];
        } else {
            s->v[k] = fmaf(a, h, s->v[k]);
        }
    }
}

static void drift(cs_md *s)
{
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < 3 * s->n; ++k) {
        const float L = s->L[k % 3];
        if (s->compensated) {
            double ex = (double)s->x[k] + s->xlo[k] + ((double)s->v[k] + s->vlo[k]) * (double)s->dt;
            ex -= L * floor(ex / L + 0.5);
            s->x[k] = (float)ex;
            s->xlo[k] = ex - (double)s->x[k];
        } else {
            float x = fmaf(s->v[k], s->dt, s->x[k]);
            s->x[k] = x - L * floorf(x / L + 0.5f);
        }
    }
}

static void settle_positions(cs_md *s)
{
    const double L[6] = {s->L[0], s->L[1], s->L[2], 0, 0, 0};
    #pragma omp parallel for schedule(static)
    for (int g = 0; g < s->n_groups; ++g) {
        double cur[9], v[9], vir[3];
        for (int a = 0; a < 3; ++a)
            for (int d = 0; d < 3; ++d) {
                size_t k = 3 * (size_t)s->groups[3 * g + a] + d;
                cur[3 * a + d] = (double)s->x[k] + s->xlo[k];
                v[3 * a + d] = (double)s->v[k] + s->vlo[k];
            }
        orc_settle_positions_one(s->snap + 9 * (size_t)g, cur, v, s->m_o, s->m_h, s->ra, s->rb,
                                 s->rcs, s->dt, L, vir);
        for (int a = 0; a < 3; ++a)
            for (int d = 0; d < 3; ++d) {
                size_t k = 3 * (size_t)s->groups[3 * g + a] + d;
                double L1 = s->L[d], xx = cur[3 * a + d];
                xx -= L1 * floor(xx / L1 + 0.5);
                s->x[k] = (float)xx;
                s->v[k] = (float)v[3 * a + d];
                if (s->compensated) {
                    s->xlo[k] = xx - (double)s->x[k];
                    s->vlo[k] = v[3 * a + d] - (double)s->v[k];
                }
            }
    }
}

static void settle_velocities(cs_md *s)
{
    const double L[6] = {s->L[0], s->L[1], s->L[2], 0, 0, 0};
    #pragma omp parallel for schedule(static)
    for (int g = 0; g < s->n_groups; ++g) {
        double pos[9], v[9];
        for (int a = 0; a < 3; ++a)
            for (int d = 0; d < 3; ++d) {
                size_t k = 3 * (size_t)s->groups[3 * g + a] + d;
                pos[3 * a + d] = (double)s->x[k] + s->xlo[k];
                v[3 * a + d] = (double)s->v[k] + s->vlo[k];
            }
        orc_settle_velocities_one(pos, v, s->m_o, s->m_h, L, 0.0, NULL);
        for (int a = 0; a < 3; ++a)
            for (int d = 0; d < 3; ++d) {
                size_t k = 3 * (size_t)s->groups[3 * g + a] + d;
                s->v[k] = (float)v[3 * a + d];
                if (s->compensated) s->vlo[k] = v[3 * a + d] - (double)s->v[k];
            }
    }
}

void *cs_create(int n, const float *posq, const float *vel, const float *mass, const int *type,
                int n_types, const float *c6, const float *c12, const float *L3, float rc,
                float r_switch, float r_skin, double alpha, double k_coulomb,
                const int *ex_off, const int *ex_par, const int *mesh, int n_groups,
                const int *groups, double m_o, double m_h, double ra, double rb, double rcs,
                float dt, int compensated)
{
    cs_md *s = calloc(1, sizeof(cs_md));
    s->n = n; s->n_types = n_types; s->n_groups = n_groups; s->compensated = compensated;
    for (int d = 0; d < 3; ++d) { s->L[d] = L3[d]; s->K[d] = mesh[d]; }
    s->x = malloc(sizeof(float) * 3 * n); s->v = malloc(sizeof(float) * 3 * n);
    s->f = calloc(3 * (size_t)n, sizeof(float)); s->q = malloc(sizeof(float) * n);
    s->m = malloc(sizeof(float) * n); s->type = malloc(sizeof(int) * n);
    s->xlo = calloc(3 * (size_t)n, sizeof(double)); s->vlo = calloc(3 * (size_t)n, sizeof(double));
    for (int i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) { s->x[3 * i + d] = posq[4 * i + d]; s->v[3 * i + d] = vel[3 * i + d]; }
        s->q[i] = posq[4 * i + 3]; s->m[i] = mass[i]; s->type[i] = type[i];
    }
    s->c6 = malloc(sizeof(float) * n_types * n_types); s->c12 = malloc(sizeof(float) * n_types * n_types);
    memcpy(s->c6, c6, sizeof(float) * n_types * n_types);
    memcpy(s->c12, c12, sizeof(float) * n_types * n_types);
    s->rc2 = rc * rc; s->rs2_sw = r_switch * r_switch;
    s->inv_delta = r_switch < rc ? 1.0f / (rc * rc - r_switch * r_switch) : 0.0f;
    if (!(r_switch < rc)) s->rs2_sw = rc * rc;
    s->kc = (float)k_coulomb; s->dt = dt;
    build_table(s, alpha, rc);
    s->ex_off = malloc(sizeof(int) * (n + 1));
    memcpy(s->ex_off, ex_off, sizeof(int) * (n + 1));
    s->ex_par = malloc(sizeof(int) * (ex_off[n] > 0 ? ex_off[n] : 1));
    memcpy(s->ex_par, ex_par, sizeof(int) * ex_off[n]);
    s->r_search = rc + r_skin;
    s->skin_half2 = 0.25f * r_skin * r_skin;
    s->xref = malloc(sizeof(float) * 3 * n);
    s->nl_off = malloc(sizeof(int) * (n + 1));
    s->n_threads = omp_get_max_threads();
    s->fthr = malloc(sizeof(float) * 3 * (size_t)n * s->n_threads);
    size_t M = (size_t)mesh[0] * mesh[1] * mesh[2];
    s->G = malloc(sizeof(float) * M);
    s->mesh = malloc(sizeof(cs_cpx) * M);
    s->kmax = mesh[0] > mesh[1] ? mesh[0] : mesh[1];
    if (mesh[2] > s->kmax) s->kmax = mesh[2];
    s->twr = malloc(sizeof(float) * s->kmax); s->twi = malloc(sizeof(float) * s->kmax);
    for (int k = 0; k < s->kmax; ++k) {
        s->twr[k] = (float)cos(2 * CS_PI * k / s->kmax);
        s->twi[k] = (float)sin(2 * CS_PI * k / s->kmax);
    }
    build_influence(s, alpha);
    s->groups = malloc(sizeof(int) * 3 * (n_groups > 0 ? n_groups : 1));
    memcpy(s->groups, groups, sizeof(int) * 3 * n_groups);
    s->m_o = m_o; s->m_h = m_h; s->ra = ra; s->rb = rb; s->rcs = rcs;
    s->snap = malloc(sizeof(double) * 9 * (n_groups > 0 ? n_groups : 1));
    build_list(s);
    s->n_rebuilds = 0;
    pair_forces(s);
    spme(s);
    return s;
}

/* n velocity-Verlet steps with SETTLE: snapshot, kick, drift, SETTLE x,
 * forces, kick, SETTLE v. */
void cs_step(void *h, int n_steps)
{
    cs_md *s = h;
    for (int t = 0; t < n_steps; ++t) {
        #pragma omp parallel for schedule(static)
        for (int g = 0; g < s->n_groups; ++g)
            for (int a = 0; a < 3; ++a)
                for (int d = 0; d < 3; ++d)
                    s->snap[9 * (size_t)g + 3 * a + d] = s->x[3 * s->groups[3 * g + a] + d];
        kick(s);
        drift(s);
        settle_positions(s);
        forces(s);
        kick(s);
        settle_velocities(s);
    }
}

int cs_rebuilds(void *h) { return ((cs_md *)h)->n_rebuilds; }

void cs_set_threads(int n) { omp_set_num_threads(n); }

int cs_max_threads(void) { return omp_get_max_threads(); }

void cs_forces(void *h, float *f) { memcpy(f, ((cs_md *)h)->f, sizeof(float) * 3 * ((cs_md *)h)->n); }

void cs_destroy(void *h)
{
    cs_md *s = h;
    if (!s) return;
    free(s->x); free(s->v); free(s->f); free(s->q); free(s->m); free(s->type);
    free(s->xlo); free(s->vlo); free(s->c6); free(s->c12); free(s->ex_off); free(s->ex_par);
    free(s->xref); free(s->nl_off); free(s->nl_j); free(s->fthr); free(s->G); free(s->mesh);
    free(s->twr); free(s->twi); free(s->groups); free(s->snap);
    free(s);
}
