/*
 * TEST INFRASTRUCTURE ONLY -- the oracle of the unbinding path.
 *
 * Plain-C restatement of the reference's self-bound check for FoF candidates
 * (bound_fortran.f90 module bound_test, driven from fof_fortran.f90:253-360).
 * Used only by tests/, the smoke check and the CPU-baseline leg of bench.py.
 *
 * Follows (in meaning, not text):
 *   bound_fortran.f90:291-330    bound_set_context          (context, ngrid, bn, s_15)
 *   bound_fortran.f90:1854-2040  grid levels, ladder, band candidate lists, selection
 *   bound_fortran.f90:2901-3073  bound_classify             (span, 15-Mpc rule, origin)
 *   bound_fortran.f90:2300-2388  Green function + 5^3 self kernel (alias sum in f64)
 *   bound_fortran.f90:2436-2888  bound_check_grid           (setup, passes, removal loop)
 *   bound_fortran.f90:3080-3117  continuous softening depth and C2 difference
 *   bound_fortran.f90:3125-3552  C2 cell CSR, full and incremental residual
 *   bound_fortran.f90:3563-3583  wave numbers, TSC window
 *   fof_fortran.f90:263-356      packing of candidates and bound write-back
 *
 * Differences from the reference, all documented where they occur:
 *   - The FFT is this file's own mixed-radix complex FFT in double precision
 *     (the reference links FFTW3f, single precision). PM potentials therefore
 *     differ from the reference at the float rounding level.
 *   - The two-level refinement of 256^3 halos (bound_fortran.f90:403-1517) is
 *     restated (tl_*): block ratios, inner band grids, per-target cutoffs.
 *     Its short-range pair sums gather per target over 27 cells of width
 *     3.5 sp instead of the reference's task order (f32 sum order only).
 * Everything else (direct summation order, float/double mix, removal order,
 * statuses) follows the reference statement by statement.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

#define BPI 3.141592653589793238462643383279502884197

/* ------------------------------------------------------------------ */
/* small complex FFT (double), mixed radix 2/3/4, any n = 2^a 3^b       */

typedef struct { double re, im; } cpx;

/* twiddle table exp(-2 pi i k / N), k < N, in double */
static cpx *fft_table(int N)
{
    cpx *t = (cpx *)malloc(sizeof(cpx) * (size_t)N);
    for (int k = 0; k < N; ++k) {
        double ang = -2.0 * BPI * (double)k / (double)N;
        t[k].re = cos(ang);
        t[k].im = sin(ang);
    }
    return t;
}

/* exp(sign 2 pi i k / N) from the table (sign -1 forward, +1 inverse) */
static inline cpx tw_at(const cpx *t, int N, long k, int sign)
{
    cpx w = t[k % N];
    if (sign > 0) w.im = -w.im;
    return w;
}

static void fft_rec(cpx *x, cpx *tmp, int n, int sign, const cpx *t, int N)
{
    /* in-place recursive decimation in time on x[0..n), n | N */
    if (n == 1) return;
    int r = (n % 4 == 0) ? 4 : (n % 2 == 0 ? 2 : (n % 3 == 0 ? 3 : n));
    int m = n / r;
    for (int q = 0; q < r; ++q)
        for (int k = 0; k < m; ++k) tmp[q * m + k] = x[k * r + q];
    memcpy(x, tmp, sizeof(cpx) * (size_t)n);
    for (int q = 0; q < r; ++q) fft_rec(x + q * m, tmp, m, sign, t, N);
    const int step = N / n, rstep = N / r;
    for (int k = 0; k < m; ++k) {
        cpx v[4];
        for (int q = 0; q < r; ++q) {
            cpx a = x[q * m + k];
            cpx w = tw_at(t, N, (long)q * k * step, sign);
            v[q].re = a.re * w.re - a.im * w.im;
            v[q].im = a.re * w.im + a.im * w.re;
        }
        for (int j = 0; j < r; ++j) {
            cpx acc = {0.0, 0.0};
            for (int q = 0; q < r; ++q) {
                cpx w = tw_at(t, N, (long)q * j * rstep, sign);
                acc.re += v[q].re * w.re - v[q].im * w.im;
                acc.im += v[q].re * w.im + v[q].im * w.re;
            }
            tmp[j * m + k] = acc;
        }
    }
    memcpy(x, tmp, sizeof(cpx) * (size_t)n);
}

/* 3-D complex FFT of an n^3 grid (x fastest). sign -1 forward, +1 inverse
 * (unnormalised). Lines are independent: OpenMP over them. */
static void fft3(cpx *g, int n, int sign)
{
    cpx *t = fft_table(n);
    for (int axis = 0; axis < 3; ++axis) {
        const size_t st = axis == 0 ? 1 : (axis == 1 ? (size_t)n : (size_t)n * n);
        #pragma omp parallel
        {
            cpx *line = (cpx *)malloc(sizeof(cpx) * n);
            cpx *tmp = (cpx *)malloc(sizeof(cpx) * n);
            #pragma omp for schedule(static)
            for (long ab = 0; ab < (long)n * n; ++ab) {
                const int a = (int)(ab % n), b = (int)(ab / n);
                size_t base;
                if (axis == 0) base = (size_t)n * (a + (size_t)n * b);
                else if (axis == 1) base = (size_t)a + (size_t)n * n * b;
                else base = (size_t)a + (size_t)n * b;
                for (int i = 0; i < n; ++i) line[i] = g[base + st * i];
                fft_rec(line, tmp, n, sign, t, n);
                for (int i = 0; i < n; ++i) g[base + st * i] = line[i];
            }
            free(line);
            free(tmp);
        }
    }
    free(t);
}

/* ------------------------------------------------------------------ */
/* context and level selection (bound_fortran.f90:291-330, 1854-2040)   */

typedef struct {
    int ngrid, bn[4], n_max, m_max;
    double app, a, b, s15, spacing_max;
    int np_min;
    int fgl_nlev[4];
    double fgl_sp[4][32], fgl_cov[4][32];
    int ntot, tab_gid[64], tab_lev[64];
    double tab_sp[64], tab_cov[64];
} bctx_t;

static void grid_levels(const bctx_t *c, int gid, int *nlev, double *sp_out)
{
    *nlev = 0;
    if (gid < 1 || gid > c->ngrid) return;
    int m = c->bn[gid - 1] / 2 - 3;
    double sp = 1.0 / 32.0;
    while ((double)m * sp < c->s15 && *nlev < 32) {
        sp_out[(*nlev)++] = sp;
        sp = 2.0 * sp;
    }
    if (gid == c->ngrid && *nlev < 32) sp_out[(*nlev)++] = c->spacing_max;
}

static void bctx_init(bctx_t *c, double app, double a, double b, double s15, int np_min, int64_t ng)
{
    static const int cand[4] = {48, 64, 128, 256};
    memset(c, 0, sizeof(*c));
    int ngi = (int)ng;
    c->ngrid = 1 + (ngi >= 64) + (ngi >= 128) + (ngi >= 256);
    for (int g = 0; g < 4; ++g) c->bn[g] = cand[g] < ngi ? cand[g] : ngi;
    c->n_max = ngi < 256 ? ngi : 256;
    c->m_max = c->n_max / 2 - 3;
    c->app = app; c->a = a; c->b = b; c->s15 = s15; c->np_min = np_min;
    c->spacing_max = s15 / (double)c->m_max;
    for (int g = 1; g <= c->ngrid; ++g) {
        int nl;
        double sp[32];
        grid_levels(c, g, &nl, sp);
        c->fgl_nlev[g - 1] = nl;
        for (int l = 0; l < nl; ++l) {
            c->fgl_sp[g - 1][l] = sp[l];
            c->fgl_cov[g - 1][l] = (double)(c->bn[g - 1] / 2 - 3) * sp[l];
        }
    }
    /* global ladder by coverage, selection sort as bound_build_ladder */
    int nt = 0;
    for (int g = 1; g <= c->ngrid; ++g)
        for (int l = 0; l < c->fgl_nlev[g - 1]; ++l) {
            c->tab_gid[nt] = g; c->tab_lev[nt] = l + 1;
            c->tab_sp[nt] = c->fgl_sp[g - 1][l];
            c->tab_cov[nt] = (double)(c->bn[g - 1] / 2 - 3) * c->fgl_sp[g - 1][l];
            ++nt;
        }
    for (int j = 0; j < nt - 1; ++j) {
        int k = j;
        for (int i = j + 1; i < nt; ++i) if (c->tab_cov[i] < c->tab_cov[k]) k = i;
        if (k != j) {
            int t; double d;
            t = c->tab_gid[j]; c->tab_gid[j] = c->tab_gid[k]; c->tab_gid[k] = t;
            t = c->tab_lev[j]; c->tab_lev[j] = c->tab_lev[k]; c->tab_lev[k] = t;
            d = c->tab_cov[j]; c->tab_cov[j] = c->tab_cov[k]; c->tab_cov[k] = d;
            d = c->tab_sp[j]; c->tab_sp[j] = c->tab_sp[k]; c->tab_sp[k] = d;
        }
    }
    c->ntot = nt;
}

/* bound_select_grid_and_level: returns 0 ok (gid 0 = direct), 2 no combination */
int oracle_select_grid(const void *cv, int64_t np, double span, int *gid, int *lev, double *sp_out)
{
    const bctx_t *c = (const bctx_t *)cv;
    static const int ncand[3] = {11, 10, 8};
    static const int cg[3][12] = {{1, 2, 2, 3, 3, 3, 4, 4, 4, 4, 4},
                                  {3, 3, 4, 4, 4, 4, 4, 4, 4, 4},
                                  {4, 4, 4, 4, 4, 4, 4, 4}};
    static const double cs[3][12] = {
        {0.125, 0.125, 0.25, 0.125, 0.25, 0.5, 0.0625, 0.125, 0.25, 0.5, -1.0},
        {0.0625, 0.125, 0.0625, 0.125, 0.25, 0.5, 1.0, 2.0, 4.0, -1.0},
        {0.0625, 0.125, 0.25, 0.5, 1.0, 2.0, 4.0, -1.0}};
    *gid = 0; *lev = 0; *sp_out = 0.0;
    if (np <= 1000) return 0;
    int band = np <= 5000 ? 0 : (np <= 20000 ? 1 : 2);
    for (int s = 0; s < ncand[band]; ++s) {
        int g = cg[band][s];
        if (g < 1 || g > c->ngrid) continue;
        double sp = cs[band][s];
        if (sp == -1.0) sp = c->spacing_max;
        int l = 0;
        for (int k = 0; k < c->fgl_nlev[g - 1]; ++k) {
            if (fabs(c->fgl_sp[g - 1][k] - sp) <= 1e-12 * fmax(1.0, fabs(sp))) { l = k + 1; break; }
        }
        if (l == 0) continue;
        if (c->fgl_cov[g - 1][l - 1] < span) continue;
        *gid = g; *lev = l; *sp_out = c->fgl_sp[g - 1][l - 1];
        return 0;
    }
    if (span <= c->s15) {
        for (int i = 0; i < c->ntot; ++i) {
            if (c->tab_cov[i] >= span) {
                int g = c->tab_gid[i], l = c->tab_lev[i];
                if (g >= 1 && g <= c->ngrid && l >= 1) {
                    *gid = g; *lev = l; *sp_out = c->fgl_sp[g - 1][l - 1];
                    return 0;
                }
                break;
            }
        }
    }
    return 2;
}

/* exposed so tests can check the selection rule */
int32_t oracle_select(double app, double s15, int64_t ng, int64_t np, double span,
                      int32_t *gid, int32_t *lev, double *sp)
{
    bctx_t c;
    bctx_init(&c, app, 1.0, 0.0, s15, 10, ng);
    int g, l;
    int st = oracle_select_grid(&c, np, span, &g, &l, sp);
    *gid = g; *lev = l;
    return st;
}

/* ------------------------------------------------------------------ */
/* Green function and self kernel (bound_fortran.f90:2300-2388, 3563-3583) */

static double wave_number(int index_1based, int n)
{
    int mode = (((index_1based - 1 + n / 2) % n) + n) % n - n / 2;
    return 2.0 * BPI * (double)mode / (double)n;
}

static double window_sq(const double k[3])
{
    double sp = 1.0;
    for (int ax = 0; ax < 3; ++ax) {
        double x = 0.5 * k[ax];
        if (x != 0.0) sp = sp * sin(x) / x;
    }
    double s2 = sp * sp;
    return s2 * s2 * s2;  /* sinc_product**(2*2+2) */
}

static double shape_s(double y)
{
    /* 12 y^-4 (2 - 2 cos y - y sin y) */
    double y4 = y * y * y * y;
    return 12.0 * (1.0 / y4) * (2.0 - 2.0 * cos(y) - y * sin(y));
}

/* G on the half spectrum [n][n][n/2+1] (float) and the 5^3 self kernel.
 * Outer grids: apm_soft = 3.5, apm_range = 0, truncated (cutoff 1 - sin(kR)/kR). */
/* Green function and 5^3 self kernel (bound_fortran.f90:2300-2382) of a grid
 * of n, split softening `soft`, band range `range` (0: the outer grid's
 * kernel) and truncation at R = n/2 (`truncated`). */
void oracle_green_ex(int32_t n, double soft, double range, int truncated, float *G, float *self125)
{
    const double apm_soft = soft, radius = 0.5 * (double)n;
    const int nh = n / 2 + 1;
    #pragma omp parallel for schedule(dynamic)
    for (int k = 1; k <= n; ++k)
        for (int j = 1; j <= n; ++j)
            for (int i = 1; i <= nh; ++i) {
                double kv[3] = {wave_number(i, n), wave_number(j, n), wave_number(k, n)};
                double asum = 0.0, bsum = 0.0;
                for (int ka = -2; ka <= 2; ++ka)
                    for (int ja = -2; ja <= 2; ++ja)
                        for (int ia = -2; ia <= 2; ++ia) {
                            double kal[3] = {kv[0] + 2.0 * BPI * ia, kv[1] + 2.0 * BPI * ja,
                                             kv[2] + 2.0 * BPI * ka};
                            double kmag = sqrt(kal[0] * kal[0] + kal[1] * kal[1] + kal[2] * kal[2]);
                            double uk2 = window_sq(kal);
                            asum += uk2;
                            double target;
                            if (kmag == 0.0) {
                                if (range == 0.0) target = -4.0 * BPI * radius * radius / 6.0;
                                else if (truncated) target = 0.0;
                                else target = -2.0 * BPI * (range * range - apm_soft * apm_soft) / 15.0;
                            } else {
                                double ss = shape_s(apm_soft * kmag / 2.0);
                                double sr = range == 0.0 ? 0.0 : shape_s(range * kmag / 2.0);
                                double kr = kmag * radius;
                                double cut = truncated ? 1.0 - sin(kr) / kr : 1.0;
                                target = -4.0 * BPI * (ss * ss - sr * sr) * cut / (kmag * kmag);
                            }
                            bsum += uk2 * target;
                        }
                G[(size_t)(i - 1) + (size_t)nh * ((j - 1) + (size_t)n * (k - 1))] =
                    (float)(bsum / (asum * asum));
            }
    /* self kernel: inverse transform of G (Hermitian extension), / n^3 */
    cpx *g = (cpx *)calloc((size_t)n * n * n, sizeof(cpx));
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) {
                int hx = x, hy = y, hz = z;
                if (x >= nh) { hx = n - x; hy = (n - y) % n; hz = (n - z) % n; }
                g[(size_t)x + (size_t)n * (y + (size_t)n * z)].re =
                    G[(size_t)hx + (size_t)nh * (hy + (size_t)n * hz)];
            }
    fft3(g, n, +1);
    double inv = 1.0 / ((double)n * n * n);
    for (int iz = -2; iz <= 2; ++iz)
        for (int iy = -2; iy <= 2; ++iy)
            for (int ix = -2; ix <= 2; ++ix) {
                int x = (ix + n) % n, y = (iy + n) % n, z = (iz + n) % n;
                self125[(ix + 2) + 5 * ((iy + 2) + 5 * (iz + 2))] =
                    (float)(g[(size_t)x + (size_t)n * (y + (size_t)n * z)].re * inv);
            }
    free(g);
}

void oracle_green(int32_t n, float *G, float *self125)
{
    oracle_green_ex(n, 3.5, 0.0, 1, G, self125);
}

/* ------------------------------------------------------------------ */
/* depth kernels (bound_fortran.f90:3080-3117)                          */

double oracle_depth(double r, double h)
{
    if (r <= 0.0) return 208.0 / (70.0 * h);
    if (r >= h) return 1.0 / r;
    double e = 2.0 * r / h;
    double prim;
    if (e <= 1.0) {
        prim = e * e * (112.0 + e * e * (-56.0 + e * (14.0 + e * (8.0 - 3.0 * e))));
        return (208.0 - prim) / (70.0 * h);
    }
    prim = e * e * (448.0 + e * (-280.0 + e * (56.0 + e * (14.0 + e * (-8.0 + e))))) - 224.0 * e -
           12.0 / e;
    return (128.0 - prim) / (70.0 * h);
}

static double c2_delta_depth(double r, double ht, double hs)
{
    if (r <= 0.0) return (208.0 / 70.0) * (1.0 / ht - 1.0 / hs);
    if (r >= ht && r >= hs) return 0.0;
    return oracle_depth(r, ht) - oracle_depth(r, hs);
}

/* ------------------------------------------------------------------ */
/* per-halo solver (bound_fortran.f90:2436-2888)                        */

typedef struct {
    int ncell[3], nt;
    double umin[3];
    int *cell_id;     /* [n], 1-based cell */
    int *cell_start;  /* [nt+1], 0-based */
    int *cell_part;   /* [n], local indices */
} c2cells_t;

static void build_cells(c2cells_t *cc, int n, const float *q, float sp)
{
    const double cw = 3.5;
    double umin[3] = {1e30, 1e30, 1e30}, umax[3] = {-1e30, -1e30, -1e30};
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            double u = (double)q[3 * i + d] / (double)sp;
            umin[d] = fmin(umin[d], u);
            umax[d] = fmax(umax[d], u);
        }
    for (int d = 0; d < 3; ++d) {
        int nc = (int)floor((umax[d] - umin[d]) / cw) + 1;
        cc->ncell[d] = nc < 1 ? 1 : nc;
        cc->umin[d] = umin[d];
    }
    cc->nt = cc->ncell[0] * cc->ncell[1] * cc->ncell[2];
    cc->cell_id = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    cc->cell_start = (int *)calloc((size_t)cc->nt + 1, sizeof(int));
    cc->cell_part = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) {
        int ic[3];
        for (int d = 0; d < 3; ++d) {
            double u = (double)q[3 * i + d] / (double)sp;
            int v = (int)floor((u - umin[d]) / cw) + 1;
            if (v < 1) v = 1;
            if (v > cc->ncell[d]) v = cc->ncell[d];
            ic[d] = v;
        }
        int c = ic[0] + (ic[1] - 1) * cc->ncell[0] + (ic[2] - 1) * cc->ncell[0] * cc->ncell[1];
        cc->cell_id[i] = c;
        cc->cell_start[c] += 1;
    }
    /* exclusive prefix: cell_start[c-1] = start of cell c (0-based) */
    int acc = 0;
    for (int c = 0; c < cc->nt; ++c) {
        int cnt = cc->cell_start[c + 1];
        cc->cell_start[c] = acc;
        acc += cnt;
    }
    cc->cell_start[cc->nt] = acc;
    int *cur = (int *)malloc(sizeof(int) * (size_t)cc->nt);
    for (int c = 0; c < cc->nt; ++c) cur[c] = cc->cell_start[c];
    for (int i = 0; i < n; ++i) cc->cell_part[cur[cc->cell_id[i] - 1]++] = i;
    free(cur);
}

static void free_cells(c2cells_t *cc)
{
    free(cc->cell_id); free(cc->cell_start); free(cc->cell_part);
}

static const int c2_stencil[13][3] = {{-1, -1, 1}, {0, -1, 1}, {1, -1, 1}, {-1, 0, 1}, {0, 0, 1},
                                      {1, 0, 1},   {-1, 1, 1}, {0, 1, 1},  {1, 1, 1},  {-1, 1, 0},
                                      {0, 1, 0},   {1, 1, 0},  {1, 0, 0}};

static void c2_full(const c2cells_t *cc, int n, const float *q, const char *keep, const float *w,
                    float sp, float app_val, double a, float *phi_c2)
{
    for (int i = 0; i < n; ++i) phi_c2[i] = 0.0f;
    int nk = 0;
    for (int i = 0; i < n; ++i) nk += keep[i] ? 1 : 0;
    if (nk <= 1 || !(sp > 0.0f)) return;
    double ht = (double)app_val / (double)sp, hs = 3.5;
    double r2cut = 12.25 * (double)sp * (double)sp;
    float pref = (float)(a / (6.0 * BPI * (double)sp));
    const int *nc = cc->ncell;
    for (int cz = 1; cz <= nc[2]; ++cz)
        for (int cy = 1; cy <= nc[1]; ++cy)
            for (int cx = 1; cx <= nc[0]; ++cx) {
                int c1 = cx + (cy - 1) * nc[0] + (cz - 1) * nc[0] * nc[1];
                int s1 = cc->cell_start[c1 - 1], e1 = cc->cell_start[c1];
                for (int p1 = s1; p1 < e1; ++p1) {
                    int i1 = cc->cell_part[p1];
                    if (!keep[i1]) continue;
                    for (int pass = 0; pass < 14; ++pass) {
                        int b0, b1;
                        if (pass == 0) { b0 = p1 + 1; b1 = e1; }
                        else {
                            int nx = cx + c2_stencil[pass - 1][0], ny = cy + c2_stencil[pass - 1][1],
                                nz = cz + c2_stencil[pass - 1][2];
                            if (nx < 1 || nx > nc[0] || ny < 1 || ny > nc[1] || nz < 1 || nz > nc[2]) continue;
                            int c2i = nx + (ny - 1) * nc[0] + (nz - 1) * nc[0] * nc[1];
                            b0 = cc->cell_start[c2i - 1]; b1 = cc->cell_start[c2i];
                        }
                        for (int p2 = b0; p2 < b1; ++p2) {
                            int i2 = cc->cell_part[p2];
                            if (!keep[i2]) continue;
                            double dx = (double)(q[3 * i2] - q[3 * i1]);
                            double dy = (double)(q[3 * i2 + 1] - q[3 * i1 + 1]);
                            double dz = (double)(q[3 * i2 + 2] - q[3 * i1 + 2]);
                            double r2 = dx * dx + dy * dy + dz * dz;
                            if (r2 < r2cut) {
                                double rs = sqrt(r2) / (double)sp;
                                double dd = c2_delta_depth(rs, ht, hs);
                                float dphi = -pref * (float)dd;
                                phi_c2[i1] = phi_c2[i1] + w[i2] * dphi;
                                phi_c2[i2] = phi_c2[i2] + w[i1] * dphi;
                            }
                        }
                    }
                }
            }
}

static void c2_incremental(const c2cells_t *cc, int n, const float *q, const char *keep,
                           const float *w, const int *just, int njust, float sp, float app_val,
                           double a, float *phi_c2)
{
    int nk = 0;
    for (int i = 0; i < n; ++i) nk += keep[i] ? 1 : 0;
    if (njust <= 0 || nk <= 0 || !(sp > 0.0f)) return;
    double ht = (double)app_val / (double)sp, hs = 3.5;
    double r2cut = 12.25 * (double)sp * (double)sp;
    float pref = (float)(a / (6.0 * BPI * (double)sp));
    const int *nc = cc->ncell;
    for (int k = 0; k < njust; ++k) {
        int j = just[k];
        int cj = cc->cell_id[j];
        int cz = (cj - 1) / (nc[0] * nc[1]) + 1;
        int cy = ((cj - 1) % (nc[0] * nc[1])) / nc[0] + 1;
        int cx = ((cj - 1) % (nc[0] * nc[1])) % nc[0] + 1;
        for (int nz = cz - 1; nz <= cz + 1; ++nz) {
            if (nz < 1 || nz > nc[2]) continue;
            for (int ny = cy - 1; ny <= cy + 1; ++ny) {
                if (ny < 1 || ny > nc[1]) continue;
                for (int nx = cx - 1; nx <= cx + 1; ++nx) {
                    if (nx < 1 || nx > nc[0]) continue;
                    int c2i = nx + (ny - 1) * nc[0] + (nz - 1) * nc[0] * nc[1];
                    for (int p = cc->cell_start[c2i - 1]; p < cc->cell_start[c2i]; ++p) {
                        int i2 = cc->cell_part[p];
                        if (!keep[i2]) continue;
                        double dx = (double)(q[3 * i2] - q[3 * j]);
                        double dy = (double)(q[3 * i2 + 1] - q[3 * j + 1]);
                        double dz = (double)(q[3 * i2 + 2] - q[3 * j + 2]);
                        double r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 < r2cut) {
                            double rs = sqrt(r2) / (double)sp;
                            double dd = c2_delta_depth(rs, ht, hs);
                            float dphi = -pref * (float)dd;
                            phi_c2[i2] = phi_c2[i2] - w[j] * dphi;
                        }
                    }
                }
            }
        }
    }
}

/* Fortran NINT for real(4): round half away from zero */
static int nint_f(float x) { return (int)roundf(x); }

typedef struct {
    int gid, lev;
    double spacing;
    float origin[3];
    float span;
} htask_t;

typedef struct { int n; float *G; float self[125]; } green_t;

/* ------------------------------------------------------------------ */
/* two-level refinement of 256^3 halos (bound_fortran.f90:239-242,       */
/* 403-1517): blocks of 16^3 coarse cells with many members get an inner  */
/* grid (128^3 at ratio 2, 216^3 at ratio 4) carrying the band between    */
/* 3.5 sp / r and 3.5 sp; their members' short-range cutoff drops to      */
/* 3.5 sp / r. phi = C1 + phi_in + C2 (bound_fortran.f90:2744).           */

#define TL_NREG 512
#define TL_N12 3926
#define TL_N24 14688
#define TL_NR2 128
#define TL_NR4 216

typedef struct {
    int n_active;
    int region_r[TL_NREG];
    int active[TL_NREG];
    int *particle_r, *target_reg;
    int *src_start, *src_member, *src_first;
    float *src_wt;
    int *tgt_first;
    float *tgt_wt, *self_in;
} tl_t;

static void tl_free(tl_t *tl)
{
    free(tl->particle_r); free(tl->target_reg); free(tl->src_start); free(tl->src_member);
    free(tl->src_first); free(tl->src_wt); free(tl->tgt_first); free(tl->tgt_wt); free(tl->self_in);
    memset(tl, 0, sizeof(*tl));
}

/* inner band kernels, computed once per process (bound_init_inner_gk) */
static const green_t *tl_green(int r)
{
    static green_t g2, g4;
    green_t *g = r == 2 ? &g2 : &g4;
    if (!g->G) {
        const int n = r == 2 ? TL_NR2 : TL_NR4;
        g->n = n;
        g->G = (float *)malloc(sizeof(float) * (size_t)(n / 2 + 1) * n * n);
        oracle_green_ex(n, 3.5, r == 2 ? 7.0 : 14.0, 0, g->G, g->self);
    }
    return g;
}

/* bound_region_bounds (bound_fortran.f90:531-539); b is 0-based */
static void tl_region_bounds(int b, float lo[3], float hi[3])
{
    const int bb[3] = {b % 8, (b / 8) % 8, b / 64};
    for (int d = 0; d < 3; ++d) {
        lo[d] = (float)(bb[d] * 16);
        hi[d] = fminf(lo[d] + 16.0f, 127.0f);
    }
}

/* TSC on the inner grid: first node and weights; 0 if outside [1, n_in/2] */
static int tl_tsc(const float t[3], const float u0[3], int r, int n_in, int first[3], float wt[9])
{
    for (int ax = 0; ax < 3; ++ax) {
        float coord = (t[ax] - u0[ax]) * (float)r;
        int nearest = nint_f(coord);
        float delta = coord - (float)nearest;
        first[ax] = nearest;
        float a0 = 0.5f - delta, a2 = 0.5f + delta;
        wt[3 * ax + 0] = 0.5f * (a0 * a0);
        wt[3 * ax + 1] = 0.75f - delta * delta;
        wt[3 * ax + 2] = 0.5f * (a2 * a2);
        if (nearest < 1 || nearest + 2 > n_in / 2) return 0;
    }
    return 1;
}

/* bound_two_level_setup_ratios + build_halo (bound_fortran.f90:551-823).
 * Returns 0, or -1 (ierr 1) when a node leaves the inner half domain. */
static int tl_setup(tl_t *tl, int n, const float *q, float spacing, const float origin[3], float app)
{
    memset(tl, 0, sizeof(*tl));
    float *t = (float *)malloc(sizeof(float) * 3 * (size_t)n);
    int cnt[TL_NREG] = {0};
    tl->particle_r = (int *)malloc(sizeof(int) * (size_t)n);
    tl->target_reg = (int *)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; ++i) {
        int bb[3];
        for (int d = 0; d < 3; ++d) {
            t[3 * i + d] = (float)(((double)q[3 * i + d] - (double)origin[d]) / (double)spacing);
            int cell = (int)floor((double)t[3 * i + d]);
            int v = cell / 16;
            bb[d] = v < 0 ? 0 : (v > 7 ? 7 : v);
        }
        tl->target_reg[i] = bb[0] + 8 * bb[1] + 64 * bb[2];
        cnt[tl->target_reg[i]]++;
    }
    for (int b = 0; b < TL_NREG; ++b) {
        int r = cnt[b] < TL_N12 ? 1 : (cnt[b] < TL_N24 ? 2 : 4);
        if (r == 4 && 3.5 * (double)spacing / 4.0 < (double)app) r = 2;
        if (r == 2 && 3.5 * (double)spacing / 2.0 < (double)app) r = 1;
        tl->region_r[b] = r;
        if (cnt[b] > 0 && r > 1) tl->active[tl->n_active++] = b;
    }
    for (int i = 0; i < n; ++i) tl->particle_r[i] = tl->region_r[tl->target_reg[i]];
    if (tl->n_active == 0) { free(t); return 0; }
    /* sources of each active region: members within its band */
    tl->src_start = (int *)calloc((size_t)tl->n_active + 1, sizeof(int));
    int cap = 0;
    for (int k = 0; k < tl->n_active; ++k) {
        int b = tl->active[k], r = tl->region_r[b];
        float lo[3], hi[3];
        tl_region_bounds(b, lo, hi);
        float bo = 3.5f + 3.0f / (float)r;
        for (int j = 0; j < n; ++j) {
            const float *tj = t + 3 * j;
            if (tj[0] >= lo[0] - bo && tj[0] <= hi[0] + bo && tj[1] >= lo[1] - bo &&
                tj[1] <= hi[1] + bo && tj[2] >= lo[2] - bo && tj[2] <= hi[2] + bo) cap++;
        }
    }
    tl->src_member = (int *)malloc(sizeof(int) * (size_t)(cap > 0 ? cap : 1));
    tl->src_first = (int *)malloc(sizeof(int) * 3 * (size_t)(cap > 0 ? cap : 1));
    tl->src_wt = (float *)malloc(sizeof(float) * 9 * (size_t)(cap > 0 ? cap : 1));
    tl->tgt_first = (int *)calloc(3 * (size_t)n, sizeof(int));
    tl->tgt_wt = (float *)calloc(9 * (size_t)n, sizeof(float));
    tl->self_in = (float *)calloc((size_t)n, sizeof(float));
    int p = 0, bad = 0;
    for (int k = 0; k < tl->n_active && !bad; ++k) {
        int b = tl->active[k], r = tl->region_r[b];
        int n_in = r == 2 ? TL_NR2 : TL_NR4;
        float lo[3], hi[3], u0[3];
        tl_region_bounds(b, lo, hi);
        float bo = 3.5f + 3.0f / (float)r;
        for (int d = 0; d < 3; ++d) u0[d] = lo[d] - bo - (1.0f / (float)r) + (0.0f / (float)r);
        tl->src_start[k] = p;
        for (int j = 0; j < n; ++j) {
            const float *tj = t + 3 * j;
            if (!(tj[0] >= lo[0] - bo && tj[0] <= hi[0] + bo && tj[1] >= lo[1] - bo &&
                  tj[1] <= hi[1] + bo && tj[2] >= lo[2] - bo && tj[2] <= hi[2] + bo)) continue;
            tl->src_member[p] = j;
            if (!tl_tsc(tj, u0, r, n_in, tl->src_first + 3 * p, tl->src_wt + 9 * p)) { bad = 1; break; }
            ++p;
        }
        /* targets: members of the region, with the inner self term */
        const green_t *g = tl_green(r);
        for (int i = 0; i < n && !bad; ++i) {
            if (tl->target_reg[i] != b) continue;
            float *w9 = tl->tgt_wt + 9 * i;
            if (!tl_tsc(t + 3 * i, u0, r, n_in, tl->tgt_first + 3 * i, w9)) { bad = 1; break; }
            double acc = 0.0;
            for (int sz = 0; sz < 3; ++sz)
                for (int sy = 0; sy < 3; ++sy)
                    for (int sx = 0; sx < 3; ++sx) {
                        float sw = w9[sx] * w9[3 + sy] * w9[6 + sz];
                        for (int tz = 0; tz < 3; ++tz)
                            for (int ty = 0; ty < 3; ++ty)
                                for (int tx = 0; tx < 3; ++tx) {
                                    float tw = w9[tx] * w9[3 + ty] * w9[6 + tz];
                                    acc = acc + (double)(sw * tw) *
                                                    (double)g->self[(tx - sx + 2) + 5 * ((ty - sy + 2) + 5 * (tz - sz + 2))];
                                }
                    }
            tl->self_in[i] = (float)acc;
        }
    }
    tl->src_start[tl->n_active] = p;
    free(t);
    return bad ? -1 : 0;
}

/* test hook: block ratios of a member set (q f32 xyz, n members) on a
 * 256-grid placement; counts[512] and ratios[512] out; returns the number
 * of active blocks, or -1 when a stencil leaves an inner grid */
int32_t oracle_tl_info(int32_t n, const float *q, float spacing, const float *origin, float app,
                       int32_t *counts, int32_t *ratios)
{
    tl_t tl;
    int rc = tl_setup(&tl, n, q, spacing, origin, app);
    for (int b = 0; b < TL_NREG; ++b) {
        ratios[b] = tl.region_r[b];
        counts[b] = 0;
    }
    if (tl.target_reg)
        for (int i = 0; i < n; ++i) counts[tl.target_reg[i]]++;
    int na = tl.n_active;
    tl_free(&tl);
    return rc != 0 ? -1 : na;
}

/* bound_two_level_run_pm (bound_fortran.f90:825-942): the band potential of
 * the members of each active region from its inner grid */
static void tl_phi_in(const tl_t *tl, int n, const char *keep, const float *w, float spacing,
                      double a, float *phi_in)
{
    for (int i = 0; i < n; ++i) phi_in[i] = 0.0f;
    for (int k = 0; k < tl->n_active; ++k) {
        int b = tl->active[k], r = tl->region_r[b];
        const green_t *g = tl_green(r);
        const int ng = g->n, nh = ng / 2 + 1;
        const size_t nn = (size_t)ng * ng * ng;
        float *rho = (float *)calloc(nn, sizeof(float));
        cpx *grid = (cpx *)calloc(nn, sizeof(cpx));
        for (int p = tl->src_start[k]; p < tl->src_start[k + 1]; ++p) {
            int j = tl->src_member[p];
            if (!keep[j]) continue;
            const int *f = tl->src_first + 3 * p;
            const float *w9 = tl->src_wt + 9 * p;
            for (int sz = 0; sz < 3; ++sz)
                for (int sy = 0; sy < 3; ++sy)
                    for (int sx = 0; sx < 3; ++sx) {
                        size_t gi = (size_t)(f[0] + sx) + (size_t)ng * ((f[1] + sy) + (size_t)ng * (f[2] + sz));
                        rho[gi] = rho[gi] + w[j] * w9[sx] * w9[3 + sy] * w9[6 + sz];
                    }
        }
        for (size_t gi = 0; gi < nn; ++gi) grid[gi].re = rho[gi];
        fft3(grid, ng, -1);
        for (int z = 0; z < ng; ++z)
            for (int y = 0; y < ng; ++y)
                for (int x = 0; x < ng; ++x) {
                    int hx = x, hy = y, hz = z;
                    if (x >= nh) { hx = ng - x; hy = (ng - y) % ng; hz = (ng - z) % ng; }
                    double gv = g->G[(size_t)hx + (size_t)nh * (hy + (size_t)ng * hz)];
                    size_t gi = (size_t)x + (size_t)ng * (y + (size_t)ng * z);
                    grid[gi].re *= gv;
                    grid[gi].im *= gv;
                }
        fft3(grid, ng, +1);
        const float norm = (float)((double)ng * ng * ng);
        for (size_t gi = 0; gi < nn; ++gi) rho[gi] = (float)grid[gi].re / norm;
        const float pref = (float)(a * (double)r / (6.0 * BPI * (double)spacing));
        for (int i = 0; i < n; ++i) {
            if (tl->target_reg[i] != b || !keep[i]) continue;
            const int *f = tl->tgt_first + 3 * i;
            const float *w9 = tl->tgt_wt + 9 * i;
            float raw = 0.0f;
            for (int sz = 0; sz < 3; ++sz)
                for (int sy = 0; sy < 3; ++sy)
                    for (int sx = 0; sx < 3; ++sx)
                        raw = raw + rho[(size_t)(f[0] + sx) + (size_t)ng * ((f[1] + sy) + (size_t)ng * (f[2] + sz))] *
                                        w9[sx] * w9[3 + sy] * w9[6 + sz];
            phi_in[i] = pref * (raw - w[i] * tl->self_in[i]);
        }
        free(rho);
        free(grid);
    }
}

/* short-range pair term of a two-level halo for target i (cutoff
 * a_i = 3.5 sp / r_i), bound_fortran.f90:1160-1175: w_j * dphi, f32 */
static float tl_pair(const float *qi, const float *qj, float wj, double ai, double eps, double factor)
{
    double dx = (double)qj[0] - (double)qi[0], dy = (double)qj[1] - (double)qi[1],
           dz = (double)qj[2] - (double)qi[2];
    double r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < ai * ai)) return 0.0f;
    double d = sqrt(r2);
    float dphi = -(float)(factor * (oracle_depth(d, eps) - oracle_depth(d, ai)));
    return wj * dphi;
}

/* bound_two_level_pp_full / _update (bound_fortran.f90:1088-1498) as a
 * gather over the 27 cells of width 3.5 sp around each target (every pair
 * within any cutoff lies there). The reference's f32 sums run in its own
 * task order; only that rounding differs. */
static void tl_pp_full(const c2cells_t *cc, int n, const float *q, const char *keep, const float *w,
                       float spacing, float app_val, double a, const int *pr, float *phi_c2)
{
    int nk = 0;
    for (int i = 0; i < n; ++i) { phi_c2[i] = 0.0f; nk += keep[i] ? 1 : 0; }
    if (nk <= 1) return;
    const double eps = (double)app_val, factor = a / (6.0 * BPI);
    const int *nc = cc->ncell;
    for (int i = 0; i < n; ++i) {
        if (!keep[i]) continue;
        const double ai = 3.5 * (double)spacing / (double)pr[i];
        int ci = cc->cell_id[i];
        int cz = (ci - 1) / (nc[0] * nc[1]) + 1, cy = ((ci - 1) % (nc[0] * nc[1])) / nc[0] + 1,
            cx = ((ci - 1) % (nc[0] * nc[1])) % nc[0] + 1;
        float acc = 0.0f;
        for (int nz = cz - 1; nz <= cz + 1; ++nz) {
            if (nz < 1 || nz > nc[2]) continue;
            for (int ny = cy - 1; ny <= cy + 1; ++ny) {
                if (ny < 1 || ny > nc[1]) continue;
                for (int nx = cx - 1; nx <= cx + 1; ++nx) {
                    if (nx < 1 || nx > nc[0]) continue;
                    int c2 = nx + (ny - 1) * nc[0] + (nz - 1) * nc[0] * nc[1];
                    for (int p = cc->cell_start[c2 - 1]; p < cc->cell_start[c2]; ++p) {
                        int j = cc->cell_part[p];
                        if (j == i || !keep[j]) continue;
                        acc = acc + tl_pair(q + 3 * i, q + 3 * j, w[j], ai, eps, factor);
                    }
                }
            }
        }
        phi_c2[i] = acc;
    }
}

static void tl_pp_update(const c2cells_t *cc, int n, const float *q, const char *keep,
                         const float *w, const int *just, int njust, float spacing, float app_val,
                         double a, const int *pr, float *phi_c2)
{
    int nk = 0;
    for (int i = 0; i < n; ++i) nk += keep[i] ? 1 : 0;
    if (njust <= 0 || nk <= 0) return;
    const double eps = (double)app_val, factor = a / (6.0 * BPI);
    const int *nc = cc->ncell;
    for (int k = 0; k < njust; ++k) {
        int j = just[k];
        int cj = cc->cell_id[j];
        int cz = (cj - 1) / (nc[0] * nc[1]) + 1, cy = ((cj - 1) % (nc[0] * nc[1])) / nc[0] + 1,
            cx = ((cj - 1) % (nc[0] * nc[1])) % nc[0] + 1;
        for (int nz = cz - 1; nz <= cz + 1; ++nz) {
            if (nz < 1 || nz > nc[2]) continue;
            for (int ny = cy - 1; ny <= cy + 1; ++ny) {
                if (ny < 1 || ny > nc[1]) continue;
                for (int nx = cx - 1; nx <= cx + 1; ++nx) {
                    if (nx < 1 || nx > nc[0]) continue;
                    int c2 = nx + (ny - 1) * nc[0] + (nz - 1) * nc[0] * nc[1];
                    for (int p = cc->cell_start[c2 - 1]; p < cc->cell_start[c2]; ++p) {
                        int i = cc->cell_part[p];
                        if (!keep[i]) continue;
                        const double ai = 3.5 * (double)spacing / (double)pr[i];
                        phi_c2[i] = phi_c2[i] - tl_pair(q + 3 * i, q + 3 * j, w[j], ai, eps, factor);
                    }
                }
            }
        }
    }
}

/* Direct summation (bound_fortran.f90:2613-2647). Pass 1 visits every pair
 * i < j once and adds the negative depth to both members, f64 sums of f32
 * differences; keep == NULL or all kept is the reference's pass 1. A member
 * not kept contributes nothing and receives nothing. */
static void direct_full(int n, const float *q, const float *w, const char *keep, double app,
                        double G_dir, double *phid)
{
    for (int i = 0; i < n; ++i) phid[i] = 0.0;
    for (int i = 0; i < n - 1; ++i) {
        if (keep && !keep[i]) continue;
        for (int j = i + 1; j < n; ++j) {
            if (keep && !keep[j]) continue;
            double dx = (double)(q[3 * j] - q[3 * i]);
            double dy = (double)(q[3 * j + 1] - q[3 * i + 1]);
            double dz = (double)(q[3 * j + 2] - q[3 * i + 2]);
            double r = sqrt(dx * dx + dy * dy + dz * dz);
            double dd = G_dir * oracle_depth(r, app);
            phid[i] = phid[i] - (double)w[j] * dd;
            phid[j] = phid[j] - (double)w[i] * dd;
        }
    }
}

/* Later passes (bound_fortran.f90:2630-2641): the sources deleted since the
 * last refresh are added back to every kept member. */
static void direct_addback(int n, const float *q, const float *w, const char *keep, const int *just,
                           int njust, double app, double G_dir, double *phid)
{
    for (int k = 0; k < njust; ++k) {
        int j = just[k];
        for (int i = 0; i < n; ++i) {
            if (!keep[i]) continue;
            double dx = (double)(q[3 * j] - q[3 * i]);
            double dy = (double)(q[3 * j + 1] - q[3 * i + 1]);
            double dz = (double)(q[3 * j + 2] - q[3 * i + 2]);
            double r = sqrt(dx * dx + dy * dy + dz * dz);
            phid[i] = phid[i] + (double)w[j] * G_dir * oracle_depth(r, app);
        }
    }
}

/* Checker switch (not a reference feature): 1 recomputes every later pass's
 * potential from scratch over the kept members (direct_full, c2_full,
 * tl_pp_full) instead of the reference's incremental updates, so that a test
 * can compare the two paths end to end. */
static int g_force_full = 0;
void oracle_set_force_full(int32_t on) { g_force_full = on ? 1 : 0; }

/* energy ordering of bound_check_grid's heap sort: "greater" = larger energy,
 * or equal energy and smaller local index */
static const float *g_energy;
static int cmp_energy(const void *pa, const void *pb)
{
    int a = *(const int *)pa, b = *(const int *)pb;
    float ea = g_energy[a], eb = g_energy[b];
    if (ea < eb) return -1;
    if (ea > eb) return 1;
    return a > b ? -1 : (a < b ? 1 : 0);  /* smaller index is "greater" */
}

/*
 * One candidate. q/v/w are in pid order (the reference heap-sorts members by
 * pid). keep (out) marks members kept. Returns status (0 ok, 1 too few,
 * 5 max iter) or a negative ierr (-4 non-finite phi, -1 non-finite energy,
 * -3 allocation). phi_first (optional) receives the first pass's potential.
 */
static int bound_check_halo(const bctx_t *c, const htask_t *t, const green_t *gr, int n,
                            const float *q, const float *v, const float *w, char *keep,
                            int *np_bound_out, float *phi_first)
{
    const int max_passes = 20;
    const double batch_fraction = 0.02, refresh_fraction = 0.20;
    const int nmin = c->np_min;
    int np_bound = n, n_just = 0, status = 5;
    float *phi = (float *)calloc((size_t)n, sizeof(float));
    float *energy = (float *)calloc((size_t)n, sizeof(float));
    double *phid = (double *)calloc((size_t)n, sizeof(double));
    int *order = (int *)malloc(sizeof(int) * (size_t)n);
    int *just = (int *)malloc(sizeof(int) * (size_t)n);
    float *self = NULL, *wt = NULL, *phi_c1 = NULL, *phi_c2 = NULL;
    int *first = NULL;
    c2cells_t cc = {{0, 0, 0}, 0, {0, 0, 0}, NULL, NULL, NULL};
    cpx *grid = NULL;
    tl_t tl;
    memset(&tl, 0, sizeof(tl));
    float *phi_in = NULL;
    const int gid = t->gid;
    const int ng = gid >= 1 ? c->bn[gid - 1] : 0;
    float spacing = (float)t->spacing;
    for (int i = 0; i < n; ++i) keep[i] = 1;
    if (gid >= 1) {
        self = (float *)malloc(sizeof(float) * n);
        wt = (float *)malloc(sizeof(float) * 9 * (size_t)n);
        first = (int *)malloc(sizeof(int) * 3 * (size_t)n);
        phi_c1 = (float *)calloc((size_t)n, sizeof(float));
        phi_c2 = (float *)calloc((size_t)n, sizeof(float));
        grid = (cpx *)malloc(sizeof(cpx) * (size_t)ng * ng * ng);
        for (int i = 0; i < n; ++i) {
            for (int ax = 0; ax < 3; ++ax) {
                float tt = (q[3 * i + ax] - t->origin[ax]) / spacing;
                if (!(tt >= 0.0f) || !(tt < (float)(ng / 2 - 1))) { status = -1; goto done; }
                int nearest = nint_f(tt);
                float delta = tt - (float)nearest;
                first[3 * i + ax] = nearest;
                float a0 = 0.5f - delta, a2 = 0.5f + delta;
                wt[9 * i + 3 * ax + 0] = 0.5f * (a0 * a0);
                wt[9 * i + 3 * ax + 1] = 0.75f - delta * delta;
                wt[9 * i + 3 * ax + 2] = 0.5f * (a2 * a2);
            }
            double acc = 0.0;
            for (int sz = 0; sz < 3; ++sz)
                for (int sy = 0; sy < 3; ++sy)
                    for (int sx = 0; sx < 3; ++sx) {
                        float sw = wt[9 * i + sx] * wt[9 * i + 3 + sy] * wt[9 * i + 6 + sz];
                        for (int tz = 0; tz < 3; ++tz)
                            for (int ty = 0; ty < 3; ++ty)
                                for (int tx = 0; tx < 3; ++tx) {
                                    float tw = wt[9 * i + tx] * wt[9 * i + 3 + ty] * wt[9 * i + 6 + tz];
                                    acc = acc + (double)(sw * tw) *
                                                    (double)gr->self[(tx - sx + 2) + 5 * ((ty - sy + 2) + 5 * (tz - sz + 2))];
                                }
                    }
            self[i] = (float)acc;
        }
        build_cells(&cc, n, q, spacing);
        /* two-level refinement of 256^3 halos (bound_fortran.f90:2603-2606) */
        if (gid == 4) {
            if (tl_setup(&tl, n, q, spacing, t->origin, (float)c->app) != 0) { status = -1; goto done; }
            if (tl.n_active > 0) phi_in = (float *)calloc((size_t)n, sizeof(float));
        }
    }
    const double G_dir = c->a / (6.0 * BPI);
    for (int pass = 1; pass <= max_passes; ++pass) {
        int n_source = 0;
        for (int i = 0; i < n; ++i) n_source += keep[i] ? 1 : 0;
        if (gid == 0) {
            if (pass == 1 || g_force_full) direct_full(n, q, w, keep, c->app, G_dir, phid);
            else direct_addback(n, q, w, keep, just, n_just, c->app, G_dir, phid);
            n_just = 0;
            for (int i = 0; i < n; ++i) if (keep[i]) phi[i] = (float)phid[i];
        } else {
            const int full_c2 = pass == 1 || g_force_full;
            const size_t nn = (size_t)ng * ng * ng;
            for (size_t g = 0; g < nn; ++g) { grid[g].re = 0.0; grid[g].im = 0.0; }
            /* deposit in float into a float grid (kept as double storage of float values) */
            float *rho = (float *)calloc(nn, sizeof(float));
            /* stencil nodes nearest..nearest+2 (bound_fortran.f90:2660-2668) */
            for (int i = 0; i < n; ++i) {
                if (!keep[i]) continue;
                for (int sz = 0; sz < 3; ++sz) {
                    int iz = first[3 * i + 2] + sz;
                    for (int sy = 0; sy < 3; ++sy) {
                        int iy = first[3 * i + 1] + sy;
                        for (int sx = 0; sx < 3; ++sx) {
                            int ix = first[3 * i] + sx;
                            size_t gi = (size_t)ix + (size_t)ng * (iy + (size_t)ng * iz);
                            rho[gi] = rho[gi] + w[i] * wt[9 * i + sx] * wt[9 * i + 3 + sy] * wt[9 * i + 6 + sz];
                        }
                    }
                }
            }
            for (size_t g = 0; g < nn; ++g) grid[g].re = rho[g];
            fft3(grid, ng, -1);
            const int nh = ng / 2 + 1;
            for (int z = 0; z < ng; ++z)
                for (int y = 0; y < ng; ++y)
                    for (int x = 0; x < ng; ++x) {
                        int hx = x, hy = y, hz = z;
                        if (x >= nh) { hx = ng - x; hy = (ng - y) % ng; hz = (ng - z) % ng; }
                        double gv = gr->G[(size_t)hx + (size_t)nh * (hy + (size_t)ng * hz)];
                        size_t gi = (size_t)x + (size_t)ng * (y + (size_t)ng * z);
                        grid[gi].re *= gv;
                        grid[gi].im *= gv;
                    }
            fft3(grid, ng, +1);
            const float norm = (float)((double)ng * ng * ng);
            for (size_t g = 0; g < nn; ++g) rho[g] = (float)grid[g].re / norm;
            const float radius_r = (float)ng / 2.0f;
            double mkeep = 0.0;
            for (int i = 0; i < n; ++i) if (keep[i]) mkeep += (double)w[i];
            const float c1pref = (float)(c->a / (6.0 * BPI * (double)spacing));
            for (int i = 0; i < n; ++i) {
                if (!keep[i]) continue;
                float dz0 = (float)fmax(0.0, mkeep - (double)w[i]) / radius_r;
                float raw = 0.0f;
                for (int sz = 0; sz < 3; ++sz) {
                    int iz = first[3 * i + 2] + sz;
                    for (int sy = 0; sy < 3; ++sy) {
                        int iy = first[3 * i + 1] + sy;
                        for (int sx = 0; sx < 3; ++sx) {
                            int ix = first[3 * i] + sx;
                            raw = raw + rho[(size_t)ix + (size_t)ng * (iy + (size_t)ng * iz)] *
                                            wt[9 * i + sx] * wt[9 * i + 3 + sy] * wt[9 * i + 6 + sz];
                        }
                    }
                }
                phi_c1[i] = c1pref * ((raw - w[i] * self[i]) - dz0);
                if (full_c2) phi_c2[i] = 0.0f;
            }
            free(rho);
            if (tl.n_active > 0) {
                tl_phi_in(&tl, n, keep, w, spacing, c->a, phi_in);
                if (full_c2) tl_pp_full(&cc, n, q, keep, w, spacing, (float)c->app, c->a, tl.particle_r, phi_c2);
                else tl_pp_update(&cc, n, q, keep, w, just, n_just, spacing, (float)c->app, c->a,
                                  tl.particle_r, phi_c2);
            } else if (full_c2) {
                c2_full(&cc, n, q, keep, w, spacing, (float)c->app, c->a, phi_c2);
            } else {
                c2_incremental(&cc, n, q, keep, w, just, n_just, spacing, (float)c->app, c->a, phi_c2);
            }
            n_just = 0;
        }
        for (int i = 0; i < n; ++i) {
            if (!keep[i]) continue;
            if (gid >= 1) phi[i] = tl.n_active > 0 ? phi_c1[i] + phi_in[i] + phi_c2[i] : phi_c1[i] + phi_c2[i];
            if (np_bound == 1 || n_source == 1) phi[i] = 0.0f;
            if (!isfinite(phi[i])) { status = -4; goto done; }
        }
        if (pass == 1 && phi_first) for (int i = 0; i < n; ++i) phi_first[i] = phi[i];
        int deleted = 0;
        for (;;) {
            double wsum = 0.0, qc[3] = {0, 0, 0}, vc[3] = {0, 0, 0};
            for (int i = 0; i < n; ++i) {
                if (!keep[i]) continue;
                double wi = (double)w[i];
                wsum = wsum + wi;
                for (int d = 0; d < 3; ++d) {
                    qc[d] = qc[d] + wi * (double)q[3 * i + d];
                    vc[d] = vc[d] + wi * (double)v[3 * i + d];
                }
            }
            for (int d = 0; d < 3; ++d) { qc[d] = qc[d] / wsum; vc[d] = vc[d] / wsum; }
            const float vc32[3] = {(float)vc[0], (float)vc[1], (float)vc[2]};
            const float qc32[3] = {(float)qc[0], (float)qc[1], (float)qc[2]};
            const float b32 = (float)c->b;
            int nu = 0;
            for (int i = 0; i < n; ++i) {
                if (!keep[i]) continue;
                float dv[3];
                for (int d = 0; d < 3; ++d) dv[d] = (v[3 * i + d] - vc32[d]) + b32 * (q[3 * i + d] - qc32[d]);
                float s = dv[0] * dv[0];
                s = s + dv[1] * dv[1];
                s = s + dv[2] * dv[2];
                energy[i] = s / 2.0f + phi[i];
                if (!isfinite(energy[i])) { status = -1; goto done; }
                if (energy[i] >= 0.0f) order[nu++] = i;
            }
            if (nu == 0) {
                if (deleted == 0) { status = 0; goto done; }
                break;
            }
            g_energy = energy;
            qsort(order, (size_t)nu, sizeof(int), cmp_energy);
            int nb = (int)ceil(batch_fraction * (double)np_bound);
            if (nb < 1) nb = 1;
            if (nb > nu) nb = nu;
            for (int k = nu - 1; k >= nu - nb; --k) {
                keep[order[k]] = 0;
                just[n_just++] = order[k];
            }
            np_bound -= nb;
            deleted += nb;
            if (np_bound < nmin) { status = 1; goto done; }
            if ((double)deleted >= refresh_fraction * (double)np_bound) break;
        }
    }
    status = 5;
done:
    *np_bound_out = np_bound;
    free(phi); free(energy); free(phid); free(order); free(just);
    free(self); free(wt); free(first); free(phi_c1); free(phi_c2); free(grid);
    tl_free(&tl);
    free(phi_in);
    if (gid >= 1) free_cells(&cc);
    return status;
}

/*
 * FoF + unbinding as fof_fortran with bound_check = 1 (fof_fortran.f90:253-356).
 * Returns ierr. Per candidate (global FoF order, up to cap entries): status,
 * np_bound, method gid, spacing. phi_first (optional, total packed members):
 * first-pass potential of every candidate member, candidates in order, members
 * in pid order.
 */
int32_t oracle_fof_bound(const double *xp, const double *vp, const double *mp,
                         const int32_t *np_halo, int32_t nhalo, double L, double b, int32_t np_min,
                         double app, double a_val, double b_val, double s15, int64_t ng,
                         int32_t *in_halo, int32_t *halo_id, double *ang_mom, int32_t cap,
                         int32_t *status_out, int32_t *nb_out, int32_t *gid_out, double *sp_out,
                         float *phi_first, int64_t phi_cap)
{
    int32_t total = 0;
    for (int32_t h = 0; h < nhalo; ++h) total += np_halo[h];
    for (int32_t i = 0; i < total; ++i) { in_halo[i] = 0; halo_id[i] = 0; }
    for (int32_t i = 0; i < 3 * nhalo; ++i) ang_mom[i] = 0.0;
    if (nhalo <= 0 || total <= 0) return 0;
    double Lh = 0.5 * L;
    /* candidates per region, CRS order of members */
    region_cands_t *rc = (region_cands_t *)calloc((size_t)nhalo, sizeof(region_cands_t));
    int32_t start = 0;
    for (int32_t h = 0; h < nhalo; ++h) {
        if (np_halo[h] >= np_min && np_halo[h] > 0) {
            if (oracle_process_region(xp, start, np_halo[h], L, Lh, b, np_min, 1, &rc[h], NULL, NULL)) {
                free(rc);
                return 3;
            }
        }
        start += np_halo[h];
    }
    bctx_t ctx;
    bctx_init(&ctx, app, a_val, b_val, s15, np_min, ng);
    int32_t ierr = 0;
    int32_t nc = 0;
    for (int32_t h = 0; h < nhalo; ++h) nc += rc[h].ncand;
    /* without candidates the batch is never entered (fof_fortran.f90:180-184) */
    if (nc > 0) {
        if (ctx.ngrid != 4) ierr = 1;  /* bound_fortran.f90:1559-1563 */
        if (!(a_val > 0.0) || !isfinite(a_val) || !isfinite(b_val)) ierr = 1;
    }
    /* pack + classify all candidates first (the batch classifies before solving) */
    htask_t *task = (htask_t *)calloc((size_t)(nc > 0 ? nc : 1), sizeof(htask_t));
    int32_t **mem_pid = (int32_t **)calloc((size_t)(nc > 0 ? nc : 1), sizeof(int32_t *));
    int32_t *cand_reg = (int32_t *)calloc((size_t)(nc > 0 ? nc : 1), sizeof(int32_t));
    int32_t *cand_n = (int32_t *)calloc((size_t)(nc > 0 ? nc : 1), sizeof(int32_t));
    int32_t *cand_anchor = (int32_t *)calloc((size_t)(nc > 0 ? nc : 1), sizeof(int32_t));
    int32_t hc = 0;
    for (int32_t h = 0; h < nhalo && ierr == 0; ++h)
        for (int32_t g = 0; g < rc[h].ncand && ierr == 0; ++g, ++hc) {
            const int32_t *mem = rc[h].members + rc[h].member_ptr[g];
            int32_t n = rc[h].cand_np[g];
            cand_reg[hc] = h;
            cand_n[hc] = n;
            cand_anchor[hc] = mem[0];
            /* span in float from the anchor-relative float positions */
            float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
            for (int32_t k = 0; k < n; ++k)
                for (int d = 0; d < 3; ++d) {
                    float qv = (float)(oracle_modulo(xp[3 * (size_t)mem[k] + d] - xp[3 * (size_t)mem[0] + d] + Lh, L) - Lh);
                    if (qv < lo[d]) lo[d] = qv;
                    if (qv > hi[d]) hi[d] = qv;
                }
            float span = hi[0] - lo[0];
            if (hi[1] - lo[1] > span) span = hi[1] - lo[1];
            if (hi[2] - lo[2] > span) span = hi[2] - lo[2];
            task[hc].span = span;
            if (n > 1000 && span > (float)s15) { ierr = 2; break; }
            int gid, lev;
            double sp;
            if (oracle_select_grid(&ctx, n, (double)span, &gid, &lev, &sp) != 0) { ierr = 1; break; }
            task[hc].gid = gid;
            task[hc].lev = lev;
            if (gid >= 1) {
                task[hc].spacing = (double)(float)sp;
                for (int d = 0; d < 3; ++d) task[hc].origin[d] = lo[d] - (float)sp;
                if ((double)(ctx.bn[gid - 1] / 2 - 3) * task[hc].spacing < (double)span) { ierr = 1; break; }
            }
            /* members sorted by pid (global index) */
            int32_t *mp2 = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
            memcpy(mp2, mem, sizeof(int32_t) * (size_t)n);
            for (int32_t i = 1; i < n; ++i) {  /* insertion sort: indices are distinct */
                int32_t key = mp2[i], j = i - 1;
                while (j >= 0 && mp2[j] > key) { mp2[j + 1] = mp2[j]; --j; }
                mp2[j + 1] = key;
            }
            mem_pid[hc] = mp2;
        }
    /* Green functions depend on n only: computed once per process and size */
    static green_t greens_cache[4];
    green_t *greens = greens_cache;
    int32_t valid = 0;
    int64_t phi_off = 0;
    if (ierr == 0) {
        for (int32_t h = 0; h < nc; ++h) {
            int gid = task[h].gid;
            if (gid >= 1 && (!greens[gid - 1].G || greens[gid - 1].n != ctx.bn[gid - 1])) {
                int n = ctx.bn[gid - 1];
                free(greens[gid - 1].G);
                greens[gid - 1].n = n;
                greens[gid - 1].G = (float *)malloc(sizeof(float) * (size_t)(n / 2 + 1) * n * n);
                oracle_green(n, greens[gid - 1].G, greens[gid - 1].self);
            }
        }
        for (int32_t h = 0; h < nc && ierr == 0; ++h) {
            int32_t n = cand_n[h];
            float *q = (float *)malloc(sizeof(float) * 3 * (size_t)n);
            float *v = (float *)malloc(sizeof(float) * 3 * (size_t)n);
            float *w = (float *)malloc(sizeof(float) * (size_t)n);
            char *keep = (char *)malloc((size_t)n);
            const int32_t *mem = mem_pid[h];
            const size_t an = (size_t)cand_anchor[h];
            for (int32_t i = 0; i < n; ++i) {
                size_t o = (size_t)mem[i];
                for (int d = 0; d < 3; ++d) {
                    /* xv = real(dx_rel, 4) relative to the anchor; q = xv - xv(anchor) = xv */
                    q[3 * i + d] = (float)(oracle_modulo(xp[3 * o + d] - xp[3 * an + d] + Lh, L) - Lh);
                    v[3 * i + d] = vp ? (float)vp[3 * o + d] : 0.0f;
                }
                w[i] = mp ? (float)mp[o] : 1.0f;
            }
            int nb = 0;
            float *pf = (phi_first && phi_off + n <= phi_cap) ? phi_first + phi_off : NULL;
            int st = bound_check_halo(&ctx, &task[h], task[h].gid >= 1 ? &greens[task[h].gid - 1] : NULL,
                                      n, q, v, w, keep, &nb, pf);
            phi_off += n;
            if (st < 0) {
                ierr = st == -4 ? 4 : 1;
            } else {
                if (h < cap) {
                    if (status_out) status_out[h] = st;
                    if (nb_out) nb_out[h] = nb;
                    if (gid_out) gid_out[h] = task[h].gid;
                    if (sp_out) sp_out[h] = task[h].spacing;
                }
                if (st == 0 && nb >= np_min) {
                    ++valid;
                    /* centre and angular momentum over kept members (f64, input values) */
                    double wsum = 0.0, qc[3] = {0, 0, 0}, vc[3] = {0, 0, 0};
                    for (int32_t i = 0; i < n; ++i) {
                        if (!keep[i]) continue;
                        size_t o = (size_t)mem[i];
                        double wi = mp ? mp[o] : 1.0;
                        wsum += wi;
                        for (int d = 0; d < 3; ++d) {
                            qc[d] += wi * (oracle_modulo(xp[3 * o + d] - xp[3 * an + d] + Lh, L) - Lh);
                            vc[d] += wi * (vp ? vp[3 * o + d] : 0.0);
                        }
                    }
                    int reg = cand_reg[h];
                    if (wsum > 0.0) {
                        for (int d = 0; d < 3; ++d) { qc[d] /= wsum; vc[d] /= wsum; }
                    }
                    for (int32_t i = 0; i < n; ++i) {
                        if (!keep[i]) continue;
                        size_t o = (size_t)mem[i];
                        in_halo[o] = 1;
                        halo_id[o] = valid;
                        if (!(wsum > 0.0)) continue;
                        double wi = mp ? mp[o] : 1.0, dx[3], dv[3];
                        for (int d = 0; d < 3; ++d) {
                            dx[d] = (oracle_modulo(xp[3 * o + d] - xp[3 * an + d] + Lh, L) - Lh) - qc[d];
                            dv[d] = (vp ? vp[3 * o + d] : 0.0) - vc[d];
                        }
                        ang_mom[3 * reg + 0] += wi * (dx[1] * dv[2] - dx[2] * dv[1]);
                        ang_mom[3 * reg + 1] += wi * (dx[2] * dv[0] - dx[0] * dv[2]);
                        ang_mom[3 * reg + 2] += wi * (dx[0] * dv[1] - dx[1] * dv[0]);
                    }
                }
            }
            free(q); free(v); free(w); free(keep);
        }
    }
    if (ierr != 0) {
        for (int32_t i = 0; i < total; ++i) { in_halo[i] = 0; halo_id[i] = 0; }
        for (int32_t i = 0; i < 3 * nhalo; ++i) ang_mom[i] = 0.0;
    }
    for (int32_t h = 0; h < nc; ++h) free(mem_pid[h]);
    for (int32_t h = 0; h < nhalo; ++h) { free(rc[h].cand_np); free(rc[h].member_ptr); free(rc[h].members); }
    free(rc); free(task); free(mem_pid); free(cand_reg); free(cand_n); free(cand_anchor);
    return ierr;
}

/* ------------------------------------------------------------------ */
/* Checker hooks: one removal batch, the reference's incremental update   */
/* against a full recomputation over the members still kept.              */
/* q: (n, 3) f32 positions, w: f32 masses, keep0: members kept before the */
/* batch, rem: the nrem members the batch deletes (kept in keep0).        */
/* phi_inc: full sum over keep0 then the incremental update for rem;      */
/* phi_full: full sum over keep0 minus rem. Returns 0, or 1 on bad input. */

static char *apply_removal(int n, const char *keep0, int nrem, const int32_t *rem)
{
    char *k1 = (char *)malloc((size_t)(n > 0 ? n : 1));
    memcpy(k1, keep0, (size_t)n);
    for (int k = 0; k < nrem; ++k) {
        if (rem[k] < 0 || rem[k] >= n || !k1[rem[k]]) { free(k1); return NULL; }
        k1[rem[k]] = 0;
    }
    return k1;
}

int32_t oracle_c2_check(int32_t n, const float *q, const float *w, const char *keep0, int32_t nrem,
                        const int32_t *rem, float sp, float app, double a, float *phi_inc,
                        float *phi_full)
{
    char *k1 = apply_removal(n, keep0, nrem, rem);
    if (!k1) return 1;
    c2cells_t cc;
    build_cells(&cc, n, q, sp);
    c2_full(&cc, n, q, keep0, w, sp, app, a, phi_inc);
    c2_incremental(&cc, n, q, k1, w, rem, nrem, sp, app, a, phi_inc);
    c2_full(&cc, n, q, k1, w, sp, app, a, phi_full);
    free_cells(&cc);
    free(k1);
    return 0;
}

/* pr: per-member refinement ratio (1, 2 or 4); cutoff a_i = 3.5 sp / pr_i */
int32_t oracle_tl_pp_check(int32_t n, const float *q, const float *w, const char *keep0, int32_t nrem,
                           const int32_t *rem, float sp, float app, double a, const int32_t *pr,
                           float *phi_inc, float *phi_full)
{
    char *k1 = apply_removal(n, keep0, nrem, rem);
    if (!k1) return 1;
    c2cells_t cc;
    build_cells(&cc, n, q, sp);
    tl_pp_full(&cc, n, q, keep0, w, sp, app, a, pr, phi_inc);
    tl_pp_update(&cc, n, q, k1, w, rem, nrem, sp, app, a, pr, phi_inc);
    tl_pp_full(&cc, n, q, k1, w, sp, app, a, pr, phi_full);
    free_cells(&cc);
    free(k1);
    return 0;
}

int32_t oracle_direct_check(int32_t n, const float *q, const float *w, const char *keep0, int32_t nrem,
                            const int32_t *rem, double app, double a, double *phi_inc,
                            double *phi_full)
{
    char *k1 = apply_removal(n, keep0, nrem, rem);
    if (!k1) return 1;
    const double G_dir = a / (6.0 * BPI);
    direct_full(n, q, w, keep0, app, G_dir, phi_inc);
    direct_addback(n, q, w, k1, rem, nrem, app, G_dir, phi_inc);
    direct_full(n, q, w, k1, app, G_dir, phi_full);
    free(k1);
    return 0;
}
