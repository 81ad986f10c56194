/*
 * TEST INFRASTRUCTURE ONLY -- the oracle.
 *
 * Plain-C restatement of the reference Friends-of-Friends path, used by the
 * tests, the smoke check and the CPU baseline leg of the benchmark as the
 * checker. Nothing in the product (py_fof_amd) links, loads or calls this.
 *
 * Follows, line by line in meaning (not in text):
 *   fof_fortran.f90:10-362   driver: regions, candidates, write-back, ang. mom.
 *   fof_fortran.f90:369-379  compute_fof_parallel_config (thread rule)
 *   fof_fortran.f90:424-478  serial DSU union / find-root with compression
 *   fof_fortran.f90:483-491  dist2_pbc (periodic predicate)
 *   fof_fortran.f90:496-947  process_single_region (wrap, bbox, geometry,
 *                            two-pass CRS, linking loops, candidates)
 *   atomic_dsu.c:10-39       lock-free union with path halving (used when the
 *                            reference's thread rule gives 2 linking threads)
 *
 * Arithmetic contract (bit-exact with an FMA-free x86-64 gfortran build):
 *   - compiled with -ffp-contract=off, no -ffast-math;
 *   - real MODULO(a, p) = fmod(a, p) (+p when non-zero and sign(a) != sign(p));
 *   - (vol/np)**(1/3) = pow(., 1.0/3.0) from libm, as gfortran emits;
 *   - x_crs = (float)x_rel; periodic r2 in double from float differences;
 *     non-periodic r2 = float sum ((dx*dx + dy*dy) + dz*dz) widened to double.
 *
 * Parity pins: tests/test_oracle_fof.py checks this file against the
 * reference's constructive known-answer tests (test_fof.py:361-458, 653-729,
 * 804-866), against an O(N^2) brute-force connected-components search with
 * the same predicate, and against scipy cKDTree + connected_components.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* gfortran MODULO for reals: fmod, then shift into the sign of p. */
double oracle_modulo(double a, double p)
{
    double r = fmod(a, p);
    if (r != 0.0 && ((r < 0.0) != (p < 0.0))) r += p;
    return r;
}

static double dmax(double a, double b) { return a > b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

/* fof_fortran.f90:483-491 */
static double dist2_pbc(const float *a, const float *b, double L, double Lh)
{
    double dx = fabs((double)(a[0] - b[0])); if (dx > Lh) dx = L - dx;
    double dy = fabs((double)(a[1] - b[1])); if (dy > Lh) dy = L - dy;
    double dz = fabs((double)(a[2] - b[2])); if (dz > Lh) dz = L - dz;
    return dx * dx + dy * dy + dz * dz;
}

/* fof_fortran.f90:636 -- sum((x1-x2)**2) over real(4), widened to real(8) */
static double dist2_flat(const float *a, const float *b)
{
    float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    float s = dx * dx;
    s = s + dy * dy;
    s = s + dz * dz;
    return (double)s;
}

/* fof_fortran.f90:424-458, 1-based parent with -1 for roots */
static void dsu_union_serial(int32_t *parent, int32_t i, int32_t j)
{
    int32_t ra = i, rb = j, cur, nxt;
    while (parent[ra - 1] > 0) ra = parent[ra - 1];
    cur = i;
    while (cur != ra) { nxt = parent[cur - 1]; parent[cur - 1] = ra; cur = nxt; }
    while (parent[rb - 1] > 0) rb = parent[rb - 1];
    cur = j;
    while (cur != rb) { nxt = parent[cur - 1]; parent[cur - 1] = rb; cur = nxt; }
    if (ra == rb) return;
    if (ra > rb) parent[ra - 1] = rb; else parent[rb - 1] = ra;
}

/* atomic_dsu.c:10-23 */
static int32_t dsu_root_atomic(int32_t *parent, int32_t p)
{
    for (;;) {
        int32_t next = __atomic_load_n(&parent[p - 1], __ATOMIC_RELAXED);
        if (next < 0) return p;
        int32_t after = __atomic_load_n(&parent[next - 1], __ATOMIC_RELAXED);
        if (after > 0) {
            int32_t expected = next;
            (void)__atomic_compare_exchange_n(&parent[p - 1], &expected, after, 0,
                                              __ATOMIC_RELAXED, __ATOMIC_RELAXED);
        }
        p = next;
    }
}

/* atomic_dsu.c:25-39 */
static void dsu_union_atomic(int32_t *parent, int32_t i, int32_t j)
{
    for (;;) {
        int32_t a = dsu_root_atomic(parent, i), b = dsu_root_atomic(parent, j);
        if (a == b) return;
        int32_t lo = a < b ? a : b, hi = a < b ? b : a, expected = -1;
        if (__atomic_compare_exchange_n(&parent[hi - 1], &expected, lo, 0,
                                        __ATOMIC_RELAXED, __ATOMIC_RELAXED))
            return;
    }
}

/* fof_fortran.f90:463-478 */
static int32_t dsu_find_root(int32_t *parent, int32_t i)
{
    int32_t r = i, cur = i, nxt;
    while (parent[r - 1] > 0) r = parent[r - 1];
    while (cur != r) { nxt = parent[cur - 1]; parent[cur - 1] = r; cur = nxt; }
    return r;
}

/* Geometry of one region, fof_fortran.f90:518-570. Exposed for tests. */
typedef struct {
    double origin[3], cell[3], bmin[3], bmax[3];
    int32_t n[3];
    int32_t periodic[3];
    int32_t is_periodic;
} oracle_geom_t;

void oracle_region_geometry(const double *bmin, const double *bmax, int32_t np,
                            double L, double b, oracle_geom_t *g)
{
    double range[3];
    for (int d = 0; d < 3; ++d) {
        g->bmin[d] = bmin[d];
        g->bmax[d] = bmax[d];
        range[d] = bmax[d] - bmin[d];
        g->periodic[d] = (range[d] + 2.0 * b >= L);
    }
    g->is_periodic = g->periodic[0] || g->periodic[1] || g->periodic[2];
    double vol = dmax(range[0], b) * dmax(range[1], b) * dmax(range[2], b);
    double lmesh = dmax(b, pow(vol / (double)imax(1, np), 1.0 / 3.0));
    for (int pass = 0; pass < 2; ++pass) {
        for (int d = 0; d < 3; ++d) {
            if (g->periodic[d]) {
                g->origin[d] = -0.5 * L;
                g->n[d] = imax(1, imin((int32_t)(L / b), (int32_t)(L / lmesh)));
                g->cell[d] = L / (double)g->n[d];
            } else {
                g->origin[d] = bmin[d];
                g->n[d] = imax(1, (int32_t)ceil(range[d] / lmesh));
                g->cell[d] = lmesh;
            }
        }
        if (pass == 1) break;
        /* refinement test, int32 arithmetic as in the reference */
        int32_t four_np = (int32_t)(4u * (uint32_t)np);
        if ((double)g->n[0] * (double)g->n[1] * (double)g->n[2] > (double)imax(262144, four_np)) {
            int32_t two_np = (int32_t)(2u * (uint32_t)imax(1, np));
            lmesh = pow(vol / (double)two_np, 1.0 / 3.0);
            lmesh = dmax(b, lmesh);
        } else {
            break;
        }
    }
}

static int32_t cell_coord(double x, double o, double c, int32_t n)
{
    int32_t i = (int32_t)floor((x - o) / c) + 1;
    return imin(n, imax(1, i));
}

#include "oracle_internal.h"

/*
 * process_single_region, fof_fortran.f90:496-947 (minus the unused provisional
 * centre of mass, :922-944). istart is 0-based. label_out (optional) receives
 * for each particle of the region the 0-based global index of its component's
 * root (first member in CRS order).
 */
int oracle_process_region(const double *xp, int32_t istart, int32_t np, double L,
                          double Lh, double b, int32_t np_min, int32_t nest,
                          region_cands_t *out, int32_t *label_out, void *geom_vp)
{
    out->ncand = 0;
    out->cand_np = out->member_ptr = out->members = NULL;
    double rp2 = b * b;
    const double *xa = xp + 3 * (size_t)istart;
    double *xr = (double *)malloc(sizeof(double) * 3 * (size_t)np);
    if (!xr) return 3;
    double bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
    for (int32_t i = 0; i < np; ++i) {
        for (int d = 0; d < 3; ++d) {
            double v = oracle_modulo(xp[3 * (size_t)(istart + i) + d] - xa[d] + Lh, L) - Lh;
            xr[3 * (size_t)i + d] = v;
            if (i == 0 || v < bmin[d]) bmin[d] = v;
            if (i == 0 || v > bmax[d]) bmax[d] = v;
        }
    }
    oracle_geom_t g;
    oracle_region_geometry(bmin, bmax, np, L, b, &g);
    if (geom_vp) *(oracle_geom_t *)geom_vp = g;
    const int32_t Nx = g.n[0], Ny = g.n[1], Nz = g.n[2];
    const int32_t ncell = Nx * Ny * Nz;

    int32_t *cell_ptr = (int32_t *)calloc((size_t)ncell + 1, sizeof(int32_t));
    int32_t *cursor = (int32_t *)calloc((size_t)ncell, sizeof(int32_t));
    int32_t *icell_of = (int32_t *)malloc(sizeof(int32_t) * (size_t)np);
    float *xc = (float *)malloc(sizeof(float) * 3 * (size_t)np);
    int32_t *pid = (int32_t *)malloc(sizeof(int32_t) * (size_t)np);
    int32_t *parent = (int32_t *)malloc(sizeof(int32_t) * (size_t)np);
    if (!cell_ptr || !cursor || !icell_of || !xc || !pid || !parent) {
        free(xr); free(cell_ptr); free(cursor); free(icell_of); free(xc); free(pid); free(parent);
        return 3;
    }
    /* pass 1: counts (:580-586) */
    for (int32_t i = 0; i < np; ++i) {
        int32_t ix = cell_coord(xr[3 * (size_t)i + 0], g.origin[0], g.cell[0], Nx);
        int32_t iy = cell_coord(xr[3 * (size_t)i + 1], g.origin[1], g.cell[1], Ny);
        int32_t iz = cell_coord(xr[3 * (size_t)i + 2], g.origin[2], g.cell[2], Nz);
        int32_t ic = ix + Nx * (iy - 1) + Nx * Ny * (iz - 1);
        icell_of[i] = ic;
        cursor[ic - 1] += 1;
    }
    /* scan (:588-593), 1-based positions */
    cell_ptr[0] = 1;
    for (int32_t c = 0; c < ncell; ++c) {
        cell_ptr[c + 1] = cell_ptr[c] + cursor[c];
        cursor[c] = cell_ptr[c];
    }
    /* pass 2: stable scatter (:602-613) */
    for (int32_t i = 0; i < np; ++i) {
        int32_t ic = icell_of[i];
        int32_t p = cursor[ic - 1];
        for (int d = 0; d < 3; ++d) xc[3 * (size_t)(p - 1) + d] = (float)xr[3 * (size_t)i + d];
        pid[p - 1] = istart + i;
        cursor[ic - 1] = p + 1;
    }
    free(xr);
    for (int32_t i = 0; i < np; ++i) parent[i] = -1;

    /* linking (:625-846). The partition is independent of the union order. */
    static const int st[13][3] = {
        {-1, -1, -1}, {0, -1, -1}, {1, -1, -1}, {-1, 0, -1}, {0, 0, -1},
        {1, 0, -1},   {-1, 1, -1}, {0, 1, -1},  {1, 1, -1},  {-1, -1, 0},
        {0, -1, 0},   {1, -1, 0},  {-1, 0, 0}};
    const int use_atomic = nest > 1;
    (void)use_atomic;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nest > 1 ? nest : 1) if (nest > 1)
#endif
    for (int32_t iz = 1; iz <= Nz; ++iz) {
        int32_t nbr[27];
        for (int32_t iy = 1; iy <= Ny; ++iy) {
            for (int32_t ix = 1; ix <= Nx; ++ix) {
                int32_t ic = ix + Nx * (iy - 1) + Nx * Ny * (iz - 1);
                int32_t a0 = cell_ptr[ic - 1], a1 = cell_ptr[ic];
                if (a0 >= a1) continue;
                int32_t nn = 0;
                if (!g.is_periodic) {
                    for (int k = 0; k < 13; ++k) {
                        int32_t jx = ix + st[k][0], jy = iy + st[k][1], jz = iz + st[k][2];
                        if (jx < 1 || jx > Nx || jy < 1 || jy > Ny || jz < 1 || jz > Nz) continue;
                        nbr[nn++] = jx + Nx * (jy - 1) + Nx * Ny * (jz - 1);
                    }
                } else {
                    for (int dz = -1; dz <= 1; ++dz) {
                        if (!g.periodic[2] && (iz + dz < 1 || iz + dz > Nz)) continue;
                        int32_t jz = (int32_t)((((iz + dz - 1) % Nz) + Nz) % Nz) + 1;
                        for (int dy = -1; dy <= 1; ++dy) {
                            if (!g.periodic[1] && (iy + dy < 1 || iy + dy > Ny)) continue;
                            int32_t jy = (int32_t)((((iy + dy - 1) % Ny) + Ny) % Ny) + 1;
                            for (int dx = -1; dx <= 1; ++dx) {
                                if (!g.periodic[0] && (ix + dx < 1 || ix + dx > Nx)) continue;
                                int32_t jx = (int32_t)((((ix + dx - 1) % Nx) + Nx) % Nx) + 1;
                                int32_t jc = jx + Nx * (jy - 1) + Nx * Ny * (jz - 1);
                                if (jc <= ic) continue;
                                int dup = 0;
                                for (int k = 0; k < nn; ++k) if (nbr[k] == jc) { dup = 1; break; }
                                if (!dup) nbr[nn++] = jc;
                            }
                        }
                    }
                }
                /* own cell, upper triangle */
                for (int32_t p1 = a0; p1 < a1; ++p1) {
                    for (int32_t p2 = p1 + 1; p2 < a1; ++p2) {
                        const float *x1 = xc + 3 * (size_t)(p1 - 1), *x2 = xc + 3 * (size_t)(p2 - 1);
                        double r2 = g.is_periodic ? dist2_pbc(x1, x2, L, Lh) : dist2_flat(x1, x2);
                        if (r2 <= rp2) {
                            if (nest > 1) dsu_union_atomic(parent, p1, p2);
                            else dsu_union_serial(parent, p1, p2);
                        }
                    }
                }
                for (int k = 0; k < nn; ++k) {
                    int32_t jc = nbr[k];
                    int32_t b0 = cell_ptr[jc - 1], b1 = cell_ptr[jc];
                    if (b0 >= b1) continue;
                    for (int32_t p1 = a0; p1 < a1; ++p1) {
                        for (int32_t p2 = b0; p2 < b1; ++p2) {
                            const float *x1 = xc + 3 * (size_t)(p1 - 1), *x2 = xc + 3 * (size_t)(p2 - 1);
                            double r2 = g.is_periodic ? dist2_pbc(x1, x2, L, Lh) : dist2_flat(x1, x2);
                            if (r2 <= rp2) {
                                if (nest > 1) dsu_union_atomic(parent, p1, p2);
                                else dsu_union_serial(parent, p1, p2);
                            }
                        }
                    }
                }
            }
        }
    }

    /* candidates (:848-920) */
    int32_t *root_size = (int32_t *)calloc((size_t)np, sizeof(int32_t));
    int32_t *root_map = (int32_t *)calloc((size_t)np, sizeof(int32_t));
    for (int32_t p = 1; p <= np; ++p) root_size[dsu_find_root(parent, p) - 1] += 1;
    if (label_out) {
        for (int32_t p = 1; p <= np; ++p)
            label_out[pid[p - 1] - istart] = pid[dsu_find_root(parent, p) - 1];
    }
    int32_t ncand = 0, nmem = 0;
    for (int32_t r = 1; r <= np; ++r) {
        if (root_size[r - 1] >= np_min) {
            root_map[r - 1] = ++ncand;
            nmem += root_size[r - 1];
        }
    }
    out->ncand = ncand;
    if (ncand > 0) {
        out->cand_np = (int32_t *)malloc(sizeof(int32_t) * (size_t)ncand);
        out->member_ptr = (int32_t *)malloc(sizeof(int32_t) * ((size_t)ncand + 1));
        out->members = (int32_t *)malloc(sizeof(int32_t) * (size_t)nmem);
        int32_t *cur = (int32_t *)malloc(sizeof(int32_t) * (size_t)ncand);
        int32_t c = 0;
        out->member_ptr[0] = 0;
        for (int32_t r = 1; r <= np; ++r) {
            if (root_size[r - 1] >= np_min) {
                out->cand_np[c] = root_size[r - 1];
                out->member_ptr[c + 1] = out->member_ptr[c] + root_size[r - 1];
                ++c;
            }
        }
        memcpy(cur, out->member_ptr, sizeof(int32_t) * (size_t)ncand);
        for (int32_t p = 1; p <= np; ++p) {
            int32_t gidx = root_map[dsu_find_root(parent, p) - 1];
            if (gidx > 0) out->members[cur[gidx - 1]++] = pid[p - 1];
        }
        free(cur);
    }
    free(root_size); free(root_map);
    free(cell_ptr); free(cursor); free(icell_of); free(xc); free(pid); free(parent);
    return 0;
}

/* fof_fortran.f90:369-379 */
void oracle_parallel_config(int32_t T, int32_t H, int32_t nmax, int32_t *team, int32_t *nest)
{
    *nest = (T >= 2 && H <= 10 * T && nmax > 10000) ? 2 : 1;
    int32_t t = T / *nest;
    *team = imin(H, imax(1, t));
}

/* angular momentum of one candidate about its f64 centre of mass, :207-250 */
static void add_ang_mom(const double *xp, const double *vp, const double *mp,
                        const int32_t *mem, int32_t n, int32_t anchor, double L,
                        double Lh, double *am)
{
    double w_sum = 0.0, qc[3] = {0, 0, 0}, vc[3] = {0, 0, 0};
    const double *xa = xp + 3 * (size_t)anchor;
    for (int32_t k = 0; k < n; ++k) {
        int32_t o = mem[k];
        double wi = mp[o], dx[3];
        w_sum += wi;
        for (int d = 0; d < 3; ++d) {
            dx[d] = oracle_modulo(xp[3 * (size_t)o + d] - xa[d] + Lh, L) - Lh;
            qc[d] += wi * dx[d];
            vc[d] += wi * vp[3 * (size_t)o + d];
        }
    }
    if (!(w_sum > 0.0)) return;
    for (int d = 0; d < 3; ++d) { qc[d] /= w_sum; vc[d] /= w_sum; }
    for (int32_t k = 0; k < n; ++k) {
        int32_t o = mem[k];
        double wi = mp[o], dx[3], dv[3];
        for (int d = 0; d < 3; ++d) {
            dx[d] = (oracle_modulo(xp[3 * (size_t)o + d] - xa[d] + Lh, L) - Lh) - qc[d];
            dv[d] = vp[3 * (size_t)o + d] - vc[d];
        }
        am[0] += wi * (dx[1] * dv[2] - dx[2] * dv[1]);
        am[1] += wi * (dx[2] * dv[0] - dx[0] * dv[2]);
        am[2] += wi * (dx[0] * dv[1] - dx[1] * dv[0]);
    }
}

/*
 * FoF without unbinding: fof_fortran.f90:10-250 with bound_check = 0.
 * xp, vp: (3, total_np) Fortran order (x0,y0,z0,x1,...), vp may be NULL (zeros).
 * mp may be NULL (unit masses). ang_mom: (3, nhalo) Fortran order.
 * label_out (optional, int32[total_np]): component root (0-based global index)
 * per particle, -1 for particles of skipped regions.
 * Returns ierr (0, 3 or 10) with the reference's meanings.
 */
/* Checker switch (not a reference feature): k > 0 links every region with k
 * threads instead of the reference's rule. The union always hooks the larger
 * root under the smaller, so the partition and its labels do not depend on
 * the thread count; only the time to build a large fixture does. */
static int32_t g_link_threads = 0;
void oracle_set_link_threads(int32_t k) { g_link_threads = k > 0 ? k : 0; }

int32_t oracle_fof(const double *xp, const double *vp, const double *mp,
                   const int32_t *np_halo, int32_t nhalo, double L, double b,
                   int32_t np_min, int32_t n_threads, int32_t total_np,
                   int32_t *in_halo, int32_t *halo_id, double *ang_mom,
                   int32_t *label_out)
{
    for (int32_t i = 0; i < total_np; ++i) { in_halo[i] = 0; halo_id[i] = 0; }
    if (label_out) for (int32_t i = 0; i < total_np; ++i) label_out[i] = -1;
    for (int32_t i = 0; i < 3 * nhalo; ++i) ang_mom[i] = 0.0;
    if (n_threads > 256) return 10;
    if (nhalo <= 0 || total_np <= 0) return 0;
    int32_t T;
#ifdef _OPENMP
    T = n_threads <= 0 ? imin(omp_get_max_threads(), 256) : imin(n_threads, 256);
#else
    T = n_threads <= 0 ? 1 : imin(n_threads, 256);
#endif
    int32_t nmax = 0;
    for (int32_t h = 0; h < nhalo; ++h) nmax = imax(nmax, np_halo[h]);
    int32_t team, nest;
    oracle_parallel_config(T, nhalo, nmax, &team, &nest);
    if (g_link_threads > 0) nest = g_link_threads;
    double Lh = 0.5 * L;
    region_cands_t *rc = (region_cands_t *)calloc((size_t)nhalo, sizeof(region_cands_t));
    int32_t *start = (int32_t *)malloc(sizeof(int32_t) * (size_t)nhalo);
    if (!rc || !start) { free(rc); free(start); return 3; }
    start[0] = 0;
    for (int32_t h = 1; h < nhalo; ++h) start[h] = start[h - 1] + np_halo[h - 1];
    int32_t ierr = 0;
    for (int32_t h = 0; h < nhalo; ++h) {
        int32_t np = np_halo[h];
        if (np >= np_min && np > 0) {
            int e = oracle_process_region(xp, start[h], np, L, Lh, b, np_min, nest, &rc[h],
                                   label_out ? label_out + start[h] : NULL, NULL);
            if (e) ierr = e;
        }
    }
    if (ierr == 0) {
        double *ones = NULL;
        const double *m = mp;
        if (vp && !m) {
            ones = (double *)malloc(sizeof(double) * (size_t)total_np);
            for (int32_t i = 0; i < total_np; ++i) ones[i] = 1.0;
            m = ones;
        }
        int32_t hcount = 0;
        for (int32_t h = 0; h < nhalo; ++h) {
            for (int32_t g = 0; g < rc[h].ncand; ++g) {
                ++hcount;
                const int32_t *mem = rc[h].members + rc[h].member_ptr[g];
                int32_t n = rc[h].cand_np[g];
                for (int32_t k = 0; k < n; ++k) { in_halo[mem[k]] = 1; halo_id[mem[k]] = hcount; }
                /* a NULL vp stands for zero velocities: the moment is exactly 0 */
                if (vp) add_ang_mom(xp, vp, m, mem, n, mem[0], L, Lh, ang_mom + 3 * (size_t)h);
            }
        }
        free(ones);
    }
    for (int32_t h = 0; h < nhalo; ++h) { free(rc[h].cand_np); free(rc[h].member_ptr); free(rc[h].members); }
    free(rc); free(start);
    return ierr;
}

/*
 * Candidate list as the reference builds it before unbinding
 * (fof_fortran.f90:263-308): for each candidate in global order, its members
 * in CRS order. Returns the number of candidates; fills cand_np/members when
 * the buffers are large enough (cap_cand, cap_mem), else only counts.
 */
int32_t oracle_fof_candidates(const double *xp, const int32_t *np_halo, int32_t nhalo,
                              double L, double b, int32_t np_min, int32_t total_np,
                              int32_t cap_cand, int32_t cap_mem, int32_t *cand_np,
                              int32_t *cand_region, int32_t *members, int32_t *n_members)
{
    (void)total_np;
    double Lh = 0.5 * L;
    int32_t start = 0, nc = 0, nm = 0;
    for (int32_t h = 0; h < nhalo; ++h) {
        int32_t np = np_halo[h];
        if (np >= np_min && np > 0) {
            region_cands_t rc;
            if (oracle_process_region(xp, start, np, L, Lh, b, np_min, 1, &rc, NULL, NULL)) return -3;
            for (int32_t g = 0; g < rc.ncand; ++g) {
                if (nc < cap_cand) { cand_np[nc] = rc.cand_np[g]; cand_region[nc] = h; }
                for (int32_t k = 0; k < rc.cand_np[g]; ++k) {
                    if (nm < cap_mem) members[nm] = rc.members[rc.member_ptr[g] + k];
                    ++nm;
                }
                ++nc;
            }
            free(rc.cand_np); free(rc.member_ptr); free(rc.members);
        }
        start += np;
    }
    *n_members = nm;
    return nc;
}

/* Region geometry from a region's particles, for tests of the geometry rule. */
int32_t oracle_geometry_of(const double *xp, int32_t np, double L, double b, oracle_geom_t *g)
{
    if (np <= 0) return 1;
    double Lh = 0.5 * L, bmin[3], bmax[3];
    for (int32_t i = 0; i < np; ++i)
        for (int d = 0; d < 3; ++d) {
            double v = oracle_modulo(xp[3 * (size_t)i + d] - xp[d] + Lh, L) - Lh;
            if (i == 0 || v < bmin[d]) bmin[d] = v;
            if (i == 0 || v > bmax[d]) bmax[d] = v;
        }
    oracle_region_geometry(bmin, bmax, np, L, b, g);
    return 0;
}
