// FoF device kernels (binning, linking with a lock-free union-find, relabel).
#pragma once
#include "common.hpp"
#include "fof_geom.hpp"

namespace pyfof {

struct Chunk {
    int32_t region, begin, end;
};

struct LinkParams {
    double L, Lh;   // box length and half length
    double b, rp2;  // linking length and its square (f64, as the reference)
    double R, m;    // pruning query half-width (b + m) and rounding margin m
    int32_t np_min, nregions;
    uint32_t N, ncell;
    int32_t sub_stride;
    int32_t intra_stage;  // dense-cell members staged in LDS by k_dense_intra
    // f32 screen of the periodic predicate: squared distances below scr_lo
    // link and above scr_hi do not (b^2 (1 -+ 1e-6) rounded to float); pairs
    // with a coordinate difference of scr_wrap (Lh / 2) or more use f64
    float scr_lo, scr_hi, scr_wrap;
};

constexpr int kDenseMin = 256;    // cells with at least this many particles get sub-cells
constexpr int kMaxSubPerDim = 16; // sub-cells per dim; at most 4096 per cell

// gfortran real MODULO: exact fmod, shifted into the sign of p. For p > 0
// and -p < a < 2p the result is found without fmod, identically: a itself
// on [0, p); a - p on [p, 2p), exact by Sterbenz's lemma, as fmod is; and on
// (-p, 0) fmod returns a, which is then shifted by the same rounded a + p.
__device__ __forceinline__ double f_modulo(double a, double p) {
    if (p > 0.0) {
        if (a >= 0.0 && a < p) return a;
        if (a >= p && a < 2.0 * p) return a - p;
        if (a < 0.0 && a > -p) return a + p;
    }
    double r = fmod(a, p);
    if (r != 0.0 && ((r < 0.0) != (p < 0.0))) r += p;
    return r;
}

// fof_fortran.f90:525 -- x_rel = MODULO(x - x_anchor + L/2, L) - L/2
__device__ __forceinline__ double rel_coord(double x, double xa, double L, double Lh) {
    return f_modulo((x - xa) + Lh, L) - Lh;
}

// fof_fortran.f90:581-583 -- min(N, max(1, floor((x - o)/c) + 1)), 1-based
__device__ __forceinline__ int32_t ref_cell_coord(double x, double o, double c, int32_t n) {
    double f = floor((x - o) / c);
    f = fmin(fmax(f, -4.0e9), 4.0e9);
    long long i = (long long)f + 1;
    i = i < 1 ? 1 : i;
    return (int32_t)(i > n ? n : i);
}

// 0-based sub-cell coordinate inside a cell whose lower corner is lo.
__device__ __forceinline__ int32_t sub_coord(double x, double lo, double s, int32_t k) {
    double f = floor((x - lo) / s);
    f = fmin(fmax(f, -1.0), (double)k);
    int32_t i = (int32_t)f;
    return i < 0 ? 0 : (i >= k ? k - 1 : i);
}

__device__ __forceinline__ uint32_t region_of_cell(const uint32_t* __restrict__ cell_off, int32_t nreg,
                                                   uint32_t c) {
    int32_t lo = 0, hi = nreg - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (cell_off[mid] <= c) lo = mid; else hi = mid - 1;
    }
    return (uint32_t)lo;
}

// Periodic predicate, fof_fortran.f90:483-491: float differences, double
// minimum image, (dx*dx + dy*dy) + dz*dz without contraction.
__device__ __forceinline__ bool pred_periodic(const float4& a, const float4& q, double L, double Lh,
                                              double rp2) {
    double dx = fabs((double)(a.x - q.x));
    double dy = fabs((double)(a.y - q.y));
    double dz = fabs((double)(a.z - q.z));
    if (dx > Lh) dx = L - dx;
    if (dy > Lh) dy = L - dy;
    if (dz > Lh) dz = L - dz;
    const double r2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
    return r2 <= rp2;
}

// Non-periodic predicate, fof_fortran.f90:636: sum over real(4) of squared
// float differences, accumulated in float, widened to double.
__device__ __forceinline__ bool pred_flat(const float4& a, const float4& q, double rp2) {
    const float dx = a.x - q.x, dy = a.y - q.y, dz = a.z - q.z;
    float s = __fmul_rn(dx, dx);
    s = __fadd_rn(s, __fmul_rn(dy, dy));
    s = __fadd_rn(s, __fmul_rn(dz, dz));
    return (double)s <= rp2;
}

// Union-find over CRS positions; parent[x] == x marks a root. Parents always
// point to smaller positions, so a component's root is its first CRS member
// (atomic_dsu.c:31-35): the root of every tree is its minimum position.
//
// dsu_find reads at agent scope (bypassing the per-CU cache): every value is
// the current one. The relabel uses it, after linking, for the exact roots.
__device__ __forceinline__ int32_t dsu_find(int32_t* parent, int32_t x) {
    for (;;) {
        const int32_t px = atomic_load_relaxed(&parent[x]);
        if (px == x) return x;
        const int32_t gp = atomic_load_relaxed(&parent[px]);
        if (gp == px) return px;
        __hip_atomic_store(&parent[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = gp;
    }
}

// A relaxed load at workgroup scope: served by the CU's vector L1 when the
// line is there, so it may return an older value than another CU's latest
// store to *p.
__device__ __forceinline__ int32_t load_near(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// dsu_find for the linking kernels, through the L1. An older parent value is
// still a smaller position of the same component (components only merge and
// pointers only decrease), so the result is a node of x's component, at most
// x, that was a root when its line was cached; path halving with such values
// still stores pointers to smaller positions of the component.
__device__ __forceinline__ int32_t dsu_find_near(int32_t* parent, int32_t x) {
    for (;;) {
        const int32_t px = load_near(&parent[x]);
        if (px == x) return x;
        const int32_t gp = load_near(&parent[px]);
        if (gp == px) return px;
        __hip_atomic_store(&parent[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = gp;
    }
}

// Joins the components of a and b; returns a node of the joined component.
// The hook is a CAS at L2 that succeeds only on a current root. Equal finds
// mean one component. A hook of a current root hi under lo < hi is always
// valid: hi is its component's minimum, so lo lies in the other component.
// A failed CAS returns hi's current parent (read at L2), and the search goes
// on from there: every failure lowers a + b, so the loop ends even when the
// L1 still holds older values.
__device__ __forceinline__ int32_t dsu_unite(int32_t* parent, int32_t a, int32_t b) {
    for (;;) {
        a = dsu_find_near(parent, a);
        b = dsu_find_near(parent, b);
        if (a == b) return a;
        const int32_t lo = a < b ? a : b, hi = a < b ? b : a;
        const int32_t old = atomicCAS(&parent[hi], hi, lo);
        if (old == hi) return lo;
        if (a == hi) a = old;
        else b = old;
    }
}

}  // namespace pyfof
