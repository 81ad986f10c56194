// Region geometry: the reference's cell rule plus this library's sub-cell plan.
#pragma once
#include <math.h>
#include <stdint.h>

namespace pyfof {

// One input region as the linking kernels see it. The reference's cell grid
// (origin, cell, n, periodic) is reproduced exactly (fof_fortran.f90:518-570);
// k/sub describe the sub-cells this library adds inside dense cells.
struct RegionDev {
    double origin[3];
    double cell[3];
    double sub[3];      // sub-cell side per dim
    int32_t n[3];       // reference cells per dim
    int32_t k[3];       // sub-cells per reference cell per dim (1 = none)
    int32_t periodic[3];
    int32_t is_periodic;
    int32_t skip;       // region not linked (np < np_halo_min or np == 0)
    int32_t clique;     // every pair inside one sub-cell passes the predicate
    int32_t prune;      // geometric pruning allowed (all periodic dims have n >= 4)
    uint32_t cell_off;  // first global cell id of this region
    uint32_t ncell;
    int32_t start, np;
};

// Reference geometry, fof_fortran.f90:518-570, host side. Uses libm pow like
// gfortran's (vol/np)**(1/3) and int32 arithmetic for the refinement test.
inline void reference_geometry(const double bmin[3], const double bmax[3], int32_t np, double L,
                               double b, RegionDev& g) {
    auto dmax = [](double x, double y) { return x > y ? x : y; };
    auto imax = [](int32_t x, int32_t y) { return x > y ? x : y; };
    auto imin = [](int32_t x, int32_t y) { return x < y ? x : y; };
    double range[3];
    for (int d = 0; d < 3; ++d) {
        range[d] = bmax[d] - bmin[d];
        g.periodic[d] = (range[d] + 2.0 * b >= L) ? 1 : 0;
    }
    g.is_periodic = (g.periodic[0] || g.periodic[1] || g.periodic[2]) ? 1 : 0;
    const double vol = dmax(range[0], b) * dmax(range[1], b) * dmax(range[2], b);
    double lmesh = dmax(b, pow(vol / (double)imax(1, np), 1.0 / 3.0));
    for (int pass = 0; pass < 2; ++pass) {
        for (int d = 0; d < 3; ++d) {
            if (g.periodic[d]) {
                g.origin[d] = -0.5 * L;
                g.n[d] = imax(1, imin((int32_t)(L / b), (int32_t)(L / lmesh)));
                g.cell[d] = L / (double)g.n[d];
            } else {
                g.origin[d] = bmin[d];
                g.n[d] = imax(1, (int32_t)ceil(range[d] / lmesh));
                g.cell[d] = lmesh;
            }
        }
        if (pass == 1) break;
        const int32_t four_np = (int32_t)(4u * (uint32_t)np);
        if ((double)g.n[0] * (double)g.n[1] * (double)g.n[2] > (double)imax(262144, four_np)) {
            const int32_t two_np = (int32_t)(2u * (uint32_t)imax(1, np));
            lmesh = dmax(b, pow(vol / (double)two_np, 1.0 / 3.0));
        } else {
            break;
        }
    }
}

// Unit in the last place of a float at magnitude |a|.
inline double ulp_f32(double a) {
    a = fabs(a);
    if (!(a > 1e-30)) return 1e-45;
    int e;
    frexp(a, &e);  // a = f * 2^e, f in [0.5, 1)
    return ldexp(1.0, e - 24);
}

}  // namespace pyfof
