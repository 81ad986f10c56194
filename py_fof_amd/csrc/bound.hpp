// Gravitational unbinding of FoF candidates on the device (bound_check=True).
#pragma once
#include "ctx.hpp"

namespace pyfof {

struct BoundArgs {
    // physical context (bound_fortran.f90:291-330)
    double app, a, b, s15;
    int64_t ng;
    double ng_global = 512.0;  // standard-grid size, for the span error report
    int32_t np_min, verbose;
    double L, Lh;
    // particle data in HBM (vp/mp may be null: zero velocities, unit masses)
    const double* xp;
    const double* vp;
    const double* mp;
    uint32_t N;
    // FoF candidates: nh halos, members grouped by halo (ascending index)
    uint32_t nh;
    uint32_t nmem;
    const uint32_t* hoff;     // inclusive member ends per halo
    const uint32_t* members;  // particle index per member slot
    const uint32_t* mhalo;    // halo of each member slot
    const uint32_t* anchor;   // first CRS member of each halo
    // outputs rewritten for bound members only
    int32_t* in_halo;
    int32_t* halo_id;
    // solver form: halos of up to direct_max members use direct summation
    // (1000 = the reference), in FP64 (reference) or f32 pair terms
    // (direct_f32); all halos of a kind per launch (batched) or one halo at
    // a time; phi_rec (optional, nmem floats) gets the first-pass potential
    int32_t direct_max = 1000;
    int32_t direct_f32 = 0;
    int32_t batched = 1;
    bool multi_wg = true;  // large candidates on several workgroups (k_pm_big)
    float* phi_rec = nullptr;
};

using BoundResult = BoundStats;

// Runs the unbinding on the context's stream. Returns an ierr code
// (0, 1, 2, 3, 4 as the reference); fills `res`. keep_out (device, nmem
// bytes) receives 1 for members that stay bound in a valid halo.
int32_t run_bound(Ctx& C, const BoundArgs& A, uint8_t* keep_out, BoundResult& res);

}  // namespace pyfof
