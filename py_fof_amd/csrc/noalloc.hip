// Allocation-free entry point (pyfof_fof_noalloc): plain FoF on device
// arrays with every buffer carved from one static scratch array per device.
// The scratch's device address is looked up once per device, with that
// device current, and only that address is handed to kernels and copies.
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pyfof.h"
#include "common.hpp"
#include "ctx.hpp"

namespace pyfof {

static __device__ __align__(256) unsigned char scratch[(size_t)PYFOF_NOALLOC_SCRATCH_MIB << 20];

struct NoAllocDev {
    Ctx* ctx = nullptr;   // stream, events and the carved buffers
    Arena arena;
    std::mutex mu;        // one call at a time owns the scratch
};
static std::mutex g_noalloc_mu;
static std::vector<NoAllocDev*> g_noalloc;

static NoAllocDev* noalloc_dev() {
    int dev = 0;
    PYFOF_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if ((int)g_noalloc.size() <= dev) g_noalloc.resize(dev + 1, nullptr);
    if (!g_noalloc[dev]) {
        void* ptr = nullptr;
        PYFOF_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        auto* d = new NoAllocDev;
        d->ctx = new Ctx(dev);
        d->arena.base = static_cast<unsigned char*>(ptr);
        d->arena.size = (size_t)PYFOF_NOALLOC_SCRATCH_MIB << 20;
        g_noalloc[dev] = d;
    }
    return g_noalloc[dev];
}


}  // namespace pyfof

using namespace pyfof;

extern "C" {

int32_t pyfof_fof_noalloc(const double* xp, const double* vp, const double* mp,
                          const int32_t* np_halo, int32_t nhalo, double L_Box,
                          double linking_length, int32_t np_halo_min, int32_t bound_check,
                          int32_t verbose, double softening, double a_val, double b_val,
                          double s_15_val, int32_t n_threads, int32_t total_np, int64_t ng_val,
                          double ng_global_val, int32_t* in_halo, int32_t* halo_id,
                          double* ang_mom) {
    CallArgs A{xp, vp, mp, np_halo, nhalo, L_Box, linking_length, np_halo_min, bound_check,
               verbose, softening, a_val, b_val, s_15_val, ng_val, total_np, in_halo, halo_id,
               ang_mom, nullptr, PYFOF_MEM_DEVICE};
    A.ng_global = ng_global_val;
    NoAllocDev* D = nullptr;
    try {
        D = noalloc_dev();
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return IERR_HIP;
    }
    if (bound_check) {
        set_last_error("pyfof_fof_noalloc runs plain FoF only (bound_check = 0)");
        std::lock_guard<std::mutex> lk(D->ctx->mu);
        clear_outputs(D->ctx, A);
        return IERR_INVALID;
    }
    if (total_np > PYFOF_NOALLOC_MAX_N) {
        set_last_error("pyfof_fof_noalloc takes at most " + std::to_string(PYFOF_NOALLOC_MAX_N) +
                  " particles");
        std::lock_guard<std::mutex> lk(D->ctx->mu);
        clear_outputs(D->ctx, A);
        return IERR_TOO_LARGE;
    }
    // every device buffer of the call is carved from this device's scratch;
    // the call ends with its stream drained, so the next one may reuse it
    std::lock_guard<std::mutex> lk(D->mu);
    ArenaScope scope(&D->arena);
    return fof_entry(D->ctx, A, n_threads);
}

}  // extern "C"
