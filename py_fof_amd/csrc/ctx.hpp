// Per-device context: stream, events and the reusable HBM workspace.
#pragma once
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "fft.hpp"



namespace pyfof {

// Uploads the constant tables of the FoF kernels to the current device.
void fof_init_device_tables();

// PYFOF_OPT_POISON_ALLOC: 0, or 1 + the byte every new allocation is filled with
inline std::atomic<int> g_poison_alloc{0};

// A bump allocator over caller-provided device memory. While a thread has
// one installed (ArenaScope), every DevBuf::get on that thread carves from it
// instead of calling hipMalloc; a buffer asked for again within the same
// epoch (one call) keeps its memory. Carved buffers are never freed.
struct Arena {
    unsigned char* base = nullptr;
    size_t size = 0, off = 0;
    uint64_t epoch = 0;
    void reset() { off = 0; ++epoch; }
};
inline thread_local Arena* tl_arena = nullptr;

struct ArenaScope {
    explicit ArenaScope(Arena* a) { a->reset(); tl_arena = a; }
    ~ArenaScope() { tl_arena = nullptr; }
};

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    uint64_t epoch = 0;    // arena epoch of a carved buffer
    bool carved = false;   // p points into an arena, not a hipMalloc block
    void* carve(Arena& a, size_t bytes) {
        if (carved && epoch == a.epoch && cap >= bytes) return p;
        const size_t at = (a.off + 255) & ~(size_t)255;
        if (at > a.size || bytes > a.size - at)
            throw HipError(hipErrorOutOfMemory,
                           ("the input needs more than the " + std::to_string(a.size >> 20) +
                            " MiB of static scratch of the allocation-free entry point").c_str());
        a.off = at + bytes;
        p = a.base + at;
        cap = bytes;
        epoch = a.epoch;
        carved = true;
        return p;
    }
    void* get(size_t bytes) {
        if (bytes == 0) bytes = 16;
        if (tl_arena) return carve(*tl_arena, bytes);
        if (carved) { p = nullptr; cap = 0; carved = false; }
        if (bytes > cap) {
            if (p) PYFOF_HIP(hipFree(p));
            p = nullptr;
            cap = 0;
            const size_t want = bytes + bytes / 8 + 256;
            hipError_t e = hipMalloc(&p, want);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                p = nullptr;
                throw HipError(hipErrorOutOfMemory, "device allocation failed");
            }
            cap = want;
            const int poison = g_poison_alloc.load();
            if (poison) {
                // completes before any stream uses the buffer
                PYFOF_HIP(hipMemset(p, poison - 1, want));
                PYFOF_HIP(hipDeviceSynchronize());
            }
        }
        return p;
    }
    template <class T>
    T* as(size_t count) { return static_cast<T*>(get(count * sizeof(T))); }
    void release() {
        if (p && !carved) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        carved = false;
    }
};

// Outcome of the last unbinding (mirrors BoundResult).
struct BoundStats {
    uint32_t nvalid = 0;
    uint32_t n_ok = 0, n_too_few = 0, n_max_iter = 0;
    uint32_t n_direct = 0, n_pm[4] = {0, 0, 0, 0};
    int32_t passes = 0;
    uint32_t n_direct_large = 0;
    std::string err_msg;  // reason of a failed call, empty on success
};

// Green function of one isolated grid size (depends on n only), cached.
struct GreenCache {
    int n = 0;
    DevBuf G;      // full-spectrum real G, n^3 floats [z][y][x]
    DevBuf tw;     // FFT twiddles
    DevBuf self;   // 5^3 self kernel (device)
    FftPlan plan;
    bool ready = false;
};

enum { kEvStart = 0, kEvBbox, kEvBin, kEvLink, kEvOut, kEvBound, kEvCat, kNumEv };

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[kNumEv] = {};
    bool ev_recorded[kNumEv] = {};
    DevBuf xp, vp, mp, chunks, regions, regstart, cell_off, bbox, keys0, keys1, vals0, vals1, ends,
        xc, parent, label, size, flag, cidx, aux, dense, mid, sub_ptr, sub_list, subs, ncomp, croot, needs_sub, temp, out_in, out_id,
        out_root, anchor, hregion, scal, cat_keys, cat_vals, cat_keys2, cat_vals2, hoff, am_h, am,
        pin;
    // last call's sizes and catalog
    uint64_t last_n = 0, last_ncell = 0, last_ndense = 0;
    bool want_catalog = false;
    // unbinding form (PYFOF_OPT_DIRECT_MAX / _DIRECT_FP32 / _BATCHED / _RECORD_PHI)
    int32_t direct_max = 1000;
    int32_t direct_f32 = 0;
    int32_t batched = 1;
    bool record_phi = false;
    // large candidates may use multi-workgroup cooperative launches; off for
    // contexts that share the GPU with others (the request server's workers),
    // where the co-residency those launches need is not assured
    bool solo = true;
    uint64_t rec_n = 0;  // recorded first-pass potentials of the last call
    DevBuf rec_phi, rec_idx;
    BoundStats last_bound;  // build the fof_with_info catalog on every call
    uint32_t cat_nh = 0;
    uint64_t cat_nmem = 0;
    uint32_t* cat_members = nullptr;  // device, grouped by halo
    uint32_t* cat_hoff = nullptr;     // device, inclusive ends
    std::mutex mu;
    std::map<std::string, DevBuf> named;  // workspace of the unbinding solver
    std::map<int, GreenCache*> greens;
    DevBuf& buf(const std::string& name) { return named[name]; }

    explicit Ctx(int dev) : device(dev) {
        PYFOF_HIP(hipSetDevice(dev));
        fof_init_device_tables();
        PYFOF_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        for (auto& e : ev) PYFOF_HIP(hipEventCreate(&e));
    }
    ~Ctx() {
        (void)hipSetDevice(device);
        for (auto& e : ev) if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
        for (DevBuf* b : {&xp, &vp, &mp, &chunks, &regions, &regstart, &cell_off, &bbox, &keys0, &keys1, &vals0,
                          &vals1, &ends, &xc, &parent, &label, &size, &flag, &cidx, &aux, &dense,
                          &sub_ptr, &sub_list, &subs, &ncomp, &croot, &needs_sub, &temp, &out_in, &out_id, &out_root, &anchor,
                          &hregion, &scal, &cat_keys, &cat_vals, &cat_keys2, &cat_vals2, &hoff,
                          &am_h, &am, &pin})
            b->release();
        for (auto& kv : named) kv.second.release();
        rec_phi.release();
        rec_idx.release();
        for (auto& kv : greens) {
            kv.second->G.release();
            kv.second->tw.release();
            kv.second->self.release();
            delete kv.second;
        }
    }
    void mark(int e) {
        PYFOF_HIP(hipEventRecord(ev[e], stream));
        ev_recorded[e] = true;
    }
};


// Arguments of one FoF (+ unbinding) call, as pyfof_fof takes them.
struct CallArgs {
    const double* xp;
    const double* vp;
    const double* mp;
    const int32_t* np_halo;
    int32_t nhalo;
    double L, b;
    int32_t np_min;
    int32_t bound_check;
    int32_t verbose;
    double softening, a_val, b_val, s_15;
    int64_t ng;
    int32_t total_np;
    int32_t* in_halo;
    int32_t* halo_id;
    double* ang_mom;
    int32_t* root_out;
    int32_t mem_kind;
    double ng_global = 512.0;  // only read by the unbinding's error report
};

// Runs one call on C (validation, then the pipeline); returns ierr. The
// message of a failure is last_error_message() on the calling thread.
int32_t fof_entry(Ctx* C, const CallArgs& A, int32_t n_threads);
const std::string& last_error_message();
void set_last_error(const std::string& m);
// Zeroes in_halo, halo_id and ang_mom of a call that ends without a result.
void clear_outputs(Ctx* C, const CallArgs& A);

}  // namespace pyfof
