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
    // further streams of the linking phase: the sparse-particle kernels run
    // on stream2 and the mid-cell kernels on stream3, beside the dense-cell
    // kernels on stream (fork after binning; ev_dense / ev_mid: the local
    // components of dense / mid cells are written; joins before the relabel).
    // In the binning, the coordinate gather runs on stream2 beside the cell
    // starts and classes (ev_sort: keys sorted; ev_gather: gather done).
    hipStream_t stream2 = nullptr, stream3 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_dense = nullptr, ev_mid = nullptr, ev_join = nullptr, ev_pair = nullptr,
               ev_sort = nullptr, ev_gather = nullptr;
    hipEvent_t ev[kNumEv] = {};
    bool ev_recorded[kNumEv] = {};
    DevBuf xp, vp, mp, chunks, regions, regstart, cell_off, bbox, keys0, keys1, vals0, vals1, ends,
        xc, parent, label, size, flag, aux, dense, mid, sub_ptr, sub_list, subs, ncomp, croot, needs_sub, sub_root, temp, out_in, out_id,
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
    // FFT convolution lines per workgroup for the grids 48, 64, 128, 216 and
    // 256 (PYFOF_OPT_FFT_LINES; a launch shape, the results do not change)
    int fft_lines[5] = {8, 8, 8, 8, 8};
    static int fft_size_index(int n) {
        switch (n) {
            case 48: return 0;
            case 64: return 1;
            case 128: return 2;
            case 216: return 3;
            case 256: return 4;
            default: return -1;
        }
    }
    int fft_lines_for(int n) const {
        const int i = fft_size_index(n);
        return i < 0 ? 8 : fft_lines[i];
    }
    // large candidates may use multi-workgroup cooperative launches; off for
    // contexts that share the GPU with others (the request server's workers),
    // where the co-residency those launches need is not assured
    bool solo = true;
    // PYFOF_OPT_TIME_FFT: HIP events around every FFT convolution launch of
    // the unbinding (pairs, kept between calls), their summed device time and
    // the convolutions' algorithmic bytes (48 n^3 per slot: forward and
    // inverse transforms, three axis passes, each reading and writing 8 B per
    // node of the n^3 complex grid)
    bool time_fft = false;
    std::vector<hipEvent_t> fft_ev;
    size_t fft_ev_used = 0;
    double fft_ms = 0.0, fft_bytes = 0.0;
    int64_t fft_slots = 0;
    void fft_begin() {
        if (!time_fft) return;
        while (fft_ev.size() < fft_ev_used + 2) {
            hipEvent_t e = nullptr;
            PYFOF_HIP(hipEventCreate(&e));
            fft_ev.push_back(e);
        }
        PYFOF_HIP(hipEventRecord(fft_ev[fft_ev_used], stream));
    }
    void fft_end(int n, int nslots) {
        if (!time_fft) return;
        PYFOF_HIP(hipEventRecord(fft_ev[fft_ev_used + 1], stream));
        fft_ev_used += 2;
        fft_bytes += 48.0 * (double)n * (double)n * (double)n * (double)nslots;
        fft_slots += nslots;
    }
    // after the stream has drained
    void fft_collect() {
        fft_ms = 0.0;
        for (size_t k = 0; k + 1 < fft_ev_used; k += 2) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, fft_ev[k], fft_ev[k + 1]) == hipSuccess) fft_ms += t;
        }
    }
    void fft_reset() { fft_ev_used = 0; fft_ms = fft_bytes = 0.0; fft_slots = 0; }
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
        PYFOF_HIP(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
        PYFOF_HIP(hipStreamCreateWithFlags(&stream3, hipStreamNonBlocking));
        for (auto& e : ev) PYFOF_HIP(hipEventCreate(&e));
        for (hipEvent_t* e : {&ev_fork, &ev_dense, &ev_mid, &ev_join, &ev_pair, &ev_sort, &ev_gather})
            PYFOF_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
    }
    ~Ctx() {
        (void)hipSetDevice(device);
        for (auto& e : ev) if (e) (void)hipEventDestroy(e);
        for (auto& e : fft_ev) if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : {ev_fork, ev_dense, ev_mid, ev_join, ev_pair, ev_sort, ev_gather})
            if (e) (void)hipEventDestroy(e);
        if (stream3) (void)hipStreamDestroy(stream3);
        if (stream2) (void)hipStreamDestroy(stream2);
        if (stream) (void)hipStreamDestroy(stream);
        release_workspace();
    }
    template <class F>
    void for_each_buf(F f) {
        for (DevBuf* b : {&xp, &vp, &mp, &chunks, &regions, &regstart, &cell_off, &bbox, &keys0, &keys1, &vals0,
                          &vals1, &ends, &xc, &parent, &label, &size, &flag, &aux, &dense,
                          &sub_ptr, &sub_list, &subs, &ncomp, &croot, &needs_sub, &sub_root, &temp, &out_in, &out_id, &out_root, &anchor,
                          &hregion, &scal, &cat_keys, &cat_vals, &cat_keys2, &cat_vals2, &hoff,
                          &am_h, &am, &pin, &rec_phi, &rec_idx})
            f(*b);
        for (auto& kv : named) f(kv.second);
        for (auto& kv : greens) {
            f(kv.second->G);
            f(kv.second->tw);
            f(kv.second->self);
        }
    }
    // Device bytes this context holds between calls (its workspace and the
    // cached Green functions).
    size_t resident_bytes() {
        size_t t = 0;
        for_each_buf([&](DevBuf& b) { if (b.p && !b.carved) t += b.cap; });
        return t;
    }
    // Frees the workspace and the caches; the next call allocates again. The
    // context's stream must be idle.
    void release_workspace() {
        for_each_buf([](DevBuf& b) { b.release(); });
        for (auto& kv : greens) delete kv.second;
        greens.clear();
        named.clear();
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
