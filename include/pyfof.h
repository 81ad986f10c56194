/*
 * C-ABI of libpyfof_hip.so -- the MI355X (gfx950) FoF + unbinding library.
 *
 * Plain pointers and sizes only. Error handling follows the reference's
 * integer ierr convention (fof_fortran.f90:116-119, py_fof/__init__.py:338-348):
 *   0 ok, 1 invalid input/configuration, 2 candidate span > s_15,
 *   3 allocation failure, 4 non-finite values, 10 n_threads > 256,
 *   20 HIP runtime failure, 21 size beyond the device layout.
 * Nothing throws across this boundary; pyfof_last_error() gives the message.
 */
#ifndef PYFOF_H
#define PYFOF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Where the particle and output arrays of a call live. */
#define PYFOF_MEM_HOST 0   /* host memory: the library stages through HBM */
#define PYFOF_MEM_DEVICE 1 /* HBM of the context's device: used in place */

typedef struct pyfof_ctx pyfof_ctx;

/*
 * Replaces the f2py entry point fof_fortran.fof_fortran(...) that
 * py_fof/__init__.py:331-336 calls (declared at fof_fortran.f90:10-58):
 * same argument order and meaning, same ierr codes, caller-allocated outputs.
 *   xp, vp      (3, total_np) f64, Fortran order (x0,y0,z0,x1,...); vp may be
 *               NULL when bound_check == 0 (zero velocities).
 *   mp          (total_np) f64; may be NULL (unit masses).
 *   np_halo     (nhalo) int32 region sizes, always host memory.
 *   in_halo, halo_id  (total_np) int32 outputs.
 *   ang_mom     (3, nhalo) f64 Fortran order output; may be NULL to skip the
 *               per-region angular momentum (it is then not computed).
 *   mem_kind    PYFOF_MEM_HOST or PYFOF_MEM_DEVICE for xp/vp/mp and outputs.
 * Runs on the calling thread's current HIP device with a per-device default
 * context; calls on one device are serialised.
 */
int32_t pyfof_fof(const double* xp, const double* vp, const double* mp, const int32_t* np_halo,
                  int32_t nhalo, double L_Box, double linking_length, int32_t np_halo_min,
                  int32_t bound_check, int32_t verbose, double softening, double a_val,
                  double b_val, double s_15_val, int32_t n_threads, int32_t total_np,
                  int64_t ng_val, double ng_global_val, int32_t* in_halo, int32_t* halo_id,
                  double* ang_mom, int32_t mem_kind);

/* Frees the per-device default contexts pyfof_fof uses (streams and HBM
 * workspaces), as bound_release (bound_fortran.f90:1816-1840) frees the
 * reference's module buffers and FFT plans. A later pyfof_fof creates them
 * again. Must not overlap a pyfof_fof call. Call it before the process
 * exits, while the HIP runtime is still up. */
void pyfof_release(void);

/*
 * Allocation-free variant of pyfof_fof, for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops, harnesses
 * that forbid hipMalloc). It has no counterpart in the reference. Same
 * arguments and ierr codes as pyfof_fof, with every array (xp, vp, mp,
 * in_halo, halo_id, ang_mom) in HBM of the calling thread's current device;
 * np_halo stays in host memory. Plain FoF only: bound_check != 0 returns
 * ierr 1. At most PYFOF_NOALLOC_MAX_N particles (ierr 21 beyond); every
 * buffer of the call comes from one static scratch array of
 * PYFOF_NOALLOC_SCRATCH_MIB MiB per device, held until the process ends, and
 * an input whose cell grid does not fit it returns ierr 3. The call makes no
 * hipMalloc or hipFree (the first call on a device looks up the scratch
 * address and creates the stream it runs on). Calls on one device share that
 * scratch, which keeps the last call's contents: the library serialises them
 * (a call returns with its device work finished), so overlapping calls wait
 * instead of racing on it. Other entry points never use the scratch.
 */
#define PYFOF_NOALLOC_MAX_N 524288
#define PYFOF_NOALLOC_SCRATCH_MIB 128
int32_t pyfof_fof_noalloc(const double* xp, const double* vp, const double* mp,
                          const int32_t* np_halo, int32_t nhalo, double L_Box,
                          double linking_length, int32_t np_halo_min, int32_t bound_check,
                          int32_t verbose, double softening, double a_val, double b_val,
                          double s_15_val, int32_t n_threads, int32_t total_np, int64_t ng_val,
                          double ng_global_val, int32_t* in_halo, int32_t* halo_id,
                          double* ang_mom);

/* Context API: explicit device, own stream and reusable HBM workspace. */
pyfof_ctx* pyfof_ctx_create(int32_t device);
void pyfof_ctx_destroy(pyfof_ctx* ctx);

/* Same as pyfof_fof on an explicit context. root_out (optional, total_np
 * int32, same memory kind) receives for every particle the 0-based index of
 * the first member, in the reference's cell-sorted (CRS) order, of its
 * friends-of-friends component: the canonical label of the partition
 * (fof_fortran.f90:452-457, atomic_dsu.c:31-35). -1 for skipped regions. */
int32_t pyfof_fof_ctx(pyfof_ctx* ctx, const double* xp, const double* vp, const double* mp,
                      const int32_t* np_halo, int32_t nhalo, double L_Box, double linking_length,
                      int32_t np_halo_min, int32_t bound_check, int32_t verbose, double softening,
                      double a_val, double b_val, double s_15_val, int32_t n_threads,
                      int32_t total_np, int64_t ng_val, double ng_global_val, int32_t* in_halo,
                      int32_t* halo_id, double* ang_mom, int32_t* root_out, int32_t mem_kind);

/* Context options (no counterpart in the reference, whose solver form is
 * fixed; the defaults reproduce it):
 *   PYFOF_OPT_CATALOG (default 0): every call also groups halo members for
 *     pyfof_catalog_* (fof_with_info).
 *   PYFOF_OPT_DIRECT_MAX (default 1000, range [1000, 65536]): candidates of
 *     up to this many members are unbound by direct summation; larger ones
 *     by the particle mesh. 1000 is the reference's rule
 *     (bound_fortran.f90:1976-1980).
 *   PYFOF_OPT_DIRECT_FP32 (default 0): direct-summation pair terms and
 *     potential sums in f32; centre-of-mass and bulk-velocity sums stay f64.
 *     0 sums in FP64 exactly as the reference (bound_fortran.f90:2613-2647).
 *   PYFOF_OPT_BATCHED (default 1): all candidates of a kind advance together,
 *     one launch per step; 0 unbinds one candidate at a time (plain form).
 *   PYFOF_OPT_RECORD_PHI (default 0): keep the first-pass potential of every
 *     candidate member for pyfof_ctx_potential.
 *   PYFOF_OPT_TIME_FFT (default 0): HIP events around every FFT
 *     convolution launch of the unbinding; pyfof_ctx_timings [6] and
 *     pyfof_ctx_stats [16..17] report their device time, slot count and
 *     algorithmic bytes.
 *   PYFOF_OPT_FFT_LINES (value 1000 * n + lines; default 8 lines for every
 *     n): grid lines per workgroup (4, 8 or 16) of the FFT convolutions on
 *     n^3 grids, n in {48, 64, 128, 216, 256}. A launch shape: results do
 *     not depend on it. py_fof_amd.tune.tune_fft_lines measures the best.
 *   PYFOF_OPT_POISON_ALLOC (default 0; process-wide, not per context): a
 *     debugging aid. 0 = off; 1 + b fills every device allocation made from
 *     then on with the byte b (0..255), so that a read of memory no kernel
 *     has written changes the result. Allocation becomes slower.
 * Returns 0, or 1 for an unknown option or an out-of-range value. */
#define PYFOF_OPT_CATALOG 1
#define PYFOF_OPT_DIRECT_MAX 2
#define PYFOF_OPT_DIRECT_FP32 3
#define PYFOF_OPT_BATCHED 4
#define PYFOF_OPT_RECORD_PHI 5
#define PYFOF_OPT_POISON_ALLOC 6
#define PYFOF_OPT_TIME_FFT 7
#define PYFOF_OPT_FFT_LINES 8
int32_t pyfof_ctx_set_option(pyfof_ctx* ctx, int32_t option, int64_t value);

/* First-pass potentials recorded by the last bound_check call on ctx (with
 * PYFOF_OPT_RECORD_PHI): phi (cube units, f32) and the particle index of
 * every FoF-candidate member, candidates in halo order, members ascending.
 * With phi and index both NULL returns the count; otherwise cap must be at
 * least the count. Returns the count, or -1 on error. */
int64_t pyfof_ctx_potential(pyfof_ctx* ctx, float* phi, int32_t* index, int64_t cap, int32_t mem_kind);

/*
 * Catalog of the last call on ctx, as fof_with_info builds it on the host
 * (py_fof/__init__.py:481-507): halos in halo_id order, each with its member
 * indices ascending. pyfof_catalog_size gives nhalo_found and the total
 * member count; pyfof_catalog_copy fills offsets (nhalo_found + 1) and
 * indices (total members), in memory of kind mem_kind.
 */
int32_t pyfof_catalog_size(pyfof_ctx* ctx, int32_t* nhalo_found, int64_t* n_members);
int32_t pyfof_catalog_copy(pyfof_ctx* ctx, int64_t* offsets, int32_t* indices, int32_t mem_kind);

/* Device time in milliseconds of the phases of the last call on ctx
 * (0: bbox+geometry, 1: binning, 2: linking, 3: relabel/outputs,
 * 4: unbinding, 5: catalog+angular momentum; 6: the unbinding's FFT
 * convolutions, with PYFOF_OPT_TIME_FFT), measured with HIP events on the
 * context's stream. Returns the number of values written. */
int32_t pyfof_ctx_timings(pyfof_ctx* ctx, double* ms, int32_t n);

/* Sizes of the last call on ctx: [0] particles, [1] reference cells,
 * [2] dense cells given sub-cells, [3] halos in the output, [4] members of
 * those halos; with bound_check: [5] valid bound halos, [6..8] candidates
 * ending OK / TOO_FEW / MAX_ITER, [9] direct-summation candidates,
 * [10..13] candidates on the 48/64/128/256 grids, [14] potential passes,
 * [15] direct-summation candidates above 1000 members; with
 * PYFOF_OPT_TIME_FFT [16] FFT convolutions (slots) and [17] their
 * algorithmic bytes (48 n^3 each).
 * Returns the number of values written. */
int32_t pyfof_ctx_stats(pyfof_ctx* ctx, int64_t* out, int32_t n);

/* HIP stream (hipStream_t) the context launches on. */
void* pyfof_ctx_stream(pyfof_ctx* ctx);

/*
 * Request server (continuous batching on one GPU). The reference has no
 * counterpart: its single entry point (fof_fortran.f90:10-58) keeps module
 * state and must be called serially (bound_fortran.f90:1524, 1564-1568).
 * A server runs `slots` workers, each with its own context (stream and
 * HBM-resident workspace). Submitted jobs queue in arrival order; the head
 * starts when a worker is free and its estimated HBM footprint
 * (pyfof_job_footprint) fits in hbm_budget_bytes beside the running jobs.
 * Jobs on different workers overlap on the device. Each job has its own ierr
 * and message. The arrays a job points to (xp, vp, mp, outputs) must stay
 * valid until pyfof_poll/pyfof_wait report it done; np_halo is copied.
 */
typedef struct pyfof_server pyfof_server;

/* One job: the arguments of pyfof_fof, by name. */
typedef struct pyfof_job_desc {
    const double* xp;
    const double* vp;
    const double* mp;
    const int32_t* np_halo;
    int32_t* in_halo;
    int32_t* halo_id;
    double* ang_mom;
    double L_Box, linking_length, softening, a_val, b_val, s_15_val, ng_global_val;
    int64_t ng_val;
    int32_t nhalo, np_halo_min, bound_check, verbose, n_threads, total_np, mem_kind, reserved;
} pyfof_job_desc;

/* slots in [1, 16]; NULL on failure (pyfof_server_last_error). Destroying a
 * server waits for every queued and running job. */
pyfof_server* pyfof_server_create(int32_t device, int64_t hbm_budget_bytes, int32_t slots);
void pyfof_server_destroy(pyfof_server* server);

/* Estimated device bytes of a job, as used for admission. */
int64_t pyfof_job_footprint(const pyfof_job_desc* job);

/* Queues a job; returns its id (> 0), or -1 (message in
 * pyfof_server_last_error) when the job is malformed or its footprint
 * exceeds the whole budget. */
int64_t pyfof_submit(pyfof_server* server, const pyfof_job_desc* job);

/* pyfof_poll: 1 done (ierr, message and times written, the id released),
 * 0 still queued or running, -1 unknown id. pyfof_wait blocks until done.
 * times (optional, 2 values): ms queued before start, ms running. */
int32_t pyfof_poll(pyfof_server* server, int64_t id, int32_t* ierr, char* msg, int32_t msg_cap,
                   double* times);
int32_t pyfof_wait(pyfof_server* server, int64_t id, int32_t* ierr, char* msg, int32_t msg_cap,
                   double* times);

/* [0] queued jobs, [1] running jobs, [2] footprint of the running jobs,
 * [3] budget, [4] bytes the workers' contexts hold between jobs, [5] bytes
 * claimed in all (per worker the larger of [4]'s share and its job's
 * footprint; admission keeps it within the budget). */
int32_t pyfof_server_status(pyfof_server* server, int64_t* out, int32_t n);
const char* pyfof_server_last_error(void);

/* Message of the last error on this thread ("" if none). */
const char* pyfof_last_error(void);

/* Library version string. */
const char* pyfof_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PYFOF_H */
