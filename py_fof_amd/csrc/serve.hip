// Request server: FoF (+ unbinding) jobs submitted as they arrive, run
// concurrently on one GPU.
//
// A server owns `slots` worker threads, each with its own context (its own
// HIP stream and a workspace that stays resident in HBM between jobs).
// Jobs wait in a FIFO queue; the head is admitted as soon as a worker is
// free and its estimated HBM footprint fits the budget next to the jobs
// already running. Jobs on different workers overlap on the device. Every
// job keeps its own ierr and message; one failing job does not affect the
// others. The reference has no counterpart: its entry point is not
// reentrant (bound_fortran.f90:1524, 1564-1568).
#include <chrono>
#include <condition_variable>
#include <algorithm>
#include <deque>
#include <memory>
#include <string.h>
#include <thread>
#include <unordered_map>

#include "../../include/pyfof.h"
#include "ctx.hpp"

namespace pyfof {

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    int64_t id = 0;
    CallArgs a{};
    int32_t n_threads = 0;
    int64_t footprint = 0;
    std::vector<int32_t> np_halo;  // copy: the caller may free its array
    bool done = false;
    int32_t ierr = 0;
    std::string msg;
    Clock::time_point t_submit, t_start, t_end;
};

}  // namespace

// Estimated device bytes of one job: the FoF workspace per particle (sort
// keys and values with their ping-pong copies 16 B, the input-order and the
// CRS float4 copies 32 B, DSU, labels and numbering 20 B, dense-cell tables
// up to 20 B, outputs, anchors and catalog 40 B, cells about 8 B, and the
// 1/8 growth slack of the workspace buffers), the staging of host arrays,
// and for the unbinding the per-member state (about 120 B) plus one 256^3
// grid pair (2 x 128 MiB) for the largest halo.
int64_t job_footprint(const CallArgs& a) {
    const int64_t n = a.total_np;
    int64_t b = n * 160 + (16ll << 20);
    if (a.mem_kind == PYFOF_MEM_HOST) b += n * (24 + (a.vp ? 24 : 0) + (a.mp ? 8 : 0) + 12);
    if (a.bound_check) b += n * 120 + (256ll << 20);
    return b;
}

// Admission counts what every worker holds, not only the jobs in flight: a
// worker's claim on HBM is the larger of the workspace its context keeps
// between jobs and the footprint of the job it runs. The head of the queue
// starts on an idle worker k when the other workers' claims plus the larger
// of k's resident workspace and the job's footprint fit the budget. When
// only resident workspaces stand in the way, the idle workers holding them
// free them (k its own at once; the others when asked), and a worker whose
// workspace grew past what the budget allows after a job frees it as well.
struct Server {
    int device = 0;
    int64_t budget = 0;
    int64_t next_id = 1;
    bool closing = false;
    bool trim = false;  // idle workers free their workspaces
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Job>> queue;
    std::unordered_map<int64_t, std::shared_ptr<Job>> jobs;
    std::vector<Ctx*> ctxs;
    std::vector<int64_t> resident;  // bytes each worker's context holds
    std::vector<int64_t> running;   // footprint of each worker's job, 0 when idle
    std::vector<std::thread> workers;

    Server(int dev, int64_t hbm_budget, int slots) : device(dev), budget(hbm_budget) {
        PYFOF_HIP(hipSetDevice(dev));
        for (int k = 0; k < slots; ++k) {
            ctxs.push_back(new Ctx(dev));
            ctxs.back()->solo = slots == 1;
        }
        resident.assign(slots, 0);
        running.assign(slots, 0);
        for (int k = 0; k < slots; ++k) workers.emplace_back([this, k] { run(k); });
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> lk(mu);
            closing = true;
        }
        cv_work.notify_all();
        for (auto& t : workers) t.join();
        for (Ctx* c : ctxs) delete c;
    }

    int64_t claim(size_t w) const { return std::max(resident[w], running[w]); }
    // claims of the workers other than k; busy_only leaves out idle ones
    int64_t others(size_t k, bool busy_only) const {
        int64_t t = 0;
        for (size_t w = 0; w < resident.size(); ++w)
            if (w != k && (!busy_only || running[w] > 0)) t += claim(w);
        return t;
    }
    // bytes held or claimed on the device by all workers (under mu)
    int64_t claimed() const { return others(resident.size(), false); }

    enum Step { kWait, kTake, kFreeOwn, kExit };
    Step next_step(size_t k) {
        if (queue.empty()) return closing ? kExit : (trim && resident[k] > 0 ? kFreeOwn : kWait);
        const int64_t f = queue.front()->footprint;
        if (others(k, false) + std::max(resident[k], f) <= budget) return kTake;
        if (resident[k] > 0 && others(k, false) + f <= budget) return kFreeOwn;
        if (others(k, true) + f <= budget) {
            if (!trim) {
                trim = true;  // idle workers' workspaces are in the way: wake them
                cv_work.notify_all();
            }
            if (resident[k] > 0) return kFreeOwn;
        }
        return kWait;
    }

    void free_workspace(size_t k, std::unique_lock<std::mutex>& lk) {
        lk.unlock();
        ctxs[k]->release_workspace();  // the worker's stream is idle
        lk.lock();
        resident[k] = 0;
        cv_work.notify_all();
    }

    void run(size_t k) {
        (void)hipSetDevice(device);
        for (;;) {
            std::shared_ptr<Job> j;
            {
                std::unique_lock<std::mutex> lk(mu);
                Step st;
                while ((st = next_step(k)) == kWait) cv_work.wait(lk);
                if (st == kExit) return;  // closing and drained
                if (st == kFreeOwn) {
                    free_workspace(k, lk);
                    continue;
                }
                j = queue.front();
                queue.pop_front();
                trim = false;
                running[k] = j->footprint;
                j->t_start = Clock::now();
            }
            cv_work.notify_all();  // the next head may fit as well
            const int32_t e = fof_entry(ctxs[k], j->a, j->n_threads);
            const int64_t held = (int64_t)ctxs[k]->resident_bytes();
            {
                std::unique_lock<std::mutex> lk(mu);
                j->ierr = e;
                j->msg = e ? last_error_message() : std::string();
                j->t_end = Clock::now();
                j->done = true;
                running[k] = 0;
                resident[k] = held;
                // a workspace larger than the budget leaves room for
                if (claimed() > budget) free_workspace(k, lk);
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

}  // namespace pyfof

using namespace pyfof;

struct pyfof_server {
    Server* s;
};

namespace {

thread_local std::string g_server_error;

void copy_msg(const std::string& m, char* buf, int32_t cap) {
    if (!buf || cap <= 0) return;
    const size_t n = std::min<size_t>(m.size(), (size_t)cap - 1);
    memcpy(buf, m.data(), n);
    buf[n] = '\0';
}

CallArgs args_of(const pyfof_job_desc* d) {
    CallArgs a{};
    a.xp = d->xp;
    a.vp = d->vp;
    a.mp = d->mp;
    a.np_halo = d->np_halo;
    a.nhalo = d->nhalo;
    a.L = d->L_Box;
    a.b = d->linking_length;
    a.np_min = d->np_halo_min;
    a.bound_check = d->bound_check;
    a.verbose = d->verbose;
    a.softening = d->softening;
    a.a_val = d->a_val;
    a.b_val = d->b_val;
    a.s_15 = d->s_15_val;
    a.ng = d->ng_val;
    a.ng_global = d->ng_global_val;
    a.total_np = d->total_np;
    a.in_halo = d->in_halo;
    a.halo_id = d->halo_id;
    a.ang_mom = d->ang_mom;
    a.root_out = nullptr;
    a.mem_kind = d->mem_kind;
    return a;
}

// 1 done (record released), 0 pending, -1 unknown id
int32_t finish(Server* s, std::unique_lock<std::mutex>& lk, int64_t id, bool block, int32_t* ierr,
               char* msg, int32_t cap, double* times) {
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return -1;
    std::shared_ptr<Job> j = it->second;
    if (!j->done) {
        if (!block) return 0;
        s->cv_done.wait(lk, [&] { return j->done; });
    }
    if (ierr) *ierr = j->ierr;
    copy_msg(j->msg, msg, cap);
    if (times) {
        times[0] = std::chrono::duration<double, std::milli>(j->t_start - j->t_submit).count();
        times[1] = std::chrono::duration<double, std::milli>(j->t_end - j->t_start).count();
    }
    s->jobs.erase(id);
    return 1;
}

}  // namespace

extern "C" {

pyfof_server* pyfof_server_create(int32_t device, int64_t hbm_budget_bytes, int32_t slots) {
    if (slots < 1 || slots > 16 || hbm_budget_bytes <= 0) {
        g_server_error = "slots must be in [1, 16] and the HBM budget positive";
        return nullptr;
    }
    try {
        auto* h = new pyfof_server;
        h->s = new Server(device, hbm_budget_bytes, slots);
        return h;
    } catch (const std::exception& e) {
        g_server_error = e.what();
        return nullptr;
    }
}

void pyfof_server_destroy(pyfof_server* server) {
    if (!server) return;
    delete server->s;
    delete server;
}

int64_t pyfof_job_footprint(const pyfof_job_desc* job) {
    if (!job) return -1;
    return job_footprint(args_of(job));
}

int64_t pyfof_submit(pyfof_server* server, const pyfof_job_desc* job) {
    if (!server || !job || !job->np_halo || job->nhalo < 0) {
        g_server_error = "null server/job/np_halo or negative nhalo";
        return -1;
    }
    Server* s = server->s;
    auto j = std::make_shared<Job>();
    j->np_halo.assign(job->np_halo, job->np_halo + job->nhalo);
    j->a = args_of(job);
    j->a.np_halo = j->np_halo.data();
    j->n_threads = job->n_threads;
    j->footprint = job_footprint(j->a);
    if (j->footprint > s->budget) {
        g_server_error = "job footprint exceeds the server's HBM budget";
        return -1;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->closing) {
            g_server_error = "server is closing";
            return -1;
        }
        j->id = s->next_id++;
        j->t_submit = Clock::now();
        s->jobs[j->id] = j;
        s->queue.push_back(j);
    }
    s->cv_work.notify_all();
    return j->id;
}

int32_t pyfof_poll(pyfof_server* server, int64_t id, int32_t* ierr, char* msg, int32_t msg_cap,
                   double* times) {
    if (!server) return -1;
    std::unique_lock<std::mutex> lk(server->s->mu);
    return finish(server->s, lk, id, false, ierr, msg, msg_cap, times);
}

int32_t pyfof_wait(pyfof_server* server, int64_t id, int32_t* ierr, char* msg, int32_t msg_cap,
                   double* times) {
    if (!server) return -1;
    std::unique_lock<std::mutex> lk(server->s->mu);
    return finish(server->s, lk, id, true, ierr, msg, msg_cap, times);
}

int32_t pyfof_server_status(pyfof_server* server, int64_t* out, int32_t n) {
    if (!server) return 0;
    Server* s = server->s;
    std::lock_guard<std::mutex> lk(s->mu);
    int64_t running = 0, inflight = 0, resident = 0;
    for (size_t w = 0; w < s->running.size(); ++w) {
        running += s->running[w] > 0;
        inflight += s->running[w];
        resident += s->resident[w];
    }
    const int64_t v[6] = {(int64_t)s->queue.size(), running, inflight, s->budget, resident,
                          s->claimed()};
    int32_t w = 0;
    for (; w < n && w < 6; ++w) out[w] = v[w];
    return w;
}

const char* pyfof_server_last_error(void) { return g_server_error.c_str(); }

}  // extern "C"
