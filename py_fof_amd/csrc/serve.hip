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

// Estimated device bytes of one job: the FoF workspace per particle
// (sort keys and values with their ping-pong copies 16 B, CRS float4 16 B,
// DSU and labels 20 B, outputs and catalog 24 B, cells about 8 B), the
// staging of host arrays, and for the unbinding the per-member state
// (about 120 B) plus one 256^3 grid pair (2 x 128 MiB) for the largest halo.
int64_t job_footprint(const CallArgs& a) {
    const int64_t n = a.total_np;
    int64_t b = n * 88 + (16ll << 20);
    if (a.mem_kind == PYFOF_MEM_HOST) b += n * (24 + (a.vp ? 24 : 0) + (a.mp ? 8 : 0) + 12);
    if (a.bound_check) b += n * 120 + (256ll << 20);
    return b;
}

struct Server {
    int device = 0;
    int64_t budget = 0;
    int64_t inflight = 0;
    int64_t next_id = 1;
    bool closing = false;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Job>> queue;
    std::unordered_map<int64_t, std::shared_ptr<Job>> jobs;
    std::vector<Ctx*> ctxs;
    std::vector<std::thread> workers;

    Server(int dev, int64_t hbm_budget, int slots) : device(dev), budget(hbm_budget) {
        PYFOF_HIP(hipSetDevice(dev));
        for (int k = 0; k < slots; ++k) {
            ctxs.push_back(new Ctx(dev));
            ctxs.back()->solo = slots == 1;
        }
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

    void run(int k) {
        (void)hipSetDevice(device);
        for (;;) {
            std::shared_ptr<Job> j;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] {
                    if (queue.empty()) return closing;
                    return inflight + queue.front()->footprint <= budget;
                });
                if (queue.empty()) return;  // closing and drained
                j = queue.front();
                queue.pop_front();
                inflight += j->footprint;
                j->t_start = Clock::now();
            }
            cv_work.notify_all();  // the next head may fit as well
            const int32_t e = fof_entry(ctxs[k], j->a, j->n_threads);
            {
                std::lock_guard<std::mutex> lk(mu);
                j->ierr = e;
                j->msg = e ? last_error_message() : std::string();
                j->t_end = Clock::now();
                j->done = true;
                inflight -= j->footprint;
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
    int64_t running = 0;
    for (auto& kv : s->jobs)
        if (!kv.second->done && kv.second->t_start != Clock::time_point()) ++running;
    const int64_t v[4] = {(int64_t)s->queue.size(), running, s->inflight, s->budget};
    int32_t w = 0;
    for (; w < n && w < 4; ++w) out[w] = v[w];
    return w;
}

const char* pyfof_server_last_error(void) { return g_server_error.c_str(); }

}  // extern "C"
