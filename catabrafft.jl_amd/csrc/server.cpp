// server.cpp -- serving loop with continuous batching of FFT requests of mixed sizes.
//
// Requests (n, batch, dtype, direction, buffers) are queued by cfft_server_submit.  A worker
// thread repeatedly
//   1. retires finished rounds (HIP event per round): results are final, device bytes return;
//   2. admits queued requests in arrival order while the resident device bytes (request
//      buffers, convolution work, resident plans) stay under the HBM cap;
//   3. turns the admitted requests into one round: requests of the same (n, dtype) share one
//      resident plan (twiddles, Rader permutations, kernel spectra stay on the device) and their
//      rows are packed together into the same work-list launches, transforms of other lengths
//      go into the same launches too, and the round is enqueued on one of `nstreams` streams.
// A request that arrives while earlier rounds run is admitted into the next round as soon as
// memory and a stream slot free up, so its kernels overlap the running ones on the device.
// Each transform's arithmetic depends only on its own plan, so a result is the same alone and in
// any mix.  In the steady state the loop allocates and frees no device memory (hipFree
// synchronizes the device and would serialize the rounds): each stream slot keeps its work-list
// tables, pinned staging and completion event, and request buffers come from a pool.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "catabrafft.h"
#include "library.h"
#include "worklist.h"

namespace cfft {

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Device buffers recycled across requests, in power-of-two size classes of at least 64 KiB.
struct DevPool {
    std::multimap<size_t, void*> idle;
    size_t idle_bytes = 0;
    DevPool() = default;
    DevPool(const DevPool&) = delete;
    DevPool& operator=(const DevPool&) = delete;
    ~DevPool() { trim(); }
    static size_t size_class(size_t n) {
        size_t c = (size_t)64 << 10;
        while (c < n) c <<= 1;
        return c;
    }
    void* get(size_t n, size_t* cls) {
        const size_t c = size_class(n);
        *cls = c;
        auto it = idle.find(c);
        if (it != idle.end()) {
            void* p = it->second;
            idle.erase(it);
            idle_bytes -= c;
            return p;
        }
        void* p = nullptr;
        if (hipMalloc(&p, c) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        return p;
    }
    void put(void* p, size_t c) {
        if (!p) return;
        idle.emplace(c, p);
        idle_bytes += c;
    }
    void trim() {
        for (auto& kv : idle) (void)hipFree(kv.second);
        idle.clear();
        idle_bytes = 0;
    }
};

struct PoolBuf {
    void* p = nullptr;
    size_t cls = 0;
};

struct SrvRequest {
    int64_t id = 0;
    int64_t n = 0, batch = 0;
    int dtype = 0, dir = -1, norm = 0;
    const void* in = nullptr;
    void* out = nullptr;
    bool on_device = false;
    double t_submit = 0, t_done = 0;
    size_t bytes = 0;
    int status = 1;  // 1 queued / running, 0 done, < 0 error
    // an on-device request with in == out whose plan runs out of place (four-step): the result
    // goes to the pooled dout and is copied into `out` on the device
    bool via_dout = false;
    PoolBuf din, dout, work;
    std::string err;
};

struct SrvRound {
    std::vector<SrvRequest*> reqs;
    WorkList wl;
    std::vector<MultiLaunch> launches;
    int slot = 0;
};

// A stream with the resources its rounds reuse (one round at a time per slot).
struct SrvSlot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    DevBuf dplans, ditems, dtable;
    HostStage stage;
};

struct Server {
    Ctx* ctx = nullptr;
    size_t cap = 0;
    std::vector<SrvSlot> slots;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<SrvRequest*> queue;
    std::map<int64_t, std::unique_ptr<SrvRequest>> reqs;
    std::map<std::pair<int64_t, int>, std::unique_ptr<Plan>> plans;
    std::deque<std::unique_ptr<SrvRound>> inflight;
    DevPool pool;
    size_t resident = 0;  // plans + admitted requests (the pool's idle buffers are counted apart)
    int64_t next_id = 1;
    bool stop = false;
    bool trace = false;  // CFFT_SERVER_TRACE: one line per round on stderr
    // statistics
    std::vector<double> lat;
    double first_submit = -1, last_done = 0, flops = 0;
    int64_t transforms = 0;
};

static size_t esz_of(int dtype) { return dtype == CFFT_C32 ? 8 : 16; }

// Returns the resident plan for (n, dtype), creating it (and counting its bytes) if needed.
static int srv_plan(Server* s, int64_t n, int dtype, Plan** out, size_t* new_bytes) {
    *new_bytes = 0;
    auto key = std::make_pair(n, dtype);
    auto it = s->plans.find(key);
    if (it != s->plans.end()) {
        *out = it->second.get();
        return CFFT_OK;
    }
    std::unique_ptr<Plan> p(new Plan());
    p->ctx = s->ctx;
    p->n = n;
    p->batch = 256;  // nominal: lets the planner pick full workgroups of rows
    p->idist = p->odist = n;
    p->dtype = dtype;
    p->prec = dtype == CFFT_C32 ? Prec::F32 : Prec::F64;
    int rc;
    if (n == 1) {
        if ((rc = make_line(p->line, 1, p->prec, 16, 1024))) return rc;
        p->kind = Kind::Direct;
        p->C = 256;
    } else {
        PlanOptions o;
        o.alloc_work = false;  // a large convolution's work buffers come with each request
        if ((rc = plan_build(s->ctx, *p, o))) return rc;
        if (p->kind == Kind::Direct || p->kind == Kind::ConvSmall) {
            // these run in work lists: lines planned for the work-list kernels' capacity, so
            // that a workgroup has the threads those kernels use
            p.reset(new Plan());
            p->ctx = s->ctx;
            p->n = n;
            p->batch = 256;
            p->idist = p->odist = n;
            p->dtype = dtype;
            p->prec = dtype == CFFT_C32 ? Prec::F32 : Prec::F64;
            o.E = kMultiEHost;
            if ((rc = plan_build(s->ctx, *p, o))) return rc;
        }
        // six-step plans (or convolutions over them) own one scratch: rounds on several streams
        // would race
        if (p->kind == Kind::Large || (p->sub && p->sub->kind == Kind::Large))
            return fail(CFFT_ERR_UNSUPPORTED, "the server takes lengths the four-step covers");
    }
    *new_bytes = plan_device_bytes(*p);
    *out = p.get();
    s->plans[key] = std::move(p);
    return CFFT_OK;
}

static size_t request_bytes(const SrvRequest& r, const Plan& p) {
    const size_t es = esz_of(r.dtype);
    size_t b = 2 * (size_t)r.batch * r.n * es;  // input + output resident while in flight
    if (p.kind == Kind::ConvLarge) {
        const size_t ces = p.cprec == Prec::F32 ? 8 : 16;
        b += (size_t)r.batch * (2 * (size_t)p.L + 2) * ces;
    }
    return b;
}

// Returns a request's pooled buffers and its share of the HBM cap.  Called with the lock held.
static void srv_release(Server* s, SrvRequest* r) {
    s->pool.put(r->din.p, r->din.cls);
    s->pool.put(r->dout.p, r->dout.cls);
    s->pool.put(r->work.p, r->work.cls);
    r->din = r->dout = r->work = PoolBuf{};
    s->resident -= r->bytes;
    r->bytes = 0;
}

// Builds and enqueues one round on slot `slot`.  Called with the lock held; the lock is released
// while the work list is built and launched (that touches only the round, the slot's buffers and
// the immutable plans, all owned by the server thread).
static int srv_launch_round(Server* s, std::vector<SrvRequest*>& admitted, int slot,
                            std::unique_lock<std::mutex>& lk) {
    std::unique_ptr<SrvRound> R(new SrvRound());
    R->slot = slot;
    hipStream_t st = s->slots[slot].stream;
    int rc = CFFT_OK;
    // group transforms by plan; large transforms (four-step, large convolution) run as their
    // plan's own batched launches instead (their work needs no per-workgroup items)
    std::map<const Plan*, std::vector<XformRef>> groups;
    std::map<const Plan*, PlanEntry> entries;
    struct Direct {
        const Plan* p;
        SrvRequest* r;
        const void* in;
        void* out;
        ExecShape sh;
    };
    std::vector<Direct> direct;
    for (SrvRequest* r : admitted) {
        Plan* p = s->plans[std::make_pair(r->n, r->dtype)].get();
        const size_t es = esz_of(r->dtype);
        const size_t bytes = (size_t)r->batch * r->n * es;
        const void* din = r->in;
        void* dout = r->out;
        if (r->on_device && r->in == r->out && p->kind == Kind::FourStep) {
            r->dout.p = s->pool.get(bytes, &r->dout.cls);
            if (!r->dout.p) {
                r->status = CFFT_ERR_NOMEM;
                r->err = "device allocation failed";
                continue;
            }
            r->via_dout = true;
            dout = r->dout.p;
        }
        if (!r->on_device) {
            r->din.p = s->pool.get(bytes, &r->din.cls);
            r->dout.p = s->pool.get(bytes, &r->dout.cls);
            if (!r->din.p || !r->dout.p) {
                r->status = CFFT_ERR_NOMEM;
                r->err = "device allocation failed";
                continue;
            }
            if (hipMemcpyAsync(r->din.p, r->in, bytes, hipMemcpyHostToDevice, st) != hipSuccess) {
                r->status = CFFT_ERR_HIP;
                r->err = "host to device copy failed";
                continue;
            }
            din = r->din.p;
            dout = r->dout.p;
        }
        size_t ces = 0;
        if (p->kind == Kind::ConvLarge) {
            ces = p->cprec == Prec::F32 ? 8 : 16;
            r->work.p = s->pool.get((size_t)r->batch * (2 * (size_t)p->L + 2) * ces, &r->work.cls);
            if (!r->work.p) {
                r->status = CFFT_ERR_NOMEM;
                r->err = "device allocation failed";
                continue;
            }
        }
        if (p->kind == Kind::FourStep || p->kind == Kind::ConvLarge) {
            Direct d{p, r, din, dout, ExecShape{}};
            d.sh.batch = r->batch;
            if (p->kind == Kind::ConvLarge) {
                const size_t wl_bytes = (size_t)r->batch * p->L * ces;
                d.sh.w1 = r->work.p;
                d.sh.w2 = (char*)r->work.p + wl_bytes;
                d.sh.dc = (char*)r->work.p + 2 * wl_bytes;
            }
            direct.push_back(d);
            R->reqs.push_back(r);
            continue;
        }
        if (!entries.count(p)) entries[p] = R->wl.add_plan(*p);
        for (int64_t b = 0; b < r->batch; ++b) {
            XformRef x;
            x.in = (uint64_t)din + (uint64_t)b * r->n * es;
            x.out = (uint64_t)dout + (uint64_t)b * r->n * es;
            x.launch_dir = false;
            x.sgn = r->dir > 0 ? -1 : 1;
            x.scale = r->norm ? 1.0 / (double)r->n : 1.0;
            if (p->kind == Kind::ConvLarge) {
                const uint64_t w = (uint64_t)r->work.p + (uint64_t)b * (2 * p->L + 2) * ces;
                x.w1 = w;
                x.w2 = w + (uint64_t)p->L * ces;
                x.dc = w + 2 * (uint64_t)p->L * ces;
            }
            groups[p].push_back(x);
        }
        R->reqs.push_back(r);
    }
    // Requests that could not be staged leave now.  Copies of theirs may already be queued on
    // the slot's stream: they finish before the buffers return to the pool.
    bool staged_fail = false;
    for (SrvRequest* r : admitted) staged_fail = staged_fail || r->status < 0;
    if (staged_fail) {
        (void)hipStreamSynchronize(st);
        for (SrvRequest* r : admitted)
            if (r->status < 0) srv_release(s, r);
        s->cv_done.notify_all();
    }
    SrvSlot& sl = s->slots[slot];
    sl.busy = true;  // claimed while the round is built without the lock
    lk.unlock();
    rc = CFFT_OK;
    for (auto& kv : groups) R->wl.add_transforms(entries[kv.first], kv.second);
    rc = R->wl.finalize(sl.dplans, sl.ditems, sl.dtable, R->launches, st, true, &sl.stage);
    MultiArgs ma{0, 0, 1, 0};
    for (size_t i = 0; rc == CFFT_OK && i < R->launches.size(); ++i)
        if (launch_multi(R->launches[i], ma, st) != hipSuccess)
            rc = fail(CFFT_ERR_HIP, "work-list launch failed");
    // A plan-launched request that fails is recorded on that request alone; the round goes on.
    for (size_t i = 0; rc == CFFT_OK && i < direct.size(); ++i) {
        const Direct& d = direct[i];
        const int drc = plan_exec_shaped(*d.p, d.in, d.out, d.r->dir, d.r->norm, st, d.sh);
        if (drc != CFFT_OK) {
            d.r->status = drc;
            d.r->err = cfft_last_error();
        }
    }
    for (SrvRequest* r : R->reqs) {
        if (rc != CFFT_OK || r->status <= 0) continue;
        const size_t bytes = (size_t)r->batch * r->n * esz_of(r->dtype);
        if (!r->on_device)
            (void)hipMemcpyAsync(r->out, r->dout.p, bytes, hipMemcpyDeviceToHost, st);
        else if (r->via_dout && hipMemcpyAsync(r->out, r->dout.p, bytes, hipMemcpyDeviceToDevice,
                                               st) != hipSuccess) {
            r->status = CFFT_ERR_HIP;
            r->err = "device to device copy failed";
        }
    }
    if (rc == CFFT_OK && hipEventRecord(sl.done, st) != hipSuccess)
        rc = fail(CFFT_ERR_HIP, "event record failed");
    lk.lock();
    if (rc != CFFT_OK) {
        // nothing of this round may still run when its buffers return to the pool
        (void)hipStreamSynchronize(st);
        sl.busy = false;
        return rc;
    }
    s->inflight.push_back(std::move(R));
    return CFFT_OK;
}

static void srv_retire(Server* s) {
    for (auto it = s->inflight.begin(); it != s->inflight.end();) {
        SrvRound* R = it->get();
        if (hipEventQuery(s->slots[R->slot].done) != hipSuccess) {
            ++it;
            continue;
        }
        const double t = now_s();
        for (SrvRequest* r : R->reqs) {
            srv_release(s, r);
            r->t_done = t;
            if (r->status < 0) continue;  // failed on its own: no latency or throughput sample
            r->status = 0;
            s->lat.push_back(t - r->t_submit);
            s->last_done = t;
            s->transforms += r->batch;
            s->flops += 5.0 * r->n * std::log2((double)std::max<int64_t>(r->n, 2)) * r->batch;
        }
        s->slots[R->slot].busy = false;
        it = s->inflight.erase(it);
        s->cv_done.notify_all();
    }
}

static void srv_loop(Server* s) {
    (void)hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        srv_retire(s);
        if (s->stop && s->queue.empty() && s->inflight.empty()) break;
        // a free stream slot?
        int slot = -1;
        for (size_t k = 0; k < s->slots.size(); ++k)
            if (!s->slots[k].busy) {
                slot = (int)k;
                break;
            }
        std::vector<SrvRequest*> admitted;
        if (slot >= 0) {
            size_t round_bytes = 0;
            while (!s->queue.empty() && admitted.size() < 4096) {
                SrvRequest* r = s->queue.front();
                Plan* p = nullptr;
                size_t pb = 0;
                int rc = srv_plan(s, r->n, r->dtype, &p, &pb);
                s->resident += pb;
                if (rc) {
                    r->status = rc;
                    r->err = "cannot plan this length";
                    s->queue.pop_front();
                    s->cv_done.notify_all();
                    continue;
                }
                const size_t need = request_bytes(*r, *p);
                if (need > s->cap) {
                    r->status = CFFT_ERR_ADMISSION;
                    r->err = "request larger than the HBM cap";
                    s->queue.pop_front();
                    s->cv_done.notify_all();
                    continue;
                }
                if (s->resident + s->pool.idle_bytes + need > s->cap && s->pool.idle_bytes)
                    s->pool.trim();  // give idle pooled buffers back first
                if (s->resident + s->pool.idle_bytes + need > s->cap) break;  // wait for memory
                r->bytes = need;
                s->resident += need;
                round_bytes += need;
                admitted.push_back(r);
                s->queue.pop_front();
                if (round_bytes > ((size_t)8 << 30)) break;  // keep rounds short for latency
            }
        }
        if (!admitted.empty()) {
            const double t0 = now_s();
            const int rc = srv_launch_round(s, admitted, slot, lk);
            if (s->trace) {
                int64_t nt = 0;
                for (SrvRequest* r : admitted) nt += r->batch;
                const SrvRound* R = s->inflight.empty() ? nullptr : s->inflight.back().get();
                std::fprintf(stderr, "round slot=%d requests=%zu transforms=%lld launches=%zu "
                             "build_ms=%.3f queued=%zu inflight=%zu\n", slot, admitted.size(),
                             (long long)nt, R ? R->launches.size() : (size_t)0,
                             (now_s() - t0) * 1e3, s->queue.size(), s->inflight.size());
            }
            if (rc) {
                for (SrvRequest* r : admitted) {
                    if (r->status > 0) r->status = rc;
                    r->err = cfft_last_error();
                    srv_release(s, r);
                }
                s->cv_done.notify_all();
            }
            continue;
        }
        if (s->inflight.empty())
            s->cv_work.wait(lk, [s] { return s->stop || !s->queue.empty(); });
        else
            s->cv_work.wait_for(lk, std::chrono::microseconds(20));
    }
}

}  // namespace cfft

using namespace cfft;

struct cfft_server_s {
    Server* s;
};

extern "C" {

int cfft_server_create(cfft_ctx ctx, int64_t hbm_cap_bytes, int nstreams, cfft_server* srv) {
    if (!ctx || !srv || hbm_cap_bytes <= 0) return fail(CFFT_ERR_INVALID, "bad argument");
    if (nstreams < 1) nstreams = 1;
    if (nstreams > 8) nstreams = 8;
    if (hipSetDevice(ctx->c->device) != hipSuccess) return fail(CFFT_ERR_HIP, "hipSetDevice failed");
    Server* s = new Server();
    s->ctx = ctx->c;
    s->cap = (size_t)hbm_cap_bytes;
    s->trace = std::getenv("CFFT_SERVER_TRACE") != nullptr;
    s->slots = std::vector<SrvSlot>(nstreams);
    for (auto& sl : s->slots)
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
            for (auto& x : s->slots) {
                if (x.stream) (void)hipStreamDestroy(x.stream);
                if (x.done) (void)hipEventDestroy(x.done);
            }
            delete s;
            return fail(CFFT_ERR_HIP, "stream creation failed");
        }
    s->worker = std::thread(srv_loop, s);
    *srv = new cfft_server_s{s};
    return CFFT_OK;
}

int cfft_server_submit(cfft_server srv, int64_t n, int64_t batch, int dtype, int direction,
                       int normalize, const void* in, void* out, int on_device, int64_t* ticket) {
    if (!srv || !ticket || !in || !out) return fail(CFFT_ERR_INVALID, "null argument");
    if (n < 1 || batch < 1) return fail(CFFT_ERR_INVALID, "n and batch must be >= 1");
    if (dtype != CFFT_C32 && dtype != CFFT_C64) return fail(CFFT_ERR_INVALID, "bad dtype");
    if (direction != CFFT_FORWARD && direction != CFFT_BACKWARD)
        return fail(CFFT_ERR_INVALID, "direction must be -1 or +1");
    Server* s = srv->s;
    std::unique_ptr<SrvRequest> r(new SrvRequest());
    r->n = n;
    r->batch = batch;
    r->dtype = dtype;
    r->dir = direction;
    r->norm = normalize ? 1 : 0;
    r->in = in;
    r->out = out;
    r->on_device = on_device != 0;
    r->t_submit = now_s();
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) return fail(CFFT_ERR_INVALID, "server is shutting down");
        r->id = s->next_id++;
        if (s->first_submit < 0) s->first_submit = r->t_submit;
        *ticket = r->id;
        s->queue.push_back(r.get());
        s->reqs[r->id] = std::move(r);
    }
    s->cv_work.notify_one();
    return CFFT_OK;
}

static int ticket_status(Server* s, int64_t ticket, bool wait) {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(CFFT_ERR_INVALID, "unknown ticket");
    SrvRequest* r = it->second.get();
    if (wait) s->cv_done.wait(lk, [r] { return r->status <= 0; });
    if (r->status > 0) return CFFT_ERR_NOTREADY;
    if (r->status < 0) return fail(r->status, r->err);
    return CFFT_OK;
}

int cfft_server_poll(cfft_server srv, int64_t ticket) {
    if (!srv) return fail(CFFT_ERR_INVALID, "null server");
    return ticket_status(srv->s, ticket, false);
}

int cfft_server_wait(cfft_server srv, int64_t ticket) {
    if (!srv) return fail(CFFT_ERR_INVALID, "null server");
    return ticket_status(srv->s, ticket, true);
}

int cfft_server_stats(cfft_server srv, double* p50, double* p99, double* transforms_per_s,
                      double* gflops) {
    if (!srv || !p50 || !p99 || !transforms_per_s || !gflops)
        return fail(CFFT_ERR_INVALID, "null argument");
    Server* s = srv->s;
    std::lock_guard<std::mutex> g(s->mu);
    std::vector<double> v = s->lat;
    *p50 = *p99 = *transforms_per_s = *gflops = 0;
    if (v.empty()) return CFFT_OK;
    std::sort(v.begin(), v.end());
    auto q = [&v](double f) {
        const double pos = f * (double)(v.size() - 1);
        const size_t lo = (size_t)std::floor(pos), hi = std::min(v.size() - 1, lo + 1);
        return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
    };
    *p50 = q(0.50);
    *p99 = q(0.99);
    const double span = s->last_done - s->first_submit;
    if (span > 0) {
        *transforms_per_s = (double)s->transforms / span;
        *gflops = s->flops / span / 1e9;
    }
    return CFFT_OK;
}

int cfft_server_reset_stats(cfft_server srv) {
    if (!srv) return fail(CFFT_ERR_INVALID, "null server");
    Server* s = srv->s;
    std::lock_guard<std::mutex> g(s->mu);
    s->lat.clear();
    s->first_submit = -1;
    s->last_done = 0;
    s->flops = 0;
    s->transforms = 0;
    return CFFT_OK;
}

int cfft_server_destroy(cfft_server srv) {
    if (!srv) return fail(CFFT_ERR_INVALID, "null server");
    Server* s = srv->s;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->ctx->device);
    for (auto& sl : s->slots) {
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
    }
    delete s;
    delete srv;
    return CFFT_OK;
}

}  // extern "C"
