// Continuous-batching circuit server (ev_server_* in include/everett.h).
//
// submit() validates the circuit on the caller's thread, stamps the request's
// submit time (its latency runs from there, planning included) and hands a
// copy to a few planner threads, which plan it (fused passes, as for a single
// run) and queue it; a burst of submissions is planned in parallel and never
// holds up the caller.  The scheduler thread:
//   1. admits queued requests in arrival order while their statevectors fit
//      the HBM budget: allocates the state (from a size-keyed cache of device
//      blocks), uploads the plan on a side stream, resets the state;
//   2. runs steps: ONE launch (the batch kernel, a workgroup per tile) over a
//      run of tiles of the current pass of every admitted request.  A step
//      holds about kStepTiles tiles: requests with the least work left come
//      first and take their whole pass, larger ones take what is left of the
//      budget (at least kMinChunk tiles each, so none starves).  A step's
//      duration is thus bounded however large the requests beside it are,
//      and small requests never queue behind large ones.  (Measured on the
//      served stream: least-work-first gives a lower overall p99 than taking
//      the large requests in arrival order, 251 against 352 ms.)  Tiles of
//      one pass are independent, and the next pass of a request starts in a
//      later launch on the same stream;
//   3. behind the step that ran a request's last tiles it queues the
//      request's f64 renormalisation (complex64), its Born-rule sampling (f64
//      CDF, host-drawn uniforms from Rng::seed_from_u64(seed)) and the copy of
//      the samples, then an event; a request is done when its event is.
// Nothing drains the stream: admission, steps and finishing requests overlap
// on the device, and new requests join at the next step.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "fused_kernels.hpp"
#include "kernels.hpp"
#include "planner.hpp"

namespace ev {
namespace {

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Tiles per step (a few milliseconds of device time at most) and the least a
// request with work left takes of every step.
constexpr uint64_t kStepTiles = uint64_t(1) << 15;
constexpr uint64_t kMinChunk = uint64_t(1) << 11;

struct Request {
    uint64_t id = 0;
    int n = 0;
    uint64_t shots = 0, seed = 0, gates = 0;
    std::unique_ptr<Circuit> circuit;  // until planned
    FusedConfig cfg{};
    Plan plan;
    size_t state_bytes = 0;
    void *psi = nullptr;
    void *desc = nullptr;
    size_t desc_bytes = 0;
    std::vector<unsigned char> desc_host;  // kept until the request is done
    size_t pass_off = 0, stage_off = 0, op_off = 0, coef_off = 0;
    size_t next_pass = 0;
    uint64_t next_tile = 0;   // first tile of the current pass not yet launched
    uint64_t tiles_left = 0;  // tiles of all passes not yet launched
    // finishing: device scratch, pinned uniforms / samples, completion event
    void *scr = nullptr;
    size_t scr_bytes = 0;
    uint64_t *pinned = nullptr;
    size_t pinned_bytes = 0;
    hipEvent_t done_ev = nullptr;
    std::vector<uint64_t> samples;
    ev_request_info info{};
    bool done = false;
    uint64_t pass_tiles() const { return uint64_t(1) << (n - plan.passes[next_pass].t); }
};

// Device blocks cached by exact size: serving reuses the same few state sizes.
class BlockCache {
  public:
    void *get(size_t bytes)
    {
        auto it = free_.find(bytes);
        if (it != free_.end()) {
            void *p = it->second;
            free_.erase(it);
            return p;
        }
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {  // give the cached blocks back and retry once
            (void)hipGetLastError();
            release_all();
            EV_HIP(hipMalloc(&p, bytes));
        }
        return p;
    }
    void put(void *p, size_t bytes)
    {
        if (p) free_.emplace(bytes, p);
    }
    void release_all()
    {
        for (auto &kv : free_) (void)hipFree(kv.second);
        free_.clear();
    }
    ~BlockCache() { release_all(); }

  private:
    std::multimap<size_t, void *> free_;
};

// Pinned host blocks cached by exact size (uniforms in, samples out).
class PinnedCache {
  public:
    void *get(size_t bytes)
    {
        auto it = free_.find(bytes);
        if (it != free_.end()) {
            void *p = it->second;
            free_.erase(it);
            return p;
        }
        void *p = nullptr;
        EV_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault));
        return p;
    }
    void put(void *p, size_t bytes)
    {
        if (p) free_.emplace(bytes, p);
    }
    void release_all()
    {
        for (auto &kv : free_) (void)hipHostFree(kv.second);
        free_.clear();
    }
    ~PinnedCache() { release_all(); }

  private:
    std::multimap<size_t, void *> free_;
};

double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace
}  // namespace ev

struct ev_server {
    ev_server_config cfg{};
    bool c64 = false;
    size_t amp_bytes = 16;
    uint64_t budget = 0;
    hipStream_t stream = nullptr;     // steps, resets, finishing work
    hipStream_t up_stream = nullptr;  // plan uploads from pageable host memory
    hipEvent_t up_ev = nullptr;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_plan;
    std::deque<std::shared_ptr<ev::Request>> plan_queue;  // submitted, not yet planned
    size_t planning = 0;                                  // requests submitted and not yet queued or failed
    std::vector<std::thread> planners;
    std::deque<std::shared_ptr<ev::Request>> queue;
    std::vector<std::shared_ptr<ev::Request>> active;     // scheduler thread only: passes left to launch
    std::vector<std::shared_ptr<ev::Request>> finishing;  // scheduler thread only: sampling queued
    std::map<uint64_t, std::shared_ptr<ev::Request>> reqs;
    bool stop = false;
    uint64_t next_id = 1;
    ev_server_stats st{};
    ev::BlockCache cache;                              // scheduler thread only
    ev::PinnedCache pinned;                            // scheduler thread only
    // Two step slots (scheduler thread only): the host fills slot k & 1
    // while the device may still run the step of the other slot.
    ev::fused::BatchEntry *h_ent[2] = {nullptr, nullptr};  // pinned
    void *d_ent[2] = {nullptr, nullptr};
    size_t ent_cap[2] = {0, 0};
    hipEvent_t ev_a[2] = {nullptr, nullptr}, ev_b[2] = {nullptr, nullptr};
    uint64_t step_no = 0;
    bool inflight = false;                             // step step_no - 1 not yet waited for
    std::thread sched;
};

namespace ev {
namespace {

void fail(ev_server *s, const std::shared_ptr<Request> &r, int code)
{
    std::lock_guard<std::mutex> lk(s->mu);
    r->info.status = code;
    r->info.done_ms = ms_since(s->t0);
    r->done = true;
    ++s->st.failed;
    s->cv_done.notify_all();
}

// Returns a request's device and pinned blocks to the caches.  Only once the
// stream has passed every use of them (its done event, or a synchronisation).
void release(ev_server *s, Request &r)
{
    s->cache.put(r.psi, r.state_bytes);
    s->cache.put(r.desc, r.desc_bytes);
    s->cache.put(r.scr, r.scr_bytes);
    s->pinned.put(r.pinned, r.pinned_bytes);
    r.psi = r.desc = r.scr = nullptr;
    r.pinned = nullptr;
    if (r.done_ev) (void)hipEventDestroy(r.done_ev);
    r.done_ev = nullptr;
    r.desc_host.clear();
    r.desc_host.shrink_to_fit();
    std::lock_guard<std::mutex> lk(s->mu);
    s->st.bytes_in_use -= r.state_bytes;
}

// Uploads the plan (side stream; the main stream waits for it) and resets
// the state (scheduler thread).
void admit(ev_server *s, Request &r)
{
    const Plan &P = r.plan;
    size_t off = 0;
    r.pass_off = off;
    off += align256(P.passes.size() * sizeof(fused::Pass));
    r.stage_off = off;
    off += align256(P.stages.size() * sizeof(fused::Stage));
    r.op_off = off;
    off += align256(std::max<size_t>(1, P.devops.size()) * sizeof(fused::DevOp));
    r.coef_off = off;
    off += align256(std::max<size_t>(1, P.coef.size()) * s->amp_bytes);
    r.desc_bytes = off;
    std::vector<unsigned char> &host = r.desc_host;
    host.assign(off, 0);
    if (!P.passes.empty()) std::memcpy(host.data() + r.pass_off, P.passes.data(), P.passes.size() * sizeof(fused::Pass));
    if (!P.stages.empty())
        std::memcpy(host.data() + r.stage_off, P.stages.data(), P.stages.size() * sizeof(fused::Stage));
    if (!P.devops.empty()) std::memcpy(host.data() + r.op_off, P.devops.data(), P.devops.size() * sizeof(fused::DevOp));
    if (s->c64) fused::inline_coef_to_c64(reinterpret_cast<fused::DevOp *>(host.data() + r.op_off), P.devops.size());
    if (!s->c64) {
        if (!P.coef.empty()) std::memcpy(host.data() + r.coef_off, P.coef.data(), P.coef.size() * sizeof(c128));
    } else {
        float *f = reinterpret_cast<float *>(host.data() + r.coef_off);
        for (size_t k = 0; k < P.coef.size(); ++k) {
            f[2 * k] = float(P.coef[k].re);
            f[2 * k + 1] = float(P.coef[k].im);
        }
    }
    r.tiles_left = 0;
    for (const fused::Pass &ps : P.passes) r.tiles_left += uint64_t(1) << (r.n - ps.t);
    r.psi = s->cache.get(r.state_bytes);
    r.desc = s->cache.get(r.desc_bytes);
    EV_HIP(hipMemcpyAsync(r.desc, host.data(), off, hipMemcpyHostToDevice, s->up_stream));
    EV_HIP(hipEventRecord(s->up_ev, s->up_stream));
    EV_HIP(hipStreamWaitEvent(s->stream, s->up_ev, 0));
    kern::init_basis(r.psi, s->c64 ? kern::C64 : kern::C128, r.n, 0, s->stream);
}

// Queues a request's end on the stream, behind the launch that ran its last
// tiles: the f64 renormalisation every complex64 run ends with (the same
// device reduction and scaling as a run on its own, so a request samples
// exactly as it would alone), the Born-rule shots (uniforms from the
// request's own stream, f64 CDF on the device) and the copy of the samples
// to pinned memory, then the request's done event.
void enqueue_finish(ev_server *s, Request &r)
{
    const size_t nscr = std::max<size_t>(kern::sample_scratch_doubles(r.n), 2 * kern::kReduceBlocks + 1);
    const size_t b_scr = align256(nscr * sizeof(double)), b_u = align256(std::max<uint64_t>(1, r.shots) * 8);
    r.scr_bytes = b_scr + 2 * b_u;
    r.scr = s->cache.get(r.scr_bytes);
    r.pinned_bytes = size_t(std::max<uint64_t>(1, r.shots)) * 8;
    r.pinned = static_cast<uint64_t *>(s->pinned.get(r.pinned_bytes));
    Rng rng = Rng::seed_from_u64(r.seed);
    double *u = reinterpret_cast<double *>(r.pinned);
    for (uint64_t k = 0; k < r.shots; ++k) u[k] = rng.next_f64();
    unsigned char *scr = static_cast<unsigned char *>(r.scr);
    const kern::Prec prec = s->c64 ? kern::C64 : kern::C128;
    if (s->c64) {
        double *partials = reinterpret_cast<double *>(scr);
        kern::normalize(r.psi, prec, r.n, partials, partials + 2 * kern::kReduceBlocks, s->stream);
    }
    if (r.shots) EV_HIP(hipMemcpyAsync(scr + b_scr, r.pinned, r.shots * 8, hipMemcpyHostToDevice, s->stream));
    kern::sample(r.psi, prec, r.n, reinterpret_cast<double *>(scr + b_scr), r.shots, reinterpret_cast<double *>(scr),
                 reinterpret_cast<uint64_t *>(scr + b_scr + b_u), s->stream);
    if (r.shots)
        EV_HIP(hipMemcpyAsync(r.pinned, scr + b_scr + b_u, r.shots * 8, hipMemcpyDeviceToHost, s->stream));
    EV_HIP(hipEventCreateWithFlags(&r.done_ev, hipEventDisableTiming));
    EV_HIP(hipEventRecord(r.done_ev, s->stream));
}

// A finishing request whose done event has passed: hands over its samples.
void complete(ev_server *s, const std::shared_ptr<Request> &r)
{
    r->samples.assign(r->pinned, r->pinned + r->shots);
    release(s, *r);
    std::lock_guard<std::mutex> lk(s->mu);
    r->info.status = EV_OK;
    r->info.done_ms = ms_since(s->t0);
    r->done = true;
    ++s->st.completed;
    s->cv_done.notify_all();
}

// Completes the finishing requests whose events have passed; with `block`,
// waits for the oldest one first.
void reap(ev_server *s, bool block)
{
    if (s->finishing.empty()) return;
    if (block) EV_HIP(hipEventSynchronize(s->finishing.front()->done_ev));
    std::vector<std::shared_ptr<Request>> still;
    for (auto &r : s->finishing) {
        const hipError_t q = hipEventQuery(r->done_ev);
        if (q == hipSuccess) {
            complete(s, r);
        } else if (q == hipErrorNotReady) {
            still.push_back(r);
        } else {
            (void)hipGetLastError();
            EV_HIP(hipStreamSynchronize(s->stream));  // nothing in flight uses its blocks
            release(s, *r);
            fail(s, r, EV_E_HIP);
        }
    }
    s->finishing.swap(still);
}

// Waits for the step in flight, if any, and adds its device time.
void drain(ev_server *s)
{
    if (!s->inflight) return;
    const int slot = int((s->step_no - 1) & 1);
    s->inflight = false;
    EV_HIP(hipEventSynchronize(s->ev_b[slot]));
    float ms = 0;
    EV_HIP(hipEventElapsedTime(&ms, s->ev_a[slot], s->ev_b[slot]));
    std::lock_guard<std::mutex> lk(s->mu);
    s->st.busy_ms += ms;
    ++s->st.steps;
}

// One step: ONE launch over a run of tiles of the current pass of every
// active request (see the top of the file for how the tiles are shared).
// Steps are pipelined two deep: this one is queued before the previous one
// is waited for, so the host's preparation of a step overlaps the device's
// run of the one before.  Returns the requests whose last tiles it ran.
std::vector<std::shared_ptr<Request>> step(ev_server *s)
{
    const int slot = int(s->step_no & 1);
    const size_t na = s->active.size();
    if (!s->ev_a[slot]) {
        EV_HIP(hipEventCreate(&s->ev_a[slot]));
        EV_HIP(hipEventCreate(&s->ev_b[slot]));
    }
    if (na > s->ent_cap[slot]) {  // this slot's last step (two ago) has completed
        if (s->h_ent[slot]) EV_HIP(hipHostFree(s->h_ent[slot]));
        if (s->d_ent[slot]) EV_HIP(hipFree(s->d_ent[slot]));
        s->h_ent[slot] = nullptr;
        s->d_ent[slot] = nullptr;
        const size_t cap = std::max<size_t>(64, 2 * na);
        EV_HIP(hipHostMalloc(reinterpret_cast<void **>(&s->h_ent[slot]), cap * sizeof(fused::BatchEntry),
                             hipHostMallocDefault));
        EV_HIP(hipMalloc(&s->d_ent[slot], cap * sizeof(fused::BatchEntry)));
        s->ent_cap[slot] = cap;
    }
    // least work left first (ties: admission order)
    std::vector<size_t> order(na);
    for (size_t i = 0; i < na; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return s->active[a]->tiles_left < s->active[b]->tiles_left; });
    uint64_t tiles = 0;
    int max_t = fused::kMinTile;
    std::vector<uint64_t> took(na, 0);
    for (size_t k = 0; k < na; ++k) {
        const size_t i = order[k];
        Request &r = *s->active[i];
        const uint64_t rem = r.pass_tiles() - r.next_tile;
        const uint64_t room = tiles < kStepTiles ? kStepTiles - tiles : 0;
        const uint64_t take = std::min(rem, std::max(room, kMinChunk));
        unsigned char *d = static_cast<unsigned char *>(r.desc);
        fused::BatchEntry &e = s->h_ent[slot][k];
        e.psi = r.psi;
        e.pass = reinterpret_cast<const fused::Pass *>(d + r.pass_off) + r.next_pass;
        e.stages = reinterpret_cast<const fused::Stage *>(d + r.stage_off);
        e.ops = reinterpret_cast<const fused::DevOp *>(d + r.op_off);
        e.coef = d + r.coef_off;
        e.tile_begin = tiles;
        e.tile_off = r.next_tile;
        tiles += take;
        took[i] = take;
        max_t = std::max(max_t, int(r.plan.passes[r.next_pass].t));
    }
    EV_HIP(hipMemcpyAsync(s->d_ent[slot], s->h_ent[slot], na * sizeof(fused::BatchEntry), hipMemcpyHostToDevice,
                          s->stream));
    EV_HIP(hipEventRecord(s->ev_a[slot], s->stream));
    fused::launch_batch(s->c64, static_cast<const fused::BatchEntry *>(s->d_ent[slot]), int(na), tiles, max_t,
                        s->stream);
    EV_HIP(hipEventRecord(s->ev_b[slot], s->stream));
    std::vector<std::shared_ptr<Request>> ended;
    for (size_t i = 0; i < na; ++i) {
        Request &r = *s->active[i];
        r.next_tile += took[i];
        r.tiles_left -= took[i];
        if (r.next_tile == r.pass_tiles()) {
            r.next_tile = 0;
            ++r.next_pass;
        }
        if (r.next_pass >= r.plan.passes.size()) ended.push_back(s->active[i]);
    }
    drain(s);  // the previous step
    ++s->step_no;
    s->inflight = true;
    return ended;
}

// Planner thread: plans submitted requests and queues them for admission.
// On shutdown it plans what was already submitted, then returns.
void planner(ev_server *s)
{
    for (;;) {
        std::shared_ptr<Request> r;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_plan.wait(lk, [&] { return s->stop || !s->plan_queue.empty(); });
            if (s->plan_queue.empty()) return;
            r = s->plan_queue.front();
            s->plan_queue.pop_front();
        }
        int code = EV_OK;
        try {
            r->plan = plan_circuit(*r->circuit, r->n, r->cfg);
        } catch (const Error &e) {
            code = e.code;
        } catch (const std::bad_alloc &) {
            code = EV_E_OOM;
        }
        r->circuit.reset();
        if (code != EV_OK) fail(s, r, code);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (code == EV_OK) {
                r->info.passes = r->plan.passes.size();
                s->queue.push_back(r);
                s->st.queued = s->queue.size();
            }
            --s->planning;
        }
        s->cv_work.notify_all();
    }
}

void scheduler(ev_server *s)
{
    for (;;) {
        std::vector<std::shared_ptr<Request>> admitted;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] {
                return (s->stop && s->planning == 0) || !s->queue.empty() || !s->active.empty() ||
                       !s->finishing.empty();
            });
            if (s->stop && s->planning == 0 && s->queue.empty() && s->active.empty() && s->finishing.empty())
                return;
            // Admission in the order planning completes, while the states fit the budget.
            while (!s->queue.empty()) {
                auto &r = s->queue.front();
                if (s->cfg.max_active && s->active.size() + admitted.size() >= s->cfg.max_active) break;
                if (s->st.bytes_in_use + r->state_bytes > s->budget) break;
                s->st.bytes_in_use += r->state_bytes;
                s->st.bytes_peak = std::max(s->st.bytes_peak, s->st.bytes_in_use);
                r->info.admit_ms = ms_since(s->t0);
                admitted.push_back(r);
                s->queue.pop_front();
            }
            s->st.queued = s->queue.size();
        }
        for (auto &r : admitted) {
            try {
                admit(s, *r);
                if (r->plan.passes.empty()) {  // no gates: sample |0..0>
                    enqueue_finish(s, *r);
                    s->finishing.push_back(r);
                } else {
                    s->active.push_back(r);
                }
            } catch (const Error &e) {
                (void)hipStreamSynchronize(s->stream);
                release(s, *r);
                fail(s, r, e.code);
            }
        }
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->st.active = s->active.size();
        }
        try {
            if (!s->active.empty()) {
                const std::vector<std::shared_ptr<Request>> ended = step(s);
                if (!ended.empty()) {
                    std::vector<std::shared_ptr<Request>> still;
                    for (auto &r : s->active)
                        if (r->next_pass < r->plan.passes.size()) still.push_back(r);
                    s->active.swap(still);
                    for (auto &r : ended) {
                        enqueue_finish(s, *r);
                        s->finishing.push_back(r);
                    }
                }
                reap(s, false);
            } else {
                drain(s);
                reap(s, !s->finishing.empty());
            }
        } catch (const Error &e) {
            s->inflight = false;
            (void)hipStreamSynchronize(s->stream);
            for (auto *list : {&s->active, &s->finishing}) {
                for (auto &r : *list) {
                    release(s, *r);
                    fail(s, r, e.code);
                }
                list->clear();
            }
            continue;
        }
    }
}

}  // namespace

ev_server *server_create(const ev_server_config &cfg)
{
    auto *s = new ev_server;
    s->cfg = cfg;
    s->c64 = cfg.precision == EV_C64;
    s->amp_bytes = s->c64 ? 8 : 16;
    s->t0 = std::chrono::steady_clock::now();
    try {
        EV_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        EV_HIP(hipStreamCreateWithFlags(&s->up_stream, hipStreamNonBlocking));
        EV_HIP(hipEventCreateWithFlags(&s->up_ev, hipEventDisableTiming));
        size_t free_b = 0, total_b = 0;
        EV_HIP(hipMemGetInfo(&free_b, &total_b));
        s->budget = cfg.memory_bytes ? cfg.memory_bytes : uint64_t(0.8 * double(free_b));
    } catch (...) {
        if (s->up_ev) (void)hipEventDestroy(s->up_ev);
        if (s->up_stream) (void)hipStreamDestroy(s->up_stream);
        if (s->stream) (void)hipStreamDestroy(s->stream);
        delete s;
        throw;
    }
    s->sched = std::thread(scheduler, s);
    const unsigned np = std::max(1u, std::min(4u, std::thread::hardware_concurrency() / 2));
    for (unsigned k = 0; k < np; ++k) s->planners.emplace_back(planner, s);
    return s;
}

void server_destroy(ev_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_plan.notify_all();
    for (auto &t : s->planners)
        if (t.joinable()) t.join();
    s->cv_work.notify_all();
    if (s->sched.joinable()) s->sched.join();
    (void)hipStreamSynchronize(s->stream);
    for (int k = 0; k < 2; ++k) {
        if (s->h_ent[k]) (void)hipHostFree(s->h_ent[k]);
        if (s->d_ent[k]) (void)hipFree(s->d_ent[k]);
        if (s->ev_a[k]) (void)hipEventDestroy(s->ev_a[k]);
        if (s->ev_b[k]) (void)hipEventDestroy(s->ev_b[k]);
    }
    (void)hipStreamSynchronize(s->up_stream);
    s->cache.release_all();
    s->pinned.release_all();
    (void)hipEventDestroy(s->up_ev);
    (void)hipStreamDestroy(s->up_stream);
    (void)hipStreamDestroy(s->stream);
    delete s;
}

uint64_t server_submit(ev_server *s, const Circuit &c, uint64_t shots, uint64_t seed)
{
    const double t_submit = ms_since(s->t0);
    c.validate();
    const int n = int(c.num_qubits());
    EV_REQUIRE(n >= fused::kMinTile && n <= 30, EV_E_LIMIT, "served circuits need 8..30 qubits");
    for (const Op &op : c.ops())
        EV_REQUIRE(op.kind == OpKind::Apply1 || op.kind == OpKind::Apply2 || op.kind == OpKind::Controlled, EV_E_ARG,
                   "served circuits must be measurement-free (shots are drawn at the end)");
    auto r = std::make_shared<Request>();
    r->n = n;
    r->shots = shots;
    r->seed = seed;
    r->gates = c.size();
    r->state_bytes = (size_t(1) << n) * s->amp_bytes;
    if (r->state_bytes > s->budget)
        throw Error(EV_E_ADMISSION, "a " + std::to_string(n) + "-qubit state exceeds the server's memory budget");
    ev_run_options o{};
    o.precision = s->cfg.precision;
    o.tile_qubits = s->cfg.tile_qubits;
    o.max_gates_per_pass = s->cfg.max_gates_per_pass;
    r->cfg = fused_config(n, s->c64, o);  // a bad configuration fails here, on the caller's thread
    r->circuit = std::make_unique<Circuit>(c);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        EV_REQUIRE(!s->stop, EV_E_STATE, "server is shutting down");
        r->id = s->next_id++;
        r->info.id = r->id;
        r->info.num_qubits = uint64_t(n);
        r->info.gates = r->gates;
        r->info.shots = shots;
        r->info.submit_ms = t_submit;
        s->reqs[r->id] = r;
        s->plan_queue.push_back(r);
        ++s->planning;
        ++s->st.submitted;
    }
    s->cv_plan.notify_one();
    return r->id;
}

bool server_poll(ev_server *s, uint64_t id)
{
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    EV_REQUIRE(it != s->reqs.end(), EV_E_ARG, "unknown request id");
    return it->second->done;
}

ev_request_info server_wait(ev_server *s, uint64_t id, uint64_t *samples)
{
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    EV_REQUIRE(it != s->reqs.end(), EV_E_ARG, "unknown request id");
    auto r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
    if (samples && !r->samples.empty()) std::memcpy(samples, r->samples.data(), r->samples.size() * sizeof(uint64_t));
    s->reqs.erase(id);
    return r->info;
}

ev_server_stats server_stats(ev_server *s)
{
    std::lock_guard<std::mutex> lk(s->mu);
    return s->st;
}

}  // namespace ev
