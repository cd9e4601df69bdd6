// Batched execution of many independent statevector circuits.
//
// All states live in one HBM arena, packed back to back.  Circuits run in
// pipelined groups: each group is planned into fused passes (in parallel on
// host threads) while the device runs the previous group; round r of a
// group is ONE launch over every tile of the r-th pass of every circuit of
// the group that has one, so small and large states share the GPU.  Each
// group's rounds can run as one captured HIP graph.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "batch_kernels.hpp"
#include "fused_kernels.hpp"
#include "planner.hpp"

namespace ev {

namespace {

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Device layout of one circuit's plan inside the descriptor buffer.
struct PlanLoc {
    size_t pass_off, stage_off, op_off, coef_off;
    uint32_t npass;
};

}  // namespace

void batch_free(ev_batch *b)
{
    if (!b) return;
    if (b->stream) {
        (void)hipStreamSynchronize(b->stream);
        for (void *d : b->descs) (void)hipFreeAsync(d, b->stream);  // stream-ordered allocations
        (void)hipStreamSynchronize(b->stream);
    }
    for (hipGraphExec_t x : b->gexecs) (void)hipGraphExecDestroy(x);
    for (hipGraph_t x : b->graphs) (void)hipGraphDestroy(x);
    for (void *h : b->staging) (void)hipHostFree(h);
    if (b->arena) (void)hipFree(b->arena);
    if (b->d_meta) (void)hipFree(b->d_meta);
    if (b->d_part) (void)hipFree(b->d_part);
    if (b->h_out) (void)hipHostFree(b->h_out);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

// Runs fn(i) for i in [begin, end) on up to 16 host threads.
template <class F> void parallel_for(size_t begin, size_t end, F fn)
{
    std::atomic<size_t> next{begin};
    const unsigned nt = unsigned(std::max<size_t>(1, std::min<size_t>({16, std::thread::hardware_concurrency(), end - begin})));
    std::vector<std::thread> pool;
    std::vector<std::exception_ptr> errs(nt);
    for (unsigned w = 0; w < nt; ++w)
        pool.emplace_back([&, w] {
            try {
                for (size_t i; (i = next.fetch_add(1)) < end;) fn(i);
            } catch (...) {
                errs[w] = std::current_exception();
            }
        });
    for (auto &th : pool) th.join();
    for (auto &e : errs)
        if (e) std::rethrow_exception(e);
}

// The circuits run in groups of consecutive indices, pipelined: while the
// device runs group g's rounds, the host plans group g+1 and stages its
// descriptors in pinned memory, so only the first group's planning and
// upload are exposed.  A nominal group holds max(kGroupAmps, amps /
// kMaxGroups) amplitudes (enough tiles to fill the device in its early
// rounds).
constexpr uint64_t kGroupAmps = uint64_t(1) << 28;
constexpr size_t kMaxGroups = 8;

ev_batch *batch_run(const Circuit *const *cs, size_t count, const ev_run_options &opt)
{
    EV_REQUIRE(count > 0, EV_E_ARG, "empty batch");
    EV_REQUIRE(count < (1u << 24), EV_E_LIMIT, "batch too large");
    const bool c64 = opt.precision == EV_C64;
    const size_t ab = c64 ? 8 : 16;
    ev_batch *b = new ev_batch;
    b->prec = c64 ? kern::C64 : kern::C128;
    std::vector<void *> pinned;
    std::vector<hipEvent_t> events;
    std::vector<hipGraph_t> graphs;
    std::vector<hipGraphExec_t> gexecs;
    hipStream_t cap = nullptr;  // an idle stream the groups' rounds are captured on
    auto cleanup = [&] {
        if (b->stream) (void)hipStreamSynchronize(b->stream);
        for (void *h : pinned) (void)hipHostFree(h);
        pinned.clear();
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        events.clear();
        for (hipGraphExec_t x : gexecs) (void)hipGraphExecDestroy(x);
        for (hipGraph_t x : graphs) (void)hipGraphDestroy(x);
        gexecs.clear();
        graphs.clear();
        if (cap) (void)hipStreamDestroy(cap);
        cap = nullptr;
    };
    try {
        // ---- validate (on host threads; the first failing circuit in index
        // order is the one reported) and size
        b->n.resize(count);
        b->aoff.resize(count + 1);
        b->gates.resize(count);
        {
            std::vector<std::exception_ptr> verr(count);
            parallel_for(0, count, [&](size_t i) {
                try {
                    const Circuit &c = *cs[i];
                    c.validate();
                    const uint64_t n = c.num_qubits();
                    EV_REQUIRE(n >= uint64_t(fused::kMinTile) && n <= 30, EV_E_LIMIT,
                               "batched circuits need 8..30 qubits");
                    for (const Op &op : c.ops())
                        EV_REQUIRE(op.kind == OpKind::Apply1 || op.kind == OpKind::Apply2 ||
                                       op.kind == OpKind::Controlled,
                                   EV_E_ARG, "batched circuits must be measurement-free");
                } catch (...) {
                    verr[i] = std::current_exception();
                }
            });
            for (auto &e : verr)
                if (e) std::rethrow_exception(e);
        }
        uint64_t amps = 0;
        for (size_t i = 0; i < count; ++i) {
            const Circuit &c = *cs[i];
            const uint64_t n = c.num_qubits();
            b->n[i] = int(n);
            b->gates[i] = c.size();
            b->aoff[i] = amps;
            amps += uint64_t(1) << n;  // powers of two >= 256, so every offset stays a multiple of kChunk
        }
        b->aoff[count] = amps;
        b->total_amps = amps;

        // ---- groups of consecutive circuits: a sixteenth of the nominal
        // size (max(2^28, amps / 8) amplitudes), then 1.5x the previous one
        // up to nominal, so the device starts after planning a small first
        // group and the host, whose planning and staging of a group take
        // about 0.8x the device time of the same circuits, keeps ahead of
        // it.  (Doubling from a quarter left the device idle between the
        // first groups: measured on c3, 12 ms more wall time of 380.)
        std::vector<size_t> gbeg{0};
        {
            const uint64_t nominal = std::max<uint64_t>(kGroupAmps, amps / kMaxGroups);
            uint64_t target = 0, size = amps >= 2 * kGroupAmps ? nominal / 16 : amps;
            for (size_t i = 0; i < count;) {
                target = std::min<uint64_t>(amps, target + size);
                while (i < count && b->aoff[i] < target) ++i;
                if (i == gbeg.back()) ++i;  // at least one circuit per group
                gbeg.push_back(i);
                size = std::min<uint64_t>(nominal, size + size / 2);
            }
        }
        const size_t ngroups = gbeg.size() - 1;

        // ---- device memory, and every state set to |0..0>
        EV_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        EV_HIP(hipMalloc(&b->arena, size_t(amps) * ab));
        EV_HIP(hipMalloc(&b->d_meta, sizeof(uint64_t) * (2 * (count + 1))));
        const uint64_t nchunks = amps / batchk::kChunk;
        EV_HIP(hipMalloc(&b->d_part, sizeof(double) * (5 * nchunks + 5 * count + count)));
        EV_HIP(hipHostMalloc(reinterpret_cast<void **>(&b->h_out), sizeof(double) * 5 * count, hipHostMallocDefault));
        {
            std::vector<uint64_t> meta(2 * (count + 1));
            for (size_t i = 0; i <= count; ++i) {
                meta[i] = b->aoff[i];
                meta[count + 1 + i] = b->aoff[i] / batchk::kChunk;
            }
            EV_HIP(hipMemcpy(b->d_meta, meta.data(), meta.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
        }
        batchk::init(b->arena, c64, amps, b->d_meta, int(count), b->stream);

        char *abase = static_cast<char *>(b->arena);
        // Each group's rounds can run as one captured HIP graph.  They are
        // captured on an idle side stream (capture records the launches
        // without running them) and the graph is launched on the batch's
        // stream, so capturing group g+1 does not wait for group g to finish.
        const bool graph_ok = opt.use_graph != 0;
        if (graph_ok) EV_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        for (size_t g = 0; g < ngroups; ++g) {
            const size_t i0 = gbeg[g], i1 = gbeg[g + 1], gc = i1 - i0;
            if (gc == 0) continue;

            // ---- plan the group's circuits on host threads
            const double t_plan = now_ms();
            std::vector<Plan> plans(gc);
            parallel_for(0, gc, [&](size_t k) {
                const size_t i = i0 + k;
                plans[k] = plan_circuit(*cs[i], b->n[i], fused_config(b->n[i], c64, opt));
            });
            b->stats.plan_ms += now_ms() - t_plan;

            // ---- descriptor layout: per-circuit plans, then the round tables
            const double t_up = now_ms();
            std::vector<PlanLoc> loc(gc);
            size_t off = 0;
            uint32_t max_passes = 0;
            for (size_t k = 0; k < gc; ++k) {
                const Plan &P = plans[k];
                loc[k].pass_off = off;
                off += align256(P.passes.size() * sizeof(fused::Pass));
                loc[k].stage_off = off;
                off += align256(P.stages.size() * sizeof(fused::Stage));
                loc[k].op_off = off;
                off += align256(std::max<size_t>(1, P.devops.size()) * sizeof(fused::DevOp));
                loc[k].coef_off = off;
                off += align256(std::max<size_t>(1, P.coef.size()) * ab);
                loc[k].npass = uint32_t(P.passes.size());
                max_passes = std::max(max_passes, loc[k].npass);
                b->stats.passes += P.passes.size();
                b->stats.pass_bytes += double(P.passes.size()) * 2.0 * double(uint64_t(1) << b->n[i0 + k]) * double(ab);
                b->stats.pass_flops += plan_flops_per_amp(P) * double(uint64_t(1) << b->n[i0 + k]);
            }
            const size_t entries_off = off;
            std::vector<size_t> round_off(max_passes), round_n(max_passes);
            std::vector<uint64_t> round_tiles(max_passes);
            std::vector<int> round_t(max_passes, fused::kMinTile);
            size_t nent = 0;
            for (uint32_t r = 0; r < max_passes; ++r) {
                round_off[r] = nent;
                for (size_t k = 0; k < gc; ++k)
                    if (loc[k].npass > r) ++nent;
                round_n[r] = nent - round_off[r];
            }
            off += align256(std::max<size_t>(1, nent) * sizeof(fused::BatchEntry));
            const size_t desc_bytes = off;

            // ---- pinned host image of the descriptors (device pointers in the
            // entries); it stays alive until the stream has copied it
            void *dev = nullptr;
            EV_HIP(hipMallocAsync(&dev, desc_bytes, b->stream));
            b->descs.push_back(dev);
            unsigned char *host = nullptr;
            EV_HIP(hipHostMalloc(reinterpret_cast<void **>(&host), desc_bytes, hipHostMallocDefault));
            pinned.push_back(host);
            unsigned char *dbase = static_cast<unsigned char *>(dev);
            parallel_for(0, gc, [&](size_t k) {
                const Plan &P = plans[k];
                if (!P.passes.empty())
                    std::memcpy(host + loc[k].pass_off, P.passes.data(), P.passes.size() * sizeof(fused::Pass));
                if (!P.stages.empty())
                    std::memcpy(host + loc[k].stage_off, P.stages.data(), P.stages.size() * sizeof(fused::Stage));
                if (!P.devops.empty())
                    std::memcpy(host + loc[k].op_off, P.devops.data(), P.devops.size() * sizeof(fused::DevOp));
                if (c64)
                    fused::inline_coef_to_c64(reinterpret_cast<fused::DevOp *>(host + loc[k].op_off), P.devops.size());
                if (!c64) {
                    if (!P.coef.empty())
                        std::memcpy(host + loc[k].coef_off, P.coef.data(), P.coef.size() * sizeof(c128));
                } else {
                    float *f = reinterpret_cast<float *>(host + loc[k].coef_off);
                    for (size_t j = 0; j < P.coef.size(); ++j) {
                        f[2 * j] = float(P.coef[j].re);
                        f[2 * j + 1] = float(P.coef[j].im);
                    }
                }
            });
            fused::BatchEntry *ent = reinterpret_cast<fused::BatchEntry *>(host + entries_off);
            for (uint32_t r = 0; r < max_passes; ++r) {
                uint64_t tiles = 0;
                size_t e = round_off[r];
                for (size_t k = 0; k < gc; ++k) {
                    if (loc[k].npass <= r) continue;
                    const size_t i = i0 + k;
                    const fused::Pass &ps = plans[k].passes[r];
                    fused::BatchEntry &en = ent[e++];
                    en.psi = abase + b->aoff[i] * ab;
                    en.pass = reinterpret_cast<const fused::Pass *>(dbase + loc[k].pass_off) + r;
                    en.stages = reinterpret_cast<const fused::Stage *>(dbase + loc[k].stage_off);
                    en.ops = reinterpret_cast<const fused::DevOp *>(dbase + loc[k].op_off);
                    en.coef = dbase + loc[k].coef_off;
                    en.tile_begin = tiles;
                    en.tile_off = 0;
                    tiles += uint64_t(1) << (b->n[i] - ps.t);
                    round_t[r] = std::max(round_t[r], int(ps.t));
                }
                round_tiles[r] = tiles;
                b->stats.tiles += tiles;
            }
            plans.clear();
            EV_HIP(hipMemcpyAsync(dev, host, desc_bytes, hipMemcpyHostToDevice, b->stream));
            b->stats.upload_ms += now_ms() - t_up;

            // ---- the group's rounds (one captured graph per group when asked)
            const fused::BatchEntry *dent = reinterpret_cast<const fused::BatchEntry *>(dbase + entries_off);
            auto enqueue = [&](hipStream_t st) {
                for (uint32_t r = 0; r < max_passes; ++r)
                    fused::launch_batch(c64, dent + round_off[r], int(round_n[r]), round_tiles[r], round_t[r], st);
            };
            hipEvent_t e0, e1;
            EV_HIP(hipEventCreate(&e0));
            events.push_back(e0);
            EV_HIP(hipEventCreate(&e1));
            events.push_back(e1);
            hipGraphExec_t gexec = nullptr;
            if (graph_ok && max_passes > 0) {
                hipGraph_t graph = nullptr;
                EV_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
                enqueue(cap);
                EV_HIP(hipStreamEndCapture(cap, &graph));
                graphs.push_back(graph);
                EV_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
                gexecs.push_back(gexec);
                b->stats.graph = 1;
            }
            EV_HIP(hipEventRecord(e0, b->stream));
            if (gexec) EV_HIP(hipGraphLaunch(gexec, b->stream));
            else enqueue(b->stream);
            EV_HIP(hipEventRecord(e1, b->stream));
            b->stats.rounds += max_passes;
            ++b->stats.groups;
        }
        // ---- f64 renormalisation of every state (c64 drift over deep circuits),
        // queued behind the last group: f64 norms, 1/sqrt on the device, scale
        hipEvent_t rn[2] = {nullptr, nullptr};
        if (opt.renormalize) {
            EV_HIP(hipEventCreate(&rn[0]));
            events.push_back(rn[0]);  // destroyed with the group events
            EV_HIP(hipEventCreate(&rn[1]));
            events.push_back(rn[1]);
            double *res = b->d_part + 5 * nchunks, *d_scale = res + 5 * count;
            EV_HIP(hipEventRecord(rn[0], b->stream));
            batchk::norms(b->arena, c64, amps, b->d_meta + count + 1, int(count), b->d_part, res, b->stream);
            batchk::inv_sqrt(res, int(count), d_scale, b->stream);
            batchk::scale(b->arena, c64, amps, b->d_meta + count + 1, int(count), d_scale, b->stream);
            EV_HIP(hipEventRecord(rn[1], b->stream));
        }
        EV_HIP(hipStreamSynchronize(b->stream));
        for (size_t k = 0; k + 1 < events.size(); k += 2) {
            float ms = 0;
            EV_HIP(hipEventElapsedTime(&ms, events[k], events[k + 1]));
            (opt.renormalize && k + 2 == events.size() ? b->stats.renorm_ms : b->stats.device_ms) += ms;
        }
        // the staging buffers and graphs stay with the batch (batch_free)
        b->staging.swap(pinned);
        b->graphs.swap(graphs);
        b->gexecs.swap(gexecs);
        cleanup();
        b->stats.circuits = count;
        for (auto g : b->gates) b->stats.logical_gates += g;


    } catch (...) {
        cleanup();
        batch_free(b);
        throw;
    }
    return b;
}

void batch_norms(ev_batch *b, double *out)
{
    const size_t count = b->n.size();
    const uint64_t nchunks = b->total_amps / batchk::kChunk;
    double *res = b->d_part + 5 * nchunks;
    batchk::norms(b->arena, b->prec == kern::C64, b->total_amps, b->d_meta + count + 1, int(count), b->d_part, res,
                  b->stream);
    EV_HIP(hipMemcpyAsync(b->h_out, res, count * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    EV_HIP(hipStreamSynchronize(b->stream));
    std::memcpy(out, b->h_out, count * sizeof(double));
}

void batch_compare(ev_batch *b, ev_batch *ref, double *fid, double *maxabs)
{
    const size_t count = b->n.size();
    EV_REQUIRE(ref->n == b->n, EV_E_ARG, "batches hold different circuit sizes");
    EV_REQUIRE(ref->prec == kern::C128, EV_E_ARG, "the reference batch must be c128");
    EV_HIP(hipStreamSynchronize(ref->stream));
    const uint64_t nchunks = b->total_amps / batchk::kChunk;
    double *res = b->d_part + 5 * nchunks;
    batchk::compare(b->arena, b->prec == kern::C64, ref->arena, b->total_amps, b->d_meta + count + 1, int(count), b->d_part,
                    res, b->stream);
    EV_HIP(hipMemcpyAsync(b->h_out, res, 5 * count * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    EV_HIP(hipStreamSynchronize(b->stream));
    for (size_t i = 0; i < count; ++i) {
        const double *r = b->h_out + 5 * i;
        const double ov = r[0] * r[0] + r[1] * r[1];
        fid[i] = (r[2] > 0 && r[3] > 0) ? ov / (r[2] * r[3]) : 0.0;
        maxabs[i] = r[4];
    }
}

void batch_read(ev_batch *b, size_t i, uint64_t offset, c128 *host, uint64_t len)
{
    EV_REQUIRE(i < b->n.size(), EV_E_ARG, "batch index out of range");
    const uint64_t dim = uint64_t(1) << b->n[i];
    EV_REQUIRE(offset <= dim && len <= dim - offset, EV_E_ARG, "amplitude range out of bounds");
    if (len == 0) return;
    const char *src = static_cast<const char *>(b->arena) + (b->aoff[i] + offset) * kern::amp_bytes(b->prec);
    if (b->prec == kern::C128) {
        EV_HIP(hipMemcpyAsync(host, src, len * 16, hipMemcpyDeviceToHost, b->stream));
        EV_HIP(hipStreamSynchronize(b->stream));
        return;
    }
    void *tmp = nullptr;
    EV_HIP(hipMalloc(&tmp, len * sizeof(c128)));
    kern::c64_to_c128(src, tmp, len, b->stream);
    const hipError_t e = hipMemcpyAsync(host, tmp, len * sizeof(c128), hipMemcpyDeviceToHost, b->stream);
    const hipError_t e2 = hipStreamSynchronize(b->stream);
    (void)hipFree(tmp);
    EV_HIP(e);
    EV_HIP(e2);
}

}  // namespace ev
