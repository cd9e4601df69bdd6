// Statevector handle (device memory, stream, reductions) and the per-op
// executor that mirrors src/backend/statevector.rs and src/backend/mod.rs.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "engine.hpp"
#include "planner.hpp"

namespace ev {

[[noreturn]] void throw_hip(int err, const char *what, const char *file, int line)
{
    const int code = (err == hipErrorOutOfMemory || err == hipErrorMemoryAllocation) ? EV_E_OOM : EV_E_HIP;
    throw Error(code, std::string(hipGetErrorString(static_cast<hipError_t>(err))) + " in " + what + " (" + file + ":" +
                          std::to_string(line) + ")");
}

RunStats &last_stats()
{
    static thread_local RunStats s;
    return s;
}

ev_sv *sv_alloc(int n, kern::Prec p)
{
    EV_REQUIRE(n >= 0 && uint64_t(n) <= kMaxQubits, EV_E_LIMIT,
               "statevector of " + std::to_string(n) + " qubits exceeds this build's limit of 34");
    ev_sv *s = new ev_sv;
    s->n = n;
    s->prec = p;
    try {
        EV_HIP(hipGetDevice(&s->device));
        EV_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        EV_HIP(hipMalloc(&s->psi, s->bytes()));
        EV_HIP(hipMalloc(&s->d_red, sizeof(double) * (kRedPartials + kRedSlots)));
        EV_HIP(hipHostMalloc(reinterpret_cast<void **>(&s->h_red), sizeof(double) * kRedSlots, hipHostMallocDefault));
    } catch (...) {
        sv_destroy(s);
        throw;
    }
    return s;
}

ev_sv *sv_create(int n, kern::Prec p)
{
    ev_sv *s = sv_alloc(n, p);
    kern::init_basis(s->psi, p, n, 0, s->stream);  // state.rs:38-44
    return s;
}

void sv_destroy(ev_sv *s)
{
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->psi) (void)hipFree(s->psi);
    if (s->d_red) (void)hipFree(s->d_red);
    if (s->plan_buf) (void)hipFree(s->plan_buf);
    if (s->plan_host) (void)hipHostFree(s->plan_host);
    if (s->h_red) (void)hipHostFree(s->h_red);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

void sv_sync(ev_sv *s) { EV_HIP(hipStreamSynchronize(s->stream)); }

static double *slots(ev_sv *s) { return s->d_red + kRedPartials; }

static void fetch_slots(ev_sv *s, int count)
{
    EV_HIP(hipMemcpyAsync(s->h_red, slots(s), sizeof(double) * count, hipMemcpyDeviceToHost, s->stream));
    sv_sync(s);
}

double sv_norm_sqr(ev_sv *s)
{
    kern::reduce_norm(s->psi, s->prec, s->n, s->d_red, slots(s), s->stream);
    fetch_slots(s, 1);
    return s->h_red[0];
}

double sv_p1(ev_sv *s, int q)
{
    kern::reduce_p1(s->psi, s->prec, s->n, q, s->d_red, slots(s), s->stream);
    fetch_slots(s, 1);
    return s->h_red[0];
}

c128 sv_overlap(ev_sv *s, ev_sv *other)
{
    EV_REQUIRE(s->n == other->n, EV_E_DIM, "states must have equal qubit counts");
    EV_REQUIRE(s->prec == other->prec, EV_E_ARG, "states must have equal precision");
    sv_sync(other);
    kern::reduce_overlap(s->psi, other->psi, s->prec, s->n, s->d_red, slots(s), s->stream);
    fetch_slots(s, 2);
    return {s->h_red[0], s->h_red[1]};
}

void sv_bloch(ev_sv *s, int q, double out[4])
{
    kern::reduce_bloch(s->psi, s->prec, s->n, q, s->d_red, slots(s), s->stream);
    fetch_slots(s, 4);
    for (int i = 0; i < 4; ++i) out[i] = s->h_red[i];
}

// src/measure.rs:15-51.  A c64 state is first rescaled to unit f64 norm, so
// the Born probabilities do not inherit the f32 drift of the gates before.
bool sv_measure(ev_sv *s, int q, double u)
{
    if (s->prec == kern::C64) sv_normalize(s);
    const double p1 = sv_p1(s, q);
    const bool outcome = u < p1;
    const double surviving = outcome ? p1 : 1.0 - p1;
    const double scale = surviving > 0.0 ? 1.0 / std::sqrt(surviving) : 0.0;
    kern::collapse(s->psi, s->prec, s->n, q, outcome ? 1 : 0, scale, s->stream);
    return outcome;
}

// src/state.rs:103-111
void sv_normalize(ev_sv *s) { kern::normalize(s->psi, s->prec, s->n, s->d_red, slots(s), s->stream); }

void sv_read(ev_sv *s, c128 *host, uint64_t len)
{
    EV_REQUIRE(len == s->len(), EV_E_DIM, "read buffer length does not match 2^n");
    if (s->prec == kern::C128) {
        EV_HIP(hipMemcpyAsync(host, s->psi, s->bytes(), hipMemcpyDeviceToHost, s->stream));
        sv_sync(s);
        return;
    }
    void *tmp = nullptr;
    EV_HIP(hipMalloc(&tmp, len * sizeof(c128)));
    kern::c64_to_c128(s->psi, tmp, len, s->stream);
    const hipError_t e = hipMemcpyAsync(host, tmp, len * sizeof(c128), hipMemcpyDeviceToHost, s->stream);
    const hipError_t e2 = hipStreamSynchronize(s->stream);
    (void)hipFree(tmp);
    EV_HIP(e);
    EV_HIP(e2);
}

void sv_read_range(ev_sv *s, uint64_t offset, c128 *host, uint64_t len)
{
    EV_REQUIRE(offset <= s->len() && len <= s->len() - offset, EV_E_ARG, "amplitude range out of bounds");
    if (len == 0) return;
    if (s->prec == kern::C128) {
        EV_HIP(hipMemcpyAsync(host, static_cast<const char *>(s->psi) + offset * 16, len * 16, hipMemcpyDeviceToHost,
                              s->stream));
        sv_sync(s);
        return;
    }
    void *tmp = nullptr;
    EV_HIP(hipMalloc(&tmp, len * sizeof(c128)));
    kern::c64_to_c128(static_cast<const char *>(s->psi) + offset * 8, tmp, len, s->stream);
    const hipError_t e = hipMemcpyAsync(host, tmp, len * sizeof(c128), hipMemcpyDeviceToHost, s->stream);
    const hipError_t e2 = hipStreamSynchronize(s->stream);
    (void)hipFree(tmp);
    EV_HIP(e);
    EV_HIP(e2);
}

void sv_write(ev_sv *s, const c128 *host, uint64_t len)
{
    EV_REQUIRE(len == s->len(), EV_E_DIM, "write buffer length does not match 2^n");
    if (s->prec == kern::C128) {
        EV_HIP(hipMemcpyAsync(s->psi, host, s->bytes(), hipMemcpyHostToDevice, s->stream));
        sv_sync(s);
        return;
    }
    void *tmp = nullptr;
    EV_HIP(hipMalloc(&tmp, len * sizeof(c128)));
    hipError_t e = hipMemcpyAsync(tmp, host, len * sizeof(c128), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) kern::c128_to_c64(tmp, s->psi, len, s->stream);
    const hipError_t e2 = hipStreamSynchronize(s->stream);
    (void)hipFree(tmp);
    EV_HIP(e);
    EV_HIP(e2);
}

void sv_sample(ev_sv *s, uint64_t shots, uint64_t seed, uint64_t *out)
{
    Rng rng = Rng::seed_from_u64(seed);
    std::vector<double> u(shots);
    for (auto &x : u) x = rng.next_f64();  // one draw per shot, in shot order
    const size_t nscr = kern::sample_scratch_doubles(s->n);
    double *d_scr = nullptr, *d_u = nullptr;
    uint64_t *d_out = nullptr;
    auto cleanup = [&] {
        (void)hipStreamSynchronize(s->stream);
        if (d_scr) (void)hipFree(d_scr);
        if (d_u) (void)hipFree(d_u);
        if (d_out) (void)hipFree(d_out);
    };
    try {
        EV_HIP(hipMalloc(&d_scr, nscr * sizeof(double)));
        if (shots) {
            EV_HIP(hipMalloc(&d_u, shots * sizeof(double)));
            EV_HIP(hipMalloc(&d_out, shots * sizeof(uint64_t)));
            EV_HIP(hipMemcpyAsync(d_u, u.data(), shots * sizeof(double), hipMemcpyHostToDevice, s->stream));
        }
        kern::sample(s->psi, s->prec, s->n, d_u, shots, d_scr, d_out, s->stream);
        if (shots) EV_HIP(hipMemcpyAsync(out, d_out, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream));
        sv_sync(s);
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

// sv_sample's result without a device allocation: the uniforms are drawn
// into the caller's output buffer (one double per shot, in shot order), which
// the library's fixed scratch then takes as staging in both directions.
void sv_sample_noalloc(ev_sv *s, uint64_t shots, uint64_t seed, uint64_t *out)
{
    EV_REQUIRE(shots <= kern::kNoAllocMaxShots, EV_E_LIMIT,
               "ev_sv_sample_noalloc takes at most " + std::to_string(kern::kNoAllocMaxShots) + " shots per call");
    Rng rng = Rng::seed_from_u64(seed);
    for (uint64_t i = 0; i < shots; ++i) {
        const double u = rng.next_f64();
        std::memcpy(out + i, &u, sizeof u);
    }
    kern::sample_noalloc(s->psi, s->prec, s->n, shots, out, s->stream, s->device);
}

void sv_apply1(ev_sv *s, int k, const c128 m[4]) { kern::apply1(s->psi, s->prec, s->n, k, m, s->stream); }
void sv_apply2(ev_sv *s, int a, int b, const c128 m[16]) { kern::apply2(s->psi, s->prec, s->n, a, b, m, s->stream); }
void sv_ctrl1(ev_sv *s, const uint64_t *controls, int nc, int target, const c128 m[4])
{
    kern::ctrl1(s->psi, s->prec, s->n, controls, nc, target, m, s->stream);
}

// --------------------------------------------------------------- profiling

Profiler::~Profiler()
{
    for (auto &r : recs_) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    if (open_) (void)hipEventDestroy(open_);
}

void Profiler::begin()
{
    if (!on_) return;
    EV_HIP(hipEventCreate(&open_));
    EV_HIP(hipEventRecord(open_, st_));
}

void Profiler::end(const char *name, double bytes)
{
    if (!on_) return;
    hipEvent_t b;
    EV_HIP(hipEventCreate(&b));
    EV_HIP(hipEventRecord(b, st_));
    recs_.push_back(Rec{open_, b, name, bytes});
    open_ = nullptr;
}

std::vector<KernelStat> Profiler::collect()
{
    std::vector<KernelStat> out;
    if (!on_) return out;
    EV_HIP(hipStreamSynchronize(st_));
    for (auto &r : recs_) {
        float ms = 0;
        EV_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        KernelStat *k = nullptr;
        for (auto &x : out)
            if (x.name == r.name) k = &x;
        if (!k) {
            out.push_back(KernelStat{r.name});
            k = &out.back();
        }
        ++k->launches;
        k->total_ms += ms;
        k->bytes += r.bytes;
    }
    return out;
}

const char *k1_name(K1 k)
{
    switch (k) {
    case K1::Dense: return "apply1_dense";
    case K1::Diag: return "apply1_diag";
    case K1::Phase: return "apply1_phase";
    }
    return "?";
}

const char *k2_name(K2 k)
{
    switch (k) {
    case K2::Cnot: return "cnot";
    case K2::Cz: return "cz";
    case K2::Swap: return "swap";
    case K2::Dense: return "apply2_dense";
    }
    return "?";
}

double k1_bytes(K1 k, int n, size_t s) { return k == K1::Phase ? full_pass_bytes(n - 1, s) : full_pass_bytes(n, s); }

double k2_bytes(K2 k, int n, size_t s)
{
    switch (k) {
    case K2::Cnot:
    case K2::Swap: return full_pass_bytes(n - 1, s);
    case K2::Cz: return full_pass_bytes(n - 2, s);
    case K2::Dense: return full_pass_bytes(n, s);
    }
    return 0;
}

// ------------------------------------------------------------- executors

namespace {

struct EventTimer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t st;
    explicit EventTimer(hipStream_t s) : st(s)
    {
        EV_HIP(hipEventCreate(&a));
        EV_HIP(hipEventCreate(&b));
        EV_HIP(hipEventRecord(a, st));
    }
    double stop()
    {
        EV_HIP(hipEventRecord(b, st));
        EV_HIP(hipEventSynchronize(b));
        float ms = 0;
        EV_HIP(hipEventElapsedTime(&ms, a, b));
        return ms;
    }
    ~EventTimer()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

}  // namespace

// backend/mod.rs:51-72, one launch per gate.
size_t sv_plain_op(ev_sv *s, const Circuit &c, size_t i, Rng &rng, uint8_t *classical, bool execute, RunStats &st,
                Profiler &prof)
{
    const Op &op = c.ops()[i];
    const size_t ab = kern::amp_bytes(s->prec);
    switch (op.kind) {
    case OpKind::Apply1:
        if (execute) {
            const K1 k = classify1(c.mat(op));
            prof.begin();
            sv_apply1(s, int(op.q0), c.mat(op));
            prof.end(k1_name(k), k1_bytes(k, s->n, ab));
            ++st.logical_gates;
            ++st.launches;
        }
        return i + 1;
    case OpKind::Apply2:
        if (execute) {
            const K2 k = classify2(c.mat(op));
            prof.begin();
            sv_apply2(s, int(op.q0), int(op.q1), c.mat(op));
            prof.end(k2_name(k), k2_bytes(k, s->n, ab));
            ++st.logical_gates;
            ++st.launches;
        }
        return i + 1;
    case OpKind::Controlled:
        if (execute) {
            prof.begin();
            sv_ctrl1(s, c.ctrls(op), int(op.nctrl), int(op.q0), c.mat(op));
            prof.end("ctrl1", ctrl_bytes(s->n, int(op.nctrl), ab));
            ++st.logical_gates;
            ++st.launches;
        }
        return i + 1;
    case OpKind::Measure:
        if (execute) classical[op.q1] = sv_measure(s, int(op.q0), rng.next_f64()) ? 1 : 0;
        return i + 1;
    case OpKind::If: return sv_plain_op(s, c, i + 1, rng, classical, execute && classical[op.q0] != 0, st, prof);
    }
    throw Error(EV_E_ARG, "corrupt op record");
}

void sv_execute(ev_sv *s, const Circuit &c, Rng &rng, uint8_t *classical, const ev_run_options &opt)
{
    RunStats st;
    Profiler prof(s->stream, opt.profile != 0);
    EventTimer timer(s->stream);
    if (opt.plain) {
        size_t i = 0;
        while (i < c.size()) i = sv_plain_op(s, c, i, rng, classical, true, st, prof);
    } else {
        execute_planned(s, c, rng, classical, opt, st, prof);
    }
    if (s->prec == kern::C64) sv_normalize(s);  // f64 renormalisation of the f32 bulk state
    st.device_ms = timer.stop();
    st.kernels = prof.collect();
    last_stats() = st;
}

void sv_run_into(ev_sv *s, const Circuit &c, uint64_t seed, uint8_t *classical, const ev_run_options &opt)
{
    EV_REQUIRE(uint64_t(s->n) == c.num_qubits(), EV_E_DIM, "state and circuit qubit counts differ");
    EV_REQUIRE(opt.precision == (s->prec == kern::C128 ? EV_C128 : EV_C64), EV_E_ARG,
               "run precision differs from the state's precision");
    kern::init_basis(s->psi, s->prec, s->n, 0, s->stream);  // state.rs:38-44
    Rng rng = Rng::seed_from_u64(seed);                      // statevector.rs:81
    for (uint64_t b = 0; b < c.num_classical(); ++b) classical[b] = 0;
    sv_execute(s, c, rng, classical, opt);
}

}  // namespace ev
