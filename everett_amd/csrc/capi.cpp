// extern "C" entry points of libeverett.so (declared in include/everett.h).
// Every call converts C++ exceptions into an ev_status code and records the
// message as the calling thread's last error.
#include <hip/hip_runtime_api.h>

#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "batch.hpp"
#include "dm_ops.hpp"
#include "engine.hpp"
#include "fused.hpp"
#include "planner.hpp"
#include "qasm.hpp"
#include "server.hpp"
#include "tune.hpp"

struct ev_circuit {
    ev::Circuit c;
};

namespace {

thread_local std::string g_last_error;

template <class F> int guard(F &&f, ev_error *err = nullptr)
{
    try {
        f();
        if (err) std::memset(err, 0, sizeof *err);
        return EV_OK;
    } catch (const ev::Error &e) {
        g_last_error = e.what();
        if (err) {
            err->code = e.code;
            err->a = e.a;
            err->b = e.b;
            std::strncpy(err->message, e.what(), sizeof err->message - 1);
            err->message[sizeof err->message - 1] = 0;
        }
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        if (err) {
            std::memset(err, 0, sizeof *err);
            err->code = EV_E_OOM;
        }
        return EV_E_OOM;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        if (err) {
            std::memset(err, 0, sizeof *err);
            err->code = EV_E_ARG;
        }
        return EV_E_ARG;
    }
}

#define NONNULL(p) EV_REQUIRE((p) != nullptr, EV_E_ARG, #p " is null")

ev::kern::Prec prec_of(ev_precision p)
{
    EV_REQUIRE(p == EV_C128 || p == EV_C64, EV_E_ARG, "unknown precision");
    return p == EV_C128 ? ev::kern::C128 : ev::kern::C64;
}

// n from a power-of-two length (src/state.rs:51-62 DimensionMismatch otherwise).
int qubits_of_len(size_t len)
{
    if (len == 0 || (len & (len - 1)) != 0) {
        size_t next = 1;
        while (next < len) next <<= 1;
        throw ev::Error(EV_E_DIM, "amplitude count " + std::to_string(len) + " is not a power of two", len, next);
    }
    return __builtin_ctzll(len);
}

void check_q(uint64_t q, int n)
{
    if (q >= uint64_t(n))
        throw ev::Error(EV_E_QUBIT_RANGE, "qubit " + std::to_string(q) + " out of range for " + std::to_string(n) + " qubits",
                        q, uint64_t(n));
}

void check_controls(const uint64_t *controls, size_t nc, uint64_t target, int n)
{
    EV_REQUIRE(nc == 0 || controls != nullptr, EV_E_ARG, "controls is null");
    check_q(target, n);
    for (size_t i = 0; i < nc; ++i) {
        check_q(controls[i], n);
        if (controls[i] == target)
            throw ev::Error(EV_E_DUP_QUBIT, "target qubit " + std::to_string(target) + " is also a control", target);
    }
}

const ev::c128 *C(const ev_c128 *m) { return reinterpret_cast<const ev::c128 *>(m); }

}  // namespace

extern "C" {

const char *ev_last_error(void) { return g_last_error.c_str(); }

// ------------------------------------------------------------------ circuit

int ev_circuit_create(uint64_t num_qubits, uint64_t num_classical, ev_circuit **out)
{
    return guard([&] {
        NONNULL(out);
        *out = new ev_circuit{ev::Circuit(num_qubits, num_classical)};
    });
}

void ev_circuit_destroy(ev_circuit *c) { delete c; }

int ev_circuit_append(ev_circuit *c, const ev_op *ops, size_t count)
{
    return guard([&] {
        NONNULL(c);
        EV_REQUIRE(count == 0 || ops != nullptr, EV_E_ARG, "ops is null");
        for (size_t i = 0; i < count;) i = c->c.append(ops, count, i);
    });
}

int ev_circuit_compose(ev_circuit *dst, const ev_circuit *src)
{
    return guard([&] {
        NONNULL(dst);
        NONNULL(src);
        ev::Circuit copy = src->c;  // also correct when dst == src
        dst->c.compose(copy);
    });
}

size_t ev_circuit_len(const ev_circuit *c) { return c ? c->c.record_count() : 0; }
uint64_t ev_circuit_num_qubits(const ev_circuit *c) { return c ? c->c.num_qubits() : 0; }
uint64_t ev_circuit_num_classical(const ev_circuit *c) { return c ? c->c.num_classical() : 0; }

int ev_circuit_export(const ev_circuit *c, ev_op *out, size_t cap, size_t *len)
{
    return guard([&] {
        NONNULL(c);
        NONNULL(len);
        const size_t nrec = c->c.record_count();
        *len = nrec;
        EV_REQUIRE(out == nullptr || cap >= nrec, EV_E_ARG, "export buffer too small");
        if (out)
            for (size_t i = 0, r = 0; i < c->c.size(); ++i) r += c->c.to_ev(i, out + r);
    });
}

int ev_circuit_validate(const ev_circuit *c, ev_error *err)
{
    return guard(
        [&] {
            NONNULL(c);
            c->c.validate();
        },
        err);
}

static int make_circuit(ev_circuit **out, ev::Circuit &&c)
{
    *out = new ev_circuit{std::move(c)};
    return EV_OK;
}

int ev_circuit_from_qasm(const char *src, size_t len, ev_circuit **out, ev_error *err)
{
    return guard(
        [&] {
            NONNULL(out);
            EV_REQUIRE(src != nullptr || len == 0, EV_E_ARG, "src is null");
            make_circuit(out, ev::parse_qasm(src ? src : "", len));
        },
        err);
}

int ev_circuit_to_qasm(const ev_circuit *c, char *buf, size_t cap, size_t *len, ev_error *err)
{
    return guard(
        [&] {
            NONNULL(c);
            NONNULL(len);
            const std::string text = ev::emit_qasm(c->c);
            *len = text.size();
            EV_REQUIRE(buf != nullptr && cap > text.size(), EV_E_LIMIT,
                       "buffer of " + std::to_string(cap) + " bytes cannot hold " + std::to_string(text.size() + 1));
            std::memcpy(buf, text.c_str(), text.size() + 1);
        },
        err);
}

void ev_rng_seed(ev_rng *r, uint64_t seed)
{
    const ev::Rng x = ev::Rng::seed_from_u64(seed);
    std::memcpy(r->s, x.s, sizeof r->s);
}

uint64_t ev_rng_next_u64(ev_rng *r)
{
    ev::Rng x;
    std::memcpy(x.s, r->s, sizeof x.s);
    const uint64_t v = x.next_u64();
    std::memcpy(r->s, x.s, sizeof x.s);
    return v;
}

double ev_rng_next_f64(ev_rng *r)
{
    ev::Rng x;
    std::memcpy(x.s, r->s, sizeof x.s);
    const double v = x.next_f64();
    std::memcpy(r->s, x.s, sizeof x.s);
    return v;
}

int ev_circuit_qft(uint64_t n, ev_circuit **out)
{
    return guard([&] {
        NONNULL(out);
        EV_REQUIRE(n <= 31, EV_E_LIMIT, "qft angle table needs n <= 31 (1u32 << m)");
        make_circuit(out, ev::gen_qft(n));
    });
}

int ev_circuit_iqft(uint64_t n, ev_circuit **out)
{
    return guard([&] {
        NONNULL(out);
        EV_REQUIRE(n <= 31, EV_E_LIMIT, "qft angle table needs n <= 31 (1u32 << m)");
        make_circuit(out, ev::gen_iqft(n));
    });
}

static ev::Rng rng_in(const ev_rng *r)
{
    ev::Rng x;
    std::memcpy(x.s, r->s, sizeof x.s);
    return x;
}
static void rng_out(ev_rng *r, const ev::Rng &x) { std::memcpy(r->s, x.s, sizeof x.s); }

int ev_circuit_brickwall(uint64_t n, uint64_t depth, ev_rng *rng, ev_circuit **out)
{
    return guard([&] {
        NONNULL(out);
        NONNULL(rng);
        ev::Rng r = rng_in(rng);
        make_circuit(out, ev::gen_brickwall(n, depth, r));
        rng_out(rng, r);
    });
}

int ev_circuit_dm_random(uint64_t n, uint64_t depth, ev_rng *rng, ev_circuit **out)
{
    return guard([&] {
        NONNULL(out);
        NONNULL(rng);
        EV_REQUIRE(n >= 1, EV_E_ARG, "dm_random needs n >= 1");
        ev::Rng r = rng_in(rng);
        make_circuit(out, ev::gen_dm_random(n, depth, r));
        rng_out(rng, r);
    });
}

// ------------------------------------------------------------- statevector

int ev_sv_create(uint64_t n, ev_precision p, ev_sv **out)
{
    return guard([&] {
        NONNULL(out);
        EV_REQUIRE(n <= ev::kMaxQubits, EV_E_LIMIT, "n exceeds 34");
        *out = ev::sv_create(int(n), prec_of(p));
        ev::sv_sync(*out);
    });
}

int ev_sv_from_amplitudes(const ev_c128 *amps, size_t len, ev_precision p, ev_sv **out)
{
    return guard([&] {
        NONNULL(out);
        NONNULL(amps);
        const int n = qubits_of_len(len);
        ev_sv *s = ev::sv_alloc(n, prec_of(p));
        try {
            ev::sv_write(s, C(amps), len);
        } catch (...) {
            ev::sv_destroy(s);
            throw;
        }
        *out = s;
    });
}

void ev_sv_destroy(ev_sv *s) { ev::sv_destroy(s); }
uint64_t ev_sv_num_qubits(const ev_sv *s) { return s ? uint64_t(s->n) : 0; }
ev_precision ev_sv_precision(const ev_sv *s) { return s && s->prec == ev::kern::C64 ? EV_C64 : EV_C128; }

int ev_sv_read(const ev_sv *s, ev_c128 *host, size_t len)
{
    return guard([&] {
        NONNULL(s);
        NONNULL(host);
        ev::sv_read(const_cast<ev_sv *>(s), reinterpret_cast<ev::c128 *>(host), len);
    });
}

int ev_sv_read_range(const ev_sv *s, uint64_t offset, ev_c128 *host, size_t len)
{
    return guard([&] {
        NONNULL(s);
        EV_REQUIRE(len == 0 || host != nullptr, EV_E_ARG, "host is null");
        ev::sv_read_range(const_cast<ev_sv *>(s), offset, reinterpret_cast<ev::c128 *>(host), len);
    });
}

int ev_sv_write(ev_sv *s, const ev_c128 *host, size_t len)
{
    return guard([&] {
        NONNULL(s);
        NONNULL(host);
        ev::sv_write(s, C(host), len);
    });
}

int ev_sv_copy(const ev_sv *src, ev_sv **out)
{
    return guard([&] {
        NONNULL(src);
        NONNULL(out);
        ev_sv *s = ev::sv_alloc(src->n, src->prec);
        ev::sv_sync(const_cast<ev_sv *>(src));
        const hipError_t e = hipMemcpyAsync(s->psi, src->psi, src->bytes(), hipMemcpyDeviceToDevice, s->stream);
        const hipError_t e2 = hipStreamSynchronize(s->stream);
        if (e != hipSuccess || e2 != hipSuccess) {
            ev::sv_destroy(s);
            EV_HIP(e != hipSuccess ? e : e2);
        }
        *out = s;
    });
}

int ev_sv_apply_1q(ev_sv *s, uint64_t k, const ev_c128 m[4])
{
    return guard([&] {
        NONNULL(s);
        NONNULL(m);
        check_q(k, s->n);
        ev::sv_apply1(s, int(k), C(m));
    });
}

int ev_sv_apply_2q(ev_sv *s, uint64_t a, uint64_t b, const ev_c128 m[16])
{
    return guard([&] {
        NONNULL(s);
        NONNULL(m);
        check_q(a, s->n);
        check_q(b, s->n);
        if (a == b) throw ev::Error(EV_E_DUP_QUBIT, "qubit " + std::to_string(a) + " used twice", a);
        ev::sv_apply2(s, int(a), int(b), C(m));
    });
}

int ev_sv_apply_controlled_1q(ev_sv *s, const uint64_t *controls, size_t nc, uint64_t target, const ev_c128 m[4])
{
    return guard([&] {
        NONNULL(s);
        NONNULL(m);
        check_controls(controls, nc, target, s->n);
        ev::sv_ctrl1(s, controls, int(nc), int(target), C(m));
    });
}

int ev_sv_measure(ev_sv *s, uint64_t qubit, double u, int *outcome)
{
    return guard([&] {
        NONNULL(s);
        NONNULL(outcome);
        check_q(qubit, s->n);
        *outcome = ev::sv_measure(s, int(qubit), u) ? 1 : 0;
        ev::sv_sync(s);
    });
}

int ev_sv_norm_sqr(const ev_sv *s, double *out)
{
    return guard([&] {
        NONNULL(s);
        NONNULL(out);
        *out = ev::sv_norm_sqr(const_cast<ev_sv *>(s));
    });
}

int ev_sv_normalize(ev_sv *s)
{
    return guard([&] {
        NONNULL(s);
        ev::sv_normalize(s);
        ev::sv_sync(s);
    });
}

int ev_sv_probability(const ev_sv *s, uint64_t basis, double *out)
{
    return guard([&] {
        NONNULL(s);
        NONNULL(out);
        EV_REQUIRE(basis < s->len(), EV_E_ARG, "basis state " + std::to_string(basis) + " out of range");
        ev_sv *m = const_cast<ev_sv *>(s);
        if (s->prec == ev::kern::C128) {
            ev::c128 a;
            EV_HIP(hipMemcpyAsync(&a, static_cast<const char *>(s->psi) + basis * 16, 16, hipMemcpyDeviceToHost, m->stream));
            ev::sv_sync(m);
            *out = __builtin_fma(a.re, a.re, a.im * a.im);
        } else {
            float a[2];
            EV_HIP(hipMemcpyAsync(a, static_cast<const char *>(s->psi) + basis * 8, 8, hipMemcpyDeviceToHost, m->stream));
            ev::sv_sync(m);
            const double re = a[0], im = a[1];
            *out = __builtin_fma(re, re, im * im);
        }
    });
}

int ev_sv_overlap(const ev_sv *s, const ev_sv *other, ev_c128 *out)
{
    return guard([&] {
        NONNULL(s);
        NONNULL(other);
        NONNULL(out);
        const ev::c128 v = ev::sv_overlap(const_cast<ev_sv *>(s), const_cast<ev_sv *>(other));
        out->re = v.re;
        out->im = v.im;
    });
}

int ev_sv_fidelity(const ev_sv *s, const ev_sv *other, double *out)
{
    return guard([&] {
        NONNULL(s);
        NONNULL(other);
        NONNULL(out);
        const ev::c128 v = ev::sv_overlap(const_cast<ev_sv *>(s), const_cast<ev_sv *>(other));
        *out = __builtin_fma(v.re, v.re, v.im * v.im);  // state.rs:149-151
    });
}

int ev_sv_bloch_vector(const ev_sv *s, uint64_t qubit, double out[3])
{
    return guard([&] {
        NONNULL(s);
        NONNULL(out);
        check_q(qubit, s->n);
        double r[4];
        ev::sv_bloch(const_cast<ev_sv *>(s), int(qubit), r);
        out[0] = 2.0 * r[2];  // state.rs:184
        out[1] = 2.0 * r[3];
        out[2] = r[0] - r[1];
    });
}

int ev_sv_sample(const ev_sv *s, uint64_t shots, uint64_t seed, uint64_t *out)
{
    return guard([&] {
        NONNULL(s);
        EV_REQUIRE(shots == 0 || out != nullptr, EV_E_ARG, "out is null");
        ev::sv_sample(const_cast<ev_sv *>(s), shots, seed, out);
    });
}

int ev_sv_sample_noalloc(const ev_sv *s, uint64_t shots, uint64_t seed, uint64_t *out)
{
    return guard([&] {
        NONNULL(s);
        EV_REQUIRE(shots == 0 || out != nullptr, EV_E_ARG, "out is null");
        ev::sv_sample_noalloc(const_cast<ev_sv *>(s), shots, seed, out);
    });
}

int ev_sv_run(const ev_circuit *c, uint64_t seed, const ev_run_options *opt, ev_sv **out_state, uint8_t *classical,
              size_t num_classical, ev_error *err)
{
    return guard(
        [&] {
            NONNULL(c);
            NONNULL(out_state);
            *out_state = nullptr;
            c->c.validate();  // statevector.rs:78
            EV_REQUIRE(num_classical >= c->c.num_classical() && (classical || c->c.num_classical() == 0), EV_E_ARG,
                       "classical buffer smaller than the circuit's classical register");
            EV_REQUIRE(c->c.num_qubits() <= ev::kMaxQubits, EV_E_LIMIT, "n exceeds 34");
            ev_run_options o{};
            if (opt) o = *opt;
            std::vector<uint8_t> bits(c->c.num_classical(), 0);
            ev_sv *s = ev::sv_create(int(c->c.num_qubits()), prec_of(o.precision));
            try {
                ev::Rng rng = ev::Rng::seed_from_u64(seed);
                ev::sv_execute(s, c->c, rng, bits.data(), o);
                ev::sv_sync(s);
            } catch (...) {
                ev::sv_destroy(s);
                throw;
            }
            if (!bits.empty()) std::memcpy(classical, bits.data(), bits.size());
            *out_state = s;
        },
        err);
}

int ev_sv_run_into(ev_sv *s, const ev_circuit *c, uint64_t seed, const ev_run_options *opt, uint8_t *classical,
                   size_t num_classical, ev_error *err)
{
    return guard(
        [&] {
            NONNULL(s);
            NONNULL(c);
            c->c.validate();
            EV_REQUIRE(num_classical >= c->c.num_classical() && (classical || c->c.num_classical() == 0), EV_E_ARG,
                       "classical buffer smaller than the circuit's classical register");
            ev_run_options o{};
            if (opt) o = *opt;
            std::vector<uint8_t> bits(c->c.num_classical(), 0);
            ev::sv_run_into(s, c->c, seed, bits.data(), o);
            ev::sv_sync(s);
            if (!bits.empty()) std::memcpy(classical, bits.data(), bits.size());
        },
        err);
}

int ev_sv_reset(ev_sv *s)
{
    return guard([&] {
        NONNULL(s);
        ev::kern::init_basis(s->psi, s->prec, s->n, 0, s->stream);
        ev::sv_sync(s);
    });
}

// Host-slice shims: src/kernel.rs:76, :192, :299 signatures.
static void with_slice(ev_c128 *amps, size_t len, const std::function<void(ev_sv *)> &f)
{
    NONNULL(amps);
    const int n = qubits_of_len(len);
    ev_sv *s = ev::sv_alloc(n, ev::kern::C128);
    try {
        ev::sv_write(s, C(amps), len);
        f(s);
        ev::sv_read(s, reinterpret_cast<ev::c128 *>(amps), len);
    } catch (...) {
        ev::sv_destroy(s);
        throw;
    }
    ev::sv_destroy(s);
}

int ev_kernel_apply_1q(ev_c128 *amps, size_t len, size_t k, const ev_c128 m[4])
{
    return guard([&] {
        NONNULL(m);
        with_slice(amps, len, [&](ev_sv *s) {
            check_q(k, s->n);
            ev::sv_apply1(s, int(k), C(m));
        });
    });
}

int ev_kernel_apply_2q(ev_c128 *amps, size_t len, size_t a, size_t b, const ev_c128 m[16])
{
    return guard([&] {
        NONNULL(m);
        with_slice(amps, len, [&](ev_sv *s) {
            check_q(a, s->n);
            check_q(b, s->n);
            if (a == b) throw ev::Error(EV_E_DUP_QUBIT, "qubit used twice", a);
            ev::sv_apply2(s, int(a), int(b), C(m));
        });
    });
}

int ev_kernel_apply_controlled_1q(ev_c128 *amps, size_t len, const uint64_t *controls, size_t nc, size_t target,
                                  const ev_c128 m[4])
{
    return guard([&] {
        NONNULL(m);
        with_slice(amps, len, [&](ev_sv *s) {
            check_controls(controls, nc, target, s->n);
            ev::sv_ctrl1(s, controls, int(nc), int(target), C(m));
        });
    });
}

// ------------------------------------------------------------ density matrix

int ev_channel_kraus(const ev_channel *ch, ev_c128 *out, size_t cap, size_t *count)
{
    return guard([&] {
        NONNULL(ch);
        NONNULL(count);
        const auto ops = ev::channel_kraus(*ch);
        *count = ops.size();
        EV_REQUIRE(out == nullptr || cap >= ops.size(), EV_E_ARG, "Kraus output buffer too small");
        if (out)
            for (size_t i = 0; i < ops.size(); ++i) std::memcpy(out + 4 * i, ops[i].data(), 4 * sizeof(ev_c128));
    });
}

int ev_dm_create(uint64_t n, ev_dm **out)
{
    return guard([&] {
        NONNULL(out);
        EV_REQUIRE(n <= ev::kMaxQubits / 2, EV_E_LIMIT, "density matrix limited to 17 qubits");
        *out = ev::dm_create(int(n));
        ev::sv_sync((*out)->vec);
    });
}

void ev_dm_destroy(ev_dm *d) { ev::dm_destroy(d); }
uint64_t ev_dm_num_qubits(const ev_dm *d) { return d ? uint64_t(d->n) : 0; }

int ev_dm_read(const ev_dm *d, ev_c128 *host, size_t len)
{
    return guard([&] {
        NONNULL(d);
        NONNULL(host);
        ev::sv_read(d->vec, reinterpret_cast<ev::c128 *>(host), len);
    });
}

int ev_dm_trace(const ev_dm *d, double *out)
{
    return guard([&] {
        NONNULL(d);
        NONNULL(out);
        *out = ev::dm_diag_sum(const_cast<ev_dm *>(d), 0, 0);
    });
}

int ev_dm_purity(const ev_dm *d, double *out)  // density.rs:130-140: sum |rho_ij|^2
{
    return guard([&] {
        NONNULL(d);
        NONNULL(out);
        *out = ev::sv_norm_sqr(d->vec);
    });
}

int ev_dm_probability(const ev_dm *d, uint64_t basis, double *out)
{
    return guard([&] {
        NONNULL(d);
        NONNULL(out);
        const uint64_t dim = uint64_t(1) << d->n;
        EV_REQUIRE(basis < dim, EV_E_ARG, "basis state " + std::to_string(basis) + " out of range");
        ev::c128 e;
        const uint64_t at = (basis << d->n) | basis;
        EV_HIP(hipMemcpyAsync(&e, static_cast<const char *>(d->vec->psi) + at * 16, 16, hipMemcpyDeviceToHost,
                              d->vec->stream));
        ev::sv_sync(d->vec);
        *out = e.re;
    });
}

int ev_dm_partial_trace_1q(const ev_dm *d, uint64_t qubit, ev_c128 out[4])
{
    return guard([&] {
        NONNULL(d);
        NONNULL(out);
        check_q(qubit, d->n);
        ev::dm_partial_trace(const_cast<ev_dm *>(d), int(qubit), reinterpret_cast<ev::c128 *>(out));
    });
}

// density.rs:237-261: Tr(P rho) = sum_{iq,jq} P[iq][jq] * reduced[jq][iq], real part.
int ev_dm_expectation_pauli(const ev_dm *d, uint64_t qubit, char pauli, double *out)
{
    return guard([&] {
        NONNULL(d);
        NONNULL(out);
        check_q(qubit, d->n);
        ev::Mat2 P;
        switch (pauli) {
        case 'I': P = ev::gate::id(); break;
        case 'X': P = ev::gate::x(); break;
        case 'Y': P = ev::gate::y(); break;
        case 'Z': P = ev::gate::z(); break;
        default: throw ev::Error(EV_E_ARG, std::string("unknown Pauli '") + pauli + "'; expected I, X, Y, or Z");
        }
        ev::c128 r[4];
        ev::dm_partial_trace(const_cast<ev_dm *>(d), int(qubit), r);
        ev::c128 acc{0, 0};
        for (int iq = 0; iq < 2; ++iq)
            for (int jq = 0; jq < 2; ++jq) acc = ev::cadd(acc, ev::cmul(P[iq * 2 + jq], r[jq * 2 + iq]));
        *out = acc.re;
    });
}

int ev_dm_run(const ev_circuit *c, const ev_noise_model *noise, uint64_t seed, const ev_run_options *opt, ev_dm **out,
              uint8_t *classical, size_t num_classical, ev_error *err)
{
    return guard(
        [&] {
            NONNULL(c);
            NONNULL(out);
            *out = nullptr;
            c->c.validate();  // density_matrix.rs:102
            EV_REQUIRE(num_classical >= c->c.num_classical() && (classical || c->c.num_classical() == 0), EV_E_ARG,
                       "classical buffer smaller than the circuit's classical register");
            EV_REQUIRE(c->c.num_qubits() <= ev::kMaxQubits / 2, EV_E_LIMIT, "density matrix limited to 17 qubits");
            if (noise) {
                for (const ev_channel *ch : {&noise->after_1q, &noise->after_2q})
                    EV_REQUIRE(ch->kind == EV_CH_CUSTOM || ch->kind == EV_CH_NONE || (ch->p >= 0.0 && ch->p <= 1.0),
                               EV_E_ARG, "noise probability not in [0, 1]");  // noise.rs:62-65
            }
            ev_run_options o{};
            if (opt) o = *opt;
            std::vector<uint8_t> bits(c->c.num_classical(), 0);
            ev_dm *d = ev::dm_create(int(c->c.num_qubits()));
            try {
                ev::Rng rng = ev::Rng::seed_from_u64(seed);
                ev::dm_execute(d, c->c, noise, rng, bits.data(), o);
                ev::sv_sync(d->vec);
            } catch (...) {
                ev::dm_destroy(d);
                throw;
            }
            if (!bits.empty()) std::memcpy(classical, bits.data(), bits.size());
            *out = d;
        },
        err);
}

// --------------------------------------------------------------- profiling

int ev_last_run_stats(ev_run_stats *out)
{
    return guard([&] {
        NONNULL(out);
        const ev::RunStats &s = ev::last_stats();
        out->logical_gates = s.logical_gates;
        out->passes = s.passes;
        out->launches = s.launches;
        out->device_ms = s.device_ms;
        out->pass_ms = s.pass_ms;
        out->pass_launches = s.pass_launches;
        out->plan_ms = s.plan_ms;
        out->pass_flops = s.pass_flops;
    });
}

int ev_batch_run(const ev_circuit *const *circuits, size_t count, const ev_run_options *opt, ev_batch **out,
                 ev_error *err)
{
    return guard(
        [&] {
            NONNULL(circuits);
            NONNULL(out);
            *out = nullptr;
            ev_run_options o{};
            if (opt) o = *opt;
            (void)prec_of(o.precision);
            std::vector<const ev::Circuit *> cs(count);
            for (size_t i = 0; i < count; ++i) {
                NONNULL(circuits[i]);
                cs[i] = &circuits[i]->c;
            }
            *out = ev::batch_run(cs.data(), count, o);
        },
        err);
}

void ev_batch_destroy(ev_batch *b) { ev::batch_free(b); }
size_t ev_batch_count(const ev_batch *b) { return b ? b->n.size() : 0; }
uint64_t ev_batch_num_qubits(const ev_batch *b, size_t i) { return (b && i < b->n.size()) ? uint64_t(b->n[i]) : 0; }

int ev_batch_read(const ev_batch *b, size_t i, uint64_t offset, ev_c128 *host, size_t len)
{
    return guard([&] {
        NONNULL(b);
        EV_REQUIRE(len == 0 || host != nullptr, EV_E_ARG, "host is null");
        ev::batch_read(const_cast<ev_batch *>(b), i, offset, reinterpret_cast<ev::c128 *>(host), len);
    });
}

int ev_batch_norms(const ev_batch *b, double *out)
{
    return guard([&] {
        NONNULL(b);
        NONNULL(out);
        ev::batch_norms(const_cast<ev_batch *>(b), out);
    });
}

int ev_batch_compare(const ev_batch *b, const ev_batch *ref, double *fidelity, double *max_abs)
{
    return guard([&] {
        NONNULL(b);
        NONNULL(ref);
        NONNULL(fidelity);
        NONNULL(max_abs);
        ev::batch_compare(const_cast<ev_batch *>(b), const_cast<ev_batch *>(ref), fidelity, max_abs);
    });
}

int ev_batch_get_stats(const ev_batch *b, ev_batch_stats *out)
{
    return guard([&] {
        NONNULL(b);
        NONNULL(out);
        *out = b->stats;
    });
}

int ev_server_create(const ev_server_config *cfg, ev_server **out)
{
    return guard([&] {
        NONNULL(out);
        ev_server_config c{};
        if (cfg) c = *cfg;
        (void)prec_of(c.precision);
        *out = ev::server_create(c);
    });
}

void ev_server_destroy(ev_server *s) { ev::server_destroy(s); }

int ev_server_submit(ev_server *s, const ev_circuit *c, uint64_t shots, uint64_t seed, uint64_t *request_id)
{
    return guard([&] {
        NONNULL(s);
        NONNULL(c);
        NONNULL(request_id);
        *request_id = ev::server_submit(s, c->c, shots, seed);
    });
}

int ev_server_submit_qasm(ev_server *s, const char *src, size_t len, uint64_t shots, uint64_t seed,
                          uint64_t *request_id, ev_error *err)
{
    return guard(
        [&] {
            NONNULL(s);
            NONNULL(request_id);
            EV_REQUIRE(src != nullptr || len == 0, EV_E_ARG, "src is null");
            const ev::Circuit c = ev::parse_qasm(src ? src : "", len);
            *request_id = ev::server_submit(s, c, shots, seed);
        },
        err);
}

int ev_server_poll(ev_server *s, uint64_t request_id, int *done)
{
    return guard([&] {
        NONNULL(s);
        NONNULL(done);
        *done = ev::server_poll(s, request_id) ? 1 : 0;
    });
}

int ev_server_wait(ev_server *s, uint64_t request_id, uint64_t *samples, ev_request_info *info)
{
    return guard([&] {
        NONNULL(s);
        const ev_request_info r = ev::server_wait(s, request_id, samples);
        if (info) *info = r;
    });
}

int ev_server_get_stats(ev_server *s, ev_server_stats *out)
{
    return guard([&] {
        NONNULL(s);
        NONNULL(out);
        *out = ev::server_stats(s);
    });
}

int ev_autotune(uint64_t n, ev_precision p, uint32_t depth, ev_tune_entry *out)
{
    return guard([&] {
        (void)prec_of(p);
        const ev_tune_entry e = ev::autotune(int(n), p == EV_C64, int(depth));
        if (out) *out = e;
    });
}

int ev_tune_get(uint64_t n, ev_precision p, ev_tune_entry *out, int *found)
{
    return guard([&] {
        NONNULL(out);
        NONNULL(found);
        (void)prec_of(p);
        *found = ev::tune_lookup(int(n), p == EV_C64, out) ? 1 : 0;
    });
}

int ev_tune_set(uint64_t n, ev_precision p, const ev_tune_entry *e)
{
    return guard([&] {
        NONNULL(e);
        (void)prec_of(p);
        ev::tune_store(int(n), p == EV_C64, *e);
    });
}

int ev_compute_rates_measure(ev_precision p, ev_compute_rates *out)
{
    return guard([&] {
        NONNULL(out);
        const bool f32 = prec_of(p) == ev::kern::C64;
        void *scr = nullptr;
        hipStream_t st = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto cleanup = [&] {
            if (st) (void)hipStreamSynchronize(st);
            if (scr) (void)hipFree(scr);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (st) (void)hipStreamDestroy(st);
        };
        try {
            EV_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            EV_HIP(hipMalloc(&scr, ev::kern::compute_rates_scratch_bytes()));
            EV_HIP(hipEventCreate(&e0));
            EV_HIP(hipEventCreate(&e1));
            const ev::kern::ComputeRates r = ev::kern::compute_rates(f32, scr, st, e0, e1);
            out->valu_tflops = r.valu_tflops;
            out->mfma_tflops = r.mfma_tflops;
            out->mixed_tflops = r.mixed_tflops;
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    });
}

int ev_copy_bandwidth(size_t bytes, uint32_t reps, double *gbps)
{
    return guard([&] {
        NONNULL(gbps);
        EV_REQUIRE(bytes >= 32 && bytes % 32 == 0 && reps > 0, EV_E_ARG, "bytes must be a positive multiple of 32");
        void *a = nullptr, *b = nullptr;
        hipStream_t st = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto cleanup = [&] {
            if (st) (void)hipStreamSynchronize(st);
            if (a) (void)hipFree(a);
            if (b) (void)hipFree(b);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (st) (void)hipStreamDestroy(st);
        };
        try {
            EV_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            EV_HIP(hipMalloc(&a, bytes));
            EV_HIP(hipMalloc(&b, bytes));
            EV_HIP(hipEventCreate(&e0));
            EV_HIP(hipEventCreate(&e1));
            ev::kern::copy(a, b, bytes, st);  // warm-up (also touches both buffers)
            EV_HIP(hipEventRecord(e0, st));
            for (uint32_t r = 0; r < reps; ++r) ev::kern::copy(a, b, bytes, st);
            EV_HIP(hipEventRecord(e1, st));
            EV_HIP(hipEventSynchronize(e1));
            float ms = 0;
            EV_HIP(hipEventElapsedTime(&ms, e0, e1));
            *gbps = 2.0 * double(bytes) * reps / (double(ms) * 1e-3) / 1e9;
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    });
}

int ev_debug_plan(const ev_circuit *c, const ev_run_options *opt, void *passes, void *stages, void *ops,
                  ev_c128 *coef, ev_plan_info *info)
{
    return guard([&] {
        NONNULL(c);
        NONNULL(info);
        c->c.validate();
        ev_run_options o{};
        if (opt) o = *opt;
        const ev::PlanBlob b = ev::debug_plan(c->c, o);
        info->num_passes = b.np;
        info->num_stages = b.ns;
        info->num_ops = b.no;
        info->num_coef = uint32_t(b.coef.size());
        info->pass_bytes = sizeof(ev::fused::Pass);
        info->stage_bytes = sizeof(ev::fused::Stage);
        info->op_bytes = sizeof(ev::fused::DevOp);
        info->tile_qubits = b.t;
        EV_REQUIRE(b.swap_tail.size() <= 32, EV_E_LIMIT, "swap tail longer than ev_plan_info holds");
        info->num_swap_pairs = uint32_t(b.swap_tail.size());
        for (size_t k = 0; k < b.swap_tail.size(); ++k) {
            info->swap_a[k] = uint32_t(b.swap_tail[k].first);
            info->swap_b[k] = uint32_t(b.swap_tail[k].second);
        }
        if (passes && !b.passes.empty()) std::memcpy(passes, b.passes.data(), b.passes.size());
        if (stages && !b.stages.empty()) std::memcpy(stages, b.stages.data(), b.stages.size());
        if (ops && !b.ops.empty()) std::memcpy(ops, b.ops.data(), b.ops.size());
        if (coef && !b.coef.empty()) std::memcpy(coef, b.coef.data(), b.coef.size() * sizeof(ev_c128));
    });
}

int ev_last_run_kernel_stats(ev_kernel_stat *out, size_t cap, size_t *count)
{
    return guard([&] {
        NONNULL(count);
        const auto &ks = ev::last_stats().kernels;
        *count = ks.size();
        EV_REQUIRE(out == nullptr || cap >= ks.size(), EV_E_ARG, "kernel stat buffer too small");
        if (!out) return;
        for (size_t i = 0; i < ks.size(); ++i) {
            std::memset(&out[i], 0, sizeof out[i]);
            std::strncpy(out[i].name, ks[i].name.c_str(), sizeof out[i].name - 1);
            out[i].launches = ks[i].launches;
            out[i].total_ms = ks[i].total_ms;
            out[i].bytes = ks[i].bytes;
        }
    });
}

}  // extern "C"
