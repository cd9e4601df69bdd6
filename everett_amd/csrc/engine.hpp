// Device-resident state handles and the circuit executors behind the C ABI.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <string>
#include <vector>

#include "circuit.hpp"
#include "common.hpp"
#include "kernels.hpp"

// The C ABI's opaque statevector: src/state.rs:23-29 State { amps, n }, with
// the amplitudes in HBM.
struct ev_sv {
    int n = 0;
    void *plan_buf = nullptr;       // device ring of fused pass descriptor blocks
    void *plan_host = nullptr;      // its pinned host staging (same offsets)
    size_t plan_cap = 0;            // bytes in each
    size_t plan_off = 0;            // next free offset in the ring
    ev::kern::Prec prec = ev::kern::C128;
    void *psi = nullptr;            // 2^n amplitudes
    hipStream_t stream = nullptr;   // all work on this handle is ordered here
    int device = 0;                 // the device current when the handle was made
    double *d_red = nullptr;        // reduction partials + result slots
    double *h_red = nullptr;        // pinned host copy of the result slots
    uint64_t len() const { return uint64_t(1) << n; }
    size_t bytes() const { return size_t(len()) * ev::kern::amp_bytes(prec); }
};

// The C ABI's opaque density matrix (src/density.rs:31-36): a row-major
// 2^n x 2^n c128 matrix, i.e. a 2n-qubit vector whose bit n+t is row qubit t
// and bit t is column qubit t.
struct ev_dm {
    ev_sv *vec = nullptr;  // 2n-qubit c128 vector
    int n = 0;
};

namespace ev {

// Reduction scratch layout inside ev_sv::d_red.
constexpr int kRedPartials = 8 * kern::kReduceBlocks;
constexpr int kRedSlots = 8;

ev_sv *sv_alloc(int n, kern::Prec p);        // uninitialised amplitudes
ev_sv *sv_create(int n, kern::Prec p);       // |0..0>
void sv_destroy(ev_sv *s);
void sv_sync(ev_sv *s);

double sv_norm_sqr(ev_sv *s);
double sv_p1(ev_sv *s, int q);
c128 sv_overlap(ev_sv *s, ev_sv *other);
void sv_bloch(ev_sv *s, int q, double out[4]);
// src/measure.rs:15-51 with the host-drawn uniform u.
bool sv_measure(ev_sv *s, int q, double u);
void sv_normalize(ev_sv *s);
void sv_read(ev_sv *s, c128 *host, uint64_t len);
void sv_read_range(ev_sv *s, uint64_t offset, c128 *host, uint64_t len);
void sv_write(ev_sv *s, const c128 *host, uint64_t len);
void sv_sample(ev_sv *s, uint64_t shots, uint64_t seed, uint64_t *out);
void sv_sample_noalloc(ev_sv *s, uint64_t shots, uint64_t seed, uint64_t *out);

// Per-op device dispatch (src/backend/statevector.rs:91-107).
void sv_apply1(ev_sv *s, int k, const c128 m[4]);
void sv_apply2(ev_sv *s, int a, int b, const c128 m[16]);
void sv_ctrl1(ev_sv *s, const uint64_t *controls, int nc, int target, const c128 m[4]);

struct KernelStat {
    std::string name;
    uint64_t launches = 0;
    double total_ms = 0, bytes = 0;
};

// Optional per-launch timing with HIP events on the handle's stream.
class Profiler {
  public:
    Profiler(hipStream_t st, bool on) : st_(st), on_(on) {}
    ~Profiler();
    bool on() const { return on_; }
    void begin();
    void end(const char *name, double bytes);
    // Synchronises and folds the recorded launches into per-name totals.
    std::vector<KernelStat> collect();

  private:
    struct Rec { hipEvent_t a, b; const char *name; double bytes; };
    hipStream_t st_;
    bool on_;
    hipEvent_t open_ = nullptr;
    std::vector<Rec> recs_;
};

struct RunStats {
    uint64_t logical_gates = 0, passes = 0, launches = 0, pass_launches = 0;
    double device_ms = 0, pass_ms = 0, plan_ms = 0;
    double pass_flops = 0;  // algorithmic flops of the fused pass launches
    std::vector<KernelStat> kernels;
};
RunStats &last_stats();

// Algorithmic HBM bytes of one launch (reads + writes of the amplitudes the
// launch must touch) for a 2^n state of s-byte amplitudes.
inline double full_pass_bytes(int n, size_t s) { return 2.0 * double(uint64_t(1) << n) * double(s); }
const char *k1_name(K1 k);
const char *k2_name(K2 k);
double k1_bytes(K1 k, int n, size_t s);
double k2_bytes(K2 k, int n, size_t s);
inline double ctrl_bytes(int n, int nc, size_t s) { return 2.0 * double(uint64_t(1) << (n - nc)) * double(s); }

// backend/mod.rs:51-72 for record i with one launch per gate; returns the
// index after the op (an If record consumes the op it guards).
size_t sv_plain_op(ev_sv *s, const Circuit &c, size_t i, Rng &rng, uint8_t *classical, bool execute, RunStats &st,
                   Profiler &prof);

// Resets s to |0..0> and drives c on it (validation is the caller's job).
void sv_run_into(ev_sv *s, const Circuit &c, uint64_t seed, uint8_t *classical, const ev_run_options &opt);

// src/backend/statevector.rs:77-88 + backend/mod.rs:43-72: drive the ops on
// the device, one launch per gate (plain) or through fused passes.
void sv_execute(ev_sv *s, const Circuit &c, Rng &rng, uint8_t *classical, const ev_run_options &opt);

// src/backend/density_matrix.rs:97-113 on the device.
void dm_execute(ev_dm *d, const Circuit &c, const ev_noise_model *noise, Rng &rng, uint8_t *classical,
                const ev_run_options &opt);
ev_dm *dm_create(int n);
void dm_destroy(ev_dm *d);

}  // namespace ev
