/*
 * everett.h -- C ABI of the MI355X-native everett engine (libeverett.so).
 *
 * This is the drop-in boundary for hz2/everett 0.1.2's simulation paths.  Each
 * entry point names the reference interface it replaces (paths relative to the
 * crate root).  A Rust crate binds these with `extern "C"` declarations of the
 * same names; the Python package everett_amd binds them with ctypes.
 *
 * Conventions
 *  - Every function returning int returns EV_OK (0) or a negative ev_status.
 *    The thread's last error text is available from ev_last_error().
 *  - Opaque handles own their device memory; host arrays are borrowed for the
 *    duration of the call only.  A handle is not thread-safe (it mirrors the
 *    reference's `&mut self`); distinct handles may be used from distinct
 *    threads.
 *  - The measurement RNG never runs on the device: the host draws uniforms
 *    from the reference stream (src/rng.rs) in the reference's order.
 *  - Qubit 0 is the least-significant bit of a basis index (src/qubit.rs:11-13).
 */
#ifndef EVERETT_H
#define EVERETT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ types */

/* src/complex.rs:22-29: #[repr(C)] Complex64 { re: f64, im: f64 } */
typedef struct { double re, im; } ev_c128;

/* Storage precision of a state.  EV_C128 is the reference's Complex64 (two
 * f64).  EV_C64 stores two f32 per amplitude and keeps every norm,
 * probability and sampling reduction in f64. */
typedef enum { EV_C128 = 0, EV_C64 = 1 } ev_precision;

typedef enum {
    EV_OK = 0,
    EV_E_QUBIT_RANGE = -1, /* Error::QubitOutOfRange        src/error.rs:12-17  a=qubit b=num_qubits */
    EV_E_CBIT_RANGE = -2,  /* Error::ClassicalBitOutOfRange src/error.rs:19-24  a=bit   b=num_classical */
    EV_E_DUP_QUBIT = -3,   /* Error::DuplicateQubit         src/error.rs:26-29  a=qubit */
    EV_E_DIM = -4,         /* Error::DimensionMismatch      src/error.rs:32-37  a=len   b=expected */
    EV_E_ARG = -5,         /* malformed argument (null pointer, unknown op kind, dangling If prefix, ...) */
    EV_E_OOM = -6,         /* device memory exhausted */
    EV_E_HIP = -7,         /* HIP runtime error */
    EV_E_ADMISSION = -8,   /* serving: the request can never fit the configured memory */
    EV_E_LIMIT = -9,       /* beyond a documented size limit of this build */
    EV_E_STATE = -10,      /* the handle is not in a state that allows the call */
    EV_E_QASM = -11        /* Error::Qasm                   src/error.rs:44-52  a=line  b=col (1-based; 0 = n/a) */
} ev_status;

/* Structured error (mirrors src/error.rs:8-53). */
typedef struct {
    int32_t code;
    uint64_t a, b;
    char message[256];
} ev_error;

/* Text of the last error raised on the calling thread ("" if none). */
const char *ev_last_error(void);

/* ------------------------------------------------------------------- ops */

/* src/op.rs:14-58.  Op::IfClassic { bit, then } is encoded as an EV_OP_IF
 * record (target = classical bit) guarding the record that follows it; nested
 * IfClassic is a run of EV_OP_IF records.
 *
 * Op::Controlled's `controls: Vec<QubitId>` (src/op.rs:34-36) has no length
 * limit, and neither has this encoding: an EV_OP_CONTROLLED record's
 * num_controls is the op's total count, its controls[] holds the first
 * EV_MAX_CONTROLS of them, and the rest follow in EV_OP_CONTROLS records
 * (EV_MAX_CONTROLS per record, num_controls = entries used in that record)
 * directly after it.  A guarded controlled op is EV_OP_IF, EV_OP_CONTROLLED,
 * then its EV_OP_CONTROLS records.  Repeated controls are allowed, as in the
 * reference: they form one condition. */
typedef enum {
    EV_OP_APPLY1 = 1,     /* Op::Apply1 { gate, target }               m[0..4)  */
    EV_OP_APPLY2 = 2,     /* Op::Apply2 { gate, a=target, b }          m[0..16) */
    EV_OP_CONTROLLED = 3, /* Op::Controlled { controls, gate, target } m[0..4)  */
    EV_OP_MEASURE = 4,    /* Op::Measure { qubit=target, into=b } */
    EV_OP_IF = 5,         /* Op::IfClassic { bit=target, then=<next record> } */
    EV_OP_CONTROLS = 6    /* more controls of the EV_OP_CONTROLLED record before it */
} ev_op_kind;

/* Controls carried per record (not a limit on an op's controls). */
#define EV_MAX_CONTROLS 16

typedef struct {
    uint32_t kind;         /* ev_op_kind */
    uint32_t num_controls; /* EV_OP_CONTROLLED: the op's controls in all; EV_OP_CONTROLS: entries in this record */
    uint64_t target;
    uint64_t b;
    uint64_t controls[EV_MAX_CONTROLS];
    ev_c128 m[16];         /* row-major gate matrix (src/gate.rs:21-35) */
} ev_op;

/* --------------------------------------------------------------- circuit */

/* src/circuit.rs:25-30 Circuit { ops, num_qubits, num_classical } */
typedef struct ev_circuit ev_circuit;

int ev_circuit_create(uint64_t num_qubits, uint64_t num_classical, ev_circuit **out); /* circuit.rs:35-52 */
void ev_circuit_destroy(ev_circuit *c);
int ev_circuit_append(ev_circuit *c, const ev_op *ops, size_t count);  /* circuit.rs:91-116, 200-236 */
int ev_circuit_compose(ev_circuit *dst, const ev_circuit *src);        /* circuit.rs:124-127 */
size_t ev_circuit_len(const ev_circuit *c);                            /* records, If and control continuations included */
uint64_t ev_circuit_num_qubits(const ev_circuit *c);                   /* circuit.rs:55-58 */
uint64_t ev_circuit_num_classical(const ev_circuit *c);                /* circuit.rs:62-65 */
int ev_circuit_export(const ev_circuit *c, ev_op *out, size_t cap, size_t *len);
int ev_circuit_validate(const ev_circuit *c, ev_error *err);           /* circuit.rs:245-306 */

/* OpenQASM 3.0 text (the reference's subset: stdgates h x y z s t id rx ry
 * rz p, cx cz swap, ctrl(k) @, both measure forms, if (c[i] == 1)).
 * ev_circuit_from_qasm replaces Circuit::from_qasm (src/circuit.rs:324-326,
 * src/qasm/parse.rs:811-819): on a parse error it returns EV_E_QASM with
 * err->a = line, err->b = col and err->message = "OpenQASM 3 error at
 * <line>:<col>: <message>" (src/error.rs:82-84).  ev_circuit_to_qasm replaces
 * Circuit::to_qasm (src/circuit.rs:315-317, src/qasm/emit.rs:17-37): it
 * writes the text and a NUL into buf and its length (without the NUL) into
 * *len; EV_E_LIMIT (with *len set) when cap is too small, EV_E_QASM for a
 * gate with no OpenQASM 3 name or a nested IfClassic. */
int ev_circuit_from_qasm(const char *src, size_t len, ev_circuit **out, ev_error *err);
int ev_circuit_to_qasm(const ev_circuit *c, char *buf, size_t cap, size_t *len, ev_error *err);

/* src/rng.rs:9-56 (SplitMix64-seeded xoshiro256++), host only. */
typedef struct { uint64_t s[4]; } ev_rng;
void ev_rng_seed(ev_rng *r, uint64_t seed);
uint64_t ev_rng_next_u64(ev_rng *r);
double ev_rng_next_f64(ev_rng *r);

/* Circuit generators.  ev_circuit_qft follows src/algorithms/qft.rs:54-62;
 * the others define the benchmark workloads (docs/configs.md in the
 * repository gives the exact draw order). */
int ev_circuit_qft(uint64_t n, ev_circuit **out);
int ev_circuit_iqft(uint64_t n, ev_circuit **out);                      /* qft.rs:81-91 */
int ev_circuit_brickwall(uint64_t n, uint64_t depth, ev_rng *rng, ev_circuit **out);
int ev_circuit_dm_random(uint64_t n, uint64_t depth, ev_rng *rng, ev_circuit **out);

/* ----------------------------------------------------------- statevector */

/* Device-resident statevector: src/state.rs:23-29 State { amps, n }. */
typedef struct ev_sv ev_sv;

int ev_sv_create(uint64_t n, ev_precision p, ev_sv **out);              /* state.rs:38-44 State::zero */
int ev_sv_from_amplitudes(const ev_c128 *amps, size_t len, ev_precision p, ev_sv **out); /* state.rs:51-62 */
void ev_sv_destroy(ev_sv *s);
uint64_t ev_sv_num_qubits(const ev_sv *s);                              /* state.rs:66-70 */
ev_precision ev_sv_precision(const ev_sv *s);
int ev_sv_read(const ev_sv *s, ev_c128 *host, size_t len);              /* state.rs:81-83 amplitudes() (c64 widened) */
int ev_sv_write(ev_sv *s, const ev_c128 *host, size_t len);
/* Copies amplitudes [offset, offset + len) to the host (c64 widened). */
int ev_sv_read_range(const ev_sv *s, uint64_t offset, ev_c128 *host, size_t len);
int ev_sv_copy(const ev_sv *src, ev_sv **out);                          /* State: Clone */

/* Per-gate device kernels (src/kernel.rs:76, :192, :299).  Preconditions are
 * checked and reported as EV_E_QUBIT_RANGE / EV_E_DUP_QUBIT / EV_E_ARG instead
 * of the reference's debug_assert!. */
int ev_sv_apply_1q(ev_sv *s, uint64_t k, const ev_c128 m[4]);
int ev_sv_apply_2q(ev_sv *s, uint64_t a, uint64_t b, const ev_c128 m[16]);
int ev_sv_apply_controlled_1q(ev_sv *s, const uint64_t *controls, size_t num_controls, uint64_t target,
                              const ev_c128 m[4]);
/* src/measure.rs:15-51 with the uniform `u` drawn by the caller (Rng::next_f64). */
int ev_sv_measure(ev_sv *s, uint64_t qubit, double u, int *outcome);

/* FP64 observables (src/state.rs:94-185). */
int ev_sv_norm_sqr(const ev_sv *s, double *out);                        /* state.rs:94-96 */
int ev_sv_normalize(ev_sv *s);                                          /* state.rs:103-111 */
int ev_sv_probability(const ev_sv *s, uint64_t basis, double *out);     /* state.rs:119-122 */
int ev_sv_overlap(const ev_sv *s, const ev_sv *other, ev_c128 *out);    /* state.rs:130-138 */
int ev_sv_fidelity(const ev_sv *s, const ev_sv *other, double *out);    /* state.rs:149-151 */
int ev_sv_bloch_vector(const ev_sv *s, uint64_t qubit, double out[3]);  /* state.rs:163-185 */
/* Born-rule sampling: `shots` basis indices, one Rng::next_f64 per shot from
 * Rng::seed_from_u64(seed), inverted through the f64 inclusive CDF of |amp|^2. */
int ev_sv_sample(const ev_sv *s, uint64_t shots, uint64_t seed, uint64_t *out);
/* ev_sv_sample without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  Same samples as
 * ev_sv_sample.  The chunk totals, uniforms and indices live in a fixed
 * per-GPU scratch of 1.5 MiB (held until the process ends); `out` is also
 * the host staging buffer for the uniforms.  At most
 * EV_SAMPLE_NOALLOC_MAX_SHOTS shots per call (EV_E_LIMIT beyond); the state's
 * device must be current (EV_E_STATE otherwise).  Every call on one GPU
 * shares that GPU's scratch, so calls that overlap on one GPU race; the call
 * returns after its stream has synchronised. */
#define EV_SAMPLE_NOALLOC_MAX_SHOTS 65536
int ev_sv_sample_noalloc(const ev_sv *s, uint64_t shots, uint64_t seed, uint64_t *out);

/* Execution options (not in the reference; defaults = zero-initialised). */
typedef struct {
    ev_precision precision; /* storage precision of the run */
    uint32_t plain;         /* 1: one launch per op (unfused); 0: planned fused passes */
    uint32_t tile_qubits;   /* fused pass tile width, 8..13 (13 = 512-thread workgroups); 0 = default */
    uint32_t max_gates_per_pass; /* 0 = default */
    uint32_t profile;       /* 1: time every kernel launch with HIP events (ev_last_run_kernel_stats) */
    uint32_t use_graph;     /* batch: 1 = capture each pipelined group's rounds in a HIP graph and
                               launch it once */
    uint32_t renormalize;   /* batch: rescale every state to unit f64 norm at the end (c64 drift) */
    uint32_t low_qubits;    /* lowest qubits always in a tile (contiguous rows); 0 = tuned or default */
    uint32_t mfma_u2;       /* complex64 single runs: 1 = dense 4x4 gates on the matrix cores
                               (v_mfma_f32_4x4x1_16b_f32) instead of VALU FMA chains; 0 = VALU.
                               Slower on MI355X at 30 qubits: the kernel then holds 116 VGPRs
                               (4 waves per SIMD instead of 6) */
} ev_run_options;

/* src/backend/statevector.rs:68-88: validate, |0..0>, Rng::seed_from_u64(seed),
 * drive the ops.  On success *out_state owns the final state and classical[]
 * (num_classical bytes, 0/1) holds the classical register.  On a validation
 * failure *err (may be NULL) receives the first error found. */
int ev_sv_run(const ev_circuit *c, uint64_t seed, const ev_run_options *opt, ev_sv **out_state,
              uint8_t *classical, size_t num_classical, ev_error *err);

/* Same as ev_sv_run but on an existing state, which is first reset to
 * |0..0> (State::zero without a new allocation). */
int ev_sv_run_into(ev_sv *s, const ev_circuit *c, uint64_t seed, const ev_run_options *opt, uint8_t *classical,
                   size_t num_classical, ev_error *err);
/* Resets s to |0..0> in place. */
int ev_sv_reset(ev_sv *s);

/* Host-slice shims with the exact signatures of src/kernel.rs's pub fns.
 * They copy `amps` to the device, apply one gate and copy back. */
int ev_kernel_apply_1q(ev_c128 *amps, size_t len, size_t k, const ev_c128 m[4]);
int ev_kernel_apply_2q(ev_c128 *amps, size_t len, size_t a, size_t b, const ev_c128 m[16]);
int ev_kernel_apply_controlled_1q(ev_c128 *amps, size_t len, const uint64_t *controls, size_t num_controls,
                                  size_t target, const ev_c128 m[4]);

/* ------------------------------------------------------- density matrix */

/* src/noise.rs:160-174 Channel; kraus/num_kraus are used by EV_CH_CUSTOM
 * (2x2 row-major operators, 4 entries each). */
typedef enum {
    EV_CH_NONE = 0,
    EV_CH_DEPOLARIZING = 1,
    EV_CH_AMPLITUDE_DAMPING = 2,
    EV_CH_PHASE_DAMPING = 3,
    EV_CH_BIT_FLIP = 4,
    EV_CH_DEPHASING = 5,
    EV_CH_CUSTOM = 6
} ev_channel_kind;

typedef struct {
    int32_t kind;
    double p;
    const ev_c128 *kraus;
    size_t num_kraus;
} ev_channel;

/* src/noise.rs:31-40 NoiseModel { after_1q, after_2q, readout_error } */
typedef struct {
    ev_channel after_1q;
    ev_channel after_2q;
    double readout_error;
} ev_noise_model;

/* src/noise.rs:179-234 Channel::kraus_ops: writes up to cap operators. */
int ev_channel_kraus(const ev_channel *ch, ev_c128 *out, size_t cap, size_t *count);

/* Device-resident density matrix (src/density.rs:31-36), row-major dim x dim. */
typedef struct ev_dm ev_dm;
int ev_dm_create(uint64_t n, ev_dm **out);                              /* density.rs:45-55 */
void ev_dm_destroy(ev_dm *d);
uint64_t ev_dm_num_qubits(const ev_dm *d);
int ev_dm_read(const ev_dm *d, ev_c128 *host, size_t len);              /* density.rs:101-103 */
int ev_dm_trace(const ev_dm *d, double *out);                           /* density.rs:121-124 */
int ev_dm_purity(const ev_dm *d, double *out);                          /* density.rs:130-140 */
int ev_dm_probability(const ev_dm *d, uint64_t basis, double *out);     /* density.rs:150-158 */
int ev_dm_expectation_pauli(const ev_dm *d, uint64_t qubit, char pauli, double *out); /* density.rs:237-261 */
int ev_dm_partial_trace_1q(const ev_dm *d, uint64_t qubit, ev_c128 out[4]);         /* density.rs:270-288 */

/* src/backend/density_matrix.rs:97-113 run_with_noise_seeded.  noise may be
 * NULL (NoiseModel::ideal, density_matrix.rs:76-78). */
int ev_dm_run(const ev_circuit *c, const ev_noise_model *noise, uint64_t seed, const ev_run_options *opt,
              ev_dm **out, uint8_t *classical, size_t num_classical, ev_error *err);

/* ------------------------------------------------------------------ batch */

/* Many independent measurement-free circuits (8..30 qubits each), each from
 * |0..0>, with all statevectors resident in one HBM arena.  Every circuit is
 * planned into fused passes.  Circuits run in groups of consecutive indices
 * (a small first group, then about an eighth of the amplitudes each),
 * pipelined: the host plans group g+1 while the device runs group g.  Round
 * r of a group is one launch over the r-th pass of every circuit of the
 * group that has one (no reference counterpart: the reference runs
 * one circuit per call, src/backend/statevector.rs:77-88). */
typedef struct ev_batch ev_batch;
typedef struct {
    uint64_t circuits, logical_gates, rounds, passes, tiles;
    double plan_ms, upload_ms, device_ms, renorm_ms; /* renorm_ms: device time of the f64 renormalisation */
    uint32_t graph;  /* 1 if the groups' rounds ran as captured HIP graphs (one per group) */
    uint32_t groups; /* pipelined groups: the host plans group g+1 while the device runs group g */
    double pass_bytes; /* algorithmic HBM bytes of all fused passes: per pass, every amplitude of its
                          state read and written once (2 * 2^n * amplitude size) */
    double pass_flops; /* algorithmic floating-point operations of all fused passes (an FMA counts 2) */
} ev_batch_stats;
int ev_batch_run(const ev_circuit *const *circuits, size_t count, const ev_run_options *opt, ev_batch **out,
                 ev_error *err);
void ev_batch_destroy(ev_batch *b);
size_t ev_batch_count(const ev_batch *b);
uint64_t ev_batch_num_qubits(const ev_batch *b, size_t i);
int ev_batch_read(const ev_batch *b, size_t i, uint64_t offset, ev_c128 *host, size_t len);
/* f64 <psi_i|psi_i> for every state. */
int ev_batch_norms(const ev_batch *b, double *out);
/* Per state i: fidelity |<ref_i|b_i>|^2 / (<b_i|b_i><ref_i|ref_i>) and max-abs
 * |b_i - ref_i|, against a c128 batch of the same circuits. */
int ev_batch_compare(const ev_batch *b, const ev_batch *ref, double *fidelity, double *max_abs);
int ev_batch_get_stats(const ev_batch *b, ev_batch_stats *out);

/* ---------------------------------------------------------------- serving */

/* A continuous-batching server (no reference counterpart).  Requests are
 * measurement-free circuits ending in `shots` Born-rule samples.  A
 * scheduler thread admits queued requests in arrival order while their
 * statevectors fit the HBM budget (2^n * 8 B for c64, 2^n * 16 B for c128),
 * then advances every admitted request by one fused pass per launch (one
 * workgroup per tile of every request's next pass), so requests of
 * different sizes overlap on the device.  A request
 * that completes is sampled (f64 CDF, one Rng::next_f64 per shot from
 * Rng::seed_from_u64(seed)) and its memory is released.  Submit, poll and
 * wait are thread-safe. */
typedef struct ev_server ev_server;
typedef struct {
    ev_precision precision;
    uint32_t max_active;        /* admitted requests at most (0 = no limit) */
    uint64_t memory_bytes;      /* statevector budget (0 = 80% of free device memory) */
    uint32_t tile_qubits;       /* 0 = default */
    uint32_t max_gates_per_pass;/* 0 = default */
    uint32_t reserved[4];
} ev_server_config;
typedef struct {
    uint64_t id, num_qubits, gates, passes, shots;
    double submit_ms, admit_ms, done_ms; /* server clock (ms since creation); submit_ms is taken when
                                            ev_server_submit is entered, before the circuit is planned */
    int32_t status;                      /* EV_OK, or the error that ended the request */
    int32_t pad;
} ev_request_info;
typedef struct {
    uint64_t submitted, completed, failed, steps, active, queued;
    uint64_t bytes_in_use, bytes_peak;
    double busy_ms;   /* device time of the step launches */
} ev_server_stats;
int ev_server_create(const ev_server_config *cfg, ev_server **out);
void ev_server_destroy(ev_server *s);   /* drains in-flight work, then frees */
/* Validates and copies the circuit (planner threads plan the copy); returns
 * EV_E_ADMISSION if it can never fit the budget.  A planning failure ends the
 * request with that status. */
int ev_server_submit(ev_server *s, const ev_circuit *c, uint64_t shots, uint64_t seed, uint64_t *request_id);
/* Serving wire format: parses OpenQASM 3 text as ev_circuit_from_qasm does
 * and submits the circuit as ev_server_submit does; a parse error returns
 * EV_E_QASM before anything is queued.  (The reference has no server; its
 * text entry point is Circuit::from_qasm, src/circuit.rs:324-326.) */
int ev_server_submit_qasm(ev_server *s, const char *src, size_t len, uint64_t shots, uint64_t seed,
                          uint64_t *request_id, ev_error *err);
/* *done = 1 once the request has completed (or failed). */
int ev_server_poll(ev_server *s, uint64_t request_id, int *done);
/* Blocks until the request completes; copies its samples (shots entries) and
 * info, then forgets the request. */
int ev_server_wait(ev_server *s, uint64_t request_id, uint64_t *samples, ev_request_info *info);
int ev_server_get_stats(ev_server *s, ev_server_stats *out);

/* ------------------------------------------------------------- autotuner */

/* Fused-pass parameters for one state size: tile width, contiguous low
 * qubits per tile and gate budget per pass, with the measured cost. */
typedef struct {
    uint32_t tile_qubits, low_qubits, max_gates_per_pass, pad;
    double ms_per_gate;
} ev_tune_entry;
/* Times the fused path on a synthetic brick-wall circuit of n qubits (depth
 * `depth`, 0 = 8) followed by QFT(n) over a grid of tile widths, low-qubit
 * counts and gate budgets, and records the fastest for (n, precision): runs
 * that leave tile_qubits, low_qubits and max_gates_per_pass at 0 use it. */
int ev_autotune(uint64_t n, ev_precision p, uint32_t depth, ev_tune_entry *out);
/* Reads or sets the recorded entry (e.g. restored from a file). */
int ev_tune_get(uint64_t n, ev_precision p, ev_tune_entry *out, int *found);
int ev_tune_set(uint64_t n, ev_precision p, const ev_tune_entry *e);

/* Measured HBM copy bandwidth: a plain read-once/write-once kernel over two
 * device buffers of `bytes` each, `reps` launches timed with HIP events.
 * Reports read + write bytes per second (the achievable-peak reference for
 * the roofline). */
int ev_copy_bandwidth(size_t bytes, uint32_t reps, double *gbps);

/* Measured arithmetic rates of this GPU in a precision's arithmetic (f64 for
 * EV_C128, f32 for EV_C64), in TFLOP/s with a multiply-add counted as 2
 * flops: independent vector-ALU multiply-add chains alone, matrix-core
 * accumulations alone (v_mfma_f64_4x4x4f64 / v_mfma_f32_4x4x1f32, the shapes
 * a 4x4 gate maps onto), and both interleaved at equal flops.  A mixed rate
 * near the larger of the other two means the units share one budget; near
 * their sum, that they run side by side.  The achievable-peak reference for
 * the compute roofline, as ev_copy_bandwidth is for HBM.  No counterpart in
 * the reference. */
typedef struct {
    double valu_tflops;
    double mfma_tflops;
    double mixed_tflops;
} ev_compute_rates;
int ev_compute_rates_measure(ev_precision p, ev_compute_rates *out);

/* ---------------------------------------------------------- introspection */

/* The fused pass plan ev_sv_run would execute for a measurement-free circuit
 * (descriptor layouts in everett_amd/csrc/fused.hpp).  Call once with NULL
 * arrays to get the counts, then with arrays of at least those sizes.  coef
 * is in c128 (a c64 run converts it to f32 on upload). */
typedef struct {
    uint32_t num_passes, num_stages, num_ops, num_coef;
    uint32_t pass_bytes, stage_bytes, op_bytes;
    uint32_t tile_qubits;
    /* a trailing network of disjoint SWAPs that the run executes after the
       passes as one bit-permutation pass: qubit pairs (swap_a[i], swap_b[i]) */
    uint32_t num_swap_pairs;
    uint32_t swap_a[32], swap_b[32];
} ev_plan_info;
int ev_debug_plan(const ev_circuit *c, const ev_run_options *opt, void *passes, void *stages, void *ops,
                  ev_c128 *coef, ev_plan_info *info);

/* ------------------------------------------------------------- profiling */

/* Statistics of the most recent ev_sv_run / ev_dm_run on the calling thread. */
typedef struct {
    uint64_t logical_gates; /* gate ops applied (unfused count) */
    uint64_t passes;        /* full-state kernel passes launched */
    uint64_t launches;      /* kernel launches */
    double device_ms;       /* device time of the op stream (HIP events) */
    double pass_ms;         /* summed device time of the fused pass kernels */
    uint64_t pass_launches; /* fused pass kernel launches */
    double plan_ms;         /* host planning time */
    double pass_flops;      /* algorithmic flops of the fused pass launches: per op, its
                               definition's complex arithmetic (8n - 2 flops for an n-term
                               complex dot, 6 for a complex scale) over the amplitudes it
                               touches; no counterpart in the reference */
} ev_run_stats;
int ev_last_run_stats(ev_run_stats *out);

/* Per-kernel timing of the most recent run made with ev_run_options.profile
 * set: launches, summed device time (HIP events on the handle's stream) and
 * the algorithmic HBM bytes those launches move. */
typedef struct {
    char name[32];
    uint64_t launches;
    double total_ms;
    double bytes;
} ev_kernel_stat;
int ev_last_run_kernel_stats(ev_kernel_stat *out, size_t cap, size_t *count);

#ifdef __cplusplus
}
#endif
#endif /* EVERETT_H */
