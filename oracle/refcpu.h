/*
 * refcpu.h -- CPU oracle for the everett hot path (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C restatement of the reference's FP64 algorithms, used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
 * Nothing in the product library links or calls this code.
 *
 * Every function names the reference file:line it follows (paths relative to
 * the hz2/everett crate root).  Rounding contract: Complex64 multiplication is
 * (fma(a.re,b.re,-(a.im*b.im)), fma(a.re,b.im,a.im*b.re)) exactly as
 * src/complex.rs:115-125 writes it with mul_add; the file is compiled with
 * -ffp-contract=off so no other contraction happens.
 *
 * Parity pins (what checks each part; see tests/test_oracle.py):
 *   complex          known answers of src/complex.rs:207-340
 *   rng              xoshiro256++ published vector (state {1,2,3,4}) and the
 *                    SplitMix64(0) published first outputs
 *   gates            identities of src/gate.rs:297-360 and tests/properties.rs
 *   kernels          brute-force Kronecker products (numpy) for n <= 8, the
 *                    sample_state fixture of src/kernel.rs:631-635, closed forms
 *   measure          collapse/renormalise invariants (src/measure.rs:59-106)
 *   density/noise    brute-force U rho U^dag and sum K rho K^dag for n <= 5,
 *                    Kraus completeness (src/noise.rs:241-310)
 *   generators       qft closed form (tests/algorithms_qft.rs:52-74)
 */
#ifndef EVERETT_REFCPU_H
#define EVERETT_REFCPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/complex.rs:22-29: #[repr(C)] { re: f64, im: f64 } */
typedef struct { double re, im; } rc128;

/* Same field layout as the product's ev_op so one test-side buffer feeds both.
 * kind: 1 Apply1, 2 Apply2, 3 Controlled, 4 Measure, 5 IfClassic-prefix
 * (guards the next op; src/op.rs:14-58), 6 more controls of the Controlled
 * record before it (a Controlled record's num_controls is its total; controls
 * past the first REF_MAX_CONTROLS follow in kind-6 records). */
#define REF_MAX_CONTROLS 16
typedef struct {
    uint32_t kind;
    uint32_t num_controls;
    uint64_t target; /* Apply1 target / Apply2 a / Controlled target / Measure qubit / If bit */
    uint64_t b;      /* Apply2 b / Measure into */
    uint64_t controls[REF_MAX_CONTROLS];
    rc128 m[16];
} ref_op;

/* ---- rng.rs ---- */
typedef struct { uint64_t s[4]; } ref_rng;
void ref_rng_seed(ref_rng *r, uint64_t seed);         /* src/rng.rs:18-30 */
uint64_t ref_rng_next_u64(ref_rng *r);                /* src/rng.rs:34-48 */
double ref_rng_next_f64(ref_rng *r);                  /* src/rng.rs:52-56 */
void ref_rng_set_state(ref_rng *r, const uint64_t s[4]);

/* ---- complex.rs ---- */
rc128 ref_cmul(rc128 a, rc128 b);                     /* src/complex.rs:115-125 */
double ref_norm_sqr(rc128 a);                         /* src/complex.rs:80-82 */

/* ---- gate.rs (2x2 row-major, 4x4 row-major |q_a q_b>) ---- */
void ref_gate_h(rc128 m[4]);
void ref_gate_x(rc128 m[4]);
void ref_gate_y(rc128 m[4]);
void ref_gate_z(rc128 m[4]);
void ref_gate_s(rc128 m[4]);
void ref_gate_t(rc128 m[4]);
void ref_gate_rx(double theta, rc128 m[4]);
void ref_gate_ry(double theta, rc128 m[4]);
void ref_gate_rz(double theta, rc128 m[4]);
void ref_gate_phase(double lambda, rc128 m[4]);
void ref_gate_compose(const rc128 a[4], const rc128 b[4], rc128 out[4]); /* gate.rs:153-162 */
void ref_gate_adjoint(const rc128 a[4], rc128 out[4]);                   /* gate.rs:142-150 */
void ref_gate_cnot(rc128 m[16]);
void ref_gate_cz(rc128 m[16]);
void ref_gate_swap(rc128 m[16]);
void ref_gate_controlled(const rc128 u[4], rc128 m[16]);                /* gate.rs:240-247 */

/* ---- kernel.rs ---- */
void ref_index_pair(size_t k, size_t i, size_t out[2]);                 /* kernel.rs:37-44 */
void ref_index_quad(size_t a, size_t b, size_t i, size_t out[4]);       /* kernel.rs:52-66 */
void ref_set_threads(int nthreads); /* 1 = literal serial loops; >1 = OpenMP over the same
                                       reduction-free loops (bit-identical results) */
void ref_apply_1q(rc128 *amps, size_t len, size_t k, const rc128 m[4]);  /* kernel.rs:76-183 */
void ref_apply_2q(rc128 *amps, size_t len, size_t a, size_t b, const rc128 m[16]); /* kernel.rs:192-289 */
void ref_apply_controlled_1q(rc128 *amps, size_t len, const uint64_t *controls, size_t nc,
                             size_t target, const rc128 m[4]);           /* kernel.rs:299-437 */

/* ---- state.rs / measure.rs ---- */
double ref_state_norm_sqr(const rc128 *amps, size_t len);               /* state.rs:94-96 */
void ref_state_normalize(rc128 *amps, size_t len);                      /* state.rs:103-111 */
rc128 ref_state_overlap(const rc128 *self_, const rc128 *other, size_t len); /* state.rs:130-138 */
int ref_measure_qubit(rc128 *amps, size_t len, size_t k, ref_rng *rng); /* measure.rs:15-51 */

/* ---- backend/statevector.rs + backend/mod.rs ---- */
/* Runs ops from |0..0> (amps must hold 2^n entries); classical gets ncl bytes.
 * Returns 0, or -1 on a malformed op array. */
int ref_sv_run(const ref_op *ops, size_t nops, size_t n, uint64_t seed,
               rc128 *amps, uint8_t *classical, size_t ncl);
/* Applies ops to an existing state (no reset); rng state is caller-owned. */
int ref_sv_apply_ops(const ref_op *ops, size_t nops, rc128 *amps, size_t len,
                     ref_rng *rng, uint8_t *classical, size_t ncl);

/* ---- density.rs / noise.rs / backend/density_matrix.rs ---- */
/* rho: row-major dim x dim, dim = 2^n. */
void ref_dm_apply_unitary(rc128 *rho, size_t n, const rc128 *u, const size_t *targets, size_t m); /* density.rs:164-191 */
void ref_dm_apply_channel(rc128 *rho, size_t n, const rc128 *kraus, size_t nk, size_t target);   /* density.rs:198-216 */
double ref_dm_trace(const rc128 *rho, size_t n);                        /* density.rs:121-124 */
double ref_dm_purity(const rc128 *rho, size_t n);                       /* density.rs:130-140 */
void ref_dm_renormalize(rc128 *rho, size_t n);                          /* density.rs:219-227 */
/* channel kinds: 0 none, 1 depolarizing, 2 amplitude damping, 3 phase damping,
 * 4 bit flip, 5 dephasing, 6 custom (ops given). Returns the op count written. */
size_t ref_channel_kraus(int kind, double p, const rc128 *custom, size_t ncustom, rc128 *out); /* noise.rs:179-234 */
typedef struct {
    int kind_1q; double p_1q; const rc128 *custom_1q; size_t ncustom_1q;
    int kind_2q; double p_2q; const rc128 *custom_2q; size_t ncustom_2q;
    double readout_error;
} ref_noise;
int ref_dm_run(const ref_op *ops, size_t nops, size_t n, const ref_noise *noise, uint64_t seed,
               rc128 *rho, uint8_t *classical, size_t ncl);             /* density_matrix.rs:97-113 */
/* The same run with every gate and channel as a pass over rho viewed as a
 * 2n-qubit vector (superoperators); equal to ref_dm_run to rounding, and
 * parallel (ref_set_threads), for sizes the literal loops cannot reach. */
int ref_dm_run_superop(const ref_op *ops, size_t nops, size_t n, const ref_noise *noise, uint64_t seed,
                       rc128 *rho, uint8_t *classical, size_t ncl);

/* ---- algorithms/qft.rs and the workload generators ---- */
/* Each returns the number of ops written (call with out=NULL to count). */
size_t ref_gen_qft(size_t n, ref_op *out);                              /* qft.rs:54-62, 94-105 */
size_t ref_gen_brickwall(size_t n, size_t depth, ref_rng *rng, ref_op *out);
size_t ref_gen_dm_random(size_t n, size_t depth, ref_rng *rng, ref_op *out);

/* Born-rule sampler definition shared with the product (see refcpu.c). */
void ref_sample(const rc128 *amps, size_t n, const double *u, size_t shots, uint64_t *out);

#ifdef __cplusplus
}
#endif
#endif
