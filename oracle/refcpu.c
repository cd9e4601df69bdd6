/*
 * refcpu.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see refcpu.h).
 *
 * Literal C restatement of hz2/everett 0.1.2's FP64 algorithms.  Loop order,
 * operand order of every complex product and the mul_add placement follow the
 * cited Rust lines, so with -ffp-contract=off the results are those of the
 * reference's scalar path; the reference's AVX2 paths (kernel.rs:125-183,
 * 367-437) compute fmaddsub(mr, v, mi*swap(v)), which rounds identically.
 */
#include "refcpu.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static int g_threads = 1;
void ref_set_threads(int nthreads) { g_threads = nthreads < 1 ? 1 : nthreads; }

/* ------------------------------------------------------------------ rng.rs */

/* src/rng.rs:18-30 -- SplitMix64 expands the seed into the four state words. */
void ref_rng_seed(ref_rng *r, uint64_t seed)
{
    for (int w = 0; w < 4; ++w) {
        seed += 0x9E3779B97F4A7C15ull;
        uint64_t z = seed;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        r->s[w] = z ^ (z >> 31);
    }
}

void ref_rng_set_state(ref_rng *r, const uint64_t s[4]) { memcpy(r->s, s, sizeof r->s); }

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* src/rng.rs:34-48 -- xoshiro256++ */
uint64_t ref_rng_next_u64(ref_rng *r)
{
    uint64_t *s = r->s;
    uint64_t result = rotl(s[0] + s[3], 23) + s[0];
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

/* src/rng.rs:52-56 -- top 53 bits times 2^-53 */
double ref_rng_next_f64(ref_rng *r)
{
    uint64_t bits = ref_rng_next_u64(r) >> 11;
    return (double)bits * (1.0 / (double)(1ull << 53));
}

/* -------------------------------------------------------------- complex.rs */

static inline rc128 C(double re, double im) { rc128 z = {re, im}; return z; }
static inline rc128 cadd(rc128 a, rc128 b) { return C(a.re + b.re, a.im + b.im); }
static inline rc128 cconj(rc128 a) { return C(a.re, -a.im); }
static inline rc128 cscale(rc128 a, double s) { return C(a.re * s, a.im * s); } /* complex.rs:127-133 */

/* src/complex.rs:115-125 */
rc128 ref_cmul(rc128 a, rc128 b)
{
    return C(fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re));
}
#define cmul ref_cmul

/* src/complex.rs:80-82 */
double ref_norm_sqr(rc128 a) { return fma(a.re, a.re, a.im * a.im); }

/* ----------------------------------------------------------------- gate.rs */

static void set4(rc128 m[4], rc128 a, rc128 b, rc128 c, rc128 d) { m[0] = a; m[1] = b; m[2] = c; m[3] = d; }
static const double FRAC_1_SQRT_2 = 0.70710678118654752440;
static const double FRAC_PI_4 = 0.78539816339744830962;
static const double TAU = 6.28318530717958647692;

void ref_gate_h(rc128 m[4]) { double s = FRAC_1_SQRT_2; set4(m, C(s, 0), C(s, 0), C(s, 0), C(-s, 0)); } /* gate.rs:81-85 */
void ref_gate_x(rc128 m[4]) { set4(m, C(0, 0), C(1, 0), C(1, 0), C(0, 0)); }                           /* gate.rs:62-64 */
void ref_gate_y(rc128 m[4]) { set4(m, C(0, 0), C(0, -1), C(0, 1), C(0, 0)); }                          /* gate.rs:68-71 */
void ref_gate_z(rc128 m[4]) { set4(m, C(1, 0), C(0, 0), C(0, 0), C(-1, 0)); }                          /* gate.rs:75-77 */
void ref_gate_s(rc128 m[4]) { set4(m, C(1, 0), C(0, 0), C(0, 0), C(0, 1)); }                           /* gate.rs:89-91 */
/* gate.rs:95-102: expi(pi/4) = (cos, sin) complex.rs:63-65 */
void ref_gate_t(rc128 m[4]) { set4(m, C(1, 0), C(0, 0), C(0, 0), C(cos(FRAC_PI_4), sin(FRAC_PI_4))); }
/* gate.rs:106-110 */
void ref_gate_rx(double theta, rc128 m[4])
{
    double s = sin(theta / 2.0), co = cos(theta / 2.0);
    set4(m, C(co, 0), C(0, -s), C(0, -s), C(co, 0));
}
/* gate.rs:114-118 */
void ref_gate_ry(double theta, rc128 m[4])
{
    double s = sin(theta / 2.0), co = cos(theta / 2.0);
    set4(m, C(co, 0), C(-s, 0), C(s, 0), C(co, 0));
}
/* gate.rs:122-130 */
void ref_gate_rz(double theta, rc128 m[4])
{
    double a = -theta / 2.0, b = theta / 2.0;
    set4(m, C(cos(a), sin(a)), C(0, 0), C(0, 0), C(cos(b), sin(b)));
}
/* gate.rs:134-136 */
void ref_gate_phase(double lambda, rc128 m[4]) { set4(m, C(1, 0), C(0, 0), C(0, 0), C(cos(lambda), sin(lambda))); }

/* gate.rs:153-162: self * other */
void ref_gate_compose(const rc128 a[4], const rc128 b[4], rc128 out[4])
{
    rc128 r0 = cadd(cmul(a[0], b[0]), cmul(a[1], b[2]));
    rc128 r1 = cadd(cmul(a[0], b[1]), cmul(a[1], b[3]));
    rc128 r2 = cadd(cmul(a[2], b[0]), cmul(a[3], b[2]));
    rc128 r3 = cadd(cmul(a[2], b[1]), cmul(a[3], b[3]));
    set4(out, r0, r1, r2, r3);
}

/* gate.rs:142-150: conjugate transpose */
void ref_gate_adjoint(const rc128 a[4], rc128 out[4])
{
    rc128 r0 = C(a[0].re, -a[0].im), r1 = C(a[2].re, -a[2].im), r2 = C(a[1].re, -a[1].im), r3 = C(a[3].re, -a[3].im);
    set4(out, r0, r1, r2, r3);
}

static void eye4(rc128 m[16]) { memset(m, 0, 16 * sizeof(rc128)); for (int i = 0; i < 4; ++i) m[i * 5] = C(1, 0); }
/* gate.rs:190-198 */
void ref_gate_cnot(rc128 m[16]) { eye4(m); m[10] = C(0, 0); m[11] = C(1, 0); m[14] = C(1, 0); m[15] = C(0, 0); }
/* gate.rs:202-221 */
void ref_gate_cz(rc128 m[16]) { eye4(m); m[15] = C(-1, 0); }
/* gate.rs:225-232 */
void ref_gate_swap(rc128 m[16]) { eye4(m); m[5] = C(0, 0); m[6] = C(1, 0); m[9] = C(1, 0); m[10] = C(0, 0); }
/* gate.rs:240-247: diag(I_2, u) in the |q_a q_b> basis */
void ref_gate_controlled(const rc128 u[4], rc128 m[16])
{
    eye4(m);
    m[10] = u[0]; m[11] = u[1]; m[14] = u[2]; m[15] = u[3];
}

/* --------------------------------------------------------------- kernel.rs */

/* kernel.rs:37-44 */
void ref_index_pair(size_t k, size_t i, size_t out[2])
{
    size_t low_mask = ((size_t)1 << k) - 1;
    size_t i0 = ((i >> k) << (k + 1)) | (i & low_mask);
    out[0] = i0;
    out[1] = i0 | ((size_t)1 << k);
}

/* kernel.rs:52-66 */
void ref_index_quad(size_t a, size_t b, size_t i, size_t out[4])
{
    size_t lo = a < b ? a : b, hi = a < b ? b : a;
    size_t low_mask = ((size_t)1 << lo) - 1;
    size_t t = ((i >> lo) << (lo + 1)) | (i & low_mask);
    size_t mid_mask = ((size_t)1 << hi) - 1;
    size_t base = ((t >> hi) << (hi + 1)) | (t & mid_mask);
    size_t bit_a = (size_t)1 << a, bit_b = (size_t)1 << b;
    out[0] = base;
    out[1] = base | bit_b;
    out[2] = base | bit_a;
    out[3] = base | bit_a | bit_b;
}

/* kernel.rs:100-119 (and the AVX2 path 125-183, which rounds the same way) */
void ref_apply_1q(rc128 *amps, size_t len, size_t k, const rc128 m[4])
{
    const long long pairs = (long long)(len / 2);
#pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
    for (long long ii = 0; ii < pairs; ++ii) {
        size_t p[2];
        ref_index_pair(k, (size_t)ii, p);
        rc128 a0 = amps[p[0]], a1 = amps[p[1]];
        amps[p[0]] = cadd(cmul(m[0], a0), cmul(m[1], a1));
        amps[p[1]] = cadd(cmul(m[2], a0), cmul(m[3], a1));
    }
}

static int eq16(const rc128 *a, const rc128 *b)
{
    for (int i = 0; i < 16; ++i)
        if (a[i].re != b[i].re || a[i].im != b[i].im) return 0;
    return 1;
}

/* kernel.rs:192-289: exact-equality dispatch, then the dense loop. */
void ref_apply_2q(rc128 *amps, size_t len, size_t a, size_t b, const rc128 mat[16])
{
    rc128 g[16];
    const long long groups = (long long)(len / 4);
    ref_gate_cnot(g);
    if (eq16(mat, g)) { /* kernel.rs:248-260 */
#pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
        for (long long i = 0; i < groups; ++i) {
            size_t idx[4];
            ref_index_quad(a, b, (size_t)i, idx);
            rc128 tmp = amps[idx[2]];
            amps[idx[2]] = amps[idx[3]];
            amps[idx[3]] = tmp;
        }
        return;
    }
    ref_gate_cz(g);
    if (eq16(mat, g)) { /* kernel.rs:263-274 */
#pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
        for (long long i = 0; i < groups; ++i) {
            size_t idx[4];
            ref_index_quad(a, b, (size_t)i, idx);
            amps[idx[3]] = C(-amps[idx[3]].re, -amps[idx[3]].im);
        }
        return;
    }
    ref_gate_swap(g);
    if (eq16(mat, g)) { /* kernel.rs:277-289 */
#pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
        for (long long i = 0; i < groups; ++i) {
            size_t idx[4];
            ref_index_quad(a, b, (size_t)i, idx);
            rc128 tmp = amps[idx[1]];
            amps[idx[1]] = amps[idx[2]];
            amps[idx[2]] = tmp;
        }
        return;
    }
    /* kernel.rs:214-243 */
#pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
    for (long long i = 0; i < groups; ++i) {
        size_t idx[4];
        ref_index_quad(a, b, (size_t)i, idx);
        rc128 amp[4] = {amps[idx[0]], amps[idx[1]], amps[idx[2]], amps[idx[3]]};
        for (int row = 0; row < 4; ++row) {
            const rc128 *r = mat + row * 4;
            rc128 v = cadd(cadd(cadd(cmul(r[0], amp[0]), cmul(r[1], amp[1])), cmul(r[2], amp[2])),
                           cmul(r[3], amp[3]));
            amps[idx[row]] = v;
        }
    }
}

/* kernel.rs:299-358 (AVX2 path 367-437 rounds identically) */
void ref_apply_controlled_1q(rc128 *amps, size_t len, const uint64_t *controls, size_t nc,
                             size_t target, const rc128 m[4])
{
    size_t mask = 0;
    for (size_t c = 0; c < nc; ++c) mask |= (size_t)1 << controls[c];
    const long long pairs = (long long)(len / 2);
#pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
    for (long long ii = 0; ii < pairs; ++ii) {
        size_t p[2];
        ref_index_pair(target, (size_t)ii, p);
        if ((p[0] & mask) == mask) {
            rc128 a0 = amps[p[0]], a1 = amps[p[1]];
            amps[p[0]] = cadd(cmul(m[0], a0), cmul(m[1], a1));
            amps[p[1]] = cadd(cmul(m[2], a0), cmul(m[3], a1));
        }
    }
}

/* ------------------------------------------------------- state.rs / measure */

/* state.rs:94-96 -- sequential sum */
double ref_state_norm_sqr(const rc128 *amps, size_t len)
{
    double s = 0.0;
    for (size_t j = 0; j < len; ++j) s += ref_norm_sqr(amps[j]);
    return s;
}

/* state.rs:103-111 */
void ref_state_normalize(rc128 *amps, size_t len)
{
    double norm = sqrt(ref_state_norm_sqr(amps, len));
    if (norm > 0.0) {
        double inv = 1.0 / norm;
        for (size_t j = 0; j < len; ++j) amps[j] = cscale(amps[j], inv);
    }
}

/* state.rs:130-138: sum_j conj(other_j) * self_j */
rc128 ref_state_overlap(const rc128 *self_, const rc128 *other, size_t len)
{
    rc128 acc = C(0, 0);
    for (size_t j = 0; j < len; ++j) acc = cadd(acc, cmul(cconj(other[j]), self_[j]));
    return acc;
}

/* measure.rs:15-51 */
int ref_measure_qubit(rc128 *amps, size_t len, size_t k, ref_rng *rng)
{
    size_t bit = (size_t)1 << k;
    double p1 = 0.0;
    for (size_t j = 0; j < len; ++j)
        if (j & bit) p1 += ref_norm_sqr(amps[j]);
    int outcome = ref_rng_next_f64(rng) < p1;
    double surviving = outcome ? p1 : 1.0 - p1;
    double scale = surviving > 0.0 ? 1.0 / sqrt(surviving) : 0.0;
    for (size_t j = 0; j < len; ++j) {
        int keep = ((j & bit) != 0) == outcome;
        amps[j] = keep ? cscale(amps[j], scale) : C(0, 0);
    }
    return outcome;
}

/* ------------------------------------------ backend/mod.rs drive + apply_op */

typedef struct {
    void (*apply_1q)(void *ctx, const rc128 m[4], size_t target);
    void (*apply_2q)(void *ctx, const rc128 m[16], size_t a, size_t b);
    void (*apply_controlled)(void *ctx, const uint64_t *controls, size_t nc, const rc128 m[4], size_t target);
    int (*measure)(void *ctx, size_t qubit);
} backend_vt;

/* Returns the index after op i (an If prefix consumes its guarded op), or 0 on
 * a malformed array.  backend/mod.rs:51-72. */
static size_t apply_op(const backend_vt *vt, void *ctx, const ref_op *ops, size_t i, size_t nops,
                       uint8_t *classical, int execute)
{
    if (i >= nops) return 0;
    const ref_op *op = &ops[i];
    switch (op->kind) {
    case 1:
        if (execute) vt->apply_1q(ctx, op->m, op->target);
        return i + 1;
    case 2:
        if (execute) vt->apply_2q(ctx, op->m, op->target, op->b);
        return i + 1;
    case 3: {
        /* Op::Controlled's controls Vec (op.rs:34-36): past REF_MAX_CONTROLS
         * they continue in kind-6 records after this one. */
        size_t total = op->num_controls, got = total < REF_MAX_CONTROLS ? total : REF_MAX_CONTROLS, j = i + 1;
        uint64_t *cs = (uint64_t *)malloc((total ? total : 1) * sizeof(uint64_t));
        if (!cs) return 0;
        memcpy(cs, op->controls, got * sizeof(uint64_t));
        while (got < total) {
            if (j >= nops || ops[j].kind != 6 || ops[j].num_controls < 1 || ops[j].num_controls > REF_MAX_CONTROLS ||
                got + ops[j].num_controls > total) {
                free(cs);
                return 0;
            }
            memcpy(cs + got, ops[j].controls, ops[j].num_controls * sizeof(uint64_t));
            got += ops[j].num_controls;
            ++j;
        }
        if (execute) vt->apply_controlled(ctx, cs, total, op->m, op->target);
        free(cs);
        return j;
    }
    case 4:
        if (execute) classical[op->b] = (uint8_t)vt->measure(ctx, op->target);
        return i + 1;
    case 5:
        return apply_op(vt, ctx, ops, i + 1, nops, classical, execute && classical[op->target]);
    default:
        return 0;
    }
}

/* backend/mod.rs:43-49 */
static int drive(const backend_vt *vt, void *ctx, const ref_op *ops, size_t nops, uint8_t *classical, size_t ncl)
{
    memset(classical, 0, ncl);
    size_t i = 0;
    while (i < nops) {
        size_t nx = apply_op(vt, ctx, ops, i, nops, classical, 1);
        if (nx == 0) return -1;
        i = nx;
    }
    return 0;
}

typedef struct { rc128 *amps; size_t len; ref_rng *rng; } sv_ctx;
static void sv_1q(void *c, const rc128 m[4], size_t t) { sv_ctx *s = c; ref_apply_1q(s->amps, s->len, t, m); }
static void sv_2q(void *c, const rc128 m[16], size_t a, size_t b) { sv_ctx *s = c; ref_apply_2q(s->amps, s->len, a, b, m); }
static void sv_ctl(void *c, const uint64_t *cs, size_t nc, const rc128 m[4], size_t t)
{
    sv_ctx *s = c;
    ref_apply_controlled_1q(s->amps, s->len, cs, nc, t, m);
}
static int sv_meas(void *c, size_t q) { sv_ctx *s = c; return ref_measure_qubit(s->amps, s->len, q, s->rng); }
static const backend_vt SV_VT = {sv_1q, sv_2q, sv_ctl, sv_meas};

int ref_sv_apply_ops(const ref_op *ops, size_t nops, rc128 *amps, size_t len, ref_rng *rng,
                     uint8_t *classical, size_t ncl)
{
    sv_ctx ctx = {amps, len, rng};
    return drive(&SV_VT, &ctx, ops, nops, classical, ncl);
}

/* backend/statevector.rs:77-88 (validation is the caller's job here) */
int ref_sv_run(const ref_op *ops, size_t nops, size_t n, uint64_t seed, rc128 *amps, uint8_t *classical, size_t ncl)
{
    size_t len = (size_t)1 << n;
    memset(amps, 0, len * sizeof(rc128));
    amps[0] = C(1, 0); /* state.rs:38-44 */
    ref_rng rng;
    ref_rng_seed(&rng, seed);
    return ref_sv_apply_ops(ops, nops, amps, len, &rng, classical, ncl);
}

/* -------------------------------------------------------------- density.rs */

/* density.rs:426-454 -- group g's indices; bit p of the local index maps to targets[p]. */
static void index_group(const size_t *targets, size_t m, size_t g, size_t *out)
{
    size_t sorted[64];
    memcpy(sorted, targets, m * sizeof(size_t));
    for (size_t i = 1; i < m; ++i) { /* sort_unstable (ascending) */
        size_t v = sorted[i], j = i;
        while (j > 0 && sorted[j - 1] > v) { sorted[j] = sorted[j - 1]; --j; }
        sorted[j] = v;
    }
    size_t base = g;
    for (size_t i = 0; i < m; ++i) {
        size_t t = sorted[i];
        size_t low_mask = ((size_t)1 << t) - 1;
        base = ((base >> t) << (t + 1)) | (base & low_mask);
    }
    size_t gd = (size_t)1 << m;
    for (size_t k = 0; k < gd; ++k) {
        size_t idx = base;
        for (size_t p = 0; p < m; ++p)
            if ((k >> p) & 1) idx |= (size_t)1 << targets[p];
        out[k] = idx;
    }
}

/* density.rs:330-360 */
static void left_multiply(const rc128 *rho, rc128 *out, const rc128 *u, const size_t *targets, size_t m, size_t n)
{
    const size_t dim = (size_t)1 << n, gd = (size_t)1 << m, ngroups = dim / gd;
#pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
    for (long long col = 0; col < (long long)dim; ++col) {
        size_t *rows = malloc(gd * sizeof(size_t));
        rc128 *in = malloc(gd * sizeof(rc128));
        for (size_t g = 0; g < ngroups; ++g) {
            index_group(targets, m, g, rows);
            for (size_t c = 0; c < gd; ++c) in[c] = rho[rows[c] * dim + (size_t)col];
            for (size_t r = 0; r < gd; ++r) {
                rc128 acc = C(0, 0);
                for (size_t c = 0; c < gd; ++c) acc = cadd(acc, cmul(u[r * gd + c], in[c]));
                out[rows[r] * dim + (size_t)col] = acc;
            }
        }
        free(rows);
        free(in);
    }
}

/* density.rs:363-391 (add=0) and 394-419 (add=1) */
static void right_multiply_adjoint(const rc128 *tmp, rc128 *dst, const rc128 *u, const size_t *targets, size_t m,
                                   size_t n, int add)
{
    const size_t dim = (size_t)1 << n, gd = (size_t)1 << m, ngroups = dim / gd;
#pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
    for (long long row = 0; row < (long long)dim; ++row) {
        size_t *cols = malloc(gd * sizeof(size_t));
        rc128 *in = malloc(gd * sizeof(rc128));
        for (size_t g = 0; g < ngroups; ++g) {
            index_group(targets, m, g, cols);
            for (size_t k = 0; k < gd; ++k) in[k] = tmp[(size_t)row * dim + cols[k]];
            for (size_t ci = 0; ci < gd; ++ci) {
                rc128 acc = C(0, 0);
                for (size_t k = 0; k < gd; ++k) acc = cadd(acc, cmul(in[k], cconj(u[ci * gd + k])));
                size_t at = (size_t)row * dim + cols[ci];
                dst[at] = add ? cadd(dst[at], acc) : acc;
            }
        }
        free(cols);
        free(in);
    }
}

/* density.rs:164-191 */
void ref_dm_apply_unitary(rc128 *rho, size_t n, const rc128 *u, const size_t *targets, size_t m)
{
    size_t dim = (size_t)1 << n;
    rc128 *out = calloc(dim * dim, sizeof(rc128));
    left_multiply(rho, out, u, targets, m, n);
    right_multiply_adjoint(out, rho, u, targets, m, n, 0);
    free(out);
}

/* density.rs:198-216 */
void ref_dm_apply_channel(rc128 *rho, size_t n, const rc128 *kraus, size_t nk, size_t target)
{
    size_t dim = (size_t)1 << n;
    rc128 *acc = calloc(dim * dim, sizeof(rc128));
    rc128 *tmp = malloc(dim * dim * sizeof(rc128));
    for (size_t k = 0; k < nk; ++k) {
        memset(tmp, 0, dim * dim * sizeof(rc128));
        left_multiply(rho, tmp, kraus + 4 * k, &target, 1, n);
        right_multiply_adjoint(tmp, acc, kraus + 4 * k, &target, 1, n, 1);
    }
    memcpy(rho, acc, dim * dim * sizeof(rc128));
    free(acc);
    free(tmp);
}

/* density.rs:121-124 */
double ref_dm_trace(const rc128 *rho, size_t n)
{
    size_t dim = (size_t)1 << n;
    double s = 0.0;
    for (size_t i = 0; i < dim; ++i) s += rho[i * dim + i].re;
    return s;
}

/* density.rs:130-140 */
double ref_dm_purity(const rc128 *rho, size_t n)
{
    size_t dim = (size_t)1 << n;
    double s = 0.0;
    for (size_t i = 0; i < dim * dim; ++i) s += ref_norm_sqr(rho[i]);
    return s;
}

/* density.rs:219-227 */
void ref_dm_renormalize(rc128 *rho, size_t n)
{
    size_t dim = (size_t)1 << n;
    double t = ref_dm_trace(rho, n);
    if (t > 0.0) {
        double inv = 1.0 / t;
        for (size_t i = 0; i < dim * dim; ++i) rho[i] = cscale(rho[i], inv);
    }
}

/* ---------------------------------------------------------------- noise.rs */

static void K4(rc128 *o, rc128 a, rc128 b, rc128 c, rc128 d) { o[0] = a; o[1] = b; o[2] = c; o[3] = d; }

/* noise.rs:179-234 */
size_t ref_channel_kraus(int kind, double p, const rc128 *custom, size_t ncustom, rc128 *out)
{
    rc128 z = C(0, 0);
    switch (kind) {
    case 1: { /* Depolarizing */
        double a = sqrt(1.0 - p), b = sqrt(p / 3.0);
        K4(out + 0, C(a, 0), z, z, C(a, 0));
        K4(out + 4, z, C(b, 0), C(b, 0), z);
        K4(out + 8, z, C(0, -b), C(0, b), z);
        K4(out + 12, C(b, 0), z, z, C(-b, 0));
        return 4;
    }
    case 2: { /* AmplitudeDamping */
        double k0 = sqrt(1.0 - p), k1 = sqrt(p);
        K4(out + 0, C(1, 0), z, z, C(k0, 0));
        K4(out + 4, z, C(k1, 0), z, z);
        return 2;
    }
    case 3: { /* PhaseDamping */
        double k0 = sqrt(1.0 - p), k1 = sqrt(p);
        K4(out + 0, C(1, 0), z, z, C(k0, 0));
        K4(out + 4, z, z, z, C(k1, 0));
        return 2;
    }
    case 4: { /* BitFlip */
        double a = sqrt(1.0 - p), b = sqrt(p);
        K4(out + 0, C(a, 0), z, z, C(a, 0));
        K4(out + 4, z, C(b, 0), C(b, 0), z);
        return 2;
    }
    case 5: { /* Dephasing */
        double a = sqrt(1.0 - p), b = sqrt(p);
        K4(out + 0, C(a, 0), z, z, C(a, 0));
        K4(out + 4, C(b, 0), z, z, C(-b, 0));
        return 2;
    }
    case 6:
        memcpy(out, custom, ncustom * 4 * sizeof(rc128));
        return ncustom;
    default:
        return 0;
    }
}

/* ----------------------------------------------- backend/density_matrix.rs */

typedef struct {
    rc128 *rho; size_t n; ref_rng *rng; const ref_noise *noise;
    rc128 k1[4 * 64]; size_t nk1; rc128 k2[4 * 64]; size_t nk2;
} dm_ctx;

/* density_matrix.rs:117-125 */
static void dm_1q(void *c, const rc128 m[4], size_t t)
{
    dm_ctx *d = c;
    ref_dm_apply_unitary(d->rho, d->n, m, &t, 1);
    if (d->nk1) ref_dm_apply_channel(d->rho, d->n, d->k1, d->nk1, t);
}

/* density_matrix.rs:127-138: targets [b, a]; channel on a, then b */
static void dm_2q(void *c, const rc128 m[16], size_t a, size_t b)
{
    dm_ctx *d = c;
    size_t tg[2] = {b, a};
    ref_dm_apply_unitary(d->rho, d->n, m, tg, 2);
    if (d->nk2) {
        ref_dm_apply_channel(d->rho, d->n, d->k2, d->nk2, a);
        ref_dm_apply_channel(d->rho, d->n, d->k2, d->nk2, b);
    }
}

/* density_matrix.rs:212-247 */
static rc128 *controlled_unitary(const uint64_t *controls, size_t nc, const rc128 m[4], size_t target, size_t n)
{
    size_t dim = (size_t)1 << n, mask = 0;
    rc128 *u = calloc(dim * dim, sizeof(rc128));
    for (size_t c = 0; c < nc; ++c) mask |= (size_t)1 << controls[c];
    for (size_t col = 0; col < dim; ++col) {
        if ((col & mask) == mask) {
            size_t col_t = (col >> target) & 1, partner = col ^ ((size_t)1 << target);
            if (col_t == 0) {
                u[col * dim + col] = m[0];
                u[partner * dim + col] = m[2];
            } else {
                u[partner * dim + col] = m[1];
                u[col * dim + col] = m[3];
            }
        } else {
            u[col * dim + col] = C(1, 0);
        }
    }
    return u;
}

/* density_matrix.rs:140-153 */
static void dm_ctl(void *c, const uint64_t *cs, size_t nc, const rc128 m[4], size_t t)
{
    dm_ctx *d = c;
    rc128 *u = controlled_unitary(cs, nc, m, t, d->n);
    size_t all[64];
    for (size_t q = 0; q < d->n; ++q) all[q] = q;
    ref_dm_apply_unitary(d->rho, d->n, u, all, d->n);
    free(u);
    if (d->nk1) {
        ref_dm_apply_channel(d->rho, d->n, d->k1, d->nk1, t);
        for (size_t i = 0; i < nc; ++i) ref_dm_apply_channel(d->rho, d->n, d->k1, d->nk1, cs[i]);
    }
}

/* density_matrix.rs:155-207 */
static int dm_meas(void *c, size_t q)
{
    dm_ctx *d = c;
    size_t dim = (size_t)1 << d->n, bit = (size_t)1 << q;
    double p1 = 0.0;
    for (size_t j = 0; j < dim; ++j)
        if (j & bit) p1 += d->rho[j * dim + j].re;
    int raw = ref_rng_next_f64(d->rng) < p1;
    for (size_t i = 0; i < dim; ++i) {
        int im = ((i & bit) != 0) == raw;
        for (size_t j = 0; j < dim; ++j) {
            int jm = ((j & bit) != 0) == raw;
            if (!im || !jm) d->rho[i * dim + j] = C(0, 0);
        }
    }
    ref_dm_renormalize(d->rho, d->n);
    if (d->noise->readout_error > 0.0 && ref_rng_next_f64(d->rng) < d->noise->readout_error) return !raw;
    return raw;
}
static const backend_vt DM_VT = {dm_1q, dm_2q, dm_ctl, dm_meas};

/* density_matrix.rs:97-113 */
int ref_dm_run(const ref_op *ops, size_t nops, size_t n, const ref_noise *noise, uint64_t seed, rc128 *rho,
               uint8_t *classical, size_t ncl)
{
    size_t dim = (size_t)1 << n;
    memset(rho, 0, dim * dim * sizeof(rc128));
    rho[0] = C(1, 0); /* density.rs:45-55 */
    ref_rng rng;
    ref_rng_seed(&rng, seed);
    dm_ctx *d = calloc(1, sizeof(dm_ctx));
    d->rho = rho; d->n = n; d->rng = &rng; d->noise = noise;
    if (noise->ncustom_1q > 64 || noise->ncustom_2q > 64) { free(d); return -1; }
    d->nk1 = ref_channel_kraus(noise->kind_1q, noise->p_1q, noise->custom_1q, noise->ncustom_1q, d->k1);
    d->nk2 = ref_channel_kraus(noise->kind_2q, noise->p_2q, noise->custom_2q, noise->ncustom_2q, d->k2);
    int rc = drive(&DM_VT, d, ops, nops, classical, ncl);
    free(d);
    return rc;
}

/* Superoperator form of the same backend (not in the reference: a faster
 * restatement for large n, checked against ref_dm_run at small n).  rho's
 * row-major storage is the 2n-qubit vector whose bit n+t is row qubit t and
 * bit t column qubit t, so U rho U^dag is U on the row bits then conj(U) on
 * the column bits (density.rs:164-191 entry by entry), and a channel
 * sum_k K rho K^dag (density.rs:198-216) is the 4x4 S = sum_k K (x) conj(K) on
 * bits (n+t, t): S[(r,c),(r',c')] = sum_k K[r][r'] conj(K[c][c']).  Sums run
 * in another order than the literal loops, so results agree to rounding. */
static void conj_n(const rc128 *m, size_t cnt, rc128 *out)
{
    for (size_t i = 0; i < cnt; ++i) out[i] = cconj(m[i]);
}

/* S = sum_k A_k (x) conj(B_k) on (row bit, column bit), the 4x4 acting on
 * vec(rho) as rho -> sum_k A_k rho B_k^dag (A = B = K for a channel, A = B =
 * U with one term for a unitary). */
static void superop(const rc128 *a, const rc128 *b, size_t nk, rc128 s[16])
{
    for (int i = 0; i < 16; ++i) s[i] = C(0, 0);
    for (size_t k = 0; k < nk; ++k) {
        const rc128 *A = a + 4 * k, *B = b + 4 * k;
        for (int r = 0; r < 2; ++r)
            for (int c = 0; c < 2; ++c)
                for (int r2 = 0; r2 < 2; ++r2)
                    for (int c2 = 0; c2 < 2; ++c2)
                        s[(2 * r + c) * 4 + (2 * r2 + c2)] =
                            cadd(s[(2 * r + c) * 4 + (2 * r2 + c2)], cmul(A[2 * r + r2], cconj(B[2 * c + c2])));
    }
}

static void mat4_mul(const rc128 x[16], const rc128 y[16], rc128 out[16])
{
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            rc128 acc = C(0, 0);
            for (int k = 0; k < 4; ++k) acc = cadd(acc, cmul(x[4 * r + k], y[4 * k + c]));
            out[4 * r + c] = acc;
        }
}

static void dmv_channel(dm_ctx *d, const rc128 *kraus, size_t nk, size_t t)
{
    rc128 s[16];
    superop(kraus, kraus, nk, s);
    ref_apply_2q(d->rho, (size_t)1 << (2 * d->n), d->n + t, t, s);
}

/* A one-qubit gate and the channel after it are one pass: S_K (U (x) conj(U)). */
static void dmv_1q(void *c, const rc128 m[4], size_t t)
{
    dm_ctx *d = c;
    rc128 su[16], sk[16], s[16];
    superop(m, m, 1, su);
    if (d->nk1) {
        superop(d->k1, d->k1, d->nk1, sk);
        mat4_mul(sk, su, s);
    } else {
        memcpy(s, su, sizeof s);
    }
    ref_apply_2q(d->rho, (size_t)1 << (2 * d->n), d->n + t, t, s);
}

static void dmv_2q(void *c, const rc128 m[16], size_t a, size_t b)
{
    dm_ctx *d = c;
    const size_t len = (size_t)1 << (2 * d->n);
    rc128 mc[16];
    conj_n(m, 16, mc);
    ref_apply_2q(d->rho, len, d->n + a, d->n + b, m);
    ref_apply_2q(d->rho, len, a, b, mc);
    if (d->nk2) {
        dmv_channel(d, d->k2, d->nk2, a);
        dmv_channel(d, d->k2, d->nk2, b);
    }
}

static void dmv_ctl(void *c, const uint64_t *cs, size_t nc, const rc128 m[4], size_t t)
{
    dm_ctx *d = c;
    const size_t len = (size_t)1 << (2 * d->n);
    uint64_t rows[64];
    rc128 mc[4];
    if (nc > 64) return;
    for (size_t i = 0; i < nc; ++i) rows[i] = cs[i] + d->n;
    conj_n(m, 4, mc);
    ref_apply_controlled_1q(d->rho, len, rows, nc, d->n + t, m);
    ref_apply_controlled_1q(d->rho, len, cs, nc, t, mc);
    if (d->nk1) {
        dmv_channel(d, d->k1, d->nk1, t);
        for (size_t i = 0; i < nc; ++i) dmv_channel(d, d->k1, d->nk1, cs[i]);
    }
}
static const backend_vt DMV_VT = {dmv_1q, dmv_2q, dmv_ctl, dm_meas};

int ref_dm_run_superop(const ref_op *ops, size_t nops, size_t n, const ref_noise *noise, uint64_t seed, rc128 *rho,
                       uint8_t *classical, size_t ncl)
{
    size_t dim = (size_t)1 << n;
    memset(rho, 0, dim * dim * sizeof(rc128));
    rho[0] = C(1, 0);
    ref_rng rng;
    ref_rng_seed(&rng, seed);
    dm_ctx *d = calloc(1, sizeof(dm_ctx));
    d->rho = rho; d->n = n; d->rng = &rng; d->noise = noise;
    if (noise->ncustom_1q > 64 || noise->ncustom_2q > 64) { free(d); return -1; }
    d->nk1 = ref_channel_kraus(noise->kind_1q, noise->p_1q, noise->custom_1q, noise->ncustom_1q, d->k1);
    d->nk2 = ref_channel_kraus(noise->kind_2q, noise->p_2q, noise->custom_2q, noise->ncustom_2q, d->k2);
    int rc = drive(&DMV_VT, d, ops, nops, classical, ncl);
    free(d);
    return rc;
}

/* -------------------------------------------- algorithms/qft.rs, generators */

static ref_op *put1(ref_op *out, size_t *cnt, const rc128 m[4], size_t t)
{
    if (out) {
        ref_op *o = &out[*cnt];
        memset(o, 0, sizeof *o);
        o->kind = 1; o->target = t;
        memcpy(o->m, m, 4 * sizeof(rc128));
    }
    ++*cnt;
    return out;
}
static void put2(ref_op *out, size_t *cnt, const rc128 m[16], size_t a, size_t b)
{
    if (out) {
        ref_op *o = &out[*cnt];
        memset(o, 0, sizeof *o);
        o->kind = 2; o->target = a; o->b = b;
        memcpy(o->m, m, 16 * sizeof(rc128));
    }
    ++*cnt;
}
static void putc1(ref_op *out, size_t *cnt, size_t ctrl, const rc128 m[4], size_t t)
{
    if (out) {
        ref_op *o = &out[*cnt];
        memset(o, 0, sizeof *o);
        o->kind = 3; o->target = t; o->num_controls = 1; o->controls[0] = ctrl;
        memcpy(o->m, m, 4 * sizeof(rc128));
    }
    ++*cnt;
}

/* qft.rs:54-62 with apply_qft_no_swap 94-105 */
size_t ref_gen_qft(size_t n, ref_op *out)
{
    size_t cnt = 0;
    rc128 m[4], sw[16];
    for (size_t jj = n; jj-- > 0;) {
        ref_gate_h(m);
        put1(out, &cnt, m, jj);
        for (size_t k = jj; k-- > 0;) {
            uint32_t mm = (uint32_t)(jj - k + 1);
            double angle = TAU / (double)(1u << mm);
            ref_gate_phase(angle, m);
            putc1(out, &cnt, k, m, jj);
        }
    }
    ref_gate_swap(sw);
    for (size_t i = 0; i < n / 2; ++i) put2(out, &cnt, sw, i, n - 1 - i);
    return cnt;
}

/* Brick-wall layer generator (workload definition, not in the reference):
 * per layer, for q = 0..n-1 draw thx, thy, thz = TAU*next_f64() in that order
 * and append rx, ry, rz on q; then CNOT(a, a+1) for a = (layer % 2), +2, ... */
size_t ref_gen_brickwall(size_t n, size_t depth, ref_rng *rng, ref_op *out)
{
    size_t cnt = 0;
    rc128 m[4], cx[16];
    ref_gate_cnot(cx);
    for (size_t l = 0; l < depth; ++l) {
        for (size_t q = 0; q < n; ++q) {
            double tx = TAU * ref_rng_next_f64(rng);
            double ty = TAU * ref_rng_next_f64(rng);
            double tz = TAU * ref_rng_next_f64(rng);
            ref_gate_rx(tx, m); put1(out, &cnt, m, q);
            ref_gate_ry(ty, m); put1(out, &cnt, m, q);
            ref_gate_rz(tz, m); put1(out, &cnt, m, q);
        }
        for (size_t a = l % 2; a + 1 < n; a += 2) put2(out, &cnt, cx, a, a + 1);
    }
    return cnt;
}

static void gen_dm_1q(ref_rng *rng, uint64_t kind, size_t q, ref_op *out, size_t *cnt)
{
    rc128 m[4];
    switch (kind) {
    case 0: ref_gate_h(m); break;
    case 1: ref_gate_s(m); break;
    case 2: ref_gate_t(m); break;
    case 3: ref_gate_rx(TAU * ref_rng_next_f64(rng), m); break;
    default: ref_gate_rz(TAU * ref_rng_next_f64(rng), m); break;
    }
    put1(out, cnt, m, q);
}

/* Noisy density-matrix workload (workload definition, not in the reference):
 * per layer, Fisher-Yates shuffle of 0..n-1 (i from n-1 down to 1, j =
 * next_u64() % (i+1)); for each consecutive pair (a, b) draw r = next_u64() % 6:
 * r == 5 -> CNOT(a, b); else 1q kind r on a, then kind next_u64() % 5 on b.
 * An odd qubit left over gets kind next_u64() % 5.  Kinds: 0 H, 1 S, 2 T,
 * 3 Rx(TAU*next_f64()), 4 Rz(TAU*next_f64()). */
size_t ref_gen_dm_random(size_t n, size_t depth, ref_rng *rng, ref_op *out)
{
    size_t cnt = 0;
    rc128 cx[16];
    ref_gate_cnot(cx);
    size_t perm[64];
    for (size_t l = 0; l < depth; ++l) {
        for (size_t q = 0; q < n; ++q) perm[q] = q;
        for (size_t i = n - 1; i >= 1; --i) {
            size_t j = (size_t)(ref_rng_next_u64(rng) % (uint64_t)(i + 1));
            size_t t = perm[i]; perm[i] = perm[j]; perm[j] = t;
        }
        size_t p = 0;
        for (; p + 1 < n; p += 2) {
            uint64_t r = ref_rng_next_u64(rng) % 6;
            if (r == 5) {
                put2(out, &cnt, cx, perm[p], perm[p + 1]);
            } else {
                gen_dm_1q(rng, r, perm[p], out, &cnt);
                gen_dm_1q(rng, ref_rng_next_u64(rng) % 5, perm[p + 1], out, &cnt);
            }
        }
        if (p < n) gen_dm_1q(rng, ref_rng_next_u64(rng) % 5, perm[p], out, &cnt);
    }
    return cnt;
}

/* ------------------------------------------- Born-rule sampler (workload) */

/* Definition shared with the product's sampler (no reference counterpart:
 * the reference only measures one qubit at a time, measure.rs:15-51).
 * Chunks of C = 2^cl amplitudes, cl = n for n < 10, else max(10, n - 16)
 * (at most 2^16 chunks); total[c] = sequential sum of
 * norm_sqr over the chunk; prefix = sequential inclusive prefix of totals.
 * A shot with uniform u targets u * prefix[last], picks the first chunk with
 * prefix > target (the last chunk if none), then the first index in it whose
 * running sum prefix[c-1] + sequential in-chunk sum exceeds the target; if
 * none does, the last index with nonzero weight in the chunk, else the
 * chunk's last index. */
void ref_sample(const rc128 *amps, size_t n, const double *u, size_t shots, uint64_t *out)
{
    size_t cl = n < 10 ? n : (n > 26 ? n - 16 : 10), csize = (size_t)1 << cl, nch = (size_t)1 << (n - cl);
    double *pre = malloc(nch * sizeof(double));
    for (size_t c = 0; c < nch; ++c) {
        double s = 0.0;
        for (size_t j = 0; j < csize; ++j) s += ref_norm_sqr(amps[c * csize + j]);
        pre[c] = s;
    }
    double run = 0.0;
    for (size_t c = 0; c < nch; ++c) { run += pre[c]; pre[c] = run; }
    for (size_t s = 0; s < shots; ++s) {
        double target = u[s] * pre[nch - 1];
        size_t lo = 0, hi = nch - 1;
        while (lo < hi) {
            size_t mid = (lo + hi) / 2;
            if (pre[mid] > target) hi = mid; else lo = mid + 1;
        }
        double r = lo > 0 ? pre[lo - 1] : 0.0;
        uint64_t pick = lo * csize + csize - 1, last_nz = UINT64_MAX;
        int found = 0;
        for (size_t j = 0; j < csize; ++j) {
            double w = ref_norm_sqr(amps[lo * csize + j]);
            if (w > 0.0) last_nz = lo * csize + j;
            r += w;
            if (r > target) { pick = lo * csize + j; found = 1; break; }
        }
        if (!found && last_nz != UINT64_MAX) pick = last_nz;
        out[s] = pick;
    }
    free(pre);
}
