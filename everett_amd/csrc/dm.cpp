// Density-matrix backend on the device (src/backend/density_matrix.rs).
//
// rho (row-major dim x dim, dim = 2^n) is stored as a 2n-qubit c128 vector:
// element (i, j) at (i << n) | j, so row qubit t is vector bit n+t and column
// qubit t is vector bit t.  Then
//   U rho U^dag on qubit t      = U on bit n+t and conj(U) on bit t,
//   sum_k K rho K^dag on t      = S = sum_k K (x) conj(K) on bits (n+t, t),
// and a one-qubit gate followed by its channel is the single 4x4
// superoperator sum_k (K_k U) (x) conj(K_k U): one HBM pass instead of the
// reference's 2 + 2K sweeps with K+1 fresh 4^n buffers (density.rs:164-216).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <string>
#include <vector>

#include "dm_kernels.hpp"
#include "dm_ops.hpp"
#include "engine.hpp"
#include "planner.hpp"

namespace ev {

// A (x) conj(A) in the |row col> basis (row bit = MSB):
// S[(2r+c), (2r'+c')] = A[r][r'] * conj(A[c][c']).
Mat4 superop_of(const Mat2 &A)
{
    Mat4 S;
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c)
            for (int rp = 0; rp < 2; ++rp)
                for (int cp = 0; cp < 2; ++cp) S[(2 * r + c) * 4 + (2 * rp + cp)] = cmul(A[2 * r + rp], cconj(A[2 * c + cp]));
    return S;
}

// sum_k (K_k U) (x) conj(K_k U); with no Kraus operators, U (x) conj(U).
Mat4 gate_channel_superop(const Mat2 &U, const std::vector<Mat2> &kraus)
{
    if (kraus.empty()) return superop_of(U);
    Mat4 S{};
    for (const Mat2 &K : kraus) {
        const Mat4 t = superop_of(gate::compose(K, U));
        for (int i = 0; i < 16; ++i) S[i] = cadd(S[i], t[i]);
    }
    return S;
}

Mat4 channel_superop(const std::vector<Mat2> &kraus) { return gate_channel_superop(gate::id(), kraus); }

Mat4 conj4(const c128 *m)
{
    Mat4 o;
    for (int i = 0; i < 16; ++i) o[i] = cconj(m[i]);
    return o;
}

Mat2 conj2(const c128 *m) { return {cconj(m[0]), cconj(m[1]), cconj(m[2]), cconj(m[3])}; }

ev_dm *dm_create(int n)
{
    EV_REQUIRE(n >= 0 && 2 * uint64_t(n) <= kMaxQubits, EV_E_LIMIT,
               "density matrix of " + std::to_string(n) + " qubits exceeds this build's limit of 17");
    ev_dm *d = new ev_dm;
    d->n = n;
    try {
        d->vec = sv_create(2 * n, kern::C128);  // density.rs:45-55: rho[0][0] = 1
    } catch (...) {
        delete d;
        throw;
    }
    return d;
}

void dm_destroy(ev_dm *d)
{
    if (!d) return;
    sv_destroy(d->vec);
    delete d;
}

static double *dslots(ev_dm *d) { return d->vec->d_red + kRedPartials; }

static void dfetch(ev_dm *d, int count)
{
    ev_sv *s = d->vec;
    EV_HIP(hipMemcpyAsync(s->h_red, dslots(d), sizeof(double) * count, hipMemcpyDeviceToHost, s->stream));
    sv_sync(s);
}

double dm_diag_sum(ev_dm *d, uint64_t mask, uint64_t want)
{
    kern::dm_diag_sum(d->vec->psi, d->n, mask, want, d->vec->d_red, dslots(d), d->vec->stream);
    dfetch(d, 1);
    return d->vec->h_red[0];
}

void dm_partial_trace(ev_dm *d, int q, c128 out[4])
{
    kern::dm_partial_trace(d->vec->psi, d->n, q, d->vec->d_red, dslots(d), d->vec->stream);
    dfetch(d, 8);
    for (int i = 0; i < 4; ++i) out[i] = {d->vec->h_red[2 * i], d->vec->h_red[2 * i + 1]};
}

// density.rs:219-227
void dm_renormalize(ev_dm *d)
{
    const double t = dm_diag_sum(d, 0, 0);
    if (t > 0.0) kern::scale(d->vec->psi, kern::C128, 2 * d->n, 1.0 / t, d->vec->stream);
}

// density_matrix.rs:155-169: p1 from the diagonal, one draw, collapse,
// renormalise, then the readout flip (second draw only if readout_error > 0).
// The trace after the collapse is the sum of the kept diagonal entries, taken
// before it in the same order (a zeroed entry adds +0 to a partial sum and
// changes nothing), so the collapse and the 1/trace scaling run as one pass.
// For outcome 1 that sum is p1 itself.
bool dm_measure(ev_dm *d, int q, Rng &rng, double readout_error)
{
    const uint64_t bit = uint64_t(1) << q;
    const double p1 = dm_diag_sum(d, bit, bit);
    const bool raw = rng.next_f64() < p1;
    const double t = raw ? p1 : dm_diag_sum(d, bit, 0);
    if (t > 0.0) kern::dm_collapse_scale(d->vec->psi, d->n, q, raw ? 1 : 0, 1.0 / t, d->vec->stream);
    else kern::dm_collapse(d->vec->psi, d->n, q, raw ? 1 : 0, d->vec->stream);
    if (readout_error > 0.0 && rng.next_f64() < readout_error) return !raw;
    return raw;
}

// A run of consecutive Measure ops (density_matrix.rs:155-169 for each, in
// order).  Each measurement's probabilities come from a device copy of the
// diagonal that is collapsed and rescaled as the matrix would be (the same
// partial sums, so the same p1, draws and outcomes as measuring one by one),
// and the matrix itself takes one pass for the whole run: the entries that
// match every outcome, times the scales in order.  Returns the index after
// the run.
size_t dm_measure_run(ev_dm *d, const Circuit &c, size_t i, Rng &rng, double readout_error, uint8_t *classical)
{
    ev_sv *v = d->vec;
    const int n = d->n;
    const uint64_t dim = uint64_t(1) << n;
    double *diag = nullptr;
    EV_HIP(hipMallocAsync(reinterpret_cast<void **>(&diag), sizeof(double) * dim, v->stream));
    try {
        kern::dm_extract_diag(v->psi, n, diag, v->stream);
        kern::DmScales sc{};
        uint64_t keep_mask = 0, keep_val = 0;
        bool none = false;
        auto diag_sum = [&](uint64_t mask, uint64_t want) {
            kern::diag_vec_sum(diag, n, mask, want, v->d_red, dslots(d), v->stream);
            dfetch(d, 1);
            return v->h_red[0];
        };
        while (i < c.size() && c.ops()[i].kind == OpKind::Measure && sc.count < kern::kMaxBlockMeasures) {
            const Op &op = c.ops()[i];
            const int q = int(op.q0);
            const uint64_t bit = uint64_t(1) << q;
            const double p1 = diag_sum(bit, bit);
            const bool raw = rng.next_f64() < p1;
            const double t = raw ? p1 : diag_sum(bit, 0);
            const int o = raw ? 1 : 0;
            const uint64_t both = bit | (bit << n), val = raw ? both : 0;
            if ((keep_mask & both) && (keep_val & both) != val) none = true;  // an outcome contradicting an earlier one
            keep_mask |= both;
            keep_val = (keep_val & ~both) | val;
            if (t > 0.0) sc.s[sc.count++] = 1.0 / t;
            kern::diag_collapse(diag, n, q, o, t > 0.0 ? 1.0 / t : 1.0, t > 0.0, v->stream);
            bool bitv = raw;
            if (readout_error > 0.0 && rng.next_f64() < readout_error) bitv = !raw;
            classical[op.q1] = bitv ? 1 : 0;
            ++i;
        }
        kern::dm_collapse_block(v->psi, n, keep_mask, keep_val, none, sc, v->stream);
    } catch (...) {
        (void)hipFreeAsync(diag, v->stream);
        throw;
    }
    EV_HIP(hipFreeAsync(diag, v->stream));
    return i;
}

DmNoise make_dm_noise(const ev_noise_model *noise)
{
    DmNoise nz;
    if (noise) {
        nz.k1 = channel_kraus(noise->after_1q);
        nz.k2 = channel_kraus(noise->after_2q);
        nz.readout_error = noise->readout_error;
    }
    nz.s1 = channel_superop(nz.k1);
    nz.s2 = channel_superop(nz.k2);
    return nz;
}

// backend/mod.rs:51-72 over the density-matrix primitives, one pass per
// superoperator.
size_t dm_plain_op(ev_dm *d, const Circuit &c, size_t i, const DmNoise &nz, Rng &rng, uint8_t *classical, bool execute,
                   RunStats &st)
{
    const Op &op = c.ops()[i];
    ev_sv *v = d->vec;
    const int n = d->n;
    switch (op.kind) {
    case OpKind::Apply1:  // density_matrix.rs:117-125
        if (execute) {
            const c128 *m = c.mat(op);
            const Mat4 S = gate_channel_superop(Mat2{m[0], m[1], m[2], m[3]}, nz.k1);
            sv_apply2(v, n + int(op.q0), int(op.q0), S.data());
            ++st.logical_gates;
            ++st.launches;
        }
        return i + 1;
    case OpKind::Apply2:  // density_matrix.rs:127-138
        if (execute) {
            const int a = int(op.q0), b = int(op.q1);
            sv_apply2(v, n + a, n + b, c.mat(op));
            sv_apply2(v, a, b, conj4(c.mat(op)).data());
            st.launches += 2;
            if (!nz.k2.empty()) {
                sv_apply2(v, n + a, a, nz.s2.data());
                sv_apply2(v, n + b, b, nz.s2.data());
                st.launches += 2;
            }
            ++st.logical_gates;
        }
        return i + 1;
    case OpKind::Controlled:  // density_matrix.rs:140-153
        if (execute) {
            const uint64_t *cs = c.ctrls(op);
            std::vector<uint64_t> rows(cs, cs + op.nctrl);
            for (auto &q : rows) q += uint64_t(n);
            sv_ctrl1(v, rows.data(), int(op.nctrl), n + int(op.q0), c.mat(op));
            sv_ctrl1(v, cs, int(op.nctrl), int(op.q0), conj2(c.mat(op)).data());
            st.launches += 2;
            if (!nz.k1.empty()) {
                sv_apply2(v, n + int(op.q0), int(op.q0), nz.s1.data());
                for (uint32_t k = 0; k < op.nctrl; ++k) sv_apply2(v, n + int(cs[k]), int(cs[k]), nz.s1.data());
                st.launches += 1 + op.nctrl;
            }
            ++st.logical_gates;
        }
        return i + 1;
    case OpKind::Measure:
        if (execute) classical[op.q1] = dm_measure(d, int(op.q0), rng, nz.readout_error) ? 1 : 0;
        return i + 1;
    case OpKind::If: return dm_plain_op(d, c, i + 1, nz, rng, classical, execute && classical[op.q0] != 0, st);
    }
    throw Error(EV_E_ARG, "corrupt op record");
}

void dm_execute(ev_dm *d, const Circuit &c, const ev_noise_model *noise, Rng &rng, uint8_t *classical,
                const ev_run_options &opt)
{
    const DmNoise nz = make_dm_noise(noise);
    RunStats st;
    hipEvent_t a, b;
    EV_HIP(hipEventCreate(&a));
    EV_HIP(hipEventCreate(&b));
    EV_HIP(hipEventRecord(a, d->vec->stream));
    try {
        if (opt.plain) {
            size_t i = 0;
            while (i < c.size()) i = dm_plain_op(d, c, i, nz, rng, classical, true, st);
        } else {
            execute_planned_dm(d, c, nz, rng, classical, opt, st);
        }
        EV_HIP(hipEventRecord(b, d->vec->stream));
        EV_HIP(hipEventSynchronize(b));
        float ms = 0;
        EV_HIP(hipEventElapsedTime(&ms, a, b));
        st.device_ms = ms;
    } catch (...) {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        throw;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    last_stats() = st;
}

}  // namespace ev
