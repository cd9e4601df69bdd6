// Host mirror of src/rng.rs, src/gate.rs, src/circuit.rs, src/noise.rs and
// src/algorithms/qft.rs, plus the benchmark workload generators.
#include "circuit.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

namespace ev {

// ------------------------------------------------------------------ rng.rs

Rng Rng::seed_from_u64(uint64_t seed)
{
    Rng r{};
    for (auto &w : r.s) {  // SplitMix64, rng.rs:20-26
        seed += 0x9E3779B97F4A7C15ull;
        uint64_t z = seed;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        w = z ^ (z >> 31);
    }
    return r;
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t Rng::next_u64()
{
    const uint64_t result = rotl64(s[0] + s[3], 23) + s[0];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

double Rng::next_f64() { return static_cast<double>(next_u64() >> 11) * (1.0 / static_cast<double>(1ull << 53)); }

// ----------------------------------------------------------------- gate.rs

namespace gate {
static constexpr double kFrac1Sqrt2 = 0.70710678118654752440;
static constexpr double kFracPi4 = 0.78539816339744830962;

Mat2 id() { return {c128{1, 0}, c128{0, 0}, c128{0, 0}, c128{1, 0}}; }
Mat2 x() { return {c128{0, 0}, c128{1, 0}, c128{1, 0}, c128{0, 0}}; }
Mat2 y() { return {c128{0, 0}, c128{0, -1}, c128{0, 1}, c128{0, 0}}; }
Mat2 z() { return {c128{1, 0}, c128{0, 0}, c128{0, 0}, c128{-1, 0}}; }
Mat2 h()
{
    const double s = kFrac1Sqrt2;
    return {c128{s, 0}, c128{s, 0}, c128{s, 0}, c128{-s, 0}};
}
Mat2 s() { return {c128{1, 0}, c128{0, 0}, c128{0, 0}, c128{0, 1}}; }
static c128 expi(double th) { return {std::cos(th), std::sin(th)}; }  // complex.rs:63-65
Mat2 t() { return {c128{1, 0}, c128{0, 0}, c128{0, 0}, expi(kFracPi4)}; }
Mat2 rx(double theta)
{
    const double sn = std::sin(theta / 2.0), co = std::cos(theta / 2.0);
    return {c128{co, 0}, c128{0, -sn}, c128{0, -sn}, c128{co, 0}};
}
Mat2 ry(double theta)
{
    const double sn = std::sin(theta / 2.0), co = std::cos(theta / 2.0);
    return {c128{co, 0}, c128{-sn, 0}, c128{sn, 0}, c128{co, 0}};
}
Mat2 rz(double theta) { return {expi(-theta / 2.0), c128{0, 0}, c128{0, 0}, expi(theta / 2.0)}; }
Mat2 phase(double lambda) { return {c128{1, 0}, c128{0, 0}, c128{0, 0}, expi(lambda)}; }
Mat2 compose(const Mat2 &a, const Mat2 &b)
{
    return {cadd(cmul(a[0], b[0]), cmul(a[1], b[2])), cadd(cmul(a[0], b[1]), cmul(a[1], b[3])),
            cadd(cmul(a[2], b[0]), cmul(a[3], b[2])), cadd(cmul(a[2], b[1]), cmul(a[3], b[3]))};
}
static Mat4 eye4()
{
    Mat4 m{};
    for (int i = 0; i < 4; ++i) m[i * 5] = c128{1, 0};
    return m;
}
Mat4 cnot()
{
    Mat4 m = eye4();
    m[10] = {0, 0};
    m[11] = {1, 0};
    m[14] = {1, 0};
    m[15] = {0, 0};
    return m;
}
Mat4 cz()
{
    Mat4 m = eye4();
    m[15] = {-1, 0};
    return m;
}
Mat4 swap()
{
    Mat4 m = eye4();
    m[5] = {0, 0};
    m[6] = {1, 0};
    m[9] = {1, 0};
    m[10] = {0, 0};
    return m;
}
}  // namespace gate

K1 classify1(const c128 m[4])
{
    const bool diag = m[1].re == 0 && m[1].im == 0 && m[2].re == 0 && m[2].im == 0;
    if (!diag) return K1::Dense;
    return (m[0].re == 1 && m[0].im == 0) ? K1::Phase : K1::Diag;
}

K2 classify2(const c128 m[16])
{
    static const Mat4 cx = gate::cnot(), z = gate::cz(), sw = gate::swap();
    auto same = [&](const Mat4 &g) {
        for (int i = 0; i < 16; ++i)
            if (!ceq(g[i], m[i])) return false;
        return true;
    };
    if (same(cx)) return K2::Cnot;
    if (same(z)) return K2::Cz;
    if (same(sw)) return K2::Swap;
    return K2::Dense;
}

// -------------------------------------------------------------- circuit.rs

void Circuit::apply1(const Mat2 &m, uint64_t target)
{
    ops_.push_back(Op{OpKind::Apply1, target, 0, 0, 0, mat_pool_.size()});
    mat_pool_.insert(mat_pool_.end(), m.begin(), m.end());
}

void Circuit::apply2(const Mat4 &m, uint64_t a, uint64_t b)
{
    ops_.push_back(Op{OpKind::Apply2, a, b, 0, 0, mat_pool_.size()});
    mat_pool_.insert(mat_pool_.end(), m.begin(), m.end());
}

void Circuit::controlled(const uint64_t *controls, uint32_t nc, const Mat2 &m, uint64_t target)
{
    ops_.push_back(Op{OpKind::Controlled, target, 0, static_cast<uint32_t>(ctrl_pool_.size()), nc, mat_pool_.size()});
    ctrl_pool_.insert(ctrl_pool_.end(), controls, controls + nc);
    mat_pool_.insert(mat_pool_.end(), m.begin(), m.end());
}

void Circuit::measure(uint64_t qubit, uint64_t into) { ops_.push_back(Op{OpKind::Measure, qubit, into, 0, 0, 0}); }

void Circuit::if_bit(uint64_t bit) { ops_.push_back(Op{OpKind::If, bit, 0, 0, 0, 0}); }

size_t Circuit::append(const ev_op *ops, size_t count, size_t i)
{
    const ev_op &op = ops[i];
    const c128 *m = reinterpret_cast<const c128 *>(op.m);
    switch (op.kind) {
    case EV_OP_APPLY1: apply1(Mat2{m[0], m[1], m[2], m[3]}, op.target); return i + 1;
    case EV_OP_APPLY2: {
        Mat4 g;
        for (int k = 0; k < 16; ++k) g[k] = m[k];
        apply2(g, op.target, op.b);
        return i + 1;
    }
    case EV_OP_CONTROLLED: {
        // Op::Controlled { controls: Vec<QubitId>, .. } (src/op.rs:34-36) has no
        // length limit: controls past the first EV_MAX_CONTROLS continue in the
        // EV_OP_CONTROLS records that follow.
        const size_t total = op.num_controls;
        std::vector<uint64_t> cs(op.controls, op.controls + std::min<size_t>(total, EV_MAX_CONTROLS));
        size_t j = i + 1;
        while (cs.size() < total) {
            EV_REQUIRE(j < count && ops[j].kind == EV_OP_CONTROLS, EV_E_ARG,
                       "controlled op with " + std::to_string(total) + " controls: only " + std::to_string(cs.size()) +
                           " before the next non-EV_OP_CONTROLS record");
            const size_t k = ops[j].num_controls;
            EV_REQUIRE(k >= 1 && k <= EV_MAX_CONTROLS && cs.size() + k <= total, EV_E_ARG,
                       "EV_OP_CONTROLS record with a control count that does not fit its op");
            cs.insert(cs.end(), ops[j].controls, ops[j].controls + k);
            ++j;
        }
        EV_REQUIRE(total <= UINT32_MAX, EV_E_LIMIT, "more than 2^32 - 1 controls on one op");
        controlled(cs.data(), static_cast<uint32_t>(total), Mat2{m[0], m[1], m[2], m[3]}, op.target);
        return j;
    }
    case EV_OP_MEASURE: measure(op.target, op.b); return i + 1;
    case EV_OP_IF: if_bit(op.target); return i + 1;
    case EV_OP_CONTROLS: throw Error(EV_E_ARG, "EV_OP_CONTROLS record that does not follow a controlled op");
    default: throw Error(EV_E_ARG, "unknown op kind " + std::to_string(op.kind));
    }
}

void Circuit::compose(const Circuit &sub)
{
    const uint64_t moff = mat_pool_.size();
    const uint32_t coff = static_cast<uint32_t>(ctrl_pool_.size());
    for (Op op : sub.ops_) {
        op.mat_off += moff;
        op.ctrl_off += coff;
        ops_.push_back(op);
    }
    mat_pool_.insert(mat_pool_.end(), sub.mat_pool_.begin(), sub.mat_pool_.end());
    ctrl_pool_.insert(ctrl_pool_.end(), sub.ctrl_pool_.begin(), sub.ctrl_pool_.end());
}

size_t Circuit::records(size_t i) const
{
    const Op &op = ops_[i];
    if (op.kind != OpKind::Controlled || op.nctrl <= EV_MAX_CONTROLS) return 1;
    return 1 + (op.nctrl - EV_MAX_CONTROLS + EV_MAX_CONTROLS - 1) / EV_MAX_CONTROLS;
}

size_t Circuit::record_count() const
{
    size_t n = 0;
    for (size_t i = 0; i < ops_.size(); ++i) n += records(i);
    return n;
}

size_t Circuit::to_ev(size_t i, ev_op *out) const
{
    const size_t nrec = records(i);
    std::memset(out, 0, nrec * sizeof(ev_op));
    const Op &op = ops_[i];
    out[0].kind = static_cast<uint32_t>(op.kind);
    out[0].target = op.q0;
    out[0].b = op.q1;
    size_t nm = op.kind == OpKind::Apply2 ? 16 : (op.kind == OpKind::Apply1 || op.kind == OpKind::Controlled ? 4 : 0);
    if (nm) std::memcpy(out[0].m, mat(op), nm * sizeof(c128));
    if (op.kind == OpKind::Controlled && op.nctrl) {
        out[0].num_controls = op.nctrl;  // the op's total; the records carry them EV_MAX_CONTROLS at a time
        const uint64_t *cs = ctrls(op);
        size_t done = std::min<size_t>(op.nctrl, EV_MAX_CONTROLS);
        std::memcpy(out[0].controls, cs, done * sizeof(uint64_t));
        for (size_t r = 1; r < nrec; ++r) {
            const size_t k = std::min<size_t>(op.nctrl - done, EV_MAX_CONTROLS);
            out[r].kind = EV_OP_CONTROLS;
            out[r].num_controls = static_cast<uint32_t>(k);
            std::memcpy(out[r].controls, cs + done, k * sizeof(uint64_t));
            done += k;
        }
    }
    return nrec;
}

static void check_qubit(uint64_t q, uint64_t nq)  // circuit.rs:288-296
{
    if (q >= nq)
        throw Error(EV_E_QUBIT_RANGE,
                    "qubit " + std::to_string(q) + " out of range for a " + std::to_string(nq) + "-qubit register", q, nq);
}

static void check_classical(uint64_t b, uint64_t ncl)  // circuit.rs:298-306
{
    if (b >= ncl)
        throw Error(EV_E_CBIT_RANGE,
                    "classical bit " + std::to_string(b) + " out of range for a " + std::to_string(ncl) +
                        "-bit classical register",
                    b, ncl);
}

// circuit.rs:252-285 validate_op
size_t Circuit::validate_at(size_t i) const
{
    EV_REQUIRE(i < ops_.size(), EV_E_ARG, "IfClassic record at the end of the op list guards nothing");
    const Op &op = ops_[i];
    switch (op.kind) {
    case OpKind::Apply1: check_qubit(op.q0, nq_); return i + 1;
    case OpKind::Apply2:
        check_qubit(op.q0, nq_);
        check_qubit(op.q1, nq_);
        if (op.q0 == op.q1) throw Error(EV_E_DUP_QUBIT, "qubit " + std::to_string(op.q0) + " used twice", op.q0);
        return i + 1;
    case OpKind::Controlled: {
        check_qubit(op.q0, nq_);
        const uint64_t *cs = ctrls(op);
        for (uint32_t c = 0; c < op.nctrl; ++c) {
            check_qubit(cs[c], nq_);
            if (cs[c] == op.q0) throw Error(EV_E_DUP_QUBIT, "qubit " + std::to_string(cs[c]) + " used twice", cs[c]);
        }
        return i + 1;
    }
    case OpKind::Measure:
        check_qubit(op.q0, nq_);
        check_classical(op.q1, ncl_);
        return i + 1;
    case OpKind::If:
        check_classical(op.q0, ncl_);
        return validate_at(i + 1);
    }
    throw Error(EV_E_ARG, "corrupt op record");
}

void Circuit::validate() const
{
    size_t i = 0;
    while (i < ops_.size()) i = validate_at(i);
}

// ------------------------------------------------------ algorithms/qft.rs

static constexpr double kTau = 6.28318530717958647692;

Circuit gen_qft(uint64_t n)
{
    Circuit c(n, 0);
    for (uint64_t j = n; j-- > 0;) {  // qft.rs:94-105
        c.apply1(gate::h(), j);
        for (uint64_t k = j; k-- > 0;) {
            const uint32_t m = static_cast<uint32_t>(j - k + 1);
            const double angle = kTau / static_cast<double>(1u << m);
            c.controlled(&k, 1, gate::phase(angle), j);
        }
    }
    for (uint64_t i = 0; i < n / 2; ++i) c.apply2(gate::swap(), i, n - 1 - i);  // qft.rs:57-60
    return c;
}

Circuit gen_iqft(uint64_t n)
{
    Circuit c(n, 0);
    for (uint64_t i = 0; i < n / 2; ++i) c.apply2(gate::swap(), i, n - 1 - i);  // qft.rs:85-88
    for (uint64_t j = 0; j < n; ++j) {                                          // qft.rs:108-118
        for (uint64_t k = 0; k < j; ++k) {
            const uint32_t m = static_cast<uint32_t>(j - k + 1);
            const double angle = -kTau / static_cast<double>(1u << m);
            c.controlled(&k, 1, gate::phase(angle), j);
        }
        c.apply1(gate::h(), j);
    }
    return c;
}

// Brick-wall workload: per layer, for q = 0..n-1 draw thx, thy, thz =
// TAU * next_f64() (in that order) and append rx, ry, rz on q; then
// CNOT(a, a+1) for a = layer % 2, layer % 2 + 2, ...
Circuit gen_brickwall(uint64_t n, uint64_t depth, Rng &rng)
{
    Circuit c(n, 0);
    const Mat4 cx = gate::cnot();
    for (uint64_t l = 0; l < depth; ++l) {
        for (uint64_t q = 0; q < n; ++q) {
            const double tx = kTau * rng.next_f64();
            const double ty = kTau * rng.next_f64();
            const double tz = kTau * rng.next_f64();
            c.apply1(gate::rx(tx), q);
            c.apply1(gate::ry(ty), q);
            c.apply1(gate::rz(tz), q);
        }
        for (uint64_t a = l % 2; a + 1 < n; a += 2) c.apply2(cx, a, a + 1);
    }
    return c;
}

static void dm_1q(Circuit &c, Rng &rng, uint64_t kind, uint64_t q)
{
    switch (kind) {
    case 0: c.apply1(gate::h(), q); break;
    case 1: c.apply1(gate::s(), q); break;
    case 2: c.apply1(gate::t(), q); break;
    case 3: c.apply1(gate::rx(kTau * rng.next_f64()), q); break;
    default: c.apply1(gate::rz(kTau * rng.next_f64()), q); break;
    }
}

// Noisy density-matrix workload: per layer, a Fisher-Yates shuffle of 0..n-1
// (i = n-1 down to 1, j = next_u64() % (i+1)); each consecutive pair (a, b)
// draws r = next_u64() % 6: r == 5 gives CNOT(a, b), otherwise 1q kind r on a
// and then kind next_u64() % 5 on b; a leftover qubit gets kind
// next_u64() % 5.  Kinds: 0 H, 1 S, 2 T, 3 Rx(TAU*u), 4 Rz(TAU*u).
Circuit gen_dm_random(uint64_t n, uint64_t depth, Rng &rng)
{
    Circuit c(n, 0);
    const Mat4 cx = gate::cnot();
    std::vector<uint64_t> perm(n);
    for (uint64_t l = 0; l < depth; ++l) {
        for (uint64_t q = 0; q < n; ++q) perm[q] = q;
        for (uint64_t i = n - 1; i >= 1 && n > 0; --i) {
            const uint64_t j = rng.next_u64() % (i + 1);
            std::swap(perm[i], perm[j]);
        }
        uint64_t p = 0;
        for (; p + 1 < n; p += 2) {
            const uint64_t r = rng.next_u64() % 6;
            if (r == 5) {
                c.apply2(cx, perm[p], perm[p + 1]);
            } else {
                dm_1q(c, rng, r, perm[p]);
                const uint64_t kb = rng.next_u64() % 5;
                dm_1q(c, rng, kb, perm[p + 1]);
            }
        }
        if (p < n) {
            const uint64_t k = rng.next_u64() % 5;
            dm_1q(c, rng, k, perm[p]);
        }
    }
    return c;
}

// ---------------------------------------------------------------- noise.rs

std::vector<Mat2> channel_kraus(const ev_channel &ch)
{
    const c128 z{0, 0};
    const double p = ch.p;
    switch (ch.kind) {
    case EV_CH_NONE: return {};
    case EV_CH_DEPOLARIZING: {
        const double a = std::sqrt(1.0 - p), b = std::sqrt(p / 3.0);
        return {Mat2{c128{a, 0}, z, z, c128{a, 0}}, Mat2{z, c128{b, 0}, c128{b, 0}, z},
                Mat2{z, c128{0, -b}, c128{0, b}, z}, Mat2{c128{b, 0}, z, z, c128{-b, 0}}};
    }
    case EV_CH_AMPLITUDE_DAMPING: {
        const double k0 = std::sqrt(1.0 - p), k1 = std::sqrt(p);
        return {Mat2{c128{1, 0}, z, z, c128{k0, 0}}, Mat2{z, c128{k1, 0}, z, z}};
    }
    case EV_CH_PHASE_DAMPING: {
        const double k0 = std::sqrt(1.0 - p), k1 = std::sqrt(p);
        return {Mat2{c128{1, 0}, z, z, c128{k0, 0}}, Mat2{z, z, z, c128{k1, 0}}};
    }
    case EV_CH_BIT_FLIP: {
        const double a = std::sqrt(1.0 - p), b = std::sqrt(p);
        return {Mat2{c128{a, 0}, z, z, c128{a, 0}}, Mat2{z, c128{b, 0}, c128{b, 0}, z}};
    }
    case EV_CH_DEPHASING: {
        const double a = std::sqrt(1.0 - p), b = std::sqrt(p);
        return {Mat2{c128{a, 0}, z, z, c128{a, 0}}, Mat2{c128{b, 0}, z, z, c128{-b, 0}}};
    }
    case EV_CH_CUSTOM: {
        EV_REQUIRE(ch.num_kraus == 0 || ch.kraus != nullptr, EV_E_ARG, "custom channel without Kraus operators");
        std::vector<Mat2> out(ch.num_kraus);
        const c128 *k = reinterpret_cast<const c128 *>(ch.kraus);
        for (size_t i = 0; i < ch.num_kraus; ++i) out[i] = Mat2{k[4 * i], k[4 * i + 1], k[4 * i + 2], k[4 * i + 3]};
        return out;
    }
    default: throw Error(EV_E_ARG, "unknown channel kind " + std::to_string(ch.kind));
    }
}

}  // namespace ev
