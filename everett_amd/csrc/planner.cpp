// Fused execution: the gate planner and the pass executor.
//
// 1. Lowering.  Circuit ops become planner ops (GOp).  Consecutive one-qubit
//    gates on a qubit are multiplied into one 2x2 in f64 (gate::compose), a
//    diagonal product becomes a phase op that needs no tile residency, CNOT
//    and multi-controlled X become register swaps, CZ and controlled phases
//    become conditional diagonals.
// 2. Passes.  Greedy in program order: a pass owns a set Q of t tile qubits
//    (always containing the lowest qubits); an op joins the pass if the qubits
//    it mixes fit in Q, none of its qubits was touched by an op already
//    deferred, and the pass's compute budget allows; otherwise it is deferred.
//    Ops on disjoint qubits commute, so the deferred ones keep their order
//    relative to everything they touch.
// 3. Stages.  The same greedy rule inside a pass with 4 register qubits.
//
// The result runs as one kernel launch per pass (fused_kernels.hip).
#include "planner.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <set>
#include <vector>

#include "dm_ops.hpp"
#include "fused.hpp"
#include "fused_kernels.hpp"
#include "tune.hpp"

namespace ev {

namespace {

using fused::DevOp;
using fused::kRegBits;
using fused::kRegs;
using fused::Pass;
using fused::Stage;

enum GKind : uint8_t { G_U1, G_CU1, G_CX, G_DIAG, G_U2, G_SWAP };

struct GOp {
    GKind kind;
    int q0, q1;      // target(s); U2: q0 = MSB of the gate basis
    uint64_t cmask;  // control qubits
    uint32_t coef;   // offset into the c128 pool (U1/CU1: 4, DIAG: 2, U2: 16 entries)
    float cost;
    uint64_t touched() const
    {
        uint64_t m = cmask | (uint64_t(1) << q0);
        if (kind == G_U2 || kind == G_SWAP) m |= uint64_t(1) << q1;
        return m;
    }
    // Qubits that must be tile (and register) qubits when the op runs.
    uint64_t need() const
    {
        switch (kind) {
        case G_DIAG: return 0;
        case G_U2:
        case G_SWAP: return (uint64_t(1) << q0) | (uint64_t(1) << q1);
        default: return uint64_t(1) << q0;
        }
    }
    // need() plus, for a singly-controlled op, its control: with both in
    // registers the pair can fuse with its neighbours into one 4x4.  Taken
    // where it fits, else the op runs on need() alone.
    uint64_t want() const
    {
        if ((kind == G_CX || kind == G_CU1) && __builtin_popcountll(cmask) == 1) return need() | cmask;
        return need();
    }
};

// Compute cost in units of one dense 2x2 over the tile.
float op_cost(GKind k, bool d0_one)
{
    switch (k) {
    case G_U1:
    case G_CU1: return 1.0f;
    case G_U2: return 2.0f;
    case G_DIAG: return d0_one ? 0.25f : 0.5f;
    default: return 0.1f;
    }
}

// Lowers circuit ops into GOps, merging one-qubit runs per qubit.
class Lowering {
  public:
    explicit Lowering(int n) : pend_(n), has_(n, false) {}

    std::vector<GOp> ops;
    std::vector<c128> pool;
    uint64_t logical = 0;

    void gate1(int q, const c128 *m)
    {
        const Mat2 g{m[0], m[1], m[2], m[3]};
        pend_[q] = has_[q] ? gate::compose(g, pend_[q]) : g;
        has_[q] = true;
    }

    void flush(int q)
    {
        if (!has_[q]) return;
        has_[q] = false;
        const Mat2 &m = pend_[q];
        const Mat2 id = gate::id();
        bool is_id = true;
        for (int i = 0; i < 4; ++i) is_id = is_id && ceq(m[i], id[i]);
        if (is_id) return;
        if (classify1(m.data()) != K1::Dense) {
            diag(q, 0, m[0], m[3]);
        } else {
            push(G_U1, q, 0, 0, m.data(), 4);
        }
    }

    void flush_all()
    {
        for (int q = 0; q < int(has_.size()); ++q) flush(q);
    }

    void gate2(int a, int b, const c128 *m)
    {
        flush(a);
        flush(b);
        switch (classify2(m)) {
        case K2::Cnot: push(G_CX, b, 0, uint64_t(1) << a, nullptr, 0); break;
        case K2::Cz: diag(b, uint64_t(1) << a, c128{1, 0}, c128{-1, 0}); break;
        case K2::Swap: push(G_SWAP, std::max(a, b), std::min(a, b), 0, nullptr, 0); break;
        case K2::Dense: push(G_U2, a, b, 0, m, 16); break;
        }
    }

    void controlled(const uint64_t *cs, int nc, int t, const c128 *m)
    {
        flush(t);
        uint64_t mask = 0;
        for (int i = 0; i < nc; ++i) {
            flush(int(cs[i]));
            mask |= uint64_t(1) << cs[i];
        }
        const Mat2 x = gate::x();
        bool is_x = true;
        for (int i = 0; i < 4; ++i) is_x = is_x && ceq(m[i], x[i]);
        if (classify1(m) != K1::Dense) {
            diag(t, mask, m[0], m[3]);
        } else if (is_x) {
            push(G_CX, t, 0, mask, nullptr, 0);
        } else {
            push(mask ? G_CU1 : G_U1, t, 0, mask, m, 4);
        }
    }

    // Row/column superoperator (DM): dense 4x4 on (a = MSB, b).
    void dense2(int a, int b, const c128 *m)
    {
        flush(a);
        flush(b);
        push(G_U2, a, b, 0, m, 16);
    }

  private:
    void diag(int q, uint64_t mask, c128 d0, c128 d1)
    {
        GOp g{G_DIAG, q, 0, mask, uint32_t(pool.size()), 0};
        pool.push_back(d0);
        pool.push_back(d1);
        g.cost = op_cost(G_DIAG, d0.re == 1 && d0.im == 0);
        ops.push_back(g);
    }
    void push(GKind k, int q0, int q1, uint64_t mask, const c128 *m, int nm)
    {
        GOp g{k, q0, q1, mask, uint32_t(pool.size()), op_cost(k, false)};
        if (nm) pool.insert(pool.end(), m, m + nm);
        ops.push_back(g);
    }
    std::vector<Mat2> pend_;
    std::vector<bool> has_;
};

int popc(uint64_t x) { return __builtin_popcountll(x); }

int rank_in(uint64_t mask, int bit) { return popc(mask & ((uint64_t(1) << bit) - 1)); }

// Builds the pass list for `g` on an n-qubit state.
// ---- pair fusion inside a stage ------------------------------------------
//
// Every op in a stage costs a dispatch and register shuffling in the kernel
// beyond its arithmetic, so ops confined to two register qubits are fused
// into one dense 4x4 (G_U2) when that does not add arithmetic: a 4x4 costs 16
// FMA per amplitude, the same as two 2x2s, so U1(a) U1(b) CX(a,b) -- the
// brick-wall pattern -- becomes one op.  Greedy in stage order, with one
// "open" block per register qubit (the latest block touching it): an op on
// qubits whose open blocks can take it is multiplied into them; ops that do
// not touch a qubit commute with it, so absorbing moves nothing across a
// dependency.  Products are formed in f64 on the host.

// Cost of an op in the kernel, in FMA per amplitude plus a per-op overhead.
constexpr float kOpOverhead = 4.0f;
float fused_cost(GKind k)
{
    switch (k) {
    case G_U1:
    case G_CU1: return 8.0f + kOpOverhead;
    case G_U2: return 16.0f + kOpOverhead;
    case G_DIAG: return 4.0f + kOpOverhead;
    default: return 2.0f + kOpOverhead;  // register moves
    }
}

Mat4 mul4(const Mat4 &a, const Mat4 &b)  // a * b
{
    Mat4 r{};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            c128 acc{0, 0};
            for (int k = 0; k < 4; ++k) acc = cadd(acc, cmul(a[i * 4 + k], b[k * 4 + j]));
            r[i * 4 + j] = acc;
        }
    return r;
}

// A 2x2 on the MSB (hi = true) or LSB of a pair, as a 4x4.
Mat4 lift(const Mat2 &m, bool hi)
{
    Mat4 r{};
    for (int ra = 0; ra < 2; ++ra)
        for (int rb = 0; rb < 2; ++rb)
            for (int ca = 0; ca < 2; ++ca)
                for (int cb = 0; cb < 2; ++cb) {
                    c128 v{0, 0};
                    if (hi && rb == cb) v = m[ra * 2 + ca];
                    if (!hi && ra == ca) v = m[rb * 2 + cb];
                    r[(ra * 2 + rb) * 4 + ca * 2 + cb] = v;
                }
    return r;
}

struct Block {
    enum Kind { Single, Pair, Opaque, Dead } kind;
    int a = -1, b = -1;        // Single: a; Pair: a = MSB, b = LSB of the basis
    Mat2 m2{};
    Mat4 m4{};
    std::vector<int> members;  // original ops, in program order
};

class StageFuser {
  public:
    StageFuser(const std::vector<GOp> &g, const std::vector<c128> &pool, uint64_t regq) : g_(g), pool_(pool), regq_(regq)
    {
    }

    void add(int id) { ids_.push_back(id); }

    // CNOTs and SWAPs between register qubits that the block choice leaves
    // bare (neither opening a block nor merging into the block of the same
    // pair just before them), in stage order.
    //
    // For this question one-qubit runs that end the stage on their qubit
    // (no later pair op takes them) count for nothing: when such a CNOT
    // moves to the exchange, they wait for the next stage, where they open
    // the next layer's blocks instead.
    std::vector<int> bare_perms()
    {
        choose_centers(true);
        std::vector<int> out;
        for (int id : ids_) {
            const GOp &op = g_[id];
            if ((op.kind != G_CX && op.kind != G_SWAP) || center_.count(id)) continue;
            Mat4 m4;
            int a, b;
            if (!as_pair(op, &a, &b, &m4)) continue;
            auto it = same_pair_.find(id);
            if (it != same_pair_.end() && center_.count(it->second)) continue;  // it multiplies into that block
            out.push_back(id);
        }
        return out;
    }

    // The fused stage as planner ops; new dense 4x4 and 2x2 coefficients go to
    // pool.  A pair block becomes one dense 4x4 when its member ops cost at
    // least as much; otherwise its members run as they are.
    std::vector<GOp> finish(std::vector<c128> &pool)
    {
        choose_centers();
        for (int id : ids_) add_now(id);
        std::vector<GOp> out;
        for (const Block &bl : kron_singles()) {
            switch (bl.kind) {
            case Block::Dead: break;
            case Block::Opaque: out.push_back(g_[bl.members[0]]); break;
            case Block::Single: {
                if (bl.members.size() == 1) {  // untouched: keep the op as lowered
                    out.push_back(g_[bl.members[0]]);
                    break;
                }
                if (is_identity2(bl.m2)) break;
                GOp o{};
                o.q0 = bl.a;
                o.coef = uint32_t(pool.size());
                if (classify1(bl.m2.data()) != K1::Dense) {
                    o.kind = G_DIAG;
                    pool.push_back(bl.m2[0]);
                    pool.push_back(bl.m2[3]);
                } else {
                    o.kind = G_U1;
                    pool.insert(pool.end(), bl.m2.begin(), bl.m2.end());
                }
                out.push_back(o);
                break;
            }
            case Block::Pair: {
                float parts = 0;
                bool dense = false;
                for (int id : bl.members) {
                    parts += fused_cost(g_[id].kind);
                    dense = dense || g_[id].kind == G_U2;
                }
                if (!dense && parts < fused_cost(G_U2)) {
                    for (int id : bl.members) out.push_back(g_[id]);
                    break;
                }
                GOp o{};
                o.kind = G_U2;
                o.q0 = bl.a;
                o.q1 = bl.b;
                o.coef = uint32_t(pool.size());
                pool.insert(pool.end(), bl.m4.begin(), bl.m4.end());
                out.push_back(o);
                break;
            }
            }
        }
        return out;
    }

  private:
    void add_now(int id)
    {
        const GOp &op = g_[id];
        Mat2 m2;
        Mat4 m4;
        int a, b;
        if (as_single(op, &a, &m2)) return add_single(id, a, m2);
        if (as_pair(op, &a, &b, &m4)) {
            const int oa = open_of(a);
            const bool same_pair = oa >= 0 && oa == open_of(b) && blocks_[oa].kind == Block::Pair;
            if (same_pair || center_.count(id)) return add_pair(id, a, b, m4);
        }
        add_opaque(id);
    }

    // Which two-qubit ops open a dense 4x4 block.  A block costs a 4x4 (16
    // FMA per amplitude) whatever it absorbs, so it pays off by taking the
    // one-qubit gates next to it: on each register qubit a run of one-qubit
    // gates between two pair ops can join either neighbour.  In a brick-wall
    // every run sits between two CNOTs of alternating partners; making every
    // CNOT a block leaves each block two runs, while making every other one a
    // block (the rest run as register swaps) gives each block four.  Greedy
    // weighted cover: a candidate is chosen while the runs it would newly
    // absorb cost more than the block adds over running the op bare (dense
    // 4x4s are always blocks: they cost the same either way).
    void choose_centers(bool free_tails = false)
    {
        struct Edge {
            int prev, next;
            float w;
            bool covered;
        };
        std::vector<Edge> edges;
        std::map<int, std::vector<int>> inc;  // candidate -> incident edges
        std::map<int, int> last;              // qubit -> last candidate on it (absent or -1: none)
        std::map<int, int> run;               // qubit -> open run edge (absent: none)
        auto close_run = [&](int q, int next) {
            auto it = run.find(q);
            if (it == run.end()) return;
            edges[size_t(it->second)].next = next;
            if (next >= 0) inc[next].push_back(it->second);
            run.erase(it);
        };
        std::vector<int> cands;
        for (int id : ids_) {
            const GOp &op = g_[id];
            Mat2 m2;
            Mat4 m4;
            int a, b;
            if (as_single(op, &a, &m2)) {
                const float w = classify1(m2.data()) == K1::Dense ? fused_cost(G_U1) : fused_cost(G_DIAG);
                auto it = run.find(a);
                if (it != run.end()) {
                    edges[size_t(it->second)].w = std::max(edges[size_t(it->second)].w, w);
                    continue;
                }
                auto lt = last.find(a);
                const int prev = lt == last.end() ? -1 : lt->second;
                run[a] = int(edges.size());
                if (prev >= 0) inc[prev].push_back(int(edges.size()));
                edges.push_back(Edge{prev, -1, w, false});
                continue;
            }
            if (as_pair(op, &a, &b, &m4)) {
                cands.push_back(id);
                inc[id];
                auto la = last.find(a), lb = last.find(b);
                if (la != last.end() && lb != last.end() && la->second >= 0 && la->second == lb->second)
                    same_pair_[id] = la->second;  // nothing but one-qubit runs since the previous op on this pair
                for (int q : {a, b}) {
                    close_run(q, id);
                    last[q] = id;
                }
                if (op.kind == G_U2) center_.insert(id);
                continue;
            }
            for (uint64_t t = op.touched() & regq_; t; t &= t - 1) {  // a boundary: nothing moves across it
                const int q = __builtin_ctzll(t);
                close_run(q, -1);
                last[q] = -1;
            }
        }
        const float block = fused_cost(G_U2);
        auto cover = [&](int c) {
            for (int e : inc[c]) edges[size_t(e)].covered = true;
        };
        for (int c : center_) cover(c);
        for (;;) {
            int best = -1;
            float best_gain = 0;
            for (int c : cands) {
                if (center_.count(c)) continue;
                float w = 0;
                for (int e : inc[c])
                    if (!edges[size_t(e)].covered && !(free_tails && edges[size_t(e)].next < 0)) w += edges[size_t(e)].w;
                const float gain = w - (block - fused_cost(g_[c].kind));
                if (gain > best_gain) {
                    best_gain = gain;
                    best = c;
                }
            }
            if (best < 0) break;
            center_.insert(best);
            cover(best);
        }
    }

    bool dense_single(const Block &bl) const
    {
        return bl.kind == Block::Single && !is_identity2(bl.m2) && classify1(bl.m2.data()) == K1::Dense;
    }

    uint64_t block_touches(const Block &bl) const
    {
        switch (bl.kind) {
        case Block::Single: return uint64_t(1) << bl.a;
        case Block::Pair: return (uint64_t(1) << bl.a) | (uint64_t(1) << bl.b);
        case Block::Opaque: return g_[bl.members[0]].touched();
        default: return 0;
        }
    }

    // Two dense one-qubit blocks on different qubits cost what one dense 4x4
    // costs (their Kronecker product), so pairing them saves a whole op.  The
    // later block moves back to the earlier one when no block in between
    // touches its qubit.
    std::vector<Block> kron_singles() const
    {
        std::vector<Block> bl = blocks_;
        for (size_t i = 0; i < bl.size(); ++i) {
            if (!dense_single(bl[i])) continue;
            uint64_t between = 0;
            for (size_t k = i + 1; k < bl.size(); ++k) {
                if (dense_single(bl[k]) && bl[k].a != bl[i].a && !((between >> bl[k].a) & 1)) {
                    Block &p = bl[i];
                    p.kind = Block::Pair;
                    p.b = bl[k].a;  // p.a stays the basis MSB
                    p.m4 = mul4(lift(p.m2, true), lift(bl[k].m2, false));
                    p.members.insert(p.members.end(), bl[k].members.begin(), bl[k].members.end());
                    std::sort(p.members.begin(), p.members.end());
                    bl[k].kind = Block::Dead;
                    break;
                }
                between |= block_touches(bl[k]);
            }
        }
        return bl;
    }

    static bool is_identity2(const Mat2 &m)
    {
        return ceq(m[0], c128{1, 0}) && ceq(m[1], c128{0, 0}) && ceq(m[2], c128{0, 0}) && ceq(m[3], c128{1, 0});
    }
    bool reg(int q) const { return q >= 0 && ((regq_ >> q) & 1); }

    // One-qubit ops on a register qubit: dense or diagonal 2x2.
    bool as_single(const GOp &op, int *q, Mat2 *m) const
    {
        if (op.cmask || !reg(op.q0)) return false;
        const c128 *c = pool_.data() + op.coef;
        if (op.kind == G_U1) *m = Mat2{c[0], c[1], c[2], c[3]};
        else if (op.kind == G_DIAG) *m = Mat2{c[0], c128{0, 0}, c128{0, 0}, c[1]};
        else return false;
        *q = op.q0;
        return true;
    }

    // Two-qubit ops confined to register qubits, as a 4x4 in the basis |a b>.
    bool as_pair(const GOp &op, int *a, int *b, Mat4 *m) const
    {
        const c128 *c = pool_.data() + op.coef;
        const c128 one{1, 0}, zero{0, 0};
        *m = Mat4{};
        if (op.kind == G_U2 || op.kind == G_SWAP) {
            if (!reg(op.q0) || !reg(op.q1) || op.cmask) return false;
            *a = op.q0;
            *b = op.q1;
            if (op.kind == G_U2) std::copy(c, c + 16, m->begin());
            else *m = gate::swap();
            return true;
        }
        if (__builtin_popcountll(op.cmask) != 1 || !reg(op.q0)) return false;
        const int ctl = __builtin_ctzll(op.cmask);
        if (!reg(ctl)) return false;
        *a = ctl;  // control = MSB of the basis
        *b = op.q0;
        Mat2 t;
        switch (op.kind) {
        case G_CX: t = Mat2{zero, one, one, zero}; break;
        case G_CU1: t = Mat2{c[0], c[1], c[2], c[3]}; break;
        case G_DIAG: t = Mat2{c[0], zero, zero, c[1]}; break;
        default: return false;
        }
        for (int i = 0; i < 2; ++i) (*m)[i * 5] = one;  // control 0: identity on the target
        for (int r = 0; r < 2; ++r)
            for (int cc = 0; cc < 2; ++cc) (*m)[(2 + r) * 4 + 2 + cc] = t[r * 2 + cc];
        return true;
    }

    int open_of(int q) const
    {
        auto it = open_.find(q);
        return it == open_.end() ? -1 : it->second;
    }

    void add_single(int id, int q, const Mat2 &m)
    {
        const int o = open_of(q);
        if (o >= 0 && blocks_[o].kind == Block::Single) {
            blocks_[o].m2 = gate::compose(m, blocks_[o].m2);
            blocks_[o].members.push_back(id);
            return;
        }
        if (o >= 0 && blocks_[o].kind == Block::Pair) {
            Block &bl = blocks_[o];
            bl.m4 = mul4(lift(m, q == bl.a), bl.m4);
            bl.members.push_back(id);
            return;
        }
        Block bl{Block::Single};
        bl.a = q;
        bl.m2 = m;
        bl.members.push_back(id);
        open_[q] = int(blocks_.size());
        blocks_.push_back(bl);
    }

    void add_pair(int id, int a, int b, const Mat4 &m)
    {
        const int oa = open_of(a), ob = open_of(b);
        if (oa >= 0 && oa == ob && blocks_[oa].kind == Block::Pair) {  // same pair: multiply in
            Block &bl = blocks_[oa];
            bl.m4 = mul4(bl.a != a ? reorder(m) : m, bl.m4);
            bl.members.push_back(id);
            return;
        }
        // A new pair block takes over the open one-qubit blocks of a and b
        // (nothing touched those qubits since, so they move forward freely).
        Block bl{Block::Pair};
        bl.a = a;
        bl.b = b;
        Mat4 pre = lift(gate::id(), true);
        for (int o : {oa, ob}) {
            if (o < 0 || blocks_[o].kind != Block::Single) continue;
            pre = mul4(lift(blocks_[o].m2, blocks_[o].a == a), pre);
            bl.members.insert(bl.members.end(), blocks_[o].members.begin(), blocks_[o].members.end());
            blocks_[o].kind = Block::Dead;
        }
        std::sort(bl.members.begin(), bl.members.end());  // ids are in program order
        bl.m4 = mul4(m, pre);
        bl.members.push_back(id);
        open_[a] = open_[b] = int(blocks_.size());
        blocks_.push_back(bl);
    }

    void add_opaque(int id)
    {
        Block bl{Block::Opaque};
        bl.members.push_back(id);
        const int idx = int(blocks_.size());
        for (uint64_t t = g_[id].touched(); t; t &= t - 1) open_[__builtin_ctzll(t)] = idx;
        blocks_.push_back(bl);
    }

    // 4x4 in the basis |b a> -> |a b>.
    static Mat4 reorder(const Mat4 &m)
    {
        auto sw = [](int x) { return ((x & 1) << 1) | (x >> 1); };
        Mat4 r{};
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) r[sw(i) * 4 + sw(j)] = m[i * 4 + j];
        return r;
    }

    const std::vector<GOp> &g_;
    const std::vector<c128> &pool_;
    uint64_t regq_;
    std::vector<int> ids_;
    std::set<int> center_;          // pair ops that open a block
    std::map<int, int> same_pair_;  // pair op -> the previous pair op on the same two qubits
    std::vector<Block> blocks_;
    std::map<int, int> open_;
};

// A 4x4 (basis |b_hi b_lo>, row-major) whose entries other than m00 m03 m11
// m22 m30 m33 are exactly zero: it mixes |00> with |11> and scales |01>
// and |10> (fused::kU2X).
bool x_form(const c128 *m)
{
    static constexpr int kZero[] = {1, 2, 4, 6, 7, 8, 9, 11, 13, 14};
    for (int k : kZero)
        if (m[k].re != 0 || m[k].im != 0) return false;
    return true;
}

// Lowers one planner op into a device op for register set S (local
// positions); regbit maps a global qubit to its register bit or -1.  With
// sparse_u2, a 4x4 of the x_form shape becomes a kU2X.
template <class RegBit>
DevOp lower_op(const GOp &op, const std::vector<c128> &pool, uint32_t S, const RegBit &regbit, std::vector<c128> &coef,
               bool sparse_u2)
{
    DevOp d{};
    uint32_t kind = 0, b0 = 0, b1 = 0, flags = 0;
    for (uint64_t m = op.cmask; m; m &= m - 1) {
        const int q = __builtin_ctzll(m);
        const int rb = regbit(S, q);
        if (rb >= 0) d.creg |= 1u << rb;
        else d.cglob |= uint64_t(1) << q;
    }
    d.coef = uint32_t(coef.size());
    switch (op.kind) {
    case G_U1:
    case G_CU1:
        kind = (op.kind == G_U1 && op.cmask == 0) ? fused::kU1 : fused::kCU1;
        b0 = uint32_t(regbit(S, op.q0));
        for (int e = 0; e < 4; ++e) {
            d.m[2 * e] = pool[op.coef + e].re;
            d.m[2 * e + 1] = pool[op.coef + e].im;
        }
        break;
    case G_CX:
        kind = fused::kCX;
        b0 = uint32_t(regbit(S, op.q0));
        break;
    case G_DIAG: {
        const c128 d0 = pool[op.coef], d1 = pool[op.coef + 1];
        const int rb = regbit(S, op.q0);
        if (rb >= 0) {
            kind = fused::kDiagR;
            b0 = uint32_t(rb);
        } else {
            kind = fused::kDiagT;
            d.tq = uint32_t(op.q0);
        }
        if (d0.re == 1 && d0.im == 0) flags |= fused::kD0IsOne;
        d.m[0] = d0.re;
        d.m[1] = d0.im;
        d.m[2] = d1.re;
        d.m[3] = d1.im;
        break;
    }
    case G_U2: {
        const int ra = regbit(S, op.q0), rb = regbit(S, op.q1);
        const c128 *m = pool.data() + op.coef;
        b0 = uint32_t(std::max(ra, rb));
        b1 = uint32_t(std::min(ra, rb));
        // the basis |b0 b1>: as stored when q_a has the higher register bit,
        // else reordered from |q_a q_b> to |q_b q_a>
        auto sw = [&](int x) { return ra > rb ? x : ((x & 1) << 1) | (x >> 1); };
        c128 mm[16];
        for (int r = 0; r < 4; ++r)
            for (int cc = 0; cc < 4; ++cc) mm[sw(r) * 4 + sw(cc)] = m[r * 4 + cc];
        if (sparse_u2 && x_form(mm)) {
            kind = fused::kU2X;
            for (int k : {0, 3, 5, 10, 12, 15}) coef.push_back(mm[k]);
        } else {
            kind = fused::kU2;
            coef.insert(coef.end(), mm, mm + 16);
        }
        break;
    }
    case G_SWAP: {
        kind = fused::kSwap;
        const int ra = regbit(S, op.q0), rb = regbit(S, op.q1);
        b0 = uint32_t(std::max(ra, rb));
        b1 = uint32_t(std::min(ra, rb));
        break;
    }
    }
    d.code = fused::op_code(kind, b0, b1, flags);
    return d;
}

// Merges a stage's conditioned phase ops.  A diagonal op whose factor on
// the 0 side is 1 multiplies, on the lanes whose global bits meet its
// condition, either the registers with one register bit set (kDiagR) or
// every register (kDiagT, whose target bit joins the condition).  Diagonal
// ops commute with each other, and ops inside a stage never change a lane's
// global bits, so such an op moves back to an earlier phase op of the same
// register bit as long as no op in between targets that bit (kDiagT terms
// commute with every op of the stage).  Each merged op computes the lane's
// product of factors once and scales its registers once.
std::vector<DevOp> merge_phases(const std::vector<DevOp> &in, std::vector<c128> &coef)
{
    // A group is keyed by its register bit b0 (kPhaseAll: every register)
    // and a register condition creg: it scales the registers j with bit b0
    // set and (j & creg) == creg.  A controlled phase with its control in a
    // register and its target outside is such a term on the control's bit.
    struct Group {
        size_t at;  // position in out
        uint32_t b0, creg;
        std::vector<std::pair<uint64_t, c128>> terms;
    };
    std::vector<DevOp> out;
    std::vector<Group> groups;
    std::map<uint32_t, int> open;  // (b0 << 8 | creg) -> group
    for (const DevOp &d : in) {
        const uint32_t kind = d.code & 0xff, b0 = (d.code >> 8) & 0xff, b1 = (d.code >> 16) & 0xff;
        const bool one = ((d.code >> 24) & fused::kD0IsOne) != 0;
        bool term = false;
        uint32_t gb = 0, gc = 0;
        uint64_t mask = 0;
        if (kind == fused::kDiagR && one) {
            term = true;
            gb = b0;
            gc = d.creg;
            mask = d.cglob;
        } else if (kind == fused::kDiagT && one) {
            term = true;
            mask = d.cglob | (uint64_t(1) << d.tq);
            if (d.creg == 0) {
                gb = fused::kPhaseAll;
            } else {
                gb = uint32_t(__builtin_ctz(d.creg));
                gc = d.creg & ~(1u << gb);
            }
        }
        if (term) {
            const c128 f{d.m[2], d.m[3]};
            const uint32_t key = (gb << 8) | gc;
            auto it = open.find(key);
            if (it == open.end() || groups[it->second].terms.size() == fused::kMaxPhaseTerms) {
                open[key] = int(groups.size());
                groups.push_back({out.size(), gb, gc, {}});
                out.push_back(d);  // placeholder: rewritten below, or kept when the group stays one term
            }
            groups[open[key]].terms.push_back({mask, f});
            continue;
        }
        uint32_t targets = 0;  // an op that targets a register bit closes the groups that read it
        switch (kind) {
        case fused::kU1:
        case fused::kCU1:
        case fused::kCX: targets = 1u << b0; break;
        case fused::kU2:
        case fused::kU2X:
        case fused::kSwap: targets = (1u << b0) | (1u << b1); break;
        default: break;
        }
        for (auto it = open.begin(); it != open.end();) {
            const Group &gr = groups[it->second];
            const uint32_t reads = gr.creg | (gr.b0 == fused::kPhaseAll ? 0u : 1u << gr.b0);
            if (reads & targets) it = open.erase(it);
            else ++it;
        }
        out.push_back(d);
    }
    for (const Group &gr : groups) {
        if (gr.terms.size() == 1) continue;  // the original op stays
        DevOp m{};
        m.code = fused::op_code(fused::kPhase, gr.b0, uint32_t(gr.terms.size()), 0);
        m.creg = gr.creg;
        m.coef = uint32_t(coef.size());
        for (size_t i = 0; i < gr.terms.size(); ++i) {
            std::memcpy(&m.m[i], &gr.terms[i].first, sizeof(uint64_t));
            coef.push_back(gr.terms[i].second);
        }
        out[gr.at] = m;
    }
    return out;
}

// Register positions for a stage that needs `S`: padded to 4 with free
// positions, never from `avoid` when the stage must stay coalesced, and
// otherwise scored so that the 16 lanes of an LDS quarter-wave fall on
// distinct 16-byte slots (thread-bit positions distinct mod 4), then away
// from the low positions, then as high as possible.
uint32_t pad_registers(uint32_t S, int t, uint32_t avoid, bool must_avoid)
{
    const int need = kRegBits - popc(S);
    if (need <= 0) return S;
    int freep[16], nf = 0;
    for (int l = 0; l < t; ++l)
        if (!((S >> l) & 1) && !(must_avoid && ((avoid >> l) & 1))) freep[nf++] = l;
    uint32_t best = 0;
    long best_score = -1;
    // every subset of `need` free positions (at most C(12, 4) = 495)
    int idx[4];
    for (int k = 0; k < need; ++k) idx[k] = k;
    while (nf >= need) {
        uint32_t cand = S;
        for (int k = 0; k < need; ++k) cand |= 1u << freep[idx[k]];
        int p[4], np = 0;
        for (int l = 0; l < t && np < 4; ++l)
            if (!((cand >> l) & 1)) p[np++] = l;
        int seen = 0, distinct = 0;
        for (int k = 0; k < np; ++k)
            if (!((seen >> (p[k] & 3)) & 1)) {
                seen |= 1 << (p[k] & 3);
                ++distinct;
            }
        long score = long(distinct) * 1000000 + (popc(cand & avoid) ? 0 : 100000);
        for (int k = 0; k < need; ++k) score += freep[idx[k]];
        if (score > best_score) {
            best_score = score;
            best = cand;
        }
        int k = need - 1;  // next combination
        while (k >= 0 && idx[k] == nf - need + k) --k;
        if (k < 0) break;
        ++idx[k];
        for (int r = k + 1; r < need; ++r) idx[r] = idx[r - 1] + 1;
    }
    EV_REQUIRE(best_score >= 0, EV_E_ARG, "no register positions left for a stage");
    return best;
}

// Algorithmic floating-point operations per amplitude of the state that one
// device op needs by its definition: an n-term complex dot product is 8n - 2
// flops (an FMA counts 2), a complex scale 6, register moves none; an op
// conditioned on register or global bits touches 2^-(condition bits) of the
// amplitudes.
double op_flops_per_amp(const DevOp &d)
{
    const uint32_t kind = d.code & 0xff, b0 = (d.code >> 8) & 0xff, flags = d.code >> 24;
    double f = 0, share = 1;
    switch (kind) {
    case fused::kU1:
    case fused::kCU1: f = 14; break;
    case fused::kU2: f = 30; break;
    case fused::kU2X: f = 10; break;  // two 2-term dot products (14 each) and two complex scales (6 each) per 4
    case fused::kDiagR: f = (flags & fused::kD0IsOne) ? 3 : 6; break;
    case fused::kDiagT: f = 6; break;
    case fused::kPhase: f = b0 == fused::kPhaseAll ? 6 : 3; break;
    default: return 0;
    }
    share = std::ldexp(1.0, -(__builtin_popcount(d.creg) + __builtin_popcountll(d.cglob)));
    return f * share;
}

// Lanes of one wave that meet in one LDS cycle collide when their slots
// share a bank: 2^(bits - rank) lanes per bank, where rank is the GF(2) rank
// of the low `bits` slot bits that the lanes' first `bits` thread-id bits
// produce (thread-id bit i puts v[i] into the slot).
int bank_degree(const uint32_t *v, int nv, int bits, uint32_t a, uint32_t b)
{
    const int k = std::min(nv, bits);
    uint32_t basis[6] = {};  // basis[h]: the reduced vector whose highest set bit is h
    int rank = 0;
    for (int i = 0; i < k; ++i) {
        uint32_t x = fused::swz(v[i], a, b) & ((1u << bits) - 1);
        for (int h = bits - 1; h >= 0 && x; --h) {
            if (!((x >> h) & 1)) continue;
            if (!basis[h]) {
                basis[h] = x;
                ++rank;
                x = 0;
            } else {
                x ^= basis[h];
            }
        }
    }
    return 1 << (k - rank);
}

// Chooses the pass's slot swizzle and applies it to the exchange tables of
// its stages (the planner fills them unswizzled).  Every exchange of a pass
// shares one swizzle, so a stage without relabelling writes exactly the
// slots it read and needs no barrier between its read and its write.  The
// score is the bank degree of the writes and the reads of every exchange,
// for 32 lanes per cycle (8-byte slots), then for 64 (4-byte slots).
void pass_swizzle(Pass &ps, Stage *st, bool wide64)
{
    const int ns = ps.nstages, np = ps.t - kRegBits;
    uint32_t best_a = 6, best_b = fused::kSwzNone;
    if (ns > 1) {
        std::vector<std::array<uint32_t, 6>> wr, rd;
        for (int k = 0; k + 1 < ns; ++k) {
            std::array<uint32_t, 6> w{}, r{};
            for (int i = 0; i < std::min(np, 6); ++i) {
                w[i] = 1u << st[k].p[i];
                r[i] = st[k + 1].relabel ? st[k + 1].rimg[i] : 1u << st[k + 1].p[i];
            }
            wr.push_back(w);
            rd.push_back(r);
        }
        // Candidates in order of how often they win on brick-wall and
        // random-circuit plans; on those, this list scores as well as every
        // (a, b) with 2 <= a < b <= 12.
        static constexpr uint32_t cand[][2] = {{3, 4}, {2, 4}, {3, 5}, {2, 5}, {2, 3}, {4, 7},
                                               {4, 5}, {3, 6}, {2, 6}, {4, 6}, {5, 6}, {6, fused::kSwzNone}};
        // the lanes that share an LDS cycle: 32 with 8-byte slots (complex128),
        // all 64 with 4-byte slots (complex64); their bank degree counts first
        const int major = wide64 ? 6 : 5, minor = wide64 ? 5 : 6;
        long best = -1;
        const long ideal = 2L * (ns - 1) * 1000 + 2L * (ns - 1);
        for (const auto &c : cand) {
            long sc = 0;
            for (size_t k = 0; k < wr.size(); ++k)
                sc += 1000L * (bank_degree(wr[k].data(), np, major, c[0], c[1]) +
                               bank_degree(rd[k].data(), np, major, c[0], c[1])) +
                      bank_degree(wr[k].data(), np, minor, c[0], c[1]) + bank_degree(rd[k].data(), np, minor, c[0], c[1]);
            if (best < 0 || sc < best) {
                best = sc;
                best_a = c[0];
                best_b = c[1];
            }
            if (best == ideal) break;
        }
    }
    ps.swz_a = int32_t(best_a);
    ps.swz_b = int32_t(best_b);
    for (int k = 0; k < ns; ++k) {
        Stage &s = st[k];
        for (int j = 0; j < kRegs; ++j) {
            s.roff[j] = fused::swz(s.roff[j], best_a, best_b);
            s.rroff[j] = fused::swz(s.rroff[j], best_a, best_b);
        }
        for (int i = 0; i < fused::kTileSlots; ++i) s.rimg[i] = fused::swz(s.rimg[i], best_a, best_b);
        for (int i = 0; i < s.naff; ++i) s.aoff[i] = fused::swz(s.aoff[i], best_a, best_b);
    }
}

// Builds the pass list for `g` on an n-qubit state.  `coalesce` is the number
// of low qubits whose amplitudes make one 128-byte run (3 for c128, 4 for
// c64): the first and last stage of every pass keep their register bits off
// those positions (where the tile holds the state's lowest qubits), so the
// kernel's global gather and scatter move whole runs.
//
// `on_pass`, when given, is called with the plan so far each time a pass is
// complete (later passes never change it), so a caller can launch pass p
// while pass p + 1 is planned.
using PassSink = std::function<void(const Plan &, size_t)>;
Plan plan(const std::vector<GOp> &g, const std::vector<c128> &pool_in, int n, int t, int low, float budget,
          int coalesce, bool sparse_u2, const PassSink &on_pass = {})
{
    Plan P;
    std::vector<c128> pool = pool_in;  // grows with fused coefficients
    std::vector<int> next_same(g.size(), -1);  // per pass: see the stage loop
    const uint64_t all = (n >= 64) ? ~uint64_t(0) : ((uint64_t(1) << n) - 1);
    const uint64_t lowmask = (uint64_t(1) << low) - 1;
    std::vector<int> rem(g.size());
    for (size_t i = 0; i < g.size(); ++i) rem[i] = int(i);

    // One pass's ops: in program order an op joins when the qubits it needs
    // fit the tile and it does not depend on a deferred op.  `grow` lets the
    // tile take new qubits up to t (program-order greedy); otherwise the tile
    // is the fixed set Q0.
    // defer holds the ops deferred before the scan stopped at rem[stop]; the
    // rest of rem follows them (appended for the chosen tile only: copying it
    // for every candidate cost more than planning itself).
    struct Formed {
        uint64_t Q;
        float work;
        std::vector<int> take, defer;
        size_t stop;
    };
    auto form = [&](uint64_t Q0, bool grow) {
        Formed f{Q0, 0.f, {}, {}, 0};
        uint64_t blocked = 0;
        float cost = 0;
        size_t i = 0;
        for (; i < rem.size(); ++i) {
            const GOp &op = g[rem[i]];
            const uint64_t tq = op.touched();
            if (tq & blocked) {
                blocked |= tq;
                f.defer.push_back(rem[i]);
            } else {
                uint64_t nq = f.Q | op.want();
                if (!grow || popc(nq) > t) nq = f.Q | op.need();
                const bool fits = grow ? popc(nq) <= t : nq == f.Q;
                if (fits && (f.take.empty() || cost + op.cost <= budget)) {
                    f.Q = nq;
                    cost += op.cost;
                    f.take.push_back(rem[i]);
                } else {
                    blocked |= tq;
                    f.defer.push_back(rem[i]);
                }
            }
            if ((blocked & all) == all) {
                ++i;
                break;
            }
        }
        f.stop = i;
        f.work = cost + 0.01f * float(f.take.size());
        return f;
    };

    while (!rem.empty()) {
        // Candidate tiles: the program-order greedy one, and every window of
        // t - low consecutive qubits above the low ones.  The greedy tile
        // follows the oldest deferred ops, which on nearest-neighbour
        // circuits are spread over the whole register at different depths
        // (each pass then advances scattered pairs by a layer); a window
        // advances one block by many layers.  The tile that takes the most
        // work wins.
        Formed best = form(lowmask, true);
        const int width = t - low;
        for (int a = low; a + width <= n; ++a) {
            const uint64_t win = (width >= 64 ? ~uint64_t(0) : ((uint64_t(1) << width) - 1)) << a;
            Formed f = form(lowmask | win, false);
            if (f.work > best.work) best = std::move(f);
        }
        uint64_t Q = best.Q;
        std::vector<int> take = std::move(best.take);
        best.defer.insert(best.defer.end(), rem.begin() + ptrdiff_t(best.stop), rem.end());
        rem.swap(best.defer);

        // Pad Q to exactly t qubits with the lowest unused ones.
        for (int q = 0; q < n && popc(Q) < t; ++q) Q |= uint64_t(1) << q;

        Pass ps{};
        ps.t = t;
        ps.ncomp = n - t;
        int k = 0, c = 0;
        for (int q = 0; q < n; ++q) {
            if ((Q >> q) & 1) ps.q[k++] = q;
            else ps.comp[c++] = q;
        }
        ps.stage_begin = int(P.stages.size());
        auto local = [&](int q) { return rank_in(Q, q); };

        // register bit of global qubit q under register set S (or -1)
        auto regbit = [&](uint32_t S, int q) -> int {
            if (!((Q >> q) & 1)) return -1;
            const int l = local(q);
            if (!((S >> l) & 1)) return -1;
            return rank_in(S, l);
        };
        // relabel: the composed inverse permutation A (linear part: image of
        // each local bit; affine part: (global bit, offset) terms) that this
        // stage's LDS read applies, or null
        struct Relabel {
            std::vector<uint32_t> lin;
            std::vector<std::pair<int, uint32_t>> aff;
        };
        auto regq_of = [&](uint32_t S) {  // global qubits held in registers
            uint64_t regq = 0;
            for (int l = 0; l < t; ++l)
                if ((S >> l) & 1) regq |= uint64_t(1) << ps.q[l];
            return regq;
        };
        auto emit_stage = [&](uint32_t S, const std::vector<int> &stake, const Relabel *relabel) {
            StageFuser fuser(g, pool, regq_of(S));
            for (int id : stake) fuser.add(id);
            const std::vector<GOp> fused = fuser.finish(pool);
            Stage st{};
            st.op_begin = int(P.devops.size());
            st.np = t - kRegBits;
            int sp[kRegBits], ns = 0, np = 0;
            for (int l = 0; l < t; ++l) {
                if ((S >> l) & 1) sp[ns++] = l;
                else {
                    st.p[np] = l;
                    st.gpos[np] = ps.q[l];
                    ++np;
                }
            }
            for (int j = 0; j < kRegs; ++j) {
                uint32_t off = 0;
                uint64_t goff = 0;
                for (int b = 0; b < kRegBits; ++b)
                    if ((j >> b) & 1) {
                        off |= 1u << sp[b];
                        goff |= uint64_t(1) << ps.q[sp[b]];
                    }
                st.roff[j] = off;  // swizzled with the pass (pass_swizzle)
                st.goff[j] = goff;
            }
            st.relabel = relabel ? 1 : 0;
            auto image = [&](uint32_t x) {  // linear part of A at x, a local index
                if (!relabel) return x;
                uint32_t r = 0;
                for (; x; x &= x - 1) r ^= relabel->lin[__builtin_ctz(x)];
                return r;
            };
            if (relabel) {
                st.naff = int(relabel->aff.size());
                for (size_t k2 = 0; k2 < relabel->aff.size(); ++k2) {
                    st.abit[k2] = uint32_t(relabel->aff[k2].first);
                    st.aoff[k2] = relabel->aff[k2].second;
                }
            }
            for (int i = 0; i < np; ++i) st.rimg[i] = image(1u << st.p[i]);
            for (int j = 0; j < kRegs; ++j) {
                uint32_t off = 0;
                for (int b = 0; b < kRegBits; ++b)
                    if ((j >> b) & 1) off |= 1u << sp[b];
                st.rroff[j] = image(off);
            }
            std::vector<DevOp> lowered;
            for (const GOp &op : fused) lowered.push_back(lower_op(op, pool, S, regbit, P.coef, sparse_u2));
            for (const DevOp &d : merge_phases(lowered, P.coef)) P.devops.push_back(d);
            st.op_end = int(P.devops.size());
            for (int o = st.op_begin; o < st.op_end; ++o) {
                const DevOp &d = P.devops[size_t(o)];
                const uint32_t kind = d.code & 0xff;
                if (d.cglob != 0 || kind == fused::kDiagT || kind == fused::kPhase) st.gneed = 1;
            }
            P.stages.push_back(st);
        };

        // Stages: greedy with 4 register qubits (local positions).  The
        // first stage takes only ops whose register bits avoid the coalesced
        // low positions (it may take none: then it only gathers the tile).
        int contiguous = 0;  // leading tile positions that are the state's lowest qubits
        while (contiguous < t && ps.q[contiguous] == contiguous) ++contiguous;
        const uint32_t lowc = (1u << std::min(coalesce, contiguous)) - 1;
        //
        // A CNOT or SWAP between tile qubits whose qubits do not fit the
        // stage's registers is not run as an op: it is a linear relabelling
        // of local indices, which the exchange after the stage applies by
        // reading through its inverse (relabel tables of the next stage).
        // Later ops on its qubits then wait for the next stage.
        std::vector<int> srem = take;
        // the next op of the pass on the qubit of each one-qubit op (-1: none)
        {
            int lastq[64];
            std::fill(lastq, lastq + 64, -1);
            for (size_t a = take.size(); a-- > 0;) {
                const int id = take[a];
                const GOp &op = g[size_t(id)];
                if (op.kind == G_U1 && op.cmask == 0) next_same[size_t(id)] = lastq[op.q0];
                for (uint64_t m = op.touched(); m; m &= m - 1) lastq[__builtin_ctzll(m)] = id;
            }
        }
        uint32_t lastS = 0;
        bool first = true;
        Relabel ident, a_in, a_out;
        ident.lin.resize(t);
        for (int k2 = 0; k2 < t; ++k2) ident.lin[k2] = 1u << k2;
        bool in_pending = false;
        while (!srem.empty()) {
            uint32_t S = 0;
            uint64_t sblocked = 0;
            std::vector<int> stake, sdefer;
            a_out = ident;
            bool out_pending = false;
            uint64_t out_touched = 0;  // qubits of the relabellings composed into a_out
            uint64_t stouched = 0;     // qubits of the ops this stage takes
            for (size_t si = 0; si < srem.size(); ++si) {
                const int id = srem[si];
                const GOp &op = g[id];
                const uint64_t tq = op.touched();
                if ((sblocked & Q) == Q) {  // every op touches a tile qubit: the rest all wait
                    sdefer.insert(sdefer.end(), srem.begin() + ptrdiff_t(si), srem.end());
                    break;
                }
                if (tq & sblocked) {
                    sblocked |= tq;
                    sdefer.push_back(id);
                    continue;
                }
                auto lmask = [&](uint64_t m) {
                    uint32_t r = 0;
                    for (; m; m &= m - 1)
                        if ((Q >> __builtin_ctzll(m)) & 1) r |= 1u << local(__builtin_ctzll(m));
                    return r;
                };
                const uint32_t wantm = lmask(op.want());
                const bool fits_want = popc(S | wantm) <= kRegBits && !(first && (wantm & lowc)) &&
                                       popc(wantm) == popc(op.want());
                const bool cx1 = op.kind == G_CX && __builtin_popcountll(op.cmask) == 1;
                const bool perm = (cx1 || op.kind == G_SWAP) && (tq & ~Q) == 0;
                // a CNOT controlled from outside the tile: affine, flips its target in half the tiles
                bool affine = cx1 && ((Q >> op.q0) & 1) && !(op.cmask & Q);
                if (affine) {
                    const int cb = __builtin_ctzll(op.cmask);
                    bool known = false;
                    for (const auto &term : a_out.aff) known = known || term.first == cb;
                    affine = known || int(a_out.aff.size()) < fused::kMaxAffine;
                }
                if ((perm || affine) && !fits_want) {
                    // compose A <- A o pi^-1 (pi is its own inverse)
                    if (affine) {
                        const int cb = __builtin_ctzll(op.cmask);
                        const uint32_t u = a_out.lin[local(op.q0)];
                        bool merged = false;
                        for (auto &term : a_out.aff)
                            if (term.first == cb) {
                                term.second ^= u;
                                merged = true;
                            }
                        if (!merged) a_out.aff.push_back({cb, u});
                    } else if (op.kind == G_CX) {
                        const int lc = local(__builtin_ctzll(op.cmask)), lt = local(op.q0);
                        a_out.lin[lc] ^= a_out.lin[lt];
                    } else {
                        std::swap(a_out.lin[local(op.q0)], a_out.lin[local(op.q1)]);
                    }
                    out_pending = true;
                    out_touched |= tq;
                    // later ops on its qubits wait for the next stage (no op of
                    // the pass targets an outside control)
                    sblocked |= affine ? (uint64_t(1) << op.q0) : tq;
                    continue;
                }
                if (op.kind == G_U1 && op.cmask == 0 && !(stouched & tq)) {
                    // A one-qubit op that would open its qubit in this stage, and
                    // whose next op on that qubit is a pair op of this pass that
                    // cannot join this stage, waits for the stage that takes the pair:
                    // taken now, it would run alone or as half of a Kronecker 4x4, and
                    // the pair's block would cost a second 4x4.  (Past the first stage
                    // the earliest op left is never held back: nothing is blocked and
                    // the registers are empty when it is examined.)
                    const uint64_t qb = uint64_t(1) << op.q0;
                    bool wait = false;
                    if (next_same[size_t(id)] >= 0) {  // still ahead in srem: it waits for this op
                        const GOp &nx = g[size_t(next_same[size_t(id)])];
                        const uint64_t pm = nx.want() & ~qb;
                        if (popc(nx.want()) == 2 && popc(pm) == 1 && (pm & Q) == pm) {
                            const uint32_t both = lmask(qb | pm);
                            wait = (pm & sblocked) != 0 || (first && (both & lowc)) || popc(S | both) > kRegBits;
                        }
                    }
                    if (wait) {
                        sblocked |= tq;
                        sdefer.push_back(id);
                        continue;
                    }
                }
                uint32_t need = wantm;
                if (!fits_want) need = lmask(op.need());
                if (popc(S | need) <= kRegBits && !(first && (need & lowc))) {
                    S |= need;
                    stouched |= tq;
                    stake.push_back(id);
                } else {
                    sblocked |= tq;
                    sdefer.push_back(id);
                }
            }
            // A CNOT or SWAP between register qubits that would run bare costs a
            // register shuffle; the exchange after the stage applies it for free
            // when the stage's later ops on its qubits wait for the next stage.
            // In a brick-wall window this leaves each block the rotations on both
            // sides of its own CNOT (four per 4x4) instead of sharing them with
            // the CNOTs of the next layer.  A move is taken only when no
            // relabelling already composed for this exchange (a later op) touches
            // its qubits or the qubits of the ops it sends on, so the composed
            // order stays the program order.
            for (;;) {
                StageFuser probe(g, pool, regq_of(S));
                for (int id : stake) probe.add(id);
                int chosen = -1;
                std::vector<int> keep, moved;
                for (int c : probe.bare_perms()) {
                    uint64_t blk = g[c].touched();
                    keep.clear();
                    moved.clear();
                    bool after = false;
                    for (int id : stake) {
                        if (id == c) {
                            after = true;
                        } else if (after && (g[id].touched() & blk)) {
                            blk |= g[id].touched();
                            moved.push_back(id);
                        } else {
                            keep.push_back(id);
                        }
                    }
                    if (!(blk & out_touched)) {
                        chosen = c;
                        break;
                    }
                }
                if (chosen < 0) break;
                const GOp &op = g[chosen];
                if (op.kind == G_CX) {
                    const int lc = local(__builtin_ctzll(op.cmask)), lt = local(op.q0);
                    a_out.lin[lc] ^= a_out.lin[lt];
                } else {
                    std::swap(a_out.lin[local(op.q0)], a_out.lin[local(op.q1)]);
                }
                out_pending = true;
                out_touched |= op.touched();
                stake.swap(keep);
                std::vector<int> merged(sdefer.size() + moved.size());
                std::merge(sdefer.begin(), sdefer.end(), moved.begin(), moved.end(), merged.begin());
                sdefer.swap(merged);  // ids are in program order
            }
            srem.swap(sdefer);
            S = pad_registers(S, t, lowc, first);
            emit_stage(S, stake, in_pending ? &a_in : nullptr);
            a_in = a_out;
            in_pending = out_pending;
            lastS = S;
            first = false;
        }
        // The last stage scatters the tile: if its register bits sit on the
        // coalesced positions, or relabellings follow it, one more (op-free)
        // stage does the scatter.
        if ((lastS & lowc) || in_pending) emit_stage(pad_registers(0, t, lowc, true), {}, in_pending ? &a_in : nullptr);
        ps.nstages = int(P.stages.size()) - ps.stage_begin;
        P.stages.back().gneed = 1;  // the last stage scatters to global addresses
        pass_swizzle(ps, P.stages.data() + ps.stage_begin, coalesce >= 4);
        P.passes.push_back(ps);
        if (on_pass) on_pass(P, P.passes.size() - 1);
    }
    return P;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

// Tile width for sizes the tuning table does not cover: 256 threads x 16
// amplitudes, the widest tile that keeps 12 waves per CU at complex128
// (13-qubit tiles take 8-wave workgroups; the tuner picks them per size).
constexpr int kDefaultTile = 12;

FusedConfig fused_config(int n, bool c64, const ev_run_options &opt)
{
    FusedConfig f{};
    f.t = kDefaultTile;
    f.low = c64 ? 5 : 4;  // 256-byte contiguous rows
    f.budget = 20.0f;
    ev_tune_entry te;
    if (!opt.tile_qubits && !opt.low_qubits && !opt.max_gates_per_pass && tune_lookup(n, c64, &te)) {
        f.t = int(te.tile_qubits);
        f.low = int(te.low_qubits);
        f.budget = float(te.max_gates_per_pass);
    }
    if (opt.tile_qubits) f.t = int(opt.tile_qubits);
    if (opt.low_qubits) f.low = int(opt.low_qubits);
    if (opt.max_gates_per_pass) f.budget = float(opt.max_gates_per_pass);
    EV_REQUIRE(f.t >= fused::kMinTile && f.t <= fused::kMaxTile, EV_E_ARG,
               "tile_qubits must be in [8, " + std::to_string(fused::kMaxTile) + "]");
    EV_REQUIRE(f.low >= 1 && f.low <= 6, EV_E_ARG, "low_qubits must be in [1, 6]");
    f.t = std::min(f.t, n);
    f.low = std::min(f.low, f.t - 1);
    f.coalesce = c64 ? 4 : 3;  // 128-byte runs
    EV_REQUIRE(c64 || !opt.mfma_u2, EV_E_ARG, "mfma_u2 applies to complex64 runs");
    f.mfma_u2 = opt.mfma_u2 != 0;
    return f;
}

namespace {

// Streams a plan's passes to the device as the planner completes them: each
// pass goes up as a self-contained block (its Pass, stages, ops and
// coefficients, indices rebased to the block) through the handle's pinned
// staging ring, on the handle's stream, and is launched at once, so the
// device runs pass p while the host plans pass p + 1.  Blocks take
// consecutive ring offsets; when the next one does not fit, the stream is
// drained (every earlier copy has landed) and the ring starts over.
class PassStreamer {
  public:
    PassStreamer(ev_sv *s, bool mfma_u2, RunStats &st, Profiler &prof) : s_(s), mfma_(mfma_u2), st_(st), prof_(prof) {}

    void operator()(const Plan &P, size_t p)
    {
        const bool c64 = s_->prec == kern::C64;
        const size_t ab = kern::amp_bytes(s_->prec);
        const Pass &ps = P.passes[p];
        EV_REQUIRE(ps.nstages > 0, EV_E_STATE, "fused pass without stages");
        const Stage *stg = P.stages.data() + ps.stage_begin;
        const int op0 = stg[0].op_begin, op1 = stg[ps.nstages - 1].op_end;
        const size_t coef0 = coef_done_, coef1 = P.coef.size();
        coef_done_ = coef1;
        const size_t nops = size_t(op1 - op0), ncoef = coef1 - coef0;
        const size_t o_stage = align256(sizeof(Pass));
        const size_t o_ops = o_stage + align256(size_t(ps.nstages) * sizeof(Stage));
        const size_t o_coef = o_ops + align256(std::max<size_t>(1, nops) * sizeof(DevOp));
        const size_t total = o_coef + align256(std::max<size_t>(1, ncoef) * ab);
        unsigned char *host = nullptr, *dev = nullptr;
        reserve(total, &host, &dev);

        Pass *hp = reinterpret_cast<Pass *>(host);
        *hp = ps;
        hp->stage_begin = 0;
        Stage *hs = reinterpret_cast<Stage *>(host + o_stage);
        for (int k = 0; k < ps.nstages; ++k) {
            hs[k] = stg[k];
            hs[k].op_begin -= op0;
            hs[k].op_end -= op0;
        }
        DevOp *ho = reinterpret_cast<DevOp *>(host + o_ops);
        double fpa = 0;
        bool sparse = false;  // the pass needs the kernel with the kU2X arm
        for (size_t i = 0; i < nops; ++i) {
            const DevOp &d = P.devops[size_t(op0) + i];
            fpa += op_flops_per_amp(d);
            sparse = sparse || (d.code & 0xff) == fused::kU2X;
            ho[i] = d;
            ho[i].coef = d.coef - uint32_t(coef0);  // every op of the pass was lowered after coef0
        }
        if (c64) fused::inline_coef_to_c64(ho, nops);
        if (!c64) {
            if (ncoef) std::memcpy(host + o_coef, P.coef.data() + coef0, ncoef * sizeof(c128));
        } else {
            float *f = reinterpret_cast<float *>(host + o_coef);
            for (size_t i = 0; i < ncoef; ++i) {
                f[2 * i] = float(P.coef[coef0 + i].re);
                f[2 * i + 1] = float(P.coef[coef0 + i].im);
            }
        }
        EV_HIP(hipMemcpyAsync(dev, host, total, hipMemcpyHostToDevice, s_->stream));
        prof_.begin();
        fused::launch_pass(s_->psi, c64, s_->n, ps.t, reinterpret_cast<const Pass *>(dev),
                           reinterpret_cast<const Stage *>(dev + o_stage), reinterpret_cast<const DevOp *>(dev + o_ops),
                           dev + o_coef, mfma_, sparse, s_->stream);
        prof_.end("fused_pass", full_pass_bytes(s_->n, ab));
        st_.pass_flops += fpa * std::ldexp(1.0, s_->n);
        ++st_.passes;
        ++st_.pass_launches;
        ++st_.launches;
    }

  private:
    // Ring space for a block of `bytes`: grows the ring (after draining the
    // stream) for a block larger than it, starts over when it is full.
    void reserve(size_t bytes, unsigned char **host, unsigned char **dev)
    {
        ev_sv *s = s_;
        if (bytes > s->plan_cap) {
            sv_sync(s);
            if (s->plan_buf) EV_HIP(hipFree(s->plan_buf));
            if (s->plan_host) EV_HIP(hipHostFree(s->plan_host));
            s->plan_buf = s->plan_host = nullptr;
            s->plan_cap = s->plan_off = 0;
            const size_t cap = std::max(bytes * 2, size_t(s->n >= 20 ? 4 : 1) << 20);
            EV_HIP(hipMalloc(&s->plan_buf, cap));
            EV_HIP(hipHostMalloc(&s->plan_host, cap, hipHostMallocDefault));
            s->plan_cap = cap;
        } else if (s->plan_off + bytes > s->plan_cap) {
            sv_sync(s);  // the ring's earlier copies have all landed
            s->plan_off = 0;
        }
        *host = static_cast<unsigned char *>(s->plan_host) + s->plan_off;
        *dev = static_cast<unsigned char *>(s->plan_buf) + s->plan_off;
        s->plan_off += bytes;
    }

    ev_sv *s_;
    bool mfma_;
    RunStats &st_;
    Profiler &prof_;
    size_t coef_done_ = 0;
};

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// A segment that ends in pairwise-disjoint SWAPs (a QFT's bit reversal) on
// qubits spread wider than one tile would spend whole passes moving
// amplitudes.  Those SWAPs whose qubits are all above the coalesced low ones
// run instead as one in-place bit-permutation pass after the plan (they
// commute with each other, and nothing follows them in the segment); the
// others stay in the plan.  Returns the peeled pairs (empty if not worth it).
std::vector<std::pair<int, int>> peel_swap_tail(Lowering &low, const FusedConfig &cfg, int n)
{
    std::vector<std::pair<int, int>> tail;
    std::vector<GOp> kept;
    uint64_t used = 0;
    size_t keep = low.ops.size();
    while (keep > 0) {
        const GOp &g = low.ops[keep - 1];
        if (g.kind != G_SWAP || g.cmask) break;
        const uint64_t m = (uint64_t(1) << g.q0) | (uint64_t(1) << g.q1);
        if (m & used) break;
        used |= m;
        if (std::min(g.q0, g.q1) >= cfg.coalesce) tail.emplace_back(g.q0, g.q1);
        else kept.push_back(g);
        --keep;
    }
    const bool worth = n > cfg.t && int(2 * tail.size()) > cfg.t - cfg.low && int(tail.size()) <= kern::kMaxSwapPairs;
    if (!worth) return {};
    low.ops.resize(keep);
    for (auto it = kept.rbegin(); it != kept.rend(); ++it) low.ops.push_back(*it);
    return tail;
}

void flush_segment(ev_sv *s, Lowering &low, const FusedConfig &cfg, RunStats &st, Profiler &prof)
{
    low.flush_all();
    if (low.ops.empty()) return;
    const double t0 = now_ms();
    const std::vector<std::pair<int, int>> tail = peel_swap_tail(low, cfg, s->n);
    const float budget = (cfg.t >= s->n) ? 1e30f : cfg.budget;  // one tile holds the state: one pass
    if (!low.ops.empty()) {
        PassStreamer launch(s, cfg.mfma_u2, st, prof);
        double in_sink = 0;  // host time spent uploading and launching, not planning
        plan(low.ops, low.pool, s->n, cfg.t, cfg.low, budget, cfg.coalesce, cfg.sparse_u2, [&](const Plan &P, size_t p) {
            const double a = now_ms();
            launch(P, p);
            in_sink += now_ms() - a;
        });
        st.plan_ms += now_ms() - t0 - in_sink;
    }
    if (!tail.empty()) {
        std::vector<int> a, b;
        for (const auto &pr : tail) {
            a.push_back(pr.first);
            b.push_back(pr.second);
        }
        prof.begin();
        kern::swap_bits(s->psi, s->prec, s->n, a.data(), b.data(), int(tail.size()), s->stream);
        prof.end("swap_bits", full_pass_bytes(s->n, kern::amp_bytes(s->prec)));
        ++st.passes;
        ++st.launches;
    }
    low.ops.clear();
    low.pool.clear();
}

}  // namespace

void execute_planned(ev_sv *s, const Circuit &c, Rng &rng, uint8_t *classical, const ev_run_options &opt,
                     RunStats &st, Profiler &prof)
{
    if (s->n < fused::kMinTile) {  // below one tile: per-gate kernels
        size_t i = 0;
        while (i < c.size()) i = sv_plain_op(s, c, i, rng, classical, true, st, prof);
        return;
    }
    const FusedConfig cfg = fused_config(s->n, s->prec == kern::C64, opt);
    Lowering low(s->n);
    size_t i = 0;
    while (i < c.size()) {
        const Op &op = c.ops()[i];
        switch (op.kind) {
        case OpKind::Apply1:
            low.gate1(int(op.q0), c.mat(op));
            ++st.logical_gates;
            ++i;
            break;
        case OpKind::Apply2:
            low.gate2(int(op.q0), int(op.q1), c.mat(op));
            ++st.logical_gates;
            ++i;
            break;
        case OpKind::Controlled:
            low.controlled(c.ctrls(op), int(op.nctrl), int(op.q0), c.mat(op));
            ++st.logical_gates;
            ++i;
            break;
        case OpKind::Measure:
        case OpKind::If:
            flush_segment(s, low, cfg, st, prof);
            i = sv_plain_op(s, c, i, rng, classical, true, st, prof);
            break;
        }
    }
    flush_segment(s, low, cfg, st, prof);
}

double plan_flops_per_amp(const Plan &P)
{
    double f = 0;
    for (const DevOp &d : P.devops) f += op_flops_per_amp(d);
    return f;
}

Plan plan_circuit(const Circuit &c, int n, const FusedConfig &cfg, bool peel_swaps)
{
    EV_REQUIRE(n >= fused::kMinTile, EV_E_ARG, "fused plans need at least 8 qubits");
    Lowering low(n);
    for (size_t i = 0; i < c.size(); ++i) {
        const Op &op = c.ops()[i];
        switch (op.kind) {
        case OpKind::Apply1: low.gate1(int(op.q0), c.mat(op)); break;
        case OpKind::Apply2: low.gate2(int(op.q0), int(op.q1), c.mat(op)); break;
        case OpKind::Controlled: low.controlled(c.ctrls(op), int(op.nctrl), int(op.q0), c.mat(op)); break;
        default: throw Error(EV_E_ARG, "fused plans take measurement-free circuits");
        }
    }
    low.flush_all();
    std::vector<std::pair<int, int>> tail;
    if (peel_swaps) tail = peel_swap_tail(low, cfg, n);
    const float budget = (cfg.t >= n) ? 1e30f : cfg.budget;
    Plan P = low.ops.empty() ? Plan{} : plan(low.ops, low.pool, n, cfg.t, cfg.low, budget, cfg.coalesce, cfg.sparse_u2);
    P.swap_tail = std::move(tail);
    return P;
}

PlanBlob debug_plan(const Circuit &c, const ev_run_options &opt)
{
    const int n = int(c.num_qubits());
    const FusedConfig cfg = fused_config(n, opt.precision == EV_C64, opt);
    const Plan P = plan_circuit(c, n, cfg, true);  // the plan a single run executes
    PlanBlob b;
    b.swap_tail = P.swap_tail;
    auto bytes = [](const void *p, size_t len) {
        const unsigned char *c = static_cast<const unsigned char *>(p);
        return std::vector<unsigned char>(c, c + len);
    };
    b.passes = bytes(P.passes.data(), P.passes.size() * sizeof(Pass));
    b.stages = bytes(P.stages.data(), P.stages.size() * sizeof(Stage));
    b.ops = bytes(P.devops.data(), P.devops.size() * sizeof(DevOp));
    b.coef = P.coef;
    b.np = uint32_t(P.passes.size());
    b.ns = uint32_t(P.stages.size());
    b.no = uint32_t(P.devops.size());
    b.t = uint32_t(cfg.t);
    return b;
}

// Density matrix as a 2n-qubit vector (element (i, j) at (i << n) | j):
// U rho U^dag = U on row bit n+t and conj(U) on column bit t; a channel is
// the 4x4 superoperator on (n+t, t).
void execute_planned_dm(ev_dm *d, const Circuit &c, const DmNoise &nz, Rng &rng, uint8_t *classical,
                        const ev_run_options &opt, RunStats &st)
{
    ev_sv *v = d->vec;
    const int n = d->n;
    Profiler prof(v->stream, opt.profile != 0);
    if (2 * n < fused::kMinTile) {
        size_t i = 0;
        while (i < c.size()) i = dm_plain_op(d, c, i, nz, rng, classical, true, st);
        return;
    }
    // Superoperator passes pair bit t with bit n+t, so a tile holds whole
    // qubit pairs and a pass is bounded by the tile, not by a gate budget:
    // the statevector tuning table does not apply.  Measured on MI355X at
    // n = 13: no budget, 12-bit tiles, 2 contiguous low bits (t and low
    // change the run time by under 8%; a 24-gate budget costs 1.7x).
    ev_run_options dopt = opt;
    if (!dopt.tile_qubits && !dopt.low_qubits && !dopt.max_gates_per_pass) {
        dopt.tile_qubits = uint32_t(kDefaultTile);
        dopt.low_qubits = 2;
        dopt.max_gates_per_pass = 1u << 30;
    }
    FusedConfig cfg = fused_config(2 * n, false, dopt);
    cfg.sparse_u2 = true;  // channel superoperators on a (row, column) pair mix only |00> and |11>
    Lowering low(2 * n);
    size_t i = 0;
    while (i < c.size()) {
        const Op &op = c.ops()[i];
        switch (op.kind) {
        case OpKind::Apply1: {  // density_matrix.rs:117-125
            const c128 *m = c.mat(op);
            const int t = int(op.q0);
            if (nz.k1.empty()) {
                low.gate1(n + t, m);
                low.gate1(t, conj2(m).data());
            } else {
                low.dense2(n + t, t, gate_channel_superop(Mat2{m[0], m[1], m[2], m[3]}, nz.k1).data());
            }
            ++st.logical_gates;
            ++i;
            break;
        }
        case OpKind::Apply2: {  // density_matrix.rs:127-138
            const int a = int(op.q0), b = int(op.q1);
            low.gate2(n + a, n + b, c.mat(op));
            low.gate2(a, b, conj4(c.mat(op)).data());
            if (!nz.k2.empty()) {
                low.dense2(n + a, a, nz.s2.data());
                low.dense2(n + b, b, nz.s2.data());
            }
            ++st.logical_gates;
            ++i;
            break;
        }
        case OpKind::Controlled: {  // density_matrix.rs:140-153
            const uint64_t *cs = c.ctrls(op);
            std::vector<uint64_t> rows(cs, cs + op.nctrl);
            for (auto &q : rows) q += uint64_t(n);
            const int t = int(op.q0);
            low.controlled(rows.data(), int(op.nctrl), n + t, c.mat(op));
            low.controlled(cs, int(op.nctrl), t, conj2(c.mat(op)).data());
            if (!nz.k1.empty()) {
                low.dense2(n + t, t, nz.s1.data());
                for (uint32_t k = 0; k < op.nctrl; ++k) low.dense2(n + int(cs[k]), int(cs[k]), nz.s1.data());
            }
            ++st.logical_gates;
            ++i;
            break;
        }
        case OpKind::Measure:
            flush_segment(v, low, cfg, st, prof);
            i = dm_measure_run(d, c, i, rng, nz.readout_error, classical);
            break;
        case OpKind::If:
            flush_segment(v, low, cfg, st, prof);
            i = dm_plain_op(d, c, i, nz, rng, classical, true, st);
            break;
        }
    }
    flush_segment(v, low, cfg, st, prof);
    st.kernels = prof.collect();
}

}  // namespace ev
