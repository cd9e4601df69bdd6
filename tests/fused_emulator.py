"""CPU emulation of the fused pass kernel (everett_amd/csrc/fused_kernels.hip).

Replays a plan from ev_debug_plan with the kernel's exact index arithmetic
(tile base by zero-bit insertion, the global gather in the first stage's
register mapping, XOR-swizzled LDS slots between stages, the global scatter
in the last stage's mapping, conditions) on a numpy state, asserting that
every HBM index is in range and every mapping is a permutation of the tile,
so the plan and the kernel's indexing are checked on the CPU before a GPU
runs them.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from everett_amd import _ffi
from everett_amd.errors import check

PASS_DTYPE = np.dtype([("t", "<i4"), ("nstages", "<i4"), ("stage_begin", "<i4"), ("ncomp", "<i4"),
                       ("q", "<i4", (16,)), ("comp", "<i4", (40,)), ("swz_a", "<i4"), ("swz_b", "<i4")])
STAGE_DTYPE = np.dtype([("op_begin", "<i4"), ("op_end", "<i4"), ("np", "<i4"), ("relabel", "<i4"),
                        ("roff", "<u4", (16,)), ("goff", "<u8", (16,)), ("p", "<i4", (12,)), ("gpos", "<i4", (12,)),
                        ("rimg", "<u4", (12,)), ("rroff", "<u4", (16,)), ("naff", "<i4"), ("abit", "<u4", (8,)),
                        ("aoff", "<u4", (8,)), ("gneed", "<i4")])
OP_DTYPE = np.dtype([("code", "<u4"), ("creg", "<u4"), ("cglob", "<u8"), ("coef", "<u4"), ("tq", "<u4"),
                     ("pad", "<u4", (2,)), ("m", "<c16", (4,))])


def decode(o):
    c = int(o["code"])
    return c & 0xFF, (c >> 8) & 0xFF, (c >> 16) & 0xFF, c >> 24

U1, CU1, CX, DIAGR, DIAGT, U2, SWAP, PHASE = range(8)
PHASE_ALL = 4
MAX_PHASE_TERMS = 8


def get_plan(circuit, precision="c128", tile_qubits=0, max_gates_per_pass=0, low_qubits=0, with_tail=False):
    """The plan a single run executes: passes, stages, ops, coefficients and
    the tile width, and with with_tail the qubit pairs of the trailing SWAP
    network it runs as one permutation pass after the passes."""
    from everett_amd.backend import run_options

    opt = run_options(precision, False, tile_qubits, max_gates_per_pass, low_qubits=low_qubits)
    info = _ffi.EvPlanInfo()
    check(_ffi.lib().ev_debug_plan(circuit.handle, C.byref(opt), None, None, None, None, C.byref(info)))
    assert info.pass_bytes == PASS_DTYPE.itemsize
    assert info.stage_bytes == STAGE_DTYPE.itemsize
    assert info.op_bytes == OP_DTYPE.itemsize
    passes = np.zeros(info.num_passes, dtype=PASS_DTYPE)
    stages = np.zeros(info.num_stages, dtype=STAGE_DTYPE)
    ops = np.zeros(max(info.num_ops, 1), dtype=OP_DTYPE)
    coef = np.zeros(max(info.num_coef, 1), dtype=np.complex128)
    check(_ffi.lib().ev_debug_plan(circuit.handle, C.byref(opt), _ffi.ptr(passes), _ffi.ptr(stages), _ffi.ptr(ops),
                                   _ffi.ptr(coef), C.byref(info)))
    out = (passes, stages, ops[: info.num_ops], coef[: info.num_coef], info.tile_qubits)
    if with_tail:
        out += ([(int(info.swap_a[k]), int(info.swap_b[k])) for k in range(info.num_swap_pairs)],)
    return out


def swz(l, a, b):
    """The pass's LDS slot swizzle: the low 6 bits XORed with the bits a and b
    places above them."""
    return l ^ ((l >> a) & 63) ^ ((l >> b) & 63)


def _coef_span(kind, b1=0):
    return {U2: 16, PHASE: b1}.get(kind, 0)


def phase_masks(o, k):
    """The k term masks of a PHASE op (64-bit patterns stored in m)."""
    return np.frombuffer(o["m"].tobytes(), dtype="<u8")[:k]


def check_plan(n, passes, stages, ops, coef, coalesce=3):
    """Structural bounds of a plan, independent of any state.

    The first and last stage of a pass must keep their register bits off the
    `coalesce` lowest local positions (3 for c128, 4 for c64) wherever those
    are the state's lowest qubits, so the global gather and scatter move
    128-byte runs."""
    for P in passes:
        t0 = int(P["t"])
        low = 0
        while low < t0 and int(P["q"][low]) == low:
            low += 1
        lowc = (1 << min(coalesce, low)) - 1
        for s in (0, int(P["nstages"]) - 1):
            st = stages[P["stage_begin"] + s]
            regpos = 0
            for j in range(4):
                gq = int(st["goff"][1 << j]).bit_length() - 1
                regpos |= 1 << list(P["q"][:t0]).index(gq)
            assert not (regpos & lowc), "first/last stage registers must leave the coalesced low qubits to lanes"
            assert [int(x) for x in st["p"][: min(coalesce, low)]] == list(range(min(coalesce, low)))
        t = int(P["t"])
        sa, sb = int(P["swz_a"]), int(P["swz_b"])
        assert 1 <= sa < sb <= 31, "the swizzle folds bits from above"
        assert 8 <= t <= 13 and t <= n and P["ncomp"] == n - t
        qs = sorted(list(P["q"][:t]) + list(P["comp"][: n - t]))
        assert qs == list(range(n)), "tile and complement must partition the qubits"
        assert list(P["q"][:t]) == sorted(P["q"][:t])
        assert P["nstages"] >= 1 and P["stage_begin"] + P["nstages"] <= len(stages)
        W = 1 << (t - 4)
        tid = np.arange(W)
        for s in range(P["nstages"]):
            st = stages[P["stage_begin"] + s]
            assert st["np"] == t - 4
            assert 0 <= st["op_begin"] <= st["op_end"] <= len(ops)
            lb = np.zeros(W, dtype=np.int64)
            for i in range(t - 4):
                lb |= ((tid >> i) & 1) << int(st["p"][i])
                assert P["q"][st["p"][i]] == st["gpos"][i]
            slots = swz(lb, sa, sb)[:, None] ^ st["roff"][None, :].astype(np.int64)
            assert slots.min() >= 0 and slots.max() < (1 << t)
            assert len(np.unique(slots)) == (1 << t), "stage mapping must be a permutation of the tile"
            rslots = _read_slots(st, tid, t, lb, sa, sb)
            assert rslots.min() >= 0 and rslots.max() < (1 << t)
            assert len(np.unique(rslots)) == (1 << t), "relabelled read must be a permutation of the tile"
            if not int(st["relabel"]):
                # the kernel writes a stage's slots without a barrier after its
                # reads only when each thread writes exactly the slots it read
                assert all(set(rslots[w].tolist()) == set(slots[w].tolist()) for w in range(W))
            if s == 0:
                assert not int(st["relabel"]), "the first stage reads HBM, not a relabelled exchange"
            if not int(st["relabel"]):
                assert (st["rroff"] == st["roff"]).all() and int(st["naff"]) == 0
            assert 0 <= int(st["naff"]) <= 8
            for k in range(int(st["naff"])):
                b = int(st["abit"][k])
                assert b < n and b not in set(int(x) for x in P["q"][:t0]), "affine terms test tile-index bits"
                assert int(st["aoff"][k]) < (1 << t0)
            regq = set()
            for j in range(4):
                gj = int(st["goff"][1 << j])
                assert gj and gj & (gj - 1) == 0
                regq.add(gj.bit_length() - 1)
            if s == P["nstages"] - 1:
                assert int(st["gneed"]) == 1, "the scattering stage needs the lanes' global index bits"
            for o in ops[st["op_begin"]: st["op_end"]]:
                k, b0, b1, _ = decode(o)
                assert k <= PHASE and o["creg"] < 16
                if int(o["cglob"]) or k in (DIAGT, PHASE):
                    assert int(st["gneed"]) == 1, "an op that tests global index bits needs them computed"
                if k == PHASE:
                    assert b0 <= PHASE_ALL and 2 <= b1 <= MAX_PHASE_TERMS and o["cglob"] == 0
                    assert b0 == PHASE_ALL or not (int(o["creg"]) >> b0) & 1
                    for mk in phase_masks(o, b1):
                        assert int(mk) < (1 << n) and all(not (int(mk) >> q) & 1 for q in regq)
                else:
                    assert b0 < 4 and b1 < 4
                if k in (U2, SWAP):
                    assert b0 > b1
                assert o["coef"] + _coef_span(k, b1) <= len(coef)
                cg = int(o["cglob"])
                assert cg < (1 << n) and all(not (cg >> q) & 1 for q in regq)
                if k == DIAGT:
                    assert o["tq"] < n and o["tq"] not in regq


def _apply(o, v, gfix, coef):
    k, B, BL, flags = decode(o)
    creg, cglob = int(o["creg"]), np.uint64(o["cglob"])
    tok = (gfix & cglob) == cglob
    m = coef[int(o["coef"]):] if k == U2 else o["m"]
    condj = [((j & creg) == creg) for j in range(16)]
    if k in (U1, CU1):
        for j in range(16):
            if j & (1 << B):
                continue
            j1 = j | (1 << B)
            rows = tok if k == CU1 else np.ones_like(tok)
            if k == CU1 and not condj[j]:
                continue
            a0, a1 = v[:, j].copy(), v[:, j1].copy()
            v[:, j] = np.where(rows, m[0] * a0 + m[1] * a1, a0)
            v[:, j1] = np.where(rows, m[2] * a0 + m[3] * a1, a1)
    elif k == CX:
        for j in range(16):
            if j & (1 << B) or not condj[j]:
                continue
            j1 = j | (1 << B)
            a0, a1 = v[:, j].copy(), v[:, j1].copy()
            v[:, j] = np.where(tok, a1, a0)
            v[:, j1] = np.where(tok, a0, a1)
    elif k == DIAGR:
        one = bool(flags & 1)
        for j in range(16):
            if not condj[j]:
                continue
            if j & (1 << B):
                v[:, j] = np.where(tok, m[1] * v[:, j], v[:, j])
            elif not one:
                v[:, j] = np.where(tok, m[0] * v[:, j], v[:, j])
    elif k == DIAGT:
        d = np.where(((gfix >> np.uint64(o["tq"])) & np.uint64(1)) == 1, m[1], m[0])
        for j in range(16):
            if condj[j]:
                v[:, j] = np.where(tok, d * v[:, j], v[:, j])
    elif k == PHASE:
        f = np.ones(len(gfix), dtype=np.complex128)
        for i, mk in enumerate(phase_masks(o, BL)):
            mk = np.uint64(mk)
            f = np.where((gfix & mk) == mk, f * coef[int(o["coef"]) + i], f)
        for j in range(16):
            if (B == PHASE_ALL or j & (1 << B)) and condj[j]:
                v[:, j] = f * v[:, j]
    elif k == U2:
        M = m[:16].reshape(4, 4)
        for j in range(16):
            if j & ((1 << B) | (1 << BL)):
                continue
            idx = [j, j | (1 << BL), j | (1 << B), j | (1 << B) | (1 << BL)]
            a = v[:, idx].copy()
            v[:, idx] = a @ M.T
    else:
        for j in range(16):
            if j & ((1 << B) | (1 << BL)):
                continue
            a, b = j | (1 << BL), j | (1 << B)
            t = v[:, a].copy()
            v[:, a] = v[:, b]
            v[:, b] = t


def tile_base(T, q):
    """The kernel's tile base: a zero bit inserted at every tile qubit (ascending)."""
    b = T
    for k in q:
        lo = b & ((1 << k) - 1)
        b = ((b >> k) << (k + 1)) | lo
    return b


def _read_slots(st, tid, t, lb, sa, sb, base=0):
    """LDS slots a stage reads: its own mapping, or through the relabelling
    tables (linear images of the thread bits, affine terms on tile-index bits)."""
    if not int(st["relabel"]):
        return swz(lb, sa, sb)[:, None] ^ st["rroff"][None, :].astype(np.int64)
    rb = np.zeros(len(tid), dtype=np.int64)
    for i in range(t - 4):
        rb ^= ((tid >> i) & 1) * int(st["rimg"][i])
    for k in range(int(st["naff"])):
        if (base >> int(st["abit"][k])) & 1:
            rb ^= int(st["aoff"][k])
    return rb[:, None] ^ st["rroff"][None, :].astype(np.int64)


def _lane(st, tid, t, base):
    lb = np.zeros(len(tid), dtype=np.int64)
    gfix = np.full(len(tid), base, dtype=np.uint64)
    for i in range(t - 4):
        bit = ((tid >> i) & 1)
        lb |= bit << int(st["p"][i])
        gfix |= bit.astype(np.uint64) << np.uint64(st["gpos"][i])
    return lb, gfix


def emulate(psi, n, passes, stages, ops, coef, swap_tail=()):
    """Runs the plan on psi (complex128, length 2^n) in place, then the SWAP
    tail as the permutation pass does (amplitude x trades places with pi(x),
    pi exchanging the bits of every pair)."""
    for P in passes:
        t = int(P["t"])
        W = 1 << (t - 4)
        tid = np.arange(W, dtype=np.int64)
        q = [int(x) for x in P["q"][:t]]
        comp = [int(x) for x in P["comp"][: n - t]]
        nst = int(P["nstages"])
        sa, sb = int(P["swz_a"]), int(P["swz_b"])
        for T in range(1 << (n - t)):
            base = tile_base(T, q)
            want = 0
            for i, c in enumerate(comp):
                want |= ((T >> i) & 1) << c
            assert base == want, "tile base must deposit the tile index into the non-tile qubits"
            lds = np.full(1 << t, np.nan, dtype=np.complex128)
            tile_idx = None
            for s in range(nst):
                st = stages[P["stage_begin"] + s]
                lb, gfix = _lane(st, tid, t, base)
                if s == 0:
                    gidx = (gfix[:, None] | st["goff"][None, :].astype(np.uint64)).astype(np.int64)
                    assert gidx.max() < (1 << n) and len(np.unique(gidx)) == (1 << t)
                    tile_idx = set(gidx.ravel().tolist())
                    v = psi[gidx]
                else:
                    v = lds[_read_slots(st, tid, t, lb, sa, sb, base)]
                    assert not np.isnan(v).any(), "a stage read an LDS slot nobody wrote"
                for o in ops[st["op_begin"]: st["op_end"]]:
                    _apply(o, v, gfix, coef)
                if s + 1 < nst:
                    slots = swz(lb, sa, sb)[:, None] ^ st["roff"][None, :].astype(np.int64)
                    lds[:] = np.nan
                    lds[slots] = v
                else:
                    gidx = (gfix[:, None] | st["goff"][None, :].astype(np.uint64)).astype(np.int64)
                    assert set(gidx.ravel().tolist()) == tile_idx, "scatter must cover the gathered tile"
                    psi[gidx] = v
    if swap_tail:
        x = np.arange(1 << n, dtype=np.int64)
        y = x.copy()
        for a, b in swap_tail:
            d = ((x >> a) ^ (x >> b)) & 1
            y ^= (d << a) | (d << b)
        psi[:] = psi[y]
    return psi
