"""Host-side checks that need no GPU: the C-ABI library loads and exports
every entry point include/everett.h declares, circuit validation follows
src/circuit.rs:245-306, and the native generators, RNG, gate constructors and
Kraus tables agree bit for bit with the oracle."""
import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

import everett_amd as ev
import oracle
from everett_amd import _ffi

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    text = (ROOT / "include" / "everett.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ev_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _ffi.lib()
    names = header_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []
    # and the Python binding declares a prototype for each of them
    assert sorted(set(names) - set(_ffi.PROTOTYPES)) == []


def test_op_layout_matches_header():
    assert _ffi.OP_DTYPE.itemsize == 408
    assert _ffi.OP_DTYPE.fields["m"][1] == 152


STRUCT_MIRRORS = {  # C struct in include/everett.h -> its ctypes mirror in everett_amd/_ffi.py
    "ev_error": "EvError", "ev_rng": "EvRng", "ev_run_options": "EvRunOptions", "ev_channel": "EvChannel",
    "ev_noise_model": "EvNoiseModel", "ev_run_stats": "EvRunStats", "ev_kernel_stat": "EvKernelStat",
    "ev_batch_stats": "EvBatchStats", "ev_server_config": "EvServerConfig", "ev_request_info": "EvRequestInfo",
    "ev_server_stats": "EvServerStats", "ev_tune_entry": "EvTuneEntry", "ev_plan_info": "EvPlanInfo",
    "ev_compute_rates": "EvComputeRates",
}


def test_struct_mirrors_match_header(tmp_path):
    # sizeof and every field offset as the C compiler lays the header out, against ctypes
    import subprocess
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "everett.h"', "int main(void) {"]
    for cname, pyname in STRUCT_MIRRORS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in getattr(_ffi, pyname)._fields_:
            lines.append(f'printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(line.split()[:2]): int(line.split()[2]) for line in out if line}
    for cname, pyname in STRUCT_MIRRORS.items():
        cls = getattr(_ffi, pyname)
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for f in cls._fields_:
            assert got[(cname, f[0])] == getattr(cls, f[0]).offset, (cname, f[0])


# ------------------------------------------------------------- validation

def test_validate_order_and_errors():  # circuit.rs:252-306 and its tests
    c = ev.Circuit(2)
    c.h(0).cnot(0, 1)
    c.validate()
    c = ev.Circuit(2)
    c.h(2)
    with pytest.raises(ev.QubitOutOfRange) as e:
        c.validate()
    assert (e.value.qubit, e.value.num_qubits) == (2, 2)
    c = ev.Circuit(3)
    c.cnot(1, 1)
    with pytest.raises(ev.DuplicateQubit) as e:
        c.validate()
    assert e.value.qubit == 1
    # a out of range is reported before a == b
    c = ev.Circuit(3)
    c.cnot(5, 5)
    with pytest.raises(ev.QubitOutOfRange):
        c.validate()
    # Controlled: target checked first, then each control (range before duplicate)
    c = ev.Circuit(3)
    c.controlled([0, 2], ev.Gate1.x(), 2)
    with pytest.raises(ev.DuplicateQubit):
        c.validate()
    c = ev.Circuit(3)
    c.controlled([7], ev.Gate1.x(), 9)
    with pytest.raises(ev.QubitOutOfRange) as e:
        c.validate()
    assert e.value.qubit == 9
    # Measure: qubit before classical bit
    c = ev.Circuit.with_classical(1, 1)
    c.measure(0, 1)
    with pytest.raises(ev.ClassicalBitOutOfRange) as e:
        c.validate()
    assert (e.value.bit, e.value.num_classical) == (1, 1)
    # IfClassic: bit first, then the guarded op
    c = ev.Circuit.with_classical(1, 1)
    c.x_if(0, 3)
    with pytest.raises(ev.QubitOutOfRange):
        c.validate()
    c = ev.Circuit.with_classical(1, 1)
    c.x_if(2, 3)
    with pytest.raises(ev.ClassicalBitOutOfRange):
        c.validate()
    # the first failing op wins
    c = ev.Circuit(2)
    c.h(9).cnot(1, 1)
    with pytest.raises(ev.QubitOutOfRange):
        c.validate()


def test_dangling_if_is_rejected():
    c = ev.Circuit.with_classical(1, 1)
    c.if_bit(0)
    with pytest.raises(ev.InvalidArgument):
        c.validate()


def test_compose_appends_ops():  # circuit.rs:124-127
    a = ev.Circuit(3)
    a.h(0)
    b = ev.Circuit(3)
    b.cnot(0, 1).swap(1, 2)
    a.compose(b)
    assert len(a) == 3 and list(a.ops()["kind"]) == [1, 2, 2]
    a.compose(a)
    assert len(a) == 6


# --------------------------------------------------- bit-exact host parity

def test_rng_matches_oracle():
    for seed in (0, 1, 42, 2**63 + 5):
        a, b = ev.Rng.seed_from_u64(seed), oracle.Rng(seed)
        assert a.state == b.state
        assert [a.next_u64() for _ in range(64)] == [b.next_u64() for _ in range(64)]
        assert [a.next_f64() for _ in range(64)] == [b.next_f64() for _ in range(64)]


def test_python_gates_match_oracle_constructors():
    lib = oracle.lib()
    for theta in (0.0, 0.7, -1.3, 2.1, 5.5, 1e-9, 123.456):
        for name in ("rx", "ry", "rz", "phase"):
            out = np.zeros(4, dtype=np.complex128)
            getattr(lib, f"ref_gate_{name}").argtypes = [C.c_double, C.c_void_p]
            getattr(lib, f"ref_gate_{name}")(theta, C.c_void_p(out.ctypes.data))
            assert np.array_equal(getattr(ev.Gate1, name)(theta).m, out), (name, theta)
    for name in ("h", "x", "y", "z", "s", "t"):
        out = np.zeros(4, dtype=np.complex128)
        getattr(lib, f"ref_gate_{name}").argtypes = [C.c_void_p]
        getattr(lib, f"ref_gate_{name}")(C.c_void_p(out.ctypes.data))
        assert np.array_equal(getattr(ev.Gate1, name)().m, out), name


def _same_ops(a, b):
    assert len(a) == len(b)
    for f in ("kind", "num_controls", "target", "b", "controls"):
        assert np.array_equal(a[f], b[f]), f
    assert np.array_equal(a["m"].view(np.uint64), b["m"].view(np.uint64))  # bit for bit


@pytest.mark.parametrize("n", [1, 2, 3, 8, 30])
def test_qft_generator_matches_oracle(n):
    _same_ops(ev.algorithms.qft(n).ops(), oracle.gen_qft(n))


@pytest.mark.parametrize("n,depth,seed", [(1, 3, 0), (5, 4, 9), (20, 20, 20), (30, 2, 30)])
def test_brickwall_generator_matches_oracle(n, depth, seed):
    r1, r2 = ev.Rng.seed_from_u64(seed), oracle.Rng(seed)
    _same_ops(ev.workloads.brickwall(n, depth, r1).ops(), oracle.gen_brickwall(n, depth, r2))
    assert r1.state == r2.state


@pytest.mark.parametrize("n,depth,seed", [(1, 4, 1), (2, 3, 2), (7, 6, 7), (13, 30, 13)])
def test_dm_random_generator_matches_oracle(n, depth, seed):
    r1, r2 = ev.Rng.seed_from_u64(seed), oracle.Rng(seed)
    _same_ops(ev.workloads.dm_random(n, depth, r1).ops(), oracle.gen_dm_random(n, depth, r2))
    assert r1.state == r2.state


def test_iqft_is_inverse_structure():  # qft.rs:81-91
    ops = ev.algorithms.iqft(3).ops()
    assert list(ops["kind"]) == [2, 1, 3, 1, 3, 3, 1]
    assert abs(ops[2]["m"][3] - complex(math.cos(-math.pi / 2), math.sin(-math.pi / 2))) == 0


@pytest.mark.parametrize("kind,p", [("depolarizing", 0.013), ("amplitude_damping", 0.3), ("phase_damping", 0.7),
                                    ("bit_flip", 0.2), ("dephasing", 0.9)])
def test_kraus_tables_match_oracle(kind, p):
    ch = ev.Channel(kind, p)
    got = np.array([k.m for k in ch.kraus_ops()])
    want = oracle.channel_kraus(kind, p)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_noise_model_panics_become_value_errors():  # noise.rs:62-65
    with pytest.raises(ValueError):
        ev.NoiseModel.uniform_depolarizing(1.5)
    with pytest.raises(ValueError):
        ev.NoiseModel.amplitude_damping(-0.1)
    assert ev.NoiseModel.phase_damping(0.1).after_2q is None


def test_composite_noise_kraus_is_complete():
    ks = ev.workloads.composite_noise_kraus()
    assert len(ks) == 8
    acc = sum(k.m.reshape(2, 2).conj().T @ k.m.reshape(2, 2) for k in ks)
    assert np.max(np.abs(acc - np.eye(2))) < 1e-14


def test_workload_sizes():
    assert len(ev.workloads.config0()) == 1390
    assert len(ev.workloads.config1()) == 4660
    c2, noise = ev.workloads.config2()
    assert c2.num_qubits() == 13 and c2.num_classical() == 13 and noise.readout_error == 0.01


def test_tune_table_roundtrip_and_plan_uses_it():
    from tests.fused_emulator import get_plan

    assert ev.tune_get(31, "c64") is None  # the shipped table covers 8..30
    ev.tune_set(31, "c64", {"tile_qubits": 9, "low_qubits": 3, "max_gates_per_pass": 7})
    e = ev.tune_get(31, "c64")
    assert (e["tile_qubits"], e["low_qubits"], e["max_gates_per_pass"]) == (9, 3, 7)
    c = ev.workloads.config0(31, 3)
    _, _, _, _, t = get_plan(c, precision="c64")
    assert t == 9
    _, _, _, _, t = get_plan(c, precision="c64", tile_qubits=11)  # explicit options win
    assert t == 11
    with pytest.raises(ev.InvalidArgument):
        ev.tune_set(31, "c64", {"tile_qubits": 14, "low_qubits": 3, "max_gates_per_pass": 7})


@pytest.mark.parametrize("low", [1, 2, 3, 5, 6])
def test_low_qubit_option_in_plans(low):
    from tests.fused_emulator import check_plan, emulate, get_plan

    c = ev.workloads.config1(10, 5)
    p, s, o, cf, t, tail = get_plan(c, tile_qubits=8, low_qubits=low, with_tail=True)
    check_plan(10, p, s, o, cf)
    for P in p:
        assert list(P["q"][:low]) == list(range(low))
    rng = np.random.default_rng(low)
    psi = rng.standard_normal(1 << 10) + 1j * rng.standard_normal(1 << 10)
    want = psi.copy()
    oracle.sv_apply_ops(c.ops(), want)
    assert np.max(np.abs(emulate(psi, 10, p, s, o, cf, tail) - want)) < 1e-12
