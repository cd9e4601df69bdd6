"""AddressSanitizer + UndefinedBehaviorSanitizer runs of the host code, on the
CPU (the checks the reference gets from Miri on its unsafe kernel code,
.github/workflows/ci.yml:28-41, restated for this build's C++ and C).

- The library's host side (circuits and their control continuation records,
  validation, the OpenQASM reader and writer on the reference corpus and on
  malformed text, the planner through ev_debug_plan, the tuning table, the
  generators) built with `everett_amd/build.py --sanitize` and loaded under
  clang's ASan runtime.
- The oracle (`make -C oracle sanitize`) under gcc's ASan and UBSan runtimes:
  every kernel, the backends, the density matrix with noise, the sampler and
  the generators.

Each driver runs in a child process with the runtime preloaded; any report
aborts it (halt_on_error), which fails the test.
"""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SAN_LIB = ROOT / "everett_amd" / "_lib" / "libeverett_san.so"
ORACLE_SAN = ROOT / "oracle" / "build" / "liboracle_san.so"


def _newer_sources(lib: Path, sources) -> bool:
    return not lib.exists() or any(p.stat().st_mtime > lib.stat().st_mtime for p in sources)


def _clang_asan_runtime() -> str:
    libs = sorted((Path("/opt/rocm/llvm/lib/clang")).glob("*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not libs:
        pytest.skip("clang's ASan runtime is not installed")
    return str(libs[-1])


def _gcc_runtime(name: str) -> str:
    out = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    if not out or not Path(out).exists():
        pytest.skip(f"gcc's {name} is not installed")
    return out


def _run(code: str, env: dict, timeout: int = 600):
    e = dict(os.environ)
    e.update(env)
    e["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=1"
    e["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    res = subprocess.run([sys.executable, "-c", code], cwd=str(ROOT), env=e, capture_output=True, text=True,
                         timeout=timeout)
    assert res.returncode == 0 and "runtime error" not in res.stderr and "ERROR: AddressSanitizer" not in res.stderr, \
        (res.returncode, res.stdout[-2000:], res.stderr[-6000:])
    return res.stdout


LIB_DRIVER = r'''
import numpy as np
import everett_amd as ev
from everett_amd import _ffi
from tests.fused_emulator import get_plan
from tests.test_qasm import CORPUS, ERRORS
assert str(_ffi.LIB_PATH).endswith("libeverett_san.so")
# circuits, continuation records, validation
for nc in (0, 1, 16, 17, 33, 70):
    c = ev.Circuit(64)
    c.controlled([(5 * i) % 63 + 1 for i in range(nc)], ev.Gate1.x(), 0)
    c.h(3).cnot(3, 4).measure(5, 0) if False else None
    recs = c.ops()
    c2 = ev.Circuit(64).append_ops(recs)
    assert np.array_equal(c2.ops(), recs)
    c.validate()
for bad in (lambda c: c.h(70), lambda c: c.cnot(2, 2), lambda c: c.measure(1, 9)):
    c = ev.Circuit.with_classical(8, 2)
    bad(c)
    try:
        c.validate()
        raise SystemExit("validate missed an error")
    except ev.EverettError:
        pass
recs = _ffi.controlled_records(0, np.array([0, 1, 1, 0]), list(range(1, 21)))
for part in (recs[:1], recs[1:]):
    try:
        ev.Circuit(24).append_ops(part)
        raise SystemExit("malformed records accepted")
    except ev.EverettError:
        pass
# OpenQASM: the corpus, its round trip, the error cases, and truncated inputs
n_ok = 0
for name, src, *_ in CORPUS:
    try:
        c = ev.Circuit.from_qasm(src)
    except ev.QasmError:
        continue
    n_ok += 1
    try:
        ev.Circuit.from_qasm(c.to_qasm())
    except ev.QasmError:
        pass
    for cut in range(0, len(src), 7):
        try:
            ev.Circuit.from_qasm(src[:cut])
        except ev.QasmError:
            pass
for case in ERRORS:
    try:
        ev.Circuit.from_qasm(case[0])
    except ev.QasmError:
        pass
for junk in (b"\xff\xfe", b"OPENQASM 3.0;\nqubit[2] q;\nctrl(99999999999) @ x q[0], q[1];\n", b"\"abc"):
    try:
        ev.Circuit.from_qasm(junk)
    except ev.QasmError:
        pass
assert n_ok > 5
# the planner (host) on the benchmark families, both precisions, several tile widths
for n, depth in ((10, 4), (16, 6), (20, 3)):
    c = ev.workloads.config1(n, depth)
    for prec in ("c128", "c64"):
        for tile in (0, 8, 12, 13):
            get_plan(c, prec, tile if tile <= n else 0)
c = ev.Circuit(12)
for q in range(11):
    c.controlled(list(range(q + 1, 12)), ev.Gate1.h(), q)
get_plan(c)
cd, _ = ev.workloads.config2(5, 3)
# tuning table and generators
ev.tune_set(31, "c64", {"tile_qubits": 9, "low_qubits": 3, "max_gates_per_pass": 7})
assert ev.tune_get(31, "c64")["tile_qubits"] == 9
for i in range(8):
    ev.workloads.config3_circuit(i)
    ev.workloads.config4_request(i)
ev.algorithms.qft(17).to_qasm()
print("lib host paths clean")
'''

ORACLE_DRIVER = r'''
import numpy as np
import oracle
from tests.helpers import gate_set, ops_array, random_state
assert str(oracle.LIB_PATH).endswith("liboracle_san.so")
rng = np.random.default_rng(1)
for n in (1, 3, 6):
    for name, g in gate_set().items():
        for k in range(n):
            a = random_state(rng, n)
            oracle.apply_1q(a, k, g)
            if n >= 2:
                oracle.apply_2q(a, k, (k + 1) % n, np.kron(g.reshape(2, 2), np.eye(2)).reshape(-1))
                oracle.apply_controlled_1q(a, [(k + 1) % n], k, g)
ops = oracle.gen_qft(7)
amps, _ = oracle.sv_run(ops, 7)
ops = ops_array([(1, 0, 0, oracle.gate1("h"), None), (4, 0, 0, None, None), (5, 0, 0, None, None),
                 (1, 1, 0, oracle.gate1("x"), None), (3, 2, 0, oracle.gate1("x"), [0, 1] * 10)])
oracle.sv_run(ops, 4, 3, num_classical=1)
b = oracle.gen_brickwall(8, 3, oracle.Rng(2))
oracle.sv_run(b, 8)
for kind in ("depolarizing", "amplitude_damping", "phase_damping", "bit_flip", "dephasing"):
    k = oracle.channel_kraus(kind, 0.1)
    oracle.dm_run(oracle.gen_dm_random(3, 2, oracle.Rng(5)), 3,
                  {"custom_1q": k, "custom_2q": k, "readout_error": 0.05}, 3, 0)
amps = random_state(rng, 9)
oracle.sample(amps, 9, np.linspace(0, 0.999, 64))
oracle.state_normalize(amps)
print("oracle clean")
'''


def test_library_host_code_under_asan_ubsan():
    srcs = list((ROOT / "everett_amd" / "csrc").glob("*"))
    if _newer_sources(SAN_LIB, srcs):
        subprocess.run([sys.executable, str(ROOT / "everett_amd" / "build.py"), "--sanitize"], check=True,
                       capture_output=True, timeout=900)
    out = _run(LIB_DRIVER, {"LD_PRELOAD": _clang_asan_runtime(), "EVERETT_AMD_LIB": str(SAN_LIB)})
    assert "lib host paths clean" in out


def test_oracle_under_asan_ubsan():
    subprocess.run(["make", "-C", str(ROOT / "oracle"), "-s", "sanitize"], check=True, capture_output=True,
                   timeout=300)
    pre = f"{_gcc_runtime('libasan.so')}:{_gcc_runtime('libubsan.so')}"
    out = _run(ORACLE_DRIVER, {"LD_PRELOAD": pre, "ORACLE_LIB": str(ORACLE_SAN)})
    assert "oracle clean" in out
