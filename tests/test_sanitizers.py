"""Sanitizer builds of the CPU code: the oracle under AddressSanitizer + UndefinedBehaviorSanitizer
and under ThreadSanitizer (oracle/Makefile targets run-asan and run-tsan, driving every oracle
entry point through oracle/sanitize_check.c, its threaded commitment loop included), and the host
build of the device arithmetic and transcript-word routines (tests/host_arith_check.cpp) under
AddressSanitizer + UndefinedBehaviorSanitizer. Each run must end clean and with consistent
results."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ORACLE = os.path.join(ROOT, "oracle")
CSRC = os.path.join(ROOT, "damascus000_amd", "csrc")
Q = 2**112 - 75


@pytest.mark.parametrize("target", ["run-asan", "run-tsan"])
def test_oracle_under_sanitizer(target):
    if not shutil.which("make") or not shutil.which("cc"):
        pytest.skip("no C toolchain")
    r = subprocess.run(["make", "-s", "-C", ORACLE, target], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "all checks passed" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "WARNING: ThreadSanitizer" not in r.stderr
    assert "runtime error" not in r.stderr  # UndefinedBehaviorSanitizer


@pytest.fixture(scope="module")
def host_arith_asan(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("host_arith_asan") / "host_arith_check")
    # host-only build; -Xarch_host keeps the sanitizers on the host side of the compiler driver
    subprocess.run([hipcc, "--offload-host-only", "-O1", "-g", "-std=c++17", "-fno-omit-frame-pointer",
                    "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                    "-Xarch_host", "-fno-sanitize-recover=all", "-I", CSRC, "-o", exe,
                    os.path.join(HERE, "host_arith_check.cpp")], check=True, capture_output=True)
    return exe


def _run(exe, args, inp=""):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe] + args, input=inp, capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    return r.stdout


def test_host_arith_under_asan_ubsan(host_arith_asan):
    xs = [1, 2, 75, Q - 1, 2**111, 0] + [pow(3, k, Q) for k in range(1, 60, 7)]
    out = _run(host_arith_asan, ["inv"], "".join(f"{x:x}\n" for x in xs)).split()
    assert [int(v, 16) for v in out] == [pow(x, Q - 2, Q) for x in xs]
    assert "mismatches 0" in _run(host_arith_asan, ["random", "20000", "7"])
    # one transcript message (label, round, epoch seed, file id, three module lengths, coefficients)
    ns = (3, 2, 2)
    inp = f"poly 5 {bytes(range(32)).hex()} {bytes(range(32, 64)).hex()} {ns[0]} {ns[1]} {ns[2]}\n"
    inp += "".join(f"{(7 * i + 1) % Q:x}\n" for i in range(8 * sum(ns)))
    assert len(bytes.fromhex(_run(host_arith_asan, ["tx"], inp).strip())) > 0
