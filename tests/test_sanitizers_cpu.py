"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host code: the FP64 oracle
(oracle/ot_oracle.cpp, gcc) and the cell / conformity-check kernel code built for the host
(tests/hostemu, hipcc host pass).  Each build runs its own test module in a child process with the
matching sanitizer runtime preloaded; an invalid memory access or undefined behaviour aborts the
child (leak detection is off: the interpreter keeps its own allocations until exit)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _gcc_runtime(name):
    try:
        p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True,
                           check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return p if os.path.isabs(p) and os.path.exists(p) else None


def _clang_asan():
    base = "/opt/rocm/llvm/lib/clang"
    if not os.path.isdir(base):
        return None
    for v in sorted(os.listdir(base), reverse=True):
        p = os.path.join(base, v, "lib", "linux", "libclang_rt.asan-x86_64.so")
        if os.path.exists(p):
            return p
    return None


def _run_tests(modules, env_extra, preload):
    env = dict(os.environ)
    env.update(env_extra)
    env["LD_PRELOAD"] = preload
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        *modules], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=1200)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "passed" in r.stdout, tail


def test_oracle_under_asan_ubsan(tmp_path):
    asan, ubsan = _gcc_runtime("libasan.so"), _gcc_runtime("libubsan.so")
    if not (asan and ubsan):
        pytest.skip("gcc sanitizer runtimes not installed")
    lib = str(tmp_path / "liboracle_san.so")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san", f"SAN_OUT={lib}"],
                   check=True)
    _run_tests(["tests/test_oracle.py", "tests/test_shipped_run_cpu.py"],
               {"COT_ORACLE_LIB": lib}, f"{asan}:{ubsan}")


def test_hostemu_under_asan_ubsan(tmp_path):
    asan = _clang_asan()
    if not asan:
        pytest.skip("clang ASan runtime not installed")
    lib = str(tmp_path / "libhostemu_san.so")
    subprocess.run([os.path.join(ROOT, "tests", "hostemu", "build.sh"), lib], check=True,
                   env=dict(os.environ, SAN="1"), capture_output=True)
    _run_tests(["tests/test_hostemu_cpu.py"], {"COT_HOSTEMU_LIB": lib}, asan)
