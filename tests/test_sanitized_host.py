"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer: the library (geometry,
history, adapters' plans, setup tables, validation and the host half of the ABI) and the oracle
are built with both sanitizers (tests/sanitize/Makefile), and the CPU tests that drive them run
against those builds in a child process; any invalid access, leak-free heap misuse or undefined
behaviour aborts it."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "sanitize", "_build")
RUNTIME = "/opt/rocm/llvm/lib/clang"

SUITES = ["test_setup_pins.py", "test_geometry_pins.py", "test_history.py", "test_aero_plan.py",
          "test_aero2d_plan.py", "test_oracle_pins.py", "test_oracle_2d.py", "test_runner.py", "test_abi.py",
          "test_aa_index_map.py"]


def asan_runtime():
    out = subprocess.run(["/opt/rocm/llvm/bin/clang++", "-print-file-name=libclang_rt.asan-x86_64.so"],
                         capture_output=True, text=True)
    path = out.stdout.strip()
    return path if os.path.exists(path) else None


@pytest.mark.skipif(not os.path.exists("/opt/rocm/llvm/bin/clang++"), reason="ROCm clang absent")
def test_host_code_is_clean_under_asan_and_ubsan():
    build = subprocess.run(["make", "-s", "-j8", "-C", os.path.join(HERE, "sanitize")], capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-2000:]
    runtime = asan_runtime()
    assert runtime is not None
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": runtime,
        # The interpreter itself is not instrumented: its allocations are outside the check.
        "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
        "ORRERY_LBM_LIBRARY": os.path.join(BUILD, "liborrery_lbm.so"),
        "ORRERY_ORACLE_LIBRARY": os.path.join(BUILD, "liborrery_oracle.so"),
        "OMP_NUM_THREADS": "4",
    })
    # the builds are instrumented, and the child maps them (not the in-tree builds)
    for lib in ("liborrery_lbm.so", "liborrery_oracle.so"):
        syms = subprocess.run(["nm", "-D", os.path.join(BUILD, lib)], capture_output=True, text=True).stdout
        assert "__asan_report" in syms and "__ubsan_handle" in syms, lib
    probe = ("import sys; sys.path.insert(0, 'tests'); from orrery_amd import abi; import oracle; abi.load(); "
             "oracle.load(); print(open('/proc/self/maps').read())")
    maps = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=env, capture_output=True, text=True)
    assert maps.returncode == 0, maps.stderr[-2000:]
    assert os.path.join(BUILD, "liborrery_lbm.so") in maps.stdout
    assert os.path.join(BUILD, "liborrery_oracle.so") in maps.stdout
    suites = [os.path.join(HERE, s) for s in SUITES if os.path.exists(os.path.join(HERE, s))]
    run = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                          *suites], cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    tail = (run.stdout + run.stderr)[-4000:]
    print(tail)
    assert run.returncode == 0, tail
    assert "runtime error" not in run.stdout + run.stderr
    assert "AddressSanitizer" not in run.stdout + run.stderr
