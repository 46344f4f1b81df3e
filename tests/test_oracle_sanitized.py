"""The oracle's CPU tests under AddressSanitizer and UndefinedBehaviorSanitizer: an instrumented
build of oracle/om_oracle.cpp (oracle/liboracle_san.so from `make liboracle_san.so`, g++ only,
no device code) is loaded by a child pytest with the ASan runtime preloaded; any invalid access or
undefined behaviour aborts the child."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
SELECTED = ["tests/test_oracle_pinning.py", "tests/test_oracle_rounding.py", "tests/test_oracle_near_boundary.py",
            "tests/test_oracle_failure_paths.py"]


def test_oracle_tests_pass_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", ORACLE, "liboracle_san.so"], check=True)
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], check=True, capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ORACLE_LIBRARY=os.path.join(ORACLE, "liboracle_san.so"),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               OMP_NUM_THREADS="4")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider", *SELECTED],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=1800)
    tail = (r.stdout[-4000:] + "\n" + r.stderr[-4000:])
    print(tail)
    assert r.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr
    # the child really ran the instrumented build
    probe = subprocess.run([sys.executable, "-c", "from oracle import oracle as O; import ctypes; "
                            "print(O.lib()._name); print(hasattr(O.lib(), '__asan_init') or "
                            "ctypes.CDLL(None).__getattr__('__asan_init') is not None)"],
                           cwd=ROOT, env=env, capture_output=True, text=True, timeout=120)
    assert probe.returncode == 0 and "liboracle_san.so" in probe.stdout, probe.stdout + probe.stderr
