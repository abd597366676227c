"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (gcc):

* the C++ header's substrate-file parser and RPSLatticeInitializer, through
  the host-only check program tests/host/compat_check.cpp;
* the CPU oracle (``make -C oracle sanitize``), loaded by a Python process
  started with both sanitizer runtimes preloaded, running the oracle's own
  CPU tests;
* a positive control: the same flags stop a program with a planted
  out-of-bounds read and a planted signed overflow, so a clean run above
  means something.

Any sanitizer finding is fatal (-fno-sanitize-recover=all).  CPU only.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined,float-cast-overflow",
       "-fno-sanitize-recover=all"]
ORACLE_TESTS = ["tests/test_rng.py", "tests/test_oracle_physics.py", "tests/test_oracle_permeable.py",
                "tests/test_unit_cell.py", "tests/test_geometry.py", "tests/test_mag_oracle.py",
                "tests/test_substrates.py", "tests/test_analysis.py"]


def _runtimes():
    libs = []
    for name in ("libasan.so", "libubsan.so"):
        p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
        if not os.path.isabs(p) or not os.path.exists(p):
            pytest.skip(f"gcc has no {name}")
        libs.append(p)
    return libs


def _env(**extra):
    env = dict(os.environ)
    env.update(ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    env.update(extra)
    return env


def test_sanitizers_stop_planted_errors(tmp_path):
    _runtimes()
    src = tmp_path / "planted.c"
    src.write_text(r'''
#include <stdio.h>
#include <stdlib.h>
#include <limits.h>
int main(int argc, char** argv) {
    int* a = malloc(4 * sizeof(int));
    for (int i = 0; i < 4; ++i) a[i] = i;
    int k = atoi(argv[1]);
    if (k == 0) printf("%d\n", a[4]);              /* one past the end */
    else { int m = INT_MAX - 1 + k; printf("%d\n", m); }   /* overflows for k >= 2 */
    free(a);
    return 0;
}
''')
    exe = tmp_path / "planted"
    subprocess.run(["gcc", *SAN, str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), "0"], capture_output=True, text=True, env=_env())
    # whichever sees it first: ASan's redzone or UBSan's object-size check
    assert r.returncode != 0 and ("heap-buffer-overflow" in r.stderr or "insufficient space" in r.stderr)
    r = subprocess.run([str(exe), "2"], capture_output=True, text=True, env=_env())
    assert r.returncode != 0 and "signed integer overflow" in r.stderr
    r = subprocess.run([str(exe), "1"], capture_output=True, text=True, env=_env())
    assert r.returncode == 0


def test_compat_parser_and_initializer_are_clean(tmp_path):
    _runtimes()
    exe = tmp_path / "compat_check"
    subprocess.run(["g++", "-std=c++17", *SAN, "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "host", "compat_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, env=_env(), timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert r.stdout.strip().splitlines()[-1].endswith(" 0 failed")


def test_oracle_tests_pass_on_the_sanitized_oracle():
    libs = _runtimes()
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    so = os.path.join(ROOT, "oracle", "_san", "liboracle.so")
    ldd = subprocess.run(["ldd", so], capture_output=True, text=True).stdout
    assert "libasan" in ldd and "libubsan" in ldd
    env = _env(LD_PRELOAD=":".join(libs), AXX_ORACLE_LIB=so, PYTHONDONTWRITEBYTECODE="1")
    probe = subprocess.run([sys.executable, "-c",
                            "import os, oracle.pyoracle as po; "
                            "assert po.lib()._name == os.environ['AXX_ORACLE_LIB']; print('sanitized oracle')"],
                           cwd=ROOT, capture_output=True, text=True, env=env, timeout=120)
    assert probe.returncode == 0 and "sanitized oracle" in probe.stdout, probe.stderr
    r = subprocess.run([sys.executable, "-m", "pytest", *ORACLE_TESTS, "-q", "-x", "-m", "not gpu",
                        "-p", "no:cacheprovider"], cwd=ROOT, capture_output=True, text=True, env=env, timeout=1200)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
