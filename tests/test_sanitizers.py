"""Host code and the oracle under sanitizers (CPU only): the library's host logic
(csrc/hc_host.cpp, csrc/hc_lcs.cpp; tests/cpp/sanitize_host.cpp) and the oracle's C restatement
(oracle/hc_oracle.c; tests/cpp/sanitize_oracle.c) built with AddressSanitizer +
UndefinedBehaviorSanitizer (any report is fatal), and with ThreadSanitizer - the host logic's frames
computed from several std::threads, the oracle's OpenMP loops with LLVM's OpenMP runtime and its
Archer tool (which makes ThreadSanitizer understand OpenMP synchronisation)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = [os.path.join(ROOT, "tests", "cpp", "sanitize_host.cpp"), os.path.join(ROOT, "discamb_amd", "csrc", "hc_host.cpp"),
        os.path.join(ROOT, "discamb_amd", "csrc", "hc_lcs.cpp")]
ORACLE = os.path.join(ROOT, "tests", "cpp", "sanitize_oracle.c")
LLVM = "/opt/rocm/llvm"


def _build_and_run(cmd, exe, env_extra):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, OMP_NUM_THREADS="4", **env_extra)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.stdout[-2000:], r.stderr[-5000:])
    assert "Sanitizer" not in r.stderr, r.stderr[-5000:]


def _need(tool):
    if shutil.which(tool) is None and not os.path.exists(tool):
        pytest.skip("%s not available" % tool)


def test_host_logic_asan_ubsan(tmp_path):
    _need("g++")
    exe = str(tmp_path / "host_asan")
    _build_and_run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-I" + ROOT] + HOST + ["-pthread", "-o", exe], exe,
                   {"ASAN_OPTIONS": "detect_leaks=1"})


def test_host_logic_threads_tsan(tmp_path):
    _need("g++")
    exe = str(tmp_path / "host_tsan")
    _build_and_run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-I" + ROOT] + HOST + ["-pthread", "-o", exe],
                   exe, {"TSAN_OPTIONS": "halt_on_error=1"})


def test_oracle_asan_ubsan(tmp_path):
    _need("gcc")
    exe = str(tmp_path / "oracle_asan")
    _build_and_run(["gcc", "-std=c99", "-O1", "-g", "-fno-omit-frame-pointer", "-fopenmp",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", ORACLE, "-lm", "-o", exe], exe,
                   {"ASAN_OPTIONS": "detect_leaks=1"})


def test_oracle_openmp_tsan(tmp_path):
    clang = os.path.join(LLVM, "bin", "clang")
    archer = os.path.join(LLVM, "lib", "libarcher.so")
    if not (os.path.exists(clang) and os.path.exists(archer)):
        pytest.skip("LLVM clang with the Archer OpenMP tool not available")
    exe = str(tmp_path / "oracle_tsan")
    _build_and_run([clang, "-std=c99", "-O1", "-g", "-fopenmp", "-fsanitize=thread", ORACLE, "-lm", "-o", exe,
                    "-Wl,-rpath," + os.path.join(LLVM, "lib")], exe,
                   {"TSAN_OPTIONS": "ignore_noninstrumented_modules=1 halt_on_error=1",
                    "OMP_TOOL_LIBRARIES": archer})
