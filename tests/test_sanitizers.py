"""Host code under the sanitizers (CPU): the host-buffer entry's threaded copy
pool (ustb_amd/csrc/host_copy.h) under ThreadSanitizer and under
AddressSanitizer + UndefinedBehaviorSanitizer, and the MATLAB MEX gateway
(ustb_amd/mex/das_hip.cpp, with the test stand-in of the MEX API) on das_c's
argument errors under AddressSanitizer + UndefinedBehaviorSanitizer. GPU code
is not instrumented."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "ustb_amd", "csrc")


def _build_and_run(tmp_path, flags, name):
    exe = str(tmp_path / name)
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", *flags, "-I", CSRC,
                    os.path.join(ROOT, "tests", "sanitize", "copy_pool_test.cpp"), "-o", exe, "-lpthread"],
                   check=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="halt_on_error=1:detect_leaks=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert p.returncode == 0 and p.stdout.strip() == "ok", p.stdout + p.stderr


def test_copy_pool_thread_sanitizer(tmp_path):
    _build_and_run(tmp_path, ["-fsanitize=thread"], "copy_pool_tsan")


def test_copy_pool_address_and_undefined_sanitizers(tmp_path):
    _build_and_run(tmp_path, ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"], "copy_pool_asan")


def test_mex_gateway_argument_errors_under_asan(tmp_path):
    """The gateway's argument handling (das_c's checks through
    ustbx_das_check_mex_args) with every error case of the driver, its own
    code instrumented."""
    obj = str(tmp_path / "oracle.o")
    exe = str(tmp_path / "driver_asan")
    san = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g"]
    subprocess.run(["gcc", "-O1", "-std=gnu11", "-ffp-contract=off", *san, "-c",
                    os.path.join(ROOT, "oracle", "das_c_oracle.c"), "-o", obj], check=True)
    lib = os.path.join(ROOT, "ustb_amd")
    subprocess.run(["g++", "-O1", "-std=c++17", *san, "-I", os.path.join(ROOT, "tests", "mexshim"), "-I",
                    os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "mexshim", "driver.cpp"),
                    os.path.join(ROOT, "ustb_amd", "mex", "das_hip.cpp"), obj, "-L", lib, "-lustbx",
                    f"-Wl,-rpath,{lib}", "-lpthread", "-lm", "-o", exe], check=True)
    env = dict(os.environ, ASAN_OPTIONS="halt_on_error=1:detect_leaks=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    p = subprocess.run([exe, "errors"], capture_output=True, text=True, timeout=120, env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    assert len(p.stdout.strip().splitlines()) == 8, p.stdout
