"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

tests/cpp/sanitize_main.cpp drives csrc/nm_scene.cpp, csrc/nm_nvdb.cpp
(NanoVDB files, damaged ones included), the oracle and the header-only
Config/SensorData parser on small inputs.  It is built here with
g++ (no GPU code; the oracle keeps its -frounding-math -ffp-contract=off
flags) and run with every sanitizer report fatal.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "sanitize_main.cpp")


def _build(out):
    sources = [SRC, os.path.join(ROOT, "nanomap_amd", "csrc", "nm_scene.cpp"),
               os.path.join(ROOT, "nanomap_amd", "csrc", "nm_nvdb.cpp"),
               os.path.join(ROOT, "oracle", "nm_oracle.cpp"),
               os.path.join(ROOT, "oracle", "nm_soft_oracle.cpp")]
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-frounding-math", "-ffp-contract=off", "-fno-fast-math",
           "-I" + os.path.join(ROOT, "nanomap_amd", "host"),
           "-I" + os.path.join(ROOT, "include"), "-pthread"] + sources + \
          ["-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        if "cannot find" in r.stderr and "asan" in r.stderr:
            pytest.skip("the sanitizer runtimes are not installed")
        raise RuntimeError("sanitizer build failed:\n%s\n%s" % (" ".join(cmd), r.stderr))


def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "sanitize_main")
    _build(exe)
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True,
                       env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed checks" in r.stdout
    assert "runtime error" not in r.stderr
    assert "AddressSanitizer" not in r.stderr
